// launch.h -- the one kernel-launch macro of the library.
//
// While resolve_only_flag() is set on the calling thread, FFTAMD_LAUNCH looks
// the kernel up (hipFuncGetAttributes, which loads its code object onto the
// current device) instead of launching it.  The planner runs a new plan once
// in that mode, so the code objects and twiddle tables a plan needs exist
// before its first timed call, as the reference builds its program at
// initialisation (api/src/fftfpga.c:116-126) and times only kernel execution
// (api/src/fft3d.c:336-340).
#pragma once
#include <hip/hip_runtime.h>

namespace fftamd {
bool& resolve_only_flag();

// One per kernel source file: looks one of its kernels up, which loads the
// file's code object onto the current device.  The runtime calls them all when
// it initialises a device, so that no later call -- one under stream capture
// included -- has a code object to load.
void preload_pass();
void preload_cube();
void preload_c512();
void preload_xy();
void preload_col();
void preload_plane();
void preload_multi();
void preload_flow();
inline void preload_kernel(const void* f) {
  hipFuncAttributes fa;
  (void)hipFuncGetAttributes(&fa, f);
}
}  // namespace fftamd

#define FFTAMD_LAUNCH(K, G, B, SHM, S, ...)                                       \
  do {                                                                            \
    if (::fftamd::resolve_only_flag()) {                                          \
      hipFuncAttributes fa_;                                                      \
      (void)hipFuncGetAttributes(&fa_, reinterpret_cast<const void*>(&K));        \
    } else {                                                                      \
      hipLaunchKernelGGL(K, G, B, SHM, S, __VA_ARGS__);                           \
    }                                                                             \
  } while (0)

// api.cpp -- the drop-in C ABI (include/fftfpga/fftfpga.h) over the HIP engine.
//
// This translation unit sees the reference's float2/double2 structs and no
// HIP header; the engine behind engine.h sees only plain pointers.  Argument
// validation mirrors the reference entry point by entry point (citations per
// function); runtime failures return valid = false with a message on stderr
// instead of the reference's exit() (api/src/opencl_utils.c:422-440).
#include "fftfpga/fftfpga.h"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>

#include "engine.h"

using fftamd::Timing;

namespace {

// Platform names this library answers to.  The two Intel names are the ones
// the reference's own harness passes (examples/fft.cpp:14-18), so callers
// written for the FPGA library initialise unchanged.
const char* const kPlatforms[] = {
    "AMD HIP (gfx950)",
    "Intel(R) FPGA SDK for OpenCL(TM)",
    "Intel(R) FPGA Emulation Platform for OpenCL(TM)",
};

std::string lower(const char* s) {
  std::string r(s);
  for (auto& ch : r) ch = (char)std::tolower((unsigned char)ch);
  return r;
}

// Case-insensitive substring search, as api/src/opencl_utils.c:24-85.
bool platform_matches(const char* name) {
  if (name == nullptr || std::strlen(name) == 0) return false;
  const std::string needle = lower(name);
  for (const char* p : kPlatforms)
    if (lower(p).find(needle) != std::string::npos) return true;
  return false;
}

bool pow2_ok(unsigned N) { return N != 0 && (N & (N - 1)) == 0; }

fpga_t invalid() {
  fpga_t t;
  std::memset(&t, 0, sizeof(t));
  t.valid = false;
  return t;
}

fpga_t to_fpga_t(const Timing& tm) {
  fpga_t t = invalid();
  t.pcie_read_t = tm.d2h_ms;
  t.pcie_write_t = tm.h2d_ms;
  t.exec_t = tm.exec_ms;
  t.svm_copyin_t = tm.stage_in_ms;
  t.svm_copyout_t = tm.stage_out_ms;
  t.valid = true;
  return t;
}

fpga_t run_host(int rank, unsigned N, size_t batch, const void* inp, void* out, bool inv, unsigned flags,
                int prec) {
  size_t dims[3] = {N, N, N};
  Timing tm;
  if (fftamd::host_c2c(rank, dims, batch, inp, out, inv, flags, prec, &tm) != 0) return invalid();
  return to_fpga_t(tm);
}

fpga_t run_mapped(int rank, unsigned N, size_t batch, const void* inp, void* out, bool inv,
                  unsigned flags) {
  size_t dims[3] = {N, N, N};
  Timing tm;
  if (fftamd::host_c2c_mapped(rank, dims, batch, inp, out, inv, flags, &tm) != 0) return invalid();
  return to_fpga_t(tm);
}

}  // namespace

extern "C" {

// api/src/fftfpga.c:72-130
int fpga_initialize(const char* platform_name, const char* path, const bool use_svm) {
  if (path == nullptr || std::strlen(path) == 0) return -1;
  if (!platform_matches(platform_name)) return -2;
  if (fftamd::runtime_device_count() <= 0) return -3;
  if (use_svm && !fftamd::runtime_can_map_host(0)) return -5;
  if (access(path, R_OK) != 0) {
    fftamd::runtime_final();
    return -4;
  }
  if (fftamd::runtime_init(0, use_svm) != 0) return -3;
  return 0;
}

int fftfpga_init_fpga(const char* platform_name, const char* path, const bool use_svm) {
  return fpga_initialize(platform_name, path, use_svm);
}

// api/src/fftfpga.c:135-142
void fpga_final(void) { fftamd::runtime_final(); }

// api/src/fftfpga.c:37-58 -> api/src/opencl_utils.c:205-213 (posix_memalign, 64 B)
void* fftfpga_complex_malloc(const size_t sz) {
  if (sz == 0) return nullptr;
  void* p = nullptr;
  if (posix_memalign(&p, 64, sz) != 0) return nullptr;
  return p;
}

void* fftfpgaf_complex_malloc(const size_t sz) {
  if (sz == 0) return nullptr;
  void* p = nullptr;
  if (posix_memalign(&p, 64, sz) != 0) return nullptr;
  return p;
}

// api/src/fft1d.c:27-137 (declared double, broken in the reference; implemented in FP64 here)
fpga_t fftfpga_c2c_1d(const unsigned N, const double2* inp, double2* out, const bool inv,
                      const unsigned iter) {
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || iter == 0) return invalid();
  return run_host(1, N, iter, inp, out, inv, fftamd::FLAG_BITREV_1D, 1);
}

// api/src/fft1d.c:148-247
fpga_t fftfpgaf_c2c_1d(const unsigned N, const float2* inp, float2* out, const bool inv,
                       const unsigned iter) {
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || iter == 0) return invalid();
  return run_host(1, N, iter, inp, out, inv, fftamd::FLAG_BITREV_1D, 0);
}

// api/src/fft1d.c:258-373 (requires use_svm at init, :264)
fpga_t fftfpgaf_c2c_1d_svm(const unsigned N, const float2* inp, float2* out, const bool inv,
                           const unsigned batch) {
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || !fftamd::runtime_svm() || batch == 0)
    return invalid();
  return run_mapped(1, N, batch, inp, out, inv, fftamd::FLAG_BITREV_1D);
}

// api/src/fft2d.c:170-338
fpga_t fftfpgaf_c2c_2d_bram(const unsigned N, const float2* inp, float2* out, const bool inv,
                            const bool interleaving, const unsigned how_many) {
  (void)interleaving;  // FPGA bank interleaving has no GPU counterpart
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || how_many == 0) return invalid();
  return run_host(2, N, how_many, inp, out, inv, 0u, 0);
}

// api/src/fft2d.c:349-505 (requires use_svm, :358)
fpga_t fftfpgaf_c2c_2d_bram_svm(const unsigned N, const float2* inp, float2* out, const bool inv,
                                const unsigned how_many) {
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || !fftamd::runtime_svm() || how_many == 0)
    return invalid();
  return run_mapped(2, N, how_many, inp, out, inv, 0u);
}

// api/src/fft2d.c:27-159
fpga_t fftfpgaf_c2c_2d_ddr(const unsigned N, const float2* inp, float2* out, const bool inv) {
  if (inp == nullptr || out == nullptr || !pow2_ok(N)) return invalid();
  return run_host(2, N, 1, inp, out, inv, 0u, 0);
}

// api/src/fft3d.c:30-196
fpga_t fftfpgaf_c2c_3d_bram(const unsigned N, const float2* inp, float2* out, const bool inv,
                            const bool interleaving) {
  (void)interleaving;
  if (inp == nullptr || out == nullptr || !pow2_ok(N)) return invalid();
  return run_host(3, N, 1, inp, out, inv, 0u, 0);
}

// api/src/fft3d.c:206-384 -- the hot path
fpga_t fftfpgaf_c2c_3d_ddr(const unsigned N, const float2* inp, float2* out, const bool inv) {
  if (inp == nullptr || out == nullptr || !pow2_ok(N)) return invalid();
  return run_host(3, N, 1, inp, out, inv, 0u, 0);
}

// api/src/fft3d.c:396-913 (how_many > 1, :403; exec_t is wall clock incl. copies, :489,861)
fpga_t fftfpgaf_c2c_3d_ddr_batch(const unsigned N, const float2* inp, float2* out, const bool inv,
                                 const bool interleaving, const unsigned how_many) {
  (void)interleaving;
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || how_many <= 1) return invalid();
  size_t dims[3] = {N, N, N};
  Timing tm;
  if (fftamd::host_c2c_pipelined(3, dims, how_many, inp, out, inv, &tm) != 0) return invalid();
  fpga_t t = invalid();
  t.exec_t = tm.exec_ms;
  t.valid = true;
  return t;
}

// api/src/fft3d_svm.c:31-232 (requires use_svm, :40)
fpga_t fftfpgaf_c2c_3d_ddr_svm(const unsigned N, const float2* inp, float2* out, const bool inv,
                               const bool interleaving) {
  (void)interleaving;
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || !fftamd::runtime_svm()) return invalid();
  return run_mapped(3, N, 1, inp, out, inv, 0u);
}

// api/src/fft3d_svm.c:244-519 (how_many > 0 and use_svm, :251)
fpga_t fftfpgaf_c2c_3d_ddr_svm_batch(const unsigned N, const float2* inp, float2* out, const bool inv,
                                     const unsigned how_many) {
  if (inp == nullptr || out == nullptr || !pow2_ok(N) || how_many == 0 || !fftamd::runtime_svm())
    return invalid();
  return run_mapped(3, N, how_many, inp, out, inv, 0u);
}

}  // extern "C"

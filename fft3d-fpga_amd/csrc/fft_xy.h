// fft_xy.h -- arguments of the fused x+y pass over 256 x 256 planes.
#pragma once

namespace fftamd {

struct XYArgs {
  const void* in;
  void* out;
  unsigned long long nplanes;    // planes in the array
  long long plane_stride;        // complex elements between planes
  const void* tw;                // Stockham twiddles of the 256-point line
  const void* tw_full;           // W_256^j, j = 0..255
  const float* w_load;           // optional real weights (see PassArgs)
  const float* w_store;
  unsigned long long w_mask;
  int inverse;
};

// 0 on success, -1 when the grid does not fit, -2 on a launch error.
// Out of place only: in and out must not overlap.
int launch_fft_xy256(const XYArgs& a, void* stream);

}  // namespace fftamd

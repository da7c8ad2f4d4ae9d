// fft_plane.h -- arguments of the plane passes: the 2D transform of every
// (LY x LX) plane of an array, both axes in one pass over HBM.
#pragma once

namespace fftamd {

struct PlaneArgs {
  const void* in;
  void* out;
  unsigned long long nplanes;    // planes in the array
  long long plane_stride;        // complex elements between planes
  const void* tw_x;              // Stockham twiddles of the LX-point rows
  const void* tw_y;              // Stockham twiddles of the LY-point columns
  const float* w_load;           // optional real weights (see PassArgs)
  const float* w_store;
  unsigned long long w_mask;
  int inverse;
  // optional transform table (planes up to 128 x 128 only): plane b belongs
  // to transform b / ppt, whose first input and output elements are
  // tab[2 t] and tab[2 t + 1] (complex elements from in / out); NULL: plane b
  // starts at b * plane_stride in both
  const unsigned long long* tab;
  unsigned ppt;
};

// 256 x 256 planes, four workgroups per plane (fft_xy.hip).  0 on success, -1
// when the grid does not fit, -2 on a launch error.  Out of place only: in
// and out must not overlap.
int launch_fft_xy256(const PlaneArgs& a, void* stream);

// LY x LX planes with 16 <= LY, LX <= 128 and LY * LX >= 1024, one workgroup
// per plane holding it in LDS (fft_plane.hip).  split: radix-16 stages on FP16
// matrix cores with hi+lo operands.  0 on success, -1 for an unsupported
// shape or grid, -2 on a launch error.  In place allowed.
int launch_fft_plane(int ly, int lx, const PlaneArgs& a, void* stream, bool split = false);
bool plane_shape_supported(long long ly, long long lx);

// The 4096-point transform along the second-fastest axis of blocks of 4096
// rows of lx elements (fft_col.hip): element (o, row, x) at o * in_o + row *
// lx + x.  Four workgroups per group of 16 adjacent columns.
struct ColArgs {
  const void* in;
  void* out;
  long long lx;                  // row length in complex elements (multiple of 16)
  long long in_o;                // elements between blocks (4096 * lx)
  unsigned long long nitems;     // blocks x column groups
  unsigned ngroups;              // column groups per block (lx / 16)
  const void* tw;                // Stockham twiddles of 256-point lines
  const void* w256;              // W256^m, m = 0..255
  const void* ptw_lo;            // W4096^t = ptw_hi[t >> ptw_hb] * ptw_lo[t & (2^ptw_hb - 1)]
  const void* ptw_hi;
  unsigned ptw_hb;
  const float* w_load;           // optional real weights (see PassArgs)
  const float* w_store;
  unsigned long long w_mask;
  int inverse;
};

// 0 on success, -1 for an unsupported shape or grid, -2 on a launch error.
// Out of place only: in and out must not overlap.
int launch_fft_col4096(const ColArgs& a, void* stream);
bool col4096_supported(long long lx, unsigned long long nouter);

// Arrays of n x n x n blocks (n = 512, 256 or 128) in two passes with the y
// axis split between them: (n/4) x 4 for n = 256 and 128 (fft_cube.hip),
// 64 x 8 for n = 512 (fft_cube512.hip).  Pass A: x and the larger y stage of
// every plane, out of place only.  Pass B: the smaller y stage and z of every
// block, in place allowed.  Pass A takes load weights only, pass B store
// weights only (see PassArgs).  0 on success, -1 for an unsupported shape or
// grid, -2 on a launch error.
struct CubeArgs {
  const void* in;
  void* out;
  int n;                         // block side: 512, 256 or 128
  unsigned long long nplanes;    // n x n planes in the array (a multiple of n for pass B)
  const void* tw;                // Stockham twiddles of n-point lines (n = 512: of the
                                 // 64-point lines of pass A, the 128-point lines of pass B)
  const void* w256;              // W256^m, m = 0..255
  const void* w512;              // W512^m, m = 0..511 (n = 512)
  const float* w_load;
  const float* w_store;
  unsigned long long w_mask;
  int inverse;
  int split;                     // radix-16 stages on FP16 MFMA with hi + lo operands (no weights)
  // optional transform table (n = 128 only, no weights): block t starts at
  // tab[2 t] in `in` and tab[2 t + 1] in `out` (complex elements); NULL:
  // block t at t n^3 in both
  const unsigned long long* tab;
};
int launch_fft_cube_a(const CubeArgs& a, void* stream);
int launch_fft_cube_b(const CubeArgs& a, void* stream);
int launch_fft_c512_a(const CubeArgs& a, void* stream);  // n = 512 (fft_cube512.hip)
int launch_fft_c512_b(const CubeArgs& a, void* stream);

}  // namespace fftamd

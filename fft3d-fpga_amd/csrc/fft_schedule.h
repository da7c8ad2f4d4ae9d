// fft_schedule.h -- compile-time radix schedule and launch geometry of the
// Stockham line-FFT pass kernels.  Shared by the device kernels and by the
// host code that builds twiddle tables and launch grids, so both always agree.
//
// A "line" is one 1D transform of L = 2^lg points.  A pass kernel holds C
// lines per workgroup, PPT points of each line per thread (TPL threads per
// line), and runs NP radix stages: the first of radix 2^(lg - 4(NP-1)), the
// rest radix 16.  Stage s sees sub-transforms of span NS(s) = prod of the
// radices before it (Stockham autosort: natural order in, natural order out).
#pragma once

namespace fftamd {

// MODE_PLANE: fused x+y pass over whole planes (fft_xy.hip for 256 x 256,
// fft_plane.hip up to 128 x 128); MODE_COL4096: the 4096-point second-fastest
// axis in one pass (fft_col.hip); MODE_CUBE_A / MODE_CUBE_B: the two passes of
// 256 x 256 x 256 blocks with the y axis split 64 x 4 (fft_cube.hip)
enum PassMode { MODE_CONTIG = 0, MODE_STRIDED = 1, MODE_PLANE = 2, MODE_COL4096 = 3, MODE_CUBE_A = 4,
                MODE_CUBE_B = 5 };
// EPI_TWIDDLE multiplies point k of column c by W_M^{(c >> cshift) * k}
// (four-step split of a long axis); EPI_BITREV stores point k at bitrev(k).
enum PassEpilogue { EPI_NONE = 0, EPI_TWIDDLE = 1, EPI_BITREV = 2 };
enum Precision { PREC_F32 = 0, PREC_F64 = 1 };

constexpr int ilog2c(long long x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }

// ---- radix schedule ------------------------------------------------------
constexpr int sched_np(int L) { return L <= 16 ? 1 : (ilog2c(L) + 3) / 4; }
constexpr int sched_radix(int L, int s) {
  return s == 0 ? (1 << (ilog2c(L) - 4 * (sched_np(L) - 1))) : 16;
}
constexpr int sched_ns(int L, int s) {
  int n = 1;
  for (int i = 0; i < s; ++i) n *= sched_radix(L, i);
  return n;
}
// Stages s >= 1 are radix 16.  Their twiddles sit at [twoff(s) + 4k + q],
// k < NS(s), holding W_{16 NS}^{k * 2^q} for q = 0..3; the other powers
// W^{k r} (r = 1..15) are formed by at most three complex products.
constexpr int TW_PER_K = 4;
constexpr int sched_twoff(int L, int s) {
  int o = 0;
  for (int i = 1; i < s; ++i) o += sched_ns(L, i) * TW_PER_K;
  return o;
}
constexpr int sched_twsize(int L) { return sched_twoff(L, sched_np(L)); }

// ---- geometry -------------------------------------------------------------
constexpr int geom_ppt(int L) { return L < 16 ? L : 16; }
constexpr int geom_tpl(int L) { return L / geom_ppt(L); }
// Lines per workgroup.  Contiguous lines: 256 threads per workgroup (one line
// per workgroup from L = 4096 up).  Strided lines: at least 16 adjacent
// columns (128-byte row segments; narrower segments measured far slower) and
// 256 threads up to L = 256; 32 columns (256-byte segments, 1024 threads) at
// L = 512, where rows lie 4 KiB to 2 MiB apart (measured 7-18% faster than 16
// on the y and z passes of a 512^3 array; slower at L = 256); 16 columns
// (1024 threads) at L = 1024, the four-step passes of long 1D axes.
constexpr int geom_cols(int L, int mode) {
  return mode == MODE_CONTIG ? (geom_tpl(L) >= 256 ? 1 : 256 / geom_tpl(L))
                             : (L >= 1024 ? 16 : L >= 512 ? 32 : L >= 256 ? 16 : 256 / geom_tpl(L));
}
constexpr int geom_threads(int L, int mode) { return geom_cols(L, mode) * geom_tpl(L); }
// Columns per workgroup of a line pass in either precision mode: split-
// precision strided lines of 512 points and more keep 16 columns (two
// workgroups per CU; with the FP32 layout's 32 columns, one, the split 512^3
// line passes took 1.80 against 1.59 ms).
constexpr int geom_cols_for(int L, int mode, bool split) {
  return split && mode == MODE_STRIDED && L >= 512 ? 16 : geom_cols(L, mode);
}
constexpr int geom_threads_for(int L, int mode, bool split) { return geom_cols_for(L, mode, split) * geom_tpl(L); }

// Largest line held by one workgroup (LDS budget).  Longer axes are split
// four-step into two passes.
constexpr int lmax_contig(int prec) { return prec == PREC_F32 ? 8192 : 4096; }
constexpr int lmax_strided(int prec) { return prec == PREC_F32 ? 512 : 256; }
// The strided (first) pass of a four-step split of a long contiguous axis may
// hold 1024-point lines (16 columns, 1024 threads, 139 KiB of LDS): 2^22 points
// as 1024 x 4096 measured 38.1 against 41.4 us for 512 x 8192.  Strided axes
// of 2D and 3D arrays keep lmax_strided.
constexpr int lmax_strided_fourstep(int prec) { return prec == PREC_F32 ? 1024 : 256; }

}  // namespace fftamd

// fft_plane.hip -- the 2D transform of every LY x LX plane of an array in one
// pass over HBM, for planes that fit one workgroup's LDS (LY, LX <= 128).
//
// One workgroup per plane, LY * LX / 16 threads, 16 points per thread in both
// phases.  Rows: each thread loads its first-stage butterfly inputs of one row
// straight from HBM (coalesced along the row), runs the Stockham stages with
// LDS exchanges ([row][x + x/16] layout), and writes the row's results into
// the column layout ([y][col], one pad column).  Columns: the same stages
// along y, and the last stage writes natural-order rows to HBM (coalesced
// along x).  A 3D transform of 128^3-sized planes then takes two passes over
// HBM instead of three, and the plane is read and written exactly once.
//
// The radix stages are the line passes' (fft_pass_dev.h), instantiated with a
// whole plane's rows or columns per workgroup.  128 x 128 FP32 planes use
// k_fft_plane_halves below (two workgroups per CU).
#include <hip/hip_runtime.h>

#include "launch.h"

#include "fft_pass_dev.h"
#include "fft_plane.h"
#include "fft_plane_dev.h"

namespace fftamd {

// SPLIT: radix-16 stages on FP16 matrix cores with hi+lo operands (fft_split.h).
template <int LY, int LX, bool INV, bool SPLIT>
__global__ void __launch_bounds__(LY * LX / 16) k_fft_plane(const PlaneArgs a) {
  __shared__ cpx<float> lds[PlaneGeom<LY, LX, SPLIT>::LDS];
  plane_tile<LY, LX, INV, SPLIT>(a, blockIdx.x, lds, (int)threadIdx.x);
}

// The 128 x 128 plane (128 KiB) in two halves: 512 threads hold the plane in
// registers, 32 points each, and LDS stages half of it at a time (rows 0..63,
// then 64..127; then columns 0..63, then 64..127).  LDS is then 68 KiB, so
// two workgroups share a CU and one loads or stores while the other
// computes; the whole-plane kernel above needs 136 KiB and runs one
// workgroup per CU.
template <int L, bool SPLIT>
struct HalfGeom {
  using GX = Geom<L, MODE_CONTIG, L / 2>;  // rows: L/2 lines of L points
  using GY = Geom<L, MODE_STRIDED, L / 2>; // columns: L/2 lines of L points
  static constexpr int NT = L * L / 32;
  static_assert(GX::NT == NT && GY::NT == NT, "16 points per thread per half");
  static constexpr int LDS = GX::LDS_ELEMS > GY::LDS_ELEMS ? GX::LDS_ELEMS : GY::LDS_ELEMS;
  // last row stage: radix 16, one butterfly per thread, so point k = tau + TPL r
  static_assert(sched_radix(L, GX::NP - 1) == 16 && sched_ns(L, GX::NP - 1) == GX::TPL,
                "points tau + TPL r: r < 8 are columns 0 .. L/2 - 1");
};

// WL, WS: the launch carries load or store weights.
template <int L, bool INV, bool SPLIT, bool WL, bool WS>
__global__ void __launch_bounds__(L * L / 32, 4) k_fft_plane_halves(const PlaneArgs a) {
  using HG = HalfGeom<L, SPLIT>;
  using GX = typename HG::GX;
  using GY = typename HG::GY;
  constexpr int H = L / 2;
  __shared__ cpx<float> lds[HG::LDS];
  const unsigned long long plane = blockIdx.x;
  long long base = (long long)plane * a.plane_stride, obase = base;
  if (a.tab) {  // planes of a batch of transforms placed anywhere
    const unsigned t = (unsigned)(plane / a.ppt), z = (unsigned)(plane % a.ppt);
    base = (long long)a.tab[2 * t] + (long long)z * (L * L);
    obase = (long long)a.tab[2 * t + 1] + (long long)z * (L * L);
  }
  const cpx<float>* __restrict__ in = reinterpret_cast<const cpx<float>*>(a.in) + base;
  cpx<float>* __restrict__ out = reinterpret_cast<cpx<float>*>(a.out) + obase;
  const int tid = threadIdx.x;
  cpx<float> rr[2][16];  // row results of row halves 0 and 1: point tau + TPL r of row c + H h

  // ---- rows, in two halves: thread (row c, tau); stage 0 reads from HBM
  const int rc = tid / GX::TPL, rtau = tid % GX::TPL;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    constexpr int RAD = sched_radix(L, 0), NB = 16 / RAD;
    cpx<float>* v = rr[h];
    const int row = rc + H * h;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = rtau + b * GX::TPL;
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        const int e = row * L + j + r * (L / RAD);
        cpx<float> x = in[e];
        if (WL) x = cscale(x, a.w_load[(unsigned long long)(base + e) & a.w_mask]);
        v[b * RAD + r] = INV ? cswap(x) : x;
      }
    }
    if (h) __syncthreads();  // the first half's last stage has read the buffer
    run_stages<float, L, MODE_CONTIG, 0, SPLIT, H>(v, lds, rc, rtau, reinterpret_cast<const cpx<float>*>(a.tw_x));
  }

  // ---- columns, in two halves: thread (column c, tau), adjacent threads on adjacent columns
  const int cc = tid % H, ctau = tid / H;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();  // the previous stage has read the buffer
    // points tau + TPL r with r in [8h, 8h + 8) are the columns of half h
#pragma unroll
    for (int rh = 0; rh < 2; ++rh)
#pragma unroll
      for (int r = 0; r < 8; ++r) lds[GY::idx(rtau + GX::TPL * r, rc + H * rh)] = rr[rh][8 * h + r];
    __syncthreads();
    cpx<float> v[16];
    stage_load_lds<float, L, MODE_STRIDED, 0, H>(v, lds, cc, ctau);
    __syncthreads();  // every thread holds its stage-0 inputs before stage 0 stores
    run_stages<float, L, MODE_STRIDED, 0, SPLIT, H>(v, lds, cc, ctau, reinterpret_cast<const cpx<float>*>(a.tw_y));
    constexpr int NPY = GY::NP;
    constexpr int RL = sched_radix(L, NPY - 1), NSL = sched_ns(L, NPY - 1), NBL = 16 / RL;
#pragma unroll
    for (int b = 0; b < NBL; ++b) {
      const int j = ctau + b * GY::TPL;
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int e = (j + r * NSL) * L + H * h + cc;
        cpx<float> x = v[b * RL + r];
        if (INV) x = cswap(x);
        if (WS) x = cscale(x, a.w_store[(unsigned long long)(obase + e) & a.w_mask]);
        out[e] = x;
      }
    }
  }
}

template <int L, bool INV, bool SPLIT>
static void launch_halves(const PlaneArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.nplanes), block(HalfGeom<L, SPLIT>::NT);
  if (a.w_load) {
    if (a.w_store) FFTAMD_LAUNCH((k_fft_plane_halves<L, INV, SPLIT, true, true>), grid, block, 0, s, a);
    else FFTAMD_LAUNCH((k_fft_plane_halves<L, INV, SPLIT, true, false>), grid, block, 0, s, a);
  } else {
    if (a.w_store) FFTAMD_LAUNCH((k_fft_plane_halves<L, INV, SPLIT, false, true>), grid, block, 0, s, a);
    else FFTAMD_LAUNCH((k_fft_plane_halves<L, INV, SPLIT, false, false>), grid, block, 0, s, a);
  }
}

bool plane_shape_supported(long long ly, long long lx) {
  auto p2 = [](long long v) { return v >= 16 && v <= 128 && (v & (v - 1)) == 0; };
  return p2(ly) && p2(lx) && ly * lx >= 1024;
}

template <int LY, int LX, bool SPLIT>
static void launch_shape(const PlaneArgs& a, hipStream_t s) {
  constexpr int NT = PlaneGeom<LY, LX, SPLIT>::NT;
  // two workgroups per CU (FP32 stages; split-precision stages would spill
  // registers at 512 threads and keep the whole-plane kernel)
  if constexpr (LY == 128 && LX == 128 && !SPLIT) {
    if (a.inverse) launch_halves<128, true, SPLIT>(a, s);
    else launch_halves<128, false, SPLIT>(a, s);
  } else {
    if (a.inverse)
      FFTAMD_LAUNCH((k_fft_plane<LY, LX, true, SPLIT>), dim3((unsigned)a.nplanes), dim3(NT), 0, s, a);
    else
      FFTAMD_LAUNCH((k_fft_plane<LY, LX, false, SPLIT>), dim3((unsigned)a.nplanes), dim3(NT), 0, s, a);
  }
}

template <int LY, bool SPLIT>
static bool dispatch_lx(int lx, const PlaneArgs& a, hipStream_t s) {
  switch (lx) {
    case 16: if constexpr (LY * 16 >= 1024) { launch_shape<LY, 16, SPLIT>(a, s); return true; } else return false;
    case 32: if constexpr (LY * 32 >= 1024) { launch_shape<LY, 32, SPLIT>(a, s); return true; } else return false;
    case 64: launch_shape<LY, 64, SPLIT>(a, s); return true;
    case 128: launch_shape<LY, 128, SPLIT>(a, s); return true;
    default: return false;
  }
}

template <bool SPLIT>
static bool dispatch_ly(int ly, int lx, const PlaneArgs& a, hipStream_t s) {
  switch (ly) {
    case 16: return dispatch_lx<16, SPLIT>(lx, a, s);
    case 32: return dispatch_lx<32, SPLIT>(lx, a, s);
    case 64: return dispatch_lx<64, SPLIT>(lx, a, s);
    case 128: return dispatch_lx<128, SPLIT>(lx, a, s);
    default: return false;
  }
}

void preload_plane() { preload_kernel(reinterpret_cast<const void*>(&k_fft_plane<64, 64, false, false>)); }

int launch_fft_plane(int ly, int lx, const PlaneArgs& a, void* stream, bool split) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!plane_shape_supported(ly, lx) || a.nplanes == 0 || a.nplanes > 0xFFFFFFFFull / 1024 ||
      (a.tab && a.ppt == 0))
    return -1;
  const bool ok = split ? dispatch_ly<true>(ly, lx, a, s) : dispatch_ly<false>(ly, lx, a, s);
  if (!ok) return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fftamd

// fft_plane_dev.h -- the work of one plane of the plane passes (the method is
// described in fft_plane.hip) as a device function, so that the
// one-plane-per-workgroup kernel and the persistent multi-transform kernel
// (fft_flow.hip) run the same code.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_pass_dev.h"
#include "fft_plane.h"

namespace fftamd {

template <int LY, int LX, bool SPLIT>
struct PlaneGeom {
  static_assert(LY >= 16 && LX >= 16 && LY <= 128 && LX <= 128, "plane sides 16..128");
  using GX = Geom<LX, MODE_CONTIG, LY>;    // rows: LY lines of LX points
  using GY = Geom<LY, MODE_STRIDED, LX>;   // columns: LX lines of LY points
  static constexpr int NT = LY * LX / 16;
  static_assert(GX::NT == NT && GY::NT == NT, "16 points per thread in both phases");
  static constexpr int LDS = GX::LDS_ELEMS > GY::LDS_ELEMS ? GX::LDS_ELEMS : GY::LDS_ELEMS;
};

// One plane (index `plane` of the array, or of the transform table) by
// threads tid = 0 .. NT - 1 with PlaneGeom::LDS elements of LDS at lds.
// SPLIT: radix-16 stages on FP16 matrix cores with hi+lo operands
// (fft_split.h; a tile is then made of whole waves).
template <int LY, int LX, bool INV, bool SPLIT>
__device__ __forceinline__ void plane_tile(const PlaneArgs& a, unsigned long long plane, cpx<float>* lds,
                                           int tid) {
  using PG = PlaneGeom<LY, LX, SPLIT>;
  using GX = typename PG::GX;
  using GY = typename PG::GY;
  long long base = (long long)plane * a.plane_stride, obase = base;
  if (a.tab) {  // planes of a batch of transforms placed anywhere
    const unsigned t = (unsigned)(plane / a.ppt), z = (unsigned)(plane % a.ppt);
    base = (long long)a.tab[2 * t] + (long long)z * (LY * LX);
    obase = (long long)a.tab[2 * t + 1] + (long long)z * (LY * LX);
  }
  const cpx<float>* __restrict__ in = reinterpret_cast<const cpx<float>*>(a.in) + base;
  cpx<float>* __restrict__ out = reinterpret_cast<cpx<float>*>(a.out) + obase;
  cpx<float> v[16];

  // ---- rows: thread (row c, tau); stage 0 reads its butterfly inputs from HBM
  {
    constexpr int RAD = sched_radix(LX, 0), NB = 16 / RAD;
    const int c = tid / GX::TPL, tau = tid % GX::TPL;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = tau + b * GX::TPL;
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        const int e = c * LX + j + r * (LX / RAD);
        cpx<float> x = in[e];
        if (a.w_load) x = cscale(x, a.w_load[(unsigned long long)(base + e) & a.w_mask]);
        v[b * RAD + r] = INV ? cswap(x) : x;
      }
    }
    run_stages<float, LX, MODE_CONTIG, 0, SPLIT, LY>(v, lds, c, tau,
                                                     reinterpret_cast<const cpx<float>*>(a.tw_x));
    // last stage: point k = j + r*NS of row c -> column layout
    constexpr int NPX = GX::NP;
    constexpr int RL = sched_radix(LX, NPX - 1), NSL = sched_ns(LX, NPX - 1), NBL = 16 / RL;
    __syncthreads();  // every row stage has read its LDS inputs
#pragma unroll
    for (int b = 0; b < NBL; ++b) {
      const int j = tau + b * GX::TPL;
#pragma unroll
      for (int r = 0; r < RL; ++r) lds[GY::idx(j + r * NSL, c)] = v[b * RL + r];
    }
  }
  __syncthreads();

  // ---- columns: thread (column c, tau), adjacent threads on adjacent columns
  {
    const int c = tid % LX, tau = tid / LX;
    stage_load_lds<float, LY, MODE_STRIDED, 0, LX>(v, lds, c, tau);
    __syncthreads();  // every thread holds its stage-0 inputs before stage 0 stores
    run_stages<float, LY, MODE_STRIDED, 0, SPLIT, LX>(v, lds, c, tau,
                                                      reinterpret_cast<const cpx<float>*>(a.tw_y));
    constexpr int NPY = GY::NP;
    constexpr int RL = sched_radix(LY, NPY - 1), NSL = sched_ns(LY, NPY - 1), NBL = 16 / RL;
#pragma unroll
    for (int b = 0; b < NBL; ++b) {
      const int j = tau + b * GY::TPL;
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int e = (j + r * NSL) * LX + c;
        cpx<float> x = v[b * RL + r];
        if (INV) x = cswap(x);
        if (a.w_store) x = cscale(x, a.w_store[(unsigned long long)(obase + e) & a.w_mask]);
        out[e] = x;
      }
    }
  }
}

}  // namespace fftamd

// fft_pass_dev.h -- device code of the Stockham line-FFT passes: geometry,
// radix stages, and pass_block(), the work of one workgroup tile.
//
// Addressing of point p of line (o, k, col) of a pass:
//   in  + o*in_o  + k*in_k  + col*in_c  + p*in_p
//   out + o*out_o + kk*out_k + col*out_c + p*out_p   (kk = bitrev(k) if kbits)
// MODE_CONTIG: in_p == 1, threads of a line are adjacent (coalesced along p).
// MODE_STRIDED: in_c == 1, threads of a row are adjacent (coalesced along col).
//
// A tile loads its C lines straight from HBM into VGPRs (stage 0 reads the
// first stage's butterfly inputs directly), exchanges through LDS between
// radix stages, and the last stage writes straight to HBM in natural order.
// LDS rows are padded so the strided exchanges avoid bank conflicts.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_device.h"
#include "fft_pass.h"
#include "fft_schedule.h"
#include "fft_split.h"

namespace fftamd {

// CC: lines per workgroup (the pass kernels' default; the plane kernels hold
// a whole plane's rows or columns)
template <int L, int MODE, int CC = geom_cols(L, MODE)>
struct Geom {
  static constexpr int PPT = geom_ppt(L);
  static constexpr int TPL = geom_tpl(L);
  static constexpr int C = CC;
  static constexpr int NT = C * TPL;
  static constexpr int NP = sched_np(L);
  // contiguous: [line][p + p/16]; strided: [p][col] with one pad column
  static constexpr int LP = L + L / 16;
  static constexpr int LDS_ELEMS = MODE == MODE_CONTIG ? C * LP : L * (C + 1);
  static __device__ __forceinline__ int idx(int c, int p) {
    if constexpr (MODE == MODE_CONTIG) return c * LP + p + (p >> 4);
    else return p * (C + 1) + c;
  }
};

// TAB: split stages read their constant operands from the LDS table `atab`
template <typename R, int L, int MODE, int S, bool SPLIT = false, int CC = geom_cols(L, MODE), bool TAB = false>
__device__ __forceinline__ void stage_compute(cpx<R>* v, int tau, const cpx<R>* __restrict__ tw,
                                              cpx<R>* lds = nullptr, const split_half8* atab = nullptr) {
  using G = Geom<L, MODE, CC>;
  constexpr int RAD = sched_radix(L, S);
  constexpr int NS = sched_ns(L, S);
  constexpr int NB = G::PPT / RAD;
  if constexpr (S > 0) {
    static_assert(RAD == 16, "stages after the first are radix 16");
    constexpr int OFF = sched_twoff(L, S);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = tau + b * G::TPL;
      const int k = j & (NS - 1);
      // W^k, W^2k, W^4k, W^8k from the table; W^{kr} for the other r by products
      const cpx<R>* t = tw + OFF + TW_PER_K * k;
      const cpx<R> w1 = t[0], w2 = t[1], w4 = t[2], w8 = t[3];
      const cpx<R> w3 = cmul(w1, w2), w12 = cmul(w4, w8);
      cpx<R> w[16];
      w[1] = w1;
      w[2] = w2;
      w[3] = w3;
      w[4] = w4;
      w[5] = cmul(w4, w1);
      w[6] = cmul(w4, w2);
      w[7] = cmul(w4, w3);
      w[8] = w8;
      w[9] = cmul(w8, w1);
      w[10] = cmul(w8, w2);
      w[11] = cmul(w8, w3);
      w[12] = w12;
      w[13] = cmul(w12, w1);
      w[14] = cmul(w12, w2);
      w[15] = cmul(w12, w3);
#pragma unroll
      for (int r = 1; r < 16; ++r) v[b * RAD + r] = cmul(v[b * RAD + r], w[r]);
    }
  }
  if constexpr (SPLIT && RAD == 16 && sizeof(R) == 4) {
    static_assert(NB == 1, "split radix-16 stages hold one butterfly per lane");
    (void)lds;
    dft16_split_perm<TAB>(reinterpret_cast<cpx<float>*>(v), atab);
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b) dft<R, RAD>(v + b * RAD);
  }
}

// write stage-S results to LDS at their Stockham destinations
template <typename R, int L, int MODE, int S, int CC = geom_cols(L, MODE)>
__device__ __forceinline__ void stage_store_lds(const cpx<R>* v, cpx<R>* lds, int c, int tau) {
  using G = Geom<L, MODE, CC>;
  constexpr int RAD = sched_radix(L, S);
  constexpr int NS = sched_ns(L, S);
  constexpr int NB = G::PPT / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = tau + b * G::TPL;
    const int base = (j / NS) * NS * RAD + (j & (NS - 1));
#pragma unroll
    for (int r = 0; r < RAD; ++r) lds[G::idx(c, base + r * NS)] = v[b * RAD + r];
  }
}

// read stage-S butterfly inputs from LDS
template <typename R, int L, int MODE, int S, int CC = geom_cols(L, MODE)>
__device__ __forceinline__ void stage_load_lds(cpx<R>* v, const cpx<R>* lds, int c, int tau) {
  using G = Geom<L, MODE, CC>;
  constexpr int RAD = sched_radix(L, S);
  constexpr int NB = G::PPT / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = tau + b * G::TPL;
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[b * RAD + r] = lds[G::idx(c, j + r * (L / RAD))];
  }
}

template <typename R, int L, int MODE, int S, bool SPLIT = false, int CC = geom_cols(L, MODE)>
__device__ __forceinline__ void run_stages(cpx<R>* v, cpx<R>* lds, int c, int tau,
                                           const cpx<R>* __restrict__ tw) {
  using G = Geom<L, MODE, CC>;
  if constexpr (S < G::NP) {
    if constexpr (S > 0) {
      __syncthreads();  // previous stage's LDS writes are visible
      stage_load_lds<R, L, MODE, S, CC>(v, lds, c, tau);
    }
    stage_compute<R, L, MODE, S, SPLIT, CC>(v, tau, tw, lds);
    if constexpr (S + 1 < G::NP) {
      if constexpr (S > 0) __syncthreads();  // every thread finished reading this stage
      stage_store_lds<R, L, MODE, S, CC>(v, lds, c, tau);
      run_stages<R, L, MODE, S + 1, SPLIT, CC>(v, lds, c, tau, tw);
    }
  }
}

// Where this thread's line of tile g sits (tile = column block, middle and
// outer line index decoded from g).
struct TileLoc {
  int c, tau;
  unsigned col;
  bool active;
  long long in_off, out_off;
};

template <int L, int MODE, int CC = geom_cols(L, MODE)>
__device__ __forceinline__ TileLoc tile_locate(const PassArgs& a, unsigned long long g, int tid) {
  using G = Geom<L, MODE, CC>;
  TileLoc t;
  t.c = MODE == MODE_CONTIG ? tid / G::TPL : tid % G::C;
  t.tau = MODE == MODE_CONTIG ? tid % G::TPL : tid / G::C;
  const unsigned cb = (unsigned)(g % a.ncb);
  g /= a.ncb;
  const unsigned kk = (unsigned)(g % a.K);
  const unsigned long long o = g / a.K + a.outer0;
  t.col = cb * G::C + t.c;
  t.active = t.col < a.ncol;
  const unsigned ko = a.kbits ? brev_bits(kk, a.kbits) : kk;
  const long long ib = a.tab ? (long long)a.tab[2 * o] : (long long)o * a.in_o;
  const long long ob = a.tab ? (long long)a.tab[2 * o + 1] : (long long)o * a.out_o;
  t.in_off = ib + (long long)kk * a.in_k + (long long)t.col * a.in_c;
  t.out_off = ob + (long long)ko * a.out_k + (long long)t.col * a.out_c;
  return t;
}

// Raw stage-0 inputs of this thread straight from HBM (no arithmetic, so a
// pipelined kernel can issue them one tile ahead).
template <typename R, int L, int MODE, int CC = geom_cols(L, MODE)>
__device__ __forceinline__ void tile_load(const PassArgs& a, const TileLoc& t, cpx<R>* v) {
  using G = Geom<L, MODE, CC>;
  constexpr int RAD = sched_radix(L, 0);
  constexpr int NB = G::PPT / RAD;
  const cpx<R>* __restrict__ in = reinterpret_cast<const cpx<R>*>(a.in) + t.in_off;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t.tau + b * G::TPL;
#pragma unroll
    for (int r = 0; r < RAD; ++r) {
      const long long p = j + r * (L / RAD);
      cpx<R> x = {R(0), R(0)};
      if (t.active) x = in[p * a.in_p];
      v[b * RAD + r] = x;
    }
  }
}

// Input weights and direction, radix stages and the epilogue of one tile
// whose raw inputs are in v.  DIR: 0 forward, 1 inverse (the swaps of
// swap(DFT(swap(x))) are register renames), -1 taken from a.inverse at run
// time (per-element selects).
template <typename R, int L, int MODE, int EPI, bool SPLIT = false, int DIR = -1, int CC = geom_cols(L, MODE)>
__device__ __forceinline__ void tile_finish(const PassArgs& a, const TileLoc& t, cpx<R>* v, cpx<R>* lds) {
  const bool inv = DIR < 0 ? a.inverse != 0 : DIR == 1;
  using G = Geom<L, MODE, CC>;
  constexpr int PPT = G::PPT, TPL = G::TPL;
  constexpr int NP = G::NP;
  const int c = t.c, tau = t.tau;
  const unsigned col = t.col;
  const bool active = t.active;
  const long long in_off = t.in_off, out_off = t.out_off;
  cpx<R>* __restrict__ out = reinterpret_cast<cpx<R>*>(a.out) + out_off;
  const cpx<R>* __restrict__ tw = reinterpret_cast<const cpx<R>*>(a.tw);
  {
    constexpr int RAD = sched_radix(L, 0);
    constexpr int NB = PPT / RAD;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = tau + b * TPL;
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        cpx<R> x = v[b * RAD + r];
        if (a.w_load && active) {
          const long long p = j + r * (L / RAD);
          x = cscale(x, (R)a.w_load[(unsigned long long)(in_off + p * a.in_p) & a.w_mask]);
        }
        v[b * RAD + r] = inv ? cswap(x) : x;
      }
    }
  }

  run_stages<R, L, MODE, 0, SPLIT, CC>(v, lds, c, tau, tw);

  // ---- last stage: point k = j + r*NS of the line is in v[b*RAD + r]
  constexpr int RL = sched_radix(L, NP - 1);
  constexpr int NSL = sched_ns(L, NP - 1);
  constexpr int NBL = PPT / RL;
  if constexpr (EPI == EPI_BITREV) {
    // Point k = j + r*NSL (r < RL) lands at bitrev(k) = bitrev(r) + RL*bitrev(j):
    // each butterfly's RL outputs form one contiguous run of the output line.
    constexpr unsigned LGR = ilog2c(RL), LGJ = ilog2c(L) - ilog2c(RL);
    if constexpr (NP > 1) {
      // through LDS: the runs are placed at their permuted positions, then the
      // line is written in order, each store instruction one contiguous span
      // (stores straight from registers would give every lane its own line)
      __syncthreads();  // the last stage has read its inputs
#pragma unroll
      for (int b = 0; b < NBL; ++b) {
        const int j = tau + b * TPL;
        const int run = RL * (int)brev_bits((unsigned)j, LGJ);
#pragma unroll
        for (int r = 0; r < RL; ++r) lds[G::idx(c, run + (int)brev_bits((unsigned)r, LGR))] = v[b * RL + r];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < PPT; ++i) {
        const long long q = tau + i * TPL;
        cpx<R> x = lds[G::idx(c, (int)q)];
        if (inv) x = cswap(x);
        if (a.w_store) x = cscale(x, (R)a.w_store[(unsigned long long)(out_off + q * a.out_p) & a.w_mask]);
        if (active) out[q * a.out_p] = x;
      }
    } else {
#pragma unroll
      for (int b = 0; b < NBL; ++b) {
        const int j = tau + b * TPL;
        const long long run = (long long)RL * brev_bits((unsigned)j, LGJ);
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const long long q = run + brev_bits((unsigned)r, LGR);
          cpx<R> x = v[b * RL + r];
          if (inv) x = cswap(x);
          if (a.w_store) x = cscale(x, (R)a.w_store[(unsigned long long)(out_off + q * a.out_p) & a.w_mask]);
          if (active) out[q * a.out_p] = x;
        }
      }
    }
  } else {
#pragma unroll
    for (int b = 0; b < NBL; ++b) {
      const int j = tau + b * TPL;
      // four-step post-twiddle W_M^{cc k}, cc = sub-axis index of this column,
      // k = j + r*NSL: W^{cc j} times the table powers W^{cc NSL m}, m = 1, 2,
      // 4, 8 (each from the two-level table: two loads and one product)
      cpx<R> base{}, s1{}, s2{}, s4{}, s8{};
      if constexpr (EPI == EPI_TWIDDLE) {
        const unsigned cc = col >> a.ptw_cshift;
        const cpx<R>* lo = reinterpret_cast<const cpx<R>*>(a.ptw_lo);
        const cpx<R>* hi = reinterpret_cast<const cpx<R>*>(a.ptw_hi);
        const unsigned lomask = (1u << a.ptw_hb) - 1u;
        auto w2 = [&](unsigned t) {
          t &= a.ptw_mmask;
          return cmul(hi[t >> a.ptw_hb], lo[t & lomask]);
        };
        base = w2(cc * (unsigned)j);
        const unsigned st = cc * (unsigned)NSL;
        if constexpr (RL > 1) s1 = w2(st);
        if constexpr (RL > 2) s2 = w2(2u * st);
        if constexpr (RL > 4) s4 = w2(4u * st);
        if constexpr (RL > 8) s8 = w2(8u * st);
      }
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int k = j + r * NSL;
        cpx<R> x = v[b * RL + r];
        if constexpr (EPI == EPI_TWIDDLE) {
          cpx<R> w = base;
          if (r & 1) w = cmul(w, s1);
          if (r & 2) w = cmul(w, s2);
          if (r & 4) w = cmul(w, s4);
          if (r & 8) w = cmul(w, s8);
          x = cmul(x, w);
        }
        if (inv) x = cswap(x);
        if (active) {
          if (a.w_store) x = cscale(x, (R)a.w_store[(unsigned long long)(out_off + (long long)k * a.out_p) & a.w_mask]);
          out[(long long)k * a.out_p] = x;
        }
      }
    }
  }
}

// One workgroup tile of a pass (block index g), run by threads tid = 0 ..
// Geom::NT - 1.  Used by k_fft_pass (one tile per workgroup) and by the
// persistent multi-transform kernels (which may run two tiles side by side in
// one workgroup, each half with its own tid range and LDS, the halves made of
// whole waves).
template <typename R, int L, int MODE, int EPI, bool SPLIT = false, int DIR = -1,
          int CC = geom_cols(L, MODE)>
__device__ __forceinline__ void pass_block(const PassArgs& a, unsigned long long g, cpx<R>* lds,
                                           int tid) {
  cpx<R> v[Geom<L, MODE, CC>::PPT];
  const TileLoc t = tile_locate<L, MODE, CC>(a, g, tid);
  tile_load<R, L, MODE, CC>(a, t, v);
  tile_finish<R, L, MODE, EPI, SPLIT, DIR, CC>(a, t, v, lds);
}

}  // namespace fftamd

// fft_cube128_dev.h -- the two passes of the y-split pair for 128^3 blocks
// as device functions (one tile each), so that the pair's kernels
// (fft_cube.hip) and the persistent multi-transform kernel (fft_flow.hip)
// run the same code.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_pass_dev.h"
#include "fft_plane.h"
#include "fft_split.h"

namespace fftamd {

// ---- 128 x 128 x 128 blocks: the same split, y = a + 4b (b < 32), ky = k + 32d.
// Pass A holds 32 rows of 128 (32 KiB) in LDS with 256 threads of 16 points
// (four workgroups per CU); pass B 32 columns x 4 rows x 128 planes (128 KiB)
// with 512 threads of 32 points (two per CU).
constexpr int C1_L = 128;
constexpr int C1_NT = 256;
constexpr long long C1_PLANE = (long long)C1_L * C1_L;
constexpr long long C1_BLOCK = C1_PLANE * C1_L;
using C1Row = Geom<C1_L, MODE_CONTIG, 32>;   // pass A rows: 32 rows of 128, 8 threads each
using C1Col = Geom<32, MODE_STRIDED, 128>;   // pass A columns: 128 columns of 32, 2 threads each
constexpr int C1B_NT = 512;
using C1Z = Geom<C1_L, MODE_STRIDED, 64>;    // pass B: 64 lines (2 rows x 32 columns) of 128
static_assert(C1Row::NT == C1_NT && C1Col::NT == C1_NT && C1Z::NT == C1B_NT, "256 / 512 threads");
static_assert(sched_np(C1_L) == 2 && sched_np(32) == 2, "two stages per line");
constexpr int C1A_LDS = C1Row::LDS_ELEMS > C1Col::LDS_ELEMS ? C1Row::LDS_ELEMS : C1Col::LDS_ELEMS;

constexpr int C1A_BUF = C1A_LDS;       // pass A tile's buffer (FP32), then 8 x 15 twiddles
constexpr int C1A_TWS = 8 * 15;

// Pass A tile `id` (plane id / 4, residue class id % 4) by threads tid < 256:
// buf holds C1A_LDS elements, tws 120.
template <bool INV, bool WL, bool SPLIT>
__device__ __forceinline__ void cube128_a_tile(const CubeArgs& a, unsigned long long id, cpx<float>* buf,
                                               cpx<float>* tws, int tid) {
  // tws: row stage-1 twiddles W128^{k r} at [15 k + r - 1]
  const unsigned long long plane = id >> 2;
  const int res = (int)(id & 3);  // residue class a of the input rows
  if (plane >= a.nplanes) return;  // uniform over the workgroup
  const long long base = (long long)plane * C1_PLANE;  // also the weights' element index
  long long ibase = base, obase = base;
  if (a.tab) {  // blocks of a batch of transforms placed anywhere
    const unsigned long long t = plane / C1_L;
    const long long zo = (long long)(plane % C1_L) * C1_PLANE;
    ibase = (long long)a.tab[2 * t] + zo;
    obase = (long long)a.tab[2 * t + 1] + zo;
  }
  const cpx<float>* in = reinterpret_cast<const cpx<float>*>(a.in) + ibase;
  cpx<float>* out = reinterpret_cast<cpx<float>*>(a.out) + obase;
  const cpx<float>* __restrict__ wdir = reinterpret_cast<const cpx<float>*>(a.w256);
  if (tid < 8 * 15) tws[tid] = wdir[2 * (tid / 15) * (tid % 15 + 1)];  // read after the first barrier

  // ---- x: thread (c, tau) transforms row b = c (y = res + 4c); its stage-0
  // inputs come straight from HBM; it ends with point kx = tau + 8 r
  const int c = tid >> 3, tau = tid & 7;
  cpx<float> v[16];
  {
    constexpr int RAD = sched_radix(C1_L, 0), NB = 16 / RAD;
    const long long row = (long long)(res + 4 * c) * C1_L;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb)
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        const long long e = row + tau + C1Row::TPL * bb + r * (C1_L / RAD);
        cpx<float> s = in[e];
        if (WL) s = cscale(s, a.w_load[(unsigned long long)(base + e) & a.w_mask]);
        v[bb * RAD + r] = INV ? cswap(s) : s;
      }
  }
  static_assert(sched_np(C1_L) == 2 && sched_ns(C1_L, 1) == C1Row::TPL, "radix 8, then radix 16 ending with point tau + 8 r");
  stage_compute<float, C1_L, MODE_CONTIG, 0, false, 32>(v, tau, nullptr);  // radix 8, no twiddles
  stage_store_lds<float, C1_L, MODE_CONTIG, 0, 32>(v, buf, c, tau);
  __syncthreads();
  stage_load_lds<float, C1_L, MODE_CONTIG, 1, 32>(v, buf, c, tau);
  // radix-16 stage over sub-transforms of span 8: twiddles W128^{tau r} from
  // the LDS table (lanes of one tau read one address)
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], tws[15 * tau + r - 1]);
  wg_dft16<SPLIT>(v);

  // ---- y over b: 128 columns of 32 points; thread (cc, tq) ends with point
  // k = tq + 2 r of its column
  __syncthreads();  // every thread has read its last row-stage inputs
#pragma unroll
  for (int r = 0; r < 16; ++r) buf[C1Col::idx(tau + 8 * r, c)] = v[r];
  __syncthreads();
  const int cc = tid & 127;
  const int tq = __builtin_amdgcn_readfirstlane(tid >> 7);  // uniform over each wave
  stage_load_lds<float, 32, MODE_STRIDED, 0, 128>(v, buf, cc, tq);
  __syncthreads();  // every thread holds its stage-0 inputs before stage 0 stores
  stage_compute<float, 32, MODE_STRIDED, 0, false, 128>(v, tq, nullptr);  // eight radix-2, no twiddles
  stage_store_lds<float, 32, MODE_STRIDED, 0, 128>(v, buf, cc, tq);
  __syncthreads();
  stage_load_lds<float, 32, MODE_STRIDED, 1, 128>(v, buf, cc, tq);
  // radix-16 stage over sub-transforms of span 2: twiddles W32^{tq r} = W256^{8 tq r}
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], wdir[8 * tq * r]);
  wg_dft16<SPLIT>(v);
  // column DFT point k = tq + 2 r to row 32 res + k (pass B applies W128^{res k})
#pragma unroll
  for (int r = 0; r < 16; ++r) out[(long long)(32 * res + tq + 2 * r) * C1_L + cc] = v[r];
}

// Pass B of 128^3 blocks: 32 adjacent columns (256-byte row segments, as the
// 128-point z line pass; 16 columns measured 102.6 against 98.9 us) x 4 rows
// x 128 planes (128 KiB) with 512 threads of 32 points, two workgroups per CU.
// Pass B tile b by threads tid < 512: buf holds C1Z::LDS_ELEMS elements;
// split: atab holds the DFT operand table, filled before the tile's first
// barrier.
template <bool INV, bool WS, bool SPLIT>
__device__ __forceinline__ void cube128_b_tile(const CubeArgs& a, unsigned b, cpx<float>* buf,
                                               const split_half8* atab, int tid) {
  const int xg = b & 3;                // group of 32 columns
  const int k = (b >> 2) & 31;         // rows k + 32 a
  const unsigned long long blk = b >> 7;
  if (blk >= a.nplanes / C1_L) return;  // uniform over the workgroup
  const long long tile = (long long)k * C1_L + 32 * xg;
  const long long base = (long long)blk * C1_BLOCK + tile;  // also the weights' element index
  const long long ibase = a.tab ? (long long)a.tab[2 * blk] + tile : base;
  const long long obase = a.tab ? (long long)a.tab[2 * blk + 1] + tile : base;
  const cpx<float>* in = reinterpret_cast<const cpx<float>*>(a.in) + ibase;
  cpx<float>* out = reinterpret_cast<cpx<float>*>(a.out) + obase;
  const cpx<float>* __restrict__ wdir = reinterpret_cast<const cpx<float>*>(a.w256);

  // ---- y over a: thread (x, zs) holds planes z = zs + 16 j; t[j][d] = X[k + 32 d]
  const int x = tid & 31, zs = tid >> 5;
  cpx<float> t[8][4];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const long long zo = (long long)(zs + 16 * j) * C1_PLANE + x;
#pragma unroll
    for (int aa = 0; aa < 4; ++aa) t[j][aa] = in[zo + 32 * C1_L * aa];
  }
  // T_a[k] = W128^{a k} = W256^{2 a k} times pass A's column DFT (uniform)
  const cpx<float> w1 = wdir[2 * k], w2 = wdir[4 * k], w3 = wdir[6 * k];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    t[j][1] = cmul(t[j][1], w1);
    t[j][2] = cmul(t[j][2], w2);
    t[j][3] = cmul(t[j][3], w3);
    dft4(t[j]);
  }
  // ---- z: the 64 lines of rows d = 2 e, 2 e + 1 at a time; line cz = 32 (d - 2 e) + x.
  // A thread's planes zs + 16 j are the first z stage's radix-8 inputs
  // (points j0 + 16 r, j0 = zs): that stage runs in registers and its outputs
  // go to LDS at their Stockham positions 8 zs + r, so the z transform needs
  // one LDS exchange
  static_assert(sched_radix(C1_L, 0) == 8 && sched_ns(C1_L, 1) == 8, "radix 8, then radix 16");
  const int cz = tid & 63;
  const int tz = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform over each wave
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    if (e) __syncthreads();  // the first half's last stage has read the buffer
#pragma unroll
    for (int dd = 0; dd < 2; ++dd) {
      cpx<float> p8[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) p8[j] = t[j][2 * e + dd];
      dft8(p8);
#pragma unroll
      for (int r = 0; r < 8; ++r) buf[C1Z::idx(32 * dd + x, 8 * zs + r)] = p8[r];
    }
    __syncthreads();
    cpx<float> v[16];
    stage_load_lds<float, C1_L, MODE_STRIDED, 1, 64>(v, buf, cz, tz);
    // radix-16 stage over sub-transforms of span 8: twiddles W128^{tz r} =
    // W256^{2 tz r}, uniform over the wave (scalar loads)
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], wdir[2 * tz * r]);
    wg_dft16<SPLIT, true>(v, atab);
    // point kz = tz + 8 r of row k + 32 d, column 32 xg + (cz & 31)
    const long long ob = (long long)(32 * (2 * e + (cz >> 5))) * C1_L + (cz & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long long o = ob + (long long)(tz + 8 * r) * C1_PLANE;
      cpx<float> s = INV ? cswap(v[r]) : v[r];
      if (WS) s = cscale(s, a.w_store[(unsigned long long)(base + o) & a.w_mask]);
      __builtin_nontemporal_store(__builtin_bit_cast(long long, s), reinterpret_cast<long long*>(out + o));
    }
  }
}

}  // namespace fftamd

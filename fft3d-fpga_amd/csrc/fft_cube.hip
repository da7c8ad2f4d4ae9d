// fft_cube.hip -- n x n x n blocks (n = 256 or 128) in two passes over HBM
// with the y axis split (n/4) x 4 between them, so that no tile is larger than
// one workgroup's registers and LDS and no input is read twice.
//
// y = a + 4 b (a < 4, b < n/4), ky = k + (n/4) d (k < n/4, d < 4):
//   X[k + (n/4) d] = sum_a W4^{a d} T_a[k],
//   T_a[k]         = Wn^{a k} sum_b x[a + 4 b] W(n/4)^{b k}.
//
// Pass A (one workgroup per plane and residue a): reads the n/4 rows
// y = a + 4 b of a plane (whole rows), transforms each along x, then each
// column along b, and writes the column DFT (T_a[k] without its twiddle) to
// row (n/4) a + k.  Out of place: the rows it writes are read by the other
// three residues.
//
// Pass B (one workgroup per (block, k, group of adjacent columns)): reads rows
// k + (n/4) a, a = 0..3, of its columns in all n planes, multiplies row
// k + (n/4) a by Wn^{a k}, forms X[k + (n/4) d] with a 4-point DFT over a,
// transforms the resulting lines along z and writes them back to the same
// rows and columns (in place allowed).  A thread loads all four rows of one
// column in 8 planes, so the 4-point DFT is in registers, and adjacent lanes
// read or write whole 128- or 256-byte row segments.
//
// n = 256: pass A's tile is 64 rows (128 KiB), pass B's 16 columns x 4 rows x
// 256 planes (128 KiB); 512 threads hold 32 points each in registers and LDS
// stages half a tile at a time (68 KiB), so two workgroups share a CU.  Against
// the 256 x 256 plane pass, whose four workgroups per plane each read the whole
// 512 KiB plane, nothing is read twice.  n = 128: see the second pair below.
// Inverse: swap(DFT(swap(x))), the swaps applied at pass A's loads and pass
// B's stores.
#include <hip/hip_runtime.h>

#include "launch.h"

#include "fft_pass_dev.h"
#include "fft_plane.h"
#include "fft_cube128_dev.h"

namespace fftamd {

constexpr int CB_L = 256;
constexpr int CB_NT = 512;
constexpr long long CB_PLANE = (long long)CB_L * CB_L;   // elements per plane
constexpr long long CB_BLOCK = CB_PLANE * CB_L;          // elements per 256^3 block

using CaRow = Geom<CB_L, MODE_CONTIG, 32>;     // pass A rows: 32 rows of 256, 16 threads each
using CaCol = Geom<64, MODE_STRIDED, 128>;     // pass A columns: 128 columns of 64, 4 threads each
using CbZ = Geom<CB_L, MODE_STRIDED, 32>;      // pass B: 32 lines (2 rows x 16 columns) of 256
static_assert(CaRow::NT == CB_NT && CaCol::NT == CB_NT && CbZ::NT == CB_NT, "512 threads");
static_assert(sched_np(CB_L) == 2 && sched_np(64) == 2, "two stages per line");
constexpr int CA_LDS = CaRow::LDS_ELEMS > CaCol::LDS_ELEMS ? CaRow::LDS_ELEMS : CaCol::LDS_ELEMS;

// Two workgroups per CU: at most 128 VGPRs (4 waves per SIMD), <= 74 KiB LDS each.
template <bool INV, bool WL, bool SPLIT>
__global__ void __launch_bounds__(CB_NT, 4) k_fft_cube_a(const CubeArgs a) {
  __shared__ cpx<float> buf[CA_LDS];
  __shared__ cpx<float> tws[16 * 15];  // row stage-1 twiddles W256^{k r} at [15 k + r - 1]
  __shared__ split_half8 atab[SPLIT ? SPLIT_ATAB : 1];  // split DFT operands (read after a barrier)
  const unsigned long long plane = blockIdx.x >> 2;
  const int res = blockIdx.x & 3;  // residue class a of the input rows
  if (plane >= a.nplanes) return;  // uniform over the workgroup
  const long long base = (long long)plane * CB_PLANE;
  const cpx<float>* in = reinterpret_cast<const cpx<float>*>(a.in) + base;
  cpx<float>* out = reinterpret_cast<cpx<float>*>(a.out) + base;
  const cpx<float>* __restrict__ wdir = reinterpret_cast<const cpx<float>*>(a.w256);
  const int tid = threadIdx.x;
  if (tid < 16 * 15) tws[tid] = wdir[(tid / 15) * (tid % 15 + 1)];  // read after the first barrier
  if constexpr (SPLIT) split_atab_fill(atab);

  // ---- x: rows b = 32 h + c (y = res + 4 b); thread (c, tau) ends with
  // point kx = tau + 16 r of its row in r0[r] (h = 0) and r1[r] (h = 1)
  const int c = tid >> 4, tau = tid & 15;
  cpx<float> r0[16], r1[16];
  // both halves' loads are issued first: the second half arrives while the
  // first is transformed (r1 holds it either way, so this costs no registers)
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    cpx<float>* v = h ? r1 : r0;
    const long long row = (long long)(res + 4 * (32 * h + c)) * CB_L;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long long e = row + tau + 16 * r;
      cpx<float> s = in[e];
      if (WL) s = cscale(s, a.w_load[(unsigned long long)(base + e) & a.w_mask]);
      v[r] = INV ? cswap(s) : s;
    }
  }
  __builtin_amdgcn_sched_barrier(0);  // keep the second half's loads ahead of the first half's arithmetic
  if constexpr (SPLIT) __syncthreads();  // the DFT operand table is written
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    cpx<float>* v = h ? r1 : r0;
    stage_compute<float, CB_L, MODE_CONTIG, 0, SPLIT, 32, SPLIT>(v, tau, nullptr, buf, atab);  // radix 16, no twiddles
    if (h) __syncthreads();  // every thread has read the first half's stage-1 inputs
    stage_store_lds<float, CB_L, MODE_CONTIG, 0, 32>(v, buf, c, tau);
    __syncthreads();
    stage_load_lds<float, CB_L, MODE_CONTIG, 1, 32>(v, buf, c, tau);
    // radix-16 stage over sub-transforms of span 16: twiddles W256^{tau r}
    // from the LDS table (lanes of one tau read one address)
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], tws[15 * tau + r - 1]);
    wg_dft16<SPLIT, true>(v, atab);
  }

  // ---- y over b: columns kx = 128 hc + cc, 64-point lines; thread (cc, tq)
  // ends with point k = tq + 4 r of its column.  The first radix-2 step of the
  // 4-point stage is done on the way into LDS (each thread holds rows b and
  // b + 32), so the column stage reads its radix-16 inputs straight from that
  // buffer: one LDS exchange and two barriers per column half instead of two
  // and four (measured 53.4 -> 52.8 us)
  const int cc = tid & 127;
  const int tq = __builtin_amdgcn_readfirstlane(tid >> 7);  // uniform over each wave
  // b = b0 + 16 b1, b1 = lo + 2 hi: thread c holds b0 = c & 15, lo = c >> 4 and
  // both hi (rows c and c + 32)
  const int b0 = c & 15, lo = c >> 4;
#pragma unroll
  for (int hc = 0; hc < 2; ++hc) {
    __syncthreads();  // the buffer's previous contents have been read
    // E_lo[m] = x[lo] + (-1)^m x[lo + 2] to LDS row b0 + 16 (m + 2 lo), E_1[1]
    // times -i (lo is uniform over the wave)
    if (lo == 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const cpx<float> u = r0[8 * hc + r], w = r1[8 * hc + r];
        buf[CaCol::idx(tau + 16 * r, b0)] = cadd(u, w);
        buf[CaCol::idx(tau + 16 * r, b0 + 16)] = csub(u, w);
      }
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const cpx<float> u = r0[8 * hc + r], w = r1[8 * hc + r];
        buf[CaCol::idx(tau + 16 * r, b0 + 32)] = cadd(u, w);
        buf[CaCol::idx(tau + 16 * r, b0 + 48)] = cpx<float>{u.y - w.y, w.x - u.x};
      }
    }
    __syncthreads();
    // thread (cc, tq), k1 = tq: X4[b0] = E_0[m] + W4^{k1} E_1[m], m = k1 & 1, then
    // W64^{k1 b0} and the 16-point DFT over b0.
    // W4^{tq} E_1[m] = +-E'_1[m] with E'_1[1] = -i E_1[1] (stored so by the
    // writer), the sign uniform over the wave
    cpx<float> v[16];
    const int m = tq & 1;
    const float sg = (tq & 2) ? -1.f : 1.f;
#pragma unroll
    for (int bb = 0; bb < 16; ++bb) {
      const cpx<float> e0 = buf[CaCol::idx(cc, bb + 16 * m)];
      const cpx<float> e1 = buf[CaCol::idx(cc, bb + 16 * (m + 2))];
      v[bb] = {__builtin_fmaf(sg, e1.x, e0.x), __builtin_fmaf(sg, e1.y, e0.y)};  // exact: e0 +- e1
    }
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], wdir[4 * tq * r]);
    wg_dft16<SPLIT, true>(v, atab);
    // column DFT point k = tq + 4 r to row 64 res + k (pass B applies W256^{res k})
#pragma unroll
    for (int r = 0; r < 16; ++r)
      out[(long long)(64 * res + tq + 4 * r) * CB_L + 128 * hc + cc] = v[r];
  }
}

template <bool INV, bool WS, bool SPLIT>
__global__ void __launch_bounds__(CB_NT, 4) k_fft_cube_b(const CubeArgs a) {
  __shared__ cpx<float> buf[CbZ::LDS_ELEMS];
  __shared__ split_half8 atab[SPLIT ? SPLIT_ATAB : 1];  // split DFT operands (read after a barrier)
  // k varies fastest over blockIdx: the tiles in flight together then cover
  // all 64 rows k (measured on this access pattern alone, in place after pass
  // A: 50.3 against 53.4 us with the column group fastest)
  const unsigned b = blockIdx.x;
  const int k = b & 63;                // rows k + 64 a
  const int xg = (b >> 6) & 15;        // group of 16 columns
  const unsigned long long blk = b >> 10;
  if (blk >= a.nplanes / CB_L) return;  // uniform over the workgroup
  const long long base = (long long)blk * CB_BLOCK + (long long)k * CB_L + 16 * xg;
  const cpx<float>* in = reinterpret_cast<const cpx<float>*>(a.in) + base;
  cpx<float>* out = reinterpret_cast<cpx<float>*>(a.out) + base;
  const cpx<float>* __restrict__ tw256 = reinterpret_cast<const cpx<float>*>(a.tw);
  const int tid = threadIdx.x;
  if constexpr (SPLIT) split_atab_fill(atab);

  // ---- y over a: thread (x, zs) holds planes z = zs + 32 j; t[j][d] = X[k + 64 d].
  // Sixteen adjacent lanes read one 128-byte row segment.
  const int x = tid & 15, zs = tid >> 4;
  cpx<float> t[8][4];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const long long zo = (long long)(zs + 32 * j) * CB_PLANE + x;
#pragma unroll
    for (int aa = 0; aa < 4; ++aa) t[j][aa] = in[zo + 64 * CB_L * aa];
  }
  // T_a[k] = W256^{a k} times pass A's column DFT (k is uniform over the workgroup)
  const cpx<float>* __restrict__ wdir = reinterpret_cast<const cpx<float>*>(a.w256);
  const cpx<float> w1 = wdir[k], w2 = wdir[2 * k], w3 = wdir[3 * k];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    t[j][1] = cmul(t[j][1], w1);
    t[j][2] = cmul(t[j][2], w2);
    t[j][3] = cmul(t[j][3], w3);
    dft4(t[j]);
  }

  // ---- z: the 32 lines of rows d = 2 e, 2 e + 1 at a time; line cz = 16 (d - 2 e) + x.
  // z = zq + 16 r with r = ep + 2 j: the first z stage's radix-16 over r is an
  // 8-point DFT over j, which a thread holding planes zs + 32 j (zs = zq + 16 ep)
  // does in registers, P_ep[zq][m]; the stage-1 thread for k0 finishes it as
  // Y[zq][k0] = P_0[zq][k0 & 7] + W16^{k0} P_1[zq][k0 & 7] while reading LDS, so
  // the z transform needs one LDS exchange
  const int zq = zs & 15, ep = zs >> 4;
  const int cz = tid & 31;
  const int tz = tid >> 5;                 // k0
  const cpx<float> w16 = wdir[16 * tz];    // W16^{k0}
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
      for (int m = 0; m < 8; ++m) buf[CbZ::idx(16 * dd + x, zq + 16 * ep + 32 * m)] = p8[m];
    }
    __syncthreads();
    cpx<float> v[16];
    const int m = tz & 7;
#pragma unroll
    for (int q = 0; q < 16; ++q)
      v[q] = cadd(buf[CbZ::idx(cz, q + 32 * m)], cmul(buf[CbZ::idx(cz, q + 16 + 32 * m)], w16));
    stage_compute<float, CB_L, MODE_STRIDED, 1, SPLIT, 32, SPLIT>(v, tz, tw256, buf, atab);  // W256^{k0 zq}, DFT over zq
    // point kz = tz + 16 r of row k + 64 d, column 16 xg + (cz & 15)
    const long long ob = (long long)(64 * (2 * e + (cz >> 4))) * CB_L + (cz & 15);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long long o = ob + (long long)(tz + 16 * r) * CB_PLANE;
      cpx<float> s = INV ? cswap(v[r]) : v[r];
      if (WS) s = cscale(s, a.w_store[(unsigned long long)(base + o) & a.w_mask]);
      // non-temporal: measured 2% faster over a forward + inverse round trip
      // (49.9 -> 47.9 us for this pass; non-temporal loads in pass A were slower)
      __builtin_nontemporal_store(__builtin_bit_cast(long long, s), reinterpret_cast<long long*>(out + o));
    }
  }
}

template <bool INV, bool WL, bool SPLIT>
__global__ void __launch_bounds__(C1_NT, 4) k_fft_cube128_a(const CubeArgs a) {
  __shared__ cpx<float> buf[C1A_LDS];
  __shared__ cpx<float> tws[C1A_TWS];
  cube128_a_tile<INV, WL, SPLIT>(a, blockIdx.x, buf, tws, (int)threadIdx.x);
}

template <bool INV, bool WS, bool SPLIT>
__global__ void __launch_bounds__(C1B_NT, 4) k_fft_cube128_b(const CubeArgs a) {
  __shared__ cpx<float> buf[C1Z::LDS_ELEMS];
  __shared__ split_half8 atab[SPLIT ? SPLIT_ATAB : 1];  // split DFT operands (read after a barrier)
  if constexpr (SPLIT) split_atab_fill(atab);
  cube128_b_tile<INV, WS, SPLIT>(a, blockIdx.x, buf, atab, (int)threadIdx.x);
}

void preload_cube() { preload_kernel(reinterpret_cast<const void*>(&k_fft_cube_a<false, false, false>)); }

static bool cube_grid(unsigned long long nplanes, unsigned long long per, unsigned long long* blocks) {
  const unsigned long long b = nplanes * per;
  if (b > 0xFFFFFFFFull / CB_NT) return false;  // grid size in work-items is 32-bit (512 threads bounds every kernel here)
  *blocks = b;
  return true;
}

// KER<inverse, weights, split>; the split-precision variants take no weights
#define CUBE_DISPATCH(KER, W, G, T, S, A)                                  \
  do {                                                                     \
    if ((A).split) {                                                       \
      if ((A).inverse) FFTAMD_LAUNCH((KER<true, false, true>), G, T, 0, S, A); \
      else FFTAMD_LAUNCH((KER<false, false, true>), G, T, 0, S, A);        \
    } else if ((A).inverse) {                                              \
      if (W) FFTAMD_LAUNCH((KER<true, true, false>), G, T, 0, S, A);       \
      else FFTAMD_LAUNCH((KER<true, false, false>), G, T, 0, S, A);        \
    } else {                                                               \
      if (W) FFTAMD_LAUNCH((KER<false, true, false>), G, T, 0, S, A);      \
      else FFTAMD_LAUNCH((KER<false, false, false>), G, T, 0, S, A);       \
    }                                                                      \
  } while (0)

int launch_fft_cube_a(const CubeArgs& a, void* stream) {
  if (a.n == 512) return launch_fft_c512_a(a, stream);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  unsigned long long blocks;
  if ((a.n != CB_L && a.n != C1_L) || a.nplanes == 0 || !cube_grid(a.nplanes, 4, &blocks) || a.w_store ||
      (a.tab && (a.n != C1_L || a.w_load)) || (a.split && a.w_load))
    return -1;
  if (a.n == C1_L) {
    const dim3 g((unsigned)blocks), t(C1_NT);
    CUBE_DISPATCH(k_fft_cube128_a, a.w_load, g, t, s, a);
  } else {
    const dim3 g((unsigned)blocks), t(CB_NT);
    CUBE_DISPATCH(k_fft_cube_a, a.w_load, g, t, s, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_fft_cube_b(const CubeArgs& a, void* stream) {
  if (a.n == 512) return launch_fft_c512_b(a, stream);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  unsigned long long blocks;
  if ((a.n != CB_L && a.n != C1_L) || a.nplanes == 0 || a.nplanes % a.n || a.w_load ||
      (a.tab && (a.n != C1_L || a.w_store)) || (a.split && a.w_store))
    return -1;
  if (a.n == C1_L) {
    if (!cube_grid(a.nplanes / C1_L, 128, &blocks)) return -1;
    const dim3 g((unsigned)blocks), t(C1B_NT);
    CUBE_DISPATCH(k_fft_cube128_b, a.w_store, g, t, s, a);
  } else {
    if (!cube_grid(a.nplanes / CB_L, 1024, &blocks)) return -1;
    const dim3 g((unsigned)blocks), t(CB_NT);
    CUBE_DISPATCH(k_fft_cube_b, a.w_store, g, t, s, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fftamd

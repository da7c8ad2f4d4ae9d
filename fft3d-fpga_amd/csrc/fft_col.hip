// fft_col.hip -- the 4096-point transform along the second-fastest axis of an
// array (the columns of an R x 4096 x LX block, LX >= 16) in one pass over HBM.
//
// A line pass holds whole lines in LDS, and 4096-point columns of 16 adjacent
// columns (the narrowest coalesced group, 128-byte row segments) are 512 KiB,
// so otherwise the axis is split four-step into two passes.  Here four
// workgroups share a group of 16 columns, each owning a quarter of the output
// rows, as the 256 x 256 plane pass shares a plane: 4096 = 256 x 16, n = n1 +
// 256 n2, k = k1 + 16 k2, and workgroup q owns the rows with k1 = q mod 4
// (k1 = q + 4u, u = 0..3, k = q + 4u + 16 k2).  It reads all 4096 rows of the
// group, evaluates only its four outputs of each first-stage 16-point DFT over
// n2 (one residue class: four rotated sums, three fixed twiddles, a 4-point
// DFT), then for each of its 4 x 16 lines (u, column) multiplies by
// W4096^{n1 k1} and runs the 256-point transform over n1 with the line
// passes' strided stages.  The four workgroups of a group are consecutive
// blocks of one XCD, so the group comes from HBM once and from that XCD's L2
// for the other three.
//
// 512 threads; a thread keeps half of its first-stage outputs in registers
// while LDS holds the other half as 32 lines of 256 points (67.6 KiB), so two
// workgroups share a CU.  Rows are addressed through buffer resources with one
// VGPR offset per lane and scalar row offsets.  Out of place only.
#include <hip/hip_runtime.h>

#include "launch.h"

#include "fft_pass_dev.h"
#include "fft_plane.h"

namespace fftamd {

constexpr int CL_L = 4096;           // column length
constexpr int CL_C = 16;             // columns per group
constexpr int CL_Q = 4;              // workgroups per group
constexpr int CL_NT = 512;           // threads per workgroup
constexpr int CL_N1 = 256;           // n = n1 + 256 n2
constexpr int CL_NG = CL_NT / CL_C;  // first-stage n1 groups: thread handles n1 = gi + 32 i
constexpr int CL_PER = CL_N1 / CL_NG;
using ColGeom = Geom<CL_N1, MODE_STRIDED, 2 * CL_C>;  // 32 lines (two u, 16 columns) of 256 points
static_assert(ColGeom::NT == CL_NT, "16 threads per 256-point line, 32 lines");
static_assert(CL_PER == 8, "eight n1 per thread");

// x * (-i)^M
template <int M>
__device__ __forceinline__ cpx<float> col_rot(cpx<float> x) {
  constexpr int m = M & 3;
  if constexpr (m == 0) return x;
  else if constexpr (m == 1) return {x.y, -x.x};
  else if constexpr (m == 2) return {-x.x, -x.y};
  else return {-x.y, x.x};
}

// x * W16^P
template <int P>
__device__ __forceinline__ cpx<float> col_w16(cpx<float> x) {
  constexpr int p = P & 15;
  if constexpr ((p & 3) == 0) return col_rot<p / 4>(x);
  else if constexpr (p == 2) return mul_w8_1(x);
  else if constexpr (p == 6) return mul_w8_3(x);
  else if constexpr (p == 10) return col_rot<2>(mul_w8_1(x));
  else if constexpr (p == 14) return col_rot<2>(mul_w8_3(x));
  else {  // odd p: cos(pi p / 8) - i sin(pi p / 8)
    constexpr float c = (p == 1 || p == 15) ? K<float>::C16 : (p == 3 || p == 13) ? K<float>::S16
                        : (p == 5 || p == 11) ? -K<float>::S16 : -K<float>::C16;
    constexpr float sn = (p == 1 || p == 7) ? K<float>::S16 : (p == 3 || p == 5) ? K<float>::C16
                         : (p == 9 || p == 15) ? -K<float>::S16 : -K<float>::C16;
    return cmul(x, cpx<float>{c, -sn});
  }
}

// Outputs G, G + 4, G + 8, G + 12 of the 16-point DFT of v[] (n = a + 4 b:
// X[G + 4u] = sum_a (-i)^{a u} W16^{a G} sum_b v[a + 4 b] (-i)^{b G}).
template <int G>
__device__ __forceinline__ void col_residue(const cpx<float>* v, cpx<float>* o) {
  cpx<float> s[4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
    s[a] = cadd(cadd(v[a], col_rot<G>(v[a + 4])), cadd(col_rot<2 * G>(v[a + 8]), col_rot<3 * G>(v[a + 12])));
  s[1] = col_w16<G>(s[1]);
  s[2] = col_w16<2 * G>(s[2]);
  s[3] = col_w16<3 * G>(s[3]);
  dft4(s);
#pragma unroll
  for (int u = 0; u < 4; ++u) o[u] = s[u];
}

typedef unsigned col_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t col_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

// Stage 2 of lines (u = 2h + l, column c16) for l = 0, 1 held in LDS: the
// twiddle W4096^{n1 k1}, the 256-point transform over n1 and the stores of
// rows k = k1 + 16 k2.  Thread (line c = 16 l + c16, tau).
template <int Q, bool INV, bool WS>
__device__ __forceinline__ void col_stage2(const ColArgs& a, __amdgpu_buffer_rsrc_t rout,
                                           __amdgpu_buffer_rsrc_t rws, int h, cpx<float>* lds,
                                           const cpx<float>* tws) {
  const int tid = threadIdx.x;
  const int c = tid % ColGeom::C, tau = tid / ColGeom::C;
  const int k1 = Q + 4 * (2 * h + (c >> 4));
  cpx<float> v[16];
  stage_load_lds<float, CL_N1, MODE_STRIDED, 0, ColGeom::C>(v, lds, c, tau);
  // point n1 = tau + 16 r: W4096^{n1 k1} = W4096^{tau k1} W256^{r k1}
  {
    const cpx<float>* lo = reinterpret_cast<const cpx<float>*>(a.ptw_lo);
    const cpx<float>* hi = reinterpret_cast<const cpx<float>*>(a.ptw_hi);
    const unsigned t = (unsigned)(tau * k1);  // < 256
    const cpx<float> base = cmul(hi[t >> a.ptw_hb], lo[t & ((1u << a.ptw_hb) - 1u)]);
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = cmul(v[r], base);
    // W256^{r k1} from the LDS table (k1 < 16)
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], tws[15 * k1 + r - 1]);
  }
  __syncthreads();  // every thread holds its stage-0 inputs before stage 0 stores
  stage_compute<float, CL_N1, MODE_STRIDED, 0, false, ColGeom::C>(v, tau, nullptr);  // radix 16, no twiddles
  stage_store_lds<float, CL_N1, MODE_STRIDED, 0, ColGeom::C>(v, lds, c, tau);
  __syncthreads();
  stage_load_lds<float, CL_N1, MODE_STRIDED, 1, ColGeom::C>(v, lds, c, tau);
  // radix-16 stage over sub-transforms of span 16: twiddles W256^{tau r}
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], tws[15 * tau + r - 1]);
  dft16(v);
  // last stage: point k2 = tau + 16 r of the line; row k1 + 16 k2, 16 lanes on
  // 16 adjacent columns (128 contiguous bytes)
  static_assert(sched_np(CL_N1) == 2 && ColGeom::TPL == 16, "256 = 16 x 16, one butterfly per lane");
  const unsigned lx8 = (unsigned)a.lx * 8u;
  const unsigned voff = (unsigned)(k1 + 16 * tau) * lx8 + (unsigned)(c & 15) * 8u;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    cpx<float> x = INV ? cswap(v[r]) : v[r];
    const unsigned soff = (unsigned)(256 * r) * lx8;
    if (WS) x = cscale(x, __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rws, voff / 2u, soff / 2u, 0)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(col_u2, x), rout, voff, soff, 0);
  }
}

template <int Q, bool INV, bool WL, bool WS>
__device__ __forceinline__ void col_quarter(const ColArgs& a, unsigned long long o, unsigned grp, cpx<float>* lds,
                                            const cpx<float>* tws) {
  // keep the first loads inside this part's code: hoisted above the switch
  // over the part, they were spilled across it
  asm volatile("" ::: "memory");
  const int tid = threadIdx.x;
  const int c16 = tid & (CL_C - 1);
  // n1 group: 16 threads per group, four groups per wave (not wave-uniform)
  const int gi = tid >> 4;
  const unsigned lx8 = (unsigned)a.lx * 8u;
  const unsigned long long first = o * (unsigned long long)a.in_o + (unsigned long long)grp * CL_C;
  const unsigned bytes = (unsigned)(CL_L - 1) * lx8 + CL_C * 8u;  // the group's span
  const __amdgpu_buffer_rsrc_t rin = col_rsrc(reinterpret_cast<const cpx<float>*>(a.in) + first, bytes);
  const __amdgpu_buffer_rsrc_t rout = col_rsrc(reinterpret_cast<cpx<float>*>(a.out) + first, bytes);
  // weights of element e of the block are w[(o * in_o) & w_mask + e]: the
  // weight period is a multiple of the block (checked by the launcher)
  const unsigned long long wfirst = ((o * (unsigned long long)a.in_o) & a.w_mask) + (unsigned long long)grp * CL_C;
  const __amdgpu_buffer_rsrc_t rwl = col_rsrc(a.w_load + wfirst, bytes / 2u);
  const __amdgpu_buffer_rsrc_t rws = col_rsrc(a.w_store + wfirst, bytes / 2u);
  const unsigned voff = (unsigned)gi * lx8 + (unsigned)c16 * 8u;

  // ---- first stage: thread (c16, gi) for n1 = gi + 32 i: the 16-point DFT
  // over n2 (rows n1 + 256 n2), outputs k1 = Q + 4u; u = 0, 1 go to LDS as
  // lines 16 u + c16, u = 2, 3 stay in registers
  cpx<float> held[CL_PER][2];
  const int rot = (int)(grp & 7u);
#pragma unroll
  for (int i = 0; i < CL_PER; ++i) {
    const int ir = (i + rot) & 7;
    cpx<float> v[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const unsigned soff = (unsigned)(CL_NG * ir + CL_N1 * n2) * lx8;
      cpx<float> s = __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(rin, voff, soff, 0));
      if (WL) s = cscale(s, __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rwl, voff / 2u, soff / 2u, 0)));
      v[n2] = INV ? cswap(s) : s;
    }
    cpx<float> out4[4];
    col_residue<Q>(v, out4);
    const int n1 = gi + CL_NG * ir;
    lds[ColGeom::idx(c16, n1)] = out4[0];
    lds[ColGeom::idx(CL_C + c16, n1)] = out4[1];
    held[i][0] = out4[2];
    held[i][1] = out4[3];
    // the registers are needed for the values held: keep the next group's
    // loads below this group's arithmetic
    __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
  col_stage2<Q, INV, WS>(a, rout, rws, 0, lds, tws);
  __syncthreads();  // the last stage has read the buffer
#pragma unroll
  for (int i = 0; i < CL_PER; ++i) {
    const int n1 = gi + CL_NG * ((i + rot) & 7);
    lds[ColGeom::idx(c16, n1)] = held[i][0];
    lds[ColGeom::idx(CL_C + c16, n1)] = held[i][1];
  }
  __syncthreads();
  col_stage2<Q, INV, WS>(a, rout, rws, 1, lds, tws);
}

// Two workgroups per CU: at most 128 VGPRs, 67.6 KiB LDS each.
template <bool INV, bool WL, bool WS>
__global__ void __launch_bounds__(CL_NT, 4) k_fft_col4096(const ColArgs a) {
  __shared__ cpx<float> lds[ColGeom::LDS_ELEMS];
  __shared__ cpx<float> tws[16 * 15];  // W256^{k r} at [15 k + r - 1]: stage-2 twiddles (read after barriers)
  if (threadIdx.x < 16 * 15)
    tws[threadIdx.x] = reinterpret_cast<const cpx<float>*>(a.w256)[(threadIdx.x / 15) * (threadIdx.x % 15 + 1)];
  // XCD-aware mapping as in the 256 x 256 plane pass: block b runs on XCD
  // b % 8; the k-th block of an XCD takes part k % 4 of item (k / 4) * 8 +
  // XCD, so the four parts of a column group are consecutive blocks of one XCD
  const unsigned b = blockIdx.x;
  const unsigned xcd = b & 7u, k = b >> 3;
  const unsigned q = k & 3u;
  const unsigned long long item = (unsigned long long)(k >> 2) * 8ull + xcd;
  if (item >= a.nitems) return;  // uniform over the workgroup
  const unsigned grp = (unsigned)(item % a.ngroups);
  const unsigned long long o = item / a.ngroups;
  switch (q) {  // uniform: the part is a compile-time constant inside
    case 0: col_quarter<0, INV, WL, WS>(a, o, grp, lds, tws); break;
    case 1: col_quarter<1, INV, WL, WS>(a, o, grp, lds, tws); break;
    case 2: col_quarter<2, INV, WL, WS>(a, o, grp, lds, tws); break;
    default: col_quarter<3, INV, WL, WS>(a, o, grp, lds, tws); break;
  }
}

template <bool INV, bool WL, bool WS>
static void launch_colw(const ColArgs& a, hipStream_t s, unsigned blocks) {
  FFTAMD_LAUNCH((k_fft_col4096<INV, WL, WS>), dim3(blocks), dim3(CL_NT), 0, s, a);
}

template <bool INV>
static void launch_coli(const ColArgs& a, hipStream_t s, unsigned blocks) {
  if (a.w_load) {
    if (a.w_store) launch_colw<INV, true, true>(a, s, blocks);
    else launch_colw<INV, true, false>(a, s, blocks);
  } else {
    if (a.w_store) launch_colw<INV, false, true>(a, s, blocks);
    else launch_colw<INV, false, false>(a, s, blocks);
  }
}

bool col4096_supported(long long lx, unsigned long long nouter) {
  // 16-column groups; the group's span (4095 rows + 16 columns) addressed
  // with 32-bit byte offsets
  return lx >= CL_C && lx % CL_C == 0 && (unsigned long long)(CL_L - 1) * (unsigned long long)lx * 8ull +
                                             CL_C * 8ull < (1ull << 31) && nouter > 0;
}

void preload_col() { preload_kernel(reinterpret_cast<const void*>(&k_fft_col4096<false, false, false>)); }

int launch_fft_col4096(const ColArgs& a, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!col4096_supported(a.lx, a.nitems / (a.ngroups ? a.ngroups : 1)) || a.ngroups != (unsigned)(a.lx / CL_C) ||
      a.in_o != (long long)CL_L * a.lx)
    return -1;
  // weights: element e of block o has weight index (o * in_o + e) & w_mask,
  // read as ((o * in_o) & w_mask) + e, which needs a period that is a
  // multiple of the block
  if ((a.w_load || a.w_store) && (a.w_mask + 1ull) % (unsigned long long)a.in_o != 0) return -1;
  const unsigned long long items8 = (a.nitems + 7ull) / 8ull * 8ull;
  const unsigned long long blocks = items8 * CL_Q;
  if (blocks > 0xFFFFFFFFull / CL_NT) return -1;
  if (a.inverse) launch_coli<true>(a, s, (unsigned)blocks);
  else launch_coli<false>(a, s, (unsigned)blocks);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fftamd

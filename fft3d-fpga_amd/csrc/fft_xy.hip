// fft_xy.hip -- fused x+y pass over 256 x 256 planes: the 2D transform of
// every plane of an array in one pass over HBM.
//
// A plane is 512 KiB, more than one CU's LDS, so four workgroups share it,
// each owning 64 of the 256 output rows.  Column transform first: 256 =
// 16 x 16, y = n1 + 16 n2, ky = k1 + 16 k2, and workgroup q owns the rows with
// k1 = q mod 4 (k1 = q + 4u, ky = 16 k2 + 4u + q).  It reads the whole plane
// in row segments, evaluates only the four k1 outputs of each first-stage
// 16-point DFT and only its four second-stage DFTs per column, so no
// arithmetic is repeated across the four workgroups; four outputs of one
// residue class cost a third of the arithmetic of a whole 16-point DFT.  Its 64 rows then get
// their 256-point row transforms (x) and are written as whole output rows.
// The four workgroups of a plane are mapped to one XCD and dispatched back to
// back (workgroups go to XCDs round-robin by index), so the plane comes from
// HBM once and from that XCD's L2 for the other three.  A 3D transform then
// takes two passes over HBM (this one and the z pass) instead of three.
//
// The workgroup's 64 x 256 results stay in registers (32 points per thread)
// and LDS only stages half of them at a time: the column stages run over one
// half of the columns at a time, the row transforms over 32 rows at a time.
// That keeps LDS at 68 KiB, so two workgroups share a CU and one loads or
// stores while the other computes.
//
// The pass reads the whole input plane and writes only its rows of the
// output, so it runs out of place.
#include <hip/hip_runtime.h>

#include "launch.h"

#include "fft_pass_dev.h"
#include "fft_plane.h"

namespace fftamd {

constexpr int XY_L = 256;
constexpr int XY_Q = 4;             // workgroups per plane
constexpr int XY_QR = XY_L / XY_Q;  // output rows per workgroup
constexpr int XY_NT = 512;          // threads per workgroup
constexpr int XY_XH = XY_L / 2;     // columns per column half
constexpr int XY_RH = XY_QR / 2;    // rows per row half
using RowGeom = Geom<XY_L, MODE_CONTIG, XY_RH>;  // 32 rows in LDS
static_assert(RowGeom::NT == XY_NT, "16 threads per row, 32 rows");
static_assert(XY_NT == 4 * XY_XH, "column lanes: 128 columns x 4 groups");

// LDS buffer, in turn:
//  (1) first-stage column results of one column half, Y[n1][u][x] at
//      (4 n1 + u) * 128 + x (16 x 4 x 128 points; a wave reads or writes 64
//      consecutive x);
//  (2) 32 of the workgroup's rows, RowGeom layout [row][x + x/16] (local row
//      r = 4 k2 + u holds output row ky = 16 k2 + 4u + q).
constexpr int XY_LDS = RowGeom::LDS_ELEMS > 16 * 4 * XY_XH ? RowGeom::LDS_ELEMS : 16 * 4 * XY_XH;

// Raw buffer resource over one plane (stride 0, byte range, 32-bit data
// format as CDNA requires for untyped buffer loads).  Its loads take a VGPR
// byte offset and a scalar byte offset, so the sixteen row loads of a column
// share one address VGPR instead of sixteen 64-bit addresses.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const cpx<float>* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<cpx<float>*>(base), (short)0,
                                           XY_L * XY_L * (int)sizeof(cpx<float>), 0x00020000);
}

__device__ __forceinline__ cpx<float> plane_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Weights of one plane: plane_weights() is the resource over the weights of
// element 0 .. 65535 of a plane whose first element has weight index `first`
// (indices are (plane * stride + e) & w_mask; the launcher checks that this
// is first + e).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_weights(const float* w, unsigned long long first) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(w) + first, (short)0,
                                           XY_L * XY_L * (int)sizeof(float), 0x00020000);
}

__device__ __forceinline__ float weight_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// output row of local row r of workgroup q
__device__ __forceinline__ int xy_grow(int r, int q) { return 16 * (r >> 2) + 4 * (r & 3) + q; }

// x * (-i)^M
template <int M>
__device__ __forceinline__ cpx<float> rot_negi(cpx<float> x) {
  constexpr int m = M & 3;
  if constexpr (m == 0) return x;
  else if constexpr (m == 1) return {x.y, -x.x};
  else if constexpr (m == 2) return {-x.x, -x.y};
  else return {-x.y, x.x};
}

// x * W16^P, W16 = e^{-2 pi i / 16}
template <int P>
__device__ __forceinline__ cpx<float> mul_w16(cpx<float> x) {
  constexpr int p = P & 15;
  if constexpr ((p & 3) == 0) return rot_negi<p / 4>(x);
  else if constexpr (p == 2) return mul_w8_1(x);
  else if constexpr (p == 6) return mul_w8_3(x);
  else if constexpr (p == 10) return rot_negi<2>(mul_w8_1(x));
  else if constexpr (p == 14) return rot_negi<2>(mul_w8_3(x));
  else {  // odd p: cos(pi p / 8) - i sin(pi p / 8) from cos(pi/8), sin(pi/8)
    constexpr float c = (p == 1 || p == 15) ? K<float>::C16 : (p == 3 || p == 13) ? K<float>::S16
                        : (p == 5 || p == 11) ? -K<float>::S16 : -K<float>::C16;
    constexpr float sn = (p == 1 || p == 7) ? K<float>::S16 : (p == 3 || p == 5) ? K<float>::C16
                         : (p == 9 || p == 15) ? -K<float>::S16 : -K<float>::C16;
    return cmul(x, cpx<float>{c, -sn});
  }
}

// Outputs G, G + 4, G + 8, G + 12 of the 16-point DFT of v[] into o[0..3].
// With n = a + 4 b: X[G + 4u] = sum_a (-i)^{a u} W16^{a G} S_a, S_a = sum_b
// v[a + 4 b] (-i)^{b G}: four 4-term sums with free rotations, three fixed
// twiddles and one 4-point DFT.
template <int G>
__device__ __forceinline__ void dft16_residue(const cpx<float>* v, cpx<float>* o) {
  cpx<float> s[4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
    s[a] = cadd(cadd(v[a], rot_negi<G>(v[a + 4])), cadd(rot_negi<2 * G>(v[a + 8]), rot_negi<3 * G>(v[a + 12])));
  s[1] = mul_w16<G>(s[1]);
  s[2] = mul_w16<2 * G>(s[2]);
  s[3] = mul_w16<3 * G>(s[3]);
  dft4(s);
#pragma unroll
  for (int u = 0; u < 4; ++u) o[u] = s[u];
}

// v[r] *= W_256^{k r} from the radix-16 stage table (W^k, W^2k, W^4k, W^8k)
__device__ __forceinline__ void twiddle256(cpx<float>* v, int k, const cpx<float>* __restrict__ tw) {
  const cpx<float>* t = tw + sched_twoff(XY_L, 1) + TW_PER_K * k;
  const cpx<float> w1 = t[0], w2 = t[1], w4 = t[2], w8 = t[3];
  const cpx<float> w3 = cmul(w1, w2), w12 = cmul(w4, w8);
  v[1] = cmul(v[1], w1);
  v[2] = cmul(v[2], w2);
  v[3] = cmul(v[3], w3);
  v[4] = cmul(v[4], w4);
  v[5] = cmul(v[5], cmul(w4, w1));
  v[6] = cmul(v[6], cmul(w4, w2));
  v[7] = cmul(v[7], cmul(w4, w3));
  v[8] = cmul(v[8], w8);
  v[9] = cmul(v[9], cmul(w8, w1));
  v[10] = cmul(v[10], cmul(w8, w2));
  v[11] = cmul(v[11], cmul(w8, w3));
  v[12] = cmul(v[12], w12);
  v[13] = cmul(v[13], cmul(w12, w1));
  v[14] = cmul(v[14], cmul(w12, w2));
  v[15] = cmul(v[15], cmul(w12, w3));
}

// Column transforms of column half H of the plane for the rows of workgroup Q: lane
// (x, u) of the second stage returns the 16 points k2 = 0..15 of local rows
// 4 k2 + u at column 128 H + x in res[].  INV: inverse transform as
// swap(DFT(swap(x))), the swaps being register renames.  WL: inputs are
// scaled by the load weights.
template <int Q, bool INV, bool WL>
__device__ __forceinline__ void column_half(const PlaneArgs& a, const cpx<float>* __restrict__ in,
                                            unsigned long long plane, int h, cpx<float>* buf,
                                            cpx<float>* res) {
  const int tid = threadIdx.x;
  const int x = tid & (XY_XH - 1);
  // group of four n1: 128 threads per group, so it is uniform over each wave
  // and the row addresses below are scalar
  const int g = __builtin_amdgcn_readfirstlane(tid >> 7);
  const int col = h * XY_XH + x;
  const __amdgpu_buffer_rsrc_t rin = plane_rsrc(in);
  const unsigned voff = (unsigned)col * (unsigned)sizeof(cpx<float>);
  const __amdgpu_buffer_rsrc_t rw = plane_weights(a.w_load, (plane * a.plane_stride) & a.w_mask);
  // ---- first stage: for n1 = 4g .. 4g+3, the 16-point DFT over n2 of
  // column col at rows n1 + 16 n2, four outputs k1 = Q + 4u
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int n1 = 4 * g + m;
    cpx<float> v[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const unsigned rowe = (unsigned)((n1 + 16 * n2) * XY_L);  // first element of the row
      cpx<float> s = plane_load(rin, voff, rowe * (unsigned)sizeof(cpx<float>));
      if (WL) s = cscale(s, weight_load(rw, (unsigned)col * 4u, rowe * 4u));
      v[n2] = INV ? cswap(s) : s;
    }
    cpx<float> o[4];
    dft16_residue<Q>(v, o);
#pragma unroll
    for (int u = 0; u < 4; ++u) buf[(4 * n1 + u) * XY_XH + x] = o[u];
    // keep the scheduler from hoisting the next group's 16 loads above this
    // one's arithmetic: the registers are needed for the results held
    __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
  // ---- second stage: lane (x, u = g): X[k1 + 16 k2] = DFT16_n1(Y[n1][k1] W256^{n1 k1}), k1 = Q + 4u
  const int u = g;
#pragma unroll
  for (int n1 = 0; n1 < 16; ++n1) res[n1] = buf[(4 * n1 + u) * XY_XH + x];
  twiddle256(res, Q + 4 * u, reinterpret_cast<const cpx<float>*>(a.tw_y));
  dft16(res);
  __syncthreads();  // every first-stage value has been read before the buffer is reused
}

// Row transforms of local rows [32 j, 32 j + 32): lane (x, u) puts its
// k2 = 8 j .. 8 j + 7 points of both column halves into LDS, then thread
// (row c, tau) runs the line passes' stages along the row and writes output
// row xy_grow(32 j + c, q); 16 lanes write 128 contiguous bytes.  WS:
// outputs are scaled by the store weights.
template <bool INV, bool WS>
__device__ __forceinline__ void row_half(const PlaneArgs& a, cpx<float>* __restrict__ out,
                                         unsigned long long plane, unsigned q, int j, cpx<float>* buf,
                                         const cpx<float>* r0, const cpx<float>* r1) {
  const int tid = threadIdx.x;
  {
    const int x = tid & (XY_XH - 1), u = __builtin_amdgcn_readfirstlane(tid >> 7);
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      buf[RowGeom::idx(4 * kk + u, x)] = r0[8 * j + kk];
      buf[RowGeom::idx(4 * kk + u, XY_XH + x)] = r1[8 * j + kk];
    }
  }
  __syncthreads();
  const int c = tid / RowGeom::TPL, tau = tid % RowGeom::TPL;
  cpx<float> v[16];
  stage_load_lds<float, XY_L, MODE_CONTIG, 0, XY_RH>(v, buf, c, tau);
  __syncthreads();  // every thread holds its stage-0 inputs before stage 0 stores
  run_stages<float, XY_L, MODE_CONTIG, 0, false, XY_RH>(v, buf, c, tau,
                                                        reinterpret_cast<const cpx<float>*>(a.tw_x));
  // last stage: point kx = tau + 16 r of the row
  static_assert(sched_np(XY_L) == 2 && RowGeom::TPL == 16, "256 = 16 x 16, one butterfly per lane");
  const int row = xy_grow(XY_RH * j + c, (int)q) * XY_L;
  const __amdgpu_buffer_rsrc_t rw = plane_weights(a.w_store, (plane * a.plane_stride) & a.w_mask);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    cpx<float> s = INV ? cswap(v[r]) : v[r];
    const int e = row + tau + 16 * r;
    if (WS) s = cscale(s, weight_load(rw, (unsigned)(row + tau) * 4u, 64u * r));
    out[e] = s;
  }
}

template <int Q, bool INV, bool WL, bool WS>
__device__ __forceinline__ void xy_quarter(const PlaneArgs& a, const cpx<float>* __restrict__ in,
                                           cpx<float>* __restrict__ out, unsigned long long plane,
                                           cpx<float>* buf) {
  cpx<float> r0[16], r1[16];  // second-stage results of column halves 0 and 1
  column_half<Q, INV, WL>(a, in, plane, 0, buf, r0);
  column_half<Q, INV, WL>(a, in, plane, 1, buf, r1);
  row_half<INV, WS>(a, out, plane, Q, 0, buf, r0, r1);
  __syncthreads();  // the last row stage has read the buffer
  row_half<INV, WS>(a, out, plane, Q, 1, buf, r0, r1);
}

// Two workgroups per CU: at most 128 VGPRs (4 waves per SIMD), 68 KiB LDS each.
template <bool INV, bool WL, bool WS>
__global__ void __launch_bounds__(XY_NT, 4) k_fft_xy256(const PlaneArgs a) {
  __shared__ cpx<float> buf[XY_LDS];
  // XCD-aware mapping: block b runs on XCD b % 8 (dispatch is round-robin
  // over XCDs); the k-th block of an XCD takes quarter k % 4 of plane
  // (k / 4) * 8 + XCD, so the four quarters of a plane are consecutive blocks
  // of one XCD.
  const unsigned b = blockIdx.x;
  const unsigned xcd = b & 7u, k = b >> 3;
  const unsigned q = k & 3u;
  const unsigned long long plane = (unsigned long long)(k >> 2) * 8ull + xcd;
  if (plane >= a.nplanes) return;  // uniform over the workgroup
  const cpx<float>* __restrict__ in = reinterpret_cast<const cpx<float>*>(a.in) + plane * a.plane_stride;
  cpx<float>* __restrict__ out = reinterpret_cast<cpx<float>*>(a.out) + plane * a.plane_stride;
  switch (q) {  // uniform: the quarter is a compile-time constant inside
    case 0: xy_quarter<0, INV, WL, WS>(a, in, out, plane, buf); break;
    case 1: xy_quarter<1, INV, WL, WS>(a, in, out, plane, buf); break;
    case 2: xy_quarter<2, INV, WL, WS>(a, in, out, plane, buf); break;
    default: xy_quarter<3, INV, WL, WS>(a, in, out, plane, buf); break;
  }
}

template <bool INV, bool WL, bool WS>
static void launch_xyw(const PlaneArgs& a, hipStream_t s, unsigned blocks) {
  FFTAMD_LAUNCH((k_fft_xy256<INV, WL, WS>), dim3(blocks), dim3(XY_NT), 0, s, a);
}

template <bool INV>
static void launch_xyi(const PlaneArgs& a, hipStream_t s, unsigned blocks) {
  if (a.w_load) {
    if (a.w_store) launch_xyw<INV, true, true>(a, s, blocks);
    else launch_xyw<INV, true, false>(a, s, blocks);
  } else {
    if (a.w_store) launch_xyw<INV, false, true>(a, s, blocks);
    else launch_xyw<INV, false, false>(a, s, blocks);
  }
}

static void launch_xy(const PlaneArgs& a, hipStream_t s, unsigned blocks) {
  if (a.inverse) launch_xyi<true>(a, s, blocks);
  else launch_xyi<false>(a, s, blocks);
}

void preload_xy() { preload_kernel(reinterpret_cast<const void*>(&k_fft_xy256<false, false, false>)); }

int launch_fft_xy256(const PlaneArgs& a, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // 4 blocks per plane, planes rounded up to a multiple of 8 (one per XCD)
  const unsigned long long planes8 = (a.nplanes + 7ull) / 8ull * 8ull;
  const unsigned long long blocks = planes8 * XY_Q;
  if (blocks > 0xFFFFFFFFull / XY_NT) return -1;  // grid size in work-items is 32-bit
  // weights are read per plane from index (plane * stride) & w_mask on, which
  // holds when planes start on 65536-element boundaries of the weight period
  constexpr unsigned long long PL = (unsigned long long)XY_L * XY_L;
  if ((a.w_load || a.w_store) && ((a.w_mask + 1ull) % PL != 0 || a.plane_stride % (long long)PL != 0))
    return -1;
  launch_xy(a, s, (unsigned)blocks);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fftamd

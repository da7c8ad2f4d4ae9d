// fft_xy.hip -- fused x+y pass over 256 x 256 planes: the 2D transform of
// every plane of an array in one pass over HBM.
//
// A plane is 512 KiB, more than one CU's LDS, so four workgroups share it.
// Row transform: 256 = 16 x 16, n = n1 + 16 n2, k = k1 + 16 k2.  Workgroup q
// owns the output columns with k1 in [4q, 4q+4) (k = 16 k2 + 4q + u): it reads
// the whole plane, evaluates only the four k1 outputs of each first-stage
// 16-point DFT and only its four second-stage DFTs per row, so no arithmetic
// is repeated across the four workgroups.  Its 64 columns stay in a 128 KiB
// LDS buffer; then it runs the 256-point column transforms (y) on them and
// writes them to the output plane.  The four workgroups of a plane are mapped
// to one XCD and dispatched back to back (workgroups go to XCDs round-robin by
// index), so the plane comes from HBM once and from that XCD's L2 for the
// other three, and the four 32-byte pieces of each 128-byte output line meet
// in the same L2.  A 3D transform then takes two passes over HBM (this one and
// the z pass) instead of three.
//
// The pass reads the whole input plane and writes only its columns of the
// output, so it runs out of place.
#include <hip/hip_runtime.h>

#include "fft_pass_dev.h"
#include "fft_plane.h"

namespace fftamd {

constexpr int XY_L = 256;
constexpr int XY_Q = 4;                 // workgroups per plane
constexpr int XY_QC = XY_L / XY_Q;      // columns per workgroup
constexpr int XY_NT = 1024;             // threads per workgroup
constexpr int XY_TPR = 8;               // first row stage: threads per row
constexpr int XY_ROWS = XY_NT / XY_TPR; // rows per first-stage chunk

// The 128 KiB LDS buffer holds, in turn:
//  (1) the first-stage results, row y: 4 blocks (u) of 16 points (n1), read
//      and written in 16-byte granules (two n1); granule G = 8u + n1/2 is
//      XOR-swizzled inside its block so that the second stage's reads (16
//      lanes = 4 rows x 4 blocks per step) hit distinct banks;
//  (2) the row-transformed columns, [256 rows][64 columns], column index
//      XOR-swizzled for the second stage's stores (4 rows x 4 columns per 16
//      lanes); the column stage reads and writes whole rows.
// Both maps are bijections within a row.
__device__ __forceinline__ int xg_idx(int row, int G) {
  return row * 32 + (G ^ (2 * (row & 3) + ((G >> 4) & 1)));
}
__device__ __forceinline__ int cb_idx(int row, int col) {
  return row * XY_QC + (col ^ ((row & 7) << 2) ^ ((col >> 5) << 3));
}

// output column of local column c of workgroup q (c = 4 k2 + u -> 16 k2 + 4 q + u)
__device__ __forceinline__ int xy_gcol(int c, int q) { return 16 * (c >> 2) + 4 * q + (c & 3); }

// x * (-i)^M
template <int M>
__device__ __forceinline__ cpx<float> rot_negi(cpx<float> x) {
  constexpr int m = M & 3;
  if constexpr (m == 0) return x;
  else if constexpr (m == 1) return {x.y, -x.x};
  else if constexpr (m == 2) return {-x.x, -x.y};
  else return {-x.y, x.x};
}

// Outputs 4 G .. 4 G + 3 of the 16-point DFT of v[] into o[0..3]: the second
// half of dft16 evaluated for one output quarter only,
// X[k1 + 4 G] = sum_n1 Y'[n1][k1] (-i)^{n1 G}.
template <int G>
__device__ __forceinline__ void dft16_quarter(cpx<float>* v, cpx<float>* o) {
  dft16_head(v);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
    o[k1] = cadd(cadd(v[4 * k1], rot_negi<G>(v[4 * k1 + 1])),
                 cadd(rot_negi<2 * G>(v[4 * k1 + 2]), rot_negi<3 * G>(v[4 * k1 + 3])));
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

// First-stage input of lane t for row y: complex 2t and 2t + 1 of the
// 16-point groups n2 = N0 .. N0 + N - 1, in 16-byte loads.
template <int N0, int N>
__device__ __forceinline__ void load_rows(const cpx<float>* __restrict__ in, int y, int t, float4* p) {
#pragma unroll
  for (int i = 0; i < N; ++i)
    p[i] = *reinterpret_cast<const float4*>(in + (long long)y * XY_L + 16 * (N0 + i) + 2 * t);
}

// First stage for one chunk of rows: the 16-point DFTs over n2 of n1 = 2t
// and 2t + 1, four outputs k1 = 4Q + u each, into the LDS granules.
template <int Q, bool INV>
__device__ __forceinline__ void first_stage(const PlaneArgs& a, const float4* p, unsigned long long plane, int y,
                                            int t, float4* gbuf) {
  cpx<float> x0[16], x1[16];
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) {
    cpx<float> v0 = {p[n2].x, p[n2].y}, v1 = {p[n2].z, p[n2].w};
    if (a.w_load) {
      const unsigned long long g = plane * a.plane_stride + (unsigned long long)(y * XY_L + 16 * n2 + 2 * t);
      v0 = cscale(v0, a.w_load[g & a.w_mask]);
      v1 = cscale(v1, a.w_load[(g + 1) & a.w_mask]);
    }
    x0[n2] = INV ? cswap(v0) : v0;
    x1[n2] = INV ? cswap(v1) : v1;
  }
  cpx<float> o0[4], o1[4];
  dft16_quarter<Q>(x0, o0);  // Y[2t][4Q + u]
  dft16_quarter<Q>(x1, o1);  // Y[2t + 1][4Q + u]
#pragma unroll
  for (int u = 0; u < 4; ++u)
    gbuf[xg_idx(y, 8 * u + t)] = make_float4(o0[u].x, o0[u].y, o1[u].x, o1[u].y);
}

// Row transforms of the whole plane for quarter Q; leaves the quarter's 64
// columns in buf (column layout).  INV: inverse transform as
// swap(DFT(swap(x))), the swaps being register renames.
template <int Q, bool INV>
__device__ __forceinline__ void row_stages(const PlaneArgs& a, const cpx<float>* __restrict__ in,
                                           unsigned long long plane, cpx<float>* buf) {
  const int tid = threadIdx.x;
  float4* gbuf = reinterpret_cast<float4*>(buf);
  // ---- first stage: 8 lanes per row, 128 rows per chunk
  {
    const int t = tid & (XY_TPR - 1), rl = tid / XY_TPR;
#pragma unroll 1
    for (int ch = 0; ch < XY_L / XY_ROWS; ++ch) {
      float4 p[16];
      load_rows<0, 16>(in, ch * XY_ROWS + rl, t, p);
      first_stage<Q, INV>(a, p, plane, ch * XY_ROWS + rl, t, gbuf);
    }
  }
  __syncthreads();
  // ---- second stage: lane (y, u) runs X[k1 + 16 k2] = DFT16_n1(Y[n1][k1] W256^{n1 k1}), k1 = 4Q + u
  const int y = tid >> 2, u = tid & 3;
  cpx<float> v[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float4 p = gbuf[xg_idx(y, 8 * u + i)];
    v[2 * i] = {p.x, p.y};
    v[2 * i + 1] = {p.z, p.w};
  }
  twiddle256(v, 4 * Q + u, reinterpret_cast<const cpx<float>*>(a.tw_x));
  dft16(v);
  __syncthreads();  // every first-stage granule has been read
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) buf[cb_idx(y, 4 * k2 + u)] = v[k2];
}

template <bool INV>
__global__ void __launch_bounds__(XY_NT) k_fft_xy256(const PlaneArgs a) {
  __shared__ __align__(16) cpx<float> buf[XY_L * XY_QC];
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
  const cpx<float>* __restrict__ tw = reinterpret_cast<const cpx<float>*>(a.tw_x);
  const int tid = threadIdx.x;

  switch (q) {  // uniform: the quarter is a compile-time constant inside
    case 0: row_stages<0, INV>(a, in, plane, buf); break;
    case 1: row_stages<1, INV>(a, in, plane, buf); break;
    case 2: row_stages<2, INV>(a, in, plane, buf); break;
    default: row_stages<3, INV>(a, in, plane, buf); break;
  }
  __syncthreads();

  // ---- y: 64 columns x 256 points; thread (column c, butterfly tau)
  const int c = tid & (XY_QC - 1), tau = tid >> 6;
  cpx<float> v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = buf[cb_idx(tau + 16 * r, c)];
  dft16(v);
  __syncthreads();  // every thread read its column inputs
#pragma unroll
  for (int r = 0; r < 16; ++r) buf[cb_idx(16 * tau + r, c)] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = buf[cb_idx(tau + 16 * r, c)];
  twiddle256(v, tau, tw);
  dft16(v);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    cpx<float> x = INV ? cswap(v[r]) : v[r];
    const long long e = (long long)(tau + 16 * r) * XY_L + xy_gcol(c, (int)q);
    if (a.w_store) x = cscale(x, a.w_store[(unsigned long long)(plane * a.plane_stride + e) & a.w_mask]);
    out[e] = x;
  }
}

int launch_fft_xy256(const PlaneArgs& a, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // 4 blocks per plane, planes rounded up to a multiple of 8 (one per XCD)
  const unsigned long long planes8 = (a.nplanes + 7ull) / 8ull * 8ull;
  const unsigned long long blocks = planes8 * XY_Q;
  if (blocks > 0xFFFFFFFFull / XY_NT) return -1;  // grid size in work-items is 32-bit
  if (a.inverse) hipLaunchKernelGGL(k_fft_xy256<true>, dim3((unsigned)blocks), dim3(XY_NT), 0, s, a);
  else hipLaunchKernelGGL(k_fft_xy256<false>, dim3((unsigned)blocks), dim3(XY_NT), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fftamd

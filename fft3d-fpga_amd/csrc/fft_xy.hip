// fft_xy.hip -- fused x+y pass over 256 x 256 planes: the 2D transform of
// every plane of an array in one pass over HBM.
//
// A plane is 512 KiB, more than one CU's LDS, so four workgroups share it,
// each owning 64 of the 256 output rows.  Column transform first: 256 =
// 16 x 16, y = n1 + 16 n2, ky = k1 + 16 k2, and workgroup q owns the rows with
// k1 in [4q, 4q+4) (ky = 16 k2 + 4q + u).  It reads the whole plane in
// 512-byte row segments, evaluates only the four k1 outputs of each
// first-stage 16-point DFT and only its four second-stage DFTs per column, so
// no arithmetic is repeated across the four workgroups.  Its 64 rows stay in
// LDS, where it runs their 256-point row transforms (x) and writes them as
// whole output rows.  The four workgroups of a plane are mapped to one XCD and
// dispatched back to back (workgroups go to XCDs round-robin by index), so the
// plane comes from HBM once and from that XCD's L2 for the other three.  A 3D
// transform then takes two passes over HBM (this one and the z pass) instead
// of three.
//
// The pass reads the whole input plane and writes only its rows of the
// output, so it runs out of place.
#include <hip/hip_runtime.h>

#include "fft_pass_dev.h"
#include "fft_plane.h"

namespace fftamd {

constexpr int XY_L = 256;
constexpr int XY_Q = 4;             // workgroups per plane
constexpr int XY_QR = XY_L / XY_Q;  // output rows per workgroup
constexpr int XY_NT = 1024;         // threads per workgroup
using RowGeom = Geom<XY_L, MODE_CONTIG, XY_QR>;  // the workgroup's 64 rows in LDS
static_assert(RowGeom::NT == XY_NT, "16 threads per row, 64 rows");

// LDS buffer, in turn:
//  (1) first-stage column results Y[n1][u][x] at (4 n1 + u) * 256 + x (16 x 4
//      x 256 points; a wave reads or writes 64 consecutive x);
//  (2) the 64 rows, RowGeom layout [row][x + x/16] (row r = 4 k2 + u holds
//      output row ky = 16 k2 + 4q + u).
constexpr int XY_LDS = RowGeom::LDS_ELEMS > 16 * 4 * XY_L ? RowGeom::LDS_ELEMS : 16 * 4 * XY_L;

// output row of local row r of workgroup q
__device__ __forceinline__ int xy_grow(int r, int q) { return 16 * (r >> 2) + 4 * q + (r & 3); }

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

// Column transforms of the whole plane for row quarter Q; leaves the
// quarter's 64 rows in buf (RowGeom layout).  INV: inverse transform as
// swap(DFT(swap(x))), the swaps being register renames.
template <int Q, bool INV>
__device__ __forceinline__ void column_stages(const PlaneArgs& a, const cpx<float>* __restrict__ in,
                                              unsigned long long plane, cpx<float>* buf) {
  const int tid = threadIdx.x;
  const int x = tid & (XY_L - 1), g = tid >> 8;  // column, group of four n1 (uniform per wave)
  // ---- first stage: for n1 = 4g .. 4g+3, the 16-point DFT over n2 of
  // column x at rows n1 + 16 n2, four outputs k1 = 4Q + u
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int n1 = 4 * g + m;
    cpx<float> v[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const int e = (n1 + 16 * n2) * XY_L + x;
      cpx<float> s = in[e];
      if (a.w_load) s = cscale(s, a.w_load[(unsigned long long)(plane * a.plane_stride + e) & a.w_mask]);
      v[n2] = INV ? cswap(s) : s;
    }
    cpx<float> o[4];
    dft16_quarter<Q>(v, o);
#pragma unroll
    for (int u = 0; u < 4; ++u) buf[(4 * n1 + u) * XY_L + x] = o[u];
  }
  __syncthreads();
  // ---- second stage: lane (x, u = g): X[k1 + 16 k2] = DFT16_n1(Y[n1][k1] W256^{n1 k1}), k1 = 4Q + u
  const int u = g;
  cpx<float> v[16];
#pragma unroll
  for (int n1 = 0; n1 < 16; ++n1) v[n1] = buf[(4 * n1 + u) * XY_L + x];
  twiddle256(v, 4 * Q + u, reinterpret_cast<const cpx<float>*>(a.tw_y));
  dft16(v);
  __syncthreads();  // every first-stage value has been read
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) buf[RowGeom::idx(4 * k2 + u, x)] = v[k2];
}

template <bool INV>
__global__ void __launch_bounds__(XY_NT) k_fft_xy256(const PlaneArgs a) {
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
    case 0: column_stages<0, INV>(a, in, plane, buf); break;
    case 1: column_stages<1, INV>(a, in, plane, buf); break;
    case 2: column_stages<2, INV>(a, in, plane, buf); break;
    default: column_stages<3, INV>(a, in, plane, buf); break;
  }
  __syncthreads();

  // ---- rows: thread (local row c, tau), the line passes' stages on LDS rows
  const int tid = threadIdx.x;
  const int c = tid / RowGeom::TPL, tau = tid % RowGeom::TPL;
  cpx<float> v[16];
  stage_load_lds<float, XY_L, MODE_CONTIG, 0, XY_QR>(v, buf, c, tau);
  __syncthreads();  // every thread holds its stage-0 inputs before stage 0 stores
  run_stages<float, XY_L, MODE_CONTIG, 0, false, XY_QR>(v, buf, c, tau,
                                                        reinterpret_cast<const cpx<float>*>(a.tw_x));
  // last stage: point kx = tau + 16 r of the row; 16 lanes write 128 contiguous bytes
  static_assert(sched_np(XY_L) == 2 && RowGeom::TPL == 16, "256 = 16 x 16, one butterfly per lane");
  const long long row = (long long)xy_grow(c, (int)q) * XY_L;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    cpx<float> s = INV ? cswap(v[r]) : v[r];
    const long long e = row + tau + 16 * r;
    if (a.w_store) s = cscale(s, a.w_store[(unsigned long long)(plane * a.plane_stride + e) & a.w_mask]);
    out[e] = s;
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

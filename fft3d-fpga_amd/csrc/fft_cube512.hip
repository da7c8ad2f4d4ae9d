// fft_cube512.hip -- 512 x 512 x 512 blocks in two passes over HBM with the y
// axis split 8 x 64 between them (the 256^3 and 128^3 pairs split it 4 x n/4,
// fft_cube.hip).  Three line passes would move the array through HBM three
// times; here every point is read and written twice.
//
// y = a + 8 b (a < 8, b < 64), ky = k + 64 d (k < 64, d < 8):
//   X[k + 64 d] = sum_a W8^{a d} W512^{a k} T_a[k],
//   T_a[k]      = sum_b x[a + 8 b] W64^{b k}.
//
// Any split of the three axes into two passes needs a tile of 2^15 points on
// one side and 2^16 on the other once the second pass keeps whole 128-byte
// row segments (9 bits of x in the first pass, 4 of x plus all 9 of z in the
// second: 27 + 4 bits over two passes).  Both tiles here are 2^15 points
// (256 KiB): one workgroup of 1024 threads per CU holds a tile in registers
// (32 points per thread) and exchanges half of it at a time through LDS; the
// second pass takes 8 columns (64-byte row segments) and the two halves of
// each 128-byte segment are consecutive tiles of one XCD, so the L2 sees
// whole lines.
//
// Pass A (one workgroup per plane and residue a, out of place): the 64 rows
// y = a + 8 b, each transformed along x (radix 32 in registers, then radix 16
// through LDS in two rounds of 16 sub-transforms), then the 512 columns along
// b (64 = 4 x 16, two column halves through LDS); writes T_a[k] (without its
// twiddle) to row 64 a + k.
//
// Pass B (one workgroup per (block, k, 8 columns), in place): reads rows
// k + 64 a, a = 0..7, of its columns in all 512 planes (a thread holds the
// eight rows of one column in four planes), multiplies row k + 64 a by
// W512^{a k}, forms X[k + 64 d] with an 8-point DFT over a in registers, then
// z as 512 = 4 x 128: the 4-point DFT over the thread's planes in registers,
// the twiddle W512^{zs kq}, and 128-point lines (8 x 16) through LDS, half of
// the rows d at a time.
// Inverse: swap(DFT(swap(x))), the swaps at pass A's loads and pass B's stores.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "fft_pass_dev.h"
#include "fft_plane.h"

namespace fftamd {

constexpr int C5_L = 512;
constexpr int C5_NT = 1024;
constexpr long long C5_PLANE = (long long)C5_L * C5_L;
constexpr long long C5_BLOCK = C5_PLANE * C5_L;

// pass A, x rounds: [row 64][sub-transform 16][tau 16 + 1 pad]
constexpr int C5X_K = 17;
constexpr int C5X_LDS = 64 * 16 * C5X_K;
using C5Col = Geom<64, MODE_STRIDED, 256>;   // pass A columns: 256 columns of 64, 4 threads each
using C5Z = Geom<128, MODE_STRIDED, 128>;    // pass B: 128 lines of 128 points, 8 threads each
static_assert(C5Col::NT == C5_NT && C5Z::NT == C5_NT, "1024 threads");
static_assert(sched_np(64) == 2 && sched_radix(64, 0) == 4, "64 = 4 x 16");
static_assert(sched_np(128) == 2 && sched_radix(128, 0) == 8, "128 = 8 x 16");
constexpr int C5A_LDS = C5X_LDS > C5Col::LDS_ELEMS ? C5X_LDS : C5Col::LDS_ELEMS;

// One workgroup per CU: 1024 threads (128 VGPRs each), <= 148 KiB of LDS.
template <bool INV, bool WL, bool SPLIT>
__global__ void __launch_bounds__(C5_NT, 1) k_fft_c512_a(const CubeArgs a) {
  __shared__ cpx<float> buf[C5A_LDS];
  __shared__ cpx<float> w512[C5_L];  // W512^m
  __shared__ split_half8 atab[SPLIT ? SPLIT_ATAB : 1];  // split DFT operands (read after a barrier)
  const unsigned long long plane = blockIdx.x >> 3;
  const int res = blockIdx.x & 7;  // residue class a of the input rows
  if (plane >= a.nplanes) return;  // uniform over the workgroup
  const long long base = (long long)plane * C5_PLANE;
  const cpx<float>* in = reinterpret_cast<const cpx<float>*>(a.in) + base;
  cpx<float>* out = reinterpret_cast<cpx<float>*>(a.out) + base;
  const cpx<float>* __restrict__ wdir = reinterpret_cast<const cpx<float>*>(a.w512);
  const cpx<float>* __restrict__ tw64 = reinterpret_cast<const cpx<float>*>(a.tw);
  const int tid = threadIdx.x;
  if (tid < C5_L) w512[tid] = wdir[tid];  // read after the first barrier
  if constexpr (SPLIT) split_atab_fill(atab);

  // ---- x: thread (c, tau) transforms row b (y = res + 8 b); stage 0 is the
  // radix-32 DFT of points tau + 16 r straight from HBM.  b = c rotated by
  // 8 (plane mod 8): the workgroups in flight together (consecutive planes)
  // then read different rows at the same moment instead of the same row
  // offset of every plane (measured on this access pattern alone: 382 against
  // 455 us per 1 GiB)
  const int c = tid >> 4, tau = tid & 15;
  const int b = (c + 8 * (int)(plane & 7)) & 63;
  cpx<float> v[32];
  {
    const long long row = (long long)(res + 8 * b) * C5_L;
#pragma unroll
    for (int r = 0; r < 32; ++r) {
      const long long e = row + tau + 16 * r;
      cpx<float> s = in[e];
      if (WL) s = cscale(s, a.w_load[(unsigned long long)(base + e) & a.w_mask]);
      v[r] = INV ? cswap(s) : s;
    }
  }
  dft32(v);  // v[k1] = Y_tau[k1]
  // stage 1, radix 16 over tau with twiddles W512^{tau k1}: sub-transforms
  // k1 = 16 R + tau' in round R; thread (c, tau') ends with X[k1 + 32 m] in o[R][m]
  cpx<float> o[2][16];
#pragma unroll
  for (int R = 0; R < 2; ++R) {
    __syncthreads();  // R = 0: the W512 table is written; R = 1: round 0 has read the buffer
#pragma unroll
    for (int k = 0; k < 16; ++k) buf[(c * 16 + k) * C5X_K + tau] = v[16 * R + k];
    __syncthreads();
    const int k1 = 16 * R + tau;
    cpx<float>* y = o[R];
#pragma unroll
    for (int t = 0; t < 16; ++t) y[t] = buf[(c * 16 + tau) * C5X_K + t];
#pragma unroll
    for (int t = 1; t < 16; ++t) y[t] = cmul(y[t], w512[(k1 * t) & (C5_L - 1)]);
    wg_dft16<SPLIT, true>(y, atab);
  }

  // ---- y over b: half R holds columns kx = 16 R + tau + 32 m (column index
  // cc = tau + 16 m in LDS); thread (cc, tq) ends with point k = tq + 4 r
  const int cc = tid & 255;
  const int tq = __builtin_amdgcn_readfirstlane(tid >> 8);  // uniform over each wave
#pragma unroll
  for (int R = 0; R < 2; ++R) {
    __syncthreads();  // the buffer's previous contents have been read
#pragma unroll
    for (int m = 0; m < 16; ++m) buf[C5Col::idx(tau + 16 * m, b)] = o[R][m];
    __syncthreads();
    cpx<float> u[16];
    stage_load_lds<float, 64, MODE_STRIDED, 0, 256>(u, buf, cc, tq);
    stage_compute<float, 64, MODE_STRIDED, 0, false, 256>(u, tq, nullptr);  // four radix-4, no twiddles
    __syncthreads();  // every thread holds its stage-0 inputs
    stage_store_lds<float, 64, MODE_STRIDED, 0, 256>(u, buf, cc, tq);
    __syncthreads();
    stage_load_lds<float, 64, MODE_STRIDED, 1, 256>(u, buf, cc, tq);
    stage_compute<float, 64, MODE_STRIDED, 1, SPLIT, 256, SPLIT>(u, tq, tw64, buf, atab);  // radix 16, W64 twiddles
    // column DFT point k = tq + 4 r to row 64 res + k (pass B applies W512^{res k})
    const int kx = 16 * R + (cc & 15) + 32 * (cc >> 4);
#pragma unroll
    for (int r = 0; r < 16; ++r) out[(long long)(64 * res + tq + 4 * r) * C5_L + kx] = u[r];
  }
}

template <bool INV, bool WS, bool SPLIT>
__global__ void __launch_bounds__(C5_NT, 1) k_fft_c512_b(const CubeArgs a) {
  __shared__ cpx<float> buf[C5Z::LDS_ELEMS];
  __shared__ cpx<float> w512[C5_L];
  __shared__ split_half8 atab[SPLIT ? SPLIT_ATAB : 1];  // split DFT operands (read after a barrier)
  // the two 8-column groups of one 128-byte row segment are blocks b and
  // b + 8: consecutive tiles of one XCD (workgroups go to XCDs round-robin).
  // Every tile's own addresses differ only in multiples of 256 KiB, so the
  // pair index P interleaves k and the column pair: bits 0-3 and 7-8 of P
  // give k, bits 4-6 and 9-10 the pair, and the 128 pairs in flight together
  // (one workgroup per CU) cover 16 rows k x 8 column pairs (address bits
  // 8-15) instead of 4 rows x all 32 pairs (measured on this access pattern
  // alone: 499 against 563 us per 1 GiB; all 64 rows x 2 pairs: 510-520 us)
  const unsigned b = blockIdx.x;
  const unsigned P = (b >> 4) * 8 + (b & 7);  // pair index
  const int k = (int)((P & 15u) | (((P >> 7) & 3u) << 4));                             // rows k + 64 a
  const int xg = 2 * (int)(((P >> 4) & 7u) | (((P >> 9) & 3u) << 3)) + (int)((b >> 3) & 1);  // 8 columns
  const unsigned long long blk = P >> 11;
  if (blk >= a.nplanes / C5_L) return;  // uniform over the workgroup
  const long long base = (long long)blk * C5_BLOCK + (long long)k * C5_L + 8 * xg;
  const cpx<float>* in = reinterpret_cast<const cpx<float>*>(a.in) + base;
  cpx<float>* out = reinterpret_cast<cpx<float>*>(a.out) + base;
  const cpx<float>* __restrict__ wdir = reinterpret_cast<const cpx<float>*>(a.w512);
  const cpx<float>* __restrict__ tw128 = reinterpret_cast<const cpx<float>*>(a.tw);
  const int tid = threadIdx.x;
  if (tid < C5_L) w512[tid] = wdir[tid];  // read after the first barrier
  if constexpr (SPLIT) split_atab_fill(atab);

  // ---- y over a: thread (x, zs) holds planes z = zs + 128 j; t[j][d] = X[k + 64 d]
  const int x = tid & 7, zs = tid >> 3;
  cpx<float> t[4][8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long long zo = (long long)(zs + 128 * j) * C5_PLANE + x;
#pragma unroll
    for (int aa = 0; aa < 8; ++aa) t[j][aa] = in[zo + 64 * C5_L * aa];
  }
  {
    // T_a[k] times W512^{a k} (k is uniform over the workgroup)
    cpx<float> wa[8];
#pragma unroll
    for (int aa = 1; aa < 8; ++aa) wa[aa] = wdir[aa * k];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int aa = 1; aa < 8; ++aa) t[j][aa] = cmul(t[j][aa], wa[aa]);
      dft8(t[j]);
    }
  }
  __syncthreads();  // the W512 table is written
  // ---- z = zs + 128 j: the 4-point DFT over j in registers, then W512^{zs kq};
  // t[kq][d] then holds the input zs of line (d, kq, x) of 128 points
  {
    const cpx<float> wz1 = w512[zs], wz2 = w512[2 * zs], wz3 = w512[3 * zs];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      cpx<float> q[4] = {t[0][d], t[1][d], t[2][d], t[3][d]};
      dft4(q);
      t[0][d] = q[0];
      t[1][d] = cmul(q[1], wz1);
      t[2][d] = cmul(q[2], wz2);
      t[3][d] = cmul(q[3], wz3);
    }
  }
  // 128-point lines over zs, rows d = 4 hh .. 4 hh + 3 at a time: line
  // cz = 32 dd + 8 kq + x; thread (cz, tz) ends with point m = tz + 8 r,
  // i.e. kz = kq + 4 m
  const int cz = tid & 127;
  const int tz = __builtin_amdgcn_readfirstlane(tid >> 7);  // uniform over each wave
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    if (hh) __syncthreads();  // the first half's last stage has read the buffer
#pragma unroll
    for (int dd = 0; dd < 4; ++dd)
#pragma unroll
      for (int kq = 0; kq < 4; ++kq) buf[C5Z::idx(32 * dd + 8 * kq + x, zs)] = t[kq][4 * hh + dd];
    __syncthreads();
    cpx<float> v[16];
    stage_load_lds<float, 128, MODE_STRIDED, 0, 128>(v, buf, cz, tz);
    stage_compute<float, 128, MODE_STRIDED, 0, false, 128>(v, tz, nullptr);  // two radix-8, no twiddles
    __syncthreads();  // every thread holds its stage-0 inputs
    stage_store_lds<float, 128, MODE_STRIDED, 0, 128>(v, buf, cz, tz);
    __syncthreads();
    stage_load_lds<float, 128, MODE_STRIDED, 1, 128>(v, buf, cz, tz);
    stage_compute<float, 128, MODE_STRIDED, 1, SPLIT, 128, SPLIT>(v, tz, tw128, buf, atab);  // radix 16, W128 twiddles
    const int xx = cz & 7, kq = (cz >> 3) & 3, dd = cz >> 5;
    const long long ob = (long long)(64 * (4 * hh + dd)) * C5_L + xx;  // row k + 64 d
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long long o = ob + (long long)(kq + 4 * (tz + 8 * r)) * C5_PLANE;
      cpx<float> s = INV ? cswap(v[r]) : v[r];
      if (WS) s = cscale(s, a.w_store[(unsigned long long)(base + o) & a.w_mask]);
      out[o] = s;
    }
  }
}

void preload_c512() { preload_kernel(reinterpret_cast<const void*>(&k_fft_c512_a<false, false, false>)); }

// KER<inverse, weights, split>; the split-precision variants take no weights
#define C512_DISPATCH(KER, W, G, T, S, A)                                  \
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

int launch_fft_c512_a(const CubeArgs& a, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (a.n != C5_L || a.nplanes == 0 || a.w_store || a.tab || !a.w512 || !a.tw || (a.split && a.w_load)) return -1;
  const unsigned long long blocks = a.nplanes * 8ull;
  if (blocks > 0xFFFFFFFFull / C5_NT) return -1;  // grid size in work-items is 32-bit
  const dim3 g((unsigned)blocks), t(C5_NT);
  C512_DISPATCH(k_fft_c512_a, a.w_load, g, t, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_fft_c512_b(const CubeArgs& a, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (a.n != C5_L || a.nplanes == 0 || a.nplanes % C5_L || a.w_load || a.tab || !a.w512 || !a.tw ||
      (a.split && a.w_store))
    return -1;
  const unsigned long long blocks = (a.nplanes / C5_L) * 4096ull;  // 64 k x 64 column groups per block
  if (blocks > 0xFFFFFFFFull / C5_NT) return -1;
  const dim3 g((unsigned)blocks), t(C5_NT);
  C512_DISPATCH(k_fft_c512_b, a.w_store, g, t, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fftamd

// fft_device.h -- complex arithmetic and in-register radix-2/4/8/16 DFT
// butterflies (forward sign, e^{-2 pi i nk/R}) for the gfx950 pass kernels.
//
// The inverse transform is obtained as swap(DFT(swap(x))) with swap(a+ib) =
// b+ia, the identity the reference engine uses (kernels/common/fft_8.cl:276-278,
// 341-343); a swap is a register rename, so one butterfly set serves both
// directions.
#pragma once
#include <hip/hip_runtime.h>

namespace fftamd {

template <typename R>
struct __attribute__((aligned(2 * sizeof(R)))) cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i
template <typename R>
__device__ __forceinline__ cpx<R> cmnegi(cpx<R> a) { return {a.y, -a.x}; }
template <typename R>
__device__ __forceinline__ cpx<R> cswap(cpx<R> a) { return {a.y, a.x}; }
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }

template <typename R> struct K {
  static constexpr R C8 = R(0.70710678118654752440084436210484904);   // cos(pi/4)
  static constexpr R C16 = R(0.92387953251128675612818318939678829);  // cos(pi/8)
  static constexpr R S16 = R(0.38268343236508977172845998403039887);  // sin(pi/8)
};

// (a+ib) * W8^1 = (a+ib)(c - ic), c = sqrt(2)/2
template <typename R>
__device__ __forceinline__ cpx<R> mul_w8_1(cpx<R> a) {
  return {K<R>::C8 * (a.x + a.y), K<R>::C8 * (a.y - a.x)};
}
// (a+ib) * W8^3 = (a+ib)(-c - ic)
template <typename R>
__device__ __forceinline__ cpx<R> mul_w8_3(cpx<R> a) {
  return {K<R>::C8 * (a.y - a.x), -K<R>::C8 * (a.x + a.y)};
}

template <typename R>
__device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

// in-place 4-point DFT on v[0], v[s], v[2s], v[3s]
template <typename R, int S = 1>
__device__ __forceinline__ void dft4s(cpx<R>* v) {
  cpx<R> t0 = cadd(v[0], v[2 * S]), t1 = csub(v[0], v[2 * S]);
  cpx<R> t2 = cadd(v[S], v[3 * S]), t3 = cmnegi(csub(v[S], v[3 * S]));
  v[0] = cadd(t0, t2);
  v[2 * S] = csub(t0, t2);
  v[S] = cadd(t1, t3);
  v[3 * S] = csub(t1, t3);
}

template <typename R>
__device__ __forceinline__ void dft4(cpx<R>* v) { dft4s<R, 1>(v); }

template <typename R>
__device__ __forceinline__ void dft8(cpx<R>* v) {
  // even/odd split: E = DFT4(x0,x2,x4,x6), O = DFT4(x1,x3,x5,x7)
  cpx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e);
  dft4(o);
  o[1] = mul_w8_1(o[1]);
  o[2] = cmnegi(o[2]);
  o[3] = mul_w8_3(o[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

// 16-point DFT as 4 x 4: n = 4*n2 + n1, k = k1 + 4*k2.
//   Y[n1][k1] = DFT4_{n2}(x[4 n2 + n1]);  Y'[n1][k1] = Y[n1][k1] W16^{n1 k1}
//   X[k1 + 4 k2] = DFT4_{n1}(Y'[n1][k1])
// First half of dft16: leaves Y'[n1][k1] in v[n1 + 4 k1].
template <typename R>
__device__ __forceinline__ void dft16_head(cpx<R>* v) {
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) dft4s<R, 4>(v + n1);  // v[n1 + 4 k1] = Y[n1][k1]
  const cpx<R> w1 = {K<R>::C16, -K<R>::S16};
  const cpx<R> w3 = {K<R>::S16, -K<R>::C16};
  const cpx<R> w9 = {-K<R>::C16, K<R>::S16};
  // n1 = 1: k1 = 1,2,3 -> W^1, W^2, W^3
  v[5] = cmul(v[5], w1);
  v[9] = mul_w8_1(v[9]);
  v[13] = cmul(v[13], w3);
  // n1 = 2: W^2, W^4, W^6
  v[6] = mul_w8_1(v[6]);
  v[10] = cmnegi(v[10]);
  v[14] = mul_w8_3(v[14]);
  // n1 = 3: W^3, W^6, W^9
  v[7] = cmul(v[7], w3);
  v[11] = mul_w8_3(v[11]);
  v[15] = cmul(v[15], w9);
}

template <typename R>
__device__ __forceinline__ void dft16(cpx<R>* v) {
  dft16_head(v);
  cpx<R> t[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<R> q[4] = {v[4 * k1 + 0], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]};
    dft4(q);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) t[k1 + 4 * k2] = q[k2];
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = t[i];
}

// 32-point DFT as 2 x 16: X[k + 16 f] = E[k] + (-1)^f W32^k O[k] with E, O the
// 16-point DFTs of the even and odd points.  W32^k = cos(pi k / 16) - i
// sin(pi k / 16); the odd k use cos and sin of pi/16 and 3 pi/16.
template <typename R>
__device__ __forceinline__ void dft32(cpx<R>* v) {
  cpx<R> e[16], o[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft16(e);
  dft16(o);
  constexpr R c1 = R(0.98078528040323044912618223613424268L);  // cos(pi/16)
  constexpr R s1 = R(0.19509032201612826784828486847702224L);  // sin(pi/16)
  constexpr R c3 = R(0.83146961230254523707878837761790576L);  // cos(3 pi/16)
  constexpr R s3 = R(0.55557023301960222474283081394853287L);  // sin(3 pi/16)
  o[1] = cmul(o[1], cpx<R>{c1, -s1});
  o[2] = cmul(o[2], cpx<R>{K<R>::C16, -K<R>::S16});
  o[3] = cmul(o[3], cpx<R>{c3, -s3});
  o[4] = mul_w8_1(o[4]);
  o[5] = cmul(o[5], cpx<R>{s3, -c3});
  o[6] = cmul(o[6], cpx<R>{K<R>::S16, -K<R>::C16});
  o[7] = cmul(o[7], cpx<R>{s1, -c1});
  o[8] = cmnegi(o[8]);
  o[9] = cmul(o[9], cpx<R>{-s1, -c1});
  o[10] = cmul(o[10], cpx<R>{-K<R>::S16, -K<R>::C16});
  o[11] = cmul(o[11], cpx<R>{-s3, -c3});
  o[12] = mul_w8_3(o[12]);
  o[13] = cmul(o[13], cpx<R>{-c3, -s3});
  o[14] = cmul(o[14], cpx<R>{-K<R>::C16, -K<R>::S16});
  o[15] = cmul(o[15], cpx<R>{-c1, -s1});
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 16] = csub(e[k], o[k]);
  }
}

template <typename R, int RAD>
__device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (RAD == 2) dft2(v);
  else if constexpr (RAD == 4) dft4(v);
  else if constexpr (RAD == 8) dft8(v);
  else if constexpr (RAD == 16) dft16(v);
  else static_assert(RAD == 1, "unsupported radix");
}

__device__ __forceinline__ unsigned brev_bits(unsigned x, unsigned bits) {
  return bits ? (__builtin_bitreverse32(x) >> (32u - bits)) : 0u;
}

}  // namespace fftamd

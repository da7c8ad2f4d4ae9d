// fft_split.h -- split-precision radix-16 butterflies on gfx950 FP16 matrix
// cores (v_mfma_f32_16x16x32_f16), FP32 accumulation.
//
// A radix-16 stage of a wave is 64 independent 16-point DFTs (one per lane).
// They are computed 16 at a time as real matrix products
//   [Yr] = [ C  S] [Xr]      C[i][t] = cos(2 pi i t / 16), S = sin(...)
//   [Yi]   [-S  C] [Xi]
// i.e. D(16x16) = A(16x32) B(32x16) with B holding 16 vectors as columns.
// FP16 carries 11 significant bits, so every operand is split as
// x = hi + lo (hi = fp16(x), lo = fp16(x - hi)) and the product is
// A_hi B_hi + A_hi B_lo + A_lo B_hi (the dropped A_lo B_lo term is ~2^-23
// relative): 6 MFMAs per 16 vectors.  Before splitting, each vector (a column
// of B, and of the product) is scaled by a power of two so its largest
// component lies in [2^14, 2^15): FP16 cannot overflow, the scale is exact,
// and a vector's precision does not depend on the other vectors' magnitudes.
//
// Operand layout of v_mfma_f32_16x16x32_f16 (checked on the device with an
// integer-valued probe): lane l holds A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16]
// for j = 0..7, and receives D[4(l/16)+e][l%16] for e = 0..3.
//
// Data moves between the "one vector per lane" form used by the radix stages
// and the operand form in registers: block transposes across the wave's four
// 16-lane rows with gfx950's v_permlane32_swap / v_permlane16_swap, so the
// stages need no LDS scratch and no barrier.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_device.h"

namespace fftamd {

typedef _Float16 split_half8 __attribute__((ext_vector_type(8)));
typedef float split_float4 __attribute__((ext_vector_type(4)));

// cos(2 pi m / 16) for m mod 16, from the exact constants of the radix-16
// butterfly (selects only: no indexed local array, which would live in scratch)
__device__ __forceinline__ float cos16(int m) {
  m &= 15;
  // cos is even and cos(pi - x) = -cos(x)
  const int a = m <= 8 ? m : 16 - m;  // 0..8
  const int b = a <= 4 ? a : 8 - a;   // 0..4
  const float v = b == 0 ? 1.0f : b == 1 ? K<float>::C16 : b == 2 ? K<float>::C8 : b == 3 ? K<float>::S16 : 0.0f;
  return a <= 4 ? v : -v;
}

__device__ __forceinline__ void split2(float x, _Float16* hi, _Float16* lo) {
  const _Float16 h = (_Float16)x;
  *hi = h;
  *lo = (_Float16)(x - (float)h);
}

// The constant DFT operands of lane `lane`: A1 = [C | S], A2 = [-S | C],
// split hi/lo (the A layout of the product above).
__device__ __forceinline__ void split_a_operands(int lane, split_half8& a1h, split_half8& a1l, split_half8& a2h,
                                                 split_half8& a2l) {
  const int row = lane & 15, part = lane >> 4;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int t = 8 * (part & 1) + j;  // point index inside the re (part 0,1) or im (2,3) half
    const int m = row * t;
    const float c = cos16(m), s = cos16(m - 4);  // sin(x) = cos(x - pi/2)
    const float e1 = part < 2 ? c : s;
    const float e2 = part < 2 ? -s : c;
    _Float16 h, l;
    split2(e1, &h, &l);
    a1h[j] = h;
    a1l[j] = l;
    split2(e2, &h, &l);
    a2h[j] = h;
    a2l[j] = l;
  }
}

// A workgroup's table of those operands in LDS, SPLIT_ATAB entries: lane l's
// four at [4 l .. 4 l + 3].  Filled by the first wave; every reader must be
// behind a workgroup barrier that follows the fill.
constexpr int SPLIT_ATAB = 4 * 64;
__device__ __forceinline__ void split_atab_fill(split_half8* atab) {
  if (threadIdx.x < 64) {
    split_half8 a1h, a1l, a2h, a2l;
    split_a_operands((int)threadIdx.x, a1h, a1l, a2h, a2l);
    atab[4 * threadIdx.x] = a1h;
    atab[4 * threadIdx.x + 1] = a1l;
    atab[4 * threadIdx.x + 2] = a2h;
    atab[4 * threadIdx.x + 3] = a2l;
  }
}

// ---- operands moved between lanes in registers ---------------------------------
// A wave's four 16-lane rows exchange 4 x 4 blocks of registers with gfx950's
// v_permlane32_swap (the upper two rows of one VGPR with the lower two of
// another) and v_permlane16_swap (the odd rows of one with the even rows of
// another): with groups G0..G3 of W VGPRs, the swaps on (G0, G2), (G1, G3) and
// then (G0, G1), (G2, G3) leave in group g of row R what group R of row g held
// (the lane layout was checked on the device with integer data).
template <int W>
__device__ __forceinline__ void split_xpose(unsigned (&x)[4 * W]) {
#pragma unroll
  for (int e = 0; e < W; ++e) {
    const auto a = __builtin_amdgcn_permlane32_swap(x[e], x[2 * W + e], false, false);
    x[e] = a[0];
    x[2 * W + e] = a[1];
    const auto b = __builtin_amdgcn_permlane32_swap(x[W + e], x[3 * W + e], false, false);
    x[W + e] = b[0];
    x[3 * W + e] = b[1];
  }
#pragma unroll
  for (int e = 0; e < W; ++e) {
    const auto a = __builtin_amdgcn_permlane16_swap(x[e], x[W + e], false, false);
    x[e] = a[0];
    x[W + e] = a[1];
    const auto b = __builtin_amdgcn_permlane16_swap(x[2 * W + e], x[3 * W + e], false, false);
    x[2 * W + e] = b[0];
    x[3 * W + e] = b[1];
  }
}

typedef _Float16 split_half2 __attribute__((ext_vector_type(2)));
typedef unsigned split_u4 __attribute__((ext_vector_type(4)));

// Radix-16 DFT (forward sign) of the 16 points v[0..15] of every lane of the
// wave, in place: each lane scales its own vector and splits it into hi and
// lo FP16 halves, packed as the four 8-value B slices [re 0..7 | re 8..15 |
// im 0..7 | im 8..15]; one block transpose hands lane 16 R + n the slice R of
// vector 16 g + n for product g; the four products' results (4 re + 4 im
// points of one vector per lane) go back through one more transpose and are
// unscaled by the lane's own factor.  All 64 lanes must call it together (no
// lane inactive).  TAB: the constant operands are read from `atab` (an LDS
// table from split_atab_fill) at every call through a lane index the compiler
// cannot see through, so nothing of one call stays live until the next;
// without, they are formed in registers (and a kernel keeps them live across
// its calls).
template <bool TAB>
__device__ __forceinline__ void dft16_split_perm(cpx<float>* v, const split_half8* atab) {
  int lane = threadIdx.x & 63;
  if constexpr (TAB) asm volatile("" : "+v"(lane));
  float mx = 0.0f;
#pragma unroll
  for (int t = 0; t < 16; ++t) mx = fmaxf(mx, fmaxf(fabsf(v[t].x), fabsf(v[t].y)));
  int ex = 0;
  (void)frexpf(mx, &ex);  // mx = f * 2^ex, f in [0.5, 1)
  const int sh = mx > 0.0f ? 15 - ex : 0;
  const float sc = ldexpf(1.0f, sh);
  const float inv_s = ldexpf(1.0f, -sh);
  unsigned hb[16], lb[16];  // slice j in words 4 j .. 4 j + 3 (values 2 d, 2 d + 1 in word 4 j + d)
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int t = 8 * (j & 1) + 2 * d;
      const float x0 = (j < 2 ? v[t].x : v[t].y) * sc;
      const float x1 = (j < 2 ? v[t + 1].x : v[t + 1].y) * sc;
      // hi = fp16(x) for both values in one packed convert; the residuals
      // x - hi are formed from that packed word (made opaque, so that the
      // compiler does not convert each value a second time to get hi back)
      unsigned hp = __builtin_bit_cast(unsigned, split_half2{(_Float16)x0, (_Float16)x1});
      asm volatile("" : "+v"(hp));
      const split_half2 h = __builtin_bit_cast(split_half2, hp);
      const float r0 = x0 - (float)h[0];
      const float r1 = x1 - (float)h[1];
      hb[4 * j + d] = hp;
      lb[4 * j + d] = __builtin_bit_cast(unsigned, split_half2{(_Float16)r0, (_Float16)r1});
    }
  split_xpose<4>(hb);
  split_xpose<4>(lb);
  split_half8 a1h, a1l, a2h, a2l;
  if constexpr (TAB) {
    int li = lane;
    asm volatile("" : "+v"(li));
    a1h = atab[4 * li];
    a1l = atab[4 * li + 1];
    a2h = atab[4 * li + 2];
    a2l = atab[4 * li + 3];
  } else {
    split_a_operands(lane, a1h, a1l, a2h, a2l);
  }
  unsigned o[32];  // product g: words 8 g .. 8 g + 3 real, 8 g + 4 .. 8 g + 7 imaginary
  const split_float4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const split_half8 bh = __builtin_bit_cast(split_half8, split_u4{hb[4 * g], hb[4 * g + 1], hb[4 * g + 2], hb[4 * g + 3]});
    const split_half8 bl = __builtin_bit_cast(split_half8, split_u4{lb[4 * g], lb[4 * g + 1], lb[4 * g + 2], lb[4 * g + 3]});
    split_float4 r = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, bh, z, 0, 0, 0);
    split_float4 q = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2h, bh, z, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, bl, r, 0, 0, 0);
    q = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2h, bl, q, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1l, bh, r, 0, 0, 0);
    q = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2l, bh, q, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      // elements copied out first (a bit cast of a vector element expression
      // read element 0 in this compiler, found with a device probe)
      const float re = r[e], im = q[e];
      o[8 * g + e] = __builtin_bit_cast(unsigned, re);
      o[8 * g + 4 + e] = __builtin_bit_cast(unsigned, im);
    }
  }
  split_xpose<8>(o);  // now words 8 j .. 8 j + 7: points 4 j .. 4 j + 3 of this lane's vector
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int e = 0; e < 4; ++e)
      v[4 * j + e] = {__builtin_bit_cast(float, o[8 * j + e]) * inv_s, __builtin_bit_cast(float, o[8 * j + 4 + e]) * inv_s};
}

// Radix-16 DFT of v[0..15] in a workgroup kernel: with SPLIT on the matrix
// cores (every thread of the wave must make the call; TAB: the constant
// operands come from the LDS table `atab`), without, the FP32 butterfly.
template <bool SPLIT, bool TAB = false>
__device__ __forceinline__ void wg_dft16(cpx<float>* v, const split_half8* atab = nullptr) {
  if constexpr (SPLIT) {
    dft16_split_perm<TAB>(v, atab);
  } else {
    dft16(v);
  }
}

}  // namespace fftamd

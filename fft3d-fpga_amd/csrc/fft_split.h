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
// and the operand form through a per-wave LDS scratch of 64 rows x 36 floats
// (32 values + 4 pad floats so the 16 lanes reading one operand slice hit
// distinct banks).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_device.h"

namespace fftamd {

typedef _Float16 split_half8 __attribute__((ext_vector_type(8)));
typedef float split_float4 __attribute__((ext_vector_type(4)));

constexpr int SPLIT_ROW = 36;                      // floats per scratch row
constexpr int SPLIT_WAVE_FLOATS = 64 * SPLIT_ROW;  // per wave

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
// four at [4 l .. 4 l + 3].  Filled by the first wave; read after a barrier.
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

// Radix-16 DFT (forward sign) of the 16 points v[0..15] held by every lane of
// the wave, in place.  All 64 lanes must call it together; `scratch` is this
// wave's private LDS region of SPLIT_WAVE_FLOATS floats.  TAB: the constant
// operands are read from `atab` (an LDS table from split_atab_fill) at every
// call, and every index of the call is one the compiler cannot see through, so
// nothing of one call (operands, scratch addresses) stays live until the next;
// without, they are formed in registers (and a kernel keeps them live across
// its calls).
template <bool TAB>
__device__ __forceinline__ void dft16_split_mfma_impl(cpx<float>* v, float* scratch, const split_half8* atab) {
  int lane = threadIdx.x & 63;
  if constexpr (TAB) asm volatile("" : "+v"(lane));
  const int row = lane & 15, part = lane >> 4;
  // ---- this lane's vector -> scratch row `lane`
  float* my = scratch + lane * SPLIT_ROW;
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    my[t] = v[t].x;
    my[16 + t] = v[t].y;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // ---- constant operands, formed once this lane's points are out of registers
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

  // Group g reads rows 16g..16g+15 and writes its results back into the same
  // rows: the wave issues every read of the group before any of its writes,
  // and one wave's LDS operations complete in order, so each group is done in
  // place and only one group's values are live at a time.
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    // operand slice: vector (16g + row), values 8*part .. 8*part+7 of [re 0..15 | im 0..15]
    float* rowp = scratch + (16 * g + row) * SPLIT_ROW;
    float x[8];
    float mx = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      x[j] = rowp[8 * part + j];
      mx = fmaxf(mx, fabsf(x[j]));
    }
    // the vector's maximum over its four lanes (parts): each output column
    // depends on its own vector only, so every vector takes its own scale
    mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    // power-of-two scale bringing the vector's maximum into [2^14, 2^15)
    int ex = 0;
    (void)frexpf(mx, &ex);  // mx = f * 2^ex, f in [0.5, 1)
    const int sh = mx > 0.0f ? 15 - ex : 0;
    const float s = ldexpf(1.0f, sh);
    const float inv_s = ldexpf(1.0f, -sh);
    split_half8 bh, bl;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      _Float16 h, l;
      split2(x[j] * s, &h, &l);
      bh[j] = h;
      bl[j] = l;
    }
    const split_float4 z = {0.f, 0.f, 0.f, 0.f};
    split_float4 r = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, bh, z, 0, 0, 0);
    split_float4 q = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2h, bh, z, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, bl, r, 0, 0, 0);
    q = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2h, bl, q, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1l, bh, r, 0, 0, 0);
    q = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2l, bh, q, 0, 0, 0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // this lane now holds points 4*part+e of vector 16g+row
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      rowp[4 * part + e] = r[e] * inv_s;
      rowp[16 + 4 * part + e] = q[e] * inv_s;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int t = 0; t < 16; ++t) v[t] = {my[t], my[16 + t]};
}

__device__ __forceinline__ void dft16_split_mfma(cpx<float>* v, float* scratch) {
  dft16_split_mfma_impl<false>(v, scratch, nullptr);
}
__device__ __forceinline__ void dft16_split_mfma_tab(cpx<float>* v, float* scratch, const split_half8* atab) {
  dft16_split_mfma_impl<true>(v, scratch, atab);
}

// Complex elements of a stage buffer of `base` elements that split-precision
// variants (wg_dft16) also borrow as operand scratch for `nt` threads.
constexpr int split_lds_cpx(int base, int nt) {
  return base > (nt / 64) * SPLIT_WAVE_FLOATS / 2 ? base : (nt / 64) * SPLIT_WAVE_FLOATS / 2;
}

// Radix-16 DFT of v[0..15] in a workgroup kernel whose stage buffer `buf` is
// free at this point (every thread has read what it needs from it) and holds
// at least SPLIT_WAVE_FLOATS floats per wave: with SPLIT each wave borrows its
// slice as operand scratch between two barriers, so every thread of the
// workgroup must make the call (TAB: the constant operands come from the LDS
// table `atab`); without, the FP32 butterfly.
template <bool SPLIT, bool TAB = false>
__device__ __forceinline__ void wg_dft16(cpx<float>* v, cpx<float>* buf, const split_half8* atab = nullptr) {
  if constexpr (SPLIT) {
    __syncthreads();  // every wave finished reading the buffer it is about to borrow
    float* scratch = reinterpret_cast<float*>(buf) + (threadIdx.x >> 6) * SPLIT_WAVE_FLOATS;
    if constexpr (TAB) dft16_split_mfma_tab(v, scratch, atab);
    else dft16_split_mfma(v, scratch);
    __syncthreads();  // scratch is released before the buffer is written again
  } else {
    dft16(v);
  }
}

}  // namespace fftamd

// fft_pass.h -- arguments of one Stockham line-FFT pass launch (host/device).
#pragma once

namespace fftamd {

struct PassArgs {
  const void* in;
  void* out;
  // element strides (complex elements, not bytes)
  long long in_p, out_p;  // between points of a line
  long long in_c, out_c;  // between adjacent columns (lines) of a block
  long long in_k, out_k;  // middle line index
  long long in_o, out_o;  // outer line index
  unsigned ncol;          // columns per (outer, middle)
  unsigned ncb;           // column blocks = ceil(ncol / C)
  unsigned K;             // middle extent
  unsigned kbits;         // nonzero: output middle offset uses bitrev(k, kbits)
  unsigned long long outer0;  // first outer index of this launch
  const void* tw;         // Stockham stage twiddles for this line length
  // four-step post-twiddle W_M^t: t = ((col >> cshift) * k) & (M-1) = hi*2^hb + lo
  const void* ptw_lo;
  const void* ptw_hi;
  unsigned ptw_hb;
  unsigned ptw_mmask;
  unsigned ptw_cshift;
  int inverse;
  // optional real weights, one per point of a transform (broadcast over the
  // batch): the input is multiplied by w_load[e & w_mask] when read, the
  // output by w_store[e & w_mask] when written (e = element offset)
  const float* w_load;
  const float* w_store;
  unsigned long long w_mask;
  // optional transform table: outer index o starts at tab[2 o] in `in` and
  // tab[2 o + 1] in `out` (complex elements) instead of o * in_o, o * out_o
  const unsigned long long* tab;
};

// Launches one pass; returns 0, -1 for an unsupported (prec, lg, mode, epi), -2 on a launch error.
// split: radix-16 stages on FP16 matrix cores with hi+lo operands (FP32 only).
int launch_fft_pass(int prec, int lg, int mode, int epi, const PassArgs& a, unsigned long long blocks,
                    void* stream, bool split = false);

}  // namespace fftamd

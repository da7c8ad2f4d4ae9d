/*
 * oracle.c -- FP64 CPU restatement of FFTFPGA's C2C transform semantics.
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library.  The
 * shipped library never links or calls it.  See oracle.h for the reference
 * citations each function follows.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------- */
/* splitmix64 (S. Vigna, public-domain reference algorithm)                */
/* ---------------------------------------------------------------------- */
uint64_t orc_splitmix64_next(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_fill_uniform(float *ri, size_t npts, uint64_t seed) {
  uint64_t s = seed;
  for (size_t i = 0; i < 2 * npts; ++i) {
    uint64_t z = orc_splitmix64_next(&s);
    float u = (float)(z >> 40) * (1.0f / 16777216.0f); /* exact: 24-bit integer * 2^-24 */
    ri[i] = 2.0f * u - 1.0f;                            /* exact in FP32 */
  }
}

void orc_fill_gauss(float *ri, size_t npts, uint64_t seed) {
  uint64_t s = seed;
  for (size_t i = 0; i < npts; ++i) {
    uint64_t a = orc_splitmix64_next(&s);
    uint64_t b = orc_splitmix64_next(&s);
    double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0); /* (0,1] */
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);         /* [0,1) */
    double r = sqrt(-2.0 * log(u1));
    double t = 2.0 * M_PI * u2;
    ri[2 * i] = (float)(r * cos(t));
    ri[2 * i + 1] = (float)(r * sin(t));
  }
}

/* examples/helper.cpp:10-18: reverse the low `bits` bits of x */
unsigned orc_bit_reversed(unsigned x, unsigned bits) {
  unsigned y = 0;
  for (unsigned i = 0; i < bits; i++) {
    y = (y << 1) | (x & 1u);
    x >>= 1;
  }
  return y;
}

static int ilog2_exact(size_t n) {
  if (n == 0 || (n & (n - 1)) != 0) return -1;
  int k = 0;
  while (((size_t)1 << k) < n) ++k;
  return k;
}

/* ---------------------------------------------------------------------- */
/* FP64 iterative radix-2 DIT FFT.  Twiddles are evaluated directly with   */
/* cos/sin of 2*pi*k/n for every k (no recurrences), so the error stays    */
/* O(eps * log n).                                                          */
/* ---------------------------------------------------------------------- */
int orc_fft1d_f64(double *ri, size_t n, int sign) {
  int logn = ilog2_exact(n);
  if (logn < 0) return -1;
  if (n == 1) return 0;
  /* bit-reversal permutation */
  for (size_t i = 0; i < n; ++i) {
    size_t j = orc_bit_reversed((unsigned)i, (unsigned)logn);
    if (j > i) {
      double tr = ri[2 * i], ti = ri[2 * i + 1];
      ri[2 * i] = ri[2 * j];
      ri[2 * i + 1] = ri[2 * j + 1];
      ri[2 * j] = tr;
      ri[2 * j + 1] = ti;
    }
  }
  double *wc = (double *)malloc(sizeof(double) * n);
  if (!wc) return -2;
  double *ws = wc + n / 2;
  for (size_t k = 0; k < n / 2; ++k) {
    double a = 2.0 * M_PI * (double)k / (double)n;
    wc[k] = cos(a);
    ws[k] = (double)sign * sin(a);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    size_t half = len >> 1, step = n / len;
    for (size_t base = 0; base < n; base += len) {
      for (size_t k = 0; k < half; ++k) {
        double c = wc[k * step], s = ws[k * step];
        double *a = ri + 2 * (base + k), *b = ri + 2 * (base + k + half);
        double br = b[0] * c - b[1] * s;
        double bi = b[0] * s + b[1] * c;
        b[0] = a[0] - br;
        b[1] = a[1] - bi;
        a[0] += br;
        a[1] += bi;
      }
    }
  }
  free(wc);
  return 0;
}

/* Apply the 1D transform along every axis of each transform in the batch. */
static int fftnd_inplace(double *buf, int rank, const size_t *dims, size_t batch, int sign) {
  size_t tot = 1, maxd = 1;
  for (int d = 0; d < rank; ++d) {
    if (ilog2_exact(dims[d]) < 0) return -1;
    tot *= dims[d];
    if (dims[d] > maxd) maxd = dims[d];
  }
  double *line = (double *)malloc(sizeof(double) * 2 * maxd);
  if (!line) return -2;
  for (size_t b = 0; b < batch; ++b) {
    double *x = buf + 2 * b * tot;
    size_t stride = tot;
    for (int d = 0; d < rank; ++d) {
      size_t n = dims[d];
      stride /= n;                       /* elements between consecutive points on axis d */
      size_t outer = tot / (n * stride); /* independent blocks above axis d */
      for (size_t o = 0; o < outer; ++o) {
        for (size_t i = 0; i < stride; ++i) {
          double *p = x + 2 * (o * n * stride + i);
          for (size_t k = 0; k < n; ++k) {
            line[2 * k] = p[2 * k * stride];
            line[2 * k + 1] = p[2 * k * stride + 1];
          }
          orc_fft1d_f64(line, n, sign);
          for (size_t k = 0; k < n; ++k) {
            p[2 * k * stride] = line[2 * k];
            p[2 * k * stride + 1] = line[2 * k + 1];
          }
        }
      }
    }
  }
  free(line);
  return 0;
}

int orc_fftnd_f32in(const float *in, double *out, int rank, const size_t *dims,
                    size_t batch, int inverse) {
  size_t tot = batch;
  for (int d = 0; d < rank; ++d) tot *= dims[d];
  for (size_t i = 0; i < 2 * tot; ++i) out[i] = (double)in[i];
  return fftnd_inplace(out, rank, dims, batch, inverse ? +1 : -1);
}

int orc_fftnd_f64(const double *in, double *out, int rank, const size_t *dims,
                  size_t batch, int inverse) {
  size_t tot = batch;
  for (int d = 0; d < rank; ++d) tot *= dims[d];
  if (out != in) memcpy(out, in, sizeof(double) * 2 * tot);
  return fftnd_inplace(out, rank, dims, batch, inverse ? +1 : -1);
}

/* Brute-force d-dimensional DFT in long double: every output is the full sum
 * over all inputs of x[n] * exp(sign * 2 pi i * sum_d k_d n_d / N_d).  The phase
 * is reduced exactly in integers before the trig call. */
int orc_dftnd_naive(const double *in, double *out, int rank, const size_t *dims, int inverse) {
  if (rank < 1 || rank > 3) return -1;
  size_t n[3] = {1, 1, 1};
  for (int d = 0; d < rank; ++d) n[3 - rank + d] = dims[d];
  size_t tot = n[0] * n[1] * n[2];
  const long double twopi = 6.283185307179586476925286766559L;
  long double sgn = inverse ? 1.0L : -1.0L;
  for (size_t k0 = 0; k0 < n[0]; ++k0)
    for (size_t k1 = 0; k1 < n[1]; ++k1)
      for (size_t k2 = 0; k2 < n[2]; ++k2) {
        long double sr = 0.0L, si = 0.0L;
        for (size_t a = 0; a < n[0]; ++a)
          for (size_t b = 0; b < n[1]; ++b)
            for (size_t c = 0; c < n[2]; ++c) {
              /* phase as an exact fraction of a full turn with denominator n0*n1*n2 */
              size_t num = ((k0 * a) % n[0]) * (n[1] * n[2]) + ((k1 * b) % n[1]) * (n[0] * n[2]) +
                           ((k2 * c) % n[2]) * (n[0] * n[1]);
              num %= tot;
              long double ang = sgn * twopi * (long double)num / (long double)tot;
              long double cr = cosl(ang), ci = sinl(ang);
              size_t idx = (a * n[1] + b) * n[2] + c;
              long double xr = in[2 * idx], xi = in[2 * idx + 1];
              sr += xr * cr - xi * ci;
              si += xr * ci + xi * cr;
            }
        size_t o = (k0 * n[1] + k1) * n[2] + k2;
        out[2 * o] = (double)sr;
        out[2 * o + 1] = (double)si;
      }
  return 0;
}

/* Expected output of fftfpgaf_c2c_1d: each N-point transform is left in
 * bit-reversed order (kernels/fft1d/fft1d.cl:238-250); the reference harness
 * un-permutes with tmp[i] = fpgaout[bit_reversed(i)] (examples/helper.cpp:143-161),
 * hence out[bitrev(k)] = X[k]. */
int orc_c2c_1d_expected(const float *in, double *out, size_t n, size_t batch, int inverse) {
  int logn = ilog2_exact(n);
  if (logn < 0) return -1;
  double *tmp = (double *)malloc(sizeof(double) * 2 * n);
  if (!tmp) return -2;
  for (size_t b = 0; b < batch; ++b) {
    for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (double)in[2 * b * n + i];
    orc_fft1d_f64(tmp, n, inverse ? +1 : -1);
    for (size_t k = 0; k < n; ++k) {
      size_t br = orc_bit_reversed((unsigned)k, (unsigned)logn);
      out[2 * (b * n + br)] = tmp[2 * k];
      out[2 * (b * n + br) + 1] = tmp[2 * k + 1];
    }
  }
  free(tmp);
  return 0;
}

double orc_rel_l2(const float *got, const double *ref, size_t npts) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < 2 * npts; ++i) {
    double d = (double)got[i] - ref[i];
    num += d * d;
    den += ref[i] * ref[i];
  }
  return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

/* examples/helper.cpp:164-198 computes 10*log10(sum|X|^2 / sum|X-Y|^2) and
 * passes above 120 dB; accumulated in double here. */
double orc_snr_db(const float *got, const double *ref, size_t npts) {
  double mag = 0.0, noise = 0.0;
  for (size_t i = 0; i < 2 * npts; ++i) {
    double d = ref[i] - (double)got[i];
    mag += ref[i] * ref[i];
    noise += d * d;
  }
  if (noise == 0.0) return INFINITY;
  return 10.0 * log10(mag / noise);
}

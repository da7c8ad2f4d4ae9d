/*
 * oracle.h -- CPU restatement of FFTFPGA's single-precision C2C transform
 * semantics, used ONLY as test infrastructure (the parity checker).
 *
 * Nothing in the shipped library links or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * What it restates (reference paths relative to the FFTFPGA source tree):
 *   - forward transform e^{-2 pi i k n / N} per axis, inverse unnormalised
 *     e^{+2 pi i k n / N} (kernels/common/fft_8.cl:272-346, the re/im swap
 *     trick at :276-278 and :341-343 yields exactly the unnormalised backward
 *     DFT; examples/helper.cpp:133-141 plans FFTW_FORWARD / FFTW_BACKWARD)
 *   - 2D/3D outputs in natural row-major order
 *     (kernels/fft3d/fft3d_ddr.cl:466-485, kernels/fft2d/fft2d_ddr.cl:259-285)
 *   - 1D output bit-reversed within each transform: out[b] = X[bitrev(b)]
 *     (kernels/fft1d/fft1d.cl:238-250, un-permuted by examples/helper.cpp:143-161)
 *   - bit_reversed() (examples/helper.cpp:10-18,
 *     kernels/matrixTranspose/diagonal_bitrev.cl:100-109)
 *   - SNR check (examples/helper.cpp:164-198)
 *
 * The arithmetic reference is FP64 (the reference verified against FFTW;
 * FFTW is not present here, so the FP64 FFT below is pinned against numpy's
 * pocketfft, a long-double naive DFT and closed forms in tests/).
 *
 * Seeded input generator: splitmix64 (Vigna), draw order re then im per point.
 */
#ifndef FFTFPGA_ORACLE_H
#define FFTFPGA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded generators ------------------------------------------------ */
uint64_t orc_splitmix64_next(uint64_t *state);
/* re, im i.i.d. U[-1,1): u = (z >> 40) * 2^-24, value = 2u - 1 (exact in FP32) */
void orc_fill_uniform(float *ri, size_t npts, uint64_t seed);
/* re, im i.i.d. N(0,1): Box-Muller in FP64 on two draws per point, rounded to FP32 */
void orc_fill_gauss(float *ri, size_t npts, uint64_t seed);

/* ---- index maps ---------------------------------------------------------- */
unsigned orc_bit_reversed(unsigned x, unsigned bits);

/* ---- FP64 transforms (interleaved re,im doubles) ------------------------- */
/* in-place radix-2 FFT of n = 2^k points; sign -1 forward, +1 backward (unnormalised) */
int orc_fft1d_f64(double *ri, size_t n, int sign);
/* batched out-of-place N-d transform of FP32 input, FP64 output, row-major,
 * last dim contiguous, transforms contiguous (distance prod(dims)).
 * inverse != 0 -> unnormalised backward. */
int orc_fftnd_f32in(const float *in, double *out, int rank, const size_t *dims,
                    size_t batch, int inverse);
/* same on FP64 input */
int orc_fftnd_f64(const double *in, double *out, int rank, const size_t *dims,
                  size_t batch, int inverse);
/* naive O(N^2) DFT of a full rank-d array in long double (brute force for tiny sizes) */
int orc_dftnd_naive(const double *in, double *out, int rank, const size_t *dims, int inverse);

/* ---- drop-in output semantics --------------------------------------------- */
/* 1D: out[j*N + b] = X_j[bitrev(b)]  (expected output of fftfpgaf_c2c_1d) */
int orc_c2c_1d_expected(const float *in, double *out, size_t n, size_t batch, int inverse);

/* ---- error metrics --------------------------------------------------------- */
double orc_rel_l2(const float *got, const double *ref, size_t npts);
double orc_snr_db(const float *got, const double *ref, size_t npts);

#ifdef __cplusplus
}
#endif
#endif

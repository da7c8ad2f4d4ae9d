/*
 * fftfpga.h -- drop-in C ABI of the FFTFPGA single-precision C2C transform
 * library, implemented on one AMD Instinct MI355X (gfx950) with HIP.
 *
 * Every declaration below keeps the name, argument list, argument meaning and
 * struct layout of the reference header api/include/fftfpga/fftfpga.h, so an
 * existing caller recompiles (or relinks) unchanged.  Citations give the
 * reference declaration each entry point replaces.
 *
 * Calls are synchronous and take host pointers, like the reference: the input
 * is copied to HBM, transformed by hand-written gfx950 kernels and copied
 * back.  fpga_t reports the host->device copy, the device execution and the
 * device->host copy in milliseconds, measured with HIP events.
 *
 * Deviation from the reference: a runtime (HIP) failure prints to stderr and
 * returns fpga_t{valid = false} instead of calling exit().  Argument errors
 * return valid = false exactly where the reference does.
 */
#ifndef FFTFPGA_H
#define FFTFPGA_H

#include <stdbool.h>
#include <stddef.h>

/* Interleaved single-precision complex value (reference fftfpga.h:16-19). */
typedef struct {
  float x; /* real part */
  float y; /* imaginary part */
} float2;

/* Interleaved double-precision complex value (reference fftfpga.h:24-27). */
typedef struct {
  double x; /* real part */
  double y; /* imaginary part */
} double2;

/* Per-call timings in milliseconds plus the validity flag; same field order
 * and types as the reference (fftfpga.h:32-39).  svm_copyin_t/svm_copyout_t
 * time the staging copies of the *_svm entry points. */
typedef struct fpga_timing {
  double pcie_read_t;   /* device -> host copy */
  double pcie_write_t;  /* host -> device copy */
  double exec_t;        /* transform execution on the GPU */
  double svm_copyin_t;  /* staging copy into host-mapped memory (svm variants) */
  double svm_copyout_t; /* staging copy out of host-mapped memory (svm variants) */
  bool valid;           /* true when the transform ran */
} fpga_t;

#ifdef __cplusplus
extern "C" {
#endif

/* Select the GPU and prepare the library (reference fftfpga.h:56,
 * api/src/fftfpga.c:72-130).  Return codes are the reference's:
 *    0 success
 *   -1 path is NULL or empty
 *   -2 platform_name matches none of the platform names this library answers
 *      to (case-insensitive substring, as the reference's findPlatform):
 *      "AMD HIP", "Intel(R) FPGA SDK for OpenCL(TM)",
 *      "Intel(R) FPGA Emulation Platform for OpenCL(TM)"
 *   -3 no HIP device is visible
 *   -4 path does not name a readable file (the reference's bitstream path;
 *      here its contents are not interpreted)
 *   -5 use_svm requested but the device cannot map host memory */
extern int fpga_initialize(const char *platform_name, const char *path, const bool use_svm);

/* Same as fpga_initialize (name used by the FFTW-like API documentation). */
extern int fftfpga_init_fpga(const char *platform_name, const char *path, const bool use_svm);

/* Release device buffers, plans and twiddle tables (reference fftfpga.h:61). */
extern void fpga_final(void);

/* 64-byte aligned host allocation, NULL for sz == 0; release with free()
 * (reference fftfpga.h:68,75, api/src/fftfpga.c:37-58). */
extern void *fftfpga_complex_malloc(const size_t sz);
extern void *fftfpgaf_complex_malloc(const size_t sz);

/* Batched double-precision 1D transform; output bit-reversed per transform
 * (reference fftfpga.h:86). */
extern fpga_t fftfpga_c2c_1d(const unsigned N, const double2 *inp, double2 *out, const bool inv,
                             const unsigned iter);

/* Batched single-precision 1D transform of `iter` contiguous N-point signals.
 * Each output transform is left in bit-reversed order, as the reference's
 * engine writes it: out[j*N + b] = X_j[bit_reversed(b, log2 N)]
 * (reference fftfpga.h:97, kernels/fft1d/fft1d.cl:238-250). */
extern fpga_t fftfpgaf_c2c_1d(const unsigned N, const float2 *inp, float2 *out, const bool inv,
                              const unsigned iter);

/* As fftfpgaf_c2c_1d with host-mapped staging; requires use_svm at init
 * (reference fftfpga.h:108). */
extern fpga_t fftfpgaf_c2c_1d_svm(const unsigned N, const float2 *inp, float2 *out, const bool inv,
                                  const unsigned batch);

/* `how_many` contiguous N x N transforms, natural order (reference fftfpga.h:120). */
extern fpga_t fftfpgaf_c2c_2d_bram(const unsigned N, const float2 *inp, float2 *out, const bool inv,
                                   const bool interleaving, const unsigned how_many);

/* As fftfpgaf_c2c_2d_bram with host-mapped staging (reference fftfpga.h:131). */
extern fpga_t fftfpgaf_c2c_2d_bram_svm(const unsigned N, const float2 *inp, float2 *out,
                                       const bool inv, const unsigned how_many);

/* One N x N transform, natural order (reference fftfpga.h:141). */
extern fpga_t fftfpgaf_c2c_2d_ddr(const unsigned N, const float2 *inp, float2 *out, const bool inv);

/* One N x N x N transform, natural order (reference fftfpga.h:152). */
extern fpga_t fftfpgaf_c2c_3d_bram(const unsigned N, const float2 *inp, float2 *out, const bool inv,
                                   const bool interleaving);

/* One N x N x N transform, natural [z][y][x] order, x contiguous; the hot
 * path (reference fftfpga.h:162, api/src/fft3d.c:206-384). */
extern fpga_t fftfpgaf_c2c_3d_ddr(const unsigned N, const float2 *inp, float2 *out, const bool inv);

/* `how_many` (> 1) contiguous N^3 transforms; host->device copies, transforms
 * and device->host copies are pipelined, and exec_t is the wall-clock time of
 * the whole call including the copies, as in the reference
 * (fftfpga.h:164, api/src/fft3d.c:396-913). */
extern fpga_t fftfpgaf_c2c_3d_ddr_batch(const unsigned N, const float2 *inp, float2 *out,
                                        const bool inv, const bool interleaving,
                                        const unsigned how_many);

/* Host-mapped staging variants of the 3D transform (reference fftfpga.h:175,186). */
extern fpga_t fftfpgaf_c2c_3d_ddr_svm(const unsigned N, const float2 *inp, float2 *out,
                                      const bool inv, const bool interleaving);
extern fpga_t fftfpgaf_c2c_3d_ddr_svm_batch(const unsigned N, const float2 *inp, float2 *out,
                                            const bool inv, const unsigned how_many);

#ifdef __cplusplus
}
#endif

#endif /* FFTFPGA_H */

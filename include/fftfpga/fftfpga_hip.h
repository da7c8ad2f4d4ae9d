/*
 * fftfpga_hip.h -- ABI-additive extensions of the FFTFPGA C API for callers
 * that already hold data in GPU memory (PyTorch tensors, HIP applications).
 *
 * The reference has no counterpart for these entry points: it only accepts
 * host pointers and copies over PCIe on every call (api/src/fft3d.c:242-264).
 * They run the same gfx950 kernels as fftfpga.h, asynchronously on the
 * caller's HIP stream (a hipStream_t passed as void*; NULL is the HIP null
 * stream, as everywhere in HIP).  All return 0 on success and a negative code on
 * error, with the reason in fftamd_last_error().
 *
 * Layout conventions are those of fftfpga.h: interleaved complex, row-major,
 * last dimension contiguous, transforms of a batch contiguous.
 * d_in == d_out (in place) is allowed; input and output ranges that overlap
 * without being the same buffer are refused (-1).
 *
 * The first call for a shape (dims, batch, flags, precision) builds its plan,
 * its twiddle tables and the code objects of its kernels; later calls for it
 * allocate nothing and may be captured into a HIP graph (stream capture) and
 * replayed.  Multi-transform plans make their tables when created.
 *
 * Concurrency: calls are serialised on the host by one library lock, but run
 * asynchronously on the caller's stream.  Some in-place shapes (256^3 blocks
 * through the y-split pair, 256 x 256 planes, transposing four-step passes)
 * run through a temporary buffer that belongs to the shape's plan: in-place
 * calls of one shape must not overlap in time on different streams (order
 * them on one stream, or use separate output buffers).
 *
 * Lifetime: fpga_final (and a new fpga_initialize) releases every plan and
 * twiddle table.  Multi-transform plans and graphs made before then are stale:
 * fftamd_multi_exec, fftamd_multi_tune, fftamd_graph_capture and
 * fftamd_graph_launch refuse them with -1; destroy and create them again.
 */
#ifndef FFTFPGA_HIP_H
#define FFTFPGA_HIP_H

#include "fftfpga.h"

#ifdef __cplusplus
extern "C" {
#endif

/* flags */
#define FFTAMD_BITREV_1D 0x1u  /* rank 1: leave each transform bit-reversed, as fftfpgaf_c2c_1d */
/* Split precision: radix-16 stages as 16x16 complex DFT products on FP16 matrix
 * cores (v_mfma_f32_16x16x32_f16) with hi+lo split operands, power-of-two block
 * scaling and FP32 accumulation; other stages and twiddles stay FP32.  Single
 * precision only; lines of 16..4096 points use it, and so do the y-split pairs
 * of 128^3, 256^3 and 512^3 blocks (transforms without weights). */
#define FFTAMD_SPLIT_MFMA 0x2u
/* Single-precision transforms whose two fastest axes form planes up to
 * 128 x 128, or 256 x 256 (from 32 planes up), run both axes in one fused pass
 * over each plane (a 3D transform then reads and writes HBM twice, not three
 * times); a second-fastest axis of 4096 points (from 64 groups of 16 columns
 * up) runs in one pass instead of a four-step pair.  This flag asks for the
 * separate line passes instead (it also turns off the y-split pair below). */
#define FFTAMD_NO_PLANE_FUSION 0x4u
/* Single-precision arrays whose three fastest axes are 256 x 256 x 256 or
 * 128 x 128 x 128 run in two passes with the y axis split (n/4) x 4 between
 * them (x and (n/4)-point y over one residue class of rows per workgroup; then
 * the 4-point y stage and z), so that every tile fits one workgroup and no
 * input is read twice; 512 x 512 x 512 likewise with y split 64 x 8.  This
 * flag keeps the plane pass and a separate z pass instead (three line passes
 * for 512^3). */
#define FFTAMD_NO_CUBE_SPLIT 0x8u

/* Single precision, rank 1..3, every dimension a power of two, `batch`
 * contiguous transforms.  inverse != 0: unnormalised backward transform. */
int fftamd_c2c_dev(int rank, const unsigned *dims, unsigned batch, const float2 *d_in, float2 *d_out,
                   int inverse, unsigned flags, void *stream);

/* fftamd_c2c_dev that allocates no device memory on any call, its first
 * included.  No counterpart in the reference: it is for callers that cannot
 * allocate -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Every FP32 twiddle table and the temporary buffer of
 * in-place shapes live in one static device array of 64 MiB per GPU (held
 * until the process ends), set up when the library initialises on a device
 * (fpga_initialize, or the first device call).  Shapes whose in-place
 * temporary buffer exceeds the array (about 63.8 MiB: e.g. 256^3 in place)
 * are refused with -1; call them out of place or through fftamd_c2c_dev.
 * Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap in time on one GPU (different streams) race.
 * The host keeps one small plan record per new shape; a repeated shape
 * allocates nothing at all.  Results equal fftamd_c2c_dev's. */
int fftamd_c2c_dev_noalloc(int rank, const unsigned *dims, unsigned batch, const float2 *d_in,
                           float2 *d_out, int inverse, unsigned flags, void *stream);

/* Double precision twin of fftamd_c2c_dev. */
int fftamd_z2z_dev(int rank, const unsigned *dims, unsigned batch, const double2 *d_in,
                   double2 *d_out, int inverse, unsigned flags, void *stream);

/* Weighted single-precision transform: out = w_out .* F(w_in .* in), where F is
 * the forward (inverse == 0) or unnormalised backward transform and w_in,
 * w_out are real FP32 device arrays of one transform's shape (prod(dims)
 * points, broadcast over the batch), either may be NULL.  The weights are
 * applied inside the first pass's loads and the last pass's stores, so they
 * cost no extra pass over memory.  With w_in = W and inverse = 1 this is the
 * adjoint of out = W .* F(in), which the autograd binding uses as its
 * backward. */
int fftamd_c2c_weighted_dev(int rank, const unsigned *dims, unsigned batch, const float2 *d_in,
                            float2 *d_out, int inverse, const float *w_in, const float *w_out,
                            void *stream);

/* Device-pointer twin of fftfpgaf_c2c_3d_ddr / _3d_ddr_batch: how_many
 * contiguous N^3 transforms (how_many >= 1). */
int fftfpgaf_c2c_3d_ddr_dev(unsigned N, const float2 *d_in, float2 *d_out, bool inv,
                            unsigned how_many, void *stream);

/* ---- transforms of uneven sizes in one persistent launch --------------------------
 *
 * A batch of 3D transforms of different power-of-two shapes (each axis <= 256)
 * placed anywhere in two device buffers.  The reference's batched call takes one
 * N per call (api/src/fft3d.c:396-913); this plan runs every pass of every
 * transform in a single persistent kernel: per phase (axis), 256-thread
 * workgroups take chunks of tiles from a device-side work queue, largest
 * transforms first, with a grid barrier between phases; or, for batches of
 * 32^3, 64^3 and 128^3 cubes, one dataflow launch of both passes of every
 * transform (plane tiles and z lines, or the y-split pair for 128^3) from one
 * work queue per XCD, where a transform's second pass waits only for its own
 * first pass.  Schedules (single launch, one launch per phase, plane and
 * z-pass kernels per shape group, or the dataflow launch; workgroups per CU,
 * tiles per queue ticket) are picked by fftamd_multi_tune and cached per shape
 * signature. */
typedef struct fftamd_multi_s *fftamd_multi_t;
typedef struct fftamd_graph_s *fftamd_graph_t;

typedef struct {
  unsigned n[3];                 /* dims (slowest .. fastest), x = n[2] contiguous */
  unsigned reserved;             /* must be 0 */
  unsigned long long in_offset;  /* first element, in complex elements from d_in */
  unsigned long long out_offset; /* first element, in complex elements from d_out */
} fftamd_xform3d;

/* flags: 0 or FFTAMD_SPLIT_MFMA. */
int fftamd_multi_create(fftamd_multi_t *plan, unsigned count, const fftamd_xform3d *xf,
                        const float2 *d_in, float2 *d_out, int inverse, unsigned flags);
/* Enqueue the plan on a stream (asynchronous). */
int fftamd_multi_exec(fftamd_multi_t plan, void *stream);
/* Synchronise the stream and report whether a wait of the last launch (grid
 * barrier in mode 0, dependency in mode 3) exceeded its bound (-1 with a
 * message) or not (0). */
int fftamd_multi_check(fftamd_multi_t plan, void *stream);
/* Time every schedule candidate on the plan's buffers (reps launches each),
 * keep the fastest and append it to cache_path (NULL: no cache).  Returns 0
 * after tuning, 1 when the schedule came from the cache, -1 on error. */
int fftamd_multi_tune(fftamd_multi_t plan, void *stream, const char *cache_path, int reps,
                      double *best_us);
/* mode 0: one launch with grid barriers; mode 1: one launch per phase;
 * mode 2 (when every transform has three axes of >= 16 points with x-y planes
 * of at most 128 x 128): per shape group, the
 * x and y passes as one plane-kernel launch (a whole plane per workgroup, in
 * LDS) and the z passes as one line-pass launch, both over a table of the
 * group's transforms (128^3 groups whose inputs and outputs do not overlap:
 * the y-split pair instead; grid and chunk do not apply); mode 3 (FP32 plans
 * whose transforms are all 32^3, 64^3 or 128^3, 128^3 out of place): the
 * same tiles in one persistent launch of 512-thread workgroups, ordered by a
 * counter per transform instead of barriers, from one work queue per XCD (a
 * transform's tiles all in one queue, its second pass after its first), so
 * that it completes with any number of resident workgroups.  chunk: tiles
 * per queue ticket in modes 0, 1 and 3, or 0 for a static round-robin
 * assignment of tiles to workgroups (no tickets; modes 0 and 1).  grid:
 * workgroups of the persistent kernel, at most those resident at once (0:
 * all of them).  The tuner considers every available mode. */
int fftamd_multi_schedule(fftamd_multi_t plan, int *mode, unsigned *grid, unsigned *chunk,
                          unsigned *items, unsigned *phases);
int fftamd_multi_set_schedule(fftamd_multi_t plan, int mode, unsigned grid, unsigned chunk);
void fftamd_multi_destroy(fftamd_multi_t plan);
/* Bound on the polls of a grid barrier or dependency wait (default 2^22); a
 * tiny value forces the timeout path so callers can test their error
 * handling. */
void fftamd_multi_set_spin_limit(unsigned polls);

/* Capture plans[0..n-1] executed in order (e.g. a forward/inverse pair) into a
 * HIP graph on a non-null stream; replay with fftamd_graph_launch.  A plan's
 * schedule is captured as it is at capture time (an untuned plan runs mode 2
 * where available, else mode 1; neither waits on a grid barrier).  Only
 * fftamd_multi_check reports a mode-0 barrier timeout. */
int fftamd_graph_capture(fftamd_graph_t *graph, fftamd_multi_t *plans, int n, void *stream);
int fftamd_graph_launch(fftamd_graph_t graph, void *stream);
void fftamd_graph_destroy(fftamd_graph_t graph);

/* Last error message of the calling thread ("" if none). */
const char *fftamd_last_error(void);

/* Drop every cached plan (twiddle tables stay until fpga_final). */
void fftamd_plan_cache_clear(void);

/* Number of cached plans. */
size_t fftamd_plan_cache_size(void);

/* Per-kernel timing: after fftamd_profile_begin, every kernel launch of the
 * library is bracketed by HIP events on its own stream (up to max_records
 * launches).  fftamd_profile_end synchronises, stops recording and writes a
 * JSON object {"<kernel label>": {"count": n, "total_ms": t}, ...} into
 * json_out (truncated to cap bytes); returns the full JSON length or -1. */
int fftamd_profile_begin(int max_records);
int fftamd_profile_end(char *json_out, int cap);

/* Describe, without touching the device, the passes the library would run
 * for a transform: 22 int64 fields per pass (line log2 length, mode,
 * epilogue, transposing, point/column/middle/outer strides in and out,
 * column count, middle extent, middle bit-reversal bits, outer extent,
 * four-step modulus, column shift, source and destination buffer, columns
 * per workgroup, threads per workgroup) followed by one final-copy flag.
 * Modes: 0 contiguous lines, 1 strided lines, 2 plane pass (the 2D transform
 * of every plane of middle-extent rows x 2^log2-length columns; outer extent
 * = planes, outer stride = plane size).
 * prec: 0 FP32, 1 FP64.  With out == NULL returns the pass count; otherwise
 * writes at most cap values and returns the pass count, or -1 on error. */
int fftamd_describe_plan(int rank, const unsigned *dims, unsigned batch, unsigned flags, int prec,
                         int aliased, long long *out, int cap);

/* Describe, without touching the device, the work queues of the dataflow
 * schedule (mode 3) for `count` transforms of dims3[3 i .. 3 i + 2], each
 * 32^3, 64^3 or 128^3: qstart[0..8] receives the queue starts (one queue per
 * XCD) and, when out is not NULL, out[3 j .. 3 j + 2] the j-th tile as
 * (transform index, pass 0 or 1, tile within that pass).  Returns the tile
 * count, or -1 on error (a transform of another shape, cap too small). */
int fftamd_describe_flow(unsigned count, const unsigned *dims3, unsigned *qstart, unsigned *out,
                         unsigned cap);

#ifdef __cplusplus
}
#endif

#endif /* FFTFPGA_HIP_H */

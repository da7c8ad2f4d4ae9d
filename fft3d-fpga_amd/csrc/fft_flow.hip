// fft_flow.hip -- the dataflow launch of a batch of n^3 transforms (n = 32,
// 64, 128): one persistent kernel runs every tile of both passes of every
// transform, the GPU counterpart of the reference's batched 3D path
// (api/src/fft3d.c:396-913, one size per call) for batches of uneven sizes.
//
// First pass: n = 32, 64 the plane tiles (x and y of a plane, fft_plane_dev.h),
// n = 128 the y-split pair's pass A; second pass: the z lines in place on the
// output (n = 32, 64) or pass B (n = 128).  The tiles are those of the grouped
// launches, so the results are bit-identical to them.
//
// Ordering without a grid barrier: every transform has a counter of its
// finished first-pass tiles.  Finished first-pass tiles are published by an
// agent-scope release (their stores reach memory) and atomic adds on their
// transforms' counters; a second-pass tile starts by polling its transform's
// counter up to the transform's first-pass tile count, then an agent-scope
// acquire.  The poll is bounded: past spin_limit polls the workgroup sets
// ctl->err and exits.
//
// Work queues: one per XCD (FLOW_QUEUES), each consumed in order through an
// atomic ticket of `chunk` tiles.  A workgroup starts on queue blockIdx % 8
// and, when that one is drained, helps the others in turn.  The host puts all
// tiles of a transform in one queue, its second-pass tiles after all its
// first-pass ones, so the lowest unfinished tile of a queue never waits: the
// launch completes with any number of resident workgroups.
#include <hip/hip_runtime.h>

#include "launch.h"

#include "fft_cube128_dev.h"
#include "fft_multi.h"
#include "fft_plane_dev.h"

namespace fftamd {

// LDS of each tile kind (FP32), side by side where a tile holds several
using FlowZ32 = Geom<32, MODE_STRIDED>;
using FlowZ64 = Geom<64, MODE_STRIDED>;
constexpr int FLOW_P32 = PlaneGeom<32, 32, false>::LDS;
constexpr int FLOW_P64 = PlaneGeom<64, 64, false>::LDS;
constexpr int FLOW_CA = C1A_BUF + C1A_TWS;
constexpr int flow_max(int a, int b) { return a > b ? a : b; }
constexpr int FLOW_LDS = flow_max(flow_max(8 * FLOW_P32, 2 * FLOW_P64),
                                  flow_max(flow_max(2 * FlowZ32::LDS_ELEMS, 2 * FlowZ64::LDS_ELEMS),
                                           flow_max(2 * FLOW_CA, C1Z::LDS_ELEMS)));
static_assert(PlaneGeom<32, 32, false>::NT * 8 == FLOW_THREADS && PlaneGeom<64, 64, false>::NT * 2 == FLOW_THREADS,
              "eight 32 x 32 planes or two 64 x 64 planes per tile");
static_assert(FlowZ32::NT * 2 == FLOW_THREADS && FlowZ64::NT * 2 == FLOW_THREADS && C1_NT * 2 == FLOW_THREADS &&
                  C1B_NT == FLOW_THREADS,
              "two z-line or pass-A blocks, or one pass-B block, per tile");
static_assert(FLOW_LDS * 8 * 2 <= 160 * 1024, "two workgroups per CU");
static_assert(sizeof(FlowArgs) + 64 <= 4096, "kernel arguments within 4 KiB");

// One tile (every thread of the workgroup takes part; the tiles it holds side
// by side have the same barriers).
template <bool INV>
__device__ __forceinline__ void flow_tile(const FlowArgs& fa, const FlowItem& it, cpx<float>* lds, int tid) {
  constexpr int DIR = INV ? 1 : 0;
  // which side-by-side tile a thread works on is uniform over its wave (scalar)
  const int half = __builtin_amdgcn_readfirstlane(tid >> 8), h = tid & 255;
  const unsigned long long tr = it.tr, sub = it.sub;
  if (it.group == 0) {  // 32^3
    const FlowGroup& g = fa.grp[0];
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (!it.pass)
      plane_tile<32, 32, INV, false>(g.pa, tr * 32 + 8 * sub + w, lds + w * FLOW_P32, tid & 63);
    else
      pass_block<float, 32, MODE_STRIDED, EPI_NONE, false, DIR>(g.za, tr * g.za.ncb + 2 * sub + half,
                                                                 lds + half * FlowZ32::LDS_ELEMS, h);
  } else if (it.group == 1) {  // 64^3
    const FlowGroup& g = fa.grp[1];
    if (!it.pass)
      plane_tile<64, 64, INV, false>(g.pa, tr * 64 + 2 * sub + half, lds + half * FLOW_P64, h);
    else
      pass_block<float, 64, MODE_STRIDED, EPI_NONE, false, DIR>(g.za, tr * g.za.ncb + 2 * sub + half,
                                                                 lds + half * FlowZ64::LDS_ELEMS, h);
  } else {  // 128^3
    const FlowGroup& g = fa.grp[2];
    if (!it.pass)
      cube128_a_tile<INV, false, false>(g.ca, tr * 512 + 2 * sub + half, lds + half * FLOW_CA,
                                        lds + half * FLOW_CA + C1A_BUF, h);
    else
      cube128_b_tile<INV, false, false>(g.cb, (unsigned)(tr * 128 + sub), lds, nullptr, tid);
  }
}

// Finished first-pass tiles not yet counted: runs of (counter, tiles) kept
// by thread 0 until the workgroup publishes them.
constexpr int FLOW_PEND = 64;

// Publish: every wave's stores are ordered before a workgroup barrier
// (workgroup-scope release), after which the first wave alone makes them
// visible to the agent (one agent-scope release, which writes back the L2 that
// all of the workgroup's stores have reached) and adds the pending tiles to
// their transforms' counters.  One L2 writeback per publication instead of one
// per wave.
__device__ __forceinline__ void flow_publish(unsigned* done, const unsigned* pdep, const unsigned* pcnt,
                                             unsigned npend, int tid) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (tid < 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (tid == 0)
      for (unsigned i = 0; i < npend; ++i)
        __hip_atomic_fetch_add(done + (size_t)pdep[i] * FLOW_DONE_STRIDE, pcnt[i], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
  }
}

// An agent-scope release writes back the XCD's whole L2 (and an acquire
// invalidates it), so each is taken by one wave per workgroup, and a workgroup
// publishes its finished first-pass tiles only before it waits (its first
// second-pass tile after first-pass ones), before it leaves a queue, or when
// FLOW_PEND runs are pending -- never while holding unpublished tiles in a
// wait, so the ordering argument above still holds.  A run of consecutive
// second-pass tiles of one ticket waits for all their counters, then takes
// one agent-scope acquire.
template <bool INV>
__global__ void __launch_bounds__(FLOW_THREADS, 4)
    k_fft_flow(const FlowArgs fa, const FlowItem* __restrict__ items,
               const unsigned* __restrict__ qstart, MultiCtl* ctl, unsigned* done, unsigned chunk,
               unsigned spin_limit) {
  __shared__ cpx<float> lds[FLOW_LDS];
  __shared__ unsigned s_ticket;
  __shared__ int s_abort;
  __shared__ unsigned s_pdep[FLOW_PEND], s_pcnt[FLOW_PEND];  // written and read by thread 0 only
  const int tid = threadIdx.x;
  const unsigned home = blockIdx.x % FLOW_QUEUES;
  unsigned npend = 0, last_dep = ~0u;  // uniform over the workgroup
  for (unsigned qi = 0; qi < FLOW_QUEUES; ++qi) {
    const unsigned q = (home + qi) % FLOW_QUEUES;
    const unsigned base = qstart[q];
    const unsigned count = qstart[q + 1] - base;
    for (;;) {
      __syncthreads();  // the previous tile's LDS reads and s_ticket reads are done
      // the counter only grows: once a plain read shows the queue drained, no
      // ticket is taken (a stale read just falls through to the atomic)
      if (tid == 0)
        s_ticket = __hip_atomic_load(&ctl->qhead[q][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= count
                       ? count
                       : atomicAdd(&ctl->qhead[q][0], chunk);
      __syncthreads();
      const unsigned t0 = __builtin_amdgcn_readfirstlane(s_ticket);
      if (t0 >= count) break;
      const unsigned t1 = min(t0 + chunk, count);
      bool acquired = false;  // for the current run of second-pass tiles
      for (unsigned t = t0; t < t1; ++t) {
        const FlowItem it = items[base + t];
        if (it.pass) {
          if (!acquired) {
            if (npend) {
              flow_publish(done, s_pdep, s_pcnt, npend, tid);
              npend = 0;
              last_dep = ~0u;
            }
            // wait for the counters of this run of second-pass tiles
            if (tid == 0) {
              int abort = 0;
              for (unsigned u = t; u < t1 && !abort; ++u) {
                const FlowItem iu = items[base + u];
                if (!iu.pass) break;
                const FlowGroup& gu = fa.grp[iu.group];
                const unsigned* cu = done + (size_t)(gu.dep0 + iu.tr) * FLOW_DONE_STRIDE;
                unsigned polls = 0;
                while (__hip_atomic_load(cu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gu.need) {
                  if (++polls > spin_limit) {
                    __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    abort = 1;
                    break;
                  }
                  __builtin_amdgcn_s_sleep(2);
                }
              }
              s_abort = abort;
            }
            // the first wave takes the agent-scope acquire after the polls (it
            // invalidates the CU's L1 and the L2 for the whole workgroup); the
            // barrier then orders the other waves' loads after it
            if (tid < 64) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __syncthreads();
            if (s_abort) return;
            acquired = true;
          } else {
            __syncthreads();  // the previous tile's LDS reads are done
          }
        } else {
          acquired = false;  // a later second-pass tile of this ticket may need this one
          if (t != t0) __syncthreads();  // the previous tile's LDS reads are done
        }
        // the thread index made opaque per tile, so that no address the tile
        // derives from it is hoisted out of the loop (and held in registers
        // across every tile)
        int ltid = tid;
        asm volatile("" : "+v"(ltid));
        flow_tile<INV>(fa, it, lds, ltid);
        if (!it.pass) {
          const unsigned dep = fa.grp[it.group].dep0 + it.tr;
          if (dep == last_dep) {
            if (tid == 0) ++s_pcnt[npend - 1];
          } else {
            if (tid == 0) {
              s_pdep[npend] = dep;
              s_pcnt[npend] = 1;
            }
            ++npend;
            last_dep = dep;
          }
          if (npend == FLOW_PEND) {
            flow_publish(done, s_pdep, s_pcnt, npend, tid);
            npend = 0;
            last_dep = ~0u;
          }
        }
      }
    }
    if (npend) {  // before leaving the queue
      flow_publish(done, s_pdep, s_pcnt, npend, tid);
      npend = 0;
      last_dep = ~0u;
    }
  }
}

__global__ void k_flow_reset(MultiCtl* ctl, unsigned* done, unsigned ndone) {
  unsigned* w = reinterpret_cast<unsigned*>(ctl);
  const unsigned n = (unsigned)(sizeof(MultiCtl) / sizeof(unsigned));
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0u;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < ndone; i += gridDim.x * blockDim.x)
    done[(size_t)i * FLOW_DONE_STRIDE] = 0u;
}

void preload_flow() { preload_kernel(reinterpret_cast<const void*>(&k_fft_flow<false>)); }

int launch_flow(const FlowArgs& fa, const FlowItem* items, const unsigned* qstart, MultiCtl* ctl,
                unsigned* done, unsigned ndone, unsigned grid, unsigned chunk, unsigned spin_limit,
                bool inverse, void* stream) {
  if (grid == 0 || chunk == 0) return -1;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  FFTAMD_LAUNCH(k_flow_reset, dim3(16), dim3(256), 0, s, ctl, done, ndone);
  if (inverse)
    FFTAMD_LAUNCH(k_fft_flow<true>, dim3(grid), dim3(FLOW_THREADS), 0, s, fa, items, qstart, ctl, done,
                  chunk, spin_limit);
  else
    FFTAMD_LAUNCH(k_fft_flow<false>, dim3(grid), dim3(FLOW_THREADS), 0, s, fa, items, qstart, ctl, done,
                  chunk, spin_limit);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int flow_max_blocks_per_cu(int* out) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fft_flow<false>, FLOW_THREADS, 0) != hipSuccess)
    return -1;
  *out = n;
  return 0;
}

}  // namespace fftamd

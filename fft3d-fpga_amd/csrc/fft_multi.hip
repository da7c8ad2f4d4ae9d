// fft_multi.hip -- persistent kernel for a batch of transforms of uneven
// power-of-two sizes (the GPU counterpart of the reference's batched 3D path,
// api/src/fft3d.c:396-913, which handles one size per call).
//
// One launch runs every pass of every transform.  Pass p of all transforms
// forms phase p.  Within a phase, 256-thread workgroups take chunks of tiles
// from a device-side work queue (one atomic ticket per chunk; tiles are
// ordered largest transform first), run them with the same pass_block() as
// the one-pass kernels, and meet at a grid barrier before the next phase:
// the workgroup publishes its stores (one agent-scope release by its first
// wave), one thread per workgroup arrives on a counter and waits for all
// workgroups (agent-scope acquire by the first wave).  The wait is bounded:
// past spin_limit polls the workgroup sets ctl->err and exits, so a launch
// that cannot make progress ends with an error instead of hanging.  The host
// sizes the grid to what is resident.
#include <hip/hip_runtime.h>

#include "launch.h"

#include "fft_multi.h"
#include "fft_pass_dev.h"

namespace fftamd {

template <int MODE, int LG, bool SPLIT>
__device__ __forceinline__ void run_path_body(const PassArgs& a, unsigned long long blk, cpx<float>* lds,
                                              int tid) {
  static_assert(geom_threads(1 << LG, MODE) == MULTI_THREADS, "multi paths use 256-thread tiles");
  // split precision applies to lines with a radix-16 stage
  pass_block<float, (1 << LG), MODE, EPI_NONE, SPLIT && (LG >= 4)>(a, blk, lds, tid);
}

// Split-precision paths run as one call per tile: out of line, each gets its
// own register allocation (inlined together they exceed the 128-VGPR budget
// of 4 waves per SIMD and spill heavily).  FP32 paths stay inlined.
template <int MODE, int LG>
__device__ __noinline__ void run_path_call(const PassArgs& a, unsigned long long blk, cpx<float>* lds, int tid) {
  run_path_body<MODE, LG, true>(a, blk, lds, tid);
}

template <int MODE, int LG, bool SPLIT>
__device__ __forceinline__ void run_path(const PassArgs& a, unsigned long long blk, cpx<float>* lds, int tid) {
  if constexpr (SPLIT) run_path_call<MODE, LG>(a, blk, lds, tid);
  else run_path_body<MODE, LG, false>(a, blk, lds, tid);
}

template <bool SPLIT>
__device__ __forceinline__ void dispatch_path(int path, const PassArgs& a, unsigned long long blk,
                                              cpx<float>* lds) {
  // the thread index made opaque per tile, so that no address a tile derives
  // from it is hoisted out of the persistent loop (and held in registers
  // across every tile)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  switch (path) {
    case 0: run_path<MODE_CONTIG, 1, SPLIT>(a, blk, lds, tid); break;
    case 1: run_path<MODE_CONTIG, 2, SPLIT>(a, blk, lds, tid); break;
    case 2: run_path<MODE_CONTIG, 3, SPLIT>(a, blk, lds, tid); break;
    case 3: run_path<MODE_CONTIG, 4, SPLIT>(a, blk, lds, tid); break;
    case 4: run_path<MODE_CONTIG, 5, SPLIT>(a, blk, lds, tid); break;
    case 5: run_path<MODE_CONTIG, 6, SPLIT>(a, blk, lds, tid); break;
    case 6: run_path<MODE_CONTIG, 7, SPLIT>(a, blk, lds, tid); break;
    case 7: run_path<MODE_CONTIG, 8, SPLIT>(a, blk, lds, tid); break;
    case 8: run_path<MODE_STRIDED, 1, SPLIT>(a, blk, lds, tid); break;
    case 9: run_path<MODE_STRIDED, 2, SPLIT>(a, blk, lds, tid); break;
    case 10: run_path<MODE_STRIDED, 3, SPLIT>(a, blk, lds, tid); break;
    case 11: run_path<MODE_STRIDED, 4, SPLIT>(a, blk, lds, tid); break;
    case 12: run_path<MODE_STRIDED, 5, SPLIT>(a, blk, lds, tid); break;
    case 13: run_path<MODE_STRIDED, 6, SPLIT>(a, blk, lds, tid); break;
    case 14: run_path<MODE_STRIDED, 7, SPLIT>(a, blk, lds, tid); break;
    case 15: run_path<MODE_STRIDED, 8, SPLIT>(a, blk, lds, tid); break;
    default: break;
  }
}

// largest LDS footprint over the 16 paths (L <= 256, 256 threads)
constexpr int multi_lds_elems() {
  int m = 1;
  for (int lg = 1; lg <= MULTI_MAX_LG; ++lg) {
    const int L = 1 << lg;
    const int c0 = geom_cols(L, MODE_CONTIG) * (L + L / 16);
    const int c1 = L * (geom_cols(L, MODE_STRIDED) + 1);
    m = c0 > m ? c0 : m;
    m = c1 > m ? c1 : m;
  }
  return m;
}

__global__ void k_multi_reset(MultiCtl* ctl) {
  unsigned* w = reinterpret_cast<unsigned*>(ctl);
  for (unsigned i = threadIdx.x; i < sizeof(MultiCtl) / sizeof(unsigned); i += blockDim.x) w[i] = 0u;
}

// 4 waves per SIMD: the 34.8 KB LDS footprint allows 4 workgroups per CU
template <bool SPLIT>
__global__ void __launch_bounds__(MULTI_THREADS, 4)
    k_fft_multi(const MultiPass* __restrict__ passes, const MultiItem* __restrict__ items,
                const unsigned* __restrict__ phase_start, unsigned phase0, unsigned nphase,
                MultiCtl* ctl, unsigned chunk, unsigned spin_limit) {
  __shared__ cpx<float> lds[multi_lds_elems()];
  __shared__ unsigned s_ticket;
  __shared__ int s_abort;
  const unsigned nwg = gridDim.x;
  for (unsigned ph = 0; ph < nphase; ++ph) {
    const unsigned gph = phase0 + ph;
    const unsigned base = phase_start[gph];
    const unsigned count = phase_start[gph + 1] - base;
    if (chunk == 0) {  // static round-robin: tile t to workgroup t mod grid, no tickets
      for (unsigned t = blockIdx.x; t < count; t += nwg) {
        __syncthreads();  // the previous tile's LDS reads are done
        const MultiItem it = items[base + t];
        const MultiPass* mp = passes + it.pass;
        dispatch_path<SPLIT>(mp->path, mp->a, it.block, lds);
      }
    }
    for (; chunk != 0;) {
      __syncthreads();  // the previous tile's LDS reads and s_ticket reads are done
      // the counter only grows: once a plain read shows the phase drained, no
      // ticket is taken (a stale read just falls through to the atomic)
      if (threadIdx.x == 0)
        s_ticket = __hip_atomic_load(&ctl->head[gph], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= count
                       ? count
                       : atomicAdd(&ctl->head[gph], chunk);
      __syncthreads();
      const unsigned t0 = s_ticket;
      if (t0 >= count) break;
      const unsigned t1 = min(t0 + chunk, count);
      for (unsigned t = t0; t < t1; ++t) {
        if (t != t0) __syncthreads();
        const MultiItem it = items[base + t];
        const MultiPass* mp = passes + it.pass;
        dispatch_path<SPLIT>(mp->path, mp->a, it.block, lds);
      }
    }
    if (ph + 1 < nphase) {
      // grid barrier between phases: the waves' stores are ordered before a
      // workgroup barrier, then the first wave alone takes the agent-scope
      // release and, after the wait, the acquire (on this chip each writes
      // back or invalidates the L2, so one per workgroup instead of one per wave)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __syncthreads();
      if (threadIdx.x < 64) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      if (threadIdx.x == 0) {
        // two levels: a workgroup arrives on its group's counter; the last of
        // its group in this phase arrives on the grid counter, which every
        // workgroup polls (one read-modify-write per group on the polled line)
        const unsigned grp = blockIdx.x % MULTI_BARRIER_GROUPS;
        const unsigned gsize = (nwg - grp + MULTI_BARRIER_GROUPS - 1) / MULTI_BARRIER_GROUPS;
        const unsigned prev = __hip_atomic_fetch_add(&ctl->arrive[grp][0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev + 1 == (ph + 1) * gsize)
          __hip_atomic_fetch_add(&ctl->barrier, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned ngroups = nwg < (unsigned)MULTI_BARRIER_GROUPS ? nwg : (unsigned)MULTI_BARRIER_GROUPS;
        const unsigned target = (ph + 1) * ngroups;
        unsigned polls = 0;
        int abort = 0;
        while (__hip_atomic_load(&ctl->barrier, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
          if (++polls > spin_limit) {
            __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            abort = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        s_abort = abort;
      }
      if (threadIdx.x < 64) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __syncthreads();
      if (s_abort) return;
    }
  }
}

void preload_multi() { preload_kernel(reinterpret_cast<const void*>(&k_multi_reset)); }

int reset_multi_ctl(MultiCtl* ctl, void* stream) {
  FFTAMD_LAUNCH(k_multi_reset, dim3(1), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), ctl);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_multi(const MultiPass* passes, const MultiItem* items, const unsigned* phase_start,
                 unsigned phase0, unsigned nphase, MultiCtl* ctl, unsigned grid, unsigned chunk,
                 unsigned spin_limit, void* stream, bool split) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  FFTAMD_LAUNCH(k_multi_reset, dim3(1), dim3(64), 0, s, ctl);
  if (split)
    FFTAMD_LAUNCH(k_fft_multi<true>, dim3(grid), dim3(MULTI_THREADS), 0, s, passes, items, phase_start,
                       phase0, nphase, ctl, chunk, spin_limit);
  else
    FFTAMD_LAUNCH(k_fft_multi<false>, dim3(grid), dim3(MULTI_THREADS), 0, s, passes, items,
                       phase_start, phase0, nphase, ctl, chunk, spin_limit);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int multi_max_blocks_per_cu(int* out, bool split) {
  int n = 0;
  hipError_t e = split ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fft_multi<true>, MULTI_THREADS, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fft_multi<false>, MULTI_THREADS, 0);
  if (e != hipSuccess) return -1;
  *out = n;
  return 0;
}

}  // namespace fftamd

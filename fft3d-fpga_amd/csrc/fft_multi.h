// fft_multi.h -- host/device types of the persistent multi-transform kernel.
#pragma once
#include "fft_pass.h"

namespace fftamd {

constexpr int MULTI_MAX_PHASES = 8;
constexpr int MULTI_THREADS = 256;
constexpr int MULTI_MAX_LG = 8;  // lines of up to 256 points per pass

// One pass of one transform inside a multi-transform launch.
struct MultiPass {
  PassArgs a;
  int path;  // mode * MULTI_MAX_LG + (lg - 1)
  int pad;
};

// Work-queue and barrier counters, zeroed before every launch.  The ticket
// counters, the barrier counter and the error flag sit on separate 256-byte
// lines, so that barrier polls do not queue behind ticket atomics.
constexpr int MULTI_BARRIER_GROUPS = 8;  // workgroups b with b % 8 = g arrive on counter g

struct MultiCtl {
  unsigned head[MULTI_MAX_PHASES];  // next item of each phase
  unsigned pad0[64 - MULTI_MAX_PHASES];
  unsigned arrive[MULTI_BARRIER_GROUPS][64];  // per group: arrivals over all phases (own line each)
  unsigned barrier;                 // groups whose last workgroup finished a phase
  unsigned pad1[63];
  unsigned err;                     // set when a barrier wait exceeded its bound
  unsigned pad2[63];
};

struct MultiItem {
  unsigned pass;   // index into the MultiPass array
  unsigned block;  // tile of that pass
};

// Launch the reset kernel and then the persistent kernel (grid workgroups,
// phases [phase0, phase0 + nphase), `chunk` items per queue ticket).
int launch_multi(const MultiPass* passes, const MultiItem* items, const unsigned* phase_start,
                 unsigned phase0, unsigned nphase, MultiCtl* ctl, unsigned grid, unsigned chunk,
                 unsigned spin_limit, void* stream, bool split);
// Zero the control block (a kernel, so the call may be captured in a graph).
int reset_multi_ctl(MultiCtl* ctl, void* stream);
// Resident 256-thread workgroups per CU for the persistent kernel.
int multi_max_blocks_per_cu(int* out, bool split);

}  // namespace fftamd

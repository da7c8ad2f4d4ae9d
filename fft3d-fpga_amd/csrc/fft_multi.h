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

// Work-queue and barrier counters, zeroed before every launch.
struct MultiCtl {
  unsigned head[MULTI_MAX_PHASES];  // next item of each phase
  unsigned barrier;                 // workgroups that finished a phase
  unsigned err;                     // set when a barrier wait exceeded its bound
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

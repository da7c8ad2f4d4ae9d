// fft_multi.h -- host/device types of the persistent multi-transform kernel.
#pragma once
#include "fft_pass.h"
#include "fft_plane.h"

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
  unsigned err;                     // set when a barrier or dependency wait exceeded its bound
  unsigned pad2[63];
  unsigned qhead[8][64];            // dataflow launch: next item of each XCD's queue (own line each)
};

struct MultiItem {
  unsigned pass;   // index into the MultiPass array
  unsigned block;  // tile of that pass
};

// ---- dataflow launch (fft_flow.hip): one persistent launch over every tile
// of both passes of every transform, no grid barrier.  A transform's second
// pass tiles wait only for its own first-pass tiles (a counter per
// transform).  Tiles sit in FLOW_QUEUES queues, one per XCD (a workgroup
// starts on queue blockIdx % 8, the XCD the dispatcher places it on, then
// helps the others); a transform's tiles all sit in one queue, its
// second-pass tiles after all its first-pass tiles, and a queue is consumed
// in order by atomic tickets, so a waiting tile only waits for tiles that
// running workgroups hold: the launch finishes with any number of resident
// workgroups.
constexpr int FLOW_THREADS = 512;
constexpr int FLOW_QUEUES = 8;
constexpr int FLOW_DONE_STRIDE = 16;  // counters 64 bytes apart

// One group of transforms of one shape n^3 (n = 32, 64 or 128).
struct FlowGroup {
  PlaneArgs pa;     // n = 32, 64: x and y of every plane (first pass)
  PassArgs za;      // n = 32, 64: z lines, in place on the output (second pass)
  CubeArgs ca, cb;  // n = 128: the y-split pair's pass A (first) and pass B (second)
  int n;
  unsigned dep0;    // counter of the group's first transform
  unsigned need;    // first-pass tiles per transform
  unsigned pad;
};

// The launch's groups, passed by value (pointers in kernel arguments are
// known to the compiler as global memory): grp[0] n = 32, grp[1] n = 64,
// grp[2] n = 128; a group without transforms is never referenced.
struct FlowArgs {
  FlowGroup grp[3];
};

// A tile of FLOW_THREADS threads: n = 32 eight planes, n = 64 two planes or
// two z-line blocks, n = 128 two pass-A blocks or one pass-B block.
struct FlowItem {
  unsigned tr;          // transform within its group
  unsigned short sub;   // tile of that transform's pass
  unsigned char group;  // 0: n = 32, 1: n = 64, 2: n = 128
  unsigned char pass;   // 0 first pass, 1 second pass
};

// Tiles per transform and pass (host side of the kernel's tile map).
inline unsigned flow_tiles(int n, int pass) {
  switch (n) {
    case 32: return 4;   // 32 planes, 8 a tile; 8 z-line blocks of 128 columns, 2 a tile
    case 64: return 32;  // 64 planes, 2 a tile; 64 z-line blocks of 64 columns, 2 a tile
    default: return pass ? 128u : 256u;  // 128 pass-B blocks, 1 a tile; 512 pass-A blocks, 2 a tile
  }
}

// Zero the control block and `ndone` counters, then launch the dataflow
// kernel (grid workgroups, `chunk` tiles per ticket).
int launch_flow(const FlowArgs& fa, const FlowItem* items, const unsigned* qstart, MultiCtl* ctl,
                unsigned* done, unsigned ndone, unsigned grid, unsigned chunk, unsigned spin_limit,
                bool inverse, void* stream);
// Resident dataflow workgroups per CU.
int flow_max_blocks_per_cu(int* out);

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

// engine.h -- internal interface between the C ABI (api.cpp, which sees the
// drop-in float2/double2 structs) and the HIP engine (which sees HIP's own
// vector types).  Only plain types cross it.
#pragma once
#include <cstddef>
#include <cstdint>

namespace fftamd {

enum : unsigned {
  FLAG_BITREV_1D = 1u,   // leave 1D output bit-reversed within each transform
  FLAG_SPLIT_MFMA = 2u,  // radix-16 stages on FP16 MFMA with hi+lo operands
  FLAG_NO_PLANE_FUSION = 4u,  // separate line passes instead of the plane and 4096-point column passes
  FLAG_NO_CUBE_SPLIT = 8u,    // 256^3 blocks: the 256 x 256 plane pass and a z pass instead of the y-split pair
};

struct Timing {
  double h2d_ms = 0, exec_ms = 0, d2h_ms = 0, stage_in_ms = 0, stage_out_ms = 0;
};

// ---- device / runtime ----------------------------------------------------------
int runtime_device_count();            // visible HIP devices (0 when none)
bool runtime_can_map_host(int dev);    // host-mapped (zero-copy) memory support
int runtime_init(int dev, bool svm);   // 0 on success
void runtime_final();
bool runtime_ready();
bool runtime_svm();
int runtime_ensure_device();           // lazy init for the device-pointer API
const char* last_error();
void set_error(const char* fmt, ...);

// ---- transforms ------------------------------------------------------------------
// prec: 0 = FP32 (8-byte points), 1 = FP64 (16-byte points)
// Host pointers, synchronous, timed: H2D, execution, D2H.
int host_c2c(int rank, const size_t* dims, size_t batch, const void* hin, void* hout, bool inverse,
             unsigned flags, int prec, Timing* t);
// Host pointers through host-mapped staging buffers (the *_svm entry points).
int host_c2c_mapped(int rank, const size_t* dims, size_t batch, const void* hin, void* hout,
                    bool inverse, unsigned flags, Timing* t);
// Host pointers, `count` transforms pipelined in chunks across copy and compute streams.
int host_c2c_pipelined(int rank, const size_t* dims, size_t count, const void* hin, void* hout,
                       bool inverse, Timing* t);
// Device pointers on the caller's stream (nullptr = the HIP null stream); asynchronous.
int dev_c2c(int rank, const size_t* dims, size_t batch, const void* din, void* dout, bool inverse,
            unsigned flags, int prec, void* stream);

// Bytes of the static device array behind dev_c2c_noalloc: every FP32 twiddle
// table (about 150 KiB) and the temporary buffer of in-place shapes.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
// FP32 dev_c2c that allocates no device memory on any call: twiddle tables and
// the temporary buffer come from that static array (see fftfpga_hip.h).
int dev_c2c_noalloc(int rank, const size_t* dims, size_t batch, const void* din, void* dout, bool inverse,
                    unsigned flags, void* stream);

// FP32 device transform with real weights (one per point of a transform,
// broadcast over the batch) fused into the first pass's load (w_load) and the
// last pass's store (w_store); either may be null.
int dev_c2c_weighted(int rank, const size_t* dims, size_t batch, const void* din, void* dout,
                     bool inverse, const float* w_load, const float* w_store, void* stream);

// ---- multi-transform plans (uneven sizes, one persistent launch) ---------------
int multi_create(void** handle, unsigned count, const unsigned* dims3, const unsigned long long* in_off,
                 const unsigned long long* out_off, const void* din, void* dout, bool inverse,
                 unsigned flags);
int multi_exec(void* handle, void* stream);
int multi_check(void* handle, void* stream);
void multi_destroy(void* handle);
int multi_tune(void* handle, void* stream, const char* cache_path, int reps, double* best_us);
int multi_get_schedule(void* handle, int* mode, unsigned* grid, unsigned* chunk, unsigned* items,
                       unsigned* nphase);
int multi_set_schedule(void* handle, int mode, unsigned grid, unsigned chunk);
void multi_set_spin_limit(unsigned polls);
int graph_capture_multi(void** out, void** handles, int n, void* stream);
int graph_launch(void* graph, void* stream);
void graph_destroy(void* graph);

// Host-only plan description (see engine.hip); returns the pass count or -1.
int describe_plan(int rank, const size_t* dims, size_t batch, unsigned flags, int prec, bool aliased,
                  long long* out, int cap);

// Host-only description of the dataflow schedule's queues (see engine.hip);
// returns the tile count or -1.
int flow_describe(unsigned count, const unsigned* dims3, unsigned* qstart, unsigned* out, unsigned cap);

// Pass profiling: while on, every pass launch is bracketed by HIP events on
// its stream; profile_end synchronises and writes a JSON summary
// {label: {count, total_ms}}.
int profile_begin(int max_records);
int profile_end(char* out, int cap);

void plan_cache_clear();
size_t plan_cache_size();

}  // namespace fftamd

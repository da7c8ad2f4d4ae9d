// ext_api.cpp -- device-pointer extension entry points (fftfpga_hip.h).
#include "fftfpga/fftfpga_hip.h"

#include <vector>

#include "engine.h"

namespace {

bool dims_ok(int rank, const unsigned* dims, unsigned batch, size_t* out) {
  if (rank < 1 || rank > 3 || dims == nullptr || batch == 0) {
    fftamd::set_error("bad rank, dims or batch");
    return false;
  }
  for (int d = 0; d < rank; ++d) {
    if (dims[d] == 0 || (dims[d] & (dims[d] - 1)) != 0) {
      fftamd::set_error("dimension %u is not a power of two", dims[d]);
      return false;
    }
    out[d] = dims[d];
  }
  return true;
}

}  // namespace

extern "C" {

int fftamd_c2c_dev(int rank, const unsigned* dims, unsigned batch, const float2* d_in, float2* d_out,
                   int inverse, unsigned flags, void* stream) {
  size_t d[3];
  if (!dims_ok(rank, dims, batch, d)) return -1;
  return fftamd::dev_c2c(rank, d, batch, d_in, d_out, inverse != 0, flags, 0, stream);
}

int fftamd_c2c_dev_noalloc(int rank, const unsigned* dims, unsigned batch, const float2* d_in, float2* d_out,
                           int inverse, unsigned flags, void* stream) {
  size_t d[3];
  if (!dims_ok(rank, dims, batch, d)) return -1;
  return fftamd::dev_c2c_noalloc(rank, d, batch, d_in, d_out, inverse != 0, flags, stream);
}

int fftamd_z2z_dev(int rank, const unsigned* dims, unsigned batch, const double2* d_in,
                   double2* d_out, int inverse, unsigned flags, void* stream) {
  size_t d[3];
  if (!dims_ok(rank, dims, batch, d)) return -1;
  return fftamd::dev_c2c(rank, d, batch, d_in, d_out, inverse != 0, flags, 1, stream);
}

int fftamd_c2c_weighted_dev(int rank, const unsigned* dims, unsigned batch, const float2* d_in,
                            float2* d_out, int inverse, const float* w_in, const float* w_out,
                            void* stream) {
  size_t d[3];
  if (!dims_ok(rank, dims, batch, d)) return -1;
  return fftamd::dev_c2c_weighted(rank, d, batch, d_in, d_out, inverse != 0, w_in, w_out, stream);
}

int fftfpgaf_c2c_3d_ddr_dev(unsigned N, const float2* d_in, float2* d_out, bool inv,
                            unsigned how_many, void* stream) {
  if (d_in == nullptr || d_out == nullptr || N == 0 || (N & (N - 1)) != 0 || how_many == 0) {
    fftamd::set_error("bad arguments");
    return -1;
  }
  size_t d[3] = {N, N, N};
  return fftamd::dev_c2c(3, d, how_many, d_in, d_out, inv, 0u, 0, stream);
}

int fftamd_multi_create(fftamd_multi_t* plan, unsigned count, const fftamd_xform3d* xf,
                        const float2* d_in, float2* d_out, int inverse, unsigned flags) {
  if (!plan || !xf || count == 0) {
    fftamd::set_error("bad multi-transform arguments");
    return -1;
  }
  std::vector<unsigned> dims(3 * (size_t)count);
  std::vector<unsigned long long> io(count), oo(count);
  for (unsigned i = 0; i < count; ++i) {
    if (xf[i].reserved != 0) {
      fftamd::set_error("fftamd_xform3d.reserved must be 0");
      return -1;
    }
    for (int a = 0; a < 3; ++a) dims[3 * i + a] = xf[i].n[a];
    io[i] = xf[i].in_offset;
    oo[i] = xf[i].out_offset;
  }
  void* h = nullptr;
  if (fftamd::multi_create(&h, count, dims.data(), io.data(), oo.data(), d_in, d_out, inverse != 0, flags))
    return -1;
  *plan = reinterpret_cast<fftamd_multi_t>(h);
  return 0;
}

int fftamd_multi_exec(fftamd_multi_t plan, void* stream) { return fftamd::multi_exec(plan, stream); }

int fftamd_multi_check(fftamd_multi_t plan, void* stream) { return fftamd::multi_check(plan, stream); }

int fftamd_multi_tune(fftamd_multi_t plan, void* stream, const char* cache_path, int reps,
                      double* best_us) {
  return fftamd::multi_tune(plan, stream, cache_path, reps, best_us);
}

int fftamd_multi_schedule(fftamd_multi_t plan, int* mode, unsigned* grid, unsigned* chunk,
                          unsigned* items, unsigned* phases) {
  return fftamd::multi_get_schedule(plan, mode, grid, chunk, items, phases);
}

int fftamd_multi_set_schedule(fftamd_multi_t plan, int mode, unsigned grid, unsigned chunk) {
  return fftamd::multi_set_schedule(plan, mode, grid, chunk);
}

void fftamd_multi_destroy(fftamd_multi_t plan) { fftamd::multi_destroy(plan); }

void fftamd_multi_set_spin_limit(unsigned polls) { fftamd::multi_set_spin_limit(polls); }

int fftamd_graph_capture(fftamd_graph_t* graph, fftamd_multi_t* plans, int n, void* stream) {
  if (!graph || !plans) {
    fftamd::set_error("bad graph arguments");
    return -1;
  }
  std::vector<void*> hs(plans, plans + (n > 0 ? n : 0));
  void* g = nullptr;
  if (fftamd::graph_capture_multi(&g, hs.data(), n, stream)) return -1;
  *graph = reinterpret_cast<fftamd_graph_t>(g);
  return 0;
}

int fftamd_graph_launch(fftamd_graph_t graph, void* stream) { return fftamd::graph_launch(graph, stream); }

void fftamd_graph_destroy(fftamd_graph_t graph) { fftamd::graph_destroy(graph); }

const char* fftamd_last_error(void) { return fftamd::last_error(); }

void fftamd_plan_cache_clear(void) { fftamd::plan_cache_clear(); }

size_t fftamd_plan_cache_size(void) { return fftamd::plan_cache_size(); }

int fftamd_profile_begin(int max_records) { return fftamd::profile_begin(max_records); }

int fftamd_profile_end(char* json_out, int cap) { return fftamd::profile_end(json_out, cap); }

int fftamd_describe_plan(int rank, const unsigned* dims, unsigned batch, unsigned flags, int prec,
                         int aliased, long long* out, int cap) {
  size_t d[3];
  if (!dims_ok(rank, dims, batch, d)) return -1;
  return fftamd::describe_plan(rank, d, batch, flags, prec, aliased != 0, out, cap);
}

int fftamd_describe_flow(unsigned count, const unsigned* dims3, unsigned* qstart, unsigned* out,
                         unsigned cap) {
  return fftamd::flow_describe(count, dims3, qstart, out, cap);
}

}  // extern "C"

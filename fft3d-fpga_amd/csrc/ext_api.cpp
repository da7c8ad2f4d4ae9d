// ext_api.cpp -- device-pointer extension entry points (fftfpga_hip.h).
#include "fftfpga/fftfpga_hip.h"

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

}  // extern "C"

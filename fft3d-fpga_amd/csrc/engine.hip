// engine.hip -- HIP runtime layer: device state, twiddle tables, the
// axis-pass planner and the host/device transform drivers.
//
// Replaces the reference's OpenCL plumbing (api/src/fftfpga.c:72-188,
// api/src/opencl_utils.c) and the per-call kernel/queue/buffer setup of
// api/src/fft{1d,2d,3d}.c with: one device, one in-order stream, grow-only
// HBM staging buffers, plans cached per shape, and twiddle tables generated
// once in FP64 and stored in the transform precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <numeric>
#include <sstream>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "engine.h"
#include "fft_multi.h"
#include "fft_pass.h"
#include "fft_plane.h"
#include "fft_schedule.h"
#include "launch.h"

namespace fftamd {

bool& resolve_only_flag() {
  static thread_local bool on = false;
  return on;
}

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  fprintf(stderr, "fftfpga(hip): %s\n", g_err);
}
const char* last_error() { return g_err; }

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return -1;                                                                       \
    }                                                                                  \
  } while (0)

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
enum Buf { B_IN = 0, B_OUT = 1, B_TMP = 2 };

struct PassDesc {
  int lg = 0, mode = MODE_CONTIG, epi = EPI_NONE;
  bool transposing = false;  // writes a different index set than it reads
  long long in_p = 1, out_p = 1, in_c = 0, out_c = 0, in_k = 0, out_k = 0, in_o = 0, out_o = 0;
  unsigned ncol = 1, K = 1, kbits = 0;
  unsigned long long nouter = 1;
  long long M = 0;
  unsigned cshift = 0;
  int src = B_IN, dst = B_OUT;
};

struct Plan {
  int prec = PREC_F32;
  bool split = false;  // radix-16 stages on FP16 matrix cores (hi+lo operands)
  std::vector<PassDesc> passes;
  size_t total = 0;  // points including batch
  size_t per = 0;    // points of one transform
  bool final_copy = false;
  bool needs_tmp = false;
  void* tmp = nullptr;
  // allocation-free plan: twiddle tables and the temporary buffer live in the
  // device's static scratch array (never freed, never allocated)
  bool noalloc = false;
};

// (rank, dims padded with 0, batch, flags, precision, aliased, allocation-free):
// trivially copyable, so a lookup allocates nothing
using PlanKey = std::tuple<int, size_t, size_t, size_t, size_t, unsigned, int, bool, bool>;

struct State {
  std::recursive_mutex mu;
  bool ready = false;
  bool svm = false;
  int dev = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  void* d_in = nullptr;
  size_t d_in_cap = 0;
  void* d_out = nullptr;
  size_t d_out_cap = 0;
  void* h_stage = nullptr;  // host-mapped staging (svm variants)
  size_t h_stage_cap = 0;
  std::map<PlanKey, Plan> plans;
  std::map<std::pair<int, int>, void*> tw_stockham;  // (prec, L)
  std::map<std::pair<int, long long>, std::pair<void*, void*>> tw_fourstep;  // (prec, M)
  std::map<long long, void*> tw_direct;  // FP32 W_M^t, t = 0..M-1
  // bumped whenever tables and plans are released (fpga_final, re-init):
  // multi-transform plans and graphs made before hold freed table pointers
  unsigned long long gen = 1;
  struct Pipe* pipe = nullptr;  // host batch pipeline buffers (grow-only)
};

static State& S() {
  static State s;
  return s;
}

// ---------------------------------------------------------------------------
// pass profiling: HIP events around every pass launch, on the launch stream
// ---------------------------------------------------------------------------
struct ProfRec {
  std::string label;
  hipEvent_t a, b;
};
struct Prof {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<ProfRec> recs;
};
static Prof& P() {
  static Prof p;
  return p;
}

// ---------------------------------------------------------------------------
// twiddles: W_M^t = exp(-2 pi i t / M), evaluated in long double with exact
// integer phase reduction, rounded once to the storage precision.
// ---------------------------------------------------------------------------
static void twiddle_ld(long long t, long long M, long double* c, long double* s) {
  t %= M;
  if (t < 0) t += M;
  // exact quarter-turn symmetries keep 1, -1, i, -i exact
  long long q = (4 * t) / M;
  if ((4 * t) % M == 0) {
    static const long double cs[4][2] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
    *c = cs[q][0];
    *s = cs[q][1];
    return;
  }
  const long double twopi = 6.283185307179586476925286766559L;
  long double a = twopi * (long double)t / (long double)M;
  *c = cosl(a);
  *s = -sinl(a);
}

template <typename R>
static std::vector<R> stockham_table(int L) {
  std::vector<R> v(2 * (size_t)std::max(1, sched_twsize(L)));
  const int np = sched_np(L);
  for (int st = 1; st < np; ++st) {
    const int rad = sched_radix(L, st), ns = sched_ns(L, st), off = sched_twoff(L, st);
    for (int k = 0; k < ns; ++k)
      for (int q = 0; q < TW_PER_K; ++q) {  // W_{16 NS}^{k 2^q}
        long double c, s;
        twiddle_ld((long long)k << q, (long long)ns * rad, &c, &s);
        size_t i = (size_t)off + (size_t)TW_PER_K * k + q;
        v[2 * i] = (R)c;
        v[2 * i + 1] = (R)s;
      }
  }
  return v;
}

static int get_stockham(int prec, int L, const void** out) {
  State& st = S();
  auto key = std::make_pair(prec, L);
  auto it = st.tw_stockham.find(key);
  if (it != st.tw_stockham.end()) {
    *out = it->second;
    return 0;
  }
  void* d = nullptr;
  if (prec == PREC_F32) {
    auto h = stockham_table<float>(L);
    HIP_TRY(hipMalloc(&d, h.size() * sizeof(float)));
    HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  } else {
    auto h = stockham_table<double>(L);
    HIP_TRY(hipMalloc(&d, h.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  st.tw_stockham[key] = d;
  *out = d;
  return 0;
}

static unsigned fourstep_hb(long long M) { return (unsigned)((ilog2c(M) + 1) / 2); }

template <typename R>
static void fourstep_tables(long long M, std::vector<R>& lo, std::vector<R>& hi) {
  const unsigned hb = fourstep_hb(M);
  const long long nlo = 1LL << hb, nhi = M >> hb;
  lo.resize(2 * nlo);
  hi.resize(2 * nhi);
  for (long long i = 0; i < nlo; ++i) {
    long double c, s;
    twiddle_ld(i, M, &c, &s);
    lo[2 * i] = (R)c;
    lo[2 * i + 1] = (R)s;
  }
  for (long long j = 0; j < nhi; ++j) {
    long double c, s;
    twiddle_ld(j << hb, M, &c, &s);
    hi[2 * j] = (R)c;
    hi[2 * j + 1] = (R)s;
  }
}

static int get_fourstep(int prec, long long M, const void** lo, const void** hi) {
  State& st = S();
  auto key = std::make_pair(prec, M);
  auto it = st.tw_fourstep.find(key);
  if (it == st.tw_fourstep.end()) {
    void *dl = nullptr, *dh = nullptr;
    if (prec == PREC_F32) {
      std::vector<float> l, h;
      fourstep_tables(M, l, h);
      HIP_TRY(hipMalloc(&dl, l.size() * sizeof(float)));
      HIP_TRY(hipMalloc(&dh, h.size() * sizeof(float)));
      HIP_TRY(hipMemcpy(dl, l.data(), l.size() * sizeof(float), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(dh, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    } else {
      std::vector<double> l, h;
      fourstep_tables(M, l, h);
      HIP_TRY(hipMalloc(&dl, l.size() * sizeof(double)));
      HIP_TRY(hipMalloc(&dh, h.size() * sizeof(double)));
      HIP_TRY(hipMemcpy(dl, l.data(), l.size() * sizeof(double), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(dh, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    it = st.tw_fourstep.emplace(key, std::make_pair(dl, dh)).first;
  }
  *lo = it->second.first;
  *hi = it->second.second;
  return 0;
}

// FP32 table of W_M^t for every t < M (one load per twiddle)
static int get_direct(long long M, const void** out) {
  State& st = S();
  auto it = st.tw_direct.find(M);
  if (it == st.tw_direct.end()) {
    std::vector<float> h(2 * (size_t)M);
    for (long long t = 0; t < M; ++t) {
      long double c, s;
      twiddle_ld(t, M, &c, &s);
      h[2 * t] = (float)c;
      h[2 * t + 1] = (float)s;
    }
    void* d = nullptr;
    HIP_TRY(hipMalloc(&d, h.size() * sizeof(float)));
    HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    it = st.tw_direct.emplace(M, d).first;
  }
  *out = it->second;
  return 0;
}

// ---------------------------------------------------------------------------
// allocation-free entry point (fftamd_c2c_dev_noalloc): a static device array
// holds every FP32 twiddle table a plan can read and the temporary buffer of
// in-place shapes, so a call allocates no device memory.  In host code the
// array's name is only a handle for hipGetSymbolAddress: the address kept per
// device (na_dev[dev].base) is what kernels and copies receive.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct NoallocDev {
  unsigned char* base = nullptr;  // device address of `scratch` on this device
  size_t stock[14] = {};          // byte offset of the Stockham table of 2^lg points
  size_t fs_lo[23] = {}, fs_hi[23] = {};  // four-step tables of modulus 2^lg
  size_t w256 = 0, w512 = 0;      // W256^t, t < 256; W512^t, t < 512
  size_t tmp_off = 0;             // the temporary buffer: [tmp_off, NOALLOC_SCRATCH_BYTES)
  bool ready = false;
};
constexpr int NOALLOC_MAX_DEV = 64;
static NoallocDev na_dev[NOALLOC_MAX_DEV];

// Once per device, with that device current: resolve the array's address and
// write every table into it (the same long-double tables as the allocating
// path, so both paths give the same bits).
static int noalloc_setup(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) {
    set_error("device %d out of range for the allocation-free scratch", dev);
    return -1;
  }
  NoallocDev& d = na_dev[dev];
  if (d.ready) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    return -1;
  }
  std::vector<float> img;
  auto put = [&](const std::vector<float>& v) {
    img.resize((img.size() + 63) / 64 * 64, 0.f);  // 256-byte aligned tables
    const size_t off = img.size() * sizeof(float);
    img.insert(img.end(), v.begin(), v.end());
    return off;
  };
  for (int lg = 1; lg <= 13; ++lg)
    if (sched_np(1 << lg) > 1) d.stock[lg] = put(stockham_table<float>(1 << lg));
  for (int lg = 2; lg <= 22; ++lg) {
    std::vector<float> lo, hi;
    fourstep_tables<float>(1LL << lg, lo, hi);
    d.fs_lo[lg] = put(lo);
    d.fs_hi[lg] = put(hi);
  }
  for (int M : {256, 512}) {
    std::vector<float> w(2 * (size_t)M);
    for (int t = 0; t < M; ++t) {
      long double c, sn;
      twiddle_ld(t, M, &c, &sn);
      w[2 * t] = (float)c;
      w[2 * t + 1] = (float)sn;
    }
    (M == 256 ? d.w256 : d.w512) = put(w);
  }
  const size_t bytes = img.size() * sizeof(float);
  d.tmp_off = (bytes + 255) / 256 * 256;
  if (d.tmp_off >= NOALLOC_SCRATCH_BYTES) {
    set_error("allocation-free scratch too small for the twiddle tables");
    return -1;
  }
  HIP_TRY(hipMemcpy(p, img.data(), bytes, hipMemcpyHostToDevice));
  d.base = static_cast<unsigned char*>(p);
  d.ready = true;
  return 0;
}

// ---------------------------------------------------------------------------
// planner: one or two passes per axis, last (contiguous) axis first
// ---------------------------------------------------------------------------
static bool is_pow2(size_t x) { return x && !(x & (x - 1)); }

static int split_contig(int prec, int lg, int* lg1, int* lg2) {
  // L = L1 (strided pass with the four-step twiddle) * L2 (contiguous pass);
  // 2^22 as 1024 x 4096 (the contiguous pass then runs four lines per CU and
  // three stages: 38.1 against 41.4 us for 512 x 8192)
  const int mc = ilog2c(lmax_contig(prec)), ms = ilog2c(lmax_strided_fourstep(prec));
  int b = std::min(mc, lg - 1);
  if (prec == PREC_F32 && lg == 22) b = 12;
  if (lg - b > ms) return -1;
  *lg2 = b;
  *lg1 = lg - b;
  return 0;
}

static int split_strided(int prec, int lg, int* lg1, int* lg2) {
  const int ms = ilog2c(lmax_strided(prec));
  int a = (lg + 1) / 2;
  if (a > ms || lg - a > ms) return -1;
  *lg1 = a;
  *lg2 = lg - a;
  return 0;
}

// Fewest planes for which a plane pass replaces the separate x and y passes:
// any count for planes up to 128 x 128 (one workgroup per plane, measured
// faster from one 64 x 64 plane up); 32 for 256 x 256 planes (four workgroups
// per plane; with 16 or fewer planes the x and y line passes are faster,
// measured 13.9 against 16.9 us at 16 planes, 20.6 against 18.1 at 32).
static constexpr unsigned long long PLANE_MIN_COUNT = 1;
static constexpr unsigned long long PLANE256_MIN_COUNT = 32;
// Fewest 16-column groups for which the one-pass 4096-point column transform
// replaces the four-step pair (four workgroups per group, as the 256 x 256
// plane pass): measured 29.5 against 31.2 us at 32 groups (4096 x 512), 39.9
// against 32.9 us at 64 (4096 x 1024), 137.6 against 103.3 us at 4096 x 4096.
static constexpr unsigned long long COL4096_MIN_GROUPS = 64;

static int build_plan(int rank, const size_t* dims, size_t batch, unsigned flags, int prec,
                      bool aliased, Plan& plan) {
  plan = Plan();
  plan.prec = prec;
  plan.split = (flags & FLAG_SPLIT_MFMA) != 0;
  if (plan.split && prec != PREC_F32) {
    set_error("split-precision mode is defined for single precision only");
    return -1;
  }
  size_t tot = 1;
  for (int d = 0; d < rank; ++d) tot *= dims[d];
  plan.total = tot * batch;
  plan.per = tot;
  const bool bitrev = (flags & FLAG_BITREV_1D) && rank == 1;
  std::vector<PassDesc> ps;

  // n x n x n blocks, n = 512, 256 or 128 (the three fastest axes): two
  // passes with the y axis split between them (64 x 8 for 512, (n/4) x 4
  // otherwise) -- x and the larger y stage (out of place), then the smaller y
  // stage and z (in place)
  int top = rank - 1;
  const long long CN = rank >= 3 ? (long long)dims[rank - 1] : 0;
  if (rank >= 3 && prec == PREC_F32 && !(flags & (FLAG_NO_PLANE_FUSION | FLAG_NO_CUBE_SPLIT)) &&
      (CN == 512 || CN == 256 || CN == 128) && (long long)dims[rank - 2] == CN && (long long)dims[rank - 3] == CN) {
    unsigned long long O = batch;
    for (int d = 0; d < rank - 3; ++d) O *= dims[d];
    PassDesc A;
    A.lg = ilog2c(CN);
    A.mode = MODE_CUBE_A;
    A.transposing = true;  // residue a's output rows are other residues' input rows
    A.in_p = A.out_p = 1;
    A.in_c = A.out_c = CN;
    A.ncol = (unsigned)CN;
    A.K = (unsigned)CN;
    A.nouter = O * (unsigned long long)CN;  // planes
    A.in_o = A.out_o = CN * CN;
    ps.push_back(A);
    PassDesc B;
    B.lg = ilog2c(CN);
    B.mode = MODE_CUBE_B;
    B.in_p = B.out_p = CN * CN;
    B.in_c = B.out_c = 1;
    B.ncol = (unsigned)(CN * CN);
    B.nouter = O;  // blocks
    B.in_o = B.out_o = CN * CN * CN;
    ps.push_back(B);
    top = rank - 4;
  }

  // planes of the two fastest axes in one fused x+y pass: 256 x 256 (four
  // workgroups per plane, out of place) or up to 128 x 128 (one workgroup
  // per plane in LDS, in place allowed)
  if (top == rank - 1 && rank >= 2 && prec == PREC_F32 && !(flags & FLAG_NO_PLANE_FUSION)) {
    const long long LY = (long long)dims[rank - 2], LX = (long long)dims[rank - 1];
    unsigned long long O = batch;
    for (int d = 0; d < rank - 2; ++d) O *= dims[d];
    const bool big = LY == 256 && LX == 256;  // no split-precision variant
    if ((big && !plan.split && O >= PLANE256_MIN_COUNT) ||
        (plane_shape_supported(LY, LX) && O >= PLANE_MIN_COUNT)) {
      PassDesc p;
      p.lg = ilog2c(LX);
      p.mode = MODE_PLANE;
      p.transposing = big;  // the four workgroups of a 256 x 256 plane all read all of it
      p.in_p = p.out_p = 1;
      p.in_c = p.out_c = LX;
      p.ncol = (unsigned)LY;
      p.K = (unsigned)LY;
      p.nouter = O;
      p.in_o = p.out_o = LY * LX;
      ps.push_back(p);
      top = rank - 3;
    }
  }

  // a 4096-point second-fastest axis in one pass (four workgroups per group of
  // 16 columns, out of place), placed first so that the contiguous pass after
  // it runs in place on the output
  int col_axis = -1;
  if (rank >= 2 && top == rank - 1 && prec == PREC_F32 && !plan.split && !(flags & FLAG_NO_PLANE_FUSION) &&
      dims[rank - 2] == 4096) {
    const long long LX = (long long)dims[rank - 1];
    unsigned long long O = batch;
    for (int d = 0; d < rank - 2; ++d) O *= dims[d];
    if (col4096_supported(LX, O) && (unsigned long long)(LX / 16) * O >= COL4096_MIN_GROUPS) {
      PassDesc p;
      p.lg = 12;
      p.mode = MODE_COL4096;
      p.transposing = true;  // the four workgroups of a column group all read all of it
      p.in_p = p.out_p = LX;
      p.in_c = p.out_c = 1;
      p.ncol = (unsigned)LX;
      p.K = 1;
      p.nouter = O;
      p.in_o = p.out_o = 4096 * LX;
      ps.push_back(p);
      col_axis = rank - 2;
    }
  }

  for (int a = top; a >= 0; --a) {
    const long long L = (long long)dims[a];
    if (L == 1 || a == col_axis) continue;
    long long s = 1;
    for (int d = a + 1; d < rank; ++d) s *= (long long)dims[d];
    unsigned long long O = batch;
    for (int d = 0; d < a; ++d) O *= dims[d];
    const int lg = ilog2c(L);
    if (s == 1) {
      if (L <= lmax_contig(prec)) {
        if (O >= (1ULL << 32)) {
          set_error("too many lines (%llu) on the contiguous axis", O);
          return -1;
        }
        PassDesc p;
        p.lg = lg;
        p.mode = MODE_CONTIG;
        p.epi = bitrev ? EPI_BITREV : EPI_NONE;
        p.in_p = p.out_p = 1;
        p.in_c = p.out_c = L;
        p.ncol = (unsigned)O;
        ps.push_back(p);
      } else {
        int lg1, lg2;
        if (split_contig(prec, lg, &lg1, &lg2)) {
          set_error("contiguous axis of 2^%d points exceeds the supported length", lg);
          return -1;
        }
        const long long L1 = 1LL << lg1, L2 = 1LL << lg2;
        PassDesc A;  // strided DFT over n1, post-twiddle W_L^{n2 k1}, in place
        A.lg = lg1;
        A.mode = MODE_STRIDED;
        A.epi = EPI_TWIDDLE;
        A.in_p = A.out_p = L2;
        A.in_c = A.out_c = 1;
        A.ncol = (unsigned)L2;
        A.nouter = O;
        A.in_o = A.out_o = L;
        A.M = L;
        A.cshift = 0;
        ps.push_back(A);
        PassDesc B;  // contiguous DFT over n2 for every k1
        B.lg = lg2;
        B.mode = MODE_CONTIG;
        B.transposing = true;
        B.in_p = 1;
        B.in_c = L2;
        B.ncol = (unsigned)L1;
        B.nouter = O;
        B.in_o = B.out_o = L;
        if (bitrev) {  // out[bitrev(k1)*L2 + bitrev(k2)] = X[k1 + L1 k2]
          B.epi = EPI_BITREV;
          B.out_p = 1;
          B.out_c = L2;
          B.K = 1;
          // column bit reversal is expressed through the middle index: one column per K
          B.ncol = 1;
          B.K = (unsigned)L1;
          B.in_k = L2;
          B.out_k = L2;
          B.in_c = B.out_c = 0;
          B.kbits = (unsigned)lg1;
        } else {  // X[k1 + L1 k2] at natural position (strided store)
          B.epi = EPI_NONE;
          B.out_p = L1;
          B.out_c = 1;
        }
        ps.push_back(B);
      }
    } else {
      if (L <= lmax_strided(prec)) {
        PassDesc p;
        p.lg = lg;
        p.mode = MODE_STRIDED;
        p.in_p = p.out_p = s;
        p.in_c = p.out_c = 1;
        if (s >= (1LL << 32)) {
          set_error("axis stride too large");
          return -1;
        }
        p.ncol = (unsigned)s;
        p.nouter = O;
        p.in_o = p.out_o = L * s;
        ps.push_back(p);
      } else {
        int lg1, lg2;
        if (split_strided(prec, lg, &lg1, &lg2)) {
          set_error("strided axis of 2^%d points exceeds the supported length", lg);
          return -1;
        }
        const long long L1 = 1LL << lg1, L2 = 1LL << lg2;
        if (L2 * s >= (1LL << 32)) {
          set_error("axis stride too large");
          return -1;
        }
        PassDesc A;
        A.lg = lg1;
        A.mode = MODE_STRIDED;
        A.epi = EPI_TWIDDLE;
        A.in_p = A.out_p = L2 * s;
        A.in_c = A.out_c = 1;
        A.ncol = (unsigned)(L2 * s);
        A.nouter = O;
        A.in_o = A.out_o = L * s;
        A.M = L;
        A.cshift = (unsigned)ilog2c(s);
        ps.push_back(A);
        PassDesc B;
        B.lg = lg2;
        B.mode = MODE_STRIDED;
        B.transposing = true;
        B.in_p = s;
        B.out_p = L1 * s;
        B.in_c = B.out_c = 1;
        B.ncol = (unsigned)s;
        B.K = (unsigned)L1;
        B.in_k = L2 * s;
        B.out_k = s;
        B.nouter = O;
        B.in_o = B.out_o = L * s;
        ps.push_back(B);
      }
    }
  }

  // buffer roles: the input is read-only unless in and out alias
  auto simulate = [&](int first_dst) {
    int cur = aliased ? B_OUT : B_IN;
    for (size_t i = 0; i < ps.size(); ++i) {
      int dst;
      if (i == 0) dst = first_dst;
      else if (ps[i].transposing) dst = cur == B_OUT ? B_TMP : B_OUT;
      else if (i + 1 == ps.size()) dst = B_OUT;  // a line pass may also run out of place
      else dst = cur;
      ps[i].src = cur;
      ps[i].dst = dst;
      cur = dst;
    }
    return cur;
  };
  plan.final_copy = false;
  if (!ps.empty()) {
    bool done = false;
    // first pass may write OUT unless it is transposing while aliased
    const bool first_to_out_ok = !(aliased && ps[0].transposing);
    if (first_to_out_ok && simulate(B_OUT) == B_OUT) done = true;
    if (!done && simulate(B_TMP) == B_OUT) done = true;
    if (!done) {
      simulate(first_to_out_ok ? B_OUT : B_TMP);
      plan.final_copy = true;
    }
  } else {
    plan.final_copy = !aliased;  // length-1 transform: identity copy
  }
  plan.passes = ps;
  plan.needs_tmp = false;
  for (auto& p : plan.passes)
    if (p.src == B_TMP || p.dst == B_TMP) plan.needs_tmp = true;
  return 0;
}

// Twiddle tables of a plan: the cached allocations, or the static scratch of
// an allocation-free plan.
static int tab_stockham(const Plan& plan, int L, const void** out) {
  if (!plan.noalloc) return get_stockham(plan.prec, L, out);
  const NoallocDev& d = na_dev[S().dev];
  const int lg = ilog2c(L);
  if (plan.prec != PREC_F32 || lg < 1 || lg > 13 || !d.ready) {
    set_error("no allocation-free Stockham table for %d points", L);
    return -1;
  }
  *out = d.base + d.stock[lg];
  return 0;
}
static int tab_fourstep(const Plan& plan, long long M, const void** lo, const void** hi) {
  if (!plan.noalloc) return get_fourstep(plan.prec, M, lo, hi);
  const NoallocDev& d = na_dev[S().dev];
  const int lg = ilog2c(M);
  if (plan.prec != PREC_F32 || lg < 2 || lg > 22 || !d.ready) {
    set_error("no allocation-free four-step table for modulus %lld", M);
    return -1;
  }
  *lo = d.base + d.fs_lo[lg];
  *hi = d.base + d.fs_hi[lg];
  return 0;
}
static int tab_direct(const Plan& plan, long long M, const void** out) {
  if (!plan.noalloc) return get_direct(M, out);
  const NoallocDev& d = na_dev[S().dev];
  if (!d.ready || (M != 256 && M != 512)) {
    set_error("no allocation-free table of W_%lld", M);
    return -1;
  }
  *out = d.base + (M == 256 ? d.w256 : d.w512);
  return 0;
}

static int run_plan(Plan& plan, const void* din, void* dout, bool inverse, hipStream_t stream,
                    const float* w_load = nullptr, const float* w_store = nullptr);

// Everything a plan's calls need before the first of them: its twiddle tables
// and the code objects of its kernels (both directions), made by running the
// plan once with the launchers in resolve-only mode.
static int plan_prepare(Plan& plan) {
  struct Resolve {
    Resolve() { resolve_only_flag() = true; }
    ~Resolve() { resolve_only_flag() = false; }
  } guard;
  static char dummy[64];
  for (int inv = 0; inv < 2; ++inv)
    if (run_plan(plan, dummy, dummy, inv != 0, nullptr)) return -1;
  return 0;
}

static int get_plan(int rank, const size_t* dims, size_t batch, unsigned flags, int prec, bool aliased,
                    Plan** out, bool noalloc = false) {
  State& st = S();
  PlanKey key(rank, dims[0], rank > 1 ? dims[1] : 0, rank > 2 ? dims[2] : 0, batch,
              flags & (FLAG_BITREV_1D | FLAG_SPLIT_MFMA | FLAG_NO_PLANE_FUSION | FLAG_NO_CUBE_SPLIT), prec,
              aliased, noalloc);
  auto it = st.plans.find(key);
  if (it == st.plans.end()) {
    Plan p;
    if (build_plan(rank, dims, batch, flags, prec, aliased, p)) return -1;
    p.noalloc = noalloc;
    if (p.needs_tmp) {
      const size_t bytes = p.total * (prec == PREC_F32 ? 8 : 16);
      if (noalloc) {
        const NoallocDev& d = na_dev[st.dev];
        if (!d.ready || bytes > NOALLOC_SCRATCH_BYTES - d.tmp_off) {
          set_error("this in-place shape needs a %zu-byte temporary buffer; the allocation-free scratch "
                    "holds %zu bytes (call out of place, or use fftamd_c2c_dev)",
                    bytes, (size_t)(NOALLOC_SCRATCH_BYTES - d.tmp_off));
          return -1;
        }
        p.tmp = d.base + d.tmp_off;
      } else {
        HIP_TRY(hipMalloc(&p.tmp, bytes));
      }
    }
    it = st.plans.emplace(key, std::move(p)).first;
    if (plan_prepare(it->second)) {
      if (it->second.tmp && !it->second.noalloc) (void)hipFree(it->second.tmp);
      st.plans.erase(it);
      return -1;
    }
  }
  *out = &it->second;
  return 0;
}

static int run_plan(Plan& plan, const void* din, void* dout, bool inverse, hipStream_t stream,
                    const float* w_load, const float* w_store) {
  const bool resolve = resolve_only_flag();
  if ((w_load || w_store) && plan.passes.empty()) {
    set_error("weighted transform of a single point is not supported");
    return -1;
  }
  const size_t esz = plan.prec == PREC_F32 ? 8 : 16;
  auto ptr = [&](int b) -> void* {
    return b == B_IN ? const_cast<void*>(din) : (b == B_OUT ? dout : plan.tmp);
  };
  int cur = B_IN;
  for (size_t pi = 0; pi < plan.passes.size(); ++pi) {
    const PassDesc& p = plan.passes[pi];
    if (p.mode == MODE_PLANE) {
      const int LX = 1 << p.lg, LY = (int)p.K;
      PlaneArgs xa{};
      xa.in = ptr(p.src);
      xa.out = ptr(p.dst);
      xa.nplanes = p.nouter;
      xa.plane_stride = p.in_o;
      if (sched_np(LX) > 1 && tab_stockham(plan, LX, &xa.tw_x)) return -1;
      if (sched_np(LY) > 1 && tab_stockham(plan, LY, &xa.tw_y)) return -1;
      xa.w_load = pi == 0 ? w_load : nullptr;
      xa.w_store = pi + 1 == plan.passes.size() ? w_store : nullptr;
      xa.w_mask = (unsigned long long)plan.per - 1ULL;
      xa.inverse = inverse ? 1 : 0;
      Prof& pf = P();
      const bool prof = pf.on && !resolve && pf.used + 2 <= pf.pool.size();
      if (prof) HIP_TRY(hipEventRecord(pf.pool[pf.used], stream));
      const int rc = (LY == 256 && LX == 256) ? launch_fft_xy256(xa, stream)
                                               : launch_fft_plane(LY, LX, xa, stream, plan.split);
      if (prof) {
        HIP_TRY(hipEventRecord(pf.pool[pf.used + 1], stream));
        char lab[64];
        snprintf(lab, sizeof(lab), "fft_plane_f32%s_%dx%d", plan.split ? "_split" : "", LY, LX);
        pf.recs.push_back({lab, pf.pool[pf.used], pf.pool[pf.used + 1]});
        pf.used += 2;
      }
      if (rc) {
        set_error("plane pass launch failed (rc=%d): %s", rc, hipGetErrorString(hipGetLastError()));
        return -1;
      }
      cur = p.dst;
      continue;
    }
    if (p.mode == MODE_CUBE_A || p.mode == MODE_CUBE_B) {
      const bool A = p.mode == MODE_CUBE_A;
      CubeArgs ca{};
      ca.in = ptr(p.src);
      ca.out = ptr(p.dst);
      ca.n = 1 << p.lg;
      ca.nplanes = A ? p.nouter : p.nouter * (unsigned long long)ca.n;
      // 512: pass A's columns are 64-point lines, pass B's z lines 4 x 128
      if (tab_stockham(plan, ca.n == 512 ? (A ? 64 : 128) : ca.n, &ca.tw)) return -1;
      if (tab_direct(plan, 256, &ca.w256)) return -1;
      if (ca.n == 512 && tab_direct(plan, 512, &ca.w512)) return -1;
      // pass A is always the first pass of its plan; pass B takes the store
      // weights when it is the last
      ca.w_load = A && pi == 0 ? w_load : nullptr;
      ca.w_store = !A && pi + 1 == plan.passes.size() ? w_store : nullptr;
      ca.w_mask = (unsigned long long)plan.per - 1ULL;
      ca.inverse = inverse ? 1 : 0;
      ca.split = plan.split ? 1 : 0;
      Prof& pf = P();
      const bool prof = pf.on && !resolve && pf.used + 2 <= pf.pool.size();
      if (prof) HIP_TRY(hipEventRecord(pf.pool[pf.used], stream));
      const int rc = A ? launch_fft_cube_a(ca, stream) : launch_fft_cube_b(ca, stream);
      if (prof) {
        HIP_TRY(hipEventRecord(pf.pool[pf.used + 1], stream));
        char lab[64];
        snprintf(lab, sizeof(lab), "fft_cube_%s_f32%s_%d", A ? "a" : "b", plan.split ? "_split" : "", ca.n);
        pf.recs.push_back({lab, pf.pool[pf.used], pf.pool[pf.used + 1]});
        pf.used += 2;
      }
      if (rc) {
        set_error("y-split pass launch failed (n=%d, rc=%d): %s", ca.n, rc, hipGetErrorString(hipGetLastError()));
        return -1;
      }
      cur = p.dst;
      continue;
    }
    if (p.mode == MODE_COL4096) {
      ColArgs ca{};
      ca.in = ptr(p.src);
      ca.out = ptr(p.dst);
      ca.lx = p.in_p;
      ca.in_o = p.in_o;
      ca.ngroups = (unsigned)(p.ncol / 16);
      ca.nitems = p.nouter * ca.ngroups;
      if (tab_stockham(plan, 256, &ca.tw)) return -1;
      if (tab_direct(plan, 256, &ca.w256)) return -1;
      if (tab_fourstep(plan, 4096, &ca.ptw_lo, &ca.ptw_hi)) return -1;
      ca.ptw_hb = fourstep_hb(4096);
      ca.w_load = pi == 0 ? w_load : nullptr;
      ca.w_store = pi + 1 == plan.passes.size() ? w_store : nullptr;
      ca.w_mask = (unsigned long long)plan.per - 1ULL;
      ca.inverse = inverse ? 1 : 0;
      Prof& pf = P();
      const bool prof = pf.on && !resolve && pf.used + 2 <= pf.pool.size();
      if (prof) HIP_TRY(hipEventRecord(pf.pool[pf.used], stream));
      const int rc = launch_fft_col4096(ca, stream);
      if (prof) {
        HIP_TRY(hipEventRecord(pf.pool[pf.used + 1], stream));
        pf.recs.push_back({"fft_col_f32_4096", pf.pool[pf.used], pf.pool[pf.used + 1]});
        pf.used += 2;
      }
      if (rc) {
        set_error("4096-point column pass launch failed (rc=%d): %s", rc, hipGetErrorString(hipGetLastError()));
        return -1;
      }
      cur = p.dst;
      continue;
    }
    PassArgs a{};
    a.w_load = pi == 0 ? w_load : nullptr;
    a.w_store = pi + 1 == plan.passes.size() ? w_store : nullptr;
    a.w_mask = (unsigned long long)plan.per - 1ULL;
    a.in = ptr(p.src);
    a.out = ptr(p.dst);
    a.in_p = p.in_p;
    a.out_p = p.out_p;
    a.in_c = p.in_c;
    a.out_c = p.out_c;
    a.in_k = p.in_k;
    a.out_k = p.out_k;
    a.in_o = p.in_o;
    a.out_o = p.out_o;
    a.ncol = p.ncol;
    const int L = 1 << p.lg;
    // split mode applies to lines with a radix-16 stage (16..4096 points)
    const bool split = plan.split && L >= 16 && L <= 4096;
    const int C = geom_cols_for(L, p.mode, split);
    a.ncb = (p.ncol + C - 1) / C;
    a.K = p.K;
    a.kbits = p.kbits;
    a.inverse = inverse ? 1 : 0;
    if (sched_np(L) > 1) {
      if (tab_stockham(plan, L, &a.tw)) return -1;
    }
    if (p.epi == EPI_TWIDDLE) {
      if (tab_fourstep(plan, p.M, &a.ptw_lo, &a.ptw_hi)) return -1;
      a.ptw_hb = fourstep_hb(p.M);
      a.ptw_mmask = (unsigned)(p.M - 1);
      a.ptw_cshift = p.cshift;
    }
    const unsigned long long per_outer = (unsigned long long)a.ncb * a.K;
    const unsigned long long max_blocks = (0xFFFFFFFFULL / (unsigned)geom_threads_for(L, p.mode, split));
    const unsigned long long outer_per_launch = std::max(1ULL, max_blocks / per_outer);
    for (unsigned long long o0 = 0; o0 < p.nouter; o0 += outer_per_launch) {
      const unsigned long long no = std::min(outer_per_launch, p.nouter - o0);
      a.outer0 = o0;
      Prof& pf = P();
      const bool prof = pf.on && !resolve && pf.used + 2 <= pf.pool.size();
      if (prof) HIP_TRY(hipEventRecord(pf.pool[pf.used], stream));
      int rc = launch_fft_pass(plan.prec, p.lg, p.mode, p.epi, a, per_outer * no, stream, split);
      if (prof) {
        HIP_TRY(hipEventRecord(pf.pool[pf.used + 1], stream));
        char lab[96];
        snprintf(lab, sizeof(lab), "fft_pass_%s%s_L%d%s%s", plan.prec == PREC_F32 ? "f32" : "f64",
                 split ? "_split" : "", L,
                 p.mode == MODE_CONTIG ? "_contig" : "_strided",
                 p.epi == EPI_TWIDDLE ? "_twiddle" : (p.epi == EPI_BITREV ? "_bitrev" : ""));
        pf.recs.push_back({lab, pf.pool[pf.used], pf.pool[pf.used + 1]});
        pf.used += 2;
      }
      if (rc) {
        set_error("pass launch failed (lg=%d mode=%d epi=%d rc=%d): %s", p.lg, p.mode, p.epi, rc,
                  hipGetErrorString(hipGetLastError()));
        return -1;
      }
    }
    cur = p.dst;
  }
  if (plan.final_copy && !resolve) {
    const void* src = plan.passes.empty() ? din : ptr(cur);
    if (src != dout) HIP_TRY(hipMemcpyAsync(dout, src, plan.total * esz, hipMemcpyDeviceToDevice, stream));
  }
  return 0;
}


static bool check_dims(int rank, const size_t* dims, size_t batch);
static double ms_between(hipEvent_t a, hipEvent_t b);
static int check_dev_buffers(int rank, const size_t* dims, size_t batch, const void* din, const void* dout,
                             size_t esz);

// ---------------------------------------------------------------------------
// multi-transform plans: transforms of uneven sizes in one persistent launch
// ---------------------------------------------------------------------------
// Transforms of one shape whose x and y passes run as the plane kernel (mode 2).
struct PlaneGroup {
  int ly = 0, lx = 0;
  unsigned ppt = 0, count = 0;  // planes per transform, transforms
  size_t tab_off = 0;           // first (in, out) pair of the group in the device tables
  bool cube = false;            // 128^3 transforms whose input and output do not overlap: the y-split pair
};

struct MultiPlan {
  std::vector<MultiPass> passes;
  std::vector<MultiItem> items;
  std::vector<unsigned> phase_start;
  unsigned nphase = 0;
  MultiPass* d_passes = nullptr;
  MultiItem* d_items = nullptr;
  unsigned* d_phase = nullptr;
  MultiCtl* d_ctl = nullptr;
  unsigned grid = 0, chunk = 4;
  // 0: one launch, grid barriers between phases; 1: one launch per phase;
  // 2: x and y of each shape group as one plane-kernel launch, then its z
  // passes as one line-pass launch (transform tables; no persistent kernel)
  int mode = 1;
  bool split = false;
  bool plane_ok = false;
  unsigned long long gen = 0;  // library generation the plan's tables belong to  // mode 2 available (every transform 3D, planes up to 128 x 128)
  std::vector<PlaneGroup> groups;
  unsigned long long* d_tab = nullptr;    // (input, output) offset per transform, by group
  unsigned long long* d_tab_z = nullptr;  // (output, output): the z passes run in place
  const void* din = nullptr;
  void* dout = nullptr;
  bool inverse = false;
  std::string sig;
  size_t points = 0;
  // mode 2 with several shape groups: one stream per group (default
  // priority) forked from and joined back to the caller's stream, so the
  // groups' launch chains overlap (also as parallel branches of a HIP graph)
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> aux_done;
  hipEvent_t fork_ev = nullptr;
  // mode 3 (one dataflow launch, fft_flow.hip): every transform n^3 with
  // n = 32, 64 or 128 (128^3 out of place), FP32
  bool flow_ok = false;
  FlowArgs flow{};
  std::vector<FlowItem> fitems;
  std::vector<unsigned> fq;        // FLOW_QUEUES + 1 queue starts
  FlowItem* d_fitems = nullptr;
  unsigned* d_fq = nullptr;
  unsigned* d_done = nullptr;      // finished first-pass tiles per transform
  unsigned ndone = 0;
  ~MultiPlan() {
    for (hipStream_t s : aux) (void)hipStreamDestroy(s);
    for (hipEvent_t e : aux_done) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (d_passes) (void)hipFree(d_passes);
    if (d_items) (void)hipFree(d_items);
    if (d_phase) (void)hipFree(d_phase);
    if (d_ctl) (void)hipFree(d_ctl);
    if (d_tab) (void)hipFree(d_tab);
    if (d_tab_z) (void)hipFree(d_tab_z);
    if (d_fitems) (void)hipFree(d_fitems);
    if (d_fq) (void)hipFree(d_fq);
    if (d_done) (void)hipFree(d_done);
  }
};

static unsigned g_multi_spin_limit = 1u << 22;

// Resident workgroups of the persistent kernel of `mode` (3: the dataflow
// kernel; else the line-tile kernel) per CU and over the device.
static int multi_default_grid(unsigned* grid, unsigned* per_cu_max, bool split, int mode = 0) {
  int ncu = 0, per = 0;
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, S().dev < 0 ? 0 : S().dev));
  if ((mode == 3 ? flow_max_blocks_per_cu(&per) : multi_max_blocks_per_cu(&per, split)) || per <= 0) {
    set_error("occupancy query for the multi-transform kernel failed");
    return -1;
  }
  *per_cu_max = (unsigned)per;
  *grid = (unsigned)(ncu * per);
  return 0;
}

static int multi_prepare_tables(const MultiPlan* mp);
static int group_cube_args(const MultiPlan* mp, const PlaneGroup& g, CubeArgs* ca, CubeArgs* cb);
static int group_plane_args(const MultiPlan* mp, const PlaneGroup& g, PlaneArgs* pa, PassArgs* za);

// ---- dataflow launch (mode 3) ----------------------------------------------------
// A transform of the launch: its kind (0: 32^3, 1: 64^3, 2: 128^3) and its
// index within its group.
struct FlowXf {
  unsigned char kind;
  unsigned tr;
};
static const int kFlowSide[3] = {32, 64, 128};
// Queues of the dataflow launch.  Transforms go to the queues largest first,
// each to the queue with the fewest bytes so far (lowest index on ties); a
// queue lists the first-pass tiles of its transforms in that order, then
// their second-pass tiles in the same order.  Every second-pass tile thus
// comes after all first-pass tiles of its transform, in the same queue.
// (Second-pass tiles interleaved among the first-pass ones, each transform's
// 256 or 2048 tiles after its last first-pass tile, measured 1.53 and 0.96 ms
// on the 448-transform batch against 0.94 ms: a workgroup publishes before
// it waits, so every switch between the passes costs a release.)
static void flow_layout(const std::vector<FlowXf>& xf, std::vector<FlowItem>* items,
                        std::vector<unsigned>* qstart) {
  std::vector<unsigned> ord(xf.size());
  std::iota(ord.begin(), ord.end(), 0u);
  std::stable_sort(ord.begin(), ord.end(), [&](unsigned a, unsigned b) { return xf[a].kind > xf[b].kind; });
  std::vector<std::vector<unsigned>> q(FLOW_QUEUES);
  std::vector<unsigned long long> load(FLOW_QUEUES, 0);
  for (unsigned i : ord) {
    const unsigned m = (unsigned)(std::min_element(load.begin(), load.end()) - load.begin());
    const unsigned long long n = (unsigned long long)kFlowSide[xf[i].kind];
    q[m].push_back(i);
    load[m] += n * n * n;
  }
  items->clear();
  qstart->assign(1, 0u);
  for (const std::vector<unsigned>& qq : q) {
    for (int pass = 0; pass < 2; ++pass)
      for (unsigned i : qq) {
        const unsigned nt = flow_tiles(kFlowSide[xf[i].kind], pass);
        for (unsigned sub = 0; sub < nt; ++sub)
          items->push_back({xf[i].tr, (unsigned short)sub, xf[i].kind, (unsigned char)pass});
      }
    qstart->push_back((unsigned)items->size());
  }
}

// The group's kind for the dataflow launch, or -1 when it has none.
static int flow_kind(const PlaneGroup& g) {
  if (g.ly != (int)g.ppt || g.lx != (int)g.ppt) return -1;
  if (g.ppt == 32) return 0;
  if (g.ppt == 64) return 1;
  if (g.ppt == 128 && g.cube) return 2;
  return -1;
}

// The dataflow launch's groups, queues and counters (FP32 plans whose groups
// all have a kind); leaves flow_ok false otherwise.
static int flow_build(MultiPlan* mp) {
  if (mp->split) return 0;
  std::vector<FlowXf> xf;
  unsigned dep = 0;
  for (const PlaneGroup& g : mp->groups)
    if (flow_kind(g) < 0) return 0;
  for (const PlaneGroup& g : mp->groups) {
    const int k = flow_kind(g);
    FlowGroup& fg = mp->flow.grp[k];
    fg.n = (int)g.ppt;
    fg.dep0 = dep;
    fg.need = flow_tiles(fg.n, 0);
    if (k == 2) {
      if (group_cube_args(mp, g, &fg.ca, &fg.cb)) return -1;
    } else {
      if (group_plane_args(mp, g, &fg.pa, &fg.za)) return -1;
      if (fg.za.ncb != 2 * flow_tiles(fg.n, 1)) {
        set_error("dataflow layout: unexpected z-line blocks per transform");
        return -1;
      }
    }
    for (unsigned tr = 0; tr < g.count; ++tr) xf.push_back({(unsigned char)k, tr});
    dep += g.count;
  }
  flow_layout(xf, &mp->fitems, &mp->fq);
  mp->ndone = dep;
  HIP_TRY(hipMalloc(&mp->d_fitems, sizeof(FlowItem) * std::max<size_t>(1, mp->fitems.size())));
  HIP_TRY(hipMalloc(&mp->d_fq, sizeof(unsigned) * mp->fq.size()));
  HIP_TRY(hipMalloc(&mp->d_done, sizeof(unsigned) * FLOW_DONE_STRIDE * std::max(1u, dep)));
  if (!mp->fitems.empty())
    HIP_TRY(hipMemcpy(mp->d_fitems, mp->fitems.data(), sizeof(FlowItem) * mp->fitems.size(),
                      hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(mp->d_fq, mp->fq.data(), sizeof(unsigned) * mp->fq.size(), hipMemcpyHostToDevice));
  mp->flow_ok = true;
  return 0;
}

// Host-only: the dataflow queues of a batch of cubes (n = 32, 64, 128), with
// each tile as (transform index, pass, tile) -- grouped by side, largest
// first, members in index order, as multi_create groups them.
int flow_describe(unsigned count, const unsigned* dims3, unsigned* qstart, unsigned* out, unsigned cap) {
  if (!dims3 || !qstart || count == 0) {
    set_error("flow_describe: bad arguments");
    return -1;
  }
  std::vector<unsigned> members[3];
  for (unsigned i = 0; i < count; ++i) {
    const unsigned n = dims3[3 * i];
    if (dims3[3 * i + 1] != n || dims3[3 * i + 2] != n || (n != 32 && n != 64 && n != 128)) {
      set_error("flow_describe: transform %u is not 32^3, 64^3 or 128^3", i);
      return -1;
    }
    members[n == 32 ? 0 : n == 64 ? 1 : 2].push_back(i);
  }
  std::vector<FlowXf> xf;
  std::vector<unsigned> orig;
  for (int k = 2; k >= 0; --k)
    for (unsigned tr = 0; tr < members[k].size(); ++tr) {
      xf.push_back({(unsigned char)k, tr});
      orig.push_back(members[k][tr]);
    }
  std::vector<FlowItem> items;
  std::vector<unsigned> qs;
  flow_layout(xf, &items, &qs);
  for (size_t i = 0; i < qs.size(); ++i) qstart[i] = qs[i];
  if (!out) return (int)items.size();
  if (items.size() > cap) {
    set_error("flow_describe: buffer too small");
    return -1;
  }
  // group-local index back to the caller's index
  std::vector<unsigned> first(3, 0);
  for (size_t j = 0; j < xf.size(); ++j)
    if (xf[j].tr == 0) first[xf[j].kind] = (unsigned)j;
  for (size_t i = 0; i < items.size(); ++i) {
    out[3 * i] = orig[first[items[i].group] + items[i].tr];
    out[3 * i + 1] = items[i].pass;
    out[3 * i + 2] = items[i].sub;
  }
  return (int)items.size();
}

int multi_create(void** handle, unsigned count, const unsigned* dims3, const unsigned long long* in_off,
                 const unsigned long long* out_off, const void* din, void* dout, bool inverse,
                 unsigned flags) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (runtime_ensure_device()) return -1;
  if (!handle || count == 0 || !dims3 || !in_off || !out_off || !din || !dout) {
    set_error("bad multi-transform arguments");
    return -1;
  }
  auto mp = std::make_unique<MultiPlan>();
  mp->split = (flags & FLAG_SPLIT_MFMA) != 0;
  struct Tr {
    unsigned idx;
    size_t pts;
    std::vector<PassDesc> ps;
  };
  std::vector<Tr> trs(count);
  std::map<std::string, unsigned> sizes;
  for (unsigned i = 0; i < count; ++i) {
    size_t d[3] = {dims3[3 * i], dims3[3 * i + 1], dims3[3 * i + 2]};
    if (!check_dims(3, d, 1)) return -1;
    for (int a = 0; a < 3; ++a)
      if (d[a] > (1u << MULTI_MAX_LG)) {
        set_error("multi-transform sizes are limited to 256 points per axis");
        return -1;
      }
    Plan p;
    const bool aliased = (const char*)din + 8 * in_off[i] == (const char*)dout + 8 * out_off[i];
    if (!aliased && check_dev_buffers(3, d, 1, (const char*)din + 8 * in_off[i], (const char*)dout + 8 * out_off[i], 8)) {
      set_error("transform %u: input and output overlap without being the same range", i);
      return -1;
    }
    // line passes only: the persistent kernel's tiles are 256-thread line tiles
    if (build_plan(3, d, 1, FLAG_NO_PLANE_FUSION, PREC_F32, aliased, p)) return -1;
    for (const PassDesc& q : p.passes)
      if (q.mode != MODE_CONTIG && q.mode != MODE_STRIDED) {
        set_error("multi-transform plan: unexpected pass kind");
        return -1;
      }
    if (p.needs_tmp || p.final_copy) {
      set_error("multi-transform plan needs a temporary buffer");
      return -1;
    }
    trs[i] = {i, p.total, p.passes};
    std::ostringstream k;
    k << d[0] << "x" << d[1] << "x" << d[2];
    sizes[k.str()] += 1;
    mp->points += p.total;
  }
  {
    std::ostringstream sg;
    sg << (mp->split ? "multi3d_split" : "multi3d");
    for (auto& kv : sizes) sg << ":" << kv.first << "*" << kv.second;
    mp->sig = sg.str();
  }
  // largest transforms first within every phase
  std::vector<unsigned> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](unsigned a, unsigned b) { return trs[a].pts > trs[b].pts; });
  unsigned nphase = 0;
  for (auto& t : trs) nphase = std::max<unsigned>(nphase, (unsigned)t.ps.size());
  if (nphase > (unsigned)MULTI_MAX_PHASES) {
    set_error("too many phases");
    return -1;
  }
  mp->nphase = nphase;
  // passes, indexed [phase][transform in order]
  std::vector<std::vector<unsigned>> pass_of(nphase, std::vector<unsigned>(count, ~0u));
  for (unsigned ph = 0; ph < nphase; ++ph) {
    for (unsigned oi = 0; oi < count; ++oi) {
      const unsigned i = order[oi];
      if (ph >= trs[i].ps.size()) continue;
      const PassDesc& p = trs[i].ps[ph];
      const int L = 1 << p.lg;
      PassArgs a{};
      const char* ib = (const char*)din + 8 * in_off[i];
      char* ob = (char*)dout + 8 * out_off[i];
      a.in = p.src == B_IN ? (const void*)ib : (const void*)ob;
      a.out = p.dst == B_IN ? (void*)ib : (void*)ob;
      a.in_p = p.in_p;
      a.out_p = p.out_p;
      a.in_c = p.in_c;
      a.out_c = p.out_c;
      a.in_k = p.in_k;
      a.out_k = p.out_k;
      a.in_o = p.in_o;
      a.out_o = p.out_o;
      a.ncol = p.ncol;
      a.ncb = (p.ncol + geom_cols(L, p.mode) - 1) / geom_cols(L, p.mode);
      a.K = p.K;
      a.kbits = p.kbits;
      a.inverse = inverse ? 1 : 0;
      if (sched_np(L) > 1 && get_stockham(PREC_F32, L, &a.tw)) return -1;
      MultiPass m{};
      m.a = a;
      m.path = p.mode * MULTI_MAX_LG + (p.lg - 1);
      pass_of[ph][oi] = (unsigned)mp->passes.size();
      mp->passes.push_back(m);
    }
  }
  mp->phase_start.push_back(0);
  for (unsigned ph = 0; ph < nphase; ++ph) {
    for (unsigned oi = 0; oi < count; ++oi) {
      const unsigned pid = pass_of[ph][oi];
      if (pid == ~0u) continue;
      const PassArgs& a = mp->passes[pid].a;
      const unsigned long long nb = (unsigned long long)a.ncb * a.K * trs[order[oi]].ps[ph].nouter;
      for (unsigned long long b = 0; b < nb; ++b) mp->items.push_back({pid, (unsigned)b});
    }
    mp->phase_start.push_back((unsigned)mp->items.size());
  }
  HIP_TRY(hipMalloc(&mp->d_passes, sizeof(MultiPass) * mp->passes.size()));
  HIP_TRY(hipMalloc(&mp->d_items, sizeof(MultiItem) * std::max<size_t>(1, mp->items.size())));
  HIP_TRY(hipMalloc(&mp->d_phase, sizeof(unsigned) * mp->phase_start.size()));
  HIP_TRY(hipMalloc(&mp->d_ctl, sizeof(MultiCtl)));
  HIP_TRY(hipMemset(mp->d_ctl, 0, sizeof(MultiCtl)));  // checked before any launch is possible
  HIP_TRY(hipMemcpy(mp->d_passes, mp->passes.data(), sizeof(MultiPass) * mp->passes.size(), hipMemcpyHostToDevice));
  if (!mp->items.empty())
    HIP_TRY(hipMemcpy(mp->d_items, mp->items.data(), sizeof(MultiItem) * mp->items.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(mp->d_phase, mp->phase_start.data(), sizeof(unsigned) * mp->phase_start.size(),
                    hipMemcpyHostToDevice));
  // mode 2: every transform has x, y and z passes and planes the plane kernel holds
  mp->din = din;
  mp->dout = dout;
  mp->inverse = inverse;
  mp->plane_ok = nphase == 3;
  for (unsigned i = 0; i < count && mp->plane_ok; ++i)
    mp->plane_ok = trs[i].ps.size() == 3 && plane_shape_supported(dims3[3 * i + 1], dims3[3 * i + 2]);
  if (mp->plane_ok) {
    std::map<std::tuple<unsigned, unsigned, unsigned>, std::vector<unsigned>> by_shape;
    for (unsigned oi = 0; oi < count; ++oi) {
      const unsigned i = order[oi];
      by_shape[std::make_tuple(dims3[3 * i], dims3[3 * i + 1], dims3[3 * i + 2])].push_back(i);
    }
    std::vector<unsigned long long> tab, tab_z;
    for (auto it = by_shape.rbegin(); it != by_shape.rend(); ++it) {  // largest shapes first
      PlaneGroup g;
      g.ppt = std::get<0>(it->first);
      g.ly = (int)std::get<1>(it->first);
      g.lx = (int)std::get<2>(it->first);
      g.count = (unsigned)it->second.size();
      g.tab_off = tab.size() / 2;
      // 128^3 transforms run through the y-split pair, whose first pass runs
      // out of place: only when no transform's input and output overlap
      g.cube = g.ppt == 128 && g.ly == 128 && g.lx == 128;
      for (unsigned i : it->second) {
        const unsigned long long bytes = 8ull * 128 * 128 * 128;
        const unsigned long long ib = (unsigned long long)(uintptr_t)din + 8ull * in_off[i];
        const unsigned long long ob = (unsigned long long)(uintptr_t)dout + 8ull * out_off[i];
        if (ib < ob + bytes && ob < ib + bytes) g.cube = false;
      }
      for (unsigned i : it->second) {
        tab.push_back(in_off[i]);
        tab.push_back(out_off[i]);
        tab_z.push_back(out_off[i]);
        tab_z.push_back(out_off[i]);
      }
      mp->groups.push_back(g);
    }
    HIP_TRY(hipMalloc(&mp->d_tab, sizeof(unsigned long long) * tab.size()));
    HIP_TRY(hipMemcpy(mp->d_tab, tab.data(), sizeof(unsigned long long) * tab.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&mp->d_tab_z, sizeof(unsigned long long) * tab_z.size()));
    HIP_TRY(hipMemcpy(mp->d_tab_z, tab_z.data(), sizeof(unsigned long long) * tab_z.size(),
                      hipMemcpyHostToDevice));
    // made now, so that a mode-2 launch allocates nothing (it may run under capture)
    if (multi_prepare_tables(mp.get())) return -1;
    if (flow_build(mp.get())) return -1;
  }
  unsigned per = 0;
  if (multi_default_grid(&mp->grid, &per, mp->split)) return -1;
  // untuned: the grouped plane launches where they apply, else one launch per
  // phase; neither waits on a grid barrier, which needs every workgroup of
  // the grid resident at once
  mp->mode = mp->plane_ok ? 2 : 1;
  mp->gen = st.gen;
  *handle = mp.release();
  return 0;
}

// every twiddle table the mode-2 launches read (cached; made outside any capture)
static int multi_prepare_tables(const MultiPlan* mp) {
  const void* t = nullptr;
  for (const PlaneGroup& g : mp->groups) {
    if (g.cube) {
      if (get_stockham(PREC_F32, 128, &t) || get_direct(256, &t)) return -1;
      continue;
    }
    if (sched_np(g.lx) > 1 && get_stockham(PREC_F32, g.lx, &t)) return -1;
    if (sched_np(g.ly) > 1 && get_stockham(PREC_F32, g.ly, &t)) return -1;
    if (sched_np((int)g.ppt) > 1 && get_stockham(PREC_F32, (int)g.ppt, &t)) return -1;
  }
  return 0;
}

// the streams and events of the mode-2 fork and its tables (created outside any capture)
static int multi_prepare_fork(MultiPlan* mp) {
  if (multi_prepare_tables(mp)) return -1;
  if (mp->groups.size() < 2 || !mp->aux.empty()) return 0;
  HIP_TRY(hipEventCreateWithFlags(&mp->fork_ev, hipEventDisableTiming));
  for (size_t i = 0; i < mp->groups.size(); ++i) {
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    mp->aux.push_back(s);
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    mp->aux_done.push_back(e);
  }
  return 0;
}

// The y-split pair's arguments for a 128^3 group: pass A out of place from
// the inputs, pass B in place on the outputs (transform tables).
static int group_cube_args(const MultiPlan* mp, const PlaneGroup& g, CubeArgs* ca, CubeArgs* cb) {
  CubeArgs c{};
  c.n = 128;
  c.nplanes = (unsigned long long)g.count * 128;
  if (get_stockham(PREC_F32, 128, &c.tw)) return -1;
  if (get_direct(256, &c.w256)) return -1;
  c.inverse = mp->inverse ? 1 : 0;
  c.split = mp->split ? 1 : 0;
  c.in = mp->din;
  c.out = mp->dout;
  c.tab = mp->d_tab + 2 * g.tab_off;
  *ca = c;
  c.in = mp->dout;  // pass B in place on the output
  c.tab = mp->d_tab_z + 2 * g.tab_off;
  *cb = c;
  return 0;
}

// A plane group's arguments: x and y of every plane from the inputs to the
// outputs, then the z lines in place on the outputs.
static int group_plane_args(const MultiPlan* mp, const PlaneGroup& g, PlaneArgs* pa, PassArgs* za) {
  PlaneArgs a{};
  a.in = mp->din;
  a.out = mp->dout;
  a.nplanes = (unsigned long long)g.count * g.ppt;
  a.plane_stride = (long long)g.ly * g.lx;
  if (sched_np(g.lx) > 1 && get_stockham(PREC_F32, g.lx, &a.tw_x)) return -1;
  if (sched_np(g.ly) > 1 && get_stockham(PREC_F32, g.ly, &a.tw_y)) return -1;
  a.inverse = mp->inverse ? 1 : 0;
  a.tab = mp->d_tab + 2 * g.tab_off;
  a.ppt = g.ppt;
  const long long L = g.ppt, s = (long long)g.ly * g.lx;
  PassArgs z{};
  z.in = mp->dout;
  z.out = mp->dout;
  z.in_p = z.out_p = s;
  z.in_c = z.out_c = 1;
  z.ncol = (unsigned)s;
  const int zc = geom_cols_for((int)L, MODE_STRIDED, mp->split && L >= 16);
  z.ncb = (unsigned)((s + zc - 1) / zc);
  z.K = 1;
  z.inverse = mp->inverse ? 1 : 0;
  z.tab = mp->d_tab_z + 2 * g.tab_off;
  if (sched_np((int)L) > 1 && get_stockham(PREC_F32, (int)L, &z.tw)) return -1;
  *pa = a;
  *za = z;
  return 0;
}

// x and y of one shape group as one plane-kernel launch, then its z axis as
// one strided line pass over all its transforms, in place
static int multi_group_exec(MultiPlan* mp, const PlaneGroup& g, hipStream_t stream) {
  if (g.cube) {
    CubeArgs ca{}, cb{};
    if (group_cube_args(mp, g, &ca, &cb)) return -1;
    if (launch_fft_cube_a(ca, stream)) {
      set_error("y-split pass A launch failed: %s", hipGetErrorString(hipGetLastError()));
      return -1;
    }
    if (launch_fft_cube_b(cb, stream)) {
      set_error("y-split pass B launch failed: %s", hipGetErrorString(hipGetLastError()));
      return -1;
    }
    return 0;
  }
  PlaneArgs a{};
  PassArgs z{};
  if (group_plane_args(mp, g, &a, &z)) return -1;
  if (launch_fft_plane(g.ly, g.lx, a, stream, mp->split)) {
    set_error("plane kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    return -1;
  }
  const long long L = g.ppt;
  // split precision applies to lines with a radix-16 stage, as in the line passes
  if (launch_fft_pass(PREC_F32, ilog2c(L), MODE_STRIDED, EPI_NONE, z, (unsigned long long)z.ncb * g.count,
                      stream, mp->split && L >= 16)) {
    set_error("z pass launch failed: %s", hipGetErrorString(hipGetLastError()));
    return -1;
  }
  return 0;
}

// A plan made before the library released its tables (fpga_final or a new
// fpga_initialize) holds pointers to freed twiddle tables: refuse it.
static bool multi_current(const MultiPlan* mp) {
  const State& st = S();
  if (!st.ready || mp->gen != st.gen) {
    set_error("multi-transform plan or graph predates fpga_final / fpga_initialize; create it again");
    return false;
  }
  return true;
}

int multi_exec(void* handle, void* stream) {
  MultiPlan* mp = static_cast<MultiPlan*>(handle);
  if (!mp) {
    set_error("null multi-transform handle");
    return -1;
  }
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (!multi_current(mp)) return -1;
  if (mp->nphase == 0) return 0;
  hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
  if (mp->mode == 2) {
    // no grid barrier in this schedule: clear what an earlier launch reported
    if (reset_multi_ctl(mp->d_ctl, stream)) {
      set_error("multi-transform reset failed: %s", hipGetErrorString(hipGetLastError()));
      return -1;
    }
    if (mp->groups.size() < 2) {
      for (const PlaneGroup& g : mp->groups)
        if (multi_group_exec(mp, g, hs)) return -1;
      return 0;
    }
    // the groups touch disjoint transforms: each chain on its own stream
    if (multi_prepare_fork(mp)) return -1;
    HIP_TRY(hipEventRecord(mp->fork_ev, hs));
    for (size_t i = 0; i < mp->groups.size(); ++i) {
      HIP_TRY(hipStreamWaitEvent(mp->aux[i], mp->fork_ev, 0));
      if (multi_group_exec(mp, mp->groups[i], mp->aux[i])) return -1;
      HIP_TRY(hipEventRecord(mp->aux_done[i], mp->aux[i]));
    }
    for (size_t i = 0; i < mp->groups.size(); ++i) HIP_TRY(hipStreamWaitEvent(hs, mp->aux_done[i], 0));
    return 0;
  }
  if (mp->mode == 3) {
    if (launch_flow(mp->flow, mp->d_fitems, mp->d_fq, mp->d_ctl, mp->d_done, mp->ndone, mp->grid, mp->chunk,
                    g_multi_spin_limit, mp->inverse, stream)) {
      set_error("dataflow launch failed: %s", hipGetErrorString(hipGetLastError()));
      return -1;
    }
    return 0;
  }
  if (mp->mode == 0) {
    if (launch_multi(mp->d_passes, mp->d_items, mp->d_phase, 0, mp->nphase, mp->d_ctl, mp->grid, mp->chunk,
                     g_multi_spin_limit, stream, mp->split)) {
      set_error("multi-transform launch failed: %s", hipGetErrorString(hipGetLastError()));
      return -1;
    }
  } else {
    for (unsigned ph = 0; ph < mp->nphase; ++ph)
      if (launch_multi(mp->d_passes, mp->d_items, mp->d_phase, ph, 1, mp->d_ctl, mp->grid, mp->chunk,
                       g_multi_spin_limit, stream, mp->split)) {
        set_error("multi-transform launch failed: %s", hipGetErrorString(hipGetLastError()));
        return -1;
      }
  }
  return 0;
}

int multi_check(void* handle, void* stream) {
  MultiPlan* mp = static_cast<MultiPlan*>(handle);
  MultiCtl c{};
  HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
  HIP_TRY(hipMemcpy(&c, mp->d_ctl, sizeof(c), hipMemcpyDeviceToHost));
  if (c.err) {
    set_error("multi-transform grid barrier wait exceeded its bound");
    return -1;
  }
  return 0;
}

void multi_destroy(void* handle) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  delete static_cast<MultiPlan*>(handle);
}

int multi_get_schedule(void* handle, int* mode, unsigned* grid, unsigned* chunk, unsigned* items,
                       unsigned* nphase) {
  MultiPlan* mp = static_cast<MultiPlan*>(handle);
  if (!mp) return -1;
  if (mode) *mode = mp->mode;
  if (grid) *grid = mp->grid;
  if (chunk) *chunk = mp->chunk;
  if (items) *items = (unsigned)mp->items.size();
  if (nphase) *nphase = mp->nphase;
  return 0;
}

int multi_set_schedule(void* handle, int mode, unsigned grid, unsigned chunk) {
  MultiPlan* mp = static_cast<MultiPlan*>(handle);
  unsigned g = 0, per = 0;
  if (!mp || multi_default_grid(&g, &per, mp->split, mode)) return -1;
  if (grid == 0) grid = g;  // every workgroup that can be resident
  if (mode < 0 || mode > 3 || (mode == 2 && !mp->plane_ok) || (mode == 3 && !mp->flow_ok) || grid > g ||
      (mode == 3 && chunk == 0)) {
    set_error("schedule out of range (modes 0, 1%s%s; grid must be 0 (all) or 1..%u resident workgroups%s)",
              mp->plane_ok ? ", 2" : "", mp->flow_ok ? ", 3" : "", g, mode == 3 ? ", chunk >= 1" : "");
    return -1;
  }
  mp->mode = mode;
  mp->grid = grid;
  mp->chunk = chunk;
  return 0;
}

void multi_set_spin_limit(unsigned polls) { g_multi_spin_limit = polls ? polls : 1u; }

// Autotuner: time every (mode, workgroups per CU, chunk) candidate on the
// plan's own buffers, keep the fastest, and record it under the plan's shape
// signature in a text cache (one line: signature mode per_cu chunk microseconds).
int multi_tune(void* handle, void* stream, const char* cache_path, int reps, double* best_us) {
  MultiPlan* mp = static_cast<MultiPlan*>(handle);
  if (!mp) return -1;
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (!multi_current(mp)) return -1;
  unsigned gmax = 0, per = 0, gflow = 0, per_flow = 0;
  if (multi_default_grid(&gmax, &per, mp->split)) return -1;
  if (mp->flow_ok && multi_default_grid(&gflow, &per_flow, mp->split, 3)) return -1;
  const unsigned ncu = gmax / per;
  const int max_mode = mp->flow_ok ? 3 : mp->plane_ok ? 2 : 1;
  auto per_of = [&](int mode) { return mode == 3 ? per_flow : per; };
  // cached?
  if (cache_path && *cache_path) {
    std::ifstream f(cache_path);
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream ls(line);
      std::string sg;
      int mode;
      unsigned pc, ch;
      double us;
      if ((ls >> sg >> mode >> pc >> ch >> us) && sg == mp->sig && mode >= 0 && mode <= max_mode &&
          (mode != 2 || mp->plane_ok) && pc >= 1 && pc <= per_of(mode) && (mode != 3 || ch >= 1)) {
        mp->mode = mode;
        mp->grid = pc * ncu;
        mp->chunk = ch;
        if (best_us) *best_us = us;
        return 1;  // from cache
      }
    }
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the events are released and the plan's schedule restored on every exit
  struct Ev {
    hipEvent_t e = nullptr;
    ~Ev() {
      if (e) (void)hipEventDestroy(e);
    }
  } e0, e1;
  const int prev_mode = mp->mode;
  const unsigned prev_grid = mp->grid, prev_chunk = mp->chunk;
  auto fail = [&]() {
    mp->mode = prev_mode;
    mp->grid = prev_grid;
    mp->chunk = prev_chunk;
    return -1;
  };
  if (hipEventCreate(&e0.e) != hipSuccess || hipEventCreate(&e1.e) != hipSuccess) {
    set_error("tuner event creation failed");
    return fail();
  }
  double best = 1e30;
  int bmode = 0;
  unsigned bpc = per, bch = 4;
  const unsigned chunks[] = {1, 2, 4, 8, 16, 0};  // 0: static round-robin, no tickets
  for (int mode = 0; mode <= max_mode; ++mode)
    for (unsigned pc = 1; pc <= per_of(mode); ++pc)
      for (unsigned ch : chunks) {
        // mode 2 launches plain kernels: grid and chunk do not apply
        if (mode == 2 && (pc != per || ch != chunks[0])) continue;
        if (mode == 2 && !mp->plane_ok) continue;
        // mode 3 takes its tiles by tickets only (its queues' order is what
        // lets a launch finish with any number of resident workgroups)
        if (mode == 3 && ch == 0) continue;
        mp->mode = mode;
        mp->grid = pc * ncu;
        mp->chunk = ch;
        if (multi_exec(mp, s) || multi_check(mp, s)) return fail();  // warm-up and error check
        if (hipEventRecord(e0.e, s) != hipSuccess) {
          set_error("tuner event record failed: %s", hipGetErrorString(hipGetLastError()));
          return fail();
        }
        for (int r = 0; r < std::max(1, reps); ++r)
          if (multi_exec(mp, s)) return fail();
        if (hipEventRecord(e1.e, s) != hipSuccess || hipEventSynchronize(e1.e) != hipSuccess) {
          set_error("tuner event record failed: %s", hipGetErrorString(hipGetLastError()));
          return fail();
        }
        const double us = 1e3 * ms_between(e0.e, e1.e) / std::max(1, reps);
        if (us < best) {
          best = us;
          bmode = mode;
          bpc = pc;
          bch = ch;
        }
      }
  if (multi_check(mp, s)) return fail();
  mp->mode = bmode;
  mp->grid = bpc * ncu;
  mp->chunk = bch;
  if (best_us) *best_us = best;
  if (cache_path && *cache_path) {
    std::ofstream f(cache_path, std::ios::app);
    if (f) f << mp->sig << " " << bmode << " " << bpc << " " << bch << " " << best << "\n";
  }
  return 0;
}

// ---- graphs: a sequence of multi-transform executions captured once ------------
struct Graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  unsigned long long gen = 0;  // library generation of the captured plans' tables
  ~Graph() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
  }
};

int graph_capture_multi(void** out, void** handles, int n, void* stream) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!out || !handles || n <= 0 || s == nullptr) {
    set_error("graph capture needs handles and a non-null stream");
    return -1;
  }
  auto gr = std::make_unique<Graph>();
  for (int i = 0; i < n; ++i) {
    if (!handles[i]) {
      set_error("null multi-transform handle");
      return -1;
    }
    if (!multi_current(static_cast<MultiPlan*>(handles[i]))) return -1;
    if (multi_prepare_fork(static_cast<MultiPlan*>(handles[i]))) return -1;  // no creation under capture
  }
  gr->gen = st.gen;
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int i = 0; i < n && !rc; ++i) rc = multi_exec(handles[i], s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return -1;
  }
  if (e != hipSuccess) {
    set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
    return -1;
  }
  gr->g = g;
  HIP_TRY(hipGraphInstantiate(&gr->x, g, nullptr, nullptr, 0));
  *out = gr.release();
  return 0;
}

int graph_launch(void* graph, void* stream) {
  Graph* g = static_cast<Graph*>(graph);
  if (!g) return -1;
  {
    State& st = S();
    std::lock_guard<std::recursive_mutex> lk(st.mu);
    if (!st.ready || g->gen != st.gen) {
      set_error("graph predates fpga_final / fpga_initialize; capture it again");
      return -1;
    }
  }
  HIP_TRY(hipGraphLaunch(g->x, reinterpret_cast<hipStream_t>(stream)));
  return 0;
}

void graph_destroy(void* graph) { delete static_cast<Graph*>(graph); }

// ---------------------------------------------------------------------------
// runtime
// ---------------------------------------------------------------------------
int runtime_device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

bool runtime_can_map_host(int dev) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeCanMapHostMemory, dev) != hipSuccess) return false;
  return v != 0;
}

static void pipe_release(Pipe* p);

static void release_locked() {
  State& st = S();
  ++st.gen;  // multi-transform plans and graphs made so far are now stale
  pipe_release(st.pipe);
  st.pipe = nullptr;
  for (auto& kv : st.plans)
    if (kv.second.tmp && !kv.second.noalloc) (void)hipFree(kv.second.tmp);
  st.plans.clear();
  for (auto& kv : st.tw_stockham) (void)hipFree(kv.second);
  st.tw_stockham.clear();
  for (auto& kv : st.tw_fourstep) {
    (void)hipFree(kv.second.first);
    (void)hipFree(kv.second.second);
  }
  st.tw_fourstep.clear();
  for (auto& kv : st.tw_direct) (void)hipFree(kv.second);
  st.tw_direct.clear();
  if (st.d_in) (void)hipFree(st.d_in);
  if (st.d_out) (void)hipFree(st.d_out);
  if (st.h_stage) (void)hipHostFree(st.h_stage);
  st.d_in = st.d_out = st.h_stage = nullptr;
  st.d_in_cap = st.d_out_cap = st.h_stage_cap = 0;
  for (auto& e : st.ev)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  if (st.stream) (void)hipStreamDestroy(st.stream);
  st.stream = nullptr;
  st.ready = false;
  st.svm = false;
}

int runtime_init(int dev, bool svm) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (st.ready) release_locked();
  HIP_TRY(hipSetDevice(dev));
  HIP_TRY(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
  for (auto& e : st.ev) HIP_TRY(hipEventCreate(&e));
  if (noalloc_setup(dev)) return -1;  // the allocation-free entry point's scratch
  // every kernel file's code object, so that no later call (one under stream
  // capture included) has one to load
  preload_pass();
  preload_cube();
  preload_c512();
  preload_xy();
  preload_col();
  preload_plane();
  preload_multi();
  preload_flow();
  st.dev = dev;
  st.svm = svm;
  st.ready = true;
  return 0;
}

void runtime_final() {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (st.ready) {
    (void)hipSetDevice(st.dev);
    (void)hipStreamSynchronize(st.stream);
  }
  release_locked();
}

bool runtime_ready() { return S().ready; }
bool runtime_svm() { return S().ready && S().svm; }

int runtime_ensure_device() {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (st.ready) return 0;
  if (runtime_device_count() <= 0) {
    set_error("no HIP device visible");
    return -1;
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  return runtime_init(dev, false);
}

void plan_cache_clear() {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (st.ready) (void)hipDeviceSynchronize();
  for (auto& kv : st.plans)
    if (kv.second.tmp && !kv.second.noalloc) (void)hipFree(kv.second.tmp);
  st.plans.clear();
}

size_t plan_cache_size() { return S().plans.size(); }

int profile_begin(int max_records) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (runtime_ensure_device()) return -1;
  Prof& pf = P();
  const size_t need = 2 * (size_t)std::max(1, max_records);
  while (pf.pool.size() < need) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    pf.pool.push_back(e);
  }
  pf.used = 0;
  pf.recs.clear();
  pf.on = true;
  return 0;
}

int profile_end(char* out, int cap) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  Prof& pf = P();
  pf.on = false;
  std::map<std::string, std::pair<long long, double>> agg;
  for (auto& r : pf.recs) {
    HIP_TRY(hipEventSynchronize(r.b));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    auto& e = agg[r.label];
    e.first += 1;
    e.second += ms;
  }
  std::string js = "{";
  bool first = true;
  for (auto& kv : agg) {
    char buf[192];
    snprintf(buf, sizeof(buf), "%s\"%s\": {\"count\": %lld, \"total_ms\": %.6f}", first ? "" : ", ",
             kv.first.c_str(), kv.second.first, kv.second.second);
    js += buf;
    first = false;
  }
  js += "}";
  pf.recs.clear();
  pf.used = 0;
  if (out && cap > 0) {
    snprintf(out, (size_t)cap, "%s", js.c_str());
  }
  return (int)js.size();
}

static bool check_dims(int rank, const size_t* dims, size_t batch);

// Host-only description of a plan (no device needed): 22 int64 fields per
// pass, then one final-copy flag.
int describe_plan(int rank, const size_t* dims, size_t batch, unsigned flags, int prec, bool aliased,
                  long long* out, int cap) {
  if (!check_dims(rank, dims, batch)) return -1;
  Plan p;
  if (build_plan(rank, dims, batch, flags, prec, aliased, p)) return -1;
  const int NF = 22;
  const int n = (int)p.passes.size();
  if (out == nullptr) return n;
  if (cap < NF * n + 1) {
    set_error("describe_plan: buffer too small");
    return -1;
  }
  auto sp = [&](const PassDesc& d) { return p.split && d.lg >= 4 && d.lg <= 12; };
  // threads per workgroup of a plane pass: 512 for 256 x 256 (four per plane)
  // and for FP32 128 x 128 (two halves), LY * LX / 16 otherwise
  auto plane_threads = [&](const PassDesc& d) -> long long {
    if (d.K == 256 || (d.K == 128 && d.lg == 7 && !p.split)) return 512;
    return ((long long)d.K << d.lg) / 16;
  };
  for (int i = 0; i < n; ++i) {
    const PassDesc& d = p.passes[i];
    long long* f = out + NF * i;
    long long v[NF] = {d.lg, d.mode, d.epi, d.transposing, d.in_p, d.out_p, d.in_c, d.out_c,
                       d.in_k, d.out_k, d.in_o, d.out_o, d.ncol, d.K, d.kbits,
                       (long long)d.nouter, d.M, d.cshift, d.src, d.dst,
                       d.mode == MODE_PLANE ? (d.K == 256 ? 64 : (long long)d.K)
                       : d.mode == MODE_COL4096 ? 16
                       : d.mode == MODE_CUBE_A ? (d.lg == 9 ? 64 : (1LL << d.lg) / 4)   // rows per workgroup
                       : d.mode == MODE_CUBE_B ? (d.lg == 9 ? 8 : d.lg == 8 ? 16 : 32)   // columns per workgroup (x 8 or 4 rows x n planes)
                       : geom_cols_for(1 << d.lg, d.mode, sp(d)),
                       d.mode == MODE_PLANE ? plane_threads(d)
                       : d.mode == MODE_COL4096 ? 512
                       : d.mode == MODE_CUBE_A ? (d.lg == 9 ? 1024 : d.lg == 8 ? 512 : 256)
                       : d.mode == MODE_CUBE_B ? (d.lg == 9 ? 1024 : 512)
                       : geom_threads_for(1 << d.lg, d.mode, sp(d))};
    for (int k = 0; k < NF; ++k) f[k] = v[k];
  }
  out[NF * n] = p.final_copy ? 1 : 0;
  return n;
}

static int ensure_buf(void** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return 0;
  if (*p) HIP_TRY(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  HIP_TRY(hipMalloc(p, bytes));
  *cap = bytes;
  return 0;
}

static bool check_dims(int rank, const size_t* dims, size_t batch) {
  if (rank < 1 || rank > 3 || batch == 0) {
    set_error("bad rank/batch");
    return false;
  }
  for (int d = 0; d < rank; ++d)
    if (!is_pow2(dims[d])) {
      set_error("dimension %zu is not a power of two", dims[d]);
      return false;
    }
  return true;
}

static double ms_between(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
  return (double)ms;
}

int host_c2c(int rank, const size_t* dims, size_t batch, const void* hin, void* hout, bool inverse,
             unsigned flags, int prec, Timing* t) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (!st.ready) {
    set_error("library not initialised (call fpga_initialize first)");
    return -1;
  }
  if (!check_dims(rank, dims, batch)) return -1;
  HIP_TRY(hipSetDevice(st.dev));
  Plan* plan = nullptr;
  if (get_plan(rank, dims, batch, flags, prec, false, &plan)) return -1;
  const size_t bytes = plan->total * (prec == PREC_F32 ? 8 : 16);
  if (ensure_buf(&st.d_in, &st.d_in_cap, bytes) || ensure_buf(&st.d_out, &st.d_out_cap, bytes))
    return -1;
  HIP_TRY(hipEventRecord(st.ev[0], st.stream));
  HIP_TRY(hipMemcpyAsync(st.d_in, hin, bytes, hipMemcpyHostToDevice, st.stream));
  HIP_TRY(hipEventRecord(st.ev[1], st.stream));
  if (run_plan(*plan, st.d_in, st.d_out, inverse, st.stream)) return -1;
  HIP_TRY(hipEventRecord(st.ev[2], st.stream));
  HIP_TRY(hipMemcpyAsync(hout, st.d_out, bytes, hipMemcpyDeviceToHost, st.stream));
  HIP_TRY(hipEventRecord(st.ev[3], st.stream));
  HIP_TRY(hipStreamSynchronize(st.stream));
  if (t) {
    t->h2d_ms = ms_between(st.ev[0], st.ev[1]);
    t->exec_ms = ms_between(st.ev[1], st.ev[2]);
    t->d2h_ms = ms_between(st.ev[2], st.ev[3]);
  }
  return 0;
}

int host_c2c_mapped(int rank, const size_t* dims, size_t batch, const void* hin, void* hout,
                    bool inverse, unsigned flags, Timing* t) {
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (!st.ready) {
    set_error("library not initialised (call fpga_initialize first)");
    return -1;
  }
  if (!check_dims(rank, dims, batch)) return -1;
  HIP_TRY(hipSetDevice(st.dev));
  Plan* plan = nullptr;
  if (get_plan(rank, dims, batch, flags, PREC_F32, false, &plan)) return -1;
  const size_t bytes = plan->total * 8;
  if (ensure_buf(&st.d_in, &st.d_in_cap, bytes) || ensure_buf(&st.d_out, &st.d_out_cap, bytes))
    return -1;
  if (st.h_stage_cap < bytes) {
    if (st.h_stage) HIP_TRY(hipHostFree(st.h_stage));
    st.h_stage = nullptr;
    st.h_stage_cap = 0;
    HIP_TRY(hipHostMalloc(&st.h_stage, bytes, hipHostMallocMapped));
    st.h_stage_cap = bytes;
  }
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  std::memcpy(st.h_stage, hin, bytes);
  auto t1 = clk::now();
  HIP_TRY(hipEventRecord(st.ev[0], st.stream));
  HIP_TRY(hipMemcpyAsync(st.d_in, st.h_stage, bytes, hipMemcpyHostToDevice, st.stream));
  HIP_TRY(hipEventRecord(st.ev[1], st.stream));
  if (run_plan(*plan, st.d_in, st.d_out, inverse, st.stream)) return -1;
  HIP_TRY(hipEventRecord(st.ev[2], st.stream));
  HIP_TRY(hipMemcpyAsync(st.h_stage, st.d_out, bytes, hipMemcpyDeviceToHost, st.stream));
  HIP_TRY(hipEventRecord(st.ev[3], st.stream));
  HIP_TRY(hipStreamSynchronize(st.stream));
  auto t2 = clk::now();
  std::memcpy(hout, st.h_stage, bytes);
  auto t3 = clk::now();
  if (t) {
    t->stage_in_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    t->h2d_ms = ms_between(st.ev[0], st.ev[1]);
    t->exec_ms = ms_between(st.ev[1], st.ev[2]);
    t->d2h_ms = ms_between(st.ev[2], st.ev[3]);
    t->stage_out_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
  }
  return 0;
}

// Pipelined host batch (the GPU counterpart of the reference's 4-buffer
// rotation, api/src/fft3d.c:551-705): chunks of whole transforms flow
//   host memcpy -> pinned[slot] -> H2D (s_h2d) -> transform (s_exec) -> D2H (s_d2h)
//   -> pinned[slot] -> host memcpy
// with two slots, so the copies of chunk i+1 and i-1 overlap the transform of
// chunk i.  exec_ms is the wall-clock time of the whole call, as the
// reference reports it (api/src/fft3d.c:489,861).
struct Pipe {
  hipStream_t s[3] = {nullptr, nullptr, nullptr};
  hipEvent_t h2d[2] = {}, ex[2] = {}, d2h[2] = {};
  void* pin_in[2] = {nullptr, nullptr};
  void* pin_out[2] = {nullptr, nullptr};
  void* d_in[2] = {nullptr, nullptr};
  void* d_out[2] = {nullptr, nullptr};
  size_t cap = 0;  // bytes of each chunk buffer
  ~Pipe() {
    for (auto st : s)
      if (st) (void)hipStreamSynchronize(st);
    for (int i = 0; i < 2; ++i) {
      if (pin_in[i]) (void)hipHostFree(pin_in[i]);
      if (pin_out[i]) (void)hipHostFree(pin_out[i]);
      if (d_in[i]) (void)hipFree(d_in[i]);
      if (d_out[i]) (void)hipFree(d_out[i]);
      if (h2d[i]) (void)hipEventDestroy(h2d[i]);
      if (ex[i]) (void)hipEventDestroy(ex[i]);
      if (d2h[i]) (void)hipEventDestroy(d2h[i]);
    }
    for (auto st : s)
      if (st) (void)hipStreamDestroy(st);
  }
};

static void pipe_release(Pipe* p) { delete p; }

int host_c2c_pipelined(int rank, const size_t* dims, size_t count, const void* hin, void* hout,
                       bool inverse, Timing* t) {
  using clk = std::chrono::steady_clock;
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (!st.ready) {
    set_error("library not initialised (call fpga_initialize first)");
    return -1;
  }
  if (!check_dims(rank, dims, count)) return -1;
  HIP_TRY(hipSetDevice(st.dev));
  size_t per = 1;
  for (int d = 0; d < rank; ++d) per *= dims[d];
  const size_t tbytes = per * 8;
  // chunks of whole transforms, about 64 MiB each
  const size_t chunk = std::max<size_t>(1, std::min<size_t>(count, (64u << 20) / std::max<size_t>(tbytes, 1)));
  const size_t nchunks = (count + chunk - 1) / chunk;
  const size_t cbytes = chunk * tbytes;
  // streams, events and the pinned and device chunk buffers live across calls
  // (released by fpga_final); the buffers only grow
  if (!st.pipe) {
    auto np = std::make_unique<Pipe>();
    for (auto& s : np->s) HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
      HIP_TRY(hipEventCreateWithFlags(&np->h2d[i], hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&np->ex[i], hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&np->d2h[i], hipEventDisableTiming));
    }
    st.pipe = np.release();
  }
  Pipe& pp = *st.pipe;
  if (pp.cap < cbytes) {
    for (int i = 0; i < 2; ++i) {
      if (pp.pin_in[i]) HIP_TRY(hipHostFree(pp.pin_in[i]));
      if (pp.pin_out[i]) HIP_TRY(hipHostFree(pp.pin_out[i]));
      if (pp.d_in[i]) HIP_TRY(hipFree(pp.d_in[i]));
      if (pp.d_out[i]) HIP_TRY(hipFree(pp.d_out[i]));
      pp.pin_in[i] = pp.pin_out[i] = pp.d_in[i] = pp.d_out[i] = nullptr;
    }
    pp.cap = 0;
    for (int i = 0; i < 2; ++i) {
      HIP_TRY(hipHostMalloc(&pp.pin_in[i], cbytes, hipHostMallocDefault));
      HIP_TRY(hipHostMalloc(&pp.pin_out[i], cbytes, hipHostMallocDefault));
      HIP_TRY(hipMalloc(&pp.d_in[i], cbytes));
      HIP_TRY(hipMalloc(&pp.d_out[i], cbytes));
    }
    pp.cap = cbytes;
  }
  // plans (and their tables and code objects) for the full chunk and a
  // partial last chunk, before the clock starts -- the reference also creates
  // its buffers before timing a batch (api/src/fft3d.c:431-489)
  Plan* plan_full = nullptr;
  Plan* plan_last = nullptr;
  const size_t last_n = count - (nchunks - 1) * chunk;
  if (get_plan(rank, dims, chunk, 0u, PREC_F32, false, &plan_full)) return -1;
  if (get_plan(rank, dims, last_n, 0u, PREC_F32, false, &plan_last)) return -1;
  auto t0 = clk::now();
  const char* src = static_cast<const char*>(hin);
  char* dst = static_cast<char*>(hout);
  auto nb = [&](size_t i) { return std::min(chunk, count - i * chunk) * tbytes; };
  auto drain = [&](size_t i) -> int {  // copy chunk i out of its pinned slot
    const int slot = (int)(i & 1);
    HIP_TRY(hipEventSynchronize(pp.d2h[slot]));
    std::memcpy(dst + i * cbytes, pp.pin_out[slot], nb(i));
    return 0;
  };
  for (size_t i = 0; i < nchunks; ++i) {
    const int slot = (int)(i & 1);
    const size_t bytes = nb(i);
    const size_t ntr = bytes / tbytes;
    if (i >= 2) HIP_TRY(hipEventSynchronize(pp.h2d[slot]));  // pinned input slot free
    std::memcpy(pp.pin_in[slot], src + i * cbytes, bytes);
    if (i >= 2) HIP_TRY(hipStreamWaitEvent(pp.s[0], pp.ex[slot], 0));  // device input slot free
    HIP_TRY(hipMemcpyAsync(pp.d_in[slot], pp.pin_in[slot], bytes, hipMemcpyHostToDevice, pp.s[0]));
    HIP_TRY(hipEventRecord(pp.h2d[slot], pp.s[0]));
    HIP_TRY(hipStreamWaitEvent(pp.s[1], pp.h2d[slot], 0));
    if (i >= 2) HIP_TRY(hipStreamWaitEvent(pp.s[1], pp.d2h[slot], 0));  // device output slot free
    Plan* plan = ntr == chunk ? plan_full : plan_last;
    if (run_plan(*plan, pp.d_in[slot], pp.d_out[slot], inverse, pp.s[1])) return -1;
    HIP_TRY(hipEventRecord(pp.ex[slot], pp.s[1]));
    if (i >= 2 && drain(i - 2)) return -1;  // pinned output slot free
    HIP_TRY(hipStreamWaitEvent(pp.s[2], pp.ex[slot], 0));
    HIP_TRY(hipMemcpyAsync(pp.pin_out[slot], pp.d_out[slot], bytes, hipMemcpyDeviceToHost, pp.s[2]));
    HIP_TRY(hipEventRecord(pp.d2h[slot], pp.s[2]));
  }
  for (size_t i = nchunks >= 2 ? nchunks - 2 : 0; i < nchunks; ++i)
    if (drain(i)) return -1;
  auto t1 = clk::now();
  if (t) {
    *t = Timing();
    t->exec_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  return 0;
}

// Device buffers of one call: null pointers and input/output ranges that
// overlap without being the same buffer are refused (the first pass of a plan
// may run out of place, so a partial overlap would corrupt the input).
static int check_dev_buffers(int rank, const size_t* dims, size_t batch, const void* din, const void* dout,
                             size_t esz) {
  if (!din || !dout) {
    set_error("null device pointer");
    return -1;
  }
  size_t bytes = batch * esz;
  for (int d = 0; d < rank; ++d) bytes *= dims[d];
  const uintptr_t a = reinterpret_cast<uintptr_t>(din), b = reinterpret_cast<uintptr_t>(dout);
  if (a != b && a < b + bytes && b < a + bytes) {
    set_error("input and output buffers overlap without being the same buffer (in place needs d_in == d_out)");
    return -1;
  }
  return 0;
}

int dev_c2c_weighted(int rank, const size_t* dims, size_t batch, const void* din, void* dout,
                     bool inverse, const float* w_load, const float* w_store, void* stream) {
  if (!check_dims(rank, dims, batch) || check_dev_buffers(rank, dims, batch, din, dout, 8)) return -1;
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (runtime_ensure_device()) return -1;
  Plan* plan = nullptr;
  if (get_plan(rank, dims, batch, 0u, PREC_F32, din == dout, &plan)) return -1;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);  // NULL = the HIP null stream
  return run_plan(*plan, din, dout, inverse, s, w_load, w_store);
}

int dev_c2c_noalloc(int rank, const size_t* dims, size_t batch, const void* din, void* dout, bool inverse,
                    unsigned flags, void* stream) {
  if (!check_dims(rank, dims, batch) || check_dev_buffers(rank, dims, batch, din, dout, 8)) return -1;
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (runtime_ensure_device()) return -1;
  if (noalloc_setup(st.dev)) return -1;
  Plan* plan = nullptr;
  if (get_plan(rank, dims, batch, flags, PREC_F32, din == dout, &plan, true)) return -1;
  return run_plan(*plan, din, dout, inverse, reinterpret_cast<hipStream_t>(stream));
}

int dev_c2c(int rank, const size_t* dims, size_t batch, const void* din, void* dout, bool inverse,
            unsigned flags, int prec, void* stream) {
  if (!check_dims(rank, dims, batch) ||
      check_dev_buffers(rank, dims, batch, din, dout, prec == PREC_F32 ? 8 : 16))
    return -1;
  State& st = S();
  std::lock_guard<std::recursive_mutex> lk(st.mu);
  if (runtime_ensure_device()) return -1;
  Plan* plan = nullptr;
  const bool aliased = din == dout;
  if (get_plan(rank, dims, batch, flags, prec, aliased, &plan)) return -1;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);  // NULL = the HIP null stream
  return run_plan(*plan, din, dout, inverse, s);
}

}  // namespace fftamd

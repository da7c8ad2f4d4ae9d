// fft.cpp -- command-line harness over the drop-in C API, the counterpart of
// the reference's examples/fft.cpp + examples/helper.cpp.
//
// Same flags and dispatch (examples/helper.cpp:45-57, examples/fft.cpp:40-75):
//   -n/--num N  -d/--dim D  -b/--back  -i/--iter I  -p/--path P  -y/--noverify
//   -c/--batch B  -t/--burst  -m/--use_bram  -s/--use_usm  -e/--emulate  -h/--help
// plus --seed S (input from splitmix64(S), U[-1,1) re/im; without it the
// reference's unseeded rand()/RAND_MAX in [0,1] is used, examples/helper.cpp:26-34).
//
// Verification follows examples/helper.cpp:114-199: an FFTW single-precision
// plan_many_dft on the same input, 1D output un-bit-reversed, SNR > 120 dB to
// pass.  FFTW's API is loaded at run time from libfftw3f, or from MKL's
// FFTW3-compatible interface (libmkl_rt); without either, verification is
// reported as unavailable.
//
// Measurements follow examples/helper.cpp:208-256 but print both unit
// conventions: the reference's (2^30 flop, 2^20 "GB" = MiB) and SI (10^9).
#include <dlfcn.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fftfpga/fftfpga.h"

namespace {

struct Config {
  unsigned num = 64, dim = 3, iter = 1, batch = 1;
  bool inv = false, noverify = false, burst = false, use_bram = false, use_usm = false,
       emulate = false;
  std::string path;
  bool seeded = false;
  uint64_t seed = 0;
};

void usage() {
  std::puts(
      "Offloading FFT on an MI355X through the FFTFPGA C API\n"
      "  -n, --num N       points per dimension (default 64)\n"
      "  -d, --dim D       dimensions 1..3 (default 3)\n"
      "  -b, --back        backward (inverse) transform\n"
      "  -i, --iter I      iterations (default 1)\n"
      "  -p, --path P      path given to fpga_initialize (any readable file)\n"
      "  -y, --noverify    skip the FFTW comparison\n"
      "  -c, --batch B     transforms per call (default 1)\n"
      "  -t, --burst       interleaving flag of the _bram/_batch calls\n"
      "  -m, --use_bram    use the *_bram entry points\n"
      "  -s, --use_usm     use the *_svm entry points (host-mapped staging)\n"
      "  -e, --emulate     pass the emulation platform name\n"
      "      --seed S      seeded U[-1,1) input instead of rand()\n"
      "  -h, --help        this text");
}

bool parse(int argc, char** argv, Config& c) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    std::string val;
    auto eq = a.find('=');
    if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
      val = a.substr(eq + 1);
      a = a.substr(0, eq);
    }
    auto need = [&](const char* name) -> const char* {
      if (!val.empty()) return val.c_str();
      if (i + 1 >= argc) {
        std::fprintf(stderr, "Error parsing options: %s needs a value\n", name);
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "-n" || a == "--num") c.num = (unsigned)std::strtoul(need("num"), nullptr, 10);
    else if (a == "-d" || a == "--dim") c.dim = (unsigned)std::strtoul(need("dim"), nullptr, 10);
    else if (a == "-b" || a == "--back") c.inv = true;
    else if (a == "-i" || a == "--iter") c.iter = (unsigned)std::strtoul(need("iter"), nullptr, 10);
    else if (a == "-p" || a == "--path") c.path = need("path");
    else if (a == "-y" || a == "--noverify") c.noverify = true;
    else if (a == "-c" || a == "--batch") c.batch = (unsigned)std::strtoul(need("batch"), nullptr, 10);
    else if (a == "-t" || a == "--burst") c.burst = true;
    else if (a == "-m" || a == "--use_bram") c.use_bram = true;
    else if (a == "-s" || a == "--use_usm") c.use_usm = true;
    else if (a == "-e" || a == "--emulate") c.emulate = true;
    else if (a == "--seed") {
      c.seeded = true;
      c.seed = std::strtoull(need("seed"), nullptr, 10);
    } else if (a == "-h" || a == "--help") {
      usage();
      std::exit(0);
    } else {
      std::fprintf(stderr, "Error parsing options: unknown option %s\n", argv[i]);
      return false;
    }
  }
  if (c.path.empty()) {
    std::fprintf(stderr, "Error parsing options: please input path to bitstream. Exiting! \n");
    return false;
  }
  if (c.dim < 1 || c.dim > 3 || c.iter == 0 || c.batch == 0) {
    std::fprintf(stderr, "Error parsing options: dim must be 1..3, iter and batch >= 1\n");
    return false;
  }
  return true;
}

void print_config(const Config& c) {
  std::printf("\n------------------------------------------\n");
  std::printf("FFT CONFIGURATION: \n");
  std::printf("--------------------------------------------\n");
  std::printf("Type               : Complex to Complex\n");
  std::printf("Points             : %u%s \n", c.num, c.dim == 1 ? "" : c.dim == 2 ? "^2" : "^3");
  std::printf("Direction          : %s \n", c.inv ? "Backward" : "Forward");
  std::printf("Placement          : Out of Place \n");
  std::printf("Batch              : %u \n", c.batch);
  std::printf("Iterations         : %u \n", c.iter);
  std::printf("Transpose3D        : %s \n", c.use_bram ? "BRAM" : "DDR");
  std::printf("Burst Interleaving : %s \n", c.burst ? "Yes" : "No");
  std::printf("Emulation          : %s \n", c.emulate ? "Yes" : "No");
  std::printf("USM Feature        : %s \n", c.use_usm ? "Yes" : "No");
  std::printf("--------------------------------------------\n\n");
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void create_data(float2* inp, size_t n, const Config& c) {
  if (c.seeded) {
    uint64_t s = c.seed;
    for (size_t i = 0; i < n; ++i) {
      inp[i].x = 2.0f * ((float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f)) - 1.0f;
      inp[i].y = 2.0f * ((float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f)) - 1.0f;
    }
  } else {
    for (size_t i = 0; i < n; ++i) {
      inp[i].x = (float)rand() / (float)RAND_MAX;
      inp[i].y = (float)rand() / (float)RAND_MAX;
    }
  }
}

unsigned bit_reversed(unsigned x, unsigned bits) {
  unsigned y = 0;
  for (unsigned i = 0; i < bits; i++) {
    y = (y << 1) | (x & 1u);
    x >>= 1;
  }
  return y;
}

// ---- FFTW single-precision API, resolved at run time -----------------------------
struct Fftw {
  void* h = nullptr;
  const char* label = "";
  void* (*plan_many)(int, const int*, int, float*, const int*, int, int, float*, const int*, int, int, int,
                     unsigned) = nullptr;
  void (*execute)(void*) = nullptr;
  void (*destroy)(void*) = nullptr;
  bool load() {
    const char* cands[][2] = {{"libfftw3f.so.3", "fftwf (FFTW3)"},
                              {"libfftw3f.so", "fftwf (FFTW3)"},
                              {"libmkl_rt.so", "fftwf API (MKL)"},
                              {"/opt/conda/lib/libmkl_rt.so", "fftwf API (MKL)"}};
    for (auto& c : cands) {
      h = dlopen(c[0], RTLD_NOW | RTLD_LOCAL);
      if (!h) continue;
      plan_many = reinterpret_cast<decltype(plan_many)>(dlsym(h, "fftwf_plan_many_dft"));
      execute = reinterpret_cast<decltype(execute)>(dlsym(h, "fftwf_execute"));
      destroy = reinterpret_cast<decltype(destroy)>(dlsym(h, "fftwf_destroy_plan"));
      if (plan_many && execute && destroy) {
        label = c[1];
        return true;
      }
      dlclose(h);
      h = nullptr;
    }
    return false;
  }
};

// examples/helper.cpp:114-199; returns 1 pass, 0 fail, -1 unavailable
int verify_fftwf(const float2* inp, const float2* out, const Config& c, double* snr) {
  static Fftw fw;
  static bool tried = false, ok = false;
  if (!tried) {
    ok = fw.load();
    tried = true;
    if (ok) std::printf("-- verification with %s\n", fw.label);
  }
  if (!ok) return -1;
  const size_t sz = (size_t)std::pow((double)c.num, (double)c.dim);
  const size_t total = sz * c.batch;
  std::vector<float2> ref(inp, inp + total);
  std::vector<int> n(c.dim, (int)c.num);
  const int FFTW_FORWARD = -1, FFTW_BACKWARD = 1;
  const unsigned FFTW_ESTIMATE = 1u << 6;
  void* plan = fw.plan_many((int)c.dim, n.data(), (int)c.batch, reinterpret_cast<float*>(ref.data()), nullptr, 1,
                            (int)sz, reinterpret_cast<float*>(ref.data()), nullptr, 1, (int)sz,
                            c.inv ? FFTW_BACKWARD : FFTW_FORWARD, FFTW_ESTIMATE);
  if (!plan) return -1;
  fw.execute(plan);
  fw.destroy(plan);
  std::vector<float2> got(out, out + total);
  if (c.dim == 1) {  // the 1D output is bit-reversed per transform
    const unsigned lg = (unsigned)std::lround(std::log2((double)c.num));
    for (unsigned j = 0; j < c.batch; ++j)
      for (unsigned i = 0; i < c.num; ++i) got[(size_t)j * c.num + i] = out[(size_t)j * c.num + bit_reversed(i, lg)];
  }
  double mag = 0, noise = 0;
  for (size_t i = 0; i < total; ++i) {
    const double dx = (double)ref[i].x - got[i].x, dy = (double)ref[i].y - got[i].y;
    mag += (double)ref[i].x * ref[i].x + (double)ref[i].y * ref[i].y;
    noise += dx * dx + dy * dy;
  }
  *snr = noise > 0 ? 10.0 * std::log10(mag / noise) : INFINITY;
  return *snr > 120.0 ? 1 : 0;
}

void perf_measures(const Config& c, const std::vector<fpga_t>& rt) {
  double ex = 0, rd = 0, wr = 0;
  for (auto& t : rt) {
    ex += t.exec_t;
    rd += t.pcie_read_t;
    wr += t.pcie_write_t;
  }
  const double n = (double)rt.size();
  ex /= n;
  rd /= n;
  wr /= n;
  double vex = 0, vrd = 0, vwr = 0;
  for (auto& t : rt) {
    vex += (t.exec_t - ex) * (t.exec_t - ex);
    vrd += (t.pcie_read_t - rd) * (t.pcie_read_t - rd);
    vwr += (t.pcie_write_t - wr) * (t.pcie_write_t - wr);
  }
  const double pts = (double)c.batch * std::pow((double)c.num, (double)c.dim);
  const double flop = (double)c.batch * c.dim * 5.0 * std::pow((double)c.num, (double)c.dim) * std::log2((double)c.num);
  const double s = ex * 1e-3;
  std::printf("\n\n------------------------------------------\n");
  std::printf("Measurements \n");
  std::printf("--------------------------------------------\n");
  if (c.iter > 1) std::printf("Average Measurements of iterations\n");
  std::printf("PCIe Write          = %.4lfms\n", wr);
  std::printf("Kernel Execution    = %.4lfms\n", ex);
  std::printf("Kernel Exec/Batch   = %.4lfms\n", ex / c.batch);
  std::printf("PCIe Read           = %.4lfms\n", rd);
  std::printf("Total               = %.4lfms\n", ex + wr + rd);
  std::printf("Throughput          = %.4lf GFLOP/s (10^9) | %.4lf GB/s (10^9 B, read+write)\n", flop / s / 1e9,
              16.0 * pts / s / 1e9);
  std::printf("Reference units     = %.4lf GFLOPS/s (2^30) | %.4lf \"GB/s\" (points*8 / 2^20)\n",
              flop / s / (1024.0 * 1024 * 1024), pts * 8 / s / (1024.0 * 1024));
  if (c.iter > 1) {
    std::printf("\nDeviation of runtimes among iterations\n");
    std::printf("PCIe Write          = %.4lfms\n", std::sqrt(vwr / n));
    std::printf("Kernel Execution    = %.4lfms\n", std::sqrt(vex / n));
    std::printf("PCIe Read           = %.4lfms\n", std::sqrt(vrd / n));
  }
}

}  // namespace

int main(int argc, char** argv) {
  Config c;
  if (!parse(argc, argv, c)) return 1;
  print_config(c);
  const char* platform = c.emulate ? "Intel(R) FPGA Emulation Platform for OpenCL(TM)"
                                   : "Intel(R) FPGA SDK for OpenCL(TM)";
  const int rc = fpga_initialize(platform, c.path.c_str(), c.use_usm);
  if (rc != 0) {
    std::fprintf(stderr, "FPGA initialization error (%d)\n", rc);
    return EXIT_FAILURE;
  }
  const size_t sz = (size_t)c.batch * (size_t)std::llround(std::pow((double)c.num, (double)c.dim));
  float2* inp = static_cast<float2*>(fftfpgaf_complex_malloc(sizeof(float2) * sz));
  float2* out = static_cast<float2*>(fftfpgaf_complex_malloc(sizeof(float2) * sz));
  if (!inp || !out) {
    std::fprintf(stderr, "allocation failed\n");
    return EXIT_FAILURE;
  }
  create_data(inp, sz, c);
  std::vector<fpga_t> rt;
  int status = EXIT_SUCCESS;
  for (unsigned it = 0; it < c.iter; ++it) {
    std::printf("%u: Calculating FFT - \n", it);
    fpga_t t{};
    switch (c.dim) {
      case 1:
        t = c.use_usm ? fftfpgaf_c2c_1d_svm(c.num, inp, out, c.inv, c.batch)
                      : fftfpgaf_c2c_1d(c.num, inp, out, c.inv, c.batch);
        break;
      case 2:
        if (c.use_bram && c.use_usm) t = fftfpgaf_c2c_2d_bram_svm(c.num, inp, out, c.inv, c.batch);
        else if (c.use_bram) t = fftfpgaf_c2c_2d_bram(c.num, inp, out, c.inv, c.burst, c.batch);
        else t = fftfpgaf_c2c_2d_ddr(c.num, inp, out, c.inv);
        break;
      default:
        if (c.use_bram) t = fftfpgaf_c2c_3d_bram(c.num, inp, out, c.inv, c.burst);
        else if (!c.use_usm && c.batch > 1) t = fftfpgaf_c2c_3d_ddr_batch(c.num, inp, out, c.inv, c.burst, c.batch);
        else if (c.use_usm && c.batch > 1) t = fftfpgaf_c2c_3d_ddr_svm_batch(c.num, inp, out, c.inv, c.batch);
        else if (c.use_usm) t = fftfpgaf_c2c_3d_ddr_svm(c.num, inp, out, c.inv, c.burst);
        else t = fftfpgaf_c2c_3d_ddr(c.num, inp, out, c.inv);
        break;
    }
    if (!t.valid) {
      std::fprintf(stderr, "Iter %u: transform call returned valid = false\n", it);
      status = EXIT_FAILURE;
      break;
    }
    rt.push_back(t);
    if (!c.noverify) {
      double snr = 0;
      const int v = verify_fftwf(inp, out, c, &snr);
      if (v == 0) {
        std::fprintf(stderr, "Iter %u: result incorrect in comparison to FFTW (SNR %.1f dB)\n", it, snr);
        status = EXIT_FAILURE;
        break;
      }
      if (v == 1) std::printf("\tSignal to noise ratio vs FFTW: %.1f dB --> PASSED\n", snr);
      if (v < 0) std::printf("\tFFTW not available: verification skipped\n");
    }
  }
  fpga_final();
  if (!rt.empty()) perf_measures(c, rt);
  std::free(inp);
  std::free(out);
  return status;
}

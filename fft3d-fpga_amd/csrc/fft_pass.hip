// fft_pass.hip -- the gfx950 Stockham line-FFT pass kernels.
//
// One launch transforms a batch of lines of L = 2^lg points along one axis of
// an array in HBM.  This replaces the reference's per-axis engine pipelines
// (kernels/fft3d/fft3d_ddr.cl fft3da/b/c, kernels/fft2d/fft2d_ddr.cl fft2d,
// kernels/fft1d/fft1d.cl fft1d, all built on kernels/common/fft_8.cl
// fft_step) and its fetch/transpose/store index-map kernels: on the GPU the
// strided axes are read directly in 128-byte row segments, so no separate
// transpose buffer or bit-reversal kernel exists.
//
// The tile work (addressing, radix stages, epilogues) is pass_block() in
// fft_pass_dev.h; this file holds the one-tile-per-workgroup kernels and the
// launch table over (precision, line length, mode, epilogue).
#include <hip/hip_runtime.h>

#include "launch.h"

#include "fft_pass_dev.h"

namespace fftamd {

template <typename R, int L, int MODE, int EPI, bool SPLIT, int DIR>
__global__ void __launch_bounds__(geom_threads_for(L, MODE, SPLIT))
    k_fft_pass(const PassArgs a) {
  constexpr int CC = geom_cols_for(L, MODE, SPLIT);
  using G = Geom<L, MODE, CC>;
  constexpr bool NEED_LDS = G::NP > 1;
  __shared__ cpx<R> lds[NEED_LDS ? G::LDS_ELEMS : 1];
  pass_block<R, L, MODE, EPI, SPLIT, DIR, CC>(a, blockIdx.x, lds, (int)threadIdx.x);
}

// ---- launch table ------------------------------------------------------------

template <typename R, int MODE, int EPI, int LG, bool SPLIT>
static void launch_one(const PassArgs& a, unsigned long long blocks, hipStream_t s) {
  constexpr int L = 1 << LG;
  // one kernel per direction: the direction swaps become register renames
  constexpr int NT = geom_threads_for(L, MODE, SPLIT);
  if (a.inverse)
    FFTAMD_LAUNCH((k_fft_pass<R, L, MODE, EPI, SPLIT, 1>), dim3((unsigned)blocks), dim3(NT), 0, s, a);
  else
    FFTAMD_LAUNCH((k_fft_pass<R, L, MODE, EPI, SPLIT, 0>), dim3((unsigned)blocks), dim3(NT), 0, s, a);
}

template <typename R, int MODE, int EPI, int LG, bool SPLIT>
static bool dispatch_lg(int lg, const PassArgs& a, unsigned long long blocks, hipStream_t s) {
  if constexpr (LG > 13) {
    return false;
  } else {
    if (lg == LG) {
      constexpr int L = 1 << LG;
      constexpr int prec = sizeof(R) == 4 ? PREC_F32 : PREC_F64;
      // split mode exists for FP32 lines with at least one radix-16 stage
      constexpr bool split_ok = !SPLIT || (sizeof(R) == 4 && L >= 16 && L <= 4096);
      constexpr bool ok = split_ok &&
                          (MODE == MODE_CONTIG ? L <= lmax_contig(prec)
                                                : L <= (EPI == EPI_TWIDDLE ? lmax_strided_fourstep(prec)
                                                                           : lmax_strided(prec))) &&
                          !(MODE == MODE_CONTIG && EPI == EPI_TWIDDLE) &&
                          !(MODE == MODE_STRIDED && EPI == EPI_BITREV);
      if constexpr (ok) {
        launch_one<R, MODE, EPI, LG, SPLIT>(a, blocks, s);
        return true;
      } else {
        return false;
      }
    }
    return dispatch_lg<R, MODE, EPI, LG + 1, SPLIT>(lg, a, blocks, s);
  }
}

template <typename R, bool SPLIT>
static bool dispatch_prec(int lg, int mode, int epi, const PassArgs& a, unsigned long long blocks,
                          hipStream_t s) {
  if (mode == MODE_CONTIG) {
    if (epi == EPI_NONE) return dispatch_lg<R, MODE_CONTIG, EPI_NONE, 1, SPLIT>(lg, a, blocks, s);
    if (epi == EPI_BITREV) return dispatch_lg<R, MODE_CONTIG, EPI_BITREV, 1, SPLIT>(lg, a, blocks, s);
  } else {
    if (epi == EPI_NONE) return dispatch_lg<R, MODE_STRIDED, EPI_NONE, 1, SPLIT>(lg, a, blocks, s);
    if (epi == EPI_TWIDDLE) return dispatch_lg<R, MODE_STRIDED, EPI_TWIDDLE, 1, SPLIT>(lg, a, blocks, s);
  }
  return false;
}

void preload_pass() { preload_kernel(reinterpret_cast<const void*>(&k_fft_pass<float, 256, MODE_CONTIG, EPI_NONE, false, 0>)); }

int launch_fft_pass(int prec, int lg, int mode, int epi, const PassArgs& a, unsigned long long blocks,
                    void* stream, bool split) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  bool ok;
  if (prec == PREC_F32)
    ok = split ? dispatch_prec<float, true>(lg, mode, epi, a, blocks, s)
               : dispatch_prec<float, false>(lg, mode, epi, a, blocks, s);
  else
    ok = !split && dispatch_prec<double, false>(lg, mode, epi, a, blocks, s);
  if (!ok) return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fftamd

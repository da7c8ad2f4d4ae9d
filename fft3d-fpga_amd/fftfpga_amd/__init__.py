"""fftfpga_amd -- MI355X-native (gfx950) drop-in for FFTFPGA's single-precision
C2C transform path.

The native library lib/libfftfpga.so exports the reference's C ABI
(include/fftfpga/fftfpga.h); this package is its Python mirror plus
device-pointer helpers for torch tensors.
"""
from ._lib import LIB_PATH, SIGNATURES, FpgaT, last_error, load
from .api import (
    FFTAMD_BITREV_1D,
    FFTAMD_NO_CUBE_SPLIT,
    FFTAMD_NO_PLANE_FUSION,
    FFTAMD_SPLIT_MFMA,
    c2c_dev,
    fftfpga_c2c_1d,
    fftfpga_complex_malloc,
    fftfpga_init_fpga,
    fftfpgaf_c2c_1d,
    fftfpgaf_c2c_1d_svm,
    fftfpgaf_c2c_2d_bram,
    fftfpgaf_c2c_2d_bram_svm,
    fftfpgaf_c2c_2d_ddr,
    fftfpgaf_c2c_3d_bram,
    fftfpgaf_c2c_3d_ddr,
    fftfpgaf_c2c_3d_ddr_batch,
    fftfpgaf_c2c_3d_ddr_svm,
    fftfpgaf_c2c_3d_ddr_svm_batch,
    fftfpgaf_complex_malloc,
    fpga_final,
    fpga_initialize,
)

__all__ = [name for name in dir() if not name.startswith("_")]

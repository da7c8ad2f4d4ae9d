"""Drop-in contract of the C ABI, runnable without a GPU.

Ports the reference's gtest suite (tests/test_fft_setup.cpp,
tests/test_fft1d_fpga.cpp, tests/test_fft2d_fpga.cpp, tests/test_fft3d_fpga.cpp):
init return codes, malloc(0) -> NULL, and valid == false for NULL pointers,
non-power-of-two N and bad batch counts.  Argument checks happen before any
device call, so these run on a machine without a GPU.  Also checks that the
library exports every symbol include/fftfpga/*.h declares, that the struct
layouts seen by C and by ctypes agree, and that a C caller compiled against
the header links and runs the same contract.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
HEADERS = [os.path.join(INC, "fftfpga", h) for h in ("fftfpga.h", "fftfpga_hip.h")]


def declared_functions():
    names = set()
    # return type of one or more words and stars, then the name and "("
    pat = re.compile(r"^\s*(?:extern\s+)?(?:[A-Za-z_]\w*[\s*]+)+([A-Za-z_]\w*)\s*\(")
    for h in HEADERS:
        for line in open(h):
            if line.lstrip().startswith(("*", "/*", "//", "#")):
                continue
            m = pat.match(line)
            if m and m.group(1) not in ("if", "return", "sizeof"):
                names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(fl):
    lib = fl.load()
    names = declared_functions()
    assert "fftfpgaf_c2c_3d_ddr" in names and "fftamd_c2c_dev" in names
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared in include/ but not exported"
    assert names == set(fl.SIGNATURES), "ctypes signature table out of sync with the headers"


def test_no_reference_or_oracle_symbols_in_library(fl):
    out = subprocess.run(["nm", "-D", "--defined-only", fl.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert "orc_" not in out  # the oracle is never linked into the product


def _cc(tmp_path, src, name, link=False):
    c = tmp_path / f"{name}.c"
    c.write_text(src)
    exe = tmp_path / name
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror", "-I", INC, str(c), "-o", str(exe)]
    if link:
        cmd += ["-L", os.path.dirname(fl_lib_path()), "-lfftfpga",
                "-Wl,-rpath," + os.path.dirname(fl_lib_path())]
    subprocess.run(cmd, check=True)
    return exe


def fl_lib_path():
    import fftfpga_amd
    return fftfpga_amd.LIB_PATH


def test_struct_layout_matches_ctypes(tmp_path, fl):
    exe = _cc(tmp_path, r"""
#include <stdio.h>
#include <stddef.h>
#include "fftfpga/fftfpga.h"
int main(void){
  printf("%zu %zu %zu %zu %zu\n", sizeof(fpga_t), offsetof(fpga_t, exec_t),
         offsetof(fpga_t, valid), sizeof(float2), sizeof(double2));
  return 0;
}""", "layout")
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals == [ctypes.sizeof(fl.FpgaT), fl.FpgaT.exec_t.offset, fl.FpgaT.valid.offset, 8, 16]


def test_c_caller_links_and_sees_contract(tmp_path, fl):
    """A C program written against the header, as examples/fft.cpp is."""
    exe = _cc(tmp_path, r"""
#include <stdio.h>
#include <stdlib.h>
#include "fftfpga/fftfpga.h"
int main(int argc, char **argv){
  (void)argc;
  float2 *buf = (float2*)fftfpgaf_complex_malloc(sizeof(float2) * 64 * 64 * 64);
  int bad = 0;
  bad += fftfpgaf_complex_malloc(0) != NULL;
  bad += fftfpga_complex_malloc(0) != NULL;
  bad += ((size_t)buf % 64) != 0;
  bad += fpga_initialize("Intel(R) FPGA", "", false) != -1;
  bad += fpga_initialize("TEST", argv[0], false) != -2;
  bad += fftfpgaf_c2c_3d_ddr(64, NULL, buf, false).valid;
  bad += fftfpgaf_c2c_3d_ddr(64, buf, NULL, false).valid;
  bad += fftfpgaf_c2c_3d_ddr(63, buf, buf, false).valid;
  bad += fftfpgaf_c2c_1d(64, NULL, buf, false, 1).valid;
  bad += fftfpgaf_c2c_2d_ddr(63, buf, buf, false).valid;
  bad += fftfpgaf_c2c_3d_ddr_batch(64, buf, buf, false, false, 1).valid;
  printf("%d %d\n", bad, fpga_initialize("Intel(R) FPGA Emulation Platform for OpenCL(TM)",
                                          argv[0], false));
  fpga_final();
  free(buf);
  return 0;
}""", "caller", link=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out[0] == "0"
    assert int(out[1]) in (0, -3)  # 0 with a GPU, -3 (no device) without


# ---- tests/test_fft_setup.cpp -----------------------------------------------------------

def test_valid_init_codes(fl):
    assert fl.fpga_initialize("Intel(R) FPGA", "", False) == -1
    assert fl.fpga_initialize("Intel(R) FPGA", None, False) == -1
    assert fl.fpga_initialize("TEST", "fft1d_emulate.aocx", False) == -2
    assert fl.fpga_initialize("", fl.LIB_PATH, False) == -2
    assert fl.fpga_initialize(None, fl.LIB_PATH, False) == -2
    # names are matched case-insensitively as substrings (opencl_utils.c:24-85)
    rc = fl.fpga_initialize("intel(r) fpga emulation", fl.LIB_PATH, False)
    assert rc in (0, -3)
    fl.fpga_final()


def test_malloc_zero_is_null(fl):
    assert fl.fftfpgaf_complex_malloc(0) is None
    assert fl.fftfpga_complex_malloc(0) is None
    p = fl.fftfpgaf_complex_malloc(1000)
    assert p is not None and p % 64 == 0
    ctypes.CDLL(None).free(ctypes.c_void_p(p))


# ---- tests/test_fft{1,2,3}d_fpga.cpp InputValidity ----------------------------------------

def _bufs(npts):
    return np.zeros(npts, np.complex64), np.zeros(npts, np.complex64)


def test_input_validity_1d(fl):
    N = 64
    a, b = _bufs(N)
    assert not fl.fftfpgaf_c2c_1d(N, None, b, False, 1).valid
    assert not fl.fftfpgaf_c2c_1d(N, a, None, False, 1).valid
    assert not fl.fftfpgaf_c2c_1d(N - 1, a, a, False, 1).valid
    assert not fl.fftfpgaf_c2c_1d(0, a, a, False, 1).valid
    assert not fl.fftfpgaf_c2c_1d(N, a, b, False, 0).valid
    d = np.zeros(N, np.complex128)
    assert not fl.fftfpga_c2c_1d(N, None, d, False, 1).valid
    assert not fl.fftfpga_c2c_1d(N - 1, d, d, False, 1).valid


def test_input_validity_1d_svm(fl):
    N = 64
    a, _ = _bufs(N)
    # svm not enabled
    assert not fl.fftfpgaf_c2c_1d_svm(N, a, a, False, 1).valid
    assert not fl.fftfpgaf_c2c_1d_svm(N, None, a, False, 1).valid
    assert not fl.fftfpgaf_c2c_1d_svm(N - 1, a, a, False, 1).valid


def test_input_validity_2d(fl):
    N = 64
    a, _ = _bufs(N * N)
    assert not fl.fftfpgaf_c2c_2d_bram(64, None, a, False, False, 1).valid
    assert not fl.fftfpgaf_c2c_2d_bram(64, a, None, False, False, 1).valid
    assert not fl.fftfpgaf_c2c_2d_bram(63, a, a, False, False, 1).valid
    assert not fl.fftfpgaf_c2c_2d_ddr(64, None, a, False).valid
    assert not fl.fftfpgaf_c2c_2d_ddr(64, a, None, False).valid
    assert not fl.fftfpgaf_c2c_2d_ddr(63, a, a, False).valid
    assert not fl.fftfpgaf_c2c_2d_bram_svm(64, a, a, False, 1).valid


def test_input_validity_3d(fl):
    N = 64
    a, _ = _bufs(2 * N ** 3)
    assert not fl.fftfpgaf_c2c_3d_bram(64, None, a, False, False).valid
    assert not fl.fftfpgaf_c2c_3d_bram(64, a, None, False, False).valid
    assert not fl.fftfpgaf_c2c_3d_bram(63, a, a, False, False).valid
    assert not fl.fftfpgaf_c2c_3d_ddr(64, None, a, False).valid
    assert not fl.fftfpgaf_c2c_3d_ddr(64, a, None, False).valid
    assert not fl.fftfpgaf_c2c_3d_ddr(63, a, a, False).valid
    # batch: how_many must exceed 1 (api/src/fft3d.c:403)
    assert not fl.fftfpgaf_c2c_3d_ddr_batch(64, a, a, False, False, 1).valid
    assert not fl.fftfpgaf_c2c_3d_ddr_batch(64, a, a, False, False, 0).valid
    assert not fl.fftfpgaf_c2c_3d_ddr_batch(63, a, a, False, False, 2).valid
    # svm batch (tests/test_fft3d_fpga.cpp:66-95), including how_many = -1 as unsigned
    assert not fl.fftfpgaf_c2c_3d_ddr_svm_batch(64, None, a, False, 2).valid
    assert not fl.fftfpgaf_c2c_3d_ddr_svm_batch(64, a, None, False, 2).valid
    assert not fl.fftfpgaf_c2c_3d_ddr_svm_batch(63, a, a, False, 2).valid
    assert not fl.fftfpgaf_c2c_3d_ddr_svm_batch(63, a, a, False, 0).valid
    assert not fl.fftfpgaf_c2c_3d_ddr_svm(64, a, a, False, False).valid


def test_uninitialised_call_is_invalid_not_fatal(fl):
    """The reference exits inside checkError; here the call returns valid=0."""
    fl.fpga_final()
    a, b = _bufs(8 ** 3)
    assert not fl.fftfpgaf_c2c_3d_ddr(8, a, b, False).valid
    assert "initialis" in fl.last_error()


def test_python_side_size_check(fl):
    a = np.zeros(10, np.complex64)
    with pytest.raises(ValueError):
        fl.fftfpgaf_c2c_3d_ddr(8, a, a, False)
    with pytest.raises(TypeError):
        fl.fftfpgaf_c2c_3d_ddr(2, np.zeros(8, np.complex128), a, False)


def test_partially_overlapping_device_buffers_refused(fl):
    """Device entry points refuse input and output ranges that overlap without
    being the same buffer (the first pass may run out of place), before any
    device work: the pointers below are never dereferenced."""
    import ctypes
    lib = fl.load()
    dims = (ctypes.c_uint * 3)(64, 64, 64)
    base = 0x10000000
    for lo, hi in ((base, base + 8), (base + 8 * 64 ** 3 - 8, base)):
        rc = lib.fftamd_c2c_dev(3, dims, 1, ctypes.c_void_p(lo), ctypes.c_void_p(hi), 0, 0, None)
        assert rc == -1 and "overlap" in fl.last_error()
        rc = lib.fftamd_c2c_weighted_dev(3, dims, 1, ctypes.c_void_p(lo), ctypes.c_void_p(hi), 0,
                                         None, None, None)
        assert rc == -1 and "overlap" in fl.last_error()
    # FP64: 16-byte points, so offsets within 16 * N^3 bytes overlap
    rc = lib.fftamd_z2z_dev(3, dims, 1, ctypes.c_void_p(base), ctypes.c_void_p(base + 8 * 64 ** 3), 0, 0,
                            None)
    assert rc == -1 and "overlap" in fl.last_error()

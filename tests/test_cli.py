"""The `fft` harness (fft3d-fpga_amd/apps/fft.cpp), counterpart of the
reference's examples/fft: same flags and dispatch, FFTW-based verification
(SNR > 120 dB, examples/helper.cpp:164-198), measurement printout."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "fft3d-fpga_amd", "bin", "fft")
LIB = os.path.join(ROOT, "fft3d-fpga_amd", "lib", "libfftfpga.so")


def run(*args, timeout=120):
    if not os.path.exists(APP):
        pytest.fail("harness not built: make -C fft3d-fpga_amd")
    return subprocess.run([APP, *args], capture_output=True, text=True, timeout=timeout)


def test_help_and_missing_path():
    r = run("-h")
    assert r.returncode == 0 and "--num" in r.stdout and "--use_usm" in r.stdout
    r = run("-n", "64")
    assert r.returncode == 1 and "path" in r.stderr


def test_bad_arguments_rejected():
    assert run("-d", "4", "-p", LIB).returncode == 1
    assert run("--bogus", "-p", LIB).returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ("-n", "64", "-d", "3", "--seed", "0"),
    ("-n", "64", "-d", "3", "-b"),
    ("-n", "256", "-d", "2", "-i", "3"),
    ("-n", "1024", "-d", "1", "-c", "5"),
    ("-n", "32", "-d", "3", "-c", "4"),
    ("-n", "32", "-d", "3", "-m", "-t"),
    ("-n", "64", "-d", "2", "-m", "-c", "3"),
    ("-n", "32", "-d", "3", "-s"),
    ("-n", "32", "-d", "3", "-s", "-c", "2"),
    ("-n", "128", "-d", "1", "-s", "-c", "3"),
    ("-n", "64", "-d", "3", "-e", "--seed", "7"),
])
def test_harness_runs_and_verifies(args):
    r = run(*args, "-p", LIB)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Kernel Execution" in r.stdout
    if _fftwf_api_loads():
        # the harness finds the same library: its SNR check must run and pass
        assert "PASSED" in r.stdout, r.stdout
    else:
        assert "verification skipped" in r.stdout, r.stdout


def _fftwf_api_loads():
    """Whether the FFTW single-precision API loads on this host from the
    libraries the harness tries (apps/fft.cpp: libfftw3f, then MKL's FFTW3
    interface in libmkl_rt)."""
    import ctypes
    for name in ("libfftw3f.so.3", "libfftw3f.so", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so"):
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        if all(hasattr(lib, f) for f in ("fftwf_plan_many_dft", "fftwf_execute", "fftwf_destroy_plan")):
            return True
    return False


def test_fftwf_probe_matches_harness_search_order():
    """The probe above tries the harness's candidates in the harness's order."""
    src = open(os.path.join(ROOT, "fft3d-fpga_amd", "apps", "fft.cpp")).read()
    for name in ("libfftw3f.so.3", "libfftw3f.so", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so"):
        assert f'"{name}"' in src

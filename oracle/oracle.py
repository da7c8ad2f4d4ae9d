"""ctypes view of the C oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  This module is the parity checker: tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py may load
it; the shipped library and its Python binding never do.

Every function mirrors one in oracle.h; the reference citations live there.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with gcc (seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        szp = ctypes.POINTER(ctypes.c_size_t)
        L.orc_splitmix64_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.orc_splitmix64_next.restype = ctypes.c_uint64
        L.orc_fill_uniform.argtypes = [f32p, ctypes.c_size_t, ctypes.c_uint64]
        L.orc_fill_gauss.argtypes = [f32p, ctypes.c_size_t, ctypes.c_uint64]
        L.orc_bit_reversed.argtypes = [ctypes.c_uint, ctypes.c_uint]
        L.orc_bit_reversed.restype = ctypes.c_uint
        L.orc_fft1d_f64.argtypes = [f64p, ctypes.c_size_t, ctypes.c_int]
        L.orc_fftnd_f32in.argtypes = [f32p, f64p, ctypes.c_int, szp, ctypes.c_size_t, ctypes.c_int]
        L.orc_fftnd_f64.argtypes = [f64p, f64p, ctypes.c_int, szp, ctypes.c_size_t, ctypes.c_int]
        L.orc_dftnd_naive.argtypes = [f64p, f64p, ctypes.c_int, szp, ctypes.c_int]
        L.orc_c2c_1d_expected.argtypes = [f32p, f64p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
        L.orc_rel_l2.argtypes = [f32p, f64p, ctypes.c_size_t]
        L.orc_rel_l2.restype = ctypes.c_double
        L.orc_snr_db.argtypes = [f32p, f64p, ctypes.c_size_t]
        L.orc_snr_db.restype = ctypes.c_double
        for fn in ("orc_fft1d_f64", "orc_fftnd_f32in", "orc_fftnd_f64", "orc_dftnd_naive",
                   "orc_c2c_1d_expected"):
            getattr(L, fn).restype = ctypes.c_int
        _lib = L
    return _lib


def _dims(dims):
    arr = (ctypes.c_size_t * len(dims))(*dims)
    return arr


def _as_ri32(x: np.ndarray) -> np.ndarray:
    """complex64 array -> contiguous float32 view (re, im interleaved)."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    return x.view(np.float32).reshape(-1)


def uniform(npts: int, seed: int) -> np.ndarray:
    """complex64[npts], re/im i.i.d. U[-1,1) from splitmix64(seed)."""
    buf = np.empty(2 * npts, np.float32)
    lib().orc_fill_uniform(buf, npts, seed)
    return buf.view(np.complex64)


def gauss(npts: int, seed: int) -> np.ndarray:
    """complex64[npts], re/im i.i.d. N(0,1) (Box-Muller in FP64) from splitmix64(seed)."""
    buf = np.empty(2 * npts, np.float32)
    lib().orc_fill_gauss(buf, npts, seed)
    return buf.view(np.complex64)


def bit_reversed(x: int, bits: int) -> int:
    return int(lib().orc_bit_reversed(x, bits))


def fftnd(x: np.ndarray, dims, batch: int = 1, inverse: bool = False) -> np.ndarray:
    """FP64 N-d transform of complex64 (or complex128) input; returns complex128
    with shape (batch, *dims).  inverse=True is the unnormalised backward DFT."""
    dims = [int(d) for d in dims]
    tot = batch * int(np.prod(dims))
    out = np.empty(2 * tot, np.float64)
    if np.asarray(x).dtype == np.complex128:
        src = np.ascontiguousarray(x).view(np.float64).reshape(-1)
        rc = lib().orc_fftnd_f64(src, out, len(dims), _dims(dims), batch, int(inverse))
    else:
        rc = lib().orc_fftnd_f32in(_as_ri32(x), out, len(dims), _dims(dims), batch, int(inverse))
    if rc != 0:
        raise ValueError(f"oracle fftnd rc={rc}")
    return out.view(np.complex128).reshape((batch, *dims))


def fft1d_f64(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    buf = np.ascontiguousarray(x, dtype=np.complex128).copy()
    rc = lib().orc_fft1d_f64(buf.view(np.float64).reshape(-1), buf.size, 1 if inverse else -1)
    if rc != 0:
        raise ValueError(f"oracle fft1d rc={rc}")
    return buf


def dft_naive(x: np.ndarray, dims, inverse: bool = False) -> np.ndarray:
    """Brute-force long-double DFT of one array of shape dims (rank 1..3)."""
    dims = [int(d) for d in dims]
    src = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    out = np.empty(2 * src.size, np.float64)
    rc = lib().orc_dftnd_naive(src.view(np.float64), out, len(dims), _dims(dims), int(inverse))
    if rc != 0:
        raise ValueError(f"oracle naive rc={rc}")
    return out.view(np.complex128).reshape(dims)


def c2c_1d_expected(x: np.ndarray, n: int, batch: int, inverse: bool = False) -> np.ndarray:
    """Expected output of fftfpgaf_c2c_1d (bit-reversed within each transform)."""
    out = np.empty(2 * n * batch, np.float64)
    rc = lib().orc_c2c_1d_expected(_as_ri32(x), out, n, batch, int(inverse))
    if rc != 0:
        raise ValueError(f"oracle 1d rc={rc}")
    return out.view(np.complex128).reshape(batch, n)


def rel_l2(got: np.ndarray, ref: np.ndarray) -> float:
    g = _as_ri32(got)
    r = np.ascontiguousarray(ref, dtype=np.complex128).view(np.float64).reshape(-1)
    if g.size != r.size:
        raise ValueError("size mismatch")
    return float(lib().orc_rel_l2(g, r, g.size // 2))


def snr_db(got: np.ndarray, ref: np.ndarray) -> float:
    g = _as_ri32(got)
    r = np.ascontiguousarray(ref, dtype=np.complex128).view(np.float64).reshape(-1)
    return float(lib().orc_snr_db(g, r, g.size // 2))

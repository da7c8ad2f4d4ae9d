"""Python mirror of the FFTFPGA C API (include/fftfpga/fftfpga.h) and of the
device-pointer extensions (include/fftfpga/fftfpga_hip.h).

Host entry points take numpy arrays (complex64, or complex128 for the double
1D call) or None, and return the library's fpga_t exactly as the C ABI does:
argument errors give ``valid == False`` like the reference
(api/src/fft3d.c:214-216).  Array sizes are checked here before the pointer
crosses into C, since C cannot see a numpy array's length.

Device entry points take torch tensors that already live on the GPU and run
on the current torch stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import FpgaT, last_error, load

FFTAMD_BITREV_1D = 0x1
FFTAMD_SPLIT_MFMA = 0x2
FFTAMD_NO_PLANE_FUSION = 0x4
FFTAMD_NO_CUBE_SPLIT = 0x8


def _ptr(a, dtype, need: int):
    if a is None:
        return None
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"expected a C-contiguous numpy array of {np.dtype(dtype).name}")
    if a.size < need:
        raise ValueError(f"array holds {a.size} points, the transform needs {need}")
    return a.ctypes.data_as(ctypes.c_void_p)


def _need(N: int, dims: int, count: int) -> int:
    return int(N) ** dims * max(int(count), 0) if N > 0 and (N & (N - 1)) == 0 else 0


# ---- setup ---------------------------------------------------------------------

def fpga_initialize(platform_name: str | None, path: str | None, use_svm: bool = False) -> int:
    p = platform_name.encode() if platform_name is not None else None
    q = path.encode() if path is not None else None
    return int(load().fpga_initialize(p, q, bool(use_svm)))


fftfpga_init_fpga = fpga_initialize


def fpga_final() -> None:
    load().fpga_final()


def fftfpgaf_complex_malloc(sz: int) -> int | None:
    """Returns the raw address (release with libc free()), None for sz == 0."""
    return load().fftfpgaf_complex_malloc(sz)


def fftfpga_complex_malloc(sz: int) -> int | None:
    return load().fftfpga_complex_malloc(sz)


# ---- host-pointer transforms -----------------------------------------------------

def fftfpgaf_c2c_1d(N, inp, out, inv=False, iter=1) -> FpgaT:
    n = _need(N, 1, iter)
    return load().fftfpgaf_c2c_1d(N, _ptr(inp, np.complex64, n), _ptr(out, np.complex64, n),
                                  bool(inv), iter)


def fftfpga_c2c_1d(N, inp, out, inv=False, iter=1) -> FpgaT:
    n = _need(N, 1, iter)
    return load().fftfpga_c2c_1d(N, _ptr(inp, np.complex128, n), _ptr(out, np.complex128, n),
                                 bool(inv), iter)


def fftfpgaf_c2c_1d_svm(N, inp, out, inv=False, batch=1) -> FpgaT:
    n = _need(N, 1, batch)
    return load().fftfpgaf_c2c_1d_svm(N, _ptr(inp, np.complex64, n), _ptr(out, np.complex64, n),
                                      bool(inv), batch)


def fftfpgaf_c2c_2d_bram(N, inp, out, inv=False, interleaving=False, how_many=1) -> FpgaT:
    n = _need(N, 2, how_many)
    return load().fftfpgaf_c2c_2d_bram(N, _ptr(inp, np.complex64, n), _ptr(out, np.complex64, n),
                                       bool(inv), bool(interleaving), how_many)


def fftfpgaf_c2c_2d_bram_svm(N, inp, out, inv=False, how_many=1) -> FpgaT:
    n = _need(N, 2, how_many)
    return load().fftfpgaf_c2c_2d_bram_svm(N, _ptr(inp, np.complex64, n),
                                           _ptr(out, np.complex64, n), bool(inv), how_many)


def fftfpgaf_c2c_2d_ddr(N, inp, out, inv=False) -> FpgaT:
    n = _need(N, 2, 1)
    return load().fftfpgaf_c2c_2d_ddr(N, _ptr(inp, np.complex64, n), _ptr(out, np.complex64, n),
                                      bool(inv))


def fftfpgaf_c2c_3d_bram(N, inp, out, inv=False, interleaving=False) -> FpgaT:
    n = _need(N, 3, 1)
    return load().fftfpgaf_c2c_3d_bram(N, _ptr(inp, np.complex64, n), _ptr(out, np.complex64, n),
                                       bool(inv), bool(interleaving))


def fftfpgaf_c2c_3d_ddr(N, inp, out, inv=False) -> FpgaT:
    n = _need(N, 3, 1)
    return load().fftfpgaf_c2c_3d_ddr(N, _ptr(inp, np.complex64, n), _ptr(out, np.complex64, n),
                                      bool(inv))


def fftfpgaf_c2c_3d_ddr_batch(N, inp, out, inv=False, interleaving=False, how_many=2) -> FpgaT:
    n = _need(N, 3, how_many)
    return load().fftfpgaf_c2c_3d_ddr_batch(N, _ptr(inp, np.complex64, n),
                                            _ptr(out, np.complex64, n), bool(inv),
                                            bool(interleaving), how_many & 0xFFFFFFFF)


def fftfpgaf_c2c_3d_ddr_svm(N, inp, out, inv=False, interleaving=False) -> FpgaT:
    n = _need(N, 3, 1)
    return load().fftfpgaf_c2c_3d_ddr_svm(N, _ptr(inp, np.complex64, n),
                                          _ptr(out, np.complex64, n), bool(inv),
                                          bool(interleaving))


def fftfpgaf_c2c_3d_ddr_svm_batch(N, inp, out, inv=False, how_many=2) -> FpgaT:
    n = _need(N, 3, how_many & 0xFFFFFFFF)
    return load().fftfpgaf_c2c_3d_ddr_svm_batch(N, _ptr(inp, np.complex64, n),
                                                _ptr(out, np.complex64, n), bool(inv),
                                                how_many & 0xFFFFFFFF)


# ---- device-pointer transforms (torch tensors on the GPU) -------------------------

def _stream_handle(t) -> ctypes.c_void_p:
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def c2c_dev(x, rank: int, inverse: bool = False, bitrev_1d: bool = False, out=None,
            split: bool = False, plane_fusion: bool = True, cube_split: bool = True):
    """Transform the last `rank` dims of a contiguous complex64/complex128 GPU
    tensor; leading dims are the batch.  Returns `out` (allocated if None).
    split=True selects the split-precision FP16 matrix-core radix-16 stages;
    plane_fusion=False runs separate line passes instead of the fused ones
    (the plane pass over the two fastest axes, the one-pass 4096-point column
    transform, the y-split pairs of 128^3, 256^3 and 512^3 blocks);
    cube_split=False runs such blocks without the y-split pair (the plane pass
    and a z pass for 128^3 and 256^3, three line passes for 512^3)."""
    import torch
    if not x.is_cuda:
        raise ValueError("c2c_dev needs a GPU tensor")
    if x.dtype not in (torch.complex64, torch.complex128):
        raise TypeError("complex64 or complex128 expected")
    if not x.is_contiguous():
        raise ValueError("contiguous tensor expected")
    if rank < 1 or rank > 3 or x.dim() < rank:
        raise ValueError("bad rank")
    dims = [int(d) for d in x.shape[x.dim() - rank:]]
    batch = 1
    for d in x.shape[: x.dim() - rank]:
        batch *= int(d)
    if out is None:
        out = torch.empty_like(x)
    elif out.shape != x.shape or out.dtype != x.dtype or not out.is_contiguous():
        raise ValueError("out must match x")
    if batch == 0:
        return out
    cd = (ctypes.c_uint * rank)(*dims)
    fn = load().fftamd_c2c_dev if x.dtype == torch.complex64 else load().fftamd_z2z_dev
    flags = ((FFTAMD_BITREV_1D if bitrev_1d else 0) | (FFTAMD_SPLIT_MFMA if split else 0)
             | (0 if plane_fusion else FFTAMD_NO_PLANE_FUSION) | (0 if cube_split else FFTAMD_NO_CUBE_SPLIT))
    rc = fn(rank, cd, batch, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            1 if inverse else 0, flags, _stream_handle(x))
    if rc != 0:
        raise RuntimeError(f"fftamd transform failed: {last_error()}")
    return out

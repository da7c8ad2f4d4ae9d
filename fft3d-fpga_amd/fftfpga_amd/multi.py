"""Batches of 3D transforms of uneven power-of-two sizes as one plan
(fftamd_multi_* in include/fftfpga/fftfpga_hip.h), over torch tensors.

    plan = MultiPlan(shapes, d_in, d_out, offsets_in, offsets_out, inverse=False)
    plan.tune(cache_path)          # optional: pick and cache the schedule
    plan()                         # enqueue on the current stream
    g = Graph([fwd, inv])          # hipGraph of a forward/inverse pair
    g()

Schedules (plan.schedule()["mode"]): 0 one persistent launch with grid
barriers between axes, 1 one persistent launch per axis, 2 per shape group a
plane-kernel launch (x and y) and a line-pass launch (z); mode 2 exists when
every transform has three axes of at least 16 points and x-y planes of at most
128 x 128.  3 one persistent dataflow launch of the mode-2 tiles (FP32 batches
of 32^3, 64^3 and 128^3 cubes, 128^3 out of place): a transform's z tiles wait
on a counter of its own plane tiles, no grid barrier, one work queue per XCD.
An untuned plan runs mode 2 where it exists, else mode 1; neither waits on a
grid barrier.  Only check() reports a timed-out wait (mode 0 barrier, mode 3
dependency).

Plans and graphs made before fpga_final (or a new fpga_initialize) are stale:
calling them raises RuntimeError; create them again.
"""
from __future__ import annotations

import ctypes
import os

import torch

from ._lib import XForm3D, last_error, load

DEFAULT_TUNE_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "fftamd", "tune.txt")


def _stream(dev=None):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def packed_offsets(shapes):
    """Element offsets of transforms packed back to back in one buffer."""
    offs, o = [], 0
    for s in shapes:
        offs.append(o)
        o += int(s[0]) * int(s[1]) * int(s[2])
    return offs, o


class MultiPlan:
    def __init__(self, shapes, d_in, d_out, in_offsets=None, out_offsets=None, inverse=False,
                 split=False):
        if d_in.dtype != torch.complex64 or d_out.dtype != torch.complex64:
            raise TypeError("complex64 buffers expected")
        if not (d_in.is_cuda and d_out.is_cuda and d_in.is_contiguous() and d_out.is_contiguous()):
            raise ValueError("contiguous GPU buffers expected")
        offs, total = packed_offsets(shapes)
        in_offsets = offs if in_offsets is None else list(in_offsets)
        out_offsets = offs if out_offsets is None else list(out_offsets)
        for s, oi, oo in zip(shapes, in_offsets, out_offsets):
            n = int(s[0]) * int(s[1]) * int(s[2])
            if oi + n > d_in.numel() or oo + n > d_out.numel():
                raise ValueError("a transform lies outside its buffer")
        arr = (XForm3D * len(shapes))()
        for i, (s, oi, oo) in enumerate(zip(shapes, in_offsets, out_offsets)):
            arr[i].n[0], arr[i].n[1], arr[i].n[2] = (int(v) for v in s)
            arr[i].reserved = 0
            arr[i].in_offset = oi
            arr[i].out_offset = oo
        self._lib = load()
        self._h = ctypes.c_void_p()
        rc = self._lib.fftamd_multi_create(ctypes.byref(self._h), len(shapes), ctypes.cast(arr, ctypes.c_void_p),
                                           ctypes.c_void_p(d_in.data_ptr()),
                                           ctypes.c_void_p(d_out.data_ptr()), 1 if inverse else 0,
                                           0x2 if split else 0)
        if rc != 0:
            raise RuntimeError(f"fftamd_multi_create failed: {last_error()}")
        self._keep = (d_in, d_out)  # the plan holds raw pointers into these tensors
        self.device = d_in.device

    def __call__(self):
        if self._lib.fftamd_multi_exec(self._h, _stream(self.device)) != 0:
            raise RuntimeError(f"fftamd_multi_exec failed: {last_error()}")

    def check(self):
        if self._lib.fftamd_multi_check(self._h, _stream(self.device)) != 0:
            raise RuntimeError(f"multi-transform launch reported: {last_error()}")

    def tune(self, cache_path: str | None = DEFAULT_TUNE_CACHE, reps: int = 3):
        if cache_path:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        us = ctypes.c_double(0)
        rc = self._lib.fftamd_multi_tune(self._h, _stream(self.device),
                                         cache_path.encode() if cache_path else None, reps,
                                         ctypes.byref(us))
        if rc < 0:
            raise RuntimeError(f"fftamd_multi_tune failed: {last_error()}")
        return {"from_cache": rc == 1, "best_us": us.value, **self.schedule()}

    def schedule(self):
        m, g, c, it, ph = ctypes.c_int(), ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        self._lib.fftamd_multi_schedule(self._h, ctypes.byref(m), ctypes.byref(g), ctypes.byref(c),
                                        ctypes.byref(it), ctypes.byref(ph))
        return {"mode": m.value, "grid": g.value, "chunk": c.value, "items": it.value,
                "phases": ph.value}

    def set_schedule(self, mode: int, grid: int, chunk: int):
        """mode 0, 1, 2 or 3 (see the module doc); grid (resident workgroups, 0
        = all) and chunk (tiles per queue ticket; 0 = static round-robin, no
        tickets, modes 0 and 1 only) apply to the persistent modes."""
        if self._lib.fftamd_multi_set_schedule(self._h, mode, grid, chunk) != 0:
            raise ValueError(last_error())

    def close(self):
        if self._h:
            self._lib.fftamd_multi_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """HIP graph of several multi-transform plans run in order on a side stream."""

    def __init__(self, plans, stream: torch.cuda.Stream | None = None):
        self._lib = load()
        self.stream = stream or torch.cuda.Stream()
        hs = (ctypes.c_void_p * len(plans))(*[p._h.value for p in plans])
        self._g = ctypes.c_void_p()
        torch.cuda.current_stream().synchronize()
        rc = self._lib.fftamd_graph_capture(ctypes.byref(self._g), hs, len(plans),
                                            ctypes.c_void_p(self.stream.cuda_stream))
        if rc != 0:
            raise RuntimeError(f"graph capture failed: {last_error()}")
        self._plans = plans

    def __call__(self):
        """Replay on the graph's stream, ordered after the current stream's work."""
        cur = torch.cuda.current_stream()
        self.stream.wait_stream(cur)
        if self._lib.fftamd_graph_launch(self._g, ctypes.c_void_p(self.stream.cuda_stream)) != 0:
            raise RuntimeError(f"graph launch failed: {last_error()}")
        cur.wait_stream(self.stream)

    def close(self):
        if self._g:
            self._lib.fftamd_graph_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_spin_limit(polls: int) -> None:
    load().fftamd_multi_set_spin_limit(polls)

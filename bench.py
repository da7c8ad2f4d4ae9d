"""Benchmark of the MI355X C2C transform path (one JSON line on stdout).

Default workload: 256^3 complex64 forward then unnormalised inverse 3D
transform (one step = both), inputs resident in HBM, through the library's
device-pointer entry point on the current torch stream.

value  = whole-job GFLOP/s at 5 N log2 N per transform (10^9 convention; the
         reference's perf_measures divides by 2^30, examples/helper.cpp:237).
roofline = the dominant kernel: algorithmic bytes of one launch (16 B per point:
         one read and one write of the complex64 array it transforms) over its
         mean launch time from HIP events recorded inside the timed region, as a
         fraction of MI355X's 8 TB/s.  transform_roofline gives the same
         fraction for the whole transform (16 B per point per transform).
cpu_baseline = the reference's CPU path, single-precision FFTW (libfftw3f, or
         MKL's FFTW3 API), on one forward transform of the same seeded input,
         threaded over this host's affinity cores with FFTW_ESTIMATE and
         FFTW_MEASURE plans, plus one thread; cpu_oracle_fp64 times the FP64
         oracle (a scalar C restatement) on the same transform.
accuracy = relative L2 error of this run's forward output (first transform)
         against that oracle output, and of the round trip (y / N - x).
Further objects on the same line (the default workload's run only): plane_pass_path (the same
steps with the 256 x 256 plane pass and a z pass instead of the y-split
pair), unfused_passes (the same
steps with 256 x 256 planes as separate x and y passes), torch_fft_rocfft
(torch.fft on the same data), batched (448 uneven transforms: persistent /
grouped-plane launch against a per-transform loop), split_512 (split-precision
512^3), autograd_16x128 (differentiable forward/backward), host_api_pcie (the
drop-in call on host buffers: copy-in, transform, copy-out times),
other_workloads (64^3, 4096^2 and 2^22-point 1D forward transforms).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "fft3d-fpga_amd")
for _p in (ROOT, PKG_DIR):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK = 8.0e12
METRIC = "3D C2C GFLOP/s (5N·log2N) + HBM GB/s at 256³ fp32; batched vs plain, 1 GPU"

WORKLOADS = {
    # name: (dims, batch, seed, round_trip); 1D output is bit-reversed per
    # transform, as fftfpgaf_c2c_1d returns it
    "3d_256_fwd_inv": ([256, 256, 256], 1, 1, True),
    "3d_64_fwd": ([64, 64, 64], 1, 0, False),
    "2d_4096_fwd": ([4096, 4096], 1, 1, False),
    "1d_4m_fwd": ([1 << 22], 1, 1, False),
    "3d_512_fwd": ([512, 512, 512], 1, 3, False),
    "3d_128_b16_fwd_inv": ([128, 128, 128], 16, 4, True),
}


def flops_per_transform(dims):
    n = math.prod(dims)
    return 5.0 * n * math.log2(n)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--cpu-fftw-child", action="store_true",
                    help="only time the FFTW-API CPU baseline of the workload and print it (run by the "
                         "bench itself in a child process that never loads torch)")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="3d_256_fwd_inv", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the torch.fft comparison")
    ap.add_argument("--no-kernel-events", action="store_true",
                    help="time the steps without per-kernel HIP events (diagnostic)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 10 plain device copies of one workload-sized array (counter calibration)")
    ap.add_argument("--no-plane-fusion", action="store_true",
                    help="run the x and y axes as separate line passes (default: one fused plane pass)")
    ap.add_argument("--no-cube-split", action="store_true",
                    help="256^3: the 256 x 256 plane pass and a z pass instead of the y-split pair")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the uneven-size batched vs plain per-transform comparison")
    ap.add_argument("--no-split", action="store_true",
                    help="skip the split-precision 512^3 (spiked field) measurement")
    ap.add_argument("--no-autograd", action="store_true",
                    help="skip the differentiable 16 x 128^3 forward/backward measurement")
    ap.add_argument("--no-cpu-fftw", action="store_true",
                    help="skip the FFTW-API CPU baseline (libfftw3f or MKL's fftwf wrappers)")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the other single-transform configurations (64^3, 4096^2, 2^22 1D)")
    ap.add_argument("--no-host-api", action="store_true",
                    help="skip the drop-in host-buffer call (PCIe-inclusive timing)")
    return ap.parse_args()


def cpu_oracle(dims, seed, gpu_forward=None, bitrev=False):
    """The FP64 oracle (a scalar C restatement, one thread) on one forward
    transform of the same seeded input: its time, and -- with gpu_forward (the
    GPU's forward output of the first transform, complex64) -- that output's
    relative L2 error against the oracle's (bit-reversed order for the 1D
    drop-in)."""
    from oracle import oracle
    x = oracle.uniform(math.prod(dims), seed)
    oracle.fftnd(x, dims)  # warm-up (page-in)
    ts = []
    t_end = time.perf_counter() + 20.0
    ref = None
    while len(ts) < 3 and (not ts or time.perf_counter() < t_end):
        t0 = time.perf_counter()
        ref = oracle.fftnd(x, dims)
        ts.append(time.perf_counter() - t0)
    t = sorted(ts)[len(ts) // 2]
    out = {"value": round(flops_per_transform(dims) / t / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
           "kind": "port",
           "sample": f"{len(ts)} forward {'x'.join(map(str, dims))} transforms, FP64 oracle, "
                     f"median {t * 1e3:.1f} ms"}
    err = None
    if gpu_forward is not None:
        if bitrev:
            ref = oracle.c2c_1d_expected(x, dims[0], 1)
        err = oracle.rel_l2(gpu_forward, ref)
    return out, err


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        return ""


def cpu_fftwf(dims, seed, reps=5):
    """The reference's CPU path: single-precision FFTW (examples/helper.cpp:119-141
    verifies with fftwf_plan_many_dft / FFTW_ESTIMATE), here out of place on this
    host's cores.  libfftw3f(_threads) if present, else MKL's FFTW3-compatible
    API (libmkl_rt).  Threaded on every core of the affinity mask with plans
    made under FFTW_ESTIMATE (the reference's flag) and FFTW_MEASURE, plus one
    thread with FFTW_ESTIMATE; plans are made outside the timed region, one
    untimed execution, then the median of `reps` executions.  `value` is the
    fastest of the three, `cores` its thread count."""
    import numpy as np
    from fftfpga_amd import datagen
    lib = label = None
    for name, lab in (("libfftw3f_threads.so.3", "fftwf"), ("libfftw3f.so.3", "fftwf"),
                      ("libfftw3f.so", "fftwf"), ("libmkl_rt.so", "fftwf API (MKL)"),
                      ("/opt/conda/lib/libmkl_rt.so", "fftwf API (MKL)")):
        try:
            cand = ctypes.CDLL(name)
        except OSError:
            continue
        if hasattr(cand, "fftwf_plan_dft") and hasattr(cand, "fftwf_execute"):
            lib, label = cand, lab
            break
    if lib is None:
        return {"available": False, "reason": "no libfftw3f or libmkl_rt on this host"}
    ncores = len(os.sched_getaffinity(0))
    vp, ci = ctypes.c_void_p, ctypes.c_int
    plan_dft = lib.fftwf_plan_dft
    plan_dft.restype = vp
    plan_dft.argtypes = [ci, ctypes.POINTER(ci), vp, vp, ci, ctypes.c_uint]
    lib.fftwf_execute.argtypes = [vp]
    lib.fftwf_destroy_plan.argtypes = [vp]
    threaded = False
    if hasattr(lib, "fftwf_init_threads") and hasattr(lib, "fftwf_plan_with_nthreads"):
        lib.fftwf_plan_with_nthreads.argtypes = [ci]
        threaded = lib.fftwf_init_threads() != 0
    if hasattr(lib, "fftwf_set_timelimit"):
        lib.fftwf_set_timelimit.argtypes = [ctypes.c_double]
        lib.fftwf_set_timelimit(20.0)  # bounds FFTW_MEASURE planning
    n = math.prod(dims)
    data = datagen.uniform(n, seed)
    x = np.empty_like(data)
    y = np.empty_like(data)
    nd = (ci * len(dims))(*dims)
    FFTW_FORWARD, FFTW_MEASURE, FFTW_ESTIMATE = -1, 0, 1 << 6
    flops = flops_per_transform(dims)

    def run(nthreads, flag, flag_name):
        if threaded:
            lib.fftwf_plan_with_nthreads(nthreads)
        t0 = time.perf_counter()
        plan = plan_dft(len(dims), nd, x.ctypes.data, y.ctypes.data, FFTW_FORWARD, flag)
        t_plan = time.perf_counter() - t0
        if not plan:
            return {"available": False, "reason": f"{label}: plan creation failed ({flag_name})"}
        x[:] = data  # FFTW_MEASURE overwrites the arrays while planning
        lib.fftwf_execute(plan)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            lib.fftwf_execute(plan)
            ts.append(time.perf_counter() - t0)
        lib.fftwf_destroy_plan(plan)
        t = sorted(ts)[len(ts) // 2]
        return {"value": round(flops / t / 1e9, 3), "unit": "GFLOP/s", "ms": round(t * 1e3, 3),
                "threads": nthreads if threaded else 1, "plan": flag_name, "plan_s": round(t_plan, 3)}

    est = run(ncores, FFTW_ESTIMATE, "FFTW_ESTIMATE")
    meas = run(ncores, FFTW_MEASURE, "FFTW_MEASURE")
    one = run(1, FFTW_ESTIMATE, "FFTW_ESTIMATE")
    # value: the fastest of the three runs (MKL's threading can lose to one
    # thread on some hosts), with the thread count that reached it
    best = max((r for r in (est, meas, one) if "value" in r), key=lambda r: r["value"], default=None)
    if best is None:
        return {"available": False, "reason": est.get("reason", "fftwf runs failed")}
    ref = np.fft.fftn(data.reshape(dims).astype(np.complex128)).reshape(-1) if n <= (1 << 24) else None
    return {"value": best["value"], "unit": "GFLOP/s", "cores": best["threads"], "kind": "fftwf",
            "lib": label, "plan": best["plan"], "threads_affinity": ncores, "cpu": _cpu_model(),
            "estimate": est, "measure": meas, "single_thread": one,
            "rel_l2_vs_numpy_fp64": (float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
                                     if ref is not None else None),
            "sample": f"forward {'x'.join(map(str, dims))} complex64, out of place, seed {seed}; "
                      f"median of {reps} executions per plan after one untimed execution"}


def split_512(steps):
    """Split-precision mode at 512^3 on U[-1,1) plus +-1e-3 spikes on 0.1% of
    the entries (splitmix64 seed 3): forward transform time of FP32 and of the
    split path (radix-16 stages on FP16 matrix cores), each through the y-split
    pair and through the three line passes, and the split result's relative L2
    distance from the FP32 path.  (Accuracy against the FP64 oracle is checked
    in the GPU tests.)"""
    import torch

    import fftfpga_amd as fl
    from fftfpga_amd import datagen
    N = 512
    x = torch.from_numpy(datagen.uniform_with_spikes(N ** 3, 3)).cuda().view(N, N, N)
    yp = torch.empty_like(x)
    ys = torch.empty_like(x)

    def run(split, out, cube_split=True):
        for _ in range(2):
            fl.c2c_dev(x, 3, out=out, split=split, cube_split=cube_split)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fl.c2c_dev(x, 3, out=out, split=split, cube_split=cube_split)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms_l = run(False, ys, cube_split=False)
    ms_sl = run(True, ys, cube_split=False)
    ms_p = run(False, yp)
    ms_s = run(True, ys)
    rel = (torch.linalg.vector_norm(ys - yp) / torch.linalg.vector_norm(yp)).item()
    alg = 16.0 * N ** 3
    fl_ = flops_per_transform([N, N, N])
    del x, yp, ys
    torch.cuda.empty_cache()
    return {"workload": "512^3 C2C forward, U[-1,1) + +-1e-3 spikes on 0.1% of entries, seed 3",
            "ms_fp32": round(ms_p, 4), "ms_split": round(ms_s, 4), "split_over_fp32": round(ms_s / ms_p, 3),
            "ms_fp32_line_passes": round(ms_l, 4), "ms_split_line_passes": round(ms_sl, 4),
            "split_over_fp32_line_passes": round(ms_s / ms_l, 3),
            "gflops_fp32": round(fl_ / (ms_p * 1e-3) / 1e9, 1), "gflops_split": round(fl_ / (ms_s * 1e-3) / 1e9, 1),
            "roofline_frac_fp32": round(alg / (ms_p * 1e-3) / HBM_PEAK, 4),
            "roofline_frac_split": round(alg / (ms_s * 1e-3) / HBM_PEAK, 4),
            "rel_l2_split_vs_fp32_path": rel}


def autograd_16x128(steps):
    """Differentiable transform y = W . FFT3(x) over a batch of 16 128^3
    complex64 arrays (x ~ N(0,1), splitmix64 seed 4; W = |Re U[-1,1)| seed 5,
    real 128^3): forward and hand-written adjoint backward (one weighted
    inverse transform each), timed separately."""
    import numpy as np
    import torch

    from fftfpga_amd import datagen
    from fftfpga_amd.autograd import fft3
    B, N = 16, 128
    x = torch.from_numpy(datagen.gauss(B * N ** 3, 4).reshape(B, N, N, N)).cuda().requires_grad_(True)
    w = torch.from_numpy(np.abs(datagen.uniform(N ** 3, 5).real).reshape(N, N, N).astype(np.float32)).cuda()
    g = torch.from_numpy(datagen.gauss(B * N ** 3, 6).reshape(B, N, N, N)).cuda()
    for _ in range(2):
        y = fft3(x, w)
        y.backward(g)
    torch.cuda.synchronize()
    # steps back to back with events between them: the queue stays busy, so an
    # event fires when the previous GPU work ends (not when Python gets to the
    # next launch); one untimed step first puts work ahead of the first event
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    x.grad = None
    y = fft3(x, w)
    y.backward(g)
    for e0, e1, e2 in evs:
        x.grad = None
        e0.record()
        y = fft3(x, w)
        e1.record()
        y.backward(g)
        e2.record()
    torch.cuda.synchronize()
    ms_f = sum(e0.elapsed_time(e1) for e0, e1, _ in evs) / steps
    ms_b = sum(e1.elapsed_time(e2) for _, e1, e2 in evs) / steps
    alg_f, alg_b = 16.0 * B * N ** 3, 16.0 * B * N ** 3 + 4.0 * N ** 3
    del x, w, g, y
    torch.cuda.empty_cache()
    return {"workload": "16 x 128^3, y = W . FFT3(x), x ~ N(0,1) seed 4, W = |Re U[-1,1)| seed 5",
            "ms_forward": round(ms_f, 4), "ms_backward": round(ms_b, 4),
            "hbm_gbs_forward": round(alg_f / (ms_f * 1e-3) / 1e9, 1),
            "hbm_gbs_backward": round(alg_b / (ms_b * 1e-3) / 1e9, 1),
            "roofline_frac_forward": round(alg_f / (ms_f * 1e-3) / HBM_PEAK, 4),
            "roofline_frac_backward": round(alg_b / (ms_b * 1e-3) / HBM_PEAK, 4),
            "timing": "events on the current stream around the forward call and around backward(), "
                      "steps issued back to back"}


def other_workloads(steps, warmup):
    """The other single-transform configurations, forward, HBM resident, through
    the device entry point (64^3 seed 0; 4096^2 and 2^22-point 1D with
    bit-reversed output, seed 1): wall-clock ms per transform over `steps`
    back-to-back calls, and each kernel's mean launch time from HIP events in a
    second pass over the same steps."""
    import torch

    import fftfpga_amd as fl
    from fftfpga_amd import datagen
    lib = fl.load()
    res = {}
    for name in ("3d_64_fwd", "2d_4096_fwd", "1d_4m_fwd"):
        dims, batch, seed, _ = WORKLOADS[name]
        rank = len(dims)
        npts = math.prod(dims) * batch
        x = torch.from_numpy(datagen.uniform(npts, seed)).cuda().reshape(batch, *dims)
        X = torch.empty_like(x)
        bitrev = rank == 1

        def step():
            fl.c2c_dev(x, rank, out=X, bitrev_1d=bitrev)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        if lib.fftamd_profile_begin(steps * 4 + 16):
            raise RuntimeError(fl.last_error())
        for _ in range(steps):
            step()
        buf = ctypes.create_string_buffer(1 << 16)
        if lib.fftamd_profile_end(buf, len(buf)) < 0:
            raise RuntimeError(fl.last_error())
        kern = json.loads(buf.value.decode())
        gbs = 16.0 * npts / (ms * 1e-3)
        res[name] = {"dims": dims, "ms_per_transform": round(ms, 5),
                     "gflops": round(flops_per_transform(dims) * batch / (ms * 1e-3) / 1e9, 1),
                     "hbm_gbs_algorithmic": round(gbs / 1e9, 1),
                     "roofline_frac": round(gbs / HBM_PEAK, 4),
                     "kernels_us": {k: round(1e3 * v["total_ms"] / v["count"], 3) for k, v in kern.items()},
                     "output_order": "bit-reversed per transform" if bitrev else "natural"}
        del x, X
    torch.cuda.empty_cache()
    return res


def host_api_pcie(N=256, seed=1, reps=3):
    """The drop-in fftfpgaf_c2c_3d_ddr on pageable host buffers (numpy), as a
    caller of the reference uses it: its fpga_t timings (HIP events around the
    copy in, the transform and the copy out), median of `reps` calls."""
    import numpy as np

    import fftfpga_amd as fl
    from fftfpga_amd import datagen
    if fl.fpga_initialize("AMD HIP", fl.LIB_PATH, False) != 0:
        return {"available": False, "reason": fl.last_error()}
    x = datagen.uniform(N ** 3, seed)
    out = np.empty_like(x)
    fl.fftfpgaf_c2c_3d_ddr(N, x, out, False)  # plan, staging buffers
    ts = []
    for _ in range(reps):
        t = fl.fftfpgaf_c2c_3d_ddr(N, x, out, False)
        if not t.valid:
            return {"available": False, "reason": fl.last_error()}
        ts.append((t.pcie_write_t, t.exec_t, t.pcie_read_t))
    w, e, r = (sorted(v)[len(v) // 2] for v in zip(*ts))
    tot = w + e + r
    return {"call": f"fftfpgaf_c2c_3d_ddr N={N} forward, pageable numpy buffers",
            "pcie_write_ms": round(w, 3), "exec_ms": round(e, 4), "pcie_read_ms": round(r, 3),
            "gflops_pcie_inclusive": round(flops_per_transform([N, N, N]) / (tot * 1e-3) / 1e9, 1),
            "gflops_exec_only": round(flops_per_transform([N, N, N]) / (e * 1e-3) / 1e9, 1)}


BATCH_COUNTS = {32: 256, 64: 160, 128: 32}


def batched_vs_plain(steps, warmup):
    """448 transforms of uneven sizes (256 x 32^3, 160 x 64^3, 32 x 128^3), shuffled
    with splitmix64 seed 2, N(0,1) data: one persistent launch (autotuned schedule,
    replayed from a HIP graph) against a plain loop of per-transform calls."""
    import numpy as np
    import torch

    import fftfpga_amd as fl
    from fftfpga_amd import datagen
    from fftfpga_amd.multi import Graph, MultiPlan, packed_offsets

    sizes = datagen.batch_order(BATCH_COUNTS, 2)
    shapes = [(n, n, n) for n in sizes]
    offs, total = packed_offsets(shapes)
    host = np.empty(total, np.complex64)
    for i, (sh, o) in enumerate(zip(shapes, offs)):
        n = sh[0] ** 3
        host[o:o + n] = datagen.gauss(n, datagen.member_seed(2, i))
    d_in = torch.from_numpy(host).cuda()
    d_out = torch.empty_like(d_in)
    d_ref = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    tune = plan.tune(reps=3)
    graph = Graph([plan])
    views = [(d_in[o:o + sh[0] ** 3].view(*sh), d_ref[o:o + sh[0] ** 3].view(*sh))
             for sh, o in zip(shapes, offs)]

    def plain():
        for a, b in views:
            fl.c2c_dev(a, 3, out=b)

    def run(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms_graph = run(graph)
    ms_eager = run(plan)
    ms_plain = run(plain)
    plan.check()
    # the whole batch as one persistent dataflow launch (mode 3), whatever the tuner chose
    single = None
    if tune["mode"] != 3:
        sch = plan.schedule()
        best = None
        for per_cu, chunk in ((2, 4), (2, 8)):
            ncu = torch.cuda.get_device_properties(0).multi_processor_count
            try:
                plan.set_schedule(3, per_cu * ncu, chunk)
            except ValueError:
                break
            ms = run(plan)
            plan.check()
            if best is None or ms < best[0]:
                best = (ms, per_cu, chunk)
        plan.set_schedule(sch["mode"], sch["grid"], sch["chunk"])
        if best:
            single = {"ms_eager": round(best[0], 4), "workgroups_per_cu": best[1], "chunk": best[2],
                      "over_tuned_graph": round(best[0] / ms_graph, 3)}
    else:
        single = {"ms_eager": round(ms_eager, 4), "over_tuned_graph": round(ms_eager / ms_graph, 3)}
    # the same batch with split-precision (FP16 matrix-core) radix-16 stages
    d_split = torch.empty_like(d_in)
    splan = MultiPlan(shapes, d_in, d_split, split=True)
    stune = splan.tune(reps=3)
    sgraph = Graph([splan])
    ms_split = run(sgraph)
    splan.check()
    split_rel = (torch.linalg.vector_norm(d_split - d_out) / torch.linalg.vector_norm(d_out)).item()
    flops = sum(flops_per_transform(sh) for sh in shapes)
    alg = 16.0 * total
    # the per-transform path with the same kernels as the tuned schedule
    # (modes 2 and 3: plane kernels, and the y-split pair for the 128^3 members;
    # modes 0, 1: line passes) must give the same bits
    for a, b in views:
        fl.c2c_dev(a, 3, out=b, plane_fusion=(tune["mode"] in (2, 3)))
    torch.cuda.synchronize()
    same = bool(torch.equal(d_out, d_ref))
    return {
        "workload": "448 uneven 3D C2C forward: 256x32^3 + 160x64^3 + 32x128^3, shuffled seed 2, N(0,1)",
        "transforms": len(shapes),
        "ms_batched_graph": round(ms_graph, 4),
        "ms_batched_eager": round(ms_eager, 4),
        "ms_plain_loop": round(ms_plain, 4),
        "speedup_vs_plain": round(ms_plain / ms_graph, 2),
        "gflops_batched": round(flops / (ms_graph * 1e-3) / 1e9, 1),
        "gflops_plain": round(flops / (ms_plain * 1e-3) / 1e9, 1),
        "hbm_gbs_algorithmic_batched": round(alg / (ms_graph * 1e-3) / 1e9, 1),
        "roofline_frac_batched": round(alg / (ms_graph * 1e-3) / HBM_PEAK, 4),
        "schedule": {k: tune[k] for k in ("mode", "grid", "chunk", "items", "phases", "from_cache")},
        "outputs_identical_to_per_transform_path": same,
        "single_persistent_launch": single,
        "split_precision": {"ms_batched_graph": round(ms_split, 4),
                            "speedup_vs_plain": round(ms_plain / ms_split, 2),
                            "rel_l2_vs_fp32_path": split_rel,
                            "schedule": {k: stune[k] for k in ("mode", "grid", "chunk")}},
    }


def cpu_fftwf_isolated(workload, dims, seed):
    """cpu_fftwf in a child process of this script that never imports torch:
    the FFTW API's threads (MKL's OpenMP runtime, when the API comes from MKL)
    then do not share the process with torch's own OpenMP pool.  Falls back to
    measuring in this process if the child fails."""
    import subprocess
    try:
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--workload", workload, "--cpu-fftw-child"],
                           capture_output=True, text=True, timeout=600)
        r = json.loads(p.stdout.strip().splitlines()[-1])
        r["process"] = "child process without torch"
        return r
    except Exception as e:  # noqa: BLE001 -- any failure of the child falls back
        r = cpu_fftwf(dims, seed)
        r["process"] = f"bench process (child failed: {type(e).__name__})"
        return r


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not a bench line)"}))
        sys.exit(2)
    if args.cpu_fftw_child:
        dims, _, seed, _ = WORKLOADS[args.workload]
        print(json.dumps(cpu_fftwf(dims, seed)), flush=True)
        return
    import torch

    import fftfpga_amd as fl
    from fftfpga_amd import datagen

    dims, batch, seed, round_trip = WORKLOADS[args.workload]
    rank = len(dims)
    npts = math.prod(dims) * batch
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        a = torch.ones(npts, dtype=torch.complex64, device=dev)
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": 8 * npts, "bytes_written": 8 * npts,
                          "launches": 10}))
        return
    x = torch.from_numpy(datagen.uniform(npts, seed)).to(dev).reshape(batch, *dims)
    X = torch.empty_like(x)
    y = torch.empty_like(x)
    lib = fl.load()

    bitrev = rank == 1
    fusion = {"on": not args.no_plane_fusion, "cube": not args.no_cube_split}
    fusable = rank >= 2  # plane passes / one-pass 4096-point columns where the shape has them
    cubic = rank == 3 and dims[0] == dims[1] == dims[2] and dims[0] in (128, 256, 512)  # y-split pair

    def step():
        fl.c2c_dev(x, rank, inverse=False, out=X, bitrev_1d=bitrev, plane_fusion=fusion["on"],
                   cube_split=fusion["cube"])
        if round_trip:
            fl.c2c_dev(X, rank, inverse=True, out=y, bitrev_1d=bitrev, plane_fusion=fusion["on"],
                       cube_split=fusion["cube"])

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    per_step = 2 if round_trip else 1
    max_passes = 4

    def timed(events: bool):
        if events and lib.fftamd_profile_begin(args.steps * per_step * max_passes + 16):
            raise RuntimeError(fl.last_error())
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        tb = time.perf_counter()
        kern = None
        if events:
            buf = ctypes.create_string_buffer(1 << 16)
            if lib.fftamd_profile_end(buf, len(buf)) < 0:
                raise RuntimeError(fl.last_error())
            kern = json.loads(buf.value.decode())
        return (tb - ta) * 1e3 / args.steps, kern

    # timed region 1: the K steps as a user runs them (no instrumentation) -> value
    ms, _ = timed(False)
    if args.no_kernel_events:
        print(json.dumps({"workload": args.workload, "ms_per_step": ms}))
        return
    # timed region 2: the same K steps with HIP events around every kernel launch,
    # recorded on the launch stream -> per-kernel durations for the roofline
    ms_ev, kernels = timed(True)

    transforms_per_step = per_step * batch
    flops = flops_per_transform(dims) * transforms_per_step
    gflops = flops / (ms * 1e-3) / 1e9
    alg_bytes_step = 16.0 * npts * per_step
    gbs = alg_bytes_step / (ms * 1e-3) / 1e9

    # dominant kernel
    dom = max(kernels.items(), key=lambda kv: kv[1]["total_ms"])
    dom_name, dom_rec = dom
    dom_avg_ms = dom_rec["total_ms"] / dom_rec["count"]
    dom_alg = 16.0 * npts
    achieved = dom_alg / (dom_avg_ms * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(dom_name)
    result = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": f"synthetic: splitmix64 seed {seed}, re/im i.i.d. U[-1,1), complex64",
        "config": {"workload": args.workload, "dims": dims, "batch": batch,
                   "round_trip": round_trip, "placement": "out-of-place, HBM resident",
                   "plane_fusion": fusable and fusion["on"],
                   "cube_split": cubic and fusion["on"] and fusion["cube"],
                   "inverse": "unnormalised",
                   "output_order": "bit-reversed per transform" if bitrev else "natural"},
        "hbm_gbs_algorithmic": round(gbs, 1),
        "transform_roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK / 1e9,
                               "unit": "GB/s", "frac": round(gbs * 1e9 / HBM_PEAK, 4)},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "traffic_source": (f"profiles/traffic_{args.workload}.json: rocprofv3 FETCH_SIZE + "
                                        "WRITE_SIZE of this kernel per launch, scaled by the counters of a "
                                        "plain device copy, from a separate profiled run of this workload "
                                        "(not measured in this run)") if traffic is not None else None,
                     "avg_launch_us": round(dom_avg_ms * 1e3, 3),
                     "timing": "HIP events around each launch, on its stream, in a second timed "
                               "region of the same K steps",
                     "ms_per_step_with_events": round(ms_ev, 5)},
        "kernels": {k: {"count": v["count"], "avg_us": round(1e3 * v["total_ms"] / v["count"], 3)}
                    for k, v in kernels.items()},
    }
    # accuracy of this run's outputs: the first transform's forward output (checked
    # against the FP64 oracle with the CPU baseline below) and the round trip
    fwd_first = X.reshape(batch, -1)[0].cpu().numpy()
    accuracy = {"bar": 2e-6}
    if round_trip:
        n_tot = math.prod(dims)
        accuracy["rel_l2_round_trip"] = float(torch.linalg.vector_norm(y / n_tot - x) /
                                              torch.linalg.vector_norm(x))
    if cubic and fusion["on"] and fusion["cube"]:
        # the same K steps through the 256 x 256 plane pass and a z pass
        fusion["cube"] = False
        for _ in range(args.warmup):
            step()
        ms_pl, _ = timed(False)
        fusion["cube"] = True
        result["plane_pass_path"] = {"ms_per_step": round(ms_pl, 5),
                                     "cube_split_speedup": round(ms_pl / ms, 3)}
    if fusable and fusion["on"]:
        # the same K steps with the x and y axes as two separate passes
        fusion["on"] = False
        for _ in range(args.warmup):
            step()
        ms_unf, _ = timed(False)
        fusion["on"] = True
        result["unfused_passes"] = {"ms_per_step": round(ms_unf, 5),
                                    "fusion_speedup": round(ms_unf / ms, 3)}
    if not args.no_compare:
        # the vendor library on the same device and data, for reference only
        for _ in range(3):
            torch.fft.fftn(x, dim=tuple(range(1, rank + 1)))
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            Xr = torch.fft.fftn(x, dim=tuple(range(1, rank + 1)))
            if round_trip:
                torch.fft.ifftn(Xr, dim=tuple(range(1, rank + 1)), norm="forward")
        torch.cuda.synchronize()
        tb = time.perf_counter()
        ms_t = (tb - ta) * 1e3 / args.steps
        result["torch_fft_rocfft"] = {"ms_per_step": round(ms_t, 5),
                                      "gflops": round(flops / (ms_t * 1e-3) / 1e9, 2),
                                      "speedup_vs_torch": round(ms_t / ms, 3)}
    # the other configurations ride along with the default workload only
    default = args.workload == "3d_256_fwd_inv"
    if default and not args.no_batched:
        result["batched"] = batched_vs_plain(max(5, min(args.steps, 20)), 2)
    del x, X, y
    torch.cuda.empty_cache()
    if default and not args.no_split:
        result["split_512"] = split_512(max(3, min(args.steps, 10)))
    if default and not args.no_autograd:
        result["autograd_16x128"] = autograd_16x128(max(3, min(args.steps, 10)))
    if default and not args.no_other:
        result["other_workloads"] = other_workloads(max(5, min(args.steps, 50)), args.warmup)
    if default and not args.no_host_api:
        result["host_api_pcie"] = host_api_pcie()
    if not args.no_cpu_baseline:
        result["cpu_oracle_fp64"], err = cpu_oracle(dims, seed, fwd_first, bitrev)
        accuracy["rel_l2_forward_vs_fp64_oracle"] = err
        if not args.no_cpu_fftw:
            # the reference's own CPU path (FFTW single precision) on this host
            result["cpu_baseline"] = cpu_fftwf_isolated(args.workload, dims, seed)
    result["accuracy"] = accuracy
    result["cpu_reference_emulation"] = {
        "available": False,
        "reason": "the reference's CPU run (fft --num=64 --dim=3 --emulate) needs the Intel FPGA "
                  "SDK for OpenCL emulator, which is not installed"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

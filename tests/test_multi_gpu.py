"""Transforms of uneven sizes in one persistent launch (fftamd_multi_*).

Bars: every transform of a mixed batch within rel-L2 2e-6 of the FP64
oracle; each schedule runs the same tile code as one per-transform path
(modes 0 and 1: the line passes; modes 2 and 3: the plane kernel and the z
line pass, or the y-split pair for 128^3), so on the 448-transform workload
its output must equal that path's output bit for bit; HIP-graph replays must
equal the eager result on every replay; a grid-barrier or dependency wait
forced past its bound from the host must be reported, not hang.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 2e-6


@pytest.fixture(autouse=True)
def announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()


def _pack(shapes, seed):
    import torch
    from fftfpga_amd import datagen
    from fftfpga_amd.multi import packed_offsets
    offs, total = packed_offsets(shapes)
    host = np.empty(total, np.complex64)
    for i, (s, o) in enumerate(zip(shapes, offs)):
        n = int(np.prod(s))
        host[o:o + n] = datagen.gauss(n, datagen.member_seed(seed, i))
    return host, torch.from_numpy(host).cuda(), offs


MIXED = [(16, 16, 16), (64, 64, 64), (8, 8, 8), (128, 128, 128), (32, 32, 32), (4, 8, 16),
         (64, 64, 64), (2, 2, 2), (32, 32, 32), (256, 16, 8), (1, 64, 32), (16, 16, 16)]


@pytest.mark.parametrize("mode,per_cu,chunk", [(0, None, 4), (1, None, 1), (0, 1, 16), (1, 2, 3), (0, None, 0), (1, 2, 0)])
@pytest.mark.parametrize("inverse", [False, True])
def test_mixed_batch_vs_oracle(gpu_lib, orc, mode, per_cu, chunk, inverse):
    import torch
    from fftfpga_amd.multi import MultiPlan
    host, d_in, offs = _pack(MIXED, 5)
    d_out = torch.zeros_like(d_in)
    plan = MultiPlan(MIXED, d_in, d_out, inverse=inverse)
    sch = plan.schedule()
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    grid = sch["grid"] if per_cu is None else min(sch["grid"], per_cu * ncu)
    plan.set_schedule(mode, grid, chunk)
    plan()
    plan.check()
    out = d_out.cpu().numpy()
    for s, o in zip(MIXED, offs):
        n = int(np.prod(s))
        ref = orc.fftnd(host[o:o + n], list(s), inverse=inverse)
        err = orc.rel_l2(out[o:o + n], ref)
        assert err <= TOL, (s, err)
    plan.close()


def test_mixed_batch_in_place(gpu_lib, orc):
    from fftfpga_amd.multi import MultiPlan
    host, d, offs = _pack(MIXED, 6)
    plan = MultiPlan(MIXED, d, d)
    plan()
    plan.check()
    out = d.cpu().numpy()
    for s, o in zip(MIXED, offs):
        n = int(np.prod(s))
        assert orc.rel_l2(out[o:o + n], orc.fftnd(host[o:o + n], list(s))) <= TOL


PLANES = [(32, 32, 32)] * 5 + [(64, 64, 64)] * 3 + [(16, 128, 128)] * 2 + [(128, 64, 32), (64, 32, 128)] \
    + [(128, 128, 128)] * 2  # out of place: the y-split pair through the transform table


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("in_place", [False, True])
def test_plane_schedule_vs_oracle(gpu_lib, orc, inverse, in_place):
    """Mode 2 (plane kernels per shape group, then table-driven z passes) on
    transforms placed in reverse order with gaps between them."""
    import torch
    from fftfpga_amd import datagen
    from fftfpga_amd.multi import MultiPlan
    sizes = [int(np.prod(s)) for s in PLANES]
    offs, o = [], 0
    for n in reversed(sizes):  # last transform first, 13-element gaps
        offs.append(o)
        o += n + 13
    offs = offs[::-1]
    host = np.zeros(o, np.complex64)
    for i, (n, off) in enumerate(zip(sizes, offs)):
        host[off:off + n] = datagen.gauss(n, datagen.member_seed(8, i))
    d_in = torch.from_numpy(host).cuda()
    d_out = d_in if in_place else torch.zeros_like(d_in)
    plan = MultiPlan(PLANES, d_in, d_out, in_offsets=offs, out_offsets=offs, inverse=inverse)
    sch = plan.schedule()
    plan.set_schedule(2, sch["grid"], sch["chunk"])
    plan()
    plan.check()
    out = d_out.cpu().numpy()
    for s, n, off in zip(PLANES, sizes, offs):
        ref = orc.fftnd(host[off:off + n], list(s), inverse=inverse)
        assert orc.rel_l2(out[off:off + n], ref) <= TOL, s
    if not in_place:  # gaps untouched
        gaps = np.ones(o, bool)
        for n, off in zip(sizes, offs):
            gaps[off:off + n] = False
        assert not out[gaps].any()
    plan.close()


def test_plane_schedule_unavailable_is_rejected(gpu_lib):
    import torch
    from fftfpga_amd.multi import MultiPlan
    _, d_in, _ = _pack(MIXED, 3)
    plan = MultiPlan(MIXED, d_in, torch.empty_like(d_in))  # 8^3 planes: no plane kernel
    sch = plan.schedule()
    with pytest.raises(ValueError):
        plan.set_schedule(2, sch["grid"], sch["chunk"])
    plan.close()


def _config3_shapes():
    from fftfpga_amd import datagen
    sizes = datagen.batch_order({32: 256, 64: 160, 128: 32}, 2)
    return [(n, n, n) for n in sizes]


@pytest.mark.parametrize("mode,chunk", [(0, None), (1, None), (2, None), (0, 0), (3, 1), (3, 4)])
def test_config3_448_uneven_matches_per_transform_bitwise(gpu_lib, orc, mode, chunk):
    import torch
    from fftfpga_amd import c2c_dev
    from fftfpga_amd.multi import MultiPlan
    shapes = _config3_shapes()
    assert len(shapes) == 448
    host, d_in, offs = _pack(shapes, 2)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    sch = plan.schedule()
    plan.set_schedule(mode, 0 if mode == 3 else sch["grid"], sch["chunk"] if chunk is None else chunk)
    plan()
    plan.check()
    ref = torch.empty_like(d_in)
    for s, o in zip(shapes, offs):
        n = int(np.prod(s))
        # modes 2 and 3 (out of place) run the same tiles as a single call: plane kernels
        # for 32^3 and 64^3, the y-split pair for 128^3
        c2c_dev(d_in[o:o + n].view(*s), 3, out=ref[o:o + n].view(*s), plane_fusion=(mode in (2, 3)))
    torch.cuda.synchronize()
    assert torch.equal(d_out, ref)
    out = d_out.cpu().numpy()
    for i in (0, 1, 2, 100, 447):
        s, o = shapes[i], offs[i]
        n = int(np.prod(s))
        assert orc.rel_l2(out[o:o + n], orc.fftnd(host[o:o + n], list(s))) <= TOL
    plan.close()


def test_graph_forward_inverse_pair_replays(gpu_lib):
    import torch
    from fftfpga_amd.multi import Graph, MultiPlan
    shapes = [(32, 32, 32)] * 8 + [(64, 64, 64)] * 4 + [(128, 128, 128)] * 2
    host, d_in, offs = _pack(shapes, 9)
    mid = torch.empty_like(d_in)
    back = torch.empty_like(d_in)
    fwd = MultiPlan(shapes, d_in, mid)
    inv = MultiPlan(shapes, mid, back, inverse=True)
    fwd()
    inv()
    fwd.check()
    eager = back.clone()
    g = Graph([fwd, inv])
    for r in range(3):
        back.zero_()
        mid.mul_(0)
        d_in.add_(0)  # eager work between replays
        torch.cuda.synchronize()
        g()
        torch.cuda.synchronize()
        assert torch.equal(back, eager), f"replay {r} differs from the eager result"
    # round trip returns N^3 x
    out = back.cpu().numpy()
    for s, o in zip(shapes, offs):
        n = int(np.prod(s))
        err = np.linalg.norm(out[o:o + n] / n - host[o:o + n]) / np.linalg.norm(host[o:o + n])
        assert err <= TOL
    g.close()


def test_graph_plane_schedule_pair_replays(gpu_lib):
    """Mode 2 with three shape groups runs each group's launches on its own
    forked stream: captured into a graph they are parallel branches.  A
    forward/inverse pair replayed from the graph equals the eager pair on
    every replay, with eager work between the replays."""
    import torch
    from fftfpga_amd.multi import Graph, MultiPlan
    shapes = [(32, 32, 32)] * 6 + [(64, 64, 64)] * 3 + [(128, 128, 128)] * 2
    host, d_in, offs = _pack(shapes, 10)
    mid = torch.empty_like(d_in)
    back = torch.empty_like(d_in)
    fwd = MultiPlan(shapes, d_in, mid)
    inv = MultiPlan(shapes, mid, back, inverse=True)
    for p in (fwd, inv):
        sch = p.schedule()
        p.set_schedule(2, sch["grid"], sch["chunk"])
    fwd()
    inv()
    inv.check()
    eager_mid, eager = mid.clone(), back.clone()
    g = Graph([fwd, inv])
    for r in range(3):
        back.zero_()
        mid.mul_(0)
        d_in.add_(0)  # eager work between replays
        torch.cuda.synchronize()
        g()
        torch.cuda.synchronize()
        assert torch.equal(mid, eager_mid), f"replay {r}: forward differs from the eager result"
        assert torch.equal(back, eager), f"replay {r}: inverse differs from the eager result"
    out = back.cpu().numpy()
    for s, o in zip(shapes, offs):
        n = int(np.prod(s))
        err = np.linalg.norm(out[o:o + n] / n - host[o:o + n]) / np.linalg.norm(host[o:o + n])
        assert err <= TOL
    g.close()
    fwd.close()
    inv.close()


def test_graph_capture_before_any_launch(gpu_lib):
    """A mode-2 pair captured before either plan ever ran: the capture must not
    allocate (every table a mode-2 launch reads is made when the plan is
    created).  The library is reset first so that no table is cached from
    earlier work.  Replays equal the eager pair run afterwards."""
    import torch
    from fftfpga_amd.multi import Graph, MultiPlan
    gpu_lib.fpga_final()  # drops every cached twiddle table
    assert gpu_lib.fpga_initialize("AMD HIP", gpu_lib.LIB_PATH, False) == 0
    shapes = [(32, 32, 32)] * 4 + [(64, 64, 64)] * 2 + [(128, 128, 128)] * 2
    host, d_in, offs = _pack(shapes, 12)
    mid = torch.empty_like(d_in)
    back = torch.empty_like(d_in)
    fwd = MultiPlan(shapes, d_in, mid)
    inv = MultiPlan(shapes, mid, back, inverse=True)
    for p in (fwd, inv):
        sch = p.schedule()
        p.set_schedule(2, sch["grid"], sch["chunk"])
    g = Graph([fwd, inv])  # nothing has run yet
    replays = []
    for _ in range(2):
        back.zero_()
        torch.cuda.synchronize()
        g()
        torch.cuda.synchronize()
        replays.append((mid.clone(), back.clone()))
    mid.zero_()
    back.zero_()
    fwd()
    inv()
    inv.check()
    for r, (m, b) in enumerate(replays):
        assert torch.equal(m, mid), f"replay {r}: forward differs from the eager result"
        assert torch.equal(b, back), f"replay {r}: inverse differs from the eager result"
    out = back.cpu().numpy()
    for s, o in zip(shapes, offs):
        n = int(np.prod(s))
        err = np.linalg.norm(out[o:o + n] / n - host[o:o + n]) / np.linalg.norm(host[o:o + n])
        assert err <= TOL
    g.close()
    fwd.close()
    inv.close()


def test_plane_schedule_single_group(gpu_lib, orc):
    """Mode 2 with one shape group (no stream fork) against the oracle."""
    import torch
    from fftfpga_amd.multi import MultiPlan
    shapes = [(64, 64, 64)] * 3
    host, d_in, offs = _pack(shapes, 11)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    sch = plan.schedule()
    plan.set_schedule(2, sch["grid"], sch["chunk"])
    plan()
    plan.check()
    out = d_out.cpu().numpy()
    for s, o in zip(shapes, offs):
        n = int(np.prod(s))
        assert orc.rel_l2(out[o:o + n], orc.fftnd(host[o:o + n], list(s))) <= TOL
    plan.close()


def test_barrier_bound_is_reported_not_hung(gpu_lib):
    import torch
    from fftfpga_amd.multi import MultiPlan, set_spin_limit
    shapes = [(64, 64, 64)] * 16
    _, d_in, _ = _pack(shapes, 3)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    plan.set_schedule(0, plan.schedule()["grid"], 1)
    set_spin_limit(1)  # the host forces the bounded wait to expire
    try:
        plan()
        with pytest.raises(RuntimeError, match="bound"):
            plan.check()
    finally:
        set_spin_limit(1 << 22)
    plan()
    plan.check()  # a normal launch afterwards is fine


def test_autotuner_picks_and_caches(gpu_lib, tmp_path):
    import torch
    from fftfpga_amd.multi import MultiPlan
    shapes = [(32, 32, 32)] * 32 + [(64, 64, 64)] * 8
    _, d_in, _ = _pack(shapes, 4)
    d_out = torch.empty_like(d_in)
    cache = str(tmp_path / "tune.txt")
    plan = MultiPlan(shapes, d_in, d_out)
    r1 = plan.tune(cache, reps=2)
    assert not r1["from_cache"] and r1["best_us"] > 0 and r1["mode"] in (0, 1, 2, 3)
    plan2 = MultiPlan(shapes, d_in, d_out)
    r2 = plan2.tune(cache, reps=2)
    assert r2["from_cache"] and (r2["mode"], r2["grid"], r2["chunk"]) == (r1["mode"], r1["grid"], r1["chunk"])
    plan2()
    plan2.check()


@pytest.mark.parametrize("inverse", [False, True])
def test_mixed_batch_split_precision(gpu_lib, orc, inverse):
    import torch
    from fftfpga_amd.multi import MultiPlan
    host, d_in, offs = _pack(MIXED, 8)
    d_out = torch.zeros_like(d_in)
    plan = MultiPlan(MIXED, d_in, d_out, inverse=inverse, split=True)
    plan()
    plan.check()
    out = d_out.cpu().numpy()
    for s, o in zip(MIXED, offs):
        n = int(np.prod(s))
        err = orc.rel_l2(out[o:o + n], orc.fftnd(host[o:o + n], list(s), inverse=inverse))
        assert err <= TOL, (s, err)


def test_plane_schedule_split_matches_per_transform(gpu_lib, orc):
    """Split precision in mode 2: the plane kernels and z passes with FP16
    matrix-core radix-16 stages, bit-identical to the per-transform split path
    and within the FP32 bar of the oracle."""
    import torch
    from fftfpga_amd import c2c_dev
    from fftfpga_amd.multi import MultiPlan
    host, d_in, offs = _pack(PLANES, 12)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(PLANES, d_in, d_out, split=True)
    sch = plan.schedule()
    plan.set_schedule(2, sch["grid"], sch["chunk"])
    plan()
    plan.check()
    ref = torch.empty_like(d_in)
    for s, o in zip(PLANES, offs):
        n = int(np.prod(s))
        c2c_dev(d_in[o:o + n].view(*s), 3, out=ref[o:o + n].view(*s), split=True)
    torch.cuda.synchronize()
    assert torch.equal(d_out, ref)
    out = d_out.cpu().numpy()
    for s, o in zip(PLANES, offs):
        n = int(np.prod(s))
        assert orc.rel_l2(out[o:o + n], orc.fftnd(host[o:o + n], list(s))) <= TOL, s
    plan.close()


def test_stale_plan_and_graph_refused_after_final(gpu_lib):
    """fpga_final releases the twiddle tables a plan and its graph point to:
    calling either afterwards is refused with an error instead of reading freed
    memory; a plan made after re-initialisation works."""
    import torch
    from fftfpga_amd.multi import Graph, MultiPlan
    shapes = [(32, 32, 32)] * 2 + [(64, 64, 64)]
    host, d_in, offs = _pack(shapes, 13)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    g = Graph([plan])
    plan()
    torch.cuda.synchronize()
    gpu_lib.fpga_final()
    try:
        with pytest.raises(RuntimeError, match="predates"):
            plan()
        with pytest.raises(RuntimeError, match="predates"):
            g()
        with pytest.raises(RuntimeError, match="predates"):
            plan.tune(None, reps=1)
    finally:
        assert gpu_lib.fpga_initialize("AMD HIP", gpu_lib.LIB_PATH, False) == 0
    with pytest.raises(RuntimeError, match="predates"):
        plan()  # still stale after a new initialisation
    fresh = MultiPlan(shapes, d_in, d_out)
    fresh()
    fresh.check()
    g.close()
    plan.close()
    fresh.close()


def test_untuned_plan_has_no_grid_barrier(gpu_lib):
    """A new plan runs mode 2 where it exists, else mode 1 (no grid barrier)."""
    import torch
    from fftfpga_amd.multi import MultiPlan
    shapes = [(32, 32, 32)] * 2 + [(64, 64, 64)]
    _, d_in, _ = _pack(shapes, 14)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    assert plan.schedule()["mode"] == 2
    plan.close()
    shapes = [(8, 8, 8)] * 3  # planes below the plane kernel's 1024 points: no mode 2
    _, d_in, _ = _pack(shapes, 15)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    assert plan.schedule()["mode"] == 1
    plan()
    plan.check()
    plan.close()


# ---- mode 3: one dataflow launch (a counter per transform, one queue per XCD) ----
CUBES = [(64, 64, 64), (32, 32, 32), (128, 128, 128), (32, 32, 32), (64, 64, 64), (128, 128, 128),
         (32, 32, 32), (32, 32, 32), (64, 64, 64), (32, 32, 32), (128, 128, 128)]


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("grid,chunk", [(0, 1), (0, 4), (3, 2)])
def test_flow_schedule_vs_oracle(gpu_lib, orc, inverse, grid, chunk):
    """Mode 3 against the oracle, with every resident workgroup and with three
    (fewer than the eight queues: the launch must still drain every queue)."""
    import torch
    from fftfpga_amd.multi import MultiPlan
    host, d_in, offs = _pack(CUBES, 11)
    d_out = torch.zeros_like(d_in)
    plan = MultiPlan(CUBES, d_in, d_out, inverse=inverse)
    plan.set_schedule(3, grid, chunk)
    plan()
    plan.check()
    out = d_out.cpu().numpy()
    for s, o in zip(CUBES, offs):
        n = int(np.prod(s))
        assert orc.rel_l2(out[o:o + n], orc.fftnd(host[o:o + n], list(s), inverse=inverse)) <= TOL, s
    plan.close()


def test_flow_schedule_in_place(gpu_lib, orc):
    """32^3 and 64^3 transforms in place (their first pass runs plane by plane)."""
    import torch
    from fftfpga_amd.multi import MultiPlan
    shapes = [s for s in CUBES if s[0] < 128]
    host, d, offs = _pack(shapes, 12)
    plan = MultiPlan(shapes, d, d)
    plan.set_schedule(3, 0, 2)
    plan()
    plan.check()
    out = d.cpu().numpy()
    for s, o in zip(shapes, offs):
        n = int(np.prod(s))
        assert orc.rel_l2(out[o:o + n], orc.fftnd(host[o:o + n], list(s))) <= TOL, s
    plan.close()


def test_flow_schedule_unavailable_is_rejected(gpu_lib):
    import torch
    from fftfpga_amd.multi import MultiPlan
    _, d_in, _ = _pack(PLANES, 3)  # not all cubes of 32, 64 or 128
    plan = MultiPlan(PLANES, d_in, torch.empty_like(d_in))
    with pytest.raises(ValueError):
        plan.set_schedule(3, 0, 1)
    plan.close()
    shapes = [(128, 128, 128)]
    _, d, _ = _pack(shapes, 3)
    plan = MultiPlan(shapes, d, d)  # 128^3 in place: the pair's pass A needs another buffer
    with pytest.raises(ValueError):
        plan.set_schedule(3, 0, 1)
    plan.close()
    _, d_in, _ = _pack(CUBES, 3)
    plan = MultiPlan(CUBES, d_in, torch.empty_like(d_in), split=True)  # FP32 tiles only
    with pytest.raises(ValueError):
        plan.set_schedule(3, 0, 1)
    plan.set_schedule(2, 0, 1)  # the grouped launches take split plans
    plan.close()


def test_flow_graph_pair_replays(gpu_lib):
    """A mode-3 forward/inverse pair captured in a HIP graph: every replay
    equals the eager pair, with eager work between the replays."""
    import torch
    from fftfpga_amd.multi import Graph, MultiPlan
    host, d_in, offs = _pack(CUBES, 13)
    d_mid = torch.zeros_like(d_in)
    d_back = torch.zeros_like(d_in)
    fwd = MultiPlan(CUBES, d_in, d_mid)
    inv = MultiPlan(CUBES, d_mid, d_back, inverse=True)
    for p in (fwd, inv):
        p.set_schedule(3, 0, 2)
    fwd()
    inv()
    torch.cuda.synchronize()
    eager_mid, eager_back = d_mid.clone(), d_back.clone()
    g = Graph([fwd, inv])
    for r in range(3):
        d_mid.zero_()
        d_back.mul_(0)
        d_in.add_(0)  # eager work between replays
        torch.cuda.synchronize()
        g()
        torch.cuda.synchronize()
        assert torch.equal(d_mid, eager_mid) and torch.equal(d_back, eager_back), f"replay {r}"
        fwd.check()
        inv.check()
    for s, o in zip(CUBES, offs):
        n = int(np.prod(s))
        back = d_back[o:o + n].cpu().numpy() / n
        assert np.linalg.norm(back - host[o:o + n]) <= 2e-6 * np.linalg.norm(host[o:o + n])
    g.close()
    fwd.close()
    inv.close()


def test_flow_dependency_bound_is_reported_not_hung(gpu_lib):
    """One 128^3 transform: its pass-B tiles are taken while its pass-A tiles
    run, so with the poll bound forced to one from the host they give up,
    report through check() and the launch still ends."""
    import torch
    from fftfpga_amd.multi import MultiPlan, set_spin_limit
    shapes = [(128, 128, 128)]
    _, d_in, _ = _pack(shapes, 3)
    d_out = torch.empty_like(d_in)
    plan = MultiPlan(shapes, d_in, d_out)
    plan.set_schedule(3, 0, 1)
    set_spin_limit(1)
    try:
        plan()
        with pytest.raises(RuntimeError, match="bound"):
            plan.check()
    finally:
        set_spin_limit(1 << 22)
    plan()
    plan.check()
    plan.close()

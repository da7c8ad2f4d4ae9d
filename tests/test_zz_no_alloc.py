"""The allocation-free entry point fftamd_c2c_dev_noalloc against fftamd_c2c_dev.

Both run the same kernels with the same twiddle values (the allocation-free
path's tables are written into the static scratch array from the same
long-double generator), so the outputs must be identical -- stricter than the
rel-L2 <= 2e-6 bar.  Each case makes two calls in a row, the second on a
smaller input, since every call shares the device's scratch array (the
temporary buffer of in-place shapes keeps the previous call's contents).
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()


def _call(lib, fn, dims, batch, x, out, inverse, flags=0):
    import torch
    d = (ctypes.c_uint * len(dims))(*dims)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(lib, fn)(len(dims), d, batch, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                          1 if inverse else 0, flags, s)
    return rc


# (dims, batch, in place): in-place 128^3 and 64^3 blocks run their first pass
# through the scratch's temporary buffer; 256^3 out of place needs none
CASES = [
    ([[128, 128, 128], 1, True], [[64, 64, 64], 2, True]),
    ([[256, 256, 256], 1, False], [[32, 32, 32], 3, False]),
    ([[1 << 20], 2, False], [[1 << 16], 1, False]),
    ([[2048, 2048], 1, True], [[512, 512], 2, True]),
]


@pytest.mark.parametrize("first,second", CASES)
@pytest.mark.parametrize("inverse", [False, True])
def test_noalloc_matches_allocating_entry(fl, orc, first, second, inverse):
    import torch
    lib = fl.load()
    for k, (dims, batch, inplace) in enumerate((first, second)):
        n = int(np.prod(dims)) * batch
        host = orc.uniform(n, 40 + 7 * k + (3 if inverse else 0))
        x = torch.from_numpy(host).cuda()
        ref = torch.empty_like(x)
        assert _call(lib, "fftamd_c2c_dev", dims, batch, x, ref, inverse) == 0, fl.last_error()
        if inplace:
            y = x.clone()
            rc = _call(lib, "fftamd_c2c_dev_noalloc", dims, batch, y, y, inverse)
        else:
            y = torch.empty_like(x)
            rc = _call(lib, "fftamd_c2c_dev_noalloc", dims, batch, x, y, inverse)
        assert rc == 0, fl.last_error()
        torch.cuda.synchronize()
        got, want = y.cpu().numpy(), ref.cpu().numpy()
        err = orc.rel_l2(got, want.astype(np.complex128))
        print(f"dims={dims} batch={batch} in_place={inplace} inverse={inverse}: rel_l2={err:.3e} "
              f"identical={np.array_equal(got, want)}")
        assert err <= 2e-6
        assert np.array_equal(got, want)


@pytest.mark.parametrize("inverse", [False, True])
def test_noalloc_split_matches_allocating_entry(fl, orc, inverse):
    """Split precision (FFTAMD_SPLIT_MFMA) through the allocation-free entry
    point: the y-split pairs' split stages keep their operand table in LDS, so
    nothing is allocated; 256^3 then 2 x 128^3, identical to fftamd_c2c_dev."""
    import torch
    lib = fl.load()
    for k, (dims, batch) in enumerate(([[256, 256, 256], 1], [[128, 128, 128], 2])):
        n = int(np.prod(dims)) * batch
        x = torch.from_numpy(orc.uniform(n, 90 + k + (5 if inverse else 0))).cuda()
        ref = torch.empty_like(x)
        y = torch.empty_like(x)
        assert _call(lib, "fftamd_c2c_dev", dims, batch, x, ref, inverse, flags=2) == 0, fl.last_error()
        assert _call(lib, "fftamd_c2c_dev_noalloc", dims, batch, x, y, inverse, flags=2) == 0, fl.last_error()
        torch.cuda.synchronize()
        assert torch.equal(y, ref), f"call {k}: {dims} x {batch}"


def test_noalloc_rejects_scratch_overflow(fl):
    """An in-place shape whose temporary buffer exceeds the scratch is refused
    with a clear error (256^3 in place needs 128 MiB; the scratch has 64 MiB)."""
    import torch
    lib = fl.load()
    x = torch.zeros(256 ** 3, dtype=torch.complex64, device="cuda")
    rc = _call(lib, "fftamd_c2c_dev_noalloc", [256, 256, 256], 1, x, x, False)
    assert rc == -1 and "temporary buffer" in fl.last_error()


def test_noalloc_calls_allocate_no_device_memory(fl, orc):
    """After the library is set up, no call of the entry point changes the
    device memory in use (first call of a new shape included)."""
    import torch
    lib = fl.load()
    dims, batch = [64, 64, 64], 4
    n = int(np.prod(dims)) * batch
    x = torch.from_numpy(orc.uniform(n, 77)).cuda()
    y = x.clone()
    ref = torch.empty_like(x)
    # library set up and the shape's kernels loaded (loading a code object
    # takes device memory of its own) through the allocating entry point
    assert _call(lib, "fftamd_c2c_dev", dims, batch, x, ref, False) == 0, fl.last_error()
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        assert _call(lib, "fftamd_c2c_dev_noalloc", dims, batch, y, y, False) == 0, fl.last_error()
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


@pytest.mark.parametrize("dims,batch,inplace", [([32, 64, 128], 3, False), ([128, 128, 128], 2, True)])
def test_noalloc_first_call_under_graph_capture(fl, orc, dims, batch, inplace):
    """The first call for a shape may itself be under stream capture (a torch
    CUDA graph, default capture mode): the call allocates nothing, so the
    capture succeeds, and every replay -- with eager work between replays --
    equals fftamd_c2c_dev's result.  The shapes are planned by no other test
    on this entry point (2 x 128^3 in place runs through the scratch's
    temporary buffer)."""
    import torch
    lib = fl.load()
    n = int(np.prod(dims)) * batch
    host = orc.uniform(n, 91 + batch)
    x = torch.from_numpy(host).cuda()
    ref = torch.empty_like(x)
    assert _call(lib, "fftamd_c2c_dev", dims, batch, x, ref, False) == 0, fl.last_error()
    y = x.clone() if inplace else torch.empty_like(x)
    src = y if inplace else x
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        rc = _call(lib, "fftamd_c2c_dev_noalloc", dims, batch, src, y, False)
    assert rc == 0, fl.last_error()
    for r in range(3):
        if inplace:
            y.copy_(x)  # the graph transforms y in place: restore its input
        else:
            y.zero_()
        x.add_(0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, ref), f"replay {r} differs from fftamd_c2c_dev"
    del g

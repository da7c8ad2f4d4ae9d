"""Split-precision mode (FFTAMD_SPLIT_MFMA): radix-16 stages on FP16 matrix
cores with hi+lo operands and FP32 accumulation.  Bar: the same rel-L2 <= 2e-6
against the FP64 oracle as the plain FP32 path, on every path that uses
radix-16 stages, and on the 512^3 field with a sparse 1e-3 spike component."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 2e-6


@pytest.fixture(autouse=True)
def announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()


@pytest.mark.parametrize("shape,rank,cube_split", [((16, 16, 16), 3, True), ((2, 64, 64, 64), 3, True),
                                                   ((256, 256, 256), 3, True), ((256, 256, 256), 3, False),
                                                   ((2, 128, 128, 128), 3, True), ((128, 128, 128), 3, False),
                                                   ((32, 128, 64), 3, True), ((4096, 4096), 2, True),
                                                   ((1 << 22,), 1, True), ((8, 4096), 1, True),
                                                   ((3, 512, 1024), 2, True), ((4, 128, 128), 2, True),
                                                   ((2, 16, 128, 128), 3, True)])
@pytest.mark.parametrize("inv", [False, True])
def test_split_vs_oracle(gpu_lib, orc, shape, rank, cube_split, inv):
    """128^3 and 256^3 blocks run the y-split pair with split radix-16 stages
    (cube_split=False: the plane pass and z pass, as other 3D shapes)."""
    import torch
    from fftfpga_amd import c2c_dev
    n = int(np.prod(shape))
    x = orc.uniform(n, 600 + n % 991).reshape(shape)
    dims = list(shape[len(shape) - rank:])
    batch = n // int(np.prod(dims))
    xt = torch.from_numpy(x).cuda()
    y = c2c_dev(xt, rank, inverse=inv, split=True, cube_split=cube_split)
    yp = c2c_dev(xt, rank, inverse=inv)
    torch.cuda.synchronize()
    ref = orc.fftnd(x, dims, batch=batch, inverse=inv)
    err = orc.rel_l2(y.cpu().numpy(), ref)
    errp = orc.rel_l2(yp.cpu().numpy(), ref)
    print(f"split {shape} rank={rank} cube_split={cube_split} inv={inv} rel_l2={err:.3e} (plain fp32 {errp:.3e})")
    assert err <= TOL


def test_split_bit_reversed_1d(gpu_lib, orc):
    import torch
    from fftfpga_amd import c2c_dev
    N, B = 4096, 3
    x = orc.uniform(N * B, 17).reshape(B, N)
    y = c2c_dev(torch.from_numpy(x).cuda(), 1, bitrev_1d=True, split=True)
    torch.cuda.synchronize()
    assert orc.rel_l2(y.cpu().numpy(), orc.c2c_1d_expected(x, N, B)) <= TOL


def test_split_dynamic_range_512cubed_with_spikes(gpu_lib, orc):
    """512^3, U[-1,1) plus +-1e-3 spikes on 0.1% of the entries (seed 3)."""
    import torch
    from fftfpga_amd import c2c_dev, datagen
    N = 512
    x = datagen.uniform_with_spikes(N ** 3, 3)
    xt = torch.from_numpy(x).cuda().view(N, N, N)
    y = c2c_dev(xt, 3, split=True).cpu().numpy().reshape(-1)
    yp = c2c_dev(xt, 3).cpu().numpy().reshape(-1)
    ref = orc.fftnd(x, [N, N, N]).reshape(-1)
    err = orc.rel_l2(y, ref)
    errp = orc.rel_l2(yp, ref)
    # the spike component alone: transform of (x - uniform part) compared through linearity
    print(f"512^3 spikes: split rel_l2={err:.3e}, plain fp32 rel_l2={errp:.3e}")
    assert err <= TOL and errp <= TOL


def test_split_constant_input_does_not_overflow(gpu_lib):
    """A constant 256^3 input drives the DC term to 2^24, far past FP16's range;
    the per-vector power-of-two scaling keeps every operand representable."""
    import torch
    from fftfpga_amd import c2c_dev
    N = 256
    x = torch.full((N, N, N), 1.0 + 0.5j, dtype=torch.complex64, device="cuda")
    y = c2c_dev(x, 3, split=True)
    torch.cuda.synchronize()
    assert torch.isfinite(y.real).all() and torch.isfinite(y.imag).all()
    dc = y[0, 0, 0].item()
    assert abs(dc - (1.0 + 0.5j) * N ** 3) <= 1e-6 * abs(dc)
    y[0, 0, 0] = 0
    assert torch.abs(y).max().item() <= 1e-6 * abs(dc)


def test_split_lines_of_very_different_magnitudes(gpu_lib, orc):
    """64-point lines put four lines in each group of 16 vectors of a split
    stage; with lines alternately 1e8 and 1 in magnitude, every line must keep
    the bar on its own (each vector is scaled by its own maximum, so the small
    lines' low parts do not fall into FP16 subnormals)."""
    import torch
    from fftfpga_amd import c2c_dev
    L, B = 64, 64
    x = orc.uniform(L * B, 77).reshape(B, L)
    x[0::2] *= 1e8
    y = c2c_dev(torch.from_numpy(x).cuda(), 1, split=True).cpu().numpy()
    ref = orc.fftnd(x.reshape(-1), [L], batch=B).reshape(B, L)
    errs = [orc.rel_l2(y[i], ref[i]) for i in range(B)]
    print(f"per-line rel_l2: large lines max {max(errs[0::2]):.3e}, small lines max {max(errs[1::2]):.3e}")
    assert max(errs) <= TOL


@pytest.mark.parametrize("shape", [(2, 128, 128, 128), (256, 256, 256), (512, 512, 512)])
def test_split_in_place_matches_out_of_place(gpu_lib, orc, shape):
    """The split y-split pairs in place (pass A into the plan's temporary
    buffer, pass B from it into the array) give the out-of-place result bit for
    bit, in both directions."""
    import torch
    from fftfpga_amd import c2c_dev
    n = int(np.prod(shape))
    x = torch.from_numpy(orc.uniform(n, 310 + n % 97)).cuda().view(shape)
    for inv in (False, True):
        ref = c2c_dev(x, 3, inverse=inv, split=True)
        y = x.clone()
        c2c_dev(y, 3, inverse=inv, split=True, out=y)
        torch.cuda.synchronize()
        assert torch.equal(y, ref), f"{shape} inv={inv}"
        del ref, y
    del x
    torch.cuda.empty_cache()

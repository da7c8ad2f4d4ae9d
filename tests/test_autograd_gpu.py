"""Differentiable transform: hand-written adjoint backward vs FP64 central
finite differences and vs torch.fft's autograd.

Loss (the differentiable workload): L(x) = sum_b sum |W . FFT3(x_b)|^2 with
real W.  L is quadratic in x, so the central difference
(L(x + h v) - L(x - h v)) / 2h equals the directional derivative
Re sum conj(grad) . v exactly up to rounding; L is evaluated in FP64.
Bars: directional derivatives within 1e-5 relative of the FP64 finite
differences (the gradient itself is computed in FP32); gradient within 2e-6
relative L2 of torch.fft's complex64 autograd gradient is not expected
(that gradient is itself FP32), so the bar there is 1e-5.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()


def _data(B, N, seed):
    import torch
    from fftfpga_amd import datagen
    x = torch.from_numpy(datagen.gauss(B * N ** 3, seed).reshape(B, N, N, N)).cuda()
    w = torch.from_numpy(np.abs(datagen.uniform(N ** 3, seed + 1).real).reshape(N, N, N)
                         .astype(np.float32)).cuda()
    return x, w


def _loss64(x, w):
    import torch
    X = torch.fft.fftn(x.to(torch.complex128), dim=(-3, -2, -1))
    return ((w.to(torch.float64) * X).abs() ** 2).sum()


@pytest.mark.parametrize("B,N,ndir", [(2, 16, 64), (16, 128, 64)])
def test_adjoint_vs_fp64_finite_differences(gpu_lib, B, N, ndir):
    import torch
    from fftfpga_amd.autograd import fft3
    x, w = _data(B, N, 4)
    xr = x.clone().requires_grad_(True)
    loss = (fft3(xr, w).abs() ** 2).sum()
    loss.backward()
    g = xr.grad.to(torch.complex128)
    gen = torch.Generator(device="cuda").manual_seed(1234)
    h = 1.0  # L is quadratic: the central difference has no truncation error
    fds, dds = [], []
    for _ in range(ndir):
        v = torch.randn(x.shape, dtype=torch.complex128, generator=gen, device="cuda")
        fds.append(((_loss64(x + h * v, w) - _loss64(x - h * v, w)) / (2 * h)).item())
        dds.append((torch.conj(g) * v).real.sum().item())
    fds, dds = np.array(fds), np.array(dds)
    # errors are measured against the RMS directional derivative: single random
    # directions can be nearly orthogonal to the gradient (derivative ~ 0)
    scale = np.sqrt(np.mean(fds ** 2))
    worst = np.max(np.abs(dds - fds)) / scale
    med = np.median(np.abs(dds - fds) / np.abs(fds))
    print(f"adjoint FD B={B} N={N} worst/rms={worst:.3e} median rel={med:.3e}")
    assert worst < 2e-6 and med < 2e-6


def test_gradient_matches_torch_fft_autograd(gpu_lib):
    import torch
    from fftfpga_amd.autograd import fft3
    x, w = _data(16, 128, 4)
    a = x.clone().requires_grad_(True)
    (fft3(a, w).abs() ** 2).sum().backward()
    b = x.clone().requires_grad_(True)
    ((w * torch.fft.fftn(b, dim=(-3, -2, -1))).abs() ** 2).sum().backward()
    rel = (torch.linalg.vector_norm(a.grad - b.grad) / torch.linalg.vector_norm(b.grad)).item()
    print(f"grad vs torch.fft rel={rel:.3e}")
    assert rel < 1e-5


def test_forward_value_and_weight_gradient(gpu_lib, orc):
    import torch
    from fftfpga_amd.autograd import fftn
    x, w = _data(2, 16, 7)
    ref = orc.fftnd(x.cpu().numpy(), [16, 16, 16], batch=2) * w.cpu().numpy().astype(np.float64)
    y = fftn(x, 3, w)
    assert orc.rel_l2(y.cpu().numpy(), ref) <= 2e-6
    # inverse direction and its adjoint
    yi = fftn(x, 3, w, inverse=True)
    refi = orc.fftnd(x.cpu().numpy(), [16, 16, 16], batch=2, inverse=True) * w.cpu().numpy()
    assert orc.rel_l2(yi.cpu().numpy(), refi) <= 2e-6
    # weight gradient against torch autograd on the reference formula
    wr = w.clone().requires_grad_(True)
    (fftn(x, 3, wr).abs() ** 2).sum().backward()
    wt = w.clone().requires_grad_(True)
    ((wt * torch.fft.fftn(x, dim=(-3, -2, -1))).abs() ** 2).sum().backward()
    rel = (torch.linalg.vector_norm(wr.grad - wt.grad) / torch.linalg.vector_norm(wt.grad)).item()
    assert rel < 1e-5


def test_unweighted_adjoint_is_unnormalised_inverse(gpu_lib, orc):
    import torch
    from fftfpga_amd.autograd import fftn
    x, _ = _data(1, 32, 9)
    xr = x.clone().requires_grad_(True)
    y = fftn(xr, 3)
    g = torch.from_numpy(orc.uniform(32 ** 3, 10).reshape(1, 32, 32, 32)).cuda()
    y.backward(g)
    ref = orc.fftnd(g.cpu().numpy(), [32, 32, 32], inverse=True)
    assert orc.rel_l2(xr.grad.cpu().numpy(), ref) <= 2e-6

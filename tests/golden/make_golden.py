"""Writes the committed fixtures in tests/golden/*.npz.

Expected outputs come from numpy's pocketfft in FP64 (np.fft), an FFT
implementation independent of both the oracle and the HIP library; the 1D
bit-reversed order is applied with a plain-Python bit reversal.  The
reference (pc2/fft3d-fpga) ships no golden vectors, so these fixtures are the
build's own; they pin the oracle in tests/test_oracle.py and the GPU path in
tests/test_parity_gpu.py.

Inputs come from the seeded splitmix64 generator of the package
(fftfpga_amd.datagen); small fixtures store the input itself, the sampled
ones store the seed, the first 64 input words and outputs at 256 indices.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "fft3d-fpga_amd"))
from fftfpga_amd import datagen  # noqa: E402


def bitrev(i, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def fftn(x, inverse):
    x = np.asarray(x, np.complex128)
    axes = tuple(range(1, x.ndim))
    return np.fft.ifftn(x, axes=axes) * np.prod([x.shape[a] for a in axes]) if inverse \
        else np.fft.fftn(x, axes=axes)


def nd(name, dims, batch, seed, inverse, gen="uniform"):
    n = int(np.prod(dims))
    x = getattr(datagen, gen)(n * batch, seed).reshape(batch, *dims)
    y = fftn(x, inverse)
    np.savez(os.path.join(HERE, name), kind="nd", dims=np.array(dims), batch=batch, seed=seed,
             inverse=inverse, x=x, y=y)


def one_d(name, n, batch, seed, inverse):
    x = datagen.uniform(n * batch, seed).reshape(batch, n)
    X = fftn(x, inverse)
    lg = n.bit_length() - 1
    y = np.empty_like(X)
    for b in range(n):
        y[:, b] = X[:, bitrev(b, lg)]
    np.savez(os.path.join(HERE, name), kind="1d_bitrev", n=n, batch=batch, seed=seed,
             inverse=inverse, x=x, y=y)


def sampled(name, dims, seed, inverse, nsamp=256):
    n = int(np.prod(dims))
    x = datagen.uniform(n, seed).reshape(dims)
    y = fftn(x[None], inverse)[0].reshape(-1)
    rng = np.random.default_rng(seed + 1000)
    idx = np.sort(rng.choice(n, nsamp, replace=False)).astype(np.int64)
    idx[0] = 0
    np.savez(os.path.join(HERE, name), kind="sampled", dims=np.array(dims), seed=seed,
             inverse=inverse, x_head=x.reshape(-1).view(np.uint32)[:64].copy(), idx=idx,
             y_at=y[idx], energy=np.vdot(y, y).real)


def sampled_1d_bitrev(name, n, seed, inverse, nsamp=256):
    """One n-point transform in the library's 1D output order (bit-reversed):
    y_at[i] = X[bitrev(idx[i])] at 256 sampled output positions idx."""
    x = datagen.uniform(n, seed)
    X = fftn(x[None], inverse)[0]
    lg = n.bit_length() - 1
    rng = np.random.default_rng(seed + 2000)
    idx = np.sort(rng.choice(n, nsamp, replace=False)).astype(np.int64)
    idx[0] = 0
    nat = np.array([bitrev(int(i), lg) for i in idx], np.int64)
    np.savez(os.path.join(HERE, name), kind="sampled_1d_bitrev", n=n, seed=seed, inverse=inverse,
             x_head=x.view(np.uint32)[:64].copy(), idx=idx, y_at=X[nat], energy=np.vdot(X, X).real)


def main():
    nd("nd_3d_8_fwd_seed0.npz", [8, 8, 8], 1, 0, False)
    nd("nd_3d_16_inv_seed1.npz", [16, 16, 16], 1, 1, True)
    nd("nd_3d_4x8x16_fwd.npz", [4, 8, 16], 1, 7, False)
    nd("nd_2d_16_fwd_gauss_batch2.npz", [16, 16], 2, 4, False, gen="gauss")
    nd("nd_2d_32_inv.npz", [32, 32], 1, 5, True)
    one_d("1d_bitrev_64_fwd_batch3.npz", 64, 3, 6, False)
    one_d("1d_bitrev_16_inv.npz", 16, 1, 8, True)
    sampled("sampled_3d_64_fwd_seed0.npz", [64, 64, 64], 0, False)
    sampled("sampled_3d_256_fwd_seed1.npz", [256, 256, 256], 1, False)
    sampled("sampled_2d_4096_fwd_seed1.npz", [4096, 4096], 1, False)
    sampled_1d_bitrev("sampled_1d_4m_fwd_seed1.npz", 1 << 22, 1, False)


if __name__ == "__main__":
    main()

"""Pins the CPU oracle before it is trusted as the parity checker.

The reference ships no golden vectors (its tests pin only argument/init
contracts; numerical parity was FFTW with SNR > 120 dB, examples/helper.cpp:
164-198).  The oracle is therefore pinned against (a) numpy's pocketfft in
FP64, an independent implementation; (b) a brute-force long-double DFT;
(c) closed forms and invariants; (d) the committed fixtures in tests/golden,
written by tests/golden/make_golden.py from numpy alone.  The last test
plants plausible mistakes (wrong sign, transposed axes, missing bit reversal,
normalised inverse) in a copy of the oracle's output and checks that the same
checks reject each one.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def np_fftn(x, inverse=False):
    x = np.asarray(x, dtype=np.complex128)
    return np.fft.ifftn(x) * x.size if inverse else np.fft.fftn(x)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


# ---- (a) against pocketfft --------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 1024, 4096, 1 << 15])
@pytest.mark.parametrize("inverse", [False, True])
def test_oracle_1d_vs_pocketfft(orc, n, inverse):
    x = orc.uniform(n * 3, 11 + n).reshape(3, n)
    got = orc.fftnd(x, [n], batch=3, inverse=inverse)
    ref = np.stack([np_fftn(r, inverse) for r in x])
    assert rel(got, ref) < 1e-14


@pytest.mark.parametrize("dims", [[2, 2], [16, 16], [64, 32], [8, 8, 8], [16, 16, 16], [32, 4, 8],
                                  [4, 64, 2]])
@pytest.mark.parametrize("inverse", [False, True])
def test_oracle_nd_vs_pocketfft(orc, dims, inverse):
    x = orc.gauss(2 * int(np.prod(dims)), 7).reshape(2, *dims)
    got = orc.fftnd(x, dims, batch=2, inverse=inverse)
    ref = np.stack([np_fftn(b, inverse) for b in x])
    assert rel(got, ref) < 1e-14


def test_oracle_64cubed_seed0_vs_pocketfft(orc):
    x = orc.uniform(64 ** 3, 0).reshape(64, 64, 64)
    assert rel(orc.fftnd(x, [64, 64, 64])[0], np.fft.fftn(x.astype(np.complex128))) < 1e-14


# ---- (b) brute force ------------------------------------------------------------------

@pytest.mark.parametrize("dims", [[1], [2], [8], [16], [64], [4, 4], [8, 2], [4, 4, 4], [8, 8, 8],
                                  [2, 4, 8]])
@pytest.mark.parametrize("inverse", [False, True])
def test_oracle_vs_naive_long_double(orc, dims, inverse):
    x = orc.uniform(int(np.prod(dims)), 3).astype(np.complex128).reshape(dims)
    got = orc.fftnd(x, dims, inverse=inverse)[0]
    ref = orc.dft_naive(x, dims, inverse=inverse)
    assert rel(got, ref) < 1e-14


# ---- (c) closed forms and invariants -------------------------------------------------

def tone(dims, k):
    grids = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    ph = sum(g * kk / d for g, kk, d in zip(grids, k, dims))
    return np.exp(2j * np.pi * ph)


@pytest.mark.parametrize("dims", [[16], [8, 16], [8, 8, 8], [4, 16, 8]])
def test_closed_forms(orc, dims):
    n = int(np.prod(dims))
    d = len(dims)
    imp = np.zeros(dims, np.complex128)
    imp[(0,) * d] = 1
    assert np.allclose(orc.fftnd(imp, dims)[0], 1.0, atol=1e-15)
    const = np.ones(dims, np.complex128)
    out = orc.fftnd(const, dims)[0]
    expect = np.zeros(dims)
    expect[(0,) * d] = n
    assert np.allclose(out, expect, atol=1e-12)
    # a complex exponential e^{+2 pi i k.n/N} lands on exactly index k (forward sign is -)
    k = tuple((3 * (a + 1)) % dd for a, dd in enumerate(dims))
    out = orc.fftnd(tone(dims, k), dims)[0]
    assert np.unravel_index(np.argmax(np.abs(out)), dims) == k
    expect = np.zeros(dims)
    expect[k] = n
    assert np.allclose(out, expect, atol=1e-9)
    # the backward transform of that tone lands on -k
    out = orc.fftnd(tone(dims, k), dims, inverse=True)[0]
    mk = tuple((-kk) % dd for kk, dd in zip(k, dims))
    assert np.unravel_index(np.argmax(np.abs(out)), dims) == mk
    # shifted impulse -> phase ramp
    s = tuple(1 + a for a in range(d))
    sh = np.zeros(dims, np.complex128)
    sh[s] = 1
    out = orc.fftnd(sh, dims)[0]
    assert np.allclose(out, np.conj(tone(dims, s)), atol=1e-12)


@pytest.mark.parametrize("dims", [[256], [32, 64], [16, 16, 16]])
def test_invariants(orc, dims):
    n = int(np.prod(dims))
    x = orc.gauss(n, 5).reshape(dims).astype(np.complex128)
    y = orc.gauss(n, 6).reshape(dims).astype(np.complex128)
    X = orc.fftnd(x, dims)[0]
    Y = orc.fftnd(y, dims)[0]
    # Parseval
    assert abs(np.vdot(X, X).real - n * np.vdot(x, x).real) < 1e-10 * n * np.vdot(x, x).real
    # round trip: backward(forward(x)) = n x
    assert rel(orc.fftnd(X, dims, inverse=True)[0], n * x) < 1e-14
    # linearity
    assert rel(orc.fftnd(2.5 * x - 1j * y, dims)[0], 2.5 * X - 1j * Y) < 1e-14
    # conjugate symmetry: F(conj x)[k] = conj(F(x)[-k])
    Xc = orc.fftnd(np.conj(x), dims)[0]
    # X[-k mod n] along every axis = roll(flip(X), 1)
    neg = np.conj(np.roll(np.flip(X), 1, axis=tuple(range(len(dims)))))
    assert rel(Xc, neg) < 1e-14
    # swap identity used by the engine: backward(x) = swap(forward(swap(x)))
    sw = lambda z: z.imag + 1j * z.real
    assert rel(orc.fftnd(x, dims, inverse=True)[0], sw(orc.fftnd(sw(x), dims)[0])) < 1e-14
    # separability: the 3D/2D transform equals 1D transforms along each axis in turn
    Z = x.copy()
    for ax in range(len(dims)):
        Z = np.apply_along_axis(lambda v: orc.fft1d_f64(v), ax, Z)
    assert rel(Z, X) < 1e-14


# ---- 1D output order of fftfpgaf_c2c_1d -----------------------------------------------

@pytest.mark.parametrize("n", [2, 8, 64, 512])
def test_1d_expected_is_bit_reversed(orc, n):
    lg = n.bit_length() - 1
    # helper.cpp:10-18 on known values
    assert [orc.bit_reversed(i, 3) for i in range(8)] == [0, 4, 2, 6, 1, 5, 3, 7]
    assert orc.bit_reversed(1, 6) == 32
    # x = delta at n0 => X[k] = exp(-2 pi i n0 k / n); engine order puts X[bitrev(b)] at b
    n0 = 3 % n
    x = np.zeros(n, np.complex64)
    x[n0] = 1
    out = orc.c2c_1d_expected(x, n, 1)[0]
    b = np.arange(n)
    brev = np.array([orc.bit_reversed(int(i), lg) for i in b])
    assert np.allclose(out, np.exp(-2j * np.pi * n0 * brev / n), atol=1e-12)
    # and un-permuting as examples/helper.cpp:143-161 recovers the natural-order DFT
    y = orc.uniform(2 * n, 9)
    out = orc.c2c_1d_expected(y, n, 2)
    nat = np.empty_like(out)
    nat[:, b] = out[:, brev]
    assert rel(nat, np.stack([np.fft.fft(r.astype(np.complex128)) for r in y.reshape(2, n)])) < 1e-14


# ---- generators and metrics -------------------------------------------------------------

def test_splitmix64_known_answer(orc):
    import ctypes
    s = ctypes.c_uint64(0)
    first = orc.lib().orc_splitmix64_next(ctypes.byref(s))
    assert first == 0xE220A8397B1DCDAF  # published first output for seed 0


def test_package_generator_matches_oracle(orc, fl):
    from fftfpga_amd import datagen
    assert int(datagen.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    for seed in (0, 1, 2, 12345):
        a = datagen.uniform(5000, seed)
        b = orc.uniform(5000, seed)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        g1 = datagen.gauss(5000, seed).view(np.float32)
        g2 = orc.gauss(5000, seed).view(np.float32)
        ulp = np.abs(g1.view(np.int32).astype(np.int64) - g2.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1
    u = datagen.uniform(100000, 1)
    assert u.real.min() >= -1 and u.real.max() < 1 and abs(u.real.mean()) < 0.01


def test_gauss_generator_is_standard_normal(orc):
    """orc_fill_gauss pinned by what N(0,1) fixes, independently of its Python
    twin: over 10^6 draws per component the mean and variance lie within 5
    standard errors of 0 and 1, the fourth moment of 3, re and im are
    uncorrelated, and a Kolmogorov-Smirnov test against the normal CDF does not
    reject (a wrong Box-Muller scale, a dropped sqrt(-2 ln u) or a cosine used
    for both components fails these)."""
    from scipy import stats
    n = 1_000_000
    g = orc.gauss(n, 17)
    for comp in (g.real.astype(np.float64), g.imag.astype(np.float64)):
        se = 1.0 / np.sqrt(n)
        assert abs(comp.mean()) < 5 * se
        assert abs(comp.var() - 1.0) < 5 * np.sqrt(2.0) * se
        assert abs(np.mean(comp ** 4) - 3.0) < 5 * np.sqrt(96.0) * se
        assert stats.kstest(comp[:200_000], "norm").pvalue > 1e-4
    assert abs(np.corrcoef(g.real, g.imag)[0, 1]) < 5.0 / np.sqrt(n)


def test_error_metrics(orc):
    ref = orc.gauss(4096, 2).astype(np.complex128)
    got = (ref * (1 + 1e-6)).astype(np.complex64)
    r = orc.rel_l2(got, ref)
    assert abs(r - 1e-6) < 2e-7
    assert abs(orc.snr_db(got, ref) - (-20 * np.log10(r))) < 1e-6


# ---- (d) committed fixtures ---------------------------------------------------------------

def _fixtures():
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("name", _fixtures())
def test_oracle_matches_golden(orc, name):
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    kind = str(z["kind"])
    inv = bool(z["inverse"])
    if kind == "nd":
        dims = [int(d) for d in z["dims"]]
        got = orc.fftnd(z["x"], dims, batch=int(z["batch"]), inverse=inv)
        assert rel(got.reshape(z["y"].shape), z["y"]) < 1e-13
    elif kind == "1d_bitrev":
        n, b = int(z["n"]), int(z["batch"])
        got = orc.c2c_1d_expected(z["x"], n, b, inverse=inv)
        assert rel(got.reshape(z["y"].shape), z["y"]) < 1e-13
    elif kind == "sampled":
        dims = [int(d) for d in z["dims"]]
        x = orc.uniform(int(np.prod(dims)), int(z["seed"]))
        assert np.array_equal(x.view(np.uint32)[: z["x_head"].size], z["x_head"])
        got = orc.fftnd(x, dims, inverse=inv).reshape(-1)
        assert rel(got[z["idx"]], z["y_at"]) < 1e-13
        assert abs(np.vdot(got, got).real - float(z["energy"])) < 1e-10 * float(z["energy"])
    elif kind == "sampled_1d_bitrev":
        n = int(z["n"])
        x = orc.uniform(n, int(z["seed"]))
        assert np.array_equal(x.view(np.uint32)[: z["x_head"].size], z["x_head"])
        got = orc.c2c_1d_expected(x, n, 1, inverse=inv).reshape(-1)
        assert rel(got[z["idx"]], z["y_at"]) < 1e-13
        assert abs(np.vdot(got, got).real - float(z["energy"])) < 1e-10 * float(z["energy"])
    else:
        pytest.fail(f"unknown fixture kind {kind}")


# ---- planted mistakes are caught ---------------------------------------------------------------

def test_planted_mistakes_are_rejected(orc):
    dims = [8, 4, 16]
    x = orc.uniform(int(np.prod(dims)), 21).reshape(dims)
    good = orc.fftnd(x, dims)[0]
    ref = np.fft.fftn(x.astype(np.complex128))
    assert rel(good, ref) < 1e-14
    wrong_sign = orc.fftnd(x, dims, inverse=True)[0]
    transposed = np.transpose(orc.fftnd(np.transpose(x, (2, 1, 0)), dims[::-1])[0], (2, 1, 0))
    swapped_axes = orc.fftnd(x.reshape(4, 8, 16), [4, 8, 16])[0].reshape(dims)
    normalised_inverse = orc.fftnd(x, dims, inverse=True)[0] / x.size
    dropped_term = good.copy()
    dropped_term.reshape(-1)[5] = 0
    for bad in (wrong_sign, swapped_axes, normalised_inverse, dropped_term):
        assert rel(bad, ref) > 1e-6
    # a transform computed on transposed storage is still the right transform (control)
    assert rel(transposed, ref) < 1e-14
    # 1D: forgetting the bit reversal is caught by the closed-form check
    n = 64
    xx = orc.uniform(n, 4)
    natural = np.fft.fft(xx.astype(np.complex128))
    assert rel(natural, orc.c2c_1d_expected(xx, n, 1)[0]) > 0.1

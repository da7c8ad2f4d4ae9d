"""CPU emulation of the GPU pass structure, before and beside the GPU runs.

1. Stockham stages: the per-thread index maps of k_fft_pass (stage loads at
   j + r*L/R, twiddle W_{NS R}^{(j mod NS) r}, stores at
   (j/NS)*NS*R + (j mod NS) + r*NS, radix schedule of fft_schedule.h) are
   replayed in numpy for every supported line length and must give the DFT.
2. Pass plans: the library's planner (fftamd_describe_plan, host-only) is
   replayed pass by pass with the kernel's address formulas; every read and
   write must stay inside the buffer, each pass must write every output
   element exactly once, and the composition must equal the oracle's
   transform (four-step twiddles, bit-reversed 1D order and buffer roles
   included).
3. The fused 256 x 256 plane pass (fft_xy.hip): its lane, shuffle and LDS
   index maps and its workgroup-to-plane mapping are replayed and must give
   the 2D DFT of every plane with every output written once.
"""
import ctypes

import numpy as np
import pytest

NF = 22


def sched(L):
    lg = L.bit_length() - 1
    np_ = 1 if L <= 16 else (lg + 3) // 4
    rad = [(1 << (lg - 4 * (np_ - 1)))] + [16] * (np_ - 1)
    ns = [int(np.prod(rad[:s])) for s in range(np_)]
    return np_, rad, ns


def emulate_line(x):
    """Replays one line through the kernel's stage index maps (forward)."""
    L = x.size
    if L == 1:
        return x.copy()
    NP, RAD, NS = sched(L)
    PPT = min(L, 16)
    TPL = L // PPT
    v = np.zeros((TPL, PPT), np.complex128)
    lds = np.full(L, np.nan + 0j)
    for S in range(NP):
        R, ns = RAD[S], NS[S]
        nb = PPT // R
        for tau in range(TPL):
            for b in range(nb):
                j = tau + b * TPL
                for r in range(R):
                    v[tau, b * R + r] = x[j + r * (L // R)] if S == 0 else lds[j + r * (L // R)]
                if S > 0:
                    k = j & (ns - 1)
                    for r in range(1, R):
                        v[tau, b * R + r] *= np.exp(-2j * np.pi * k * r / (ns * R))
                v[tau, b * R:(b + 1) * R] = np.fft.fft(v[tau, b * R:(b + 1) * R])
        if S + 1 < NP:
            new = np.full(L, np.nan + 0j)
            for tau in range(TPL):
                for b in range(nb):
                    j = tau + b * TPL
                    base = (j // ns) * ns * R + (j & (ns - 1))
                    for r in range(R):
                        assert np.isnan(new[base + r * ns])  # written once
                        new[base + r * ns] = v[tau, b * R + r]
            lds = new
    R, ns = RAD[-1], NS[-1]
    X = np.full(L, np.nan + 0j)
    for tau in range(TPL):
        for b in range(PPT // R):
            j = tau + b * TPL
            for r in range(R):
                X[j + r * ns] = v[tau, b * R + r]
    return X


@pytest.mark.parametrize("lg", range(0, 14))
def test_stockham_stage_maps(lg):
    L = 1 << lg
    rng = np.random.default_rng(lg)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    X = emulate_line(x)
    assert not np.isnan(X).any()
    assert np.allclose(X, np.fft.fft(x), rtol=1e-12, atol=1e-10)


def describe(fl, dims, batch, flags=0, prec=0, aliased=False):
    lib = fl.load()
    rank = len(dims)
    cd = (ctypes.c_uint * rank)(*dims)
    n = lib.fftamd_describe_plan(rank, cd, batch, flags, prec, int(aliased), None, 0)
    assert n >= 0, fl.last_error()
    buf = (ctypes.c_longlong * (NF * n + 1))()
    assert lib.fftamd_describe_plan(rank, cd, batch, flags, prec, int(aliased), buf, len(buf)) == n
    arr = np.frombuffer(buf, dtype=np.int64)
    keys = ["lg", "mode", "epi", "transposing", "in_p", "out_p", "in_c", "out_c", "in_k", "out_k",
            "in_o", "out_o", "ncol", "K", "kbits", "nouter", "M", "cshift", "src", "dst", "C", "NT"]
    return [dict(zip(keys, (int(v) for v in arr[NF * i:NF * (i + 1)]))) for i in range(n)], \
        bool(arr[NF * n])


def brev(x, bits):
    r = np.zeros_like(x)
    for i in range(bits):
        r |= ((x >> i) & 1) << (bits - 1 - i)
    return r


def run_plan(passes, final_copy, x, inverse, aliased=False):
    total = x.size
    bufs = {0: x.astype(np.complex128).copy(), 1: np.zeros(total, np.complex128),
            2: np.zeros(total, np.complex128)}
    if aliased:
        bufs[1] = bufs[0]
    cur = 1 if aliased else 0
    for p in passes:
        L = 1 << p["lg"]
        C = p["C"]
        if p["mode"] == 2:  # fused plane pass: 2D transform of every (K x 2^lg) plane
            LY, LX = p["K"], 1 << p["lg"]
            assert p["transposing"] == (LY == 256 and LX == 256)  # the 256 x 256 pass runs out of place
            if p["transposing"]:
                assert p["src"] != p["dst"]
            assert p["nouter"] * p["in_o"] == total and p["in_o"] == LY * LX
            planes = bufs[p["src"]].reshape(p["nouter"], LY, LX)
            X = np.fft.ifft2(planes) * (LY * LX) if inverse else np.fft.fft2(planes)
            bufs[p["dst"]][:] = X.reshape(-1)
            cur = p["dst"]
            continue
        if p["mode"] in (4, 5):  # n^3 blocks, y split q x A (y = a + A b, ky = k + q d)
            def dft(v, ax):
                return np.fft.ifft(v, axis=ax) * v.shape[ax] if inverse else np.fft.fft(v, axis=ax)
            assert p["src"] == cur
            n = 1 << p["lg"]
            assert n in (128, 256, 512)
            A = 8 if n == 512 else 4
            q = n // A
            if p["mode"] == 4:  # x, then over b; the column DFT of residue a to row q a + k
                assert p["transposing"] and p["src"] != p["dst"]
                assert p["in_o"] == n * n and p["nouter"] * n * n == total
                pl = bufs[p["src"]].reshape(p["nouter"], n, n)
                res = np.empty_like(pl)
                for a in range(A):
                    res[:, q * a:q * a + q, :] = dft(dft(pl[:, a::A, :], 2), 1)
            else:  # row k + q a times Wn^{a k}, over a to rows k + q d, then z
                assert p["in_o"] == n ** 3 and p["nouter"] * n ** 3 == total
                blocks = bufs[p["src"]].reshape(p["nouter"], n, A, q, n)
                ak = np.arange(A)[:, None] * np.arange(q)[None, :]
                w = np.exp((2j if inverse else -2j) * np.pi * ak / n)
                res = dft(dft(blocks * w[None, None, :, :, None], 2), 1)
            bufs[p["dst"]][:] = res.reshape(-1)
            cur = p["dst"]
            continue
        ncb = (p["ncol"] + C - 1) // C
        # every active line (o, k, col); the grid covers col < ncb*C, inactive ones store nothing
        o, k, col = np.meshgrid(np.arange(p["nouter"]), np.arange(p["K"]), np.arange(ncb * C),
                                indexing="ij")
        o, k, col = o.reshape(-1), k.reshape(-1), col.reshape(-1)
        act = col < p["ncol"]
        o, k, col = o[act], k[act], col[act]
        ko = brev(k, p["kbits"]) if p["kbits"] else k
        pts = np.arange(L)
        iidx = (o * p["in_o"] + k * p["in_k"] + col * p["in_c"])[:, None] + pts[None, :] * p["in_p"]
        oidx = (o * p["out_o"] + ko * p["out_k"] + col * p["out_c"])[:, None] + pts[None, :] * p["out_p"]
        assert iidx.min() >= 0 and iidx.max() < total
        assert oidx.min() >= 0 and oidx.max() < total
        assert np.unique(oidx).size == oidx.size == total, "each output written exactly once"
        assert p["src"] == cur
        src = bufs[p["src"]]
        lines = src[iidx]
        if inverse:
            X = np.fft.ifft(lines, axis=1) * L
        else:
            X = np.fft.fft(lines, axis=1)
        if p["epi"] == 1:  # four-step post-twiddle, applied in the forward (swapped) domain
            t = ((col >> p["cshift"])[:, None] * pts[None, :]) % p["M"]
            w = np.exp(-2j * np.pi * t / p["M"])
            X = np.conj(w) * X if inverse else w * X
        if p["epi"] == 2:  # point k stored at bitrev(k)
            X = X[:, brev(pts, p["lg"])]
        if p["src"] == p["dst"]:
            assert not p["transposing"]
            assert np.array_equal(np.sort(iidx.reshape(-1)), np.sort(oidx.reshape(-1)))
        dst = bufs[p["dst"]]
        dst[oidx] = X
        cur = p["dst"]
    if final_copy:
        bufs[1][:] = bufs[cur]
        cur = 1
    assert cur == 1
    return bufs[1]


CASES = [([8, 8, 8], 1, 0), ([64, 64, 64], 1, 0), ([4, 16, 32], 2, 0), ([16, 16], 3, 0),
         ([1024, 8], 1, 0), ([2048, 32], 1, 0), ([8, 4096], 1, 0), ([1, 1, 1], 2, 0),
         ([2, 1, 4], 1, 0), ([256], 5, 1), ([1 << 14], 2, 1), ([1 << 14], 2, 0), ([1 << 16], 1, 1),
         ([4096, 4], 1, 0), ([1 << 10, 1 << 10], 1, 0), ([2, 256, 256], 1, 0), ([256, 256], 2, 0),
         ([256, 256], 1, 4), ([2, 256, 256], 1, 4), ([4, 128, 128], 1, 0), ([128, 64], 3, 0),
         ([64, 256, 256], 1, 0), ([256, 256], 65, 0), ([256, 256, 256], 1, 0), ([256, 256, 256], 1, 8),
         ([128, 128, 128], 2, 0), ([128, 128, 128], 1, 8),
         ([2, 16, 64], 2, 0), ([32, 32, 32], 1, 0), ([64, 64, 64], 1, 4), ([2, 512, 64], 1, 2),
         ([4, 128, 128], 1, 2), ([4096, 1024], 1, 0), ([4, 4096, 256], 1, 0), ([4096, 16], 64, 0)]


@pytest.mark.parametrize("dims,batch,flags", CASES)
@pytest.mark.parametrize("inverse", [False, True])
def test_plan_replay_matches_oracle(fl, orc, dims, batch, flags, inverse):
    passes, final_copy = describe(fl, dims, batch, flags)
    n = int(np.prod(dims)) * batch
    x = orc.uniform(n, 77 + n % 101)
    got = run_plan(passes, final_copy, x, inverse)
    if flags & 1 and len(dims) == 1:
        ref = orc.c2c_1d_expected(x, dims[0], batch, inverse=inverse).reshape(-1)
    else:
        ref = orc.fftnd(x, dims, batch=batch, inverse=inverse).reshape(-1)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("dims,batch,flags", [([64, 64, 64], 1, 0), ([2048, 32], 1, 0),
                                              ([1 << 14], 1, 1), ([1 << 14], 1, 0),
                                              ([2, 256, 256], 1, 0), ([256, 256], 2, 0),
                                              ([64, 256, 256], 1, 0), ([4096, 1024], 1, 0),
                                              ([256, 256, 256], 1, 0), ([128, 128, 128], 2, 0)])
def test_plan_replay_in_place(fl, orc, dims, batch, flags):
    passes, final_copy = describe(fl, dims, batch, flags, aliased=True)
    n = int(np.prod(dims)) * batch
    x = orc.uniform(n, 5)
    got = run_plan(passes, final_copy, x, False, aliased=True)
    ref = (orc.c2c_1d_expected(x, dims[0], batch) if flags & 1 else
           orc.fftnd(x, dims, batch=batch)).reshape(-1)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_plan_shapes_of_headline_configs(fl):
    # 256^3: the y-split pair (x and y over b, then y over a and z), no temporary buffer
    p, fc = describe(fl, [256, 256, 256], 1)
    assert [q["mode"] for q in p] == [4, 5] and not fc and all(q["dst"] == 1 for q in p)
    assert [(q["src"], q["dst"]) for q in p] == [(0, 1), (1, 1)]
    # in place: pass A writes the temporary, pass B writes back
    p, fc = describe(fl, [256, 256, 256], 1, aliased=True)
    assert [(q["src"], q["dst"]) for q in p] == [(1, 2), (2, 1)] and not fc
    # batches of 256^3 blocks keep the pair
    assert [q["mode"] for q in describe(fl, [256, 256, 256], 3)[0]] == [4, 5]
    # without the y split: the fused x+y plane pass, then the strided z pass
    p, fc = describe(fl, [256, 256, 256], 1, flags=8)
    assert [q["mode"] for q in p] == [2, 1] and not fc and all(q["dst"] == 1 for q in p)
    p, fc = describe(fl, [256, 256, 256], 1, flags=8, aliased=True)
    assert [(q["src"], q["dst"]) for q in p] == [(1, 2), (2, 1)] and not fc
    # other depths over 256 x 256 planes keep the plane pass
    assert [q["mode"] for q in describe(fl, [128, 256, 256], 1)[0]] == [2, 1]
    # 128^3 blocks (the differentiable 16 x 128^3 configuration): the y-split pair
    p, fc = describe(fl, [128, 128, 128], 16)
    assert [(q["mode"], q["lg"], q["NT"]) for q in p] == [(4, 7, 256), (5, 7, 512)] and not fc
    assert [q["mode"] for q in describe(fl, [128, 128, 128], 16, flags=8)[0]] == [2, 1]
    assert [q["mode"] for q in describe(fl, [64, 128, 128], 1)[0]] == [2, 1]
    # without plane fusion: three single passes (x contiguous, y and z strided)
    p, fc = describe(fl, [256, 256, 256], 1, flags=4)
    assert [q["mode"] for q in p] == [0, 1, 1] and not fc and all(q["dst"] == 1 for q in p)
    # 256 x 256 planes go through the plane pass from 32 planes up
    assert [q["mode"] for q in describe(fl, [256, 256], 32)[0]] == [2]
    assert [q["mode"] for q in describe(fl, [256, 256], 16)[0]] == [0, 1]
    # double precision keeps the separate passes; split mode runs the y-split
    # pairs with split radix-16 stages, and without them the three line passes
    assert [q["mode"] for q in describe(fl, [256, 256, 256], 1, prec=1)[0]] == [0, 1, 1]
    for n in (128, 256, 512):
        assert [q["mode"] for q in describe(fl, [n, n, n], 1, flags=2)[0]] == [4, 5]
    assert [q["mode"] for q in describe(fl, [256, 256, 256], 1, flags=2 | 8)[0]] == [0, 1, 1]
    # 512^3 blocks: the y-split pair with y = a + 8 b, 1024-thread workgroups
    p, fc = describe(fl, [512, 512, 512], 1)
    assert [(q["mode"], q["lg"], q["NT"], q["C"]) for q in p] == [(4, 9, 1024, 64), (5, 9, 1024, 8)] and not fc
    assert [(q["src"], q["dst"]) for q in p] == [(0, 1), (1, 1)]
    p, fc = describe(fl, [512, 512, 512], 1, aliased=True)
    assert [(q["src"], q["dst"]) for q in p] == [(1, 2), (2, 1)] and not fc
    assert [q["mode"] for q in describe(fl, [512, 512, 512], 1, flags=8)[0]] == [0, 1, 1]
    # 1D 2^22 bit-reversed: four-step, strided 1024 (16 columns) then contiguous 4096
    p, fc = describe(fl, [1 << 22], 1, 1)
    assert [(q["lg"], q["mode"], q["epi"], q["C"], q["NT"]) for q in p] == [(10, 1, 1, 16, 1024), (12, 0, 2, 1, 256)]
    # 2^21 keeps 256 x 8192; 2^23 fits as 1024 x 8192
    assert [q["lg"] for q in describe(fl, [1 << 21], 1, 1)[0]] == [8, 13]
    assert [q["lg"] for q in describe(fl, [1 << 23], 1, 1)[0]] == [10, 13]
    # strided axes of 3D arrays keep 512-point lines (1024 as a 32 x 32 four-step pair)
    assert [q["lg"] for q in describe(fl, [1024, 1024, 1024], 1)[0]] == [10, 5, 5, 5, 5]
    # 2D 4096^2: the 4096-point column pass first (out of place), then the
    # contiguous 4096-point rows in place on the output; no temporary buffer
    p, fc = describe(fl, [4096, 4096], 1)
    assert [(q["lg"], q["mode"]) for q in p] == [(12, 3), (12, 0)] and not fc
    assert [(q["src"], q["dst"]) for q in p] == [(0, 1), (1, 1)]
    # without fused passes: contiguous 4096, then the strided axis split 64 x 64
    p, fc = describe(fl, [4096, 4096], 1, flags=4)
    assert [(q["lg"], q["mode"]) for q in p] == [(12, 0), (6, 1), (6, 1)]
    # fewer than 64 column groups keep the four-step pair
    assert [q["mode"] for q in describe(fl, [4096, 512], 1)[0]] == [0, 1, 1]
    assert [q["mode"] for q in describe(fl, [4096, 1024], 1)[0]] == [3, 0]
    # every workgroup stays within 1024 threads
    for dims in ([2, 2, 2], [8192], [512, 512], [1 << 22]):
        for q in describe(fl, dims, 1)[0]:
            assert q["NT"] <= 1024


@pytest.mark.parametrize("dims,batch", [([1024, 1024, 1024], 1), ([16384, 32768], 1), ([256, 256, 256], 64)])
def test_plan_bounds_of_large_arrays(fl, dims, batch):
    """Arrays past 2^31 bytes (the GPU single-tone cases): every pass's first and
    last line address stays inside the array (addresses are linear in the line
    and point indices with non-negative strides, so the extremes bound them),
    the grids fit the launch limits, and each pass covers the array once."""
    passes, _ = describe(fl, dims, batch)
    total = int(np.prod(dims)) * batch
    assert total * 8 >= 1 << 32
    for p in passes:
        L = 1 << p["lg"]
        if p["mode"] in (2, 4, 5):
            assert p["nouter"] * p["in_o"] == total
            continue
        for s, o in (("in_p", "in_o"), ("out_p", "out_o")):
            k, c = s.replace("_p", "_k"), s.replace("_p", "_c")
            assert min(p[s], p[o], p[k], p[c]) >= 0
            hi = (p["nouter"] - 1) * p[o] + (p["K"] - 1) * p[k] + (p["ncol"] - 1) * p[c] + (L - 1) * p[s]
            assert hi < total, (p, s)
        assert p["nouter"] * p["K"] * p["ncol"] * L == total
        blocks = p["nouter"] * p["K"] * ((p["ncol"] + p["C"] - 1) // p["C"])
        assert blocks * p["NT"] < 1 << 32


@pytest.mark.parametrize("lg", range(1, 14))
def test_bitrev_epilogue_runs(lg):
    """The bit-reversed epilogue writes point k = j + r*NS of the last stage to
    RL*bitrev(j) + bitrev(r); that must equal bitrev(k) over log2 L bits."""
    L = 1 << lg
    _, rad, ns = sched(L)
    RL, NSL = rad[-1], ns[-1]
    lgr = RL.bit_length() - 1
    j = np.arange(NSL)[:, None]
    r = np.arange(RL)[None, :]
    k = j + r * NSL
    got = RL * brev(j, lg - lgr) + brev(r, lgr)
    assert np.array_equal(got, brev(k, lg))
    assert np.unique(got).size == L


# ---- fused 256 x 256 plane pass ------------------------------------------------

def _row_idx(r, p):
    """LDS index of point p of local row r (the line passes' [row][p + p/16] layout)."""
    return r * 272 + p + (p >> 4)


def _grow(r, q):
    return 16 * (r >> 2) + 4 * (r & 3) + q


def _mul_w16(x, p):
    return x * np.exp(-2j * np.pi * (p % 16) / 16)


def _dft16_residue(v, g):
    """fft_xy.hip's dft16_residue: outputs g, g+4, g+8, g+12 of a 16-point DFT
    as four rotated 4-term sums, fixed twiddles W16^{a g}, one 4-point DFT."""
    rot = lambda x, m: x * (-1j) ** (m % 4)  # noqa: E731
    s = [v[..., a] + rot(v[..., a + 4], g) + rot(v[..., a + 8], 2 * g) + rot(v[..., a + 12], 3 * g)
         for a in range(4)]
    s = [_mul_w16(s[a], a * g) for a in range(4)]
    return np.stack([sum(s[a] * (-1j) ** ((a * u) % 4) for a in range(4)) for u in range(4)], axis=-1)


@pytest.mark.parametrize("g", range(4))
def test_dft16_residue(g):
    rng = np.random.default_rng(g)
    v = rng.standard_normal((5, 16)) + 1j * rng.standard_normal((5, 16))
    assert np.allclose(_dft16_residue(v, g), np.fft.fft(v, axis=1)[:, g::4], atol=1e-12)


def emulate_plane256(plane):
    """Replays k_fft_xy256's index maps for one plane (all four workgroups of
    512 threads): column stages per column half, row transforms per 32 rows."""
    out = np.full((256, 256), np.nan + 0j)
    tid = np.arange(512)
    w256 = lambda j: np.exp(-2j * np.pi * (j % 256) / 256)  # noqa: E731
    r16 = np.arange(16)
    x, g = tid & 127, tid >> 7
    for q in range(4):
        res = []
        for h in range(2):
            # column first stage: lane (x, g), n1 = 4g + m, outputs k1 = q + 4u
            col = 128 * h + x
            ybuf = np.full(16 * 4 * 128, np.nan + 0j)
            for m in range(4):
                n1 = 4 * g + m
                rowe = (n1[:, None] + 16 * r16[None, :]) * 256
                # buffer-load byte offsets stay inside the 512 KiB plane
                assert (col[:, None] * 8 + rowe * 8).max() < 256 * 256 * 8
                v = plane[n1[:, None] + 16 * r16[None, :], col[:, None]]
                o = _dft16_residue(v, q)
                for u in range(4):
                    idx = (4 * n1 + u) * 128 + x
                    assert np.isnan(ybuf[idx]).all()
                    ybuf[idx] = o[:, u]
            assert not np.isnan(ybuf).any()
            # column second stage: lane (x, u = g), k1 = q + 4u
            u = g
            v = ybuf[(4 * r16[None, :] + u[:, None]) * 128 + x[:, None]]
            res.append(np.fft.fft(v * w256((q + 4 * u)[:, None] * r16[None, :]), axis=1))
        for j in range(2):
            # lane (x, u) stores its points k2 = 8j .. 8j+7 of both halves as
            # local rows 4 kk + u; then the line-pass stages run on the 32 rows
            rows = np.full(32 * 272, np.nan + 0j)
            u = g
            for kk in range(8):
                for h in range(2):
                    idx = _row_idx(4 * kk + u, 128 * h + x)
                    assert np.isnan(rows[idx]).all()
                    rows[idx] = res[h][:, 8 * j + kk]
            pts = np.arange(256)
            for c in range(32):
                line = rows[_row_idx(c, pts)]
                assert not np.isnan(line).any()
                assert np.isnan(out[_grow(32 * j + c, q)]).all()
                out[_grow(32 * j + c, q)] = np.fft.fft(line)
    return out


def test_plane256_index_maps():
    rng = np.random.default_rng(3)
    plane = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    got = emulate_plane256(plane)
    assert not np.isnan(got).any()
    assert np.allclose(got, np.fft.fft2(plane), rtol=1e-12, atol=1e-9)
    # 32 LDS rows fit the buffer and do not overlap; the buffer (the larger of
    # the row block and the 16 x 4 x 128 first-stage block) lets two
    # workgroups share a CU's 160 KiB
    r, p = np.meshgrid(np.arange(32), np.arange(256), indexing="ij")
    idx = _row_idx(r, p)
    assert idx.max() < 32 * 272 and np.unique(idx).size == idx.size
    assert 2 * max(32 * 272, 16 * 4 * 128) * 8 <= 160 * 1024
    assert sorted(_grow(r, q) for q in range(4) for r in range(64)) == list(range(256))


def emulate_plane_halves(plane):
    """Replays k_fft_plane_halves' index maps for one 128 x 128 plane (512
    threads): row transforms of rows 0..63 then 64..127 held in registers,
    then the columns in two halves of 64 through the strided LDS layout."""
    L, H, TPL = 128, 64, 8
    tid = np.arange(512)
    rc, rtau = tid // TPL, tid % TPL
    r16 = np.arange(16)
    # after the row stages, thread (rc, rtau) holds point rtau + TPL r of its row
    rr = [np.fft.fft(plane[rc + H * h], axis=1)[np.arange(512)[:, None], rtau[:, None] + TPL * r16[None, :]]
          for h in range(2)]
    out = np.full((L, L), np.nan + 0j)
    cc, ctau = tid % H, tid // H
    for h in range(2):
        lds = np.full(L * (H + 1), np.nan + 0j)   # [y][column] with one pad column
        for rh in range(2):
            for r in range(8):
                col, y = rtau + TPL * r, rc + H * rh
                idx = y * (H + 1) + col
                assert np.isnan(lds[idx]).all()
                lds[idx] = rr[rh][:, 8 * h + r]
        assert np.isnan(lds.reshape(L, H + 1)[:, H]).all()  # pad column unused
        cols = lds.reshape(L, H + 1)[:, :H]
        assert not np.isnan(cols).any()
        # column stages (the line passes' strided stages), then point y of
        # column cc goes to e = y * L + H h + cc
        spec = np.fft.fft(cols, axis=0)
        for t in range(512):
            if ctau[t] == 0:
                y = np.arange(L)
                e = y * L + H * h + cc[t]
                assert np.isnan(out.reshape(-1)[e]).all()
                out.reshape(-1)[e] = spec[y, cc[t]]
    return out


def test_plane_halves_index_maps():
    rng = np.random.default_rng(4)
    plane = rng.standard_normal((128, 128)) + 1j * rng.standard_normal((128, 128))
    got = emulate_plane_halves(plane)
    assert not np.isnan(got).any()
    assert np.allclose(got, np.fft.fft2(plane), rtol=1e-12, atol=1e-9)
    # the larger of the 64-row block ([row][x + x/16]) and the 64-column block
    # fits twice in a CU's 160 KiB
    assert 2 * max(64 * (128 + 8), 128 * 65) * 8 <= 160 * 1024


def emulate_col4096(block, lx):
    """Replays k_fft_col4096's index maps for one 16-column group of a 4096 x lx
    block (all four workgroups of 512 threads), with its byte offsets."""
    out = np.full((4096, 16), np.nan + 0j)
    tid = np.arange(512)
    c16, gi = tid & 15, tid >> 4
    r16 = np.arange(16)
    lx8 = lx * 8
    span = 4095 * lx8 + 16 * 8
    w4096 = lambda t: np.exp(-2j * np.pi * (t % 4096) / 4096)  # noqa: E731
    for q in range(4):
        # first stage: thread (c16, gi), n1 = gi + 32 i, rows n1 + 256 n2
        lines = {}
        for i in range(8):
            n1 = gi + 32 * i
            voff = gi * lx8 + c16 * 8
            soff = (32 * i + 256 * r16) * lx8
            byte = voff[:, None] + soff[None, :]
            assert byte.max() + 8 <= span and (byte % 8 == 0).all()
            rows = byte // lx8
            cols = (byte % lx8) // 8
            assert np.array_equal(rows, n1[:, None] + 256 * r16[None, :])
            v = block[rows, cols]
            o = _dft16_residue(v, q)
            for u in range(4):
                for t in range(512):
                    lines.setdefault((u, int(c16[t])), np.full(256, np.nan + 0j))[n1[t]] = o[t, u]
        # stage 2 per half h: LDS lines c = 16 l + c16 (u = 2h + l), point n1 at n1 * 33 + c
        for h in range(2):
            lds = np.full(256 * 33, np.nan + 0j)
            for l in range(2):
                for c in range(16):
                    idx = np.arange(256) * 33 + 16 * l + c
                    assert np.isnan(lds[idx]).all()
                    lds[idx] = lines[(2 * h + l, c)]
            c, tau = tid % 32, tid // 32
            k1 = q + 4 * (2 * h + (c >> 4))
            for t in range(512):
                if tau[t]:
                    continue  # one thread per line replays the line
                line = lds[np.arange(256) * 33 + c[t]]
                assert not np.isnan(line).any()
                spec = np.fft.fft(line * w4096(np.arange(256) * k1[t]))
                for tt in range(16):  # thread tau stores k2 = tau + 16 r at row k1 + 16 k2
                    voff = (k1[t] + 16 * tt) * lx8 + (c[t] & 15) * 8
                    byte = voff + 256 * r16 * lx8
                    assert byte.max() + 8 <= span
                    rows, cols = byte // lx8, (byte % lx8) // 8
                    assert np.isnan(out[rows, cols]).all()
                    out[rows, cols] = spec[tt + 16 * r16]
    return out


@pytest.mark.parametrize("lx", [16, 64])
def test_col4096_index_maps(lx):
    rng = np.random.default_rng(lx)
    block = rng.standard_normal((4096, 16)) + 1j * rng.standard_normal((4096, 16))
    got = emulate_col4096(block, lx)
    assert not np.isnan(got).any()
    assert np.allclose(got, np.fft.fft(block, axis=0), rtol=1e-10, atol=1e-8)
    # 32 lines of 256 points with one pad column fit twice in a CU's 160 KiB
    assert 2 * 256 * 33 * 8 <= 160 * 1024


@pytest.mark.parametrize("nplanes", [1, 7, 8, 9, 256, 1000])
def test_plane256_block_mapping(nplanes):
    """Grid of ceil(P/8)*8*4 workgroups: block b is XCD b%8, quarter (b/8)%4 of
    plane (b/32)*8 + b%8; every (plane, quarter) is taken exactly once."""
    blocks = (nplanes + 7) // 8 * 8 * 4
    b = np.arange(blocks)
    xcd, k = b & 7, b >> 3
    q, plane = k & 3, (k >> 2) * 8 + xcd
    live = plane < nplanes
    pairs = plane[live] * 4 + q[live]
    assert np.array_equal(np.sort(pairs), np.arange(nplanes * 4))
    # the four quarters of a plane share an XCD and are consecutive on it
    for p in range(min(nplanes, 16)):
        bs = b[live & (plane == p)]
        assert len(set(bs & 7)) == 1 and np.array_equal(np.diff(np.sort(bs)), [8, 8, 8])


# ---- the y-split pair of n^3 blocks (fft_cube.hip): thread, LDS and store maps ----------

def _cube_geom(n):
    """Per-side constants of k_fft_cube_a/b (n = 256) and k_fft_cube128_a/b (n = 128)."""
    if n == 256:
        return dict(nt_a=512, tpl_row=16, halves=2, rows_per_half=32, ldsa_pad=129, col_tpl=4,
                    nt_b=512, cols_b=16, zs_b=32, ldsb_pad=33, tpl_z=16)
    return dict(nt_a=256, tpl_row=8, halves=1, rows_per_half=32, ldsa_pad=129, col_tpl=2,
                nt_b=512, cols_b=32, zs_b=16, ldsb_pad=65, tpl_z=8)


def emulate_cube_a(blocks, n):
    """Pass A over planes [..., n, n]: thread (c, tau) holds point kx = tau + TPL r of
    row b = rows_per_half h + c (y = res + 4b); the column halves go to LDS at
    [row][col] (one pad column) -- for n = 256 as the radix-2 partial sums of the
    4-point stage, for n = 128 as is; column thread (cc, tq) stores point
    k = tq + col_tpl r of its column to row (n/4) res + k."""
    g = _cube_geom(n)
    q = n // 4
    planes = blocks.reshape(-1, n, n)
    out = np.full(planes.shape, np.nan + 0j)
    tid = np.arange(g["nt_a"])
    c, tau = tid // g["tpl_row"], tid % g["tpl_row"]
    r = np.arange(n // g["tpl_row"])
    ncol_half = n // (2 if g["halves"] == 2 else 1)
    cc, tq = tid % ncol_half, tid // ncol_half

    def dft_a(v, ax):
        return np.fft.fft(v, axis=ax)

    for res in range(4):
        rowfft = dft_a(planes[:, res::4, :], 2)  # [plane, b, kx]
        regs = []
        for h in range(g["halves"]):
            b = g["rows_per_half"] * h + c
            regs.append(rowfft[:, b[:, None], (tau[:, None] + g["tpl_row"] * r[None, :])])  # [plane, tid, r]
        for hc in range(g["halves"]):
            lds = np.full((planes.shape[0], q * g["ldsa_pad"]), np.nan + 0j)
            per = len(r) // g["halves"]
            p = np.arange(q)
            if n == 256:
                # b = b0 + 16 (lo + 2 hi); thread c holds b0 = c & 15, lo = c >> 4,
                # hi = 0 (r0) and 1 (r1); E_lo[m] = r0 + (-1)^m r1 to row b0 + 16 (m + 2 lo),
                # E_1[1] stored times -i
                b0, lo = c & 15, c >> 4
                for rr in range(per):
                    col = tau + g["tpl_row"] * rr
                    u, w = regs[0][:, :, per * hc + rr], regs[1][:, :, per * hc + rr]
                    rot = np.where(lo == 1, -1j, 1)[None, :]
                    for m, val in ((0, u + w), (1, (u - w) * rot)):
                        idx = (b0 + 16 * (m + 2 * lo)) * g["ldsa_pad"] + col
                        assert np.isnan(lds[0, idx]).all(), "LDS slot written twice"
                        lds[:, idx] = val
                rows_ = lds[:, p[None, :] * g["ldsa_pad"] + np.arange(ncol_half)[:, None]]  # [plane, col, row]
                assert not np.isnan(rows_).any(), "LDS slot left empty"
                colfft = np.empty_like(rows_)
                sgn = -1
                for k1 in range(4):  # X4 = E_0[m] +- E'_1[m] (minus for k1 >= 2), twiddle W64^{b0 k1}, DFT16 over b0
                    m = k1 & 1
                    x4 = rows_[:, :, 16 * m:16 * m + 16] + (-1 if k1 & 2 else 1) * rows_[:, :, 16 * (m + 2):16 * (m + 2) + 16]
                    x4 = x4 * np.exp(sgn * 2j * np.pi * k1 * np.arange(16) / 64)
                    colfft[:, :, k1::4] = dft_a(x4, 2)
            else:
                for h in range(g["halves"]):
                    for rr in range(per):
                        col = tau + g["tpl_row"] * rr
                        idx = (c + g["rows_per_half"] * h) * g["ldsa_pad"] + col
                        assert np.isnan(lds[0, idx]).all(), "LDS slot written twice"
                        lds[:, idx] = regs[h][:, :, per * hc + rr]
                lines = lds[:, p[None, :] * g["ldsa_pad"] + np.arange(ncol_half)[:, None]]  # [plane, col, p]
                assert not np.isnan(lines).any(), "LDS column slot left empty"
                colfft = dft_a(lines, 2)  # [plane, col, k]
            kk = tq[:, None] + g["col_tpl"] * np.arange(q // g["col_tpl"])[None, :]
            orow = q * res + kk
            ocol = ncol_half * hc + np.broadcast_to(cc[:, None], kk.shape)
            assert np.isnan(out[0, orow, ocol]).all(), "output written twice"
            out[:, orow, ocol] = colfft[:, cc[:, None], kk]
    assert not np.isnan(out).any()
    return out.reshape(blocks.shape)


def emulate_cube_b(arr, n):
    """Pass B over blocks [..., n, n, n] in place: thread (x, zs) loads rows
    k + (n/4) a of column cols xg + x in planes zs + zs_b j, applies Wn^{a k} and
    the 4-point DFT, and does an 8-point DFT over j for each line of rows
    d = 2e, 2e + 1 into LDS ([slot][line], one pad column): for n = 256 the
    first half of the z transform's radix-16 stage, for n = 128 its radix-8
    stage; line thread (cz, tz) finishes the z transform and stores
    kz = tz + TPL r."""
    g = _cube_geom(n)
    q, cb = n // 4, g["cols_b"]
    blk = arr.reshape(-1, n, n, n).copy()
    out = np.full(blk.shape, np.nan + 0j)
    tid = np.arange(g["nt_b"])
    x, zs = tid % cb, tid // cb
    j = np.arange(8)
    z = zs[:, None] + g["zs_b"] * j[None, :]  # [tid, j]
    assert np.array_equal(np.sort(z.reshape(-1)), np.repeat(np.arange(n), cb))
    cz, tz = tid % (2 * cb), tid // (2 * cb)
    for k in range(q):
        w = np.exp(-2j * np.pi * np.arange(4) * k / n)
        for xg in range(n // cb):
            cols = cb * xg + x
            t = np.stack([blk[:, z, (k + q * a), cols[:, None]] for a in range(4)], axis=-1)  # [blk, tid, j, a]
            t = np.fft.fft(t * w, axis=-1)  # d
            for e in range(2):
                lds = np.full((blk.shape[0], n * g["ldsb_pad"]), np.nan + 0j)
                for dd in range(2):
                    # the thread's 8-point DFT over j (planes zs + zs_b j) to its LDS slots
                    p8 = np.fft.fft(t[..., 2 * e + dd], axis=2)  # [blk, tid, m]
                    if n == 256:  # slot zq + 16 ep + 32 m, zs = zq + 16 ep
                        slot = zs[:, None] + 32 * np.arange(8)[None, :]
                    else:  # Stockham stage-0 position 8 zs + r
                        slot = 8 * zs[:, None] + np.arange(8)[None, :]
                    idx = slot * g["ldsb_pad"] + (x + cb * dd)[:, None]
                    assert np.isnan(lds[0, idx]).all(), "LDS slot written twice"
                    lds[:, idx] = p8
                zz = np.arange(n)
                lines = lds[:, zz[None, :] * g["ldsb_pad"] + np.arange(2 * cb)[:, None]]  # [blk, line, slot]
                assert not np.isnan(lines).any(), "LDS slot left empty"
                zf = np.empty_like(lines)
                nk0 = g["tpl_z"]
                for k0 in range(nk0):  # the stage-1 thread for k0: its 16 inputs, twiddles, DFT16
                    if n == 256:
                        m = k0 & 7
                        v = lines[:, :, 32 * m:32 * m + 16] + np.exp(-2j * np.pi * k0 / 16) * lines[:, :, 32 * m + 16:32 * m + 32]
                    else:
                        v = lines[:, :, k0::8]
                    v = v * np.exp(-2j * np.pi * k0 * np.arange(16) / n)
                    zf[:, :, k0::nk0] = np.fft.fft(v, axis=2)
                kz = tz[:, None] + g["tpl_z"] * np.arange(n // g["tpl_z"])[None, :]
                orow = k + q * (2 * e + cz // cb)
                ocol = cb * xg + cz % cb
                sel = out[0, kz, orow[:, None], ocol[:, None]]
                assert np.isnan(sel).all(), "output written twice"
                out[:, kz, orow[:, None], ocol[:, None]] = zf[:, cz[:, None], kz]
    assert not np.isnan(out).any()
    return out.reshape(arr.shape)


@pytest.mark.parametrize("n", [128, 256])
def test_cube_split_index_maps(n):
    """The two kernels' maps, replayed: every LDS slot and output element is
    written exactly once, and pass B after pass A is the 3D DFT."""
    rng = np.random.default_rng(11 + n)
    x = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    got = emulate_cube_b(emulate_cube_a(x, n), n)
    ref = np.fft.fftn(x)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_cube256_pass_b_block_mapping():
    """256^3 pass B, blockIdx -> (block, k, column group) with k fastest: every
    tile of two blocks once, and the 512 workgroups in flight together (two per
    CU) cover all 64 rows k (and 8 of the 16 column groups) of the first block."""
    b = np.arange(2 * 1024)
    k, xg, blk = b & 63, (b >> 6) & 15, b >> 10
    assert np.array_equal(np.sort((blk * 64 + k) * 16 + xg), b)
    assert len(np.unique(k[:512])) == 64 and len(np.unique(xg[:512])) == 8
    assert len(np.unique(xg[:1024])) == 16



# ---- 512^3 blocks: the y-split pair with y = a + 8 b (fft_cube512.hip) --------------------

def test_c512_split_formula():
    """The factorisation both kernels implement, on 512-point lines:
    X[k + 64 d] = sum_a W8^{a d} W512^{a k} T_a[k], T_a[k] = sum_b x[a + 8 b] W64^{b k}."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal(512) + 1j * rng.standard_normal(512)
    T = np.fft.fft(x.reshape(64, 8).T, axis=1)  # [a, k]
    a, k = np.arange(8)[:, None], np.arange(64)[None, :]
    X = np.fft.fft(T * np.exp(-2j * np.pi * a * k / 512), axis=0)  # [d, k]
    assert np.allclose(X.reshape(-1), np.fft.fft(x), atol=1e-10)  # index k + 64 d


def emulate_c512_a(plane, res, pidx=0):
    """Pass A on one 512 x 512 plane (index pidx) for residue res: thread
    (c, tau) loads points tau + 16 r of row y = res + 8 b, b = (c + 8 (pidx mod
    8)) mod 64, and does the radix-32 stage; round R writes its sub-transforms
    16 R .. 16 R + 15 to LDS [(c 16 + k) 17 + tau] and thread (c, tau') runs
    radix 16 over tau for k1 = 16 R + tau' (ending with kx = k1 + 32 m); column
    half R puts point (row b, column tau + 16 m) at LDS [b 257 + tau + 16 m] and
    column thread (cc, tq) stores point k = tq + 4 r of column kx = 16 R +
    (cc & 15) + 32 (cc >> 4) to row 64 res + k.  Returns the 64 output rows
    (NaN where not written)."""
    tid = np.arange(1024)
    c, tau = tid >> 4, tid & 15
    b = (c + 8 * (pidx & 7)) & 63
    r = np.arange(32)
    v = plane[(res + 8 * b)[:, None], tau[:, None] + 16 * r[None, :]]  # [tid, r]
    Y = np.fft.fft(v, axis=1)  # Y_tau[k1]
    o = np.empty((2, 1024, 16), complex)
    for R in range(2):
        lds = np.full(64 * 16 * 17, np.nan + 0j)
        for k in range(16):
            idx = (c * 16 + k) * 17 + tau
            assert np.isnan(lds[idx]).all(), "x-round LDS slot written twice"
            lds[idx] = Y[:, 16 * R + k]
        y = lds[(c * 16 + tau)[:, None] * 17 + np.arange(16)[None, :]]  # [tid, t]
        assert not np.isnan(y).any(), "x-round LDS slot read before written"
        k1 = 16 * R + tau
        y = y * np.exp(-2j * np.pi * (k1[:, None] * np.arange(16)[None, :]) / 512)
        o[R] = np.fft.fft(y, axis=1)  # X[k1 + 32 m]
    out = np.full((64, 512), np.nan + 0j)
    cc, tq = tid & 255, tid >> 8
    for R in range(2):
        lds = np.full(64 * 257, np.nan + 0j)
        for m in range(16):
            idx = b * 257 + tau + 16 * m
            assert np.isnan(lds[idx]).all(), "column LDS slot written twice"
            lds[idx] = o[R][:, m]
        cols = lds[np.arange(64)[None, :] * 257 + np.arange(256)[:, None]]  # [col, b]
        assert not np.isnan(cols).any(), "column LDS slot left empty"
        colfft = np.fft.fft(cols, axis=1)  # [col, k]
        kk = tq[:, None] + 4 * np.arange(16)[None, :]
        kx = 16 * R + (cc & 15) + 32 * (cc >> 4)
        assert np.isnan(out[kk, kx[:, None]]).all(), "output written twice"
        out[kk, kx[:, None]] = colfft[cc[:, None], kk]
    return out


def test_c512_pass_a_maps():
    rng = np.random.default_rng(21)
    for pidx in (0, 5):  # row rotation 0 and 40
        plane = rng.standard_normal((512, 512)) + 1j * rng.standard_normal((512, 512))
        full = np.full((512, 512), np.nan + 0j)
        for res in range(8):
            full[64 * res:64 * res + 64] = emulate_c512_a(plane, res, pidx)
        assert not np.isnan(full).any()
        # row 64 a + k holds T_a[k] = sum_b FFT_x(x[a + 8 b]) W64^{b k}
        ref = np.fft.fft(np.fft.fft(plane, axis=1).reshape(64, 8, 512), axis=0)  # [k, a, kx]
        ref = ref.transpose(1, 0, 2).reshape(512, 512)
        assert np.linalg.norm(full - ref) <= 1e-12 * np.linalg.norm(ref)


def emulate_c512_b(tin, k):
    """Pass B for one tile: tin[z, a, x] is row k + 64 a, column x of the tile's
    8 columns in plane z.  Thread (x, zs) loads planes zs + 128 j of its column,
    applies W512^{a k}, the 8-point DFT over a, the 4-point DFT over j and
    W512^{zs kq}; half hh writes line 32 dd + 8 kq + x (d = 4 hh + dd), slot zs
    to LDS [zs 129 + line]; line thread (cz, tz) ends with point m = tz + 8 r,
    stored at plane kq + 4 m of row k + 64 d.  Returns out[kz, d, x]."""
    tid = np.arange(1024)
    x, zs = tid & 7, tid >> 3
    j = np.arange(4)
    t = tin[(zs[:, None] + 128 * j[None, :]), :, x[:, None]]  # [tid, j, a]
    t = t * np.exp(-2j * np.pi * np.arange(8) * k / 512)[None, None, :]
    t = np.fft.fft(t, axis=2)  # [tid, j, d]
    t = np.fft.fft(t, axis=1)  # [tid, kq, d]
    t = t * np.exp(-2j * np.pi * (zs[:, None] * np.arange(4)[None, :]) / 512)[:, :, None]
    out = np.full((512, 8, 8), np.nan + 0j)
    cz, tz = tid & 127, tid >> 7
    for hh in range(2):
        lds = np.full(128 * 129, np.nan + 0j)
        for dd in range(4):
            for kq in range(4):
                idx = zs * 129 + 32 * dd + 8 * kq + x
                assert np.isnan(lds[idx]).all(), "z LDS slot written twice"
                lds[idx] = t[:, kq, 4 * hh + dd]
        lines = lds[np.arange(128)[None, :] * 129 + np.arange(128)[:, None]]  # [line, zs]
        assert not np.isnan(lines).any(), "z LDS slot left empty"
        zf = np.fft.fft(lines, axis=1)  # [line, m]
        xx, kq, dd = cz & 7, (cz >> 3) & 3, cz >> 5
        m = tz[:, None] + 8 * np.arange(16)[None, :]
        kz = kq[:, None] + 4 * m
        sel = out[kz, (4 * hh + dd)[:, None], xx[:, None]]
        assert np.isnan(sel).all(), "output written twice"
        out[kz, (4 * hh + dd)[:, None], xx[:, None]] = zf[cz[:, None], m]
    return out


@pytest.mark.parametrize("k", [0, 37])
def test_c512_pass_b_maps(k):
    rng = np.random.default_rng(31 + k)
    tin = rng.standard_normal((512, 8, 8)) + 1j * rng.standard_normal((512, 8, 8))
    got = emulate_c512_b(tin, k)
    assert not np.isnan(got).any()
    # X[kz][k + 64 d] = sum_z W512^{z kz} sum_a W8^{a d} W512^{a k} T[z][k + 64 a]
    ref = np.fft.fft(np.fft.fft(tin * np.exp(-2j * np.pi * np.arange(8) * k / 512)[None, :, None], axis=1),
                     axis=0)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_c512_pass_b_block_mapping():
    """blockIdx -> (k, column group): every tile of a block once, the two
    8-column groups of a 128-byte segment on one XCD (blocks b and b + 8), and
    the first 256 workgroups (one per CU) spread over 16 rows k x 16 column
    groups."""
    b = np.arange(2 * 4096)
    P = (b >> 4) * 8 + (b & 7)
    k = (P & 15) | (((P >> 7) & 3) << 4)
    xg = 2 * (((P >> 4) & 7) | (((P >> 9) & 3) << 3)) + ((b >> 3) & 1)
    blk = P >> 11
    assert len(np.unique(k[:256])) == 16 and len(np.unique(xg[:256])) == 16
    key = (blk * 64 + k) * 64 + xg
    assert np.array_equal(np.sort(key), np.arange(2 * 4096))
    seg = (blk * 64 + k) * 32 + xg // 2
    for s in np.unique(seg)[:200]:
        members = b[seg == s]
        assert len(members) == 2 and members[1] - members[0] == 8 and members[0] % 8 == members[1] % 8


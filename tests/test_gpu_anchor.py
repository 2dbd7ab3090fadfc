"""The anchored ring products of the fp16x3 form (engine.RING_ANCHOR): every RING_ANCHOR-th step
computes W_A = Wc + 1/2 v t2_A and Spp_A = t2_A W_A with split operands and keeps them with t2_A's
ring image and W_A; the steps in between compute W = W_A + 1/2 v FP16(t2 - t2_A) (one partial
product; FP16, or FP8 with engine.RING_DELTA_FP8) and Spp = Spp_A + [t2_hi | FP16(t2 - t2_A)] [FP16(W - W_A) | FP16(W_A)]
(one product over the concatenated K, both W halves on W_A's per-row scales).

Pinned by: nl_split_delta16 against a torch statement of its arithmetic (bit for bit, with and
without the hi half); nl_convert_split_delta with given scales against the same difference
converted with those scales in torch (bit for bit); the FP8 conversions against torch's E4M3
conversion (bit for bit) and the FP8 product against FP32 of its decoded operands; the
anchored W and Spp against the full products at the same amplitudes, with errors that shrink in
proportion to the amplitudes' difference from the anchor (the property the fixed point relies on: at
convergence the difference, hence the error, vanishes); and whole solves whose converged energy
equals the unanchored one to 1e-8 (the L=5 A=28 clause itself is tests/test_gpu_split.py).
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import _lib, ccm, engine, ops  # noqa: E402
from nulattice_amd.planewave import planewave_device_hamiltonian  # noqa: E402


def P(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def split16(x):
    """[hi | lo] FP16 rows of an FP32 array (x = hi + lo to FP16 pairs)."""
    hi = x.half()
    lo = (x - hi.float()).half()
    return torch.cat([hi, lo], dim=1).contiguous()


@pytest.mark.parametrize("with_hi", [0, 1])
@pytest.mark.parametrize("rows,K,kp", [(37, 96, 128), (200, 824, 832), (5, 64, 64)])
def test_split_delta16_is_its_fp32_statement(rows, K, kp, with_hi):
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(rows, K, generator=g, device="cuda") * 300
    xa = x + torch.randn(rows, K, generator=g, device="cuda") * 0.01
    X, XA = split16(x), split16(xa)
    ldd = (1 + with_hi) * kp + 8                   # a pitch past the written columns
    D = torch.full((rows * ldd,), -1, dtype=torch.int16, device="cuda")
    assert _lib.load().nl_split_delta16(rows, K, 2 * K, P(X), P(XA), P(D), ldd, kp, with_hi,
                                        S()) == 0
    torch.cuda.synchronize()
    a = X[:, :K].float() + X[:, K:].float()
    b = XA[:, :K].float() + XA[:, K:].float()
    ref = (a - b).half().view(torch.int16)
    got = D.view(rows, ldd)
    off = kp if with_hi else 0
    assert torch.equal(got[:, off:off + K], ref)
    assert bool((got[:, off + K:off + kp] == 0).all())        # the pad to the K step is zero
    if with_hi:
        assert torch.equal(got[:, :K], X[:, :K].contiguous().view(torch.int16))
        assert bool((got[:, K:kp] == 0).all())
    assert bool((got[:, off + kp:] == -1).all())              # nothing past the written columns


def test_split_delta16_rejects_unaligned_shapes():
    X = torch.zeros(64, dtype=torch.int16, device="cuda")
    f = _lib.load().nl_split_delta16
    assert f(2, 12, 24, P(X), P(X), P(X), 16, 16, 0, S()) == -1    # K % 8
    assert f(2, 8, 12, P(X), P(X), P(X), 8, 8, 0, S()) == -1       # ldx < 2K
    assert f(2, 8, 16, P(X), P(X), P(X), 16, 16, 1, S()) == -1     # ldd < 2 kp with the hi half
    assert f(2, 16, 32, P(X), P(X), P(X), 16, 8, 0, S()) == -1     # kp < K
    assert f(2, 8, 16, P(X), P(X), P(X), 16, 8, 2, S()) == -1      # with_hi not 0 or 1


@pytest.mark.parametrize("p,h", [(6, 4), (4, 8), (9, 12)])
def test_convert_split_delta_with_given_scales(p, h):
    """FP32 rows (b,j) of W - W_A read through the ring's [c][k][b][j] strides, converted on
    given power-of-two inverse scales (no amax pass): FP16((W - W_A) / s_r) bit for bit, zero pad
    from p h to kp, the other half of each 2 kp row untouched (h < 8: one workgroup per row; h >=
    8: the transposing 64x64 tiles)."""
    ph = p * h
    kp = (ph + 63) // 64 * 64
    g = torch.Generator(device="cuda").manual_seed(5)
    W = torch.randn(p, h, p, h, generator=g, device="cuda")            # [c][k][b][j]
    WA = W + torch.randn(p, h, p, h, generator=g, device="cuda") * 1e-3
    iw = 2.0 ** torch.randint(-12, -6, (ph,), generator=g, device="cuda").float()
    out = torch.full((ph, 2 * kp), -1, dtype=torch.int16, device="cuda")
    s4 = (ctypes.c_int64 * 4)(h, 1, p * h * h, p * h)
    assert _lib.load().nl_convert_split_delta(
        ops.FP16 | (ops.D_F32 << 8), p, h, p, h, s4, P(W), P(WA), P(out), 2 * kp, kp, 1, 0,
        None, P(iw), ctypes.c_float(ops.lp_target(ops.FP16, False)), ctypes.c_double(1.0),
        S()) == 0
    torch.cuda.synchronize()
    rows = (W.double() - WA.double()).permute(2, 3, 0, 1).reshape(ph, ph)   # rows (b,j)
    ref = (rows / iw.double()[:, None]).float().half().view(torch.int16)
    assert torch.equal(out[:, :ph], ref)
    assert bool((out[:, ph:kp] == 0).all())
    assert bool((out[:, kp:] == -1).all())


def _fp8(t):
    """Decode E4M3 bytes (uint8) to float32."""
    return t.view(torch.float8_e4m3fn).float()


@pytest.mark.parametrize("rows,K,kp", [(37, 96, 128), (300, 1000, 1024)])
def test_rows8_conversions_are_their_torch_statement(rows, K, kp):
    """nl_convert_rows8 (FP64 rows) and nl_split_delta8 (difference of two split images): per-row
    power-of-two scales with the row maximum in (128, 256], E4M3 bytes equal to torch's
    round-to-nearest-even conversion of x s, zero pad from K to kp."""
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(rows, K, generator=g, dtype=torch.float64, device="cuda")
    x = x * torch.logspace(-6, 3, rows, dtype=torch.float64, device="cuda")[:, None]
    x[3] = 0.0                                                   # an all-zero row
    ldd = kp + 16
    lib = _lib.load()
    for mode in ("rows", "delta"):
        D = torch.full((rows, ldd), 7, dtype=torch.uint8, device="cuda")
        isc = torch.empty(rows, dtype=torch.float32, device="cuda")
        if mode == "rows":
            assert lib.nl_convert_rows8(rows, K, K, P(x), P(D), ldd, kp, P(isc), S()) == 0
            val = x.float()
        else:
            xf = x.float()
            X, XA = split16(xf), split16(0.5 * xf)
            assert lib.nl_split_delta8(rows, K, 2 * K, P(X), P(XA), P(D), ldd, kp, P(isc),
                                       S()) == 0
            val = (X[:, :K].float() + X[:, K:].float()) - (XA[:, :K].float() + XA[:, K:].float())
        torch.cuda.synchronize()
        sc = 1.0 / isc
        m, e = torch.frexp(sc)
        assert bool((m == 0.5).all())                            # powers of two
        amax = val.abs().amax(1)
        nz = amax > 0
        assert bool((amax[nz] * sc[nz] > 128).all() and (amax[nz] * sc[nz] <= 256).all())
        ref = (val * sc[:, None]).to(torch.float8_e4m3fn).view(torch.uint8)
        assert torch.equal(D[:, :K], ref), mode
        assert bool((D[:, K:kp] == 0).all()) and bool((D[:, kp:] == 7).all())


@pytest.mark.parametrize("M,N,K,splits", [(4096, 4096, 384, 1), (4000, 4100, 1280, 1),
                                          (2048, 2048, 2048, 4)])
def test_fp8_gemm_is_fp32_of_its_operands(M, N, K, splits):
    """nl_gemm_lp with FP8 operands: alpha ia[m] (A B^T) ib[n] + D against the FP32 product of
    the decoded E4M3 operands.  The matrix core's sum of the 128 products of one instruction is
    not a sequence of IEEE FP32 additions (measured: 3e-5 of the output's largest element at
    K = 384), so the bound is 1e-4 of that element: far below what a mismatched operand layout
    gives (errors of the elements' own size)."""
    g = torch.Generator(device="cuda").manual_seed(5)
    lib = _lib.load()
    ops_ = []
    for R in (M, N):
        x = torch.randn(R, K, generator=g, dtype=torch.float64, device="cuda")
        img = torch.empty(R, K, dtype=torch.uint8, device="cuda")
        isc = torch.empty(R, dtype=torch.float32, device="cuda")
        assert lib.nl_convert_rows8(R, K, K, P(x), P(img), K, K, P(isc), S()) == 0
        ops_.append((img, isc))
    (A, ia), (B, ib) = ops_
    Dd = torch.randn(M, N, generator=g, device="cuda")
    C = torch.empty(M, N, device="cuda")
    ws = torch.empty(splits * M * N if splits > 1 else 1, dtype=torch.float64, device="cuda")
    io = (ops.C_F32 | ops.D_F32) << 8
    assert lib.nl_gemm_lp(2 | io, M, N, K, ctypes.c_double(0.75), P(A), K, 0, P(ia), 0, P(B), K,
                          0, P(ib), 0, ctypes.c_double(1.0), P(Dd), P(C), N, 0, 1, splits,
                          P(ws) if splits > 1 else None, S()) == 0
    torch.cuda.synchronize()
    ref = 0.75 * (_fp8(A) * ia[:, None]) @ (_fp8(B) * ib[:, None]).T + Dd
    err = float((C - ref).abs().max())
    print(f"max error {err:.3e} of max |ref| {float(ref.abs().max()):.3e}")
    assert err <= 1e-4 * float(ref.abs().max()), err


def test_fp8_gemm_rejects_unsupported_shapes():
    X = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda")
    C = torch.zeros(1 << 18, dtype=torch.float32, device="cuda")
    f = _lib.load().nl_gemm_lp
    io = (ops.C_F32 | ops.D_F32) << 8
    args = lambda M, N, K: (2 | io, M, N, K, ctypes.c_double(1.0), P(X), K, 0, None, 0, P(X), K,
                            0, None, 0, ctypes.c_double(0.0), None, P(C), N, 0, 1, 1, None, S())
    assert f(*args(4096, 4096, 96)) == -1          # K not a multiple of 128
    assert f(*args(128, 4096, 128)) == -1          # M below one tile
    assert f(*args(512, 512, 128)) == -1           # too few tiles to fill the GPU


def _w_after(H, B, t1, t2, prec):
    """W and Spp (FP32) after the T1 and T2 parts of a step at (t1, t2), with the split images
    made as engine.update makes them."""
    with ops.lp_reuse(t2):
        dt = engine._fast_split(H, prec)
        engine._t2_split_images(H, t2, dt, B)
        engine.t1_part(H, t1, t2, B, prec)
        engine.t2_part(H, t1, t2, B, prec)
        B.t2_split = B.t2_pairs = B.N56r = B.t2_ring = None
    return B.acc32().W.clone(), B.acc32().Spp.clone()


def test_anchored_w_error_shrinks_with_the_difference():
    """Plane-wave L=3 A=28: anchor at t2_A, then W and Spp at t2_A + eps d for eps = 1e-1, 1e-2,
    1e-4, against the full split products at the same amplitudes.  Each error stays below 2^-9
    (W) / 2^-8 (Spp: two roundings) of the difference's own contribution plus the FP32 storage
    floor (16 ulp of the largest element), 2^-3 / 2^-3 with the FP8 W product (E4M3's 2^-4 on
    both of its operands), and falls with eps about in proportion until it reaches that
    floor."""
    _, _, H = planewave_device_hamiltonian(3, 28, -4.0, 1.0, device="cuda")
    prec = ccm._PREC["fp16x3"]
    p, h = H.p, H.h
    B = engine.Buffers(p, h, H.device)
    t1 = torch.zeros(p, h, dtype=torch.float64, device="cuda")
    t2 = torch.zeros(p, p, h, h, dtype=torch.float64, device="cuda")
    engine.init_amplitudes(H, B, t1, t2, 0.0)
    g = torch.Generator(device="cuda").manual_seed(7)
    d = torch.randn(p, p, h, h, generator=g, dtype=torch.float64, device="cuda")
    d = d - d.transpose(0, 1)
    d = (d - d.transpose(2, 3)) * float(t2.abs().max())
    anchor = engine.RING_ANCHOR
    try:
        engine.RING_ANCHOR = 5
        B.ring_step = 0
        WA, SA = _w_after(H, B, t1, t2, prec)                      # the anchor step (full)
        assert B.ring_step == 1 and B.ring_s_ok
        errs, errs_s = [], []
        for eps in (1e-1, 1e-2, 1e-4):
            tt = t2 + eps * d
            B.ring_step = 1
            Wan, San = _w_after(H, B, t1, tt, prec)                # anchored
            assert B.ring_step == 2 and B.ring_mode == "delta"
            engine.RING_ANCHOR = 0
            Bf = engine.Buffers(p, h, H.device)
            Wfull, Sfull = _w_after(H, Bf, t1, tt, prec)           # full split products
            engine.RING_ANCHOR = 5
            f8 = engine.RING_DELTA_FP8 and p * h <= engine._FP8_MAXK
            rw, rs = (2.0 ** -3, 2.0 ** -3) if f8 else (2.0 ** -9, 2.0 ** -8)
            for X, XA, Xan, Xfull, rel, out in (("W", WA, Wan, Wfull, rw, errs),
                                                ("Spp", SA, San, Sfull, rs, errs_s)):
                dX = float((Xfull - XA).abs().max())               # the difference's part
                err = float((Xan - Xfull).abs().max())
                floor = 16 * 2.0 ** -24 * float(Xfull.abs().max())
                out.append(err)
                print(f"eps {eps:g} {X}: |X - X_A| {dX:.3e}, anchored error {err:.3e}, "
                      f"floor {floor:.1e}")
                assert err <= rel * dX + floor
        for e in (errs, errs_s):
            assert e[1] < 0.2 * e[0]                               # ~1/10 per decade of eps
    finally:
        engine.RING_ANCHOR = anchor


@pytest.mark.parametrize("fp8", [False, True])
@pytest.mark.parametrize("period", [2, 5])
def test_anchored_solve_converges_to_the_unanchored_energy(period, fp8, monkeypatch):
    """Plane-wave L=3 A=28: the fp16x3 fixed point with the anchored ring products equals the
    unanchored one to 1e-8 relative (the split products' fixed point itself moves by ~1e-9 with
    the convergence threshold and step count: their rounding noise)."""
    monkeypatch.setattr(engine, "RING_DELTA_FP8", fp8)
    _, _, H = planewave_device_hamiltonian(3, 28, -4.0, 1.0, device="cuda")
    es = {}
    anchor = engine.RING_ANCHOR
    try:
        for n in (0, period):
            engine.RING_ANCHOR = n
            e, _, _ = ccm.ccsd_solver(None, None, eps=1e-11, maxSteps=400, max_diis=0,
                                      hamiltonian=H, precision="fp16x3", refine=False)
            es[n] = e
    finally:
        engine.RING_ANCHOR = anchor
    assert abs(es[period] - es[0]) <= 1e-8 * abs(es[0]), es

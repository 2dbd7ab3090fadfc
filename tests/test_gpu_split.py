"""GPU parity of the split-operand precision form ("bf16x3") through the C ABI.

BF16 rounding is restated in NumPy (round to nearest even on the FP32 bit pattern), so the
hi/lo images are checked bit for bit; the split GEMM is checked on integer-valued operands,
whose products and FP32 sums are exact; the solver is checked against the FP64 fixed point at
the headline size.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import _lib, ccm, engine, ops  # noqa: E402
from nulattice_amd.planewave import PlaneWave, planewave_hamiltonian, shell_holes  # noqa: E402

L_ = _lib.load


def P(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def i64(*v):
    return (ctypes.c_int64 * len(v))(*v)


def bf16_bits(x32):
    """Round-to-nearest-even FP32 -> BF16 bit patterns (finite inputs)."""
    u = np.ascontiguousarray(x32, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split_ref(x, fp16=False):
    """hi = 16-bit(FP32(x)), lo = 16-bit(FP32(x - hi)) as the kernels define the split (BF16, or
    FP16 with numpy's round-to-nearest-even FP32 -> FP16)."""
    if fp16:
        hi = x.astype(np.float32).astype(np.float16)
        lo = (x - hi.astype(np.float64)).astype(np.float32).astype(np.float16)
        return hi.view(np.uint16), lo.view(np.uint16)
    hi = bf16_bits(x.astype(np.float32))
    lo = bf16_bits((x - bf16_value(hi)).astype(np.float32))
    return hi, lo


def u16(t):
    return t.detach().cpu().numpy().view(np.uint16)


def check_split(hi_k, lo_k, y, fp16):
    """The kernel's split (hi_k, lo_k bit patterns) of y.  BF16: bit for bit split_ref.  FP16: hi
    is FP16 nearest-even of y rounded via FP32 or directly (the two differ only where the FP32
    value ties between two FP16 values, about once in 8192 elements; both are nearest
    roundings), and lo is the FP16 of y - hi for the kernel's own hi, rounded the same two
    ways."""
    if not fp16:
        hi, lo = split_ref(y)
        np.testing.assert_array_equal(hi_k, hi)
        np.testing.assert_array_equal(lo_k, lo)
        return
    hi2, _ = split_ref(y, fp16=True)
    hi1 = y.astype(np.float16).view(np.uint16)
    assert np.all((hi_k == hi1) | (hi_k == hi2))
    hk = np.ascontiguousarray(hi_k).view(np.float16).astype(np.float64)
    lo2 = (y - hk).astype(np.float32).astype(np.float16).view(np.uint16)
    lo1 = (y - hk).astype(np.float16).view(np.uint16)
    assert np.all((lo_k == lo1) | (lo_k == lo2))


def test_bf16_reference_rounding():
    """The NumPy restatement rounds like torch's own FP32 -> BF16 cast (ties to even)."""
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100000) * 10.0 ** rng.integers(-20, 20, 100000)).astype(np.float32)
    x[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 1.00390625]   # exact ties
    t = torch.tensor(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(bf16_bits(x), t)


@pytest.mark.parametrize("transposed", [False, True])
def test_convert_split_blocks(transposed):
    """nl_convert_split writes [hi | lo | hi] rows (lomask 0b010) bit for bit, zero padding,
    and hi + lo reproduces x to 2^-16 relative; both the row kernel and the transposing tile
    kernel (source rows unit-stride)."""
    rng = np.random.default_rng(3)
    R, C, kpad = 70, 100, 104
    x = rng.standard_normal((R, C)) * 10.0 ** rng.integers(-6, 3, (R, C))
    src = torch.tensor(x.T.copy() if transposed else x, device="cuda")
    strides = (0, 1, 0, R) if transposed else (0, C, 0, 1)
    out = torch.full((R * 3 * kpad,), -1, dtype=torch.int16, device="cuda")
    rc = L_().nl_convert_split(0, 1, R, 1, C, i64(*strides), P(src), P(out), 3 * kpad, kpad, 3,
                               0b010, None, ctypes.c_float(0.0), ctypes.c_double(1.0), S())
    assert rc == 0
    torch.cuda.synchronize()
    o = u16(out).reshape(R, 3, kpad)
    hi, lo = split_ref(x)
    np.testing.assert_array_equal(o[:, 0, :C], hi)
    np.testing.assert_array_equal(o[:, 1, :C], lo)
    np.testing.assert_array_equal(o[:, 2, :C], hi)
    assert not o[:, :, C:].any()
    rec = bf16_value(hi) + bf16_value(lo)
    assert np.max(np.abs(rec - x) / np.abs(x)) < 2.0 ** -16


def test_convert_split_fp16_fixed_scale():
    """FP16 split of x * 2^12 (the amplitude images' fixed scale): bit for bit, and hi + lo
    reproduces the scaled value to 2^-21 for elements whose lo part is a normal FP16 number."""
    rng = np.random.default_rng(4)
    R, C = 64, 96
    x = rng.standard_normal((R, C)) * 10.0 ** rng.integers(-4, 0, (R, C))
    src = torch.tensor(x, device="cuda")
    out = torch.empty(R * 2 * C, dtype=torch.int16, device="cuda")
    assert L_().nl_convert_split(1, 1, R, 1, C, i64(0, C, 0, 1), P(src), P(out), 2 * C, C, 2,
                                 0b10, None, ctypes.c_float(0.0), ctypes.c_double(4096.0),
                                 S()) == 0
    torch.cuda.synchronize()
    o = u16(out).reshape(R, 2, C)
    check_split(o[:, 0], o[:, 1], x * 4096.0, True)
    hi, lo = o[:, 0].copy(), o[:, 1].copy()
    rec = hi.view(np.float16).astype(np.float64) + lo.view(np.float16).astype(np.float64)
    big = np.abs(x) > 1e-2
    assert np.max(np.abs(rec - 4096.0 * x)[big] / (4096.0 * np.abs(x[big]))) < 2.0 ** -21


def test_convert_exact_scales_make_contact_blocks_exact():
    """target < 0: rows whose nonzero elements are +-g (a contact interaction's blocks) convert
    to exactly +-1, the magnitude g carried by the FP32 inverse scale."""
    g = -4.1234567 / 125.0
    rng = np.random.default_rng(5)
    x = g * rng.integers(-1, 2, (40, 64)).astype(np.float64)
    x[3] = 0.0
    src = torch.tensor(x, device="cuda")
    out = torch.empty(40 * 64, dtype=torch.int16, device="cuda")
    isc = torch.empty(40, dtype=torch.float32, device="cuda")
    assert L_().nl_convert_rows(0, 1, 40, 1, 64, i64(0, 64, 0, 1), P(src), P(out), 64, P(isc),
                                ctypes.c_float(-1.0), S()) == 0
    torch.cuda.synchronize()
    v = bf16_value(u16(out)).reshape(40, 64)
    np.testing.assert_array_equal(v, np.sign(x))
    s = isc.cpu().numpy().astype(np.float64)
    assert s[3] == 1.0
    rows = [r for r in range(40) if r != 3]
    np.testing.assert_allclose(s[rows], abs(g), rtol=2 ** -23)


def _bf16(a):
    return torch.tensor(a, dtype=torch.float32, device="cuda").to(torch.bfloat16)


def _fp16(a):
    return torch.tensor(a, dtype=torch.float32, device="cuda").to(torch.float16)


@pytest.mark.parametrize("M,N,K", [(300, 513, 96), (4000, 4200, 600), (4100, 4000, 472),
                                   (4096, 4097, 1024)])
@pytest.mark.parametrize("parts", ["2x2", "1x2", "2x1"])
@pytest.mark.parametrize("dtype", [0, 1])
def test_split_gemm_exact(M, N, K, parts, dtype):
    """nl_gemm_lp_split = alpha (Ah Bh^T + Ah Bl^T + Al Bh^T) + beta D on integer operands
    (exact in BF16 and in FP32 sums): the 256x256 multi-stage kernel (large shapes, K not a
    multiple of 32 in the two-stage form) and the fallback of three products (small shape)."""
    rng = np.random.default_rng(M + N + K)
    Ah, Al, Bh, Bl = (rng.integers(-8, 9, s).astype(np.float64)
                      for s in ((M, K), (M, K), (N, K), (N, K)))
    use_al, use_bl = parts[0] == "2", parts[2] == "2"
    ref = Ah @ Bh.T + (Ah @ Bl.T if use_bl else 0) + (Al @ Bh.T if use_al else 0)
    D = rng.integers(-5, 6, (M, N)).astype(np.float64)
    cv = _bf16 if dtype == 0 else _fp16
    ta, tal, tb, tbl = cv(Ah), cv(Al), cv(Bh), cv(Bl)
    C = torch.full((M, N), float("nan"), dtype=torch.float64, device="cuda")
    Dd = torch.tensor(D, device="cuda")
    rc = L_().nl_gemm_lp_split(dtype, M, N, K, ctypes.c_double(0.5), P(ta), P(tal if use_al else None),
                               K, 0, None, 0, P(tb), P(tbl if use_bl else None), K, 0, None, 0,
                               ctypes.c_double(2.0), P(Dd), P(C), N, 0, 1, 1, None, S())
    assert rc == 0
    np.testing.assert_array_equal(C.cpu().numpy(), 0.5 * ref + 2.0 * D)


def test_split_gemm_strided_output_and_split_k():
    """The strided-output split GEMM with a split-K reduction writes C[(m1,m2),(n1,n2)] at
    two-run strides exactly."""
    rng = np.random.default_rng(11)
    m1, m2, n1, n2, K = 40, 28, 40, 28, 4096      # 5 x 5 tiles of 256, 16 K splits: 400 groups
    M, N = m1 * m2, n1 * n2
    Ah, Al, Bh, Bl = (rng.integers(-4, 5, s).astype(np.float64)
                      for s in ((M, K), (M, K), (N, K), (N, K)))
    ref = (Ah @ Bh.T + Ah @ Bl.T + Al @ Bh.T).reshape(m1, m2, n1, n2)
    out = torch.zeros((m1, n1, m2, n2), dtype=torch.float64, device="cuda")   # [m1][n1][m2][n2]
    ws = torch.empty(16 * M * N, dtype=torch.float64, device="cuda")
    sc = i64(n1 * m2 * n2, n2, m2 * n2, 1)
    ta, tal, tb, tbl = _bf16(Ah), _bf16(Al), _bf16(Bh), _bf16(Bl)   # kept alive for the launch
    rc = L_().nl_gemm_lp_strided_split(0, K, ctypes.c_double(1.0), P(ta), P(tal), K, None,
                                       P(tb), P(tbl), K, None, ctypes.c_double(0.0), None,
                                       P(out), i64(m1, m2, n1, n2), sc, 16, P(ws), S())
    assert rc == 0
    np.testing.assert_array_equal(out.cpu().numpy().transpose(0, 2, 1, 3), ref)


@pytest.mark.parametrize("dtype,scale", [(0, 1.0), (1, 4096.0)])
def test_tau_split(dtype, scale):
    """nl_tau_split: FP64 tau = t2 + t1 t1 - t1 t1 and the [hi | lo] rows of tau * scale bit for
    bit (BF16 unscaled, FP16 at the amplitude images' scale)."""
    rng = np.random.default_rng(7)
    p, h = 24, 6
    t1 = rng.standard_normal((p, h)) * 0.1
    t2 = rng.standard_normal((p, p, h, h)) * 0.01
    tau = t2 + np.einsum("ai,bj->abij", t1, t1) - np.einsum("aj,bi->abij", t1, t1)
    g1, g2 = torch.tensor(t1, device="cuda"), torch.tensor(t2, device="cuda")
    tau_d = torch.empty_like(g2)
    img = torch.empty(2 * p * p * h * h, dtype=torch.int16, device="cuda")
    assert L_().nl_tau_split(p, h, P(g1), P(g2), P(tau_d), P(img), dtype, ctypes.c_double(scale),
                             S()) == 0
    torch.cuda.synchronize()
    np.testing.assert_allclose(tau_d.cpu().numpy(), tau, rtol=1e-14, atol=1e-17)
    o = u16(img).reshape(p * p, 2, h * h)
    check_split(o[:, 0], o[:, 1], tau_d.cpu().numpy().reshape(p * p, h * h) * scale, dtype == 1)


@pytest.mark.parametrize("kind,scale", [(1, 1.0), (2, 4096.0)])
def test_assemble_ext_diagonals_and_split_image(kind, scale):
    """nl_assemble_t2_ext: H2 = S0 + pAB(Sab) + pIJ(Sij) + pIJ pAB(Spp) + (dpp_a + dpp_b + dhh_i +
    dhh_j) t2, the Jacobi step with the dressed denominators, and the [hi | lo] image of t2n."""
    rng = np.random.default_rng(9)
    p, h = 20, 6
    S0, Sab, Sij, t2, v = (rng.standard_normal((p, p, h, h)) * 0.1 for _ in range(5))
    Spp = rng.standard_normal((p, h, p, h)) * 0.1
    Xpp = rng.standard_normal((p, p)) + np.diag(8 + rng.random(p))
    Xhh = rng.standard_normal((h, h)) - np.diag(8 + rng.random(h))
    dpp, dhh = rng.standard_normal(p), rng.standard_normal(h)
    t1n = rng.standard_normal((p, h)) * 0.1
    pAB = lambda x: x - x.transpose(1, 0, 2, 3)
    pIJ = lambda x: x - x.transpose(0, 1, 3, 2)
    H = (S0 + pAB(Sab) + pIJ(Sij) + pIJ(pAB(Spp.transpose(0, 2, 1, 3)))
         + (dpp[:, None, None, None] + dpp[None, :, None, None] + dhh[None, None, :, None]
            + dhh[None, None, None, :]) * t2)
    dp, dh = np.diag(Xpp), np.diag(Xhh)
    den = -(dp[:, None, None, None] + dp[None, :, None, None] + dh[None, None, :, None]
            + dh[None, None, None, :])
    m = 0.3
    tn = m * t2 + (1 - m) * (t2 + H / den)
    d = lambda x: torch.tensor(x, device="cuda")
    g = [d(x) for x in (S0, Sab, Sij, Spp, Xpp, Xhh, dpp, dhh, t2, v, t1n)]
    t2n, H2 = torch.empty_like(g[0]), torch.empty_like(g[0])
    part = torch.empty(2 * int(L_().nl_partials(p * p * h * h)) + 8, dtype=torch.float64,
                       device="cuda")
    img = torch.empty(2 * p * p * h * h, dtype=torch.int16, device="cuda")
    rc = L_().nl_assemble_t2_ext(p, h, *[P(x) for x in g], ctypes.c_double(m), P(t2n), P(H2),
                                 P(part), P(img), kind, ctypes.c_double(scale), 0, S())
    assert rc == 0
    torch.cuda.synchronize()
    np.testing.assert_allclose(H2.cpu().numpy(), H, rtol=1e-13, atol=1e-13)
    tnd = t2n.cpu().numpy()
    np.testing.assert_allclose(tnd, tn, rtol=1e-12, atol=1e-14)
    o = u16(img).reshape(p, 2, -1)
    check_split(o[:, 0], o[:, 1], tnd.reshape(p, -1) * scale, kind == 2)


def _steps(H, prec, n):
    p, h = H.p, H.h
    B = engine.Buffers(p, h, H.device)
    t1 = torch.zeros(p, h, dtype=torch.float64, device="cuda")
    t2 = torch.zeros(p, p, h, h, dtype=torch.float64, device="cuda")
    engine.init_amplitudes(H, B, t1, t2, 0.0)
    t1n, t2n = torch.empty_like(t1), torch.empty_like(t2)
    img = engine.T2Image()
    es = []
    for _ in range(n):
        engine.update(H, t1, t2, B, t1n, t2n, 0.5, prec, img=img)
        es.append(float(B.E.item()))
        t1, t1n, t2, t2n = t1n, t1, t2n, t2
    return np.array(es), t2.clone()


@pytest.mark.parametrize("prec,tol", [("bf16x3", 1e-12), ("fp16x3", 1e-7)])
def test_split_images_equal_per_product_conversions(prec, tol):
    """Plane-wave L=3, A=28 (p=80, h=28): the shared split images of t2 (written by the
    assembly, permuted in 16 bits) and of tau give the same steps as converting every operand
    per product: BF16 to summation order (1e-12: the same roundings); FP16 to the split's
    precision (1e-7: the shared images use the fixed scale 2^12, the per-product conversions
    per-row scales, so the roundings differ), with the shared-image path's accumulators in FP64
    as the per-product path's; with its FP32 accumulators, with those and the hole-pair
    products, and with those and t2's image over the hole pairs (the default), to 1e-6.  The
    ring intermediate runs unanchored here (the anchored form's steps between anchors differ
    by the rounding of the amplitudes' difference: tests/test_gpu_anchor.py)."""
    E0, fock, v = planewave_hamiltonian(3, 28, -4.0, 1.0)
    H = engine.Hamiltonian(fock, v, device="cuda")
    anchor, engine.RING_ANCHOR = engine.RING_ANCHOR, 0
    fast = ccm._PREC[prec]
    assert engine._fast_split(H, fast) is not None
    slow = dict(fast, ring_W=prec)               # any sub-family key disables the shared images
    assert engine._fast_split(H, slow) is None
    es, ts = _steps(H, slow, 4)
    scale = float(ts.abs().max())
    for acc32, pairs, imgp, tl in ((False, False, False, tol), (True, False, False, 1e-6),
                                   (True, True, False, 1e-6), (True, True, True, 1e-6)):
        engine.ACC32, engine.HOLE_PAIRS, engine.IMG_PAIRS = acc32, pairs, imgp
        try:
            ef, tf = _steps(H, fast, 4)
        finally:
            engine.ACC32, engine.HOLE_PAIRS, engine.IMG_PAIRS = True, True, True
        np.testing.assert_allclose(ef, es, rtol=tl)
        assert float((tf - ts).abs().max()) <= tl * scale
    engine.RING_ANCHOR = anchor


def test_pair_image_equals_full_image_with_split_sab(monkeypatch):
    """Plane-wave L=3, A=28 with the 16-bit products' size floor lowered to 64 (so that the
    split Sab product runs at p = 80): t2's image over the hole pairs (ring rows expanded from
    it, A19 and Sab over the pairs) gives the steps of the full image to 1e-7 (A19's FP32
    accumulation order and the rounding of t2's antisymmetric partner elements differ; a wrong
    sign or slot moves them by far more)."""
    monkeypatch.setattr(ops, "LP_MIN_DIM", 64)
    E0, fock, v = planewave_hamiltonian(3, 28, -4.0, 1.0)
    H = engine.Hamiltonian(fock, v, device="cuda")
    fast = ccm._PREC["fp16x3"]
    assert engine._sab_pairs(H, engine._fast_split(H, fast))
    out = []
    for imgp in (False, True):
        monkeypatch.setattr(engine, "IMG_PAIRS", imgp)
        out.append(_steps(H, fast, 4))
    (ef, tf), (ep, tp) = out
    np.testing.assert_allclose(ep, ef, rtol=1e-7)
    assert float((tp - tf).abs().max()) <= 1e-7 * float(tf.abs().max())


def test_precision_form_L5_A28_meets_1e6():
    """Headline size (L=5, A=28, C0 from seed 0, D0=1): the fp16x3 and bf16x3 iterations converge
    to within 1e-6 relative of the FP64 energy (the symmetry-sector FP64 solve, itself equal to
    the dense FP64 solve to 1e-12), with no FP64 refinement."""
    from nulattice_amd.sectors import SectorHamiltonian
    C0 = float(np.random.default_rng(0).uniform(-4.5, -3.5))
    SH = SectorHamiltonian.from_planewave(PlaneWave(5, shell_holes(5, 28), C0, 1.0))
    e64, _, _ = ccm.ccsd_solver(None, None, eps=1e-13, maxSteps=400, max_diis=0, hamiltonian=SH)
    del SH
    E0, fock, v = planewave_hamiltonian(5, 28, C0, 1.0)
    H = engine.Hamiltonian(fock, v, device="cuda")
    for prec in ("fp16x3", "bf16x3"):
        info = {}
        e, _, _ = ccm.ccsd_solver(fock, v, eps=1e-10, maxSteps=120, max_diis=0, hamiltonian=H,
                                  precision=prec, refine=False, info=info)
        rel = abs(e - e64) / abs(e64)
        print(f"L=5 A=28 seed 0: {prec} |dE|/|E| = {rel:.3e} after {info['steps_lowprec']} "
              f"steps")
        assert info["converged"]
        assert rel <= 1e-6
    # the single-rounding forms for the record (their fixed points miss 1e-6)
    for prec in ("bf16", "fp16"):
        el, _, _ = ccm.ccsd_solver(fock, v, eps=1e-9, maxSteps=60, max_diis=0, hamiltonian=H,
                                   precision=prec, refine=False)
        print(f"  {prec}: |dE|/|E| = {abs(el - e64) / abs(e64):.3e}")


@pytest.mark.parametrize("nb,kb,ncol", [(3, 37, 40), (5, 236, 784), (1, 1, 4)])
def test_gather_rows16_split_equals_two_gathers(nb, kb, ncol):
    """nl_gather_rows16_split (both parts of [hi | lo] rows in one pass) writes exactly what two
    nl_gather_rows16_into calls write: the transposed gathered rows at part offsets 0 and kpad,
    zeros for padding rows (idx < 0) and for k in [kb, kpad)."""
    g = torch.Generator(device="cuda").manual_seed(nb * 1000 + kb)
    nsrc = 3 * kb + 5
    src = torch.randint(-30000, 30000, (nsrc, 2 * ncol), device="cuda", generator=g,
                        dtype=torch.int16)
    idx = torch.randint(-1, nsrc, (nb * kb,), device="cuda", generator=g)
    kpad = (kb + 31) // 32 * 32
    ldk = 2 * kpad
    a = torch.full((nb * ncol * ldk,), 7, dtype=torch.int16, device="cuda")
    b = a.clone()
    st = ops._stream()
    _lib.call("nl_gather_rows16_split", nb, kb, ncol, ops._p(idx), ops._p(src), 2 * ncol,
              ops._p(a), kpad, ldk, st)
    for q in range(2):
        _lib.call("nl_gather_rows16_into", nb, kb, ncol, ops._p(idx), ops._off16(src, q * ncol),
                  2 * ncol, ops._off16(b, q * kpad), kpad, ldk, st)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    # and the content: part q of column n of block bb, row k = src[idx[bb*kb + k], q*ncol + n]
    bb, n, q, k = nb - 1, ncol - 1, 1, kb - 1
    r = int(idx[bb * kb + k])
    want = 0 if r < 0 else int(src[r, q * ncol + n])
    assert int(a[(bb * ncol + n) * ldk + q * kpad + k]) == want

"""GPU parity of the individual kernels (through the C ABI) against float64 references."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import _lib, ops  # noqa: E402


def dev(x):
    return torch.tensor(np.asarray(x, np.float64), dtype=torch.float64, device="cuda")


def host(t):
    return t.detach().cpu().numpy()


SPECS = [
    # (spec, shapes of A, B) covering batch, broadcast batch, split-K, transposed operands
    ("ij,jk->ik", (37, 53), (53, 29)),
    ("ji,jk->ik", (53, 37), (53, 29)),
    ("ij,kj->ik", (37, 53), (29, 53)),
    ("ji,kj->ki", (53, 37), (29, 53)),
    ("bij,bjk->bik", (3, 17, 40), (3, 40, 19)),
    ("akci,ck->ai", (9, 4, 9, 4), (9, 4)),
    ("cdkl,cdij->klij", (9, 9, 4, 4), (9, 9, 4, 4)),
    ("abkl,klij->abij", (9, 9, 4, 4), (4, 4, 4, 4)),
    ("bkij,ak->abij", (9, 4, 4, 4), (9, 4)),
    ("idkl,adkj->ilaj", (4, 9, 4, 4), (9, 9, 4, 4)),
    ("cjkl,bl->ckbj", (9, 4, 4, 4), (9, 4)),
    ("li,ablj->abij", (4, 4), (9, 9, 4, 4)),
    ("dckl,dakl->ac", (9, 9, 4, 4), (9, 9, 4, 4)),
    ("cdlk,cdli->ki", (9, 9, 4, 4), (9, 9, 4, 4)),
    ("ij,jk->ik", (5, 4000), (4000, 3)),       # tall K: split-K path
    # skinny kernel: small N, small M (operand swap), huge K with small M and N, batched
    ("mk,kn->mn", (3000, 40), (40, 7)),
    ("km,kn->mn", (5000, 20), (5000, 30)),
    ("ab,bc->ac", (20, 3000), (3000, 900)),
    ("bmk,bkn->bmn", (4, 700, 33), (4, 33, 9)),
    ("mk,nk->mn", (1000, 50), (64, 50)),
    # small outputs: 4x4-blocked kernel with a wavefront split-K reduction (>= 64 splits), the
    # general small-output kernel (M, N in (32, 64]), strided small outputs
    ("ki,kj->ij", (20000, 7), (20000, 5)),
    ("ki,kj->ij", (3000, 40), (3000, 50)),
    ("abkl,abki->li", (40, 30, 6, 5), (40, 30, 6, 7)),
    # low K (<= 32) with large M and N: lanes along n, along m (unit-stride j of A's group),
    # and an operand layout that needs the copy fallback
    ("mk,kn->mn", (500, 16), (16, 700)),
    ("cjkl,bl->ckbj", (70, 6, 5, 20), (90, 20)),
    ("ak,bkij->aibj", (300, 20), (200, 20, 6, 5)),
    # no summed index (outer products of the adjoint sweep)
    ("ai,ck->acik", (9, 4), (9, 4)),
    ("dl,ai->dali", (9, 4), (9, 4)),
]


@pytest.mark.parametrize("spec,sa,sb", SPECS)
def test_contract_matches_einsum(spec, sa, sb):
    rng = np.random.default_rng(hash(spec) % 2 ** 32)
    A, B = rng.standard_normal(sa), rng.standard_normal(sb)
    ref = np.einsum(spec, A, B)
    C0 = rng.standard_normal(ref.shape)
    out = dev(C0)
    ops.contract(spec, dev(A), dev(B), out, alpha=0.7, beta=0.3)
    np.testing.assert_allclose(host(out), 0.7 * ref + 0.3 * C0, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("M,N,K,batch", [(1, 1, 1, 1), (130, 257, 65, 2), (128, 128, 16, 1),
                                         (7, 300, 1000, 3)])
@pytest.mark.parametrize("ak,bn", [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_gemm_f64_layouts(M, N, K, batch, ak, bn):
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((batch, M, K) if ak else (batch, K, M))
    B = rng.standard_normal((batch, K, N) if bn else (batch, N, K))
    Am = A if ak else A.transpose(0, 2, 1)
    Bm = B if bn else B.transpose(0, 2, 1)
    ref = Am @ Bm
    C = dev(np.zeros((batch, M, N)))
    Ad, Bd = dev(A), dev(B)      # keep the device buffers alive across the launch
    for splits in (1, 3):
        ws = torch.empty(batch * splits * M * N, dtype=torch.float64, device="cuda")
        _lib.call("nl_gemm_f64", ak, bn, M, N, K, ctypes.c_double(1.0), ops._p(Ad),
                  K if ak else M, A[0].size, ops._p(Bd), N if bn else K, B[0].size,
                  ctypes.c_double(0.0), ops._p(C), N, M * N, batch, splits, ops._p(ws),
                  ops._stream())
        np.testing.assert_allclose(host(C), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("spec,sa,sb", [
    # the four operand layouts, edge tiles in M and N, a partial last K step
    ("ik,kj->ij", (2000, 334), (334, 4000)),
    ("ki,kj->ij", (333, 2000), (333, 4000)),
    ("ik,jk->ij", (2000, 334), (4000, 334)),
    ("ki,jk->ij", (334, 2000), (4000, 334)),
    # two-run (strided) index groups: runs of 16 and of 14 (< the K step of 16), and a batch
    ("acik,ckbj->aibj", (180, 180, 16, 16), (180, 16, 180, 16)),
    ("cdkl,dblj->ckbj", (210, 210, 14, 14), (210, 210, 14, 14)),
    ("bik,bkj->bij", (8, 1024, 200), (8, 200, 1024)),
])
def test_contract_f64_large_tile_path_exact(spec, sa, sb):
    """Shapes that take the 256x128 FP64 tiles (LDS DMA, walkers along two-run k indices):
    small-integer operands make every product and sum exact, so the result must equal the
    float64 reference bit for bit, including alpha and the addend."""
    g = torch.Generator(device="cuda").manual_seed(len(spec) + sa[0])
    A = torch.randint(-3, 4, sa, device="cuda", generator=g).double()
    B = torch.randint(-3, 4, sb, device="cuda", generator=g).double()
    ref = torch.einsum(spec, A, B)
    D = torch.randint(-3, 4, ref.shape, device="cuda", generator=g).double()
    C = D.clone()
    ops.contract(spec, A, B, C, alpha=2.0, beta=0.5)
    assert float((C - (2.0 * ref + 0.5 * D)).abs().max()) == 0.0


@pytest.mark.parametrize("dt", [0, 1])
def test_gemm_lp_exact_on_small_integers(dt):
    """BF16/FP16 inputs that are small integers are exact; FP32 accumulation of products
    below 2^24 is exact, so the MFMA operand/result lane layout is checked bit for bit."""
    rng = np.random.default_rng(5 + dt)
    M, N, K = 150, 70, 200
    A = rng.integers(-4, 5, (M, K)).astype(np.float64)
    B = rng.integers(-4, 5, (N, K)).astype(np.float64)
    out = dev(np.zeros((M, N)))
    ops.contract("mk,nk->mn", dev(A), dev(B), out, beta=0.0, precision=("bf16", "fp16")[dt])
    np.testing.assert_array_equal(host(out), A @ B.T)


@pytest.mark.parametrize("dt", [0, 1])
@pytest.mark.parametrize("M,N,K,splits", [(4000, 4100, 4136, 1), (4096, 4352, 1024, 1),
                                          (1024, 1000, 4104, 16), (4000, 4100, 4160, 1),
                                          (1000, 1030, 2048, 16)])
def test_gemm_lp_large_tile_path_exact(dt, M, N, K, splits):
    """Shapes that take the 256x256-tile kernel (LDS DMA, swizzled LDS rows): edge tiles in M
    and N, a partial last K step and an odd step count (K = 4136 = 64.6 x 64), split-K, per-row
    scales and the addend, all exact on small-integer inputs."""
    g = torch.Generator(device="cuda").manual_seed(M + K + dt)
    tdt = (torch.bfloat16, torch.float16)[dt]
    A = torch.randint(-4, 5, (M, K), device="cuda", generator=g).to(tdt)
    B = torch.randint(-4, 5, (N, K), device="cuda", generator=g).to(tdt)
    ia = torch.full((M,), 0.5, device="cuda")
    ib = torch.full((N,), 2.0, device="cuda")
    D = torch.randn(M, N, dtype=torch.float64, device="cuda", generator=g)
    C = torch.empty(M, N, dtype=torch.float64, device="cuda")
    ws = torch.empty(splits * M * N, dtype=torch.float64, device="cuda") if splits > 1 else None
    _lib.call("nl_gemm_lp", dt, M, N, K, ctypes.c_double(3.0), ops._p(A), K, 0, ops._p(ia), 0,
              ops._p(B), K, 0, ops._p(ib), 0, ctypes.c_double(0.5), ops._p(D), ops._p(C), N, 0,
              1, splits, ops._p(ws), ops._stream())
    ref = 3.0 * (A.double() @ B.double().t()) + 0.5 * D
    assert float((C - ref).abs().max()) == 0.0


@pytest.mark.parametrize("prec,tol", [("bf16", 2e-2), ("fp16", 4e-3)])
def test_gemm_lp_scaled_accuracy(prec, tol):
    """Per-row scaling keeps tiny-magnitude rows (amplitude-like values ~1e-6) accurate."""
    rng = np.random.default_rng(9)
    A = rng.standard_normal((96, 320)) * np.logspace(-8, 0, 96)[:, None]
    B = rng.standard_normal((320, 80)) * 1e-3
    out = dev(np.zeros((96, 80)))
    ops.contract("mk,kn->mn", dev(A), dev(B), out, beta=0.0, precision=prec)
    ref = A @ B
    rel = np.abs(host(out) - ref).max(1) / np.abs(ref).max(1)
    assert rel.max() < tol


def test_permute_tau_segsum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 4, 5, 6))
    out = torch.empty((5, 3, 6, 4), dtype=torch.float64, device="cuda")
    ops.permute_into(dev(x).permute(2, 0, 3, 1), out, alpha=2.0)
    np.testing.assert_array_equal(host(out), 2.0 * x.transpose(2, 0, 3, 1))
    p, h = 5, 3
    t1, t2 = rng.standard_normal((p, h)), rng.standard_normal((p, p, h, h))
    tau = torch.empty((p, p, h, h), dtype=torch.float64, device="cuda")
    t1d, t2d = dev(t1), dev(t2)
    _lib.call("nl_tau", p, h, ops._p(t1d), ops._p(t2d), ops._p(tau), ops._stream())
    ref = t2 + np.einsum("ai,bj->abij", t1, t1) - np.einsum("aj,bi->abij", t1, t1)
    np.testing.assert_allclose(host(tau), ref, rtol=1e-14, atol=1e-14)
    # segsum: out[dst[s] + x*2] += 0.5 * sum val[e] * src[so[e] + x]
    src = rng.standard_normal(50)
    ptr = np.array([0, 2, 2, 5]); so = np.array([0, 10, 3, 7, 20]); val = rng.standard_normal(5)
    dst = np.array([0, 1, 30])
    o = np.zeros(40)
    oref = o.copy()
    for s in range(3):
        for x in range(4):
            oref[dst[s] + 2 * x] += 0.5 * sum(val[e] * src[so[e] + x] for e in range(ptr[s],
                                                                                    ptr[s + 1]))
    od = dev(o)
    L = lambda a: torch.tensor(a, dtype=torch.int64, device="cuda")
    keep = [L(ptr), L(so), dev(val), L(dst), dev(src)]   # alive until after the launch
    _lib.call("nl_segsum", 3, 4, ops._p(keep[0]), ops._p(keep[1]), ops._p(keep[2]),
              ops._p(keep[3]), 1, 2, ctypes.c_double(0.5), ops._p(keep[4]), ops._p(od), 0,
              ops._stream())
    torch.cuda.synchronize()
    np.testing.assert_allclose(host(od), oref, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("X", [1, 3, 28, 70])
def test_segsum_lane_packing(X):
    """Segments share a wavefront (G = 1 ... 64 lanes each, one X = 70 pass of two rounds);
    uneven segment lengths (0 to 7 entries) and strided sources and destinations."""
    rng = np.random.default_rng(X)
    S = 300
    lens = rng.integers(0, 8, S)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    E = int(ptr[-1])
    so = rng.integers(0, 50, E).astype(np.int64)
    val = rng.standard_normal(E)
    ss, ds = 2, 3
    dst = (np.arange(S) * X * ds).astype(np.int64)
    src = rng.standard_normal(50 + X * ss)
    out0 = rng.standard_normal(S * X * ds)
    ref = out0.copy()
    for s in range(S):
        for x in range(X):
            ref[dst[s] + x * ds] += 0.7 * sum(val[e] * src[so[e] + x * ss]
                                              for e in range(ptr[s], ptr[s + 1]))
    L = lambda a: torch.tensor(a, dtype=torch.int64, device="cuda")
    keep = [L(ptr), L(so), dev(val), L(dst), dev(src), dev(out0)]
    _lib.call("nl_segsum", S, X, ops._p(keep[0]), ops._p(keep[1]), ops._p(keep[2]),
              ops._p(keep[3]), ss, ds, ctypes.c_double(0.7), ops._p(keep[4]), ops._p(keep[5]), 0,
              ops._stream())
    np.testing.assert_allclose(host(keep[5]), ref, rtol=1e-13, atol=1e-13)


def test_gather_convert_matches_gather_then_cast():
    """Fused row gather + transposing BF16 cast of the packed-block products: bit-equal to
    torch's round-to-nearest-even cast of the gathered, transposed rows; idx < 0 rows and the
    k padding up to ldk are zero."""
    g = torch.Generator(device="cuda").manual_seed(3)
    nb, kb, ncol, ldk, nsrc = 5, 37, 70, 40, 90
    src = torch.randn(nsrc, ncol, dtype=torch.float64, device="cuda", generator=g)
    idx = torch.randint(-1, nsrc, (nb * kb,), device="cuda", generator=g)
    out = torch.full((nb, ncol, ldk), 7, dtype=torch.int16, device="cuda")
    _lib.call("nl_gather_convert", nb, kb, ncol, ops._p(idx), ops._p(src), ncol, ops._p(out), ldk,
              ops._stream())
    G = torch.where(idx[:, None] >= 0, src[idx.clamp(min=0)], torch.zeros_like(src[:1]))
    ref = torch.zeros(nb, ncol, ldk, dtype=torch.bfloat16, device="cuda")
    ref[:, :, :kb] = G.view(nb, kb, ncol).transpose(1, 2).float().to(torch.bfloat16)
    assert torch.equal(out, ref.view(torch.int16))


@pytest.mark.parametrize("M,N", [(40, 70), (300, 784)])
def test_gemm_lp_rows_scatter_accumulates_exactly(M, N):
    """Row-scatter LP GEMM: block rows land on C rows crow[b*M + m] (several batches, dropped
    rows), accumulated onto C's values; small integers keep every sum exact."""
    g = torch.Generator(device="cuda").manual_seed(M)
    nbat, K, nrow = 3, 64, 1000
    A = torch.randint(-3, 4, (nbat, M, K), device="cuda", generator=g).to(torch.bfloat16)
    B = torch.randint(-3, 4, (nbat, N, K), device="cuda", generator=g).to(torch.bfloat16)
    perm = torch.randperm(nrow, device="cuda", generator=g)[: nbat * M]
    crow = torch.where(torch.rand(nbat * M, device="cuda", generator=g) < 0.2,
                       torch.full_like(perm, -1), perm)
    C0 = torch.randint(-5, 6, (nrow, N), device="cuda", generator=g).double()
    C = C0.clone()
    _lib.call("nl_gemm_lp_rows", 0, M, N, K, ctypes.c_double(0.5), ops._p(A), K, M * K, None, 0,
              ops._p(B), K, N * K, None, 0, ops._p(C), N, ops._p(crow), nbat, ops._stream())
    R = 0.5 * torch.bmm(A.double(), B.double().transpose(1, 2).contiguous()).reshape(nbat * M, N)
    ref = C0.clone()
    keep = crow >= 0
    ref[crow[keep]] += R[keep]
    assert float((C - ref).abs().max()) == 0.0


def test_tau16_transpose16_gather16():
    """BF16 tau is the FP64 tau rounded (via FP32, as every FP64 -> BF16 cast here); the 16-bit
    transpose and row gather move bits unchanged."""
    g = torch.Generator(device="cuda").manual_seed(11)
    p, h = 37, 6
    t1 = torch.randn(p, h, dtype=torch.float64, device="cuda", generator=g)
    t2 = torch.randn(p, p, h, h, dtype=torch.float64, device="cuda", generator=g)
    tau = torch.empty_like(t2)
    tau16 = torch.empty(p * p * h * h, dtype=torch.int16, device="cuda")
    _lib.call("nl_tau", p, h, ops._p(t1), ops._p(t2), ops._p(tau), ops._stream())
    _lib.call("nl_tau16", p, h, ops._p(t1), ops._p(t2), ops._p(tau16), ops._stream())
    ref16 = tau.float().to(torch.bfloat16).view(torch.int16).reshape(-1)
    assert torch.equal(tau16, ref16)
    R, C, ldo = p * p, h * h, p * p + 7                       # rows padded with zeros
    out = torch.full((C, ldo), 5, dtype=torch.int16, device="cuda")
    _lib.call("nl_transpose16", R, C, ops._p(tau16), C, ops._p(out), ldo, ops._stream())
    assert torch.equal(out[:, :R], tau16.view(R, C).t())
    assert int(out[:, R:].abs().max()) == 0
    nb, kb, ldk = 4, 9, 16
    idx = torch.randint(-1, R, (nb * kb,), device="cuda", generator=g)
    Gt = torch.empty(nb, C, ldk, dtype=torch.int16, device="cuda")
    Gf = torch.empty(nb, C, ldk, dtype=torch.int16, device="cuda")
    _lib.call("nl_gather_rows16", nb, kb, C, ops._p(idx), ops._p(tau16), C, ops._p(Gt), ldk,
              ops._stream())
    _lib.call("nl_gather_convert", nb, kb, C, ops._p(idx), ops._p(tau), C, ops._p(Gf), ldk,
              ops._stream())
    assert torch.equal(Gt, Gf)


@pytest.mark.parametrize("h", [4, 6])
def test_permute16_and_batched_transpose16(h):
    """16-bit re-layouts move bits unchanged: the [a][c][i][k] -> [a][i][c][k] permute (4-element
    vectors when h % 4 == 0, single elements otherwise) and the per-a transpose
    [a][(d,k)][j] -> [a][j][(d,k)]."""
    g = torch.Generator(device="cuda").manual_seed(h)
    p = 13
    S = torch.randint(-30000, 30000, (p, p, h, h), device="cuda", generator=g).to(torch.int16)
    out = torch.empty(p, h, p, h, dtype=torch.int16, device="cuda")
    shp = (ctypes.c_int64 * 4)(p, h, p, h)
    istr = (ctypes.c_int64 * 4)(p * h * h, h, h * h, 1)
    _lib.call("nl_permute16", 4, shp, istr, ops._p(S), ops._p(out), ops._stream())
    assert torch.equal(out, S.permute(0, 2, 1, 3))
    T = torch.empty(p, h, p * h, dtype=torch.int16, device="cuda")
    _lib.call("nl_transpose16_batched", p, p * h, h, ops._p(S), h, p * h * h, ops._p(T), p * h,
              p * h * h, ops._stream())
    assert torch.equal(T, S.reshape(p, p * h, h).transpose(1, 2))


@pytest.mark.parametrize("h", [6, 8])
def test_fused_sab_bf16_product_is_exact_on_small_integers(h):
    """The P(ab) group as one BF16 GEMM over K = (k | d): t1 and N56 columns, then Mpp and the
    transposed BF16 image of t2 (column-range casts and transposes into padded operand rows; h=6
    takes the 2-byte transpose, h=8 the 8-byte one).  Small integers keep every product and FP32
    sum exact."""
    from types import SimpleNamespace
    from nulattice_amd import engine
    g = torch.Generator(device="cuda").manual_seed(17 + h)
    p = 40
    ri = lambda *s: torch.randint(-3, 4, s, device="cuda", generator=g).double()
    t1, Mpp, N56, t2 = ri(p, h), ri(p, p), ri(p, h, h, h), ri(p, p, h, h)
    B = SimpleNamespace(Mpp=Mpp, N56=N56,
                        t2_16=t2.float().to(torch.bfloat16).view(torch.int16).reshape(-1))
    Sab = torch.full((p, p, h, h), 9.0, dtype=torch.float64, device="cuda")
    engine._sab_bf16(SimpleNamespace(p=p, h=h), t1, B, Sab)
    ref = torch.einsum("ak,bkij->abij", t1, N56) + torch.einsum("ad,dbij->abij", Mpp, t2)
    assert float((Sab - ref).abs().max()) == 0.0


def test_error_codes():
    """Bad arguments come back as -1, never as a fault."""
    lib = _lib.load()
    assert lib.nl_gemm_f64(1, 1, -1, 1, 1, 1.0, None, 1, 0, None, 1, 0, 0.0, None, 1, 0, 1, 1,
                           None, None) == -1
    assert lib.nl_gemm_lp(0, 8, 8, 7, 1.0, None, 8, 0, None, 0, None, 8, 0, None, 0, 0.0, None,
                          None, 8, 0, 1, 1, None, None) == -1
    assert lib.nl_contract_f64(None, 1.0, None, None, None, None, 0.0, None, None, None, 1, None,
                               0, None) == -1


@pytest.mark.parametrize("k,n", [(1, 1000), (5, 1), (10, 300_001), (12, 4096 * 256 + 3),
                                 (14, 5000), (20, 777)])
def test_gram_matches_float64_products(k, n):
    """nl_gram_part / nl_gram_finalize (through ops.gram): every pair a <= b of up to 12 vectors in
    one pass, larger sets as pairs of groups; equals V V^T of torch FP64 to rounding, symmetric,
    and bitwise the same on a second call."""
    g = torch.Generator(device="cuda").manual_seed(k * 7 + n)
    vs = [torch.randn(n, generator=g, dtype=torch.float64, device="cuda") for _ in range(k)]
    G = ops.gram(vs)
    V = torch.stack(vs)
    np.testing.assert_allclose(G, host(V @ V.T), rtol=1e-12, atol=1e-13 * n)
    assert np.array_equal(G, G.T)
    assert np.array_equal(G, ops.gram(vs))


def test_gram_pair_order_on_exact_values():
    """Integer vectors (exact sums): G[a, b] = (a+1)(b+1) n + a b, so a mixed-up pair index fails."""
    n, k = 10_000, 12
    ar = torch.arange(n, device="cuda", dtype=torch.float64)
    vs = [(a + 1) * torch.ones(n, dtype=torch.float64, device="cuda") + a * (ar == 5).double()
          for a in range(k)]
    G = ops.gram(vs)
    ref = np.array([[(a + 1) * (b + 1) * n + a * (b + 1) + b * (a + 1) + a * b
                     for b in range(k)] for a in range(k)], dtype=np.float64)
    assert np.array_equal(G, ref)


@pytest.mark.parametrize("k,n", [(1, 5), (7, 100_003), (16, 4097), (17, 999), (33, 2049)])
def test_lincomb_matches_float64_sum(k, n):
    """nl_lincomb (through ops.lincomb): sum_k c_k x_k in one pass over up to 16 vectors, longer
    sums in chained chunks; equals the torch FP64 sum to rounding."""
    g = torch.Generator(device="cuda").manual_seed(k + 3 * n)
    vs = [torch.randn(n, generator=g, dtype=torch.float64, device="cuda") for _ in range(k)]
    c = np.linspace(-1.5, 2.0, k)
    out = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
    ops.lincomb(c, vs, out)
    ref = sum(float(ck) * v for ck, v in zip(c, vs))
    np.testing.assert_allclose(host(out), host(ref), rtol=1e-13, atol=1e-13 * k)


def test_gram_and_lincomb_argument_errors():
    lib = _lib.load()
    assert lib.nl_gram_part(10, 13, None, None, None) == -1
    assert lib.nl_gram_part(10, 0, None, None, None) == -1
    assert lib.nl_gram_finalize(0, 1, None, None, None) == -1
    assert lib.nl_lincomb(10, 17, None, None, None, None) == -1
    assert lib.nl_lincomb(10, 0, None, None, None, None) == -1


def test_copy_sub_is_a_copy_and_a_difference():
    """nl_copy_sub (a DIIS step's history entry and error vector in one pass): bitwise the copy
    and the FP64 difference."""
    g = torch.Generator(device="cuda").manual_seed(5)
    for n in (1, 1000, 1 << 20):
        x = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
        y = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
        hc, c = torch.empty_like(x), torch.empty_like(x)
        _lib.call("nl_copy_sub", n, ops._p(x), ops._p(y), ops._p(hc), ops._p(c), ops._stream())
        assert torch.equal(hc, x) and torch.equal(c, x - y)
    assert _lib.load().nl_copy_sub(-1, None, None, None, None, None) == -1


def test_concat_rows():
    """nl_concat_rows: [A | B] rows at pitch ldo, exactly; the columns beyond n1 + n2 untouched."""
    A = torch.randn(37, 9, dtype=torch.float64, device="cuda")
    Bm = torch.randn(37, 50, dtype=torch.float64, device="cuda")
    out = torch.full((37, 64), 7.0, dtype=torch.float64, device="cuda")
    _lib.call("nl_concat_rows", 37, 5, ops._p(A), 9, 50, ops._p(Bm), 50, ops._p(out), 64,
              ops._stream())
    ref = torch.cat([A[:, :5], Bm, torch.full((37, 9), 7.0, dtype=torch.float64,
                                               device="cuda")], 1)
    assert torch.equal(out, ref)
    assert _lib.load().nl_concat_rows(3, 5, None, 5, 2, None, 2, None, 7, None) == -1

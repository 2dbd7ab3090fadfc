"""FP32 outputs and addends of the product kernels (the io flags of the C ABI) and the assembly
over FP32 accumulators.

Every kernel forms its value in FP64 (FP32 inside the 16-bit MFMA products) and rounds it once
on store, so an FP32 output must equal, bit for bit, the FP64 output of the same launch rounded
to nearest even, when the FP64 launch starts from the same (FP32-representable) addend.  Each
kernel class of nl_contract_f64 (128x128 and 256x128 MFMA tiles, split-K with both reductions,
skinny, small-output, low-K), the 16-bit GEMMs (256x256 single-rounding, split, row scatter),
the segment sums and the split conversion of an FP32 source are checked that way; the assembly
with NL_ACC_F32 | NL_ADD_V against the FP64 assembly with v_pphh folded into S0.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import _lib, ops  # noqa: E402

L_ = _lib.load
F32, F64 = torch.float32, torch.float64


def P(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def i64(*v):
    return (ctypes.c_int64 * len(v))(*v)


def rnd(shape, seed, scale=1.0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(shape, generator=g, dtype=F64, device="cuda") * scale


CONTRACTIONS = [
    # (spec, A shape, B shape, kernel class)
    ("bkij,ak->abij", (96, 8, 8, 8), (96, 8), "low-K"),
    ("ilaj,bl->aibj", (8, 8, 96, 8), (96, 8), "low-K, strided C"),
    ("kj,abik->abij", (8, 8), (40, 40, 8, 8), "low-K"),
    ("bkci,cj->bkij", (64, 8, 120, 8), (120, 8), "skinny"),
    ("cikl,cakl->ai", (100, 8, 8, 8), (100, 90, 8, 8), "skinny, split-K"),
    ("cdlk,cdli->ki", (60, 60, 8, 8), (60, 60, 8, 8), "small output, split-K (wave reduce)"),
    ("ij,jk->ik", (300, 200), (200, 260), "128x128 tiles"),
    ("ij,jk->ik", (200, 3000), (3000, 300), "128x128 tiles, split-K"),
    ("ij,jk->ik", (4096, 256), (256, 2048), "256x128 tiles"),
]


@pytest.mark.parametrize("spec,sa,sb,what", CONTRACTIONS)
@pytest.mark.parametrize("mode", ["beta0", "rmw", "fp64_addend"])
def test_contract_fp32_output_is_rounded_fp64(spec, sa, sb, what, mode):
    """ops.contract into an FP32 output: bitwise the FP64 result of the same launch rounded to
    FP32 (beta = 0; beta = 1 reading the FP32 output; an FP64 addend with an FP32 output)."""
    A, B = rnd(sa, 1), rnd(sb, 2)
    la, lb, lo = spec.replace("->", ",").split(",")
    size = dict(zip(la, sa)) | dict(zip(lb, sb))
    oshape = tuple(size[c] for c in lo)
    c0 = rnd(oshape, 3).to(F32)                         # FP32-representable start values
    out32 = c0.clone()
    out64 = c0.to(F64)
    if mode == "beta0":
        ops.contract(spec, A, B, out32, alpha=0.75, beta=0.0)
        ops.contract(spec, A, B, out64, alpha=0.75, beta=0.0)
    elif mode == "rmw":
        ops.contract(spec, A, B, out32, alpha=-0.5, beta=1.0)
        ops.contract(spec, A, B, out64, alpha=-0.5, beta=1.0)
    else:
        add = rnd(oshape, 4)                            # FP64 values, not FP32-representable
        ops.contract(spec, A, B, out32, alpha=0.5, beta=1.0, addend=add)
        ops.contract(spec, A, B, out64, alpha=0.5, beta=1.0, addend=add)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out32.cpu().numpy(), out64.cpu().numpy().astype(np.float32))


def _split_rows(x, dtype):
    """[hi | lo] 16-bit rows of x (FP64 rows, FP16 at scale 1)."""
    R, K = x.shape
    out = torch.empty(R * 2 * K, dtype=torch.int16, device="cuda")
    assert L_().nl_convert_split(dtype, 1, R, 1, K, i64(0, K, 0, 1), P(x), P(out), 2 * K, K, 2,
                                 0b10, None, ctypes.c_float(0.0), ctypes.c_double(1.0), S()) == 0
    return out


def check_rounded(out32, out64, exact):
    """out32 is out64 rounded to FP32 (exact), or within the roundings of three FP32 passes."""
    a, b = out32.cpu().numpy(), out64.cpu().numpy()
    if exact:
        np.testing.assert_array_equal(a, b.astype(np.float32))
    else:
        np.testing.assert_allclose(a, b, rtol=4 * 2.0 ** -24, atol=1e-7 * np.abs(b).max())


@pytest.mark.parametrize("M,N,K", [(4096, 4096, 64), (300, 513, 96)])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_split_gemm_fp32_output(M, N, K, beta):
    """nl_gemm_lp_split with dtype | NL_C_F32 << 8 (| NL_D_F32 << 8 when it reads C back): the
    256x256 split kernel's FP32 output is its FP64 output rounded, bit for bit; the fallback for
    small shapes (three products, each adding to C) rounds after each product."""
    fp16 = 1
    A, B = rnd((M, K), 5, 0.1), rnd((N, K), 6, 0.1)
    As, Bs = _split_rows(A, fp16), _split_rows(B, fp16)
    c0 = rnd((M, N), 7).to(F32)
    out32, out64 = c0.clone(), c0.to(F64)
    lo = lambda t: ctypes.c_void_p(t.data_ptr() + 2 * K)
    io = ops.C_F32 | (ops.D_F32 if beta else 0)
    for out, dt in ((out32, fp16 | (io << 8)), (out64, fp16)):
        rc = L_().nl_gemm_lp_split(dt, M, N, K, ctypes.c_double(0.5), P(As), lo(As), 2 * K, 0,
                                   None, 0, P(Bs), lo(Bs), 2 * K, 0, None, 0,
                                   ctypes.c_double(beta), P(out) if beta else None, P(out), N, 0,
                                   1, 1, None, S())
        assert rc == 0
    torch.cuda.synchronize()
    check_rounded(out32, out64, M * N >= 256 * 256 * 256)


def test_big_gemm_fp32_output_fp64_addend():
    """nl_gemm_lp (256x256 single-rounding kernel, K a multiple of 64) with an FP32 output and an
    FP64 addend: the ring intermediate's form, bitwise the FP64 launch rounded."""
    M, N, K = 4096, 4096, 512
    A = rnd((M, K), 8, 0.1).to(torch.float16)
    B = rnd((N, K), 9, 0.1).to(torch.float16)
    add = rnd((M, N), 10)
    out32 = torch.empty((M, N), dtype=F32, device="cuda")
    out64 = torch.empty((M, N), dtype=F64, device="cuda")
    for out, dt in ((out32, 1 | (ops.C_F32 << 8)), (out64, 1)):
        assert L_().nl_gemm_lp(dt, M, N, K, ctypes.c_double(0.5), P(A), K, 0, None, 0, P(B), K,
                               0, None, 0, ctypes.c_double(1.0), P(add), P(out), N, 0, 1, 1, None,
                               S()) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out32.cpu().numpy(), out64.cpu().numpy().astype(np.float32))


@pytest.mark.parametrize("nb", [6, 256])
def test_rows_split_fp32_output(nb):
    """nl_gemm_lp_rows_split into FP32 rows (the pp ladder's scatter into S0): bitwise the FP64
    rows rounded, rows not named left untouched (three-product fallback for 6 blocks, the split
    kernel's row-scatter epilogue for 256)."""
    mb, K, N = 40, 64, 96
    V = rnd((nb * mb, K), 11, 0.1)
    G = rnd((nb * N, K), 12, 0.1)
    Vs, Gs = _split_rows(V, 1), _split_rows(G, 1)
    rng = np.random.default_rng(3)
    nrows = nb * mb + 100
    rows = rng.permutation(nrows)[:nb * mb].astype(np.int64)
    rows[5] = -1
    crow = torch.tensor(rows, device="cuda")
    c0 = rnd((nrows, N), 13).to(F32)
    out32, out64 = c0.clone(), c0.to(F64)
    lo = lambda t: ctypes.c_void_p(t.data_ptr() + 2 * K)
    for out, dt in ((out32, 1 | (ops.C_F32 << 8)), (out64, 1)):
        assert L_().nl_gemm_lp_rows_split(dt, mb, N, K, ctypes.c_double(0.25), P(Vs), lo(Vs),
                                          2 * K, mb * 2 * K, None, 0, P(Gs), lo(Gs), 2 * K,
                                          N * 2 * K, None, 0, P(out), N, P(crow), nb, S()) == 0
    torch.cuda.synchronize()
    check_rounded(out32, out64, nb >= 256)


def test_segsum_fp32_output():
    """nl_segsum with io = NL_C_F32: each output is read as FP32, the segment sum added in FP64
    and rounded once; bitwise the FP64 launch over the same start values, rounded."""
    rng = np.random.default_rng(17)
    S_, X = 200, 28
    lens = rng.integers(0, 9, S_)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    E = int(ptr[-1])
    so = rng.integers(0, 60, E).astype(np.int64)
    dst = (np.arange(S_) * X).astype(np.int64)
    L = lambda a: torch.tensor(a, dtype=torch.int64, device="cuda")
    keep = [L(ptr), L(so), rnd(E, 18), L(dst), rnd(60 + X, 19)]
    c0 = rnd(S_ * X, 20).to(F32)
    out32, out64 = c0.clone(), c0.to(F64)
    for out, io in ((out32, ops.C_F32), (out64, 0)):
        _lib.call("nl_segsum", S_, X, P(keep[0]), P(keep[1]), P(keep[2]), P(keep[3]), 1, 1,
                  ctypes.c_double(0.7), P(keep[4]), P(out), io, S())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out32.cpu().numpy(), out64.cpu().numpy().astype(np.float32))


@pytest.mark.parametrize("transposing", [False, True])
def test_convert_split_fp32_source(transposing):
    """nl_convert_split of an FP32 source (dtype | NL_D_F32 << 8): bitwise the conversion of the
    same values held in FP64 (row-wise and transposing-tile kernels, per-row scales)."""
    x32 = rnd((96, 80), 21).to(F32)
    x64 = x32.to(F64)
    R, C = (80, 96) if transposing else (96, 80)
    s4 = i64(0, 1, 0, 80) if transposing else i64(0, 80, 0, 1)
    outs, iscs = [], []
    for src, dt in ((x32, 1 | (ops.D_F32 << 8)), (x64, 1)):
        out = torch.empty(R * 2 * C, dtype=torch.int16, device="cuda")
        isc = torch.empty(R, dtype=torch.float32, device="cuda")
        assert L_().nl_convert_split(dt, 1, R, 1, C, s4, P(src), P(out), 2 * C, C, 2, 0b10,
                                     P(isc), ctypes.c_float(2.0 ** 14), ctypes.c_double(1.0),
                                     S()) == 0
        outs.append(out)
        iscs.append(isc)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(outs[0].cpu().numpy(), outs[1].cpu().numpy())
    np.testing.assert_array_equal(iscs[0].cpu().numpy(), iscs[1].cpu().numpy())


def test_assemble_fp32_accumulators_add_v():
    """nl_assemble_t2_ext with NL_ACC_F32 | NL_ADD_V (FP32 S0, Sab, Sij, Spp; v_pphh added by the
    assembly) equals the FP64 assembly with the same accumulator values and v_pphh in S0
    (only the order of the FP64 sums differs)."""
    p, h = 20, 6
    n = p * p * h * h
    acc32 = [rnd((p, p, h, h), 30 + k, 0.1).to(F32) for k in range(4)]
    v = rnd((p, p, h, h), 40, 0.05)
    t2, t1n = rnd((p, p, h, h), 41, 0.01), rnd((p, h), 42, 0.01)
    Xpp = rnd((p, p), 43) + 10.0 * torch.eye(p, dtype=F64, device="cuda")
    Xhh = rnd((h, h), 44) - 10.0 * torch.eye(h, dtype=F64, device="cuda")
    npart = 2 * int(L_().nl_partials(n)) + 8
    res = []
    for io in (1 | 2, 0):
        acc = acc32 if io else [acc32[0].to(F64) + v] + [a.to(F64) for a in acc32[1:]]
        t2n, H2 = torch.empty_like(t2), torch.empty_like(t2)
        part = torch.empty(npart, dtype=F64, device="cuda")
        assert L_().nl_assemble_t2_ext(p, h, *[P(a) for a in acc], P(Xpp), P(Xhh), None, None,
                                       P(t2), P(v), P(t1n), ctypes.c_double(0.3), P(t2n), P(H2),
                                       P(part), None, 0, ctypes.c_double(1.0), io, S()) == 0
        res.append((t2n, H2, part))
    torch.cuda.synchronize()
    for k in range(2):
        np.testing.assert_allclose(res[0][k].cpu().numpy(), res[1][k].cpu().numpy(), rtol=1e-13,
                                   atol=1e-15)
    assert L_().nl_assemble_t2_ext(p, h, *[P(a) for a in acc32], P(Xpp), P(Xhh), None, None,
                                   P(t2), None, None, ctypes.c_double(0.3), P(t2), None, P(part),
                                   None, 0, ctypes.c_double(1.0), 2, S()) == -1   # ADD_V, no v


def _pairs(h):
    return [(i, j) for i in range(h) for j in range(i + 1, h)]


@pytest.mark.parametrize("h,w", [(6, 1), (28, 1), (5, 7)])
def test_hole_pairs_reduce_expand(h, w):
    """nl_hole_pairs: reduce keeps the pairs i < j in row order (slots beyond P untouched);
    expand rebuilds the antisymmetric block (zero diagonal)."""
    R, npr = 3, h * (h - 1) // 2
    ldp = npr + 5
    x = np.random.default_rng(h).standard_normal((R, h, h, w))
    x = x - x.transpose(0, 2, 1, 3)
    xin = torch.tensor(x, device="cuda")
    red = torch.full((R, ldp, w), 9.0, dtype=F64, device="cuda")
    assert L_().nl_hole_pairs(R, h, ldp, w, P(xin), P(red), 0, S()) == 0
    back = torch.empty_like(xin)
    assert L_().nl_hole_pairs(R, h, ldp, w, P(red), P(back), 1, S()) == 0
    torch.cuda.synchronize()
    ref = np.stack([x[:, i, j] for i, j in _pairs(h)], 1)
    r = red.cpu().numpy()
    np.testing.assert_array_equal(r[:, :npr], ref)
    assert np.all(r[:, npr:] == 9.0)
    np.testing.assert_array_equal(back.cpu().numpy(), x)


@pytest.mark.parametrize("dtype,scale", [(0, 1.0), (1, 4096.0)])
def test_tau_split_pairs_equals_full_tau_columns(dtype, scale):
    """nl_tau_split_pairs: the [hi | lo] columns of the pairs i < j are those of nl_tau_split at
    (i, j), bit for bit, and the slots from P on are zero."""
    p, h = 12, 6
    npr, Pp = h * (h - 1) // 2, 32
    t1 = rnd((p, h), 50, 0.1)
    t2 = rnd((p, p, h, h), 51, 0.01)
    t2 = t2 - t2.transpose(2, 3)
    full = torch.empty(2 * p * p * h * h, dtype=torch.int16, device="cuda")
    red = torch.full((2 * p * p * Pp,), 7, dtype=torch.int16, device="cuda")
    assert L_().nl_tau_split(p, h, P(t1), P(t2), None, P(full), dtype, ctypes.c_double(scale),
                             S()) == 0
    assert L_().nl_tau_split_pairs(p, h, P(t1), P(t2), P(red), Pp, 1, dtype,
                                   ctypes.c_double(scale), S()) == 0
    torch.cuda.synchronize()
    f = full.cpu().numpy().reshape(p * p, 2, h, h)
    r = red.cpu().numpy().reshape(p * p, 2, Pp)
    ref = np.stack([f[:, :, i, j] for i, j in _pairs(h)], 2)
    np.testing.assert_array_equal(r[:, :, :npr], ref)
    assert np.all(r[:, :, npr:] == 0)


def test_assemble_s0_pairs_equals_full_s0():
    """nl_assemble_t2_ext with NL_S0_PAIRS: S0 given over the hole pairs (rows of P32 slots)
    gives the assembly of the antisymmetric full S0, bit for bit."""
    p, h = 14, 6
    npr, Pp = h * (h - 1) // 2, 32
    n = p * p * h * h
    S0 = rnd((p, p, h, h), 60, 0.1).to(F32)
    S0 = S0 - S0.transpose(2, 3)
    red = torch.zeros((p * p, Pp), dtype=F32, device="cuda")
    for q, (i, j) in enumerate(_pairs(h)):
        red[:, q] = S0[:, :, i, j].reshape(-1)
    acc = [rnd((p, p, h, h), 61 + k, 0.1).to(F32) for k in range(2)] + \
          [rnd((p, h, p, h), 63, 0.1).to(F32)]
    v = rnd((p, p, h, h), 64, 0.05)
    t2, t1n = rnd((p, p, h, h), 65, 0.01), rnd((p, h), 66, 0.01)
    Xpp = rnd((p, p), 67) + 10.0 * torch.eye(p, dtype=F64, device="cuda")
    Xhh = rnd((h, h), 68) - 10.0 * torch.eye(h, dtype=F64, device="cuda")
    npart = 2 * int(L_().nl_partials(n)) + 8
    out = []
    for s0, io in ((red, 1 | 2 | 4), (S0, 1 | 2)):
        t2n, H2 = torch.empty_like(t2), torch.empty_like(t2)
        part = torch.empty(npart, dtype=F64, device="cuda")
        assert L_().nl_assemble_t2_ext(p, h, P(s0), *[P(a) for a in acc], P(Xpp), P(Xhh), None,
                                       None, P(t2), P(v), P(t1n), ctypes.c_double(0.3), P(t2n),
                                       P(H2), P(part), None, 0, ctypes.c_double(1.0), io,
                                       S()) == 0
        out.append((t2n, H2))
    torch.cuda.synchronize()
    for k in range(2):
        assert torch.equal(out[0][k], out[1][k])


@pytest.mark.parametrize("dtype,scale", [(0, 1.0), (1, 4096.0)])
def test_tau_split_pairs_grouped_image_of_t2(dtype, scale):
    """nl_tau_split_pairs with grp = p and no t1: row a holds [hi((b,q)) | lo((b,q))], the same
    bits as the grp = 1 rows (a,b) of tau with t1 = 0 (tau = t2), slots from P on zero."""
    p, h = 12, 6
    npr, Pp = h * (h - 1) // 2, 32
    t2 = rnd((p, p, h, h), 52, 0.01)
    t2 = t2 - t2.transpose(2, 3)
    z1 = torch.zeros(p, h, dtype=F64, device="cuda")
    rows = torch.full((2 * p * p * Pp,), 7, dtype=torch.int16, device="cuda")
    grp = torch.full((2 * p * p * Pp,), 7, dtype=torch.int16, device="cuda")
    assert L_().nl_tau_split_pairs(p, h, P(z1), P(t2), P(rows), Pp, 1, dtype,
                                   ctypes.c_double(scale), S()) == 0
    assert L_().nl_tau_split_pairs(p, h, None, P(t2), P(grp), Pp, p, dtype,
                                   ctypes.c_double(scale), S()) == 0
    assert L_().nl_tau_split_pairs(p, h, None, P(t2), P(grp), Pp, 2, dtype,
                                   ctypes.c_double(scale), S()) == -1    # grp neither 1 nor p
    torch.cuda.synchronize()
    r = rows.cpu().numpy().reshape(p, p, 2, Pp).transpose(0, 2, 1, 3)
    g = grp.cpu().numpy().reshape(p, 2, p, Pp)
    np.testing.assert_array_equal(g, r)
    assert np.all(g[:, :, :, npr:] == 0) and np.any(g[:, 0, :, :npr] != 0)


@pytest.mark.parametrize("p,h", [(6, 5), (10, 28), (18, 28), (20, 6), (3, 2)])
def test_ring_from_pairs(p, h):
    """nl_ring_from_pairs: row (a,i), part, column (c,k) of the ring image is the pair slot of
    (min(i,k), max(i,k)) of row a, part, block c, with the sign bit flipped for i > k and zero
    for i = k (integer reference over every element)."""
    npr = h * (h - 1) // 2
    Pp = (npr + 31) // 32 * 32
    rng = np.random.default_rng(h)
    src = rng.integers(-32768, 32768, size=(p, 2, p, Pp), dtype=np.int64).astype(np.int16)
    out = torch.full((p * h * 2 * p * h,), 5, dtype=torch.int16, device="cuda")
    assert L_().nl_ring_from_pairs(p, h, Pp, P(torch.tensor(src, device="cuda")), P(out),
                                   S()) == 0
    assert L_().nl_ring_from_pairs(p, h, npr - 1, P(out), P(out), S()) == -1   # Pp < P
    torch.cuda.synchronize()
    q = np.zeros((h, h), dtype=np.int64)
    sgn = np.zeros((h, h), dtype=np.int64)
    for n_, (i, j) in enumerate(_pairs(h)):
        q[i, j] = q[j, i] = n_
        sgn[i, j], sgn[j, i] = 1, -1
    u = src.view(np.uint16)
    ref = np.zeros((p, h, 2, p, h), dtype=np.uint16)
    for i in range(h):
        for k in range(h):
            if i == k:
                continue
            v = u[:, :, :, q[i, k]]                       # [a][part][c]
            ref[:, i, :, :, k] = v if sgn[i, k] > 0 else v ^ 0x8000
    got = out.cpu().numpy().view(np.uint16).reshape(p, h, 2, p, h)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kernel", ["pair", "flat"])
def test_assemble_image_and_sab_pairs(kernel):
    """nl_assemble_t2_ext with NL_SAB_PAIRS | NL_IMG_PAIRS: Sab given over the hole pairs gives
    the assembly of the antisymmetric full Sab bit for bit, and the split image written over the
    pairs holds the full image's bits at every pair (i < j) and leaves the slots from P on
    untouched.  Both assembly kernels: the pair kernel (Sij and Spp given) and the flat one."""
    p, h = 14, 6
    npr, Pp = h * (h - 1) // 2, 32
    n = p * p * h * h
    S0 = rnd((p, p, h, h), 70, 0.1).to(F32)
    Sab = rnd((p, p, h, h), 71, 0.1).to(F32)
    Sab = Sab - Sab.transpose(2, 3)
    red = torch.zeros((p * p, Pp), dtype=F32, device="cuda")
    for q, (i, j) in enumerate(_pairs(h)):
        red[:, q] = Sab[:, :, i, j].reshape(-1)
    if kernel == "pair":
        Sij, Spp = rnd((p, p, h, h), 72, 0.1).to(F32), rnd((p, h, p, h), 73, 0.1).to(F32)
    else:
        Sij = Spp = None
    v = rnd((p, p, h, h), 74, 0.05)
    v = v - v.transpose(2, 3)
    t2, t1n = rnd((p, p, h, h), 75, 0.01), rnd((p, h), 76, 0.01)
    t2 = t2 - t2.transpose(2, 3)
    Xpp = rnd((p, p), 77) + 10.0 * torch.eye(p, dtype=F64, device="cuda")
    Xhh = rnd((h, h), 78) - 10.0 * torch.eye(h, dtype=F64, device="cuda")
    npart = 2 * int(L_().nl_partials(n)) + 8
    out = []
    for sab, nimg, io in ((Sab, 2 * n, 1 | 2), (red, 2 * p * p * Pp, 1 | 2 | 8 | 16)):
        t2n, H2 = torch.empty_like(t2), torch.empty_like(t2)
        part = torch.empty(npart, dtype=F64, device="cuda")
        img = torch.full((nimg,), 3, dtype=torch.int16, device="cuda")
        assert L_().nl_assemble_t2_ext(p, h, P(S0), P(sab), P(Sij), P(Spp), P(Xpp), P(Xhh),
                                       None, None, P(t2), P(v), P(t1n), ctypes.c_double(0.3),
                                       P(t2n), P(H2), P(part), P(img), 2,
                                       ctypes.c_double(4096.0), io, S()) == 0
        out.append((t2n, H2, img))
    # the pair image needs a split image kind
    assert L_().nl_assemble_t2_ext(p, h, P(S0), P(red), P(Sij), P(Spp), P(Xpp), P(Xhh), None,
                                   None, P(t2), P(v), P(t1n), ctypes.c_double(0.3), P(t2n),
                                   None, P(part), P(img), 0, ctypes.c_double(1.0), 1 | 2 | 8,
                                   S()) == -1
    torch.cuda.synchronize()
    for k in range(2):
        assert torch.equal(out[0][k], out[1][k])
    full = out[0][2].cpu().numpy().reshape(p, 2, p, h, h)
    pim = out[1][2].cpu().numpy().reshape(p, 2, p, Pp)
    ref = np.stack([full[:, :, :, i, j] for i, j in _pairs(h)], 3)
    np.testing.assert_array_equal(pim[:, :, :, :npr], ref)
    assert np.all(pim[:, :, :, npr:] == 3)

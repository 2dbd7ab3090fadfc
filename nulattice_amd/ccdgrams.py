"""NuLattice's CCSD diagram library (``NuLattice.CCM.ccDgrams``, ccDgrams.py:13-578) on MI355X.

Same function names, argument order and coefficients (the code's, where a docstring disagrees:
ccDgrams.py:307 vs :316, :361 vs :370).  Arguments may be NumPy arrays or torch tensors; every
result is a new float64 HIP tensor computed by libnlccsd.so kernels, and no input is modified.
Products of three or more tensors are evaluated as a chain of two-operand contractions
(``ops.contract``, one MFMA launch each), the antisymmetrisers by the permute kernel, and the COO
diagrams through the packed plans of ``plans.py`` (segment sums and pair-block GEMMs).

These are the reference's building blocks for callers that use them one by one (the unfused
``t1Iter`` / ``t2Iter`` of ccm.py); ``ccsd_solver`` runs the fused residual of engine.py.
"""
import ctypes

import torch

from . import _lib, engine, ops
from .ops import contract, permute_into
from .plans import PairBlocks, PpphPlans, coo_arrays

F64 = torch.float64


def _device():
    if not torch.cuda.is_available():
        raise _lib.NLError("no HIP device visible (the diagram kernels have no CPU fallback)")
    return torch.device("cuda")


def _d(x):
    """Device float64 view or copy of x (never written to)."""
    if isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == F64:
        return x.detach()
    return engine._dev(x, _device())


def _c(spec, alpha, *xs):
    """alpha * einsum(spec, *xs) as a new tensor, left to right over two-operand contractions.
    Every intermediate keeps the indices still needed by a later operand or the output."""
    lhs, out = spec.replace(" ", "").split("->")
    labs = lhs.split(",")
    xs = [_d(x) for x in xs]
    size = {}
    for lab, x in zip(labs, xs):
        size.update(zip(lab, x.shape))
    acc, la = xs[0], labs[0]
    for n in range(1, len(xs)):
        lb = labs[n]
        later = set(out).union(*labs[n + 1:])
        keep = [c for c in la + lb if c in later or (c in la) != (c in lb)]
        lo = out if n == len(xs) - 1 else "".join(dict.fromkeys(keep))
        res = torch.empty([size[c] for c in lo], dtype=F64, device=acc.device)
        contract(f"{la},{lb}->{lo}", acc, xs[n], res, alpha=alpha if n == len(xs) - 1 else 1.0,
                 beta=0.0)
        acc, la = res, lo
    return acc


def _perm_sum(x, terms):
    """sum_k coef_k * x.permute(perm_k) into a new contiguous tensor."""
    out = torch.empty(x.shape, dtype=F64, device=x.device)
    for n, (coef, perm) in enumerate(terms):
        permute_into(x.permute(*perm), out, coef, 0.0 if n == 0 else 1.0)
    return out


def pAB(x, scale=1.0):
    """x_abij - x_baij (ccDgrams.py:193-202)."""
    return _perm_sum(_d(x), [(scale, (0, 1, 2, 3)), (-scale, (1, 0, 2, 3))])


def pIJ(x, scale=1.0):
    """x_abij - x_abji (ccDgrams.py:204-213)."""
    return _perm_sum(_d(x), [(scale, (0, 1, 2, 3)), (-scale, (0, 1, 3, 2))])


def _pp(x, scale=1.0):
    """pIJ(pAB(x)) = pAB(pIJ(x))."""
    return _perm_sum(x, [(scale, (0, 1, 2, 3)), (-scale, (1, 0, 2, 3)), (-scale, (0, 1, 3, 2)),
                         (scale, (1, 0, 3, 2))])


# ------------------------------------------------------------------ COO scatter diagrams
def v_ppph_dgrams(v_ppph, t1, t2):
    """The seven v_ppph intermediates (ccDgrams.py:13-45) from COO <cd||ak> entries:
    r0[a,i] = -1/2 v t2[c,d,k,i], r1[a,d] = -v t1[c,k], r2[c,d,i,k] = v t1[a,i],
    r3[d,a] = v t1[c,k], r4[a,d,i,k] = v t1[c,i], r5[a,i,j,k] = v t2[c,d,i,j],
    r6[a,i,j,k] = v t1[c,i] t1[d,j] (summed over entries)."""
    t1, t2 = _d(t1), _d(t2)
    p, h = t1.shape
    dev = t1.device
    idx, val = coo_arrays(v_ppph, p, h, "ppph")
    z = lambda *s: torch.zeros(s, dtype=F64, device=dev)
    r0, r1, r3 = z(p, h), z(p, p), z(p, p)
    r2, r4 = z(p, p, h, h), z(p, p, h, h)
    r5, r6 = z(p, h, h, h), z(p, h, h, h)
    if val.size == 0:
        return r0, r1, r2, r3, r4, r5, r6
    pl = PpphPlans(idx, val, p, h, dev)
    t1c, t2c = ops.copy(t1), ops.copy(t2)
    T0 = z(p, h, h)
    engine._segsum(pl.ret0, h, t2c, T0, 1.0)
    contract("aki,k->ai", T0, torch.ones(h, dtype=F64, device=dev), r0, alpha=-0.5, beta=0.0)
    engine._segsum(pl.ret1, 1, t1c, r1, -1.0)
    permute_into(r1.permute(1, 0), r3, -1.0)
    S = z(p, h, p, h)                       # ret2 lands as [c][i][d][k]
    engine._segsum(pl.ret2, h, t1c, S, 1.0, 1, p * h)
    permute_into(S.permute(0, 2, 1, 3), r2)
    W = z(p, h, p, h)                       # ret4 lands as [d][k][a][i]
    engine._segsum(pl.ret4, h, t1c, W, 1.0)
    permute_into(W.permute(2, 0, 3, 1), r4)
    tt = _c("ci,dj->cdij", 1.0, t1c, t1c)
    for src, dst in ((t2c, r5), (tt, r6)):
        R = z(p, h, h, h)                   # [(a,k)][(i,j)]
        engine._block_gemm(pl.r56, src, R, 1.0, h * h, "ppph_r56")
        permute_into(R.permute(0, 2, 3, 1), dst)
    return r0, r1, r2, r3, r4, r5, r6


def v_pppp_dgrams(v_pppp, t1, t2):
    """pp ladder over COO <ab||cd> (ccDgrams.py:215-237): (sum v t2[c,d], sum v t1[c,i] t1[d,j])
    as (p,p,h,h) tensors, one pair-block GEMM each."""
    t1, t2 = _d(t1), _d(t2)
    p, h = t1.shape
    dev = t1.device
    idx, val = coo_arrays(v_pppp, p, h, "pppp")
    r1 = torch.zeros((p, p, h, h), dtype=F64, device=dev)
    r2 = torch.zeros((p, p, h, h), dtype=F64, device=dev)
    if val.size == 0:
        return r1, r2
    pb = PairBlocks(idx[:, 0] * p + idx[:, 1], idx[:, 2] * p + idx[:, 3], val, p * p, p * p, dev)
    engine._block_gemm(pb, ops.copy(t2), r1, 1.0, h * h, "pppp_lad")
    engine._block_gemm(pb, _c("ci,dj->cdij", 1.0, t1, t1), r2, 1.0, h * h, "pppp_lad")
    return r1, r2


# ------------------------------------------------------------ T1 diagrams (ccDgrams.py:48-191)
def dgram_akci_ck(v_phph, t1):          return _c("akci,ck->ai", -1.0, v_phph, t1)
def dgram_ck_acik(f_ph, t2):            return _c("ck,acik->ai", 1.0, f_ph, t2)
def dgram_cikl_cakl(v_phhh, t2):        return _c("cikl,cakl->ai", -0.5, v_phhh, t2)
def dgram_cdkl_ck_dali(v_pphh, t1, t2): return _c("cdkl,ck,dali->ai", 1.0, v_pphh, t1, t2)
def dgram_ck_ci(f_ph, t1):              return _c("ck,ci->ki", -0.5, f_ph, t1)
def dgram_ck_ak(f_ph, t1):              return _c("ck,ak->ac", -0.5, f_ph, t1)
def dgram_bijk_bj(v_phhh, t1):          return _c("bijk,bj->ki", -1.0, v_phhh, t1)
def dgram_cdlk_cdli(v_pphh, t2):        return _c("cdlk,cdli->ki", -0.5, v_pphh, t2)
def dgram_dckl_dakl(v_pphh, t2):        return _c("dckl,dakl->ac", -0.5, v_pphh, t2)
def dgram_cdlk_cl_di(v_pphh, t1):       return _c("cdlk,cl,di->ki", -0.5, v_pphh, t1, t1)
def dgram_cdkl_dk_al(v_pphh, t1):       return _c("cdkl,dk,al->ac", 0.5, v_pphh, t1, t1)


# ----------------------------------------------------------- T2 diagrams (ccDgrams.py:240-578)
def dgram_klij_abkl(v_hhhh, t2):        return _c("klij,abkl->abij", 0.5, v_hhhh, t2)
def dgram_bkcj_acik(v_phph, t2):        return _pp(_c("bkcj,acik->abij", 1.0, v_phph, t2), -1.0)
def dgram_bkij_ak(v_phhh, t1):          return pAB(_c("bkij,ak->abij", 1.0, v_phhh, t1))


def dgram_cdkl_acik_dblj(v_pphh, t2):
    # v[c,d,k,l] t2[d,b,l,j] first: the (c,k),(b,j) ring intermediate
    return _pp(_c("cdkl,dblj,acik->abij", 1.0, v_pphh, t2, t2), 0.5)


def dgram_cdkl_cdij_abkl(v_pphh, t2):
    return _c("cdkl,cdij,abkl->abij", 0.25, v_pphh, t2, t2)


def dgram_klij_ak_bl(v_hhhh, t1):
    return pAB(_c("klij,ak,bl->abij", 1.0, v_hhhh, t1, t1), 0.5)


def dgram_bkci_ak_cj(v_phph, t1):
    return _pp(_c("bkci,cj,ak->abij", 1.0, v_phph, t1, t1), -1.0)


def dgram_cikl_ck_ablj(v_phhh, t1, t2):
    return pIJ(_c("cikl,ck,ablj->abij", 1.0, v_phhh, t1, t2), -1.0)


def dgram_da_dbij(r3, t2):              return pAB(_c("da,dbij->abij", 1.0, r3, t2), -1.0)
def dgram_acik_bcjk(r4, t2):            return _pp(_c("acik,bcjk->abij", 1.0, r4, t2))


def dgram_cikl_al_bcjk(v_phhh, t1, t2):
    return _pp(_c("cikl,al,bcjk->abij", 1.0, v_phhh, t1, t2), -1.0)


def dgram_cjkl_ci_abkl(v_phhh, t1, t2):
    return pIJ(_c("cjkl,ci,abkl->abij", 1.0, v_phhh, t1, t2), 0.5)


def dgram_bijk_ak1(r5, t1):             return pAB(_c("bijk,ak->abij", 1.0, r5, t1), 0.5)
def dgram_bijk_ak2(r6, t1):             return _pp(_c("bijk,ak->abij", 1.0, r6, t1), 0.5)


def dgram_cjkl_ci_ak_bl(v_phhh, t1):
    return _pp(_c("cjkl,ci,ak,bl->abij", 1.0, v_phhh, t1, t1, t1), 0.5)


def dgram_cdkl_ci_dj_abkl(v_pphh, t1, t2):
    return pIJ(_c("cdkl,ci,dj,abkl->abij", 1.0, v_pphh, t1, t1, t2), 0.25)


def dgram_cdkl_ak_bl_cdij(v_pphh, t1, t2):
    return pAB(_c("cdkl,cdij,ak,bl->abij", 1.0, v_pphh, t2, t1, t1), 0.25)


def dgram_cdkl_ci_bl_adkj(v_pphh, t1, t2):
    return _pp(_c("cdkl,ci,adkj,bl->abij", 1.0, v_pphh, t1, t2, t1))


def dgram_cdkl_ci_ak_dj_bl(v_pphh, t1):
    return _pp(_c("cdkl,ci,dj,ak,bl->abij", 1.0, v_pphh, t1, t1, t1, t1), 0.25)


def dgram_cdkl_bdkl(v_pphh, t2):        return _c("cdkl,bdkl->bc", -0.5, v_pphh, t2)
def dgram_cdkl_cdjl(v_pphh, t2):        return _c("cdkl,cdjl->kj", -0.5, v_pphh, t2)
def dgram_ck_bk(f_ph, t1):              return _c("ck,bk->bc", -1.0, f_ph, t1)
def dgram_ck_cj(f_ph, t1):              return _c("ck,cj->kj", -1.0, f_ph, t1)
def dgram_cdlk_cl_dj(v_pphh, t1):       return _c("cdlk,cl,dj->kj", -1.0, v_pphh, t1, t1)
def dgram_cdlk_dk_bl(v_pphh, t1):       return _c("cdlk,dk,bl->bc", -1.0, v_pphh, t1, t1)


# ------------------------------------------------------------------ helpers for ccm.t1Iter/t2Iter
def accumulate(acc, *terms, scale=1.0):
    """acc += scale * sum(terms) in place (acc is a contiguous tensor this module created; the
    terms may be any arrays or tensors of acc's shape)."""
    for t in terms:
        permute_into(_d(t), acc, scale, 1.0)
    return acc


def start(x, scale=1.0):
    """A new contiguous device tensor scale * x (the accumulator a sum starts from)."""
    x = _d(x)
    out = torch.empty(x.shape, dtype=F64, device=x.device)
    return permute_into(x, out, scale, 0.0)


def jacobi_t1(t1, H1, X_pp, X_hh):
    """t1 + H1 / (-(X_pp[a,a] + X_hh[i,i])) (coupled_cluster.py:410-414)."""
    p, h = t1.shape
    out = torch.empty((p, h), dtype=F64, device=t1.device)
    _lib.call("nl_update_t1", p, h, ops._p(t1), ops._p(H1), ops._p(X_pp), ops._p(X_hh),
              ctypes.c_double(0.0), ops._p(out), ops._stream())
    return out


def jacobi_t2(t2, H2, X_pp, X_hh):
    """t2 + H2 / (-(X_aa + X_bb + X_ii + X_jj)) (coupled_cluster.py:539-545)."""
    p, _, h, _ = t2.shape
    out = torch.empty(t2.shape, dtype=F64, device=t2.device)
    part = torch.empty(2 * int(_lib.load().nl_partials(p * p * h * h)) + 8, dtype=F64,
                       device=t2.device)
    _lib.call("nl_assemble_t2", p, h, ops._p(H2), None, None, None, ops._p(X_pp), ops._p(X_hh),
              ops._p(t2), None, None, ctypes.c_double(0.0), ops._p(out), None, ops._p(part),
              ops._stream())
    return out

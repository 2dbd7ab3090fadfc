"""Hartree-Fock reference and the normal-ordered Hamiltonian in the Hartree-Fock basis.

The reference normal-orders its lattice Hamiltonian with respect to a determinant of lattice
states (``get_norm_ordered_ham``, coupled_cluster.py:752-832) and runs CCSD in that bare basis;
its Hartree-Fock solver (``solve_HF``, HF/hartree_fock.py:339-376) is standalone.  This module
connects the two: it solves HF as that solver does, forms the Fock matrix, the NO2B two-body
interaction and E_ref of the HF determinant in the lattice basis from the density matrix
rho = C_occ C_occ^T, and rotates them into the HF orbitals with four FP64 quarter
transformations on the GPU (``ops.contract``, the library's MFMA contraction), returning the
reference's block format with the first A HF orbitals as holes:

  F      = h + sum_tu (V[p,t,q,u] + 1/2 V3[p,t,q,u]) rho[t,u]
  V_eff  = V + V3,   V3[p,q,r,s] = sum_tu W[p,q,t,r,s,u] rho[t,u]
  E_ref  = tr(h rho) + 1/2 sum V rho rho + 1/6 sum W rho rho rho
  f' = C^T F C,   v'[a,b,c,d] = sum C[p,a] C[q,b] C[r,c] C[s,d] V_eff[p,q,r,s].

V and W are the fully antisymmetric tensors of the reference's operator lists: every two-body
entry [p,q,r,s,val] stands for its four signed images (hartree_fock.py:95-104) and every
three-body entry [a,b,c,d,e,f,val] for its 36 (three_body_utils.py:6-156).  With a permutation
matrix for C these are the reference's lattice-basis blocks (get_all_interactions,
get_fock_matrices, get_3NF_fock, get_3NF_tbme, get_ref_energy).  The HF Fock build here uses
the full antisymmetric W; HF/hartree_fock.py:129-207 forms the (bca),(fde) term of its
three-body contraction with the indices of the (bca),(def) term, so with a three-body force
the two HF solutions differ (without one they coincide).
"""
from itertools import permutations

import numpy as np
import torch

from . import _lib, ops
from .hamiltonian import _split

F64 = torch.float64


def _perm_sign(p):
    s = 1.0
    p = list(p)
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            if p[i] > p[j]:
                s = -s
    return s


_S2 = [(p, _perm_sign(p)) for p in permutations(range(2))]
_S3 = [(p, _perm_sign(p)) for p in permutations(range(3))]


def one_body_matrix(op1, nstat):
    """Dense h[p,q] of a one-body list [[p, q, val], ...], repeated (p, q) entries summed as the
    CC Fock build does (get_fock_matrices, coupled_cluster.py:21-97: at L=2 the +x and -x hops
    of the periodic kinetic list are the same pair).  get_1body_matrix (hartree_fock.py:41-67)
    keeps the last repeat instead; the two differ only when an entry repeats."""
    idx, val = _split(op1, 2)
    h = np.zeros((nstat, nstat))
    np.add.at(h, (idx[:, 0], idx[:, 1]), val)
    return h


def two_body_tensor(op2, nstat):
    """Dense antisymmetric V[p,q,r,s]: the four signed images of every entry, summed
    (hartree_fock.py:95-104)."""
    V = np.zeros((nstat,) * 4)
    if op2 is None or len(op2) == 0:
        return V
    idx, val = _split(op2, 4)
    for pk, sk in _S2:
        for pb, sb in _S2:
            np.add.at(V, (idx[:, pk[0]], idx[:, pk[1]], idx[:, 2 + pb[0]], idx[:, 2 + pb[1]]),
                      sk * sb * val)
    return V


def _three_body_images(op3):
    idx, val = _split(op3, 6)
    for pk, sk in _S3:
        for pb, sb in _S3:
            ket = idx[:, list(pk)]
            bra = idx[:, [3 + x for x in pb]]
            yield ket, bra, sk * sb * val


def three_body_no2b(op3, rho):
    """V3[p,q,r,s] = sum_tu W[p,q,t,r,s,u] rho[t,u] (dense, antisymmetric)."""
    n = rho.shape[0]
    V3 = np.zeros((n,) * 4)
    if op3 is None or len(op3) == 0:
        return V3
    for ket, bra, v in _three_body_images(op3):
        np.add.at(V3, (ket[:, 0], ket[:, 1], bra[:, 0], bra[:, 1]), v * rho[ket[:, 2], bra[:, 2]])
    return V3


def fock_pieces(op2, op3, rho):
    """(Gamma, Omega): Gamma[p,q] = sum_tu V[p,t,q,u] rho[t,u] and Omega[p,q] = sum W[p,t,t',q,u,u']
    rho[t,u] rho[t',u'], straight from the operator lists (the quantities of contract_2nf /
    contract_3nf, hartree_fock.py:70-213)."""
    n = rho.shape[0]
    G = np.zeros((n, n))
    if op2 is not None and len(op2):
        idx, val = _split(op2, 4)
        for pk, sk in _S2:
            for pb, sb in _S2:
                k = idx[:, list(pk)]
                b = idx[:, [2 + x for x in pb]]
                np.add.at(G, (k[:, 0], b[:, 0]), sk * sb * val * rho[k[:, 1], b[:, 1]])
    O = np.zeros((n, n))
    if op3 is not None and len(op3):
        for ket, bra, v in _three_body_images(op3):
            np.add.at(O, (ket[:, 0], bra[:, 0]),
                      v * rho[ket[:, 1], bra[:, 1]] * rho[ket[:, 2], bra[:, 2]])
    return G, O


def init_density(nstat, hole):
    """Projector onto the given lattice states (hartree_fock.py:237-256)."""
    d = np.zeros((nstat, nstat))
    d[list(hole), list(hole)] = 1.0
    return d


def hf_energy(op1, op2, op3, dens):
    """tr((h + 1/2 Gamma + 1/6 Omega) rho) (HF_energy, hartree_fock.py:259-280)."""
    h = one_body_matrix(op1, dens.shape[0])
    G, O = fock_pieces(op2, op3, dens)
    return float(np.sum((h + 0.5 * G + O / 6.0) * dens.T))


def solve_hf(op1, op2, op3, dens, mix=0.5, eps=1e-8, max_iter=100, verbose=False):
    """Self-consistent field by density mixing (solve_HF, hartree_fock.py:339-376): each step
    diagonalises h + Gamma + 1/2 Omega in FP64, occupies the lowest A = round(tr rho) orbitals and
    mixes the densities as mix * new + (1 - mix) * old; stops when sum |rho_new - rho| < eps after
    the second step.  Returns (energy, vecs, converged) with the HF orbitals as the columns of
    vecs (occupied first).  (HF_iter, hartree_fock.py:282-301, passes mix=0.5 on whatever the
    caller gave; here ``mix`` is used as documented, and the default is the same 0.5.)"""
    dens = np.array(dens, dtype=np.float64)
    nstat = dens.shape[0]
    npart = int(round(np.trace(dens)))
    h = one_body_matrix(op1, nstat)
    erg0, erg, vecs, converged = 0.0, 0.0, None, False
    for i in range(max_iter):
        G, O = fock_pieces(op2, op3, dens)
        erg = float(np.sum((h + 0.5 * G + O / 6.0) * dens.T))
        _, vecs = np.linalg.eigh(h + G + 0.5 * O)
        occ = vecs[:, :npart]
        new = occ @ occ.T
        new = mix * new + (1.0 - mix) * dens if mix != 1.0 else new
        diff = float(np.sum(np.abs(new - dens)))
        if verbose:
            print(i, "E=", erg, ", Delta E=", abs(erg - erg0), ", Delta rho =", diff)
        if diff < eps and i > 1:
            converged = True
            break
        erg0 = erg
        dens = new
    return erg, vecs, converged


def normal_order(op1, op2, op3, rho):
    """(E_ref, F, V_eff) of the determinant with density rho, in the lattice basis (host)."""
    n = rho.shape[0]
    h = one_body_matrix(op1, n)
    G, O = fock_pieces(op2, op3, rho)
    e_ref = float(np.sum((h + 0.5 * G + O / 6.0) * rho.T))
    V = two_body_tensor(op2, n) + three_body_no2b(op3, rho)
    return e_ref, h + G + 0.5 * O, V


def pair_support(op2, op3, rho, n):
    """The two-body interaction V_eff = V + V3 of ``normal_order`` on its support, without an
    n^4 array: (pairs (2, m) int64, Vr (m, m)) with V_eff[p,q,r,s] = Vr[i,j] for the ordered
    lattice-state pairs (p,q) = pairs[:, i], (r,s) = pairs[:, j], and zero for every (p,q,r,s)
    outside pairs x pairs.  The support is every ordered pair that occurs in a signed image of a
    two-body entry or as the first two indices of a three-body image (contact interactions: the
    pairs of states on one site)."""
    kk, kb, w = [], [], []
    if op2 is not None and len(op2):
        idx, val = _split(op2, 4)
        for pk, sk in _S2:
            for pb, sb in _S2:
                kk.append(idx[:, pk[0]] * n + idx[:, pk[1]])
                kb.append(idx[:, 2 + pb[0]] * n + idx[:, 2 + pb[1]])
                w.append(sk * sb * val)
    if op3 is not None and len(op3):
        for ket, bra, v in _three_body_images(op3):
            kk.append(ket[:, 0] * n + ket[:, 1])
            kb.append(bra[:, 0] * n + bra[:, 1])
            w.append(v * rho[ket[:, 2], bra[:, 2]])
    if not w:
        return np.zeros((2, 0), np.int64), np.zeros((0, 0))
    kk, kb, w = np.concatenate(kk), np.concatenate(kb), np.concatenate(w)
    keys = np.unique(np.concatenate([kk, kb]))
    Vr = np.zeros((keys.size, keys.size))
    np.add.at(Vr, (np.searchsorted(keys, kk), np.searchsorted(keys, kb)), w)
    return np.stack([keys // n, keys % n]).astype(np.int64), Vr


def hf_basis_engine(op1, op2, op3, C, A, device="cuda"):
    """(E_ref, fock_mats, engine.Hamiltonian) of the NO2B Hamiltonian normal-ordered with respect
    to the determinant of the first A columns of C, in the basis of C's columns (holes = columns
    0..A-1), built on the GPU without any n^4 array on the host: the interaction on its pair
    support (``pair_support``, m pairs) is rotated by two FP64 GEMMs per block,
      v'[(ab),(cd)] = P_ab^T (Vr P_cd),   P_ab[r, (a,b)] = C[p_r, a] C[q_r, b]  (nl_pair_outer),
    straight into the engine's layouts: the four dense blocks, v_pppp as one dense pair block
    and v_ppph as its device COO list (segment plans sorted on the device) with one dense R56
    block.  The Fock matrix, E_ref and the HF density are host arrays of size n^2."""
    import torch as _t
    from . import engine
    from .plans import PairBlocks, PpphPlans
    device = _t.device(device)
    if device.type != "cuda":
        raise _lib.NLError("hf_basis_engine: a HIP device is required (no CPU fallback)")
    _lib.load()
    C = np.asarray(C, dtype=np.float64)
    n = C.shape[0]
    if C.shape != (n, n) or not 0 < A < n:
        raise ValueError("C must be square and 0 < A < its size")
    rho = C[:, :A] @ C[:, :A].T
    h1 = one_body_matrix(op1, n)
    G, O = fock_pieces(op2, op3, rho)
    e_ref = float(np.sum((h1 + 0.5 * G + O / 6.0) * rho.T))
    Fr = C.T @ (h1 + G + 0.5 * O) @ C
    fock = [np.ascontiguousarray(x) for x in (Fr[A:, A:], Fr[A:, :A], Fr[:A, :A])]
    pairs, Vr = pair_support(op2, op3, rho, n)
    p, h, m = n - A, A, pairs.shape[1]
    # nl_pair_outer reads C row-major; torch.tensor would keep an F-ordered array's strides
    Cd = _t.tensor(np.ascontiguousarray(C), dtype=F64, device=device)
    Vd = _t.tensor(np.ascontiguousarray(Vr), dtype=F64, device=device)
    pr = _t.tensor(pairs[0], dtype=_t.int64, device=device)
    qr = _t.tensor(pairs[1], dtype=_t.int64, device=device)
    span = {"h": (0, h), "p": (A, p)}

    def images(xy):
        (c1, n1), (c2, n2) = span[xy[0]], span[xy[1]]
        P = _t.empty((m, n1 * n2), dtype=F64, device=device)
        _lib.call("nl_pair_outer", m, ops._p(pr), ops._p(qr), ops._p(Cd), n, c1, n1, c2, n2,
                  ops._p(P), ops._stream())
        return P

    P = {xy: images(xy) for xy in ("pp", "ph", "hh")}
    Q = {}
    for zw in ("pp", "ph", "hh"):
        Q[zw] = _t.empty_like(P[zw])
        ops.contract("rs,sx->rx", Vd, P[zw], Q[zw], beta=0.0)
    size = {"p": p, "h": h}

    def block(xy, zw):
        out = _t.empty((size[xy[0]] * size[xy[1]], size[zw[0]] * size[zw[1]]), dtype=F64,
                       device=device)
        ops.contract("rx,ry->xy", P[xy], Q[zw], out, beta=0.0)
        return out.view(size[xy[0]], size[xy[1]], size[zw[0]], size[zw[1]])

    dense = [block("pp", "hh"), block("ph", "ph"), block("ph", "hh"), block("hh", "hh")]
    v_ppph = block("pp", "ph")
    v_pppp = block("pp", "pp")
    del P, Q
    allp = np.arange(p * p)
    pppp = PairBlocks.from_components([allp], [allp],
                                      lambda rows, cols, nb, mb, kb: v_pppp.view(1, mb, kb),
                                      device)
    mask = v_ppph != 0
    idx = mask.nonzero().t().contiguous()
    val = v_ppph[mask]
    del mask
    r56 = PairBlocks.from_components(
        [np.arange(p * h)], [allp],
        lambda rows, cols, nb, mb, kb: ops.copy(v_ppph.permute(2, 3, 0, 1)).view(1, mb, kb),
        device)
    ppph = PpphPlans(idx, val, p, h, device, r56=r56)
    del v_ppph
    H = engine.Hamiltonian(fock, [pppp, ppph] + dense, device=device)
    return e_ref, fock, H


def _blocks(fock, V, A, sparse):
    h, p = slice(0, A), slice(A, None)
    f_pp, f_ph, f_hh = fock[p, p], fock[p, h], fock[h, h]
    v_pppp = V[p, p, p, p]
    v_ppph = V[p, p, p, h]
    out = [np.ascontiguousarray(x) for x in (v_pppp, v_ppph, V[p, p, h, h], V[p, h, p, h],
                                             V[p, h, h, h], V[h, h, h, h])]
    if sparse:
        for k in (0, 1):
            nz = np.nonzero(out[k])
            out[k] = (np.stack(nz, 1).astype(np.int64), out[k][nz])
    return [np.ascontiguousarray(x) for x in (f_pp, f_ph, f_hh)], out


def hf_basis_hamiltonian(op1, op2, op3, C, A, device="cuda", sparse=True):
    """(E_ref, fock_mats, two_body_int) of the NO2B Hamiltonian normal-ordered with respect to
    the determinant of the first A columns of C, in the basis of C's columns (holes = columns
    0..A-1 in order, particles = the rest): the reference's block format, ready for
    ccm.ccsd_solver.  The four quarter transformations run on the GPU in FP64."""
    device = torch.device(device)
    if device.type != "cuda":
        raise _lib.NLError("hf_basis_hamiltonian: a HIP device is required (no CPU fallback)")
    C = np.asarray(C, dtype=np.float64)
    n = C.shape[0]
    if C.shape != (n, n) or not 0 <= A <= n:
        raise ValueError("C must be square and 0 <= A <= its size")
    rho = C[:, :A] @ C[:, :A].T
    e_ref, F, V = normal_order(op1, op2, op3, rho)
    Cd = torch.tensor(C, dtype=F64, device=device)
    X = torch.tensor(V, dtype=F64, device=device)
    Y = torch.empty_like(X)
    # v'[a,b,c,d]: transform one index per pass, alternating the two buffers
    ops.contract("pa,pqrs->aqrs", Cd, X, Y, beta=0.0)
    ops.contract("qb,aqrs->abrs", Cd, Y, X, beta=0.0)
    ops.contract("rc,abrs->abcs", Cd, X, Y, beta=0.0)
    ops.contract("sd,abcs->abcd", Cd, Y, X, beta=0.0)
    Fd = torch.tensor(F, dtype=F64, device=device)
    T = torch.empty_like(Fd)
    ops.contract("pa,pq->aq", Cd, Fd, T, beta=0.0)
    ops.contract("aq,qb->ab", T, Cd, Fd, beta=0.0)
    fock, two = _blocks(Fd.cpu().numpy(), X.cpu().numpy(), A, sparse)
    return e_ref, fock, two

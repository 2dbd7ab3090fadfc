"""CPU-only checks of the package: the C ABI library exports every declared symbol, the
plane-wave generator is the Fourier transform of the position-space operators, and the
host-side plans reproduce the COO arithmetic."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from nulattice_amd import _lib
from nulattice_amd.planewave import PlaneWave, planewave_hamiltonian, shell_holes
from nulattice_amd.plans import PairBlocks, PpphPlans, coo_arrays
from oracle import ccsd_ref as cc
from oracle import hamiltonian_ref as ham_ref
from oracle import lattice_ref as lat
from tests.helpers import random_amplitudes, random_antisym_hamiltonian, to_coo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\s*(?:int|int64_t)\s+(nl_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnlccsd.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(_lib.SIGNATURES), "ctypes signatures and header disagree"


def _fourier_dense(L):
    """Position-space kinetic + contact operators transformed with the plane-wave unitary."""
    n = 4 * L ** 3
    T = np.zeros((n, n))
    for p, q, v in lat.Tkin(L):
        T[p, q] += v
    V = np.zeros((n, n, n, n))
    for p, q, r, s, v in lat.contacts(-3.0, -3.0, L):
        V[p, q, r, s] = v; V[q, p, r, s] = -v; V[p, q, s, r] = -v; V[q, p, s, r] = v
    x = np.array([[i, j, k] for i in range(L) for j in range(L) for k in range(L)])
    kv = x.copy()
    phase = np.exp(2j * np.pi * x @ kv.T / L) / L ** 1.5      # U[x, k]
    U = np.kron(phase, np.eye(4))                            # orbitals (site, sigma)
    Tk = U.conj().T @ T @ U
    Vk = np.einsum("xp,yq,xyzw,zr,ws->pqrs", U.conj(), U.conj(), V, U, U, optimize=True)
    return Tk, Vk


def test_planewave_is_fourier_transform_of_lattice_operators():
    L = 2
    Tk, Vk = _fourier_dense(L)
    pw = PlaneWave(L, shell_holes(L, 4), -3.0)
    n = 4 * L ** 3
    np.testing.assert_allclose(Tk.imag, 0, atol=1e-12)
    np.testing.assert_allclose(np.diag(Tk.real), np.repeat(pw.eps, 4), atol=1e-12)
    np.testing.assert_allclose(Tk - np.diag(np.diag(Tk)), 0, atol=1e-12)
    allo = np.arange(n)
    np.testing.assert_allclose(Vk.real, pw.block(allo, allo, allo, allo), atol=1e-12)
    np.testing.assert_allclose(Vk.imag, 0, atol=1e-12)


def test_planewave_normal_ordering_matches_reference_sorting():
    """Feeding the plane-wave contact list to the oracle's get_norm_ordered_ham (the reference's
    block sorting + Fock + E_ref) gives the generator's blocks and closed forms."""
    L = 2
    pw = PlaneWave(L, shell_holes(L, 4), -3.0)
    n = 4 * L ** 3
    allo = np.arange(n)
    Vfull = pw.block(allo, allo, allo, allo)
    lst = [[p, q, r, s, Vfull[p, q, r, s]] for p in range(n) for q in range(p + 1, n)
           for r in range(n) for s in range(r + 1, n) if Vfull[p, q, r, s] != 0]
    tk = [[q, q, pw.eps[q // 4]] for q in range(n)]
    holes_states = [[q // 4 // 4, (q // 4 // 2) % 2, (q // 4) % 2, (q % 4) // 2, q % 2]
                    for q in pw.holes]
    vac, fock, blocks = ham_ref.get_norm_ordered_ham(L, holes_states, tk, lst, None, sparse=False)
    E, fk, bl = pw.hamiltonian(sparse=False)
    assert abs(vac - E) < 1e-12
    assert abs(E - 6 * (-3.0) / L ** 3) < 1e-12        # E_ref(A=4) = 6 C0 / L^3, eps(0) = 0
    for a, b in zip(fock, fk):
        np.testing.assert_allclose(a, b, atol=1e-12)
    for a, b in zip(blocks, bl):
        np.testing.assert_allclose(a, b, atol=1e-12)


def test_planewave_three_body_no2b_matches_reference_normal_ordering():
    """D0 normal ordering: the momentum-space three-body contact list run through the oracle's
    get_3NF / get_3NF_fock / get_3NF_tbme / get_ref_energy (three_body_utils.py) equals the
    generator's closed forms (L=2, A=4)."""
    L, C0, D0 = 2, -3.0, 1.3
    pw = PlaneWave(L, shell_holes(L, 4), C0, D0)
    n = 4 * L ** 3
    kv = pw._kv[pw._orb_k]
    s = pw._orb_s
    trip = [(a, b, c) for a in range(n) for b in range(a + 1, n) for c in range(b + 1, n)
            if len({s[a], s[b], s[c]}) == 3]
    tk = {}
    for t in trip:
        K = tuple((kv[t[0]] + kv[t[1]] + kv[t[2]]) % L)
        tk.setdefault(K, []).append(t)
    hs = set(pw.holes)
    w3 = []
    for K, ts in tk.items():
        for x in ts:
            if not any(i in hs for i in x):
                continue
            for y in ts:
                if not any(i in hs for i in y):
                    continue
                # <x||y> = D0/L^6 * sum_perm sign [sigma(x) == perm sigma(y)]
                sx, sy = [s[i] for i in x], [s[i] for i in y]
                val = 0.0
                for perm, sg in (((0, 1, 2), 1), ((1, 0, 2), -1), ((2, 1, 0), -1),
                                 ((0, 2, 1), -1), ((1, 2, 0), 1), ((2, 0, 1), 1)):
                    if all(sx[m] == sy[perm[m]] for m in range(3)):
                        val += sg
                if val:
                    w3.append([x[0], x[1], x[2], y[0], y[1], y[2], val * D0 / L ** 6])
    allo = np.arange(n)
    pw2 = PlaneWave(L, shell_holes(L, 4), C0, 0.0)
    V2 = pw2.block(allo, allo, allo, allo)
    lst = [[p, q, r, t, V2[p, q, r, t]] for p in range(n) for q in range(p + 1, n)
           for r in range(n) for t in range(r + 1, n) if V2[p, q, r, t] != 0]
    tkin = [[q, q, pw.eps[q // 4]] for q in range(n)]
    holes_states = [[q // 16, (q // 8) % 2, (q // 4) % 2, (q % 4) // 2, q % 2] for q in pw.holes]
    vac, fock, blocks = ham_ref.get_norm_ordered_ham(L, holes_states, tkin, lst, w3, sparse=False)
    E, fk, bl = pw.hamiltonian(sparse=False)
    assert abs(vac - E) < 1e-12
    for a, b in zip(fock, fk):
        np.testing.assert_allclose(a, b, atol=1e-12)
    # blocks that contain at least one hole pair are complete in the filtered list
    for a, b in zip(blocks[2:], bl[2:]):
        np.testing.assert_allclose(a, b, atol=1e-12)


def test_planewave_counts_match_structure_formulas():
    E, f, v = planewave_hamiltonian(3, 4, -4.0)
    assert len(v[0][1]) == 406224 and len(v[1][1]) == 15600
    assert np.all(f[1] == 0)
    assert abs(E - 6 * (-4.0) / 27) < 1e-12          # E_ref(A=4) = 6 C0 / L^3 (eps(0) = 0)


def _segsum_host(plan, X, src, out, alpha, ss=1, ds=1):
    ptr, so, val, dst = (x.cpu().numpy() for x in (plan.ptr, plan.src, plan.val, plan.dst))
    src = src.ravel(); o = out.ravel()
    for s in range(plan.S):
        for x in range(X):
            o[dst[s] + x * ds] += alpha * sum(val[e] * src[so[e] + x * ss]
                                              for e in range(ptr[s], ptr[s + 1]))


def test_ppph_plans_reproduce_v_ppph_dgrams():
    p, h = 5, 3
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=2)
    t1, t2 = random_amplitudes(p, h, seed=3)
    coo = to_coo(dense[1])
    idx, val = coo_arrays(coo, p, h, "ppph")
    pl = PpphPlans(idx, val, p, h, "cpu")
    r0, r1, r2, r3, r4, r5, r6 = cc.v_ppph_dgrams(coo, t1, t2)
    o = np.zeros((p, h, h)); _segsum_host(pl.ret0, h, t2, o, -0.5)
    np.testing.assert_allclose(o.sum(1), r0, atol=1e-13)
    o = np.zeros((p, p)); _segsum_host(pl.ret1, 1, t1, o, -1.0)
    np.testing.assert_allclose(o, r1, atol=1e-13)
    np.testing.assert_allclose(-o.T, r3, atol=1e-13)
    o = np.zeros((p, h, p, h)); _segsum_host(pl.ret2, h, t1, o, 1.0, 1, p * h)
    np.testing.assert_allclose(o.transpose(0, 2, 1, 3), r2, atol=1e-13)
    # ret2p: the P(ij) form with a unit-stride destination differs from ret2 element by element
    # but has the same P(ij) image
    q = np.zeros((p, p, h, h)); _segsum_host(pl.ret2p, h, t1, q, 1.0)     # [c][d][k][x]
    assert np.abs(q - r2).max() > 1e-3
    np.testing.assert_allclose(cc.pIJ(q), cc.pIJ(r2), atol=1e-13)
    W = np.zeros((p, h, p, h)); _segsum_host(pl.ret4, h, t1, W, 1.0)
    np.testing.assert_allclose(W, r4.transpose(1, 3, 0, 2), atol=1e-13)
    tau = t2 + np.einsum("ci,dj->cdij", t1, t1) - np.einsum("cj,di->cdij", t1, t1)
    o = _blocks_host(pl.r56, tau.reshape(p * p, h * h), p * h).reshape(p, h, h, h)
    np.testing.assert_allclose(o, (r5 + r6 - r6.transpose(0, 2, 1, 3)).transpose(0, 3, 1, 2),
                               atol=1e-13)


def _blocks_host(pb, src, nrow):
    cols, rows = pb.cols.numpy(), pb.rows.numpy()
    G = np.where(cols[:, None] >= 0, src[np.maximum(cols, 0)], 0.0)
    R = pb.V.numpy() @ G.reshape(pb.nb, pb.kb, src.shape[1])
    out = np.zeros((nrow, src.shape[1]))
    for r, i in zip(R.reshape(-1, src.shape[1]), rows):
        if i >= 0:
            out[i] += r
    return out


def test_pair_blocks_reproduce_pp_ladder():
    p, h = 6, 3
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=7)
    vp = dense[0].copy()
    vp[np.abs(vp) < 0.3] = 0.0          # make it sparse with several components
    t1, t2 = random_amplitudes(p, h, seed=8)
    coo = to_coo(vp)
    idx, val = coo_arrays(coo, p, h, "pppp")
    pb = PairBlocks(idx[:, 0] * p + idx[:, 1], idx[:, 2] * p + idx[:, 3], val, p * p, p * p,
                    "cpu", min_rows=4)
    tau = (t2 + np.einsum("ci,dj->cdij", t1, t1) - np.einsum("cj,di->cdij", t1, t1)).reshape(
        p * p, h * h)
    G = np.where(pb.cols.numpy()[:, None] >= 0, tau[np.maximum(pb.cols.numpy(), 0)], 0.0)
    R = pb.V.numpy() @ G.reshape(pb.nb, pb.kb, h * h)
    out = np.zeros((p * p, h * h))
    for r, i in zip(R.reshape(-1, h * h), pb.rows.numpy()):
        if i >= 0:
            out[i] += r
    l1, l2 = cc.v_pppp_dgrams(coo, t1, t2)
    np.testing.assert_allclose(out.reshape(p, p, h, h), l1 + cc.pIJ(l2), atol=1e-12)


class _TwoBodyLike:
    """The attribute interface of _torch_types.TwoBodyOperator (_torch_types.py:169-199):
    indices (N, 4) int64 tensor, values (N,) tensor."""

    def __init__(self, indices, values):
        self.indices, self.values = indices, values


@pytest.mark.parametrize("n", [4, 5, 37])
def test_coo_arrays_containers(n):
    """Every container of a sparse block gives the same (N,4) rows: the legacy list, a
    TwoBodyOperator-like object, this package's CooBlock, and the (4, N) pair to_soa_sparse
    returns (soa/coupled_cluster.py:18-46), including n = 4, where (4, N) and (N, 4) have the
    same shape and only the container fixes the layout."""
    from nulattice_amd.hamiltonian import CooList
    from nulattice_amd.plans import CooBlock
    rng = np.random.default_rng(n)
    p, h = 9, 4
    idx = np.stack([rng.integers(0, p, n), rng.integers(0, p, n), rng.integers(0, p, n),
                    rng.integers(0, h, n)], 1)
    val = rng.standard_normal(n)
    ref_idx, ref_val = coo_arrays([list(map(int, i)) + [float(v)] for i, v in zip(idx, val)],
                                  p, h, "ppph")
    np.testing.assert_array_equal(ref_idx, idx)
    soa = (torch.tensor(idx.T, dtype=torch.long), torch.tensor(val))       # (4, N) as SoA
    for c in (soa, _TwoBodyLike(torch.tensor(idx), torch.tensor(val)), CooBlock(idx, val),
              CooList(idx, val)):
        i2, v2 = coo_arrays(c, p, h, "ppph")
        np.testing.assert_array_equal(i2, idx)
        np.testing.assert_array_equal(v2, val)
    if n != 4:
        i2, _ = coo_arrays((idx, val), p, h, "ppph")          # unambiguous (N, 4) pair
        np.testing.assert_array_equal(i2, idx)


def test_coo_arrays_rejects_mismatched_layouts():
    from nulattice_amd.plans import CooBlock
    idx = np.zeros((5, 4), np.int64)
    with pytest.raises(ValueError):
        coo_arrays((idx, np.zeros(6)), 9, 4, "ppph")
    with pytest.raises(ValueError):
        coo_arrays(CooBlock(idx.T.copy(), np.zeros(5)), 9, 4, "ppph")


def test_host_blocks_reach_the_device_row_major():
    """An F-ordered NumPy block (a column slice, an eigenvector matrix) becomes a row-major
    tensor: torch.tensor alone keeps the column-major strides, which raw kernels would misread."""
    import numpy as np
    from nulattice_amd import engine
    x = np.asfortranarray(np.arange(12.0).reshape(3, 4))
    t = engine._dev(x, "cpu")
    assert t.is_contiguous() and t.stride() == (4, 1)
    assert np.array_equal(t.numpy(), x)

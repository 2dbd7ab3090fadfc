"""Host side of the HF-basis Hamiltonian (no GPU): the density-matrix normal ordering against the
oracle's lattice-basis blocks, and the HF fixed point."""
import json
import os

import numpy as np
import pytest

from nulattice_amd import hfbasis as hb
from oracle import hamiltonian_ref as ham
from oracle import lattice_ref as lat

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "trajectories.json")))


def _lists(L, vT1, vS1, w3):
    return (lat.Tkin(L), lat.contacts(vT1, vS1, L), lat.NNNcontact(w3, L) if w3 else None)


def permuted_blocks(op1, op2, op3, holes, parts, A):
    """normal_order at the lattice determinant, rotated by the permutation (holes, particles)."""
    n = len(holes) + len(parts)
    order = list(holes) + list(parts)
    rho = hb.init_density(n, holes)
    e, F, V = hb.normal_order(op1, op2, op3, rho)
    F = F[np.ix_(order, order)]
    V = V[np.ix_(order, order, order, order)]
    return (e,) + tuple(hb._blocks(F, V, A, sparse=False))


@pytest.mark.parametrize("w3", [0.0, 6.0])
def test_lattice_determinant_equals_oracle_blocks(w3):
    """With C a permutation (the reference's holes, then its particles) every block, the Fock
    matrices and E_ref equal the oracle's get_norm_ordered_ham (16O, L=2, with and without the
    three-body contact), so the density-matrix form of the normal ordering is the reference's."""
    g = GOLD["16O_L2"]
    L = 2
    op1, op2, op3 = _lists(L, g["vT1"], g["vS1"], w3)
    e_ref, fock, v = ham.get_norm_ordered_ham(L, lat.ref_16O_gs, op1, op2, op3, sparse=False)
    holes, parts = lat.states2PHSpace(lat.ref_16O_gs, L)
    e, fk, two = permuted_blocks(op1, op2, op3, holes, parts, len(holes))
    assert e == pytest.approx(e_ref, abs=1e-12)
    for x, y in zip(fk, fock):
        np.testing.assert_allclose(x, y, atol=1e-13)
    for x, y in zip(two, v):
        np.testing.assert_allclose(x, np.asarray(y), atol=1e-13)


def test_hf_fixed_point():
    """solve_hf on 16O at L=2 (two- and three-body contacts): the HF Fock matrix has no
    occupied-virtual block (Brillouin), E_HF is E_ref of the HF determinant, and it lies below
    the lattice determinant's E_ref (variational)."""
    g = GOLD["16O_L2"]
    L = 2
    op1, op2, op3 = _lists(L, g["vT1"], g["vS1"], g["w3"])
    holes, _ = lat.states2PHSpace(lat.ref_16O_gs, L)
    n = 4 * L ** 3
    A = len(holes)
    # start from a slightly perturbed density so that the iteration has work to do
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(np.eye(n) + 0.05 * rng.standard_normal((n, n)))
    C0 = Q[:, :0]
    d0 = hb.init_density(n, holes)
    occ = np.linalg.qr(d0[:, list(holes)] + 0.05 * rng.standard_normal((n, A)))[0]
    erg, C, conv = hb.solve_hf(op1, op2, op3, occ @ occ.T, eps=1e-11, max_iter=500)
    assert conv
    rho = C[:, :A] @ C[:, :A].T
    e_ref, F, _ = hb.normal_order(op1, op2, op3, rho)
    Fp = C.T @ F @ C
    assert np.abs(Fp[A:, :A]).max() < 1e-6
    assert erg == pytest.approx(e_ref, abs=1e-8)
    e_lat = hb.normal_order(op1, op2, op3, d0)[0]
    assert erg <= e_lat + 1e-10
    del C0


def test_two_and_three_body_tensors_are_antisymmetric():
    L = 2
    op1, op2, op3 = _lists(L, -9.0, -7.0, 6.0)
    n = 4 * L ** 3
    V = hb.two_body_tensor(op2, n)
    np.testing.assert_array_equal(V, -V.transpose(1, 0, 2, 3))
    np.testing.assert_array_equal(V, -V.transpose(0, 1, 3, 2))
    np.testing.assert_allclose(V, V.transpose(2, 3, 0, 1), atol=0)
    rho = hb.init_density(n, range(8))
    V3 = hb.three_body_no2b(op3, rho)
    np.testing.assert_allclose(V3, -V3.transpose(1, 0, 2, 3), atol=1e-15)
    # Gamma from the list equals the contraction of the dense tensor
    G, O = hb.fock_pieces(op2, op3, rho)
    np.testing.assert_allclose(G, np.einsum("ptqu,tu->pq", V, rho), atol=1e-13)
    np.testing.assert_allclose(O, np.einsum("ptqu,tu->pq", V3, rho), atol=1e-13)


@pytest.mark.parametrize("w3", [0.0, 6.0])
def test_pair_support_is_the_dense_interaction(w3):
    """pair_support carries V_eff = V + V3 on its ordered-pair support: scattered back to n^4 it
    equals two_body_tensor + three_body_no2b at a random density, and nothing outside the support
    is nonzero (16O lists, L=2)."""
    g = GOLD["16O_L2"]
    L = 2
    op1, op2, op3 = _lists(L, g["vT1"], g["vS1"], w3)
    n = 4 * L ** 3
    rng = np.random.default_rng(5)
    X = rng.standard_normal((n, 16))
    rho = X @ X.T / n
    pairs, Vr = hb.pair_support(op2, op3, rho, n)
    assert pairs.shape[1] == 12 * L ** 3            # ordered pairs of distinct states per site
    V = hb.two_body_tensor(op2, n) + hb.three_body_no2b(op3, rho)
    W = np.zeros_like(V)
    p, q = pairs
    W[p[:, None], q[:, None], p[None, :], q[None, :]] = Vr
    np.testing.assert_allclose(W, V, rtol=0, atol=1e-13)

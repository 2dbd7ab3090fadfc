"""GPU side of the HF-basis Hamiltonian: the quarter transformations against NumPy, the lattice
determinant against the oracle's blocks, and the invariance of E_ref and of the CCSD energy under
rotations that keep the occupied and virtual spaces."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from nulattice_amd import ccm  # noqa: E402
from nulattice_amd import hfbasis as hb  # noqa: E402
from oracle import hamiltonian_ref as ham  # noqa: E402
from oracle import lattice_ref as lat  # noqa: E402

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "trajectories.json")))
L = 2


def _lists(w3=True):
    g = GOLD["16O_L2"]
    return (lat.Tkin(L), lat.contacts(g["vT1"], g["vS1"], L),
            lat.NNNcontact(g["w3"], L) if w3 else None)


def test_permutation_basis_equals_oracle_blocks():
    op1, op2, op3 = _lists()
    e_ref, fock, v = ham.get_norm_ordered_ham(L, lat.ref_16O_gs, op1, op2, op3, sparse=False)
    holes, parts = lat.states2PHSpace(lat.ref_16O_gs, L)
    n = 4 * L ** 3
    C = np.eye(n)[:, list(holes) + list(parts)]
    e, fk, two = hb.hf_basis_hamiltonian(op1, op2, op3, C, len(holes), sparse=False)
    assert e == pytest.approx(e_ref, abs=1e-12)
    for x, y in zip(fk, fock):
        np.testing.assert_allclose(x, y, atol=1e-13)
    for x, y in zip(two, v):
        np.testing.assert_allclose(x, np.asarray(y), atol=1e-13)


def test_quarter_transformations_match_numpy():
    op1, op2, op3 = _lists()
    n = 4 * L ** 3
    rng = np.random.default_rng(1)
    C, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = 16
    e, fk, two = hb.hf_basis_hamiltonian(op1, op2, op3, C, A, sparse=False)
    e0, F, V = hb.normal_order(op1, op2, op3, C[:, :A] @ C[:, :A].T)
    Fr = C.T @ F @ C
    Vr = np.einsum("pa,qb,rc,sd,pqrs->abcd", C, C, C, C, V, optimize=True)
    ref_f, ref_v = hb._blocks(Fr, Vr, A, sparse=False)
    assert e == e0
    for x, y in zip(fk, ref_f):
        np.testing.assert_allclose(x, y, atol=1e-12)
    for x, y in zip(two, ref_v):
        np.testing.assert_allclose(x, y, atol=1e-12)


def test_hf_basis_ccsd_is_invariant_under_occupied_and_virtual_rotations():
    """16O at L=2 with the three-body contact: HF, then CCSD in the HF basis, and again after
    rotating C by random orthogonal matrices inside the occupied and inside the virtual space:
    E_ref is unchanged (1e-12) and so is the CCSD correlation energy (1e-8 relative)."""
    op1, op2, op3 = _lists()
    holes, _ = lat.states2PHSpace(lat.ref_16O_gs, L)
    n, A = 4 * L ** 3, len(holes)
    # from the lattice determinant the mixing iteration oscillates between degenerate orbitals;
    # a slightly perturbed start converges
    rng = np.random.default_rng(0)
    occ = np.linalg.qr(hb.init_density(n, holes)[:, list(holes)]
                       + 0.05 * rng.standard_normal((n, A)))[0]
    erg, C, conv = hb.solve_hf(op1, op2, op3, occ @ occ.T, eps=1e-11, max_iter=300)
    assert conv
    rng = np.random.default_rng(2)
    Uo, _ = np.linalg.qr(rng.standard_normal((A, A)))
    Uv, _ = np.linalg.qr(rng.standard_normal((n - A, n - A)))
    C2 = C.copy()
    C2[:, :A] = C[:, :A] @ Uo
    C2[:, A:] = C[:, A:] @ Uv
    res = []
    for c in (C, C2):
        e_ref, fock, two = hb.hf_basis_hamiltonian(op1, op2, op3, c, A)
        assert np.abs(fock[1]).max() < 1e-6                      # Brillouin: f_ph = 0
        e, _, _ = ccm.ccsd_solver(fock, two, eps=1e-12, maxSteps=500, diis="pulsed")
        res.append((e_ref, e))
    assert res[0][0] == pytest.approx(erg, abs=1e-8)
    assert res[1][0] == pytest.approx(res[0][0], abs=1e-12)
    assert res[1][1] == pytest.approx(res[0][1], rel=1e-8)

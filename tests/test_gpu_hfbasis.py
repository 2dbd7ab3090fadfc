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


def _dense_ppph(H):
    """v_ppph[c,d,a,k] back from the engine's single R56 block [(a,k)][(c,d)]."""
    p, h = H.p, H.h
    return H.ppph.r56.V.view(p, h, p, p).permute(2, 3, 0, 1).cpu().numpy()


def test_engine_build_equals_host_rotation():
    """hf_basis_engine (pair support, two GEMMs per block, no n^4 host array) against the dense
    rotation of hf_basis_hamiltonian: every block, the Fock matrices and E_ref (16O lists with
    the three-body contact, L=2, random orthogonal C), 1e-12."""
    op1, op2, op3 = _lists()
    n, A = 4 * L ** 3, 16
    rng = np.random.default_rng(7)
    C, _ = np.linalg.qr(rng.standard_normal((n, n)))
    C = np.asfortranarray(C)                # column-major input, as eigh returns it
    e0, f0, v0 = hb.hf_basis_hamiltonian(op1, op2, op3, C, A, sparse=False)
    e1, f1, H = hb.hf_basis_engine(op1, op2, op3, C, A)
    assert e1 == e0
    for x, y in zip(f1, f0):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-12)
    p = n - A
    got = [H.pppp.V.view(p, p, p, p).cpu().numpy(), _dense_ppph(H)] + \
        [t.cpu().numpy() for t in (H.v_pphh, H.v_phph, H.v_phhh, H.v_hhhh)]
    for name, x, y in zip(("pppp", "ppph", "pphh", "phph", "phhh", "hhhh"), got, v0):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-12, err_msg=name)
    assert H.ppph.n > 0.9 * v0[1].size           # the rotated block is dense


def test_engine_L4_lattice_determinant_reproduces_notebook():
    """16O at L=4 (contacts -8, -8, three-body 5.5; 256 states, A=16, p=240): with C the
    permutation (the reference's holes, then its particles) the GPU build is the reference's
    bare-basis Hamiltonian, so E_ref and the CCSD steps 0-10 equal examples/Example_CCSD2.ipynb
    cell 3 (1e-12): v_pppp is one dense 57600^2 block (26.5 GB) and v_ppph a 221M-entry device
    COO list, built from the 768-pair support on the GPU."""
    from nulattice_amd import hamiltonian as hm
    g = GOLD["16O_L4"]
    L4 = 4
    op1, op2, op3 = hm.Tkin(L4), hm.contacts(g["vT1"], g["vS1"], L4), hm.NNNcontact(g["w3"], L4)
    holes, parts = lat.states2PHSpace(lat.ref_16O_gs, L4)
    n = 4 * L4 ** 3
    C = np.eye(n)[:, list(holes) + list(parts)]
    print("L=4 build", flush=True)
    e_ref, fock, H = hb.hf_basis_engine(op1, op2, op3, C, len(holes))
    old_unit = 0.5 * 197.3269631 ** 2 / (lat.MASS * g["a_lat"] ** 2)
    assert abs(e_ref * old_unit - g["E_ref_MeV"]) < 1e-12 * abs(g["E_ref_MeV"])
    hist = []
    ccm.ccsd_solver(None, None, eps=1e-8, maxSteps=10, max_diis=10, hamiltonian=H, history=hist)
    ref = np.asarray(g["energies"][:11])
    assert len(hist) == 11
    np.testing.assert_allclose(hist, ref, rtol=1e-12)


def test_engine_L4_hf_basis_ccsd_invariant_under_rotations():
    """16O at L=4 with the three-body contact: HF (host, n = 256), the HF-basis Hamiltonian built
    on the GPU, CCSD to 1e-13; again after random orthogonal rotations inside the occupied (16)
    and virtual (240) spaces: E_ref equals the HF energy and is unchanged, and so is the CCSD
    correlation energy (1e-8 relative)."""
    from nulattice_amd import hamiltonian as hm
    g = GOLD["16O_L4"]
    L4 = 4
    op1, op2, op3 = hm.Tkin(L4), hm.contacts(g["vT1"], g["vS1"], L4), hm.NNNcontact(g["w3"], L4)
    holes, _ = lat.states2PHSpace(lat.ref_16O_gs, L4)
    n, A = 4 * L4 ** 3, len(holes)
    rng = np.random.default_rng(0)
    occ = np.linalg.qr(hb.init_density(n, holes)[:, list(holes)]
                       + 0.05 * rng.standard_normal((n, A)))[0]
    print("L=4 HF", flush=True)
    erg, C, conv = hb.solve_hf(op1, op2, op3, occ @ occ.T, eps=1e-11, max_iter=500)
    assert conv
    rng = np.random.default_rng(2)
    Uo, _ = np.linalg.qr(rng.standard_normal((A, A)))
    Uv, _ = np.linalg.qr(rng.standard_normal((n - A, n - A)))
    C2 = C.copy()
    C2[:, :A] = C[:, :A] @ Uo
    C2[:, A:] = C[:, A:] @ Uv
    res = []
    for c in (C, C2):
        print("L=4 HF-basis build + CCSD", flush=True)
        e_ref, fock, H = hb.hf_basis_engine(op1, op2, op3, c, A)
        assert np.abs(fock[1]).max() < 1e-6                      # Brillouin: f_ph = 0
        info = {}
        e, _, _ = ccm.ccsd_solver(None, None, eps=1e-13, maxSteps=400, diis="pulsed",
                                  hamiltonian=H, info=info)
        res.append((e_ref, e))
        print(f"E_ref {e_ref!r} E_corr {e!r} steps {info.get('steps_fp64')}", flush=True)
        del H
    assert res[0][0] == pytest.approx(erg, abs=1e-9)
    assert res[1][0] == pytest.approx(res[0][0], abs=1e-10)
    assert res[1][1] == pytest.approx(res[0][1], rel=1e-8)

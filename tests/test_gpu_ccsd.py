"""GPU parity of the CCSD iteration against the CPU oracle and the reference's stored energies."""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import ccm, engine, ops  # noqa: E402
from nulattice_amd.planewave import planewave_hamiltonian  # noqa: E402
from oracle import ccsd_ref as cc  # noqa: E402
from oracle import hamiltonian_ref as ham_ref  # noqa: E402
from oracle import lattice_ref as lat  # noqa: E402
from tests.helpers import random_amplitudes, random_antisym_hamiltonian, to_coo  # noqa: E402

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "trajectories.json")))


def host(t):
    return t.detach().cpu().numpy()


def dev(x):
    return torch.tensor(np.asarray(x, np.float64), dtype=torch.float64, device="cuda")


def _coo_list(c):
    idx, val = c
    return [[int(a), int(b), int(x), int(y), float(v)] for (a, b, x, y), v in zip(idx, val)]


@pytest.mark.parametrize("seed", [0, 1])
def test_one_jacobi_step_matches_oracle(seed):
    """One full update (t1Iter, t2Iter, mixing, energy) on a random antisymmetric Hamiltonian
    with nonzero t1 and every v_ppph / v_pppp term active: agrees with the oracle to 1e-12."""
    p, h = 7, 4
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=seed)
    f_pp, f_ph, f_hh = fock
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = dense
    t1, t2 = random_amplitudes(p, h, seed=seed + 10)
    vp, v3 = to_coo(v_pppp), to_coo(v_ppph)
    res = cc.v_ppph_dgrams(v3, t1, t2)
    m = 0.3
    t1_ref = m * t1 + (1 - m) * cc.t1Iter(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, res)
    t2_ref = m * t2 + (1 - m) * cc.t2Iter(t1, t2, f_ph, f_hh, f_pp, vp, v_phph, v_phhh, v_pphh,
                                          res, v_hhhh)
    e_ref = cc.ccsd_energy(f_ph, v_pphh, t2_ref, t1_ref)

    H = engine.Hamiltonian(fock, [vp, v3, v_pphh, v_phph, v_phhh, v_hhhh], device="cuda",
                           min_block_rows=8)
    B = engine.Buffers(p, h, H.device)
    t1n = torch.empty((p, h), dtype=torch.float64, device="cuda")
    t2n = torch.empty((p, p, h, h), dtype=torch.float64, device="cuda")
    engine.update(H, dev(t1), dev(t2), B, t1n, t2n, m)
    np.testing.assert_allclose(host(t1n), t1_ref, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(host(t2n), t2_ref, rtol=1e-12, atol=1e-13)
    assert abs(B.E.item() - e_ref) <= 1e-12 * abs(e_ref)


def test_init_and_energy_match_oracle():
    p, h = 6, 3
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=4)
    f_pp, f_ph, f_hh = fock
    for delta in (0.0, 0.7):
        t1 = ccm.t1Init(f_ph, f_pp, f_hh, delta)
        t2 = ccm.t2Init(f_pp, f_hh, dense[2], delta)
        np.testing.assert_allclose(host(t1), cc.t1Init(f_ph, f_pp, f_hh, delta), rtol=1e-14)
        np.testing.assert_allclose(host(t2), cc.t2Init(f_pp, f_hh, dense[2], delta), rtol=1e-14,
                                   atol=1e-15)
        e = ccm.ccsd_energy(f_ph, dense[2], t2, t1)
        assert abs(e - cc.ccsd_energy(f_ph, dense[2], host(t2), host(t1))) < 1e-13 * abs(e)


def test_golden_16O_L2_trajectory_on_gpu():
    """examples/Example_CCSD.ipynb (L=2 16O, three-body contact): the GPU solver with the pulsed
    DIIS reproduces the stored per-step energies (steps 0-10 plain Jacobi to 1e-12; the DIIS
    steps to the conditioning of the DIIS solve, 1e-6; final energy to 1e-9)."""
    g = GOLD["16O_L2"]
    vac, fock, v = ham_ref.get_norm_ord_int(2, lat.ref_16O_gs, g["vT1"], g["vS1"], g["w3"], True)
    hist = []
    e, t1, t2 = ccm.ccsd_solver(fock, v, eps=1e-8, maxSteps=500, max_diis=10, mixing=0.5,
                                sparse=True, diis="pulsed", history=hist)
    ref = np.array(g["energies"])
    assert len(hist) == len(ref)
    rel = np.abs(np.array(hist) - ref) / np.abs(ref)
    assert rel[:11].max() < 1e-12
    assert rel.max() < 1e-6
    assert abs(e - ref[-1]) < 1e-9 * abs(ref[-1])


def test_legacy_solver_dense_input_matches_oracle():
    """sparse=False inputs (dense v_pppp / v_ppph) and the legacy no-DIIS schedule."""
    g = GOLD["16O_L2"]
    _, fock, v = ham_ref.get_norm_ord_int(2, lat.ref_16O_gs, g["vT1"], g["vS1"], g["w3"], False)
    hist = []
    ccm.ccsd_solver(fock, v, maxSteps=10, max_diis=10, sparse=False, history=hist)
    ref = np.array(g["energies"][:11])          # steps 0-10: before any DIIS step
    assert len(hist) == 11
    assert np.max(np.abs(np.array(hist) - ref) / np.abs(ref)) < 1e-12


def test_planewave_L3_A4_matches_oracle():
    """Plane-wave L=3, A=4, C0=-4 (108 orbitals): 6 legacy steps agree with the oracle to 1e-12
    relative; T1 stays exactly zero; T2 stays antisymmetric."""
    E0, fock, v = planewave_hamiltonian(3, 4, -4.0)
    vl = [_coo_list(v[0]), _coo_list(v[1])] + v[2:]
    href = []
    cc.ccsd_solver(fock, vl, maxSteps=6, max_diis=0, history=href)
    hg = []
    e, t1, t2 = ccm.ccsd_solver(fock, v, maxSteps=6, max_diis=0, history=hg)
    np.testing.assert_allclose(hg, href, rtol=1e-12)
    assert float(t1.abs().max()) == 0.0
    t2h = host(t2)
    np.testing.assert_allclose(t2h, -t2h.transpose(1, 0, 2, 3), atol=1e-15)


def test_planewave_L4_A28_matches_oracle_fixture():
    """Plane-wave L=4, A=28 (256 orbitals, p=228, h=28; C0=-4, D0=1): the GPU FP64 iteration on
    the Hamiltonian written on the GPU follows the oracle's first 20 steps
    (tests/golden/planewave_L4_A28.json, written by tests/golden/make_planewave_L4.py) to 1e-10
    relative per step, from the same E_ref."""
    from nulattice_amd.planewave import planewave_device_hamiltonian
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "planewave_L4_A28.json")))
    E0, fock, H = planewave_device_hamiltonian(g["L"], g["A"], g["C0"], g["D0"])
    assert abs(E0 - g["E_ref"]) <= 1e-13 * abs(g["E_ref"])
    hist = []
    ccm.ccsd_solver(None, None, eps=1e-14, maxSteps=g["steps"], max_diis=g["max_diis"],
                    mixing=g["mixing"], hamiltonian=H, history=hist)
    assert len(hist) == len(g["energies"]) == g["steps"] + 1
    np.testing.assert_allclose(hist, g["energies"], rtol=1e-10)


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_precision_form_refines_to_fp64(prec):
    """Low-precision contractions converge near the FP64 fixed point; the FP64 refinement lands
    on it (|dE| <= 1e-9 relative at eps = 1e-11)."""
    E0, fock, v = planewave_hamiltonian(3, 4, -4.0)
    H = engine.Hamiltonian(fock, v, device="cuda")
    e64, _, _ = ccm.ccsd_solver(fock, v, eps=1e-11, maxSteps=400, max_diis=0, hamiltonian=H)
    info = {}
    e_lp, _, _ = ccm.ccsd_solver(fock, v, eps=1e-11, maxSteps=400, max_diis=0, hamiltonian=H,
                                 precision=prec, info=info)
    assert abs(info["E_lowprec"] - e64) < 1e-3 * abs(e64)
    assert abs(e_lp - e64) < 1e-9 * abs(e64)
    assert info["steps_lowprec"] > 0


def test_bf16_images_are_bitwise_the_per_product_conversions():
    """With every tau product in BF16 (L=3, A=28: h*h, p*p above the low-precision threshold),
    forming tau directly in BF16, and taking every BF16 layout of t2 from one cast of t2, round
    exactly as converting the FP64 tensors per product does, so two BF16 iterations give
    bit-identical amplitudes and energy either way."""
    E0, fock, v = planewave_hamiltonian(3, 28, -4.0)        # p = 80, h = 28: p, p*h multiples of 8
    H = engine.Hamiltonian(fock, v, device="cuda")
    p, h = H.p, H.h
    assert p % 8 == 0 and (h * h) % 8 == 0 and p * p >= ops.LP_MIN_DIM
    res = {}
    # (images on, carried by the assembly): the carried image is the cast's bit for bit
    for flag, carry in ((True, False), (False, False), (True, True)):
        engine.TAU16 = engine.T2_IMAGES = flag
        try:
            B = engine.Buffers(p, h, H.device)
            t1 = torch.zeros(p, h, dtype=torch.float64, device="cuda")
            t2 = torch.zeros(p, p, h, h, dtype=torch.float64, device="cuda")
            engine.init_amplitudes(H, B, t1, t2, 0.0)
            t1n, t2n = torch.empty_like(t1), torch.empty_like(t2)
            img = engine.T2Image() if carry else None
            for _ in range(3):
                engine.update(H, t1, t2, B, t1n, t2n, 0.5, ccm._PREC["bf16"], img=img)
                t1, t1n, t2, t2n = t1n, t1, t2n, t2
            if carry:
                assert img.src is t2
            res[(flag, carry)] = (t2.clone(), float(B.E.item()))
        finally:
            engine.TAU16 = engine.T2_IMAGES = True
    for key in ((False, False), (True, True)):
        assert torch.equal(res[(True, False)][0], res[key][0])
        assert res[(True, False)][1] == res[key][1]


def test_product_hamiltonian_on_gpu_solver():
    """nulattice_amd.hamiltonian (the package's own input step) + the GPU solver reproduce the
    notebook trajectory's steps 0-10 (sparse CooList blocks) and the dense form agrees."""
    from nulattice_amd import hamiltonian as hm
    g = GOLD["16O_L2"]
    for sparse in (True, False):
        vac, fock, v = hm.get_norm_ord_int(2, lat.ref_16O_gs, g["vT1"], g["vS1"], g["w3"],
                                           sparse)
        hist = []
        ccm.ccsd_solver(fock, v, maxSteps=10, sparse=sparse, history=hist)
        np.testing.assert_allclose(hist, g["energies"][:11], rtol=1e-12)


def test_bf16_operands_carried_by_the_assembly_are_bitwise_the_conversions():
    """Random Hamiltonian with a T1 source, p = 32, h = 16, 262,144 amplitudes (enough for the
    rare FP64 values whose FP32 rounding ties between two BF16 values, which a single FP64 ->
    BF16 rounding would round differently): three BF16 steps that take t2's BF16 image from the
    previous step's assembly (engine.T2Image, nl_assemble_t2_16) give bit-identical amplitudes
    and energy to casting t2 at the start of each step."""
    p, h = 32, 16
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=21, scale=0.1)
    H = engine.Hamiltonian(fock, dense, device="cuda")
    t1_0, t2_0 = random_amplitudes(p, h, seed=22, scale=0.02)
    res = {}
    for carry in (False, True):
        B = engine.Buffers(p, h, H.device)
        t1, t2 = dev(t1_0), dev(t2_0)
        t1n, t2n = torch.empty_like(t1), torch.empty_like(t2)
        img = engine.T2Image() if carry else None
        for _ in range(3):
            engine.update(H, t1, t2, B, t1n, t2n, 0.5, ccm._PREC["bf16"], img=img)
            t1, t1n, t2, t2n = t1n, t1, t2n, t2
        if carry:
            assert img.src is t2
        res[carry] = (t2.clone(), t1.clone(), float(B.E.item()))
    assert float(res[True][1].abs().max()) > 0
    assert torch.equal(res[False][0], res[True][0])
    assert torch.equal(res[False][1], res[True][1])
    assert res[False][2] == res[True][2]

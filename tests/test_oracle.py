"""Pins of the CPU oracle (no GPU): reference golden trajectories, an unfactorised CCSD residual,
A=2 exactness, dense/sparse agreement and operator-list counts."""
import json
import os

import numpy as np
import pytest

from oracle import ccsd_ref as cc
from oracle import hamiltonian_ref as ham
from oracle import lattice_ref as lat
from tests.helpers import random_amplitudes, random_antisym_hamiltonian, to_coo

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "trajectories.json")))


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) / np.abs(b)


def test_golden_16O_L2_full_trajectory():
    """examples/Example_CCSD.ipynb cell 5: all 33 printed energies.

    Steps 0-10 are plain damped Jacobi and must agree to round-off.  Steps 11-32 follow a
    pulsed-DIIS extrapolation; the DIIS solve is ill-conditioned and amplifies round-off (the
    reference's own dense and sparse runs differ by 1.2e-7 right after the first DIIS step,
    benchmarks/CCSDBenchmark.ipynb cell 2), so they are held to 1e-6 and the converged energy
    to 1e-9.
    """
    g = GOLD["16O_L2"]
    vac, fock, v = ham.get_norm_ord_int(2, lat.ref_16O_gs, g["vT1"], g["vS1"], g["w3"], True)
    hist = []
    e, _, _ = cc.ccsd_solver(fock, v, eps=1e-8, maxSteps=500, max_diis=10, mixing=0.5,
                             sparse=True, diis="pulsed", history=hist)
    ref = g["energies"]
    assert len(hist) == len(ref)
    assert _rel(hist[:11], ref[:11]).max() < 1e-13
    assert _rel(hist[11:], ref[11:]).max() < 1e-6
    assert _rel(e, ref[-1]) < 1e-9
    # the published MeV value (HBARC of that version was 197.3269631 MeV fm)
    old_unit = 0.5 * 197.3269631 ** 2 / (lat.MASS * g["a_lat"] ** 2)
    assert abs((e + vac) * old_unit - g["E_gs_MeV"]) < 1e-6


def test_golden_16O_L2_legacy_solver_matches_first_steps():
    """The current legacy solver (DIIS never entered) reproduces steps 0-10 exactly."""
    g = GOLD["16O_L2"]
    _, fock, v = ham.get_norm_ord_int(2, lat.ref_16O_gs, g["vT1"], g["vS1"], g["w3"], True)
    hist = []
    cc.ccsd_solver(fock, v, eps=1e-8, maxSteps=10, max_diis=10, sparse=True, diis="legacy",
                   history=hist)
    assert _rel(hist, g["energies"][:11]).max() < 1e-13


def test_golden_16O_L4_first_steps():
    """examples/Example_CCSD2.ipynb cell 3 (L=4, three-body contact 5.5): steps 0-2, and E_ref."""
    g = GOLD["16O_L4"]
    vac, fock, v = ham.get_norm_ordered_ham(4, lat.ref_16O_gs, lat.Tkin(4),
                                            lat.contacts(-8.0, -8.0, 4),
                                            lat.NNNcontact(5.5, 4), sparse=True)
    old_unit = 0.5 * 197.3269631 ** 2 / (lat.MASS * g["a_lat"] ** 2)
    assert _rel(vac * old_unit, g["E_ref_MeV"]) < 1e-12
    hist = []
    cc.ccsd_solver(fock, v, eps=1e-8, maxSteps=2, max_diis=10, sparse=True, diis="pulsed",
                   history=hist)
    assert _rel(hist, g["energies"][:3]).max() < 1e-13


def test_operator_counts_match_notebook():
    """Example_CCSD2.ipynb cell 2 prints 1792 / 512 / 256 elements at L=4."""
    assert len(lat.Tkin(4)) == 1792
    assert len(lat.contacts(-8.0, -8.0, 4)) == 512
    assert len(lat.NNNcontact(5.5, 4)) == 256


def test_laplacian_spectrum():
    """Tkin is the periodic Laplacian: eigenvalues 6 - 2 sum_d cos(2 pi k_d / L), 4-fold."""
    for L in (2, 3):
        n = 4 * L ** 3
        T = np.zeros((n, n))
        for p, q, val in lat.Tkin(L):
            T[p, q] += val
        ks = np.arange(L)
        eps = np.array([6 - 2 * (np.cos(2 * np.pi * a / L) + np.cos(2 * np.pi * b / L)
                                 + np.cos(2 * np.pi * c / L))
                        for a in ks for b in ks for c in ks])
        np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(T)), np.sort(np.repeat(eps, 4)),
                                   atol=1e-12)


# ------------------------------------------------------------------ unfactorised residual
def _c(s, *ops):
    return np.einsum(s, *ops, optimize=True)


def textbook_h1(f_pp, f_ph, f_hh, v_ppph, v_pphh, v_phph, v_phhh, t1, t2):
    """Unfactorised spin-orbital CCSD T1 residual (Crawford & Schaefer), as restated from
    benchmarks/DeuteronBenchmark.py:108-124 - an implementation independent of the factorised
    t1Iter."""
    P = _c
    return (f_ph + P("ac,ci->ai", f_pp, t1) - P("ki,ak->ai", f_hh, t1)
            - P("akci,ck->ai", v_phph, t1) + P("ck,acik->ai", f_ph, t2)
            - 0.5 * P("cdak,cdki->ai", v_ppph, t2) - 0.5 * P("cikl,cakl->ai", v_phhh, t2)
            - P("ck,ci,ak->ai", f_ph, t1, t1) - P("cikl,ck,al->ai", v_phhh, t1, t1)
            - P("cdak,ck,di->ai", v_ppph, t1, t1)
            - P("cdkl,ck,di,al->ai", v_pphh, t1, t1, t1)
            + P("cdkl,ck,dali->ai", v_pphh, t1, t2)
            - 0.5 * P("cdkl,cdki,al->ai", v_pphh, t2, t1)
            - 0.5 * P("cdkl,cakl,di->ai", v_pphh, t2, t1))


def textbook_h2(f_pp, f_ph, f_hh, v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh, t1, t2):
    """Unfactorised spin-orbital CCSD T2 residual, restated from
    benchmarks/DeuteronBenchmark.py:126-161."""
    A, I = cc.pAB, cc.pIJ
    P = _c
    h = v_pphh.copy()
    h += A(P("bc,acij->abij", f_pp, t2)) - I(P("kj,abik->abij", f_hh, t2))
    h += 0.5 * P("klij,abkl->abij", v_hhhh, t2) + 0.5 * P("abcd,cdij->abij", v_pppp, t2)
    h += -A(I(P("bkcj,acik->abij", v_phph, t2)))
    h += I(P("abcj,ci->abij", v_ppph, t1)) + A(P("bkij,ak->abij", v_phhh, t1))
    h += 0.5 * I(A(P("cdkl,acik,dblj->abij", v_pphh, t2, t2)))
    h += 0.25 * P("cdkl,cdij,abkl->abij", v_pphh, t2, t2)
    h += -0.5 * A(P("cdkl,acij,bdkl->abij", v_pphh, t2, t2))
    h += -0.5 * I(P("cdkl,abik,cdjl->abij", v_pphh, t2, t2))
    h += 0.5 * A(P("klij,ak,bl->abij", v_hhhh, t1, t1))
    h += 0.5 * I(P("abcd,ci,dj->abij", v_pppp, t1, t1))
    h += -I(A(P("bkci,ak,cj->abij", v_phph, t1, t1)))
    h += A(P("ck,ak,bcij->abij", f_ph, t1, t2)) + I(P("ck,ci,abjk->abij", f_ph, t1, t2))
    h += -I(P("cikl,ck,ablj->abij", v_phhh, t1, t2))
    h += -A(P("cdak,ck,dbij->abij", v_ppph, t1, t2))
    h += I(A(P("dcak,di,bcjk->abij", v_ppph, t1, t2)))
    h += -I(A(P("cikl,al,bcjk->abij", v_phhh, t1, t2)))
    h += 0.5 * I(P("cjkl,ci,abkl->abij", v_phhh, t1, t2))
    h += 0.5 * A(P("cdbk,ak,cdij->abij", v_ppph, t1, t2))
    h += 0.5 * I(A(P("cdbk,ci,ak,dj->abij", v_ppph, t1, t1, t1)))
    h += 0.5 * I(A(P("cjkl,ci,ak,bl->abij", v_phhh, t1, t1, t1)))
    h += -I(P("cdkl,ck,di,ablj->abij", v_pphh, t1, t1, t2))
    h += -A(P("cdkl,ck,al,dbij->abij", v_pphh, t1, t1, t2))
    h += 0.25 * I(P("cdkl,ci,dj,abkl->abij", v_pphh, t1, t1, t2))
    h += 0.25 * A(P("cdkl,ak,bl,cdij->abij", v_pphh, t1, t1, t2))
    h += I(A(P("cdkl,ci,bl,adkj->abij", v_pphh, t1, t1, t2)))
    h += 0.25 * I(A(P("cdkl,ci,ak,dj,bl->abij", v_pphh, t1, t1, t1, t1)))
    return h


@pytest.mark.parametrize("sparse", [True, False])
def test_factorised_residual_equals_unfactorised(sparse):
    """t1Iter/t2Iter's factorised H1/H2 equal the unfactorised residual at arbitrary amplitudes."""
    p, h = 6, 4
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=3)
    f_pp, f_ph, f_hh = fock
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = dense
    t1, t2 = random_amplitudes(p, h, seed=4)
    if sparse:
        res = cc.v_ppph_dgrams(to_coo(v_ppph), t1, t2)
        vp = to_coo(v_pppp)
    else:
        res, vp = v_ppph, v_pppp
    H1, _, _ = cc.t1_residual(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, res, sparse)
    H2, _, _ = cc.t2_residual(t1, t2, f_ph, f_hh, f_pp, vp, v_phph, v_phhh, v_pphh, res, v_hhhh,
                              sparse)
    np.testing.assert_allclose(H1, textbook_h1(f_pp, f_ph, f_hh, v_ppph, v_pphh, v_phph, v_phhh,
                                               t1, t2), atol=1e-12)
    np.testing.assert_allclose(H2, textbook_h2(f_pp, f_ph, f_hh, v_pppp, v_ppph, v_pphh, v_phph,
                                               v_phhh, v_hhhh, t1, t2), atol=1e-12)


def test_planted_sign_error_is_caught():
    """A sign flip in one diagram (planted in a copy of the oracle on the CPU) breaks the pin."""
    p, h = 5, 3
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=5)
    f_pp, f_ph, f_hh = fock
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = dense
    t1, t2 = random_amplitudes(p, h, seed=6)
    good = textbook_h2(f_pp, f_ph, f_hh, v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh, t1, t2)
    orig = cc.dgram_cikl_al_bcjk
    try:
        cc.dgram_cikl_al_bcjk = lambda v, a, b: -orig(v, a, b)
        H2, _, _ = cc.t2_residual(t1, t2, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh,
                                  v_ppph, v_hhhh, sparse=False)
    finally:
        cc.dgram_cikl_al_bcjk = orig
    assert np.abs(H2 - good).max() > 1e-4


# ------------------------------------------------------------------ A = 2 exactness
def _fci_two_body_ground(L, holes, vT1, vS1):
    """Lowest eigenvalue of the two-particle Hamiltonian in the (Tz, Sz) sector of ``holes``."""
    n = 4 * L ** 3
    T = np.zeros((n, n))
    for p, q, val in lat.Tkin(L):
        T[p, q] += val
    V = np.zeros((n, n, n, n))
    for p, q, r, s, val in lat.contacts(vT1, vS1, L):
        V[p, q, r, s] = val; V[q, p, r, s] = -val; V[p, q, s, r] = -val; V[q, p, s, r] = val
    tz = lambda x: (x % 4) // 2
    sz = lambda x: x % 2
    hs = [lat.state2index(x, L) for x in holes]
    TZ, SZ = sum(tz(x) for x in hs), sum(sz(x) for x in hs)
    pairs = [(p, q) for p in range(n) for q in range(p + 1, n) if tz(p) + tz(q) == TZ
             and sz(p) + sz(q) == SZ]
    H = np.zeros((len(pairs), len(pairs)))
    for A_, (p, q) in enumerate(pairs):
        for B_, (r, s) in enumerate(pairs):
            one = (T[p, r] * (q == s) + T[q, s] * (p == r) - T[p, s] * (q == r)
                   - T[q, r] * (p == s))
            H[A_, B_] = one + V[p, q, r, s]
    return np.linalg.eigvalsh(H)[0]


def test_ccsd_is_exact_for_two_particles():
    """CCSD is exact for A=2 (cf. benchmarks/DeuteronBenchmark.py:322-324): E_ref + E_corr
    equals exact diagonalisation.  Pins get_norm_ord_int, E_ref and the whole iteration."""
    L, vT1, vS1 = 2, -9.0, -9.0
    vac, fock, v = ham.get_norm_ord_int(L, lat.ref_2H_gs, vT1, vS1, 0, sparse=True)
    # delta=1 only shifts the starting guess: the particle and hole levels of the half-filled
    # site are degenerate, so the plain perturbative guess is 0/0.
    e, _, _ = cc.ccsd_solver(fock, v, eps=1e-13, maxSteps=3000, max_diis=10, mixing=0.5,
                             sparse=True, diis="pulsed", delta=1.0)
    assert abs(e + vac - _fci_two_body_ground(L, lat.ref_2H_gs, vT1, vS1)) < 1e-8


def test_dense_and_sparse_solvers_agree():
    """sparse=True (COO ladders) and sparse=False (dense einsums) give the same trajectory."""
    g = GOLD["16O_L2"]
    _, fs, vs = ham.get_norm_ord_int(2, lat.ref_16O_gs, g["vT1"], g["vS1"], g["w3"], True)
    _, fd, vd = ham.get_norm_ord_int(2, lat.ref_16O_gs, g["vT1"], g["vS1"], g["w3"], False)
    hs, hd = [], []
    cc.ccsd_solver(fs, vs, maxSteps=6, max_diis=0, sparse=True, history=hs)
    cc.ccsd_solver(fd, vd, maxSteps=6, max_diis=0, sparse=False, history=hd)
    assert _rel(hs, hd).max() < 1e-13


def _pw_lists(L, A, C0):
    from nulattice_amd.planewave import planewave_hamiltonian
    E0, fock, v = planewave_hamiltonian(L, A, C0)
    vl = [[[int(x) for x in i] + [float(val)] for i, val in zip(*v[0])],
          [[int(x) for x in i] + [float(val)] for i, val in zip(*v[1])]] + v[2:]
    return fock, v, vl


def test_oracle_failure_paths(capsys):
    """The oracle's failure semantics are the reference driver's (coupled_cluster.py:675-687):
    |E| > 1e10 prints "Iteration Diverged" and returns None (forced by over-relaxation,
    mixing = -20, which makes t <- -20 t + 21 t_new); a NaN energy prints "nan error" and returns
    None (forced by a degenerate Fock diagonal f_aa = f_ii with delta = 0: 0/0 in the initial
    amplitudes).  The GPU drivers are tested on the same two inputs."""
    fock, _, vl = _pw_lists(2, 4, -4.0)
    h = []
    assert cc.ccsd_solver(fock, vl, max_diis=0, maxSteps=100, mixing=-20.0, history=h) is None
    assert abs(h[-1]) > 1e10 and len(h) < 100
    assert "Iteration Diverged" in capsys.readouterr().out
    f2 = [x.copy() for x in fock]
    f2[0][0, 0] = f2[2][0, 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        assert cc.ccsd_solver(f2, vl, max_diis=0, maxSteps=100) is None
    assert "nan error" in capsys.readouterr().out


@pytest.mark.parametrize("soa", [False, True])
def test_oracle_diis_window_is_the_constrained_least_squares_fit(soa):
    """On a linear fixed-point map x -> G x + b, the oracle's DIIS step returns sum_k c_k x_k
    with c minimising |sum_k c_k e_k| subject to sum_k c_k = 1, computed here in closed form
    (c proportional to (E^T E)^-1 1) over exactly the kept window: the last max_diis - 1 errors
    for the pulsed variant (coupled_cluster.py:635-672), all max_diis for the SoA variant
    (soa/coupled_cluster.py:520-567).  The other window gives a different vector, so a wrong
    window or pairing of errors with amplitudes fails the check."""
    rng = np.random.default_rng(17)
    n, nd = 30, 6
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    G = Q @ np.diag(rng.uniform(-0.85, 0.85, n)) @ Q.T
    b = rng.standard_normal(n)
    x = rng.standard_normal(n)
    split = lambda v: (v[:6].reshape(3, 2).copy(), v[6:].reshape(2, 2, 3, 2).copy())
    x1, x2 = split(x)
    h1, h2, errs, xs, es = [x1], [x2], [], [], []
    ext = None
    for _ in range(nd):
        xn = G @ x + b
        a1, a2 = split(xn)
        h1.append(a1); h2.append(a2); errs.append(xn - x)
        xs.append(xn); es.append(xn - x)
        ext = cc.diis_window_step(h1, h2, errs, nd, soa)
        x = xn
    assert ext is not None
    got = np.concatenate([ext[0].ravel(), ext[1].ravel()])

    def fit(k):
        E, X = np.stack(es[-k:], 1), np.stack(xs[-k:], 1)
        c = np.linalg.solve(E.T @ E, np.ones(k))
        return X @ (c / c.sum())
    keep = nd if soa else nd - 1
    np.testing.assert_allclose(got, fit(keep), rtol=1e-9, atol=1e-12)
    other = fit(nd - 1 if soa else nd)
    assert np.max(np.abs(other - got)) > 1e-6 * np.max(np.abs(got))

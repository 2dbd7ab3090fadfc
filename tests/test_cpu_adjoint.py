"""CPU pins of the differentiation form's host pieces: the derivative Hamiltonian of the
plane-wave contact theory, the transposed sparse-term plans, and the sensitivity of the
complex-step Jacobian pin used by the GPU adjoint tests."""
import numpy as np
import torch

from nulattice_amd.planewave import PlaneWave, shell_holes
from nulattice_amd.plans import PairBlocks, PpphPlans
from oracle import ccsd_ref as cc
from tests.helpers import random_amplitudes, random_antisym_hamiltonian


def test_derivative_hamiltonian_matches_planewave_fd():
    """PlaneWave.coupling_derivative (Fock shift, two-body scale, dE_ref) equals the central
    difference of the plane-wave blocks in C0 and D0 (exact for blocks affine in C0, D0)."""
    L, A, C0, D0, hs = 3, 28, -4.0, 1.0, 0.5
    pw = PlaneWave(L, shell_holes(L, A), C0, D0)
    _, _, v0 = pw.hamiltonian()
    for name in ("C0", "D0"):
        dC, dD = (hs, 0.0) if name == "C0" else (0.0, hs)
        Ep, fp, vp = PlaneWave(L, shell_holes(L, A), C0 + dC, D0 + dD).hamiltonian()
        Em, fm, vm = PlaneWave(L, shell_holes(L, A), C0 - dC, D0 - dD).hamiltonian()
        dEref, dfock, scale = pw.coupling_derivative(name)
        assert abs(dEref - (Ep - Em) / (2 * hs)) < 1e-12 * max(1.0, abs(dEref))
        for a, b, c in zip(dfock, fp, fm):
            np.testing.assert_allclose(a, (b - c) / (2 * hs), atol=1e-13)
        for k in range(2, 6):
            np.testing.assert_allclose(scale * v0[k], (vp[k] - vm[k]) / (2 * hs), atol=1e-13)
        for k in range(2):
            np.testing.assert_array_equal(vp[k][0], v0[k][0])
            np.testing.assert_allclose(scale * v0[k][1], (vp[k][1] - vm[k][1]) / (2 * hs),
                                       atol=1e-13)


def _segmap(plan, X, ss, ds, nout, nin):
    ptr, so = plan.ptr.numpy(), plan.src.numpy()
    val, dst = plan.val.numpy(), plan.dst.numpy()
    M = np.zeros((nout, nin))
    for s in range(plan.S):
        for e in range(ptr[s], ptr[s + 1]):
            for x in range(X):
                M[dst[s] + x * ds, so[e] + x * ss] += val[e]
    return M


def _blockmap(pb, nout, nin):
    M = np.zeros((nout, nin))
    V, rows, cols = pb.V.numpy(), pb.rows.numpy().reshape(pb.nb, -1), pb.cols.numpy().reshape(
        pb.nb, -1)
    for b in range(pb.nb):
        for m, r in enumerate(rows[b]):
            for k, c in enumerate(cols[b]):
                if r >= 0 and c >= 0:
                    M[r, c] += V[b, m, k]
    return M


def test_transposed_ppph_plans_are_exact_transposes():
    """Each adjoint plan of the v_ppph terms is the transpose of its forward map (dense maps
    built from the plan arrays), including ret0's sum over k and the -1/2, -1 factors."""
    p, h = 5, 3
    rng = np.random.default_rng(0)
    n = 60
    idx = np.stack([rng.integers(0, p, n), rng.integers(0, p, n), rng.integers(0, p, n),
                    rng.integers(0, h, n)], 1)
    idx = np.unique(idx, axis=0)
    val = rng.standard_normal(len(idx))
    pl = PpphPlans(idx, val, p, h, "cpu")
    adj = pl.adjoint(lambda V: V.transpose(1, 2).contiguous())
    ph, pphh = p * h, p * p * h * h
    # ret0: H1[a,i] = -1/2 sum_k T0[a,k,i],  T0 = ret0(t2)
    T0 = _segmap(pl.ret0, h, 1, 1, p * h * h, pphh)
    S = np.zeros((p * h, p * h * h))
    for a in range(p):
        for k in range(h):
            for i in range(h):
                S[a * h + i, (a * h + k) * h + i] = 1.0
    np.testing.assert_allclose(-0.5 * _segmap(adj["ret0T"], h, 1, 1, pphh, ph),
                               (-0.5 * S @ T0).T, atol=1e-14)
    np.testing.assert_allclose(_segmap(adj["ret1T"], 1, 1, 1, ph, p * p),
                               _segmap(pl.ret1, 1, 1, 1, p * p, ph).T, atol=1e-14)
    np.testing.assert_allclose(_segmap(adj["ret2T"], h, ph, 1, ph, pphh),
                               _segmap(pl.ret2, h, 1, ph, pphh, ph).T, atol=1e-14)
    np.testing.assert_allclose(_segmap(adj["ret4T"], h, 1, 1, ph, pphh),
                               _segmap(pl.ret4, h, 1, 1, pphh, ph).T, atol=1e-14)
    np.testing.assert_allclose(_blockmap(adj["r56T"], p * p, p * h),
                               _blockmap(pl.r56, p * h, p * p).T, atol=1e-14)


def test_transposed_pair_blocks():
    rng = np.random.default_rng(1)
    n = 40
    row, col = rng.integers(0, 30, n), rng.integers(0, 25, n)
    key = np.unique(row * 100 + col)
    row, col = key // 100, key % 100
    val = rng.standard_normal(len(key))
    pb = PairBlocks(row, col, val, 30, 25, "cpu", min_rows=4)
    pt = pb.transposed(lambda V: V.transpose(1, 2).contiguous())
    np.testing.assert_allclose(_blockmap(pt, 25, 30), _blockmap(pb, 30, 25).T)
    ref = np.zeros((30, 25))
    ref[row, col] = val
    np.testing.assert_allclose(_blockmap(pb, 30, 25), ref)
    # each column pair appears in exactly one block: the transposed scatter has no conflicts
    c = pt.rows.numpy()
    c = c[c >= 0]
    assert len(c) == len(np.unique(c))


def _oracle_jvp(fock, dense, t1, t2, d1, d2, l1, l2, eps=1e-20):
    f_pp, f_ph, f_hh = (x.astype(complex) for x in fock)
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = (x.astype(complex) for x in dense)
    t1c, t2c = t1 + 1j * eps * d1, t2 + 1j * eps * d2
    H1 = cc.t1_residual(t1c, t2c, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, v_ppph,
                        sparse=False)[0]
    H2 = cc.t2_residual(t1c, t2c, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh, v_ppph,
                        v_hhhh, sparse=False)[0]
    return (l1 * H1.imag / eps).sum() + (l2 * H2.imag / eps).sum()


def test_complex_step_pin_matches_fd_and_sees_a_planted_sign_error():
    """The complex-step <l, J d> of the oracle agrees with a central difference of the real
    residual, and flipping the sign of one diagram moves it far beyond the 1e-12 tolerance the
    GPU transpose test uses."""
    p, h = 6, 3
    fock, dense, _, _ = random_antisym_hamiltonian(p, h, seed=5)
    t1, t2 = random_amplitudes(p, h, seed=15)
    l1, l2 = random_amplitudes(p, h, seed=25)
    d1, d2 = random_amplitudes(p, h, seed=35)
    good = _oracle_jvp(fock, dense, t1, t2, d1, d2, l1, l2)

    def res(s):
        H1 = cc.t1_residual(t1 + s * d1, t2 + s * d2, fock[1], fock[0], fock[2], dense[3],
                            dense[4], dense[2], dense[1], sparse=False)[0]
        H2 = cc.t2_residual(t1 + s * d1, t2 + s * d2, fock[1], fock[2], fock[0], dense[0],
                            dense[3], dense[4], dense[2], dense[1], dense[5], sparse=False)[0]
        return (l1 * H1).sum() + (l2 * H2).sum()
    fd = (res(1e-5) - res(-1e-5)) / 2e-5
    assert abs(fd - good) < 1e-7 * abs(good)
    orig = cc.dgram_cdkl_ci_bl_adkj
    try:
        cc.dgram_cdkl_ci_bl_adkj = lambda *a: -orig(*a)
        bad = _oracle_jvp(fock, dense, t1, t2, d1, d2, l1, l2)
    finally:
        cc.dgram_cdkl_ci_bl_adkj = orig
    assert abs(bad - good) > 1e-6 * abs(good)

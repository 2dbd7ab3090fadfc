"""Host index structures of the symmetry-sector form (no GPU): the packed layout, the crossed-
channel maps and the grouped-GEMM problem lists, stepped by a NumPy statement of the sector
kernels, against the oracle's dense t1Iter / t2Iter on plane-wave Hamiltonians."""
import numpy as np
import pytest

from nulattice_amd.planewave import PlaneWave, shell_holes
from nulattice_amd.sectors import SectorLayout, SectorViolation, pack_hamiltonian
from oracle import ccsd_ref as cc
from tests import sector_emulator as emu


def _setup(L, A, C0=-4.0, D0=1.0):
    pw = PlaneWave(L, shell_holes(L, A), C0, D0)
    E0, fock, v = pw.hamiltonian(sparse=False)
    lay = SectorLayout(*pw.labels())
    return pw, fock, v, lay


def _random_t2(lay, seed):
    """Random antisymmetric amplitudes inside the pair sectors."""
    rng = np.random.default_rng(seed)
    p, h = lay.p, lay.h
    R = lay.unpack(rng.standard_normal(lay.n) * 1e-2)
    R = R - R.transpose(1, 0, 2, 3)
    return R - R.transpose(0, 1, 3, 2)


@pytest.mark.parametrize("L,A", [(2, 4), (3, 28)])
def test_layout_is_a_bijection_onto_the_sectors(L, A):
    pw, fock, v, lay = _setup(L, A)
    p, h = lay.p, lay.h
    assert np.unique(lay.dense_index).size == lay.n
    # the sectors hold exactly the elements allowed by the labels (a < b, i < j)
    lp, lh, mod = pw.labels()
    tot = (lp[:, None, None, None] + lp[None, :, None, None] - lh[None, None, :, None]
           - lh[None, None, None, :])
    tot[..., :3] %= L
    allowed = ~np.any(tot != 0, axis=-1)
    ar, hr = np.arange(p), np.arange(h)
    allowed &= (ar[:, None] < ar[None, :])[:, :, None, None]
    allowed &= (hr[:, None] < hr[None, :])[None, None, :, :]
    assert allowed.sum() == lay.n
    assert allowed.reshape(-1)[lay.dense_index].all()
    # element() inverts dense_index
    a, b, i, j = lay.orb
    np.testing.assert_array_equal(lay.element(a, b, i, j), np.arange(lay.n))
    x = np.random.default_rng(0).standard_normal(lay.n)
    np.testing.assert_array_equal(lay.pack(lay.unpack(x)), x)
    # every Spp gather lands in ring storage
    assert lay.sg.min() >= 0 and lay.sg.max() < lay.nring
    # dressing chunks tile every segment, in order, at most DRESS_CHUNK entries each
    cs, cp, seg = lay.cstart, lay.cptr, lay.seg
    assert cs[0] == 0 and cs[-1] == seg[-1] and np.all(np.diff(cs) > 0)
    assert np.diff(cs).max() <= lay.DRESS_CHUNK
    np.testing.assert_array_equal(cs[cp], seg)
    # the launch lists of each product cover every output element exactly once
    from nulattice_amd.sectors import MODE_TILE
    for name in ("pp", "whh", "hh", "w", "s"):
        dsc = getattr(lay, "desc_" + name)
        covered = 0
        for mode, sub, til in getattr(lay, "launch_" + name):
            tm, tn = MODE_TILE[mode]
            M, N = sub[til[:, 0], 0], sub[til[:, 0], 1]
            assert (til[:, 1] < M).all() and (til[:, 2] < N).all()
            covered += int((np.minimum(tm, M - til[:, 1]) * np.minimum(tn, N - til[:, 2])).sum())
        assert covered == int((dsc[:, 0] * dsc[:, 1]).sum())
    # the t2 of the plane-wave Hamiltonian lies in the sectors
    v_pphh = v[2]
    assert 4 * np.count_nonzero(v_pphh.reshape(-1)[lay.dense_index]) == np.count_nonzero(v_pphh)
    np.testing.assert_array_equal(lay.unpack(lay.pack(v_pphh)), v_pphh)
    # signed_element agrees with the reference-layout images
    x = np.random.default_rng(1).standard_normal(lay.n)
    T = lay.unpack(x)
    e = lay.signed_element(a, b, i, j)
    np.testing.assert_array_equal(np.where(e >= 0, x[np.maximum(e, 0)],
                                           np.where(e <= -2, -x[np.maximum(-e - 2, 0)], 0.0)),
                                  T[a, b, i, j])
    e = lay.signed_element(b, a, i, j)
    np.testing.assert_array_equal(np.where(e >= 0, x[np.maximum(e, 0)],
                                           np.where(e <= -2, -x[np.maximum(-e - 2, 0)], 0.0)),
                                  T[b, a, i, j])


@pytest.mark.parametrize("L,A,seed", [(2, 4, 0), (3, 4, 1), (3, 28, 2)])
def test_sector_step_matches_oracle_dense_step(L, A, seed):
    """One damped Jacobi step of the sector statement equals the oracle's dense t2Iter step at
    random antisymmetric sector amplitudes (1e-12), and the oracle keeps t1 = 0 and t2 inside
    the sectors."""
    pw, fock, v, lay = _setup(L, A)
    f_pp, f_ph, f_hh = fock
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = v
    hst = pack_hamiltonian(lay, fock, v)
    t2 = _random_t2(lay, seed)
    t1 = np.zeros((lay.p, lay.h))
    new1 = cc.t1Iter(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, v_ppph, sparse=False)
    assert not new1.any()
    new2 = cc.t2Iter(t1, t2, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh, v_ppph, v_hhhh,
                     sparse=False)
    ref = 0.5 * t2 + 0.5 * new2
    np.testing.assert_array_equal(lay.unpack(lay.pack(ref)), ref)    # antisymmetric, in sectors
    tn, e, _ = emu.step(lay, hst, lay.pack(t2), 0.5)
    np.testing.assert_allclose(tn, lay.pack(ref), rtol=1e-12, atol=1e-15 * np.abs(ref).max())
    assert e == pytest.approx(cc.ccsd_energy(f_ph, v_pphh, ref, t1), rel=1e-12)


def test_sector_solve_matches_oracle_trajectory():
    """Six legacy solver steps in the sector statement against the oracle (L=3, A=4)."""
    pw, fock, v, lay = _setup(3, 4)
    hst = pack_hamiltonian(lay, fock, v)
    href = []
    cc.ccsd_solver(fock, v, maxSteps=6, max_diis=0, sparse=False, history=href)
    # t2Init in packed form
    f = hst["fdiag"]
    a, b, i, j = lay.orb
    t = hst["V"] / (-(f[a] + f[b] + f[lay.p + i] + f[lay.p + j]))
    hist = [(hst["V"] * t).sum()]
    for _ in range(6):
        t, e, _ = emu.step(lay, hst, t, 0.5)
        hist.append(e)
    np.testing.assert_allclose(hist, href, rtol=1e-12)


def test_violations_are_refused():
    pw, fock, v, lay = _setup(2, 4)
    # a v_pphh element outside every pair sector
    a, b, i, j = (x.reshape(-1) for x in np.meshgrid(*(np.arange(n) for n in
                                                       (lay.p, lay.p, lay.h, lay.h)),
                                                     indexing="ij"))
    out = np.nonzero((lay.element(a, b, i, j) < 0) & (a < b) & (i < j))[0][0]
    bad = [x.copy() for x in v]
    bad[2][a[out], b[out], i[out], j[out]] = 1e-3
    with pytest.raises(SectorViolation, match="v_pphh"):
        pack_hamiltonian(lay, fock, bad)
    # a v_phhh element that does not conserve the labels
    lp, lh, mod = pw.labels()
    tot = (lp[:, None, None, None] + lh[None, :, None, None] - lh[None, None, :, None]
           - lh[None, None, None, :])
    tot[..., :3] %= 2
    viol = tuple(np.argwhere(np.any(tot != 0, axis=-1))[0])
    bad = [x.copy() for x in v]
    bad[4][viol] = 1e-3
    with pytest.raises(SectorViolation, match="v_phhh"):
        pack_hamiltonian(lay, fock, bad)
    f2 = [fock[0].copy(), fock[1].copy(), fock[2].copy()]
    f2[0][0, 1] = 0.1
    with pytest.raises(SectorViolation, match="diagonal"):
        pack_hamiltonian(lay, f2, v)
    f3 = [fock[0], fock[1] + 0.1, fock[2]]
    with pytest.raises(SectorViolation, match="f_ph"):
        pack_hamiltonian(lay, f3, v)
    lp, lh, mod = pw.labels()
    with pytest.raises(ValueError, match="identify"):
        SectorLayout(lp[:, :3], lh[:, :3], mod[:3])      # momentum alone does not


@pytest.mark.parametrize("L,A", [(2, 4), (3, 28)])
def test_contact_fill_specs_reproduce_the_packed_hamiltonian(L, A):
    """The direct plane-wave build (nl_contact_fill's statement over contact_block_specs) equals
    the packing of the generator's dense blocks, element for element."""
    from nulattice_amd.sectors import contact_block_specs
    pw, fock, v, lay = _setup(L, A)
    hst = pack_hamiltonian(lay, fock, v)
    specs = contact_block_specs(lay, pw)
    for name in ("V", "VPP", "VHH", "VPH"):
        got = emu.contact_fill(*specs[name], pw.g)
        np.testing.assert_array_equal(got, hst[name])


@pytest.mark.parametrize("L,A", [(3, 4), (3, 28)])
def test_fused_ring_statement_equals_the_grouped_products(L, A):
    """nl_sector_ring's pair table reproduces the two grouped ring products (W, then Spp) on
    every crossed sector, self-conjugate ones included."""
    pw, fock, v, lay = _setup(L, A)
    hst = pack_hamiltonian(lay, fock, v)
    assert lay.ring_fused
    t = lay.pack(_random_t2(lay, 5))
    Tr = emu.gather(lay.ring_map, t)
    VPr = emu.gather(lay.ring_map, hst["V"])
    W = np.zeros(lay.nw)
    emu.gemm(lay.desc_w, 0.5, VPr, Tr, -1.0, W, hst["VPH"])
    S = np.zeros(lay.nring)
    emu.gemm(lay.desc_s, 1.0, Tr, W, 0.0, S)
    S2 = emu.ring_fused(lay.ring_pairs, lay.ring_map, t, VPr, hst["VPH"], np.zeros(lay.nring))
    np.testing.assert_allclose(S2, S, rtol=1e-13, atol=1e-16)
    covered = np.zeros(lay.nring, bool)
    for rx, ry, wx, wy, n1, n2, self_ in lay.ring_pairs:
        covered[rx:rx + n1 * n2] = True
        covered[ry:ry + n1 * n2] = True
    assert covered.all()


@pytest.mark.parametrize("L,A", [(2, 4), (3, 28)])
def test_reverse_sweep_is_the_transpose_of_the_residual(L, A):
    """<lam, J d> = <J^T lam, d> for the packed residual (J d by central differences, exact for
    the quadratic residual up to rounding), at random sector amplitudes and directions."""
    pw, fock, v, lay = _setup(L, A)
    hst = pack_hamiltonian(lay, fock, v)
    rng = np.random.default_rng(7)
    t = lay.pack(_random_t2(lay, 8))
    d = rng.standard_normal(lay.n)
    lam = rng.standard_normal(lay.n)
    eps = 1e-3
    Jd = (emu.step(lay, hst, t + eps * d, 0.5)[2] - emu.step(lay, hst, t - eps * d, 0.5)[2]) / (2 * eps)
    a, b = emu.residual_vjp(lay, hst, t, lam)
    lhs, rhs = lam @ Jd, (a + b) @ d
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1e-300)
    # reverse-sweep maps: each packed element has four crossed images, each Spp position is read
    # by at most one assembly term
    dec = lambda m: np.where(m >= 0, m, -m - 2)
    assert np.all(lay.ring_map[dec(lay.ring_inv)] != -1)
    used = lay.sg_inv[lay.sg_inv != -1]
    assert np.unique(dec(used)).size * 4 == used.size


def _random_labelled_hamiltonian(p, h, seed, mod=4):
    """A random antisymmetric Hamiltonian conserving random state-identifying labels
    (component 0 taken mod ``mod``, component 1 unbounded), holes first."""
    from tests.helpers import random_antisym_hamiltonian
    rng = np.random.default_rng(seed)
    n = p + h
    while True:
        lab = np.stack([rng.integers(0, mod, n), rng.integers(-2, 3, n)], 1)
        if len(np.unique(lab, axis=0)) == n:
            break
    _, _, V, F = random_antisym_hamiltonian(p, h, seed=seed + 100)
    tot = lab[:, None, None, None, :] + lab[None, :, None, None, :] \
        - lab[None, None, :, None, :] - lab[None, None, None, :, :]
    tot[..., 0] %= mod
    V = np.where(np.any(tot != 0, axis=-1), 0.0, V)
    F = np.diag(np.diag(F))
    o, v = slice(0, h), slice(h, n)
    fock = [F[v, v].copy(), F[v, o].copy(), F[o, o].copy()]
    dense = [V[v, v, v, v].copy(), V[v, v, v, o].copy(), V[v, v, o, o].copy(),
             V[v, o, v, o].copy(), V[v, o, o, o].copy(), V[o, o, o, o].copy()]
    return fock, dense, lab[h:], lab[:h], np.array([mod, 0])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sector_step_on_random_labelled_hamiltonians(seed):
    """Beyond plane waves: random label-conserving Hamiltonians (labels mod 4 and unbounded,
    self-conjugate crossed sectors included) stepped by the sector statement against the
    oracle's dense t2Iter (1e-12), t1 staying exactly 0 in the oracle."""
    p, h = 10, 4
    fock, v, lp, lh, mod = _random_labelled_hamiltonian(p, h, seed)
    lay = SectorLayout(lp, lh, mod)
    hst = pack_hamiltonian(lay, fock, v)
    t2 = _random_t2(lay, seed + 7)
    t1 = np.zeros((p, h))
    f_pp, f_ph, f_hh = fock
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = v
    assert not cc.t1Iter(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, v_ppph,
                         sparse=False).any()
    new2 = cc.t2Iter(t1, t2, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh, v_ppph, v_hhhh,
                     sparse=False)
    ref = 0.5 * t2 + 0.5 * new2
    np.testing.assert_allclose(lay.unpack(lay.pack(ref)), ref, atol=1e-15)
    tn, e, _ = emu.step(lay, hst, lay.pack(t2), 0.5)
    np.testing.assert_allclose(tn, lay.pack(ref), rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    assert e == pytest.approx(cc.ccsd_energy(f_ph, v_pphh, ref, t1), rel=1e-12)
    # and the reverse sweep is the transpose there too
    rng = np.random.default_rng(seed)
    d, lam = rng.standard_normal(lay.n), rng.standard_normal(lay.n)
    eps = 1e-3
    tp = lay.pack(t2)
    Jd = (emu.step(lay, hst, tp + eps * d, 0.5)[2] - emu.step(lay, hst, tp - eps * d, 0.5)[2]) / (2 * eps)
    a, b = emu.residual_vjp(lay, hst, tp, lam)
    assert abs(lam @ Jd - (a + b) @ d) <= 1e-10 * abs(lam @ Jd)

"""GPU parity of the symmetry-sector form: its kernels against the NumPy statement of their
C-ABI contracts, and the sector solver against the oracle and the dense GPU solver."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import _lib, ccm, ops  # noqa: E402
from nulattice_amd.planewave import PlaneWave, planewave_hamiltonian, shell_holes  # noqa: E402
from nulattice_amd.sectors import SectorHamiltonian, SectorLayout, _tiles  # noqa: E402
from oracle import ccsd_ref as cc  # noqa: E402
from tests import sector_emulator as emu  # noqa: E402


def dev(x, dt=torch.float64):
    return torch.tensor(np.ascontiguousarray(x), dtype=dt, device="cuda")


def host(t):
    return t.detach().cpu().numpy()


def _problems(rng, nprob, integer, small=False):
    """Random problem list over three flat buffers, with both unit-stride orientations of A and
    B, sizes straddling the 64 x 32 x 32 tile (small: outputs up to 32 x 32, K up to 300)."""
    desc, oa, ob, oc = [], 0, 0, 0
    for _ in range(nprob):
        M, N, K = (int(rng.integers(1, x)) for x in ((33, 33, 300) if small else (150, 70, 90)))
        ta, tb = rng.integers(0, 2, 2)
        sAm, sAk = (K, 1) if ta else (1, M)
        sBk, sBn = (N, 1) if tb else (1, K)
        desc.append([M, N, K, oa, sAm, sAk, ob, sBk, sBn, oc, N, 1])
        oa += M * K + int(rng.integers(0, 5))
        ob += K * N + int(rng.integers(0, 5))
        oc += M * N + int(rng.integers(0, 5))
    desc = np.array(desc, np.int64)
    gen = (lambda n: rng.integers(-4, 5, n).astype(np.float64)) if integer else \
        (lambda n: rng.standard_normal(n))
    return desc, gen(oa), gen(ob), gen(oc), gen(oc)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("integer,beta,use_d", [(True, 0.0, False), (True, -1.0, True),
                                                (False, 0.5, False), (False, 1.0, True)])
def test_grouped_gemm(integer, beta, use_d, mode):
    """Every kernel mode (64x32 tiles; 32x32 and 64x8 tiles with K split over wavefronts) on
    the same random problem lists."""
    rng = np.random.default_rng(7)
    desc, A, B, C, D = _problems(rng, 40, integer, mode in (1, 3))
    til = _tiles(desc, mode)
    alpha = 1.0 if integer else 0.75
    # every device buffer stays bound until the call has been read back: a temporary freed
    # before the launch hands its memory to the next upload, which would overwrite it first
    Cg, Dg, Ag, Bg = dev(C), dev(D), dev(A), dev(B)
    descg, tilg = dev(desc, torch.int64), dev(til, torch.int32)
    _lib.call("nl_gemm_grouped_f64", mode, til.shape[0], ops._p(descg), ops._p(tilg),
              ctypes.c_double(alpha), ops._p(Ag), ops._p(Bg), ctypes.c_double(beta),
              ops._p(Dg if use_d else None), ops._p(Cg), ops._stream())
    torch.cuda.synchronize()
    ref = C.copy()
    emu.gemm(desc, alpha, A, B, beta, ref, D if use_d else None)
    if integer:
        np.testing.assert_array_equal(host(Cg), ref)      # exact: every partial sum is an integer
    else:
        np.testing.assert_allclose(host(Cg), ref, rtol=1e-13, atol=1e-13)


def _pw(L, A, C0=-4.0, D0=1.0, sparse=True):
    pw = PlaneWave(L, shell_holes(L, A), C0, D0)
    E0, fock, v = pw.hamiltonian(sparse=sparse)
    return pw, fock, v


def test_sector_kernels_match_their_statement():
    """Dressing, ladder and ring products, crossed-channel gather and the fused assembly of one
    step (L=3, A=28) against tests/sector_emulator.py at random sector amplitudes."""
    pw, fock, v = _pw(3, 28, sparse=False)
    SH = SectorHamiltonian(fock, v, *pw.labels())
    lay = SH.layout
    from nulattice_amd.sectors import pack_hamiltonian
    hst = pack_hamiltonian(lay, fock, v)
    rng = np.random.default_rng(3)
    R = lay.unpack(rng.standard_normal(lay.n) * 1e-2)
    R = R - R.transpose(1, 0, 2, 3)
    t = lay.pack(R - R.transpose(0, 1, 3, 2))
    Bf = SH.buffers()
    tn = torch.empty(lay.n, dtype=torch.float64, device="cuda")
    H2 = torch.empty_like(tn)
    SH.update(dev(t), tn, Bf, 0.5, H2out=H2)
    ref_tn, ref_e, ref_H = emu.step(lay, hst, t, 0.5)
    np.testing.assert_allclose(host(H2), ref_H, rtol=1e-12, atol=1e-14 * np.abs(ref_H).max())
    np.testing.assert_allclose(host(tn), ref_tn, rtol=1e-12, atol=1e-14 * np.abs(ref_tn).max())
    assert float(Bf["E"].item()) == pytest.approx(ref_e, rel=1e-12)
    # round trip through the reference layout
    np.testing.assert_array_equal(host(SH.pack(SH.unpack(dev(t)))), t)
    np.testing.assert_array_equal(host(SH.unpack(dev(t))), lay.unpack(t))


@pytest.mark.parametrize("L,A", [(2, 4), (3, 4)])
def test_sector_solver_matches_oracle(L, A):
    """Six legacy steps of the sector solver agree with the oracle's dense solver (1e-12)."""
    pw, fock, v = _pw(L, A, sparse=False)
    href = []
    cc.ccsd_solver(fock, v, maxSteps=6, max_diis=0, sparse=False, history=href)
    hist = []
    e, t1, t2 = ccm.ccsd_solver(fock, v, maxSteps=6, max_diis=0, sparse=False, history=hist,
                                sectors=pw.labels())
    np.testing.assert_allclose(hist, href, rtol=1e-12)
    assert float(t1.abs().max()) == 0.0
    assert tuple(t2.shape) == (pw.p, pw.p, pw.h, pw.h)


def test_sector_solver_matches_dense_gpu_solver_with_diis():
    """L=3, A=28, pulsed DIIS to convergence: per-step energies agree with the dense GPU solver
    (1e-12 through the Jacobi steps; after the first extrapolation the ill-conditioned DIIS
    solve amplifies round-off, as between the reference's own dense and sparse runs, so 1e-6),
    and both land on the same fixed point."""
    pw, fock, v = _pw(3, 28)
    hd, hs = [], []
    ed, _, t2d = ccm.ccsd_solver(fock, v, eps=1e-10, diis="pulsed", history=hd)
    es, t1s, t2s = ccm.ccsd_solver(fock, v, eps=1e-10, diis="pulsed", history=hs,
                                   sectors=pw.labels())
    n = min(len(hd), len(hs))
    np.testing.assert_allclose(hs[:11], hd[:11], rtol=1e-12)
    np.testing.assert_allclose(hs[:n], hd[:n], rtol=1e-6)
    assert es == pytest.approx(ed, rel=1e-8)
    assert float((t2s - t2d).abs().max()) < 1e-6 * float(t2d.abs().max())


def test_sector_solver_L5_A28_against_dense_fp64():
    """The benchmark workload (L=5, A=28, 472 particles): five legacy steps agree with the dense
    GPU FP64 solver to 1e-12, and the amplitudes to 1e-12 of their scale."""
    E0, fock, v = planewave_hamiltonian(5, 28, -4.0, 1.0)
    pw = PlaneWave(5, shell_holes(5, 28), -4.0, 1.0)
    hd, hs = [], []
    _, _, t2d = ccm.ccsd_solver(fock, v, maxSteps=5, max_diis=0, history=hd)
    info = {}
    _, _, t2s = ccm.ccsd_solver(fock, v, maxSteps=5, max_diis=0, history=hs,
                                sectors=pw.labels(), info=info)
    np.testing.assert_allclose(hs, hd, rtol=1e-12)
    assert float((t2s - t2d).abs().max()) < 1e-12 * float(t2d.abs().max())
    assert info["sector_amplitudes"] == 49268


def test_sector_ccs_and_t1initial():
    pw, fock, v = _pw(2, 4)
    e, t1, t2 = ccm.ccsd_solver(fock, v, ccs=True, sectors=pw.labels())
    assert e == 0.0 and float(t2.abs().max()) == 0.0
    with pytest.raises(ValueError):
        ccm.ccsd_solver(fock, v, t1initial=np.ones((pw.p, pw.h)), sectors=pw.labels())
    with pytest.raises(ValueError):
        ccm.ccsd_solver(fock, v, precision="bf16", sectors=pw.labels())


def test_graph_replays_equal_kernel_launches():
    """The HIP-graph step (two captured graphs, replayed alternately, with the DIIS copies, dots
    and host read-backs between replays) gives bit-identical trajectories and amplitudes to the
    same kernels launched one by one (L=3, A=28, pulsed DIIS, 40 steps)."""
    from nulattice_amd import sectors
    pw, fock, v = _pw(3, 28)
    SH = SectorHamiltonian(fock, v, *pw.labels())
    res = {}
    for graph in (False, True):
        hist = []
        e, _, t2 = sectors.solve(SH, eps=1e-14, maxSteps=40, diis="pulsed", history=hist,
                                 graph=graph)
        res[graph] = (hist, host(t2))
    assert res[True][0] == res[False][0]
    np.testing.assert_array_equal(res[True][1], res[False][1])


@pytest.mark.parametrize("L,A", [(3, 28), (5, 28)])
def test_direct_planewave_build_equals_packed_blocks(L, A):
    """SectorHamiltonian.from_planewave (blocks written on the GPU) equals the packing of the
    generator's dense blocks element for element, and its solve gives the same energies."""
    pw, fock, v = _pw(L, A)
    SH = SectorHamiltonian(fock, v, *pw.labels())
    SD = SectorHamiltonian.from_planewave(pw)
    for name in ("V", "VPP", "VHH", "VPHPH", "VPr", "base"):
        assert torch.equal(getattr(SD, name), getattr(SH, name)), name
    h1, h2 = [], []
    ccm.ccsd_solver(None, None, maxSteps=4, max_diis=0, hamiltonian=SH, history=h1)
    ccm.ccsd_solver(None, None, maxSteps=4, max_diis=0, hamiltonian=SD, history=h2)
    assert h1 == h2


def test_sector_reverse_sweep_matches_its_statement():
    """The GPU reverse sweep (J^T lam without its diagonal part) equals the NumPy statement at
    random amplitudes and Lambda vectors (L=3, A=28)."""
    from nulattice_amd import sectors
    from nulattice_amd.sectors import pack_hamiltonian
    pw, fock, v = _pw(3, 28)
    SH = SectorHamiltonian.from_planewave(pw)
    lay = SH.layout
    hst = pack_hamiltonian(lay, fock, v)
    rng = np.random.default_rng(4)
    R = lay.unpack(rng.standard_normal(lay.n) * 1e-2)
    t = lay.pack(R)
    lam = rng.standard_normal(lay.n)
    Bf = sectors.lambda_buffers(SH)
    td, ld = dev(t), dev(lam)
    SH.residual(td, Bf)
    out = torch.empty_like(td)
    sectors.residual_vjp(SH, td, ld, Bf, out)
    ref, _ = emu.residual_vjp(lay, hst, t, lam)
    np.testing.assert_allclose(host(out), ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_sector_gradient_matches_dense_adjoint():
    """dE/dC0 and dE/dD0 of the sector form (packed Lambda equations) equal the dense adjoint's
    (which is pinned to finite differences) at L=3, A=28."""
    from nulattice_amd import adjoint
    res = {}
    for form in ("dense", "sectors"):
        C0 = torch.tensor(-4.0, dtype=torch.float64, requires_grad=True)
        D0 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
        E = adjoint.planewave_ccsd_energy(C0, D0, 3, 28, eps=1e-13, form=form,
                                          lambda_eps=1e-12)
        E.backward()
        res[form] = (float(E.detach()), float(C0.grad), float(D0.grad))
    d, s = res["dense"], res["sectors"]
    assert s[0] == pytest.approx(d[0], rel=1e-11)
    assert s[1] == pytest.approx(d[1], rel=1e-8)
    assert s[2] == pytest.approx(d[2], rel=1e-8)


def test_sector_gradient_L5_A28_against_finite_differences():
    """The benchmark size: the sector-form adjoint gradient against central differences of
    converged sector-form energies (h = 1e-4; E is smooth in C0, D0): 1e-6 relative."""
    from nulattice_amd import adjoint
    C0 = torch.tensor(-4.0, dtype=torch.float64, requires_grad=True)
    D0 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    info = {}
    E = adjoint.planewave_ccsd_energy(C0, D0, 5, 28, eps=1e-13, form="sectors", info=info)
    E.backward()
    h = 1e-4

    def energy(c, d):
        return float(adjoint.planewave_ccsd_energy(c, d, 5, 28, eps=1e-13, form="sectors"))
    fd_c = (energy(-4.0 + h, 1.0) - energy(-4.0 - h, 1.0)) / (2 * h)
    fd_d = (energy(-4.0, 1.0 + h) - energy(-4.0, 1.0 - h)) / (2 * h)
    assert float(C0.grad) == pytest.approx(fd_c, rel=1e-6)
    assert float(D0.grad) == pytest.approx(fd_d, rel=1e-6)
    assert info["lambda_residual"] <= 1e-10


def test_grouped_ring_path_equals_fused_ring():
    """The ring through the grouped GEMMs (the path for crossed sectors wider than 32 pairs)
    and through the fused kernel give the same step (L=3, A=28)."""
    pw, fock, v = _pw(3, 28)
    SH = SectorHamiltonian.from_planewave(pw)
    lay = SH.layout
    rng = np.random.default_rng(9)
    t = dev(lay.pack(lay.unpack(rng.standard_normal(lay.n) * 1e-2)))
    out = []
    for fused in (True, False):
        if not fused:
            SH.ring_pairs = None
        Bf = SH.buffers()
        tn = torch.empty_like(t)
        SH.update(t, tn, Bf, 0.5)
        out.append((host(tn), float(Bf["E"].item())))
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=1e-13, atol=1e-17)
    assert out[1][1] == pytest.approx(out[0][1], rel=1e-13)


@pytest.mark.parametrize("diis,delta", [("soa", 0.0), ("legacy", 0.3)])
def test_sector_solver_options_match_dense_solver(diis, delta):
    """The SoA DIIS variant and a nonzero delta (shifted initial denominators) give the dense
    solver's trajectory (L=3, A=4): 1e-9 through the Jacobi steps, 1e-6 once DIIS extrapolates
    (its ill-conditioned solve amplifies round-off), and the same fixed point."""
    pw, fock, v = _pw(3, 4)
    hd, hs = [], []
    ed, _, _ = ccm.ccsd_solver(fock, v, eps=1e-10, maxSteps=40, diis=diis, delta=delta,
                               history=hd)
    es, _, _ = ccm.ccsd_solver(fock, v, eps=1e-10, maxSteps=40, diis=diis, delta=delta,
                               history=hs, sectors=pw.labels())
    n = min(len(hd), len(hs))
    np.testing.assert_allclose(hs[:11], hd[:11], rtol=1e-9)
    np.testing.assert_allclose(hs[:n], hd[:n], rtol=1e-6)
    assert es == pytest.approx(ed, rel=1e-8)
    assert abs(len(hd) - len(hs)) <= 2

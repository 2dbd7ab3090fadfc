"""GPU build of the dense-form plane-wave Hamiltonian (nl_pw_dense, nl_pw_pair_fill, nl_pw_coo
and the device-sorted segment plans) against the host generator plus the host plan builder:
every block and every plan must be the same bits, and the CCSD trajectories equal."""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import adjoint, ccm, engine  # noqa: E402
from nulattice_amd.planewave import PlaneWave, shell_holes  # noqa: E402


def same_bits(a, b, what):
    assert a.dtype == b.dtype and tuple(a.shape) == tuple(b.shape), what
    if a.dtype == torch.float64:
        a, b = a.view(torch.int64), b.view(torch.int64)
    assert torch.equal(a, b), what


def check_pairblocks(x, y, what):
    assert (x.nb, x.mb, x.kb, x.nnz) == (y.nb, y.mb, y.kb, y.nnz), what
    for f in ("V", "rows", "cols"):
        same_bits(getattr(x, f), getattr(y, f), f"{what}.{f}")


def check_seg(x, y, what):
    assert x.S == y.S, what
    for f in ("ptr", "src", "val", "dst"):
        same_bits(getattr(x, f), getattr(y, f), f"{what}.{f}")


def build_both(L, A, C0, D0):
    pw = PlaneWave(L, shell_holes(L, A), C0, D0)
    E0, fock, v = pw.hamiltonian()
    Hh = engine.Hamiltonian(fock, v, device="cuda")
    E0d, fockd, Hd = pw.device_hamiltonian("cuda")
    torch.cuda.synchronize()
    assert E0 == E0d
    return Hh, Hd


@pytest.mark.parametrize("L,A,C0,D0", [(3, 28, -4.0, 1.0), (3, 4, -3.7, 0.0), (4, 28, -4.2, 0.8)])
def test_device_build_bitwise(L, A, C0, D0):
    print(f"device build L={L} A={A}", flush=True)
    Hh, Hd = build_both(L, A, C0, D0)
    for name in ("f_pp", "f_ph", "f_hh", "v_pphh", "v_phph", "v_phhh", "v_hhhh", "vP", "Wc",
                 "v_phph_bkic", "v_pphh_T"):
        same_bits(getattr(Hh, name), getattr(Hd, name), name)
    check_pairblocks(Hh.pppp, Hd.pppp, "pppp")
    ph, pd = Hh.ppph, Hd.ppph
    assert ph.n == pd.n and ph.n > 0
    for name in ("ret0", "ret1", "ret2", "ret2p", "ret4"):
        check_seg(getattr(ph, name), getattr(pd, name), name)
    check_pairblocks(ph.r56, pd.r56, "r56")
    # the Lambda-equation plans built from the device COO list
    ah, ad = ph.adjoint(adjoint._transpose_blocks), pd.adjoint(adjoint._transpose_blocks)
    for name in ("ret0T", "ret1T", "ret2T", "ret4T"):
        check_seg(ah[name], ad[name], name)
    check_pairblocks(ah["r56T"], ad["r56T"], "r56T")


def test_device_build_trajectory():
    """Same Hamiltonian bits, same kernels: the FP64 iteration energies agree exactly."""
    Hh, Hd = build_both(3, 28, -4.1, 1.2)
    runs = []
    for H in (Hh, Hd):
        hist = []
        e, _, _ = ccm.ccsd_solver(None, None, eps=1e-12, maxSteps=12, max_diis=0, hamiltonian=H,
                                  history=hist)
        runs.append(hist)
    assert len(runs[0]) == len(runs[1]) > 5
    assert runs[0] == runs[1]


def test_device_build_L5_time():
    """L=5 A=28 (p=472, h=28): the whole dense-form Hamiltonian, 6.3 GB of blocks and plans,
    built on the device; the build is timed after a warm-up build at L=3."""
    PlaneWave(3, shell_holes(3, 28), -4.0, 1.0).device_hamiltonian("cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pw = PlaneWave(5, shell_holes(5, 28), -4.0, 1.0)
    E0, fock, H = pw.device_hamiltonian("cuda")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"L=5 device build {dt:.3f} s, pppp nnz {H.pppp.nnz}, ppph nnz {H.ppph.n}", flush=True)
    assert H.pppp.nnz == 37227120 and H.ppph.n > 0
    # spot check against the host generator: the dense blocks at sampled rows
    rng = np.random.default_rng(0)
    P, Hs = pw.parts, pw.holes
    rows = rng.integers(0, pw.p, 6)
    ref = pw.block([P[r] for r in rows], P, Hs, Hs)
    same_bits(H.v_pphh[torch.as_tensor(rows, device="cuda")].cpu(),
              torch.from_numpy(ref), "v_pphh rows")
    assert dt < 2.0

"""The anchored ring intermediate of the fp16x3 form (engine.RING_ANCHOR): every RING_ANCHOR-th
step computes W_A = Wc + 1/2 v t2_A with split operands and keeps it with t2_A's ring image; the
steps in between compute W = W_A + 1/2 v FP16(t2 - t2_A) (one partial product).

Pinned by: nl_split_delta16 against a torch statement of its arithmetic (bit for bit); the
anchored W against the full W at the same amplitudes, with an error that shrinks in proportion
to the amplitudes' difference from the anchor (the property the fixed point relies on: at
convergence the difference, hence the error, vanishes); and whole solves whose converged energy
equals the unanchored one to 1e-9 (the L=5 A=28 clause itself is tests/test_gpu_split.py).
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import _lib, ccm, engine, ops  # noqa: E402
from nulattice_amd.planewave import planewave_device_hamiltonian  # noqa: E402


def P(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def split16(x):
    """[hi | lo] FP16 rows of an FP32 array (x = hi + lo to FP16 pairs)."""
    hi = x.half()
    lo = (x - hi.float()).half()
    return torch.cat([hi, lo], dim=1).contiguous()


@pytest.mark.parametrize("rows,K,ldd", [(37, 96, 128), (200, 824, 832)])
def test_split_delta16_is_its_fp32_statement(rows, K, ldd):
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(rows, K, generator=g, device="cuda") * 300
    xa = x + torch.randn(rows, K, generator=g, device="cuda") * 0.01
    X, XA = split16(x), split16(xa)
    D = torch.full((rows * ldd,), -1, dtype=torch.int16, device="cuda")
    assert _lib.load().nl_split_delta16(rows, K, 2 * K, P(X), P(XA), P(D), ldd, S()) == 0
    torch.cuda.synchronize()
    a = X[:, :K].float() + X[:, K:].float()
    b = XA[:, :K].float() + XA[:, K:].float()
    ref = (a - b).half().view(torch.int16)
    got = D.view(rows, ldd)
    assert torch.equal(got[:, :K], ref)
    assert bool((got[:, K:] == 0).all())          # the pad to the K step is zero


def test_split_delta16_rejects_unaligned_shapes():
    X = torch.zeros(64, dtype=torch.int16, device="cuda")
    assert _lib.load().nl_split_delta16(2, 12, 24, P(X), P(X), P(X), 16, S()) == -1   # K % 8
    assert _lib.load().nl_split_delta16(2, 8, 12, P(X), P(X), P(X), 8, S()) == -1     # ldx < 2K


def _w_after(H, B, t1, t2, prec):
    """W (FP32, as the Spp product reads it) after the T1 and T2 parts of a step at (t1, t2),
    with the split images made as engine.update makes them."""
    with ops.lp_reuse(t2):
        dt = engine._fast_split(H, prec)
        engine._t2_split_images(H, t2, dt, B)
        engine.t1_part(H, t1, t2, B, prec)
        engine.t2_part(H, t1, t2, B, prec)
        B.t2_split = B.t2_pairs = B.N56r = B.t2_ring = None
    return B.acc32().W.clone()


def test_anchored_w_error_shrinks_with_the_difference():
    """Plane-wave L=3 A=28: anchor at t2_A, then W at t2_A + eps d for eps = 1e-1, 1e-2, 1e-4,
    against the full split product at the same amplitudes.  The error stays below 2^-9 of the
    difference's own contribution plus the FP32 storage floor of W (16 ulp of its largest
    element), and falls with eps about in proportion until it reaches that floor."""
    _, _, H = planewave_device_hamiltonian(3, 28, -4.0, 1.0, device="cuda")
    prec = ccm._PREC["fp16x3"]
    p, h = H.p, H.h
    B = engine.Buffers(p, h, H.device)
    t1 = torch.zeros(p, h, dtype=torch.float64, device="cuda")
    t2 = torch.zeros(p, p, h, h, dtype=torch.float64, device="cuda")
    engine.init_amplitudes(H, B, t1, t2, 0.0)
    g = torch.Generator(device="cuda").manual_seed(7)
    d = torch.randn(p, p, h, h, generator=g, dtype=torch.float64, device="cuda")
    d = d - d.transpose(0, 1)
    d = (d - d.transpose(2, 3)) * float(t2.abs().max())
    anchor = engine.RING_ANCHOR
    try:
        engine.RING_ANCHOR = 5
        B.ring_step = 0
        WA = _w_after(H, B, t1, t2, prec)                          # the anchor step (full)
        assert B.ring_step == 1
        errs, floor = [], 0.0
        for eps in (1e-1, 1e-2, 1e-4):
            tt = t2 + eps * d
            B.ring_step = 1
            Wan = _w_after(H, B, t1, tt, prec)                     # anchored
            assert B.ring_step == 2
            engine.RING_ANCHOR = 0
            Bf = engine.Buffers(p, h, H.device)
            Wfull = _w_after(H, Bf, t1, tt, prec)                  # full split product
            engine.RING_ANCHOR = 5
            dW = float((Wfull - WA).abs().max())                   # the difference's product
            err = float((Wan - Wfull).abs().max())
            floor = 16 * 2.0 ** -24 * float(Wfull.abs().max())
            errs.append(err)
            print(f"eps {eps:g}: |W - W_A| {dW:.3e}, anchored error {err:.3e}, floor {floor:.1e}")
            assert err <= 2.0 ** -9 * dW + floor
        assert errs[1] < 0.2 * errs[0]                           # ~1/10 per decade of eps
        assert errs[2] <= floor + 2.0 ** -9 * 1e-4 / 1e-1 * errs[0] / 2.0 ** -12
    finally:
        engine.RING_ANCHOR = anchor


@pytest.mark.parametrize("period", [2, 5])
def test_anchored_solve_converges_to_the_unanchored_energy(period):
    """Plane-wave L=3 A=28: the fp16x3 fixed point with the anchored ring intermediate equals the
    unanchored one to 1e-9 relative (the split products' own precision is ~1e-8)."""
    _, _, H = planewave_device_hamiltonian(3, 28, -4.0, 1.0, device="cuda")
    es = {}
    anchor = engine.RING_ANCHOR
    try:
        for n in (0, period):
            engine.RING_ANCHOR = n
            e, _, _ = ccm.ccsd_solver(None, None, eps=1e-11, maxSteps=400, max_diis=0,
                                      hamiltonian=H, precision="fp16x3", refine=False)
            es[n] = e
    finally:
        engine.RING_ANCHOR = anchor
    assert abs(es[period] - es[0]) <= 1e-9 * abs(es[0]), es

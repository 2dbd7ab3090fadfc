"""Differentiation form: CC Lambda equations and dE/d(coupling) on the GPU.

At converged amplitudes t the CCSD residual H(t; h) vanishes (H = [H1, H2], the raw residuals of
t1Iter/t2Iter, coupled_cluster.py:342-545) and the energy is E = E_ref(h) + E_corr(t; h)
(ccsd_energy, coupled_cluster.py:265-286).  Every term of H and of E_corr carries exactly one
Hamiltonian matrix element, so both are linear in h.  For a parameter theta of h,

    dE/dtheta = dE_ref/dtheta + E_corr(t; h_theta) + lambda . H(t; h_theta),
    J^T lambda = -dE_corr/dt          (the Lambda equations, J = dH/dt at t),

with h_theta = dh/dtheta.  For the lattice contact theory every normal-ordered block is affine in
(C0, D0), so h_theta is the Hamiltonian of ``PlaneWave.coupling_derivative``.

T2 amplitudes are antisymmetric in (a,b) and in (i,j); the Lambda equations are solved on that
subspace (lambda2 antisymmetric, residual projected), where the Jacobi denominators of the
amplitude iteration are the diagonal of J.

J^T lambda is applied by ``residual_vjp``: a hand-written reverse sweep over the same term
regrouping as engine.t1_part / engine.t2_part, where each contraction's adjoint is again one
contraction launch (operand roles transposed) that accumulates into the gradient buffer, the
sparse v_ppph / v_pppp terms use transposed segment plans and transposed pair blocks, and the
T2 antisymmetrisers use their own transpose kernel (nl_assemble_t2_adj).  The sweep reads the
forward intermediates of engine.Buffers at the converged amplitudes; nothing is recorded by a
tracing autograd engine.
"""
import ctypes
import math
import time

import torch

from . import _lib, engine, ops
from .ops import contract, permute_into

F64 = torch.float64


def _transpose_blocks(V):
    return ops.copy(V.permute(0, 2, 1))


def _pppp_T(ham):
    t = getattr(ham.pppp, "_T", None)
    if t is None:
        t = ham.pppp.transposed(_transpose_blocks)
        ham.pppp._T = t
    return t


class AdjointBuffers:
    """Lambda amplitudes and the adjoints of the residual's intermediates (one set per size)."""

    def __init__(self, p, h, device):
        z = lambda *s: torch.zeros(s, dtype=F64, device=device)
        self.p, self.h = p, h
        self.l1, self.l1n, self.gt1 = z(p, h), z(p, h), z(p, h)
        self.l2, self.l2n, self.gt2 = z(p, p, h, h), z(p, p, h, h), z(p, p, h, h)
        self.R2 = z(p, h, p, h)          # 1/4 of the T2 residual, as [a][i][b][j]
        self.gTr = z(p, h, p, h)         # adjoint of t2 through its [a][i][c][k] layout
        self.gSab = z(p, p, h, h)
        self.gSij, self.gSpp = z(p, h, p, h), z(p, h, p, h)
        self.gtau, self.gW = z(p, p, h, h), z(p, h, p, h)
        self.gX1pp, self.gX1hh, self.gX2pp, self.gX2hh = z(p, p), z(h, h), z(p, p), z(h, h)
        self.gret1 = z(p, p)
        self.gs, self.gu, self.gw, self.gz = z(p, h), z(p, h), z(p, h), z(p, h)
        self.gq = z(h, h)
        self.gWhh, self.gzz = z(h, h, h, h), z(h, h, h, h)
        self.gN56, self.gU = z(p, h, h, h), z(p, h, h, h)
        self.gY, self.gXd = z(h, h, p, h), z(h, p, h, h)
        self.part = z(2 * int(_lib.load().nl_partials(p * p * h * h)) + 8)
        self.out = z(1)


def energy_gradient(ham, t1, t2, G):
    """G.gt1, G.gt2 = dE_corr/dt1, dE_corr/dt2 of E = f.t1 + 1/4 v.t2 + 1/2 v t1 t1."""
    permute_into(ham.f_ph, G.gt1)
    contract("abij,bj->ai", ham.v_pphh, t1, G.gt1, alpha=0.5)
    contract("baji,bj->ai", ham.v_pphh, t1, G.gt1, alpha=0.5)
    permute_into(ham.v_pphh, G.gt2, alpha=0.25)


def residual_vjp(ham, t1, t2, B, l1, l2, G):
    """G.gt1 += (J^T l)_1 and G.gt2 += (J^T l)_2 for the residual H(t) of engine.t1_part and
    engine.t2_part, with B holding their FP64 intermediates at (t1, t2).

    Forward term -> adjoint contributions (g* = adjoint of an intermediate):
      H2 = S0 + pAB Sab + pIJ Sij + pIJ pAB Spp       gS0 = l2, gSab = pAB l2, ... (transposed)
      Spp = t2.W + Y.t1 - t1.U                         ring GEMMs with W^T and t2^T
      W = Wc + 1/2 v.t2 + ret4(t1) - v_phhh.t1         gt2 += 1/2 v^T gW, ...
      S0 = 1/2 tau.Whh + 1/2 V.tau                     gtau, gWhh, transposed pair blocks
      Sab, Sij, X dressings, H1                        one transposed contraction per term
      tau = t2 + P(ij) t1 t1                           gt2 += gtau, gt1 += 4 contractions
    """
    p, h = ham.p, ham.h
    ph = p * h
    pl = ham.ppph
    adj = pl.adjoint(_transpose_blocks) if pl.n else None
    gt1, gt2 = G.gt1, G.gt2
    st = ops._stream()
    _lib.call("nl_assemble_t2_adj", p, h, ops._p(l2), ops._p(G.gSab), ops._p(G.gSij),
              ops._p(G.gSpp), st)
    # ---- Spp[a,i,b,j] = t2[a,c,i,k] W[c,k,b,j] + Y[i,l,a,j] t1[b,l] - t1[a,k] U[b,k,i,j]
    gSpp = G.gSpp
    contract("aibj,ckbj->aick", gSpp, B.W, G.gTr, beta=0.0, tag="adj_ring_t")
    contract("acik,aibj->ckbj", t2, gSpp, G.gW, beta=0.0, tag="adj_ring_W")
    contract("aibj,bl->ilaj", gSpp, t1, G.gY, beta=0.0)
    contract("ilaj,aibj->bl", B.Yilaj, gSpp, gt1)
    contract("aibj,bkij->ak", gSpp, B.Ubkij, gt1, alpha=-1.0)
    contract("ak,aibj->bkij", t1, gSpp, G.gU, alpha=-1.0, beta=0.0)
    # W[c,k,b,j] = Wc + 1/2 v[c,d,k,l] t2[d,b,l,j] + ret4 - v_phhh[c,j,k,l] t1[b,l]
    contract("ckdl,ckbj->dlbj", ham.vP, G.gW, G.gTr, alpha=0.5, tag="adj_ring_v")
    contract("cjkl,ckbj->bl", ham.v_phhh, G.gW, gt1, alpha=-1.0)
    if pl.n:
        engine._segsum(adj["ret4T"], h, G.gW, gt1, 1.0)
    # Y[i,l,a,j] = Xd[i,d,k,l] t2[a,d,k,j],  Xd[i,d,k,l] = v[c,d,k,l] t1[c,i]
    contract("ilaj,adkj->idkl", G.gY, t2, G.gXd, beta=0.0)
    contract("idkl,ilaj->adkj", B.Xidkl, G.gY, gt2)
    contract("cdkl,idkl->ci", ham.v_pphh, G.gXd, gt1)
    # U[b,k,i,j] = v_phph[b,k,c,i] t1[c,j]
    contract("bkci,bkij->cj", ham.v_phph, G.gU, gt1)
    # ---- Sij[a,b,i,j] = -q[l,i] t2[a,b,l,j] + X2hh[k,j] t2[a,b,i,k] + ret2(t1)
    gSij = G.gSij.permute(0, 2, 1, 3)
    contract("abij,ablj->li", gSij, t2, G.gq, alpha=-1.0, beta=0.0)
    contract("li,abij->ablj", B.q_hh, gSij, gt2, alpha=-1.0)
    contract("abij,abik->kj", gSij, t2, G.gX2hh, beta=0.0)
    contract("kj,abij->abik", B.X2hh, gSij, gt2)
    if pl.n:
        engine._segsum(adj["ret2T"], h, G.gSij, gt1, 1.0, ph, 1)
    contract("cikl,li->ck", ham.v_phhh, G.gq, gt1)          # q[l,i] = v_phhh[c,i,k,l] t1[c,k]
    # ---- Sab[a,b,i,j] = N56[b,k,i,j] t1[a,k] + ret1[a,d] t2[d,b,i,j] + X2pp[b,c] t2[a,c,i,j]
    gSab = G.gSab
    contract("abij,ak->bkij", gSab, t1, G.gN56, beta=0.0)
    contract("abij,bkij->ak", gSab, B.N56, gt1)
    if pl.n:
        contract("abij,dbij->ad", gSab, t2, G.gret1, beta=0.0)
        contract("ad,abij->dbij", B.ret1, gSab, gt2)
    contract("abij,acij->bc", gSab, t2, G.gX2pp, beta=0.0)
    contract("bc,abij->acij", B.X2pp, gSab, gt2)
    # ---- S0 = 1/2 tau[a,b,k,l] Whh[k,l,i,j] + 1/2 V[ab,cd] tau[cd,ij]
    contract("abij,klij->abkl", l2, B.Whh, G.gtau, alpha=0.5, beta=0.0)
    contract("abkl,abij->klij", B.tau, l2, G.gWhh, alpha=0.5, beta=0.0)
    engine._block_gemm(_pppp_T(ham), l2, G.gtau, 0.5, h * h, "adj_ladder")
    # Whh = v_hhhh + zz - zz^(ij) + 1/2 v[c,d,k,l] tau[c,d,i,j],
    # zz[k,l,i,j] = t1[c,i] v_phhh[c,j,k,l]
    contract("cdkl,klij->cdij", ham.v_pphh, G.gWhh, G.gtau, alpha=0.5)
    permute_into(G.gWhh, G.gzz)
    permute_into(G.gWhh.permute(0, 1, 3, 2), G.gzz, -1.0, 1.0)
    contract("cjkl,klij->ci", ham.v_phhh, G.gzz, gt1)
    # N56 = v_phhh + 1/2 R56(tau)
    if pl.n:
        engine._block_gemm(adj["r56T"], G.gN56, G.gtau, 0.5, h * h, "adj_r56")
    # ---- X2hh[k,j] = -f_hh - 1/2 v[c,d,k,l] t2[c,d,j,l] - f_ph[c,k] t1[c,j] - u[d,k] t1[d,j]
    contract("cdkl,kj->cdjl", ham.v_pphh, G.gX2hh, gt2, alpha=-0.5)
    contract("ck,kj->cj", ham.f_ph, G.gX2hh, gt1, alpha=-1.0)
    contract("kj,dj->dk", G.gX2hh, t1, G.gu, alpha=-1.0, beta=0.0)
    contract("dk,kj->dj", B.u_ph, G.gX2hh, gt1, alpha=-1.0)
    # X2pp[b,c] = f_pp - 1/2 v[c,d,k,l] t2[b,d,k,l] - f_ph[c,k] t1[b,k] - z[c,l] t1[b,l]
    contract("bc,cdkl->bdkl", G.gX2pp, ham.v_pphh, gt2, alpha=-0.5)
    contract("bc,ck->bk", G.gX2pp, ham.f_ph, gt1, alpha=-1.0)
    contract("bc,bl->cl", G.gX2pp, t1, G.gz, alpha=-1.0, beta=0.0)
    contract("bc,cl->bl", G.gX2pp, B.z_ph, gt1, alpha=-1.0)
    # ---- H1[a,i]
    contract("akci,ai->ck", ham.v_phph, l1, gt1, alpha=-1.0)            # A11
    contract("ai,ck->acik", l1, ham.f_ph, gt2)                           # A12
    contract("cikl,ai->cakl", ham.v_phhh, l1, gt2, alpha=-0.5)          # A13
    contract("dali,ai->dl", t2, l1, G.gs, beta=0.0)                      # A14
    contract("dl,ai->dali", B.s_ph, l1, gt2)
    contract("cdkl,dl->ck", ham.v_pphh, G.gs, gt1)
    if pl.n:
        engine._segsum(adj["ret0T"], h, l1, gt2, -0.5)                  # ret0
    contract("ai,ci->ac", l1, t1, G.gX1pp, beta=0.0)                     # X1pp t1
    contract("ac,ai->ci", B.X1pp, l1, gt1)
    contract("ai,ak->ki", l1, t1, G.gX1hh, beta=0.0)                     # X1hh t1
    contract("ki,ai->ak", B.X1hh, l1, gt1)
    # X1hh[k,i] = -f_hh - 1/2 f_ph[c,k] t1[c,i] - v_phhh[b,i,j,k] t1[b,j]
    #             - 1/2 v[c,d,l,k] t2[c,d,l,i] - 1/2 u[d,k] t1[d,i]
    contract("ck,ki->ci", ham.f_ph, G.gX1hh, gt1, alpha=-0.5)
    contract("bijk,ki->bj", ham.v_phhh, G.gX1hh, gt1, alpha=-1.0)
    contract("cdlk,ki->cdli", ham.v_pphh, G.gX1hh, gt2, alpha=-0.5)
    contract("ki,di->dk", G.gX1hh, t1, G.gu, alpha=-0.5)
    contract("dk,ki->di", B.u_ph, G.gX1hh, gt1, alpha=-0.5)
    # X1pp[a,c] = f_pp - 1/2 f_ph[c,k] t1[a,k] - 1/2 v[d,c,k,l] t2[d,a,k,l]
    #             + 1/2 w[c,l] t1[a,l] + ret1[a,c]
    contract("ac,ck->ak", G.gX1pp, ham.f_ph, gt1, alpha=-0.5)
    contract("dckl,ac->dakl", ham.v_pphh, G.gX1pp, gt2, alpha=-0.5)
    contract("ac,al->cl", G.gX1pp, t1, G.gw, alpha=0.5, beta=0.0)
    contract("ac,cl->al", G.gX1pp, B.w_ph, gt1, alpha=0.5)
    if pl.n:
        ops.axpby(1.0, G.gX1pp, 1.0, G.gret1)
    # ---- first-order intermediates of t1: u, w, z, s (above), ret1
    contract("cdlk,dk->cl", ham.v_pphh, G.gu, gt1)       # u[d,k] = v[c,d,l,k] t1[c,l]
    contract("cdkl,cl->dk", ham.v_pphh, G.gw, gt1)       # w[c,l] = v[c,d,k,l] t1[d,k]
    contract("cdlk,cl->dk", ham.v_pphh, G.gz, gt1)       # z[c,l] = v[c,d,l,k] t1[d,k]
    if pl.n:
        engine._segsum(adj["ret1T"], 1, G.gret1, gt1, -1.0)
    # ---- tau[a,b,i,j] = t2 + t1[a,i] t1[b,j] - t1[a,j] t1[b,i]
    ops.axpby(1.0, G.gtau, 1.0, gt2)
    contract("abij,bj->ai", G.gtau, t1, gt1)
    contract("baji,bj->ai", G.gtau, t1, gt1)
    contract("abji,bj->ai", G.gtau, t1, gt1, alpha=-1.0)
    contract("baij,bj->ai", G.gtau, t1, gt1, alpha=-1.0)
    # t2 adjoint gathered through the [a][i][c][k] layout
    permute_into(G.gTr.permute(0, 2, 1, 3), gt2, 1.0, 1.0)


def forward_intermediates(ham, t1, t2, B):
    """FP64 residual intermediates of engine.t1_part / engine.t2_part at (t1, t2) into B."""
    engine.t1_part(ham, t1, t2, B)
    engine.t2_part(ham, t1, t2, B)


def _dot(x, y, G):
    """x.y over all elements on the GPU (fixed summation order), returned to the host."""
    n = x.numel()
    _lib.call("nl_dot_part", n, ops._p(x), ops._p(y), ops._p(G.part), ops._stream())
    _lib.call("nl_finalize", int(_lib.load().nl_partials(n)), ops._p(G.part), 0, None, None,
              ops._p(G.out), ops._stream())
    return float(G.out.item())


def solve_lambda(ham, t1, t2, eps=1e-10, maxSteps=500, mixing=0.5, recompute=False, B=None,
                 G=None, info=None, verbose=False, max_diis=10):
    """Solve J^T lambda = -dE_corr/dt at converged amplitudes (t1, t2) by damped Jacobi.

    lambda_new = m lambda + (1-m) (lambda + P(J^T lambda + g) / D) with the dressed denominators
    D of the amplitude iteration at t and P the antisymmetriser of the T2 block; stops when
    |P(J^T lambda + g)| <= eps |g|.  With ``max_diis`` > 0 the steps are extrapolated by the
    pulsed DIIS of the amplitude solver every ``max_diis`` steps (the equations are linear, so
    this is a restarted Krylov acceleration of the same fixed point).  ``recompute=True``
    re-derives the forward intermediates from (t1, t2) in every iteration instead of keeping
    them between iterations.  Returns (l1, l2) (the tensors G.l1, G.l2)."""
    from .ccm import _Diis
    p, h = ham.p, ham.h
    dev = t1.device
    B = B if B is not None else engine.Buffers(p, h, dev)
    G = G if G is not None else AdjointBuffers(p, h, dev)
    st = ops._stream()
    nparts = int(_lib.load().nl_partials(p * p * h * h))
    forward_intermediates(ham, t1, t2, B)
    energy_gradient(ham, t1, t2, G)
    gnorm = math.sqrt(_dot(G.gt1, G.gt1, G) + _dot(G.gt2, G.gt2, G))
    G.l1.zero_()
    G.l2.zero_()
    dstate = _Diis("pulsed", max_diis, G.l1, G.l2, False) if max_diis > 0 else None
    t0 = time.perf_counter()
    res, steps = math.inf, 0
    for steps in range(1, maxSteps + 1):
        if recompute and steps > 1:
            forward_intermediates(ham, t1, t2, B)
        energy_gradient(ham, t1, t2, G)
        residual_vjp(ham, t1, t2, B, G.l1, G.l2, G)
        r1 = _dot(G.gt1, G.gt1, G)
        # T2 block: R2 = 1/4 residual as [a][i][b][j]; the assembly's pIJ pAB makes it P(res)
        permute_into(G.gt2.permute(0, 2, 1, 3), G.R2, 0.25)
        _lib.call("nl_update_t1", p, h, ops._p(G.l1), ops._p(G.gt1), ops._p(B.X1pp),
                  ops._p(B.X1hh), ctypes.c_double(mixing), ops._p(G.l1n), st)
        _lib.call("nl_assemble_t2", p, h, None, None, None, ops._p(G.R2), ops._p(B.X2pp),
                  ops._p(B.X2hh), ops._p(G.l2), None, None, ctypes.c_double(mixing),
                  ops._p(G.l2n), None, ops._p(G.part), st)
        _lib.call("nl_finalize", nparts, ops._p(G.part[nparts:]), 0, None, None, ops._p(G.out),
                  st)
        r2 = float(G.out.item())
        res = math.sqrt(r1 + r2) / max(gnorm, 1e-300)
        if verbose:
            print(f"lambda step {steps}: |residual|/|g| = {res:.3e}")
        if res <= eps or not math.isfinite(res):
            break        # the residual was taken at the current lambda: keep it
        G.l1, G.l1n = G.l1n, G.l1
        G.l2, G.l2n = G.l2n, G.l2
        if dstate is not None:
            ext = dstate.step(G.l1, G.l2, G.l1n, G.l2n, B)
            if ext is not None:
                permute_into(ext[0], G.l1)
                permute_into(ext[1], G.l2)
    if info is not None:
        info.update(lambda_steps=steps, lambda_residual=res,
                    lambda_seconds=time.perf_counter() - t0)
    if not math.isfinite(res):
        raise _lib.NLError("Lambda iteration diverged")
    return G.l1, G.l2


def residual(ham, t1, t2, B, H2out):
    """Raw residual H(t; ham): H1 into B.H1 and H2 into H2out (FP64)."""
    p, h = ham.p, ham.h
    engine.t1_part(ham, t1, t2, B)
    engine.t2_part(ham, t1, t2, B)
    _lib.call("nl_assemble_t2", p, h, ops._p(B.S0), ops._p(B.Sab), ops._p(B.Sij), ops._p(B.Spp),
              ops._p(B.X2pp), ops._p(B.X2hh), ops._p(t2), None, None, ctypes.c_double(0.0), None,
              ops._p(H2out), ops._p(B.part), ops._stream())
    return B.H1, H2out


def correlation_energy(ham, t1, t2, G):
    """E_corr(t; ham) = f.t1 + 1/4 v.t2 + 1/2 v t1 t1."""
    p, h = ham.p, ham.h
    st = ops._stream()
    _lib.call("nl_energy_part", p, h, ops._p(ham.v_pphh), ops._p(t2), ops._p(t1),
              ops._p(G.part), st)
    _lib.call("nl_finalize", int(_lib.load().nl_partials(p * p * h * h)), ops._p(G.part), p * h,
              ops._p(ham.f_ph), ops._p(t1), ops._p(G.out), st)
    return float(G.out.item())


def coupling_gradient(ham, t1, t2, derivatives, B=None, G=None, info=None, **lambda_kw):
    """dE/dtheta for each (dE_ref/dtheta, derivative Hamiltonian) in ``derivatives``:
    dE_ref/dtheta + E_corr(t; h_theta) + lambda . H(t; h_theta)."""
    p, h = ham.p, ham.h
    dev = t1.device
    B = B if B is not None else engine.Buffers(p, h, dev)
    G = G if G is not None else AdjointBuffers(p, h, dev)
    l1, l2 = solve_lambda(ham, t1, t2, B=B, G=G, info=info, **lambda_kw)
    out = []
    for dEref, hd in derivatives:
        H1, H2 = residual(hd, t1, t2, B, G.gt2)      # gt2 is free once the sweep is done
        out.append(dEref + correlation_energy(hd, t1, t2, G) + _dot(l1, H1, G)
                   + _dot(l2, H2, G))
    return out


# ------------------------------------------------------------------ autograd Function
class PlaneWaveCCSDEnergy(torch.autograd.Function):
    """E(C0, D0) = E_ref + E_corr of closed-shell plane-wave CCSD at (L, A); ``backward`` returns
    dE/dC0 and dE/dD0 from the Lambda equations.

    opts: eps (amplitude convergence, default 1e-12), maxSteps, mixing, max_diis (pulsed DIIS
    of the amplitude and Lambda iterations, default 10; 0 = plain damped Jacobi), precision
    ("fp64", "bf16", "fp16": low-precision iteration followed by FP64 refinement), lambda_eps,
    lambda_maxSteps, recompute (default True: the converged T2 is not kept between forward and
    backward; the backward re-solves the amplitude equations, whose kernels are deterministic,
    to regain it), device, info (dict receiving energies, steps and timings), form ("dense",
    default, or "sectors": the symmetry-sector form of the amplitude and Lambda iterations, FP64,
    with ``precision`` ignored)."""

    @staticmethod
    def forward(ctx, C0, D0, L, A, opts):
        from . import ccm
        from .planewave import PlaneWave, shell_holes
        dev = ccm._device(opts.get("device"))
        pw = PlaneWave(L, shell_holes(L, A), float(C0), float(D0))
        info = opts.get("info")
        if opts.get("form", "dense") == "sectors":
            from . import sectors
            SH = sectors.SectorHamiltonian.from_planewave(pw, device=dev)
            E0 = pw.ref_energy()
            e, t2 = _solve_sectors(sectors, SH, opts, info, "forward")
            ctx.state = dict(pw=pw, SH=SH, opts=opts, e=e,
                             t=t2 if not opts.get("recompute", True) else None)
            if info is not None:
                info.update(E_ref=E0, E_corr=e)
            return torch.tensor(E0 + e, dtype=F64)
        E0, fock, ham = pw.device_hamiltonian(dev)       # blocks and plans written on the GPU
        v = None
        e, t1, t2 = _solve(ccm, ham, fock, v, opts, info, "forward")
        keep = not opts.get("recompute", True)
        ctx.state = dict(pw=pw, ham=ham, fock=fock, v=v, opts=opts, e=e,
                         t=(t1, t2) if keep else None)
        if info is not None:
            info.update(E_ref=E0, E_corr=e)
        return torch.tensor(E0 + e, dtype=F64)

    @staticmethod
    def backward(ctx, gE):
        from . import ccm
        s = ctx.state
        if "SH" in s:
            return _backward_sectors(ctx, gE)
        ham, pw, opts = s["ham"], s["pw"], s["opts"]
        info = opts.get("info")
        if s["t"] is None:
            e, t1, t2 = _solve(ccm, ham, s["fock"], s["v"], opts, info, "backward")
            if info is not None:
                info["recomputed_E_corr_diff"] = abs(e - s["e"])
        else:
            t1, t2 = s["t"]
        derivs = []
        for n, name in enumerate(("C0", "D0")):
            if ctx.needs_input_grad[n]:
                dEref, dfock, scale = pw.coupling_derivative(name)
                derivs.append((dEref, ham.rescaled(dfock, scale)))
            else:
                derivs.append(None)
        lam = {}
        vals = iter(coupling_gradient(ham, t1, t2, [d for d in derivs if d is not None],
                                      info=lam, eps=opts.get("lambda_eps", 1e-10),
                                      maxSteps=opts.get("lambda_maxSteps", 500),
                                      mixing=opts.get("mixing", 0.5),
                                      max_diis=opts.get("max_diis", 10)))
        grads = [None if d is None else next(vals) for d in derivs]
        if info is not None:
            info.update(lam)
            info["dE_dC0"], info["dE_dD0"] = grads
        g = float(gE)
        out = [None if x is None else torch.tensor(g * x, dtype=F64) for x in grads]
        return out[0], out[1], None, None, None


def _solve(ccm, ham, fock, v, opts, info, phase):
    sinfo = {}
    out = ccm.ccsd_solver(fock, v, eps=opts.get("eps", 1e-12), maxSteps=opts.get("maxSteps", 1000),
                          max_diis=opts.get("max_diis", 10), diis="pulsed",
                          mixing=opts.get("mixing", 0.5), hamiltonian=ham,
                          precision=opts.get("precision", "fp64"), info=sinfo)
    if out is None or not sinfo.get("converged"):
        raise _lib.NLError("CCSD amplitude iteration did not converge: no gradient at an "
                           "unconverged point")
    e, t1, t2 = out
    if info is not None:
        info[phase] = sinfo
    return e, t1, t2


def _solve_sectors(sectors, SH, opts, info, phase):
    sinfo = {}
    out = sectors.solve(SH, eps=opts.get("eps", 1e-12), maxSteps=opts.get("maxSteps", 1000),
                        max_diis=opts.get("max_diis", 10), diis="pulsed",
                        mixing=opts.get("mixing", 0.5), info=sinfo, unpack=False)
    if out is None or not sinfo.get("converged"):
        raise _lib.NLError("CCSD amplitude iteration did not converge: no gradient at an "
                           "unconverged point")
    if info is not None:
        info[phase] = {k: v for k, v in sinfo.items() if k != "sector_t2"}
    return out[0], sinfo["sector_t2"]


def _backward_sectors(ctx, gE):
    from . import sectors
    s = ctx.state
    SH, pw, opts = s["SH"], s["pw"], s["opts"]
    info = opts.get("info")
    if s["t"] is None:
        e, t2 = _solve_sectors(sectors, SH, opts, info, "backward")
        if info is not None:
            info["recomputed_E_corr_diff"] = abs(e - s["e"])
    else:
        t2 = s["t"]
    names = [n for k, n in enumerate(("C0", "D0")) if ctx.needs_input_grad[k]]
    derivs = [sectors.planewave_derivative(SH, pw, n) for n in names]
    lam = {}
    vals = dict(zip(names, sectors.coupling_gradient(
        SH, t2, derivs, info=lam, eps=opts.get("lambda_eps", 1e-10),
        maxSteps=opts.get("lambda_maxSteps", 500), mixing=opts.get("mixing", 0.5),
        max_diis=opts.get("max_diis", 10))))
    grads = [vals.get("C0"), vals.get("D0")]
    if info is not None:
        info.update(lam)
        info["dE_dC0"], info["dE_dD0"] = grads
    g = float(gE)
    out = [None if x is None else torch.tensor(g * x, dtype=F64) for x in grads]
    return out[0], out[1], None, None, None


def planewave_ccsd_energy(C0, D0, L, A, **opts):
    """Differentiable CCSD total energy of the closed-shell plane-wave contact Hamiltonian.

    C0, D0: floats or 0-dim tensors (with requires_grad set, backward gives dE/dC0, dE/dD0)."""
    C0 = C0 if isinstance(C0, torch.Tensor) else torch.tensor(float(C0), dtype=F64)
    D0 = D0 if isinstance(D0, torch.Tensor) else torch.tensor(float(D0), dtype=F64)
    return PlaneWaveCCSDEnergy.apply(C0, D0, L, A, opts)

"""Drop-in for NuLattice's CCSD amplitude iteration (``NuLattice.CCM.coupled_cluster``) on MI355X.

Same entry points, argument order and meaning, return types and failure behaviour as the
reference (src/NuLattice/CCM/coupled_cluster.py:265-687, ccDgrams.py), with the arithmetic in
libnlccsd.so.  Inputs may be NumPy arrays, torch tensors, reference COO lists
([[i0,i1,i2,i3,val], ...]) or (indices, values) pairs; amplitudes are returned as float64 HIP
device tensors in the reference layout t1[a,i], t2[a,b,i,j].

Extensions (no reference counterpart), all keyword-only with reference-compatible defaults:
  diis       "legacy" (default: the current reference, whose DIIS block never runs),
             "pulsed" (the pulsed DIIS that produced the reference notebooks' trajectories) or
             "soa" (src/NuLattice/CCM/soa/coupled_cluster.py:520-567)
  precision  "fp64" (default) | "bf16" | "fp16": ring, ladder and dressing contractions on
             low-precision MFMA with FP32 accumulation; with ``refine=True`` the iteration then
             continues in FP64 from the low-precision fixed point until ``eps`` is met.
  device, history (list receiving every step's energy), info (dict receiving timings/steps).
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib, engine, ops
from .engine import Buffers, Hamiltonian

F64 = torch.float64

_PREC = {
    "fp64": engine.DEFAULT_PREC,
    "bf16": {"ring": "bf16", "hh": "bf16", "dress": "bf16", "ladder": "bf16"},
    "fp16": {"ring": "fp16", "hh": "fp16", "dress": "fp16", "ladder": "fp16"},
    "bf16x3": {"ring": "bf16x3", "hh": "bf16x3", "dress": "bf16x3", "ladder": "bf16x3"},
    "fp16x3": {"ring": "fp16x3", "hh": "fp16x3", "dress": "fp16x3", "ladder": "fp16x3"},
}


def _device(device):
    if device is None:
        if not torch.cuda.is_available():
            raise _lib.NLError("no HIP device visible (the CCSD kernels have no CPU fallback)")
        device = "cuda"
    return torch.device(device)


def _t(x, device):
    return engine._dev(x, device)


class _Diis:
    """Pulsed DIIS on device vectors; B-matrix solve in FP64 on the host (a 10x10 system).
    ``store`` is the dtype of the kept history entries and error vectors: float64, or float32
    as the SoA driver keeps them with dtype=torch.float32 (soa/coupled_cluster.py:453-460,
    520-567); products, sums and the extrapolated amplitudes are FP64 either way."""

    def __init__(self, mode, max_diis, t1, t2, ccs, store=torch.float64):
        self.mode, self.n, self.ccs = mode, max_diis, ccs
        self.f32 = store == torch.float32
        self.h1, self.h2, self.e1, self.e2 = [self._entry(t1)], [self._entry(t2)], [], []

    def _entry(self, t):
        if not self.f32:
            return ops.copy(t)
        hc = torch.empty(t.shape, dtype=torch.float32, device=t.device)
        junk = torch.empty_like(hc)
        _lib.call("nl_copy_sub_f32", t.numel(), ops._p(t), ops._p(t), ops._p(hc), ops._p(junk),
                  ops._stream())
        return hc

    def step(self, t1, t2, o1, o2, B=None):
        """Record (t, t - old); return extrapolated (t1, t2) or None.  ``B`` is unused (kept for
        callers that pass reduction buffers; the Gram kernel allocates its own partials)."""
        soa = self.mode == "soa"
        # history entry (copy of t) and error vector (t - old) in one pass each
        for t, o, hist, errs in ((t1, o1, self.h1, self.e1), (t2, o2, self.h2, self.e2)):
            if not (t.is_contiguous() and o.is_contiguous()) or t.numel() != o.numel():
                raise _lib.NLError("DIIS vectors must be contiguous and of one size")
            dt = torch.float32 if self.f32 else torch.float64
            hc = torch.empty(t.shape, dtype=dt, device=t.device)
            e = torch.empty_like(hc)
            _lib.call("nl_copy_sub_f32" if self.f32 else "nl_copy_sub", t.numel(), ops._p(t),
                      ops._p(o), ops._p(hc), ops._p(e), ops._stream())
            hist.append(hc); errs.append(e)
        if soa and len(self.e1) > self.n:
            for L in (self.h1, self.h2, self.e1, self.e2):
                L.pop(0)
        if len(self.e1) != self.n:
            return None
        if not soa:
            for L in (self.h1, self.h2, self.e1, self.e2):
                L.pop(0)
        n = len(self.e1)
        # B[a, b] = e_a . e_b over the concatenated (t1, t2) errors, each list read in one pass
        Bm = ops.gram(self.e1) + ops.gram(self.e2)
        Bm = Bm / (np.abs(Bm).max() + (1e-16 if soa else 0.0))
        A = -np.ones((n + 1, n + 1)); A[:n, :n] = Bm; A[n, n] = 0.0
        rhs = np.zeros(n + 1); rhs[n] = -1.0
        out = None
        try:
            c = np.linalg.solve(A, rhs)[:n]
            n1 = ops.lincomb(c, self.h1[1:], torch.empty_like(t1))
            n2 = t2 if self.ccs else ops.lincomb(c, self.h2[1:], torch.empty_like(t2))
            out = (n1, n2)
        except np.linalg.LinAlgError:
            pass
        r1, r2 = out if out is not None else (t1, t2)
        self.h1, self.h2, self.e1, self.e2 = [self._entry(r1)], [self._entry(r2)], [], []
        return out


def t1Iter(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, v_ppph_results, sparse=True):
    """One unfused T1 Jacobi step, t1 + H1 / (-(X_pp[a,a] + X_hh[i,i])) (coupled_cluster.py:
    342-414), built from the diagram functions of ccdgrams.  ``v_ppph_results`` is the 7-tuple of
    ``v_ppph_dgrams`` (sparse=True) or the dense v_ppph[a,b,c,i] (sparse=False).  Returns a new
    float64 HIP tensor; inputs are not modified."""
    from . import ccdgrams as dg
    t1, t2 = dg._d(t1), dg._d(t2)
    H1 = dg.start(f_ph)
    dg.accumulate(H1, dg.dgram_akci_ck(v_phph, t1), dg.dgram_ck_acik(f_ph, t2),
                  dg.dgram_cikl_cakl(v_phhh, t2), dg.dgram_cdkl_ck_dali(v_pphh, t1, t2))
    X_hh = dg.start(f_hh, -1.0)
    X_pp = dg.start(f_pp)
    dg.accumulate(X_hh, dg.dgram_ck_ci(f_ph, t1), dg.dgram_bijk_bj(v_phhh, t1),
                  dg.dgram_cdlk_cdli(v_pphh, t2), dg.dgram_cdlk_cl_di(v_pphh, t1))
    dg.accumulate(X_pp, dg.dgram_ck_ak(f_ph, t1), dg.dgram_dckl_dakl(v_pphh, t2),
                  dg.dgram_cdkl_dk_al(v_pphh, t1))
    if sparse:
        dg.accumulate(H1, v_ppph_results[0])
        dg.accumulate(X_pp, v_ppph_results[1])
    else:
        dg.accumulate(H1, dg._c("cdak,cdki->ai", -0.5, v_ppph_results, t2))
        dg.accumulate(X_pp, dg._c("cdak,ck->ad", -1.0, v_ppph_results, t1))
    dg.accumulate(H1, dg._c("ac,ci->ai", 1.0, X_pp, t1), dg._c("ki,ak->ai", 1.0, X_hh, t1))
    return dg.jacobi_t1(t1, H1, X_pp, X_hh)


def t2Iter(t1, t2, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh, v_ppph_results, v_hhhh,
           sparse=True):
    """One unfused T2 Jacobi step, t2 + H2 / (-(X_aa + X_bb + X_ii + X_jj)) (coupled_cluster.py:
    442-545; note the reference's argument order f_ph, f_hh, f_pp), built from the diagram
    functions of ccdgrams.  Sparse: v_pppp is COO and ``v_ppph_results`` the 7-tuple of
    ``v_ppph_dgrams``; dense: both blocks are arrays.  Returns a new float64 HIP tensor."""
    from . import ccdgrams as dg
    t1, t2 = dg._d(t1), dg._d(t2)
    H2 = dg.start(v_pphh)
    dg.accumulate(H2, dg.dgram_klij_abkl(v_hhhh, t2), dg.dgram_bkcj_acik(v_phph, t2),
                  dg.dgram_bkij_ak(v_phhh, t1), dg.dgram_cdkl_acik_dblj(v_pphh, t2),
                  dg.dgram_cdkl_cdij_abkl(v_pphh, t2), dg.dgram_klij_ak_bl(v_hhhh, t1),
                  dg.dgram_bkci_ak_cj(v_phph, t1), dg.dgram_cikl_ck_ablj(v_phhh, t1, t2),
                  dg.dgram_cikl_al_bcjk(v_phhh, t1, t2), dg.dgram_cjkl_ci_abkl(v_phhh, t1, t2),
                  dg.dgram_cjkl_ci_ak_bl(v_phhh, t1), dg.dgram_cdkl_ci_dj_abkl(v_pphh, t1, t2),
                  dg.dgram_cdkl_ak_bl_cdij(v_pphh, t1, t2),
                  dg.dgram_cdkl_ci_bl_adkj(v_pphh, t1, t2), dg.dgram_cdkl_ci_ak_dj_bl(v_pphh, t1))
    X_hh = dg.start(f_hh, -1.0)
    X_pp = dg.start(f_pp)
    dg.accumulate(X_pp, dg.dgram_cdkl_bdkl(v_pphh, t2), dg.dgram_ck_bk(f_ph, t1),
                  dg.dgram_cdlk_dk_bl(v_pphh, t1))
    dg.accumulate(X_hh, dg.dgram_cdkl_cdjl(v_pphh, t2), dg.dgram_ck_cj(f_ph, t1),
                  dg.dgram_cdlk_cl_dj(v_pphh, t1))
    if sparse:
        r = v_ppph_results
        dg.accumulate(H2, dg.pIJ(r[2]), dg.dgram_da_dbij(r[3], t2), dg.dgram_acik_bcjk(r[4], t2),
                      dg.dgram_bijk_ak1(r[5], t1), dg.dgram_bijk_ak2(r[6], t1))
        lad1, lad2 = dg.v_pppp_dgrams(v_pppp, t1, t2)
        dg.accumulate(H2, lad1, dg.pIJ(lad2), scale=0.5)
    else:
        v3 = dg._d(v_ppph_results)
        dg.accumulate(H2, dg.pIJ(dg._c("abcj,ci->abij", 1.0, v3, t1)),
                      dg.pAB(dg._c("cdak,ck,dbij->abij", 1.0, v3, t1, t2), -1.0),
                      dg._pp(dg._c("dcak,di,bcjk->abij", 1.0, v3, t1, t2)),
                      dg.pAB(dg._c("cdbk,cdij,ak->abij", 1.0, v3, t2, t1), 0.5),
                      dg._pp(dg._c("cdbk,ci,dj,ak->abij", 1.0, v3, t1, t1, t1), 0.5),
                      dg._c("abcd,cdij->abij", 0.5, v_pppp, t2),
                      dg.pIJ(dg._c("abcd,ci,dj->abij", 1.0, v_pppp, t1, t1), 0.5))
    dg.accumulate(H2, dg.pAB(dg._c("bc,acij->abij", 1.0, X_pp, t2)),
                  dg.pIJ(dg._c("kj,abik->abij", 1.0, X_hh, t2)))
    return dg.jacobi_t2(t2, H2, X_pp, X_hh)


def v_ppph_dgrams(v_ppph, t1, t2):
    """The seven v_ppph intermediates (ccDgrams.py:13-45) on the GPU; see ccdgrams."""
    from . import ccdgrams as dg
    return dg.v_ppph_dgrams(v_ppph, t1, t2)


def v_pppp_dgrams(v_pppp, t1, t2):
    """pp ladder intermediates (ccDgrams.py:215-237) on the GPU; see ccdgrams."""
    from . import ccdgrams as dg
    return dg.v_pppp_dgrams(v_pppp, t1, t2)


def ccsd_energy(f_ph, v_pphh, t2, t1, device=None):
    """E = sum f_ai t_ai + 1/4 sum v_abij t_abij + 1/2 sum v_abij t_ai t_bj
    (coupled_cluster.py:265-286), evaluated on the GPU; returns a Python float."""
    dev = _device(device if device is not None else getattr(t2, "device", None))
    f_ph, v_pphh, t1, t2 = (_t(x, dev) for x in (f_ph, v_pphh, t1, t2))
    p, h = f_ph.shape
    part = torch.empty(int(_lib.load().nl_partials(p * p * h * h)) + 1, dtype=F64, device=dev)
    E = torch.empty(1, dtype=F64, device=dev)
    _lib.call("nl_energy_part", p, h, ops._p(v_pphh), ops._p(t2), ops._p(t1), ops._p(part),
              ops._stream())
    _lib.call("nl_finalize", int(_lib.load().nl_partials(p * p * h * h)), ops._p(part), p * h,
              ops._p(f_ph), ops._p(t1), ops._p(E), ops._stream())
    return float(E.item())


def t1Init(f_ph, f_pp, f_hh, delta, device=None):
    """t1 = f_ph / (f_ii - f_aa + delta) (coupled_cluster.py:319-340)."""
    dev = _device(device)
    ham = _MiniHam(dev, f_pp=f_pp, f_ph=f_ph, f_hh=f_hh)
    t1 = torch.empty_like(ham.f_ph)
    Xhh = torch.empty_like(ham.f_hh)
    ops.permute_into(ham.f_hh, Xhh, alpha=-1.0)
    if delta:
        engine._diag_add(Xhh, -delta)
    zero = torch.zeros_like(t1)
    _lib.call("nl_update_t1", ham.p, ham.h, ops._p(zero), ops._p(ham.f_ph), ops._p(ham.f_pp),
              ops._p(Xhh), ctypes.c_double(0.0), ops._p(t1), ops._stream())
    return t1


def t2Init(f_pp, f_hh, v_pphh, delta, device=None):
    """t2 = v_pphh / (f_ii + f_jj - f_aa - f_bb + delta) (coupled_cluster.py:416-439)."""
    dev = _device(device)
    ham = _MiniHam(dev, f_pp=f_pp, f_hh=f_hh, v_pphh=v_pphh)
    p, h = ham.f_pp.shape[0], ham.f_hh.shape[0]
    t2 = torch.empty((p, p, h, h), dtype=F64, device=dev)
    Xhh = torch.empty_like(ham.f_hh)
    ops.permute_into(ham.f_hh, Xhh, alpha=-1.0)
    if delta:
        engine._diag_add(Xhh, -0.5 * delta)
    zero = torch.zeros_like(t2)
    part = torch.empty(2 * int(_lib.load().nl_partials(p * p * h * h)) + 1, dtype=F64, device=dev)
    _lib.call("nl_assemble_t2", p, h, ops._p(ham.v_pphh), None, None, None, ops._p(ham.f_pp),
              ops._p(Xhh), ops._p(zero), None, None, ctypes.c_double(0.0), ops._p(t2), None,
              ops._p(part), ops._stream())
    return t2


class _MiniHam:
    def __init__(self, dev, **kw):
        for k, v in kw.items():
            setattr(self, k, _t(v, dev))
        if hasattr(self, "f_ph"):
            self.p, self.h = self.f_ph.shape


def _solver_ham(fock_mats, two_body_int, device):
    return Hamiltonian(fock_mats, two_body_int, device=device)


def ccsd_solver(fock_mats, two_body_int, t1initial=None, eps=1e-8, maxSteps=1000, max_diis=10,
                delta=0, mixing=0.5, verbose=False, sparse=True, ccs=False, *, device=None,
                diis="legacy", precision="fp64", refine=True, eps_lowprec=None, history=None,
                info=None, hamiltonian=None, adjoint=False, lambda_eps=1e-10, sectors=None,
                storage=torch.float64, _soa_rules=False):
    """Damped-Jacobi CCSD (coupled_cluster.py:547-687) on one MI355X.

    Returns (E_corr, t1, t2); like the reference, returns None after a divergence or NaN.
    ``sparse`` only says how v_pppp / v_ppph are given (COO lists or dense arrays); both are
    packed into the same block plans.  ``hamiltonian`` may pass a prebuilt engine.Hamiltonian,
    or a sectors.SectorHamiltonian for the symmetry-sector form; ``sectors=(part_labels,
    hole_labels, moduli)`` (e.g. ``PlaneWave.labels()``) builds the latter from the blocks.  The
    sector form is FP64 and returns the same (E_corr, t1, t2) in the reference layout.
    With a low ``precision`` the iteration first runs to |dE|/|E| < ``eps_lowprec`` (default
    max(eps, 1e-7)) or until |dE| stops decreasing (the low-precision fixed point is only defined
    to the rounding level of the contractions) and then, if ``refine``, continues in FP64 to
    ``eps``.  ``adjoint=True`` also solves the CC Lambda equations at the returned amplitudes
    (to |residual| <= ``lambda_eps`` |dE/dt|) and stores (lambda1, lambda2) in info["lambda"].
    ``storage=torch.float32`` keeps the DIIS history in FP32 and returns t1, t2 as float32
    (the working amplitudes, energy and denominators stay FP64).
    ``_soa_rules`` applies the SoA driver's stopping and failure rules instead (see soa.py).
    """
    if adjoint and info is None:
        raise ValueError("adjoint=True needs an info dict to receive the Lambda amplitudes")
    if diis not in ("legacy", "pulsed", "soa"):
        raise ValueError(f"diis must be 'legacy', 'pulsed' or 'soa', not {diis!r}")
    if precision not in _PREC:
        raise ValueError(f"precision must be one of {sorted(_PREC)}")
    dev = _device(device)
    from . import sectors as sec
    if sectors is not None and hamiltonian is None:
        hamiltonian = sec.SectorHamiltonian(fock_mats, two_body_int, *sectors, device=dev)
    if isinstance(hamiltonian, sec.SectorHamiltonian):
        if _soa_rules:
            raise ValueError("the sector form follows the legacy driver's rules")
        if precision != "fp64" or adjoint:
            raise ValueError("the sector form runs in FP64 without the adjoint")
        return sec.solve(hamiltonian, t1initial, eps, maxSteps, max_diis, delta, mixing, verbose,
                         ccs, diis, history, info)
    ham = hamiltonian if hamiltonian is not None else _solver_ham(fock_mats, two_body_int, dev)
    p, h = ham.p, ham.h
    B = Buffers(p, h, dev)
    t1 = torch.empty((p, h), dtype=F64, device=dev)
    t2 = torch.empty((p, p, h, h), dtype=F64, device=dev)
    t1n, t2n = torch.empty_like(t1), torch.empty_like(t2)
    if t1initial is None:
        engine.init_amplitudes(ham, B, t1, t2, delta)
    else:
        ops.permute_into(_t(t1initial, dev), t1)
    if ccs or t1initial is not None:
        t2.zero_()
    pulsed = diis != "legacy" and max_diis > 0
    if storage not in (torch.float64, torch.float32):
        raise ValueError("storage must be torch.float64 or torch.float32")
    dstate = _Diis(diis, max_diis, t1, t2, ccs, storage) if pulsed else None
    out = (lambda e, a, b: (e, a, b)) if storage == torch.float64 else \
        (lambda e, a, b: (e, _as_f32(a), _as_f32(b)))
    img = engine.T2Image()      # BF16 image of t2 handed from one step's assembly to the next
    engine.energy(ham, t1, t2, B)
    prev = float(B.E.item())
    if history is not None:
        history.append(prev)
    if verbose:
        print(f"Step {0}: {prev}")
    prec = _PREC[precision]
    lp_phase = precision != "fp64"
    eps_lp = max(eps, 1e-7) if eps_lowprec is None else eps_lowprec
    best_rel, since_best = math.inf, 0
    steps_lp = steps_fp64 = 0
    t_start = time.perf_counter()
    energy = prev
    i = 0
    while i < maxSteps:
        if ccs:
            engine.t1_part(ham, t1, t2, B, prec)
            _lib.call("nl_update_t1", p, h, ops._p(t1), ops._p(B.H1), ops._p(B.X1pp),
                      ops._p(B.X1hh), ctypes.c_double(mixing), ops._p(t1n), ops._stream())
            ops.permute_into(t2, t2n)
            engine.energy(ham, t1n, t2n, B)
        else:
            engine.update(ham, t1, t2, B, t1n, t2n, mixing, prec, img=img)
        energy = float(B.E.item())
        if lp_phase:
            steps_lp += 1
        else:
            steps_fp64 += 1
        if history is not None:
            history.append(energy)
        if verbose:
            print(f"Step {i + 1}: {energy}", "difference =", abs(energy - prev) / abs(energy))
        old1, old2 = t1, t2
        t1, t1n = t1n, t1
        t2, t2n = t2n, t2
        # legacy: |dE|/|E| (coupled_cluster.py:629);
        # SoA: |dE|/max(1, |E|) (soa/coupled_cluster.py:517)
        den = max(1.0, abs(energy)) if _soa_rules else abs(energy)
        rel = abs(energy - prev) / den if den != 0 else 0.0
        converged = (energy == 0 and prev == 0) or rel < eps
        if lp_phase:
            # the low-precision iteration settles into rounding jitter instead of converging:
            # stop it at eps_lowprec or when |dE| has not halved for 6 steps
            if rel < 0.5 * best_rel:
                best_rel, since_best = rel, 0
            else:
                since_best += 1
        if lp_phase and refine and ((energy == 0 and prev == 0) or rel < eps_lp
                                    or since_best >= 6):
            # the low-precision fixed point is reached: continue in FP64 from here
            lp_phase, prec = False, _PREC["fp64"]
            if info is not None:
                info["E_lowprec"] = energy
            prev = energy
            i += 1
            continue
        if converged:
            if verbose and not _soa_rules:
                print(f"Energy found in {i + 1} iterations")
            _fill_info(info, steps_lp, steps_fp64, t_start, energy, True)
            if adjoint:
                _lambda(ham, t1, t2, B, mixing, lambda_eps, info)
            return out(energy, t1, t2)
        if dstate is not None and dstate.f32 and rel < 1e-7:
            # an FP32 history cannot carry the amplitudes below its rounding (each
            # extrapolation would re-inject ~6e-8 relative noise): the last steps run as plain
            # damped Jacobi in FP64
            dstate = None
        if dstate is not None:
            ext = dstate.step(t1, t2, old1, old2, B)
            if ext is not None:
                t1, t2 = ext
                t1n, t2n = torch.empty_like(t1), torch.empty_like(t2)
        if _soa_rules and (abs(energy) > 1e10 or math.isnan(energy)):
            # SoA (soa/coupled_cluster.py:569-575): report, leave the loop, return the last
            # energy and amplitudes
            print("Diverged.")
            break
        if abs(energy) > 1e10:
            print("Iteration Diverged")
            _fill_info(info, steps_lp, steps_fp64, t_start, energy)
            return None
        if math.isnan(energy):
            print("nan error")
            _fill_info(info, steps_lp, steps_fp64, t_start, energy)
            return None
        prev = energy
        i += 1
    print("Max iterations reached." if _soa_rules else "Max Iterations Reached")
    _fill_info(info, steps_lp, steps_fp64, t_start, energy, False)
    if adjoint:
        _lambda(ham, t1, t2, B, mixing, lambda_eps, info)
    return out(energy, t1, t2)


def _as_f32(t):
    """float32 copy of an FP64 device tensor (the library's cast: the copy half of
    nl_copy_sub_f32)."""
    o = torch.empty(t.shape, dtype=torch.float32, device=t.device)
    junk = torch.empty_like(o)
    _lib.call("nl_copy_sub_f32", t.numel(), ops._p(t), ops._p(t), ops._p(o), ops._p(junk),
              ops._stream())
    return o


def _fill_info(info, steps_lp, steps_fp64, t_start, energy, converged=False):
    if info is not None:
        info.update(steps_lowprec=steps_lp, steps_fp64=steps_fp64,
                    seconds=time.perf_counter() - t_start, energy=energy, converged=converged)


def _lambda(ham, t1, t2, B, mixing, eps, info):
    from . import adjoint as adj
    l1, l2 = adj.solve_lambda(ham, t1, t2, eps=eps, mixing=mixing, B=B, info=info)
    info["lambda"] = (l1.clone(), l2.clone())

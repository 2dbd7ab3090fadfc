"""CCSD iterations/sec at L=5, A=28 on one MI355X: mixed precision vs FP64, |dE| vs FP64.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--L 5] [--A 28] [--seed 0]
                    [--prec bf16x3|fp16x3|bf16|fp16] [--no-converge]
                    [--cpu-sample sample|full|none] [--form all|sectors|diis] [--grad on|off]

Workload: plane-wave lattice contact theory, L=5 (500 orbitals), A=28 closed shell (k = 0 and
the first momentum shell filled; h=28, p=472), SU(4) two-body contact C0 ~ U[-4.5,-3.5] from
numpy.random.default_rng(seed), three-body contact D0=1.0 normal-ordered at the reference,
dense T2 (p,p,h,h) as in the reference's layout.  A step is one CCSD Jacobi iteration (T1 and T2
residuals, update with mixing 0.5, energy read back to the host), timed over K steps after W
warm-ups with the device synchronised on both sides.  ``value`` is the mixed-precision
iteration rate (default fp16x3: FP16 matrix cores on hi/lo split operands, FP32 accumulation,
FP64 energy, denominators and update); the FP64 rate, the speed-up and |dE| of the
mixed-precision fixed point against the converged FP64 energy (seeds 0..4) are reported beside
it, with the single-rounding BF16 / FP16 forms for reference, the roofline of the dominant
kernel (the ring GEMM), the executed matrix-core work per iteration against the peaks, the
wall time to the converged energy, the cost of one DIIS extrapolation, the symmetry-sector
form of the same iteration, the adjoint gradient against central differences, and the CPU
oracle's rate on this host.  ``--form sectors`` and ``--form diis`` time only the sector form
or only the DIIS extrapolation.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=5)
    ap.add_argument("--A", type=int, default=28)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--D0", type=float, default=None,
                    help="three-body contact (default 1.0; drawn from U[0.5, 1.5] after C0 "
                         "from the seed in the combined form)")
    ap.add_argument("--prec", default="fp16x3", choices=["fp16x3", "bf16x3", "bf16", "fp16"])
    ap.add_argument("--no-converge", action="store_true", help="skip the |dE| convergence runs")
    ap.add_argument("--cpu-sample", default="sample", choices=["sample", "full", "none"],
                    help="sample: one oracle iteration at the bench size with every dense "
                         "contraction in full and the two COO loops over every 16th entry "
                         "(their time scaled by 16); full: the whole iteration")
    ap.add_argument("--form", default="all",
                    choices=["all", "sectors", "diis", "combined", "iter"],
                    help="sectors: time only the symmetry-sector form and print its line; "
                         "diis: time only one DIIS extrapolation over 10 T2-sized vectors; "
                         "combined (the default for L >= 6): the largest-size line, FP32 "
                         "storage through the SoA driver, mixed-precision residual, adjoint "
                         "gradient, seeded C0 and D0; iter: only the FP64 and the precision "
                         "form's iterations (the launches a kernel profile covers)")
    ap.add_argument("--grad", default="on", choices=["on", "off"],
                    help="adjoint dE/dC0, dE/dD0 at C0=-4, D0=1 checked against central FD")
    return ap.parse_args()


def canonical_flops(p, h, nnz_pppp, nnz_ppph):
    """Reference term list work per iteration (best pairwise order), FLOP: dense MACs
    (4 p^3h^3 ring-type, 6 p^2h^4 ladder-type, 5 p^3h^2 dressing-type, p^2h^3 rest) + COO."""
    dense = 4 * p**3 * h**3 + 6 * p**2 * h**4 + 5 * p**3 * h**2 + 20 * p**2 * h**3
    coo = 2 * nnz_pppp * h * h + nnz_ppph * (3 * h + 2 * h * h + 2)
    return 2.0 * (dense + coo)


def time_iterations(ccm_mod, engine, ops, H, prec, steps, warmup, torch):
    p, h = H.p, H.h
    dev = H.device
    B = engine.Buffers(p, h, dev)
    t1 = torch.empty((p, h), dtype=torch.float64, device=dev)
    t2 = torch.empty((p, p, h, h), dtype=torch.float64, device=dev)
    t1n, t2n = torch.empty_like(t1), torch.empty_like(t2)
    engine.init_amplitudes(H, B, t1, t2, 0.0)
    P = ccm_mod._PREC[prec]
    img = engine.T2Image()        # as the solver does: the assembly writes the next BF16 image
    energies = []
    for _ in range(warmup):
        engine.update(H, t1, t2, B, t1n, t2n, 0.5, P, img=img)
        energies.append(float(B.E.item()))
        t1, t1n, t2, t2n = t1n, t1, t2n, t2
    torch.cuda.synchronize()
    ops.TIMERS.clear(); ops.TIMERS["on"] = True
    ops.FLOPS.update(on=True, fp64=0.0, lp=0.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        engine.update(H, t1, t2, B, t1n, t2n, 0.5, P, img=img)
        energies.append(float(B.E.item()))
        t1, t1n, t2, t2n = t1n, t1, t2n, t2
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ops.TIMERS["on"] = False
    ops.FLOPS["on"] = False
    kern_flops = {"fp64": ops.FLOPS["fp64"] / steps, "lp": ops.FLOPS["lp"] / steps}
    kern = {}
    for tag, evs in ops.TIMERS.items():
        if tag == "on":
            continue
        kern[tag] = sum(s.elapsed_time(e) for s, e in evs) / len(evs) / 1e3   # seconds
    kern["flops"] = kern_flops
    return dt / steps, kern, energies


def time_sectors(SH, steps, warmup, torch, graph=True):
    """Per-iteration time of the symmetry-sector form (same step: update, mixing, energy read
    back to the host), with the step replayed as a HIP graph or launched kernel by kernel."""
    from nulattice_amd.sectors import SectorStepper
    Bf = SH.buffers()
    st = SectorStepper(SH, Bf, SH.init_amplitudes(Bf), 0.5, graph=graph)
    for _ in range(warmup):
        st.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st.step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def profile_value(name, key):
    """A figure recorded in a committed profile summary (``key = value`` line), or None."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as f:
            m = re.search(rf"^{re.escape(key)}\s*=\s*(\S+)", f.read(), re.M)
    except OSError:
        return None
    return float(m.group(1)) if m else None


def combined(a, C0, D0, torch, ccm, engine, ops, adjoint):
    """The combined configuration at the largest single-GPU size (L=6, A=28: T2 836^2 x 28^2 =
    5.5e8 amplitudes): the dense mixed-precision iteration rate (and the FP64 one), the SoA
    driver with FP32 storage (DIIS history and returned amplitudes in FP32) to convergence with
    |dE| against the FP64 symmetry-sector solve, the matrix-core work per iteration against the
    peaks, the assembly's HBM rate, and dE/dC0, dE/dD0 from the Lambda equations (sector form)
    against central finite differences of converged energies."""
    from nulattice_amd import soa
    from nulattice_amd.planewave import PlaneWave, planewave_device_hamiltonian, shell_holes
    from nulattice_amd.sectors import SectorHamiltonian
    t0 = time.perf_counter()
    E0, fock, H = planewave_device_hamiltonian(a.L, a.A, C0, D0)
    v = None                  # the blocks exist only on the device
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    p, h = H.p, H.h
    print(f"# combined L={a.L} A={a.A} p={p} h={h} C0={C0:.6f} D0={D0:.6f}: setup {setup:.1f} s",
          file=sys.stderr, flush=True)
    slp, klp, _ = time_iterations(ccm, engine, ops, H, a.prec, a.steps, a.warmup, torch)
    print(f"# {a.prec}: {slp * 1e3:.1f} ms/iter", file=sys.stderr, flush=True)
    s64, k64, _ = time_iterations(ccm, engine, ops, H, "fp64", 2, 1, torch)
    print(f"# fp64: {s64 * 1e3:.1f} ms/iter", file=sys.stderr, flush=True)
    n = p * p * h * h
    asm = klp.get("assemble")
    # assembly: reads S0, Sab, Sij, Spp, t2, v_pphh (FP64), writes t2n (FP64) and its 16-bit
    # image (4 B for the split image, 2 B for a BF16 one), each element once
    img_b = 4 if a.prec.endswith("x3") else 2
    asm_gbs = (6 * 8 + 8 + img_b) * n / asm / 1e9 if asm else None
    fl = klp["flops"]
    out = {"metric": f"CCSD iters/sec at L={a.L} A={a.A}: mixed-prec, FP32 storage, "
                     f"|dE| vs FP64, adjoint gradient",
           "value": 1.0 / slp, "unit": "iters/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": slp * 1e3, "higher_is_better": True,
           "dtype": f"{a.prec}+fp64, fp32 storage",
           "config": {"workload": f"CCSD L={a.L} A={a.A} plane-wave, dense T2 {p}x{p}x{h}x{h}",
                      "L": a.L, "A": a.A, "p": p, "h": h, "C0": C0, "D0": D0, "seed": a.seed},
           "fp64_ms_per_step": s64 * 1e3, "speedup_vs_fp64": s64 / slp,
           "mfma_frac_executed": (fl["lp"] / 2.5e15 + fl["fp64"] / 78.6e12) / slp,
           "executed_flop": fl, "assembly_ms": asm * 1e3 if asm else None,
           "assembly_hbm_GBps": asm_gbs, "setup_s": setup}
    SH = SectorHamiltonian.from_planewave(PlaneWave(a.L, shell_holes(a.L, a.A), C0, D0))
    e64, _, _ = ccm.ccsd_solver(None, None, eps=1e-12, maxSteps=400, max_diis=0, hamiltonian=SH)
    del SH
    info = {}
    e32, t1, t2 = soa.ccsd_solver(fock, v, eps=1e-10, maxSteps=200, dtype=torch.float32,
                                  hamiltonian=H, info=info)
    out.update(E_corr_fp64=e64, E_corr_fp32_storage=e32, dE_rel=abs(e32 - e64) / abs(e64),
               amplitudes_dtype=str(t2.dtype).replace("torch.", ""),
               steps=info.get("steps_lowprec", 0) + info.get("steps_fp64", 0),
               seconds_to_solution=info.get("seconds"))
    print(f"# FP32-storage SoA solve: E {e32!r} vs FP64 {e64!r}: |dE|/|E| {out['dE_rel']:.2e} "
          f"({out['steps']} steps, {out['seconds_to_solution']:.1f} s)", file=sys.stderr,
          flush=True)
    del t1, t2, H
    torch.cuda.empty_cache()
    # adjoint gradient (symmetry-sector Lambda equations) against central differences
    c = torch.tensor(C0, dtype=torch.float64, requires_grad=True)
    d = torch.tensor(D0, dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    E = adjoint.planewave_ccsd_energy(c, d, a.L, a.A, eps=1e-13, form="sectors")
    E.backward()
    tg = time.perf_counter() - t0
    g = {"C0": float(c.grad), "D0": float(d.grad)}
    hstep, fd = 1e-4, {}
    for name in ("C0", "D0"):
        Es = []
        for sgn in (1.0, -1.0):
            cc0 = C0 + (sgn * hstep if name == "C0" else 0.0)
            dd0 = D0 + (sgn * hstep if name == "D0" else 0.0)
            Es.append(float(adjoint.planewave_ccsd_energy(
                torch.tensor(cc0, dtype=torch.float64), torch.tensor(dd0, dtype=torch.float64),
                a.L, a.A, eps=1e-13, form="sectors")))
        fd[name] = (Es[0] - Es[1]) / (2 * hstep)
    out["adjoint_gradient"] = {"form": "symmetry-sector Lambda equations", "adjoint": g,
                               "fd": fd, "fd_step": hstep, "seconds": tg,
                               "max_rel_err": max(abs(g[k] - fd[k]) / abs(fd[k]) for k in g)}
    print(f"# gradient: {out['adjoint_gradient']}", file=sys.stderr, flush=True)
    return out


def sectors_only(a, C0, torch, ccm):
    """The symmetry-sector form alone: iteration rate, work and the converged energy."""
    from nulattice_amd.planewave import PlaneWave, shell_holes
    from nulattice_amd.sectors import SectorHamiltonian
    t0 = time.perf_counter()
    SH = SectorHamiltonian.from_planewave(PlaneWave(a.L, shell_holes(a.L, a.A), C0, a.D0))
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    s = time_sectors(SH, a.steps, a.warmup, torch)
    s_eager = time_sectors(SH, a.steps, a.warmup, torch, graph=False)
    out = {"metric": f"CCSD iters/sec at L={a.L} A={a.A}, symmetry-sector form (FP64)",
           "value": 1.0 / s, "unit": "iters/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": s * 1e3, "higher_is_better": True,
           "dtype": "f64", "amplitudes": SH.layout.n, "mac_per_iter": SH.layout.macs(),
           "ms_per_step_without_graph": s_eager * 1e3, "setup_s": setup}
    if not a.no_converge:
        e, _, _ = ccm.ccsd_solver(None, None, eps=1e-10, maxSteps=300, max_diis=0,
                                  hamiltonian=SH)
        out["E_corr"] = e
    return out


def gradient_check(L, A, prec, engine, ccm, adjoint):
    """dE/dC0 and dE/dD0 at (L, A, C0=-4, D0=1) from the Lambda equations, against central
    finite differences (h=1e-4) of converged total energies.  The normal-ordered blocks are
    affine in (C0, D0), so the shifted Hamiltonians h(theta +- h) = h +- h dh/dtheta are built
    from the derivative Hamiltonian (exact; the derivative blocks themselves are pinned against
    differences of freshly generated blocks in the tests)."""
    from nulattice_amd.planewave import PlaneWave, shell_holes
    import torch
    C0, D0, hstep, eps = -4.0, 1.0, 1e-4, 1e-13
    t0 = time.perf_counter()
    pw = PlaneWave(L, shell_holes(L, A), C0, D0)
    E0, fock, H = pw.device_hamiltonian("cuda")
    v = None
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0

    def solve(ham):
        inf = {}
        out = ccm.ccsd_solver(fock, v, eps=eps, maxSteps=600, max_diis=10, diis="pulsed",
                              hamiltonian=ham, precision=prec, info=inf)
        if out is None or not inf.get("converged"):
            raise RuntimeError("CCSD did not converge in the gradient check")
        return out, inf

    t0 = time.perf_counter()
    (e, t1, t2), finfo = solve(H)
    t_fwd = time.perf_counter() - t0
    parts = {n: pw.coupling_derivative(n) for n in ("C0", "D0")}
    derivs = [(dEref, H.rescaled(dfock, scale)) for dEref, dfock, scale in parts.values()]
    lam = {}
    t0 = time.perf_counter()
    grads = adjoint.coupling_gradient(H, t1, t2, derivs, info=lam, eps=1e-10)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    del derivs, t1, t2
    print(f"# gradient L={L} A={A}: adjoint {grads} ({t_grad:.1f} s, "
          f"{lam['lambda_steps']} Lambda steps)", file=sys.stderr, flush=True)
    res = {}
    for (name, (dEref, dfock, scale)), g in zip(parts.items(), grads):
        Es = []
        for sgn in (1.0, -1.0):
            Hs = H.rescaled([f + sgn * hstep * df for f, df in zip(fock, dfock)],
                            1.0 + sgn * hstep * scale)
            (es, _, _), _ = solve(Hs)
            Es.append(E0 + sgn * hstep * dEref + es)
            del Hs
        fd = (Es[0] - Es[1]) / (2 * hstep)
        res[name] = {"adjoint": g, "fd": fd, "rel_err": abs(g - fd) / abs(fd)}
        print(f"# d/d{name}: adjoint {g:.12g} fd {fd:.12g} rel {res[name]['rel_err']:.2e}",
              file=sys.stderr, flush=True)
    return {"L": L, "A": A, "C0": C0, "D0": D0, "fd_step": hstep, "E_total": E0 + e,
            "grad": res, "max_rel_err": max(r["rel_err"] for r in res.values()),
            "forward_prec": prec, "forward_s": t_fwd, "forward_steps":
            (finfo.get("steps_lowprec"), finfo.get("steps_fp64")),
            "lambda_steps": lam["lambda_steps"], "lambda_residual": lam["lambda_residual"],
            "lambda_s": lam["lambda_seconds"], "gradient_s": t_grad, "setup_s": setup}


def diis_cost(p, h, ccm, torch, n=10):
    """Wall time of one pulsed-DIIS extrapolation over n T2-sized vectors (the n-th recorded
    step: its copy and error vector, the n(n+1)/2 B-matrix entries in one Gram pass over the
    error vectors, the host solve, the one-pass linear combination and the history restart),
    which the converging runs pay once every n steps."""
    dev = "cuda"
    t1 = torch.zeros((p, h), dtype=torch.float64, device=dev)
    t2 = torch.full((p, p, h, h), 1e-3, dtype=torch.float64, device=dev)
    d = ccm._Diis("pulsed", n, t1, t2, False)
    cur = t2.clone()
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    for k in range(n):
        new = cur + 1e-4 * 0.5 ** k * torch.rand(cur.shape, generator=gen, dtype=cur.dtype,
                                                 device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ext = d.step(t1, new, t1, cur)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        cur = new
    assert ext is not None
    return dt * 1e3


def lowprec_solve(ccm, fock, v, H, prec, info=None):
    """The mixed-precision fixed point without FP64 refinement: split forms converge to
    |dE|/|E| < 1e-10 like FP64 (their rounding sits below that), the single-rounding forms stop
    at 1e-7 or after 40 steps (their fixed point jitters at the rounding level)."""
    split = prec.endswith("x3")
    e, _, _ = ccm.ccsd_solver(fock, v, eps=1e-10 if split else 1e-7,
                              maxSteps=150 if split else 40, max_diis=0, hamiltonian=H,
                              precision=prec, refine=False, info=info)
    return e


def seed_errors(a, ccm, engine, torch, e64_seed0, lp_seed0):
    """|dE| of the mixed-precision fixed point for C0 from seeds 0..4: the dense mixed-precision
    iteration against the converged FP64 energy, which for seeds 1..4 comes from the
    symmetry-sector form (the same FP64 fixed point to rounding)."""
    import numpy as np
    from nulattice_amd.planewave import PlaneWave, planewave_device_hamiltonian, shell_holes
    from nulattice_amd.sectors import SectorHamiltonian
    out = [{"seed": a.seed, "C0": float(np.random.default_rng(a.seed).uniform(-4.5, -3.5)),
            "E_corr_fp64": e64_seed0, "dE_rel_lowprec": abs(lp_seed0 - e64_seed0) / abs(e64_seed0)}]
    for seed in range(5):
        if seed == a.seed:
            continue
        C0 = float(np.random.default_rng(seed).uniform(-4.5, -3.5))
        SH = SectorHamiltonian.from_planewave(PlaneWave(a.L, shell_holes(a.L, a.A), C0, a.D0))
        e64, _, _ = ccm.ccsd_solver(None, None, eps=1e-12, maxSteps=400, max_diis=0,
                                    hamiltonian=SH)
        del SH
        E0, fock, H = planewave_device_hamiltonian(a.L, a.A, C0, a.D0)
        elp = lowprec_solve(ccm, fock, None, H, a.prec)
        del H
        torch.cuda.empty_cache()
        out.append({"seed": seed, "C0": C0, "E_corr_fp64": e64,
                    "dE_rel_lowprec": abs(elp - e64) / abs(e64)})
        print(f"# seed {seed}: {out[-1]}", file=sys.stderr, flush=True)
    return out


def sector_gradient(L, A, adjoint, torch, dense=None):
    """E and dE/dC0, dE/dD0 at (C0=-4, D0=1) in the symmetry-sector form (packed Lambda
    equations), timed end to end (Hamiltonian build, amplitude solve, Lambda solve, derivative
    residuals), compared with the dense adjoint and its finite differences when given."""
    C0 = torch.tensor(-4.0, dtype=torch.float64, requires_grad=True)
    D0 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    info = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E = adjoint.planewave_ccsd_energy(C0, D0, L, A, eps=1e-13, form="sectors", info=info)
    E.backward()
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    g = {"C0": float(C0.grad), "D0": float(D0.grad)}
    out = {"E_total": float(E.detach()), "dE_dC0": g["C0"], "dE_dD0": g["D0"], "seconds": sec,
           "lambda_steps": info.get("lambda_steps"), "lambda_residual": info.get("lambda_residual")}
    if dense is not None:
        out["rel_vs_dense_adjoint"] = max(abs(g[n] - dense["grad"][n]["adjoint"]) /
                                          abs(dense["grad"][n]["adjoint"]) for n in g)
        out["rel_vs_fd"] = max(abs(g[n] - dense["grad"][n]["fd"]) / abs(dense["grad"][n]["fd"])
                               for n in g)
    print(f"# sector gradient: {out}", file=sys.stderr, flush=True)
    return out


def cpu_baseline(mode, seed, D0, L, A):
    """The oracle (the CPU restatement of the reference iteration, numpy FP64) timed on this
    host at the bench size: one iteration as the reference's solver runs it (v_ppph_dgrams, then
    t1Iter, t2Iter with its v_pppp_dgrams, the mixing and ccsd_energy; coupled_cluster.py:
    601-623).  ``sample``: every dense contraction runs in full; the two COO loops (Python loops
    over every v_pppp / v_ppph entry, whose cost per entry does not depend on the entry) run
    over every 16th entry and their loop time is scaled by 16.  ``full``: everything in full."""
    import numpy as np
    from nulattice_amd.planewave import planewave_hamiltonian
    from oracle import ccsd_ref as cc
    cores = len(os.sched_getaffinity(0))
    if mode == "none":
        return None
    stride = 16 if mode == "sample" else 1
    C0 = float(np.random.default_rng(seed).uniform(-4.5, -3.5))
    E0, fock, v = planewave_hamiltonian(L, A, C0, D0)
    sub = lambda blk: np.concatenate([blk[0][::stride], blk[1][::stride, None]], 1).tolist()
    vpppp, vppph = sub(v[0]), sub(v[1])
    f_pp, f_ph, f_hh = fock
    _, _, v_pphh, v_phph, v_phhh, v_hhhh = v
    t1 = cc.t1Init(f_ph, f_pp, f_hh, 0)
    t2 = cc.t2Init(f_pp, f_hh, v_pphh, 0)

    def timed(f, *args):
        t0 = time.perf_counter()
        r = f(*args)
        return time.perf_counter() - t0, r

    # COO loops: (sample) - (empty list) isolates the loop from the intermediates' allocation
    t_ppph, res = timed(cc.v_ppph_dgrams, vppph, t1, t2)
    t_ppph0, _ = timed(cc.v_ppph_dgrams, [], t1, t2)
    t_pppp, _ = timed(cc.v_pppp_dgrams, vpppp, t1, t2)
    t_pppp0, _ = timed(cc.v_pppp_dgrams, [], t1, t2)
    loops = stride * (max(t_ppph - t_ppph0, 0.0) + max(t_pppp - t_pppp0, 0.0))

    def dense_part():
        t1n = 0.5 * t1 + 0.5 * cc.t1Iter(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, res)
        t2n = 0.5 * t2 + 0.5 * cc.t2Iter(t1, t2, f_ph, f_hh, f_pp, [], v_phph, v_phhh, v_pphh,
                                         res, v_hhhh)
        return cc.ccsd_energy(f_ph, v_pphh, t2n, t1n)
    t_dense, _ = timed(dense_part)       # includes v_pppp_dgrams' setup over an empty list
    sec = t_dense + t_ppph0 + loops
    return {"value": 1.0 / sec, "unit": "iters/s", "cores": cores, "kind": "port",
            "sample": (f"one oracle iteration at L={L} A={A} (numpy FP64): dense contractions "
                       f"{t_dense:.1f} s in full, the v_ppph/v_pppp COO loops "
                       + (f"timed over every {stride}th of their {len(v[1][1])}/{len(v[0][1])} "
                          f"entries and scaled by {stride} ({loops:.1f} s)" if stride > 1 else
                          f"in full ({loops:.1f} s)") + f"; total {sec:.1f} s per iteration"),
            "seconds_per_iter": sec, "dense_s": t_dense, "coo_loops_s": loops,
            "host_cpu": _cpu_model(), "os_cpu_count": os.cpu_count(), "affinity_cores": cores,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "torch_threads": _torch_threads()}


def _torch_threads():
    import torch
    return torch.get_num_threads()


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {a.gpus}); the CCSD path here is "
              "single-GPU and multi-GPU sharding is a design note only", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from nulattice_amd import adjoint, ccm, engine, ops
    from nulattice_amd.planewave import planewave_device_hamiltonian

    rng = np.random.default_rng(a.seed)
    C0 = float(rng.uniform(-4.5, -3.5))
    if a.form == "all" and a.L >= 6:
        a.form = "combined"
    if a.D0 is None:
        a.D0 = float(rng.uniform(0.5, 1.5)) if a.form == "combined" else 1.0
    t0 = time.perf_counter()
    if a.form == "combined":
        print(json.dumps(combined(a, C0, a.D0, torch, ccm, engine, ops, adjoint)))
        return
    if a.form == "sectors":
        print(json.dumps(sectors_only(a, C0, torch, ccm)))
        return
    if a.form == "diis":
        p, h = 4 * a.L ** 3 - a.A, a.A
        ms = diis_cost(p, h, ccm, torch)
        print(json.dumps({"diis_ms_per_extrapolation": ms, "vectors": 10, "L": a.L, "A": a.A,
                          "t2_bytes": p * p * h * h * 8}))
        return
    torch.zeros(1, device="cuda")      # the HIP context, outside the Hamiltonian's setup time
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E0, fock, H = planewave_device_hamiltonian(a.L, a.A, C0, a.D0)
    v = None                            # every block and plan is written on the device
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    p, h = H.p, H.h
    nnz_pppp, nnz_ppph = H.pppp.nnz, H.ppph.n
    print(f"# setup L={a.L} A={a.A} p={p} h={h} C0={C0:.6f} D0={a.D0}: {setup_s:.3f} s "
          f"(nnz v_pppp {nnz_pppp}, v_ppph {nnz_ppph})", file=sys.stderr, flush=True)

    s64, k64, _ = time_iterations(ccm, engine, ops, H, "fp64", max(2, a.steps // 2), 1, torch)
    print(f"# fp64: {s64 * 1e3:.1f} ms/iter", file=sys.stderr, flush=True)
    slp, klp, _ = time_iterations(ccm, engine, ops, H, a.prec, a.steps, a.warmup, torch)
    print(f"# {a.prec}: {slp * 1e3:.1f} ms/iter", file=sys.stderr, flush=True)
    if a.form == "iter":
        print(json.dumps({"form": "iter", "L": a.L, "A": a.A, "prec": a.prec,
                          "ms_per_step": slp * 1e3, "fp64_ms_per_step": s64 * 1e3,
                          "speedup_vs_fp64": s64 / slp,
                          "ring_ms": (klp.get("ring") or 0) * 1e3,
                          "ring_delta_ms": (klp.get("ring_delta") or 0) * 1e3}))
        return
    # the single-rounding 16-bit forms, for reference (their fixed points miss 1e-6)
    singles = {}
    for other in ("bf16", "fp16"):
        if other != a.prec:
            so, _, _ = time_iterations(ccm, engine, ops, H, other, max(2, a.steps // 2), 1,
                                       torch)
            singles[other] = {"ms_per_step": so * 1e3}
            print(f"# {other}: {so * 1e3:.1f} ms/iter", file=sys.stderr, flush=True)

    dE = {}
    if not a.no_converge:
        info64, infolp = {}, {}
        e64, _, _ = ccm.ccsd_solver(fock, v, eps=1e-10, maxSteps=300, max_diis=0, hamiltonian=H,
                                    info=info64)
        elp = lowprec_solve(ccm, fock, v, H, a.prec, infolp)
        dE = {"E_corr_fp64": e64,
              "dE_rel_lowprec": abs(elp - e64) / abs(e64),
              "steps_fp64_converge": info64.get("steps_fp64"),
              "steps_lowprec_converge": infolp.get("steps_lowprec"),
              # wall time to the converged energy (eps 1e-10, plain Jacobi, no refinement of
              # the mixed-precision form)
              "seconds_fp64_converge": info64.get("seconds"),
              "seconds_lowprec_converge": infolp.get("seconds")}
        if info64.get("seconds") and infolp.get("seconds"):
            dE["time_to_solution_speedup"] = info64["seconds"] / infolp["seconds"]
        print(f"# convergence: {dE}", file=sys.stderr, flush=True)
        for other in singles:
            eo = lowprec_solve(ccm, fock, v, H, other)
            singles[other]["dE_rel_lowprec"] = abs(eo - e64) / abs(e64)
            print(f"# {other} fixed point: |dE|/|E| = {singles[other]['dE_rel_lowprec']:.2e}",
                  file=sys.stderr, flush=True)
        dE["seeds"] = seed_errors(a, ccm, engine, torch, e64_seed0=e64, lp_seed0=elp)
        worst = max(r["dE_rel_lowprec"] for r in dE["seeds"])
        dE["dE_rel_lowprec_max_over_seeds"] = worst
        dE["meets_target"] = bool(worst <= 1e-6 and s64 / slp >= 5.0)
        dE["diis_ms_per_extrapolation"] = diis_cost(p, h, ccm, torch)
        print(f"# DIIS extrapolation: {dE['diis_ms_per_extrapolation']:.1f} ms",
              file=sys.stderr, flush=True)

    # the symmetry-sector form of the same iteration (FP64, block-sparse T2)
    from nulattice_amd.planewave import PlaneWave, shell_holes
    from nulattice_amd.sectors import SectorHamiltonian
    t0 = time.perf_counter()
    SH = SectorHamiltonian.from_planewave(PlaneWave(a.L, shell_holes(a.L, a.A), C0, a.D0))
    torch.cuda.synchronize()
    sec_setup = time.perf_counter() - t0
    ssec = time_sectors(SH, max(a.steps, 20), a.warmup, torch)
    sector = {"ms_per_step": ssec * 1e3, "iters_per_s": 1.0 / ssec, "hip_graph": True,
              "speedup_vs_dense_fp64": s64 / ssec, "speedup_vs_dense_lowprec": slp / ssec,
              "amplitudes": SH.layout.n, "dense_amplitudes": p * p * h * h,
              "mac_per_iter": SH.layout.macs(), "setup_s": sec_setup}
    if not a.no_converge:
        es, _, _ = ccm.ccsd_solver(fock, v, eps=1e-10, maxSteps=300, max_diis=0, hamiltonian=SH)
        sector.update(E_corr=es, dE_rel_vs_dense_fp64=abs(es - dE["E_corr_fp64"]) /
                      abs(dE["E_corr_fp64"]))
    del SH
    print(f"# sectors: {sector}", file=sys.stderr, flush=True)

    # roofline of the dominant kernel: the T2 ring GEMM Spp = T[(a,i),(c,k)] W[(c,k),(b,j)],
    # timed by events on its launch stream inside the timed loop; its algorithmic work is
    # 2 (ph)^3 FLOP per launch.  The fp16x3 form computes it as split products (3 partial
    # products) every engine.RING_ANCHOR-th step and as one single-rounded product over the
    # concatenated K = 2 kp, [t2_hi | t2 - t2_A] x [W - W_A | W_A], on the steps in between: that
    # launch (tag "ring_delta") runs on RING_ANCHOR - 1 of every RING_ANCHOR steps and is the
    # dominant kernel.
    ph = p * h
    kp = (ph + 63) // 64 * 64
    ring_flop = 2.0 * ph ** 3
    full_s, delta_s = klp.get("ring"), klp.get("ring_delta")
    nprod = 3 if a.prec.endswith("x3") else 1
    peak = 2.5e15
    roof = None
    fp65 = (a.L, a.A, a.prec) == (5, 28, "fp16x3")
    if delta_s:
        ach = ring_flop / delta_s
        exe = 2.0 * ph * ph * 2 * kp
        traffic = (profile_value("ccsd_L5_A28_ring_hbm_traffic.md", "ring_delta_traffic_bytes")
                   if fp65 else None)
        roof = {"bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": ach / peak, "traffic": traffic,
                "traffic_source": "profiles/ccsd_L5_A28_ring_hbm_traffic.md (calibrated "
                                  "FETCH_SIZE + WRITE_SIZE of this launch)" if traffic else None,
                "executed_achieved": exe / delta_s / 1e12, "executed_frac": exe / delta_s / peak,
                "kernel": "ring GEMM Spp between anchors (one FP16 product over K = 2 kp = "
                          f"{2 * kp}: [t2_hi | t2 - t2_A] x [W - W_A | W_A])",
                "kernel_ms": delta_s * 1e3, "algorithmic_flop": ring_flop, "executed_flop": exe,
                "anchor_period": engine.RING_ANCHOR}
    if full_s:
        ach = ring_flop / full_s
        full = {"kernel": f"ring GEMM Spp ({a.prec}: {nprod} partial products of 2(ph)^3 FLOP)",
                "kernel_ms": full_s * 1e3, "achieved": ach / 1e12, "frac": ach / peak,
                "executed_achieved": nprod * ach / 1e12, "executed_frac": nprod * ach / peak,
                "traffic": (profile_value("ccsd_L5_A28_ring_hbm_traffic.md",
                                          "ring_spp_traffic_bytes") if fp65 else None)}
        if roof is None:
            roof = {"bound": "mfma", "achieved": full["achieved"], "peak": peak / 1e12,
                    "unit": "TFLOP/s", "frac": full["frac"], "traffic": full["traffic"],
                    "traffic_source": "profiles/ccsd_L5_A28_ring_hbm_traffic.md (calibrated "
                                      "FETCH_SIZE + WRITE_SIZE of this launch)"
                                      if full["traffic"] else None,
                    "executed_achieved": full["executed_achieved"],
                    "executed_frac": full["executed_frac"], "kernel": full["kernel"],
                    "kernel_ms": full_s * 1e3, "algorithmic_flop": ring_flop,
                    "executed_flop": nprod * ring_flop}
        else:
            roof["anchor_step_ring"] = full
    if roof is not None:
        roof["fp64_ring_ms"] = (k64.get("ring") or 0) * 1e3
        roof["fp64_ring_frac_of_78.6TF"] = (ring_flop / k64["ring"] / 78.6e12
                                            if k64.get("ring") else None)
    grad = None
    if a.grad == "on":
        del H
        # amplitudes and the finite-difference energies from the FP64 form (converged to 1e-13;
        # the finite differences divide energy differences by 2e-4)
        grad = gradient_check(a.L, a.A, "fp64", engine, ccm, adjoint)
        sector["adjoint_gradient"] = sector_gradient(a.L, a.A, adjoint, torch, grad)
    cpu = cpu_baseline(a.cpu_sample, a.seed, a.D0, a.L, a.A)
    if cpu is not None:
        print(f"# cpu baseline: {cpu['sample']}", file=sys.stderr, flush=True)
    # matrix-core work executed per iteration (counted at every GEMM launch) against the peaks,
    # over the measured iteration time; the rest of the iteration is HBM-bound
    fl, fl64 = klp["flops"], k64["flops"]
    whole = {"lowprec": {"lp_flop": fl["lp"], "fp64_flop": fl["fp64"],
                         "mfma_frac": (fl["lp"] / 2.5e15 + fl["fp64"] / 78.6e12) / slp},
             "fp64": {"fp64_flop": fl64["fp64"], "mfma_frac": fl64["fp64"] / 78.6e12 / s64},
             "canonical_flop_reference_terms": canonical_flops(p, h, nnz_pppp, nnz_ppph),
             "note": "executed FLOP of the regrouped schedule (the ring family is two (ph)^3 "
                     "products instead of the reference's four; split products count each "
                     "partial product); canonical_flop is the reference term list's work"}
    out = {
        "metric": f"CCSD iters/sec at L={a.L} A={a.A}: mixed-prec vs FP64 GPU, |dE| vs FP64 oracle",
        "value": 1.0 / slp, "unit": "iters/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": slp * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": f"{a.prec}+fp64",
        "data": "synthetic (plane-wave contact Hamiltonian, seeded C0)",
        "config": {"workload": f"CCSD L={a.L} A={a.A} plane-wave, dense T2 {p}x{p}x{h}x{h}, "
                               f"{a.prec} ring/ladder/dressing MFMA (FP32 accumulation) + "
                               f"FP64 update/energy/denominators",
                   "L": a.L, "A": a.A, "p": p, "h": h, "C0": C0, "D0": a.D0, "seed": a.seed},
        "fp64_iters_per_s": 1.0 / s64, "fp64_ms_per_step": s64 * 1e3,
        "speedup_vs_fp64": s64 / slp,
        "single_rounding_forms": singles,
        "whole_iteration": whole,
        "roofline": roof, "cpu_baseline": cpu, "adjoint_gradient": grad,
        "sector_form": sector,
        "setup_s": setup_s,
    }
    out.update(dE)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""CCSD diagram library, T1/T2 iteration, energy and solver (oracle; test infrastructure only).

Restates NuLattice ``src/NuLattice/CCM/ccDgrams.py`` and
``src/NuLattice/CCM/coupled_cluster.py:265-687``.  ``opt_einsum.contract`` (opt-einsum>=3.4,
pyproject.toml:9, not installed here) is replaced by ``numpy.einsum(..., optimize=True)``; both
compute the same sums, only the floating-point summation order differs.  Every coefficient and
sign follows the code, not the docstrings (ccDgrams.py:307 vs :316, :361 vs :370).

``diis="legacy"`` reproduces the current legacy solver, whose DIIS block is never entered
(coupled_cluster.py:590-593, 635-641); ``diis="pulsed"`` is the pulsed DIIS that produced the
reference's stored notebook trajectories (the legacy block run every ``max_diis`` steps);
``diis="soa"`` follows ``src/NuLattice/CCM/soa/coupled_cluster.py:520-567``.
"""
import numpy as np


def contract(spec, *ops):
    return np.einsum(spec.replace(" ", ""), *ops, optimize=True)


def pAB(x):
    """x_abij - x_baij (ccDgrams.py:193-202)."""
    return x - x.transpose(1, 0, 2, 3)


def pIJ(x):
    """x_abij - x_abji (ccDgrams.py:204-213)."""
    return x - x.transpose(0, 1, 3, 2)


# ------------------------------------------------------------------ COO scatter diagrams
def v_ppph_dgrams(v_ppph, t1, t2):
    """Seven intermediates from one pass over COO <cd||ak> (ccDgrams.py:13-45)."""
    p, h = t1.shape
    r0 = np.zeros((p, h)); r1 = np.zeros((p, p)); r2 = np.zeros((p, p, h, h))
    r3 = np.zeros((p, p)); r4 = np.zeros((p, p, h, h)); r5 = np.zeros((p, h, h, h))
    r6 = np.zeros((p, h, h, h))
    tt = contract("ci,dj->cdij", t1, t1)
    for c, d, a, k, v in v_ppph:
        c, d, a, k = int(c), int(d), int(a), int(k)
        r1[a, d] -= v * t1[c, k]
        r3[d, a] += v * t1[c, k]
        r0[a, :] -= 0.5 * v * t2[c, d, k, :]
        r2[c, d, :, k] += v * t1[a, :]
        r4[a, d, :, k] += v * t1[c, :]
        r5[a, :, :, k] += v * t2[c, d, :, :]
        r6[a, :, :, k] += v * tt[c, d, :, :]
    return r0, r1, r2, r3, r4, r5, r6


def v_pppp_dgrams(v_pppp, t1, t2):
    """pp ladder over COO <ab||cd>: returns (sum v t2[cd], sum v t1 t1[cd]) (ccDgrams.py:215-237)."""
    p, h = t1.shape
    r1 = np.zeros((p, p, h, h)); r2 = np.zeros((p, p, h, h))
    tt = contract("ci,dj->cdij", t1, t1)
    for a, b, c, d, v in v_pppp:
        a, b, c, d = int(a), int(b), int(c), int(d)
        r1[a, b] += v * t2[c, d]
        r2[a, b] += v * tt[c, d]
    return r1, r2


# ------------------------------------------------------------------ T1 diagrams (ccDgrams.py:48-191)
def dgram_akci_ck(v_phph, t1):          return -contract("akci,ck->ai", v_phph, t1)
def dgram_ck_acik(f_ph, t2):            return contract("ck,acik->ai", f_ph, t2)
def dgram_cikl_cakl(v_phhh, t2):        return -0.5 * contract("cikl,cakl->ai", v_phhh, t2)
def dgram_cdkl_ck_dali(v_pphh, t1, t2): return contract("cdkl,ck,dali->ai", v_pphh, t1, t2)
def dgram_ck_ci(f_ph, t1):              return -0.5 * contract("ck,ci->ki", f_ph, t1)
def dgram_ck_ak(f_ph, t1):              return -0.5 * contract("ck,ak->ac", f_ph, t1)
def dgram_bijk_bj(v_phhh, t1):          return -contract("bijk,bj->ki", v_phhh, t1)
def dgram_cdlk_cdli(v_pphh, t2):        return -0.5 * contract("cdlk,cdli->ki", v_pphh, t2)
def dgram_dckl_dakl(v_pphh, t2):        return -0.5 * contract("dckl,dakl->ac", v_pphh, t2)
def dgram_cdlk_cl_di(v_pphh, t1):       return -0.5 * contract("cdlk,cl,di->ki", v_pphh, t1, t1)
def dgram_cdkl_dk_al(v_pphh, t1):       return 0.5 * contract("cdkl,dk,al->ac", v_pphh, t1, t1)


# ------------------------------------------------------------------ T2 diagrams (ccDgrams.py:240-578)
def dgram_klij_abkl(v_hhhh, t2):        return 0.5 * contract("klij,abkl->abij", v_hhhh, t2)
def dgram_bkcj_acik(v_phph, t2):        return -pIJ(pAB(contract("bkcj,acik->abij", v_phph, t2)))
def dgram_bkij_ak(v_phhh, t1):          return pAB(contract("bkij,ak->abij", v_phhh, t1))
def dgram_cdkl_acik_dblj(v_pphh, t2):
    return 0.5 * pIJ(pAB(contract("cdkl,acik,dblj->abij", v_pphh, t2, t2)))
def dgram_cdkl_cdij_abkl(v_pphh, t2):
    return 0.25 * contract("cdkl,cdij,abkl->abij", v_pphh, t2, t2)
def dgram_klij_ak_bl(v_hhhh, t1):       return 0.5 * pAB(contract("klij,ak,bl->abij", v_hhhh, t1, t1))
def dgram_bkci_ak_cj(v_phph, t1):       return -pAB(pIJ(contract("bkci,ak,cj->abij", v_phph, t1, t1)))
def dgram_cikl_ck_ablj(v_phhh, t1, t2): return -pIJ(contract("cikl,ck,ablj->abij", v_phhh, t1, t2))
def dgram_da_dbij(r3, t2):              return -pAB(contract("da,dbij->abij", r3, t2))
def dgram_acik_bcjk(r4, t2):            return pIJ(pAB(contract("acik,bcjk->abij", r4, t2)))
def dgram_cikl_al_bcjk(v_phhh, t1, t2):
    return -pIJ(pAB(contract("cikl,al,bcjk->abij", v_phhh, t1, t2)))
def dgram_cjkl_ci_abkl(v_phhh, t1, t2):
    return 0.5 * pIJ(contract("cjkl,ci,abkl->abij", v_phhh, t1, t2))
def dgram_bijk_ak1(r5, t1):             return 0.5 * pAB(contract("bijk,ak->abij", r5, t1))
def dgram_bijk_ak2(r6, t1):             return 0.5 * pIJ(pAB(contract("bijk,ak->abij", r6, t1)))
def dgram_cjkl_ci_ak_bl(v_phhh, t1):
    return 0.5 * pIJ(pAB(contract("cjkl,ci,ak,bl->abij", v_phhh, t1, t1, t1)))
def dgram_cdkl_ci_dj_abkl(v_pphh, t1, t2):
    return 0.25 * pIJ(contract("cdkl,ci,dj,abkl->abij", v_pphh, t1, t1, t2))
def dgram_cdkl_ak_bl_cdij(v_pphh, t1, t2):
    return 0.25 * pAB(contract("cdkl,ak,bl,cdij->abij", v_pphh, t1, t1, t2))
def dgram_cdkl_ci_bl_adkj(v_pphh, t1, t2):
    return pIJ(pAB(contract("cdkl,ci,bl,adkj->abij", v_pphh, t1, t1, t2)))
def dgram_cdkl_ci_ak_dj_bl(v_pphh, t1):
    return 0.25 * pIJ(pAB(contract("cdkl,ci,ak,dj,bl->abij", v_pphh, t1, t1, t1, t1)))
def dgram_cdkl_bdkl(v_pphh, t2):        return -0.5 * contract("cdkl,bdkl->bc", v_pphh, t2)
def dgram_cdkl_cdjl(v_pphh, t2):        return -0.5 * contract("cdkl,cdjl->kj", v_pphh, t2)
def dgram_ck_bk(f_ph, t1):              return -contract("ck,bk->bc", f_ph, t1)
def dgram_ck_cj(f_ph, t1):              return -contract("ck,cj->kj", f_ph, t1)
def dgram_cdlk_cl_dj(v_pphh, t1):       return -contract("cdlk,cl,dj->kj", v_pphh, t1, t1)
def dgram_cdlk_dk_bl(v_pphh, t1):       return -contract("cdlk,dk,bl->bc", v_pphh, t1, t1)


# ------------------------------------------------------------------ iteration
def ccsd_energy(f_ph, v_pphh, t2, t1):
    """E = f.t1 + 1/4 v.t2 + 1/2 v t1 t1 (coupled_cluster.py:265-286)."""
    return (contract("ai,ai->", f_ph, t1) + contract("abij,abij->", v_pphh, t2) * 0.25
            + contract("abij,ai,bj->", v_pphh, t1, t1) * 0.5)


def t1Init(f_ph, f_pp, f_hh, delta):
    """t1 = f_ph / (f_ii - f_aa + delta) (coupled_cluster.py:319-340)."""
    return f_ph / (np.add.outer(-np.diag(f_pp), np.diag(f_hh)) + delta)


def t2Init(f_pp, f_hh, v_pphh, delta):
    """t2 = v_pphh / (f_ii + f_jj - f_aa - f_bb + delta) (coupled_cluster.py:416-439)."""
    dh, dp = np.diag(f_hh), np.diag(f_pp)
    return v_pphh / (np.add.outer(-np.add.outer(dp, dp), np.add.outer(dh, dh)) + delta)


def t1_intermediates(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, v_ppph_results,
                     sparse=True):
    """(H1 without the X.t1 terms, X_pp, X_hh) of t1Iter (coupled_cluster.py:372-406)."""
    H1 = (f_ph + dgram_akci_ck(v_phph, t1) + dgram_ck_acik(f_ph, t2)
          + dgram_cikl_cakl(v_phhh, t2) + dgram_cdkl_ck_dali(v_pphh, t1, t2))
    X_hh = -f_hh.copy()
    X_pp = f_pp.copy()
    X_hh += dgram_ck_ci(f_ph, t1)
    X_pp += dgram_ck_ak(f_ph, t1)
    X_hh += dgram_bijk_bj(v_phhh, t1)
    X_hh += dgram_cdlk_cdli(v_pphh, t2)
    X_pp += dgram_dckl_dakl(v_pphh, t2)
    X_hh += dgram_cdlk_cl_di(v_pphh, t1)
    X_pp += dgram_cdkl_dk_al(v_pphh, t1)
    if sparse:
        H1 = H1 + v_ppph_results[0]
        X_pp += v_ppph_results[1]
    else:
        H1 = H1 - 0.5 * contract("cdak,cdki->ai", v_ppph_results, t2)
        X_pp -= contract("cdak,ck->ad", v_ppph_results, t1)
    return H1, X_pp, X_hh


def t1_residual(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, v_ppph_results, sparse=True):
    """Full T1 residual H1 and the dressed diagonals used by the Jacobi update."""
    H1, X_pp, X_hh = t1_intermediates(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh,
                                      v_ppph_results, sparse)
    H1 = H1 + contract("ac,ci->ai", X_pp, t1) + contract("ki,ak->ai", X_hh, t1)
    return H1, X_pp, X_hh


def t1Iter(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh, v_ppph_results, sparse=True):
    """t1 + H1 / (-(X_aa + X_ii)) (coupled_cluster.py:342-414)."""
    H1, X_pp, X_hh = t1_residual(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh, v_pphh,
                                 v_ppph_results, sparse)
    denom = -np.add.outer(np.diag(X_pp), np.diag(X_hh))
    return t1 + H1 / denom


def t2_residual(t1, t2, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh, v_ppph_results,
                v_hhhh, sparse=True):
    """Full T2 residual H2 and the dressed X_pp, X_hh (coupled_cluster.py:475-536)."""
    H2 = v_pphh.copy()
    H2 += dgram_klij_abkl(v_hhhh, t2)
    H2 += dgram_bkcj_acik(v_phph, t2)
    H2 += dgram_bkij_ak(v_phhh, t1)
    H2 += dgram_cdkl_acik_dblj(v_pphh, t2)
    H2 += dgram_cdkl_cdij_abkl(v_pphh, t2)
    H2 += dgram_klij_ak_bl(v_hhhh, t1)
    H2 += dgram_bkci_ak_cj(v_phph, t1)
    H2 += dgram_cikl_ck_ablj(v_phhh, t1, t2)
    H2 += dgram_cikl_al_bcjk(v_phhh, t1, t2)
    H2 += dgram_cjkl_ci_abkl(v_phhh, t1, t2)
    H2 += dgram_cjkl_ci_ak_bl(v_phhh, t1)
    H2 += dgram_cdkl_ci_dj_abkl(v_pphh, t1, t2)
    H2 += dgram_cdkl_ak_bl_cdij(v_pphh, t1, t2)
    H2 += dgram_cdkl_ci_bl_adkj(v_pphh, t1, t2)
    H2 += dgram_cdkl_ci_ak_dj_bl(v_pphh, t1)
    X_hh = -f_hh.copy()
    X_pp = f_pp.copy()
    X_pp += dgram_cdkl_bdkl(v_pphh, t2)
    X_hh += dgram_cdkl_cdjl(v_pphh, t2)
    X_pp += dgram_ck_bk(f_ph, t1)
    X_hh += dgram_ck_cj(f_ph, t1)
    X_hh += dgram_cdlk_cl_dj(v_pphh, t1)
    X_pp += dgram_cdlk_dk_bl(v_pphh, t1)
    if sparse:
        r = v_ppph_results
        H2 += pIJ(r[2])
        H2 += dgram_da_dbij(r[3], t2)
        H2 += dgram_acik_bcjk(r[4], t2)
        H2 += dgram_bijk_ak1(r[5], t1)
        H2 += dgram_bijk_ak2(r[6], t1)
        lad1, lad2 = v_pppp_dgrams(v_pppp, t1, t2)
        H2 += 0.5 * lad1
        H2 += 0.5 * pIJ(lad2)
    else:
        v3 = v_ppph_results
        H2 += pIJ(contract("abcj,ci->abij", v3, t1))
        H2 += -pAB(contract("cdak,ck,dbij->abij", v3, t1, t2))
        H2 += pIJ(pAB(contract("dcak,di,bcjk->abij", v3, t1, t2)))
        H2 += 0.5 * pAB(contract("cdbk,ak,cdij->abij", v3, t1, t2))
        H2 += 0.5 * pIJ(pAB(contract("cdbk,ci,ak,dj->abij", v3, t1, t1, t1)))
        H2 += 0.5 * contract("abcd,cdij->abij", v_pppp, t2)
        H2 += 0.5 * pIJ(contract("abcd,ci,dj->abij", v_pppp, t1, t1))
    H2 += pAB(contract("bc,acij->abij", X_pp, t2)) + pIJ(contract("kj,abik->abij", X_hh, t2))
    return H2, X_pp, X_hh


def t2Iter(t1, t2, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh, v_ppph_results, v_hhhh,
           sparse=True):
    """t2 + H2 / (-(X_aa + X_bb + X_ii + X_jj)) (coupled_cluster.py:442-545)."""
    H2, X_pp, X_hh = t2_residual(t1, t2, f_ph, f_hh, f_pp, v_pppp, v_phph, v_phhh, v_pphh,
                                 v_ppph_results, v_hhhh, sparse)
    dh, dp = np.diag(X_hh), np.diag(X_pp)
    denom = -np.add.outer(np.add.outer(dp, dp), np.add.outer(dh, dh))
    return t2 + H2 / denom


def diis_window_step(hist1, hist2, errs, max_diis, soa, ccs=False):
    """One step of the pulsed DIIS window, after the newest amplitudes (hist1/hist2[-1]) and
    error vector (errs[-1]) were appended.  Returns None while the window fills; when it holds
    max_diis errors it returns the extrapolation sum_k c_k hist[k+1] with c minimising
    |sum_k c_k e_k| subject to sum_k c_k = 1 (bordered system on B_ab = e_a . e_b / max|B|):
      * pulsed (coupled_cluster.py:635-672, with errors appended every step): the oldest entry
        of the window is dropped before the solve, so n = max_diis - 1 errors enter;
      * soa (soa/coupled_cluster.py:520-567): all max_diis errors enter, B / (max|B| + 1e-16).
    A singular solve returns the newest amplitudes unchanged.  The caller restarts the window
    from the returned amplitudes."""
    if soa and len(errs) > max_diis:
        hist1.pop(0); hist2.pop(0); errs.pop(0)
    if len(errs) != max_diis:
        return None
    if not soa:
        # the window drops its oldest entry before the solve (coupled_cluster.py:642-647)
        hist1.pop(0); hist2.pop(0); errs.pop(0)
    n = len(errs)
    B = np.empty((n, n))
    for a in range(n):
        for b in range(a, n):
            B[a, b] = B[b, a] = np.dot(errs[a], errs[b])
    B = B / (np.abs(B).max() + (1e-16 if soa else 0.0))
    A = -np.ones((n + 1, n + 1)); A[:n, :n] = B; A[n, n] = 0.0
    rhs = np.zeros(n + 1); rhs[n] = -1.0
    try:
        c = np.linalg.solve(A, rhs)[:n]
    except np.linalg.LinAlgError:
        return hist1[-1], hist2[-1]
    t1 = sum(c[k] * hist1[k + 1] for k in range(n))
    t2 = hist2[-1] if ccs else sum(c[k] * hist2[k + 1] for k in range(n))
    return t1, t2


def ccsd_solver(fock_mats, two_body_int, t1initial=None, eps=1e-8, maxSteps=1000, max_diis=10,
                delta=0, mixing=0.5, verbose=False, sparse=True, ccs=False, diis="legacy",
                history=None):
    """Damped-Jacobi CCSD (coupled_cluster.py:547-687).

    diis="legacy": the current reference solver (DIIS never runs; divergence/NaN -> None).
    diis="pulsed": the pulsed DIIS that produced the reference's stored notebook trajectories:
        every ``max_diis`` steps the window [t_0..t_n], [e_1..e_n] drops its oldest entry and the
        remaining n-1 error vectors give the extrapolation (the block of
        coupled_cluster.py:635-672 run with errors appended every step).
    diis="soa": the SoA variant, which keeps all ``max_diis`` vectors and normalises B by
        max|B| + 1e-16 (soa/coupled_cluster.py:520-567).
    ``history`` (a list) receives the energy of every step.
    """
    f_pp, f_ph, f_hh = fock_mats
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = two_body_int
    t1 = t1Init(f_ph, f_pp, f_hh, delta) if t1initial is None else np.array(t1initial, float)
    t2 = np.zeros_like(v_pphh) if (ccs or t1initial is not None) else t2Init(f_pp, f_hh, v_pphh,
                                                                           delta)
    pulsed = diis in ("pulsed", "soa") and max_diis > 0
    soa = diis == "soa"
    if pulsed:
        hist1, hist2, errs = [t1.copy()], [t2.copy()], []
    prev = ccsd_energy(f_ph, v_pphh, t2, t1)
    if history is not None:
        history.append(prev)
    if verbose:
        print(f"Step 0: {prev}")
    energy = prev
    for it in range(maxSteps):
        oldT1, oldT2 = t1.copy(), t2.copy()
        res = v_ppph_dgrams(v_ppph, t1, t2) if sparse else v_ppph
        t1 = mixing * t1 + (1.0 - mixing) * t1Iter(t1, t2, f_ph, f_pp, f_hh, v_phph, v_phhh,
                                                   v_pphh, res, sparse)
        if not ccs:
            t2 = mixing * t2 + (1.0 - mixing) * t2Iter(oldT1, t2, f_ph, f_hh, f_pp, v_pppp,
                                                       v_phph, v_phhh, v_pphh, res, v_hhhh,
                                                       sparse)
        energy = ccsd_energy(f_ph, v_pphh, t2, t1)
        if history is not None:
            history.append(energy)
        if verbose:
            print(f"Step {it + 1}: {energy}")
        if (energy == 0 and prev == 0) or abs(energy - prev) / abs(energy) < eps:
            return energy, t1, t2
        if pulsed:
            hist1.append(t1.copy()); hist2.append(t2.copy())
            errs.append(np.concatenate([(t1 - oldT1).ravel(), (t2 - oldT2).ravel()]))
            ext = diis_window_step(hist1, hist2, errs, max_diis, soa, ccs)
            if ext is not None:
                t1, t2 = ext
                hist1, hist2, errs = [t1.copy()], [t2.copy()], []
        if abs(energy) > 1e10:
            print("Iteration Diverged")
            return None
        if np.isnan(energy):
            print("nan error")
            return None
        prev = energy
    print("Max Iterations Reached")
    return energy, t1, t2

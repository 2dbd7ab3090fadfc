"""Plane-wave (momentum-basis) Hamiltonian of the lattice contact theory, emitted directly in the
reference's particle/hole block layout (the input step of the CCSD path).

Basis: orbital index ((kx*L + ky)*L + kz)*4 + sigma, sigma = tz*2 + sz, the same numbering as
NuLattice ``lattice.state2index`` (lattice.py:119-151) with (i,j,k) read as (kx,ky,kz).

Operators (the plane-wave transforms of the reference's position-space operators):
  kinetic      eps(k) = 6 - 2 sum_d cos(2 pi k_d / L)      (Fourier transform of lattice.Tkin)
  2-body       <pq||rs> = (C/L^3) delta(k_p + k_q = k_r + k_s mod L)
                          [d(s_p,s_r) d(s_q,s_s) - d(s_p,s_s) d(s_q,s_r)]
               (lattice.contacts with vT1 = vS1 = C0 is this SU(4) contact on every site)
  3-body       W = D0 sum_x sum_{s1<s2<s3} n_{x s1} n_{x s2} n_{x s3}   (lattice.NNNcontact)
normal-ordered with respect to the closed-shell plane-wave determinant (n_k filled momenta,
all four sigma each, rho = n_k / L^3) in the NO2B form of coupled_cluster.get_norm_ordered_ham
(coupled_cluster.py:752-832):
  v  -> contact strength C0 + 2 D0 rho                (sum_i <p q i||r s i>)
  f_pp = eps + 3 n_k C0 / L^3 + 3 D0 rho^2            (T + sum_i <p i||p i> + 1/2 sum_ij w)
  E_ref = sum_i f_ii - 1/2 sum_ij v_ijij + 4 D0 n_k^3 / L^6
f_ph = 0 (momentum conservation), so T1 = 0 exactly for these references.
"""
import numpy as np

from .plans import CooBlock


def kvectors(L):
    k = np.arange(L)
    return np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)


def shell_holes(L, A):
    """Closed-shell hole list in the standard order: k = 0 (A=4), or k in
    [0, +x, -x, +y, -y, +z, -z] (A=28), each with sigma = 0..3."""
    if A == 4:
        ks = [(0, 0, 0)]
    elif A == 28:
        m = L - 1
        ks = [(0, 0, 0), (1, 0, 0), (m, 0, 0), (0, 1, 0), (0, m, 0), (0, 0, 1), (0, 0, m)]
        if L < 3:
            raise ValueError("A=28 needs L >= 3 (the +k and -k shells coincide at L=2)")
    else:
        raise ValueError("closed shells A=4 or A=28")
    return [((kx * L + ky) * L + kz) * 4 + s for (kx, ky, kz) in ks for s in range(4)]


def eps_k(L):
    kv = kvectors(L)
    return 6.0 - 2.0 * np.cos(2 * np.pi * kv / L).sum(1)


class PlaneWave:
    """Momentum-basis Hamiltonian blocks for (L, holes, C0, D0)."""

    def __init__(self, L, holes, C0, D0=0.0):
        self.L, self.C0, self.D0 = L, float(C0), float(D0)
        n = 4 * L ** 3
        self.holes = list(holes)
        hs = set(self.holes)
        self.parts = [q for q in range(n) if q not in hs]
        self.h, self.p = len(self.holes), len(self.parts)
        kv = kvectors(L)
        orb_k = np.repeat(np.arange(L ** 3), 4)
        orb_s = np.tile(np.arange(4), L ** 3)
        occ_k = sorted({q // 4 for q in self.holes})
        for kk in occ_k:   # closed shell: every filled momentum carries all four sigma
            if not all(kk * 4 + s in hs for s in range(4)):
                raise ValueError("holes must fill all four spin-isospin states of each momentum")
        self.nk = len(occ_k)
        self.rho = self.nk / L ** 3
        self.g = (self.C0 + 2.0 * self.D0 * self.rho) / L ** 3      # NO2B contact strength
        self._kv, self._orb_k, self._orb_s = kv, orb_k, orb_s
        self._kd = [np.ascontiguousarray(kv[:, d]) for d in range(3)]
        self.eps = eps_k(L)
        e_h = self.eps[orb_k[self.holes]]
        e_p = self.eps[orb_k[self.parts]]
        shift = 3.0 * self.nk * self.C0 / L ** 3 + 3.0 * self.D0 * self.rho ** 2
        self.f_pp = np.diag(e_p + shift)
        self.f_hh = np.diag(e_h + shift)
        self.f_ph = np.zeros((self.p, self.h))

    def labels(self):
        """(particle labels (p, 5), hole labels (h, 5), moduli): the additive conserved labels
        (kx, ky, kz, sz, tz) of every state, momenta taken mod L.  They identify each state, so
        the Hamiltonian is block diagonal in them (``nulattice_amd.sectors``)."""
        def lab(orbs):
            o = np.asarray(orbs)
            s = self._orb_s[o]
            return np.column_stack([self._kv[self._orb_k[o]], s % 2, s // 2]).astype(np.int64)
        L = self.L
        return lab(self.parts), lab(self.holes), np.array([L, L, L, 0, 0], np.int64)

    # -------------------------------------------------------------- helpers
    def _K(self, a, b):
        """Encoded total momentum (k_a + k_b mod L) for orbital arrays a, b (broadcasting)."""
        L = self.L
        ka, kb = self._orb_k[a], self._orb_k[b]
        out = 0
        for kd in self._kd:          # one axis at a time, on flat integer arrays
            t = kd[ka] + kd[kb]
            t -= L * (t >= L)        # = (k_a + k_b) mod L for momenta in [0, L)
            out = out * L + t
        return out

    def block(self, rows1, rows2, cols1, cols2):
        """Dense <r1 r2 || c1 c2> over orbital lists."""
        r1, r2, c1, c2 = (np.asarray(x) for x in (rows1, rows2, cols1, cols2))
        Kb = self._K(r1[:, None], r2[None, :])
        Kk = self._K(c1[:, None], c2[None, :])
        s = self._orb_s
        sr1, sr2, sc1, sc2 = s[r1], s[r2], s[c1], s[c2]
        spin = ((sr1[:, None, None, None] == sc1[None, None, :, None])
                & (sr2[None, :, None, None] == sc2[None, None, None, :])).astype(np.int8)
        spin -= ((sr1[:, None, None, None] == sc2[None, None, None, :])
                 & (sr2[None, :, None, None] == sc1[None, None, :, None])).astype(np.int8)
        mom = Kb[:, :, None, None] == Kk[None, None, :, :]
        return np.where(mom, spin * self.g, 0.0)

    # -------------------------------------------------------------- reference blocks
    def dense_blocks(self):
        P, H = self.parts, self.holes
        v_pphh = self.block(P, P, H, H)
        v_phph = self.block(P, H, P, H)
        v_phhh = self.block(P, H, H, H)
        v_hhhh = self.block(H, H, H, H)
        return v_pphh, v_phph, v_phhh, v_hhhh

    def coo_pppp(self):
        """All nonzeros of v_pppp[a,b,c,d] as a CooBlock (idx (N,4) int64, val (N,))."""
        return CooBlock(*self._coo(self.parts, self.parts))

    def coo_ppph(self):
        """All nonzeros of v_ppph[a,b,c,i] as a CooBlock (idx (N,4) int64, val (N,))."""
        return CooBlock(*self._coo(self.parts, self.holes))

    def _coo(self, cols1, cols2):
        P = np.asarray(self.parts)
        c1, c2 = np.asarray(cols1), np.asarray(cols2)
        s = self._orb_s
        # row pairs (a,b) of particles with s_a != s_b, column pairs with s != s
        ra, rb = np.meshgrid(np.arange(len(P)), np.arange(len(P)), indexing="ij")
        ra, rb = ra.ravel(), rb.ravel()
        keep = s[P[ra]] != s[P[rb]]
        ra, rb = ra[keep], rb[keep]
        ca, cb = np.meshgrid(np.arange(len(c1)), np.arange(len(c2)), indexing="ij")
        ca, cb = ca.ravel(), cb.ravel()
        keep = s[c1[ca]] != s[c2[cb]]
        ca, cb = ca[keep], cb[keep]
        # group by (total momentum, unordered sigma pair): only equal groups couple
        def key(x, y):
            sx, sy = s[x], s[y]
            lo, hi = np.minimum(sx, sy), np.maximum(sx, sy)
            return self._K(x, y) * 16 + lo * 4 + hi
        rk = key(P[ra], P[rb])
        ck = key(c1[ca], c2[cb])
        ro, co = np.argsort(rk, kind="stable"), np.argsort(ck, kind="stable")
        rk_s, ck_s = rk[ro], ck[co]
        ur, rs = np.unique(rk_s, return_index=True)
        re = np.append(rs[1:], rk_s.size)
        uc, cs = np.unique(ck_s, return_index=True)
        ce = np.append(cs[1:], ck_s.size)
        cmap = {int(u): (a, b) for u, a, b in zip(uc, cs, ce)}
        out_i, out_v = [], []
        for u, a0, a1 in zip(ur, rs, re):
            if int(u) not in cmap:
                continue
            b0, b1 = cmap[int(u)]
            R = ro[a0:a1]; C = co[b0:b1]
            x1, x2 = P[ra[R]], P[rb[R]]
            y1, y2 = c1[ca[C]], c2[cb[C]]
            # sign: +1 if (s_x1, s_x2) == (s_y1, s_y2), -1 if swapped
            same = (s[x1][:, None] == s[y1][None, :]).astype(np.float64) * 2.0 - 1.0
            I = np.stack(np.broadcast_arrays(ra[R][:, None], rb[R][:, None], ca[C][None, :],
                                             cb[C][None, :]), -1).reshape(-1, 4)
            out_i.append(I)
            out_v.append((same * self.g).reshape(-1))
        if not out_i:
            return np.zeros((0, 4), np.int64), np.zeros(0)
        return np.concatenate(out_i).astype(np.int64), np.concatenate(out_v)

    # -------------------------------------------------------------- device build
    def codes(self, orbs):
        """Orbital codes kx | ky << 8 | kz << 16 | sigma << 24 (int32) of the kernels that write
        the blocks on the GPU."""
        o = np.asarray(orbs)
        k = self._kv[self._orb_k[o]]
        return (k[:, 0] | (k[:, 1] << 8) | (k[:, 2] << 16) | (self._orb_s[o] << 24)).astype(
            np.int32)

    def pair_groups(self, first, second):
        """Pairs (x, y) of ``first`` x ``second`` with sigma_x != sigma_y, grouped by the key
        (total momentum, unordered sigma pair) that the contact conserves: (keys ascending,
        start, end, pairs), pairs = x * len(second) + y ascending inside each group.  These are
        the row / column classes of ``_coo`` (the two-body matrix couples a row pair only to
        column pairs of the same key)."""
        A, B = np.asarray(first), np.asarray(second)
        s = self._orb_s
        ia = np.repeat(np.arange(len(A)), len(B))
        ib = np.tile(np.arange(len(B)), len(A))
        keep = s[A[ia]] != s[B[ib]]
        ia, ib = ia[keep], ib[keep]
        sx, sy = s[A[ia]], s[B[ib]]
        key = self._K(A[ia], B[ib]) * 16 + np.minimum(sx, sy) * 4 + np.maximum(sx, sy)
        o = np.argsort(key, kind="stable")
        key, pair = key[o], (ia * len(B) + ib)[o]
        u, start = np.unique(key, return_index=True)
        return u, start, np.append(start[1:], key.size), pair

    def device_hamiltonian(self, device="cuda", min_block_rows=128):
        """(E_ref, [f_pp, f_ph, f_hh], engine.Hamiltonian) with every two-body block written on
        the GPU: the four dense blocks by ``nl_pw_dense``, the v_pppp pair blocks by
        ``nl_pw_pair_fill`` over the analytically known components, the v_ppph COO list by
        ``nl_pw_coo`` (then its segment plans sorted on the device).  The result equals
        ``engine.Hamiltonian(*hamiltonian()[1:])`` bit for bit (same blocks, same plans)."""
        import torch
        from . import _lib, engine, ops
        from .plans import PairBlocks, PpphPlans
        dev = torch.device(device)
        if dev.type != "cuda":
            raise _lib.NLError("device_hamiltonian: a HIP device is required (no CPU fallback)")
        _lib.load()
        L, g, p, h = self.L, self.g, self.p, self.h
        if L > 255:
            raise ValueError("device_hamiltonian supports L <= 255")
        F64 = torch.float64
        cP = torch.tensor(self.codes(self.parts), device=dev)
        cH = torch.tensor(self.codes(self.holes), device=dev)
        dims = {"p": (p, cP), "h": (h, cH)}
        dense = []
        for pattern in ("pphh", "phph", "phhh", "hhhh"):
            ns, cs = zip(*(dims[x] for x in pattern))
            out = torch.empty(ns, dtype=F64, device=dev)
            _lib.call("nl_pw_dense", *ns, *(ops._p(c) for c in cs), L, g, ops._p(out),
                      ops._stream())
            dense.append(out)

        def filler(rdiv, r0, r1, cdiv, q0, q1):
            def fill(rows, cols, nb, mb, kb):
                V = torch.empty((nb, mb, kb), dtype=F64, device=dev)
                _lib.call("nl_pw_pair_fill", nb, mb, kb, ops._p(rows), ops._p(cols), rdiv, cdiv,
                          ops._p(r0), ops._p(r1), ops._p(q0), ops._p(q1), L, g, ops._p(V),
                          ops._stream())
                return V
            return fill

        P, H = self.parts, self.holes
        ku, ks, ke, kpair = self.pair_groups(P, P)            # row classes of v_pppp and v_ppph
        pp_groups = {int(u): kpair[a:b] for u, a, b in zip(ku, ks, ke)}
        # v_pppp: every class couples its pairs with themselves; components ordered by first row
        comp = sorted(pp_groups.values(), key=lambda r: int(r[0]))
        pppp = PairBlocks.from_components(comp, comp, filler(p, cP, cP, p, cP, cP), dev,
                                          min_block_rows)
        # v_ppph = <cd||ak>: rows (c,d) of a class, columns (a,k) of the same class
        hu, hs_, he, hpair = self.pair_groups(P, H)
        ph_groups = {int(u): hpair[a:b] for u, a, b in zip(hu, hs_, he)}
        common = [u for u in (int(x) for x in ku) if u in ph_groups]
        if common:
            R = [pp_groups[u] for u in common]
            C = [ph_groups[u] for u in common]
            nr = np.array([len(r) for r in R], np.int64)
            nc = np.array([len(c) for c in C], np.int64)
            eoff = np.concatenate([[0], np.cumsum(nr * nc)]).astype(np.int64)
            roff = np.concatenate([[0], np.cumsum(nr)]).astype(np.int64)
            coff = np.concatenate([[0], np.cumsum(nc)]).astype(np.int64)
            n = int(eoff[-1])
            T = lambda x: torch.tensor(x, dtype=torch.int64, device=dev)  # noqa: E731
            eoff_d, roff_d, coff_d = T(eoff), T(roff), T(coff)
            rp_d, cp_d = T(np.concatenate(R)), T(np.concatenate(C))
            idx = torch.empty((4, n), dtype=torch.int64, device=dev)
            val = torch.empty(n, dtype=F64, device=dev)
            _lib.call("nl_pw_coo", len(common), ops._p(eoff_d), ops._p(roff_d), ops._p(coff_d),
                      ops._p(rp_d), ops._p(cp_d), p, h, ops._p(cP), ops._p(cP), ops._p(cP),
                      ops._p(cH), L, g, n, ops._p(idx), ops._p(val), ops._stream())
            # R56 rows (a,k), columns (c,d): components ordered by first (a,k) row
            order = sorted(range(len(common)), key=lambda q: int(C[q][0]))
            r56 = PairBlocks.from_components([C[q] for q in order], [R[q] for q in order],
                                             filler(h, cP, cH, p, cP, cP), dev)
            ppph = PpphPlans(idx, val, p, h, dev, r56=r56)
        else:
            ppph = PpphPlans(np.zeros((0, 4), np.int64), np.zeros(0), p, h, dev)
        fock = [self.f_pp, self.f_ph, self.f_hh]
        H = engine.Hamiltonian(fock, [pppp, ppph] + dense, device=dev,
                               min_block_rows=min_block_rows)
        return self.ref_energy(), fock, H

    def ref_energy(self):
        H = self.holes
        v_hhhh = self.block(H, H, H, H)
        hh = np.arange(self.h)
        e = np.trace(self.f_hh) - 0.5 * v_hhhh[hh[:, None], hh[None, :], hh[:, None],
                                                hh[None, :]].sum()
        return e + 4.0 * self.D0 * self.nk ** 3 / self.L ** 6

    def coupling_derivative(self, name):
        """d/dC0 or d/dD0 of the normal-ordered Hamiltonian: (dE_ref, [df_pp, df_ph, df_hh], scale).

        Every block is affine in (C0, D0): the two-body blocks are g times a fixed pattern with
        g = (C0 + 2 D0 rho) / L^3, the Fock diagonals carry shift = 3 n_k C0 / L^3 + 3 D0 rho^2,
        and E_ref = sum_i (eps_i + shift) - 1/2 sum_ij v_ijij + 4 D0 n_k^3 / L^6.  The derivative
        Hamiltonian therefore has constant Fock diagonals d(shift) and two-body blocks equal to
        ``scale`` = (dg/dtheta) / g times this Hamiltonian's."""
        L, nk, rho = self.L, self.nk, self.rho
        if name == "C0":
            dg, dshift, de3 = 1.0 / L ** 3, 3.0 * nk / L ** 3, 0.0
        elif name == "D0":
            dg, dshift, de3 = 2.0 * rho / L ** 3, 3.0 * rho ** 2, 4.0 * nk ** 3 / L ** 6
        else:
            raise ValueError(f"unknown coupling {name!r} (C0 or D0)")
        if self.g == 0.0:
            raise ValueError("the normal-ordered contact strength is zero: the derivative "
                             "two-body blocks are not a multiple of the Hamiltonian's")
        scale = dg / self.g
        H = self.holes
        hh = np.arange(self.h)
        vsum = self.block(H, H, H, H)[hh[:, None], hh[None, :], hh[:, None], hh[None, :]].sum()
        dEref = self.h * dshift - 0.5 * scale * vsum + de3
        fock = [dshift * np.eye(self.p), np.zeros((self.p, self.h)), dshift * np.eye(self.h)]
        return dEref, fock, scale

    def hamiltonian(self, sparse=True):
        """(E_ref, [f_pp, f_ph, f_hh], [v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh]) with the
        two COO blocks as (indices, values) pairs (sparse=True) or dense arrays."""
        v_pphh, v_phph, v_phhh, v_hhhh = self.dense_blocks()
        if sparse:
            v_pppp, v_ppph = self.coo_pppp(), self.coo_ppph()
        else:
            P, H = self.parts, self.holes
            v_pppp, v_ppph = self.block(P, P, P, P), self.block(P, P, P, H)
        return (self.ref_energy(), [self.f_pp, self.f_ph, self.f_hh],
                [v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh])


def planewave_hamiltonian(L, A, C0, D0=0.0, sparse=True):
    """Hamiltonian of the closed-shell plane-wave reference (A = 4 or 28)."""
    return PlaneWave(L, shell_holes(L, A), C0, D0).hamiltonian(sparse)


def planewave_device_hamiltonian(L, A, C0, D0=0.0, device="cuda"):
    """(E_ref, fock_mats, engine.Hamiltonian) of the closed-shell plane-wave reference with the
    two-body blocks and plans written on the GPU (``PlaneWave.device_hamiltonian``)."""
    return PlaneWave(L, shell_holes(L, A), C0, D0).device_hamiltonian(device)

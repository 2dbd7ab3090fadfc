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
        ka, kb = self._kv[self._orb_k[a]], self._kv[self._orb_k[b]]
        t = (ka + kb) % L
        return (t[..., 0] * L + t[..., 1]) * L + t[..., 2]

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

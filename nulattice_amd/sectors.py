"""Symmetry-sector (block-sparse) CCSD iteration.

When every single-particle state carries additive conserved labels that identify it -- for plane
waves the momentum k mod L and the spin and isospin projections (``PlaneWave.labels()``) -- the
normal-ordered Hamiltonian conserves the summed labels of each pair and:

* t2[a,b,i,j] is nonzero only in the *pair sector* l_a + l_b = l_i + l_j; the packed amplitudes
  hold, sector after sector, the dense [(a,b)][(i,j)] block over the pairs a < b, i < j of that
  sector, each antisymmetric element once (L=5 A=28: 49,268 amplitudes instead of the 1.75e8
  of the dense t2);
* the ring products of t2Iter are block diagonal in the *crossed sector* l_a - l_i;
* the dressings X_pp, X_hh and the Fock matrices are diagonal, and t1 stays exactly zero: every
  term of the T1 residual (t1Iter, coupled_cluster.py:342-414) needs l_a = l_i, which labels that
  identify the states forbid.

With t1 = 0 the T2 residual of t2Iter (coupled_cluster.py:475-545) is
  H2 = v_pphh + 1/2 v_pppp t2 (v_pppp_dgrams, ccDgrams.py:215-237)
       + t2 (1/2 v_hhhh + 1/4 v_pphh^T t2)          (A21 :240-251, A25 :292-303)
       + P(ij)P(ab) t2 (-v_phph + 1/2 v_pphh t2)   (A22 :253-264, A24 :279-290; crossed channel)
       + (X_aa + X_bb + X_ii + X_jj) t2             (the X_pp / X_hh terms, coupled_cluster.py:536)
with every sum over an ordered pair (c,d) or (k,l) equal to twice the sum over c < d or k < l, so
on the stored elements the ladders are v_pppp t2 + t2 (v_hhhh + v_pphh^T t2) over a < b, i < j
blocks and the dressings X_aa = f_aa - sum over stored elements containing a of v t2 (and
X_ii likewise); the ring keeps ordered crossed pairs, read from the packed amplitudes through
signed maps.
and every product is a set of small dense GEMMs, one per sector, run by grouped-GEMM launches
(one per kernel shape: short K in one pass, long K split over the wavefronts of a workgroup,
64x8 tiles for the narrow sectors).  An iteration is the dressing, the ladder and ring products,
the crossed-channel gather, the fused assembly and the energy reduction, all FP64, replayed as
one HIP graph; results equal the dense path's to rounding (only summation order differs).

``SectorLayout`` is the host-side index structure (NumPy only); ``SectorHamiltonian`` packs a
reference-format Hamiltonian onto the GPU; ``ccm.ccsd_solver(..., hamiltonian=SectorHamiltonian)``
runs the reference solver loop on it and returns t2 in the reference layout.
"""
import ctypes
import math
import time
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib, ops
from .plans import _np, coo_arrays

F64 = torch.float64
I32 = torch.int32
I64 = torch.int64


def _norm(x, moduli):
    x = np.array(x, dtype=np.int64, copy=True)
    m = moduli > 0
    if m.any():
        x[..., m] = np.mod(x[..., m], moduli[m])
    return x


def _group(keys, nkeys):
    """(position of each item within its key's group (-1 for key -1), group sizes, items sorted
    by key (stable), group start offsets)."""
    valid = keys >= 0
    order = np.argsort(np.where(valid, keys, nkeys), kind="stable")
    order = order[valid[order]]
    counts = np.bincount(keys[valid], minlength=nkeys).astype(np.int64)
    start = np.cumsum(counts) - counts
    pos = np.full(keys.size, -1, dtype=np.int64)
    pos[order] = np.arange(order.size) - start[keys[order]]
    return pos, counts, order, start


def _outer(nr, nc):
    """For blocks of nr[s] x nc[s] laid out back to back (row-major): (block, row, col) of every
    entry."""
    sizes = nr * nc
    blk = np.repeat(np.arange(nr.size), sizes)
    local = np.arange(int(sizes.sum()), dtype=np.int64) - (np.cumsum(sizes) - sizes)[blk]
    ncb = nc[blk]
    return blk, local // np.maximum(ncb, 1), local % np.maximum(ncb, 1)


# output tile (rows, columns) of the grouped-GEMM kernel modes: 0 = 64x32 tiles, 1 = 32x32 and
# 2 = 64x8 tiles with K split over the four wavefronts of a workgroup, 3 = 32x32 tiles with K
# split over eight wavefronts (lists of only a few long-K tiles, whose K chain sets their time)
MODE_TILE = {0: (64, 32), 1: (32, 32), 2: (64, 8), 3: (32, 32)}


def _mode(M, N, K):
    """Kernel mode for one product: short K in one pass, long K split, narrow N in 64x8 tiles."""
    if K <= 32:
        return 0
    return 2 if N <= 8 else 1


def _tiles(desc, mode=0):
    """Tile table (problem, m0, n0) of a grouped-GEMM problem list for one kernel mode."""
    tm, tn = MODE_TILE[mode]
    out = []
    for q, d in enumerate(desc):
        M, N = int(d[0]), int(d[1])
        for m0 in range(0, M, tm):
            for n0 in range(0, N, tn):
                out.append((q, m0, n0))
    return np.asarray(out, dtype=np.int32).reshape(-1, 3)


def _launch_lists(desc):
    """The problem list split by kernel mode: [(mode, desc, tiles), ...]."""
    if len(desc) == 0:
        return []
    modes = np.array([_mode(*map(int, d[:3])) for d in desc])
    out = []
    for m in (0, 1, 2):
        sub = desc[modes == m]
        if len(sub):
            til = _tiles(sub, m)
            out.append((3 if m == 1 and len(til) < 64 else m, sub, til))
    return out


class SectorLayout:
    """Pair and crossed sectors, packed offsets, index maps and grouped-GEMM problem lists for
    particles with labels ``part_labels`` (p, d) and holes with ``hole_labels`` (h, d); label
    component c is taken modulo ``moduli[c]`` (0 = no modulus).

    Attributes: n (packed amplitudes, a < b and i < j), orb (4, n) = (a, b, i, j) of each packed
    element, dense_index (n,) = ((a*p + b)*h + i)*h + j, image_maps (the four signed reference-
    layout images), ring_map (ring image -> signed packed element),
    sg (4, n) ring-image offsets of Spp[a,i,b,j], Spp[a,j,b,i], Spp[b,i,a,j], Spp[b,j,a,i],
    seg / elem (dressing segments: elements by a and b, then by i and j), and the problem lists
    ``desc_pp``, ``desc_whh``, ``desc_hh``, ``desc_w``, ``desc_s``.
    """

    DRESS_CHUNK = 256

    def __init__(self, part_labels, hole_labels, moduli=None):
        lp = np.asarray(part_labels, dtype=np.int64)
        lh = np.asarray(hole_labels, dtype=np.int64)
        lp = lp.reshape(lp.shape[0], -1)
        lh = lh.reshape(lh.shape[0], -1)
        if lp.shape[1] != lh.shape[1]:
            raise ValueError("particle and hole labels need the same number of components")
        d = lp.shape[1]
        moduli = np.zeros(d, np.int64) if moduli is None else np.asarray(moduli, np.int64)
        if moduli.shape != (d,) or (moduli < 0).any():
            raise ValueError("moduli: one non-negative integer per label component")
        lp, lh = _norm(lp, moduli), _norm(lh, moduli)
        allx = np.concatenate([lp, lh])
        if len(np.unique(allx, axis=0)) != len(allx):
            raise ValueError("the labels must identify every single-particle state (no two "
                             "states may share all label components)")
        p, h = len(lp), len(lh)
        self.p, self.h, self.d = p, h, d
        self.labels_p, self.labels_h, self.moduli = lp, lh, moduli

        # ---------------- pair sectors (l_a + l_b = l_i + l_j), pairs a < b, i < j
        pl = _norm(lp[:, None, :] + lp[None, :, :], moduli).reshape(p * p, d)
        hl = _norm(lh[:, None, :] + lh[None, :, :], moduli).reshape(h * h, d)
        u, inv = np.unique(np.concatenate([pl, hl]).reshape(-1, d), axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        kp, kh = inv[:p * p].copy(), inv[p * p:].copy()
        pa = np.arange(p * p)
        kp[pa // p >= pa % p] = -1                       # keep a < b
        ha = np.arange(h * h)
        kh[ha // h >= ha % h] = -1                       # keep i < j
        present = np.intersect1d(kp[kp >= 0], kh[kh >= 0])
        S = present.size
        sec_of = np.full(len(u), -1, np.int64)
        sec_of[present] = np.arange(S)
        self.sp = np.where(kp >= 0, sec_of[np.maximum(kp, 0)], -1)     # sector of pair a*p+b
        self.sh = np.where(kh >= 0, sec_of[np.maximum(kh, 0)], -1)     # sector of pair i*h+j
        self.pp_pos, self.nP, self.pp_list, self.pp_start = _group(self.sp, S)
        self.hh_pos, self.nH, self.hh_list, self.hh_start = _group(self.sh, S)
        self.S = S
        sizes = self.nP * self.nH
        self.toff = np.cumsum(sizes) - sizes
        self.n = int(sizes.sum())
        self.ppoff = np.cumsum(self.nP ** 2) - self.nP ** 2
        self.npp = int((self.nP ** 2).sum())
        self.hhoff = np.cumsum(self.nH ** 2) - self.nH ** 2
        self.nhh = int((self.nH ** 2).sum())
        blk, r, c = _outer(self.nP, self.nH)
        ab = self.pp_list[self.pp_start[blk] + r]
        ij = self.hh_list[self.hh_start[blk] + c]
        a, b, i, j = ab // p, ab % p, ij // h, ij % h
        self.orb = np.stack([a, b, i, j]).astype(np.int32)
        self.dense_index = ((a * p + b) * h + i) * h + j
        # the four antisymmetric images of each stored element in the reference layout, as
        # signed maps (m >= 0: +x at m; m <= -2: -x at -m-2)
        dn = lambda w, x, y, z: ((w * p + x) * h + y) * h + z
        self.image_maps = np.stack([dn(a, b, i, j), -dn(b, a, i, j) - 2, -dn(a, b, j, i) - 2,
                                    dn(b, a, j, i)]).astype(np.int64)

        # ---------------- crossed sectors (l_a - l_i)
        ql = _norm(lp[:, None, :] - lh[None, :, :], moduli).reshape(p * h, d)
        nql = _norm(-ql, moduli)
        u2, inv2 = np.unique(np.concatenate([ql, nql]).reshape(-1, d), axis=0,
                             return_inverse=True)
        inv2 = inv2.reshape(-1)
        kq, knq = inv2[:p * h], inv2[p * h:]
        neg = np.full(len(u2), -1, np.int64)
        neg[kq] = knq
        self.rpos, nR_all, self.r_list, self.r_start = _group(kq, len(u2))
        self.kq = kq
        # ring storage: one block [R_x][R_neg(x)] per crossed sector x with pairs
        xs = np.nonzero(nR_all)[0]
        nx = neg[xs]
        nRx = nR_all[xs]
        nRn = nR_all[nx]
        self.roff_all = np.full(len(u2), -1, np.int64)
        rsz = nRx * nRn
        self.roff_all[xs] = np.cumsum(rsz) - rsz
        self.nring = int(rsz.sum())
        self.woff_all = np.full(len(u2), -1, np.int64)
        wsz = nRx * nRx
        self.woff_all[xs] = np.cumsum(wsz) - wsz
        self.nw = int(wsz.sum())
        self.nR_all, self.neg = nR_all, neg
        # ring image: Tr[(a,i),(c,k)] = t2[a,c,i,k]
        blk, r, c = _outer(nRx, nRn)
        ai = self.r_list[self.r_start[xs[blk]] + r]
        ck = self.r_list[self.r_start[nx[blk]] + c]
        ra, ri, rc, rk = ai // h, ai % h, ck // h, ck % h
        self.ring_map = self.signed_element(ra, rc, ri, rk)
        # W image of v_phph: W[(c,k),(b,j)] = v_phph[b,k,c,j]
        blk, r, c = _outer(nRx, nRx)
        ck = self.r_list[self.r_start[xs[blk]] + r]
        bj = self.r_list[self.r_start[xs[blk]] + c]
        wc, wk, wb, wj = ck // h, ck % h, bj // h, bj % h
        self.phph_index = ((wb * h + wk) * p + wc) * h + wj
        # assembly gathers of Spp[x,y,z,w] = Tr-layout offset of row (x,y), column (z,w)
        def ro(x, y, z, w):
            return (self.roff_all[kq[x * h + y]] + self.rpos[x * h + y] * nR_all[kq[z * h + w]]
                    + self.rpos[z * h + w])
        self.sg = np.stack([ro(a, i, b, j), ro(a, j, b, i), ro(b, i, a, j),
                            ro(b, j, a, i)]).astype(np.int32)

        # ---------------- dressing segments: elements by a and by b (X_pp), by i and by j (X_hh)
        both = np.concatenate([np.arange(self.n), np.arange(self.n)])
        _, cnt_a, ord_a, _ = _group(np.concatenate([a, b]).astype(np.int64), p)
        _, cnt_i, ord_i, _ = _group(np.concatenate([i, j]).astype(np.int64), h)
        self.seg = np.concatenate([[0], np.cumsum(np.concatenate([cnt_a, cnt_i]))]).astype(
            np.int64)
        self.elem = np.concatenate([both[ord_a], both[ord_i]]).astype(np.int32)
        # chunks of at most DRESS_CHUNK consecutive entries, never crossing a segment
        nch = -(-np.diff(self.seg) // self.DRESS_CHUNK)
        self.cptr = np.concatenate([[0], np.cumsum(nch)]).astype(np.int64)
        owner = np.repeat(np.arange(p + h), nch)
        first = self.seg[owner] + (np.arange(int(nch.sum())) - self.cptr[owner]) * \
            self.DRESS_CHUNK
        self.cstart = np.concatenate([first, [self.seg[-1]]]).astype(np.int64)

        # ---------------- grouped-GEMM problem lists (M N K oA sAm sAk oB sBk sBn oC sCm sCn)
        nP, nH, to, po, ho = self.nP, self.nH, self.toff, self.ppoff, self.hhoff
        one = np.ones(S, np.int64)
        self.desc_pp = np.stack([nP, nH, nP, po, nP, one, to, nH, one, to, nH, one], 1)
        self.desc_whh = np.stack([nH, nH, nP, to, one, nH, to, nH, one, ho, nH, one], 1)
        self.desc_hh = np.stack([nP, nH, nH, to, nH, one, ho, nH, one, to, nH, one], 1)
        # reverse sweep: VPP^T x and x Whh^T
        self.desc_ppT = np.stack([nP, nH, nP, po, one, nP, to, nH, one, to, nH, one], 1)
        self.desc_hhT = np.stack([nP, nH, nH, to, nH, one, ho, one, nH, to, nH, one], 1)
        keep = (nRx > 0) & (nRn > 0)
        xk, nk = xs[keep], nx[keep]
        Rx, Rn = nR_all[xk], nR_all[nk]
        ro_x, ro_n = self.roff_all[xk], self.roff_all[nk]
        wo_x, wo_n = self.woff_all[xk], self.woff_all[nk]
        o = np.ones(xk.size, np.int64)
        # W_x = 1/2 VPr_x Tr_neg(x) - VPHPH_x: [R_x][R_x], K = |R_neg(x)|
        self.desc_w = np.stack([Rx, Rx, Rn, ro_x, Rn, o, ro_n, Rx, o, wo_x, Rx, o], 1)
        # Spp_x = Tr_x W_neg(x): [R_x][R_neg(x)], K = |R_neg(x)|
        self.desc_s = np.stack([Rx, Rn, Rn, ro_x, Rn, o, wo_n, Rn, o, ro_x, Rn, o], 1)
        # the fused ring kernel: one entry per pair {x, neg(x)} of crossed sectors
        first = xk <= nk
        self.ring_pairs = np.stack([ro_x, ro_n, wo_x, wo_n, Rx, Rn, (xk == nk).astype(np.int64)],
                                   1)[first].astype(np.int64).reshape(-1, 7)
        self.ring_rows = int(max(Rx.max(), Rn.max())) if len(Rx) else 0
        self.ring_fused = self.ring_rows <= 32
        # reverse-sweep maps: the packed element (signed) whose assembly reads each Spp position,
        # and the four signed crossed-image positions of each packed element
        self.sg_inv = np.full(self.nring, -1, np.int64)
        for k, sign in enumerate((1, -1, -1, 1)):
            pos = self.sg[k].astype(np.int64)
            if np.any(self.sg_inv[pos] != -1):
                raise AssertionError("an Spp position is read by two assembly terms")
            e = np.arange(self.n, dtype=np.int64)
            self.sg_inv[pos] = e if sign > 0 else -e - 2
        rm = self.ring_map
        pos = np.nonzero(rm != -1)[0]
        enc = rm[pos]
        elem_of = np.where(enc >= 0, enc, -enc - 2)
        sgn = np.where(enc >= 0, 1, -1)
        order = np.argsort(elem_of, kind="stable")
        if not np.array_equal(np.bincount(elem_of, minlength=self.n), np.full(self.n, 4)):
            raise AssertionError("a packed element does not have four crossed images")
        p4 = pos[order].reshape(self.n, 4)
        s4 = sgn[order].reshape(self.n, 4)
        self.ring_inv = np.where(s4 > 0, p4, -p4 - 2).T.copy().astype(np.int64)
        for name in ("desc_pp", "desc_whh", "desc_hh", "desc_w", "desc_s", "desc_ppT",
                     "desc_hhT"):
            dsc = getattr(self, name).astype(np.int64).reshape(-1, 12)
            dsc = dsc[(dsc[:, 0] > 0) & (dsc[:, 1] > 0)]
            setattr(self, name, dsc)
            setattr(self, name.replace("desc", "launch"), _launch_lists(dsc))

    def element(self, a, b, i, j):
        """Packed index of t2[a,b,i,j] for a < b, i < j (-1 for other orderings and where the
        element lies outside every sector)."""
        p, h = self.p, self.h
        s1 = self.sp[a * p + b]
        s2 = self.sh[i * h + j]
        ok = (s1 >= 0) & (s1 == s2)
        s = np.maximum(s1, 0)
        e = self.toff[s] + self.pp_pos[a * p + b] * self.nH[s] + self.hh_pos[i * h + j]
        return np.where(ok, e, -1).astype(np.int64)

    def signed_element(self, a, b, i, j):
        """t2[a,b,i,j] in any index order as a signed packed reference: e (the value is +T[e]),
        -e-2 (-T[e]) or -1 (zero: a == b, i == j or outside every sector)."""
        a, b, i, j = (np.asarray(x, np.int64) for x in (a, b, i, j))
        sg = np.where(a > b, -1, 1) * np.where(i > j, -1, 1)
        e = self.element(np.minimum(a, b), np.maximum(a, b), np.minimum(i, j), np.maximum(i, j))
        e = np.where((a == b) | (i == j), -1, e)
        return np.where(e < 0, -1, np.where(sg > 0, e, -e - 2)).astype(np.int64)

    def pack(self, t2):
        """Packed form (a < b, i < j) of a reference-layout t2[a,b,i,j] (NumPy)."""
        return np.asarray(t2, np.float64).reshape(-1)[self.dense_index]

    def unpack(self, x):
        """Reference-layout antisymmetric t2[a,b,i,j] of packed amplitudes (NumPy)."""
        out = np.zeros(self.p * self.p * self.h * self.h)
        for m in self.image_maps:
            out[np.where(m >= 0, m, -m - 2)] = np.where(m >= 0, x, -x)
        return out.reshape(self.p, self.p, self.h, self.h)

    def macs(self):
        """Multiply-adds of one iteration's sector products."""
        tot = 0
        for dsc in (self.desc_pp, self.desc_whh, self.desc_hh, self.desc_w, self.desc_s):
            tot += int((dsc[:, 0] * dsc[:, 1] * dsc[:, 2]).sum())
        return tot


class SectorViolation(ValueError):
    """The Hamiltonian has a nonzero element that does not conserve the labels."""


def _conserves(labels_of, idx, val, moduli, sign=(1, 1, -1, -1)):
    """True where sum_k sign_k l(idx_k) == 0 (mod) or val == 0."""
    tot = sum(s * labels_of[k](idx[:, k]) for k, s in enumerate(sign))
    tot = _norm(tot, moduli)
    return (~np.any(tot != 0, axis=1)) | (val == 0)


def pack_hamiltonian(lay, fock_mats, two_body_int):
    """Host (NumPy) sector images of a reference-format Hamiltonian: dict with V (packed
    v_pphh, a < b, i < j), VPP (v_pppp sector blocks [(a<b)][(c<d)]), VHH (v_hhhh sector blocks
    [(k<l)][(i<j)]), VPH (W_x[(c,k),(b,j)] = v_phph[b,k,c,j]) and fdiag = (f_aa, -f_ii).
    Raises SectorViolation if any nonzero element does not conserve the labels, if a stored block
    is not antisymmetric, if a Fock matrix is not diagonal or if f_ph is not zero."""
    f_pp, f_ph, f_hh = (np.asarray(_np(x), np.float64) for x in fock_mats)
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = two_body_int
    p, h = f_ph.shape
    if (lay.p, lay.h) != (p, h):
        raise ValueError(f"labels are for p={lay.p}, h={lay.h}; the Hamiltonian has p={p}, "
                         f"h={h}")
    if np.any(f_ph != 0):
        raise SectorViolation("f_ph must vanish (t1 = 0 in the sector form)")
    for name, f in (("f_pp", f_pp), ("f_hh", f_hh)):
        if np.count_nonzero(f - np.diag(np.diag(f))):
            raise SectorViolation(f"{name} must be diagonal for state-identifying labels")
    v_pphh = np.asarray(_np(v_pphh), np.float64)
    v_phph = np.asarray(_np(v_phph), np.float64)
    v_phhh = np.asarray(_np(v_phhh), np.float64)
    v_hhhh = np.asarray(_np(v_hhhh), np.float64)
    V = lay.pack(v_pphh)
    flat = v_pphh.reshape(-1)
    for m in lay.image_maps[1:]:
        if not np.array_equal(np.where(m >= 0, flat[np.maximum(m, 0)], -flat[-np.minimum(m, -2) - 2]),
                              V):
            raise SectorViolation("v_pphh is not antisymmetric")
    if 4 * np.count_nonzero(V) != np.count_nonzero(v_pphh):
        raise SectorViolation("v_pphh has nonzero elements outside the pair sectors")
    VPH = v_phph.reshape(-1)[lay.phph_index]
    if np.count_nonzero(VPH) != np.count_nonzero(v_phph):
        raise SectorViolation("v_phph has nonzero elements outside the crossed sectors")
    blk, r, c = _outer(lay.nH, lay.nH)
    kl = lay.hh_list[lay.hh_start[blk] + r]
    ij = lay.hh_list[lay.hh_start[blk] + c]
    VHH = v_hhhh.reshape(-1)[kl * h * h + ij]
    k_, l_, i_, j_ = kl // h, kl % h, ij // h, ij % h
    for x1, x2, y1, y2, sg in ((l_, k_, i_, j_, -1), (k_, l_, j_, i_, -1), (l_, k_, j_, i_, 1)):
        if not np.array_equal(sg * v_hhhh[x1, x2, y1, y2], VHH):
            raise SectorViolation("v_hhhh is not antisymmetric")
    lpf = lambda x: lay.labels_p[x]
    lhf = lambda x: lay.labels_h[x]
    # (elements of hole-pair sectors without particle pairs never meet t2 and are not stored)
    nz = np.nonzero(v_hhhh)
    if nz[0].size and not _conserves([lhf, lhf, lhf, lhf], np.stack(nz, 1), v_hhhh[nz],
                                     lay.moduli).all():
        raise SectorViolation("v_hhhh does not conserve the labels")
    nz = np.nonzero(v_phhh)
    if nz[0].size and not _conserves([lpf, lhf, lhf, lhf], np.stack(nz, 1), v_phhh[nz],
                                     lay.moduli).all():
        raise SectorViolation("v_phhh does not conserve the labels")
    idx, val = coo_arrays(v_ppph, p, h, "ppph")
    if idx.size and not _conserves([lpf, lpf, lpf, lhf], idx, val, lay.moduli).all():
        raise SectorViolation("v_ppph does not conserve the labels")
    idx, val = coo_arrays(v_pppp, p, h, "pppp")
    if idx.size and not _conserves([lpf, lpf, lpf, lpf], idx, val, lay.moduli).all():
        raise SectorViolation("v_pppp does not conserve the labels")
    if (((idx[:, 0] == idx[:, 1]) | (idx[:, 2] == idx[:, 3])) & (val != 0)).any():
        raise SectorViolation("v_pppp is not antisymmetric (nonzero <aa||cd> or <ab||cc>)")
    # each entry lands, with the sign of its reordering, on the stored [(a<b)][(c<d)] element;
    # a list holding all four images of an element (the reference's lists do) must give the
    # same block from each image class, a list of a < b, c < d entries alone gives it directly
    a0, b0 = np.minimum(idx[:, 0], idx[:, 1]), np.maximum(idx[:, 0], idx[:, 1])
    c0, d0 = np.minimum(idx[:, 2], idx[:, 3]), np.maximum(idx[:, 2], idx[:, 3])
    sgn = np.where(idx[:, 0] > idx[:, 1], -1.0, 1.0) * np.where(idx[:, 2] > idx[:, 3], -1.0, 1.0)
    cls = (idx[:, 0] > idx[:, 1]).astype(np.int64) * 2 + (idx[:, 2] > idx[:, 3])
    rows, cols = a0 * p + b0, c0 * p + d0
    s1, s2 = lay.sp[rows], lay.sp[cols]
    use = (s1 >= 0) & (s1 == s2)
    s = s1[use]
    flat = lay.ppoff[s] + lay.pp_pos[rows[use]] * lay.nP[s] + lay.pp_pos[cols[use]]
    part = [np.bincount(flat[cls[use] == k], weights=(sgn * val)[use][cls[use] == k],
                        minlength=lay.npp) for k in range(4)]
    VPP = part[0]
    present = [k for k in (1, 2, 3) if np.any(cls[use] == k)]
    if present and not all(np.array_equal(part[k], VPP) for k in (1, 2, 3)):
        raise SectorViolation("v_pppp is not antisymmetric")
    fdiag = np.concatenate([np.diag(f_pp), -np.diag(f_hh)])
    return {"V": V, "VPP": VPP, "VHH": VHH, "VPH": VPH, "fdiag": fdiag}


def contact_block_specs(lay, pw):
    """Arguments of ``nl_contact_fill`` for the four sector images of a plane-wave contact
    Hamiltonian: {name: (off (nblk+1), ncol, roff, coff, row codes, column codes, crossed)} for
    V (pp x hh pair blocks), VPP (pp x pp), VHH (hh x hh) and VPH (crossed ph x ph, <b k||c j> at
    [(c,k),(b,j)]); a pair's code is 4 s_first + s_second."""
    sp = pw._orb_s[np.asarray(pw.parts)]
    sh = pw._orb_s[np.asarray(pw.holes)]
    p, h = lay.p, lay.h
    pp, hh, rr = lay.pp_list, lay.hh_list, lay.r_list
    code = {"pp": (4 * sp[pp // p] + sp[pp % p]).astype(np.uint8),
            "hh": (4 * sh[hh // h] + sh[hh % h]).astype(np.uint8),
            "ph": (4 * sp[rr // h] + sh[rr % h]).astype(np.uint8)}
    xs = np.nonzero(lay.nR_all)[0]
    nR = lay.nR_all[xs]

    def spec(nr, nc, roff, coff, rk, ck, crossed):
        off = np.concatenate([[0], np.cumsum(nr * nc)]).astype(np.int64)
        return (off, np.asarray(nc, np.int64), np.asarray(roff, np.int64),
                np.asarray(coff, np.int64), code[rk], code[ck], crossed)

    return {"V": spec(lay.nP, lay.nH, lay.pp_start, lay.hh_start, "pp", "hh", 0),
            "VPP": spec(lay.nP, lay.nP, lay.pp_start, lay.pp_start, "pp", "pp", 0),
            "VHH": spec(lay.nH, lay.nH, lay.hh_start, lay.hh_start, "hh", "hh", 0),
            "VPH": spec(nR, nR, lay.r_start[xs], lay.r_start[xs], "ph", "ph", 1)}


class SectorHamiltonian:
    """A normal-ordered Hamiltonian in the reference's block format (fock_mats, two_body_int; the
    sparse blocks as COO lists / (indices, values) or dense) packed into symmetry sectors on one
    GPU.  ``part_labels`` / ``hole_labels`` give every particle / hole state its conserved labels
    (see ``PlaneWave.labels``).  Raises SectorViolation if any nonzero element does not conserve
    them, if a Fock matrix is not diagonal or f_ph is not zero."""

    def __init__(self, fock_mats, two_body_int, part_labels, hole_labels, moduli=None,
                 device="cuda"):
        device = torch.device(device)
        if device.type != "cuda":
            raise _lib.NLError("SectorHamiltonian: a HIP device is required (no CPU fallback)")
        _lib.load()
        lay = SectorLayout(part_labels, hole_labels, moduli)
        hst = pack_hamiltonian(lay, fock_mats, two_body_int)
        t = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=F64, device=device)
        self._finish(lay, device, t(hst["V"]), t(hst["VPP"]), t(hst["VHH"]), t(hst["VPH"]),
                     hst["fdiag"])

    @classmethod
    def from_planewave(cls, pw, device="cuda"):
        """The sector Hamiltonian of a ``planewave.PlaneWave`` written directly on the GPU: every
        element of a sector block conserves momentum, so it is the contact's spin factor times
        the NO2B strength g (``nl_contact_fill``); no dense block or COO list is formed on the
        host.  Equal to ``SectorHamiltonian(*pw.hamiltonian(), *pw.labels())`` element for
        element."""
        device = torch.device(device)
        if device.type != "cuda":
            raise _lib.NLError("SectorHamiltonian: a HIP device is required (no CPU fallback)")
        _lib.load()
        self = cls.__new__(cls)
        lay = SectorLayout(*pw.labels())
        specs = contact_block_specs(lay, pw)
        out = {}
        for name, (off, ncol, roff, coff, rcode, ccode, crossed) in specs.items():
            buf = torch.empty(max(int(off[-1]), 1), dtype=F64, device=device)
            arrs = [torch.tensor(x, device=device) for x in (off, ncol, roff, coff, rcode, ccode)]
            _lib.call("nl_contact_fill", len(off) - 1, *(ops._p(x) for x in arrs),
                      int(crossed), ctypes.c_double(pw.g), int(off[-1]), ops._p(buf),
                      ops._stream())
            torch.cuda.current_stream().synchronize()      # before arrs is released
            out[name] = buf
        fdiag = np.concatenate([np.diag(pw.f_pp), -np.diag(pw.f_hh)])
        self._finish(lay, device, out["V"], out["VPP"], out["VHH"], out["VPH"], fdiag)
        return self

    def _finish(self, lay, device, V, VPP, VHH, VPH, fdiag):
        p, h = lay.p, lay.h
        self.layout, self.p, self.h, self.device = lay, p, h, device
        t = lambda x, dt=F64: torch.tensor(np.ascontiguousarray(x), dtype=dt, device=device)
        self.V, self.VPP, self.VHH, self.VPHPH = V, VPP, VHH, VPH
        self.orb = t(lay.orb, I32)
        self.sg = t(lay.sg, I32)
        self.ring_map = t(lay.ring_map, I64)
        self.dense_index = t(lay.dense_index, I64)
        self.image_maps = [t(m, I64) for m in lay.image_maps]
        self.elem = t(lay.elem, I32)
        self.cptr = t(lay.cptr, I64)
        self.cstart = t(lay.cstart, I64)
        self.fdiag = fdiag
        self.base = t(self.fdiag)
        self.VPr = torch.empty(lay.nring, dtype=F64, device=device)
        self._gather(self.ring_map, self.V, self.VPr)
        self.ring_pairs = t(lay.ring_pairs, I64) if lay.ring_fused else None
        self.sg_inv = t(lay.sg_inv, I64)
        self.ring_inv = [t(m, I64) for m in lay.ring_inv]
        self.problems = {}
        for name in ("pp", "whh", "hh", "w", "s", "ppT", "hhT"):
            self.problems[name] = [(mode, t(dsc, I64), t(til, I32), til.shape[0])
                                   for mode, dsc, til in getattr(lay, "launch_" + name)]
        self._side = (torch.cuda.Stream(device=device), torch.cuda.Stream(device=device))
        self.nparts = int(_lib.load().nl_partials(lay.n))
        self._dot_parts = max(self.nparts, int(_lib.load().nl_partials(p * h)))

    def scaled(self, scale, fdiag):
        """The Hamiltonian whose two-body blocks are ``scale`` times these and whose Fock
        diagonals are ``fdiag`` = (f_aa, -f_ii) (the derivative Hamiltonian of a coupling that
        multiplies the two-body part and shifts the Fock diagonals); shares layout and maps."""
        new = SectorHamiltonian.__new__(SectorHamiltonian)
        new.__dict__.update(self.__dict__)
        for name in ("V", "VPP", "VHH", "VPHPH", "VPr"):
            x = getattr(self, name)
            y = torch.zeros_like(x)
            _lib.call("nl_axpby", x.numel(), ctypes.c_double(scale), ops._p(x),
                      ctypes.c_double(0.0), ops._p(y), ops._stream())
            setattr(new, name, y)
        new.fdiag = np.asarray(fdiag, np.float64)
        new.base = torch.tensor(new.fdiag, dtype=F64, device=self.device)
        return new

    # ------------------------------------------------------------------ launches
    def _gather(self, idx, src, out):
        _lib.call("nl_sector_gather", idx.numel(), ops._p(idx), ops._p(src), ops._p(out),
                  ops._stream())

    def _gemm(self, name, alpha, A, B, beta, C, D=None):
        for mode, dsc, til, nt in self.problems[name]:
            _lib.call("nl_gemm_grouped_f64", mode, nt, ops._p(dsc), ops._p(til),
                      ctypes.c_double(alpha), ops._p(A), ops._p(B), ctypes.c_double(beta),
                      ops._p(D), ops._p(C), ops._stream())

    def buffers(self):
        lay, dev = self.layout, self.device
        e = lambda n: torch.empty(max(n, 1), dtype=F64, device=dev)
        return {"L": e(lay.n), "Whh": e(lay.nhh), "Tr": e(lay.nring), "W": e(lay.nw),
                "S": e(lay.nring), "X": e(self.p + self.h), "Xc": e(len(lay.cstart)),
                "part": e(2 * self._dot_parts + 1), "E": e(1)}

    def init_amplitudes(self, Bf, delta=0.0):
        """t2 = v_pphh / (f_ii + f_jj - f_aa - f_bb + delta) in packed form (t2Init,
        coupled_cluster.py:416-439) and its energy in Bf["E"]."""
        lay = self.layout
        X = self.fdiag.copy()
        X[self.p:] -= 0.5 * delta
        Bf["X"].copy_(torch.tensor(X, dtype=F64))
        zero = torch.zeros(max(lay.n, 1), dtype=F64, device=self.device)
        t2 = torch.empty(max(lay.n, 1), dtype=F64, device=self.device)
        _lib.call("nl_sector_assemble", lay.n, self.p, ops._p(self.orb), None, ops._p(self.V),
                  None, None, ops._p(Bf["X"]), ops._p(zero), ctypes.c_double(0.0),
                  ctypes.c_double(1.0), ops._p(t2), None, ops._p(Bf["part"]), ops._stream())
        self._energy(Bf)
        return t2

    def _energy(self, Bf):
        _lib.call("nl_finalize", self.nparts, ops._p(Bf["part"]), 0, None, None, ops._p(Bf["E"]),
                  ops._stream())

    def energy(self, t2, Bf):
        """E = 1/4 sum v_pphh t2 over all orderings = sum over the stored a < b, i < j elements
        (ccsd_energy with t1 = 0, coupled_cluster.py:265-286)."""
        _lib.call("nl_dot_part", self.layout.n, ops._p(self.V), ops._p(t2), ops._p(Bf["part"]),
                  ops._stream())
        _lib.call("nl_finalize", self.nparts, ops._p(Bf["part"]), 0, None, None, ops._p(Bf["E"]),
                  ops._stream())
        return float(Bf["E"].item())

    def residual(self, t2, Bf):
        """Everything of H2 except v_pphh, the X t2 terms and the antisymmetrisers: Bf["L"]
        (ladders), Bf["S"] (ring, crossed layout) and Bf["X"] (dressed diagonals).

        The three independent chains -- dressing, pp ladder and hh ladder; the Whh dressing; the
        crossed-channel gather and the two ring products -- run on the caller's stream and two
        side streams (default priority), joined before returning, so the small launches overlap
        instead of queueing behind each other's tails."""
        main = torch.cuda.current_stream()
        sa, sb = self._side
        sa.wait_stream(main)
        sb.wait_stream(main)
        with torch.cuda.stream(sa):
            self.chain_whh(t2, Bf)
        with torch.cuda.stream(sb):
            self.chain_ring(t2, Bf)
        self.chain_pp(t2, Bf)
        main.wait_stream(sa)
        self._gemm("hh", 1.0, t2, Bf["Whh"], 1.0, Bf["L"])
        main.wait_stream(sb)
        return self.layout

    # the independent chains of one step
    def chain_pp(self, t2, Bf):
        """Dressed diagonals X and the pp ladder L = v_pppp t2."""
        lay = self.layout
        _lib.call("nl_sector_dress", self.p + self.h, ops._p(self.cptr), len(lay.cstart) - 1,
                  ops._p(self.cstart), ops._p(self.elem), ops._p(self.V), ops._p(t2),
                  ops._p(self.base), ctypes.c_double(-1.0), ops._p(Bf["Xc"]), ops._p(Bf["X"]),
                  ops._stream())
        self._gemm("pp", 1.0, self.VPP, t2, 0.0, Bf["L"])

    def chain_whh(self, t2, Bf):
        """Whh = v_hhhh + v_pphh^T t2."""
        self._gemm("whh", 1.0, self.V, t2, 1.0, Bf["Whh"], self.VHH)

    def chain_ring(self, t2, Bf):
        """Crossed-channel image of t2, W = 1/2 v_pphh t2 - v_phph, Spp = t2 W (one fused
        launch when every crossed sector has at most 32 pairs)."""
        if self.ring_pairs is not None:
            _lib.call("nl_sector_ring", self.layout.ring_pairs.shape[0], self.layout.ring_rows,
                      ops._p(self.ring_pairs),
                      ops._p(self.ring_map), ops._p(t2), ops._p(self.VPr), ops._p(self.VPHPH),
                      ops._p(Bf["S"]), ops._stream())
            return
        self._gather(self.ring_map, t2, Bf["Tr"])
        self._gemm("w", 0.5, self.VPr, Bf["Tr"], -1.0, Bf["W"], self.VPHPH)
        self._gemm("s", 1.0, Bf["Tr"], Bf["W"], 0.0, Bf["S"])

    def update(self, t2, t2n, Bf, mixing, H2out=None):
        """One damped Jacobi step t2n = mixing t2 + (1 - mixing) t2Iter(t2) (coupled_cluster.py:
        618-620) and its energy in Bf["E"]."""
        self.residual(t2, Bf)
        self._assemble(t2, t2n, Bf, mixing, H2out)

    def _assemble(self, t2, t2n, Bf, mixing, H2out=None):
        lay = self.layout
        _lib.call("nl_sector_assemble", lay.n, self.p, ops._p(self.orb), ops._p(self.sg),
                  ops._p(self.V), ops._p(Bf["L"]), ops._p(Bf["S"]), ops._p(Bf["X"]), ops._p(t2),
                  ctypes.c_double(mixing), ctypes.c_double(1.0), ops._p(t2n), ops._p(H2out),
                  ops._p(Bf["part"]), ops._stream())
        self._energy(Bf)

    def unpack(self, t2):
        """Reference-layout t2[a,b,i,j] (p, p, h, h) HIP tensor of packed amplitudes."""
        out = torch.zeros((self.p, self.p, self.h, self.h), dtype=F64, device=self.device)
        for m in self.image_maps:
            _lib.call("nl_sector_scatter", self.layout.n, ops._p(m), ops._p(t2), ops._p(out),
                      ops._stream())
        return out

    def pack(self, t2):
        """Packed amplitudes of a reference-layout (p, p, h, h) tensor or array."""
        src = t2 if isinstance(t2, torch.Tensor) else torch.tensor(np.asarray(t2), dtype=F64)
        src = src.to(device=self.device, dtype=F64).contiguous()
        out = torch.empty(max(self.layout.n, 1), dtype=F64, device=self.device)
        _lib.call("nl_sector_gather", self.layout.n, ops._p(self.dense_index), ops._p(src),
                  ops._p(out), ops._stream())
        return out


class SectorStepper:
    """Damped Jacobi steps on a fixed pair of packed amplitude buffers.  With ``graph`` each
    direction of the step (buffer 0 -> 1 and 1 -> 0) is captured once as a HIP graph
    (torch.cuda.CUDAGraph around the library's launches, which allocate nothing and record no
    memset; the side-stream fork and join of ``residual`` become graph edges), so a step is one
    graph launch plus the energy read-back instead of thirteen kernel launches from Python.
    (Replaying the three independent chains as separate graphs on three streams measured
    slower: the launches then wait on the host instead of on each other.)"""

    def __init__(self, SH, Bf, t_init, mixing, graph=True):
        self.SH, self.Bf, self.mixing = SH, Bf, mixing
        self.T = [t_init, torch.empty_like(t_init)]
        self.cur = 0
        self.graphs = None
        if graph:
            self.graphs = []
            for k in (0, 1):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    SH.update(self.T[k], self.T[1 - k], Bf, mixing)
                self.graphs.append(g)

    @property
    def t(self):
        """The current amplitudes."""
        return self.T[self.cur]

    @property
    def previous(self):
        """The amplitudes before the last step (valid until the next step)."""
        return self.T[1 - self.cur]

    def step(self):
        """Advance the current amplitudes by one step; returns the new energy."""
        if self.graphs is not None:
            self.graphs[self.cur].replay()
        else:
            self.SH.update(self.T[self.cur], self.T[1 - self.cur], self.Bf, self.mixing)
        self.cur ^= 1
        return float(self.Bf["E"].item())

    def set(self, t):
        """Replace the current amplitudes (a DIIS extrapolation)."""
        self.T[self.cur].copy_(t)


def solve(SH, t1initial=None, eps=1e-8, maxSteps=1000, max_diis=10, delta=0, mixing=0.5,
          verbose=False, ccs=False, diis="legacy", history=None, info=None, graph=True,
          unpack=True):
    """The reference solver loop (coupled_cluster.py:547-687) on a SectorHamiltonian: same
    stopping rule, messages and return values (E_corr, t1, t2), t2 in the reference layout
    (``unpack=False``: t2 is returned packed, a < b, i < j)."""
    from .ccm import _Diis, _fill_info
    p, h, dev = SH.p, SH.h, SH.device
    Bf = SH.buffers()
    t1 = torch.zeros((p, h), dtype=F64, device=dev)
    if t1initial is not None:
        t1i = np.asarray(_np(t1initial), np.float64)
        if t1i.shape != (p, h):
            raise ValueError(f"t1initial has shape {t1i.shape}, expected {(p, h)}")
        if np.any(t1i != 0):
            raise SectorViolation("a nonzero t1initial breaks the sector structure (t1 = 0)")
    n = max(SH.layout.n, 1)
    if ccs or t1initial is not None:
        t2 = torch.zeros(n, dtype=F64, device=dev)
        Bf["E"].zero_()
        prev = 0.0
    else:
        t2 = SH.init_amplitudes(Bf, delta)
        prev = float(Bf["E"].item())
    pulsed = diis != "legacy" and max_diis > 0
    dstate = _Diis(diis, max_diis, t1, t2, ccs) if pulsed else None
    stepper = SectorStepper(SH, Bf, t2, mixing, graph=graph and not ccs)
    if history is not None:
        history.append(prev)
    if verbose:
        print(f"Step {0}: {prev}")
    t_start = time.perf_counter()
    energy = prev
    i = 0
    while i < maxSteps:
        if ccs:
            energy = 0.0            # t1 = 0 and t2 = 0 stay fixed
            stepper.previous.copy_(stepper.t)
            stepper.cur ^= 1
        else:
            energy = stepper.step()
        if history is not None:
            history.append(energy)
        if verbose:
            print(f"Step {i + 1}: {energy}", "difference =",
                  abs(energy - prev) / abs(energy) if energy != 0 else 0.0)
        t2 = stepper.t
        rel = abs(energy - prev) / abs(energy) if energy != 0 else 0.0
        if (energy == 0 and prev == 0) or rel < eps:
            if verbose:
                print(f"Energy found in {i + 1} iterations")
            _fill_info(info, 0, i + 1, t_start, energy, True)
            _sector_info(info, SH, t2)
            return energy, t1, SH.unpack(t2) if unpack else t2
        if dstate is not None:
            ext = dstate.step(t1, t2, t1, stepper.previous,
                              SimpleNamespace(part=Bf["part"], E=Bf["E"]))
            if ext is not None:
                stepper.set(ext[1])
                t2 = stepper.t
        if abs(energy) > 1e10:
            print("Iteration Diverged")
            _fill_info(info, 0, i + 1, t_start, energy)
            return None
        if math.isnan(energy):
            print("nan error")
            _fill_info(info, 0, i + 1, t_start, energy)
            return None
        prev = energy
        i += 1
    print("Max Iterations Reached")
    _fill_info(info, 0, maxSteps, t_start, energy, False)
    _sector_info(info, SH, t2)
    return energy, t1, SH.unpack(t2) if unpack else t2


def _sector_info(info, SH, t2):
    if info is not None:
        info["sector_t2"] = t2
        info["sector_amplitudes"] = SH.layout.n


# ------------------------------------------------------------------ Lambda equations (reverse)
def residual_vjp(SH, t2, lam, Bf, out):
    """out = J^T lam of the packed residual at t2 without its diagonal part D lam (D = X_a + X_b
    + X_i + X_j at t2, added by the assembly), each forward product transposed:
      v_pppp^T lam + lam Whh^T + v_pphh (t^T lam)     (pp and hh ladders; Whh = v_hhhh + v_pphh^T t)
      + the ring's reverse sweep gathered back onto the packed elements
      - v * (gX_a + gX_b + gX_i + gX_j),  gX = sums of lam t over the elements of each state
    Needs Bf["Whh"] of the forward step at t2 (``forward_intermediates``)."""
    lay = SH.layout
    SH._gemm("ppT", 1.0, SH.VPP, lam, 0.0, out)
    SH._gemm("whh", 1.0, t2, lam, 0.0, Bf["G1"])
    SH._gemm("hh", 1.0, SH.V, Bf["G1"], 1.0, out)
    SH._gemm("hhT", 1.0, lam, Bf["Whh"], 1.0, out)
    if SH.ring_pairs is None:
        raise _lib.NLError("the sector reverse sweep needs crossed sectors of at most 32 pairs")
    SH._gather(SH.sg_inv, lam, Bf["gS"])
    _lib.call("nl_sector_ring_adj", lay.ring_pairs.shape[0], lay.ring_rows, ops._p(SH.ring_pairs),
              ops._p(SH.ring_map), ops._p(t2), ops._p(SH.VPr), ops._p(SH.VPHPH), ops._p(Bf["gS"]),
              ops._p(Bf["gTr"]), ops._stream())
    for m in SH.ring_inv:
        _lib.call("nl_sector_gather_add", lay.n, ops._p(m), ops._p(Bf["gTr"]),
                  ctypes.c_double(1.0), ops._p(out), ops._stream())
    _lib.call("nl_sector_dress", SH.p + SH.h, ops._p(SH.cptr), len(lay.cstart) - 1,
              ops._p(SH.cstart), ops._p(SH.elem), ops._p(lam), ops._p(t2), ops._p(Bf["zero_x"]),
              ctypes.c_double(1.0), ops._p(Bf["Xc"]), ops._p(Bf["gX"]), ops._stream())
    _lib.call("nl_sector_diag_axpy", lay.n, SH.p, ops._p(SH.orb), ops._p(SH.V), ops._p(Bf["gX"]),
              ctypes.c_double(-1.0), ops._p(out), ops._stream())


def lambda_buffers(SH):
    lay, dev = SH.layout, SH.device
    Bf = SH.buffers()
    e = lambda n: torch.empty(max(n, 1), dtype=F64, device=dev)
    Bf.update(G1=e(lay.nhh), gS=e(lay.nring), gTr=e(lay.nring), gX=e(SH.p + SH.h),
              zero_x=torch.zeros(SH.p + SH.h, dtype=F64, device=dev), JT=e(lay.n),
              R=e(lay.n))
    return Bf


def solve_lambda(SH, t2, eps=1e-10, maxSteps=500, mixing=0.5, max_diis=10, info=None,
                 verbose=False):
    """Packed Lambda amplitudes: J^T lam = -dE/dt at converged packed amplitudes t2 (dE/dt = v,
    E = sum v t over the stored elements).  Damped Jacobi with the amplitude iteration's dressed
    denominators, lam_new = m lam + (1-m) (lam + (J^T lam + v) / -D), pulsed DIIS every
    ``max_diis`` steps; stops when |J^T lam + v| <= eps |v|."""
    from .ccm import _Diis
    lay = SH.layout
    Bf = lambda_buffers(SH)
    SH.residual(t2, Bf)                          # forward X and Whh at t2
    n = max(lay.n, 1)
    lam = torch.zeros(n, dtype=F64, device=SH.device)
    lamn = torch.empty_like(lam)
    z1 = torch.zeros((SH.p, SH.h), dtype=F64, device=SH.device)
    ns = SimpleNamespace(part=Bf["part"], E=Bf["E"])
    _lib.call("nl_dot_part", lay.n, ops._p(SH.V), ops._p(SH.V), ops._p(Bf["part"]),
              ops._stream())
    _lib.call("nl_finalize", SH.nparts, ops._p(Bf["part"]), 0, None, None, ops._p(Bf["E"]),
              ops._stream())
    gnorm = math.sqrt(float(Bf["E"].item()))
    dstate = _Diis("pulsed", max_diis, z1, lam, False) if max_diis > 0 else None
    t0 = time.perf_counter()
    res, steps = math.inf, 0
    for steps in range(1, maxSteps + 1):
        residual_vjp(SH, t2, lam, Bf, Bf["JT"])
        _lib.call("nl_sector_assemble", lay.n, SH.p, ops._p(SH.orb), None, ops._p(SH.V),
                  ops._p(Bf["JT"]), None, ops._p(Bf["X"]), ops._p(lam), ctypes.c_double(mixing),
                  ctypes.c_double(0.0), ops._p(lamn), ops._p(Bf["R"]), ops._p(Bf["part"]),
                  ops._stream())
        _lib.call("nl_finalize", SH.nparts, ops._p(Bf["part"][SH.nparts:]), 0, None, None,
                  ops._p(Bf["E"]), ops._stream())
        res = math.sqrt(float(Bf["E"].item())) / max(gnorm, 1e-300)
        if verbose:
            print(f"lambda step {steps}: |residual|/|v| = {res:.3e}")
        if res <= eps or not math.isfinite(res):
            break
        lam, lamn = lamn, lam
        if dstate is not None:
            ext = dstate.step(z1, lam, z1, lamn, ns)
            if ext is not None:
                lam.copy_(ext[1])
    if info is not None:
        info.update(lambda_steps=steps, lambda_residual=res,
                    lambda_seconds=time.perf_counter() - t0)
    if not math.isfinite(res):
        raise _lib.NLError("Lambda iteration diverged")
    return lam


def coupling_gradient(SH, t2, derivatives, info=None, **lambda_kw):
    """dE/dtheta = dE_ref/dtheta + E_corr(t; h_theta) + lam . H(t; h_theta) for each
    (dE_ref/dtheta, derivative SectorHamiltonian) in ``derivatives``, at converged packed t2."""
    lam = solve_lambda(SH, t2, info=info, **lambda_kw)
    out = []
    for dEref, SD in derivatives:
        Bd = SD.buffers()
        tmp = torch.empty_like(t2)
        H = torch.empty_like(t2)
        SD.update(t2, tmp, Bd, 0.0, H2out=H)
        out.append(dEref + SD.energy(t2, Bd) + _vdot(lam, H, Bd, SD))
    return out


def _vdot(x, y, Bf, SH):
    _lib.call("nl_dot_part", SH.layout.n, ops._p(x), ops._p(y), ops._p(Bf["part"]), ops._stream())
    _lib.call("nl_finalize", SH.nparts, ops._p(Bf["part"]), 0, None, None, ops._p(Bf["E"]),
              ops._stream())
    return float(Bf["E"].item())


def planewave_derivative(SH, pw, name):
    """(dE_ref/dtheta, derivative SectorHamiltonian) of the plane-wave coupling ``name`` ("C0" or
    "D0"): two-body blocks times (dg/dtheta)/g, Fock diagonals shifted by d(shift)/dtheta."""
    dEref, fock, scale = pw.coupling_derivative(name)
    fdiag = np.concatenate([np.diag(fock[0]), -np.diag(fock[2])])
    return dEref, SH.scaled(scale, fdiag)

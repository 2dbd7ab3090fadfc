"""Normal-ordered lattice Hamiltonian in particle/hole blocks: the CCM input step
(NuLattice ``CCM/coupled_cluster.py:21-317, 690-832``, ``CCM/three_body_utils.py:6-377`` and the
``lattice.py`` operator lists they consume), vectorised over the operator lists.

Same function names, arguments and semantics as the reference, including its quirks:
* ``get_fock_matrices`` adds the two-body trace only where the kinetic list has an entry
  (coupled_cluster.py:61-67); duplicate kinetic entries are summed (:41-47);
* dense two-body blocks are assigned, not accumulated, per antisymmetric image (:151-260);
* in ``get_3NF`` a (p,h,h) ket with an (h,h,h) bra lands in w_pph_hhh, so w_phh_hhh stays
  empty (three_body_utils.py:112-116).
Sparse blocks (v_pppp, v_ppph, three-body lists) are ``CooList`` objects: they iterate, index,
``len()`` and ``tolist()`` like the reference's ``[[i0, ..., val], ...]`` lists, with entries in
the reference's order, and carry ``indices`` / ``values`` arrays that the solver packs directly.
"""
from itertools import combinations

import numpy as np


class CooList:
    """Reference-format sparse list backed by arrays: row n is [*indices[n], values[n]]."""

    def __init__(self, indices, values):
        self.indices = np.asarray(indices, dtype=np.int64)
        self.values = np.asarray(values, dtype=np.float64)

    def __len__(self):
        return self.values.size

    def __getitem__(self, n):
        return [int(x) for x in self.indices[n]] + [float(self.values[n])]

    def __iter__(self):
        for n in range(len(self)):
            yield self[n]

    def tolist(self):
        return [list(map(int, i)) + [float(v)] for i, v in zip(self.indices, self.values)]

    def __iadd__(self, other):
        idx, val = _split(other, self.indices.shape[1] if self.indices.ndim == 2 else 4)
        self.indices = np.concatenate([self.indices.reshape(-1, idx.shape[1]), idx])
        self.values = np.concatenate([self.values, val])
        return self


def _split(lst, nidx):
    """(indices (N, nidx) int64, values (N,)) of a reference list, an (N, nidx+1) array, a
    CooList or an (indices, values) pair."""
    if isinstance(lst, CooList):
        return lst.indices.reshape(-1, nidx), lst.values
    if isinstance(lst, tuple) and len(lst) == 2:
        return np.asarray(lst[0], np.int64).reshape(-1, nidx), np.asarray(lst[1], np.float64)
    arr = np.asarray(lst, dtype=np.float64).reshape(-1, nidx + 1)
    return arr[:, :nidx].astype(np.int64), arr[:, nidx].copy()


# ------------------------------------------------------------------ lattice operator lists
def state2index(state, L, spin=2, isospin=2):
    """((i*L + j)*L + k)*isospin*spin + tz*spin + sz (lattice.py:119-151)."""
    i, j, k, tz, sz = state
    return (((i * L + j) * L + k) * isospin + tz) * spin + sz


def states2PHSpace(holeList, L):
    """(hole indices in the given order, particle indices ascending) (lattice.py:937-954)."""
    holes = tuple(int(state2index(s, L)) for s in holeList)
    mask = np.ones(4 * L ** 3, bool)
    mask[list(holes)] = False
    return holes, tuple(int(x) for x in np.nonzero(mask)[0])


def Tkin(L, spin=2, isospin=2):
    """Negative dimensionless lattice Laplacian as rows [p, q, value] (lattice.py:306-418): 6
    on the diagonal, -1 to the right neighbour in each direction and its transpose (on L=2 a
    pair appears twice; consumers sum duplicates)."""
    nloc = spin * isospin
    n = L ** 3 * nloc
    rows = [np.stack([np.arange(n), np.arange(n), np.full(n, 6)], 1)]
    site = np.arange(L ** 3)
    xyz = np.stack(np.unravel_index(site, (L, L, L)), 1)
    for d in range(3):
        nb = xyz.copy()
        nb[:, d] = (nb[:, d] + 1) % L
        q_site = np.ravel_multi_index(nb.T, (L, L, L))
        p = (site[:, None] * nloc + np.arange(nloc)).reshape(-1)
        q = (q_site[:, None] * nloc + np.arange(nloc)).reshape(-1)
        hop = np.empty((2 * p.size, 3), np.int64)
        hop[0::2] = np.stack([p, q, -np.ones_like(p)], 1)
        hop[1::2] = np.stack([q, p, -np.ones_like(p)], 1)
        rows.append(hop)
    return np.concatenate(rows)


def _contact_channels(vT1, vS1, spin=2, isospin=2):
    nloc = spin * isospin
    out = []
    for p, q in combinations(range(nloc), 2):
        tzp, szp = divmod(p, spin)
        tzq, szq = divmod(q, spin)
        for r, s in combinations(range(nloc), 2):
            tzr, szr = divmod(r, spin)
            tzs, szs = divmod(s, spin)
            if tzp + tzq != tzr + tzs or szp + szq != szr + szs:
                continue
            if tzp == tzq:
                val = vT1
            elif szp == szq:
                val = vS1
            elif p == r:
                val = 0.5 * (vS1 + vT1)
            else:
                val = 0.5 * (vS1 - vT1)
            out.append((p, q, r, s, val))
    return out


def contacts(vT1, vS1, L, spin=2, isospin=2):
    """On-site antisymmetrised two-body contact <pq||rs>, p<q, r<s, as CooList rows
    (lattice.py:555-627): same isospin projection vT1, same spin projection vS1, the mixed
    channel (vS1 +- vT1)/2; total tz and sz conserved."""
    nloc = spin * isospin
    ch = _contact_channels(vT1, vS1, spin, isospin)
    loc = np.array([c[:4] for c in ch], np.int64)
    val = np.array([c[4] for c in ch], np.float64)
    off = (np.arange(L ** 3) * nloc)[:, None, None]
    idx = (off + loc[None]).reshape(-1, 4)
    return CooList(idx, np.tile(val, L ** 3))


def NNNcontact(v3NF, L, spin=2, isospin=2):
    """On-site three-body contact <pqr||pqr> = v3NF, p<q<r on one site (lattice.py:678-752)."""
    nloc = spin * isospin
    trip = np.array(list(combinations(range(nloc), 3)), np.int64)
    off = (np.arange(L ** 3) * nloc)[:, None, None]
    t = (off + trip[None]).reshape(-1, 3)
    return CooList(np.concatenate([t, t], 1), np.full(t.shape[0], float(v3NF)))


# ------------------------------------------------------------------ two-body blocks
_PERM4 = (((0, 1, 2, 3), 1.0), ((1, 0, 2, 3), -1.0), ((0, 1, 3, 2), -1.0), ((1, 0, 3, 2), 1.0))


def _ph_lookup(part, hole):
    n = max(max(part, default=-1), max(hole, default=-1)) + 1
    is_h = np.zeros(n, bool)
    loc = np.zeros(n, np.int64)
    is_h[list(hole)] = True
    loc[list(part)] = np.arange(len(part))
    loc[list(hole)] = np.arange(len(hole))
    return is_h, loc


def _pair(x, y, is_h, loc):
    """(kind 0=pp 1=ph 2=hh 3=unused, first, second, sign): particle before hole
    (coupled_cluster.py:151-218)."""
    hx, hy = is_h[x], is_h[y]
    swap = hx & ~hy
    first = np.where(swap, loc[y], loc[x])
    second = np.where(swap, loc[x], loc[y])
    kind = hx.astype(np.int64) + hy.astype(np.int64)
    return kind, first, second, np.where(swap, -1.0, 1.0)


def get_all_interactions(part, hole, mycontact, sparse=False):
    """Sort a p<q, r<s two-body list into [v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh]
    (coupled_cluster.py:99-263).  Blocks with a hole bra and particle ket are not stored (the
    interaction is real symmetric)."""
    pnum, hnum = len(part), len(hole)
    idx, val = _split(mycontact, 4)
    is_h, loc = _ph_lookup(part, hole)
    kk, a, b, sk = _pair(idx[:, 0], idx[:, 1], is_h, loc)
    kb, c, d, sb = _pair(idx[:, 2], idx[:, 3], is_h, loc)
    base = np.stack([a, b, c, d], 1)
    x = val * sk * sb
    shapes = {"pppp": (pnum,) * 4, "ppph": (pnum, pnum, pnum, hnum), "pphh": (pnum, pnum, hnum, hnum),
              "phph": (pnum, hnum, pnum, hnum), "phhh": (pnum, hnum, hnum, hnum),
              "hhhh": (hnum,) * 4}
    rules = (("pppp", 0, 0, _PERM4), ("ppph", 0, 1, _PERM4[:2]), ("pphh", 0, 2, _PERM4),
             ("phph", 1, 1, _PERM4[:1]), ("phhh", 1, 2, (_PERM4[0], _PERM4[2])),
             ("hhhh", 2, 2, _PERM4))
    out = {}
    for name, k1, k2, perms in rules:
        m = (kk == k1) & (kb == k2)
        bi, xv = base[m], x[m]
        pidx = np.stack([bi[:, list(pm)] for pm, _ in perms], 1).reshape(-1, 4)
        pval = (xv[:, None] * np.array([sg for _, sg in perms])[None]).reshape(-1)
        if sparse and name in ("pppp", "ppph"):
            out[name] = CooList(pidx, pval)
        else:
            blk = np.zeros(shapes[name])
            blk[tuple(pidx.T)] = pval
            out[name] = blk
    return tuple(out[n] for n in ("pppp", "ppph", "pphh", "phph", "phhh", "hhhh"))


def get_fock_matrices(part, hole, myTkin, v_phph, v_phhh, v_hhhh):
    """[f_pp, f_ph, f_hh] with f_pq = T_pq + sum_i <p i||q i> wherever the kinetic list has an
    entry for (p, q) (coupled_cluster.py:21-97)."""
    pnum, hnum = len(part), len(hole)
    tidx, tval = _split(myTkin, 2)
    is_h, loc = _ph_lookup(part, hole)
    out = []
    for rows_h, cols_h, v in ((False, False, v_phph), (False, True, v_phhh), (True, True, v_hhhh)):
        nr, nc = (hnum if rows_h else pnum), (hnum if cols_h else pnum)
        m = (is_h[tidx[:, 0]] == rows_h) & (is_h[tidx[:, 1]] == cols_h)
        T = np.zeros((nr, nc))
        has = np.zeros((nr, nc), bool)
        r, c = loc[tidx[m, 0]], loc[tidx[m, 1]]
        np.add.at(T, (r, c), tval[m])
        has[r, c] = True
        tr = np.einsum("aibi->ab", np.asarray(v)) if hnum else np.zeros((nr, nc))
        out.append(np.where(has, T + tr, 0.0))
    return out[0], out[1], out[2]


def get_ref_energy(no_1b_hh, no_2b_hhhh, w_hhh_hhh=None):
    """E_ref = sum_i f_ii - 1/2 sum_ij v_ijij + sum over diagonal w_hhh_hhh / 6
    (coupled_cluster.py:288-317)."""
    f = np.asarray(no_1b_hh)
    v = np.asarray(no_2b_hhhh)
    en = float(np.trace(f) - 0.5 * np.einsum("ijij->", v)) if f.size else 0.0
    if w_hhh_hhh is not None and len(w_hhh_hhh):
        idx, val = _split(w_hhh_hhh, 6)
        diag = (idx[:, :3] == idx[:, 3:]).all(1)
        en += float(val[diag].sum()) / 6.0
    return en


# ------------------------------------------------------------------ three-body normal ordering
_P3 = (((0, 1, 2), 1.0), ((1, 0, 2), -1.0), ((2, 1, 0), -1.0), ((0, 2, 1), -1.0),
       ((1, 2, 0), 1.0), ((2, 0, 1), 1.0))
_P2_01 = (((0, 1, 2), 1.0), ((1, 0, 2), -1.0))
_P2_12 = (((0, 1, 2), 1.0), ((0, 2, 1), -1.0))
_W_NAMES = ("ppp_pph", "ppp_phh", "pph_pph", "ppp_hhh", "pph_phh", "pph_hhh", "phh_phh",
            "phh_hhh", "hhh_hhh")
# (ket kind, bra kind) -> (block, ket permutations, bra permutations); kinds count holes
_W_TABLE = {(0, 1): ("ppp_pph", _P3, _P2_01), (0, 2): ("ppp_phh", _P3, _P2_12),
            (0, 3): ("ppp_hhh", _P3, _P3), (1, 1): ("pph_pph", _P2_01, _P2_01),
            (1, 2): ("pph_phh", _P2_01, _P2_12), (1, 3): ("pph_hhh", _P2_01, _P3),
            (2, 2): ("phh_phh", _P2_12, _P2_12), (2, 3): ("pph_hhh", _P2_12, _P3),
            (3, 3): ("hhh_hhh", _P3, _P3)}


def _order3(i1, i2, i3, is_h, loc):
    """A 3-index state with particles first (three_body_utils.py:158-226): (holes, sign,
    local indices (N,3)).  p-h-p and h-p-p become p-p-h, h-p-h and h-h-p become p-h-h, each with
    sign -1 and the pair order of the reference."""
    h1, h2, h3 = is_h[i1], is_h[i2], is_h[i3]
    l1, l2, l3 = loc[i1], loc[i2], loc[i3]
    nh = h1.astype(np.int64) + h2 + h3
    out = np.stack([l1, l2, l3], 1)
    sign = np.ones(i1.size)
    for (m1, m2, m3), order in (((0, 1, 0), (0, 2, 1)), ((1, 0, 0), (2, 1, 0)),
                                ((1, 0, 1), (1, 0, 2)), ((1, 1, 0), (2, 1, 0))):
        m = (h1 == m1) & (h2 == m2) & (h3 == m3)
        out[m] = np.stack([l1, l2, l3], 1)[m][:, list(order)]
        sign[m] = -1.0
    return nh, sign, out


def get_3NF(part, hole, my3body):
    """Sort a three-body list into the nine p/h blocks with all antisymmetric images
    (three_body_utils.py:6-156); returns nine CooLists [w_ppp_pph, w_ppp_phh, w_pph_pph,
    w_ppp_hhh, w_pph_phh, w_pph_hhh, w_phh_phh, w_phh_hhh, w_hhh_hhh]."""
    idx, val = _split(my3body, 6)
    is_h, loc = _ph_lookup(part, hole)
    nk, sk, ki = _order3(idx[:, 0], idx[:, 1], idx[:, 2], is_h, loc)
    nb, sb, bi = _order3(idx[:, 3], idx[:, 4], idx[:, 5], is_h, loc)
    # per entry, the block it feeds, in input order (several (ket, bra) kinds share a block)
    pieces = {n: [] for n in _W_NAMES}
    code = nk * 4 + nb
    for (k1, k2), (name, kp, bp) in _W_TABLE.items():
        m = np.nonzero(code == k1 * 4 + k2)[0]
        if m.size == 0:
            continue
        rows = []
        for pk, s1 in kp:
            for pb, s2 in bp:
                rows.append((np.concatenate([ki[m][:, list(pk)], bi[m][:, list(pb)]], 1),
                             val[m] * sk[m] * sb[m] * s1 * s2))
        I = np.stack([r[0] for r in rows], 1).reshape(-1, 6)
        V = np.stack([r[1] for r in rows], 1).reshape(-1)
        order = np.repeat(m, len(rows))
        pieces[name].append((order, I, V))
    out = []
    for n in _W_NAMES:
        if not pieces[n]:
            out.append(CooList(np.zeros((0, 6), np.int64), np.zeros(0)))
            continue
        order = np.concatenate([p[0] for p in pieces[n]])
        I = np.concatenate([p[1] for p in pieces[n]])
        V = np.concatenate([p[2] for p in pieces[n]])
        o = np.argsort(order, kind="stable")        # entry order of the input list
        out.append(CooList(I[o], V[o]))
    return tuple(out)


def get_3NF_Eref(w_hhh_hhh):
    """sum over diagonal w_hhh_hhh entries / 6 (three_body_utils.py:229-246)."""
    return get_ref_energy(np.zeros((0, 0)), np.zeros((0, 0, 0, 0)), w_hhh_hhh)


def get_3NF_fock(hnum, pnum, w_phh_phh, w_phh_hhh, w_hhh_hhh):
    """f_pq += 1/2 sum_ij <p i j||q i j> (three_body_utils.py:248-284)."""
    out = []
    for shape, lst in (((pnum, pnum), w_phh_phh), ((pnum, hnum), w_phh_hhh),
                       ((hnum, hnum), w_hhh_hhh)):
        f = np.zeros(shape)
        idx, val = _split(lst, 6)
        m = (idx[:, 1] == idx[:, 4]) & (idx[:, 2] == idx[:, 5])
        np.add.at(f, (idx[m, 0], idx[m, 3]), 0.5 * val[m])
        out.append(f)
    return tuple(out)


def get_3NF_tbme(w_pph_pph, w_pph_phh, w_pph_hhh, w_phh_phh, w_phh_hhh, w_hhh_hhh, pnum, hnum,
                 sparse=True):
    """NO2B two-body parts v_pqrs += sum_i <p q i||r s i> (three_body_utils.py:287-377):
    [v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh]."""
    def traced(lst, keep, ci, cj):
        idx, val = _split(lst, 6)
        m = idx[:, ci] == idx[:, cj]
        return idx[m][:, list(keep)], val[m]
    out = []
    i4, v4 = traced(w_pph_pph, (0, 1, 3, 4), 2, 5)
    i3, v3 = traced(w_pph_phh, (0, 1, 3, 4), 2, 5)
    if sparse:
        out += [CooList(i4, v4), CooList(i3, v3)]
    else:
        for shape, (I, V) in (((pnum,) * 4, (i4, v4)), ((pnum, pnum, pnum, hnum), (i3, v3))):
            blk = np.zeros(shape)
            np.add.at(blk, tuple(I.T), V)
            out.append(blk)
    for shape, lst in (((pnum, pnum, hnum, hnum), w_pph_hhh), ((pnum, hnum, pnum, hnum), w_phh_phh),
                       ((pnum, hnum, hnum, hnum), w_phh_hhh), ((hnum,) * 4, w_hhh_hhh)):
        I, V = traced(lst, (0, 1, 3, 4), 2, 5)
        blk = np.zeros(shape)
        np.add.at(blk, tuple(I.T), V)
        out.append(blk)
    return tuple(out)


# ------------------------------------------------------------------ assembled Hamiltonian
def get_norm_ordered_ham(thisL, holes, myTkin, mycontact, my3body=None, sparse=True, NO2B=True):
    """(E_ref, [f_pp, f_ph, f_hh], [v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh]) of the
    lattice Hamiltonian normal-ordered at the reference state ``holes``, with the three-body
    list in the NO2B approximation (coupled_cluster.py:752-832).  With my3body and
    NO2B=False the nine three-body blocks are returned too: ((E_ref, fock, two_body), w)."""
    hole, part = states2PHSpace(holes, thisL)
    hnum, pnum = len(hole), len(part)
    two = list(get_all_interactions(part, hole, mycontact, sparse))
    fock = list(get_fock_matrices(part, hole, myTkin, two[3], two[4], two[5]))
    if my3body is None:
        return get_ref_energy(fock[2], two[5], None), fock, two
    w = get_3NF(part, hole, my3body)
    for n, x in enumerate(get_3NF_fock(hnum, pnum, w[6], w[7], w[8])):
        fock[n] = fock[n] + x
    dum = get_3NF_tbme(w[2], w[4], w[5], w[6], w[7], w[8], pnum, hnum, sparse)
    for n in range(6):
        if sparse and n < 2:
            two[n] += dum[n]
        else:
            two[n] = two[n] + dum[n]
    res = get_ref_energy(fock[2], two[5], w[8]), fock, two
    return res if NO2B else (res, list(w))


def get_norm_ord_int(thisL, holes, vT1, vS1, str_3NF=0, sparse=True):
    """Lattice Hamiltonian with on-site contacts (coupled_cluster.py:690-750): kinetic term,
    two-body contact (vT1, vS1) and, if str_3NF != 0, the three-body contact in NO2B."""
    my3 = NNNcontact(str_3NF, thisL) if str_3NF != 0 else None
    return get_norm_ordered_ham(thisL, holes, Tkin(thisL), contacts(vT1, vS1, thisL), my3,
                                sparse)

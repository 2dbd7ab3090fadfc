"""Normal-ordered Hamiltonian in particle/hole blocks (oracle; test infrastructure only).

Restates NuLattice ``src/NuLattice/CCM/coupled_cluster.py:21-317,690-832`` and
``src/NuLattice/CCM/three_body_utils.py:6-377``.  Block conventions (coupled_cluster.py:151-260):

    v_pppp[a,b,c,d] = <ab||cd>   (COO list [a,b,c,d,val] with all 4 antisymmetric perms when sparse)
    v_ppph[a,b,c,i] = <ab||ci>   (COO list with the 2 perms (a,b),(b,a) when sparse)
    v_pphh[a,b,i,j], v_phph[a,i,b,j], v_phhh[a,i,j,k], v_hhhh[i,j,k,l] dense.

The blocks hp/pp and hh/pp are not stored (the contact is real symmetric).
"""
import numpy as np

from . import lattice_ref as lat


def _blocks_of(ket):
    return tuple("h" if x else "p" for x in ket)


def get_all_interactions(part, hole, mycontact, sparse=False):
    """Sort a p<q, r<s contact list into the six blocks (coupled_cluster.py:99-263)."""
    pnum, hnum = len(part), len(hole)
    lp = {s: i for i, s in enumerate(part)}
    lh = {s: i for i, s in enumerate(hole)}
    hs = set(hole)
    if sparse:
        v_pppp, v_ppph = [], []
    else:
        v_pppp = np.zeros((pnum, pnum, pnum, pnum))
        v_ppph = np.zeros((pnum, pnum, pnum, hnum))
    v_pphh = np.zeros((pnum, pnum, hnum, hnum))
    v_phph = np.zeros((pnum, hnum, pnum, hnum))
    v_phhh = np.zeros((pnum, hnum, hnum, hnum))
    v_hhhh = np.zeros((hnum, hnum, hnum, hnum))

    def norm_pair(x, y):
        # returns (kind, first, second, sign) with a particle always before a hole
        kx, ky = (x in hs), (y in hs)
        if not kx and not ky:
            return "pp", lp[x], lp[y], 1.0
        if not kx and ky:
            return "ph", lp[x], lh[y], 1.0
        if kx and ky:
            return "hh", lh[x], lh[y], 1.0
        return "ph", lp[y], lh[x], -1.0          # (h,p) -> -(p,h)

    perms4 = (((0, 1, 2, 3), 1.0), ((1, 0, 2, 3), -1.0), ((0, 1, 3, 2), -1.0), ((1, 0, 3, 2), 1.0))
    for i1, i2, i3, i4, val in mycontact:
        kk, a, b, sk = norm_pair(i1, i2)
        kb, c, d, sb = norm_pair(i3, i4)
        base = (a, b, c, d)
        target, perms, sparse_block = None, None, False
        if kk == "pp":
            if kb == "pp":
                target, perms, sparse_block = v_pppp, perms4, sparse
            elif kb == "ph":
                target, perms, sparse_block = v_ppph, perms4[:2], sparse
            else:
                target, perms = v_pphh, perms4
        elif kk == "ph":
            if kb == "ph":
                target, perms = v_phph, perms4[:1]
            elif kb == "hh":
                target, perms = v_phhh, (perms4[0], perms4[2])
        elif kk == "hh" and kb == "hh":
            target, perms = v_hhhh, perms4
        if target is None:
            continue
        for perm, sg in perms:
            idx = tuple(base[k] for k in perm)
            x = val * sk * sb * sg
            if sparse_block:
                target.append([idx[0], idx[1], idx[2], idx[3], x])
            else:
                target[idx] = x
    return v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh


def get_fock_matrices(part, hole, myTkin, v_phph, v_phhh, v_hhhh):
    """f = T + sum_i <p i||q i> per block (coupled_cluster.py:21-97).

    Faithful quirk: the two-body sum is only added where the kinetic list has an entry
    for the pair (coupled_cluster.py:61-67).
    """
    t = {}
    for p, q, val in myTkin:
        t[(p, q)] = t.get((p, q), 0) + val
    pnum, hnum = len(part), len(hole)
    f_pp = np.zeros((pnum, pnum))
    f_ph = np.zeros((pnum, hnum))
    f_hh = np.zeros((hnum, hnum))
    for blk, rows, cols, v in ((f_pp, part, part, v_phph), (f_ph, part, hole, v_phhh),
                               (f_hh, hole, hole, v_hhhh)):
        for a, ka in enumerate(rows):
            for b, kb in enumerate(cols):
                val = t.get((ka, kb))
                if val is None:
                    continue
                blk[a, b] = val + np.sum([v[a, i, b, i] for i in range(hnum)])
    return f_pp, f_ph, f_hh


def get_ref_energy(no_1b_hh, no_2b_hhhh, w_hhh_hhh=None):
    """E_ref = sum_i f_ii - 1/2 sum_ij v_ijij + sum_diag w/6 (coupled_cluster.py:288-317)."""
    hnum = len(no_1b_hh)
    en = 0.0
    for i in range(hnum):
        en += no_1b_hh[i, i]
        for j in range(hnum):
            en -= 0.5 * no_2b_hhhh[i, j, i, j]
    if w_hhh_hhh is not None:
        for m, i, j, n, k, l, val in w_hhh_hhh:
            if (m, i, j) == (n, k, l):
                en += val / 6.0
    return en


# ---------------------------------------------------------------- three-body normal ordering
_P3 = (((0, 1, 2), 1), ((1, 0, 2), -1), ((2, 1, 0), -1), ((0, 2, 1), -1), ((1, 2, 0), 1),
       ((2, 0, 1), 1))
_P2_01 = (((0, 1, 2), 1), ((1, 0, 2), -1))
_P2_12 = (((0, 1, 2), 1), ((0, 2, 1), -1))


def _order_state(kind, lp, lh, i1, i2, i3):
    """Bring a 3-index state to p-before-h order (three_body_utils.py:158-226)."""
    k = "".join(kind)
    if k in ("ppp", "hhh", "pph", "phh"):
        lk = [lp if c == "p" else lh for c in k]
        return k, 1.0, [lk[0][i1], lk[1][i2], lk[2][i3]]
    if k == "php":
        return "pph", -1.0, [lp[i1], lp[i3], lh[i2]]
    if k == "hpp":
        return "pph", -1.0, [lp[i3], lp[i2], lh[i1]]
    if k == "hph":
        return "phh", -1.0, [lp[i2], lh[i1], lh[i3]]
    if k == "hhp":
        return "phh", -1.0, [lp[i3], lh[i2], lh[i1]]
    raise ValueError(k)


def get_3NF(part, hole, my3body):
    """Sort a 3-body list into the nine p/h blocks with all antisymmetric permutations
    (three_body_utils.py:6-156).

    Faithful quirk kept: a (p,h,h) ket with an (h,h,h) bra is appended to w_pph_hhh, so
    w_phh_hhh stays empty (three_body_utils.py:112-116).
    """
    lp = {s: i for i, s in enumerate(part)}
    lh = {s: i for i, s in enumerate(hole)}
    hs = set(hole)
    names = ("ppp_pph", "ppp_phh", "pph_pph", "ppp_hhh", "pph_phh", "pph_hhh", "phh_phh",
             "phh_hhh", "hhh_hhh")
    w = {n: [] for n in names}
    table = {
        ("ppp", "pph"): ("ppp_pph", _P3, _P2_01), ("ppp", "phh"): ("ppp_phh", _P3, _P2_12),
        ("ppp", "hhh"): ("ppp_hhh", _P3, _P3),
        ("pph", "pph"): ("pph_pph", _P2_01, _P2_01), ("pph", "phh"): ("pph_phh", _P2_01, _P2_12),
        ("pph", "hhh"): ("pph_hhh", _P2_01, _P3),
        ("phh", "phh"): ("phh_phh", _P2_12, _P2_12), ("phh", "hhh"): ("pph_hhh", _P2_12, _P3),
        ("hhh", "hhh"): ("hhh_hhh", _P3, _P3),
    }
    for i1, i2, i3, i4, i5, i6, val in my3body:
        kk = tuple("h" if x in hs else "p" for x in (i1, i2, i3))
        kb = tuple("h" if x in hs else "p" for x in (i4, i5, i6))
        ket, sk, ki = _order_state(kk, lp, lh, i1, i2, i3)
        bra, sb, bi = _order_state(kb, lp, lh, i4, i5, i6)
        ent = table.get((ket, bra))
        if ent is None:
            continue
        name, kp, bp = ent
        for pk, s1 in kp:
            for pb, s2 in bp:
                w[name].append([ki[pk[0]], ki[pk[1]], ki[pk[2]], bi[pb[0]], bi[pb[1]], bi[pb[2]],
                                val * sk * sb * s1 * s2])
    return tuple(w[n] for n in names)


def get_3NF_fock(hnum, pnum, w_phh_phh, w_phh_hhh, w_hhh_hhh):
    """f_pq += 1/2 sum_ij <p i j||q i j> (three_body_utils.py:248-284)."""
    f_pp = np.zeros((pnum, pnum))
    f_ph = np.zeros((pnum, hnum))
    f_hh = np.zeros((hnum, hnum))
    for blk, lst in ((f_pp, w_phh_phh), (f_ph, w_phh_hhh), (f_hh, w_hhh_hhh)):
        for a, i, j, b, k, l, val in lst:
            if i == k and j == l:
                blk[a, b] += 0.5 * val
    return f_pp, f_ph, f_hh


def get_3NF_tbme(w_pph_pph, w_pph_phh, w_pph_hhh, w_phh_phh, w_phh_hhh, w_hhh_hhh,
                 pnum, hnum, sparse=True):
    """v_pqrs += sum_i <p q i||r s i> (three_body_utils.py:287-377)."""
    v_phph = np.zeros((pnum, hnum, pnum, hnum))
    v_pphh = np.zeros((pnum, pnum, hnum, hnum))
    v_phhh = np.zeros((pnum, hnum, hnum, hnum))
    v_hhhh = np.zeros((hnum, hnum, hnum, hnum))
    if sparse:
        v_pppp, v_ppph = [], []
    else:
        v_pppp = np.zeros((pnum,) * 4)
        v_ppph = np.zeros((pnum, pnum, pnum, hnum))
    for a, b, i, c, d, j, val in w_pph_pph:
        if i == j:
            if sparse:
                v_pppp.append([a, b, c, d, val])
            else:
                v_pppp[a, b, c, d] += val
    for a, b, i, c, k, j, val in w_pph_phh:
        if i == j:
            if sparse:
                v_ppph.append([a, b, c, k, val])
            else:
                v_ppph[a, b, c, k] += val
    for a, b, i, m, k, j, val in w_pph_hhh:
        if i == j:
            v_pphh[a, b, m, k] += val
    for a, n, i, b, k, j, val in w_phh_phh:
        if i == j:
            v_phph[a, n, b, k] += val
    for a, n, i, l, k, j, val in w_phh_hhh:
        if i == j:
            v_phhh[a, n, l, k] += val
    for m, n, i, l, k, j, val in w_hhh_hhh:
        if i == j:
            v_hhhh[m, n, l, k] += val
    return v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh


def _add_no2b(two_body, dum, sparse):
    for n in range(6):
        if sparse and n < 2:
            two_body[n].extend(list(dum[n]))
        else:
            two_body[n] += dum[n]


def get_norm_ordered_ham(thisL, holes, myTkin, mycontact, my3body=None, sparse=True):
    """Normal-ordered Hamiltonian from operator lists (coupled_cluster.py:752-832), NO2B form.

    Returns (E_ref, [f_pp, f_ph, f_hh], [v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh]).
    """
    hole, part = lat.states2PHSpace(holes, thisL)
    hnum, pnum = len(hole), len(part)
    two_body = list(get_all_interactions(part, hole, mycontact, sparse))
    fock = list(get_fock_matrices(part, hole, myTkin, two_body[3], two_body[4], two_body[5]))
    w_hhh_hhh = None
    if my3body is not None:
        w = get_3NF(part, hole, my3body)
        (w_ppp_pph, w_ppp_phh, w_pph_pph, w_ppp_hhh, w_pph_phh, w_pph_hhh, w_phh_phh,
         w_phh_hhh, w_hhh_hhh) = w
        for n, x in enumerate(get_3NF_fock(hnum, pnum, w_phh_phh, w_phh_hhh, w_hhh_hhh)):
            fock[n] += x
        dum = get_3NF_tbme(w_pph_pph, w_pph_phh, w_pph_hhh, w_phh_phh, w_phh_hhh, w_hhh_hhh,
                           pnum, hnum, sparse)
        _add_no2b(two_body, dum, sparse)
    vac = get_ref_energy(fock[2], two_body[5], w_hhh_hhh)
    return vac, fock, two_body


def get_norm_ord_int(thisL, holes, vT1, vS1, str_3NF=0, sparse=True):
    """Lattice Hamiltonian with contacts (coupled_cluster.py:690-750)."""
    my3 = lat.NNNcontact(str_3NF, thisL) if str_3NF != 0 else None
    return get_norm_ordered_ham(thisL, holes, lat.Tkin(thisL), lat.contacts(vT1, vS1, thisL),
                                my3, sparse)

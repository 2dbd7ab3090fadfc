"""Position-space lattice basis and operator lists (oracle; test infrastructure only).

Restates the parts of NuLattice ``src/NuLattice/lattice.py`` that feed the CCSD path:
orbital indexing (lattice.py:119-151), the kinetic hopping list (lattice.py:306-418),
the on-site two-body contact list (lattice.py:421-450, 555-627), the on-site three-body
contact list (lattice.py:678-752), the hole/particle split (lattice.py:937-954) and the
energy unit (lattice.py:57-61, constants.py:5-8).
"""
from itertools import combinations

import numpy as np

HBARC = 197.3269804                       # MeV fm         (constants.py:5)
MASS = 0.5 * (938.27208943 + 939.56542194)  # MeV          (constants.py:6-8)


def phys_unit(a_lat):
    """Lattice energy unit in MeV, hbar^2/(2 m a^2) (lattice.py:57-61)."""
    return 0.5 * HBARC ** 2 / (MASS * a_lat ** 2)


def state2index(state, L, spin=2, isospin=2):
    """((i*L + j)*L + k)*isospin*spin + tz*spin + sz (lattice.py:119-151)."""
    i, j, k, tz, sz = state
    return (((i * L + j) * L + k) * isospin + tz) * spin + sz


def sp_basis(L, spin=2, isospin=2):
    """All [i, j, k, tz, sz] in index order (lattice.py:82-98)."""
    return [[i, j, k, tz, sz] for i in range(L) for j in range(L) for k in range(L)
            for tz in range(isospin) for sz in range(spin)]


def Tkin(L, spin=2, isospin=2):
    """Negative dimensionless Laplacian as [p, q, value] (lattice.py:306-418).

    Diagonal 6 on every orbital, then for each direction a -1 hop to the right neighbour and
    its transpose.  On L=2 the right and left neighbours coincide, so a pair appears twice;
    consumers accumulate duplicates (coupled_cluster.py:41-47).  Values are ints
    (lattice.py:418).
    """
    nloc = spin * isospin
    out = [[p, p, 6] for p in range(L ** 3 * nloc)]
    for d in range(3):
        for i in range(L):
            for j in range(L):
                for k in range(L):
                    site = [i, j, k]
                    nb = list(site)
                    nb[d] = (nb[d] + 1) % L
                    for loc in range(nloc):
                        p = (((i * L + j) * L + k) * nloc) + loc
                        q = (((nb[0] * L + nb[1]) * L + nb[2]) * nloc) + loc
                        out.append([p, q, -1])
                        out.append([q, p, -1])
    return out


def contacts(vT1, vS1, L, spin=2, isospin=2):
    """On-site antisymmetrised two-body contact <pq||rs>, p<q, r<s (lattice.py:555-627).

    Channel values: same isospin projection -> vT1; same spin projection -> vS1; mixed
    (Tz=0, Sz=0): (vS1+vT1)/2 when p==r, (vS1-vT1)/2 otherwise.  Total tz and sz conserved.
    """
    nloc = spin * isospin
    chans = []
    for p in range(nloc):
        for q in range(p + 1, nloc):
            tzp, szp = divmod(p, spin)
            tzq, szq = divmod(q, spin)
            for r in range(nloc):
                for s in range(r + 1, nloc):
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
                    chans.append((p, q, r, s, val))
    out = []
    for site in range(L ** 3):
        off = site * nloc
        for p, q, r, s, val in chans:
            out.append([off + p, off + q, off + r, off + s, val])
    return out


def NNNcontact(v3NF, L, spin=2, isospin=2):
    """On-site three-body contact, diagonal elements <pqr||pqr> = v3NF, p<q<r on one site
    (lattice.py:678-752)."""
    nloc = spin * isospin
    out = []
    for site in range(L ** 3):
        off = site * nloc
        for p, q, r in combinations(range(nloc), 3):
            out.append([off + p, off + q, off + r, off + p, off + q, off + r, float(v3NF)])
    return out


def states2PHSpace(holeList, L):
    """Hole indices in the given order, particle indices ascending (lattice.py:937-954)."""
    holes = [state2index(h, L) for h in holeList]
    hs = set(holes)
    parts = tuple(p for p in range(L ** 3 * 4) if p not in hs)
    return tuple(holes), parts


def _filled_sites(sites):
    # the deprecated references.py lists: for each site, (tz, sz) in order 00, 10, 01, 11
    return [[x, y, z, tz, sz] for (x, y, z) in sites for sz in (0, 1) for tz in (0, 1)]


# Reference states of src/NuLattice/references.py:40-149 (tz, sz order as stored there).
ref_2H_gs = [[0, 0, 0, 0, 0], [0, 0, 0, 1, 0]]
ref_3He_gs = [[0, 0, 0, 0, 0], [0, 0, 0, 1, 0], [0, 0, 0, 0, 1]]
ref_4He_gs = _filled_sites([(0, 0, 0)])
ref_8Be_gs = _filled_sites([(0, 0, 0), (1, 0, 0)])
ref_12C_gs = _filled_sites([(0, 0, 0), (1, 0, 0), (0, 1, 0)])
ref_16O_gs = _filled_sites([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)])

"""Text Hamiltonian files and golden energy fixtures.

``read_text_hamiltonian`` reads the normal-ordered text format of the reference's benchmark data
(``benchmarks/H2_data/*.dat``, parsed by the script ``benchmarks/CCSDBenchmark.py:9-118``): one
element per line, ``val`` (0B: the reference energy), ``p q val`` (1B) and ``p q r s val`` (2B),
single-particle indices from 0.  It returns the blocks the solver takes, dense and with the
sparse (COO list) ``v_pppp`` / ``v_ppph`` of ``sparse=True``, classified and indexed as that
script does:

* the number of states is one more than the largest index in the 1B file;
* particles are the remaining states in increasing order, holes in the order given;
* a 1B element with one hole index fills ``f_ph[a, i]``; 2B elements fill only the six canonical
  orderings pppp, ppph, pphh, phph, phhh, hhhh (the file holds every antisymmetric permutation);
* the COO lists keep the first occurrence of each identical ``[a, b, c, d, val]`` entry.

Parsing is vectorised (one ``numpy.loadtxt`` per file; ``.gz`` files are read transparently), so
the 75k-line deuteron file loads in well under a second instead of the script's quadratic list
membership tests.

``write_energy_fixture`` / ``read_energy_fixture`` keep per-step energy trajectories (the
``history`` a solver run records) as named JSON entries, the form of every known-answer fixture
the tests compare the solvers against.
"""
import json
import os

import numpy as np

from .hamiltonian import CooList

__all__ = ["read_text_hamiltonian", "write_energy_fixture", "read_energy_fixture"]


def _rows(path, width):
    data = np.loadtxt(path, ndmin=2)
    if data.shape[1] != width:
        raise ValueError(f"{path}: expected {width} columns, found {data.shape[1]}")
    return data


def _first_unique_rows(m):
    """Rows of m in order of first occurrence, rows equal as numbers dropped (+ 0.0 maps -0.0
    to 0.0, which Python's list comparison treats as equal)."""
    if len(m) == 0:
        return m
    _, first = np.unique(m + 0.0, axis=0, return_index=True)
    return m[np.sort(first)]


def read_text_hamiltonian(file0, file1, file2, holes):
    """(E_ref, fock_mats, dense_blocks, sparse_blocks) of a normal-ordered text Hamiltonian.

    fock_mats = [f_pp, f_ph, f_hh]; dense_blocks = [v_pppp, v_ppph, v_pphh, v_phph, v_phhh,
    v_hhhh] (all dense arrays); sparse_blocks = the same with v_pppp and v_ppph as COO lists
    ``[[a, b, c, d, val], ...]`` (``CooList``).  ``holes`` is the tuple of hole states, e.g.
    ``(0, 2)`` for the deuteron files.
    """
    holes = tuple(int(x) for x in holes)
    e_ref = float(np.loadtxt(file0))
    one = _rows(file1, 3)
    two = _rows(file2, 5)
    nstat = int(one[:, :2].max()) + 1 if len(one) else 0
    if len(two) and int(two[:, :4].max()) >= nstat:
        raise ValueError("2B file names a state beyond the 1B file's range")
    if any(x >= nstat or x < 0 for x in holes):
        raise ValueError("hole state outside the file's state range")
    part = tuple(s for s in range(nstat) if s not in holes)
    p, h = len(part), len(holes)
    is_hole = np.zeros(nstat, dtype=bool)
    is_hole[list(holes)] = True
    local = np.zeros(nstat, dtype=np.int64)
    local[list(part)] = np.arange(p)
    local[list(holes)] = np.arange(h)

    f_pp, f_ph, f_hh = np.zeros((p, p)), np.zeros((p, h)), np.zeros((h, h))
    idx = one[:, :2].astype(np.int64)
    val = one[:, 2]
    hp, hq = is_hole[idx[:, 0]], is_hole[idx[:, 1]]
    lp, lq = local[idx[:, 0]], local[idx[:, 1]]
    # in file order, so a repeated element keeps its last value as in the script (88 lines)
    for k in range(len(val)):
        if hp[k]:
            if hq[k]:
                f_hh[lp[k], lq[k]] = val[k]
            else:
                f_ph[lq[k], lp[k]] = val[k]
        elif hq[k]:
            f_ph[lp[k], lq[k]] = val[k]
        else:
            f_pp[lp[k], lq[k]] = val[k]

    idx2 = two[:, :4].astype(np.int64)
    v2 = two[:, 4]
    hmask = is_hole[idx2]                      # (n, 4): True where the index is a hole
    loc = local[idx2]
    pattern = {"pppp": (0, 0, 0, 0), "ppph": (0, 0, 0, 1), "pphh": (0, 0, 1, 1),
               "phph": (0, 1, 0, 1), "phhh": (0, 1, 1, 1), "hhhh": (1, 1, 1, 1)}
    dims = {"p": p, "h": h}
    dense = {}
    coo = {}
    for name, pat in pattern.items():
        sel = np.all(hmask == np.array(pat, dtype=bool), axis=1)
        arr = np.zeros(tuple(dims[c] for c in name))
        rows, vals = loc[sel], v2[sel]
        # assignment keeps the last of repeated elements, as the script's loop does
        arr[tuple(rows.T)] = vals
        dense[name] = arr
        if name in ("pppp", "ppph"):
            m = np.concatenate([rows.astype(np.float64), vals[:, None]], axis=1)
            m = _first_unique_rows(m)
            coo[name] = CooList(m[:, :4].astype(np.int64), m[:, 4])
    fock = [f_pp, f_ph, f_hh]
    dense_blocks = [dense[n] for n in ("pppp", "ppph", "pphh", "phph", "phhh", "hhhh")]
    sparse_blocks = [coo["pppp"], coo["ppph"]] + dense_blocks[2:]
    return e_ref, fock, dense_blocks, sparse_blocks


def write_energy_fixture(path, name, energies, **meta):
    """Store ``energies`` (per-step E_corr, step 0 first) under ``name`` in the JSON file at
    ``path`` (created or updated), with any extra JSON-serialisable ``meta`` fields."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    entry = {"energies": [float(e) for e in energies]}
    entry.update(meta)
    data[name] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)
    return entry


def read_energy_fixture(path, name):
    """The entry ``name`` of a fixture file written by ``write_energy_fixture``."""
    with open(path) as f:
        return json.load(f)[name]

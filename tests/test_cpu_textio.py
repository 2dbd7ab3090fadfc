"""Text Hamiltonian loader (no GPU): the deuteron benchmark files through the oracle reproduce
the energies the reference printed for them, and the parsing rules on a small hand-made file."""
import gzip
import os

import numpy as np
import pytest

from nulattice_amd.textio import read_energy_fixture, read_text_hamiltonian, write_energy_fixture
from oracle import ccsd_ref as cc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIXTURE = os.path.join(GOLDEN, "deuteron_ccsd.json")


def deuteron():
    f = [os.path.join(GOLDEN, f"deuteron_HNO_basis1_{k}.dat.gz") for k in ("0B", "1B", "2B")]
    return read_text_hamiltonian(*f, holes=(0, 2))


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) / np.abs(b)


def check_trajectory(hist, ref, final=None):
    """Steps 0-10 are damped Jacobi and agree to round-off.  From step 11 on the pulsed DIIS
    extrapolation amplifies round-off (the reference's own dense and sparse runs of these files
    differ by 1.2e-7 at step 11), so those steps and the converged energy are held to 1e-6."""
    n = min(len(hist), len(ref))
    assert n >= 11
    assert _rel(hist[:11], ref[:11]).max() < 1e-13
    if n > 11:
        assert _rel(hist[11:n], ref[11:n]).max() < 1e-6
    if final is not None:
        assert _rel(final, ref[-1]) < 1e-6


def test_deuteron_blocks():
    e_ref, fock, dense, sparse = deuteron()
    assert e_ref == pytest.approx(0.769126172995, abs=0)
    assert [f.shape for f in fock] == [(38, 38), (38, 2), (2, 2)]
    assert [v.shape for v in dense] == [(38,) * 4, (38, 38, 38, 2), (38, 38, 2, 2),
                                        (38, 2, 38, 2), (38, 2, 2, 2), (2,) * 4]
    # the COO lists hold the same elements as the dense blocks they replace
    for coo, full in zip(sparse[:2], dense[:2]):
        rebuilt = np.zeros_like(full)
        rebuilt[tuple(coo.indices.T)] = coo.values
        assert np.array_equal(rebuilt, full)
        assert len(np.unique(coo.indices, axis=0)) == len(coo)
    # every block is antisymmetric where the file lists both orderings
    v_pppp, v_pphh = dense[0], dense[2]
    assert np.array_equal(v_pppp, -v_pppp.transpose(1, 0, 2, 3))
    assert np.array_equal(v_pphh, -v_pphh.transpose(0, 1, 3, 2))
    assert np.array_equal(fock[0], fock[0].T)


def test_deuteron_dense_trajectory_through_oracle():
    """benchmarks/CCSDBenchmark.ipynb cell 2, dense run (33 energies)."""
    _, fock, dense, _ = deuteron()
    hist = []
    e, _, _ = cc.ccsd_solver(fock, dense, sparse=False, diis="pulsed", history=hist)
    check_trajectory(hist, read_energy_fixture(FIXTURE, "dense")["energies"], e)


def test_deuteron_sparse_trajectory_through_oracle():
    """The same notebook's sparse run, through the first DIIS extrapolation (13 energies: the
    oracle's COO loop is a pure-Python pass over 63k elements per step)."""
    _, fock, _, sparse = deuteron()
    hist = []
    cc.ccsd_solver(fock, sparse, sparse=True, diis="pulsed", history=hist, maxSteps=12)
    check_trajectory(hist, read_energy_fixture(FIXTURE, "sparse")["energies"])


def _write(path, rows, gz=False):
    text = "".join(" ".join(repr(x) for x in r) + "\n" for r in rows)
    if gz:
        with gzip.open(path, "wt") as f:
            f.write(text)
    else:
        with open(path, "w") as f:
            f.write(text)


def test_parsing_rules(tmp_path):
    """Five states with holes (3, 1): particles (0, 2, 4) ascending, holes in the given order."""
    f0, f1, f2 = (str(tmp_path / n) for n in ("0B.dat", "1B.dat.gz", "2B.dat"))
    _write(f0, [[-1.5]])
    _write(f1, [[0, 0, 1.0], [3, 3, -2.0], [3, 1, 0.25], [1, 4, 0.5], [4, 1, 0.75],
                [2, 4, 0.125], [2, 4, 0.375]], gz=True)
    _write(f2, [[0, 2, 4, 0, 1.0], [0, 2, 4, 0, 1.0], [0, 2, 4, 0, 2.0],   # pppp, repeated
                [4, 2, 0, 3, 3.0], [0, 4, 3, 1, 4.0], [2, 1, 4, 3, 5.0],   # ppph pphh phph
                [4, 3, 1, 3, 6.0], [1, 3, 3, 1, 7.0],                      # phhh hhhh
                [1, 0, 2, 4, 9.0], [0, 3, 3, 0, 9.0]])                     # not canonical
    e_ref, (f_pp, f_ph, f_hh), dense, sparse = read_text_hamiltonian(f0, f1, f2, holes=(3, 1))
    assert e_ref == -1.5
    a = {0: 0, 2: 1, 4: 2}
    i = {3: 0, 1: 1}
    assert f_hh[i[3], i[3]] == -2.0 and f_hh[i[3], i[1]] == 0.25
    assert f_ph[a[4], i[1]] == 0.75            # (p, q) = (4, 1): particle row, hole column
    assert f_ph.sum() == 0.75                  # (1, 4) wrote the same element first
    assert f_pp[a[0], a[0]] == 1.0 and f_pp[a[2], a[4]] == 0.375   # the last repeat wins
    v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = dense
    assert v_pppp[a[0], a[2], a[4], a[0]] == 2.0 and np.count_nonzero(v_pppp) == 1
    assert v_ppph[a[4], a[2], a[0], i[3]] == 3.0 and np.count_nonzero(v_ppph) == 1
    assert v_pphh[a[0], a[4], i[3], i[1]] == 4.0 and np.count_nonzero(v_pphh) == 1
    assert v_phph[a[2], i[1], a[4], i[3]] == 5.0 and np.count_nonzero(v_phph) == 1
    assert v_phhh[a[4], i[3], i[1], i[3]] == 6.0 and np.count_nonzero(v_phhh) == 1
    assert v_hhhh[i[1], i[3], i[3], i[1]] == 7.0 and np.count_nonzero(v_hhhh) == 1
    # the COO list keeps each distinct (indices, value) once, in file order
    assert sparse[0].tolist() == [[0, 1, 2, 0, 1.0], [0, 1, 2, 0, 2.0]]
    assert sparse[1].tolist() == [[2, 1, 0, 0, 3.0]]
    assert sparse[2] is dense[2]


def test_parsing_errors(tmp_path):
    f0, f1, f2 = (str(tmp_path / n) for n in ("0B.dat", "1B.dat", "2B.dat"))
    _write(f0, [[0.0]])
    _write(f1, [[0, 0, 1.0], [2, 2, 1.0]])
    _write(f2, [[0, 1, 2, 3, 1.0]])
    with pytest.raises(ValueError, match="beyond"):
        read_text_hamiltonian(f0, f1, f2, holes=(0,))
    _write(f2, [[0, 1, 2, 1.0]])
    with pytest.raises(ValueError, match="columns"):
        read_text_hamiltonian(f0, f1, f2, holes=(0,))
    _write(f2, [[0, 1, 2, 2, 1.0]])
    with pytest.raises(ValueError, match="hole"):
        read_text_hamiltonian(f0, f1, f2, holes=(5,))


def test_energy_fixture_round_trip(tmp_path):
    path = str(tmp_path / "e.json")
    write_energy_fixture(path, "a", [1.0, 0.5], holes=[0, 2])
    write_energy_fixture(path, "b", np.array([-1e-17, 3.0]))
    assert read_energy_fixture(path, "a") == {"energies": [1.0, 0.5], "holes": [0, 2]}
    assert read_energy_fixture(path, "b")["energies"] == [-1e-17, 3.0]

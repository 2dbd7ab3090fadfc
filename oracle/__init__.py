"""CPU FP64 oracle for the NuLattice CCSD amplitude iteration.

TEST INFRASTRUCTURE ONLY.  Nothing in ``nulattice_amd`` imports this package.
It may be imported by ``tests/``, by ``__graft_entry__.smoke()`` (to verify its result)
and by ``bench.py``'s ``cpu_baseline`` leg (timed as the CPU reference), never as
the thing measured or shipped.

Modules
-------
lattice_ref      position-space lattice basis and operator lists
                 (follows NuLattice ``src/NuLattice/lattice.py``)
hamiltonian_ref  sorting of operator lists into particle/hole blocks, Fock matrices,
                 three-body normal ordering, reference energy
                 (follows ``src/NuLattice/CCM/coupled_cluster.py:21-317,690-832`` and
                 ``src/NuLattice/CCM/three_body_utils.py``)
ccsd_ref         the diagram library, T1/T2 iteration, energy and solver
                 (follows ``src/NuLattice/CCM/ccDgrams.py`` and
                 ``src/NuLattice/CCM/coupled_cluster.py:265-687``; the DIIS window step
                 ``diis_window_step`` follows ``coupled_cluster.py:635-672`` and
                 ``src/NuLattice/CCM/soa/coupled_cluster.py:520-567``)

How each function is pinned (tests/test_oracle.py unless another file is named):
  * ccsd_ref (all diagrams, t1Iter, t2Iter, ccsd_energy, ccsd_solver with max_diis=0 and
    pulsed DIIS) and hamiltonian_ref (get_norm_ord_int / get_norm_ordered_ham with the
    three-body contact) are pinned by the per-iteration energies stored in the reference's
    notebooks (examples/Example_CCSD.ipynb cell 5, examples/Example_CCSD2.ipynb cell 3),
    16 digits, steps 0-32, committed as tests/golden/trajectories.json.
  * Every diagram is additionally checked against an independent, unfactorised CCSD residual
    restated from benchmarks/DeuteronBenchmark.py:108-161: a dropped term, wrong sign, index or
    coefficient in either fails it (a planted sign error is shown to fail).
  * ccsd_solver is checked on A=2 against exact diagonalisation (CCSD is exact for two
    particles).
  * diis_window_step (the pulsed and the SoA window) is checked on a linear fixed-point map
    against the closed-form constrained least-squares combination over the kept window.
  * lattice_ref is pinned by the operator counts stored in Example_CCSD2.ipynb
    (1792 kinetic, 512 contact, 256 three-body elements at L=4) and by the Laplacian spectrum.
  * The package's plane-wave generator, which feeds the oracle at the benchmark sizes, is
    checked against the position-space operators by a Fourier transform and against
    hamiltonian_ref's sorting (tests/test_cpu_package.py).
"""

"""MI355X-native CCSD amplitude iteration for NuLattice's CCM path.

``nulattice_amd.ccm`` mirrors ``NuLattice.CCM.coupled_cluster`` (ccsd_solver, t1Iter, t2Iter,
t1Init, t2Init, ccsd_energy, v_ppph_dgrams, v_pppp_dgrams), ``nulattice_amd.ccdgrams`` mirrors
``NuLattice.CCM.ccDgrams`` and ``nulattice_amd.hamiltonian`` the Hamiltonian input step
(get_norm_ord_int, get_norm_ordered_ham, get_all_interactions, get_fock_matrices,
get_ref_energy, the three-body NO2B helpers and the lattice operator lists).  The arithmetic runs
in libnlccsd.so (hand-written HIP for gfx950, C ABI in include/nulattice_ccsd.h); build it with
``python -m nulattice_amd.build``.  ``nulattice_amd.adjoint`` adds the differentiation form
(Lambda equations, dE/dC0 and dE/dD0) and ``nulattice_amd.planewave`` the momentum-basis
Hamiltonian.
"""
from . import _lib  # noqa: F401

__all__ = ["adjoint", "ccdgrams", "ccm", "engine", "hamiltonian", "hfbasis", "ops", "planewave",
           "plans", "sectors", "textio"]

"""MI355X-native CCSD amplitude iteration for NuLattice's CCM path.

``nulattice_amd.ccm`` mirrors ``NuLattice.CCM.coupled_cluster`` (ccsd_solver, t1Iter, t2Iter,
t1Init, t2Init, ccsd_energy); the arithmetic runs in libnlccsd.so (hand-written HIP for gfx950,
C ABI in include/nulattice_ccsd.h).  Build it with ``python -m nulattice_amd.build``.
"""
from . import _lib  # noqa: F401

__all__ = ["ccm", "engine", "ops", "planewave", "plans"]

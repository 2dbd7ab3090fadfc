"""Argument rules of the SoA driver interface (no GPU needed: every case is rejected before a
kernel runs)."""
import pytest
import torch

from nulattice_amd import _lib, soa


def test_soa_driver_rejects_what_it_cannot_run():
    with pytest.raises(ValueError):
        soa.ccsd_solver(None, None, dtype=torch.float16)
    with pytest.raises(_lib.NLError):
        soa.ccsd_solver(None, None, device="cpu")
    with pytest.raises(TypeError):
        soa.ccsd_solver(None, None, device="cuda", diis="pulsed")
    with pytest.raises(TypeError):
        soa.ccsd_solver(None, None, device="cuda", sectors=([], [], []))


def test_soa_module_exposes_the_soa_function_names():
    """soa/coupled_cluster.py:260-437 and :440: the amplitude functions and the driver."""
    for name in ("ccsd_solver", "ccsd_energy", "get_ref_energy", "t1Init", "t2Init", "t1Iter",
                 "t2Iter"):
        assert callable(getattr(soa, name))

"""The allocation-free energy entry point (nl_ccsd_energy_noalloc, static device scratch)
against the workspace entry points it mirrors (nl_energy_part + nl_finalize)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from nulattice_amd import _lib  # noqa: E402


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_noalloc_energy_equals_workspace_energy():
    """Two calls in a row, the second smaller: p=120, h=12 (2.07e6 amplitudes, 4096 partials:
    the static array's full size), then p=24, h=4 (36 partials, over the first call's leftovers).
    Each result equals the workspace entry points' bit for bit and the energy formula
    (coupled_cluster.py:265-286) to 1e-12."""
    lib = _lib.load()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for p, h, seed in ((120, 12, 1), (24, 4, 2)):
        rng = np.random.default_rng(seed)
        f = rng.standard_normal((p, h))
        v = rng.standard_normal((p, p, h, h))
        t2 = rng.standard_normal((p, p, h, h)) * 0.01
        t1 = rng.standard_normal((p, h)) * 0.1
        ref = (np.sum(f * t1) + 0.25 * np.sum(v * t2)
               + 0.5 * np.einsum("abij,ai,bj->", v, t1, t1))
        d = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
        gf, gv, gt2, gt1 = d(f), d(v), d(t2), d(t1)
        n = p * p * h * h
        nparts = int(lib.nl_partials(n))
        part = torch.empty(nparts, dtype=torch.float64, device="cuda")
        e1 = torch.empty(1, dtype=torch.float64, device="cuda")
        e2 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        assert lib.nl_energy_part(p, h, _p(gv), _p(gt2), _p(gt1), _p(part), st) == 0
        assert lib.nl_finalize(nparts, _p(part), p * h, _p(gf), _p(gt1), _p(e1), st) == 0
        assert lib.nl_ccsd_energy_noalloc(p, h, _p(gf), _p(gv), _p(gt2), _p(gt1), _p(e2),
                                          st) == 0
        torch.cuda.synchronize()
        assert float(e2.item()) == float(e1.item())
        assert abs(float(e2.item()) - ref) <= 1e-12 * abs(ref)

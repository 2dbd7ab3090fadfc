"""The SoA CCSD driver's interface on the MI355X path.

NuLattice has two CCSD drivers with the same arguments: the legacy one
(src/NuLattice/CCM/coupled_cluster.py:547-687, mirrored by ``ccm.ccsd_solver``) and the SoA one
(src/NuLattice/CCM/soa/coupled_cluster.py:440-575).  The SoA driver runs the same damped Jacobi
iteration with three different rules, which ``ccsd_solver`` below applies:

* it stops when |E - E_prev| / max(1, |E|) < eps (soa/coupled_cluster.py:517), not |dE|/|E|;
* pulsed DIIS is on whenever max_diis > 0 (soa/coupled_cluster.py:520-567): the newest max_diis
  error vectors, B normalised by max|B| + 1e-16, a singular solve skipped;
* on divergence (|E| > 1e10) or NaN it prints "Diverged.", leaves the loop, prints "Max
  iterations reached." and returns (E, t1, t2) of the last step instead of None
  (soa/coupled_cluster.py:569-575); the same message and return end a run at maxSteps.

The amplitude functions have the SoA module's signatures, which are the legacy ones
(soa/coupled_cluster.py:260-437), and are the same GPU implementations as in ``ccm``.
"""
import torch

from . import _lib, ccm
from .ccm import ccsd_energy, t1Init, t1Iter, t2Init, t2Iter, v_ppph_dgrams  # noqa: F401
from .hamiltonian import get_ref_energy  # noqa: F401


def ccsd_solver(fock_mats, two_body_int, t1initial=None, eps=1e-8, maxSteps=1000, max_diis=10,
                delta=0, mixing=0.5, verbose=False, sparse=True, ccs=False, device=None,
                dtype=torch.float64, **options):
    """CCSD with the SoA driver's rules (soa/coupled_cluster.py:440-575).  Returns (E_corr, t1,
    t2), also after a divergence or NaN.  ``device`` must be a HIP device (None: the current
    one).  ``dtype`` is the storage type, as in the reference (soa/coupled_cluster.py:453-460):
    with torch.float32 the DIIS history (the reference's "major memory hog", :521) is kept in
    FP32 and t1, t2 are returned as float32, and the contractions default to the split-operand
    precision form (precision="fp16x3", FP32-class accuracy, no FP64 refinement unless
    refine=True is passed); energy, denominators, the update and the DIIS solve stay FP64.
    Other keyword options are those of ``ccm.ccsd_solver`` (history, info, precision,
    hamiltonian, ...), except ``diis`` and ``sectors``, which the SoA rules fix."""
    if dtype not in (torch.float64, torch.float32):
        raise ValueError("dtype must be torch.float64 or torch.float32")
    if dtype == torch.float32:
        options.setdefault("precision", "fp16x3")
        options.setdefault("refine", False)
    if device is not None and torch.device(device).type != "cuda":
        raise _lib.NLError("soa.ccsd_solver: a HIP device is required (no CPU fallback)")
    for k in ("diis", "sectors"):
        if k in options:
            raise TypeError(f"soa.ccsd_solver fixes {k!r}")
    return ccm.ccsd_solver(fock_mats, two_body_int, t1initial, eps, maxSteps, max_diis, delta,
                           mixing, verbose, sparse, ccs, device=device, diis="soa",
                           storage=dtype, _soa_rules=True, **options)

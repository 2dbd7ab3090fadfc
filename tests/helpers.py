"""Shared test utilities: random physically consistent Hamiltonians and amplitudes."""
import numpy as np


def random_antisym_hamiltonian(p, h, seed=0, scale=0.3):
    """Full antisymmetric, real-symmetric two-body tensor and symmetric one-body matrix over
    n = h + p orbitals (holes first), sliced into the reference's blocks.

    Returns (fock=[f_pp,f_ph,f_hh], dense=[v_pppp,v_ppph,v_pphh,v_phph,v_phhh,v_hhhh], full V, F).
    """
    rng = np.random.default_rng(seed)
    n = p + h
    V = rng.standard_normal((n, n, n, n)) * scale
    V = V - V.transpose(1, 0, 2, 3)
    V = V - V.transpose(0, 1, 3, 2)
    V = V + V.transpose(2, 3, 0, 1)
    F = rng.standard_normal((n, n)) * scale
    F = F + F.T
    F[np.diag_indices(n)] += np.concatenate([-3.0 - rng.random(h), 3.0 + rng.random(p)])
    o, v = slice(0, h), slice(h, n)
    fock = [F[v, v].copy(), F[v, o].copy(), F[o, o].copy()]
    dense = [V[v, v, v, v].copy(), V[v, v, v, o].copy(), V[v, v, o, o].copy(),
             V[v, o, v, o].copy(), V[v, o, o, o].copy(), V[o, o, o, o].copy()]
    return fock, dense, V, F


def to_coo(t):
    """Dense tensor -> reference COO list [[i0,i1,i2,i3,val], ...] of its nonzeros."""
    idx = np.argwhere(t != 0)
    return [[int(a), int(b), int(c), int(d), float(t[a, b, c, d])] for a, b, c, d in idx]


def random_amplitudes(p, h, seed=1, scale=0.05, antisym=True):
    rng = np.random.default_rng(seed)
    t1 = rng.standard_normal((p, h)) * scale
    t2 = rng.standard_normal((p, p, h, h)) * scale
    if antisym:
        t2 = t2 - t2.transpose(1, 0, 2, 3)
        t2 = t2 - t2.transpose(0, 1, 3, 2)
    return t1, t2

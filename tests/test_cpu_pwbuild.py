"""Host side of the GPU plane-wave build: the analytically known components give the same
pair-block packing as the connected-component search over the COO list, and the group-major
entry order (what nl_pw_coo writes) is the COO list's order."""
import numpy as np
import pytest
import torch

from nulattice_amd.planewave import PlaneWave, shell_holes
from nulattice_amd.plans import PairBlocks, coo_arrays


def zeros_fill(rows, cols, nb, mb, kb):
    return torch.zeros((nb, mb, kb), dtype=torch.float64)


@pytest.mark.parametrize("L,A", [(3, 28), (3, 4), (2, 4)])
def test_components_pack_like_coo(L, A):
    pw = PlaneWave(L, shell_holes(L, A), -4.0, 1.0)
    p, h = pw.p, pw.h
    ku, ks, ke, kpair = pw.pair_groups(pw.parts, pw.parts)
    # v_pppp
    idx, val = coo_arrays(pw.coo_pppp(), p, h, "pppp")
    host = PairBlocks(idx[:, 0] * p + idx[:, 1], idx[:, 2] * p + idx[:, 3], val, p * p, p * p,
                      "cpu")
    comp = sorted([kpair[a:b] for a, b in zip(ks, ke)], key=lambda r: int(r[0]))
    new = PairBlocks.from_components(comp, comp, zeros_fill, "cpu")
    assert (host.nb, host.mb, host.kb, host.nnz) == (new.nb, new.mb, new.kb, new.nnz)
    assert torch.equal(host.rows, new.rows) and torch.equal(host.cols, new.cols)
    # v_ppph: group-major order of (rows x columns) is the COO list's order, and R56's packing
    idx3, val3 = coo_arrays(pw.coo_ppph(), p, h, "ppph")
    hu, hs, he, hpair = pw.pair_groups(pw.parts, pw.holes)
    pp = {int(u): kpair[a:b] for u, a, b in zip(ku, ks, ke)}
    ph = {int(u): hpair[a:b] for u, a, b in zip(hu, hs, he)}
    common = [int(u) for u in ku if int(u) in ph]
    rows = np.concatenate([np.repeat(pp[u], len(ph[u])) for u in common])
    cols = np.concatenate([np.tile(ph[u], len(pp[u])) for u in common])
    assert np.array_equal(rows, idx3[:, 0] * p + idx3[:, 1])
    assert np.array_equal(cols, idx3[:, 2] * h + idx3[:, 3])
    c, d, a, k = (idx3[:, n] for n in range(4))
    host = PairBlocks(a * h + k, c * p + d, val3, p * h, p * p, "cpu")
    order = sorted(common, key=lambda u: int(ph[u][0]))
    new = PairBlocks.from_components([ph[u] for u in order], [pp[u] for u in order], zeros_fill,
                                     "cpu")
    assert (host.nb, host.mb, host.kb, host.nnz) == (new.nb, new.mb, new.kb, new.nnz)
    assert torch.equal(host.rows, new.rows) and torch.equal(host.cols, new.cols)


def test_codes():
    pw = PlaneWave(5, shell_holes(5, 28), -4.0, 0.0)
    orbs = np.arange(4 * 125)
    c = pw.codes(orbs).astype(np.int64)
    k = np.stack([c & 255, (c >> 8) & 255, (c >> 16) & 255], 1)
    assert np.array_equal(((k[:, 0] * 5 + k[:, 1]) * 5 + k[:, 2]) * 4 + (c >> 24), orbs)


def test_ppph_specs_rebuild_the_host_plans():
    """The linear key / source / destination forms that nl_segplan_sort and nl_segplan_finish
    evaluate on the device (ppph_specs), applied here with a stable NumPy sort in their place,
    give every v_ppph plan (forward and transposed) of the host builder exactly (plane-wave
    L=3 A=28 list)."""
    from nulattice_amd.plans import PpphPlans, ppph_specs
    pw = PlaneWave(3, shell_holes(3, 28), -4.0, 1.0)
    p, h = pw.p, pw.h
    idx, val = coo_arrays(pw.coo_ppph(), p, h, "ppph")
    host = PpphPlans(idx, val, p, h, "cpu")
    adj = host.adjoint(lambda V: V.transpose(1, 2).contiguous())
    for name, (wk, ws, sign, kmax, spec) in ppph_specs(p, h).items():
        ref = adj[name] if name.endswith("T") else getattr(host, name)
        key, src = idx @ np.array(wk), idx @ np.array(ws)
        assert key.max() < kmax
        o = np.argsort(key, kind="stable")
        uk, start = np.unique(key[o], return_index=True)
        dst = sum(((uk // dv) % md if md else uk // dv) * mul for dv, md, mul in spec)
        assert ref.S == uk.size, name
        assert np.array_equal(ref.ptr.numpy(), np.append(start, key.size)), name
        assert np.array_equal(ref.src.numpy(), src[o]), name
        assert np.array_equal(ref.val.numpy(), sign * val[o]), name
        assert np.array_equal(ref.dst.numpy(), dst), name

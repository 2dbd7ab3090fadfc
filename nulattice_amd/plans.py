"""Host-side packing of the sparse two-body blocks into device plans (built once per Hamiltonian).

v_pppp (COO <ab||cd>) becomes *pair blocks*: the bipartite graph between row pairs (a,b) and
column pairs (c,d) is split into connected components, each a dense block V_b; small blocks are
packed block-diagonally and all are padded to one (rows, cols) shape so the pp ladder is one
batched MFMA GEMM between a row gather and a row scatter.  For plane-wave inputs the components
are the (total momentum, spin-isospin pair) sectors; for position-space contacts they are sites.

v_ppph (COO <cd||ak>) becomes *segment plans* for nl_segsum: for each intermediate of
v_ppph_dgrams (ccDgrams.py:13-45) the entries are grouped by output row so that every output row
is owned by one wavefront (no atomics).
"""
import ctypes
from typing import NamedTuple

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import torch


class CooBlock(NamedTuple):
    """A sparse two-body block as rows: indices (N, 4) int64 [a, b, c, d], values (N,) -- the
    row layout of _torch_types.TwoBodyOperator (_torch_types.py:169-199), as a named pair."""
    indices: np.ndarray
    values: np.ndarray


def coo_arrays(v, pnum, hnum, kind):
    """Normalise a reference-format sparse block to (idx (N,4) int64, val (N,) float64).

    Accepts, by container:
      * the legacy list [[i0,i1,i2,i3,val], ...] (coupled_cluster.py:256-258) or an (N,5) array;
      * an object with ``.indices`` (N,4) and ``.values`` (N,): _torch_types.TwoBodyOperator
        (_torch_types.py:169-199), this package's CooBlock and hamiltonian.CooList;
      * a bare (indices, values) pair as soa/coupled_cluster.py:18-46 (to_soa_sparse) returns
        it: indices (4, N); (N, 4) indices are accepted only when N != 4, where the two
        layouts cannot be confused;
      * a dense array (sparse=False path), whose nonzeros are taken."""
    shape = (pnum, pnum, pnum, pnum if kind == "pppp" else hnum)
    if hasattr(v, "indices") and hasattr(v, "values"):
        idx = np.asarray(_np(v.indices), dtype=np.int64)
        val = np.asarray(_np(v.values), dtype=np.float64).reshape(-1)
        if val.size == 0:
            idx = idx.reshape(0, 4)
        if idx.shape != (val.size, 4):
            raise ValueError(f"{kind}: container indices have shape {idx.shape}, expected "
                             f"({val.size}, 4)")
    elif isinstance(v, tuple) and len(v) == 2:
        idx = np.asarray(_np(v[0]), dtype=np.int64)
        val = np.asarray(_np(v[1]), dtype=np.float64).reshape(-1)
        n = val.size
        if idx.size == 0 and n == 0:
            idx = idx.reshape(0, 4)
        elif idx.shape == (4, n):
            idx = idx.T                      # to_soa_sparse layout (including n == 4)
        elif idx.shape != (n, 4):
            raise ValueError(f"{kind}: (indices, values) pair with indices {idx.shape} and "
                             f"{n} values; expected (4, {n}) as to_soa_sparse returns")
    else:
        arr = _np(v)
        arr = np.asarray(arr, dtype=np.float64)
        if arr.ndim == 4:
            if arr.shape != shape:
                raise ValueError(f"dense {kind} block has shape {arr.shape}, expected {shape}")
            nz = np.nonzero(arr)
            idx = np.stack(nz, axis=1).astype(np.int64)
            val = arr[nz]
        else:
            arr = arr.reshape(-1, 5) if arr.size else np.zeros((0, 5))
            idx = arr[:, :4].astype(np.int64)
            val = arr[:, 4].astype(np.float64)
    idx = np.ascontiguousarray(idx)
    if idx.size:
        hi = np.array(shape)
        if (idx < 0).any() or (idx >= hi).any():
            raise ValueError(f"{kind} COO index out of range for p={pnum}, h={hnum}")
    return idx, val


def _np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return x


def _group(items, labels):
    """{label: items with that label} via one sort."""
    o = np.argsort(labels, kind="stable")
    lab, it = labels[o], items[o]
    u, start = np.unique(lab, return_index=True)
    ends = np.append(start[1:], lab.size)
    return {int(x): it[s:e] for x, s, e in zip(u, start, ends)}


def _pack_bins(r_list, c_list, min_rows):
    """Greedy block-diagonal packing of components (given as ascending row / column index
    arrays, in component order) into bins of >= ``min_rows`` rows, largest components first
    (ties in component order).  Returns (bins, rows (nb, mb), cols (nb, kb)), -1 padded."""
    order = sorted(range(len(r_list)), key=lambda q: -len(r_list[q]))
    bins, cur, cur_rows = [], [], 0
    for q in order:
        cur.append(q); cur_rows += len(r_list[q])
        if cur_rows >= min_rows:
            bins.append(cur); cur, cur_rows = [], 0
    if cur:
        bins.append(cur)
    mb = max(1, max(sum(len(r_list[q]) for q in b) for b in bins))
    kb = max(1, max(sum(len(c_list[q]) for q in b) for b in bins))
    rows = -np.ones((len(bins), mb), np.int64)
    cols = -np.ones((len(bins), kb), np.int64)
    for bi, b in enumerate(bins):
        r = np.concatenate([r_list[q] for q in b]); cc = np.concatenate([c_list[q] for q in b])
        rows[bi, :len(r)] = r; cols[bi, :len(cc)] = cc
    return bins, rows, cols


class PairBlocks:
    """Padded pair-block form of a COO (N,4) matrix over (row pair, col pair)."""

    def __init__(self, row, col, val, nrow, ncol, device, min_rows=128):
        self.nb = 0
        if row.size == 0:
            return
        g = sp.coo_matrix((np.ones(row.size), (row, nrow + col)), shape=(nrow + ncol,) * 2)
        ncomp, lab = csg.connected_components(g, directed=False)
        rlab = lab[row]
        comps = [int(x) for x in np.unique(rlab)]
        # rows / cols of each component
        ur = np.unique(row)
        uc = np.unique(col)
        r_of = _group(ur, lab[ur])
        c_of = _group(uc, lab[nrow + uc])
        bins, rows, cols = _pack_bins([r_of[c] for c in comps], [c_of[c] for c in comps],
                                      min_rows)
        nb, mb = rows.shape
        kb = cols.shape[1]
        rpos = np.empty(nrow, np.int64)   # row pair -> (bin, local row)
        cpos = np.empty(ncol, np.int64)
        rbin = np.empty(nrow, np.int64)
        for bi in range(nb):
            r = rows[bi][rows[bi] >= 0]; cc = cols[bi][cols[bi] >= 0]
            rpos[r] = np.arange(len(r)); cpos[cc] = np.arange(len(cc)); rbin[r] = bi
        flat = (rbin[row] * mb + rpos[row]) * kb + cpos[col]
        V = np.bincount(flat, weights=val, minlength=nb * mb * kb).reshape(nb, mb, kb)
        self.nb, self.mb, self.kb = nb, mb, kb
        self.nnz = row.size
        self.V = torch.tensor(V, dtype=torch.float64, device=device)
        self.rows = torch.tensor(rows.reshape(-1), dtype=torch.int64, device=device)
        self.cols = torch.tensor(cols.reshape(-1), dtype=torch.int64, device=device)

    @classmethod
    def from_components(cls, r_list, c_list, fill, device, min_rows=128):
        """Pair blocks whose connected components are known: component q couples exactly the
        rows ``r_list[q]`` and columns ``c_list[q]`` (ascending, components ordered by their
        smallest row, the order a graph search from row 0 upwards labels them), so the packing
        equals the one found from the COO list.  ``fill(rows, cols, nb, mb, kb)`` returns V as a
        (nb, mb, kb) device tensor from the device index tensors."""
        new = cls.__new__(cls)
        new.nb = 0
        if not r_list:
            return new
        _, rows, cols = _pack_bins(r_list, c_list, min_rows)
        nb, mb = rows.shape
        kb = cols.shape[1]
        new.nb, new.mb, new.kb = nb, mb, kb
        new.nnz = int(sum(len(r) * len(c) for r, c in zip(r_list, c_list)))
        new.rows = torch.tensor(rows.reshape(-1), dtype=torch.int64, device=device)
        new.cols = torch.tensor(cols.reshape(-1), dtype=torch.int64, device=device)
        new.V = fill(new.rows, new.cols, nb, mb, kb)
        return new

    def _like(self, V, rows, cols, mb, kb):
        new = PairBlocks.__new__(PairBlocks)
        new.nb, new.mb, new.kb, new.nnz = self.nb, mb, kb, getattr(self, "nnz", 0)
        new.V, new.rows, new.cols = V, rows, cols
        return new

    def transposed(self, transpose_blocks):
        """Blocks of the transposed matrix (rows and columns swapped, V_b -> V_b^T); a column
        pair belongs to exactly one block, so the transposed scatter is conflict-free too.
        ``transpose_blocks(V)`` returns V (nb, mb, kb) as a contiguous (nb, kb, mb) tensor."""
        if self.nb == 0:
            return self
        return self._like(transpose_blocks(self.V), self.cols, self.rows, self.kb, self.mb)

    def scaled(self, scale, scale_values):
        """The same blocks times ``scale``; ``scale_values(V, s)`` returns s * V."""
        if self.nb == 0:
            return self
        return self._like(scale_values(self.V, scale), self.rows, self.cols, self.mb, self.kb)


class SegPlan:
    """Segments of (key -> [(src_off, val)]) for nl_segsum, with per-segment destination offset."""

    def __init__(self, key, src_off, val, dst_of_key, device):
        order = np.argsort(key, kind="stable")
        key, src_off, val = key[order], src_off[order], val[order]
        uk, start = np.unique(key, return_index=True)
        ptr = np.append(start, key.size).astype(np.int64)
        self.S = int(uk.size)
        self.ptr = torch.tensor(ptr, dtype=torch.int64, device=device)
        self.src = torch.tensor(src_off.astype(np.int64), dtype=torch.int64, device=device)
        self.val = torch.tensor(val, dtype=torch.float64, device=device)
        self.dst = torch.tensor(np.asarray(dst_of_key(uk), dtype=np.int64), dtype=torch.int64,
                                device=device)

    @classmethod
    def on_device(cls, idx, val, wkey, wsrc, sign, kmax, dst_spec):
        """The same plan for a device COO list, built by nl_segplan_sort / nl_segplan_finish: the
        entries (c, d, a, k) = idx[0..3] (a (4, N) int64 tensor), key = wkey . (c, d, a, k) <
        kmax, src_off = wsrc . (c, d, a, k), value sign * val, and destination offset
        sum_t ((u // div_t) % mod_t) * mul_t of key u for dst_spec = ((div, mod, mul) x 3), mod 0
        meaning none.  A stable radix sort orders equal keys as the list does, so the plan equals
        the NumPy one built from the same list."""
        from . import _lib
        n = int(val.numel())
        dev = val.device
        arr = lambda v: (ctypes.c_int64 * len(v))(*[int(x) for x in v])  # noqa: E731
        ws_bytes = int(_lib.load().nl_segplan_ws(n))
        if ws_bytes < 0:
            raise _lib.NLError(f"nl_segplan_ws: unsupported list length {n}")
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        new = cls.__new__(cls)
        new.src = torch.empty(n, dtype=torch.int64, device=dev)
        new.val = torch.empty(n, dtype=torch.float64, device=dev)
        nseg = torch.empty(1, dtype=torch.int64, device=dev)
        st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        _lib.call("nl_segplan_sort", n, p(idx), arr(wkey), arr(wsrc), p(val),
                  ctypes.c_double(sign), max(1, int(kmax - 1).bit_length()), p(ws), ws_bytes,
                  p(new.src), p(new.val), p(nseg), st)
        new.S = int(nseg.item())
        new.ptr = torch.empty(new.S + 1, dtype=torch.int64, device=dev)
        new.dst = torch.empty(new.S, dtype=torch.int64, device=dev)
        _lib.call("nl_segplan_finish", new.S, n, p(ws), ws_bytes,
                  arr([x for t in dst_spec for x in t]), p(new.ptr), p(new.dst), st)
        return new

    def scaled(self, scale, scale_values):
        new = SegPlan.__new__(SegPlan)
        new.S, new.ptr, new.src, new.dst = self.S, self.ptr, self.src, self.dst
        new.val = scale_values(self.val, scale)
        return new


def ppph_specs(p, h):
    """The linear forms of the v_ppph segment plans over an entry (c, d, a, k) of <cd||ak>:
    name -> (key weights, source-offset weights, value sign, key bound, destination spec), with
    destination offset sum_t ((u // div_t) % mod_t) * mul_t of key u for spec ((div, mod, mul)
    x 3, mod 0 = none).  These are the keys, sources and destinations PpphPlans builds from a
    NumPy list (forward plans, then the transposed ones of ``adjoint``)."""
    hh, ph = h * h, p * h
    one = (1, 0, 0)
    return {
        "ret0": ((0, 0, h, 1), (p * hh, hh, 0, h), 1.0, ph, ((1, 0, h), one, one)),
        "ret1": ((0, 1, p, 0), (h, 0, 0, 1), 1.0, p * p, ((1, 0, 1), one, one)),
        "ret2": ((ph, h, 0, 1), (0, 0, h, 0), 1.0, p * ph,
                 ((ph, 0, h * ph), (h, p, h), (1, h, 1))),
        "ret2p": ((ph, h, 0, 1), (0, 0, h, 0), -1.0, p * ph, ((1, 0, h), one, one)),
        "ret4": ((0, h, ph, 1), (h, 0, 0, 0), 1.0, p * ph,
                 ((h, p, h * ph), (1, h, ph), (ph, 0, h))),
        "ret0T": ((ph, h, 0, 1), (0, 0, h, 0), 1.0, p * ph, ((1, 0, h), one, one)),
        "ret1T": ((h, 0, 0, 1), (0, 1, p, 0), 1.0, ph, ((1, 0, 1), one, one)),
        "ret2T": ((0, 0, 1, 0), (h * ph, h, 0, 1), 1.0, p, ((1, 0, h), one, one)),
        "ret4T": ((1, 0, 0, 0), (0, h * ph, h, ph), 1.0, p, ((1, 0, h), one, one)),
    }


class PpphPlans:
    """The five segment plans that replace v_ppph_dgrams' scalar loop (ccDgrams.py:36-44).

    Entry (c, d, a, k, v) of <cd||ak>:
      ret0[a,i]      -= 1/2 v t2[c,d,k,i]   key (a,k)    src (c,d,k,:) of t2 (then sum over k)
      ret1[a,d]      -= v t1[c,k]           key (a,d)    src (c,k) of t1
      ret2[c,d,i,k]  += v t1[a,i]           key (c,d,k)  src (a,:) of t1   (dest [c][i][d][k])
      ret2p: ret2 as P(ij) sees it, -v t1[a,x] at [c][d][k][x] (unit-stride destination)
      ret4[a,d,i,k]  += v t1[c,i]           key (a,d,k)  src (c,:) of t1   (into the ring W)
      R56[a,k,i,j]   += v tau[c,d,i,j]      pair blocks over rows (a,k), columns (c,d)
    """

    def __init__(self, idx, val, p, h, device, r56=None):
        """``idx`` (N, 4) and ``val`` (N,) as NumPy arrays, or ``idx`` a (4, N) int64 device
        tensor and ``val`` a device tensor (plans then sorted on the device by nl_segplan_sort).
        ``r56`` may pass the R56 pair blocks when their components are known
        (``PairBlocks.from_components``)."""
        on_dev = isinstance(val, torch.Tensor)
        self.n = int(val.numel()) if on_dev else val.size
        self.p, self.h, self.device = p, h, device
        self._adj = None
        if self.n == 0:
            return
        hh, ph = h * h, p * h
        if on_dev:
            self._coo = None
            self._dcoo = (idx, val.to(torch.float64))
            specs = ppph_specs(p, h)
            for name in ("ret0", "ret1", "ret2", "ret2p", "ret4"):
                setattr(self, name, SegPlan.on_device(*self._dcoo, *specs[name]))
            if r56 is None:
                raise ValueError("PpphPlans on a device list needs its R56 pair blocks")
            self.r56 = r56
            return
        c, d, a, k = (idx[:, n].astype(np.int64) for n in range(4))
        val = np.asarray(val, dtype=np.float64)
        self._coo = (c, d, a, k, val)
        # ret0 partials T0[a][k][i] (segments (a,k)); ret0[a,i] = sum_k T0[a,k,i]
        self.ret0 = SegPlan(a * h + k, ((c * p + d) * h + k) * h, val, lambda u: u * h, device)
        self.ret1 = SegPlan(a * p + d, c * h + k, val, lambda u: u, device)
        # ret2[c,d,i,k] lands in Sij stored as [c][i][d][k]: offset c*h*ph + i*ph + d*h + k
        self.ret2 = SegPlan((c * p + d) * h + k, a * h, val,
                            lambda u: (u // (p * h)) * h * ph + ((u // h) % p) * h + u % h,
                            device)
        # the same term as the residual uses it, under P(ik): -v t1[a,x] at Sij[c,d,k,x] of the
        # [a][b][i][j] accumulator, so x runs along its unit stride (offset ((c*p+d)*h+k)*h + x)
        self.ret2p = SegPlan((c * p + d) * h + k, a * h, -val, lambda u: u * h, device)
        # ret4[a,d,i,k] lands in W[(d,k),(a,i)] = W[(d*h + k)*ph + a*h + i]
        self.ret4 = SegPlan((a * p + d) * h + k, c * h, val,
                            lambda u: ((u // h) % p * h + u % h) * ph + (u // (p * h)) * h,
                            device)
        # R56[(a,k),(i,j)] = sum_(c,d) v[(c,d),(a,k)] tau[(c,d),(i,j)]: a pair-block GEMM
        self.r56 = r56 if r56 is not None else PairBlocks(a * h + k, c * p + d, val, p * h,
                                                           p * p, device)

    def scaled(self, scale, scale_values):
        """The same plans with every value times ``scale`` (adjoint plans are rebuilt on use
        from the scaled list)."""
        new = PpphPlans.__new__(PpphPlans)
        new.n, new.p, new.h, new.device, new._adj = self.n, self.p, self.h, self.device, None
        if self.n:
            if self._coo is None:
                new._coo = None
                new._dcoo = (self._dcoo[0], scale_values(self._dcoo[1], scale))
            else:
                new._coo = self._coo[:4] + (self._coo[4] * scale,)
            for name in ("ret0", "ret1", "ret2", "ret2p", "ret4"):
                setattr(new, name, getattr(self, name).scaled(scale, scale_values))
            new.r56 = self.r56.scaled(scale, scale_values)
        return new

    def adjoint(self, transpose_blocks):
        """Transposed plans for the Lambda equations (built once, on first use).

        Entry (c, d, a, k, v); l1 / gS* / gW are the adjoints of H1 / Sij / W / ret1:
          ret0T: gt2[c,d,k,i]  += -1/2 v l1[a,i]           key (c,d,k)  src (a,:) of l1
          ret1T: gt1[c,k]      -= v gret1[a,d]             key (c,k)    src (a,d) of gret1
          ret2T: gt1[a,i]      += v gSij[c][i][d][k]       key a        src (c,.,d,k), stride ph
          ret4T: gt1[c,i]      += v gW[(d,k),(a,i)]        key c        src ((d,k),(a,.))
          r56T:  gtau[(c,d),:] += 1/2 v gN56[(a,k),:]      transposed pair blocks
        """
        if self._adj is None and self.n:
            p, h, dev = self.p, self.h, self.device
            ph = p * h
            if self._coo is None:
                specs = ppph_specs(p, h)
                self._adj = {name: SegPlan.on_device(*self._dcoo, *specs[name])
                             for name in ("ret0T", "ret1T", "ret2T", "ret4T")}
                self._adj["r56T"] = self.r56.transposed(transpose_blocks)
                return self._adj
            c, d, a, k, val = self._coo
            self._adj = {
                "ret0T": SegPlan((c * p + d) * h + k, a * h, val, lambda u: u * h, dev),
                "ret1T": SegPlan(c * h + k, a * p + d, val, lambda u: u, dev),
                "ret2T": SegPlan(a, c * h * ph + d * h + k, val, lambda u: u * h, dev),
                "ret4T": SegPlan(c, (d * h + k) * ph + a * h, val, lambda u: u * h, dev),
                "r56T": self.r56.transposed(transpose_blocks),
            }
        return self._adj

"""Tensor-level wrappers over libnlccsd.so.

Device tensors come from PyTorch (allocation, streams); every arithmetic operation is a kernel of
libnlccsd.so.  ``contract`` maps a two-operand einsum onto one MFMA GEMM launch: indices are
classified into batch / M / N / K groups, each operand is used in place when its strides allow
(the group merges into one stride and K or M/N is unit-stride), otherwise it is first laid out by
the permute kernel; the result is written straight into the destination when its strides allow,
else through a scratch tensor and a permute-accumulate.
"""
import ctypes
import itertools
import math

import numpy as np
import torch

from . import _lib

BF16, FP16 = 0, 1


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _off16(t, n):
    """Pointer to 16-bit element n of t."""
    return ctypes.c_void_p(t.data_ptr() + 2 * int(n))


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(t, name, dtype=torch.float64):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise _lib.NLError(f"{name}: expected a HIP device tensor (no CPU fallback)")
    if t.dtype not in (dtype if isinstance(dtype, tuple) else (dtype,)):
        raise _lib.NLError(f"{name}: expected {dtype}, got {t.dtype}")


# Element-type flags of a product's output C and addend D (include/nulattice_ccsd.h): float
# instead of double.  The precision form keeps its T2-sized accumulators in FP32.
C_F32, D_F32 = 1, 2
F64_OR_F32 = (torch.float64, torch.float32)


def _io(out, addend, beta):
    """io flags of out = alpha A.B + beta (addend or out)."""
    io = C_F32 if out.dtype == torch.float32 else 0
    d = addend if addend is not None else out
    if beta != 0.0 and d.dtype == torch.float32:
        io |= D_F32
    return io


class Workspace:
    """Grow-only scratch buffers keyed by name (one set per device)."""

    def __init__(self):
        self.bufs = {}

    def get(self, name, numel, dtype=torch.float64, device=None):
        b = self.bufs.get(name)
        if b is None or b.numel() < numel or b.dtype != dtype:
            b = torch.empty(max(numel, 1), dtype=dtype, device=device)
            self.bufs[name] = b
        return b[:numel]


WS = Workspace()

# Low-precision MFMA is used only when both output dimensions are at least this large; thinner
# products (the h x h dressings) are bandwidth-bound and stay FP64.
LP_MIN_DIM = 256

# Optional kernel timers: when TIMERS["on"] is set, a contract() call with tag=T records a pair
# of events on the launch stream around its GEMM launch into TIMERS[T].
TIMERS = {"on": False}

# Optional count of the matrix-core work launched: when FLOPS["on"] is set, every GEMM launch adds
# its executed FLOP (2 M N K per product; a split product counts each of its partial products)
# to FLOPS["fp64"] or FLOPS["lp"].
FLOPS = {"on": False, "fp64": 0.0, "lp": 0.0}


def count_flops(kind, flop):
    if FLOPS["on"]:
        FLOPS[kind] += float(flop)


def _timed(tag):
    if TIMERS.get("all"):
        tag = tag or _timed.spec
    if not (tag and TIMERS["on"]):
        return None
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    TIMERS.setdefault(tag, []).append((s, e))
    return s, e


# ----------------------------------------------------------------------------- permute
def permute_into(src, out, alpha=1.0, beta=0.0):
    """out = alpha * src + beta * out, where ``src`` may be any strided view with out's shape
    and ``out`` is contiguous."""
    _check(src, "src"); _check(out, "out")
    if tuple(src.shape) != tuple(out.shape) or not out.is_contiguous():
        raise _lib.NLError("permute_into: shape mismatch or non-contiguous output")
    if out.numel() == 0:
        return out
    shape = [int(s) for s in out.shape] or [1]
    strides = [int(s) for s in src.stride()] or [1]
    # drop size-1 dims and merge adjacent dims that are contiguous in both views
    dims = [(sz, st) for sz, st in zip(shape, strides) if sz != 1] or [(1, 1)]
    merged = [list(dims[0])]
    for sz, st in dims[1:]:
        psz, pst = merged[-1]
        if pst == st * sz:
            merged[-1] = [psz * sz, st]
        else:
            merged.append([sz, st])
    if len(merged) > 6:
        raise _lib.NLError("permute_into: more than 6 non-mergeable dims")
    sh = (ctypes.c_int64 * len(merged))(*[m[0] for m in merged])
    st = (ctypes.c_int64 * len(merged))(*[m[1] for m in merged])
    base = src.data_ptr()
    _lib.call("nl_permute", len(merged), sh, st, ctypes.c_double(alpha), ctypes.c_void_p(base),
              ctypes.c_double(beta), _p(out), _stream())
    return out


def copy(src):
    out = torch.empty(src.shape, dtype=torch.float64, device=src.device)
    return permute_into(src, out)


def axpby(a, x, b, y):
    """y = a*x + b*y (contiguous, same numel)."""
    _check(x, "x"); _check(y, "y")
    _lib.call("nl_axpby", y.numel(), ctypes.c_double(a), _p(x), ctypes.c_double(b), _p(y),
              _stream())
    return y


GRAM_MAX, LINCOMB_MAX = 12, 16


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def gram(vs):
    """Host FP64 matrix G[a, b] = vs[a] . vs[b] of contiguous device vectors of one size and
    dtype (float64, or float32 for a history kept in FP32; FP64 products and sums either way):
    one nl_gram_part pass over up to 12 vectors; larger sets run as pairs of groups of 6 (each
    pair one pass over its 12 vectors).  Fixed summation order: bitwise reproducible."""
    k = len(vs)
    if k == 0:
        return np.zeros((0, 0))
    n = vs[0].numel()
    f32 = vs[0].dtype == torch.float32
    for v in vs:
        _check(v, "v", torch.float32 if f32 else torch.float64)
        if v.numel() != n or not v.is_contiguous():
            raise ValueError("gram: vectors must be contiguous and of one size")
    G = np.empty((k, k))
    if k <= GRAM_MAX:
        sets = [list(range(k))]
    else:
        g = [list(range(s, min(s + GRAM_MAX // 2, k))) for s in range(0, k, GRAM_MAX // 2)]
        sets = [g[i] + g[j] for i in range(len(g)) for j in range(i + 1, len(g))]
    nparts = int(_lib.load().nl_partials(n))
    for idx in sets:
        m = len(idx)
        npair = m * (m + 1) // 2
        part = torch.empty(npair * nparts, dtype=torch.float64, device=vs[0].device)
        out = torch.empty(npair, dtype=torch.float64, device=vs[0].device)
        _lib.call("nl_gram_part_f32" if f32 else "nl_gram_part", n, m,
                  _ptrs([vs[i] for i in idx]), _p(part), _stream())
        _lib.call("nl_gram_finalize", nparts, npair, _p(part), _p(out), _stream())
        o = out.cpu().numpy()
        q = 0
        for a in range(m):
            for b in range(a, m):
                G[idx[a], idx[b]] = G[idx[b], idx[a]] = o[q]
                q += 1
    return G


def lincomb(coef, vs, out):
    """out (float64) = sum_k coef[k] * vs[k] (contiguous, same numel; vs float64, or float32
    for a history kept in FP32): one nl_lincomb pass over up to 16 vectors; longer sums continue
    in chunks that add to out."""
    _check(out, "out")
    vs = list(vs)
    if len(coef) != len(vs) or not vs:
        raise ValueError("lincomb: need one coefficient per vector")
    n = out.numel()
    if vs[0].dtype == torch.float32:
        for v in vs:
            _check(v, "v", torch.float32)
            if v.numel() != n or not v.is_contiguous():
                raise ValueError("lincomb: vectors must be contiguous and of one size")
        tmp = None
        for k0 in range(0, len(vs), LINCOMB_MAX):
            xs = vs[k0:k0 + LINCOMB_MAX]
            cs = [float(c) for c in coef[k0:k0 + LINCOMB_MAX]]
            dst = out if k0 == 0 else tmp
            _lib.call("nl_lincomb_f32", n, len(xs), (ctypes.c_double * len(cs))(*cs), _ptrs(xs),
                      _p(dst), _stream())
            if k0:
                axpby(1.0, tmp, 1.0, out)
            elif len(vs) > LINCOMB_MAX:
                tmp = torch.empty_like(out)
        return out
    for v in vs + [out]:
        _check(v, "v")
        if v.numel() != n or not v.is_contiguous():
            raise ValueError("lincomb: vectors must be contiguous and of one size")
    k0 = 0
    while k0 < len(vs):
        if k0 == 0:
            xs, cs = list(vs[:LINCOMB_MAX]), [float(c) for c in coef[:LINCOMB_MAX]]
        else:
            xs = [out] + list(vs[k0:k0 + LINCOMB_MAX - 1])
            cs = [1.0] + [float(c) for c in coef[k0:k0 + LINCOMB_MAX - 1]]
        _lib.call("nl_lincomb", n, len(xs), (ctypes.c_double * len(cs))(*cs), _ptrs(xs), _p(out),
                  _stream())
        k0 += len(xs) - (0 if k0 == 0 else 1)
    return out


# ----------------------------------------------------------------------------- contract
def _parse(spec):
    lhs, out = spec.replace(" ", "").split("->")
    a, b = lhs.split(",")
    return a, b, out


def _splits(M, N, K, batch):
    """Split-K factor so that a launch has enough workgroups (deterministic reduction).  Tiny
    outputs with a long K (the h x h dressings: one tile, K = p^2 h) get up to 1024 splits: each
    workgroup's K loop is a chain of dependent steps, so their count, not the bytes, sets the
    time."""
    tiles = math.ceil(M / 128) * math.ceil(N / 128) * batch
    if tiles >= 512 or K <= 512:
        return 1
    return int(max(1, min(1024, 2048 // tiles, K // 256)))


def _lp_splits(M, N, K, batch):
    """Split-K for nl_gemm_lp, which runs its 256x256 tiles when tiles x splits >= 256: for a
    long K (>= 4096) whose 256x256 tiles alone do not fill the 256 CUs, the smallest split giving
    >= 256 workgroups with the last round >= 80% busy (each split keeping >= 512 of K)."""
    big = math.ceil(M / 256) * math.ceil(N / 256) * batch
    if M >= 256 and N >= 256 and K >= 4096 and big < 256:
        for s in range(2, 65):
            w = big * s
            if K // s < 512:
                break
            if w >= 256 and w >= 0.8 * 256 * math.ceil(w / 256):
                return s
    return _splits(M, N, K, batch)


def _order(t, labels, group):
    """Indices of ``group`` sorted by decreasing stride of ``t`` (its memory order)."""
    return sorted(group, key=lambda c: (-t.stride(labels.index(c)), labels.index(c)))


def _splits_at(t, labels, order):
    """Positions i (0 < i < len) where order[i-1], order[i] cannot merge in t."""
    out = set()
    for i in range(1, len(order)):
        x, y = labels.index(order[i - 1]), labels.index(order[i])
        if t.shape[x] == 1 or t.shape[y] == 1:
            continue
        if t.stride(x) != t.stride(y) * t.shape[y]:
            out.add(i)
    return out


def _group_runs(order, tensors, size):
    """Split a group into <= 2 runs valid for every tensor.  tensors: [(t, labels)] holding all
    of the group's labels, or (None, None) for a tensor the group broadcasts over.
    Returns (n1, n2, [(s1, s2) per tensor]) or None."""
    cuts = set()
    for t, lab in tensors:
        if t is not None:
            cuts |= _splits_at(t, lab, order)
    if len(cuts) > 1:
        return None
    cut = min(cuts) if cuts else 0
    r1, r2 = order[:cut], order[cut:]
    n1 = math.prod(size[c] for c in r1)
    n2 = math.prod(size[c] for c in r2)
    strides = []
    for t, lab in tensors:
        if t is None:
            strides.append((0, 0))
            continue
        st = lambda run: t.stride(lab.index(run[-1])) if run else 0
        strides.append((st(r1), st(r2)))
    return n1, n2, strides


def _best_runs(group, tensors, size):
    """Try every order of a (small) group; prefer orders whose inner run is unit-stride in the
    first tensor."""
    best = None
    for order in (itertools.permutations(group) if len(group) <= 4 else [tuple(group)]):
        r = _group_runs(list(order), tensors, size)
        if r is None:
            continue
        score = 0 if (tensors[0][0] is not None and r[2][0][1] == 1) else 1
        if best is None or score < best[0]:
            best = (score, list(order), r)
            if score == 0:
                break
    return None if best is None else (best[1], best[2])


def _plan(A, la, B, lb, out, lo, size):
    bat = [c for c in lo if c in la and c in lb]
    Mi = [c for c in lo if c in la and c not in lb]
    Ni = [c for c in lo if c in lb and c not in la]
    Ki = [c for c in la if c in lb and c not in lo]
    opt = [(bat, Mi, Ni)]
    # an outer M (or N) index may become a broadcast batch index when it cannot merge
    if not bat:
        if len(Mi) > 1:
            opt += [([c], [x for x in Mi if x != c], Ni) for c in Mi]
        if len(Ni) > 1:
            opt += [([c], Mi, [x for x in Ni if x != c]) for c in Ni]
    for bat_, M_, N_ in opt:
        if math.prod(size[c] for c in bat_) > 65535:
            continue
        tb = lambda t, lab: (t, lab) if all(c in lab for c in bat_) else (None, None)
        rb = _best_runs(bat_, [tb(out, lo), tb(A, la), tb(B, lb)], size) if bat_ else \
            ([], (1, 1, [(0, 0)] * 3))
        rk = _best_runs(Ki, [(A, la), (B, lb)], size) if Ki else ([], (1, 1, [(0, 0)] * 2))
        rm = _best_runs(M_, [(out, lo), (A, la)], size) if M_ else ([], (1, 1, [(0, 0)] * 2))
        rn = _best_runs(N_, [(out, lo), (B, lb)], size) if N_ else ([], (1, 1, [(0, 0)] * 2))
        if None in (rb, rk, rm, rn):
            continue
        (b1, b2, bs), (k1, k2, ks), (m1, m2, ms), (n1, n2, ns) = rb[1], rk[1], rm[1], rn[1]
        dims = [m1, m2, n1, n2, k1, k2, b1, b2]
        sa = [ms[1][0], ms[1][1], ks[0][0], ks[0][1], bs[1][0], bs[1][1]]
        sb = [ks[1][0], ks[1][1], ns[1][0], ns[1][1], bs[2][0], bs[2][1]]
        sc = [ms[0][0], ms[0][1], ns[0][0], ns[0][1], bs[0][0], bs[0][1]]
        # coalescing score: unit inner strides for A (k or m), B (n or k) and C (n)
        score = (sa[3] != 1 and sa[1] != 1) + (sb[3] != 1 and sb[1] != 1) + 2 * (sc[3] != 1)
        return dict(dims=dims, sa=sa, sb=sb, sc=sc, score=score)
    return None


def _copy_layout(t, labels, order):
    perm = [labels.index(c) for c in order]
    return copy(t.permute(*perm)), "".join(order)


def contract(spec, A, B, out, alpha=1.0, beta=1.0, precision="fp64", tag=None, addend=None):
    """out = alpha * einsum(spec, A, B) + beta * (addend if given else out)  (one MFMA launch).

    Every operand may be any strided view: the index groups (batch, M, N, K) are split into at
    most two stride runs per tensor, which the contraction kernel addresses directly.  An operand
    that cannot be expressed so is first laid out contiguously by the permute kernel.
    precision: "fp64" (MFMA f64), "bf16" or "fp16" (MFMA with FP32 accumulation, per-row
    power-of-two scaling of both operands; FP64 output; used only when both output dimensions
    are >= LP_MIN_DIM).  ``addend`` must have ``out``'s shape.
    """
    la, lb, lo = _parse(spec)
    _timed.spec = spec
    _check(A, "A", F64_OR_F32); _check(B, "B", F64_OR_F32); _check(out, "out", F64_OR_F32)
    if addend is not None:
        _check(addend, "addend", F64_OR_F32)
        if tuple(addend.shape) != tuple(out.shape) or addend.stride() != out.stride():
            raise _lib.NLError(f"contract {spec}: addend must match out's shape and strides")
    if len(la) != A.dim() or len(lb) != B.dim() or len(lo) != out.dim():
        raise _lib.NLError(f"contract {spec}: rank mismatch {A.shape} {B.shape} {out.shape}")
    size = {}
    for lab, t in ((la, A), (lb, B), (lo, out)):
        for c, n in zip(lab, t.shape):
            if size.setdefault(c, n) != n:
                raise _lib.NLError(f"contract {spec}: size mismatch for index {c}")
    for c in set(la) | set(lb):
        if c not in lo and not (c in la and c in lb):
            raise _lib.NLError(f"contract {spec}: index {c} summed in one operand only")
    if math.prod(size[c] for c in lo) == 0:
        return out
    if precision != "fp64":
        Mi = [c for c in lo if c in la and c not in lb]
        Ni = [c for c in lo if c in lb and c not in la]
        M = math.prod(size[c] for c in Mi); N = math.prod(size[c] for c in Ni)
        if min(M, N) < LP_MIN_DIM or any(c in la and c in lb for c in lo):
            precision = "fp64"      # thin or batched products stay exact
    if precision == "fp64" and (A.dtype != torch.float64 or B.dtype != torch.float64):
        raise _lib.NLError(f"contract {spec}: FP32 operands are read by the 16-bit products only")
    if ABSENT_PTRS and precision == "fp64" and (A.data_ptr() in ABSENT_PTRS or
                                                B.data_ptr() in ABSENT_PTRS):
        raise _lib.NLError(f"contract {spec}: FP64 read of a tensor whose FP64 values are absent")
    if precision != "fp64":
        Ki = _order(A, la, [c for c in la if c in lb and c not in lo])
        K = math.prod(size[c] for c in Ki)
        return _contract_lp(A, la, B, lb, out, lo, Mi, Ni, Ki, M, N, K, alpha, beta, precision,
                            tag, addend)
    io = _io(out, addend, beta)
    # choose the operand order with the best coalescing among the copy-free plans
    cands = []
    for (X, lx, Y, ly) in ((A, la, B, lb), (B, lb, A, la)):
        pl = _plan(X, lx, Y, ly, out, lo, size)
        if pl is not None:
            cands.append((pl["score"], X, Y, pl))
    direct = True
    if cands:
        cands.sort(key=lambda c: c[0])
        _, X, Y, pl = cands[0]
    else:
        # lay out the operands contiguously as [bat][M][K] and [bat][K][N]
        bat = [c for c in lo if c in la and c in lb]
        Mi = [c for c in lo if c in la and c not in lb]
        Ni = [c for c in lo if c in lb and c not in la]
        Ki = [c for c in la if c in lb and c not in lo]
        X, lx = _copy_layout(A, la, bat + Mi + Ki)
        Y, ly = _copy_layout(B, lb, bat + Ki + Ni)
        pl = _plan(X, lx, Y, ly, out, lo, size)
        if pl is None:
            if not out.is_contiguous():
                raise _lib.NLError(f"contract {spec}: destination must be contiguous")
            order = bat + Mi + Ni
            tmp = WS.get("contract_out", out.numel(), device=out.device).view(
                *[size[c] for c in order])
            tv = tmp.permute(*[order.index(c) for c in lo])
            pl = _plan(X, lx, Y, ly, tv, lo, size)
            direct = False
            if io:
                raise _lib.NLError(f"contract {spec}: an FP32 output must be addressable directly")
    dims, sa, sb, sc = pl["dims"], pl["sa"], pl["sb"], pl["sc"]
    arr = lambda v: (ctypes.c_int64 * len(v))(*[int(x) for x in v])
    nws = int(_lib.load().nl_contract_ws(arr(dims)))
    ws = WS.get("splitk", nws, device=out.device) if nws > 0 else None
    splits = -1           # kernel and K split chosen by the library
    C = out if direct else tv
    D = (addend if addend is not None else out) if (direct and beta != 0.0) else None
    ev = _timed(tag)
    if ev:
        ev[0].record()
    count_flops("fp64", 2.0 * math.prod(dims))
    _lib.call("nl_contract_f64", arr(dims), ctypes.c_double(alpha), _p(X), arr(sa), _p(Y),
              arr(sb), ctypes.c_double(beta if direct else 0.0), _p(D), _p(C), arr(sc), splits,
              _p(ws), io, _stream())
    if ev:
        ev[1].record()
    if not direct:
        if addend is not None and beta != 0.0:
            permute_into(addend, out, beta, 0.0)
            permute_into(tv, out, 1.0, 1.0)
        else:
            permute_into(tv, out, 1.0, beta)
    return out


def _view4(t, labels, rows, cols):
    """(R1, R2, C1, C2, strides4) of t seen as rows x cols with <= 2 merged parts each."""
    def parts(group):
        # split group (in given order) into maximal mergeable runs
        runs = []
        for c in group:
            d = labels.index(c)
            if t.shape[d] == 1:
                continue
            if runs and runs[-1][1] == t.stride(d) * t.shape[d]:
                runs[-1] = [runs[-1][0] * t.shape[d], t.stride(d)]
            else:
                runs.append([t.shape[d], t.stride(d)])
        while len(runs) < 2:
            runs.insert(0, [1, 0])
        return runs
    pr, pc = parts(rows), parts(cols)
    if len(pr) > 2 or len(pc) > 2:
        return None
    return pr[0][0], pr[1][0], pc[0][0], pc[1][0], [pr[0][1], pr[1][1], pc[0][1], pc[1][1]]


# Tensors that stay constant for the life of a solve (Hamiltonian blocks): their low-precision
# conversions are cached by (storage pointer, view, layout, dtype).
CONST_PTRS = set()
# Hamiltonian blocks among them: converted with exact per-row scales (target < 0 in
# nl_convert_rows), so blocks whose nonzero elements share one magnitude are exact in 16 bits.
EXACT_PTRS = set()
_LP_CACHE = {}
# Storage whose FP64 values are not formed (only 16-bit images registered by lp_provide): any
# other read of it raises instead of using stale data.
ABSENT_PTRS = set()


def mark_constant(*ts, exact=True):
    for t in ts:
        CONST_PTRS.add(t.data_ptr())
        if exact:
            EXACT_PTRS.add(t.data_ptr())


def forget_ptrs(ptrs):
    """Drop constant registrations (and cached conversions) of freed storage."""
    for ptr in ptrs:
        CONST_PTRS.discard(ptr)
        EXACT_PTRS.discard(ptr)
        for k in [k for k in _LP_CACHE if k[0] == ptr]:
            del _LP_CACHE[k]


class lp_reuse:
    """Context in which the given tensors do not change: their 16-bit conversions are cached as
    those of constants are, and dropped on exit.  Lets two products that read one operand in the
    same layout (the two ring GEMMs and t2) share a single conversion."""

    def __init__(self, *ts, absent=False):
        self.ptrs = [t.data_ptr() for t in ts if t.data_ptr() not in CONST_PTRS]
        self.absent = absent

    def __enter__(self):
        CONST_PTRS.update(self.ptrs)
        if self.absent:
            ABSENT_PTRS.update(self.ptrs)
        return self

    def __exit__(self, *exc):
        forget_ptrs(self.ptrs)
        ABSENT_PTRS.difference_update(self.ptrs)
        return False


def lp_provide(t, labels, rows, cols, buf, ldk, dtype=BF16, parts=(1, 0), scale=1.0):
    """Register ``buf`` ([R][ldk] 16-bit rows of t * scale, no row scales; with parts=(2, 0b10)
    each row is [hi | lo] of the split) as the conversion of t seen as rows x cols, for products
    inside an ``lp_reuse(t)`` block: they use it without reading t, so t's FP64 values need not
    exist.  ``scale`` (a power of two) is divided out of the products through alpha."""
    v = _view4(t, labels, rows, cols)
    if v is None or t.data_ptr() not in CONST_PTRS:
        raise _lib.NLError("lp_provide: the view must be two-run and t under lp_reuse")
    key = (t.data_ptr(), v[0], v[1], v[2], v[3], tuple(v[4]), dtype)
    if parts != (1, 0):
        key = key + (parts,)
    _LP_CACHE[key] = (buf, None, ldk, 1.0 / scale)


def lp_target(dtype, exact):
    """nl_convert_rows target: Hamiltonian blocks (exact) get exact per-row scales (a contact
    interaction's +-g elements become +-1 in BF16, +-2^14 in FP16: its magnitude is not rounded
    to 8 or 11 bits, which would bias every term the block enters by the same relative amount);
    other FP16 operands power-of-two scales; other BF16 operands a plain cast."""
    if exact:
        return -1.0 if dtype == BF16 else -(2.0 ** 14)
    return 2.0 ** 14 if dtype != BF16 else 0.0


def lp_kind(precision):
    """(dtype, split) of a low-precision name: "bf16", "fp16" (one rounding of each operand) or
    "bf16x3", "fp16x3" (hi/lo split operands, three products hi.hi + lo.hi + hi.lo)."""
    if precision not in ("bf16", "fp16", "bf16x3", "fp16x3", "bf16x2a", "bf16x2b"):
        raise ValueError(f"unknown precision {precision!r}")
    return (BF16 if precision.startswith("bf16") else FP16), precision[4:] or False


def split_parts(exact_a, exact_b):
    """Operand images of a split product, as (nblk, lomask) K-blocks per row: a rounded operand
    is [hi | lo]; an exact (Hamiltonian) operand is its hi part alone, or [hi | hi] when it meets
    a rounded one, so that the K-concatenated product is hi.hi + hi.lo.  Two rounded operands
    stay [hi | lo] each and go to the split GEMM (hi.hi + hi.lo + lo.hi in one kernel)."""
    if exact_a and exact_b:
        return (1, 0), (1, 0)
    if exact_a:
        return (2, 0), (2, 0b10)
    if exact_b:
        return (2, 0b10), (2, 0)
    return (2, 0b10), (2, 0b10)


def to_lp_rows(t, labels, rows, cols, dtype, tag, parts=(1, 0)):
    """Convert view t (rows x cols) to a 16-bit [R][ldk] buffer: Hamiltonian blocks with exact
    per-row scales, other FP16 operands with per-row power-of-two scales (returned as inverse
    scales), other BF16 operands by a plain cast (inverse scales None).  Conversions of constant
    storage are cached by the memory view they read (pointer, extents, strides), so differently
    labelled views of the same rows share one."""
    v = _view4(t, labels, rows, cols)
    key = None
    if v is not None and t.data_ptr() in CONST_PTRS:
        key = (t.data_ptr(), v[0], v[1], v[2], v[3], tuple(v[4]), dtype)
        if parts != (1, 0):
            key = key + (parts,)
        hit = _LP_CACHE.get(key)
        if hit is not None:
            return hit          # (buf, row inverse scales or None, ldk, tensor inverse scale)
    if t.data_ptr() in ABSENT_PTRS:
        raise _lib.NLError("to_lp_rows: no 16-bit image registered for this view of a tensor "
                           "whose FP64 values are not formed")
    if v is None:
        t, labels = _copy_layout(t, labels, rows + cols)
        v = _view4(t, labels, rows, cols)
    R1, R2, C1, C2, st = v
    R, Ck = R1 * R2, C1 * C2
    kpad = (Ck + 7) // 8 * 8
    nblk, mask = parts
    ldk = nblk * kpad
    target = lp_target(dtype, key is not None and t.data_ptr() in EXACT_PTRS)
    scaled = target != 0.0          # BF16 has FP32's exponent range: a plain cast suffices
    if key is not None:
        buf = torch.empty(R * ldk, dtype=torch.int16, device=t.device)
        isc = torch.empty(R, dtype=torch.float32, device=t.device) if scaled else None
    else:
        buf = WS.get(tag, R * ldk, dtype=torch.int16, device=t.device)
        isc = WS.get(tag + "_s", R, dtype=torch.float32, device=t.device) if scaled else None
    s4 = (ctypes.c_int64 * 4)(*st)
    if parts == (1, 0):
        if t.dtype != torch.float64:
            raise _lib.NLError("to_lp_rows: the single-rounding conversion reads FP64")
        _lib.call("nl_convert_rows", dtype, R1, R2, C1, C2, s4, ctypes.c_void_p(t.data_ptr()),
                  _p(buf), ldk, _p(isc), ctypes.c_float(target), _stream())
    else:
        src_t = dtype | ((D_F32 << 8) if t.dtype == torch.float32 else 0)   # FP32 source
        _lib.call("nl_convert_split", src_t, R1, R2, C1, C2, s4, ctypes.c_void_p(t.data_ptr()),
                  _p(buf), ldk, kpad, nblk, mask, _p(isc), ctypes.c_float(target),
                  ctypes.c_double(1.0), _stream())
    if key is not None:
        _LP_CACHE[key] = (buf, isc, ldk, 1.0)
    return buf, isc, ldk, 1.0


def _sizes(t, labels):
    return {c: int(n) for c, n in zip(labels, t.shape)}


def _rowmajor(t, labels, Mi, Ni):
    """ldc if t seen as [M][N] (M = Mi merged, N = Ni merged and unit-stride), else None."""
    if _splits_at(t, labels, Mi) or _splits_at(t, labels, Ni):
        return None
    nn = [c for c in Ni if t.shape[labels.index(c)] != 1]
    mm = [c for c in Mi if t.shape[labels.index(c)] != 1]
    if nn and t.stride(labels.index(nn[-1])) != 1:
        return None
    return t.stride(labels.index(mm[-1])) if mm else 1


def _contract_lp(A, la, B, lb, out, lo, Mi, Ni, Ki, M, N, K, alpha, beta, precision, tag=None,
                 addend=None):
    dt, split = lp_kind(precision)
    pa = pb = (1, 0)
    if split:
        ex = lambda t, lab, rows: (t.data_ptr() in EXACT_PTRS
                                   and _view4(t, lab, rows, Ki) is not None)
        pa, pb = split_parts(ex(A, la, Mi) or split == "x2b", ex(B, lb, Ni) or split == "x2a")
    Ab, ia, lda, sa = to_lp_rows(A, la, Mi, Ki, dt, "lpA", pa)
    Bb, ib, ldb, sb = to_lp_rows(B, lb, Ni, Ki, dt, "lpB", pb)
    alpha = alpha * sa * sb            # tensor scales of provided images (powers of two)
    # both operands rounded: [hi | lo] rows of each, lo parts passed to the split GEMM
    both = pa == (2, 0b10) and pb == (2, 0b10)
    Kg = lda // 2 if both else lda
    lo_a = ctypes.c_void_p(Ab.data_ptr() + 2 * Kg) if both else None
    lo_b = ctypes.c_void_p(Bb.data_ptr() + 2 * Kg) if both else None
    ldc = _rowmajor(out, lo, Mi, Ni)
    runs = None
    if ldc is None:
        # the output as <= 2 stride runs per group (in the group orders the operands used)
        rm = _group_runs(Mi, [(out, lo)], _sizes(out, lo))
        rn = _group_runs(Ni, [(out, lo)], _sizes(out, lo))
        if rm is not None and rn is not None:
            runs = (rm, rn)
    direct = ldc is not None or runs is not None
    io = _io(out, addend, beta)
    if io and not direct:
        raise _lib.NLError("contract: an FP32 output must be addressable directly")
    C = out if direct else WS.get("contract_out", M * N, device=out.device).view(M, N)
    D = (addend if addend is not None else out) if (direct and beta != 0.0) else None
    dt = dt | (io << 8)             # C / D element types ride in the dtype argument
    splits = _lp_splits(M, N, Kg, 1)
    ws = WS.get("splitk", splits * M * N, device=out.device) if splits > 1 else None
    count_flops("lp", 2.0 * M * N * (3 * Kg if both else lda))
    ev = _timed(tag)
    if ev:
        ev[0].record()
    arr = lambda v: (ctypes.c_int64 * 4)(*[int(x) for x in v])
    if runs is not None:
        (m1, m2, (sm,)), (n1, n2, (sn,)) = runs
        dims, strides = arr([m1, m2, n1, n2]), arr([sm[0], sm[1], sn[0], sn[1]])
        if both:
            _lib.call("nl_gemm_lp_strided_split", dt, Kg, ctypes.c_double(alpha), _p(Ab), lo_a,
                      lda, _p(ia), _p(Bb), lo_b, ldb, _p(ib), ctypes.c_double(beta), _p(D),
                      _p(C), dims, strides, splits, _p(ws), _stream())
        else:
            _lib.call("nl_gemm_lp_strided", dt, lda, ctypes.c_double(alpha), _p(Ab), lda, _p(ia),
                      _p(Bb), ldb, _p(ib), ctypes.c_double(beta), _p(D), _p(C), dims, strides,
                      splits, _p(ws), _stream())
    elif both:
        _lib.call("nl_gemm_lp_split", dt, M, N, Kg, ctypes.c_double(alpha), _p(Ab), lo_a, lda, 0,
                  _p(ia), 0, _p(Bb), lo_b, ldb, 0, _p(ib), 0,
                  ctypes.c_double(beta if direct else 0.0), _p(D), _p(C),
                  max(ldc if ldc is not None else N, 1), 0, 1, splits, _p(ws), _stream())
    else:
        _lib.call("nl_gemm_lp", dt, M, N, lda, ctypes.c_double(alpha), _p(Ab), lda, 0, _p(ia), 0,
                  _p(Bb), ldb, 0, _p(ib), 0, ctypes.c_double(beta if direct else 0.0), _p(D),
                  _p(C), max(ldc if ldc is not None else N, 1), 0, 1, splits, _p(ws), _stream())
    if ev:
        ev[1].record()
    if not direct:
        shp = [out.shape[lo.index(c)] for c in Mi + Ni]
        perm = [(Mi + Ni).index(c) for c in lo]
        tv = C.view(*shp).permute(*perm)
        if addend is not None and beta != 0.0:
            permute_into(addend, out, beta, 0.0)
            permute_into(tv, out, 1.0, 1.0)
        else:
            permute_into(tv, out, 1.0, beta)
    return out

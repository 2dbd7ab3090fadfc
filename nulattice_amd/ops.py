"""Tensor-level wrappers over libnlccsd.so.

Device tensors come from PyTorch (allocation, streams); every arithmetic operation is a kernel of
libnlccsd.so.  ``contract`` maps a two-operand einsum onto one MFMA GEMM launch: indices are
classified into batch / M / N / K groups, each operand is used in place when its strides allow
(the group merges into one stride and K or M/N is unit-stride), otherwise it is first laid out by
the permute kernel; the result is written straight into the destination when its strides allow,
else through a scratch tensor and a permute-accumulate.
"""
import ctypes
import math

import torch

from . import _lib

BF16, FP16 = 0, 1


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(t, name, dtype=torch.float64):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise _lib.NLError(f"{name}: expected a HIP device tensor (no CPU fallback)")
    if t.dtype != dtype:
        raise _lib.NLError(f"{name}: expected {dtype}, got {t.dtype}")


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

# Optional kernel timers: when TIMERS["on"] is set, a contract() call with tag=T records a pair
# of events on the launch stream around its GEMM launch into TIMERS[T].
TIMERS = {"on": False}


def _timed(tag):
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


# ----------------------------------------------------------------------------- contract
def _parse(spec):
    lhs, out = spec.replace(" ", "").split("->")
    a, b = lhs.split(",")
    return a, b, out


def _merge(t, dims):
    """(size, stride) of dims merged into one index, or None if the strides do not allow it.
    Empty group -> (1, 0)."""
    dims = [d for d in dims if t.shape[d] != 1]
    if not dims:
        return 1, 0
    size = 1
    for x, y in zip(dims[:-1], dims[1:]):
        if t.stride(x) != t.stride(y) * t.shape[y]:
            return None
    for d in dims:
        size *= t.shape[d]
    return size, t.stride(dims[-1])


def _order(t, labels, group):
    """Indices of ``group`` sorted by decreasing stride of ``t`` (its memory order)."""
    return sorted(group, key=lambda c: (-t.stride(labels.index(c)), labels.index(c)))


def _layout(t, labels, outer, g1, g2):
    """Try t as [outer][g1][g2] with g2 unit-stride. Returns (s_outer, s_g1) or None."""
    pos = lambda g: [labels.index(c) for c in g]
    mo, m1, m2 = _merge(t, pos(outer)), _merge(t, pos(g1)), _merge(t, pos(g2))
    if mo is None or m1 is None or m2 is None:
        return None
    n2 = math.prod(t.shape[labels.index(c)] for c in g2)
    if n2 > 1 and m2[1] != 1:
        return None
    return mo[1], m1[1]


def _as_layout(t, labels, order):
    """Contiguous copy of t with its dims in ``order``."""
    perm = [labels.index(c) for c in order]
    return copy(t.permute(*perm))


def _splits(M, N, K, batch):
    tiles = math.ceil(M / 128) * math.ceil(N / 128) * batch
    if tiles >= 512 or K <= 512:
        return 1
    return int(max(1, min(256, 1024 // tiles, K // 256)))


def _operand(t, labels, bat, g1, g2):
    """Strides of operand t as [bat][g1][g2] (g2 unit-stride) or [bat][g2][g1] (g1 unit-stride).

    Returns (inner_is_g2, s_bat, ld) or None.  Batch labels absent from t get stride 0 (only a
    single broadcast batch label is allowed)."""
    present = [c for c in bat if c in labels]
    if len(present) != len(bat) and len(bat) != 1:
        return None
    for first, second, flag in ((g1, g2, True), (g2, g1, False)):
        lay = _layout(t, labels, present, first, second)
        if lay is not None:
            s_bat = lay[0] if present else 0
            return flag, s_bat, lay[1]
    return None


def _plan(A, la, B, lb, out, lo, size):
    """GEMM plan for out[lo] = sum A[la] B[lb]: groups, operand layouts and the bytes that would
    have to be copied (operands that need a permute, an out that needs a scratch)."""
    bat = [c for c in lo if c in la and c in lb]
    Mi = [c for c in lo if c in la and c not in lb]
    Ni = [c for c in lo if c in lb and c not in la]
    Ki = [c for c in la if c in lb and c not in lo]
    big, lbig = (A, la) if A.numel() >= B.numel() else (B, lb)
    bat = _order(big, lbig, bat)
    Ki = _order(big, lbig, Ki)
    Mi = _order(A, la, Mi)
    Ni = _order(B, lb, Ni)
    opA = _operand(A, la, bat, Mi, Ki)
    opB = _operand(B, lb, bat, Ki, Ni)
    # an unmergeable outer M (or N) index becomes a broadcast batch index
    if not bat and (opA is None or opB is None):
        cand = None
        if opA is None and len(Mi) > 1:
            cand = ([Mi[0]], Mi[1:], Ni)
        elif opB is None and len(Ni) > 1:
            cand = ([Ni[0]], Mi, Ni[1:])
        if cand is not None:
            b2, M2, N2 = cand
            oa, ob = _operand(A, la, b2, M2, Ki), _operand(B, lb, b2, Ki, N2)
            if oa is not None and ob is not None and size[b2[0]] <= 65535:
                bat, Mi, Ni, opA, opB = b2, M2, N2, oa, ob
    nbat = math.prod(size[c] for c in bat)
    outok = _layout(out, lo, bat, Mi, Ni) is not None and nbat <= 65535
    cost = ((A.numel() if opA is None else 0) + (B.numel() if opB is None else 0)
            + (0 if outok else 2 * out.numel()))
    return dict(A=A, la=la, B=B, lb=lb, bat=bat, Mi=Mi, Ni=Ni, Ki=Ki, opA=opA, opB=opB,
                outok=outok, cost=cost)


def contract(spec, A, B, out, alpha=1.0, beta=1.0, precision="fp64", tag=None):
    """out[spec] = alpha * einsum(spec, A, B) + beta * out   (one GEMM launch).

    precision: "fp64" (MFMA f64), "bf16" or "fp16" (MFMA with FP32 accumulation, per-row
    power-of-two scaling of both operands; FP64 output).  The operand order is chosen so that the
    fewest bytes need a layout change.  ``out`` must be contiguous unless its strides already
    match the GEMM's output layout.
    """
    la, lb, lo = _parse(spec)
    _check(A, "A"); _check(B, "B"); _check(out, "out")
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
        Ki = _order(A, la, [c for c in la if c in lb and c not in lo])
        if any(c in la and c in lb for c in lo):
            raise _lib.NLError("contract: batched low-precision contraction not supported")
        M = math.prod(size[c] for c in Mi); N = math.prod(size[c] for c in Ni)
        K = math.prod(size[c] for c in Ki)
        return _contract_lp(A, la, B, lb, out, lo, Mi, Ni, Ki, M, N, K, alpha, beta, precision,
                            tag)
    p1 = _plan(A, la, B, lb, out, lo, size)
    p2 = _plan(B, lb, A, la, out, lo, size)
    pl = p1 if p1["cost"] <= p2["cost"] else p2
    A, la, B, lb = pl["A"], pl["la"], pl["B"], pl["lb"]
    bat, Mi, Ni, Ki, opA, opB = pl["bat"], pl["Mi"], pl["Ni"], pl["Ki"], pl["opA"], pl["opB"]
    if opA is None:
        present = [c for c in bat if c in la]
        A = _as_layout(A, la, present + Mi + Ki); la = "".join(present + Mi + Ki)
        opA = _operand(A, la, bat, Mi, Ki)
    if opB is None:
        present = [c for c in bat if c in lb]
        B = _as_layout(B, lb, present + Ki + Ni); lb = "".join(present + Ki + Ni)
        opB = _operand(B, lb, bat, Ki, Ni)
    a_k, sA, lda = opA          # a_k: A is [M][K] (K unit-stride)
    b_n, sB, ldb = opB          # b_n: B is [K][N] (N unit-stride)
    M = math.prod(size[c] for c in Mi); N = math.prod(size[c] for c in Ni)
    K = math.prod(size[c] for c in Ki); nb = math.prod(size[c] for c in bat)
    if nb > 65535:
        raise _lib.NLError(f"contract {spec}: batch {nb} too large")
    if pl["outok"]:
        C, (sC, ldc) = out, _layout(out, lo, bat, Mi, Ni)
        direct = True
    else:
        if not out.is_contiguous():
            raise _lib.NLError(f"contract {spec}: destination must be contiguous")
        C = WS.get("contract_out", nb * M * N, device=out.device)
        sC, ldc, direct = M * N, N, False
    lda = max(lda, 1); ldb = max(ldb, 1); ldc = max(ldc, 1)
    splits = _splits(M, N, K, nb)
    ws = WS.get("splitk", nb * splits * M * N, device=out.device) if splits > 1 else None
    ev = _timed(tag)
    if ev:
        ev[0].record()
    _lib.call("nl_gemm_f64", int(a_k), int(b_n), M, N, K, ctypes.c_double(alpha), _p(A), lda, sA,
              _p(B), ldb, sB, ctypes.c_double(beta if direct else 0.0), _p(C), ldc, sC, nb,
              splits, _p(ws), _stream())
    if ev:
        ev[1].record()
    if not direct:
        order = bat + Mi + Ni
        Cv = C.view(*[size[c] for c in order])
        permute_into(Cv.permute(*[order.index(c) for c in lo]), out, 1.0, beta)
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
_LP_CACHE = {}


def mark_constant(*ts):
    for t in ts:
        CONST_PTRS.add(t.data_ptr())


def forget_ptrs(ptrs):
    """Drop constant registrations (and cached conversions) of freed storage."""
    for ptr in ptrs:
        CONST_PTRS.discard(ptr)
        for k in [k for k in _LP_CACHE if k[0] == ptr]:
            del _LP_CACHE[k]


def to_lp_rows(t, labels, rows, cols, dtype, tag):
    """Convert view t (rows x cols) to a 16-bit [R][ldk] buffer with per-row scales."""
    key = None
    if t.data_ptr() in CONST_PTRS:
        key = (t.data_ptr(), tuple(t.shape), tuple(t.stride()), labels, tuple(rows), tuple(cols),
               dtype)
        hit = _LP_CACHE.get(key)
        if hit is not None:
            return hit
    v = _view4(t, labels, rows, cols)
    if v is None:
        t = _as_layout(t, labels, rows + cols); labels = "".join(rows + cols)
        v = _view4(t, labels, rows, cols)
    R1, R2, C1, C2, st = v
    R, Ck = R1 * R2, C1 * C2
    ldk = (Ck + 7) // 8 * 8
    if key is not None:
        buf = torch.empty(R * ldk, dtype=torch.int16, device=t.device)
        isc = torch.empty(R, dtype=torch.float32, device=t.device)
    else:
        buf = WS.get(tag, R * ldk, dtype=torch.int16, device=t.device)
        isc = WS.get(tag + "_s", R, dtype=torch.float32, device=t.device)
    target = 1.0 if dtype == BF16 else 2.0 ** 14
    s4 = (ctypes.c_int64 * 4)(*st)
    _lib.call("nl_convert_rows", dtype, R1, R2, C1, C2, s4, ctypes.c_void_p(t.data_ptr()),
              _p(buf), ldk, _p(isc), ctypes.c_float(target), _stream())
    if key is not None:
        _LP_CACHE[key] = (buf, isc, ldk)
    return buf, isc, ldk


def _contract_lp(A, la, B, lb, out, lo, Mi, Ni, Ki, M, N, K, alpha, beta, precision, tag=None):
    dt = BF16 if precision == "bf16" else FP16
    Ab, ia, lda = to_lp_rows(A, la, Mi, Ki, dt, "lpA")
    Bb, ib, ldb = to_lp_rows(B, lb, Ni, Ki, dt, "lpB")
    lay = _layout(out, lo, [], Mi, Ni)
    if lay is not None:
        C, (_, ldc), direct = out, lay, True
    else:
        C = WS.get("contract_out", M * N, device=out.device).view(M, N)
        ldc, direct = N, False
    splits = _splits(M, N, lda, 1)
    ws = WS.get("splitk", splits * M * N, device=out.device) if splits > 1 else None
    ev = _timed(tag)
    if ev:
        ev[0].record()
    _lib.call("nl_gemm_lp", dt, M, N, lda, ctypes.c_double(alpha), _p(Ab), lda, 0, _p(ia), 0,
              _p(Bb), ldb, 0, _p(ib), 0, ctypes.c_double(beta if direct else 0.0), _p(C),
              max(ldc, 1), 0, 1, splits, _p(ws), _stream())
    if ev:
        ev[1].record()
    if not direct:
        shp = [out.shape[lo.index(c)] for c in Mi + Ni]
        perm = [(Mi + Ni).index(c) for c in lo]
        permute_into(C.view(*shp).permute(*perm), out, 1.0, beta)
    return out

"""Device-resident CCSD residual of NuLattice's CCM path, evaluated with libnlccsd.so kernels.

``Hamiltonian`` holds the normal-ordered blocks on the GPU in the reference layout
(f_pp[a,b], f_ph[a,i], f_hh[i,j], v_pphh[a,b,i,j], v_phph[a,i,b,j], v_phhh[a,i,j,k],
v_hhhh[i,j,k,l]; coupled_cluster.py:99-263) plus the derived layouts the kernels want, and the
packed plans of the sparse blocks v_pppp / v_ppph.

``t1_part``/``t2_part``/``update`` compute, for amplitudes (t1, t2), exactly the sums of the
reference's t1Iter and t2Iter (coupled_cluster.py:342-545, ccDgrams.py) regrouped for the GPU:

* T2 terms are accumulated into four buffers by antisymmetriser, S0 (none), Sab (P(ab)),
  Sij (P(ij)) and Spp (P(ij)P(ab), stored as [(a,i)][(b,j)]); one fused kernel applies the
  antisymmetrisers, the dressed denominator, the Jacobi step, the mixing and the energy.
* Ring family (A22 v_phph ring, A24 quadratic ring, A37 ppph ring, A29):
      Spp[(a,i),(b,j)] += sum_(c,k) t2[a,c,i,k] W[(c,k),(b,j)],
      W = -v_phph[b,k,c,j] + 1/2 sum_(d,l) v_pphh[c,d,k,l] t2[d,b,l,j] + ret4[b,c,j,k]
          - sum_l v_phhh[c,j,k,l] t1[b,l]
  (P(ij)P(ab) is invariant under (a,i)<->(b,j), which moves every ring term onto t2[a,c,i,k]).
* hh-ladder family (A21, A25, A26, A30-A33, A35):  S0 += 1/2 sum_kl tau[a,b,k,l] Whh[k,l,i,j],
      Whh = v_hhhh + P(ij) sum_c v_phhh[c,j,k,l] t1[c,i] + 1/2 sum_cd v_pphh[c,d,k,l] tau[c,d,i,j],
      tau = t2 + P(ij) t1 t1.
* pp ladder (A9): S0 += 1/2 sum_cd v_pppp[a,b,c,d] tau[c,d,i,j], block-dense batched GEMM.
* v_ppph consumers (A10, A36-A39): segment plans; A39's two terms become one contraction with
  R56[b,k,i,j] = sum_cd v_ppph[c,d,b,k] tau[c,d,i,j].
Every regrouping is an exact identity of the reference's sums; only floating-point summation
order changes.

Precision form: ``prec`` maps the families "ring", "hh", "dress", "ladder" to "fp64", "bf16" or
"fp16"; low-precision contractions use BF16/FP16 MFMA with FP32 accumulation (FP16 operands with
per-row power-of-two scales, BF16 operands plainly rounded), everything else (Fock, denominators,
update, energy, DIIS) stays FP64.
"""
import contextlib
import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from . import ops
from .ops import contract, permute_into
from .plans import PairBlocks, PpphPlans, coo_arrays

F64 = torch.float64

# nl_assemble_t2_ext flags: FP32 accumulators; v_pphh added to H2 by the assembly; S0 over the
# hole pairs i < j; the split image of t2n written over the hole pairs; Sab over the hole pairs
ACC_F32, ADD_V, S0_PAIRS, IMG_PAIRS_IO, SAB_PAIRS = 1, 2, 4, 8, 16


def _dev(x, device):
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=device, dtype=F64).contiguous()
    # torch.tensor keeps a NumPy array's strides (an F-ordered block stays column-major), and
    # the kernels read row-major memory
    return torch.tensor(np.ascontiguousarray(x, dtype=np.float64), dtype=F64, device=device)


class Hamiltonian:
    """Normal-ordered Hamiltonian blocks on one GPU (reference layout + derived layouts)."""

    def __init__(self, fock_mats, two_body_int, device="cuda", min_block_rows=128):
        """``two_body_int`` in the reference's order (v_pppp, v_ppph, v_pphh, v_phph, v_phhh,
        v_hhhh); v_pppp may also be given as built PairBlocks and v_ppph as built PpphPlans
        (``planewave.PlaneWave.device_hamiltonian`` writes both on the device)."""
        device = torch.device(device)
        if device.type != "cuda":
            raise _lib.NLError("Hamiltonian: a HIP device is required (no CPU fallback)")
        _lib.load()
        f_pp, f_ph, f_hh = fock_mats
        v_pppp, v_ppph, v_pphh, v_phph, v_phhh, v_hhhh = two_body_int
        self.device = device
        self.f_pp, self.f_ph, self.f_hh = (_dev(x, device) for x in (f_pp, f_ph, f_hh))
        self.p, self.h = self.f_ph.shape
        p, h = self.p, self.h
        self.v_pphh = _dev(v_pphh, device)
        self.v_phph = _dev(v_phph, device)
        self.v_phhh = _dev(v_phhh, device)
        self.v_hhhh = _dev(v_hhhh, device)
        for name, t, shp in (("f_pp", self.f_pp, (p, p)), ("f_hh", self.f_hh, (h, h)),
                             ("v_pphh", self.v_pphh, (p, p, h, h)),
                             ("v_phph", self.v_phph, (p, h, p, h)),
                             ("v_phhh", self.v_phhh, (p, h, h, h)),
                             ("v_hhhh", self.v_hhhh, (h, h, h, h))):
            if tuple(t.shape) != shp:
                raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shp}")
        # v_pphh with (c,k),(d,l) adjacent: physical [c][k][d][l]
        self.vP = ops.copy(self.v_pphh.permute(0, 2, 1, 3))
        # constant part of the ring intermediate: -v_phph[b,k,c,j] at W[c][k][b][j]
        self.Wc = torch.empty((p, h, p, h), dtype=F64, device=device)
        permute_into(self.v_phph.permute(2, 1, 0, 3), self.Wc, alpha=-1.0)
        # v_phph as [b][k][i][c] (A27) and v_pphh as [c][d][k][l] = v_pphh[d,c,k,l] (A19)
        self.v_phph_bkic = ops.copy(self.v_phph.permute(0, 1, 3, 2))
        self.v_pphh_T = ops.copy(self.v_pphh.permute(1, 0, 2, 3))
        self.ones_h = torch.ones(h, dtype=F64, device=device)
        self.eye_h = torch.eye(h, dtype=F64, device=device)
        consts = (self.f_pp, self.f_ph, self.f_hh, self.v_pphh, self.v_phph, self.v_phhh,
                  self.v_hhhh, self.vP, self.Wc, self.v_phph_bkic, self.v_pphh_T)
        ops.mark_constant(*consts)
        weakref.finalize(self, ops.forget_ptrs, [t.data_ptr() for t in consts])
        if isinstance(v_pppp, PairBlocks):      # plans already built on the device
            self.pppp = v_pppp
        else:
            idx, val = coo_arrays(v_pppp, p, h, "pppp")
            self.pppp = PairBlocks(idx[:, 0] * p + idx[:, 1], idx[:, 2] * p + idx[:, 3], val,
                                   p * p, p * p, device, min_block_rows)
        if isinstance(v_ppph, PpphPlans):
            self.ppph = v_ppph
        else:
            idx3, val3 = coo_arrays(v_ppph, p, h, "ppph")
            self.ppph = PpphPlans(idx3, val3, p, h, device)

    def vpphh_view(self):
        """v_pphh[c,d,k,l] as a view of the [c][k][d][l] copy."""
        return self.vP.permute(0, 2, 1, 3)

    _TWO_BODY = ("v_pphh", "v_phph", "v_phhh", "v_hhhh", "vP", "Wc", "v_phph_bkic", "v_pphh_T")

    def rescaled(self, fock_mats, scale):
        """Hamiltonian with Fock blocks ``fock_mats`` and every two-body block ``scale`` times this
        one's (same sparsity plans).  For a contact theory whose couplings enter the normal-ordered
        blocks affinely this is the derivative Hamiltonian dh/dtheta."""
        new = Hamiltonian.__new__(Hamiltonian)
        new.device, new.p, new.h = self.device, self.p, self.h
        f_pp, f_ph, f_hh = fock_mats
        new.f_pp, new.f_ph, new.f_hh = (_dev(x, self.device) for x in (f_pp, f_ph, f_hh))
        for name, t, shp in (("f_pp", new.f_pp, (self.p, self.p)),
                             ("f_ph", new.f_ph, (self.p, self.h)),
                             ("f_hh", new.f_hh, (self.h, self.h))):
            if tuple(t.shape) != shp:
                raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shp}")
        for name in self._TWO_BODY:
            setattr(new, name, _scaled(getattr(self, name), scale))
        new.ones_h, new.eye_h = self.ones_h, self.eye_h
        new.pppp = self.pppp.scaled(scale, _scaled)
        new.ppph = self.ppph.scaled(scale, _scaled)
        consts = [new.f_pp, new.f_ph, new.f_hh] + [getattr(new, n) for n in self._TWO_BODY]
        ops.mark_constant(*consts)
        weakref.finalize(new, ops.forget_ptrs, [t.data_ptr() for t in consts])
        return new


def _scaled(t, s):
    out = torch.empty_like(t)
    ops.axpby(float(s), t, 0.0, out)
    return out


class Buffers:
    """Per-iteration device buffers (allocated once per problem size)."""

    def __init__(self, p, h, device):
        z = lambda *s: torch.zeros(s, dtype=F64, device=device)
        self.H1, self.X1pp, self.X1hh = z(p, h), z(p, p), z(h, h)
        self.X2pp, self.X2hh = z(p, p), z(h, h)
        self.s_ph, self.u_ph, self.w_ph, self.z_ph = z(p, h), z(p, h), z(p, h), z(p, h)
        self.q_hh = z(h, h)
        self.Dpp, self.Dhh = z(p, p), z(h, h)     # v.t2 dressings shared by both parts
        self.Mpp, self.Mhh = z(p, p), z(h, h)     # combined Sab / Sij dressing matrices
        self.dMpp = z(p)             # Mpp's diagonal when the Sab product runs in low precision
        self.dpp = None              # the diagonal the assembly adds in FP64 (None: none)
        self.ret1 = z(p, p)
        self.T0 = z(p, h, h)
        self.tau = z(p, p, h, h)
        self.S0, self.Sab = z(p, p, h, h), z(p, p, h, h)
        self.Sij = z(p, p, h, h)     # P(ij) group as [a][b][i][j]
        self.Spp = z(p, h, p, h)
        self.W = z(p, h, p, h)       # ring intermediate [c][k][b][j]
        self.Whh, self.zz = z(h, h, h, h), z(h, h, h, h)
        self.N56 = z(p, h, h, h)
        self.Xidkl = z(h, p, h, h)
        self.Yilaj = z(h, h, p, h)
        self.Zilaj = z(h, h, p, h)   # Yilaj - U[a,l,j,i]: the two Spp tails as one operand
        self.F_ph = z(p, h)          # f_ph + s: the A12 + A14 dressing of H1
        self.Ubkij = z(p, h, h, h)
        self.part = z(2 * int(_lib.load().nl_partials(p * p * h * h)) + 8)
        self.E = z(1)
        self.p, self.h, self.device = p, h, device
        self._acc32 = None
        self._sab_img = None
        self.t2_16 = self.t2_split = self.t2_pairs = None
        self.sab_pairs = False       # Sab of this step over the hole pairs (NL_SAB_PAIRS)

    def s0_pairs(self, Pp):
        """S0 of the split form over the hole pairs: [p*p][Pp] FP32 (slots from P on unused)."""
        n = self.p * self.p * Pp
        b = getattr(self, "_s0p", None)
        if b is None or b.numel() != n:
            b = self._s0p = torch.zeros(n, dtype=torch.float32, device=self.device)
        return b

    def pair_operand(self, Pp):
        """[Pp][hi(Pp) | lo(Pp)] 16-bit rows of Whh^T over the hole pairs (zero-initialised: the
        rows and columns from P on stay zero) and their FP32 row scales (ones initially)."""
        b = getattr(self, "_whhT", None)
        if b is None or b[0].numel() != Pp * 2 * Pp:
            b = self._whhT = (torch.zeros(Pp * 2 * Pp, dtype=torch.int16, device=self.device),
                              torch.ones(Pp, dtype=torch.float32, device=self.device))
        return b

    def n56_pairs(self, Pp):
        """N56 over the hole pairs: [(b,k)][Pp] FP64, zero-initialised (slots from P on stay
        zero: the split Sab product reads them)."""
        n = self.p * self.h * Pp
        b = getattr(self, "_n56p", None)
        if b is None or b.numel() != n:
            b = self._n56p = torch.zeros(n, dtype=torch.float64, device=self.device)
        return b

    def sab_image(self, kp, rows=None):
        """The K-concatenated [N56 | t2^T] operand of the split Sab product: ``rows`` (p*h*h, or
        p*Pp over the hole pairs) rows of [hi(kp) | lo(kp)] 16-bit elements, zero-initialised
        once (its pad columns stay zero)."""
        n = (self.p * self.h * self.h if rows is None else rows) * 2 * kp
        if self._sab_img is None or self._sab_img.numel() != n:
            self._sab_img = torch.zeros(n, dtype=torch.int16, device=self.device)
        return self._sab_img

    def acc32(self):
        """The FP32 accumulators of the split precision form (allocated on first use): S0, Sab,
        Sij, Spp and the ring intermediate W, with the FP64 buffers' layouts.  Their products
        accumulate in FP32 (16-bit MFMA) or FP64 (the rest) and are rounded once per pass, so
        each read-modify-write pass over them moves half the bytes of an FP64 one."""
        if self._acc32 is None:
            p, h = self.p, self.h
            z = lambda *s: torch.zeros(s, dtype=torch.float32, device=self.device)
            # W is the ring product's operand: FP32 only where that product runs on the 16-bit
            # matrix cores (which read it through a split conversion); thinner rings are FP64
            W = z(p, h, p, h) if p * h >= ops.LP_MIN_DIM else self.W
            self._acc32 = Acc(z(p, p, h, h), z(p, p, h, h), z(p, p, h, h), z(p, h, p, h), W)
        return self._acc32


class Acc:
    """A set of T2-sized accumulators (S0, Sab, Sij as [a][b][i][j]; Spp, W as [a][i][b][j])."""

    def __init__(self, S0, Sab, Sij, Spp, W):
        self.S0, self.Sab, self.Sij, self.Spp, self.W = S0, Sab, Sij, Spp, W


def _segsum(plan, X, src, out, alpha, src_stride=1, dst_stride=1):
    if plan is None:
        return
    io = ops.C_F32 if out.dtype == torch.float32 else 0
    _lib.call("nl_segsum", plan.S, X, ops._p(plan.ptr), ops._p(plan.src), ops._p(plan.val),
              ops._p(plan.dst), src_stride, dst_stride, ctypes.c_double(alpha), ops._p(src),
              ops._p(out), io, ops._stream())


DEFAULT_PREC = {"ring": "fp64", "hh": "fp64", "dress": "fp64", "ladder": "fp64"}

# The diagonal of the Sab dressing matrix is applied in FP64 by the assembly when the Sab
# product runs in low precision (exact regrouping: X t2 = diag(X) t2 + offdiag(X) t2).
DIAG_SPLIT = True

# Form tau directly in BF16 when every product that reads it runs in BF16 (False: always FP64).
TAU16 = True

# The split form (shared split images) keeps S0, Sab, Sij, Spp and W in FP32 (False: FP64, for
# checks that isolate the images from the accumulators' rounding).
ACC32 = True

# The split form runs the hh ladder, the pp ladder and R56 on the hole pairs i < j (tau, Whh,
# S0 and R56 are antisymmetric in their hole pairs; False: all (i, j), for comparison checks).
HOLE_PAIRS = True


def _pairs(ham, split):
    """Whether this split-form step uses the hole-pair products."""
    return split is not None and ACC32 and HOLE_PAIRS and ham.h >= 2


# With the hole-pair products, t2's split image is carried over the hole pairs i < j too (half
# the full image): the ring operand is expanded from it with the antisymmetry's signs, and the
# A19 and Sab products read it directly (False: the full image, for comparison checks).
IMG_PAIRS = True


def _img_pairs(ham, split):
    """Whether this split-form step carries t2's split image over the hole pairs."""
    return _pairs(ham, split) and IMG_PAIRS


def _sab_pairs(ham, split):
    """Whether this step's split Sab product runs over the hole pairs (output columns (b, q))."""
    return _img_pairs(ham, split) and ham.p >= ops.LP_MIN_DIM


def t1_part(ham, t1, t2, B, prec=DEFAULT_PREC):
    """H1, X1pp, X1hh of t1Iter (coupled_cluster.py:372-408) into B.H1/B.X1pp/B.X1hh."""
    vpphh = ham.vpphh_view()
    H1, Xpp, Xhh = B.H1, B.X1pp, B.X1hh
    permute_into(ham.f_ph, H1)
    # U[b,k,i,j] = v_phph[b,k,c,i] t1[c,j] (A27) and X[i,d,k,l] = v_pphh[c,d,k,l] t1[c,i] (A34)
    # serve the T2 part; A11 is U's (k = j) trace and A14's s its X's (i = k) trace, so each of
    # the two blocks is read once per iteration
    contract("bkci,cj->bkij", ham.v_phph_bkic.permute(0, 1, 3, 2), t1, B.Ubkij, beta=0.0)  # A27
    contract("akij,kj->ai", B.Ubkij, ham.eye_h, H1, alpha=-1.0)              # A11
    contract("cikl,cakl->ai", ham.v_phhh, t2, H1, alpha=-0.5)                # A13
    contract("cdkl,ci->idkl", ham.v_pphh, t1, B.Xidkl, beta=0.0)             # A34
    # A14: s[d,l] = v[c,d,k,l] t1[c,k].  With v antisymmetric in (c,d) and in (k,l), the other
    # v_pphh.t1 products of both parts are the same matrix: u (A20, hh) = s, w (A20, pp) = -s,
    # z (A46) = s; and s[d,l] t2[d,a,l,i] = s[d,l] t2[a,d,i,l], so A12 and A14 are one product
    # with f_ph + s.
    contract("idkl,ik->dl", B.Xidkl, ham.eye_h, B.s_ph, beta=0.0)
    permute_into(B.s_ph, B.u_ph)
    permute_into(B.s_ph, B.w_ph, alpha=-1.0)
    permute_into(B.s_ph, B.z_ph)
    permute_into(ham.f_ph, B.F_ph)
    ops.axpby(1.0, B.s_ph, 1.0, B.F_ph)
    contract("ck,acik->ai", B.F_ph, t2, H1)                                  # A12 + A14
    pl = ham.ppph
    if pl.n:
        T0 = B.T0
        T0.zero_()
        _segsum(pl.ret0, ham.h, t2, T0, 1.0)                                 # ret0 partials
        contract("aki,k->ai", T0, ham.ones_h, H1, alpha=-0.5)               # ret0
        B.ret1.zero_()
        _segsum(pl.ret1, 1, t1, B.ret1, -1.0)                                # ret1
    permute_into(ham.f_hh, Xhh, alpha=-1.0)
    contract("ck,ci->ki", ham.f_ph, t1, Xhh, alpha=-0.5)                     # A15
    contract("bijk,bj->ki", ham.v_phhh, t1, Xhh, alpha=-1.0)                 # A17
    # A18 into Dhh: the same matrix is A42 of the T2 part (v_pphh antisymmetric in k,l and t2
    # in i,j make -1/2 v[c,d,k,l] t2[c,d,j,l] = -1/2 v[c,d,l,k] t2[c,d,l,j])
    contract("cdlk,cdli->ki", ham.v_pphh, t2, B.Dhh, alpha=-0.5, beta=0.0,
             precision=prec["dress"])
    ops.axpby(1.0, B.Dhh, 1.0, Xhh)
    contract("dk,di->ki", B.u_ph, t1, Xhh, alpha=-0.5)                      # A20 (hh)
    permute_into(ham.f_pp, Xpp)
    contract("ck,ak->ac", ham.f_ph, t1, Xpp, alpha=-0.5)                     # A16
    # A19 into Dpp: equal to A41 of the T2 part (antisymmetry of v_pphh in c,d and of t2 in a,b)
    if B.t2_pairs is not None:
        _dpp_pairs(ham, B)
    else:
        contract("dckl,adkl->ac", ham.v_pphh_T.permute(1, 0, 2, 3), t2, B.Dpp, alpha=0.5,
                 beta=0.0, precision=prec["dress"])     # t2[d,a,k,l] = -t2[a,d,k,l]
    ops.axpby(1.0, B.Dpp, 1.0, Xpp)
    contract("cl,al->ac", B.w_ph, t1, Xpp, alpha=0.5)                       # A20 (pp)
    if pl.n:
        ops.axpby(1.0, B.ret1, 1.0, Xpp)                                     # X_pp += ret1
    contract("ac,ci->ai", Xpp, t1, H1)
    contract("ki,ak->ai", Xhh, t1, H1)


SPLIT = (2, 0b10)       # [hi | lo] rows of a split operand image


def _v_dpp_image(ham, dt):
    """The exact 16-bit image of v_pphh[d,c,k,l] over the hole pairs k < l: rows c, columns
    (d, q(k,l)) (Pp slots per d, zero from P on), per-row scales that make a contact
    interaction's elements exact; built once per Hamiltonian and type."""
    hit = getattr(ham, "_v_dpp", None)
    if hit is not None and hit[0] == dt:
        return hit[1], hit[2]
    p, h = ham.p, ham.h
    Pp = _pair_dims(ham)[1]
    dev, st = ham.v_pphh.device, ops._stream()
    K = p * Pp
    vr = torch.zeros(p * K, dtype=F64, device=dev)                 # [(c,d)][q(k,l)]
    _lib.call("nl_hole_pairs", p * p, h, Pp, 1, ops._p(ham.v_pphh_T), ops._p(vr), 0, st)
    img = torch.empty(p * K, dtype=torch.int16, device=dev)
    iv = torch.empty(p, dtype=torch.float32, device=dev)
    _lib.call("nl_convert_rows", dt, 1, p, 1, K, (ctypes.c_int64 * 4)(0, K, 0, 1), ops._p(vr),
              ops._p(img), K, ops._p(iv), ctypes.c_float(ops.lp_target(dt, True)), st)
    del vr                          # (stream-ordered: its memory is reused after these kernels)
    ham._v_dpp = (dt, img, iv)
    return img, iv


def _dpp_pairs(ham, B):
    """A19 from the hole-pair image of t2: Dpp[a,c] = 1/2 sum_{d,k,l} v[d,c,k,l] t2[a,d,k,l]
    = sum_{d, k<l} v[d,c,k,l] t2[a,d,k,l] (both antisymmetric in (k, l)), as the split product
    t2_hi v + t2_lo v over K = (d, q): rows a of the pair image [hi | lo], rows c of v's exact
    pair image."""
    p = ham.p
    dt = B.t2_pairs_dt
    Pp = _pair_dims(ham)[1]
    K = p * Pp
    NR = B.t2_pairs
    vimg, iv = _v_dpp_image(ham, dt)
    splits = ops._lp_splits(p, p, K, 1)
    ws = ops.WS.get("splitk", splits * p * p, device=NR.device) if splits > 1 else None
    ops.count_flops("lp", 2 * 2.0 * p * p * K)
    _lib.call("nl_gemm_lp_split", dt, p, p, K, ctypes.c_double(1.0 / AMP_SCALE[dt]),
              ops._p(NR), ops._off16(NR, K), 2 * K, 0, None, 0, ops._p(vimg), None, K, 0,
              ops._p(iv), 0, ctypes.c_double(0.0), None, ops._p(B.Dpp), p, 0, 1, splits,
              ops._p(ws), ops._stream())

# Fixed scale of the shared FP16 split images of the amplitudes (t2, tau): 2^12 keeps
# |t| < 16 inside FP16's range and the elements down to ~1e-6 (far below the ones that set the
# sums) in normal numbers; a power of two, divided out exactly through alpha.  Larger amplitudes
# overflow to inf and the solver stops with its NaN report.
AMP_SCALE = {ops.BF16: 1.0, ops.FP16: 2.0 ** 12}


def _fast_split(ham, prec):
    """Every family on split operands of one type ("bf16x3" or "fp16x3"): t2's and tau's split
    images are made once per iteration (t2's by the previous step's assembly) and shared by the
    products instead of one conversion per product.  Returns the 16-bit type or None."""
    p, h = ham.p, ham.h
    kinds = {prec.get(k) for k in ("ring", "hh", "dress", "ladder")}
    if (len(kinds) != 1 or kinds.pop() not in ("bf16x3", "fp16x3")
            or any(k in prec for k in ("ring_W", "ring_S", "hh_W", "hh_S"))
            or p * p * h * h * 2 >= 2 ** 31 - 1):
        return None
    return ops.lp_kind(prec["ring"])[0]


def _lp_tau(ham, prec):
    """tau feeds only the hh products, the pp ladder and R56: when all of them run in BF16 (and
    are large enough not to fall back to FP64) it is formed directly in BF16."""
    p, h = ham.p, ham.h
    return (TAU16 and prec["hh"] == "bf16" and prec.get("ladder", "fp64") == "bf16"
            and min(p * p, h * h) >= ops.LP_MIN_DIM and (h * h) % 8 == 0)


def t2_part(ham, t1, t2, B, prec=DEFAULT_PREC):
    """S0/Sab/Sij/Spp and X2pp/X2hh of t2Iter (coupled_cluster.py:476-536).

    Requires B.ret1, B.u_ph, B.Dpp, B.Dhh, B.Ubkij and B.Xidkl from t1_part at the same
    amplitudes.  Uses that t2
    is antisymmetric in (a,b) and in (i,j) and that Sab / Sij enter the residual only through
    P(ab) / P(ij): X_pp[b,c] t2[a,c,i,j] and X_pp[a,c] t2[c,b,i,j] have the same P(ab) image, as
    have -q[l,i] t2[a,b,l,j] and -q[k,j] t2[a,b,i,k] under P(ij), so each group's two dressings
    run as one contraction with the summed matrix."""
    p, h = ham.p, ham.h
    vpphh = ham.vpphh_view()
    pl = ham.ppph
    # tau feeds only the hh products, the pp ladder and R56: when all of them run in BF16 (and
    # are large enough not to fall back to FP64) it is formed directly in BF16, as [ab][ij] rows
    # and, for the Whh product (summed over c,d), as [ij][cd] rows; the FP64 tau is not written
    tau = B.tau
    split = _fast_split(ham, prec)
    pairs = _pairs(ham, split)
    lp_tau = (_lp_tau(ham, prec) or split is not None) and not pairs
    tau16 = None
    if pairs:
        pass                        # tau is formed over the hole pairs by _hh_pairs
    elif split is not None:
        # tau only as its split image: [ab][hi(ij) | lo(ij)] rows, and [ij][hi(cd) | lo(cd)]
        # rows (the Whh product sums over c, d) by two 16-bit transposes
        n, hh, pp = p * p * h * h, h * h, p * p
        tau16 = ops.WS.get("tau_split", 2 * n, dtype=torch.int16, device=t2.device)
        tauT = ops.WS.get("tau_splitT", 2 * n, dtype=torch.int16, device=t2.device)
        st = ops._stream()
        _lib.call("nl_tau_split", p, h, ops._p(t1), ops._p(t2), None, ops._p(tau16), split,
                  ctypes.c_double(AMP_SCALE[split]), st)
        for q in range(2):
            _lib.call("nl_transpose16_into", 1, pp, hh, ops._off16(tau16, q * hh), 2 * hh, 0,
                      ops._off16(tauT, q * pp), 2 * pp, 0, pp, st)
    elif lp_tau:
        ldT = (p * p + 7) // 8 * 8                   # [ij][cd] rows, zero-padded to 8
        tau16 = ops.WS.get("tau16", p * p * h * h, dtype=torch.int16, device=t2.device)
        tauT = ops.WS.get("tau16T", h * h * ldT, dtype=torch.int16, device=t2.device)
        _lib.call("nl_tau16", p, h, ops._p(t1), ops._p(t2), ops._p(tau16), ops._stream())
        _lib.call("nl_transpose16", p * p, h * h, ops._p(tau16), h * h, ops._p(tauT), ldT,
                  ops._stream())
    else:
        _lib.call("nl_tau", p, h, ops._p(t1), ops._p(t2), ops._p(tau), ops._stream())
    # ---- dressed Fock matrices
    Xpp, Xhh = B.X2pp, B.X2hh
    permute_into(ham.f_pp, Xpp)
    ops.axpby(1.0, B.Dpp, 1.0, Xpp)                                          # A41 (= A19)
    contract("ck,bk->bc", ham.f_ph, t1, Xpp, alpha=-1.0)                     # A43
    contract("cl,bl->bc", B.z_ph, t1, Xpp, alpha=-1.0)                      # A46
    permute_into(ham.f_hh, Xhh, alpha=-1.0)
    ops.axpby(1.0, B.Dhh, 1.0, Xhh)                                          # A42 (= A18)
    contract("ck,cj->kj", ham.f_ph, t1, Xhh, alpha=-1.0)                     # A44
    contract("dk,dj->kj", B.u_ph, t1, Xhh, alpha=-1.0)                      # A45
    # ---- S0: v_pphh + 1/2 tau.Whh + 1/2 pp ladder(tau).  The split form accumulates S0, Sab,
    # Sij, Spp and W in FP32 and leaves v_pphh out of S0: the assembly adds it in FP64
    acc32 = split is not None and ACC32
    acc = B.acc32() if acc32 else B
    S0 = acc.S0
    permute_into(ham.v_hhhh, B.Whh)
    contract("ci,cjkl->klij", t1, ham.v_phhh, B.zz, beta=0.0)
    permute_into(B.zz, B.Whh, 1.0, 1.0)
    permute_into(B.zz.permute(0, 1, 3, 2), B.Whh, -1.0, 1.0)
    Sab = acc.Sab
    if pairs:
        _hh_pairs(ham, t1, t2, B, split, prec)
    with (ops.lp_reuse(tau, absent=True) if lp_tau else contextlib.nullcontext()) \
            if not pairs else contextlib.nullcontext():
        if pairs:
            pass
        elif split is not None:
            sc = AMP_SCALE[split]
            ops.lp_provide(tau, "cdij", ["i", "j"], ["c", "d"], tauT, 2 * p * p, split, SPLIT, sc)
            ops.lp_provide(tau, "abkl", ["a", "b"], ["k", "l"], tau16, 2 * h * h, split, SPLIT,
                           sc)
        elif lp_tau:
            ops.lp_provide(tau, "cdij", ["i", "j"], ["c", "d"], tauT, ldT)
            ops.lp_provide(tau, "abkl", ["a", "b"], ["k", "l"], tau16, h * h)
        if not pairs:
            contract("cdkl,cdij->klij", ham.v_pphh, tau, B.Whh, alpha=0.5,
                     precision=prec.get("hh_W", prec["hh"]))
            if acc32:
                contract("abkl,klij->abij", tau, B.Whh, S0, alpha=0.5, beta=0.0,
                         precision=prec.get("hh_S", prec["hh"]))
            else:
                contract("abkl,klij->abij", tau, B.Whh, S0, alpha=0.5,
                         precision=prec.get("hh_S", prec["hh"]), addend=ham.v_pphh)
            _pp_ladder(ham, tau, S0, prec, tau16)
            # ---- Sab: P(ab) group
            permute_into(ham.v_phhh, B.N56)
            if pl.n:
                _block_gemm(pl.r56, tau, B.N56, 0.5, h * h, "r56", prec.get("ladder", "fp64"),
                            tau16)
    permute_into(Xpp, B.Mpp)                                                 # X_pp t2 (P(ab) image)
    if pl.n:
        ops.axpby(1.0, B.ret1, 1.0, B.Mpp)                                   # + A38
    # Mpp's diagonal (the Fock diagonal: the largest term of the residual, the one the Jacobi
    # denominator divides out) leaves a low-precision Sab product: the assembly adds
    # (Mpp[a,a] + Mpp[b,b]) t2[a,b,i,j] in FP64, the P(ab) image of the diagonal part
    B.dpp = None
    if prec["dress"] != "fp64" and DIAG_SPLIT:
        _lib.call("nl_take_diag", p, ops._p(B.Mpp), ops._p(B.dMpp), ops._stream())
        B.dpp = B.dMpp
    B.sab_pairs = False
    if split is not None and p >= ops.LP_MIN_DIM:
        _sab_split(ham, t1, B, Sab, split)
        B.sab_pairs = B.t2_pairs is not None
    elif getattr(B, "t2_16", None) is not None and p >= ops.LP_MIN_DIM:
        _sab_bf16(ham, t1, B, Sab)
    else:
        contract("bkij,ak->abij", B.N56, t1, Sab, beta=0.0)                  # A23 + A39
        contract("ad,dbij->abij", B.Mpp, t2, Sab, precision=prec["dress"])
    # ---- Sij: P(ij) group, [a][b][i][j]
    Sij = acc.Sij
    contract("cikl,ck->li", ham.v_phhh, t1, B.q_hh, beta=0.0)                # A28
    _lib.call("nl_sub", h * h, ops._p(Xhh), ops._p(B.q_hh), ops._p(B.Mhh), ops._stream())
    contract("kj,abik->abij", B.Mhh, t2, Sij, beta=0.0)                      # X_hh t2 - A28 t2
    if pl.n:
        _segsum(pl.ret2p, h, t1, Sij, 1.0)                                   # A36
    # ---- Spp: P(ij)P(ab) group in [(a,i)][(b,j)] layout
    W = acc.W
    # t2[d,b,l,j] = t2[b,d,j,l]: ring_W reads t2 in the ring's own [(b,j)][(d,l)] rows, so the
    # two ring GEMMs share one 16-bit conversion of it
    # the FP64 ring GEMMs read t2 as [(a,i)][(c,k)] rows of a permuted copy (long contiguous K
    # runs); the BF16 ones read t2's 16-bit image, so there t2 itself is the operand
    t2ring = t2
    if prec["ring"] == "fp64":
        Tr = ops.WS.get("Tr", p * p * h * h, device=t2.device).view(p, h, p, h)
        permute_into(t2.permute(0, 2, 1, 3), Tr)
        t2ring = Tr.permute(0, 2, 1, 3)
    with ops.lp_reuse(t2ring):
        if not _ring_w_anchored(ham, B, W):
            contract("cdkl,bdjl->ckbj", vpphh, t2ring, W, alpha=0.5,
                     precision=prec.get("ring_W", prec["ring"]), tag="ring_W",
                     addend=ham.Wc)                                          # A24 + A22
            _ring_w_anchor(ham, B, W)
        if pl.n:
            _segsum(pl.ret4, h, t1, W, 1.0)                                  # A37
        contract("cjkl,bl->ckbj", ham.v_phhh, t1, W, alpha=-1.0)             # A29
        if not _ring_s_anchored(ham, B, W, acc.Spp):
            contract("acik,ckbj->aibj", t2ring, W, acc.Spp, beta=0.0,
                     precision=prec.get("ring_S", prec["ring"]), tag="ring")
            _ring_s_anchor(ham, B, W, acc.Spp)
    if split is not None:
        # t2[a,d,k,j] = -t2[a,d,j,k]: the Y product reads the ring image (rows (a,j), columns
        # (d,k)) with the sign in alpha, so it needs no image of t2 of its own
        contract("idkl,adjk->ilaj", B.Xidkl, t2, B.Yilaj, alpha=-1.0, beta=0.0,
                 precision=prec["hh"])                                                 # A34
    else:
        contract("idkl,adkj->ilaj", B.Xidkl, t2, B.Yilaj, beta=0.0, precision=prec["hh"])
    # Spp enters only through P(ij)P(ab), which is invariant under (a,i) <-> (b,j): the A27 term
    # -t1[a,k] U[b,k,i,j] may stand as -U[a,l,j,i] t1[b,l], so both tails are one product
    permute_into(B.Yilaj, B.Zilaj)
    permute_into(B.Ubkij.permute(3, 1, 0, 2), B.Zilaj, alpha=-1.0, beta=1.0)
    contract("ilaj,bl->aibj", B.Zilaj, t1, acc.Spp)


# Anchored ring products (fp16x3 form): every RING_ANCHOR-th step runs the split products
# W_A = Wc + 1/2 v t2_A (two partial products) and Spp_A = t2_A W_A (three) and keeps them with
# t2_A's ring image, W_A and W_A's FP16 image; the steps in between form W = W_A + 1/2 v (t2 -
# t2_A) (one partial product on FP16(t2 - t2_A)) and Spp = Spp_A + t2 (W - W_A) + (t2 - t2_A) W_A
# (one single-rounded product over the concatenated K: [t2_hi | FP16(t2 - t2_A)] against
# [FP16(W - W_A) | FP16(W_A)], both W halves on W_A's per-row scales).  Each rounding is 2^-11 of
# a difference from the anchor, which shrinks to nothing as the amplitudes converge, so the fixed
# point is the split products'.  0: off.
RING_ANCHOR = 8
_KSTEP = 64                     # the 256x256 16-bit kernel's K step (a zero-padded K)
# The W product between anchors on FP8 (E4M3) operands: v's and (t2 - t2_A)'s rows on per-row
# power-of-two scales.  Its roundings (2^-4) multiply t2 - t2_A, so they too vanish at
# convergence and leave the fixed point where it is (L=5 A=28, seeds 0-4: |dE|/|E| <= 3.8e-8 in
# the same step counts).  Off by default: the FP8 kernel (16x16x128 MFMAs fed from the FP16
# kernel's LDS tiles) takes 5.37 ms on that product against 4.27 ms for FP16.
RING_DELTA_FP8 = False
_KSTEP8 = 128                   # the FP8 kernel's K step
_FP8_MAXK = 16384               # longest row the FP8 conversion stages in LDS


def _ring_anchor_on(ham, B, W):
    ring = getattr(B, "t2_ring", None)
    return (RING_ANCHOR > 0 and ring is not None and ring[1] == ops.FP16
            and W.dtype == torch.float32 and W.is_contiguous()
            and ham.p * ham.h >= 256)


def _ring_kp(ham):
    ph = ham.p * ham.h
    return ph, (ph + _KSTEP - 1) // _KSTEP * _KSTEP


def _v_ring_image(ham, dt, kp):
    """v_pphh as [(c,k)][(d,l)] rows (ham.vP), its exact 16-bit image with rows of kp slots
    (zero from p*h on) and per-row inverse scales; built once per Hamiltonian and type."""
    hit = getattr(ham, "_v_ring", None)
    if hit is not None and hit[0] == (dt, kp):
        return hit[1], hit[2]
    ph = ham.p * ham.h
    dev, st = ham.vP.device, ops._stream()
    img = torch.zeros(ph * kp, dtype=torch.int16, device=dev)
    iv = torch.empty(ph, dtype=torch.float32, device=dev)
    _lib.call("nl_convert_rows", dt, 1, ph, 1, ph, (ctypes.c_int64 * 4)(0, ph, 0, 1),
              ops._p(ham.vP), ops._p(img), kp, ops._p(iv),
              ctypes.c_float(ops.lp_target(dt, True)), st)
    ham._v_ring = ((dt, kp), img, iv)
    return img, iv


def _v_ring_image8(ham, kp8):
    """v_pphh as [(c,k)][(d,l)] rows in FP8 (E4M3, per-row power-of-two scales), rows of kp8
    bytes (zero from p*h on); built once per Hamiltonian."""
    hit = getattr(ham, "_v_ring8", None)
    if hit is not None and hit[0] == kp8:
        return hit[1], hit[2]
    ph = ham.p * ham.h
    dev, st = ham.vP.device, ops._stream()
    img = torch.empty(ph * kp8, dtype=torch.uint8, device=dev)
    iv = torch.empty(ph, dtype=torch.float32, device=dev)
    _lib.call("nl_convert_rows8", ph, ph, ph, ops._p(ham.vP), ops._p(img), kp8, kp8,
              ops._p(iv), st)
    ham._v_ring8 = (kp8, img, iv)
    return img, iv


def _ring_w_anchor(ham, B, W):
    """After a full W product: keep W (Wc + 1/2 v t2, FP32) and t2's ring image as the anchor."""
    if not _ring_anchor_on(ham, B, W):
        B.ring_step = 0
        B.ring_mode = None
        return
    R, _ = B.t2_ring
    ph = ham.p * ham.h
    WA = ops.WS.get("ring_WA", ph * ph, dtype=torch.float32, device=W.device)
    RA = ops.WS.get("ring_RA", 2 * ph * ph, dtype=torch.int16, device=W.device)
    WA.copy_(W.view(-1))
    RA.copy_(R[:2 * ph * ph])
    B.ring_step = 1
    B.ring_mode = "anchor"


def _ring_w_anchored(ham, B, W):
    """W = W_A + 1/2 v (t2 - t2_A) on the anchor's steps in between; False when this step takes
    the full product (the first step, every RING_ANCHOR-th step, other forms).  Also writes the
    Spp delta product's left operand [t2_hi | FP16(t2 - t2_A)] (rows (a,i), 2 kp columns), whose
    second half is this product's right operand."""
    step = getattr(B, "ring_step", 0)
    B.ring_mode = None
    if not _ring_anchor_on(ham, B, W) or step == 0 or step >= RING_ANCHOR:
        return False
    R, dt = B.t2_ring
    ph, kp = _ring_kp(ham)
    dev, st = W.device, ops._stream()
    WA = ops.WS.get("ring_WA", ph * ph, dtype=torch.float32, device=dev)
    RA = ops.WS.get("ring_RA", 2 * ph * ph, dtype=torch.int16, device=dev)
    TD = ops.WS.get("ring_TD", ph * 2 * kp, dtype=torch.int16, device=dev)
    _lib.call("nl_split_delta16", ph, ph, 2 * ph, ops._p(R), ops._p(RA), ops._p(TD), 2 * kp, kp,
              1, st)
    io = (ops.C_F32 | ops.D_F32) << 8
    fp8 = RING_DELTA_FP8 and ph <= _FP8_MAXK
    if fp8:
        kp8 = (ph + _KSTEP8 - 1) // _KSTEP8 * _KSTEP8
        D8 = ops.WS.get("ring_dT8", ph * kp8, dtype=torch.uint8, device=dev)
        i8 = ops.WS.get("ring_dT8_s", ph, dtype=torch.float32, device=dev)
        _lib.call("nl_split_delta8", ph, ph, 2 * ph, ops._p(R), ops._p(RA), ops._p(D8), kp8, kp8,
                  ops._p(i8), st)
        vimg, iv = _v_ring_image8(ham, kp8)
        ops.count_flops("lp", 2.0 * ph * ph * kp8)
    else:
        vimg, iv = _v_ring_image(ham, dt, kp)
        ops.count_flops("lp", 2.0 * ph * ph * kp)
    ev = ops._timed("ring_W")
    if ev:
        ev[0].record()
    if fp8:
        # the FP8 kernel takes 256x256 tiles only: below 256 of them, split K to fill the GPU
        tiles = ((ph + 255) // 256) ** 2
        splits = 1 if tiles >= 256 else -(-256 // tiles)
        ws = (ops.WS.get("ring_ws8", splits * ph * ph, dtype=torch.float64, device=dev)
              if splits > 1 else None)
        _lib.call("nl_gemm_lp", 2 | io, ph, ph, kp8, ctypes.c_double(0.5 / AMP_SCALE[dt]),
                  ops._p(vimg), kp8, 0, ops._p(iv), 0, ops._p(D8), kp8, 0, ops._p(i8), 0,
                  ctypes.c_double(1.0), ops._p(WA), ops._p(W), ph, 0, 1, splits, ops._p(ws), st)
    else:
        _lib.call("nl_gemm_lp", dt | io, ph, ph, kp, ctypes.c_double(0.5 / AMP_SCALE[dt]),
                  ops._p(vimg), kp, 0, ops._p(iv), 0, ops._off16(TD, kp), 2 * kp, 0, None, 0,
                  ctypes.c_double(1.0), ops._p(WA), ops._p(W), ph, 0, 1, 1, None, st)
    if ev:
        ev[1].record()
    B.ring_step = step + 1
    B.ring_mode = "delta"
    return True


# W as rows (b,j) of columns (c,k): the [c][k][b][j] FP32 buffer seen transposed
def _w_rows(ham):
    p, h = ham.p, ham.h
    return p, h, p, h, (ctypes.c_int64 * 4)(h, 1, p * h * h, p * h)


def _ring_s_anchor(ham, B, W, Spp):
    """After a full Spp product on an anchor step: keep Spp (before its T1 tails), W and W's
    single-rounded FP16 image (rows (b,j), K padded to the product's step, per-row scales) as the
    second half of the delta product's right operand."""
    if getattr(B, "ring_mode", None) != "anchor" or Spp.dtype != torch.float32:
        B.ring_s_ok = False
        return
    ph, kp = _ring_kp(ham)
    dev, st = W.device, ops._stream()
    SA = ops.WS.get("ring_SA", ph * ph, dtype=torch.float32, device=dev)
    WF = ops.WS.get("ring_WF", ph * ph, dtype=torch.float32, device=dev)
    WB = ops.WS.get("ring_WB", ph * 2 * kp, dtype=torch.int16, device=dev)
    iw = ops.WS.get("ring_WB_s", ph, dtype=torch.float32, device=dev)
    SA.copy_(Spp.view(-1))
    WF.copy_(W.view(-1))
    R1, R2, C1, C2, s4 = _w_rows(ham)
    _lib.call("nl_convert_split", ops.FP16 | (ops.D_F32 << 8), R1, R2, C1, C2, s4, ops._p(W),
              ops._off16(WB, kp), 2 * kp, kp, 1, 0, ops._p(iw),
              ctypes.c_float(ops.lp_target(ops.FP16, False)), ctypes.c_double(1.0), st)
    B.ring_s_ok = True


def _ring_s_anchored(ham, B, W, Spp):
    """Spp = Spp_A + t2 (W - W_A) + (t2 - t2_A) W_A on the steps between anchors, as one
    single-rounded product over the concatenated K, [t2_hi | FP16(t2 - t2_A)] against
    [FP16(W - W_A) | FP16(W_A)] (the W halves on W_A's per-row scales: W - W_A is small against
    W_A, so the shared scale cannot overflow and only narrows a rounding already far below
    W_A's): the roundings scale with the differences from the anchor.  False when this step
    takes the full product."""
    if getattr(B, "ring_mode", None) != "delta" or not getattr(B, "ring_s_ok", False):
        return False
    if Spp.dtype != torch.float32 or not Spp.is_contiguous():
        return False
    _, dt = B.t2_ring
    ph, kp = _ring_kp(ham)
    dev, st = W.device, ops._stream()
    SA = ops.WS.get("ring_SA", ph * ph, dtype=torch.float32, device=dev)
    WF = ops.WS.get("ring_WF", ph * ph, dtype=torch.float32, device=dev)
    WB = ops.WS.get("ring_WB", ph * 2 * kp, dtype=torch.int16, device=dev)
    iw = ops.WS.get("ring_WB_s", ph, dtype=torch.float32, device=dev)
    TD = ops.WS.get("ring_TD", ph * 2 * kp, dtype=torch.int16, device=dev)
    R1, R2, C1, C2, s4 = _w_rows(ham)
    _lib.call("nl_convert_split_delta", dt | (ops.D_F32 << 8), R1, R2, C1, C2, s4, ops._p(W),
              ops._p(WF), ops._p(WB), 2 * kp, kp, 1, 0, None, ops._p(iw),
              ctypes.c_float(ops.lp_target(dt, False)), ctypes.c_double(1.0), st)
    io = (ops.C_F32 | ops.D_F32) << 8
    ops.count_flops("lp", 2.0 * ph * ph * 2 * kp)
    ev = ops._timed("ring_delta")
    if ev:
        ev[0].record()
    _lib.call("nl_gemm_lp", dt | io, ph, ph, 2 * kp, ctypes.c_double(1.0 / AMP_SCALE[dt]),
              ops._p(TD), 2 * kp, 0, None, 0, ops._p(WB), 2 * kp, 0, ops._p(iw), 0,
              ctypes.c_double(1.0), ops._p(SA), ops._p(Spp), ph, 0, 1, 1, None, st)
    if ev:
        ev[1].record()
    return True


def _sab_bf16(ham, t1, B, Sab):
    """Sab[a,(b,i,j)] = sum_k t1[a,k] N56[b,k,i,j] + sum_d Mpp[a,d] t2[d,b,i,j] as ONE BF16 GEMM
    over K = (k | d): A = [t1 | Mpp] rows a, B = [N56 | t2^T] rows (b,i,j) (t2^T from the
    iteration's BF16 image of t2), written once with beta = 0 instead of a low-K FP64 pass that
    writes Sab and a BF16 GEMM that reads and rewrites it."""
    p, h = ham.p, ham.h
    ldk = (h + p + 7) // 8 * 8
    dev, st = t1.device, ops._stream()
    A16 = ops.WS.get("sab_A16", p * ldk, dtype=torch.int16, device=dev)
    B16 = ops.WS.get("sab_B16", p * h * h * ldk, dtype=torch.int16, device=dev)
    i4 = lambda *v: (ctypes.c_int64 * 4)(*v)
    off = lambda t, n: ctypes.c_void_p(t.data_ptr() + 2 * n)       # 16-bit element offset
    _lib.call("nl_convert_rows_into", 1, p, 1, h, i4(0, h, 0, 1), ops._p(t1), ops._p(A16), ldk,
              ldk, st)                                              # [t1 | 0 ...]
    _lib.call("nl_convert_rows_into", 1, p, 1, p, i4(0, p, 0, 1), ops._p(B.Mpp), off(A16, h),
              ldk, p, st)                                           # [t1 | Mpp | 0]
    _lib.call("nl_convert_rows_into", p, h * h, 1, h, i4(h ** 3, 1, 0, h * h), ops._p(B.N56),
              ops._p(B16), ldk, ldk, st)                            # [N56 | 0 ...]
    _lib.call("nl_transpose16_into", 1, p, p * h * h, ops._p(B.t2_16), p * h * h, 0,
              off(B16, h), ldk, 0, p, st)                           # [N56 | t2^T | 0]
    splits = ops._lp_splits(p, p * h * h, ldk, 1)
    ws = ops.WS.get("splitk", splits * p * p * h * h, device=dev) if splits > 1 else None
    ops.count_flops("lp", 2.0 * p * p * h * h * ldk)
    _lib.call("nl_gemm_lp", ops.BF16, p, p * h * h, ldk, ctypes.c_double(1.0), ops._p(A16), ldk,
              0, None, 0, ops._p(B16), ldk, 0, None, 0, ctypes.c_double(0.0), None, ops._p(Sab),
              p * h * h, 0, 1, splits, ops._p(ws), st)


def _pair_dims(ham):
    """P hole pairs i < j, their slots per row Pp (P rounded up to 32), and the (c, d) pair
    count rounded up to 32 (the K extent of the Whh product)."""
    p, h = ham.p, ham.h
    P = h * (h - 1) // 2
    return P, (P + 31) // 32 * 32, (p * p + 31) // 32 * 32


def _v_pairs_image(ham, dt):
    """The exact 16-bit image of v_pphh[c,d,k,l] over the hole pairs k < l: rows q(k,l), columns
    (c, d) (zero beyond p^2), per-row scales that make a contact interaction's elements exact;
    built once per Hamiltonian and type."""
    hit = getattr(ham, "_v_pairs", None)
    if hit is not None and hit[0] == dt:
        return hit[1], hit[2]
    p, h = ham.p, ham.h
    P, Pp, Kc = _pair_dims(ham)
    dev, st = ham.v_pphh.device, ops._stream()
    vr = torch.empty(p * p * Pp, dtype=F64, device=dev)            # [(c,d)][q(k,l)]
    _lib.call("nl_hole_pairs", p * p, h, Pp, 1, ops._p(ham.v_pphh), ops._p(vr), 0, st)
    img = torch.empty(P * Kc, dtype=torch.int16, device=dev)
    iv = torch.empty(P, dtype=torch.float32, device=dev)
    _lib.call("nl_convert_rows", dt, 1, P, 1, p * p, (ctypes.c_int64 * 4)(0, 1, 0, Pp),
              ops._p(vr), ops._p(img), Kc, ops._p(iv), ctypes.c_float(ops.lp_target(dt, True)),
              st)
    del vr                          # (stream-ordered: its memory is reused after these kernels)
    ham._v_pairs = (dt, img, iv)
    return img, iv


def _hh_pairs(ham, t1, t2, B, dt, prec):
    """The hh ladder, the pp ladder and R56 of the split form on the hole pairs i < j.  tau, Whh,
    S0 and R56 are antisymmetric in their hole pairs (ij) and (kl), so the products need only
    the P = h(h-1)/2 pairs i < j as output columns, and the hh ladder's sum over (k, l) is twice
    its sum over k < l:
        Whh[kl,ij] (k<l, i<j) = its v_hhhh and t1 terms + 1/2 sum_cd v[c,d,k,l] tau[c,d,i,j]
        S0[ab,ij]  (i<j)      = sum_{k<l} tau[a,b,k,l] Whh[kl,ij] + 1/2 sum_cd V[ab,cd] tau[cd,ij]
        N56[bk,ij] (i<j)      = v_phhh + 1/2 sum_cd v_ppph[c,d,b,k] tau[c,d,i,j]
    S0 stays in the pair form (the assembly expands it); N56 too when the Sab product runs over
    the pairs (B.N56r), else it is expanded for that product."""
    p, h = ham.p, ham.h
    pp = p * p
    P, Pp, Kc = _pair_dims(ham)
    dev, st = t2.device, ops._stream()
    sc = AMP_SCALE[dt]
    i16 = torch.int16
    off = ops._off16
    # tau over the pairs: [ab][hi(Pp) | lo(Pp)], and [q(i,j)][hi(Kc) | lo(Kc)] (summed over cd)
    tauR = ops.WS.get("tau_pairs", pp * 2 * Pp, dtype=i16, device=dev)
    _lib.call("nl_tau_split_pairs", p, h, ops._p(t1), ops._p(t2), ops._p(tauR), Pp, 1, dt,
              ctypes.c_double(sc), st)
    tauRT = ops.WS.get("tau_pairsT", P * 2 * Kc, dtype=i16, device=dev)
    for q in range(2):
        _lib.call("nl_transpose16_into", 1, pp, P, off(tauR, q * Pp), 2 * Pp, 0,
                  off(tauRT, q * Kc), 2 * Kc, 0, Kc, st)
    # Whh over the pairs, [q(k,l)][q(i,j)] at pitch Pp: B.Whh's v_hhhh and t1 terms reduced in
    # both pairs, then + 1/2 v tau (exact v rows x split tau rows, split-K)
    Wt = ops.WS.get("whh_pairs_t", h * h * Pp, device=dev)
    _lib.call("nl_hole_pairs", h * h, h, Pp, 1, ops._p(B.Whh), ops._p(Wt), 0, st)
    Wr = ops.WS.get("whh_pairs", P * Pp, device=dev)
    _lib.call("nl_hole_pairs", 1, h, P, Pp, ops._p(Wt), ops._p(Wr), 0, st)
    vimg, iv = _v_pairs_image(ham, dt)
    splits = ops._lp_splits(P, P, Kc, 1)
    ws = ops.WS.get("splitk", splits * P * P, device=dev) if splits > 1 else None
    ops.count_flops("lp", 2 * 2.0 * P * P * Kc)
    _lib.call("nl_gemm_lp_split", dt, P, P, Kc, ctypes.c_double(0.5 / sc), ops._p(vimg), None,
              Kc, 0, ops._p(iv), 0, ops._p(tauRT), off(tauRT, Kc), 2 * Kc, 0, None, 0,
              ctypes.c_double(1.0), ops._p(Wr), ops._p(Wr), Pp, 0, 1, splits, ops._p(ws), st)
    # S0 over the pairs = sum_{k<l} tau[ab,kl] Whh[kl,ij]: Whh^T rows q(i,j) split with per-row
    # scales (rows from P on stay zero), FP32 output
    WT, iw = B.pair_operand(Pp)
    target = ops.lp_target(dt, False)
    _lib.call("nl_convert_split", dt, 1, P, 1, P, (ctypes.c_int64 * 4)(0, 1, 0, Pp), ops._p(Wr),
              ops._p(WT), 2 * Pp, Pp, 2, 0b10, ops._p(iw) if target != 0.0 else None,
              ctypes.c_float(target), ctypes.c_double(1.0), st)
    S0 = B.s0_pairs(Pp)
    ops.count_flops("lp", 3 * 2.0 * pp * P * Pp)
    ev = ops._timed("hh_S")
    if ev:
        ev[0].record()
    _lib.call("nl_gemm_lp_split", dt | (ops.C_F32 << 8), pp, P, Pp, ctypes.c_double(1.0 / sc),
              ops._p(tauR), off(tauR, Pp), 2 * Pp, 0, None, 0, ops._p(WT), off(WT, Pp), 2 * Pp,
              0, ops._p(iw) if target != 0.0 else None, 0, ctypes.c_double(0.0), None,
              ops._p(S0), Pp, 0, 1, 1, None, st)
    if ev:
        ev[1].record()
    # pp ladder into S0's pair rows (columns from P on receive zeros)
    lad = prec.get("ladder", "fp64")
    _block_gemm(ham.pppp, t2, S0, 0.5, Pp, "ladder", lad, tauR)
    # R56 over the pairs: N56 = v_phhh + 1/2 R56 as [(b,k)][q(i,j)], expanded to [b][k][i][j]
    # unless the Sab product reads the pairs
    N56r = B.n56_pairs(Pp)
    _lib.call("nl_hole_pairs", p * h, h, Pp, 1, ops._p(ham.v_phhh), ops._p(N56r), 0, st)
    if ham.ppph.n:
        _block_gemm(ham.ppph.r56, t2, N56r, 0.5, Pp, "r56", lad, tauR)
    if _sab_pairs(ham, dt):
        B.N56r = N56r
    else:
        _lib.call("nl_hole_pairs", p * h, h, Pp, 1, ops._p(N56r), ops._p(B.N56), 1, st)


def _sab_split(ham, t1, B, Sab, dt):
    """Sab[a,(b,i,j)] = sum_k t1[a,k] N56[b,k,i,j] + sum_d Mpp[a,d] t2[d,b,i,j] as ONE split GEMM
    over K = (k | d) (kp = h + p padded to 32): A = [t1 | Mpp] rows a as [hi(kp) | lo(kp)] with
    per-row scales, B = [N56 | t2^T] rows (b,i,j) at the amplitude images' fixed scale; the
    t2^T columns are 16-bit transposes of this step's split image of t2 (B.t2_split).  Sab is
    written once (beta = 0) instead of by a low-K FP64 pass and a product that reads it back.
    Over the hole pairs (B.t2_pairs set): Sab is antisymmetric in (i, j) as N56 and t2 are, so
    the product runs on the rows (b, q(i<j)) of N56's and t2's pair forms and writes Sab as
    [a][(b, q)] (Pp slots per b, the assembly's NL_SAB_PAIRS layout)."""
    p, h = ham.p, ham.h
    kp = (h + p + 31) // 32 * 32
    ld = 2 * kp
    dev, st = t1.device, ops._stream()
    sc = AMP_SCALE[dt]
    i4 = lambda *v: (ctypes.c_int64 * 4)(*v)
    if B.t2_pairs is not None:
        Pp = _pair_dims(ham)[1]
        row = p * Pp
        img = B.sab_image(kp, row)
        # N56 over the pairs [(b,k)][q] as rows (b,q), columns k
        n56, n56s = B.N56r, i4(h * Pp, 1, 0, Pp)
        R2, N = Pp, B.t2_pairs      # [d][hi(b,q) | lo(b,q)]
    else:
        row = p * h * h
        img = B.sab_image(kp)
        # N56[b,k,i,j] as rows (b,i,j), columns k
        n56, n56s = B.N56, i4(h ** 3, 1, 0, h * h)
        R2, N = h * h, B.t2_split   # [d][hi(b,i,j) | lo(b,i,j)]
    # (img is zero-initialised: the columns [h + p, kp) stay zero)
    # N56: hi into columns [0, h), lo into [kp, kp + h)
    for off, mask in ((0, 0), (kp, 1)):
        _lib.call("nl_convert_split", dt, p, R2, 1, h, n56s, ops._p(n56), ops._off16(img, off),
                  ld, h, 1, mask, None, ctypes.c_float(0.0), ctypes.c_double(sc), st)
    for q in range(2):
        _lib.call("nl_transpose16_into", 1, p, row, ops._off16(N, q * row), 2 * row, 0,
                  ops._off16(img, q * kp + h), ld, 0, p, st)
    A64 = ops.WS.get("sab_A64", p * (h + p), device=dev)
    _lib.call("nl_concat_rows", p, h, ops._p(t1), h, p, ops._p(B.Mpp), p, ops._p(A64), h + p, st)
    A16 = ops.WS.get("sab_A16", p * ld, dtype=torch.int16, device=dev)
    target = ops.lp_target(dt, False)
    ia = ops.WS.get("sab_ia", p, dtype=torch.float32, device=dev) if target != 0.0 else None
    _lib.call("nl_convert_split", dt, 1, p, 1, h + p, i4(0, h + p, 0, 1), ops._p(A64),
              ops._p(A16), ld, kp, 2, 0b10, ops._p(ia), ctypes.c_float(target),
              ctypes.c_double(1.0), st)
    io = ops.C_F32 if Sab.dtype == torch.float32 else 0
    ops.count_flops("lp", 3 * 2.0 * p * row * kp)
    ev = ops._timed("sab")
    if ev:
        ev[0].record()
    _lib.call("nl_gemm_lp_split", dt | (io << 8), p, row, kp, ctypes.c_double(1.0 / sc),
              ops._p(A16), ops._off16(A16, kp), ld, 0, ops._p(ia), 0, ops._p(img),
              ops._off16(img, kp), ld, 0, None, 0, ctypes.c_double(0.0), None, ops._p(Sab), row,
              0, 1, 1, None, st)
    if ev:
        ev[1].record()


def _pp_ladder(ham, tau, S0, prec, tau16=None):
    """S0 += 1/2 sum_cd v_pppp[a,b,c,d] tau[c,d,i,j] over packed pair blocks."""
    _block_gemm(ham.pppp, tau, S0, 0.5, ham.h * ham.h, "ladder", prec.get("ladder", "fp64"),
                tau16)


def _block_gemm(pb, src, out, alpha, ncol, tag, precision="fp64", src16=None):
    """out[rows_b] += alpha * V_b src[cols_b] for every packed block b (row gather, batched MFMA
    GEMM, row scatter; rows of src/out have ncol contiguous elements).  With a low precision
    the blocks V_b (constant, converted once with exact per-row scales) and the gathered rows
    (transposed per block) are 16-bit and the batched GEMM accumulates in FP32: BF16 gathers and
    casts in one pass (or, given src16, the BF16 rows of src, gathers them) and adds the block
    rows straight into out from the GEMM; FP16 (per-row scales) and the split forms go through
    FP64 gathered rows and block results (split: [V | V] x [hi | lo] of the gathered rows)."""
    if pb.nb == 0:
        return
    st = ops._stream()
    ev = ops._timed(tag)
    if ev:
        ev[0].record()
    # split operands with src16 given: src16 rows are [hi(ncol) | lo(ncol)] of src * its
    # fixed scale (AMP_SCALE)
    split16 = precision in ("bf16x3", "fp16x3") and src16 is not None
    out32 = (ops.C_F32 << 8) if out.dtype == torch.float32 else 0
    if out32 and precision != "bf16" and not split16:
        raise _lib.NLError("_block_gemm: an FP32 output takes only the row-scatter products")
    if precision != "bf16" and not split16:
        G = ops.WS.get(tag + "_G", pb.nb * pb.kb * ncol, device=src.device)
        R = ops.WS.get(tag + "_R", pb.nb * pb.mb * ncol, device=src.device)
        _lib.call("nl_gather_rows", pb.nb * pb.kb, ncol, ops._p(pb.cols), ops._p(src), ncol,
                  ops._p(G), st)
    if precision == "fp64":
        ops.count_flops("fp64", 2.0 * pb.nb * pb.mb * ncol * pb.kb)
        _lib.call("nl_gemm_f64", 1, 1, pb.mb, ncol, pb.kb, ctypes.c_double(1.0), ops._p(pb.V),
                  pb.kb, pb.mb * pb.kb, ops._p(G), ncol, pb.kb * ncol, ctypes.c_double(0.0),
                  ops._p(R), ncol, pb.mb * ncol, pb.nb, 1, ctypes.c_void_p(0), st)
    else:
        dt, split = ops.lp_kind(precision)
        nblk = 2 if split else 1
        # split images pad K to 32, the split GEMM's step (zero columns contribute nothing)
        kpad = (pb.kb + 31) // 32 * 32 if split else (pb.kb + 7) // 8 * 8
        ldk = nblk * kpad
        lp = getattr(pb, "lp", None)
        if lp is None or lp[0] != (dt, split):
            # the blocks are Hamiltonian elements: exact per-row scales in both 16-bit types
            # (split: the block twice, [V | V], against [hi | lo] of the gathered rows)
            Vb = torch.empty(pb.nb * pb.mb * ldk, dtype=torch.int16, device=src.device)
            iv = torch.empty(pb.nb * pb.mb, dtype=torch.float32, device=src.device)
            s4 = (ctypes.c_int64 * 4)(pb.mb * pb.kb, pb.kb, 0, 1)
            tgt = ctypes.c_float(ops.lp_target(dt, True))
            if split:
                _lib.call("nl_convert_split", dt, pb.nb, pb.mb, 1, pb.kb, s4, ops._p(pb.V),
                          ops._p(Vb), ldk, kpad, 2, 0, ops._p(iv), tgt, ctypes.c_double(1.0), st)
            else:
                _lib.call("nl_convert_rows", dt, pb.nb, pb.mb, 1, pb.kb, s4, ops._p(pb.V),
                          ops._p(Vb), ldk, ops._p(iv), tgt, st)
            pb.lp = lp = ((dt, split), Vb, iv)
        _, Vb, iv = lp
        ops.count_flops("lp", 2.0 * pb.nb * pb.mb * ncol * ldk)
        target = ops.lp_target(dt, False)
        scaled = target != 0.0
        Gt = ops.WS.get(tag + "_Gt", pb.nb * ncol * ldk, dtype=torch.int16, device=src.device)
        if precision == "bf16" or split16:
            # BF16: gather + transposing cast in one pass, block rows added straight into out
            # (split: the hi and lo halves of the gathered rows, [V | V] x [hi | lo])
            if split16:
                _lib.call("nl_gather_rows16_split", pb.nb, pb.kb, ncol, ops._p(pb.cols),
                          ops._p(src16), 2 * ncol, ops._p(Gt), kpad, ldk, st)
            elif src16 is not None:
                _lib.call("nl_gather_rows16", pb.nb, pb.kb, ncol, ops._p(pb.cols), ops._p(src16),
                          ncol, ops._p(Gt), ldk, st)
            else:
                _lib.call("nl_gather_convert", pb.nb, pb.kb, ncol, ops._p(pb.cols), ops._p(src),
                          ncol, ops._p(Gt), ldk, st)
            a_eff = alpha / AMP_SCALE[dt] if split16 else alpha
            if split16:
                # V (exact, its first K block) x [hi | lo] of the gathered rows: one split GEMM
                # with the row-scatter epilogue (V.hi + V.lo per K step)
                _lib.call("nl_gemm_lp_rows_split", dt | out32, pb.mb, ncol, kpad,
                          ctypes.c_double(a_eff), ops._p(Vb), None, ldk, pb.mb * ldk,
                          ops._p(iv), pb.mb, ops._p(Gt), ops._off16(Gt, kpad), ldk, ncol * ldk,
                          None, 0, ops._p(out), ncol, ops._p(pb.rows), pb.nb, st)
            else:
                _lib.call("nl_gemm_lp_rows", dt | out32, pb.mb, ncol, ldk, ctypes.c_double(a_eff),
                          ops._p(Vb), ldk, pb.mb * ldk, ops._p(iv), pb.mb, ops._p(Gt), ldk,
                          ncol * ldk, None, 0, ops._p(out), ncol, ops._p(pb.rows), pb.nb, st)
            if ev:
                ev[1].record()
            return
        ig = (ops.WS.get(tag + "_Gs", pb.nb * ncol, dtype=torch.float32, device=src.device)
              if scaled else None)
        s4 = (ctypes.c_int64 * 4)(pb.kb * ncol, 1, 0, ncol)
        if split:
            _lib.call("nl_convert_split", dt, pb.nb, ncol, 1, pb.kb, s4, ops._p(G), ops._p(Gt),
                      ldk, kpad, 2, 0b10, ops._p(ig), ctypes.c_float(target),
                      ctypes.c_double(1.0), st)
        else:
            _lib.call("nl_convert_rows", dt, pb.nb, ncol, 1, pb.kb, s4, ops._p(G), ops._p(Gt),
                      ldk, ops._p(ig), ctypes.c_float(target), st)
        if split:
            # V (exact, its first K block) x [hi | lo]: the split GEMM, as in the shared-image path
            _lib.call("nl_gemm_lp_split", dt, pb.mb, ncol, kpad, ctypes.c_double(1.0),
                      ops._p(Vb), None, ldk, pb.mb * ldk, ops._p(iv), pb.mb, ops._p(Gt),
                      ops._off16(Gt, kpad), ldk, ncol * ldk, ops._p(ig), ncol,
                      ctypes.c_double(0.0), None, ops._p(R), ncol, pb.mb * ncol, pb.nb, 1, None,
                      st)
        else:
            _lib.call("nl_gemm_lp", dt, pb.mb, ncol, ldk, ctypes.c_double(1.0), ops._p(Vb), ldk,
                      pb.mb * ldk, ops._p(iv), pb.mb, ops._p(Gt), ldk, ncol * ldk, ops._p(ig),
                      ncol, ctypes.c_double(0.0), None, ops._p(R), ncol, pb.mb * ncol, pb.nb, 1,
                      None, st)
    _lib.call("nl_scatter_add_rows", pb.nb * pb.mb, ncol, ops._p(pb.rows), ops._p(R),
              ctypes.c_double(alpha), ops._p(out), ncol, st)
    if ev:
        ev[1].record()


class T2Image:
    """The BF16 image of t2 carried from one step's assembly to the next step's BF16 products
    (pass the same object to every ``update`` of a solve).  Two buffers alternate: the one the
    current step reads and the one its assembly writes.  ``src`` is the t2 tensor object whose
    image the current buffer holds; any other t2 (a DIIS extrapolation, a fresh start) is cast
    as before.  (Carrying tau the same way was measured and left out: the two extra t1 reads
    per element cost the assembly more than the tau pass they replace.)"""

    def __init__(self):
        self.bufs = [None, None]
        self.cur = 0
        self.src = None
        self.kind = 0           # 0: BF16 image, 1 / 2: [hi | lo] BF16 / FP16 split image
                                # (nl_assemble_t2_ext's image kinds), 3 / 4: the same split
                                # over the hole pairs i < j (NL_IMG_PAIRS)
        self.zeroed = [False, False]

    def buffer(self, k, n, dev, zero=False):
        """Buffer k of n 16-bit elements; ``zero``: zero-initialised (an image that leaves slots
        unwritten, whose zeros the products read) and not since written by a full image."""
        b = self.bufs[k]
        if b is None or b.numel() < n or b.device != dev or (zero and not self.zeroed[k]):
            b = (torch.zeros if zero else torch.empty)(n, dtype=torch.int16, device=dev)
            self.bufs[k] = b
        self.zeroed[k] = zero
        return b[:n]

    def holds(self, t2, kind=0):
        return self.src is t2 and self.kind == kind


def update(ham, t1, t2, B, t1n, t2n, mixing, prec=DEFAULT_PREC, H2out=None, img=None):
    """One Jacobi step (coupled_cluster.py:602-623): writes t1n, t2n and B.E[0] = E(t1n, t2n).
    ``img`` (a T2Image) lets the assembly write the BF16 image of t2n that the next step's BF16
    products read, instead of that step casting t2 again."""
    p, h = ham.p, ham.h
    with ops.lp_reuse(t2):
        B.t2_16 = None
        try:
            dt = _fast_split(ham, prec)
            if dt is not None:
                _t2_split_images(ham, t2, dt, B, img)
            else:
                _t2_images(ham, t2, B, prec, img)
            t1_part(ham, t1, t2, B, prec)
            _lib.call("nl_update_t1", p, h, ops._p(t1), ops._p(B.H1), ops._p(B.X1pp),
                      ops._p(B.X1hh), ctypes.c_double(mixing), ops._p(t1n), ops._stream())
            t2_part(ham, t1, t2, B, prec)
        finally:
            B.t2_16 = None
            B.t2_split = B.t2_pairs = B.N56r = B.t2_ring = None
    assemble(ham, t2, B, t1n, t2n, mixing, H2out, img=img, prec=prec)


# BF16 images of t2 for the products that read it in BF16 (all from one cast of t2):
#   A19 (dress)  t2 rows [a][(d,k,l)]   the cast itself
#   Sab          t2 rows [(b,i,j)][d]   16-bit transpose into the fused Sab operand
#   Y (hh)       t2 rows [(a,j)][(d,k)] batched 16-bit transpose
#   ring (both)  t2 rows [(a,i)][(c,k)] 16-bit permute
T2_IMAGES = True


def _images_wanted(ham, prec):
    p, h = ham.p, ham.h
    use = any(prec.get(k) == "bf16" for k in ("dress", "hh", "ring"))
    return bool(T2_IMAGES and use) and not (p % 8 or (p * h) % 8
                                            or p * p * h * h >= 2 ** 31 - 1)


def _t2_images(ham, t2, B, prec, img=None):
    """Register the BF16 images of t2 that this iteration's BF16 products read, made from a single
    FP64 -> BF16 pass over t2 instead of one conversion of t2 per product (or taken from ``img``
    when the previous step's assembly wrote this t2's image); the FP64 t2 stays valid for the
    FP64 products.  Needs t2 under ops.lp_reuse."""
    p, h = ham.p, ham.h
    if not _images_wanted(ham, prec):
        return
    use = {k for k in ("dress", "hh", "ring") if prec.get(k) == "bf16"}
    n, ph, hh = p * p * h * h, p * h, h * h
    dev, st = t2.device, ops._stream()
    if img is not None and img.holds(t2, 0):
        S = img.buffer(img.cur, n, dev)                               # [a][d][k][l]
    else:
        S = ops.WS.get("t2_bf16", n, dtype=torch.int16, device=dev)   # [a][d][k][l]
        s4 = (ctypes.c_int64 * 4)(0, p * hh, 0, 1)
        _lib.call("nl_convert_rows", ops.BF16, 1, p, 1, p * hh, s4, ops._p(t2), ops._p(S),
                  p * hh, None, ctypes.c_float(0.0), st)
    if "dress" in use:
        ops.lp_provide(t2, "adkl", ["a"], ["d", "k", "l"], S, p * hh)
        B.t2_16 = S                      # the fused Sab product transposes it into its operand
    if "hh" in use:
        T2 = ops.WS.get("t2_bf16_ajdk", n, dtype=torch.int16, device=dev)
        _lib.call("nl_transpose16_batched", p, ph, h, ops._p(S), h, p * hh, ops._p(T2), ph,
                  p * hh, st)
        ops.lp_provide(t2, "adkj", ["a", "j"], ["d", "k"], T2, ph)
    if "ring" in use:
        T3 = ops.WS.get("t2_bf16_aick", n, dtype=torch.int16, device=dev)   # [a][i][c][k]
        shp = (ctypes.c_int64 * 4)(p, h, p, h)
        istr = (ctypes.c_int64 * 4)(p * hh, h, hh, 1)                 # out [a][i][c][k]
        _lib.call("nl_permute16", 4, shp, istr, ops._p(S), ops._p(T3), st)
        ops.lp_provide(t2, "acik", ["a", "i"], ["c", "k"], T3, ph)


def _t2_split_images(ham, t2, dt, B, img=None):
    """Register the [hi | lo] split images of t2 that the bf16x3 products read: the rows
    [a][(d,k,l)] (A19; written by the previous step's assembly, or split here from FP64; kept
    as B.t2_split for the Sab product), and from them by a 16-bit permute [(a,i)][(c,k)] (both
    ring products and, with a sign, Y).  Needs t2 under ops.lp_reuse.
    Over the hole pairs (_img_pairs): the image is [a][hi((d,q)) | lo((d,q))] (Pp slots per d,
    zero from P on), kept as B.t2_pairs for the A19 and Sab products, and the ring rows are
    expanded from it with the antisymmetry's signs (nl_ring_from_pairs)."""
    p, h = ham.p, ham.h
    n, row, ph = p * p * h * h, p * h * h, p * h
    dev, st = t2.device, ops._stream()
    sc = AMP_SCALE[dt]
    R = ops.WS.get("t2_split_ring", 2 * n, dtype=torch.int16, device=dev)
    if _img_pairs(ham, dt):
        Pp = _pair_dims(ham)[1]
        if img is not None and img.holds(t2, 3 + dt):
            NR = img.buffer(img.cur, 2 * p * p * Pp, dev, zero=True)
        else:
            NR = ops.WS.get("t2_split_pairs", 2 * p * p * Pp, dtype=torch.int16, device=dev)
            _lib.call("nl_tau_split_pairs", p, h, None, ops._p(t2), ops._p(NR), Pp, p, dt,
                      ctypes.c_double(sc), st)
        _lib.call("nl_ring_from_pairs", p, h, Pp, ops._p(NR), ops._p(R), st)
        ops.lp_provide(t2, "acik", ["a", "i"], ["c", "k"], R, 2 * ph, dt, SPLIT, sc)
        B.t2_pairs, B.t2_pairs_dt = NR, dt
        B.t2_ring = (R, dt)
        return
    kind = 1 + dt
    if img is not None and img.holds(t2, kind):
        N = img.buffer(img.cur, 2 * n, dev)
    else:
        N = ops.WS.get("t2_split", 2 * n, dtype=torch.int16, device=dev)
        s4 = (ctypes.c_int64 * 4)(0, row, 0, 1)
        _lib.call("nl_convert_split", dt, 1, p, 1, row, s4, ops._p(t2), ops._p(N), 2 * row,
                  row, 2, 0b10, None, ctypes.c_float(0.0), ctypes.c_double(sc), st)
    ops.lp_provide(t2, "adkl", ["a"], ["d", "k", "l"], N, 2 * row, dt, SPLIT, sc)
    # ring operand rows (a,i), per part the columns (c,k): out (a, i, part, c, k)
    shp = (ctypes.c_int64 * 5)(p, h, 2, p, h)
    istr = (ctypes.c_int64 * 5)(2 * row, h, row, h * h, 1)
    _lib.call("nl_permute16", 5, shp, istr, ops._p(N), ops._p(R), st)
    ops.lp_provide(t2, "acik", ["a", "i"], ["c", "k"], R, 2 * ph, dt, SPLIT, sc)
    # the Sab product's t2 columns are transposed from N into its own operand (_sab_split)
    B.t2_split = N
    B.t2_ring = (R, dt)


def assemble(ham, t2, B, t1n, t2n, mixing, H2out=None, img=None, prec=None):
    """The assembly (nl_assemble_t2_ext, with the diagonal t2_part split off in B.dpp).  With
    ``img`` and ``prec`` it also writes the BF16 image of t2n into ``img`` when ``prec`` reads t2
    in BF16."""
    p, h = ham.p, ham.h
    st = ops._stream()
    dt = _fast_split(ham, prec) if prec is not None else None
    # the split form's FP32 accumulators, with v_pphh added here (S0 does not hold it)
    acc, io = (B.acc32(), ACC_F32 | ADD_V) if dt is not None and ACC32 else (B, 0)
    S0 = acc.S0
    if _pairs(ham, dt):
        S0 = B.s0_pairs(_pair_dims(ham)[1])      # S0 over the hole pairs (expanded here)
        io |= S0_PAIRS
    if B.sab_pairs:
        io |= SAB_PAIRS                          # Sab over the hole pairs (from _sab_split)
        B.sab_pairs = False
    args = (p, h, ops._p(S0), ops._p(acc.Sab), ops._p(acc.Sij), ops._p(acc.Spp),
            ops._p(B.X2pp),
            ops._p(B.X2hh), ops._p(getattr(B, "dpp", None)), None, ops._p(t2),
            ops._p(ham.v_pphh), ops._p(t1n), ctypes.c_double(mixing), ops._p(t2n),
            ops._p(H2out), ops._p(B.part))
    buf, kind, sc = None, 0, 1.0
    if img is not None and t2n is not None and dt is not None and _img_pairs(ham, dt):
        # the split image over the hole pairs (zero-initialised: slots from P on stay zero)
        nxt = 1 - img.cur
        buf = img.buffer(nxt, 2 * p * p * _pair_dims(ham)[1], t2n.device, zero=True)
        kind, sc = 1 + dt, AMP_SCALE[dt]
        img.cur, img.src, img.kind = nxt, t2n, 3 + dt
        io |= IMG_PAIRS_IO
    elif img is not None and t2n is not None and dt is not None:
        nxt = 1 - img.cur
        buf, kind, sc = img.buffer(nxt, 2 * p * p * h * h, t2n.device), 1 + dt, AMP_SCALE[dt]
        img.cur, img.src, img.kind = nxt, t2n, kind
    elif (img is not None and prec is not None and t2n is not None
            and _images_wanted(ham, prec)):
        nxt = 1 - img.cur
        buf = img.buffer(nxt, p * p * h * h, t2n.device)
        img.cur, img.src, img.kind = nxt, t2n, 0
    elif img is not None:
        img.src = None
    ev = ops._timed("assemble")
    if ev:
        ev[0].record()
    _lib.call("nl_assemble_t2_ext", *args, ops._p(buf), kind, ctypes.c_double(sc), io, st)
    if ev:
        ev[1].record()
    nparts = int(_lib.load().nl_partials(p * p * h * h))
    _lib.call("nl_finalize", nparts, ops._p(B.part), p * h, ops._p(ham.f_ph), ops._p(t1n),
              ops._p(B.E), st)


def energy(ham, t1, t2, B):
    """ccsd_energy (coupled_cluster.py:265-286) into B.E[0]."""
    p, h = ham.p, ham.h
    st = ops._stream()
    _lib.call("nl_energy_part", p, h, ops._p(ham.v_pphh), ops._p(t2), ops._p(t1), ops._p(B.part),
              st)
    nparts = int(_lib.load().nl_partials(p * p * h * h))
    _lib.call("nl_finalize", nparts, ops._p(B.part), p * h, ops._p(ham.f_ph), ops._p(t1),
              ops._p(B.E), st)


def init_amplitudes(ham, B, t1, t2, delta):
    """t1Init / t2Init (coupled_cluster.py:319-340, 416-439) through the update kernels."""
    p, h = ham.p, ham.h
    st = ops._stream()
    zero1 = torch.zeros_like(t1)
    permute_into(ham.f_hh, B.X1hh, alpha=-1.0)
    if delta:
        _diag_add(B.X1hh, -delta)
    _lib.call("nl_update_t1", p, h, ops._p(zero1), ops._p(ham.f_ph), ops._p(ham.f_pp),
              ops._p(B.X1hh), ctypes.c_double(0.0), ops._p(t1), st)
    permute_into(ham.f_hh, B.X2hh, alpha=-1.0)
    if delta:
        _diag_add(B.X2hh, -0.5 * delta)
    zero2 = ops.WS.get("zero_t2", p * p * h * h, device=t2.device).view(p, p, h, h)
    zero2.zero_()
    _lib.call("nl_assemble_t2", p, h, ops._p(ham.v_pphh), None, None, None, ops._p(ham.f_pp),
              ops._p(B.X2hh), ops._p(zero2), None, None, ctypes.c_double(0.0), ops._p(t2), None,
              ops._p(B.part), st)


def _diag_add(X, c):
    n = X.shape[0]
    eye = torch.eye(n, dtype=F64, device=X.device)
    ops.axpby(c, eye, 1.0, X)

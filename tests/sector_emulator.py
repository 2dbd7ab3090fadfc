"""NumPy statement of the sector-form C-ABI calls (nl_gemm_grouped_f64, nl_sector_gather,
nl_sector_dress, nl_sector_assemble), following their header comments, for testing the host
index structures of nulattice_amd.sectors without a GPU."""
import numpy as np


def gemm(desc, alpha, A, B, beta, C, D=None):
    D = C if D is None else D
    out = C.copy()
    for M, N, K, oA, sAm, sAk, oB, sBk, sBn, oC, sCm, sCn in desc:
        m = np.arange(M)[:, None]
        n = np.arange(N)[None, :]
        k = np.arange(K)
        a = A[oA + m * sAm + k[None, :] * sAk]                  # (M, K)
        b = B[oB + k[:, None] * sBk + n * sBn]                  # (K, N)
        c = oC + m * sCm + n * sCn
        out[c] = alpha * (a @ b) + beta * D[c]
    C[:] = out


def gather(idx, src):
    """Signed map: m >= 0 -> +src[m], m <= -2 -> -src[-m-2], m = -1 -> 0."""
    pos = src[np.maximum(idx, 0)]
    neg = -src[np.maximum(-idx - 2, 0)]
    return np.where(idx >= 0, pos, np.where(idx <= -2, neg, 0.0))


def dress(seg, elem, V, T, base, alpha):
    prod = V * T
    return np.array([base[o] + alpha * prod[elem[seg[o]:seg[o + 1]]].sum()
                     for o in range(len(base))])


def step(lay, hst, t, mixing):
    """One packed Jacobi step: (t_new, energy, H2)."""
    V, n = hst["V"], lay.n
    X = dress(lay.seg, lay.elem, V, t, hst["fdiag"], -1.0)
    L = np.zeros(n)
    Whh = np.zeros(lay.nhh)
    gemm(lay.desc_pp, 1.0, hst["VPP"], t, 0.0, L)
    gemm(lay.desc_whh, 1.0, V, t, 1.0, Whh, hst["VHH"])
    gemm(lay.desc_hh, 1.0, t, Whh, 1.0, L)
    Tr = gather(lay.ring_map, t)
    VPr = gather(lay.ring_map, V)
    W = np.zeros(lay.nw)
    gemm(lay.desc_w, 0.5, VPr, Tr, -1.0, W, hst["VPH"])
    S = np.zeros(lay.nring)
    gemm(lay.desc_s, 1.0, Tr, W, 0.0, S)
    a, b, i, j = lay.orb
    p = lay.p
    H = V + L + S[lay.sg[0]] - S[lay.sg[1]] - S[lay.sg[2]] + S[lay.sg[3]]
    dsum = X[a] + X[b] + X[p + i] + X[p + j]
    H = H + dsum * t
    tn = mixing * t + (1.0 - mixing) * (t + H / (-dsum))
    return tn, (V * tn).sum(), H


def contact_fill(off, ncol, roff, coff, rcode, ccode, crossed, g):
    """nl_contact_fill: block b = out[off[b]:off[b+1]], row-major, ncol[b] columns."""
    out = np.zeros(int(off[-1]))
    for b in range(len(off) - 1):
        nc = int(ncol[b])
        if nc == 0:
            continue
        nr = (int(off[b + 1]) - int(off[b])) // nc
        rc = rcode[roff[b]:roff[b] + nr].astype(np.int64)[:, None]
        cc = ccode[coff[b]:coff[b] + nc].astype(np.int64)[None, :]
        r1, r2, c1, c2 = rc >> 2, rc & 3, cc >> 2, cc & 3
        p, q, r, s = (c1, r2, r1, c2) if crossed else (r1, r2, c1, c2)
        out[off[b]:off[b + 1]] = (g * (((p == r) & (q == s)).astype(float)
                                      - ((p == s) & (q == r)).astype(float))).reshape(-1)
    return out


def ring_fused(pairs, ring_map, T, VPr, VPH, S):
    """nl_sector_ring: per pair {x, y}, W_x = 1/2 VPr_x Tr_y - VPH_x, Spp_x = Tr_x W_y (and the
    same with x and y exchanged unless self-conjugate), written into S at the Tr offsets."""
    S = S.copy()
    Tr = gather(ring_map, T)
    for rx, ry, wx, wy, n1, n2, self_ in pairs:
        Tx = Tr[rx:rx + n1 * n2].reshape(n1, n2)
        Px = VPr[rx:rx + n1 * n2].reshape(n1, n2)
        if self_:
            Wx = 0.5 * Px @ Tx - VPH[wx:wx + n1 * n1].reshape(n1, n1)
            S[rx:rx + n1 * n2] = (Tx @ Wx).reshape(-1)
            continue
        Ty = Tr[ry:ry + n2 * n1].reshape(n2, n1)
        Py = VPr[ry:ry + n2 * n1].reshape(n2, n1)
        Wx = 0.5 * Px @ Ty - VPH[wx:wx + n1 * n1].reshape(n1, n1)
        Wy = 0.5 * Py @ Tx - VPH[wy:wy + n2 * n2].reshape(n2, n2)
        S[rx:rx + n1 * n2] = (Tx @ Wy).reshape(-1)
        S[ry:ry + n2 * n1] = (Ty @ Wx).reshape(-1)
    return S


def ring_adj(pairs, ring_map, T, VPr, VPH, gS):
    """nl_sector_ring_adj: gTr_x = gS_x W_y^T + 1/2 VPr_y^T (Tr_x^T gS_x) per crossed pair."""
    Tr = gather(ring_map, T)
    gTr = np.zeros_like(gS)
    for rx, ry, wx, wy, n1, n2, self_ in pairs:
        Tx = Tr[rx:rx + n1 * n2].reshape(n1, n2)
        Px = VPr[rx:rx + n1 * n2].reshape(n1, n2)
        Gx = gS[rx:rx + n1 * n2].reshape(n1, n2)
        if self_:
            W = 0.5 * Px @ Tx - VPH[wx:wx + n1 * n1].reshape(n1, n1)
            gTr[rx:rx + n1 * n2] = (Gx @ W.T + 0.5 * Px.T @ (Tx.T @ Gx)).reshape(-1)
            continue
        Ty = Tr[ry:ry + n2 * n1].reshape(n2, n1)
        Py = VPr[ry:ry + n2 * n1].reshape(n2, n1)
        Gy = gS[ry:ry + n2 * n1].reshape(n2, n1)
        Wx = 0.5 * Px @ Ty - VPH[wx:wx + n1 * n1].reshape(n1, n1)
        Wy = 0.5 * Py @ Tx - VPH[wy:wy + n2 * n2].reshape(n2, n2)
        gTr[rx:rx + n1 * n2] = (Gx @ Wy.T + 0.5 * Py.T @ (Tx.T @ Gx)).reshape(-1)
        gTr[ry:ry + n2 * n1] = (Gy @ Wx.T + 0.5 * Px.T @ (Ty.T @ Gy)).reshape(-1)
    return gTr


def residual_vjp(lay, hst, t, lam):
    """J^T lam of the packed residual H(t) (every term but the D t diagonal's D lam part is
    returned in the first value; that part, D lam with D = X_a + X_b + X_i + X_j at t, in the
    second)."""
    V, n, p = hst["V"], lay.n, lay.p
    X = dress(lay.seg, lay.elem, V, t, hst["fdiag"], -1.0)
    Whh = np.zeros(lay.nhh)
    gemm(lay.desc_whh, 1.0, V, t, 1.0, Whh, hst["VHH"])
    out = np.zeros(n)
    gemm(lay.desc_ppT, 1.0, hst["VPP"], lam, 0.0, out)
    G1 = np.zeros(lay.nhh)
    gemm(lay.desc_whh, 1.0, t, lam, 0.0, G1)
    gemm(lay.desc_hh, 1.0, V, G1, 1.0, out)
    gemm(lay.desc_hhT, 1.0, lam, Whh, 1.0, out)
    gS = gather(lay.sg_inv, lam)
    VPr = gather(lay.ring_map, V)
    gTr = ring_adj(lay.ring_pairs, lay.ring_map, t, VPr, hst["VPH"], gS)
    for k in range(4):
        out += gather(lay.ring_inv[k], gTr)
    gX = dress(lay.seg, lay.elem, lam, t, np.zeros(len(X)), 1.0)
    a, b, i, j = lay.orb
    out -= V * (gX[a] + gX[b] + gX[p + i] + gX[p + j])
    return out, (X[a] + X[b] + X[p + i] + X[p + j]) * lam

// Symmetry-sector (block-sparse) kernels of the CCSD iteration on gfx950.
//
// When every single-particle state carries additive conserved labels that identify it (plane
// waves: momentum mod L, spin and isospin projections), T2 and every two-body block are block
// diagonal: t2[a,b,i,j] lives in the pair sector l_a + l_b = l_i + l_j, the ring products in the
// crossed sectors l_a - l_i, and the one-body dressings are diagonal.  The iteration is then a
// set of small dense products per sector (a grouped GEMM over a device-side problem list), a
// gather into the crossed-channel image, a segmented dressing reduction and one fused
// assembly / Jacobi update / energy pass over the packed amplitudes.
#include "nl_common.h"

extern "C" int64_t nl_partials(int64_t n);

namespace {

constexpr int NT = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline double block_sum(double v) {
  __shared__ double part[NT / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += part[i];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------- grouped GEMM
// Problem q: C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * D(m,n) with
// A(m,k) = A[oA + m*sAm + k*sAk], B(k,n) = B[oB + k*sBk + n*sBn], C/D(m,n) = X[oC + m*sCm + n*sCn].
// desc holds 12 int64 per problem: M N K oA sAm sAk oB sBk sBn oC sCm sCn.
// tiles holds 3 int32 per workgroup: problem, m0, n0.
//
// The sector products are small (M <= a few hundred, N <= 32) and many, so a workgroup's time is
// its chain of K steps: each step's global loads are issued into registers one step ahead
// (they are in flight while the previous step is computed from LDS), and each step covers 32
// values of k.  Operands are read along whichever index is unit-stride.
struct GDesc {
  int64_t M, N, K, oA, sAm, sAk, oB, sBk, sBn, oC, sCm, sCn;
};

__device__ inline GDesc load_desc(const int64_t* d) {
  return GDesc{d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11]};
}

__device__ inline void grouped_epilogue(const GDesc& g, int64_t m, int64_t n, double acc,
                                        double alpha, double beta, const double* D, double* C) {
  if (m >= g.M || n >= g.N) return;
  const int64_t off = g.oC + m * g.sCm + n * g.sCn;
  double v = alpha * acc;
  if (beta != 0.0) v += beta * (D ? D : C)[off];
  C[off] = v;
}

// 64 x 32 output tile, 256 threads of 4 x 2 outputs, K steps of 32.
constexpr int GTM = 64, GTN = 32, GTK = 32;

__global__ __launch_bounds__(NT) void gemm_grouped_f64_kernel(const int64_t* __restrict__ desc,
                                                              const int* __restrict__ tiles,
                                                              double alpha, const double* A,
                                                              const double* B, double beta,
                                                              const double* D, double* C) {
  __shared__ double As[GTK][GTM + 1];
  __shared__ double Bs[GTK][GTN + 1];
  const int q = tiles[3 * blockIdx.x], m0 = tiles[3 * blockIdx.x + 1],
            n0 = tiles[3 * blockIdx.x + 2];
  const GDesc g = load_desc(desc + 12 * (int64_t)q);
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const bool a_k1 = g.sAk == 1, b_n1 = g.sBn == 1;
  double ra[8], rb[4];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int idx = tid + NT * r;
      const int kk = a_k1 ? (idx & 31) : (idx >> 6);
      const int mm = a_k1 ? (idx >> 5) : (idx & 63);
      const int64_t m = m0 + mm, k = k0 + kk;
      ra[r] = (m < g.M && k < g.K) ? A[g.oA + m * g.sAm + k * g.sAk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + NT * r;
      const int nn = idx & 31, kk = idx >> 5;           // b_n1: n fastest
      const int nb = b_n1 ? nn : kk, kb = b_n1 ? kk : nn;
      const int64_t n = n0 + nb, k = k0 + kb;
      rb[r] = (n < g.N && k < g.K) ? B[g.oB + k * g.sBk + n * g.sBn] : 0.0;
    }
  };
  double acc[4][2];
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r][0] = acc[r][1] = 0.0;
  if (g.K > 0) fetch(0);
  for (int64_t k0 = 0; k0 < g.K; k0 += GTK) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int idx = tid + NT * r;
      const int kk = a_k1 ? (idx & 31) : (idx >> 6);
      const int mm = a_k1 ? (idx >> 5) : (idx & 63);
      As[kk][mm] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + NT * r;
      const int nn = idx & 31, kk = idx >> 5;
      Bs[b_n1 ? kk : nn][b_n1 ? nn : kk] = rb[r];
    }
    __syncthreads();
    if (k0 + GTK < g.K) fetch(k0 + GTK);
#pragma unroll 8
    for (int kk = 0; kk < GTK; ++kk) {
      const double b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double a = As[kk][ty + 16 * r];
        acc[r][0] = fma(a, b0, acc[r][0]);
        acc[r][1] = fma(a, b1, acc[r][1]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
      grouped_epilogue(g, m0 + ty + 16 * r, n0 + tx + 16 * c, acc[r][c], alpha, beta, D, C);
}

// Split-K tiles for the many small-N products (most pair sectors hold 1-8 hole pairs): a
// TM x TN output tile per workgroup, whose NW wavefronts each own 1/NW of K (RM x 4 outputs per
// lane, K steps of 16 with the same one-step register prefetch); the partial tiles are summed
// through LDS in a fixed order.  (TM, TN, NW) = (32, 32, 4), (64, 8, 4), or (32, 32, 8) for the
// few widest products, whose time is the length of their K chain.
constexpr int STK = 16;

template <int TM, int TN, int NW>
__global__ __launch_bounds__(64 * NW) void gemm_grouped_split_kernel(const int64_t* __restrict__ desc,
                                                                const int* __restrict__ tiles,
                                                                double alpha, const double* A,
                                                                const double* B, double beta,
                                                                const double* D, double* C) {
  constexpr int NL = TN / 4, ML = 64 / NL, RM = TM / ML;   // lanes along N, along M; rows/lane
  constexpr int AE = TM * STK / 64, BE = STK * TN / 64;    // A, B elements per lane per step
  constexpr int NTH = 64 * NW;
  __shared__ double As[NW][STK][TM + 1];
  __shared__ double Bs[NW][STK][TN + 1];
  __shared__ double Red[NW][TM][TN + 1];
  const int q = tiles[3 * blockIdx.x], m0 = tiles[3 * blockIdx.x + 1],
            n0 = tiles[3 * blockIdx.x + 2];
  const GDesc g = load_desc(desc + 12 * (int64_t)q);
  const int tid = threadIdx.x, grp = tid >> 6, lane = tid & 63, ly = lane / NL, lx = lane % NL;
  const bool a_k1 = g.sAk == 1, b_n1 = g.sBn == 1;
  const int64_t kq = ((g.K + NW - 1) / NW + STK - 1) / STK * STK;   // K range per wavefront
  const int64_t kb0 = grp * kq, kb1 = (kb0 + kq < g.K) ? kb0 + kq : g.K;
  double ra[AE], rb[BE];
  auto a_at = [&](int idx, int& mm, int& kk) {
    if (a_k1) { kk = idx & (STK - 1); mm = idx / STK; } else { mm = idx % TM; kk = idx / TM; }
  };
  auto b_at = [&](int idx, int& nn, int& kk) {
    if (b_n1) { nn = idx % TN; kk = idx / TN; } else { kk = idx & (STK - 1); nn = idx / STK; }
  };
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < AE; ++r) {
      int mm, kk;
      a_at(lane + 64 * r, mm, kk);
      const int64_t m = m0 + mm, k = k0 + kk;
      ra[r] = (m < g.M && k < kb1) ? A[g.oA + m * g.sAm + k * g.sAk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < BE; ++r) {
      int nn, kk;
      b_at(lane + 64 * r, nn, kk);
      const int64_t n = n0 + nn, k = k0 + kk;
      rb[r] = (n < g.N && k < kb1) ? B[g.oB + k * g.sBk + n * g.sBn] : 0.0;
    }
  };
  double acc[RM][4];
#pragma unroll
  for (int r = 0; r < RM; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
  const int64_t steps = kq / STK;                         // the same count in every wavefront
  if (steps > 0) fetch(kb0);
  for (int64_t s = 0; s < steps; ++s) {
    const int64_t k0 = kb0 + s * STK;
#pragma unroll
    for (int r = 0; r < AE; ++r) {
      int mm, kk;
      a_at(lane + 64 * r, mm, kk);
      As[grp][kk][mm] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < BE; ++r) {
      int nn, kk;
      b_at(lane + 64 * r, nn, kk);
      Bs[grp][kk][nn] = rb[r];
    }
    __syncthreads();
    if (s + 1 < steps) fetch(k0 + STK);
#pragma unroll
    for (int kk = 0; kk < STK; ++kk) {
      double a[RM], b[4];
#pragma unroll
      for (int r = 0; r < RM; ++r) a[r] = As[grp][kk][ly + ML * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[grp][kk][lx + NL * c];
#pragma unroll
      for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < RM; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) Red[grp][ly + ML * r][lx + NL * c] = acc[r][c];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < TM * TN / NTH; ++r) {
    const int idx = tid + NTH * r;
    const int m = idx / TN, n = idx % TN;
    double v = Red[0][m][n];
#pragma unroll
    for (int w = 1; w < NW; ++w) v += Red[w][m][n];
    grouped_epilogue(g, m0 + m, n0 + n, v, alpha, beta, D, C);
  }
}

// ---------------------------------------------------------------- gathers
__global__ __launch_bounds__(NT) void sector_gather_kernel(int64_t n, const int64_t* __restrict__ idx,
                                                           const double* __restrict__ in,
                                                           double* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < n; t += (int64_t)gridDim.x * NT) {
    const int64_t s = idx[t];
    out[t] = s >= 0 ? in[s] : (s <= -2 ? -in[-s - 2] : 0.0);
  }
}

__global__ __launch_bounds__(NT) void sector_scatter_kernel(int64_t n,
                                                            const int64_t* __restrict__ idx,
                                                            const double* __restrict__ in,
                                                            double* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < n; t += (int64_t)gridDim.x * NT) {
    const int64_t s = idx[t];
    if (s >= 0) out[s] = in[t];
    else if (s <= -2) out[-s - 2] = -in[t];
  }
}

// ---------------------------------------------------------------- dressing
// X[o] = base[o] + alpha * sum_{e in segment o} V[elem[e]] * T[elem[e]].  Segments are split into
// chunks of consecutive elem entries (chunk c = cstart[c]..cstart[c+1], chunks of output o =
// cptr[o]..cptr[o+1]); one wavefront sums a chunk into ws[c], then one lane per output adds its
// chunks in order (the hole segments are ~p^2 long, so a wavefront per output would serialise).
__global__ __launch_bounds__(NT) void sector_dress_chunks_kernel(int64_t nchunk,
                                                                 const int64_t* __restrict__ cstart,
                                                                 const int* __restrict__ elem,
                                                                 const double* __restrict__ V,
                                                                 const double* __restrict__ T,
                                                                 double* __restrict__ ws) {
  const int64_t c = blockIdx.x * (int64_t)(NT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= nchunk) return;
  double acc = 0.0;
  for (int64_t e = cstart[c] + lane; e < cstart[c + 1]; e += 64) {
    const int s = elem[e];
    acc = fma(V[s], T[s], acc);
  }
  acc = wave_sum(acc);
  if (lane == 0) ws[c] = acc;
}

__global__ __launch_bounds__(NT) void sector_dress_sum_kernel(int64_t no,
                                                              const int64_t* __restrict__ cptr,
                                                              const double* __restrict__ ws,
                                                              const double* __restrict__ base,
                                                              double alpha, double* __restrict__ X) {
  const int64_t o = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (o >= no) return;
  double acc = 0.0;
  for (int64_t c = cptr[o]; c < cptr[o + 1]; ++c) acc += ws[c];
  X[o] = base[o] + alpha * acc;
}

// ---------------------------------------------------------------- crossed-channel ring
// One workgroup per pair of crossed sectors {x, y = -x} (|R_x|, |R_y| <= 32): the images
// Tr_x [R_x][R_y] and Tr_y [R_y][R_x] are gathered from the packed amplitudes through the signed
// map, and with VPr (the same images of v_pphh) and VPH staged in LDS it forms
//   W_x = 1/2 VPr_x Tr_y - VPH_x,  W_y = 1/2 VPr_y Tr_x - VPH_y   (in LDS only)
//   Spp_x = Tr_x W_y,  Spp_y = Tr_y W_x                          (written to S)
// so the ring intermediate never reaches HBM.  A self-conjugate sector (y = x) is its own pair.
constexpr int RR = 32;

__device__ inline double signed_at(const int64_t* map, const double* v, int64_t k) {
  const int64_t m = map[k];
  return m >= 0 ? v[m] : (m <= -2 ? -v[-m - 2] : 0.0);
}

__global__ __launch_bounds__(NT) void sector_ring_kernel(const int64_t* __restrict__ pairs,
                                                         const int64_t* __restrict__ rmap,
                                                         const double* __restrict__ T,
                                                         const double* __restrict__ VPr,
                                                         const double* __restrict__ VPH,
                                                         double* __restrict__ S) {
  __shared__ double Tx[RR][RR + 1], Ty[RR][RR + 1], Px[RR][RR + 1], Py[RR][RR + 1];
  __shared__ double Wx[RR][RR + 1], Wy[RR][RR + 1];
  const int64_t* q = pairs + 7 * (int64_t)blockIdx.x;
  const int64_t rx = q[0], ry = q[1], wx = q[2], wy = q[3];
  // the host checks |R| <= RR (max_rows); the clamp only keeps a bad table inside the LDS arrays
  const int n1 = (int)min(q[4], (int64_t)RR), n2 = (int)min(q[5], (int64_t)RR);
  const bool self = q[6] != 0;
  const int tid = threadIdx.x;
  for (int e = tid; e < n1 * n2; e += NT) {
    const int r = e / n2, c = e % n2;
    Tx[r][c] = signed_at(rmap, T, rx + e);
    Px[r][c] = VPr[rx + e];
  }
  for (int e = tid; e < n1 * n1; e += NT) Wx[e / n1][e % n1] = VPH[wx + e];
  if (!self) {
    for (int e = tid; e < n2 * n1; e += NT) {
      const int r = e / n1, c = e % n1;
      Ty[r][c] = signed_at(rmap, T, ry + e);
      Py[r][c] = VPr[ry + e];
    }
    for (int e = tid; e < n2 * n2; e += NT) Wy[e / n2][e % n2] = VPH[wy + e];
  }
  __syncthreads();
  // W_x (n1 x n1, K = n2) and W_y (n2 x n2, K = n1); each thread owns up to 4 + 4 entries
  double wxv[4], wyv[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tid + NT * u;
    wxv[u] = 0.0;
    if (e < n1 * n1) {
      const int r = e / n1, c = e % n1;
      double acc = 0.0;
      for (int k = 0; k < n2; ++k) acc = fma(Px[r][k], self ? Tx[k][c] : Ty[k][c], acc);
      wxv[u] = 0.5 * acc - Wx[r][c];
    }
    wyv[u] = 0.0;
    if (!self && e < n2 * n2) {
      const int r = e / n2, c = e % n2;
      double acc = 0.0;
      for (int k = 0; k < n1; ++k) acc = fma(Py[r][k], Tx[k][c], acc);
      wyv[u] = 0.5 * acc - Wy[r][c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tid + NT * u;
    if (e < n1 * n1) Wx[e / n1][e % n1] = wxv[u];
    if (!self && e < n2 * n2) Wy[e / n2][e % n2] = wyv[u];
  }
  __syncthreads();
  // Spp_x = Tr_x W_y (n1 x n2), Spp_y = Tr_y W_x (n2 x n1)
  for (int e = tid; e < n1 * n2; e += NT) {
    const int r = e / n2, c = e % n2;
    double acc = 0.0;
    for (int k = 0; k < n2; ++k) acc = fma(Tx[r][k], self ? Wx[k][c] : Wy[k][c], acc);
    S[rx + e] = acc;
  }
  if (!self) {
    for (int e = tid; e < n2 * n1; e += NT) {
      const int r = e / n1, c = e % n1;
      double acc = 0.0;
      for (int k = 0; k < n1; ++k) acc = fma(Ty[r][k], Wx[k][c], acc);
      S[ry + e] = acc;
    }
  }
}

// Reverse sweep of the ring kernel (the Lambda equations): given the adjoint gS of Spp in the
// crossed layout, recompute W_x, W_y in LDS and form
//   gW_y = Tr_x^T gS_x,  gW_x = Tr_y^T gS_y,
//   gTr_x = gS_x W_y^T + 1/2 VPr_y^T gW_y,  gTr_y = gS_y W_x^T + 1/2 VPr_x^T gW_x,
// written to gTr at the offsets of Tr_x and Tr_y (self-conjugate: y = x, one set).
__global__ __launch_bounds__(NT) void sector_ring_adj_kernel(const int64_t* __restrict__ pairs,
                                                             const int64_t* __restrict__ rmap,
                                                             const double* __restrict__ T,
                                                             const double* __restrict__ VPr,
                                                             const double* __restrict__ VPH,
                                                             const double* __restrict__ gS,
                                                             double* __restrict__ gTr) {
  __shared__ double Tx[RR][RR + 1], Ty[RR][RR + 1], Px[RR][RR + 1], Py[RR][RR + 1];
  __shared__ double Wx[RR][RR + 1], Wy[RR][RR + 1], Gx[RR][RR + 1], Gy[RR][RR + 1];
  __shared__ double Hx[RR][RR + 1], Hy[RR][RR + 1];          // gW_x, gW_y
  const int64_t* q = pairs + 7 * (int64_t)blockIdx.x;
  const int64_t rx = q[0], ry = q[1], wx = q[2], wy = q[3];
  // the host checks |R| <= RR (max_rows); the clamp only keeps a bad table inside the LDS arrays
  const int n1 = (int)min(q[4], (int64_t)RR), n2 = (int)min(q[5], (int64_t)RR);
  const bool self = q[6] != 0;
  const int tid = threadIdx.x;
  // with y = x the y-side arrays alias the x-side ones
  double (*TY)[RR + 1] = self ? Tx : Ty;
  double (*PY)[RR + 1] = self ? Px : Py;
  double (*WY)[RR + 1] = self ? Wx : Wy;
  double (*GY)[RR + 1] = self ? Gx : Gy;
  double (*HY)[RR + 1] = self ? Hx : Hy;
  for (int e = tid; e < n1 * n2; e += NT) {
    const int r = e / n2, c = e % n2;
    Tx[r][c] = signed_at(rmap, T, rx + e);
    Px[r][c] = VPr[rx + e];
    Gx[r][c] = gS[rx + e];
  }
  if (!self) {
    for (int e = tid; e < n2 * n1; e += NT) {
      const int r = e / n1, c = e % n1;
      Ty[r][c] = signed_at(rmap, T, ry + e);
      Py[r][c] = VPr[ry + e];
      Gy[r][c] = gS[ry + e];
    }
  }
  __syncthreads();
  // W_x = 1/2 P_x T_y - VPH_x (n1 x n1), W_y = 1/2 P_y T_x - VPH_y (n2 x n2)
  for (int e = tid; e < n1 * n1; e += NT) {
    const int r = e / n1, c = e % n1;
    double acc = 0.0;
    for (int k = 0; k < n2; ++k) acc = fma(Px[r][k], TY[k][c], acc);
    Wx[r][c] = 0.5 * acc - VPH[wx + e];
  }
  if (!self)
    for (int e = tid; e < n2 * n2; e += NT) {
      const int r = e / n2, c = e % n2;
      double acc = 0.0;
      for (int k = 0; k < n1; ++k) acc = fma(Py[r][k], Tx[k][c], acc);
      Wy[r][c] = 0.5 * acc - VPH[wy + e];
    }
  // gW_y = T_x^T gS_x (n2 x n2), gW_x = T_y^T gS_y (n1 x n1)
  for (int e = tid; e < n2 * n2; e += NT) {
    const int r = e / n2, c = e % n2;
    double acc = 0.0;
    for (int k = 0; k < n1; ++k) acc = fma(Tx[k][r], Gx[k][c], acc);
    HY[r][c] = acc;
  }
  if (!self)
    for (int e = tid; e < n1 * n1; e += NT) {
      const int r = e / n1, c = e % n1;
      double acc = 0.0;
      for (int k = 0; k < n2; ++k) acc = fma(Ty[k][r], Gy[k][c], acc);
      Hx[r][c] = acc;
    }
  __syncthreads();
  // gTr_x = gS_x W_y^T + 1/2 P_y^T gW_y (n1 x n2)
  for (int e = tid; e < n1 * n2; e += NT) {
    const int r = e / n2, c = e % n2;
    double a1 = 0.0, a2 = 0.0;
    for (int k = 0; k < n2; ++k) {
      a1 = fma(Gx[r][k], WY[c][k], a1);
      a2 = fma(PY[k][r], HY[k][c], a2);
    }
    gTr[rx + e] = a1 + 0.5 * a2;
  }
  if (!self)
    for (int e = tid; e < n2 * n1; e += NT) {
      const int r = e / n1, c = e % n1;
      double a1 = 0.0, a2 = 0.0;
      for (int k = 0; k < n1; ++k) {
        a1 = fma(Gy[r][k], Wx[c][k], a1);
        a2 = fma(Px[k][r], Hx[k][c], a2);
      }
      gTr[ry + e] = a1 + 0.5 * a2;
    }
}

// out[e] += alpha * (signed in[idx[e]])
__global__ __launch_bounds__(NT) void sector_gather_add_kernel(int64_t n, const int64_t* __restrict__ idx,
                                                               const double* __restrict__ in,
                                                               double alpha,
                                                               double* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < n; t += (int64_t)gridDim.x * NT) {
    const int64_t s = idx[t];
    if (s >= 0) out[t] += alpha * in[s];
    else if (s <= -2) out[t] -= alpha * in[-s - 2];
  }
}

// out[e] += alpha * W[e] * (G[a] + G[b] + G[p+i] + G[p+j])
__global__ __launch_bounds__(NT) void sector_diag_axpy_kernel(int64_t n, int64_t p,
                                                              const int* __restrict__ orb,
                                                              const double* __restrict__ W,
                                                              const double* __restrict__ G,
                                                              double alpha, double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += (int64_t)gridDim.x * NT) {
    const int a = orb[e], b = orb[n + e], i = orb[2 * n + e], j = orb[3 * n + e];
    out[e] += alpha * W[e] * (G[a] + G[b] + G[p + i] + G[p + j]);
  }
}

// ---------------------------------------------------------------- contact blocks
// Sector blocks of the SU(4) contact <pq||rs> = g [d(s_p,s_r) d(s_q,s_s) - d(s_p,s_s) d(s_q,s_r)]
// (momentum conservation holds inside every sector by construction).  Block b occupies
// out[off[b] .. off[b+1]) as a row-major nrow x ncol matrix; row r carries the spin code
// rs[roff[b] + r], column c the code cs[coff[b] + c] (code = 4 s_first + s_second).  crossed = 0:
// bra = the row pair, ket = the column pair; crossed = 1 (W[(c,k),(b,j)] = <b k||c j>): bra =
// (column first, row second), ket = (row first, column second).
__global__ __launch_bounds__(NT) void contact_fill_kernel(int64_t nblk, const int64_t* __restrict__ off,
                                                          const int64_t* __restrict__ ncol,
                                                          const int64_t* __restrict__ roff,
                                                          const int64_t* __restrict__ coff,
                                                          const unsigned char* __restrict__ rs,
                                                          const unsigned char* __restrict__ cs,
                                                          int crossed, double g,
                                                          double* __restrict__ out) {
  const int64_t total = off[nblk];
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * NT) {
    int64_t lo = 0, hi = nblk - 1;                  // block b with off[b] <= t < off[b+1]
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (off[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int64_t local = t - off[lo], nc = ncol[lo];
    const int r = (int)(local / nc), c = (int)(local % nc);
    const int rc = rs[roff[lo] + r], cc = cs[coff[lo] + c];
    const int r1 = rc >> 2, r2 = rc & 3, c1 = cc >> 2, c2 = cc & 3;
    const int p = crossed ? c1 : r1, q = r2, r_ = crossed ? r1 : c1, s_ = c2;
    out[t] = g * (double)((p == r_ && q == s_) - (p == s_ && q == r_));
  }
}

// ---------------------------------------------------------------- plane-wave blocks
// An orbital code packs (kx, ky, kz, sigma) as kx | ky << 8 | kz << 16 | sigma << 24, L <= 255.
__device__ __forceinline__ int pw_ksum(int a, int b, int L) {     // k_a + k_b mod L, per axis
  int x = (a & 255) + (b & 255), y = ((a >> 8) & 255) + ((b >> 8) & 255),
      z = ((a >> 16) & 255) + ((b >> 16) & 255);
  x -= x >= L ? L : 0; y -= y >= L ? L : 0; z -= z >= L ? L : 0;
  return x | (y << 8) | (z << 16);
}

// <r1 r2||c1 c2> = g delta(K) [d(s1,s3) d(s2,s4) - d(s1,s4) d(s2,s3)], written as the product of
// the small-integer spin factor and g so each element is the same double the host forms.
__global__ __launch_bounds__(NT) void pw_dense_kernel(int64_t n0, int n1, int n2, int n3,
                                                      const int* __restrict__ c0,
                                                      const int* __restrict__ c1,
                                                      const int* __restrict__ c2,
                                                      const int* __restrict__ c3, int L, double g,
                                                      double* __restrict__ out) {
  const int64_t inner = (int64_t)n2 * n3, total = n0 * n1 * inner;
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * NT) {
    const int64_t row = t / inner;
    const int col = (int)(t - row * inner);
    const int i0 = (int)(row / n1), i1 = (int)(row - (int64_t)i0 * n1);
    const int i2 = col / n3, i3 = col - i2 * n3;
    const int a = c0[i0], b = c1[i1], c = c2[i2], d = c3[i3];
    const int sa = a >> 24, sb = b >> 24, sc = c >> 24, sd = d >> 24;
    const int spin = (sa == sc && sb == sd) - (sa == sd && sb == sc);
    out[t] = pw_ksum(a, b, L) == pw_ksum(c, d, L) ? (double)spin * g : 0.0;
  }
}

// Value of <x y||u w> for pairs with s_x != s_y, s_u != s_w: g when (s_u, s_w) = (s_x, s_y),
// -g when swapped, 0 unless total momentum and the unordered spin pair agree.
__device__ __forceinline__ bool pw_pair_val(int x, int y, int u, int w, int L, double g,
                                            double* v) {
  const int sx = x >> 24, sy = y >> 24, su = u >> 24, sw = w >> 24;
  if (sx == sy || su == sw) return false;
  if (!((sx == su && sy == sw) || (sx == sw && sy == su))) return false;
  if (pw_ksum(x, y, L) != pw_ksum(u, w, L)) return false;
  *v = (sx == su ? 1.0 : -1.0) * g;
  return true;
}

struct PwPairArgs {
  int64_t nb, mb, kb;
  const int64_t* rows;   // [nb][mb] row pair index (-1 = padding), pair = first * rdiv + second
  const int64_t* cols;   // [nb][kb] column pair index (-1 = padding)
  int64_t rdiv, cdiv;
  const int* r0; const int* r1; const int* q0; const int* q1;   // codes of the pair members
  int L;
  double g;
  double* out;           // [nb][mb][kb]
};

__global__ __launch_bounds__(NT) void pw_pair_fill_kernel(PwPairArgs a) {
  const int64_t total = a.nb * a.mb * a.kb;
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * NT) {
    const int64_t brow = t / a.kb, c = t - brow * a.kb;
    const int64_t b = brow / a.mb;
    const int64_t ra = a.rows[brow], ca = a.cols[b * a.kb + c];
    double v = 0.0, e;
    if (ra >= 0 && ca >= 0 &&
        pw_pair_val(a.r0[ra / a.rdiv], a.r1[ra % a.rdiv], a.q0[ca / a.cdiv], a.q1[ca % a.cdiv],
                    a.L, a.g, &e))
      v = 0.0 + e;                    // the host scatter adds each value to a zero once
    a.out[t] = v;
  }
}

struct PwCooArgs {
  int64_t ngrp;
  const int64_t* eoff;   // [ngrp+1] first entry of each group
  const int64_t* roff;   // [ngrp+1] rows of group q: rpair[roff[q]..roff[q+1])
  const int64_t* coff;   // [ngrp+1]
  const int64_t* rpair; const int64_t* cpair;
  int64_t rdiv, cdiv;
  const int* r0; const int* r1; const int* q0; const int* q1;
  int L;
  double g;
  int64_t* idx;          // [4][n]: rpair / rdiv, rpair % rdiv, cpair / cdiv, cpair % cdiv
  double* val;
  int64_t n;
};

__global__ __launch_bounds__(NT) void pw_coo_kernel(PwCooArgs a) {
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < a.n;
       t += (int64_t)gridDim.x * NT) {
    int64_t lo = 0, hi = a.ngrp - 1;                // group q with eoff[q] <= t < eoff[q+1]
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (a.eoff[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int64_t nc = a.coff[lo + 1] - a.coff[lo], local = t - a.eoff[lo];
    const int64_t r = local / nc, c = local - r * nc;
    const int64_t rp = a.rpair[a.roff[lo] + r], cp = a.cpair[a.coff[lo] + c];
    const int64_t x = rp / a.rdiv, y = rp % a.rdiv, u = cp / a.cdiv, w = cp % a.cdiv;
    a.idx[t] = x; a.idx[a.n + t] = y; a.idx[2 * a.n + t] = u; a.idx[3 * a.n + t] = w;
    a.val[t] = ((a.r0[x] >> 24) == (a.q0[u] >> 24) ? 1.0 : -1.0) * a.g;
  }
}

// ---------------------------------------------------------------- assembly
struct SectorAsm {
  int64_t n, p;
  const int* orb;        // [4][n]: a, b (particles), i, j (holes) of packed element e
  const int* sg;         // [4][n]: ring-image offsets of S[a,i,b,j], S[a,j,b,i], S[b,i,a,j],
                         //         S[b,j,a,i]; null = no ring term
  const double *V, *L, *S, *X, *t;
  double mixing, escale;
  double *tn, *H2out, *part;
};

__global__ __launch_bounds__(NT) void sector_assemble_kernel(SectorAsm g) {
  double e_acc = 0.0, r_acc = 0.0;
  const int64_t n = g.n;
  const double* Xh = g.X + g.p;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += (int64_t)gridDim.x * NT) {
    const int a = g.orb[e], b = g.orb[n + e], i = g.orb[2 * n + e], j = g.orb[3 * n + e];
    const double v = g.V[e], told = g.t[e];
    double H = v;
    if (g.L) H += g.L[e];
    if (g.S)
      H += g.S[g.sg[e]] - g.S[g.sg[n + e]] - g.S[g.sg[2 * n + e]] + g.S[g.sg[3 * n + e]];
    const double dsum = g.X[a] + g.X[b] + Xh[i] + Xh[j];
    H += dsum * told;
    const double nt = told + H / (-dsum);
    const double tn = g.mixing * told + (1.0 - g.mixing) * nt;
    if (g.tn) g.tn[e] = tn;
    if (g.H2out) g.H2out[e] = H;
    e_acc += g.escale * v * tn;
    r_acc += H * H;
  }
  const double es = block_sum(e_acc);
  const double rs = block_sum(r_acc);
  if (threadIdx.x == 0) {
    g.part[blockIdx.x] = es;
    g.part[gridDim.x + blockIdx.x] = rs;
  }
}

inline int grid_of(int64_t n, int cap) {
  int64_t g = nl_cdiv(n, NT);
  if (g < 1) g = 1;
  return (int)(g < cap ? g : cap);
}

}  // namespace

extern "C" {

int nl_gemm_grouped_f64(int mode, int64_t ntile, const int64_t* desc, const int* tiles,
                        double alpha, const double* A, const double* B, double beta,
                        const double* D, double* C, void* stream) {
  if (ntile < 0 || ntile > INT32_MAX || mode < 0 || mode > 3) return NL_EARG;
  if (ntile == 0) return NL_OK;
  if (!desc || !tiles || !A || !B || !C) return NL_EARG;
  hipStream_t st = (hipStream_t)stream;
  if (mode == 0)
    hipLaunchKernelGGL(gemm_grouped_f64_kernel, dim3((unsigned)ntile), dim3(NT), 0, st, desc,
                       tiles, alpha, A, B, beta, D, C);
  else if (mode == 1)
    hipLaunchKernelGGL((gemm_grouped_split_kernel<32, 32, 4>), dim3((unsigned)ntile), dim3(NT),
                       0, st, desc, tiles, alpha, A, B, beta, D, C);
  else if (mode == 2)
    hipLaunchKernelGGL((gemm_grouped_split_kernel<64, 8, 4>), dim3((unsigned)ntile), dim3(NT), 0,
                       st, desc, tiles, alpha, A, B, beta, D, C);
  else
    hipLaunchKernelGGL((gemm_grouped_split_kernel<32, 32, 8>), dim3((unsigned)ntile),
                       dim3(512), 0, st, desc, tiles, alpha, A, B, beta, D, C);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_gather(int64_t n, const int64_t* idx, const double* in, double* out, void* stream) {
  if (n < 0) return NL_EARG;
  if (n == 0) return NL_OK;
  if (!idx || !in || !out) return NL_EARG;
  hipLaunchKernelGGL(sector_gather_kernel, dim3(grid_of(n, 8192)), dim3(NT), 0,
                     (hipStream_t)stream, n, idx, in, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_scatter(int64_t n, const int64_t* idx, const double* in, double* out, void* stream) {
  if (n < 0) return NL_EARG;
  if (n == 0) return NL_OK;
  if (!idx || !in || !out) return NL_EARG;
  hipLaunchKernelGGL(sector_scatter_kernel, dim3(grid_of(n, 8192)), dim3(NT), 0,
                     (hipStream_t)stream, n, idx, in, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_dress(int64_t no, const int64_t* cptr, int64_t nchunk, const int64_t* cstart,
                    const int* elem, const double* V, const double* T, const double* base,
                    double alpha, double* ws, double* X, void* stream) {
  if (no < 0 || nchunk < 0) return NL_EARG;
  if (no == 0) return NL_OK;
  if (!cptr || !base || !X || (nchunk > 0 && (!cstart || !elem || !V || !T || !ws)))
    return NL_EARG;
  const int64_t g1 = nl_cdiv(nchunk, NT / 64), g2 = nl_cdiv(no, NT);
  if (g1 > INT32_MAX || g2 > INT32_MAX) return NL_EARG;
  if (nchunk > 0)
    hipLaunchKernelGGL(sector_dress_chunks_kernel, dim3((unsigned)g1), dim3(NT), 0,
                       (hipStream_t)stream, nchunk, cstart, elem, V, T, ws);
  hipLaunchKernelGGL(sector_dress_sum_kernel, dim3((unsigned)g2), dim3(NT), 0,
                     (hipStream_t)stream, no, cptr, ws, base, alpha, X);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_ring(int64_t npair, int64_t max_rows, const int64_t* pairs, const int64_t* rmap,
                   const double* T, const double* VPr, const double* VPH, double* S,
                   void* stream) {
  if (npair < 0 || npair > INT32_MAX || max_rows < 0 || max_rows > RR) return NL_EARG;
  if (npair == 0) return NL_OK;
  if (!pairs || !rmap || !T || !VPr || !VPH || !S) return NL_EARG;
  hipLaunchKernelGGL(sector_ring_kernel, dim3((unsigned)npair), dim3(NT), 0, (hipStream_t)stream,
                     pairs, rmap, T, VPr, VPH, S);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_ring_adj(int64_t npair, int64_t max_rows, const int64_t* pairs,
                       const int64_t* rmap, const double* T, const double* VPr, const double* VPH,
                       const double* gS, double* gTr, void* stream) {
  if (npair < 0 || npair > INT32_MAX || max_rows < 0 || max_rows > RR) return NL_EARG;
  if (npair == 0) return NL_OK;
  if (!pairs || !rmap || !T || !VPr || !VPH || !gS || !gTr) return NL_EARG;
  hipLaunchKernelGGL(sector_ring_adj_kernel, dim3((unsigned)npair), dim3(NT), 0,
                     (hipStream_t)stream, pairs, rmap, T, VPr, VPH, gS, gTr);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_gather_add(int64_t n, const int64_t* idx, const double* in, double alpha,
                         double* out, void* stream) {
  if (n < 0) return NL_EARG;
  if (n == 0) return NL_OK;
  if (!idx || !in || !out) return NL_EARG;
  hipLaunchKernelGGL(sector_gather_add_kernel, dim3(grid_of(n, 8192)), dim3(NT), 0,
                     (hipStream_t)stream, n, idx, in, alpha, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_diag_axpy(int64_t n, int64_t p, const int* orb, const double* W, const double* G,
                        double alpha, double* out, void* stream) {
  if (n < 0 || p < 0) return NL_EARG;
  if (n == 0) return NL_OK;
  if (!orb || !W || !G || !out) return NL_EARG;
  hipLaunchKernelGGL(sector_diag_axpy_kernel, dim3(grid_of(n, 8192)), dim3(NT), 0,
                     (hipStream_t)stream, n, p, orb, W, G, alpha, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_contact_fill(int64_t nblk, const int64_t* off, const int64_t* ncol, const int64_t* roff,
                    const int64_t* coff, const unsigned char* rs, const unsigned char* cs,
                    int crossed, double g, int64_t total, double* out, void* stream) {
  if (nblk < 0 || total < 0) return NL_EARG;
  if (nblk == 0 || total == 0) return NL_OK;
  if (!off || !ncol || !roff || !coff || !rs || !cs || !out) return NL_EARG;
  hipLaunchKernelGGL(contact_fill_kernel, dim3(grid_of(total, 16384)), dim3(NT), 0,
                     (hipStream_t)stream, nblk, off, ncol, roff, coff, rs, cs, crossed, g, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_pw_dense(int64_t n0, int64_t n1, int64_t n2, int64_t n3, const int* c0, const int* c1,
                const int* c2, const int* c3, int L, double g, double* out, void* stream) {
  if (n0 < 0 || n1 < 0 || n2 < 0 || n3 < 0 || L < 1 || L > 255) return NL_EARG;
  if (n1 > INT32_MAX || n2 > INT32_MAX || n3 > INT32_MAX || n2 * n3 > INT32_MAX) return NL_EARG;
  const int64_t total = n0 * n1 * n2 * n3;
  if (total == 0) return NL_OK;
  if (!c0 || !c1 || !c2 || !c3 || !out) return NL_EARG;
  hipLaunchKernelGGL(pw_dense_kernel, dim3(grid_of(total, 16384)), dim3(NT), 0,
                     (hipStream_t)stream, n0, (int)n1, (int)n2, (int)n3, c0, c1, c2, c3, L, g,
                     out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_pw_pair_fill(int64_t nb, int64_t mb, int64_t kb, const int64_t* rows, const int64_t* cols,
                    int64_t rdiv, int64_t cdiv, const int* r0, const int* r1, const int* q0,
                    const int* q1, int L, double g, double* out, void* stream) {
  if (nb < 0 || mb < 0 || kb < 0 || rdiv < 1 || cdiv < 1 || L < 1 || L > 255) return NL_EARG;
  const int64_t total = nb * mb * kb;
  if (total == 0) return NL_OK;
  if (!rows || !cols || !r0 || !r1 || !q0 || !q1 || !out) return NL_EARG;
  PwPairArgs a{nb, mb, kb, rows, cols, rdiv, cdiv, r0, r1, q0, q1, L, g, out};
  hipLaunchKernelGGL(pw_pair_fill_kernel, dim3(grid_of(total, 16384)), dim3(NT), 0,
                     (hipStream_t)stream, a);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_pw_coo(int64_t ngrp, const int64_t* eoff, const int64_t* roff, const int64_t* coff,
              const int64_t* rpair, const int64_t* cpair, int64_t rdiv, int64_t cdiv,
              const int* r0, const int* r1, const int* q0, const int* q1, int L, double g,
              int64_t n, int64_t* idx, double* val, void* stream) {
  if (ngrp < 0 || n < 0 || rdiv < 1 || cdiv < 1 || L < 1 || L > 255) return NL_EARG;
  if (n == 0) return NL_OK;
  if (ngrp == 0 || !eoff || !roff || !coff || !rpair || !cpair || !r0 || !r1 || !q0 || !q1 ||
      !idx || !val)
    return NL_EARG;
  PwCooArgs a{ngrp, eoff, roff, coff, rpair, cpair, rdiv, cdiv, r0, r1, q0, q1, L, g, idx, val,
              n};
  hipLaunchKernelGGL(pw_coo_kernel, dim3(grid_of(n, 16384)), dim3(NT), 0, (hipStream_t)stream,
                     a);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_assemble(int64_t n, int64_t p, const int* orb, const int* sg, const double* V,
                       const double* L, const double* S, const double* X, const double* t,
                       double mixing, double escale, double* tn, double* H2out, double* part,
                       void* stream) {
  if (n < 0 || p < 0 || !orb || !V || !X || !t || !part || (S && !sg)) return NL_EARG;
  SectorAsm g{n, p, orb, sg, V, L, S, X, t, mixing, escale, tn, H2out, part};
  hipLaunchKernelGGL(sector_assemble_kernel, dim3((unsigned)nl_partials(n)), dim3(NT), 0,
                     (hipStream_t)stream, g);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

}  // extern "C"

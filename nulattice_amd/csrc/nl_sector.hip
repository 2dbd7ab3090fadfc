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
// tiles holds 3 int32 per workgroup: problem, m0, n0 (64 x 32 output tile).
constexpr int GTM = 64, GTN = 32, GTK = 16;

__global__ __launch_bounds__(NT) void gemm_grouped_f64_kernel(const int64_t* __restrict__ desc,
                                                              const int* __restrict__ tiles,
                                                              double alpha, const double* A,
                                                              const double* B, double beta,
                                                              const double* D, double* C) {
  __shared__ double As[GTK][GTM + 1];
  __shared__ double Bs[GTK][GTN + 1];
  const int q = tiles[3 * blockIdx.x], m0 = tiles[3 * blockIdx.x + 1],
            n0 = tiles[3 * blockIdx.x + 2];
  const int64_t* d = desc + 12 * (int64_t)q;
  const int64_t M = d[0], N = d[1], K = d[2];
  const int64_t oA = d[3], sAm = d[4], sAk = d[5], oB = d[6], sBk = d[7], sBn = d[8];
  const int64_t oC = d[9], sCm = d[10], sCn = d[11];
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const bool a_k1 = sAk == 1, b_n1 = sBn == 1;
  double acc[4][2];
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r][0] = acc[r][1] = 0.0;
  for (int64_t k0 = 0; k0 < K; k0 += GTK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + NT * r;
      const int kk = a_k1 ? (idx & 15) : (idx >> 6);
      const int mm = a_k1 ? (idx >> 4) : (idx & 63);
      const int64_t m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < M && k < K) ? A[oA + m * sAm + k * sAk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int idx = tid + NT * r;
      const int nn = b_n1 ? (idx & 31) : (idx >> 4);
      const int kk = b_n1 ? (idx >> 5) : (idx & 15);
      const int64_t n = n0 + nn, k = k0 + kk;
      Bs[kk][nn] = (n < N && k < K) ? B[oB + k * sBk + n * sBn] : 0.0;
    }
    __syncthreads();
#pragma unroll
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
  const double* Dp = D ? D : C;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t m = m0 + ty + 16 * r;
    if (m >= M) continue;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int64_t n = n0 + tx + 16 * c;
      if (n >= N) continue;
      const int64_t off = oC + m * sCm + n * sCn;
      double v = alpha * acc[r][c];
      if (beta != 0.0) v += beta * Dp[off];
      C[off] = v;
    }
  }
}

// ---------------------------------------------------------------- gathers
__global__ __launch_bounds__(NT) void sector_gather_kernel(int64_t n, const int64_t* __restrict__ idx,
                                                           const double* __restrict__ in,
                                                           double* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < n; t += (int64_t)gridDim.x * NT) {
    const int64_t s = idx[t];
    out[t] = s >= 0 ? in[s] : 0.0;
  }
}

__global__ __launch_bounds__(NT) void sector_scatter_kernel(int64_t n,
                                                            const int64_t* __restrict__ idx,
                                                            const double* __restrict__ in,
                                                            double* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < n; t += (int64_t)gridDim.x * NT) {
    const int64_t s = idx[t];
    if (s >= 0) out[s] = in[t];
  }
}

// ---------------------------------------------------------------- dressing
// X[o] = base[o] + alpha * sum_{e in segment o} V[elem[e]] * T[elem[e]]; one wavefront per o.
__global__ __launch_bounds__(NT) void sector_dress_kernel(int64_t no, const int64_t* __restrict__ seg,
                                                          const int* __restrict__ elem,
                                                          const double* __restrict__ V,
                                                          const double* __restrict__ T,
                                                          const double* __restrict__ base,
                                                          double alpha, double* __restrict__ X) {
  const int64_t o = blockIdx.x * (int64_t)(NT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (o >= no) return;
  double acc = 0.0;
  for (int64_t e = seg[o] + lane; e < seg[o + 1]; e += 64) {
    const int s = elem[e];
    acc = fma(V[s], T[s], acc);
  }
  acc = wave_sum(acc);
  if (lane == 0) X[o] = base[o] + alpha * acc;
}

// ---------------------------------------------------------------- assembly
struct SectorAsm {
  int64_t n, p;
  const int* orb;        // [4][n]: a, b (particles), i, j (holes) of packed element e
  const int* sg;         // [4][n]: ring-image offsets of S[a,i,b,j], S[a,j,b,i], S[b,i,a,j],
                         //         S[b,j,a,i]; null = no ring term
  const double *V, *L, *S, *X, *t;
  double mixing;
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
    e_acc += 0.25 * v * tn;
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

int nl_gemm_grouped_f64(int64_t ntile, const int64_t* desc, const int* tiles, double alpha,
                        const double* A, const double* B, double beta, const double* D, double* C,
                        void* stream) {
  if (ntile < 0 || ntile > INT32_MAX) return NL_EARG;
  if (ntile == 0) return NL_OK;
  if (!desc || !tiles || !A || !B || !C) return NL_EARG;
  hipLaunchKernelGGL(gemm_grouped_f64_kernel, dim3((unsigned)ntile), dim3(NT), 0,
                     (hipStream_t)stream, desc, tiles, alpha, A, B, beta, D, C);
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

int nl_sector_dress(int64_t no, const int64_t* seg, const int* elem, const double* V,
                    const double* T, const double* base, double alpha, double* X, void* stream) {
  if (no < 0) return NL_EARG;
  if (no == 0) return NL_OK;
  if (!seg || !V || !T || !base || !X) return NL_EARG;
  const int64_t grid = nl_cdiv(no, NT / 64);
  if (grid > INT32_MAX) return NL_EARG;
  hipLaunchKernelGGL(sector_dress_kernel, dim3((unsigned)grid), dim3(NT), 0, (hipStream_t)stream,
                     no, seg, elem, V, T, base, alpha, X);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sector_assemble(int64_t n, int64_t p, const int* orb, const int* sg, const double* V,
                       const double* L, const double* S, const double* X, const double* t,
                       double mixing, double* tn, double* H2out, double* part, void* stream) {
  if (n < 0 || p < 0 || !orb || !V || !X || !t || !part || (S && !sg)) return NL_EARG;
  SectorAsm g{n, p, orb, sg, V, L, S, X, t, mixing, tn, H2out, part};
  hipLaunchKernelGGL(sector_assemble_kernel, dim3((unsigned)nl_partials(n)), dim3(NT), 0,
                     (hipStream_t)stream, g);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

}  // extern "C"

// MFMA GEMMs for the CCSD contractions on gfx950.
//
//  * nl_gemm_f64: C = alpha * op(A) op(B) + beta * C in FP64 on v_mfma_f64_16x16x4f64.
//    Batched and strided; A may be K- or M-contiguous, B may be N- or K-contiguous, so most
//    tensor contractions map onto it after at most one permute.  Split-K with a deterministic
//    workspace reduction serves the tall-K shapes (Fock dressings, hh-ladder intermediates).
//  * nl_gemm_lp: C (FP64) = alpha * diag(ia) (A B^T) diag(ib) + beta * C with A, B in BF16 or
//    FP16 (K-contiguous rows, per-row power-of-two scales folded out through ia/ib) on
//    v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulation.
//
// Both use a 128x128 workgroup tile, 4 waves of 64x64, LDS staging with a register prefetch of
// the next K-tile, and an XCD-aware tile order: consecutive workgroup ids land on different XCDs,
// so the id is remapped such that each XCD walks a contiguous band of tiles (shared A rows / B
// columns stay in that XCD's L2), then grouped by 8 tile-rows.
#include "nl_common.h"

#include <type_traits>

namespace {

constexpr int TM = 128, TN = 128;

struct TileCoord { int tm, tn; };

__device__ inline TileCoord tile_of(int wg, int nwg, int tiles_m, int tiles_n) {
  // XCD-aware remap: hardware dispatches workgroup ids round-robin over the 8 XCDs.
  if ((nwg & 7) == 0) wg = (wg & 7) * (nwg >> 3) + (wg >> 3);
  const int G = 8;
  int per_group = G * tiles_n;
  int gid = wg / per_group;
  int first_m = gid * G;
  int gsize = min(tiles_m - first_m, G);
  int in = wg % per_group;
  TileCoord c;
  c.tm = first_m + in % gsize;
  c.tn = in / gsize;
  return c;
}

// ---------------------------------------------------------------------------------- FP64
// Strided tensor contraction: every index group is up to two runs with arbitrary strides,
//   m = m1*M2 + m2, n = n1*N2 + n2, k = k1*K2 + k2, batch = b1*B2 + b2,
//   A(m,k) = A[m1*am1 + m2*am2 + k1*ak1 + k2*ak2 (+ b1*ab1 + b2*ab2)], B and C alike,
// so a 4-index contraction maps onto the MFMA tiles without a transpose.
constexpr int BK64 = 16;
constexpr int PAD64 = 4;

// Extent of the inner run and strides of the outer / inner run, with the multiply-shift divisor of
// the extent (as FastDiv: i / n2 = (i * mg) >> sh for 0 <= i < 2^31).
struct Idx2 { int n2; int64_t s1, s2; uint64_t mg; int sh; };

__device__ inline int div2(int i, const Idx2& d) {
  return (int)(((uint64_t)(uint32_t)i * d.mg) >> d.sh);
}

__device__ inline int64_t off2(int i, const Idx2& d) {
  int q = div2(i, d);
  return q * d.s1 + (int64_t)(i - q * d.n2) * d.s2;
}

// Walks consecutive indices i, i+1, ... of a two-run index, carrying into the outer run.
struct Walk2 {
  int q, r;
  int64_t off;
  __device__ inline Walk2(int i, const Idx2& d) {
    q = div2(i, d); r = i - q * d.n2; off = q * d.s1 + (int64_t)r * d.s2;
  }
  __device__ inline void next(const Idx2& d) {
    ++r; off += d.s2;
    if (r == d.n2) { r = 0; ++q; off = q * d.s1; }
  }
};

struct GemmF64Args {
  int M, N, K, nbatch;
  double alpha, beta;
  const double* A; Idx2 am, ak, ab;
  const double* B; Idx2 bk, bn, bb;
  AnyOut C; Idx2 cm, cn, cb;                // FP64 or FP32 output
  AnyIn D;                                  // addend (same layout as C); may equal C or be null
  int tiles_m, tiles_n, splits, kchunk;
  double* ws;  // split-K partials [batch][splits][M][N] (only when splits > 1)
};

// AK: a thread loads 8 consecutive k of one row (A unit-stride in k); else 8 consecutive m.
// BNC: a thread loads 8 consecutive n of one k-row (B unit-stride in n); else 8 consecutive k.
template <bool AK, bool BNC>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmF64Args g) {
  __shared__ double As[BK64][TM + PAD64];
  __shared__ double Bs[BK64][TN + PAD64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int batch = blockIdx.z, split = blockIdx.y;
  TileCoord tc = tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n);
  const int m0 = tc.tm * TM, n0 = tc.tn * TN;
  const double* A = g.A + off2(batch, g.ab);
  const double* B = g.B + off2(batch, g.bb);
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);

  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

  // per-thread fixed row (AK) / column (BNC) offsets
  const int ar = AK ? (tid >> 1) : 0;
  const int bc = BNC ? 0 : (tid >> 1);
  const int64_t a_row = (AK && m0 + ar < g.M) ? off2(m0 + ar, g.am) : 0;
  const int64_t b_col = (!BNC && n0 + bc < g.N) ? off2(n0 + bc, g.bn) : 0;

  // per-thread index walkers: the m (or n) run of a non-AK (BNC) thread is fixed per kernel
  const int am0 = AK ? 0 : m0 + (tid & 15) * 8;
  const int bn0 = BNC ? n0 + (tid & 15) * 8 : 0;
  const Walk2 wa(am0, g.am), wb(bn0, g.bn);

  double ra[8], rb[8];
  auto load = [&](int k0) {
    if (AK) {
      int k = k0 + (tid & 1) * 8, m = m0 + ar;
      Walk2 wk(k, g.ak);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        ra[e] = (m < g.M && k + e < kend) ? A[a_row + wk.off] : 0.0;
        wk.next(g.ak);
      }
    } else {
      int k = k0 + (tid >> 4);
      int64_t ko = (k < kend) ? off2(k, g.ak) : 0;
      Walk2 wm = wa;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        ra[e] = (am0 + e < g.M && k < kend) ? A[ko + wm.off] : 0.0;
        wm.next(g.am);
      }
    }
    if (BNC) {
      int k = k0 + (tid >> 4);
      int64_t ko = (k < kend) ? off2(k, g.bk) : 0;
      Walk2 wn = wb;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        rb[e] = (bn0 + e < g.N && k < kend) ? B[ko + wn.off] : 0.0;
        wn.next(g.bn);
      }
    } else {
      int k = k0 + (tid & 1) * 8, n = n0 + bc;
      Walk2 wk(k, g.bk);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        rb[e] = (n < g.N && k + e < kend) ? B[b_col + wk.off] : 0.0;
        wk.next(g.bk);
      }
    }
  };
  auto store = [&]() {
    if (AK) {
      int kk = (tid & 1) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) As[kk + e][ar] = ra[e];
    } else {
      int kk = tid >> 4, r = (tid & 15) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) As[kk][r + e] = ra[e];
    }
    if (BNC) {
      int kk = tid >> 4, c = (tid & 15) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) Bs[kk][c + e] = rb[e];
    } else {
      int kk = (tid & 1) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) Bs[kk + e][bc] = rb[e];
    }
  };

  if (kbeg < kend) {
    load(kbeg);
    store();
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += BK64) {
      const bool more = k0 + BK64 < kend;
      if (more) load(k0 + BK64);
#pragma unroll
      for (int kk = 0; kk < BK64 / 4; ++kk) {
        const int kr = kk * 4 + (lane >> 4);
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      __syncthreads();
      if (more) {
        store();
        __syncthreads();
      }
    }
  }

  // f64 16x16x4 accumulator layout: register r of lane l holds D[4r + l/16][l%16]
  if (g.splits > 1) {
    double* W = g.ws + ((int64_t)batch * g.splits + split) * (int64_t)g.M * g.N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = m0 + wm * 64 + i * 16 + r * 4 + (lane >> 4);
          int col = n0 + wn * 64 + j * 16 + (lane & 15);
          if (row < g.M && col < g.N) W[(int64_t)row * g.N + col] = g.alpha * acc[i][j][r];
        }
    return;
  }
  const int64_t cb = off2(batch, g.cb);
  const bool addend = g.D && g.beta != 0.0;
  int64_t coff[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int col = n0 + wn * 64 + j * 16 + (lane & 15);
    coff[j] = col < g.N ? off2(col, g.cn) : -1;
  }
  // per 16-row group: all addend loads first, then the stores (D may be C; each element is read
  // and written by this lane only), so the loads overlap instead of waiting on each store
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int64_t ro[4];
    double dv[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + wm * 64 + i * 16 + r * 4 + (lane >> 4);
      ro[r] = row < g.M ? cb + off2(row, g.cm) : -1;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        dv[r][j] = (addend && ro[r] >= 0 && coff[j] >= 0) ? g.D.get(ro[r] + coff[j]) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (ro[r] < 0) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (coff[j] >= 0) g.C.put(ro[r] + coff[j], g.alpha * acc[i][j][r] + g.beta * dv[r][j]);
    }
  }
}

// Skinny contraction (N <= NB <= 64): one output row per thread, the N outputs in registers,
// K streamed through LDS in tiles of 16 (A tile 256 x 16 loaded along its unit-stride index).
// Bandwidth-bound shapes (Fock dressings, h x h products, matrix-vector terms) run here instead
// of on 128x128 MFMA tiles that would be mostly padding.
constexpr int SKM = 256, SKK = 16;

template <int NB, bool AM>   // AM: A unit-stride along m (threads along m), else along k
__global__ __launch_bounds__(256) void skinny_f64_kernel(GemmF64Args g) {
  __shared__ double As[SKK][SKM + 1];
  __shared__ double Bs[SKK][NB];
  const int tid = threadIdx.x;
  const int batch = blockIdx.z, split = blockIdx.y;
  const int m0 = blockIdx.x * SKM;
  const double* A = g.A + off2(batch, g.ab);
  const double* B = g.B + off2(batch, g.bb);
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  double acc[NB];
#pragma unroll
  for (int n = 0; n < NB; ++n) acc[n] = 0.0;
  // AM mapping: thread owns row m0+tid for the loads too; else thread loads k = tid%16 of rows
  // m0 + tid/16 + 16e (e < 16), whose offsets are fixed for the whole kernel.
  const int64_t arow = (AM && m0 + tid < g.M) ? off2(m0 + tid, g.am) : 0;
  int64_t rowoff[AM ? 1 : 16];
  if (!AM) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      int m = m0 + tid / 16 + 16 * e;
      rowoff[e] = m < g.M ? off2(m, g.am) : 0;
    }
  }
  // register prefetch: the next tile's loads are in flight while this tile is multiplied
  constexpr int NBQ = (SKK * NB + 255) / 256;
  double ra[16], rb[NBQ];
  auto load = [&](int k0) {
    if (AM) {
      Walk2 wk(k0, g.ak);
#pragma unroll
      for (int e = 0; e < SKK; ++e) {
        ra[e] = (m0 + tid < g.M && k0 + e < kend) ? A[arow + wk.off] : 0.0;
        wk.next(g.ak);
      }
    } else {
      const int k = k0 + (tid & 15);
      const int64_t ko = k < kend ? off2(k, g.ak) : 0;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int mm = tid / 16 + 16 * e;
        ra[e] = (m0 + mm < g.M && k < kend) ? A[rowoff[e] + ko] : 0.0;
      }
    }
#pragma unroll
    for (int j = 0; j < NBQ; ++j) {
      const int q = tid + 256 * j, kk = q / NB, n = q - kk * NB, k = k0 + kk;
      rb[j] = (q < SKK * NB && n < g.N && k < kend) ? B[off2(k, g.bk) + off2(n, g.bn)] : 0.0;
    }
  };
  auto store = [&]() {
    if (AM) {
#pragma unroll
      for (int e = 0; e < SKK; ++e) As[e][tid] = ra[e];
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) As[tid & 15][tid / 16 + 16 * e] = ra[e];
    }
#pragma unroll
    for (int j = 0; j < NBQ; ++j) {
      const int q = tid + 256 * j;
      if (q < SKK * NB) Bs[q / NB][q % NB] = rb[j];
    }
  };
  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += SKK) {
    store();
    __syncthreads();
    if (k0 + SKK < kend) load(k0 + SKK);
#pragma unroll
    for (int kk = 0; kk < SKK; ++kk) {
      const double a = As[kk][tid];
#pragma unroll
      for (int n = 0; n < NB; ++n) acc[n] = fma(a, Bs[kk][n], acc[n]);
    }
    __syncthreads();
  }
  // Outputs leave through LDS in SR-row slices (the A tile's storage is free now): a thread
  // holds one row, but the stores run with consecutive lanes on consecutive columns, so rows of
  // a row-major output (and the split-K partials) are written as contiguous runs.
  constexpr int SR = NB <= 32 ? 64 : 32;
  static_assert(SR * (NB + 1) <= SKK * (SKM + 1), "output staging must fit in the A tile");
  double (*St)[NB + 1] = reinterpret_cast<double (*)[NB + 1]>(&As[0][0]);
  const bool partial = g.splits > 1;
  double* W = partial ? g.ws + ((int64_t)batch * g.splits + split) * (int64_t)g.M * g.N : nullptr;
  const int64_t cbo = off2(batch, g.cb);
  for (int sl = 0; sl < SKM / SR; ++sl) {
    __syncthreads();
    if (tid / SR == sl) {
#pragma unroll
      for (int n = 0; n < NB; ++n) St[tid % SR][n] = acc[n];
    }
    __syncthreads();
    for (int q = tid; q < SR * NB; q += 256) {
      const int r = q / NB, n = q - r * NB, m = m0 + sl * SR + r;
      if (m >= g.M || n >= g.N) continue;
      double v = g.alpha * St[r][n];
      if (partial) {
        W[(int64_t)m * g.N + n] = v;
      } else {
        const int64_t o = cbo + off2(m, g.cm) + off2(n, g.cn);
        if (g.D && g.beta != 0.0) v += g.beta * g.D.get(o);
        g.C.put(o, v);
      }
    }
  }
}

// Small output (M, N <= 64): all M*N outputs of a K-chunk in one workgroup, each thread owning
// up to 16 of them; K streamed through LDS in tiles of 32 with coalesced tile loads.  Serves the
// tall-K h x h dressings (sum over p*p*h elements) at streaming bandwidth.
constexpr int SMK = 32;

template <bool AM, bool BN>   // AM: A unit-stride along m; BN: B unit-stride along n
__global__ __launch_bounds__(256) void smallmn_f64_kernel(GemmF64Args g) {
  __shared__ double As[SMK][65];
  __shared__ double Bs[SMK][65];
  const int tid = threadIdx.x;
  const int batch = blockIdx.z, split = blockIdx.y;
  const double* A = g.A + off2(batch, g.ab);
  const double* B = g.B + off2(batch, g.bb);
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const int MN = g.M * g.N;
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  for (int k0 = kbeg; k0 < kend; k0 += SMK) {
    // all tile loads are issued before any LDS store, so their latencies overlap
    constexpr int SQ = SMK * 64 / 256;
    double ra[SQ], rb[SQ];
#pragma unroll
    for (int u = 0; u < SQ; ++u) {
      const int q = tid + 256 * u;
      int k, m, n;
      if (AM) { k = q >> 6; m = q & 63; } else { m = q / SMK; k = q - m * SMK; }
      ra[u] = (m < g.M && k0 + k < kend) ? A[off2(m, g.am) + off2(k0 + k, g.ak)] : 0.0;
      if (BN) { k = q >> 6; n = q & 63; } else { n = q / SMK; k = q - n * SMK; }
      rb[u] = (n < g.N && k0 + k < kend) ? B[off2(k0 + k, g.bk) + off2(n, g.bn)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SQ; ++u) {
      const int q = tid + 256 * u;
      int k, m, n;
      if (AM) { k = q >> 6; m = q & 63; } else { m = q / SMK; k = q - m * SMK; }
      As[k][m] = ra[u];
      if (BN) { k = q >> 6; n = q & 63; } else { n = q / SMK; k = q - n * SMK; }
      Bs[k][n] = rb[u];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      int o = tid + 256 * j;
      if (o < MN) {
        int m = o / g.N, n = o - m * g.N;
        double s = acc[j];
#pragma unroll 8
        for (int kk = 0; kk < SMK; ++kk) s = fma(As[kk][m], Bs[kk][n], s);
        acc[j] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    int o = tid + 256 * j;
    if (o >= MN) continue;
    int m = o / g.N, n = o - m * g.N;
    if (g.splits > 1) {
      g.ws[((int64_t)batch * g.splits + split) * (int64_t)MN + o] = g.alpha * acc[j];
    } else {
      const int64_t off = off2(batch, g.cb) + off2(m, g.cm) + off2(n, g.cn);
      double v = g.alpha * acc[j];
      if (g.D && g.beta != 0.0) v += g.beta * g.D.get(off);
      g.C.put(off, v);
    }
  }
}

// Small output with M, N <= 32 (the h x h dressings): each thread owns a 4x4 register block of
// the 32x32 output, so a k step costs 8 LDS reads for 16 FMAs (the kernel above reads two LDS
// values per FMA), and the four 64-thread groups of the workgroup take interleaved quarters of
// each 32-wide K tile; their partial blocks are added in a fixed order at the end.
template <bool AM, bool BN>
__global__ __launch_bounds__(256) void smallmn32_f64_kernel(GemmF64Args g) {
  __shared__ double As[SMK][33];
  __shared__ double Bs[SMK][33];
  __shared__ double Red[3][32][33];
  const int tid = threadIdx.x, grp = tid >> 6, lt = tid & 63;
  const int bm = (lt >> 3) * 4, bn = (lt & 7) * 4;
  const int batch = blockIdx.z, split = blockIdx.y;
  const double* A = g.A + off2(batch, g.ab);
  const double* B = g.B + off2(batch, g.bb);
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  // load mapping (4 elements per thread per tile and operand): along the unit-stride index
  const int lo = tid & 31, hi = tid >> 5;                  // hi + 8*it, it < 4
  int64_t amo[4], bno[4];                                  // fixed m / n offsets
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const int m = AM ? lo : hi + 8 * it, n = BN ? lo : hi + 8 * it;
    amo[it] = m < g.M ? off2(m, g.am) : 0;
    bno[it] = n < g.N ? off2(n, g.bn) : 0;
  }
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
  // register prefetch: the next K tile's loads are in flight while this one is multiplied
  double ra[4], rb[4];
  auto load = [&](int k0) {
    const int64_t kfix = (k0 + lo < kend) ? off2(k0 + lo, g.ak) : 0;
    const int64_t kfixb = (k0 + lo < kend) ? off2(k0 + lo, g.bk) : 0;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int kv = hi + 8 * it;                          // k of the "along m/n" mappings
      if (AM) {
        const bool ok = lo < g.M && k0 + kv < kend;
        ra[it] = ok ? A[amo[it] + off2(k0 + kv, g.ak)] : 0.0;
      } else {
        const bool ok = kv < g.M && k0 + lo < kend;
        ra[it] = ok ? A[amo[it] + kfix] : 0.0;
      }
      if (BN) {
        const bool ok = lo < g.N && k0 + kv < kend;
        rb[it] = ok ? B[off2(k0 + kv, g.bk) + bno[it]] : 0.0;
      } else {
        const bool ok = kv < g.N && k0 + lo < kend;
        rb[it] = ok ? B[kfixb + bno[it]] : 0.0;
      }
    }
  };
  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += SMK) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int kv = hi + 8 * it;
      if (AM) As[kv][lo] = ra[it]; else As[lo][kv] = ra[it];
      if (BN) Bs[kv][lo] = rb[it]; else Bs[lo][kv] = rb[it];
    }
    __syncthreads();
    if (k0 + SMK < kend) load(k0 + SMK);
#pragma unroll
    for (int kq = 0; kq < SMK / 4; ++kq) {
      const int kk = kq * 4 + grp;
      double a[4], b[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) { a[x] = As[kk][bm + x]; b[x] = Bs[kk][bn + x]; }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = fma(a[x], b[y], acc[x][y]);
    }
    __syncthreads();
  }
  if (grp > 0) {
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) Red[grp - 1][bm + x][bn + y] = acc[x][y];
  }
  __syncthreads();
  if (grp != 0) return;
  const int MN = g.M * g.N;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int m = bm + x, n = bn + y;
      if (m >= g.M || n >= g.N) continue;
      const double v = g.alpha * (((acc[x][y] + Red[0][m][n]) + Red[1][m][n]) + Red[2][m][n]);
      if (g.splits > 1) {
        g.ws[((int64_t)batch * g.splits + split) * (int64_t)MN + (int64_t)m * g.N + n] = v;
      } else {
        const int64_t off = off2(batch, g.cb) + off2(m, g.cm) + off2(n, g.cn);
        g.C.put(off, (g.D && g.beta != 0.0) ? v + g.beta * g.D.get(off) : v);
      }
    }
}

__device__ inline double wave_sum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// C[b](m,n) = sum_s ws[b][s][m][n] + beta * D[b](m,n) with one wavefront per output: lanes take
// the splits in a fixed interleave and a shuffle tree adds them (for large split counts, where
// a thread-per-output loop would walk thousands of dependent loads).
__global__ __launch_bounds__(256) void splitk_reduce_wave_kernel(int M, int N, int splits,
                                                                 int nbatch, const double* ws,
                                                                 double beta, AnyIn D,
                                                                 AnyOut C, Idx2 cm, Idx2 cn,
                                                                 Idx2 cb) {
  const int lane = threadIdx.x & 63;
  const int64_t MN = (int64_t)M * N, total = (int64_t)nbatch * MN;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t t = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; t < total; t += nw) {
    const int64_t b = t / MN, mn = t - b * MN;
    const double* W = ws + b * splits * MN + mn;
    double s = 0.0;
    for (int q = lane; q < splits; q += 64) s += W[(int64_t)q * MN];
    s = wave_sum64(s);
    if (lane == 0) {
      const int64_t m = mn / N, n = mn - m * N;
      const int64_t o = off2((int)b, cb) + off2((int)m, cm) + off2((int)n, cn);
      C.put(o, (D && beta != 0.0) ? s + beta * D.get(o) : s);
    }
  }
}

// Low-K products (K <= 32 with M, N > 64: outer-product-like accumulations into T2-sized
// outputs).  Bandwidth lies in reading and writing C, so a workgroup stages the whole K of a 64x64
// output tile's A and B slices in LDS once, and each thread owns 4 x 4 outputs: 4 consecutive
// along C's unit-stride index (LN: n, else m), so its addend loads and stores are one 16-byte
// (FP32) or two 16-byte (FP64) vector accesses where the four are contiguous and aligned, and
// 4 spaced 16 apart along the other index; 16 lanes cover 64 consecutive elements.
constexpr int LKT = 64, LKK = 32;

template <typename T>   // 4 consecutive elements at p (16-byte aligned when T is float)
__device__ inline void ld4(const T* p, double (&v)[4]) {
  if constexpr (sizeof(T) == 4) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
    const double2 q0 = reinterpret_cast<const double2*>(p)[0];
    const double2 q1 = reinterpret_cast<const double2*>(p)[1];
    v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y;
  }
}
template <typename T>
__device__ inline void st4(T* p, const double (&v)[4]) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
  } else {
    reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
  }
}

template <bool LN, typename TC, typename TD>   // TC, TD: element types of C and D
__global__ __launch_bounds__(256) void lowk_f64_kernel(GemmF64Args g) {
  TC* const Cp = static_cast<TC*>(g.C.p);
  const TD* const Dp = static_cast<const TD*>(g.D.p);
  __shared__ double As[LKK][LKT + 2];
  __shared__ double Bs[LKK][LKT + 2];
  const int tid = threadIdx.x;
  const int batch = blockIdx.z;
  TileCoord tc = tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n);
  const int m0 = tc.tm * LKT, n0 = tc.tn * LKT;
  const double* A = g.A + off2(batch, g.ab);
  const double* B = g.B + off2(batch, g.bb);
  const int K = g.K;
  // operand tiles: along k when that is the operand's unit stride, else along m / n
  const bool ak1 = g.ak.n2 > 1 ? g.ak.s2 == 1 : g.ak.s1 == 1;
  const bool bk1 = g.bk.n2 > 1 ? g.bk.s2 == 1 : g.bk.s1 == 1;
  // tile rows / columns of this thread's outputs
  const int tu = tid & 15, tv = tid >> 4;
  int rm[4], rn[4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    rm[x] = LN ? tv + 16 * x : 4 * tu + x;
    rn[x] = LN ? 4 * tu + x : tv + 16 * x;
  }
  // output offsets and the addend are read before the operand staging, so their latency overlaps
  // it (each output is read and written by this thread only, so D may be C).  Group q holds the
  // four outputs consecutive along C's unit-stride index: (row x = q, col y = t) for LN, else
  // (row x = t, col y = q).
  const int64_t cb = off2(batch, g.cb);
  int64_t roff[4], coff[4];
  bool rok[4], cok[4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    rok[x] = m0 + rm[x] < g.M;
    cok[x] = n0 + rn[x] < g.N;
    roff[x] = rok[x] ? cb + off2(m0 + rm[x], g.cm) : 0;
    coff[x] = cok[x] ? off2(n0 + rn[x], g.cn) : 0;
  }
  auto off = [&](int q, int t) { return LN ? roff[q] + coff[t] : roff[t] + coff[q]; };
  auto ok = [&](int q, int t) { return LN ? rok[q] && cok[t] : rok[t] && cok[q]; };
  const bool addend = g.D && g.beta != 0.0;
  // a group is one vector access when its four offsets are consecutive, all four in range and
  // the addresses in C (and D) 16-byte aligned
  bool vec[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    vec[q] = ok(q, 3) && (reinterpret_cast<uintptr_t>(Cp + off(q, 0)) & 15) == 0 &&
             (!addend || (reinterpret_cast<uintptr_t>(Dp + off(q, 0)) & 15) == 0);
#pragma unroll
    for (int t = 1; t < 4; ++t) vec[q] = vec[q] && ok(q, t) && off(q, t) == off(q, 0) + t;
  }
  double dv[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (addend && vec[q]) {
      ld4(Dp + off(q, 0), dv[q]);
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) dv[q][t] = (addend && ok(q, t)) ? (double)Dp[off(q, t)] : 0.0;
    }
  }
  // all staging loads are issued before any LDS store, so their latencies overlap; lanes run
  // along each operand's unit-stride index (k: 32 lanes x 8 rows, else 64 lanes x 4 k)
  constexpr int LQ = LKK * LKT / 256;
  double ra[LQ], rb[LQ];
#pragma unroll
  for (int u = 0; u < LQ; ++u) {
    const int ka = ak1 ? (tid & 31) : (tid >> 6) + 4 * u, m = ak1 ? (tid >> 5) + 8 * u : tid & 63;
    const int kb = bk1 ? (tid & 31) : (tid >> 6) + 4 * u, n = bk1 ? (tid >> 5) + 8 * u : tid & 63;
    ra[u] = (ka < K && m0 + m < g.M) ? A[off2(m0 + m, g.am) + off2(ka, g.ak)] : 0.0;
    rb[u] = (kb < K && n0 + n < g.N) ? B[off2(kb, g.bk) + off2(n0 + n, g.bn)] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < LQ; ++u) {
    const int ka = ak1 ? (tid & 31) : (tid >> 6) + 4 * u, m = ak1 ? (tid >> 5) + 8 * u : tid & 63;
    const int kb = bk1 ? (tid & 31) : (tid >> 6) + 4 * u, n = bk1 ? (tid >> 5) + 8 * u : tid & 63;
    As[ka][m] = ra[u];
    Bs[kb][n] = rb[u];
  }
  __syncthreads();
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
  for (int k = 0; k < K; ++k) {
    double a[4], b[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) { a[x] = As[k][rm[x]]; b[x] = Bs[k][rn[x]]; }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) acc[x][y] = fma(a[x], b[y], acc[x][y]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double v[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      v[t] = g.alpha * (LN ? acc[q][t] : acc[t][q]);
      if (addend) v[t] += g.beta * dv[q][t];
    }
    if (vec[q]) {
      st4(Cp + off(q, 0), v);
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (ok(q, t)) Cp[off(q, t)] = (TC)v[t];
    }
  }
}

// C[b](m,n) = sum_s ws[b][s][m][n] + beta * D[b](m,n)
__global__ void splitk_reduce_strided_kernel(int M, int N, int splits, int nbatch,
                                             const double* ws, double beta, AnyIn D,
                                             AnyOut C, Idx2 cm, Idx2 cn, Idx2 cb) {
  int64_t total = (int64_t)nbatch * M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = t / ((int64_t)M * N);
    int64_t mn = t - b * (int64_t)M * N;
    int64_t m = mn / N, n = mn - m * N;
    const double* W = ws + b * splits * (int64_t)M * N + mn;
    double s = 0.0;
    for (int q = 0; q < splits; ++q) s += W[(int64_t)q * M * N];
    int64_t o = off2(b, cb) + off2(m, cm) + off2(n, cn);
    if (D && beta != 0.0) s += beta * D.get(o);
    C.put(o, s);
  }
}


// ---------------------------------------------------------------------------------- BF16/FP16
constexpr int BKL = 64;
constexpr int PADL = 8;

struct GemmLPArgs {
  int M, N, K;
  double alpha, beta;
  const uint16_t* A; int64_t lda, sA;   // [M][lda] K-contiguous, 16-bit elements
  const uint16_t* B; int64_t ldb, sB;   // [N][ldb] K-contiguous
  const uint16_t* Alo;                  // split operands: lo parts with A's / B's layout (or
  const uint16_t* Blo;                  // null); the product is hi.hi + hi.lo + lo.hi
  const float* ia; int64_t sia;         // per-row inverse scale of A (or null)
  const float* ib; int64_t sib;         // per-row inverse scale of B (or null)
  AnyOut C; Idx2 cm, cn; int64_t sC;   // C(m,n) = C[off2(m,cm) + off2(n,cn)] (+ batch * sC),
                                        // FP64 or FP32
  AnyIn D;                              // addend with C's layout (may be C or null)
  int tiles_m, tiles_n, splits, kchunk;
  double* ws;                           // split-K partials [batch][splits][M][N]
  const int64_t* crow;                  // row scatter: row m of batch b goes to C row
                                        // crow[b*M + m] (< 0: dropped), no batch offset
};

// Epilogue of the 16-bit GEMMs: C = alpha ia[m] acc ib[n] + beta D for this wave's NI x 4 MFMA
// tiles (rows m0 + wm*16*NI + i*16 + (lane/16)*4 + r, columns n0 + wn*64 + j*16 + lane%16), into
// the split-K partials, scattered rows (crow) or C's two-run layout.  The element types of C and
// D are dispatched once, outside the loops, so the addend loads of a row group issue back to
// back (a type test per access would split them into dependent blocks).
template <int NI>
__device__ __attribute__((always_inline)) inline void lp_epilogue(const GemmLPArgs& g, const v4f (&acc)[NI][4], int m0, int n0,
                                   int wm, int wn, int lane, int batch, int split) {
  const bool partial = g.splits > 1;
  const int64_t cbo = g.crow ? 0 : (int64_t)batch * g.sC;
  const double beta = partial ? 0.0 : g.beta;
  const bool addend = !partial && g.D && beta != 0.0;
  const float* ia = g.ia ? g.ia + (int64_t)batch * g.sia : nullptr;
  const float* ib = g.ib ? g.ib + (int64_t)batch * g.sib : nullptr;
  double sc[4];
  int64_t coff[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int col = n0 + wn * 64 + j * 16 + (lane & 15);
    sc[j] = (ib && col < g.N) ? (double)ib[col] : 1.0;
    coff[j] = col < g.N ? (partial ? col : off2(col, g.cn)) : -1;
  }
  // (always inlined: a call would pass the accumulators through scratch memory)
  auto body = [&](auto* C, const auto* D) __attribute__((always_inline)) {
    using TC = std::remove_pointer_t<decltype(C)>;
    // per 16-row group: all addend loads first, then the stores (D may be C; each element is
    // read and written by this lane only), so the loads overlap instead of waiting on each store
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      int64_t ro[4];
      double dv[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 16 * NI + i * 16 + (lane >> 4) * 4 + r;
        ro[r] = row >= g.M ? -1
                : partial ? (int64_t)row * g.N
                : g.crow ? (g.crow[(int64_t)batch * g.M + row] >= 0
                                ? g.crow[(int64_t)batch * g.M + row] * g.cm.s1 : -1)
                : off2(row, g.cm);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          dv[r][j] = (addend && ro[r] >= 0 && coff[j] >= 0) ? (double)D[ro[r] + coff[j]] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (ro[r] < 0) continue;
        const int row = m0 + wm * 16 * NI + i * 16 + (lane >> 4) * 4 + r;
        const double sr = g.alpha * (ia ? (double)ia[row] : 1.0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (coff[j] >= 0)
            C[ro[r] + coff[j]] = (TC)(sr * (double)acc[i][j][r] * sc[j] + beta * dv[r][j]);
      }
    }
  };
  if (partial) {
    body(g.ws + ((int64_t)batch * g.splits + split) * (int64_t)g.M * g.N, (const double*)nullptr);
  } else {
    const AnyOut C = g.C.at(cbo);
    const AnyIn D = g.D ? g.D.at(cbo) : AnyIn{nullptr, false};
    if (C.f32) {
      if (D.f32) body(static_cast<float*>(C.p), static_cast<const float*>(D.p));
      else body(static_cast<float*>(C.p), static_cast<const double*>(D.p));
    } else {
      if (D.f32) body(static_cast<double*>(C.p), static_cast<const float*>(D.p));
      else body(static_cast<double*>(C.p), static_cast<const double*>(D.p));
    }
  }
}

template <bool IS_BF16>
__global__ __launch_bounds__(256) void gemm_lp_kernel(GemmLPArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[TM][BKL + PADL];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][BKL + PADL];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int batch = blockIdx.z;
  TileCoord tc = tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n);
  const int m0 = tc.tm * TM, n0 = tc.tn * TN;
  const uint16_t* A = g.A + (int64_t)batch * g.sA;
  const uint16_t* B = g.B + (int64_t)batch * g.sB;
  const int split = blockIdx.y;
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};

  // 128 rows x 64 k = 1024 chunks of 8 elements (16 B); 4 chunks per thread per operand.
  uint4 ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int c = tid + 256 * q, r = c >> 3, kc = (c & 7) * 8, k = k0 + kc;
      int m = m0 + r, n = n0 + r;
      // K, kchunk, lda and ldb are multiples of 8 (checked / arranged by the host entry point),
      // so a chunk is either fully inside or fully outside [kbeg, kend).
      ra[q] = (m < g.M && k < kend) ? *(const uint4*)(A + (int64_t)m * g.lda + k)
                                    : make_uint4(0, 0, 0, 0);
      rb[q] = (n < g.N && k < kend) ? *(const uint4*)(B + (int64_t)n * g.ldb + k)
                                    : make_uint4(0, 0, 0, 0);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int c = tid + 256 * q, r = c >> 3, kc = (c & 7) * 8;
      *(uint4*)&As[r][kc] = ra[q];
      *(uint4*)&Bs[r][kc] = rb[q];
    }
  };

  if (kbeg < kend) {
  load(kbeg);
  store();
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += BKL) {
    const bool more = k0 + BKL < kend;
    if (more) load(k0 + BKL);
#pragma unroll
    for (int s = 0; s < BKL / 32; ++s) {
      const int kc = s * 32 + (lane >> 4) * 8;
      uint4 a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = *(const uint4*)&As[wm * 64 + i * 16 + (lane & 15)][kc];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = *(const uint4*)&Bs[wn * 64 + j * 16 + (lane & 15)][kc];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (IS_BF16)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                __builtin_bit_cast(v8bf, a[i]), __builtin_bit_cast(v8bf, b[j]), acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                __builtin_bit_cast(v8h, a[i]), __builtin_bit_cast(v8h, b[j]), acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
    if (more) {
      store();
      __syncthreads();
    }
  }
  }

  lp_epilogue<4>(g, acc, m0, n0, wm, wn, lane, batch, split);
}

// Large-tile BF16/FP16 GEMM for the big square products (the two ring GEMMs): 256x256 block
// tile, eight waves (two per SIMD) each owning a 128x64 piece (8x4 MFMA 16x16x32 tiles, 128 FP32
// accumulators per lane), K staged in 64-wide steps through two LDS buffers.  Global data goes
// straight to LDS (global_load_lds_dwordx4: each wave instruction fills 1 KB = 8 rows of 128 B),
// so no VGPRs hold staged tiles and the next step's loads run under the current step's MFMAs.
// The 16-byte chunk c of tile row r is stored at position c ^ ((r >> 1) & 7) of its row: the 16
// lanes of a fragment load (16 consecutive rows, one logical chunk) then hit 16 distinct bank
// quads.  Rows beyond M or N keep stale LDS data, which only reaches outputs that are not
// stored; K beyond the range is zero-filled (only the last step of a split can be partial).
constexpr int BT = 256;
constexpr int BKD = 64;

__device__ inline void lds_dma16(const void* g, uint16_t* lds_row_base) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_row_base, 16,
                                   0, 0);
}

using BigTile = uint16_t[2][BT][BKD];        // [A/B][tile row][k] of one K step

// One K step of MFMAs for this wave's 128x64 piece from LDS buffer T.
template <bool IS_BF16>
__device__ inline void big_step(const BigTile& T, v4f (&acc)[8][4], int wm, int wn, int lane) {
#pragma unroll
  for (int s = 0; s < BKD / 32; ++s) {
    const int cl = s * 4 + (lane >> 4);
    const int ra = wm * 128 + (lane & 15), rb = wn * 64 + (lane & 15);
    uint4 a[8], b[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = ra + i * 16;
      a[i] = *(const uint4*)&T[0][r][(cl ^ ((r >> 1) & 7)) * 8];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = rb + j * 16;
      b[j] = *(const uint4*)&T[1][r][(cl ^ ((r >> 1) & 7)) * 8];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (IS_BF16)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(v8bf, a[i]), __builtin_bit_cast(v8bf, b[j]), acc[i][j], 0, 0, 0);
        else
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
              __builtin_bit_cast(v8h, a[i]), __builtin_bit_cast(v8h, b[j]), acc[i][j], 0, 0, 0);
      }
    // keep the next slice's fragment reads below these MFMAs: one slice of fragments (48
    // VGPRs) live next to the 128 accumulators fits the two-waves-per-SIMD register budget
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The same K step for FP8 (OCP E4M3) operands: a 128-byte tile row holds 128 elements, one
// v_mfma_scale_f32_16x16x128_f8f6f4 per 16x16 output tile (unit scales).  Lane group q = lane >> 4
// feeds the MFMA the 32 bytes of logical chunks q and 4 + q of its row, for A and B alike: the
// instruction pairs A's and B's operand bytes position by position, so whichever k each position
// stands for, every k of the row meets its own k once and the sum over the step is the full dot
// product (the FP16 step's two 32-wide slices, fused).
typedef int v8i __attribute__((ext_vector_type(8)));
__device__ inline v8i frag8(const uint16_t (&T)[BT][BKD], int r, int q) {
  const uint4 x = *(const uint4*)&T[r][(q ^ ((r >> 1) & 7)) * 8];
  const uint4 y = *(const uint4*)&T[r][((q + 4) ^ ((r >> 1) & 7)) * 8];
  return (v8i){(int)x.x, (int)x.y, (int)x.z, (int)x.w, (int)y.x, (int)y.y, (int)y.z, (int)y.w};
}

__device__ inline void big_step8(const BigTile& T, v4f (&acc)[8][4], int wm, int wn, int lane) {
  const int q = lane >> 4;
  const int ra = wm * 128 + (lane & 15), rb = wn * 64 + (lane & 15);
  v8i b[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = frag8(T[1], rb + j * 16, q);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const v8i a = frag8(T[0], ra + i * 16, q);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b[j], acc[i][j], 0, 0, 0,
                                                                   127, 0, 127);
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <bool IS_BF16, bool KAL = false, bool F8 = false>
__global__ __launch_bounds__(512) void gemm_lp_big_kernel(GemmLPArgs g) {
  // two distinct LDS objects (not one array indexed by a step parity), so the compiler can tell
  // the buffer being read from the buffer the in-flight loads are filling
  __shared__ __attribute__((aligned(1024))) BigTile S0;
  __shared__ __attribute__((aligned(1024))) BigTile S1;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 2, wn = w & 3;
  const int batch = blockIdx.z;
  TileCoord tc = tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n);
  const int m0 = tc.tm * BT, n0 = tc.tn * BT;
  const uint16_t* A = g.A + (int64_t)batch * g.sA;
  const uint16_t* B = g.B + (int64_t)batch * g.sB;
  const int split = blockIdx.y;
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);

  v4f acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};

  // this lane's place in a DMA instruction: row (lane >> 3) of an 8-row group, position lane & 7
  const int drow = lane >> 3, dpos = lane & 7;
  // wave w fills rows [32 w, 32 w + 32) of both operand tiles, 4 instructions of 8 rows each
  auto issue = [&](BigTile& T, int k0) {
    if constexpr (KAL) {
      // K range a multiple of 64: a step is wholly inside or wholly beyond it (a uniform
      // branch), and rows beyond M or N read a clamped row, whose products reach only outputs
      // that are not stored, so no lane branches around a DMA
      if (k0 < kend) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r0 = w * 32 + q * 8, r = r0 + drow;
          const int k = k0 + (dpos ^ ((r >> 1) & 7)) * 8;
          const int m = min(m0 + r, g.M - 1), n = min(n0 + r, g.N - 1);
          lds_dma16(A + (int64_t)m * g.lda + k, &T[0][r0][0]);
          lds_dma16(B + (int64_t)n * g.ldb + k, &T[1][r0][0]);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = w * 32 + q * 8 + drow;
          *(uint4*)&T[0][r][dpos * 8] = make_uint4(0, 0, 0, 0);
          *(uint4*)&T[1][r][dpos * 8] = make_uint4(0, 0, 0, 0);
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r0 = w * 32 + q * 8, r = r0 + drow;         // tile row
      const int c = dpos ^ ((r >> 1) & 7);                  // logical chunk stored here
      const int k = k0 + c * 8;
      const bool kin = k < kend;
      const int m = m0 + r, n = n0 + r;
      if (m < g.M && kin) lds_dma16(A + (int64_t)m * g.lda + k, &T[0][r0][0]);
      if (n < g.N && kin) lds_dma16(B + (int64_t)n * g.ldb + k, &T[1][r0][0]);
      if (!kin) {
        *(uint4*)&T[0][r][dpos * 8] = make_uint4(0, 0, 0, 0);
        *(uint4*)&T[1][r][dpos * 8] = make_uint4(0, 0, 0, 0);
      }
    }
  };

  // Steps alternate S0, S1, in pairs (an odd step count gets one zero-filled step, so the loop
  // has a single exit).  At the top of a step: wait for this wave's loads, then a barrier
  // (everyone's loads of this step landed, everyone's reads of the other buffer are done), then
  // start the next step's loads into the other buffer and compute this one.
  const int npairs = kbeg < kend ? (kend - kbeg + 2 * BKD - 1) / (2 * BKD) : 0;
  if (npairs > 0) issue(S0, kbeg);
  for (int p = 0; p < npairs; ++p) {
    const int k0 = kbeg + p * 2 * BKD;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    issue(S1, k0 + BKD);
    if constexpr (F8) big_step8(S0, acc, wm, wn, lane);
    else big_step<IS_BF16>(S0, acc, wm, wn, lane);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (p + 1 < npairs) issue(S0, k0 + 2 * BKD);
    if constexpr (F8) big_step8(S1, acc, wm, wn, lane);
    else big_step<IS_BF16>(S1, acc, wm, wn, lane);
  }

  lp_epilogue<8>(g, acc, m0, n0, wm, wn, lane, batch, split);
}

// ---------------------------------------------------------------------------------- split GEMM
// 256x256 tiles like gemm_lp_big_kernel, for operands given as hi/lo pairs (x = hi + lo, both
// 16-bit): C = alpha diag(ia) (Ahi Bhi^T + Ahi Blo^T + Alo Bhi^T) diag(ib) + beta D, with PA / PB
// = 1 for an operand without a lo part (an exact Hamiltonian block).  One K step is 32 wide and
// holds the PA + PB operand tiles; NS steps are in flight through NS distinct LDS stages:
// at the top of step s a wave waits (vmcnt) for its own loads of stage s only, leaving the
// later stages' loads running, and after the barrier issues stage s + NS - 1 into the buffer
// step s - 1 read.  Every lane issues every DMA instruction (rows beyond M or N read a clamped
// row, whose products reach only outputs that are not stored), so each wave's count of loads
// per stage is the same constant and the vmcnt thresholds are exact.  K and the split-K chunk
// are multiples of 32 (the host checks), so no step is partial.  Rows of a stage tile are 64 B
// (four 16-byte chunks); chunk c of row r is stored at c ^ ((r >> 2) & 3), so the 16 lanes of a
// fragment read (16 consecutive rows, one logical chunk) hit 16 distinct bank quads.
constexpr int SBK = 32;

template <int NP>
using SStage = uint16_t[NP][BT][SBK];

// s_waitcnt with vmcnt = N (gfx9 encoding: vmcnt bits [3:0] and [15:14]; expcnt and lgkmcnt
// left at their maxima, so only vector-memory operations are waited for).
template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <bool IS_BF16>
__device__ inline void mfma_block(v4f (&acc)[8][4], const uint4 (&a)[8], const uint4 (&b)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (IS_BF16)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(v8bf, a[i]), __builtin_bit_cast(v8bf, b[j]), acc[i][j], 0, 0, 0);
      else
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
            __builtin_bit_cast(v8h, a[i]), __builtin_bit_cast(v8h, b[j]), acc[i][j], 0, 0, 0);
    }
}

template <bool IS_BF16, int PA, int PB, int NS, bool ZFILL = false>
__global__ __launch_bounds__(512) void gemm_lp_split_kernel(GemmLPArgs g) {
  constexpr int NP = PA + PB;
  constexpr int LPS = 2 * NP;          // DMA instructions per wave per stage
  __shared__ __attribute__((aligned(1024))) SStage<NP> S0;
  __shared__ __attribute__((aligned(1024))) SStage<NP> S1;
  // (stages beyond NS are never referenced, and unreferenced LDS objects are not allocated)
  __shared__ __attribute__((aligned(1024))) SStage<NP> S2[1];
  __shared__ __attribute__((aligned(1024))) SStage<NP> S3[1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 2, wn = w & 3;
  const int batch = blockIdx.z;
  TileCoord tc = tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n);
  const int m0 = tc.tm * BT, n0 = tc.tn * BT;
  const int64_t oA = (int64_t)batch * g.sA, oB = (int64_t)batch * g.sB;
  const uint16_t* Ah = g.A + oA;
  const uint16_t* Al = PA == 2 ? g.Alo + oA : nullptr;
  const uint16_t* Bh = g.B + oB;
  const uint16_t* Bl = PB == 2 ? g.Blo + oB : nullptr;
  const int split = blockIdx.y;
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const int nsteps = kbeg < kend ? (kend - kbeg + SBK - 1) / SBK : 0;

  v4f acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};

  // this lane's place in a DMA instruction: row (lane >> 2) of a 16-row group, position lane & 3
  const int drow = lane >> 2, dpos = lane & 3;
  auto issue = [&](SStage<NP>& T, int k0) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const bool isA = q < PA;
      const uint16_t* base = isA ? (q == 0 ? Ah : Al) : (q == PA ? Bh : Bl);
      const int64_t ld = isA ? g.lda : g.ldb;
      const int lim = (isA ? g.M : g.N) - 1, o0 = isA ? m0 : n0;
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int r0 = w * 32 + hh * 16, r = r0 + drow;
        const int c = dpos ^ ((r >> 2) & 3);
        const int row = min(o0 + r, lim);
        if (!ZFILL || k0 + c * 8 < kend) {
          lds_dma16(base + (int64_t)row * ld + k0 + c * 8, &T[q][r0][0]);
        } else {
          // (ZFILL: two stages only, K not a multiple of 32) the last step of the K range is
          // zero-filled beyond it; every wait there is a full one, so the count of DMA
          // instructions does not matter.  Without ZFILL no lane branches here, which keeps
          // the per-step issue free of exec-mask changes.
          *(uint4*)&T[q][r][dpos * 8] = make_uint4(0, 0, 0, 0);
        }
      }
    }
  };
  const int cl = lane >> 4;
  const int ra = wm * 128 + (lane & 15), rb = wn * 64 + (lane & 15);
  auto compute = [&](const SStage<NP>& T) {
    uint4 a[8], b[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = ra + i * 16;
      a[i] = *(const uint4*)&T[0][r][(cl ^ ((r >> 2) & 3)) * 8];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = rb + j * 16;
      b[j] = *(const uint4*)&T[PA][r][(cl ^ ((r >> 2) & 3)) * 8];
    }
    mfma_block<IS_BF16>(acc, a, b);                         // hi . hi
    if constexpr (PB == 2) {
      uint4 bl[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = rb + j * 16;
        bl[j] = *(const uint4*)&T[PA + 1][r][(cl ^ ((r >> 2) & 3)) * 8];
      }
      mfma_block<IS_BF16>(acc, a, bl);                      // hi . lo
    }
    if constexpr (PA == 2) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int r = ra + i * 16;
        a[i] = *(const uint4*)&T[1][r][(cl ^ ((r >> 2) & 3)) * 8];
      }
      mfma_block<IS_BF16>(acc, a, b);                       // lo . hi
    }
  };
  // top of step s: stage s's loads done (the later issued stages may still be in flight), all
  // waves past step s - 1 (so its buffer may be refilled), then issue stage s + NS - 1
  auto step = [&](SStage<NP>& Tc, SStage<NP>& Ti, int s) {
    if (s >= nsteps) return;
    const int rem = nsteps - 1 - s;
    if constexpr (NS == 4) {
      if (rem >= 2) wait_vmcnt<2 * LPS>();
      else if (rem == 1) wait_vmcnt<LPS>();
      else wait_vmcnt<0>();
    } else if constexpr (NS == 3) {
      if (rem >= 1) wait_vmcnt<LPS>();
      else wait_vmcnt<0>();
    } else {
      __builtin_amdgcn_s_waitcnt(0);     // loads and zero-fill stores of stage s landed
    }
    asm volatile("s_barrier" ::: "memory");
    if (s + NS - 1 < nsteps) issue(Ti, kbeg + (s + NS - 1) * SBK);
    compute(Tc);
  };
  if (0 < nsteps) issue(S0, kbeg);
  if (NS > 2 && 1 < nsteps) issue(S1, kbeg + SBK);
  if constexpr (NS > 3) {
    if (2 < nsteps) issue(S2[0], kbeg + 2 * SBK);
  }
  for (int s = 0; s < nsteps; s += NS) {
    if constexpr (NS == 2) {
      step(S0, S1, s);
      step(S1, S0, s + 1);
    } else if constexpr (NS == 3) {
      step(S0, S2[0], s);
      step(S1, S0, s + 1);
      step(S2[0], S1, s + 2);
    } else {
      step(S0, S3[0], s);
      step(S1, S0, s + 1);
      step(S2[0], S1, s + 2);
      step(S3[0], S2[0], s + 3);
    }
  }

  lp_epilogue<8>(g, acc, m0, n0, wm, wn, lane, batch, split);
}

// ---------------------------------------------------------------------------------- FP64, large
// 256x128 FP64 tiles for the large contractions (the ring products of the plain form, the hh
// ladder product): 8 waves of 64x64 (4x4 v_mfma_f64_16x16x4f64 tiles each, two waves per SIMD).
// K advances in steps of 16, staged by LDS DMA (global_load_lds_dwordx4: 2 doubles per lane, no
// register staging) into two distinct LDS buffers, so step s+1's loads run under step s's MFMAs.
// Each operand tile is held along the operand's unit-stride index:
//   k rows  (unit stride in k):        [row][16 k], 128-byte rows, 16-byte chunk c stored at
//                                      position c ^ ((row >> 1) & 7) of its row;
//   k-major (unit stride in m or n):   [k][rows], chunk c of k-row kr stored at c ^ ((kr & 1) << 3).
// With 64 banks of 4 bytes, both let the 32 lanes of a fragment read half (16 rows x 2 k) hit 32
// distinct 8-byte bank pairs.  The host checks that every 16-byte chunk is contiguous and aligned.
// Rows beyond M or N keep stale LDS data, which reaches only outputs that are not stored; k beyond
// the range is zero-filled in both operands.
constexpr int FTM = 256, FTN = 128, FBK = 16;
constexpr int FA_ELEMS = FTM * FBK, FB_ELEMS = FTN * FBK;
using F64Tile = double[FA_ELEMS + FB_ELEMS];   // A tile, then B tile, of one K step

__device__ inline void lds_dma16d(const double* g, double* lds) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
}

// LDS index of (row, k) in a k-rows tile, and in a k-major tile of ROWS rows.
__device__ inline int krows_at(int r, int k) {
  return r * FBK + ((((k >> 1) ^ ((r >> 1) & 7)) << 1) | (k & 1));
}
template <int ROWS>
__device__ inline int kmajor_at(int r, int k) {
  return k * ROWS + ((((r >> 1) ^ ((k & 1) << 3)) << 1) | (r & 1));
}

// Wave w's share of one operand tile per K step: ROWS/64 DMA instructions of 1 KB covering rows
// [r0, r0 + ROWS) of an operand with nrows rows (index dr) and index k (dk).  Row offsets are
// fixed per lane; the k offsets come from two walkers that advance by FBK per step along the
// (possibly two-run) k index, so the loop carries no integer division.
template <bool KROWS, int ROWS>
struct F64Fill {
  static constexpr int NI = ROWS / 64;             // instructions per wave and step
  static constexpr int PER_K = ROWS / 128;         // k-major: instructions per k-row
  int64_t roff[NI];
  bool rin[NI];
  int wk[2], wr[2];                                // walker: k, position in its run
  int64_t woff[2];                                 // walker: offset of k
  __device__ inline F64Fill(const Idx2& dr, const Idx2& dk, int r0, int nrows, int k0, int w,
                            int lane) {
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const int ins = w * NI + q;
      int r;
      if (KROWS) r = ins * 8 + (lane >> 3);
      else r = 2 * (((ins % PER_K) * 64 + lane) ^ (((ins / PER_K) & 1) << 3));
      rin[q] = r0 + r < nrows;
      roff[q] = rin[q] ? off2(r0 + r, dr) : 0;
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      int k;
      if (KROWS) {
        const int r = (w * NI + v) * 8 + (lane >> 3);
        k = k0 + 2 * ((lane & 7) ^ ((r >> 1) & 7));
      } else {
        k = k0 + (w * NI + v * PER_K) / PER_K;
      }
      const int qk = div2(k, dk);
      wk[v] = k; wr[v] = k - qk * dk.n2;
      woff[v] = qk * dk.s1 + (int64_t)wr[v] * dk.s2;
    }
  }
  __device__ inline void issue(double* T, const double* P, const Idx2& dk, int kend, int w,
                               int lane) {
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const int ins = w * NI + q;
      const int v = KROWS ? (q & 1) : q / PER_K;
      if (KROWS) {
        const int rb = ins * 8;
        if (wk[v] < kend) {
          if (rin[q]) lds_dma16d(P + roff[q] + woff[v], T + rb * FBK);
        } else {
          double* z = T + (rb + (lane >> 3)) * FBK + 2 * (lane & 7);
          z[0] = 0.0; z[1] = 0.0;
        }
      } else {
        const int kr = ins / PER_K, part = ins % PER_K;
        if (wk[v] < kend) {
          if (rin[q]) lds_dma16d(P + woff[v] + roff[q], T + kr * ROWS + part * 128);
        } else {
          double* z = T + kr * ROWS + 2 * (part * 64 + lane);
          z[0] = 0.0; z[1] = 0.0;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      wk[v] += FBK;
      if (dk.n2 == 1) {
        woff[v] += FBK * dk.s1;
      } else {
        wr[v] += FBK; woff[v] += FBK * dk.s2;
        while (wr[v] >= dk.n2) { wr[v] -= dk.n2; woff[v] += dk.s1 - (int64_t)dk.n2 * dk.s2; }
      }
    }
  }
};

// One K step (16) of this wave's 64x64 piece from LDS buffer T.
template <bool AK, bool BN>
__device__ inline void f64_step(const F64Tile& T, v4d (&acc)[4][4], int wm, int wn, int lane) {
  const double* TA = T;
  const double* TB = T + FA_ELEMS;
#pragma unroll
  for (int kk = 0; kk < FBK / 4; ++kk) {
    const int k = kk * 4 + (lane >> 4);
    double a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = wm * 64 + i * 16 + (lane & 15);
      a[i] = TA[AK ? krows_at(r, k) : kmajor_at<FTM>(r, k)];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = wn * 64 + j * 16 + (lane & 15);
      b[j] = TB[BN ? kmajor_at<FTN>(r, k) : krows_at(r, k)];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

template <bool AK, bool BN>
__global__ __launch_bounds__(512) void gemm_f64_big_kernel(GemmF64Args g) {
  // two distinct LDS objects, so the compiler can tell the buffer being read from the buffer
  // that the in-flight DMA is filling
  __shared__ __attribute__((aligned(1024))) F64Tile S0;
  __shared__ __attribute__((aligned(1024))) F64Tile S1;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int batch = blockIdx.z, split = blockIdx.y;
  TileCoord tc = tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n);
  const int m0 = tc.tm * FTM, n0 = tc.tn * FTN;
  const double* A = g.A + off2(batch, g.ab);
  const double* B = g.B + off2(batch, g.bb);
  const int kbeg = split * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);

  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

  F64Fill<AK, FTM> fa(g.am, g.ak, m0, g.M, kbeg, w, lane);
  F64Fill<!BN, FTN> fb(g.bn, g.bk, n0, g.N, kbeg, w, lane);
  // successive calls fill successive K steps starting at kbeg
  auto issue = [&](F64Tile& T, int) {
    fa.issue(T, A, g.ak, kend, w, lane);
    fb.issue(T + FA_ELEMS, B, g.bk, kend, w, lane);
  };

  // Steps alternate S0, S1 in pairs (an odd step count gets one zero-filled step, so the loop
  // has a single exit).  At the top of a step: wait for this wave's loads, a barrier (everyone's
  // loads of this step landed, everyone's reads of the other buffer are done), then start the
  // next step's loads into the other buffer and compute this one.
  const int npairs = kbeg < kend ? (kend - kbeg + 2 * FBK - 1) / (2 * FBK) : 0;
  if (npairs > 0) issue(S0, kbeg);
  for (int p = 0; p < npairs; ++p) {
    const int k0 = kbeg + p * 2 * FBK;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    issue(S1, k0 + FBK);
    f64_step<AK, BN>(S0, acc, wm, wn, lane);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (p + 1 < npairs) issue(S0, k0 + 2 * FBK);
    f64_step<AK, BN>(S1, acc, wm, wn, lane);
  }

  // f64 16x16x4 accumulator layout: register r of lane l holds D[4r + l/16][l%16]
  if (g.splits > 1) {
    double* W = g.ws + ((int64_t)batch * g.splits + split) * (int64_t)g.M * g.N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = m0 + wm * 64 + i * 16 + r * 4 + (lane >> 4);
          int col = n0 + wn * 64 + j * 16 + (lane & 15);
          if (row < g.M && col < g.N) W[(int64_t)row * g.N + col] = g.alpha * acc[i][j][r];
        }
    return;
  }
  const int64_t cb = off2(batch, g.cb);
  const bool addend = g.D && g.beta != 0.0;
  int64_t coff[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int col = n0 + wn * 64 + j * 16 + (lane & 15);
    coff[j] = col < g.N ? off2(col, g.cn) : -1;
  }
  // per 16-row group: all addend loads first, then the stores (D may be C; each element is read
  // and written by this lane only), so the loads overlap instead of waiting on each store
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int64_t ro[4];
    double dv[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + wm * 64 + i * 16 + r * 4 + (lane >> 4);
      ro[r] = row < g.M ? cb + off2(row, g.cm) : -1;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        dv[r][j] = (addend && ro[r] >= 0 && coff[j] >= 0) ? g.D.get(ro[r] + coff[j]) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (ro[r] < 0) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (coff[j] >= 0) g.C.put(ro[r] + coff[j], g.alpha * acc[i][j][r] + g.beta * dv[r][j]);
    }
  }
}

}  // namespace

extern "C" {

static inline Idx2 mk(int64_t n2, int64_t s1, int64_t s2) {
  const FastDiv f = nl_fastdiv(n2);
  Idx2 d; d.n2 = f.d; d.s1 = s1; d.s2 = s2; d.mg = f.mg; d.sh = f.sh;
  return d;
}

// Kernel choice and split-K factor for a contraction of the given extents; returns the number of
// workspace doubles nl_contract_f64 needs (0 if no split).
static void contract_plan(int64_t M, int64_t N, int64_t K, int64_t nb, int* kind, int64_t* splits) {
  int64_t tiles;
  if (K <= LKK && M > 64 && N > 64) {
    *kind = 3;   // low K: 64x64 output tiles, whole K in LDS, lanes along C's unit stride
    *splits = 1;
  } else if (N <= 64 && M <= 64) {
    *kind = 2;   // small output: split K widely
    int64_t s = K <= 256 ? 1 : std::min<int64_t>(2048 / std::max<int64_t>(nb, 1), K / 256);
    *splits = std::max<int64_t>(1, std::min<int64_t>(s, 4096));
  } else if (N <= 64 || M <= 64) {
    *kind = 1;   // skinny (operands swapped when M is the small side)
    int64_t big = N <= 64 ? M : N;
    tiles = nl_cdiv(big, SKM) * nb;
    int64_t s = (tiles >= 2048 || K <= 256) ? 1 : std::min<int64_t>(2048 / tiles, K / 128);
    *splits = std::max<int64_t>(1, std::min<int64_t>(s, 4096));
  } else {
    *kind = 0;
    tiles = nl_cdiv(M, TM) * nl_cdiv(N, TN) * nb;
    int64_t s = (tiles >= 512 || K <= 512) ? 1 : std::min<int64_t>(1024 / tiles, K / 256);
    *splits = std::max<int64_t>(1, std::min<int64_t>(s, 256));
  }
}

int64_t nl_contract_ws(const int64_t* dims) {
  if (!dims) return -1;
  const int64_t M = dims[0] * dims[1], N = dims[2] * dims[3], K = dims[4] * dims[5];
  const int64_t nb = dims[6] * dims[7];
  int kind; int64_t splits;
  contract_plan(M, N, K, nb, &kind, &splits);
  return splits > 1 ? nb * splits * M * N : 0;
}

// General strided FP64 contraction (see the kernel comment).  dims[8] = {M1, M2, N1, N2, K1,
// K2, B1, B2}; sa[6] = A strides {m1, m2, k1, k2, b1, b2}; sb[6] = B {k1, k2, n1, n2, b1, b2};
// sc[6] = C (and D) {m1, m2, n1, n2, b1, b2}.  C = alpha*A*B + beta*D (D may be C or null).
// Split-K reduction of the [batch][splits][M][N] partials into C (layout cm, cn, cb).
static void reduce_splits(hipStream_t st, int M, int N, int splits, int nb, const double* ws,
                          double beta, AnyIn D, AnyOut C, Idx2 cm, Idx2 cn, Idx2 cb) {
  const int64_t total = (int64_t)nb * M * N;
  if (splits >= 64) {
    int blocks = (int)std::min<int64_t>(nl_cdiv(total * 64, 256), 16384);
    hipLaunchKernelGGL(splitk_reduce_wave_kernel, dim3(blocks), dim3(256), 0, st, M, N, splits,
                       nb, ws, beta, D, C, cm, cn, cb);
  } else {
    int blocks = (int)std::min<int64_t>(nl_cdiv(total, 256), 8192);
    hipLaunchKernelGGL(splitk_reduce_strided_kernel, dim3(blocks), dim3(256), 0, st, M, N,
                       splits, nb, ws, beta, D, C, cm, cn, cb);
  }
}

// The 16-byte chunks (2 doubles) along an operand's unit-stride index group (extents n1, n2,
// strides s1, s2) are contiguous and, given a 16-byte aligned base and even strides elsewhere,
// aligned: the DMA condition of gemm_f64_big_kernel.
static bool chunk_group(int64_t n1, int64_t n2, int64_t s1, int64_t s2) {
  if (n2 > 1) return s2 == 1 && n2 % 2 == 0 && (n1 <= 1 || s1 % 2 == 0);
  return n1 > 1 && s1 == 1 && n1 % 2 == 0;
}
static bool even_stride(int64_t n, int64_t s) { return n <= 1 || s % 2 == 0; }

int nl_contract_f64(const int64_t* dims, double alpha, const double* A, const int64_t* sa,
                    const double* B, const int64_t* sb, double beta, const void* D, void* C,
                    const int64_t* sc, int64_t splits, double* ws, int io, void* stream) {
  if (!dims || !sa || !sb || !sc || (io & ~(NL_C_F32 | NL_D_F32))) return NL_EARG;
  int kind = 0;
  const bool auto_plan = splits <= 0;
  if (auto_plan) {   // automatic kernel choice and split (ws sized by nl_contract_ws)
    contract_plan(dims[0] * dims[1], dims[2] * dims[3], dims[4] * dims[5], dims[6] * dims[7],
                  &kind, &splits);
  }
  for (int d = 0; d < 8; ++d)
    if (dims[d] < 0) return NL_EARG;
  const int64_t M = dims[0] * dims[1], N = dims[2] * dims[3], K = dims[4] * dims[5];
  const int64_t nb = dims[6] * dims[7];
  if (M == 0 || N == 0 || nb == 0) return NL_OK;
  if (!A || !B || !C || splits < 1 || (splits > 1 && !ws)) return NL_EARG;
  if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX || nb > 65535 || splits > 65535)
    return NL_EARG;
  if (M + TM > INT32_MAX || N + TN > INT32_MAX || K + BK64 > INT32_MAX) return NL_EARG;
  hipStream_t st = (hipStream_t)stream;
  GemmF64Args g;
  g.M = (int)M; g.N = (int)N; g.K = (int)K; g.nbatch = (int)nb; g.alpha = alpha;
  g.beta = beta; g.D = AnyIn{D, (io & NL_D_F32) != 0};
  g.A = A; g.am = mk(dims[1], sa[0], sa[1]); g.ak = mk(dims[5], sa[2], sa[3]);
  g.ab = mk(dims[7], sa[4], sa[5]);
  g.B = B; g.bk = mk(dims[5], sb[0], sb[1]); g.bn = mk(dims[3], sb[2], sb[3]);
  g.bb = mk(dims[7], sb[4], sb[5]);
  g.C = AnyOut{C, (io & NL_C_F32) != 0}; g.cm = mk(dims[1], sc[0], sc[1]); g.cn = mk(dims[3], sc[2], sc[3]);
  g.cb = mk(dims[7], sc[4], sc[5]);
  g.tiles_m = (int)nl_cdiv(M, TM); g.tiles_n = (int)nl_cdiv(N, TN);
  int64_t kchunk = nl_cdiv(nl_cdiv(K > 0 ? K : 1, splits), BK64) * BK64;
  g.splits = K == 0 ? 1 : (int)splits; g.kchunk = (int)kchunk; g.ws = ws;
  if (kind == 2) {
    const bool am = (g.am.n2 > 1 ? g.am.s2 == 1 : g.am.s1 == 1);
    const bool bn = (g.bn.n2 > 1 ? g.bn.s2 == 1 : g.bn.s1 == 1);
    g.kchunk = (int)(nl_cdiv(nl_cdiv(K > 0 ? K : 1, g.splits), SMK) * SMK);
    dim3 sg(1u, (unsigned)g.splits, (unsigned)nb);
#define NL_SM(KERN)                                                                          \
    if (am && bn)   hipLaunchKernelGGL((KERN<true, true>), sg, dim3(256), 0, st, g);       \
    else if (am)    hipLaunchKernelGGL((KERN<true, false>), sg, dim3(256), 0, st, g);      \
    else if (bn)    hipLaunchKernelGGL((KERN<false, true>), sg, dim3(256), 0, st, g);      \
    else            hipLaunchKernelGGL((KERN<false, false>), sg, dim3(256), 0, st, g);
    if (M <= 32 && N <= 32) { NL_SM(smallmn32_f64_kernel) } else { NL_SM(smallmn_f64_kernel) }
#undef NL_SM
    NL_LAUNCH_CHECK();
    if (g.splits > 1) {
      reduce_splits(st, (int)M, (int)N, g.splits, (int)nb, ws, beta, g.D, g.C, g.cm, g.cn, g.cb);
      NL_LAUNCH_CHECK();
    }
    return NL_OK;
  }
  if (kind == 3) {
    g.tiles_m = (int)nl_cdiv(M, LKT); g.tiles_n = (int)nl_cdiv(N, LKT);
    g.splits = 1;
    const bool ln = (g.cn.n2 > 1 ? g.cn.s2 == 1 : g.cn.s1 == 1) ||
                    !(g.cm.n2 > 1 ? g.cm.s2 == 1 : g.cm.s1 == 1);
    dim3 lg((unsigned)(g.tiles_m * g.tiles_n), 1u, (unsigned)nb);
#define NL_LK(TC, TD)                                                                        \
    if (ln) hipLaunchKernelGGL((lowk_f64_kernel<true, TC, TD>), lg, dim3(256), 0, st, g);    \
    else    hipLaunchKernelGGL((lowk_f64_kernel<false, TC, TD>), lg, dim3(256), 0, st, g);
    if (g.C.f32) { if (g.D.f32) { NL_LK(float, float) } else { NL_LK(float, double) } }
    else { if (g.D.f32) { NL_LK(double, float) } else { NL_LK(double, double) } }
#undef NL_LK
    NL_LAUNCH_CHECK();
    return NL_OK;
  }
  if (kind == 1) {
    if (N > 64) {   // M is the small side: C^T = B^T A^T
      std::swap(g.M, g.N); std::swap(g.tiles_m, g.tiles_n);
      std::swap(g.A, g.B);
      Idx2 t = g.am; g.am = g.bn; g.bn = t;
      t = g.ak; g.ak = g.bk; g.bk = t;
      t = g.ab; g.ab = g.bb; g.bb = t;
      t = g.cm; g.cm = g.cn; g.cn = t;
    }
    const bool am = (g.am.n2 > 1 ? g.am.s2 == 1 : g.am.s1 == 1);
    dim3 sg((unsigned)nl_cdiv(g.M, SKM), (unsigned)g.splits, (unsigned)nb);
    const int n = g.N;
#define NL_SK(NB_)                                                                        \
    if (am) hipLaunchKernelGGL((skinny_f64_kernel<NB_, true>), sg, dim3(256), 0, st, g);  \
    else    hipLaunchKernelGGL((skinny_f64_kernel<NB_, false>), sg, dim3(256), 0, st, g);
    if (n <= 8) { NL_SK(8) } else if (n <= 16) { NL_SK(16) } else if (n <= 32) { NL_SK(32) }
    else { NL_SK(64) }
#undef NL_SK
    NL_LAUNCH_CHECK();
    if (g.splits > 1) {
      reduce_splits(st, g.M, g.N, g.splits, (int)nb, ws, beta, g.D, g.C, g.cm, g.cn, g.cb);
      NL_LAUNCH_CHECK();
    }
    return NL_OK;
  }
  // thread mapping: unit-stride inner k run of A -> load along k; of B's n -> load along n
  const bool ak = (dims[5] > 1 ? sa[3] == 1 : dims[4] > 1 ? sa[2] == 1 : true) ||
                  !(dims[1] > 1 ? sa[1] == 1 : dims[0] > 1 ? sa[0] == 1 : false);
  const bool bn = (dims[3] > 1 ? sb[3] == 1 : dims[2] > 1 ? sb[2] == 1 : true) ||
                  !(dims[5] > 1 ? sb[1] == 1 : dims[4] > 1 ? sb[0] == 1 : false);
  // 256x128 DMA tiles when every chunk is aligned and they alone give each CU a workgroup
  const int64_t big_tiles = nl_cdiv(M, FTM) * nl_cdiv(N, FTN);
  const bool a_dma = (((uintptr_t)A) & 15) == 0 && even_stride(dims[6], sa[4]) &&
                     even_stride(dims[7], sa[5]) &&
                     (ak ? chunk_group(dims[4], dims[5], sa[2], sa[3]) &&
                               even_stride(dims[0], sa[0]) && even_stride(dims[1], sa[1])
                         : chunk_group(dims[0], dims[1], sa[0], sa[1]) &&
                               even_stride(dims[4], sa[2]) && even_stride(dims[5], sa[3]));
  const bool b_dma = (((uintptr_t)B) & 15) == 0 && even_stride(dims[6], sb[4]) &&
                     even_stride(dims[7], sb[5]) &&
                     (bn ? chunk_group(dims[2], dims[3], sb[2], sb[3]) &&
                               even_stride(dims[4], sb[0]) && even_stride(dims[5], sb[1])
                         : chunk_group(dims[4], dims[5], sb[0], sb[1]) &&
                               even_stride(dims[2], sb[2]) && even_stride(dims[3], sb[3]));
  if (auto_plan && g.splits == 1 && M >= FTM && N >= FTN && big_tiles * nb >= 256 && a_dma &&
      b_dma) {
    g.tiles_m = (int)nl_cdiv(M, FTM); g.tiles_n = (int)nl_cdiv(N, FTN);
    dim3 bg((unsigned)big_tiles, 1u, (unsigned)nb);
    if (ak && bn)   hipLaunchKernelGGL((gemm_f64_big_kernel<true, true>), bg, dim3(512), 0, st, g);
    else if (ak)    hipLaunchKernelGGL((gemm_f64_big_kernel<true, false>), bg, dim3(512), 0, st, g);
    else if (bn)    hipLaunchKernelGGL((gemm_f64_big_kernel<false, true>), bg, dim3(512), 0, st, g);
    else            hipLaunchKernelGGL((gemm_f64_big_kernel<false, false>), bg, dim3(512), 0, st, g);
    NL_LAUNCH_CHECK();
    return NL_OK;
  }
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), (unsigned)g.splits, (unsigned)nb);
  if (ak && bn)   hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, dim3(256), 0, st, g);
  else if (ak)    hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, dim3(256), 0, st, g);
  else if (bn)    hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, dim3(256), 0, st, g);
  else            hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, dim3(256), 0, st, g);
  NL_LAUNCH_CHECK();
  if (g.splits > 1) {
    reduce_splits(st, (int)M, (int)N, g.splits, (int)nb, ws, beta, g.D, g.C, g.cm, g.cn, g.cb);
    NL_LAUNCH_CHECK();
  }
  return NL_OK;
}

// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b], row-major C (ldc), FP64.
//   a_kcontig = 1: A(m,k) = A[m*lda + k];  0: A(m,k) = A[k*lda + m]
//   b_ncontig = 1: B(k,n) = B[k*ldb + n];  0: B(k,n) = B[n*ldb + k]
//   splits > 1 splits K; ws must hold batch*splits*M*N doubles.
int nl_gemm_f64(int a_kcontig, int b_ncontig, int64_t M, int64_t N, int64_t K, double alpha,
                const double* A, int64_t lda, int64_t strideA, const double* B, int64_t ldb,
                int64_t strideB, double beta, double* C, int64_t ldc, int64_t strideC,
                int64_t batch, int64_t splits, double* ws, void* stream) {
  if (M < 0 || N < 0 || K < 0 || batch < 0) return NL_EARG;
  const int64_t dims[8] = {1, M, 1, N, 1, K, 1, batch};
  const int64_t sa[6] = {0, a_kcontig ? lda : 1, 0, a_kcontig ? 1 : lda, 0, strideA};
  const int64_t sb[6] = {0, b_ncontig ? ldb : 1, 0, b_ncontig ? 1 : ldb, 0, strideB};
  const int64_t sc[6] = {0, ldc, 0, 1, 0, strideC};
  return nl_contract_f64(dims, alpha, A, sa, B, sb, beta, C, C, sc, splits, ws, 0, stream);
}

// C[b] (FP64, row-major) = alpha * diag(ia) * (A[b] * B[b]^T) * diag(ib) + beta * C[b]
// A: [M][lda], B: [N][ldb], 16-bit elements (dtype 0 = bf16, 1 = fp16), K-contiguous rows.
// Shared body of the two LP entry points; C/D addressed through (cm, cn) and batch stride sC.
// K, lda, ldb must be multiples of 8 and A, B 16-byte aligned.
static int gemm_lp_impl(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                        int64_t lda, int64_t strideA, const float* ia, int64_t strideIa,
                        const void* B, int64_t ldb, int64_t strideB, const float* ib,
                        int64_t strideIb, double beta, const void* D, void* C, Idx2 cm,
                        Idx2 cn, int64_t strideC, int64_t batch, int64_t splits, double* ws,
                        void* stream, const int64_t* crow = nullptr, const void* Alo = nullptr,
                        const void* Blo = nullptr, bool staged = false) {
  // dtype: operand type (bits 0-7) | io flags << 8 (NL_C_F32, NL_D_F32: FP32 output / addend)
  const int io = dtype >> 8;
  dtype &= 0xff;
  if (io & ~(NL_C_F32 | NL_D_F32)) return NL_EARG;
  if (M < 0 || N < 0 || K < 0 || batch < 0 || dtype < 0 || dtype > 2) return NL_EARG;
  if (M == 0 || N == 0 || batch == 0) return NL_OK;
  if (dtype == 2) {
    // FP8 (E4M3) operands: one byte per element; the 256x256 kernel only, read as 16-bit units
    // (a 16-byte DMA moves 16 elements, a 64-unit K step holds 128).  K, the leading dimensions
    // and the batch strides in elements: K a multiple of 128, the rest of 16.
    if (Alo || Blo || crow || staged) return NL_EARG;
    if ((K % 128) || (lda % 16) || (ldb % 16) || lda < K || ldb < K || (strideA % 16) ||
        (strideB % 16) || !A || !B || !C || (((uintptr_t)A) & 15) || (((uintptr_t)B) & 15))
      return NL_EARG;
    if (M < BT || N < BT || nl_cdiv(M, BT) * nl_cdiv(N, BT) * splits * batch < 256)
      return NL_EARG;
    if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX || batch > 65535) return NL_EARG;
    if (splits < 1 || splits > 65535 || (splits > 1 && !ws)) return NL_EARG;
    GemmLPArgs g;
    g.M = (int)M; g.N = (int)N; g.K = (int)(K / 2); g.alpha = alpha; g.beta = beta;
    g.A = (const uint16_t*)A; g.lda = lda / 2; g.sA = strideA / 2;
    g.B = (const uint16_t*)B; g.ldb = ldb / 2; g.sB = strideB / 2;
    g.Alo = nullptr; g.Blo = nullptr;
    g.ia = ia; g.sia = strideIa; g.ib = ib; g.sib = strideIb;
    g.C = AnyOut{C, (io & NL_C_F32) != 0}; g.cm = cm; g.cn = cn; g.sC = strideC;
    g.D = AnyIn{D, (io & NL_D_F32) != 0}; g.crow = nullptr;
    g.tiles_m = (int)nl_cdiv(M, BT); g.tiles_n = (int)nl_cdiv(N, BT);
    g.splits = (int)splits; g.ws = ws;
    g.kchunk = (int)(nl_cdiv(nl_cdiv(g.K, splits), BKD) * BKD);
    const dim3 grid((unsigned)(g.tiles_m * g.tiles_n), (unsigned)splits, (unsigned)batch);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL((gemm_lp_big_kernel<false, true, true>), grid, dim3(512), 0, st, g);
    NL_LAUNCH_CHECK();
    if (splits > 1) {
      reduce_splits(st, (int)M, (int)N, (int)splits, (int)batch, ws, beta, g.D, g.C, cm, cn,
                    mk(1, strideC, 0));
      NL_LAUNCH_CHECK();
    }
    return NL_OK;
  }
  if (!A || !B || !C) return NL_EARG;
  if ((K % 8) || (lda % 8) || (ldb % 8) || lda < K || ldb < K) return NL_EARG;
  if ((((uintptr_t)A) & 15) || (((uintptr_t)B) & 15) || (strideA % 8) || (strideB % 8))
    return NL_EARG;
  if ((((uintptr_t)Alo) & 15) || (((uintptr_t)Blo) & 15)) return NL_EARG;
  if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX || batch > 65535) return NL_EARG;
  if (splits < 1 || splits > 65535 || (splits > 1 && !ws)) return NL_EARG;
  if (K == 0) return NL_EARG;   // nothing to accumulate: no caller needs C = beta * C here
  GemmLPArgs g;
  g.M = (int)M; g.N = (int)N; g.K = (int)K; g.alpha = alpha; g.beta = beta;
  g.A = (const uint16_t*)A; g.lda = lda; g.sA = strideA;
  g.B = (const uint16_t*)B; g.ldb = ldb; g.sB = strideB;
  g.Alo = (const uint16_t*)Alo; g.Blo = (const uint16_t*)Blo;
  g.ia = ia; g.sia = strideIa; g.ib = ib; g.sib = strideIb;
  g.C = AnyOut{C, (io & NL_C_F32) != 0}; g.cm = cm; g.cn = cn; g.sC = strideC;
  g.D = AnyIn{D, (io & NL_D_F32) != 0}; g.crow = crow;
  if (crow && splits != 1) return NL_EARG;
  g.tiles_m = (int)nl_cdiv(M, TM); g.tiles_n = (int)nl_cdiv(N, TN);
  g.splits = (int)splits; g.ws = ws;
  g.kchunk = (int)(nl_cdiv(nl_cdiv(K, splits), BKL) * BKL);   // multiple of 64 (hence of 8)
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), (unsigned)splits, (unsigned)batch);
  hipStream_t st = (hipStream_t)stream;
  // 256x256 tiles when they alone give every CU a workgroup, else 128x128 tiles
  const int64_t big_tiles = nl_cdiv(M, BT) * nl_cdiv(N, BT);
  const bool fills = big_tiles * splits * batch >= 256;
  const bool big = M >= BT && N >= BT && fills;
  const bool split_ops = Alo || Blo;
  // the split kernel clamps its operand rows, so M or N below one tile is fine there (the many
  // small pair blocks of the pp ladder, batched) as long as the tiles fill the GPU
  if (fills && (split_ops || staged) && (K % SBK == 0 || (Alo && Blo))) {
    // hi/lo operands (or the staged pipeline on request): the multi-stage split kernel
    g.tiles_m = (int)nl_cdiv(M, BT); g.tiles_n = (int)nl_cdiv(N, BT);
    grid = dim3((unsigned)(g.tiles_m * g.tiles_n), (unsigned)splits, (unsigned)batch);
    const int pa = Alo ? 2 : 1, pb = Blo ? 2 : 1;
#define NL_SPLIT(BF, PA_, PB_, NS_, ZF_)                                                     \
  hipLaunchKernelGGL((gemm_lp_split_kernel<BF, PA_, PB_, NS_, ZF_>), grid, dim3(512), 0, st, g)
    // only the two-stage form takes a K that is not a multiple of 32 (zero-filled last step)
    const bool zf = K % SBK != 0;
    if (dtype == 0) {
      if (pa == 2 && pb == 2) {
        if (zf) NL_SPLIT(true, 2, 2, 2, true); else NL_SPLIT(true, 2, 2, 2, false);
      } else if (pa == 2) NL_SPLIT(true, 2, 1, 3, false);
      else if (pb == 2) NL_SPLIT(true, 1, 2, 3, false);
      else NL_SPLIT(true, 1, 1, 4, false);
    } else {
      if (pa == 2 && pb == 2) {
        if (zf) NL_SPLIT(false, 2, 2, 2, true); else NL_SPLIT(false, 2, 2, 2, false);
      } else if (pa == 2) NL_SPLIT(false, 2, 1, 3, false);
      else if (pb == 2) NL_SPLIT(false, 1, 2, 3, false);
      else NL_SPLIT(false, 1, 1, 4, false);
    }
#undef NL_SPLIT
  } else if (split_ops) {
    // shapes the split kernel does not take: the three products in turn, each accumulating
    // into C (the first applies beta and D)
    // (the later products add to C itself, so their addend has C's element type)
    const int t0 = dtype | (io << 8);
    const int tc = dtype | ((io & NL_C_F32) ? (NL_C_F32 | NL_D_F32) << 8 : 0);
    int rc = gemm_lp_impl(t0, M, N, K, alpha, A, lda, strideA, ia, strideIa, B, ldb, strideB,
                          ib, strideIb, beta, D, C, cm, cn, strideC, batch, splits, ws, stream,
                          crow);
    if (rc == NL_OK && Blo)
      rc = gemm_lp_impl(tc, M, N, K, alpha, A, lda, strideA, ia, strideIa, Blo, ldb, strideB,
                        ib, strideIb, 1.0, C, C, cm, cn, strideC, batch, splits, ws,
                        stream, crow);
    if (rc == NL_OK && Alo)
      rc = gemm_lp_impl(tc, M, N, K, alpha, Alo, lda, strideA, ia, strideIa, B, ldb, strideB,
                        ib, strideIb, 1.0, C, C, cm, cn, strideC, batch, splits, ws,
                        stream, crow);
    return rc;
  } else if (big) {
    g.tiles_m = (int)nl_cdiv(M, BT); g.tiles_n = (int)nl_cdiv(N, BT);
    grid = dim3((unsigned)(g.tiles_m * g.tiles_n), (unsigned)splits, (unsigned)batch);
    // every split's K range a multiple of 64 (kchunk is): steps never straddle its end
    const bool kal = K % BKD == 0;
    if (dtype == 0) {
      if (kal) hipLaunchKernelGGL((gemm_lp_big_kernel<true, true>), grid, dim3(512), 0, st, g);
      else hipLaunchKernelGGL((gemm_lp_big_kernel<true, false>), grid, dim3(512), 0, st, g);
    } else {
      if (kal) hipLaunchKernelGGL((gemm_lp_big_kernel<false, true>), grid, dim3(512), 0, st, g);
      else hipLaunchKernelGGL((gemm_lp_big_kernel<false, false>), grid, dim3(512), 0, st, g);
    }
  } else if (dtype == 0) {
    hipLaunchKernelGGL((gemm_lp_kernel<true>), grid, dim3(256), 0, st, g);
  } else {
    hipLaunchKernelGGL((gemm_lp_kernel<false>), grid, dim3(256), 0, st, g);
  }
  NL_LAUNCH_CHECK();
  if (splits > 1) {
    reduce_splits(st, (int)M, (int)N, (int)splits, (int)batch, ws, beta, g.D, g.C, cm, cn,
                  mk(1, strideC, 0));
    NL_LAUNCH_CHECK();
  }
  return NL_OK;
}

int nl_gemm_lp(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
               int64_t lda, int64_t strideA, const float* ia, int64_t strideIa, const void* B,
               int64_t ldb, int64_t strideB, const float* ib, int64_t strideIb, double beta,
               const void* D, void* C, int64_t ldc, int64_t strideC, int64_t batch,
               int64_t splits, double* ws, void* stream) {
  return gemm_lp_impl(dtype, M, N, K, alpha, A, lda, strideA, ia, strideIa, B, ldb, strideB, ib,
                      strideIb, beta, D, C, mk(1, ldc, 0), mk(1, 1, 0), strideC, batch, splits,
                      ws, stream);
}

int nl_gemm_lp_rows(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                    int64_t lda, int64_t strideA, const float* ia, int64_t strideIa, const void* B,
                    int64_t ldb, int64_t strideB, const float* ib, int64_t strideIb, void* C,
                    int64_t ldc, const int64_t* crow, int64_t batch, void* stream) {
  if (!crow || ldc < N) return NL_EARG;
  if (dtype & (NL_C_F32 << 8)) dtype |= NL_D_F32 << 8;       // the rows add to C itself
  return gemm_lp_impl(dtype, M, N, K, alpha, A, lda, strideA, ia, strideIa, B, ldb, strideB, ib,
                      strideIb, 1.0, C, C, mk(1, ldc, 0), mk(1, 1, 0), 0, batch, 1, nullptr,
                      stream, crow);
}

int nl_gemm_lp_strided(int dtype, int64_t K, double alpha, const void* A, int64_t lda,
                       const float* ia, const void* B, int64_t ldb, const float* ib, double beta,
                       const void* D, void* C, const int64_t* cdims, const int64_t* sc,
                       int64_t splits, double* ws, void* stream) {
  if (!cdims || !sc) return NL_EARG;
  for (int d = 0; d < 4; ++d)
    if (cdims[d] < 1) return NL_EARG;
  if (cdims[1] > INT32_MAX || cdims[3] > INT32_MAX) return NL_EARG;
  return gemm_lp_impl(dtype, cdims[0] * cdims[1], cdims[2] * cdims[3], K, alpha, A, lda, 0, ia, 0,
                      B, ldb, 0, ib, 0, beta, D, C, mk(cdims[1], sc[0], sc[1]),
                      mk(cdims[3], sc[2], sc[3]), 0, 1, splits, ws, stream);
}

int nl_gemm_lp_split(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                     const void* Alo, int64_t lda, int64_t strideA, const float* ia,
                     int64_t strideIa, const void* B, const void* Blo, int64_t ldb,
                     int64_t strideB, const float* ib, int64_t strideIb, double beta,
                     const void* D, void* C, int64_t ldc, int64_t strideC, int64_t batch,
                     int64_t splits, double* ws, void* stream) {
  return gemm_lp_impl(dtype, M, N, K, alpha, A, lda, strideA, ia, strideIa, B, ldb, strideB, ib,
                      strideIb, beta, D, C, mk(1, ldc, 0), mk(1, 1, 0), strideC, batch, splits,
                      ws, stream, nullptr, Alo, Blo, true);
}

int nl_gemm_lp_strided_split(int dtype, int64_t K, double alpha, const void* A, const void* Alo,
                             int64_t lda, const float* ia, const void* B, const void* Blo,
                             int64_t ldb, const float* ib, double beta, const void* D,
                             void* C, const int64_t* cdims, const int64_t* sc, int64_t splits,
                             double* ws, void* stream) {
  if (!cdims || !sc) return NL_EARG;
  for (int d = 0; d < 4; ++d)
    if (cdims[d] < 1) return NL_EARG;
  if (cdims[1] > INT32_MAX || cdims[3] > INT32_MAX) return NL_EARG;
  return gemm_lp_impl(dtype, cdims[0] * cdims[1], cdims[2] * cdims[3], K, alpha, A, lda, 0, ia, 0,
                      B, ldb, 0, ib, 0, beta, D, C, mk(cdims[1], sc[0], sc[1]),
                      mk(cdims[3], sc[2], sc[3]), 0, 1, splits, ws, stream, nullptr, Alo, Blo,
                      true);
}

int nl_gemm_lp_rows_split(int dtype, int64_t M, int64_t N, int64_t K, double alpha,
                          const void* A, const void* Alo, int64_t lda, int64_t strideA,
                          const float* ia, int64_t strideIa, const void* B, const void* Blo,
                          int64_t ldb, int64_t strideB, const float* ib, int64_t strideIb,
                          void* C, int64_t ldc, const int64_t* crow, int64_t batch,
                          void* stream) {
  if (!crow || ldc < N) return NL_EARG;
  if (dtype & (NL_C_F32 << 8)) dtype |= NL_D_F32 << 8;       // the rows add to C itself
  return gemm_lp_impl(dtype, M, N, K, alpha, A, lda, strideA, ia, strideIa, B, ldb, strideB, ib,
                      strideIb, 1.0, C, C, mk(1, ldc, 0), mk(1, 1, 0), 0, batch, 1, nullptr,
                      stream, crow, Alo, Blo, true);
}

}  // extern "C"

// Bandwidth-bound kernels of the CCSD iteration on gfx950: layout changes, precision conversion
// with per-row scaling, COO (segment) accumulations, the fused T2 assembly/Jacobi update/energy,
// and deterministic reductions.  Every kernel is a grid-stride loop over coalesced output
// elements; reductions use wavefront shuffles and a fixed-order second pass (bitwise
// reproducible run to run).
#include <algorithm>
#include <mutex>

#include "nl_common.h"

namespace {

constexpr int NT = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide sum; result valid in thread 0.
__device__ inline double block_sum(double v) {
  __shared__ double part[NT / 64];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += part[i];
  __syncthreads();
  return s;
}

struct PermArgs {
  int rank;
  int64_t shape[6];    // output shape (row-major, contiguous output)
  int64_t istr[6];     // input element stride of each output dimension
  FastDiv fd[6];       // divisors of the shape (used when total < 2^31)
  int64_t total;
  double alpha, beta;
  const double* in;
  double* out;
};

template <bool SMALL>   // SMALL: total < 2^31, coordinates by multiply-shift division
__global__ void permute_kernel(PermArgs a) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < a.total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t off = 0;
    if (SMALL) {
      int rem = (int)t;
      for (int d = a.rank - 1; d >= 0; --d) {
        const int q = fdiv(rem, a.fd[d]);
        off += (int64_t)(rem - q * a.fd[d].d) * a.istr[d];
        rem = q;
      }
    } else {
      int64_t rem = t;
      for (int d = a.rank - 1; d >= 0; --d) {
        int64_t q = rem / a.shape[d];
        off += (rem - q * a.shape[d]) * a.istr[d];
        rem = q;
      }
    }
    double v = a.alpha * a.in[off];
    a.out[t] = (a.beta != 0.0) ? v + a.beta * a.out[t] : v;
  }
}

struct ConvArgs {
  int64_t R2, C, C2, ldo;   // rows = R1*R2 (grid), cols = C1*C2 = C, output row pitch ldo
  int64_t cend;             // columns written per row: the C values, then zeros up to cend
  FastDiv fr2, fc2;         // divisors R2, C2 (rows and columns are below 2^31)
  int64_t s[4];             // strides of (r1, r2, c1, c2)
  const double* in;
  uint16_t* out;
  float* iscale;            // per-row inverse scale (power of two)
  float target;             // row max is scaled to at most this
  int bf16;
};

__device__ inline uint16_t to_bf16_bits(float f) {
  __bf16 b = (__bf16)f;  // round to nearest even
  return __builtin_bit_cast(uint16_t, b);
}

// BF16 of an FP64 value rounded twice, FP64 -> FP32 -> BF16 (each to nearest even), as the cast
// kernels round it.  The FP32 value passes through an empty asm so that the compiler cannot fold
// the two truncations into one FP64 -> BF16 rounding, which differs in rare ties-after-rounding
// cases (an FP64 value whose FP32 rounding lands exactly halfway between two BF16 values).
__device__ inline uint16_t to_bf16_bits_2r(double x) {
  float f = (float)x;
  asm volatile("" : "+v"(f));
  return to_bf16_bits(f);
}
__device__ inline uint16_t to_f16_bits(float f) {
  _Float16 h = (_Float16)f;
  return __builtin_bit_cast(uint16_t, h);
}

// Largest e with 2^e <= target/amax (clamped), the per-row power-of-two scale exponent.
__device__ inline int scale_exp(float amax, float target) {
  int e = 0;
  if (amax > 0.f && isfinite(amax)) {
    int ex;
    frexpf(target / amax, &ex);   // target/amax = f * 2^ex, f in [0.5, 1)
    e = max(-126, min(126, ex - 1));
  }
  return e;
}

// Per-row scale and inverse scale.  target > 0: the power of two 2^e <= target/amax (exact
// rescaling, the row max lands in [target/2, target]); target < 0: exactly |target|/amax, so a row
// whose nonzero elements share one magnitude (a contact interaction's blocks: every element is
// +-g) converts to exactly +-|target| and its magnitude lives in the FP32 inverse scale.
__device__ inline float row_scale(float amax, float target) {
  if (target > 0.f) return ldexpf(1.f, scale_exp(amax, target));
  return (amax > 0.f && isfinite(amax)) ? -target / amax : 1.f;
}
__device__ inline float row_iscale(float amax, float target) {
  if (target > 0.f) return ldexpf(1.f, -scale_exp(amax, target));
  return (amax > 0.f && isfinite(amax)) ? amax / -target : 1.f;
}

// One workgroup per output row: amax pass, then scaled conversion (per-row power-of-two scale).
// Used when the row's elements are unit-stride or when the rows themselves are strided.
__global__ __launch_bounds__(NT) void convert_rows_kernel(ConvArgs a) {
  __shared__ float red[NT / 64];
  int64_t r = blockIdx.x;
  int64_t r1 = fdiv((int)r, a.fr2), r2 = r - r1 * a.R2;
  const double* base = a.in + r1 * a.s[0] + r2 * a.s[1];
  float m = 0.f;
  for (int64_t c = threadIdx.x; c < a.C; c += NT) {
    int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
    m = fmaxf(m, (float)fabs(base[c1 * a.s[2] + c2 * a.s[3]]));
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  float amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  const float sc = row_scale(amax, a.target);
  if (threadIdx.x == 0) a.iscale[r] = row_iscale(amax, a.target);
  uint16_t* o = a.out + r * a.ldo;
  for (int64_t c = threadIdx.x; c < a.cend; c += NT) {
    float v = 0.f;
    if (c < a.C) {
      int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
      v = (float)base[c1 * a.s[2] + c2 * a.s[3]] * sc;
    }
    o[c] = a.bf16 ? to_bf16_bits(v) : to_f16_bits(v);
  }
}

// Unscaled cast (BF16 keeps FP32's exponent range, so its relative precision does not depend on
// a row scale): a 2D grid of (row, 2048-column chunk), so long rows are spread over workgroups.
constexpr int CAST_CHUNK = 2048;
__global__ __launch_bounds__(NT) void cast_rows_kernel(ConvArgs a) {
  const int64_t r = blockIdx.x, r1 = fdiv((int)r, a.fr2), r2 = r - r1 * a.R2;
  const double* base = a.in + r1 * a.s[0] + r2 * a.s[1];
  uint16_t* o = a.out + r * a.ldo;
  const int64_t c0 = (int64_t)blockIdx.y * CAST_CHUNK, c1e = min(a.cend, c0 + CAST_CHUNK);
  for (int64_t c = c0 + threadIdx.x; c < c1e; c += NT) {
    float v = 0.f;
    if (c < a.C) {
      int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
      v = (float)base[c1 * a.s[2] + c2 * a.s[3]];
    }
    o[c] = a.bf16 ? to_bf16_bits(v) : to_f16_bits(v);
  }
}

// ---- hi/lo split conversion (precision form with split operands)
// x*sc = hi + lo + r with hi = 16-bit(x*sc), lo = 16-bit(x*sc - hi): |r| <= 2^-9 |lo| (BF16) or
// 2^-12 |lo| (FP16), so the pair carries about 17 (BF16) or 23 (FP16) significant bits.  A GEMM
// over K-concatenated blocks [hi | lo | hi] x [hi | hi | lo] then sums hi.hi + lo.hi + hi.lo, the
// product to that precision without the lo.lo term (Henry, Tang & Heinecke, ARITH 2019 for BF16;
// the same split in FP16 with per-row scales).
struct SplitArgs {
  int64_t R2, C, C2, ldo, kpad;   // rows = R1*R2, cols C = C1*C2, row pitch ldo, block pitch kpad
  FastDiv fr2, fc2;
  int64_t s[4];
  AnyIn in;                       // FP64 or FP32 source
  AnyIn sub;                      // (SUB kernels) subtrahend with the source's type and layout
  uint16_t* out;
  float* iscale;                  // per-row inverse scale (null: unscaled)
  const float* given;             // per-row inverse scales to use (no amax pass), or null
  float target;
  int nblk, lomask, bf16;         // block q holds the lo part when lomask bit q is set, else hi
  double scale;                   // fixed factor applied before the split (with the row scale)
};

__device__ inline float from16(uint16_t b, int bf16) {
  return bf16 ? __uint_as_float((unsigned)b << 16)
              : (float)__builtin_bit_cast(_Float16, b);
}

__device__ inline void split16(double xs, int bf16, uint16_t& hi, uint16_t& lo) {
  // FP64 -> FP32 -> 16 bits, each step a hardware conversion: the FP32 values pass through an
  // empty asm so that the compiler cannot fold the two roundings into one FP64 -> FP16 rounding
  // (a long software sequence on this target)
  float f = (float)xs;
  asm volatile("" : "+v"(f));
  hi = bf16 ? to_bf16_bits(f) : to_f16_bits(f);
  float r = (float)(xs - (double)from16(hi, bf16));
  asm volatile("" : "+v"(r));
  lo = bf16 ? to_bf16_bits(r) : to_f16_bits(r);
}

// Element o of the value a split converts: the source, less the subtrahend (SUB; exact in FP64
// for an FP32 pair).
template <bool SUB>
__device__ inline double split_value(const AnyIn& base, const AnyIn& sbase, int64_t o) {
  return SUB ? base.get(o) - sbase.get(o) : base.get(o);
}

// One workgroup per output row: amax pass (when scaled), then the split written to every block.
template <bool SUB>
__global__ __launch_bounds__(NT) void split_rows_kernel(SplitArgs a) {
  __shared__ float red[NT / 64];
  const int64_t r = blockIdx.x;
  const int64_t r1 = fdiv((int)r, a.fr2), r2 = r - r1 * a.R2;
  const AnyIn base = a.in.at(r1 * a.s[0] + r2 * a.s[1]);
  const AnyIn sbase = SUB ? a.sub.at(r1 * a.s[0] + r2 * a.s[1]) : base;
  float sc = 1.f;
  if (a.given) {
    sc = 1.f / a.given[r];
  } else if (a.iscale) {
    float m = 0.f;
    for (int64_t c = threadIdx.x; c < a.C; c += NT) {
      const int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
      m = fmaxf(m, (float)fabs(split_value<SUB>(base, sbase, c1 * a.s[2] + c2 * a.s[3])));
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    const float amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    sc = row_scale(amax, a.target);
    if (threadIdx.x == 0) a.iscale[r] = row_iscale(amax, a.target);
  }
  uint16_t* o = a.out + r * a.ldo;
  for (int64_t c = threadIdx.x; c < a.kpad; c += NT) {
    uint16_t hi = 0, lo = 0;
    if (c < a.C) {
      const int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
      split16(split_value<SUB>(base, sbase, c1 * a.s[2] + c2 * a.s[3]) * ((double)sc * a.scale),
              a.bf16, hi, lo);
    }
    for (int q = 0; q < a.nblk; ++q) o[q * a.kpad + c] = ((a.lomask >> q) & 1) ? lo : hi;
  }
}

// Transposing split (output rows are the unit-stride direction of the source): 64x64 tiles
// through LDS, per-row scales from amax bits gathered by amax_split_kernel.
template <bool SUB>
__global__ __launch_bounds__(NT) void amax_split_kernel(SplitArgs a, int64_t R, unsigned* bits) {
  __shared__ float red[4][64];
  const int tr = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int64_t r = blockIdx.x * 64 + tr, c0 = blockIdx.y * 64;
  float m = 0.f;
  if (r < R) {
    const int64_t r1 = fdiv((int)r, a.fr2), r2 = r - r1 * a.R2;
    const AnyIn base = a.in.at(r1 * a.s[0] + r2 * a.s[1]);
    const AnyIn sbase = SUB ? a.sub.at(r1 * a.s[0] + r2 * a.s[1]) : base;
    for (int cc = tg; cc < 64; cc += 4) {
      const int64_t c = c0 + cc;
      if (c < a.C) {
        const int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
        m = fmaxf(m, (float)fabs(split_value<SUB>(base, sbase, c1 * a.s[2] + c2 * a.s[3])));
      }
    }
  }
  red[tg][tr] = m;
  __syncthreads();
  if (tg == 0 && r < R) {
    m = fmaxf(fmaxf(red[0][tr], red[1][tr]), fmaxf(red[2][tr], red[3][tr]));
    atomicMax(&bits[r], __float_as_uint(m));
  }
}

template <bool SUB>
__global__ __launch_bounds__(NT) void split_tiles_kernel(SplitArgs a, int64_t R,
                                                         const unsigned* bits) {
  __shared__ uint16_t thi[64][66], tlo[64][66];
  const int tr = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int64_t r0 = blockIdx.x * 64, c0 = blockIdx.y * 64, r = r0 + tr;
  if (r < R) {
    const float sc = a.given ? 1.f / a.given[r]
                     : bits ? row_scale(__uint_as_float(bits[r]), a.target) : 1.f;
    const int64_t r1 = fdiv((int)r, a.fr2), r2 = r - r1 * a.R2;
    const AnyIn base = a.in.at(r1 * a.s[0] + r2 * a.s[1]);
    const AnyIn sbase = SUB ? a.sub.at(r1 * a.s[0] + r2 * a.s[1]) : base;
    for (int cc = tg; cc < 64; cc += 4) {
      const int64_t c = c0 + cc;
      uint16_t hi = 0, lo = 0;
      if (c < a.C) {
        const int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
        split16(split_value<SUB>(base, sbase, c1 * a.s[2] + c2 * a.s[3]) * ((double)sc * a.scale),
                a.bf16, hi, lo);
      }
      thi[tr][cc] = hi;
      tlo[tr][cc] = lo;
    }
  }
  __syncthreads();
  for (int rr = tg; rr < 64; rr += 4) {
    const int64_t row = r0 + rr, c = c0 + tr;
    if (row < R && c < a.kpad)
      for (int q = 0; q < a.nblk; ++q)
        a.out[row * a.ldo + q * a.kpad + c] = ((a.lomask >> q) & 1) ? tlo[rr][tr] : thi[rr][tr];
  }
}

__global__ __launch_bounds__(NT) void bits_to_iscale_kernel(int64_t R, float target,
                                                            float* iscale) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < R;
       t += (int64_t)gridDim.x * blockDim.x)
    iscale[t] = row_iscale(__uint_as_float(__float_as_uint(iscale[t])), target);
}

// Transposing variant (the output rows are the unit-stride direction of the source): 64x64
// tiles through LDS so that both the FP64 reads and the 16-bit writes are coalesced.
//   pass 1: per-row amax (as float bits, atomicMax; |x| >= 0 orders like its bit pattern)
//   pass 2: scaled conversion, transposed through LDS
//   pass 3: amax bits -> inverse scales
__global__ __launch_bounds__(NT) void zero_u32_kernel(int64_t n, unsigned* x) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    x[t] = 0u;
}

__global__ __launch_bounds__(NT) void amax_tiles_kernel(ConvArgs a, int64_t R, unsigned* bits) {
  __shared__ float red[4][64];
  const int tr = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int64_t r = blockIdx.x * 64 + tr, c0 = blockIdx.y * 64;
  float m = 0.f;
  if (r < R) {
    int64_t r1 = fdiv((int)r, a.fr2), r2 = r - r1 * a.R2;
    const double* base = a.in + r1 * a.s[0] + r2 * a.s[1];
    for (int cc = tg; cc < 64; cc += 4) {
      int64_t c = c0 + cc;
      if (c < a.C) {
        int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
        m = fmaxf(m, (float)fabs(base[c1 * a.s[2] + c2 * a.s[3]]));
      }
    }
  }
  red[tg][tr] = m;
  __syncthreads();
  if (tg == 0 && r < R) {
    m = fmaxf(fmaxf(red[0][tr], red[1][tr]), fmaxf(red[2][tr], red[3][tr]));
    atomicMax(&bits[r], __float_as_uint(m));
  }
}

__global__ __launch_bounds__(NT) void convert_tiles_kernel(ConvArgs a, int64_t R,
                                                           const unsigned* bits) {
  __shared__ uint16_t tile[64][66];
  const int tr = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int64_t r0 = blockIdx.x * 64, c0 = blockIdx.y * 64, r = r0 + tr;
  if (r < R) {
    float sc = bits ? row_scale(__uint_as_float(bits[r]), a.target) : 1.f;
    int64_t r1 = fdiv((int)r, a.fr2), r2 = r - r1 * a.R2;
    const double* base = a.in + r1 * a.s[0] + r2 * a.s[1];
    for (int cc = tg; cc < 64; cc += 4) {
      int64_t c = c0 + cc;
      float v = 0.f;
      if (c < a.C) {
        int64_t c1 = fdiv((int)c, a.fc2), c2 = c - c1 * a.C2;
        v = (float)base[c1 * a.s[2] + c2 * a.s[3]] * sc;
      }
      tile[tr][cc] = a.bf16 ? to_bf16_bits(v) : to_f16_bits(v);
    }
  }
  __syncthreads();
  for (int rr = tg; rr < 64; rr += 4) {
    int64_t row = r0 + rr, c = c0 + tr;
    if (row < R && c < a.cend) a.out[row * a.ldo + c] = tile[rr][tr];
  }
}

// out[b][n][k] = bf16(src[idx[b*kb + k]*lds + n]) (0 where idx < 0 or k >= kb, k < ldk):
// the row gather of a packed-block product fused with its transposing BF16 conversion.  Rows
// r = b*ncol + n of the output are tiled 64 x 64 with k through LDS, so the FP64 reads run along
// n (source rows) and the 16-bit writes along k.
template <typename S>   // source rows: double (cast here) or uint16_t (BF16 already)
__global__ __launch_bounds__(NT) void gather_convert_kernel(int64_t R, int64_t ncol, FastDiv fn,
                                                            int64_t kb, int64_t ldk,
                                                            const int64_t* idx, const S* src,
                                                            int64_t lds, uint16_t* out,
                                                            int64_t ldo) {
  __shared__ uint16_t tile[64][66];
  const int tr = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int64_t r0 = blockIdx.x * 64, c0 = blockIdx.y * 64, r = r0 + tr;
  if (r < R) {
    const int64_t b = fdiv((int)r, fn), n = r - b * ncol;
    for (int cc = tg; cc < 64; cc += 4) {
      const int64_t k = c0 + cc;
      uint16_t v = 0;
      if (k < kb) {
        const int64_t row = idx[b * kb + k];
        if (row >= 0) {
          if constexpr (sizeof(S) == 8) v = to_bf16_bits((float)src[row * lds + n]);
          else v = src[row * lds + n];
        }
      }
      tile[tr][cc] = v;
    }
  }
  __syncthreads();
  for (int rr = tg; rr < 64; rr += 4) {
    const int64_t row = r0 + rr, c = c0 + tr;
    if (row < R && c < ldk) out[row * ldo + c] = tile[rr][tr];
  }
}

// out[c*ldo + r] = in[r*ldi + c] for 16-bit elements (r < R, c < C), and 0 for R <= r < rlen:
// 64 x 64 tiles through LDS, reads along c and writes along r coalesced.
__global__ __launch_bounds__(NT) void transpose16_kernel(int64_t R, int64_t C, const uint16_t* in,
                                                         int64_t ldi, int64_t si, uint16_t* out,
                                                         int64_t ldo, int64_t so, int64_t rlen) {
  __shared__ uint16_t tile[64][66];
  const int tc = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int64_t r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  in += blockIdx.z * si;
  out += blockIdx.z * so;
  for (int rr = tg; rr < 64; rr += 4) {
    const int64_t r = r0 + rr, c = c0 + tc;
    tile[rr][tc] = (r < R && c < C) ? in[r * ldi + c] : (uint16_t)0;
  }
  __syncthreads();
  for (int cc = tg; cc < 64; cc += 4) {
    const int64_t c = c0 + cc, r = r0 + tc;
    if (c < C && r < rlen) out[c * ldo + r] = tile[tc][cc];
  }
}

// BF16 source rows: out[b][n][k] = src[idx[b*kb + k]*lds + n].  Tiles of 64 k x 256 n: each
// thread reads 4 consecutive n (8 bytes) of one source row, and writes 4 consecutive k of one
// output row (8 bytes) from the LDS tile.  Needs ncol % 4 == 0, lds % 4 == 0 and 8-byte aligned
// rows (checked on the host).
// NPART = 2: the source rows are [hi(ncol) | lo(ncol)] and part q of column n lands at
// out[(b*ncol + n)*ldo + q*kpart + k] (the [hi(kpart) | lo(kpart)] rows of a split operand).
template <int NPART>
__global__ __launch_bounds__(NT) void gather16_kernel(int64_t ncol, int64_t kb, int64_t ldk,
                                                      const int64_t* idx, const uint16_t* src,
                                                      int64_t lds, uint16_t* out, int64_t ldo,
                                                      int64_t kpart) {
  __shared__ uint16_t tile[64][256 + 4];
  const int b = blockIdx.z, tid = threadIdx.x;
  const int64_t n0 = (int64_t)blockIdx.x * 256, k0 = (int64_t)blockIdx.y * 64;
  const int64_t nall = NPART * ncol;
  const int nl = (tid & 63) * 4;                 // 4 n per thread, 64 threads per k row
#pragma unroll 4
  for (int kk = tid >> 6; kk < 64; kk += 4) {
    const int64_t k = k0 + kk, n = n0 + nl;
    uint2 v = make_uint2(0u, 0u);
    if (k < kb && n < nall) {
      const int64_t row = idx[(int64_t)b * kb + k];
      if (row >= 0) v = *(const uint2*)(src + row * lds + n);
    }
    *(uint2*)&tile[kk][nl] = v;
  }
  __syncthreads();
  const int kq = (tid & 15) * 4;                 // 4 k per thread, 16 threads per n row
#pragma unroll 4
  for (int nn = tid >> 4; nn < 256; nn += 16) {
    int64_t n = n0 + nn;
    const int64_t k = k0 + kq;
    if (n >= nall || k >= ldk) continue;
    int64_t o = k;
    if (NPART == 2 && n >= ncol) { n -= ncol; o += kpart; }
    uint16_t w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = tile[kq + u][nn];
    *(uint2*)(out + ((int64_t)b * ncol + n) * ldo + o) =
        make_uint2(w[0] | ((unsigned)w[1] << 16), w[2] | ((unsigned)w[3] << 16));
  }
}

// transpose16 with 8-byte accesses (R, C, ldi, ldo, strides multiples of 4; 8-byte aligned):
// reads 4 consecutive c of one row per lane, writes 4 consecutive r of one output row per lane.
__global__ __launch_bounds__(NT) void transpose16v_kernel(int64_t R, int64_t C, const uint16_t* in,
                                                          int64_t ldi, int64_t si, uint16_t* out,
                                                          int64_t ldo, int64_t so, int64_t rlen) {
  __shared__ uint16_t tile[64][64 + 4];
  const int tid = threadIdx.x, q4 = (tid & 15) * 4, row = tid >> 4;
  const int64_t r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  in += blockIdx.z * si;
  out += blockIdx.z * so;
#pragma unroll
  for (int rr = row; rr < 64; rr += 16) {
    const int64_t r = r0 + rr, c = c0 + q4;
    uint2 v = make_uint2(0u, 0u);
    if (r < R && c < C) v = *(const uint2*)(in + r * ldi + c);
    tile[rr][q4 + 0] = (uint16_t)(v.x & 0xffffu);
    tile[rr][q4 + 1] = (uint16_t)(v.x >> 16);
    tile[rr][q4 + 2] = (uint16_t)(v.y & 0xffffu);
    tile[rr][q4 + 3] = (uint16_t)(v.y >> 16);
  }
  __syncthreads();
#pragma unroll
  for (int cc = row; cc < 64; cc += 16) {
    const int64_t c = c0 + cc, r = r0 + q4;
    if (c >= C || r >= rlen) continue;
    *(uint2*)(out + c * ldo + r) =
        make_uint2(tile[q4][cc] | ((unsigned)tile[q4 + 1][cc] << 16),
                   tile[q4 + 2][cc] | ((unsigned)tile[q4 + 3][cc] << 16));
  }
}

// 16-bit permute: out (contiguous, shape fd[].d) = in[sum_d idx_d * istr[d]], total < 2^31.
// V = 4: elements move as 4-element (8-byte) vectors along a last dimension that is unit-stride
// in both (extent and strides divided by 4 on the host).
template <int V>
struct Perm16Args {
  int rank;
  int64_t istr[6];
  FastDiv fd[6];
  int64_t total;
  const uint16_t* in;
  uint16_t* out;
};

template <int V>
__global__ __launch_bounds__(NT) void permute16_kernel(Perm16Args<V> a) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < a.total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int rem = (int)t;
    int64_t off = 0;
    for (int d = a.rank - 1; d >= 0; --d) {
      const int q = fdiv(rem, a.fd[d]);
      off += (int64_t)(rem - q * a.fd[d].d) * a.istr[d];
      rem = q;
    }
    if (V == 4) *(uint2*)(a.out + t * 4) = *(const uint2*)(a.in + off * 4);
    else a.out[t] = a.in[off];
  }
}

__global__ __launch_bounds__(NT) void iscale_kernel(int64_t R, float target, float* iscale) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < R;
       t += (int64_t)gridDim.x * blockDim.x) {
    float amax = __uint_as_float(__float_as_uint(iscale[t]));
    iscale[t] = row_iscale(amax, target);
  }
}

// tau[a,b,i,j] = t2[a,b,i,j] + t1[a,i] t1[b,j] - t1[a,j] t1[b,i] with the two fused
// multiply-adds written out, so its rounding does not depend on how the compiler contracts the
// expression.
__device__ inline double tau_value(double x, double ai, double bj, double aj, double bi) {
  return fma(-aj, bi, fma(ai, bj, x));
}

// The tau kernel: workgroups walk the pairs (a, b), threads the h x h block (contiguous in t2
// and tau).
template <typename T>   // double, or uint16_t for a BF16 tau (the rounded FP64 value)
__global__ __launch_bounds__(NT) void tau_kernel(int p, int h, FastDiv fp, FastDiv fh,
                                                 const double* t1, const double* t2, T* tau) {
  // the pair's two t1 rows are staged in LDS (h <= NT), so an element costs one global load
  __shared__ double ta[NT], tb[NT];
  const int hh = h * h;
  for (int q = blockIdx.x; q < p * p; q += gridDim.x) {
    const int a = fdiv(q, fp), b = q - a * p;
    const double* x = t2 + (int64_t)q * hh;
    T* y = tau + (int64_t)q * hh;
    __syncthreads();
    if (threadIdx.x < h) {
      ta[threadIdx.x] = t1[a * h + threadIdx.x];
      tb[threadIdx.x] = t1[b * h + threadIdx.x];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < hh; e += NT) {
      const int i = fdiv(e, fh), j = e - i * h;
      const double v = tau_value(x[e], ta[i], tb[j], ta[j], tb[i]);
      if constexpr (sizeof(T) == 8) y[e] = v;
      else y[e] = to_bf16_bits_2r(v);
    }
  }
}

// tau as the hi/lo split of the precision form with split operands: row (a,b) of out holds
// [hi(h*h) | lo(h*h)] (row pitch 2*h*h), optionally with the FP64 tau beside it.
__global__ __launch_bounds__(NT) void tau_split_kernel(int p, int h, FastDiv fp, FastDiv fh,
                                                       const double* t1, const double* t2,
                                                       double* tau, uint16_t* out, int bf16,
                                                       double scale) {
  __shared__ double ta[NT], tb[NT];
  const int hh = h * h;
  for (int q = blockIdx.x; q < p * p; q += gridDim.x) {
    const int a = fdiv(q, fp), b = q - a * p;
    const double* x = t2 + (int64_t)q * hh;
    uint16_t* y = out + (int64_t)q * 2 * hh;
    __syncthreads();
    if (threadIdx.x < h) {
      ta[threadIdx.x] = t1[a * h + threadIdx.x];
      tb[threadIdx.x] = t1[b * h + threadIdx.x];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < hh; e += NT) {
      const int i = fdiv(e, fh), j = e - i * h;
      const double v = tau_value(x[e], ta[i], tb[j], ta[j], tb[i]);
      if (tau) tau[(int64_t)q * hh + e] = v;
      uint16_t hi, lo;
      split16(v * scale, bf16, hi, lo);
      y[e] = hi;
      y[hh + e] = lo;
    }
  }
}

// Hole pairs i < j of an h x h block, numbered row by row: q(i,j) = i*h - i*(i+1)/2 + j - i - 1.
// pair_of(q) = (i, j) by walking the rows (at most h steps).
__device__ inline void pair_of(int q, int h, int& i, int& j) {
  int r = 0, start = 0;
  while (start + (h - 1 - r) <= q) { start += h - 1 - r; ++r; }
  i = r;
  j = r + 1 + (q - start);
}

// The tau of tau_split_kernel over the hole pairs i < j only: row (a, b) of out is
// [hi(P) 0... | lo(P) 0...] at pitch 2*Pp (P = h(h-1)/2 pairs, columns P <= q < Pp zero).
// tau is antisymmetric in (i, j), so these rows carry all of it.
__global__ __launch_bounds__(NT) void tau_split_pairs_kernel(int64_t pp, int p, int h, int P,
                                                             int Pp, int grp, FastDiv fp,
                                                             FastDiv fPp, FastDiv fgrp,
                                                             const double* t1, const double* t2,
                                                             uint16_t* out, int bf16,
                                                             double scale) {
  const int64_t total = pp * Pp;                         // < 2^31 (checked by the host)
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * NT) {
    const int64_t row = fdiv((int)t, fPp);
    const int q = (int)(t - row * Pp);
    uint16_t hi = 0, lo = 0;
    if (q < P) {
      int i, j;
      pair_of(q, h, i, j);
      const int a = fdiv((int)row, fp), b = (int)row - a * p;
      const double x = t2[row * h * h + i * h + j];
      const double v = t1 ? tau_value(x, t1[a * h + i], t1[b * h + j], t1[a * h + j],
                                      t1[b * h + i]) : x;
      split16(v * scale, bf16, hi, lo);
    }
    // rows grouped by grp: group g = row / grp holds [hi(grp * Pp) | lo(grp * Pp)]
    const int64_t gq = fdiv((int)row, fgrp), rin = row - gq * grp;
    uint16_t* y = out + gq * 2 * grp * Pp + rin * Pp + q;
    y[0] = hi;
    y[(int64_t)grp * Pp] = lo;
  }
}

// The ring image of t2 from its hole-pair image: out row (a,i) = [hi((c,k)) | lo((c,k))] of
// t2[a,c,i,k] = +-in[a][part][(c, q(min(i,k), max(i,k)))] (0 for i = k): in row a holds
// [hi((c,q)) | lo((c,q))], Pp slots per c.  A 16-bit value is negated by its sign bit.
// One workgroup per (a, block of ct columns c): both parts' slots of those c are read once into
// LDS with 16-byte loads, then every output row (a,i) receives its two contiguous segments of
// ct*h elements, 8 per thread and store (16 bytes when VEC: segment starts 8-aligned).
template <bool VEC>
__global__ __launch_bounds__(NT) void ring_from_pairs_kernel(int p, int h, int Pp, int ct,
                                                             int nct, const uint16_t* in,
                                                             uint16_t* out) {
  extern __shared__ uint4 ring_lds_raw[];
  uint16_t* lds = reinterpret_cast<uint16_t*>(ring_lds_raw);     // [part][cl][Pp]
  const int a = blockIdx.x / nct, c0 = (blockIdx.x - a * nct) * ct;
  const int cta = min(ct, p - c0);
  const int64_t half = (int64_t)p * Pp, ph = (int64_t)p * h;
  const int seg = cta * Pp;                                       // slots per part, mult of 32
  for (int v = threadIdx.x; v < 2 * seg / 8; v += NT) {
    const int part = v >= seg / 8, o = (v - part * (seg / 8)) * 8;
    const uint4 x = *reinterpret_cast<const uint4*>(in + a * 2 * half + part * half +
                                                    (int64_t)c0 * Pp + o);
    *reinterpret_cast<uint4*>(lds + part * ct * Pp + o) = x;
  }
  __syncthreads();
  const int len = cta * h, nch = (len + 7) / 8;                   // a segment's 8-chunks
  for (int w = threadIdx.x; w < h * 2 * nch; w += NT) {
    const int i = w / (2 * nch), rem = w - i * 2 * nch;
    const int part = rem >= nch, ch = rem - part * nch;
    const int s0 = ch * 8;
    int cl = s0 / h, k = s0 - cl * h;
    const uint16_t* src = lds + part * ct * Pp;
    uint16_t val[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      uint16_t v = 0;
      if (s0 + u < len && i != k) {
        const int lo = i < k ? i : k, hi = i < k ? k : i;
        v = src[cl * Pp + lo * h - lo * (lo + 1) / 2 + hi - lo - 1];
        if (i > k) v ^= 0x8000;
      }
      val[u] = v;
      if (++k == h) { k = 0; ++cl; }
    }
    uint16_t* dst = out + ((int64_t)a * h + i) * 2 * ph + part * ph + (int64_t)c0 * h + s0;
    if (VEC && s0 + 8 <= len) {
      uint4 x;
      x.x = val[0] | ((uint32_t)val[1] << 16); x.y = val[2] | ((uint32_t)val[3] << 16);
      x.z = val[4] | ((uint32_t)val[5] << 16); x.w = val[6] | ((uint32_t)val[7] << 16);
      *reinterpret_cast<uint4*>(dst) = x;
    } else {
      for (int u = 0; u < 8 && s0 + u < len; ++u) dst[u] = val[u];
    }
  }
}

// Reduce: out[r][q][w] = in[r][i][j][w] for the pairs q = (i < j); expand: the antisymmetric
// block, out[r][i][j][w] = in[r][q(i,j)][w] (i < j), -in[r][q(j,i)][w] (i > j), 0 (i = j).
// Pair-form rows hold ldp >= P pair slots (the slots beyond P are neither read nor written).
__global__ __launch_bounds__(NT) void hole_pairs_kernel(int64_t R, int h, int P, int64_t ldp,
                                                        int64_t w, const double* in, double* out,
                                                        int expand) {
  const int64_t per = (expand ? (int64_t)h * h : P) * w, total = R * per;
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * NT) {
    const int64_t r = t / per, rem = t - r * per, x = rem / w, c = rem - x * w;
    if (expand) {
      const int i = (int)(x / h), j = (int)(x - (int64_t)i * h);
      double v = 0.0;
      if (i < j) v = in[(r * ldp + (i * h - i * (i + 1) / 2 + j - i - 1)) * w + c];
      else if (i > j) v = -in[(r * ldp + (j * h - j * (j + 1) / 2 + i - j - 1)) * w + c];
      out[t] = v;
    } else {
      int i, j;
      pair_of((int)x, h, i, j);
      out[(r * ldp + x) * w + c] = in[((r * h + i) * h + j) * w + c];
    }
  }
}

// out[dst[s] + x*ds] += alpha * sum_{e in [ptr[s], ptr[s+1])} val[e] * src[so[e] + x*ss]
// G lanes per segment (G = the power of two >= X, at most 64), so a wavefront serves 64/G
// segments; each lane keeps two partial sums (entries alternate between them, added in a fixed
// order), which keeps two independent loads in flight along a segment.
template <typename T>   // output elements: double, or float for an FP32 accumulator
__global__ __launch_bounds__(NT) void segsum_kernel(int64_t S, int64_t X, int G,
                                                  const int64_t* ptr, const int64_t* so,
                                                  const double* val, const int64_t* dst,
                                                  int64_t ss, int64_t ds, double alpha,
                                                  const double* src, T* out) {
  const int lane = threadIdx.x & 63, sub = lane / G, xl = lane - sub * G, per = 64 / G;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t wv = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; wv * per < S;
       wv += nw) {
    const int64_t s = wv * per + sub;
    if (s >= S) continue;
    const int64_t e0 = ptr[s], e1 = ptr[s + 1], d = dst[s];
    for (int64_t x = xl; x < X; x += G) {
      double acc0 = 0.0, acc1 = 0.0;
      int64_t e = e0;
      for (; e + 1 < e1; e += 2) {
        acc0 += val[e] * src[so[e] + x * ss];
        acc1 += val[e + 1] * src[so[e + 1] + x * ss];
      }
      if (e < e1) acc0 += val[e] * src[so[e] + x * ss];
      out[d + x * ds] = (T)((double)out[d + x * ds] + alpha * (acc0 + acc1));
    }
  }
}

// out[r][c] = idx[r] >= 0 ? src[idx[r]*lds + c] : 0   (r < R, c < Cn), out contiguous
__global__ __launch_bounds__(NT) void gather_rows_kernel(int64_t R, int64_t Cn,
                                                         const int64_t* idx, const double* src,
                                                         int64_t lds, double* out) {
  for (int64_t r = blockIdx.x; r < R; r += gridDim.x) {
    const int64_t k = idx[r];
    double* o = out + r * Cn;
    for (int64_t c = threadIdx.x; c < Cn; c += NT) o[c] = (k >= 0) ? src[k * lds + c] : 0.0;
  }
}

// out[r][a][b] = C[pr[r]*ldc + c1 + a] * C[qr[r]*ldc + c2 + b]  (r < m, a < n1, b < n2): the
// orbital-pair images of the rows of a basis-change matrix, one row per workgroup pass
__global__ __launch_bounds__(NT) void pair_outer_kernel(int64_t m, const int64_t* pr,
                                                        const int64_t* qr, const double* C,
                                                        int64_t ldc, int64_t c1, int64_t n1,
                                                        int64_t c2, int64_t n2, double* out) {
  const int64_t per = n1 * n2;
  for (int64_t r = blockIdx.x; r < m; r += gridDim.x) {
    const double* x = C + pr[r] * ldc + c1;
    const double* y = C + qr[r] * ldc + c2;
    double* o = out + r * per;
    for (int64_t a = 0; a < n1; ++a) {
      const double xa = x[a];
      for (int64_t b = threadIdx.x; b < n2; b += NT) o[a * n2 + b] = xa * y[b];
    }
  }
}

// out[idx[r]*ldo + c] += alpha * src[r*Cn + c]   (rows with idx < 0 skipped; idx unique)
__global__ __launch_bounds__(NT) void scatter_add_rows_kernel(int64_t R, int64_t Cn,
                                                              const int64_t* idx,
                                                              const double* src, double alpha,
                                                              double* out, int64_t ldo) {
  for (int64_t r = blockIdx.x; r < R; r += gridDim.x) {
    const int64_t k = idx[r];
    if (k < 0) continue;
    const double* x = src + r * Cn;
    double* o = out + k * ldo;
    for (int64_t c = threadIdx.x; c < Cn; c += NT) o[c] += alpha * x[c];
  }
}

// out[r*ldo + c] = c < n1 ? A[r*lda + c] : B[r*ldb + c - n1]   (r < R, c < n1 + n2)
__global__ __launch_bounds__(NT) void concat_rows_kernel(int64_t R, int64_t n1, const double* A,
                                                         int64_t lda, int64_t n2, const double* B,
                                                         int64_t ldb, double* out, int64_t ldo) {
  const int64_t n = n1 + n2;
  for (int64_t r = blockIdx.x; r < R; r += gridDim.x)
    for (int64_t c = threadIdx.x; c < n; c += NT)
      out[r * ldo + c] = c < n1 ? A[r * lda + c] : B[r * ldb + c - n1];
}

// t1n = m*t1 + (1-m)*(t1 + H1/(-(Xpp[a,a] + Xhh[i,i])))
__global__ void update_t1_kernel(int64_t p, int64_t h, const double* t1, const double* H1,
                                 const double* Xpp, const double* Xhh, double mixing,
                                 double* t1n) {
  int64_t total = p * h;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t a = t / h, i = t - a * h;
    double denom = -(Xpp[a * p + a] + Xhh[i * h + i]);
    double nt = t1[t] + H1[t] / denom;
    t1n[t] = mixing * t1[t] + (1.0 - mixing) * nt;
  }
}

struct AssembleArgs {
  int64_t p, h;
  FastDiv fp, fh;                       // divisors p, h
  const void *S0, *Sab, *Sij, *Spp;     // S0, Sab, Sij [a][b][i][j]; Spp [(a,i)][(b,j)]; any null
                                        // (all four double, or all four float: the kernels'
                                        // template type T)
  const double *Xpp, *Xhh;              // dressed diagonals (full p x p, h x h matrices)
  const double *dpp, *dhh;              // optional: (dpp[a] + dpp[b] + dhh[i] + dhh[j]) t2[a,b,i,j]
                                        // is added to H2 in FP64 (the diagonal of a dressing
                                        // product that ran in low precision without it)
  const double *t2, *vpphh, *t1n;       // old t2, v_pphh, new t1 (for the fused energy)
  double mixing;
  double* t2n;                          // new t2 (may alias nothing)
  double* part;                         // per-block partials: [3][gridDim]: E2, |H2|^2, spare
  double* H2out;                        // optional: raw residual H2 (for adjoint / diagnostics)
  uint16_t* t2n16;                      // optional: BF16 image of t2n (same layout, RNE via FP32)
  int split16;                          // t2n16 holds the hi/lo split instead: row a of t2n as
                                        // [hi(p*h*h) | lo(p*h*h)] (row pitch 2*p*h*h);
                                        // 1: BF16, 2: FP16 of t2n * img_scale
  double img_scale;
  int addv;                             // H2 += v_pphh (when S0 does not hold it)
  int s0pairs;                          // S0 holds the hole pairs i < j only: row (a,b) of
                                        // s0ld >= h(h-1)/2 slots, S0[a,b,i,j] = +-S0[(a,b)][q]
  int64_t s0ld;
  int sabpairs;                         // Sab over the hole pairs too (rows (a,b) of s0ld slots)
  int imgpairs;                         // the split image of t2n over the hole pairs only: row a
                                        // as [hi((b,q)) | lo((b,q))], b of s0ld slots each
};

// S0[a,b,i,j] from S0's hole-pair rows (row = a*p + b): the antisymmetric expansion.
template <typename T>
__device__ inline double s0_pair(const T* S0, int64_t row, int64_t ld, int h, int i, int j) {
  if (i < j) return (double)S0[row * ld + (i * h - i * (i + 1) / 2 + j - i - 1)];
  if (i > j) return -(double)S0[row * ld + (j * h - j * (j + 1) / 2 + i - j - 1)];
  return 0.0;
}

// The split image over the hole pairs: t2n[a,b,i,j] (i < j, pair slot q) at row a of
// [hi(p * s0ld) | lo(p * s0ld)].
__device__ inline void write_image_pair(const AssembleArgs& g, int64_t a, int64_t b, int64_t q,
                                        double tn) {
  const int64_t half = g.p * g.s0ld, o = a * 2 * half + b * g.s0ld + q;
  uint16_t hi, lo;
  split16(tn * g.img_scale, g.split16 == 1, hi, lo);
  g.t2n16[o] = hi;
  g.t2n16[o + half] = lo;
}

__device__ inline void write_image(const AssembleArgs& g, int64_t t, int64_t a, double tn) {
  if (!g.split16) {
    g.t2n16[t] = to_bf16_bits_2r(tn);
    return;
  }
  const int64_t row = g.p * g.h * g.h, o = t + a * row;
  uint16_t hi, lo;
  split16(tn * g.img_scale, g.split16 == 1, hi, lo);
  g.t2n16[o] = hi;
  g.t2n16[o + row] = lo;
}

template <typename T>   // accumulator elements: double, or float (the split form's)
__global__ __launch_bounds__(NT) void assemble_t2_kernel(AssembleArgs g) {
  const T* const S0t = static_cast<const T*>(g.S0);
  const T* const Sabt = static_cast<const T*>(g.Sab);
  const T* const Sijt = static_cast<const T*>(g.Sij);
  const T* const Sppt = static_cast<const T*>(g.Spp);
  auto S0p = [&](int64_t o) { return (double)S0t[o]; };
  auto Sabp = [&](int64_t o) { return (double)Sabt[o]; };
  auto Sijp = [&](int64_t o) { return (double)Sijt[o]; };
  auto Sppp = [&](int64_t o) { return (double)Sppt[o]; };
  const int64_t p = g.p, h = g.h, hh = h * h, total = p * p * hh;
  double e_acc = 0.0, r_acc = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = t % h, r = t / h, i = r % h;
    r /= h;
    int64_t b = r % p, a = r / p;
    double H = !g.S0 ? 0.0 : g.s0pairs ? s0_pair(S0t, a * p + b, g.s0ld, (int)h, (int)i, (int)j)
                                        : S0p(t);
    if (g.Sab && g.sabpairs)
      H += s0_pair(Sabt, a * p + b, g.s0ld, (int)h, (int)i, (int)j) -
           s0_pair(Sabt, b * p + a, g.s0ld, (int)h, (int)i, (int)j);
    else if (g.Sab) H += Sabp(t) - Sabp(((b * p + a) * h + i) * h + j);
    const int64_t ph = p * h;
    if (g.Sij) H += Sijp(t) - Sijp(((a * p + b) * h + j) * h + i);
    if (g.Spp) {
      H += Sppp((a * h + i) * ph + b * h + j) - Sppp((b * h + i) * ph + a * h + j)
         - Sppp((a * h + j) * ph + b * h + i) + Sppp((b * h + j) * ph + a * h + i);
    }
    double denom = -(g.Xpp[a * p + a] + g.Xpp[b * p + b] + g.Xhh[i * h + i] + g.Xhh[j * h + j]);
    double told = g.t2[t];
    if (g.addv) H += g.vpphh[t];
    if (g.dpp) H += (g.dpp[a] + g.dpp[b]) * told;
    if (g.dhh) H += (g.dhh[i] + g.dhh[j]) * told;
    double nt = told + H / denom;
    double tn = g.mixing * told + (1.0 - g.mixing) * nt;
    if (g.t2n) g.t2n[t] = tn;
    if (g.t2n16 && g.imgpairs) {
      if (i < j) write_image_pair(g, a, b, i * h - i * (i + 1) / 2 + j - i - 1, tn);
    } else if (g.t2n16) {
      write_image(g, t, a, tn);
    }
    if (g.H2out) g.H2out[t] = H;
    if (g.vpphh) {
      double v = g.vpphh[t];
      e_acc += 0.25 * v * tn + 0.5 * v * g.t1n[a * h + i] * g.t1n[b * h + j];
    }
    r_acc += H * H;
  }
  double e = block_sum(e_acc);
  double rr = block_sum(r_acc);
  if (threadIdx.x == 0) {
    g.part[blockIdx.x] = e;
    g.part[gridDim.x + blockIdx.x] = rr;
  }
}

// The same assembly with one workgroup per pair {a, b} (a <= b), for h <= AH.  The h x h blocks
// Sij[a,b,:,:], Sij[b,a,:,:] (contiguous) and Spp[(a,i),(b,j)], Spp[(b,i),(a,j)] (h rows of h
// contiguous elements) are staged in LDS, where the transposed (i <-> j) reads of the
// antisymmetrisers are served; S0, Sab, t2, v_pphh, t2n are contiguous h x h blocks at (a,b) and
// (b,a).  Every tensor is therefore read once with row-contiguous accesses (the flat kernel
// above reads two of the four Spp terms at a stride of p*h elements).
// H(b,a) uses the same staged blocks: its Spp part is minus that of H(a,b).
// One thread per element (i, j) of the block (blockDim = h*h rounded up to wavefronts): all
// twelve loads of a pair are issued together, and the next pair's loads are in flight while
// this pair is finished.  The staged blocks alternate between two LDS buffers, so one barrier
// per pair separates a buffer's stores from its transposed reads: before a buffer is written
// again every thread has passed the barrier of the pair in between, hence finished its reads.
constexpr int AH = 32;
constexpr int ALD = AH + 1;              // LDS row pitch of a staged block

template <int NW>   // block-wide sum over NW wavefronts; result valid in thread 0
__device__ inline double block_sum_w(double v) {
  __shared__ double part[NW];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += part[i];
  __syncthreads();
  return s;
}

template <typename T>   // accumulator elements: double, or float (the split form's)
__global__ __launch_bounds__(AH * AH) void assemble_pair_kernel(AssembleArgs g) {
  const T* const S0t = static_cast<const T*>(g.S0);
  const T* const Sabt = static_cast<const T*>(g.Sab);
  const T* const Sijt = static_cast<const T*>(g.Sij);
  const T* const Sppt = static_cast<const T*>(g.Spp);
  __shared__ T Si[2][2][AH * ALD], Sp[2][2][AH * ALD];     // [buffer][(a,b) / (b,a)][i][j]
  // hole-pair rows of S0 and Sab ((a,b) and (b,a) each), read coalesced (slot e by thread e)
  // and staged here; each thread then takes its slot with its sign (FP32 accumulators only)
  constexpr int PZ = sizeof(T) == 4 ? AH * (AH - 1) / 2 : 1;
  __shared__ T Pz[2][4][PZ];                                // [buffer][S0 ab, ba, Sab ab, ba]
  const int64_t p = g.p, ph = p * g.h, pp = p * p;
  const int h = (int)g.h, hh = h * h, npair = h * (h - 1) / 2;
  const int e = threadIdx.x;
  const bool act = e < hh;
  const int i = act ? fdiv(e, g.fh) : 0, j = act ? e - i * h : 0;
  // this thread's slot and sign in the hole-pair rows (fixed over all pairs a, b)
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const int s0q = i != j ? lo * h - lo * (lo + 1) / 2 + hi - lo - 1 : 0;
  const T s0s = i < j ? T(1) : i > j ? T(-1) : T(0);
  const bool pz = sizeof(T) == 4 && (g.s0pairs || g.sabpairs);
  const bool slot = e < npair;                              // (slots < npair imply act)
  double e_acc = 0.0, r_acc = 0.0;
  auto finish = [&](int64_t a, int64_t b, int64_t t, double H, double told, double v) {
    double denom = -(g.Xpp[a * p + a] + g.Xpp[b * p + b] + g.Xhh[i * h + i] + g.Xhh[j * h + j]);
    if (g.addv) H += v;
    if (g.dpp) H += (g.dpp[a] + g.dpp[b]) * told;
    if (g.dhh) H += (g.dhh[i] + g.dhh[j]) * told;
    double tn = g.mixing * told + (1.0 - g.mixing) * (told + H / denom);
    if (g.t2n) g.t2n[t] = tn;
    if (g.t2n16 && g.imgpairs) {
      if (i < j) write_image_pair(g, a, b, s0q, tn);
    } else if (g.t2n16) {
      write_image(g, t, a, tn);
    }
    if (g.H2out) g.H2out[t] = H;
    if (g.vpphh) e_acc += 0.25 * v * tn + 0.5 * v * g.t1n[a * h + i] * g.t1n[b * h + j];
    r_acc += H * H;
  };
  // the pairs of this workgroup: q = blockIdx.x + k * gridDim.x with a <= b
  auto next = [&](int64_t q) {
    for (; q < pp; q += gridDim.x) {
      const int64_t a = fdiv((int)q, g.fp);
      if (a <= q - a * p) break;
    }
    return q;
  };
  struct Pair { T sa, sb, pa, pb, s0a, s0b, xa, xb; double ta, tb, va, vb; };
  auto load = [&](int64_t q, Pair& r) {
    const int64_t a = fdiv((int)q, g.fp), b = q - a * p;
    const int64_t tab = (a * p + b) * hh + e, tba = (b * p + a) * hh + e;
    const int64_t oab = (a * g.h + i) * ph + b * g.h + j, oba = (b * g.h + i) * ph + a * g.h + j;
    r.sa = g.Sij ? Sijt[tab] : T(0); r.sb = g.Sij ? Sijt[tba] : T(0);
    r.pa = g.Spp ? Sppt[oab] : T(0); r.pb = g.Spp ? Sppt[oba] : T(0);
    // pair rows: slot e (staged through LDS; pz) or this thread's slot with its sign
    const int64_t qab = (a * p + b) * g.s0ld + (pz ? e : s0q);
    const int64_t qba = (b * p + a) * g.s0ld + (pz ? e : s0q);
    const T sg = pz ? T(1) : s0s;
    const bool take = pz ? slot : true;
    if (g.s0pairs) {
      r.s0a = take ? sg * S0t[qab] : T(0); r.s0b = take ? sg * S0t[qba] : T(0);
    } else {
      r.s0a = g.S0 ? S0t[tab] : T(0); r.s0b = g.S0 ? S0t[tba] : T(0);
    }
    if (g.sabpairs) {
      r.xa = take ? sg * Sabt[qab] : T(0); r.xb = take ? sg * Sabt[qba] : T(0);
    } else {
      r.xa = g.Sab ? Sabt[tab] : T(0); r.xb = g.Sab ? Sabt[tba] : T(0);
    }
    r.ta = g.t2[tab]; r.tb = g.t2[tba];
    r.va = g.vpphh ? g.vpphh[tab] : 0.0; r.vb = g.vpphh ? g.vpphh[tba] : 0.0;
  };
  int64_t q = next(blockIdx.x);
  Pair cur;
  if (act && q < pp) load(q, cur);
  int buf = 0;
  while (q < pp) {
    const int64_t a = fdiv((int)q, g.fp), b = q - a * p;
    T* const SiA = Si[buf][0]; T* const SiB = Si[buf][1];
    T* const SpA = Sp[buf][0]; T* const SpB = Sp[buf][1];
    if (act) {
      SiA[i * ALD + j] = cur.sa; SiB[i * ALD + j] = cur.sb;
      SpA[i * ALD + j] = cur.pa; SpB[i * ALD + j] = cur.pb;
    }
    if (pz && slot) {
      Pz[buf][0][e] = cur.s0a; Pz[buf][1][e] = cur.s0b;
      Pz[buf][2][e] = cur.xa; Pz[buf][3][e] = cur.xb;
    }
    const int64_t qn = next(q + gridDim.x);
    Pair nxt;
    if (act && qn < pp) load(qn, nxt);
    // this buffer's stores complete, then the barrier (a compiler barrier too: the LDS reads
    // below stay after it; the next pair's global loads stay in flight)
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (act) {
      const int64_t tab = (a * p + b) * hh + e, tba = (b * p + a) * hh + e;
      // a staged row's slot (0 on the diagonal i = j, which has none)
      auto staged = [&](int k) { return i != j ? s0s * Pz[buf][k][s0q] : T(0); };
      double Hab = (double)(pz && g.s0pairs ? staged(0) : cur.s0a);
      double Hba = (double)(pz && g.s0pairs ? staged(1) : cur.s0b);
      const double x = pz && g.sabpairs ? staged(2) : cur.xa;
      const double y = pz && g.sabpairs ? staged(3) : cur.xb;
      Hab += x - y;
      Hba += y - x;
      Hab += (double)SiA[i * ALD + j] - (double)SiA[j * ALD + i];
      Hba += (double)SiB[i * ALD + j] - (double)SiB[j * ALD + i];
      const double sp = (double)SpA[i * ALD + j] - (double)SpB[i * ALD + j] -
                        (double)SpA[j * ALD + i] + (double)SpB[j * ALD + i];
      finish(a, b, tab, Hab + sp, cur.ta, cur.va);
      if (a != b) finish(b, a, tba, Hba - sp, cur.tb, cur.vb);
    }
    cur = nxt;
    q = qn;
    buf ^= 1;
  }
  double es = block_sum_w<AH * AH / 64>(e_acc);
  double rs = block_sum_w<AH * AH / 64>(r_acc);
  if (threadIdx.x == 0) {
    g.part[blockIdx.x] = es;
    g.part[gridDim.x + blockIdx.x] = rs;
  }
}

// Transpose of the assembly's antisymmetrisers: for the adjoint L of H2,
//   gSab[a,b,i,j] = L[a,b,i,j] - L[b,a,i,j]                         (layout [a][b][i][j])
//   gSij[a,i,b,j] = L[a,b,i,j] - L[a,b,j,i]                         (layout [a][i][b][j])
//   gSpp[a,i,b,j] = L[a,b,i,j] - L[b,a,i,j] - L[a,b,j,i] + L[b,a,j,i] (layout [a][i][b][j])
// Threads walk the [a][i][b][j] order; j is innermost in every output and in two of the reads.
template <bool SMALL>   // SMALL: p*p*h*h < 2^31, coordinates by multiply-shift division
__global__ __launch_bounds__(NT) void assemble_adj_kernel(int64_t p, int64_t h, FastDiv fp,
                                                          FastDiv fh, const double* L,
                                                          double* gSab, double* gSij,
                                                          double* gSpp) {
  const int64_t total = p * p * h * h;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t a, i, b, j;
    if (SMALL) {
      const int r = fdiv((int)t, fh), r2 = fdiv(r, fp), a_ = fdiv(r2, fh);
      j = t - (int64_t)r * h; b = r - (int64_t)r2 * p; a = a_; i = r2 - (int64_t)a_ * h;
    } else {
      j = t % h; int64_t r = t / h; b = r % p; r /= p; i = r % h; a = r / h;
    }
    double v0 = L[((a * p + b) * h + i) * h + j];
    double v1 = L[((b * p + a) * h + i) * h + j];
    double v2 = L[((a * p + b) * h + j) * h + i];
    double v3 = L[((b * p + a) * h + j) * h + i];
    if (gSab) gSab[((a * p + b) * h + i) * h + j] = v0 - v1;
    if (gSij) gSij[t] = v0 - v2;
    if (gSpp) gSpp[t] = v0 - v1 - v2 + v3;
  }
}

// E2 partials for E = f.t1 + 1/4 v.t2 + 1/2 v t1 t1 (the f.t1 part is added by finalize).
template <bool SMALL>   // SMALL: p*p*h*h < 2^31, coordinates by multiply-shift division
__global__ __launch_bounds__(NT) void energy_part_kernel(int64_t p, int64_t h, FastDiv fp,
                                                         FastDiv fh, const double* v,
                                                         const double* t2, const double* t1,
                                                         double* part) {
  const int64_t total = p * p * h * h;
  double acc = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t a, i, b, j;
    if (SMALL) {
      const int r = fdiv((int)t, fh), r2 = fdiv(r, fh), a_ = fdiv(r2, fp);
      j = t - (int64_t)r * h; i = r - (int64_t)r2 * h; a = a_; b = r2 - (int64_t)a_ * p;
    } else {
      j = t % h; int64_t r = t / h; i = r % h; r /= h; b = r % p; a = r / p;
    }
    double x = v[t];
    acc += 0.25 * x * t2[t] + 0.5 * x * t1[a * h + i] * t1[b * h + j];
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[0] = sum(part[0:n]) + sum_k x[k]*y[k] over nx elements (fixed order, one workgroup)
__global__ __launch_bounds__(NT) void finalize_kernel(int64_t n, const double* part, int64_t nx,
                                                      const double* x, const double* y,
                                                      double* out) {
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < nx; k += NT) acc += x[k] * y[k];
  double a = block_sum(acc);
  double b = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += NT) b += part[k];
  b = block_sum(b);
  if (threadIdx.x == 0) out[0] = a + b;
}

__global__ __launch_bounds__(NT) void dot_part_kernel(int64_t n, const double* x, const double* y,
                                                      double* part) {
  double acc = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    acc += x[t] * y[t];
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Gram matrix of K vectors in one pass: every element of every vector is read once and feeds
// all K(K+1)/2 products.  part[q * gridDim.x + blockIdx.x] is this workgroup's share of
// x_a . x_b, with q running over a <= b row by row (a = 0: b = 0..K-1, then a = 1, ...).
constexpr int GRAM_MAX = 12;
constexpr int LINCOMB_MAX = 16;

// T: the stored vectors' type (double, or float for a history kept in FP32); the products and
// sums are FP64 either way.
template <typename T>
struct GramArgs {
  const T* x[GRAM_MAX];
};

template <int K, typename T>
__global__ __launch_bounds__(NT) void gram_part_kernel(int64_t n, GramArgs<T> g, double* part) {
  constexpr int NP = K * (K + 1) / 2;
  double acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double v[K];
#pragma unroll
    for (int a = 0; a < K; ++a) v[a] = (double)g.x[a][t];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int b = a; b < K; ++b) acc[a * K - a * (a - 1) / 2 + (b - a)] += v[a] * v[b];
  }
  __shared__ double red[NT / 64][NP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const double s = wave_sum(acc[q]);
    if (lane == 0) red[w][q] = s;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < NP; q += NT) {
    double s = 0.0;
    for (int i = 0; i < NT / 64; ++i) s += red[i][q];
    part[(int64_t)q * gridDim.x + blockIdx.x] = s;
  }
}

// out[q] = sum(part[q*nparts : (q+1)*nparts]) in a fixed order; one workgroup per q.
__global__ __launch_bounds__(NT) void gram_finalize_kernel(int64_t nparts, const double* part,
                                                           double* out) {
  const double* p = part + (int64_t)blockIdx.x * nparts;
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < nparts; k += NT) acc += p[k];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

template <typename T>
struct LincombArgs {
  const T* x[LINCOMB_MAX];
  double c[LINCOMB_MAX];
};

// y = sum_k c[k] x_k, one pass (reads K vectors, writes one FP64 vector)
template <int K, typename T>
__global__ __launch_bounds__(NT) void lincomb_kernel(int64_t n, LincombArgs<T> g, double* y) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double s = g.c[0] * (double)g.x[0][t];
#pragma unroll
    for (int k = 1; k < K; ++k) s += g.c[k] * (double)g.x[k][t];
    y[t] = s;
  }
}

// hcopy = x, c = x - y in one pass (a DIIS step's history entry and error vector), stored as T
template <typename T>
__global__ void copy_sub_kernel(int64_t n, const double* x, const double* y, T* hcopy, T* c) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[t];
    hcopy[t] = (T)v;
    c[t] = (T)(v - y[t]);
  }
}

// y = a*x + b*y
__global__ void axpby_kernel(int64_t n, double a, const double* x, double b, double* y) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double v = a * x[t];
    y[t] = (b != 0.0) ? v + b * y[t] : v;
  }
}

// c = x - y (difference vectors for DIIS errors)
__global__ void sub_kernel(int64_t n, const double* x, const double* y, double* c) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    c[t] = x[t] - y[t];
}

// Delta image of a split operand against an anchor: for rows r < rows and columns k < K of the
// split images X, Xa (row r at r*ldx: [hi(K) | lo(K)], hi at column k, lo at K + k)
//   D[r*ldd + off + k] = FP16((hi + lo) - (hi_a + lo_a))   (each pair summed in FP32),
// zero for K <= k < kp (a pad to the product's K step); off = kp when D also takes X's hi part
// (D[r*ldd + k] = hi, zero from K to kp: the [t2_hi | t2 - t2_A] operand of a K-concatenated
// product), else 0.  Eight columns per thread (16-byte loads and stores; K, kp, ldx, ldd
// multiples of 8, the host checks).
template <bool HI>
__global__ __launch_bounds__(NT) void split_delta16_kernel(int64_t rows, int64_t K, int64_t ldx,
                                                           const uint16_t* X,
                                                           const uint16_t* Xa, uint16_t* D,
                                                           int64_t ldd, int64_t kp) {
  const int64_t per_row = kp / 8, total = rows * per_row;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / per_row, k = (t - r * per_row) * 8;
    uint4 out = make_uint4(0, 0, 0, 0), hic = make_uint4(0, 0, 0, 0);
    if (k < K) {
      const int64_t o = r * ldx + k;
      const uint4 hi = *(const uint4*)(X + o), lo = *(const uint4*)(X + o + K);
      const uint4 ha = *(const uint4*)(Xa + o), la = *(const uint4*)(Xa + o + K);
      const v8h h0 = __builtin_bit_cast(v8h, hi), l0 = __builtin_bit_cast(v8h, lo);
      const v8h h1 = __builtin_bit_cast(v8h, ha), l1 = __builtin_bit_cast(v8h, la);
      v8h d;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float a = (float)h0[e] + (float)l0[e], b = (float)h1[e] + (float)l1[e];
        d[e] = (_Float16)(a - b);
      }
      out = __builtin_bit_cast(uint4, d);
      hic = hi;
    }
    if (HI) *(uint4*)(D + r * ldd + k) = hic;
    *(uint4*)(D + r * ldd + (HI ? kp : 0) + k) = out;
  }
}

inline int grid_for(int64_t n, int cap = 16384) {
  int64_t g = nl_cdiv(n, NT);
  if (g < 1) g = 1;
  return (int)(g < cap ? g : cap);
}

// FP8 (OCP E4M3) rows with a per-row power-of-two scale: x s with |x s| in (128, 256] for the
// row maximum (below E4M3's 448), rounded to nearest even by the hardware conversion; zero from
// K to kp.  One workgroup per row; the row's values wait in LDS (K floats) between the amax
// pass and the conversion.  F: (X, Xa) split 16-bit images [hi(K) | lo(K)] at pitch ldx, value
// (hi + lo) - (hi_a + lo_a) in FP32; else X is FP64 at pitch ldx.
__device__ inline unsigned pack_fp8x4(float a, float b, float c, float d) {
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
  return (unsigned)w;
}

template <bool F>
__global__ __launch_bounds__(NT) void rows_fp8_kernel(int64_t K, int64_t ldx, const void* X,
                                                      const void* Xa, uint8_t* D, int64_t ldd,
                                                      int64_t kp, float* iscale) {
  extern __shared__ float vals[];
  __shared__ float red[NT / 64];
  const int64_t r = blockIdx.x;
  float m = 0.f;
  for (int64_t k = threadIdx.x * 8; k < K; k += NT * 8) {
    float v[8];
    if constexpr (F) {
      const uint16_t* x = (const uint16_t*)X + r * ldx + k;
      const uint16_t* xa = (const uint16_t*)Xa + r * ldx + k;
      const v8h h0 = __builtin_bit_cast(v8h, *(const uint4*)x);
      const v8h l0 = __builtin_bit_cast(v8h, *(const uint4*)(x + K));
      const v8h h1 = __builtin_bit_cast(v8h, *(const uint4*)xa);
      const v8h l1 = __builtin_bit_cast(v8h, *(const uint4*)(xa + K));
#pragma unroll
      for (int e = 0; e < 8; ++e)
        v[e] = ((float)h0[e] + (float)l0[e]) - ((float)h1[e] + (float)l1[e]);
    } else {
      const double* x = (const double*)X + r * ldx + k;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = (float)x[e];
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      vals[k + e] = v[e];
      m = fmaxf(m, fabsf(v[e]));
    }
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  const float amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  const float sc = row_scale(amax, 256.f);
  if (threadIdx.x == 0) iscale[r] = row_iscale(amax, 256.f);
  for (int64_t k = threadIdx.x * 8; k < kp; k += NT * 8) {
    uint2 o = make_uint2(0, 0);
    if (k < K) {
      const float* v = vals + k;
      o.x = pack_fp8x4(v[0] * sc, v[1] * sc, v[2] * sc, v[3] * sc);
      o.y = pack_fp8x4(v[4] * sc, v[5] * sc, v[6] * sc, v[7] * sc);
    }
    *(uint2*)(D + r * ldd + k) = o;
  }
}


}  // namespace

extern "C" {

int nl_permute(int rank, const int64_t* out_shape, const int64_t* in_strides, double alpha,
               const double* in, double beta, double* out, void* stream) {
  if (rank < 1 || rank > 6 || !out_shape || !in_strides || !in || !out) return NL_EARG;
  PermArgs a;
  a.rank = rank;
  a.total = 1;
  for (int d = 0; d < rank; ++d) {
    if (out_shape[d] < 0) return NL_EARG;
    a.shape[d] = out_shape[d];
    a.istr[d] = in_strides[d];
    a.fd[d] = nl_fastdiv(out_shape[d] < INT32_MAX ? out_shape[d] : 1);
    a.total *= out_shape[d];
  }
  if (a.total == 0) return NL_OK;
  a.alpha = alpha; a.beta = beta; a.in = in; a.out = out;
  if (a.total < INT32_MAX)
    hipLaunchKernelGGL((permute_kernel<true>), dim3(grid_for(a.total)), dim3(NT), 0,
                       (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL((permute_kernel<false>), dim3(grid_for(a.total)), dim3(NT), 0,
                       (hipStream_t)stream, a);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_convert_rows(int dtype, int64_t R1, int64_t R2, int64_t C1, int64_t C2,
                    const int64_t* strides4, const double* in, void* out, int64_t ldo,
                    float* iscale, float target, void* stream) {
  const bool scaled = target != 0.f;
  if ((dtype != 0 && dtype != 1) || !strides4 || !in || !out || (scaled && !iscale))
    return NL_EARG;
  if (R1 < 0 || R2 < 1 || C1 < 0 || C2 < 1 || ldo < C1 * C2 || (ldo % 8)) return NL_EARG;
  if (R1 * R2 == 0) return NL_OK;
  if (R1 * R2 > INT32_MAX || ldo > INT32_MAX - CAST_CHUNK) return NL_EARG;
  ConvArgs a;
  a.R2 = R2; a.C = C1 * C2; a.C2 = C2; a.ldo = ldo; a.cend = ldo;
  a.fr2 = nl_fastdiv(R2); a.fc2 = nl_fastdiv(C2);
  for (int d = 0; d < 4; ++d) a.s[d] = strides4[d];
  a.in = in; a.out = (uint16_t*)out; a.iscale = iscale; a.target = target; a.bf16 = dtype == 0;
  hipStream_t st = (hipStream_t)stream;
  const int64_t R = R1 * R2;
  if (a.s[1] == 1 && R2 >= 8 && a.s[3] != 1 && nl_cdiv(ldo, 64) <= 65535) {
    // rows are the unit-stride direction of the source: transpose through LDS tiles
    if (!scaled) {
      dim3 g2((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(ldo, 64));
      hipLaunchKernelGGL(convert_tiles_kernel, g2, dim3(NT), 0, st, a, R, (const unsigned*)nullptr);
      NL_LAUNCH_CHECK();
      return NL_OK;
    }
    unsigned* bits = (unsigned*)iscale;   // holds amax bits until the last pass
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(R)), dim3(NT), 0, st, R, bits);
    NL_LAUNCH_CHECK();
    dim3 g1((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(a.C, 64));
    hipLaunchKernelGGL(amax_tiles_kernel, g1, dim3(NT), 0, st, a, R, bits);
    NL_LAUNCH_CHECK();
    dim3 g2((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(ldo, 64));
    hipLaunchKernelGGL(convert_tiles_kernel, g2, dim3(NT), 0, st, a, R, (const unsigned*)bits);
    NL_LAUNCH_CHECK();
    hipLaunchKernelGGL(iscale_kernel, dim3(grid_for(R)), dim3(NT), 0, st, R, target, iscale);
    NL_LAUNCH_CHECK();
    return NL_OK;
  }
  if (!scaled) {
    if (nl_cdiv(ldo, CAST_CHUNK) > 65535) return NL_EARG;
    dim3 g((unsigned)R, (unsigned)nl_cdiv(ldo, CAST_CHUNK));
    hipLaunchKernelGGL(cast_rows_kernel, g, dim3(NT), 0, st, a);
    NL_LAUNCH_CHECK();
    return NL_OK;
  }
  hipLaunchKernelGGL(convert_rows_kernel, dim3((unsigned)R), dim3(NT), 0, st, a);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

static int convert_split_impl(int dtype, int64_t R1, int64_t R2, int64_t C1, int64_t C2,
                              const int64_t* strides4, const void* in, const void* sub,
                              void* out, int64_t ldo, int64_t kpad, int nblk, int lomask,
                              float* iscale, float target, double scale, void* stream,
                              const float* given = nullptr) {
  // dtype: 16-bit type (bits 0-7) | NL_D_F32 << 8 for an FP32 source (else FP64)
  const bool in32 = (dtype >> 8) == NL_D_F32;
  if (dtype >> 8 && !in32) return NL_EARG;
  dtype &= 0xff;
  const bool scaled = target != 0.f && !given;
  if ((dtype != 0 && dtype != 1) || !strides4 || !in || !out || (scaled && !iscale))
    return NL_EARG;
  if (nblk < 1 || nblk > 4 || lomask < 0 || lomask >= (1 << nblk)) return NL_EARG;
  if (R1 < 0 || R2 < 1 || C1 < 0 || C2 < 1 || kpad < C1 * C2 || ldo < nblk * kpad) return NL_EARG;
  if (R1 * R2 == 0) return NL_OK;
  if (R1 * R2 > INT32_MAX || kpad > INT32_MAX || C1 * C2 > INT32_MAX) return NL_EARG;
  SplitArgs a;
  a.R2 = R2; a.C = C1 * C2; a.C2 = C2; a.ldo = ldo; a.kpad = kpad;
  a.fr2 = nl_fastdiv(R2); a.fc2 = nl_fastdiv(C2);
  for (int d = 0; d < 4; ++d) a.s[d] = strides4[d];
  a.in = AnyIn{in, in32}; a.sub = AnyIn{sub, in32};
  a.out = (uint16_t*)out; a.iscale = scaled ? iscale : nullptr; a.given = given;
  a.target = target;
  a.nblk = nblk; a.lomask = lomask; a.bf16 = dtype == 0; a.scale = scale;
  hipStream_t st = (hipStream_t)stream;
  const int64_t R = R1 * R2;
  if (a.s[1] == 1 && R2 >= 8 && a.s[3] != 1 && nl_cdiv(kpad, 64) <= 65535) {
    unsigned* bits = scaled ? (unsigned*)iscale : nullptr;
    if (scaled) {
      hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(R)), dim3(NT), 0, st, R, bits);
      NL_LAUNCH_CHECK();
      dim3 g1((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(a.C, 64));
      if (sub) hipLaunchKernelGGL(amax_split_kernel<true>, g1, dim3(NT), 0, st, a, R, bits);
      else hipLaunchKernelGGL(amax_split_kernel<false>, g1, dim3(NT), 0, st, a, R, bits);
      NL_LAUNCH_CHECK();
    }
    dim3 g2((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(kpad, 64));
    if (sub)
      hipLaunchKernelGGL(split_tiles_kernel<true>, g2, dim3(NT), 0, st, a, R,
                         (const unsigned*)bits);
    else
      hipLaunchKernelGGL(split_tiles_kernel<false>, g2, dim3(NT), 0, st, a, R,
                         (const unsigned*)bits);
    NL_LAUNCH_CHECK();
    if (scaled) {
      hipLaunchKernelGGL(bits_to_iscale_kernel, dim3(grid_for(R)), dim3(NT), 0, st, R, target,
                         iscale);
      NL_LAUNCH_CHECK();
    }
    return NL_OK;
  }
  if (sub) hipLaunchKernelGGL(split_rows_kernel<true>, dim3((unsigned)R), dim3(NT), 0, st, a);
  else hipLaunchKernelGGL(split_rows_kernel<false>, dim3((unsigned)R), dim3(NT), 0, st, a);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_convert_split(int dtype, int64_t R1, int64_t R2, int64_t C1, int64_t C2,
                     const int64_t* strides4, const void* in, void* out, int64_t ldo,
                     int64_t kpad, int nblk, int lomask, float* iscale, float target,
                     double scale, void* stream) {
  return convert_split_impl(dtype, R1, R2, C1, C2, strides4, in, nullptr, out, ldo, kpad, nblk,
                            lomask, iscale, target, scale, stream);
}

int nl_convert_split_delta(int dtype, int64_t R1, int64_t R2, int64_t C1, int64_t C2,
                           const int64_t* strides4, const void* in, const void* sub, void* out,
                           int64_t ldo, int64_t kpad, int nblk, int lomask, float* iscale,
                           const float* iscale_in, float target, double scale, void* stream) {
  if (!sub) return NL_EARG;
  return convert_split_impl(dtype, R1, R2, C1, C2, strides4, in, sub, out, ldo, kpad, nblk,
                            lomask, iscale, target, scale, stream, iscale_in);
}

int nl_gather_convert(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx, const double* src,
                      int64_t lds, void* out, int64_t ldk, void* stream) {
  if (nb < 0 || kb < 0 || ncol < 0 || ldk < kb || (ldk % 8) || lds < ncol) return NL_EARG;
  const int64_t R = nb * ncol;
  if (R == 0 || ldk == 0) return NL_OK;
  if (!idx || !src || !out || R >= INT32_MAX || nl_cdiv(ldk, 64) > 65535) return NL_EARG;
  dim3 g((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(ldk, 64));
  hipLaunchKernelGGL((gather_convert_kernel<double>), g, dim3(NT), 0, (hipStream_t)stream, R,
                     ncol, nl_fastdiv(ncol), kb, ldk, idx, src, lds, (uint16_t*)out, ldk);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

static int gather_rows16_impl(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx,
                              const void* src, int64_t lds, void* out, int64_t ldk, int64_t ldo,
                              void* stream) {
  if (nb < 0 || kb < 0 || ncol < 0 || ldk < kb || (ldk % 8) || lds < ncol || ldo < ldk)
    return NL_EARG;
  const int64_t R = nb * ncol;
  if (R == 0 || ldk == 0) return NL_OK;
  if (!idx || !src || !out || R >= INT32_MAX || nl_cdiv(ldk, 64) > 65535) return NL_EARG;
  if (ncol % 4 == 0 && lds % 4 == 0 && ldo % 4 == 0 && (((uintptr_t)src) & 7) == 0 &&
      (((uintptr_t)out) & 7) == 0 && nb <= 65535) {
    dim3 g((unsigned)nl_cdiv(ncol, 256), (unsigned)nl_cdiv(ldk, 64), (unsigned)nb);
    hipLaunchKernelGGL(gather16_kernel<1>, g, dim3(NT), 0, (hipStream_t)stream, ncol, kb, ldk,
                       idx, (const uint16_t*)src, lds, (uint16_t*)out, ldo, (int64_t)0);
    NL_LAUNCH_CHECK();
    return NL_OK;
  }
  dim3 g((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(ldk, 64));
  hipLaunchKernelGGL((gather_convert_kernel<uint16_t>), g, dim3(NT), 0, (hipStream_t)stream, R,
                     ncol, nl_fastdiv(ncol), kb, ldk, idx, (const uint16_t*)src, lds,
                     (uint16_t*)out, ldo);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_gather_rows16(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx, const void* src,
                     int64_t lds, void* out, int64_t ldk, void* stream) {
  return gather_rows16_impl(nb, kb, ncol, idx, src, lds, out, ldk, ldk, stream);
}

int nl_gather_rows16_into(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx,
                          const void* src, int64_t lds, void* out, int64_t ldk, int64_t ldo,
                          void* stream) {
  return gather_rows16_impl(nb, kb, ncol, idx, src, lds, out, ldk, ldo, stream);
}

int nl_gather_rows16_split(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx,
                           const void* src, int64_t lds, void* out, int64_t kpad, int64_t ldo,
                           void* stream) {
  if (nb < 0 || kb < 0 || ncol < 0 || kpad < kb || (kpad % 8) || lds < 2 * ncol ||
      ldo < 2 * kpad)
    return NL_EARG;
  if (nb * ncol == 0 || kpad == 0) return NL_OK;
  if (!idx || !src || !out || nb > 65535 || nb * ncol >= INT32_MAX ||
      nl_cdiv(kpad, 64) > 65535)
    return NL_EARG;
  if (ncol % 4 || lds % 4 || ldo % 4 || (((uintptr_t)src) & 7) || (((uintptr_t)out) & 7))
    return NL_EARG;
  dim3 g((unsigned)nl_cdiv(2 * ncol, 256), (unsigned)nl_cdiv(kpad, 64), (unsigned)nb);
  hipLaunchKernelGGL(gather16_kernel<2>, g, dim3(NT), 0, (hipStream_t)stream, ncol, kb, kpad,
                     idx, (const uint16_t*)src, lds, (uint16_t*)out, ldo, kpad);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_tau_split(int64_t p, int64_t h, const double* t1, const double* t2, double* tau,
                 void* out, int dtype, double scale, void* stream) {
  if (p < 0 || h < 0 || !t1 || !t2 || !out || (dtype != 0 && dtype != 1)) return NL_EARG;
  if (p * p * h * h == 0) return NL_OK;
  if (p * p > INT32_MAX || h * h > INT32_MAX || p * h > INT32_MAX || h > NT) return NL_EARG;
  hipLaunchKernelGGL(tau_split_kernel, dim3((unsigned)std::min<int64_t>(p * p, 65536)),
                     dim3(NT), 0, (hipStream_t)stream, (int)p, (int)h, nl_fastdiv(p),
                     nl_fastdiv(h), t1, t2, tau, (uint16_t*)out, dtype == 0, scale);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_tau_split_pairs(int64_t p, int64_t h, const double* t1, const double* t2, void* out,
                       int64_t Pp, int64_t grp, int dtype, double scale, void* stream) {
  const int64_t P = h * (h - 1) / 2;
  if (p < 0 || h < 0 || !t2 || !out || (dtype != 0 && dtype != 1) || Pp < P) return NL_EARG;
  if (grp != 1 && grp != p) return NL_EARG;
  if (p * p * Pp == 0) return NL_OK;
  if (p * p * Pp >= INT32_MAX || p * h > INT32_MAX) return NL_EARG;
  hipLaunchKernelGGL(tau_split_pairs_kernel, dim3(grid_for(p * p * Pp)), dim3(NT), 0,
                     (hipStream_t)stream, p * p, (int)p, (int)h, (int)P, (int)Pp, (int)grp,
                     nl_fastdiv(p), nl_fastdiv(Pp), nl_fastdiv(grp), t1, t2, (uint16_t*)out,
                     dtype == 0, scale);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_split_delta16(int64_t rows, int64_t K, int64_t ldx, const void* X, const void* Xa,
                     void* D, int64_t ldd, int64_t kp, int with_hi, void* stream) {
  if (rows < 0 || K < 0 || kp < K || ldx < 2 * K || ldd < (with_hi ? 2 : 1) * kp || (K % 8) ||
      (kp % 8) || (ldx % 8) || (ldd % 8) || (with_hi != 0 && with_hi != 1))
    return NL_EARG;
  if (rows == 0 || kp == 0) return NL_OK;
  if (!X || !Xa || !D || (((uintptr_t)X) & 15) || (((uintptr_t)Xa) & 15) ||
      (((uintptr_t)D) & 15))
    return NL_EARG;
  const dim3 grid(grid_for(rows * (kp / 8)));
  hipStream_t st = (hipStream_t)stream;
  if (with_hi)
    hipLaunchKernelGGL(split_delta16_kernel<true>, grid, dim3(NT), 0, st, rows, K, ldx,
                       (const uint16_t*)X, (const uint16_t*)Xa, (uint16_t*)D, ldd, kp);
  else
    hipLaunchKernelGGL(split_delta16_kernel<false>, grid, dim3(NT), 0, st, rows, K, ldx,
                       (const uint16_t*)X, (const uint16_t*)Xa, (uint16_t*)D, ldd, kp);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

constexpr int64_t kFp8MaxK = 16384;      // NL_FP8_MAXK: the row's FP32 values in 64 KB of LDS

static int rows_fp8(bool split, int64_t rows, int64_t K, int64_t ldx, const void* X,
                    const void* Xa, void* D, int64_t ldd, int64_t kp, float* iscale,
                    void* stream) {
  if (rows < 0 || K < 0 || kp < K || (K % 8) || (kp % 8) || (ldd % 8) || ldd < kp ||
      ldx < (split ? 2 : 1) * K || (split && (ldx % 8)) || K > kFp8MaxK)
    return NL_EARG;
  if (rows == 0 || kp == 0) return NL_OK;
  if (!X || !D || !iscale || (split && !Xa) || (((uintptr_t)D) & 7) ||
      (split && ((((uintptr_t)X) & 15) || (((uintptr_t)Xa) & 15))) ||
      (!split && (((uintptr_t)X) & 7)) || rows > INT32_MAX)
    return NL_EARG;
  const size_t lds = (size_t)nl_cdiv(K, NT * 8) * NT * 8 * sizeof(float);
  hipStream_t st = (hipStream_t)stream;
  if (split)
    hipLaunchKernelGGL(rows_fp8_kernel<true>, dim3((unsigned)rows), dim3(NT), lds, st, K, ldx, X,
                       Xa, (uint8_t*)D, ldd, kp, iscale);
  else
    hipLaunchKernelGGL(rows_fp8_kernel<false>, dim3((unsigned)rows), dim3(NT), lds, st, K, ldx,
                       X, nullptr, (uint8_t*)D, ldd, kp, iscale);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_split_delta8(int64_t rows, int64_t K, int64_t ldx, const void* X, const void* Xa,
                    void* D, int64_t ldd, int64_t kp, float* iscale, void* stream) {
  return rows_fp8(true, rows, K, ldx, X, Xa, D, ldd, kp, iscale, stream);
}

int nl_convert_rows8(int64_t rows, int64_t K, int64_t ldx, const double* X, void* D,
                     int64_t ldd, int64_t kp, float* iscale, void* stream) {
  return rows_fp8(false, rows, K, ldx, X, nullptr, D, ldd, kp, iscale, stream);
}

int nl_ring_from_pairs(int64_t p, int64_t h, int64_t Pp, const void* in, void* out,
                       void* stream) {
  const int64_t P = h * (h - 1) / 2;
  if (p < 0 || h < 0 || Pp < P || (Pp % 32) || !in || !out) return NL_EARG;
  if (p * h == 0) return NL_OK;
  if ((((uintptr_t)in) & 15) || p > INT32_MAX || h > 65535) return NL_EARG;
  // columns c per workgroup: 16, fewer when both parts' slots would pass 64 KB of LDS
  const int64_t ct = std::max<int64_t>(1, std::min<int64_t>(16, 65536 / (4 * std::max<int64_t>(Pp, 32))));
  if (4 * Pp * ct > 65536) return NL_EARG;                        // h above 181
  const int64_t nct = nl_cdiv(p, ct), blocks = p * nct;
  if (blocks > INT32_MAX) return NL_EARG;
  // 16-byte stores when every segment start (a,i,part,c0) is 8-aligned
  const bool vec = ((p * h) % 8 == 0) && ((ct * h) % 8 == 0) && !(((uintptr_t)out) & 15);
  const size_t lds = (size_t)(2 * ct * Pp * 2);
  hipStream_t st = (hipStream_t)stream;
  if (vec)
    hipLaunchKernelGGL(ring_from_pairs_kernel<true>, dim3((unsigned)blocks), dim3(NT), lds, st,
                       (int)p, (int)h, (int)Pp, (int)ct, (int)nct, (const uint16_t*)in,
                       (uint16_t*)out);
  else
    hipLaunchKernelGGL(ring_from_pairs_kernel<false>, dim3((unsigned)blocks), dim3(NT), lds, st,
                       (int)p, (int)h, (int)Pp, (int)ct, (int)nct, (const uint16_t*)in,
                       (uint16_t*)out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_hole_pairs(int64_t R, int64_t h, int64_t ldp, int64_t w, const double* in, double* out,
                  int expand, void* stream) {
  const int64_t P = h * (h - 1) / 2;
  if (R < 0 || h < 0 || w < 1 || ldp < P || !in || !out || (expand != 0 && expand != 1))
    return NL_EARG;
  if (R * P * w == 0 && !expand) return NL_OK;
  if (R * h * h * w == 0) return NL_OK;
  if (h > 65535) return NL_EARG;
  hipLaunchKernelGGL(hole_pairs_kernel, dim3(grid_for(R * h * h * w)), dim3(NT), 0,
                     (hipStream_t)stream, R, (int)h, (int)P, ldp, w, in, out, expand);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_tau16(int64_t p, int64_t h, const double* t1, const double* t2, void* tau16,
             void* stream) {
  if (p < 0 || h < 0 || !t1 || !t2 || !tau16) return NL_EARG;
  if (p * p * h * h == 0) return NL_OK;
  if (p * p > INT32_MAX || h * h > INT32_MAX || p * h > INT32_MAX || h > NT) return NL_EARG;
  hipLaunchKernelGGL((tau_kernel<uint16_t>), dim3((unsigned)std::min<int64_t>(p * p, 65536)),
                     dim3(NT), 0, (hipStream_t)stream, (int)p, (int)h, nl_fastdiv(p),
                     nl_fastdiv(h), t1, t2, (uint16_t*)tau16);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_transpose16_batched(int64_t batch, int64_t R, int64_t C, const void* in, int64_t ldi,
                           int64_t stride_in, void* out, int64_t ldo, int64_t stride_out,
                           void* stream);

int nl_transpose16_into(int64_t batch, int64_t R, int64_t C, const void* in, int64_t ldi,
                        int64_t stride_in, void* out, int64_t ldo, int64_t stride_out,
                        int64_t rlen, void* stream);

int nl_transpose16(int64_t R, int64_t C, const void* in, int64_t ldi, void* out, int64_t ldo,
                   void* stream) {
  return nl_transpose16_into(1, R, C, in, ldi, 0, out, ldo, 0, ldo, stream);
}

int nl_transpose16_batched(int64_t batch, int64_t R, int64_t C, const void* in, int64_t ldi,
                           int64_t stride_in, void* out, int64_t ldo, int64_t stride_out,
                           void* stream) {
  return nl_transpose16_into(batch, R, C, in, ldi, stride_in, out, ldo, stride_out, ldo, stream);
}

int nl_transpose16_into(int64_t batch, int64_t R, int64_t C, const void* in, int64_t ldi,
                        int64_t stride_in, void* out, int64_t ldo, int64_t stride_out,
                        int64_t rlen, void* stream) {
  if (batch < 0 || R < 0 || C < 0 || ldi < C || rlen < R || rlen > ldo) return NL_EARG;
  if (batch == 0 || C == 0 || rlen == 0) return NL_OK;
  if (!in || !out || nl_cdiv(rlen, 64) > 65535 || batch > 65535) return NL_EARG;
  dim3 g((unsigned)nl_cdiv(C, 64), (unsigned)nl_cdiv(rlen, 64), (unsigned)batch);
  const bool vec = C % 4 == 0 && ldi % 4 == 0 && ldo % 4 == 0 && rlen % 4 == 0 &&
                   stride_in % 4 == 0 && stride_out % 4 == 0 && (((uintptr_t)in) & 7) == 0 &&
                   (((uintptr_t)out) & 7) == 0;
  if (vec)
    hipLaunchKernelGGL(transpose16v_kernel, g, dim3(NT), 0, (hipStream_t)stream, R, C,
                       (const uint16_t*)in, ldi, stride_in, (uint16_t*)out, ldo, stride_out,
                       rlen);
  else
    hipLaunchKernelGGL(transpose16_kernel, g, dim3(NT), 0, (hipStream_t)stream, R, C,
                       (const uint16_t*)in, ldi, stride_in, (uint16_t*)out, ldo, stride_out,
                       rlen);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_permute16(int rank, const int64_t* out_shape, const int64_t* in_strides, const void* in,
                 void* out, void* stream) {
  if (rank < 1 || rank > 6 || !out_shape || !in_strides || !in || !out) return NL_EARG;
  int64_t total = 1;
  for (int d = 0; d < rank; ++d) {
    if (out_shape[d] < 0) return NL_EARG;
    total *= out_shape[d];
  }
  if (total == 0) return NL_OK;
  if (total >= INT32_MAX) return NL_EARG;
  // 4-element vectors when the last dimension is unit-stride, of extent and strides divisible
  // by 4, and both buffers 8-byte aligned
  bool vec = in_strides[rank - 1] == 1 && out_shape[rank - 1] % 4 == 0 &&
             (((uintptr_t)in) & 7) == 0 && (((uintptr_t)out) & 7) == 0;
  for (int d = 0; d + 1 < rank; ++d) vec = vec && in_strides[d] % 4 == 0;
  if (vec) {
    Perm16Args<4> a;
    a.rank = rank; a.total = total / 4;
    for (int d = 0; d < rank; ++d) {
      const int64_t ext = d + 1 < rank ? out_shape[d] : out_shape[d] / 4;
      a.istr[d] = d + 1 < rank ? in_strides[d] / 4 : 1;
      a.fd[d] = nl_fastdiv(ext > 0 ? ext : 1);
    }
    a.in = (const uint16_t*)in; a.out = (uint16_t*)out;
    hipLaunchKernelGGL((permute16_kernel<4>), dim3(grid_for(a.total)), dim3(NT), 0,
                       (hipStream_t)stream, a);
    NL_LAUNCH_CHECK();
    return NL_OK;
  }
  Perm16Args<1> a;
  a.rank = rank;
  a.total = total;
  for (int d = 0; d < rank; ++d) {
    a.istr[d] = in_strides[d];
    a.fd[d] = nl_fastdiv(out_shape[d] > 0 ? out_shape[d] : 1);
  }
  a.in = (const uint16_t*)in; a.out = (uint16_t*)out;
  hipLaunchKernelGGL((permute16_kernel<1>), dim3(grid_for(a.total)), dim3(NT), 0,
                     (hipStream_t)stream, a);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_convert_rows_into(int64_t R1, int64_t R2, int64_t C1, int64_t C2, const int64_t* strides4,
                         const double* in, void* out, int64_t ldo, int64_t cend, void* stream) {
  if (!strides4 || !in || !out) return NL_EARG;
  if (R1 < 0 || R2 < 1 || C1 < 0 || C2 < 1 || cend < C1 * C2 || ldo < cend) return NL_EARG;
  if (R1 * R2 == 0 || cend == 0) return NL_OK;
  if (R1 * R2 > INT32_MAX || ldo > INT32_MAX - CAST_CHUNK) return NL_EARG;
  ConvArgs a;
  a.R2 = R2; a.C = C1 * C2; a.C2 = C2; a.ldo = ldo; a.cend = cend;
  a.fr2 = nl_fastdiv(R2); a.fc2 = nl_fastdiv(C2);
  for (int d = 0; d < 4; ++d) a.s[d] = strides4[d];
  a.in = in; a.out = (uint16_t*)out; a.iscale = nullptr; a.target = 0.f; a.bf16 = 1;
  hipStream_t st = (hipStream_t)stream;
  const int64_t R = R1 * R2;
  if (a.s[1] == 1 && R2 >= 8 && a.s[3] != 1 && nl_cdiv(cend, 64) <= 65535) {
    dim3 g2((unsigned)nl_cdiv(R, 64), (unsigned)nl_cdiv(cend, 64));
    hipLaunchKernelGGL(convert_tiles_kernel, g2, dim3(NT), 0, st, a, R, (const unsigned*)nullptr);
  } else {
    if (nl_cdiv(cend, CAST_CHUNK) > 65535) return NL_EARG;
    dim3 g((unsigned)R, (unsigned)nl_cdiv(cend, CAST_CHUNK));
    hipLaunchKernelGGL(cast_rows_kernel, g, dim3(NT), 0, st, a);
  }
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_tau(int64_t p, int64_t h, const double* t1, const double* t2, double* tau, void* stream) {
  if (p < 0 || h < 0 || !t1 || !t2 || !tau) return NL_EARG;
  int64_t n = p * p * h * h;
  if (n == 0) return NL_OK;
  if (p * p > INT32_MAX || h * h > INT32_MAX || p * h > INT32_MAX || h > NT) return NL_EARG;
  hipLaunchKernelGGL((tau_kernel<double>), dim3((unsigned)std::min<int64_t>(p * p, 65536)), dim3(NT),
                     0, (hipStream_t)stream, (int)p, (int)h, nl_fastdiv(p), nl_fastdiv(h), t1, t2,
                     tau);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_segsum(int64_t S, int64_t X, const int64_t* ptr, const int64_t* src_off, const double* val,
              const int64_t* dst_off, int64_t src_stride, int64_t dst_stride, double alpha,
              const double* src, void* out, int io, void* stream) {
  if (S < 0 || X < 0 || (io & ~NL_C_F32)) return NL_EARG;
  if (S == 0 || X == 0) return NL_OK;
  if (!ptr || !src_off || !val || !dst_off || !src || !out) return NL_EARG;
  int G = 1;
  while (G < X && G < 64) G *= 2;
  const int64_t waves = (S + 64 / G - 1) / (64 / G);
  int blocks = grid_for(waves * 64);
  if (io & NL_C_F32)
    hipLaunchKernelGGL(segsum_kernel<float>, dim3(blocks), dim3(NT), 0, (hipStream_t)stream, S,
                       X, G, ptr, src_off, val, dst_off, src_stride, dst_stride, alpha, src,
                       (float*)out);
  else
    hipLaunchKernelGGL(segsum_kernel<double>, dim3(blocks), dim3(NT), 0, (hipStream_t)stream, S,
                       X, G, ptr, src_off, val, dst_off, src_stride, dst_stride, alpha, src,
                       (double*)out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_pair_outer(int64_t m, const int64_t* pr, const int64_t* qr, const double* C, int64_t ldc,
                  int64_t c1, int64_t n1, int64_t c2, int64_t n2, double* out, void* stream) {
  if (m < 0 || n1 < 0 || n2 < 0 || c1 < 0 || c2 < 0 || c1 + n1 > ldc || c2 + n2 > ldc)
    return NL_EARG;
  if (m == 0 || n1 == 0 || n2 == 0) return NL_OK;
  if (!pr || !qr || !C || !out) return NL_EARG;
  hipLaunchKernelGGL(pair_outer_kernel, dim3((unsigned)std::min<int64_t>(m, 65536)), dim3(NT), 0,
                     (hipStream_t)stream, m, pr, qr, C, ldc, c1, n1, c2, n2, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_gather_rows(int64_t R, int64_t Cn, const int64_t* idx, const double* src, int64_t lds,
                   double* out, void* stream) {
  if (R < 0 || Cn < 0) return NL_EARG;
  if (R * Cn == 0) return NL_OK;
  if (!idx || !src || !out) return NL_EARG;
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)std::min<int64_t>(R, 65536)), dim3(NT), 0,
                     (hipStream_t)stream, R, Cn, idx, src, lds, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_concat_rows(int64_t R, int64_t n1, const double* A, int64_t lda, int64_t n2,
                   const double* B, int64_t ldb, double* out, int64_t ldo, void* stream) {
  if (R < 0 || n1 < 0 || n2 < 0 || lda < n1 || ldb < n2 || ldo < n1 + n2) return NL_EARG;
  if (R == 0 || n1 + n2 == 0) return NL_OK;
  if ((n1 && !A) || (n2 && !B) || !out) return NL_EARG;
  hipLaunchKernelGGL(concat_rows_kernel, dim3((unsigned)std::min<int64_t>(R, 65535)), dim3(NT),
                     0, (hipStream_t)stream, R, n1, A, lda, n2, B, ldb, out, ldo);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_scatter_add_rows(int64_t R, int64_t Cn, const int64_t* idx, const double* src, double alpha,
                        double* out, int64_t ldo, void* stream) {
  if (R < 0 || Cn < 0) return NL_EARG;
  if (R * Cn == 0) return NL_OK;
  if (!idx || !src || !out) return NL_EARG;
  hipLaunchKernelGGL(scatter_add_rows_kernel, dim3((unsigned)std::min<int64_t>(R, 65536)),
                     dim3(NT), 0, (hipStream_t)stream, R, Cn, idx, src, alpha, out, ldo);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_update_t1(int64_t p, int64_t h, const double* t1, const double* H1, const double* Xpp,
                 const double* Xhh, double mixing, double* t1n, void* stream) {
  if (p < 0 || h < 0 || !t1 || !H1 || !Xpp || !Xhh || !t1n) return NL_EARG;
  if (p * h == 0) return NL_OK;
  hipLaunchKernelGGL(update_t1_kernel, dim3(grid_for(p * h)), dim3(NT), 0, (hipStream_t)stream, p,
                     h, t1, H1, Xpp, Xhh, mixing, t1n);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

// Number of partial-sum slots (per quantity) nl_assemble_t2 / nl_energy / nl_dot use.
int64_t nl_partials(int64_t n) { return grid_for(n, 4096); }

static int assemble_t2_impl(int64_t p, int64_t h, const void* S0, const void* Sab,
                            const void* Sij, const void* Spp, const double* Xpp,
                            const double* Xhh, const double* t2, const double* vpphh,
                            const double* t1n, double mixing, double* t2n, double* H2out,
                            double* part, uint16_t* t2n16, void* stream,
                            const double* dpp = nullptr, const double* dhh = nullptr,
                            int split16 = 0, double img_scale = 1.0, int io = 0) {
  if (p < 0 || h < 0 || !Xpp || !Xhh || !t2 || !part) return NL_EARG;
  if (vpphh && !t1n) return NL_EARG;
  if ((io & ~(NL_ACC_F32 | NL_ADD_V | NL_S0_PAIRS | NL_IMG_PAIRS | NL_SAB_PAIRS)) ||
      ((io & NL_ADD_V) && !vpphh) || ((io & NL_S0_PAIRS) && !S0) ||
      ((io & NL_SAB_PAIRS) && !Sab) || ((io & NL_IMG_PAIRS) && (!t2n16 || split16 == 0)))
    return NL_EARG;
  int64_t n = p * p * h * h;
  if (n == 0) return NL_OK;
  AssembleArgs g;
  const bool f32 = (io & NL_ACC_F32) != 0;
  g.p = p; g.h = h; g.S0 = S0; g.Sab = Sab; g.Sij = Sij; g.Spp = Spp;
  g.Xpp = Xpp; g.Xhh = Xhh; g.addv = (io & NL_ADD_V) != 0;
  g.s0pairs = (io & NL_S0_PAIRS) != 0;
  g.sabpairs = (io & NL_SAB_PAIRS) != 0;
  g.imgpairs = (io & NL_IMG_PAIRS) != 0;
  g.s0ld = nl_cdiv(h * (h - 1) / 2, 32) * 32;
  g.t2 = t2; g.vpphh = vpphh; g.t1n = t1n; g.mixing = mixing; g.t2n = t2n; g.part = part;
  g.H2out = H2out; g.t2n16 = t2n16; g.dpp = dpp; g.dhh = dhh; g.split16 = split16;
  g.img_scale = img_scale;
  g.fp = nl_fastdiv(p); g.fh = nl_fastdiv(h);
  const dim3 grid(grid_for(n, 4096));
  hipStream_t st = (hipStream_t)stream;
  if (h <= AH && (Sij || Spp) && p * p < INT32_MAX) {
    const dim3 block((unsigned)nl_cdiv(h * h, 64) * 64);      // one thread per block element
    if (f32) hipLaunchKernelGGL(assemble_pair_kernel<float>, grid, block, 0, st, g);
    else hipLaunchKernelGGL(assemble_pair_kernel<double>, grid, block, 0, st, g);
  } else {
    if (f32) hipLaunchKernelGGL(assemble_t2_kernel<float>, grid, dim3(NT), 0, st, g);
    else hipLaunchKernelGGL(assemble_t2_kernel<double>, grid, dim3(NT), 0, st, g);
  }
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_assemble_t2(int64_t p, int64_t h, const double* S0, const double* Sab, const double* Sij,
                   const double* Spp, const double* Xpp, const double* Xhh, const double* t2,
                   const double* vpphh, const double* t1n, double mixing, double* t2n,
                   double* H2out, double* part, void* stream) {
  return assemble_t2_impl(p, h, S0, Sab, Sij, Spp, Xpp, Xhh, t2, vpphh, t1n, mixing, t2n, H2out,
                          part, nullptr, stream);
}

int nl_assemble_t2_16(int64_t p, int64_t h, const double* S0, const double* Sab,
                      const double* Sij, const double* Spp, const double* Xpp, const double* Xhh,
                      const double* t2, const double* vpphh, const double* t1n, double mixing,
                      double* t2n, double* H2out, double* part, void* t2n16, void* stream) {
  if (!t2n16) return NL_EARG;
  return assemble_t2_impl(p, h, S0, Sab, Sij, Spp, Xpp, Xhh, t2, vpphh, t1n, mixing, t2n, H2out,
                          part, (uint16_t*)t2n16, stream);
}

int nl_assemble_t2_ext(int64_t p, int64_t h, const void* S0, const void* Sab,
                       const void* Sij, const void* Spp, const double* Xpp, const double* Xhh,
                       const double* dpp, const double* dhh, const double* t2,
                       const double* vpphh, const double* t1n, double mixing, double* t2n,
                       double* H2out, double* part, void* t2n16, int image_kind,
                       double image_scale, int io, void* stream) {
  if (image_kind < 0 || image_kind > 2) return NL_EARG;
  return assemble_t2_impl(p, h, S0, Sab, Sij, Spp, Xpp, Xhh, t2, vpphh, t1n, mixing, t2n, H2out,
                          part, (uint16_t*)t2n16, stream, dpp, dhh, image_kind, image_scale, io);
}

__global__ void take_diag_kernel(int64_t n, double* X, double* d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    d[i] = X[i * n + i];
    X[i * n + i] = 0.0;
  }
}

int nl_take_diag(int64_t n, double* X, double* d, void* stream) {
  if (n < 0 || (n > 0 && (!X || !d))) return NL_EARG;
  if (n == 0) return NL_OK;
  hipLaunchKernelGGL(take_diag_kernel, dim3(grid_for(n)), dim3(NT), 0, (hipStream_t)stream, n,
                     X, d);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_assemble_t2_adj(int64_t p, int64_t h, const double* L, double* gSab, double* gSij,
                       double* gSpp, void* stream) {
  if (p < 0 || h < 0 || !L) return NL_EARG;
  int64_t n = p * p * h * h;
  if (n == 0) return NL_OK;
  if (n < INT32_MAX)
    hipLaunchKernelGGL((assemble_adj_kernel<true>), dim3(grid_for(n)), dim3(NT), 0,
                       (hipStream_t)stream, p, h, nl_fastdiv(p), nl_fastdiv(h), L, gSab, gSij,
                       gSpp);
  else
    hipLaunchKernelGGL((assemble_adj_kernel<false>), dim3(grid_for(n)), dim3(NT), 0,
                       (hipStream_t)stream, p, h, nl_fastdiv(1), nl_fastdiv(1), L, gSab, gSij,
                       gSpp);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_energy_part(int64_t p, int64_t h, const double* vpphh, const double* t2, const double* t1,
                   double* part, void* stream) {
  if (p < 0 || h < 0 || !vpphh || !t2 || !t1 || !part) return NL_EARG;
  int64_t n = p * p * h * h;
  if (n == 0) return NL_OK;
  if (n < INT32_MAX)
    hipLaunchKernelGGL((energy_part_kernel<true>), dim3(grid_for(n, 4096)), dim3(NT), 0,
                       (hipStream_t)stream, p, h, nl_fastdiv(p), nl_fastdiv(h), vpphh, t2, t1,
                       part);
  else
    hipLaunchKernelGGL((energy_part_kernel<false>), dim3(grid_for(n, 4096)), dim3(NT), 0,
                       (hipStream_t)stream, p, h, nl_fastdiv(1), nl_fastdiv(1), vpphh, t2, t1,
                       part);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_finalize(int64_t nparts, const double* part, int64_t nx, const double* x, const double* y,
                double* out, void* stream) {
  if (nparts < 0 || nx < 0 || !out || (nparts && !part) || (nx && (!x || !y))) return NL_EARG;
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(NT), 0, (hipStream_t)stream, nparts, part, nx,
                     x, y, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

// ---- allocation-free energy: nl_energy_part + nl_finalize with the partials in a file-scope
// device array instead of a caller workspace.  `scratch` holds NL_SCRATCH_DOUBLES doubles (32 KiB)
// on every GPU the library runs on, until the process ends.  Host code never passes the array's
// name to a kernel: its device address is resolved with hipGetSymbolAddress once per device (with
// that device current, on the first call) and kept in g_scratch.  Every call overwrites each
// partial it reads, so nothing needs clearing; calls that overlap on one GPU share the array and
// race.
}  // extern "C"

constexpr int64_t NL_SCRATCH_DOUBLES = 4096;     // >= nl_partials(n) = grid_for(n, 4096)
static __device__ __align__(256) double scratch[NL_SCRATCH_DOUBLES];
constexpr int NL_MAX_DEVICES = 64;
static double* g_scratch[NL_MAX_DEVICES];
static std::mutex g_scratch_lock;

static int scratch_for_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NL_MAX_DEVICES) return NL_EHIP;
  std::lock_guard<std::mutex> lk(g_scratch_lock);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return NL_EHIP;
    g_scratch[dev] = (double*)ptr;
  }
  *out = g_scratch[dev];
  return NL_OK;
}

extern "C" {

int nl_ccsd_energy_noalloc(int64_t p, int64_t h, const double* f_ph, const double* vpphh,
                           const double* t2, const double* t1, double* E, void* stream) {
  if (p < 0 || h < 0 || !f_ph || !vpphh || !t2 || !t1 || !E) return NL_EARG;
  const int64_t n = p * p * h * h;
  const int64_t nparts = n ? grid_for(n, 4096) : 0;
  if (nparts > NL_SCRATCH_DOUBLES) return NL_EARG;    // beyond the static partials array
  double* part = nullptr;
  const int rc = scratch_for_current_device(&part);
  if (rc != NL_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (n > 0) {
    if (n < INT32_MAX)
      hipLaunchKernelGGL((energy_part_kernel<true>), dim3(nparts), dim3(NT), 0, st, p, h,
                         nl_fastdiv(p), nl_fastdiv(h), vpphh, t2, t1, part);
    else
      hipLaunchKernelGGL((energy_part_kernel<false>), dim3(nparts), dim3(NT), 0, st, p, h,
                         nl_fastdiv(1), nl_fastdiv(1), vpphh, t2, t1, part);
    NL_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(NT), 0, st, nparts, part, p * h, f_ph, t1,
                     E);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_dot_part(int64_t n, const double* x, const double* y, double* part, void* stream) {
  if (n < 0 || !x || !y || !part) return NL_EARG;
  if (n == 0) return NL_OK;
  hipLaunchKernelGGL(dot_part_kernel, dim3(grid_for(n, 4096)), dim3(NT), 0, (hipStream_t)stream,
                     n, x, y, part);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_axpby(int64_t n, double a, const double* x, double b, double* y, void* stream) {
  if (n < 0 || !x || !y) return NL_EARG;
  if (n == 0) return NL_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(NT), 0, (hipStream_t)stream, n, a, x, b,
                     y);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

int nl_sub(int64_t n, const double* x, const double* y, double* c, void* stream) {
  if (n < 0 || !x || !y || !c) return NL_EARG;
  if (n == 0) return NL_OK;
  hipLaunchKernelGGL(sub_kernel, dim3(grid_for(n)), dim3(NT), 0, (hipStream_t)stream, n, x, y, c);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

extern "C++" {
template <typename T>
static int copy_sub_impl(int64_t n, const double* x, const double* y, T* hcopy, T* c,
                         void* stream) {
  if (n < 0 || !x || !y || !hcopy || !c) return NL_EARG;
  if (n == 0) return NL_OK;
  hipLaunchKernelGGL((copy_sub_kernel<T>), dim3(grid_for(n)), dim3(NT), 0, (hipStream_t)stream,
                     n, x, y, hcopy, c);
  NL_LAUNCH_CHECK();
  return NL_OK;
}
}  // extern "C++"

int nl_copy_sub(int64_t n, const double* x, const double* y, double* hcopy, double* c,
                void* stream) {
  return copy_sub_impl(n, x, y, hcopy, c, stream);
}

int nl_copy_sub_f32(int64_t n, const double* x, const double* y, float* hcopy, float* c,
                    void* stream) {
  return copy_sub_impl(n, x, y, hcopy, c, stream);
}

extern "C++" {
template <typename T>
static int gram_part_impl(int64_t n, int k, const T* const* x, double* part, void* stream) {
  if (n < 0 || k < 1 || k > GRAM_MAX || !x || !part) return NL_EARG;
  GramArgs<T> g = {};
  for (int a = 0; a < k; ++a) {
    if (n > 0 && !x[a]) return NL_EARG;
    g.x[a] = x[a];
  }
  const dim3 grid(grid_for(n, 4096)), block(NT);
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
#define NL_GRAM_CASE(K) \
    case K: hipLaunchKernelGGL((gram_part_kernel<K, T>), grid, block, 0, s, n, g, part); break;
    NL_GRAM_CASE(1) NL_GRAM_CASE(2) NL_GRAM_CASE(3) NL_GRAM_CASE(4) NL_GRAM_CASE(5)
    NL_GRAM_CASE(6) NL_GRAM_CASE(7) NL_GRAM_CASE(8) NL_GRAM_CASE(9) NL_GRAM_CASE(10)
    NL_GRAM_CASE(11) NL_GRAM_CASE(12)
#undef NL_GRAM_CASE
  }
  NL_LAUNCH_CHECK();
  return NL_OK;
}
}  // extern "C++"

int nl_gram_part(int64_t n, int k, const double* const* x, double* part, void* stream) {
  return gram_part_impl(n, k, x, part, stream);
}

int nl_gram_part_f32(int64_t n, int k, const float* const* x, double* part, void* stream) {
  return gram_part_impl(n, k, x, part, stream);
}

int nl_gram_finalize(int64_t nparts, int npair, const double* part, double* out, void* stream) {
  if (nparts < 1 || npair < 1 || !part || !out) return NL_EARG;
  hipLaunchKernelGGL(gram_finalize_kernel, dim3(npair), dim3(NT), 0, (hipStream_t)stream, nparts,
                     part, out);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

extern "C++" {
template <typename T>
static int lincomb_impl(int64_t n, int k, const double* coef, const T* const* x, double* y,
                        void* stream) {
  if (n < 0 || k < 1 || k > LINCOMB_MAX || !coef || !x || !y) return NL_EARG;
  if (n == 0) return NL_OK;
  LincombArgs<T> g = {};
  for (int a = 0; a < k; ++a) {
    if (!x[a]) return NL_EARG;
    g.x[a] = x[a];
    g.c[a] = coef[a];
  }
  const dim3 grid(grid_for(n)), block(NT);
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
#define NL_LINCOMB_CASE(K) \
    case K: hipLaunchKernelGGL((lincomb_kernel<K, T>), grid, block, 0, s, n, g, y); break;
    NL_LINCOMB_CASE(1) NL_LINCOMB_CASE(2) NL_LINCOMB_CASE(3) NL_LINCOMB_CASE(4)
    NL_LINCOMB_CASE(5) NL_LINCOMB_CASE(6) NL_LINCOMB_CASE(7) NL_LINCOMB_CASE(8)
    NL_LINCOMB_CASE(9) NL_LINCOMB_CASE(10) NL_LINCOMB_CASE(11) NL_LINCOMB_CASE(12)
    NL_LINCOMB_CASE(13) NL_LINCOMB_CASE(14) NL_LINCOMB_CASE(15) NL_LINCOMB_CASE(16)
#undef NL_LINCOMB_CASE
  }
  NL_LAUNCH_CHECK();
  return NL_OK;
}
}  // extern "C++"

int nl_lincomb(int64_t n, int k, const double* coef, const double* const* x, double* y,
               void* stream) {
  return lincomb_impl(n, k, coef, x, y, stream);
}

int nl_lincomb_f32(int64_t n, int k, const double* coef, const float* const* x, double* y,
                   void* stream) {
  return lincomb_impl(n, k, coef, x, y, stream);
}

}  // extern "C"

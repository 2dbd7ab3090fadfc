// Shared definitions for the NuLattice CCSD HIP kernels (gfx950 / MI355X).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// Status codes returned by every extern "C" entry point (0 = ok).
#define NL_OK 0
#define NL_EARG -1     // invalid argument (shape, null pointer, unsupported layout)
#define NL_EHIP -2     // a HIP runtime call failed (launch error)

#define NL_LAUNCH_CHECK()                                   \
  do {                                                      \
    hipError_t e_ = hipGetLastError();                      \
    if (e_ != hipSuccess) return NL_EHIP;                   \
  } while (0)

static inline int64_t nl_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Division by a fixed divisor d >= 1 for dividends 0 <= i < 2^31 without the GPU's multi-
// instruction integer divide: i / d = (i * mg) >> sh with sh = 31 + ceil(log2 d) and
// mg = ceil(2^sh / d) (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994: mg * d lies in [2^sh, 2^sh + 2^(sh-31)), so the quotient is exact
// for every 31-bit dividend).
struct FastDiv { int d; int sh; uint64_t mg; };

static inline FastDiv nl_fastdiv(int64_t d) {
  FastDiv f;
  f.d = d > 0 ? (int)d : 1;
  int l = 0;
  while ((int64_t(1) << l) < f.d) ++l;
  f.sh = 31 + l;
  f.mg = ((uint64_t(1) << f.sh) + (uint64_t)f.d - 1) / (uint64_t)f.d;
  return f;
}

__device__ inline int fdiv(int i, const FastDiv& f) {
  return (int)(((uint64_t)(uint32_t)i * f.mg) >> f.sh);
}

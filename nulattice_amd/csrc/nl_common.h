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

// Element types of a product's output C and addend D (the "io" flags of the entry points):
// FP64 by default; FP32 for the accumulators of the precision form, whose read-modify-write
// passes then move half the bytes.  A set flag selects float for that array.
#define NL_C_F32 1
#define NL_D_F32 2
// Flags of the T2 assembly: its four accumulators are FP32; it adds v_pphh to the residual.
#define NL_ACC_F32 1
#define NL_ADD_V 2
#define NL_S0_PAIRS 4   // S0 holds the hole pairs i < j of each (a,b) row only
#define NL_IMG_PAIRS 8  // the split image of the new t2 over the hole pairs only
#define NL_SAB_PAIRS 16 // Sab holds the hole pairs i < j of each (a,b) row only

// An array of double or float elements addressed by element offset (a wave-uniform type).
struct AnyOut {
  void* p; bool f32;
  __device__ inline void put(int64_t o, double v) const {
    if (f32) static_cast<float*>(p)[o] = (float)v; else static_cast<double*>(p)[o] = v;
  }
  __device__ inline double get(int64_t o) const {
    return f32 ? (double)static_cast<const float*>(p)[o] : static_cast<const double*>(p)[o];
  }
  __device__ inline AnyOut at(int64_t o) const {
    AnyOut r; r.f32 = f32;
    r.p = f32 ? (void*)(static_cast<float*>(p) + o) : (void*)(static_cast<double*>(p) + o);
    return r;
  }
};
struct AnyIn {
  const void* p; bool f32;
  __device__ inline double get(int64_t o) const {
    return f32 ? (double)static_cast<const float*>(p)[o] : static_cast<const double*>(p)[o];
  }
  __device__ inline AnyIn at(int64_t o) const {
    AnyIn r; r.f32 = f32;
    r.p = f32 ? (const void*)(static_cast<const float*>(p) + o)
              : (const void*)(static_cast<const double*>(p) + o);
    return r;
  }
  __device__ inline explicit operator bool() const { return p != nullptr; }
};

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

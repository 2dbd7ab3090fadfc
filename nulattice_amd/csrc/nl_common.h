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

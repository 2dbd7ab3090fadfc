// Segment plans of the v_ppph intermediates built on the device (setup, not the iteration).
//
// A plan groups the COO entries (c, d, a, k, v) of <cd||ak> (the loop of v_ppph_dgrams,
// ccDgrams.py:13-45) by an output key so that nl_segsum can give every output row to one
// wavefront: key and source offset are linear in the four indices, the entries are ordered by key
// with ties in list order (a stable radix sort of the keys), and every segment gets its start and
// its destination offset.  Two calls: nl_segplan_sort (sort, gather, run-length encode; the number
// of segments S is left in device memory for the caller to read) and nl_segplan_finish (segment
// starts and destination offsets into caller buffers of S+1 and S entries).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>   // host memset, which rocprim's texture iterator needs declared first
#include <rocprim/rocprim.hpp>

#include "nl_common.h"

namespace {

constexpr int NT = 256;

inline int grid_for(int64_t n) { return (int)std::min<int64_t>((n + NT - 1) / NT, 16384); }

// key[e] = sum_j idx[j][e] * wk[j], src[e] = sum_j idx[j][e] * ws[j], perm[e] = e
__global__ __launch_bounds__(NT) void plan_keys_kernel(int64_t n, const int64_t* __restrict__ idx,
                                                       int64_t w0, int64_t w1, int64_t w2,
                                                       int64_t w3, int64_t s0, int64_t s1,
                                                       int64_t s2, int64_t s3,
                                                       uint64_t* __restrict__ key,
                                                       int64_t* __restrict__ src,
                                                       uint32_t* __restrict__ perm) {
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += (int64_t)gridDim.x * NT) {
    const int64_t c = idx[e], d = idx[n + e], a = idx[2 * n + e], k = idx[3 * n + e];
    key[e] = (uint64_t)(c * w0 + d * w1 + a * w2 + k * w3);
    src[e] = c * s0 + d * s1 + a * s2 + k * s3;
    perm[e] = (uint32_t)e;
  }
}

// out_src[e] = src[perm[e]], out_val[e] = sign * val[perm[e]]
__global__ __launch_bounds__(NT) void plan_gather_kernel(int64_t n, const uint32_t* __restrict__ perm,
                                                         const int64_t* __restrict__ src,
                                                         const double* __restrict__ val,
                                                         double sign, int64_t* __restrict__ osrc,
                                                         double* __restrict__ oval) {
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += (int64_t)gridDim.x * NT) {
    const uint32_t q = perm[e];
    osrc[e] = src[q];
    oval[e] = sign * val[q];
  }
}

// dst[s] = sum_t ((u / div[t]) % mod[t]) * mul[t] (mod 0: no modulo), u = ukey[s]; ptr[S] = n
struct DstSpec { int64_t v[9]; };   // (div, mod, mul) x 3

__global__ __launch_bounds__(NT) void plan_dst_kernel(int64_t S, int64_t n,
                                                      const uint64_t* __restrict__ ukey,
                                                      DstSpec spec, int64_t* __restrict__ dst,
                                                      int64_t* __restrict__ ptr) {
  for (int64_t s = blockIdx.x * (int64_t)NT + threadIdx.x; s < S; s += (int64_t)gridDim.x * NT) {
    const int64_t u = (int64_t)ukey[s];
    int64_t o = 0;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      int64_t x = u / spec.v[3 * t];
      if (spec.v[3 * t + 1] > 0) x %= spec.v[3 * t + 1];
      o += x * spec.v[3 * t + 2];
    }
    dst[s] = o;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ptr[S] = n;
}

// Workspace layout (bytes, each part 256-aligned): keys in/out (2 x 8n), perm in/out (2 x 4n),
// src (8n), unique keys (8n), counts (4n), runs (8), rocPRIM temporary storage.
struct PlanWs {
  uint64_t *kin, *kout, *ukey;
  uint32_t *pin, *pout, *cnt;
  int64_t* src;
  uint64_t* runs;
  void* tmp;
  size_t tmp_bytes, total;
};

inline size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

static hipError_t plan_layout(int64_t n, char* base, PlanWs* w) {
  size_t t_sort = 0, t_rle = 0, t_scan = 0;
  const unsigned un = (unsigned)n;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, t_sort, (uint64_t*)nullptr,
                                           (uint64_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, un, 0, 64);
  if (e != hipSuccess) return e;
  e = rocprim::run_length_encode(nullptr, t_rle, (uint64_t*)nullptr, un, (uint64_t*)nullptr,
                                 (uint32_t*)nullptr, (uint64_t*)nullptr);
  if (e != hipSuccess) return e;
  e = rocprim::exclusive_scan(nullptr, t_scan, (uint32_t*)nullptr, (int64_t*)nullptr, (int64_t)0,
                              un, rocprim::plus<int64_t>());
  if (e != hipSuccess) return e;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += al(bytes); return base + o; };
  w->kin = (uint64_t*)take(8 * (size_t)n);
  w->kout = (uint64_t*)take(8 * (size_t)n);
  w->pin = (uint32_t*)take(4 * (size_t)n);
  w->pout = (uint32_t*)take(4 * (size_t)n);
  w->src = (int64_t*)take(8 * (size_t)n);
  w->ukey = (uint64_t*)take(8 * (size_t)n);
  w->cnt = (uint32_t*)take(4 * (size_t)n);
  w->runs = (uint64_t*)take(8);
  w->tmp_bytes = std::max(t_sort, std::max(t_rle, t_scan));
  w->tmp = take(w->tmp_bytes);
  w->total = off;
  return hipSuccess;
}

}  // namespace

extern "C" {

int64_t nl_segplan_ws(int64_t n) {
  if (n < 1 || n > (int64_t)UINT32_MAX) return -1;
  PlanWs w;
  if (plan_layout(n, nullptr, &w) != hipSuccess) return -1;
  return (int64_t)w.total;
}

int nl_segplan_sort(int64_t n, const int64_t* idx, const int64_t* wkey, const int64_t* wsrc,
                    const double* val, double sign, int kbits, void* ws, int64_t ws_bytes,
                    int64_t* src_out, double* val_out, int64_t* nseg, void* stream) {
  if (n < 1 || n > (int64_t)UINT32_MAX || kbits < 1 || kbits > 64) return NL_EARG;
  if (!idx || !wkey || !wsrc || !val || !ws || !src_out || !val_out || !nseg) return NL_EARG;
  PlanWs w;
  if (plan_layout(n, (char*)ws, &w) != hipSuccess) return NL_EHIP;
  if ((int64_t)w.total > ws_bytes) return NL_EARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(plan_keys_kernel, dim3(grid_for(n)), dim3(NT), 0, st, n, idx, wkey[0],
                     wkey[1], wkey[2], wkey[3], wsrc[0], wsrc[1], wsrc[2], wsrc[3], w.kin, w.src,
                     w.pin);
  NL_LAUNCH_CHECK();
  size_t tb = w.tmp_bytes;
  if (rocprim::radix_sort_pairs(w.tmp, tb, w.kin, w.kout, w.pin, w.pout, (unsigned)n, 0,
                                (unsigned)kbits, st) != hipSuccess)
    return NL_EHIP;
  hipLaunchKernelGGL(plan_gather_kernel, dim3(grid_for(n)), dim3(NT), 0, st, n, w.pout, w.src,
                     val, sign, src_out, val_out);
  NL_LAUNCH_CHECK();
  tb = w.tmp_bytes;
  if (rocprim::run_length_encode(w.tmp, tb, w.kout, (unsigned)n, w.ukey, w.cnt, w.runs, st) !=
      hipSuccess)
    return NL_EHIP;
  if (hipMemcpyAsync(nseg, w.runs, 8, hipMemcpyDeviceToDevice, st) != hipSuccess) return NL_EHIP;
  return NL_OK;
}

int nl_segplan_finish(int64_t S, int64_t n, void* ws, int64_t ws_bytes, const int64_t* dst_spec,
                      int64_t* ptr, int64_t* dst, void* stream) {
  if (S < 1 || S > n || n > (int64_t)UINT32_MAX || !ws || !dst_spec || !ptr || !dst)
    return NL_EARG;
  for (int t = 0; t < 3; ++t)
    if (dst_spec[3 * t] < 1 || dst_spec[3 * t + 1] < 0) return NL_EARG;
  PlanWs w;
  if (plan_layout(n, (char*)ws, &w) != hipSuccess) return NL_EHIP;
  if ((int64_t)w.total > ws_bytes) return NL_EARG;
  hipStream_t st = (hipStream_t)stream;
  DstSpec spec;
  for (int t = 0; t < 9; ++t) spec.v[t] = dst_spec[t];
  size_t tb = w.tmp_bytes;
  if (rocprim::exclusive_scan(w.tmp, tb, w.cnt, ptr, (int64_t)0, (unsigned)S,
                              rocprim::plus<int64_t>(), st) != hipSuccess)
    return NL_EHIP;
  hipLaunchKernelGGL(plan_dst_kernel, dim3(grid_for(S)), dim3(NT), 0, st, S, n, w.ukey, spec, dst,
                     ptr);
  NL_LAUNCH_CHECK();
  return NL_OK;
}

}  // extern "C"

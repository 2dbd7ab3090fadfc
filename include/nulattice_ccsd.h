/*
 * C ABI of libnlccsd.so: the MI355X (gfx950) kernels behind the CCSD amplitude iteration of
 * NuLattice's CCM module.
 *
 * The reference is pure Python (src/NuLattice/CCM/coupled_cluster.py, ccDgrams.py) and has no
 * FFI of its own; every entry point below replaces the array arithmetic of one or more of its
 * contraction sites.  A binding layer (ctypes; see nulattice_amd/_lib.py) keeps the reference's
 * Python signatures (ccsd_solver, t1Iter, t2Iter, ccsd_energy, ...) above this ABI.
 *
 * Conventions: all pointers are device pointers to caller-allocated buffers; sizes and strides
 * are in elements; every call is asynchronous on the given hipStream_t (null = default stream)
 * and returns 0 (NL_OK), -1 (invalid argument) or -2 (HIP launch error).  No call allocates,
 * frees or synchronises; no call throws.
 */
#ifndef NULATTICE_CCSD_H
#define NULATTICE_CCSD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of a product's output C and addend D ("io" flags): double unless the flag is
 * set, then float.  The precision form keeps its T2-sized accumulators in FP32, which halves the
 * bytes of every read-modify-write pass over them (no reference counterpart: the reference
 * stores everything in FP64).  nl_contract_f64, nl_segsum take them as an argument; the 16-bit
 * GEMMs (nl_gemm_lp*) as dtype | io << 8. */
#define NL_C_F32 1
#define NL_D_F32 2
/* Flags of nl_assemble_t2_ext: the four accumulators are FP32; v_pphh is added to H2 (S0 does
 * not hold it). */
#define NL_ACC_F32 1
#define NL_ADD_V 2
/* ... and S0 holds the hole pairs i < j only (see nl_hole_pairs): row a*p + b of P32 slots
 * (P = h(h-1)/2 rounded up to a multiple of 32; slots from P on unused), S0[a,b,i,j] =
 * S0[(a,b)][q(i,j)], -S0[(a,b)][q(j,i)] or 0. */
#define NL_S0_PAIRS 4
/* ... the split image of the new t2 (image_kind 1 or 2) is written over the hole pairs only:
 * row a as [hi((b, q)) | lo((b, q))] with P32 slots per b (slots from P on untouched); and Sab
 * comes over the hole pairs like S0 (NL_SAB_PAIRS). */
#define NL_IMG_PAIRS 8
#define NL_SAB_PAIRS 16

/* C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b]  (FP64, MFMA 16x16x4).
 * Replaces every dense opt_einsum pairwise contraction of the diagram library
 * (ccDgrams.py:48-578, coupled_cluster.py:282-284,405-408,528-536) after a layout change.
 * a_kcontig: A(m,k)=A[m*lda+k] (1) or A[k*lda+m] (0); b_ncontig: B(k,n)=B[k*ldb+n] (1) or
 * B[n*ldb+k] (0).  splits>1 splits K and needs ws of batch*splits*M*N doubles. */
int nl_gemm_f64(int a_kcontig, int b_ncontig, int64_t M, int64_t N, int64_t K, double alpha,
                const double* A, int64_t lda, int64_t strideA, const double* B, int64_t ldb,
                int64_t strideB, double beta, double* C, int64_t ldc, int64_t strideC,
                int64_t batch, int64_t splits, double* ws, void* stream);

/* General strided FP64 tensor contraction on the same MFMA tiles: every index group is at most
 * two runs with arbitrary element strides,
 *   m = m1*M2+m2, n = n1*N2+n2, k = k1*K2+k2, batch = b1*B2+b2,
 *   C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * D(m,n)   (D has C's layout; may be C or null)
 * dims = {M1,M2,N1,N2,K1,K2,B1,B2}; sa = A strides {m1,m2,k1,k2,b1,b2};
 * sb = B strides {k1,k2,n1,n2,b1,b2}; sc = C/D strides {m1,m2,n1,n2,b1,b2}.
 * This is how every dense einsum of the diagram library (ccDgrams.py:48-578) runs without an
 * operand transpose.  splits>1 forces the MFMA path with that K split (ws: batch*splits*M*N
 * doubles); splits<=0 chooses automatically between the 128x128 MFMA tiles and the skinny
 * (N<=64 or M<=64) bandwidth kernel and their split, with ws of nl_contract_ws(dims) doubles.
 * io: NL_C_F32 / NL_D_F32 for a float C / D (products accumulate in FP64 either way). */
int64_t nl_contract_ws(const int64_t* dims);
int nl_contract_f64(const int64_t* dims, double alpha, const double* A, const int64_t* sa,
                    const double* B, const int64_t* sb, double beta, const void* D, void* C,
                    const int64_t* sc, int64_t splits, double* ws, int io, void* stream);

/* C[b] = alpha * diag(ia) (A[b] B[b]^T) diag(ib) + beta * D[b] with BF16 (dtype 0) or
 * FP16 (dtype 1) operands, FP32 accumulation (MFMA 16x16x32).  The precision form of the ring
 * and ladder contractions (ccDgrams.py:240-303); no reference counterpart for the precision.
 * K, lda, ldb multiples of 8; A, B 16-byte aligned.  dtype NL_FP8 (2): OCP E4M3 operands, one
 * byte each (MFMA 16x16x128), 256x256 tiles only: K a multiple of 128, lda, ldb and the batch
 * strides of 16, M, N >= 256 with at least 256 tiles x splits x batch.  D (addend, C's layout) may be C or null.
 * C and D are double unless dtype carries NL_C_F32 << 8 / NL_D_F32 << 8 (float), in every
 * nl_gemm_lp* entry point.
 * splits>1 splits K (FP64 partials in ws, batch*splits*M*N doubles, fixed-order reduction). */
int nl_gemm_lp(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
               int64_t lda, int64_t strideA, const float* ia, int64_t strideIa, const void* B,
               int64_t ldb, int64_t strideB, const float* ib, int64_t strideIb, double beta,
               const void* D, void* C, int64_t ldc, int64_t strideC, int64_t batch,
               int64_t splits, double* ws, void* stream);

/* Row-scatter form of nl_gemm_lp: for each batch b and row m with crow[b*M + m] >= 0,
 *   C[crow[b*M + m]*ldc + n] += alpha * ia[b][m] * (A[b] B[b]^T)[m][n] * ib[b][n]   (n < N).
 * The packed pair-block products of the pp ladder and of R56 (ccDgrams.py:215-237 and :13-45,
 * regrouped by connected pair blocks) add their rows straight into S0 / N56 with it, with no
 * FP64 block result and no separate scatter pass.  crow must not name a row twice (over all
 * batches); ia / ib may be null. */
int nl_gemm_lp_rows(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                    int64_t lda, int64_t strideA, const float* ia, int64_t strideIa, const void* B,
                    int64_t ldb, int64_t strideB, const float* ib, int64_t strideIb, void* C,
                    int64_t ldc, const int64_t* crow, int64_t batch, void* stream);

/* nl_gemm_lp (batch 1) with C and D addressed like nl_contract_f64's output: cdims = {M1, M2,
 * N1, N2} (M = M1*M2, N = N1*N2), sc = {m1, m2, n1, n2} strides, C(m,n) at
 * C[(m/M2)*sc0 + (m%M2)*sc1 + (n/N2)*sc2 + (n%N2)*sc3].  Lets the precision form accumulate
 * straight into strided T2 accumulators (the antisymmetriser groups of t2Iter,
 * coupled_cluster.py:476-536) without a row-major staging copy. */
int nl_gemm_lp_strided(int dtype, int64_t K, double alpha, const void* A, int64_t lda,
                       const float* ia, const void* B, int64_t ldb, const float* ib, double beta,
                       const void* D, void* C, const int64_t* cdims, const int64_t* sc,
                       int64_t splits, double* ws, void* stream);

/* out (contiguous, out_shape) = alpha * in[sum_d idx_d * in_strides[d]] + beta * out.
 * Replaces the transposes inside pAB/pIJ (ccDgrams.py:193-213) and opt_einsum's operand
 * permutations. */
int nl_permute(int rank, const int64_t* out_shape, const int64_t* in_strides, double alpha,
               const double* in, double beta, double* out, void* stream);

/* nl_gemm_lp with split operands: A = Ahi + Alo, B = Bhi + Blo (16-bit each, the lo part at the
 * hi part's layout; either lo may be null), C = alpha diag(ia) (Ahi Bhi^T + Ahi Blo^T +
 * Alo Bhi^T) diag(ib) + beta D.  The lo.lo term is below the split's rounding.  The precision
 * form's ring, ladder and dressing products (ccDgrams.py:240-303, coupled_cluster.py:476-536)
 * with operands carried to ~17 bits. */
int nl_gemm_lp_split(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                     const void* Alo, int64_t lda, int64_t strideA, const float* ia,
                     int64_t strideIa, const void* B, const void* Blo, int64_t ldb,
                     int64_t strideB, const float* ib, int64_t strideIb, double beta,
                     const void* D, void* C, int64_t ldc, int64_t strideC, int64_t batch,
                     int64_t splits, double* ws, void* stream);

/* nl_gemm_lp_strided with split operands (see nl_gemm_lp_split). */
int nl_gemm_lp_strided_split(int dtype, int64_t K, double alpha, const void* A, const void* Alo,
                             int64_t lda, const float* ia, const void* B, const void* Blo,
                             int64_t ldb, const float* ib, double beta, const void* D,
                             void* C, const int64_t* cdims, const int64_t* sc, int64_t splits,
                             double* ws, void* stream);

/* nl_gemm_lp_rows with split operands (see nl_gemm_lp_split). */
int nl_gemm_lp_rows_split(int dtype, int64_t M, int64_t N, int64_t K, double alpha,
                          const void* A, const void* Alo, int64_t lda, int64_t strideA,
                          const float* ia, int64_t strideIa, const void* B, const void* Blo,
                          int64_t ldb, int64_t strideB, const float* ib, int64_t strideIb,
                          void* C, int64_t ldc, const int64_t* crow, int64_t batch,
                          void* stream);

/* Row-wise FP64 -> BF16 (dtype 0) / FP16 (dtype 1) conversion of a 4-index view
 * (rows r1*R2+r2, cols c1*C2+c2, element in[r1*s0+r2*s1+c1*s2+c2*s3]) with a per-row scale;
 * iscale[r] receives the inverse scale; columns [C1*C2, ldo) are zero-filled.
 *   target > 0:  power-of-two scale bringing the row max to <= target;
 *   target < 0:  scale exactly |target| / row max (a row whose nonzero elements share one
 *                magnitude, as every block of a contact interaction, converts exactly);
 *   target == 0: plain rounding cast, no scales (iscale unused, may be null).
 * Precision-form operand preparation (no reference counterpart). */
int nl_convert_rows(int dtype, int64_t R1, int64_t R2, int64_t C1, int64_t C2,
                    const int64_t* strides4, const double* in, void* out, int64_t ldo,
                    float* iscale, float target, void* stream);

/* tau[a,b,i,j] = t2 + t1[a,i] t1[b,j] - t1[a,j] t1[b,i] (ccDgrams.py:215-237, 444-472) as the
 * hi/lo split (dtype 0 BF16, 1 FP16) of tau * scale for the precision form with split operands:
 * row (a,b) of out is [hi(h*h) | lo(h*h)]; tau (FP64, may be null) receives the value too. */
int nl_tau_split(int64_t p, int64_t h, const double* t1, const double* t2, double* tau,
                 void* out, int dtype, double scale, void* stream);

/* nl_tau_split over the hole pairs i < j only: tau[a,b,i,j] * scale at slot q(i,j) =
 * i*h - i(i+1)/2 + j - i - 1 (P = h(h-1)/2 pairs; slots P <= q < Pp zero).  grp = 1: row (a,b)
 * of out is [hi(Pp) | lo(Pp)]; grp = p: row a is [hi((b, q)) | lo((b, q))] (the layout of
 * NL_IMG_PAIRS).  t1 may be null (then the image is that of t2).  tau is antisymmetric in
 * (i, j) (ccDgrams.py:215-237 applies it through pIJ), so the hh-ladder, pp-ladder and R56
 * products (coupled_cluster.py:476-536) run on these P columns and, summed over (k, l), on P
 * terms with a factor 2. */
int nl_tau_split_pairs(int64_t p, int64_t h, const double* t1, const double* t2, void* out,
                       int64_t Pp, int64_t grp, int dtype, double scale, void* stream);

/* FP16 delta of a split image against an anchor image (both FP16 [hi(K) | lo(K)] rows, pitch
 * ldx): D[r*ldd + off + k] = FP16((hi + lo) - (hi_a + lo_a)) for k < K (each pair summed in FP32),
 * 0 for K <= k < kp; with with_hi = 1 also D[r*ldd + k] = hi (0 from K to kp) and off = kp, else
 * off = 0.  The anchored ring products of the precision form: W = W(t_A) + the single-rounded
 * product with t2 - t_A, Spp = Spp(t_A) + [t2_hi | t2 - t_A] x [W - W_A | W_A], whose roundings
 * shrink with the difference (no reference counterpart: the precision form has none).  K, kp,
 * ldx, ldd multiples of 8; X, Xa, D 16-byte aligned. */
int nl_split_delta16(int64_t rows, int64_t K, int64_t ldx, const void* X, const void* Xa,
                     void* D, int64_t ldd, int64_t kp, int with_hi, void* stream);
/* FP8 (OCP E4M3) operands of the anchored W product (no reference counterpart): per-row
 * power-of-two scale s (the row maximum times s in (128, 256]), iscale[r] = 1/s, rounding to
 * nearest even, zero from K to kp (D: bytes, pitch ldd).  nl_split_delta8 writes
 * (hi + lo) - (hi_a + lo_a) of two split images (as nl_split_delta16), nl_convert_rows8 an FP64
 * matrix (pitch ldx).  K <= NL_FP8_MAXK; K, kp, ldd multiples of 8.  The FP8 product is
 * nl_gemm_lp with dtype NL_FP8. */
#define NL_FP8 2
#define NL_FP8_MAXK 16384
int nl_split_delta8(int64_t rows, int64_t K, int64_t ldx, const void* X, const void* Xa,
                    void* D, int64_t ldd, int64_t kp, float* iscale, void* stream);
int nl_convert_rows8(int64_t rows, int64_t K, int64_t ldx, const double* X, void* D,
                     int64_t ldd, int64_t kp, float* iscale, void* stream);


/* The ring image of t2 from its hole-pair split image (NL_IMG_PAIRS layout, Pp slots per b):
 * row (a,i) of out is [hi((c,k)) | lo((c,k))] of t2[a,c,i,k] (the ring operand of t2Iter's
 * ccDgrams.py:240-303 products, coupled_cluster.py:476-536), each element the pair slot of
 * (min(i,k), max(i,k)) with the sign of (i,k) (0 on i = k).  Pp a multiple of 32 (at most
 * 16384: h <= 181), in 16-byte aligned. */
int nl_ring_from_pairs(int64_t p, int64_t h, int64_t Pp, const void* in, void* out,
                       void* stream);

/* Hole-pair form of an array antisymmetric in an index pair (i, j) < h: expand = 0:
 * out[r][q][c] = in[r][i][j][c] for the P = h(h-1)/2 pairs q = (i < j), numbered row by row;
 * expand = 1: out[r][i][j][c] = in[r][q(i,j)][c] (i < j), -in[r][q(j,i)][c] (i > j), 0 (i = j).
 * r < R, c < w (FP64); pair-form rows have ldp >= P slots (those from P on untouched).  The
 * hh-ladder intermediate Whh and the R56 result of the precision form (coupled_cluster.py:
 * 476-536) pass between the two forms with it. */
int nl_hole_pairs(int64_t R, int64_t h, int64_t ldp, int64_t w, const double* in, double* out,
                  int expand, void* stream);

/* nl_gather_rows16 writing k < ldk of each output row at row pitch ldo >= ldk (the hi and lo
 * halves of a split operand row are two gathers into one [hi | lo] row). */
int nl_gather_rows16_into(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx,
                          const void* src, int64_t lds, void* out, int64_t ldk, int64_t ldo,
                          void* stream);

/* Both parts of split rows in one pass: the source rows are [hi(ncol) | lo(ncol)] (row pitch
 * lds) and part q of gathered row (b, k) lands transposed at out[(b*ncol + n)*ldo + q*kpad + k]
 * (k < kb; zeros for k in [kb, kpad) and for idx < 0): the [hi(kpad) | lo(kpad)] rows of the
 * pair-block products' right operand (ccDgrams.py:215-237 and :13-45 regrouped by pair
 * blocks), as two nl_gather_rows16_into calls would write them. */
int nl_gather_rows16_split(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx,
                           const void* src, int64_t lds, void* out, int64_t kpad, int64_t ldo,
                           void* stream);

/* nl_convert_rows writing a hi/lo split: with y = x * scale * (per-row scale as in
 * nl_convert_rows; target 0: none), hi = 16-bit(y), lo = 16-bit(y - hi), so hi + lo carries ~17 (BF16) or ~23
 * (FP16) significant bits.  Row r gets nblk blocks of kpad columns, block q at
 * out[r*ldo + q*kpad + c] holding lo when bit q of lomask is set and hi otherwise (columns
 * C1*C2 <= c < kpad zero).  GEMM operands [hi|lo|hi] x [hi|hi|lo] over the K-concatenated blocks
 * sum hi.hi + lo.hi + hi.lo.  Precision form with split operands (no reference counterpart).
 * The source is double, or float with dtype | NL_D_F32 << 8 (an FP32 accumulator). */
int nl_convert_split(int dtype, int64_t R1, int64_t R2, int64_t C1, int64_t C2,
                     const int64_t* strides4, const void* in, void* out, int64_t ldo,
                     int64_t kpad, int nblk, int lomask, float* iscale, float target,
                     double scale, void* stream);

/* nl_convert_split of in - sub (sub: the source's element type and layout; each difference
 * exact in FP64 for FP32 operands), with the per-row inverse scales either computed (iscale_in
 * null: written to iscale as nl_convert_split does) or given (iscale_in: powers of two, no amax
 * pass; iscale may be null): the difference of the ring intermediate from its anchor, scaled
 * like the anchor's own image, in the anchored precision form (no reference counterpart). */
int nl_convert_split_delta(int dtype, int64_t R1, int64_t R2, int64_t C1, int64_t C2,
                           const int64_t* strides4, const void* in, const void* sub, void* out,
                           int64_t ldo, int64_t kpad, int nblk, int lomask, float* iscale,
                           const float* iscale_in, float target, double scale, void* stream);

/* out[b][n][k] (BF16) = src[idx[b*kb + k]*lds + n] for n < ncol, k < kb (0 where idx < 0), and 0
 * for kb <= k < ldk (ldk a multiple of 8): the row gather of the packed pair-block products
 * (ccDgrams.py:215-237, :13-45) fused with the transposing BF16 conversion of their operand. */
int nl_gather_convert(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx, const double* src,
                      int64_t lds, void* out, int64_t ldk, void* stream);

/* nl_gather_convert with BF16 source rows (src: 16-bit elements, row stride lds): the same
 * gather + transpose, no rounding. */
int nl_gather_rows16(int64_t nb, int64_t kb, int64_t ncol, const int64_t* idx, const void* src,
                     int64_t lds, void* out, int64_t ldk, void* stream);

/* tau of nl_tau, rounded to BF16 and stored as [a][b][i][j] 16-bit elements: when every
 * product that reads tau (ccDgrams.py:215-237 ladder, the hh intermediate of
 * coupled_cluster.py:524-526, R56) runs in BF16, the FP64 tau is never formed. */
int nl_tau16(int64_t p, int64_t h, const double* t1, const double* t2, void* tau16, void* stream);

/* out[c*ldo + r] = in[r*ldi + c] for 16-bit elements (r < R, c < C), 0 for R <= r < ldo:
 * re-lays a BF16 operand so that the contracted index is the unit-stride one. */
int nl_transpose16(int64_t R, int64_t C, const void* in, int64_t ldi, void* out, int64_t ldo,
                   void* stream);

/* nl_transpose16 for batch matrices at element strides stride_in / stride_out apart. */
int nl_transpose16_batched(int64_t batch, int64_t R, int64_t C, const void* in, int64_t ldi,
                           int64_t stride_in, void* out, int64_t ldo, int64_t stride_out,
                           void* stream);

/* nl_transpose16_batched writing rows of rlen elements (zeros for R <= r < rlen) at pitch ldo:
 * fills a column range of wider operand rows (the t2 part of the fused Sab operand). */
int nl_transpose16_into(int64_t batch, int64_t R, int64_t C, const void* in, int64_t ldi,
                        int64_t stride_in, void* out, int64_t ldo, int64_t stride_out,
                        int64_t rlen, void* stream);

/* BF16 cast of a 4-index view (as nl_convert_rows, target <= 0) into rows of pitch ldo, writing
 * columns [0, C1*C2) and zeros up to cend: builds operands whose K is a concatenation (the
 * fused Sab product of t2Iter's P(ab) group, coupled_cluster.py:476-536). */
int nl_convert_rows_into(int64_t R1, int64_t R2, int64_t C1, int64_t C2, const int64_t* strides4,
                         const double* in, void* out, int64_t ldo, int64_t cend, void* stream);

/* out (contiguous, out_shape, 16-bit elements) = in[sum_d idx_d * in_strides[d]]: re-lays the
 * BF16 image of t2 for the products that read it in another order (nl_permute's 16-bit form;
 * fewer than 2^31 elements). */
int nl_permute16(int rank, const int64_t* out_shape, const int64_t* in_strides, const void* in,
                 void* out, void* stream);

/* tau[a,b,i,j] = t2[a,b,i,j] + t1[a,i] t1[b,j] - t1[a,j] t1[b,i]: the t2 + pIJ(t1 t1) combination
 * that the pp ladder applies to (ccDgrams.py:215-237 with coupled_cluster.py:524-526). */
int nl_tau(int64_t p, int64_t h, const double* t1, const double* t2, double* tau, void* stream);

/* Segmented COO accumulation: for each segment s,
 *   out[dst_off[s] + x*dst_stride] += alpha * sum_{e=ptr[s]}^{ptr[s+1]-1} val[e] * src[src_off[e] + x*src_stride]
 * for x < X.  Replaces the scalar loops of v_ppph_dgrams (ccDgrams.py:13-45) for the
 * intermediates ret0..ret4 with precomputed segment plans (conflict-free, no atomics).
 * io: NL_C_F32 for a float out (sums in FP64, one rounding per output element). */
int nl_segsum(int64_t S, int64_t X, const int64_t* ptr, const int64_t* src_off, const double* val,
              const int64_t* dst_off, int64_t src_stride, int64_t dst_stride, double alpha,
              const double* src, void* out, int io, void* stream);

/* out[r][:] = src[idx[r]][:] (0 where idx<0) and out[idx[r]][:] += alpha*src[r][:]:
 * pair-row gather/scatter around the block-dense pp ladder and v_ppph contractions
 * (ccDgrams.py:13-45, 215-237 regrouped by connected pair blocks). */
int nl_gather_rows(int64_t R, int64_t Cn, const int64_t* idx, const double* src, int64_t lds,
                   double* out, void* stream);
int nl_scatter_add_rows(int64_t R, int64_t Cn, const int64_t* idx, const double* src, double alpha,
                        double* out, int64_t ldo, void* stream);

/* out[r][c] = A[r][c] for c < n1, B[r][c - n1] for n1 <= c < n1 + n2 (FP64, row pitches lda,
 * ldb, ldo): the rows [t1 | X_pp] of the K-concatenated operand that runs t2Iter's two P(ab)
 * products (coupled_cluster.py:519-536, the A23/A39 and X_pp dressings) as one GEMM. */
int nl_concat_rows(int64_t R, int64_t n1, const double* A, int64_t lda, int64_t n2,
                   const double* B, int64_t ldb, double* out, int64_t ldo, void* stream);

/* t1n = m*t1 + (1-m)*(t1 + H1 / (-(Xpp[a,a] + Xhh[i,i])))
 * (coupled_cluster.py:410-414 and the mixing of :614). */
int nl_update_t1(int64_t p, int64_t h, const double* t1, const double* H1, const double* Xpp,
                 const double* Xhh, double mixing, double* t1n, void* stream);

/* Fused T2 assembly, Jacobi update, mixing and energy partials:
 *   H2  = S0 + pAB(Sab) + pIJ(Sij) + pIJ(pAB(Spp))   (S0, Sab, Sij as [a][b][i][j], Spp as
 *                                                     [(a,i)][(b,j)])
 *   t2n = m*t2 + (1-m)*(t2 + H2 / (-(Xpp_aa + Xpp_bb + Xhh_ii + Xhh_jj)))
 *   part[blk] = partial of 1/4 v.t2n + 1/2 v t1n t1n; part[nblk+blk] = partial of |H2|^2
 * (coupled_cluster.py:536-545, :618, :265-286).  nblk = nl_partials(p*p*h*h). */
int nl_assemble_t2(int64_t p, int64_t h, const double* S0, const double* Sab, const double* Sij,
                   const double* Spp, const double* Xpp, const double* Xhh, const double* t2,
                   const double* vpphh, const double* t1n, double mixing, double* t2n,
                   double* H2out, double* part, void* stream);

/* nl_assemble_t2 that also writes t2n16[t] = BF16(t2n[t]) (16-bit elements, t2n's layout,
 * rounded FP64 -> FP32 -> BF16, each to nearest even: bit for bit the cast of nl_convert_rows):
 * the BF16 image of the new amplitudes that the next iteration's BF16 products read
 * (coupled_cluster.py:602-623 feeding :442-545), written while the value is in a register
 * instead of by a later cast pass over t2. */
int nl_assemble_t2_16(int64_t p, int64_t h, const double* S0, const double* Sab,
                      const double* Sij, const double* Spp, const double* Xpp, const double* Xhh,
                      const double* t2, const double* vpphh, const double* t1n, double mixing,
                      double* t2n, double* H2out, double* part, void* t2n16, void* stream);

/* nl_assemble_t2 (+ an image of t2n when t2n16 is not null) with two optional FP64 diagonal
 * vectors dpp (p) and dhh (h): H2 += (dpp[a] + dpp[b] + dhh[i] + dhh[j]) t2[a,b,i,j] before the
 * update.  The precision form runs the X_pp t2 dressing (coupled_cluster.py:536) in low precision
 * without X_pp's diagonal, whose product with t2 is the largest term of the residual and is
 * added here exactly.  image_kind 0: t2n16 is the BF16 image of nl_assemble_t2_16; 1 (BF16) or 2
 * (FP16): the hi/lo split of t2n * image_scale, row a as [hi(p*h*h) | lo(p*h*h)].
 * io: NL_ACC_F32 when S0, Sab, Sij, Spp are float (the precision form's FP32 accumulators);
 * NL_ADD_V adds v_pphh (the vpphh argument, required then) to H2 in FP64, for an S0 that does
 * not hold the bare interaction term of t2Iter (coupled_cluster.py:477). */
int nl_assemble_t2_ext(int64_t p, int64_t h, const void* S0, const void* Sab,
                       const void* Sij, const void* Spp, const double* Xpp, const double* Xhh,
                       const double* dpp, const double* dhh, const double* t2,
                       const double* vpphh, const double* t1n, double mixing, double* t2n,
                       double* H2out, double* part, void* t2n16, int image_kind,
                       double image_scale, int io, void* stream);

/* d[i] = X[i,i], then X[i,i] = 0, for a row-major n x n FP64 matrix (the diagonal split of a
 * dressing matrix before its low-precision product; coupled_cluster.py:536). */
int nl_take_diag(int64_t n, double* X, double* d, void* stream);

/* Transpose of nl_assemble_t2's antisymmetrisers, for the Lambda (adjoint) equations of t2Iter
 * (coupled_cluster.py:536-545): given the adjoint L[a,b,i,j] of H2,
 *   gSab = pAB(L) as [a][b][i][j];  gSij = pIJ(L), gSpp = pIJ(pAB(L)) as [(a,i)][(b,j)].
 * Any output may be null. */
int nl_assemble_t2_adj(int64_t p, int64_t h, const double* L, double* gSab, double* gSij,
                       double* gSpp, void* stream);

/* Number of partial slots used by nl_assemble_t2 / nl_energy_part / nl_dot_part for n elements. */
int64_t nl_partials(int64_t n);

/* Partials of 1/4 v.t2 + 1/2 v t1 t1 (ccsd_energy, coupled_cluster.py:265-286). */
int nl_energy_part(int64_t p, int64_t h, const double* vpphh, const double* t2, const double* t1,
                   double* part, void* stream);

/* out[0] = sum(part[0:nparts]) + x.y (nx elements), fixed summation order. */
int nl_finalize(int64_t nparts, const double* part, int64_t nx, const double* x, const double* y,
                double* out, void* stream);

/* ccsd_energy without a workspace: out E[0] = f_ph.t1 + 1/4 v_pphh.t2 + 1/2 v_pphh t1 t1, the
 * result of nl_energy_part + nl_finalize bit for bit, with the partial sums in a static device
 * array of the library (4096 doubles per GPU, held until the process ends) instead of a
 * caller-provided workspace.  No counterpart in the reference: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It calls
 * no allocating routine; the array is resolved once per device on the first call.  Calls that
 * overlap on one GPU share the array and race. */
int nl_ccsd_energy_noalloc(int64_t p, int64_t h, const double* f_ph, const double* vpphh,
                           const double* t2, const double* t1, double* E, void* stream);

/* Partials of x.y (DIIS B-matrix, soa/coupled_cluster.py:536-543). */
int nl_dot_part(int64_t n, const double* x, const double* y, double* part, void* stream);

/* y = a*x + b*y and c = x - y (mixing, DIIS extrapolation and error vectors). */
int nl_axpby(int64_t n, double a, const double* x, double b, double* y, void* stream);
int nl_sub(int64_t n, const double* x, const double* y, double* c, void* stream);

/* hcopy = x and c = x - y in one pass: a DIIS step's history entry (the deepcopy of t) and error
 * vector (t - oldT) (coupled_cluster.py:636-641, soa/coupled_cluster.py:523-526). */
int nl_copy_sub(int64_t n, const double* x, const double* y, double* hcopy, double* c,
                void* stream);

/* Gram matrix of k <= 12 vectors of n elements in one pass over each: the DIIS B matrix
 * (coupled_cluster.py:648-656, soa/coupled_cluster.py:536-543), which the reference forms pair by
 * pair.  x is a host array of k device pointers; part holds k(k+1)/2 * nl_partials(n) doubles,
 * pair q (a <= b, row by row: q = a*k - a*(a-1)/2 + b - a) at part[q*nl_partials(n)].
 * nl_gram_finalize sums each pair's partials in a fixed order into out[0:npair]. */
int nl_gram_part(int64_t n, int k, const double* const* x, double* part, void* stream);
int nl_gram_finalize(int64_t nparts, int npair, const double* part, double* out, void* stream);

/* y = sum_k coef[k] * x_k for k <= 16 vectors in one pass: the DIIS extrapolation
 * (coupled_cluster.py:663-669).  coef: host array of k doubles; x: host array of k device
 * pointers. */
int nl_lincomb(int64_t n, int k, const double* coef, const double* const* x, double* y,
               void* stream);

/* The DIIS history kept in FP32, as the SoA driver keeps it with dtype=torch.float32
 * (soa/coupled_cluster.py:453-460, 520-567): nl_copy_sub storing the entry and error vector as
 * float, and the Gram pass and the extrapolation reading float vectors (products, sums and the
 * extrapolated amplitudes in FP64). */
int nl_copy_sub_f32(int64_t n, const double* x, const double* y, float* hcopy, float* c,
                    void* stream);
int nl_gram_part_f32(int64_t n, int k, const float* const* x, double* part, void* stream);
int nl_lincomb_f32(int64_t n, int k, const double* coef, const float* const* x, double* y,
                   void* stream);

/* ---- Symmetry-sector (block-sparse) form of the same iteration ----
 * For Hamiltonians whose single-particle states carry additive conserved labels that identify
 * them (plane waves: momentum mod L, spin and isospin projections), t2[a,b,i,j] is nonzero only
 * in the pair sector l_a + l_b = l_i + l_j and every contraction of t2Iter
 * (coupled_cluster.py:442-545; ccDgrams.py:240-578) is block diagonal; t1 and f_ph vanish.  The
 * packed amplitudes hold the sectors' [ab][ij] blocks back to back. */

/* Grouped FP64 GEMM over a device-side problem list: for problem q,
 *   C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * D(m,n)   (D has C's layout; null = C)
 * with A(m,k) = A[oA + m*sAm + k*sAk], B(k,n) = B[oB + k*sBk + n*sBn], C(m,n) = C[oC + m*sCm +
 * n*sCn].  desc: 12 int64 per problem (M N K oA sAm sAk oB sBk sBn oC sCm sCn); tiles: 3 int32
 * per workgroup (problem, m0, n0) of the output tile of ``mode``: 0 = 64x32 tiles; 1 = 32x32 and
 * 2 = 64x8 tiles with K split over the workgroup's four wavefronts (small-N products and long
 * K); 3 = 32x32 tiles with K split over eight wavefronts (512 threads).  Replaces the per-sector products of the pp ladder (v_pppp_dgrams,
 * ccDgrams.py:215-237), the hh ladders (dgram_klij_abkl :240-251, dgram_cdkl_cdij_abkl
 * :292-303) and the ring terms (dgram_bkcj_acik :253-264, dgram_cdkl_acik_dblj :279-290). */
int nl_gemm_grouped_f64(int mode, int64_t ntile, const int64_t* desc, const int* tiles,
                        double alpha, const double* A, const double* B, double beta,
                        const double* D, double* C, void* stream);

/* Signed index maps (m >= 0: +x at m; m <= -2: -x at -m-2; m = -1: zero / nothing):
 * gather out[e] = in[idx[e]] with that sign (crossed-channel image of the packed t2, whose
 * stored elements are a < b, i < j; packed form of reference-layout tensors); scatter
 * out[idx[e]] = in[e] with that sign (the packed t2 back to the reference layout t2[a,b,i,j]
 * that ccsd_solver returns, coupled_cluster.py:687: four calls for the four antisymmetric
 * images). */
int nl_sector_gather(int64_t n, const int64_t* idx, const double* in, double* out, void* stream);
int nl_sector_scatter(int64_t n, const int64_t* idx, const double* in, double* out, void* stream);

/* Diagonal dressings X_pp[a,a], X_hh[i,i] of t2Iter (coupled_cluster.py:501-515; dgram_cdkl_bdkl
 * and dgram_cdkl_cdjl, ccDgrams.py:502-526): X[o] = base[o] + alpha * sum over the elem entries
 * of output o of V[elem[e]] * T[elem[e]].  The entries are cut into nchunk chunks of consecutive
 * entries (chunk c = cstart[c]..cstart[c+1]; chunks of output o = cptr[o]..cptr[o+1]), summed
 * into ws[0:nchunk] and then added per output in chunk order. */
int nl_sector_dress(int64_t no, const int64_t* cptr, int64_t nchunk, const int64_t* cstart,
                    const int* elem, const double* V, const double* T, const double* base,
                    double alpha, double* ws, double* X, void* stream);

/* The ring terms of t2Iter (dgram_bkcj_acik, dgram_cdkl_acik_dblj: ccDgrams.py:253-264,
 * 279-290) for crossed sectors of at most 32 pairs, one workgroup per pair of sectors {x, -x}:
 * pairs holds 7 int64 per workgroup (offsets of Tr_x and Tr_-x in the crossed layout, of VPH_x and
 * VPH_-x, |R_x|, |R_-x|, self-conjugate flag).  Tr is gathered from the packed amplitudes T
 * through the signed map rmap, W_x = 1/2 VPr_x Tr_-x - VPH_x stays in LDS, and
 * S_x = Tr_x W_-x is written at the offset of Tr_x.  max_rows is the largest |R| in the table:
 * the kernel holds each sector's images in 32 x 32 LDS tiles, so max_rows > 32 is rejected
 * (NL_EARG; such layouts take the grouped-GEMM path instead). */
int nl_sector_ring(int64_t npair, int64_t max_rows, const int64_t* pairs, const int64_t* rmap,
                   const double* T, const double* VPr, const double* VPH, double* S,
                   void* stream);

/* Reverse sweep of nl_sector_ring for the Lambda equations of the sector form (no reference
 * counterpart; the transpose of the ring terms of t2Iter): from the adjoint gS of Spp in the
 * crossed layout, gTr_x = gS_x W_-x^T + 1/2 VPr_-x^T (Tr_x^T gS_x) for every crossed sector,
 * written at the offsets of Tr_x (W recomputed in LDS).  Same 32-pair limit (max_rows). */
int nl_sector_ring_adj(int64_t npair, int64_t max_rows, const int64_t* pairs,
                       const int64_t* rmap, const double* T, const double* VPr, const double* VPH,
                       const double* gS, double* gTr, void* stream);

/* out[e] += alpha * (signed in[idx[e]]) (signed map as above): the reverse sweep's gathers. */
int nl_sector_gather_add(int64_t n, const int64_t* idx, const double* in, double alpha,
                         double* out, void* stream);

/* out[e] += alpha * W[e] * (G[a] + G[b] + G[p+i] + G[p+j]), (a,b,i,j) = orb[0..3][e]: the
 * transpose of the dressed-diagonal terms of t2Iter (coupled_cluster.py:536-545) in the
 * reverse sweep. */
int nl_sector_diag_axpy(int64_t n, int64_t p, const int* orb, const double* W, const double* G,
                        double alpha, double* out, void* stream);

/* Sector blocks of the plane-wave SU(4) contact, written on the GPU instead of gathered from
 * host arrays (the NO2B two-body blocks of coupled_cluster.py:99-263 / 752-832 for the
 * plane-wave transform of lattice.contacts): block b is out[off[b]..off[b+1]) (off has nblk+1
 * entries, off[nblk] = total), row-major with ncol[b] columns; rows / columns carry the spin codes
 * rs[roff[b]+r], cs[coff[b]+c] (4 s_1 + s_2) of their pairs, and
 *   out = g [d(s_p,s_r) d(s_q,s_s) - d(s_p,s_s) d(s_q,s_r)]
 * with (p,q) = the row pair, (r,s) = the column pair (crossed = 0), or (p,q) = (column first, row
 * second), (r,s) = (row first, column second) (crossed = 1: W[(c,k),(b,j)] = <b k||c j>). */
int nl_contact_fill(int64_t nblk, const int64_t* off, const int64_t* ncol, const int64_t* roff,
                    const int64_t* coff, const unsigned char* rs, const unsigned char* cs,
                    int crossed, double g, int64_t total, double* out, void* stream);

/* Segment plans of the v_ppph intermediates built on the device (the per-output-row grouping of
 * the COO loop of v_ppph_dgrams, ccDgrams.py:13-45, that nl_segsum consumes).  For a (4, n) list
 * idx of entries (c, d, a, k) with values val: key = wkey . (c,d,a,k) < 2^kbits, src_out = wsrc .
 * (c,d,a,k) and val_out = sign * val, both ordered by key with equal keys in list order (stable
 * radix sort); the number of distinct keys S is written to nseg (device).  nl_segplan_finish then
 * writes ptr[0..S] (segment starts, ptr[S] = n) and dst[s] = sum_t ((u / div_t) % mod_t) * mul_t
 * for the s-th key u, dst_spec = {div, mod, mul} x 3 (mod 0: none).  wkey, wsrc and dst_spec
 * are host arrays; ws is a device workspace of nl_segplan_ws(n) bytes shared by the two calls;
 * n < 2^32. */
int64_t nl_segplan_ws(int64_t n);
int nl_segplan_sort(int64_t n, const int64_t* idx, const int64_t* wkey, const int64_t* wsrc,
                    const double* val, double sign, int kbits, void* ws, int64_t ws_bytes,
                    int64_t* src_out, double* val_out, int64_t* nseg, void* stream);
int nl_segplan_finish(int64_t S, int64_t n, void* ws, int64_t ws_bytes, const int64_t* dst_spec,
                      int64_t* ptr, int64_t* dst, void* stream);

/* Orbital-pair images of a basis change, out[r][a][b] = C[pr[r]][c1+a] * C[qr[r]][c2+b] for the m
 * lattice-state pairs (pr[r], qr[r]) on which the two-body interaction is supported (C row-major
 * with ldc columns).  The quarter transformations of a two-body operator into the Hartree-Fock
 * basis (the rotation of the V_eff that HF/hartree_fock.py:70-213 contracts, applied to the
 * coupled_cluster.py:752-832 blocks) become two GEMMs over the support: v'[(ab),(cd)] =
 * P_ab^T (V_r P_cd). */
int nl_pair_outer(int64_t m, const int64_t* pr, const int64_t* qr, const double* C, int64_t ldc,
                  int64_t c1, int64_t n1, int64_t c2, int64_t n2, double* out, void* stream);

/* Plane-wave contact blocks in the reference's dense layout, written on the GPU (the v_pphh,
 * v_phph, v_phhh, v_hhhh blocks that get_norm_ordered_ham returns, coupled_cluster.py:752-832
 * with the sorting of coupled_cluster.py:99-263, for the plane-wave transform of
 * lattice.contacts).  Orbital codes pack kx | ky << 8 | kz << 16 | sigma << 24 (L <= 255);
 *   out[i0][i1][i2][i3] = g delta(k0 + k1 = k2 + k3 mod L)
 *                           [d(s0,s2) d(s1,s3) - d(s0,s3) d(s1,s2)]
 * for codes c0[i0], c1[i1], c2[i2], c3[i3]; each element is (double)(spin factor) * g. */
int nl_pw_dense(int64_t n0, int64_t n1, int64_t n2, int64_t n3, const int* c0, const int* c1,
                const int* c2, const int* c3, int L, double g, double* out, void* stream);

/* Padded pair blocks of the plane-wave contact (the packed form of the v_pppp / v_ppph COO lists
 * of coupled_cluster.py:219-263 that the pp ladder and the ppph GEMM consume): out[b][r][c] =
 * <x y||u w> for row pair rows[b*mb+r] = x*rdiv + y (codes r0[x], r1[y]) and column pair
 * cols[b*kb+c] = u*cdiv + w (codes q0[u], q1[w]), 0 for padding (-1) and for pairs of different
 * total momentum or spin pair; nonzero values are +-g (0.0 + value, as a scatter-add forms it). */
int nl_pw_pair_fill(int64_t nb, int64_t mb, int64_t kb, const int64_t* rows, const int64_t* cols,
                    int64_t rdiv, int64_t cdiv, const int* r0, const int* r1, const int* q0,
                    const int* q1, int L, double g, double* out, void* stream);

/* The plane-wave contact's COO list (coupled_cluster.py:219-260 form, every antisymmetric
 * permutation present) in group-major order: group q holds rows rpair[roff[q]..roff[q+1]) x
 * columns cpair[coff[q]..coff[q+1]), row-major, from entry eoff[q]; entry e is written as
 * idx[0..3][e] = (x, y, u, w) with (x*rdiv + y, u*cdiv + w) its pairs and val[e] = +g when
 * sigma_x = sigma_u, else -g.  Groups are (total momentum, unordered spin pair) classes. */
int nl_pw_coo(int64_t ngrp, const int64_t* eoff, const int64_t* roff, const int64_t* coff,
              const int64_t* rpair, const int64_t* cpair, int64_t rdiv, int64_t cdiv,
              const int* r0, const int* r1, const int* q0, const int* q1, int L, double g,
              int64_t n, int64_t* idx, double* val, void* stream);

/* Fused H2 assembly, Jacobi step, mixing and energy partials on the packed amplitudes
 * (coupled_cluster.py:475-545, 618-620, 623, 265-286): with (a,b,i,j) = orb[0..3][e],
 *   H = V + L + S[sg0] - S[sg1] - S[sg2] + S[sg3] + (X[a] + X[b] + X[p+i] + X[p+j]) t,
 *   tn = mixing t + (1 - mixing) (t + H / -(X[a] + X[b] + X[p+i] + X[p+j])),
 * part[0:g] = partial sums of escale V tn (1 when each antisymmetric element is stored once),
 * part[g:2g] = of H^2, g = nl_partials(n).  L, S (then sg), tn and H2out may be null. */
int nl_sector_assemble(int64_t n, int64_t p, const int* orb, const int* sg, const double* V,
                       const double* L, const double* S, const double* X, const double* t,
                       double mixing, double escale, double* tn, double* H2out, double* part,
                       void* stream);

#ifdef __cplusplus
}
#endif
#endif

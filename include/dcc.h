/*
 * dcc.h -- C ABI of libdcc.so, the MI355X (gfx950) triples engine.
 *
 * Plain C: device pointers are `void*` / `double*`, sizes are int64_t, streams
 * are opaque `dcc_stream_t` handles (a hipStream_t underneath; NULL = the
 * default stream).  Every function returns 0 on success and a nonzero code on
 * failure; `dcc_last_error()` then returns a message for the calling thread.
 * All buffers are caller-owned device memory (from dcc_malloc or any HIP
 * allocator); no entry point allocates except where stated.
 *
 * Reference interfaces each entry point replaces are cited as file:line of
 * jinyuchem/Distributed-CC.  The reference binds `void f(...)` C helpers via
 * ctypes (distr_cc/_lib.py:19, distr_cc/_c_rccsdt.py:23-175); here the same
 * role is played by device-side entry points taking whole canonical-block
 * arrays instead of one v^3 block per call.
 */
#ifndef DCC_H
#define DCC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *dcc_stream_t;
typedef void *dcc_event_t;

/* ---------------------------------------------------------------- runtime */

/* Library version, e.g. 1 for this ABI. */
int dcc_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *dcc_last_error(void);

/* Number of visible HIP devices (0 when none; never an error). */
int dcc_device_count(int *count);

/* Select the device for this thread and report its gfx arch name (>= 32 bytes). */
int dcc_set_device(int device, char *arch_name, int64_t arch_len);

/* Free and total device memory of the current device (bytes). */
int dcc_mem_get_info(int64_t *free_bytes, int64_t *total_bytes);

int dcc_malloc(void **ptr, int64_t bytes);
int dcc_free(void *ptr);
int dcc_memcpy_htod(void *dst, const void *src, int64_t bytes, dcc_stream_t stream);
int dcc_memcpy_dtoh(void *dst, const void *src, int64_t bytes, dcc_stream_t stream);
int dcc_memcpy_dtod(void *dst, const void *src, int64_t bytes, dcc_stream_t stream);
/* Fill with a double value using a kernel (safe inside captured graphs). */
int dcc_fill_f64(double *dst, int64_t n, double value, dcc_stream_t stream);
int dcc_fill_f32(float *dst, int64_t n, float value, dcc_stream_t stream);
int dcc_stream_create(dcc_stream_t *stream);
int dcc_stream_destroy(dcc_stream_t stream);
int dcc_stream_synchronize(dcc_stream_t stream);
int dcc_device_synchronize(void);
int dcc_event_create(dcc_event_t *event);
int dcc_event_destroy(dcc_event_t event);
int dcc_event_record(dcc_event_t event, dcc_stream_t stream);
int dcc_event_elapsed_ms(float *ms, dcc_event_t start, dcc_event_t stop);

/* ------------------------------------------------- contraction (MFMA FP64) */

/*
 * Index map of one GEMM dimension into one operand: an index x in [0, extent)
 * maps to the element offset
 *     table ? table[batch * table_bstride + x]
 *           : (x / (e1*e2)) * s0 + ((x / e2) % e1) * s1 + (x % e2) * s2
 * i.e. either a lookup table or a three-level mixed-radix decomposition.
 */
typedef struct dcc_idxmap {
    const int64_t *table;
    int64_t table_bstride;
    int64_t e1, e2;
    int64_t s0, s1, s2;
} dcc_idxmap;

/* One batch entry of a contraction: base offsets and the row range it owns. */
typedef struct dcc_batch {
    int64_t a_off, b_off, c_off;
    int64_t m_begin, m_end;
} dcc_batch;

/*
 * Batched gather-GEMM on FP64 MFMA (v_mfma_f64_16x16x4f64, LDS-staged tiles):
 *   C[c_off + cm(m) + cn(n)] = beta*C[...] + alpha * sum_k A[a_off + am(m) + ak(k)]
 *                                                        * B[b_off + bk(k) + bn(n)]
 * for every batch entry b and m in [m_begin, m_end), n < N.  `batches` is a
 * device array of nbatch entries.  Entries must write disjoint C elements.
 * Replaces the per-block `einsum` calls of distr_cc/rccsdt.py:385-426,
 * :540-596 and :726-731 (NumPy/BLAS GEMMs on the CPU).
 */
int dcc_contract_f64(int64_t M, int64_t N, int64_t K, double alpha, double beta,
                     const double *A, dcc_idxmap am, dcc_idxmap ak,
                     const double *B, dcc_idxmap bk, dcc_idxmap bn,
                     double *C, dcc_idxmap cm, dcc_idxmap cn,
                     const dcc_batch *batches, int64_t nbatch, int64_t max_rows,
                     dcc_stream_t stream);

/*
 * Same contraction with FP32 operands on FP32 MFMA (v_mfma_f32_16x16x4f32,
 * FP32 accumulation); alpha/beta and the C update are applied in FP64 and
 * rounded to FP32 once.  The mixed-precision form of dcc_contract_f64.
 */
int dcc_contract_f32(int64_t M, int64_t N, int64_t K, double alpha, double beta,
                     const float *A, dcc_idxmap am, dcc_idxmap ak,
                     const float *B, dcc_idxmap bk, dcc_idxmap bn,
                     float *C, dcc_idxmap cm, dcc_idxmap cn,
                     const dcc_batch *batches, int64_t nbatch, int64_t max_rows,
                     dcc_stream_t stream);

/*
 * dcc_contract_f64 / _f32 with an explicit workgroup tile: tile_shape 0 = 64x64, 1 = 128x64,
 * 2 = 64x128, 3 = 128x128 (M x N) with a two-stage register prefetch of the K tiles, 4..7 = the
 * same shapes with a one-stage prefetch; chosen per launch by the engines' autotuner.
 */
int dcc_contract_tiled_f64(int64_t M, int64_t N, int64_t K, double alpha, double beta, const double *A,
                           dcc_idxmap am, dcc_idxmap ak, const double *B, dcc_idxmap bk, dcc_idxmap bn, double *C,
                           dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches, int64_t nbatch,
                           int64_t max_rows, int64_t tile_shape, dcc_stream_t stream);
int dcc_contract_tiled_f32(int64_t M, int64_t N, int64_t K, double alpha, double beta, const float *A,
                           dcc_idxmap am, dcc_idxmap ak, const float *B, dcc_idxmap bk, dcc_idxmap bn, float *C,
                           dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches, int64_t nbatch,
                           int64_t max_rows, int64_t tile_shape, dcc_stream_t stream);

/*
 * dcc_contract_f32 with split-BF16x3 products: each FP32 operand x is split into hi = bf16(x) and
 * lo = bf16(x - hi) (round to nearest even) and every product is accumulated as
 * hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 in FP32 (relative product error ~2^-16).
 * Same arguments and semantics as dcc_contract_f32; the reference has no counterpart (its
 * contractions are FP64 NumPy einsums, distr_cc/rccsdt.py:368-433) -- this is the compensated
 * low-precision variant of the triples contractions.
 */
int dcc_contract_f32_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, double beta, const float *A,
                            dcc_idxmap am, dcc_idxmap ak, const float *B, dcc_idxmap bk, dcc_idxmap bn, float *C,
                            dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches, int64_t nbatch,
                            int64_t max_rows, dcc_stream_t stream);
/* Split-BF16x3 packing of an FP32 tensor: dst[i] = hi | lo << 16 with hi = bf16(src[i]) and
 * lo = bf16(src[i] - hi), both rounded to nearest even; dst may alias src (in-place packing).
 * No reference counterpart (the reference is FP64 throughout). */
int dcc_pack_bf16x3(const float *src, uint32_t *dst, int64_t n, dcc_stream_t stream);

/* dcc_contract_tiled_f32_bf16x3 on packed operands: A and B hold dcc_pack_bf16x3 words (same index
 * maps, element offsets as for the FP32 tensors), so the GEMM does no splitting; C is FP32.
 * No reference counterpart (as dcc_contract_f32_bf16x3). */
int dcc_contract_packed_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, double beta, const uint32_t *A,
                               dcc_idxmap am, dcc_idxmap ak, const uint32_t *B, dcc_idxmap bk, dcc_idxmap bn,
                               float *C, dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches, int64_t nbatch,
                               int64_t max_rows, int64_t tile_shape, dcc_stream_t stream);

/* dcc_contract_f32_bf16x3 with an explicit workgroup tile: tile_shape 0 = 64x64, 1 = 128x64,
 * 2 = 64x128, 3 = 128x128 (256 threads), 4 = 256x128, 5 = 128x256, 6 = 256x256 (512 threads) (M x N) on
 * v_mfma_f32_16x16x32_bf16, 7 = 256x256 (512 threads) on v_mfma_f32_32x32x16_bf16; 8 = 256x256, 9 = 256x128,
 * 10 = 128x256, 11 = 256x256 on 32x32x16, 12 and 13 = 256x256 on four waves (128x128 per wave, AGPR
 * accumulators) on 16x16x32 and 32x32x16, with operands staged by direct global -> LDS requests (two, three,
 * three, two, two and two stages) when both packed operands are K-contiguous in 16-byte quads and K is a
 * positive multiple of 32 x stages (otherwise the register-staged tile 6, 4, 5, 7, 6 or 7 runs); -1 picks
 * 128 on each side unless that side has at most 64 rows.
 * No reference counterpart (as dcc_contract_f32_bf16x3). */
int dcc_contract_tiled_f32_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, double beta, const float *A,
                                  dcc_idxmap am, dcc_idxmap ak, const float *B, dcc_idxmap bk, dcc_idxmap bn,
                                  float *C, dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches, int64_t nbatch,
                                  int64_t max_rows, int64_t tile_shape, dcc_stream_t stream);

/* ---------------------------------------- S3 block transforms (HBM-bound) */

/*
 * One job of the S3 transform: out block at `dst`, source block at `src`
 * (element offsets), and six coefficients c[p] for the six axis
 * permutations p of (a,b,c) in the order
 *   (0,1,2) (0,2,1) (1,0,2) (1,2,0) (2,0,1) (2,1,0)
 * (the reference's tp_t3 order, src/rccsdt.c:32-39), meaning
 *   out[x] = beta*out[x] + sum_p c[p] * src.transpose(p)[x].
 */
typedef struct dcc_s3_job {
    int64_t dst, src;
    double coef[6];
} dcc_s3_job;

/*
 * Apply njobs S3 transforms to v^3 blocks.  Each workgroup computes one 16x16x16
 * output tile from the permuted source tiles, staged one at a time in LDS, so
 * that every HBM read and write is along the contiguous axis.  Replaces take_t3_ijk_single_
 * (src/rccsdt.c:119-129), t3_single_spin_summation_inplace_
 * (src/rccsdt.c:131-233), t3_spin_summation_triple_sym_ (src/rccsdt.c:245-310)
 * and t3_transpose_add_ (src/rccsdt.c:320-375); out and src must not overlap.
 */
int dcc_s3_transform_f64(double *out, const double *src, int64_t nvir, double beta,
                         const dcc_s3_job *jobs, int64_t njobs, dcc_stream_t stream);
/* FP32 storage, FP64 arithmetic. */
int dcc_s3_transform_f32(float *out, const float *src, int64_t nvir, double beta,
                         const dcc_s3_job *jobs, int64_t njobs, dcc_stream_t stream);
/* dcc_s3_transform_f32 with beta = 0 whose output is written as split-BF16x3 words (dcc_pack_bf16x3
 * of each float result): builds a packed GEMM operand in one pass instead of a transform and a
 * packing pass.  No reference counterpart (fuses take_t3_ijk_single_ with the split). */
int dcc_s3_transform_pack_f32(uint32_t *out, const float *src, int64_t nvir, const dcc_s3_job *jobs,
                              int64_t njobs, dcc_stream_t stream);

/*
 * Multi-source S3 transform: each job adds sum_p c[p] * (block at src[p]).transpose(p) to the out
 * block at dst (src[p] ignored where c[p] == 0), reading the out block once.  Used for the
 * symmetrisation sum_pi X[pi(ijk)][pi(abc)] of the T2->T3 base term and for placing the ring
 * operator, which the reference evaluates as separate permuted einsums (distr_cc/rccsdt.py:388-399,
 * 540-596).  Jobs of one call must have distinct dst blocks; out and src must not overlap.
 */
typedef struct dcc_s3_gather_job {
    int64_t dst;
    int64_t src[6];
    double coef[6];
} dcc_s3_gather_job;
int dcc_s3_gather_f64(double *out, const double *src, int64_t nvir, double beta,
                      const dcc_s3_gather_job *jobs, int64_t njobs, dcc_stream_t stream);
int dcc_s3_gather_f32(float *out, const float *src, int64_t nvir, double beta,
                      const dcc_s3_gather_job *jobs, int64_t njobs, dcc_stream_t stream);

/*
 * Fused end of the triples update (distr_cc/rccsdt.py:947-971):
 *   r = (R - (1/6) sum_p R.transpose(p)) / (eia[i,a] + eia[j,b] + eia[k,c])
 *   t3 += r;  optionally r3_out = r;
 *   partial[blk, tile] = (sum r^2, sum over a,b,c < nvir_diis of r^2)
 * R and t3 are (nblk, v, v, v); blk_ijk is a device int32 array (nblk, 3) of
 * the canonical triples; eia is (nocc, v) = e_occ - e_vir - level_shift.
 * `partials` must hold 2 * nblk * dcc_r3_finalize_tiles(nvir) doubles.
 * Replaces PySCF's batched P3_full spin summation (rccsdt.py:947),
 * r3_tri_divide_e_ (rccsdt.py:734-744) and the norms/update (rccsdt.py:957-971).
 */
int64_t dcc_r3_finalize_tiles(int64_t nvir);
int dcc_r3_finalize_f64(const double *R, double *t3, double *r3_out, int64_t nblk, int64_t nvir,
                        const int32_t *blk_ijk, const double *eia, int64_t nocc, int64_t nvir_diis,
                        double *partials, dcc_stream_t stream);
/* FP32 storage of R and t3; projection, division, update and norms in FP64. */
int dcc_r3_finalize_f32(const float *R, float *t3, float *r3_out, int64_t nblk, int64_t nvir,
                        const int32_t *blk_ijk, const double *eia, int64_t nocc, int64_t nvir_diis,
                        double *partials, dcc_stream_t stream);

/*
 * dcc_r3_finalize_f64 followed by dcc_reduce_pairs_f64 of its partials into norms[0..1] (device
 * memory: sum r^2, and the nvir_diis sub-norm), with the partials in a file-scope device array of
 * the library instead of a caller workspace: this entry point allocates nothing (no hipMalloc, no
 * hipFree, no allocating library call).  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  Limit:
 * nblk * dcc_r3_finalize_tiles(nvir) <= 1572864 (24 MiB of partials per device, held until the
 * process ends; nocc=16, nvir=160 needs 1256640), otherwise status 2 and nothing is launched.
 * Every partial it reads is written by the same call, so nothing has to be cleared between calls;
 * calls that overlap on one device (several streams or threads) share the array and race.
 */
int dcc_r3_finalize_noalloc_f64(const double *R, double *t3, double *r3_out, int64_t nblk, int64_t nvir,
                                const int32_t *blk_ijk, const double *eia, int64_t nocc, int64_t nvir_diis,
                                double *norms, dcc_stream_t stream);

/* Deterministic fixed-order sum of `n` pairs (x[2i], x[2i+1]) into out[0..1]. */
int dcc_reduce_pairs_f64(const double *x, int64_t n, double *out, dcc_stream_t stream);

/* ------------------------------------------------ GEMM operand builders */

/*
 * pfs[s][p][q] (6 x nblk x nblk): coefficient of source canonical block p,
 * transposed by permutation s, in target canonical block q for the W_oooo and
 * F_oo terms.  Same coefficients as the prefactor loop of
 * compute_r3_Foo_Woooo_contribution_batch_ (distr_cc/rccsdt.py:616-723),
 * computed for all source/target pairs at once on the device.
 */
int dcc_oooo_prefactors_f64(const double *W_oooo, const double *F_oo, int64_t nocc, const int32_t *blk_ijk,
                            int64_t nblk, double *pfs, dcc_stream_t stream);
int dcc_oooo_prefactors_f32(const float *W_oooo, const float *F_oo, int64_t nocc, const int32_t *blk_ijk,
                            int64_t nblk, float *pfs, dcc_stream_t stream);

/*
 * Ring-term weight matrices ((nocc*nvir) x (nocc*nvir), rows (r,x), cols (l,d)):
 *   WR1 = -W_ovov[l,x,r,d], WR3a = W_ovvo[l,x,d,r],
 *   WR3b = -(W_ovov[l,x,r,d] + W_ovvo[l,x,d,r]) / 2
 * operands of the W_ovov/W_ovvo x T3 terms of distr_cc/rccsdt.py:540-596.
 */
int dcc_ring_weights_f64(const double *W_ovov, const double *W_ovvo, int64_t nocc, int64_t nvir, double *WR1,
                         double *WR3a, double *WR3b, dcc_stream_t stream);
int dcc_ring_weights_f32(const float *W_ovov, const float *W_ovvo, int64_t nocc, int64_t nvir, float *WR1,
                         float *WR3a, float *WR3b, dcc_stream_t stream);

/* ------------------------------------------------------ (Q)/[Q] energy */

/*
 * out[a,b,c,i,j,k] (a,b,c < vpad) = T[ijk][abc] for a,b,c < nvir and 0 in the
 * padding, from the canonical i<=j<=k block array t3_tri.  The one-GPU form of
 * the IJK -> ABC redistribution (distr_cc/distribute_t3.py:1544-1642) and of
 * the per-task gather fill_t3_from_ptr_array (src/rccsdt_q.c:665-828).
 */
int dcc_t3_to_abc_f64(const double *t3_tri, int64_t nocc, int64_t nvir, int64_t vpad, double *out,
                      dcc_stream_t stream);

/*
 * IJK -> ABC at one GPU: t3_abc[lin(a,b,c), i, j, k] = T[ijk][abc] for every a <= b <= c (lexicographic
 * canonical index over nvir) and every ordered (i,j,k), from the canonical i<=j<=k block array.  The
 * layout of DistributedT3ABC.t3_local at one rank; replaces redistribute_ijk_to_abc
 * (distr_cc/distribute_t3.py:1544-1642) and its packing helpers (:1228-1420).
 */
int dcc_t3_tri_to_abc_packed_f64(const double *t3_tri, int64_t nocc, int64_t nvir, double *t3_abc,
                                 dcc_stream_t stream);

/*
 * out[a,b,c,i,j,k] (a,b,c < vpad) = T_{ijk}^{abc} for a,b,c < nvir and 0 in the padding, from an ABC
 * array t3_abc[lin(a<=b<=c), i, j, k] (the ABC-prepared tamps rccsdt_q.kernel accepts,
 * distr_cc/rccsdt_q.py:63-96); the per-task gather of fill_t3_from_ptr_array (src/rccsdt_q.c:665-828)
 * for all tasks at once.
 */
int dcc_t3_abc_packed_to_dense_f64(const double *t3_abc, int64_t nocc, int64_t nvir, int64_t vpad, double *out,
                                   dcc_stream_t stream);

/*
 * Per (task, a, b, c, d) of a chunk of (Q) tasks (tasks: int32 (ntask, 4) block
 * starts a0,b0,c0,d0): t4 <- (t4 + z4) / D with D -> 0 when |D| <= 1e-15,
 * P4_444 on the occupied indices, times the 24/12/6/0 factor; then
 * partials[wg] = (<t4, z4>, <t4, z4_sq>) with z4_sq built from vvoo (vpad,vpad,o,o)
 * and t2 (vpad,vpad,o,o).  t4/z4 are (ntask, bs, bs, bs, bs, o, o, o, o) and are
 * not modified.  ws = NULL keeps the P4_444 intermediates in LDS (nocc <= 9);
 * otherwise ws holds ntask * bs^4 * nocc^4 doubles of scratch and any nocc works.
 * Replaces t4_add_, e_abcdijkl_division_, t4_spin_summation_inplace_,
 * t4_multiply_factor_ and the dot products of distr_cc/rccsdt_q.py:1131-1153.
 */
int dcc_q_finalize_f64(const double *t4, const double *z4, const int32_t *tasks, int64_t ntask, int64_t blksize,
                       int64_t nocc, int64_t nvir, int64_t vpad, const double *e_occ, const double *e_vir,
                       const double *vvoo, const double *t2, double *ws, double *partials, dcc_stream_t stream);

/* ------------------------------------- lower-order iteration (around T3) */

/*
 * Strided index space of up to 6 axes; strides are in elements.  Used for the
 * block views, transposes and packing of the singles/doubles update, which
 * the reference writes as NumPy slicing and .transpose (distr_cc/rccsdt.py:
 * 909-933, the intermediates of :935 and the DIIS packing of
 * distr_cc/_runtime.py:151-186).
 */
typedef struct dcc_strided {
    int64_t ndim;
    int64_t shape[6];
    int64_t dst_stride[6];
    int64_t src_stride[6];
} dcc_strided;

/* dst[i] = beta * dst[i] + alpha * src[i] over the index space (dst is not read when beta == 0). */
int dcc_tensor_axpby_f64(double *dst, const double *src, const dcc_strided *layout, double alpha, double beta,
                         dcc_stream_t stream);
int dcc_tensor_axpby_f32(float *dst, const float *src, const dcc_strided *layout, double alpha, double beta,
                         dcc_stream_t stream);
/* Converting forms (FP64 arithmetic, rounded to dst): the FP32-storage triples path exchanges T2, the
 * intermediates and the T3 partial sums with the FP64 lower-order iteration through these. */
int dcc_tensor_axpby_f32_f64(float *dst, const double *src, const dcc_strided *layout, double alpha, double beta,
                             dcc_stream_t stream);
int dcc_tensor_axpby_f64_f32(double *dst, const float *src, const dcc_strided *layout, double alpha, double beta,
                             dcc_stream_t stream);

/*
 * Operands of the T2 -> T3 base term (distr_cc/rccsdt.py:388-399) as one product over K = (d, l) of
 * extent nvir + nocc, for the first occupied indices i0 .. i0+n-1 (i < n below):
 *   Ap[i,b,j,a,K] = W_ovvv[j,a,b,K] (K < nvir), -W_oovo[i0+i,j,a,K-nvir] (K >= nvir)
 *   Bp[i,b,K,k,c] = t2[i0+i,k,K,c]   (K < nvir), t2[K-nvir,k,b,c]        (K >= nvir)
 * with W_ovvv = W_vvvo[a,b,d,j] stored as [j,a,b,d] (nocc, nvir, nvir, nvir), W_oovo = W_vooo[a,l,i,j]
 * stored as [i,j,a,l] (nocc, nocc, nvir, nocc) and t2 (nocc, nocc, nvir, nvir); Ap holds
 * n*nvir*nocc*nvir*(nvir+nocc) and Bp n*nvir*(nvir+nocc)*nocc*nvir elements.  Then
 * X[i0+i,j,k,a,b,c] = sum_K Ap[i,b,j,a,K] Bp[i,b,K,k,c] is the base term of rccsdt.py:388-399 for every
 * ordered (i,j,k), one batched GEMM over (i,b).  The _packed_f32 form writes split-BF16x3 words
 * (dcc_pack_bf16x3 of each float).  No reference counterpart (the reference evaluates the two einsums
 * of the base term separately).
 */
int dcc_t2t3_operands_f64(double *Ap, double *Bp, const double *W_ovvv, const double *W_oovo, const double *t2,
                          int64_t nocc, int64_t nvir, int64_t i0, int64_t n, dcc_stream_t stream);
int dcc_t2t3_operands_f32(float *Ap, float *Bp, const float *W_ovvv, const float *W_oovo, const float *t2,
                          int64_t nocc, int64_t nvir, int64_t i0, int64_t n, dcc_stream_t stream);
int dcc_t2t3_operands_packed_f32(uint32_t *Ap, uint32_t *Bp, const float *W_ovvv, const float *W_oovo,
                                 const float *t2, int64_t nocc, int64_t nvir, int64_t i0, int64_t n,
                                 dcc_stream_t stream);

/*
 * Pair-symmetric W_vvvv x T3 (distr_cc/rccsdt.py:424-426).  Each of the three terms is
 * r[..x..y..] += sum_{d,e} W1[x,y,d,e] T[..d..e..] with the output pair (x, y) and the T3 pair (d, e) at
 * the same two positions of a v^3 block (pair_axes 0 = positions (0,1), 1 = (0,2), 2 = (1,2)) and the
 * third index z shared.  For a pair-symmetric W1 (W1[x,y,d,e] = W1[y,x,e,d], as every dressed W_vvvv
 * is) the term is two products over triangular pairs, of half the flops of the plain one:
 *   US[p(x<=y)] = sum_q WS[p,q] TS[q],  UA[p(x<y)] = sum_q WA[p,q] TA[q],
 *   r[xy] += (US + UA)/2, r[yx] += (US - UA)/2, r[xx] += US/2,
 * with TS[q(d<=e)] = T[de] + T[ed], TA[q(d<e)] = T[de] - T[ed], WS[p,q(d<e)] = W1[p,de] + W1[p,ed],
 * WS[p,q(d=d)] = W1[p,dd], WA[p,q] = W1[p,de] - W1[p,ed].  Pairs are numbered row-major over the upper
 * triangle: q(d<=e) = d v - d(d-1)/2 + e - d (nS = v(v+1)/2 pairs), q(d<e) = d(v-1) - d(d-1)/2 + e-d-1
 * (nA = v(v-1)/2).  kpad >= nS is the padded K of both products.  Layouts:
 *   Wp [nS + nA][kpad]                          WS rows then WA rows, zero past nS / nA columns
 *   Tp [2][nblk][kpad][v] (pair_axes 0), [2][nblk][v][kpad] (1, 2)   TS then TA, zero padded
 *   Up [2][nblk][nS][v]   (pair_axes 0), [2][nblk][v][nS]   (1, 2)   US then UA (UA in the first nA)
 * dcc_vvvv_pair_weights_* builds Wp from W1 (nvir^4); dcc_vvvv_pair_split_* builds Tp from nblk
 * consecutive blocks of T3; the caller forms Up with one batched GEMM (two entries: WS x TS, WA x TA);
 * dcc_vvvv_pair_unfold_* adds it into nblk consecutive blocks of R.  The _packed_f32 forms write
 * split-BF16x3 words.  No reference counterpart (the reference runs the plain v^2 x v^2 einsums).
 */
int dcc_vvvv_pair_weights_f64(const double *W1, int64_t nvir, int64_t kpad, double *Wp, dcc_stream_t stream);
int dcc_vvvv_pair_weights_f32(const float *W1, int64_t nvir, int64_t kpad, float *Wp, dcc_stream_t stream);
int dcc_vvvv_pair_weights_packed_f32(const float *W1, int64_t nvir, int64_t kpad, uint32_t *Wp,
                                     dcc_stream_t stream);
int dcc_vvvv_pair_split_f64(const double *t3, int64_t nblk, int64_t nvir, int32_t pair_axes, int64_t kpad,
                            double *Tp, dcc_stream_t stream);
int dcc_vvvv_pair_split_f32(const float *t3, int64_t nblk, int64_t nvir, int32_t pair_axes, int64_t kpad,
                            float *Tp, dcc_stream_t stream);
int dcc_vvvv_pair_split_packed_f32(const float *t3, int64_t nblk, int64_t nvir, int32_t pair_axes, int64_t kpad,
                                   uint32_t *Tp, dcc_stream_t stream);
int dcc_vvvv_pair_unfold_f64(const double *Up, int64_t nblk, int64_t nvir, int32_t pair_axes, double *R,
                             dcc_stream_t stream);
int dcc_vvvv_pair_unfold_f32(const float *Up, int64_t nblk, int64_t nvir, int32_t pair_axes, float *R,
                             dcc_stream_t stream);

/*
 * dst[i] = beta * dst[i] + alpha * sum_{s < nslices} src[s * n + i] (FP64 sum in slice order): the
 * reduction of split-K partial products of the T3 -> r1/r2/W contractions (distr_cc/rccsdt.py:253-366).
 */
int dcc_sum_slices_f64(double *dst, const double *src, int64_t nslices, int64_t n, double alpha, double beta,
                       dcc_stream_t stream);
int dcc_sum_slices_f32(float *dst, const float *src, int64_t nslices, int64_t n, double alpha, double beta,
                       dcc_stream_t stream);

/* Number of doubles the `partials` scratch of dcc_dot_* must hold. */
int64_t dcc_dot_partials(void);

/*
 * out[0] = sum_i x[i] y[i] in FP64, summed in a fixed order (deterministic).
 * Replaces the np.linalg.norm / np.dot calls of distr_cc/rccsdt.py:929,957-963
 * and distr_cc/diis.py:122-128.
 */
int dcc_dot_f64(const double *x, const double *y, int64_t n, double *partials, double *out, dcc_stream_t stream);
int dcc_dot_f32(const float *x, const float *y, int64_t n, double *partials, double *out, dcc_stream_t stream);

/*
 * rank 1: x[i,a] /= eia[i,a]; rank 2: x[i,j,a,b] /= eia[i,a] + eia[j,b], with
 * eia = e_i - e_a - level_shift (the r1r2_divide_e_ call of distr_cc/rccsdt.py:924).
 */
int dcc_denom_divide_f64(double *x, int64_t nocc, int64_t nvir, int64_t rank, const double *eia,
                         dcc_stream_t stream);

/* ------------------------------------------ batched ensembles of small systems */

/*
 * One contraction of a grouped launch: the arguments of dcc_contract_f64 in descriptor form.
 * tiles_m, tiles_n, tile_begin and flags are filled by dcc_grouped_prepare.  The reference has no
 * counterpart (it runs one system per MPI job); these entry points pack many small RCCSDT systems
 * (uneven nocc/nvir) into one persistent launch per residual term.
 */
typedef struct dcc_gemm_problem {
    int64_t N, K;
    double alpha, beta;
    const void *A;
    const void *B;
    void *C;
    dcc_idxmap am, ak, bk, bn, cm, cn;
    const dcc_batch *batches;
    int64_t nbatch, max_rows;
    int64_t tiles_m, tiles_n, tile_begin, flags;
} dcc_gemm_problem;

typedef void *dcc_graph_t;

/* Host side: tile every problem for tile shape 0..7 (as dcc_contract_tiled_f64: 0..3 = 64x64, 128x64,
 * 64x128, 128x128 with two-stage register prefetch, 4..7 the same with one stage), fill the tiling
 * fields and flags, and return the total tile count and the kernel variant to launch with. */
int dcc_grouped_prepare(dcc_gemm_problem *probs, int64_t nprob, int64_t tile_shape, int64_t *total_tiles,
                        int64_t *variant);

/* Persistent grouped FP64 gather-GEMM over a device array of prepared problems whose C regions are
 * disjoint; variant from dcc_grouped_prepare; nworkgroups = 0 launches as many workgroups as are
 * resident at once.  Every workgroup exits when the tile list is exhausted. */
int dcc_contract_grouped_f64(const dcc_gemm_problem *probs, int64_t nprob, int64_t total_tiles,
                             int64_t variant, int64_t nworkgroups, dcc_stream_t stream);

/* HIP graph capture of everything issued on a (non-default) stream, and its replay. */
int dcc_graph_capture_begin(dcc_stream_t stream);
int dcc_graph_capture_end(dcc_stream_t stream, dcc_graph_t *graph);
int dcc_graph_launch(dcc_graph_t graph, dcc_stream_t stream);
int dcc_graph_destroy(dcc_graph_t graph);

#ifdef __cplusplus
}
#endif

#endif /* DCC_H */

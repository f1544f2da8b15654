/*
 * f2jk.h -- C ABI of the MI355X density-fitted J/K builder (libf2jk.so).
 *
 * Replaces the numpy/BLAS body of forte2's FockBuilder:
 *   forte2/jkbuilder/jkbuilder.py:209-251  FockBuilder.build_JK(C)  -> f2jk_build_jk
 *   forte2/jkbuilder/jkbuilder.py:161-166  FockBuilder.build_J(D)   -> f2jk_build_j
 *   forte2/jkbuilder/jkbuilder.py:184-207  FockBuilder.build_K(C)   -> f2jk_build_k
 *   forte2/jkbuilder/jkbuilder.py:52-80    cached B_Pmn / B_nPm     -> f2jk_set_B (borrowed, resident in HBM)
 *
 * Conventions
 *   - B is the density-fitted 3-index tensor B[P][m][n] (naux x nbf x nbf,
 *     row-major, symmetric in m,n), exactly the reference's B_Pmn layout.
 *     The library never copies or frees it; the caller keeps it alive.
 *   - C_k is an nbf x nocc_k row-major FP64 matrix with row stride ldc_k
 *     (a column slice C[:, :nocc] of an nbf x nmo array has ldc = nmo).
 *   - D, J, K are nbf x nbf row-major FP64 (D with row stride ldd, J/K dense).
 *   - All pointers are device pointers.  Every call is stream-ordered on the
 *     given hipStream_t (NULL = default stream) and never synchronises.
 *   - Scratch memory is supplied by the caller (ws, ws_bytes); query its size
 *     with f2jk_workspace_size.  The library allocates device memory only in
 *     f2jk_ctx_create / f2jk_set_B / f2jk_set_tuning("j_sparse") (small
 *     index tables).
 *   - Reductions are deterministic: a result depends only on the inputs and
 *     the context's shapes, never on concurrent work.
 *   - Return value: 0 on success, a negative f2jk_status otherwise; the
 *     message of the last failure on the calling thread is f2jk_last_error().
 */
#ifndef F2JK_H
#define F2JK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct f2jk_ctx f2jk_ctx;

enum f2jk_status {
    F2JK_OK = 0,
    F2JK_EINVAL = -1,      /* bad argument (shape, pointer, enum)            */
    F2JK_EHIP = -2,        /* a HIP runtime call or kernel launch failed     */
    F2JK_ENOMEM = -3,      /* workspace too small / allocation failed        */
    F2JK_EUNSUPPORTED = -4,/* combination not implemented                    */
    F2JK_ESTATE = -5       /* context not ready (B not set for this dtype)   */
};

/* storage type of a B copy held by the context */
enum f2jk_dtype { F2JK_F64 = 0, F2JK_F32 = 1 };

/* arithmetic of a build.  FP64: B64, FP64 MFMA, FP64 accumulation.
 * FP32: B32 for J (FP64 accumulation) and FP32 MFMA for K (FP32 accumulation
 * within a P block, FP64 across blocks).
 * BF16X3: B32 for J; K operands split into bf16 hi/lo pairs, three bf16 MFMA
 * products (hi*hi + hi*lo + lo*hi) with FP32 accumulation (over the P range of
 * one split in the contraction, FP64 across splits). */
enum f2jk_precision { F2JK_PREC_FP64 = 0, F2JK_PREC_FP32 = 1, F2JK_PREC_BF16X3 = 2 };

/* what f2jk_build_jk computes; F2JK_TIMING records per-phase HIP events */
enum f2jk_flags { F2JK_WANT_J = 1, F2JK_WANT_K = 2, F2JK_TIMING = 16 };

/* phases reported by f2jk_last_timings (milliseconds, in this order) */
enum f2jk_phase {
    F2JK_PH_PREP = 0,      /* C padding, D = C C^T                           */
    F2JK_PH_JGAMMA = 1,    /* J pass 1: gamma_P = sum_mn B_Pmn D_mn          */
    F2JK_PH_JACC = 2,      /* J pass 2: J_mn = sum_P gamma_P B_Pmn (+reduce) */
    F2JK_PH_KX = 3,        /* K half-transform X_Pmi = sum_n B_Pmn C_ni      */
    F2JK_PH_KS = 4,        /* K contraction K_mn += sum_Pi X_Pmi X_Pni       */
    F2JK_PH_KRED = 5,      /* K split reduction + mirror                     */
    F2JK_PH_TOTAL = 6,
    F2JK_NPHASE = 7
};

/* Create a context on HIP device `device`.  Replaces the construction of
 * FockBuilder (forte2/jkbuilder/jkbuilder.py:46-50). */
int f2jk_ctx_create(int device, f2jk_ctx** out);
int f2jk_ctx_destroy(f2jk_ctx* ctx);

/* Register a resident copy of B (borrowed device pointer) of the given dtype.
 * A context may hold one F64 and one F32 copy of the same tensor; both must
 * have the same naux and nbf.  Replaces FockBuilder.B_Pmn
 * (forte2/jkbuilder/jkbuilder.py:52-72) and B_nPm (:74-80); the transposed
 * copy is not needed. */
int f2jk_set_B(f2jk_ctx* ctx, int dtype, const void* B, int64_t naux, int64_t nbf);

/* Shapes of the registered B. */
int f2jk_get_shape(const f2jk_ctx* ctx, int64_t* naux, int64_t* nbf);

/* Launch-shape knobs (no counterpart in the reference, whose BLAS calls have
 * none): "xchunk_bytes" caps the half-transformed chunk X held in the
 * workspace (default 4 GiB); "target_blocks" is the number of workgroups a
 * split launch aims for (default 2048); "j_nontemporal" (0/1, default 1)
 * streams B through non-temporal loads in the J passes; "overlap_jk" (0/1,
 * default 0) runs the J passes of a build_jk on the context's side stream
 * concurrently with K (joined back before the call returns its work to the
 * caller's stream); "bf16_nbw_max" (2, 4, 5, 6, 8, 10) caps the BF16x3
 * half-transform tile at 32 * value occupied columns (tiles of up to 5 run two
 * workgroups per CU); "kx_waves" (4 or 8, default 8) is the number of waves of
 * a BF16x3 half-transform workgroup (8: two per SIMD sharing the tile; tiles
 * of an even number of 32-column units); "ks_tile" (0, 128, 256; default 0 =
 * 256 when nbf >= 768) is the lower-tile edge of the single-job BF16x3
 * contraction (256: one 8-wave workgroup per CU; grouped builds use 128);
 * "kx_wide" (-1, 0, 1; default -1 = above 128 occupied columns) runs the
 * FP64/FP32 half-transform in 128 x 320 tiles (8 waves), which read each
 * row block of B once instead of once per 128-column tile (the same products,
 * bitwise); "kx_rows" (0, 128, 256; default 0 = 256 when nbf >= 768 and nbf % 4 == 0)
 * is the row count of a single-job BF16x3 half-transform tile (256: 8 waves,
 * even widths, coefficient tiles loaded straight into LDS); "ks_dma" (-1, 0,
 * 1; default -1 = with 256-wide tiles) loads the BF16x3 contraction's X
 * tiles straight into LDS instead of staging them through registers (the same
 * products, bitwise); "j_sparse" (0/1, default 0): when 1, every registered B
 * copy is scanned once (now, and at each later f2jk_set_B) for the
 * lower-triangle vector elements where it is nonzero for some P, and the J
 * passes read only those (density-fitted B of a molecule is exactly zero on
 * its screened shell pairs: ~75% of the FP64 elements of (H2O)_64); J
 * elsewhere is exactly zero, as in the full sum.  B's zero pattern must not
 * change while registered (register it again to rescan).  The scan
 * synchronises the device.
 * Set by the autotuner. */
int f2jk_set_tuning(f2jk_ctx* ctx, const char* key, int64_t value);

/* Scratch bytes needed by f2jk_build_jk / _j / _k for these arguments.
 * nocc may be NULL when flags has no F2JK_WANT_K and no F2JK_WANT_J with C. */
int f2jk_workspace_size(const f2jk_ctx* ctx, int ndens, const int64_t* nocc,
                        int precision, int flags, size_t* bytes);

/* build_JK (forte2/jkbuilder/jkbuilder.py:209-251): for k < ndens,
 *   D_k = C_k C_k^T, J_k = J[D_k], K_k = sum_P (B_P C_k)(B_P C_k)^T.
 * J or K may be NULL when the corresponding flag is absent. */
int f2jk_build_jk(f2jk_ctx* ctx, int ndens, const double* const* C,
                  const int64_t* nocc, const int64_t* ldc, double* const* J,
                  double* const* K, int precision, int flags, void* ws,
                  size_t ws_bytes, void* stream);

/* Bytes of the scratch array that f2jk_build_jk_noalloc works in (64 MiB per
 * device, held by the library from the first f2jk_ctx_create on that device
 * until the process ends). */
#define F2JK_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* f2jk_build_jk without a caller workspace and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The build works in the library's per-device scratch
 * array of F2JK_NOALLOC_SCRATCH_BYTES; it calls no hipMalloc/hipFree and no
 * allocating library routine.  Launch shapes are the context's, lowered (X
 * chunk, split count) until the build fits the array; a build that does not
 * fit even at one P per chunk and one split fails with F2JK_ENOMEM before
 * anything is launched (one density of a cc-pVDZ water cluster fits up to
 * nbf 1152, (H2O)_48; nbf 1536, (H2O)_64, does not).  Results equal
 * f2jk_build_jk's to its tolerances, bitwise when the context's own launch
 * shapes already fit.  Every call on a device shares that device's array:
 * calls that overlap on one device (other streams, other threads) race, so
 * callers serialise them. */
int f2jk_build_jk_noalloc(f2jk_ctx* ctx, int ndens, const double* const* C,
                          const int64_t* nocc, const int64_t* ldc, double* const* J,
                          double* const* K, int precision, int flags, void* stream);

/* build_J (forte2/jkbuilder/jkbuilder.py:161-166): J_k = sum_P B_P <B_P, D_k>
 * for symmetric D_k (only the lower triangle of D_k is read). */
int f2jk_build_j(f2jk_ctx* ctx, int ndens, const double* const* D, int64_t ldd,
                 double* const* J, int precision, int flags, void* ws,
                 size_t ws_bytes, void* stream);

/* build_K (forte2/jkbuilder/jkbuilder.py:184-207, the B_nPm path):
 * K_k = sum_{P,i} X_Pmi X_Pni with X_P = B_P C_k. */
int f2jk_build_k(f2jk_ctx* ctx, int ndens, const double* const* C,
                 const int64_t* nocc, const int64_t* ldc, double* const* K,
                 int precision, int flags, void* ws, size_t ws_bytes,
                 void* stream);

/* Per-phase times (ms) of the last build issued with F2JK_TIMING; waits for
 * that build's events.  n <= F2JK_NPHASE values are written. */
int f2jk_last_timings(f2jk_ctx* ctx, double* ms, int n);

/* Number of timed (start, end) intervals per phase in the last build issued
 * with F2JK_TIMING (one per kernel launch group: e.g. one KX interval per X
 * chunk), so that an average launch duration is ms[phase] / counts[phase]. */
int f2jk_last_launches(const f2jk_ctx* ctx, int* counts, int n);

/* Algorithmic work of one build with these arguments (for roofline
 * reporting): bytes of B streamed by the J passes and K flops
 * (2*Q*N^2*o half-transform + N(N+1)*Q*o symmetric contraction). */
int f2jk_work(const f2jk_ctx* ctx, int ndens, const int64_t* nocc, int precision,
              int flags, double* j_bytes, double* k_flops);

/* Grouped builds for the serving layer (no counterpart in the reference,
 * which builds one job at a time): build_JK for njobs independent jobs, each
 * with its own context (B), 1 or 2 densities and its own workspace, in one
 * persistent launch per phase.  Arrays C, nocc, ldc, J, K are flattened over
 * jobs (sum of ndens entries); ws/ws_bytes have one entry per job, sized by
 * f2jk_group_workspace_size.  precision: FP64, FP32 or BF16X3 (one per
 * group); flags: WANT_J|WANT_K (J from the C of each density) or WANT_K alone.
 * A job's results do not depend on what else is in the group (bitwise).  In
 * FP64 they are bitwise those of f2jk_build_jk.  In the precision forms K
 * is bitwise that of f2jk_build_jk with the same target_blocks and ks_tile
 * 128 (grouped builds contract in 128-wide tiles), and J differs from it in
 * summation order (the group reads B with a 2-wide vector).  nbf even; the job's X must fit
 * one chunk (EUNSUPPORTED otherwise). */
int f2jk_group_workspace_size(f2jk_ctx* const* ctxs, int njobs, const int* ndens,
                              const int64_t* nocc, int precision, int flags,
                              size_t* job_bytes);
int f2jk_group_build_jk(f2jk_ctx* const* ctxs, int njobs, const int* ndens,
                        const double* const* C, const int64_t* nocc, const int64_t* ldc,
                        double* const* J, double* const* K, int precision, int flags,
                        void* const* ws, const size_t* ws_bytes, void* stream);

/* Half-transform X[P-p0][m][i] = sum_n B[P][m][n] C[n][i] for p0 <= P < p1,
 * i < ncol, into the caller's X ((p1-p0)*nbf rows, row stride ldx), FP64.  The first step of
 * FockBuilder.B_tensor_gen_block (forte2/jkbuilder/jkbuilder.py:521-552),
 * whose einsum "Pmn,mi,nj->Pij" is finished by a GEMM with C1 on the caller's
 * side; two_electron_integrals_gen_block (:382-420) contracts two such blocks
 * over P.  Workspace: f2jk_workspace_size(ctx, 1, &ncol, FP64, F2JK_WANT_K).  With ldx equal
 * to ncol rounded up to 16 the kernel writes X in place (no copy). */
int f2jk_half_transform(f2jk_ctx* ctx, const double* C, int64_t ncol, int64_t ldc, int64_t p0,
                        int64_t p1, double* X, int64_t ldx, int precision, void* ws, size_t ws_bytes,
                        void* stream);

/* Message of the last failure on this thread ("" if none). */
const char* f2jk_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* F2JK_H */

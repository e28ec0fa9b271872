/*
 * alglin_amd -- C ABI of the MI355X bordered cyclic-reduction BABD solver.
 *
 * Two groups of entry points:
 *
 *  1. The reference's C interface for BorderedCR (BABD_*).  Its header
 *     (Alglin/BABD_C_interface.h) is absent from the reference snapshot; the
 *     signatures below are the ones src_tests/test9-Cinterface.c calls:
 *       BABD_factorize            test9-Cinterface.c:90-91
 *       BABD_factorize_bordered   test9-Cinterface.c:171-174
 *       BABD_solve                test9-Cinterface.c:97
 *       BABD_get_last_error       test9-Cinterface.c:92
 *     mat_fact: 0 or 2 = cyclic reduction + LU, 3 = + QR, 4 = + QRP (the
 *     BABD_Choice values of src/Alglin/BABD.hxx:66-71; 1 = Diaz is not part of
 *     this solver and is rejected).  last_block_fact: the
 *     BORDERED_LAST_Choice values (BABD_BorderedCR.hxx:118-127), 0 = LU.
 *     Return 0 on success.
 *
 *  2. A handle API mirroring alglin::BorderedCR<T>
 *     (src/Alglin/BABD_BorderedCR.hxx:104-1176): create/allocate
 *     (hxx:373-442), the block loaders (cc:237-793), select* (hxx:458-540),
 *     factorize (cc:803-839), solve and multi-RHS solve (cc:1489-1605),
 *     Mv (cc:2111-2166), sparse export/import (cc:2178-2482), last_error
 *     (hxx:445), info (cc:208-235); plus device-resident variants whose data
 *     pointers are HIP device pointers (inputs already in HBM).
 *
 * Element type of a handle: dtype 0 = double, 1 = float; every `void*` data
 * argument points to that type.  Matrices are column-major with a leading
 * dimension, as in the reference.  Handle calls return 0 on success and
 * nonzero on failure (message in alglin_bcr_last_error).
 */
#ifndef ALGLIN_AMD_H
#define ALGLIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t BABD_intType;
typedef double BABD_realType;

/* ---- reference C interface (test9-Cinterface.c) ---------------------------- */
int BABD_factorize(BABD_intType mat_id, BABD_intType mat_fact, BABD_intType last_block_fact, BABD_intType nblock,
                   BABD_intType n, BABD_intType qr, BABD_intType qx, BABD_realType const DE[], BABD_intType ldDE,
                   BABD_realType const H0[], BABD_intType ldH0, BABD_realType const HN[], BABD_intType ldHN,
                   BABD_realType const Hq[], BABD_intType ldHq);
int BABD_factorize_bordered(BABD_intType mat_id, BABD_intType mat_fact, BABD_intType last_block_fact,
                            BABD_intType nblock, BABD_intType n, BABD_intType qr, BABD_intType nr, BABD_intType qx,
                            BABD_intType nx, BABD_realType const DE[], BABD_intType ldDE, BABD_realType const H0[],
                            BABD_intType ldH0, BABD_realType const HN[], BABD_intType ldHN, BABD_realType const Hq[],
                            BABD_intType ldHq, BABD_realType const B[], BABD_intType ldB, BABD_realType const C[],
                            BABD_intType ldC, BABD_realType const D[], BABD_intType ldD);
int BABD_solve(BABD_intType mat_id, BABD_realType rhs_sol[]);
char const* BABD_get_last_error(void);
/* releases the GPU memory of mat_id (no counterpart in test9; callers that
   reuse many ids need it) */
int BABD_free(BABD_intType mat_id);

/* ---- handle API (BorderedCR<T> surface) ------------------------------------ */
typedef struct alglin_bcr* alglin_bcr_t;

alglin_bcr_t alglin_bcr_create(int dtype);                                   /* hxx:373-384 */
void alglin_bcr_destroy(alglin_bcr_t h);
int alglin_bcr_allocate(alglin_bcr_t h, int32_t nblock, int32_t n, int32_t qr, int32_t qx, int32_t nr,
                        int32_t nx);                                         /* cc:42-170 */
int alglin_bcr_select(alglin_bcr_t h, char const* internal, char const* last, char const* last2); /* hxx:458-540 */
int alglin_bcr_dims(alglin_bcr_t h, int32_t out[10]);  /* N n qr qx nr nx nrows ncols Nr Nc (hxx:617-650) */

/* loaders (cc:237-793): `what` is one of "B","C","D","E","DE","DEB" (block nbl),
   "F","Cq","CqF","H" (nbl ignored), "add_B","add_C","add_C2","add_C2F","add_F" */
int alglin_bcr_load(alglin_bcr_t h, char const* what, int32_t nbl, void const* A, int32_t ld);
int alglin_bcr_load_bottom(alglin_bcr_t h, void const* H0, int32_t ld0, void const* HN, int32_t ldN, void const* Hq,
                           int32_t ldQ, void const* Hp, int32_t ldP);   /* cc:289-305 */
int alglin_bcr_load_bottom2(alglin_bcr_t h, void const* C0, int32_t ld0, void const* CN, int32_t ldN, void const* Cq,
                            int32_t ldCq, void const* F, int32_t ldF); /* cc:380-395 */
/* all blocks at once, packed in the reference's internal layout (hxx:972-998):
   D,E: N n*n blocks; B: N n*nx; C: (N+1) nr*n; H: (n+qr) x Nc; Cq: nr x qx; F: nr x nx */
int alglin_bcr_load_packed(alglin_bcr_t h, void const* D, void const* E, void const* B, void const* C,
                           void const* H, void const* Cq, void const* F);

int alglin_bcr_factorize(alglin_bcr_t h);                                    /* cc:803-839 */
int alglin_bcr_solve(alglin_bcr_t h, void* x);                               /* cc:1489-1542 */
int alglin_bcr_solve_n(alglin_bcr_t h, int32_t nrhs, void* X, int32_t ld);   /* cc:1546-1605 */
int alglin_bcr_mv(alglin_bcr_t h, void const* x, void* y);                   /* cc:2111-2116 */
int32_t alglin_bcr_sparse_nnz(alglin_bcr_t h);                               /* hxx:1079-1088 */
int alglin_bcr_sparse_pattern(alglin_bcr_t h, int32_t* I, int32_t* J, int32_t offs); /* cc:2342-2378 */
int alglin_bcr_sparse_values(alglin_bcr_t h, void* V);                       /* cc:2382-2413 */
int alglin_bcr_sparse_load(alglin_bcr_t h, void const* V, int32_t const* I, int32_t r_offs, int32_t const* J,
                           int32_t c_offs, int32_t nnz, int do_error);       /* cc:2417-2482 */
int alglin_bcr_check_matrix(alglin_bcr_t h);                                 /* cc:2538-2659 */
char const* alglin_bcr_last_error(alglin_bcr_t h);                           /* hxx:445 */
char const* alglin_bcr_info(alglin_bcr_t h);                                 /* cc:208-235 */
/* debug mode (hxx:1172): factorize scans the blocks for NaN / Inf first
   (check_matrix, cc:2538-2659) and fails on the first one */
int alglin_bcr_set_debug(alglin_bcr_t h, int on);
/* elimination order (hxx:407-423, cc:139-158): npb block ranges reduced on
   their own, then the chain of their end rows; 1 = pure cyclic reduction.
   Applies from the next factorization. */
int alglin_bcr_set_num_parallel_block(alglin_bcr_t h, int32_t npb);
/* the partition count in use after the reference's shrink rule (2P <= N) */
int32_t alglin_bcr_used_parallel_block(alglin_bcr_t h);
/* copy of another handle's loaded blocks and partition (cc:181-198) */
int alglin_bcr_dup(alglin_bcr_t h, alglin_bcr_t src);
/* transposed solve: not defined in the reference either, fails with its
   message (cc:2089-2101) */
int alglin_bcr_t_solve(alglin_bcr_t h, void* x);
/* print_matlab_script (cc:2493-2534) into buf (len bytes, NUL-terminated,
   truncated); returns the length the whole script needs, -1 on error */
int64_t alglin_bcr_matlab_script(alglin_bcr_t h, char* buf, int64_t len);
/* host copies of the loaded blocks, packed as in alglin_bcr_load_packed;
   any pointer may be null */
int alglin_bcr_get_packed(alglin_bcr_t h, void* D, void* E, void* B, void* C, void* H, void* Cq, void* F);

/* ---- device-resident variants (HIP device pointers) ------------------------ */
/* use caller-owned device arrays (packed layout as above) as the operator */
int alglin_bcr_bind_device(alglin_bcr_t h, void const* D, void const* E, void const* B, void const* C,
                           void const* H, void const* Cq, void const* F);
/* enqueue the factorization of the bound operator; sync=1 waits and returns status */
int alglin_bcr_factorize_device(alglin_bcr_t h, int sync);
/* in-place solve of a device block X (ld x nrhs, ld >= max(nrows, ncols)
   when nrhs > 1; a single column needs max(nrows, ncols) entries); fails
   when the object is not factorized */
int alglin_bcr_solve_device(alglin_bcr_t h, void* X, int32_t nrhs, int64_t ld, int sync);
/* FP64 residual r = b - A x on the operator the handle factors (double
   device vectors; the mixed engine's copy after a MIXED factorize of host
   blocks) */
int alglin_bcr_residual_device(alglin_bcr_t h, double const* x, double const* b, double* r, int sync);
int alglin_bcr_sync(alglin_bcr_t h);
/* the HIP stream the handle launches on (hipStream_t) */
void* alglin_bcr_stream(alglin_bcr_t h);
/* factor arrays (host copies) for tests: T panels, Perm, D, E, B, C after
   factorize; any pointer may be null */
int alglin_bcr_get_factors(alglin_bcr_t h, void* T, int32_t* perm, void* D, void* E, void* B, void* C,
                           double* last_block, double* pending_sumsq);
/* 1 when the last factorization used the register-resident LU kernel */
int alglin_bcr_fast_path(alglin_bcr_t h);
/* LU pair kernel: 0 (default) the matrix-core pair for 16 < n where it
   applies (FP32 factors n <= 64, FP64 factors n <= 48; nr <= 8, nx <= 16),
   else register kernels for n <= 32, the column-split-waves kernel up to
   n = 64 and the register-tiled kernel beyond; 3 the register-tiled kernel
   for every n; 4 the column-split-waves kernel for every n <= 64; 5 the
   matrix-core pair where it applies (n + nr <= 48 / 64 / 80 for n <= 32 / 48
   / 64; other shapes: the register-tiled kernel) */
int alglin_bcr_set_pair_kernel(alglin_bcr_t h, int variant);

/* ---- precision (extension of the reference surface) -------------------------- */
/* mode "FP64" (default, the reference's arithmetic) or "MIXED" (double handles
   only): FP32 cyclic-reduction factors + FP64 iterative refinement on the GPU
   until ||b - A x||_2 / ||b||_2 <= tol or max_iter correction sweeps */
int alglin_bcr_set_precision(alglin_bcr_t h, char const* mode, double tol, int32_t max_iter);
/* MIXED: the columns whose refinement stalls (or hits max_iter) are solved
   again by FP64 cyclic reduction with column-pivoted QR on the original
   operator, then refined in FP64 (on by default); off = the solve fails with
   "iterative refinement did not reach the requested residual" */
int alglin_bcr_set_fallback(alglin_bcr_t h, int on);
/* last refined solve, worst over the columns: out = {sweeps, rel. residual,
   residual after the first solve, status, componentwise backward error
   max_i |b - A x|_i / (|A||x| + |b|)_i where evaluated (-1: never)}.
   status: 0 converged to tol; 1 stalled (no progress, backward error above
   64 u); 2 max_iter; 3 the FP64 fallback factorization failed; 4 solved to
   tol by the FP64 QRP fallback; 5 stopped at the FP64 residual floor
   (backward error <= 64 u: no FP64 method gets the residual lower); 6 the
   FP64 QRP fallback still missed tol.  0, 4 and 5 count as solved. */
int alglin_bcr_refine_report(alglin_bcr_t h, double out[5]);

/* ---- per-phase GPU timing (HIP events on the handle's stream) --------------- */
/* on=1 records an event pair around every launch group from now on */
int alglin_bcr_set_timing(alglin_bcr_t h, int on);
/* accumulated milliseconds since the last reset and launch counts:
   out[0] pair eliminations  out[1] forward sweeps  out[2] backward sweeps
   out[3] last block (assemble+LU+solve)  out[4] reductions/copies
   out[5] residual SpMV  out[6] launches of pair-elimination kernels
   out[7] launches of sweep kernels; reset=1 clears after reading */
int alglin_bcr_timing(alglin_bcr_t h, double out[8], int reset);

/* ---- allocation-free factor + solve -------------------------------------------
   No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  The same FP64 cyclic reduction (LU eliminations, LU
   last block, pure cyclic-reduction order) and the same kernels as
   alglin_bcr_bind_device + alglin_bcr_factorize_device +
   alglin_bcr_solve_device, with every working array a slice of one fixed
   64 MiB device array of the library (one per GPU, held until the process
   ends): no hipMalloc, no hipFree.  dims = {N, n, qr, qx, nr, nx}
   (qr + nr == qx + nx, n <= 64, nr <= 16); D..F: the operator in the packed
   layout of alglin_bcr_load_packed (device pointers, not modified); X: ld x
   nrhs device block, right-hand sides in, solutions out (ld >= max(nrows,
   ncols) when nrhs > 1); stream: a hipStream_t (0 = the null stream).
   sync = 1 waits and reports the factorization status; sync = 0 only enqueues.
   Returns 0 ok, 1 zero pivot, 2 singular last block, 3 the system's working
   set exceeds the 64 MiB array, -1 bad arguments; err (errlen bytes, may be
   null) receives the message.  Calls on one GPU share the array: calls that
   overlap on one GPU race. */
int alglin_bcr_factor_solve_noalloc(int32_t const dims[6], double const* D, double const* E, double const* B,
                                    double const* C, double const* H, double const* Cq, double const* F, double* X,
                                    int32_t nrhs, int64_t ld, void* stream, int sync, char* err, int32_t errlen);

/* ---- device memory (no other runtime needed) -------------------------------- */
void* alglin_dev_malloc(int64_t bytes);
int alglin_dev_free(void* p);
/* plain copies return once the data is on the device (or host): work queued
   afterwards on any stream, including a solver's stream, sees it */
int alglin_memcpy_htod(void* dst, void const* src, int64_t bytes);
int alglin_memcpy_dtoh(void* dst, void const* src, int64_t bytes);
int alglin_memcpy_dtod(void* dst, void const* src, int64_t bytes);
/* device copy queued on `stream` (e.g. alglin_bcr_stream(h)), ordered with the
   solver's own work on that stream */
int alglin_memcpy_dtod_async(void* dst, void const* src, int64_t bytes, void* stream);
int alglin_device_synchronize(void);
/* device name, compute units, HBM bytes */
int alglin_device_info(char* name, int32_t len, int32_t* cus, int64_t* hbm_bytes);

/* ---- batched systems (extension: no counterpart in the reference) ------------
   Many independent BABD systems of uneven shape solved together: one launch
   per reduction level and kernel class for all systems, the launch sequence
   optionally captured once into a HIP graph and replayed.  Each system is the
   single-system algorithm above (BorderedCR at num_parallel_block = 1) with
   LU last block.  mixed=1: FP32 factors + FP64 refinement per system.

   dims: 7 int32 per system: N n qr qx nr nx method (0 LU, 1 QR, 2 QRP for the
   pair eliminations); qr + nr == qx + nx.  Packed device arrays: kind
   0 = D and E, 1 = B, 2 = C, 3 = H, 4 = Cq, 5 = F, 6 = x (right-hand sides and
   solutions); system s starts at alglin_batch_offset(h, kind, s). */
typedef struct alglin_batch* alglin_batch_t;
alglin_batch_t alglin_batch_create(int mixed);
void alglin_batch_destroy(alglin_batch_t h);
int alglin_batch_plan(alglin_batch_t h, int32_t nsys, int32_t const* dims);
int64_t alglin_batch_size(alglin_batch_t h, int32_t kind);
int64_t alglin_batch_offset(alglin_batch_t h, int32_t kind, int32_t s);
int alglin_batch_bind_device(alglin_batch_t h, double const* D, double const* E, double const* B, double const* C,
                             double const* H, double const* Cq, double const* F);
/* fill the bound arrays with seeded Gaussian systems: D_k = shift I + G/sqrt(n),
   E_k = G/sqrt(n), the other blocks G (counter-based generator on the GPU) */
int alglin_batch_generate(alglin_batch_t h, uint64_t seed, double shift);
/* as above; for system s with 1 <= sing_k[s] < N the pair column block
   [E_{k-1}; D_k] (k = sing_k[s]) is made rank deficient up to eps: its last
   column is the mean of its first two plus eps G/sqrt(n) (n >= 3) */
int alglin_batch_generate_singular(alglin_batch_t h, uint64_t seed, double shift, int32_t nsys,
                                   int32_t const* sing_k, double eps);
/* X = A^{-1} B for every system (device pointers, kind 6 layout); returns 0
   when every system reached ||b - A x||/||b|| <= tol, 2 when some did not
   (see alglin_batch_report), 1 on error */
int alglin_batch_solve(alglin_batch_t h, double const* B, double* X, double tol, int32_t max_iter, int use_graph);
/* per system: status (bit 0 zero pivot, bit 1 singular last block, bit 2 tol
   not reached, bit 3 stopped at the FP64 residual floor: the componentwise
   backward error max_i |b - A x|_i / (|A||x| + |b|)_i is <= 64 u, so no FP64
   method gets the residual lower and the FP64 fallback is not tried),
   relative residual, refinement sweeps; summary = {sweeps, max residual,
   failed systems (status != 0), wall ms of the last solve}; any may be null */
int alglin_batch_report(alglin_batch_t h, int32_t* status, double* rel_residual, int32_t* sweeps, double summary[4]);
char const* alglin_batch_last_error(alglin_batch_t h);
void* alglin_batch_stream(alglin_batch_t h);
/* One eager factorization (FP32 factors for a mixed batch) with HIP events
   around every pair-elimination launch.  Row i of `rows` (6 doubles, at most
   max_rows rows) = level, kernel class, workgroups, milliseconds,
   algorithmic HBM bytes, flops of launch i.  Returns the number of launches
   (may exceed max_rows), -1 on error.  No counterpart in the reference: the
   measurement hook of the batched extension. */
int alglin_batch_time_factorize(alglin_batch_t h, double* rows, int32_t max_rows);

/* Kernels of the 16 < n <= 32 and 32 < n <= 64 LU pair classes: 0 default
   (= 5 where it applies, else 4), 4 the rank-1 kernels (two-wave split for
   n <= 32, column-split waves beyond: every column updated per step), 5 the
   matrix-core pair (FP32 or FP64 factors, n <= 48, nr <= 8, nx <= 16,
   n + nr <= 48 for n <= 32 and <= 64 beyond; other shapes take the rank-1
   kernels).
   Drops captured graphs.  No counterpart in the reference (a tuning hook). */
int alglin_batch_set_pair_kernel(alglin_batch_t h, int variant);

/* 1: every launch of the batch on its stream, one after the other; 2 (the
   default): the 32 < n <= 64 LU class factorized and swept on a second
   stream beside the other classes, from reduction level fork_level on (the
   levels before it one launch after the other).  Drops captured graphs.  No
   counterpart in the reference (a tuning hook). */
int alglin_batch_set_streams(alglin_batch_t h, int streams, int fork_level);

/* Tuned kernel of one LU pair-elimination class (no counterpart in the
   reference: the hook the autotuner, alglin_amd/autotune.py, drives).
   fp32_factors: 1 the FP32-factor engine of a mixed batch, 0 the FP64 one
   (plain batches, and the FP64 fallback of a mixed one).  cls: 3 for
   16 < n <= 32, 4 for 32 < n <= 64.  kernel: -1 default, 0 the matrix-core
   pair (FP32 factors), 1 the rank-1 kernels (two-wave split for class 3,
   column-split for class 4 with `waves` waves per pair: 4, 8 or 16 for
   n <= 48, 8 or 16 beyond, 0 = default), 2 the register-tiled kernel
   (class 4).  Drops captured graphs. */
int alglin_batch_set_class_kernel(alglin_batch_t h, int fp32_factors, int cls, int kernel, int waves);

/* per-system pair elimination method for alglin_batch_solve (0 LU, 1 QR, 2 QRP) */
int alglin_batch_set_methods(alglin_batch_t h, int32_t nsys, int32_t const* methods);

/* ---- learned LU-vs-QRP selector (extension: no counterpart in the reference)
   A 12 -> 64 -> 64 -> 2 ReLU perceptron, FP32, trained and evaluated on the
   GPU.  Features (12 per system) come from the FP32 LU factors and the first
   residual (alglin_batch_selection); class 1 = eliminate with QRP.  Parameter
   vector of alglin_mlp_num_params() floats: W1 (12x64 row-major), b1, W2
   (64x64), b2, W3 (64x2), b3; inputs standardized by (x - mu) / sig. */
typedef struct alglin_mlp* alglin_mlp_t;
int32_t alglin_mlp_num_params(void);
alglin_mlp_t alglin_mlp_create(uint64_t seed);
void alglin_mlp_destroy(alglin_mlp_t h);
char const* alglin_mlp_last_error(alglin_mlp_t h);
int alglin_mlp_get_params(alglin_mlp_t h, float* params, float* mu, float* sig);
int alglin_mlp_set_params(alglin_mlp_t h, float const* params, float const* mu, float const* sig);
/* weights file: 12 mu, 12 sig, then the parameters (float32) */
int alglin_mlp_load(alglin_mlp_t h, char const* path);
int alglin_mlp_save(alglin_mlp_t h, char const* path);
/* mean cross-entropy of n samples (X: n x 12 device floats, y: n device
   int32 labels) and its gradient (host, alglin_mlp_num_params floats) */
int alglin_mlp_loss_grad(alglin_mlp_t h, float const* X, int32_t const* y, int32_t n, float* loss, float* grad);
/* full-batch Adam for `epochs` steps; losses (host, epochs floats) may be null */
int alglin_mlp_train(alglin_mlp_t h, float const* X, int32_t const* y, int32_t n, int32_t epochs, float lr,
                     float* losses);
/* probability of class 1 per sample (device output) */
int alglin_mlp_predict(alglin_mlp_t h, float const* X, int32_t n, float* p1);

/* mixed batch solve with per-system method choice: FP32 LU first; the
   selector (may be null: none) scores every system from that first solve.
   With fallback=1, the systems it scores above threshold go straight to the
   FP64 path (LU, then FP64 QRP for what LU misses), the others are refined
   in FP64 and join that path if they still miss tol.  With fallback=0 the
   flagged systems are refactored with FP32 QRP instead, and every system is
   refined in FP64.  Return codes as alglin_batch_solve. */
int alglin_batch_solve_selected(alglin_batch_t h, double const* B, double* X, double tol, int32_t max_iter,
                                alglin_mlp_t selector, float threshold, int fallback);
/* of the last solve_selected: features (12 per system), systems sent to QRP,
   FP64 fallback stage per system (0 none, 1 FP64 LU, 2 FP64 QRP) (host
   arrays; any may be null) */
int alglin_batch_selection(alglin_batch_t h, float* features, int32_t* chose_qrp, int32_t* fell_back);

#ifdef __cplusplus
}
#endif

#endif /* ALGLIN_AMD_H */

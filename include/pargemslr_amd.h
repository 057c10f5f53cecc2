/*
 * pargemslr_amd: C ABI of the MI355X GeMSLR-preconditioned FGMRES solve path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status:
 * 0 on success, the reference's PARGEMSLR_ERROR_* numbering for the shared
 * conditions (ParGeMSLR/INC/pargemslr.hpp:9-21) and PGS_ERR_* (>= 200) for
 * conditions of this library.  Nothing throws across the ABI.
 *
 * Numbering: values are double (real operators) or interleaved (re, im)
 * double pairs (complex operators); device pointers are hipMalloc'd memory on
 * the device the operator lives on; `stream` is a hipStream_t (NULL = default
 * stream).  Multi-RHS vectors are row-interleaved: entry (row i, column c) of
 * an n x nrhs block is at [i * nrhs + c].
 *
 * Ownership: setup and upload copy the caller's arrays; x, b, v, z are caller
 * owned.  A device operator may be shared by several workspaces; calls that
 * use distinct workspaces may run concurrently on distinct streams.
 */
#ifndef PARGEMSLR_AMD_H
#define PARGEMSLR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (reference values: ParGeMSLR/INC/pargemslr.hpp:9-21) */
#define PGS_SUCCESS 0
#define PGS_ERROR_INVALID_OPTION 100
#define PGS_ERROR_INVALID_PARAM 101
#define PGS_ERROR_IO 102
#define PGS_ERROR_ILU_EMPTY_ROW 103
#define PGS_ERROR_FUNCTION_CALL 106
#define PGS_ERROR_MEMORY_LOCATION 107
#define PGS_ERR_HIP 200
#define PGS_ERR_OOM 201
#define PGS_ERR_NO_DEVICE 202
#define PGS_ERR_BREAKDOWN 203
#define PGS_ERR_NOT_CONVERGED 204
#define PGS_ERR_UNSUPPORTED 205
#define PGS_ERR_NONFINITE 206 /* a column met NaN or Inf and was retired; the other columns are unaffected */
#define PGS_ERR_SCRATCH_LIMIT 207 /* the call needs more scratch memory than pgs_apply_noalloc holds */

/* parameter array layout: the reference's PARGEMSLR_IO_* indices
 * (ParGeMSLR/SRC/utils/utils.hpp:77-297); only the entries listed here are read */
#define PGS_IO_SIZE 180
#define PGS_IO_SOLVER_KDIM 11
#define PGS_IO_SOLVER_MAXITS 12
#define PGS_IO_SOLVER_TOL 13
#define PGS_IO_SOLVER_ATOL 14
#define PGS_IO_PREPOSS_NLEV_GLOBAL 20
#define PGS_IO_PREPOSS_NCOMP_GLOBAL 22
#define PGS_IO_PREPOSS_PARTITION_GLOBAL 24
#define PGS_IO_PREPOSS_MINSEP 26
#define PGS_IO_PREPOSS_KMIN_GLOBAL 27
#define PGS_IO_PREPOSS_KFACTOR_GLOBAL 29
#define PGS_IO_PREPOSS_VTXSEP_GLOBAL 31
#define PGS_IO_ILU_DROPTOL_B_GLOBAL 40
#define PGS_IO_ILU_DROPTOL_C_GLOBAL 44
#define PGS_IO_ILU_ROWNNZ_B_GLOBAL 48
#define PGS_IO_ILU_ROWNNZ_C_GLOBAL 50
#define PGS_IO_ILU_PERM_OPTION_GLOBAL 58
#define PGS_IO_LR_ARNOLDI_OPTION1_GLOBAL 70
#define PGS_IO_LR_ARNOLDI_OPTION2_GLOBAL 71
#define PGS_IO_LR_RANK1_GLOBAL 74
#define PGS_IO_LR_RANK2_GLOBAL 75
#define PGS_IO_LR_RANK_FACTOR1_GLOBAL 78
#define PGS_IO_LR_RANK_FACTOR2_GLOBAL 79
#define PGS_IO_LR_ARNOLDI_FACTOR1_GLOBAL 82
#define PGS_IO_LR_ARNOLDI_FACTOR2_GLOBAL 83
#define PGS_IO_LR_TOL_EIG1_GLOBAL 86
#define PGS_IO_LR_TOL_EIG2_GLOBAL 87
#define PGS_IO_LR_MAXITS1_GLOBAL 90
#define PGS_IO_LR_MAXITS2_GLOBAL 91
#define PGS_IO_LR_TR_FACTOR 94
#define PGS_IO_LR_TOL_ORTH 95
#define PGS_IO_LR_RAND_INIT_GUESS 97
#define PGS_IO_SCHUR_ENABLE 130
#define PGS_IO_SCHUR_ITER_TOL 131
#define PGS_IO_SCHUR_MAXITS 132
#define PGS_IO_ADVANCED_GLOBAL_SOLVE 141 /* SolveLevel variant: 0 = LU (kGemslrLUSolve), 1 = U (kGemslrUSolve) */
/* most inner Schur iterations per apply (PGS_IO_SCHUR_MAXITS) this library takes */
#define PGS_SCHUR_MAXITS_LIMIT 32

/* precision of the preconditioner apply */
#define PGS_PREC_FP64 1 /* apply in the operator's own precision (double / complex128) */
#define PGS_PREC_FP32 2 /* apply in float / complex64 with float accumulation */

/* Krylov orthogonalization */
#define PGS_ORTH_MGS 0  /* modified Gram-Schmidt, as FlexGmresClass (fgmres.hpp:424-426) */
#define PGS_ORTH_CGS2 1 /* classical Gram-Schmidt with one re-orthogonalization pass */
#define PGS_ORTH_DCGS2 2 /* CGS2 with the re-orthogonalization delayed into the next step
                            (two basis sweeps per step instead of three); real bases of any
                            Krylov dimension, complex bases up to 64 vectors (larger ones use
                            PGS_ORTH_CGS2) */

typedef struct pgs_gemslr_s* pgs_gemslr_t; /* host-resident factored operator (setup output) */
typedef struct pgs_dev_op_s* pgs_dev_op_t; /* operator resident in device memory */
typedef struct pgs_ws_s* pgs_ws_t;         /* per-stream solve workspace */

/* ---------------- host setup (GemslrClass::Setup) ---------------- */

/* Fills params[PGS_IO_SIZE] with the reference defaults
 * (PargemslrSetDefaultParameterArray, ParGeMSLR/SRC/utils/utils.cpp:31-127). */
int pgs_params_default(double* params);

/* GemslrClass::SetMatrix + SetWithParameterArray + Setup
 * (ParGeMSLR/SRC/preconditioners/gemslr.hpp:2069-2161, gemslr.cpp:1602-1711).
 * CSR (0-based, int32), `a` holds n values per row entry: double (is_complex=0)
 * or (re, im) pairs (is_complex=1).  seed selects the deterministic random
 * streams of the setup (partition tie-breaking, Arnoldi start vector). */
int pgs_gemslr_setup(int n, const int* ia, const int* ja, const double* a, int is_complex, const double* params,
                     uint64_t seed, pgs_gemslr_t* out);
int pgs_gemslr_destroy(pgs_gemslr_t op);
/* The same setup with the low-rank corrections built on a device (the setup
 * acceleration of the reference's low-rank build, GemslrClass::Setup ->
 * SetupLowRank*, ParGeMSLR/SRC/preconditioners/gemslr.cpp:2782-3257, which runs
 * on the host there): partition, orderings and ILUT on the host, then per
 * level, bottom first, the thick-restart Arnoldi process with its basis, the
 * applies of G (E B^-1 F and C against the device solve of the levels below)
 * and the orthogonalization on `device`, its small Hessenberg / Schur work on
 * the host.  FP64 (complex<double>) arithmetic.  The result is a host operator
 * like pgs_gemslr_setup's (its low-rank bases equal to rounding, not bit for
 * bit); upload it with pgs_dev_op_create.  PGS_ERR_NO_DEVICE without a device. */
int pgs_gemslr_setup_device(int n, const int* ia, const int* ja, const double* a, int is_complex,
                            const double* params, uint64_t seed, int device, pgs_gemslr_t* out);

typedef struct {
  int n;
  int nlev;
  int is_complex;
  int64_t nnz_a;
  int64_t nnz_ilu;    /* sum over blocks of n_b + nnz(L_b) + nnz(U_b) */
  int64_t nnz_lr;     /* sum over levels of 2 * m_l * k_l */
  double setup_seconds[4]; /* partition+ordering, ILUT, low-rank, total */
  int schur_maxits;        /* inner Schur iteration steps in every apply; 0 when off */
  double schur_tol;        /* its relative tolerance */
  int global_solve;        /* SolveLevel variant of the apply: 0 = LU, 1 = U (PGS_IO_ADVANCED_GLOBAL_SOLVE) */
} pgs_gemslr_info_t;
int pgs_gemslr_info(pgs_gemslr_t op, pgs_gemslr_info_t* info);

/* One level of the factored operator, level-local numbering.  The factors of
 * all ncomp diagonal blocks are concatenated into one block-diagonal CSR
 * (L strict lower, U strict upper, Dinv = inverse pivots).  E is nrem x nloc,
 * F is nloc x nrem, W and WH are nrem x rank column-major.  C (nrem x nrem,
 * the lower-right block of P A P^T) is kept on level 0 when the inner Schur
 * iteration is on (PGS_IO_SCHUR_ENABLE), else C_ia is NULL.  Pointers stay
 * valid until pgs_gemslr_destroy. */
typedef struct {
  int lev_start, nloc, nrem, ncomp, rank;
  const int* dom_ptr; /* ncomp+1, level-local */
  const int *L_ia, *L_ja, *U_ia, *U_ja;
  const double *L_a, *U_a, *Dinv;
  const int *E_ia, *E_ja, *F_ia, *F_ja;
  const double *E_a, *F_a;
  const double *W, *WH;
  const int *C_ia, *C_ja;
  const double* C_a;
} pgs_level_view_t;
int pgs_gemslr_level_view(pgs_gemslr_t op, int level, pgs_level_view_t* view);
/* perm[i]: (P b)[i] = b[perm[i]] */
int pgs_gemslr_perm(pgs_gemslr_t op, const int** perm);
/* The matrix the operator was set up on (0-based CSR, n x n; values as in
 * pgs_gemslr_setup).  Valid until pgs_gemslr_destroy. */
int pgs_gemslr_matrix(pgs_gemslr_t op, int* n, const int** ia, const int** ja, const double** a);

/* Operator files: the complete setup output (matrix, permutation, levels,
 * ILUT factors, couplings, low-rank terms, parameters) in one little-endian
 * binary file closed by a 64-bit FNV-1a checksum, so that a restarted process
 * uploads an operator without repeating GemslrClass::Setup.  The reference
 * keeps its operators in memory only: these two calls have no counterpart
 * there.  save writes path + ".tmp" and renames it to path.  load returns
 * PGS_ERROR_IO for a missing, truncated, corrupted or structurally invalid
 * file (CSR shapes and index ranges, strictly triangular factors, a true
 * permutation, level and block offsets are all checked), else a handle that
 * behaves exactly like the one that was saved (same factors, bit for bit). */
int pgs_gemslr_save(pgs_gemslr_t op, const char* path);
int pgs_gemslr_load(const char* path, pgs_gemslr_t* out);

/* Host-resident apply z = M^{-1} v in FP64 (GemslrClass::Solve on kMemoryHost,
 * gemslr.cpp:3431-3467), with the inner Schur iteration on level 0 when it is
 * on (gemslr.cpp:3556-3563). */
int pgs_gemslr_host_apply(pgs_gemslr_t op, const double* v, double* z);

/* ---------------- device ---------------- */

int pgs_device_count(int* count);

/* Device memory and streams of the library's own HIP runtime, for callers
 * that do not hold one (tests, benchmarks, the request server).  No
 * counterpart in the reference, whose memory layer is
 * ParGeMSLR/SRC/utils/memory.cpp:93-503. */
int pgs_dev_malloc(int device, int64_t bytes, void** ptr);
int pgs_dev_free(int device, void* ptr);
int pgs_memcpy_htod(int device, void* dst, const void* src, int64_t bytes, void* stream);
int pgs_memcpy_dtoh(int device, void* dst, const void* src, int64_t bytes, void* stream);
int pgs_memcpy_dtod(int device, void* dst, const void* src, int64_t bytes, void* stream);
int pgs_memset_zero(int device, void* ptr, int64_t bytes, void* stream);
int pgs_stream_create(int device, void** stream);
int pgs_stream_destroy(int device, void* stream);
int pgs_stream_synchronize(int device, void* stream);
int pgs_device_synchronize(int device);

/* ---------------- device vectors and matrices of the class surface ---------------- */

/* Single-vector operations on device arrays of n values (double, or (re, im)
 * pairs with is_complex = 1); alpha / value are one double or one pair.
 * SequentialVectorClass::Axpy / Scale / Fill / Dot / Norm2
 * (ParGeMSLR/SRC/vectors/sequential_vector.cpp:706-732, 1037-1172 over
 * vectorops.cpp:93-146, 633-680): y += alpha x; x *= alpha; x = value;
 * result = sum_i conj(x_i) y_i (one double or one pair, written to host
 * memory after the stream has drained; summed in a fixed order, so repeated
 * calls give the same bits). */
int pgs_vec_axpy(int device, int64_t n, int is_complex, const double* alpha, const double* x, double* y,
                 void* stream);
int pgs_vec_scale(int device, int64_t n, int is_complex, const double* alpha, double* x, void* stream);
int pgs_vec_fill(int device, int64_t n, int is_complex, const double* value, double* x, void* stream);
int pgs_vec_dot(int device, int64_t n, int is_complex, const double* x, const double* y, double* result,
                void* stream);

/* A square CSR matrix resident on the device by itself (CsrMatrixClass::
 * MoveData(kMemoryDevice), csr_matrix.cpp), for MatVec without a
 * preconditioner: y = alpha A x + beta y with the SpMV kernels of the solve
 * (CsrMatrixClass::MatVec, csr_matrix.cpp:1340-1365).  ia (n + 1, from 0),
 * ja (0-based columns < n) and a (double or (re, im) pairs) are copied. */
typedef struct pgs_dev_csr_s* pgs_dev_csr_t;
int pgs_dev_csr_create(int device, int n, const int* ia, const int* ja, const double* a, int is_complex,
                       pgs_dev_csr_t* out);
int pgs_dev_csr_destroy(pgs_dev_csr_t m);
int pgs_dev_csr_spmv(pgs_dev_csr_t m, const double* alpha, const double* x, const double* beta, double* y,
                     void* stream);

/* Uploads the operator once (A, permutation, level factors in level-set
 * order, E, F, W, WH) for the precisions in prec_mask (PGS_PREC_* bits). */
int pgs_dev_op_create(pgs_gemslr_t op, int device, int prec_mask, pgs_dev_op_t* out);
int pgs_dev_op_destroy(pgs_dev_op_t dop);
/* device bytes held by the operator */
int pgs_dev_op_bytes(pgs_dev_op_t dop, int64_t* bytes);
/* order n, device and value type (is_complex: 1 for complex128 operators,
 * whose vectors are (re, im) double pairs) of the operator */
int pgs_dev_op_dims(pgs_dev_op_t dop, int* n, int* device, int* is_complex);

/* Workspace for up to max_nrhs right-hand sides and Krylov dimension kdim. */
int pgs_ws_create(pgs_dev_op_t dop, int max_nrhs, int kdim, pgs_ws_t* out);
int pgs_ws_destroy(pgs_ws_t ws);
/* device bytes pgs_ws_create would allocate for these arguments */
int pgs_ws_estimate(pgs_dev_op_t dop, int max_nrhs, int kdim, int64_t* bytes);

/* y = alpha * A * x + beta * y  (CsrMatrixClass::MatVec('N', ...),
 * ParGeMSLR/SRC/matrices/csr_matrix.cpp:1340-1365), FP64 / complex128. */
int pgs_spmv(pgs_dev_op_t dop, pgs_ws_t ws, const double* alpha, const double* x, const double* beta, double* y,
             int nrhs, void* stream);

/* z = M^{-1} v on the device (GemslrClass::Solve, gemslr.cpp:3431-3467).
 * v and z are in the operator's precision (double / complex128); `prec`
 * selects the arithmetic of the apply (PGS_PREC_FP64 or PGS_PREC_FP32). */
int pgs_apply(pgs_dev_op_t dop, pgs_ws_t ws, const double* v, double* z, int nrhs, int prec, void* stream);

/* z = M^{-1} v as pgs_apply, without allocating device memory and without a
 * workspace.  No counterpart in the reference: for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  The apply's buffers (permuted copies, level
 * temporaries, low-rank partial sums) live in a file-scope device array of
 * PGS_APPLY_NOALLOC_SCRATCH bytes that each device the library runs on holds
 * until the process ends; a call whose buffers do not fit returns
 * PGS_ERR_SCRATCH_LIMIT (about n * nrhs * 3 values of the operator's type: at
 * most 2^21 unknowns at one double column).  Operators with the inner Schur
 * iteration return PGS_ERR_UNSUPPORTED.  Every call on a device shares its
 * array: calls that overlap on one device (other streams or threads) race;
 * serialize them.  Results equal pgs_apply's bit for bit. */
#define PGS_APPLY_NOALLOC_SCRATCH (64 * 1024 * 1024)
int pgs_apply_noalloc(pgs_dev_op_t dop, const double* v, double* z, int nrhs, int prec, void* stream);

typedef struct {
  int kdim;         /* restart length m */
  int maxits;       /* total iteration cap */
  double tol;       /* stop when ||r|| <= tol * ||b|| (or <= tol if absolute) */
  int absolute_tol;
  int prec;         /* PGS_PREC_* of the preconditioner apply */
  int orth;         /* PGS_ORTH_* */
  int use_precond;  /* 0: no preconditioner (z = v) */
  int check_every;  /* host convergence poll interval in iterations (0 = 1) */
  int use_graph;    /* 1: each inner step j of a cycle is captured once as a HIP graph
                       (per workspace and settings) and replayed; results are the
                       same as with direct launches */
} pgs_fgmres_opts_t;

typedef struct {
  int status;       /* 0 converged, PGS_ERR_NOT_CONVERGED, PGS_ERR_BREAKDOWN, PGS_ERR_NONFINITE
                       (NaN or Inf in b, x0 or the iteration: the column stops at once,
                       x keeps its value from the start of the cycle) */
  int iters;
  double rel_res;   /* Givens estimate ||r||/||b|| at exit (FlexGmresClass::_rel_res) */
} pgs_fgmres_result_t;

/* Flexible GMRES(m), right preconditioned (FlexGmresClass::Solve,
 * ParGeMSLR/SRC/solvers/fgmres.hpp:263-558).  x holds the initial guess on
 * entry.  For nrhs > 1 every column runs its own FGMRES (own Hessenberg,
 * rotations, stopping test); kernels process all columns per launch.
 * res must hold nrhs entries. */
int pgs_fgmres(pgs_dev_op_t dop, pgs_ws_t ws, const double* b, double* x, int nrhs, const pgs_fgmres_opts_t* opts,
               pgs_fgmres_result_t* res, void* stream);

/* Kernel timing for the roofline report (no counterpart in the reference).
 * When enabled, the workspace records a HIP event pair (on the launch stream)
 * around every launch of the solve path it issues and accumulates, per kernel
 * family, the launches, their time and their algorithmic bytes (s = value
 * bytes of the launch, vb = n * nrhs * s of the Krylov scalar):
 *   PGS_PROF_TRSV0 / PGS_PROF_TRSV   block triangular solves of level 0 / the levels below:
 *       (nnz(L) + nnz(U)) * (s + 4) + 2 * (n_l + 1) * 4 + n_l * s + 2 * n_l * nrhs * s
 *   PGS_PROF_SPMV_A / PGS_PROF_SPMV_EF   y = A x (+ r) / the coupling blocks E, F:
 *       nnz * (s + 4) + (rows + 1) * 4 + cols * nrhs * s_x + rows * nrhs * s_y (x2 when y or r is read)
 *   PGS_PROF_LOWRANK    (2 k + 3 nrhs) * m * s
 *   PGS_PROF_PERM       n * nrhs * (s_in + s_out) + 4 n (gather or scatter)
 *   PGS_PROF_ORTH_DOTS / PGS_PROF_ORTH_UPDATE   basis sweeps: (vectors read + written) * vb
 *   PGS_PROF_SCALE      2 vb;  PGS_PROF_UPDATE_X  (i * s_z / s + 2) * vb
 *   PGS_PROF_STATE      the one-workgroup Givens / Hessenberg kernels (0 bytes)
 * Timed solves launch directly (no graph replay).  Reading synchronizes the
 * recorded events. */
#define PGS_PROF_TRSV0 0
#define PGS_PROF_TRSV 1
#define PGS_PROF_SPMV_A 2
#define PGS_PROF_SPMV_EF 3
#define PGS_PROF_LOWRANK 4
#define PGS_PROF_PERM 5
#define PGS_PROF_ORTH_DOTS 6
#define PGS_PROF_ORTH_UPDATE 7
#define PGS_PROF_SCALE 8
#define PGS_PROF_UPDATE_X 9
#define PGS_PROF_STATE 10
#define PGS_PROFILE_FAMILIES 11
typedef struct {
  int64_t launches;
  double total_ms;
  double total_bytes;
} pgs_kernel_profile_t;
int pgs_ws_profile(pgs_ws_t ws, int enable);
/* the level-0 block triangular solve (family PGS_PROF_TRSV0) */
int pgs_ws_profile_read(pgs_ws_t ws, pgs_kernel_profile_t* out);
/* every family: out[PGS_PROFILE_FAMILIES], indexed by PGS_PROF_* */
int pgs_ws_profile_families(pgs_ws_t ws, pgs_kernel_profile_t* out);

const char* pgs_strerror(int code);
/* message of the last HIP error seen by this thread */
const char* pgs_last_hip_error(void);

/* ---------------- request server (continuous batching) ----------------
 * No counterpart in the reference (one right-hand side per Solve call).
 * Several uploaded operators stay resident; requests of 1..max_batch
 * right-hand sides are queued in arrival order.  Each of `slots` workers (one
 * stream and one host thread each) takes the oldest request plus every later
 * queued request for the same operator that fits in max_batch columns and
 * runs them as one batch of columns, one FGMRES restart cycle (pgs_fgmres,
 * at most kdim iterations) at a time: at each restart, columns that have
 * converged or failed leave (a request completes when all of its columns
 * have), and queued requests for the same operator take the free columns.
 * Work of several batches overlaps on the device.  The iteration cap maxits
 * is applied at restart boundaries.  A worker keeps one workspace per
 * operator it has served; it admits a new one only when the device's free
 * memory minus mem_reserve holds it, dropping its own idle workspaces and
 * then waiting for other batches if needed. */
typedef struct pgs_server_s* pgs_server_t;
typedef struct {
  int max_batch;       /* most right-hand sides per batch (power of two <= 64) */
  int slots;           /* batches in flight at once (1..16) */
  int64_t mem_reserve; /* device bytes left free when admitting a workspace */
  int kdim, maxits;    /* FGMRES(kdim), iteration cap per column */
  double tol;          /* relative tolerance on ||b - A x|| / ||b|| */
  int prec, orth;      /* PGS_PREC_* of the apply, PGS_ORTH_* */
  int check_every;     /* as in pgs_fgmres_opts_t */
  int use_graph;       /* as in pgs_fgmres_opts_t */
} pgs_server_opts_t;
typedef struct {
  int status;          /* 0 when every column converged, else the first column's error code */
  int nrhs;            /* columns of the request */
  int batch_nrhs;      /* columns of the batch it was solved in */
  double queue_ms;     /* submit to batch start */
  double solve_ms;     /* batch start to results written */
  double total_ms;     /* submit to results written */
  int joined_cycle;    /* restart cycle of its batch at which it joined (0: the batch started with it) */
} pgs_request_info_t;
int pgs_server_create(int device, const pgs_server_opts_t* opts, pgs_server_t* out);
/* the operator must outlive the server */
int pgs_server_add_operator(pgs_server_t srv, pgs_dev_op_t dop, int* op_id);
/* b, x: host arrays n x nrhs, column c at [c * n] (complex operators: (re, im)
 * double pairs, column c at [2 * c * n]); iters: nrhs ints or NULL.
 * All three stay owned by the caller and must stay valid until
 * pgs_server_wait returns for the request; x and iters are written then. */
int pgs_server_submit(pgs_server_t srv, int op_id, int nrhs, const double* b, double* x, int* iters, int64_t* req);
/* blocks until the request is done; a request can be waited for once */
int pgs_server_wait(pgs_server_t srv, int64_t req, pgs_request_info_t* info);
/* finishes every queued request, then stops the workers */
int pgs_server_destroy(pgs_server_t srv);

/* ---------------- problem generators (test and benchmark inputs) ---------------- */

/* CooMatrixClass::Laplacian (ParGeMSLR/SRC/matrices/coo_matrix.cpp:519-640):
 * 5-point (one of nx, ny, nz == 1) or 7-point stencil, Dirichlet boundary,
 * diagonal (2D ? 4 : 6) - shift, +direction neighbour -1 - alpha, -direction
 * neighbour -1 + alpha.  is_complex selects complex values (shift_im used).
 * Arrays are malloc'd; release with pgs_free. */
int pgs_gen_laplacian(int nx, int ny, int nz, double shift_re, double shift_im, double ax, double ay, double az,
                      int is_complex, int* n, int64_t* nnz, int** ia, int** ja, double** a);
/* 3D first-order upwind convection-diffusion: diagonal 6 + 3c, upstream
 * (-x, -y, -z) neighbours -1 - c, downstream neighbours -1. */
int pgs_gen_convdiff(int nx, int ny, int nz, double c, int* n, int64_t* nnz, int** ia, int** ja, double** a);
void pgs_free(void* p);

/* ---------------- parameter files ---------------- */

#define PGS_INPUTS_SEQUENTIAL 0 /* the sequential drivers' 38-line `inputs` */
#define PGS_INPUTS_HYPRE 1      /* the hypre interface's 31-line `inputs` */
/* Reads a parameter file of the reference's drivers into params (a
 * PGS_IO_SIZE array, typically pgs_params_default's) line by line:
 * read_inputs_from_file of ParGeMSLR/TESTS/sequential/io.cpp:336-700
 * (PGS_INPUTS_SEQUENTIAL) or of ParGeMSLR/TESTS/hypre/
 * pargemslr_hypre_interface.cpp:391-820 (PGS_INPUTS_HYPRE), with the same
 * keywords, line order and GLOBAL-to-LOCAL copies.  Returns PGS_ERROR_IO for a
 * missing file or an unknown keyword (the lines before it stay applied).
 * Options the setup does not implement (B solves other than ILUT, ND
 * partitions, AMD/ND local orderings, global solve options other than the
 * LU and U variants of the level solve) are read faithfully; pgs_gemslr_setup then refuses them with
 * PGS_ERROR_INVALID_OPTION. */
int pgs_params_read_inputs(const char* path, int format, double* params);

/* ---------------- hypre-style preconditioner object (one process) ---------------- */

/* The life cycle of the reference's hypre interface
 * (ParGeMSLR/TESTS/hypre/HYPRE_pargemslr.hpp:23-58: HYPRE_GEMSLRCreate /
 * CreateFromFile, Setup, Solve, Destroy) without hypre's types.
 * create reads inputs_file (NULL: "inputs" in the working directory) in the
 * hypre layout and, like HYPRE_GEMSLRCreateFromFile, keeps the defaults for
 * what it cannot read.  setup takes the diagonal block of the ParCSR matrix
 * (hypre_CSRMatrixI/J/Data of hypre_ParCSRMatrixDiag; at one process it is
 * the whole matrix), runs the GeMSLR setup with those parameters and uploads
 * the operator to `device`.  solve applies the preconditioner once, x = M^{-1} b,
 * on host arrays of n values (hypre_VectorData of the local vectors): what
 * hypre's Krylov solvers call HYPRE_GEMSLRSolve for.  The apply is FP64 unless
 * set_precision selects PGS_PREC_FP32.  At one process the operator is the
 * sequential GemslrClass's (RKway levels), not ParallelGemslrClass's. */
typedef struct pgs_hypre_gemslr_s* pgs_hypre_gemslr_t;
int pgs_hypre_gemslr_create(const char* inputs_file, pgs_hypre_gemslr_t* out);
int pgs_hypre_gemslr_params(pgs_hypre_gemslr_t s, double* params);
int pgs_hypre_gemslr_set_precision(pgs_hypre_gemslr_t s, int prec);
int pgs_hypre_gemslr_setup(pgs_hypre_gemslr_t s, int n, const int* ia, const int* ja, const double* a, int device);
int pgs_hypre_gemslr_solve(pgs_hypre_gemslr_t s, const double* b, double* x);
int pgs_hypre_gemslr_destroy(pgs_hypre_gemslr_t s);

/* ---------------- MatrixMarket ingest ---------------- */

/* CsrMatrixClass::ReadFromMMFile (ParGeMSLR/SRC/matrices/csr_matrix.cpp:
 * 1706-1714 via CooMatrixReadFromFile, matrixops.cpp:11134-11453, and
 * CooMatrixClass::ToCsr, coo_matrix.cpp:479-512): a "matrix coordinate" file
 * with indices based at idxin (1 for standard files, 0 for the reference's
 * "zerobase" option) into a CSR with rows in file order.  is_complex = 0 reads
 * like the reference's double reader (real and integer data; off-diagonal
 * entries mirrored for symmetric storage only); is_complex = 1 like its
 * complexd reader (also complex data; symmetric, hermitian and skew-symmetric
 * storage mirrored) and returns (re, im) pairs.  Returns PGS_ERROR_IO for a
 * missing file, an unsupported type, a short entry list or an index outside
 * the declared size.  Arrays are malloc'd; release with pgs_free. */
int pgs_read_mm_csr(const char* path, int idxin, int is_complex, int* nrows, int* ncols, int64_t* nnz, int** ia,
                    int** ja, double** a);
/* SequentialVectorClass::ReadFromMMFile (ParGeMSLR/SRC/vectors/
 * sequential_vector.cpp:1192-1195 via SequentialVectorReadFromFile,
 * vectorops.cpp:1116-1336): a coordinate file of M rows into a dense vector
 * of length n = M, zero where the file has no entry; entry (I, J) sets v[I].
 * The complex reader requires one column.  Release v with pgs_free. */
int pgs_read_mm_vector(const char* path, int idxin, int is_complex, int* n, double** v);
/* counter-based splitmix64 streams: uniform [lo, hi) and standard normal */
int pgs_rng_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
int pgs_rng_normal(uint64_t seed, int64_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif /* PARGEMSLR_AMD_H */

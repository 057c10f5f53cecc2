/*
 * Oracle: a plain-C, FP64 / complex128, single-threaded restatement of the
 * reference's GeMSLR-FGMRES solve path.  TEST INFRASTRUCTURE ONLY: it is the checker the
 * tests, the smoke check and the benchmark's CPU-baseline leg compare against;
 * the product solve path never calls it.
 *
 * Each function follows the reference line ranges cited next to it
 * (paths relative to ParGeMSLR/).  Inputs are the factored operator produced
 * by the library's setup (exported level by level), so the oracle checks the
 * solve phase: SpMV, ILU triangular solves, the multilevel recursion with the
 * low-rank correction, permutations, MGS / CGS2 and FGMRES with Givens.
 *
 * The functions are written once (oracle_impl.h) and compiled for double
 * (orc_*) and for double complex (orc_*_z), as the reference instantiates its
 * templates for both.
 *
 * How each function is pinned (tests/test_oracle.py; the complex versions by
 * the same checks on complex operators, plus the real/complex consistency of
 * a complex operator with zero imaginary parts):
 *   orc_csr_matvec     closed form of the Laplacian row sums, scipy.sparse
 *   orc_ilu_solve      exact-LU special case (drop 0, unlimited fill) vs
 *                      numpy.linalg.solve; residual L(D+U)x = rhs otherwise
 *   orc_lowrank_apply  dense numpy W (Hk (W^H z)) with the setup's own Hk
 *   orc_gemslr_apply   dense block formula on tiny problems; exact ILU +
 *                      full-rank correction gives A^{-1} (closed form)
 *   orc_mgs / orc_cgs2 Q^T Q = I, V H = A V relations vs numpy
 *   orc_fgmres         unpreconditioned residual history vs a numpy GMRES
 *                      (unique in exact arithmetic); exact preconditioner
 *                      converges in one iteration; Givens LSQ vs lstsq
 *   inner Schur        (orc_gemslr_apply with inner_maxits > 0) exact ILU +
 *   iteration          enough inner steps gives A^{-1} even without a low-rank
 *                      term; one and K inner steps equal a dense numpy GMRES
 *                      on S = C - E B^{-1} F with the preconditioner of the
 *                      levels below (tests/test_oracle_schur.py)
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EPS 2.220446049250313e-16

/* value arrays hold double or double complex, as the operator */
typedef struct {
  int lev_start, nloc, nrem, rank;
  const int *L_ia, *L_ja, *U_ia, *U_ja;
  const void *L_a, *U_a, *Dinv;
  const int *E_ia, *E_ja, *F_ia, *F_ja;
  const void *E_a, *F_a;
  const void *W, *WH; /* nrem x rank, column-major */
  const int *C_ia, *C_ja; /* level 0 with the inner iteration: nrem x nrem lower-right block */
  const void* C_a;
  int ncomp;             /* diagonal blocks of the level (0: solve the level as one block) */
  const int* dom_ptr;    /* ncomp + 1 block row offsets within the level */
} orc_level;

typedef struct {
  int n, nlev;
  const int* perm;
  const orc_level* levels;
  int inner_maxits;  /* > 0: inner FGMRES on the level-0 Schur complement */
  double inner_tol;
  int usolve;        /* 1: the U variant of SolveLevel (no block L solve on the way down) */
} orc_gemslr;

/* threads over which a level's diagonal blocks are solved (the CPU
 * baseline's OpenMP-over-blocks variant); 1 = serial.  Every row is computed
 * in the same order either way, so the results are the same bits. */
static int orc_threads = 1;
void orc_set_threads(int t) { orc_threads = t < 1 ? 1 : t; }

/* workspace size (values) of orc_gemslr_apply / orc_gemslr_apply_z */
long orc_gemslr_work_size(const orc_gemslr* op) {
  long s = 2L * op->n;
  int l;
  for (l = 0; l < op->nlev; ++l) {
    const orc_level* L = &op->levels[l];
    s += (long)L->nloc + 2L * L->nrem + (L->rank > 0 ? L->rank : 0);
  }
  return s + 16;
}

#define ORC_T double
#define ORC_FN(x) x
#define ORC_CONJ(v) (v)
#define ORC_ABS(v) fabs(v)
#include "oracle_impl.h"
#undef ORC_T
#undef ORC_FN
#undef ORC_CONJ
#undef ORC_ABS

#define ORC_T double complex
#define ORC_FN(x) x##_z
#define ORC_CONJ(v) conj(v)
#define ORC_ABS(v) cabs(v)
#include "oracle_impl.h"
#undef ORC_T
#undef ORC_FN
#undef ORC_CONJ
#undef ORC_ABS

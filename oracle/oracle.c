/*
 * Oracle: a plain-C, FP64, single-threaded restatement of the reference's
 * GeMSLR-FGMRES solve path.  TEST INFRASTRUCTURE ONLY: it is the checker the
 * tests, the smoke check and the benchmark's CPU-baseline leg compare against;
 * the product solve path never calls it.
 *
 * Each function follows the reference line ranges cited next to it
 * (paths relative to ParGeMSLR/).  Inputs are the factored operator produced
 * by the library's setup (exported level by level), so the oracle checks the
 * solve phase: SpMV, ILU triangular solves, the multilevel recursion with the
 * low-rank correction, permutations, MGS / CGS2 and FGMRES with Givens.
 *
 * How each function is pinned (tests/test_oracle.py):
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
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EPS 2.220446049250313e-16

typedef struct {
  int lev_start, nloc, nrem, rank;
  const int *L_ia, *L_ja, *U_ia, *U_ja;
  const double *L_a, *U_a, *Dinv;
  const int *E_ia, *E_ja, *F_ia, *F_ja;
  const double *E_a, *F_a;
  const double *W, *WH; /* nrem x rank, column-major */
} orc_level;

typedef struct {
  int n, nlev;
  const int* perm;
  const orc_level* levels;
} orc_gemslr;

/* y = beta*y; y += alpha*A*x  (CsrMatrixPMatVecHostTemplate, 'N',
 * SRC/matrices/matrixops.cpp:5139-5330; x != y assumed) */
void orc_csr_matvec(int nrows, const int* ia, const int* ja, const double* a, double alpha, const double* x,
                    double beta, double* y) {
  int i, j;
  for (i = 0; i < nrows; ++i) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
  if (alpha == 0.0) return;
  for (i = 0; i < nrows; ++i) {
    double r = 0.0;
    for (j = ia[i]; j < ia[i + 1]; ++j) r += a[j] * x[ja[j]];
    y[i] += alpha * r;
  }
}

/* x = U^{-1} L^{-1} rhs, L unit lower (strict part stored), U strictly upper,
 * D stored inverted (IluClass::SolveHost, no-permutation branch,
 * SRC/preconditioners/ilu.cpp:3752-3783) */
void orc_ilu_solve(int n, const int* L_ia, const int* L_ja, const double* L_a, const int* U_ia, const int* U_ja,
                   const double* U_a, const double* Dinv, const double* rhs, double* x) {
  int i, j;
  for (i = 0; i < n; ++i) x[i] = rhs[i];
  for (i = 0; i < n; ++i)
    for (j = L_ia[i]; j < L_ia[i + 1]; ++j) x[i] -= L_a[j] * x[L_ja[j]];
  for (i = n - 1; i >= 0; --i) {
    for (j = U_ia[i]; j < U_ia[i + 1]; ++j) x[i] -= U_a[j] * x[U_ja[j]];
    x[i] *= Dinv[i];
  }
}

/* x = WH * (W^H * z): DenseMatrixPMatVecTemplate 'C' then 'N'
 * (SRC/matrices/matrixops.cpp:101-332), as GemslrClass::SolveApplyLowRankLevel
 * (SRC/preconditioners/gemslr.cpp:3681-3719).  t: workspace of `rank`. */
void orc_lowrank_apply(int m, int rank, const double* W, const double* WH, const double* z, double* x, double* t) {
  int i, j;
  for (j = 0; j < rank; ++j) {
    double s = 0.0;
    for (i = 0; i < m; ++i) s += W[(size_t)j * m + i] * z[i];
    t[j] = s;
  }
  for (i = 0; i < m; ++i) x[i] = 0.0;
  for (j = 0; j < rank; ++j)
    for (i = 0; i < m; ++i) x[i] += WH[(size_t)j * m + i] * t[j];
}

/* GemslrClass::SolveLevel (SRC/preconditioners/gemslr.cpp:3474-3625) without
 * the inner Schur iteration.  rhs, x: length n - lev_start.  work: per level
 * buffers carved from `work` (z of length n - lev_start, xlr of length nrem,
 * t of length rank). */
static void solve_level(const orc_gemslr* op, int l, const double* rhs, double* x, double* work) {
  const orc_level* L = &op->levels[l];
  const int nl = L->nloc, m = L->nrem;
  double *z, *zl, *xlr, *t;
  int i;
  if (l == op->nlev - 1) {
    orc_ilu_solve(nl, L->L_ia, L->L_ja, L->L_a, L->U_ia, L->U_ja, L->U_a, L->Dinv, rhs, x);
    return;
  }
  z = work;
  zl = z + nl;
  xlr = z + nl + m;
  t = xlr + m;
  /* step 1: z_u = B^{-1} rhs_u (block-diagonal ILU, blocks are independent) */
  orc_ilu_solve(nl, L->L_ia, L->L_ja, L->L_a, L->U_ia, L->U_ja, L->U_a, L->Dinv, rhs, z);
  /* steps 2-3: z_l = -E z_u; z_l += rhs_l */
  orc_csr_matvec(m, L->E_ia, L->E_ja, L->E_a, -1.0, z, 0.0, zl);
  for (i = 0; i < m; ++i) zl[i] += rhs[nl + i];
  /* step 4: low-rank correction, then the next level */
  if (L->rank > 0) {
    orc_lowrank_apply(m, L->rank, L->W, L->WH, zl, xlr, t);
    for (i = 0; i < m; ++i) zl[i] += xlr[i];
  }
  solve_level(op, l + 1, zl, x + nl, t + (L->rank > 0 ? L->rank : 0));
  /* steps 5-6: z_u = -F x_l; z_u += rhs_u */
  orc_csr_matvec(nl, L->F_ia, L->F_ja, L->F_a, -1.0, x + nl, 0.0, z);
  for (i = 0; i < nl; ++i) z[i] += rhs[i];
  /* step 7: x_u = B^{-1} z_u */
  orc_ilu_solve(nl, L->L_ia, L->L_ja, L->L_a, L->U_ia, L->U_ja, L->U_a, L->Dinv, z, x);
}

/* workspace size (doubles) of orc_gemslr_apply */
long orc_gemslr_work_size(const orc_gemslr* op) {
  long s = 2L * op->n;
  int l;
  for (l = 0; l < op->nlev; ++l) {
    const orc_level* L = &op->levels[l];
    s += (long)L->nloc + 2L * L->nrem + (L->rank > 0 ? L->rank : 0);
  }
  return s + 16;
}

/* z = M^{-1} v: GatherPerm, SolveLevel(0), ScatterRperm
 * (GemslrClass::Solve, SRC/preconditioners/gemslr.cpp:3431-3467;
 *  IntVectorClass::GatherPerm/ScatterRperm, SRC/vectors/int_vector.cpp:629-696) */
void orc_gemslr_apply(const orc_gemslr* op, const double* v, double* z, double* work) {
  const int n = op->n;
  double* r = work;
  double* x = work + n;
  int i;
  for (i = 0; i < n; ++i) r[i] = v[op->perm[i]];
  solve_level(op, 0, r, x, work + 2 * n);
  for (i = 0; i < n; ++i) z[op->perm[i]] = x[i];
}

static double dot(int n, const double* x, const double* y) {
  double s = 0.0;
  int i;
  for (i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

/* PargemslrMgs without re-orthogonalization (tol_reorth < 0, as FGMRES calls
 * it: SRC/solvers/fgmres.hpp:426; SRC/matrices/matrixops.cpp:10728-10807).
 * V column-major n x (k+2) with ld n; w = V(:,k+1).  H(j,k) at H[j + k*ldh]. */
void orc_mgs(int n, double* V, int k, double* H, int ldh, double* t_out) {
  double* w = V + (size_t)(k + 1) * n;
  int i, r;
  double t;
  for (i = 0; i <= k; ++i) {
    const double* v = V + (size_t)i * n;
    double t1 = dot(n, v, w);
    H[i + (size_t)k * ldh] = t1;
    for (r = 0; r < n; ++r) w[r] -= t1 * v[r];
  }
  t = sqrt(fabs(dot(n, w, w)));
  H[(k + 1) + (size_t)k * ldh] = t;
  for (r = 0; r < n; ++r) w[r] *= 1.0 / t;
  *t_out = t;
}

/* PargemslrCgs2 (SRC/matrices/matrixops.cpp:10588-10709). */
void orc_cgs2(int n, double* V, int k, double* H, int ldh, double* t_out) {
  double* w = V + (size_t)(k + 1) * n;
  double* s = (double*)malloc(sizeof(double) * (k + 1));
  int i, r;
  double t;
  for (i = 0; i <= k; ++i) H[i + (size_t)k * ldh] = dot(n, V + (size_t)i * n, w);
  for (i = 0; i <= k; ++i) {
    const double* v = V + (size_t)i * n;
    double h = H[i + (size_t)k * ldh];
    for (r = 0; r < n; ++r) w[r] += -h * v[r];
  }
  for (i = 0; i <= k; ++i) s[i] = dot(n, V + (size_t)i * n, w);
  for (i = 0; i <= k; ++i) {
    const double* v = V + (size_t)i * n;
    for (r = 0; r < n; ++r) w[r] += -s[i] * v[r];
    H[i + (size_t)k * ldh] += s[i];
  }
  t = sqrt(fabs(dot(n, w, w)));
  H[(k + 1) + (size_t)k * ldh] = t;
  for (r = 0; r < n; ++r) w[r] *= 1.0 / t;
  free(s);
  *t_out = t;
}

/* FlexGmresClass::Solve (SRC/solvers/fgmres.hpp:263-558), real FP64.
 * prec == NULL: no preconditioner.  orth: 0 MGS, 1 CGS2.
 * hist (optional, maxits+1 entries): relative residual estimates.
 * Returns 0 converged, 1 not converged, 2 breakdown. */
int orc_fgmres(int n, const int* ia, const int* ja, const double* a, const orc_gemslr* prec, const double* b,
               double* x, int kdim, int maxits, double tol, int absolute, int orth, int* iters_out,
               double* rel_res_out, double* hist) {
  const int m = kdim;
  double *V, *Z, *H, *c, *s, *rs, *work = NULL;
  double normb, normr, tolr, t, gam, hii, hii1, inormr;
  int i = 0, j, k, iter = 0, status = 1, r;
  V = (double*)calloc((size_t)n * (m + 1), sizeof(double));
  Z = (double*)calloc((size_t)n * (m + 1), sizeof(double));
  H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
  c = (double*)calloc(m, sizeof(double));
  s = (double*)calloc(m, sizeof(double));
  rs = (double*)calloc(m + 1, sizeof(double));
  if (prec) work = (double*)malloc(sizeof(double) * orc_gemslr_work_size(prec));
#define Hm(ii, jj) H[(ii) + (size_t)(jj) * (m + 1)]
  normb = sqrt(fabs(dot(n, b, b)));
  if (normb < ORC_EPS) {
    for (r = 0; r < n; ++r) x[r] = 0.0;
    *iters_out = 0;
    *rel_res_out = 0.0;
    if (hist) hist[0] = 0.0;
    status = 0;
    goto done;
  }
  /* v = b - A x */
  memcpy(V, b, sizeof(double) * n);
  orc_csr_matvec(n, ia, ja, a, -1.0, x, 1.0, V);
  normr = sqrt(fabs(dot(n, V, V)));
  if (normr < ORC_EPS) {
    *iters_out = 0;
    *rel_res_out = 0.0;
    if (hist) hist[0] = 0.0;
    status = 0;
    goto done;
  }
  tolr = absolute ? tol : tol * normb;
  if (hist) hist[0] = normr / normb;
  while (iter < maxits) {
    rs[0] = normr;
    inormr = 1.0 / normr;
    for (r = 0; r < n; ++r) V[r] *= inormr;
    i = 0;
    while (i < m && iter < maxits) {
      double* v;
      double* z;
      i++;
      iter++;
      v = V + (size_t)(i - 1) * n;
      z = Z + (size_t)(i - 1) * n;
      if (prec) orc_gemslr_apply(prec, v, z, work);
      else memcpy(z, v, sizeof(double) * n);
      orc_csr_matvec(n, ia, ja, a, 1.0, z, 0.0, V + (size_t)i * n);
      if (orth == 1) orc_cgs2(n, V, i - 1, H, m + 1, &t);
      else orc_mgs(n, V, i - 1, H, m + 1, &t);
      for (j = 1; j < i; j++) {
        hii = Hm(j - 1, i - 1);
        Hm(j - 1, i - 1) = c[j - 1] * hii + s[j - 1] * Hm(j, i - 1);
        Hm(j, i - 1) = -s[j - 1] * hii + c[j - 1] * Hm(j, i - 1);
      }
      hii = Hm(i - 1, i - 1);
      hii1 = Hm(i, i - 1);
      gam = sqrt(hii * hii + hii1 * hii1);
      if (fabs(gam) < ORC_EPS) {
        status = 2;
        goto label;
      }
      c[i - 1] = hii / gam;
      s[i - 1] = hii1 / gam;
      rs[i] = -s[i - 1] * rs[i - 1];
      rs[i - 1] = c[i - 1] * rs[i - 1];
      Hm(i - 1, i - 1) = c[i - 1] * hii + s[i - 1] * hii1;
      normr = fabs(rs[i]);
      if (hist) hist[iter] = normr / normb;
      if (normr <= tolr) break;
    }
    rs[i - 1] /= Hm(i - 1, i - 1);
    for (k = i - 2; k >= 0; k--) {
      for (j = k + 1; j < i; j++) rs[k] -= Hm(k, j) * rs[j];
      rs[k] /= Hm(k, k);
    }
    for (j = 0; j < i; j++) {
      const double* z = Z + (size_t)j * n;
      for (r = 0; r < n; ++r) x[r] += rs[j] * z[r];
    }
    if (normr <= tolr) {
      status = 0;
      break;
    }
    /* restart: v = b - A x, keeping the Givens estimate as normr */
    memcpy(V, b, sizeof(double) * n);
    orc_csr_matvec(n, ia, ja, a, 1.0, x, 0.0, V + (size_t)n);
    for (r = 0; r < n; ++r) V[r] -= V[(size_t)n + r];
  }
label:
  *iters_out = iter;
  *rel_res_out = normr / normb;
done:
#undef Hm
  free(V);
  free(Z);
  free(H);
  free(c);
  free(s);
  free(rs);
  free(work);
  return status;
}

/*
 * Body of the oracle, written once for a scalar type and included by
 * oracle.c for double (ORC_SFX empty) and for double complex (ORC_SFX _z).
 * TEST INFRASTRUCTURE ONLY (see oracle.c).
 *
 * Macros set by the includer:
 *   ORC_T        scalar type
 *   ORC_FN(x)    name of function x for this scalar type
 *   ORC_CONJ(v)  complex conjugate (identity for real)
 *   ORC_ABS(v)   modulus, as PargemslrAbs (fabs / cabs)
 */

/* y = beta*y; y += alpha*A*x  (CsrMatrixPMatVecHostTemplate, 'N',
 * SRC/matrices/matrixops.cpp:5139-5330; x != y assumed) */
void ORC_FN(orc_csr_matvec)(int nrows, const int* ia, const int* ja, const ORC_T* a, double alpha, const ORC_T* x,
                            double beta, ORC_T* y) {
  int i, j;
  for (i = 0; i < nrows; ++i) y[i] = (beta == 0.0) ? (ORC_T)0.0 : beta * y[i];
  if (alpha == 0.0) return;
  for (i = 0; i < nrows; ++i) {
    ORC_T r = 0.0;
    for (j = ia[i]; j < ia[i + 1]; ++j) r += a[j] * x[ja[j]];
    y[i] += alpha * r;
  }
}

/* x = U^{-1} L^{-1} rhs, L unit lower (strict part stored), U strictly upper,
 * D stored inverted (IluClass::SolveHost, no-permutation branch,
 * SRC/preconditioners/ilu.cpp:3752-3783) */
void ORC_FN(orc_ilu_solve)(int n, const int* L_ia, const int* L_ja, const ORC_T* L_a, const int* U_ia,
                           const int* U_ja, const ORC_T* U_a, const ORC_T* Dinv, const ORC_T* rhs, ORC_T* x) {
  int i, j;
  for (i = 0; i < n; ++i) x[i] = rhs[i];
  for (i = 0; i < n; ++i)
    for (j = L_ia[i]; j < L_ia[i + 1]; ++j) x[i] -= L_a[j] * x[L_ja[j]];
  for (i = n - 1; i >= 0; --i) {
    for (j = U_ia[i]; j < U_ia[i + 1]; ++j) x[i] -= U_a[j] * x[U_ja[j]];
    x[i] *= Dinv[i];
  }
}

/* The block-diagonal ILU solve of one level, one diagonal block at a time
 * (GemslrClass::SolveB, SRC/preconditioners/gemslr.cpp:3632-3674, each block
 * by IluClass::SolveHost, ilu.cpp:3677-3786), the blocks spread over
 * orc_threads OpenMP threads as SolveB's loop (gemslr.cpp:3653-3655).  A
 * block's rows reference only its own rows, and every row is computed as in
 * orc_ilu_solve. */
static void ORC_FN(level_ilu_solve)(const orc_level* L, const ORC_T* rhs, ORC_T* x) {
  const ORC_T* La = (const ORC_T*)L->L_a;
  const ORC_T* Ua = (const ORC_T*)L->U_a;
  const ORC_T* Dinv = (const ORC_T*)L->Dinv;
  int b;
  if (!L->dom_ptr || L->ncomp <= 1 || orc_threads <= 1) {
    ORC_FN(orc_ilu_solve)(L->nloc, L->L_ia, L->L_ja, La, L->U_ia, L->U_ja, Ua, Dinv, rhs, x);
    return;
  }
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_threads)
  for (b = 0; b < L->ncomp; ++b) {
    const int r0 = L->dom_ptr[b], r1 = L->dom_ptr[b + 1];
    int i, j;
    for (i = r0; i < r1; ++i) x[i] = rhs[i];
    for (i = r0; i < r1; ++i)
      for (j = L->L_ia[i]; j < L->L_ia[i + 1]; ++j) x[i] -= La[j] * x[L->L_ja[j]];
    for (i = r1 - 1; i >= r0; --i) {
      for (j = L->U_ia[i]; j < L->U_ia[i + 1]; ++j) x[i] -= Ua[j] * x[L->U_ja[j]];
      x[i] *= Dinv[i];
    }
  }
}

/* x = WH * (W^H * z): DenseMatrixPMatVecTemplate 'C' then 'N'
 * (SRC/matrices/matrixops.cpp:101-332; the 'C' product conjugates W,
 * matrixops.cpp:260), as GemslrClass::SolveApplyLowRankLevel
 * (SRC/preconditioners/gemslr.cpp:3681-3719).  t: workspace of `rank`. */
void ORC_FN(orc_lowrank_apply)(int m, int rank, const ORC_T* W, const ORC_T* WH, const ORC_T* z, ORC_T* x,
                               ORC_T* t) {
  int i, j;
  for (j = 0; j < rank; ++j) {
    ORC_T s = 0.0;
    for (i = 0; i < m; ++i) s += ORC_CONJ(W[(size_t)j * m + i]) * z[i];
    t[j] = s;
  }
  for (i = 0; i < m; ++i) x[i] = 0.0;
  for (j = 0; j < rank; ++j)
    for (i = 0; i < m; ++i) x[i] += WH[(size_t)j * m + i] * t[j];
}

static void ORC_FN(solve_level)(const orc_gemslr* op, int l, const ORC_T* rhs, ORC_T* x, ORC_T* work);
void ORC_FN(orc_mgs)(int n, ORC_T* V, int k, ORC_T* H, int ldh, double* t_out);

/* x^H y (SequentialVectorClass::Dot -> zdotc, SRC/vectors/vectorops.cpp:100-120) */
static ORC_T ORC_FN(dot)(int n, const ORC_T* x, const ORC_T* y) {
  ORC_T s = 0.0;
  int i;
  for (i = 0; i < n; ++i) s += ORC_CONJ(x[i]) * y[i];
  return s;
}

/* Norm2: sqrt(|x^H x|) */
static double ORC_FN(norm2)(int n, const ORC_T* x) { return sqrt(ORC_ABS(ORC_FN(dot)(n, x, x))); }

/* y = S x on level 0, S = C - E B^{-1} F with B^{-1} the block ILU solve
 * (GemslrClass::SchurMatVec, SRC/preconditioners/gemslr.cpp:3799-3955 with
 * alpha = 1, beta = 0: y = -E B^{-1} F x, then y += C x).  u, b: nloc values. */
static void ORC_FN(schur_matvec)(const orc_gemslr* op, const ORC_T* x, ORC_T* y, ORC_T* u, ORC_T* b) {
  const orc_level* L = &op->levels[0];
  const int nl = L->nloc, m = L->nrem;
  int i, k;
  ORC_FN(orc_csr_matvec)(nl, L->F_ia, L->F_ja, (const ORC_T*)L->F_a, 1.0, x, 0.0, u);
  ORC_FN(orc_ilu_solve)(nl, L->L_ia, L->L_ja, (const ORC_T*)L->L_a, L->U_ia, L->U_ja, (const ORC_T*)L->U_a,
                        (const ORC_T*)L->Dinv, u, b);
  for (i = 0; i < m; ++i) {
    ORC_T e = 0.0, c = 0.0;
    for (k = L->E_ia[i]; k < L->E_ia[i + 1]; ++k) e += ((const ORC_T*)L->E_a)[k] * b[L->E_ja[k]];
    for (k = L->C_ia[i]; k < L->C_ia[i + 1]; ++k) c += ((const ORC_T*)L->C_a)[k] * x[L->C_ja[k]];
    y[i] = -e + c;
  }
}

/* z = SolveLevel(1, v + W H W^H v): the preconditioner of the inner
 * iteration (GemslrSchurMatrixClass::Solve, gemslr.cpp:292-320).
 * t: m values, xlr: m values, c: rank values, work: the level-1 buffers. */
static void ORC_FN(schur_precond)(const orc_gemslr* op, const ORC_T* v, ORC_T* z, ORC_T* t, ORC_T* xlr, ORC_T* c,
                                  ORC_T* work) {
  const orc_level* L = &op->levels[0];
  const int m = L->nrem;
  int i;
  if (L->rank > 0) {
    ORC_FN(orc_lowrank_apply)(m, L->rank, (const ORC_T*)L->W, (const ORC_T*)L->WH, v, xlr, c);
    for (i = 0; i < m; ++i) t[i] = xlr[i] + v[i];
  } else {
    for (i = 0; i < m; ++i) t[i] = v[i];
  }
  ORC_FN(solve_level)(op, 1, t, z, work);
}

/* Inner FGMRES on the level-0 Schur complement with x0 = 0, the solve phase
 * of SolveLevel with inner iterations (gemslr.cpp:3556-3563): FlexGmresClass::
 * Solve (SRC/solvers/fgmres.hpp:263-558) with Krylov dimension = iteration
 * limit = inner_maxits, relative tolerance inner_tol, PargemslrMgs.  The
 * restart residual the reference computes after the last cycle does not
 * change x and is skipped.  b, x: nrem values; work: the level-1 buffers. */
static void ORC_FN(schur_inner)(const orc_gemslr* op, const ORC_T* b, ORC_T* x, ORC_T* work) {
  const orc_level* L = &op->levels[0];
  const int m = L->nrem, nl = L->nloc, K = op->inner_maxits;
  ORC_T *V, *Z, *H, *c, *s, *rs, *t, *xlr, *lc, *u, *bb;
  ORC_T hii, hii1;
  double normb, normr, tolr, tn, gam;
  int i = 0, j, k, r;
  for (r = 0; r < m; ++r) x[r] = 0.0;
  normb = ORC_FN(norm2)(m, b);
  if (normb < ORC_EPS) return;
  V = (ORC_T*)calloc((size_t)m * (K + 1), sizeof(ORC_T));
  Z = (ORC_T*)calloc((size_t)m * K, sizeof(ORC_T));
  H = (ORC_T*)calloc((size_t)(K + 1) * K, sizeof(ORC_T));
  c = (ORC_T*)calloc(K, sizeof(ORC_T));
  s = (ORC_T*)calloc(K, sizeof(ORC_T));
  rs = (ORC_T*)calloc(K + 1, sizeof(ORC_T));
  t = (ORC_T*)calloc(m, sizeof(ORC_T));
  xlr = (ORC_T*)calloc(m, sizeof(ORC_T));
  lc = (ORC_T*)calloc(L->rank > 0 ? L->rank : 1, sizeof(ORC_T));
  u = (ORC_T*)calloc(nl > 0 ? nl : 1, sizeof(ORC_T));
  bb = (ORC_T*)calloc(nl > 0 ? nl : 1, sizeof(ORC_T));
#define Hi(ii, jj) H[(ii) + (size_t)(jj) * (K + 1)]
  /* x0 = 0: the residual is b, normr = normb */
  normr = normb;
  tolr = op->inner_tol * normb;
  rs[0] = normr;
  for (r = 0; r < m; ++r) V[r] = b[r] * (1.0 / normr);
  while (i < K) {
    ORC_T* w;
    i++;
    w = V + (size_t)i * m;
    ORC_FN(schur_precond)(op, V + (size_t)(i - 1) * m, Z + (size_t)(i - 1) * m, t, xlr, lc, work);
    ORC_FN(schur_matvec)(op, Z + (size_t)(i - 1) * m, w, u, bb);
    ORC_FN(orc_mgs)(m, V, i - 1, H, K + 1, &tn);
    for (j = 1; j < i; j++) {
      hii = Hi(j - 1, i - 1);
      Hi(j - 1, i - 1) = ORC_CONJ(c[j - 1]) * hii + s[j - 1] * Hi(j, i - 1);
      Hi(j, i - 1) = -s[j - 1] * hii + c[j - 1] * Hi(j, i - 1);
    }
    hii = Hi(i - 1, i - 1);
    hii1 = Hi(i, i - 1);
    gam = sqrt(ORC_ABS(hii) * ORC_ABS(hii) + ORC_ABS(hii1) * ORC_ABS(hii1));
    if (fabs(gam) < ORC_EPS) goto out; /* breakdown: x stays at the initial guess */
    c[i - 1] = hii / gam;
    s[i - 1] = hii1 / gam;
    rs[i] = -s[i - 1] * rs[i - 1];
    rs[i - 1] = ORC_CONJ(c[i - 1]) * rs[i - 1];
    Hi(i - 1, i - 1) = ORC_CONJ(c[i - 1]) * hii + s[i - 1] * hii1;
    normr = ORC_ABS(rs[i]);
    if (normr <= tolr) break;
  }
  rs[i - 1] /= Hi(i - 1, i - 1);
  for (k = i - 2; k >= 0; k--) {
    for (j = k + 1; j < i; j++) rs[k] -= Hi(k, j) * rs[j];
    rs[k] /= Hi(k, k);
  }
  for (j = 0; j < i; j++) {
    const ORC_T* z = Z + (size_t)j * m;
    for (r = 0; r < m; ++r) x[r] += rs[j] * z[r];
  }
out:
#undef Hi
  free(V);
  free(Z);
  free(H);
  free(c);
  free(s);
  free(rs);
  free(t);
  free(xlr);
  free(lc);
  free(u);
  free(bb);
}

/* GemslrClass::SolveLevel (SRC/preconditioners/gemslr.cpp:3474-3625); on
 * level 0 with inner_maxits > 0 the low-rank correction and the next level
 * are replaced by the inner FGMRES (3556-3563).  rhs, x: length n -
 * lev_start.  work: per level buffers carved from `work` (z of length n -
 * lev_start, xlr of length nrem, t of length rank). */
static void ORC_FN(solve_level)(const orc_gemslr* op, int l, const ORC_T* rhs, ORC_T* x, ORC_T* work) {
  const orc_level* L = &op->levels[l];
  const int nl = L->nloc, m = L->nrem;
  ORC_T *z, *zl, *xlr, *t;
  int i;
  if (l == op->nlev - 1) {
    ORC_FN(level_ilu_solve)(L, rhs, x);
    return;
  }
  z = work;
  zl = z + nl;
  xlr = z + nl + m;
  t = xlr + m;
  if (op->usolve) {
    /* the U variant (kGemslrUSolve, gemslr.cpp:3527-3532): step 1 is z = rhs,
     * steps 2-3 are skipped */
    for (i = 0; i < m; ++i) zl[i] = rhs[nl + i];
  } else {
    /* step 1: z_u = B^{-1} rhs_u (block-diagonal ILU, blocks are independent) */
    ORC_FN(level_ilu_solve)(L, rhs, z);
    /* steps 2-3: z_l = -E z_u; z_l += rhs_l */
    ORC_FN(orc_csr_matvec)(m, L->E_ia, L->E_ja, (const ORC_T*)L->E_a, -1.0, z, 0.0, zl);
    for (i = 0; i < m; ++i) zl[i] += rhs[nl + i];
  }
  /* step 4: low-rank correction, then the next level (or the inner iteration) */
  if (l == 0 && op->inner_maxits > 0) {
    ORC_FN(schur_inner)(op, zl, x + nl, t + (L->rank > 0 ? L->rank : 0));
  } else {
    if (L->rank > 0) {
      ORC_FN(orc_lowrank_apply)(m, L->rank, (const ORC_T*)L->W, (const ORC_T*)L->WH, zl, xlr, t);
      for (i = 0; i < m; ++i) zl[i] += xlr[i];
    }
    ORC_FN(solve_level)(op, l + 1, zl, x + nl, t + (L->rank > 0 ? L->rank : 0));
  }
  /* steps 5-6: z_u = -F x_l; z_u += rhs_u */
  ORC_FN(orc_csr_matvec)(nl, L->F_ia, L->F_ja, (const ORC_T*)L->F_a, -1.0, x + nl, 0.0, z);
  for (i = 0; i < nl; ++i) z[i] += rhs[i];
  /* step 7: x_u = B^{-1} z_u */
  ORC_FN(level_ilu_solve)(L, z, x);
}

/* z = M^{-1} v: GatherPerm, SolveLevel(0), ScatterRperm
 * (GemslrClass::Solve, SRC/preconditioners/gemslr.cpp:3431-3467;
 *  IntVectorClass::GatherPerm/ScatterRperm, SRC/vectors/int_vector.cpp:629-696).
 * work: orc_gemslr_work_size(op) values of ORC_T. */
void ORC_FN(orc_gemslr_apply)(const orc_gemslr* op, const ORC_T* v, ORC_T* z, ORC_T* work) {
  const int n = op->n;
  ORC_T* r = work;
  ORC_T* x = work + n;
  int i;
  for (i = 0; i < n; ++i) r[i] = v[op->perm[i]];
  ORC_FN(solve_level)(op, 0, r, x, work + 2 * n);
  for (i = 0; i < n; ++i) z[op->perm[i]] = x[i];
}

/* PargemslrMgs without re-orthogonalization (tol_reorth < 0, as FGMRES calls
 * it: SRC/solvers/fgmres.hpp:426; SRC/matrices/matrixops.cpp:10728-10807).
 * V column-major n x (k+2) with ld n; w = V(:,k+1).  H(j,k) at H[j + k*ldh]. */
void ORC_FN(orc_mgs)(int n, ORC_T* V, int k, ORC_T* H, int ldh, double* t_out) {
  ORC_T* w = V + (size_t)(k + 1) * n;
  int i, r;
  double t;
  for (i = 0; i <= k; ++i) {
    const ORC_T* v = V + (size_t)i * n;
    ORC_T t1 = ORC_FN(dot)(n, v, w);
    H[i + (size_t)k * ldh] = t1;
    for (r = 0; r < n; ++r) w[r] -= t1 * v[r];
  }
  t = ORC_FN(norm2)(n, w);
  H[(k + 1) + (size_t)k * ldh] = t;
  for (r = 0; r < n; ++r) w[r] *= 1.0 / t;
  *t_out = t;
}

/* PargemslrCgs2 (SRC/matrices/matrixops.cpp:10588-10709). */
void ORC_FN(orc_cgs2)(int n, ORC_T* V, int k, ORC_T* H, int ldh, double* t_out) {
  ORC_T* w = V + (size_t)(k + 1) * n;
  ORC_T* s = (ORC_T*)malloc(sizeof(ORC_T) * (k + 1));
  int i, r;
  double t;
  for (i = 0; i <= k; ++i) H[i + (size_t)k * ldh] = ORC_FN(dot)(n, V + (size_t)i * n, w);
  for (i = 0; i <= k; ++i) {
    const ORC_T* v = V + (size_t)i * n;
    ORC_T h = H[i + (size_t)k * ldh];
    for (r = 0; r < n; ++r) w[r] += -h * v[r];
  }
  for (i = 0; i <= k; ++i) s[i] = ORC_FN(dot)(n, V + (size_t)i * n, w);
  for (i = 0; i <= k; ++i) {
    const ORC_T* v = V + (size_t)i * n;
    for (r = 0; r < n; ++r) w[r] += -s[i] * v[r];
    H[i + (size_t)k * ldh] += s[i];
  }
  t = ORC_FN(norm2)(n, w);
  H[(k + 1) + (size_t)k * ldh] = t;
  for (r = 0; r < n; ++r) w[r] *= 1.0 / t;
  free(s);
  *t_out = t;
}

/* FlexGmresClass::Solve (SRC/solvers/fgmres.hpp:263-558); the rotations use
 * conj(c) as fgmres.hpp:428-447 does.  prec == NULL: no preconditioner.
 * orth: 0 MGS, 1 CGS2.  hist (optional, maxits+1 entries): relative residual
 * estimates.  Returns 0 converged, 1 not converged, 2 breakdown. */
int ORC_FN(orc_fgmres)(int n, const int* ia, const int* ja, const ORC_T* a, const orc_gemslr* prec, const ORC_T* b,
                       ORC_T* x, int kdim, int maxits, double tol, int absolute, int orth, int* iters_out,
                       double* rel_res_out, double* hist) {
  const int m = kdim;
  ORC_T *V, *Z, *H, *c, *s, *rs, *work = NULL;
  ORC_T hii, hii1;
  double normb, normr, tolr, t, gam, inormr;
  int i = 0, j, k, iter = 0, status = 1, r;
  V = (ORC_T*)calloc((size_t)n * (m + 1), sizeof(ORC_T));
  Z = (ORC_T*)calloc((size_t)n * (m + 1), sizeof(ORC_T));
  H = (ORC_T*)calloc((size_t)(m + 1) * m, sizeof(ORC_T));
  c = (ORC_T*)calloc(m, sizeof(ORC_T));
  s = (ORC_T*)calloc(m, sizeof(ORC_T));
  rs = (ORC_T*)calloc(m + 1, sizeof(ORC_T));
  if (prec) work = (ORC_T*)malloc(sizeof(ORC_T) * orc_gemslr_work_size(prec));
#define Hm(ii, jj) H[(ii) + (size_t)(jj) * (m + 1)]
  normb = ORC_FN(norm2)(n, b);
  if (normb < ORC_EPS) {
    for (r = 0; r < n; ++r) x[r] = 0.0;
    *iters_out = 0;
    *rel_res_out = 0.0;
    if (hist) hist[0] = 0.0;
    status = 0;
    goto done;
  }
  /* v = b - A x */
  memcpy(V, b, sizeof(ORC_T) * n);
  ORC_FN(orc_csr_matvec)(n, ia, ja, a, -1.0, x, 1.0, V);
  normr = ORC_FN(norm2)(n, V);
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
      ORC_T* v;
      ORC_T* z;
      i++;
      iter++;
      v = V + (size_t)(i - 1) * n;
      z = Z + (size_t)(i - 1) * n;
      if (prec) ORC_FN(orc_gemslr_apply)(prec, v, z, work);
      else memcpy(z, v, sizeof(ORC_T) * n);
      ORC_FN(orc_csr_matvec)(n, ia, ja, a, 1.0, z, 0.0, V + (size_t)i * n);
      if (orth == 1) ORC_FN(orc_cgs2)(n, V, i - 1, H, m + 1, &t);
      else ORC_FN(orc_mgs)(n, V, i - 1, H, m + 1, &t);
      for (j = 1; j < i; j++) {
        hii = Hm(j - 1, i - 1);
        Hm(j - 1, i - 1) = ORC_CONJ(c[j - 1]) * hii + s[j - 1] * Hm(j, i - 1);
        Hm(j, i - 1) = -s[j - 1] * hii + c[j - 1] * Hm(j, i - 1);
      }
      hii = Hm(i - 1, i - 1);
      hii1 = Hm(i, i - 1);
      gam = sqrt(ORC_ABS(hii) * ORC_ABS(hii) + ORC_ABS(hii1) * ORC_ABS(hii1));
      if (fabs(gam) < ORC_EPS) {
        status = 2;
        goto label;
      }
      c[i - 1] = hii / gam;
      s[i - 1] = hii1 / gam;
      rs[i] = -s[i - 1] * rs[i - 1];
      rs[i - 1] = ORC_CONJ(c[i - 1]) * rs[i - 1];
      Hm(i - 1, i - 1) = ORC_CONJ(c[i - 1]) * hii + s[i - 1] * hii1;
      normr = ORC_ABS(rs[i]);
      if (hist) hist[iter] = normr / normb;
      if (normr <= tolr) break;
    }
    rs[i - 1] /= Hm(i - 1, i - 1);
    for (k = i - 2; k >= 0; k--) {
      for (j = k + 1; j < i; j++) rs[k] -= Hm(k, j) * rs[j];
      rs[k] /= Hm(k, k);
    }
    for (j = 0; j < i; j++) {
      const ORC_T* z = Z + (size_t)j * n;
      for (r = 0; r < n; ++r) x[r] += rs[j] * z[r];
    }
    if (normr <= tolr) {
      status = 0;
      break;
    }
    /* restart: v = b - A x, keeping the Givens estimate as normr */
    memcpy(V, b, sizeof(ORC_T) * n);
    ORC_FN(orc_csr_matvec)(n, ia, ja, a, 1.0, x, 0.0, V + (size_t)n);
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

// Host half of the C ABI: parameter array, GeMSLR setup, inspection and the
// host-resident apply.
#include <cstring>
#include <new>

#include "../../../include/pargemslr_amd.h"
#include "../device/capi_internal.hpp"
#include "gemslr_host.hpp"

using namespace pgs;

extern "C" {

int pgs_params_default(double* params) {
  if (!params) return PGS_ERROR_INVALID_PARAM;
  params_default_array(params);
  return 0;
}

static int gemslr_setup_impl(int n, const int* ia, const int* ja, const double* a, int is_complex,
                             const double* params, uint64_t seed, int lowrank_device, pgs_gemslr_t* out) {
  if (!out) return PGS_ERROR_INVALID_PARAM;
  *out = nullptr;
  if (n < 0 || (n > 0 && (!ia || !ja || !a)) || !params) return PGS_ERROR_INVALID_PARAM;
  if (n > 0 && ia[0] != 0) return PGS_ERROR_INVALID_PARAM;
  for (int i = 0; i < n; ++i)
    if (ia[i + 1] < ia[i]) return PGS_ERROR_INVALID_PARAM;
  const int64_t nnz = n > 0 ? ia[n] : 0;
  for (int64_t k = 0; k < nnz; ++k)
    if (ja[k] < 0 || ja[k] >= n) return PGS_ERROR_INVALID_PARAM;
  GemslrParams prm;
  params_from_array(params, prm);
  prm.seed = seed;
  prm.lowrank_device = lowrank_device;
  // inner Schur iteration: at least one and at most kSchurMaxIts steps (the
  // device keeps its Krylov basis in one workspace allocation)
  if (prm.schur_enable != 0 && (prm.schur_maxits < 1 || prm.schur_maxits > PGS_SCHUR_MAXITS_LIMIT ||
                                !(prm.schur_tol >= 0.0)))
    return PGS_ERROR_INVALID_OPTION;
  pgs_gemslr_s* h = new (std::nothrow) pgs_gemslr_s();
  if (!h) return PGS_ERR_OOM;
  h->is_complex = is_complex ? 1 : 0;
  int err = 0;
  try {
    if (!is_complex) {
      Csr<double> A;
      A.nrows = A.ncols = n;
      A.ia.assign(ia, ia + n + 1);
      A.ja.assign(ja, ja + nnz);
      A.a.assign(a, a + nnz);
      h->op_d.reset(new HostOp<double>());
      err = gemslr_setup(A, prm, *h->op_d, &h->msg);
      h->work_d.reset(new HostApplyWork<double>());
    } else {
      Csr<zdouble> A;
      A.nrows = A.ncols = n;
      A.ia.assign(ia, ia + n + 1);
      A.ja.assign(ja, ja + nnz);
      A.a.resize(nnz);
      std::memcpy((void*)A.a.data(), a, sizeof(zdouble) * nnz);
      h->op_z.reset(new HostOp<zdouble>());
      err = gemslr_setup(A, prm, *h->op_z, &h->msg);
      h->work_z.reset(new HostApplyWork<zdouble>());
    }
  } catch (const std::bad_alloc&) {
    err = PGS_ERR_OOM;
  } catch (...) {
    err = PGS_ERROR_FUNCTION_CALL;
  }
  if (err) {
    delete h;
    return err;
  }
  *out = reinterpret_cast<pgs_gemslr_t>(h);
  return 0;
}

int pgs_gemslr_setup(int n, const int* ia, const int* ja, const double* a, int is_complex, const double* params,
                     uint64_t seed, pgs_gemslr_t* out) {
  return gemslr_setup_impl(n, ia, ja, a, is_complex, params, seed, -1, out);
}

int pgs_gemslr_setup_device(int n, const int* ia, const int* ja, const double* a, int is_complex,
                            const double* params, uint64_t seed, int device, pgs_gemslr_t* out) {
  if (device < 0) {
    if (out) *out = nullptr;
    return PGS_ERROR_INVALID_PARAM;
  }
  return gemslr_setup_impl(n, ia, ja, a, is_complex, params, seed, device, out);
}

int pgs_gemslr_destroy(pgs_gemslr_t op) {
  delete reinterpret_cast<pgs_gemslr_s*>(op);
  return 0;
}

}  // extern "C"

namespace {
template <typename T>
void fill_info(const HostOp<T>& H, pgs_gemslr_info_t* info) {
  info->n = H.n;
  info->nlev = H.nlev;
  info->nnz_a = H.A.nnz();
  info->nnz_ilu = 0;
  info->nnz_lr = 0;
  for (int l = 0; l < H.nlev; ++l) {
    for (const auto& f : H.levels[l].ilu) info->nnz_ilu += f.nnz();
    info->nnz_lr += 2LL * H.nrem(l) * H.levels[l].rank;
  }
  for (int i = 0; i < 4; ++i) info->setup_seconds[i] = H.setup_seconds[i];
  // the inner iteration runs on level 0 when there is a level below it
  const bool inner = H.prm.schur_enable && H.nlev > 1;
  info->schur_maxits = inner ? H.prm.schur_maxits : 0;
  info->schur_tol = inner ? H.prm.schur_tol : 0.0;
  info->global_solve = H.prm.global_solve;
}

template <typename T>
void fill_view(const HostOp<T>& H, int l, pgs_level_view_t* v) {
  const HostLevel<T>& L = H.levels[l];
  v->lev_start = H.lev_ptr[l];
  v->nloc = H.nloc(l);
  v->nrem = H.nrem(l);
  v->ncomp = L.ncomp;
  v->rank = L.rank;
  v->dom_ptr = nullptr;
  v->L_ia = L.Lcat.ia.data();
  v->L_ja = L.Lcat.ja.data();
  v->L_a = reinterpret_cast<const double*>(L.Lcat.a.data());
  v->U_ia = L.Ucat.ia.data();
  v->U_ja = L.Ucat.ja.data();
  v->U_a = reinterpret_cast<const double*>(L.Ucat.a.data());
  v->Dinv = reinterpret_cast<const double*>(L.Dcat.data());
  v->E_ia = L.E.ia.data();
  v->E_ja = L.E.ja.data();
  v->E_a = reinterpret_cast<const double*>(L.E.a.data());
  v->F_ia = L.F.ia.data();
  v->F_ja = L.F.ja.data();
  v->F_a = reinterpret_cast<const double*>(L.F.a.data());
  v->W = L.rank > 0 ? reinterpret_cast<const double*>(L.W.data()) : nullptr;
  v->WH = L.rank > 0 ? reinterpret_cast<const double*>(L.WH.data()) : nullptr;
  const bool has_c = l == 0 && H.prm.schur_enable && !L.C.ia.empty();
  v->C_ia = has_c ? L.C.ia.data() : nullptr;
  v->C_ja = has_c ? L.C.ja.data() : nullptr;
  v->C_a = has_c ? reinterpret_cast<const double*>(L.C.a.data()) : nullptr;
}
}  // namespace

extern "C" {

int pgs_gemslr_info(pgs_gemslr_t op, pgs_gemslr_info_t* info) {
  if (!op || !info) return PGS_ERROR_INVALID_PARAM;
  pgs_gemslr_s* h = reinterpret_cast<pgs_gemslr_s*>(op);
  std::memset(info, 0, sizeof(*info));
  info->is_complex = h->is_complex;
  if (h->is_complex) fill_info(*h->op_z, info);
  else fill_info(*h->op_d, info);
  return 0;
}

int pgs_gemslr_level_view(pgs_gemslr_t op, int level, pgs_level_view_t* view) {
  if (!op || !view) return PGS_ERROR_INVALID_PARAM;
  pgs_gemslr_s* h = reinterpret_cast<pgs_gemslr_s*>(op);
  const int nlev = h->is_complex ? h->op_z->nlev : h->op_d->nlev;
  if (level < 0 || level >= nlev) return PGS_ERROR_INVALID_PARAM;
  if (h->is_complex) fill_view(*h->op_z, level, view);
  else fill_view(*h->op_d, level, view);
  // level-local block offsets
  static thread_local std::vector<int> dp;
  const std::vector<int>& g = h->is_complex ? h->op_z->levels[level].dom_ptr : h->op_d->levels[level].dom_ptr;
  dp.assign(g.begin(), g.end());
  for (int& x : dp) x -= view->lev_start;
  view->dom_ptr = dp.data();
  return 0;
}

int pgs_gemslr_perm(pgs_gemslr_t op, const int** perm) {
  if (!op || !perm) return PGS_ERROR_INVALID_PARAM;
  pgs_gemslr_s* h = reinterpret_cast<pgs_gemslr_s*>(op);
  *perm = h->is_complex ? h->op_z->perm.data() : h->op_d->perm.data();
  return 0;
}

int pgs_gemslr_save(pgs_gemslr_t op, const char* path) {
  if (!op || !path) return PGS_ERROR_INVALID_PARAM;
  pgs_gemslr_s* h = reinterpret_cast<pgs_gemslr_s*>(op);
  try {
    return h->is_complex ? op_save(*h->op_z, path) : op_save(*h->op_d, path);
  } catch (const std::bad_alloc&) {
    return PGS_ERR_OOM;
  }
}

int pgs_gemslr_load(const char* path, pgs_gemslr_t* out) {
  if (!out) return PGS_ERROR_INVALID_PARAM;
  *out = nullptr;
  if (!path) return PGS_ERROR_INVALID_PARAM;
  int kind = 0;
  int err = op_file_kind(path, &kind);
  if (err) return err;
  pgs_gemslr_s* h = new (std::nothrow) pgs_gemslr_s();
  if (!h) return PGS_ERR_OOM;
  h->is_complex = kind;
  try {
    if (!kind) {
      h->op_d.reset(new HostOp<double>());
      err = op_load(path, *h->op_d, &h->msg);
      h->work_d.reset(new HostApplyWork<double>());
    } else {
      h->op_z.reset(new HostOp<zdouble>());
      err = op_load(path, *h->op_z, &h->msg);
      h->work_z.reset(new HostApplyWork<zdouble>());
    }
  } catch (const std::bad_alloc&) {
    err = PGS_ERR_OOM;
  }
  if (err) {
    delete h;
    return err;
  }
  *out = reinterpret_cast<pgs_gemslr_t>(h);
  return 0;
}

int pgs_gemslr_matrix(pgs_gemslr_t op, int* n, const int** ia, const int** ja, const double** a) {
  if (!op || !n || !ia || !ja || !a) return PGS_ERROR_INVALID_PARAM;
  pgs_gemslr_s* h = reinterpret_cast<pgs_gemslr_s*>(op);
  if (h->is_complex) {
    const Csr<zdouble>& A = h->op_z->A;
    *n = A.nrows;
    *ia = A.ia.data();
    *ja = A.ja.data();
    *a = reinterpret_cast<const double*>(A.a.data());
  } else {
    const Csr<double>& A = h->op_d->A;
    *n = A.nrows;
    *ia = A.ia.data();
    *ja = A.ja.data();
    *a = A.a.data();
  }
  return 0;
}

int pgs_gemslr_host_apply(pgs_gemslr_t op, const double* v, double* z) {
  if (!op || !v || !z) return PGS_ERROR_INVALID_PARAM;
  pgs_gemslr_s* h = reinterpret_cast<pgs_gemslr_s*>(op);
  if (h->is_complex)
    host_apply(*h->op_z, reinterpret_cast<const zdouble*>(v), reinterpret_cast<zdouble*>(z), *h->work_z);
  else
    host_apply(*h->op_d, v, z, *h->work_d);
  return 0;
}

const char* pgs_strerror(int code) {
  switch (code) {
    case PGS_SUCCESS: return "success";
    case PGS_ERROR_INVALID_OPTION: return "invalid option";
    case PGS_ERROR_INVALID_PARAM: return "invalid parameter";
    case PGS_ERROR_IO: return "I/O error";
    case PGS_ERROR_ILU_EMPTY_ROW: return "ILUT: empty row";
    case PGS_ERROR_FUNCTION_CALL: return "function call error";
    case PGS_ERROR_MEMORY_LOCATION: return "memory location error";
    case PGS_ERR_HIP: return "HIP runtime error";
    case PGS_ERR_OOM: return "out of device memory";
    case PGS_ERR_NO_DEVICE: return "no HIP device";
    case PGS_ERR_BREAKDOWN: return "FGMRES breakdown";
    case PGS_ERR_NOT_CONVERGED: return "FGMRES did not converge";
    case PGS_ERR_UNSUPPORTED: return "unsupported operation";
    case PGS_ERR_NONFINITE: return "NaN or Inf in the right-hand side, the initial guess or the iteration";
    case PGS_ERR_SCRATCH_LIMIT: return "the apply needs more scratch memory than pgs_apply_noalloc holds";
    default: return "unknown error";
  }
}

}  // extern "C"

// Preconditioner object with the life cycle of the reference's hypre
// interface (HYPRE_GEMSLRCreate / CreateFromFile / Setup / Solve / Destroy,
// ParGeMSLR/TESTS/hypre/HYPRE_pargemslr.cpp:33-232) at one process, written
// over this library's own C ABI.  hypre's types are not needed: at one
// process a ParCSR matrix is its diagonal block, so setup takes that block's
// CSR arrays, and solve takes the local vectors' data.
#include <cstdint>
#include <new>

#include "../../../include/pargemslr_amd.h"

struct pgs_hypre_gemslr_s {
  double params[PGS_IO_SIZE];
  int prec = PGS_PREC_FP64;
  int device = 0;
  int n = 0;
  pgs_gemslr_t op = nullptr;
  pgs_dev_op_t dop = nullptr;
  pgs_ws_t ws = nullptr;
  double* dv = nullptr;
  double* dz = nullptr;
  void release() {
    if (dv) pgs_dev_free(device, dv);
    if (dz) pgs_dev_free(device, dz);
    if (ws) pgs_ws_destroy(ws);
    if (dop) pgs_dev_op_destroy(dop);
    if (op) pgs_gemslr_destroy(op);
    dv = dz = nullptr;
    ws = nullptr;
    dop = nullptr;
    op = nullptr;
    n = 0;
  }
};

extern "C" {

int pgs_hypre_gemslr_create(const char* inputs_file, pgs_hypre_gemslr_t* out) {
  if (!out) return PGS_ERROR_INVALID_PARAM;
  pgs_hypre_gemslr_s* s = new (std::nothrow) pgs_hypre_gemslr_s();
  if (!s) return PGS_ERR_OOM;
  pgs_params_default(s->params);
  // HYPRE_GEMSLRCreateFromFile reads "inputs" when no file is named and
  // keeps the defaults for whatever the file does not set, without failing
  (void)pgs_params_read_inputs(inputs_file ? inputs_file : "inputs", PGS_INPUTS_HYPRE, s->params);
  *out = s;
  return 0;
}

int pgs_hypre_gemslr_params(pgs_hypre_gemslr_t s, double* params) {
  if (!s || !params) return PGS_ERROR_INVALID_PARAM;
  for (int i = 0; i < PGS_IO_SIZE; ++i) params[i] = s->params[i];
  return 0;
}

int pgs_hypre_gemslr_set_precision(pgs_hypre_gemslr_t s, int prec) {
  if (!s || (prec != PGS_PREC_FP64 && prec != PGS_PREC_FP32)) return PGS_ERROR_INVALID_PARAM;
  s->prec = prec;
  return 0;
}

int pgs_hypre_gemslr_setup(pgs_hypre_gemslr_t s, int n, const int* ia, const int* ja, const double* a,
                           int device) {
  if (!s) return PGS_ERROR_INVALID_PARAM;
  s->release();
  s->device = device;
  int err = pgs_gemslr_setup(n, ia, ja, a, 0, s->params, 0, &s->op);
  if (!err) err = pgs_dev_op_create(s->op, device, PGS_PREC_FP64 | PGS_PREC_FP32, &s->dop);
  if (!err) err = pgs_ws_create(s->dop, 1, 1, &s->ws);
  if (!err) err = pgs_dev_malloc(device, (int64_t)sizeof(double) * n, (void**)&s->dv);
  if (!err) err = pgs_dev_malloc(device, (int64_t)sizeof(double) * n, (void**)&s->dz);
  if (err) {
    s->release();
    return err;
  }
  s->n = n;
  return 0;
}

int pgs_hypre_gemslr_solve(pgs_hypre_gemslr_t s, const double* b, double* x) {
  if (!s || !b || !x) return PGS_ERROR_INVALID_PARAM;
  if (!s->dop) return PGS_ERROR_FUNCTION_CALL;  // solve before setup
  const int64_t bytes = (int64_t)sizeof(double) * s->n;
  int err = pgs_memcpy_htod(s->device, s->dv, b, bytes, nullptr);
  if (!err) err = pgs_apply(s->dop, s->ws, s->dv, s->dz, 1, s->prec, nullptr);
  if (!err) err = pgs_memcpy_dtoh(s->device, x, s->dz, bytes, nullptr);
  return err;
}

int pgs_hypre_gemslr_destroy(pgs_hypre_gemslr_t s) {
  if (!s) return 0;
  s->release();
  delete s;
  return 0;
}

}  // extern "C"

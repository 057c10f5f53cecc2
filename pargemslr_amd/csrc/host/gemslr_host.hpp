// GeMSLR factored operator on the host: the output of the setup phase and the
// input of the device upload.
//
// Data model follows GemslrClass / GemslrLevelClass
// (ParGeMSLR/SRC/preconditioners/gemslr.hpp:1103-1282, 1511-3190): a symmetric
// permutation P, levels l = 0..nlev-1 with row ranges [lev_ptr[l], lev_ptr[l+1])
// of P A P^T, ncomp diagonal blocks per level factored by ILUT, coupling blocks
// E_l (rows below the level x level columns) and F_l (level rows x columns
// below), and a rank-k correction (W_l, WH_l = W_l ((I - H_l)^{-1} - I)).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "csr.hpp"
#include "rcm_ilut.hpp"

namespace pgs {

// Typed view of the PARGEMSLR_IO_* entries this library reads
// (ParGeMSLR/SRC/utils/utils.hpp:77-297, defaults utils.cpp:31-127).
struct GemslrParams {
  int bsolve_top = 0;    // B solve of the top levels: 0 = ILUT (the only option implemented)
  int bsolve_lower = 0;  // B solve of the other levels: 0 = ILUT
  int global_solve = 0;  // SolveLevel variant of the solve phase: 0 = LU, 1 = U (skip the block L solve)
  int nlev = 5;
  int ncomp = 4;
  int partition = 1;     // 1 = RKway (the only option implemented)
  int minsep = 10;
  int kmin = 1;
  int kfactor = 1;
  bool vertexsep = true;
  double droptol_B = 1e-2;
  double droptol_C = 1e-3;
  int rownnz_B = 100;
  int rownnz_C = 200;
  int perm_option = 1;   // 0 none, 1 RCM
  int lr_option1 = 1, lr_option2 = 1;  // 1 = thick-restart Arnoldi
  int lr_rank1 = 100, lr_rank2 = 50;
  double lr_rank_factor1 = 1.0, lr_rank_factor2 = 1.0;
  double lr_arnoldi_factor1 = 2.0, lr_arnoldi_factor2 = 2.0;
  double lr_tol_eig1 = 1e-9, lr_tol_eig2 = 1e-9;
  int lr_maxits1 = 3, lr_maxits2 = 3;
  double lr_tr_factor = 0.25;
  double lr_tol_orth = 1e-14;
  bool lr_rand_init = true;
  int schur_enable = 1;
  double schur_tol = 1e-2;
  int schur_maxits = 5;
  uint64_t seed = 0;
  int lowrank_device = -1;  // >= 0: the low-rank corrections are built on this device
};

template <typename T>
struct HostLevel {
  int ncomp = 0;
  std::vector<int> dom_ptr;          // ncomp+1 global offsets (permuted numbering)
  std::vector<IluFactor<T>> ilu;     // one per block
  Csr<T> E, F;                       // E: m x nloc, F: nloc x m (levels < nlev-1)
  Csr<T> C;                          // m x m lower-right block (setup; kept on level 0 for the inner iteration)
  int rank = 0;                      // k
  std::vector<T> W, WH;              // m x k column-major
  std::vector<T> Hk;                 // k x k: (I - H)^{-1} - I
  // factors of all blocks concatenated (block-diagonal, level-local numbering)
  Csr<T> Lcat, Ucat;
  std::vector<T> Dcat;
};

template <typename T>
struct HostOp {
  int n = 0;
  Csr<T> A;                          // original matrix
  std::vector<int> perm;             // (P b)[i] = b[perm[i]]
  int nlev = 0;
  std::vector<int> lev_ptr;          // nlev+1
  std::vector<HostLevel<T>> levels;
  GemslrParams prm;
  bool solve_phase = false;  // set when the setup is complete: SolveLevel may use the inner iteration
  double setup_seconds[4] = {0, 0, 0, 0};  // partition, ilut, lowrank, total
  int nloc(int l) const { return lev_ptr[l + 1] - lev_ptr[l]; }
  int nrem(int l) const { return n - lev_ptr[l + 1]; }
};

// GemslrClass::Setup (ParGeMSLR/SRC/preconditioners/gemslr.cpp:1602-1711).
// Returns 0 or an error code (PARGEMSLR_ERROR_* numbering).
template <typename T>
int gemslr_setup(const Csr<T>& A, const GemslrParams& prm, HostOp<T>& op, std::string* msg = nullptr);

// The m-length vector work of the thick-restart Arnoldi low-rank build of
// level l: basis columns V(:, 0..ncols-1) of length m = nrem(l) and the
// operator G x = (I - C Ct^{-1}) x + E B^{-1} F Ct^{-1} x
// (GemslrClass::EBFCMatVec, gemslr.cpp:3737-3792).  The setup's host code and
// the device build supply it; lowrank_level_with drives it (the Hessenberg
// matrix, Schur forms and restart choices stay on the host).
template <typename T>
struct LrBackend {
  virtual ~LrBackend() = default;
  virtual void begin(int l, int m, int ncols) = 0;
  virtual void set_col(int j, const T* x) = 0;            // V(:,j) = x (host array of m values)
  virtual void apply(int j) = 0;                          // V(:,j+1) = G V(:,j)
  virtual void dots(int j, T* h) = 0;                     // h[q] = V(:,q)^H V(:,j+1), q = 0..j (host array)
  virtual void update(int j, const T* h) = 0;             // V(:,j+1) -= V(:,0:j) h
  virtual double norm(int j) = 0;                         // ||V(:,j)||
  virtual void divide(int j, double d) = 0;               // V(:,j) /= d
  virtual void restart(int mcur, int p, const T* Y) = 0;  // V(:,0:p) = V(:,0:mcur) Y, then V(:,p) = V(:,mcur)
  // W = V(:,0:mcur) Y and WH = W R, column-major m x r host arrays
  virtual void finish(int mcur, int r, const T* Y, const T* R, std::vector<T>& W, std::vector<T>& WH) = 0;
};
template <typename T>
int lowrank_level_with(HostOp<T>& op, int l, LrBackend<T>& be);

// Low-rank build on a device (installed by the device code when the library
// has it): the corrections of every level of a setup with lowrank_device >= 0,
// bottom level first.  Returns 0 or an error code.
using LowrankDeviceFn = int (*)(HostOp<double>&, int, std::string*);
using LowrankDeviceFnZ = int (*)(HostOp<zdouble>&, int, std::string*);
extern LowrankDeviceFn g_lowrank_device;
extern LowrankDeviceFnZ g_lowrank_device_z;

// Lcat / Ucat / Dcat of level l from its per-block factors (part of the setup;
// also run when an operator is read back from a file).
template <typename T>
void build_level_cat(HostOp<T>& op, int l);

// Operator files: the complete output of the setup (matrix, permutation,
// levels, per-block ILUT factors, couplings, low-rank terms, parameters) in
// a little-endian binary layout closed by a 64-bit checksum.  save writes to
// path + ".tmp" and renames, so a reader never sees a partial file.  load
// checks the checksum and the structure (CSR shapes and index ranges, strict
// triangles, permutation, level and block offsets) before it returns 0.
template <typename T>
int op_save(const HostOp<T>& op, const char* path);
// Reads the value kind stored in a file: 0 double, 1 complex.
int op_file_kind(const char* path, int* is_complex);
template <typename T>
int op_load(const char* path, HostOp<T>& op, std::string* msg = nullptr);

// Host apply z = M^{-1} v (GemslrClass::Solve, gemslr.cpp:3431-3467); used by
// the setup's Arnoldi process.  Not the product solve path.
template <typename T>
struct HostApplyWork {
  std::vector<std::vector<T>> z;    // per level, length n - lev_ptr[l]
  std::vector<T> r0, x0, t;
};
template <typename T>
void host_solve_level(const HostOp<T>& op, int lev, const T* rhs, T* x, HostApplyWork<T>& w);
template <typename T>
void host_apply(const HostOp<T>& op, const T* v, T* z, HostApplyWork<T>& w);

// Reads the reference-layout parameter array (double[180]).
void params_from_array(const double* p, GemslrParams& g);
void params_default_array(double* p);

}  // namespace pgs

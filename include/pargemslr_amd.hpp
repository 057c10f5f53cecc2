// C++ class surface of pargemslr_amd, header-only over the C ABI of
// pargemslr_amd.h.  Mirrors the reference classes a sequential (np = 1)
// driver uses, with the same template parameters and typedef names, so that
// such a driver switches by changing its include and namespace:
//   SequentialVectorClass<T>  ParGeMSLR/SRC/vectors/sequential_vector.hpp (typedefs :575-579)
//   CsrMatrixClass<T>         ParGeMSLR/SRC/matrices/csr_matrix.hpp:26-716 (MatVec 562-575, typedefs 718-721)
//   GemslrClass<M, V, T>      ParGeMSLR/SRC/preconditioners/gemslr.hpp:1511-3190 (typedefs 3193-3196)
//   IluClass<M, V, T>         ParGeMSLR/SRC/preconditioners/ilu.hpp:118-1206 (typedefs 1208-1211)
//   FlexGmresClass<M, V, T>   ParGeMSLR/SRC/solvers/fgmres.hpp:27-688 (typedefs 690-693)
// with the SolverClass contract (ParGeMSLR/SRC/solvers/solver.hpp:45-534):
// Setup(x, rhs) / Solve(x, rhs) / SetSolveLocation / SetWithParameterArray,
// int return codes, non-owning matrix and preconditioner pointers.
//
// DataType is double or complexd (std::complex<double>, the layout of the
// reference's complexd: real part, then imaginary part).  A vector lives on
// the host (kMemoryHost) or on the device (kMemoryDevice); the solver's
// location decides where the solve runs: kMemoryDevice runs the HIP path,
// kMemoryHost runs the library's host apply (GemslrClass::Solve only).
#ifndef PARGEMSLR_AMD_HPP
#define PARGEMSLR_AMD_HPP

#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "pargemslr_amd.h"

namespace pargemslr_amd {

typedef std::complex<double> complexd;

// memory locations (ParGeMSLR/SRC/utils/memory.hpp:21-27)
enum MemoryLocationEnum { kMemoryHost = 0, kMemoryDevice = 1 };

template <typename T>
struct ValueTraits {
  static_assert(std::is_same<T, double>::value || std::is_same<T, complexd>::value,
                "values are double or complexd");
  static constexpr int is_complex = std::is_same<T, complexd>::value ? 1 : 0;
  // the C ABI passes (re, im) pairs of doubles
  static double* raw(T* p) { return reinterpret_cast<double*>(p); }
  static const double* raw(const T* p) { return reinterpret_cast<const double*>(p); }
};

template <typename T>
class SequentialVectorClass {
 public:
  SequentialVectorClass() = default;
  SequentialVectorClass(const SequentialVectorClass&) = delete;
  SequentialVectorClass& operator=(const SequentialVectorClass&) = delete;
  ~SequentialVectorClass() { Clear(); }

  int Setup(int n, int location = kMemoryHost, bool setzero = true) {
    Clear();
    n_ = n;
    location_ = location;
    own_ = true;
    if (location == kMemoryHost) {
      host_.assign((size_t)n, T(0.0));
      data_ = host_.data();
      return 0;
    }
    void* p = nullptr;
    int err = pgs_dev_malloc(device_, (int64_t)n * (int64_t)sizeof(T), &p);
    if (err) return err;
    data_ = static_cast<T*>(p);
    return setzero ? pgs_memset_zero(device_, data_, (int64_t)n * (int64_t)sizeof(T), nullptr) : 0;
  }
  // non-owning view (SequentialVectorClass::SetupPtr, sequential_vector.cpp:337-395)
  int SetupPtr(T* data, int n, int location) {
    Clear();
    data_ = data;
    n_ = n;
    location_ = location;
    own_ = false;
    return 0;
  }
  void Clear() {
    if (own_ && location_ == kMemoryDevice && data_) pgs_dev_free(device_, data_);
    host_.clear();
    data_ = nullptr;
    n_ = 0;
    own_ = false;
  }
  T* GetData() const { return data_; }
  int GetLengthLocal() const { return n_; }
  int GetDataLocation() const { return location_; }
  int Fill(T v) {
    if (location_ == kMemoryHost) {
      for (int i = 0; i < n_; ++i) data_[i] = v;
      return 0;
    }
    std::vector<T> tmp((size_t)n_, v);
    return pgs_memcpy_htod(device_, data_, tmp.data(), (int64_t)n_ * (int64_t)sizeof(T), nullptr);
  }
  // copies to / from host arrays regardless of location
  int CopyFromHost(const T* src) {
    if (location_ == kMemoryHost) {
      std::memcpy(static_cast<void*>(data_), src, sizeof(T) * n_);
      return 0;
    }
    return pgs_memcpy_htod(device_, data_, src, (int64_t)n_ * (int64_t)sizeof(T), nullptr);
  }
  int CopyToHost(T* dst) const {
    if (location_ == kMemoryHost) {
      std::memcpy(static_cast<void*>(dst), data_, sizeof(T) * n_);
      return 0;
    }
    return pgs_memcpy_dtoh(device_, dst, data_, (int64_t)n_ * (int64_t)sizeof(T), nullptr);
  }
  // SequentialVectorClass::ReadFromMMFile (sequential_vector.cpp:1192-1195):
  // a MatrixMarket coordinate vector, on the host
  int ReadFromMMFile(const char* vecfile, int idxin) {
    int n = 0;
    double* v = nullptr;
    int err = pgs_read_mm_vector(vecfile, idxin, ValueTraits<T>::is_complex, &n, &v);
    if (err) return err;
    err = Setup(n, kMemoryHost);
    if (!err) std::memcpy(static_cast<void*>(data_), v, sizeof(T) * (size_t)n);
    pgs_free(v);
    return err;
  }
  // SequentialVectorClass::MoveData
  int MoveData(int location) {
    if (location == location_) return 0;
    std::vector<T> tmp((size_t)n_);
    int err = CopyToHost(tmp.data());
    if (err) return err;
    const int n = n_;
    err = Setup(n, location, false);
    if (err) return err;
    return CopyFromHost(tmp.data());
  }

 private:
  T* data_ = nullptr;
  int n_ = 0;
  int location_ = kMemoryHost;
  int device_ = 0;
  bool own_ = false;
  std::vector<T> host_;
};

template <class MatrixType, class VectorType, typename DataType>
class GemslrClass;

template <typename T>
class CsrMatrixClass {
 public:
  // copies a 0-based CSR (host arrays)
  int Setup(int n, const int* ia, const int* ja, const T* a) {
    n_ = n;
    ia_.assign(ia, ia + n + 1);
    ja_.assign(ja, ja + ia[n]);
    a_.assign(a, a + ia[n]);
    return 0;
  }
  // CooMatrixClass::Laplacian (coo_matrix.cpp:519-640), argument order of the
  // reference; shift and the convection coefficients are T
  int Laplacian(int nx, int ny, int nz, T alphax, T alphay, T alphaz, T shift) {
    int n = 0;
    int64_t nnz = 0;
    int *ia = nullptr, *ja = nullptr;
    double* a = nullptr;
    const complexd s(shift), ax(alphax), ay(alphay), az(alphaz);
    // the generator takes real convection coefficients
    if (ax.imag() != 0.0 || ay.imag() != 0.0 || az.imag() != 0.0) return PGS_ERROR_INVALID_PARAM;
    int err = pgs_gen_laplacian(nx, ny, nz, s.real(), s.imag(), ax.real(), ay.real(), az.real(),
                                ValueTraits<T>::is_complex, &n, &nnz, &ia, &ja, &a);
    if (err) return err;
    err = Setup(n, ia, ja, reinterpret_cast<const T*>(a));
    pgs_free(ia);
    pgs_free(ja);
    pgs_free(a);
    return err;
  }
  // CsrMatrixClass::ReadFromMMFile (csr_matrix.cpp:1706-1714): a MatrixMarket
  // coordinate matrix, indices based at idxin; the matrix must be square
  int ReadFromMMFile(const char* matfile, int idxin) {
    int nr = 0, nc = 0;
    int64_t nnz = 0;
    int *ia = nullptr, *ja = nullptr;
    double* a = nullptr;
    int err = pgs_read_mm_csr(matfile, idxin, ValueTraits<T>::is_complex, &nr, &nc, &nnz, &ia, &ja, &a);
    if (err) return err;
    err = nr == nc ? Setup(nr, ia, ja, reinterpret_cast<const T*>(a)) : PGS_ERROR_INVALID_PARAM;
    pgs_free(ia);
    pgs_free(ja);
    pgs_free(a);
    return err;
  }
  int GetNumRowsLocal() const { return n_; }
  int64_t GetNumNonzeros() const { return (int64_t)ja_.size(); }
  const int* GetI() const { return ia_.data(); }
  const int* GetJ() const { return ja_.data(); }
  const T* GetData() const { return a_.data(); }
  // y = alpha * A * x + beta * y, trans 'N' (CsrMatrixClass::MatVec,
  // csr_matrix.cpp:1340-1365).  Host vectors run on the host; device vectors
  // need a device operator from a GemslrClass set up on this matrix.
  int MatVec(char trans, T alpha, const SequentialVectorClass<T>& x, T beta, SequentialVectorClass<T>& y) const {
    if (trans != 'N' && trans != 'n') return PGS_ERROR_INVALID_OPTION;
    if (x.GetLengthLocal() != n_ || y.GetLengthLocal() != n_) return PGS_ERROR_INVALID_PARAM;
    if (x.GetDataLocation() == kMemoryDevice && y.GetDataLocation() == kMemoryDevice) {
      if (!dop_) return PGS_ERROR_MEMORY_LOCATION;
      const complexd a(alpha), b(beta);
      const double ab[2] = {a.real(), a.imag()}, bb[2] = {b.real(), b.imag()};
      return pgs_spmv(dop_, ws_, ab, ValueTraits<T>::raw(x.GetData()), bb, ValueTraits<T>::raw(y.GetData()), 1,
                      nullptr);
    }
    if (x.GetDataLocation() != kMemoryHost || y.GetDataLocation() != kMemoryHost) return PGS_ERROR_MEMORY_LOCATION;
    const T* xv = x.GetData();
    T* yv = y.GetData();
    for (int i = 0; i < n_; ++i) {
      T s(0.0);
      for (int k = ia_[i]; k < ia_[i + 1]; ++k) s += a_[k] * xv[ja_[k]];
      yv[i] = (beta == T(0.0) ? T(0.0) : beta * yv[i]) + alpha * s;
    }
    return 0;
  }

 private:
  template <class, class, typename>
  friend class GemslrClass;
  int n_ = 0;
  std::vector<int> ia_, ja_;
  std::vector<T> a_;
  mutable pgs_dev_op_t dop_ = nullptr;  // set by GemslrClass::Setup on the device
  mutable pgs_ws_t ws_ = nullptr;
};

// Preconditioner slot (SolverClass): Setup builds the operator on the host
// and uploads it; Solve applies z = M^{-1} rhs.
template <class MatrixType, class VectorType, typename DataType>
class GemslrClass {
 public:
  using VT = ValueTraits<DataType>;
  GemslrClass() { pgs_params_default(params_); }
  GemslrClass(const GemslrClass&) = delete;
  GemslrClass& operator=(const GemslrClass&) = delete;
  virtual ~GemslrClass() { Clear(); }
  void Clear() {
    if (matrix_ && matrix_->dop_ == dop_) {
      matrix_->dop_ = nullptr;
      matrix_->ws_ = nullptr;
    }
    if (ws_) pgs_ws_destroy(ws_);
    if (dop_) pgs_dev_op_destroy(dop_);
    if (host_) pgs_gemslr_destroy(host_);
    ws_ = nullptr;
    dop_ = nullptr;
    host_ = nullptr;
    ready_ = false;
  }
  int SetMatrix(MatrixType& A) {
    Clear();
    matrix_ = &A;
    return 0;
  }
  // GemslrClass::SetWithParameterArray (gemslr.hpp:2069-2161)
  virtual int SetWithParameterArray(const double* params) {
    std::memcpy(params_, params, sizeof(params_));
    return 0;
  }
  int SetSolveLocation(int location) {
    location_ = location;
    return 0;
  }
  // PGS_PREC_FP64 (default: double / complex128) or PGS_PREC_FP32 (float /
  // complex64) arithmetic for the device apply
  int SetApplyPrecision(int prec) {
    prec_ = prec;
    return 0;
  }
  void SetSeed(uint64_t seed) { seed_ = seed; }
  // Builds the low-rank corrections of the next Setup on this device
  // (pgs_gemslr_setup_device); -1 (the default) builds them on the host.
  void SetLowRankDevice(int device) { lr_device_ = device; }
  // Writes the host operator (after Setup or Load) to an operator file.
  int Save(const char* path) const {
    if (!host_) return PGS_ERROR_FUNCTION_CALL;
    return pgs_gemslr_save(host_, path);
  }
  // Reads an operator file written by Save for the matrix given to SetMatrix
  // (the file's matrix must equal it); the next Setup uploads it without
  // running the host setup.
  int Load(const char* path) {
    if (!matrix_) return PGS_ERROR_INVALID_PARAM;
    Clear();
    int err = pgs_gemslr_load(path, &host_);
    if (err) return err;
    pgs_gemslr_info_t info;
    int n = 0;
    const int *ia = nullptr, *ja = nullptr;
    const double* a = nullptr;
    err = pgs_gemslr_info(host_, &info);
    if (!err) err = pgs_gemslr_matrix(host_, &n, &ia, &ja, &a);
    const MatrixType& A = *matrix_;
    const bool same = !err && info.is_complex == VT::is_complex && n == A.n_ && ia[n] == A.ia_[n] &&
                      std::equal(ia, ia + n + 1, A.ia_.begin()) && std::equal(ja, ja + ia[n], A.ja_.begin()) &&
                      std::memcmp(a, A.a_.data(), sizeof(DataType) * (size_t)ia[n]) == 0;
    if (!same) {
      pgs_gemslr_destroy(host_);
      host_ = nullptr;
      return err ? err : PGS_ERROR_INVALID_PARAM;
    }
    return 0;
  }
  virtual int Setup(VectorType&, VectorType&) {
    if (ready_) return 0;
    if (!matrix_) return PGS_ERROR_INVALID_PARAM;
    int err = 0;
    if (!host_ && lr_device_ < 0)
      err = pgs_gemslr_setup(matrix_->n_, matrix_->ia_.data(), matrix_->ja_.data(), VT::raw(matrix_->a_.data()),
                             VT::is_complex, params_, seed_, &host_);
    else if (!host_)
      err = pgs_gemslr_setup_device(matrix_->n_, matrix_->ia_.data(), matrix_->ja_.data(),
                                    VT::raw(matrix_->a_.data()), VT::is_complex, params_, seed_, lr_device_, &host_);
    if (err) return err;
    if (location_ == kMemoryDevice) {
      err = pgs_dev_op_create(host_, 0, PGS_PREC_FP64 | PGS_PREC_FP32, &dop_);
      if (err) return err;
      err = pgs_ws_create(dop_, 1, kdim_ws_, &ws_);
      if (err) return err;
      matrix_->dop_ = dop_;
      matrix_->ws_ = ws_;
    }
    ready_ = true;
    return 0;
  }
  // z = M^{-1} rhs (GemslrClass::Solve, gemslr.cpp:3431-3467)
  virtual int Solve(VectorType& x, VectorType& rhs) {
    if (!ready_) return PGS_ERROR_FUNCTION_CALL;
    if (x.GetLengthLocal() != matrix_->n_ || rhs.GetLengthLocal() != matrix_->n_) return PGS_ERROR_INVALID_PARAM;
    if (x.GetDataLocation() == kMemoryDevice && rhs.GetDataLocation() == kMemoryDevice) {
      if (!dop_) return PGS_ERROR_MEMORY_LOCATION;
      return pgs_apply(dop_, ws_, VT::raw(rhs.GetData()), VT::raw(x.GetData()), 1, prec_, nullptr);
    }
    if (x.GetDataLocation() == kMemoryHost && rhs.GetDataLocation() == kMemoryHost)
      return pgs_gemslr_host_apply(host_, VT::raw(rhs.GetData()), VT::raw(x.GetData()));
    return PGS_ERROR_MEMORY_LOCATION;
  }
  // nnz of the ILU factors and of the low-rank terms (GemslrClass::GetNumNonzeros)
  int64_t GetNumNonzeros(int64_t& nnz_ilu, int64_t& nnz_lr) const {
    pgs_gemslr_info_t inf;
    if (!host_ || pgs_gemslr_info(host_, &inf)) return 0;
    nnz_ilu = inf.nnz_ilu;
    nnz_lr = inf.nnz_lr;
    return inf.nnz_ilu + inf.nnz_lr;
  }
  pgs_gemslr_t GetHostOperator() const { return host_; }
  pgs_dev_op_t GetDeviceOperator() const { return dop_; }
  int GetApplyPrecision() const { return prec_; }
  void SetWorkspaceKrylovDimension(int kdim) { kdim_ws_ = kdim; }

 protected:
  MatrixType* matrix_ = nullptr;
  double params_[PGS_IO_SIZE];
  int location_ = kMemoryDevice;
  int prec_ = PGS_PREC_FP64;
  int kdim_ws_ = 50;
  uint64_t seed_ = 0;
  int lr_device_ = -1;
  bool ready_ = false;
  pgs_gemslr_t host_ = nullptr;
  pgs_dev_op_t dop_ = nullptr;
  pgs_ws_t ws_ = nullptr;
};

// ILUT preconditioner of the whole matrix (IluClass with kIluOptionILUT):
// a one-level GeMSLR operator, i.e. a single ILUT of P A P^T with the RCM
// ordering P.  Droptol / max row nnz follow PGS_IO_ILU_DROPTOL_B_GLOBAL /
// PGS_IO_ILU_ROWNNZ_B_GLOBAL.
template <class MatrixType, class VectorType, typename DataType>
class IluClass : public GemslrClass<MatrixType, VectorType, DataType> {
  using Base = GemslrClass<MatrixType, VectorType, DataType>;

 public:
  IluClass() {
    this->params_[PGS_IO_PREPOSS_NLEV_GLOBAL] = 1;
    this->params_[PGS_IO_SCHUR_ENABLE] = 0;
  }
  int SetDropTolerance(double tol) {
    this->params_[PGS_IO_ILU_DROPTOL_B_GLOBAL] = tol;
    this->params_[PGS_IO_ILU_DROPTOL_C_GLOBAL] = tol;
    return 0;
  }
  int SetMaxNnzPerRow(int lfil) {
    this->params_[PGS_IO_ILU_ROWNNZ_B_GLOBAL] = lfil;
    this->params_[PGS_IO_ILU_ROWNNZ_C_GLOBAL] = lfil;
    return 0;
  }
  int SetWithParameterArray(const double* params) override {
    Base::SetWithParameterArray(params);
    this->params_[PGS_IO_PREPOSS_NLEV_GLOBAL] = 1;
    this->params_[PGS_IO_SCHUR_ENABLE] = 0;
    this->params_[PGS_IO_ILU_DROPTOL_C_GLOBAL] = this->params_[PGS_IO_ILU_DROPTOL_B_GLOBAL];
    this->params_[PGS_IO_ILU_ROWNNZ_C_GLOBAL] = this->params_[PGS_IO_ILU_ROWNNZ_B_GLOBAL];
    return 0;
  }
};

// Flexible GMRES(m) with right preconditioning (FlexGmresClass,
// fgmres.hpp:230-558).  Runs on the device (x, rhs device vectors).
template <class MatrixType, class VectorType, typename DataType>
class FlexGmresClass {
 public:
  using VT = ValueTraits<DataType>;
  using PrecondType = GemslrClass<MatrixType, VectorType, DataType>;
  int SetMatrix(MatrixType& A) {
    matrix_ = &A;
    ready_ = false;
    return 0;
  }
  int SetPreconditioner(PrecondType& P) {
    precond_ = &P;
    ready_ = false;
    return 0;
  }
  // FlexGmresClass::SetWithParameterArray (fgmres.hpp:595-603)
  int SetWithParameterArray(const double* params) {
    tol_ = params[PGS_IO_SOLVER_TOL];
    maxits_ = (int)params[PGS_IO_SOLVER_MAXITS];
    kdim_ = (int)params[PGS_IO_SOLVER_KDIM];
    absolute_tol_ = params[PGS_IO_SOLVER_ATOL] != 0.0;
    return 0;
  }
  int SetKrylovSubspaceDimension(int kdim) {
    kdim_ = kdim;
    return 0;
  }
  int SetMaxNumberIterations(int maxits) {
    maxits_ = maxits;
    return 0;
  }
  int SetTolerance(double tol) {
    tol_ = tol;
    return 0;
  }
  int SetAbsoluteTol(bool abs) {
    absolute_tol_ = abs;
    return 0;
  }
  int SetSolveLocation(int location) {
    location_ = location;
    return 0;
  }
  // PGS_ORTH_MGS (the reference's modified Gram-Schmidt) or PGS_ORTH_CGS2
  int SetOrthogonalization(int orth) {
    orth_ = orth;
    return 0;
  }
  int Setup(VectorType& x, VectorType& rhs) {
    if (ready_) return 0;
    if (!matrix_) return PGS_ERROR_INVALID_PARAM;
    if (precond_) {
      precond_->SetSolveLocation(location_);
      precond_->SetWorkspaceKrylovDimension(kdim_);
      int err = precond_->Setup(x, rhs);
      if (err) return err;
    }
    ready_ = true;
    return 0;
  }
  int Solve(VectorType& x, VectorType& rhs) {
    if (!ready_) return PGS_ERROR_FUNCTION_CALL;
    if (location_ != kMemoryDevice || x.GetDataLocation() != kMemoryDevice ||
        rhs.GetDataLocation() != kMemoryDevice)
      return PGS_ERROR_MEMORY_LOCATION;
    pgs_dev_op_t dop = precond_ ? precond_->GetDeviceOperator() : nullptr;
    if (!dop) return PGS_ERROR_FUNCTION_CALL;
    if (!ws_ || ws_kdim_ < kdim_) {
      if (ws_) pgs_ws_destroy(ws_);
      ws_ = nullptr;
      int err = pgs_ws_create(dop, 1, kdim_, &ws_);
      if (err) return err;
      ws_kdim_ = kdim_;
    }
    pgs_fgmres_opts_t o;
    o.kdim = kdim_;
    o.maxits = maxits_;
    o.tol = tol_;
    o.absolute_tol = absolute_tol_ ? 1 : 0;
    o.prec = precond_->GetApplyPrecision();
    o.orth = orth_;
    o.use_precond = 1;
    o.check_every = 1;
    o.use_graph = 0;
    pgs_fgmres_result_t r;
    int err = pgs_fgmres(dop, ws_, VT::raw(rhs.GetData()), VT::raw(x.GetData()), 1, &o, &r, nullptr);
    if (err) return err;
    iters_ = r.iters;
    rel_res_ = r.rel_res;
    status_ = r.status;
    return 0;
  }
  int GetNumberIterations() const { return iters_; }
  double GetFinalRelativeResidual() const { return rel_res_; }
  int GetStatus() const { return status_; }
  ~FlexGmresClass() {
    if (ws_) pgs_ws_destroy(ws_);
  }

 private:
  MatrixType* matrix_ = nullptr;
  PrecondType* precond_ = nullptr;
  int kdim_ = 50, maxits_ = 100;
  double tol_ = 1e-6;
  bool absolute_tol_ = false;
  int location_ = kMemoryDevice;
  int orth_ = PGS_ORTH_CGS2;
  bool ready_ = false;
  pgs_ws_t ws_ = nullptr;
  int ws_kdim_ = 0;
  int iters_ = 0, status_ = 0;
  double rel_res_ = 0.0;
};

// the reference's typedef names (double and complexd instantiations)
typedef SequentialVectorClass<double> vector_seq_double;
typedef SequentialVectorClass<complexd> vector_seq_complexd;
typedef CsrMatrixClass<double> matrix_csr_double;
typedef CsrMatrixClass<complexd> matrix_csr_complexd;
typedef GemslrClass<CsrMatrixClass<double>, SequentialVectorClass<double>, double> precond_gemslr_csr_seq_double;
typedef GemslrClass<CsrMatrixClass<complexd>, SequentialVectorClass<complexd>, complexd>
    precond_gemslr_csr_seq_complexd;
typedef IluClass<CsrMatrixClass<double>, SequentialVectorClass<double>, double> precond_ilu_csr_seq_double;
typedef IluClass<CsrMatrixClass<complexd>, SequentialVectorClass<complexd>, complexd> precond_ilu_csr_seq_complexd;
typedef FlexGmresClass<CsrMatrixClass<double>, SequentialVectorClass<double>, double> fgmres_csr_seq_double;
typedef FlexGmresClass<CsrMatrixClass<complexd>, SequentialVectorClass<complexd>, complexd> fgmres_csr_seq_complexd;

}  // namespace pargemslr_amd

#endif  // PARGEMSLR_AMD_HPP

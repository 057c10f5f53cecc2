// C++ class surface of pargemslr_amd, header-only over the C ABI of
// pargemslr_amd.h.  Mirrors the reference classes a one-process (np = 1)
// driver uses, with the same template parameters, method names, argument
// meanings, defaults and typedef names, so that such a driver switches by
// changing its include and namespace:
//   SolverClass<M, V, T>            ParGeMSLR/SRC/solvers/solver.hpp:45-534
//   SequentialVectorClass<T>        ParGeMSLR/SRC/vectors/sequential_vector.hpp (typedefs :575-579)
//   CsrMatrixClass<T>               ParGeMSLR/SRC/matrices/csr_matrix.hpp:26-716 (MatVec 562-575, typedefs 718-721)
//   GemslrClass<M, V, T>            ParGeMSLR/SRC/preconditioners/gemslr.hpp:1511-3190 (typedefs 3193-3196)
//   IluClass<M, V, T>               ParGeMSLR/SRC/preconditioners/ilu.hpp:118-1206 (typedefs 1208-1211)
//   FlexGmresClass<M, V, T>         ParGeMSLR/SRC/solvers/fgmres.hpp:27-688 (typedefs 690-693)
//   ParallelGemslrClass<M, V, T>    ParGeMSLR/SRC/preconditioners/parallel_gemslr.hpp:813 (one process)
//   ParallelCsrMatrixClass<T>, ParallelVectorClass<T> (one process)
// plus the driver utilities the reference's drivers call (pargemslr_global,
// parallel_log, PargemslrInit / PargemslrFinalize, PARGEMSLR_PRINT and the
// timing macros of ParGeMSLR/SRC/utils/{utils,parallel}.hpp).
//
// DataType is double or complexd (std::complex<double>, the layout of the
// reference's complexd: real part, then imaginary part).  A vector lives on
// the host (kMemoryHost) or on the device (kMemoryDevice).  A solver's
// location decides where its solve runs:
// - GemslrClass::Solve on device vectors is the HIP apply, on host vectors the
//   library's host apply;
// - FlexGmresClass::Solve on device vectors with a GemslrClass-family
//   preconditioner runs the whole FGMRES on the device (pgs_fgmres); with any
//   other SolverClass preconditioner, or on host vectors, it runs the
//   reference's FGMRES loop (fgmres.hpp:263-558) in this header over the
//   vectors' own operations (device kernels for device vectors).
#ifndef PARGEMSLR_AMD_HPP
#define PARGEMSLR_AMD_HPP

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <type_traits>
#include <vector>

#include "pargemslr_amd.h"

// status names of the reference (ParGeMSLR/INC/pargemslr.hpp:9-21), same values
#ifndef PARGEMSLR_SUCCESS
#define PARGEMSLR_SUCCESS PGS_SUCCESS
#define PARGEMSLR_ERROR_INVALED_OPTION PGS_ERROR_INVALID_OPTION
#define PARGEMSLR_ERROR_INVALED_PARAM PGS_ERROR_INVALID_PARAM
#define PARGEMSLR_ERROR_IO_ERROR PGS_ERROR_IO
#define PARGEMSLR_ERROR_ILU_EMPTY_ROW PGS_ERROR_ILU_EMPTY_ROW
#define PARGEMSLR_ERROR_FUNCTION_CALL_ERR PGS_ERROR_FUNCTION_CALL
#define PARGEMSLR_ERROR_MEMORY_LOCATION PGS_ERROR_MEMORY_LOCATION
#endif

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
  static void pair(const T& v, double out[2]) {
    const complexd z(v);
    out[0] = z.real();
    out[1] = z.imag();
  }
  static T from_pair(const double in[2]) {
    if constexpr (is_complex) return T(in[0], in[1]);
    else return in[0];
  }
};

// ---------------------------------------------------------------------------
// driver utilities
// ---------------------------------------------------------------------------
namespace detail {
inline double* default_params() {
  static double p[PGS_IO_SIZE];
  pgs_params_default(p);
  return p;
}
}  // namespace detail

// Process-wide settings of the reference's drivers (pargemslr_global,
// ParGeMSLR/SRC/utils/utils.hpp:370-390, utils.cpp:13-29): the parameter
// array and the random engine of Rand (mt19937 with
// uniform_int_distribution<int>(0, INT_MAX), the reference's generator, so a
// driver that seeds it with 0 draws the reference's values; the reference
// seeds it from std::random_device at start-up, this one with 0).
struct pargemslr_global {
  static inline double* _params = detail::default_params();
  static inline std::mt19937 _mersenne_twister_engine{0u};
  static inline std::uniform_int_distribution<int> _uniform_int_distribution{0, INT_MAX};
  static inline int _dash_line_width = 32;
  static inline FILE* _out_file = stdout;
};

// PargemslrValueRandHost (utils.cpp:221-245): U[0, 1] per real value, two
// draws (real, then imaginary part) per complex value
inline int PargemslrValueRandHost(double& a) {
  a = pargemslr_global::_uniform_int_distribution(pargemslr_global::_mersenne_twister_engine) /
      (double)pargemslr_global::_uniform_int_distribution.max();
  return PGS_SUCCESS;
}
inline int PargemslrValueRandHost(complexd& a) {
  double re = 0.0, im = 0.0;
  PargemslrValueRandHost(re);
  PargemslrValueRandHost(im);
  a = complexd(re, im);
  return PGS_SUCCESS;
}

#define PARGEMSLR_PRINT(...) std::fprintf(pargemslr_amd::pargemslr_global::_out_file, __VA_ARGS__)

// timer slots of the reference (ParGeMSLR/SRC/utils/parallel.hpp:28-65)
#define PARGEMSLR_TIMES_NUM 60
#define PARGEMSLR_BUILDTIME_PARTITION 0
#define PARGEMSLR_TOTAL_GEN_MAT_TIME 50
#define PARGEMSLR_TOTAL_SETUP_TIME 51
#define PARGEMSLR_TOTAL_SOLVE_TIME 52

// The one-process state of the reference's parallel_log (ParallelLogClass,
// ParGeMSLR/SRC/utils/parallel.hpp): one rank, plus the timer slots of the
// drivers' timing macros.  There is no MPI here; GetMpiInfo reports rank 0 of
// 1 and leaves the communicator argument as it is.
class ParallelLogClass {
 public:
  static inline int _gsize = 1;
  static inline int _grank = 0;
  static inline std::vector<double> _times = std::vector<double>(PARGEMSLR_TIMES_NUM, 0.0);
  template <class Comm>
  int GetMpiInfo(int& np, int& myid, Comm&) const {
    np = 1;
    myid = 0;
    return PGS_SUCCESS;
  }
  static double Wtime() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
};
typedef ParallelLogClass parallel_log;

// one process, no MPI: nothing to start or stop; the parameter array keeps
// what the driver has set
inline int PargemslrInit(int*, char***) { return PGS_SUCCESS; }
inline int PargemslrFinalize() { return PGS_SUCCESS; }
inline int PargemslrPrintParallelInfo() {
  PARGEMSLR_PRINT("Total number of MPI ranks: 1\n");
  return PGS_SUCCESS;
}
inline int PargemslrPrintDashLine(int width) {
  for (int i = 0; i < width; ++i) PARGEMSLR_PRINT("-");
  PARGEMSLR_PRINT("\n");
  return PGS_SUCCESS;
}

// the drivers' timing macros (parallel.hpp:73-200): wall time of a call added
// to a timer slot; the solve calls return after their device work has ended
#define PARGEMSLR_LOCAL_FIRM_TIME_CALL(num, ...)                                  \
  {                                                                               \
    const double _pgs_t0 = pargemslr_amd::ParallelLogClass::Wtime();              \
    (__VA_ARGS__);                                                                \
    pargemslr_amd::ParallelLogClass::_times[num] +=                               \
        pargemslr_amd::ParallelLogClass::Wtime() - _pgs_t0;                       \
  }
#define PARGEMSLR_GLOBAL_FIRM_TIME_CALL(num, ...) PARGEMSLR_LOCAL_FIRM_TIME_CALL(num, __VA_ARGS__)
#define PARGEMSLR_RESET_TIME \
  std::fill(pargemslr_amd::ParallelLogClass::_times.begin(), pargemslr_amd::ParallelLogClass::_times.end(), 0.0);
#define PARGEMSLR_PRINT_TIMING_RESULT(print_level, ...)                                                     \
  {                                                                                                         \
    if (__VA_ARGS__) {                                                                                      \
      const std::vector<double>& _pgs_t = pargemslr_amd::ParallelLogClass::_times;                          \
      PARGEMSLR_PRINT("\n");                                                                                \
      pargemslr_amd::PargemslrPrintDashLine(pargemslr_amd::pargemslr_global::_dash_line_width);             \
      PARGEMSLR_PRINT("Time info:\n");                                                                      \
      PARGEMSLR_PRINT("\tLoad matrix time:   %fs\n", _pgs_t[PARGEMSLR_TOTAL_GEN_MAT_TIME]);                \
      PARGEMSLR_PRINT("\tSetup time:         %fs\n", _pgs_t[PARGEMSLR_TOTAL_SETUP_TIME]);                  \
      PARGEMSLR_PRINT("\tSolve time:         %fs\n", _pgs_t[PARGEMSLR_TOTAL_SOLVE_TIME]);                  \
      PARGEMSLR_PRINT("\tTotal time:         %fs\n",                                                       \
                      _pgs_t[PARGEMSLR_TOTAL_SETUP_TIME] + _pgs_t[PARGEMSLR_TOTAL_SOLVE_TIME]);             \
      pargemslr_amd::PargemslrPrintDashLine(pargemslr_amd::pargemslr_global::_dash_line_width);             \
      PARGEMSLR_PRINT("\n");                                                                                \
      (void)(print_level);                                                                                  \
    }                                                                                                       \
  }

// ---------------------------------------------------------------------------
// vectors
// ---------------------------------------------------------------------------
template <typename T>
class SequentialVectorClass {
 public:
  using VT = ValueTraits<T>;
  SequentialVectorClass() = default;
  SequentialVectorClass(const SequentialVectorClass&) = delete;
  SequentialVectorClass& operator=(const SequentialVectorClass&) = delete;
  virtual ~SequentialVectorClass() { Clear(); }

  // Setup(length[, location][, setzero]) (sequential_vector.hpp): an owned
  // vector of n values on the host or the device
  int Setup(int n) { return Setup(n, kMemoryHost, true); }
  int Setup(int n, bool setzero) { return Setup(n, kMemoryHost, setzero); }
  int Setup(int n, int location, bool setzero) {
    Clear();
    if (n < 0 || (location != kMemoryHost && location != kMemoryDevice)) return PGS_ERROR_INVALID_PARAM;
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
    if (err) {
      n_ = 0;
      own_ = false;
      return err;
    }
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
  // re-points a view (UpdatePtr, sequential_vector.cpp)
  int UpdatePtr(T* data, int location) {
    if (own_) return PGS_ERROR_FUNCTION_CALL;
    data_ = data;
    location_ = location;
    return 0;
  }
  virtual int Clear() {
    if (own_ && location_ == kMemoryDevice && data_) pgs_dev_free(device_, data_);
    host_.clear();
    host_.shrink_to_fit();
    data_ = nullptr;
    n_ = 0;
    own_ = false;
    location_ = kMemoryHost;
    return 0;
  }
  T* GetData() const { return data_; }
  int GetLengthLocal() const { return n_; }
  int64_t GetLengthGlobal() const { return n_; }
  int GetDataLocation() const { return location_; }
  int GetDevice() const { return device_; }

  // x = v (Fill, sequential_vector.cpp:706-732)
  int Fill(const T& v) {
    if (location_ == kMemoryHost) {
      std::fill(data_, data_ + n_, v);
      return 0;
    }
    double p[2];
    VT::pair(v, p);
    return pgs_vec_fill(device_, n_, VT::is_complex, p, VT::raw(data_), nullptr);
  }
  // U[0, 1] values from the process-wide engine, drawn on the host in index
  // order (Rand, sequential_vector.cpp:739-748)
  int Rand() {
    if (location_ == kMemoryHost) {
      for (int i = 0; i < n_; ++i) PargemslrValueRandHost(data_[i]);
      return 0;
    }
    std::vector<T> tmp((size_t)n_);
    for (int i = 0; i < n_; ++i) PargemslrValueRandHost(tmp[i]);
    return CopyFromHost(tmp.data());
  }
  // x *= alpha (Scale, sequential_vector.cpp:1037-1040)
  int Scale(const T& alpha) {
    if (location_ == kMemoryHost) {
      for (int i = 0; i < n_; ++i) data_[i] *= alpha;
      return 0;
    }
    double p[2];
    VT::pair(alpha, p);
    return pgs_vec_scale(device_, n_, VT::is_complex, p, VT::raw(data_), nullptr);
  }
  // this += alpha x (Axpy, sequential_vector.cpp:1046-1060; vectorops.cpp:633-680)
  int Axpy(const T& alpha, const SequentialVectorClass& x) {
    if (x.n_ != n_) return PGS_ERROR_INVALID_PARAM;
    if (location_ != x.location_) return PGS_ERROR_MEMORY_LOCATION;
    if (location_ == kMemoryHost) {
      for (int i = 0; i < n_; ++i) data_[i] += alpha * x.data_[i];
      return 0;
    }
    double p[2];
    VT::pair(alpha, p);
    return pgs_vec_axpy(device_, n_, VT::is_complex, p, VT::raw(x.data_), VT::raw(data_), nullptr);
  }
  // t = sum_i conj(this_i) y_i (Dot, vectorops.cpp:93-146)
  int Dot(const SequentialVectorClass& y, T& t) const {
    if (y.n_ != n_) return PGS_ERROR_INVALID_PARAM;
    if (location_ != y.location_) return PGS_ERROR_MEMORY_LOCATION;
    if (location_ == kMemoryHost) {
      T s(0.0);
      for (int i = 0; i < n_; ++i) {
        if constexpr (VT::is_complex) s += std::conj(data_[i]) * y.data_[i];
        else s += data_[i] * y.data_[i];
      }
      t = s;
      return 0;
    }
    double r[2] = {0.0, 0.0};
    int err = pgs_vec_dot(device_, n_, VT::is_complex, VT::raw(data_), VT::raw(y.data_), r, nullptr);
    if (!err) t = VT::from_pair(r);
    return err;
  }
  // t = ||x||_2 (Norm2)
  int Norm2(double& t) const {
    T d(0.0);
    int err = Dot(*this, d);
    t = std::sqrt(std::max(0.0, std::real(d)));
    return err;
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
  // this = src, any pair of locations (PARGEMSLR_MEMCPY)
  int CopyFrom(const SequentialVectorClass& src) {
    if (src.n_ != n_) return PGS_ERROR_INVALID_PARAM;
    const int64_t bytes = (int64_t)n_ * (int64_t)sizeof(T);
    if (location_ == kMemoryHost && src.location_ == kMemoryHost) {
      std::memcpy(static_cast<void*>(data_), src.data_, (size_t)bytes);
      return 0;
    }
    if (location_ == kMemoryHost) return pgs_memcpy_dtoh(device_, data_, src.data_, bytes, nullptr);
    if (src.location_ == kMemoryHost) return pgs_memcpy_htod(device_, data_, src.data_, bytes, nullptr);
    return pgs_memcpy_dtod(device_, data_, src.data_, bytes, nullptr);
  }
  // SequentialVectorClass::ReadFromMMFile (sequential_vector.cpp:1192-1195):
  // a MatrixMarket coordinate vector, on the host
  int ReadFromMMFile(const char* vecfile, int idxin) {
    int n = 0;
    double* v = nullptr;
    int err = pgs_read_mm_vector(vecfile, idxin, VT::is_complex, &n, &v);
    if (err) return err;
    err = Setup(n, kMemoryHost, false);
    if (!err) std::memcpy(static_cast<void*>(data_), v, sizeof(T) * (size_t)n);
    pgs_free(v);
    return err;
  }
  // One value per line, "%16.12f" (real) or "%16.12f %16.12f" (complex); host
  // vectors only (WriteToDisk, sequential_vector.cpp:984-1029)
  int WriteToDisk(const char* path) const {
    if (location_ == kMemoryDevice) return PGS_ERROR_MEMORY_LOCATION;
    FILE* f = std::fopen(path, "w");
    if (!f) return PGS_ERROR_IO;
    for (int i = 0; i < n_; ++i) {
      if constexpr (VT::is_complex) std::fprintf(f, "%16.12f %16.12f\n", data_[i].real(), data_[i].imag());
      else std::fprintf(f, "%16.12f \n", data_[i]);
    }
    return std::fclose(f) == 0 ? 0 : PGS_ERROR_IO;
  }
  // SequentialVectorClass::MoveData
  int MoveData(int location) {
    if (location == location_) return 0;
    if (!own_) return PGS_ERROR_FUNCTION_CALL;  // a view does not own its storage
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

// ---------------------------------------------------------------------------
// CSR matrix
// ---------------------------------------------------------------------------
template <typename T>
class CsrMatrixClass {
 public:
  using VT = ValueTraits<T>;
  CsrMatrixClass() = default;
  CsrMatrixClass(const CsrMatrixClass&) = delete;
  CsrMatrixClass& operator=(const CsrMatrixClass&) = delete;
  virtual ~CsrMatrixClass() { Clear(); }

  // copies a 0-based CSR (host arrays)
  int Setup(int n, const int* ia, const int* ja, const T* a) {
    Clear();
    if (n < 0 || !ia) return PGS_ERROR_INVALID_PARAM;
    n_ = n;
    ia_.assign(ia, ia + n + 1);
    ja_.assign(ja, ja + ia[n]);
    a_.assign(a, a + ia[n]);
    return 0;
  }
  // CooMatrixClass::Laplacian (coo_matrix.cpp:519-640), argument order of the
  // reference; shift and the convection coefficients are T.  rand_perturb
  // subtracts |shift| * Rand() from each diagonal entry, one draw per row in
  // row order from the process-wide engine, as the reference does.
  int Laplacian(int nx, int ny, int nz, T alphax, T alphay, T alphaz, T shift, bool rand_perturb = false) {
    int n = 0;
    int64_t nnz = 0;
    int *ia = nullptr, *ja = nullptr;
    double* a = nullptr;
    const complexd s(shift), ax(alphax), ay(alphay), az(alphaz);
    // the generator takes real convection coefficients
    if (ax.imag() != 0.0 || ay.imag() != 0.0 || az.imag() != 0.0) return PGS_ERROR_INVALID_PARAM;
    int err = pgs_gen_laplacian(nx, ny, nz, s.real(), s.imag(), ax.real(), ay.real(), az.real(), VT::is_complex, &n,
                                &nnz, &ia, &ja, &a);
    if (err) return err;
    err = Setup(n, ia, ja, reinterpret_cast<const T*>(a));
    pgs_free(ia);
    pgs_free(ja);
    pgs_free(a);
    if (err || !rand_perturb) return err;
    const double shift_abs = std::abs(s);
    for (int i = 0; i < n_; ++i) {
      double p = 0.0;
      PargemslrValueRandHost(p);
      for (int k = ia_[i]; k < ia_[i + 1]; ++k)
        if (ja_[k] == i) a_[k] -= T(shift_abs * p);
    }
    return 0;
  }
  // CsrMatrixClass::ReadFromMMFile (csr_matrix.cpp:1706-1714): a MatrixMarket
  // coordinate matrix, indices based at idxin; the matrix must be square
  int ReadFromMMFile(const char* matfile, int idxin) {
    int nr = 0, nc = 0;
    int64_t nnz = 0;
    int *ia = nullptr, *ja = nullptr;
    double* a = nullptr;
    int err = pgs_read_mm_csr(matfile, idxin, VT::is_complex, &nr, &nc, &nnz, &ia, &ja, &a);
    if (err) return err;
    err = nr == nc ? Setup(nr, ia, ja, reinterpret_cast<const T*>(a)) : PGS_ERROR_INVALID_PARAM;
    pgs_free(ia);
    pgs_free(ja);
    pgs_free(a);
    return err;
  }
  virtual int Clear() {
    if (dcsr_) pgs_dev_csr_destroy(dcsr_);
    dcsr_ = nullptr;
    dop_ = nullptr;
    ws_ = nullptr;
    n_ = 0;
    ia_.assign(1, 0);
    ja_.clear();
    a_.clear();
    location_ = kMemoryHost;
    return 0;
  }
  int GetNumRowsLocal() const { return n_; }
  int GetNumColsLocal() const { return n_; }
  int64_t GetNumRowsGlobal() const { return n_; }
  int64_t GetNumNonzeros() const { return (int64_t)ja_.size(); }
  const int* GetI() const { return ia_.data(); }
  const int* GetJ() const { return ja_.data(); }
  const T* GetData() const { return a_.data(); }
  int GetDataLocation() const { return location_; }
  // CsrMatrixClass::MoveData: kMemoryDevice keeps a device copy for MatVec on
  // device vectors (the host arrays stay: the setup reads them); kMemoryHost
  // releases it
  int MoveData(int location) {
    if (location == kMemoryHost) {
      if (dcsr_) pgs_dev_csr_destroy(dcsr_);
      dcsr_ = nullptr;
      location_ = kMemoryHost;
      return 0;
    }
    if (location != kMemoryDevice) return PGS_ERROR_INVALID_PARAM;
    if (!dcsr_) {
      int err = pgs_dev_csr_create(0, n_, ia_.data(), ja_.data(), VT::raw(a_.data()), VT::is_complex, &dcsr_);
      if (err) return err;
    }
    location_ = kMemoryDevice;
    return 0;
  }
  // y = alpha * A * x + beta * y, trans 'N' (CsrMatrixClass::MatVec,
  // csr_matrix.cpp:1340-1365).  Host vectors run on the host; device vectors
  // use the matrix's device copy (MoveData) or, without one, the operator of
  // a GemslrClass set up on this matrix at kMemoryDevice.
  int MatVec(char trans, const T& alpha, const SequentialVectorClass<T>& x, const T& beta,
             SequentialVectorClass<T>& y) const {
    if (trans != 'N' && trans != 'n') return PGS_ERROR_INVALID_OPTION;
    if (x.GetLengthLocal() != n_ || y.GetLengthLocal() != n_) return PGS_ERROR_INVALID_PARAM;
    if (x.GetDataLocation() == kMemoryDevice && y.GetDataLocation() == kMemoryDevice) {
      double ab[2], bb[2];
      VT::pair(alpha, ab);
      VT::pair(beta, bb);
      if (dcsr_) return pgs_dev_csr_spmv(dcsr_, ab, VT::raw(x.GetData()), bb, VT::raw(y.GetData()), nullptr);
      if (dop_) return pgs_spmv(dop_, ws_, ab, VT::raw(x.GetData()), bb, VT::raw(y.GetData()), 1, nullptr);
      return PGS_ERROR_MEMORY_LOCATION;
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
  // Set by a GemslrClass whose operator (which holds A) is on the device, and
  // cleared when that operator goes away.
  void BindDeviceOperator(pgs_dev_op_t dop, pgs_ws_t ws) const {
    dop_ = dop;
    ws_ = ws;
  }
  pgs_dev_op_t BoundDeviceOperator() const { return dop_; }

 private:
  int n_ = 0;
  std::vector<int> ia_{0}, ja_;
  std::vector<T> a_;
  int location_ = kMemoryHost;
  pgs_dev_csr_t dcsr_ = nullptr;
  mutable pgs_dev_op_t dop_ = nullptr;
  mutable pgs_ws_t ws_ = nullptr;
};

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------

// The solver / preconditioner contract (SolverClass, solver.hpp:45-534):
// Setup(x, rhs), Solve(x, rhs), SetSolveLocation, SetWithParameterArray;
// non-owning matrix and preconditioner pointers unless SetOwnMatrix /
// SetOwnPreconditioner(true); SetMatrix drops the preconditioner.
template <class MatrixType, class VectorType, typename DataType>
class SolverClass {
 public:
  SolverClass() = default;
  SolverClass(const SolverClass&) = delete;
  SolverClass& operator=(const SolverClass&) = delete;
  virtual ~SolverClass() { SolverClass::Clear(); }
  virtual int Clear() {
    if (matrix_ && own_matrix_) delete matrix_;
    matrix_ = nullptr;
    if (precond_ && own_precond_) delete precond_;
    precond_ = nullptr;
    own_matrix_ = own_precond_ = false;
    ready_ = false;
    return 0;
  }
  virtual int Setup(VectorType& x, VectorType& rhs) = 0;
  virtual int Solve(VectorType& x, VectorType& rhs) = 0;
  virtual int SetSolveLocation(const int& location) = 0;
  virtual int SetWithParameterArray(double*) { return PGS_ERROR_FUNCTION_CALL; }
  virtual long int GetNumNonzeros() { return PGS_ERROR_FUNCTION_CALL; }
  MatrixType* GetMatrix() { return matrix_; }
  SolverClass* GetPreconditioner() { return precond_; }
  int SetMatrix(MatrixType& matrix) { return SetMatrixP(&matrix); }
  int SetMatrixP(MatrixType* matrix) {
    if (matrix == matrix_) return 0;
    if (matrix_ && own_matrix_) delete matrix_;
    if (precond_ && own_precond_) delete precond_;
    matrix_ = matrix;
    precond_ = nullptr;
    own_matrix_ = own_precond_ = false;
    ready_ = false;
    OnMatrixChanged();
    return 0;
  }
  int SetPreconditioner(SolverClass& precond) { return SetPreconditionerP(&precond); }
  int SetPreconditionerP(SolverClass* precond) {
    if (precond_ && own_precond_ && precond_ != precond) delete precond_;
    precond_ = precond;
    own_precond_ = false;
    return 0;
  }
  bool IsReady() const { return ready_; }
  int SetOwnMatrix(bool own) {
    own_matrix_ = own;
    return 0;
  }
  int SetOwnPreconditioner(bool own) {
    own_precond_ = own;
    return 0;
  }
  int SetPrintOption(int print_option) {
    print_option_ = print_option;
    return 0;
  }

 protected:
  virtual void OnMatrixChanged() {}
  MatrixType* matrix_ = nullptr;
  SolverClass* precond_ = nullptr;
  bool own_matrix_ = false, own_precond_ = false;
  bool ready_ = false;
  int print_option_ = 0;
};

// Preconditioner slot: Setup builds the operator on the host and uploads it
// when the solve location is the device; Solve applies z = M^{-1} rhs.
template <class MatrixType, class VectorType, typename DataType>
class GemslrClass : public SolverClass<MatrixType, VectorType, DataType> {
 public:
  using VT = ValueTraits<DataType>;
  GemslrClass() { pgs_params_default(params_); }
  ~GemslrClass() override { Release(); }
  int Clear() override {
    Release();
    return SolverClass<MatrixType, VectorType, DataType>::Clear();
  }
  // GemslrClass::SetWithParameterArray (gemslr.hpp:2069-2161)
  int SetWithParameterArray(double* params) override {
    std::memcpy(params_, params, sizeof(params_));
    return 0;
  }
  int SetSolveLocation(const int& location) override {
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
    if (!this->matrix_) return PGS_ERROR_INVALID_PARAM;
    Release();
    int err = pgs_gemslr_load(path, &host_);
    if (err) return err;
    pgs_gemslr_info_t info;
    int n = 0;
    const int *ia = nullptr, *ja = nullptr;
    const double* a = nullptr;
    err = pgs_gemslr_info(host_, &info);
    if (!err) err = pgs_gemslr_matrix(host_, &n, &ia, &ja, &a);
    const MatrixType& A = *this->matrix_;
    const int nA = A.GetNumRowsLocal();
    const bool same = !err && info.is_complex == VT::is_complex && n == nA && ia[n] == A.GetI()[nA] &&
                      std::equal(ia, ia + n + 1, A.GetI()) && std::equal(ja, ja + ia[n], A.GetJ()) &&
                      std::memcmp(a, A.GetData(), sizeof(DataType) * (size_t)ia[n]) == 0;
    if (!same) {
      pgs_gemslr_destroy(host_);
      host_ = nullptr;
      return err ? err : PGS_ERROR_INVALID_PARAM;
    }
    return 0;
  }
  int Setup(VectorType&, VectorType&) override {
    if (this->ready_) return 0;
    if (!this->matrix_) return PGS_ERROR_INVALID_PARAM;
    const MatrixType& A = *this->matrix_;
    int err = 0;
    if (!host_ && lr_device_ < 0)
      err = pgs_gemslr_setup(A.GetNumRowsLocal(), A.GetI(), A.GetJ(), VT::raw(A.GetData()), VT::is_complex, params_,
                             seed_, &host_);
    else if (!host_)
      err = pgs_gemslr_setup_device(A.GetNumRowsLocal(), A.GetI(), A.GetJ(), VT::raw(A.GetData()), VT::is_complex,
                                    params_, seed_, lr_device_, &host_);
    if (err) return err;
    if (location_ == kMemoryDevice) {
      err = pgs_dev_op_create(host_, 0, PGS_PREC_FP64 | PGS_PREC_FP32, &dop_);
      if (err) return err;
      err = pgs_ws_create(dop_, 1, kdim_ws_, &ws_);
      if (err) return err;
      A.BindDeviceOperator(dop_, ws_);
    }
    this->ready_ = true;
    return 0;
  }
  // z = M^{-1} rhs (GemslrClass::Solve, gemslr.cpp:3431-3467)
  int Solve(VectorType& x, VectorType& rhs) override {
    if (!this->ready_) return PGS_ERROR_FUNCTION_CALL;
    const int n = this->matrix_->GetNumRowsLocal();
    if (x.GetLengthLocal() != n || rhs.GetLengthLocal() != n) return PGS_ERROR_INVALID_PARAM;
    if (x.GetDataLocation() == kMemoryDevice && rhs.GetDataLocation() == kMemoryDevice) {
      if (!dop_) return PGS_ERROR_MEMORY_LOCATION;
      return pgs_apply(dop_, ws_, VT::raw(rhs.GetData()), VT::raw(x.GetData()), 1, prec_, nullptr);
    }
    if (x.GetDataLocation() == kMemoryHost && rhs.GetDataLocation() == kMemoryHost)
      return pgs_gemslr_host_apply(host_, VT::raw(rhs.GetData()), VT::raw(x.GetData()));
    return PGS_ERROR_MEMORY_LOCATION;
  }
  // nnz of the ILU factors and of the low-rank terms (GemslrClass::GetNumNonzeros)
  long int GetNumNonzeros(long int& nnz_ilu, long int& nnz_lr) const {
    pgs_gemslr_info_t inf;
    if (!host_ || pgs_gemslr_info(host_, &inf)) return 0;
    nnz_ilu = (long int)inf.nnz_ilu;
    nnz_lr = (long int)inf.nnz_lr;
    return (long int)(inf.nnz_ilu + inf.nnz_lr);
  }
  long int GetNumNonzeros() override {
    long int a = 0, b = 0;
    return GetNumNonzeros(a, b);
  }
  pgs_gemslr_t GetHostOperator() const { return host_; }
  pgs_dev_op_t GetDeviceOperator() const { return dop_; }
  int GetApplyPrecision() const { return prec_; }
  void SetWorkspaceKrylovDimension(int kdim) { kdim_ws_ = kdim; }

 protected:
  void OnMatrixChanged() override { Release(); }
  void Release() {
    if (this->matrix_ && this->matrix_->BoundDeviceOperator() == dop_ && dop_)
      this->matrix_->BindDeviceOperator(nullptr, nullptr);
    if (ws_) pgs_ws_destroy(ws_);
    if (dop_) pgs_dev_op_destroy(dop_);
    if (host_) pgs_gemslr_destroy(host_);
    ws_ = nullptr;
    dop_ = nullptr;
    host_ = nullptr;
    this->ready_ = false;
  }
  double params_[PGS_IO_SIZE];
  int location_ = kMemoryHost;
  int prec_ = PGS_PREC_FP64;
  int kdim_ws_ = 50;
  uint64_t seed_ = 0;
  int lr_device_ = -1;
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
  int SetWithParameterArray(double* params) override {
    Base::SetWithParameterArray(params);
    this->params_[PGS_IO_PREPOSS_NLEV_GLOBAL] = 1;
    this->params_[PGS_IO_SCHUR_ENABLE] = 0;
    this->params_[PGS_IO_ILU_DROPTOL_C_GLOBAL] = this->params_[PGS_IO_ILU_DROPTOL_B_GLOBAL];
    this->params_[PGS_IO_ILU_ROWNNZ_C_GLOBAL] = this->params_[PGS_IO_ILU_ROWNNZ_B_GLOBAL];
    return 0;
  }
};

// Flexible GMRES(m) with right preconditioning (FlexGmresClass,
// fgmres.hpp:80-558): defaults kdim 50, maxits 200, tol 1e-8 relative,
// modified Gram-Schmidt, host location, as the reference.
template <class MatrixType, class VectorType, typename DataType>
class FlexGmresClass : public SolverClass<MatrixType, VectorType, DataType> {
 public:
  using VT = ValueTraits<DataType>;
  using Solver = SolverClass<MatrixType, VectorType, DataType>;
  using GemslrType = GemslrClass<MatrixType, VectorType, DataType>;
  ~FlexGmresClass() override { ReleaseWorkspace(); }
  int Clear() override {
    ReleaseWorkspace();
    Solver::Clear();
    location_ = kMemoryHost;
    iters_ = 0;
    status_ = 0;
    rel_res_ = 0.0;
    rel_res_vector_.clear();
    return 0;
  }
  // FlexGmresClass::SetWithParameterArray (fgmres.hpp:595-603)
  int SetWithParameterArray(double* params) override {
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
  int SetSolveLocation(const int& location) override {
    location_ = location;
    if (this->precond_) this->precond_->SetSolveLocation(location);
    return 0;
  }
  // PGS_ORTH_MGS (the reference's modified Gram-Schmidt, the default),
  // PGS_ORTH_CGS2 or PGS_ORTH_DCGS2 for the device solve
  int SetOrthogonalization(int orth) {
    orth_ = orth;
    return 0;
  }
  int Setup(VectorType& x, VectorType& rhs) override {
    if (this->ready_) return 0;
    if (!this->matrix_) return PGS_ERROR_INVALID_PARAM;
    if (this->precond_) {
      this->precond_->SetSolveLocation(location_);
      if (GemslrType* g = dynamic_cast<GemslrType*>(this->precond_)) g->SetWorkspaceKrylovDimension(kdim_);
      int err = this->precond_->Setup(x, rhs);
      if (err) return err;
    }
    this->ready_ = true;
    return 0;
  }
  int Solve(VectorType& x, VectorType& rhs) override {
    if (!this->ready_) return PGS_ERROR_FUNCTION_CALL;
    const int n = this->matrix_->GetNumRowsLocal();
    if (x.GetLengthLocal() != n || rhs.GetLengthLocal() != n) return PGS_ERROR_INVALID_PARAM;
    if (x.GetDataLocation() != rhs.GetDataLocation()) return PGS_ERROR_MEMORY_LOCATION;
    GemslrType* g = dynamic_cast<GemslrType*>(this->precond_);
    if (x.GetDataLocation() == kMemoryDevice && g && g->GetDeviceOperator() && g->GetMatrix() == this->matrix_)
      return SolveDevice(*g, x, rhs);
    return SolveGeneric(x, rhs);
  }
  int GetNumberIterations() const { return iters_; }
  double GetFinalRelativeResidual() const { return rel_res_; }
  // Relative residual estimates per iteration (GetRelativeResidual): every
  // iteration for the loop in this header; the device solve reports the
  // final estimate only.
  const std::vector<double>& GetRelativeResidual() const { return rel_res_vector_; }
  // 0, PGS_ERR_NOT_CONVERGED, PGS_ERR_BREAKDOWN or PGS_ERR_NONFINITE (the
  // reference returns success in every case and leaves the check to the caller)
  int GetStatus() const { return status_; }

 private:
  void ReleaseWorkspace() {
    if (ws_) pgs_ws_destroy(ws_);
    ws_ = nullptr;
    ws_kdim_ = 0;
  }

  // the whole FGMRES on the device, with the preconditioner's operator
  int SolveDevice(GemslrType& g, VectorType& x, VectorType& rhs) {
    pgs_dev_op_t dop = g.GetDeviceOperator();
    if (ws_ && (ws_kdim_ < kdim_ || ws_op_ != dop)) ReleaseWorkspace();
    if (!ws_) {
      int err = pgs_ws_create(dop, 1, kdim_, &ws_);
      if (err) return err;
      ws_kdim_ = kdim_;
      ws_op_ = dop;
    }
    pgs_fgmres_opts_t o;
    o.kdim = kdim_;
    o.maxits = maxits_;
    o.tol = tol_;
    o.absolute_tol = absolute_tol_ ? 1 : 0;
    o.prec = g.GetApplyPrecision();
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
    rel_res_vector_.assign(1, r.rel_res);
    return 0;
  }

  // FlexGmresClass::Solve (fgmres.hpp:263-558) over the vectors' own
  // operations: any SolverClass preconditioner (or none), host or device
  // vectors.  Modified Gram-Schmidt without re-orthogonalization
  // (PargemslrMgs with tol_reorth < 0, matrixops.cpp:10728-10807), Givens
  // rotations on the host; a restart keeps the Givens estimate as the
  // residual norm, as the reference does; a breakdown (|gamma| < eps) leaves
  // x as it was at the start of the cycle.
  int SolveGeneric(VectorType& x, VectorType& rhs) {
    using T = DataType;
    const int n = x.GetLengthLocal();
    const int loc = x.GetDataLocation();
    const int m = std::max(1, kdim_);
    const double eps = std::numeric_limits<double>::epsilon();
    const MatrixType& A = *this->matrix_;
    iters_ = 0;
    status_ = 0;
    rel_res_ = 0.0;
    rel_res_vector_.clear();

    // basis V and preconditioned basis Z: (m + 1) vectors each, in x's location
    VectorType Vs, Zs;
    int err = Vs.Setup(n * (m + 1), loc, true);
    if (!err) err = Zs.Setup(n * (m + 1), loc, true);
    if (err) return err;
    auto col = [n](VectorType& big, int j, VectorType& view) {
      view.SetupPtr(big.GetData() + (int64_t)n * j, n, big.GetDataLocation());
    };
    VectorType v, z, w;
    std::vector<T> H((size_t)(m + 1) * m, T(0.0)), c(m, T(0.0)), s(m, T(0.0)), rs(m + 1, T(0.0));
    auto h = [&H, m](int i, int j) -> T& { return H[(size_t)j * (m + 1) + i]; };

    double normb = 0.0, normr = 0.0;
    if ((err = rhs.Norm2(normb))) return err;
    if (!std::isfinite(normb)) {
      status_ = PGS_ERR_NONFINITE;
      return 0;
    }
    if (normb < eps) {
      rel_res_vector_.assign(1, 0.0);
      return x.Fill(T(0.0));
    }
    col(Vs, 0, v);
    if ((err = v.CopyFrom(rhs))) return err;
    if ((err = A.MatVec('N', T(-1.0), x, T(1.0), v))) return err;
    if ((err = v.Norm2(normr))) return err;
    if (normr < eps) {
      rel_res_vector_.assign(1, 0.0);
      return 0;
    }
    const double tolr = absolute_tol_ ? tol_ : tol_ * normb;
    rel_res_vector_.push_back(normr / normb);
    int i = 0;
    bool breakdown = false;
    while (iters_ < maxits_ && !breakdown) {
      rs[0] = T(normr);
      col(Vs, 0, v);
      if ((err = v.Scale(T(1.0 / normr)))) return err;
      i = 0;
      while (i < m && iters_ < maxits_) {
        ++i;
        ++iters_;
        col(Vs, i - 1, v);
        col(Zs, i - 1, z);
        col(Vs, i, w);
        if (this->precond_) {
          if ((err = this->precond_->Solve(z, v))) return err;
          if ((err = A.MatVec('N', T(1.0), z, T(0.0), w))) return err;
        } else {
          if ((err = z.CopyFrom(v))) return err;
          if ((err = A.MatVec('N', T(1.0), v, T(0.0), w))) return err;
        }
        // modified Gram-Schmidt of w against V(:, 0..i-1)
        VectorType vj;
        for (int j = 0; j < i; ++j) {
          col(Vs, j, vj);
          T t1(0.0);
          if ((err = vj.Dot(w, t1))) return err;
          h(j, i - 1) = t1;
          if ((err = w.Axpy(-t1, vj))) return err;
        }
        double t = 0.0;
        if ((err = w.Norm2(t))) return err;
        h(i, i - 1) = T(t);
        if ((err = w.Scale(T(1.0) / T(t)))) return err;
        // previous rotations on the new column, then its own rotation
        for (int j = 1; j < i; ++j) {
          const T hii = h(j - 1, i - 1);
          h(j - 1, i - 1) = conj_(c[j - 1]) * hii + s[j - 1] * h(j, i - 1);
          h(j, i - 1) = -s[j - 1] * hii + c[j - 1] * h(j, i - 1);
        }
        const T hii = h(i - 1, i - 1), hii1 = h(i, i - 1);
        const double gam = std::sqrt(std::norm(hii) + std::norm(hii1));
        if (!std::isfinite(gam)) {
          status_ = PGS_ERR_NONFINITE;
          breakdown = true;
          break;
        }
        if (std::abs(gam) < eps) {
          status_ = PGS_ERR_BREAKDOWN;
          breakdown = true;
          break;
        }
        c[i - 1] = hii / gam;
        s[i - 1] = hii1 / gam;
        rs[i] = -s[i - 1] * rs[i - 1];
        rs[i - 1] = conj_(c[i - 1]) * rs[i - 1];
        h(i - 1, i - 1) = conj_(c[i - 1]) * hii + s[i - 1] * hii1;
        normr = std::abs(rs[i]);
        rel_res_vector_.push_back(normr / normb);
        if (normr <= tolr) break;
      }
      if (breakdown) break;
      // x += Z y with H(0:i, 0:i) y = rs(0:i)
      rs[i - 1] /= h(i - 1, i - 1);
      for (int k = i - 2; k >= 0; --k) {
        for (int j = k + 1; j < i; ++j) rs[k] -= h(k, j) * rs[j];
        rs[k] /= h(k, k);
      }
      for (int j = 0; j < i; ++j) {
        col(Zs, j, z);
        if ((err = x.Axpy(rs[j], z))) return err;
      }
      if (normr <= tolr) break;
      // restart (fgmres.hpp:500-514): V(:,0) = rhs - A x; the next cycle
      // scales it by the Givens estimate, not by its own norm
      col(Vs, 0, v);
      col(Vs, i, w);
      if ((err = v.CopyFrom(rhs))) return err;
      if ((err = A.MatVec('N', T(1.0), x, T(0.0), w))) return err;
      if ((err = v.Axpy(T(-1.0), w))) return err;
    }
    rel_res_ = normr / normb;
    if (!status_ && normr > tolr) status_ = PGS_ERR_NOT_CONVERGED;
    return 0;
  }
  static DataType conj_(const DataType& v) {
    if constexpr (VT::is_complex) return std::conj(v);
    else return v;
  }

  int kdim_ = 50, maxits_ = 200;
  double tol_ = 1e-8;
  bool absolute_tol_ = false;
  int location_ = kMemoryHost;
  int orth_ = PGS_ORTH_MGS;
  pgs_ws_t ws_ = nullptr;
  pgs_dev_op_t ws_op_ = nullptr;
  int ws_kdim_ = 0;
  int iters_ = 0, status_ = 0;
  double rel_res_ = 0.0;
  std::vector<double> rel_res_vector_;
};

// ---------------------------------------------------------------------------
// the parallel classes at one process
// ---------------------------------------------------------------------------

// ParallelVectorClass at one process: the local part is the whole vector.
template <typename T>
class ParallelVectorClass : public SequentialVectorClass<T> {
 public:
  using SequentialVectorClass<T>::Setup;
  // Setup(n_local, location, setzero, parlog) (parallel_vector.hpp)
  int Setup(int n_local, int location, bool setzero, const parallel_log&) {
    return SequentialVectorClass<T>::Setup(n_local, location, setzero);
  }
  int Setup(int n_local, const parallel_log&) { return SequentialVectorClass<T>::Setup(n_local, kMemoryHost, true); }
};

// ParallelCsrMatrixClass at one process: the diagonal block is the matrix and
// there is no off-processor part.  The Laplacian generators take a processor
// grid that must be 1 x 1 x 1 here.
template <typename T>
class ParallelCsrMatrixClass : public CsrMatrixClass<T> {
 public:
  int Laplacian(int nx, int ny, int nz, int dx, int dy, int dz, T alphax, T alphay, T alphaz, T shift,
                const parallel_log&, bool rand_perturb = false) {
    if (dx * dy * dz != 1) return PGS_ERROR_INVALID_PARAM;
    return CsrMatrixClass<T>::Laplacian(nx, ny, nz, alphax, alphay, alphaz, shift, rand_perturb);
  }
  int LaplacianWithPartition(int nx, int ny, int nz, int dx, int dy, int dz, int d2x, int d2y, int d2z, T alphax,
                             T alphay, T alphaz, T shift, const parallel_log& parlog, bool rand_perturb = false) {
    if (d2x * d2y * d2z != 1) return PGS_ERROR_INVALID_PARAM;
    return Laplacian(nx, ny, nz, dx, dy, dz, alphax, alphay, alphaz, shift, parlog, rand_perturb);
  }
  // the communication pattern of MatVec: nothing to set up at one process
  int SetupMatvec() { return 0; }
  int GetMpiInfo(int& np, int& myid) const {
    np = 1;
    myid = 0;
    return 0;
  }
};

// ParallelGemslrClass (parallel_gemslr.hpp:813; Solve parallel_gemslr.cpp:5496)
// at one process: the sequential GeMSLR operator (RKway levels, ILUT blocks,
// Arnoldi low-rank corrections) and apply behind the parallel class's name,
// Setup / Solve / GetNumNonzeros and parameter array.  The reference's
// distributed setup (ParMETIS partition across ranks, distributed Schur
// levels) has no counterpart; at one process it is a GeMSLR over the whole
// matrix, which this class builds with the sequential setup.
template <class MatrixType, class VectorType, typename DataType>
class ParallelGemslrClass : public GemslrClass<MatrixType, VectorType, DataType> {};

// the reference's typedef names (double and complexd instantiations)
typedef SequentialVectorClass<double> vector_seq_double;
typedef SequentialVectorClass<complexd> vector_seq_complexd;
typedef ParallelVectorClass<double> vector_par_double;
typedef ParallelVectorClass<complexd> vector_par_complexd;
typedef CsrMatrixClass<double> matrix_csr_double;
typedef CsrMatrixClass<complexd> matrix_csr_complexd;
typedef ParallelCsrMatrixClass<double> matrix_csr_par_double;
typedef ParallelCsrMatrixClass<complexd> matrix_csr_par_complexd;
typedef GemslrClass<CsrMatrixClass<double>, SequentialVectorClass<double>, double> precond_gemslr_csr_seq_double;
typedef GemslrClass<CsrMatrixClass<complexd>, SequentialVectorClass<complexd>, complexd>
    precond_gemslr_csr_seq_complexd;
typedef ParallelGemslrClass<ParallelCsrMatrixClass<double>, ParallelVectorClass<double>, double>
    precond_gemslr_csr_par_double;
typedef ParallelGemslrClass<ParallelCsrMatrixClass<complexd>, ParallelVectorClass<complexd>, complexd>
    precond_gemslr_csr_par_complexd;
typedef IluClass<CsrMatrixClass<double>, SequentialVectorClass<double>, double> precond_ilu_csr_seq_double;
typedef IluClass<CsrMatrixClass<complexd>, SequentialVectorClass<complexd>, complexd> precond_ilu_csr_seq_complexd;
typedef FlexGmresClass<CsrMatrixClass<double>, SequentialVectorClass<double>, double> fgmres_csr_seq_double;
typedef FlexGmresClass<CsrMatrixClass<complexd>, SequentialVectorClass<complexd>, complexd> fgmres_csr_seq_complexd;
typedef FlexGmresClass<ParallelCsrMatrixClass<double>, ParallelVectorClass<double>, double> fgmres_csr_par_double;
typedef FlexGmresClass<ParallelCsrMatrixClass<complexd>, ParallelVectorClass<complexd>, complexd>
    fgmres_csr_par_complexd;

}  // namespace pargemslr_amd

#endif  // PARGEMSLR_AMD_HPP

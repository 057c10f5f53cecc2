// Sequential GeMSLR-FGMRES driver for a matrix read from a MatrixMarket file,
// following the flow of the reference's general-matrix drivers
// (ParGeMSLR/TESTS/sequential/driver_gen_gemslr_seq.cpp:160-331 for double,
// driver_gen_gemslrz_seq.cpp for complexd): CsrMatrixClass::ReadFromMMFile,
// right-hand side from a vector file or b = A * 1 ("."), zero initial guess,
// FlexGmresClass + GemslrClass, Setup, Solve.  An operator file may be
// written after the setup ("save") or read in place of it ("load").
//
// usage: driver_gen matfile rhsfile|. complex(0|1) nlev ncomp droptol rownnz rank kdim tol prec(1|2)
//                   [save|load opfile]
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pargemslr_amd.hpp"

using namespace pargemslr_amd;

template <typename T>
int run(const char* matfile, const char* rhsfile, double* params, int prec, const char* mode,
        const char* opfile) {
  CsrMatrixClass<T> A;
  int err = A.ReadFromMMFile(matfile, 1);
  if (err) {
    std::fprintf(stderr, "reading %s failed: %s\n", matfile, pgs_strerror(err));
    return 3;
  }
  const int n = A.GetNumRowsLocal();
  SequentialVectorClass<T> x, b;
  if (std::strcmp(rhsfile, ".") == 0) {
    // x* = 1, b = A x*
    SequentialVectorClass<T> ones;
    ones.Setup(n, kMemoryHost);
    ones.Fill(T(1.0));
    b.Setup(n, kMemoryHost);
    A.MatVec('N', T(1.0), ones, T(0.0), b);
  } else if ((err = b.ReadFromMMFile(rhsfile, 1)) || b.GetLengthLocal() != n) {
    std::fprintf(stderr, "reading %s failed: %s\n", rhsfile, err ? pgs_strerror(err) : "length");
    return 3;
  }
  std::vector<T> bh(b.GetData(), b.GetData() + n);
  if (b.MoveData(kMemoryDevice)) return 4;
  x.Setup(n, kMemoryDevice, true);

  FlexGmresClass<CsrMatrixClass<T>, SequentialVectorClass<T>, T> solver;
  GemslrClass<CsrMatrixClass<T>, SequentialVectorClass<T>, T> precond;
  solver.SetMatrix(A);
  solver.SetWithParameterArray(params);
  solver.SetSolveLocation(kMemoryDevice);
  precond.SetMatrix(A);
  precond.SetWithParameterArray(params);
  precond.SetSolveLocation(kMemoryDevice);
  precond.SetApplyPrecision(prec == 2 ? PGS_PREC_FP32 : PGS_PREC_FP64);
  if (mode && std::strcmp(mode, "load") == 0 && (err = precond.Load(opfile))) {
    std::fprintf(stderr, "loading %s failed: %s\n", opfile, pgs_strerror(err));
    return 5;
  }
  solver.SetPreconditioner(precond);
  if ((err = solver.Setup(x, b))) {
    std::fprintf(stderr, "setup failed: %s\n", pgs_strerror(err));
    return 5;
  }
  if (mode && std::strcmp(mode, "save") == 0 && (err = precond.Save(opfile))) {
    std::fprintf(stderr, "saving %s failed: %s\n", opfile, pgs_strerror(err));
    return 5;
  }
  if ((err = solver.Solve(x, b))) {
    std::fprintf(stderr, "solve failed: %s\n", pgs_strerror(err));
    return 6;
  }
  std::vector<T> xh(n);
  x.CopyToHost(xh.data());
  SequentialVectorClass<T> xv, r;
  xv.SetupPtr(xh.data(), n, kMemoryHost);
  r.Setup(n, kMemoryHost);
  A.MatVec('N', T(1.0), xv, T(0.0), r);
  double nr = 0.0, nb = 0.0;
  for (int i = 0; i < n; ++i) {
    nr += std::norm(bh[i] - r.GetData()[i]);
    nb += std::norm(bh[i]);
  }
  const double rel = std::sqrt(nr / nb);
  std::printf("Matrix %s: n %d nnz %lld\nNumber of iterations: %d\nFinal rel res (Givens): %.17e\n"
              "True rel res: %e\n",
              matfile, n, (long long)A.GetNumNonzeros(), solver.GetNumberIterations(),
              solver.GetFinalRelativeResidual(), rel);
  return (solver.GetStatus() == 0 && rel <= params[PGS_IO_SOLVER_TOL]) ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc < 12) {
    std::fprintf(stderr,
                 "usage: %s matfile rhsfile|. complex nlev ncomp droptol rownnz rank kdim tol prec [save|load opfile]\n",
                 argv[0]);
    return 2;
  }
  double params[PGS_IO_SIZE];
  pgs_params_default(params);
  params[PGS_IO_PREPOSS_NLEV_GLOBAL] = std::atoi(argv[4]);
  params[PGS_IO_PREPOSS_NCOMP_GLOBAL] = std::atoi(argv[5]);
  params[PGS_IO_ILU_DROPTOL_B_GLOBAL] = params[PGS_IO_ILU_DROPTOL_C_GLOBAL] = std::atof(argv[6]);
  params[PGS_IO_ILU_ROWNNZ_B_GLOBAL] = params[PGS_IO_ILU_ROWNNZ_C_GLOBAL] = std::atoi(argv[7]);
  params[PGS_IO_LR_RANK1_GLOBAL] = params[PGS_IO_LR_RANK2_GLOBAL] = std::atoi(argv[8]);
  params[PGS_IO_LR_TOL_EIG1_GLOBAL] = params[PGS_IO_LR_TOL_EIG2_GLOBAL] = 1e-3;
  params[PGS_IO_LR_MAXITS1_GLOBAL] = params[PGS_IO_LR_MAXITS2_GLOBAL] = 10;
  params[PGS_IO_SOLVER_KDIM] = std::atoi(argv[9]);
  params[PGS_IO_SOLVER_MAXITS] = 1000;
  params[PGS_IO_SOLVER_TOL] = std::atof(argv[10]);
  const int prec = std::atoi(argv[11]);
  const bool cplx = std::atoi(argv[3]) != 0;
  const char* mode = argc > 13 ? argv[12] : nullptr;
  const char* opfile = argc > 13 ? argv[13] : nullptr;
  if (cplx) return run<complexd>(argv[1], argv[2], params, prec, mode, opfile);
  return run<double>(argv[1], argv[2], params, prec, mode, opfile);
}

// Sequential GeMSLR-FGMRES driver on the C++ class surface, following the
// flow of the reference's drivers (ParGeMSLR/TESTS/sequential/
// driver_laplacian_gemslr_seq.cpp:160-331 for double and
// driver_laplacian_gemslrz_seq.cpp for complexd): Laplacian, x* = 1,
// b = A x*, zero initial guess, FlexGmresClass + GemslrClass, Setup, Solve.
//
// usage: driver_laplacian nx ny nz nlev ncomp droptol rownnz rank kdim tol prec(1|2)
//                         [shift_re shift_im complex(0|1)]
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pargemslr_amd.hpp"

using namespace pargemslr_amd;

template <typename T>
int run(int nx, int ny, int nz, double* params, int prec, T shift) {
  CsrMatrixClass<T> A;
  if (A.Laplacian(nx, ny, nz, T(0.0), T(0.0), T(0.0), shift)) return 3;
  const int n = A.GetNumRowsLocal();
  // b = A * 1 on the host, then moved to the device
  SequentialVectorClass<T> x, b, ones;
  ones.Setup(n, kMemoryHost);
  ones.Fill(T(1.0));
  b.Setup(n, kMemoryHost);
  A.MatVec('N', T(1.0), ones, T(0.0), b);
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
  // PGS_LOWRANK_DEVICE=d: the low-rank corrections are built on device d
  if (const char* d = std::getenv("PGS_LOWRANK_DEVICE")) precond.SetLowRankDevice(std::atoi(d));
  solver.SetPreconditioner(precond);
  int err = solver.Setup(x, b);
  if (err) {
    std::fprintf(stderr, "setup failed: %s\n", pgs_strerror(err));
    return 5;
  }
  err = solver.Solve(x, b);
  if (err) {
    std::fprintf(stderr, "solve failed: %s\n", pgs_strerror(err));
    return 6;
  }
  // true residual on the host
  std::vector<T> xh(n);
  x.CopyToHost(xh.data());
  SequentialVectorClass<T> xv, r;
  xv.SetupPtr(xh.data(), n, kMemoryHost);
  r.Setup(n, kMemoryHost);
  A.MatVec('N', T(1.0), xv, T(0.0), r);
  double nr = 0.0, nb = 0.0, ne = 0.0;
  for (int i = 0; i < n; ++i) {
    nr += std::norm(bh[i] - r.GetData()[i]);
    nb += std::norm(bh[i]);
    ne += std::norm(xh[i] - T(1.0));
  }
  const double rel = std::sqrt(nr / nb);
  int64_t nnz_ilu = 0, nnz_lr = 0;
  precond.GetNumNonzeros(nnz_ilu, nnz_lr);
  std::printf("Number of iterations: %d\nFinal rel res (Givens): %e\nTrue rel res: %e\nError vs x*: %e\n"
              "Preconditioner fill level: ILU %f; Low-rank %f\n",
              solver.GetNumberIterations(), solver.GetFinalRelativeResidual(), rel, std::sqrt(ne / n),
              (double)nnz_ilu / A.GetNumNonzeros(), (double)nnz_lr / A.GetNumNonzeros());
  return (solver.GetStatus() == 0 && rel <= params[PGS_IO_SOLVER_TOL]) ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc < 12) {
    std::fprintf(stderr,
                 "usage: %s nx ny nz nlev ncomp droptol rownnz rank kdim tol prec [shift_re shift_im complex]\n",
                 argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]);
  double params[PGS_IO_SIZE];
  pgs_params_default(params);
  params[PGS_IO_PREPOSS_NLEV_GLOBAL] = std::atoi(argv[4]);
  params[PGS_IO_PREPOSS_NCOMP_GLOBAL] = std::atoi(argv[5]);
  params[PGS_IO_ILU_DROPTOL_B_GLOBAL] = params[PGS_IO_ILU_DROPTOL_C_GLOBAL] = std::atof(argv[6]);
  params[PGS_IO_ILU_ROWNNZ_B_GLOBAL] = params[PGS_IO_ILU_ROWNNZ_C_GLOBAL] = std::atoi(argv[7]);
  params[PGS_IO_LR_RANK1_GLOBAL] = params[PGS_IO_LR_RANK2_GLOBAL] = std::atoi(argv[8]);
  params[PGS_IO_LR_TOL_EIG1_GLOBAL] = params[PGS_IO_LR_TOL_EIG2_GLOBAL] = 1e-3;
  params[PGS_IO_LR_MAXITS1_GLOBAL] = params[PGS_IO_LR_MAXITS2_GLOBAL] = 10;
  params[PGS_IO_SCHUR_ENABLE] = 0;
  params[PGS_IO_SOLVER_KDIM] = std::atoi(argv[9]);
  params[PGS_IO_SOLVER_MAXITS] = 1000;
  params[PGS_IO_SOLVER_TOL] = std::atof(argv[10]);
  const int prec = std::atoi(argv[11]);
  const double sre = argc > 12 ? std::atof(argv[12]) : 0.0;
  const double sim = argc > 13 ? std::atof(argv[13]) : 0.0;
  const bool cplx = argc > 14 && std::atoi(argv[14]) != 0;
  if (cplx) return run<complexd>(nx, ny, nz, params, prec, complexd(sre, sim));
  return run<double>(nx, ny, nz, params, prec, sre);
}

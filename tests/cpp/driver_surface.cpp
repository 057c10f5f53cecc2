// The reference's one-process driver flow on the class surface, calling every
// method its sequential Laplacian driver calls (ParGeMSLR/TESTS/sequential/
// driver_laplacian_gemslr_seq.cpp:160-351: engine seed, the 8-argument
// Laplacian, MoveData, Setup / Rand / Fill / MatVec for the solution and
// initial-guess options, SetMatrix / SetWithParameterArray / SetSolveLocation
// / SetPreconditioner, Setup and Solve under the timing macros,
// GetNumNonzeros, the iteration and residual getters, WriteToDisk and Clear),
// and the same with the parallel classes at one process
// (TESTS/parallel/driver_laplacian_gemslr_par.cpp:247-454:
// LaplacianWithPartition, SetupMatvec, Setup with a parallel_log,
// SetMatrixP / SetPreconditionerP), an IluClass preconditioner, a
// user-written SolverClass preconditioner (Jacobi) and no preconditioner.
//
// usage: driver_surface mode location nx ny nz nlev ncomp droptol rownnz rank kdim tol sol_opt init_opt
//                       rand_perturb [shift [solfile]]
//   mode: seq | par | ilu | jacobi | none | rand;  location: host | device
//   (rand: seeds the engine with 0, fills a vector of nx*ny*nz values with
//   Rand on `location` and writes it to solfile; with rand_perturb also
//   writes the diagonal of the perturbed Laplacian to solfile.diag)
// Prints the iteration count, the Givens estimate, the true relative residual
// and the error against x* (sol_opt 0 or 1); exit 0 when the solve reached tol.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pargemslr_amd.hpp"

using namespace pargemslr_amd;

namespace {

// z = D^{-1} v: a preconditioner written against SolverClass only
template <class M, class V, typename T>
class JacobiClass : public SolverClass<M, V, T> {
 public:
  int Setup(V&, V&) override {
    const M& A = *this->GetMatrix();
    const int n = A.GetNumRowsLocal();
    dinv_.assign(n, T(1.0));
    for (int i = 0; i < n; ++i)
      for (int k = A.GetI()[i]; k < A.GetI()[i + 1]; ++k)
        if (A.GetJ()[k] == i) dinv_[i] = T(1.0) / A.GetData()[k];
    this->ready_ = true;
    return 0;
  }
  int Solve(V& x, V& rhs) override {
    const int n = rhs.GetLengthLocal();
    std::vector<T> h(n);
    int err = rhs.CopyToHost(h.data());
    if (err) return err;
    for (int i = 0; i < n; ++i) h[i] *= dinv_[i];
    return x.CopyFromHost(h.data());
  }
  int SetSolveLocation(const int&) override { return 0; }

 private:
  std::vector<T> dinv_;
};

struct Args {
  std::string mode, loc;
  int nx, ny, nz, nlev, ncomp, rownnz, rank, kdim, sol_opt, init_opt;
  double droptol, tol;
  bool rand_perturb;
  double shift;
  const char* solfile;
};

void set_params(double* params, const Args& a) {
  params[PGS_IO_PREPOSS_NLEV_GLOBAL] = a.nlev;
  params[PGS_IO_PREPOSS_NCOMP_GLOBAL] = a.ncomp;
  params[PGS_IO_ILU_DROPTOL_B_GLOBAL] = params[PGS_IO_ILU_DROPTOL_C_GLOBAL] = a.droptol;
  params[PGS_IO_ILU_ROWNNZ_B_GLOBAL] = params[PGS_IO_ILU_ROWNNZ_C_GLOBAL] = a.rownnz;
  params[PGS_IO_LR_RANK1_GLOBAL] = params[PGS_IO_LR_RANK2_GLOBAL] = a.rank;
  params[PGS_IO_LR_TOL_EIG1_GLOBAL] = params[PGS_IO_LR_TOL_EIG2_GLOBAL] = 1e-3;
  params[PGS_IO_LR_MAXITS1_GLOBAL] = params[PGS_IO_LR_MAXITS2_GLOBAL] = 10;
  params[PGS_IO_SCHUR_ENABLE] = 0;
  params[PGS_IO_SOLVER_KDIM] = a.kdim;
  params[PGS_IO_SOLVER_MAXITS] = 1000;
  params[PGS_IO_SOLVER_TOL] = a.tol;
}

// right-hand side and initial guess as the reference driver's sol_opt /
// init_opt switches (driver_laplacian_gemslr_seq.cpp:196-264)
template <class M, class V>
void make_problem(const Args& a, M& A, V& x, V& b) {
  const double one = 1.0, zero = 0.0;
  switch (a.sol_opt) {
    case 1:
      x.Rand();
      b.Fill(zero);
      A.MatVec('N', one, x, zero, b);
      break;
    case 2:
      b.Fill(one);
      break;
    case 3:
      b.Rand();
      break;
    default:
      x.Fill(one);
      b.Fill(zero);
      A.MatVec('N', one, x, zero, b);
      break;
  }
  switch (a.init_opt) {
    case 1:
      x.Fill(one);
      break;
    case 2:
      x.Rand();
      break;
    default:
      x.Fill(zero);
      break;
  }
}

// true relative residual and error against x* on the host
template <class M, class V>
void report(const M& A, V& x, const std::vector<double>& bh, const std::vector<double>& xs, int iters,
            double givens, int status, double tol, double nnz_ilu, double nnz_lr) {
  const int n = A.GetNumRowsLocal();
  std::vector<double> xh(n);
  x.CopyToHost(xh.data());
  V xv, r;
  xv.SetupPtr(xh.data(), n, kMemoryHost);
  r.Setup(n, kMemoryHost, true);
  A.MatVec('N', 1.0, xv, 0.0, r);
  double nr = 0.0, nb = 0.0, ne = 0.0, nx = 0.0;
  for (int i = 0; i < n; ++i) {
    nr += (bh[i] - r.GetData()[i]) * (bh[i] - r.GetData()[i]);
    nb += bh[i] * bh[i];
    if (!xs.empty()) {
      ne += (xh[i] - xs[i]) * (xh[i] - xs[i]);
      nx += xs[i] * xs[i];
    }
  }
  const double rel = std::sqrt(nr / nb);
  PARGEMSLR_PRINT("\tNumber of iterations: %d\n", iters);
  PARGEMSLR_PRINT("\tFinal rel res: %e\n", givens);
  PARGEMSLR_PRINT("\tTrue rel res: %e\n", rel);
  if (!xs.empty()) PARGEMSLR_PRINT("\tError vs x*: %e\n", std::sqrt(ne / nx));
  PARGEMSLR_PRINT("\tStatus: %d\n", status);
  PARGEMSLR_PRINT("\tPreconditioner fill level: ILU: %f; Low-rank: %f; Total: %f\n", nnz_ilu, nnz_lr,
                  nnz_ilu + nnz_lr);
  std::fflush(stdout);
  (void)tol;
}

template <template <class, class, typename> class Precond>
int run_seq(const Args& a, double* params, int location) {
  typedef CsrMatrixClass<double> M;
  typedef SequentialVectorClass<double> V;
  M csr_mat;
  V x, b;
  FlexGmresClass<M, V, double> solver;
  Precond<M, V, double> precond;

  pargemslr_global::_mersenne_twister_engine.seed(0);
  PARGEMSLR_LOCAL_FIRM_TIME_CALL(PARGEMSLR_TOTAL_GEN_MAT_TIME,
                                 (csr_mat.Laplacian(a.nx, a.ny, a.nz, 0.0, 0.0, 0.0, a.shift, a.rand_perturb)));
  const int n = csr_mat.GetNumRowsLocal();
  if (csr_mat.MoveData(location)) return 3;
  if (x.Setup(n, location, false) || b.Setup(n, location, false)) return 3;
  make_problem(a, csr_mat, x, b);
  std::vector<double> bh(n), xs;
  b.CopyToHost(bh.data());
  if (a.sol_opt == 0) xs.assign(n, 1.0);
  if (a.sol_opt == 1) {  // x holds x* until the initial guess overwrote it: regenerate the same draws
    pargemslr_global::_mersenne_twister_engine.seed(0);
    if (a.rand_perturb) for (int i = 0; i < n; ++i) { double t; PargemslrValueRandHost(t); }
    xs.resize(n);
    for (int i = 0; i < n; ++i) PargemslrValueRandHost(xs[i]);
  }

  solver.SetMatrix(csr_mat);
  solver.SetWithParameterArray(params);
  solver.SetSolveLocation(location);
  precond.SetMatrix(csr_mat);
  precond.SetWithParameterArray(params);
  precond.SetSolveLocation(location);
  solver.SetPreconditioner(precond);
  int err = 0;
  PARGEMSLR_LOCAL_FIRM_TIME_CALL(PARGEMSLR_TOTAL_SETUP_TIME, (err = solver.Setup(x, b)));
  if (err) {
    std::fprintf(stderr, "setup failed: %s\n", pgs_strerror(err));
    return 5;
  }
  PARGEMSLR_LOCAL_FIRM_TIME_CALL(PARGEMSLR_TOTAL_SOLVE_TIME, (err = solver.Solve(x, b)));
  if (err) {
    std::fprintf(stderr, "solve failed: %s\n", pgs_strerror(err));
    return 6;
  }
  long int nnzILU = 0, nnzLR = 0;
  const long int nnzA = csr_mat.GetNumNonzeros();
  const long int nnzM = precond.GetNumNonzeros(nnzILU, nnzLR);
  if (nnzM != nnzILU + nnzLR || precond.GetNumNonzeros() != nnzM) return 7;
  PargemslrPrintDashLine(pargemslr_global::_dash_line_width);
  PARGEMSLR_PRINT("Solution info:\n");
  report(csr_mat, x, bh, xs, solver.GetNumberIterations(), solver.GetFinalRelativeResidual(), solver.GetStatus(),
         a.tol, (double)nnzILU / nnzA, (double)nnzLR / nnzA);
  PARGEMSLR_PRINT_TIMING_RESULT(params[PGS_IO_SOLVER_KDIM] > 0 ? 0 : 1, 1);
  PARGEMSLR_RESET_TIME;
  const int status = solver.GetStatus();
  if (a.solfile) {
    if (x.MoveData(kMemoryHost) || x.WriteToDisk(a.solfile)) return 8;
  }
  x.Clear();
  b.Clear();
  csr_mat.Clear();
  solver.Clear();
  precond.Clear();
  return status == 0 ? 0 : 1;
}

// a SolverClass preconditioner of the driver's own, or none
int run_seq_plain(const Args& a, double* params, int location, bool jacobi) {
  typedef CsrMatrixClass<double> M;
  typedef SequentialVectorClass<double> V;
  M csr_mat;
  V x, b;
  FlexGmresClass<M, V, double> solver;
  JacobiClass<M, V, double> precond;
  pargemslr_global::_mersenne_twister_engine.seed(0);
  csr_mat.Laplacian(a.nx, a.ny, a.nz, 0.0, 0.0, 0.0, a.shift, a.rand_perturb);
  const int n = csr_mat.GetNumRowsLocal();
  if (csr_mat.MoveData(location)) return 3;
  if (x.Setup(n, location, false) || b.Setup(n, location, false)) return 3;
  make_problem(a, csr_mat, x, b);
  std::vector<double> bh(n), xs;
  b.CopyToHost(bh.data());
  if (a.sol_opt == 0) xs.assign(n, 1.0);
  solver.SetMatrix(csr_mat);
  solver.SetWithParameterArray(params);
  solver.SetSolveLocation(location);
  if (jacobi) {
    precond.SetMatrix(csr_mat);
    solver.SetPreconditioner(precond);
  }
  int err = solver.Setup(x, b);
  if (!err) err = solver.Solve(x, b);
  if (err) {
    std::fprintf(stderr, "solve failed: %s\n", pgs_strerror(err));
    return 6;
  }
  PARGEMSLR_PRINT("Solution info:\n");
  report(csr_mat, x, bh, xs, solver.GetNumberIterations(), solver.GetFinalRelativeResidual(), solver.GetStatus(),
         a.tol, 0.0, 0.0);
  const std::vector<double>& hist = solver.GetRelativeResidual();
  PARGEMSLR_PRINT("\tHistory length: %zu\n", hist.size());
  return solver.GetStatus() == 0 ? 0 : 1;
}

// the parallel classes at one process (driver_laplacian_gemslr_par.cpp:247-454)
int run_par(const Args& a, double* params, int location) {
  typedef ParallelCsrMatrixClass<double> M;
  typedef ParallelVectorClass<double> V;
  parallel_log parlog;
  int np = 0, myid = -1, comm = 0;
  parlog.GetMpiInfo(np, myid, comm);
  if (np != 1 || myid != 0) return 9;
  M* parcsr_mat = new M();
  V* x = new V();
  V* b = new V();
  FlexGmresClass<M, V, double>* solverp = new FlexGmresClass<M, V, double>();
  ParallelGemslrClass<M, V, double>* precondp = new ParallelGemslrClass<M, V, double>();
  pargemslr_global::_mersenne_twister_engine.seed(0);
  int err = parcsr_mat->LaplacianWithPartition(a.nx, a.ny, a.nz, 1, 1, 1, 1, 1, 1, 0.0, 0.0, 0.0, a.shift, parlog,
                                               a.rand_perturb);
  if (err) return 3;
  const int n_local = parcsr_mat->GetNumRowsLocal();
  if ((int64_t)n_local != parcsr_mat->GetNumRowsGlobal()) return 3;
  parcsr_mat->SetupMatvec();
  parcsr_mat->MoveData(location);
  x->Setup(n_local, location, false, parlog);
  b->Setup(n_local, location, false, parlog);
  make_problem(a, *parcsr_mat, *x, *b);
  std::vector<double> bh(n_local), xs;
  b->CopyToHost(bh.data());
  if (a.sol_opt == 0) xs.assign(n_local, 1.0);
  solverp->SetMatrixP(parcsr_mat);
  solverp->SetWithParameterArray(params);
  solverp->SetSolveLocation(location);
  precondp->SetMatrixP(parcsr_mat);
  precondp->SetWithParameterArray(params);
  precondp->SetSolveLocation(location);
  solverp->SetPreconditionerP(precondp);
  PARGEMSLR_GLOBAL_FIRM_TIME_CALL(PARGEMSLR_TOTAL_SETUP_TIME, (err = solverp->Setup(*x, *b)));
  if (!err) PARGEMSLR_GLOBAL_FIRM_TIME_CALL(PARGEMSLR_TOTAL_SOLVE_TIME, (err = solverp->Solve(*x, *b)));
  if (err) {
    std::fprintf(stderr, "solve failed: %s\n", pgs_strerror(err));
    return 6;
  }
  long int nnzILU = 0, nnzLR = 0;
  const long int nnzA = parcsr_mat->GetNumNonzeros();
  precondp->GetNumNonzeros(nnzILU, nnzLR);
  PARGEMSLR_PRINT("Solution info:\n");
  report(*parcsr_mat, *x, bh, xs, solverp->GetNumberIterations(), solverp->GetFinalRelativeResidual(),
         solverp->GetStatus(), a.tol, (double)nnzILU / nnzA, (double)nnzLR / nnzA);
  const int status = solverp->GetStatus();
  x->Clear();
  b->Clear();
  parcsr_mat->Clear();
  solverp->Clear();
  precondp->Clear();
  delete x;
  delete b;
  delete solverp;
  delete precondp;
  delete parcsr_mat;
  return status == 0 ? 0 : 1;
}

// Rand on either location, written with WriteToDisk; the perturbed
// Laplacian's diagonal beside it
int run_rand(const Args& a, int location) {
  const int n = a.nx * a.ny * a.nz;
  SequentialVectorClass<double> v;
  pargemslr_global::_mersenne_twister_engine.seed(0);
  if (v.Setup(n, location, false) || v.Rand() || v.MoveData(kMemoryHost)) return 3;
  if (!a.solfile || v.WriteToDisk(a.solfile)) return 8;
  if (a.rand_perturb) {
    CsrMatrixClass<double> A;
    pargemslr_global::_mersenne_twister_engine.seed(0);
    if (A.Laplacian(a.nx, a.ny, a.nz, 0.0, 0.0, 0.0, a.shift, true)) return 3;
    SequentialVectorClass<double> d;
    d.Setup(n, kMemoryHost, true);
    for (int i = 0; i < n; ++i)
      for (int k = A.GetI()[i]; k < A.GetI()[i + 1]; ++k)
        if (A.GetJ()[k] == i) d.GetData()[i] = A.GetData()[k];
    if (d.WriteToDisk((std::string(a.solfile) + ".diag").c_str())) return 8;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 16) {
    std::fprintf(stderr,
                 "usage: %s mode location nx ny nz nlev ncomp droptol rownnz rank kdim tol sol_opt init_opt "
                 "rand_perturb [shift [solfile]]\n",
                 argv[0]);
    return 2;
  }
  PargemslrInit(&argc, &argv);
  Args a;
  a.mode = argv[1];
  a.loc = argv[2];
  a.nx = std::atoi(argv[3]);
  a.ny = std::atoi(argv[4]);
  a.nz = std::atoi(argv[5]);
  a.nlev = std::atoi(argv[6]);
  a.ncomp = std::atoi(argv[7]);
  a.droptol = std::atof(argv[8]);
  a.rownnz = std::atoi(argv[9]);
  a.rank = std::atoi(argv[10]);
  a.kdim = std::atoi(argv[11]);
  a.tol = std::atof(argv[12]);
  a.sol_opt = std::atoi(argv[13]);
  a.init_opt = std::atoi(argv[14]);
  a.rand_perturb = std::atoi(argv[15]) != 0;
  a.shift = argc > 16 ? std::atof(argv[16]) : 0.0;
  a.solfile = argc > 17 ? argv[17] : nullptr;
  double* params = pargemslr_global::_params;
  set_params(params, a);
  const int location = a.loc == "device" ? kMemoryDevice : kMemoryHost;
  PargemslrPrintParallelInfo();
  int rc = 2;
  if (a.mode == "seq") rc = run_seq<GemslrClass>(a, params, location);
  else if (a.mode == "ilu") rc = run_seq<IluClass>(a, params, location);
  else if (a.mode == "par") rc = run_par(a, params, location);
  else if (a.mode == "jacobi") rc = run_seq_plain(a, params, location, true);
  else if (a.mode == "none") rc = run_seq_plain(a, params, location, false);
  else if (a.mode == "rand") rc = run_rand(a, location);
  PargemslrFinalize();
  return rc;
}

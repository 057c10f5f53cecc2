"""Test problems and helpers shared by the CPU and GPU tests."""
import os

import numpy as np

import pargemslr_amd as pg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_mtx(name):
    """MatrixMarket fixture from the reference's test data (tests/golden;
    a .gz name is read through gzip)."""
    import gzip
    import scipy.io as sio
    path = os.path.join(GOLDEN, name)
    if name.endswith(".gz"):
        with gzip.open(path) as f:
            return sio.mmread(f)
    return sio.mmread(path)


def csr_from_scipy(M):
    M = M.tocsr()
    M.sort_indices()
    dt = np.complex128 if np.iscomplexobj(M.data) else np.float64
    return pg.CsrHost(M.shape[0], M.indptr, M.indices, M.data.astype(dt))


def dense(A):
    return A.to_scipy().toarray()


def setup(A, nlev, ncomp, droptol, rownnz, rank, seed=0, lr_maxits=10, lr_tol_eig=1e-3, perm=1, schur=None,
          usolve=False):
    p = pg.config_params(nlev, ncomp, droptol, rownnz, rank, lr_maxits=lr_maxits, lr_tol_eig=lr_tol_eig, perm=perm,
                         schur=schur, usolve=usolve)
    return pg.Gemslr(A, p, seed=seed)


def oracle_of(G):
    """The oracle on G's factors, with G's inner Schur iteration settings."""
    from oracle import oracle as orc
    return orc.GemslrOracle(G.A.n, G.perm(), G.levels(), schur=G.schur(), usolve=G.usolve())


# configurations (sizes scaled so the oracle finishes in seconds)
C1 = dict(grid=(64, 64, 1), nlev=2, ncomp=4, droptol=1e-2, rownnz=10, rank=0, kdim=30)
C2_SMALL = dict(grid=(32, 32, 32), nlev=3, ncomp=8, droptol=1e-3, rownnz=20, rank=16, kdim=50)
C2_MID = dict(grid=(48, 48, 48), nlev=3, ncomp=64, droptol=1e-3, rownnz=20, rank=16, kdim=50)


def build(cfg, shift=0.0, alphas=(0.0, 0.0, 0.0)):
    A = pg.laplacian(*cfg["grid"], shift=shift, alphas=alphas)
    G = setup(A, cfg["nlev"], cfg["ncomp"], cfg["droptol"], cfg["rownnz"], cfg["rank"])
    return A, G

"""Low-rank corrections built on the device (pgs_gemslr_setup_device) against
the host build (pgs_gemslr_setup) of the same operator, parameters and seed.
Partition, orderings, ILUT and couplings come from the same host code and are
identical; the Arnoldi process runs the same driver over device vectors, so
its bases agree with the host's to rounding: same ranks, orthonormal bases
spanning the host's subspaces, preconditioners whose applies agree to far
below the solve tolerance, and solves that converge in the same iteration
counts.  The device-built operator is also checked against the oracle on its
own factors (the device solve path is the same as for host-built operators)."""
import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P

pytestmark = pytest.mark.gpu


def _dev(x, dev):
    dt = np.complex128 if np.iscomplexobj(x) else np.float64
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=dt), device=dev)


def _host(t):
    pg.synchronize(t.device)
    return t.numpy()


def _rhs(A, seed, nr):
    if A.is_complex:
        return np.stack([pg.rng_normal(seed + c, 2 * A.n).view(np.complex128) for c in range(nr)], axis=1)
    return np.stack([pg.rng_normal(seed + c, A.n) for c in range(nr)], axis=1)


CASES = {
    "laplacian": dict(grid=20, shift=0.0, alphas=(0.0, 0.0, 0.0), cplx=False, schur=None),
    "convdiff_schur": dict(grid=20, shift=0.3, alphas=(0.05, -0.02, 0.01), cplx=False, schur=(5, 1e-2)),
    "complex": dict(grid=14, shift=0.2 + 0.02j, alphas=(0.05, -0.02, 0.01), cplx=True, schur=None),
}


def _pair(name, gpu):
    c = CASES[name]
    g = c["grid"]
    A = pg.laplacian(g, g, g, shift=c["shift"], alphas=c["alphas"], complex_=c["cplx"])
    p = pg.config_params(3, 8, 1e-3, 20, 16, lr_maxits=10, lr_tol_eig=1e-3, schur=c["schur"])
    print(f"case {name}: n={A.n}", flush=True)
    return A, pg.Gemslr(A, p, seed=5), pg.Gemslr(A, p, seed=5, lowrank_device=gpu)


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_bases_match_the_host_build(name, gpu):
    A, Gh, Gd = _pair(name, gpu)
    assert Gd.info().nlev == Gh.info().nlev
    assert np.array_equal(Gd.perm(), Gh.perm())
    for l, (lh, ld) in enumerate(zip(Gh.levels(), Gd.levels())):
        for k in lh:
            if k in ("W", "WH", "rank"):
                continue
            a, b = lh[k], ld[k]
            if isinstance(a, np.ndarray):
                assert np.array_equal(a, b), (l, k)
            elif isinstance(a, tuple):
                assert all(np.array_equal(x, y) for x, y in zip(a, b)), (l, k)
            else:
                assert a == b, (l, k)
        assert lh["rank"] == ld["rank"], l
        if lh["rank"] == 0:
            continue
        Wh, Wd = lh["W"], ld["W"]
        r = Wd.shape[1]
        # orthonormal columns: W is V Y with V and Y orthonormal to rounding
        ortho = np.linalg.norm(Wd.conj().T @ Wd - np.eye(r))
        # the same invariant subspace: the part of Wd outside span(Wh)
        outside = np.linalg.norm(Wd - Wh @ (Wh.conj().T @ Wd))
        print(f"level {l}: rank {r}, ||Wd^H Wd - I|| = {ortho:.2e}, ||(I - Wh Wh^H) Wd|| = {outside:.2e}", flush=True)
        assert ortho <= 1e-10
        assert outside <= 1e-6


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_built_operator_solves_like_the_host_one(name, gpu):
    A, Gh, Gd = _pair(name, gpu)
    nr = 4
    B = _rhs(A, 80, nr)
    outs = []
    for G in (Gh, Gd):
        D = pg.DevOp(G)
        W = pg.Workspace(D, max_nrhs=nr, kdim=50)
        Z = _dev(np.zeros_like(B), gpu)
        W.apply(_dev(B, gpu), Z, nrhs=nr, prec=pg.PGS_PREC_FP64)
        X = _dev(np.zeros_like(B), gpu)
        res = W.fgmres(_dev(B, gpu), X, nrhs=nr, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP64)
        outs.append((_host(Z), _host(X), [r[:2] for r in res]))
    (Zh, Xh, rh), (Zd, Xd, rd) = outs
    dz = np.linalg.norm(Zd - Zh) / np.linalg.norm(Zh)
    print(f"apply difference {dz:.2e}; iterations host {[r[1] for r in rh]} device {[r[1] for r in rd]}", flush=True)
    assert dz <= 1e-6
    for (sh, ih), (sd, idv) in zip(rh, rd):
        assert sh == 0 and sd == 0
        assert abs(ih - idv) <= 1
    for c in range(nr):
        ax = A.matvec(Xd[:, c])
        assert np.linalg.norm(B[:, c] - ax) <= 1e-8 * np.linalg.norm(B[:, c]) * 1.0001


def test_device_built_operator_matches_the_oracle(gpu):
    A, Gh, Gd = _pair("laplacian", gpu)
    v = pg.rng_normal(9, A.n)
    ref = P.oracle_of(Gd).apply(v)
    D = pg.DevOp(Gd)
    W = pg.Workspace(D, max_nrhs=1, kdim=4)
    Z = _dev(np.zeros(A.n), gpu)
    W.apply(_dev(v, gpu), Z, nrhs=1, prec=pg.PGS_PREC_FP64)
    z = _host(Z).reshape(-1)
    assert np.linalg.norm(z - ref) <= 1e-12 * np.linalg.norm(ref)


def test_setup_time_is_reported(gpu):
    A, Gh, Gd = _pair("laplacian", gpu)
    th, td = Gh.info().setup_seconds, Gd.info().setup_seconds
    print(f"low-rank build: host {th[2]:.3f} s, device {td[2]:.3f} s", flush=True)
    assert td[2] > 0 and abs(td[3] - (td[0] + td[1] + td[2])) <= 1e-6 * max(1.0, td[3])

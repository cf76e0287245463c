"""Host side of the implicit method (cfd2d/implicit.py) on the CPU: the
restarted GMRES, the block-row to BSR conversion, and the run loop driven by
the oracle's matrix fill (the GPU fill is compared in test_gpu_implicit.py)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from cfd2d import case as cases
from cfd2d import implicit


def oracle_assembler(case):
    o = oracle.OracleSolver(case)

    def assemble(U, steady, cfl, tau):
        o.set_state(U)
        return o.implicit_assemble(steady=steady, cfl=cfl, tau=tau)
    return assemble


def random_system(n, seed, dominance=4.0):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.05, random_state=seed, format="csr")
    A = A + sp.diags(dominance + rng.uniform(0, 1, n))
    return A.tocsr(), rng.normal(size=n)


def test_gmres_solves_to_tolerance():
    A, b = random_system(300, 1)
    x, it, rel = implicit.gmres(A, b, tol=1e-10, max_iter=200, kdim=30)
    assert rel < 1e-10 and it < 200
    np.testing.assert_allclose(x, spla.spsolve(A.tocsc(), b), rtol=1e-8, atol=1e-10)
    # restarts: a Krylov space of 5 still converges, in more iterations
    x5, it5, rel5 = implicit.gmres(A, b, tol=1e-10, max_iter=400, kdim=5)
    assert rel5 < 1e-10 and it5 >= it


def test_gmres_iteration_cap():
    A, b = random_system(300, 2, dominance=0.5)
    x, it, rel = implicit.gmres(A, b, tol=1e-12, max_iter=7)
    assert it == 7 and rel > 1e-12
    # the residual it reports is the true one
    assert abs(rel - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) < 1e-14


def test_bsr_from_rows_sums_repeated_columns():
    rng = np.random.default_rng(3)
    nc = 5
    blocks = rng.normal(size=(nc, 4, 4, 4))
    cols = np.array([[0, 1, -1, 2], [1, 0, 0, 4], [2, -1, -1, -1], [3, 4, 2, 1], [4, 3, 1, -1]], np.int32)
    A = implicit.bsr_from_rows(blocks, cols).toarray()
    D = np.zeros((4 * nc, 4 * nc))
    for i in range(nc):
        for k in range(4):
            j = cols[i, k]
            if j >= 0:
                D[4 * i:4 * i + 4, 4 * j:4 * j + 4] += blocks[i, k]
    np.testing.assert_array_equal(A, D)


def bump(mach=0.6, flux="LAX", **kw):
    return cases.bump_channel(8, height=0.06, mach=mach, outlet="BOUND_OUTLET", flux=flux, limiter="NONE", **kw)


@pytest.mark.parametrize("flux", ["LAX", "GODUNOV"])
def test_steady_run_drives_the_residual_down(flux):
    c = bump(flux=flux)
    s = implicit.ImplicitSolver(c, steady=True, cfl=1.0, max_lim_cells=50, assemble=oracle_assembler(c))
    o = oracle.OracleSolver(c)

    def rhs_norm():
        o.set_state(s.U)
        return np.abs(o.implicit_assemble(steady=True, cfl=1.0)[2]).max()

    r0 = rhs_norm()
    s.run(30)
    assert s.step_count == 30 and s.solver_err == 0
    assert rhs_norm() < 0.15 * r0
    assert np.all(np.isfinite(s.U)) and not (s.flags & implicit.CELL_FLAG_LIM).any()


def test_cfl_ramp():
    c = bump()
    s = implicit.ImplicitSolver(c, steady=True, cfl=0.5, cfl_scale=1.5, cfl_max=1.0, cfl_step=2,
                                assemble=oracle_assembler(c))
    seen = []
    for _ in range(4):
        assert s.step() == implicit.RESULT_OK
        seen.append(s.cfl)
    assert seen == [0.5, 0.75, 0.75, 1.0]


def test_unsteady_time_and_solver_failure_path():
    c = bump(mach=0.5)
    s = implicit.ImplicitSolver(c, steady=False, tau=1e-3, assemble=oracle_assembler(c))
    s.run(5)
    assert s.step_count == 5 and abs(s.t - 5e-3) < 1e-15
    # a solve that cannot meet the tolerance leaves the state untouched
    s2 = implicit.ImplicitSolver(c, steady=False, tau=1e-1, eps=1e-300, max_iter=2, assemble=oracle_assembler(c))
    U0 = s2.state()
    assert s2.step() != implicit.RESULT_OK
    np.testing.assert_array_equal(s2.state(), U0)
    # steady: the failed step is retried at a lower CFL, the run gives up loudly
    s3 = implicit.ImplicitSolver(c, steady=True, cfl=4.0, eps=1e-300, max_iter=2, assemble=oracle_assembler(c))
    with pytest.raises(RuntimeError):
        s3.run(3)
    assert s3.cfl == 0.1 and s3.step_count == 1


def test_limits_clamp_flag_and_remediate():
    c = bump(mach=0.5, p_min=0.5)
    s = implicit.ImplicitSolver(c, steady=False, tau=1e-2, assemble=oracle_assembler(c))
    nb = s._topology()
    S = s._S
    i = int(np.nonzero((nb >= 0).all(axis=1))[0][5])  # an interior cell
    U0 = s.state()
    x = np.zeros((s.nc, 4))
    x[i, 3] = -0.4 * U0[3, i]  # energy drop: p falls below p_min
    x[i + 1, 0] = 1e-3
    s.flags[i + 2] |= implicit.CELL_FLAG_LIM  # frozen: its increment is ignored
    x[i + 2, 0] = 1.0
    s._apply(x.reshape(-1))
    assert s.flags[i] & implicit.CELL_FLAG_LIM and not s.flags[i + 1] & implicit.CELL_FLAG_LIM
    r = s.U[0, i]
    u, v = s.U[1, i] / r, s.U[2, i] / r
    p = (1.4 - 1.0) * (s.U[3, i] - 0.5 * r * (u * u + v * v))
    assert r == U0[0, i] and abs(p - 0.5) <= 1e-15
    assert s.U[0, i + 1] == U0[0, i + 1] + 1e-3 and s.U[0, i + 2] == U0[0, i + 2]
    # remediation: S-weighted mean of the edge neighbours, flag counter +0x10000
    s.flags[i + 2] = implicit.CELL_FLAG_LIM | 0x1F0000  # its 32nd flagged step: the flag clears
    Ub = s.state()
    s._remediate()
    js = nb[i]
    np.testing.assert_allclose(s.U[:, i], (Ub[:, js] * S[js]).sum(axis=1) / S[js].sum(), rtol=1e-15)
    assert s.flags[i] == implicit.CELL_FLAG_LIM | 0x10000
    assert s.flags[i + 2] == 0

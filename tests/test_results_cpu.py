"""BatchSolve: the array form of a batch's results and its per-job SolveResults
(host logic only)."""
import ctypes

import numpy as np

import fastercap_amd as fc
from fastercap_amd import _ffi


def _stats(n):
    sts = (_ffi.SolveStats * n)()
    for i in range(n):
        sts[i].n_rhs = 2
        sts[i].iterations[0] = 10 + i
        sts[i].iterations[1] = 20 + i
        sts[i].max_iterations = 20 + i
        sts[i].final_residual = 1e-7 * (i + 1)
        sts[i].converged = i % 2
        sts[i].cycles = i + 1
        sts[i].t_solve_s = 0.5 * i
        sts[i].n_mvp = 3 * i
    return sts


def test_results_follow_the_arrays():
    # job 0: 2 x 3 real, job 1: 1 x 2 complex (real parts, then imaginary), job 2: 2 x 2 real
    rows = np.array([2, 1, 2], dtype=np.int32)
    cols = np.array([3, 2, 2], dtype=np.int32)
    cplx = [False, True, False]
    caps = np.arange(6 + 4 + 4, dtype=np.float64)
    sts = _stats(3)
    b = fc.BatchSolve(caps, rows, cols, cplx, np.frombuffer(sts, dtype=np.ctypeslib.as_array(_ffi.SolveStats()).dtype))
    assert len(b) == 3
    res = b.results()
    assert np.array_equal(res[0].cap, np.arange(6.0).reshape(2, 3))
    assert np.array_equal(res[1].cap, np.array([[6.0 + 8.0j, 7.0 + 9.0j]]))
    assert np.array_equal(res[2].cap, np.arange(10.0, 14.0).reshape(2, 2))
    for i, r in enumerate(res):
        assert r.iterations == [10 + i, 20 + i][: rows[i]]
        assert r.residual == 1e-7 * (i + 1)
        assert r.converged == bool(i % 2)
        assert r.cycles == i + 1
        assert r.t_solve_s == 0.5 * i
        assert r.n_mvp == 3 * i
        assert r.t_mvp_s == 0.0 and r.charges is None and r.residuals is None


def test_solve_result_defaults():
    r = fc.SolveResult(cap=np.zeros((1, 1)))
    assert r.iterations is None and r.mem is None and r.job_mem is None and r.converged
    assert ctypes.sizeof(_ffi.SolveStats) > 0

"""The refinement loop of hodlr_solve driven from the device: the convergence
check after every FP64 residual runs as a kernel and gates the next sweep, so
a captured solve (fused pass, first residual, k gated sweeps) is one graph
replay.  The graph and eager paths must give bitwise the same solution and
sweep count, and gated-off sweeps must leave no trace."""
import numpy as np
import pytest

from fastsolver_amd import gen, hodlr
from tests.gpu_util import host_system, oracle_solve, rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _setup(prec):
    spec = gen.config(3, scale=16)[0]          # N = 16384, leaves of 1024, rank 100, 8 rhs
    X, V, K = host_system(spec)
    Xo = oracle_solve(spec, X, V, K)
    ds = hodlr.DeviceSystem.from_host(spec, X, V, K)
    return spec, ds, Xo[:, :spec.nrhs].copy(order="F"), hodlr.Plan([ds], prec)


@pytest.mark.parametrize("prec", [hodlr.PREC_FP32, hodlr.PREC_BF16])
def test_graph_solve_equals_eager_solve(prec):
    spec, ds, xo, plan = _setup(prec)
    eager = plan.solve(TOL, 30)
    x_eager = ds.solution()
    assert eager.converged and eager.iterations >= 1
    plan.use_graph(True)
    for rep_i in range(3):       # 1st: sweeps replayed one by one; then one replay holding them all
        rep = plan.solve(TOL, 30)
        x = ds.solution()
        print(f"  replay {rep_i}: {rep.iterations} sweeps, relres {rep.relres:.3e}", flush=True)
        assert rep.iterations == eager.iterations
        assert rep.relres == eager.relres and rep.relres_first == eager.relres_first
        assert np.array_equal(x, x_eager)
    assert rel(x_eager, xo) <= TOL
    plan.close()


def test_gated_sweeps_leave_no_trace():
    """A graph captured with three sweeps, replayed for a looser tolerance that
    one sweep meets: the check switches the other two off, and the result is
    the eager one-sweep result bit for bit."""
    spec, ds, xo, plan = _setup(hodlr.PREC_BF16)
    loose = 1e-6
    one = plan.solve(loose, 30)
    x_one = ds.solution()
    assert one.converged and one.iterations == 1, one.iterations
    full = plan.solve(TOL, 30)
    assert full.iterations >= 3
    plan.use_graph(True)
    plan.solve(TOL, 30)                         # captures the solve with `full.iterations` sweeps
    rep = plan.solve(TOL, 30)
    assert rep.iterations == full.iterations
    rep = plan.solve(loose, 30)                 # same graph, loose tolerance
    assert rep.iterations == 1 and rep.relres == one.relres
    assert np.array_equal(ds.solution(), x_one)
    plan.close()


def test_sweep_budget_and_resume():
    """maxiter caps the sweeps (not converged, reported as such); the next
    solve replays the shorter graph and finishes with one-sweep replays."""
    spec, ds, xo, plan = _setup(hodlr.PREC_BF16)
    ref = plan.solve(TOL, 30)
    x_ref = ds.solution()
    plan.use_graph(True)
    capped = plan.solve(TOL, 1, raise_on_fail=False)
    assert capped.iterations == 1 and not capped.converged
    rep = plan.solve(TOL, 30)
    assert rep.converged and rep.iterations == ref.iterations
    assert np.array_equal(ds.solution(), x_ref)
    assert rel(x_ref, xo) <= TOL
    plan.close()

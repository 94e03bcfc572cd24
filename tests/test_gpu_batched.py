"""The batched-uneven workload family at full size: 512 independent HODLR
systems, N uniform in [4096, 65536] in steps of 256, leaves <= 256, per-level
ranks in [8, 64], 1..16 right-hand sides, all in ONE plan (one launch per phase
per height).  FP32 and BF16 factors + FP64 refinement and plain FP64, checked
by residual on every system and against the oracle on a sample of 16 systems
(every 32nd)."""
import numpy as np
import pytest

from fastsolver_amd import gen, hodlr
from tests.gpu_util import host_system, oracle_solve, rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.slow
def test_uneven_batch_512_systems():
    specs = gen.config(4)
    assert len(specs) == 512
    print(f"512 systems, N in [{min(s.n for s in specs)}, {max(s.n for s in specs)}], "
          f"sum N {sum(s.n for s in specs)}", flush=True)
    dss, checked = [], {}
    for i, s in enumerate(specs):
        X, V, K = host_system(s)
        dss.append(hodlr.DeviceSystem.from_host(s, X, V, K))
        if i % 32 == 0:            # oracle on every 32nd system
            checked[i] = oracle_solve(s, X, V, K)[:, :s.nrhs].copy(order="F")
        del X, V, K
    print(f"  uploaded; oracle on {len(checked)} systems", flush=True)
    p32 = hodlr.Plan(dss, hodlr.PREC_FP32)
    rep = p32.solve(tol=TOL, maxiter=10)
    print(f"  fp32+refine: first {rep.relres_first:.2e}, {rep.iterations} sweeps, relres {rep.relres:.2e}", flush=True)
    assert rep.converged and rep.relres <= TOL
    _, per = p32.residual()
    assert np.all(per <= TOL)
    for i, xo in checked.items():
        assert rel(dss[i].solution(), xo) <= TOL, (i, specs[i])
    p32.close()
    p16 = hodlr.Plan(dss, hodlr.PREC_BF16)
    rep = p16.solve(tol=TOL, maxiter=10)
    print(f"  bf16+refine: first {rep.relres_first:.2e}, {rep.iterations} sweeps, relres {rep.relres:.2e}", flush=True)
    assert rep.converged and rep.relres <= TOL
    assert 1e-6 < rep.relres_first < 1e-2
    _, per = p16.residual()
    assert np.all(per <= TOL)
    for i, xo in checked.items():
        assert rel(dss[i].solution(), xo) <= TOL, (i, specs[i])
    p16.close()
    p64 = hodlr.Plan(dss, hodlr.PREC_FP64)
    p64.factor_solve()
    rc, _, _ = p64.status()
    assert rc == 0
    relres, per = p64.residual()
    print(f"  fp64: relres {relres:.2e}", flush=True)
    assert relres <= TOL
    for i, xo in checked.items():
        assert rel(dss[i].solution(), xo) <= TOL, (i, specs[i])


def test_uneven_batch_after_b_changes():
    """||b|| is taken at every residual: a plan created before its right-hand
    side is filled, or reused with a new b, still reports the true relative
    residual and converges to the new solution."""
    specs = gen.config(4, scale=64)
    hosts = [host_system(s) for s in specs]
    dss = [hodlr.DeviceSystem.from_host(s, *h) for s, h in zip(specs, hosts)]
    plan = hodlr.Plan(dss, hodlr.PREC_FP32)
    rep = plan.solve(tol=TOL, maxiter=10)
    assert rep.converged
    rng = np.random.default_rng(3)
    for s, h, d in zip(specs, hosts, dss):
        X = h[0].copy(order="F")
        X[:, : s.nrhs] = 100.0 * rng.random((s.n, s.nrhs))
        d.X.upload(X)
        h[0][:] = X
    rep = plan.solve(tol=TOL, maxiter=10)
    assert rep.converged and rep.relres <= TOL
    mx, per = plan.residual()
    assert mx <= TOL
    for s, h, d in zip(specs, hosts, dss):
        xo = oracle_solve(s, *h)[:, : s.nrhs]
        assert rel(d.solution(), xo) <= TOL, s


@pytest.mark.parametrize("prec", [hodlr.PREC_FP32, hodlr.PREC_BF16])
def test_leaf_sizes_with_partial_row_groups(prec):
    """Leaves of 129..256 rows whose last 64-row block ends inside a 16-row
    group or leaves whole 16-row groups empty (the BF16 TRSM's waves past a
    leaf's last row must contribute nothing), in one plan with the oracle on
    every system."""
    sizes = [129, 161, 174, 193, 205, 207, 215, 220, 223, 239, 241, 250, 256]
    specs = [gen.SystemSpec(4 * m, 256, 1 + (i % 5), [24 + i, 9 + i], seed=1245667 + i, d_scale=1 / 16)
             for i, m in enumerate(sizes)]
    hosts = [host_system(s) for s in specs]
    dss = [hodlr.DeviceSystem.from_host(s, *h) for s, h in zip(specs, hosts)]
    plan = hodlr.Plan(dss, prec)
    rep = plan.solve(tol=TOL, maxiter=10)
    print(f"  prec {prec}: first {rep.relres_first:.2e}, {rep.iterations} sweeps, relres {rep.relres:.2e}", flush=True)
    assert rep.converged and rep.relres <= TOL
    for s, h, d in zip(specs, hosts, dss):
        assert gen.tree_info(s)[0] == 4
        xo = oracle_solve(s, *h)[:, : s.nrhs]
        assert rel(d.solution(), xo) <= TOL, s.n

"""The threaded reference-structure CPU iteration (oracle/cpu_ref.c, timed as
bench.py's CPU baseline) against the serial oracle: the same design
iterations (KE computed twice, K re-assembled, a smoothed-aggregation
multigrid rebuilt every iteration under FGMRES(30), Jacobi-CG filter solves,
the oracle's MMA) give the oracle's compliance, volume constraint and design
change to the solvers' tolerance, with the multigrid taking a small fraction
of the oracle's Jacobi-CG iterations."""
import numpy as np
import pytest


@pytest.mark.parametrize("ke_cache", [0, 1])
def test_cpu_ref_matches_oracle(O, ke_cache):
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.1875, rtol_el=1e-10, rtol_f=1e-10, ke_cache=ke_cache)
    a = O.cantilever(8, 16, 8, 8, jitter=0.15, seed=7, **prm)
    b = O.cantilever(8, 16, 8, 8, jitter=0.15, seed=7, **prm)
    a.init_design()
    b.init_design()
    R = b.cpu_ref(2)
    assert R.threads == 2
    for it in range(1, 4):
        fx, gx, ch, rc = a.design_iteration(it)
        r = R.design_iteration(it)
        assert rc == 0 and r["rc"] == 0
        assert abs(r["fx"] - fx) <= 1e-8 * fx, (it, r["fx"], fx)
        assert abs(r["gx"] - gx[0]) <= 1e-10
        assert abs(r["ch"] - ch) <= 1e-7
        assert r["relres"] <= 1e-10
        assert r["levels"] >= 2
        assert 0 < r["krylov"] < a.diag()["it_el"] / 4
        assert set(r["times"]) == set(O.CpuRef.PHASES) and all(v >= 0.0 for v in r["times"].values())
    assert np.abs(a.vec("X") - b.vec("X")).max() <= 1e-7
    assert R.ke_seconds() > 0.0

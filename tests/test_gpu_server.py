"""Serving form through the C-ABI: requests submitted while others run (Poisson arrivals from a
second thread) are admitted as slots and arena space free up; each result equals the same solve run
alone (bitwise, same kernels) and, in the FP32-element form, stays within 1e-6 of the FP64 oracle."""
import threading
import time

import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from mps_gpu import configs

pytestmark = pytest.mark.gpu


def request(i, rng):
    nx, ny, nz = (int(x) for x in rng.integers(2, 7, 3))
    m = configs.perturb(configs.box(nx, ny, nz, lx=nx / 4, ly=ny / 4, lz=nz / 4), 0.1, 500 + i)
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.05)
    return dict(mesh=m, E=float(10 ** rng.uniform(5, 7)), nu=float(rng.uniform(0.2, 0.45)), dir_dof=dof,
                dir_val=val, n_steps=int(rng.integers(40, 160)))


def alone(r, precision):
    with G.Solver(r["mesh"], E=r["E"], nu=r["nu"], dir_dof=r["dir_dof"], dir_val=r["dir_val"],
                  n_steps=r["n_steps"], precision=precision) as s:
        rc, rep = s.run(r["n_steps"])
        assert rc == 0
        return s.get_state()[0]


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_server_results_equal_alone_under_concurrent_arrivals(precision):
    rng = np.random.default_rng(7)
    reqs = [request(i, rng) for i in range(24)]
    ids = [None] * len(reqs)
    # small arena and few slots so that admission has to wait for retirements
    with G.Server(arena_bytes=6 << 20, max_instances=6, precision=precision, steps_per_launch=8) as srv:
        def producer():
            for i, r in enumerate(reqs):
                ids[i] = srv.submit(r["mesh"], E=r["E"], nu=r["nu"], dir_dof=r["dir_dof"], dir_val=r["dir_val"],
                                    n_steps=r["n_steps"])
                time.sleep(rng.exponential(0.002))
        th = threading.Thread(target=producer)
        th.start()
        th.join()
        assert srv.drain(120.0) == 0
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["p99_ms"] >= st["p50_ms"] > 0
        served = []
        for i, r in enumerate(reqs):
            rc, u, lat, qw = srv.result(ids[i])
            assert rc == 0 and u is not None and lat >= qw >= 0
            assert np.array_equal(u, alone(r, precision)), i
            served.append(u)
    for i in (3, 11):  # served result against the FP64 oracle, to the bar of the precision form
        r = reqs[i]
        ref = O.OracleProblem(r["mesh"], E=r["E"], nu=r["nu"], dir_dof=r["dir_dof"], dir_val=r["dir_val"],
                              n_steps=r["n_steps"]).run()["u"]
        tol = 1e-12 if precision == G.PREC_FP64 else 1e-6
        assert np.abs(served[i] - ref).max() <= tol * np.abs(ref).max()


def test_server_rejects_invalid_and_survives_failed_solve():
    rng = np.random.default_rng(9)
    with G.Server(arena_bytes=8 << 20, max_instances=4, precision=G.PREC_FP64) as srv:
        good = request(0, rng)
        bad_mesh = dict(good["mesh"])
        bad_mesh["conn"] = good["mesh"]["conn"].copy()
        bad_mesh["conn"][0] = -1
        with pytest.raises(G.MpsError):
            srv.submit(bad_mesh, n_steps=10)
        # a crushing solve (x faces, -4000 m/s) fails with non-finite u; its neighbour is unaffected
        m = configs.box(3, 3, 3)
        N = m["n_nodes"]
        X = m["X"].reshape(3, -1)
        lo = np.nonzero(X[0] < 1e-12)[0]
        hi = np.nonzero(X[0] > 1 - 1e-12)[0]
        dof = np.concatenate([lo, lo + N, lo + 2 * N, hi]).astype(np.int32)
        val = np.concatenate([np.zeros(3 * lo.size), np.full(hi.size, -4000.0)])
        rid_bad = srv.submit(m, dir_dof=dof, dir_val=val, n_steps=400, scale=0.9)
        rid_good = srv.submit(good["mesh"], E=good["E"], nu=good["nu"], dir_dof=good["dir_dof"],
                              dir_val=good["dir_val"], n_steps=good["n_steps"])
        assert srv.drain(60.0) == 0
        rc_bad, u_bad, _, _ = srv.result(rid_bad)
        rc_good, u_good, _, _ = srv.result(rid_good)
        assert rc_bad == G.MPS_ERR_NAN and u_bad is None
        assert rc_good == 0 and np.array_equal(u_good, alone(good, G.PREC_FP64))

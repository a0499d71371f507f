"""The batched and serving workloads at their own sizes, against the FP64 CPU oracle:

- batched: 16 real instances of the 512-cube workload (side n from 6 to 40 across the drawn sides, per-instance
  E and nu, 1000 steps each) in one batch, each instance against its own oracle solve, in both forms, and bitwise
  equal to the same solve run alone;
- serving: 64 requests of the serving workload (hex8 boxes with N_e log-uniform in [1e3, 2e5], aspect ratio
  1..4, 500-5000 steps) submitted under Poisson arrivals while others run; the 8 requests with the smallest
  N_e x steps (so that the CPU oracle stays within seconds) are checked against the oracle and against the same
  solve run alone.
"""
import os
import threading
import time

import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from mps_gpu import configs

pytestmark = pytest.mark.gpu

TOL = {G.PREC_FP64: 1e-12, G.PREC_FP32_ELEM: 1e-6}


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def oracle_u(c):
    p = O.OracleProblem(c["mesh"], E=c["E"], nu=c["nu"], dir_dof=c["dir_dof"], dir_val=c["dir_val"],
                        bc_mode=O.BC_VELOCITY, vel_ramp_steps=100, scale=0.5, n_steps=c["n_steps"],
                        sum_order=O.SUM_ELEMENT, n_threads=threads())
    r = p.run()
    assert r["rc"] == 0 and r["steps"] == c["n_steps"]
    return r["u"]


def alone(c, precision):
    with G.Solver(c["mesh"], E=c["E"], nu=c["nu"], dir_dof=c["dir_dof"], dir_val=c["dir_val"],
                  n_steps=c["n_steps"], precision=precision) as s:
        rc, rep = s.run(c["n_steps"])
        assert rc == 0 and rep["status"] == G.MPS_DONE
        return s.get_state()[0]


def spread_instances(count=16):
    """`count` instances of the 512-cube workload whose sides spread over the drawn range (smallest and largest
    included)."""
    sols = configs.batch_meshes(512, steps=1000)
    order = sorted(range(len(sols)), key=lambda i: sols[i]["mesh"]["nx"])
    pick = sorted({order[int(round(q * (len(order) - 1)))] for q in np.linspace(0.0, 1.0, count)})
    return [sols[i] for i in pick]


@pytest.fixture(scope="module")
def cfg3_cases():
    cases = spread_instances(16)
    sides = [c["mesh"]["nx"] for c in cases]
    assert len(cases) == 16 and min(sides) == 6 and max(sides) == 40, sides
    print(f"cfg3 sides {sides}; oracle on {threads()} threads", flush=True)
    return cases, [oracle_u(c) for c in cases]


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_cfg3_instances_against_their_oracle(precision, cfg3_cases):
    cases, refs = cfg3_cases
    with G.Batch(arena_bytes=2 << 30, max_instances=32, precision=precision, steps_per_launch=32) as b:
        slots = [b.add(c["mesh"], E=c["E"], nu=c["nu"], dir_dof=c["dir_dof"], dir_val=c["dir_val"],
                       n_steps=c["n_steps"]) for c in cases]
        b.autotune()
        b.step(1000)
        worst = 0.0
        for s, c, ref in zip(slots, cases, refs):
            rep = b.query(s)
            assert rep["status"] == G.MPS_DONE and rep["steps"] == 1000, rep
            u = b.get_state(s)[0]
            err = rel(u, ref)
            worst = max(worst, err)
            assert err <= TOL[precision], (c["mesh"]["nx"], err)
            assert np.array_equal(u, alone(c, precision)), c["mesh"]["nx"]
    print(f"cfg3 16 instances, precision {precision}: worst rel err vs oracle {worst:.3e}", flush=True)


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_cfg4_served_requests_against_the_oracle(precision):
    reqs = configs.serving_requests(64)
    cost = [r["mesh"]["n_elem"] * r["n_steps"] for r in reqs]
    check = sorted(np.argsort(cost)[:8].tolist())
    assert min(r["mesh"]["n_elem"] for r in reqs) >= 800 and max(r["n_steps"] for r in reqs) >= 4000
    ids = [None] * len(reqs)
    rng = np.random.default_rng(4)
    with G.Server(arena_bytes=4 << 30, max_instances=256, precision=precision, steps_per_launch=32) as srv:
        def producer():
            for i, r in enumerate(reqs):
                ids[i] = srv.submit(r["mesh"], E=r["E"], nu=r["nu"], dir_dof=r["dir_dof"], dir_val=r["dir_val"],
                                    n_steps=r["n_steps"])
                time.sleep(rng.exponential(1.0 / 200.0))  # Poisson arrivals at 200 requests/s
        th = threading.Thread(target=producer)
        th.start()
        th.join()
        assert srv.drain(300.0) == 0
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["tunings"] >= 1 and st["elem_block"] in (64, 128, 256) and 1 <= st["grid_mult"] <= 16, st
        served = {}
        for i in range(len(reqs)):
            rc, u, lat, qw = srv.result(ids[i])
            assert rc == 0 and u is not None and lat >= qw >= 0
            if i in check:
                served[i] = u
    print(f"cfg4: {st['solves_per_s']:.1f} solves/s, p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms; "
          f"checking requests {check}", flush=True)
    worst = 0.0
    for i in check:
        r = reqs[i]
        assert np.array_equal(served[i], alone(r, precision)), i
        err = rel(served[i], oracle_u(r))
        worst = max(worst, err)
        assert err <= TOL[precision], (i, r["mesh"]["n_elem"], r["n_steps"], err)
    print(f"cfg4 precision {precision}: worst rel err vs oracle {worst:.3e}", flush=True)

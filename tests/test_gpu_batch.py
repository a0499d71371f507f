"""Batched form through the C-ABI: every solve in a batch must give the same result as the same
solve run alone (single-solve path, bitwise: same kernels, same arithmetic), whatever else is in the
batch and whenever it was admitted; and the same as the CPU oracle to the FP64 / FP32 bars."""
import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from mps_gpu import configs

pytestmark = pytest.mark.gpu


def case(n, seed, steps, tet=False, per_elem=False):
    rng = np.random.default_rng(seed)
    m = configs.box(n, n, max(2, n - 1), tet=tet)
    m = configs.perturb(m, 0.1, seed)
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.1)
    if per_elem:
        E, nu = configs.random_materials(m["n_elem"], seed)
    else:
        E, nu = float(10 ** rng.uniform(5, 7)), float(rng.uniform(0.2, 0.45))
    return dict(mesh=m, E=E, nu=nu, dir_dof=dof, dir_val=val, n_steps=steps)


def alone(c, precision=G.PREC_FP64):
    with G.Solver(c["mesh"], E=c["E"], nu=c["nu"], dir_dof=c["dir_dof"], dir_val=c["dir_val"],
                  n_steps=c["n_steps"], precision=precision) as s:
        rc, rep = s.run(c["n_steps"])
        assert rc == 0 and rep["status"] == G.MPS_DONE
        return s.get_state()[0]


def oracle(c):
    p = O.OracleProblem(c["mesh"], E=c["E"], nu=c["nu"], dir_dof=c["dir_dof"], dir_val=c["dir_val"],
                        n_steps=c["n_steps"])
    return p.run()["u"]


def add(b, c):
    return b.add(c["mesh"], E=c["E"], nu=c["nu"], dir_dof=c["dir_dof"], dir_val=c["dir_val"], n_steps=c["n_steps"])


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_batch_equals_alone_bitwise_and_oracle():
    cases = [case(n, 100 + n, 40 + 17 * n) for n in (3, 4, 5, 6, 7, 9)]
    with G.Batch(arena_bytes=64 << 20, steps_per_launch=16) as b:
        slots = [add(b, c) for c in cases]
        b.step(max(c["n_steps"] for c in cases))
        for s, c in zip(slots, cases):
            rep = b.query(s)
            assert rep["status"] == G.MPS_DONE and rep["steps"] == c["n_steps"], rep
            u, _, _ = b.get_state(s)
            assert np.array_equal(u, alone(c))
        assert rel(b.get_state(slots[2])[0], oracle(cases[2])) <= 1e-12


def test_continuous_batching_admission_and_retirement():
    cases = [case(n, 200 + i, 30 + 11 * i) for i, n in enumerate((4, 6, 3, 8, 5, 7, 4, 6))]
    ref = [alone(c) for c in cases]
    with G.Batch(arena_bytes=48 << 20, steps_per_launch=8, max_instances=4) as b:
        pending = list(range(len(cases)))
        live = {}
        results = {}
        while pending or live:
            while pending and len(live) < 4:  # admit while slots are free
                i = pending.pop(0)
                live[add(b, cases[i])] = i
            b.launch(1)
            b.sync()
            for s in list(live):
                rep = b.query(s)
                if rep["status"] == G.MPS_DONE:
                    results[live[s]] = b.get_state(s)[0]
                    b.remove(s)
                    del live[s]
        for i in range(len(cases)):
            assert np.array_equal(results[i], ref[i]), i


def test_batch_fp32_tet4_heterogeneous_against_fp64_oracle():
    cases = [case(n, 300 + n, 60, tet=True, per_elem=True) for n in (3, 4, 5)]
    with G.Batch(arena_bytes=64 << 20, precision=G.PREC_FP32_ELEM) as b:
        slots = [add(b, c) for c in cases]
        b.step(60)
        for s, c in zip(slots, cases):
            u = b.get_state(s)[0]
            assert np.array_equal(u, alone(c, G.PREC_FP32_ELEM))
            assert rel(u, oracle(c)) <= 1e-6


def test_autotune_keeps_state_and_results():
    cases = [case(n, 400 + n, 90) for n in (5, 6, 7, 8)]
    with G.Batch(arena_bytes=64 << 20, steps_per_launch=10) as b:
        slots = [add(b, c) for c in cases]
        b.step(30)
        bs, gm = b.autotune()
        assert bs in (64, 128, 256) and 1 <= gm <= 16
        assert all(b.query(s)["steps"] == 30 for s in slots)
        b.step(60)
        for s, c in zip(slots, cases):
            assert np.array_equal(b.get_state(s)[0], alone(c))


def test_capacity_errors_are_codes():
    c = case(6, 1, 10)
    with G.Batch(arena_bytes=1 << 20, max_instances=2) as b:
        with pytest.raises(G.MpsError) as ei:
            add(b, case(20, 2, 10))
        assert ei.value.code == G.MPS_ERR_CAPACITY
        add(b, case(3, 3, 10))
        add(b, case(3, 4, 10))
        with pytest.raises(G.MpsError) as ei:
            add(b, case(3, 5, 10))
        assert ei.value.code == G.MPS_ERR_CAPACITY
    assert c

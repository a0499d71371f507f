"""Energies and the energy balance on the device, against the oracle (FP64):

- DEFO_ENERGY, KIN_ENERGY, VOLUME (reference definitions, dataExtraction.cpp:569-776), EXTERNAL_ENERGY u . fext
  (:867-875), the lumped kinetic energy and the trapezoidal work of the loads and of the supports, step by step,
  in the element kernels (FP64 hex8 and tet4) and in the standalone pass (FP32 element form, viscosity);
- energies change nothing in the trajectory, and give the same history with graph replay, direct launches and
  any chunking, also when computed every 3rd step;
- the balance lumped kinetic + deformation energy - work of the loads and supports holds on the 10^3 cube over
  1000 steps (1e-6 of the work at the end, 1e-4 after step 50; the oracle gives 3.4e-8 and 2.5e-5).
"""
import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from meshgen import box, clamp_and_pull, perturb
from mps_gpu import configs

pytestmark = pytest.mark.gpu

KEYS = ("e_def", "e_kin", "volume", "e_ext", "e_kin_lumped", "w_ext", "w_react")


def device_run(mesh, steps, every=1, chunks=None, graphs=True, precision=G.PREC_FP64, loads=None, visc=None, **kw):
    with G.Solver(mesh, n_steps=steps, energy_every=every, history_cap=max(steps, 1), precision=precision, **kw) as s:
        if loads is not None:
            s.set_loads(loads=loads)
        if visc is not None:
            s.set_viscosity(G.VISC_BULK, *visc)
        s.set_graphs(graphs)
        for c in chunks or [steps]:
            rc, rep = s.run(c)
            assert rc == 0, rep
        assert rep["steps"] == steps
        h = s.history(0, steps)
        h.update(s.energy_balance(0, steps))
        return s.get_state()[0], h


def oracle_run(mesh, steps, every=1, loads=None, visc=None, **kw):
    extra = {}
    if visc is not None:
        extra = dict(visc=O.VISC_BULK, visc_c1=visc[0], visc_cl=visc[1])
    p = O.OracleProblem(mesh, n_steps=steps, energy_every=every, loads=loads, **extra, **kw)
    r = p.run(hist_cap=steps)
    assert r["rc"] == 0
    return r["u"], r["hist"]


def compare(h, ho, steps, every, rtol):
    idx = np.array([k for k in range(steps) if (k + 1) % every == 0])
    for key in KEYS:
        rows = np.arange(steps) if key in ("w_ext", "w_react") else idx
        a, b = h[key][rows], ho[key][rows]
        scale = max(np.abs(b).max() if key not in ("w_ext", "w_react") else
                    np.abs(ho["w_ext"]).max() + np.abs(ho["w_react"]).max(), 1e-300)
        assert np.all(np.isfinite(a)), key
        assert np.abs(a - b).max() <= rtol * scale, (key, np.abs(a - b).max() / scale)


@pytest.mark.parametrize("tet", [False, True])
def test_energies_and_work_against_the_oracle(tet):
    m = perturb(box(6, 5, 5, tet=tet), 0.1, 71 + tet)
    dd, dv = clamp_and_pull(m, vel=0.2)
    kw = dict(dir_dof=dd, dir_val=dv, bc_mode=G.BC_VELOCITY, vel_ramp_steps=30)
    u, h = device_run(m, 150, **kw)
    uo, ho = oracle_run(m, 150, **kw)
    assert np.abs(u - uo).max() <= 1e-12 * np.abs(uo).max()
    compare(h, ho, 150, 1, 1e-10)
    assert ho["w_react"][-1] > 0 and np.all(h["w_ext"] == 0.0)


def test_loads_energies_against_the_oracle():
    """Nodal forces, gravity and a pressure ramp: EXTERNAL_ENERGY and the work of the loads (the flagged nodes
    keep their previous force; gravity is constant)."""
    m = perturb(box(4, 4, 5), 0.1, 9)
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    conn = m["conn"].reshape(-1, 8)
    bot = [n for n in range(N) if X[2, n] < 1e-12]
    top_nodes = [n for n in range(N) if X[2, n] > 1 - 1e-12]
    top = [e for e in range(conn.shape[0]) if np.all(X[2, conn[e, 4:]] > 1 - 1e-12)]
    dd = np.array([n + N * d for n in bot for d in range(3)], np.int32)
    loads = [("force", top_nodes[:3], [[200.0, -100.0, 400.0]] * 3, 0.0, 0.02), ("gravity", 9.81),
             ("pressure_ramp", 4e4, 0.0, 0.015, top, [2] * len(top))]
    kw = dict(dir_dof=dd, dir_val=np.zeros(dd.size), bc_mode=G.BC_RAMP, t0=0.0, tf=0.03)
    u, h = device_run(m, 120, loads=loads, **kw)
    uo, ho = oracle_run(m, 120, loads=loads, **kw)
    assert np.abs(u - uo).max() <= 1e-12 * np.abs(uo).max()
    compare(h, ho, 120, 1, 1e-10)
    assert abs(ho["w_ext"][-1]) > 0 and abs(ho["e_ext"][-1]) > 0


@pytest.mark.parametrize("case", ["fp32", "visc"])
def test_standalone_energy_pass(case):
    """The FP32 element form and the viscous form take their energies from the standalone FP64 pass, placed
    between the element and node kernels: same definitions (FP32: to the form's bar)."""
    m = perturb(box(5, 5, 4), 0.1, 13)
    dd, dv = clamp_and_pull(m, vel=0.1)
    kw = dict(dir_dof=dd, dir_val=dv, bc_mode=G.BC_VELOCITY, vel_ramp_steps=20)
    if case == "fp32":
        u, h = device_run(m, 100, precision=G.PREC_FP32_ELEM, **kw)
        uo, ho = oracle_run(m, 100, **kw)
        compare(h, ho, 100, 1, 2e-6)
    else:
        u, h = device_run(m, 100, visc=(0.5, 0.06), **kw)
        uo, ho = oracle_run(m, 100, visc=(0.5, 0.06), **kw)
        assert np.abs(u - uo).max() <= 1e-12 * np.abs(uo).max()
        compare(h, ho, 100, 1, 1e-10)


def test_energies_leave_the_trajectory_and_replay_unchanged():
    """Energies every step or every 3rd step do not change u; the history is the same with graph replay, with
    direct launches, and over different chunkings (the kinetic energy of a chunk's last step comes from the pass
    after it, the others from the next step's element kernel)."""
    m = perturb(box(6, 6, 5), 0.1, 5)
    dd, dv = clamp_and_pull(m)
    kw = dict(dir_dof=dd, dir_val=dv, bc_mode=G.BC_VELOCITY, vel_ramp_steps=40)
    u0, _ = device_run(m, 90, every=0, **kw)
    for every in (1, 3):
        ua, ha = device_run(m, 90, every=every, **kw)
        ub, hb = device_run(m, 90, every=every, graphs=False, chunks=[7, 50, 33], **kw)
        uc, hc = device_run(m, 90, every=every, chunks=[1, 1, 2, 86], **kw)
        assert np.array_equal(ua, u0) and np.array_equal(ub, u0) and np.array_equal(uc, u0)
        for key in ("time", "dt") + KEYS:
            if key == "e_kin":  # element kernel or end-of-launch pass: the same formulas, compiled apart
                np.testing.assert_allclose(ha[key], hb[key], rtol=1e-14, equal_nan=True)
                np.testing.assert_allclose(ha[key], hc[key], rtol=1e-14, equal_nan=True)
                continue
            assert np.array_equal(ha[key], hb[key], equal_nan=True), (every, key)
            assert np.array_equal(ha[key], hc[key], equal_nan=True), (every, key)
        uo, ho = oracle_run(m, 90, every=every, **kw)
        compare(ha, ho, 90, every, 1e-10)
        off = [k for k in range(90) if (k + 1) % every != 0]
        assert np.all(np.isnan(ha["e_def"][off])) and np.all(np.isnan(ha["e_kin"][off]))


def test_time_window_run_ends_mid_launch():
    """Time-window solve that ends inside a launch: the kinetic energy of its last step is still written."""
    m = box(4, 4, 4)
    dd, dv = clamp_and_pull(m, vel=0.002)
    kw = dict(dir_dof=dd, dir_val=dv, bc_mode=G.BC_RAMP, t0=0.0, tf=0.05)
    with G.Solver(m, n_steps=0, energy_every=1, history_cap=4096, **kw) as s:
        rc, rep = s.run(10 ** 6)
        assert rc == 0 and rep["status"] == G.MPS_DONE
        n = rep["steps"]
        h = s.history(0, n)
    o = O.OracleProblem(m, n_steps=0, energy_every=1, **kw)
    ho = o.run(hist_cap=4096)["hist"]
    assert n > 10 and n % 256 != 0 and np.isfinite(h["e_kin"]).all()
    np.testing.assert_allclose(h["e_kin"], ho["e_kin"][:n], rtol=1e-10)


def test_cfg0_energy_balance():
    cfg = configs.config(0)
    kw = dict(dir_dof=cfg["dir_dof"], dir_val=cfg["dir_val"], bc_mode=G.BC_VELOCITY, vel_ramp_steps=100)
    u, h = device_run(cfg["mesh"], 1000, **kw)
    E = h["e_def"] + h["e_kin_lumped"]
    W = h["w_ext"] + h["w_react"]
    B = (E - E[0]) - W
    scale = np.maximum.accumulate(np.maximum(np.abs(W), np.abs(h["e_def"] - h["e_def"][0]) + h["e_kin_lumped"]))
    rel = np.abs(B) / scale
    print(f"cfg0 balance: end {rel[-1]:.3e}, max after step 50 {rel[49:].max():.3e}, work {W[-1]:.6e} J", flush=True)
    assert rel[-1] <= 1e-6 and rel[49:].max() <= 1e-4

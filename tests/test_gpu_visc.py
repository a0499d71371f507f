"""Bulk artificial viscosity on the HIP path (mps_set_viscosity) against the oracle.

Every Gauss point adds q cof(F) to its PK1, q from tr(F - F_prev)/dt (bulkViscosity,
solverExplicit.cpp:505-542); the device takes F - F_prev from the displacement increment of the step, the
reference (and the oracle) as the difference of the two F, so the two differ by rounding only.
Tolerances as in test_gpu_parity: FP64 <= 1e-12, FP32-element <= 1e-6 (relative displacement).
"""
import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from meshgen import box, clamp_and_pull, perturb

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def visc_pair(m, steps, visc, c1, cl, h=0.0, c0=0.0, precision=G.PREC_FP64, bc_mode=O.BC_VELOCITY, tf=1.0,
              dirs=None, state0=None, law=0):
    dd, dv = dirs if dirs is not None else clamp_and_pull(m, vel=0.5)
    op = O.OracleProblem(m, law=law, dir_dof=dd, dir_val=dv, bc_mode=bc_mode, vel_ramp_steps=20, t0=0.0, tf=tf,
                         scale=0.5, n_steps=steps, sum_order=O.SUM_REFERENCE, visc=visc, visc_c1=c1, visc_cl=cl,
                         visc_h=h, visc_c0=c0)
    u0, v0 = state0 if state0 is not None else (None, None)
    ro = op.run(u0, v0)
    with G.Solver(m, law=law, dir_dof=dd, dir_val=dv, bc_mode=bc_mode, vel_ramp_steps=20, t0=0.0, tf=tf, scale=0.5,
                  n_steps=steps, precision=precision) as s:
        s.set_viscosity(visc, c1, cl, h, c0)
        if state0 is not None:
            s.set_state(u0, v0)
        rc, rep = s.run(steps if steps else 10 ** 9)
        u = s.get_state()[0]
    return ro, rc, rep, u


@pytest.mark.parametrize("tet", [False, True])
def test_bulk_viscosity_fp64_fp32(tet):
    m = perturb(box(4, 4, 4, tet=tet), 0.1, 41)
    base = visc_pair(m, 200, O.VISC_NONE, 0.0, 0.0)
    ro, rc, rep, u = visc_pair(m, 200, O.VISC_BULK, 0.5, 0.06)
    assert ro["rc"] == 0 and rc == 0 and rep["steps"] == 200
    err = relerr(u, ro["u"])
    # the viscosity changes the motion well beyond the tolerance
    assert relerr(ro["u"], base[0]["u"]) > 1e-6
    _ro, rc32, _rep, u32 = visc_pair(m, 200, O.VISC_BULK, 0.5, 0.06, precision=G.PREC_FP32_ELEM)
    err32 = relerr(u32, ro["u"])
    print(f"tet={tet}: fp64 {err:.2e} fp32 {err32:.2e}; viscous vs inviscid {relerr(ro['u'], base[0]['u']):.2e}")
    assert err <= TOL64 and rc32 == 0 and err32 <= TOL32


def test_dev_viscosity_changes_nothing_and_explicit_size():
    m = perturb(box(4, 3, 3), 0.1, 43)
    ro, rc, rep, u = visc_pair(m, 150, O.VISC_NONE, 0.0, 0.0)
    _r, rc2, _rep, ud = visc_pair(m, 150, O.VISC_DEV, 1.0, 0.5)
    assert rc == 0 and rc2 == 0 and np.array_equal(u, ud)
    ro, rc, rep, u = visc_pair(m, 150, O.VISC_BULK, 0.3, 0.05, h=0.2, c0=40.0)
    assert rc == 0 and relerr(u, ro["u"]) <= TOL64


def test_bulk_viscosity_time_window_and_handover():
    """Reference time-window loop with DISPLACEMENT RAMP, and a solver handed a state (the previous state of its
    first step is the handed-over u)."""
    m = box(4, 4, 4)
    dd, dv = clamp_and_pull(m, vel=0.01)
    ro, rc, rep, u = visc_pair(m, 0, O.VISC_BULK, 0.5, 0.06, bc_mode=O.BC_RAMP, tf=0.02, dirs=(dd, dv))
    assert rc == 0 and rep["steps"] == ro["steps"] and relerr(u, ro["u"]) <= TOL64
    rng = np.random.default_rng(3)
    N = m["n_nodes"]
    u0 = rng.uniform(-1e-3, 1e-3, 3 * N)
    v0 = rng.uniform(-0.05, 0.05, 3 * N)
    ro, rc, rep, u = visc_pair(m, 80, O.VISC_BULK, 0.5, 0.06, dirs=(dd, dv), state0=(u0, v0))
    assert rc == 0 and relerr(u, ro["u"]) <= TOL64


def test_set_viscosity_errors():
    m = box(2, 2, 2)
    with G.Solver(m, n_steps=3) as s:
        with pytest.raises(G.MpsError) as ei:
            s.set_viscosity(5)
        assert ei.value.code == G.MPS_ERR_ARG
        with pytest.raises(G.MpsError):
            s.set_viscosity(G.VISC_BULK, np.nan)
        s.set_viscosity(G.VISC_BULK)
        rc, _ = s.run(1)
        assert rc == 0
        with pytest.raises(G.MpsError) as ei:
            s.set_viscosity(G.VISC_NONE)
        assert ei.value.code == G.MPS_ERR_STATE

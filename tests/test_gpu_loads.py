"""Neumann loads on the HIP path (mps_set_loads / mps_external_force) against the oracle.

The node kernel adds the external force of each step -- nodal force entries (*BOUNDARY, TYPE=FORCE),
pressure surface forces (*BOUNDARY, TYPE=PRESSURE RAMP|INST, computed by a surface kernel from the step's
displacement) and gravity (*GRAVITY), in input order (NeumannBCManagement, boundaryConditions.cpp:75-140) --
before the lumped-mass solve.  Gravity is summed per node once at set-up, in the reference's element and
Gauss-point order, so it is bitwise the oracle's wherever no force entry precedes it on the same
node; where one does, the two differ by the rounding of the order only (1e-14 of the largest load).

Tolerances: FP64 <= 1e-12 (relative displacement), as in test_gpu_parity.  The FP32-element form is held to
1e-5 here, not 1e-6: these loads stress the body to strains of order 1e-2, and each element's forces carry
the FP32 rounding of that stress (relative 6e-8), which the assembled net nodal forces -- differences of
large element contributions -- do not cancel.  Measured after 250 steps: 1.0-1.8e-6 for hex8, 1-4e-7 for
tet4; the unloaded parity workloads stay below 1e-6.
"""
import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from meshgen import box, face_nodes, perturb

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32_LOADED = 1e-5


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def clamp_bottom(mesh):
    N = mesh["n_nodes"]
    bot = face_nodes(mesh, 2, mesh["X"].reshape(3, -1)[2].min())
    dd = np.array([n + N * d for n in bot for d in range(3)], np.int32)
    return dd, np.zeros(dd.size)


def top_forces(mesh, t0, tf, scale=2e3, seed=1):
    """Force entries on the top face in input order, one node listed twice, a window per entry."""
    rng = np.random.default_rng(seed)
    top = face_nodes(mesh, 2, mesh["X"].reshape(3, -1)[2].max())
    nodes = np.concatenate([top, top[:1]])
    vec = rng.uniform(-1, 1, size=(nodes.size, 3)) * scale
    w0 = np.full(nodes.size, t0)
    w1 = np.full(nodes.size, tf)
    w1[::3] = 0.5 * (t0 + tf)  # some entries stop half-way: 0 after tf, as getTracAmount gives
    return nodes.astype(np.int32), vec, w0, w1


def loaded_pair(mesh, steps, forces, gravity, gravity_pos, precision=G.PREC_FP64, dirs=None, bc_mode=O.BC_VELOCITY,
                t0=0.0, tf=1.0, law=0, rho=1000.0, loads=None):
    dd, dv = dirs if dirs is not None else clamp_bottom(mesh)
    op = O.OracleProblem(mesh, law=law, rho=rho, dir_dof=dd, dir_val=dv, bc_mode=bc_mode, t0=t0, tf=tf, scale=0.5,
                         n_steps=steps, sum_order=O.SUM_REFERENCE, gravity=gravity, forces=forces,
                         gravity_pos=gravity_pos, loads=loads)
    ro = op.run(hist_cap=max(steps, 1) if steps else 20000)
    with G.Solver(mesh, law=law, rho=rho, dir_dof=dd, dir_val=dv, bc_mode=bc_mode, t0=t0, tf=tf, scale=0.5,
                  n_steps=steps, precision=precision) as s:
        s.set_loads(forces=forces, gravity=gravity, gravity_pos=gravity_pos, loads=loads)
        rc, rep = s.run(steps if steps else 10 ** 9)
        u, v, a = s.get_state()
    return ro, rc, rep, u, v


@pytest.mark.parametrize("tet", [False, True])
def test_external_force_matches_oracle(tet):
    m = perturb(box(4, 3, 5, tet=tet), 0.12, 21)
    N = m["n_nodes"]
    forces = top_forces(m, 0.1, 0.9)
    for gpos in (0, 2, -1):
        op = O.OracleProblem(m, rho=1100.0, n_steps=1, gravity=9.81, forces=forces, gravity_pos=gpos)
        with G.Solver(m, rho=1100.0, n_steps=1) as s:
            assert np.all(s.external_force(0.3) == 0.0)
            s.set_loads(forces=forces, gravity=9.81, gravity_pos=gpos)
            for t in (0.0, 0.1, 0.3, 0.5, 0.9, 0.95):
                fd, fo = s.external_force(t), op.external_force(t)
                if gpos == 0:
                    assert np.array_equal(fd, fo), (gpos, t)
                else:
                    np.testing.assert_allclose(fd, fo, rtol=0, atol=1e-14 * np.abs(fo).max())
                    # nodes without force entries: gravity alone, bitwise
                    free = np.setdiff1d(np.arange(N), forces[0])
                    assert np.array_equal(fd[2 * N + free], fo[2 * N + free])
    # gravity only: -g times the lumped mass, bitwise the oracle's GP-order sum
    op = O.OracleProblem(m, rho=1100.0, n_steps=1, gravity=-3.5)
    with G.Solver(m, rho=1100.0, n_steps=1) as s:
        s.set_loads(gravity=-3.5)
        fd = s.external_force(0.0)
        assert np.array_equal(fd, op.external_force(0.0))
        np.testing.assert_allclose(fd[2 * N:], 3.5 * s.mass(), rtol=1e-12)


@pytest.mark.parametrize("gpos", [0, -1])
@pytest.mark.parametrize("tet", [False, True])
def test_loaded_cantilever_fp64_fp32(tet, gpos):
    m = perturb(box(3, 3, 6, lz=2.0, tet=tet), 0.1, 5)
    forces = top_forces(m, 0.0, 0.4)
    steps = 300
    ro, rc, rep, u, _v = loaded_pair(m, steps, forces, 9.81, gpos)
    assert ro["rc"] == 0 and rc == 0 and rep["steps"] == steps
    err = relerr(u, ro["u"])
    print(f"tet={tet} gpos={gpos} fp64 rel err {err:.3e}, time {rep['time']:.4f}")
    assert err <= TOL64
    _ro, rc32, _rep, u32, _v32 = loaded_pair(m, steps, forces, 9.81, gpos, precision=G.PREC_FP32_ELEM)
    err32 = relerr(u32, ro["u"])
    print(f"fp32-element rel err {err32:.3e}")
    assert rc32 == 0 and err32 <= TOL32_LOADED
    # the loads matter: without them the clamped body stays at rest
    assert np.abs(u).max() > 1e-6


def test_time_window_run_with_ramp_and_forces():
    """Reference time-window loop with DISPLACEMENT RAMP and force windows that end before tf."""
    m = box(4, 4, 4)
    N = m["n_nodes"]
    dd, dv = clamp_bottom(m)
    side = np.setdiff1d(face_nodes(m, 0, 1.0), face_nodes(m, 2, 0.0))
    dd = np.concatenate([dd, (side + N * 0).astype(np.int32)])
    dv = np.concatenate([dv, np.full(side.size, 0.001)])
    forces = top_forces(m, 0.01, 0.03, scale=5e3, seed=3)
    ro, rc, rep, u, _v = loaded_pair(m, 0, forces, 9.81, 1, dirs=(dd, dv), bc_mode=O.BC_RAMP, t0=0.0, tf=0.05)
    assert ro["rc"] == 0 and rc == 0
    assert rep["status"] == G.MPS_DONE and rep["steps"] == ro["steps"] and rep["time"] == ro["time"]
    assert relerr(u, ro["u"]) <= TOL64


def test_free_fall():
    """Gravity on an unconstrained body: every node accelerates at -g, so u_z follows the scalar central-difference
    recurrence with a_0 = 0 over the run's dt sequence, and x, y stay at rest."""
    m = perturb(box(3, 3, 3), 0.1, 8)
    N = m["n_nodes"]
    steps, g = 50, 9.81
    ro, rc, rep, u, v = loaded_pair(m, steps, None, g, -1, dirs=(np.zeros(0, np.int32), np.zeros(0)))
    assert rc == 0 and relerr(u, ro["u"]) <= TOL64
    uz, vz, az = 0.0, 0.0, 0.0
    for dt in ro["hist"]["dt"][:steps]:
        uz += dt * vz + 0.5 * dt * dt * az
        vz += dt * 0.5 * (az - g)
        az = -g
    np.testing.assert_allclose(u[2 * N:], uz, rtol=1e-10)
    np.testing.assert_allclose(v[2 * N:], vz, rtol=1e-10)
    assert np.abs(u[:2 * N]).max() < 1e-10 * abs(uz)


def test_set_loads_errors():
    m = box(2, 2, 2)
    N = m["n_nodes"]
    with G.Solver(m, n_steps=5) as s:
        for bad in (([N], [[1.0, 0, 0]], [0.0], [1.0]),      # node out of range
                    ([0], [[1.0, 0, 0]], [0.5], [0.5]),      # empty window
                    ([0], [[np.nan, 0, 0]], [0.0], [1.0])):  # non-finite force
            with pytest.raises(G.MpsError) as ei:
                s.set_loads(forces=bad)
            assert ei.value.code == G.MPS_ERR_ARG
        with pytest.raises(G.MpsError):
            s.set_loads(gravity=np.inf)
        s.set_loads(gravity=1.0)
        s.set_loads()  # replaces: no loads
        assert np.all(s.external_force(0.5) == 0.0)
        rc, _rep = s.run(1)
        assert rc == 0
        with pytest.raises(G.MpsError) as ei:
            s.set_loads(gravity=1.0)
        assert ei.value.code == G.MPS_ERR_STATE


def boundary_faces(mesh, axis=2, value=None):
    """(element, face S) of the exterior faces of a box mesh, optionally only those on the plane x_axis = value."""
    nen = mesh["nen"]
    conn = mesh["conn"].reshape(-1, nen)
    X = mesh["X"].reshape(3, -1)
    faces = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]] if nen == 8 else \
        [[0, 1, 2], [0, 3, 1], [1, 3, 2], [0, 2, 3]]
    seen = {}
    for e in range(conn.shape[0]):
        for S, f in enumerate(faces):
            seen.setdefault(tuple(sorted(conn[e, f])), []).append((e, S + 1, f))
    el, fc = [], []
    for v in seen.values():
        if len(v) != 1:
            continue
        e, S, f = v[0]
        if value is not None and np.abs(X[axis, conn[e, f]] - value).max() > 1e-9:
            continue
        el.append(e)
        fc.append(S)
    return np.array(el, np.int32), np.array(fc, np.int32)


@pytest.mark.parametrize("tet", [False, True])
def test_pressure_external_force_matches_oracle(tet):
    """Surface forces at a deformed state, every face of the body, both pressure and both traction types."""
    m = perturb(box(3, 2, 2, tet=tet), 0.1, 17)
    N = m["n_nodes"]
    el, fc = boundary_faces(m)
    rng = np.random.default_rng(5)
    u = rng.uniform(-0.02, 0.02, 3 * N)
    blocks = [("pressure_ramp", 2e3, 0.1, 0.7, el[::2], fc[::2]), ("gravity", 9.81),
              ("force", [3, 7], [[1.0, 2.0, 3.0], [-4.0, 0.5, 0.0]], 0.0, 1.0),
              ("pressure_inst", -7e2, 0.2, 0.9, el[1::2], fc[1::2]),
              ("traction_ramp", (1e2, -3e2, 2e2), 0.0, 0.8, el[::3], fc[::3]),
              ("traction_inst", (-5e1, 0.0, 4e2), 0.3, 0.9, el[1::3], fc[1::3])]
    op = O.OracleProblem(m, rho=1100.0, n_steps=1, loads=blocks)
    with G.Solver(m, rho=1100.0, n_steps=1) as s:
        s.set_state(u)
        s.set_loads(loads=blocks)
        worst = 0.0
        for t in (0.05, 0.1, 0.4, 0.7, 0.8, 0.95):
            fd, fo = s.external_force(t), op.external_force(t, u)
            scale = np.abs(fo).max()
            worst = max(worst, np.abs(fd - fo).max() / scale)
            np.testing.assert_allclose(fd, fo, rtol=0, atol=1e-13 * scale)
        print(f"tet={tet}: max |device - oracle| / max|f| = {worst:.2e}")


@pytest.mark.parametrize("kind", ["pressure_ramp", "pressure_inst"])
@pytest.mark.parametrize("tet", [False, True])
def test_pressure_cantilever_fp64_fp32(tet, kind):
    """A clamped block under a follower pressure on its top face (and a side face), FP64 and FP32-element forms."""
    m = perturb(box(3, 3, 5, lz=1.5, tet=tet), 0.08, 23)
    top = boundary_faces(m, 2, 1.5)
    side = boundary_faces(m, 0, 1.0)
    blocks = [(kind, 4e4, 0.0, 0.3, top[0], top[1]), ("force", [int(m["n_nodes"]) - 1], [[0.0, 300.0, 0.0]], 0.0, 1.0),
              (kind, -1.5e4, 0.05, 0.25, side[0], side[1])]
    steps = 250
    dd, dv = clamp_bottom(m)
    op = O.OracleProblem(m, dir_dof=dd, dir_val=dv, bc_mode=O.BC_VELOCITY, scale=0.5, n_steps=steps,
                         sum_order=O.SUM_REFERENCE, loads=blocks)
    ro = op.run()
    out = {}
    for prec in (G.PREC_FP64, G.PREC_FP32_ELEM):
        with G.Solver(m, dir_dof=dd, dir_val=dv, bc_mode=O.BC_VELOCITY, scale=0.5, n_steps=steps, precision=prec) as s:
            s.set_loads(loads=blocks)
            rc, rep = s.run(steps)
            assert rc == 0 and rep["steps"] == steps
            out[prec] = s.get_state()[0]
    e64, e32 = relerr(out[G.PREC_FP64], ro["u"]), relerr(out[G.PREC_FP32_ELEM], ro["u"])
    print(f"tet={tet} {kind}: fp64 {e64:.2e}, fp32 {e32:.2e}, max|u| {np.abs(ro['u']).max():.3e}")
    assert ro["rc"] == 0 and e64 <= TOL64 and e32 <= TOL32_LOADED
    assert np.abs(ro["u"]).max() > 1e-5


def test_pressure_time_window_run():
    """Reference time-window loop, DISPLACEMENT RAMP and a pressure ramp that ends before tf."""
    m = box(4, 4, 3)
    N = m["n_nodes"]
    dd, dv = clamp_bottom(m)
    top = boundary_faces(m, 2, 1.0)
    blocks = [("pressure_ramp", 1e4, 0.005, 0.02, top[0], top[1]), ("gravity", 9.81)]
    ro, rc, rep, u, _v = loaded_pair(m, 0, None, 0.0, -1, dirs=(dd, dv), bc_mode=O.BC_RAMP, t0=0.0, tf=0.04,
                                     loads=blocks)
    assert ro["rc"] == 0 and rc == 0 and rep["steps"] == ro["steps"] and rep["time"] == ro["time"]
    assert relerr(u, ro["u"]) <= TOL64 and np.abs(u).max() > 0


@pytest.mark.parametrize("tet", [False, True])
def test_traction_run_fp64(tet):
    """A clamped block with a shear traction ramp on its top face and an inward traction on one side."""
    m = perturb(box(3, 3, 4, tet=tet), 0.08, 29)
    top = boundary_faces(m, 2, 1.0)
    side = boundary_faces(m, 0, 1.0)
    blocks = [("traction_ramp", (3e4, 1e4, 0.0), 0.0, 0.2, top[0], top[1]),
              ("traction_inst", (-2e4, 0.0, 0.0), 0.05, 0.15, side[0], side[1])]
    ro, rc, rep, u, _v = loaded_pair(m, 200, None, 0.0, -1, loads=blocks)
    err = relerr(u, ro["u"])
    print(f"tet={tet} traction fp64 {err:.2e}, max|u| {np.abs(ro['u']).max():.3e}")
    assert ro["rc"] == 0 and rc == 0 and err <= TOL64 and np.abs(ro["u"]).max() > 1e-5


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_dof_series_recorded_inside_the_steps(precision):
    """mps_set_dof_series: u, v, a, internal and external force at chosen DOFs (free and constrained), recorded
    by the device inside every step with graph replay and any chunking, equal to what the host reads after each
    step of a run launched one step at a time (u, v, a and the internal force bit for bit; the external force
    against mps_external_force at the step's time)."""
    m = perturb(box(3, 3, 5, lz=1.5), 0.08, 23)
    N = m["n_nodes"]
    top = boundary_faces(m, 2, 1.5)
    blocks = [("pressure_ramp", 4e4, 0.0, 0.3, top[0], top[1]), ("gravity", 9.81),
              ("force", [N - 1], [[0.0, 300.0, 0.0]], 0.0, 1.0)]
    dd, dv = clamp_bottom(m)
    steps = 90
    nodes = [0, N // 2, N - 1, 5]
    entries = [(kind, n + N * d) for n in nodes for kind in range(5) for d in range(3)]
    which = np.array([e[0] for e in entries], dtype=np.int32)
    dof = np.array([e[1] for e in entries], dtype=np.int32)
    kw = dict(dir_dof=dd, dir_val=dv, bc_mode=G.BC_VELOCITY, vel_ramp_steps=30, scale=0.5, n_steps=steps,
              precision=precision, history_cap=steps)
    ref = np.zeros((steps, len(dof)))
    with G.Solver(m, **kw) as s:
        s.set_loads(loads=blocks)
        for k in range(steps):
            rc, rep = s.run(1)
            assert rc == 0
            vals = {0: s.get_dofs(0, dof), 1: s.get_dofs(1, dof), 2: s.get_dofs(2, dof),
                    3: s.internal_force_dofs(dof), 4: s.external_force(rep["time"])[dof]}
            ref[k] = [vals[w][i] for i, w in enumerate(which)]
    for chunks in ([steps], [37, 53], [1, 2, 87]):
        with G.Solver(m, **kw) as s:
            s.set_loads(loads=blocks)
            s.set_dof_series(which, dof)
            for c in chunks:
                rc, rep = s.run(c)
                assert rc == 0
            got = s.dof_series(0, steps)
        exact = which != G.SERIES_FEXT
        assert np.array_equal(got[:, exact], ref[:, exact]), chunks
        fe = ~exact
        np.testing.assert_allclose(got[:, fe], ref[:, fe], rtol=0, atol=1e-13 * np.abs(ref[:, fe]).max())
        assert np.abs(ref[:, which == G.SERIES_FINT]).max() > 0 and np.abs(ref[:, fe]).max() > 0

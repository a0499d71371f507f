"""HIP path (through the C-ABI) against the CPU oracle on the same inputs.

Tolerances (relative, infinity norm over all DOFs, normalised by max |u_oracle|):
  FP64 form:            <= 1e-12  (stated bar for the FP64 GPU path)
  FP32-element form:    <= 1e-6   (stated bar for the precision form, against the FP64 oracle)
"""
import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from mps_gpu import configs
from meshgen import box, clamp_and_pull, perturb

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run_pair(mesh, steps, precision=G.PREC_FP64, law=0, E=1e6, nu=0.3, rho=1000.0, bc_mode=O.BC_VELOCITY,
             dirs=None, t0=0.0, tf=1.0, scale=0.5, energy_every=0, state0=None, chunks=None):
    dd, dv = dirs if dirs is not None else clamp_and_pull(mesh)
    op = O.OracleProblem(mesh, law=law, E=E, nu=nu, rho=rho, dir_dof=dd, dir_val=dv, bc_mode=bc_mode,
                         t0=t0, tf=tf, scale=scale, n_steps=steps, sum_order=O.SUM_REFERENCE,
                         energy_every=energy_every)
    u0 = v0 = a0 = None
    if state0 is not None:
        u0, v0, a0 = state0
    ro = op.run(u0, v0, a0, hist_cap=max(steps, 1) if steps else 20000)
    with G.Solver(mesh, law=law, E=E, nu=nu, rho=rho, dir_dof=dd, dir_val=dv, bc_mode=bc_mode, t0=t0, tf=tf,
                  scale=scale, n_steps=steps, precision=precision, energy_every=energy_every,
                  history_cap=20000) as s:
        if state0 is not None:
            s.set_state(u0, v0, a0)
        total = steps if steps else 10 ** 9
        if chunks:
            rep = None
            for c in chunks:
                rc, rep = s.run(c)
        else:
            rc, rep = s.run(total)
        u, v, a = s.get_state()
        nh = int(rep["steps"])
        hist = s.history(0, nh) if nh else None
    return ro, (rc, rep, u, v, a, hist)


def test_cfg0_fp64_1000_steps_energies():
    m = box(10, 10, 10)
    ro, (rc, rep, u, v, a, hist) = run_pair(m, 1000, energy_every=1)
    assert ro["rc"] == 0 and rc == 0
    assert rep["steps"] == 1000 and ro["steps"] == 1000
    assert rep["status"] == G.MPS_DONE
    assert relerr(u, ro["u"]) <= TOL64
    assert relerr(v, ro["v"]) <= 1e-10
    np.testing.assert_allclose(hist["dt"], ro["hist"]["dt"], rtol=1e-13, atol=0)
    np.testing.assert_allclose(hist["time"], ro["hist"]["time"], rtol=1e-13, atol=0)
    for k in ("e_def", "e_kin", "volume"):
        np.testing.assert_allclose(hist[k], ro["hist"][k], rtol=1e-11, atol=0, err_msg=k)


def test_cfg0_first_dt_bitwise():
    m = box(10, 10, 10)
    dd, dv = clamp_and_pull(m)
    op = O.OracleProblem(m, dir_dof=dd, dir_val=dv, n_steps=1)
    with G.Solver(m, dir_dof=dd, dir_val=dv, n_steps=1) as s:
        assert s.critical_dt() == op.critical_dt()
        rng = np.random.default_rng(5)
        u = rng.uniform(-1e-3, 1e-3, 3 * m["n_nodes"])
        assert s.critical_dt(u) == op.critical_dt(u)


def test_mass_and_internal_force():
    m = perturb(box(6, 5, 4), 0.1, 3)
    dd, dv = clamp_and_pull(m)
    op = O.OracleProblem(m, dir_dof=dd, dir_val=dv, n_steps=1)
    ope = O.OracleProblem(m, dir_dof=dd, dir_val=dv, n_steps=1, sum_order=O.SUM_ELEMENT)
    with G.Solver(m, dir_dof=dd, dir_val=dv, n_steps=1) as s:
        mg = s.mass()
        me = ope.mass()
        assert np.abs(mg - me).max() == 0.0, np.abs(mg - me).max()  # same association: bitwise
        assert relerr(mg, op.mass()) <= 1e-15
        rng = np.random.default_rng(7)
        u = rng.uniform(-1e-3, 1e-3, 3 * m["n_nodes"])
        fo = op.internal_force(u)
        fg = s.internal_force(u)
        assert relerr(fg, fo) <= 1e-12


@pytest.mark.parametrize("precision,law,drop_first", [(G.PREC_FP64, 0, False), (G.PREC_FP64, 2, False),
                                                     (G.PREC_FP32_ELEM, 0, False), (G.PREC_FP32_ELEM, 2, False),
                                                     (G.PREC_FP64, 0, True), (G.PREC_FP32_ELEM, 0, True)])
def test_tet_node_staging_bitwise(precision, law, drop_first, monkeypatch):
    """The tet element kernel that stages each patch's distinct nodes in LDS gives bit for bit the
    results of the kernel that gathers every element slot (MPS_TET_STAGING=0): displacements, dt
    history and the standalone internal force, over a ragged last patch (1260 tets = 4 x 256 + 236;
    without the first tet 1259 = 4 x 256 + 235, an odd count whose contribution list ends inside a
    16-byte piece), heterogeneous materials, and both match the oracle to the form's bar."""
    m = perturb(box(7, 6, 5, tet=True), 0.12, 9 + law)
    if drop_first:  # every node of the first tet belongs to other tets too (the hex's main diagonal)
        m = dict(m, conn=m["conn"][4:].copy(), n_elem=m["n_elem"] - 1)
        assert np.unique(m["conn"]).size == m["n_nodes"]
    assert m["n_elem"] % 256 != 0
    rng = np.random.default_rng(5)
    E = rng.uniform(5e5, 2e6, m["n_elem"])
    nu = rng.uniform(0.2, 0.4, m["n_elem"])
    rho = rng.uniform(800.0, 1200.0, m["n_elem"])
    dd, dv = clamp_and_pull(m)
    uf = rng.uniform(-1e-3, 1e-3, 3 * m["n_nodes"])
    out = {}
    for staged in (True, False):
        if staged:
            monkeypatch.delenv("MPS_TET_STAGING", raising=False)
        else:
            monkeypatch.setenv("MPS_TET_STAGING", "0")
        with G.Solver(m, law=law, E=E, nu=nu, rho=rho, dir_dof=dd, dir_val=dv, n_steps=150,
                      precision=precision, history_cap=200) as s:
            f = s.internal_force(uf)
            rc, rep = s.run(150)
            assert rc == 0 and rep["steps"] == 150
            u, v, a = s.get_state()
            out[staged] = (u, v, f, s.history(0, 150)["dt"])
    for k in range(4):
        assert np.array_equal(out[True][k], out[False][k]), k
    op = O.OracleProblem(m, law=law, E=E, nu=nu, rho=rho, dir_dof=dd, dir_val=dv, n_steps=150,
                         sum_order=O.SUM_REFERENCE)
    ro = op.run(hist_cap=150)
    tol = TOL64 if precision == G.PREC_FP64 else TOL32
    assert relerr(out[True][0], ro["u"]) <= tol
    assert relerr(out[True][2], op.internal_force(uf)) <= (1e-12 if precision == G.PREC_FP64 else 1e-5)


@pytest.mark.parametrize("precision,law,energies", [(G.PREC_FP64, 0, False), (G.PREC_FP64, 1, False),
                                                     (G.PREC_FP64, 0, True), (G.PREC_FP32_ELEM, 0, False)])
def test_hex_node_gathers_by_offset_bitwise(precision, law, energies, monkeypatch):
    """The hex8 element kernels that address the gathered nodes by 32-bit offsets from scalar bases (meshes of
    at most 2^29 nodes) give bit for bit the results of the 64-bit-address kernels (MPS_GATHER32=0, the ones
    larger meshes run): displacements, velocities, dt history, the standalone internal force and, with
    energies every step, the energy history; over a ragged last patch and per-element materials."""
    m = perturb(box(9, 8, 7), 0.12, 21 + law)
    assert m["n_elem"] % 256 != 0
    rng = np.random.default_rng(6)
    E = rng.uniform(5e5, 2e6, m["n_elem"])
    nu = rng.uniform(0.2, 0.4, m["n_elem"])
    dd, dv = clamp_and_pull(m)
    uf = rng.uniform(-1e-3, 1e-3, 3 * m["n_nodes"])
    keys = ("dt", "e_def", "e_kin", "volume") if energies else ("dt",)
    out = {}
    for o32 in (True, False):
        if o32:
            monkeypatch.delenv("MPS_GATHER32", raising=False)
        else:
            monkeypatch.setenv("MPS_GATHER32", "0")
        with G.Solver(m, law=law, E=E, nu=nu, dir_dof=dd, dir_val=dv, n_steps=120, precision=precision,
                      energy_every=1 if energies else 0, history_cap=200) as s:
            f = s.internal_force(uf)
            rc, rep = s.run(120)
            assert rc == 0 and rep["steps"] == 120
            u, v, _a = s.get_state()
            h = s.history(0, 120)
            out[o32] = [u, v, f] + [h[k] for k in keys]
    for k in range(len(out[True])):
        assert np.array_equal(out[True][k], out[False][k], equal_nan=True), k
    if precision == G.PREC_FP64 and not energies:
        op = O.OracleProblem(m, law=law, E=E, nu=nu, dir_dof=dd, dir_val=dv, n_steps=120, sum_order=O.SUM_REFERENCE)
        assert relerr(out[True][0], op.run(hist_cap=120)["u"]) <= TOL64


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_tet_without_shared_nodes_uses_the_gathering_kernel(precision):
    """Tets that share no nodes: a patch of 256 tets has 1024 distinct nodes, whose staging (48 KB) and
    force slots do not fit the 64 KB of LDS a workgroup may use, so the set-up keeps the per-element
    gathers.  The solve must run and match the oracle to the form's bar."""
    base = perturb(box(6, 5, 4, tet=True), 0.1, 77)
    X0 = base["X"].reshape(3, -1)
    conn0 = base["conn"].reshape(-1, 4)
    ne = conn0.shape[0]
    X = X0[:, conn0.ravel()].copy()  # node 4*t + a is a private copy of the tet's node a
    m = dict(base, X=X.ravel(), conn=np.arange(4 * ne, dtype=np.int32), n_nodes=4 * ne)
    assert ne > 256
    tol = TOL64 if precision == G.PREC_FP64 else TOL32
    ro, (rc, rep, u, v, a, hist) = run_pair(m, 150, precision=precision)
    assert rc == 0 and rep["steps"] == 150
    assert np.abs(ro["u"]).max() > 0
    assert relerr(u, ro["u"]) <= tol, relerr(u, ro["u"])


@pytest.mark.parametrize("law", [0, 1])
def test_perturbed_hex_fp64_and_fp32(law):
    m = perturb(box(12, 12, 12), 0.1, 11)
    ro, (rc, rep, u, v, a, hist) = run_pair(m, 200, law=law)
    assert rc == 0 and relerr(u, ro["u"]) <= TOL64, relerr(u, ro["u"])
    ro2, (rc2, rep2, u2, *_r) = run_pair(m, 200, law=law, precision=G.PREC_FP32_ELEM)
    err = relerr(u2, ro2["u"])
    print(f"law {law} fp32-element rel err {err:.3e}")
    assert rc2 == 0 and err <= TOL32


@pytest.mark.parametrize("law", [G.LAW_SVK, G.LAW_LINEAR_ELASTIC])
@pytest.mark.parametrize("tet", [False, True])
def test_svk_and_linear_elastic(law, tet):
    """St-Venant-Kirchhoff and linear elastic laws (sound speed sqrt(E/rho)) in both forms, hex and tet."""
    m = perturb(box(8, 7, 6, tet=tet), 0.1, 31 + law)
    ro, (rc, rep, u, v, a, hist) = run_pair(m, 150, law=law, energy_every=1)
    assert rc == 0 and relerr(u, ro["u"]) <= TOL64, relerr(u, ro["u"])
    ed = ro["hist"]["e_def"][:150]
    assert np.abs(hist["e_def"] - ed).max() <= 1e-10 * np.abs(ed).max()
    ro2, (rc2, rep2, u2, *_r) = run_pair(m, 150, law=law, precision=G.PREC_FP32_ELEM)
    assert rc2 == 0 and relerr(u2, ro2["u"]) <= TOL32, relerr(u2, ro2["u"])


@pytest.mark.parametrize("law", [0, 1, 2, 3])
@pytest.mark.parametrize("tet", [False, True])
def test_less_than_one_wave_of_elements(law, tet):
    """A mesh smaller than one wavefront (8 hexes, or 24 tets): one partial patch whose elements do not
    fill a wave, in both forms and every law, against the oracle."""
    m = perturb(box(2, 2, 2, tet=False) if not tet else box(2, 2, 1, tet=True), 0.1, 51 + law)
    assert m["n_elem"] < 64
    ro, (rc, rep, u, v, a, hist) = run_pair(m, 120, law=law)
    assert rc == 0 and relerr(u, ro["u"]) <= TOL64, relerr(u, ro["u"])
    np.testing.assert_allclose(hist["dt"], ro["hist"]["dt"], rtol=1e-13, atol=0)
    ro2, (rc2, rep2, u2, *_r) = run_pair(m, 120, law=law, precision=G.PREC_FP32_ELEM)
    assert rc2 == 0 and relerr(u2, ro2["u"]) <= TOL32, relerr(u2, ro2["u"])


@pytest.mark.parametrize("law", [0, 1, 2, 3])
@pytest.mark.parametrize("tet", [False, True])
def test_element_outputs(law, tet):
    """VTK element fields (GP-averaged Cauchy, strain, VMS, volumetric/equivalent strain) of the device state
    against the oracle evaluated on the same displacement."""
    m = perturb(box(6, 5, 4, tet=tet), 0.1, 41 + law)
    dd, dv = clamp_and_pull(m, vel=0.5)
    with G.Solver(m, law=law, dir_dof=dd, dir_val=dv, n_steps=80) as s:
        rc, rep = s.run(80)
        assert rc == 0
        u = s.get_state()[0]
        q = s.element_outputs()
    qo = O.OracleProblem(m, law=law, n_steps=1).element_outputs(u)
    assert np.abs(qo[:, :9]).max() > 0
    for cols in (slice(0, 9), slice(9, 18), slice(18, 19), slice(19, 21)):
        scale = max(np.abs(qo[:, cols]).max(), 1e-300)
        assert np.abs(q[:, cols] - qo[:, cols]).max() <= 1e-10 * scale, cols


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_mixed_orientation_hex(precision):
    """Half of the elements (random) use a mirrored local numbering (xi reversed), so det J < 0 there
    and waves hold both signs; the reference integrates with w|det J| (classGPs.cpp:26-27)."""
    m = perturb(box(9, 8, 7), 0.1, 21)
    conn = m["conn"].reshape(-1, 8).copy()
    flip = np.random.default_rng(3).random(conn.shape[0]) < 0.5
    conn[flip] = conn[flip][:, [1, 0, 3, 2, 5, 4, 7, 6]]
    m = dict(m, conn=np.ascontiguousarray(conn.ravel()))
    ro, (rc, rep, u, *_r) = run_pair(m, 150, precision=precision)
    tol = TOL64 if precision == G.PREC_FP64 else TOL32
    assert rc == 0 and relerr(u, ro["u"]) <= tol, relerr(u, ro["u"])


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_tet4_reflected_splits(precision):
    """The 6-tet split in alternating mirror images (conforming, half of the tets negatively oriented, node
    valences up to 48 instead of 24: longer and more uneven CSR rows and patch lists), through the node-staging
    kernel, against the oracle."""
    m = configs.perturb(configs.box(9, 8, 7, tet=True, reflected=True), 0.15, 41)
    ro, (rc, rep, u, *_r) = run_pair(m, 150, precision=precision)
    tol = TOL64 if precision == G.PREC_FP64 else TOL32
    assert rc == 0 and relerr(u, ro["u"]) <= tol, relerr(u, ro["u"])


def test_tet4_heterogeneous_fp64_fp32():
    m = perturb(box(6, 6, 6, tet=True), 0.15, 2)
    rng = np.random.default_rng(3)
    E = 10 ** rng.uniform(5, 7, m["n_elem"])
    nu = rng.uniform(0.2, 0.45, m["n_elem"])
    ro, (rc, rep, u, *_r) = run_pair(m, 150, E=E, nu=nu)
    assert rc == 0 and relerr(u, ro["u"]) <= TOL64, relerr(u, ro["u"])
    ro2, (rc2, rep2, u2, *_r) = run_pair(m, 150, E=E, nu=nu, precision=G.PREC_FP32_ELEM)
    err = relerr(u2, ro2["u"])
    print(f"tet fp32-element rel err {err:.3e}")
    assert rc2 == 0 and err <= TOL32


@pytest.mark.parametrize("mode", [O.BC_RAMP, O.BC_INST])
def test_reference_time_window_mode(mode):
    # reference-expressible run: *Time 0, tf with DISPLACEMENT RAMP / INST on the top face
    m = box(5, 5, 5)
    dd, dv = clamp_and_pull(m, vel=0.002)
    ro, (rc, rep, u, v, a, hist) = run_pair(m, 0, bc_mode=mode, dirs=(dd, dv), t0=0.0, tf=0.05)
    assert ro["rc"] == 0 and rc == 0
    assert rep["status"] == G.MPS_DONE
    assert rep["steps"] == ro["steps"]
    assert rep["time"] == ro["time"]
    assert relerr(u, ro["u"]) <= TOL64


def test_ramp_given_as_rate_equals_ramp():
    """MPS_BC_RAMP_RATE takes UU = U/(tf-t0) as classDirichlet stores it; the run is bitwise the RAMP run."""
    m = box(4, 4, 4)
    dd, dv = clamp_and_pull(m, vel=0.002)
    tf = 0.03
    runs = []
    for mode, vals in ((G.BC_RAMP, dv), (G.BC_RAMP_RATE, dv / (tf - 0.0))):
        with G.Solver(m, dir_dof=dd, dir_val=vals, bc_mode=mode, t0=0.0, tf=tf, n_steps=0) as s:
            rc, rep = s.run(10 ** 9)
            assert rc == 0 and rep["status"] == G.MPS_DONE
            runs.append(s.get_state()[0])
    assert np.array_equal(runs[0], runs[1])


def test_chained_solver_handoff_and_chunks():
    m = box(4, 4, 6)
    ro1, (rc1, rep1, u1, v1, a1, _h) = run_pair(m, 120, chunks=[7, 50, 63])
    assert relerr(u1, ro1["u"]) <= TOL64
    # second solver starting at t0 > 0 from the handed-over state (main.cpp:161-182)
    ro2, (rc2, rep2, u2, *_r) = run_pair(m, 80, t0=0.3, state0=(ro1["u"], ro1["v"], ro1["a"]))
    assert rc2 == 0 and relerr(u2, ro2["u"]) <= TOL64


def test_nan_error_matches_oracle():
    # x = max face driven into the body at -4000 m/s: elements invert, log(J<=0) gives NaN,
    # which reaches DOF 1 (node 1 is free with x-faces constrained): solverExplicit.cpp:306-308
    m = box(3, 3, 3)
    dd, dv = clamp_and_pull(m, vel=-4000.0, axis=0)
    ro, (rc, rep, *_r) = run_pair(m, 400, dirs=(dd, dv), scale=0.9)
    assert ro["rc"] == O.ORC_ERR_NAN
    assert rc == G.MPS_ERR_NAN
    assert rep["steps"] == ro["steps"]


def test_prepared_graphs_give_the_same_run():
    """mps_prepare_run captures the chunk graphs a later mps_run replays: the run that uses them, the run that
    captures as it goes and the run launched kernel by kernel are bitwise the same."""
    m = perturb(box(6, 6, 5), 0.1, 61)
    dd, dv = clamp_and_pull(m)
    out = []
    for mode in ("prepared", "lazy", "plain"):
        with G.Solver(m, dir_dof=dd, dir_val=dv, n_steps=90, history_cap=100) as s:
            if mode == "plain":
                s.set_graphs(False)
            rc, rep = s.run(13)
            assert rc == 0
            if mode == "prepared":
                s.prepare_run(77)
                rc, rep = s.run(0)  # the capture ran nothing
                assert rc == 0 and rep["steps"] == 13
            rc, rep = s.run(77)
            assert rc == 0 and rep["steps"] == 90 and rep["status"] == G.MPS_DONE
            u, v, _ = s.get_state()
            out.append((u, v, s.history(0, 90)["dt"]))
    for other in out[1:]:
        for k in range(3):
            assert np.array_equal(out[0][k], other[k]), k

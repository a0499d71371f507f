"""Host driver end to end on the device: .inp in the reference grammar -> mps_explicit -> final
displacement and CSV extractions, against the CPU oracle on the same problem."""
import csv
import os
import subprocess

import numpy as np
import pytest

import inpgen
import mps_gpu as G
import oracle_ctypes as O
from mps_gpu import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "v.1.1_amd", "lib", "mps_explicit")


def run(inp, outdir, *extra):
    dump = os.path.join(outdir, "u.txt")
    r = subprocess.run([EXE, "-input", str(inp), "-outputDir", str(outdir), "-dumpU", dump, *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return np.loadtxt(dump), r.stdout


def read_csv(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], np.array([[float(x) for x in r] for r in rows[1:]])


def g16(x):
    """Values as the CSV writer prints them (%.16g, dataExtraction.cpp:289-300)."""
    return np.array([float(f"{v:.16g}") for v in np.ravel(x)])


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_reference_ramp_time_window_with_extractions(tmp_path):
    m = configs.perturb(configs.box(4, 4, 4), 0.1, 11)
    N = m["n_nodes"]
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.002)  # top: total ramp displacement 2 mm
    bc = inpgen.bc_triples(dof, val, N)
    probe = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=1.0, scale=0.5)
    tf = 60.5 * probe.critical_dt()  # about 60 steps of the time-window loop
    top = [int(n) for n in np.arange(N)[-5:]]
    bottom = list(range(5))
    inp = tmp_path / "cube.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=tf, outputs=3, bc=bc)],
                     node_extractions=[("Top", "Unknown", "Rough", 2, top),
                                       ("Bot", "InternalForce", "Sum", 2, bottom)],
                     element_extractions=[("E", "DEFO_ENERGY", "Sum"), ("K", "KIN_ENERGY", "Sum"),
                                          ("V", "VOLUME", "Sum")])
    u, log = run(inp, tmp_path)
    assert "Solver used is EXPLICIT" in log and "done solving" in log
    o = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=tf, scale=0.5,
                        energy_every=1)
    r = o.run(hist_cap=4096)
    steps = r["steps"]
    assert steps > 50
    assert rel(u, r["u"]) <= 1e-12
    # energies and volume: one row per step at the oracle's times.  dt is computed in strict FP64 (bitwise the
    # oracle's for the same u); u agrees with the oracle's to rounding, so a step's dt may differ in its last bit.
    h = r["hist"]
    for name, key in (("All_DEFO_ENERGY_Sum", "e_def"), ("All_KIN_ENERGY_Sum", "e_kin"), ("All_VOLUME_Sum", "volume")):
        head, rows = read_csv(tmp_path / f"{name}.csv")
        assert head == ["Time", "Value"] and rows.shape == (steps, 2)
        np.testing.assert_allclose(rows[:, 0], g16(h["time"][:steps]), rtol=1e-14, atol=0)
        assert np.allclose(rows[:, 1], h[key][:steps], rtol=1e-11, atol=1e-12 * np.abs(h[key][:steps]).max())
    head, rows = read_csv(tmp_path / "Top_Unknown2_Rough.csv")
    assert head == ["Time"] + [f"Node{n + 1}Comp2" for n in top] and rows.shape == (steps, 6)
    assert np.array_equal(rows[-1, 1:], g16(u[np.array(top) + 2 * N]))
    assert np.abs(rows[-1, 1:] - r["u"][np.array(top) + 2 * N]).max() <= 1e-12 * np.abs(r["u"]).max()
    _, rows = read_csv(tmp_path / "Bot_InternalForce2_Sum.csv")
    f = o.internal_force(r["u"])
    ref = sum(f[n + 2 * N] for n in bottom)
    assert abs(rows[-1, 1] - ref) <= 1e-10 * np.abs(f).max()


def test_chained_solvers_hand_over_state(tmp_path):
    """Two *SOLVER, EXPLICIT blocks: the second starts at the first one's tf from its u, v, a
    (initValues / getFinalValues); its first dt uses u = 0 (t0 > 0), as the reference does."""
    m = configs.box(3, 3, 4)
    N = m["n_nodes"]
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.001)
    bc = inpgen.bc_triples(dof, val, N)
    dt0 = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, tf=1.0, scale=0.5).critical_dt()
    t1, t2 = 20.5 * dt0, 45.5 * dt0
    inp = tmp_path / "chain.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=t1, outputs=1, bc=bc),
                      dict(scale=0.4, t0=t1, tf=t2, outputs=1, bc=bc)])
    u, _ = run(inp, tmp_path)
    a = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=t1, scale=0.5).run()
    b = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=t1, tf=t2, scale=0.4).run(
        u0=a["u"], v0=a["v"], a0=a["a"])
    assert rel(u, b["u"]) <= 1e-12


def test_fixed_steps_velocity_fp32_and_materials(tmp_path):
    """Extensions: fixed step count, velocity BC ramp, FP32 element form; two materials by element list."""
    m = configs.perturb(configs.box(4, 3, 3), 0.1, 5)
    N = m["n_nodes"]
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.05)
    bc = inpgen.bc_triples(dof, val, N)
    hard = list(range(0, m["n_elem"], 3))
    inp = tmp_path / "ext.inp"
    inpgen.write_inp(inp, m, [dict(name="A", rho=1000.0, E=1e6, nu=0.3),
                              dict(name="B", rho=1100.0, E=4e6, nu=0.35, elems=hard)],
                     [dict(scale=0.5, outputs=1, bc=bc, bc_type="VELOCITY",
                           extra_lines=["*MPS Fixed Steps", "80", "*MPS Velocity Ramp Steps", "20",
                                        "*MPS Precision", "FP32_ELEM"])])
    u32, _ = run(inp, tmp_path)
    u64, _ = run(inp, tmp_path, "-precision", "FP64")
    E = np.full(m["n_elem"], 1e6)
    nu = np.full(m["n_elem"], 0.3)
    rho = np.full(m["n_elem"], 1000.0)
    E[hard], nu[hard], rho[hard] = 4e6, 0.35, 1100.0
    r = O.OracleProblem(m, E=E, nu=nu, rho=rho, dir_dof=dof, dir_val=val, bc_mode=O.BC_VELOCITY,
                        vel_ramp_steps=20, n_steps=80, scale=0.5).run()
    assert r["steps"] == 80
    assert rel(u64, r["u"]) <= 1e-12
    assert rel(u32, r["u"]) <= 1e-6


def test_force_and_gravity_loads_from_input(tmp_path):
    """*BOUNDARY, TYPE=FORCE blocks and *GRAVITY read from the file drive the device loads in input order
    (two entries on one node with gravity between them); the ExternalForce and EXTERNAL_ENERGY extractions
    follow them (dataExtraction.cpp:812-823, 866-874)."""
    m = configs.perturb(configs.box(3, 3, 4), 0.1, 7)
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    bot = [n for n in range(N) if X[2, n] < 1e-12]
    top = [n for n in range(N) if abs(X[2, n] - X[2].max()) < 1e-12]
    bc = [(n, d, 0.0) for n in bot for d in range(3)]
    dof = np.array([n + N * d for n, d, _ in bc], np.int32)
    val = np.zeros(dof.size)
    dt0 = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, tf=1.0, scale=0.5).critical_dt()
    tf = 80.5 * dt0
    rng = np.random.default_rng(4)
    F1, F2 = rng.uniform(-1e3, 1e3, (4, 3)), rng.uniform(-1e3, 1e3, (4, 3))
    n1, n2 = top[:4], top[2:6]
    w1, w2 = (0.0, 0.6 * tf, 2.0), (0.2 * tf, tf, 1.0)
    lines = ["*BOUNDARY, TYPE=FORCE", f"{w1[0]:.17g}, {w1[1]:.17g}, {w1[2]:.17g}"]
    lines += [f"{n + 1}, {f[0]:.17g}, {f[1]:.17g}, {f[2]:.17g}" for n, f in zip(n1, F1)]
    lines += ["*GRAVITY", "9.81", "*BOUNDARY, TYPE=FORCE", f"{w2[0]:.17g}, {w2[1]:.17g}, {w2[2]:.17g}"]
    lines += [f"{n + 1}, {f[0]:.17g}, {f[1]:.17g}, {f[2]:.17g}" for n, f in zip(n2, F2)]
    inp = tmp_path / "loads.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=tf, outputs=2, bc=bc)],
                     node_extractions=[("Fz", "ExternalForce", "Sum", 2, n2), ("Fx", "ExternalForce", "Rough", 0, n1)],
                     element_extractions=[("W", "EXTERNAL_ENERGY", "Sum")], extra="\n".join(lines))
    u, _ = run(inp, tmp_path, "-noVtk")
    forces = (np.array(n1 + n2, np.int32), np.vstack([F1 * w1[2], F2 * w2[2]]),
              np.array([w1[0]] * 4 + [w2[0]] * 4), np.array([w1[1]] * 4 + [w2[1]] * 4))
    o = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=tf, scale=0.5, gravity=9.81,
                        forces=forces, gravity_pos=4)
    r = o.run(hist_cap=4096)
    steps = r["steps"]
    assert steps > 70 and rel(u, r["u"]) <= 1e-12 and np.abs(u).max() > 0
    times = r["hist"]["time"][:steps]
    fext = [o.external_force(t) for t in times]
    _, rows = read_csv(tmp_path / "Fz_ExternalForce2_Sum.csv")
    assert rows.shape == (steps, 2)
    np.testing.assert_allclose(rows[:, 0], g16(times), rtol=1e-14, atol=0)  # dt to its last bit, as above
    ref = np.array([sum(f[n + 2 * N] for n in n2) for f in fext])
    np.testing.assert_allclose(rows[:, 1], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    head, rows = read_csv(tmp_path / "Fx_ExternalForce0_Rough.csv")
    assert head == ["Time"] + [f"Node{n + 1}Comp0" for n in n1]
    np.testing.assert_allclose(rows[:, 1:], g16(np.array([[f[n] for n in n1] for f in fext])).reshape(steps, 4),
                               rtol=1e-15, atol=0)
    _, rows = read_csv(tmp_path / "All_EXTERNAL_ENERGY_Sum.csv")
    assert rows.shape == (steps, 2)
    assert rows[-1, 1] == pytest.approx(float(np.dot(r["u"], fext[-1])), rel=1e-10)


def test_pressure_loads_from_input(tmp_path):
    """*BOUNDARY, TYPE=PRESSURE RAMP / INST blocks ("t0, tf, P, S" and element ids) on the top face of a clamped
    block, with *GRAVITY between them, against the oracle's surface-force restatement."""
    m = configs.box(3, 3, 3)
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    conn = m["conn"].reshape(-1, 8)
    bot = [n for n in range(N) if X[2, n] < 1e-12]
    bc = [(n, d, 0.0) for n in bot for d in range(3)]
    dof = np.array([n + N * d for n, d, _ in bc], np.int32)
    val = np.zeros(dof.size)
    top = [e for e in range(conn.shape[0]) if np.all(X[2, conn[e, 4:]] > 1 - 1e-12)]  # face 2 = nodes 4,7,6,5
    side = [e for e in range(conn.shape[0]) if np.all(X[0, conn[e, [1, 5, 6, 2]]] > 1 - 1e-12)]  # face 4
    dt0 = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, tf=1.0, scale=0.5).critical_dt()
    tf = 60.5 * dt0
    lines = ["*BOUNDARY, TYPE=PRESSURE RAMP", f"0, {0.8 * tf:.17g}, 5e4, 2"] + [str(e + 1) for e in top]
    lines += ["*GRAVITY", "9.81", "*BOUNDARY, TYPE=PRESSURE INST", f"{0.2 * tf:.17g}, {tf:.17g}, -2e4, 4"]
    lines += [f"{e + 1}, ignored" for e in side]
    inp = tmp_path / "press.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=tf, outputs=1, bc=bc)],
                     extra="\n".join(lines))
    u, _ = run(inp, tmp_path, "-noVtk")
    loads = [("pressure_ramp", 5e4, 0.0, 0.8 * tf, top, [2] * len(top)), ("gravity", 9.81),
             ("pressure_inst", -2e4, 0.2 * tf, tf, side, [4] * len(side))]
    r = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=tf, scale=0.5,
                        loads=loads).run()
    assert r["rc"] == 0 and r["steps"] > 50
    assert rel(u, r["u"]) <= 1e-12 and np.abs(u).max() > 1e-6


def test_loads_with_displacement_extraction(tmp_path):
    """FORCE, PRESSURE and GRAVITY loads together with an Unknown (displacement) node extraction: the per-step
    displacement gather must leave the load tables alone (regression: growing the gather buffers once released
    them, so every later step read freed memory)."""
    m = configs.perturb(configs.box(3, 3, 4), 0.1, 17)
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    conn = m["conn"].reshape(-1, 8)
    bot = [n for n in range(N) if X[2, n] < 1e-12]
    top_nodes = [n for n in range(N) if abs(X[2, n] - X[2].max()) < 1e-12]
    bc = [(n, d, 0.0) for n in bot for d in range(3)]
    dof = np.array([n + N * d for n, d, _ in bc], np.int32)
    val = np.zeros(dof.size)
    top = [e for e in range(conn.shape[0]) if np.all(X[2, conn[e, 4:]] > X[2].max() - 1e-12)]
    dt0 = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, tf=1.0, scale=0.5).critical_dt()
    tf = 70.5 * dt0
    F = np.array([[300.0, -200.0, 500.0], [-100.0, 400.0, 250.0]])
    fn = top_nodes[:2]
    lines = ["*BOUNDARY, TYPE=FORCE", f"0, {0.7 * tf:.17g}, 1.0"]
    lines += [f"{n + 1}, {f[0]:.17g}, {f[1]:.17g}, {f[2]:.17g}" for n, f in zip(fn, F)]
    lines += ["*GRAVITY", "9.81", "*BOUNDARY, TYPE=PRESSURE RAMP", f"0, {0.8 * tf:.17g}, 3e4, 2"]
    lines += [str(e + 1) for e in top]
    inp = tmp_path / "loads_u.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=tf, outputs=1, bc=bc)],
                     node_extractions=[("U", "Unknown", "Rough", 2, top_nodes), ("V", "Unknown", "Max", 0, top_nodes[:3])],
                     extra="\n".join(lines))
    u, _ = run(inp, tmp_path, "-noVtk")
    loads = [("force", np.array(fn, np.int32), F, 0.0, 0.7 * tf), ("gravity", 9.81),
             ("pressure_ramp", 3e4, 0.0, 0.8 * tf, top, [2] * len(top))]
    r = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=tf, scale=0.5,
                        loads=loads).run()
    assert r["rc"] == 0 and r["steps"] > 60
    assert rel(u, r["u"]) <= 1e-12 and np.abs(u).max() > 1e-6
    head, rows = read_csv(tmp_path / "U_Unknown2_Rough.csv")
    assert head == ["Time"] + [f"Node{n + 1}Comp2" for n in top_nodes] and rows.shape[0] == r["steps"]
    np.testing.assert_allclose(rows[-1, 1:], r["u"][np.array(top_nodes) + 2 * N], rtol=0,
                               atol=1e-12 * np.abs(r["u"]).max())


def parse_vtk(path):
    """Sections of a legacy ASCII VTK file written by the series writer: name -> list of value rows."""
    lines = open(path).read().split("\n")
    head = lines[:6]
    sec, cur = {}, None
    for ln in lines[5:]:
        t = ln.strip()
        if not t or t == "LOOKUP_TABLE default":
            continue
        w = t.split()
        if w[0] in ("POINTS", "CELLS", "CELL_TYPES", "POINT_DATA", "CELL_DATA"):
            cur = w[0] if w[0] not in ("POINT_DATA", "CELL_DATA") else None
            if cur:
                sec[cur] = []
            continue
        if w[0] in ("VECTORS", "TENSORS", "SCALARS"):
            cur = w[1]
            sec[cur] = []
            continue
        sec[cur].append(w)
    return head, sec, lines


def test_vtk_series_matches_state_and_oracle(tmp_path):
    """VTK series of the reference writer (output.cpp:296-918): file numbering and .vtk.series across two
    chained solvers, header and sections, POINTS/CELLS, nodal fields equal to the final state printed with 6
    significant digits, element fields (GP-averaged Cauchy, Green-Lagrange strain, VMS, volumetric and
    equivalent strain, pressure) against the oracle on the same displacement."""
    m = configs.perturb(configs.box(3, 3, 4), 0.1, 13)
    N, Ne = m["n_nodes"], m["n_elem"]
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.002)
    bc = inpgen.bc_triples(dof, val, N)
    dt0 = O.OracleProblem(m, law=1, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, tf=1.0, scale=0.5).critical_dt()
    t1, t2 = 20.5 * dt0, 31.5 * dt0
    inp = tmp_path / "vtk.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3, law=1)],
                     [dict(scale=0.5, t0=0.0, tf=t1, outputs=3, bc=bc), dict(scale=0.5, t0=t1, tf=t2, outputs=2, bc=bc)])
    inp.write_text(inp.read_text().replace("*Outputs\n3\n", "*Outputs\n3\nAll, 2\n"))
    u, _ = run(inp, tmp_path)
    files = sorted(tmp_path.glob("MuPhiSimOutputs_0_.*.vtk"), key=lambda p: int(p.name.split(".")[1]))
    assert [p.name for p in files] == [f"MuPhiSimOutputs_0_.{k}.vtk" for k in range(len(files))]
    assert len(files) >= 5  # t0 + >= 3 for the first solver + >= 1 for the second
    series = (tmp_path / "allOutputs_0.vtk.series").read_text().split("\n")
    assert series[:3] == ["{", '"file-series-version" : "1.0",', '"files" : [']
    entries = [ln for ln in series if ln.startswith("{ \"name\"")]
    assert len(entries) == len(files)
    times = [float(e.split(":")[-1].rstrip("},")) for e in entries]
    assert times[0] == 0.0 and all(b > a for a, b in zip(times, times[1:]))
    assert abs(times[-1] - t2) <= 1e-5 * t2
    assert all(ln.endswith("},") for ln in entries[:-1]) and entries[-1].endswith("}")
    # initial file: zero fields
    head0, s0, _ = parse_vtk(files[0])
    assert all(float(x) == 0.0 for r in s0["Displacement"] for x in r)
    assert all(float(x) == 0.0 for r in s0["Cauchy"] for x in r)
    # last file = final state of the second solver
    head, s, lines = parse_vtk(files[-1])
    assert head == ["# vtk DataFile Version 1.0", "3D Unstructured Grid of C3D8 elements", "ASCII", "",
                    "DATASET UNSTRUCTURED_GRID", f"POINTS {N} float"]
    X = m["X"].reshape(3, N)
    assert s["POINTS"] == [[f"{X[d, n]:.6g}" for d in range(3)] for n in range(N)]
    conn = m["conn"].reshape(Ne, 8)
    assert s["CELLS"] == [["8"] + [str(c) for c in conn[e]] for e in range(Ne)]
    assert s["CELL_TYPES"] == [["12"]] * Ne
    U = u.reshape(3, N)
    assert s["Displacement"] == [[f"{U[d, n]:.6g}" for d in range(3)] for n in range(N)]
    assert lines[lines.index("VECTORS Displacement float ") + 1].endswith(" ")
    q = O.OracleProblem(m, law=1, n_steps=1).element_outputs(u)

    def close(rows, ref):
        got = np.array([[float(x) for x in r] for r in rows]).reshape(ref.shape)
        return np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max() + 1e-300

    assert close(s["Cauchy"], q[:, :9])
    assert close(s["E"], q[:, 9:18])
    assert close(s["VMS"], q[:, 18])
    assert close(s["VolumetricStrain"], q[:, 19])
    assert close(s["EquivalentStrain"], q[:, 20])
    assert close(s["Pressure"], (q[:, 0] + q[:, 4] + q[:, 8]) / 3)
    assert not any(k.startswith("IntVars") for k in s)  # the second *Outputs names no internal variables
    _, s1, _ = parse_vtk(files[1])  # the first solver's *Outputs asked for variables 1 and 2
    assert s1["IntVars1"] == [["0"]] * Ne and s1["IntVars2"] == [["0"]] * Ne


def test_vtk_on_a_mesh_formatted_in_parallel(tmp_path):
    """A mesh above the writer's one-thread limit (20,000 nodes or elements per section): the sections are
    formatted on several threads and the geometry once per run, from device snapshots taken while the solve
    goes on.  Every file has the same POINTS and CELLS text, the numbers of each section in order, and the last
    one holds the final state."""
    m = configs.perturb(configs.box(30, 28, 26), 0.1, 17)
    N, Ne = m["n_nodes"], m["n_elem"]
    assert N > 20000 and Ne > 20000
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.05)
    bc = inpgen.bc_triples(dof, val, N)
    dt0 = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, tf=1.0, scale=0.5).critical_dt()
    inp = tmp_path / "big.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=40.5 * dt0, outputs=6, bc=bc)])
    u, _ = run(inp, tmp_path)
    files = sorted(tmp_path.glob("MuPhiSimOutputs_0_.*.vtk"), key=lambda p: int(p.name.split(".")[1]))
    assert len(files) >= 4
    X = m["X"].reshape(3, N)
    conn = m["conn"].reshape(Ne, 8)
    points = [[f"{X[d, n]:.6g}" for d in range(3)] for n in range(N)]
    cells = [["8"] + [str(c) for c in conn[e]] for e in range(Ne)]
    for f in files:
        _, s, _ = parse_vtk(f)
        assert s["POINTS"] == points and s["CELLS"] == cells, f.name
        assert len(s["Displacement"]) == N and len(s["Cauchy"]) == 3 * Ne and len(s["Pressure"]) == Ne
    _, s, _ = parse_vtk(files[-1])
    U = u.reshape(3, N)
    assert s["Displacement"] == [[f"{U[d, n]:.6g}" for d in range(3)] for n in range(N)]
    vms = np.array([float(r[0]) for r in s["VMS"]])
    assert np.all(np.isfinite(vms)) and vms.max() > 0


def test_snapshots_taken_while_stepping():
    """mps_snapshot_*: a snapshot holds the state of the step it was begun at although the solve went on before
    it was read; two slots; a slot is begun again only after its release."""
    m = configs.perturb(configs.box(6, 5, 4), 0.1, 3)
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.05)
    with G.Solver(m, dir_dof=dof, dir_val=val, n_steps=200) as s:
        s.run(30)
        u1, v1, a1 = s.get_state()
        e1 = s.element_outputs()
        s.snapshot_begin(0)
        s.run(40)
        u2, v2, a2 = s.get_state()
        s.snapshot_begin(1, with_elements=False)
        with pytest.raises(G.MpsError):
            s.snapshot_begin(0)
        su, sv, sa, se = s.snapshot_wait(0)
        assert np.array_equal(su, u1) and np.array_equal(sv, v1) and np.array_equal(sa, a1)
        assert np.array_equal(se, e1)
        s.snapshot_release(0)
        tu, tv, ta, te = s.snapshot_wait(1)
        assert np.array_equal(tu, u2) and np.array_equal(tv, v2) and np.array_equal(ta, a2) and te is None
        assert not np.array_equal(u1, u2)
        s.snapshot_release(1)
        s.snapshot_begin(0)
        s.run(10)
        assert np.array_equal(s.snapshot_wait(0)[0], u2)
        s.snapshot_release(0)


def test_forces_output_against_the_oracle(tmp_path):
    """*Forces inside an EXPLICIT block (forceOutManagement, inputFile.cpp:4830-4887; printForcesOut,
    output.cpp:153-279): externalForces<t0>-<tf>.csv gets one row per listed (node, direction) at every step whose
    time reaches iterations*max(period, dt) + t0, with that step's internal force, std::to_string formatted."""
    m = configs.perturb(configs.box(3, 3, 4), 0.1, 29)
    N = m["n_nodes"]
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.02)
    bc = inpgen.bc_triples(dof, val, N)
    probe = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=1.0, scale=0.5)
    dt0 = probe.critical_dt()
    tf = 60.5 * dt0
    period = 7.3 * dt0
    pairs = [(N - 1, 2), (0, 0), (int(np.argmin(np.abs(m["X"][:N] - 0.5) + np.abs(m["X"][N:2 * N] - 0.5)
                                                  + np.abs(m["X"][2 * N:] - 0.5))), 1)]
    lines = ["*Forces", f"{1.0 / period:.17g}"] + [f"{n + 1},{d}" for n, d in pairs]
    inp = tmp_path / "forces.inp"
    inpgen.write_inp(inp, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=tf, outputs=1, bc=bc, extra_lines=lines)])
    u, _ = run(inp, tmp_path, "-noVtk")
    o = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=tf, scale=0.5)
    r = o.run(hist_cap=4096)
    assert r["rc"] == 0 and rel(u, r["u"]) <= 1e-12
    steps = r["steps"]
    times, dts = r["hist"]["time"][:steps], r["hist"]["dt"][:steps]
    due, it = [], 0
    for k in range(steps):  # the test of printForcesOut, step by step
        p = max(period, dts[k])
        if times[k] >= it * p + 0.0:
            due.append(k)
            it += 1
    assert len(due) >= 7
    path = tmp_path / f"externalForces{0.0:f}-{tf:f}.csv"
    rows = path.read_text().split("\n")
    assert rows[0] == "Node,Direction,Force,Time,"
    body = [x for x in rows[1:] if x]
    assert len(body) == len(due) * len(pairs)
    fscale = 0.0
    for j, k in enumerate(due):
        uk = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_RAMP, t0=0.0, tf=tf, scale=0.5,
                             n_steps=k + 1).run()["u"]
        f = o.internal_force(uk)
        fscale = max(fscale, np.abs(f).max())
        for q, (n, d) in enumerate(pairs):
            node, direction, force, time, tail = body[j * len(pairs) + q].split(",")
            assert (node, direction, tail) == (str(n + 1), str(d), "")
            assert time == f"{times[k]:f}"
            assert abs(float(force) - f[n + d * N]) <= 5e-7 + 1e-12 * abs(f[n + d * N]), (k, n, d)
    assert fscale > 1.0  # forces of order newtons: the six printed decimals carry the comparison


def test_internal_force_dofs_equal_the_full_gather():
    """mps_get_internal_force_dofs: the step's partial forces summed at the listed DOFs, bitwise the
    corresponding entries of mps_internal_force at the same displacement; before the first step, the force of
    the initial state."""
    m = configs.perturb(configs.box(5, 4, 4), 0.1, 33)
    N = m["n_nodes"]
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.05)
    idx = np.arange(0, 3 * N, 5, dtype=np.int32)
    import mps_gpu as G
    for prec in (G.PREC_FP64, G.PREC_FP32_ELEM):
        with G.Solver(m, dir_dof=dof, dir_val=val, n_steps=50, precision=prec) as s:
            assert np.array_equal(s.internal_force_dofs(idx), np.zeros(idx.size))  # u0 = 0
            rc, rep = s.run(37)
            assert rc == 0
            u = s.get_state()[0]
            full = s.internal_force(u)  # its own partial buffer: the step's stay valid
            got = s.internal_force_dofs(idx)
            assert np.array_equal(got, full[idx]), prec
            assert np.abs(got).max() > 0

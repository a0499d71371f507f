"""Host driver (mps_explicit) on the CPU: the .inp reader's view of a model (-check prints it as JSON,
no device call) and the reference's input errors, each of which the reader turns into a message."""
import json
import os
import subprocess

import numpy as np
import pytest

import inpgen
from mps_gpu import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "v.1.1_amd", "lib", "mps_explicit")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="mps_explicit not built (run __graft_entry__.build())")


def check(path):
    r = subprocess.run([EXE, "-check", "-input", str(path)], capture_output=True, text=True, timeout=60)
    return r.returncode, (json.loads(r.stdout) if r.returncode == 0 else None), r.stderr


def cube(n=3):
    m = configs.box(n, n, n)
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.01)
    return m, inpgen.bc_triples(dof, val, m["n_nodes"])


def test_reads_model_solvers_materials_extractions(tmp_path):
    m, bc = cube(3)
    p = tmp_path / "a.inp"
    half = list(range(0, m["n_elem"], 2))
    inpgen.write_inp(p, m,
                     [dict(name="Soft", rho=1000.0, E=1e6, nu=0.3),
                      dict(name="Hard", rho=1200.0, E=5e6, nu=0.25, elems=half)],
                     [dict(scale=0.5, t0=0.0, tf=0.01, outputs=4, bc=bc),
                      dict(scale=0.25, t0=0.01, tf=0.02, outputs=2, bc=bc, extra_lines=["*MPS Fixed Steps", "7"])],
                     node_extractions=[("U", "Unknown", "Rough", 2, [62, 63]), ("F", "InternalForce", "Sum", 2, [0])],
                     element_extractions=[("E", "DEFO_ENERGY", "Sum"), ("K", "KIN_ENERGY", "Sum")])
    rc, j, err = check(p)
    assert rc == 0, err
    assert j["nodes"] == 64 and j["elements"] == 27 and j["type"] == "C3D8"
    assert [x["name"] for x in j["materials"]] == ["Soft", "Hard"]
    assert j["materials"][1] == dict(name="Hard", law="3D-Compressible-Neo-Hookean", rho=1200.0, E=5e6, nu=0.25)
    s0, s1 = j["solvers"]
    assert (s0["scale"], s0["t0"], s0["tf"], s0["outputs"], s0["dirichlet"]) == (0.5, 0.0, 0.01, 4, len(bc))
    assert s0["bc_mode"] == 0 and s0["fixed_steps"] == 0
    assert (s1["scale"], s1["t0"], s1["fixed_steps"]) == (0.25, 0.01, 7)
    assert j["node_extractions"] == ["F_InternalForce2_Sum", "U_Unknown2_Rough"]
    assert j["element_extractions"] == ["All_DEFO_ENERGY_Sum", "All_KIN_ENERGY_Sum"]


def test_carry_over_between_solver_blocks(tmp_path):
    """Scale factor, time window, *Outputs and the instantaneous flag are function-scope variables of
    solverManagement (inputFile.cpp:2253-2830): a later block that does not set them inherits them."""
    m, bc = cube(2)
    p = tmp_path / "b.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.3, t0=0.0, tf=0.01, outputs=5, bc=bc, bc_type="DISPLACEMENT INST"),
                      dict(scale=0.3, t0=0.01, tf=0.02, outputs=None, bc=bc)])
    rc, j, err = check(p)
    assert rc == 0, err
    s0, s1 = j["solvers"]
    assert s0["bc_mode"] == 1 and s1["bc_mode"] == 1  # INST persists into the RAMP block
    assert s1["outputs"] == 5


def test_velocity_extension_and_precision(tmp_path):
    m, bc = cube(2)
    p = tmp_path / "c.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3, law=1)],
                     [dict(scale=0.5, outputs=1, bc=bc, bc_type="VELOCITY",
                           extra_lines=["*MPS Fixed Steps", "100", "*MPS Velocity Ramp Steps", "10",
                                        "*MPS Precision", "FP32_ELEM"])])
    rc, j, err = check(p)
    assert rc == 0, err
    s = j["solvers"][0]
    assert (s["bc_mode"], s["fixed_steps"], s["vel_ramp_steps"], s["precision"]) == (2, 100, 10, 1)
    assert j["materials"][0]["law"] == "Neo-Hookean"


@pytest.mark.parametrize("line,expect", [("All, 3", [0, 1, 2]), ("5", [5]), ("4, 1, 3", [0, 2]), (None, [])])
def test_outputs_internal_variable_line(tmp_path, line, expect):
    """outManagement (inputFile.cpp:4769-4829): "All, n" -> 0..n-1; a single k -> k as is; "max, i, j" ->
    i-1, j-1."""
    m, bc = cube(2)
    p = tmp_path / "ov.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=0.01, outputs=2, bc=bc)])
    if line is not None:
        p.write_text(p.read_text().replace("*Outputs\n2\n", f"*Outputs\n2\n{line}\n"))
    rc, j, err = check(p)
    assert rc == 0, err
    assert j["solvers"][0]["out_vars"] == expect


def test_reads_svk_and_linear_elastic(tmp_path):
    m, bc = cube(2)
    p = tmp_path / "laws.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc)])
    for law_line, name in (("*HyperElastic, St-Venant-Kirchhoff", "St-Venant-Kirchhoff"),
                           ("*Linear-Elastic", "Linear-Elastic")):
        p.write_text(p.read_text().replace("*HyperElastic, 3D-Compressible-Neo-Hookean", law_line)
                     .replace("*HyperElastic, St-Venant-Kirchhoff", law_line))
        rc, j, err = check(p)
        assert rc == 0, err
        assert j["materials"][0]["law"] == name


@pytest.mark.parametrize("case,needle", [
    ("bc_before_time", "*Time must precede *BOUNDARY"),
    ("no_outputs", "*Outputs is mandatory"),
    ("duplicate_dof", "duplicate Dirichlet DOF"),
    ("no_dirichlet", "at least one mechanical Dirichlet DOF"),
    ("implicit", "is not on this path"),
    ("stvk", "has no device descriptor"),
    ("neumann", "TRACTION needs a"),
    ("pressure_face", "face S outside 1..6"),
    ("force_window", "tf > t0"),
    ("force_entry", "node, Fx, Fy, Fz"),
    ("two_gravity", "one gravity load"),
    ("unassigned_element", "has no material"),
    ("bad_direction", "direction"),
    ("cps3", "element type"),
])
def test_input_errors(tmp_path, case, needle):
    m, bc = cube(2)
    mats = [dict(rho=1000.0, E=1e6, nu=0.3)]
    sol = dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc)
    extra = ""
    if case == "bc_before_time":
        sol = dict(scale=0.5, outputs=1, bc=bc, extra_lines=[])
        sol.pop("t0", None)
    elif case == "no_outputs":
        sol["outputs"] = None
    elif case == "duplicate_dof":
        sol["bc"] = bc + [bc[0]]
    elif case == "no_dirichlet":
        sol["bc"] = []
    elif case == "unassigned_element":
        mats = [dict(rho=1000.0, E=1e6, nu=0.3, elems=[0, 1])]
    elif case == "bad_direction":
        sol["bc"] = [(0, 3, 0.0)]
    elif case == "neumann":
        extra = "*BOUNDARY, TYPE=TRACTION INST\n0,2,1e3,2\n4"  # four fields: a pressure line
    elif case == "pressure_face":
        extra = "*BOUNDARY, TYPE=PRESSURE INST\n0, 1, 1e3, 7\n1"
    elif case == "force_window":
        extra = "*BOUNDARY, TYPE=FORCE\n0.5, 0.5, 1\n1, 1, 0, 0"
    elif case == "force_entry":
        extra = "*BOUNDARY, TYPE=FORCE\n0, 1, 1\n1, 1, 0"
    elif case == "two_gravity":
        extra = "*GRAVITY\n9.81\n*GRAVITY\n1.0"
    p = tmp_path / f"{case}.inp"
    inpgen.write_inp(p, m, mats, [sol], extra=extra)
    text = p.read_text()
    if case == "bc_before_time":
        text = text.replace("*END SOLVER", "*Time\n0, 0.01\n*END SOLVER")
    if case == "implicit":
        text = text.replace("*SOLVER, EXPLICIT", "*SOLVER, IMPLICIT STATIC")
    if case == "stvk":
        text = text.replace("3D-Compressible-Neo-Hookean", "St-Venant-Kirchhoff-ThermoMechanics")
    if case == "cps3":
        text = text.replace("type=C3D8", "type=CPS3")
    p.write_text(text)
    rc, _, err = check(p)
    assert rc != 0
    assert needle in err, err


def test_reads_force_and_gravity(tmp_path):
    """NeumannCommonManagement reads the whole file in order, FORCE blocks inside *SOLVER too (the loads are
    global); each entry is F * Ff over its block's window (forceBCManagement, inputFile.cpp:2841-2941), and
    *GRAVITY takes its place among them."""
    m, bc = cube(2)
    p = tmp_path / "loads.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc,
                           extra_lines=["*BOUNDARY, TYPE=FORCE", "0, 1, 1", "1, 0, 0, 1"])],
                     extra="\n".join(["*BOUNDARY, TYPE=FORCE", "0.0, 0.5, 2.0", "3, 1.0, -2.0, 0.5", "5, 0, 0, 4",
                                      "*GRAVITY", "9.81", "*BOUNDARY, TYPE=FORCE", "0.1, 0.3, 1.0", "3, 7, 0, 0"]))
    rc, j, err = check(p)
    assert rc == 0, err
    ld = j["loads"]
    assert [b["type"] for b in ld] == ["FORCE", "FORCE", "GRAVITY", "FORCE"]
    assert ld[0]["entries"] == [[0, 0, 0, 1]] and (ld[0]["t0"], ld[0]["tf"]) == (0, 1)
    assert ld[1]["entries"] == [[2, 2, -4, 1], [4, 0, 0, 8]] and (ld[1]["t0"], ld[1]["tf"]) == (0, 0.5)
    assert ld[2]["value"] == 9.81
    assert ld[3]["entries"] == [[2, 7, 0, 0]] and (ld[3]["t0"], ld[3]["tf"]) == (0.1, 0.3)
    assert j["solvers"][0]["dirichlet"] == len(bc)
    # pressure blocks: "t0, tf, P, S" then element ids (their leading integer)
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc)],
                     extra="\n".join(["*BOUNDARY, TYPE=PRESSURE RAMP", "0.0, 0.5, 3e3, 2", "5", "6, 1", "*GRAVITY", "2.0",
                                      "*BOUNDARY, TYPE=PRESSURE INST", "0.25, 0.25, -1e3, 6", "8"]))
    rc, j, err = check(p)
    assert rc == 0, err
    assert j["loads"] == [dict(type="PRESSURE RAMP", t0=0, tf=0.5, value=3e3, entries=[[4, 2], [5, 2]]),
                          dict(type="GRAVITY", t0=0, tf=0, value=2.0, entries=[]),
                          dict(type="PRESSURE INST", t0=0.25, tf=0.25, value=-1e3, entries=[[7, 6]])]
    # traction blocks: "t0, tf, Tx, Ty, Tz, S" then element ids
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc)],
                     extra="\n".join(["*BOUNDARY, TYPE=TRACTION RAMP", "0.0, 0.5, 1, -2, 3e2, 2", "5", "6"]))
    rc, j, err = check(p)
    assert rc == 0, err
    assert j["loads"] == [dict(type="TRACTION RAMP", t0=0, tf=0.5, value=0, traction=[1, -2, 300],
                               entries=[[4, 2], [5, 2]])]
    # no loads
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc)])
    rc, j, err = check(p)
    assert rc == 0 and j["loads"] == []


@pytest.mark.parametrize("lines,expect", [
    (["*Bulk viscosity", "0.3, 2.5"], (1, 0.3, 2.5)),
    (["*Bulk viscosity", "0.7"], (1, 0.06, 0.7)),         # one number closes cl (the last character)
    (["*Bulk viscosity"], (1, 0.06, 1.2)),                # defaults; the next keyword is still read
    (["*Dev viscosity", "2, 0.25"], (2, 2.0, 0.25)),
    ([], (0, 0.0, 0.0)),
])
def test_viscosity_keywords(tmp_path, lines, expect):
    """solverManagement (inputFile.cpp:2462-2510): *Bulk viscosity / *Dev viscosity with an optional "c1, cl" line,
    per solver (reset at every *SOLVER, :2316-2318)."""
    m, bc = cube(2)
    p = tmp_path / "v.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc, pre_lines=lines),
                      dict(scale=0.5, t0=0.01, tf=0.02, outputs=1, bc=bc)])
    rc, j, err = check(p)
    assert rc == 0, err
    s0, s1 = j["solvers"]
    assert (s0["visc"], s0["visc_c1"], s0["visc_cl"]) == expect
    assert s1["visc"] == 0
    assert (s0["outputs"], s0["t0"]) == (1, 0.0)


def test_coordinates_round_trip(tmp_path):
    """The reader gets every coordinate bit for bit (written with %.17g) and every connectivity entry."""
    m = configs.perturb(configs.box(6, 5, 4), 0.1, 7)
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.01)
    p = tmp_path / "d.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=inpgen.bc_triples(dof, val, m["n_nodes"]))])
    rc, j, err = check(p)
    assert rc == 0, err
    assert j["nodes"] == m["n_nodes"] and j["elements"] == m["n_elem"]
    xor = int(np.bitwise_xor.reduce(np.ascontiguousarray(m["X"]).view(np.uint64)))
    assert int(j["x_xor"], 16) == xor
    assert j["conn_sum"] == int(m["conn"].astype(np.int64).sum())


def test_forces_block(tmp_path):
    """*Forces: frequency line, then "node, direction" lines (direction 0-based, < ndim); a later *Forces of the
    same block replaces the earlier one; the block belongs to its solver only."""
    m, bc = cube(2)
    p = tmp_path / "f.inp"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                     [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc,
                           extra_lines=["*Forces", "4", "27,0", "*Forces", "20", "1, 2", "5,1"]),
                      dict(scale=0.5, t0=0.01, tf=0.02, outputs=1, bc=bc)])
    rc, j, err = check(p)
    assert rc == 0, err
    assert j["solvers"][0]["forces"] == dict(period=0.05, entries=[[0, 2], [4, 1]])
    assert j["solvers"][1]["forces"] is None
    for bad, msg in ((["*Forces", "10", "1,3"], "exceeds the simulation dimension"),
                     (["*Forces", "10", "28,0"], "outside the mesh"),
                     (["*Forces", "0", "1,0"], "positive output frequency")):
        inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)],
                         [dict(scale=0.5, t0=0.0, tf=0.01, outputs=1, bc=bc, extra_lines=bad)])
        rc, _, err = check(p)
        assert rc != 0 and msg in err, err

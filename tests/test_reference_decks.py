"""The reader on the reference's own input decks, read in place (nothing from them is kept in this repository;
the tests skip where the reference tree is absent, as on a machine that has only this repository):

- applications/examples/3Dcube.inp: the whole deck is refused with a message naming its IMPLICIT STATIC solver;
  with that block removed, its EXPLICIT block, PRESSURE INST load and St-Venant-Kirchhoff material read as
  written;
- the EXPLICIT blocks of applications/examples/example2.inp, example7.inp and example8.inp (whose meshes are
  2-D or whose materials are off this path) spliced after a 3-D mesh: scale factors, *Outputs with their
  internal-variable lines, time windows, Dirichlet lists, *Bulk viscosity (reset per EXPLICIT block,
  inputFile.cpp:2312-2318), global PRESSURE INST blocks and the *Forces block;
- the *EXTRACTION lines of applications/tools/python-gmsh/examples/Flexion/input.inp (15 node and 3 element
  extractions inside its solver block), with its solver read as EXPLICIT.
"""
import json
import os
import re
import subprocess

import pytest

import inpgen
from mps_gpu import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "v.1.1_amd", "lib", "mps_explicit")
REF = os.environ.get("MPS_REFERENCE_DIR", "/root/reference")
EXAMPLES = os.path.join(REF, "applications", "examples")
FLEXION = os.path.join(REF, "applications", "tools", "python-gmsh", "examples", "Flexion", "input.inp")

pytestmark = [pytest.mark.skipif(not os.path.exists(EXE), reason="mps_explicit not built"),
              pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="reference decks not present")]


def check(path):
    r = subprocess.run([EXE, "-check", "-input", str(path)], capture_output=True, text=True, timeout=60)
    return r.returncode, (json.loads(r.stdout) if r.returncode == 0 else None), r.stderr


def lines_of(name):
    with open(os.path.join(EXAMPLES, name)) as f:
        return f.read().split("\n")


def blocks(lines, head, end="*END SOLVER"):
    """Every block of `lines` from a line starting with `head` to the line starting with `end` (inclusive)."""
    out, cur = [], None
    for ln in lines:
        if cur is None and ln.startswith(head):
            cur = [ln]
        elif cur is not None:
            cur.append(ln)
            if ln.startswith(end):
                out.append(cur)
                cur = None
    return out


def keyword_blocks(lines, head):
    """Blocks that start with `head` and run to the next keyword line."""
    out = []
    for i, ln in enumerate(lines):
        if ln.startswith(head):
            j = i + 1
            while j < len(lines) and not lines[j].startswith("*"):
                j += 1
            out.append(lines[i:j])
    return out


def test_3dcube_full_deck_names_its_implicit_solver(tmp_path):
    rc, _, err = check(os.path.join(EXAMPLES, "3Dcube.inp"))
    assert rc != 0 and "IMPLICIT STATIC" in err, err


def test_3dcube_explicit_block(tmp_path):
    lines = lines_of("3Dcube.inp")
    implicit = blocks(lines, "*SOLVER, IMPLICIT STATIC")
    assert len(implicit) == 1
    start = lines.index(implicit[0][0])
    kept = lines[:start] + lines[start + len(implicit[0]):]
    p = tmp_path / "cube_explicit.inp"
    p.write_text("\n".join(kept))
    rc, j, err = check(p)
    assert rc == 0, err
    assert (j["nodes"], j["elements"], j["type"]) == (9, 12, "C3D4")
    assert j["materials"] == [dict(name="Region1", law="St-Venant-Kirchhoff", rho=1000.0, E=1e9, nu=0.3)]
    (s,) = j["solvers"]
    assert (s["scale"], s["outputs"], s["t0"], s["tf"], s["dirichlet"], s["bc_mode"]) == (0.1, 100, 1.0, 2.0, 9, 0)
    assert s["forces"] is None and s["visc"] == 0
    (ld,) = j["loads"]
    assert (ld["type"], ld["t0"], ld["tf"], ld["value"]) == ("PRESSURE INST", 0.0, 2.0, 1e8)
    assert ld["entries"] == [[3, 2], [11, 2]]  # elements 4 and 12 (0-based), face 2


def splice(tmp_path, name, mesh_n=7):
    """A 3-D hex mesh large enough for every node and element id of the deck's solver and load blocks, then the
    deck's EXPLICIT blocks and global PRESSURE blocks as they are written, then one material."""
    lines = lines_of(name)
    sol = blocks(lines, "*SOLVER, EXPLICIT")
    inside = set()
    for b in blocks(lines, "*SOLVER"):
        inside.update(id(x) for x in b)
    press = [b for b in keyword_blocks(lines, "*BOUNDARY, TYPE=PRESSURE")]
    m = configs.box(mesh_n, mesh_n, mesh_n)
    p = tmp_path / f"spliced_{name}"
    inpgen.write_inp(p, m, [dict(rho=1000.0, E=1e6, nu=0.3)], [])
    text = p.read_text().split("\n")
    at = text.index("*MATERIAL, name=Mat1")
    body = [ln for b in sol for ln in b] + [ln for b in press for ln in b]
    p.write_text("\n".join(text[:at] + body + text[at:]))
    return sol, check(p)


def test_example2_explicit_block(tmp_path):
    sol, (rc, j, err) = splice(tmp_path, "example2.inp")
    assert rc == 0, err
    (s,) = j["solvers"]
    n_bc = len(sol[0]) - sol[0].index("*BOUNDARY, TYPE=DISPLACEMENT RAMP") - 2
    assert (s["scale"], s["outputs"], s["t0"], s["tf"], s["dirichlet"]) == (0.1, 500, 0.0, 1.0, n_bc)
    assert s["out_vars"] == list(range(22))  # "All, 22"


def test_example7_three_explicit_blocks(tmp_path):
    sol, (rc, j, err) = splice(tmp_path, "example7.inp")
    assert rc == 0, err
    s = j["solvers"]
    assert len(s) == 3
    assert [(x["t0"], x["tf"]) for x in s] == [(0.0, 2.5), (2.5, 5.0), (5.0, 7.5)]
    assert [(x["visc"], x["visc_c1"], x["visc_cl"]) for x in s] == [(1, 2.0, 3.0), (0, 0.0, 0.0), (1, 3.0, 3.0)]
    # "All, 0": atoi of the text from column 4 is 0, so no internal variable is printed (outManagement)
    assert all(x["scale"] == 0.1 and x["outputs"] == 5000 and x["out_vars"] == [] for x in s)
    assert [x["dirichlet"] for x in s] == [len(b) - b.index("*BOUNDARY, TYPE=DISPLACEMENT RAMP") - 2 for b in sol]
    loads = j["loads"]
    assert [(ld["type"], ld["t0"], ld["tf"], ld["value"]) for ld in loads] == \
        [("PRESSURE INST", 0.0, 1.25, -5e6), ("PRESSURE INST", 2.5, 3.75, -5e6), ("PRESSURE INST", 5.0, 6.25, -5e6)]
    assert all(ld["entries"] == [[e - 1, 2] for e in (195, 167, 139, 111, 83, 55, 27)] for ld in loads)


def test_example8_explicit_block_with_forces(tmp_path):
    sol, (rc, j, err) = splice(tmp_path, "example8.inp")
    assert rc == 0, err
    (s,) = j["solvers"]
    assert (s["scale"], s["outputs"], s["t0"], s["tf"], s["dirichlet"]) == (0.1, 1000, 1.0, 4.0, 6)
    assert s["forces"] == dict(period=0.1, entries=[[0, 1]])  # frequency 10; node 1, direction 1


@pytest.mark.skipif(not os.path.exists(FLEXION), reason="Flexion deck not present")
def test_flexion_extraction_lines(tmp_path):
    with open(FLEXION) as f:
        text = f.read()
    assert "*SOLVER, IMPLICIT\n" in text
    p = tmp_path / "flexion_explicit.inp"
    p.write_text(text.replace("*SOLVER, IMPLICIT\n", "*SOLVER, EXPLICIT\n", 1))
    rc, j, err = check(p)
    assert rc == 0, err
    lines = text.split("\n")
    node_names, elem_names = set(), set()
    for i, ln in enumerate(lines):
        if ln.startswith("*EXTRACTION, NODE"):
            name, typ, op, comp = lines[i + 1].split(",")
            node_names.add(f"{name}_{typ}{int(comp)}_{op}")
        elif ln.startswith("*EXTRACTION, ELEMENT"):
            name, typ, op = lines[i + 1].split(",")
            elem_names.add(("All" if lines[i + 2].startswith("All") else name) + f"_{typ}_{op}")
    assert len(node_names) == 15 and len(elem_names) == 3
    assert set(j["node_extractions"]) == node_names
    assert set(j["element_extractions"]) == elem_names
    assert j["solvers"][0]["out_vars"] == []  # "All,max": atoi("max") = 0 variables
    assert j["type"] == "C3D4" and j["materials"][0]["law"] == "Linear-Elastic"
    assert [ld["type"] for ld in j["loads"]] == ["PRESSURE RAMP", "PRESSURE RAMP"]
    assert re.search(r"\d+", str(j["warnings"]))  # *Number Of Steps is read and reported as ignored

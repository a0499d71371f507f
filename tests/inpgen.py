"""Writes .inp files in the reference grammar (node and element ids 1-based, Dirichlet direction 0-based,
keywords as read by inputFile.cpp) for the host-driver tests."""
import numpy as np


def write_inp(path, mesh, materials, solvers, node_extractions=(), element_extractions=(), extra=""):
    """materials: list of dict(name, law, rho, E, nu, elems=None|list of 0-based ids).
    solvers: list of dict(scale, t0, tf, outputs, bc=list of (node0, dir, value), bc_type, extra_lines).
    node_extractions: list of (name, type, op, comp, nodes0); element_extractions: (name, type, op)."""
    N, Ne, nen = mesh["n_nodes"], mesh["n_elem"], mesh["nen"]
    X = mesh["X"].reshape(3, N)
    conn = mesh["conn"].reshape(Ne, nen)
    out = ["*Node"]
    out += [f"{i + 1}, {X[0, i]:.17g}, {X[1, i]:.17g}, {X[2, i]:.17g}" for i in range(N)]
    out.append("*Element, type=" + ("C3D8" if nen == 8 else "C3D4"))
    out += [f"{e + 1}, " + ", ".join(str(int(c) + 1) for c in conn[e]) for e in range(Ne)]
    out += ["*FEM Nodes", "All", "*FEM Elements", "All"]
    for s in solvers:
        out += ["*SOLVER, EXPLICIT", "*Scale Factor", f"{s.get('scale', 0.5):.17g}"]
        out += s.get("pre_lines", [])
        if s.get("outputs", 1) is not None:
            out += ["*Outputs", str(s.get("outputs", 1))]
        if "t0" in s:
            out += ["*Time", f"{s['t0']:.17g}, {s['tf']:.17g}"]
        out += s.get("extra_lines", [])
        out.append("*BOUNDARY, TYPE=" + s.get("bc_type", "DISPLACEMENT RAMP"))
        out += [f"{n + 1},{d},{v:.17g}" for n, d, v in s["bc"]]
        out.append("*END SOLVER")
    for name, typ, op, comp, nodes in node_extractions:
        out += ["*EXTRACTION, NODE", f"{name},{typ},{op},{comp}"] + [str(n + 1) for n in nodes]
    for name, typ, op in element_extractions:
        out += ["*EXTRACTION, ELEMENT", f"{name},{typ},{op}", "All"]
    if extra:
        out += extra.strip("\n").split("\n")
    for m in materials:
        law = "3D-Compressible-Neo-Hookean" if m.get("law", 0) == 0 else "Neo-Hookean"
        out += [f"*MATERIAL, name={m.get('name', 'Mat1')}", "*Density", f"{m['rho']:.17g}",
                f"*HyperElastic, {law}", f"{m['E']:.17g}, {m['nu']:.17g}", "*List of Elements"]
        if m.get("elems") is None:
            out.append("All")
        else:
            out += [str(e + 1) for e in m["elems"]]
    out.append("*END")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def bc_triples(dir_dof, dir_val, n_nodes):
    dof = np.asarray(dir_dof)
    return [(int(d % n_nodes), int(d // n_nodes), float(v)) for d, v in zip(dof, dir_val)]

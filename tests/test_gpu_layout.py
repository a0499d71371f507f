"""Element storage order.  The library stores hex8 elements sorted along a space-filling curve of their
centroids (compact patches: fewer partial forces for the element kernel to write and the node kernel to
read) and maps element ids at its boundary: per-element materials, face loads, element outputs; tet4
elements stay in input order.  Checked here: the order is a permutation, the patch tables follow it (the
partial count is the distinct nodes of each run of 256 stored elements), and a hex solve in curve order
matches the solve in input order (MPS_ELEM_ORDER=input) to summation-order rounding, with the lumped mass
bitwise the same (its CSR adds the element terms in input order either way).  The other hex GPU tests run
in curve order against the oracle."""
import numpy as np
import pytest

import mps_gpu as G
from meshgen import box, clamp_and_pull, perturb
from test_gpu_loads import boundary_faces

pytestmark = pytest.mark.gpu

PATCH = 256


def partials(conn):
    return int(sum(np.unique(conn[s:s + PATCH]).size for s in range(0, conn.shape[0], PATCH)))


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("tet", [False, True])
def test_element_order_is_a_compact_permutation(tet, monkeypatch):
    m = perturb(box(32, 16, 16, tet=tet), 0.1, 11)
    ne, nen = m["n_elem"], m["nen"]
    conn = m["conn"].reshape(ne, nen)
    monkeypatch.delenv("MPS_ELEM_ORDER", raising=False)
    with G.Solver(m, n_steps=1) as s:
        lay = s.layout()
    order = lay["order"]
    assert np.array_equal(np.sort(order), np.arange(ne))
    assert lay["n_patch"] == -(-ne // PATCH)
    assert lay["n_part"] == partials(conn[order])
    monkeypatch.setenv("MPS_ELEM_ORDER", "input")
    with G.Solver(m, n_steps=1) as s:
        lin = s.layout()
    assert np.array_equal(lin["order"], np.arange(ne))
    assert lin["n_part"] == partials(conn)
    print(f"tet={tet}: partials per element {lin['n_part'] / ne:.3f} (input) -> {lay['n_part'] / ne:.3f} (stored)")
    if tet:
        assert np.array_equal(order, np.arange(ne))
    else:
        assert lay["n_part"] < 0.8 * lin["n_part"]


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_FP32_ELEM])
def test_curve_order_matches_input_order(precision, monkeypatch):
    """Hex8: per-element E and nu, a follower pressure and a traction on the top face, 120 steps: displacement,
    history, element outputs and the mass of the two storage orders."""
    m = perturb(box(12, 10, 8, lz=0.8), 0.08, 29)
    ne = m["n_elem"]
    rng = np.random.default_rng(3)
    E = 10 ** rng.uniform(5, 7, ne)
    nu = rng.uniform(0.2, 0.45, ne)
    top = boundary_faces(box(12, 10, 8, lz=0.8), 2, 0.8)  # same ids and faces on the perturbed mesh
    blocks = [("pressure_ramp", 3e4, 0.0, 0.4, top[0][::2], top[1][::2]),
              ("traction_inst", (2e3, 0.0, -1e3), 0.1, 0.9, top[0][1::2], top[1][1::2])]
    dd, dv = clamp_and_pull(m)
    steps = 120
    out = {}
    for mode in ("curve", "input"):
        if mode == "input":
            monkeypatch.setenv("MPS_ELEM_ORDER", "input")
        else:
            monkeypatch.delenv("MPS_ELEM_ORDER", raising=False)
        with G.Solver(m, E=E, nu=nu, dir_dof=dd, dir_val=dv, n_steps=steps, precision=precision,
                      energy_every=1, history_cap=steps) as s:
            s.set_loads(loads=blocks)
            rc, rep = s.run(steps)
            assert rc == 0 and rep["steps"] == steps
            hist = s.history(0, steps)
            hist.update(s.energy_balance(0, steps))
            out[mode] = dict(u=s.get_state()[0], mass=s.mass(), eo=s.element_outputs(), hist=hist,
                             order=s.layout()["order"])
    a, b = out["curve"], out["input"]
    assert not np.array_equal(a["order"], np.arange(ne))
    assert np.array_equal(a["mass"], b["mass"])
    tol = 1e-12 if precision == G.PREC_FP64 else 2e-6  # FP32 form: each order within 1e-6 of FP64
    err = relerr(a["u"], b["u"])
    print(f"precision={precision}: max rel |u_curve - u_input| = {err:.2e}")
    assert err <= tol
    for j in range(a["eo"].shape[1]):  # each output field on its own scale
        assert relerr(a["eo"][:, j], b["eo"][:, j]) <= (1e-9 if precision == G.PREC_FP64 else 1e-4), j
    for k in ("time", "dt"):
        np.testing.assert_allclose(a["hist"][k], b["hist"][k], rtol=1e-13 if precision == G.PREC_FP64 else 1e-6)
    for k in ("e_def", "e_kin", "volume", "e_ext"):
        np.testing.assert_allclose(a["hist"][k], b["hist"][k], rtol=1e-10 if precision == G.PREC_FP64 else 1e-5,
                                   atol=1e-14 * np.abs(b["hist"][k]).max(), err_msg=k)


@pytest.mark.parametrize("case", ["flat", "one", "two", "coincident"])
def test_element_order_edge_meshes(case, monkeypatch):
    """Meshes whose centroids do not fill a box: one element thick (the key's thin axis), one and two
    elements, and every element duplicated in place (pairs of equal keys, kept in input order).  The order
    is a permutation, ties keep input order, and a short solve matches the input-order solve."""
    if case == "flat":
        m = perturb(box(40, 30, 1, lz=1.0 / 30), 0.05, 5)
    elif case == "one":
        m = box(1, 1, 1)
    elif case == "two":
        m = box(2, 1, 1)
    else:
        b = box(8, 6, 4)
        conn = b["conn"].reshape(-1, 8)
        m = dict(b, conn=np.ascontiguousarray(np.repeat(conn, 2, axis=0).ravel()), n_elem=2 * b["n_elem"])
    ne = m["n_elem"]
    dd, dv = clamp_and_pull(m)
    out = {}
    for mode in ("curve", "input"):
        if mode == "input":
            monkeypatch.setenv("MPS_ELEM_ORDER", "input")
        else:
            monkeypatch.delenv("MPS_ELEM_ORDER", raising=False)
        with G.Solver(m, dir_dof=dd, dir_val=dv, n_steps=40) as s:
            order = s.layout()["order"]
            rc, rep = s.run(40)
            assert rc == 0 and rep["steps"] == 40
            out[mode] = (order, s.get_state()[0], s.element_outputs())
    order = out["curve"][0]
    assert np.array_equal(np.sort(order), np.arange(ne))
    if case == "coincident":  # each duplicate right after its original
        assert np.all(order[1::2] == order[0::2] + 1) and np.all(order[0::2] % 2 == 0)
    assert relerr(out["curve"][1], out["input"][1]) <= 1e-12
    np.testing.assert_allclose(out["curve"][2], out["input"][2], rtol=0, atol=1e-9 * np.abs(out["input"][2]).max())

"""CPU-side checks of the product library and its host logic (no GPU calls)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mps_gpu as G
from mps_gpu import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = G.load()
    declared = G.declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert missing == []


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["readelf", "-d", G.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", G.LIB_PATH], capture_output=True, text=True).stdout
    assert " orc_" not in syms


def test_c_header_compiles_as_plain_c():
    src = os.path.join(ROOT, "include", "mps_gpu.h")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_create_rejects_invalid_input_without_a_gpu():
    # argument validation happens before any device call
    m = configs.box(2, 2, 2)
    bad = dict(m)
    bad["conn"] = m["conn"].copy()
    bad["conn"][5] = 10 ** 6
    with pytest.raises(G.MpsError) as ei:
        G.Solver(bad, n_steps=10)
    assert ei.value.code == G.MPS_ERR_ARG and "outside" in str(ei.value)
    dof = np.array([3, 3], np.int32)
    with pytest.raises(G.MpsError) as ei:
        G.Solver(m, dir_dof=dof, dir_val=np.zeros(2), n_steps=10)
    assert "duplicate" in str(ei.value)
    with pytest.raises(G.MpsError) as ei:
        G.Solver(m, nu=0.5, n_steps=10)
    assert "material" in str(ei.value)
    with pytest.raises(G.MpsError):
        G.Solver(m, n_steps=0, t0=1.0, tf=1.0)


def test_splitmix64_reference_values():
    # SplitMix64 (Steele, Lea, Flood 2014) stream from seed 0: first outputs are published constants
    z = configs.splitmix64(0, np.arange(3))
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_box_generator_orientation_and_counts():
    for tet in (False, True):
        m = configs.box(3, 4, 5, tet=tet)
        nen = 4 if tet else 8
        assert m["n_nodes"] == 4 * 5 * 6 and m["n_elem"] == 60 * (6 if tet else 1)
        X = m["X"].reshape(3, -1)
        conn = m["conn"].reshape(-1, nen)
        vol = 0.0
        for e in conn:
            p = X[:, e].T
            if tet:
                v = abs(np.linalg.det(np.array([p[1] - p[0], p[2] - p[0], p[3] - p[0]]))) / 6
            else:
                v = np.linalg.det(np.array([p[1] - p[0], p[3] - p[0], p[4] - p[0]]))
                assert v > 0  # reference node order: right-handed
            vol += v
        assert vol == pytest.approx(1.0, rel=1e-12)


def test_perturbation_is_deterministic_and_interior_only():
    m = configs.box(6, 6, 6)
    a = configs.perturb(m, 0.1, configs.SEED_CFG1)
    b = configs.perturb(m, 0.1, configs.SEED_CFG1)
    assert np.array_equal(a["X"], b["X"])
    d = (a["X"] - m["X"]).reshape(3, -1)
    X = m["X"].reshape(3, -1)
    boundary = np.any((X < 1e-12) | (X > 1 - 1e-12), axis=0)
    assert np.all(d[:, boundary] == 0) and np.all(np.abs(d) <= 0.1 / 6 + 1e-15)
    assert np.abs(d[:, ~boundary]).max() > 0.05 / 6


def test_batch_instances_ranges():
    inst = configs.batch_instances(512)
    n = np.array([i["n"] for i in inst])
    assert n.min() >= 6 and n.max() <= 40 and len(set(n)) > 20
    E = np.array([i["E"] for i in inst])
    assert E.min() >= 1e5 and E.max() <= 1e7

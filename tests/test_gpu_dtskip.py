"""Stable-dt skip of the FP64 hex8 and tet4 element kernels.  An element whose dt is certainly above the previous step's
min by more than a margin skips the strict characteristic length; the kernel's last workgroup certifies that
the step's min over the exactly computed elements is at most that threshold, and otherwise adds every element's
exact dt itself.  The min, and so every step, is the same bit for bit with the skip off (MPS_DT_SKIP=0), on
(default margin 1/64), and with a negative margin (MPS_DT_SKIP=-0.5: every element skips, each certificate
fails and the fallback runs, then skipping pauses for 64 steps and fails again)."""
import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from meshgen import box, clamp_and_pull, perturb

pytestmark = pytest.mark.gpu


def run(m, steps, E, nu, dd, dv, energy_every, chunks):
    with G.Solver(m, E=E, nu=nu, dir_dof=dd, dir_val=dv, n_steps=steps, energy_every=energy_every,
                  history_cap=steps) as s:
        done = 0
        for c in chunks:
            rc, rep = s.run(c)
            assert rc == 0
            done += c
        assert rep["steps"] == steps
        u, v, a = s.get_state()
        return u, v, a, s.history(0, steps)


@pytest.mark.parametrize("energy_every", [0, 1])
@pytest.mark.parametrize("tet", [False, True])
def test_dt_skip_bitwise(tet, energy_every, monkeypatch):
    """Hex8 (FP64 element kernel) and tet4 (node-staging FP64 kernel), per-element materials."""
    m = perturb(box(20, 14, 10, lx=1.0, ly=0.7, lz=0.5, tet=tet), 0.12, 31)
    ne = m["n_elem"]
    rng = np.random.default_rng(9)
    E = 10 ** rng.uniform(5.5, 6.5, ne)
    nu = rng.uniform(0.25, 0.4, ne)
    dd, dv = clamp_and_pull(m, vel=-0.5)  # compression: the elements' lengths, and so dt, fall below l0/c
    steps = 300
    out = {}
    for mode in ("0", None, "-0.5"):
        if mode is None:
            monkeypatch.delenv("MPS_DT_SKIP", raising=False)
        else:
            monkeypatch.setenv("MPS_DT_SKIP", mode)
        out[mode] = run(m, steps, E, nu, dd, dv, energy_every, [7, 150, steps - 157])
    ref = out["0"]
    for mode in (None, "-0.5"):
        for i in range(3):
            assert np.array_equal(out[mode][i], ref[i]), (mode, i)
        for k in ref[3]:
            assert np.array_equal(out[mode][3][k], ref[3][k], equal_nan=True), (mode, k)
    # the dt history falls over the run (the skip threshold follows it)
    dt = ref[3]["dt"]
    assert dt.min() < dt.max()


def test_dt_skip_matches_oracle(monkeypatch):
    """Default skip against the oracle's dt history and displacement (the FP64 bar)."""
    monkeypatch.delenv("MPS_DT_SKIP", raising=False)
    m = perturb(box(12, 10, 8), 0.1, 4)
    dd, dv = clamp_and_pull(m, vel=0.5)
    steps = 200
    op = O.OracleProblem(m, dir_dof=dd, dir_val=dv, n_steps=steps, sum_order=O.SUM_REFERENCE)
    ro = op.run(hist_cap=steps)
    with G.Solver(m, dir_dof=dd, dir_val=dv, n_steps=steps, history_cap=steps) as s:
        rc, rep = s.run(steps)
        assert rc == 0 and rep["steps"] == steps
        u = s.get_state()[0]
        h = s.history(0, steps)
    assert np.abs(u - ro["u"]).max() <= 1e-12 * np.abs(ro["u"]).max()
    np.testing.assert_allclose(h["dt"], ro["hist"]["dt"], rtol=1e-13, atol=0)

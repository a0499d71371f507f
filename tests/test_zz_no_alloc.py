"""DOF gather without allocation (mps_get_dofs_static) against the allocating gather (mps_get_dofs), through the
C-ABI: bitwise equal on two calls in a row (the second smaller), for u, v and a, and the size limit."""
import numpy as np
import pytest

import mps_gpu as G
from meshgen import box, clamp_and_pull, perturb

pytestmark = pytest.mark.gpu


def test_static_gather_equals_allocating_gather():
    m = perturb(box(7, 6, 5), 0.1, 91)
    dd, dv = clamp_and_pull(m)
    ndof = 3 * m["n_nodes"]
    rng = np.random.default_rng(12)
    with G.Solver(m, dir_dof=dd, dir_val=dv, n_steps=60) as s:
        rc, rep = s.run(35)
        assert rc == 0 and rep["steps"] == 35
        big = rng.integers(0, ndof, G.DOFS_STATIC_MAX).astype(np.int32)  # the largest call allowed
        for which in (0, 1, 2):
            ref = s.get_dofs(which, big)
            got = s.get_dofs_static(which, big)
            assert np.abs(ref).max() > 0 or which == 2
            assert np.array_equal(got, ref), which
        rc, rep = s.run(25)
        assert rc == 0 and rep["steps"] == 60
        small = rng.integers(0, ndof, 37).astype(np.int32)
        u = s.get_state()[0]
        for which in (0, 1, 2):
            got = s.get_dofs_static(which, small)  # second call, smaller: no stale entries of the first
            assert np.array_equal(got, s.get_dofs(which, small)), which
        assert np.array_equal(s.get_dofs_static(0, small), u[small])
        with pytest.raises(G.MpsError) as ei:
            s.get_dofs_static(0, np.zeros(G.DOFS_STATIC_MAX + 1, np.int32))
        assert ei.value.code == G.MPS_ERR_CAPACITY and "at most" in str(ei.value)
        with pytest.raises(G.MpsError) as ei:
            s.get_dofs_static(0, np.array([ndof], np.int32))
        assert ei.value.code == G.MPS_ERR_ARG


def test_static_gather_on_two_solvers():
    """Two solvers on the device share the staging array; calls in turn each see their own values."""
    a = box(4, 4, 4)
    b = perturb(box(5, 4, 3), 0.1, 4)
    with G.Solver(a, dir_dof=clamp_and_pull(a)[0], dir_val=clamp_and_pull(a)[1], n_steps=30) as sa, \
            G.Solver(b, dir_dof=clamp_and_pull(b)[0], dir_val=clamp_and_pull(b)[1], n_steps=30) as sb:
        assert sa.run(30)[0] == 0 and sb.run(30)[0] == 0
        ia = np.arange(0, 3 * a["n_nodes"], 3, dtype=np.int32)
        ib = np.arange(1, 3 * b["n_nodes"], 2, dtype=np.int32)
        ua, ub = sa.get_state()[0], sb.get_state()[0]
        for _ in range(2):
            assert np.array_equal(sa.get_dofs_static(0, ia), ua[ia])
            assert np.array_equal(sb.get_dofs_static(0, ib), ub[ib])

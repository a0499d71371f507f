"""Full-size checks on the device: both forms of the 1M-hex8 workload (200 steps) and the 3.07M-tet4 workload
(500 steps) against the threaded CPU oracle over their whole horizons (tens of seconds on 16 host threads), and
otherwise through size-independent properties (the oracle pins the same kernels at small sizes in
test_gpu_parity.py):

- the 1M-hex8 workload: bitwise determinism across graph replay, direct launches and chunking; mirror symmetry
  (x <-> y) of the unperturbed cube;
- a 4.1M-hex8 mesh (160^3): int32 index ranges of the patch tables hold, FP32 within 1e-6 of FP64;
- one element alone (hex8 and tet4) against the oracle.
"""
import numpy as np
import pytest

import mps_gpu as G
import oracle_ctypes as O
from mps_gpu import configs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def solve(cfg, steps, precision=G.PREC_FP64, graphs=True, chunks=None):
    m = cfg["mesh"]
    print(f"solve {m['n_elem']} elements, {steps} steps, precision {precision}, graphs {graphs}", flush=True)
    with G.Solver(m, law=cfg.get("law", 0), E=cfg["E"], nu=cfg["nu"], rho=cfg.get("rho", 1000.0),
                  dir_dof=cfg["dir_dof"], dir_val=cfg["dir_val"], bc_mode=G.BC_VELOCITY, vel_ramp_steps=100,
                  scale=cfg.get("scale", 0.5), n_steps=steps, precision=precision) as s:
        s.set_graphs(graphs)
        for c in chunks or [steps]:
            rc, rep = s.run(c)
            assert rc == 0, rep
        assert rep["steps"] == steps and rep["status"] == G.MPS_DONE
        return s.get_state()[0]


def oracle_threads():
    import os
    return max(1, min(16, os.cpu_count() or 1))


@pytest.mark.parametrize("k,steps", [(1, 200), (2, 500)])
def test_full_size_against_the_oracle(k, steps):
    """The 1M-hex8 workload (200 steps) and the 3.07M-tet4 workload (per-element materials, 500 steps), each over
    its whole horizon: the FP64 form against the oracle (element-order sums, on the host's threads) to the FP64
    bar 1e-12, and the FP32-element form against the same oracle to its bar 1e-6."""
    cfg = configs.config(k)
    m = cfg["mesh"]
    assert cfg["n_steps"] == steps
    u = solve(cfg, steps)
    u32 = solve(cfg, steps, precision=G.PREC_FP32_ELEM)
    print(f"oracle on {oracle_threads()} threads, cfg{k}, {steps} steps", flush=True)
    o = O.OracleProblem(m, law=cfg.get("law", 0), E=cfg["E"], nu=cfg["nu"], rho=cfg.get("rho", 1000.0),
                        dir_dof=cfg["dir_dof"], dir_val=cfg["dir_val"], bc_mode=O.BC_VELOCITY, vel_ramp_steps=100,
                        scale=cfg.get("scale", 0.5), n_steps=steps, sum_order=O.SUM_ELEMENT,
                        n_threads=oracle_threads()).run()
    assert o["rc"] == 0 and o["steps"] == steps
    err = rel(u, o["u"])
    err32 = rel(u32, o["u"])
    print(f"cfg{k} vs oracle after {steps} steps: fp64 {err:.3e}, fp32-element {err32:.3e}", flush=True)
    assert err <= 1e-12
    assert err32 <= 1e-6


def test_cfg1_determinism():
    cfg = configs.config(1)
    u64 = solve(cfg, 200)
    assert np.array_equal(solve(cfg, 200, graphs=False, chunks=[1, 63, 136]), u64)


def test_cfg1_unperturbed_mirror_symmetry():
    """x <-> y mirror of the clamped/pulled cube: u_x(i,j,k) = u_y(j,i,k).  The mirrored elements see a
    reflected local numbering, so only rounding differs."""
    n = 100
    m = configs.box(n, n, n)
    dof, val = configs.clamp_bottom_pull_top(m)
    u = solve(dict(mesh=m, E=1e6, nu=0.3, dir_dof=dof, dir_val=val), 200)
    N = m["n_nodes"]
    ux, uy, uz = u[:N].reshape(n + 1, n + 1, n + 1), u[N:2 * N].reshape(n + 1, n + 1, n + 1), \
        u[2 * N:].reshape(n + 1, n + 1, n + 1)   # [k, j, i]
    scale = np.abs(u).max()
    assert scale > 0
    assert np.abs(ux - uy.transpose(0, 2, 1)).max() <= 1e-10 * scale
    assert np.abs(uz - uz.transpose(0, 2, 1)).max() <= 1e-10 * scale


def test_4m_hex_index_ranges():
    m = configs.perturb(configs.box(160, 160, 160), 0.1, 7)
    dof, val = configs.clamp_bottom_pull_top(m)
    cfg = dict(mesh=m, E=1e6, nu=0.3, dir_dof=dof, dir_val=val)
    u64 = solve(cfg, 20)
    u32 = solve(cfg, 20, precision=G.PREC_FP32_ELEM)
    assert np.isfinite(u64).all()
    assert rel(u32, u64) <= 1e-6


@pytest.mark.parametrize("tet", [False, True])
def test_single_element(tet):
    m = configs.box(1, 1, 1, tet=tet)
    if tet:  # one tet of the split
        m = dict(m, n_elem=1, conn=np.ascontiguousarray(m["conn"][:4]))
    dof, val = configs.clamp_bottom_pull_top(m, vel=0.05)
    used = np.unique(m["conn"])
    keep = np.isin(np.asarray(dof) % m["n_nodes"], used)
    dof, val = np.ascontiguousarray(np.asarray(dof)[keep]), np.ascontiguousarray(np.asarray(val)[keep])
    o = O.OracleProblem(m, dir_dof=dof, dir_val=val, n_steps=300).run()
    with G.Solver(m, dir_dof=dof, dir_val=val, n_steps=300) as s:
        rc, rep = s.run(300)
        assert rc == 0 and rep["steps"] == 300
        u = s.get_state()[0]
    assert rel(u, o["u"]) <= 1e-12

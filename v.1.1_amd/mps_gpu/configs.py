"""Synthetic workloads of the explicit hex8/tet4 neo-Hookean path.

Meshes are structured boxes with the reference node order
(classLinearHexahedra.cpp:346-353); tet meshes split every hex into 6 tets around
the 0-6 diagonal, the pattern of the reference's Hex2Tet tool
(applications/tools/Hex2Tet-Parallel/Hex2Tet-Partitioning.py:87-109).

Random numbers: counter-based SplitMix64, k-th draw = mix(seed + (k+1)*0x9E3779B97F4A7C15),
uniform = (draw >> 11) * 2^-53.  Node perturbations use draw index 3*node + dir; per-element
materials use the seed xor 0x5A5A5A5A5A5A5A5A with draw index 2*elem (E) and 2*elem+1 (nu).
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED_CFG1 = 0x4D50530001
SEED_CFG2 = 0x4D50530002
SEED_CFG3 = 0x4D50530003
SEED_CFG4 = 0x4D50530004
MAT_XOR = 0x5A5A5A5A5A5A5A5A

HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)
TET_SPLIT = np.array([[0, 2, 3, 6], [0, 1, 2, 6], [3, 0, 6, 7],
                      [0, 4, 6, 7], [0, 4, 5, 6], [1, 0, 5, 6]], dtype=np.int64)


def splitmix64(seed, index):
    """Draws number `index` (uint64 array) of the SplitMix64 stream started at `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (np.asarray(index, dtype=np.uint64) + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, index):
    return (splitmix64(seed, index) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, tet=False, reflected=False):
    """Structured box; node id = i + (nx+1)*(j + (ny+1)*k), elements x-fastest.

    reflected: each cell's local corner numbering mirrored along x, y, z where the cell index along that axis is
    odd.  Hex8 cells then alternate their orientation (det J < 0 on an odd number of mirrors); tet4 cells get
    the 6-tet split in alternating mirror images, which stays conforming (two cells sharing a face differ in one
    axis' parity, so both split it along the same diagonal) and gives the nodes valences other than 24."""
    N = (nx + 1) * (ny + 1) * (nz + 1)
    ids = np.arange(N, dtype=np.int64)
    i = ids % (nx + 1)
    j = (ids // (nx + 1)) % (ny + 1)
    k = ids // ((nx + 1) * (ny + 1))
    X = np.empty(3 * N)
    X[:N] = i * (lx / nx)
    X[N:2 * N] = j * (ly / ny)
    X[2 * N:] = k * (lz / nz)
    e = np.arange(nx * ny * nz, dtype=np.int64)
    ei = e % nx
    ej = (e // nx) % ny
    ek = e // (nx * ny)
    hexes = np.empty((e.size, 8), dtype=np.int64)
    fi, fj, fk = (ei & 1, ej & 1, ek & 1) if reflected else (0, 0, 0)
    for a, (ci, cj, ck) in enumerate(HEX_CORNERS):
        hexes[:, a] = (ei + (ci ^ fi)) + (nx + 1) * ((ej + (cj ^ fj)) + (ny + 1) * (ek + (ck ^ fk)))
    if tet:
        conn = hexes[:, TET_SPLIT].reshape(-1, 4)
        nen = 4
    else:
        conn = hexes
        nen = 8
    return dict(X=X, conn=np.ascontiguousarray(conn.ravel(), dtype=np.int32), n_nodes=N, n_elem=conn.shape[0],
                nen=nen, nx=nx, ny=ny, nz=nz, h=(lx / nx, ly / ny, lz / nz))


def perturb(mesh, frac, seed):
    """Interior nodes (not on the box boundary) moved by U(-frac*h, frac*h) per direction."""
    N = mesh["n_nodes"]
    nx, ny, nz = mesh["nx"], mesh["ny"], mesh["nz"]
    ids = np.arange(N, dtype=np.int64)
    i = ids % (nx + 1)
    j = (ids // (nx + 1)) % (ny + 1)
    k = ids // ((nx + 1) * (ny + 1))
    interior = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
    X = mesh["X"].copy()
    for d in range(3):
        r = uniform01(seed, 3 * ids + d) * 2.0 - 1.0
        X[d * N:(d + 1) * N] += np.where(interior, r * (frac * mesh["h"][d]), 0.0)
    out = dict(mesh)
    out["X"] = X
    return out


def clamp_bottom_pull_top(mesh, vel=0.1):
    """z = 0 face fixed in x, y, z; z = max face prescribed along z with amplitude `vel`."""
    N = mesh["n_nodes"]
    nx, ny, nz = mesh["nx"], mesh["ny"], mesh["nz"]
    per_layer = (nx + 1) * (ny + 1)
    bot = np.arange(per_layer, dtype=np.int64)
    top = bot + nz * per_layer
    dof = np.concatenate([bot, bot + N, bot + 2 * N, top + 2 * N]).astype(np.int32)
    val = np.concatenate([np.zeros(3 * per_layer), np.full(per_layer, vel)])
    return dof, val


def random_materials(n, seed, e_lo=1e5, e_hi=1e7, nu_lo=0.2, nu_hi=0.45):
    s = seed ^ MAT_XOR
    idx = np.arange(n, dtype=np.int64)
    E = 10.0 ** (np.log10(e_lo) + (np.log10(e_hi) - np.log10(e_lo)) * uniform01(s, 2 * idx))
    nu = nu_lo + (nu_hi - nu_lo) * uniform01(s, 2 * idx + 1)
    return E, nu


def config(k, scale_down=1):
    """Workload k of the explicit path as a dict of solver inputs.

    scale_down > 1 shrinks the mesh (same recipe) for CPU-sized checks.
    """
    if k == 0:
        m = box(10, 10, 10)
        dof, val = clamp_bottom_pull_top(m)
        return dict(name="cfg0: 10^3 hex8 unit cube, 1000 steps", mesh=m, law=0, E=1e6, nu=0.3, rho=1000.0,
                    dir_dof=dof, dir_val=val, n_steps=1000, scale=0.5)
    if k == 1:
        n = 100 // scale_down
        m = perturb(box(n, n, n), 0.10, SEED_CFG1)
        dof, val = clamp_bottom_pull_top(m)
        return dict(name=f"cfg1: {n}^3 hex8 perturbed 10%, 200 steps", mesh=m, law=0, E=1e6, nu=0.3, rho=1000.0,
                    dir_dof=dof, dir_val=val, n_steps=200, scale=0.5)
    if k == 2:
        n = 80 // scale_down
        m = perturb(box(n, n, n, tet=True), 0.15, SEED_CFG2)
        E, nu = random_materials(m["n_elem"], SEED_CFG2)
        dof, val = clamp_bottom_pull_top(m)
        return dict(name=f"cfg2: {n}^3x6 tet4 perturbed 15%, per-element E/nu, 500 steps", mesh=m, law=0, E=E,
                    nu=nu, rho=1000.0, dir_dof=dof, dir_val=val, n_steps=500, scale=0.5)
    raise ValueError(k)


def batch_instances(count=512, seed=SEED_CFG3, n_lo=6, n_hi=40, steps=1000):
    """Workload of the batched form: `count` clamped-bottom / top-loaded cubes of side n ~ U{n_lo..n_hi},
    E log-uniform [1e5,1e7], nu uniform [0.2,0.45], rho 1000."""
    idx = np.arange(count, dtype=np.int64)
    sides = n_lo + (uniform01(seed, 3 * idx) * (n_hi - n_lo + 1)).astype(np.int64)
    E = 10.0 ** (5.0 + 2.0 * uniform01(seed, 3 * idx + 1))
    nu = 0.2 + 0.25 * uniform01(seed, 3 * idx + 2)
    return [dict(n=int(s), E=float(e), nu=float(v), steps=steps) for s, e, v in zip(sides, E, nu)]


def serving_requests(count, seed=SEED_CFG4, ne_lo=1e3, ne_hi=2e5, steps_lo=500, steps_hi=5000):
    """Requests of the serving workload: hex8 boxes with N_e log-uniform in [ne_lo, ne_hi], aspect ratio
    (long / short side) uniform in [1, 4] along z, element size 1/n_z (longest side 1 m), material E
    log-uniform [1e5, 1e7] Pa, nu uniform [0.2, 0.45], rho 1000, bottom clamped, top pulled along z at
    0.1 m/s (ramp over 100 steps), step count uniform in {steps_lo..steps_hi}."""
    idx = np.arange(count, dtype=np.int64)
    u = [uniform01(seed, 6 * idx + j) for j in range(6)]
    ne = 10.0 ** (np.log10(ne_lo) + (np.log10(ne_hi) - np.log10(ne_lo)) * u[0])
    asp = 1.0 + 3.0 * u[1]
    out = []
    for i in range(count):
        s = (ne[i] / asp[i]) ** (1.0 / 3.0)
        nx = max(1, int(round(s)))
        nz = max(1, int(round(asp[i] * s)))
        h = 1.0 / nz
        m = box(nx, nx, nz, lx=nx * h, ly=nx * h, lz=1.0)
        dof, val = clamp_bottom_pull_top(m, vel=0.1)
        out.append(dict(mesh=m, E=float(10.0 ** (5.0 + 2.0 * u[2][i])), nu=float(0.2 + 0.25 * u[3][i]),
                        dir_dof=dof, dir_val=val,
                        n_steps=int(steps_lo + int(u[4][i] * (steps_hi - steps_lo + 1)))))
    return out


def batch_meshes(count=512, seed=SEED_CFG3, steps=1000):
    """Solves of the batched workload (unit cubes n^3, see batch_instances), ready for submission."""
    out = []
    for spec in batch_instances(count, seed, steps=steps):
        m = box(spec["n"], spec["n"], spec["n"])
        dof, val = clamp_bottom_pull_top(m, vel=0.1)
        out.append(dict(mesh=m, E=spec["E"], nu=spec["nu"], dir_dof=dof, dir_val=val, n_steps=steps))
    return out

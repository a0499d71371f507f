"""Small structured meshes for tests (numpy).  The product's own generators are
native (C++); this module builds the same node order for CPU-side checks.

Hex node order follows classLinearHexahedra.cpp:346-353:
(-,-,-), (+,-,-), (+,+,-), (-,+,-), (-,-,+), (+,-,+), (+,+,+), (-,+,+).
"""
import numpy as np

HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
# 6-tet split of a hex sharing the 0-6 diagonal (pattern of the reference's
# Hex2Tet tool, applications/tools/Hex2Tet-Parallel/Hex2Tet-Partitioning.py:87-109).
TET_SPLIT = np.array([[0, 2, 3, 6], [0, 1, 2, 6], [3, 0, 6, 7],
                      [0, 4, 6, 7], [0, 4, 5, 6], [1, 0, 5, 6]])


def node_id(i, j, k, nx, ny):
    return i + (nx + 1) * (j + (ny + 1) * k)


def box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, tet=False):
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    # node id = i + (nx+1)*(j + (ny+1)*k): order x fastest
    N = (nx + 1) * (ny + 1) * (nz + 1)
    X = np.zeros((3, N))
    ids = node_id(ii, jj, kk, nx, ny).ravel()
    X[0, ids] = (ii.ravel() * (lx / nx))
    X[1, ids] = (jj.ravel() * (ly / ny))
    X[2, ids] = (kk.ravel() * (lz / nz))
    hexes = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                hexes.append([node_id(i + c[0], j + c[1], k + c[2], nx, ny) for c in HEX_CORNERS])
    hexes = np.array(hexes, dtype=np.int32)
    if tet:
        conn = hexes[:, TET_SPLIT].reshape(-1, 4)
        nen = 4
    else:
        conn = hexes
        nen = 8
    return dict(X=X.ravel(), conn=conn.ravel().astype(np.int32), n_nodes=N,
                n_elem=conn.shape[0], nen=nen, nx=nx, ny=ny, nz=nz)


def face_nodes(mesh, axis, value, tol=1e-12):
    X = mesh["X"].reshape(3, -1)
    return np.nonzero(np.abs(X[axis] - value) < tol)[0].astype(np.int32)


def clamp_and_pull(mesh, vel=0.1, axis=2):
    """Bottom (min axis) face clamped in x,y,z; top face prescribed along `axis`."""
    N = mesh["n_nodes"]
    X = mesh["X"].reshape(3, -1)
    lo, hi = X[axis].min(), X[axis].max()
    bot = face_nodes(mesh, axis, lo)
    top = face_nodes(mesh, axis, hi)
    dofs, vals = [], []
    for n in bot:
        for d in range(3):
            dofs.append(n + N * d)
            vals.append(0.0)
    for n in top:
        dofs.append(n + N * axis)
        vals.append(vel)
    return np.array(dofs, np.int32), np.array(vals, np.float64)


def perturb(mesh, frac, seed):
    """Interior nodes moved by U(-frac*h, frac*h) per coordinate (numpy RNG; tests only)."""
    X = mesh["X"].reshape(3, -1).copy()
    rng = np.random.default_rng(seed)
    h = np.array([X[d].max() / [mesh["nx"], mesh["ny"], mesh["nz"]][d] for d in range(3)])
    lo = X.min(axis=1, keepdims=True)
    hi = X.max(axis=1, keepdims=True)
    interior = np.all((X > lo + 1e-12) & (X < hi - 1e-12), axis=0)
    d = rng.uniform(-1, 1, size=X.shape) * (frac * h)[:, None]
    X[:, interior] += d[:, interior]
    out = dict(mesh)
    out["X"] = X.ravel()
    return out

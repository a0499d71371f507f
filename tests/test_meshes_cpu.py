"""The synthetic meshes of the workloads (mps_gpu.configs), checked on the CPU: the tet4 split is conforming
(every interior triangle shared by exactly two tets, every boundary triangle by one), fills the box, and in its
reflected form alternates orientation and widens the node valences (the CSR gather's row lengths)."""
import numpy as np
import pytest

from mps_gpu import configs


def tet_faces(conn):
    faces = {}
    for t in conn:
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            key = tuple(sorted(t[list(f)]))
            faces[key] = faces.get(key, 0) + 1
    return faces


def signed_volumes(m):
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    c = m["conn"].reshape(-1, m["nen"])
    e1, e2, e3 = (X[:, c[:, a]] - X[:, c[:, 0]] for a in (1, 2, 3))
    return np.einsum("ij,ij->j", e1, np.cross(e2.T, e3.T).T) / 6.0


@pytest.mark.parametrize("reflected", [False, True])
def test_tet_split_is_conforming_and_fills_the_box(reflected):
    nx, ny, nz = 5, 4, 3
    m = configs.box(nx, ny, nz, lx=2.0, ly=1.5, lz=1.0, tet=True, reflected=reflected)
    faces = tet_faces(m["conn"].reshape(-1, 4))
    mult = np.bincount(list(faces.values()))
    boundary = 2 * 2 * (nx * ny + ny * nz + nx * nz)  # two triangles per boundary quad
    assert mult.size == 3 and mult[1] == boundary
    assert 2 * mult[2] + mult[1] == 4 * m["n_elem"]
    vol = signed_volumes(m)
    assert abs(np.abs(vol).sum() - 3.0) < 1e-12
    if reflected:
        assert 0 < (vol < 0).sum() < vol.size
    else:
        assert np.all(vol > 0)


def test_reflected_split_widens_the_valences():
    std = configs.box(6, 6, 6, tet=True)
    ref = configs.box(6, 6, 6, tet=True, reflected=True)
    v_std = np.bincount(std["conn"], minlength=std["n_nodes"])
    v_ref = np.bincount(ref["conn"], minlength=ref["n_nodes"])
    assert v_std.sum() == v_ref.sum() and v_std.max() == 24 and v_ref.max() > 24


def test_reflected_hex_alternates_orientation():
    m = configs.box(4, 4, 4, reflected=True)
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    c = m["conn"].reshape(-1, 8)
    # orientation of the corner frame (corner 1 - 0, corner 3 - 0, corner 4 - 0): one sign per mirror count
    e1, e3, e4 = (X[:, c[:, a]] - X[:, c[:, 0]] for a in (1, 3, 4))
    det = np.einsum("ij,ij->j", e1, np.cross(e3.T, e4.T).T)
    e = np.arange(64)
    mirrors = (e % 4 & 1) + (e // 4 % 4 & 1) + (e // 16 & 1)
    assert np.array_equal(det < 0, mirrors % 2 == 1)

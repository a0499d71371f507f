"""Pins of the CPU oracle (oracle/mps_oracle.c) against closed forms, invariants and an
independent textbook computation, function by function.

The reference ships no golden vectors for this path (no fixtures, no unit tests), so each
oracle function is checked against what the mathematics fixes, chosen so that a dropped term,
a wrong sign or a transposed operand fails at least one check.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes as O
from meshgen import box, clamp_and_pull, perturb

L = O.lib()
G = 0.577350269  # Gauss abscissa literal of the reference (GPsDistribution.cpp:392-432)
HEX_S = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def pk1(law, F, E=1e6, nu=0.3):
    P = np.zeros(9)
    L.orc_pk1(law, O.dptr(arr(F).ravel()), E, nu, O.dptr(P))
    return P.reshape(3, 3)


def energy(law, F, E=1e6, nu=0.3):
    return L.orc_defo_energy(law, O.dptr(arr(F).ravel()), E, nu)


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))


def gp_data(nen, xyz):
    xw = np.zeros(32)
    ngp = L.orc_quadrature(nen, O.dptr(xw))
    out = []
    for g in range(ngp):
        phi, dphi, wJ, dJ = np.zeros(8), np.zeros(24), C.c_double(), C.c_double()
        L.orc_define_gp(nen, O.dptr(arr(xyz).ravel()), O.dptr(xw[4 * g:4 * g + 4].copy()), O.dptr(phi), O.dptr(dphi),
                        C.byref(wJ), C.byref(dJ))
        out.append((phi[:nen].copy(), dphi[:3 * nen].reshape(nen, 3).copy(), wJ.value))
    return out


def rand_F(rng, eps=0.05):
    return np.eye(3) + rng.uniform(-eps, eps, (3, 3))


# ------------------------------------------------------------------ 3x3 algebra

def test_det_inverse_against_numpy():
    rng = np.random.default_rng(0)
    for _ in range(50):
        A = rng.normal(size=(3, 3))
        inv = np.zeros(9)
        L.orc_inverse3(O.dptr(arr(A).ravel()), O.dptr(inv))
        assert L.orc_det3(O.dptr(arr(A).ravel())) == pytest.approx(np.linalg.det(A), rel=1e-12)
        np.testing.assert_allclose(inv.reshape(3, 3), np.linalg.inv(A), rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------------ quadrature, shape functions

def test_quadrature_literals():
    xw = np.zeros(32)
    assert L.orc_quadrature(8, O.dptr(xw)) == 8
    q = xw.reshape(8, 4)
    np.testing.assert_array_equal(q[:, :3], HEX_S * G)  # same order as the hex nodes
    np.testing.assert_array_equal(q[:, 3], 1.0)
    assert L.orc_quadrature(4, O.dptr(xw)) == 1
    np.testing.assert_array_equal(xw[:4], [0.25, 0.25, 0.25, 0.166666666667])


def test_hex_shape_functions_partition_and_kronecker():
    phi, nd = np.zeros(8), np.zeros(24)
    for a in range(8):  # phi_b(node a) = delta_ab
        L.orc_hex8_shape(O.dptr(arr([*HEX_S[a], 0.0])), O.dptr(phi), O.dptr(nd))
        np.testing.assert_array_equal(phi, np.eye(8)[a])
    rng = np.random.default_rng(1)
    for _ in range(10):
        xi = rng.uniform(-1, 1, 3)
        L.orc_hex8_shape(O.dptr(arr([*xi, 0.0])), O.dptr(phi), O.dptr(nd))
        assert phi.sum() == pytest.approx(1.0, abs=1e-15)
        np.testing.assert_allclose(nd.reshape(8, 3).sum(axis=0), 0.0, atol=1e-15)
        # natural derivative = finite difference of phi
        h = 1e-6
        for d in range(3):
            p1, p2, tmp = np.zeros(8), np.zeros(8), np.zeros(24)
            xp, xm = xi.copy(), xi.copy()
            xp[d] += h
            xm[d] -= h
            L.orc_hex8_shape(O.dptr(arr([*xp, 0.0])), O.dptr(p1), O.dptr(tmp))
            L.orc_hex8_shape(O.dptr(arr([*xm, 0.0])), O.dptr(p2), O.dptr(tmp))
            np.testing.assert_allclose((p1 - p2) / (2 * h), nd.reshape(8, 3)[:, d], atol=1e-9)


@pytest.mark.parametrize("nen", [8, 4])
def test_define_gp_isoparametric_reproduction(nen):
    rng = np.random.default_rng(2)
    A = np.eye(3) + rng.uniform(-0.3, 0.3, (3, 3))
    b = rng.uniform(-1, 1, 3)
    ref = (HEX_S + 1) / 2 if nen == 8 else np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    xyz = ref @ A.T + b
    gps = gp_data(nen, xyz)
    wsum = 0.0
    for phi, dphi, wJ in gps:
        assert phi.sum() == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_allclose(dphi.sum(axis=0), 0.0, atol=1e-12)
        np.testing.assert_allclose(xyz.T @ dphi, np.eye(3), atol=1e-12)  # sum_a X_a (x) dphi_a = I
        wsum += wJ
    vol = abs(np.linalg.det(A)) * (1.0 if nen == 8 else 1.0 / 6.0)
    if nen == 8:
        assert wsum == pytest.approx(vol, rel=1e-13)  # affine hex, weight 1: exact
    else:
        assert wsum == pytest.approx(vol * 6 * 0.166666666667, rel=1e-13)  # tet weight literal (1+2e-12)


# ------------------------------------------------------------------ characteristic length

def char_len(nen, xyz):
    return L.orc_char_length(nen, O.dptr(arr(xyz).ravel()))


def test_char_length_cube_and_regular_tet():
    for h in (0.1, 1.0, 3.7):
        assert char_len(8, (HEX_S + 1) / 2 * h) == pytest.approx(h, rel=1e-15)  # r = 3h^3/(6h^2)
    a = 0.8  # regular tetrahedron of edge a: inradius a/(2 sqrt 6)
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1.0]]) * a / (2 * np.sqrt(2))
    assert char_len(4, tet) == pytest.approx(a / np.sqrt(6), rel=1e-14)


def test_char_length_rigid_invariance_and_parallelepiped_volume():
    rng = np.random.default_rng(3)
    x = (HEX_S + 1) / 2 + rng.uniform(-0.1, 0.1, (8, 3))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Q *= np.sign(np.linalg.det(Q))
    assert char_len(8, x @ Q.T + 5.0) == pytest.approx(char_len(8, x), rel=1e-13)
    # the hex "volume" is |det(p1-p0, p3-p0, p4-p0)| (geometry.cpp:348-351): moving node 6 along the
    # normal of none of the faces used by V changes l only through the face areas
    y = x.copy()
    y[6] += np.array([0.05, 0.02, 0.03])
    faces = [[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 3, 7, 4]]

    def area(p):
        return sum(0.5 * np.linalg.norm(np.cross(p[f[i]] - p[f[j]], p[f[k]] - p[f[j]]))
                   for f in faces for (i, k, j) in ((0, 1, 2), (0, 2, 3)))
    V = abs(np.linalg.det(np.array([x[1] - x[0], x[3] - x[0], x[4] - x[0]])))
    assert char_len(8, y) == pytest.approx(2 * 3 * V / area(y), rel=1e-13)


# ------------------------------------------------------------------ constitutive laws

def test_material_constants_and_sound_speed():
    for E, nu, rho in ((1e6, 0.3, 1000.0), (3e9, 0.45, 7800.0)):
        lam, mu, K = np.zeros(1), np.zeros(1), np.zeros(1)
        L.orc_material(E, nu, O.dptr(lam), O.dptr(mu), O.dptr(K))
        assert K[0] == pytest.approx(lam[0] + 2 * mu[0] / 3, rel=1e-14)
        c_p = np.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu) * rho))  # P-wave speed
        assert L.orc_sound_speed(E, nu, rho) == pytest.approx(c_p, rel=1e-14)


@pytest.mark.parametrize("law", [0, 1, 2])
def test_pk1_reference_state_and_objectivity(law):
    np.testing.assert_allclose(pk1(law, np.eye(3)), 0.0, atol=1e-9)
    rng = np.random.default_rng(4 + law)
    for _ in range(10):
        F = rand_F(rng, 0.1)
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        Q *= np.sign(np.linalg.det(Q))
        np.testing.assert_allclose(pk1(law, Q @ F), Q @ pk1(law, F), rtol=1e-9, atol=1e-6)


def test_pk1_compressible_closed_form_and_energy_gradient():
    lam, mu, _ = lame(1e6, 0.3)
    rng = np.random.default_rng(6)
    for _ in range(10):
        F = rand_F(rng, 0.2)
        Fit = np.linalg.inv(F).T
        J = np.linalg.det(F)
        np.testing.assert_allclose(pk1(0, F), mu * (F - Fit) + lam * np.log(J) * Fit, rtol=1e-9, atol=1e-6)
        h = 1e-6  # hyperelasticity: P = dW/dF
        dW = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                Fp, Fm = F.copy(), F.copy()
                Fp[i, j] += h
                Fm[i, j] -= h
                dW[i, j] = (energy(0, Fp) - energy(0, Fm)) / (2 * h)
        np.testing.assert_allclose(dW, pk1(0, F), rtol=1e-5, atol=1e-2)
    assert energy(0, np.eye(3)) == pytest.approx(1.5 * mu, rel=1e-15)  # no -3mu/2 constant
    assert energy(1, rand_F(rng)) == 0.0  # Belytschko law has no defoEnergy


def test_pk1_svk_and_linear_closed_forms_and_energy_gradients():
    """St-Venant-Kirchhoff: P = F (lam tr(E) I + 2 mu E), E = (F^T F - I)/2; linear elastic: P = lam tr(e) I
    + 2 mu e, e = sym(F) - I (not objective: a rotation is not stress free).  Both: P = dW/dF."""
    lam, mu, _ = lame(1e6, 0.3)
    rng = np.random.default_rng(12)
    for law in (2, 3):
        np.testing.assert_allclose(pk1(law, np.eye(3)), 0.0, atol=1e-9)
        for _ in range(6):
            F = rand_F(rng, 0.2)
            if law == 2:
                E = 0.5 * (F.T @ F - np.eye(3))
                ref = F @ (lam * np.trace(E) * np.eye(3) + 2 * mu * E)
            else:
                e = 0.5 * (F + F.T) - np.eye(3)
                ref = lam * np.trace(e) * np.eye(3) + 2 * mu * e
            np.testing.assert_allclose(pk1(law, F), ref, rtol=1e-9, atol=1e-6)
            h = 1e-6
            dW = np.zeros((3, 3))
            for i in range(3):
                for j in range(3):
                    Fp, Fm = F.copy(), F.copy()
                    Fp[i, j] += h
                    Fm[i, j] -= h
                    dW[i, j] = (energy(law, Fp) - energy(law, Fm)) / (2 * h)
            np.testing.assert_allclose(dW, pk1(law, F), rtol=1e-5, atol=1e-2)
    ang = 0.3
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    assert np.abs(pk1(2, R)).max() < 1e-6 * mu       # SVK: rotation is stress free
    assert np.abs(pk1(3, R)).max() > 1e-2 * mu       # small-strain law: it is not
    for law in (2, 3):  # sqrt(E/rho) (hyperElasticStVenantKirchhoff.cpp, linearElastic.cpp:49-54)
        assert L.orc_law_sound_speed(law, 2e9, 0.3, 7800.0) == np.sqrt(2e9 / 7800.0)
    assert L.orc_law_sound_speed(0, 2e9, 0.3, 7800.0) == L.orc_sound_speed(2e9, 0.3, 7800.0)


def test_pk1_belytschko_closed_form_and_pure_dilation():
    _, mu, K = lame(1e6, 0.3)
    rng = np.random.default_rng(7)
    for _ in range(10):
        F = rand_F(rng, 0.2)
        J = np.linalg.det(F)
        B = F @ F.T
        sig = mu * J ** (-5 / 3) * (B - np.trace(B) / 3 * np.eye(3)) + K * (J - 1) * np.eye(3)
        np.testing.assert_allclose(pk1(1, F), J * sig @ np.linalg.inv(F).T, rtol=1e-9, atol=1e-6)
    s = 1.02
    F = s * np.eye(3)
    J = s ** 3
    np.testing.assert_allclose(pk1(1, F), K * (J - 1) * J * np.linalg.inv(F).T, rtol=1e-12)


# ------------------------------------------------------------------ kinematics, force, mass

def element_force(nen, xyz, uloc, law=0, E=1e6, nu=0.3):
    """Element nodal forces [nen,3] through orc_kinematics/orc_pk1/orc_gp_force."""
    f = np.zeros((nen, 3))
    for phi, dphi, wJ in gp_data(nen, xyz):
        F = np.zeros(9)
        L.orc_kinematics(nen, O.dptr(arr(dphi).ravel()), O.dptr(arr(uloc).ravel()), O.dptr(F))
        P = pk1(law, F.reshape(3, 3), E, nu)
        fg = np.zeros(3 * nen)
        L.orc_gp_force(nen, wJ, O.dptr(arr(P).ravel()), O.dptr(arr(dphi).ravel()), O.dptr(fg))
        f += fg.reshape(3, nen).T  # vForce(a + i*nen)
    return f


@pytest.mark.parametrize("nen", [8, 4])
def test_kinematics_patch_test(nen):
    rng = np.random.default_rng(8)
    ref = (HEX_S + 1) / 2 if nen == 8 else np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    xyz = ref + rng.uniform(-0.1, 0.1, ref.shape)
    Gm = rand_F(rng, 0.1)
    u = xyz @ (Gm - np.eye(3)).T  # u = (G - I) X  =>  F = G at every Gauss point
    for phi, dphi, wJ in gp_data(nen, xyz):
        F = np.zeros(9)
        L.orc_kinematics(nen, O.dptr(arr(dphi).ravel()), O.dptr(arr(u).ravel()), O.dptr(F))
        np.testing.assert_allclose(F.reshape(3, 3), Gm, atol=1e-13)
        L.orc_kinematics(nen, O.dptr(arr(dphi).ravel()), O.dptr(arr(np.ones((nen, 3)) * 0.3).ravel()), O.dptr(F))
        np.testing.assert_allclose(F.reshape(3, 3), np.eye(3), atol=1e-14)


@pytest.mark.parametrize("law", [0, 1, 2])
def test_element_force_equilibrium_and_rigid_rotation(law):
    rng = np.random.default_rng(9)
    xyz = (HEX_S + 1) / 2 + rng.uniform(-0.1, 0.1, (8, 3))
    u = rng.uniform(-0.02, 0.02, (8, 3))
    f = element_force(8, xyz, u, law)
    assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).max()  # sum_a f_a = 0
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Q *= np.sign(np.linalg.det(Q))
    c = xyz.mean(axis=0)
    urot = (xyz - c) @ Q.T + c - xyz  # rigid rotation: stress free
    assert np.abs(element_force(8, xyz, urot, law)).max() <= 1e-7 * np.abs(f).max()


def element_energy(nen, xyz, uloc, E=1e6, nu=0.3, law=0):
    e = 0.0
    for phi, dphi, wJ in gp_data(nen, xyz):
        F = np.zeros(9)
        L.orc_kinematics(nen, O.dptr(arr(dphi).ravel()), O.dptr(arr(uloc).ravel()), O.dptr(F))
        e += wJ * energy(law, F.reshape(3, 3), E, nu)
    return e


@pytest.mark.parametrize("nen", [8, 4])
@pytest.mark.parametrize("law", [0, 2, 3])
def test_element_force_is_energy_gradient_at_finite_strain(nen, law):
    # f_a = d/du_a sum_gp wJ W(F): pins the P-dphi contraction order (P is not symmetric here)
    rng = np.random.default_rng(11)
    ref = (HEX_S + 1) / 2 if nen == 8 else np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    xyz = ref + rng.uniform(-0.1, 0.1, ref.shape)
    ang = 0.4
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    u = xyz @ (R @ np.diag([1.15, 0.9, 1.05]) - np.eye(3)).T + rng.uniform(-0.02, 0.02, ref.shape)
    f = element_force(nen, xyz, u, law)
    h = 1e-7
    g = np.zeros_like(u)
    for a in range(nen):
        for i in range(3):
            up, um = u.copy(), u.copy()
            up[a, i] += h
            um[a, i] -= h
            g[a, i] = (element_energy(nen, xyz, up, law=law) - element_energy(nen, xyz, um, law=law)) / (2 * h)
    np.testing.assert_allclose(f, g, rtol=1e-5, atol=1e-5 * np.abs(f).max())


def textbook_hex_stiffness(xyz, E, nu):
    """Small-strain isotropic stiffness of a hex8 (Voigt B-matrix, 2x2x2 Gauss at +-0.577350269)."""
    lam, mu, _ = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    K = np.zeros((24, 24))
    for s in HEX_S * G:
        dN = np.array([0.125 * HEX_S[a] * np.array([(1 + HEX_S[a][1] * s[1]) * (1 + HEX_S[a][2] * s[2]),
                                                    (1 + HEX_S[a][0] * s[0]) * (1 + HEX_S[a][2] * s[2]),
                                                    (1 + HEX_S[a][0] * s[0]) * (1 + HEX_S[a][1] * s[1])])
                       for a in range(8)])
        Jm = xyz.T @ dN
        dX = dN @ np.linalg.inv(Jm)
        B = np.zeros((6, 24))
        for a in range(8):
            bx, by, bz = dX[a]
            B[:, 3 * a:3 * a + 3] = [[bx, 0, 0], [0, by, 0], [0, 0, bz], [by, bx, 0], [0, bz, by], [bz, 0, bx]]
        K += B.T @ D @ B * abs(np.linalg.det(Jm))
    return K


@pytest.mark.parametrize("law", [0, 1, 2, 3])
def test_element_force_linearises_to_textbook_stiffness(law):
    rng = np.random.default_rng(10)
    xyz = (HEX_S + 1) / 2 * 0.1 + rng.uniform(-0.005, 0.005, (8, 3))
    K = textbook_hex_stiffness(xyz, 1e6, 0.3)
    r = rng.uniform(-1, 1, (8, 3))
    eps = 1e-8
    f = element_force(8, xyz, eps * r, law)
    np.testing.assert_allclose(f.ravel() / eps, K @ r.ravel(), rtol=1e-5, atol=1e-5 * np.abs(K @ r.ravel()).max())


def test_lumped_mass_totals_and_corner_mass():
    m = box(3, 2, 4)
    p = O.OracleProblem(m, rho=1000.0, n_steps=1)
    mass = p.mass()
    assert mass.sum() == pytest.approx(1000.0 * 1.0, rel=1e-13)  # rho * volume of the unit box
    h3 = (1 / 3) * (1 / 2) * (1 / 4)
    assert mass[0] == pytest.approx(1000.0 * h3 / 8, rel=1e-13)  # corner node of one element
    mt = box(2, 2, 2, tet=True)
    assert O.OracleProblem(mt, rho=1000.0, n_steps=1).mass().sum() == pytest.approx(1000.0 * 6 * 0.166666666667,
                                                                                    rel=1e-13)


def test_internal_force_homogeneous_patch_interior_nodes():
    m = perturb(box(3, 3, 3), 0.15, 5)
    Gm = np.eye(3) + np.array([[0.01, 0.003, 0], [0.002, -0.004, 0.001], [0, 0.002, 0.006]])
    X = m["X"].reshape(3, -1)
    u = ((Gm - np.eye(3)) @ X).ravel()
    f = O.OracleProblem(m, n_steps=1).internal_force(u).reshape(3, -1)
    interior = [n for n in range(m["n_nodes"]) if np.all((X[:, n] > 1e-9) & (X[:, n] < 1 - 1e-9))]
    assert len(interior) == 8
    assert np.abs(f[:, interior]).max() <= 1e-10 * np.abs(f).max()  # homogeneous stress: no interior force


# ------------------------------------------------------------------ stable dt

def test_critical_dt_cube_and_compression():
    m = box(4, 4, 4)
    c = L.orc_sound_speed(1e6, 0.3, 1000.0)
    p = O.OracleProblem(m, n_steps=1, scale=0.5)
    assert p.critical_dt() == pytest.approx(0.5 * 0.25 / c, rel=1e-15)
    X = m["X"].reshape(3, -1)
    u = np.zeros((3, m["n_nodes"]))
    u[2] = -0.1 * X[2]  # uniform 10% compression in z
    hx, hz = 0.25, 0.225
    l_def = 2 * 3 * (hx * hx * hz) / (2 * hx * hx + 4 * hx * hz)
    assert p.critical_dt(u.ravel()) == pytest.approx(0.5 * l_def / c, rel=1e-14)
    u[2] = +0.1 * X[2]  # stretching cannot raise dt above the undeformed value: min(l0, l)
    assert p.critical_dt(u.ravel()) == pytest.approx(0.5 * 0.25 / c, rel=1e-15)


# ------------------------------------------------------------------ integrator and quirks

def test_rigid_translation_integrates_prescribed_velocity_exactly():
    m = box(2, 2, 2)
    N = m["n_nodes"]
    dof = np.arange(N, dtype=np.int32)  # every node prescribed in x, velocity ramp over 10 steps
    val = np.full(N, 0.3)
    p = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_VELOCITY, vel_ramp_steps=10, n_steps=25,
                        energy_every=1)
    r = p.run(hist_cap=25)
    dts = r["hist"]["dt"]
    V = 0.3 * np.minimum(np.arange(1, 26) / 10, 1.0)
    ux = np.cumsum(V * dts)[-1]
    np.testing.assert_allclose(r["u"][:N], ux, rtol=1e-14)
    assert np.abs(r["u"][N:]).max() <= 1e-14 * ux  # stress-free: only round-off moves y, z
    np.testing.assert_allclose(r["v"][:N], 0.3)
    # uniform velocity: E_kin = rho V v^2 / 2 ; undeformed E_def = 1.5 mu V ; VOLUME = V
    _, mu, _ = lame(1e6, 0.3)
    assert r["hist"]["e_kin"][-1] == pytest.approx(0.5 * 1000.0 * 0.3 ** 2, rel=1e-13)
    assert r["hist"]["e_def"][-1] == pytest.approx(1.5 * mu, rel=1e-13)
    assert r["hist"]["volume"][-1] == pytest.approx(1.0, rel=1e-13)


def test_single_dof_central_difference_recurrence():
    m = box(1, 1, 1)
    N = m["n_nodes"]
    free = 6 + 2 * N  # node 6 (the +,+,+ corner), z
    dofs = np.array([d for d in range(3 * N) if d != free], np.int32)
    p = O.OracleProblem(m, dir_dof=dofs, dir_val=np.zeros(dofs.size), bc_mode=O.BC_VELOCITY, n_steps=60,
                        scale=0.5)
    v0 = np.zeros(3 * N)
    v0[free] = 1e-9
    r = p.run(v0=v0, hist_cap=60)
    mass = p.mass()[6]
    du = 1e-10
    e = np.zeros(3 * N)
    e[free] = du
    k = (p.internal_force(e)[free] - p.internal_force(-e)[free]) / (2 * du)
    u, v, a = 0.0, 1e-9, 0.0  # a0 = 0 (solverExplicit.cpp:143)
    for dt in r["hist"]["dt"]:
        u = u + dt * v + 0.5 * dt * dt * a
        an = -k * u / mass
        v = v + dt * (0.5 * a + 0.5 * an)
        a = an
    assert r["u"][free] == pytest.approx(u, rel=1e-6)
    assert r["v"][free] == pytest.approx(v, rel=1e-6)


def test_ramp_quirks_lagged_time_and_window():
    m = box(2, 2, 2)
    dd, dv = clamp_and_pull(m, vel=0.01)  # ramp: total U = 0.01 over [t0, tf]
    t0, tf = 0.0, 0.02
    p = O.OracleProblem(m, dir_dof=dd, dir_val=dv, bc_mode=O.BC_RAMP, t0=t0, tf=tf, scale=0.5, n_steps=0)
    r = p.run(hist_cap=10000)
    n = r["steps"]
    t, dt = r["hist"]["time"][:n], r["hist"]["dt"][:n]
    dt0 = p.critical_dt()
    lag = np.concatenate([[dt0], dt[:-1]])  # time advances by the previous dt (solverExplicit.cpp:227-232)
    np.testing.assert_allclose(t[:-1], t0 + np.cumsum(lag)[:-1], rtol=1e-13)
    assert t[-1] == tf and t[-2] < tf - 1e-12  # last time clamped to tf, window terminates
    top = dd[-1]
    UU = 0.01 / (tf - t0)
    assert r["v"][top] == pytest.approx(UU * dt[-1], rel=1e-14)  # v := UU*dt on ramp DOFs
    assert r["u"][top] == pytest.approx(UU * dt.sum(), rel=1e-12)


def test_instantaneous_bc_applied_once():
    m = box(2, 2, 2)
    dd, dv = clamp_and_pull(m, vel=0.001)
    p = O.OracleProblem(m, dir_dof=dd, dir_val=dv, bc_mode=O.BC_INST, n_steps=1)
    r1 = p.run()
    top = dd[-1]
    assert r1["u"][top] == 0.001 and r1["v"][top] == 0.001
    p5 = O.OracleProblem(m, dir_dof=dd, dir_val=dv, bc_mode=O.BC_INST, n_steps=5)
    r5 = p5.run()
    assert r5["u"][top] == 0.001 and r5["v"][top] == 0.0


def test_free_vibration_conserves_energy():
    # No BCs, smooth initial velocity: the central-difference scheme with lumped mass M keeps
    # 1/2 v^T M v + W(u) bounded to O((omega dt)^2); W(u) - W(0) from DEFO_ENERGY (W(0) = 1.5 mu V).
    m = box(6, 6, 6)
    X = m["X"].reshape(3, -1)
    v0 = np.zeros((3, m["n_nodes"]))
    v0[0] = 0.01 * np.sin(np.pi * X[2])
    v0[0] -= v0[0].mean()
    p = O.OracleProblem(m, n_steps=300, scale=0.5, energy_every=1)
    r = p.run(v0=v0.ravel(), hist_cap=300)
    mass = np.tile(p.mass(), 3)
    _, mu, _ = lame(1e6, 0.3)
    t0 = 0.5 * np.sum(mass * v0.ravel() ** 2)
    t1 = 0.5 * np.sum(mass * r["v"] ** 2)
    w1 = r["hist"]["e_def"][-1] - 1.5 * mu * 1.0
    assert abs(t1 + w1 - t0) <= 0.05 * t0
    ek = r["hist"]["e_kin"]
    assert ek.min() < 0.5 * ek.max()  # kinetic and strain energy are actually exchanged


def test_summation_orders_and_threads_agree():
    m = perturb(box(5, 4, 3), 0.1, 12)
    dd, dv = clamp_and_pull(m)
    runs = []
    for so, nt in ((O.SUM_REFERENCE, 1), (O.SUM_ELEMENT, 1), (O.SUM_ELEMENT, 4)):
        runs.append(O.OracleProblem(m, dir_dof=dd, dir_val=dv, n_steps=200, sum_order=so, n_threads=nt).run()["u"])
    assert np.abs(runs[0] - runs[1]).max() <= 1e-13 * np.abs(runs[0]).max()
    assert np.array_equal(runs[1], runs[2])
    # threads need the element-order sums: the reference order with threads is rejected, not run
    bad = O.OracleProblem(m, dir_dof=dd, dir_val=dv, n_steps=5, sum_order=O.SUM_REFERENCE, n_threads=4).run()
    assert bad["rc"] == -3 and bad["steps"] == 0


@pytest.mark.parametrize("law", [0, 1, 2, 3])
def test_element_outputs_homogeneous_deformation(law):
    """u = (G - I) X gives F = G at every GP, so each element's averages are the closed forms at G:
    Cauchy J^-1 P G^T (Belytschko: sigma), Green-Lagrange E (small strain for the linear law), von Mises,
    tr E, 0.6666 sqrt(1.5 sum E_ii^2 + 3 sum_{i<j} E_ij^2); quantities a law does not set are 0."""
    m = perturb(box(3, 2, 2), 0.15, 5)
    G = np.array([[1.03, 0.02, -0.01], [0.015, 0.98, 0.025], [-0.02, 0.01, 1.05]])
    X = m["X"].reshape(3, -1)
    u = ((G - np.eye(3)) @ X).ravel()
    q = O.OracleProblem(m, law=law, n_steps=1).element_outputs(u)
    J = np.linalg.det(G)
    P = pk1(law, G)
    cau = P @ G.T / J
    if law == 3:
        E = 0.5 * (G + G.T) - np.eye(3)
    else:
        E = 0.5 * (G.T @ G - np.eye(3))
    vms = np.sqrt(0.5 * ((cau[0, 0] - cau[1, 1]) ** 2 + (cau[1, 1] - cau[2, 2]) ** 2 + (cau[2, 2] - cau[0, 0]) ** 2
                         + 6 * (cau[0, 1] ** 2 + cau[0, 2] ** 2 + cau[1, 2] ** 2)))
    exp = np.zeros(21)
    exp[:9] = cau.ravel()
    if law != 0:
        exp[9:18] = E.ravel()
    exp[18] = vms
    if law in (1, 2):
        exp[19] = np.trace(E)
        exp[20] = 0.6666 * np.sqrt(1.5 * np.sum(np.diag(E) ** 2) + 3 * (E[0, 1] ** 2 + E[0, 2] ** 2 + E[1, 2] ** 2))
    for row in q:
        np.testing.assert_allclose(row, exp, rtol=1e-9, atol=1e-9 * np.abs(exp).max())


# ------------------------------------------------------------------ Neumann loads

def test_gravity_load_totals_and_lumped_mass():
    """BCsGravity (boundaryConditions.cpp:1146-1172): f_z = -g sum_gp w rho |J| phi_a, i.e. -g times the row-sum
    lumped mass node by node, and -g rho V in total."""
    m = perturb(box(4, 3, 3), 0.1, 3)
    for nen_tet in (False, True):
        mm = perturb(box(3, 3, 2, tet=True), 0.1, 4) if nen_tet else m
        p = O.OracleProblem(mm, rho=1200.0, n_steps=1, gravity=9.81)
        f = p.external_force(0.37)
        N = mm["n_nodes"]
        assert np.all(f[:2 * N] == 0.0)
        mass = p.mass()
        np.testing.assert_allclose(f[2 * N:], -9.81 * mass, rtol=1e-12)
        vol = 1.0 if not nen_tet else 1.0  # unit box, Sum wJ = V for the affine split as well
        assert f[2 * N:].sum() == pytest.approx(-9.81 * 1200.0 * vol * (6 * 0.166666666667 if nen_tet else 1.0),
                                                rel=1e-11)


def test_nodal_force_window_and_ramp():
    """classNeumann::getTracAmount (classNodes.h:101-108): F/(tf-t0) * t inside [t0, tf] -- t, not t - t0 --
    and 0 outside, also after tf; entries on one node add up."""
    m = box(2, 2, 2)
    N = m["n_nodes"]
    forces = ([5, 5, 9], np.array([[1.0, -2.0, 3.0], [0.5, 0.0, 0.0], [0.0, 0.0, 7.0]]), [0.2, 0.0, 0.1], [0.6, 1.0, 0.3])
    p = O.OracleProblem(m, n_steps=1, forces=forces)
    for t in (0.0, 0.1, 0.2, 0.25, 0.3, 0.6, 0.61, 1.0, 1.2):
        f = p.external_force(t)
        exp = np.zeros(3 * N)
        for node, vec, t0, tf in zip(*forces):
            if t0 <= t <= tf:
                for d in range(3):
                    exp[node + N * d] += vec[d] / (tf - t0) * t
        np.testing.assert_allclose(f, exp, rtol=1e-15, atol=0)


def test_load_order_gravity_position():
    """NeumannBCManagement (boundaryConditions.cpp:75-140) adds the loads into one vector in input order,
    gravity GP by GP onto what is there: gravity first leaves f_z = (gravity alone) + F exactly; gravity
    after the force differs only by the rounding of the order."""
    m = perturb(box(3, 2, 2), 0.1, 5)
    N = m["n_nodes"]
    forces = ([7, 11], np.array([[0.0, 1.0, 3.3], [2.0, 0.0, -1.7]]), [0.0, 0.0], [1.0, 2.0])
    t = 0.7
    grav = O.OracleProblem(m, rho=900.0, n_steps=1, gravity=9.81).external_force(t)
    frc = O.OracleProblem(m, n_steps=1, forces=forces).external_force(t)
    first = O.OracleProblem(m, rho=900.0, n_steps=1, gravity=9.81, forces=forces, gravity_pos=0).external_force(t)
    last = O.OracleProblem(m, rho=900.0, n_steps=1, gravity=9.81, forces=forces, gravity_pos=-1).external_force(t)
    middle = O.OracleProblem(m, rho=900.0, n_steps=1, gravity=9.81, forces=forces, gravity_pos=1).external_force(t)
    assert np.array_equal(first, grav + frc)
    np.testing.assert_allclose(last, grav + frc, rtol=1e-15, atol=0)
    np.testing.assert_allclose(middle, grav + frc, rtol=1e-15, atol=0)
    # with gravity after the force, node 7 accumulates F_z and then the GP contributions one by one
    assert last[7 + 2 * N] == pytest.approx(grav[7 + 2 * N] + 3.3 * t, rel=1e-15)
    loaded = {7, 11}
    for n in range(N):
        if n not in loaded:
            assert last[n + 2 * N] == grav[n + 2 * N] == middle[n + 2 * N]


def test_free_fall_and_momentum_under_nodal_forces():
    """A free body under gravity: every node accelerates at exactly -g, so u_z follows the scalar
    central-difference recurrence with a_0 = 0 (solverExplicit.cpp:143) -- -g dt^2 n(n-1)/2 at constant dt,
    not -g T^2/2 -- and x, y stay at rest; under nodal forces the momentum follows the trapezoid sum
    sum_k dt_k (F(t_{k-1}) + F(t_k)) / 2 with F(t_0) = 0."""
    m = perturb(box(3, 3, 3), 0.1, 8)
    N = m["n_nodes"]
    p = O.OracleProblem(m, n_steps=40, bc_mode=O.BC_VELOCITY, gravity=9.81)
    r = p.run(hist_cap=40)
    uz, vz, az = 0.0, 0.0, 0.0
    for dt in r["hist"]["dt"][:40]:
        uz += dt * vz + 0.5 * dt * dt * az
        vz += dt * 0.5 * (az + (-9.81))
        az = -9.81
    np.testing.assert_allclose(r["u"][2 * N:], uz, rtol=1e-10)
    np.testing.assert_allclose(r["v"][2 * N:], vz, rtol=1e-10)
    assert np.abs(r["u"][:2 * N]).max() < 1e-12 * abs(uz)
    forces = ([0, 13, 40], np.array([[1e3, 0.0, 0.0], [0.0, -2e3, 5e2], [3e2, 3e2, 3e2]]), [0.0, 0.0, 0.0],
              [1.0, 1.0, 1.0])
    p = O.OracleProblem(m, n_steps=30, bc_mode=O.BC_VELOCITY, forces=forces)
    r = p.run(hist_cap=30)
    mass = p.mass()
    mom = np.array([(mass * r["v"][d * N:(d + 1) * N]).sum() for d in range(3)])
    Ftot = lambda t: sum(vec * t for vec in forces[1])  # all windows [0, 1] contain the 30 steps
    imp = np.zeros(3)
    tprev = 0.0
    for k in range(30):
        t, dt = r["hist"]["time"][k], r["hist"]["dt"][k]
        imp += dt * 0.5 * ((Ftot(tprev) if k > 0 else 0.0) + Ftot(t))
        tprev = t
    np.testing.assert_allclose(mom, imp, rtol=1e-9)


# ---------------------------------------------------------------------------
# Pressure loads (BCsPressureRamp / BCsPressureInst, boundaryConditions.cpp:370-496)
# ---------------------------------------------------------------------------

def _boundary_faces(m):
    """(element, face S) of every exterior face of a box mesh, from the getSubElement node lists."""
    nen = m["nen"]
    conn = m["conn"].reshape(-1, nen)
    faces = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]] if nen == 8 else \
        [[0, 1, 2], [0, 3, 1], [1, 3, 2], [0, 2, 3]]
    count = {}
    for e in range(conn.shape[0]):
        for S, f in enumerate(faces):
            key = tuple(sorted(conn[e, f]))
            count.setdefault(key, []).append((e, S + 1))
    out = [v[0] for v in count.values() if len(v) == 1]
    return np.array([e for e, _ in out], np.int32), np.array([s for _, s in out], np.int32)


def test_pressure_face_geometry():
    """Surface GPs: getn0 is the unit normal pointing into the element; sum w|J_s| is the face area; the face
    shape is a partition of unity that reproduces the GP position; for a box the bulk shape at the point
    (defineShapeOnAPoint's corner-parallelepiped inverse map) equals the face shape on the face's nodes."""
    for tet in (False, True):
        m = box(2, 1, 1, lx=2.0, ly=1.5, lz=0.8, tet=tet) if not tet else perturb(box(2, 2, 2, tet=True), 0.1, 3)
        p = O.OracleProblem(m, n_steps=1)
        nen = m["nen"]
        X = m["X"].reshape(3, -1)
        conn = m["conn"].reshape(-1, nen)
        faces = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]] if not tet else \
            [[0, 1, 2], [0, 3, 1], [1, 3, 2], [0, 2, 3]]
        for e in range(0, conn.shape[0], 3):
            xe = X[:, conn[e]].T
            for S in range(1, len(faces) + 1):
                g = p.pressure_face(e, S)
                xf = xe[faces[S - 1]]
                nf = len(faces[S - 1])
                if nf == 4:
                    area = 0.5 * np.linalg.norm(np.cross(xf[2] - xf[0], xf[3] - xf[1]))
                else:
                    area = 0.5 * np.linalg.norm(np.cross(xf[1] - xf[0], xf[2] - xf[0]))
                assert g["ws"].sum() == pytest.approx(area, rel=1e-8 if nf == 4 else 1e-12)
                inward = xe.mean(axis=0) - xf.mean(axis=0)
                for k in range(len(g["ws"])):
                    n0 = g["n0"][k]
                    assert np.linalg.norm(n0) == pytest.approx(1.0, abs=1e-15)
                    assert n0 @ inward > 0
                    assert abs(n0 @ (xf[1] - xf[0])) < 1e-12 and abs(n0 @ (xf[2] - xf[0])) < 1e-12
                    ph = g["phi_s"][k][:nf]
                    assert ph.sum() == pytest.approx(1.0, abs=1e-15)
                    np.testing.assert_allclose(ph @ xf, g["xyz"][k], rtol=0, atol=1e-14)
                    if not tet:  # box: exact inverse map
                        pb = g["phi_b"][k]
                        np.testing.assert_allclose(pb[faces[S - 1]], ph, rtol=0, atol=1e-8)
                        np.testing.assert_allclose(np.delete(pb, faces[S - 1]), 0.0, atol=1e-8)
                        np.testing.assert_allclose(pb @ xe, g["xyz"][k], rtol=0, atol=1e-8)
                    else:  # the tet's shape at the natural origin: all weight on node 0
                        assert np.array_equal(g["phi_b"][k][:4], [1.0, 0.0, 0.0, 0.0])
                    # bulk gradients reproduce the element geometry: sum_a x_a (x) dphi_a = I
                    np.testing.assert_allclose(xe.T @ g["dphi_b"][k][:nen], np.eye(3), rtol=0, atol=1e-8)


@pytest.mark.parametrize("tet", [False, True])
def test_pressure_undeformed_totals_and_time_function(tet):
    """At u = 0, a pressure block on faces f gives sum_f P(t) A_f n0_f in total (inward normals: a positive P
    pushes into the body); P(t) = P/(tf-t0) (t-t0) inside the window for RAMP, P for INST, 0 outside.  INST
    spreads each face's force by the face shape (A/4 per node on a flat quad, A/3 on a triangle); RAMP uses
    the bulk shape at the point -- the same on a box, but for tets all on the element's first node."""
    m = box(2, 2, 2, tet=tet)
    N = m["n_nodes"]
    el, fc = _boundary_faces(m)
    top = [i for i in range(el.size) if p_face_z(m, el[i], fc[i]) == 1.0]
    el, fc = el[top], fc[top]
    conn = m["conn"].reshape(-1, m["nen"])
    P, t0, tf = 2.5e3, 0.2, 0.6
    for kind in ("pressure_ramp", "pressure_inst"):
        p = O.OracleProblem(m, n_steps=1, loads=[(kind, P, t0, tf, el, fc)])
        for t in (0.1, 0.2, 0.35, 0.6, 0.7):
            f = p.external_force(t).reshape(3, N)
            Pt = (P / (tf - t0) * (t - t0) if kind == "pressure_ramp" else P) if t0 <= t <= tf else 0.0
            np.testing.assert_allclose(f.sum(axis=1), [0, 0, -Pt * 1.0], rtol=1e-12, atol=1e-12 * P)
            if Pt == 0.0:
                assert np.all(f == 0.0)
                continue
            loaded = np.nonzero(np.abs(f[2]) > 1e-9 * Pt)[0]  # the inverse map leaves rounding-level weights
            if kind == "pressure_inst" or not tet:
                X = m["X"].reshape(3, -1)
                assert set(loaded) == set(np.nonzero(np.abs(X[2] - 1.0) < 1e-12)[0])
                if not tet:  # quads of area 1/4: corner nodes get Pt/16, edge nodes 2x, interior 4x
                    counts = np.round(-f[2, loaded] / (Pt * 0.25 / 4)).astype(int)
                    assert sorted(set(counts)) == [1, 2, 4]
                    np.testing.assert_allclose(-f[2, loaded], counts * Pt / 16, rtol=1e-8)
            else:
                assert set(loaded) <= set(conn[el, 0])


def p_face_z(m, e, S):
    """z of the face's first node when the face is horizontal, else None."""
    nen = m["nen"]
    faces = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]] if nen == 8 else \
        [[0, 1, 2], [0, 3, 1], [1, 3, 2], [0, 2, 3]]
    z = m["X"].reshape(3, -1)[2][m["conn"].reshape(-1, nen)[e, faces[S - 1]]]
    return z[0] if np.ptp(z) < 1e-12 else None


@pytest.mark.parametrize("tet", [False, True])
def test_pressure_follower_load_nanson(tet):
    """Under a homogeneous deformation u = (G - I) X the bulk F at every surface point is G, so each face carries
    P det(G) G^-T n0 A (Nanson's formula) and a closed surface carries none; a rigid rotation turns the load
    with the body."""
    m = perturb(box(2, 2, 2, tet=tet), 0.08, 9)
    X = m["X"].reshape(3, -1)
    N = m["n_nodes"]
    el, fc = _boundary_faces(m)
    G = np.array([[1.1, 0.05, -0.02], [0.03, 0.92, 0.04], [-0.01, 0.06, 1.05]])
    u = ((G - np.eye(3)) @ X).ravel()
    P = 1.7e3
    p = O.OracleProblem(m, n_steps=1, loads=[("pressure_inst", P, 0.0, 1.0, el, fc)])
    f = p.external_force(0.5, u).reshape(3, N)
    np.testing.assert_allclose(f.sum(axis=1), 0.0, atol=1e-10 * P)
    # one face: compare with the closed form
    e, S = int(el[0]), int(fc[0])
    g = p.pressure_face(e, S)
    one = O.OracleProblem(m, n_steps=1, loads=[("pressure_inst", P, 0.0, 1.0, [e], [S])])
    f1 = one.external_force(0.5, u).reshape(3, N)
    cof = np.linalg.det(G) * np.linalg.inv(G).T
    expect = sum(P * w * (cof @ n0) for w, n0 in zip(g["ws"], g["n0"]))
    np.testing.assert_allclose(f1.sum(axis=1), expect, rtol=1e-10)
    # rigid rotation about z by 30 degrees: the face force rotates with the body
    c, s_ = np.cos(np.pi / 6), np.sin(np.pi / 6)
    R = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1.0]])
    ur = ((R - np.eye(3)) @ X).ravel()
    fr = one.external_force(0.5, ur).reshape(3, N)
    f0 = one.external_force(0.5).reshape(3, N)
    np.testing.assert_allclose(fr.sum(axis=1), R @ f0.sum(axis=1), rtol=1e-10, atol=1e-10 * P)


def test_pressure_blocks_in_input_order_and_dynamics():
    """Blocks add in input order (a pressure block before or after a force block on the same node changes only
    the rounding), and a cube whose faces all carry the same inward pressure contracts symmetrically: the
    centre node stays at rest and the volume shrinks."""
    m = box(2, 2, 2)
    N = m["n_nodes"]
    el, fc = _boundary_faces(m)
    pr = ("pressure_inst", 1e4, 0.0, 1.0, el, fc)
    fo = ("force", [26], [[0.0, 0.0, 5.0]], 0.0, 1.0)
    a = O.OracleProblem(m, n_steps=1, loads=[pr, fo]).external_force(0.3)
    b = O.OracleProblem(m, n_steps=1, loads=[fo, pr]).external_force(0.3)
    np.testing.assert_allclose(a, b, rtol=1e-15, atol=1e-12)
    r = O.OracleProblem(m, n_steps=40, bc_mode=O.BC_VELOCITY, loads=[pr], energy_every=40).run(hist_cap=40)
    u = r["u"].reshape(3, N)
    X = m["X"].reshape(3, -1)
    centre = int(np.nonzero(np.all(np.abs(X - 0.5) < 1e-12, axis=0))[0][0])
    assert np.abs(u[:, centre]).max() < 1e-12 * np.abs(u).max()
    assert r["hist"]["volume"][39] < 1.0
    # inward motion of every corner, symmetric in x, y, z
    corners = np.nonzero(np.all((np.abs(X) < 1e-12) | (np.abs(X - 1) < 1e-12), axis=0))[0]
    d = (u[:, corners] * np.sign(0.5 - X[:, corners]))
    assert np.all(d > 0)
    np.testing.assert_allclose(d, d[0, 0], rtol=1e-9)


# ---------------------------------------------------------------------------
# Artificial viscosity (solverExplicit.cpp:381-421, 469-542)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("eps", [1e-3, -2e-3])
def test_bulk_viscosity_closed_form_unit_cube(eps):
    """A unit cube stretched uniformly from F0 = I to F1 = (1+eps) I in one step dt: trD = 3 eps/dt,
    q = rho/J h (c1^2 h |trD| min(0, trD) + cl c0 trD), P_vis = J q F^-T, and the viscous nodal force is
    P_vis int grad phi_a = J q/(1+eps) s_a/4 (s_a the corner's outward signs); expansion only uses the linear
    term, compression both."""
    m = box(1, 1, 1)
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    rho, dt, c1, cl = 1000.0, 1e-4, 0.3, 1.5
    p0 = O.OracleProblem(m, rho=rho, n_steps=1)
    pv = O.OracleProblem(m, rho=rho, n_steps=1, visc=O.VISC_BULK, visc_c1=c1, visc_cl=cl)
    u = (eps * X).ravel()
    fv = pv.internal_force_visc(u, np.zeros(3 * N), dt) - p0.internal_force(u)
    c0 = O.lib().orc_sound_speed(1e6, 0.3, rho)
    h = 1.0  # the cube's undeformed characteristic length
    J = (1 + eps) ** 3
    trD = 3 * eps / dt
    q = rho / J * h * (c1 * c1 * h * abs(trD) * min(0.0, trD) + cl * c0 * trD)
    s = np.sign(X - 0.5)
    np.testing.assert_allclose(fv.reshape(3, N), J * q / (1 + eps) * s / 4, rtol=1e-9, atol=1e-9 * abs(q))
    # a general homogeneous step F0 = I -> F1 = G: P_vis = det(G) q G^-T
    G = np.array([[1.0 + eps, 0.3 * eps, -0.2 * eps], [0.1 * eps, 1.0 - 0.5 * eps, 0.4 * eps],
                  [-0.3 * eps, 0.2 * eps, 1.0 + 2 * eps]])
    ug = ((G - np.eye(3)) @ X).ravel()
    fg = pv.internal_force_visc(ug, np.zeros(3 * N), dt) - p0.internal_force(ug)
    trG = np.trace(G - np.eye(3)) / dt
    JG = np.linalg.det(G)
    qg = rho / JG * h * (c1 * c1 * h * abs(trG) * min(0.0, trG) + cl * c0 * trG)
    Pg = JG * qg * np.linalg.inv(G).T
    np.testing.assert_allclose(fg.reshape(3, N), Pg @ s / 4, rtol=1e-8, atol=1e-8 * np.abs(Pg).max())
    # no rate, no viscous force
    assert np.array_equal(pv.internal_force_visc(u, u, dt), p0.internal_force(u))
    # explicit element size and sound speed replace the element's own
    pw = O.OracleProblem(m, rho=rho, n_steps=1, visc=O.VISC_BULK, visc_c1=c1, visc_cl=cl, visc_h=0.5, visc_c0=7.0)
    fw = pw.internal_force_visc(u, np.zeros(3 * N), dt) - p0.internal_force(u)
    q2 = rho / J * 0.5 * (c1 * c1 * 0.5 * abs(trD) * min(0.0, trD) + cl * 7.0 * trD)
    np.testing.assert_allclose(fw.reshape(3, N), J * q2 / (1 + eps) * s / 4, rtol=1e-9, atol=1e-9 * abs(q2))


def test_viscosity_dissipates_and_dev_variant_adds_nothing():
    """Free vibration from an initial velocity field: without viscosity E_def + E_kin keeps its level up to the
    scheme's per-step wobble; with bulk viscosity it loses a good part of the initial kinetic energy; the
    deviatoric variant leaves every bit unchanged (its stress never reaches the PK1).  (The reference's
    defaults c1 = 0.06, cl = 1.2 are unstable at a scale factor of 0.5 once h and c0 are the element's: the
    linear term alone needs dt <= (sqrt(1 + cl^2) - cl) h/c.)"""
    m = perturb(box(3, 3, 3), 0.1, 31)
    N = m["n_nodes"]
    X = m["X"].reshape(3, N)
    v0 = np.concatenate([0.02 * (X[0] - 0.5), -0.01 * (X[1] - 0.5), 0.03 * np.sin(np.pi * X[2])])
    runs = {}
    for visc in (O.VISC_NONE, O.VISC_BULK, O.VISC_DEV):
        p = O.OracleProblem(m, n_steps=300, bc_mode=O.BC_VELOCITY, energy_every=1, visc=visc, visc_c1=0.5,
                            visc_cl=0.06)
        runs[visc] = p.run(v0=v0, hist_cap=300)
        assert runs[visc]["rc"] == 0
    for k in ("u", "v", "a"):
        assert np.array_equal(runs[O.VISC_DEV][k], runs[O.VISC_NONE][k])
    E = {k: r["hist"]["e_def"] + r["hist"]["e_kin"] for k, r in runs.items()}
    ke0 = runs[O.VISC_NONE]["hist"]["e_kin"][0]
    assert abs(E[O.VISC_NONE][-50:].mean() - E[O.VISC_NONE][0]) < 0.1 * ke0
    assert E[O.VISC_BULK][-50:].mean() - E[O.VISC_BULK][0] < -0.25 * ke0


@pytest.mark.parametrize("tet", [False, True])
def test_traction_dead_load_totals_and_time_function(tet):
    """BCsTractionRamp / BCsTractionInst (boundaryConditions.cpp:218-271): a dead load -- the total on faces of
    area A is T(t) A whatever the displacement, spread by the face shape; T(t) = T/(tf-t0) (t-t0) for RAMP,
    T for INST inside the window, 0 outside."""
    m = perturb(box(2, 2, 3, tet=tet), 0.1, 12)
    N = m["n_nodes"]
    el, fc = _boundary_faces(m)
    top = [i for i in range(el.size) if p_face_z(m, el[i], fc[i]) == 1.0]
    el, fc = el[top], fc[top]
    T, t0, tf = np.array([3.0e2, -1.5e2, 7.0e2]), 0.1, 0.5
    u = np.random.default_rng(2).uniform(-0.05, 0.05, 3 * N)
    for kind in ("traction_ramp", "traction_inst"):
        p = O.OracleProblem(m, n_steps=1, loads=[(kind, T, t0, tf, el, fc)])
        for t in (0.05, 0.1, 0.3, 0.5, 0.6):
            Tt = (T / (tf - t0) * (t - t0) if kind == "traction_ramp" else T) if t0 <= t <= tf else 0.0 * T
            f0 = p.external_force(t).reshape(3, N)
            fu = p.external_force(t, u).reshape(3, N)
            assert np.array_equal(f0, fu)  # not a follower load
            np.testing.assert_allclose(f0.sum(axis=1), Tt * 1.0, rtol=1e-12, atol=1e-12 * np.abs(T).max())
            X = m["X"].reshape(3, -1)
            off = np.abs(X[2] - 1.0) > 1e-12
            assert np.all(f0[:, off] == 0.0)


# ------------------------------------------------------------------ energies under non-uniform fields

@pytest.mark.parametrize("n", [1, 2, 3])
def test_kinetic_energy_is_the_consistent_interpolation(n):
    """KIN_ENERGY interpolates v to each GP (sum_a phi_a v_a, dataExtraction.cpp:655-668).  For the linear field
    v = alpha X on the unit cube that is 1/2 rho alpha^2 int |X|^2 dV = 1/2 rho alpha^2 (2x2x2 Gauss integrates the
    quadratic exactly up to the 0.577350269 literal, ~1e-10).  Lumping the mass instead gives
    1/2 sum_a m_a |v_a|^2, the trapezoidal rule of |X|^2, i.e. (1 + 1/(2 n^2)) times that value on n^3 elements
    (1.5x on one element): a lumped-KE slip fails."""
    m = box(n, n, n)
    N = m["n_nodes"]
    X = m["X"]
    alpha, rho = 0.7, 1000.0
    v = alpha * X
    p = O.OracleProblem(m, rho=rho, n_steps=1)
    _, ek, vol = p.energies(np.zeros(3 * N), v)
    exact = 0.5 * rho * alpha ** 2 * 1.0
    assert ek == pytest.approx(exact, rel=1e-9)
    assert vol == pytest.approx(1.0, rel=1e-13)
    mass = p.mass()
    lumped = 0.5 * float(np.sum(mass * (v.reshape(3, N) ** 2).sum(axis=0)))
    assert lumped == pytest.approx(exact * (1 + 0.5 / n ** 2), rel=1e-9)
    assert abs(ek - lumped) > 0.4 / n ** 2 * exact
    # a field with a constant part and a gradient along one axis only: the cross term 2 c.grad integrates too
    c = np.array([0.3, -0.2, 0.1])
    v2 = (c[:, None] + np.array([0.0, 0.0, 1.5])[:, None] * X.reshape(3, N)[0][None, :]).ravel()
    _, ek2, _ = p.energies(np.zeros(3 * N), v2)
    # int |c + (0,0,1.5 x)|^2 = |c|^2 + 2 c_z 1.5 / 2 + 1.5^2 / 3
    assert ek2 == pytest.approx(0.5 * rho * (c @ c + 2 * c[2] * 1.5 * 0.5 + 1.5 ** 2 / 3), rel=1e-9)


@pytest.mark.parametrize("tet", [False, True])
def test_volume_is_det_G_under_homogeneous_deformation(tet):
    """VOLUME = sum_gp wJ det F; under u = (G - I) X, F = G at every GP, so the deformed volume is det(G) V
    (V = 1; for tets the weight literal 0.166666666667 makes it (1 + 2e-12) det G).  A volume formed from
    the undeformed Jacobian alone, or from det(F) - 1, fails."""
    m = perturb(box(3, 2, 2, tet=tet), 0.1, 23)
    N = m["n_nodes"]
    G = np.array([[1.1, 0.05, -0.02], [0.03, 0.93, 0.04], [-0.01, 0.02, 1.07]])
    X = m["X"].reshape(3, N)
    u = ((G - np.eye(3)) @ X).ravel()
    p = O.OracleProblem(m, n_steps=1)
    ed, ek, vol = p.energies(u, np.zeros(3 * N))
    w = 6 * 0.166666666667 if tet else 1.0
    assert vol == pytest.approx(np.linalg.det(G) * w, rel=1e-13)
    assert ek == 0.0
    # compressible NH energy density at F = G is uniform: E_def = W(G) V
    lam, mu, _ = lame(1e6, 0.3)
    J = np.linalg.det(G)
    W = 0.5 * lam * np.log(J) ** 2 - mu * np.log(J) + 0.5 * mu * np.trace(G.T @ G)
    assert ed == pytest.approx(W * w, rel=1e-12)


# ------------------------------------------------------------------ energy balance (build-defined)

def test_work_and_energies_in_free_fall():
    """Free fall under gravity, nothing constrained: the body translates rigidly, so EXTERNAL_ENERGY = u . fext
    = -M g u_z, the lumped kinetic energy is 1/2 M v_z^2, and the trapezoidal work of the constant weight from
    the end of step 1 is exactly -M g (u_z(k) - u_z(1))."""
    m = perturb(box(3, 2, 2), 0.1, 41)
    N = m["n_nodes"]
    g, rho = 9.81, 1000.0
    p = O.OracleProblem(m, rho=rho, n_steps=60, energy_every=1, loads=[("gravity", g)], bc_mode=O.BC_RAMP)
    r = p.run(hist_cap=60)
    h = r["hist"]
    M = rho * 1.0  # unit-volume box
    uz = r["u"][2 * N:]
    assert np.ptp(uz) <= 1e-12 * np.abs(uz).max()  # rigid: every node falls alike
    assert h["e_ext"][-1] == pytest.approx(-M * g * uz[0], rel=1e-12)
    assert h["e_kin_lumped"][-1] == pytest.approx(0.5 * M * r["v"][2 * N] ** 2, rel=1e-12)
    # u_z after step 1 from the recurrence: a0 = 0, so step 1 moves nothing; then u_z(k) from the history
    dts = h["dt"]
    vz, az, u1 = 0.0, 0.0, None
    uzk = 0.0
    for k in range(60):
        uzk += dts[k] * vz + 0.5 * dts[k] * dts[k] * az
        vz, az = vz + dts[k] * 0.5 * (az - g), -g
        if k == 0:
            u1 = uzk
    assert uzk == pytest.approx(uz[0], rel=1e-12)
    assert h["w_ext"][0] == 0.0
    assert h["w_ext"][-1] == pytest.approx(-M * g * (uzk - u1), rel=1e-10)
    assert np.all(h["w_react"] == 0.0)


def test_reaction_work_carries_the_prescribed_momentum():
    """All DOFs prescribed in a rigid x-translation whose velocity ramps up: nothing deforms, so the support
    force fint - fext vanishes and the reaction work is the kinetic energy the prescribed velocity gives the
    lumped mass, 1/2 M (V_k^2 - V_1^2).  A reaction work without that inertia term stays 0 and fails."""
    m = box(2, 2, 2)
    N = m["n_nodes"]
    dof = np.arange(3 * N, dtype=np.int32)
    val = np.concatenate([np.full(N, 0.4), np.zeros(2 * N)])
    p = O.OracleProblem(m, rho=1000.0, dir_dof=dof, dir_val=val, bc_mode=O.BC_VELOCITY, vel_ramp_steps=20,
                        n_steps=30, energy_every=1)
    r = p.run(hist_cap=30)
    h = r["hist"]
    V = 0.4 * np.minimum(np.arange(1, 31) / 20, 1.0)
    M = 1000.0
    np.testing.assert_allclose(h["w_react"], 0.5 * M * (V ** 2 - V[0] ** 2), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(h["e_kin_lumped"], 0.5 * M * V ** 2, rtol=1e-12)
    assert np.all(h["w_ext"] == 0.0)


def test_energy_balance_closes_on_a_pulled_block():
    """Lumped kinetic + deformation energy against the work of the loads and supports, from the end of step 1:
    the explicit update conserves it up to its O(dt^2) error, here to 1e-5 of the work at the end and within
    3% at every step after the first few."""
    m = perturb(box(5, 5, 5), 0.1, 17)
    dof, val = clamp_and_pull(m, vel=0.1)
    p = O.OracleProblem(m, dir_dof=dof, dir_val=val, bc_mode=O.BC_VELOCITY, vel_ramp_steps=50, n_steps=400,
                        energy_every=1)
    h = p.run(hist_cap=400)["hist"]
    E = h["e_def"] + h["e_kin_lumped"]
    W = h["w_ext"] + h["w_react"]
    B = (E - E[0]) - W
    scale = np.maximum.accumulate(np.maximum(np.abs(W), np.abs(h["e_def"] - h["e_def"][0]) + h["e_kin_lumped"]))
    assert W[-1] > 10.0
    assert abs(B[-1]) <= 1e-5 * scale[-1]
    assert np.all(np.abs(B[5:]) <= 3e-2 * scale[5:])

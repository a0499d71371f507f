"""ctypes bindings for the CPU oracle (oracle/lib/libmps_oracle.so).

Test infrastructure only: the oracle is the checker for the HIP path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("MPS_ORACLE_LIB", os.path.join(ROOT, "oracle", "lib", "libmps_oracle.so"))

LAW_NH_COMPRESSIBLE = 0
LAW_NH_BELYTSCHKO = 1
LAW_SVK = 2
LAW_LINEAR = 3
BC_RAMP, BC_INST, BC_VELOCITY = 0, 1, 2
SUM_REFERENCE, SUM_ELEMENT = 0, 1
ORC_OK, ORC_ERR_DT, ORC_ERR_NAN = 0, -1, -2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Problem(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_elem", C.c_int32), ("nen", C.c_int32), ("law", C.c_int32),
        ("X", _dp), ("conn", _ip), ("mat_per_elem", C.c_int32),
        ("E", _dp), ("nu", _dp), ("rho", _dp),
        ("n_dir", C.c_int32), ("dir_dof", _ip), ("dir_val", _dp),
        ("bc_mode", C.c_int32), ("vel_ramp_steps", C.c_int32),
        ("t0", C.c_double), ("tf", C.c_double), ("scale", C.c_double),
        ("n_steps", C.c_int64), ("sum_order", C.c_int32), ("n_threads", C.c_int32),
        ("energy_every", C.c_int32), ("_pad", C.c_int32),
        ("n_loads", C.c_int32), ("_pad2", C.c_int32), ("loads", C.c_void_p),
        ("visc", C.c_int32), ("_pad3", C.c_int32), ("visc_c1", C.c_double), ("visc_cl", C.c_double),
        ("visc_h", C.c_double), ("visc_c0", C.c_double),
    ]


VISC_NONE, VISC_BULK, VISC_DEV = 0, 1, 2


class Load(C.Structure):
    _fields_ = [("type", C.c_int32), ("n", C.c_int32), ("t0", C.c_double), ("tf", C.c_double),
                ("value", C.c_double), ("index", _ip), ("face", _ip), ("force", _dp)]


LOAD_FORCE, LOAD_GRAVITY, LOAD_PRESSURE_RAMP, LOAD_PRESSURE_INST, LOAD_TRACTION_RAMP, LOAD_TRACTION_INST = 0, 1, 2, 3, 4, 5


def load_blocks(forces=None, gravity=0.0, gravity_pos=-1, loads=None):
    """Neumann blocks in input order, as tuples:
      ("force", nodes, vec[k,3], t0, tf)                 *BOUNDARY, TYPE=FORCE (entries F*Ff, one window)
      ("gravity", g)                                     *GRAVITY
      ("pressure_ramp" | "pressure_inst", P, t0, tf, elems, faces)
      ("traction_ramp" | "traction_inst", (Tx, Ty, Tz), t0, tf, elems, faces)
    `loads` is such a list; otherwise forces = (node[k], vec[k,3], t0[k], tf[k]) gives one block per entry and
    a non-zero gravity is placed before entry gravity_pos (-1: after all)."""
    if loads is not None:
        return list(loads)
    out = []
    if forces is not None:
        fn, fv, f0, f1 = forces
        fv = np.reshape(np.asarray(fv, dtype=np.float64), (-1, 3))
        out = [("force", [int(fn[k])], fv[k:k + 1], float(f0[k]), float(f1[k])) for k in range(len(fn))]
    if gravity != 0.0:
        pos = gravity_pos if 0 <= gravity_pos <= len(out) else len(out)
        out.insert(pos, ("gravity", float(gravity)))
    return out


def pack_loads(blocks):
    """ctypes array of Load for the blocks, plus the numpy buffers it points into (keep them alive)."""
    keep = []
    arr = (Load * max(len(blocks), 1))()
    for b, blk in enumerate(blocks):
        kind = blk[0]
        L = arr[b]
        if kind == "force":
            nodes = np.ascontiguousarray(blk[1], dtype=np.int32)
            vec = np.ascontiguousarray(np.reshape(blk[2], (-1, 3)), dtype=np.float64)
            keep += [nodes, vec]
            L.type, L.n, L.t0, L.tf = LOAD_FORCE, nodes.size, blk[3], blk[4]
            L.index, L.force = iptr(nodes), dptr(vec)
        elif kind == "gravity":
            L.type, L.n, L.value = LOAD_GRAVITY, 0, blk[1]
        elif kind in ("pressure_ramp", "pressure_inst", "traction_ramp", "traction_inst"):
            el = np.ascontiguousarray(blk[4], dtype=np.int32)
            fc = np.ascontiguousarray(blk[5], dtype=np.int32)
            keep += [el, fc]
            L.type = {"pressure_ramp": LOAD_PRESSURE_RAMP, "pressure_inst": LOAD_PRESSURE_INST,
                      "traction_ramp": LOAD_TRACTION_RAMP, "traction_inst": LOAD_TRACTION_INST}[kind]
            L.n, L.t0, L.tf = el.size, blk[2], blk[3]
            L.index, L.face = iptr(el), iptr(fc)
            if kind.startswith("traction"):
                T = np.ascontiguousarray(blk[1], dtype=np.float64).reshape(3)
                keep.append(T)
                L.force = dptr(T)
            else:
                L.value = blk[1]
        else:
            raise ValueError(kind)
    return arr, keep


class Report(C.Structure):
    _fields_ = [("steps", C.c_int64), ("time", C.c_double), ("dt_last", C.c_double),
                ("status", C.c_int32), ("_pad", C.c_int32)]


HIST_KEYS = ("time", "dt", "e_def", "e_kin", "volume", "e_ext", "e_kin_lumped", "w_ext", "w_react")


class History(C.Structure):
    _fields_ = [("cap", C.c_int64)] + [(k, _dp) for k in HIST_KEYS]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.orc_det3.restype = C.c_double
        L.orc_det3.argtypes = [_dp]
        L.orc_inverse3.argtypes = [_dp, _dp]
        L.orc_quadrature.restype = C.c_int32
        L.orc_quadrature.argtypes = [C.c_int32, _dp]
        L.orc_define_gp.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_area_triangle.restype = C.c_double
        L.orc_area_triangle.argtypes = [_dp]
        L.orc_char_length.restype = C.c_double
        L.orc_char_length.argtypes = [C.c_int32, _dp]
        L.orc_material.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_sound_speed.restype = C.c_double
        L.orc_sound_speed.argtypes = [C.c_double] * 3
        L.orc_law_sound_speed.restype = C.c_double
        L.orc_law_sound_speed.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double]
        L.orc_pk1.argtypes = [C.c_int32, _dp, C.c_double, C.c_double, _dp]
        L.orc_defo_energy.restype = C.c_double
        L.orc_defo_energy.argtypes = [C.c_int32, _dp, C.c_double, C.c_double]
        L.orc_kinematics.argtypes = [C.c_int32, _dp, _dp, _dp]
        L.orc_gp_force.argtypes = [C.c_int32, C.c_double, _dp, _dp, _dp]
        L.orc_gp_lumped_mass.argtypes = [C.c_int32, C.c_double, C.c_double, _dp, _dp]
        L.orc_run.restype = C.c_int32
        L.orc_run.argtypes = [C.POINTER(Problem), _dp, _dp, _dp, C.POINTER(History), C.POINTER(Report)]
        L.orc_lumped_mass.restype = C.c_int32
        L.orc_lumped_mass.argtypes = [C.POINTER(Problem), _dp]
        L.orc_internal_force.restype = C.c_int32
        L.orc_element_outputs.restype = C.c_int32
        L.orc_external_force.restype = C.c_int32
        L.orc_external_force.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.orc_internal_force_visc.restype = C.c_int32
        L.orc_internal_force_visc.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp]
        L.orc_pressure_face.restype = C.c_int32
        L.orc_pressure_face.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_internal_force.argtypes = [C.POINTER(Problem), _dp, _dp]
        L.orc_energies.restype = C.c_int32
        L.orc_energies.argtypes = [C.POINTER(Problem), _dp, _dp, _dp]
        L.orc_critical_dt.restype = C.c_double
        L.orc_critical_dt.argtypes = [C.POINTER(Problem), _dp]
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class OracleProblem:
    """Holds numpy arrays alive for the lifetime of the ctypes struct."""

    def __init__(self, mesh, law=LAW_NH_COMPRESSIBLE, E=1e6, nu=0.3, rho=1000.0,
                 dir_dof=None, dir_val=None, bc_mode=BC_VELOCITY, vel_ramp_steps=100,
                 t0=0.0, tf=1.0, scale=0.5, n_steps=0, sum_order=SUM_REFERENCE,
                 n_threads=1, energy_every=0, gravity=0.0, forces=None, gravity_pos=-1, loads=None,
                 visc=0, visc_c1=0.06, visc_cl=1.2, visc_h=0.0, visc_c0=0.0):
        """Neumann loads: see load_blocks (forces / gravity / gravity_pos, or an explicit block list)."""
        self.X = np.ascontiguousarray(mesh["X"], dtype=np.float64)
        self.conn = np.ascontiguousarray(mesh["conn"], dtype=np.int32)
        self.E = np.atleast_1d(np.asarray(E, dtype=np.float64)).copy()
        self.nu = np.atleast_1d(np.asarray(nu, dtype=np.float64)).copy()
        self.rho = np.atleast_1d(np.asarray(rho, dtype=np.float64)).copy()
        per = int(self.E.size > 1)
        if per:
            assert self.E.size == mesh["n_elem"] and self.nu.size == mesh["n_elem"]
            if self.rho.size == 1:
                self.rho = np.full(mesh["n_elem"], self.rho[0])
        dir_dof = np.zeros(0, np.int32) if dir_dof is None else dir_dof
        dir_val = np.zeros(0, np.float64) if dir_val is None else dir_val
        self.dir_dof = np.ascontiguousarray(dir_dof, dtype=np.int32)
        self.dir_val = np.ascontiguousarray(dir_val, dtype=np.float64)
        self.n_nodes = mesh["n_nodes"]
        self.blocks = load_blocks(forces, gravity, gravity_pos, loads)
        self.load_arr, self.load_keep = pack_loads(self.blocks)
        self.p = Problem(
            n_nodes=mesh["n_nodes"], n_elem=mesh["n_elem"], nen=mesh["nen"], law=law,
            X=dptr(self.X), conn=iptr(self.conn), mat_per_elem=per,
            E=dptr(self.E), nu=dptr(self.nu), rho=dptr(self.rho),
            n_dir=self.dir_dof.size,
            dir_dof=iptr(self.dir_dof) if self.dir_dof.size else None,
            dir_val=dptr(self.dir_val) if self.dir_val.size else None,
            bc_mode=bc_mode, vel_ramp_steps=vel_ramp_steps, t0=t0, tf=tf, scale=scale,
            n_steps=n_steps, sum_order=sum_order, n_threads=n_threads,
            energy_every=energy_every, n_loads=len(self.blocks), visc=visc, visc_c1=visc_c1, visc_cl=visc_cl,
            visc_h=visc_h, visc_c0=visc_c0,
            loads=C.cast(self.load_arr, C.c_void_p) if self.blocks else None)

    def run(self, u0=None, v0=None, a0=None, hist_cap=0):
        n3 = 3 * self.n_nodes
        u = np.zeros(n3) if u0 is None else np.array(u0, dtype=np.float64)
        v = np.zeros(n3) if v0 is None else np.array(v0, dtype=np.float64)
        a = np.zeros(n3) if a0 is None else np.array(a0, dtype=np.float64)
        hist = None
        harr = {}
        if hist_cap:
            for k in HIST_KEYS:
                harr[k] = np.full(hist_cap, np.nan)
            hist = History(cap=hist_cap, **{k: dptr(v_) for k, v_ in harr.items()})
        rep = Report()
        rc = lib().orc_run(C.byref(self.p), dptr(u), dptr(v), dptr(a),
                           C.byref(hist) if hist is not None else None, C.byref(rep))
        return dict(rc=rc, u=u, v=v, a=a, steps=rep.steps, time=rep.time,
                    dt_last=rep.dt_last, hist=harr)

    def mass(self):
        m = np.zeros(self.n_nodes)
        rc = lib().orc_lumped_mass(C.byref(self.p), dptr(m))
        assert rc == 0
        return m

    def internal_force(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        f = np.zeros(3 * self.n_nodes)
        rc = lib().orc_internal_force(C.byref(self.p), dptr(u), dptr(f))
        assert rc == 0
        return f

    def internal_force_visc(self, u, uprev, dt):
        u = np.ascontiguousarray(u, dtype=np.float64)
        uprev = np.ascontiguousarray(uprev, dtype=np.float64)
        f = np.zeros(3 * self.n_nodes)
        rc = lib().orc_internal_force_visc(C.byref(self.p), dptr(u), dptr(uprev), dt, dptr(f))
        assert rc == 0
        return f

    def element_outputs(self, u):
        """Per element [n_elem, 21]: GP-averaged Cauchy (9), strain (9), VMS, volumetric, equivalent strain."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(21 * self.p.n_elem)
        rc = lib().orc_element_outputs(C.byref(self.p), dptr(u), dptr(out))
        assert rc == 0
        return out.reshape(-1, 21)

    def energies(self, u, v):
        """(DEFO_ENERGY, KIN_ENERGY, VOLUME) summed over all GPs for the given u and v."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(3)
        rc = lib().orc_energies(C.byref(self.p), dptr(u), dptr(v), dptr(out))
        assert rc == 0
        return tuple(out)

    def external_force(self, t, u=None):
        f = np.zeros(3 * self.n_nodes)
        if u is not None:
            u = np.ascontiguousarray(u, dtype=np.float64)
        rc = lib().orc_external_force(C.byref(self.p), dptr(u) if u is not None else None, t, dptr(f))
        assert rc == 0
        return f

    def pressure_face(self, elem, face):
        """Surface GPs of one face: dict of position, ws, n0, phi_s, phi_b, dphi_b (one row per point)."""
        out = dict(xyz=np.zeros(12), ws=np.zeros(4), n0=np.zeros(12), phi_s=np.zeros(16), phi_b=np.zeros(32),
                   dphi_b=np.zeros(96))
        ng = lib().orc_pressure_face(C.byref(self.p), elem, face, *[dptr(out[k]) for k in
                                                                    ("xyz", "ws", "n0", "phi_s", "phi_b", "dphi_b")])
        assert ng > 0, ng
        return dict(xyz=out["xyz"][:3 * ng].reshape(ng, 3), ws=out["ws"][:ng], n0=out["n0"][:3 * ng].reshape(ng, 3),
                    phi_s=out["phi_s"][:4 * ng].reshape(ng, 4), phi_b=out["phi_b"][:8 * ng].reshape(ng, 8),
                    dphi_b=out["dphi_b"][:24 * ng].reshape(ng, 8, 3))

    def critical_dt(self, u=None):
        if u is not None:
            u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_critical_dt(C.byref(self.p), dptr(u) if u is not None else None)

"""ctypes bindings of libmps_gpu.so (the C-ABI declared in include/mps_gpu.h).

This module is plumbing for the test-suite and the benchmark: it passes numpy
buffers through the C-ABI unchanged.  There is no fallback path: when the native
library is missing, `load()` raises.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(PKG_DIR), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmps_gpu.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(PKG_DIR)), "include", "mps_gpu.h")

MPS_OK, MPS_ERR_ARG, MPS_ERR_HIP, MPS_ERR_DT, MPS_ERR_NAN, MPS_ERR_STATE, MPS_ERR_CAPACITY = 0, -1, -2, -3, -4, -5, -6
MPS_RUNNING, MPS_DONE = 0, 1
LAW_NH_COMPRESSIBLE, LAW_NH_BELYTSCHKO, LAW_SVK, LAW_LINEAR_ELASTIC = 0, 1, 2, 3
BC_RAMP, BC_INST, BC_VELOCITY, BC_RAMP_RATE = 0, 1, 2, 3
PREC_FP64, PREC_FP32_ELEM = 0, 1
SERIES_U, SERIES_V, SERIES_A, SERIES_FINT, SERIES_FEXT = 0, 1, 2, 3, 4  # MPS_SERIES_*
DOFS_STATIC_MAX = 65536  # MPS_DOFS_STATIC_MAX

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_elem", C.c_int32), ("nen", C.c_int32), ("_pad", C.c_int32),
                ("X", _dp), ("conn", _ip)]


class Material(C.Structure):
    _fields_ = [("law", C.c_int32), ("per_element", C.c_int32), ("E", _dp), ("nu", _dp), ("rho", _dp)]


class Bcs(C.Structure):
    _fields_ = [("n_dirichlet", C.c_int32), ("mode", C.c_int32), ("vel_ramp_steps", C.c_int32),
                ("_pad", C.c_int32), ("dof", _ip), ("value", _dp)]


class SolverCfg(C.Structure):
    _fields_ = [("t0", C.c_double), ("tf", C.c_double), ("scale", C.c_double), ("n_steps", C.c_int64),
                ("precision", C.c_int32), ("energy_every", C.c_int32), ("history_cap", C.c_int32),
                ("device", C.c_int32)]


class Load(C.Structure):
    _fields_ = [("type", C.c_int32), ("n", C.c_int32), ("t0", C.c_double), ("tf", C.c_double),
                ("value", C.c_double), ("index", _ip), ("face", _ip), ("force", _dp)]


LOAD_FORCE, LOAD_GRAVITY, LOAD_PRESSURE_RAMP, LOAD_PRESSURE_INST, LOAD_TRACTION_RAMP, LOAD_TRACTION_INST = 0, 1, 2, 3, 4, 5
VISC_NONE, VISC_BULK, VISC_DEV = 0, 1, 2
_LOAD_TYPES = {"force": LOAD_FORCE, "gravity": LOAD_GRAVITY, "pressure_ramp": LOAD_PRESSURE_RAMP,
               "pressure_inst": LOAD_PRESSURE_INST, "traction_ramp": LOAD_TRACTION_RAMP,
               "traction_inst": LOAD_TRACTION_INST}


def load_blocks(forces=None, gravity=0.0, gravity_pos=-1, loads=None):
    """Neumann blocks in input order, as tuples:
      ("force", nodes, force[k,3], t0, tf)                *BOUNDARY, TYPE=FORCE (entries F*Ff, one window)
      ("gravity", g)                                      *GRAVITY
      ("pressure_ramp" | "pressure_inst", P, t0, tf, elements, faces)
      ("traction_ramp" | "traction_inst", (Tx, Ty, Tz), t0, tf, elements, faces)
    `loads` is such a list; otherwise forces = (node[k], force[k,3], t0[k], tf[k]) gives one block per entry and
    a non-zero gravity goes before entry gravity_pos (-1: after all)."""
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


class Report(C.Structure):
    _fields_ = [("steps", C.c_int64), ("time", C.c_double), ("dt", C.c_double), ("status", C.c_int32),
                ("nonfinite", C.c_int32), ("device_ms", C.c_double)]


class BatchCfg(C.Structure):
    _fields_ = [("arena_bytes", C.c_int64), ("max_instances", C.c_int32), ("precision", C.c_int32),
                ("steps_per_launch", C.c_int32), ("elem_block", C.c_int32), ("grid_mult", C.c_int32),
                ("device", C.c_int32)]


class ServerCfg(C.Structure):
    _fields_ = [("batch", BatchCfg)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("launches", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double), ("queue_p50_ms", C.c_double),
                ("span_ms", C.c_double), ("solves_per_s", C.c_double), ("busy_ms", C.c_double),
                ("elem_block", C.c_int32), ("grid_mult", C.c_int32), ("tunings", C.c_int32), ("_pad", C.c_int32)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class MpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mps error {code}: {msg}")
        self.code = code


_lib = None


def load(path=None):
    """Load libmps_gpu.so (raises OSError if it is not built).  MPS_GPU_LIB overrides the path."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("MPS_GPU_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"libmps_gpu.so not built: {p} missing (run __graft_entry__.build())")
    L = C.CDLL(p)
    H = C.c_void_p
    L.mps_create.restype = C.c_int
    L.mps_create.argtypes = [C.POINTER(Mesh), C.POINTER(Material), C.POINTER(Bcs), C.POINTER(SolverCfg), C.POINTER(H)]
    L.mps_set_state.argtypes = [H, _dp, _dp, _dp]
    L.mps_get_state.argtypes = [H, _dp, _dp, _dp]
    L.mps_run.argtypes = [H, C.c_int64, C.POINTER(Report)]
    L.mps_get_history.argtypes = [H, C.c_int64, C.c_int64, _dp, _dp, _dp, _dp, _dp]
    L.mps_get_energy_balance.argtypes = [H, C.c_int64, C.c_int64, _dp, _dp, _dp, _dp]
    L.mps_get_mass.argtypes = [H, _dp]
    L.mps_internal_force.argtypes = [H, _dp, _dp]
    L.mps_critical_dt.argtypes = [H, _dp, _dp]
    L.mps_element_outputs.argtypes = [H, _dp]
    _cdpp = C.POINTER(C.POINTER(C.c_double))
    L.mps_snapshot_begin.argtypes = [H, C.c_int32, C.c_int32]
    L.mps_set_dof_series.argtypes = [H, C.c_int64, _ip, _ip]
    L.mps_get_dof_series.argtypes = [H, C.c_int64, C.c_int64, _dp]
    L.mps_snapshot_wait.argtypes = [H, C.c_int32, _cdpp, _cdpp, _cdpp, _cdpp]
    L.mps_snapshot_release.argtypes = [H, C.c_int32]
    L.mps_set_loads.argtypes = [H, C.c_int32, C.POINTER(Load)]
    L.mps_set_viscosity.argtypes = [H, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double]
    L.mps_external_force.argtypes = [H, C.c_double, _dp]
    L.mps_get_dofs.argtypes = [H, C.c_int32, C.c_int64, _ip, _dp]
    L.mps_get_dofs_static.argtypes = [H, C.c_int32, C.c_int64, _ip, _dp]
    L.mps_get_internal_force_dofs.argtypes = [H, C.c_int64, _ip, _dp]
    L.mps_destroy.argtypes = [H]
    L.mps_last_error.restype = C.c_char_p
    L.mps_last_error.argtypes = [H]
    L.mps_device_info.argtypes = [C.c_int32, C.c_char_p, C.c_int32]
    L.mps_bandwidth_probe.argtypes = [C.c_int32, C.c_int64, C.c_int32, _dp]
    L.mps_profile.argtypes = [H, C.c_int32]
    L.mps_set_graphs.argtypes = [H, C.c_int32]
    L.mps_prepare_run.argtypes = [H, C.c_int64]
    L.mps_kernel_times.argtypes = [H, _dp, C.POINTER(C.c_int64)]
    L.mps_layout.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    if hasattr(L, "mps_batch_create"):
        B = C.c_void_p
        L.mps_batch_create.argtypes = [C.POINTER(BatchCfg), C.POINTER(B)]
        L.mps_batch_add.argtypes = [B, C.POINTER(Mesh), C.POINTER(Material), C.POINTER(Bcs), C.POINTER(SolverCfg),
                                    C.POINTER(C.c_int32)]
        L.mps_batch_step.argtypes = [B, C.c_int64]
        L.mps_batch_query.argtypes = [B, C.c_int32, C.POINTER(Report)]
        L.mps_batch_get_state.argtypes = [B, C.c_int32, _dp, _dp, _dp]
        L.mps_batch_remove.argtypes = [B, C.c_int32]
        L.mps_batch_autotune.argtypes = [B, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.mps_batch_destroy.argtypes = [B]
        L.mps_batch_launch.argtypes = [B, C.c_int64]
        L.mps_batch_sync.argtypes = [B]
        L.mps_batch_info.argtypes = [B, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    if hasattr(L, "mps_server_create"):
        S = C.c_void_p
        L.mps_server_create.argtypes = [C.POINTER(ServerCfg), C.POINTER(S)]
        L.mps_server_submit.argtypes = [S, C.POINTER(Mesh), C.POINTER(Material), C.POINTER(Bcs),
                                        C.POINTER(SolverCfg), C.POINTER(C.c_int64)]
        L.mps_server_poll.argtypes = [S, C.c_int64, C.POINTER(C.c_int32), _dp, _dp]
        L.mps_server_wait.argtypes = [S, C.c_int64, C.c_double, C.POINTER(C.c_int32)]
        L.mps_server_drain.argtypes = [S, C.c_double]
        L.mps_server_get_stats.argtypes = [S, C.POINTER(ServerStats)]
        L.mps_server_reset_stats.argtypes = [S]
        L.mps_server_destroy.argtypes = [S]
        L.mps_server_last_error.restype = C.c_char_p
        L.mps_server_last_error.argtypes = [S]
        L.mps_batch_last_error.restype = C.c_char_p
        L.mps_batch_last_error.argtypes = [B]
    _lib = L
    return L


def declared_symbols(header=HEADER):
    """Function names declared in include/mps_gpu.h."""
    import re
    txt = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*\*?\s*(mps_\w+)\s*\(", txt, re.M)))


def device_info(device=0):
    buf = C.create_string_buffer(256)
    rc = load().mps_device_info(device, buf, 256)
    return buf.value.decode() if rc == 0 else None


def bandwidth_probe(nbytes=1 << 30, reps=20, device=0):
    """Measured copy bandwidth (read+write GB/s) of the device."""
    out = C.c_double()
    rc = load().mps_bandwidth_probe(device, nbytes, reps, C.byref(out))
    if rc != 0:
        raise MpsError(rc, "bandwidth probe failed")
    return out.value


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class _Inputs:
    """Keeps the numpy inputs and the ctypes structs of one solve alive."""

    def __init__(self, mesh, law, E, nu, rho, dir_dof, dir_val, bc_mode, vel_ramp_steps,
                 t0, tf, scale, n_steps, precision, energy_every, history_cap, device):
        self.X = np.ascontiguousarray(mesh["X"], dtype=np.float64)
        self.conn = np.ascontiguousarray(mesh["conn"], dtype=np.int32)
        self.E = np.ascontiguousarray(np.atleast_1d(E), dtype=np.float64)
        self.nu = np.ascontiguousarray(np.atleast_1d(nu), dtype=np.float64)
        self.rho = np.ascontiguousarray(np.atleast_1d(rho), dtype=np.float64)
        per = int(self.E.size > 1)
        if per and self.rho.size == 1:
            self.rho = np.full(self.E.size, self.rho[0])
        self.dof = np.ascontiguousarray(np.zeros(0) if dir_dof is None else dir_dof, dtype=np.int32)
        self.val = np.ascontiguousarray(np.zeros(0) if dir_val is None else dir_val, dtype=np.float64)
        self.mesh = Mesh(n_nodes=mesh["n_nodes"], n_elem=mesh["n_elem"], nen=mesh["nen"], X=_d(self.X),
                         conn=_i(self.conn))
        self.mat = Material(law=law, per_element=per, E=_d(self.E), nu=_d(self.nu), rho=_d(self.rho))
        self.bcs = Bcs(n_dirichlet=self.dof.size, mode=bc_mode, vel_ramp_steps=vel_ramp_steps,
                       dof=_i(self.dof) if self.dof.size else None, value=_d(self.val) if self.val.size else None)
        self.cfg = SolverCfg(t0=t0, tf=tf, scale=scale, n_steps=n_steps, precision=precision,
                             energy_every=energy_every, history_cap=history_cap, device=device)
        self.n_nodes = mesh["n_nodes"]


class Solver:
    """One explicit solve on the device (mps_create ... mps_destroy)."""

    def __init__(self, mesh, law=LAW_NH_COMPRESSIBLE, E=1e6, nu=0.3, rho=1000.0, dir_dof=None, dir_val=None,
                 bc_mode=BC_VELOCITY, vel_ramp_steps=100, t0=0.0, tf=1.0, scale=0.5, n_steps=0,
                 precision=PREC_FP64, energy_every=0, history_cap=0, device=0):
        self.L = load()
        self.inp = _Inputs(mesh, law, E, nu, rho, dir_dof, dir_val, bc_mode, vel_ramp_steps, t0, tf, scale,
                           n_steps, precision, energy_every, history_cap, device)
        self.h = C.c_void_p()
        rc = self.L.mps_create(C.byref(self.inp.mesh), C.byref(self.inp.mat), C.byref(self.inp.bcs),
                               C.byref(self.inp.cfg), C.byref(self.h))
        if rc != 0:
            raise MpsError(rc, self.L.mps_last_error(None).decode())
        self.n = self.inp.n_nodes
        self.ne = int(mesh["n_elem"])

    def _chk(self, rc):
        if rc != 0:
            raise MpsError(rc, self.L.mps_last_error(self.h).decode())

    def set_state(self, u=None, v=None, a=None):
        arrs = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (u, v, a)]
        self._chk(self.L.mps_set_state(self.h, *[None if x is None else _d(x) for x in arrs]))

    def set_loads(self, forces=None, gravity=0.0, gravity_pos=-1, loads=None):
        """Neumann loads in input order: see load_blocks."""
        blocks = load_blocks(forces, gravity, gravity_pos, loads)
        arr = (Load * max(len(blocks), 1))()
        keep = []
        for b, blk in enumerate(blocks):
            kind = _LOAD_TYPES[blk[0]]
            Lb = arr[b]
            Lb.type = kind
            if kind == LOAD_FORCE:
                nodes = np.ascontiguousarray(blk[1], dtype=np.int32)
                vec = np.ascontiguousarray(np.reshape(blk[2], (-1, 3)), dtype=np.float64)
                keep += [nodes, vec]
                Lb.n, Lb.t0, Lb.tf, Lb.index, Lb.force = nodes.size, blk[3], blk[4], _i(nodes), _d(vec)
            elif kind == LOAD_GRAVITY:
                Lb.n, Lb.value = 0, blk[1]
            else:
                el = np.ascontiguousarray(blk[4], dtype=np.int32)
                fc = np.ascontiguousarray(blk[5], dtype=np.int32)
                keep += [el, fc]
                Lb.n, Lb.t0, Lb.tf, Lb.index, Lb.face = el.size, blk[2], blk[3], _i(el), _i(fc)
                if kind in (LOAD_TRACTION_RAMP, LOAD_TRACTION_INST):
                    T = np.ascontiguousarray(blk[1], dtype=np.float64).reshape(3)
                    keep.append(T)
                    Lb.force = _d(T)
                else:
                    Lb.value = blk[1]
        self._loads = keep
        self._chk(self.L.mps_set_loads(self.h, len(blocks), arr))

    def set_viscosity(self, kind, c1=0.06, cl=1.2, elem_size=0.0, c0=0.0):
        """kind: VISC_NONE / VISC_BULK / VISC_DEV (*Bulk viscosity / *Dev viscosity)."""
        self._chk(self.L.mps_set_viscosity(self.h, kind, c1, cl, elem_size, c0))

    def external_force(self, t):
        f = np.zeros(3 * self.n)
        self._chk(self.L.mps_external_force(self.h, t, _d(f)))
        return f

    def get_state(self):
        u, v, a = (np.zeros(3 * self.n) for _ in range(3))
        self._chk(self.L.mps_get_state(self.h, _d(u), _d(v), _d(a)))
        return u, v, a

    def get_dofs(self, which, dof):
        """u (which=0), v (1) or a (2) at DOF indices node + n_nodes*dir (mps_get_dofs)."""
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        out = np.zeros(dof.size)
        self._chk(self.L.mps_get_dofs(self.h, which, dof.size, _i(dof), _d(out)))
        return out

    def internal_force_dofs(self, dof):
        """Internal force of the current state at DOF indices (mps_get_internal_force_dofs)."""
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        out = np.zeros(dof.size)
        self._chk(self.L.mps_get_internal_force_dofs(self.h, dof.size, _i(dof), _d(out)))
        return out

    def get_dofs_static(self, which, dof):
        """As get_dofs, staged through the device's fixed array (mps_get_dofs_static: no allocation)."""
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        out = np.zeros(dof.size)
        self._chk(self.L.mps_get_dofs_static(self.h, which, dof.size, _i(dof), _d(out)))
        return out

    def run(self, max_steps):
        """Returns (status_code, report dict); errors of the solve are codes, not exceptions."""
        rep = Report()
        rc = self.L.mps_run(self.h, max_steps, C.byref(rep))
        if rc in (MPS_ERR_ARG, MPS_ERR_HIP, MPS_ERR_STATE):
            self._chk(rc)
        return rc, dict(steps=rep.steps, time=rep.time, dt=rep.dt, status=rep.status, nonfinite=rep.nonfinite,
                        device_ms=rep.device_ms)

    def history(self, first, count):
        out = [np.zeros(count) for _ in range(5)]
        self._chk(self.L.mps_get_history(self.h, first, count, *[_d(x) for x in out]))
        return dict(zip(("time", "dt", "e_def", "e_kin", "volume"), out))

    def energy_balance(self, first, count):
        """e_ext (u . fext), e_kin_lumped, w_ext, w_react of steps first+1 .. first+count (mps_get_energy_balance)."""
        out = [np.zeros(count) for _ in range(4)]
        self._chk(self.L.mps_get_energy_balance(self.h, first, count, *[_d(x) for x in out]))
        return dict(zip(("e_ext", "e_kin_lumped", "w_ext", "w_react"), out))

    def element_outputs(self):
        """Per element [n_elem, 21]: GP-averaged Cauchy (9), strain (9), VMS, volumetric, equivalent strain."""
        out = np.zeros(21 * self.ne)
        self._chk(self.L.mps_element_outputs(self.h, _d(out)))
        return out.reshape(-1, 21)

    def set_dof_series(self, which, dof):
        """Per-step series recorded on the device: which[i] in SERIES_* at DOF dof[i] = node + n_nodes*dir."""
        which = np.ascontiguousarray(which, dtype=np.int32)
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        assert which.size == dof.size
        self._chk(self.L.mps_set_dof_series(self.h, dof.size, _i(which), _i(dof)))
        self._nser = dof.size

    def dof_series(self, first, count):
        """Rows [first, first+count) (row = step - 1) of the series: array [count, n]."""
        n = getattr(self, "_nser", 0)
        out = np.zeros((count, n))
        if n and count:
            self._chk(self.L.mps_get_dof_series(self.h, first, count, _d(out)))
        return out

    def snapshot_begin(self, slot, with_elements=True):
        """Stage u, v, a (and the element outputs) of the last step for a copy that runs while stepping goes on."""
        self._chk(self.L.mps_snapshot_begin(self.h, slot, int(with_elements)))

    def snapshot_wait(self, slot):
        """Wait for the slot's copy: (u, v, a, element outputs [n_elem, 21] or None), as numpy copies."""
        p = [C.POINTER(C.c_double)() for _ in range(4)]
        self._chk(self.L.mps_snapshot_wait(self.h, slot, *(C.byref(x) for x in p)))
        u, v, a = (np.ctypeslib.as_array(x, shape=(3 * self.n,)).copy() for x in p[:3])
        e = np.ctypeslib.as_array(p[3], shape=(21 * self.ne,)).copy().reshape(-1, 21) if p[3] else None
        return u, v, a, e

    def snapshot_release(self, slot):
        self._chk(self.L.mps_snapshot_release(self.h, slot))

    def mass(self):
        m = np.zeros(self.n)
        self._chk(self.L.mps_get_mass(self.h, _d(m)))
        return m

    def internal_force(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        f = np.zeros(3 * self.n)
        self._chk(self.L.mps_internal_force(self.h, _d(u), _d(f)))
        return f

    def critical_dt(self, u=None):
        out = C.c_double()
        if u is not None:
            u = np.ascontiguousarray(u, dtype=np.float64)
        self._chk(self.L.mps_critical_dt(self.h, _d(u) if u is not None else None, C.byref(out)))
        return out.value

    def set_graphs(self, enable=True):
        self._chk(self.L.mps_set_graphs(self.h, int(enable)))

    def prepare_run(self, max_steps):
        """Capture the step graphs a later run(max_steps) replays, without running them."""
        self._chk(self.L.mps_prepare_run(self.h, max_steps))

    def profile(self, enable=True):
        self._chk(self.L.mps_profile(self.h, int(enable)))

    def kernel_times(self):
        ms = np.zeros(4)
        n = (C.c_int64 * 4)()
        self._chk(self.L.mps_kernel_times(self.h, _d(ms), n))
        kinds = ("element", "node", "begin", "energy")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(kinds)}

    def layout(self):
        """Patches, partial forces and the stored element order (input id of each stored element)."""
        npa, npt = C.c_int64(), C.c_int64()
        order = np.zeros(self.ne, dtype=np.int32)
        self._chk(self.L.mps_layout(self.h, C.byref(npa), C.byref(npt), order.ctypes.data_as(C.POINTER(C.c_int32))))
        return {"n_patch": npa.value, "n_part": npt.value, "order": order}

    def close(self):
        if self.h:
            self.L.mps_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched form: many independent fixed-step solves stepped together (mps_batch_*)."""

    def __init__(self, arena_bytes=1 << 30, max_instances=1024, precision=PREC_FP64, steps_per_launch=32,
                 elem_block=0, grid_mult=0, device=0):
        self.L = load()
        self.cfg = BatchCfg(arena_bytes=arena_bytes, max_instances=max_instances, precision=precision,
                            steps_per_launch=steps_per_launch, elem_block=elem_block, grid_mult=grid_mult,
                            device=device)
        self.b = C.c_void_p()
        rc = self.L.mps_batch_create(C.byref(self.cfg), C.byref(self.b))
        if rc != 0:
            raise MpsError(rc, "mps_batch_create failed")
        self.inputs = {}
        self.precision = precision

    def _chk(self, rc):
        if rc != 0:
            raise MpsError(rc, self.L.mps_batch_last_error(self.b).decode())

    def add(self, mesh, law=LAW_NH_COMPRESSIBLE, E=1e6, nu=0.3, rho=1000.0, dir_dof=None, dir_val=None,
            bc_mode=BC_VELOCITY, vel_ramp_steps=100, scale=0.5, n_steps=1000, t0=0.0, tf=1.0):
        inp = _Inputs(mesh, law, E, nu, rho, dir_dof, dir_val, bc_mode, vel_ramp_steps, t0, tf, scale, n_steps,
                      self.precision, 0, 0, self.cfg.device)
        slot = C.c_int32()
        rc = self.L.mps_batch_add(self.b, C.byref(inp.mesh), C.byref(inp.mat), C.byref(inp.bcs), C.byref(inp.cfg),
                                  C.byref(slot))
        self._chk(rc)
        self.inputs[slot.value] = inp
        return slot.value

    def step(self, max_steps):
        self._chk(self.L.mps_batch_step(self.b, max_steps))

    def launch(self, launches):
        self._chk(self.L.mps_batch_launch(self.b, launches))

    def sync(self):
        self._chk(self.L.mps_batch_sync(self.b))

    def query(self, slot):
        rep = Report()
        self._chk(self.L.mps_batch_query(self.b, slot, C.byref(rep)))
        return dict(steps=rep.steps, time=rep.time, dt=rep.dt, status=rep.status, nonfinite=rep.nonfinite)

    def get_state(self, slot):
        n = self.inputs[slot].n_nodes
        u, v, a = (np.zeros(3 * n) for _ in range(3))
        self._chk(self.L.mps_batch_get_state(self.b, slot, _d(u), _d(v), _d(a)))
        return u, v, a

    def remove(self, slot):
        self._chk(self.L.mps_batch_remove(self.b, slot))
        self.inputs.pop(slot, None)

    def info(self):
        f, l_, a = C.c_int64(), C.c_int64(), C.c_int32()
        self._chk(self.L.mps_batch_info(self.b, C.byref(f), C.byref(l_), C.byref(a)))
        return dict(free=f.value, largest=l_.value, active=a.value)

    def autotune(self):
        bs, gm = C.c_int32(), C.c_int32()
        self._chk(self.L.mps_batch_autotune(self.b, C.byref(bs), C.byref(gm)))
        return bs.value, gm.value

    def close(self):
        if self.b:
            self.L.mps_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving form (mps_server_*): continuous batching request loop on the device."""

    def __init__(self, arena_bytes=4 << 30, max_instances=1024, precision=PREC_FP32_ELEM, steps_per_launch=32,
                 elem_block=0, grid_mult=0, device=0):
        self.L = load()
        self.cfg = ServerCfg(batch=BatchCfg(arena_bytes=arena_bytes, max_instances=max_instances,
                                            precision=precision, steps_per_launch=steps_per_launch,
                                            elem_block=elem_block, grid_mult=grid_mult, device=device))
        self.s = C.c_void_p()
        rc = self.L.mps_server_create(C.byref(self.cfg), C.byref(self.s))
        if rc != 0:
            raise MpsError(rc, "mps_server_create failed")
        self.precision = precision
        self.nodes = {}

    def submit(self, mesh, law=LAW_NH_COMPRESSIBLE, E=1e6, nu=0.3, rho=1000.0, dir_dof=None, dir_val=None,
               bc_mode=BC_VELOCITY, vel_ramp_steps=100, scale=0.5, n_steps=1000):
        inp = _Inputs(mesh, law, E, nu, rho, dir_dof, dir_val, bc_mode, vel_ramp_steps, 0.0, 1.0, scale, n_steps,
                      self.precision, 0, 0, self.cfg.batch.device)
        rid = C.c_int64()
        rc = self.L.mps_server_submit(self.s, C.byref(inp.mesh), C.byref(inp.mat), C.byref(inp.bcs),
                                      C.byref(inp.cfg), C.byref(rid))
        if rc != 0:
            raise MpsError(rc, self.L.mps_server_last_error(self.s).decode())
        self.nodes[rid.value] = inp.n_nodes
        return rid.value

    def result(self, rid, timeout=600.0):
        """Waits for request rid; returns (status, u, latency_ms, queue_ms)."""
        st = C.c_int32()
        self.L.mps_server_wait(self.s, rid, timeout, C.byref(st))
        u = np.zeros(3 * self.nodes[rid])
        lat = np.zeros(2)
        rc = self.L.mps_server_poll(self.s, rid, C.byref(st), _d(u), _d(lat))
        return rc, (u if st.value == REQ_DONE else None), lat[0], lat[1]

    def drain(self, timeout=600.0):
        return self.L.mps_server_drain(self.s, timeout)

    def stats(self):
        st = ServerStats()
        self.L.mps_server_get_stats(self.s, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        self.L.mps_server_reset_stats(self.s)

    def close(self):
        if self.s:
            self.L.mps_server_destroy(self.s)
            self.s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

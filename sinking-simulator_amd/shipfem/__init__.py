"""Python binding of libshipfem.so (include/shipfem.h) through ctypes.

The compute path is the HIP library; this module only moves numpy arrays across the
C ABI. It raises ImportError-like RuntimeError when the library is missing: there is
no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libshipfem.so")

OK, NO_DEVICE, INVALID_ARG, OUT_OF_MEMORY, UNSTABLE, DEVICE_ERROR, BUSY, NOT_FOUND, UNSUPPORTED = range(9)
PIN_X, PIN_Y, PIN_Z, PIN_ALL = 0x1, 0x2, 0x4, 0x7
LOADED, CONTACT = 0x8, 0x10
DRIVE_X, DRIVE_Y, DRIVE_Z = 0x20, 0x40, 0x80


class ShipfemError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"shipfem status {status} ({status_string(status)}): {msg}")


D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
U32 = C.POINTER(C.c_uint32)


class Material(C.Structure):
    _fields_ = [("youngs_modulus", C.c_double), ("poisson_ratio", C.c_double),
                ("density", C.c_double), ("plastic", C.c_int32), ("yield_strength", C.c_double),
                ("hardening_modulus", C.c_double), ("failure_strain", C.c_double),
                ("triaxiality_sensitivity", C.c_double), ("reference_triaxiality", C.c_double),
                ("cutoff_triaxiality", C.c_double)]


# the reference's Rice-Tracey constants (engine/sim/plasticity.hpp:193-199)
RICE_TRACEY = dict(sensitivity=1.5, reference_triaxiality=1.0 / 3.0,
                   cutoff_triaxiality=-1.0 / 3.0)


class Indenter(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("radius", C.c_double), ("center", C.c_double * 3),
                ("velocity", C.c_double * 3), ("penalty", C.c_double)]


class CaseDesc(C.Structure):
    _fields_ = [("node_count", C.c_uint32), ("tet_count", C.c_uint32), ("rest_position", D),
                ("index", U32), ("node_flags", U32), ("load", D), ("prescribed", D),
                ("initial_displacement", D), ("initial_velocity", D), ("material", Material),
                ("indenter", Indenter), ("load_ramp", C.c_double), ("drive_ramp", C.c_double),
                ("gravity", C.c_double), ("damping", C.c_double), ("dt", C.c_double),
                ("dt_safety", C.c_double), ("sample_every", C.c_int32),
                ("max_steps", C.c_int64), ("mass", D), ("fracture_strain", C.c_double)]


class StepParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double), ("damping", C.c_double),
                ("youngs_modulus", C.c_double), ("poisson_ratio", C.c_double)]


class EnergyRecord(C.Structure):
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("strain", C.c_double),
                ("kinetic", C.c_double), ("dissipation", C.c_double), ("w_ext", C.c_double),
                ("w_contact", C.c_double), ("w_boundary", C.c_double),
                ("damping_loss", C.c_double), ("fracture", C.c_double),
                ("contact_force", C.c_double * 3)]


class StateView(C.Structure):
    _fields_ = [("displacement", D), ("velocity", D), ("position", F), ("eqps", F),
                ("plastic_strain", F), ("damage", F), ("step", C.c_int64), ("time", C.c_double),
                ("status", C.c_uint32)]


class Tear(C.Structure):
    _fields_ = [("tet", C.c_uint32), ("step", C.c_int64), ("normal", C.c_double * 3)]


VARIANT_TILED, VARIANT_SLOTS = 0, 1


class LaunchConfig(C.Structure):
    _fields_ = [("elem_block", C.c_int32), ("node_block", C.c_int32), ("xcd_remap", C.c_int32),
                ("chunk", C.c_int32), ("variant", C.c_int32), ("split", C.c_int32),
                ("rest_from_positions", C.c_int32), ("resident", C.c_int32)]


ROUNDS_ALL, ROUNDS_SHORTEST = 0, 1


class ServerOpts(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("max_active", C.c_uint32), ("chunk", C.c_int32),
                ("launch", LaunchConfig), ("policy", C.c_int32), ("round_work", C.c_double),
                ("starve_rounds", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("status", C.c_int), ("steps", C.c_int64),
                ("sim_time", C.c_double), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("node_count", C.c_uint32), ("tet_count", C.c_uint32),
                ("max_displacement", C.c_double), ("max_eqps", C.c_double),
                ("peak_contact_force", C.c_double), ("torn_elements", C.c_uint32),
                ("torn_volume", C.c_double), ("last", EnergyRecord)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("failed", C.c_uint64), ("refused", C.c_uint64), ("active", C.c_uint64),
                ("queued", C.c_uint64), ("bytes_in_use", C.c_uint64),
                ("bytes_budget", C.c_uint64), ("rounds", C.c_uint64),
                ("element_steps", C.c_double), ("device_ms", C.c_double)]


class Handoff(C.Structure):
    _fields_ = [("position", D), ("displacement", D), ("velocity", D), ("plastic", D),
                ("tear_step", C.POINTER(C.c_int64)), ("tear_stress", D),
                ("steps", C.c_int64), ("time", C.c_double), ("penetration", C.c_double),
                ("work", C.c_double), ("w_ext", C.c_double), ("w_contact", C.c_double),
                ("w_boundary", C.c_double), ("damping_loss", C.c_double),
                ("dissipation", C.c_double), ("fracture", C.c_double), ("torn", C.c_uint32),
                ("hex_points", D), ("hex_enhanced", D)]


ENERGY_FIELDS = ["strain", "kinetic", "dissipation", "w_ext", "w_contact", "w_boundary",
                 "damping_loss", "fracture"]
HANDOFF_SCALARS = ["steps", "time", "penetration", "work", "w_ext", "w_contact", "w_boundary",
                   "damping_loss", "dissipation", "fracture", "torn"]

# every symbol include/shipfem.h declares, with (restype, argtypes)
VP = C.c_void_p
PP = C.POINTER(C.c_void_p)
S = C.c_int
SIGNATURES = {
    "shipfem_status_string": (C.c_char_p, [S]),
    "shipfem_device_query": (S, [C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64)]),
    "shipfem_mesh_box": (S, [C.c_double] * 3 + [C.c_int32] * 3 + [PP]),
    "shipfem_mesh_jitter": (S, [VP, C.c_double, C.c_uint64]),
    "shipfem_mesh_panel": (S, [C.c_double] * 3 + [C.c_int32] + [C.c_double] * 2 + [C.c_int32] * 3 + [PP]),
    "shipfem_mesh_counts": (S, [VP, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "shipfem_mesh_arrays": (S, [VP, C.POINTER(D), C.POINTER(U32), C.POINTER(U32), C.POINTER(D)]),
    "shipfem_mesh_flag_plane": (S, [VP, C.c_int32, C.c_double, C.c_double, C.c_uint32]),
    "shipfem_mesh_flag_disc": (S, [VP, C.c_int32] + [C.c_double] * 5 + [C.c_uint32]),
    "shipfem_mesh_traction": (S, [VP, C.c_int32, C.c_double, C.c_double, D]),
    "shipfem_mesh_fill_case": (S, [VP, C.POINTER(CaseDesc)]),
    "shipfem_mesh_destroy": (None, [VP]),
    "shipfem_edge_timestep": (C.c_double, [D, U32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
    "shipfem_stable_timestep": (S, [D, U32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, D,
                                    C.c_char_p, C.c_size_t]),
    "shipfem_stable_timestep_noalloc": (S, [D, U32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                            D, C.c_char_p, C.c_size_t]),
    "shipfem_hex_failure_strain": (S, [D, C.c_double, C.c_double, D]),
    "shipfem_cpu_tet_forces": (S, [F, D, U32, F, F, C.c_uint64, C.c_float, C.c_float, F]),
    "shipfem_cpu_step": (S, [C.c_uint32, C.c_uint32, F, F, F, U32, U32, F, F, F, U32, U32, D] +
                         [C.c_float] * 5),
    "shipfem_solver_create": (S, [C.POINTER(CaseDesc), PP, C.c_char_p, C.c_size_t]),
    "shipfem_solver_timestep": (S, [VP, D, D, D]),
    "shipfem_solver_set_params": (S, [VP, C.POINTER(StepParams)]),
    "shipfem_solver_run": (S, [VP, C.c_int32, C.c_int32, D]),
    "shipfem_solver_readback": (S, [VP, C.POINTER(StateView)]),
    "shipfem_solver_energy": (S, [VP, C.POINTER(EnergyRecord), C.c_int32, C.POINTER(C.c_int32)]),
    "shipfem_solver_tet_forces": (S, [VP, F]),
    "shipfem_solver_set_launch": (S, [VP, C.POINTER(LaunchConfig)]),
    "shipfem_solver_profile": (S, [VP, C.c_int32, D, D]),
    "shipfem_solver_footprint": (S, [VP, C.POINTER(C.c_uint64)]),
    "shipfem_solver_traffic": (S, [VP, C.POINTER(C.c_uint64)]),
    "shipfem_batch_profile": (S, [VP, C.c_int32, D, D]),
    "shipfem_batch_traffic": (S, [VP, C.POINTER(C.c_uint64)]),
    "shipfem_batch_footprint": (S, [VP, C.POINTER(C.c_uint64)]),
    "shipfem_case_footprint_mesh": (C.c_uint64, [C.POINTER(CaseDesc), C.c_int32]),
    "shipfem_solver_tears": (S, [VP, C.POINTER(Tear), C.c_uint32, C.POINTER(C.c_uint32)]),
    "shipfem_regularised_failure_strain": (C.c_double, [C.c_double] * 4),
    "shipfem_torn_groups": (S, [D, U32, C.c_uint32, U32, F, C.c_double, U32, C.c_uint32,
                                C.POINTER(C.c_uint32)]),
    "shipfem_solver_export": (S, [VP, C.POINTER(Handoff)]),
    "shipfem_solver_adopt": (S, [VP, C.POINTER(Handoff)]),
    "shipfem_solver_checkpoint": (S, [VP, VP, C.POINTER(C.c_uint64)]),
    "shipfem_solver_restore": (S, [VP, VP, C.c_uint64]),
    "shipfem_apply_edge_drive": (C.c_uint64, [C.c_uint32, D, C.POINTER(C.c_uint8), U32, D]),
    "shipfem_case_encoded_size": (C.c_uint64, [C.POINTER(CaseDesc)]),
    "shipfem_case_encode": (S, [C.POINTER(CaseDesc), VP, C.c_uint64]),
    "shipfem_case_decode": (S, [VP, C.c_uint64, PP, C.POINTER(CaseDesc)]),
    "shipfem_case_blob_destroy": (None, [VP]),
    "shipfem_case_save": (S, [C.POINTER(CaseDesc), C.c_char_p]),
    "shipfem_hex_case_encoded_size": (C.c_uint64, [C.POINTER(CaseDesc)]),
    "shipfem_hex_case_encode": (S, [C.POINTER(CaseDesc), VP, C.c_uint64]),
    "shipfem_hex_case_save": (S, [C.POINTER(CaseDesc), C.c_char_p]),
    "shipfem_case_blob_corners": (C.c_uint32, [VP]),
    "shipfem_case_load": (S, [C.c_char_p, PP, C.POINTER(CaseDesc)]),
    "shipfem_server_submit_encoded": (S, [VP, VP, C.c_uint64, C.c_double, C.POINTER(C.c_uint64)]),
    "shipfem_solver_destroy": (None, [VP]),
    "shipfem_hex_stiffness": (S, [D, C.c_double, C.c_double, D]),
    "shipfem_hex_mass": (S, [D, C.c_double, D]),
    "shipfem_hex_force": (S, [D, D, D, D]),
    "shipfem_hex_rotation": (S, [D, D, D]),
    "shipfem_hex_plastic_update": (S, [D, D] + [C.c_double] * 8 + [D, D, D, D, C.POINTER(C.c_int32)]),
    "shipfem_hex_critical_timestep": (S, [D] + [C.c_double] * 4 + [D]),
    "shipfem_hex_solver_create": (S, [C.POINTER(CaseDesc), PP, C.c_char_p, C.c_size_t]),
    "shipfem_hex_plate_mesh": (S, [C.c_double] * 3 + [C.c_int32] * 3 + [C.POINTER(C.c_uint32)] * 2
                               + [D, U32]),
    "shipfem_hex_pressure_load": (S, [D, C.c_uint32, U32, C.c_uint32, C.c_double, D]),
    "shipfem_case_footprint": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_int32, C.c_int32]),
    "shipfem_batch_create": (S, [C.POINTER(CaseDesc), C.c_uint32, C.POINTER(LaunchConfig), PP,
                                 C.c_char_p, C.c_size_t]),
    "shipfem_hex_batch_create": (S, [C.POINTER(CaseDesc), C.c_uint32, C.POINTER(LaunchConfig), PP,
                                     C.c_char_p, C.c_size_t]),
    "shipfem_batch_run": (S, [VP, C.c_int32, D]),
    "shipfem_batch_tune": (S, [VP, C.c_int32, C.POINTER(LaunchConfig), D]),
    "shipfem_batch_set_config": (S, [VP, C.POINTER(LaunchConfig)]),
    "shipfem_batch_readback": (S, [VP, C.c_uint32, C.POINTER(StateView)]),
    "shipfem_batch_energy": (S, [VP, C.c_uint32, C.POINTER(EnergyRecord), C.c_int32,
                                 C.POINTER(C.c_int32)]),
    "shipfem_batch_tears": (S, [VP, C.c_uint32, C.POINTER(Tear), C.c_uint32,
                                C.POINTER(C.c_uint32)]),
    "shipfem_batch_reset": (S, [VP]),
    "shipfem_batch_destroy": (None, [VP]),
    "shipfem_server_create": (S, [C.POINTER(ServerOpts), PP, C.c_char_p, C.c_size_t]),
    "shipfem_server_submit": (S, [VP, C.POINTER(CaseDesc), C.c_double, C.POINTER(C.c_uint64)]),
    "shipfem_server_submit_hex": (S, [VP, C.POINTER(CaseDesc), C.c_double, C.POINTER(C.c_uint64)]),
    "shipfem_server_poll": (S, [VP, C.POINTER(Result), C.c_uint32, C.POINTER(C.c_uint32)]),
    "shipfem_server_keep": (S, [VP, C.c_uint64]),
    "shipfem_server_fetch": (S, [VP, C.c_uint64, C.POINTER(StateView)]),
    "shipfem_server_drain": (S, [VP, C.c_double]),
    "shipfem_server_stats_get": (S, [VP, C.POINTER(ServerStats)]),
    "shipfem_now": (C.c_double, []),
    "shipfem_server_destroy": (None, [VP]),
}

_lib = None


def lib():
    """The HIP library. Raises when it is not built: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libshipfem.so not built at {LIB_PATH}: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(s):
    try:
        return lib().shipfem_status_string(s).decode()
    except Exception:
        return "?"


def _check(status, msg=""):
    if status != OK:
        raise ShipfemError(status, msg)


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def device_info():
    """(name, hbm_bytes) or None when there is no usable device."""
    buf = C.create_string_buffer(256)
    hbm = C.c_uint64()
    s = lib().shipfem_device_query(buf, 256, C.byref(hbm))
    if s == NO_DEVICE:
        return None
    _check(s, buf.value.decode())
    return buf.value.decode(), hbm.value


def now():
    return lib().shipfem_now()


# ---------------------------------------------------------------- meshes
class Mesh:
    """A case under construction (shipfem_mesh_*). Arrays are views of library memory."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @classmethod
    def box(cls, lx, ly, lz, nx, ny, nz):
        h = C.c_void_p()
        _check(lib().shipfem_mesh_box(lx, ly, lz, nx, ny, nz, C.byref(h)), "mesh_box")
        return cls(h.value)

    @classmethod
    def panel(cls, lx, ly, t_plate, n_stiff, web_h, web_t, nx, ny_bay, nz_web):
        h = C.c_void_p()
        _check(lib().shipfem_mesh_panel(lx, ly, t_plate, n_stiff, web_h, web_t, nx, ny_bay,
                                        nz_web, C.byref(h)), "mesh_panel")
        return cls(h.value)

    def __del__(self):
        try:
            if self.h:
                lib().shipfem_mesh_destroy(self.h)
        except Exception:
            pass

    def jitter(self, amp, seed):
        _check(lib().shipfem_mesh_jitter(self.h, amp, seed), "mesh_jitter")
        return self

    def counts(self):
        n, t = C.c_uint32(), C.c_uint32()
        lib().shipfem_mesh_counts(self.h, C.byref(n), C.byref(t))
        return n.value, t.value

    def arrays(self):
        """Copies: X (N,3) f64, index (T,4) u32, flags (N,) u32, load (N,3) f64."""
        N, T = self.counts()
        x, i, f, l = D(), U32(), U32(), D()
        lib().shipfem_mesh_arrays(self.h, C.byref(x), C.byref(i), C.byref(f), C.byref(l))
        return (np.ctypeslib.as_array(x, shape=(N, 3)).copy(),
                np.ctypeslib.as_array(i, shape=(T, 4)).copy(),
                np.ctypeslib.as_array(f, shape=(N,)).copy(),
                np.ctypeslib.as_array(l, shape=(N, 3)).copy())

    def flag_plane(self, axis, value, flags, tol=1e-9):
        _check(lib().shipfem_mesh_flag_plane(self.h, axis, value, tol, flags))
        return self

    def flag_disc(self, axis, value, c0, c1, radius, flags, tol=1e-9):
        _check(lib().shipfem_mesh_flag_disc(self.h, axis, value, tol, c0, c1, radius, flags))
        return self

    def traction(self, axis, value, traction, tol=1e-9):
        t = (C.c_double * 3)(*traction)
        _check(lib().shipfem_mesh_traction(self.h, axis, value, tol, t))
        return self


class Case:
    """A case description: numpy arrays plus scalars, convertible to CaseDesc."""

    corners = 4  # nodes per element: tets

    def __init__(self, X, index, *, flags=None, load=None, prescribed=None, u0=None, v0=None,
                 E=210e9, nu=0.3, rho=7850.0, plastic=False, yield_=355e6, hardening=1e9,
                 indenter=None, load_ramp=0.0, drive_ramp=0.0, gravity=0.0, damping=1.0,
                 dt=0.0, dt_safety=0.9, sample_every=100, max_steps=0, failure=None, mass=None):
        self.X = np.ascontiguousarray(X, np.float64).reshape(-1, 3)
        self.index = np.ascontiguousarray(index, np.uint32).reshape(-1, self.corners)
        c = lambda a, dt_: None if a is None else np.ascontiguousarray(a, dt_)  # noqa: E731
        self.flags = c(flags, np.uint32)
        self.load = c(load, np.float64)
        self.prescribed = c(prescribed, np.float64)
        self.u0 = c(u0, np.float64)
        self.v0 = c(v0, np.float64)
        self.mass = c(mass, np.float64)  # lumped nodal masses in place of rho V / corners
        self.E, self.nu, self.rho = E, nu, rho
        self.plastic, self.yield_, self.hardening = plastic, yield_, hardening
        self.indenter = indenter
        self.load_ramp, self.drive_ramp = load_ramp, drive_ramp
        self.gravity, self.damping = gravity, damping
        self.dt, self.dt_safety = dt, dt_safety
        self.sample_every, self.max_steps = sample_every, max_steps
        # ductile failure: dict(failure_strain=..., [sensitivity, reference_triaxiality,
        # cutoff_triaxiality]) or None
        self.failure = dict(failure) if failure else None

    @classmethod
    def from_mesh(cls, mesh, **kw):
        X, index, flags, load = mesh.arrays()
        if not np.any(load):
            load = None
        return cls(X, index, flags=flags, load=load, **kw)

    @property
    def N(self):
        return len(self.X)

    @property
    def T(self):
        return len(self.index)

    def desc(self):
        d = CaseDesc()
        d.node_count, d.tet_count = self.N, self.T
        d.rest_position = _p(self.X, C.c_double)
        d.index = _p(self.index, C.c_uint32)
        d.node_flags = _p(self.flags, C.c_uint32)
        d.load = _p(self.load, C.c_double)
        d.prescribed = _p(self.prescribed, C.c_double)
        d.initial_displacement = _p(self.u0, C.c_double)
        d.initial_velocity = _p(self.v0, C.c_double)
        d.mass = _p(self.mass, C.c_double)
        d.fracture_strain = float((self.failure or {}).get("fracture_strain", 0.0))
        f = dict(RICE_TRACEY)
        f.update(self.failure or {})
        d.material = Material(self.E, self.nu, self.rho, 1 if self.plastic else 0, self.yield_,
                              self.hardening, f.get("failure_strain", 0.0) if self.failure else 0.0,
                              f["sensitivity"], f["reference_triaxiality"],
                              f["cutoff_triaxiality"])
        if self.indenter:
            ind = self.indenter
            d.indenter = Indenter(1, ind["radius"], (C.c_double * 3)(*ind["center"]),
                                  (C.c_double * 3)(*ind["velocity"]), ind.get("kappa", 0.1))
        d.load_ramp, d.drive_ramp = self.load_ramp, self.drive_ramp
        d.gravity, d.damping = self.gravity, self.damping
        d.dt, d.dt_safety = self.dt, self.dt_safety
        d.sample_every, d.max_steps = self.sample_every, self.max_steps
        return d

    def encode(self):
        """The case wire format (shipfem_case_encode, shipfem_hex_case_encode for a hex
        case): one self-checking blob."""
        d = self.desc()
        hx = self.corners == 8
        n = (lib().shipfem_hex_case_encoded_size if hx else lib().shipfem_case_encoded_size)(C.byref(d))
        buf = C.create_string_buffer(n)
        _check((lib().shipfem_hex_case_encode if hx else lib().shipfem_case_encode)(C.byref(d), buf, n),
               "encode")
        return buf.raw[:n]

    @classmethod
    def _from_desc(cls, d, corners=4):
        # the blob says which kind of case it holds
        cls = HexCase if corners == 8 else Case
        N, T = d.node_count, d.tet_count
        arr = lambda ptr, n, dt_: None if not ptr else np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt_).copy()  # noqa: E731
        X = arr(d.rest_position, 3 * N, np.float64).reshape(N, 3)
        index = arr(d.index, corners * T, np.uint32).reshape(T, corners)
        vec3 = lambda ptr: None if not ptr else arr(ptr, 3 * N, np.float64).reshape(N, 3)  # noqa: E731
        m, ind = d.material, d.indenter
        failure = None
        if m.plastic and m.failure_strain > 0:
            failure = dict(failure_strain=m.failure_strain, sensitivity=m.triaxiality_sensitivity,
                           reference_triaxiality=m.reference_triaxiality,
                           cutoff_triaxiality=m.cutoff_triaxiality)
            if d.fracture_strain > 0:
                failure["fracture_strain"] = d.fracture_strain
        indenter = None
        if ind.enabled:
            indenter = dict(radius=ind.radius, center=tuple(ind.center), velocity=tuple(ind.velocity),
                            kappa=ind.penalty)
        return cls(X, index, flags=arr(d.node_flags, N, np.uint32), load=vec3(d.load),
                   prescribed=vec3(d.prescribed), u0=vec3(d.initial_displacement),
                   v0=vec3(d.initial_velocity), E=m.youngs_modulus, nu=m.poisson_ratio,
                   rho=m.density, plastic=bool(m.plastic), yield_=m.yield_strength,
                   hardening=m.hardening_modulus, indenter=indenter, load_ramp=d.load_ramp,
                   drive_ramp=d.drive_ramp, gravity=d.gravity, damping=d.damping, dt=d.dt,
                   dt_safety=d.dt_safety, sample_every=d.sample_every, max_steps=d.max_steps,
                   failure=failure, mass=arr(d.mass, N, np.float64))

    @classmethod
    def decode(cls, blob):
        """A case from the wire format (shipfem_case_decode)."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        h, d = C.c_void_p(), CaseDesc()
        _check(lib().shipfem_case_decode(buf, len(blob), C.byref(h), C.byref(d)), "decode")
        try:
            return cls._from_desc(d, lib().shipfem_case_blob_corners(h))
        finally:
            lib().shipfem_case_blob_destroy(h)

    def save(self, path):
        d = self.desc()
        fn = lib().shipfem_hex_case_save if self.corners == 8 else lib().shipfem_case_save
        _check(fn(C.byref(d), os.fsencode(path)), f"save {path}")

    @classmethod
    def load(cls, path):
        h, d = C.c_void_p(), CaseDesc()
        _check(lib().shipfem_case_load(os.fsencode(path), C.byref(h), C.byref(d)), f"load {path}")
        try:
            return cls._from_desc(d, lib().shipfem_case_blob_corners(h))
        finally:
            lib().shipfem_case_blob_destroy(h)

    def oracle_kwargs(self):
        """Keyword arguments of oracle.Case describing the same problem (test use)."""
        ind = None
        if self.indenter:
            ind = dict(self.indenter)
            ind.setdefault("kappa", 0.1)
        return dict(E=self.E, nu=self.nu, rho=self.rho, flags=self.flags, load=self.load,
                    prescribed=self.prescribed, plastic=self.plastic, yield_=self.yield_,
                    hardening=self.hardening, load_ramp=self.load_ramp,
                    drive_ramp=self.drive_ramp, gravity=self.gravity, damping=self.damping,
                    indenter=ind, failure=self.failure, mass=self.mass)


class HexCase(Case):
    """A solid-shell hex case: index [T, 8], nodes 0-3 on the element's -zeta face and 4-7
    on its +zeta face (the thickness direction), each counter-clockwise seen from +zeta
    (solid_shell.hpp:115-125). Elastic, or plastic with `plastic` set."""

    corners = 8

    @classmethod
    def plate(cls, lx, ly, thickness, nx, ny, nz=1, pressure=0.0, **kw):
        """makePlateMesh (solid_shell.cpp:1152-1185), with uniformPressureLoad of
        `pressure` on the top face when nonzero (solid_shell.cpp:1505-1561)."""
        X, index = hex_plate_mesh(lx, ly, thickness, nx, ny, nz)
        load = hex_pressure_load(X, index, pressure) if pressure else None
        return cls(X, index, load=load, **kw)


def hex_plate_mesh(lx, ly, thickness, nx, ny, nz=1):
    n, e = C.c_uint32(), C.c_uint32()
    _check(lib().shipfem_hex_plate_mesh(lx, ly, thickness, nx, ny, nz, C.byref(n), C.byref(e),
                                        None, None), "hex_plate_mesh")
    X = np.zeros((n.value, 3))
    index = np.zeros((e.value, 8), np.uint32)
    _check(lib().shipfem_hex_plate_mesh(lx, ly, thickness, nx, ny, nz, None, None,
                                        _p(X, C.c_double), _p(index, C.c_uint32)))
    return X, index


def hex_pressure_load(X, index, pressure, load=None):
    X = np.ascontiguousarray(X, np.float64)
    index = np.ascontiguousarray(index, np.uint32)
    out = np.zeros((len(X), 3)) if load is None else load
    _check(lib().shipfem_hex_pressure_load(_p(X, C.c_double), len(X), _p(index, C.c_uint32),
                                           len(index), pressure, _p(out, C.c_double)))
    return out


def _nodes24(nodes):
    return np.ascontiguousarray(np.asarray(nodes, np.float64).reshape(24))


def hex_stiffness(nodes, E=210e9, nu=0.3):
    """Condensed elastic stiffness [24, 24] of one solid-shell hex (elementStiffness)."""
    out = np.zeros((24, 24))
    _check(lib().shipfem_hex_stiffness(_p(_nodes24(nodes), C.c_double), E, nu,
                                       _p(out, C.c_double)), "hex_stiffness")
    return out


def hex_mass(nodes, rho=7850.0):
    out = np.zeros(8)
    _check(lib().shipfem_hex_mass(_p(_nodes24(nodes), C.c_double), rho, _p(out, C.c_double)),
           "hex_mass")
    return out


def hex_force(K, rest, current):
    out = np.zeros(24)
    K = np.ascontiguousarray(K, np.float64)
    _check(lib().shipfem_hex_force(_p(K, C.c_double), _p(_nodes24(rest), C.c_double),
                                   _p(_nodes24(current), C.c_double), _p(out, C.c_double)))
    return out


def hex_rotation(rest, current):
    """Centre rotation [3, 3] (row-major R[i, j]) of one element (elementRotation)."""
    out = np.zeros(9)
    _check(lib().shipfem_hex_rotation(_p(_nodes24(rest), C.c_double),
                                      _p(_nodes24(current), C.c_double), _p(out, C.c_double)),
           "hex_rotation")
    return out.reshape(3, 3).T


def hex_plastic_update(rest, current, point_state, enhanced, E=210e9, nu=0.3, yield_strength=355e6,
                       hardening=1e9, failure_strain=0.0, sensitivity=1.5,
                       reference_triaxiality=1 / 3, cutoff_triaxiality=-1 / 3):
    """One element's elastoplastic update (elementPlasticUpdate). `point_state` [64] (per
    point: plastic strain 6, eqps, damage) and `enhanced` [7] are float64 arrays advanced
    in place. Returns (force [24], dissipation, failed points)."""
    for a, n in ((point_state, 64), (enhanced, 7)):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.size == n
                and a.flags.c_contiguous):
            raise ValueError(f"hex_plastic_update: need a contiguous float64 array of {n}")
    out = np.zeros(24)
    diss = C.c_double()
    failed = C.c_int32()
    _check(lib().shipfem_hex_plastic_update(
        _p(_nodes24(rest), C.c_double), _p(_nodes24(current), C.c_double), E, nu, yield_strength,
        hardening, failure_strain, sensitivity, reference_triaxiality, cutoff_triaxiality,
        _p(point_state, C.c_double), _p(enhanced, C.c_double), _p(out, C.c_double),
        C.byref(diss), C.byref(failed)), "hex_plastic_update")
    return out, diss.value, failed.value


def hex_failure_strain(nodes, uniform, fracture):
    """initialisePlasticState's regularised failure strain of one element
    (shipfem_hex_failure_strain)."""
    x = _nodes24(nodes)
    out = C.c_double()
    _check(lib().shipfem_hex_failure_strain(_p(x, C.c_double), uniform, fracture, C.byref(out)))
    return out.value


def hex_critical_timestep(nodes, E=210e9, nu=0.3, rho=7850.0, safety=0.9):
    dt = C.c_double()
    _check(lib().shipfem_hex_critical_timestep(_p(_nodes24(nodes), C.c_double), E, nu, rho, safety,
                                               C.byref(dt)), "hex_critical_timestep")
    return dt.value


def _records(arr, n):
    out = []
    for i in range(n):
        r = arr[i]
        d = {k: getattr(r, k) for k in ["step", "time"] + ENERGY_FIELDS}
        d["contact_force"] = tuple(r.contact_force)
        out.append(d)
    return out


def _tears(call):
    n = C.c_uint32()
    _check(call(None, 0, C.byref(n)), "tears")
    arr = (Tear * max(n.value, 1))()
    _check(call(arr, n.value, C.byref(n)), "tears")
    return [(arr[k].tet, arr[k].step, np.array(arr[k].normal[:])) for k in range(n.value)]


def stable_timestep(X, index, E=210e9, nu=0.3, rho=7850.0, noalloc=False):
    """2 / omega_max of a tet mesh in FP64 on the device (shipfem_stable_timestep; with
    noalloc, shipfem_stable_timestep_noalloc, which allocates nothing)."""
    X = np.ascontiguousarray(X, np.float64)
    idx = np.ascontiguousarray(index, np.uint32)
    cfl, err = C.c_double(), C.create_string_buffer(512)
    fn = lib().shipfem_stable_timestep_noalloc if noalloc else lib().shipfem_stable_timestep
    _check(fn(_p(X, C.c_double), _p(idx, C.c_uint32), len(X), len(idx), E, nu, rho,
                                         C.byref(cfl), err, 512), err.value.decode())
    return cfl.value


def cpu_tet_forces(position, index, rest_inv, rest_volume, E=210e9, nu=0.3, rest=None):
    """tetForces (fem.hpp:84-86) of every tet on the host, FP32: [T, 12]. position [N, 3]
    float32, rest_inv [T, 9] column-major and rest_volume [T] as computeRestState gives
    them; rest [N, 3] double selects the displacement form."""
    pos = np.ascontiguousarray(position, np.float32)
    idx = np.ascontiguousarray(index, np.uint32)
    ri = np.ascontiguousarray(rest_inv, np.float32)
    rv = np.ascontiguousarray(rest_volume, np.float32)
    rs = None if rest is None else np.ascontiguousarray(rest, np.float64)
    lam = np.float32(E) * np.float32(nu) / ((np.float32(1) + np.float32(nu)) * (np.float32(1) - np.float32(2) * np.float32(nu)))
    mu = np.float32(E) / (np.float32(2) * (np.float32(1) + np.float32(nu)))
    out = np.zeros((len(idx), 12), np.float32)
    for t in range(len(idx)):
        _check(lib().shipfem_cpu_tet_forces(_p(pos, C.c_float), _p(rs, C.c_double), _p(idx, C.c_uint32),
                                            _p(ri, C.c_float), _p(rv, C.c_float), t, float(lam),
                                            float(mu), _p(out[t], C.c_float)))
    return out


def cpu_step(state, dt, gravity=0.0, damping=1.0, E=210e9, nu=0.3):
    """stepCpu (fem.hpp:77-82) on the host, in place on a dict of TetMesh arrays:
    position [N, 3] and velocity [N, 3] float32, mass, fixed, index, rest_inv,
    rest_volume, tet_force, adj_offset, adj_entry and optionally rest [N, 3] double."""
    g = state
    lam = np.float32(E) * np.float32(nu) / ((np.float32(1) + np.float32(nu)) * (np.float32(1) - np.float32(2) * np.float32(nu)))
    mu = np.float32(E) / (np.float32(2) * (np.float32(1) + np.float32(nu)))
    N, T = len(g["position"]), len(g["index"])
    _check(lib().shipfem_cpu_step(N, T, _p(g["position"], C.c_float), _p(g["velocity"], C.c_float),
                                  _p(g["mass"], C.c_float), _p(g.get("fixed"), C.c_uint32),
                                  _p(g["index"], C.c_uint32), _p(g["rest_inv"], C.c_float),
                                  _p(g["rest_volume"], C.c_float), _p(g["tet_force"], C.c_float),
                                  _p(g["adj_offset"], C.c_uint32), _p(g["adj_entry"], C.c_uint32),
                                  _p(g.get("rest"), C.c_double), float(lam), float(mu), dt, gravity,
                                  damping))


def regularised_failure_strain(uniform, fracture, length, thickness):
    """uniform + (fracture - uniform) min(thickness / length, 1) (plasticity.cpp:158-169)."""
    return lib().shipfem_regularised_failure_strain(uniform, fracture, length, thickness)


def torn_groups(X, index, tet_group, damage, fraction=0.5):
    """Group ids whose torn volume reaches `fraction` of their volume (zone.cpp:1262-1275)."""
    X = np.ascontiguousarray(X, np.float64)
    index = np.ascontiguousarray(index, np.uint32)
    g = np.ascontiguousarray(tet_group, np.uint32)
    dmg = np.ascontiguousarray(damage, np.float32)
    n = C.c_uint32()
    T = len(index)
    _check(lib().shipfem_torn_groups(_p(X, C.c_double), _p(index, C.c_uint32), T,
                                     _p(g, C.c_uint32), _p(dmg, C.c_float), fraction, None, 0,
                                     C.byref(n)))
    out = np.zeros(max(n.value, 1), np.uint32)
    _check(lib().shipfem_torn_groups(_p(X, C.c_double), _p(index, C.c_uint32), T,
                                     _p(g, C.c_uint32), _p(dmg, C.c_float), fraction,
                                     _p(out, C.c_uint32), n.value, C.byref(n)))
    return out[:n.value].tolist()


def apply_edge_drive(flags, prescribed, displacement, driven):
    """coupling::applyEdgeDrive (coupling.cpp:110-125) on a case's node flags [N] and
    prescribed displacement [N, 3], in place: every DOF with driven[n, a] != 0 is driven
    to displacement[n, a] (its PIN bit cleared, DRIVE bit set); returns the DOF count."""
    assert flags.dtype == np.uint32 and flags.flags.c_contiguous
    assert prescribed.dtype == np.float64 and prescribed.flags.c_contiguous
    disp = np.ascontiguousarray(displacement, np.float64)
    drv = np.ascontiguousarray(driven, np.uint8)
    N = len(flags)
    assert prescribed.size == 3 * N and disp.size == 3 * N and drv.size == 3 * N
    return lib().shipfem_apply_edge_drive(N, _p(disp, C.c_double), _p(drv, C.c_uint8),
                                          _p(flags, C.c_uint32), _p(prescribed, C.c_double))


class State:
    def __init__(self, N, T):
        self.u = np.zeros((N, 3))
        self.v = np.zeros((N, 3))
        self.position = np.zeros((N, 3), np.float32)
        self.eqps = np.zeros(T, np.float32)
        self.plastic_strain = np.zeros((T, 6), np.float32)
        self.damage = np.zeros(T, np.float32)
        self.step = 0
        self.time = 0.0
        self.status = 0

    def view(self):
        v = StateView()
        v.displacement = _p(self.u, C.c_double)
        v.velocity = _p(self.v, C.c_double)
        v.position = _p(self.position, C.c_float)
        v.eqps = _p(self.eqps, C.c_float)
        v.plastic_strain = _p(self.plastic_strain, C.c_float)
        v.damage = _p(self.damage, C.c_float)
        return v

    def take(self, v):
        self.step, self.time, self.status = v.step, v.time, v.status
        return self


class Solver:
    """One case on the device (replaces gpu::FemGpuSolver; a HexCase runs the elastic
    solid-shell solve of zone::Solver)."""

    def __init__(self, case):
        self.case = case
        self._desc = case.desc()
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        create = lib().shipfem_hex_solver_create if case.corners == 8 else lib().shipfem_solver_create
        _check(create(C.byref(self._desc), C.byref(h), err, 512), err.value.decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().shipfem_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timestep(self):
        cfl, dt, edge = C.c_double(), C.c_double(), C.c_double()
        _check(lib().shipfem_solver_timestep(self.h, C.byref(cfl), C.byref(dt), C.byref(edge)))
        return dict(cfl=cfl.value, dt=dt.value, edge_dt=edge.value)

    @property
    def dt(self):
        return self.timestep()["dt"]

    def set_launch(self, variant=VARIANT_TILED, node_block=256, elem_block=256, xcd_remap=1,
                   rest_from_positions=0, resident=0):
        cfg = LaunchConfig(elem_block, node_block, xcd_remap, 0, variant, 0, rest_from_positions, resident)
        _check(lib().shipfem_solver_set_launch(self.h, C.byref(cfg)), "set_launch")
        return self

    def set_params(self, dt, gravity=0.0, damping=1.0, E=0.0, nu=0.0):
        p = StepParams(dt, gravity, damping, E, nu)
        _check(lib().shipfem_solver_set_params(self.h, C.byref(p)))

    def run(self, steps, mode=0, allow_unstable=False):
        ms = C.c_double()
        s = lib().shipfem_solver_run(self.h, steps, mode, C.byref(ms))
        if not (allow_unstable and s == UNSTABLE):
            _check(s, "run")
        return ms.value

    def readback(self):
        st = State(self.case.N, self.case.T)
        v = st.view()
        _check(lib().shipfem_solver_readback(self.h, C.byref(v)))
        return st.take(v)

    def energy(self):
        n = C.c_int32()
        _check(lib().shipfem_solver_energy(self.h, None, 0, C.byref(n)))
        arr = (EnergyRecord * max(n.value, 1))()
        _check(lib().shipfem_solver_energy(self.h, arr, n.value, C.byref(n)))
        return _records(arr, n.value)

    def tears(self):
        """[(tet, step, normal[3])] ordered by step then tet."""
        return _tears(lambda out, mx, n: lib().shipfem_solver_tears(self.h, out, mx, n))

    def tet_forces(self):
        """Per-element nodal forces of the last element pass: [T, 12] (tets) or [T, 24]
        (hexes)."""
        out = np.zeros((self.case.T, 3 * self.case.corners), np.float32)
        _check(lib().shipfem_solver_tet_forces(self.h, _p(out, C.c_float)))
        return out

    def footprint(self):
        b = C.c_uint64()
        _check(lib().shipfem_solver_footprint(self.h, C.byref(b)))
        return b.value

    def profile(self, steps):
        """Mean (element-pass ms, node-pass ms) over `steps` event-bracketed steps."""
        e, n = C.c_double(), C.c_double()
        _check(lib().shipfem_solver_profile(self.h, steps, C.byref(e), C.byref(n)))
        return e.value, n.value

    def traffic(self):
        """The bytes one step's passes must move: dict elem_read, elem_write, node_read,
        node_write (shipfem_solver_traffic)."""
        out = (C.c_uint64 * 4)()
        _check(lib().shipfem_solver_traffic(self.h, out))
        return dict(zip(("elem_read", "elem_write", "node_read", "node_write"), list(out)))

    def export(self):
        """The state in the shape zone::Solver::adopt takes (ZoneGpuState): dict with
        position [N,3], displacement [N,3], velocity [N,3], plastic [T,8] (eps_p 6, eqps, damage), tear_step
        [T] (-1 intact), tear_stress [T,6] and the scalars steps, time, penetration, work,
        w_ext, w_contact, w_boundary, damping_loss, dissipation, fracture, torn. A plastic hex
        case adds hex_points [T,8,8] (per point: eps_p 6, eqps, damage) and hex_enhanced
        [T,7]; its plastic rows are the per-element summary (0 x 6, largest eqps, smallest
        damage)."""
        N, T = self.case.N, self.case.T
        out = dict(position=np.zeros((N, 3)), displacement=np.zeros((N, 3)),
                   velocity=np.zeros((N, 3)),
                   plastic=np.zeros((T, 8)), tear_step=np.zeros(T, np.int64),
                   tear_stress=np.zeros((T, 6)))
        h = Handoff()
        if isinstance(self.case, HexCase) and self.case.plastic:
            out["hex_points"], out["hex_enhanced"] = np.zeros((T, 8, 8)), np.zeros((T, 7))
            h.hex_points = _p(out["hex_points"], C.c_double)
            h.hex_enhanced = _p(out["hex_enhanced"], C.c_double)
        h.position, h.velocity = _p(out["position"], C.c_double), _p(out["velocity"], C.c_double)
        h.displacement = _p(out["displacement"], C.c_double)
        h.plastic = _p(out["plastic"], C.c_double)
        h.tear_step = _p(out["tear_step"], C.c_int64)
        h.tear_stress = _p(out["tear_stress"], C.c_double)
        _check(lib().shipfem_solver_export(self.h, C.byref(h)), "export")
        for k in HANDOFF_SCALARS:
            out[k] = getattr(h, k)
        return out

    def adopt(self, state):
        """zone::Solver::adopt on this solver: continue from a state made elsewhere
        (a dict as export() returns; plastic, tear_step, tear_stress, hex_points and
        hex_enhanced optional)."""
        keep = {}
        h = Handoff()
        for k, ct in (("position", C.c_double), ("displacement", C.c_double),
                      ("velocity", C.c_double), ("plastic", C.c_double),
                      ("tear_step", C.c_int64), ("tear_stress", C.c_double),
                      ("hex_points", C.c_double), ("hex_enhanced", C.c_double)):
            a = state.get(k)
            if a is not None:
                keep[k] = np.ascontiguousarray(a, np.int64 if k == "tear_step" else np.float64)
                setattr(h, k, _p(keep[k], ct))
        for k in HANDOFF_SCALARS:
            if k in state:
                setattr(h, k, state[k])
        _check(lib().shipfem_solver_adopt(self.h, C.byref(h)), "adopt")
        return self

    def checkpoint(self):
        """Bitwise checkpoint of the evolving device state, as bytes."""
        n = C.c_uint64()
        _check(lib().shipfem_solver_checkpoint(self.h, None, C.byref(n)), "checkpoint")
        buf = C.create_string_buffer(n.value)
        _check(lib().shipfem_solver_checkpoint(self.h, buf, C.byref(n)), "checkpoint")
        return buf.raw[:n.value]

    def restore(self, blob):
        buf = C.create_string_buffer(blob, len(blob))
        _check(lib().shipfem_solver_restore(self.h, buf, len(blob)), "restore")
        return self


class Batch:
    """Many cases advancing in one launch per phase. split=0: two or more cases run as parts
    (up to four, balanced by element count) on as many streams side by side; split=-1: one
    stream."""

    def __init__(self, cases, elem_block=256, node_block=256, xcd_remap=1, chunk=0,
                 variant=VARIANT_TILED, split=0, rest_from_positions=0):
        self.cases = list(cases)
        self._descs = (CaseDesc * len(self.cases))(*[c.desc() for c in self.cases])
        cfg = LaunchConfig(elem_block, node_block, xcd_remap, chunk, variant, split, rest_from_positions)
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        kinds = {c.corners for c in self.cases}
        if len(kinds) != 1:
            raise ValueError("a batch holds one element kind (all tets or all hexes)")
        create = lib().shipfem_hex_batch_create if kinds == {8} else lib().shipfem_batch_create
        _check(create(self._descs, len(self.cases), C.byref(cfg), C.byref(h), err, 512),
               err.value.decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().shipfem_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, steps):
        ms = C.c_double()
        _check(lib().shipfem_batch_run(self.h, steps, C.byref(ms)))
        return ms.value

    def profile(self, steps):
        """Mean (element-pass ms, node-pass ms) over `steps` (whole chunks) event-bracketed
        steps (shipfem_batch_profile)."""
        e, n = C.c_double(), C.c_double()
        _check(lib().shipfem_batch_profile(self.h, steps, C.byref(e), C.byref(n)))
        return e.value, n.value

    def footprint(self):
        b = C.c_uint64()
        _check(lib().shipfem_batch_footprint(self.h, C.byref(b)))
        return b.value

    def traffic(self):
        """shipfem_batch_traffic: bytes one step's passes must move, over all cases."""
        out = (C.c_uint64 * 4)()
        _check(lib().shipfem_batch_traffic(self.h, out))
        return dict(zip(("elem_read", "elem_write", "node_read", "node_write"), list(out)))

    def tune(self, steps_per_trial=200):
        cfg = LaunchConfig()
        ms = C.c_double()
        _check(lib().shipfem_batch_tune(self.h, steps_per_trial, C.byref(cfg), C.byref(ms)))
        return dict(elem_block=cfg.elem_block, node_block=cfg.node_block,
                    xcd_remap=cfg.xcd_remap, chunk=cfg.chunk, variant=cfg.variant, split=cfg.split,
                    rest_from_positions=cfg.rest_from_positions,
                    ms_per_step=ms.value)

    def set_config(self, elem_block, node_block, xcd_remap, variant=VARIANT_TILED, split=0,
                   rest_from_positions=0):
        cfg = LaunchConfig(elem_block, node_block, xcd_remap, 0, variant, split, rest_from_positions)
        _check(lib().shipfem_batch_set_config(self.h, C.byref(cfg)))

    def reset(self):
        _check(lib().shipfem_batch_reset(self.h))

    def readback(self, i):
        c = self.cases[i]
        st = State(c.N, c.T)
        v = st.view()
        _check(lib().shipfem_batch_readback(self.h, i, C.byref(v)))
        return st.take(v)

    def energy(self, i):
        n = C.c_int32()
        _check(lib().shipfem_batch_energy(self.h, i, None, 0, C.byref(n)))
        arr = (EnergyRecord * max(n.value, 1))()
        _check(lib().shipfem_batch_energy(self.h, i, arr, n.value, C.byref(n)))
        return _records(arr, n.value)

    def tears(self, i):
        return _tears(lambda out, mx, n: lib().shipfem_batch_tears(self.h, i, out, mx, n))


class Server:
    """Resident serving loop: submit cases at any time, poll results."""

    def __init__(self, memory_budget=0, max_active=0, chunk=0, elem_block=0, node_block=0,
                 xcd_remap=1, variant=VARIANT_TILED, policy=ROUNDS_ALL, round_work=0.0, starve_rounds=0):
        opts = ServerOpts(memory_budget, max_active, chunk,
                          LaunchConfig(elem_block, node_block, xcd_remap, 0, variant), policy, round_work,
                          starve_rounds)
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        _check(lib().shipfem_server_create(C.byref(opts), C.byref(h), err, 512), err.value.decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().shipfem_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, case, run_to, keep=False):
        d = case.desc()
        rid = C.c_uint64()
        fn = lib().shipfem_server_submit_hex if case.corners == 8 else lib().shipfem_server_submit
        _check(fn(self.h, C.byref(d), run_to, C.byref(rid)), "submit")
        if keep:
            _check(lib().shipfem_server_keep(self.h, rid.value))
        return rid.value

    def submit_encoded(self, blob, run_to, keep=False):
        """submit() of a case in the wire format (Case.encode / a case file's bytes)."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        rid = C.c_uint64()
        _check(lib().shipfem_server_submit_encoded(self.h, buf, len(blob), run_to, C.byref(rid)),
               "submit_encoded")
        if keep:
            _check(lib().shipfem_server_keep(self.h, rid.value))
        return rid.value

    def poll(self, max_results=4096):
        arr = (Result * max_results)()
        n = C.c_uint32()
        _check(lib().shipfem_server_poll(self.h, arr, max_results, C.byref(n)))
        out = []
        for i in range(n.value):
            r = arr[i]
            d = {k: getattr(r, k) for k, _ in Result._fields_ if k != "last"}
            d["last"] = _records([r.last], 1)[0]
            out.append(d)
        return out

    def drain(self, timeout=60.0):
        return lib().shipfem_server_drain(self.h, timeout) == OK

    def fetch(self, rid, case):
        st = State(case.N, case.T)
        v = st.view()
        _check(lib().shipfem_server_fetch(self.h, rid, C.byref(v)), f"fetch {rid}")
        return st.take(v)

    def stats(self):
        s = ServerStats()
        _check(lib().shipfem_server_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}


def case_footprint(N, T, plastic=False, records=64):
    # plastic: False/0 elastic, True/1 J2, 2 J2 with failure
    return lib().shipfem_case_footprint(N, T, int(plastic), records)


def case_footprint_mesh(case, records=64):
    """An upper bound of a tet case's device bytes from its own mesh (node degree)."""
    d = case.desc()
    return lib().shipfem_case_footprint_mesh(C.byref(d), records)

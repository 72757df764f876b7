"""ctypes bindings of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports this.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

PIN_X, PIN_Y, PIN_Z = 1, 2, 4
PIN_ALL = 7
LOADED, CONTACT = 8, 16
DRIVE_X, DRIVE_Y, DRIVE_Z = 32, 64, 128

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
U32 = C.POINTER(C.c_uint32)
I64 = C.POINTER(C.c_int64)


class PanelSpec(C.Structure):
    _fields_ = [("lx", C.c_double), ("ly", C.c_double), ("t_plate", C.c_double),
                ("n_stiff", C.c_int), ("web_h", C.c_double), ("web_t", C.c_double),
                ("nx", C.c_int), ("ny_bay", C.c_int), ("nz_web", C.c_int)]


class J2(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("yield_", C.c_double),
                ("hardening", C.c_double), ("failure_strain", C.c_double),
                ("sensitivity", C.c_double), ("reference_triaxiality", C.c_double),
                ("cutoff_triaxiality", C.c_double)]


# Rice-Tracey constants of the reference's Failure defaults (plasticity.hpp:193-199)
RICE_TRACEY = dict(sensitivity=1.5, reference_triaxiality=1.0 / 3.0,
                   cutoff_triaxiality=-1.0 / 3.0)


class CaseStruct(C.Structure):
    _fields_ = [("N", C.c_int64), ("T", C.c_int64), ("X", D), ("index", U32), ("flags", U32),
                ("load", D), ("prescribed", D), ("E", C.c_double), ("nu", C.c_double),
                ("rho", C.c_double), ("plastic", C.c_int), ("yield_", C.c_double),
                ("hardening", C.c_double), ("load_ramp", C.c_double), ("drive_ramp", C.c_double),
                ("gravity", C.c_double), ("damping", C.c_double),
                ("failure_strain", C.c_double), ("sensitivity", C.c_double),
                ("reference_triaxiality", C.c_double), ("cutoff_triaxiality", C.c_double),
                ("indenter", C.c_int),
                ("ind_radius", C.c_double), ("ind_center", C.c_double * 3),
                ("ind_velocity", C.c_double * 3), ("ind_kappa", C.c_double), ("dt", C.c_double),
                ("rest_inv", D), ("volume", D), ("mass", D), ("adj_offset", U32),
                ("adj_entry", U32), ("corners", C.c_int), ("hex_k", D), ("hex_mass", D),
                ("hex_B", D), ("hex_G", D), ("hex_w", D), ("hex_failure", D)]


class StateStruct(C.Structure):
    _fields_ = [("u", D), ("v", D), ("eps_p", D), ("eqps", D), ("damage", D),
                ("tear_step", I64), ("tear_normal", D), ("force", D), ("step", C.c_int64),
                ("w_ext", C.c_double), ("w_contact", C.c_double), ("w_boundary", C.c_double),
                ("dissipation", C.c_double), ("damping_loss", C.c_double),
                ("fracture", C.c_double), ("hex_point", D), ("hex_enhanced", D)]


class EnergyStruct(C.Structure):
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("strain", C.c_double),
                ("kinetic", C.c_double), ("dissipation", C.c_double), ("w_ext", C.c_double),
                ("w_contact", C.c_double), ("w_boundary", C.c_double),
                ("damping_loss", C.c_double), ("fracture", C.c_double),
                ("contact_force", C.c_double * 3)]


ENERGY_FIELDS = ["strain", "kinetic", "dissipation", "w_ext", "w_contact", "w_boundary",
                 "damping_loss", "fracture"]


def _declare(L):
    L.ora_hex_force.argtypes = [D, D, D, D]
    L.ora_hex_plastic_update.argtypes = [D, D, D, D, D, C.POINTER(J2), D, D, D, D, D]
    L.ora_hex_plastic_update.restype = C.c_int
    L.ora_box_mesh.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [I64, I64, D, U32]
    L.ora_box_mesh_f32.argtypes = [C.c_float] * 3 + [C.c_int] * 3 + [F, U32]
    L.ora_uniform.argtypes = [C.c_uint64, C.c_uint64]
    L.ora_uniform.restype = C.c_double
    L.ora_box_jitter.argtypes = [D, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_uint64]
    L.ora_panel_mesh.argtypes = [C.POINTER(PanelSpec), I64, I64, D, U32]
    L.ora_rest_state.argtypes = [D, C.c_int64, U32, C.c_int64, C.c_double, D, D, D, U32, U32]
    L.ora_edge_timestep.argtypes = [D, U32, C.c_int64, C.c_double, C.c_double, C.c_double]
    L.ora_edge_timestep.restype = C.c_double
    L.ora_max_element_omega.argtypes = [D, C.c_int64, C.c_double, C.c_double, C.c_double]
    L.ora_max_element_omega.restype = C.c_double
    L.ora_face_traction.argtypes = [D, U32, C.c_int64, C.c_int, C.c_double, C.c_double, D, D]
    L.ora_tet_force.argtypes = [D, D, C.c_double, C.c_double, C.c_double, D]
    L.ora_j2_update.argtypes = [C.POINTER(J2), D, D, D, D, D]
    L.ora_j2_update.restype = C.c_int
    L.ora_j2_update_damage.argtypes = [C.POINTER(J2), D, D, D, D, D, D, D, D]
    L.ora_j2_update_damage.restype = C.c_int
    L.ora_regularised_failure_strain.argtypes = [C.c_double] * 4
    L.ora_regularised_failure_strain.restype = C.c_double
    L.ora_triaxiality_factor.argtypes = [C.c_double] * 4
    L.ora_triaxiality_factor.restype = C.c_double
    L.ora_case_init.argtypes = [C.POINTER(CaseStruct)]
    L.ora_case_init.restype = C.c_int
    L.ora_case_free.argtypes = [C.POINTER(CaseStruct)]
    L.ora_run.argtypes = [C.POINTER(CaseStruct), C.POINTER(StateStruct), C.c_int64, C.c_int64,
                          C.POINTER(EnergyStruct), C.c_int64, I64]
    L.ora_run_mt.argtypes = L.ora_run.argtypes + [C.c_int]
    L.ref32_rest_state.argtypes = [F, C.c_int64, U32, C.c_int64, C.c_float, F, F, F, U32, U32]
    L.ref32_tet_forces.argtypes = [F, U32, F, F, C.c_int64, C.c_float, C.c_float, F]
    step_args = [F, F, F, U32, C.c_int64, U32, F, F, C.c_int64, F, U32, U32, C.c_float,
                 C.c_float, C.c_float, C.c_float, C.c_float]
    L.ref32_step.argtypes = step_args
    L.ref32_step_mt.argtypes = step_args + [C.c_int]


# ---------------------------------------------------------------- meshes
def box_mesh(lx, ly, lz, nx, ny, nz):
    L = lib()
    n, t = C.c_int64(), C.c_int64()
    L.ora_box_mesh(lx, ly, lz, nx, ny, nz, C.byref(n), C.byref(t), None, None)
    X = np.zeros((n.value, 3))
    index = np.zeros((t.value, 4), np.uint32)
    L.ora_box_mesh(lx, ly, lz, nx, ny, nz, None, None, _p(X, C.c_double), _p(index, C.c_uint32))
    return X, index


def box_mesh_f32(lx, ly, lz, nx, ny, nz):
    N = (nx + 1) * (ny + 1) * (nz + 1)
    X = np.zeros((N, 3), np.float32)
    index = np.zeros((nx * ny * nz * 6, 4), np.uint32)
    lib().ora_box_mesh_f32(lx, ly, lz, nx, ny, nz, _p(X, C.c_float), _p(index, C.c_uint32))
    return X, index


def uniform(seed, counter):
    return lib().ora_uniform(seed, counter)


def box_jitter(X, nx, ny, nz, lx, ly, lz, amp, seed):
    X = np.ascontiguousarray(X, dtype=np.float64).copy()
    lib().ora_box_jitter(_p(X, C.c_double), nx, ny, nz, lx, ly, lz, amp, seed)
    return X


def panel_mesh(lx, ly, t_plate, n_stiff, web_h, web_t, nx, ny_bay, nz_web):
    spec = PanelSpec(lx, ly, t_plate, n_stiff, web_h, web_t, nx, ny_bay, nz_web)
    n, t = C.c_int64(), C.c_int64()
    L = lib()
    L.ora_panel_mesh(C.byref(spec), C.byref(n), C.byref(t), None, None)
    X = np.zeros((n.value, 3))
    index = np.zeros((t.value, 4), np.uint32)
    L.ora_panel_mesh(C.byref(spec), None, None, _p(X, C.c_double), _p(index, C.c_uint32))
    return X, index


# ---------------------------------------------------------------- rest state
def rest_state(X, index, rho):
    X = np.ascontiguousarray(X, np.float64)
    index = np.ascontiguousarray(index, np.uint32)
    N, T = len(X), len(index)
    out = dict(rest_inv=np.zeros((T, 9)), volume=np.zeros(T), mass=np.zeros(N),
               adj_offset=np.zeros(N + 1, np.uint32), adj_entry=np.zeros(4 * T, np.uint32))
    lib().ora_rest_state(_p(X, C.c_double), N, _p(index, C.c_uint32), T, rho,
                         _p(out["rest_inv"], C.c_double), _p(out["volume"], C.c_double),
                         _p(out["mass"], C.c_double), _p(out["adj_offset"], C.c_uint32),
                         _p(out["adj_entry"], C.c_uint32))
    return out


def edge_timestep(X, index, E, rho, safety=0.5):
    X = np.ascontiguousarray(X, np.float64)
    index = np.ascontiguousarray(index, np.uint32)
    return lib().ora_edge_timestep(_p(X, C.c_double), _p(index, C.c_uint32), len(index), E, rho,
                                   safety)


def max_element_omega(rest_inv, E, nu, rho):
    rest_inv = np.ascontiguousarray(rest_inv, np.float64)
    return lib().ora_max_element_omega(_p(rest_inv, C.c_double), len(rest_inv), E, nu, rho)


def face_traction(X, index, axis, value, tol, traction, load=None):
    X = np.ascontiguousarray(X, np.float64)
    index = np.ascontiguousarray(index, np.uint32)
    if load is None:
        load = np.zeros((len(X), 3))
    tr = np.ascontiguousarray(traction, np.float64)
    lib().ora_face_traction(_p(X, C.c_double), _p(index, C.c_uint32), len(index), axis, value,
                            tol, _p(tr, C.c_double), _p(load, C.c_double))
    return load


def tet_force(x12, dmi9, volume, lam, mu):
    x12 = np.ascontiguousarray(x12, np.float64)
    dmi9 = np.ascontiguousarray(dmi9, np.float64)
    out = np.zeros(12)
    lib().ora_tet_force(_p(x12, C.c_double), _p(dmi9, C.c_double), volume, lam, mu,
                        _p(out, C.c_double))
    return out


def j2_update(E, nu, yield_, hardening, strain, eps_p=None, eqps=0.0):
    m = J2(E, nu, yield_, hardening, 0.0, 0.0, 0.0, 0.0)
    strain = np.ascontiguousarray(strain, np.float64)
    ep = np.zeros(6) if eps_p is None else np.array(eps_p, np.float64)
    q = np.array([eqps], np.float64)
    stress = np.zeros(6)
    diss = np.zeros(1)
    y = lib().ora_j2_update(C.byref(m), _p(strain, C.c_double), _p(ep, C.c_double),
                            _p(q, C.c_double), _p(stress, C.c_double), _p(diss, C.c_double))
    return dict(yielded=bool(y), stress=stress, eps_p=ep, eqps=float(q[0]), dissipation=float(diss[0]))


def j2_update_damage(E, nu, yield_, hardening, failure_strain, strain, eps_p=None, eqps=0.0,
                     damage=0.0, sensitivity=1.5, reference_triaxiality=1.0 / 3.0,
                     cutoff_triaxiality=-1.0 / 3.0):
    m = J2(E, nu, yield_, hardening, failure_strain, sensitivity, reference_triaxiality,
           cutoff_triaxiality)
    strain = np.ascontiguousarray(strain, np.float64)
    ep = np.zeros(6) if eps_p is None else np.array(eps_p, np.float64)
    q = np.array([eqps], np.float64)
    dmg = np.array([damage], np.float64)
    stress, diss, nrm, eta = np.zeros(6), np.zeros(1), np.zeros(3), np.zeros(1)
    r = lib().ora_j2_update_damage(C.byref(m), _p(strain, C.c_double), _p(ep, C.c_double),
                                   _p(q, C.c_double), _p(dmg, C.c_double), _p(stress, C.c_double),
                                   _p(diss, C.c_double), _p(nrm, C.c_double), _p(eta, C.c_double))
    return dict(yielded=bool(r & 1), failed_now=bool(r & 2), stress=stress, eps_p=ep,
                eqps=float(q[0]), damage=float(dmg[0]), dissipation=float(diss[0]), normal=nrm,
                eta=float(eta[0]))


def regularised_failure_strain(uniform, fracture, length, thickness):
    return lib().ora_regularised_failure_strain(uniform, fracture, length, thickness)


def triaxiality_factor(eta, sensitivity=1.5, reference=1.0 / 3.0, cutoff=-1.0 / 3.0):
    return lib().ora_triaxiality_factor(sensitivity, reference, cutoff, eta)


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


# ---------------------------------------------------------------- solver
class Case:
    """FP64 oracle solve of one case. Arrays are numpy; kept alive on the object."""

    def __init__(self, X, index, *, E=210e9, nu=0.3, rho=7850.0, dt, flags=None, load=None,
                 prescribed=None, plastic=False, yield_=355e6, hardening=1e9, load_ramp=0.0,
                 drive_ramp=0.0, gravity=0.0, damping=1.0, indenter=None, failure=None,
                 mass=None):
        """index [T, 4] solves tets; index [T, 8] solves solid-shell hexes (elastic), their
        condensed stiffness and masses from oracle.hex. `mass` [N] replaces the lumped
        nodal masses the rest state computes from rho."""
        self.X = np.ascontiguousarray(X, np.float64)
        self.index = np.ascontiguousarray(index, np.uint32)
        self.N, self.T = len(self.X), len(self.index)
        self.corners = self.index.shape[1]
        if self.corners == 8:
            from oracle import hex as H
            nodes = self.X[self.index]
            self.hex_mass = np.ascontiguousarray(H.mass(nodes, rho))
            if plastic:
                B, G, w = H.forms(nodes)
                self.hex_B, self.hex_G = np.ascontiguousarray(B), np.ascontiguousarray(G)
                self.hex_w = np.ascontiguousarray(w)
                self.hex_k = None
            else:
                self.hex_k = np.ascontiguousarray(H.stiffness(nodes, E, nu))
            self.hex_point = np.zeros((self.T, 8, 9))
            self.hex_enhanced = np.zeros((self.T, 7))
        self.flags = None if flags is None else np.ascontiguousarray(flags, np.uint32)
        self.load = None if load is None else np.ascontiguousarray(load, np.float64)
        self.prescribed = None if prescribed is None else np.ascontiguousarray(prescribed, np.float64)
        s = CaseStruct()
        s.N, s.T = self.N, self.T
        s.X, s.index = _p(self.X, C.c_double), _p(self.index, C.c_uint32)
        s.flags = _p(self.flags, C.c_uint32)
        s.load = _p(self.load, C.c_double)
        s.prescribed = _p(self.prescribed, C.c_double)
        s.E, s.nu, s.rho = E, nu, rho
        s.plastic = 1 if plastic else 0
        s.yield_, s.hardening = yield_, hardening
        s.load_ramp, s.drive_ramp = load_ramp, drive_ramp
        s.gravity, s.damping = gravity, damping
        if failure:
            f = dict(RICE_TRACEY)
            f.update(failure)
            s.failure_strain = f["failure_strain"]
            if self.corners == 8 and plastic and f.get("fracture_strain", 0.0) > 0:
                # every element resolves its own failure strain from its size
                from oracle import hex as H
                self.hex_failure = np.ascontiguousarray(
                    H.failure_strain(self.X[self.index], f["failure_strain"], f["fracture_strain"]))
                s.hex_failure = _p(self.hex_failure, C.c_double)
            s.sensitivity = f["sensitivity"]
            s.reference_triaxiality = f["reference_triaxiality"]
            s.cutoff_triaxiality = f["cutoff_triaxiality"]
        if indenter:
            s.indenter = 1
            s.ind_radius = indenter["radius"]
            s.ind_center = (C.c_double * 3)(*indenter["center"])
            s.ind_velocity = (C.c_double * 3)(*indenter["velocity"])
            s.ind_kappa = indenter.get("kappa", 0.1)
        s.dt = dt
        if self.corners == 8:
            s.corners = 8
            s.hex_k, s.hex_mass = _p(self.hex_k, C.c_double), _p(self.hex_mass, C.c_double)
            if plastic:
                s.hex_B, s.hex_G = _p(self.hex_B, C.c_double), _p(self.hex_G, C.c_double)
                s.hex_w = _p(self.hex_w, C.c_double)
        if lib().ora_case_init(C.byref(s)) != 0:
            raise MemoryError("oracle case init")
        self.s = s
        if mass is not None:
            np.ctypeslib.as_array(s.mass, shape=(self.N,))[:] = np.asarray(mass, np.float64)
        self.mass = np.ctypeslib.as_array(s.mass, shape=(self.N,)).copy()
        self.volume = np.ctypeslib.as_array(s.volume, shape=(self.T,)).copy()
        self.u = np.zeros((self.N, 3))
        self.v = np.zeros((self.N, 3))
        self.eps_p = np.zeros((self.T, 6))
        self.eqps = np.zeros(self.T)
        self.force = np.zeros((self.T, 3 * self.corners))
        self.damage = np.zeros(self.T)
        self.tear_step = np.full(self.T, -1, np.int64)
        self.tear_normal = np.zeros((self.T, 3))
        st = StateStruct()
        st.u, st.v = _p(self.u, C.c_double), _p(self.v, C.c_double)
        st.eps_p, st.eqps = _p(self.eps_p, C.c_double), _p(self.eqps, C.c_double)
        st.damage = _p(self.damage, C.c_double)
        st.tear_step = _p(self.tear_step, C.c_int64)
        st.tear_normal = _p(self.tear_normal, C.c_double)
        st.force = _p(self.force, C.c_double)
        if self.corners == 8:
            st.hex_point = _p(self.hex_point, C.c_double)
            st.hex_enhanced = _p(self.hex_enhanced, C.c_double)
        self.st = st

    def __del__(self):
        try:
            lib().ora_case_free(C.byref(self.s))
        except Exception:
            pass

    @property
    def step(self):
        return self.st.step

    @property
    def hex_eqps(self):
        """Per hex element: the largest equivalent plastic strain of its points."""
        return self.hex_point[:, :, 6].max(1)

    @property
    def hex_damage(self):
        """Per hex element: the smallest damage of its points (1: every point failed, the
        element is torn)."""
        return self.hex_point[:, :, 7].min(1)

    def adopt(self, state):
        """Continue from a state computed elsewhere, as zone::Solver::adopt does
        (zone.cpp:1321-1368): a dict with displacement [N,3] (or position), velocity [N,3],
        optional plastic [T,8] (eps_p 6, eqps, damage) for tets or hex_points [T,8,8] and
        hex_enhanced [T,7] for hexes, tear_step [T], steps, and the account's running sums
        w_ext, w_contact, w_boundary, damping_loss, dissipation, fracture."""
        if state.get("displacement") is not None:
            self.u[:] = np.asarray(state["displacement"], np.float64).reshape(self.N, 3)
        else:
            self.u[:] = np.asarray(state["position"], np.float64).reshape(self.N, 3) - self.X
        self.v[:] = np.asarray(state["velocity"], np.float64).reshape(self.N, 3)
        if state.get("plastic") is not None:
            pl = np.asarray(state["plastic"], np.float64).reshape(self.T, 8)
            self.eps_p[:] = pl[:, :6]
            self.eqps[:] = pl[:, 6]
            self.damage[:] = pl[:, 7]
        if self.corners == 8 and state.get("hex_points") is not None:
            # the per-point history of a solid-shell hex (ElementPlasticState,
            # solid_shell.hpp:436-441); the failed flag follows from damage >= 1
            hp = np.asarray(state["hex_points"], np.float64).reshape(self.T, 8, 8)
            self.hex_point[:, :, :8] = hp
            self.hex_point[:, :, 8] = (hp[:, :, 7] >= 1.0).astype(np.float64)
            he = state.get("hex_enhanced")
            self.hex_enhanced[:] = 0.0 if he is None else np.asarray(he, np.float64).reshape(self.T, 7)
        if state.get("tear_step") is not None:
            self.tear_step[:] = state["tear_step"]
        self.st.step = int(state["steps"])
        for k in ("w_ext", "w_contact", "w_boundary", "damping_loss", "dissipation", "fracture"):
            setattr(self.st, k, float(state.get(k, 0.0)))
        return self

    def export(self):
        """The oracle's state in the same dict shape (position, velocity, plastic, steps,
        running sums)."""
        out = dict(position=self.X + self.u, displacement=self.u.copy(), velocity=self.v.copy(),
                   plastic=np.concatenate([self.eps_p, self.eqps[:, None], self.damage[:, None]], 1),
                   tear_step=self.tear_step.copy(), steps=self.st.step)
        if self.corners == 8:
            out["hex_points"] = self.hex_point[:, :, :8].copy()
            out["hex_enhanced"] = self.hex_enhanced.copy()
            out["plastic"] = np.zeros((self.T, 8))
            out["plastic"][:, 6], out["plastic"][:, 7] = self.hex_eqps, self.hex_damage
        for k in ("w_ext", "w_contact", "w_boundary", "damping_loss", "dissipation", "fracture"):
            out[k] = getattr(self.st, k)
        return out

    def run(self, steps, sample_every=0, threads=1):
        """Advance; returns the energy records (list of dicts) of sampled steps.
        threads > 1 splits the element and node loops statically: u, v and the plastic
        state are bit-identical to one thread; the energy sums differ in rounding."""
        max_rec = (steps // sample_every + 2) if sample_every > 0 else 0
        rec = (EnergyStruct * max(max_rec, 1))()
        n = C.c_int64(0)
        if threads > 1:
            lib().ora_run_mt(C.byref(self.s), C.byref(self.st), steps, sample_every, rec,
                             max_rec, C.byref(n), int(threads))
        else:
            lib().ora_run(C.byref(self.s), C.byref(self.st), steps, sample_every, rec, max_rec,
                          C.byref(n))
        out = []
        for i in range(n.value):
            r = rec[i]
            d = {k: getattr(r, k) for k in ["step", "time"] + ENERGY_FIELDS}
            d["contact_force"] = tuple(r.contact_force)
            out.append(d)
        return out


# ---------------------------------------------------------------- FP32 reference restatement
def ref32_rest_state(position, index, rho):
    position = np.ascontiguousarray(position, np.float32)
    index = np.ascontiguousarray(index, np.uint32)
    N, T = len(position), len(index)
    out = dict(rest_inv=np.zeros((T, 9), np.float32), volume=np.zeros(T, np.float32),
               mass=np.zeros(N, np.float32), adj_offset=np.zeros(N + 1, np.uint32),
               adj_entry=np.zeros(4 * T, np.uint32))
    lib().ref32_rest_state(_p(position, C.c_float), N, _p(index, C.c_uint32), T, rho,
                           _p(out["rest_inv"], C.c_float), _p(out["volume"], C.c_float),
                           _p(out["mass"], C.c_float), _p(out["adj_offset"], C.c_uint32),
                           _p(out["adj_entry"], C.c_uint32))
    return out


def ref32_tet_forces(position, index, rs, E, nu):
    lam, mu = ref32_lame(E, nu)
    position = np.ascontiguousarray(position, np.float32)
    T = len(index)
    out = np.zeros((T, 12), np.float32)
    L = lib()
    for t in range(T):
        L.ref32_tet_forces(_p(position, C.c_float), _p(index, C.c_uint32),
                           _p(rs["rest_inv"], C.c_float), _p(rs["volume"], C.c_float), t, lam, mu,
                           _p(out[t], C.c_float))
    return out


def ref32_lame(E, nu):
    E, nu = np.float32(E), np.float32(nu)
    one, two = np.float32(1), np.float32(2)
    lam = E * nu / ((one + nu) * (one - two * nu))
    mu = E / (two * (one + nu))
    return float(lam), float(mu)


class Ref32:
    """The reference's FP32 stepCpu restated (engine/sim/fem.cpp:187-215)."""

    def __init__(self, rest_position, index, fixed, E=210e9, nu=0.3, rho=7850.0):
        self.index = np.ascontiguousarray(index, np.uint32)
        rp = np.ascontiguousarray(rest_position, np.float32).copy()
        self.rs = ref32_rest_state(rp, self.index, rho)
        self.N, self.T = len(rp), len(self.index)
        self.position = rp.copy()
        self.velocity = np.zeros((self.N, 3), np.float32)
        self.fixed = np.ascontiguousarray(fixed, np.uint32)
        self.tet_force = np.zeros((self.T, 12), np.float32)
        self.E, self.nu = E, nu

    def step(self, dt, gravity=0.0, damping=1.0, threads=1):
        L = lib()
        args = [_p(self.position, C.c_float), _p(self.velocity, C.c_float),
                _p(self.rs["mass"], C.c_float), _p(self.fixed, C.c_uint32), self.N,
                _p(self.index, C.c_uint32), _p(self.rs["rest_inv"], C.c_float),
                _p(self.rs["volume"], C.c_float), self.T, _p(self.tet_force, C.c_float),
                _p(self.rs["adj_offset"], C.c_uint32), _p(self.rs["adj_entry"], C.c_uint32),
                self.E, self.nu, dt, gravity, damping]
        if threads == 1:
            L.ref32_step(*args)
        else:
            L.ref32_step_mt(*(args + [threads]))


# ---------------------------------------------------------------- compiled reference (pin only)
_ref = None


def ref_lib():
    """The compiled reference (oracle/_ref/libref.so), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_box_counts.argtypes = [C.c_int] * 3 + [I64, I64]
        R.ref_box_mesh.argtypes = [C.c_float] * 3 + [C.c_int] * 3 + [F, U32]
        R.ref_rest_state.argtypes = [F, C.c_int64, U32, C.c_int64, C.c_float, C.c_float, C.c_float,
                                     C.c_float, F, F, F, U32, U32]
        R.ref_rest_state.restype = C.c_float
        R.ref_tet_forces.argtypes = [F, F, C.c_int64, U32, C.c_int64, C.c_float, C.c_float,
                                     C.c_float, F]
        R.ref_step.argtypes = [F, F, F, U32, C.c_int64, U32, C.c_int64, C.c_float, C.c_float,
                               C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]
        R.ref_j2_update.argtypes = [C.c_double] * 4 + [D, D, D, D, D, C.POINTER(C.c_int)]
        R.ref_j2_update.restype = C.c_int
        R.ref_j2_damage_update.argtypes = [C.c_double] * 8 + [D, D, D, D, C.POINTER(C.c_int), D,
                                                              D, D, D, C.POINTER(C.c_int)]
        R.ref_j2_damage_update.restype = C.c_int
        R.ref_regularised_failure_strain.argtypes = [C.c_double] * 4
        R.ref_regularised_failure_strain.restype = C.c_double
        R.ref_triaxiality_factor.argtypes = [C.c_double] * 4
        R.ref_triaxiality_factor.restype = C.c_double
        R.ref_hex_stiffness.argtypes = [D, C.c_double, C.c_double, C.c_int, D]
        R.ref_hex_mass.argtypes = [D, C.c_double, D]
        R.ref_hex_force.argtypes = [D, D, D, D]
        R.ref_hex_rotation.argtypes = [D, D, D]
        R.ref_hex_critical_timestep.argtypes = [D] + [C.c_double] * 4
        R.ref_hex_critical_timestep.restype = C.c_double
        R.ref_hex_plastic_path.argtypes = [D, D, C.c_int] + [C.c_double] * 8 + [D] * 5
        R.ref_hex_failure_strain.argtypes = [D, C.c_double, C.c_double]
        R.ref_hex_failure_strain.restype = C.c_double
        R.ref_hex_plate_counts.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [I64, I64]
        R.ref_hex_plate.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [C.c_double, D, U32, D]
        _ref = R
    return _ref

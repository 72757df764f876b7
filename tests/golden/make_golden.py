"""Write golden vectors from the compiled reference (oracle/_ref/libref.so).

Run in a container that has the reference tree and the built oracle/_ref:
    make -C oracle && python tests/golden/make_golden.py [fem|hex|all]
Output: tests/golden/reference_fem.npz and tests/golden/reference_hex.npz (inputs and
the reference's outputs only).

What is recorded, and from which reference function:
  box_*       makeBoxTetMesh(0.4, 0.1, 0.02, 16, 4, 2)          engine/sim/fem.cpp:217-263
  c1_*        makeBoxTetMesh(1.0, 0.1, 0.1, 40, 4, 4)           (the small cantilever)
  rest_*      computeRestState + criticalTimestep(0.5) of box   engine/sim/fem.cpp:119-185
  twist_*     tetForces on the fem_spike deformed/twisted box   tools/fem_spike/main.cpp:147-159
  step1_/step10_  stepCpu 1 and 10 steps from the twisted state, root clamped, gravity
              9.80665, damping 1, dt = criticalTimestep        engine/sim/fem.cpp:187-215
  j2_*        plasticity::update along strain paths, linear hardening, state carried
              from point to point                               engine/sim/plasticity.cpp:319-455
  fail_*      regularisedFailureStrain, triaxialityFactor      engine/sim/plasticity.cpp:158-174
  dmg_*       plasticity::update with ductile failure along pull, rotated pull,
              squeeze, biaxial and mixed paths (damage, failed, failure normal)
                                                                engine/sim/plasticity.cpp:457-477

reference_hex.npz, from engine/sim/solid_shell.cpp (Formulation::SolidShell, steel
E 206 GPa, nu 0.3, rho 7850):
  el_nodes    eight elements: a 0.6 x 0.6 x 0.012 m plate element, the unit cube, a
              0.3 x 0.2 x 0.01 m element with seeded corner noise, that element rotated
              and moved to (3, -9.9, 8) m, a sheared and tapered element, a 2:1 brick
              at nu 0.45, a thick skewed element and a 0.1 x 0.05 x 0.02 m web element
  el_nu       Poisson ratio of each element (0.45 for the sixth, else 0.3)
  el_K        elementStiffness [24 x 24]                   solid_shell.cpp:500-526
  el_mass     elementMass                                   solid_shell.cpp:655-663
  el_dt       criticalTimestep(safety 0.9)                  solid_shell.cpp:1088-1134
  el_current  three current configurations per element: a small seeded deformation,
              a 35 degree rotation plus that deformation, and 5% stretch plus shear
  el_force    internalForce(el_K, rest, current)            solid_shell.cpp:735-766
  el_rotation elementRotation(rest, current), column-major  solid_shell.cpp:680-733
  plate_*     makePlateMesh(0.6, 0.3, 0.012, 3, 2, 1) and (0.4, 0.2, 0.02, 2, 2, 2) with
              uniformPressureLoad(1e5)                      solid_shell.cpp:1152-1185,
                                                            1505-1561
  pl{0,1,2}_* elementPlasticUpdate along three load paths of 40 configurations from a fresh
              state (sigma_y 355 MPa, H 1 GPa; Rice-Tracey 1.5, 1/3, -1/3): a thin plate
              element pulled in x to 25% at constant volume and bent to 1.5 /m, failure strain 0.12
              (its points fail and the element tears); the noisy element bent and rotated
              20 degrees without failure; a thick skewed element in simple shear to 0.8 with
              a twist (shear varying across it), failure strain 0.25.
              Force, point stress and state, enhanced parameters, dissipation,
              iterations, yielded/failed counts, torn   solid_shell.cpp:826-1029
  fs_*        initialisePlasticState's regularised failure strain of each test element
              (uniform 0.12, fracture 0.8)                  solid_shell.cpp:770-815
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import oracle as O  # noqa: E402


def p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref.so not built (make -C oracle)")
    E, nu, rho = 210e9, 0.3, 7850.0
    out = {}

    def box(lx, ly, lz, nx, ny, nz):
        n, t = C.c_int64(), C.c_int64()
        R.ref_box_counts(nx, ny, nz, C.byref(n), C.byref(t))
        pos = np.zeros((n.value, 3), np.float32)
        idx = np.zeros((t.value, 4), np.uint32)
        R.ref_box_mesh(lx, ly, lz, nx, ny, nz, p(pos, C.c_float), p(idx, C.c_uint32))
        return pos, idx

    pos, idx = box(0.4, 0.1, 0.02, 16, 4, 2)
    out["box_position"], out["box_index"] = pos, idx
    c1p, c1i = box(1.0, 0.1, 0.1, 40, 4, 4)
    out["c1_position"], out["c1_index"] = c1p, c1i

    N, T = len(pos), len(idx)
    ri = np.zeros((T, 9), np.float32)
    vol = np.zeros(T, np.float32)
    mass = np.zeros(N, np.float32)
    off = np.zeros(N + 1, np.uint32)
    ent = np.zeros(4 * T, np.uint32)
    dt = R.ref_rest_state(p(pos, C.c_float), N, p(idx, C.c_uint32), T, E, nu, rho, 0.5,
                          p(ri, C.c_float), p(vol, C.c_float), p(mass, C.c_float),
                          p(off, C.c_uint32), p(ent, C.c_uint32))
    out.update(rest_inv=ri, rest_volume=vol, rest_mass=mass, rest_adj_offset=off,
               rest_adj_entry=ent, rest_dt=np.float32(dt))

    # the fem_spike deformation: z += 0.02 x^2, y += 0.01 x, applied after rest state
    cur = pos.copy()
    x = cur[:, 0].copy()
    cur[:, 2] += np.float32(0.02) * x * x
    cur[:, 1] += np.float32(0.01) * x
    out["twist_position"] = cur
    forces = np.zeros((T, 12), np.float32)
    R.ref_tet_forces(p(pos, C.c_float), p(cur, C.c_float), N, p(idx, C.c_uint32), T, E, nu, rho,
                     p(forces, C.c_float))
    out["twist_forces"] = forces

    fixed = (pos[:, 0] < np.float32(1e-6)).astype(np.uint32)
    out["twist_fixed"] = fixed
    for steps in (1, 10):
        P = cur.copy()
        V = np.zeros_like(P)
        R.ref_step(p(pos, C.c_float), p(P, C.c_float), p(V, C.c_float), p(fixed, C.c_uint32), N,
                   p(idx, C.c_uint32), T, E, nu, rho, dt, 9.80665, 1.0, steps)
        out[f"step{steps}_position"], out[f"step{steps}_velocity"] = P, V

    # J2 point law along three strain paths, state carried between points
    Ej, nuj, sy, H = 206e9, 0.3, 355e6, 1e9
    ey = sy / Ej
    paths = {
        "uniaxial": [np.array([a, -nuj * a, -nuj * a, 0, 0, 0]) for a in
                     (0.5 * ey, 0.999 * ey, 1.5 * ey, 3 * ey, 10 * ey, 40 * ey, 20 * ey, -5 * ey)],
        "shear": [np.array([0, 0, 0, g, 0, 0]) for g in (0.5 * ey, 1.2 * ey, 2.0 * ey, 6 * ey, 3 * ey)],
        "mixed": [np.array([0.4, -0.1, 0.05, 0.3, -0.2, 0.15]) * s * ey for s in (1, 3, 6, 12, 8, 20)],
    }
    for name, path in paths.items():
        ep = np.zeros(6)
        q = np.zeros(1)
        strains, stresses, eps_ps, eqpss, disss, ylds, iters = [], [], [], [], [], [], []
        for s in path:
            s = np.ascontiguousarray(s, np.float64)
            stress = np.zeros(6)
            d = np.zeros(1)
            y = C.c_int()
            it = R.ref_j2_update(Ej, nuj, sy, H, p(s, C.c_double), p(ep, C.c_double),
                                 p(q, C.c_double), p(stress, C.c_double), p(d, C.c_double),
                                 C.byref(y))
            strains.append(s.copy())
            stresses.append(stress.copy())
            eps_ps.append(ep.copy())
            eqpss.append(q[0])
            disss.append(d[0])
            ylds.append(y.value)
            iters.append(it)
        out[f"j2_{name}_strain"] = np.array(strains)
        out[f"j2_{name}_stress"] = np.array(stresses)
        out[f"j2_{name}_eps_p"] = np.array(eps_ps)
        out[f"j2_{name}_eqps"] = np.array(eqpss)
        out[f"j2_{name}_dissipation"] = np.array(disss)
        out[f"j2_{name}_yielded"] = np.array(ylds)
        out[f"j2_{name}_iterations"] = np.array(iters)
    out["j2_material"] = np.array([Ej, nuj, sy, H])

    # Ductile damage on the same point law: a regularised failure strain and the
    # Rice-Tracey constants, state (damage, failed) carried along each path
    eu, efr, length, thick = 0.12, 0.8, 0.05, 0.02
    sens, eta_ref, eta_cut = 1.5, 1.0 / 3.0, -1.0 / 3.0
    ef = R.ref_regularised_failure_strain(eu, efr, length, thick)
    lengths = np.array([0.005, 0.01, 0.02, 0.05, 0.1, 1.0, 100.0])
    out["fail_reg_input"] = np.array([eu, efr, thick])
    out["fail_reg_length"] = lengths
    out["fail_reg_strain"] = np.array([R.ref_regularised_failure_strain(eu, efr, l, thick)
                                       for l in lengths])
    etas = np.array([-0.5, -1.0 / 3.0 + 1e-9, 0.0, 1.0 / 3.0, 0.5, 2.0 / 3.0, 1.0])
    out["fail_eta"] = etas
    out["fail_factor"] = np.array([R.ref_triaxiality_factor(sens, eta_ref, eta_cut, e)
                                   for e in etas])
    out["fail_material"] = np.array([Ej, nuj, sy, H, ef, sens, eta_ref, eta_cut])

    def pull(plastic, rot=None):
        sig = sy + H * plastic
        a = sig / Ej + plastic
        lat = -nuj * sig / Ej - 0.5 * plastic
        if rot is None:
            return np.array([a, lat, lat, 0, 0, 0])
        T = rot @ np.diag([a, lat, lat]) @ rot.T
        return np.array([T[0, 0], T[1, 1], T[2, 2], 2 * T[0, 1], 2 * T[1, 2], 2 * T[0, 2]])

    def biaxial(plastic):
        # sigma = (s, s, 0): flow (1/2, 1/2, -1) eqps, elastic part from Hooke
        sig = sy + H * plastic
        a = sig * (1 - nuj) / Ej + 0.5 * plastic
        c = -2 * nuj * sig / Ej - plastic
        return np.array([a, a, c, 0, 0, 0])

    ax = np.array([0.31, 0.62, -0.72])
    ax /= np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    rot = np.eye(3) + np.sin(0.9) * K + (1 - np.cos(0.9)) * K @ K
    fracs = (0.1, 0.4, 0.8, 0.99, 1.005, 1.2, 0.5)
    dpaths = {
        # uniaxial tension sits at the reference triaxiality: damage = eps_p / eps_f
        "pull": [pull(f * ef) for f in fracs],
        "rotated": [pull(f * ef, rot) for f in fracs],
        # a superposed hydrostatic squeeze keeps the triaxiality below the cutoff
        "squeeze": [np.array([-2 * f - 0.004, f - 0.004, f - 0.004, 0, 0, 0])
                    for f in np.linspace(0.01, 0.1, 8)],
        # equibiaxial plane stress: eta = 2/3, fails at exp(-1/2) of the uniaxial strain
        "biaxial": [biaxial(f * ef) for f in (0.1, 0.3, 0.5, 0.6, 0.61, 0.7)],
        "mixed": [np.array([0.4, -0.3, -0.05, 0.3, -0.2, 0.15]) * f
                  for f in np.linspace(0.004, 0.8, 12)],
    }
    for name, path in dpaths.items():
        ep = np.zeros(6)
        q = np.zeros(1)
        dmg = np.zeros(1)
        failed = C.c_int(0)
        rows = {k: [] for k in ("strain", "stress", "eps_p", "eqps", "damage", "failed",
                                "failed_now", "normal", "eta", "dissipation", "yielded")}
        for s_ in path:
            s_ = np.ascontiguousarray(s_, np.float64)
            stress, d, nrm, eta = np.zeros(6), np.zeros(1), np.zeros(3), np.zeros(1)
            y = C.c_int()
            now = R.ref_j2_damage_update(Ej, nuj, sy, H, ef, sens, eta_ref, eta_cut,
                                         p(s_, C.c_double), p(ep, C.c_double), p(q, C.c_double),
                                         p(dmg, C.c_double), C.byref(failed),
                                         p(stress, C.c_double), p(d, C.c_double),
                                         p(nrm, C.c_double), p(eta, C.c_double), C.byref(y))
            for k, v in (("strain", s_), ("stress", stress), ("eps_p", ep.copy()), ("eqps", q[0]),
                         ("damage", dmg[0]), ("failed", failed.value), ("failed_now", now),
                         ("normal", nrm), ("eta", eta[0]), ("dissipation", d[0]),
                         ("yielded", y.value)):
                rows[k].append(np.array(v, copy=True))
        for k, v in rows.items():
            out[f"dmg_{name}_{k}"] = np.array(v)

    path = os.path.join(HERE, "reference_fem.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


def hex_elements(rng):
    """The eight test elements (node order of solid_shell.hpp:115-125)."""
    unit = np.stack([[-1, 1, 1, -1, -1, 1, 1, -1], [-1, -1, 1, 1, -1, -1, 1, 1],
                     [-1, -1, -1, -1, 1, 1, 1, 1]], 1) * 0.5 + 0.5
    plate = unit * np.array([0.6, 0.6, 0.012]) - np.array([0, 0, 0.006])
    noisy = unit * np.array([0.3, 0.2, 0.01]) + rng.normal(0, 1, (8, 3)) * np.array([3e-3, 2e-3, 1e-4])
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    moved = noisy @ q.T + np.array([3.0, -9.9, 8.0])
    sheared = unit * np.array([0.4, 0.3, 0.015])
    sheared[:, 0] += 0.1 * sheared[:, 1]           # in-plane shear
    sheared[:, 1] *= 1.0 - 0.3 * unit[:, 0]        # taper along x
    sheared[:, 2] += 0.02 * unit[:, 0]             # warp
    brick = unit * np.array([0.2, 0.1, 0.05])
    thick = unit * np.array([0.1, 0.1, 0.08])
    thick[:, 0] += 0.3 * thick[:, 2]
    web = unit * np.array([0.1, 0.05, 0.02])
    return [plate, unit, noisy, moved, sheared, brick, thick, web], [0.3] * 5 + [0.45] + [0.3] * 2


def main_hex():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref.so not built (make -C oracle)")
    E, rho = 206e9, 7850.0
    rng = np.random.default_rng(20261203)
    els, nus = hex_elements(rng)
    out = {"el_nodes": np.array(els), "el_nu": np.array(nus)}
    Ks, ms, dts, curs, fs, rots = [], [], [], [], [], []
    for x, nu in zip(els, nus):
        x = np.ascontiguousarray(x)
        K = np.zeros(576)
        R.ref_hex_stiffness(p(x, C.c_double), E, nu, 2, p(K, C.c_double))
        m = np.zeros(8)
        R.ref_hex_mass(p(x, C.c_double), rho, p(m, C.c_double))
        dts.append(R.ref_hex_critical_timestep(p(x, C.c_double), E, nu, rho, 0.9))
        size = np.ptp(x, axis=0).max()
        a = np.deg2rad(35.0)
        ax = np.array([1.0, 2.0, 0.5]) / np.linalg.norm([1.0, 2.0, 0.5])
        W = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        Rot = np.eye(3) + np.sin(a) * W + (1 - np.cos(a)) * W @ W
        small = rng.normal(0, 1e-3 * size, (8, 3))
        c0 = x.mean(0)
        stretch = np.array([[1.05, 0.02, 0.0], [0.0, 0.98, 0.01], [0.0, 0.0, 1.01]])
        cur3 = [x + small, (x - c0) @ Rot.T + c0 + small, (x - c0) @ stretch.T + c0]
        fe, re = [], []
        for cur in cur3:
            cur = np.ascontiguousarray(cur)
            f = np.zeros(24)
            R.ref_hex_force(p(K, C.c_double), p(x, C.c_double), p(cur, C.c_double), p(f, C.c_double))
            r9 = np.zeros(9)
            R.ref_hex_rotation(p(x, C.c_double), p(cur, C.c_double), p(r9, C.c_double))
            fe.append(f)
            re.append(r9)
        Ks.append(K.reshape(24, 24))
        ms.append(m)
        curs.append(np.array(cur3))
        fs.append(np.array(fe))
        rots.append(np.array(re))
    out.update(el_K=np.array(Ks), el_mass=np.array(ms), el_dt=np.array(dts),
               el_current=np.array(curs), el_force=np.array(fs), el_rotation=np.array(rots),
               el_material=np.array([E, rho]))
    for name, (lx, ly, t, nx, ny, nz) in (("plate_a", (0.6, 0.3, 0.012, 3, 2, 1)),
                                          ("plate_b", (0.4, 0.2, 0.02, 2, 2, 2))):
        n, e = C.c_int64(), C.c_int64()
        R.ref_hex_plate_counts(lx, ly, t, nx, ny, nz, C.byref(n), C.byref(e))
        X = np.zeros((n.value, 3))
        idx = np.zeros((e.value, 8), np.uint32)
        load = np.zeros((n.value, 3))
        R.ref_hex_plate(lx, ly, t, nx, ny, nz, 1e5, p(X, C.c_double), p(idx, C.c_uint32),
                        p(load, C.c_double))
        out[f"{name}_spec"] = np.array([lx, ly, t, nx, ny, nz])
        out[f"{name}_position"], out[f"{name}_index"], out[f"{name}_load"] = X, idx, load
    # plastic paths
    Ej, nuj, sy, H = 206e9, 0.3, 355e6, 1e9
    out["pl_material"] = np.array([Ej, nuj, sy, H, 1.5, 1.0 / 3.0, -1.0 / 3.0])
    K = 40
    specs = [(0, 0.12), (2, 0.0), (6, 0.25)]
    for j, (k, ef) in enumerate(specs):
        x = np.ascontiguousarray(els[k])
        c0 = x.mean(0)
        size = np.ptp(x, axis=0)
        cur = []
        for step in range(1, K + 1):
            t = step / K
            if j == 0:
                # near-incompressible pull with a superposed bending curvature
                a = 0.25 * t
                F = np.diag([1 + a, 1 / np.sqrt(1 + a), 1 / np.sqrt(1 + a)])
                y = (x - c0) @ F.T + c0
                kap = 1.5 * t  # 1/m, Kirchhoff bending about y on top of the pull
                y[:, 0] -= kap * (x[:, 2] - c0[2]) * (x[:, 0] - c0[0])
                y[:, 2] += 0.5 * kap * (x[:, 0] - c0[0]) ** 2
            elif j == 1:
                a = np.deg2rad(20.0) * t
                Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
                bend = np.zeros_like(x)
                bend[:, 0] = 0.08 * t * (x[:, 2] - c0[2]) * (x[:, 0] - c0[0]) / size[2] / 10
                bend[:, 2] = -0.5 * 0.08 * t * (x[:, 0] - c0[0]) ** 2 / size[2] / 10
                y = (x - c0 + bend) @ Rz.T + c0
            else:
                # simple shear with a shear gradient through the thickness
                F = np.array([[1.0, 0.8 * t, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
                y = (x - c0) @ F.T + c0
                y[:, 0] += 0.6 * t * (x[:, 2] - c0[2]) * (x[:, 1] - c0[1]) / size[1]
            cur.append(y)
        cur = np.ascontiguousarray(np.array(cur))
        f = np.zeros((K, 24))
        st = np.zeros((K, 48))
        state = np.zeros((K, 72))
        enh = np.zeros((K, 7))
        info = np.zeros((K, 6))
        R.ref_hex_plastic_path(p(x, C.c_double), p(cur, C.c_double), K, Ej, nuj, sy, H, ef, 1.5,
                               1.0 / 3.0, -1.0 / 3.0, p(f, C.c_double), p(st, C.c_double),
                               p(state, C.c_double), p(enh, C.c_double), p(info, C.c_double))
        out.update({f"pl{j}_rest": x, f"pl{j}_current": cur, f"pl{j}_failure_strain": np.float64(ef),
                    f"pl{j}_force": f, f"pl{j}_stress": st.reshape(K, 8, 6),
                    f"pl{j}_state": state.reshape(K, 8, 9), f"pl{j}_enhanced": enh,
                    f"pl{j}_info": info})
    out["fs_strain"] = np.array([R.ref_hex_failure_strain(p(np.ascontiguousarray(x), C.c_double),
                                                          0.12, 0.8) for x in els])

    path = os.path.join(HERE, "reference_hex.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("fem", "all"):
        main()
    if what in ("hex", "all"):
        main_hex()

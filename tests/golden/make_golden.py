"""Write golden vectors from the compiled reference (oracle/_ref/libref.so).

Run in a container that has the reference tree and the built oracle/_ref:
    make -C oracle && python tests/golden/make_golden.py
Output: tests/golden/reference_fem.npz (inputs and the reference's outputs only).

What is recorded, and from which reference function:
  box_*       makeBoxTetMesh(0.4, 0.1, 0.02, 16, 4, 2)          engine/sim/fem.cpp:217-263
  c1_*        makeBoxTetMesh(1.0, 0.1, 0.1, 40, 4, 4)           (the small cantilever)
  rest_*      computeRestState + criticalTimestep(0.5) of box   engine/sim/fem.cpp:119-185
  twist_*     tetForces on the fem_spike deformed/twisted box   tools/fem_spike/main.cpp:147-159
  step1_/step10_  stepCpu 1 and 10 steps from the twisted state, root clamped, gravity
              9.80665, damping 1, dt = criticalTimestep        engine/sim/fem.cpp:187-215
  j2_*        plasticity::update along strain paths, linear hardening, state carried
              from point to point                               engine/sim/plasticity.cpp:319-455
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

    path = os.path.join(HERE, "reference_fem.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()

"""The solid-shell hexahedron on the host: the library's FP64 element (hex.cpp) and the
oracle's numpy restatement (oracle/hex.py), each pinned to golden vectors of the
compiled reference (engine/sim/solid_shell.cpp, tests/golden/reference_hex.npz), then to
what the mechanics fixes: the rigid-body null space, the constant-strain patch test,
rotation invariance of the co-rotational force, the consistent mass and the stable step
of a leapfrog. The oracle's hex stepping is pinned by its energy account and a free
translation.
"""
import os

import numpy as np
import pytest

import shipfem as sf
from oracle import hex as H
from oracle import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_hex.npz"))
E, RHO = GOLD["el_material"]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("k", range(len(GOLD["el_nodes"])))
def test_library_element_matches_compiled_reference(k):
    x, nu = GOLD["el_nodes"][k], GOLD["el_nu"][k]
    K = sf.hex_stiffness(x, E, nu)
    assert rel(K, GOLD["el_K"][k]) <= 1e-13
    assert rel(sf.hex_mass(x, RHO), GOLD["el_mass"][k]) <= 1e-14
    assert sf.hex_critical_timestep(x, E, nu, RHO, 0.9) == pytest.approx(GOLD["el_dt"][k], rel=1e-9)
    for c in range(3):
        f = sf.hex_force(GOLD["el_K"][k], x, GOLD["el_current"][k, c])
        # the moved element works at |x| ~ 13 m: FP64 cancellation in R^T x - X
        assert rel(f, GOLD["el_force"][k, c]) <= (1e-9 if k == 3 else 1e-12)


@pytest.mark.parametrize("k", range(len(GOLD["el_nodes"])))
def test_oracle_element_matches_compiled_reference(k):
    x, nu = GOLD["el_nodes"][k], GOLD["el_nu"][k]
    K = H.stiffness(x[None], E, nu)[0]
    assert rel(K, GOLD["el_K"][k]) <= 1e-12
    assert rel(H.mass(x[None], RHO)[0], GOLD["el_mass"][k]) <= 1e-14
    m = H.mass(x[None], RHO)
    assert H.critical_timestep(K[None], m)[0] == pytest.approx(GOLD["el_dt"][k], rel=1e-9)
    cur = GOLD["el_current"][k]
    f = H.force(np.repeat(GOLD["el_K"][k][None], 3, 0), np.repeat(x[None], 3, 0), cur)
    assert rel(f, GOLD["el_force"][k]) <= (1e-9 if k == 3 else 1e-12)
    R = H.rotation(np.repeat(x[None], 3, 0), cur)
    gold = GOLD["el_rotation"][k].reshape(3, 3, 3).transpose(0, 2, 1)  # column-major
    assert np.abs(R - gold).max() <= 1e-12
    # the C oracle's element force is the same function
    for c in range(3):
        out = np.zeros(24)
        O.lib().ora_hex_force(O._p(np.ascontiguousarray(GOLD["el_K"][k]), O.C.c_double),
                              O._p(np.ascontiguousarray(x), O.C.c_double),
                              O._p(np.ascontiguousarray(cur[c]), O.C.c_double),
                              O._p(out, O.C.c_double))
        assert rel(out, GOLD["el_force"][k, c]) <= (1e-9 if k == 3 else 1e-12)


@pytest.mark.parametrize("name", ["plate_a", "plate_b"])
def test_plate_mesh_and_pressure_load_match_compiled_reference(name):
    lx, ly, t, nx, ny, nz = GOLD[f"{name}_spec"]
    X, index = sf.hex_plate_mesh(lx, ly, t, int(nx), int(ny), int(nz))
    assert np.array_equal(X, GOLD[f"{name}_position"])
    assert np.array_equal(index, GOLD[f"{name}_index"])
    load = sf.hex_pressure_load(X, index, 1e5)
    assert rel(load, GOLD[f"{name}_load"]) <= 1e-14
    assert load[:, 2].sum() == pytest.approx(-1e5 * lx * ly, rel=1e-14)
    Xo, io = H.plate_mesh(lx, ly, t, int(nx), int(ny), int(nz))
    assert np.array_equal(Xo, X) and np.array_equal(io, index)


def elements():
    return [np.asarray(x) for x in GOLD["el_nodes"]]


def test_stiffness_has_exactly_the_six_rigid_modes():
    for x, nu in zip(elements(), GOLD["el_nu"]):
        K = sf.hex_stiffness(x, E, nu)
        assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
        w = np.linalg.eigvalsh(0.5 * (K + K.T))
        # six zeros at rounding; the seventh (a thin element's twist, ~ (t/L)^4 of the
        # thickness stiffness) stands orders of magnitude clear of them
        assert np.all(np.abs(w[:6]) <= 1e-14 * w[-1]) and w[6] > 1e6 * np.abs(w[:6]).max()
        # the null space is the rigid translations and infinitesimal rotations
        c = x.mean(0)
        for a in range(3):
            t = np.zeros((8, 3))
            t[:, a] = 1.0
            assert np.abs(K @ t.ravel()).max() <= 1e-9 * np.abs(K).max()
            w_ = np.zeros(3)
            w_[a] = 1.0
            r = np.cross(w_, x - c)
            assert np.abs(K @ r.ravel()).max() <= 1e-9 * np.abs(K).max() * np.ptp(x)


def affine(x):
    """The element is an affine image of the unit cube (constant Jacobian)."""
    nat = np.stack([H.XI, H.ETA, H.ZETA, np.ones(8)], 1)
    coef, *_ = np.linalg.lstsq(nat, x, rcond=None)
    return np.abs(nat @ coef - x).max() <= 1e-12 * np.ptp(x)


def test_constant_strain_patch_test():
    # a linear displacement field u = G X has constant strain sym(G); on an affine element
    # (constant Jacobian, so constant covariant strain for the ANS samples to reproduce)
    # the element stores exactly V * 1/2 eps : C : eps, the EAS modes being orthogonal to
    # constant stress. On a distorted element the ANS interpolation of covariant strains
    # that vary across it does not reproduce them: those elements are left out.
    rng = np.random.default_rng(5)
    cases = [(x, nu) for x, nu in zip(elements(), GOLD["el_nu"]) if affine(x)]
    assert len(cases) == 5
    for x, nu in cases:
        K = sf.hex_stiffness(x, E, nu)
        V = sf.hex_mass(x, 1.0).sum()
        for _ in range(3):
            G = rng.normal(size=(3, 3)) * 1e-4
            u = (x @ G.T).ravel()
            eps = 0.5 * (G + G.T)
            v = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[2, 0]])
            U = 0.5 * V * v @ H.elasticity(E, nu) @ v
            assert 0.5 * u @ K @ u == pytest.approx(U, rel=1e-9)
            # and the nodal forces are the consistent loads of that constant stress:
            # they vanish when summed with the moment arm (equilibrium)
            f = -(K @ u).reshape(8, 3)
            assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).max()
            assert np.abs(np.cross(x, f).sum(0)).max() <= 1e-9 * np.abs(f).max() * np.ptp(x)


def test_force_is_zero_for_finite_rigid_motion_and_rotates_with_it():
    rng = np.random.default_rng(8)
    for x, nu in zip(elements(), GOLD["el_nu"]):
        K = sf.hex_stiffness(x, E, nu)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        c = x.mean(0)
        moved = (x - c) @ q.T + c + np.array([0.3, -0.2, 0.1])
        f = sf.hex_force(K, x, moved)
        scale = np.abs(K).max() * np.ptp(x)
        assert np.abs(f).max() <= 1e-12 * scale
        # deform, then rotate the deformed element: the force rotates with it
        cur = x + rng.normal(0, 1e-3 * np.ptp(x), (8, 3))
        f0 = sf.hex_force(K, x, cur).reshape(8, 3)
        f1 = sf.hex_force(K, x, (cur - c) @ q.T + c).reshape(8, 3)
        assert np.abs(f1 - f0 @ q.T).max() <= 1e-9 * np.abs(f0).max()


def test_consistent_mass_sums_to_rho_V_and_stable_step_bounds_the_leapfrog():
    for x, nu in zip(elements(), GOLD["el_nu"]):
        m = sf.hex_mass(x, RHO)
        B, G, w = H.forms(x)  # w: the Gauss points' volumes
        assert m.sum() == pytest.approx(RHO * w.sum(), rel=1e-13)
        if affine(x):  # an affine element's volume is 8 det J
            J = H.jacobian(x[None], H.dshape(0, 0, 0))[0]
            assert m.sum() == pytest.approx(RHO * 8 * np.linalg.det(J), rel=1e-13)
        K = sf.hex_stiffness(x, E, nu)
        dt = sf.hex_critical_timestep(x, E, nu, RHO, 1.0)
        # leapfrog on the free element: bounded at 0.99 of 2/omega, growing at 1.02
        Minv = 1.0 / np.repeat(m, 3)
        for factor, bounded in ((0.99, True), (1.02, False)):
            h = factor * dt
            u = np.random.default_rng(1).normal(size=24) * 1e-9
            v = np.zeros(24)
            for _ in range(3000):
                v -= h * Minv * (K @ u)
                u += h * v
                if np.abs(u).max() > 1.0:
                    break
            assert (np.abs(u).max() < 1e-6) == bounded


def test_oracle_hex_solve_closes_its_energy_account_and_translates_freely():
    # Clamped plate under a ramped pressure with velocity damping. The force uses one
    # rotation per element (its centre's) while the strain varies across it, so it is not
    # exactly the gradient of 1/2 d.Kd: the account's residual is of the order of the
    # rotation, relative (about 1e-4 at 200 kPa, 2 mm deflection). At 2 kPa it closes to
    # 1e-7, and it shrinks with the load.
    X, index = H.plate_mesh(0.6, 0.3, 0.012, 6, 3, 1)
    N = len(X)
    flags = np.zeros(N, np.uint32)
    edge = (X[:, 0] < 1e-9) | (X[:, 0] > 0.6 - 1e-9)
    flags[edge] = O.PIN_ALL
    K = H.stiffness(X[index], E, 0.3)
    dt = 0.9 * H.critical_timestep(K, H.mass(X[index], RHO)).min()
    res = []
    for p in (2e4, 2e3):
        load = sf.hex_pressure_load(X, index, p)
        c = O.Case(X, index, E=E, nu=0.3, rho=RHO, dt=dt, flags=flags, load=load,
                   load_ramp=2e-3, damping=0.9995)
        rec = c.run(3000, sample_every=500)
        r = rec[-1]
        residual = r["w_ext"] - (r["strain"] + r["kinetic"] - rec[0]["strain"] - rec[0]["kinetic"]) \
            - r["damping_loss"]
        assert r["strain"] > 0 and r["damping_loss"] > 0
        res.append(abs(residual) / r["w_ext"])
    assert res[1] <= 2e-7 and res[1] < 0.2 * res[0]
    # a free plate given a uniform velocity carries no stress and moves rigidly
    f = O.Case(X, index, E=E, nu=0.3, rho=RHO, dt=dt)
    f.v[:] = np.array([1.0, -2.0, 0.5])
    rec = f.run(200, sample_every=100)
    assert np.abs(f.u - 200 * dt * np.array([1.0, -2.0, 0.5])).max() <= 1e-12
    # (rounding of R^T x - X at |u| ~ 0.5 mm: forces of micro-newtons, energy 1e-13 of
    # the kinetic)
    assert all(abs(r["strain"]) <= 1e-12 * r["kinetic"] for r in rec)
    assert np.abs(f.force).max() <= 1e-3


def test_hex_solver_refuses_a_plastic_case_before_touching_a_device():
    case = sf.HexCase.plate(0.2, 0.1, 0.01, 2, 2, 1, plastic=True)
    with pytest.raises(sf.ShipfemError) as e:
        sf.Solver(case)
    assert e.value.status == sf.UNSUPPORTED

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
        gold = GOLD["el_rotation"][k, c].reshape(3, 3).T  # column-major
        assert np.abs(sf.hex_rotation(x, GOLD["el_current"][k, c]) - gold).max() <= 1e-12


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


def plastic_update(X, x, j2, point, enh, B=None, G=None, w=None):
    if B is None:
        B, G, w = (np.ascontiguousarray(a[0]) for a in H.forms(X[None]))
    f, st, info = np.zeros(24), np.zeros(48), np.zeros(6)
    c = O.C.c_double
    O.lib().ora_hex_plastic_update(O._p(B, c), O._p(G, c), O._p(w, c),
                                   O._p(np.ascontiguousarray(X), c), O._p(np.ascontiguousarray(x), c),
                                   O.C.byref(j2), O._p(point, c), O._p(enh, c), O._p(f, c),
                                   O._p(st, c), O._p(info, c))
    return f, st.reshape(8, 6), info


@pytest.mark.parametrize("j", range(3))
def test_oracle_plastic_update_matches_compiled_reference(j):
    # elementPlasticUpdate along a path (solid_shell.cpp:826-1029): the enhanced Newton,
    # the return at eight points with damage, the rerun without enhanced modes in the step
    # a point fails, the torn element
    E_, nu_, sy, Hh, sens, er, ec = GOLD["pl_material"]
    X, cur = GOLD[f"pl{j}_rest"], GOLD[f"pl{j}_current"]
    j2 = O.J2(E_, nu_, sy, Hh, float(GOLD[f"pl{j}_failure_strain"]), sens, er, ec)
    point, enh = np.zeros(72), np.zeros(7)
    fscale = np.abs(GOLD[f"pl{j}_force"]).max()
    for k in range(len(cur)):
        f, st, info = plastic_update(X, cur[k], j2, point, enh)
        gi = GOLD[f"pl{j}_info"][k]
        assert np.abs(f - GOLD[f"pl{j}_force"][k]).max() <= 1e-10 * fscale
        assert np.abs(st - GOLD[f"pl{j}_stress"][k]).max() <= 1e-10 * sy
        assert np.abs(point.reshape(8, 9) - GOLD[f"pl{j}_state"][k]).max() <= 1e-12
        assert np.abs(enh - GOLD[f"pl{j}_enhanced"][k]).max() <= 1e-12 * max(
            np.abs(GOLD[f"pl{j}_enhanced"]).max(), 1e-30)
        assert info[0] == pytest.approx(gi[0], rel=1e-9, abs=1e-12 * sy * 1e-3)
        assert list(info[1:]) == list(gi[1:])
    info_all = GOLD[f"pl{j}_info"]
    if j != 1:  # the failing paths tear, and some step ran the enhanced Newton
        assert info_all[-1, 5] == 1 and info_all[:, 1].max() >= 2
    else:
        assert info_all[:, 1].max() == 3 and info_all[-1, 3] == 8


@pytest.mark.parametrize("j", range(3))
def test_library_plastic_update_matches_compiled_reference(j):
    # the element code the device pass runs (hex_plastic.hpp), on the host through
    # shipfem_hex_plastic_update, along the same three paths: the co-rotation, the enhanced
    # Newton, the point returns with damage, the rerun without enhanced modes, the torn
    # element
    E_, nu_, sy, Hh, sens, er, ec = GOLD["pl_material"]
    X, cur = GOLD[f"pl{j}_rest"], GOLD[f"pl{j}_current"]
    ef = float(GOLD[f"pl{j}_failure_strain"])
    point, enh = np.zeros(64), np.zeros(7)
    fscale = np.abs(GOLD[f"pl{j}_force"]).max()
    escale = max(np.abs(GOLD[f"pl{j}_enhanced"]).max(), 1e-30)
    for k in range(len(cur)):
        gi = GOLD[f"pl{j}_info"][k]
        f, diss, failed = sf.hex_plastic_update(X, cur[k], point, enh, E_, nu_, sy, Hh, ef, sens, er, ec)
        gs = GOLD[f"pl{j}_state"][k]
        assert np.abs(f - GOLD[f"pl{j}_force"][k]).max() <= 1e-10 * fscale, k
        assert np.abs(point.reshape(8, 8) - gs[:, :8]).max() <= 1e-12, k
        assert np.array_equal(point.reshape(8, 8)[:, 7] >= 1.0, gs[:, 8] == 1.0)
        # the operator is formed in another order than the reference's stored rows: the
        # enhanced parameters carry Kaa's conditioning of that rounding (about 1e-12)
        assert np.abs(enh - GOLD[f"pl{j}_enhanced"][k]).max() <= 1e-10 * escale, k
        assert diss == pytest.approx(gi[0], rel=1e-9, abs=1e-12 * sy * 1e-3)
        assert failed == gi[4]


def test_plastic_update_below_yield_is_the_condensed_elastic_force():
    # the reference's tie between the two paths: a material that never yields gives
    # internalForce with the condensed stiffness (the EAS Newton solves the linear system
    # that condensation eliminated)
    # (a deformation of 1e-6 of the element's size: far below yield. The Newton's stop is
    # scaled by sigma_y V, so the yield strength stays a real one.)
    rng = np.random.default_rng(12)
    for x, nu in zip(elements(), GOLD["el_nu"]):
        K = sf.hex_stiffness(x, E, nu)
        cur = x + rng.normal(0, 1e-6 * np.ptp(x, axis=0).max(), (8, 3))
        j2 = O.J2(E, nu, 355e6, 1e9, 0.0, 1.5, 1 / 3, -1 / 3)
        f, _, info = plastic_update(x, cur, j2, np.zeros(72), np.zeros(7))
        ref = sf.hex_force(K, x, cur)
        assert np.abs(f - ref).max() <= 1e-8 * np.abs(ref).max()
        assert info[2] == 1 and info[3] == 0


def punch_case(speed, failure):
    X, index = H.plate_mesh(0.3, 0.15, 0.012, 12, 6, 1)
    N = len(X)
    flags = np.zeros(N, np.uint32)
    edge = (X[:, 0] < 1e-9) | (X[:, 0] > 0.3 - 1e-9) | (X[:, 1] < 1e-9) | (X[:, 1] > 0.15 - 1e-9)
    flags[edge] = O.PIN_ALL
    flags[~edge] |= O.CONTACT
    ind = dict(radius=0.04, center=(0.15, 0.075, -0.006 - 0.04), velocity=(0, 0, speed), kappa=0.1)
    K = H.stiffness(X[index], E, 0.3)
    dt = 0.9 * H.critical_timestep(K, H.mass(X[index], RHO)).min()
    return O.Case(X, index, E=E, nu=0.3, rho=RHO, dt=dt, flags=flags, indenter=ind, plastic=True,
                  yield_=355e6, hardening=1e9, failure=failure)


def test_oracle_plastic_hex_punch_account_and_tearing():
    # Without failure the account closes to the order of the plastic increments and the
    # rotations (about 1e-3 of the work at 9 mm of punch travel, 6% plastic strain).
    c = punch_case(2.0, None)
    rec = c.run(2000, threads=4, sample_every=500)
    r = rec[-1]
    res = r["w_contact"] - (r["strain"] + r["kinetic"]) - r["dissipation"]
    assert r["dissipation"] > 0.5 * r["w_contact"] and c.hex_eqps.max() > 0.03
    assert abs(res) <= 5e-3 * r["w_contact"]
    # With failure, points fail, elements tear and their steps are recorded, and the stored
    # energy of the failing points leaves as fracture. (The reference reruns an element
    # without its enhanced modes once a point fails, solid_shell.cpp:848-866; a thin
    # element then carries the locked thickness stress of the plain ANS hex, so the
    # account jumps at each failure: the oracle follows the reference there.)
    c = punch_case(20.0, dict(failure_strain=0.05))
    c.run(600, threads=4)
    torn = c.hex_damage >= 1.0
    assert torn.sum() > 0 and (c.tear_step[torn] > 0).all() and (c.tear_step[~torn] == -1).all()
    assert c.st.fracture > 0 and c.st.dissipation > 0
    # a torn element is not updated again: its points stay failed, its history frozen
    frozen = c.hex_point[torn].copy()
    c.run(50, threads=4)
    assert np.array_equal(c.hex_point[torn], frozen)


def test_hex_batch_refuses_mixed_elastic_and_plastic_before_touching_a_device():
    a = sf.HexCase.plate(0.2, 0.1, 0.01, 2, 2, 1)
    b = sf.HexCase.plate(0.2, 0.1, 0.01, 2, 2, 1, plastic=True)
    with pytest.raises(sf.ShipfemError) as e:
        sf.Batch([a, b])
    assert e.value.status == sf.INVALID_ARG


def test_hex_cases_round_trip_through_the_wire_format(tmp_path):
    # the header marks the hex kind: decode and load give back a HexCase with 8 ids per
    # element and every field; a flipped payload byte fails the checksum
    from shipfem import configs
    c = configs.hex_zone(lx=0.3, ly=0.15, nx=6, ny=3, radius=0.04, speed=5.0, plastic=True,
                         yield_=355e6, hardening=1e9, failure=dict(failure_strain=0.05))
    blob = c.encode()
    assert len(blob) == 64 + 216 + 24 * c.N + 32 * c.T + 4 * c.N
    d = sf.Case.decode(blob)
    assert isinstance(d, sf.HexCase) and d.index.shape == (c.T, 8)
    assert np.array_equal(d.X, c.X) and np.array_equal(d.index, c.index)
    assert np.array_equal(d.flags, c.flags) and d.plastic and d.yield_ == c.yield_
    assert d.failure["failure_strain"] == 0.05 and d.indenter["radius"] == 0.04
    assert d.encode() == blob
    c.save(str(tmp_path / "zone.case"))
    assert sf.Case.load(str(tmp_path / "zone.case")).encode() == blob
    bad = bytearray(blob)
    bad[64 + 216 + 8] ^= 1
    with pytest.raises(sf.ShipfemError):
        sf.Case.decode(bytes(bad))
    # a tet case stays a tet case
    t = sf.Case.decode(configs.plate(lx=0.2, ly=0.1, nx=4, ny=2).encode())
    assert type(t) is sf.Case and t.index.shape[1] == 4


def test_element_failure_strain_matches_initialise_plastic_state():
    # initialisePlasticState's regularised failure strain of the eight golden elements
    # (uniform 0.12, fracture 0.8): the oracle's restatement and the library's host function
    nodes = GOLD["el_nodes"]
    want = GOLD["fs_strain"]
    got_o = H.failure_strain(nodes, 0.12, 0.8)
    assert np.allclose(got_o, want, rtol=1e-13, atol=0)
    got_l = np.array([sf.hex_failure_strain(x, 0.12, 0.8) for x in nodes])
    assert np.allclose(got_l, want, rtol=1e-13, atol=0)
    # closed forms: an l x l x t box element reads share t / l; an element thicker than it
    # is wide reads the fracture strain; a collapsed mid-surface falls to fracture
    box = np.array([[0, 0, 0], [0.4, 0, 0], [0.4, 0.4, 0], [0, 0.4, 0],
                    [0, 0, 0.01], [0.4, 0, 0.01], [0.4, 0.4, 0.01], [0, 0.4, 0.01]], np.float64)
    assert sf.hex_failure_strain(box, 0.12, 0.8) == pytest.approx(0.12 + 0.68 * 0.01 / 0.4, rel=1e-13)
    tall = box * np.array([1.0, 1.0, 100.0])
    assert sf.hex_failure_strain(tall, 0.12, 0.8) == pytest.approx(0.8, rel=1e-15)
    flat = box.copy()
    flat[:, 1] = 0.0
    assert sf.hex_failure_strain(flat, 0.12, 0.8) == 0.8
    assert H.failure_strain(flat[None], 0.12, 0.8)[0] == 0.8
    # a wrong share (thickness over area instead of over the in-plane size) would miss
    assert not np.allclose(0.12 + 0.68 * np.minimum(1, H.element_size(nodes)[1] / H.element_size(nodes)[0] ** 2),
                           want, rtol=1e-6)

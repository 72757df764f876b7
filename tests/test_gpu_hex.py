"""The solid-shell hex solve on the device (shipfem_hex_solver_create through the C ABI)
against the FP64 oracle (oracle/hex.py element, oracle.c stepping).

Tolerances as for the tet path: relative L2 displacement error <= 1e-5 and the
energy-balance residual within 1e-4 of the work. The element forces of one pass on a
bent, twisted and rotated plate are checked first, then full solves: pressure on a
clamped plate, the zone punch, the prescribed edge drive with damping, and the reference's
absolute-coordinates failure case (a patch 13 m from the origin, engine/gpu/zone_gpu.cpp
centroid shift; docs/07-fem-spike-findings.md "Precision").
"""
import sys

import numpy as np
import pytest

import shipfem as sf
from shipfem import configs
from oracle import hex as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_U = 1e-5
ENERGY = 1e-4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    if sf.device_info() is None:
        pytest.skip("no usable gfx950 device")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def residual(r, r0):
    work = r["w_ext"] + r["w_contact"] + r["w_boundary"]
    return work - (r["strain"] + r["kinetic"] - r0["strain"] - r0["kinetic"]) - \
        r["dissipation"] - r["damping_loss"] - r["fracture"]


def energy_gap(g_rec, o_rec):
    assert [r["step"] for r in g_rec] == [r["step"] for r in o_rec] and len(o_rec) > 1
    w = max(abs(o_rec[-1]["w_ext"]) + abs(o_rec[-1]["w_contact"]) + abs(o_rec[-1]["w_boundary"]),
            1e-30)
    return max(abs(residual(g, g_rec[0]) - residual(o, o_rec[0])) for g, o in zip(g_rec, o_rec)) / w


def oracle_for(case, dt):
    o = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    if case.u0 is not None:
        o.u[:] = case.u0
    if case.v0 is not None:
        o.v[:] = case.v0
    return o


def test_element_forces_of_a_bent_twisted_rotated_plate():
    # a plate bent into an arc about y, twisted about x, then rotated 25 degrees about z
    lx, ly, t = 0.6, 0.3, 0.012
    case = sf.HexCase.plate(lx, ly, t, 12, 6, 1)
    X = case.X
    kappa, twist = 0.8, 0.4
    z = X[:, 2] + 0.5 * kappa * X[:, 0] ** 2 + twist * X[:, 0] * X[:, 1]
    cur = np.stack([X[:, 0] - 0.5 * kappa ** 2 * X[:, 0] ** 3 / 3, X[:, 1], z], 1)
    a = np.deg2rad(25.0)
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    cur = cur @ Rz.T
    case.u0 = np.ascontiguousarray(cur - X)
    s = sf.Solver(case)
    s.run(1)
    f = s.tet_forces().astype(np.float64)
    K = H.stiffness(X[case.index], case.E, case.nu)
    ref = H.force(K, X[case.index], cur[case.index])
    err = rel_l2(f, ref)
    worst = float(np.abs(f - ref).max() / np.abs(ref).max())
    say(f"hex forces: rel L2 {err:.3e}, max {worst:.3e} (|f| <= {np.abs(ref).max():.3e} N)")
    assert err <= 1e-5 and worst <= 1e-5


def test_clamped_plate_under_pressure_matches_oracle():
    case = configs.hex_plate(lx=0.6, ly=0.3, nx=24, ny=12, pressure=2e5, ramp=1e-3,
                             damping=0.9995, sample_every=200)
    s = sf.Solver(case)
    ts = s.timestep()
    K = H.stiffness(case.X[case.index], case.E, case.nu)
    dt_ref = H.critical_timestep(K, H.mass(case.X[case.index], case.rho), 1.0).min()
    say(f"hex plate {case.T} elements: dt {ts['dt']:.4e}, 2/omega {ts['cfl']:.6e} "
        f"(oracle {dt_ref:.6e})")
    assert ts["cfl"] == pytest.approx(dt_ref, rel=1e-9)
    o = oracle_for(case, s.dt)
    ro = o.run(3000, sample_every=200)
    s.run(3000)
    st = s.readback()
    g = s.energy()
    err = rel_l2(st.u, o.u)
    gap = energy_gap(g, ro)
    say(f"pressure plate 3000 steps: u rel L2 {err:.3e}, max |u| {np.abs(o.u).max():.3e} m, "
        f"energy gap {gap:.3e}")
    assert err <= REL_U and gap <= ENERGY
    for a, b in zip(g[1:], ro[1:]):
        assert a["strain"] == pytest.approx(b["strain"], rel=1e-4)
        assert a["w_ext"] == pytest.approx(b["w_ext"], rel=1e-4)


def test_zone_punch_matches_oracle():
    case = configs.hex_zone(nx=32, ny=16, speed=5.0)
    s = sf.Solver(case)
    o = oracle_for(case, s.dt)
    ro = o.run(4000, sample_every=100)
    s.run(4000)
    st = s.readback()
    g = s.energy()
    err = rel_l2(st.u, o.u)
    gap = energy_gap(g, ro)
    fz_g = max(abs(r["contact_force"][2]) for r in g)
    fz_o = max(abs(r["contact_force"][2]) for r in ro)
    say(f"zone punch 4000 steps: u rel L2 {err:.3e}, contact work {ro[-1]['w_contact']:.3e} J, "
        f"peak force {fz_g:.4e} / {fz_o:.4e} N, energy gap {gap:.3e}")
    assert ro[-1]["w_contact"] > 0 and fz_o > 0
    assert err <= REL_U and gap <= ENERGY
    assert fz_g == pytest.approx(fz_o, rel=1e-4)


def test_edge_drive_with_damping_matches_oracle():
    lx, ly = 0.6, 0.3
    case = sf.HexCase.plate(lx, ly, 0.012, 24, 12, 1)
    X = case.X
    flags = np.zeros(case.N, np.uint32)
    pres = np.zeros((case.N, 3))
    xe = X[:, 0] > lx - 1e-9
    x0 = X[:, 0] < 1e-9
    flags[x0] = sf.PIN_ALL
    flags[xe] = sf.DRIVE_X | sf.DRIVE_Y | sf.DRIVE_Z
    pres[xe, 0] = 3e-5
    pres[xe, 2] = 2e-3 * (X[xe, 1] / ly - 0.5)  # a twist of the free edge
    case.flags, case.prescribed = flags, pres
    case.drive_ramp, case.damping, case.sample_every = 1e-3, 0.999, 100
    s = sf.Solver(case)
    o = oracle_for(case, s.dt)
    ro = o.run(2000, sample_every=100)
    s.run(2000)
    st = s.readback()
    err = rel_l2(st.u, o.u)
    gap = energy_gap(s.energy(), ro)
    say(f"edge drive + damping 2000 steps: u rel L2 {err:.3e}, W_bnd {ro[-1]['w_boundary']:.3e} J, "
        f"damping loss {ro[-1]['damping_loss']:.3e} J, energy gap {gap:.3e}")
    assert np.abs(st.u[xe] - pres[xe]).max() <= 1e-6 * np.abs(pres).max()
    assert err <= REL_U and gap <= ENERGY and ro[-1]["damping_loss"] > 0


def test_patch_far_from_the_origin_keeps_parity():
    # The reference's FP32 shader lost the first thousand steps' displacements against
    # |x| ~ 13 m coordinates until it shifted each patch to its centroid. This solve keeps
    # u apart from X and works on corner offsets, so a patch at (3, -9.9, 8) m matches the
    # oracle as well as one at the origin, and the two runs agree with each other.
    near = configs.hex_zone(nx=32, ny=16, speed=5.0)
    far = configs.hex_zone(nx=32, ny=16, speed=5.0, origin=(3.0, -9.9, 8.0))
    sn, sf_ = sf.Solver(near), sf.Solver(far)
    o = oracle_for(far, sf_.dt)
    o.run(1000)
    sn.run(1000)
    sf_.run(1000)
    un, uf = sn.readback().u, sf_.readback().u
    err, pair = rel_l2(uf, o.u), rel_l2(uf, un)
    say(f"patch at 13 m, 1000 steps: u rel L2 vs oracle {err:.3e}, vs the patch at the origin "
        f"{pair:.3e}")
    assert err <= REL_U and pair <= REL_U


def test_hex_solve_is_deterministic_and_graph_replay_matches_plain_launches():
    case = configs.hex_zone(nx=24, ny=12, speed=5.0)
    a, b, c = sf.Solver(case), sf.Solver(case), sf.Solver(case)
    a.run(600, mode=0)
    b.run(600, mode=0)
    c.run(600, mode=1)
    ua, ub, uc = a.readback(), b.readback(), c.readback()
    assert np.array_equal(ua.u, ub.u) and np.array_equal(ua.v, ub.v)
    assert np.array_equal(ua.u, uc.u) and np.array_equal(ua.v, uc.v)
    assert a.energy() == b.energy() == c.energy()
    # launch shape does not change the result (one lane per element; fixed gather order)
    d = sf.Solver(case).set_launch(sf.VARIANT_SLOTS, node_block=64, elem_block=128)
    d.run(600)
    assert np.array_equal(d.readback().u, ua.u)


def test_hex_batch_matches_each_case_alone_and_tunes():
    # one launch per phase over several hex cases: every case ends with the bits it has
    # when solved alone, whatever else shares the launch and whatever shape the tuner picks
    rng = np.random.default_rng(3)
    cases = [configs.hex_zone(nx=int(n), ny=int(n) // 2, speed=float(v))
             for n, v in zip(rng.integers(8, 40, 6), rng.uniform(2.0, 8.0, 6))]
    cases.append(configs.hex_plate(lx=0.6, ly=0.3, nx=20, ny=10, pressure=1e5, ramp=1e-3))
    b = sf.Batch(cases)
    b.run(400)
    alone = []
    for c in cases:
        s = sf.Solver(c)
        s.run(400)
        alone.append(s.readback().u)
    for i, u in enumerate(alone):
        assert np.array_equal(b.readback(i).u, u), i
    best = b.tune(steps_per_trial=100)
    say(f"hex batch of {len(cases)}: tuned {best}")
    b.reset()
    b.run(400)
    for i, u in enumerate(alone):
        assert np.array_equal(b.readback(i).u, u), i


# ---------------------------------------------------------------- plastic solid-shell
PLASTIC = dict(plastic=True, yield_=355e6, hardening=1e9)
ABS_EQPS = 1e-4


def plastic_zone(speed, failure=None, nx=24, ny=12, **kw):
    return configs.hex_zone(lx=0.3, ly=0.15, nx=nx, ny=ny, radius=0.04, speed=speed,
                            failure=failure, **PLASTIC, **kw)


def test_plastic_element_pass_matches_the_compiled_reference_update():
    # one element pass from a fresh history at configurations of the golden plastic paths
    # (tests/golden/reference_hex.npz): forces and the committed point history against the
    # oracle's update, which those vectors pin to the compiled elementPlasticUpdate
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_hex.npz"))
    E_, nu_, sy, Hh, sens, er, ec = g["pl_material"]
    for j in range(3):
        X = g[f"pl{j}_rest"]
        ef = float(g[f"pl{j}_failure_strain"])
        k = 30 if j == 1 else 4  # configurations past first yield on each path
        cur = g[f"pl{j}_current"][k]
        case = sf.HexCase(X, np.arange(8, dtype=np.uint32)[None], E=E_, nu=nu_, plastic=True,
                          yield_=sy, hardening=Hh, u0=cur - X,
                          failure=dict(failure_strain=ef) if ef > 0 else None)
        s = sf.Solver(case)
        s.run(1)
        f = s.tet_forces().astype(np.float64)[0]
        o = oracle_for(case, s.dt)
        o.run(1)
        err = float(np.abs(f - o.force[0]).max() / np.abs(o.force[0]).max())
        eq = float(abs(s.readback().eqps[0] - o.hex_eqps[0]))
        say(f"plastic element pass, path {j}: force max rel {err:.2e}, eqps |d| {eq:.2e} "
            f"(eqps {o.hex_eqps[0]:.3e})")
        assert err <= 1e-5 and eq <= 1e-7 and o.hex_eqps[0] > 0


def test_plastic_punch_matches_oracle():
    case = plastic_zone(5.0)
    s = sf.Solver(case)
    o = oracle_for(case, s.dt)
    ro = o.run(3000, sample_every=100, threads=8)
    s.run(3000)
    st = s.readback()
    g = s.energy()
    err = rel_l2(st.u, o.u)
    deq = float(np.abs(st.eqps - o.hex_eqps).max())
    gap = energy_gap(g, ro)
    dd = abs(g[-1]["dissipation"] - ro[-1]["dissipation"]) / ro[-1]["dissipation"]
    say(f"plastic punch 3000 steps: u rel L2 {err:.3e}, eqps |d| {deq:.2e} (max {o.hex_eqps.max():.3f}), "
        f"dissipation {ro[-1]['dissipation']:.3e} J (rel {dd:.1e}), energy gap {gap:.3e}")
    assert o.hex_eqps.max() > 0.01
    assert err <= REL_U and deq <= ABS_EQPS and gap <= ENERGY and dd <= 1e-4


def test_plastic_punch_with_tearing_matches_oracle():
    case = plastic_zone(10.0, failure=dict(failure_strain=0.06))
    s = sf.Solver(case)
    o = oracle_for(case, s.dt)
    ro = o.run(1500, sample_every=100, threads=8)
    s.run(1500)
    st = s.readback()
    torn_g = set(np.nonzero(st.damage >= 1)[0])
    torn_o = set(np.nonzero(o.hex_damage >= 1)[0])
    tg = {t: st_ for t, st_, _ in s.tears()}
    same_steps = sum(1 for t in torn_g & torn_o if tg[t] == o.tear_step[t])
    err = rel_l2(st.u, o.u)
    deq = float(np.abs(st.eqps - o.hex_eqps).max())
    say(f"plastic punch with tearing, 1500 steps: torn {len(torn_g)} / {len(torn_o)} "
        f"(same step {same_steps}), u rel L2 {err:.3e}, eqps |d| {deq:.2e}, "
        f"fracture {ro[-1]['fracture']:.3e} J")
    assert len(torn_o) > 0 and len(torn_g ^ torn_o) <= max(1, len(torn_o) // 50)
    assert err <= REL_U and deq <= ABS_EQPS
    # a checkpoint carries the eight points of every element: the restored solver
    # continues with the same bits
    blob = s.checkpoint()
    s.run(200)
    r = sf.Solver(case).restore(blob)
    r.run(200)
    a, b = s.readback(), r.readback()
    assert np.array_equal(a.u, b.u) and np.array_equal(a.eqps, b.eqps)
    # the hand-off carries the same points: a torn element has eight failed points and
    # its tear step
    ex = s.export()
    torn_x = np.nonzero((ex["hex_points"][:, :, 7] >= 1).all(1))[0]
    assert set(torn_x) == set(np.nonzero(a.damage >= 1)[0])
    assert {t: int(ex["tear_step"][t]) for t in torn_x} == {t: st_ for t, st_, _ in s.tears()}


def jittered(case, amp, seed, lx, ly):
    # in-plane jitter of every node off the clamped edges (the same offset through the
    # thickness keeps the element's thickness direction), plus a small warp out of plane:
    # trapezoidal, warped, non-affine elements
    rng = np.random.default_rng(seed)
    X = case.X.copy()
    nxy = np.unique(np.round(X[:, :2], 12), axis=0, return_inverse=True)[1].ravel()
    off = rng.uniform(-amp, amp, (nxy.max() + 1, 3))
    off[:, 2] *= 0.05
    inner = (X[:, 0] > 1e-9) & (X[:, 0] < lx - 1e-9) & (X[:, 1] > 1e-9) & (X[:, 1] < ly - 1e-9)
    X[inner] += off[nxy[inner]]
    case.X = np.ascontiguousarray(X)
    return case


def test_jittered_two_layer_plate_matches_oracle():
    # elastic: non-affine, warped elements, two through the thickness
    lx, ly = 0.6, 0.3
    case = configs.hex_plate(lx=lx, ly=ly, thickness=0.02, nx=24, ny=12, nz=2, pressure=2e5,
                             ramp=1e-3, damping=0.9995)
    case = jittered(case, 0.25 * lx / 24, 11, lx, ly)
    s = sf.Solver(case)
    o = oracle_for(case, s.dt)
    ro = o.run(2000, sample_every=100)
    s.run(2000)
    err = rel_l2(s.readback().u, o.u)
    gap = energy_gap(s.energy(), ro)
    say(f"jittered two-layer plate, {case.T} elements, 2000 steps: u rel L2 {err:.3e}, gap {gap:.3e}")
    assert err <= REL_U and gap <= ENERGY


def test_jittered_plastic_zone_matches_oracle():
    case = plastic_zone(5.0, failure=dict(failure_strain=0.08), nx=20, ny=10)
    case = jittered(case, 0.2 * 0.3 / 20, 5, 0.3, 0.15)
    s = sf.Solver(case)
    o = oracle_for(case, s.dt)
    o.run(2500, threads=8)
    s.run(2500)
    st = s.readback()
    err = rel_l2(st.u, o.u)
    deq = float(np.abs(st.eqps - o.hex_eqps).max())
    torn_g = set(np.nonzero(st.damage >= 1)[0])
    torn_o = set(np.nonzero(o.hex_damage >= 1)[0])
    say(f"jittered plastic zone, 2500 steps: u rel L2 {err:.3e}, eqps |d| {deq:.2e} "
        f"(max {o.hex_eqps.max():.3f}), torn {len(torn_g)} / {len(torn_o)}")
    assert o.hex_eqps.max() > 0.01
    assert err <= REL_U and deq <= ABS_EQPS and len(torn_g ^ torn_o) <= max(1, len(torn_o) // 50)


def test_graded_plate_tears_at_each_elements_own_failure_strain():
    # initialisePlasticState (engine/sim/zone.cpp:1248-1280) gives every element its own
    # failure strain from its size when a fracture strain is set: a plate graded along x
    # (elements three times as long under the punch as at the clamped ends, thinner than
    # they are wide) and jittered spreads it over the elements that tear
    lx, ly, nx, ny = 0.3, 0.15, 20, 10
    case = plastic_zone(10.0, failure=dict(failure_strain=0.04, fracture_strain=0.07), nx=nx, ny=ny)
    s_ = case.X[:, 0] / lx
    case.X[:, 0] = lx * (s_ - 0.5 * np.sin(2 * np.pi * s_) / (2 * np.pi))
    case = jittered(case, 0.15 * ly / ny, 3, lx, ly)
    ef = H.failure_strain(case.X[case.index], 0.04, 0.07)
    s = sf.Solver(case)
    o = oracle_for(case, s.dt)
    ro = o.run(2500, sample_every=100, threads=8)
    assert np.array_equal(o.hex_failure, ef)
    s.run(2500)
    st = s.readback()
    torn_g = set(np.nonzero(st.damage >= 1)[0])
    torn_o = set(np.nonzero(o.hex_damage >= 1)[0])
    tg = {t: st_ for t, st_, _ in s.tears()}
    same_steps = sum(1 for t in torn_g & torn_o if tg[t] == o.tear_step[t])
    err = rel_l2(st.u, o.u)
    deq = float(np.abs(st.eqps - o.hex_eqps).max())
    torn_ef = ef[sorted(torn_o)] if torn_o else ef[:0]
    say(f"graded plate, per-element failure strain {ef.min():.4f}-{ef.max():.4f}: torn {len(torn_g)} / "
        f"{len(torn_o)} (same step {same_steps}; their strains {torn_ef.min() if len(torn_ef) else 0:.4f}-"
        f"{torn_ef.max() if len(torn_ef) else 0:.4f}), u rel L2 {err:.3e}, eqps |d| {deq:.2e}, "
        f"fracture {ro[-1]['fracture']:.3e} J")
    assert len(torn_ef) > 4 and torn_ef.max() - torn_ef.min() > 0.002
    assert len(torn_o) > 0 and len(torn_g ^ torn_o) <= max(1, len(torn_o) // 50)
    assert same_steps >= len(torn_g & torn_o) - max(1, len(torn_o) // 50)
    assert err <= REL_U and deq <= ABS_EQPS


def test_plastic_hex_batch_matches_each_case_alone():
    cases = [plastic_zone(v, failure=dict(failure_strain=f), nx=n, ny=n // 2)
             for v, f, n in ((4.0, 0.1, 16), (6.0, 0.06, 20), (8.0, 0.08, 12))]
    b = sf.Batch(cases)
    b.run(600)
    for i, c in enumerate(cases):
        s = sf.Solver(c)
        s.run(600)
        a, bb = s.readback(), b.readback(i)
        assert np.array_equal(a.u, bb.u) and np.array_equal(a.eqps, bb.eqps), i
        assert np.array_equal(a.damage, bb.damage), i

"""A solve's state across the device boundary, against the FP64 oracle.

- checkpoint/restore: a restored solver continues with the same bits;
- adopt (zone::Solver::adopt, zone.cpp:1321-1368; its identity test is
  test_zone.cpp:1729-1800): adopting a run's own export reproduces the run, and does
  not advance the plastic history it was handed;
- GPU -> CPU: the oracle adopts the device's export and continues; CPU -> GPU: the
  device adopts the oracle's state and continues; both at the parity tolerances;
- the Tier-1 edge drive (coupling::applyEdgeDrive, coupling.cpp:110-125) in a solve.
"""
import sys

import numpy as np
import pytest

import shipfem as sf
from shipfem import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_U, ABS_EQPS, ENERGY = 1e-5, 1e-4, 1e-4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    if sf.device_info() is None:
        pytest.skip("no usable gfx950 device")


def tearing_panel():
    return configs.stiffened_panel(nx=40, ny_bay=6, nz_web=4, speed=10.0, radius=0.05,
                                   geometry=dict(lx=0.6, ly=0.3, n_stiff=2, web_h=0.05),
                                   failure=dict(failure_strain=0.08))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def residual(r, r0):
    work = r["w_ext"] + r["w_contact"] + r["w_boundary"]
    return work - (r["strain"] + r["kinetic"] - r0["strain"] - r0["kinetic"]) - \
        r["dissipation"] - r["damping_loss"] - r["fracture"]


def state_of(s):
    st = s.readback()
    return st.u.copy(), st.v.copy(), st.eqps.copy(), st.damage.copy(), st.step


def test_checkpoint_restore_continues_with_the_same_bits():
    case = tearing_panel()
    a = sf.Solver(case)
    a.run(600)
    blob = a.checkpoint()
    a.run(600)
    b = sf.Solver(case).restore(blob)
    assert b.readback().step == 600
    b.run(600)
    ua, va, qa, da, na = state_of(a)
    ub, vb, qb, db, nb = state_of(b)
    say(f"checkpoint {len(blob) / 1e6:.2f} MB; torn {int((da >= 1).sum())}, max eqps {qa.max():.3f}")
    assert na == nb == 1200
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)
    assert np.array_equal(qa, qb) and np.array_equal(da, db)
    ta, tb = a.tears(), b.tears()
    assert len(ta) == len(tb) > 0
    assert all(x[0] == y[0] and x[1] == y[1] and np.array_equal(x[2], y[2]) for x, y in zip(ta, tb))
    # the history restored with it: the records after the checkpoint are the same
    ea, eb = a.energy(), b.energy()
    assert ea == eb and len(ea) > 6
    assert (da >= 1).sum() > 0 and qa.max() > 1e-3  # the run tore and yielded
    # a checkpoint of another case is refused
    other = sf.Solver(configs.cantilever(nx=10, ny=2, nz=2))
    with pytest.raises(sf.ShipfemError) as e:
        other.restore(blob)
    assert e.value.status == sf.INVALID_ARG


def test_adopting_a_runs_own_export_reproduces_the_run():
    case = tearing_panel()
    a = sf.Solver(case)
    a.run(1000)
    ex = a.export()
    b = sf.Solver(case).adopt(ex)
    back = b.export()
    # the identity: the adopted solver reports what it was handed
    for k in ("steps", "time", "torn", "penetration"):
        assert back[k] == ex[k], k
    for k in ("displacement", "velocity", "plastic", "tear_step"):
        assert np.array_equal(back[k], ex[k]), k
    for k in ("w_ext", "w_contact", "w_boundary", "damping_loss", "dissipation", "fracture"):
        assert back[k] == pytest.approx(ex[k], rel=1e-14, abs=1e-300), k
    assert ex["torn"] > 0 and ex["plastic"][:, 6].max() > 1e-3 and ex["w_contact"] > 0
    # adopting did not advance the plastic history
    assert np.array_equal(b.readback().eqps, a.readback().eqps)
    # Both continue as one solve. The double value of each FP32 + compensation pair is
    # exact, but a compensation below double resolution of its value does not split
    # back into the same two floats, and the tearing and contact thresholds amplify
    # that last-bit difference: the continuation holds to the parity bar, not bitwise
    # (the checkpoint is the bitwise path).
    a.run(400)
    b.run(400)
    ua, va, qa, da, _ = state_of(a)
    ub, vb, qb, db, _ = state_of(b)
    ra = [r for r in a.energy() if r["step"] >= 1000]
    rb = b.energy()
    assert [r["step"] for r in ra] == [r["step"] for r in rb]
    w = abs(ra[-1]["w_contact"])
    gap = max(abs(residual(x, ra[0]) - residual(y, rb[0])) for x, y in zip(ra, rb)) / w
    say(f"adopted own export at 1000, +400 steps: u rel {rel_l2(ub, ua):.2e}, "
        f"eqps |d| {np.abs(qb - qa).max():.2e}, energy gap {gap:.2e}")
    assert rel_l2(ub, ua) <= REL_U and np.abs(qb - qa).max() <= ABS_EQPS and gap <= ENERGY
    # position - X instead of the exact displacement: u's last bits go to X's magnitude
    c = sf.Solver(case).adopt({k: v for k, v in ex.items() if k != "displacement"})
    err = np.abs(c.export()["displacement"] - ex["displacement"]).max()
    say(f"adopt via position: max |du| {err:.2e} m (|X| <= {np.abs(case.X).max():.2f} m)")
    assert err <= 4 * np.finfo(np.float64).eps * np.abs(case.X).max()


def test_adopting_an_elastic_runs_export_continues_to_rounding():
    # no thresholds to amplify a last-bit difference: the continuation stays at rounding
    case = configs.plate(lx=0.5, ly=0.2, nx=20, ny=8, ramp=1e-4)
    a = sf.Solver(case)
    a.run(300)
    b = sf.Solver(case).adopt(a.export())
    a.run(300)
    b.run(300)
    d = rel_l2(b.readback().u, a.readback().u)
    say(f"elastic adopt, +300 steps: u rel {d:.2e}")
    assert d <= 1e-9


def test_adopt_refuses_a_time_that_is_not_its_step_count():
    case = configs.cantilever(nx=10, ny=2, nz=2)
    s = sf.Solver(case)
    s.run(10)
    ex = s.export()
    ex["time"] *= 1.5
    with pytest.raises(sf.ShipfemError) as e:
        sf.Solver(case).adopt(ex)
    assert e.value.status == sf.INVALID_ARG


def test_gpu_state_continues_on_the_cpu_and_back():
    # GPU -> CPU: the device runs n steps, the oracle adopts its export and continues;
    # the device continues too. CPU -> GPU: the oracle's state at n adopted by a fresh
    # device solver. Both continuations against the oracle's own uninterrupted run.
    case = tearing_panel()
    n, m = 500, 500
    s = sf.Solver(case)
    dt = s.dt
    full = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    rf = full.run(n + m, sample_every=case.sample_every)
    s.run(n)
    ex = s.export()
    cpu = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs()).adopt(ex)
    rc = cpu.run(m, sample_every=case.sample_every)
    s.run(m)
    st = s.readback()
    err_cpu = rel_l2(cpu.u, full.u)
    err_gpu = rel_l2(st.u, cpu.u)
    deq = float(np.max(np.abs(st.eqps - cpu.eqps)))
    rg = [r for r in s.energy() if r["step"] >= n]
    assert [r["step"] for r in rg] == [r["step"] for r in rc] and len(rc) > 2
    w = max(abs(rf[-1]["w_contact"]), 1e-30)
    r0 = rf[0]
    gap = max(abs(residual(g, r0) - residual(c, r0)) for g, c in zip(rg, rc)) / w
    say(f"GPU->CPU at {n}: CPU continuation vs uninterrupted oracle u {err_cpu:.2e}; "
        f"GPU vs CPU continuation u {err_gpu:.2e}, eqps {deq:.2e}, energy gap {gap:.2e}")
    assert err_cpu <= REL_U and err_gpu <= REL_U and deq <= ABS_EQPS and gap <= ENERGY

    part = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    part.run(n)
    g = sf.Solver(case).adopt(dict(part.export(), time=n * dt))
    g.run(m)
    sg = g.readback()
    err = rel_l2(sg.u, full.u)
    deq = float(np.max(np.abs(sg.eqps - full.eqps)))
    torn_g, torn_o = set(np.nonzero(sg.damage >= 1)[0]), set(np.nonzero(full.damage >= 1)[0])
    say(f"CPU->GPU at {n}: u {err:.2e}, eqps {deq:.2e}, torn {len(torn_g)} / {len(torn_o)}")
    assert err <= REL_U and deq <= ABS_EQPS
    assert len(torn_g ^ torn_o) <= max(2, len(torn_o) // 50)


def test_tier1_edge_drive_matches_oracle():
    # a plate panel whose x-edges follow an interface displacement from the surrounding
    # model (a linear stretch plus a bending rotation about y), applied by the edge
    # drive over a smoothstep ramp; the y-edges the drive does not name stay clamped
    lx, ly, lz = 0.5, 0.2, 0.02
    case = configs.plate(lx=lx, ly=ly, lz=lz, nx=20, ny=8, nz=2, pressure=0.0, jitter=0.0)
    X = case.X
    flags = case.flags.copy()
    pres = np.zeros((case.N, 3))
    disp = np.zeros((case.N, 3))
    disp[:, 0] = 2e-6 * X[:, 0] / lx + 1e-4 * (X[:, 2] - lz / 2) * (X[:, 0] - lx / 2) / lx
    disp[:, 2] = -5e-5 * (X[:, 0] / lx - 0.5) ** 2
    edge_x = (X[:, 0] < 1e-9) | (X[:, 0] > lx - 1e-9)
    driven = np.zeros((case.N, 3), np.uint8)
    driven[edge_x] = 1
    n = sf.apply_edge_drive(flags, pres, disp, driven)
    assert n == 3 * edge_x.sum()
    c2 = sf.Case(X, case.index, flags=flags, prescribed=pres, drive_ramp=2e-4,
                 sample_every=100)
    s = sf.Solver(c2)
    o = O.Case(X, case.index, dt=s.dt, **c2.oracle_kwargs())
    ora = o.run(1600, sample_every=100)
    s.run(1600)
    st = s.readback()
    g = s.energy()
    err = rel_l2(st.u, o.u)
    wb = ora[-1]["w_boundary"]
    dw = max(abs(a["w_boundary"] - b["w_boundary"]) for a, b in zip(g, ora)) / abs(wb)
    y_edge = ((X[:, 1] < 1e-9) | (X[:, 1] > ly - 1e-9)) & ~edge_x
    say(f"edge drive: {n} DOF driven, u rel L2 {err:.2e}, W_bnd {wb:.3e} J (gap {dw:.2e}), "
        f"max |u| on clamped y-edges {np.abs(st.u[y_edge]).max():.1e}")
    assert np.allclose(st.u[edge_x], disp[edge_x], rtol=0, atol=1e-5 * np.abs(disp).max())
    assert np.abs(st.u[y_edge]).max() == 0.0
    assert err <= REL_U and dw <= ENERGY and wb > 0


def test_plastic_hex_history_crosses_the_boundary_both_ways():
    # a plastic solid-shell zone that tears: every element's eight points (eps_p, eqps,
    # damage) and enhanced parameters go out in the point order of the oracle
    # (solid_shell.cpp's Gauss order), a device adopting its own export continues with the
    # same bits, and each side continues the other's state at the parity tolerances
    case = configs.hex_zone(lx=0.3, ly=0.15, nx=24, ny=12, radius=0.04, speed=10.0, plastic=True,
                            yield_=355e6, hardening=1e9, failure=dict(failure_strain=0.06))
    n, m = 700, 800
    s = sf.Solver(case)
    dt = s.dt
    kw = case.oracle_kwargs()
    full = O.Case(case.X, case.index, dt=dt, **kw)
    full.run(n + m, threads=8)
    half = O.Case(case.X, case.index, dt=dt, **kw)
    half.run(n, threads=8)
    s.run(n)
    ex = s.export()
    dq = float(np.abs(ex["hex_points"][:, :, 6] - half.hex_point[:, :, 6]).max())
    dp = float(np.abs(ex["hex_points"][:, :, :6] - half.hex_point[:, :, :6]).max())
    say(f"hex export at {n}: per-point eqps |d| {dq:.2e} (max {half.hex_point[:, :, 6].max():.3f}), "
        f"eps_p |d| {dp:.2e}")
    assert half.hex_point[:, :, 6].max() > 0.01 and dq <= ABS_EQPS and dp <= ABS_EQPS
    r0 = s.readback()  # readback's per-element eqps and damage are the export's summary
    assert np.allclose(ex["plastic"][:, 6], r0.eqps, rtol=1e-6, atol=0)
    assert np.allclose(ex["plastic"][:, 7], r0.damage, rtol=1e-6, atol=0)
    b = sf.Solver(case).adopt(ex)
    cpu = O.Case(case.X, case.index, dt=dt, **kw).adopt(ex)
    cpu.run(m, threads=8)
    s.run(m)
    b.run(m)
    sa, sb = s.readback(), b.readback()
    assert np.array_equal(sa.u, sb.u) and np.array_equal(sa.v, sb.v)
    assert np.array_equal(sa.eqps, sb.eqps) and np.array_equal(sa.damage, sb.damage)
    assert b.export()["tear_step"].tolist() == s.export()["tear_step"].tolist()
    err = rel_l2(sa.u, cpu.u)
    deq = float(np.abs(sa.eqps - cpu.hex_eqps).max())
    torn_g, torn_c = set(np.nonzero(sa.damage >= 1)[0]), set(np.nonzero(cpu.hex_damage >= 1)[0])
    say(f"hex GPU->CPU at {n}: u {err:.2e}, eqps {deq:.2e}, torn {len(torn_g)} / {len(torn_c)}")
    assert len(torn_c) > 0 and err <= REL_U and deq <= ABS_EQPS
    assert len(torn_g ^ torn_c) <= max(1, len(torn_c) // 50)

    g = sf.Solver(case).adopt(dict(half.export(), time=n * dt))
    g.run(m)
    sg = g.readback()
    err = rel_l2(sg.u, full.u)
    deq = float(np.abs(sg.eqps - full.hex_eqps).max())
    torn_g, torn_o = set(np.nonzero(sg.damage >= 1)[0]), set(np.nonzero(full.hex_damage >= 1)[0])
    say(f"hex CPU->GPU at {n}: u {err:.2e}, eqps {deq:.2e}, torn {len(torn_g)} / {len(torn_o)}")
    assert err <= REL_U and deq <= ABS_EQPS
    assert len(torn_g ^ torn_o) <= max(1, len(torn_o) // 50)

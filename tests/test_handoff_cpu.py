"""Hand-off pieces that run on the host: the Tier-1 edge drive's flag rules
(coupling::applyEdgeDrive, coupling.cpp:110-125) and the oracle's own adopt, which the
GPU hand-off tests use as the CPU side (zone::Solver::adopt, zone.cpp:1321-1368).
"""
import numpy as np

import shipfem as sf
from shipfem import configs
from oracle import oracle as O


def test_apply_edge_drive_replaces_clamps_and_never_unpins():
    N = 4
    flags = np.array([sf.PIN_ALL, sf.PIN_ALL, sf.PIN_X, 0], np.uint32)
    pres = np.zeros(3 * N)
    disp = np.arange(1.0, 3 * N + 1).reshape(N, 3) * 1e-3
    driven = np.zeros((N, 3), np.uint8)
    driven[0, :] = 1          # a clamped node driven on all three axes
    driven[1, 2] = 1          # a clamped node driven along z only
    driven[3, 0] = 1          # a free node gets a drive (the drive pins it)
    n = sf.apply_edge_drive(flags, pres, disp, driven)
    assert n == 5
    assert flags[0] == sf.DRIVE_X | sf.DRIVE_Y | sf.DRIVE_Z
    # x and y of node 1 are not named: they stay clamped; z is driven
    assert flags[1] == sf.PIN_X | sf.PIN_Y | sf.DRIVE_Z
    assert flags[2] == sf.PIN_X  # untouched
    assert flags[3] == sf.DRIVE_X
    p = pres.reshape(N, 3)
    assert np.array_equal(p[0], disp[0]) and p[1, 2] == disp[1, 2] and p[3, 0] == disp[3, 0]
    assert p[1, 0] == 0 and p[1, 1] == 0 and (p[2] == 0).all()
    # a drive naming nothing writes nothing
    assert sf.apply_edge_drive(flags, pres, disp, np.zeros((N, 3), np.uint8)) == 0


def test_oracle_adopting_its_own_state_continues_the_same_run():
    # plastic indentation with failure: adopt at step 150 and continue to 300; the
    # continuation is the same FP64 run, so state and running sums are identical
    case = configs.stiffened_panel(nx=12, ny_bay=3, nz_web=2, speed=10.0, radius=0.05,
                                   geometry=dict(lx=0.3, ly=0.15, n_stiff=1, web_h=0.03),
                                   failure=dict(failure_strain=0.08))
    dt = 2e-7
    a = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    ra = a.run(300, sample_every=50)
    b = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    b.run(150)
    c = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs()).adopt(b.export())
    rc = c.run(150, sample_every=50)
    assert c.step == a.step == 300
    assert np.array_equal(c.u, a.u) and np.array_equal(c.v, a.v)
    assert np.array_equal(c.eqps, a.eqps) and np.array_equal(c.damage, a.damage)
    assert rc == [r for r in ra if r["step"] >= 150]
    assert a.eqps.max() > 0  # the run yields, so the plastic state was carried


def test_oracle_hex_adopt_carries_every_points_history():
    # a plastic solid-shell zone that tears: adopt at step 600 and continue to 1200. The
    # eight points' history and the enhanced parameters (the next update's Newton start)
    # travel in hex_points / hex_enhanced; without the enhanced parameters the
    # continuation is a different rounding of the same run, so identity needs them
    case = configs.hex_zone(lx=0.3, ly=0.15, nx=12, ny=6, radius=0.04, speed=10.0, plastic=True,
                            yield_=355e6, hardening=1e9, failure=dict(failure_strain=0.06))
    dt = 1e-6
    a = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    ra = a.run(1200, sample_every=100, threads=4)
    b = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    b.run(600, threads=4)
    ex = b.export()
    assert ex["hex_points"].shape == (case.T, 8, 8) and ex["hex_enhanced"].shape == (case.T, 7)
    assert np.array_equal(ex["plastic"][:, 6], b.hex_eqps) and np.array_equal(ex["plastic"][:, 7], b.hex_damage)
    c = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs()).adopt(ex)
    rc = c.run(600, sample_every=100, threads=4)
    assert np.array_equal(c.u, a.u) and np.array_equal(c.v, a.v)
    assert np.array_equal(c.hex_point, a.hex_point) and np.array_equal(c.tear_step, a.tear_step)
    assert [r["step"] for r in rc] == [r["step"] for r in ra if r["step"] >= 600]
    assert a.hex_eqps.max() > 0.01 and (a.hex_damage >= 1).any()
    # the enhanced parameters matter: a cold start of the Newton iterations changes bits
    d = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs()).adopt(
        {k: v for k, v in ex.items() if k != "hex_enhanced"})
    d.run(600, threads=4)
    assert not np.array_equal(d.u, a.u)

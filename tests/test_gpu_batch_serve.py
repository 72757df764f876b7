"""Many cases per launch (batch) and the resident serving loop, on the GPU.

The claims checked bit-for-bit: a case's result is the same alone and in a mix, for
any launch shape the tuner may pick, graph-replayed or not, and served or batched.
Against the FP64 oracle: one batched and one served case at the parity tolerances.
"""
import time

import numpy as np
import pytest

import shipfem as sf
from shipfem import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if sf.device_info() is None:
        pytest.skip("no usable gfx950 device")


def small_panel(speed=8.0, nx=24, **kw):
    return configs.stiffened_panel(nx=nx, ny_bay=4, nz_web=3, speed=speed, radius=0.05,
                                   geometry=dict(lx=0.36, ly=0.2, n_stiff=1, web_h=0.04), **kw)


def mix():
    return [configs.plate(lx=0.6, ly=0.24, nx=21, ny=9, ramp=1e-4),
            small_panel(),
            configs.plate(lx=0.4, ly=0.2, nx=14, ny=7, ramp=1e-4, plastic=True, yield_=150e6,
                          pressure=2e6),
            small_panel(speed=4.0, nx=30)]


def state(b, i):
    st = b.readback(i)
    return st.u.copy(), st.v.copy(), st.eqps.copy(), b.energy(i)


def same(a, b):
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and
            np.array_equal(a[2], b[2]) and a[3] == b[3])


def test_case_is_bit_identical_alone_and_in_the_mix():
    cases = mix()
    both = sf.Batch(cases)
    both.run(600)
    for i, c in enumerate(cases):
        alone = sf.Batch([c])
        alone.run(600)
        assert same(state(alone, 0), state(both, i)), i
    # and the single-case solver path gives the same bits
    s = sf.Solver(cases[1])
    s.run(600)
    st = s.readback()
    assert np.array_equal(st.u, state(both, 1)[0]) and s.energy() == state(both, 1)[3]


def test_tearing_case_is_bit_identical_alone_and_in_the_mix():
    # a case with ductile failure joins the mix: the launch then runs the damage
    # instantiation for everyone, and no case's bits may change for it
    tear = small_panel(speed=20.0, failure=dict(failure_strain=0.05))
    cases = mix() + [tear]
    both = sf.Batch(cases)
    both.run(600)
    alone = sf.Batch([tear])
    alone.run(600)
    k = len(cases) - 1
    assert same(state(alone, 0), state(both, k))
    ta, tb = alone.tears(0), both.tears(k)
    assert len(ta) > 5 and [(t, s) for t, s, _ in ta] == [(t, s) for t, s, _ in tb]
    assert all(np.array_equal(x[2], y[2]) for x, y in zip(ta, tb))
    assert np.array_equal(both.readback(k).damage, alone.readback(0).damage)
    for i, c in enumerate(mix()):
        ref = sf.Batch([c])
        ref.run(600)
        assert same(state(ref, 0), state(both, i)), i
        assert both.tears(i) == []


def test_server_reports_tears():
    tear = small_panel(speed=20.0, failure=dict(failure_strain=0.05))
    srv, ids, res, kept = serve([small_panel(), tear], 4e-4, keep={1})
    r = res[ids[1]]
    assert r["status"] == sf.OK and r["torn_elements"] > 5
    alone = sf.Batch([tear])
    alone.run(r["steps"])
    st = alone.readback(0)
    assert r["torn_elements"] == int((st.damage >= 1.0).sum())
    assert np.array_equal(kept[ids[1]].u, st.u)
    assert res[ids[0]]["torn_elements"] == 0


@pytest.mark.parametrize("variant", [sf.VARIANT_TILED, sf.VARIANT_SLOTS])
@pytest.mark.parametrize("shape", [(64, 512, 0), (512, 64, 1), (128, 128, 1)])
def test_launch_shape_does_not_change_any_bit(shape, variant):
    cases = mix()
    ref = sf.Batch(cases, 256, 256, 1, variant=variant)
    ref.run(400)
    other = sf.Batch(cases, *shape, variant=variant)
    other.run(400)
    for i in range(len(cases)):
        assert same(state(ref, i), state(other, i)), (shape, i)


@pytest.mark.parametrize("n", [2, 3, 6])
def test_parts_on_their_own_streams_keep_every_bit(n):
    # split=0 runs min(n, 4) parts, one stream each (six cases: parts of one and two cases)
    cases = (mix() + [configs.plate(lx=0.5, ly=0.2, nx=18, ny=8, ramp=1e-4), small_panel(speed=6.0)])[:n]
    one = sf.Batch(cases, split=-1)
    parts = sf.Batch(cases, split=0)
    for steps in (250, 51):  # whole chunks, then a ragged tail of plain steps
        one.run(steps)
        parts.run(steps)
    for i in range(n):
        assert same(state(one, i), state(parts, i)), (n, i)


def test_tuner_picks_a_shape_and_restores_the_state():
    cases = mix()
    b = sf.Batch(cases)
    b.run(200)
    best = b.tune(steps_per_trial=200)
    assert best["elem_block"] in (64, 128, 256, 512) and best["node_block"] in (64, 128, 256, 512)
    b.run(200)
    ref = sf.Batch(cases)
    ref.run(400)
    for i in range(len(cases)):
        assert same(state(ref, i), state(b, i)), i


def test_reset_replays_identically():
    b = sf.Batch(mix())
    b.run(300)
    first = [state(b, i) for i in range(4)]
    b.reset()
    b.run(300)
    assert all(same(first[i], state(b, i)) for i in range(4))


def test_batched_plastic_case_matches_oracle():
    cases = mix()
    b = sf.Batch(cases)
    b.run(800)
    c = cases[1]
    st = b.readback(1)
    dt = st.time / st.step
    o = O.Case(c.X, c.index, dt=dt, **c.oracle_kwargs())
    o.run(800)
    err = np.linalg.norm(st.u - o.u) / np.linalg.norm(o.u)
    assert o.eqps.max() > 1e-3
    assert err <= 1e-5
    assert np.max(np.abs(st.eqps - o.eqps)) <= 1e-4


def serve(cases, run_to, budget=2 << 30, chunk=32, keep=(), gaps=None):
    srv = sf.Server(memory_budget=budget, chunk=chunk)
    ids = []
    for k, c in enumerate(cases):
        if gaps is not None:
            time.sleep(gaps[k])
        ids.append(srv.submit(c, run_to, keep=k in keep))
    assert srv.drain(120.0)
    res = {r["request_id"]: r for r in srv.poll()}
    kept = {ids[k]: srv.fetch(ids[k], cases[k]) for k in keep}
    return srv, ids, res, kept


def test_server_completes_every_request_and_matches_alone():
    rng = np.random.default_rng(3)
    cases = [small_panel(speed=float(rng.uniform(2, 10)), nx=int(rng.integers(16, 40)))
             for _ in range(10)]
    run_to = 3e-4
    srv, ids, res, kept = serve(cases, run_to, keep={4}, gaps=rng.uniform(0, 0.02, 10))
    assert len(res) == 10
    for i in ids:
        r = res[i]
        assert r["status"] == sf.OK, r
        assert r["done_s"] >= r["admit_s"] >= r["submit_s"]
    st = srv.stats()
    assert st["completed"] == 10 and st["failed"] == 0
    # the same case served alone
    _, ids2, res2, kept2 = serve([cases[4]], run_to, keep={0})
    a, b = kept[ids[4]], kept2[ids2[0]]
    assert a.step == b.step and np.array_equal(a.u, b.u) and np.array_equal(a.eqps, b.eqps)
    # and against the FP64 oracle (the served sample)
    c = cases[4]
    dt = a.time / a.step
    o = O.Case(c.X, c.index, dt=dt, **c.oracle_kwargs())
    o.run(a.step)
    assert np.linalg.norm(a.u - o.u) / np.linalg.norm(o.u) <= 1e-5
    assert np.max(np.abs(a.eqps - o.eqps)) <= 1e-4


def test_server_admits_by_footprint_and_refuses_what_never_fits():
    cases = [small_panel(nx=24) for _ in range(6)]
    one = sf.case_footprint(cases[0].N, cases[0].T, True, 512)
    # room for about two cases at once: the rest wait, all complete
    srv, ids, res, _ = serve(cases, 2e-4, budget=int(2.5 * one))
    assert all(res[i]["status"] == sf.OK for i in ids)
    big = configs.plate(lx=2.0, ly=1.0, nx=70, ny=35)
    srv2 = sf.Server(memory_budget=int(2.5 * one), chunk=32)
    with pytest.raises(sf.ShipfemError) as e:
        srv2.submit(big, 1e-4)
    assert e.value.status == sf.BUSY
    assert srv2.stats()["refused"] == 1


def test_server_reports_a_bad_case_and_keeps_serving():
    good = small_panel()
    X = good.X.copy()
    bad = sf.Case(X, good.index.copy(), plastic=True)
    bad.index[0, 0] = len(X) + 5  # out of range node: refused by the preparation
    srv, ids, res, _ = serve([good, bad, good], 1e-4)
    assert res[ids[0]]["status"] == sf.OK and res[ids[2]]["status"] == sf.OK
    assert res[ids[1]]["status"] == sf.INVALID_ARG


def test_server_serves_hex_cases_beside_tets():
    # rounds run one element kind at a time (tets, elastic hexes, plastic hexes in turn);
    # a served hex case ends where the single-case solver takes it, bit for bit
    plastic = dict(plastic=True, yield_=355e6, hardening=1e9)
    hexes = [configs.hex_zone(lx=0.3, ly=0.15, nx=24, ny=12, radius=0.04, speed=5.0),
             configs.hex_zone(lx=0.3, ly=0.15, nx=24, ny=12, radius=0.04, speed=30.0,
                              failure=dict(failure_strain=0.02), **plastic),
             configs.hex_zone(lx=0.2, ly=0.1, nx=16, ny=8, radius=0.03, speed=8.0, **plastic)]
    cases = [small_panel(nx=20), hexes[0], hexes[1], small_panel(speed=4.0, nx=16), hexes[2]]
    run_to = 6e-4
    print("server mix:", [(type(c).__name__, c.T) for c in cases], flush=True)
    srv, ids, res, kept = serve(cases, run_to, keep=set(range(len(cases))))
    assert all(res[i]["status"] == sf.OK for i in ids), [res[i]["status"] for i in ids]
    assert srv.stats()["completed"] == len(cases)
    for k in (1, 2, 4):
        c, a = cases[k], kept[ids[k]]
        s = sf.Solver(c)
        assert a.time / a.step == pytest.approx(s.timestep()["dt"], rel=1e-15)
        s.run(a.step)
        st = s.readback()
        assert np.array_equal(a.u, st.u) and np.array_equal(a.eqps, st.eqps), k
        torn = int(np.sum(st.damage >= 1.0)) if st.damage is not None else 0
        assert res[ids[k]]["torn_elements"] == torn, k
    assert res[ids[2]]["torn_elements"] > 0 and res[ids[2]]["max_eqps"] > 0.0
    assert res[ids[4]]["max_eqps"] > 0.0 and res[ids[1]]["max_eqps"] == 0.0


def test_fetch_returns_position_damage_and_plastic_strain_in_the_callers_order():
    # the library lays a tet case out in its own tet and node orders; a kept result comes
    # back in the caller's: position is X + u node by node, and the per-tet fields equal
    # the single-case solver's readback
    c = small_panel(speed=10.0, failure=dict(failure_strain=0.08))
    run_to = 1200 * sf.Solver(c).timestep()["dt"] * (1 - 1e-9)
    print("fetch case:", c.N, "nodes", c.T, "tets", flush=True)
    srv, ids, res, kept = serve([c], run_to, keep={0})
    a = kept[ids[0]]
    assert res[ids[0]]["status"] == sf.OK
    assert np.array_equal(a.position, (c.X + a.u).astype(np.float32))
    s = sf.Solver(c)
    s.run(a.step)
    st = s.readback()
    assert np.array_equal(a.u, st.u) and np.array_equal(a.position, st.position)
    assert np.array_equal(a.damage, st.damage) and np.array_equal(a.plastic_strain, st.plastic_strain)
    torn = int(np.sum(a.damage >= 1.0))
    assert torn > 10 and res[ids[0]]["torn_elements"] == torn
    # a fetched result is handed over once
    with pytest.raises(sf.ShipfemError) as e:
        srv.fetch(ids[0], c)
    assert e.value.status == sf.NOT_FOUND


def test_shortest_rounds_finish_small_cases_first_and_keep_every_bit():
    # rounds of at most round_work tet-steps, least remaining work first: the small panels
    # all finish before the large one, which still advances through the starvation bound;
    # which round a case joins never changes its arithmetic
    small = [small_panel(speed=float(s), nx=16) for s in (3.0, 5.0, 7.0, 9.0)]
    large = small_panel(speed=6.0, nx=60)
    cases = [large] + small
    run_to = 4e-4
    srv = sf.Server(memory_budget=2 << 30, chunk=32, policy=sf.ROUNDS_SHORTEST,
                    round_work=1.5 * small[0].T * 32, starve_rounds=4)
    ids = [srv.submit(c, run_to, keep=True) for c in cases]
    assert srv.drain(120.0)
    res = {r["request_id"]: r for r in srv.poll()}
    assert all(res[i]["status"] == sf.OK for i in ids)
    print("done (large first):", [round(res[i]["done_s"] - res[ids[0]]["submit_s"], 4) for i in ids], flush=True)
    assert max(res[i]["done_s"] for i in ids[1:]) < res[ids[0]]["done_s"]
    for k in (0, 2):
        a = srv.fetch(ids[k], cases[k])
        s = sf.Solver(cases[k])
        s.run(a.step)
        st = s.readback()
        assert np.array_equal(a.u, st.u) and np.array_equal(a.eqps, st.eqps), k
    with pytest.raises(sf.ShipfemError):
        sf.Server(policy=7)

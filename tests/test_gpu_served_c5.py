"""The served configuration at its workload: Poisson arrivals of stiffened-panel indentation
cases, tets log-uniform in [5e3, 5e5], indenter speed uniform in 2-10 m/s, each run to 2 ms
of simulated time, admitted by device footprint while others are on the GPU.

Checks: every request completes; a seeded sampled case ends bit for bit where the
single-case solver takes it (served in the mix = alone); and the FP64 oracle on a seeded
5% sample (one case of the draw, chosen by the seed, not the smallest) at the parity
tolerances: u rel L2 <= 1e-5, eqps within 1e-4, energy residual gap <= 1e-4 of the work.
"""
import math
import time

import numpy as np
import pytest

import shipfem as sf
from shipfem import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RUN_TO = 2e-3
CHUNK = 64


@pytest.fixture(scope="module", autouse=True)
def device():
    if sf.device_info() is None:
        pytest.skip("no usable gfx950 device")


def draw(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        t = math.exp(rng.uniform(math.log(5e3), math.log(5e5)))
        out.append(configs.served_case(t, float(rng.uniform(2.0, 10.0))))
    return out, rng


def residual(r):
    return (r["w_ext"] + r["w_contact"] + r["w_boundary"] - r["strain"] - r["kinetic"] - r["dissipation"]
            - r["damping_loss"] - r["fracture"])


def test_served_c5_draw_completes_and_matches_alone_and_oracle():
    n = 16
    cases, rng = draw(n, seed=2026)
    print("served draw (tets):", [c.T for c in cases], flush=True)
    rate = 40.0  # requests per second, Poisson
    gaps = rng.exponential(1.0 / rate, n)
    srv = sf.Server(chunk=CHUNK)
    ids = []
    t0 = time.perf_counter()
    for k, c in enumerate(cases):
        time.sleep(gaps[k])
        ids.append(srv.submit(c, RUN_TO, keep=True))
    assert srv.drain(600.0)
    wall = time.perf_counter() - t0
    res = {r["request_id"]: r for r in srv.poll()}
    st = srv.stats()
    lat = [res[i]["done_s"] - res[i]["submit_s"] for i in ids]
    print(f"served {n} in {wall:.2f} s: p50 {np.percentile(lat, 50):.3f} s, p99 {np.percentile(lat, 99):.3f} s, "
          f"{st['element_steps'] / wall:.3e} tet-steps/s", flush=True)
    assert all(res[i]["status"] == sf.OK for i in ids), [res[i]["status"] for i in ids]
    assert st["completed"] == n and st["failed"] == 0
    kept = {i: srv.fetch(i, c) for i, c in zip(ids, cases)}
    for i, c in zip(ids, cases):
        a = kept[i]
        assert a.time >= RUN_TO * (1 - 1e-12) and a.time - a.time / a.step < RUN_TO
        assert np.all(np.isfinite(a.u))
    # the 5% sample for the FP64 oracle: one case chosen by the seed among all but the smallest
    order = np.argsort([c.T for c in cases])
    pick = int(rng.choice(order[1:]))
    # a second seeded case for the served = alone identity
    other = int(rng.choice([k for k in range(n) if k != pick]))
    for k in (other, pick):
        c, a = cases[k], kept[ids[k]]
        s = sf.Solver(c)
        assert a.time / a.step == pytest.approx(s.timestep()["dt"], rel=1e-15)
        s.run(a.step)
        alone = s.readback()
        print(f"case {k} ({c.T} tets, {a.step} steps): served == alone", flush=True)
        assert np.array_equal(a.u, alone.u) and np.array_equal(a.v, alone.v)
        assert np.array_equal(a.eqps, alone.eqps)
        s.close()
    c, a = cases[pick], kept[ids[pick]]
    print(f"oracle sample: case {pick}, {c.T} tets, {a.step} steps", flush=True)
    o = O.Case(c.X, c.index, dt=a.time / a.step, **c.oracle_kwargs())
    rec = o.run(a.step, sample_every=CHUNK, threads=16)
    err_u = np.linalg.norm(a.u - o.u) / np.linalg.norm(o.u)
    err_q = float(np.max(np.abs(a.eqps - o.eqps)))
    last = res[ids[pick]]["last"]
    ora_last = next(r for r in rec if r["step"] == last["step"])
    work = abs(ora_last["w_ext"]) + abs(ora_last["w_contact"])
    gap = abs(residual(last) - residual(ora_last)) / work
    print(f"oracle: u rel L2 {err_u:.2e}, eqps |d| {err_q:.2e} (max {o.eqps.max():.3e}), energy gap "
          f"{gap:.2e} of the work at step {last['step']}", flush=True)
    assert o.eqps.max() > 1e-3
    assert err_u <= 1e-5
    assert err_q <= 1e-4
    assert gap <= 1e-4

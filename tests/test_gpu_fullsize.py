"""The measured workloads at their full sizes.

Size-independent properties at full size: bitwise run-to-run determinism, finiteness, the
energy account's self-consistency, contact engagement, alone-vs-mix identity in a full
batch. Then the full-size configurations against the FP64 oracle itself, run on the host's
cores with a static split (bit-identical in state to one thread): the large plate for 2000
steps, the stiffened panel for 4000 steps, and a seeded sample of the 64-case batch, each at
the parity tolerances (u relative L2 1e-5, eqps 1e-4 absolute, energy residual 1e-4 of the
work).
"""
import os
import sys
import time

import numpy as np
import pytest

import shipfem as sf
from shipfem import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    if sf.device_info() is None:
        pytest.skip("no usable gfx950 device")


def residual(r, r0):
    work = r["w_ext"] + r["w_contact"] + r["w_boundary"]
    return work - (r["strain"] + r["kinetic"] - r0["strain"] - r0["kinetic"]) - \
        r["dissipation"] - r["damping_loss"] - r["fracture"]


def test_large_plate_full_size_is_deterministic_and_conserves_energy():
    case = configs.plate()
    assert case.T == 1209600 and case.N == 254475
    outs = []
    for _ in range(2):
        s = sf.Solver(case)
        s.run(2000)
        st = s.readback()
        assert st.status == 0 and st.step == 2000 and np.isfinite(st.u).all()
        outs.append((st.u, s.energy()))
        s.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]
    rec = outs[0][1]
    w = rec[-1]["w_ext"]
    r = residual(rec[-1], rec[0])
    say(f"plate full: W_ext {w:.4e} J, residual/W {r / w:.2e}, max |u| {np.abs(outs[0][0]).max():.3e} m")
    assert w > 0 and abs(r) < 1e-3 * w
    # the load points down: the loaded face moves down on average
    loaded = case.load[:, 2] != 0
    assert outs[0][0][loaded, 2].mean() < 0


def test_large_plate_quarter_scale_matches_oracle():
    # a quarter of the large plate's cells per in-plane axis, same geometry and loading
    case = configs.plate(lx=2.5, ly=1.0, nx=88, ny=36, ramp=5e-4)
    s = sf.Solver(case)
    o = O.Case(case.X, case.index, dt=s.dt, **case.oracle_kwargs())
    s.run(1000)
    o.run(1000)
    err = np.linalg.norm(s.readback().u - o.u) / np.linalg.norm(o.u)
    say(f"plate quarter scale ({case.T} tets, 1000 steps): u rel L2 {err:.3e}")
    assert err <= 1e-5


def test_stiffened_panel_full_size_engages_contact_and_yields():
    case = configs.stiffened_panel()
    assert case.T == 812196
    s = sf.Solver(case)
    s.run(1200)
    st = s.readback()
    rec = s.energy()
    assert st.status == 0 and np.isfinite(st.u).all()
    last = rec[-1]
    fz = last["contact_force"][2]
    say(f"panel full: t {st.time*1e3:.3f} ms, contact Fz {fz:.3e} N, W_con {last['w_contact']:.3e} J, "
        f"D {last['dissipation']:.3e} J, max eqps {st.eqps.max():.3e}")
    # the sphere pushes +z into the plate; it does positive work and the plate yields
    assert fz > 0 and last["w_contact"] > 0 and st.eqps.max() > 0
    assert abs(residual(last, rec[0])) < 5e-2 * last["w_contact"]


def test_full_batch_alone_equals_mixed():
    cases = configs.batch_cases(64)
    b = sf.Batch(cases, chunk=100)
    b.run(200)
    for i in (0, 37):
        alone = sf.Batch([cases[i]], chunk=100)
        alone.run(200)
        a, m = alone.readback(i * 0), b.readback(i)
        assert np.array_equal(a.u, m.u) and np.array_equal(a.eqps, m.eqps)
        assert alone.energy(0) == b.energy(i)
    assert all(b.readback(i).status == 0 for i in range(64))


THREADS = max(1, min(16, os.cpu_count() or 1))
REL_U, ABS_EQPS, ENERGY = 1e-5, 1e-4, 1e-4


def oracle_run(case, dt, steps):
    o = O.Case(case.X, case.index, dt=dt, **case.oracle_kwargs())
    t0 = time.perf_counter()
    rec = o.run(steps, sample_every=case.sample_every, threads=THREADS)
    say(f"  oracle: {case.T} tets x {steps} steps on {THREADS} threads in "
        f"{time.perf_counter() - t0:.1f} s")
    return o, rec


def energy_gap(g, o):
    assert [r["step"] for r in g] == [r["step"] for r in o]
    work = max(abs(o[-1]["w_ext"]) + abs(o[-1]["w_contact"]) + abs(o[-1]["w_boundary"]), 1e-300)
    return max(abs(residual(a, g[0]) - residual(b, o[0])) for a, b in zip(g[1:], o[1:])) / work


def test_large_plate_full_size_matches_oracle():
    case = configs.plate()
    s = sf.Solver(case)
    s.run(2000)
    st = s.readback()
    o, rec = oracle_run(case, s.dt, 2000)
    err = np.linalg.norm(st.u - o.u) / np.linalg.norm(o.u)
    gap = energy_gap(s.energy(), rec)
    say(f"plate full size, 2000 steps: u rel L2 {err:.3e}, energy gap/W {gap:.2e}")
    assert err <= REL_U and gap <= ENERGY


def test_stiffened_panel_full_size_matches_oracle():
    case = configs.stiffened_panel()
    s = sf.Solver(case)
    s.run(4000)
    st = s.readback()
    o, rec = oracle_run(case, s.dt, 4000)
    err = np.linalg.norm(st.u - o.u) / np.linalg.norm(o.u)
    deq = float(np.max(np.abs(st.eqps - o.eqps)))
    gap = energy_gap(s.energy(), rec)
    say(f"panel full size, 4000 steps: u rel L2 {err:.3e}, max |d eqps| {deq:.2e} "
        f"(max eqps {o.eqps.max():.3e}), energy gap/W {gap:.2e}")
    assert o.eqps.max() > 1e-3
    assert err <= REL_U and deq <= ABS_EQPS and gap <= ENERGY


def test_full_batch_sample_matches_oracle():
    cases = configs.batch_cases(64)
    b = sf.Batch(cases)
    b.run(1000)
    rng = np.random.default_rng(2026)
    for i in sorted(rng.choice(64, size=4, replace=False).tolist()):
        st = b.readback(i)
        o, rec = oracle_run(cases[i], st.time / st.step, 1000)
        err = np.linalg.norm(st.u - o.u) / np.linalg.norm(o.u)
        deq = float(np.max(np.abs(st.eqps - o.eqps)))
        gap = energy_gap(b.energy(i), rec)
        say(f"batch case {i} ({cases[i].T} tets), 1000 steps: u rel L2 {err:.3e}, "
            f"max |d eqps| {deq:.2e}, energy gap/W {gap:.2e}")
        assert err <= REL_U and deq <= ABS_EQPS and gap <= ENERGY

"""Benchmark of the explicit tet step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs LIST] [--csv FILE]
                    [--workload c4|plate|cantilever|panel|batch|serve|sweep|hex|hex_plastic]

Default workload (c4): the batched configuration, 64 independent clamped 20 mm steel plates
(tets log-uniform in [1e4, 3e5], pressure 50-400 kPa, yield 235-460 MPa, J2 with H = 1 GPa,
FP32 + Kahan state), one step = one element pass + one node pass over every case, replayed
from a captured graph. `value` is batched tet-steps/s over exactly K timed steps. Beside it,
measured in the same run on the same 64 cases: the plain path (one case per solver, two plain
launches per step; with this library's kernels, and with the reference's slot scheme) and the
ratios with the HBM bound; the roofline of the dominant kernel on
the bytes its layout must move, timed live with HIP events on the batch's stream; the large
plate (1,209,600 tets) on the plain path; a short served leg (Poisson arrivals of stiffened
panels) with p50/p99 latency; and the reference's FP32 CPU step on this host's cores.
Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "sinking-simulator_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per tet-step and per node-step of the two-pass design, split by
# kernel (element pass / node pass). Per tet: connectivity 16 + Dm^-1 36 + volume 4 +
# force-slot write 48 (element) and force-slot read 48 + CSR entries 16 (node); plus
# J2 state 28 read + 28 write (element) when plastic. Per node: displacement read by
# the element pass 12 (element); CSR offset 4 + X 12 + 1/m 4 + load 12 + state read
# 48 + state write 48 (node). Sum: 168 T + 140 N (+ 56 T plastic) per step.
ELEM_B_T, ELEM_B_N, NODE_B_T, NODE_B_N, PLASTIC_B_T = 104, 12, 64, 128, 56
REFERENCE_SUSTAINED = 470e6  # tet-steps/s, GTX 1070 Ti, 491,520 tets (docs/07-fem-spike-findings.md:89)


def emit(obj):
    print(json.dumps(obj), flush=True)


def workload(name):
    from shipfem import configs

    if name == "plate":
        return configs.plate(), ("plate 10x4x0.02 m, 350x144x4 Kuhn cells x6 tets, nodes jittered "
                                 "+-5%, edges clamped, 200 kPa ramped over 5 ms, elastic")
    if name == "cantilever":
        return configs.cantilever(), ("cantilever 1.0x0.1x0.1 m, 40x4x4 Kuhn cells x6 tets, x=0 "
                                      "clamped, 1 kN tip load ramped over 1 ms, elastic")
    if name == "panel":
        return configs.stiffened_panel(), ("stiffened panel 6x3 m, 12 mm plate + 4 flat bars "
                                           "200x12 mm, J2 355 MPa / H 1 GPa, r=0.5 m sphere at 5 m/s")
    raise SystemExit(f"unknown workload {name}")


def traffic_from_profiles(name):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except Exception:
        return None


def cpu_baseline(case, budget_s=6.0):
    """The reference's FP32 CPU step (fem.cpp:187-215) restated bit-exactly, on the
    same mesh, all host cores (static split) and one core; bounded sample."""
    import numpy as np

    from oracle import oracle as O

    pinned = ((case.flags & 7) == 7).astype(np.uint32)
    sim = O.Ref32(case.X.astype(np.float32), case.index, pinned)
    threads = max(1, min(16, os.cpu_count() or 1))
    dt = 1e-7
    out = {}
    for label, th in (("all", threads), ("one", 1)):
        sim.step(dt, threads=th)  # warm-up
        t0 = time.perf_counter()
        n = 0
        while n < 50 and time.perf_counter() - t0 < budget_s / 2:
            sim.step(dt, threads=th)
            n += 1
        out[label] = (case.T * n / (time.perf_counter() - t0), n)
    return dict(value=out["all"][0], unit="tet-steps/s", cores=threads, kind="port",
                sample=(f"reference stepCpu restated in C (FP32, -O2 -ffp-contract=off, bit-identical "
                        f"to the compiled reference on golden vectors) on the same {case.T}-tet mesh: "
                        f"{out['all'][1]} steps on {threads} threads; one core "
                        f"{out['one'][0]:.3e} tet-steps/s over {out['one'][1]} steps"))


def bench_batch(args, sf, info):
    """64 plate cases (tets log-uniform in [1e4, 3e5], pressure 50-400 kPa, yield
    235-460 MPa, J2, FP32 + Kahan): the plain path (one case at a time, two plain
    launches per step; with this library's tiled kernels, and with the reference's
    slot scheme) against one batched launch per phase (graph-replayed, tuned)."""
    from shipfem import configs

    cases = configs.batch_cases(64)
    tets = sum(c.T for c in cases)
    # plain: the reference's pattern, one case per solver, stepped case after case
    plain_steps = min(args.steps, 200)
    t0 = time.perf_counter()
    for c in cases:
        s = sf.Solver(c)
        s.run(plain_steps, mode=1)
        s.close()
    plain_wall_total = time.perf_counter() - t0
    # device-side plain rate excluding set-up: time the stepping only
    plain_time = 0.0
    for c in cases:
        s = sf.Solver(c)
        t1 = time.perf_counter()
        s.run(plain_steps, mode=1)
        plain_time += time.perf_counter() - t1
        s.close()
    plain_rate = tets * plain_steps / plain_time
    # the same, with the reference's own scheme: per-tet 48-byte force slots gathered
    # through the CSR node map (tet_forces.comp + node_integrate.comp)
    slots_time = 0.0
    for c in cases:
        s = sf.Solver(c).set_launch(sf.VARIANT_SLOTS)
        t1 = time.perf_counter()
        s.run(plain_steps, mode=1)
        slots_time += time.perf_counter() - t1
        s.close()
    plain_slots_rate = tets * plain_steps / slots_time
    b = sf.Batch(cases, chunk=100)
    best = b.tune(steps_per_trial=200)
    b.reset()
    b.run(args.warmup)
    b.reset()
    t0 = time.perf_counter()
    dev_ms = b.run(args.steps)
    wall = time.perf_counter() - t0
    rate = tets * args.steps / wall
    bad = [i for i in range(len(cases)) if b.readback(i).status]
    if bad:
        raise SystemExit(f"non-finite cases {bad}")
    tr = b.traffic()
    step_bytes = tr["elem_read"] + tr["elem_write"] + tr["node_read"] + tr["node_write"]
    emit(dict(metric="tet-element-steps/s", value=rate, unit="tet-steps/s", n_gpus=1,
              steps=args.steps, warmup=args.warmup, ms_per_step=wall * 1e3 / args.steps,
              higher_is_better=True, scaling="weak", vs_baseline=None, dtype="fp32",
              data="synthetic",
              config=dict(workload=f"batched: 64 plates, {tets} tets total, J2, one launch per "
                                   f"phase over all cases, graph chunk 100, tuned {best}",
                          device=info[0], plain_rate=plain_rate, plain_steps=plain_steps,
                          plain_wall_incl_setup_s=plain_wall_total,
                          batched_over_plain=rate / plain_rate,
                          plain_slots_rate=plain_slots_rate,
                          batched_over_plain_slots=rate / plain_slots_rate,
                          device_ms_per_step=dev_ms / args.steps,
                          step_bytes=step_bytes, step_hbm_gbs=step_bytes * args.steps / wall / 1e9,
                          step_hbm_frac=step_bytes * args.steps / wall / 8e12)))


def bench_serve(args, sf, info):
    """Poisson arrivals of stiffened-panel indentation cases (tets log-uniform in
    [5e3, 5e5], indenter 2-10 m/s, run to 2 ms simulated), continuous batching."""
    import math

    import numpy as np

    from shipfem import configs

    rng = np.random.default_rng(11)
    pool = []
    for _ in range(args.pool):
        t = math.exp(rng.uniform(math.log(5e3), math.log(5e5)))
        pool.append(configs.served_case(t, float(rng.uniform(2.0, 10.0))))
    rates = [float(r) for r in args.rates.split(",")]
    rows = []
    total_tet_steps = 0.0
    total_wall = 0.0
    for lam, policy in [(r, p) for p in policies for r in rates]:
        srv = sf.Server(chunk=64, policy=policy)
        n = max(4, int(lam * args.duration))
        arrivals = np.cumsum(rng.exponential(1.0 / lam, n))
        picks = rng.integers(0, len(pool), n)
        t0 = time.perf_counter()
        for k in range(n):
            wait = arrivals[k] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            srv.submit(pool[picks[k]], 2e-3)
        if not srv.drain(600.0):
            raise SystemExit("server did not drain")
        wall = time.perf_counter() - t0
        res = srv.poll(n + 16)
        st = srv.stats()
        srv.close()
        lat = np.array([r["done_s"] - r["submit_s"] for r in res])
        qwait = np.array([r["admit_s"] - r["submit_s"] for r in res])
        ok = sum(1 for r in res if r["status"] == sf.OK)
        rows.append(dict(rate=lam, policy="shortest" if policy == sf.ROUNDS_SHORTEST else "all", requests=n,
                         completed=ok, p50_s=float(np.percentile(lat, 50)),
                         p99_s=float(np.percentile(lat, 99)),
                         p99_queue_s=float(np.percentile(qwait, 99)),
                         tet_steps_per_s=st["element_steps"] / wall,
                         device_busy=st["device_ms"] * 1e-3 / wall, rounds=st["rounds"]))
        total_tet_steps += st["element_steps"]
        total_wall += wall
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    top = max(rows, key=lambda r: r["tet_steps_per_s"])
    emit(dict(metric="tet-element-steps/s", value=top["tet_steps_per_s"], unit="tet-steps/s",
              n_gpus=1, steps=None, warmup=0, ms_per_step=None, higher_is_better=True,
              scaling="weak", vs_baseline=None, dtype="fp32", data="synthetic",
              config=dict(workload="served: Poisson arrivals of stiffened-panel indentation cases, "
                                   "tets log-uniform [5e3,5e5], 2-10 m/s, run to 2 ms, "
                                   f"pool of {len(pool)} meshes",
                          device=info[0], sweep=rows)))


# the reference's fem_spike throughput meshes: makeBoxTetMesh(2.0, 0.5, 0.05, nx, ny, nz),
# root face clamped, gravity on, undamped, 200 warm-up + 2000 timed steps
# (tools/fem_spike/main.cpp:227-262; docs/07-fem-spike-findings.md:85-89)
SPIKE_MESHES = [(20, 4, 2), (60, 8, 4), (120, 16, 4), (200, 24, 8), (320, 32, 8)]
SPIKE_REFERENCE = {960: 116e6, 11520: 666e6, 46080: 587e6, 230400: 439e6, 491520: 470e6}


def bench_sweep(args, sf, info, rows):
    """The fem_spike throughput table from one run: every mesh, repeats interleaved,
    an idle check before each run, one CSV row per run, minimum over repeats."""
    from shipfem import results

    cases = []
    for nx, ny, nz in SPIKE_MESHES:
        m = sf.Mesh.box(2.0, 0.5, 0.05, nx, ny, nz)
        m.flag_plane(0, 0.0, sf.PIN_ALL, tol=1e-6)
        cases.append(sf.Case.from_mesh(m, gravity=9.81, sample_every=100))
    warm, steps = args.warmup, args.steps
    best = {}
    for rep in range(args.repeats):
        for case in cases:
            busy = results.host_busy()
            s = sf.Solver(case)
            s.run(warm)
            ms = s.run(steps)
            s.close()
            rate = case.T * steps / (ms * 1e-3)
            best[case.T] = max(best.get(case.T, 0.0), rate)
            if rows:
                rows.add(steps, workload="fem_spike box 2.0x0.5x0.05", tets=case.T, nodes=case.N,
                         repeat=rep, warmup=warm, device_ms=round(ms, 4),
                         ms_per_step=round(ms / steps, 6), tet_steps_per_s=round(rate),
                         host_busy_before=None if busy is None else round(busy, 3))
            print(json.dumps(dict(repeat=rep, tets=case.T, ms_per_step=ms / steps,
                                  tet_steps_per_s=rate)), file=sys.stderr, flush=True)
    top = cases[-1].T
    table = [dict(tets=t, tet_steps_per_s=r, reference=SPIKE_REFERENCE.get(t)) for t, r in best.items()]
    emit(dict(metric="tet-element-steps/s", value=best[top], unit="tet-steps/s", n_gpus=1,
              steps=steps, warmup=warm, ms_per_step=top / best[top] * 1e3, higher_is_better=True,
              scaling="weak", vs_baseline=best[top] / SPIKE_REFERENCE[top],
              vs_baseline_basis="fem_spike's own number on the same mesh (GTX 1070 Ti, "
                                "docs/07-fem-spike-findings.md:85-89)",
              dtype="fp32", data="synthetic",
              config=dict(workload="fem_spike throughput meshes makeBoxTetMesh(2.0,0.5,0.05,...), "
                                   "root clamped, gravity, undamped; device time by HIP events; "
                                   f"minimum over {args.repeats} interleaved repeats",
                          device=info[0], table=table, csv=args.csv)))


def hex_cpu_baseline(budget_s=8.0, plastic=False):
    """The zone step's element work (elastic: solidshell::internalForce with a 24x24 FP64
    stiffness per element; plastic: elementPlasticUpdate; zone.cpp:928-1012) and node
    update, as the oracle's C hex solve: a bounded sample, a 100 x 100-element patch of
    the same plating (plastic: 40 x 20 elements, punched)."""
    import numpy as np

    from oracle import oracle as O
    from shipfem import configs

    if plastic:
        c = configs.hex_zone(lx=0.4, ly=0.2, nx=40, ny=20, radius=0.05, speed=5.0, plastic=True,
                             yield_=355e6, hardening=1e9, failure=dict(failure_strain=0.1))
    else:
        c = configs.hex_plate(lx=2.5, ly=2.5, nx=100, ny=100, pressure=50e3)
    o = O.Case(c.X, c.index, dt=1e-7, **c.oracle_kwargs())
    threads = max(1, min(16, os.cpu_count() or 1))
    out = {}
    for label, th in (("all", threads), ("one", 1)):
        o.run(1, threads=th)
        t0 = time.perf_counter()
        n = 0
        while n < 400 and time.perf_counter() - t0 < budget_s / 2:
            o.run(5, threads=th)
            n += 5
        out[label] = (c.T * n / (time.perf_counter() - t0), n)
    del np
    what = ("the enhanced-strain Newton and eight J2 returns per element, as the reference's CPU "
            "zone step" if plastic else
            "polar rotation to convergence, full 24x24 stiffness per element as the reference's "
            "CPU zone step")
    return dict(value=out["all"][0], unit="hex-steps/s", cores=threads, kind="port",
                sample=(f"oracle hex solve in C (FP64: {what}) on a {c.T}-element patch of the same "
                        f"plating: {out['all'][1]} steps on {threads} threads; one core "
                        f"{out['one'][0]:.3e} hex-steps/s over {out['one'][1]} steps"))


def bench_hex(args, sf, info, rows, plastic=False):
    """The elastic solid-shell solve on the 1.2M-element hex plate, or (plastic) the plastic
    solve of a 12 x 6 m zone of 12 mm plating punched by a sphere."""
    from shipfem import configs

    if plastic:
        case = configs.hex_zone(lx=12.0, ly=6.0, nx=1200, ny=600, radius=0.5, speed=5.0,
                                plastic=True, yield_=355e6, hardening=1e9,
                                failure=dict(failure_strain=0.1))
        desc = ("solid-shell zone 12x6 m x 12 mm, 1200x600x1 elements, edges clamped, r=0.5 m "
                "sphere at 5 m/s, J2 355 MPa / H 1 GPa with ductile failure 0.1 (FP64 enhanced-strain "
                "Newton per element)")
    else:
        case = configs.hex_plate()
        desc = ("solid-shell hex plate 30x25 m x 12 mm, 1200x1000x1 elements, edges clamped, 50 kPa "
                "ramped over 5 ms, elastic co-rotational (FP64 stiffness block, FP32 + Kahan state)")
    T, N = case.T, case.N
    t0 = time.perf_counter()
    solver = sf.Solver(case)
    prep_s = time.perf_counter() - t0
    ts = solver.timestep()
    solver.run(args.warmup)
    t0 = time.perf_counter()
    dev_ms = solver.run(args.steps)
    wall = time.perf_counter() - t0
    st = solver.readback()
    if st.status:
        raise SystemExit("state went non-finite")
    value = T * args.steps / wall
    elem_ms, node_ms = solver.profile(args.profile_steps)
    # the bytes each pass must move by the case's layout and state (a plastic point that has
    # not yielded is read as its eqps alone)
    tr = solver.traffic()
    elem_bytes = tr["elem_read"] + tr["elem_write"]
    node_bytes = tr["node_read"] + tr["node_write"]
    ek = "k_elem_hex_plastic" if plastic else "k_elem_hex"
    if elem_ms >= node_ms:
        dom, dom_ms, dom_bytes = f"element pass ({ek})", elem_ms, elem_bytes
    else:
        dom, dom_ms, dom_bytes = "node pass (k_node)", node_ms, node_bytes
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles("hex_plastic" if plastic else "hex")
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=(traffic or {}).get(dom.split()[0]), kernel=dom,
                    kernel_ms=round(dom_ms, 5), algorithmic_bytes=dom_bytes,
                    elem_ms=round(elem_ms, 5), node_ms=round(node_ms, 5))
    step_bytes = elem_bytes + node_bytes
    step_gbs = step_bytes * args.steps / wall / 1e9
    cpu = None
    if not args.no_cpu:
        try:
            cpu = hex_cpu_baseline(plastic=plastic)
        except Exception as e:
            cpu = dict(value=None, unit="hex-steps/s", cores=0, kind="port", sample=f"failed: {e}")
    if rows:
        rows.add(args.steps, workload="hex_plastic" if plastic else "hex", elements=T, nodes=N,
                 warmup=args.warmup,
                 wall_ms_per_step=round(wall * 1e3 / args.steps, 6),
                 device_ms_per_step=round(dev_ms / args.steps, 6), hex_steps_per_s=round(value),
                 elem_ms=round(elem_ms, 6), node_ms=round(node_ms, 6))
    emit(dict(metric="hex-element-steps/s", value=value, unit="hex-steps/s", n_gpus=1,
              steps=args.steps, warmup=args.warmup, ms_per_step=wall * 1e3 / args.steps,
              higher_is_better=True, scaling="weak", vs_baseline=None, dtype="fp32+fp64",
              data="synthetic",
              config=dict(workload=desc, elements=T, nodes=N, dt=ts["dt"], cfl=ts["cfl"], device=info[0],
                          host_prep_s=round(prep_s, 2), device_ms_per_step=dev_ms / args.steps,
                          step_hbm_gbs=round(step_gbs, 1), algorithmic_bytes_per_step=step_bytes),
              roofline=roofline, cpu_baseline=cpu))


# ---------------------------------------------------------------- the default (c4) line
def host_cpus():
    """Threads the CPU legs use: the CPUs this process may run on, capped by a cgroup CPU
    quota when one is set. Returns (threads, dict describing the host)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    from shipfem import results

    threads = min(aff, quota) if quota else aff
    return threads, dict(cpu_model=results.cpu_model(), logical_cpus=os.cpu_count(), affinity=aff,
                         cgroup_quota_cpus=quota)


def cpu_baseline_c4(cases, repeats=5):
    """The reference's FP32 stepCpu (fem.cpp:187-215), restated bit-exactly in C, on the
    64 batched meshes: all threads this process may use (static split) and one core, and
    the compiled reference itself on one core for the largest case. Interleaved repeats,
    each after an idle check (/proc/stat busy fraction), reduced by the minimum time.
    The reference has no J2 on tets: the FP64 oracle (with J2, the workload's own
    algorithm) is timed beside it for information."""
    import numpy as np

    from oracle import oracle as O
    from shipfem import results

    threads, host = host_cpus()
    sims = [O.Ref32(c.X.astype(np.float32), c.index, ((c.flags & 7) == 7).astype(np.uint32)) for c in cases]
    tets = sum(c.T for c in cases)
    big = max(cases, key=lambda c: c.T)
    R = O.ref_lib()
    oracle = [O.Case(c.X, c.index, dt=1e-7, **c.oracle_kwargs()) for c in cases[:16]]
    otets = sum(c.T for c in cases[:16])
    best = dict(all=float("inf"), one=float("inf"), ref=float("inf"), oracle=float("inf"))
    busy = []
    dt = 1e-7
    for c in sims:  # warm-up: first touch of every array
        c.step(dt, threads=threads)

    def ref_step(steps):
        pos = big.X.astype(np.float32).copy()
        vel = np.zeros_like(pos)
        fixed = ((big.flags & 7) == 7).astype(np.uint32)
        t0 = time.perf_counter()
        R.ref_step(O._p(big.X.astype(np.float32), O.C.c_float), O._p(pos, O.C.c_float),
                   O._p(vel, O.C.c_float), O._p(fixed, O.C.c_uint32), big.N,
                   O._p(np.ascontiguousarray(big.index, np.uint32), O.C.c_uint32), big.T, 210e9, 0.3,
                   7850.0, dt, 0.0, 1.0, steps)
        return time.perf_counter() - t0

    for _ in range(repeats):
        for leg in ("all", "one", "ref", "oracle"):
            if leg == "ref" and R is None:
                continue
            b = results.host_busy(0.2)
            if b is not None:
                busy.append(b)
            if leg in ("all", "one"):
                th = threads if leg == "all" else 1
                t0 = time.perf_counter()
                for c in sims:
                    c.step(dt, threads=th)
                t = time.perf_counter() - t0
            elif leg == "ref":
                t = ref_step(3) - ref_step(0)  # stepCpu alone: the rest-state set-up cancels
                t /= 3
            else:
                t0 = time.perf_counter()
                for o in oracle:
                    o.run(1, threads=threads)
                t = time.perf_counter() - t0
            best[leg] = min(best[leg], t)
    rate = tets / best["all"]
    extra = dict(one_core=tets / best["one"],
                 reference_one_core=(big.T / best["ref"]) if R is not None else None,
                 oracle_fp64_j2_all_threads=otets / best["oracle"],
                 repeats=repeats, reduce="minimum", host_busy_max=max(busy) if busy else None, **host)
    return dict(value=rate, unit="tet-steps/s", cores=threads, kind="port",
                sample=(f"the reference's FP32 stepCpu restated in C (bit-identical to the compiled "
                        f"reference on golden vectors), elastic (the reference has no J2 on tets), one "
                        f"step of all 64 batched meshes ({tets} tets) per repeat on {threads} threads; "
                        f"{repeats} interleaved repeats, minimum; one-core leg the same step on one "
                        f"thread; compiled reference on one core: 3 steps of the largest case "
                        f"({big.T} tets) net of its set-up; FP64 oracle with J2: one step of the first "
                        f"16 cases ({otets} tets)"), **extra)


def roofline_entry(kernel, ms, need, model_bytes, counters=None):
    """Roofline of one pass: `need` is the bytes its layout must move (read + write)."""
    achieved = need / (ms * 1e-3) / 1e9
    r = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
             frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, kernel=kernel,
             kernel_ms=round(ms, 5), algorithmic_bytes=int(need),
             model_bytes=int(model_bytes), model_frac=round(model_bytes / (ms * 1e-3) / 8e12, 4))
    if counters:
        r["traffic"] = counters.get("bytes")
        r["traffic_source"] = counters.get("source")
        # what the pass is bound by instead, where its counter profile says (busy fractions
        # of the pass's cycles at 2.4 GHz, from the committed counter run)
        if counters.get("lds_busy_cycles_per_cu"):
            cyc = ms * 1e-3 * 2.4e9
            r["lds_busy_frac"] = round(counters["lds_busy_cycles_per_cu"] / cyc, 3)
            r["lds_bank_conflict_frac"] = round(counters["lds_conflict_cycles_per_cu"] / cyc, 3)
    return r


def counters_for(name, kernel_key):
    t = traffic_from_profiles(name)
    if not t or kernel_key not in t:
        return None
    out = dict(bytes=t[kernel_key], source=t.get("source", f"profiles/traffic_{name}.json"))
    if kernel_key == "element" and "element_lds_idx_active_per_cu" in t:
        out.update(lds_busy_cycles_per_cu=t["element_lds_idx_active_per_cu"],
                   lds_conflict_cycles_per_cu=t["element_lds_bank_conflict_per_cu"])
    return out


PRE_STEPS = 500  # untimed steps before the timed ones in the batched configuration


def chunk_for(steps):
    """Graph chunk: even, at most 100, at most the timed step count."""
    return max(2, min(100, steps - steps % 2))


def model_bytes(c):
    """The survey's per-step model (168 T + 140 N, + 56 T with J2), split by pass."""
    e = (ELEM_B_T + (PLASTIC_B_T if c.plastic else 0)) * c.T + ELEM_B_N * c.N
    n = NODE_B_T * c.T + NODE_B_N * c.N
    return e, n


def bench_headline(args, sf, info):
    import numpy as np

    from shipfem import configs

    legs = set(args.legs.split(","))
    K, W = args.steps, args.warmup
    cases = configs.batch_cases(64)
    tets = sum(c.T for c in cases)
    nodes = sum(c.N for c in cases)
    S = chunk_for(K)
    t0 = time.perf_counter()
    # each tet's Dm^-1 and V formed from its tile's rest positions: the step moves 134 MB less
    b = sf.Batch(cases, chunk=S, rest_from_positions=1)
    prep_s = time.perf_counter() - t0
    arena = b.footprint()  # creation copied this many bytes device to device once
    # past the start-up: the timed steps follow `pre` untimed ones, with every case's
    # pressure ramping (a plate's bending period is tens of ms, so over these 0.1-0.2 ms
    # of simulated time the plates load elastically: every tet's J2 trial is tested and
    # none yields; `yielded_tets` reports it)
    pre = S * max(1, round(PRE_STEPS / S))
    b.run(pre)
    best = b.tune(steps_per_trial=max(S, 100))  # restores the state it started from
    b.run(S * -(-W // S))  # warm-up, whole chunks
    b.run(0)  # the batch's stream is idle: the timed region starts from a synchronized device
    t0 = time.perf_counter()
    dev_ms = b.run(K)  # returns after the stream has drained
    wall = time.perf_counter() - t0
    states = [b.readback(i) for i in range(len(cases))]
    bad = [i for i, st in enumerate(states) if st.status]
    if bad:
        raise SystemExit(f"non-finite cases {bad}")
    value = tets * K / wall
    eqps_max = max(float(np.max(st.eqps)) for st in states)
    yielded = sum(int(np.count_nonzero(st.eqps)) for st in states)
    timed_from = int(states[0].step) - K
    del states
    # per-kernel durations, live, of the one-stream kernels (each over all 64 cases; with
    # several streams the parts' kernels share the device and no kernel has a duration of its
    # own): after the timed steps the batch switches to one stream and runs whole chunks;
    # that segment's device time per step (HIP events around the graph replay on the
    # batch's stream) is apportioned between the two passes by their durations when each
    # dispatch carries its own events (those read a few us long each: the event markers'
    # own latency, which the ratio cancels)
    b.run(-int(timed_from + K) % S)  # back onto a chunk boundary
    b.set_config(best["elem_block"], best["node_block"], best["xcd_remap"], best["variant"], split=-1,
                 rest_from_positions=best["rest_from_positions"])
    one = S * max(1, -(-200 // S))
    b.run(S)  # captures the one-stream chunk before its timing starts
    one_ms = b.run(one) / one
    ev_elem, ev_node = b.profile(S * max(1, -(-40 // S)))
    elem_ms = one_ms * ev_elem / (ev_elem + ev_node)
    node_ms = one_ms * ev_node / (ev_elem + ev_node)
    tr = b.traffic()
    elem_need = tr["elem_read"] + tr["elem_write"]
    node_need = tr["node_read"] + tr["node_write"]
    me = sum(model_bytes(c)[0] for c in cases)
    mn = sum(model_bytes(c)[1] for c in cases)
    if elem_ms >= node_ms:
        roofline = roofline_entry("element pass (k_elem_tiled, J2), all 64 cases in one launch", elem_ms,
                                  elem_need, me, counters_for("batch", "element"))
    else:
        roofline = roofline_entry("node pass (k_node), all 64 cases in one launch", node_ms, node_need, mn,
                                  counters_for("batch", "node"))
    roofline.update(elem_ms=round(elem_ms, 5), node_ms=round(node_ms, 5),
                    event_elem_ms=round(ev_elem, 5), event_node_ms=round(ev_node, 5),
                    elem_bytes=int(elem_need), node_bytes=int(node_need),
                    node_frac=round(node_need / (node_ms * 1e-3) / 8e12, 4),
                    one_stream_ms_per_step=round(one_ms, 5), one_stream_steps=one)
    step_need = elem_need + node_need
    # the ceiling of the batched rate: every step's compulsory bytes at the HBM peak
    batched_bound = tets / (step_need / 8e12)
    b.close()
    out = dict(metric="tet-element-steps/s", value=value, unit="tet-steps/s", n_gpus=1, steps=K, warmup=W,
               ms_per_step=wall * 1e3 / K, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="fp32", data="synthetic")
    cfg = dict(workload=("batched: 64 clamped steel plates 20 mm thick, 4 tets through the thickness, tets "
                         "log-uniform in [1e4, 3e5] (seed 7), 50-400 kPa ramped over 0.25 ms, J2 yield "
                         "235-460 MPa, H 1 GPa, FP32 + Kahan state; one element and one node launch per "
                         f"step over all cases, graph chunk {S}"),
               cases=64, tets=tets, nodes=nodes, device=info[0], host_prep_s=round(prep_s, 2),
               batch_arena_bytes=arena,
               tuned=best, device_ms_per_step=dev_ms / K, timed_steps_from=timed_from,
               eqps_max=eqps_max, yielded_tets=yielded,
               streams=min(4, len(cases)) if best.get("split", -1) == 0 else 1,
               step_bytes=int(step_need), step_hbm_frac=round(step_need * K / (dev_ms * 1e-3) / 8e12, 4),
               batched_bound_at_hbm_peak=batched_bound)
    if "plain" in legs:
        # the plain path on the same 64 cases: one solver per case, two plain launches per
        # step (the reference's dispatch pattern), the cases stepped one after another
        per = []
        plain_time = graph_time = slots_time = 0.0
        for c in cases:
            s = sf.Solver(c)
            s.run(timed_from - W)  # from the state the batched steps were timed from
            s.run(W, mode=1)
            t1 = time.perf_counter()
            s.run(K, mode=1)
            dt_ = time.perf_counter() - t1
            plain_time += dt_
            per.append((c.T, c.T * K / dt_))
            # the same case with its steps replayed from captured graphs (one submission of
            # many steps, as the reference records a run into one command buffer)
            s.run(S * -(-max(W, 1) // S), mode=0)
            t1 = time.perf_counter()
            s.run(K, mode=0)
            graph_time += time.perf_counter() - t1
            # and with the reference's own kernel scheme: one 48-byte force slot per tet,
            # gathered through the CSR node map (tet_forces.comp + node_integrate.comp)
            s.set_launch(sf.VARIANT_SLOTS)
            s.run(W, mode=1)
            t1 = time.perf_counter()
            s.run(K, mode=1)
            slots_time += time.perf_counter() - t1
            s.close()
        plain_rate = tets * K / plain_time
        plain_graph_rate = tets * K / graph_time
        plain_slots_rate = tets * K / slots_time
        bins = [(1e4, 3e4), (3e4, 1e5), (1e5, 3.01e5)]
        by_bin = []
        for lo, hi in bins:
            xs = [(t, r) for t, r in per if lo <= t < hi]
            if xs:
                T_ = sum(t for t, _ in xs)
                by_bin.append(dict(tets=[int(lo), int(hi)], cases=len(xs),
                                   rate=T_ / sum(t / r for t, r in xs),
                                   us_per_step=1e6 * sum(t / r for t, r in xs) / len(xs)))
        cfg.update(plain_rate=plain_rate, batched_over_plain=value / plain_rate,
                   batched_over_plain_bound=batched_bound / plain_rate, plain_by_size=by_bin,
                   plain_graph_rate=plain_graph_rate, batched_over_plain_graph=value / plain_graph_rate,
                   plain_slots_rate=plain_slots_rate, batched_over_plain_slots=value / plain_slots_rate)
    if "plate" in legs:
        case, desc = workload("plate")
        s = sf.Solver(case)
        s.run(20)
        t1 = time.perf_counter()
        n = 200
        pdev = s.run(n)
        pw = time.perf_counter() - t1
        qe, qn = s.profile(40)
        pe, pn = pdev / n * qe / (qe + qn), pdev / n * qn / (qe + qn)
        ptr = s.traffic()
        pe_need = ptr["elem_read"] + ptr["elem_write"]
        cfg["plate"] = dict(workload=desc, tets=case.T, nodes=case.N, steps=n, rate=case.T * n / pw,
                            ms_per_step=pw * 1e3 / n,
                            roofline=roofline_entry("element pass (k_elem_tiled)", pe, pe_need,
                                                    model_bytes(case)[0], counters_for("plate", "element")),
                            node_ms=round(pn, 5), node_bytes=ptr["node_read"] + ptr["node_write"],
                            node_frac=round((ptr["node_read"] + ptr["node_write"]) / (pn * 1e-3) / 8e12, 4),
                            vs_reference_sustained=case.T * n / pw / REFERENCE_SUSTAINED)
        s.close()
    if "serve" in legs:
        cfg["served"] = serve_leg(sf, rates=(20.0, 100.0, 150.0), duration=args.duration, pool_n=12)
    cpu = None
    if "cpu" in legs and not args.no_cpu:
        try:
            cpu = cpu_baseline_c4(cases)
        except Exception as e:  # reported, never fatal for the GPU number
            cpu = dict(value=None, unit="tet-steps/s", cores=0, kind="port", sample=f"failed: {e}")
    out.update(config=cfg, roofline=roofline, cpu_baseline=cpu)
    emit(out)


def serve_leg(sf, rates, duration, pool_n, seed=11, policies=None):
    """Poisson arrivals of stiffened-panel indentation cases (tets log-uniform in
    [5e3, 5e5], indenter 2-10 m/s, run to 2 ms): p50/p99 latency and throughput."""
    import math

    import numpy as np

    from shipfem import configs

    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(pool_n):
        t = math.exp(rng.uniform(math.log(5e3), math.log(5e5)))
        pool.append(configs.served_case(t, float(rng.uniform(2.0, 10.0))))
    policies = policies or [sf.ROUNDS_ALL, sf.ROUNDS_SHORTEST]
    rows = []
    # each rate's arrivals and picks are drawn once: both policies serve the same stream
    streams = {}
    for lam in rates:
        n = max(8, int(lam * duration))
        streams[lam] = (n, np.cumsum(rng.exponential(1.0 / lam, n)), rng.integers(0, len(pool), n))
    for lam, policy in [(r, p) for p in policies for r in rates]:
        srv = sf.Server(chunk=64, policy=policy)
        n, arrivals, picks = streams[lam]
        t0 = time.perf_counter()
        for k in range(n):
            wait = arrivals[k] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            srv.submit(pool[picks[k]], 2e-3)
        if not srv.drain(600.0):
            raise SystemExit("server did not drain")
        wall = time.perf_counter() - t0
        res = srv.poll(n + 16)
        st = srv.stats()
        srv.close()
        lat = np.array([r["done_s"] - r["submit_s"] for r in res])
        ok = sum(1 for r in res if r["status"] == sf.OK)
        rows.append(dict(rate=lam, policy="shortest" if policy == sf.ROUNDS_SHORTEST else "all", requests=n,
                         completed=ok, p50_s=float(np.percentile(lat, 50)),
                         p99_s=float(np.percentile(lat, 99)),
                         tet_steps_per_s=st["element_steps"] / wall,
                         device_busy=st["device_ms"] * 1e-3 / wall))
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    return dict(workload="Poisson arrivals of 6x3 m stiffened panels (12 mm plate, 4 flat bars), tets "
                         f"log-uniform [5e3, 5e5], r = 0.5 m sphere at 2-10 m/s, run to 2 ms; pool of "
                         f"{pool_n} meshes", rates=rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--legs", default="batch,plain,plate,serve,cpu",
                    help="c4 only: which legs beside the batched one to run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=40)
    ap.add_argument("--variant", default="tiled", choices=["tiled", "slots"])
    ap.add_argument("--rates", default="10,50,100,150,200")
    ap.add_argument("--duration", type=float, default=3.0)
    ap.add_argument("--pool", type=int, default=24)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--csv", default=None,
                    help="write one row per timed run, stamped with run id, commit, source digest, "
                         "time and step count")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after timing, create a one-case batch: its creation is one device-to-"
                         "device copy of a known size, the reference for rocprofv3 byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} refused); cases shard across GPUs with no communication",
              file=sys.stderr)
        sys.exit(2)

    import shipfem as sf

    info = sf.device_info()
    if info is None:
        print("no usable gfx950 device", file=sys.stderr)
        sys.exit(1)
    rows = None
    if args.csv:
        from shipfem import results

        rows = results.Rows(args.csv, results.run_identity(info[0]))
    if args.workload == "c4":
        return bench_headline(args, sf, info)
    if args.workload == "sweep":
        return bench_sweep(args, sf, info, rows)
    if args.workload == "batch":
        return bench_batch(args, sf, info)
    if args.workload == "serve":
        return bench_serve(args, sf, info)
    if args.workload in ("hex", "hex_plastic"):
        return bench_hex(args, sf, info, rows, plastic=args.workload == "hex_plastic")
    case, desc = workload(args.workload)
    T, N = case.T, case.N
    variant = sf.VARIANT_TILED if args.variant == "tiled" else sf.VARIANT_SLOTS
    solver = sf.Solver(case).set_launch(variant)
    ts = solver.timestep()
    solver.run(args.warmup)
    t0 = time.perf_counter()
    dev_ms = solver.run(args.steps)
    wall = time.perf_counter() - t0
    st = solver.readback()
    if st.status:
        raise SystemExit("state went non-finite")
    value = T * args.steps / wall
    ms_per_step = wall * 1e3 / args.steps

    elem_ms, node_ms = solver.profile(args.profile_steps)
    # the bytes each pass must move by the case's layout and state; the survey's per-step
    # model beside it, never as it
    tr = solver.traffic()
    elem_bytes = tr["elem_read"] + tr["elem_write"]
    node_bytes = tr["node_read"] + tr["node_write"]
    me, mn = model_bytes(case)
    ek = "k_elem_tiled" if args.variant == "tiled" else "k_elem"
    # committed counter profiles exist for the tiled variant of some workloads only
    prof = args.workload if args.variant == "tiled" else None
    if elem_ms >= node_ms:
        roofline = roofline_entry(f"element pass ({ek})", elem_ms, elem_bytes, me,
                                  counters_for(prof, "element") if prof else None)
    else:
        roofline = roofline_entry("node pass (k_node)", node_ms, node_bytes, mn,
                                  counters_for(prof, "node") if prof else None)
    roofline.update(elem_ms=round(elem_ms, 5), node_ms=round(node_ms, 5))
    step_bytes = elem_bytes + node_bytes
    step_gbs = step_bytes * args.steps / wall / 1e9
    copy_probe = None
    if args.copy_probe:
        foot = solver.footprint()
        solver.close()
        sf.Batch([case]).close()  # one device-to-device copy of the padded arena
        copy_probe = (foot + 4095) // 4096 * 4096

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(case)
        except Exception as e:  # reported, never fatal for the GPU number
            cpu = dict(value=None, unit="tet-steps/s", cores=0, kind="port", sample=f"failed: {e}")

    if rows:
        rows.add(args.steps, workload=args.workload, tets=T, nodes=N, variant=args.variant,
                 warmup=args.warmup, wall_ms_per_step=round(ms_per_step, 6),
                 device_ms_per_step=round(dev_ms / args.steps, 6), tet_steps_per_s=round(value),
                 elem_ms=round(elem_ms, 6), node_ms=round(node_ms, 6))
    emit(dict(metric="tet-element-steps/s", value=value, unit="tet-steps/s", n_gpus=1,
              steps=args.steps, warmup=args.warmup, ms_per_step=ms_per_step,
              higher_is_better=True, scaling="weak",
              vs_baseline=value / REFERENCE_SUSTAINED,
              vs_baseline_basis="reference sustained 470e6 tet-steps/s (GTX 1070 Ti, Vulkan, FP32, "
                                "491,520 tets; docs/07-fem-spike-findings.md:89)",
              dtype="fp32", data="synthetic",
              config=dict(workload=desc, tets=T, nodes=N, variant=args.variant, dt=ts["dt"],
                          cfl=ts["cfl"],
                          reference_edge_dt=ts["edge_dt"], device=info[0],
                          device_ms_per_step=dev_ms / args.steps, step_hbm_gbs=round(step_gbs, 1),
                          step_roofline_frac=round(step_gbs / HBM_PEAK_GBS, 4),
                          algorithmic_bytes_per_step=step_bytes, copy_probe_bytes=copy_probe),
              roofline=roofline, cpu_baseline=cpu))


if __name__ == "__main__":
    main()

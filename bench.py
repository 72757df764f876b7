"""Benchmark of the explicit tet step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--csv FILE]
                    [--workload plate|cantilever|panel|batch|serve|sweep|hex|hex_plastic]

Default workload: the large plain-path plate (10 x 4 x 0.02 m, 350x144x4 Kuhn cells,
1,209,600 tets, nodes jittered +-5%, edges clamped, 200 kPa ramped over 5 ms,
FP32 + Kahan state), one step = one element pass + one node pass over the whole
mesh. Prints ONE JSON line with tet-steps/s, the HBM roofline of the dominant kernel
(HIP events on the solver's stream) and the reference CPU path (the FP32 stepCpu
restatement, bit-identical to the compiled reference) timed on this host.
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
# Solid-shell hex pass, per element: connectivity 32 + the packed 21x21 stiffness block
# 231 x 8 + centre gradients and rest offsets 42 x 4 + force-slot write 96; per node the
# displacement and its compensation 28. Node pass per element: 8 slot reads of 12 + 8
# gather entries of 4; per node as for tets.
HEX_ELEM_B_T, HEX_ELEM_B_N, HEX_NODE_B_T = 32 + 231 * 8 + 42 * 4 + 96, 28, 8 * 16
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
    step_bytes = sum((168 + 56) * c.T + 140 * c.N for c in cases)
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
                          step_hbm_gbs_model=step_bytes * args.steps / wall / 1e9,
                          step_roofline_frac_model=step_bytes * args.steps / wall / 8e12)))


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
    for lam in rates:
        srv = sf.Server(chunk=64)
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
        rows.append(dict(rate=lam, requests=n, completed=ok, p50_s=float(np.percentile(lat, 50)),
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


# Plastic solid-shell pass, per element: connectivity 32 + rest geometry in FP64 42 x 8 (the
# strain operator is formed from it in the pass) + point history and enhanced parameters
# read 71 x 8 + enhanced parameters written 7 x 8 + force-slot write 96; per node 28. A point
# writes its history back only when it yields (elsewhere it is unchanged): those writes are
# not counted.
HEX_PL_ELEM_B_T = 32 + 42 * 8 + 71 * 8 + 7 * 8 + 96


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
    elem_bytes = (HEX_PL_ELEM_B_T if plastic else HEX_ELEM_B_T) * T + HEX_ELEM_B_N * N
    node_bytes = HEX_NODE_B_T * T + NODE_B_N * N
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="plate")
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
    plastic = bool(case.plastic)
    elem_bytes = (ELEM_B_T + (PLASTIC_B_T if plastic else 0)) * T + ELEM_B_N * N
    node_bytes = NODE_B_T * T + NODE_B_N * N
    ek = "k_elem_tiled" if args.variant == "tiled" else "k_elem"
    if elem_ms >= node_ms:
        dom, dom_ms, dom_bytes = f"element pass ({ek})", elem_ms, elem_bytes
    else:
        dom, dom_ms, dom_bytes = "node pass (k_node)", node_ms, node_bytes
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(args.workload)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=(traffic or {}).get(dom.split()[0]), kernel=dom,
                    kernel_ms=round(dom_ms, 5), algorithmic_bytes=dom_bytes,
                    elem_ms=round(elem_ms, 5), node_ms=round(node_ms, 5))
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

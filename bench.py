"""Benchmark: FP64 AMR multigrid Poisson solve on one MI355X.

Workload (one "step" = one full solve to a 1e-10 relative max-norm residual from a
zero initial guess, inputs resident in HBM): 3-level AMR, base 256^3, refinement
ratio 2, centred refined regions covering 25 % then 6 % of the parent domain
(320^3 and 400^3 fine regions), 64^3 maximum boxes, homogeneous Dirichlet BCs,
alpha = 0, beta = 1, Gaussian-bump rhs (seed 1002); V-cycles with 2 pre / 2 post
GSRB sweeps and a BiCGStab bottom solve.

Prints one JSON line.  value = Mcell-relax/s (cells x red-black GSRB iterations
over every level relaxed, per second of solve wall time).  roofline = the GSRB
kernel's algorithmic bytes (24 B per cell-relax in FP64) over its HIP-event time
inside the timed region, against 8 TB/s HBM.  cpu_baseline = the oracle (a C
restatement of the reference's FP64 path, OpenMP over boxes) built for the host CPU
and timed on one full solve of the same workload (setup excluded).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "chombo_3.2_amd"))

METRIC = "Mcell-relax/s & %HBM roofline; solves/s + p99 latency at 1e-10, MP vs FP64"
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "gsrb_amr3_256_traffic.json")


def read_traffic():
    """Measured HBM bytes / algorithmic bytes of the GSRB kernel, from the committed
    rocprofv3 PMC summary (None when absent)."""
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as f:
            return json.load(f).get("traffic_over_algorithmic")
    return None


def summarize_trace(trace_csv, out_path, names, command):
    """Markdown summary of a rocprofv3 kernel trace of a run made with --trace-marks:
    the dispatches between each pair of marker copies (k_copy_probe) are one timed
    region; per region, kernel time by kernel (templates merged) and by instantiation."""
    import csv
    import re
    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    # segments between runs of marker dispatches; the timed regions are every second one
    segs, cur, in_mark = [], [], False
    for r in rows:
        mark = "k_copy_probe" in r[2]
        if mark and not in_mark:
            segs.append(cur)
            cur = []
        if not mark:
            cur.append(r)
        in_mark = mark
    segs.append(cur)
    regions = segs[1::2]

    def base(name):
        n = name.replace("void ", "").replace("(anonymous namespace)::", "")
        n = n.split("(")[0]
        n = re.sub(r"^amrmg::(\(anonymous namespace\)::)?(ops::|vc::)?", "", n)
        return n

    out = [f"# {names[0] if names else 'Kernel time'}", "", "Command (repository root on the GPU host; the output "
           "directory is scratch), then this summary:", "", "```", command,
           f"python3 bench.py --summarize-trace {trace_csv} {out_path}", "```", ""]
    for q, reg in enumerate(regions):
        if not reg:
            continue
        title = names[q + 1] if q + 1 < len(names) else f"timed region {q + 1}"
        busy = sum(e - s for s, e, _ in reg) * 1e-6
        span = (reg[-1][1] - reg[0][0]) * 1e-6
        cls, inst = {}, {}
        for s_, e_, n in reg:
            k = re.sub(r"<.*", "", base(n))
            cls.setdefault(k, [0, 0.0])
            cls[k][0] += 1
            cls[k][1] += (e_ - s_) * 1e-6
            i = base(n)
            inst.setdefault(i, [0, 0.0])
            inst[i][0] += 1
            inst[i][1] += (e_ - s_) * 1e-6
        # the roofline kernel: every one-launch relaxation (k_gsrb_fused, k_gsrb_zmarch, k_gsrb_zmarch4)
        gs = [(e_ - s_) * 1e-6 for s_, e_, n in reg if re.match(r"k_gsrb_(fused|zmarch)", base(n))]
        gline = (f"GSRB relaxation class (the roofline kernel): {len(gs)} launches, {sum(gs):.2f} ms, "
                 f"average {1000 * sum(gs) / len(gs):.1f} us per launch.") if gs else ""
        out += [f"## {title}", "", f"{len(reg)} dispatches, span {span:.2f} ms under the profiler, kernels busy "
                f"{busy:.2f} ms.", ""] + ([gline, ""] if gline else []) + [
                "| kernel | calls | ms | % of busy |", "|---|---:|---:|---:|"]
        for k, (c, ms) in sorted(cls.items(), key=lambda kv: -kv[1][1])[:16]:
            out.append(f"| `{k}` | {c} | {ms:.2f} | {100 * ms / busy:.1f} |")
        out += ["", "| instantiation | calls | ms | avg us |", "|---|---:|---:|---:|"]
        for k, (c, ms) in sorted(inst.items(), key=lambda kv: -kv[1][1])[:16]:
            out.append(f"| `{k}` | {c} | {ms:.2f} | {1000 * ms / c:.1f} |")
        out.append("")
    with open(out_path, "w") as f:
        f.write("\n".join(out))
    log(f"wrote {out_path} ({len(regions)} regions)")


def summarize_pmc(fetch_csv, write_csv, fetch_json, copy_bytes, out_path):
    """Writes the GSRB traffic summary from two rocprofv3 --pmc runs of this script
    (FETCH_SIZE, then WRITE_SIZE; both with --steps 1 --warmup 0 --no-cpu).

    Each counter is scaled by what it reads for the plain copy kernel of copy_bytes
    (reads copy_bytes, writes copy_bytes).  The ratio of scaled GSRB bytes to the
    GSRB algorithmic bytes of the same run (taken from that run's JSON line) gives
    the HBM traffic per launch of the benchmark's launch mix."""
    import csv

    with open(fetch_json) as f:
        run = json.loads(f.read().strip().splitlines()[-1])["roofline"]
    n = run["launches"]

    def per_kernel(path):
        # dispatch order; the timed solves' relaxations are the last n one-launch
        # relaxations before the copy calibration (the autotuner's come earlier)
        rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
        first_cp = min(q for q, r in enumerate(rows) if "k_copy_probe" in r["Kernel_Name"])
        gs = [float(r["Counter_Value"]) * 1024.0 for r in rows[:first_cp]
              if "k_gsrb_fused" in r["Kernel_Name"] or "k_gsrb_zmarch" in r["Kernel_Name"]]
        cp = [float(r["Counter_Value"]) * 1024.0 for r in rows if "k_copy_probe" in r["Kernel_Name"]]
        if len(gs) < n:
            raise SystemExit(f"dispatch count mismatch: {len(gs)} relaxations before the calibration, {n} timed")
        return gs[len(gs) - n:], cp
    gf, cf = per_kernel(fetch_csv)
    gw, cw = per_kernel(write_csv)
    fscale = copy_bytes / (sum(cf) / len(cf))
    wscale = copy_bytes / (sum(cw) / len(cw))
    alg_total = run["algorithmic_bytes_per_launch"] * n
    meas_total = sum(gf) * fscale + sum(gw) * wscale
    ratio = meas_total / alg_total
    d = {
        "kernel": "GSRB relaxation class (k_gsrb_fused, k_gsrb_zmarch)",
        "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4),
        "copy_bytes": copy_bytes,
        "gsrb_dispatches": len(gf),
        "fetch_bytes_total": round(sum(gf) * fscale), "write_bytes_total": round(sum(gw) * wscale),
        "algorithmic_bytes_total": round(alg_total),
        "traffic_over_algorithmic": round(ratio, 4),
    }
    with open(out_path, "w") as f:
        json.dump(d, f, indent=1)
    print(json.dumps(d))


def _oracle_native():
    """The oracle built for this host's CPU (gcc -O3 -march=native -ffp-contract=off,
    the parity build's arithmetic); None if the compiler is unavailable (then the
    parity build, -O2, is timed)."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as ORC
    try:
        path = ORC.build_native(os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle_native.so"))
        ORC.use_library(path)
        return "gcc -O3 -march=native -ffp-contract=off -fopenmp, built on this host"
    except Exception as e:  # noqa: BLE001
        log(f"native oracle build failed ({e}); timing the parity build")
        if not os.path.exists(ORC._LIB_PATH):
            ORC.build()
        return "gcc -O2 -ffp-contract=off -fopenmp (parity build)"


def _cpu_solve(hier, rhs, sp, rp, threads):
    """One full solve on the CPU oracle: (seconds, cell-relaxations, info); setup excluded."""
    import oracle as ORC
    ORC.set_threads(threads)
    o = ORC.Oracle(hier)
    ORC.Oracle.reset_counters()
    t0 = time.perf_counter()
    _, hist, info = o.solve(rhs, sp, rp)
    dt = time.perf_counter() - t0
    cr = ORC.Oracle.cell_relax_count()
    o.close()
    return dt, cr, info


def cpu_baseline(args, hier, rhs, sp, rp, c1):
    """The oracle (a C restatement of the reference's FP64 path, OpenMP over boxes like
    the reference) on the host cores: one full solve of the benchmark's workload to
    1e-10 from zero, setup excluded.  With --cpu-full, BASELINE.md's protocol: one
    warm-up, then the median of 5 solves on all cores, one solve on 1 core, and the
    single-level 64^3 workload the same way."""
    build = _oracle_native()
    import oracle as ORC
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or ORC.get_threads()
    dt, cr, info = _cpu_solve(hier, rhs, sp, rp, threads)
    out = dict(value=cr / dt / 1e6, unit="Mcell-relax/s", cores=threads, kind="port",
               sample=f"one full solve of the same {hier.nlev}-level {hier.domain[1][0] + 1}^3 workload to 1e-10 "
                      f"({info['iterations']} V-cycles, "
                      f"{cr / 1e6:.1f} M cell-relaxations in {dt:.2f} s; oracle setup excluded); {build}",
               solves_per_s=round(1.0 / dt, 5), build=build)
    if args.cpu_full:
        import numpy as np
        times = []
        _cpu_solve(hier, rhs, sp, rp, threads)  # warm-up
        log("cpu protocol: warm-up solve done")
        for q in range(5):
            times.append(_cpu_solve(hier, rhs, sp, rp, threads)[0])
            log(f"cpu protocol: all-core solve {q} {times[-1]:.2f} s")
        # one core: the first 4 V-cycles of the same solve (a full solve takes minutes);
        # the rate is per cell-relaxation, the solve rate follows from the full count
        sp4 = sp.copy()
        sp4[4] = 4
        t1, cr1, _ = _cpu_solve(hier, rhs, sp4, rp, 1)
        log(f"cpu protocol: 1-core 4 V-cycles {t1:.2f} s")
        ch, crhs, csp, crp = c1
        c1_all = [_cpu_solve(ch, crhs, csp, crp, threads)[0] for _ in range(6)][1:]
        c1_one = _cpu_solve(ch, crhs, csp, crp, 1)[0]
        log("cpu protocol: 64^3 single level done")
        rate1 = cr1 / t1
        out["protocol"] = {
            "c2_solves_per_s_all_cores_median5": round(1.0 / float(np.median(times)), 5),
            "c2_mcell_relax_per_s_all_cores": round(cr / float(np.median(times)) / 1e6, 2),
            "c2_solves_per_s_1core": round(rate1 / cr, 6),
            "c2_mcell_relax_per_s_1core": round(rate1 / 1e6, 2),
            "c2_1core_sample": "first 4 V-cycles of the solve; solves/s = rate / cell-relaxations of the full solve",
            "c1_solves_per_s_all_cores_median5": round(1.0 / float(np.median(c1_all)), 3),
            "c1_solves_per_s_1core": round(1.0 / c1_one, 3),
            "threads": threads,
            "cpu": _cpu_model(),
        }
    ORC.set_threads(threads)
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def c1_leg(args, ctx):
    """The reference's CPU-runnable case on the GPU: single-level 64^3 Poisson,
    8 boxes of 32^3, Dirichlet 0, rhs = sin 2pi x sin 2pi y sin 2pi z + N(0, 1e-3)
    (seed 1001), V(2,2) cycles with a BiCGStab bottom on 4^3, tol 1e-10.  At 0.26 M
    cells the solve is launch- and latency-bound on the GPU."""
    import amrmg
    import hierarchy as H
    h = H.single_level(64, 32, 8)
    rhs = H.sinusoid_rhs(h, 1001)
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=100)
    s = amrmg.AMRMultiGrid(ctx, h)
    s.setSolverParameters(sp, rp)
    s.set_field(s.RHS, rhs)
    for _ in range(3):
        s.solve()
    n = max(20, args.steps)
    ctx.profile(["gsrb"])
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(n):
        hist, info = s.solve()
    wall = time.perf_counter() - t0
    prof = ctx.profile_read("gsrb")
    ctx.profile([])
    s.close()
    out = {"workload": "single-level 64^3 Poisson, 8 boxes of 32^3, Dirichlet 0, sin*sin*sin + N(0,1e-3) rhs "
                       "seed 1001, V(2,2) + BiCGStab bottom, tol 1e-10, FP64",
           "ms_per_solve": round(wall / n * 1e3, 3), "solves_per_s": round(n / wall, 2),
           "iterations": info["iterations"], "final_rel_residual": float(hist[-1] / hist[0]),
           "mcell_relax_per_s": round(prof["units"] / wall / 1e6, 2)}
    log(f"c1: {out}")
    return out, (h, rhs, sp, rp)


def mixed_leg(args, ctx, hier, rhs, sp, rp, solver64, fp64_s_per_step):
    """The same workload solved with FP32 inner V-cycles under the FP64 outer
    iteration (precision 1): solves/s against the FP64 path and the max-norm
    difference of the composite solution (cells not covered by a finer level)
    from the FP64 solution, which the parity tests hold bitwise equal to the
    oracle's."""
    import numpy as np
    import amrmg
    phi64 = solver64.get_field(solver64.PHI)
    ms = amrmg.AMRMultiGrid(ctx, hier, precision=1)
    ms.setSolverParameters(sp, rp)
    ms.set_field(ms.RHS, rhs)
    if not args.no_tune:
        ms.tune_relax()
    for _ in range(args.warmup):
        ms.solve()
    ctx.profile(["gsrb"])
    ctx.profile_reset()
    if args.trace_marks:
        ctx.copy_probe(4096, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hist, info = ms.solve()
    wall = time.perf_counter() - t0
    if args.trace_marks:
        ctx.copy_probe(4096, 1)
    prof = ctx.profile_read("gsrb")
    ctx.profile([])
    phi = ms.get_field(ms.PHI)
    offs = np.cumsum([0] + [hier.cells(l) for l in range(hier.nlev)])
    mask = np.ones(offs[-1], dtype=bool)
    for lev in range(hier.nlev - 1):  # covered cells: outside the composite residual
        cov = [((b[0][0] // 2, b[0][1] // 2, b[0][2] // 2), (b[1][0] // 2, b[1][1] // 2, b[1][2] // 2))
               for b in hier.levels[lev + 1]]
        o = offs[lev]
        for (lo, hi) in hier.levels[lev]:
            n = [hi[d] - lo[d] + 1 for d in range(3)]
            cnt = n[0] * n[1] * n[2]
            I = np.meshgrid(*[np.arange(lo[d], hi[d] + 1) for d in range(3)], indexing="ij")
            c = np.zeros(n, dtype=bool)
            for (cl, ch) in cov:
                m = np.ones(n, dtype=bool)
                for d in range(3):
                    m &= (I[d] >= cl[d]) & (I[d] <= ch[d])
                c |= m
            mask[o:o + cnt] = ~c.ravel(order="F")
            o += cnt
    err = float(np.abs(phi - phi64)[mask].max() / np.abs(phi64)[mask].max())
    cell_relax = prof["units"]
    out = {
        "dtype": "f32 inner V-cycles / f64 outer residual, update, bottom",
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "solves_per_s": round(args.steps / wall, 4),
        "speedup_vs_fp64": round(fp64_s_per_step / (wall / args.steps), 3),
        "mcell_relax_per_s": round(cell_relax / wall / 1e6, 3),
        "relax_achieved_GBs": round(prof["bytes"] / (prof["ms"] * 1e-3) / 1e9, 1) if prof["ms"] > 0 else None,
        "iterations": info["iterations"],
        "final_rel_residual": float(hist[-1] / hist[0]),
        "max_rel_err_vs_fp64_composite": err,
    }
    log(f"mixed precision: {out}")
    ms.close()
    return out


def vc_leg(args, ctx, h, rhs):
    """Variable coefficients (VCAMRPoissonOp2) on the benchmark's AMR hierarchy:
    -div(b grad phi) = rhs with a smooth positive face coefficient b of contrast 4
    on every level, solved to 1e-10 in FP64 and in mixed precision.  The GSRB
    class's algorithmic bytes are 56 B per cell-relax in FP64 (phi read and
    written, rhs, aCoef and the three face arrays read once per red-black
    iteration; lambda is a function of aCoef and bCoef)."""
    import amrmg
    import hierarchy as H
    t0 = time.time()
    a, b = H.vc_coefficients(h, seed=7, contrast=4.0)
    log(f"vc: coefficients for {h.total_cells()} cells in {time.time() - t0:.1f}s")
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=args.iter_max)
    out = {"workload": f"the 3-level AMR hierarchy above (base {args.base}^3), Dirichlet 0, alpha=0 beta=1, "
                       f"b = 1 + 3(0.5 + 0.5 sin cos cos) on faces of every level, a smooth, Gaussian-bump rhs, "
                       f"tol 1e-10"}
    for prec in (0, 1):
        s = amrmg.AMRMultiGrid(ctx, h, precision=prec)
        s.setSolverParameters(sp, rp)
        s.set_vc_coefs(a, b)
        s.set_field(s.RHS, rhs)
        for _ in range(args.warmup):
            s.solve()
        ctx.profile(["gsrb"])
        ctx.profile_reset()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            hist, info = s.solve()
        wall = time.perf_counter() - t1
        prof = ctx.profile_read("gsrb")
        ctx.profile([])
        key = "fp64" if prec == 0 else "mixed"
        out[key] = {
            "ms_per_step": round(wall / args.steps * 1e3, 3),
            "solves_per_s": round(args.steps / wall, 4),
            "iterations": info["iterations"],
            "final_rel_residual": float(hist[-1] / hist[0]),
            "mcell_relax_per_s": round(prof["units"] / wall / 1e6, 3),
            "relax_achieved_GBs": round(prof["bytes"] / (prof["ms"] * 1e-3) / 1e9, 1) if prof["ms"] > 0 else None,
            "relax_avg_launch_ms": round(prof["ms"] / max(prof["launches"], 1), 5),
            "relax_share_of_solve": round(prof["ms"] * 1e-3 / wall, 3),
        }
        log(f"vc {key}: {out[key]}")
        s.close()
    out["mixed_speedup"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    return out


def heat_leg(args, ctx, h, rhs):
    """Time stepping over the benchmark's AMR hierarchy: d phi/dt = Laplacian(phi) + S
    with the Gaussian bump as S, from phi = 0, dt = 1e-3, AMRTGA (two Helmholtz
    solves per step from phi^n) and BackwardEuler (one solve per step from phi^n),
    FP64, each solve to 1e-10.  Steps after the first are warm-started, which is the
    use the device-resident hierarchy serves."""
    import amrmg
    import hierarchy as H
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=args.iter_max)
    out = {"workload": f"the 3-level AMR hierarchy above (base {args.base}^3), dphi/dt = lap(phi) + S, "
                       f"S = Gaussian bump, phi0 = 0, dt = 1e-3, each solve to 1e-10, FP64"}
    nsteps = max(args.steps, 2)
    for scheme in ("tga", "be"):
        s = amrmg.AMRMultiGrid(ctx, h)
        s.setSolverParameters(sp, rp)
        s.set_field(s.PHI, 0.0 * rhs)
        heat = amrmg.HeatSolver(s, scheme)
        heat.set_source(rhs)
        heat.step(1e-3)  # warm-up step (not timed)
        its = []
        t0 = time.perf_counter()
        for _ in range(nsteps):
            i1, i2 = heat.step(1e-3)
            its.append((i1["iterations"], i2["iterations"]) if scheme == "tga" else (i1["iterations"],))
        wall = time.perf_counter() - t0
        out[scheme] = {"ms_per_step": round(wall / nsteps * 1e3, 3), "steps_per_s": round(nsteps / wall, 4),
                       "solve_iterations_per_step": its[-1]}
        log(f"heat {scheme}: {out[scheme]}")
        heat.close()
        s.close()
    return out


def batch_leg(args, ctx):
    """The batched workload: independent single-level Helmholtz problems of edge
    16..128 (log-uniform), solved (a) one after another with the plain FP64 solver,
    each solve timed alone with its inputs already in HBM, (b) together with the
    batched FP64 solver and (c) together with the batched mixed-precision solver,
    one launch per V-cycle stage for all problems."""
    import numpy as np
    import amrmg
    import hierarchy as H
    t0 = time.time()
    probs = H.batch_problems(args.batch, seed=1004)
    cells = sum(h.total_cells() for h, _ in probs)
    log(f"batch: {len(probs)} problems, {cells / 1e6:.1f} M cells (generated in {time.time() - t0:.1f}s)")
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=100)
    # (a) plain FP64 path: one solver per edge size, coefficients and rhs set per problem (untimed)
    solvers = {}
    t_plain, it_plain, phi_plain = 0.0, [], {}
    for q, (h, rhs) in enumerate(probs):
        n = h.domain[1][0] + 1
        if n not in solvers:
            s = amrmg.AMRMultiGrid(ctx, h)
            s.setSolverParameters(sp, rp)
            s.set_field(s.RHS, rhs)
            s.solve()  # warm-up of this size
            solvers[n] = s
        s = solvers[n]
        s.set_alpha_beta(h.alpha, h.beta)
        s.set_field(s.RHS, rhs)
        t1 = time.perf_counter()
        _, info = s.solve()
        t_plain += time.perf_counter() - t1
        it_plain.append(info["iterations"])
        if q % 16 == 0:
            phi_plain[q] = s.get_field(s.PHI)
    for s in solvers.values():
        s.close()
    log(f"plain FP64 one-by-one: {t_plain:.3f}s for {len(probs)} solves")

    def batched(prec, order=0):
        b = amrmg.BatchSolver(ctx, [h for h, _ in probs], prec, lanes=args.batch_lanes)
        b.setSolverParameters(sp, rp)
        if order:
            b.set_reduction_order(order)
        for q, (_, rhs) in enumerate(probs):
            b.set_rhs(q, rhs)
        if not args.no_tune:
            b.tune_relax()
        for _ in range(max(args.warmup, 1)):
            b.solve()  # the first V-cycle captures the graph
        # timed: profiler off, every V-cycle replayed from the captured HIP graph
        t1 = time.perf_counter()
        for _ in range(args.steps):
            it, ex = b.solve()
        wall = (time.perf_counter() - t1) / args.steps
        replays = b.set_graph(None)
        st = b.stats()
        phis = {q: b.get_phi(q) for q in phi_plain}
        hists = [b.history(q) for q in range(len(probs))]
        # one more solve with per-kernel HIP-event timing (direct launches, same kernels)
        ctx.profile(["gsrb", "residual"])
        ctx.profile_reset()
        b.solve()
        rel = ctx.profile_read("gsrb")
        res = ctx.profile_read("residual")
        ctx.profile([])
        b.close()
        return wall, it, ex, rel, res, phis, hists, st, replays

    w64, it64, ex64, rel64, _, phi64, hist64, st64, rep64 = batched(0)
    bitwise = all(np.array_equal(phi64[q], phi_plain[q]) for q in phi_plain)
    log(f"batched FP64: {w64 * 1e3:.1f} ms per batch; bitwise equal to plain on the {len(phi_plain)} sampled: {bitwise}")
    wmp, itmp, exmp, relmp, resmp, phimp, histmp, stmp, repmp = batched(1)
    err = max(float(np.abs(phimp[q] - phi64[q]).max() / np.abs(phi64[q]).max()) for q in phi64)
    conv = all(h[-1] <= 1e-10 * h[0] for h in histmp)
    # the same two batches with the bottom solves' reductions in the tree order
    wt64, _, _, _, _, phit64, histt64, _, _ = batched(0, order=1)
    wtmp, _, _, _, _, phitmp, histtmp, _, _ = batched(1, order=1)

    def max_err(phis):
        return max(float(np.abs(phis[q] - phi_plain[q]).max() / np.abs(phi_plain[q]).max()) for q in phi_plain)
    tree = {
        "batched_fp64_solves_per_s": round(len(probs) / wt64, 2),
        "batched_mixed_solves_per_s": round(len(probs) / wtmp, 2),
        "mixed_speedup_vs_fp64_reference_order": round(w64 / wtmp, 3),
        "ms_per_batch_fp64": round(wt64 * 1e3, 3),
        "ms_per_batch_mixed": round(wtmp * 1e3, 3),
        "max_rel_err_fp64_vs_plain_sampled": max_err(phit64),
        "max_rel_err_mixed_vs_plain_sampled": max_err(phitmp),
        "converged_all": all(h[-1] <= 1e-10 * h[0] for h in histt64 + histtmp),
    }
    log(f"batched, tree reduction order in the bottoms: {tree}")
    out = {
        "workload": f"{len(probs)} single-level Helmholtz problems, edge 16..128 log-uniform (seed 1004), "
                    "64^3 max boxes, alpha~U(0,1), beta~U(0.5,2), Dirichlet 0, sinusoid+N(0,1e-3) rhs, tol 1e-10",
        "problems": len(probs),
        "cells": cells,
        "plain_fp64_solves_per_s": round(len(probs) / t_plain, 2),
        "batched_fp64_solves_per_s": round(len(probs) / w64, 2),
        "batched_mixed_solves_per_s": round(len(probs) / wmp, 2),
        "speedup_batched_mixed_vs_plain_fp64": round(t_plain / wmp, 2),
        "speedup_batched_fp64_vs_plain_fp64": round(t_plain / w64, 2),
        "ms_per_batch_fp64": round(w64 * 1e3, 3),
        "ms_per_batch_mixed": round(wmp * 1e3, 3),
        "batched_fp64_bitwise_equal_plain_sampled": bitwise,
        "mixed_converged_all": conv,
        "max_rel_err_mixed_vs_fp64_sampled": err,
        "iterations_mean_plain_fp64": round(float(np.mean(it_plain)), 3),
        "iterations_mean_mixed": round(float(np.mean(itmp)), 3),
        "relax_mixed_achieved_GBs": round(relmp["bytes"] / (relmp["ms"] * 1e-3) / 1e9, 1) if relmp["ms"] else None,
        "relax_mixed_frac": round(relmp["bytes"] / (relmp["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4) if relmp["ms"] else None,
        "relax_fp64_achieved_GBs": round(rel64["bytes"] / (rel64["ms"] * 1e-3) / 1e9, 1) if rel64["ms"] else None,
        "residual_mixed_achieved_GBs": round(resmp["bytes"] / (resmp["ms"] * 1e-3) / 1e9, 1) if resmp["ms"] else None,
        "mcell_relax_per_s_mixed": round(stmp["cell_relax"] / wmp / 1e6, 1),
        "vcycles_from_hip_graph": {"fp64": rep64, "mixed": repmp},
        "tree_order_bottoms": tree,
        "lanes": args.batch_lanes,
    }
    log(f"batched: {out}")
    return out


def serve_leg(args):
    """Serving: 1-2-level AMR Poisson requests (bases 32..256^3) through the
    continuous-admission request loop.  Capacity is measured closed-loop (all
    requests submitted at once); then the same requests arrive as a Poisson
    process at 0.8x that capacity (seeded) and latency is arrival to result."""
    import threading
    import numpy as np
    import amrmg
    import hierarchy as H
    t0 = time.time()
    reqs = H.serve_requests(args.serve, seed=1005)
    log(f"serve: {len(reqs)} requests, mean {np.mean([h.total_cells() for h, _ in reqs]) / 1e6:.2f} M cells "
        f"(generated in {time.time() - t0:.1f}s)")
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=100)
    budget = int(args.serve_budget_gb * (1 << 30))

    def run(arrivals):
        eng = amrmg.ServeEngine(0, slots=args.serve_slots, budget_bytes=budget, sparams=sp, rparams=rp,
                                max_large=args.serve_max_large, large_cells=args.serve_large_cells)
        # warm the resident hierarchies (one request per distinct hierarchy)
        seen = {}
        for h, rhs in reqs:
            if id(h) not in seen:
                seen[id(h)] = eng.submit(h, rhs)
        for i in seen.values():
            eng.wait(i, want_phi=False)
        ids = [None] * len(reqs)

        def producer():
            start = time.perf_counter()
            for q, (h, rhs) in enumerate(reqs):
                if arrivals is not None:
                    delay = arrivals[q] - (time.perf_counter() - start)
                    if delay > 0:
                        time.sleep(delay)
                ids[q] = eng.submit(h, rhs)
        th = threading.Thread(target=producer)
        t1 = time.perf_counter()
        th.start()
        th.join()
        res = [eng.wait(i, want_phi=False) for i in ids]
        wall = time.perf_counter() - t1
        st = eng.stats()
        eng.close()
        return res, wall, st

    res, wall, st = run(None)
    cap = len(reqs) / wall
    log(f"serve closed loop: {cap:.2f} solves/s, stats {st}")
    rate = 0.8 * cap
    rng = np.random.default_rng(1005)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
    res2, wall2, st2 = run(arrivals)
    lat = np.array([r["latency"] for r in res2]) * 1e3
    base = np.array([h.domain[1][0] + 1 for h, _ in reqs])
    by_base = {}
    for b in np.unique(base):
        m = base == b
        by_base[str(int(b))] = {"requests": int(m.sum()),
                                "latency_ms_p50": round(float(np.percentile(lat[m], 50)), 2),
                                "latency_ms_p99": round(float(np.percentile(lat[m], 99)), 2)}
    out = {
        "workload": f"{len(reqs)} requests: 1-2-level AMR Poisson, base U{{32,64,128,256}}^3, refined fraction "
                    "U(0.05,0.25), Gaussian-bump rhs, tol 1e-10, Poisson arrivals at 0.8x closed-loop capacity "
                    "(seed 1005)",
        "slots": args.serve_slots,
        "max_large_running": args.serve_max_large,
        "large_cells": args.serve_large_cells,
        "hbm_budget_gb": args.serve_budget_gb,
        "closed_loop_solves_per_s": round(cap, 3),
        "arrival_rate_per_s": round(rate, 3),
        "open_loop_solves_per_s": round(len(reqs) / wall2, 3),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
        "latency_ms_max": round(float(lat.max()), 2),
        "latency_by_base": by_base,
        "converged_all": all(r["exit_status"] & 1 for r in res2),
        "max_running": st2["max_running"],
        "hierarchies_created": st2["created"],
        "max_hbm_gb": round(st2["max_bytes"] / (1 << 30), 2),
    }
    log(f"serve: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--base", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline protocol: median of 5 full solves on all cores, one on 1 core, and the 64^3 case")
    ap.add_argument("--no-c1", action="store_true", help="skip the single-level 64^3 leg")
    ap.add_argument("--iter-max", type=int, default=100,
                    help="cap on V-cycles per solve (a short cap is for profiling runs only)")
    ap.add_argument("--no-tune", action="store_true", help="skip the relaxation autotuner (default shapes)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-vc", action="store_true", help="skip the variable-coefficient leg")
    ap.add_argument("--no-heat", action="store_true", help="skip the heat-stepping leg")
    ap.add_argument("--batch", type=int, default=512, help="problems in the batched leg (0 = skip)")
    ap.add_argument("--batch-lanes", type=int, default=3, choices=(1, 2, 3),
                    help="2 or 3: the problems with large multigrid bottoms lockstep on further streams")
    ap.add_argument("--serve", type=int, default=2000, help="requests in the serving leg (0 = skip)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-max-large", type=int, default=4,
                    help="at most this many requests of at least --serve-large-cells cells run at once")
    ap.add_argument("--serve-large-cells", type=int, default=1 << 22)
    ap.add_argument("--serve-budget-gb", type=float, default=96.0)
    ap.add_argument("--copy-mb", type=int, default=2048,
                    help="size of the plain copy kernel run after the timed region (0 = skip)")
    ap.add_argument("--trace-marks", action="store_true",
                    help="launch a 4 KiB copy kernel before and after each timed region of the FP64 and mixed "
                         "legs (region boundaries for --summarize-trace)")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("KERNEL_TRACE_CSV", "OUT_MD"))
    ap.add_argument("--trace-names", default="Kernel time, 3-level AMR Poisson (base 256^3, 113.5 M cells);"
                    "FP64 solve (timed);mixed-precision solve (timed)")
    ap.add_argument("--trace-command", default="")
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "FETCH_RUN_JSON"),
                    help="write the GSRB traffic summary from two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(*args.summarize_trace, names=args.trace_names.split(";"), command=args.trace_command)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, copy_bytes=args.copy_mb << 20, out_path=TRAFFIC_JSON)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import amrmg
    import hierarchy as H

    hier = H.amr_hierarchy(args.base, fracs=(0.25, 0.06), max_box=64, block=8)
    cells = [hier.cells(l) for l in range(hier.nlev)]
    log(f"hierarchy: {[len(l) for l in hier.levels]} boxes, cells {cells} (total {sum(cells)})")
    t0 = time.time()
    rhs = H.bump_rhs(hier, 1002)
    log(f"rhs generated in {time.time() - t0:.1f}s")
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=args.iter_max)

    ctx = amrmg.Context(0)
    t0 = time.time()
    solver = amrmg.AMRMultiGrid(ctx, hier)
    solver.setSolverParameters(sp, rp)
    solver.set_field(solver.RHS, rhs)
    log(f"solver set up in {time.time() - t0:.1f}s (device bytes ~{solver.device_bytes() / 1e9:.1f} GB)")

    tune = None
    if not args.no_tune:
        t0 = time.time()
        tune = solver.tune_relax()
        log(f"relaxation tuned in {time.time() - t0:.2f}s: " +
            ", ".join(f"{r['kind']}{r['index']} {r['shape'][0]}x{r['shape'][1]}/kc{r['kc']} {r['ms']:.3f}ms"
                      for r in tune))
    for w in range(args.warmup):
        hist, info = solver.solve()
        log(f"warmup {w}: {info} residual {hist[0]:.3e} -> {hist[-1]:.3e}")

    ctx.profile(["gsrb"])
    ctx.profile_reset()
    if args.trace_marks:
        ctx.copy_probe(4096, 1)
    t0 = time.perf_counter()
    infos = []
    for s in range(args.steps):
        hist, info = solver.solve()
        infos.append((hist, info))
        log(f"step {s}: {info['iterations']} iterations, rel residual {hist[-1] / hist[0]:.3e}")
    t1 = time.perf_counter()
    if args.trace_marks:
        ctx.copy_probe(4096, 1)
    prof = ctx.profile_read("gsrb")
    ctx.profile([])

    wall = t1 - t0
    ms_per_step = wall / args.steps * 1e3
    cell_relax = prof["units"]                 # one cell-relax = one red + one black update of a cell
    value = cell_relax / wall / 1e6
    avg_launch_ms = prof["ms"] / max(prof["launches"], 1)
    achieved = prof["bytes"] / (prof["ms"] * 1e-3) / 1e9 if prof["ms"] > 0 else 0.0
    bytes_per_launch = prof["bytes"] / max(prof["launches"], 1)
    ratio = read_traffic()
    traffic = round(ratio * bytes_per_launch, 1) if ratio else None
    final_rel = float(infos[-1][0][-1] / infos[-1][0][0])
    ok = all(h[-1] <= 1e-10 * h[0] for h, _ in infos)
    mixed = None
    if not args.no_mixed:
        mixed = mixed_leg(args, ctx, hier, rhs, sp, rp, solver, wall / args.steps)
    vcout = None
    if not args.no_vc:
        solver.close()
        vcout = vc_leg(args, ctx, hier, rhs)
    heat = None
    if not args.no_heat:
        solver.close()
        heat = heat_leg(args, ctx, hier, rhs)
    batch = None
    if args.batch > 0:
        solver.close()
        batch = batch_leg(args, ctx)
    serve = None
    if args.serve > 0:
        solver.close()
        serve = serve_leg(args)
    c1, c1case = None, None
    if not args.no_c1:
        c1, c1case = c1_leg(args, ctx)
    copy_gbs = None
    if args.copy_mb > 0:
        nbytes = args.copy_mb << 20
        ms = ctx.copy_probe(nbytes, 10)
        copy_gbs = 2 * nbytes / (ms * 1e-3) / 1e9
        log(f"plain copy kernel ({args.copy_mb} MiB): {ms:.3f} ms, {copy_gbs:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        log("timing the CPU oracle (one full solve) ...")
        if c1case is None:
            import hierarchy as H
            ch = H.single_level(64, 32, 8)
            c1case = (ch, H.sinusoid_rhs(ch, 1001), *H.solver_params(eps=1e-10, hang=1e-15, iterMax=100))
        cpu = cpu_baseline(args, hier, rhs, sp, rp, c1case)
        log(f"cpu baseline {cpu['value']:.1f} Mcell-relax/s, {cpu['solves_per_s']:.4f} solves/s")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-relax/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3-level AMR Poisson, base {args.base}^3, ratio 2, centred refined regions 25% then 6% "
                        f"of parent, 64^3 max boxes, Dirichlet 0, alpha=0 beta=1, Gaussian-bump rhs seed 1002, "
                        f"tol 1e-10, V(2,2) + BiCGStab bottom, FP64",
            "levels": [len(l) for l in hier.levels],
            "cells": sum(cells),
        },
        "solves_per_s": round(args.steps / wall, 4),
        "iterations": infos[-1][1]["iterations"],
        "final_rel_residual": final_rel,
        "converged": ok,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            "traffic_over_algorithmic": ratio,
            "kernel": "GSRB relaxation class: k_gsrb_fused / k_gsrb_zmarch (one red+black iteration of a whole level per launch)",
            "avg_launch_ms": round(avg_launch_ms, 5),
            "algorithmic_bytes_per_launch": round(bytes_per_launch, 1),
            "launches": prof["launches"],
            "measured_copy_GBs": round(copy_gbs, 1) if copy_gbs else None,
        },
        "relax_tiles": ({f"{r['kind']}{r['index']}": f"{r['shape'][0]}x{r['shape'][1]}/kc{r['kc']}" for r in tune}
                        if tune else None),
        "mixed_precision": mixed,
        "variable_coefficients": vcout,
        "heat": heat,
        "batched": batch,
        "serving": serve,
        "single_level_64": c1,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    solver.close()
    ctx.close()


if __name__ == "__main__":
    main()

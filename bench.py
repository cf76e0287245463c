"""Throughput of the explicit finite-volume step on one MI355X.

Workload: the periodic Navier-Stokes case on the unit square, 1024x1024
quads split into 2,097,152 triangles (seeded +-10% node jitter), uniform
Mach-0.5 flow at 30 degrees plus a Gaussian density/pressure pulse and 1e-3
velocity noise; mu = 1e-4, Pr = 0.72, gamma = 1.4; least-squares gradients,
exact Godunov flux, SSP-RK3, CFL 0.4 with the spectral time step fixed at
init.  A step = one full time step (three RK stages) over all cells; the
metric is cell-updates/s with the state resident in HBM.

Prints one JSON line.  --gpus above 1 exits: the library runs on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(REPO, "cfd-2d_amd"), os.path.join(REPO, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
DOMINANT = "edge_flux"     # the kernel timed with HIP events inside the timed region
KERNEL_PASS_STEPS = 10     # untimed steps that time every other kernel
METRIC = "FP64 cell-updates/s (2.1M-tri NS) + % HBM peak; batched/serial and adjoint/fwd ratios"

# algorithmic bytes per launch (arrays touched once per pass, neighbour reuse
# perfect): gradient 156 B/cell with LSQ (164 with Green-Gauss); edge flux
# 128 B/edge (c1, c2, n, l, Gauss-point offsets read, the flux record written)
# plus, reading stored gradients, 96 B/cell (P and G records) or, forming
# them itself as unlimited steps do, 104 B/cell (P record, the 24 B of faces
# and neighbours and the 48 B of LSQ weights); update 212 B/cell (the 4 B of
# padding in each 16-byte face/neighbour record is not counted)
def alg_bytes(kernel, nc, ne, lsq=True, fused=True):
    if kernel == "gradient":
        return (156 if lsq else 164) * nc
    if kernel == "edge_flux":
        return 128 * ne + ((104 if lsq else 112) if fused else 96) * nc
    if kernel == "update":
        return 212 * nc
    return 0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(case_fn, steps, threads=1, implicit=False):
    import oracle
    c = case_fn()
    t0 = time.perf_counter()
    s = oracle.OracleSolver(c, threads=threads)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    out = dict(value=s.nc * steps / dt, unit="cell-updates/s", cores=threads, kind="port",
               sample="%d step(s) of the same %d-cell workload on %d host thread(s) (C restatement, gcc -O2)"
               % (steps, s.nc, threads), seconds=round(dt, 3), setup_seconds=round(setup, 3))
    if implicit:
        # the implicit method's matrix fill, one serial edge-loop pass as the reference makes it
        t0 = time.perf_counter()
        s.implicit_assemble(steady=True, cfl=0.8)
        out["implicit_assembly_s"] = round(time.perf_counter() - t0, 3)
    return out


def host_cpu():
    """Host CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return dict(model=model, usable_cores=usable, logical_cpus=os.cpu_count())


def reference_cpu(steps=100, opt="O2"):
    """The reference's own FVM_TVD (compiled from its sources into
    oracle/_ref) on the 8,000-cell shock tube, where it can run; opt "O0" is
    the build as its CMakeLists.txt ships it (no optimisation flags,
    CMakeLists.txt:4-6), "O2" the optimised build."""
    exe = os.path.join(REPO, "oracle", "_ref", "ref_harness" if opt == "O2" else "ref_harness_O0")
    if not os.path.exists(exe):
        return None
    from cfd2d import meshgen, unv
    with tempfile.TemporaryDirectory() as d:
        m = meshgen.rect_mesh(200, 20, 1.0, 0.2, seed=1)
        unv.write_unv(m, os.path.join(d, "mesh.unv"))
        regions = [dict(name="cells_left", Vx=0.0, Vy=0.0, T=1.0, P=1.0),
                   dict(name="cells_right", Vx=0.0, Vy=0.0, T=1.0, P=0.5)]
        bnds = [dict(name="edge_bottom", type="BOUND_WALL_SLIP"), dict(name="edge_top", type="BOUND_WALL_SLIP"),
                dict(name="edge_left", type="BOUND_OUTLET"), dict(name="edge_right", type="BOUND_OUTLET")]
        with open(os.path.join(d, "task.xml"), "w") as f:
            f.write(unv.task_xml("mesh.unv", regions=regions, boundaries=bnds, step_max=steps))
        try:
            out = subprocess.run([exe, "time", "task.xml", str(steps)], cwd=d, capture_output=True, text=True,
                                 timeout=120, check=True).stdout
        except (subprocess.SubprocessError, OSError) as e:
            log("reference timing skipped: %s" % e)
            return None
    line = [l for l in out.splitlines() if l.startswith("{")][-1]
    r = json.loads(line)
    return dict(value=r["cell_updates_per_s"], unit="cell-updates/s", cores=1, kind="reference",
                sample="%d steps of the reference FVM_TVD (GG, exact Godunov, RK2) on the 8,000-cell tube, g++ -%s%s"
                % (steps, opt, " (as shipped)" if opt == "O0" else ""))


def batched_vs_serial(n_cases, steps, seed=7):
    """Batched bump channels: n_cases channels (2k-60k cells, log-uniform,
    inflow Mach 0.2-3.0, angle -5..5 deg; Euler, exact Godunov, minmod, RK2)
    advanced together in one index space, against the same cases advanced one
    at a time (eager launches, and replayed HIP graphs)."""
    from cfd2d import batch
    from cfd2d import case as cases
    from cfd2d import meshgen
    specs = meshgen.bump_batch(n_cases, seed=seed)
    cs = [cases.bump_channel(s["ny"], height=s["height"], mach=s["mach"], angle_deg=s["angle_deg"],
                             seed=s["seed"]) for s in specs]
    cells = sum(c.mesh.n_cells for c in cs)
    b = batch.BatchSolver(cs)
    b.run(2)
    tuned = b.autotune()  # step variant of the batch (same results whichever)
    t0 = time.perf_counter()
    b.run(steps)
    t_batch = time.perf_counter() - t0
    b.close()
    from cfd2d import native
    solvers = [native.GpuSolver(c) for c in cs]
    picks = []
    for s in solvers:
        s.run(2)
        picks.append(s.autotune())  # each small case tuned for itself
    t0 = time.perf_counter()
    for s in solvers:
        s.step(steps)
        s.sync()
    t_serial = time.perf_counter() - t0
    for s in solvers:
        s.run(2, graph=True)
    t0 = time.perf_counter()
    for s in solvers:
        s.step_graph(steps)
        s.sync()
    t_graph = time.perf_counter() - t0
    for s in solvers:
        s.close()
    rate = lambda t: cells * steps / t
    return dict(cases=n_cases, cells=cells, steps=steps, batched=rate(t_batch), serial_eager=rate(t_serial),
                serial_graph=rate(t_graph), ratio_vs_eager=t_serial / t_batch, ratio_vs_graph=t_graph / t_batch,
                unit="cell-updates/s",
                autotune=dict(batch=tuned, serial_fused=sum(p["fuse"] == 1 for p in picks),
                              serial_tiled=sum(p["fuse"] == 2 for p in picks),
                              serial_graph=sum(p["graph"] for p in picks),
                              serial_stream=sum(p["stream"] for p in picks), serial_cases=len(picks)))


def adjoint_vs_forward(steps, ckpt, ny=289, flux="LAX"):
    """Adjoint workload: bump channel with ny=289 (501,126 triangles), Mach
    0.8, bump height 0.05; LAX flux (or the reference's exact Godunov flux,
    whose reverse differentiates the Newton iterations as evaluated, or
    HLLC),
    Green-Gauss gradients, Heun RK2, no limiter; J = y pressure force on the
    bump.  Wall time of cfd2d_adjoint
    (forward pass, checkpoints, recomputation, reverse sweep, gradient copied
    to the host) over the wall time of the same number of plain steps, each
    the faster of two calls.  The first call with each checkpoint interval
    (timed separately as "first_call_s") sizes the context's adjoint arena
    and records the sweep's graphs; the timed calls reuse them, as repeated
    calls do."""
    from cfd2d import case as cases
    from cfd2d import native
    c = cases.bump_channel(ny, height=0.05, mach=0.8, seed=3, flux=flux, gradient="GREEN_GAUSS", limiter="NONE",
                           rk_stages=2)
    g = native.GpuSolver(c)
    mask = g.edge_mask("bump")
    U0 = c.initial_state()
    g.run(5)

    def timed(fn, reps):
        best = None
        for _ in range(reps):
            g.set_state(U0)
            g.sync()
            t0 = time.perf_counter()
            r = fn()
            g.sync()
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        return best, r

    t_fwd, _ = timed(lambda: g.run(steps), 2)
    out = dict(cells=g.nc, edges=g.ne, steps=steps, workload="bump channel ny=%d, %s/GG/RK2, Mach 0.8" % (ny, flux),
               forward_s=round(t_fwd, 5))
    for name, ck in (("ckpt%d" % ckpt, ckpt), ("auto", 0)):
        t_first, _ = timed(lambda: g.adjoint(steps, mask, ckpt_every=ck), 1)
        t, (J, gU, gb) = timed(lambda: g.adjoint(steps, mask, ckpt_every=ck), 2)
        out[name] = dict(seconds=round(t, 5), ratio=round(t / t_fwd, 3), first_call_s=round(t_first, 5))
        out["J"] = J
    g.close()
    out["ratio"] = out["ckpt%d" % ckpt]["ratio"]
    out["ratio_auto"] = out["auto"]["ratio"]
    return out


def batched_adjoint(n_cases, steps, seed=11):
    """Batched adjoint: n_cases bump channels (5k-100k cells, log-uniform;
    Mach 0.3-2.0, angle -5..5 deg, bump height 0.02-0.1; LAX/GG/RK2, no
    limiter), J_k = y pressure force on case k's bump after `steps` steps.
    One cfd2d_adjoint call for all cases (automatic interval, arena from an
    untimed first call) against the batched forward run of the same steps;
    the returned gradients include dJ_k/dMach_k and dJ_k/dangle_k.  Then one
    cfd2d_adjoint_geometry call, which also returns every case's dJ_k/d(node
    coordinates), of which the bump nodes' y derivatives are reported."""
    from cfd2d import batch
    from cfd2d import case as cases
    from cfd2d import meshgen, sensitivity
    specs = meshgen.bump_batch(n_cases, 5000, 100000, seed=seed, mach_range=(0.3, 2.0))
    cs = [cases.bump_channel(s["ny"], height=s["height"], mach=s["mach"], angle_deg=s["angle_deg"], seed=s["seed"],
                             flux="LAX", gradient="GREEN_GAUSS", limiter="NONE", rk_stages=2) for s in specs]
    b = batch.BatchSolver(cs)
    mask = b.edge_mask("bump")
    U0 = b.ctx.state()
    b.run(2)
    b.ctx.set_state(U0)
    b.ctx.sync()
    t0 = time.perf_counter()
    b.run(steps)
    t_fwd = time.perf_counter() - t0
    b.ctx.set_state(U0)
    t0 = time.perf_counter()
    b.ctx.adjoint(steps, mask, 0)
    t_first = time.perf_counter() - t0
    b.ctx.set_state(U0)
    b.ctx.sync()
    t0 = time.perf_counter()
    res = b.adjoint(steps, mask, 0)
    t_adj = time.perf_counter() - t0
    b.ctx.set_state(U0)
    b.ctx.adjoint_geometry(steps, mask, 0)  # sizes the arena for the geometry accumulators
    b.ctx.set_state(U0)
    b.ctx.sync()
    t0 = time.perf_counter()
    gres = b.adjoint_geometry(steps, mask, 0)
    t_geo = time.perf_counter() - t0
    b.close()
    finite = all(np.isfinite(gU).all() and np.isfinite(gb).all() for _, gU, gb in res)
    finite_geo = all(np.isfinite(gx).all() for _, _, _, gx in gres)
    grads = [sensitivity.mach_angle_gradient(c, gU, gb) for c, (_, gU, gb) in zip(cs, res)]
    bump_dy = []
    for c, (_, _, _, gx) in zip(cs, gres):
        m = c.mesh
        nodes = np.unique(m.bedges[m.bedge_group == m.group_index("bump")])
        bump_dy.append(gx[nodes, 1])
    cells = sum(c.mesh.n_cells for c in cs)
    return dict(cases=n_cases, cells=cells, steps=steps, forward_s=round(t_fwd, 4), adjoint_s=round(t_adj, 4),
                first_call_s=round(t_first, 4), ratio=round(t_adj / t_fwd, 3),
                adjoint_cell_steps_per_s=cells * steps / t_adj, finite=bool(finite),
                dJ_dMach_case0=grads[0][0], dJ_dangle_case0=grads[0][1],
                geometry=dict(adjoint_s=round(t_geo, 4), ratio=round(t_geo / t_fwd, 3), finite=bool(finite_geo),
                              bump_nodes=int(sum(len(v) for v in bump_dy)),
                              dJ_dy_bump_case0_l2=float(np.linalg.norm(bump_dy[0])),
                              dJ_dy_bump_case0_max=float(np.abs(bump_dy[0]).max())))


def renumbering(n, steps=10):
    """The same periodic Navier-Stokes case (n x n quads x 2) in three cell
    orders: the generator's row order, a seeded random shuffle (what an
    unordered external mesh looks like), and the shuffle renumbered along a
    Hilbert curve (cfd2d/renumber.py).  ms per step of each."""
    from cfd2d import case as cases
    from cfd2d import native, renumber
    base = cases.periodic_pulse(n, seed=2)
    shuffled, _ = renumber.renumber_case(base, "random", seed=1)
    hilbert, _ = renumber.renumber_case(shuffled, "hilbert")
    out = dict(cells=base.mesh.n_cells, steps=steps)
    for name, c in (("generator", base), ("random", shuffled), ("hilbert", hilbert)):
        s = native.GpuSolver(c)
        s.run(2)
        t0 = time.perf_counter()
        s.run(steps)
        out[name + "_ms_per_step"] = round((time.perf_counter() - t0) * 1e3 / steps, 4)
        s.close()
    return out


def remediation_cost(s, counts=(0, 100, 10000), seed=1):
    """k_remediate (FVM_TVD::remediateLimCells) on the benchmark's context:
    for each count, that many cells chosen at random are LIM-flagged, one
    step is run and the remediation launch timed (HIP events).  Randomly
    scattered cells are the sweep's worst case: nearly every one sits in a
    256-cell block of its own.  Changes the context's state (run last)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    U0 = s.state()
    out = {}
    for k in counts:
        s.set_state(U0)
        f = np.zeros(s.nc, np.uint32)
        f[rng.choice(s.nc, k, replace=False)] = 0x2
        s.set_flags(f)
        s.timing(True, kernels=["remediate"])
        s.step(1)
        s.sync()
        ms, n = s.timing_read()["remediate"]
        s.timing(False)
        out[str(k)] = round(1e3 * ms / max(n, 1), 1)
    s.set_flags(np.zeros(s.nc, np.uint32))
    return dict(lim_cells_to_us=out, workload="LIM cells scattered at random over the benchmark's mesh, one step")


def implicit_assembly(s, reps=20, traffic=None):
    """The implicit method's matrix fill (cfd2d_implicit_assemble) on the
    context of the main workload: device-only assembly timed per kernel with
    HIP events over reps calls, then one call that also copies the block rows
    back to host memory (what a CPU solver is fed)."""
    nc, ne = s.nc, s.ne
    s.implicit_assemble(steady=True, cfl=0.8, outputs=False)
    s.timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        s.implicit_assemble(steady=True, cfl=0.8, outputs=False)
    wall = (time.perf_counter() - t0) / reps
    tim = s.timing_read()
    s.timing(False)
    alg = dict(impl_gradient=164 * nc, impl_edge=128 * ne + 96 * nc, impl_assemble=632 * nc + 56 * ne)
    per = {}
    total = 0.0
    for k, b in alg.items():
        ms, n = tim[k]
        avg = ms / max(n, 1)
        total += avg
        per[k] = dict(avg_ms=round(avg, 5), launches=n, alg_bytes=b, gbs=round(b / (avg * 1e-3) / 1e9, 1))
    bufs = s.implicit_buffers()  # page-locked host arrays, as a CPU solver would hold them
    s.implicit_assemble(steady=True, cfl=0.8, out=bufs)
    t0 = time.perf_counter()
    B, C, R, T = s.implicit_assemble(steady=True, cfl=0.8, out=bufs)
    with_copy = time.perf_counter() - t0
    out_bytes = B.nbytes + C.nbytes + R.nbytes + T.nbytes
    finite = bool(np.isfinite(B).all() and np.isfinite(R).all())
    del B, C, R, T, bufs
    a = per["impl_assemble"]
    roof = dict(bound="hbm", kernel="impl_assemble", achieved=a["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(a["gbs"] / HBM_PEAK_GBS, 4), traffic=(traffic or {}).get("impl_assemble"),
                alg_bytes_per_launch=a["alg_bytes"], avg_ms=a["avg_ms"])
    return dict(cells=nc, reps=reps, kernels=per, roofline=roof, device_ms=round(total, 4),
                wall_ms_device_only=round(wall * 1e3, 4), cells_per_s=nc / (total * 1e-3),
                with_copy_ms=round(with_copy * 1e3, 2), output_bytes=out_bytes,
                d2h_gbs=round(out_bytes / (with_copy - total * 1e-3) / 1e9, 1), finite=finite)


def load_sq(workload):
    """VALU instructions per wave and FP64 issue fraction per kernel from the
    committed rocprofv3 SQ-counter summary."""
    try:
        with open(os.path.join(REPO, "profiles", "sq_counters.json")) as f:
            return json.load(f).get(workload) or {}
    except (OSError, ValueError):
        return {}


def load_traffic(workload):
    """HBM bytes per launch of each kernel from the committed rocprofv3 PMC summary."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(workload)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024, help="quads per side of the periodic square")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay each step as a recorded HIP graph")
    ap.add_argument("--batch-cases", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-ckpt", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--badj-cases", type=int, default=64)
    ap.add_argument("--badj-steps", type=int, default=300)
    ap.add_argument("--no-badj", action="store_true")
    ap.add_argument("--no-renumber", action="store_true")
    ap.add_argument("--no-implicit", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the library runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from cfd2d import case as cases
    from cfd2d import native

    if native.device_count() < 1:
        sys.exit("bench.py: no HIP device")
    t0 = time.perf_counter()
    case_fn = lambda: cases.periodic_pulse(args.n, seed=2)
    c = case_fn()
    s = native.GpuSolver(c)
    log("setup %.1f s: %d cells, %d edges" % (time.perf_counter() - t0, s.nc, s.ne))
    s.run(args.warmup, graph=args.graph)
    tuned = s.autotune()  # step variant (same results whichever; timed steps below launch eagerly)
    # timed region: HIP events only around the dominant kernel (the edge flux
    # pass), so that the events themselves barely disturb the stream
    s.timing(True, kernels=[DOMINANT])
    s.sync()
    t0 = time.perf_counter()
    if args.graph:
        s.step_graph(args.steps)
    else:
        s.step(args.steps)
    s.sync()
    dt = time.perf_counter() - t0
    tim_timed = s.timing_read()
    # the other kernels' averages: an untimed pass with events around every launch
    s.timing(True)
    s.run(KERNEL_PASS_STEPS)
    tim = s.timing_read()
    s.timing(False)
    if tim_timed[DOMINANT][1]:
        tim[DOMINANT] = tim_timed[DOMINANT]
    hits, nonfinite = s.status()
    impl = None if args.no_implicit else implicit_assembly(s, traffic=load_traffic("periodic_ns_%d" % args.n))
    remed = remediation_cost(s)
    log("remediation: %s" % remed)
    s.close()
    value = s.nc * args.steps / dt
    per = {}
    for k, (ms, n) in tim.items():
        if n:
            avg = ms / n
            b = alg_bytes(k, s.nc, s.ne, fused=bool(tuned["fuse"]))
            per[k] = dict(avg_ms=round(avg, 5), launches=n, alg_bytes=b,
                          gbs=round(b / (avg * 1e-3) / 1e9, 1) if b else None,
                          timed_region=(k == DOMINANT and bool(tim_timed[DOMINANT][1])))
    dom = max(per, key=lambda k: per[k]["avg_ms"] * per[k]["launches"])
    d = per[dom]
    traffic = load_traffic("periodic_ns_%d" % args.n)
    roof = dict(bound="hbm", kernel=dom, achieved=d["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(d["gbs"] / HBM_PEAK_GBS, 4) if d["gbs"] else None,
                traffic=(traffic or {}).get(dom), alg_bytes_per_launch=d["alg_bytes"], avg_ms=d["avg_ms"])
    sq = load_sq("periodic_ns_%d" % args.n).get(dom)
    if sq:
        # what actually limits the kernel: FP64 instruction issue (profiles/periodic_ns_1024_sq_counters.md)
        roof["valu_insts_per_wave"] = sq["valu_insts_per_wave"]
        roof["valu_issue_frac_fp64"] = sq["valu_issue_frac_fp64"]
    # per cell-update by the canonical three-pass stage (gradient, edge, update):
    # 3 stages x (464 + 128 e), e = 1.5.  Unlimited steps here run two passes
    # (the edge pass forms the gradients) and move 3 x (316 + 128 e) = 1 524 B.
    step_bytes = 1968 * s.nc
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered periodic triangulation, Gaussian pulse + seeded velocity noise)",
        "config": {"workload": "periodic Navier-Stokes, %d triangles (%dx%d quads x 2), LSQ gradients, exact "
                               "Godunov flux, SSP-RK3, CFL 0.4 spectral dt, mu=1e-4, Pr=0.72" % (s.nc, args.n, args.n),
                   "cells": s.nc, "edges": s.ne, "graph": bool(args.graph), "autotune": tuned},
        "roofline": roof,
        "step_hbm_frac": round(step_bytes * args.steps / dt / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": per,
        "newton_cap_hits": hits, "nonfinite": nonfinite,
    }
    out["remediation"] = remed
    if impl:
        out["implicit_assembly"] = impl
        log("implicit assembly: %s" % impl)
    if not args.no_batch:
        out["batched"] = batched_vs_serial(args.batch_cases, args.batch_steps)
        log("batched: %s" % out["batched"])
    if not args.no_adjoint:
        out["adjoint"] = adjoint_vs_forward(args.adj_steps, args.adj_ckpt)
        log("adjoint: %s" % out["adjoint"])
        out["adjoint_godunov"] = adjoint_vs_forward(args.adj_steps, args.adj_ckpt, flux="GODUNOV")
        log("adjoint, exact Godunov flux: %s" % out["adjoint_godunov"])
        out["adjoint_hllc"] = adjoint_vs_forward(args.adj_steps, args.adj_ckpt, flux="HLLC")
        log("adjoint, HLLC flux: %s" % out["adjoint_hllc"])
    if not args.no_badj:
        out["batched_adjoint"] = batched_adjoint(args.badj_cases, args.badj_steps)
        log("batched adjoint: %s" % out["batched_adjoint"])
    if not args.no_renumber:
        out["renumbering"] = renumbering(args.n)
        log("renumbering: %s" % out["renumbering"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case_fn, args.cpu_steps, implicit=not args.no_implicit)
        host = host_cpu()
        # the same restatement with OpenMP on the cores this process may use
        # (capped at 16, the host share of one GPU)
        nthr = max(1, min(16, host["usable_cores"] or 1))
        out["cpu_baseline_parallel"] = cpu_baseline(case_fn, max(args.cpu_steps, 4), threads=nthr)
        out["host_cpu"] = host
        ref = reference_cpu()
        if ref:
            out["cpu_reference"] = ref
        ref0 = reference_cpu(steps=30, opt="O0")
        if ref0:
            out["cpu_reference_as_shipped"] = ref0
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

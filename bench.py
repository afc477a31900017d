#!/usr/bin/env python3
"""Nonbonded-step benchmark on one MI355X.

Workload (default): the 288,000-atom water box - 96,000 rigid 3-site waters
on a seeded jittered lattice in a 14.2 nm cubic box, rc = rlist = 1.0 nm,
potential-shifted LJ on oxygen, Ewald real-space Coulomb with
ewald_rtol = 1e-5 (the plain-C kernel's quadratic-spline correction table by
default, the faster flavour on this GPU; --ewald analytical selects the
SIMD kernels' analytical correction).
A step is one pass of the nonbonded path over resident device buffers:
coordinate conversion into the cluster layout, the cluster-pair kernel with
forces, energies and shift forces, the force reduction into atom order and
the FP64 virial (gmxcalculator.cpp:70-83 plus the energy/virial outputs).
The grid and pair search run once before timing and are reported apart.

Throughput is useful pair interactions per second: the exact number of
non-excluded atom pairs within rc (the quantity the reference benchmark
estimates as N/2 (rho 4/3 pi rc^3 + 1), bench_setup.cpp:273-276) divided by
the step time.  The kernel's roofline uses the reference's stated flop count
per pair (107 for Ewald + LJ with energies, gmxlib/nrnb.cpp:106) times the
pairs the kernel computes (16 per cluster pair in its list), divided by the
kernel's duration measured with HIP events on its stream.

The same line carries the metric's two other parts, measured in the same
run: "batched" - the 64-replica water ensemble (replica sizes log-uniform in
1,000-16,000 molecules) as one batched launch, eager and replayed from a
captured HIP graph, against the same 64 systems launched one after another
(aggregate useful pairs/s); "fwd_bwd" - the 24,000-atom, 16-type
force-matching system: the forward forces call against the adjoint call
that returns dL/dx, dL/dq, dL/dC6 and dL/dC12 (the backward of the
torch.autograd.Function).

One GPU only: --gpus above 1 exits with a message.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NB pair-interactions/s & ms/step, 288k-atom water; batched and fwd+bwd vs plain"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOPS_PER_PAIR_VF = 107  # Ewald + LJ, forces and energies (nrnb.cpp:106)
FLOPS_PER_PAIR_F = 66  # forces only (nrnb.cpp:105)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def excluded_pairs_within(system, rc):
    import numpy as np

    n = 0
    for a in range(system.natoms):
        for b in system.excl_idx[system.excl_off[a]:system.excl_off[a + 1]]:
            if b > a:
                d = system.x[a] - system.x[b]
                d -= system.box * np.round(d / system.box)
                if float(d @ d) < rc * rc:
                    n += 1
    return n


def excluded_pairs_water(system, rc):
    # all three intramolecular pairs of a rigid water are within rc
    return 3 * (system.natoms // 3)


def host_cpu():
    """The host CPU's model name and the threads this process may use (its
    affinity, capped by OMP_NUM_THREADS when that is set)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return model, max(1, n)


def cpu_baseline(system, budget_s=12.0):
    """The reference's CPU kernel scheme timed on the host cores this process
    may use: full F+E steps (coordinate conversion, kernel, force reduction)
    of the same workload, median over a bounded sample.  On AVX-512 hosts the
    oracle build's AVX-512 kernel (the reference's SIMD scheme: a 4x4 cluster
    pair per 16-lane register, exclusion masks as vector masks, analytical
    Ewald correction, geometric LJ, OpenMP over i-entries with per-thread
    outputs reduced afterwards); elsewhere its plain-C kernel (-O3 AVX2/FMA,
    the same threading).  Returns (seconds per step, set-up seconds, samples,
    threads, cpu model, kernel name)."""
    from oracle import oracle as O

    model, threads = host_cpu()
    t0 = time.perf_counter()
    o = O.Oracle("bench_f32", "ewald", rlist=system.rlist, rcoulomb=system.rcoulomb)
    o.setup(system.box, system.x, system.types, system.charges, O.nblib_nbfp(system.c6, system.c12, "f32"),
            system.ntypes, system.excl_off, system.excl_idx)
    t_setup = time.perf_counter() - t0
    if o.simd_available():
        run, kind = o.compute_simd, "AVX-512 SIMD kernel (analytical Ewald correction)"
    else:
        run, kind = o.compute_threads, "plain-C kernel (table Ewald correction), -O3 AVX2/FMA"
    run(system.x, threads)  # warm-up
    times = []
    while sum(times) < budget_s and len(times) < 50:
        t0 = time.perf_counter()
        run(system.x, threads)
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], t_setup, len(times), threads, model, kind


def measured_traffic(workload, ewald):
    """HBM bytes per launch of the F+E+shift kernel from the committed
    rocprofv3 counter summary (profiles/nonbonded_<ewald>_<workload>_pmc.md,
    a line 'traffic_bytes_per_launch: N'), or None."""
    path = os.path.join(ROOT, "profiles", f"nonbonded_{ewald}_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def _time(fn, stream, reps, warm=3):
    import torch

    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_ensemble(ewald, stream, reps=20):
    """64 water boxes: sequential per-system launches vs one batched launch
    (eager and HIP-graph replay), forces + energies."""
    import numpy as np
    import torch

    from gmxapi_amd import nbx, systems

    reps_ = systems.ensemble(64)
    b = nbx.batch_for(reps_, ewald_correction=ewald)
    b.set_stream(stream.cuda_stream)
    x = np.concatenate([r.x.astype(np.float32).reshape(-1) for r in reps_])
    xd = torch.from_numpy(x).to("cuda")
    fd = torch.zeros_like(xd)
    flags = nbx.COMPUTE_FORCES | nbx.COMPUTE_ENERGY
    offs = np.cumsum([0] + [r.natoms for r in reps_])
    ctxs = [b.system(k) for k in range(len(reps_))]
    useful = sum(c.count_pairs(r.rcoulomb) - 3 * (r.natoms // 3) for c, r in zip(ctxs, reps_))

    def sequential():
        for k, c in enumerate(ctxs):
            c.compute_device(xd.data_ptr() + 12 * int(offs[k]), fd.data_ptr() + 12 * int(offs[k]), flags)

    seq_ms = _time(sequential, stream, reps)
    chosen = b.autotune(xd.data_ptr(), fd.data_ptr(), flags, repeats=5)
    eager_ms = _time(lambda: b.compute_device(xd.data_ptr(), fd.data_ptr(), flags), stream, reps)
    b.capture_graph(xd.data_ptr(), fd.data_ptr(), flags)
    graph_ms = _time(b.launch_graph, stream, reps)
    return {
        "workload": "64 water boxes (config-0 recipe), 1,000-16,000 molecules each, log-uniform",
        "natoms": int(offs[-1]),
        "useful_pairs": int(useful),
        "sequential_ms": seq_ms,
        "batched_ms": eager_ms,
        "graph_ms": graph_ms,
        "pairs_per_s_sequential": useful / (seq_ms * 1e-3),
        "pairs_per_s_graph": useful / (graph_ms * 1e-3),
        "speedup_graph_vs_sequential": seq_ms / graph_ms,
        "speedup_batched_vs_sequential": seq_ms / eager_ms,
        "autotuned": chosen,
        "step": "F + E for all 64 systems (x convert, pair kernel, FP64 exclusion energies, per-system sums, f reduce)",
    }


def bench_forward_backward(stream, reps=20, ewald="analytical"):
    """The 24,000-atom, 16-type force-matching system: forward forces vs the
    adjoint returning dL/dx, dL/dq, dL/dC6, dL/dC12.  The analytical Ewald
    correction is the one to differentiate: its force is a smooth function of
    r, while the table's is piecewise linear."""
    import numpy as np
    import torch

    from gmxapi_amd import nbx, systems

    s = systems.config(4)
    ctx = nbx.context_for(s, ewald_correction=ewald)
    ctx.set_stream(stream.cuda_stream)
    dev = torch.device("cuda:0")
    x = torch.tensor(s.x, dtype=torch.float32, device=dev).contiguous()
    f = torch.empty_like(x)
    g = torch.tensor(np.random.default_rng(1).normal(size=s.x.shape), dtype=torch.float32, device=dev)
    gx = torch.empty_like(x)
    gq = torch.empty(s.natoms, dtype=torch.float64, device=dev)
    g6 = torch.empty((s.ntypes, s.ntypes), dtype=torch.float64, device=dev)
    g12 = torch.empty_like(g6)
    useful = ctx.count_pairs(s.rcoulomb) - excluded_pairs_within(s, s.rcoulomb)

    def fwd():
        ctx.compute_device(x.data_ptr(), f.data_ptr(), nbx.COMPUTE_FORCES)

    def bwd():
        ctx.compute_adjoint_device(x.data_ptr(), g.data_ptr(), gx.data_ptr(), gq.data_ptr(), g6.data_ptr(),
                                   g12.data_ptr())

    fwd_ms = _time(fwd, stream, reps)
    bwd_ms = _time(bwd, stream, reps)
    return {
        "workload": f"force matching: {s.natoms} atoms, 10-atom chains, 16 LJ types, {s.box[0]} nm box, "
                    f"rc {s.rcoulomb} nm, rlist {s.rlist} nm, {ewald} Ewald correction",
        "useful_pairs": int(useful),
        "fwd_ms": fwd_ms,
        "bwd_ms": bwd_ms,
        "fwd_plus_bwd_over_fwd": (fwd_ms + bwd_ms) / fwd_ms,
        "pairs_per_s_fwd": useful / (fwd_ms * 1e-3),
        "pairs_per_s_fwd_bwd": useful / ((fwd_ms + bwd_ms) * 1e-3),
        "bwd_outputs": "dL/dx (N x 3), dL/dq (N, FP64), dL/dC6 and dL/dC12 (16 x 16, FP64)",
    }


def bench_prune(stream, reps=20, parts=4):
    """Dynamic pruning on the 288k water box searched with a 0.1 nm buffer
    (rlist 1.1 nm) and pruned back to the 1.0 nm cut-off (static
    coordinates): the search (with the kernel's work layout); the prune inside
    the work layout (nbx_prune_rolling: whole layout, and one of `parts`
    rolling parts); the F+E+shift step on the layout as searched and as
    pruned; and the step loop with one rolling part pruned before every step
    (the prune amortised as in an MD loop with rolling pruning).  The full
    re-listing prune (nbx_prune_pairlist, which rebuilds the layout) is
    timed for comparison."""
    import numpy as np
    import torch

    from gmxapi_amd import nbx, systems

    s = systems.water_box(96000, 14.2, seed=systems.WATER_SEED + 1, rc=1.0, rlist=1.1)
    ctx = nbx.context_for(s)
    ctx.set_stream(stream.cuda_stream)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(s.x.astype(np.float32).reshape(-1)).to(dev)
    f = torch.empty_like(x)
    flags = nbx.COMPUTE_FORCES | nbx.COMPUTE_ENERGY | nbx.COMPUTE_SHIFT_FORCES

    def step():
        ctx.compute_device(x.data_ptr(), f.data_ptr(), flags)

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    ctx.construct_pairlist()
    ev[1].record(stream)
    torch.cuda.synchronize()
    search_ms = ev[0].elapsed_time(ev[1])
    outer_cj = ctx.info()["num_j_entries"]
    outer_ms = _time(step, stream, reps)
    tiles_outer = int(ctx.work_groups()["chunks"][:, 3].sum())
    prune_all_ms = _time(lambda: ctx.prune_rolling(s.rcoulomb), stream, reps)
    tiles_inner = int(ctx.work_groups()["chunks"][:, 3].sum())
    inner_ms = _time(step, stream, reps)
    part = [0]

    def prune_part():
        ctx.prune_rolling(s.rcoulomb, part[0] % parts, parts)
        part[0] += 1

    prune_part_ms = _time(prune_part, stream, reps)

    def rolling_step():
        prune_part()
        step()

    rolling_ms = _time(rolling_step, stream, reps)
    # the re-listing prune (the list itself pruned, layout rebuilt)
    ev[0].record(stream)
    ctx.prune_pairlist(s.rcoulomb)
    ev[1].record(stream)
    torch.cuda.synchronize()
    relist_ms = ev[0].elapsed_time(ev[1])
    inner_cj = ctx.info()["num_j_entries"]
    return {
        "workload": "water-288k searched at rlist 1.1 nm, pruned to rc 1.0 nm (static coordinates)",
        "search_ms": search_ms,
        "cluster_pairs_searched": int(outer_cj),
        "cluster_pairs_pruned": int(inner_cj),
        "kernel_tiles_searched": tiles_outer,
        "kernel_tiles_pruned": tiles_inner,
        "prune_layout_ms": prune_all_ms,
        "prune_layout_part_ms": prune_part_ms,
        "rolling_parts": parts,
        "step_ms_searched_list": outer_ms,
        "step_ms_pruned_layout": inner_ms,
        "step_ms_rolling_prune_amortised": rolling_ms,
        "prune_relist_ms": relist_ms,
        "note": "prune_layout: nbx_prune_rolling over the whole layout; rolling step: one of the parts pruned "
                "before each step; prune_relist (nbx_prune_pairlist) includes rebuilding the work layout",
    }


def bench_flavours(s, stream, reps=10):
    """The other LJ flavours on the same 288k water box: the force and
    potential switches (rvdw_switch 0.8 nm) and LJ-PME's real-space part
    (geometric grid rule), each with energies, shift forces and virial: the
    step and its pair kernel (HIP events recorded by the launch)."""
    import numpy as np
    import torch

    from gmxapi_amd import nbx

    dev = torch.device("cuda:0")
    x = torch.from_numpy(s.x.astype(np.float32).reshape(-1)).to(dev)
    f = torch.empty_like(x)
    flags = nbx.COMPUTE_FORCES | nbx.COMPUTE_ENERGY | nbx.COMPUTE_SHIFT_FORCES | nbx.COMPUTE_VIRIAL
    out = {}
    for name, opts in (("potential-shift", {}),
                       ("force-switch", {"vdw_modifier": "force-switch", "rvdw_switch": 0.8}),
                       ("potential-switch", {"vdw_modifier": "potential-switch", "rvdw_switch": 0.8}),
                       ("lj-pme-geometric", {"vdw_type": "pme"})):
        ctx = nbx.context_for(s, **opts)
        ctx.set_stream(stream.cuda_stream)

        def step():
            ctx.compute_device(x.data_ptr(), f.data_ptr(), flags)

        step_ms = _time(step, stream, reps)
        ctx.set_kernel_timing(reps)
        for _ in range(reps):
            step()
        kernel_ms, _ = ctx.kernel_timing()
        ctx.set_kernel_timing(0)
        out[name] = {"step_ms": step_ms, "kernel_ms": kernel_ms}
        ctx.close()
    base = out["potential-shift"]["kernel_ms"]
    for v in out.values():
        v["kernel_vs_potential_shift"] = v["kernel_ms"] / base
    out["workload"] = "water-288k, F+E+shift+virial step per LJ flavour (table Ewald)"
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="water288k", choices=["water288k", "water9k", "chains150k"])
    ap.add_argument("--ewald", default="table", choices=["table", "analytical"],
                    help="Ewald correction: the reference's spline table or the analytical form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--search-reps", type=int, default=5,
                    help="grid + search rebuilds timed for grid_ms / search_ms (median)")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and forward+backward parts")
    ap.add_argument("--profile-copy", action="store_true",
                    help="also run one 64 MiB device-to-device copy: the reference that scales rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a batched ensemble would "
              "shard its replicas across GPUs, a spatial decomposition the bricks of one system). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gmxapi_amd import nbx, systems

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    index = {"water9k": 0, "water288k": 1, "chains150k": 2}[args.workload]
    t0 = time.perf_counter()
    s = systems.config(index)
    log(f"[bench] workload {args.workload}: {s.natoms} atoms, box {s.box[0]} nm, rc {s.rcoulomb}, "
        f"rlist {s.rlist} (generated in {time.perf_counter() - t0:.1f} s)")

    dev = torch.device("cuda:0")
    # a stream of our own (default priority): the library and the events
    # below run on it, so the events bracket exactly the library's work
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = nbx.Context(rlist=s.rlist, rcoulomb=s.rcoulomb, ewald_correction=args.ewald, device=0)
    ctx.set_stream(stream.cuda_stream)
    m6, m12 = nbx.nblib_c6c12_matrices(s.c6, s.c12)
    ctx.set_nonbonded_parameters(m6, m12)
    ctx.set_topology(s.types, s.charges, s.excl_off, s.excl_idx)
    x_d = torch.from_numpy(s.x.astype(np.float32).reshape(-1)).to(dev)
    f_d = torch.empty_like(x_d)
    torch.cuda.synchronize()

    # grid + search (the reference's nbnxn_put_on_grid + constructPairlist)
    # the first build sizes the buffers; the rebuilds that follow are what a
    # simulation pays every list update (median of --search-reps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ctx.put_on_grid(s.box, x_device_ptr=x_d.data_ptr())
    ctx.construct_pairlist()
    torch.cuda.synchronize()
    grid_t, search_t = [], []
    for _ in range(max(1, args.search_reps)):
        ev[0].record(stream)
        ctx.put_on_grid(s.box, x_device_ptr=x_d.data_ptr())
        ev[1].record(stream)
        ctx.construct_pairlist()
        ev[2].record(stream)
        torch.cuda.synchronize()
        grid_t.append(ev[0].elapsed_time(ev[1]))
        search_t.append(ev[1].elapsed_time(ev[2]))
    grid_ms = sorted(grid_t)[len(grid_t) // 2]
    search_ms = sorted(search_t)[len(search_t) // 2]
    info = ctx.info()
    computed_pairs = 16 * info["num_j_entries"]
    within_rc = ctx.count_pairs(s.rcoulomb)
    excl = excluded_pairs_water(s, s.rcoulomb) if args.workload.startswith("water") else \
        excluded_pairs_within(s, s.rcoulomb)
    useful = within_rc - excl
    log(f"[bench] grid {grid_ms:.3f} ms, search {search_ms:.3f} ms; clusters {info['num_clusters']}, "
        f"i-entries {info['num_i_entries']}, cluster pairs {info['num_j_entries']}; "
        f"useful pairs {useful} (estimate {s.useful_pairs_estimate():.0f}), computed {computed_pairs}")

    flags = nbx.COMPUTE_FORCES | nbx.COMPUTE_ENERGY | nbx.COMPUTE_SHIFT_FORCES | nbx.COMPUTE_VIRIAL
    if args.profile_copy:
        src = torch.ones(16 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def step():
        ctx.compute_device(x_d.data_ptr(), f_d.data_ptr(), flags)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        step()
    stop.record(stream)
    torch.cuda.synchronize()
    ms_per_step = start.elapsed_time(stop) / args.steps
    # the cluster-pair kernel's duration: the same steps again, each launch
    # carrying HIP events on the library's stream (recorded by the launch
    # itself).  Kept out of the timed steps above: the event packets leave
    # the GPU idle for about 11 us around every launch.
    ctx.set_kernel_timing(args.steps)
    for _ in range(args.steps):
        step()
    kernel_ms, launches = ctx.kernel_timing()
    assert launches == args.steps, launches

    # the forces-only kernel (no energies, no shift forces), same way
    ctx.set_kernel_timing(args.steps)
    for _ in range(args.steps):
        ctx.dispatch_kernel(nbx.COMPUTE_FORCES)
    kernel_f_ms, _ = ctx.kernel_timing()
    ctx.set_kernel_timing(0)
    step()
    torch.cuda.synchronize()
    vlj, vc = ctx.energies()

    achieved = FLOPS_PER_PAIR_VF * computed_pairs / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": useful / (ms_per_step * 1e-3),
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": {"water288k": "water-288k: 96,000 rigid 3-site waters, 14.2 nm cubic box",
                         "water9k": "water-9k: 3,000 rigid 3-site waters, 4.48 nm cubic box",
                         "chains150k": "chains-150k: 15,000 10-atom chains, 20 LJ types, 11.45 nm box"}[
                args.workload] + f", rc {s.rcoulomb} nm, rlist {s.rlist} nm, Ewald rtol 1e-5 ({args.ewald} "
                "correction), LJ potential shift; step = x convert + F/E/shift-force kernel + FP64 self/exclusion "
                "energies + f reduce + FP64 virial",
            "natoms": s.natoms,
            "useful_pairs": useful,
            "computed_pairs": computed_pairs,
            "list_efficiency": useful / computed_pairs,
            "kernel_group_size": ctx.work_groups()["group_size"],
            "kernel_ms_fe": kernel_ms,
            "kernel_ms_f": kernel_f_ms,
            "grid_ms": grid_ms,
            "search_ms": search_ms,
            "energies_kj_mol": {"lj": vlj, "coulomb_real_space": vc},
        },
        "roofline": {
            "bound": "valu",
            "unit": "TFLOP/s",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic(args.workload, args.ewald),
            "kernel": "nonbondedKernel<%s, energy, shift>" % ("EwaldTable" if args.ewald == "table"
                                                               else "EwaldAnalytical"),
            "kernel_ms": kernel_ms,
            "compute_unit": "FP32 vector ALU (pair math is not a dense contraction; peak is MI355X FP32 157.3 TFLOP/s)",
            "flops_per_pair": FLOPS_PER_PAIR_VF,
            "pairs_per_launch": computed_pairs,
        },
    }
    if not args.no_extras:
        t0 = time.perf_counter()
        out["batched"] = bench_ensemble(args.ewald, stream)
        out["fwd_bwd"] = bench_forward_backward(stream)
        out["prune"] = bench_prune(stream)
        out["lj_flavours"] = bench_flavours(s, stream)
        log(f"[bench] batched and forward+backward parts in {time.perf_counter() - t0:.1f} s")
    if not args.no_cpu_baseline:
        t_step, t_setup, nrep, threads, model, kind = cpu_baseline(s)
        out["cpu_baseline"] = {
            "value": useful / t_step,
            "unit": "pair-interactions/s",
            "cores": threads,
            "kind": "port",
            "cpu": model,
            "ms_per_step": t_step * 1e3,
            "sample": f"full F+E steps (convert, kernel, reduce) of the same {s.natoms}-atom workload with the "
                      f"oracle build's {kind}, float, {threads} OpenMP threads with per-thread outputs reduced, "
                      f"median of {nrep}: {t_step * 1e3:.1f} ms per step (grid + search {t_setup:.2f} s, "
                      f"single-threaded, not included); the reference itself is not built here",
            "kernel": kind,
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

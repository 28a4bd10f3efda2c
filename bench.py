"""Benchmark: batched fixed-point FP32 nonbonded evaluations of the 256-box
water ensemble on one MI355X, against plain per-system launches, with the
error against the FP64 GPU oracle mode and a CPU baseline.

A step is one force+energy evaluation of all 256 systems: positions perturbed
by N(0, 0.002 nm) about the search positions, the nonbonded cluster-pair
kernel, fixed-point force conversion and FP64 energy / shift-force reduction.
The timed steps run as one hipGraph.  Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "gromacs-fmm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "Gpairs/s, batched fixed-point FP32 vs plain launch, at stated err vs FP64"
FP32_PEAK_TFLOPS = 157.3      # MI355X published FP32 vector peak (packed FP32)
# The same FP32 vector pipe without packed FP32 instructions, which this build
# does not use: one wave64 FMA per SIMD every 4 cycles.
FP32_NONPACKED_TFLOPS = 78.6
# HBM bytes per batched F+E launch of this workload (at most 48 pruned
# cluster pairs per work item below 1,500 atoms, 2 waves per block):
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a 1 GiB device copy) from
# two rocprofv3 --pmc runs of the current kernel,
# profiles/nonbonded_vf_ensemble256_counters.md.  Counters cannot be read
# inside the timed run, so the measured figure is carried here.
TRAFFIC_BYTES_PER_LAUNCH = 2.126e9
HBM_PEAK_GBS = 8000.0
FLOPS_PER_PAIR_VF = 107       # Ewald + LJ force+energy (src/gromacs/gmxlib/nrnb.cpp:90-103)
RC, RLIST_OUTER, RLIST_INNER, EWALD_RTOL, SIGMA = 1.0, 1.1, 1.04, 1e-5, 0.002
BUCKET_EDGES = [1500, 3500, 8000]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(systems, seconds, num_systems=16):
    """The 4xM SIMD CPU kernel (oracle/cpu4xm.cpp, AVX2+FMA, OpenMP over all
    threads the host grants) timed on a bounded sample of the ensemble: the
    same F+E evaluation, with its 4x8 list pruned at the same inner radius."""
    import nbnxm_amd as nb
    import oracle

    ic = nb.interaction_const(RC, EWALD_RTOL)
    order = np.random.default_rng(7).permutation(len(systems))[:num_systems]
    kernels, pairs, atoms_total = [], 0, 0
    for k in order:
        s = systems[int(k)]
        search, plist, atoms = nb.build(s, RLIST_OUTER, nb.VDW_CUT_COMB_GEOM)
        pairs += oracle.count_pairs(plist, atoms, RC)
        atoms_total += s.n
        kernels.append(oracle.Cpu4xm(plist, atoms, ic, RLIST_INNER))
    for kern in kernels:  # untimed pre-iteration (bench_setup.cpp:355-390)
        kern.compute(True)
    t0 = time.perf_counter()
    evals = 0
    while True:
        for kern in kernels:
            kern.compute(True)
        evals += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": pairs * evals / dt / 1e9, "unit": "Gpairs/s", "cores": kernels[0].threads, "kind": "port",
            "sample": f"{len(kernels)} ensemble boxes ({atoms_total} atoms, {pairs} useful pairs), {evals} F+E "
                      f"evaluations of each in {dt:.1f} s; 4xM AVX2+FMA kernel, {kernels[0].threads} OpenMP threads "
                      f"on {cpu_model or platform.processor()} (nproc {os.cpu_count()})"}


# Reference flop weights per pair (src/gromacs/gmxlib/nrnb.cpp:90-103): F, F+E.
FLAVOUR_WEIGHTS = {
    "ewald_analytical": (66, 107),
    "ewald_tabulated": (41, 59),
    "reaction_field": (38, 54),
    "ewald_twin": (66, 107),
    "ewald_forceswitch": (66 + 12, 107 + 22),
    "ewald_potswitch": (66 + 27, 107 + 20),
    "ewald_ljpme": (66 + 36, 107 + 33),
}
FLAVOUR_OPTIONS = {
    "ewald_analytical": {},
    "ewald_tabulated": dict(tabulated=True),
    "reaction_field": dict(coulomb="rf"),
    "ewald_twin": dict(rvdw=0.9),
    "ewald_forceswitch": dict(vdw_modifier="forceswitch", rvdw_switch=0.8),
    "ewald_potswitch": dict(vdw_modifier="potswitch", rvdw_switch=0.8),
    "ewald_ljpme": dict(lj_pme=True, ewald_rtol_lj=1e-3),
}


def water96k_setup(nb, vdw):
    s = nb.systems.config_water_96k()
    search, plist, atoms = nb.build(s, RLIST_OUTER, vdw)
    return s, search, plist, atoms


def fresh_list_table(launches):
    """One JSON line per (F / F+E, path): the first launch on a fresh list of
    the 96,000-atom water box (list 1.1 nm, fixed point), either the combined
    prune and interaction kernel (no dynamic pruning: the list is pruned to
    rlistOuter inside the interaction kernel, then the entries are sorted) or
    the prune kernel followed by the interaction kernel on the pruned, sorted
    list (dynamic pruning with rlistInner = rlistOuter).  Wall clock per
    fresh launch up to its completion (upload of the list and the force copy
    excluded) over `launches` uploads."""
    import nbnxm_amd as nb

    s, search, plist, atoms = water96k_setup(nb, nb.VDW_CUT_COMB_GEOM)
    ic = nb.interaction_const(RC, EWALD_RTOL)
    for energy in (False, True):
        for combined in (True, False):
            task = nb.GpuTask(ic, atoms, RLIST_OUTER, None, nb.ACCUM_FIXED64, dynamic_pruning=not combined)
            times = []
            for k in range(launches + 2):
                task.init_pairlist(plist)
                t0 = time.perf_counter()
                task.launch(energy, False)
                task.wait()
                if k >= 2:
                    times.append(time.perf_counter() - t0)
            print(json.dumps({"fresh_list": "combined prune kernel" if combined else "prune kernel + interaction kernel",
                              "energies": energy, "atoms": s.n, "ms_per_fresh_launch": 1e3 * float(np.median(times)),
                              "launches": launches}), flush=True)
            del task


def placement_table(launches):
    """One JSON line per (box, flavour, F / F+E, i-cluster placement): water
    boxes of 96,000, 384,000 and 3,000,000 atoms with unsplit lists, the
    i-cluster data of the list walk in registers (0) and in LDS (1), wall
    clock over `launches` launches after warm-up, both placements twice in
    alternation (the minimum is reported)."""
    import nbnxm_amd as nb

    names = ["ewald_analytical", "reaction_field", "ewald_tabulated", "ewald_forceswitch", "ewald_twin", "ewald_ljpme"]
    for nmol, seed in ((32000, 2), (128000, 3), (1000000, 5)):
        s = nb.systems.water_box(nmol, None, seed=seed)
        lists = {}
        for name in names:
            vdw = nb.VDW_CUT_TYPE_TABLE if name == "ewald_ljpme" else nb.VDW_CUT_COMB_GEOM
            if vdw not in lists:
                lists[vdw] = nb.build(s, RLIST_OUTER, vdw)[1:]
            plist, atoms = lists[vdw]
            kw = FLAVOUR_OPTIONS[name]
            ic = nb.interaction_const(RC, EWALD_RTOL) if not kw else nb.interaction_const_ex(RC, ewald_rtol=EWALD_RTOL, **kw)
            for energy in (False, True):
                ms = {0: [], 1: []}
                for mode in (0, 1, 0, 1):
                    task = nb.GpuTask(ic, atoms, RLIST_OUTER, RLIST_INNER, nb.ACCUM_FIXED64)
                    task.set_i_cluster_in_lds(mode)
                    task.init_pairlist(plist)
                    for _ in range(3):
                        task.launch(energy, False)
                    task.wait()
                    # Launches only: no force copy in the timed loop.
                    t0 = time.perf_counter()
                    for _ in range(launches):
                        task.launch(energy, False)
                    task.wait()
                    ms[mode].append((time.perf_counter() - t0) / launches * 1e3)
                    del task
                print(json.dumps({"placement": name, "atoms": s.n, "entries": int(len(plist.sci)), "energies": energy,
                                  "ms_registers": min(ms[0]), "ms_lds": min(ms[1]),
                                  "lj": "type table" if vdw == nb.VDW_CUT_TYPE_TABLE else "geometric"}), flush=True)


def flavour_table(launches):
    """One JSON line per interaction flavour, list and F / F+E: one evaluation
    of the 96,000-atom water box (rc 1.0 nm, list 1.1 nm, fixed point) per
    launch, wall clock over `launches` launches after warm-up.  Two lists: the
    unsplit one (2,539 entries) and the one split toward min_ci_balanced
    entries, as a search handed the library's `gpu_min_ci_balanced` builds it
    (target max(36, cluster pairs / min_ci_balanced), the simplified
    get_nsubpair_target of the MD step and the calculator)."""
    import nbnxm_amd as nb

    geom = water96k_setup(nb, nb.VDW_CUT_COMB_GEOM)
    table = water96k_setup(nb, nb.VDW_CUT_TYPE_TABLE)
    useful = geom[1].count_pairs(RC)
    probe = nb.GpuTask(nb.interaction_const(RC, EWALD_RTOL), geom[3], RLIST_OUTER, RLIST_INNER, nb.ACCUM_FIXED64)
    min_ci = probe.min_ci_balanced()
    del probe
    lists = {}
    for vdw, setup in ((nb.VDW_CUT_COMB_GEOM, geom), (nb.VDW_CUT_TYPE_TABLE, table)):
        target = max(36, -(-setup[2].num_cluster_pairs() // min_ci))
        lists[vdw] = {"unsplit": setup[2:], "split": nb.build(setup[0], RLIST_OUTER, vdw, target)[1:]}
    for name, kw in FLAVOUR_OPTIONS.items():
        for split in ("unsplit", "split"):
            plist, atoms = lists[nb.VDW_CUT_TYPE_TABLE if kw.get("lj_pme") else nb.VDW_CUT_COMB_GEOM][split]
            ic = nb.interaction_const(RC, EWALD_RTOL) if not kw else nb.interaction_const_ex(RC, ewald_rtol=EWALD_RTOL, **kw)
            for energy in (False, True):
                task = nb.GpuTask(ic, atoms, RLIST_OUTER, RLIST_INNER, nb.ACCUM_FIXED64)
                task.init_pairlist(plist)
                for _ in range(3):
                    task.launch(energy, False)
                task.wait()
                # Launches only: no force copy in the timed loop.
                t0 = time.perf_counter()
                for _ in range(launches):
                    task.launch(energy, False)
                task.wait()
                ms = (time.perf_counter() - t0) / launches * 1e3
                flops = FLAVOUR_WEIGHTS[name][1 if energy else 0]
                print(json.dumps({"flavour": name, "list": split, "entries": int(len(plist.sci)), "energies": energy, "ms_per_launch": ms,
                                  "gpairs_per_s": useful / ms / 1e6, "useful_pairs": useful,
                                  "algorithmic_tflops": useful * flops / (ms * 1e-3) / 1e12, "flops_per_pair": flops,
                                  "lj": "type table" if kw.get("lj_pme") else "geometric"}), flush=True)
                del task


def md_step_table(steps):
    """The nonbonded part of an MD step (nbnxm_verlet) on the 96,000-atom water
    box: search every 100 steps on the host or on the device, rolling prune,
    x -> xq, kernel (energies every 100 steps), force reduction; coordinates
    held fixed, so only the nonbonded work is timed."""
    import nbnxm_amd as nb

    s = nb.systems.config_water_96k()
    params = np.stack([6.0 * s.c6c12[..., 0], 12.0 * s.c6c12[..., 1]], axis=-1)
    ic = nb.interaction_const(RC, EWALD_RTOL)
    d_x = nb.DeviceArray(s.x.astype(np.float32))
    d_f = nb.DeviceArray.zeros(d_x.shape)
    stream = 0
    for gpu_search in (False, True):
        ver = nb.NonbondedVerlet(s.types, params, s.q, ic, nstlist=100, rlist_outer=RLIST_OUTER, rlist_inner=1.02,
                                 nstlist_prune=10, excl_ranges=s.excl_ranges, excl_elements=s.excl_elements,
                                 gpu_search=gpu_search)
        ver.step(0, s.box, d_x.ptr, d_f.ptr, stream)
        search_ms, other = [], []
        for step in range(1, steps + 1):
            t0 = time.perf_counter()
            ver.step(step, s.box, d_x.ptr, d_f.ptr, stream, compute_energy=(step % 100 == 0))
            dt = (time.perf_counter() - t0) * 1e3
            (search_ms if step % 100 == 0 else other).append(dt)
        st = ver.stats()
        total = sum(search_ms) + sum(other)
        print(json.dumps({"md_step": "nonbonded part", "search": "device" if gpu_search else "host",
                          "atoms": s.n, "steps": steps, "ms_per_step": total / steps,
                          "search_step_ms": float(np.mean(search_ms)) if search_ms else None,
                          "other_step_ms": float(np.mean(other)), "nstlist": 100, "nstlist_prune": 10,
                          "stats": st}), flush=True)
        del ver


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--split", type=str, default="",
                    help="fixed schedule instead of the autotuner: max pruned cluster pairs per work item for each "
                         "size bucket, comma-separated (0 = no split), e.g. 512,512,512,0")
    ap.add_argument("--wpb", type=int, default=4, help="waves per block with --split")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (counter calibration for HBM bytes) and exit")
    ap.add_argument("--flavours", type=int, default=0, metavar="LAUNCHES",
                    help="only print the per-flavour kernel table for the 96k-atom water box and exit")
    ap.add_argument("--placement", type=int, default=0, metavar="LAUNCHES",
                    help="only print the i-cluster placement table (registers vs LDS) on 96k-3M-atom water boxes and exit")
    ap.add_argument("--md-step", type=int, default=0, metavar="STEPS",
                    help="only time the nonbonded part of an MD step (96k-atom water) and exit")
    ap.add_argument("--fresh-list", type=int, default=0, metavar="LAUNCHES",
                    help="only time fresh-list launches, combined prune kernel vs prune + interaction (96k-atom water), and exit")
    args = ap.parse_args()
    if args.flavours:
        flavour_table(args.flavours)
        return
    if args.md_step:
        md_step_table(args.md_step)
        return
    if args.fresh_list:
        fresh_list_table(args.fresh_list)
        return
    if args.placement:
        placement_table(args.placement)
        return
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the ensemble shards by "
              "whole systems, so one GPU's throughput is the per-GPU figure", file=sys.stderr)
        sys.exit(2)

    import nbnxm_amd as nb

    t0 = time.perf_counter()
    systems = nb.systems.config_ensemble(256)
    n_atoms = sum(s.n for s in systems)
    log(f"generated 256 boxes, {n_atoms} atoms in {time.perf_counter() - t0:.1f}s")
    # Grids and lists of all boxes on the device (one pass), timed on its own.
    gs = nb.GpuPairSearch()
    gs.build(systems, RLIST_OUTER, 0)
    t0 = time.perf_counter()
    gs.build(systems, RLIST_OUTER, 0)
    gpu_list_ms = 1e3 * (time.perf_counter() - t0)
    del gs
    t0 = time.perf_counter()
    ens = nb.Ensemble(systems, RC, EWALD_RTOL, RLIST_OUTER, RLIST_INNER, nb.VDW_CUT_COMB_GEOM, nb.ACCUM_FIXED64,
                      gpu_search=True)
    setup_s = time.perf_counter() - t0
    useful, per_sys = ens.count_pairs()
    stats = ens.stats()
    log(f"device pair search {gpu_list_ms:.1f} ms; ensemble set up in {setup_s:.1f}s: {stats}, useful pairs {useful}")

    tuned = None
    if args.split:
        split = [int(v) for v in args.split.split(",")]
        if len(split) != len(BUCKET_EDGES) + 1:
            sys.exit(f"--split needs {len(BUCKET_EDGES) + 1} values")
        ens.set_split(BUCKET_EDGES, split, args.wpb)
        tuned = {"bucket_edges_atoms": BUCKET_EDGES, "max_pairs_per_item": split, "waves_per_block": args.wpb,
                 "i_cluster_in_lds": ens.launch_config()[1], "fixed": True}
    elif not args.no_autotune:
        t0 = time.perf_counter()
        split, wpb = ens.autotune(BUCKET_EDGES, compute_energy=True)
        tuned = {"bucket_edges_atoms": BUCKET_EDGES, "max_pairs_per_item": split, "waves_per_block": wpb,
                 "i_cluster_in_lds": ens.launch_config()[1]}
        log(f"autotuned in {time.perf_counter() - t0:.1f}s: {tuned}")
    stats = ens.stats()

    # Warmup, then the timed graph of K steps.
    ens.run(args.warmup, nb.MODE_BATCHED, SIGMA, 11, True, True, args.steps)
    ens.sync()
    ens.run(args.steps, nb.MODE_BATCHED, SIGMA, 11, True, True, args.steps)  # builds and warms the K-step graph
    ens.sync()
    t0 = time.perf_counter()
    ens.run(args.steps, nb.MODE_BATCHED, SIGMA, 11, True, True, args.steps)
    ens.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = useful * args.steps / wall / 1e9
    log(f"batched: {ms_per_step:.3f} ms/step, {value:.1f} Gpairs/s")

    # Nonbonded kernel duration from HIP events on the ensemble's stream.
    ev_total, kern_ms = ens.time(min(args.steps, 50), nb.MODE_BATCHED, SIGMA, 11, True, False, True)
    achieved_tflops = useful * FLOPS_PER_PAIR_VF / (kern_ms * 1e-3) / 1e12

    # Plain per-system launches (same kernels, one launch set per system).
    ens.run(1, nb.MODE_PLAIN, SIGMA, 11, True)
    ens.sync()
    t0 = time.perf_counter()
    ens.run(args.plain_steps, nb.MODE_PLAIN, SIGMA, 11, True)
    ens.sync()
    plain_wall = time.perf_counter() - t0
    plain_value = useful * args.plain_steps / plain_wall / 1e9
    log(f"plain: {1e3 * plain_wall / args.plain_steps:.3f} ms/step, {plain_value:.1f} Gpairs/s")

    # Error against the FP64 GPU oracle mode at the search positions.
    ens.run(1, nb.MODE_BATCHED, 0.0, 11, True)
    ens.sync()
    ref = nb.Ensemble(systems, RC, EWALD_RTOL, RLIST_OUTER, RLIST_INNER, nb.VDW_CUT_COMB_GEOM, nb.ACCUM_FP64)
    ref.run(1, nb.MODE_BATCHED, 0.0, 11, True)
    ref.sync()
    num = den = 0.0
    fmax_err = 0.0
    for k in range(len(systems)):
        f = ens.forces(k)
        fr = ref.forces(k)
        num += float(((f - fr) ** 2).sum())
        den += float((fr ** 2).sum())
        fmax_err = max(fmax_err, float(np.abs(f - fr).max()))
    e32, _ = ens.energies()
    e64, _ = ref.energies()
    natoms = np.array([s.n for s in systems], dtype=np.float64)
    err = {
        "force_rel_rms": float(np.sqrt(num / den)),
        "force_max_abs": fmax_err,
        # Totals over the ensemble, and the worst system per atom (a box's LJ
        # total can be near zero, so a per-box relative error is not meaningful).
        "energy_lj_total_rel": float(abs(e32[:, 0].sum() - e64[:, 0].sum()) / abs(e64[:, 0].sum())),
        "energy_coul_total_rel": float(abs(e32[:, 1].sum() - e64[:, 1].sum()) / abs(e64[:, 1].sum())),
        "energy_lj_max_abs_per_atom": float(np.max(np.abs(e32[:, 0] - e64[:, 0]) / natoms)),
        "energy_coul_max_abs_per_atom": float(np.max(np.abs(e32[:, 1] - e64[:, 1]) / natoms)),
    }
    log(f"error vs FP64: {err}")
    del ref

    cpu = None
    if not args.skip_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(systems, args.cpu_seconds)
        log(f"cpu baseline: {cpu} in {time.perf_counter() - t0:.1f}s")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "ensemble of 256 rigid 3-site water boxes, 648-24000 atoms log-uniform, 33.4 molecules/nm^3, "
                        "Ewald (analytical) + LJ cut-off, rc 1.0 nm, rlist 1.1 nm, inner prune 1.04 nm, "
                        "one force+energy evaluation of all boxes per step, hipGraph over the timed steps",
            "systems": len(systems),
            "atoms": n_atoms,
            "useful_pairs_per_step": int(useful),
            "accumulation": "int64 fixed point (2^-32), FP64 energy/shift-force reduction",
            "perturbation_nm": SIGMA,
            "list_stats": stats,
            "device_pair_search_ms": gpu_list_ms,
            "ensemble_setup_s": setup_s,
            "autotune": tuned,
            "plain_launch_gpairs_per_s": plain_value,
            "plain_ms_per_step": 1e3 * plain_wall / args.plain_steps,
            "batched_over_plain": value / plain_value,
            "error_vs_fp64": err,
            "event_ms_per_step": ev_total / min(args.steps, 50),
            "nonbonded_kernel_ms": kern_ms,
        },
        "roofline": {
            "bound": "fp32-valu",
            "achieved": achieved_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS,
            "traffic": TRAFFIC_BYTES_PER_LAUNCH,
            "nonpacked_peak": FP32_NONPACKED_TFLOPS,
            "frac_of_nonpacked": achieved_tflops / FP32_NONPACKED_TFLOPS,
            "note": "compute roofline of the FP32 vector (VALU) pipe: the kernel issues no MFMA; "
                    "algorithmic flops = useful pairs x 107 (Ewald+LJ F+E weight); the build uses no "
                    "packed FP32 instructions, so 78.6 TFLOP/s (non-packed) is its ceiling",
        },
        "cpu_baseline": cpu,
        "host": platform.node(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: subhalo-particle binding-energy evaluations per second.

One step = one RefineParticles pass over the whole batch: every subhalo is
unbound to convergence (batched mixed-precision form, inputs resident in
HBM).  Default workload: 5,000 power-law subhalos (N_p 20..2e5, 20%
contaminants) in a periodic 100 Mpc/h box.

Prints one JSON line.  Besides the headline (batched) throughput it reports
the plain per-subhalo GPU form on the same inputs (same kernels, one subhalo
at a time; the bound sets must be identical), the roofline of the dominant
kernel (the potential walk: HIP-event timed launches, FLOPs from an
instrumented pass) and the CPU baseline (the oracle restatement of the
reference, OpenMP over subhalos, on a bounded sample).

    python bench.py [--steps K] [--warmup W] [--config c2|c1|c3|c4|c5]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "subhalo-particle energy evals/s: batched mixed-prec vs plain GPU, same bound sets"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
LANE_SLOTS_PER_S = 256 * 4 * 16 * 2.4e9
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "walk_c2_traffic.md")


def walk_traffic(config, launches):
    """HBM bytes per walk launch for the default workload, from the committed
    counter summary (FETCH_SIZE and WRITE_SIZE passes of rocprofv3, each
    scaled by its reading for a copy of known size; see that file)."""
    if config != "c2" or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def make_workload(name):
    from hbtplus_amd import synth
    if name == "c2":
        cat = synth.config2()
        desc = ("5,000 power-law subhalos (dN/dN_p ~ N_p^-1.9 on [20, 2e5], log-normal c, "
                "20% contaminants at 3 v_esc), periodic 100 Mpc/h box, eps=0.5, theta=0.45")
    elif name == "c1":
        cat = synth.config1(bound_mass_precision=0.995)
        desc = "single 24k-particle NFW subhalo with contaminants, eps=0.5, theta=0.45"
    elif name == "c3":
        cat = synth.config3()
        desc = "50,000 subhalos, N_p log-uniform on [20, 4096], 15% contaminants"
    elif name == "c4":
        cat = synth.config4()
        desc = "4M-particle NFW host with 200 nested satellites (N_p 1e3..1e5), eps=1"
    elif name == "c5":
        _, _, cat = synth.config5()
        desc = "200,000 power-law subhalos (config-2 law) of one rank's share"
    else:
        raise SystemExit(f"unknown config {name}")
    return cat, desc


def cpu_baseline(cat, budget_s=20.0, threads=None, with_reference=False):
    """The oracle (CPU restatement of the reference, OpenMP over subhalos) on
    a bounded sample: subhalos taken in a fixed random order until the
    estimated work fills about budget_s seconds.  with_reference: also time
    the reference itself (oracle/_ref, built from its sources where they are
    present) on the same sample and threads; the evaluation count is the
    oracle's (the two follow the same trajectory bit for bit)."""
    import oracle
    threads = threads or min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(0)
    order = rng.permutation(len(cat.subs))
    top = cat.subs["parent"] < 0
    what = "top-level subhalos"
    if int(top.sum()) < 16:
        # one or a few large nests (a host with its satellites): sample the
        # nested subhalos, each unbound on its own, leaving out lists of more
        # than a million particles (a single serial unbinding of minutes)
        top = cat.subs["count"] <= 1_000_000
        what = "subhalos (nested ones unbound on their own, lists above 1 M particles left out)"
    order = order[top[order]]
    # calibrate on a small sample, then grow to the budget
    take, elapsed, evals = 16, 0.0, 0
    while True:
        sub = cat.subset(order[:take])
        print(f"[bench] cpu baseline: restatement on {take} {what.split(' (')[0]}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        subs, _, _ = oracle.unbind_batch(sub, nthreads=threads)
        elapsed = time.perf_counter() - t0
        evals = int(subs["n_evals"].sum())
        if elapsed > budget_s / 4 or take >= len(order):
            break
        take = min(len(order), int(take * max(2.0, budget_s / 4 / max(elapsed, 1e-3))))
    sample = (f"{take} of {len(order)} {what} ({sub.n_particles} particles, {evals} evals), "
              f"OpenMP over nests, schedule(dynamic,1)")
    port = dict(value=evals / elapsed, unit="evals/s", cores=threads, kind="port",
                sample=f"{sample}; {elapsed:.1f} s")
    if not (with_reference and oracle.have_reference()):
        return port
    # the reference's own unbinding, compiled from its sources, on the same sample
    print(f"[bench] cpu baseline: reference on the same {take}", file=sys.stderr, flush=True)
    secs = oracle.ref_time_unbind(sub, threads)
    return dict(value=evals / secs, unit="evals/s", cores=threads, kind="reference",
                sample=f"{sample}; {secs:.1f} s (evaluation count from the port, same trajectory)",
                port=port)


def bench_mask(args):
    """MaskSubhalos over 1,000 host groups with overlapping source lists
    (about 8 M list entries): GPU entries/s against the oracle restatement
    (one thread: the reference masks each group serially, groups in parallel)."""
    import torch

    from hbtplus_amd import abi, synth
    from hbtplus_amd.device import GpuUnbinder
    cols, subs, grp = synth.masking_catalogue(seed=1006, ngroups=1000, pool=4000)
    n = len(cols["id"])
    U = GpuUnbinder(abi.make_params(), device=0)
    for _ in range(args.warmup):
        U.mask_subhalos(cols, subs, grp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, cnt = U.mask_subhalos(cols, subs, grp)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    res = {"metric": "MaskSubhalos list entries/s", "value": n / (ms / 1e3), "unit": "entries/s",
           "ms_per_call": round(ms, 3), "entries": n, "subhalos": int(len(subs)), "groups": 1000,
           "note": "wall time per call including the host-to-device copy of the columns"}
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        ro, rc = oracle.mask_subhalos(cols, subs, grp)
        res["cpu_port"] = {"ms": round((time.perf_counter() - t0) * 1e3, 1), "cores": 1,
                           "identical": bool(np.array_equal(rc, cnt) and all(
                               np.array_equal(out[o:o + c], ro[o:o + c]) for o, c in zip(subs["offset"], cnt)))}
    U.close()
    print(json.dumps(res), flush=True)


def bench_merge(args):
    """MergeSubhalos after one snapshot's unbinding: 300 host groups laid out
    with sinking satellites (about 2.3 M particles, scale 3 of the merging
    test catalogue).  The state is the GPU's own FAST unbinding of the
    catalogue; a call detects the traps, merges the trapped subhalos'
    lists into their hosts and unbinds the merged hosts again.  Columns and
    lists are resident on the device; the call's wall time is timed, against
    the reference's own SubhaloSnapshot_t::MergeSubhalos on the host cores."""
    import torch

    from hbtplus_amd import abi, synth
    from hbtplus_amd.device import GpuUnbinder
    import oracle

    cat, grp = synth.merging_catalogue(seed=1010, ngroups=300, scale=3.0)
    params = cat.params.copy()
    params["precision"] = abi.HBT_PREC_REPLAY if args.replay else abi.HBT_PREC_FAST
    U = GpuUnbinder(params, device=0)
    subs, idx, _ = U.unbind_batch(cat)
    lst, sin = oracle.merge_input(cat, subs, idx)
    hm = U.upload_merge(cat.cols, lst)
    print(f"[bench] merge: {len(sin)} subhalos, {cat.n_particles} particles, {len(lst)} list entries",
          file=sys.stderr, flush=True)
    out = None
    for _ in range(args.warmup):
        got, out, used = U.merge_device(hm, sin, grp, True, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        got, out, used = U.merge_device(hm, sin, grp, True, out)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    dev_ms, evals, passes = U.last_stats()
    away = int(((got["count"] == 1) & (sin["count"] > 1)).sum())
    trapped = int((got["sink_track_id"] != sin["sink_track_id"]).sum())
    res = {"metric": "MergeSubhalos subhalos/s", "value": len(sin) / (ms / 1e3), "unit": "subhalos/s",
           "ms_per_call": round(ms, 3), "subhalos": int(len(sin)), "particles": int(cat.n_particles),
           "list_entries": int(len(lst)), "newly_trapped": trapped, "merged_away": away,
           "reunbind": {"device_ms": round(dev_ms, 3), "evals": int(evals), "passes": int(passes)},
           "dtype": "f32+f64" if not args.replay else "replay",
           "note": "wall time per call; columns and lists resident on the device"}
    if not args.no_cpu and oracle.have_reference():
        threads = min(16, os.cpu_count() or 1)
        secs = oracle.ref_merge_subhalos(params, cat.cols, lst, sin, grp, True, time_threads=threads)
        res["cpu_baseline"] = {"kind": "reference", "ms_per_call": round(secs * 1e3, 1), "cores": threads,
                               "value": len(sin) / secs, "unit": "subhalos/s"}
    U.close()
    print(json.dumps(res), flush=True)


def bench_c5(args):
    """One rank's share: 200,000 subhalos given as source-ID lists into a
    shuffled snapshot (with a uniform background).  A step resolves every list
    entry on the device (hbt_id_to_index) and unbinds all subhalos in one
    batched pass (hbt_unbind_batch_indexed).  The plain form runs on a fixed
    random sample of the subhalos (the per-subhalo loop over all 200,000 takes
    minutes); its rate is compared with the batched rate on the same sample."""
    import torch

    from hbtplus_amd import abi, synth
    from hbtplus_amd.abi import SUB_DTYPE
    from hbtplus_amd.device import GpuUnbinder

    t0 = time.perf_counter()
    snap, ids, cat = synth.config5()
    gen_s = time.perf_counter() - t0
    params = cat.params.copy()
    params["precision"] = abi.HBT_PREC_FAST
    U = GpuUnbinder(params, device=0)
    h = U.upload_snapshot(snap, ids, cat.subs)
    torch.cuda.synchronize()
    print(f"[bench] c5: {len(cat.subs)} subhalos, {len(ids)} listed particles, snapshot "
          f"{len(snap['x'])} (generated in {gen_s:.0f} s)", file=sys.stderr, flush=True)

    def step():
        idx = U.resolve_ids(h)
        return U.unbind_indexed(h, idx)

    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(len(cat.subs), size=min(2000, len(cat.subs)), replace=False))
    sub = cat.subset(sample)
    hb = U.upload(sub)
    tuning = None
    if not args.no_autotune:
        # tuned on the sample (same size distribution), before the timed region
        t = U.autotune(hb, reps=1)
        tuning = {"walk_group": [int(v) for v in t["walk_group"]],
                  "bin_edge": [int(v) for v in t["bin_edge"]], "use_graph": int(t["use_graph"]),
                  "graph_passes": int(t["graph_passes"]), "xcd_run": int(t["xcd_run"])}
        print(f"[bench] autotuned: {tuning}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    subs = out[0].cpu().numpy().view(SUB_DTYPE)
    evals = int(subs["n_evals"].sum())
    value = evals / (ms_per_step / 1e3)
    t0 = time.perf_counter()
    U.resolve_ids(h)
    torch.cuda.synchronize()
    resolve_ms = (time.perf_counter() - t0) * 1e3
    print(f"[bench] c5 batched: {ms_per_step:.1f} ms/step ({resolve_ms:.1f} ms ID resolution), "
          f"{evals} evals, {value:.4g} evals/s", file=sys.stderr, flush=True)
    result = {"metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
              "config": {"workload": "200,000 power-law subhalos (config-2 law) as source-ID lists "
                                     "into a shuffled snapshot with a 10% uniform background",
                         "subhalos": int(len(cat.subs)), "listed_particles": int(len(ids)),
                         "snapshot_particles": int(len(snap["x"])), "evals_per_step": evals,
                         "id_resolution_ms": round(resolve_ms, 2),
                         "precision": "FAST (fp32+Kahan walk, fp64 frames)", "tuning": tuning}}
    if not args.no_plain:
        U.unbind_device(hb)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ob = U.unbind_device(hb)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        op = U.unbind_device(hb, plain=True)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        ev = int(ob[0].cpu().numpy().view(SUB_DTYPE)["n_evals"].sum())
        same = all(np.array_equal(a.cpu().numpy().view(np.uint8), b.cpu().numpy().view(np.uint8))
                   for a, b in zip(ob, op))
        result["plain_sample"] = {"subhalos": int(len(sample)), "evals": ev,
                                  "plain_evals_per_s": ev / tp, "batched_evals_per_s": ev / tb,
                                  "plain_ms": round(tp * 1e3, 1), "identical_outputs": bool(same)}
        result["batched_over_plain"] = round(value / (ev / tp), 2)
        print(f"[bench] c5 plain sample: {tp * 1e3:.0f} ms for {len(sample)} subhalos", file=sys.stderr,
              flush=True)
    U.close()
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--replay", action="store_true", help="REPLAY precision instead of FAST")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default execution tuning")
    ap.add_argument("--cpu-only", type=int, default=0, metavar="THREADS",
                    help="no GPU: time the oracle and, where it is built, the reference on the "
                         "CPU sample with THREADS threads (cross-calibration)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device three times (a kernel of known "
                         "traffic that calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.cpu_only:
        cat, desc = make_workload(args.config)
        print(json.dumps({"config": args.config, "workload": desc,
                          "cpu": cpu_baseline(cat, threads=args.cpu_only, with_reference=True)}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the unbinding of one rank's share runs on one MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch

    from hbtplus_amd import abi
    from hbtplus_amd.abi import SUB_DTYPE
    from hbtplus_amd.device import GpuUnbinder

    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 4
        a = torch.arange(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: 3 x {args.copy_probe} MiB read + written", file=sys.stderr, flush=True)
        del a, b
    if args.config == "c5":
        return bench_c5(args)
    if args.config == "mask":
        return bench_mask(args)
    if args.config == "merge":
        return bench_merge(args)
    cat, desc = make_workload(args.config)
    params = cat.params.copy()
    params["precision"] = abi.HBT_PREC_REPLAY if args.replay else abi.HBT_PREC_FAST
    U = GpuUnbinder(params, device=0)
    h = U.upload(cat)
    torch.cuda.synchronize()
    print(f"[bench] {args.config}: {len(cat.subs)} subhalos, {cat.n_particles} particles", file=sys.stderr,
          flush=True)

    tuning = None
    if not args.no_autotune:
        # execution tuning (walk group sizes per size bin, passes per graph launch),
        # chosen by timing before the timed region; results are identical for every choice
        t = U.autotune(h, reps=1)
        tuning = {"walk_group": [int(v) for v in t["walk_group"]],
                  "bin_edge": [int(v) for v in t["bin_edge"]], "use_graph": int(t["use_graph"]),
                  "graph_passes": int(t["graph_passes"]), "xcd_run": int(t["xcd_run"])}
        print(f"[bench] autotuned: {tuning}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        U.unbind_device(h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(args.steps):
        out = U.unbind_device(h)
        dev_ms += U.last_stats()[0]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    subs = out[0].cpu().numpy().view(SUB_DTYPE)
    evals = int(subs["n_evals"].sum())
    passes = U.last_stats()[2]
    value = evals / (ms_per_step / 1e3)
    print(f"[bench] batched: {ms_per_step:.2f} ms/step, {evals} evals/step, {value:.4g} evals/s",
          file=sys.stderr, flush=True)

    # roofline of the walk kernel: timed launches, then an instrumented pass
    U.profile(1)
    U.unbind_device(h)
    prof_t = U.profile_read()
    U.profile(2)
    U.unbind_device(h)
    prof_c = U.profile_read()
    U.profile(0)
    I, O, T = prof_c["interactions"], prof_c["opened"], prof_c["targets"]
    lane_eff = (I + O) / max(prof_c["lane_slots"], 1)
    # op-mix ceiling: FP32 VALU issue slots (one per lane per instruction,
    # 3.93e13 lane-slots/s on 256 CUs x 4 SIMD16 at 2.4 GHz; an FMA counts 2
    # FLOPs, so 78.6 TFLOP/s is the non-packed FMA peak) for the algorithmic
    # mix: a summed source costs 3 sub + 3 (r^2) + 4 (rsq, quarter rate) + 1 mul
    # + 1 add = 12 slots for 11 FLOPs, an opened node 3 + 3 + 1 mul + 1 cmp =
    # 8 slots for 10 FLOPs, an evaluation 20 slots for 20 FLOPs
    flops = 11.0 * I + 10.0 * O + 20.0 * T
    slots = 12.0 * I + 8.0 * O + 20.0 * T
    op_mix_ceiling = LANE_SLOTS_PER_S * flops / slots / 1e12 if slots > 0 else 0.0
    walk_s = prof_t["walk_ms"] / 1e3
    launches = max(prof_t["walk_launches"], 1)
    achieved = flops / walk_s / 1e12 if walk_s > 0 else 0.0
    roofline = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / PEAK_FP32_TFLOPS, 4),
                    traffic=walk_traffic(args.config, launches),
                    op_mix_ceiling=round(op_mix_ceiling, 2),
                    frac_of_op_mix_ceiling=round(achieved / op_mix_ceiling, 4) if op_mix_ceiling else None,
                    kernel="walk_shared_kernel", launches_per_step=launches,
                    avg_launch_ms=round(prof_t["walk_ms"] / launches, 4),
                    flops_per_step=flops, interactions_per_eval=round(I / max(T, 1), 1),
                    opened_per_eval=round(O / max(T, 1), 1),
                    lane_efficiency=round(lane_eff, 3),
                    walk_share_of_step=round(prof_t["walk_ms"] / ms_per_step, 3))

    result = {"metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
              "config": {"workload": desc, "subhalos": int(len(cat.subs)),
                         "particles": int(cat.n_particles), "evals_per_step": evals,
                         "passes_per_step": int(passes),
                         "precision": "REPLAY (fp64 walk)" if args.replay else
                         "FAST (fp32+Kahan walk, fp64 frames)",
                         "device_ms_per_step": round(dev_ms / args.steps, 3),
                         "tuning": tuning},
              "roofline": roofline}

    if not args.no_plain:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pout = U.unbind_device(h, plain=True)
        torch.cuda.synchronize()
        pt = time.perf_counter() - t0
        same = all(np.array_equal(a.cpu().numpy().view(np.uint8), b.cpu().numpy().view(np.uint8))
                   for a, b in zip(out, pout))
        result["plain"] = {"value": evals / pt, "unit": "evals/s", "ms_per_step": round(pt * 1e3, 1),
                           "identical_outputs": bool(same)}
        result["batched_over_plain"] = round(value / (evals / pt), 2)
        print(f"[bench] plain: {pt * 1e3:.1f} ms, identical outputs {same}", file=sys.stderr, flush=True)

    # post-unbind properties (profile + shape) on the step's outputs
    result["properties"] = bench_properties(U, h, out, cat, args)

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cat, with_reference=True)
        print(f"[bench] cpu: {result['cpu_baseline']}", file=sys.stderr, flush=True)
    U.close()
    print(json.dumps(result), flush=True)


def bench_properties(U, h, out, cat, args):
    """hbt_subhalo_properties on the unbinding outputs: wall time per call
    (records read back for the layout, distance sort, FP64 scans and sums),
    against the oracle restatement on the host's cores (OpenMP over
    subhalos).  Bytes: each bound particle's index, position and mass read
    once (24 B) is the algorithmic minimum."""
    import torch

    from hbtplus_amd.abi import SUB_DTYPE
    subs = out[0].cpu().numpy().view(SUB_DTYPE)
    nbp = int(subs["nbound"][subs["nbound"] > 1].sum())
    U.properties_device(h, out)  # warm
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        props = U.properties_device(h, out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    res = {"ms_per_call": round(ms, 3), "subhalos": int(len(subs)), "bound_particles": nbp,
           "bound_particles_per_s": nbp / (ms / 1e3),
           "algorithmic_GBps": round(24.0 * nbp / (ms / 1e3) / 1e9, 2)}
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        idx = out[1].cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.subhalo_properties(cat.params, cat.cols, subs, idx, nthreads=threads)
        tc = time.perf_counter() - t0
        same = all(np.array_equal(np.ascontiguousarray(ref[f]).view(np.uint8),
                                  np.ascontiguousarray(props[f]).view(np.uint8)) for f in ref.dtype.names)
        res["cpu_port"] = {"ms": round(tc * 1e3, 1), "cores": threads, "identical": bool(same)}
    print(f"[bench] properties: {res}", file=sys.stderr, flush=True)
    return res


if __name__ == "__main__":
    main()

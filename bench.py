"""Benchmark: subhalo-particle binding-energy evaluations per second.

One step = one RefineParticles pass over one rank's share: every subhalo is
unbound to convergence (batched mixed-precision form, inputs resident in
HBM).  Default workload: 200,000 power-law subhalos (N_p 20..2e5, 20%
contaminants) given as source-ID lists into a shuffled 70.7 M-particle
snapshot in a periodic 100 Mpc/h box; a step resolves every listed ID on the
device and unbinds every subhalo.

Prints one JSON line.  Besides the headline (batched) throughput it reports
  plain          the plain per-subhalo GPU form over the same subhalos (same
                 kernels, one subhalo at a time; outputs must be identical),
                 and batched_over_plain on that identical work
  roofline       the dominant kernel (the potential walk: HIP-event timed
                 launches, FLOPs counted by an instrumented pass), for the
                 batched form; plain_roofline for the plain form
  potential_error  the FAST tree potential against an FP64 direct sum (and
                 the reference tree, i.e. REPLAY precision) on a sample
  cpu_baseline   the reference's own unbinding (oracle/_ref, built from its
                 sources) on the host cores over a bounded sample

    python bench.py [--steps K] [--warmup W] [--config c5|c2|c1|c3|c4|mask|merge]
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
SAMPLE = {"max_sample_size": 0}  # --max-sample: MaxSampleSizeOfPotentialEstimate for every leg


def sampling_note():
    ms = SAMPLE["max_sample_size"]
    if ms <= 0:
        return ("MaxSampleSizeOfPotentialEstimate=0 (no sampling), the configs' setting; the "
                "reference's default samples 1000 particles per full-pass tree")
    return (f"MaxSampleSizeOfPotentialEstimate={ms} with RefineMostboundParticle (the reference's default "
            f"mode): lists longer than the sample shuffled from one rand() stream, full-pass trees over "
            f"the first {ms} entries")


def apply_sampling(cat):
    if SAMPLE["max_sample_size"] > 0:
        cat.params["max_sample_size"] = SAMPLE["max_sample_size"]
        cat.params["refine_mostbound_particle"] = 1


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpus():
    """Threads for the CPU baseline and what the host offers: the process's
    affinity set, capped at 16 (the host cores a one-GPU process is given)."""
    aff = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(aff, 16), dict(affinity_cpus=aff, nproc=os.cpu_count(), cpu_model=model)


def walk_traffic(config):
    """HBM bytes per walk launch for a workload, from the committed counter
    summary profiles/walk_<config>_traffic.md (FETCH_SIZE and WRITE_SIZE
    passes of rocprofv3, each scaled by its reading for a copy of known
    size; see that file), or None.  The summaries cover the workloads without
    sampling; a sampled run reports None."""
    if SAMPLE["max_sample_size"] > 0:
        return None
    path = os.path.join(ROOT, "profiles", f"walk_{config}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
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
    cores, host = host_cpus()
    threads = threads or cores
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
        log(f"cpu baseline: restatement on {take} {what.split(' (')[0]}")
        t0 = time.perf_counter()
        subs, _, _ = oracle.unbind_batch(sub, nthreads=threads)
        elapsed = time.perf_counter() - t0
        evals = int(subs["n_evals"].sum())
        if elapsed > budget_s / 4 or take >= len(order):
            break
        take = min(len(order), int(take * max(2.0, budget_s / 4 / max(elapsed, 1e-3))))
    sample = (f"{take} of {len(order)} {what} drawn at random ({sub.n_particles} particles, {evals} "
              f"evals), OpenMP over nests, schedule(dynamic,1)")
    port = dict(value=evals / elapsed, unit="evals/s", cores=threads, kind="port",
                sample=f"{sample}; {elapsed:.1f} s", sampling=sampling_note(), **host)
    if not (with_reference and oracle.have_reference()):
        return port
    # the reference's own unbinding, compiled from its sources, on the same sample
    log(f"cpu baseline: reference on the same {take}")
    secs = oracle.ref_time_unbind(sub, threads)
    same = ("same trajectory" if SAMPLE["max_sample_size"] <= 0 else
            "the reference's threads draw rand() in scheduling order, so its shuffles and count may differ slightly")
    return dict(value=evals / secs, unit="evals/s", cores=threads, kind="reference",
                sample=f"{sample}; {secs:.1f} s (evaluation count from the port, {same})",
                sampling=sampling_note(), port=port, **host)


def walk_roofline(U, run, ms_per_step, config):
    """Roofline of the walk kernel: one run with every walk launch timed by
    HIP events on the library stream, one run with the kernel counting its
    summed sources (I), opened nodes (O) and targets (T)."""
    U.profile(1)
    run()
    prof_t = U.profile_read()
    sort_t = U.profile_sorts()
    U.profile(2)
    run()
    prof_c = U.profile_read()
    sort_c = U.profile_sorts()
    U.profile(0)
    I, O, T = prof_c["interactions"], prof_c["opened"], prof_c["targets"]
    lane_eff = (I + O) / max(prof_c["lane_slots"], 1)
    # a wave's groups step together until the slowest is done: wave steps
    # against the lane slots its groups needed (64 lanes per wave step)
    wave_eff = prof_c["lane_slots"] / 64.0 / max(prof_c["wave_steps"], 1)
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
    # algorithmic HBM bytes: 36 B of target record per evaluation (position,
    # velocity, mass, energy, list position) and every item of the trees
    # walked read once (32 B)
    alg_bytes = 36.0 * T + 32.0 * prof_c["tree_items"]
    return dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                frac=round(achieved / PEAK_FP32_TFLOPS, 4), traffic=walk_traffic(config),
                algorithmic_bytes_per_launch=int(alg_bytes / launches),
                op_mix_ceiling=round(op_mix_ceiling, 2),
                frac_of_op_mix_ceiling=round(achieved / op_mix_ceiling, 4) if op_mix_ceiling else None,
                kernel="walk_shared_kernel", launches_per_step=launches,
                avg_launch_ms=round(prof_t["walk_ms"] / launches, 4),
                walk_ms_per_step=round(prof_t["walk_ms"], 3),
                flops_per_step=flops, interactions_per_eval=round(I / max(T, 1), 1),
                opened_per_eval=round(O / max(T, 1), 1), lane_efficiency=round(lane_eff, 3),
                wave_efficiency=round(wave_eff, 3), wave_steps_per_step=prof_c["wave_steps"],
                ns_per_wave_step=round(prof_t["walk_ms"] * 1e6 / max(prof_c["wave_steps"], 1), 3),
                walk_share_of_step=round(prof_t["walk_ms"] / ms_per_step, 3),
                sorts=sort_roofline(sort_t, sort_c, ms_per_step)), flops


def sort_roofline(sort_t, sort_c, ms_per_step):
    """HBM roofline of the segmented sorts of a step (the octant keys of every
    tree, the energy order): the tile sort reads and writes each element's
    12-byte key and value once, and every merge level once more, so the
    algorithmic bytes are 24 B x (elements + elements x merge levels); timed
    with HIP events around each sort (tile sort and its merge levels)."""
    alg = 24.0 * (sort_c["elements"] + sort_c["element_levels"])
    s = sort_t["sort_ms"] / 1e3
    achieved = alg / s / 1e9 if s > 0 else 0.0
    return dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(achieved / PEAK_HBM_GBS, 4), kernels="tile_sort_kernel + merge_pass_kernel",
                sorts_per_step=sort_t["sort_calls"], ms_per_step=round(sort_t["sort_ms"], 3),
                elements_per_step=sort_c["elements"], element_levels_per_step=sort_c["element_levels"],
                algorithmic_bytes_per_step=int(alg), share_of_step=round(sort_t["sort_ms"] / ms_per_step, 3))


def potential_error(params, cols, subs, nsample=64, max_count=60000, seed=0):
    """FAST tree potential (fp32 terms summed in fp64, theta as configured) of whole
    source lists against the FP64 direct pair sum and against the reference
    tree (REPLAY precision: the reference's own arithmetic, bit for bit), on
    a sample of subhalos: the 8 largest lists of at most max_count particles
    and the rest drawn at random.  Per subhalo the relative RMS error
    sqrt(mean((a - b)^2) / mean(b^2)); median, 99th percentile and maximum
    over the sample."""
    from hbtplus_amd import abi
    from hbtplus_amd.device import GpuUnbinder
    cnt = subs["count"]
    ok = np.flatnonzero((cnt >= 2) & (cnt <= max_count))
    big = ok[np.argsort(cnt[ok])[-8:]]
    rng = np.random.default_rng(seed)
    rest = np.setdiff1d(ok, big)
    pick = np.concatenate([big, rng.choice(rest, size=min(nsample - len(big), len(rest)), replace=False)])
    pf, pr = params.copy(), params.copy()
    pf["precision"] = abi.HBT_PREC_FAST
    pr["precision"] = abi.HBT_PREC_REPLAY
    fast_d, ref_d, fast_ref = [], [], []
    pos = [np.arange(subs["offset"][s], subs["offset"][s] + cnt[s]) for s in pick]
    seg = np.zeros(len(pick) + 1, np.int64)
    seg[1:] = np.cumsum([len(p) for p in pos])
    allp = np.concatenate(pos)
    xyzm = [np.ascontiguousarray(cols[k][allp]) for k in ("x", "y", "z", "mass")]
    with GpuUnbinder(pf, device=0) as UF, GpuUnbinder(pr, device=0) as UR:
        direct = UF.direct_potential(seg, *xyzm, abi.HBT_DIRECT_FP64)
        for j in range(len(pick)):
            a, b = seg[j], seg[j + 1]
            src = [v[a:b] for v in xyzm]
            ft = UF.tree_potential(src, src[:3], src[3])
            rt = UR.tree_potential(src, src[:3], src[3])
            d = direct[a:b]
            rel = lambda u, v: float(np.sqrt(np.mean((u - v) ** 2) / np.mean(v ** 2)))  # noqa: E731
            fast_d.append(rel(ft, d))
            ref_d.append(rel(rt, d))
            fast_ref.append(rel(ft, rt))

    def stats(v):
        v = np.asarray(v)
        return {"median": float(np.median(v)), "p99": float(np.percentile(v, 99)), "max": float(v.max())}
    return {"sample": f"{len(pick)} subhalos, {int(seg[-1])} particles (lists of 2..{max_count}: the 8 "
                      f"largest and the rest at random), every particle a target of its own list",
            "fast_tree_vs_fp64_direct": stats(fast_d), "reference_tree_vs_fp64_direct": stats(ref_d),
            "fast_tree_vs_reference_tree": stats(fast_ref),
            "note": "tree vs direct is the opening-angle error of the method (the reference's own); "
                    "FAST vs the reference tree is the precision error (target <= 1e-4)"}


def apply_tuning(U, args):
    """--tuning JSON: set that execution tuning on the context (results are
    identical for every tuning); returns it, or None."""
    if not args.tuning:
        return None
    want = json.loads(args.tuning)
    t = U.get_tuning()
    for k, v in want.items():
        t[k] = v
    U.set_tuning(t)
    return want


def _record_bytes(subs):
    z = np.ascontiguousarray(subs).copy()
    z["out_offset"] = 0
    return z.view(np.uint8)


def _lists(oi, oe, offs, lens):
    lens = np.asarray(lens, np.int64)
    tot = int(lens.sum())
    if tot == 0:
        return oi[:0], oe[:0].view(np.uint32)
    start = np.repeat(np.asarray(offs, np.int64) - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
    at = start + np.arange(tot)
    return oi[at], oe[at].view(np.uint32)


def bench_c5(args):
    """One rank's share: 200,000 subhalos given as source-ID lists into a
    shuffled snapshot (with a uniform background).  A step resolves every list
    entry on the device (hbt_id_to_index) and unbinds all subhalos in one
    batched pass (hbt_unbind_batch_indexed)."""
    import torch

    from hbtplus_amd import abi, synth
    from hbtplus_amd.abi import SUB_DTYPE
    from hbtplus_amd.device import GpuUnbinder

    t0 = time.perf_counter()
    snap, ids, cat = synth.config5()
    apply_sampling(cat)
    gen_s = time.perf_counter() - t0
    params = cat.params.copy()
    params["precision"] = abi.HBT_PREC_FAST
    U = GpuUnbinder(params, device=0)
    h = U.upload_snapshot(snap, ids, cat.subs)
    torch.cuda.synchronize()
    log(f"c5: {len(cat.subs)} subhalos, {len(ids)} listed particles, snapshot {len(snap['x'])} "
        f"(generated in {gen_s:.0f} s)")

    def step():
        idx = U.resolve_ids(h)
        return U.unbind_indexed(h, idx)

    tuning = apply_tuning(U, args)
    if tuning is None and not args.no_autotune:
        # tuned on the whole share (the catalogue's gathered columns: the same
        # subhalos and lists), before the timed region
        hc = U.upload(cat)
        t = U.autotune(hc, reps=1)
        del hc
        tuning = {"walk_group": [int(v) for v in t["walk_group"]],
                  "bin_edge": [int(v) for v in t["bin_edge"]], "use_graph": int(t["use_graph"]),
                  "graph_passes": int(t["graph_passes"]), "xcd_run": int(t["xcd_run"])}
        log(f"autotuned: {tuning}")
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
    passes = U.last_stats()[2]
    value = evals / (ms_per_step / 1e3)
    t0 = time.perf_counter()
    idx = U.resolve_ids(h)
    torch.cuda.synchronize()
    resolve_ms = (time.perf_counter() - t0) * 1e3
    log(f"c5 batched: {ms_per_step:.1f} ms/step ({resolve_ms:.1f} ms ID resolution), {evals} evals, "
        f"{value:.4g} evals/s")
    roofline, flops = walk_roofline(U, step, ms_per_step, "c5")
    log(f"roofline: {roofline}")
    result = {"metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
              "config": {"workload": "200,000-subhalo share: power-law subhalos (dN/dN_p ~ N_p^-1.9 on "
                                     "[20, 2e5], 20% contaminants at 3 v_esc) as source-ID lists into a "
                                     "shuffled snapshot with a 10% uniform background, periodic 100 Mpc/h "
                                     "box, eps=0.5, theta=0.45",
                         "subhalos": int(len(cat.subs)), "listed_particles": int(len(ids)),
                         "snapshot_particles": int(len(snap["x"])), "evals_per_step": evals,
                         "passes_per_step": int(passes), "id_resolution_ms": round(resolve_ms, 2),
                         "precision": "FAST (fp32 walk terms summed in fp64, fp64 frames)", "sampling": sampling_note(),
                         "tuning": tuning},
              "roofline": roofline}
    # post-unbind properties on the step's outputs (snapshot positions)
    U.properties_device(h, out)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U.properties_device(h, out)
    torch.cuda.synchronize()
    pms = (time.perf_counter() - t0) * 1e3
    nbp = int(subs["nbound"][subs["nbound"] > 1].sum())
    result["properties"] = {"ms_per_call": round(pms, 2), "bound_particles": nbp,
                            "bound_particles_per_s": nbp / (pms / 1e3),
                            "note": "wall time per call, outputs read back to the host included"}
    ref_out = (subs.copy(), out[1].cpu().numpy(), out[2].cpu().numpy())
    del out
    if not args.no_plain:
        result.update(plain_c5(U, h, idx, ref_out, evals, value, flops, args.plain_chunks))
    U.close()
    if not args.no_error:
        log("potential error against the FP64 direct sum")
        result["potential_error"] = potential_error(cat.params, cat.cols, cat.subs)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cat, with_reference=True)
        log(f"cpu: {result['cpu_baseline']}")
    print(json.dumps(result), flush=True)


def plain_c5(U, h, idx, ref_out, evals, value, flops, nchunks):
    """The plain per-subhalo form over every subhalo of the share, in chunks
    of consecutive subhalos (one progress line each; the plain form unbinds
    one subhalo at a time, so chunking changes nothing in its work), compared
    byte for byte with the batched step's records, lists and energies.  The
    walk launches of the first chunk are timed for the plain form's roofline
    (its FLOPs per evaluation are the batched form's: same trees, same
    walks)."""
    import torch

    from hbtplus_amd.abi import SUB_DTYPE
    rs, ri, re_ = ref_out
    nsub = len(rs)
    tot_s, same, pev = 0.0, True, 0
    walk_ms, walk_launches, walk_evals = 0.0, 0, 0
    for j, sel in enumerate(np.array_split(np.arange(nsub), nchunks)):
        hs = U.subset_handle(h, sel)
        prof = j == 0
        if prof:
            U.profile(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o = U.unbind_indexed(hs, idx, plain=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if prof:
            p = U.profile_read()
            U.profile(0)
            walk_ms, walk_launches = p["walk_ms"], p["walk_launches"]
        else:
            tot_s += dt
        ps = o[0].cpu().numpy().view(SUB_DTYPE)
        pi, pe = o[1].cpu().numpy(), o[2].cpu().numpy()
        ev = int(ps["n_evals"].sum())
        if prof:
            walk_evals = ev
        else:
            pev += ev
        ok = np.array_equal(_record_bytes(ps), _record_bytes(rs[sel]))
        a = _lists(pi, pe, ps["out_offset"], ps["list_len"])
        b = _lists(ri, re_, rs["out_offset"][sel], rs["list_len"][sel])
        ok = ok and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        same = same and ok
        log(f"plain chunk {j + 1}/{nchunks}: {len(sel)} subhalos, {dt * 1e3:.0f} ms, identical {ok}"
            + (" (walk launches timed; not in the rate)" if prof else ""))
    plain_rate = pev / tot_s
    res = {"plain": {"value": plain_rate, "unit": "evals/s", "subhalos": int(nsub),
                     "ms_all_subhalos": round(tot_s * 1e3 * evals / max(pev, 1), 1),
                     "identical_outputs": bool(same),
                     "note": f"every subhalo of the share in {nchunks} chunks; the rate is over chunks "
                             f"2..{nchunks} ({pev} evals in {tot_s:.1f} s), the first chunk ran with "
                             f"walk-launch timing; every chunk's outputs compared with the batched step"},
           "batched_over_plain": round(value / plain_rate, 2)}
    if walk_launches:
        # the first chunk's walk FLOPs, at the batched pass's FLOPs per evaluation
        fl = flops * walk_evals / evals
        ach = fl / (walk_ms / 1e3) / 1e12
        res["plain_roofline"] = {"bound": "valu", "achieved": round(ach, 4), "peak": PEAK_FP32_TFLOPS,
                                 "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 5),
                                 "kernel": "walk_shared_kernel", "launches": int(walk_launches),
                                 "avg_launch_ms": round(walk_ms / walk_launches, 4),
                                 "sample": f"first chunk ({walk_evals} evals)"}
    return res


def bench_catalogue(args):
    """A catalogue workload (c1-c4): columns resident on the device, one
    hbt_unbind_batch per step."""
    import torch

    from hbtplus_amd import abi
    from hbtplus_amd.abi import SUB_DTYPE
    from hbtplus_amd.device import GpuUnbinder

    cat, desc = make_workload(args.config)
    apply_sampling(cat)
    params = cat.params.copy()
    params["precision"] = abi.HBT_PREC_REPLAY if args.replay else abi.HBT_PREC_FAST
    U = GpuUnbinder(params, device=0)
    h = U.upload(cat)
    torch.cuda.synchronize()
    log(f"{args.config}: {len(cat.subs)} subhalos, {cat.n_particles} particles")

    tuning = apply_tuning(U, args)
    if tuning is None and not args.no_autotune:
        # execution tuning (walk group sizes per size bin, passes per graph launch),
        # chosen by timing before the timed region; results are identical for every choice
        t = U.autotune(h, reps=1)
        tuning = {"walk_group": [int(v) for v in t["walk_group"]],
                  "bin_edge": [int(v) for v in t["bin_edge"]], "use_graph": int(t["use_graph"]),
                  "graph_passes": int(t["graph_passes"]), "xcd_run": int(t["xcd_run"])}
        log(f"autotuned: {tuning}")
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
    log(f"batched: {ms_per_step:.2f} ms/step, {evals} evals/step, {value:.4g} evals/s")
    roofline, flops = walk_roofline(U, lambda: U.unbind_device(h), ms_per_step, args.config)

    result = {"metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
              "config": {"workload": desc, "subhalos": int(len(cat.subs)),
                         "particles": int(cat.n_particles), "evals_per_step": evals,
                         "passes_per_step": int(passes),
                         "precision": "REPLAY (fp64 walk)" if args.replay else
                         "FAST (fp32 walk terms summed in fp64, fp64 frames)", "sampling": sampling_note(),
                         "device_ms_per_step": round(dev_ms / args.steps, 3),
                         "tuning": tuning},
              "roofline": roofline}

    if not args.no_plain:
        U.profile(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pout = U.unbind_device(h, plain=True)
        torch.cuda.synchronize()
        pt = time.perf_counter() - t0
        p = U.profile_read()
        U.profile(0)
        same = all(np.array_equal(a.cpu().numpy().view(np.uint8), b.cpu().numpy().view(np.uint8))
                   for a, b in zip(out, pout))
        result["plain"] = {"value": evals / pt, "unit": "evals/s", "ms_per_step": round(pt * 1e3, 1),
                           "identical_outputs": bool(same),
                           "note": "one run with walk-launch timing on (as in every pass of the plain "
                                   "form, which synchronises after each pass anyway)"}
        result["batched_over_plain"] = round(value / (evals / pt), 2)
        if p["walk_launches"]:
            ach = flops / (p["walk_ms"] / 1e3) / 1e12
            result["plain_roofline"] = {"bound": "valu", "achieved": round(ach, 4), "peak": PEAK_FP32_TFLOPS,
                                        "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 5),
                                        "kernel": "walk_shared_kernel", "launches": int(p["walk_launches"]),
                                        "avg_launch_ms": round(p["walk_ms"] / p["walk_launches"], 4)}
        log(f"plain: {pt * 1e3:.1f} ms, identical outputs {same}")

    # post-unbind properties (profile + shape) on the step's outputs
    result["properties"] = bench_properties(U, h, out, cat, args)
    U.close()
    if not args.no_error:
        result["potential_error"] = potential_error(cat.params, cat.cols, cat.subs)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cat, with_reference=True)
        log(f"cpu: {result['cpu_baseline']}")
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
        threads, _ = host_cpus()
        idx = out[1].cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.subhalo_properties(cat.params, cat.cols, subs, idx, nthreads=threads)
        tc = time.perf_counter() - t0
        same = all(np.array_equal(np.ascontiguousarray(ref[f]).view(np.uint8),
                                  np.ascontiguousarray(props[f]).view(np.uint8)) for f in ref.dtype.names)
        res["cpu_port"] = {"ms": round(tc * 1e3, 1), "cores": threads, "identical": bool(same)}
    log(f"properties: {res}")
    return res


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
    log(f"merge: {len(sin)} subhalos, {cat.n_particles} particles, {len(lst)} list entries")
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
        threads, host = host_cpus()
        secs = oracle.ref_merge_subhalos(params, cat.cols, lst, sin, grp, True, time_threads=threads)
        res["cpu_baseline"] = {"kind": "reference", "ms_per_call": round(secs * 1e3, 1), "cores": threads,
                               "value": len(sin) / secs, "unit": "subhalos/s", **host}
    U.close()
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c5")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-error", action="store_true", help="skip the potential-error sample")
    ap.add_argument("--plain-chunks", type=int, default=20, help="chunks of the C5 plain run")
    ap.add_argument("--replay", action="store_true", help="REPLAY precision instead of FAST")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default execution tuning")
    ap.add_argument("--max-sample", type=int, default=0, metavar="N",
                    help="MaxSampleSizeOfPotentialEstimate (with RefineMostboundParticle) for the GPU and "
                         "CPU legs; 0 (default): no sampling, the configs' setting")
    ap.add_argument("--tuning", default="", metavar="JSON",
                    help="set this execution tuning (the line's config.tuning) instead of autotuning")
    ap.add_argument("--cpu-only", type=int, default=0, metavar="THREADS",
                    help="no GPU: time the oracle and, where it is built, the reference on the "
                         "CPU sample with THREADS threads (cross-calibration)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device three times (a kernel of known "
                         "traffic that calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    SAMPLE["max_sample_size"] = max(0, args.max_sample)
    if args.cpu_only:
        if args.config == "c5":
            from hbtplus_amd import synth
            _, _, cat = synth.config5()
            desc = "C5 share (catalogue of gathered lists)"
        else:
            cat, desc = make_workload(args.config)
        apply_sampling(cat)
        print(json.dumps({"config": args.config, "workload": desc,
                          "cpu": cpu_baseline(cat, threads=args.cpu_only, with_reference=True)}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the unbinding of one rank's share runs on one MI355X; "
              "runs on several GPUs are not supported (ranks would each unbind their own share)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 4
        a = torch.arange(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        log(f"copy probe: 3 x {args.copy_probe} MiB read + written")
        del a, b
    if args.config == "c5":
        return bench_c5(args)
    if args.config == "mask":
        return bench_mask(args)
    if args.config == "merge":
        return bench_merge(args)
    return bench_catalogue(args)


if __name__ == "__main__":
    main()

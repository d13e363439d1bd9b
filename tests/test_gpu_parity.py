"""GPU parity: the HIP library through its C-ABI against the reference's own
outputs (golden fixtures) and the CPU oracle.

REPLAY precision reproduces the reference's arithmetic: records and particle
lists must be bit-identical (lists up to the order of exactly tied energies).
The one documented exception is the sub-resolution floor, where the reference
places coincident particles with the process-global drand48 stream; there the
GPU keeps such particles in one bucket cell and parity is tolerant.
FAST precision (FP32 walk terms summed in FP64, FP64 frames) meets the tolerant criteria:
bound sets identical outside the 1e-5*|E_min| band around E=0, bound counts
within the band count, most-bound IDs equal unless tied within the band,
potentials within 1e-4 relative RMS of the reference tree."""
import hashlib
import os
import sys

import numpy as np
import pytest

import oracle
from fixtures import TREE_CASES, UNBIND_CASES, load_tree, load_unbind
from hbtplus_amd import abi, synth
from helpers import as_ids, assert_exact, assert_tolerant, check_unbind_invariants, unused_slots

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_c4_full import catalogue_digest  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

EXACT_CASES = [c for c in UNBIND_CASES if c != "floor"]
SUB_FIELDS = abi.SUB_DTYPE.names


def _with(params, **kw):
    p = params.copy()
    for k, v in kw.items():
        p[k] = v
    return p


def _same_bytes(a, b):
    return all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(a, b))


# ---- potentials ------------------------------------------------------------

@pytest.mark.parametrize("name", TREE_CASES)
def test_tree_potential_replay(gpu, name):
    params, (x, y, z, m), ref = load_tree(name)
    got = gpu(_with(params, precision=abi.HBT_PREC_REPLAY)).tree_potential((x, y, z, m), (x, y, z), m)
    nbad = int(np.sum(got != ref))
    rel = np.abs(got / ref - 1)
    print(f"{name}: {nbad}/{len(ref)} differ, max rel {rel.max():.2e}")
    if name == "tree_floor":
        assert rel.max() < 1e-6  # bucketed sub-resolution cells
    else:
        assert nbad == 0


@pytest.mark.parametrize("name", TREE_CASES)
def test_tree_potential_fast(gpu, name):
    params, (x, y, z, m), ref = load_tree(name)
    got = gpu(_with(params, precision=abi.HBT_PREC_FAST)).tree_potential((x, y, z, m), (x, y, z), m)
    rms = np.sqrt(np.mean(((got - ref) / ref) ** 2))
    print(f"{name}: FAST relative RMS vs reference tree {rms:.2e}")
    assert rms < 1e-4


def test_tree_potential_external_targets(gpu):
    """targets outside the source set (target_mass = 0)"""
    cat = synth.config1(n_nfw=2000, n_cont=0)
    c = cat.cols
    src = (c["x"], c["y"], c["z"], c["mass"])
    rng = np.random.default_rng(3)
    t = rng.normal(size=(500, 3)) * 80.0
    tgt = tuple(t[:, i].astype(np.float32) for i in range(3))
    p = _with(cat.params, precision=abi.HBT_PREC_REPLAY)
    ref = oracle.tree_potential(p, src, tgt, None)
    got = gpu(p).tree_potential(src, tgt, None)
    assert np.array_equal(ref, got)


@pytest.mark.parametrize("prec", [abi.HBT_PREC_FAST, abi.HBT_PREC_REPLAY])
def test_tree_potential_targets_across_the_box_edge(gpu, prec):
    """periodic box: sources just above the z = 0 face, targets just below
    z = box (and some among the sources); the separations wrap around the
    face in both precisions"""
    rng = np.random.default_rng(5)
    box = 1.0e4
    n = 3000
    sp = rng.normal(size=(n, 3)) * 25.0 + np.array([500.0, 500.0, 150.0])
    assert sp[:, 2].min() > 0
    src = tuple(sp[:, i].astype(np.float32) for i in range(3)) + (np.full(n, 1e-4, np.float32),)
    tp = rng.normal(size=(400, 3)) * 30.0 + np.array([500.0, 500.0, box - 60.0])
    tp = np.concatenate([tp, sp[:100]]) % box
    tgt = tuple(tp[:, i].astype(np.float32) for i in range(3))
    p = abi.make_params(periodic=1, box_size=box, precision=prec)
    ref = oracle.tree_potential(p, src, tgt, None)
    got = gpu(p).tree_potential(src, tgt, None)
    if prec == abi.HBT_PREC_REPLAY:
        assert np.array_equal(ref, got)
    else:
        rel = np.abs(got / ref - 1)
        print(f"FAST across the edge: max rel {rel.max():.2e}")
        assert rel.max() < 1e-5


# ---- unbinding: golden fixtures ----------------------------------------------

@pytest.mark.parametrize("name", EXACT_CASES)
def test_unbind_replay_matches_reference(gpu, name):
    cat, ref = load_unbind(name)
    got = gpu(_with(cat.params, precision=abi.HBT_PREC_REPLAY)).unbind_batch(cat)
    assert_exact(ref, got)


def test_unbind_replay_floor_tolerant(gpu):
    cat, ref = load_unbind("floor")
    got = gpu(_with(cat.params, precision=abi.HBT_PREC_REPLAY)).unbind_batch(cat)
    print("floor, replay:", assert_tolerant(ref, got))


@pytest.mark.parametrize("name", UNBIND_CASES)
def test_unbind_fast_matches_reference(gpu, name):
    cat, ref = load_unbind(name)
    got = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST)).unbind_batch(cat)
    print(name, "fast:", assert_tolerant(ref, got))


# ---- unbinding: forms, sizes, entry points -----------------------------------

@pytest.mark.parametrize("prec", [abi.HBT_PREC_FAST, abi.HBT_PREC_REPLAY])
def test_plain_equals_batched(gpu, prec):
    cat, _ = load_unbind("c4_small")
    U = gpu(_with(cat.params, precision=prec))
    a = U.unbind_batch(cat, plain=False)
    b = U.unbind_batch(cat, plain=True)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_unbind_one_matches_oracle(gpu):
    cat, _ = load_unbind("c1_small")
    p = _with(cat.params, precision=abi.HBT_PREC_REPLAY)
    ref = oracle.unbind_batch(cat, params=p)
    rec, idx, en = gpu(p).unbind_one(cat.cols, cat.subs[:1])
    rec["out_offset"] = 0
    assert_exact(ref, (rec, idx, en))


def test_unbind_c1_full_replay(gpu):
    cat = synth.config1()
    p = _with(cat.params, precision=abi.HBT_PREC_REPLAY)
    ref = oracle.unbind_batch(cat, params=p)
    got = gpu(p).unbind_batch(cat)
    assert_exact(ref, got)


def test_unbind_c2_sample_fast_and_replay(gpu):
    cat = synth.config2(seed=1002, nsub=400)
    ref = oracle.unbind_batch(cat, nthreads=16)
    got = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST)).unbind_batch(cat)
    print("C2 sample fast:", assert_tolerant(ref, got))
    rep = gpu(_with(cat.params, precision=abi.HBT_PREC_REPLAY)).unbind_batch(cat)
    print("C2 sample replay:", assert_tolerant(ref, rep))


def test_unbind_at_tile_form_boundaries(gpu):
    """Lists of exactly 1, 2, 3, 63-65, 255-257, 511-513, 1023-1025 and 2049
    entries: the one-wave tile kernels take the segments of at most 64 slots
    (one entry per lane) and of at most 256 (four per lane), the workgroup
    kernels the rest, and a segment above 1024 needs merge levels.  FAST and
    REPLAY against the oracle, and the batched form against the plain one
    (which runs each subhalo alone, so every kernel form meets every size)."""
    base = synth.config2(seed=1021, nsub=3000)
    sizes = [1, 2, 3, 63, 64, 65, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025, 2049]
    rng = np.random.default_rng(23)
    picks, counts = [], []
    for n in sizes:
        cand = np.flatnonzero(base.subs["count"] >= n)
        for s in rng.choice(cand, size=min(3, len(cand)), replace=False):
            picks.append(int(s))
            counts.append(n)
    subs = base.subs[np.array(picks)].copy()
    subs["count"] = counts  # the first n entries of each list (lists are in random order)
    subs["nbound_prev"] = counts
    cat = synth.Catalogue(base.cols, subs, base.params.copy(), "tile-form boundaries")
    ref = oracle.unbind_batch(cat, nthreads=8)
    fast = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST))
    got = fast.unbind_batch(cat)
    print("boundaries, fast:", assert_tolerant(ref, got))
    plain = fast.unbind_batch(cat, plain=True)
    for x, y in zip(got, plain):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    rep = gpu(_with(cat.params, precision=abi.HBT_PREC_REPLAY)).unbind_batch(cat)
    print("boundaries, replay:", assert_tolerant(ref, rep))


def test_unbind_at_box_edges(gpu):
    """400 subhalos centred on the faces, edges and corners of the periodic
    box (380 wrap around it): FAST meets the tolerant criteria and the
    size-free invariants, REPLAY is bit-identical to the oracle, and the
    per-subhalo form gives the batched bytes.  A correction pass's sources
    (the block just removed) can lie on one side of a face while its targets
    lie on both."""
    cat = synth.config2(seed=1012, nsub=400, hi=2.0e4, at_edges=True)
    ref = oracle.unbind_batch(cat, nthreads=16)
    U = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST))
    fast = U.unbind_batch(cat)
    check_unbind_invariants(cat, fast)
    print("edges fast:", assert_tolerant(ref, fast))
    assert _same_bytes(fast, U.unbind_batch(cat, plain=True))
    rep = gpu(_with(cat.params, precision=abi.HBT_PREC_REPLAY)).unbind_batch(cat)
    assert_exact(ref, rep)


def test_repeated_calls_are_identical(gpu):
    cat, _ = load_unbind("c2_small")
    U = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST))
    a = U.unbind_batch(cat)
    big = synth.config2(seed=9, nsub=30, hi=4000)
    U.unbind_batch(big)
    b = U.unbind_batch(cat)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


# ---- direct sums, ID lookup ------------------------------------------------------

def test_direct_potential_modes(gpu):
    cat = synth.config3(seed=77, nsub=30)
    c = cat.cols
    seg = np.concatenate([cat.subs["offset"], [cat.n_particles]])
    U = gpu(cat.params)
    ref64 = oracle.direct_potential(cat.params, seg, c["x"], c["y"], c["z"], c["mass"],
                                    abi.HBT_DIRECT_FP64, nthreads=8)
    g64 = U.direct_potential(seg, c["x"], c["y"], c["z"], c["mass"], abi.HBT_DIRECT_FP64)
    assert np.allclose(g64, ref64, rtol=1e-12, atol=0)
    errs = {}
    for mode in (abi.HBT_DIRECT_FP32, abi.HBT_DIRECT_FP32_KAHAN):
        g = U.direct_potential(seg, c["x"], c["y"], c["z"], c["mass"], mode)
        errs[mode] = np.sqrt(np.mean(((g - ref64) / ref64) ** 2))
    print("direct relative RMS vs FP64: fp32 %.2e kahan %.2e" % (errs[0], errs[1]))
    assert errs[abi.HBT_DIRECT_FP32_KAHAN] < 1e-6 and errs[abi.HBT_DIRECT_FP32] < 1e-5


def test_id_to_index(gpu):
    rng = np.random.default_rng(8)
    table = synth.unique_ids(np.random.Generator(np.random.Philox(key=3)), 200000)
    q = np.concatenate([table[rng.integers(0, len(table), 50000)],
                        rng.integers(0, 2 ** 62, 1000), -rng.integers(1, 100, 10)])
    rng.shuffle(q)
    U = gpu(abi.make_params())
    got = U.id_to_index(table, q)
    ref = oracle.id_to_index(table, q)
    assert np.array_equal(got, ref)
    assert U.id_to_index(table, np.zeros(0, np.int64)).size == 0


def test_id_to_index_two_level_and_repeats(gpu):
    """3 M IDs take both scatter levels of the partitioned table (2048
    partitions); repeated IDs keep their smallest position; 2 M repeats of
    one ID overfill a partition and take the single-table build; an empty
    table and a table of null IDs resolve nothing."""
    rng = np.random.default_rng(21)
    U = gpu(abi.make_params())
    table = synth.unique_ids(np.random.Generator(np.random.Philox(key=5)), 3_000_000)
    table[rng.integers(0, len(table), 1000)] = -rng.integers(1, 9, 1000)  # null IDs are never entered
    q = np.concatenate([table[rng.integers(0, len(table), 400000)], rng.integers(0, 2 ** 62, 20000),
                        -rng.integers(1, 100, 10)])
    rng.shuffle(q)
    assert np.array_equal(U.id_to_index(table, q), oracle.id_to_index(table, q))
    # every ID three times, in shuffled order
    base = table[:400000]
    rep = np.concatenate([base, base, base])
    rng.shuffle(rep)
    q = np.concatenate([base[rng.integers(0, len(base), 50000)], rng.integers(0, 2 ** 62, 1000)])
    assert np.array_equal(U.id_to_index(rep, q), oracle.id_to_index(rep, q))
    # one ID two million times among unique ones: its partition overflows
    many = np.concatenate([np.full(2_000_000, 77, np.int64), table[:500000]])
    rng.shuffle(many)
    q = np.concatenate([np.array([77, 78], np.int64), table[rng.integers(0, 500000, 20000)]])
    got = U.id_to_index(many, q)
    assert np.array_equal(got, oracle.id_to_index(many, q))
    assert got[0] == int(np.flatnonzero(many == 77)[0])
    assert np.all(U.id_to_index(np.zeros(0, np.int64), q) == -1)
    assert np.all(U.id_to_index(-np.ones(5000, np.int64), q) == -1)


# ---- execution variants give identical results ---------------------------------

def _unbinder_with_env(params, **env):
    import os
    from hbtplus_amd.device import GpuUnbinder
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return GpuUnbinder(params, device=0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("prec", [abi.HBT_PREC_FAST, abi.HBT_PREC_REPLAY])
def test_walk_modes_and_graph_are_identical(prec):
    """group-shared walk with 64/32/16/8-lane groups vs per-lane walk, hipGraph
    replay vs direct launches, 1 vs 3 passes per graph launch: byte-identical
    outputs"""
    cat = synth.config2(seed=61, nsub=200, hi=20000)
    p = _with(cat.params, precision=prec)
    runs = []
    for env in (dict(HBT_WALK_MODE=0, HBT_GRAPH=1, HBT_GRAPH_PASSES=2),
                dict(HBT_WALK_MODE=1, HBT_GRAPH=1, HBT_GRAPH_PASSES=2),
                dict(HBT_WALK_MODE=0, HBT_WALK_GROUP=32),
                dict(HBT_WALK_MODE=0, HBT_WALK_GROUP=16),
                dict(HBT_WALK_MODE=0, HBT_WALK_GROUP=8),
                dict(HBT_WALK_MODE=0, HBT_GRAPH=0),
                dict(HBT_WALK_MODE=0, HBT_GRAPH=1, HBT_GRAPH_PASSES=3)):
        U = _unbinder_with_env(p, **env)
        runs.append(U.unbind_batch(cat))
        U.close()
    for other in runs[1:]:
        for x, y in zip(runs[0], other):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_indexed_batch_matches_direct(gpu):
    """Lists given as IDs, resolved on the device, unbound from a shuffled
    snapshot with background particles: same results as the direct batch."""
    snap, ids, cat = synth.config5(seed=5, nsub=300)
    U = gpu(cat.params)
    h = U.upload_snapshot(snap, ids, cat.subs)
    idx = U.resolve_ids(h)
    pos = idx.cpu().numpy()
    assert np.array_equal(snap["id"][pos], ids)
    s1, i1, e1 = U.unbind_indexed(h, idx)
    s2, i2, e2 = U.unbind_batch(cat)
    s1 = s1.cpu().numpy().view(abi.SUB_DTYPE)
    i1 = i1.cpu().numpy()[: len(i2)]
    for f in abi.OUTPUT_FIELDS:
        assert np.array_equal(s1[f], s2[f]), f
    # out_index: snapshot positions vs catalogue positions of the same particles
    assert np.array_equal(snap["id"][i1], cat.cols["id"][i2])
    assert np.array_equal(e1.cpu().numpy()[: len(e2)].view(np.uint32), e2.view(np.uint32))


def test_indexed_batch_rejects_positions_past_the_snapshot(gpu):
    """negative positions are null particles (dropped); a position at or past
    the snapshot's end is an error"""
    snap, ids, cat = synth.config5(seed=6, nsub=20)
    U = gpu(cat.params)
    h = U.upload_snapshot(snap, ids, cat.subs)
    idx = U.resolve_ids(h)
    idx[3] = len(snap["x"])
    with pytest.raises(RuntimeError, match="past the snapshot"):
        U.unbind_indexed(h, idx)


def test_tuning_api_and_bins_are_identical(gpu):
    """per-bin group sizes set through hbt_set_tuning, and the autotuner's
    choice: byte-identical outputs; invalid tunings are rejected"""
    cat = synth.config2(seed=62, nsub=150, hi=20000)
    for prec in (abi.HBT_PREC_FAST, abi.HBT_PREC_REPLAY):
        U = gpu(_with(cat.params, precision=prec))
        base = U.unbind_batch(cat)
        t = U.get_tuning()
        t["walk_group"] = (64, 32, 8)
        t["bin_edge"] = (300, 3000)
        t["graph_passes"] = 3
        t["xcd_run"] = 4
        U.set_tuning(t)
        got = U.unbind_batch(cat)
        for x, y in zip(base, got):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
        chosen = U.autotune(U.upload(cat), reps=1)
        assert int(chosen["walk_group"][2]) in (8, 16, 32, 64)
        got = U.unbind_batch(cat)
        for x, y in zip(base, got):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    for field, value in (("walk_group", (64, 12, 64)), ("xcd_run", 3), ("graph_passes", 0)):
        bad = U.get_tuning()
        bad[field] = value
        with pytest.raises(RuntimeError):
            U.set_tuning(bad)


def test_plain_equals_batched_c3_sample(gpu):
    """many small subhalos (N_p log-uniform on [20, 4096]): the batched and
    per-subhalo forms give byte-identical records, lists and energies (each
    pass's launch bounds come from the level's other subhalos in one form and
    from the subhalo alone in the other), and match the oracle"""
    cat = synth.config3(seed=1003, nsub=3000)
    U = gpu(cat.params)
    a = U.unbind_batch(cat, plain=False)
    b = U.unbind_batch(cat, plain=True)
    bad = [f for f in SUB_FIELDS if not np.array_equal(a[0][f], b[0][f])]
    rows = np.nonzero(np.any(a[0].view(np.uint8).reshape(len(a[0]), -1) !=
                             b[0].view(np.uint8).reshape(len(b[0]), -1), axis=1))[0]
    print("differing fields", bad, "rows", rows[:10], len(rows))
    for r in rows[:3]:
        print({f: (a[0][f][r], b[0][f][r]) for f in bad})
    assert not bad and len(rows) == 0
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
    print("C3 sample fast vs oracle:", assert_tolerant(oracle.unbind_batch(cat, nthreads=16), a))


# ---- full sizes ------------------------------------------------------------------

def test_unbind_c2_full_size(gpu):
    """The bench workload at full size (5,000 power-law subhalos, 1.50 M
    particles): FAST meets the tolerant criteria against the oracle on every
    subhalo and the size-free invariants, a second call is byte-identical,
    and REPLAY is bit-identical to the oracle"""
    cat = synth.config2()
    U = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST))
    got = U.unbind_batch(cat)
    nconv = check_unbind_invariants(cat, got)
    ref = oracle.unbind_batch(cat, nthreads=16)
    stats = assert_tolerant(ref, got)
    assert _same_bytes(got, U.unbind_batch(cat))
    print(f"C2 full: {nconv} of {len(cat.subs)} converged, {stats}")
    assert_exact(ref, gpu(_with(cat.params, precision=abi.HBT_PREC_REPLAY)).unbind_batch(cat))


def test_unbind_c3_full_size(gpu):
    """50,000 subhalos, 44.3 M particles (N_p log-uniform on [20, 4096], 15%
    contaminants), FAST: the size-free invariants over the whole catalogue,
    the tolerant criteria against the oracle on 1,000 subhalos drawn at
    random (a flat batch's subhalos do not interact, so the oracle runs them
    alone; lists are compared as particle IDs), and a second call
    byte-identical"""
    cat = synth.config3()
    U = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST))
    got = U.unbind_batch(cat)
    nconv = check_unbind_invariants(cat, got)
    sel = np.sort(np.random.default_rng(7).choice(len(cat.subs), 1000, replace=False))
    part = cat.subset(sel)
    rs, ri, re_ = oracle.unbind_batch(part, nthreads=16)
    stats = assert_tolerant((rs, as_ids(part.cols["id"], ri), re_), (got[0][sel], as_ids(cat.cols["id"], got[1]), got[2]))
    assert _same_bytes(got, U.unbind_batch(cat))
    print(f"C3 full: {nconv} of {len(cat.subs)} converged, sample {stats}")


def test_unbind_c4_full_size(gpu):
    """One 4 M-particle host holding 200 satellites (5.33 M particles), FAST:
    the size-free invariants over the nest (the host's list is its own
    particles plus the satellites' unbound tails), the satellites against the
    oracle run on them alone (a satellite is unbound from its own list before
    its host), the host's unused output slots hold -1 and NaN, the
    per-subhalo form gives the batched bytes, and every
    subhalo, the 4 M-particle host included, meets the tolerant criteria
    against a digest of the oracle's run of the whole catalogue."""
    cat = synth.config4()
    U = gpu(_with(cat.params, precision=abi.HBT_PREC_FAST))
    got = U.unbind_batch(cat)
    nconv = check_unbind_invariants(cat, got)
    free = unused_slots(got[0], len(got[1]))
    assert free.sum() > 0 and np.all(got[1][free] == -1) and np.all(np.isnan(got[2][free]))
    sat = np.nonzero(cat.subs["parent"] >= 0)[0]
    part = cat.subset(sat)
    rs, ri, re_ = oracle.unbind_batch(part, nthreads=16)
    stats = assert_tolerant((rs, as_ids(part.cols["id"], ri), re_), (got[0][sat], as_ids(cat.cols["id"], got[1]), got[2]))
    assert _same_bytes(got, U.unbind_batch(cat, plain=True))
    print(f"C4 full: {nconv} of {len(cat.subs)} converged, host nbound {got[0]['nbound'][0]}, satellites {stats}")
    # every subhalo, the host included, against the oracle's run of the whole
    # catalogue (tests/golden/make_c4_full.py): bound sets equal outside the
    # band, Nbound within the band count, most-bound IDs equal unless tied
    z = np.load(os.path.join(GOLDEN, "c4_full_oracle.npz"))
    assert bytes(z["catalogue_sha1"]) == catalogue_digest(cat), "catalogue generator changed"
    ids = cat.cols["id"].astype(np.uint64)
    subs, idx, _ = got
    for s in range(len(subs)):
        o, L, nb = int(subs["out_offset"][s]), int(subs["list_len"][s]), int(subs["nbound"][s])
        band = z["band_ids"][z["band_off"][s]:z["band_off"][s + 1]]
        b = ids[idx[o:o + nb]]
        outside = np.sort(b[~np.isin(b, band)])
        assert hashlib.sha1(outside.tobytes()).digest() == bytes(z["bound_digest"][s]), f"subhalo {s}: bound set"
        assert abs(nb - int(z["nbound"][s])) <= len(band), f"subhalo {s}: nbound {nb} vs {z['nbound'][s]}"
        assert ids[idx[o]] == z["mostbound_id"][s] or z["top_tie"][s], f"subhalo {s}: most-bound ID"


def test_indexed_batch_c5_tenth_size(gpu):
    """A tenth of the one-rank share (20,000 subhalos, 6.4 M listed particles
    in a shuffled 7.1 M-particle snapshot; the full share's catalogue takes
    minutes to generate and is run by the bench): ID lists resolved on the
    device and unbound from the snapshot give the direct batch's results,
    which meet the size-free invariants and, on 500 subhalos drawn at random,
    the tolerant criteria against the oracle"""
    snap, ids, cat = synth.config5(seed=1015, nsub=20000)
    U = gpu(cat.params)
    h = U.upload_snapshot(snap, ids, cat.subs)
    idx = U.resolve_ids(h)
    assert np.array_equal(snap["id"][idx.cpu().numpy()], ids)
    s1, i1, e1 = U.unbind_indexed(h, idx)
    got = U.unbind_batch(cat)
    s1 = s1.cpu().numpy().view(abi.SUB_DTYPE)
    for f in abi.OUTPUT_FIELDS:
        assert np.array_equal(s1[f], got[0][f]), f
    assert np.array_equal(snap["id"][i1.cpu().numpy()[: len(got[1])]], as_ids(cat.cols["id"], got[1]))
    assert np.array_equal(e1.cpu().numpy()[: len(got[2])].view(np.uint32), got[2].view(np.uint32))
    nconv = check_unbind_invariants(cat, got)
    sel = np.sort(np.random.default_rng(15).choice(len(cat.subs), 500, replace=False))
    part = cat.subset(sel)
    rs, ri, re_ = oracle.unbind_batch(part, nthreads=16)
    stats = assert_tolerant((rs, as_ids(part.cols["id"], ri), re_), (got[0][sel], as_ids(cat.cols["id"], got[1]), got[2]))
    print(f"C5/10: {nconv} of {len(cat.subs)} converged, sample {stats}")


def test_indexed_batch_c5_full_size(gpu):
    """The whole one-rank share, as the bench runs it (200,000 subhalos, 64.3 M
    listed particles in a shuffled 70.7 M-particle snapshot): IDs resolved on
    the device and unbound from the snapshot; the size-free invariants over
    every subhalo (lists mapped back to catalogue positions), the tolerant
    criteria against the oracle on 300 subhalos drawn at random, and a second
    call byte-identical"""
    snap, ids, cat = synth.config5()
    U = gpu(cat.params)
    h = U.upload_snapshot(snap, ids, cat.subs)
    idx = U.resolve_ids(h)
    pos = idx.cpu().numpy()
    assert np.array_equal(snap["id"][pos], ids)
    s1, i1, e1 = U.unbind_indexed(h, idx)
    subs = s1.cpu().numpy().view(abi.SUB_DTYPE).copy()
    out_i = i1.cpu().numpy()
    out_e = e1.cpu().numpy()
    # snapshot position -> catalogue position (the catalogue holds the listed
    # particles in list order, each once)
    inv = np.full(len(snap["id"]), -1, np.int64)
    inv[pos] = np.arange(len(pos))
    cat_i = np.where(out_i >= 0, inv[np.where(out_i >= 0, out_i, 0)], -1)
    got = (subs, cat_i, out_e)
    nconv = check_unbind_invariants(cat, got)
    sel = np.sort(np.random.default_rng(25).choice(len(cat.subs), 300, replace=False))
    part = cat.subset(sel)
    rs, ri, re_ = oracle.unbind_batch(part, nthreads=16)
    stats = assert_tolerant((rs, as_ids(part.cols["id"], ri), re_), (subs[sel], as_ids(cat.cols["id"], cat_i), out_e))
    s2, i2, e2 = U.unbind_indexed(h, U.resolve_ids(h))
    assert np.array_equal(s2.cpu().numpy().view(np.uint8), s1.cpu().numpy().view(np.uint8))
    assert np.array_equal(i2.cpu().numpy(), out_i)
    assert np.array_equal(e2.cpu().numpy().view(np.uint32), out_e.view(np.uint32))
    print(f"C5 full: {nconv} of {len(cat.subs)} converged, sample {stats}")


def test_instrumentation_counts_and_outputs(gpu):
    """hbt_profile modes 1 (events around the walk and the sorts) and 2 (the
    counting walk and sort counts) leave the outputs unchanged, and their
    counts hang together: every walk launch timed, interactions and opened
    nodes within the lane slots the groups spent, those within the wave steps
    (64 lanes each), the targets equal to the evaluations the records count,
    and sorts of the pass's elements with at least one merge level for lists
    above one tile"""
    cat = synth.config2(seed=66, nsub=300, hi=20000)
    U = gpu(cat.params)
    base = U.unbind_batch(cat)
    passes = U.last_stats()[2]
    U.profile(1)
    got = U.unbind_batch(cat)
    t, srt = U.profile_read(), U.profile_sorts()
    assert _same_bytes(base, got)
    assert t["walk_launches"] == passes and t["walk_ms"] > 0
    assert srt["sort_calls"] >= passes and srt["sort_ms"] > 0
    U.profile(2)
    got = U.unbind_batch(cat)
    c, cs = U.profile_read(), U.profile_sorts()
    U.profile(0)
    assert _same_bytes(base, got)
    assert c["targets"] == int(base[0]["n_evals"].sum())
    assert 0 < c["interactions"] + c["opened"] <= c["lane_slots"] <= 64 * c["wave_steps"]
    assert c["tree_items"] > 0
    assert cs["elements"] >= int(cat.subs["count"].sum())
    big = cat.subs["count"] > 1024
    assert (cs["element_levels"] > 0) == bool(big.any())
    assert _same_bytes(base, U.unbind_batch(cat))


def test_particle_type_outside_the_table_is_an_error(gpu):
    """the per-type bound counts index by type: a type outside [0, 6) is
    rejected before any unbinding"""
    cat = synth.config2(seed=64, nsub=20, hi=2000)
    cat.cols["type"] = cat.cols["type"].copy()
    cat.cols["type"][5] = 6
    with pytest.raises(RuntimeError, match="particle types outside"):
        gpu(cat.params).unbind_batch(cat)

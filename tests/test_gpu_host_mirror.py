"""The C++ host mirror (hbtplus_amd/host/hbt_host.hpp), driven by a C++
program the way a reference caller would (AoS Particle_t lists in Subhalo_t
objects), against the reference's outputs in the golden fixtures."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from fixtures import load_unbind
from hbtplus_amd import abi
from hbtplus_amd.abi import PARAMS_DTYPE, SUB_DTYPE
from helpers import canonical_order, list_of

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "hbt_host_test")


def run_host(cat, mode, host_group=None):
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.ascontiguousarray(cat.params, PARAMS_DTYPE).tobytes())
            f.write(np.array([len(cat.subs), cat.n_particles], np.int64).tobytes())
            f.write(np.ascontiguousarray(cat.cols["id"], np.int64).tobytes())
            for k in ("x", "y", "z", "vx", "vy", "vz", "mass"):
                f.write(np.ascontiguousarray(cat.cols[k], np.float32).tobytes())
            f.write(np.ascontiguousarray(cat.cols["type"], np.int32).tobytes())
            f.write(np.ascontiguousarray(cat.subs, SUB_DTYPE).tobytes())
            if host_group is not None:
                f.write(np.ascontiguousarray(host_group, np.int64).tobytes())
        subprocess.run([EXE, fin, fout, mode], check=True, timeout=300)
        raw = open(fout, "rb").read()
    out, pos = [], 0
    for _ in range(len(cat.subs)):
        nb, npart = np.frombuffer(raw, np.int64, 2, pos)
        pos += 16
        mb = np.frombuffer(raw, np.float32, 7, pos)
        pos += 28
        sink = int(np.frombuffer(raw, np.int64, 1, pos)[0])
        death, sink_snap = (int(v) for v in np.frombuffer(raw, np.int32, 2, pos + 8))
        pos += 16
        ids = np.frombuffer(raw, np.int64, int(npart), pos)
        pos += 8 * int(npart)
        out.append((int(nb), mb, ids) if host_group is None else (int(nb), mb, ids, sink, death, sink_snap))
    return out


def test_host_driver_is_built():
    assert os.path.exists(EXE), "build with __graft_entry__.build()"


@pytest.mark.gpu
@pytest.mark.parametrize("mode,name", [("batch", "c4_small"), ("plain", "c4_small"), ("batch", "samp_c4"),
                                       ("plain", "samp_c2")])
def test_host_mirror_refine_particles(mode, name):
    """RefineParticles through the mirror, REPLAY, against the reference's
    outputs; samp_*: sampling on (MaxSampleSizeOfPotentialEstimate and
    RefineMostboundParticle mapped from the configuration, the shuffles
    drawn from the context's stream from srand(1))"""
    cat, (rs, ri, re_) = load_unbind(name)
    cat.params["precision"] = abi.HBT_PREC_REPLAY
    got = run_host(cat, mode)
    ids = cat.cols["id"]
    where = {int(g): i for i, g in enumerate(ids.tolist())}
    for s, (nb, mb, gid) in enumerate(got):
        assert nb == rs["nbound"][s]
        assert mb[0] == rs["mbound"][s]
        assert np.array_equal(mb[1:4], rs["ref_pos"][s]) and np.array_equal(mb[4:7], rs["ref_vel"][s])
        ref_idx, ref_e = list_of(rs, ri, s), list_of(rs, re_, s)
        e_of = dict(zip(ref_idx.tolist(), ref_e.tolist()))
        ns = int(rs["nsource"][s])
        assert len(gid) == ns
        got_idx = np.array([where[int(g)] for g in gid.tolist()], np.int64)
        got_e = np.array([e_of[i] for i in got_idx.tolist()], np.float32)
        # exactly tied energies may come in either order (std::sort)
        assert np.array_equal(canonical_order(got_idx, got_e), canonical_order(ref_idx, ref_e)[:ns])


@pytest.mark.gpu
def test_host_mirror_refine_then_merge():
    """RefineParticles then MergeSubhalos through the mirror against the
    reference's own RecursiveUnbind/TruncateSource followed by its
    SubhaloSnapshot_t::MergeSubhalos on the same catalogue (fixture
    merge_a, pipe_* outputs)"""
    from fixtures import load_merge
    from hbtplus_amd import synth
    F = load_merge("merge_a")
    cat, grp = synth.merging_catalogue(seed=1007, periodic=True)
    assert np.array_equal(cat.cols["id"], F.cols["id"]) and np.array_equal(grp, F.host_group)
    cat.params["precision"] = abi.HBT_PREC_REPLAY
    got = run_host(cat, "merge", grp)
    rs, rl, re_ = F.pipe
    ids = cat.cols["id"]
    for s, (nb, mb, gid, sink, death, ssnap) in enumerate(got):
        o, c = int(rs["offset"][s]), int(rs["count"][s])
        assert nb == rs["nbound"][s] and len(gid) == c, s
        assert mb[0] == rs["mbound"][s]
        assert np.array_equal(mb[1:4], rs["ref_pos"][s]) and np.array_equal(mb[4:7], rs["ref_vel"][s])
        assert (sink, death, ssnap) == (rs["sink_track_id"][s], rs["snapshot_index_of_death"][s],
                                       rs["snapshot_index_of_sink"][s]), s
        ref_ids = ids[rl[o:o + c]]
        e = re_[o:o + c]
        if np.isfinite(e).any():  # unbound again: exactly tied energies in any order
            e_of = dict(zip(ref_ids.tolist(), e.tolist()))
            ge = np.array([e_of.get(int(g), np.nan) for g in gid.tolist()], np.float32)
            assert np.array_equal(canonical_order(np.asarray(gid), ge), canonical_order(ref_ids, e)), s
        else:
            assert np.array_equal(gid, ref_ids), s

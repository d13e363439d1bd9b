"""A hand-built merging configuration whose trap decisions and merged lists
follow from the definitions in src/subhalo_merge.cpp by hand.

Host H (central of group 0) holds three unit-mass particles at x = 0, 1, -1
with velocities 0, 1, -1 along x: its core is at the origin at rest with
dispersions sigma_R = sigma_V = sqrt(2/3) (BuildPosition/BuildVelocity,
:29-127), so a subhalo whose most bound particle sits at distance d with
relative speed v is trapped by H iff (d + v) / sqrt(2/3) < 2, i.e. d + v <
1.63299.  Nested in H (index order): A (d + v = 0.8, trapped), B (1.7, not
trapped), O (an orphan, 0.1: trapped though too small to merge) and, nested
in O, C (0.2: O has Nbound <= 1 and is skipped, so C's sink is H).  D, nested
in B, sits on B's core (B: particles at x = 5, 5.2, velocities 0, 0.2; core
x = 5.1, v = 0.1, dispersions 0.1) and is trapped by B.  E is the group's
second head (glued into H's nest) at H's core.  F is a field halo with G
nested in it and trapped by it.  P was trapped before (dead, skipped).  Q is
a field halo at rest (sigma_V = 0): for R, nested in Q on its core, 0/0 is
NaN and R is not trapped.

With merging on (G large, MinNumPartOfSub 2, so every list stays bound):
post-order from H merges A, then D into B, then C, then E into H; H's list
becomes its own [0 1 2] plus A's new IDs [3 4] (A's copy of column 0 is
dropped), C's [7 8] and E's new IDs [11 12] (E's copy of column 1 is
dropped).
Every merged subhalo keeps its first particle; H and B (IsMerged) and F
(IsMerged, but a field halo, so nothing merged into it) are unbound again."""
import numpy as np

from hbtplus_amd.abi import SUB_DTYPE, make_params

#              H         A           B       O    C         D          E            F            G
LISTS = [[0, 1, 2], [3, 4, 0], [5, 6], [13], [7, 8], [9, 10], [11, 12, 1], [14, 15, 16], [17, 18],
         # P          Q             R
         [19, 20], [21, 22, 23], [24, 25]]
PARENT = [-1, 0, 0, 0, 3, 2, -1, -1, 7, 0, -1, 10]
GROUP = [0, 0, 0, 0, 0, 0, 0, -1, -1, 0, -1, -1]
NBOUND = [3, 3, 2, 1, 2, 2, 3, 3, 2, 1, 3, 2]
H, A, B, O, C, D, E, F, G, P, Q, R = range(12)
SNAP = 6


def hand_merge_case(G=1.0e6):
    n = 26
    x = np.zeros((n, 3), np.float32)
    v = np.zeros((n, 3), np.float32)
    x[0:3, 0] = [0, 1, -1]
    v[0:3, 0] = [0, 1, -1]
    x[3:5, 0] = [0.5, 0.6]
    v[3:5, 0] = [0.3, 0.3]
    x[5:7, 0] = [5.0, 5.2]
    v[5:7, 0] = [0.0, 0.2]
    x[7:9, 0] = [0.2, 0.25]
    x[9:11, 0] = [5.1, 5.12]
    v[9:11, 0] = [0.1, 0.1]
    x[11:13, 0] = [0.0, 0.05]
    x[13, 0] = 0.1
    x[14:17] = [[50, 0, 0], [51, 0, 0], [49, 0, 0]]
    v[14:17, 1] = [0, 1, -1]
    x[17:19] = [[50, 0, 0], [50.1, 0, 0]]
    x[19:21, 0] = [0.0, 0.1]
    x[21:24] = [[0, 80, 0], [1, 80, 0], [-1, 80, 0]]
    x[24:26] = [[0, 80, 0], [0.1, 80, 0]]
    cols = dict(id=np.arange(1000, 1000 + n, dtype=np.int64) * 7, x=x[:, 0].copy(), y=x[:, 1].copy(),
                z=x[:, 2].copy(), vx=v[:, 0].copy(), vy=v[:, 1].copy(), vz=v[:, 2].copy(),
                mass=np.ones(n, np.float32), type=np.ones(n, np.int32))
    lst = np.concatenate([np.array(l, np.int64) for l in LISTS])
    subs = np.zeros(len(LISTS), SUB_DTYPE)
    off = 0
    for s, l in enumerate(LISTS):
        subs["offset"][s], subs["count"][s] = off, len(l)
        off += len(l)
    subs["parent"] = PARENT
    subs["nbound"] = NBOUND
    subs["nbound_prev"] = NBOUND
    subs["sink_track_id"] = -1
    subs["snapshot_index_of_death"] = -1
    subs["snapshot_index_of_sink"] = -1
    subs["mbound"] = np.array(NBOUND, np.float32)
    # most bound particle: the first of the list, except where the trap test
    # is set by hand (B sits at d + v = 1.7 from H's core)
    for s, l in enumerate(LISTS):
        subs["mostbound_pos"][s] = x[l[0]]
        subs["mostbound_vel"][s] = v[l[0]]
        subs["ref_pos"][s] = x[l].mean(axis=0)
        subs["ref_vel"][s] = v[l].mean(axis=0)
    subs["mostbound_pos"][B] = [1.0, 0, 0]
    subs["mostbound_vel"][B] = [0.7, 0, 0]
    subs["snapshot_index_of_death"][P] = SNAP - 2
    subs["sink_track_id"][P] = H
    subs["snapshot_index_of_sink"][P] = SNAP - 2
    params = make_params(G=G, softening=0.01, periodic=0, min_num_part_of_sub=2, snapshot_index=SNAP)
    return params, cols, lst, subs, np.array(GROUP, np.int64)


EXPECT_SINK = {A: H, B: -1, O: H, C: H, D: B, E: H, G: F, P: H, H: -1, F: -1, Q: -1, R: -1}
EXPECT_MERGED_AWAY = {A: 3, C: 7, D: 9, E: 11}
EXPECT_UNION = {H: {0, 1, 2, 3, 4, 7, 8, 11, 12}, B: {5, 6, 9, 10}, F: {14, 15, 16}}


def check_hand_merge(subs_in, out, lst, merge_trapped):
    """The outcome the module docstring derives."""
    lists = [lst[int(o):int(o) + int(c)] for o, c in zip(out["offset"], out["count"])]
    for s, t in EXPECT_SINK.items():
        assert out["sink_track_id"][s] == t, (s, out["sink_track_id"][s], t)
        newly = subs_in["sink_track_id"][s] == -1 and t != -1
        assert out["snapshot_index_of_sink"][s] == (SNAP if newly else subs_in["snapshot_index_of_sink"][s])
    if not merge_trapped:
        for s in range(len(out)):
            assert out["snapshot_index_of_death"][s] == subs_in["snapshot_index_of_death"][s]
            assert out["count"][s] == subs_in["count"][s] and out["nbound"][s] == subs_in["nbound"][s]
        return
    for s, first in EXPECT_MERGED_AWAY.items():
        assert list(lists[s]) == [first] and out["nbound"][s] == 1 and out["mbound"][s] == 1.0
        assert out["snapshot_index_of_death"][s] == SNAP
    assert out["snapshot_index_of_death"][O] == SNAP and out["count"][O] == 1
    assert out["snapshot_index_of_death"][P] == SNAP - 2
    for s, ids in EXPECT_UNION.items():
        assert set(lists[s].tolist()) == ids and len(lists[s]) == len(ids), (s, lists[s])
        assert out["nbound"][s] == len(ids)
    assert list(lists[G]) == [17, 18] and out["snapshot_index_of_death"][G] == -1
    assert list(lists[R]) == [24, 25] and list(lists[Q]) == [21, 22, 23]

// hbt_merge.hpp -- hbt_merge_subhalos: merging of trapped subhalos after the
// unbinding of a snapshot (SubhaloSnapshot_t::MergeSubhalos,
// src/subhalo_merge.cpp:187-232, Subhalo_t::MergeTo :244-262).  Included by
// hbt_unbind.hip.
//
// On the device:
//   merge_core_kernel  per subhalo, the core of its (at most 20) most bound
//                      particles: mass-weighted mean position (periodic about
//                      the first) and velocity with their dispersions, FP64
//                      sums in list order (SubHelper_t::BuildPosition /
//                      BuildVelocity, :29-127)
//   merge_trap_kernel  per subhalo not yet trapped, up its chain of hosts:
//                      the first host with Nbound > 1 whose core lies within
//                      phase-space distance 2 of its most bound particle
//                      becomes its sink (DetectTraps, SinkDistance, :129-161)
//   union: for every subhalo that receives merged subhalos, its own list
//          followed by the lists of what merges into it (post-order, the
//          order MergeRecursive calls MergeTo) is sorted by (ID, position);
//          an entry stays if it belongs to the receiver's own list or is the
//          first of its ID (MergeTo's unordered_set, :249-255), then the
//          kept entries are compacted in list order
//   re-unbinding of the merged hosts through hbt_unbind_batch_indexed
//          (plain Unbind semantics, then TruncateSource, :207-214)
// The host walks the nests in post-order (MergeRecursive, :222-232) to decide
// which lists merge where; that bookkeeping is O(subhalos).
#pragma once

namespace {

constexpr int MERGE_CORE_MAX = 20;       // NumPartCoreMax, src/subhalo_merge.cpp:11
constexpr float MERGE_DELTA_CRIT = 2.f;  // DeltaCrit, src/subhalo_merge.cpp:12

struct MergeIn {
  int64_t offset;   // list start in list_index
  int64_t nbound;   // Nbound after unbinding
  int64_t host;     // HostTrackId: the subhalo whose (glued) nest holds this one, -1 none
  int64_t sink;     // SinkTrackId on entry
  float mb_pos[3];  // ComovingMostBoundPosition
  float mb_vel[3];  // PhysicalMostBoundVelocity
};

struct MergeCore {
  float pos[3], vel[3];
  float sig_r, sig_v;
};

struct MergeCols {
  const int64_t *list;
  const int64_t *id;
  const float *x, *y, *z, *vx, *vy, *vz, *mass;
  const int32_t *type;
};

// SubHelper_t::BuildPosition / BuildVelocity (src/subhalo_merge.cpp:29-127).
// The list's first entry is the most bound particle.
__global__ void __launch_bounds__(BLOCK) merge_core_kernel(DevParams P, const MergeIn *__restrict__ in,
                                                           int64_t nsub, MergeCols C,
                                                           MergeCore *__restrict__ core) {
#pragma clang fp contract(off)
  for (int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x; s < nsub; s += (int64_t)gridDim.x * BLOCK) {
    const MergeIn m = in[s];
    MergeCore c;
    c.sig_r = 0.f;
    c.sig_v = 0.f;
    for (int j = 0; j < 3; j++) { c.pos[j] = 0.f; c.vel[j] = 0.f; }
    if (m.nbound == 1) {
      const int64_t p = C.list[m.offset];
      c.pos[0] = C.x[p]; c.pos[1] = C.y[p]; c.pos[2] = C.z[p];
      c.vel[0] = C.vx[p]; c.vel[1] = C.vy[p]; c.vel[2] = C.vz[p];
    } else if (m.nbound > 1) {
      const int n = (int)(m.nbound < MERGE_CORE_MAX ? m.nbound : MERGE_CORE_MAX);
      const int64_t p0 = C.list[m.offset];
      // the periodic origin is held in double: the differences and NEAREST
      // are evaluated in double, the mean is rounded to float and the origin
      // added in double (float += double)
      const double org[3] = {C.x[p0], C.y[p0], C.z[p0]};
      double sx[3] = {0., 0., 0.}, sx2[3] = {0., 0., 0.}, msum = 0.;
      double sv[3] = {0., 0., 0.}, sv2[3] = {0., 0., 0.};
      for (int i = 0; i < n; i++) {
        const int64_t p = C.list[m.offset + i];
        const float mf = C.mass[p];
        const double md = (double)mf;
        msum += md;
        const float xp[3] = {C.x[p], C.y[p], C.z[p]};
        const float vp[3] = {C.vx[p], C.vy[p], C.vz[p]};
        for (int j = 0; j < 3; j++) {
          const double dx = P.periodic ? nearest_d((double)xp[j] - org[j], P) : (double)xp[j];
          sx[j] += dx * md;
          sx2[j] += dx * dx * md;
          const double dv = (double)vp[j];
          sv[j] += dv * md;
          sv2[j] += dv * dv * md;
        }
      }
      for (int j = 0; j < 3; j++) {
        sx[j] /= msum;
        sx2[j] /= msum;
        c.pos[j] = (float)sx[j];
        if (P.periodic) c.pos[j] = (float)((double)c.pos[j] + org[j]);
        sx2[j] -= sx[j] * sx[j];
        sv[j] /= msum;
        sv2[j] /= msum;
        c.vel[j] = (float)sv[j];
        sv2[j] -= sv[j] * sv[j];
      }
      c.sig_r = (float)sqrt(sx2[0] + sx2[1] + sx2[2]);
      c.sig_v = (float)sqrt(sv2[0] + sv2[1] + sv2[2]);
    }
    core[s] = c;
  }
}

// DetectTraps (src/subhalo_merge.cpp:136-161) with SinkDistance (:129-134).
// sink_out[i] = the new sink or the entry value; merged[h] = 1 when a
// subhalo with Nbound > 1 sinks into h (IsMerged).  The host chain is
// acyclic (checked on the host); the step bound is a second guard.
__global__ void __launch_bounds__(BLOCK) merge_trap_kernel(DevParams P, const MergeIn *__restrict__ in,
                                                           const MergeCore *__restrict__ core, int64_t nsub,
                                                           int64_t *__restrict__ sink_out,
                                                           int *__restrict__ merged) {
#pragma clang fp contract(off)
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nsub; i += (int64_t)gridDim.x * BLOCK) {
    const MergeIn m = in[i];
    sink_out[i] = m.sink;
    if (m.sink != -1) continue;
    int64_t h = m.host;
    for (int64_t step = 0; h >= 0 && h < nsub && step < nsub; step++) {
      if (in[h].nbound > 1) {
        const MergeCore c = core[h];
        const float d = periodic_distance_f(P, c.pos, m.mb_pos[0], m.mb_pos[1], m.mb_pos[2]);
        const float e0 = c.vel[0] - m.mb_vel[0], e1 = c.vel[1] - m.mb_vel[1], e2 = c.vel[2] - m.mb_vel[2];
        const float v = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
        const float delta = d / c.sig_r + v / c.sig_v;
        if (delta < MERGE_DELTA_CRIT) {
          sink_out[i] = h;
          if (m.nbound > 1) merged[h] = 1;
          break;
        }
      }
      h = in[h].host;
    }
  }
}

// One own list inside a union segment.
struct MergePiece {
  int64_t offset;  // in list_index
  int64_t count;
  int seg;         // union segment
  int pstart;      // position of its first entry in the segment
};

struct MergeUnion {
  const MergePiece *piece;
  int npiece;
  const int64_t *gbase;  // first slot of each segment
  const int *gn;         // entries per segment
  const int *own;        // per segment: length of the receiver's own list
  const int64_t *ubase;  // per segment: start of its kept entries in ubuf
  int64_t nslots;
  const int *seg_of_slot;
  const int64_t *list;
  const int64_t *id;
  u64 *k0;
  u32 *v0;
  unsigned char *keep;
  u64 *scan;
  int64_t *ubuf;
  int64_t *kept;         // per segment
  int nseg;
};

__global__ void __launch_bounds__(BLOCK) merge_gather_kernel(MergeUnion U) {
  for (int q = blockIdx.x; q < U.npiece; q += gridDim.x) {
    const MergePiece pc = U.piece[q];
    const int64_t b = U.gbase[pc.seg] + pc.pstart;
    for (int64_t l = threadIdx.x; l < pc.count; l += BLOCK) {
      U.k0[b + l] = (u64)U.id[U.list[pc.offset + l]] ^ (1ull << 63);
      U.v0[b + l] = (u32)(pc.pstart + l);
    }
  }
}

// sorted position i of a segment: the entry stays if it is in the receiver's
// own list or its ID differs from position i-1
__global__ void __launch_bounds__(BLOCK) merge_keep_kernel(MergeUnion U) {
  for (int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x; s < U.nslots; s += (int64_t)gridDim.x * BLOCK) {
    const int g = U.seg_of_slot[s];
    if (g < 0) continue;
    const int64_t b = U.gbase[g];
    const int i = (int)(s - b);
    if (i >= U.gn[g]) continue;
    const u32 pos = U.v0[s];
    const bool first = i == 0 || U.k0[s - 1] != U.k0[s];
    U.keep[b + pos] = (first || (int)pos < U.own[g]) ? 1 : 0;
  }
}

__global__ void __launch_bounds__(BLOCK) merge_write_kernel(MergeUnion U) {
  for (int q = blockIdx.x; q < U.npiece; q += gridDim.x) {
    const MergePiece pc = U.piece[q];
    const int64_t b = U.gbase[pc.seg] + pc.pstart;
    for (int64_t l = threadIdx.x; l < pc.count; l += BLOCK)
      if (U.keep[b + l]) U.ubuf[U.ubase[pc.seg] + (int64_t)U.scan[b + l]] = U.list[pc.offset + l];
  }
}

__global__ void merge_kept_kernel(MergeUnion U) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < U.nseg; g += gridDim.x * blockDim.x) {
    if (U.gn[g] <= 0) { U.kept[g] = 0; continue; }
    const int64_t last = U.gbase[g] + U.gn[g] - 1;
    U.kept[g] = (int64_t)U.scan[last] + (int64_t)U.keep[last];
  }
}

struct MergeCopy {
  const int64_t *src;
  int64_t src_off, count, dst_off;
};

__global__ void __launch_bounds__(BLOCK) merge_copy_kernel(const MergeCopy *__restrict__ r, int64_t nr,
                                                           int64_t *__restrict__ dst) {
  for (int64_t q = blockIdx.x; q < nr; q += gridDim.x) {
    const MergeCopy c = r[q];
    for (int64_t l = threadIdx.x; l < c.count; l += BLOCK) dst[c.dst_off + l] = c.src[c.src_off + l];
  }
}

// mass and type of a list's first entry (CountParticles of a one-particle list)
__global__ void merge_first_kernel(const int64_t *__restrict__ list, const int64_t *__restrict__ at, int64_t n,
                                   const float *__restrict__ mass, const int32_t *__restrict__ type,
                                   float *__restrict__ m_out, int32_t *__restrict__ t_out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = list[at[i]];
    m_out[i] = mass[p];
    t_out[i] = type[p];
  }
}

}  // namespace

extern "C" int hbt_merge_subhalos(hbt_ctx *ctx, const hbt_particles *snapshot, const int64_t *list_index,
                                  int64_t n_list, hbt_subhalo *subs_io, const int64_t *host_group,
                                  const int64_t *visit_rank, int64_t nsub, int merge_trapped, int64_t *out_list,
                                  int64_t out_capacity, int64_t *out_total) {
  if (!ctx || !snapshot || n_list < 0 || nsub < 0 || out_capacity < 0 ||
      (nsub > 0 && (!subs_io || !host_group || !list_index || !out_list)))
    return fail(HBT_ERR_INVALID_ARG, "bad argument");
  if (out_total) *out_total = 0;
  if (nsub == 0) return HBT_OK;
  if (nsub >= (int64_t)INT32_MAX || n_list >= (int64_t)INT32_MAX)
    return fail(HBT_ERR_UNSUPPORTED, "too many subhalos or list entries");
  const hbt_particles &S = *snapshot;
  if (!S.id || !S.x || !S.y || !S.z || !S.vx || !S.vy || !S.vz || !S.mass || !S.type)
    return fail(HBT_ERR_INVALID_ARG, "every particle column is required");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int isnap = ctx->params.snapshot_index;
  const DevParams P = make_dev_params(ctx->params);
  // the records are worked on in a copy: a call that fails leaves them as they were
  std::vector<hbt_subhalo> work(subs_io, subs_io + nsub);
  hbt_subhalo *subs_host = work.data();

  // ---- host checks, nests and glued hosts ----------------------------------
  std::vector<std::vector<int>> nested(nsub);
  std::unordered_map<int64_t, int> gidx;
  std::vector<std::vector<int>> heads;
  for (int64_t s = 0; s < nsub; s++) {
    const hbt_subhalo &r = subs_host[s];
    if (r.count < 0 || r.offset < 0 || r.offset + r.count > n_list)
      return fail(HBT_ERR_INVALID_ARG, "subhalo list outside list_index");
    if (r.nbound < 0 || r.nbound > r.count) return fail(HBT_ERR_INVALID_ARG, "nbound outside [0, count]");
    if (r.sink_track_id != -1 && r.snapshot_index_of_death == -1 && r.nbound > 1)
      return fail(HBT_ERR_INVALID_ARG,
                  "a trapped subhalo that is alive with nbound > 1 (unbinding clears the trap of a subhalo "
                  "that stays bound and marks the others dead)");
    const int64_t p = r.parent;
    if (p >= 0) {
      if (p >= nsub || p == s || host_group[p] != host_group[s])
        return fail(HBT_ERR_INVALID_ARG, "a nested subhalo must share its parent's host group");
      nested[p].push_back((int)s);
    } else if (host_group[s] >= 0) {
      auto it = gidx.find(host_group[s]);
      if (it == gidx.end()) {
        gidx.emplace(host_group[s], (int)heads.size());
        heads.push_back({(int)s});
      } else {
        heads[it->second].push_back((int)s);
      }
    }
  }
  // siblings in the caller's visit order (the reference's SubGroups order for
  // a group's heads, its central first; NestedSubhalos order for nested ones)
  if (visit_rank) {
    auto by_rank = [&](int a, int b) { return visit_rank[a] != visit_rank[b] ? visit_rank[a] < visit_rank[b] : a < b; };
    for (auto &v : nested) std::sort(v.begin(), v.end(), by_rank);
    for (auto &v : heads) std::sort(v.begin(), v.end(), by_rank);
  }
  std::vector<int64_t> host(nsub, -1);
  for (int64_t s = 0; s < nsub; s++) host[s] = subs_host[s].parent;
  for (const auto &h : heads)
    for (size_t k = 1; k < h.size(); k++) host[h[k]] = h[0];
  // every subhalo reached exactly once from the tops of the nests: no cycles
  {
    std::vector<char> seen(nsub, 0);
    std::vector<int> stack;
    int64_t reached = 0;
    for (int64_t s = 0; s < nsub; s++) {
      if (subs_host[s].parent >= 0) continue;
      stack.push_back((int)s);
      while (!stack.empty()) {
        const int u = stack.back();
        stack.pop_back();
        if (seen[u]) return fail(HBT_ERR_INVALID_ARG, "nesting has a cycle");
        seen[u] = 1;
        reached++;
        for (int c : nested[u]) stack.push_back(c);
      }
    }
    if (reached != nsub) return fail(HBT_ERR_INVALID_ARG, "nesting has a cycle");
  }

  // every list entry must be a snapshot position
  if (n_list > 0) {
    RETCHK(ctx->active.ensure(16));
    int zero = 0, bad = 0;
    HIPCHK(hipMemcpyAsync(ctx->active.p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    check_index_kernel<<<grid_for(n_list, ctx->cus), BLOCK, 0, st>>>(list_index, n_list, S.n, ctx->active.as<int>());
    HIPCHK(hipMemcpyAsync(&bad, ctx->active.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (bad) return fail(HBT_ERR_INVALID_ARG, std::to_string(bad) + " list entries are not snapshot positions");
  }

  // ---- cores and traps on the device ----------------------------------------
  std::vector<MergeIn> mins(nsub);
  for (int64_t s = 0; s < nsub; s++) {
    const hbt_subhalo &r = subs_host[s];
    MergeIn &m = mins[s];
    m.offset = r.offset;
    m.nbound = r.nbound;
    m.host = host[s];
    m.sink = r.sink_track_id;
    for (int j = 0; j < 3; j++) { m.mb_pos[j] = r.mostbound_pos[j]; m.mb_vel[j] = r.mostbound_vel[j]; }
  }
  DBuf d_in, d_core, d_sink, d_merged;
  auto release = [&]() { d_in.release(); d_core.release(); d_sink.release(); d_merged.release(); };
  int rc = HBT_OK;
  if ((rc = d_in.ensure(nsub * sizeof(MergeIn))) != HBT_OK || (rc = d_core.ensure(nsub * sizeof(MergeCore))) != HBT_OK ||
      (rc = d_sink.ensure(nsub * 8)) != HBT_OK || (rc = d_merged.ensure(nsub * 4)) != HBT_OK) {
    release();
    return rc;
  }
  std::vector<int64_t> sink(nsub);
  std::vector<int> merged(nsub);
  MergeCols C{list_index, S.id, S.x, S.y, S.z, S.vx, S.vy, S.vz, S.mass, S.type};
  {
    hipError_t e = hipMemcpyAsync(d_in.p, mins.data(), nsub * sizeof(MergeIn), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_merged.p, 0, nsub * 4, st);
    if (e == hipSuccess) {
      const int g = grid_for(nsub, ctx->cus);
      merge_core_kernel<<<g, BLOCK, 0, st>>>(P, d_in.as<MergeIn>(), nsub, C, d_core.as<MergeCore>());
      merge_trap_kernel<<<g, BLOCK, 0, st>>>(P, d_in.as<MergeIn>(), d_core.as<MergeCore>(), nsub,
                                             d_sink.as<int64_t>(), d_merged.as<int>());
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(sink.data(), d_sink.p, nsub * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(merged.data(), d_merged.p, nsub * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      release();
      return fail(HBT_ERR_DEVICE, std::string("merge_subhalos: ") + hipGetErrorString(e));
    }
  }
  release();
  for (int64_t s = 0; s < nsub; s++) {
    if (sink[s] != subs_host[s].sink_track_id) {
      subs_host[s].sink_track_id = sink[s];
      subs_host[s].snapshot_index_of_sink = isnap;
    }
  }

  // ---- MergeRecursive: post-order from each group's central ----------------
  // pieces[x]: the own lists whose concatenation is x's list before the
  // duplicates are dropped; x_first[x]: x merged away and keeps its first entry
  std::vector<std::vector<int>> pieces(nsub);
  std::vector<char> merged_away(nsub, 0);
  std::vector<int64_t> nbound_cur(nsub);
  for (int64_t s = 0; s < nsub; s++) {
    pieces[s].push_back((int)s);
    nbound_cur[s] = subs_host[s].nbound;
  }
  if (merge_trapped) {
    for (const auto &h : heads) {
      const int central = h[0];
      std::vector<std::pair<int, size_t>> stack{{central, 0}};
      while (!stack.empty()) {
        auto &top = stack.back();
        const int x = top.first;
        const std::vector<int> &kids = nested[x];
        const size_t nk = kids.size() + (x == central ? h.size() - 1 : 0);
        if (top.second < nk) {
          const size_t i = top.second++;
          stack.push_back({i < kids.size() ? kids[i] : h[1 + (i - kids.size())], 0});
          continue;
        }
        stack.pop_back();
        hbt_subhalo &r = subs_host[x];
        if (r.sink_track_id == -1 || r.snapshot_index_of_death != -1) continue;
        if (nbound_cur[x] > 1) {  // Subhalo_t::MergeTo
          const int64_t t = r.sink_track_id;
          if (t < 0 || t >= nsub) return fail(HBT_ERR_INVALID_ARG, "sink outside the subhalos");
          std::vector<int> &dst = pieces[t];
          dst.insert(dst.end(), pieces[x].begin(), pieces[x].end());
          pieces[x].assign(1, x);
          merged_away[x] = 1;
          nbound_cur[t] += nbound_cur[x];
          nbound_cur[x] = 1;
        }
        r.snapshot_index_of_death = isnap;
      }
    }
  }

  // ---- union segments -------------------------------------------------------
  std::vector<int> seg_of(nsub, -1);
  std::vector<MergePiece> pc;
  std::vector<int64_t> gbase, ubase;
  std::vector<int> gn, own, capal, first_tile, ntile;
  std::vector<Tile> tiles;
  int64_t slots = 0, ulen = 0;
  int maxn = 0;
  for (int64_t s = 0; s < nsub; s++) {
    if (pieces[s].size() < 2) continue;
    const int k = (int)gbase.size();
    seg_of[s] = k;
    int64_t pos = 0;
    for (int q : pieces[s]) {
      pc.push_back(MergePiece{subs_host[q].offset, subs_host[q].count, k, (int)pos});
      pos += subs_host[q].count;
      if (pos >= (int64_t)INT32_MAX / 2) return fail(HBT_ERR_UNSUPPORTED, "merged list too large");
    }
    gbase.push_back(slots);
    ubase.push_back(ulen);
    gn.push_back((int)pos);
    own.push_back((int)subs_host[s].count);
    const int ca = (int)((pos + WAVE - 1) / WAVE * WAVE);
    capal.push_back(ca);
    first_tile.push_back((int)tiles.size());
    const int nt = (int)((pos + TILE - 1) / TILE);
    ntile.push_back(nt);
    for (int t = 0; t < nt; t++) tiles.push_back(Tile{k, t * TILE});
    slots += ca;
    ulen += pos;
    maxn = std::max(maxn, (int)pos);
  }
  const int nseg = (int)gbase.size();
  DBuf ubuf, d_kept, pre, d_rec, d_oi, d_oe, d_ranges, d_first;
  auto release_all = [&]() {
    for (DBuf *b : {&ubuf, &d_kept, &pre, &d_rec, &d_oi, &d_oe, &d_ranges, &d_first}) b->release();
  };
  std::vector<int64_t> kept(nseg);
  if (nseg > 0) {
    if ((rc = ensure_slots(ctx, std::max<int64_t>(slots, 1), nseg, std::max<int64_t>((int64_t)tiles.size(), 1))) !=
            HBT_OK ||
        (rc = ubuf.ensure(std::max<int64_t>(ulen, 1) * 8)) != HBT_OK ||
        (rc = d_kept.ensure(nseg * 8)) != HBT_OK) {
      release_all();
      return rc;
    }
    void *d[8];
    rc = stage_blobs(ctx, {{pc.data(), pc.size() * sizeof(MergePiece)},
                           {gbase.data(), gbase.size() * 8},
                           {gn.data(), gn.size() * 4},
                           {own.data(), own.size() * 4},
                           {ubase.data(), ubase.size() * 8},
                           {capal.data(), capal.size() * 4},
                           {tiles.data(), tiles.size() * sizeof(Tile)},
                           {first_tile.data(), first_tile.size() * 4}},
                     d);
    if (rc != HBT_OK) { release_all(); return rc; }
    MergeUnion U{};
    U.piece = (const MergePiece *)d[0];
    U.npiece = (int)pc.size();
    U.gbase = (const int64_t *)d[1];
    U.gn = (const int *)d[2];
    U.own = (const int *)d[3];
    U.ubase = (const int64_t *)d[4];
    U.nslots = slots;
    U.seg_of_slot = ctx->seg_of_slot.as<int>();
    U.list = list_index;
    U.id = S.id;
    U.k0 = ctx->k0.as<u64>();
    U.v0 = ctx->v0.as<u32>();
    U.keep = ctx->isb.as<unsigned char>();
    U.scan = ctx->scan64.as<u64>();
    U.ubuf = ubuf.as<int64_t>();
    U.kept = d_kept.as<int64_t>();
    U.nseg = nseg;
    const Tile *dt = (const Tile *)d[6];
    const int nt = (int)tiles.size();
    const int gp = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)pc.size(), (int64_t)ctx->cus * 8));
    const int gt = (int)std::max<int64_t>(1, std::min<int64_t>(nt, (int64_t)ctx->cus * 8));
    const int gm = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)nt * (TILE / BLOCK), (int64_t)ctx->cus * 16));
    fill_seg_of_slot_kernel<<<std::min(nseg, ctx->cus * 8), BLOCK, 0, st>>>(nseg, U.gbase, (const int *)d[5],
                                                                           ctx->seg_of_slot.as<int>());
    merge_gather_kernel<<<gp, BLOCK, 0, st>>>(U);
    tile_sort_kernel<<<gt, BLOCK, 0, st>>>(dt, nt, U.gbase, U.gn, U.k0, U.v0, ctx->k1.as<u64>(), ctx->v1.as<u32>());
    int level = 0;
    for (int w = TILE; w < maxn; w *= 2, level++)
      merge_pass_kernel<<<gm, BLOCK, 0, st>>>(dt, nt, U.gbase, U.gn, level, U.k0, U.v0, ctx->k1.as<u64>(),
                                              ctx->v1.as<u32>());
    merge_keep_kernel<<<grid_for(slots, ctx->cus), BLOCK, 0, st>>>(U);
    u64 *ts = ctx->tsum.as<u64>();
    scan_tiles_kernel<u64><<<gt, BLOCK, 0, st>>>(dt, nt, U.gn, ts, KeepLoad{U.gbase, U.keep}, KeepStore{U.gbase, U.scan});
    scan_add_kernel<u64><<<gt, BLOCK, 0, st>>>(dt, nt, U.gn, ts, KeepAdd{U.gbase, U.scan});
    merge_write_kernel<<<gp, BLOCK, 0, st>>>(U);
    merge_kept_kernel<<<(nseg + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(U);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(kept.data(), d_kept.p, nseg * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      release_all();
      return fail(HBT_ERR_DEVICE, std::string("merge_subhalos: ") + hipGetErrorString(e));
    }
  }

  // ---- lists after MergeRecursive (before the re-unbinding) ----------------
  std::vector<int64_t> pre_off(nsub), pre_cnt(nsub);
  std::vector<MergeCopy> ranges(nsub);
  int64_t pre_total = 0;
  for (int64_t s = 0; s < nsub; s++) {
    const hbt_subhalo &r = subs_host[s];
    MergeCopy &c = ranges[s];
    if (seg_of[s] >= 0) {
      c = MergeCopy{ubuf.as<int64_t>(), ubase[seg_of[s]], kept[seg_of[s]], pre_total};
    } else {
      c = MergeCopy{list_index, r.offset, merged_away[s] ? 1 : r.count, pre_total};
    }
    pre_off[s] = pre_total;
    pre_cnt[s] = c.count;
    pre_total += c.count;
  }
  std::vector<int> reunbind;
  if (merge_trapped)
    for (int64_t s = 0; s < nsub; s++)
      if (merged[s]) reunbind.push_back((int)s);
  // the one-particle lists of merged-away subhalos that are not re-unbound
  std::vector<int64_t> first_at;
  std::vector<int> first_sub;
  for (int64_t s = 0; s < nsub; s++)
    if (merged_away[s] && !(merge_trapped && merged[s])) {
      first_at.push_back(pre_off[s]);
      first_sub.push_back((int)s);
    }
  std::vector<float> first_m(first_at.size());
  std::vector<int32_t> first_t(first_at.size());
  if ((rc = pre.ensure(std::max<int64_t>(pre_total, 1) * 8)) != HBT_OK ||
      (rc = d_ranges.ensure(nsub * sizeof(MergeCopy))) != HBT_OK ||
      (rc = d_first.ensure(std::max<size_t>(first_at.size(), 1) * 16)) != HBT_OK) {
    release_all();
    return rc;
  }
  {
    hipError_t e = hipMemcpyAsync(d_ranges.p, ranges.data(), nsub * sizeof(MergeCopy), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      merge_copy_kernel<<<(int)std::min<int64_t>(nsub, (int64_t)ctx->cus * 8), BLOCK, 0, st>>>(
          d_ranges.as<MergeCopy>(), nsub, pre.as<int64_t>());
      e = hipGetLastError();
    }
    const int64_t nf = (int64_t)first_at.size();
    if (e == hipSuccess && nf > 0) {
      int64_t *d_at = d_first.as<int64_t>();
      float *d_m = (float *)(d_at + nf);
      int32_t *d_t = (int32_t *)(d_m + nf);
      e = hipMemcpyAsync(d_at, first_at.data(), nf * 8, hipMemcpyHostToDevice, st);
      if (e == hipSuccess) {
        merge_first_kernel<<<grid_for(nf, ctx->cus), BLOCK, 0, st>>>(pre.as<int64_t>(), d_at, nf, S.mass, S.type,
                                                                     d_m, d_t);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipMemcpyAsync(first_m.data(), d_m, nf * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipMemcpyAsync(first_t.data(), d_t, nf * 4, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      release_all();
      return fail(HBT_ERR_DEVICE, std::string("merge_subhalos: ") + hipGetErrorString(e));
    }
  }
  ubuf.release();
  // MergeTo's bookkeeping: Nbound = 1 and CountParticles of the kept particle
  for (size_t i = 0; i < first_sub.size(); i++) {
    hbt_subhalo &r = subs_host[first_sub[i]];
    const int t = first_t[i];
    if (t < 0 || t >= HBT_TYPE_MAX) { release_all(); return fail(HBT_ERR_INVALID_ARG, "particle type out of range"); }
    r.nbound = 1;
    for (int k = 0; k < HBT_TYPE_MAX; k++) { r.nbound_type[k] = 0; r.mbound_type[k] = 0.f; }
    r.nbound_type[t] = 1;
    r.mbound_type[t] = first_m[i];
    float mb = 0.f;
    for (int k = 0; k < HBT_TYPE_MAX; k++) mb += r.mbound_type[k];
    r.mbound = mb;
  }
  for (int64_t s = 0; s < nsub; s++)
    if (!merged_away[s] && seg_of[s] >= 0) subs_host[s].nbound = nbound_cur[s];

  // ---- re-unbinding of the merged hosts, then TruncateSource ---------------
  const int64_t nr = (int64_t)reunbind.size();
  std::vector<hbt_subhalo> rec(nr);
  if (nr > 0) {
    for (int64_t k = 0; k < nr; k++) {
      const int s = reunbind[k];
      rec[k] = subs_host[s];
      rec[k].offset = pre_off[s];
      rec[k].count = pre_cnt[s];
      rec[k].parent = -1;
      rec[k].nbound_prev = INT64_MAX;  // plain Unbind: no orphan rule
    }
    int64_t total = 0;
    if ((rc = hbt_plan_batch(rec.data(), nr, &total)) != HBT_OK ||
        (rc = d_rec.ensure(nr * sizeof(hbt_subhalo))) != HBT_OK ||
        (rc = d_oi.ensure(std::max<int64_t>(total, 1) * 8)) != HBT_OK ||
        (rc = d_oe.ensure(std::max<int64_t>(total, 1) * 4)) != HBT_OK) {
      release_all();
      return rc;
    }
    hipError_t e = hipMemcpyAsync(d_rec.p, rec.data(), nr * sizeof(hbt_subhalo), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) { release_all(); return fail(HBT_ERR_DEVICE, hipGetErrorString(e)); }
    rc = hbt_unbind_batch_indexed(ctx, snapshot, pre.as<int64_t>(), pre_total, d_rec.as<hbt_subhalo>(), rec.data(),
                                  nr, d_oi.as<int64_t>(), d_oe.as<float>(), 0);
    if (rc != HBT_OK) { release_all(); return rc; }
    e = hipMemcpyAsync(rec.data(), d_rec.p, nr * sizeof(hbt_subhalo), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { release_all(); return fail(HBT_ERR_DEVICE, hipGetErrorString(e)); }
  }

  // ---- final lists, compacted into out_list ---------------------------------
  std::vector<int> rpos(nsub, -1);
  for (int64_t k = 0; k < nr; k++) rpos[reunbind[k]] = (int)k;
  int64_t fin = 0;
  for (int64_t s = 0; s < nsub; s++) {
    MergeCopy &c = ranges[s];
    if (rpos[s] >= 0) {
      const hbt_subhalo &o = rec[rpos[s]];
      c = MergeCopy{d_oi.as<int64_t>(), o.out_offset, o.nsource, fin};
    } else {
      c = MergeCopy{pre.as<int64_t>(), pre_off[s], pre_cnt[s], fin};
    }
    fin += c.count;
  }
  if (out_total) *out_total = fin;
  if (fin > out_capacity) {
    release_all();
    return fail(HBT_ERR_CAPACITY, "out_list needs " + std::to_string(fin) + " entries");
  }
  {
    hipError_t e = hipMemcpyAsync(d_ranges.p, ranges.data(), nsub * sizeof(MergeCopy), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      merge_copy_kernel<<<(int)std::min<int64_t>(nsub, (int64_t)ctx->cus * 8), BLOCK, 0, st>>>(
          d_ranges.as<MergeCopy>(), nsub, out_list);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    release_all();
    if (e != hipSuccess) return fail(HBT_ERR_DEVICE, std::string("merge_subhalos: ") + hipGetErrorString(e));
  }
  for (int64_t s = 0; s < nsub; s++) {
    hbt_subhalo &r = subs_host[s];
    if (rpos[s] >= 0) {
      const hbt_subhalo &o = rec[rpos[s]];
      const int64_t parent = r.parent, nbound_prev = r.nbound_prev;
      r = o;
      r.parent = parent;
      r.nbound_prev = nbound_prev;
    }
    r.offset = ranges[s].dst_off;
    r.count = ranges[s].count;
    r.list_len = r.count;
    r.nsource = r.count;
    r.out_offset = r.offset;
  }
  std::copy(work.begin(), work.end(), subs_io);
  return HBT_OK;
}

// hbt_mask.hpp -- hbt_mask_subhalos: exclusive particle ownership inside each
// host halo (SubhaloMasker_t / SubhaloSnapshot_t::MaskSubhalos,
// src/subhalo_tracking.cpp:793-841).  Included by hbt_unbind.hip.
//
// The host orders each group's subhalos in the reference's visit order
// (post-order from the central, the other heads nested in it after its own
// nested subhalos) and lays the entries of the subhalos that claim (Nbound >
// 1) out group by group in that order, so an entry's position in its group is
// its claim order.  On the device:
//   mask_gather   key = particle ID (sign bit flipped: unsigned order =
//                 signed order), value = claim order
//   segmented sort by (ID, claim order) per group
//   mask_owner    the first entry of every run of equal IDs keeps its particle
//   segmented scan of the keep flags per group, mask_write: each subhalo's
//   kept entries in list order at the front of its range
// Subhalos that do not claim (orphans, field) keep their lists (mask_copy).
#pragma once

namespace {

struct MaskSub {
  int64_t offset;  // list range in the particle columns
  int64_t count;
  int grp;         // group segment, -1: list kept as it is
  int gstart;      // first claim position in the group segment
};

struct MaskArgs {
  const MaskSub *sub;
  int nsub;
  const int64_t *gbase;  // first slot of each group segment
  const int *gn;         // entries per group segment
  int ngrp;
  int64_t nslots;
  const int *seg_of_slot;
  const int64_t *id;
  u64 *k0;
  u32 *v0;
  unsigned char *keep;
  u64 *scan;
  int64_t *out_list;
  int64_t *new_count;
};

__global__ void __launch_bounds__(BLOCK) mask_copy_kernel(MaskArgs A) {
  for (int s = blockIdx.x; s < A.nsub; s += gridDim.x) {
    const MaskSub m = A.sub[s];
    if (m.grp >= 0) continue;
    for (int64_t l = threadIdx.x; l < m.count; l += BLOCK) A.out_list[m.offset + l] = m.offset + l;
    if (threadIdx.x == 0) A.new_count[s] = m.count;
  }
}

__global__ void __launch_bounds__(BLOCK) mask_gather_kernel(MaskArgs A) {
  for (int s = blockIdx.x; s < A.nsub; s += gridDim.x) {
    const MaskSub m = A.sub[s];
    if (m.grp < 0) continue;
    const int64_t b = A.gbase[m.grp] + m.gstart;
    for (int64_t l = threadIdx.x; l < m.count; l += BLOCK) {
      A.k0[b + l] = (u64)A.id[m.offset + l] ^ (1ull << 63);
      A.v0[b + l] = (u32)(m.gstart + l);
    }
  }
}

// sorted position i of a group: owner iff its ID differs from position i-1
__global__ void __launch_bounds__(BLOCK) mask_owner_kernel(MaskArgs A) {
  for (int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x; s < A.nslots; s += (int64_t)gridDim.x * BLOCK) {
    const int g = A.seg_of_slot[s];
    if (g < 0) continue;
    const int64_t b = A.gbase[g];
    const int i = (int)(s - b);
    if (i >= A.gn[g]) continue;
    const bool owner = i == 0 || A.k0[s - 1] != A.k0[s];
    A.keep[b + A.v0[s]] = owner ? 1 : 0;
  }
}

struct KeepLoad {
  const int64_t *gbase; const unsigned char *keep;
  __device__ u64 operator()(int g, int l) const { return (u64)keep[gbase[g] + l]; }
};
struct KeepStore {
  const int64_t *gbase; u64 *scan;
  __device__ void operator()(int g, int l, u64 x) const { scan[gbase[g] + l] = x; }
};
struct KeepAdd {
  const int64_t *gbase; u64 *scan;
  __device__ void operator()(int g, int l, u64 x) const { scan[gbase[g] + l] += x; }
};

__global__ void __launch_bounds__(BLOCK) mask_write_kernel(MaskArgs A) {
  for (int s = blockIdx.x; s < A.nsub; s += gridDim.x) {
    const MaskSub m = A.sub[s];
    if (m.grp < 0) continue;
    const int64_t b = A.gbase[m.grp] + m.gstart;
    const u64 first = A.scan[b];
    u64 kept = 0;
    for (int64_t l = threadIdx.x; l < m.count; l += BLOCK) {
      if (!A.keep[b + l]) continue;
      A.out_list[m.offset + (int64_t)(A.scan[b + l] - first)] = m.offset + l;
      kept++;
    }
    u64 tot;
    (void)block_exclusive_scan(kept, &tot);
    if (threadIdx.x == 0) A.new_count[s] = (int64_t)tot;
  }
}

}  // namespace

extern "C" int hbt_mask_subhalos(hbt_ctx *ctx, const hbt_particles *particles, const hbt_subhalo *subs_host,
                                 const int64_t *host_group, const int64_t *visit_rank, int64_t nsub, int64_t *out_list,
                                 int64_t *new_count) {
  if (!ctx || !particles || nsub < 0 || (nsub > 0 && (!subs_host || !host_group || !out_list || !new_count)))
    return fail(HBT_ERR_INVALID_ARG, "bad argument");
  if (nsub == 0) return HBT_OK;
  if (nsub >= (int64_t)INT32_MAX) return fail(HBT_ERR_UNSUPPORTED, "too many subhalos");
  if (!particles->id) return fail(HBT_ERR_INVALID_ARG, "particle IDs are required");
  HIPCHK(hipSetDevice(ctx->device));
  // nesting and heads per group
  std::vector<std::vector<int>> nested(nsub);
  std::vector<int> head_of_group;  // central per group segment
  std::unordered_map<int64_t, int> gidx;
  std::vector<std::vector<int>> heads;
  for (int64_t s = 0; s < nsub; s++) {
    const hbt_subhalo &r = subs_host[s];
    if (r.count < 0 || r.offset < 0 || r.offset + r.count > particles->n)
      return fail(HBT_ERR_INVALID_ARG, "subhalo list outside the particle columns");
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
  // visit order (post-order from the central; the other heads nested in the
  // central after its own nested subhalos), claim positions per group
  std::vector<MaskSub> ms(nsub);
  for (int64_t s = 0; s < nsub; s++) ms[s] = MaskSub{subs_host[s].offset, subs_host[s].count, -1, 0};
  std::vector<int64_t> gbase;
  std::vector<int> gn, capal, first_tile, ntile;
  std::vector<Tile> tiles;
  std::vector<char> seen(nsub, 0);
  int64_t off = 0;
  int maxn = 0;
  for (size_t g = 0; g < heads.size(); g++) {
    const int central = heads[g][0];
    int64_t pos = 0;
    // iterative post-order: (node, next child index)
    std::vector<std::pair<int, size_t>> stack{{central, 0}};
    while (!stack.empty()) {
      auto &top = stack.back();
      const int s = top.first;
      const std::vector<int> &kids = nested[s];
      const size_t nk = kids.size() + (s == central ? heads[g].size() - 1 : 0);
      if (top.second < nk) {
        const size_t i = top.second++;
        const int c = i < kids.size() ? kids[i] : heads[g][1 + (i - kids.size())];
        if (seen[c]) return fail(HBT_ERR_INVALID_ARG, "nesting has a cycle");
        seen[c] = 1;
        stack.push_back({c, 0});
        continue;
      }
      stack.pop_back();
      if (subs_host[s].nbound_prev > 1 && subs_host[s].count > 0) {
        ms[s].grp = (int)gbase.size();
        ms[s].gstart = (int)pos;
        pos += subs_host[s].count;
      }
    }
    if (pos >= (int64_t)INT32_MAX / 2) return fail(HBT_ERR_UNSUPPORTED, "host group too large");
    if (pos == 0) continue;
    const int k = (int)gbase.size();
    gbase.push_back(off);
    gn.push_back((int)pos);
    const int ca = (int)((pos + WAVE - 1) / WAVE * WAVE);
    capal.push_back(ca);
    first_tile.push_back((int)tiles.size());
    const int nt = (int)((pos + TILE - 1) / TILE);
    ntile.push_back(nt);
    for (int t = 0; t < nt; t++) tiles.push_back(Tile{k, t * TILE});
    off += ca;
    maxn = std::max(maxn, (int)pos);
  }
  // groups whose central never got a segment keep -1; map group ids for subhalos
  // that claim: ms[s].grp was set to the segment about to be created
  const int ngrp = (int)gbase.size();
  hipStream_t st = ctx->stream;
  RETCHK(ensure_slots(ctx, std::max<int64_t>(off, 1), std::max(ngrp, 1), std::max<int64_t>((int64_t)tiles.size(), 1)));
  DBuf dcnt;
  RETCHK(dcnt.ensure(nsub * 8));
  void *d[6];
  int rc = stage_blobs(ctx, {{ms.data(), ms.size() * sizeof(MaskSub)},
                             {gbase.data(), gbase.size() * 8},
                             {gn.data(), gn.size() * 4},
                             {capal.data(), capal.size() * 4},
                             {first_tile.data(), first_tile.size() * 4},
                             {tiles.data(), tiles.size() * sizeof(Tile)}},
                       d);
  if (rc != HBT_OK) { dcnt.release(); return rc; }
  MaskArgs A{};
  A.sub = (const MaskSub *)d[0];
  A.nsub = (int)nsub;
  A.gbase = (const int64_t *)d[1];
  A.gn = (const int *)d[2];
  A.ngrp = ngrp;
  A.nslots = off;
  A.seg_of_slot = ctx->seg_of_slot.as<int>();
  A.id = particles->id;
  A.k0 = ctx->k0.as<u64>();
  A.v0 = ctx->v0.as<u32>();
  A.keep = ctx->isb.as<unsigned char>();
  A.scan = ctx->scan64.as<u64>();
  A.out_list = out_list;
  A.new_count = dcnt.as<int64_t>();
  const int gsub = (int)std::min<int64_t>(nsub, (int64_t)ctx->cus * 8);
  mask_copy_kernel<<<gsub, BLOCK, 0, st>>>(A);
  if (ngrp > 0) {
    const Tile *dt = (const Tile *)d[5];
    const int nt = (int)tiles.size();
    const int gt = (int)std::max<int64_t>(1, std::min<int64_t>(nt, (int64_t)ctx->cus * 8));
    const int gm = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)nt * (TILE / BLOCK), (int64_t)ctx->cus * 16));
    fill_seg_of_slot_kernel<<<std::min(ngrp, ctx->cus * 8), BLOCK, 0, st>>>(ngrp, A.gbase, (const int *)d[3],
                                                                            ctx->seg_of_slot.as<int>());
    mask_gather_kernel<<<gsub, BLOCK, 0, st>>>(A);
    tile_sort_kernel<<<gt, BLOCK, 0, st>>>(dt, nt, A.gbase, A.gn, A.k0, A.v0, ctx->k1.as<u64>(), ctx->v1.as<u32>());
    int level = 0;
    for (int w = TILE; w < maxn; w *= 2, level++)
      merge_pass_kernel<<<gm, BLOCK, 0, st>>>(dt, nt, A.gbase, A.gn, level, A.k0, A.v0, ctx->k1.as<u64>(),
                                              ctx->v1.as<u32>());
    mask_owner_kernel<<<grid_for(off, ctx->cus), BLOCK, 0, st>>>(A);
    u64 *ts = ctx->tsum.as<u64>();
    scan_tiles_kernel<u64><<<gt, BLOCK, 0, st>>>(dt, nt, A.gn, ts, KeepLoad{A.gbase, A.keep}, KeepStore{A.gbase, A.scan});
    scan_add_kernel<u64><<<gt, BLOCK, 0, st>>>(dt, nt, A.gn, ts, KeepAdd{A.gbase, A.scan});
    mask_write_kernel<<<gsub, BLOCK, 0, st>>>(A);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(new_count, dcnt.p, nsub * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dcnt.release();
  if (e != hipSuccess) return fail(HBT_ERR_DEVICE, std::string("mask_subhalos: ") + hipGetErrorString(e));
  return HBT_OK;
}

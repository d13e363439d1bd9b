// hbt_nest.hpp -- hbt_nest_subhalos: the nest bookkeeping before masking
// (SubhaloSnapshot_t::NestSubhalos -> LevelUpDetachedSubhalos,
// Subhalo_t::LevelUpDetachedMembers, src/subhalo_tracking.cpp:669-688,
// 735-786).  Included by hbt_unbind.hip.
//
// The reference first marks the heads (Rank 0), then gives every group's
// central and every field subhalo Rank 0, then walks each head's nest
// recursively: a nested subhalo leaves its host's nest (and gets Rank 0) when
// it lies in another host halo than that host or has Rank 0.  The decision for
// one subhalo depends only on itself, its host and whether any host up its
// chain was a head on entry (the walks start at every such head), so one
// thread per subhalo decides it: it climbs its chain of hosts (as given on
// entry) to the top, then applies the rule.
#pragma once

namespace {

constexpr int NEST_DEPTH_MAX = 4096;  // longer chains are rejected (cyclic or corrupt)

struct NestArgs {
  int64_t nsub;
  const int64_t *group;      // HostHaloId
  const int32_t *central;    // 1: the group's central (SubGroups[g][0])
  const int64_t *parent0;    // nest relation on entry
  const int32_t *rank0;      // Rank on entry
  int64_t *parent;           // out
  int32_t *rank;             // out
  int *bad;                  // subhalos with a bad host index or a chain deeper than NEST_DEPTH_MAX
};

__global__ void __launch_bounds__(BLOCK) nest_kernel(NestArgs A) {
  int nbad = 0;
  for (int64_t x = (int64_t)blockIdx.x * BLOCK + threadIdx.x; x < A.nsub; x += (int64_t)gridDim.x * BLOCK) {
    const int64_t p = A.parent0[x];
    // promotion: centrals and field subhalos get Rank 0
    int32_t r = (A.central[x] || A.group[x] < 0) ? 0 : A.rank0[x];
    int64_t np = p;
    if (p < -1 || p >= A.nsub || p == x) {
      nbad++;
    } else if (p >= 0) {
      // the recursive walks start at every head on entry (Rank 0), nested or
      // not: x is reached iff one of its hosts up the chain was a head
      int64_t h = p;
      int steps = 0;
      bool reached = false, ok = true;
      while (true) {
        if (A.rank0[h] == 0) reached = true;
        const int64_t u = A.parent0[h];
        if (u < 0) break;
        if (u >= A.nsub || ++steps >= NEST_DEPTH_MAX) { ok = false; break; }
        h = u;
      }
      if (!ok) {
        nbad++;
      } else if (reached && (A.group[x] != A.group[p] || r == 0)) {
        np = -1;  // detached: leaves its host's nest and becomes a head
        r = 0;
      }
    }
    A.parent[x] = np;
    A.rank[x] = r;
  }
  if (nbad) atomicAdd(A.bad, nbad);
}

}  // namespace

extern "C" int hbt_nest_subhalos(hbt_ctx *ctx, int64_t nsub, const int64_t *host_group,
                                 const int32_t *is_central, int64_t *parent, int32_t *rank) {
  if (!ctx || nsub < 0 || (nsub > 0 && (!host_group || !is_central || !parent || !rank)))
    return fail(HBT_ERR_INVALID_ARG, "bad argument");
  if (nsub == 0) return HBT_OK;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  DBuf copy;
  RETCHK(copy.ensure((size_t)nsub * 12 + 16));
  RETCHK(ctx->active.ensure(16));
  int64_t *p0 = copy.as<int64_t>();
  int32_t *r0 = (int32_t *)(p0 + nsub);
  NestArgs A{nsub, host_group, is_central, p0, r0, parent, rank, ctx->active.as<int>()};
  int zero = 0, bad = 0;
  hipError_t e = hipMemcpyAsync(p0, parent, nsub * 8, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(r0, rank, nsub * 4, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(A.bad, &zero, sizeof(int), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    nest_kernel<<<grid_for(nsub, ctx->cus), BLOCK, 0, st>>>(A);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, A.bad, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess && bad) {  // a rejected call leaves the arrays as they were
    e = hipMemcpyAsync(parent, p0, nsub * 8, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(rank, r0, nsub * 4, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  copy.release();
  if (e != hipSuccess) return fail(HBT_ERR_DEVICE, std::string("nest_subhalos: ") + hipGetErrorString(e));
  if (bad)
    return fail(HBT_ERR_INVALID_ARG, std::to_string(bad) +
                                         " subhalos have a host outside the subhalos or a nest chain deeper than 4096");
  return HBT_OK;
}

// hbt_tree_api.hpp -- hbt_tree_potential: GravityTree_t::Build over a source
// set in the given order followed by EvaluatePotential at arbitrary targets
// (src/oct_tree.tpp:17-144, src/gravity_tree.cpp:79-164), on the same tree
// kernels as the unbinding passes.  Included by hbt_unbind.hip.
#pragma once

namespace {

__global__ void tree_only_setup_kernel(Work W, int n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Seg &S = W.seg[0];
    W.ncount[NC_SRC * W.nseg] = n;
    S.flags = F_ACTIVE;
    S.n_src = n; S.src_lo = 0; S.n_tgt = 0; S.nsort = n; S.nlast = n; S.nbound = n;
    for (int j = 0; j < 3; j++) { S.bmin[j] = 0xffffffffu; S.bmax[j] = 0u; }
  }
}

__global__ void tree_sources_kernel(Work W, int n, const float *x, const float *y, const float *z,
                                    const float *m) {
  for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    W.pos[i] = make_float4(x[i], y[i], z[i], m[i]);
    W.el[i] = i;
  }
}

__global__ void walk_points_kernel(Work W, DevParams P, int64_t nt, const float *tx, const float *ty,
                                   const float *tz, const float *tm, double *out) {
  // the targets may lie anywhere: a periodic box always takes the wrapped walk
  Seg S = W.seg[0];
  S.straddle = P.periodic;
  unsigned ni = 0, no = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nt; i += (int64_t)gridDim.x * BLOCK)
    out[i] = walk_dispatch<false>(W, P, S, tx[i], ty[i], tz[i], tm ? tm[i] : 0.f, ni, no);
}

}  // namespace

extern "C" int hbt_tree_potential(hbt_ctx *ctx, int64_t n_src, const float *sx, const float *sy,
                                  const float *sz, const float *smass, int64_t n_tgt, const float *tx,
                                  const float *ty, const float *tz, const float *target_mass,
                                  double *out_phi) {
  if (!ctx || n_src < 1 || n_tgt < 0 || !sx || !sy || !sz || !smass || (n_tgt > 0 && (!tx || !ty || !tz || !out_phi)))
    return fail(HBT_ERR_INVALID_ARG, "bad argument");
  if (n_src >= INT32_MAX / 2) return fail(HBT_ERR_UNSUPPORTED, "too many sources");
  HIPCHK(hipSetDevice(ctx->device));
  hbt_subhalo h{};
  h.count = n_src;
  h.parent = -1;
  Nest N;
  RETCHK(plan_nests(&h, 1, N));
  Level L;
  build_level(N, std::vector<int>{0}, ctx->tune, nullptr, 1, L);
  RETCHK(ensure_slots(ctx, L.nslots, 1, (int64_t)L.tiles.size()));
  RETCHK(stage_level(ctx, L));
  Runner R;
  R.c = ctx;
  R.L = &L;
  R.P = make_dev_params(ctx->params);
  R.W = make_work(ctx, L, nullptr, nullptr, nullptr, nullptr);
  R.max_n = L.maxcap;
  R.list_all();
  R.gseg = 1;
  R.gblk = 1;
  hipStream_t st = ctx->stream;
  seg_expand_kernel<<<1, BLOCK, 0, st>>>(R.W.seg, ctx->segin.as<SegIn>(), 1);
  seg_of_slot_kernel<<<grid_for(L.nslots, ctx->cus), BLOCK, 0, st>>>(1, R.W.ibase, L.nslots,
                                                                      ctx->seg_of_slot.as<int>());
  tree_only_setup_kernel<<<1, 1, 0, st>>>(R.W, (int)n_src);
  tree_sources_kernel<<<grid_for(n_src, ctx->cus), BLOCK, 0, st>>>(R.W, (int)n_src, sx, sy, sz, smass);
  RETCHK(R.build_tree(false));
  if (n_tgt > 0)
    walk_points_kernel<<<grid_for(n_tgt, ctx->cus), BLOCK, 0, st>>>(R.W, R.P, n_tgt, tx, ty, tz,
                                                                    target_mass, out_phi);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return HBT_OK;
}

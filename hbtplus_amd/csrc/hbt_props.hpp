// hbt_props.hpp -- hbt_subhalo_properties: Subhalo_t::CalculateProfileProperties
// and Subhalo_t::CalculateShape (src/subhalo.cpp:242-398) for a batch of
// subhalos after unbinding, on the bound lists resident on the device.
// Included by hbt_unbind.hip.
//
// One segment per subhalo with Nbound > 1 (slots = its bound particles):
//   props_gather   distance to the most-bound particle (PeriodicDistance,
//                  src/config_parser.h:143-156) and mass, key (distance, index)
//   segmented sort by distance (the unbinding's tile sort and merge levels)
//   segmented FP64 scan of the masses in distance order (m_cum, :296-297)
//   props_elem     per tile: first maximum of m(<r)/max(r, eps) (:299-306),
//                  the outermost radius enclosing 200 times the critical
//                  density (Snapshot_t::SphericalOverdensitySize,
//                  src/snapshot.cpp:259-280), RHalf and R2Sigma
//   props_shape    per tile: the 12 FP64 inertia sums over bound particles
//                  1..Nbound-1 in list order (:355-384)
//   props_combine  per subhalo: tiles in order, outputs
#pragma once

namespace {

struct PropSeg {
  int64_t ibase;    // first slot
  int64_t out_off;  // the caller's out_offset
  int sub;          // record index
  int nb;           // Nbound (> 1)
};

struct PropsArgs {
  const PropSeg *seg;
  int nseg;
  int64_t nslots;
  const int *seg_of_slot;
  const int64_t *ibase;
  const int *n;  // Nbound per segment
  const hbt_subhalo *subs;
  const int64_t *out_index;
  hbt_properties *props;
  const float *gx, *gy, *gz, *gm;
  u64 *k0, *k1;
  u32 *v0, *v1;
  float *r, *m;  // per slot, list order
  double *cum;   // per slot, distance order: exclusive prefix of the masses
};

constexpr int PROF_PART = 5;   // per tile: v, its position, its radius, SO position, SO mass
constexpr int SHAPE_SUMS = 12;

__device__ __forceinline__ float periodic_distance_f(const DevParams &P, const float c[3], float x,
                                                     float y, float z) {
#pragma clang fp contract(off)
  float d0 = c[0] - x, d1 = c[1] - y, d2 = c[2] - z;
  if (P.periodic) { d0 = nearest_f(d0, P); d1 = nearest_f(d1, P); d2 = nearest_f(d2, P); }
  return sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
}

// 200 times the critical density in comoving units, with the reference's
// float/double promotions (src/snapshot.cpp:268, VirialFactor = 200 at c200).
__device__ __forceinline__ float rho_c200(const DevParams &P) {
#pragma clang fp contract(off)
  const float f = 200.f * P.hz * P.hz;
  return (float)((double)f / 2.0 / (double)P.G * (double)P.a * (double)P.a * (double)P.a);
}

// Nbound <= 1: zero profile and shape (src/subhalo.cpp:262-279, 339-351).
__global__ void props_small_kernel(const hbt_subhalo *subs, int64_t nsub, hbt_properties *props) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsub; i += (int64_t)gridDim.x * blockDim.x) {
    if (subs[i].nbound > 1) continue;
    hbt_properties &o = props[i];
    o.rmax_comoving = 0.f; o.vmax_physical = 0.f; o.r2sigma_comoving = 0.f; o.rhalf_comoving = 0.f;
    o.bound_r200crit_comoving = 0.f; o.bound_m200crit = 0.f;
    for (int j = 0; j < 6; j++) { o.inertial_tensor[j] = 0.f; o.inertial_tensor_weighted[j] = 0.f; }
  }
}

__global__ void __launch_bounds__(BLOCK) props_gather_kernel(PropsArgs A, DevParams P) {
  for (int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x; s < A.nslots; s += (int64_t)gridDim.x * BLOCK) {
    const int k = A.seg_of_slot[s];
    if (k < 0) continue;
    const PropSeg &S = A.seg[k];
    const int l = (int)(s - S.ibase);
    if (l >= S.nb) continue;
    const hbt_subhalo &h = A.subs[S.sub];
    const int64_t g = A.out_index[S.out_off + l];
    const float rr = periodic_distance_f(P, h.mostbound_pos, A.gx[g], A.gy[g], A.gz[g]);
    A.r[s] = rr;
    A.m[s] = A.gm[g];
    A.k0[s] = ((u64)f2ord(rr) << 32) | (u64)(unsigned)l;
    A.v0[s] = (u32)l;
  }
}

// scan functors: masses in distance order
struct CumLoad {
  const int64_t *ibase; const u32 *v0; const float *m;
  __device__ double operator()(int k, int l) const {
    const int64_t b = ibase[k];
    return (double)m[b + v0[b + l]];
  }
};
struct CumStore {
  const int64_t *ibase; double *cum;
  __device__ void operator()(int k, int l, double x) const { cum[ibase[k] + l] = x; }
};
struct CumAdd {
  const int64_t *ibase; double *cum;
  __device__ void operator()(int k, int l, double x) const { cum[ibase[k] + l] += x; }
};

// Per tile of a segment in distance order.  v = m(<r)/max(r, eps) with
// m(<r) the float of the FP64 running sum; the first maximum wins (max_element).
__global__ void __launch_bounds__(BLOCK)
props_elem_kernel(PropsArgs A, DevParams P, const Tile *tiles, int ntiles, double *part) {
#pragma clang fp contract(off)
  __shared__ float sv[BLOCK], sr[BLOCK], sm[BLOCK];
  __shared__ int sl[BLOCK], so[BLOCK];
  const float rho = rho_c200(P);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const PropSeg &S = A.seg[tl.seg];
    const int cnt = min(TILE, S.nb - tl.start);
    if (cnt <= 0) continue;  // uniform
    const int64_t b = S.ibase;
    float bv = -1.f, br = 0.f, bm = 0.f;
    int bl = INT_MAX, bso = -1;
    for (int i = threadIdx.x; i < cnt; i += BLOCK) {
      const int ls = tl.start + i;
      const int li = (int)A.v0[b + ls];
      float rr = A.r[b + li];
      const float mc = (float)(A.cum[b + ls] + (double)A.m[b + li]);
      if (rr < P.eps) rr = P.eps;  // SofteningHalo
      const float v = mc / rr;
      if (v > bv || (v == bv && ls < bl)) { bv = v; bl = ls; br = rr; }
      if (mc > rho * rr * rr * rr && ls > bso) { bso = ls; bm = mc; }
      if (ls == S.nb / 2) A.props[S.sub].rhalf_comoving = rr;
      if (ls == (int)((int64_t)S.nb * 0.955)) A.props[S.sub].r2sigma_comoving = rr;
    }
    sv[threadIdx.x] = bv; sl[threadIdx.x] = bl; sr[threadIdx.x] = br;
    so[threadIdx.x] = bso; sm[threadIdx.x] = bm;
    __syncthreads();
    if (threadIdx.x == 0) {
      float v = -1.f, rr = 0.f, mso = 0.f;
      int l = INT_MAX, lso = -1;
      for (int q = 0; q < BLOCK; q++) {
        if (sv[q] > v || (sv[q] == v && sl[q] < l)) { v = sv[q]; l = sl[q]; rr = sr[q]; }
        if (so[q] > lso) { lso = so[q]; mso = sm[q]; }
      }
      double *p = part + (int64_t)t * PROF_PART;
      p[0] = v; p[1] = l; p[2] = rr; p[3] = lso; p[4] = mso;
    }
    __syncthreads();
  }
}

// Inertia sums of a tile in list order, particles 1..Nbound-1.
__global__ void __launch_bounds__(BLOCK)
props_shape_kernel(PropsArgs A, DevParams P, const Tile *tiles, int ntiles, double *part) {
#pragma clang fp contract(off)
  __shared__ double sh[SHAPE_SUMS][BLOCK / WAVE];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const PropSeg &S = A.seg[tl.seg];
    const int cnt = min(TILE, S.nb - tl.start);
    if (cnt <= 0) continue;  // uniform
    const float *c = A.subs[S.sub].mostbound_pos;
    double v[SHAPE_SUMS];
    for (int q = 0; q < SHAPE_SUMS; q++) v[q] = 0.;
    for (int i = threadIdx.x; i < cnt; i += BLOCK) {
      const int l = tl.start + i;
      if (l == 0) continue;
      const int64_t g = A.out_index[S.out_off + l];
      const float m = A.gm[g];
      float dx = A.gx[g] - c[0], dy = A.gy[g] - c[1], dz = A.gz[g] - c[2];
      if (P.periodic) { dx = nearest_f(dx, P); dy = nearest_f(dy, P); dz = nearest_f(dz, P); }
      const float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
      v[0] += dx2 * m; v[1] += dx * dy * m; v[2] += dx * dz * m;
      v[3] += dy2 * m; v[4] += dy * dz * m; v[5] += dz2 * m;
      float dr2 = dx2 + dy2 + dz2;
      dr2 /= m;
      v[6] += dx2 / dr2; v[7] += dx * dy / dr2; v[8] += dx * dz / dr2;
      v[9] += dy2 / dr2; v[10] += dy * dz / dr2; v[11] += dz2 / dr2;
    }
    for (int q = 0; q < SHAPE_SUMS; q++) {
      const double r = wave_reduce_sum(v[q]);
      if ((threadIdx.x & (WAVE - 1)) == 0) sh[q][threadIdx.x / WAVE] = r;
    }
    __syncthreads();
    if (threadIdx.x < SHAPE_SUMS) {
      double r = 0;
      for (int w = 0; w < BLOCK / WAVE; w++) r += sh[threadIdx.x][w];
      part[(int64_t)t * SHAPE_SUMS + threadIdx.x] = r;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BLOCK) props_combine_kernel(PropsArgs A, DevParams P, const int *first_tile,
                                     const double *prof_part, const double *shape_part) {
#pragma clang fp contract(off)
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A.nseg; k += gridDim.x * blockDim.x) {
    const PropSeg &S = A.seg[k];
    const hbt_subhalo &h = A.subs[S.sub];
    hbt_properties &o = A.props[S.sub];
    const int t0 = first_tile[k], nt = (S.nb + TILE - 1) / TILE;
    double v = -1., rr = 0., mso = 0.;
    int lso = -1;
    double sums[SHAPE_SUMS];
    for (int q = 0; q < SHAPE_SUMS; q++) sums[q] = 0.;
    // tiles in order (the first maximum stays; the sums keep their order);
    // PREF tiles are loaded ahead of their use, so a subhalo of thousands of
    // tiles does not wait for each load in turn
    constexpr int PREF = 4;
    int t = t0;
    for (; t + PREF <= t0 + nt; t += PREF) {
      double pv[PREF], pr[PREF], pl[PREF], pm[PREF], sp[PREF][SHAPE_SUMS];
#pragma unroll
      for (int u = 0; u < PREF; u++) {
        const double *p = prof_part + (int64_t)(t + u) * PROF_PART;
        pv[u] = p[0]; pr[u] = p[2]; pl[u] = p[3]; pm[u] = p[4];
#pragma unroll
        for (int q = 0; q < SHAPE_SUMS; q++) sp[u][q] = shape_part[(int64_t)(t + u) * SHAPE_SUMS + q];
      }
#pragma unroll
      for (int u = 0; u < PREF; u++) {
        if (pv[u] > v) { v = pv[u]; rr = pr[u]; }
        if ((int)pl[u] > lso) { lso = (int)pl[u]; mso = pm[u]; }
#pragma unroll
        for (int q = 0; q < SHAPE_SUMS; q++) sums[q] += sp[u][q];
      }
    }
    for (; t < t0 + nt; t++) {
      const double *p = prof_part + (int64_t)t * PROF_PART;
      if (p[0] > v) { v = p[0]; rr = p[2]; }
      if ((int)p[3] > lso) { lso = (int)p[3]; mso = p[4]; }
      for (int q = 0; q < SHAPE_SUMS; q++) sums[q] += shape_part[(int64_t)t * SHAPE_SUMS + q];
    }
    const float vel_unit = P.G / P.a;
    o.rmax_comoving = (float)rr;
    o.vmax_physical = sqrtf((float)v * vel_unit);
    if (lso >= 0) {
      const float rho = rho_c200(P);
      o.bound_m200crit = (float)mso;
      o.bound_r200crit_comoving = (float)pow((double)((float)mso / rho), 1.0 / 3);
    }
    if (o.vmax_physical >= o.last_max_vmax_physical) {
      o.snapshot_index_of_last_max_vmax = P.snap;
      o.last_max_vmax_physical = o.vmax_physical;
    }
    for (int j = 0; j < 6; j++) {  // {xx, xy, xz, yy, yz, zz}, then divided by Mbound
      float a = (float)sums[j];
      a /= h.mbound;
      o.inertial_tensor[j] = a;
      float w = (float)sums[6 + j];
      w /= h.mbound;
      o.inertial_tensor_weighted[j] = w;
    }
  }
}

}  // namespace

extern "C" int hbt_subhalo_properties(hbt_ctx *ctx, const hbt_particles *particles,
                                      const hbt_subhalo *subs, int64_t nsub, const int64_t *out_index,
                                      hbt_properties *props) {
  if (!ctx || !particles || nsub < 0 || (nsub > 0 && (!subs || !out_index || !props)))
    return fail(HBT_ERR_INVALID_ARG, "bad argument");
  if (nsub == 0) return HBT_OK;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  std::vector<hbt_subhalo> rec(nsub);
  HIPCHK(hipMemcpyAsync(rec.data(), subs, nsub * sizeof(hbt_subhalo), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // segments: subhalos with Nbound > 1
  std::vector<PropSeg> seg;
  std::vector<int64_t> ibase;
  std::vector<int> capal, first_tile, ntile, nb;
  std::vector<Tile> tiles;
  int64_t off = 0;
  int maxn = 0;
  for (int64_t i = 0; i < nsub; i++) {
    const int64_t n = rec[i].nbound;
    if (n <= 1) continue;
    if (n >= (int64_t)INT32_MAX / 2 || rec[i].out_offset < 0)
      return fail(HBT_ERR_INVALID_ARG, "bad nbound or out_offset");
    const int k = (int)seg.size();
    seg.push_back(PropSeg{off, rec[i].out_offset, (int)i, (int)n});
    ibase.push_back(off);
    const int ca = (int)((n + WAVE - 1) / WAVE * WAVE);
    capal.push_back(ca);
    first_tile.push_back((int)tiles.size());
    const int nt = (int)((n + TILE - 1) / TILE);
    ntile.push_back(nt);
    for (int t = 0; t < nt; t++) tiles.push_back(Tile{k, t * TILE});
    nb.push_back((int)n);
    off += ca;
    maxn = std::max(maxn, (int)n);
  }
  const DevParams P = make_dev_params(ctx->params);
  props_small_kernel<<<grid_for(nsub, ctx->cus), BLOCK, 0, st>>>(subs, nsub, props);
  if (seg.empty()) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return HBT_OK;
  }
  const int nseg = (int)seg.size(), ntiles = (int)tiles.size();
  RETCHK(ensure_slots(ctx, off, nseg, ntiles));
  void *d[7];
  RETCHK(stage_blobs(ctx, {{seg.data(), seg.size() * sizeof(PropSeg)},
                           {ibase.data(), ibase.size() * 8},
                           {capal.data(), capal.size() * 4},
                           {first_tile.data(), first_tile.size() * 4},
                           {ntile.data(), ntile.size() * 4},
                           {nb.data(), nb.size() * 4},
                           {tiles.data(), tiles.size() * sizeof(Tile)}},
                     d));
  PropsArgs A{};
  A.seg = (const PropSeg *)d[0];
  A.nseg = nseg;
  A.nslots = off;
  A.seg_of_slot = ctx->seg_of_slot.as<int>();
  A.ibase = (const int64_t *)d[1];
  A.n = (const int *)d[5];
  A.subs = subs;
  A.out_index = out_index;
  A.props = props;
  A.gx = particles->x; A.gy = particles->y; A.gz = particles->z; A.gm = particles->mass;
  A.k0 = ctx->k0.as<u64>(); A.k1 = ctx->k1.as<u64>();
  A.v0 = ctx->v0.as<u32>(); A.v1 = ctx->v1.as<u32>();
  A.r = ctx->E.as<float>();
  A.m = ctx->Er.as<float>();
  A.cum = (double *)ctx->scan64.as<u64>();
  const Tile *dt = (const Tile *)d[6];
  const int gs = grid_for(off, ctx->cus);
  const int gt = (int)std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)ctx->cus * 8));
  const int gm = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ntiles * (TILE / BLOCK), (int64_t)ctx->cus * 16));
  const int gk = grid_for(nseg, ctx->cus);
  double *prof_part = ctx->tpart.as<double>();
  double *shape_part = prof_part + (size_t)(ntiles + 1) * PROF_PART;
  seg_of_slot_kernel<<<gs, BLOCK, 0, st>>>(nseg, A.ibase, off, ctx->seg_of_slot.as<int>());
  props_gather_kernel<<<gs, BLOCK, 0, st>>>(A, P);
  // sort by distance
  tile_sort_kernel<<<gt, BLOCK, 0, st>>>(dt, ntiles, A.ibase, A.n, A.k0, A.v0, A.k1, A.v1);
  int level = 0;
  for (int w = TILE; w < maxn; w *= 2, level++)
    merge_pass_kernel<<<gm, BLOCK, 0, st>>>(dt, ntiles, A.ibase, A.n, level, A.k0, A.v0, A.k1, A.v1);
  // running mass in distance order
  double *ts = (double *)ctx->tsum.as<u64>();
  scan_tiles_kernel<double><<<gt, BLOCK, 0, st>>>(dt, ntiles, A.n, ts, CumLoad{A.ibase, A.v0, A.m},
                                                  CumStore{A.ibase, A.cum});
  scan_add_kernel<double><<<gt, BLOCK, 0, st>>>(dt, ntiles, A.n, ts, CumAdd{A.ibase, A.cum});
  props_elem_kernel<<<gt, BLOCK, 0, st>>>(A, P, dt, ntiles, prof_part);
  props_shape_kernel<<<gt, BLOCK, 0, st>>>(A, P, dt, ntiles, shape_part);
  props_combine_kernel<<<gk, BLOCK, 0, st>>>(A, P, (const int *)d[3], prof_part, shape_part);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return HBT_OK;
}

// hbt_direct.hpp -- hbt_direct_potential: exact pair sums with the walk's
// softened kernel (src/gravity_tree.cpp:144-160), self term excluded, for
// groups of particles (one group per subhalo).  Sources are staged through
// LDS in tiles of BLOCK particles; each thread owns one target.
// Modes: FP32, FP32 with Kahan-compensated sums, FP64.  Included by
// hbt_unbind.hip.
#pragma once

namespace {

struct DTile { int64_t seg_lo, seg_hi, tgt0; };

template <int MODE>
__global__ void __launch_bounds__(BLOCK)
direct_kernel(const DTile *tiles, int ntiles, const float *x, const float *y, const float *z,
              const float *m, DevParams P, double *out) {
  __shared__ float4 sh[BLOCK];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const DTile T = tiles[t];
    const int64_t i = T.tgt0 + threadIdx.x;
    const bool live = i < T.seg_hi;
    const float tx = live ? x[i] : 0.f, ty = live ? y[i] : 0.f, tz = live ? z[i] : 0.f;
    const double h = 2.8 * (double)P.eps, hinv = 1.0 / h;
    const float hf = 2.8f * P.eps, h2f = hf * hf, hinvf = 1.0f / hf;
    double accd = 0.0;
    Kahan<true> acc;
    acc.init(0.f);
    for (int64_t j0 = T.seg_lo; j0 < T.seg_hi; j0 += BLOCK) {
      const int64_t j = j0 + threadIdx.x;
      sh[threadIdx.x] = j < T.seg_hi ? make_float4(x[j], y[j], z[j], m[j]) : make_float4(0, 0, 0, 0);
      __syncthreads();
      const int nj = (int)min<int64_t>(BLOCK, T.seg_hi - j0);
      if (live) {
        for (int q = 0; q < nj; q++) {
          if (j0 + q == i) continue;
          const float4 s = sh[q];
          if (MODE == HBT_DIRECT_FP64) {
#pragma clang fp contract(off)
            double dx = (double)s.x - tx, dy = (double)s.y - ty, dz = (double)s.z - tz;
            if (P.periodic) { dx = nearest_d(dx, P); dy = nearest_d(dy, P); dz = nearest_d(dz, P); }
            accd += kernel_term_d((double)s.w, sqrt(dx * dx + dy * dy + dz * dz), h, hinv);
          } else {
            float dx = s.x - tx, dy = s.y - ty, dz = s.z - tz;
            if (P.periodic) { dx = nearest_f(dx, P); dy = nearest_f(dy, P); dz = nearest_f(dz, P); }
            const float term = kernel_term_f(s.w, pair_r2(dx, dy, dz), h2f, hinvf);
            if (MODE == HBT_DIRECT_FP32_KAHAN) acc.add(term);
            else acc.s += term;
          }
        }
      }
      __syncthreads();
    }
    if (live) {
      const double pot = MODE == HBT_DIRECT_FP64 ? accd : (double)acc.s;
      out[i] = pot * (double)P.G / (double)P.a;
    }
  }
}

}  // namespace

extern "C" int hbt_direct_potential(hbt_ctx *ctx, int64_t nseg, const int64_t *seg_offset,
                                    const float *x, const float *y, const float *z,
                                    const float *mass, int mode, double *out_phi) {
  if (!ctx || nseg < 0 || !seg_offset) return fail(HBT_ERR_INVALID_ARG, "bad argument");
  if (mode != HBT_DIRECT_FP32 && mode != HBT_DIRECT_FP32_KAHAN && mode != HBT_DIRECT_FP64)
    return fail(HBT_ERR_INVALID_ARG, "bad mode");
  if (nseg == 0) return HBT_OK;
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<int64_t> off(nseg + 1);
  HIPCHK(hipMemcpy(off.data(), seg_offset, (nseg + 1) * 8, hipMemcpyDeviceToHost));
  std::vector<DTile> tiles;
  for (int64_t s = 0; s < nseg; s++) {
    if (off[s + 1] < off[s]) return fail(HBT_ERR_INVALID_ARG, "seg_offset not ascending");
    for (int64_t t = off[s]; t < off[s + 1]; t += BLOCK) tiles.push_back(DTile{off[s], off[s + 1], t});
  }
  if (tiles.empty()) return HBT_OK;
  DBuf tb;
  RETCHK(tb.ensure(tiles.size() * sizeof(DTile)));
  hipStream_t st = ctx->stream;
  hipError_t e = hipMemcpyAsync(tb.p, tiles.data(), tiles.size() * sizeof(DTile), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    const DevParams P = make_dev_params(ctx->params);
    const int nt = (int)tiles.size();
    const int g = std::min(nt, ctx->cus * 16);
    if (mode == HBT_DIRECT_FP64)
      direct_kernel<HBT_DIRECT_FP64><<<g, BLOCK, 0, st>>>(tb.as<DTile>(), nt, x, y, z, mass, P, out_phi);
    else if (mode == HBT_DIRECT_FP32_KAHAN)
      direct_kernel<HBT_DIRECT_FP32_KAHAN><<<g, BLOCK, 0, st>>>(tb.as<DTile>(), nt, x, y, z, mass, P, out_phi);
    else
      direct_kernel<HBT_DIRECT_FP32><<<g, BLOCK, 0, st>>>(tb.as<DTile>(), nt, x, y, z, mass, P, out_phi);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  tb.release();
  if (e != hipSuccess) return fail(HBT_ERR_DEVICE, std::string("direct potential: ") + hipGetErrorString(e));
  return HBT_OK;
}

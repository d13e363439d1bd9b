// hbt_direct.hpp -- hbt_direct_potential: exact pair sums with the walk's
// softened kernel (src/gravity_tree.cpp:144-160), self term excluded, for
// groups of particles (one group per subhalo).  Sources are staged through
// LDS in tiles of BLOCK particles; each thread owns one target.
// Modes: FP32, FP32 with Kahan-compensated sums, FP64.  Included by
// hbt_unbind.hip.
#pragma once

namespace {

struct DTile { int64_t seg_lo, seg_hi, tgt0; };

// One tile of targets (BLOCK consecutive particles of a group) against every
// particle of its group; the workgroup stages the sources through sh.
template <int MODE>
__device__ __forceinline__ void direct_tile(const DTile T, const float *x, const float *y, const float *z,
                                            const float *m, const DevParams &P, double *out, float4 *sh) {
  {
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

template <int MODE>
__global__ void __launch_bounds__(BLOCK)
direct_kernel(const DTile *tiles, int ntiles, const float *x, const float *y, const float *z,
              const float *m, DevParams P, double *out) {
  __shared__ float4 sh[BLOCK];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) direct_tile<MODE>(tiles[t], x, y, z, m, P, out, sh);
}

// ---- the entry point that allocates nothing ---------------------------------
// Its scratch -- the first tile of every group and the tile total, then a flag
// for offsets that are not ascending -- lives in this file-scope device array.
// Each GPU holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress gives for the device (the array's
// name in host code is not a device address).
constexpr int64_t NOALLOC_MAX_GROUPS = 65536;
static __device__ __align__(256) int64_t scratch[NOALLOC_MAX_GROUPS + 2];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Once per device, with that device current (hbt_ctx_create).
int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(HBT_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return HBT_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(HBT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return HBT_OK;
}

// First tile of every group (ceil(len / BLOCK) tiles each, an exclusive scan
// in one workgroup), the total at tstart[nseg], and *bad = 1 when a group's
// offsets descend (such a group gets no tiles).
__global__ void __launch_bounds__(BLOCK) direct_tile_starts_kernel(const int64_t *__restrict__ seg_offset,
                                                                   int64_t nseg, int64_t *tstart, int *bad) {
  __shared__ int sbad;
  if (threadIdx.x == 0) sbad = 0;
  __syncthreads();
  int64_t run = 0;
  for (int64_t s0 = 0; s0 < nseg; s0 += BLOCK) {  // uniform across the block
    const int64_t s = s0 + threadIdx.x;
    int64_t nt = 0;
    if (s < nseg) {
      const int64_t len = seg_offset[s + 1] - seg_offset[s];
      if (len < 0) atomicOr(&sbad, 1);
      nt = len > 0 ? (len + BLOCK - 1) / BLOCK : 0;
    }
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_scan((unsigned long long)nt, &tot);
    if (s < nseg) tstart[s] = run + (int64_t)ex;
    run += (int64_t)tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    tstart[nseg] = run;
    *bad = sbad;
  }
}

// Tile t belongs to the last group whose first tile is at or before it.
template <int MODE>
__global__ void __launch_bounds__(BLOCK)
direct_noalloc_kernel(const int64_t *__restrict__ seg_offset, int64_t nseg, const int64_t *__restrict__ tstart,
                      const float *x, const float *y, const float *z, const float *m, DevParams P, double *out) {
  __shared__ float4 sh[BLOCK];
  const int64_t ntiles = tstart[nseg];
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int64_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (tstart[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const DTile T{seg_offset[lo], seg_offset[lo + 1], seg_offset[lo] + (t - tstart[lo]) * BLOCK};
    direct_tile<MODE>(T, x, y, z, m, P, out, sh);
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
  // ordered after whatever the caller enqueued on the context's stream
  HIPCHK(hipMemcpyAsync(off.data(), seg_offset, (nseg + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
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

// hbt_direct_potential without allocating: no hipMalloc, hipFree or other
// allocating call; the tile table goes to the device's scratch array.  Calls
// that overlap on one GPU race for it.
extern "C" int hbt_direct_potential_noalloc(hbt_ctx *ctx, int64_t nseg, const int64_t *seg_offset,
                                            const float *x, const float *y, const float *z,
                                            const float *mass, int mode, double *out_phi) {
  if (!ctx || nseg < 0 || !seg_offset) return fail(HBT_ERR_INVALID_ARG, "bad argument");
  if (mode != HBT_DIRECT_FP32 && mode != HBT_DIRECT_FP32_KAHAN && mode != HBT_DIRECT_FP64)
    return fail(HBT_ERR_INVALID_ARG, "bad mode");
  if (nseg > NOALLOC_MAX_GROUPS)
    return fail(HBT_ERR_CAPACITY, "hbt_direct_potential_noalloc takes at most 65536 groups");
  if (nseg == 0) return HBT_OK;
  HIPCHK(hipSetDevice(ctx->device));
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) ptr = g_scratch_ptr[ctx->device];
  }
  if (!ptr) return fail(HBT_ERR_DEVICE, "no scratch address for the context's device");
  int64_t *tstart = (int64_t *)ptr;
  int *bad = (int *)(tstart + NOALLOC_MAX_GROUPS + 1);
  hipStream_t st = ctx->stream;
  const DevParams P = make_dev_params(ctx->params);
  direct_tile_starts_kernel<<<1, BLOCK, 0, st>>>(seg_offset, nseg, tstart, bad);
  const int g = ctx->cus * 16;
  if (mode == HBT_DIRECT_FP64)
    direct_noalloc_kernel<HBT_DIRECT_FP64><<<g, BLOCK, 0, st>>>(seg_offset, nseg, tstart, x, y, z, mass, P, out_phi);
  else if (mode == HBT_DIRECT_FP32_KAHAN)
    direct_noalloc_kernel<HBT_DIRECT_FP32_KAHAN><<<g, BLOCK, 0, st>>>(seg_offset, nseg, tstart, x, y, z, mass, P,
                                                                     out_phi);
  else
    direct_noalloc_kernel<HBT_DIRECT_FP32><<<g, BLOCK, 0, st>>>(seg_offset, nseg, tstart, x, y, z, mass, P, out_phi);
  HIPCHK(hipGetLastError());
  int h_bad = 0;
  HIPCHK(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h_bad) return fail(HBT_ERR_INVALID_ARG, "seg_offset not ascending");
  return HBT_OK;
}

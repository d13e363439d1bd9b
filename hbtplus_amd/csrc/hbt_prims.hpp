// hbt_prims.hpp -- segmented device primitives for gfx950 (wave64).
//
// Every batch is a set of segments (one per subhalo).  Segment k owns the
// slot range [base[k], base[k] + cap[k]) of every per-particle work array;
// the number of live elements of the current operation is a per-segment
// device value (n[k]), so launches have static shapes (graph friendly) and
// work inside them is data dependent.  Segments are cut into TILE-element
// tiles (Tile table, built once per batch) for block-level work.
//
// Primitives:
//   segmented sort of (u64 key, u32 value) pairs by (key, value):
//     tile_sort_kernel    : LDS bitonic sort of each tile
//     sort_wave_kernel    : the same for segments of at most SMALL_CAP, one per wave
//     merge_pass_kernel   : one bottom-up merge level (co-rank by binary search)
//   segmented exclusive scan (load/store functors):
//     scan_tiles_kernel / scan_add_kernel, scan_wave_kernel
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hbt {

typedef unsigned long long u64;
typedef unsigned int u32;

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int TILE = 1024;           // elements per tile (4 per thread)
constexpr int ITEMS = TILE / BLOCK;

struct Tile {
  int seg;     // segment index
  int start;   // local offset of the tile inside the segment
};

__device__ __forceinline__ bool pair_less(u64 ka, u32 va, u64 kb, u32 vb) {
  return ka < kb || (ka == kb && va < vb);
}

// ---- wave / block scans ----------------------------------------------------

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    T y = __shfl_up(x, d, WAVE);
    if (lane >= d) x += y;
  }
  return x;
}

// A double moved across lanes by a DPP control (both halves).
template <int CTRL>
__device__ __forceinline__ double dpp_double(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_double(double x, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                          __builtin_amdgcn_readlane(__double2loint(x), lane));
}

// Sum of x over the wave (every lane active), returned in every lane: DPP
// butterflies within each row of 16 lanes (quad permutes, half-row and row
// mirrors), then the four row sums read from lanes 0, 16, 32 and 48.  The
// order of the additions is fixed, so the result is deterministic.
__device__ __forceinline__ double wave_total(double x) {
#pragma clang fp contract(off)
  x += dpp_double<0xb1>(x);   // quad_perm [1,0,3,2]
  x += dpp_double<0x4e>(x);   // quad_perm [2,3,0,1]
  x += dpp_double<0x141>(x);  // row_half_mirror
  x += dpp_double<0x140>(x);  // row_mirror
  return (readlane_double(x, 0) + readlane_double(x, 16)) + (readlane_double(x, 32) + readlane_double(x, 48));
}

template <class T>
__device__ __forceinline__ T wave_reduce_sum(T x) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_down(x, d, WAVE);
  return x;  // valid in lane 0
}

// Exclusive scan over the block (BLOCK threads, one value each); returns the
// exclusive prefix and writes the block total to *total.
template <class T>
__device__ __forceinline__ T block_exclusive_scan(T x, T *total) {
  __shared__ T wsum[BLOCK / WAVE];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  T inc = wave_inclusive_scan(x);
  if (lane == WAVE - 1) wsum[w] = inc;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < BLOCK / WAVE; i++) {
    if (i < w) off += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

// ---- one-wave tiles ------------------------------------------------------------
// The tile table lists a level's segments largest first, so the tiles of the
// segments of at most SMALL_CAP slots -- one tile each, most segments of a
// level -- are a suffix of it.  Their kernels give each tile one wave instead
// of one workgroup: no workgroup barriers, four tiles per workgroup, and less
// LDS per tile in flight.
constexpr int SMALL_CAP = 4 * WAVE;  // one-wave tiles: 4 elements per lane ...
constexpr int TINY_CAP = WAVE;       // ... or 1 for the segments of at most one wave

// LDS writes of the wave visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tiles t0 <= t < ntiles, one per wave (t is uniform across the wave).
#define WAVE_TILE_LOOP(t, t0, ntiles)                                                            \
  for (int t = (t0) + (int)(blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE); t < (ntiles); \
       t += (int)gridDim.x * (BLOCK / WAVE))

// ---- XCD-aware block order -----------------------------------------------------
// MI355X has 8 accelerator dies (XCDs), each with its own L2; the workgroups of
// a launch are dispatched to them round-robin (block b runs on XCD b % 8).
// xcd_block returns a virtual block index under which each XCD receives runs
// of `run` consecutive virtual blocks, so neighbouring slots (spatially close
// targets of one subhalo) share an L2.  Identity when run <= 1 or the grid is
// not a multiple of 8*run.
constexpr int NXCD = 8;

__device__ __forceinline__ int xcd_block(int run) {
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  if (run <= 1 || G % (NXCD * run) != 0) return b;
  const int x = b % NXCD, i = b / NXCD;
  return ((i / run) * NXCD + x) * run + (i % run);
}

// ---- segmented sort ----------------------------------------------------------

// Number of merge levels a segment of n elements needs after the tile sort.
__device__ __forceinline__ int merge_levels(int n) {
  int m = 0;
  for (int w = TILE; w < n; w *= 2) m++;
  return m;
}

// Sorts each tile's live elements [start, min(start+TILE, n[seg])) from
// (k0, v0) into the buffer from which the segment's merge levels end in
// (k0, v0): (k1, v1) when it needs an odd number of levels, else in place.
__global__ void __launch_bounds__(BLOCK)
tile_sort_kernel(const Tile *__restrict__ tiles, int ntiles, const int64_t *__restrict__ base,
                 const int *__restrict__ n, u64 *k0, u32 *v0, u64 *k1, u32 *v1) {
  __shared__ u64 sk[TILE];
  __shared__ u32 sv[TILE];
  __shared__ int slive[BLOCK];
  // The block's tiles (blockIdx.x + k gridDim.x) are checked BLOCK at a time,
  // one per thread, and only the non-empty ones are sorted: a sort of a few
  // entries per segment (the energy sort of a removed block) skips its empty
  // tiles without a chain of header loads.
  for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x * BLOCK) {
    const int tc = tb + (int)threadIdx.x * (int)gridDim.x;
    int live = 0;
    if (tc < ntiles) live = n[tiles[tc].seg] - tiles[tc].start > 0 ? 1 : 0;
    int nlive;
    const int at = block_exclusive_scan(live, &nlive);
    if (live) slive[at] = tc;
    __syncthreads();
    for (int li = 0; li < nlive; li++) {
      const int t = slive[li];
      const Tile tl = tiles[t];
      const int nseg = n[tl.seg];
      const int cnt = min(TILE, nseg - tl.start);
      const int64_t off = base[tl.seg] + tl.start;
      u64 *kout = (merge_levels(nseg) & 1) ? k1 : k0;
      u32 *vout = (merge_levels(nseg) & 1) ? v1 : v0;
      if (cnt == 1) {
        if (threadIdx.x == 0) { kout[off] = k0[off]; vout[off] = v0[off]; }
        continue;
      }
      int P = 1;
      while (P < cnt) P <<= 1;
      for (int i = threadIdx.x; i < P; i += BLOCK) {
        if (i < cnt) { sk[i] = k0[off + i]; sv[i] = v0[off + i]; }
        else { sk[i] = ~0ull; sv[i] = 0xffffffffu; }
      }
      __syncthreads();
      for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int p = threadIdx.x; p < (P >> 1); p += BLOCK) {
            // 2 j (p / j) + p % j, with j a power of two
            const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
            const int q = i + j;
            const bool up = (i & k) == 0;
            u64 ka = sk[i], kb = sk[q];
            u32 va = sv[i], vb = sv[q];
            if (pair_less(kb, vb, ka, va) == up) {
              sk[i] = kb; sk[q] = ka; sv[i] = vb; sv[q] = va;
            }
          }
          __syncthreads();
        }
      }
      for (int i = threadIdx.x; i < cnt; i += BLOCK) { kout[off + i] = sk[i]; vout[off + i] = sv[i]; }
      __syncthreads();
    }
    __syncthreads();  // slive is rewritten for the next batch of tiles
  }
}

// One-wave tiles of up to R * WAVE elements: lane l holds elements R l ..
// R l + R - 1 (padding sorts last), a bitonic network up to the power of two
// that covers the segment, partners less than R apart swapped in the lane and
// the others read across lanes.  Such segments need no merge level, so the
// result stays in (k0, v0).
__device__ __forceinline__ void sort_keep(bool keep_small, u64 ko, u32 vo, u64 &k, u32 &v) {
  const bool other_less = pair_less(ko, vo, k, v);
  if (keep_small ? other_less : !other_less) {
    k = ko;
    v = vo;
  }
}

// Elements a < b of one lane: the lower keeps the smaller in an ascending run.
__device__ __forceinline__ void sort_pair(bool up, u64 &ka, u32 &va, u64 &kb, u32 &vb) {
  if (pair_less(kb, vb, ka, va) == up) {
    const u64 tk = ka; ka = kb; kb = tk;
    const u32 tv = va; va = vb; vb = tv;
  }
}

template <int R>
__device__ __forceinline__ void sort_wave_tile(const Tile &tl, const int64_t *__restrict__ base,
                                               const int *__restrict__ n, u64 *k0, u32 *v0) {
  static_assert(R == 1 || R == 4, "one or four elements per lane");
  const int lane = threadIdx.x & (WAVE - 1);
  const int e0 = R * lane;
  {
    const int cnt = min(R * WAVE, n[tl.seg]);
    if (cnt <= 1) return;  // uniform across the wave
    const int64_t off = base[tl.seg];
    u64 k[R];
    u32 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int e = e0 + r;
      k[r] = e < cnt ? k0[off + e] : ~0ull;
      v[r] = e < cnt ? v0[off + e] : 0xffffffffu;
    }
    int P = 2;
    while (P < cnt) P <<= 1;
    for (int size = 2; size <= P; size <<= 1) {
      for (int j = size >> 1; j > 0; j >>= 1) {  // uniform across the wave
        if (j >= R) {
#pragma unroll
          for (int r = 0; r < R; r++) {
            const int e = e0 + r;
            const u64 ko = __shfl_xor(k[r], j / R, WAVE);
            const u32 vo = __shfl_xor(v[r], j / R, WAVE);
            // the lower element of a pair keeps the smaller one in an ascending run
            sort_keep(((e & j) == 0) == ((e & size) == 0), ko, vo, k[r], v[r]);
          }
        } else if (R == 4) {
          const bool up = (e0 & size) == 0;  // size > j: the same for the lane's pairs
          if (j == 2) {
            sort_pair(up, k[0], v[0], k[R > 2 ? 2 : 0], v[R > 2 ? 2 : 0]);
            sort_pair(up, k[R > 1 ? 1 : 0], v[R > 1 ? 1 : 0], k[R > 3 ? 3 : 0], v[R > 3 ? 3 : 0]);
          } else {
            // at size 2 the lane's two pairs run in opposite directions
            sort_pair(up, k[0], v[0], k[R > 1 ? 1 : 0], v[R > 1 ? 1 : 0]);
            sort_pair(size == 2 ? !up : up, k[R > 2 ? 2 : 0], v[R > 2 ? 2 : 0], k[R > 3 ? 3 : 0],
                      v[R > 3 ? 3 : 0]);
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int e = e0 + r;
      if (e < cnt) {
        k0[off + e] = k[r];
        v0[off + e] = v[r];
      }
    }
  }
}

// Tiles t0 <= t < tmid take four elements per lane, tmid <= t < ntiles one.
__global__ void __launch_bounds__(BLOCK)
sort_wave_kernel(const Tile *__restrict__ tiles, int t0, int tmid, int ntiles, const int64_t *__restrict__ base,
                 const int *__restrict__ n, u64 *k0, u32 *v0) {
  WAVE_TILE_LOOP(t, t0, ntiles) {
    const Tile tl = tiles[t];
    if (t < tmid) sort_wave_tile<4>(tl, base, n, k0, v0);
    else sort_wave_tile<1>(tl, base, n, k0, v0);
  }
}

// Merge level `level` (runs of width TILE << level merged pairwise).  A
// segment that needs m levels takes part in levels 0..m-1 and ping-pongs
// between the two buffers so that level m-1 writes (k0, v0); segments with
// m <= level are done and are not touched.  Merge path: one block per tile of
// output; two threads find by binary search how many of the tile's outputs come
// from the left run at its two ends, the block stages those parts of the two
// runs in LDS, and each thread finds its own split in LDS and merges
// MERGE_PER outputs.  Keys with their values are distinct, so the merged order
// is unique.
constexpr int MERGE_PER = TILE / BLOCK;

// Outputs among the first d of the merge of A (la) and B (lb) that come from A.
template <class GetA, class GetB>
__device__ __forceinline__ int merge_split(GetA a, GetB b, int la, int lb, int d) {
  int lo = d > lb ? d - lb : 0, hi = d < la ? d : la;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    u64 ka, kb;
    u32 va, vb;
    a(mid, ka, va);
    b(d - 1 - mid, kb, vb);
    if (pair_less(ka, va, kb, vb)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(BLOCK)
merge_pass_kernel(const Tile *__restrict__ tiles, int ntiles, const int64_t *__restrict__ base,
                  const int *__restrict__ n, int level, u64 *k0, u32 *v0, u64 *k1, u32 *v1) {
  __shared__ u64 sk[TILE];
  __shared__ u32 sv[TILE];
  __shared__ int ssplit[2];
  const int width = TILE << level;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const int cnt = n[tl.seg];
    if (cnt <= width || tl.start >= cnt) continue;  // uniform across the block
    const int m = merge_levels(cnt);
    // buffer holding the input of this level: the tile sort wrote parity m&1
    const bool in1 = ((m + level) & 1) != 0;
    const u64 *kin = in1 ? k1 : k0;
    const u32 *vin = in1 ? v1 : v0;
    u64 *kout = in1 ? k0 : k1;
    u32 *vout = in1 ? v0 : v1;
    const int64_t bs = base[tl.seg];
    const int a0 = tl.start / (2 * width) * (2 * width);  // start of the merged pair
    const int la = min(width, cnt - a0);
    const int lb = max(0, min(width, cnt - a0 - width));
    const u64 *ka = kin + bs + a0, *kb = ka + width;
    const u32 *va = vin + bs + a0, *vb = va + width;
    const int o0 = tl.start - a0, o1 = min(tl.start + TILE, cnt) - a0;  // outputs of this tile
    auto ga = [&](int i, u64 &k, u32 &v) { k = ka[i]; v = va[i]; };
    auto gb = [&](int i, u64 &k, u32 &v) { k = kb[i]; v = vb[i]; };
    if (threadIdx.x < 2) ssplit[threadIdx.x] = merge_split(ga, gb, la, lb, threadIdx.x == 0 ? o0 : o1);
    __syncthreads();
    const int i0 = ssplit[0], i1 = ssplit[1];
    const int j0 = o0 - i0, j1 = o1 - i1;
    const int na = i1 - i0, nb = j1 - j0;  // na + nb = o1 - o0 <= TILE
    for (int i = threadIdx.x; i < na; i += BLOCK) { sk[i] = ka[i0 + i]; sv[i] = va[i0 + i]; }
    for (int i = threadIdx.x; i < nb; i += BLOCK) { sk[na + i] = kb[j0 + i]; sv[na + i] = vb[j0 + i]; }
    __syncthreads();
    const int d = (int)threadIdx.x * MERGE_PER;
    if (d < na + nb) {
      auto la_ = [&](int i, u64 &k, u32 &v) { k = sk[i]; v = sv[i]; };
      auto lb_ = [&](int i, u64 &k, u32 &v) { k = sk[na + i]; v = sv[na + i]; };
      int i = merge_split(la_, lb_, na, nb, d), j = d - i;
      const int e = min(d + MERGE_PER, na + nb);
      u64 *ko = kout + bs + a0 + o0;
      u32 *vo = vout + bs + a0 + o0;
      for (int o = d; o < e; o++) {
        const bool take_a = j >= nb || (i < na && pair_less(sk[i], sv[i], sk[na + j], sv[na + j]));
        if (take_a) { ko[o] = sk[i]; vo[o] = sv[i]; i++; }
        else { ko[o] = sk[na + j]; vo[o] = sv[na + j]; j++; }
      }
    }
    __syncthreads();
  }
}

// ---- segmented exclusive scan ----------------------------------------------
// Load(seg, local) -> T for local < n[seg]; Store(seg, local, exclusive prefix).
// Pass 1 scans each tile and stores tile-local prefixes and the tile total;
// pass 2 adds to every tile the totals of the tiles before it in its segment.

template <class T, class Load, class Store>
__global__ void __launch_bounds__(BLOCK)
scan_tiles_kernel(const Tile *__restrict__ tiles, int ntiles, const int *__restrict__ n,
                  T *__restrict__ tile_sum, Load load, Store store) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const int cnt = min(TILE, n[tl.seg] - tl.start);
    if (cnt <= 0) {
      if (threadIdx.x == 0) tile_sum[t] = T(0);
      continue;
    }
    T v[ITEMS];
    T s = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
      const int l = threadIdx.x * ITEMS + i;
      v[i] = l < cnt ? load(tl.seg, tl.start + l) : T(0);
      s += v[i];
    }
    T total;
    T off = block_exclusive_scan(s, &total);
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
      const int l = threadIdx.x * ITEMS + i;
      if (l < cnt) store(tl.seg, tl.start + l, off);
      off += v[i];
    }
    if (threadIdx.x == 0) tile_sum[t] = total;
  }
}

// Pass 1 for one-wave tiles of up to R * WAVE elements (a segment's only
// tile: its prefixes are final); lane l holds elements R l .. R l + R - 1.
template <int R, class T, class Load, class Store>
__device__ __forceinline__ void scan_wave_tile(const Tile &tl, int t, const int *__restrict__ n,
                                               T *__restrict__ tile_sum, Load load, Store store) {
  const int lane = threadIdx.x & (WAVE - 1);
  {
    const int cnt = min(R * WAVE, n[tl.seg]);
    if (cnt <= 0) {  // uniform across the wave
      if (lane == 0) tile_sum[t] = T(0);
      return;
    }
    T v[R];
    T sum = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int e = R * lane + r;
      v[r] = e < cnt ? load(tl.seg, e) : T(0);
      sum += v[r];
    }
    const T inc = wave_inclusive_scan(sum);
    T off = inc - sum;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int e = R * lane + r;
      if (e < cnt) store(tl.seg, e, off);
      off += v[r];
    }
    if (lane == WAVE - 1) tile_sum[t] = inc;
  }
}

// Tiles t0 <= t < tmid take four elements per lane, tmid <= t < ntiles one.
template <class T, class Load, class Store>
__global__ void __launch_bounds__(BLOCK)
scan_wave_kernel(const Tile *__restrict__ tiles, int t0, int tmid, int ntiles, const int *__restrict__ n,
                 T *__restrict__ tile_sum, Load load, Store store) {
  WAVE_TILE_LOOP(t, t0, ntiles) {
    const Tile tl = tiles[t];
    if (t < tmid) scan_wave_tile<4, T>(tl, t, n, tile_sum, load, store);
    else scan_wave_tile<1, T>(tl, t, n, tile_sum, load, store);
  }
}

// Pass 2: adds to each tile the sum of the earlier tiles of its segment (the
// tiles of a segment are consecutive in the table; tile t is the
// start/TILE-th of its segment).
template <class T, class Update>
__global__ void __launch_bounds__(BLOCK)
scan_add_kernel(const Tile *__restrict__ tiles, int ntiles, const int *__restrict__ n,
                const T *__restrict__ tile_sum, Update update) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    if (tl.start == 0) continue;  // first tile of a segment has offset 0
    const int cnt = min(TILE, n[tl.seg] - tl.start);
    if (cnt <= 0) continue;
    const int t0 = t - tl.start / TILE;
    T part = 0;
    for (int u = t0 + (int)threadIdx.x; u < t; u += BLOCK) part += tile_sum[u];
    T o;
    (void)block_exclusive_scan(part, &o);
    for (int l = threadIdx.x; l < cnt; l += BLOCK) update(tl.seg, tl.start + l, o);
  }
}

}  // namespace hbt

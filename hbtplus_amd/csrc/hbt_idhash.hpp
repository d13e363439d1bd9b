// hbt_idhash.hpp -- hbt_id_to_index: particle ID -> snapshot index on the
// device (the lookup of ParticleSnapshot_t::GetIndices / MappedIndexTable_t::
// GetIndex, src/hash_remote.tpp:103-109, src/hash.tpp:47-56).  Open
// addressing with linear probing over power-of-two tables of (key, index)
// pairs; a key appearing twice keeps its smallest index.  Negative or absent
// IDs give -1.  Included by hbt_unbind.hip.
//
// The table is built partitioned: NP = 2^pb partitions of PART_SLOTS slots,
// a key's partition being the top pb bits of its hash and its home slot the
// low bits, probing wrapping inside the partition.  The (key, position) pairs
// are grouped by partition in two scatter passes (first by the top pb1 bits,
// then by the rest inside each first-level bucket), each pass staging a chunk
// in LDS so that every bucket's keys leave in one run; one workgroup per
// partition then inserts its keys with LDS atomics and writes the partition
// out whole.  HBM sees streaming reads and writes only, where inserting into
// one table of the snapshot's size costs two random atomics per key.  A
// partition with more than PART_MAX keys (many repeats of one ID, say)
// sends the whole call to the single-table build.
#pragma once

namespace {

__device__ __forceinline__ unsigned long long id_hash(long long key) {
  unsigned long long x = (unsigned long long)key;
  x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
  return x;
}

// Slot h holds its key at tab[2h] and its index at tab[2h+1]: a probe and the
// index it leads to share a cache line.
constexpr unsigned long long HASH_EMPTY = ~0ull, HASH_NO_INDEX = 0x7fffffffffffffffull;

// ---- single table (any input) ------------------------------------------------

__global__ void hash_clear_kernel(unsigned long long *tab, int64_t size) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < size; i += (int64_t)gridDim.x * BLOCK) {
    tab[2 * i] = HASH_EMPTY;
    tab[2 * i + 1] = HASH_NO_INDEX;
  }
}

__global__ void hash_insert_kernel(unsigned long long *tab, int64_t mask, const int64_t *ids, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const long long key = ids[i];
    if (key < 0) continue;
    int64_t h = (int64_t)(id_hash(key) & (unsigned long long)mask);
    while (true) {
      const unsigned long long prev = atomicCAS(&tab[2 * h], HASH_EMPTY, (unsigned long long)key);
      if (prev == HASH_EMPTY || prev == (unsigned long long)key) {
        atomicMin(&tab[2 * h + 1], (unsigned long long)i);  // a repeated key keeps its smallest index
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

__global__ void hash_query_kernel(const unsigned long long *tab, int64_t mask, const int64_t *q, int64_t n,
                                  int64_t *out) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const long long key = q[i];
    long long r = -1;
    if (key >= 0) {
      int64_t h = (int64_t)(id_hash(key) & (unsigned long long)mask);
      while (true) {
        const unsigned long long k = tab[2 * h];
        if (k == (unsigned long long)key) { r = (long long)tab[2 * h + 1]; break; }
        if (k == HASH_EMPTY) break;
        h = (h + 1) & mask;
      }
    }
    out[i] = r;
  }
}

// ---- partitioned table ---------------------------------------------------------

constexpr int PART_SLOTS = 4096;   // slots per partition (48 KB of LDS while it is built)
constexpr int PART_MAX = 3584;     // keys, repeats included, a partition may hold
constexpr int PART_MEAN = 2560;    // pb is chosen so the mean partition holds at most this many
constexpr int PART_L1_BITS = 8;    // first-level buckets: at most 256
constexpr int PART_L2_BITS = 10;   // second-level buckets per first-level bucket: at most 1024
constexpr int PART_CHUNK = 4096;   // keys a workgroup scatters at a time

__device__ __forceinline__ u32 part_of(unsigned long long h, int pb) { return pb ? (u32)(h >> (64 - pb)) : 0u; }

template <int NB>
struct ScatterLds {
  u64 key[PART_CHUNK];  // the chunk's pairs, grouped by bucket
  u32 pos[PART_CHUNK];
  u32 cnt[NB];          // keys of the chunk per bucket
  u32 off[NB];          // their first place in the grouped chunk
  u32 dst[NB];          // their first place in the output
};

// Scatters pairs c0 <= e < c1 (load(e, key, pos) gives pair e, false for a
// null ID) to the output at the buckets' cursors: counts per bucket in LDS,
// one run per bucket reserved at its cursor, the pairs grouped by bucket in
// LDS and written out run by run.  Called by the whole workgroup.
template <int NB, class Load, class Bucket>
__device__ __forceinline__ void scatter_chunk(ScatterLds<NB> &L, int nb, int64_t c0, int64_t c1, Load load,
                                              Bucket bucket, u32 *cursor, u64 *okey, u32 *opos) {
  constexpr int PER = PART_CHUNK / BLOCK;
  constexpr int PT = NB / BLOCK;  // buckets per thread in the scan
  static_assert(NB % BLOCK == 0, "bucket table must be a multiple of the block");
  for (int b = threadIdx.x; b < nb; b += BLOCK) L.cnt[b] = 0;
  __syncthreads();
  u64 key[PER];
  u32 pos[PER], rank[PER];
  int bk[PER];
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int64_t e = c0 + (int64_t)j * BLOCK + threadIdx.x;
    bk[j] = -1;
    if (e < c1 && load(e, key[j], pos[j])) {
      bk[j] = bucket(key[j]);
      rank[j] = atomicAdd(&L.cnt[bk[j]], 1u);
    }
  }
  __syncthreads();
  u32 c[PT], sum = 0;
#pragma unroll
  for (int q = 0; q < PT; q++) {
    const int b = threadIdx.x * PT + q;
    c[q] = b < nb ? L.cnt[b] : 0u;
    sum += c[q];
  }
  u32 total;
  u32 o = block_exclusive_scan(sum, &total);
#pragma unroll
  for (int q = 0; q < PT; q++) {
    const int b = threadIdx.x * PT + q;
    if (b < nb) {
      L.off[b] = o;
      L.dst[b] = c[q] ? atomicAdd(&cursor[b], c[q]) : 0u;
      o += c[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; j++)
    if (bk[j] >= 0) {
      const u32 p = L.off[bk[j]] + rank[j];
      L.key[p] = key[j];
      L.pos[p] = pos[j];
    }
  __syncthreads();
  for (u32 p = threadIdx.x; p < total; p += BLOCK) {
    const u64 k = L.key[p];
    const int b = bucket(k);
    const u32 g = L.dst[b] + (p - L.off[b]);
    okey[g] = k;
    opos[g] = L.pos[p];
  }
  __syncthreads();
}

// First level: keys per bucket (top pb1 bits of the partition).
__global__ void __launch_bounds__(BLOCK) part_l1_count_kernel(const int64_t *ids, int64_t n, int pb, int pb1,
                                                              u32 *total) {
  __shared__ u32 h[1 << PART_L1_BITS];
  const int nb = 1 << pb1;
  for (int b = threadIdx.x; b < nb; b += BLOCK) h[b] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const long long key = ids[i];
    if (key >= 0) atomicAdd(&h[part_of(id_hash(key), pb) >> (pb - pb1)], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += BLOCK)
    if (h[b]) atomicAdd(&total[b], h[b]);
}

// Bucket starts (base[0..nb]) and cursors from the counts; with one level the
// buckets are the partitions, and the largest count is kept in *maxcnt.
__global__ void __launch_bounds__(BLOCK) part_l1_base_kernel(const u32 *total, int nb, int one_level, u32 *base,
                                                             u32 *cursor, u32 *maxcnt) {
  const u32 c = (int)threadIdx.x < nb ? total[threadIdx.x] : 0u;
  u32 all;
  const u32 o = block_exclusive_scan(c, &all);
  if ((int)threadIdx.x < nb) {
    base[threadIdx.x] = o;
    cursor[threadIdx.x] = o;
    if (one_level) atomicMax(maxcnt, c);
  }
  if (threadIdx.x == 0) base[nb] = all;
}

__global__ void __launch_bounds__(BLOCK) part_l1_scatter_kernel(const int64_t *ids, int64_t n, int pb, int pb1,
                                                                u32 *cursor, u64 *okey, u32 *opos) {
  __shared__ ScatterLds<1 << PART_L1_BITS> L;
  const int nb = 1 << pb1;
  for (int64_t c0 = (int64_t)blockIdx.x * PART_CHUNK; c0 < n; c0 += (int64_t)gridDim.x * PART_CHUNK)
    scatter_chunk(
        L, nb, c0, c0 + PART_CHUNK < n ? c0 + PART_CHUNK : n,
        [&](int64_t e, u64 &k, u32 &p) {
          const long long v = ids[e];
          k = (u64)v;
          p = (u32)e;
          return v >= 0;
        },
        [&](u64 k) { return (int)(part_of(id_hash((long long)k), pb) >> (pb - pb1)); }, cursor, okey, opos);
}

// Second level: workgroup w handles share w % G of first-level bucket w / G.
__device__ __forceinline__ void l2_share(const u32 *base, int w, int G, int64_t &s0, int64_t &s1, int &q) {
  q = w / G;
  const int g = w % G;
  const int64_t lo = base[q], len = (int64_t)base[q + 1] - lo;
  s0 = lo + len * g / G;
  s1 = lo + len * (g + 1) / G;
}

__global__ void __launch_bounds__(BLOCK) part_l2_count_kernel(const u64 *ikey, const u32 *base, int pb, int pb2,
                                                              int G, u32 *count) {
  __shared__ u32 h[1 << PART_L2_BITS];
  const int nb = 1 << pb2;
  for (int b = threadIdx.x; b < nb; b += BLOCK) h[b] = 0;
  __syncthreads();
  int64_t s0, s1;
  int q;
  l2_share(base, blockIdx.x, G, s0, s1, q);
  for (int64_t e = s0 + threadIdx.x; e < s1; e += BLOCK)
    atomicAdd(&h[part_of(id_hash((long long)ikey[e]), pb) & (u32)(nb - 1)], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += BLOCK)
    if (h[b]) atomicAdd(&count[(int64_t)q * nb + b], h[b]);
}

// Partition starts and cursors inside first-level bucket blockIdx.x; the
// largest partition count goes to *maxcnt; part_off[np] = all keys.
__global__ void __launch_bounds__(BLOCK) part_l2_base_kernel(const u32 *count, const u32 *base, int pb2, int np,
                                                             u32 *part_off, u32 *cursor, u32 *maxcnt) {
  constexpr int PT = (1 << PART_L2_BITS) / BLOCK;
  const int nb = 1 << pb2, q = blockIdx.x;
  u32 c[PT], sum = 0, mx = 0;
#pragma unroll
  for (int j = 0; j < PT; j++) {
    const int b = threadIdx.x * PT + j;
    c[j] = b < nb ? count[(int64_t)q * nb + b] : 0u;
    sum += c[j];
    mx = c[j] > mx ? c[j] : mx;
  }
  u32 all;
  u32 o = base[q] + block_exclusive_scan(sum, &all);
#pragma unroll
  for (int j = 0; j < PT; j++) {
    const int b = threadIdx.x * PT + j;
    if (b < nb) {
      part_off[(int64_t)q * nb + b] = o;
      cursor[(int64_t)q * nb + b] = o;
      o += c[j];
    }
  }
  if (mx) atomicMax(maxcnt, mx);
  if (q == 0 && threadIdx.x == 0) part_off[np] = base[gridDim.x];
}

__global__ void __launch_bounds__(BLOCK) part_l2_scatter_kernel(const u64 *ikey, const u32 *ipos, const u32 *base,
                                                                int pb, int pb2, int G, u32 *cursor, u64 *okey,
                                                                u32 *opos) {
  __shared__ ScatterLds<1 << PART_L2_BITS> L;
  const int nb = 1 << pb2;
  int64_t s0, s1;
  int q;
  l2_share(base, blockIdx.x, G, s0, s1, q);
  for (int64_t c0 = s0; c0 < s1; c0 += PART_CHUNK)
    scatter_chunk(
        L, nb, c0, c0 + PART_CHUNK < s1 ? c0 + PART_CHUNK : s1,
        [&](int64_t e, u64 &k, u32 &p) {
          k = ikey[e];
          p = ipos[e];
          return true;
        },
        [&](u64 k) { return (int)(part_of(id_hash((long long)k), pb) & (u32)(nb - 1)); },
        cursor + (int64_t)q * nb, okey, opos);
}

// One workgroup per partition: its keys inserted into an LDS table (CAS on
// the key, then min on the position), the table written out whole.
__global__ void __launch_bounds__(BLOCK) part_build_kernel(const u64 *ikey, const u32 *ipos, const u32 *part_off,
                                                           int np, unsigned long long *tab) {
  __shared__ u64 sk[PART_SLOTS];
  __shared__ u32 sp[PART_SLOTS];
  constexpr int PER = (PART_MAX + BLOCK - 1) / BLOCK;
  for (int p = blockIdx.x; p < np; p += gridDim.x) {
    // the partition's pairs in one batch of loads, issued before the table is cleared
    const u32 lo = part_off[p], hi = part_off[p + 1];
    const bool fits = hi - lo <= (u32)PART_MAX;  // uniform; the host never uses a table with a larger partition
    u64 key[PER];
    u32 pos[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const u32 e = lo + j * BLOCK + threadIdx.x;
      if (fits && e < hi) {
        key[j] = ikey[e];
        pos[j] = ipos[e];
      }
    }
    for (int j = threadIdx.x; j < PART_SLOTS; j += BLOCK) {
      sk[j] = HASH_EMPTY;
      sp[j] = 0xffffffffu;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++) {
      if (!fits || lo + j * BLOCK + threadIdx.x >= hi) continue;
      u32 s = (u32)id_hash((long long)key[j]) & (PART_SLOTS - 1);
      for (int probe = 0; probe < PART_SLOTS; probe++) {
        const u64 prev = atomicCAS(&sk[s], HASH_EMPTY, key[j]);
        if (prev == HASH_EMPTY || prev == key[j]) {
          atomicMin(&sp[s], pos[j]);
          break;
        }
        s = (s + 1) & (PART_SLOTS - 1);
      }
    }
    __syncthreads();
    ulonglong2 *out = reinterpret_cast<ulonglong2 *>(tab) + (int64_t)p * PART_SLOTS;
    for (int j = threadIdx.x; j < PART_SLOTS; j += BLOCK)
      out[j] = make_ulonglong2(sk[j], sp[j] == 0xffffffffu ? HASH_NO_INDEX : (unsigned long long)sp[j]);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BLOCK) part_query_kernel(const unsigned long long *tab, int pb, const int64_t *q,
                                                           int64_t n, int64_t *out) {
  const ulonglong2 *slot = reinterpret_cast<const ulonglong2 *>(tab);
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const long long key = q[i];
    long long r = -1;
    if (key >= 0) {
      const unsigned long long h = id_hash(key);
      const int64_t base = (int64_t)part_of(h, pb) * PART_SLOTS;
      u32 s = (u32)h & (PART_SLOTS - 1);
      for (int probe = 0; probe < PART_SLOTS; probe++) {
        const ulonglong2 v = slot[base + s];
        if (v.x == (unsigned long long)key) { r = (long long)v.y; break; }
        if (v.x == HASH_EMPTY) break;
        s = (s + 1) & (PART_SLOTS - 1);
      }
    }
    out[i] = r;
  }
}

// Partitioned build and lookup; *done = false when a partition would hold
// more than PART_MAX keys (nothing is written to out_index then).
int id_to_index_partitioned(hbt_ctx *ctx, int64_t n_table, const int64_t *table_ids, int64_t n_query,
                            const int64_t *query_ids, int64_t *out_index, bool *done) {
  *done = false;
  int pb = 0;
  while (pb < PART_L1_BITS + PART_L2_BITS && (n_table >> pb) > PART_MEAN) pb++;
  if ((n_table >> pb) > PART_MEAN || n_table >= ((int64_t)1 << 31)) return HBT_OK;
  const int pb1 = pb < PART_L1_BITS ? pb : PART_L1_BITS, pb2 = pb - pb1;
  const int nb1 = 1 << pb1, np = 1 << pb;
  const int64_t n = n_table;
  // pairs: two (key, position) arrays; counts: total/base/cursor of the first
  // level, count/part_off/cursor of the second, the largest partition
  RETCHK(ctx->idpart.ensure((size_t)n * 24 + 64));
  u64 *k1 = ctx->idpart.as<u64>(), *k2 = k1 + n;
  u32 *p1 = reinterpret_cast<u32 *>(k2 + n), *p2 = p1 + n;
  const size_t ncnt = 3 * (size_t)nb1 + 1 + 3 * (size_t)np + 1 + 1;
  RETCHK(ctx->idcnt.ensure(ncnt * 4));
  u32 *total = ctx->idcnt.as<u32>(), *base = total + nb1, *cur1 = base + nb1 + 1;
  u32 *count = cur1 + nb1, *part_off = count + np, *cur2 = part_off + np + 1, *maxcnt = cur2 + np;
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemsetAsync(ctx->idcnt.p, 0, ncnt * 4, st));
  const int gcount = grid_for(n, ctx->cus);
  const int64_t chunks = (n + PART_CHUNK - 1) / PART_CHUNK;
  const int gscat = (int)std::max<int64_t>(1, std::min<int64_t>(chunks, (int64_t)ctx->cus * 8));
  part_l1_count_kernel<<<gcount, BLOCK, 0, st>>>(table_ids, n, pb, pb1, total);
  part_l1_base_kernel<<<1, BLOCK, 0, st>>>(total, nb1, pb2 == 0, base, cur1, maxcnt);
  part_l1_scatter_kernel<<<gscat, BLOCK, 0, st>>>(table_ids, n, pb, pb1, cur1, pb2 ? k1 : k2, pb2 ? p1 : p2);
  const u32 *off = base;
  if (pb2 > 0) {
    // workgroups per first-level bucket: about four chunks each
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>(64, n / ((int64_t)nb1 * 4 * PART_CHUNK)));
    part_l2_count_kernel<<<nb1 * G, BLOCK, 0, st>>>(k1, base, pb, pb2, G, count);
    part_l2_base_kernel<<<nb1, BLOCK, 0, st>>>(count, base, pb2, np, part_off, cur2, maxcnt);
    part_l2_scatter_kernel<<<nb1 * G, BLOCK, 0, st>>>(k1, p1, base, pb, pb2, G, cur2, k2, p2);
    off = part_off;
  }
  u32 mx = 0;
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&mx, maxcnt, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (mx > (u32)PART_MAX) return HBT_OK;
  RETCHK(ctx->idtab.ensure((size_t)np * PART_SLOTS * 16));
  unsigned long long *tab = ctx->idtab.as<unsigned long long>();
  const int gbuild = std::max(1, std::min(np, ctx->cus * 6));
  part_build_kernel<<<gbuild, BLOCK, 0, st>>>(k2, p2, off, np, tab);
  if (n_query > 0)
    part_query_kernel<<<grid_for(n_query, ctx->cus), BLOCK, 0, st>>>(tab, pb, query_ids, n_query, out_index);
  *done = true;
  return HBT_OK;
}

}  // namespace

extern "C" int hbt_id_to_index(hbt_ctx *ctx, int64_t n_table, const int64_t *table_ids, int64_t n_query,
                               const int64_t *query_ids, int64_t *out_index) {
  if (!ctx || n_table < 0 || n_query < 0 || (n_table > 0 && !table_ids) ||
      (n_query > 0 && (!query_ids || !out_index)))
    return fail(HBT_ERR_INVALID_ARG, "bad argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  bool done = false;
  if (n_table > 0)
    RETCHK(id_to_index_partitioned(ctx, n_table, table_ids, n_query, query_ids, out_index, &done));
  if (!done) {
    int64_t size = 1024;
    while (size < 2 * n_table) size <<= 1;
    // the table stays with the context, so repeated lookups against snapshots
    // of similar size allocate nothing
    RETCHK(ctx->idtab.ensure((size_t)size * 16));
    unsigned long long *tab = ctx->idtab.as<unsigned long long>();
    hash_clear_kernel<<<grid_for(size, ctx->cus), BLOCK, 0, st>>>(tab, size);
    if (n_table > 0)
      hash_insert_kernel<<<grid_for(n_table, ctx->cus), BLOCK, 0, st>>>(tab, size - 1, table_ids, n_table);
    if (n_query > 0)
      hash_query_kernel<<<grid_for(n_query, ctx->cus), BLOCK, 0, st>>>(tab, size - 1, query_ids, n_query,
                                                                       out_index);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(HBT_ERR_DEVICE, std::string("id_to_index: ") + hipGetErrorString(e));
  return HBT_OK;
}

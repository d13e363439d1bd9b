// hbt_idhash.hpp -- hbt_id_to_index: particle ID -> snapshot index on the
// device (the lookup of ParticleSnapshot_t::GetIndices / MappedIndexTable_t::
// GetIndex, src/hash_remote.tpp:103-109, src/hash.tpp:47-56).  Open
// addressing with linear probing over a power-of-two table of (key, index)
// pairs; a key appearing twice keeps its smallest index.  Negative or absent
// IDs give -1.  Included by hbt_unbind.hip.
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

}  // namespace

extern "C" int hbt_id_to_index(hbt_ctx *ctx, int64_t n_table, const int64_t *table_ids,
                               int64_t n_query, const int64_t *query_ids, int64_t *out_index) {
  if (!ctx || n_table < 0 || n_query < 0 || (n_table > 0 && !table_ids) ||
      (n_query > 0 && (!query_ids || !out_index)))
    return fail(HBT_ERR_INVALID_ARG, "bad argument");
  HIPCHK(hipSetDevice(ctx->device));
  int64_t size = 1024;
  while (size < 2 * n_table) size <<= 1;
  // the table stays with the context, so repeated lookups against snapshots
  // of similar size allocate nothing
  RETCHK(ctx->idtab.ensure((size_t)size * 16));
  unsigned long long *tab = ctx->idtab.as<unsigned long long>();
  hipStream_t st = ctx->stream;
  hash_clear_kernel<<<grid_for(size, ctx->cus), BLOCK, 0, st>>>(tab, size);
  if (n_table > 0)
    hash_insert_kernel<<<grid_for(n_table, ctx->cus), BLOCK, 0, st>>>(tab, size - 1, table_ids, n_table);
  if (n_query > 0)
    hash_query_kernel<<<grid_for(n_query, ctx->cus), BLOCK, 0, st>>>(tab, size - 1, query_ids, n_query, out_index);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(HBT_ERR_DEVICE, std::string("id_to_index: ") + hipGetErrorString(e));
  return HBT_OK;
}

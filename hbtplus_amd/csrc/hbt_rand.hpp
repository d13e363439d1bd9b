// hbt_rand.hpp -- glibc's rand() stream on the device, for the shuffle of
// sampled unbinding (std::random_shuffle at src/subhalo_unbind.cpp:302, which
// draws from the process's rand()).  Included by hbt_unbind.hip.
//
// glibc's rand() is random(3)'s TYPE_3 additive feedback generator: srand(s)
// sets r[0] = s (1 for 0) and r[i] = 16807 r[i-1] mod (2^31 - 1) for
// i = 1..30, r[i] = r[i-31] for i = 31..33; then r[i] = r[i-31] + r[i-3]
// mod 2^32, the values r[34..343] are discarded and call k returns
// r[344+k] >> 1.  A process that never calls srand draws from srand(1).
//
// The recurrence is linear over Z/2^32: the window of the last 31 values
// moves one step by the 31x31 companion matrix M (w'[u] = w[u+1], w'[30] =
// w[0] + w[28]).  The host keeps M^(2^k) and jumps to any call number; the
// device splits a run of draws into chunks of RCHUNK values, each wave jumping
// from the run's window with the powers M^(RCHUNK 2^k) and then stepping the
// recurrence itself.  The matrices are 31x31 row-major over Z/2^32.
#pragma once

#include <mutex>

namespace {

constexpr int RWIN = 31;                   // window of the recurrence
constexpr int RCHUNK = RWIN * 128;         // draws per wave (a multiple of 31: the ring starts aligned)
constexpr int RJUMPS = 24;                 // chunk jumps M^(RCHUNK 2^k), k < RJUMPS: 2^24 chunks per run

void rmat_mul(const uint32_t *a, const uint32_t *b, uint32_t *out) {
  uint32_t t[RWIN * RWIN];
  for (int i = 0; i < RWIN; i++)
    for (int j = 0; j < RWIN; j++) {
      uint32_t v = 0;
      for (int q = 0; q < RWIN; q++) v += a[i * RWIN + q] * b[q * RWIN + j];
      t[i * RWIN + j] = v;
    }
  memcpy(out, t, sizeof t);
}

void rmat_apply(const uint32_t *a, const uint32_t *w, uint32_t *out) {
  uint32_t t[RWIN];
  for (int i = 0; i < RWIN; i++) {
    uint32_t v = 0;
    for (int q = 0; q < RWIN; q++) v += a[i * RWIN + q] * w[q];
    t[i] = v;
  }
  memcpy(out, t, sizeof t);
}

// M^(2^k) for k < 48 (host jumps) and M^(RCHUNK 2^k) for k < RJUMPS (device).
struct RandJumps {
  std::vector<uint32_t> pow2, chunk;
  RandJumps() : pow2((size_t)48 * RWIN * RWIN, 0u), chunk((size_t)RJUMPS * RWIN * RWIN, 0u) {
    uint32_t *m = pow2.data();
    for (int u = 0; u + 1 < RWIN; u++) m[u * RWIN + u + 1] = 1u;
    m[(RWIN - 1) * RWIN + 0] = 1u;
    m[(RWIN - 1) * RWIN + 28] = 1u;
    for (int k = 1; k < 48; k++)
      rmat_mul(&pow2[(size_t)(k - 1) * RWIN * RWIN], &pow2[(size_t)(k - 1) * RWIN * RWIN],
               &pow2[(size_t)k * RWIN * RWIN]);
    // M^RCHUNK from the binary powers, then its squares
    std::vector<uint32_t> a((size_t)RWIN * RWIN, 0u);
    for (int u = 0; u < RWIN; u++) a[u * RWIN + u] = 1u;
    for (int k = 0; k < 48; k++)
      if (((uint64_t)RCHUNK >> k) & 1) rmat_mul(a.data(), &pow2[(size_t)k * RWIN * RWIN], a.data());
    memcpy(chunk.data(), a.data(), sizeof(uint32_t) * RWIN * RWIN);
    for (int k = 1; k < RJUMPS; k++)
      rmat_mul(&chunk[(size_t)(k - 1) * RWIN * RWIN], &chunk[(size_t)(k - 1) * RWIN * RWIN],
               &chunk[(size_t)k * RWIN * RWIN]);
  }
};

const RandJumps &rand_jumps() {
  static std::once_flag once;
  static RandJumps *j = nullptr;
  std::call_once(once, [] { j = new RandJumps(); });
  return *j;
}

// The window (r[m-31], ..., r[m-1]) before call number `pos` (m = 344 + pos)
// of the stream srand(seed).
void rand_window(uint32_t seed, int64_t pos, uint32_t w[RWIN]) {
  int32_t r[34];
  int32_t word = (int32_t)(seed == 0 ? 1u : seed);
  r[0] = word;
  for (int i = 1; i < 31; i++) {  // 16807 word mod (2^31 - 1), Schrage's method as glibc writes it
    const long hi = word / 127773, lo = word % 127773;
    word = (int32_t)(16807 * lo - 2836 * hi);
    if (word < 0) word += 2147483647;
    r[i] = word;
  }
  for (int i = 31; i < 34; i++) r[i] = r[i - 31];
  for (int u = 0; u < RWIN; u++) w[u] = (uint32_t)r[3 + u];  // window before index 34
  const uint64_t steps = 310ull + (uint64_t)pos;
  const RandJumps &J = rand_jumps();
  for (int k = 0; k < 48; k++)
    if ((steps >> k) & 1) rmat_apply(&J.pow2[(size_t)k * RWIN * RWIN], w, w);
}

// draws[t] = call pos + t of the stream for t < n, where w0 is the window
// before call pos.  One wave per chunk of RCHUNK draws: lanes 0..30 hold the
// chunk's window after the jumps M^(RCHUNK 2^k) of the chunk number's bits;
// every lane then holds the whole window and steps the recurrence (the same
// values in every lane), lane u storing value u of each group of 31.
__global__ void __launch_bounds__(BLOCK) rand_gen_kernel(const uint32_t *__restrict__ jumps,
                                                         const uint32_t *__restrict__ w0, int64_t n,
                                                         int *__restrict__ draws) {
  const int lane = (int)(threadIdx.x & (WAVE - 1));
  const int64_t nchunk = (n + RCHUNK - 1) / RCHUNK;
  for (int64_t c = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE; c < nchunk;
       c += (int64_t)gridDim.x * BLOCK / WAVE) {
    uint32_t cur = lane < RWIN ? w0[lane] : 0u;
    for (int k = 0; k < RJUMPS; k++) {
      if (!((c >> k) & 1)) continue;
      const uint32_t *A = jumps + (size_t)k * RWIN * RWIN;
      uint32_t v = 0u;
      for (int q = 0; q < RWIN; q++) {
        const uint32_t wq = (uint32_t)__shfl((int)cur, q, WAVE);
        if (lane < RWIN) v += A[lane * RWIN + q] * wq;
      }
      cur = v;
    }
    uint32_t w[RWIN];
#pragma unroll
    for (int q = 0; q < RWIN; q++) w[q] = (uint32_t)__shfl((int)cur, q, WAVE);
    const int64_t base = c * RCHUNK;
    for (int g = 0; g < RCHUNK / RWIN; g++) {
      uint32_t mine = 0u;
#pragma unroll
      for (int u = 0; u < RWIN; u++) {
        w[u] += w[(u + 28) % RWIN];
        mine = lane == u ? w[u] : mine;
      }
      const int64_t t = base + (int64_t)g * RWIN + lane;
      if (lane < RWIN && t < n) draws[t] = (int)(mine >> 1);
    }
  }
}

}  // namespace

/*
 * Post-processing of i-sorted pair lists on the GPU:
 *   first_neighbours  row starts of an i-sorted list
 *                     (src/libneighbours/first_neighbours.cc:15-40)
 *   count_jumps       distinct values of a sorted array, the n of
 *                     get_jump_indicies (first_neighbours.cc:42-55)
 *   triplets          pairs of pairs sharing the first atom
 *                     (src/libneighbours/triplet_list.cc:15-34)
 */
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mnl_engine.hh"

namespace mnl {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

__global__ void k_seed_fill(int32_t *seed, int64_t n, int32_t v) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) seed[t] = v;
}

/* seed[l] = first pair of row l: the pair at k starts rows i[k-1]+1 .. i[k]
   when i changes at k (rows without pairs point at the next row's start),
   rows before i[0] keep -1, rows after the last get nn. Indices outside
   [0, n] are ignored (never written out of bounds). ik = i[k], ip = i[k-1]
   (unused for k = 0). */
__device__ __forceinline__ void seed_at(int64_t k, int32_t ik, int32_t ip, int64_t nn,
                                        int32_t n, int32_t *__restrict__ seed) {
    if (k == 0) {
        if (ik >= 0 && ik <= n) seed[ik] = 0;
    } else {
        const int32_t hi = ik < n ? ik : n;
        for (int32_t l = (ip + 1 > 0 ? ip + 1 : 0); l <= hi; l++) seed[l] = static_cast<int32_t>(k);
    }
    if (k == nn - 1)
        for (int32_t l = (ik + 1 > 0 ? ik + 1 : 0); l <= n; l++) seed[l] = static_cast<int32_t>(nn);
}

/* Pairs k0 .. nn-1, one per thread. */
__global__ void k_first_neighbours(const int32_t *__restrict__ i, int64_t k0, int64_t nn,
                                   int32_t n, int32_t *__restrict__ seed) {
    const int64_t k = k0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    seed_at(k, i[k], k > 0 ? i[k - 1] : 0, nn, n, seed);
}

/* Pairs 0 .. 4 nq - 1, four per thread from one 16-byte load (i 16-byte
   aligned). */
__global__ void k_first_neighbours4(const int32_t *__restrict__ i, int64_t nq, int64_t nn,
                                    int32_t n, int32_t *__restrict__ seed) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int4 v = reinterpret_cast<const int4 *>(i)[q];
    const int64_t k = 4 * q;
    seed_at(k, v.x, k > 0 ? i[k - 1] : 0, nn, n, seed);
    seed_at(k + 1, v.y, v.x, nn, n, seed);
    seed_at(k + 2, v.z, v.y, nn, n, seed);
    seed_at(k + 3, v.w, v.z, nn, n, seed);
}

/* Number of positions where a sorted array changes value (block sums, one
   atomic per block). */
__global__ __launch_bounds__(kBlock) void k_count_jumps(const int32_t *__restrict__ s,
                                                        int64_t nn,
                                                        int32_t *__restrict__ count) {
    __shared__ int32_t s_w[kBlock / kWave];
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int v = (k + 1 < nn && s[k] != s[k + 1]) ? 1 : 0;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if (__lane_id() == 0) s_w[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kBlock / kWave; w++) t += s_w[w];
        if (t) atomicAdd(count, t);
    }
}

/* With a distance cutoff: one wave per 64 consecutive rows. Their pairs
   form one contiguous range [lo, hi), which the wave compacts chunk by
   chunk (kept: !(d >= cutoff), so NaN distances are kept, as in the
   reference's test) into vidx[lo ...] in ascending order. A row's kept pairs
   then start at vstart = lo + (kept pairs before the row's first pair),
   read off the ballot of the chunk holding that first pair; m = the next
   row's start minus its own, tcnt = m (m - 1). Rows whose first index is
   negative (before the first atom with pairs) are empty. */
__global__ __launch_bounds__(kBlock) void k_triplet_rows(const int32_t *__restrict__ first_i,
                                                         int32_t nrows, int64_t nn,
                                                         const double *__restrict__ absdist,
                                                         double cutoff, int32_t *__restrict__ m,
                                                         int64_t *__restrict__ tcnt,
                                                         int32_t *__restrict__ vstart,
                                                         int32_t *__restrict__ vidx) {
    const int64_t r0 =
        (static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave) * kWave;
    if (r0 >= nrows) return;
    const int lane = __lane_id();
    const int nr = static_cast<int>(nrows - r0 < kWave ? nrows - r0 : kWave);
    const int32_t b = lane < nr ? first_i[r0 + lane] : -1;
    const int64_t end_first = first_i[r0 + nr];
    const int64_t hi = end_first < nn ? end_first : nn;
    int32_t lo = b >= 0 ? b : INT32_MAX;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) lo = min(lo, __shfl_xor(lo, o, kWave));
    int gp = -1; /* kept pairs in [lo, b) */
    int run = 0;
    for (int64_t k0 = lo; k0 < hi; k0 += kWave) {
        const int64_t k = k0 + lane;
        const bool keep = k < hi && !(absdist[k] >= cutoff);
        const uint64_t mask = __builtin_amdgcn_ballot_w64(keep);
        const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                  __builtin_amdgcn_mbcnt_lo(
                                                      static_cast<uint32_t>(mask), 0u));
        if (keep) vidx[lo + run + pre] = static_cast<int32_t>(k);
        if (b >= k0 && b < k0 + kWave)
            gp = run + static_cast<int>(__popcll(mask & ((1ull << (b - k0)) - 1)));
        run += static_cast<int>(__popcll(mask));
    }
    if (b >= 0 && gp < 0) gp = run; /* first pair at or past hi: an empty row */
    const int gnext_l = __shfl_down(gp, 1, kWave);
    const int gnext = lane == nr - 1 ? run : gnext_l;
    if (lane < nr) {
        const int cnt = b >= 0 ? gnext - gp : 0;
        m[r0 + lane] = cnt;
        tcnt[r0 + lane] = static_cast<int64_t>(cnt) * (cnt > 0 ? cnt - 1 : 0);
        vstart[r0 + lane] = b >= 0 ? lo + gp : 0;
    }
}

/* Rows without a distance cutoff keep every pair: one thread per row,
   m = the row's length, tcnt = m (m - 1); the kept indices are the pair
   indices themselves (no compaction). */
__global__ __launch_bounds__(kBlock) void k_triplet_rows_all(const int32_t *__restrict__ first_i,
                                                             int32_t nrows, int64_t nn,
                                                             int32_t *__restrict__ m,
                                                             int64_t *__restrict__ tcnt) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= nrows) return;
    const int32_t b = first_i[r];
    const int64_t e = first_i[r + 1] < nn ? first_i[r + 1] : nn;
    const int cnt = b >= 0 && e > b ? static_cast<int>(e - b) : 0;
    m[r] = cnt;
    tcnt[r] = static_cast<int64_t>(cnt) * (cnt > 0 ? cnt - 1 : 0);
}

/* One wave per 64 consecutive rows (their m, first index and triplet
   offset arrive in one coalesced load each). Triplet t of a row is
   (kept[a], kept[c]) with a = t / (m - 1) and c the (t mod (m - 1))-th kept
   index other than a: ij outer, ik inner, both ascending - the reference's
   loop order. Lanes take t = lane, lane + 64, ...; (a, t mod (m - 1)) steps
   by (64 / (m - 1), 64 mod (m - 1)) with one carry, so there is one division
   per row, not per triplet. kept[x] is vidx[b + x] with b the row's start
   in the compacted table, or b + x with b its first pair when vidx is NULL
   (every pair kept). */
__global__ __launch_bounds__(kBlock) void k_triplet_fill(const int32_t *__restrict__ vbase,
                                                         int32_t nrows,
                                                         const int32_t *__restrict__ m,
                                                         const int64_t *__restrict__ toff,
                                                         const int32_t *__restrict__ vidx,
                                                         int32_t *__restrict__ ij,
                                                         int32_t *__restrict__ ik) {
    const int64_t r0 =
        (static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave) * kWave;
    if (r0 >= nrows) return;
    const int lane = __lane_id();
    const int64_t rl = r0 + lane;
    int m_l = 0, b_l = 0;
    int64_t base_l = 0;
    if (rl < nrows) {
        m_l = m[rl];
        b_l = vbase[rl];
        base_l = toff[rl];
    }
    const int nr = static_cast<int>(nrows - r0 < kWave ? nrows - r0 : kWave);
    for (int q = 0; q < nr; q++) {
        const int mr = __builtin_amdgcn_readlane(m_l, q);
        if (mr < 2) continue;
        const int b = __builtin_amdgcn_readlane(b_l, q);
        const int64_t base =
            static_cast<int64_t>(static_cast<uint32_t>(
                __builtin_amdgcn_readlane(static_cast<int>(base_l), q))) |
            (static_cast<int64_t>(__builtin_amdgcn_readlane(static_cast<int>(base_l >> 32), q))
             << 32);
        const int d = mr - 1;
        const int64_t total = static_cast<int64_t>(mr) * d;
        const int qa = kWave / d, rem = kWave - qa * d;
        int a = lane / d;
        int c0 = lane - a * d;
        for (int64_t t = lane; t < total; t += kWave) {
            const int c = c0 + (c0 >= a ? 1 : 0);
            ij[base + t] = vidx ? vidx[b + a] : b + a;
            ik[base + t] = vidx ? vidx[b + c] : b + c;
            a += qa;
            c0 += rem;
            if (c0 >= d) {
                c0 -= d;
                a++;
            }
        }
    }
}

}  // namespace

hipError_t launch_first_neighbours(const int32_t *i, int64_t nn, int32_t n, int32_t *seed,
                                   hipStream_t stream) {
    k_seed_fill<<<blocks_for(int64_t(n) + 1), kBlock, 0, stream>>>(seed, int64_t(n) + 1,
                                                                   nn > 0 ? -1 : 0);
    if (nn <= 0) return hipGetLastError();
    const bool aligned = (reinterpret_cast<uintptr_t>(i) & 15) == 0;
    const int64_t nq = aligned ? nn / 4 : 0;
    if (nq > 0)
        k_first_neighbours4<<<blocks_for(nq), kBlock, 0, stream>>>(i, nq, nn, n, seed);
    if (4 * nq < nn)
        k_first_neighbours<<<blocks_for(nn - 4 * nq), kBlock, 0, stream>>>(i, 4 * nq, nn, n,
                                                                            seed);
    return hipGetLastError();
}

hipError_t launch_count_jumps(const int32_t *sorted, int64_t nn, int32_t *count,
                              hipStream_t stream) {
    k_seed_fill<<<1, kWave, 0, stream>>>(count, 1, 0);
    if (nn > 1) k_count_jumps<<<blocks_for(nn), kBlock, 0, stream>>>(sorted, nn, count);
    return hipGetLastError();
}

hipError_t launch_triplet_rows(const int32_t *first_i, int32_t nrows, int64_t nn,
                               const double *absdist, double cutoff, int32_t *m, int64_t *tcnt,
                               int32_t *vstart, int32_t *vidx, hipStream_t stream) {
    if (nrows <= 0) return hipSuccess;
    if (!absdist) {
        k_triplet_rows_all<<<blocks_for(nrows), kBlock, 0, stream>>>(first_i, nrows, nn, m,
                                                                     tcnt);
        return hipGetLastError();
    }
    const int64_t rows_per_block = int64_t(kBlock / kWave) * kWave;
    const unsigned nb = static_cast<unsigned>((nrows + rows_per_block - 1) / rows_per_block);
    k_triplet_rows<<<nb, kBlock, 0, stream>>>(first_i, nrows, nn, absdist, cutoff, m, tcnt,
                                              vstart, vidx);
    return hipGetLastError();
}

hipError_t launch_triplet_fill(const int32_t *vbase, int32_t nrows, const int32_t *m,
                               const int64_t *toff, const int32_t *vidx, int32_t *ij,
                               int32_t *ik, hipStream_t stream) {
    if (nrows <= 0) return hipSuccess;
    const int64_t rows_per_block = int64_t(kBlock / kWave) * kWave;
    const unsigned nb = static_cast<unsigned>((nrows + rows_per_block - 1) / rows_per_block);
    k_triplet_fill<<<nb, kBlock, 0, stream>>>(vbase, nrows, m, toff, vidx, ij, ik);
    return hipGetLastError();
}

}  // namespace mnl

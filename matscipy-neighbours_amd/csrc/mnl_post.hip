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
   rows before i[0] keep -1, rows after the last get nn. */
__global__ void k_first_neighbours(const int32_t *__restrict__ i, int64_t nn, int32_t n,
                                   int32_t *__restrict__ seed) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    /* indices outside [0, n] are ignored (never written out of bounds) */
    const int32_t ik = i[k];
    if (k == 0) {
        if (ik >= 0 && ik <= n) seed[ik] = 0;
    } else {
        const int32_t ip = i[k - 1];
        const int32_t hi = ik < n ? ik : n;
        for (int32_t l = (ip + 1 > 0 ? ip + 1 : 0); l <= hi; l++) seed[l] = static_cast<int32_t>(k);
    }
    if (k == nn - 1)
        for (int32_t l = (ik + 1 > 0 ? ik + 1 : 0); l <= n; l++) seed[l] = static_cast<int32_t>(nn);
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

/* One wave per row: compact the row's kept pairs (!(d >= cutoff), so NaN
   distances are kept, as in the reference's test) into vidx in ascending
   order; m = kept count, tcnt = m (m - 1) triplets. A row starting at a
   negative index (rows before the first atom with pairs) is empty. */
__global__ __launch_bounds__(kBlock) void k_triplet_rows(const int32_t *__restrict__ first_i,
                                                         int32_t nrows, int64_t nn,
                                                         const double *__restrict__ absdist,
                                                         double cutoff, int32_t *__restrict__ m,
                                                         int64_t *__restrict__ tcnt,
                                                         int32_t *__restrict__ vidx) {
    const int r = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (r >= nrows) return;
    const int lane = __lane_id();
    const int32_t b = first_i[r];
    const int32_t e = static_cast<int32_t>(first_i[r + 1] < nn ? first_i[r + 1] : nn);
    int cnt = 0;
    if (b >= 0) {
        for (int32_t k0 = b; k0 < e; k0 += kWave) {
            const int32_t k = k0 + lane;
            const bool keep = k < e && !(absdist && absdist[k] >= cutoff);
            const uint64_t mask = __builtin_amdgcn_ballot_w64(keep);
            const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(
                                                          static_cast<uint32_t>(mask), 0u));
            if (keep) vidx[b + cnt + pre] = k;
            cnt += __popcll(mask);
        }
    }
    if (lane == 0) {
        m[r] = cnt;
        tcnt[r] = static_cast<int64_t>(cnt) * (cnt > 0 ? cnt - 1 : 0);
    }
}

/* One wave per row: triplet t of the row is (kept[a], kept[c]) with
   a = t / (m - 1) and c the (t mod (m - 1))-th kept index other than a:
   ij outer, ik inner, both ascending - the reference's loop order. */
__global__ __launch_bounds__(kBlock) void k_triplet_fill(const int32_t *__restrict__ first_i,
                                                         int32_t nrows,
                                                         const int32_t *__restrict__ m,
                                                         const int64_t *__restrict__ toff,
                                                         const int32_t *__restrict__ vidx,
                                                         int32_t *__restrict__ ij,
                                                         int32_t *__restrict__ ik) {
    const int r = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (r >= nrows) return;
    const int lane = __lane_id();
    const int mr = m[r];
    if (mr < 2) return;
    const int32_t b = first_i[r];
    const int64_t base = toff[r];
    const int64_t total = static_cast<int64_t>(mr) * (mr - 1);
    for (int64_t t = lane; t < total; t += kWave) {
        const int a = static_cast<int>(t / (mr - 1));
        const int c0 = static_cast<int>(t - static_cast<int64_t>(a) * (mr - 1));
        const int c = c0 + (c0 >= a ? 1 : 0);
        ij[base + t] = vidx[b + a];
        ik[base + t] = vidx[b + c];
    }
}

}  // namespace

hipError_t launch_first_neighbours(const int32_t *i, int64_t nn, int32_t n, int32_t *seed,
                                   hipStream_t stream) {
    k_seed_fill<<<blocks_for(int64_t(n) + 1), kBlock, 0, stream>>>(seed, int64_t(n) + 1,
                                                                   nn > 0 ? -1 : 0);
    if (nn > 0) k_first_neighbours<<<blocks_for(nn), kBlock, 0, stream>>>(i, nn, n, seed);
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
                               int32_t *vidx, hipStream_t stream) {
    if (nrows <= 0) return hipSuccess;
    const unsigned nb = static_cast<unsigned>((nrows + kBlock / kWave - 1) / (kBlock / kWave));
    k_triplet_rows<<<nb, kBlock, 0, stream>>>(first_i, nrows, nn, absdist, cutoff, m, tcnt,
                                              vidx);
    return hipGetLastError();
}

hipError_t launch_triplet_fill(const int32_t *first_i, int32_t nrows, const int32_t *m,
                               const int64_t *toff, const int32_t *vidx, int32_t *ij,
                               int32_t *ik, hipStream_t stream) {
    if (nrows <= 0) return hipSuccess;
    const unsigned nb = static_cast<unsigned>((nrows + kBlock / kWave - 1) / (kBlock / kWave));
    k_triplet_fill<<<nb, kBlock, 0, stream>>>(first_i, nrows, m, toff, vidx, ij, ik);
    return hipGetLastError();
}

}  // namespace mnl

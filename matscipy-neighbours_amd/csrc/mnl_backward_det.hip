/*
 * Deterministic backward of a pair list (mnl_plan_pair_backward with
 * deterministic != 0): the same gradients as k_pair_backward, with every sum
 * taken in a fixed order, so that the result does not depend on scheduling.
 *
 *   g_p            = dL/dD_p + (dL/dd_p / d_p) D_p   (d term dropped at d_p = 0)
 *   dL/dr[a]       = sum over pairs with j = a of g_p  -  sum over pairs with i = a of g_p
 *   dL/dcell[k][m] = sum_p S_pk g_pm over periodic directions k
 *
 * The pairs are grouped by j with a stable LSD radix sort of the pair
 * indices (digits of up to 10 bits, as few passes as the largest j needs;
 * per block a digit histogram, one exclusive scan of all histograms, then a
 * stable scatter that ranks equal digits lane by lane and wave by wave), so
 * each atom's j-side pairs sit contiguously in ascending pair order. Each
 * pair's g is computed once (into G), in the pass that also forms the cell
 * gradient's block partials. One wave
 * per atom then sums its j-side entries and its own row (the list is
 * i-sorted): lane l takes entries l, l + 64, ... of each in order, and the
 * 64 lane sums are combined by a fixed halving tree (lane l += lane l + h
 * for h = 32, 16, ..., 1). The cell gradient: block b of nb takes pairs
 * b * 256 + t, + 256 nb, ... per thread t in order, the threads are combined
 * by the halving tree over 256 threads; then thread t sums the block
 * partials t, t + 256, ... in order and the same tree combines them.
 * Nothing uses floating-point atomics.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mnl_engine.hh"

#pragma clang fp contract(off)

namespace mnl {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;
constexpr int kMaxBins = 1024;      /* digits of up to 10 bits */
constexpr int kItems = 16;          /* keys per thread per radix tile */
constexpr int kTile = kBlock * kItems;
constexpr int kCellBlocks = 1024;   /* fixed tiles of the cell gradient */

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

/* Key of pair p in the first radix pass: its j, or nat when j lies outside
   [0, nat) (such pairs sort last and enter no atom's sum). */
__device__ __forceinline__ int32_t key_of(const int32_t *__restrict__ keys,
                                          const int32_t *__restrict__ secnd, int32_t nat,
                                          int64_t p) {
    if (keys) return keys[p];
    const int32_t j = secnd[p];
    return (j >= 0 && j < nat) ? j : nat;
}

/* Digit histogram of each tile: hist[digit * nblocks + block]. */
__global__ __launch_bounds__(kBlock) void k_radix_hist(const int32_t *__restrict__ keys,
                                                       const int32_t *__restrict__ secnd,
                                                       int32_t nat, int64_t n, int shift,
                                                       int bins, int32_t *__restrict__ hist,
                                                       int nblocks) {
    __shared__ int s_h[kMaxBins];
    for (int t = threadIdx.x; t < bins; t += kBlock) s_h[t] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&s_h[(key_of(keys, secnd, nat, idx) >> shift) & (bins - 1)], 1);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < bins; t += kBlock)
        hist[static_cast<int64_t>(t) * nblocks + blockIdx.x] = s_h[t];
}

/* Stable scatter of one tile: keys in tile order, each to offs[digit][block]
   plus the number of earlier keys of the tile with the same digit. Within a
   round of 256 keys a lane's rank among equal digits comes from one ballot
   per digit bit (lanes below it in its wave) and the digit counts of the
   waves before it, which the lowest lane of each digit group posts in LDS;
   the rounds run in order, and after each the posting lanes add their
   counts to the digit's running base and clear their posts. vals == NULL:
   the values are the pair indices themselves. */
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const int32_t *__restrict__ keys, const int32_t *__restrict__ secnd, int32_t nat,
    const int32_t *__restrict__ vals, int64_t n, int shift, int dbits,
    const int64_t *__restrict__ offs, int nblocks, int32_t *__restrict__ keys_out,
    int32_t *__restrict__ vals_out) {
    __shared__ unsigned long long s_base[kMaxBins];
    __shared__ int s_cnt[kWaves][kMaxBins];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    const int bins = 1 << dbits;
    for (int t = tid; t < bins; t += kBlock) {
        s_base[t] = static_cast<unsigned long long>(offs[static_cast<int64_t>(t) * nblocks + blockIdx.x]);
#pragma unroll
        for (int q = 0; q < kWaves; q++) s_cnt[q][t] = 0;
    }
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
    const uint64_t below = (1ull << lane) - 1;
    for (int k = 0; k < kItems; k++) {
        const int64_t idx = base + k * kBlock + tid;
        const bool valid = idx < n;
        int32_t key = 0, val = 0;
        int d = 0;
        if (valid) {
            key = key_of(keys, secnd, nat, idx);
            val = vals ? vals[idx] : static_cast<int32_t>(idx);
            d = (key >> shift) & (bins - 1);
        }
        uint64_t peers = __builtin_amdgcn_ballot_w64(valid);
        for (int b = 0; b < dbits; b++) {
            const uint64_t bal = __builtin_amdgcn_ballot_w64(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        const int rank = __popcll(peers & below);
        const int cnt = __popcll(peers);
        const bool poster = valid && rank == 0;
        if (poster) s_cnt[w][d] = cnt;
        __syncthreads();
        if (valid) {
            int64_t pos = static_cast<int64_t>(s_base[d]) + rank;
            for (int q = 0; q < w; q++) pos += s_cnt[q][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        if (poster) {
            atomicAdd(&s_base[d], static_cast<unsigned long long>(cnt));
            s_cnt[w][d] = 0;
        }
        __syncthreads();
    }
}

/* start[a] = number of entries of the sorted array below a, a = 0..nat. */
__global__ void k_lower_bounds(const int32_t *__restrict__ sorted, int64_t n, int32_t nat,
                               int64_t *__restrict__ start) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a > nat) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (sorted[mid] < a)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[a] = lo;
}

/* g_p in float64, the same operations as k_pair_backward. */
template <typename T>
__device__ __forceinline__ void pair_g(int64_t p, const T *__restrict__ D, const T *__restrict__ d,
                                       const T *__restrict__ gD, const T *__restrict__ gd,
                                       double &g0, double &g1, double &g2) {
    g0 = g1 = g2 = 0.0;
    if (gD) {
        g0 = static_cast<double>(gD[3 * p + 0]);
        g1 = static_cast<double>(gD[3 * p + 1]);
        g2 = static_cast<double>(gD[3 * p + 2]);
    }
    if (gd) {
        const double dd = static_cast<double>(d[p]);
        if (dd > 0.0) {
            const double f = static_cast<double>(gd[p]) / dd;
            g0 += f * static_cast<double>(D[3 * p + 0]);
            g1 += f * static_cast<double>(D[3 * p + 1]);
            g2 += f * static_cast<double>(D[3 * p + 2]);
        }
    }
}

/* G[p] = g_p (float64, 3 per pair), and the cell gradient's partial of
   this block: thread t takes pairs b * 256 + t, + stride, ... in order, the
   threads are combined by the halving tree; partial[b][9] (periodic
   directions of S only). */
template <typename T>
__global__ __launch_bounds__(kBlock) void k_det_pair_g(int64_t npairs, const int32_t *__restrict__ shift,
                                                       const T *__restrict__ D,
                                                       const T *__restrict__ d,
                                                       const T *__restrict__ gD,
                                                       const T *__restrict__ gd, int pbc_mask,
                                                       double *__restrict__ G,
                                                       double *__restrict__ partial) {
    __shared__ double s_c[9][kBlock];
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; p < npairs;
         p += stride) {
        double g0, g1, g2;
        pair_g<T>(p, D, d, gD, gd, g0, g1, g2);
        G[3 * p + 0] = g0;
        G[3 * p + 1] = g1;
        G[3 * p + 2] = g2;
        const double s0 = (pbc_mask & 1) ? static_cast<double>(shift[3 * p + 0]) : 0.0;
        const double s1 = (pbc_mask & 2) ? static_cast<double>(shift[3 * p + 1]) : 0.0;
        const double s2 = (pbc_mask & 4) ? static_cast<double>(shift[3 * p + 2]) : 0.0;
        c[0] += s0 * g0; c[1] += s0 * g1; c[2] += s0 * g2;
        c[3] += s1 * g0; c[4] += s1 * g1; c[5] += s1 * g2;
        c[6] += s2 * g0; c[7] += s2 * g1; c[8] += s2 * g2;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) s_c[k][threadIdx.x] = c[k];
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h)
            for (int k = 0; k < 9; k++) s_c[k][threadIdx.x] += s_c[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 9) partial[static_cast<int64_t>(blockIdx.x) * 9 + threadIdx.x] =
        s_c[threadIdx.x][0];
}

/* The fixed halving tree over the 64 lanes: lane 0 ends with the sum. */
__device__ __forceinline__ double lane_tree(double v) {
#pragma unroll
    for (int h = kWave / 2; h > 0; h >>= 1) v += __shfl_down(v, h, kWave);
    return v;
}

/* One wave per atom: its j-side entries (ascending pair order) minus its
   own row, each as 64 strided lane sums combined by the fixed tree. Lane l
   adds entries l, l + 64, l + 128, ... of each row in that order, starting
   from 0.0. Each pass of the loop takes two such entries of both rows and
   issues the row loads and the permutation loads together before any sum,
   so a row of up to 128 entries costs three dependent memory round trips
   (starts, permutation with own row, gathered entries) rather than one
   per 64 entries and per load. */
__global__ __launch_bounds__(kBlock) void k_det_positions(int32_t nat,
                                                          const int64_t *__restrict__ jstart,
                                                          const int32_t *__restrict__ jperm,
                                                          const int64_t *__restrict__ istart,
                                                          const double *__restrict__ G,
                                                          double *__restrict__ grad_pos) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kWaves + threadIdx.x / kWave;
    if (a >= nat) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t jend = jstart[a + 1], iend = istart[a + 1];
    int64_t k = jstart[a] + lane, q = istart[a] + lane;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    while (k < jend || q < iend) {
        const bool jA = k < jend, jB = k + kWave < jend;
        const bool iA = q < iend, iB = q + kWave < iend;
        const int64_t pA = jA ? static_cast<int64_t>(jperm[k]) : 0;
        const int64_t pB = jB ? static_cast<int64_t>(jperm[k + kWave]) : 0;
        double qa0 = 0.0, qa1 = 0.0, qa2 = 0.0, qb0 = 0.0, qb1 = 0.0, qb2 = 0.0;
        if (iA) {
            qa0 = G[3 * q + 0];
            qa1 = G[3 * q + 1];
            qa2 = G[3 * q + 2];
        }
        if (iB) {
            qb0 = G[3 * (q + kWave) + 0];
            qb1 = G[3 * (q + kWave) + 1];
            qb2 = G[3 * (q + kWave) + 2];
        }
        double ga0 = 0.0, ga1 = 0.0, ga2 = 0.0, gb0 = 0.0, gb1 = 0.0, gb2 = 0.0;
        if (jA) {
            ga0 = G[3 * pA + 0];
            ga1 = G[3 * pA + 1];
            ga2 = G[3 * pA + 2];
        }
        if (jB) {
            gb0 = G[3 * pB + 0];
            gb1 = G[3 * pB + 1];
            gb2 = G[3 * pB + 2];
        }
        if (jA) {
            s0 += ga0;
            s1 += ga1;
            s2 += ga2;
        }
        if (jB) {
            s0 += gb0;
            s1 += gb1;
            s2 += gb2;
        }
        if (iA) {
            r0 += qa0;
            r1 += qa1;
            r2 += qa2;
        }
        if (iB) {
            r0 += qb0;
            r1 += qb1;
            r2 += qb2;
        }
        k += 2 * kWave;
        q += 2 * kWave;
    }
    s0 = lane_tree(s0);
    s1 = lane_tree(s1);
    s2 = lane_tree(s2);
    r0 = lane_tree(r0);
    r1 = lane_tree(r1);
    r2 = lane_tree(r2);
    if (lane == 0) {
        grad_pos[3 * a + 0] = s0 - r0;
        grad_pos[3 * a + 1] = s1 - r1;
        grad_pos[3 * a + 2] = s2 - r2;
    }
}

/* grad_cell: thread t sums the block partials t, t + 256, ... in order,
   then the halving tree over the 256 threads. One block. */
__global__ __launch_bounds__(kBlock) void k_det_cell_sum(const double *__restrict__ partial, int nb,
                                                         double *__restrict__ grad_cell) {
    __shared__ double s_c[9][kBlock];
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock)
#pragma unroll
        for (int k = 0; k < 9; k++) c[k] += partial[static_cast<int64_t>(b) * 9 + k];
#pragma unroll
    for (int k = 0; k < 9; k++) s_c[k][threadIdx.x] = c[k];
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h)
            for (int k = 0; k < 9; k++) s_c[k][threadIdx.x] += s_c[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 9) grad_cell[threadIdx.x] = s_c[threadIdx.x][0];
}

__global__ void k_zero_f64(double *p, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

/* Bits of the largest key (keys run 0..nat). */
int key_bits(int32_t nat) {
    int bits = 1;
    while (bits < 31 && (int64_t(1) << bits) <= nat) bits++;
    return bits;
}

int64_t radix_blocks(int64_t npairs) { return std::max<int64_t>(1, (npairs + kTile - 1) / kTile); }

template <typename T>
hipError_t det_impl(int64_t npairs, int32_t nat, const int32_t *first, const int32_t *secnd,
                    const int32_t *shift, const void *D, const void *d, const void *gD,
                    const void *gd, int pbc_mask, double *grad_pos, double *grad_cell,
                    void *scratch, hipStream_t stream) {
    const T *Dt = static_cast<const T *>(D), *dt = static_cast<const T *>(d);
    const T *gDt = static_cast<const T *>(gD), *gdt = static_cast<const T *>(gd);
    const int64_t nb = radix_blocks(npairs);
    char *s = static_cast<char *>(scratch);
    auto take = [&](size_t bytes) {
        char *q = s;
        s += (bytes + 255) / 256 * 256;
        return q;
    };
    int32_t *k0 = reinterpret_cast<int32_t *>(take(4 * size_t(npairs)));
    int32_t *v0 = reinterpret_cast<int32_t *>(take(4 * size_t(npairs)));
    int32_t *k1 = reinterpret_cast<int32_t *>(take(4 * size_t(npairs)));
    int32_t *v1 = reinterpret_cast<int32_t *>(take(4 * size_t(npairs)));
    int32_t *hist = reinterpret_cast<int32_t *>(take(4 * size_t(kMaxBins * nb)));
    int64_t *offs = reinterpret_cast<int64_t *>(take(8 * size_t(kMaxBins * nb + 1)));
    int64_t *part = reinterpret_cast<int64_t *>(take(8 * size_t(scan_partial_size(kMaxBins * nb))));
    int64_t *jstart = reinterpret_cast<int64_t *>(take(8 * (size_t(nat) + 1)));
    int64_t *istart = reinterpret_cast<int64_t *>(take(8 * (size_t(nat) + 1)));
    double *cpart = reinterpret_cast<double *>(take(8 * 9 * size_t(kCellBlocks)));
    double *G = reinterpret_cast<double *>(take(24 * size_t(npairs)));

    if (npairs == 0) {
        if (nat > 0) k_zero_f64<<<blocks_for(3 * int64_t(nat)), kBlock, 0, stream>>>(
            grad_pos, 3 * int64_t(nat));
        k_zero_f64<<<1, kWave, 0, stream>>>(grad_cell, 9);
        return hipGetLastError();
    }
    /* LSD radix passes; the first reads the keys from secnd and takes the
       pair indices as values */
    const int bits = key_bits(nat);
    const int passes = (bits + 9) / 10;
    const int dbits = (bits + passes - 1) / passes;
    const int bins = 1 << dbits;
    const int32_t *kin = nullptr, *vin = nullptr;
    int32_t *kout = k0, *vout = v0;
    for (int ps = 0; ps < passes; ps++) {
        const int sh = dbits * ps;
        k_radix_hist<<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(
            kin, secnd, nat, npairs, sh, bins, hist, static_cast<int>(nb));
        hipError_t e = launch_scan_i32_to_i64(hist, bins * nb, offs, part, nullptr, stream);
        if (e != hipSuccess) return e;
        k_radix_scatter<<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(
            kin, secnd, nat, vin, npairs, sh, dbits, offs, static_cast<int>(nb), kout, vout);
        kin = kout;
        vin = vout;
        kout = kout == k0 ? k1 : k0;
        vout = vout == v0 ? v1 : v0;
    }
    /* each atom's j-side range in the sorted keys, and its row in first */
    k_lower_bounds<<<blocks_for(int64_t(nat) + 1), kBlock, 0, stream>>>(kin, npairs, nat, jstart);
    k_lower_bounds<<<blocks_for(int64_t(nat) + 1), kBlock, 0, stream>>>(first, npairs, nat,
                                                                        istart);
    const int cb = static_cast<int>(std::min<int64_t>(kCellBlocks, (npairs + kBlock - 1) / kBlock));
    k_det_pair_g<T><<<cb, kBlock, 0, stream>>>(npairs, shift, Dt, dt, gDt, gdt, pbc_mask, G,
                                               cpart);
    if (nat > 0)
        k_det_positions<<<static_cast<unsigned>((int64_t(nat) + kWaves - 1) / kWaves), kBlock, 0,
                          stream>>>(nat, jstart, vin, istart, G, grad_pos);
    k_det_cell_sum<<<1, kBlock, 0, stream>>>(cpart, cb, grad_cell);
    return hipGetLastError();
}

}  // namespace

size_t backward_det_scratch_bytes(int64_t npairs, int32_t nat) {
    const int64_t nb = radix_blocks(npairs);
    auto r = [](size_t b) { return (b + 255) / 256 * 256; };
    return 4 * r(4 * size_t(npairs)) + r(4 * size_t(kMaxBins * nb)) +
           r(8 * size_t(kMaxBins * nb + 1)) + r(8 * size_t(scan_partial_size(kMaxBins * nb))) +
           2 * r(8 * (size_t(nat) + 1)) +
           r(8 * 9 * size_t(kCellBlocks)) + r(24 * size_t(npairs));
}

hipError_t launch_pair_backward_det(int64_t npairs, int32_t nat, const int32_t *first,
                                    const int32_t *secnd, const int32_t *shift, const void *D,
                                    const void *d, const void *gD, const void *gd, bool f32,
                                    int pbc_mask, double *grad_pos, double *grad_cell,
                                    void *scratch, hipStream_t stream) {
    if (f32)
        return det_impl<float>(npairs, nat, first, secnd, shift, D, d, gD, gd, pbc_mask, grad_pos,
                               grad_cell, scratch, stream);
    return det_impl<double>(npairs, nat, first, secnd, shift, D, d, gD, gd, pbc_mask, grad_pos,
                            grad_cell, scratch, stream);
}

}  // namespace mnl

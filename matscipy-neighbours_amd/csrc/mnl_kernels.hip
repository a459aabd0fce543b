/*
 * mnl HIP kernels for gfx950 (CDNA4, wave64).
 *
 * Pipeline of one build (plain path = a batch of one system):
 *   bin         raw cell per atom, fold, per-cell histogram with an atomic
 *               ticket                                     (1 thread / atom)
 *   scan        cell counts -> cell slice starts  (reduce + apply; a spine
 *               pass beyond 1024 tiles; the apply alone for one tile)
 *   scatter     atoms into their cell slice at the ticket  (1 thread / atom)
 *   rank+gather stable order inside each cell (ascending atom index, the
 *               reference's serial counting sort) + gather of the
 *               cell-sorted per-atom records               (1 thread / atom)
 *   count       neighbours per atom                        (1 wave / cell)
 *   scan        counts -> int64 pair offsets, total on device
 *   fill        i, j, D, d, S at the offsets: the count's traversal again,
 *               accepted pairs compacted in LDS and written 64 at a time
 *                                                          (1 wave / cell)
 *
 * A small single system (at most 4096 atoms and 4096 cells) is binned by one
 * block instead (k_bin_small: parameters, histogram, scan, member lists and
 * stable ranks in LDS). Beside the list: the dense neighbour matrix (the
 * fill's traversal into fixed rows), the fused Lennard-Jones consumer (k_lj:
 * forces, energy and virial from the traversal, no list written) and the
 * backward of D and d (k_pair_backward).
 *
 * Floating point: compiled with contraction off and the reference's
 * operation order everywhere a double is produced, so the cell index, the
 * bin-relative positions, D and d are bit-identical to the reference CPU
 * path (src/libneighbours/neighbour_visit.hh:63-115,
 * neighbour_list.cc:126-167, tools.hh:30-67).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "mnl_engine.hh"

#pragma clang fp contract(off)

namespace mnl {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
/* Scan tiles of 2048 values: against 4096, the steps of 100k-200k atoms
   took 9-10 us less (twice the blocks, each with half the serial work before
   its look-back), fcc 1M 1.5 us less; tiles of 1024 cost fcc 1M 9 us. */
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

inline unsigned grid_for(int64_t n) {
    return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

/* tools.hh:39-43 (result of the wrap loops: the representative in [0, n)). */
__device__ __forceinline__ int wrap_idx(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

/* tools.hh:46-52 */
__device__ __forceinline__ int clamp_idx(int i, int n) {
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

/* floor(a / n) for n > 0. */
__device__ __forceinline__ int floor_div(int a, int n) {
    int q = a / n;
    return (a % n != 0 && a < 0) ? q - 1 : q;
}

/* x / d for 0 <= x < 2^31 with the multiplier and shift of make_div_magic
   (scalar instructions when x is wave-uniform). */
__device__ __forceinline__ int div_magic(int x, uint32_t mul, int shift) {
    return static_cast<int>((static_cast<uint64_t>(static_cast<uint32_t>(x)) * mul) >> shift);
}

/* Periodic image of cell index cj, in [0, n). near: cj is within one period
   of the range (the stencil reach is below n), one conditional step. */
__device__ __forceinline__ int wrap_cell(int cj, int n, bool near) {
    if (near) {
        cj += cj < 0 ? n : 0;
        return cj >= n ? cj - n : cj;
    }
    return cj - floor_div(cj, n) * n;
}

/* Component k of (a, b, c). */
__device__ __forceinline__ int pick3(int k, int a, int b, int c) {
    return k == 0 ? a : (k == 1 ? b : c);
}

/* Dense cell index (without the system's base) of wrapped cell coordinates
   in the system's layout lay = (f, m, s): c[f] + n[f] * (c[m] + n[m] * c[s]). */
__device__ __forceinline__ int dense_index(const SysGeom &g, int c0, int c1, int c2) {
    const int f = g.lay[0], m = g.lay[1], sl = g.lay[2];
    return pick3(f, c0, c1, c2) +
           pick3(f, g.n[0], g.n[1], g.n[2]) *
               (pick3(m, c0, c1, c2) + pick3(m, g.n[0], g.n[1], g.n[2]) * pick3(sl, c0, c1, c2));
}

/* Largest s with base[s] <= idx (base holds nsys+1 ascending offsets). */
__device__ __forceinline__ int system_of(const int32_t *__restrict__ base, int nsys,
                                         int idx) {
    if (nsys <= 1) return 0;
    int lo = 0, hi = nsys;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= idx)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* Fibonacci hash of a 64-bit cell key (cell_list.hh:72-75). */
__device__ __forceinline__ int64_t cell_hash(int64_t key) {
    const uint64_t h = static_cast<uint64_t>(key) * 0x9E3779B97F4A7C15ull;
    return static_cast<int64_t>(h >> 1);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const T t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

/* Lanes below `lane` whose bit is set in mask. */
__device__ __forceinline__ int mbcnt64(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

/* XCD-aware remap: the hardware deals workgroups round-robin to the 8 XCDs.
   Hand each XCD runs of kXcdRun consecutive logical blocks (128 cells), runs
   dealt round-robin: neighbouring cells share an L2, and work that is
   concentrated in one part of the grid (a slab under vacuum) still spreads
   over all XCDs. A bijection for any grid size (the tail that does not fill
   a round of runs keeps the hardware order). */
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    constexpr int kXcd = 8;
    constexpr int kXcdRun = 32;
    const int full = nblocks / (kXcd * kXcdRun) * (kXcd * kXcdRun);
    if (b >= full) return b;
    const int x = b % kXcd, k = b / kXcd;
    return ((k / kXcdRun) * kXcd + x) * kXcdRun + (k % kXcdRun);
}

/* ---------------------------------------------------------------- fills */

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

/* ---------------------------------------------------------------- binning */

/* Raw cell of one atom: position_to_cell_index (tools.hh:30-36,55-67), the
   fractional coordinate accumulated as ((0 + m0 v0) + m1 v1) + m2 v2. */
__device__ __forceinline__ void raw_cell_of(const double *__restrict__ r, int a,
                                            const SysGeom &g, int raw[3]) {
    const double dr0 = r[3 * (int64_t)a + 0] - g.origin[0];
    const double dr1 = r[3 * (int64_t)a + 1] - g.origin[1];
    const double dr2 = r[3 * (int64_t)a + 2] - g.origin[2];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double s = 0.0;
        s += g.inv[3 * k + 0] * dr0;
        s += g.inv[3 * k + 1] * dr1;
        s += g.inv[3 * k + 2] * dr2;
        raw[k] = static_cast<int>(floor(s * g.n[k]));
    }
}

/* Histogram ticket with one atomic per distinct counter in the wave: the
   lanes sharing a counter are grouped (one ballot per distinct value), the
   group's lowest lane adds the group size, and each lane's ticket is that
   base plus its position in the group. Called by every lane of the wave;
   lanes with valid == false get no ticket and add nothing. */
__device__ __forceinline__ int wave_ticket(int32_t *__restrict__ counters, int key, bool valid) {
    const int lane = lane_id();
    uint64_t todo = __builtin_amdgcn_ballot_w64(valid);
    uint64_t grp = 0;
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const int lk = __builtin_amdgcn_readlane(key, leader);
        const uint64_t same = __builtin_amdgcn_ballot_w64(valid && key == lk) & todo;
        if ((same >> lane) & 1) grp = same;
        todo &= ~same;
    }
    const int first = valid ? __builtin_ctzll(grp) : lane;
    int base = 0;
    if (valid && lane == first) base = atomicAdd(&counters[key], __popcll(grp));
    base = __shfl(base, first, kWave);
    return base + __popcll(grp & ((1ull << lane) - 1));
}

__global__ __launch_bounds__(kBlock) void k_bin_dense(
    const double *__restrict__ r, int32_t nat, const SysGeom *__restrict__ geo,
    const int32_t *__restrict__ atom_base, int32_t nsys, int4 *__restrict__ raw4,
    int32_t *__restrict__ ticket, int32_t *__restrict__ cell_count,
    int32_t *__restrict__ order_stats) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = a < nat;
    int raw[3] = {0, 0, 0};
    int c[3] = {0, 0, 0};
    int cell = -1;
    if (valid) {
        const SysGeom &g = geo[system_of(atom_base, nsys, a)];
        raw_cell_of(r, a, g, raw);
#pragma unroll
        for (int k = 0; k < 3; k++)
            c[k] = g.pbc[k] ? wrap_idx(raw[k], g.n[k]) : clamp_idx(raw[k], g.n[k]);
        cell = g.cell_base + dense_index(g, c[0], c[1], c[2]);
    }
    if (order_stats && blockIdx.x % 16 == 0 && threadIdx.x < kWave) {
        /* A sample (the first wave of every 16th block; all waves adding to
           the same three counters would serialise): per axis, the periodic
           cell distance between atoms a and a + 32. Over 32 atoms a lattice
           in index order moves along its fastest axis only (32 is a multiple
           of the usual basis sizes), a random order along all three. */
        const SysGeom &g0 = geo[0];
        const bool pair = lane_id() < 32 && a + 32 < nat;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int cn = __shfl_down(c[k], 32, kWave);
            int dd = cn - c[k];
            if (2 * dd > g0.n[k]) dd -= g0.n[k];
            if (2 * dd < -g0.n[k]) dd += g0.n[k];
            /* capped at 64 cells: the sums stay far below 2^31 */
            const int moved = wave_sum(pair ? min(dd < 0 ? -dd : dd, 64) : 0);
            if (lane_id() == 0 && moved) atomicAdd(&order_stats[k], moved);
        }
    }
    const int tk = wave_ticket(cell_count, cell, valid);
    if (!valid) return;
    ticket[a] = tk;
    raw4[a] = make_int4(raw[0], raw[1], raw[2], cell);
}

/* Hashed compact grid (single system): open addressing keyed by the 64-bit
   linear cell index, capacity a power of two >= 2N+1 (cell_list.cc:105-149).
   Slot placement depends on arrival order; the slices it yields do not. */
__global__ __launch_bounds__(kBlock) void k_bin_hashed(
    const double *__restrict__ r, int32_t nat, const SysGeom *__restrict__ geo,
    int4 *__restrict__ raw4, int32_t *__restrict__ ticket, int64_t *__restrict__ hkey,
    int64_t hmask, int32_t *__restrict__ hcount) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = a < nat;
    int raw[3] = {0, 0, 0};
    int64_t h = -1;
    if (valid) {
        const SysGeom &g = geo[0];
        raw_cell_of(r, a, g, raw);
        int c[3];
#pragma unroll
        for (int k = 0; k < 3; k++)
            c[k] = g.pbc[k] ? wrap_idx(raw[k], g.n[k]) : clamp_idx(raw[k], g.n[k]);
        const int64_t key =
            static_cast<int64_t>(c[0]) +
            static_cast<int64_t>(g.n[0]) *
                (static_cast<int64_t>(c[1]) + static_cast<int64_t>(g.n[1]) * c[2]);
        h = cell_hash(key) & hmask;
        auto *slots = reinterpret_cast<unsigned long long *>(hkey);
        const unsigned long long empty = ~0ull;
        /* ends: at most nat keys in >= 2 nat + 1 slots */
        while (true) {
            const unsigned long long old =
                atomicCAS(&slots[h], empty, static_cast<unsigned long long>(key));
            if (old == empty || static_cast<int64_t>(old) == key) break;
            h = (h + 1) & hmask;
        }
    }
    const int tk = wave_ticket(hcount, static_cast<int>(h), valid);
    if (!valid) return;
    ticket[a] = tk;
    raw4[a] = make_int4(raw[0], raw[1], raw[2], static_cast<int>(h));
}

__global__ __launch_bounds__(kBlock) void k_scatter(const int4 *__restrict__ raw4,
                                                    const int32_t *__restrict__ ticket,
                                                    int32_t nat,
                                                    const int32_t *__restrict__ cell_first,
                                                    int32_t *__restrict__ tmp) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nat) return;
    tmp[cell_first[raw4[a].w] + ticket[a]] = a;
}

/* Stable placement + gather. The rank of atom a inside its cell is the number
   of cell members with a smaller index, which puts every cell slice in
   ascending atom order exactly like the reference's serial scatter
   (cell_list.cc:77-80). The cell-sorted record is then written as in
   neighbour_list.cc:147-166: rel = raw if periodic else clamped, rel_pos =
   ((r - rel1 b1) - rel2 b2) - rel3 b3. For the shift the record keeps, per
   direction, the raw cell (periodic) or the image count
   trunc((raw - clamp(raw)) / n) (non-periodic). */
__device__ __forceinline__ void place_atom(const double *__restrict__ r, int a, int s,
                                           const int raw[3], const SysGeom &g,
                                           const double *__restrict__ per_atom,
                                           const int32_t *__restrict__ types,
                                           double4 *__restrict__ spos, int4 *__restrict__ sinfo,
                                           int32_t *__restrict__ stype,
                                           int32_t *__restrict__ scount,
                                           int32_t *__restrict__ slot) {
    int rel[3], base[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (g.pbc[k]) {
            rel[k] = raw[k];
            base[k] = raw[k];
        } else {
            rel[k] = clamp_idx(raw[k], g.n[k]);
            base[k] = (raw[k] - rel[k]) / g.n[k];
        }
    }
    double p[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double v = r[3 * (int64_t)a + k];
        v = v - rel[0] * g.bin[0 + k];
        v = v - rel[1] * g.bin[3 + k];
        v = v - rel[2] * g.bin[6 + k];
        p[k] = v;
    }
    spos[s] = make_double4(p[0], p[1], p[2], per_atom ? per_atom[a] : 0.0);
    sinfo[s] = make_int4(base[0], base[1], base[2], a - g.atom_base);
    if (stype) stype[s] = types ? types[a] : 0;
    if (scount) scount[s] = 0; /* the half count accumulates here */
    if (slot) slot[a] = s;     /* atom -> cell-sorted index */
}

__global__ __launch_bounds__(kBlock) void k_rank_gather(
    const double *__restrict__ r, const int4 *__restrict__ raw4,
    const int32_t *__restrict__ tmp, int32_t nat, const int32_t *__restrict__ cell_first,
    const int32_t *__restrict__ cell_count, const SysGeom *__restrict__ geo,
    const int32_t *__restrict__ atom_base, int32_t nsys,
    const double *__restrict__ per_atom, const int32_t *__restrict__ types,
    double4 *__restrict__ spos, int4 *__restrict__ sinfo, int32_t *__restrict__ stype,
    int32_t *__restrict__ scount, int32_t *__restrict__ slot, int32_t *__restrict__ first_flag) {
    /* one thread per slot of the ticket-ordered member lists: the record of
       the member held there goes to a slot of the same cell, so a wave
       writes one contiguous range of records for any input order (a thread
       per atom scatters them when the input order is spatially random) */
    const int s0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (s0 >= nat) return;
    const int a = tmp[s0];
    const int4 rc = raw4[a];
    const int b = cell_first[rc.w];
    const int m = cell_count[rc.w];
    /* members read eight at a time: independent loads, one round trip */
    int rank = 0;
    for (int t0 = 0; t0 < m; t0 += 8) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = t0 + u < m ? tmp[b + t0 + u] : INT32_MAX;
#pragma unroll
        for (int u = 0; u < 8; u++) rank += v[u] < a ? 1 : 0;
    }
    const int s = b + rank;
    if (first_flag) first_flag[a] = rank == 0 ? 1 : 0;
    const int raw[3] = {rc.x, rc.y, rc.z};
    place_atom(r, a, s, raw, geo[system_of(atom_base, nsys, a)], per_atom, types, spos, sinfo,
               stype, scount, slot);
}

/* The occupied cells (hash slots) in the order of their lowest atom. */
__global__ __launch_bounds__(kBlock) void k_home_list(const int4 *__restrict__ raw4,
                                                      const int32_t *__restrict__ flag,
                                                      const int32_t *__restrict__ pos,
                                                      int32_t nat, int32_t *__restrict__ home_list) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nat) return;
    if (flag[a]) home_list[pos[a]] = raw4[a].w;
}

/* ---------------------------------------------------------------- scan */

template <typename TI>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const TI *__restrict__ in,
                                                         int64_t n,
                                                         int64_t *__restrict__ partial) {
    __shared__ int64_t s_w[kWavesPerBlock];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    int64_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        if (idx < n) v += in[idx];
    }
    v = wave_sum(v);
    if (lane_id() == 0) s_w[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < kWavesPerBlock; w++) t += s_w[w];
        partial[blockIdx.x] = t;
    }
}

/* k_scan_reduce over counts held in cell-sorted order: the value of atom a
   is scount[slot[a]], also written to counts[a] (the half count's per-atom
   counts in atom order, read by the apply phase). */
__global__ __launch_bounds__(kBlock) void k_scan_reduce_gather(const int32_t *__restrict__ scount,
                                                                const int32_t *__restrict__ slot,
                                                                int32_t *__restrict__ counts,
                                                                int64_t n,
                                                                int64_t *__restrict__ partial) {
    __shared__ int64_t s_w[kWavesPerBlock];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    int64_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        if (idx < n) {
            const int32_t c = scount[slot[idx]];
            counts[idx] = c;
            v += c;
        }
    }
    v = wave_sum(v);
    if (lane_id() == 0) s_w[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < kWavesPerBlock; w++) t += s_w[w];
        partial[blockIdx.x] = t;
    }
}

/* Single block: exclusive scan of the nb block sums in place; total out. */
__global__ __launch_bounds__(1024) void k_scan_spine(int64_t *__restrict__ partial, int64_t nb,
                                                     int64_t *__restrict__ total,
                                                     int64_t *__restrict__ total2) {
    __shared__ int64_t s_w[1024 / kWave];
    __shared__ int64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t idx = base + threadIdx.x;
        const int64_t v = idx < nb ? partial[idx] : 0;
        const int64_t incl = wave_incl_scan(v);
        if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = incl;
        __syncthreads();
        int64_t pre = s_carry;
        for (int w = 0; w < static_cast<int>(threadIdx.x / kWave); w++) pre += s_w[w];
        if (idx < nb) partial[idx] = pre + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = pre + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = s_carry;
    if (threadIdx.x == 0 && total2) *total2 = s_carry;
}

/* Scan of the tiles with their offsets: with `fused`, partial holds the
   tiles' sums and each block adds up those before it (few tiles: replaces
   the spine kernel; the last block writes the totals); otherwise partial
   holds the exclusive offsets the spine wrote. A single tile needs no
   partial sums at all. With gslot the input is gathered as
   gsrc[gslot[idx]] and also stored to gout[idx] (the first phase's gather,
   for a scan that runs without it). */
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const TI *__restrict__ in, int64_t n,
                                                        const int64_t *__restrict__ partial,
                                                        TO *__restrict__ out, int write_total,
                                                        int fused, int64_t *__restrict__ total,
                                                        int64_t *__restrict__ total2,
                                                        const int32_t *__restrict__ gsrc,
                                                        const int32_t *__restrict__ gslot,
                                                        int32_t *__restrict__ gout) {
    __shared__ int64_t s_tile[kScanTile];
    __shared__ int64_t s_w[kWavesPerBlock];
    __shared__ int64_t s_pre;
    __shared__ int64_t s_last;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    if (fused) {
        int64_t v = 0;
        for (int b = threadIdx.x; b < static_cast<int>(blockIdx.x); b += kBlock) v += partial[b];
        v = wave_sum(v);
        if (lane_id() == 0) s_w[threadIdx.x / kWave] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t t = 0;
            for (int w = 0; w < kWavesPerBlock; w++) t += s_w[w];
            s_pre = t;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        int64_t v = 0;
        if (idx < n) {
            if (gslot) {
                const int32_t c = gsrc[gslot[idx]];
                gout[idx] = c;
                v = c;
            } else {
                v = static_cast<int64_t>(in[idx]);
            }
            if (idx == n - 1) s_last = v;
        }
        s_tile[k * kBlock + threadIdx.x] = v;
    }
    __syncthreads();
    int64_t local[kScanItems];
    int64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        local[k] = sum;
        sum += s_tile[threadIdx.x * kScanItems + k];
    }
    const int64_t incl = wave_incl_scan(sum);
    if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = incl;
    __syncthreads();
    int64_t pre = fused ? s_pre : partial[blockIdx.x];
    for (int w = 0; w < static_cast<int>(threadIdx.x / kWave); w++) pre += s_w[w];
    pre += incl - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) s_tile[threadIdx.x * kScanItems + k] = pre + local[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        if (idx < n) out[idx] = static_cast<TO>(s_tile[k * kBlock + threadIdx.x]);
    }
    if (write_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) {
        /* the last element's exclusive prefix plus itself */
        const int64_t tot = s_tile[n - 1 - base] + s_last;
        out[n] = static_cast<TO>(tot);
    }
    if (fused && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1 && (total || total2)) {
        const int64_t tot = s_tile[n - 1 - base] + s_last;
        if (total) *total = tot;
        if (total2) *total2 = tot;
    }
}

/* Decoupled look-back (one wave): publishes tile t's sum, adds up the sums
   of the tiles before it until one with an inclusive prefix, publishes its
   own inclusive prefix and returns the exclusive one. Status word: bits
   62-63 = 0 not yet, 1 sum only, 2 inclusive prefix; bits 0-61 the value.
   Every tile waited on took its ticket earlier, so it is resident or done
   and publishes its sum before it waits on anything. */
__device__ __forceinline__ int64_t tile_lookback(uint64_t *status, int t, int64_t sum,
                                                 int32_t *flags) {
    constexpr uint64_t kAgg = uint64_t(1) << 62, kInc = uint64_t(2) << 62;
    constexpr uint64_t kVal = kAgg - 1;
    constexpr int kSpinLimit = 1 << 22;
    const int lane = lane_id();
    if (t == 0) {
        if (lane == 0)
            __hip_atomic_store(&status[0], kInc | static_cast<uint64_t>(sum), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if (lane == 0)
        __hip_atomic_store(&status[t], kAgg | static_cast<uint64_t>(sum), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    int64_t excl = 0;
    int j = t - 1;
    int spins = 0;
    while (true) {
        const int idx = j - lane;
        /* before the first tile: an inclusive zero */
        const uint64_t v = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                                    : kInc;
        const uint64_t flag = v >> 62;
        const uint64_t inc = __builtin_amdgcn_ballot_w64(flag == 2);
        const uint64_t zero = __builtin_amdgcn_ballot_w64(flag == 0);
        const int first = inc ? static_cast<int>(__builtin_ctzll(inc)) : kWave;
        const uint64_t need = first >= kWave - 1 ? ~uint64_t(0) : ((uint64_t(2) << first) - 1);
        if (zero & need) {
            if (++spins > kSpinLimit) {
                /* never expected: report instead of waiting for ever. The
                   flag stays set (the host checks it where it reads a pair
                   total), and the prefix 0 keeps every later write in bounds. */
                if (lane == 0) atomicOr(flags, 1);
                return 0;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += wave_sum(lane <= first ? static_cast<int64_t>(v & kVal) : int64_t(0));
        if (first < kWave) break;
        j -= kWave;
    }
    if (lane == 0)
        __hip_atomic_store(&status[t], kInc | static_cast<uint64_t>(excl + sum), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

/*
 * Single-pass exclusive scan of nb (2 .. kScanOnepassTiles) tiles: each block
 * takes a tile by ticket, scans it as k_scan_apply does, and finds the sum
 * of the tiles before it by a decoupled look-back, so one launch replaces the
 * reduce and apply kernels. status holds kScanOnepassTiles words and then the
 * ticket, the count of tiles done and an error flag (int32); all are zero
 * before the launch, and the last tile to finish its look-back (when no
 * block reads the words any more) zeroes them again for the next scan, so
 * the scan records no clearing and replays correctly in a graph.
 */
constexpr int kScanOnepassTiles = 1024;

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_scan_onepass(
    const TI *__restrict__ in, int64_t n, TO *__restrict__ out, int write_total,
    int64_t *__restrict__ total, int64_t *__restrict__ total2, const int32_t *__restrict__ gsrc,
    const int32_t *__restrict__ gslot, int32_t *__restrict__ gout, uint64_t *__restrict__ status,
    int nb) {
    __shared__ int64_t s_tile[kScanTile];
    __shared__ int64_t s_w[kWavesPerBlock];
    __shared__ int64_t s_pre;
    __shared__ int64_t s_last;
    __shared__ int s_t;
    int32_t *counters = reinterpret_cast<int32_t *>(status + kScanOnepassTiles);
    if (threadIdx.x == 0) s_t = atomicAdd(&counters[0], 1);
    __syncthreads();
    const int t = s_t;
    const int64_t base = static_cast<int64_t>(t) * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        int64_t v = 0;
        if (idx < n) {
            if (gslot) {
                const int32_t c = gsrc[gslot[idx]];
                gout[idx] = c;
                v = c;
            } else {
                v = static_cast<int64_t>(in[idx]);
            }
            if (idx == n - 1) s_last = v;
        }
        s_tile[k * kBlock + threadIdx.x] = v;
    }
    __syncthreads();
    int64_t local[kScanItems];
    int64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        local[k] = sum;
        sum += s_tile[threadIdx.x * kScanItems + k];
    }
    const int64_t incl = wave_incl_scan(sum);
    if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = incl;
    __syncthreads();
    if (threadIdx.x < kWave) {
        int64_t bsum = 0;
        for (int w = 0; w < kWavesPerBlock; w++) bsum += s_w[w];
        const int64_t pre = tile_lookback(status, t, bsum, counters + 2);
        if (threadIdx.x == 0) s_pre = pre;
        /* the last tile through its look-back clears the status words */
        int done = 0;
        if (threadIdx.x == 0) done = atomicAdd(&counters[1], 1);
        done = __shfl(done, 0, kWave);
        if (done == nb - 1) {
            for (int k = threadIdx.x; k < nb; k += kWave) status[k] = 0;
            if (threadIdx.x == 0) {
                counters[0] = 0;
                counters[1] = 0;
            }
        }
    }
    __syncthreads();
    int64_t pre = s_pre;
    for (int w = 0; w < static_cast<int>(threadIdx.x / kWave); w++) pre += s_w[w];
    pre += incl - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) s_tile[threadIdx.x * kScanItems + k] = pre + local[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        if (idx < n) out[idx] = static_cast<TO>(s_tile[k * kBlock + threadIdx.x]);
    }
    if (t == nb - 1 && threadIdx.x == kBlock - 1) {
        /* the last element's exclusive prefix plus itself */
        const int64_t tot = s_tile[n - 1 - base] + s_last;
        if (write_total) out[n] = static_cast<TO>(tot);
        if (total) *total = tot;
        if (total2) *total2 = tot;
    }
}

/* ---------------------------------------------------------------- pairs */

enum { MODE_COUNT = 0, MODE_FILL64 = 1, MODE_FILL32 = 2, MODE_MATRIX = 3 };

/* Dense slice of cell (c0,c1,c2) of a system (cell_list.hh:94-104). */
__device__ __forceinline__ void dense_slice(const PairArgs &A, int cell_base, int lf, int lm,
                                            int ls, int nf, int nm, int c0, int c1, int c2,
                                            int &b, int &len) {
    const int c = cell_base + pick3(lf, c0, c1, c2) +
                  nf * (pick3(lm, c0, c1, c2) + nm * pick3(ls, c0, c1, c2));
    b = A.cell_first[c];
    len = A.cell_count[c];
}

/* Hashed slice (cell_list.hh:107-131): probe until the key or an empty slot. */
__device__ __forceinline__ void hashed_slice(const PairArgs &A, int n0, int n1, int c0,
                                             int c1, int c2, int &b, int &len) {
    const int64_t key = static_cast<int64_t>(c0) +
                        static_cast<int64_t>(n0) *
                            (static_cast<int64_t>(c1) + static_cast<int64_t>(n1) * c2);
    int64_t h = cell_hash(key) & A.hmask;
    b = 0;
    len = 0;
    while (true) {
        const int64_t k = A.hkey[h];
        if (k == -1) return;
        if (k == key) {
            b = A.cell_first[h];
            len = A.cell_count[h];
            return;
        }
        h = (h + 1) & A.hmask;
    }
}

/* Read-only views in the constant address space: a load through them at a
   wave-uniform address becomes a scalar (SMEM) load into SGPRs. The pair
   kernels never write these arrays. */
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)(p);
}

/* A pointer into LDS, so that reads through it are ds_read (not flat). */
using lds_dptr = const __attribute__((address_space(3))) double *;

/* Kernel-side cut mode: per-type cutoffs with the table staged in LDS. */
constexpr int CUT_PER_TYPE_LDS = 3;

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

/* Element i of a read-only record array through the constant address space
   (component-wise: the vector types have no address-space-qualified copy). */
__device__ __forceinline__ double4 load_const(const double4 *p, int i) {
    cptr<double> q = (cptr<double>)(p) + 4 * static_cast<int64_t>(i);
    return make_double4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ int4 load_const(const int4 *p, int i) {
    cptr<int> q = (cptr<int>)(p) + 4 * static_cast<int64_t>(i);
    return make_int4(q[0], q[1], q[2], q[3]);
}

}  // namespace
/* The v_writelane intrinsic (no clang builtin spells it): returns `old` with
   lane `lane` replaced by the wave-uniform `val`. */
extern "C" __device__ int mnl_llvm_writelane(int val, int lane, int old) __asm(
    "llvm.amdgcn.writelane.i32");
namespace {

__device__ __forceinline__ uint32_t write_lane(uint32_t v, uint32_t s, int lane) {
    return static_cast<uint32_t>(
        mnl_llvm_writelane(static_cast<int>(s), lane, static_cast<int>(v)));
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Home cell decoded once per wave (all values wave-uniform). */
struct HomeCell {
    int m, hb;             /* atoms in the cell, first cell-sorted atom */
    int c0, c1, c2;        /* cell coordinates */
    int n0, n1, n2;
    int nx, ny, nz, sx, sy, nseg, centre;
    uint32_t msx, msy;     /* division by sx, sy (make_div_magic) */
    int hsx, hsy;
    int p0, p1, p2;
    int abase, cbase;
    int split;             /* waves sharing this cell's home atoms */
    int lf, lm, ls, nf, nm; /* dense layout: fastest, middle, slowest axis */
    double cutoff_sq;
    double b00, b01, b02, b10, b11, b12, b20, b21, b22;
    double rn0, rn1, rn2;
};

template <bool HASHED>
__device__ __forceinline__ bool load_home(const PairArgs &A, int h, HomeCell &H) {
    H.m = uniform(as_const(A.cell_count)[h]);
    if (H.m == 0) return false;
    H.hb = uniform(as_const(A.cell_first)[h]);
    int sys = 0;
    if (!HASHED) sys = system_of(A.cell_base, A.nsys, h);
    sys = uniform(sys);
    cptr<SysGeom> gp = as_const(A.geo) + sys;
    H.n0 = gp->n[0];
    H.n1 = gp->n[1];
    H.n2 = gp->n[2];
    H.lf = gp->lay[0];
    H.lm = gp->lay[1];
    H.ls = gp->lay[2];
    H.nf = pick3(H.lf, H.n0, H.n1, H.n2);
    H.nm = pick3(H.lm, H.n0, H.n1, H.n2);
    if (!HASHED) {
        /* lin = c_f + n_f (c_m + n_m c_s) */
        const int lin = h - gp->cell_base;
        const int q = div_magic(lin, gp->dmul[H.lf], gp->dshift[H.lf]);
        const int cf = lin - q * H.nf;
        const int cs = div_magic(q, gp->dmul[H.lm], gp->dshift[H.lm]);
        const int cm = q - cs * H.nm;
        H.c0 = H.lf == 0 ? cf : (H.lm == 0 ? cm : cs);
        H.c1 = H.lf == 1 ? cf : (H.lm == 1 ? cm : cs);
        H.c2 = H.lf == 2 ? cf : (H.lm == 2 ? cm : cs);
    } else {
        const int64_t key = as_const(A.hkey)[h];
        H.c0 = static_cast<int>(key % H.n0);
        H.c1 = static_cast<int>((key / H.n0) % H.n1);
        H.c2 = static_cast<int>(key / (static_cast<int64_t>(H.n0) * H.n1));
    }
    H.nx = gp->nxyz[0];
    H.ny = gp->nxyz[1];
    H.nz = gp->nxyz[2];
    H.sx = 2 * H.nx + 1;
    H.sy = 2 * H.ny + 1;
    H.msx = gp->dmul[3];
    H.hsx = gp->dshift[3];
    H.msy = gp->dmul[4];
    H.hsy = gp->dshift[4];
    H.nseg = gp->nseg;
    H.centre = (H.nz * H.sy + H.ny) * H.sx + H.nx; /* segment of offset (0,0,0) */
    H.p0 = gp->pbc[0];
    H.p1 = gp->pbc[1];
    H.p2 = gp->pbc[2];
    H.cutoff_sq = gp->cutoff_sq;
    H.b00 = gp->bin[0];
    H.b01 = gp->bin[1];
    H.b02 = gp->bin[2];
    H.b10 = gp->bin[3];
    H.b11 = gp->bin[4];
    H.b12 = gp->bin[5];
    H.b20 = gp->bin[6];
    H.b21 = gp->bin[7];
    H.b22 = gp->bin[8];
    H.rn0 = gp->rn[0];
    H.rn1 = gp->rn[1];
    H.rn2 = gp->rn[2];
    H.abase = gp->atom_base;
    H.cbase = gp->cell_base;
    H.split = A.item_base ? gp->split : A.split;
    return true;
}

/* Wave w -> home cell w / split and part w % split. A cell's home atoms
   are shared out over its split waves (each atom's row is independent), so
   that systems with few cells still give the chip thousands of waves. */
__device__ __forceinline__ void wave_home(const PairArgs &A, int w, int &h, int &part) {
    if (A.item_base) {
        /* a batch: each system's cells with the system's own split, so that
           the one-cell systems (all their atoms in one home cell, 27 images
           of it as candidates) do not leave a few waves running long after
           the rest */
        if (w >= A.items) {
            h = A.home_cells;
            part = 0;
            return;
        }
        const int s = uniform(system_of(A.item_base, A.nsys, w));
        cptr<SysGeom> gp = as_const(A.geo) + s;
        const int sp = uniform(gp->split);
        const int local = w - uniform(as_const(A.item_base)[s]);
        const int c = sp > 1 ? local / sp : local;
        h = uniform(gp->cell_base) + c;
        part = local - c * sp;
        return;
    }
    h = A.split > 1 ? w / A.split : w;
    part = w - h * A.split;
}

/* Work items of a hashed-grid pair-kernel launch: (cell, part) for every
   entry of the home list, split ways each, taken grid-stride by the
   launch's waves (the list's length is known on the device only). The
   dense grid runs one wave per (cell, part) with no loop: the loop costs
   the dense fill about 10%. */
__device__ __forceinline__ int64_t pair_items(const PairArgs &A) {
    const int64_t cells = A.home_list ? uniform(static_cast<int>(as_const(A.home_items)[0]))
                                      : static_cast<int64_t>(A.home_cells);
    return cells * A.split;
}

__device__ __forceinline__ int item_cell(const PairArgs &A, int idx) {
    return A.home_list ? uniform(as_const(A.home_list)[idx]) : idx;
}

/* Home atoms [lo, hi) of part `part` of a cell with m atoms (split parts,
   sizes differing by at most one). */
__device__ __forceinline__ void home_part(int m, int part, int split, int &lo, int &hi) {
    if (split == 1) {
        lo = 0;
        hi = m;
        return;
    }
    lo = static_cast<int>(static_cast<int64_t>(m) * part / split);
    hi = static_cast<int>(static_cast<int64_t>(m) * (part + 1) / split);
}

/* Segment table of one group of up to 64 neighbour-cell offsets, in the
   reference's traversal order. */
struct SegLds {
    double so[3][kWave]; /* offset vector off = (z b3 + y b2) + x b1 */
    int st[kWave];       /* first candidate index of each segment */
    int sb[kWave];       /* first cell-sorted atom of each segment */
};

/* Cell offset (x, y, z) of segment `seg` (x fastest, each -n..n). */
__device__ __forceinline__ void segment_offset(const HomeCell &H, int seg, int &x, int &y,
                                               int &z) {
    const int rest = div_magic(seg, H.msx, H.hsx);
    x = seg - rest * H.sx - H.nx;
    const int zq = div_magic(rest, H.msy, H.hsy);
    y = rest - zq * H.sy - H.ny;
    z = zq - H.nz;
}

/* Builds the table of segments sg .. sg+nsg-1; returns the group's candidate
   count, and in *self_start the candidate index of the home cell's first atom
   at zero offset (-1 if that segment is not in the group). */
template <bool HASHED>
__device__ __forceinline__ int build_segments(const PairArgs &A, const HomeCell &H, int sg,
                                              int nsg, int lane, SegLds &S, int *self_start) {
    int len = 0, begin = 0;
    if (lane < nsg) {
        int x, y, z;
        segment_offset(H, sg + lane, x, y, z);
        int cj0 = H.c0 + x, cj1 = H.c1 + y, cj2 = H.c2 + z;
        bool ok = true;
        if (H.p0)
            cj0 = wrap_cell(cj0, H.n0, H.nx < H.n0);
        else
            ok = ok && cj0 >= 0 && cj0 < H.n0;
        if (H.p1)
            cj1 = wrap_cell(cj1, H.n1, H.ny < H.n1);
        else
            ok = ok && cj1 >= 0 && cj1 < H.n1;
        if (H.p2)
            cj2 = wrap_cell(cj2, H.n2, H.nz < H.n2);
        else
            ok = ok && cj2 >= 0 && cj2 < H.n2;
        if (ok) {
            if (HASHED)
                hashed_slice(A, H.n0, H.n1, cj0, cj1, cj2, begin, len);
            else
                dense_slice(A, H.cbase, H.lf, H.lm, H.ls, H.nf, H.nm, cj0, cj1, cj2, begin,
                            len);
        }
        /* the offset vector, neighbour_visit.hh:67,73-74,80-81 */
        S.so[0][lane] = (z * H.b20 + y * H.b10) + x * H.b00;
        S.so[1][lane] = (z * H.b21 + y * H.b11) + x * H.b01;
        S.so[2][lane] = (z * H.b22 + y * H.b12) + x * H.b02;
    }
    const int incl = wave_incl_scan(len);
    const int total = uniform(__shfl(incl, kWave - 1, kWave));
    const int ci = H.centre - sg;
    *self_start = (ci >= 0 && ci < nsg) ? uniform(__shfl(incl - len, ci, kWave)) : -1;
    S.st[lane] = incl - len;
    S.sb[lane] = begin;
    wave_sync_lds();
    return total;
}

/* Largest index p < n with v[p] <= x, for non-decreasing v with
   v[0] <= x. Binary lifting with the probe clamped to n - 1, so every LDS
   read is unconditional (no divergent branches). */
__device__ __forceinline__ int last_at_or_below(const int *v, int n, int x) {
    int p = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const int q = min(p + step, n - 1);
        p = (v[q] <= x) ? q : p;
    }
    return p;
}

/* One lane's candidate of the current chunk, as the test needs it. */
struct Cand {
    double px, py, pz;     /* bin-relative position (px = NaN for an idle lane) */
    double pw;             /* radius (per-atom cutoffs) */
    double o0, o1, o2;     /* cell offset vector off = (z b3 + y b2) + x b1 */
    int tj;                /* type (per-type cutoffs) */
    int j;                 /* fill: atom index (system-local); count: cell-sorted index */
    int c0, c1, c2;        /* shift numerator part x - raw_j (fill only) */
};

/* Candidate c of the group of segments from sg on: its segment (the last
   one starting at or before c; empty segments never win), cell-sorted
   index, record and the segment's offset vector. */
template <int CUT, bool FILL>
__device__ __forceinline__ void load_candidate(const PairArgs &A, const HomeCell &H,
                                               const SegLds &S, int sg, int nsg, int c,
                                               int total, Cand &C) {
    const bool cv = c < total;
    const int p = last_at_or_below(S.st, nsg, c);
    const int sj = cv ? S.sb[p] + (c - S.st[p]) : H.hb;
    C.o0 = S.so[0][p];
    C.o1 = S.so[1][p];
    C.o2 = S.so[2][p];
    const double4 pj = A.spos[sj];
    C.px = cv ? pj.x : __builtin_nan("");
    C.py = pj.y;
    C.pz = pj.z;
    C.pw = pj.w;
    C.tj = (CUT == CUT_PER_TYPE || CUT == CUT_PER_TYPE_LDS) ? A.stype[sj] : 0;
    if (FILL) {
        int x, y, z;
        segment_offset(H, sg + p, x, y, z);
        const int4 bj = A.sinfo[sj];
        C.j = bj.w;
        C.c0 = x - bj.x;
        C.c1 = y - bj.y;
        C.c2 = z - bj.z;
    } else {
        C.j = sj;
    }
}

/* The distance test of neighbour_visit.hh:89-109 for a home atom at pi:
   dr = (rel_j - rel_i) + off, r2 = (dx^2 + dy^2) + dz^2, strict r2 < rc^2,
   then the per-atom (sum of radii) or per-type (matrix entry when both types
   are in range) cutoff. dr is returned for the fill. */
template <int CUT>
__device__ __forceinline__ bool pair_test(const PairArgs &A, const HomeCell &H, const Cand &C,
                                          int si, const double4 &pi, double &d0, double &d1,
                                          double &d2, lds_dptr pt_lds) {
    d0 = (C.px - pi.x) + C.o0;
    d1 = (C.py - pi.y) + C.o1;
    d2 = (C.pz - pi.z) + C.o2;
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    bool ok = r2 < H.cutoff_sq;
    if (CUT == CUT_PER_ATOM) {
        const double cc = pi.w + C.pw;
        ok = ok & (r2 < cc * cc);
    } else if (CUT == CUT_PER_TYPE || CUT == CUT_PER_TYPE_LDS) {
        const int ti = as_const(A.stype)[si];
        if (ti >= 0 && ti < A.ncut) {
            const bool tj_in = (C.tj >= 0) & (C.tj < A.ncut);
            const int e = ti * A.ncut + (tj_in ? C.tj : 0);
            const double lim = CUT == CUT_PER_TYPE_LDS ? pt_lds[e] : A.pt_sq[e];
            ok = ok & (!tj_in | (r2 < lim));
        }
    }
    return ok;
}

/*
 * Per-type cutoffs with the table in LDS, by home-atom class. The home atoms
 * of a group are taken class by class: their type, or one class for types
 * outside the table. For a class each candidate lane holds the limit
 * min(rc^2, table[type][tj]) in a register, and the test is then one
 * comparison, as for a uniform cutoff. This is pair_test's per-type rule:
 * (r2 < rc^2) & (r2 < t) is r2 < min(rc^2, t), and a pair with either type
 * outside the table keeps r2 < rc^2 alone (neighbour_visit.hh:100-108). The
 * order in which home atoms are taken does not change the output: each
 * atom's pairs keep their rank in its own row.
 */
__device__ __forceinline__ int home_class(const PairArgs &A, int ti) {
    return (ti >= 0 && ti < A.ncut) ? ti : A.ncut;
}

__device__ __forceinline__ double class_limit(const PairArgs &A, const HomeCell &H,
                                              lds_dptr pt, int v, int tj) {
    const bool in = v < A.ncut && tj >= 0 && tj < A.ncut;
    const double t = pt[in ? v * A.ncut + tj : 0];
    return in ? fmin(H.cutoff_sq, t) : H.cutoff_sq;
}

/* pair_test with the lane's limit for the home atom's class */
__device__ __forceinline__ bool pair_test_lim(const Cand &C, const double4 &pi, double lim,
                                              double &d0, double &d1, double &d2) {
    d0 = (C.px - pi.x) + C.o0;
    d1 = (C.py - pi.y) + C.o1;
    d2 = (C.pz - pi.z) + C.o2;
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    return r2 < lim;
}

/*
 * Count pass. One wave per home cell. The home cell's neighbour cells, in
 * the reference's traversal order (z outer, y, x inner, each -n..n;
 * neighbour_visit.hh:63-84), are concatenated into one candidate index
 * space: lanes take 64 consecutive candidates (coalesced loads, kept in
 * registers) and the wave loops over the home atoms, testing all 64
 * candidates against one home atom per step (the home atom's record comes
 * by a scalar load). The ballot of atom t, self pair removed
 * (neighbour_visit.hh:86), is written into lane t; its popcount is the
 * count. Home atoms are taken in groups of 64.
 */
template <int CUT, bool HASHED>
__device__ __forceinline__ void count_cell(const PairArgs &A, int h, int part, SegLds &S,
                                           lds_dptr pt_sq) {
    HomeCell H;
    if (!load_home<HASHED>(A, h, H)) return;
    const int lane = lane_id();
    int lo, hi;
    home_part(H.m, part, H.split, lo, hi);
    for (int grp0 = lo; grp0 < hi; grp0 += kWave) {
        const int mg = uniform(min(kWave, hi - grp0));
        const int hcls = CUT == CUT_PER_TYPE_LDS && lane < mg
                             ? home_class(A, A.stype[H.hb + grp0 + lane])
                             : 0;
        int acc = 0;
        for (int sg = 0; sg < H.nseg; sg += kWave) {
            const int nsg = min(kWave, H.nseg - sg);
            int self_start;
            const int total = build_segments<HASHED>(A, H, sg, nsg, lane, S, &self_start);
            const int self_base = self_start >= 0 ? self_start + grp0 : -(1 << 30);
            for (int cb = 0; cb < total; cb += kWave) {
                Cand C;
                load_candidate<CUT, false>(A, H, S, sg, nsg, cb + lane, total, C);
                const int c = cb + lane;
                uint32_t mlo = 0, mhi = 0;
                auto home = [&](int t, bool ok) {
                    ok = ok & (c != self_base + t);
                    const uint64_t mask = __builtin_amdgcn_ballot_w64(ok);
                    mlo = write_lane(mlo, static_cast<uint32_t>(mask), t);
                    mhi = write_lane(mhi, static_cast<uint32_t>(mask >> 32), t);
                };
                if constexpr (CUT == CUT_PER_TYPE_LDS) {
                    /* home atoms class by class, each lane's limit per class */
                    for (int v = 0; v <= A.ncut; v++) {
                        uint64_t todo = __builtin_amdgcn_ballot_w64(lane < mg && hcls == v);
                        if (!todo) continue;
                        const double lim = class_limit(A, H, pt_sq, v, C.tj);
                        while (todo) {
                            const int t = static_cast<int>(__builtin_ctzll(todo));
                            todo &= todo - 1;
                            const double4 pi = load_const(A.spos, H.hb + grp0 + t);
                            double d0, d1, d2;
                            home(t, pair_test_lim(C, pi, lim, d0, d1, d2));
                        }
                    }
                } else {
                    for (int t = 0; t < mg; t++) {
                        const int si = H.hb + grp0 + t;
                        const double4 pi = load_const(A.spos, si);
                        double d0, d1, d2;
                        home(t, pair_test<CUT>(A, H, C, si, pi, d0, d1, d2, pt_sq));
                    }
                }
                acc += __popc(mlo) + __popc(mhi);
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (lane < mg) A.counts[H.abase + A.sinfo[H.hb + grp0 + lane].w] = acc;
    }
}

/*
 * Count pass over half the neighbour-cell offsets. The test is symmetric:
 * for the reversed pair every operand is negated exactly (pi - pj = -(pj -
 * pi), the offset vector of -x is -off), so dr and r2 carry the same bits
 * and (j, i, -x) is accepted exactly when (i, j, x) is; the per-atom sum of
 * radii commutes, and per-type tables are used here only when symmetric.
 * The wave therefore visits the offsets from (0,0,0) on in traversal order
 * - (z, y, x) lexicographically >= 0, the segments from the centre to the
 * end - and inside the home cell only candidates after the home atom. Each
 * accepted pair counts for both atoms: the home atom's count is the ballot's
 * popcount, the candidate's the number of home atoms it passed; both are
 * added atomically into the cell-sorted counts.
 */
/* acc + (bit `lane` of mask): one add with the mask as carry-in. */
__device__ __forceinline__ int add_lane_bit(int acc, uint64_t mask) {
    uint64_t carry_out;
    asm volatile("v_addc_co_u32_e64 %0, %1, 0, %0, %2"
                 : "+v"(acc), "=s"(carry_out)
                 : "s"(mask));
    return acc;
}

/* One candidate chunk against home atoms grp0 .. grp0+mg-1 (half count):
   lane t of `icnt` gets home atom t's accepted candidates, cj the home
   atoms each candidate accepted. CENTRE: the chunk holds home-cell atoms
   (candidates c < m); home atom g takes only those with c > g, a scalar
   mask over the chunk's low lanes. */
template <int CUT, bool CENTRE>
__device__ __forceinline__ void half_sweep(const PairArgs &A, const HomeCell &H, const Cand &C,
                                           int grp0, int mg, int cb, lds_dptr pt_sq, int hcls,
                                           uint32_t &icnt, int &cj) {
    auto home = [&](int t, auto &&test) {
        const int si = H.hb + grp0 + t;
        const double4 pi = load_const(A.spos, si);
        double d0, d1, d2;
        uint64_t mask = __builtin_amdgcn_ballot_w64(test(si, pi, d0, d1, d2));
        if (CENTRE) {
            const int hi = min(grp0 + t, H.m - 1) - cb; /* last lane to drop */
            if (hi >= kWave - 1)
                mask = 0;
            else if (hi >= 0)
                mask &= ~0ull << (hi + 1);
        }
        cj = add_lane_bit(cj, mask);
        icnt = write_lane(icnt, static_cast<uint32_t>(__popcll(mask)), t);
    };
    if constexpr (CUT == CUT_PER_TYPE_LDS) {
        for (int v = 0; v <= A.ncut; v++) {
            uint64_t todo = __builtin_amdgcn_ballot_w64(lane_id() < mg && hcls == v);
            if (!todo) continue;
            const double lim = class_limit(A, H, pt_sq, v, C.tj);
            while (todo) {
                const int t = static_cast<int>(__builtin_ctzll(todo));
                todo &= todo - 1;
                home(t, [&](int, const double4 &pi, double &d0, double &d1, double &d2) {
                    return pair_test_lim(C, pi, lim, d0, d1, d2);
                });
            }
        }
    } else {
        for (int t = 0; t < mg; t++)
            home(t, [&](int si, const double4 &pi, double &d0, double &d1, double &d2) {
                return pair_test<CUT>(A, H, C, si, pi, d0, d1, d2, pt_sq);
            });
    }
}

template <int CUT, bool HASHED>
__device__ __forceinline__ void count_cell_half(const PairArgs &A, int h, int part, SegLds &S,
                                                lds_dptr pt_sq) {
    HomeCell H;
    if (!load_home<HASHED>(A, h, H)) return;
    const int lane = lane_id();
    int lo, hi;
    home_part(H.m, part, H.split, lo, hi);
    for (int grp0 = lo; grp0 < hi; grp0 += kWave) {
        const int mg = uniform(min(kWave, hi - grp0));
        const int hcls = CUT == CUT_PER_TYPE_LDS && lane < mg
                             ? home_class(A, A.stype[H.hb + grp0 + lane])
                             : 0;
        int acc = 0;
        for (int sg = H.centre; sg < H.nseg; sg += kWave) {
            const int nsg = min(kWave, H.nseg - sg);
            int self_start;
            const int total = build_segments<HASHED>(A, H, sg, nsg, lane, S, &self_start);
            for (int cb = 0; cb < total; cb += kWave) {
                Cand C;
                load_candidate<CUT, false>(A, H, S, sg, nsg, cb + lane, total, C);
                uint32_t icnt = 0;
                int cj = 0;
                /* the centre segment leads the first group: its candidates
                   c < m are the home atoms, taken only after the home atom */
                if (sg == H.centre && cb < H.m)
                    half_sweep<CUT, true>(A, H, C, grp0, mg, cb, pt_sq, hcls, icnt, cj);
                else
                    half_sweep<CUT, false>(A, H, C, grp0, mg, cb, pt_sq, hcls, icnt, cj);
                acc += static_cast<int>(icnt);
                if (cj) atomicAdd(A.scount + C.j, cj);
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (lane < mg && acc) atomicAdd(A.scount + H.hb + grp0 + lane, acc);
    }
}

/* Per-type cutoff tables up to this many entries (11 types) are staged in
   LDS once per block (the test reads one entry per candidate and home atom);
   larger ones are read from global memory. At 128 entries four fill blocks
   still fit in a CU's 160 KiB of LDS. */
constexpr int kLdsTypes = 128;

template <int CUT>
__device__ __forceinline__ lds_dptr stage_types(const PairArgs &A, double *s_pt) {
    if (CUT != CUT_PER_TYPE_LDS) return nullptr;
    const int nt = A.ncut * A.ncut;
    for (int k = threadIdx.x; k < nt; k += kBlock) s_pt[k] = A.pt_sq[k];
    __syncthreads();
    return (lds_dptr)(s_pt);
}

/* 8 waves per SIMD (the scalar registers would otherwise allow 7; the
   excess spills to vector lanes): the count is latency-bound in its per-cell
   set-up, and the extra wave hides some of it. Leaving the compiler 6 to 8
   cost the fcc 1M step 12 us. */
template <int CUT, bool HASHED, bool HALF>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_count(
    PairArgs A) {
    __shared__ SegLds s_lds[kWavesPerBlock];
    __shared__ double s_pt[CUT == CUT_PER_TYPE_LDS ? kLdsTypes : 1];
    const lds_dptr pt_sq = stage_types<CUT>(A, s_pt);
    const int wib = uniform(threadIdx.x / kWave);
    const int blk = xcd_remap(blockIdx.x, A.xcd_blocks);
    if (!HASHED) {
        int h, part;
        wave_home(A, blk * kWavesPerBlock + wib, h, part);
        if (h >= A.home_cells) return;
        if (HALF)
            count_cell_half<CUT, HASHED>(A, h, part, s_lds[wib], pt_sq);
        else
            count_cell<CUT, HASHED>(A, h, part, s_lds[wib], pt_sq);
        return;
    }
    const int64_t items = pair_items(A);
    for (int64_t v = blk * kWavesPerBlock + wib; v < items;
         v += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
        int idx, part;
        wave_home(A, static_cast<int>(v), idx, part);
        const int h = item_cell(A, idx);
        if (HALF)
            count_cell_half<CUT, HASHED>(A, h, part, s_lds[wib], pt_sq);
        else
            count_cell<CUT, HASHED>(A, h, part, s_lds[wib], pt_sq);
    }
}

/* ------------------------------------------------------- fused consumer */

/*
 * Lennard-Jones forces, energy and virial straight from the cell grid, with
 * no pair list. The traversal is the full count's: every ordered pair with
 * i a home atom and j a candidate, the pair's own test (r2 < rc^2, i != j
 * at zero offset) and its D = r_j - r_i. With s6 = (sigma^2 / r2)^3:
 *   coef = -24 eps (2 s6^2 - s6) / r2, force on i coef D, on j -coef D
 * (examples/lj_langevin/lj_langevin_cpu.cc:76-90, there with eps = sigma =
 * 1 and D recomputed from the positions). Each lane sums the force on its
 * candidate from the home atoms and adds it to the cell-sorted force arrays
 * (one atomic per lane, component and chunk): over all home cells every
 * atom receives the force of each of its neighbours once, and no sum runs
 * across lanes. Per ordered pair the lane also adds half the pair energy
 * 4 eps (s6^2 - s6) and half the virial -coef D (x) D; these are summed
 * per block into partial[block][10] (energy, then the 3x3 virial row by
 * row) and reduced in a fixed order.
 */
template <bool HASHED>
__device__ __forceinline__ void lj_cell(const PairArgs &A, int h, int part, SegLds &S,
                                        double eps4, double sig2, double *__restrict__ fx,
                                        double *__restrict__ fy, double *__restrict__ fz,
                                        double acc[7]) {
    HomeCell H;
    if (!load_home<HASHED>(A, h, H)) return;
    const int lane = lane_id();
    int lo, hi;
    home_part(H.m, part, H.split, lo, hi);
    for (int grp0 = lo; grp0 < hi; grp0 += kWave) {
        const int mg = uniform(min(kWave, hi - grp0));
        for (int sg = 0; sg < H.nseg; sg += kWave) {
            const int nsg = min(kWave, H.nseg - sg);
            int self_start;
            const int total = build_segments<HASHED>(A, H, sg, nsg, lane, S, &self_start);
            const int self_base = self_start >= 0 ? self_start + grp0 : -(1 << 30);
            for (int cb = 0; cb < total; cb += kWave) {
                Cand C;
                load_candidate<CUT_GLOBAL, false>(A, H, S, sg, nsg, cb + lane, total, C);
                const int c = cb + lane;
                double f0 = 0.0, f1 = 0.0, f2 = 0.0;
                for (int t = 0; t < mg; t++) {
                    const int si = H.hb + grp0 + t;
                    const double4 pi = load_const(A.spos, si);
                    double d0, d1, d2;
                    const bool ok = pair_test<CUT_GLOBAL>(A, H, C, si, pi, d0, d1, d2, nullptr) &
                                    (c != self_base + t);
                    if (ok) {
                        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                        const double ir2 = 1.0 / r2;
                        const double s2 = sig2 * ir2;
                        const double s6 = s2 * s2 * s2;
                        const double coef = -6.0 * eps4 * ir2 * (2.0 * s6 * s6 - s6);
                        const double g0 = coef * d0, g1 = coef * d1, g2 = coef * d2;
                        f0 -= g0;
                        f1 -= g1;
                        f2 -= g2;
                        acc[0] += 0.5 * eps4 * (s6 * s6 - s6);
                        acc[1] -= 0.5 * g0 * d0;
                        acc[2] -= 0.5 * g1 * d1;
                        acc[3] -= 0.5 * g2 * d2;
                        acc[4] -= 0.5 * g0 * d1;
                        acc[5] -= 0.5 * g0 * d2;
                        acc[6] -= 0.5 * g1 * d2;
                    }
                }
                if (c < total && (f0 != 0.0 || f1 != 0.0 || f2 != 0.0)) {
                    unsafeAtomicAdd(fx + C.j, f0);
                    unsafeAtomicAdd(fy + C.j, f1);
                    unsafeAtomicAdd(fz + C.j, f2);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

template <bool HASHED>
__global__ __launch_bounds__(kBlock) void k_lj(PairArgs A, double eps4, double sig2,
                                               double *__restrict__ fx, double *__restrict__ fy,
                                               double *__restrict__ fz,
                                               double *__restrict__ partial) {
    __shared__ SegLds s_lds[kWavesPerBlock];
    __shared__ double s_red[kWavesPerBlock][7];
    const int wib = uniform(threadIdx.x / kWave);
    const int blk = xcd_remap(blockIdx.x, A.xcd_blocks);
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!HASHED) {
        int h, part;
        wave_home(A, blk * kWavesPerBlock + wib, h, part);
        if (h < A.home_cells) lj_cell<HASHED>(A, h, part, s_lds[wib], eps4, sig2, fx, fy, fz, acc);
    } else {
        const int64_t items = pair_items(A);
        for (int64_t v = blk * kWavesPerBlock + wib; v < items;
             v += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
            int idx, part;
            wave_home(A, static_cast<int>(v), idx, part);
            lj_cell<HASHED>(A, item_cell(A, idx), part, s_lds[wib], eps4, sig2, fx, fy, fz, acc);
        }
    }
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] = wave_sum(acc[k]);
    if (lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < 7; k++) s_red[wib][k] = acc[k];
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double t = 0.0;
        for (int q = 0; q < kWavesPerBlock; q++) t += s_red[q][threadIdx.x];
        s_red[0][threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        /* energy, then the symmetric virial row by row */
        constexpr int kFrom[10] = {0, 1, 4, 5, 4, 2, 6, 5, 6, 3};
        partial[static_cast<int64_t>(blockIdx.x) * 10 + threadIdx.x] = s_red[0][kFrom[threadIdx.x]];
    }
}

/* Forces to atom order ([nat][3]: atom a's sums sit at its cell-sorted
   index), and the reduced energy and virial to the caller's pointers. */
__global__ __launch_bounds__(kBlock) void k_lj_finish(const double *__restrict__ fx,
                                                      const double *__restrict__ fy,
                                                      const double *__restrict__ fz,
                                                      const int32_t *__restrict__ slot,
                                                      int32_t nat, const double *__restrict__ sums,
                                                      double *__restrict__ forces,
                                                      double *__restrict__ energy,
                                                      double *__restrict__ virial) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a < nat) {
        const int s = slot[a];
        forces[3 * a + 0] = fx[s];
        forces[3 * a + 1] = fy[s];
        forces[3 * a + 2] = fz[s];
    }
    if (a == 0) {
        if (energy) energy[0] = sums[0];
        if (virial)
            for (int k = 0; k < 9; k++) virial[k] = sums[1 + k];
    }
}

/* ------------------------------------------------------------- fill pass */

/* Candidate chunks held in registers while the home atoms are swept (the
   default; a launch may take 5, see launch_fill_t), and the ring of
   accepted pairs waiting for a full 64-lane store. */
constexpr int kFillChunks = 4;
constexpr int kRing = 128;

/* Shift along one direction: periodic S = (raw_i + x - raw_j) / n, an exact
   quotient (raw = wrapped cell + n * image on both sides, and the visited
   cell is c_i + x wrapped), taken as rint((double)num * (1/n)), exact for
   |S| < 2^51; non-periodic S = image count of i. This is the reference's
   truncating (raw_i - rel_j + x) / n (neighbour_visit.hh:111-114). */
__device__ __forceinline__ int shift_of(int periodic, int bi, int cj, double rn) {
    return periodic ? static_cast<int>(rint(static_cast<double>(bi + cj) * rn)) : bi;
}

/* Write one pair: D = dr as the test computed it, d = sqrt(r2) with r2 in
   the test's operation order (the bits the test saw). */
template <int MODE>
__device__ __forceinline__ void write_pair(const PairArgs &A, int64_t pos, int i, int j,
                                           double dx, double dy, double dz, int s0, int s1,
                                           int s2) {
    if (pos < 0 || pos >= A.capacity) return;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (A.out_i) A.out_i[pos] = i;
    if (A.out_j) A.out_j[pos] = j;
    if (MODE == MODE_FILL64) {
        if (A.out_D) {
            double *D = static_cast<double *>(A.out_D) + 3 * pos;
            D[0] = dx;
            D[1] = dy;
            D[2] = dz;
        }
        if (A.out_d) static_cast<double *>(A.out_d)[pos] = sqrt(r2);
    } else {
        if (A.out_D) {
            float *D = static_cast<float *>(A.out_D) + 3 * pos;
            D[0] = static_cast<float>(dx);
            D[1] = static_cast<float>(dy);
            D[2] = static_cast<float>(dz);
        }
        if (A.out_d) static_cast<float *>(A.out_d)[pos] = static_cast<float>(sqrt(r2));
    }
    if (A.out_S) {
        int32_t *S = A.out_S + 3 * pos;
        S[0] = s0;
        S[1] = s1;
        S[2] = s2;
    }
}

/* Fill-pass LDS of one wave: segment table, the home atoms of the current
   group (output row, image/raw cell, index) and the ring of accepted pairs
   (dr, j, shift numerator part, home atom slot, rank in the atom's row).
   The rank has a word of its own: a row may hold up to 2^31 - 1 pairs, as
   the reference's int32 offsets allow. */
struct FillLds {
    SegLds seg;
    int64_t hpos[kWave];
    int hbx[kWave], hby[kWave], hbz[kWave], hi[kWave];
    double rdx[kRing], rdy[kRing], rdz[kRing];
    int rj[kRing], rc0[kRing], rc1[kRing], rc2[kRing], rt[kRing], rrank[kRing];
};

/* Write n (<= 64) ring entries from `tail` on, one per lane. */
template <int MODE>
__device__ __forceinline__ void flush_ring(const PairArgs &A, const HomeCell &H, FillLds &L,
                                           int tail, int n) {
    const int lane = lane_id();
    wave_sync_lds();
    if (lane < n) {
        const int e = (tail + lane) & (kRing - 1);
        const int t = L.rt[e] & (kWave - 1);
        const int rank = L.rrank[e];
        const int64_t pos = L.hpos[t] + rank;
        if (MODE == MODE_MATRIX) {
            /* row of the dense matrix: neighbour index and D, ranks < K */
            if (rank >= 0 && rank < A.max_nb) {
                A.out_j[pos] = L.rj[e];
                double *D = static_cast<double *>(A.out_D) + 3 * pos;
                D[0] = L.rdx[e];
                D[1] = L.rdy[e];
                D[2] = L.rdz[e];
            }
        } else {
            write_pair<MODE>(A, pos, L.hi[t], L.rj[e], L.rdx[e], L.rdy[e], L.rdz[e],
                             shift_of(H.p0, L.hbx[t], L.rc0[e], H.rn0),
                             shift_of(H.p1, L.hby[t], L.rc1[e], H.rn1),
                             shift_of(H.p2, L.hbz[t], L.rc2[e], H.rn2));
        }
    }
    wave_sync_lds();
}

/*
 * Fill pass, same traversal as the count. Per window of up to FC
 * candidate chunks (held in registers) the wave sweeps the home atoms and,
 * for each atom, the window's chunks in order: the test is repeated and the
 * accepted lanes append (dr, j, shift part, atom, rank) to an LDS ring - atom
 * by atom, in candidate order, i.e. in output order within each atom. Every
 * time the ring holds 64 pairs they are written with one full-width store
 * per output array, at the atom's row offset plus the pair's rank. Per-atom
 * ranks continue across windows, so an atom's row fills front to back.
 */
template <int MODE, int CUT, bool HASHED, int FC>
__device__ __forceinline__ void fill_cell(const PairArgs &A, int h, int part, FillLds &L,
                                          lds_dptr pt_sq) {
    HomeCell H;
    if (!load_home<HASHED>(A, h, H)) return;
    const int lane = lane_id();
    int lo, hi;
    home_part(H.m, part, H.split, lo, hi);
    for (int grp0 = lo; grp0 < hi; grp0 += kWave) {
        const int mg = uniform(min(kWave, hi - grp0));
        if (lane < mg) {
            const int4 inf = A.sinfo[H.hb + grp0 + lane];
            /* pair list: the atom's row offset; matrix: its row of K slots */
            L.hpos[lane] = MODE == MODE_MATRIX
                               ? static_cast<int64_t>(H.abase + inf.w) * A.max_nb
                               : A.offsets[H.abase + inf.w];
            L.hbx[lane] = inf.x;
            L.hby[lane] = inf.y;
            L.hbz[lane] = inf.z;
            L.hi[lane] = inf.w;
        }
        const int hcls = CUT == CUT_PER_TYPE_LDS && lane < mg
                             ? home_class(A, A.stype[H.hb + grp0 + lane])
                             : 0;
        wave_sync_lds();
        int acc = 0;     /* lane t: pairs of atom t emitted so far */
        int head = 0;    /* ring positions (uniform) */
        int tail = 0;
        for (int sg = 0; sg < H.nseg; sg += kWave) {
            const int nsg = min(kWave, H.nseg - sg);
            int self_start;
            const int total = build_segments<HASHED>(A, H, sg, nsg, lane, L.seg, &self_start);
            const int self_base = self_start >= 0 ? self_start + grp0 : -(1 << 30);
            for (int w0 = 0; w0 < total; w0 += kWave * FC) {
                Cand C[FC];
#pragma unroll
                for (int k = 0; k < FC; k++)
                    if (w0 + k * kWave < total)
                        load_candidate<CUT, true>(A, H, L.seg, sg, nsg, w0 + k * kWave + lane,
                                                  total, C[k]);
                if constexpr (CUT == CUT_PER_TYPE_LDS) {
                    /* home atoms class by class, each lane's limit per class */
                    for (int v = 0; v <= A.ncut; v++) {
                        uint64_t todo = __builtin_amdgcn_ballot_w64(lane < mg && hcls == v);
                        if (!todo) continue;
                        double lim[FC];
#pragma unroll
                        for (int k = 0; k < FC; k++)
                            lim[k] = class_limit(A, H, pt_sq, v, C[k].tj);
                        while (todo) {
                            const int t = static_cast<int>(__builtin_ctzll(todo));
                            todo &= todo - 1;
                            const int si = H.hb + grp0 + t;
                            const double4 pi = load_const(A.spos, si);
                            /* atom t's pairs so far: a scalar across the window */
                            int rank0 = __builtin_amdgcn_readlane(acc, t);
#pragma unroll
                            for (int k = 0; k < FC; k++) {
                                const int cb = w0 + k * kWave;
                                if (cb >= total) break;
                                double d0, d1, d2;
                                uint64_t mask = __builtin_amdgcn_ballot_w64(
                                    pair_test_lim(C[k], pi, lim[k], d0, d1, d2));
                                /* the atom itself at zero offset (neighbour_visit.hh:86) */
                                const int self_lane = self_base + t - cb;
                                if (self_lane >= 0 && self_lane < kWave) mask &= ~(1ull << self_lane);
                                if (mask) {
                                    if (__builtin_amdgcn_inverse_ballot_w64(mask)) {
                                        const int r = mbcnt64(mask);
                                        const int e = (head + r) & (kRing - 1);
                                        L.rdx[e] = d0;
                                        L.rdy[e] = d1;
                                        L.rdz[e] = d2;
                                        L.rj[e] = C[k].j;
                                        L.rc0[e] = C[k].c0;
                                        L.rc1[e] = C[k].c1;
                                        L.rc2[e] = C[k].c2;
                                        L.rt[e] = t;
                                        L.rrank[e] = rank0 + r;
                                    }
                                    const int pc = static_cast<int>(__popcll(mask));
                                    rank0 += pc;
                                    head += pc;
                                    if (head - tail >= kWave) {
                                        flush_ring<MODE>(A, H, L, tail, kWave);
                                        tail += kWave;
                                    }
                                }
                            }
                            acc = static_cast<int>(
                                write_lane(static_cast<uint32_t>(acc), static_cast<uint32_t>(rank0), t));
                        }
                    }
                } else {
                    for (int t = 0; t < mg; t++) {
                        const int si = H.hb + grp0 + t;
                        const double4 pi = load_const(A.spos, si);
                        /* atom t's pairs so far: a scalar across the window */
                        int rank0 = __builtin_amdgcn_readlane(acc, t);
#pragma unroll
                        for (int k = 0; k < FC; k++) {
                            const int cb = w0 + k * kWave;
                            if (cb >= total) break;
                            double d0, d1, d2;
                            uint64_t mask = __builtin_amdgcn_ballot_w64(
                                pair_test<CUT>(A, H, C[k], si, pi, d0, d1, d2, pt_sq));
                            /* the atom itself at zero offset (neighbour_visit.hh:86) */
                            const int self_lane = self_base + t - cb;
                            if (self_lane >= 0 && self_lane < kWave) mask &= ~(1ull << self_lane);
                            if (mask) {
                                if (__builtin_amdgcn_inverse_ballot_w64(mask)) {
                                    const int r = mbcnt64(mask);
                                    const int e = (head + r) & (kRing - 1);
                                    L.rdx[e] = d0;
                                    L.rdy[e] = d1;
                                    L.rdz[e] = d2;
                                    L.rj[e] = C[k].j;
                                    L.rc0[e] = C[k].c0;
                                    L.rc1[e] = C[k].c1;
                                    L.rc2[e] = C[k].c2;
                                    L.rt[e] = t;
                                    L.rrank[e] = rank0 + r;
                                }
                                const int pc = static_cast<int>(__popcll(mask));
                                rank0 += pc;
                                head += pc;
                                if (head - tail >= kWave) {
                                    flush_ring<MODE>(A, H, L, tail, kWave);
                                    tail += kWave;
                                }
                            }
                        }
                        acc = static_cast<int>(
                            write_lane(static_cast<uint32_t>(acc), static_cast<uint32_t>(rank0), t));
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (head > tail) flush_ring<MODE>(A, H, L, tail, head - tail);
        if (MODE == MODE_MATRIX) {
            /* the matrix has no count pass: the traversal's per-atom total
               is the row count. Each row's unused slots (from the count on)
               are zeroed here, slots the flushes above never write, so the
               matrix needs no clearing beforehand. */
            if (lane < mg) {
                A.counts[H.abase + L.hi[lane]] = acc;
                if (acc > A.max_nb && A.overflow) atomicMax(A.overflow, 1);
            }
            double *dist = static_cast<double *>(A.out_D);
            for (int t = 0; t < mg; t++) {
                const int used = min(__builtin_amdgcn_readlane(acc, t), A.max_nb);
                const int64_t row = L.hpos[t];
                for (int sl = used + lane; sl < A.max_nb; sl += kWave) {
                    A.out_j[row + sl] = 0;
                    dist[3 * (row + sl)] = 0.0;
                    dist[3 * (row + sl) + 1] = 0.0;
                    dist[3 * (row + sl) + 2] = 0.0;
                }
            }
        }
    }
}

/* Launch bounds of the fill: at least 4 waves per SIMD (at most 128 VGPRs)
   where the compiler would otherwise take a few registers more and run 3,
   and the cap costs no spills: the 5-chunk uniform-cutoff fills other than
   the dense float64 one (float32 outputs on fcc 1M: 1016 -> 936 us per
   step) and the hashed 4-chunk fills (per-type slab on the hashed grid:
   326 -> 313 us). Elsewhere the compiler's own allocation stays: capping
   the dense float64 and dense 4-chunk fills changes their code for a
   slight loss, and the 5-chunk per-atom and per-type fills would spill. */
template <int MODE, int CUT, bool HASHED, int FC>
constexpr int fill_min_waves() {
    if (FC == 4) return HASHED ? 4 : 1;
    return (CUT == CUT_GLOBAL && (MODE != MODE_FILL64 || HASHED)) ? 4 : 1;
}

template <int MODE, int CUT, bool HASHED, int FC>
__global__ __launch_bounds__(kBlock, (fill_min_waves<MODE, CUT, HASHED, FC>())) void k_fill(
    PairArgs A) {
    __shared__ FillLds s_lds[kWavesPerBlock];
    __shared__ double s_pt[CUT == CUT_PER_TYPE_LDS ? kLdsTypes : 1];
    const lds_dptr pt_sq = stage_types<CUT>(A, s_pt);
    const int wib = uniform(threadIdx.x / kWave);
    const int blk = xcd_remap(blockIdx.x, A.xcd_blocks);
    if (!HASHED) {
        int h, part;
        wave_home(A, blk * kWavesPerBlock + wib, h, part);
        if (h >= A.home_cells) return;
        fill_cell<MODE, CUT, HASHED, FC>(A, h, part, s_lds[wib], pt_sq);
        return;
    }
    const int64_t items = pair_items(A);
    for (int64_t v = blk * kWavesPerBlock + wib; v < items;
         v += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
        int idx, part;
        wave_home(A, static_cast<int>(v), idx, part);
        fill_cell<MODE, CUT, HASHED, FC>(A, item_cell(A, idx), part, s_lds[wib], pt_sq);
    }
}

/* The previous binning measured, per axis, how far atoms 32 indices apart
   lie (in cells). If one axis carries at least twice the distance of each
   other, the input's atom order runs along it: make it the fastest of the
   dense layout (then the next), so that cells processed one after another
   hold atoms with nearby indices and the output rows they write lie close
   together. */
__device__ __forceinline__ void choose_layout(int s0, int s1, int s2, SysGeom &g) {
    const int f = (s0 >= s1 && s0 >= s2) ? 0 : (s1 >= s2 ? 1 : 2);
    const int u = f == 0 ? 1 : 0, v = f == 2 ? 1 : 2; /* the other two, u < v */
    const int sf = pick3(f, s0, s1, s2), su = pick3(u, s0, s1, s2), sv = pick3(v, s0, s1, s2);
    if (sf > 2 * su && sf > 2 * sv) {
        g.lay[0] = f;
        g.lay[1] = su >= sv ? u : v;
        g.lay[2] = su >= sv ? v : u;
    }
}

/* Device-side parameter set-up (keeps the whole build free of host-to-device
   copies, so that it can be captured in a graph): the one-system parameters,
   and the cleared grid: cell (or slot) counts to 0, hash keys (if any) to
   empty. */
__global__ void k_set_params(SysGeom *__restrict__ geo, SysGeom g, int32_t *__restrict__ atom_base,
                             int32_t *__restrict__ cell_base, int32_t nat, int32_t ncells,
                             int32_t *__restrict__ cell_count, int64_t nslots,
                             int64_t *__restrict__ hkey, int32_t *__restrict__ order_stats) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < nslots) {
        cell_count[t] = 0;
        if (hkey) hkey[t] = -1;
    }
    if (t != 0) return;
    if (order_stats) {
        choose_layout(order_stats[0], order_stats[1], order_stats[2], g);
        order_stats[0] = order_stats[1] = order_stats[2] = 0;
    }
    geo[0] = g;
    atom_base[0] = 0;
    atom_base[1] = nat;
    cell_base[0] = 0;
    cell_base[1] = ncells;
}

/*
 * The whole binning of a small single system (dense grid, up to
 * kSmallAtoms atoms and kSmallCells cells) in one block, replacing
 * k_set_params, k_bin_dense, the cell scan, k_scatter and k_rank_gather
 * with the same results: the parameters (with the layout the previous
 * build's order sample selects) and the cell grid, cell slices in ascending
 * atom order, the cell-sorted records. The histogram, the scan and the
 * member lists live in LDS. The order sample for the next build is the one
 * k_bin_dense takes at this size (atoms 0..63, pairs a, a + 32).
 */
constexpr int kSmallThreads = 1024;
constexpr int kSmallAtoms = 4096;
constexpr int kSmallCells = 4096;
constexpr int kSmallPer = kSmallAtoms / kSmallThreads;

__global__ __launch_bounds__(kSmallThreads) void k_bin_small(
    const double *__restrict__ r, int32_t nat, SysGeom g0, int32_t ncells,
    SysGeom *__restrict__ geo, int32_t *__restrict__ atom_base, int32_t *__restrict__ cell_base,
    int32_t *__restrict__ order_stats, const double *__restrict__ per_atom,
    const int32_t *__restrict__ types, int32_t *__restrict__ cell_first,
    int32_t *__restrict__ cell_count, double4 *__restrict__ spos, int4 *__restrict__ sinfo,
    int32_t *__restrict__ stype, int32_t *__restrict__ scount, int32_t *__restrict__ slot) {
    __shared__ int s_cnt[kSmallCells];
    __shared__ int s_first[kSmallCells];
    __shared__ int s_tmp[kSmallAtoms];
    __shared__ int s_w[kSmallThreads / kWave];
    __shared__ SysGeom s_g;
    const int tid = threadIdx.x;
    if (tid == 0) {
        SysGeom h = g0;
        if (order_stats) choose_layout(order_stats[0], order_stats[1], order_stats[2], h);
        s_g = h;
        geo[0] = h;
        atom_base[0] = 0;
        atom_base[1] = nat;
        cell_base[0] = 0;
        cell_base[1] = ncells;
    }
    for (int c = tid; c < ncells; c += kSmallThreads) s_cnt[c] = 0;
    __syncthreads();
    const SysGeom &g = s_g;

    /* cells, and tickets in arrival order */
    int raw[kSmallPer][3], cell[kSmallPer], tk[kSmallPer];
    int c0[3] = {0, 0, 0};
#pragma unroll
    for (int u = 0; u < kSmallPer; u++) {
        const int a = tid + u * kSmallThreads;
        cell[u] = -1;
        tk[u] = 0;
        raw[u][0] = raw[u][1] = raw[u][2] = 0;
        if (a < nat) {
            raw_cell_of(r, a, g, raw[u]);
            int c[3];
#pragma unroll
            for (int k = 0; k < 3; k++)
                c[k] = g.pbc[k] ? wrap_idx(raw[u][k], g.n[k]) : clamp_idx(raw[u][k], g.n[k]);
            cell[u] = dense_index(g, c[0], c[1], c[2]);
            tk[u] = atomicAdd(&s_cnt[cell[u]], 1);
            if (u == 0) {
                c0[0] = c[0];
                c0[1] = c[1];
                c0[2] = c[2];
            }
        }
    }
    if (order_stats && tid < kWave) {
        /* all lanes take part in the shuffles and sums; the store waits for
           every thread's read of the previous sample (barriers below) */
        const bool pair = tid < 32 && tid + 32 < nat;
        int moved[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int cn = __shfl_down(c0[k], 32, kWave);
            int dd = cn - c0[k];
            if (2 * dd > g.n[k]) dd -= g.n[k];
            if (2 * dd < -g.n[k]) dd += g.n[k];
            moved[k] = wave_sum(pair ? min(dd < 0 ? -dd : dd, 64) : 0);
        }
        if (tid == 0) {
            s_w[0] = moved[0];
            s_w[1] = moved[1];
            s_w[2] = moved[2];
        }
    }
    __syncthreads();
    if (order_stats && tid == 0) {
        order_stats[0] = s_w[0];
        order_stats[1] = s_w[1];
        order_stats[2] = s_w[2];
    }
    __syncthreads();

    /* exclusive scan of the cell counts: kSmallPer consecutive cells per
       thread, then the thread sums across the block */
    int loc[kSmallPer];
    int sum = 0;
#pragma unroll
    for (int u = 0; u < kSmallPer; u++) {
        const int c = tid * kSmallPer + u;
        loc[u] = sum;
        sum += c < ncells ? s_cnt[c] : 0;
    }
    const int incl = wave_incl_scan(sum);
    if (lane_id() == kWave - 1) s_w[tid / kWave] = incl;
    __syncthreads();
    int pre = incl - sum;
    for (int w = 0; w < tid / kWave; w++) pre += s_w[w];
#pragma unroll
    for (int u = 0; u < kSmallPer; u++) {
        const int c = tid * kSmallPer + u;
        if (c < ncells) {
            s_first[c] = pre + loc[u];
            cell_first[c] = pre + loc[u];
            cell_count[c] = s_cnt[c];
        }
    }
    __syncthreads();

    /* member lists, then the stable rank: members with a smaller index */
#pragma unroll
    for (int u = 0; u < kSmallPer; u++)
        if (cell[u] >= 0) s_tmp[s_first[cell[u]] + tk[u]] = tid + u * kSmallThreads;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSmallPer; u++) {
        if (cell[u] < 0) continue;
        const int a = tid + u * kSmallThreads;
        const int b = s_first[cell[u]], m = s_cnt[cell[u]];
        int rank = 0;
        for (int t = 0; t < m; t++) rank += s_tmp[b + t] < a ? 1 : 0;
        place_atom(r, a, b + rank, raw[u], g, per_atom, types, spos, sinfo, stype, scount, slot);
    }
}

/*
 * The whole binning of a batch of small systems, one block per system (up
 * to kSysAtoms atoms and kSysCells cells each): the same results as
 * k_bin_dense, the cell scan, k_scatter and k_rank_gather over the batch.
 * Every system's atoms land in its own cells, so its cell slices start at
 * its atom base and the block's own scan places them. The histogram, the
 * scan and the member lists live in LDS.
 */
constexpr int kSysThreads = 256;
constexpr int kSysAtoms = 2048;
constexpr int kSysCells = 2048;
constexpr int kSysPer = kSysAtoms / kSysThreads;

__global__ __launch_bounds__(kSysThreads) void k_bin_systems(
    const double *__restrict__ r, const SysGeom *__restrict__ geo,
    int32_t *__restrict__ cell_first, int32_t *__restrict__ cell_count,
    double4 *__restrict__ spos, int4 *__restrict__ sinfo, int32_t *__restrict__ scount,
    int32_t *__restrict__ slot) {
    __shared__ int s_cnt[kSysCells];
    __shared__ int s_first[kSysCells];
    __shared__ int s_tmp[kSysAtoms];
    __shared__ int s_w[kSysThreads / kWave];
    __shared__ SysGeom s_g;
    const int tid = threadIdx.x;
    if (tid == 0) s_g = geo[blockIdx.x];
    __syncthreads();
    const SysGeom &g = s_g;
    const int nat = g.natoms, ncells = g.ncells, abase = g.atom_base, cbase = g.cell_base;
    for (int c = tid; c < ncells; c += kSysThreads) s_cnt[c] = 0;
    __syncthreads();

    /* cells, and tickets in arrival order */
    int raw[kSysPer][3], cell[kSysPer], tk[kSysPer];
#pragma unroll
    for (int u = 0; u < kSysPer; u++) {
        const int a = tid + u * kSysThreads;
        cell[u] = -1;
        tk[u] = 0;
        raw[u][0] = raw[u][1] = raw[u][2] = 0;
        if (a < nat) {
            raw_cell_of(r, abase + a, g, raw[u]);
            int c[3];
#pragma unroll
            for (int k = 0; k < 3; k++)
                c[k] = g.pbc[k] ? wrap_idx(raw[u][k], g.n[k]) : clamp_idx(raw[u][k], g.n[k]);
            cell[u] = dense_index(g, c[0], c[1], c[2]);
            tk[u] = atomicAdd(&s_cnt[cell[u]], 1);
        }
    }
    __syncthreads();

    /* exclusive scan of the cell counts: kSysPer consecutive cells per
       thread, then across the block */
    int loc[kSysPer];
    int sum = 0;
#pragma unroll
    for (int u = 0; u < kSysPer; u++) {
        const int c = tid * kSysPer + u;
        loc[u] = sum;
        sum += c < ncells ? s_cnt[c] : 0;
    }
    const int incl = wave_incl_scan(sum);
    if (lane_id() == kWave - 1) s_w[tid / kWave] = incl;
    __syncthreads();
    int pre = incl - sum;
    for (int w = 0; w < tid / kWave; w++) pre += s_w[w];
#pragma unroll
    for (int u = 0; u < kSysPer; u++) {
        const int c = tid * kSysPer + u;
        if (c < ncells) {
            s_first[c] = pre + loc[u];
            cell_first[cbase + c] = abase + pre + loc[u];
            cell_count[cbase + c] = s_cnt[c];
        }
    }
    __syncthreads();

    /* member lists, then the stable rank: members with a smaller index */
#pragma unroll
    for (int u = 0; u < kSysPer; u++)
        if (cell[u] >= 0) s_tmp[s_first[cell[u]] + tk[u]] = tid + u * kSysThreads;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSysPer; u++) {
        if (cell[u] < 0) continue;
        const int a = tid + u * kSysThreads;
        const int b = s_first[cell[u]], m = s_cnt[cell[u]];
        int rank = 0;
        for (int t = 0; t < m; t++) rank += s_tmp[b + t] < a ? 1 : 0;
        place_atom(r, abase + a, abase + b + rank, raw[u], g, nullptr, nullptr, spos, sinfo,
                   nullptr, scount, slot);
    }
}

/* ---------------------------------------------------------------- backward */

/* Inclusive sums of (a, b, c) over the lanes of the same key (keys are
   non-decreasing across the wave, so equal keys are contiguous): lane l adds
   lane l - s's partial when their keys match, for s = 1, 2, 4, 8 inside each
   row of 16 lanes (DPP row shifts), then the last lane of row 0 into row 1
   and of row 2 into row 3, then lane 31 into rows 2 and 3 (DPP row
   broadcasts). The lane moves run on the VALU rather than through LDS, and
   the key moves once per step for all three sums. A lane without a source
   receives 0 and no key. */
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(u & 0xffffffffu), CTRL,
                                               ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(u >> 32), CTRL, ROWS, 0xf,
                                               false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                          static_cast<uint32_t>(lo));
}

template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_step(double &a, double &b, double &c, int key) {
    constexpr int kNoKey = -2147483647 - 1; /* below every key (keys are >= -1) */
    const int k = __builtin_amdgcn_update_dpp(kNoKey, key, CTRL, ROWS, 0xf, false);
    const double ta = dpp_move<CTRL, ROWS>(a);
    const double tb = dpp_move<CTRL, ROWS>(b);
    const double tc = dpp_move<CTRL, ROWS>(c);
    if (k == key) {
        a += ta;
        b += tb;
        c += tc;
    }
}

__device__ __forceinline__ void seg_incl_sum3(double &a, double &b, double &c, int key) {
    seg_step<0x111, 0xf>(a, b, c, key); /* row_shr:1 */
    seg_step<0x112, 0xf>(a, b, c, key); /* row_shr:2 */
    seg_step<0x114, 0xf>(a, b, c, key); /* row_shr:4 */
    seg_step<0x118, 0xf>(a, b, c, key); /* row_shr:8 */
    seg_step<0x142, 0xa>(a, b, c, key); /* row_bcast:15 into rows 1, 3 */
    seg_step<0x143, 0xc>(a, b, c, key); /* row_bcast:31 into rows 2, 3 */
}

/* Per-block table of j-side sums (LDS): slots, and linear probes tried
   before a pair adds straight to global memory. */
constexpr int kHashSlots = 1024;
constexpr int kHashProbe = 8;

/*
 * Adds the block's table of j-side sums to the global ones, in ascending j.
 * The table holds its atoms in hash order; added in that order, the 64
 * atomics of a wave land on about as many cache lines of each of gx, gy,
 * gz. Sorted, consecutive lanes take consecutive atoms of the index runs a
 * block's neighbourhoods consist of (measured on bcc 131k: 333 -> 212 us for
 * the backward with the flush's atomics on consecutive addresses).
 *   1. Every thread keeps its kHashSlots / kBlock slots (key, sums) in
 *      registers.
 *   2. The slot numbers are sorted by key (bitonic, in the first sum array's
 *      LDS: key in the high word, empty slots last, slot number low).
 *   3. s_key[slot] = sorted position; each slot's owner writes its sums
 *      there, and thread t adds sorted positions t, t + kBlock, ...
 * Called by all threads of the block after the table's last update.
 */
__device__ __forceinline__ void flush_table_sorted(int *s_key, double (*s_g)[kHashSlots],
                                                   double *__restrict__ gx,
                                                   double *__restrict__ gy,
                                                   double *__restrict__ gz) {
    constexpr int kPer = kHashSlots / kBlock;
    static_assert(kPer * kBlock == kHashSlots, "whole slots per thread");
    int key[kPer];
    double v0[kPer], v1[kPer], v2[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int k = static_cast<int>(threadIdx.x) + r * kBlock;
        key[r] = s_key[k];
        v0[r] = s_g[0][k];
        v1[r] = s_g[1][k];
        v2[r] = s_g[2][k];
    }
    __syncthreads();
    uint64_t *s_sort = reinterpret_cast<uint64_t *>(s_g[0]);
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int k = static_cast<int>(threadIdx.x) + r * kBlock;
        /* an empty slot (-1) gets the largest key: it sorts last */
        s_sort[k] = (static_cast<uint64_t>(static_cast<uint32_t>(key[r])) << 32) |
                    static_cast<uint32_t>(k);
    }
    for (int size = 2; size <= kHashSlots; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int e = threadIdx.x; e < kHashSlots / 2; e += kBlock) {
                const int lo = 2 * e - (e & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t a = s_sort[lo], b = s_sort[hi];
                if ((a > b) == up) {
                    s_sort[lo] = b;
                    s_sort[hi] = a;
                }
            }
        }
    }
    __syncthreads();
    int skey[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int pos = static_cast<int>(threadIdx.x) + r * kBlock;
        const uint64_t e = s_sort[pos];
        skey[r] = static_cast<int>(static_cast<uint32_t>(e >> 32));
        s_key[static_cast<int>(e & 0xffffffffu)] = pos;
    }
    __syncthreads();
    int mine[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) mine[r] = s_key[static_cast<int>(threadIdx.x) + r * kBlock];
    /* every read of s_sort (the first sum array) ended before the barrier */
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        s_g[0][mine[r]] = v0[r];
        s_g[1][mine[r]] = v1[r];
        s_g[2][mine[r]] = v2[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int pos = static_cast<int>(threadIdx.x) + r * kBlock;
        const int j = skey[r];
        if (j >= 0) {
            unsafeAtomicAdd(&gx[j], s_g[0][pos]);
            unsafeAtomicAdd(&gy[j], s_g[1][pos]);
            unsafeAtomicAdd(&gz[j], s_g[2][pos]);
        }
    }
}

/*
 * Backward of D_p = r[j] - r[i] + S_p . cell and d_p = |D_p| over an
 * i-sorted pair list. Per pair g = dL/dD + (dL/dd) D / d (the d term is
 * dropped for coincident atoms, d = 0). Each block takes a contiguous tile
 * of pairs:
 *   dL/dr[j] += g goes through a per-block LDS table keyed by j (open
 *     addressing, CAS on the key, LDS f64 adds): the rows of a tile are
 *     consecutive atoms whose neighbours overlap, so the table holds far
 *     fewer atoms than the tile has pairs, and it is flushed with one global
 *     atomic per atom and component; a pair whose probes all fail adds
 *     straight to global memory;
 *   dL/dr[i] -= g summed first over each i-segment inside the wave (one
 *     atomic per segment per 64 pairs);
 *   dL/dcell[a][b] = sum_p S_pa g_pb over the periodic directions a (along
 *     a non-periodic one D = r_j - r_i does not involve the cell, whatever
 *     S holds there), kept per thread, reduced per block into
 *     cell_partial[block][9] (fixed tiles, fixed order: deterministic).
 * The position sums go to separate x, y, z arrays (gx, gy, gz).
 */
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pair_backward(
    int64_t npairs, int64_t tile, const int32_t *__restrict__ first,
    const int32_t *__restrict__ secnd, const int32_t *__restrict__ shift,
    const T *__restrict__ D, const T *__restrict__ d, const T *__restrict__ gD,
    const T *__restrict__ gd, int pbc_mask, double *__restrict__ gx, double *__restrict__ gy,
    double *__restrict__ gz, double *__restrict__ cell_partial) {
    __shared__ double s_red[kWavesPerBlock][9];
    __shared__ int s_key[kHashSlots];
    __shared__ double s_g[3][kHashSlots];
    for (int k = threadIdx.x; k < kHashSlots; k += kBlock) {
        s_key[k] = -1;
        s_g[0][k] = 0.0;
        s_g[1][k] = 0.0;
        s_g[2][k] = 0.0;
    }
    __syncthreads();
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = 0.0;
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * tile;
    const int64_t t1 = t0 + tile < npairs ? t0 + tile : npairs;
    /* every lane of a wave runs the same trip count so that the lane moves
       stay converged; lanes past the tile carry key -1 and zeros */
    for (int64_t wb = t0 + (threadIdx.x & ~(kWave - 1)); wb < t1; wb += kBlock) {
        const int64_t p = wb + lane_id();
        const bool valid = p < t1;
        int i = -1, j = -1;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (valid) {
            i = first[p];
            j = secnd[p];
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
            /* the cell enters D only along periodic directions */
            s0 = (pbc_mask & 1) ? static_cast<double>(shift[3 * p + 0]) : 0.0;
            s1 = (pbc_mask & 2) ? static_cast<double>(shift[3 * p + 1]) : 0.0;
            s2 = (pbc_mask & 4) ? static_cast<double>(shift[3 * p + 2]) : 0.0;
            /* j-side: the block's table, or global memory when full */
            const unsigned hsh = static_cast<unsigned>(j) * 2654435761u;
            int slot = -1;
            for (int q = 0; q < kHashProbe; q++) {
                const int k = static_cast<int>((hsh >> 22) + q) & (kHashSlots - 1);
                const int old = atomicCAS(&s_key[k], -1, j);
                if (old == -1 || old == j) {
                    slot = k;
                    break;
                }
            }
            if (slot >= 0) {
                atomicAdd(&s_g[0][slot], g0);
                atomicAdd(&s_g[1][slot], g1);
                atomicAdd(&s_g[2][slot], g2);
            } else {
                unsafeAtomicAdd(&gx[j], g0);
                unsafeAtomicAdd(&gy[j], g1);
                unsafeAtomicAdd(&gz[j], g2);
            }
        }
        c[0] += s0 * g0; c[1] += s0 * g1; c[2] += s0 * g2;
        c[3] += s1 * g0; c[4] += s1 * g1; c[5] += s1 * g2;
        c[6] += s2 * g0; c[7] += s2 * g1; c[8] += s2 * g2;
        double a0 = g0, a1 = g1, a2 = g2;
        seg_incl_sum3(a0, a1, a2, i);
        const int next = __shfl_down(i, 1, kWave);
        const bool tail = lane_id() == kWave - 1 || next != i;
        if (valid && tail) {
            unsafeAtomicAdd(&gx[i], -a0);
            unsafeAtomicAdd(&gy[i], -a1);
            unsafeAtomicAdd(&gz[i], -a2);
        }
    }
    __syncthreads();
    flush_table_sorted(s_key, s_g, gx, gy, gz);
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = wave_sum(c[k]);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < 9; k++) s_red[w][k] = c[k];
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        double t = 0.0;
        for (int q = 0; q < kWavesPerBlock; q++) t += s_red[q][threadIdx.x];
        cell_partial[static_cast<int64_t>(blockIdx.x) * 9 + threadIdx.x] = t;
    }
}

/* grad_pos[a] = (gx[a], gy[a], gz[a]): the per-component sums, accumulated
   in separate arrays so that a wave's atomics on consecutive atoms land in
   few cache lines, interleaved into the caller's [nat][3] layout. */
__global__ __launch_bounds__(kBlock) void k_soa_to_aos(const double *__restrict__ soa,
                                                       int32_t nat, double *__restrict__ out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 3 * static_cast<int64_t>(nat)) return;
    const int64_t a = t / 3, k = t - 3 * a;
    out[t] = soa[k * nat + a];
}

/* Sums of per-block partials (rows of ncomp components) in a fixed order:
   block g adds rows [g * per, min(nb, (g + 1) * per)) into row g of
   grad_cell, each thread a strided subset, then a fixed-shape tree over
   the block. One block with per >= nb sums all rows. */
__global__ __launch_bounds__(kBlock) void k_cell_reduce(const double *__restrict__ partial,
                                                        int nb, double *__restrict__ grad_cell,
                                                        int ncomp = 9, int per = INT32_MAX) {
    __shared__ double s_sum[kBlock];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * per;
    const int64_t r1 = min(static_cast<int64_t>(nb), r0 + per);
    for (int k = 0; k < ncomp; k++) {
        double t = 0.0;
        for (int64_t b = r0 + threadIdx.x; b < r1; b += kBlock)
            t += partial[b * ncomp + k];
        s_sum[threadIdx.x] = t;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (static_cast<int>(threadIdx.x) < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) grad_cell[static_cast<int64_t>(blockIdx.x) * ncomp + k] = s_sum[0];
        __syncthreads();
    }
}

__global__ void k_fill_f64(double *p, int64_t n, double v) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_system_offsets(const int64_t *__restrict__ offsets,
                                 const int32_t *__restrict__ atom_base, int32_t nsys,
                                 int64_t *__restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= nsys) out[s] = offsets[atom_base[s]];
}

__global__ void k_sqrt_probe(const double *__restrict__ x, double *__restrict__ y, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) y[t] = sqrt(x[t]);
}

}  // namespace

/* ---------------------------------------------------------------- launchers */

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_fill_i32<<<grid_for(n), kBlock, 0, stream>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_i64(int64_t *p, int64_t n, int64_t v, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_fill_i64<<<grid_for(n), kBlock, 0, stream>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_bin_dense(const double *r, int32_t nat, const SysGeom *geo,
                            const int32_t *atom_base, int32_t nsys, int4 *raw4,
                            int32_t *ticket, int32_t *cell_count, int32_t *order_stats,
                            hipStream_t stream) {
    if (nat <= 0) return hipSuccess;
    k_bin_dense<<<grid_for(nat), kBlock, 0, stream>>>(r, nat, geo, atom_base, nsys, raw4,
                                                      ticket, cell_count, order_stats);
    return hipGetLastError();
}

hipError_t launch_bin_hashed(const double *r, int32_t nat, const SysGeom *geo, int4 *raw4,
                             int32_t *ticket, int64_t *hkey, int64_t hmask,
                             int32_t *hcount, hipStream_t stream) {
    if (nat <= 0) return hipSuccess;
    k_bin_hashed<<<grid_for(nat), kBlock, 0, stream>>>(r, nat, geo, raw4, ticket, hkey,
                                                       hmask, hcount);
    return hipGetLastError();
}

hipError_t launch_scatter(const int4 *raw4, const int32_t *ticket, int32_t nat,
                          const int32_t *cell_first, int32_t *tmp, hipStream_t stream) {
    if (nat <= 0) return hipSuccess;
    k_scatter<<<grid_for(nat), kBlock, 0, stream>>>(raw4, ticket, nat, cell_first, tmp);
    return hipGetLastError();
}

hipError_t launch_rank_gather(const double *r, const int4 *raw4, const int32_t *tmp,
                              int32_t nat, const int32_t *cell_first,
                              const int32_t *cell_count, const SysGeom *geo,
                              const int32_t *atom_base, int32_t nsys,
                              const double *per_atom, const int32_t *types, double4 *spos,
                              int4 *sinfo, int32_t *stype, int32_t *scount, int32_t *slot,
                              int32_t *first_flag, hipStream_t stream) {
    if (nat <= 0) return hipSuccess;
    k_rank_gather<<<grid_for(nat), kBlock, 0, stream>>>(r, raw4, tmp, nat, cell_first,
                                                        cell_count, geo, atom_base, nsys,
                                                        per_atom, types, spos, sinfo, stype,
                                                        scount, slot, first_flag);
    return hipGetLastError();
}

hipError_t launch_home_list(const int4 *raw4, const int32_t *flag, const int32_t *pos,
                            int32_t nat, int32_t *home_list, hipStream_t stream) {
    if (nat <= 0) return hipSuccess;
    k_home_list<<<grid_for(nat), kBlock, 0, stream>>>(raw4, flag, pos, nat, home_list);
    return hipGetLastError();
}

int64_t scan_partial_size(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

int64_t scan_status_words() { return kScanOnepassTiles + 2; }

template <typename TI, typename TO>
static hipError_t scan_impl(const TI *in, int64_t n, TO *out, int64_t *partial,
                            int64_t *total, int64_t *total2, bool write_total,
                            hipStream_t stream, const int32_t *gsrc = nullptr,
                            const int32_t *gslot = nullptr, uint64_t *status = nullptr) {
    if (n <= 0) {
        if (write_total) {
            /* out[0] = 0 */
            if (sizeof(TO) == 8)
                k_fill_i64<<<1, kBlock, 0, stream>>>(reinterpret_cast<int64_t *>(out), 1, 0);
            else
                k_fill_i32<<<1, kBlock, 0, stream>>>(reinterpret_cast<int32_t *>(out), 1, 0);
        }
        if (total) k_fill_i64<<<1, kBlock, 0, stream>>>(total, 1, 0);
        if (total2) k_fill_i64<<<1, kBlock, 0, stream>>>(total2, 1, 0);
        return hipGetLastError();
    }
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    int32_t *gout = reinterpret_cast<int32_t *>(const_cast<TI *>(in));
    if (nb == 1) {
        /* one tile: the apply alone (its block has no tiles before it), with
           the gather folded in */
        k_scan_apply<TI, TO><<<1, kBlock, 0, stream>>>(in, n, partial, out,
                                                       write_total ? 1 : 0, 1, total, total2,
                                                       gslot ? gsrc : nullptr, gslot,
                                                       gslot ? gout : nullptr);
        return hipGetLastError();
    }
    if (status && nb <= kScanOnepassTiles) {
        k_scan_onepass<TI, TO><<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(
            in, n, out, write_total ? 1 : 0, total, total2, gslot ? gsrc : nullptr, gslot,
            gslot ? gout : nullptr, status, static_cast<int>(nb));
        return hipGetLastError();
    }
    if (gslot)
        k_scan_reduce_gather<<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(gsrc, gslot, gout,
                                                                               n, partial);
    else
        k_scan_reduce<TI><<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(in, n, partial);
    /* up to this many tiles, each apply block adds up the tile sums before
       it (no spine kernel) */
    constexpr int64_t kFusedTiles = 1024;
    const bool fused = nb <= kFusedTiles;
    if (!fused) k_scan_spine<<<1, 1024, 0, stream>>>(partial, nb, total, total2);
    k_scan_apply<TI, TO><<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(
        in, n, partial, out, write_total ? 1 : 0, fused ? 1 : 0, fused ? total : nullptr,
        fused ? total2 : nullptr, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_scan_i32_to_i32(const int32_t *in, int64_t n, int32_t *out,
                                  int64_t *partial, int64_t *total, hipStream_t stream,
                                  uint64_t *status) {
    return scan_impl<int32_t, int32_t>(in, n, out, partial, total, nullptr, false, stream,
                                       nullptr, nullptr, status);
}

hipError_t launch_scan_i32_to_i64(const int32_t *in, int64_t n, int64_t *out,
                                  int64_t *partial, int64_t *total, hipStream_t stream,
                                  int64_t *total2, uint64_t *status) {
    return scan_impl<int32_t, int64_t>(in, n, out, partial, total, total2, true, stream,
                                       nullptr, nullptr, status);
}

hipError_t launch_scan_sorted_counts(const int32_t *scount, const int32_t *slot,
                                     int32_t *counts, int64_t n, int64_t *out,
                                     int64_t *partial, int64_t *total, hipStream_t stream,
                                     int64_t *total2, uint64_t *status) {
    return scan_impl<int32_t, int64_t>(counts, n, out, partial, total, total2, true, stream,
                                       scount, slot, status);
}

hipError_t launch_scan_i64_to_i64(const int64_t *in, int64_t n, int64_t *out,
                                  int64_t *partial, int64_t *total, hipStream_t stream) {
    return scan_impl<int64_t, int64_t>(in, n, out, partial, total, nullptr, true, stream);
}

template <int CUT, bool HASHED>
static void launch_count_t(const PairArgs &a, unsigned nblocks, hipStream_t stream) {
    if (a.half) {
        /* per-atom counts stay in cell-sorted order: the caller scans them
           with launch_scan_sorted_counts */
        k_count<CUT, HASHED, true><<<nblocks, kBlock, 0, stream>>>(a);
    } else {
        k_count<CUT, HASHED, false><<<nblocks, kBlock, 0, stream>>>(a);
    }
}

/* Windows of 5 chunks when the host expects 257-320 candidates per home
   cell (one window then covers a typical cell instead of a 4-chunk window
   and a second sweep over the home atoms for the rest); 4 otherwise. */
template <int MODE, int CUT, bool HASHED>
static void launch_fill_t(const PairArgs &a, unsigned nblocks, hipStream_t stream) {
    if (a.fill_chunks == 5)
        k_fill<MODE, CUT, HASHED, 5><<<nblocks, kBlock, 0, stream>>>(a);
    else
        k_fill<MODE, CUT, HASHED, kFillChunks><<<nblocks, kBlock, 0, stream>>>(a);
}

template <int CUT>
static void launch_pairs_h(int mode, bool hashed, const PairArgs &a, unsigned nb,
                           hipStream_t s) {
    if (mode == MODE_COUNT) {
        if (hashed)
            launch_count_t<CUT, true>(a, nb, s);
        else
            launch_count_t<CUT, false>(a, nb, s);
    } else if (mode == MODE_FILL64) {
        if (hashed)
            launch_fill_t<MODE_FILL64, CUT, true>(a, nb, s);
        else
            launch_fill_t<MODE_FILL64, CUT, false>(a, nb, s);
    } else if (mode == MODE_FILL32) {
        if (hashed)
            launch_fill_t<MODE_FILL32, CUT, true>(a, nb, s);
        else
            launch_fill_t<MODE_FILL32, CUT, false>(a, nb, s);
    } else {
        if (hashed)
            launch_fill_t<MODE_MATRIX, CUT, true>(a, nb, s);
        else
            launch_fill_t<MODE_MATRIX, CUT, false>(a, nb, s);
    }
}

/* Waves walking a home list: at most one per listed cell and part (the
   list holds at most min(slots, atoms) cells), and no more than a few
   rounds of the chip's wave slots. */
constexpr int64_t kListWaves = 16384;

int64_t pair_waves(const PairArgs &a) {
    if (a.item_base) return a.items;
    const int64_t split = std::max(a.split, 1);
    if (!a.home_list) return static_cast<int64_t>(a.home_cells) * split;
    const int64_t cells = std::min<int64_t>(a.home_cells, std::max(a.nat, 0));
    return std::min(cells * split, kListWaves);
}

hipError_t launch_pairs(int mode, int cut, bool hashed, const PairArgs &args,
                        hipStream_t stream) {
    const int64_t waves = pair_waves(args);
    if (waves <= 0) return hipSuccess;
    const unsigned nb = static_cast<unsigned>((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    PairArgs a = args;
    a.xcd_blocks = static_cast<int32_t>(nb);
    if (cut == CUT_PER_ATOM)
        launch_pairs_h<CUT_PER_ATOM>(mode, hashed, a, nb, stream);
    else if (cut == CUT_PER_TYPE && a.ncut * a.ncut <= kLdsTypes)
        launch_pairs_h<CUT_PER_TYPE_LDS>(mode, hashed, a, nb, stream);
    else if (cut == CUT_PER_TYPE)
        launch_pairs_h<CUT_PER_TYPE>(mode, hashed, a, nb, stream);
    else
        launch_pairs_h<CUT_GLOBAL>(mode, hashed, a, nb, stream);
    return hipGetLastError();
}

bool batch_binning_fits(int64_t max_atoms, int64_t max_cells) {
    return max_atoms <= kSysAtoms && max_cells <= kSysCells;
}

hipError_t launch_bin_systems(const double *r, int32_t nsys, const SysGeom *geo,
                              int32_t *cell_first, int32_t *cell_count, double4 *spos,
                              int4 *sinfo, int32_t *scount, int32_t *slot, hipStream_t stream) {
    if (nsys <= 0) return hipSuccess;
    k_bin_systems<<<static_cast<unsigned>(nsys), kSysThreads, 0, stream>>>(
        r, geo, cell_first, cell_count, spos, sinfo, scount, slot);
    return hipGetLastError();
}

bool small_binning_fits(int64_t nat, int64_t ncells) {
    return nat > 0 && nat <= kSmallAtoms && ncells > 0 && ncells <= kSmallCells;
}

hipError_t launch_bin_small(const double *r, int32_t nat, const SysGeom &g, int32_t ncells,
                            SysGeom *geo, int32_t *atom_base, int32_t *cell_base,
                            int32_t *order_stats, const double *per_atom, const int32_t *types,
                            int32_t *cell_first, int32_t *cell_count, double4 *spos,
                            int4 *sinfo, int32_t *stype, int32_t *scount, int32_t *slot,
                            hipStream_t stream) {
    if (!small_binning_fits(nat, ncells)) return hipErrorInvalidValue;
    k_bin_small<<<1, kSmallThreads, 0, stream>>>(r, nat, g, ncells, geo, atom_base, cell_base,
                                                 order_stats, per_atom, types, cell_first,
                                                 cell_count, spos, sinfo, stype, scount, slot);
    return hipGetLastError();
}

hipError_t launch_set_params(SysGeom *geo, const SysGeom &g, int32_t *atom_base,
                             int32_t *cell_base, int32_t nat, int32_t ncells,
                             int32_t *cell_count, int64_t nslots, int64_t *hkey,
                             int32_t *order_stats, hipStream_t stream) {
    k_set_params<<<grid_for(std::max<int64_t>(nslots, 1)), kBlock, 0, stream>>>(
        geo, g, atom_base, cell_base, nat, ncells, cell_count, nslots, hkey, order_stats);
    return hipGetLastError();
}

hipError_t launch_system_offsets(const int64_t *offsets, const int32_t *atom_base,
                                 int32_t nsys, int64_t *out, hipStream_t stream) {
    if (nsys < 0) return hipSuccess;
    k_system_offsets<<<grid_for(int64_t(nsys) + 1), kBlock, 0, stream>>>(offsets, atom_base,
                                                                         nsys, out);
    return hipGetLastError();
}

/* Tiles of at least 8192 pairs (a block's table then collects many pairs
   per atom), at most kBackwardMaxBlocks of them. */
int backward_blocks(int64_t npairs) {
    const int64_t want = (npairs + 8191) / 8192;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(want, kBackwardMaxBlocks)));
}

static int64_t backward_tile(int64_t npairs, int nb) {
    const int64_t t = (npairs + nb - 1) / nb;
    return (t + kWave - 1) / kWave * kWave;
}

hipError_t launch_pair_backward(int64_t npairs, int32_t nat, const int32_t *first,
                                const int32_t *secnd, const int32_t *shift, const void *D,
                                const void *d, const void *gD, const void *gd, bool f32,
                                int pbc_mask, double *grad_pos, double *grad_cell,
                                double *cell_partial, hipStream_t stream) {
    const int nb = backward_blocks(npairs);
    /* per-component sums after the cell partials */
    double *soa = cell_partial + 9 * static_cast<int64_t>(kBackwardMaxBlocks);
    const int64_t n3 = int64_t(nat) * 3;
    if (nat > 0) k_fill_f64<<<grid_for(n3), kBlock, 0, stream>>>(soa, n3, 0.0);
    if (npairs > 0) {
        if (f32)
            k_pair_backward<float><<<nb, kBlock, 0, stream>>>(
                npairs, backward_tile(npairs, nb), first, secnd, shift, static_cast<const float *>(D),
                static_cast<const float *>(d), static_cast<const float *>(gD),
                static_cast<const float *>(gd), pbc_mask, soa, soa + nat, soa + 2 * int64_t(nat),
                cell_partial);
        else
            k_pair_backward<double><<<nb, kBlock, 0, stream>>>(
                npairs, backward_tile(npairs, nb), first, secnd, shift, static_cast<const double *>(D),
                static_cast<const double *>(d), static_cast<const double *>(gD),
                static_cast<const double *>(gd), pbc_mask, soa, soa + nat, soa + 2 * int64_t(nat),
                cell_partial);
        k_cell_reduce<<<1, kBlock, 0, stream>>>(cell_partial, nb, grad_cell);
    } else {
        k_fill_f64<<<1, kWave, 0, stream>>>(grad_cell, 9, 0.0);
    }
    if (nat > 0) k_soa_to_aos<<<grid_for(n3), kBlock, 0, stream>>>(soa, nat, grad_pos);
    return hipGetLastError();
}

/* First level of the energy/virial reduction: this many blocks, each over a
   contiguous range of the per-block partials. */
constexpr int kLjReduceBlocks = 64;

int64_t lj_scratch_doubles(const PairArgs &a) {
    const int64_t waves = pair_waves(a);
    const int64_t nb = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
    return 3 * static_cast<int64_t>(std::max(a.nat, 1)) + 10 * std::max<int64_t>(nb, 1) +
           10 * kLjReduceBlocks + 10;
}

hipError_t launch_lj(const PairArgs &args, bool hashed, double epsilon, double sigma,
                     const int32_t *slot, double *scratch, double *forces, double *energy,
                     double *virial, hipStream_t stream) {
    const int32_t nat = args.nat;
    const int64_t waves = pair_waves(args);
    const int64_t nb = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
    double *fx = scratch, *fy = scratch + nat, *fz = scratch + 2 * int64_t(nat);
    double *partial = scratch + 3 * int64_t(nat);
    double *partial2 = partial + 10 * std::max<int64_t>(nb, 1);
    double *sums = partial2 + 10 * kLjReduceBlocks;
    if (nat > 0) k_fill_f64<<<grid_for(3 * int64_t(nat)), kBlock, 0, stream>>>(scratch,
                                                                              3 * int64_t(nat), 0.0);
    if (nb > 0) {
        PairArgs a = args;
        a.xcd_blocks = static_cast<int32_t>(nb);
        const double eps4 = 4.0 * epsilon, sig2 = sigma * sigma;
        if (hashed)
            k_lj<true><<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(a, eps4, sig2, fx, fy, fz,
                                                                         partial);
        else
            k_lj<false><<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(a, eps4, sig2, fx, fy, fz,
                                                                          partial);
        const int per = static_cast<int>((nb + kLjReduceBlocks - 1) / kLjReduceBlocks);
        k_cell_reduce<<<kLjReduceBlocks, kBlock, 0, stream>>>(partial, static_cast<int>(nb),
                                                              partial2, 10, per);
        k_cell_reduce<<<1, kBlock, 0, stream>>>(partial2, kLjReduceBlocks, sums, 10);
    } else {
        k_fill_f64<<<1, kWave, 0, stream>>>(sums, 10, 0.0);
    }
    k_lj_finish<<<grid_for(std::max<int64_t>(nat, 1)), kBlock, 0, stream>>>(
        fx, fy, fz, slot, nat, sums, forces, energy, virial);
    return hipGetLastError();
}

hipError_t launch_sqrt_probe(const double *x, double *y, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_sqrt_probe<<<grid_for(n), kBlock, 0, stream>>>(x, y, n);
    return hipGetLastError();
}

}  // namespace mnl

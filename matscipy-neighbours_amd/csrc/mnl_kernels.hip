/*
 * mnl HIP kernels for gfx950 (CDNA4, wave64).
 *
 * Pipeline of one build (plain path = a batch of one system):
 *   bin         raw cell per atom, fold, per-cell histogram with an atomic
 *               ticket                                     (1 thread / atom)
 *   scan        cell counts -> cell slice starts           (3-phase scan)
 *   scatter     atoms into their cell slice at the ticket  (1 thread / atom)
 *   rank+gather stable order inside each cell (ascending atom index, the
 *               reference's serial counting sort) + gather of the
 *               cell-sorted per-atom records               (1 thread / atom)
 *   count       neighbours per atom                        (1 wave / cell)
 *   scan        counts -> int64 pair offsets, total on device
 *   fill        i, j, D, d, S at the offsets               (1 wave / cell)
 *
 * Floating point: compiled with contraction off and the reference's
 * operation order everywhere a double is produced, so the cell index, the
 * bin-relative positions, D and d are bit-identical to the reference CPU
 * path (src/libneighbours/neighbour_visit.hh:63-115,
 * neighbour_list.cc:126-167, tools.hh:30-67).
 */
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mnl_engine.hh"

#pragma clang fp contract(off)

namespace mnl {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kScanItems = 16;
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

/* XCD-aware remap: the hardware deals workgroups round-robin to the 8 XCDs;
   hand each XCD a contiguous range of logical blocks so that neighbouring
   cells share one L2. A bijection for any grid size. */
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    constexpr int kXcd = 8;
    const int q = nblocks / kXcd, r = nblocks % kXcd;
    const int x = b % kXcd, k = b / kXcd;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
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

__global__ __launch_bounds__(kBlock) void k_bin_dense(
    const double *__restrict__ r, int32_t nat, const SysGeom *__restrict__ geo,
    const int32_t *__restrict__ atom_base, int32_t nsys, int4 *__restrict__ raw4,
    int32_t *__restrict__ ticket, int32_t *__restrict__ cell_count) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nat) return;
    const SysGeom &g = geo[system_of(atom_base, nsys, a)];
    int raw[3];
    raw_cell_of(r, a, g, raw);
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; k++)
        c[k] = g.pbc[k] ? wrap_idx(raw[k], g.n[k]) : clamp_idx(raw[k], g.n[k]);
    const int cell = g.cell_base + c[0] + g.n[0] * (c[1] + g.n[1] * c[2]);
    ticket[a] = atomicAdd(&cell_count[cell], 1);
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
    if (a >= nat) return;
    const SysGeom &g = geo[0];
    int raw[3];
    raw_cell_of(r, a, g, raw);
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; k++)
        c[k] = g.pbc[k] ? wrap_idx(raw[k], g.n[k]) : clamp_idx(raw[k], g.n[k]);
    const int64_t key = static_cast<int64_t>(c[0]) +
                        static_cast<int64_t>(g.n[0]) *
                            (static_cast<int64_t>(c[1]) + static_cast<int64_t>(g.n[1]) * c[2]);
    int64_t h = cell_hash(key) & hmask;
    auto *slots = reinterpret_cast<unsigned long long *>(hkey);
    const unsigned long long empty = ~0ull;
    while (true) {
        const unsigned long long old =
            atomicCAS(&slots[h], empty, static_cast<unsigned long long>(key));
        if (old == empty || static_cast<int64_t>(old) == key) break;
        h = (h + 1) & hmask;
    }
    ticket[a] = atomicAdd(&hcount[h], 1);
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
   ((r - rel1 b1) - rel2 b2) - rel3 b3. The per-atom image count feeds the
   shift: periodic img = (raw - wrap(raw)) / n (exact), non-periodic
   img = trunc((raw - clamp(raw)) / n). */
__global__ __launch_bounds__(kBlock) void k_rank_gather(
    const double *__restrict__ r, const int4 *__restrict__ raw4,
    const int32_t *__restrict__ tmp, int32_t nat, const int32_t *__restrict__ cell_first,
    const int32_t *__restrict__ cell_count, const SysGeom *__restrict__ geo,
    const int32_t *__restrict__ atom_base, int32_t nsys,
    const double *__restrict__ per_atom, const int32_t *__restrict__ types,
    double4 *__restrict__ spos, int4 *__restrict__ sinfo, int32_t *__restrict__ stype) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nat) return;
    const int4 rc = raw4[a];
    const int b = cell_first[rc.w];
    const int m = cell_count[rc.w];
    int rank = 0;
    for (int t = 0; t < m; t++) rank += tmp[b + t] < a ? 1 : 0;
    const int s = b + rank;

    const int sys = system_of(atom_base, nsys, a);
    const SysGeom &g = geo[sys];
    const int raw[3] = {rc.x, rc.y, rc.z};
    int rel[3], img[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (g.pbc[k]) {
            rel[k] = raw[k];
            img[k] = (raw[k] - wrap_idx(raw[k], g.n[k])) / g.n[k];
        } else {
            rel[k] = clamp_idx(raw[k], g.n[k]);
            img[k] = (raw[k] - rel[k]) / g.n[k];
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
    sinfo[s] = make_int4(img[0], img[1], img[2], a - g.atom_base);
    if (stype) stype[s] = types ? types[a] : 0;
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

/* Single block: exclusive scan of the nb block sums in place; total out. */
__global__ __launch_bounds__(1024) void k_scan_spine(int64_t *__restrict__ partial, int64_t nb,
                                                     int64_t *__restrict__ total) {
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
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const TI *__restrict__ in, int64_t n,
                                                        const int64_t *__restrict__ partial,
                                                        TO *__restrict__ out, int write_total) {
    __shared__ int64_t s_tile[kScanTile];
    __shared__ int64_t s_w[kWavesPerBlock];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t idx = base + k * kBlock + threadIdx.x;
        s_tile[k * kBlock + threadIdx.x] = idx < n ? static_cast<int64_t>(in[idx]) : 0;
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
    int64_t pre = partial[blockIdx.x];
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
        const int64_t last = n - 1;
        const int64_t tot = s_tile[last - base] + static_cast<int64_t>(in[last]);
        out[n] = static_cast<TO>(tot);
    }
}

/* ---------------------------------------------------------------- pairs */

enum { MODE_COUNT = 0, MODE_FILL64 = 1, MODE_FILL32 = 2 };

/* Pack three small signed integers (|v| < 512) into 30 bits and back. */
__device__ __forceinline__ int pack3(int a, int b, int c) {
    return ((a + 512) & 1023) | (((b + 512) & 1023) << 10) | (((c + 512) & 1023) << 20);
}
__device__ __forceinline__ void unpack3(int p, int &a, int &b, int &c) {
    a = (p & 1023) - 512;
    b = ((p >> 10) & 1023) - 512;
    c = ((p >> 20) & 1023) - 512;
}

/* Dense slice of cell (c0,c1,c2) of a system (cell_list.hh:94-104). */
__device__ __forceinline__ void dense_slice(const PairArgs &A, int cell_base, int n0, int n1,
                                            int c0, int c1, int c2, int &b, int &len) {
    const int c = cell_base + c0 + n0 * (c1 + n1 * c2);
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

/*
 * One wave per home cell. The home cell's neighbour cells, in the
 * reference's traversal order (z outer, y, x inner, each -n..n;
 * neighbour_visit.hh:63-84), are concatenated into one candidate index
 * space: lanes take 64 consecutive candidates (coalesced loads, kept in
 * registers) and the wave then loops over the home atoms, testing every
 * candidate against one home atom per step. A ballot gives the accepted
 * lanes; for the fill, a lane's output slot is the home atom's cursor plus
 * the accepted lanes below it, so accepted pairs of one atom land in
 * consecutive slots in candidate order - the reference's output order - and
 * consecutive lanes write consecutive addresses.
 *
 * Home-atom state (count or write cursor) lives in lane t for home atom t;
 * home cells with more than 64 atoms are processed in groups of 64.
 *
 * The shift S = (raw_i - rel_j + x) / n (neighbour_visit.hh:111-114) is
 * computed without integer division: periodic directions give
 * S = w + img_i - img_j, with w = floor((c_i + x) / n) the wrap count of the
 * neighbour cell; non-periodic directions give S = img_i (trunc), since there
 * rel_j = c_i + x.
 */
template <int MODE, int CUT, bool HASHED>
__global__ __launch_bounds__(kBlock) void k_pairs(PairArgs A) {
    __shared__ int s_start[kWavesPerBlock][kWave];
    __shared__ int s_begin[kWavesPerBlock][kWave];
    __shared__ int s_xyz[kWavesPerBlock][kWave];
    __shared__ int s_wrap[kWavesPerBlock][kWave];

    const int lane = lane_id();
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int blk = xcd_remap(blockIdx.x, A.xcd_blocks);
    const int h = blk * kWavesPerBlock + wib;
    if (h >= A.home_cells) return;

    const int m = A.cell_count[h];
    if (m == 0) return;
    const int hb = A.cell_first[h];

    int sys = 0;
    int c0, c1, c2;
    if (!HASHED) {
        sys = system_of(A.cell_base, A.nsys, h);
    }
    const SysGeom &g = A.geo[sys];
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    if (!HASHED) {
        const int lin = h - g.cell_base;
        c0 = lin % n0;
        c1 = (lin / n0) % n1;
        c2 = lin / (n0 * n1);
    } else {
        const int64_t key = A.hkey[h];
        c0 = static_cast<int>(key % n0);
        c1 = static_cast<int>((key / n0) % n1);
        c2 = static_cast<int>(key / (static_cast<int64_t>(n0) * n1));
    }
    const int nx = g.nxyz[0], ny = g.nxyz[1], nz = g.nxyz[2];
    const int sx = 2 * nx + 1, sy = 2 * ny + 1;
    const int nseg = g.nseg;
    const int p0 = g.pbc[0], p1 = g.pbc[1], p2 = g.pbc[2];
    const double cutoff_sq = g.cutoff_sq;
    const double b00 = g.bin[0], b01 = g.bin[1], b02 = g.bin[2];
    const double b10 = g.bin[3], b11 = g.bin[4], b12 = g.bin[5];
    const double b20 = g.bin[6], b21 = g.bin[7], b22 = g.bin[8];
    const int abase = g.atom_base;

    int *const st = s_start[wib];
    int *const sb = s_begin[wib];
    int *const sxyz = s_xyz[wib];
    int *const swr = s_wrap[wib];

    for (int grp0 = 0; grp0 < m; grp0 += kWave) {
        const int mg = min(kWave, m - grp0);
        int acc = 0;
        int64_t cursor = 0;
        if (MODE != MODE_COUNT && lane < mg)
            cursor = A.offsets[abase + A.sinfo[hb + grp0 + lane].w];

        for (int sg = 0; sg < nseg; sg += kWave) {
            const int nsg = min(kWave, nseg - sg);
            int len = 0, begin = 0, pxyz = 0, pw = 0;
            if (lane < nsg) {
                const int seg = sg + lane;
                const int x = seg % sx - nx;
                const int rest = seg / sx;
                const int y = rest % sy - ny;
                const int z = rest / sy - nz;
                int cj0 = c0 + x, cj1 = c1 + y, cj2 = c2 + z;
                int w0 = 0, w1 = 0, w2 = 0;
                bool ok = true;
                if (p0) {
                    w0 = floor_div(cj0, n0);
                    cj0 -= w0 * n0;
                } else {
                    ok = ok && cj0 >= 0 && cj0 < n0;
                }
                if (p1) {
                    w1 = floor_div(cj1, n1);
                    cj1 -= w1 * n1;
                } else {
                    ok = ok && cj1 >= 0 && cj1 < n1;
                }
                if (p2) {
                    w2 = floor_div(cj2, n2);
                    cj2 -= w2 * n2;
                } else {
                    ok = ok && cj2 >= 0 && cj2 < n2;
                }
                if (ok) {
                    if (HASHED)
                        hashed_slice(A, n0, n1, cj0, cj1, cj2, begin, len);
                    else
                        dense_slice(A, g.cell_base, n0, n1, cj0, cj1, cj2, begin, len);
                }
                pxyz = pack3(x, y, z);
                pw = pack3(w0, w1, w2);
            }
            const int incl = wave_incl_scan(len);
            const int total = __shfl(incl, kWave - 1, kWave);
            st[lane] = incl - len;
            sb[lane] = begin;
            sxyz[lane] = pxyz;
            swr[lane] = pw;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

            for (int cb = 0; cb < total; cb += kWave) {
                const int c = cb + lane;
                const bool cv = c < total;
                /* segment holding candidate c: the last one starting at or
                   before c (empty segments never win: the covering one comes
                   later in order) */
                int p = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1)
                    if (p + step < nsg && st[p + step] <= c) p += step;
                const int sj = cv ? sb[p] + (c - st[p]) : hb;
                int x, y, z, w0, w1, w2;
                unpack3(sxyz[p], x, y, z);
                unpack3(swr[p], w0, w1, w2);
                /* off = (z b3 + y b2) + x b1 (neighbour_visit.hh:67,73-74,80-81) */
                const double o0 = (z * b20 + y * b10) + x * b00;
                const double o1 = (z * b21 + y * b11) + x * b01;
                const double o2 = (z * b22 + y * b12) + x * b02;
                const double4 pj = A.spos[sj];
                const int4 ij = A.sinfo[sj];
                const bool centre = (x == 0) && (y == 0) && (z == 0);
                const int tj = CUT == CUT_PER_TYPE ? A.stype[sj] : 0;
                const int ds0 = p0 ? w0 - ij.x : 0;
                const int ds1 = p1 ? w1 - ij.y : 0;
                const int ds2 = p2 ? w2 - ij.z : 0;

                for (int t = 0; t < mg; t++) {
                    const int si = hb + grp0 + t;
                    const double4 pi = A.spos[si];
                    /* dr = (rel_j - rel_i) + off; r2 = (dx^2 + dy^2) + dz^2
                       (neighbour_visit.hh:89-94) */
                    const double d0 = (pj.x - pi.x) + o0;
                    const double d1 = (pj.y - pi.y) + o1;
                    const double d2 = (pj.z - pi.z) + o2;
                    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                    bool ok = cv && (r2 < cutoff_sq) && !(centre && sj == si);
                    if (CUT == CUT_PER_ATOM) {
                        const double cc = pi.w + pj.w;
                        ok = ok && (r2 < cc * cc);
                    } else if (CUT == CUT_PER_TYPE) {
                        const int ti = A.stype[si];
                        if (ti >= 0 && ti < A.ncut && tj >= 0 && tj < A.ncut)
                            ok = ok && (r2 < A.pt_sq[ti * A.ncut + tj]);
                    }
                    const uint64_t mask = __ballot(ok);
                    const int cnt = __popcll(mask);
                    if (MODE == MODE_COUNT) {
                        acc += (lane == t) ? cnt : 0;
                    } else if (mask) {
                        const int64_t base = __shfl(cursor, t, kWave);
                        if (ok) {
                            const int64_t pos = base + mbcnt64(mask);
                            if (pos < A.capacity) {
                                const int4 ii = A.sinfo[si];
                                if (A.out_i) A.out_i[pos] = ii.w;
                                if (A.out_j) A.out_j[pos] = ij.w;
                                if (MODE == MODE_FILL64) {
                                    if (A.out_D) {
                                        double *D = static_cast<double *>(A.out_D) + 3 * pos;
                                        D[0] = d0;
                                        D[1] = d1;
                                        D[2] = d2;
                                    }
                                    if (A.out_d) static_cast<double *>(A.out_d)[pos] = sqrt(r2);
                                } else {
                                    if (A.out_D) {
                                        float *D = static_cast<float *>(A.out_D) + 3 * pos;
                                        D[0] = static_cast<float>(d0);
                                        D[1] = static_cast<float>(d1);
                                        D[2] = static_cast<float>(d2);
                                    }
                                    if (A.out_d)
                                        static_cast<float *>(A.out_d)[pos] =
                                            static_cast<float>(sqrt(r2));
                                }
                                if (A.out_S) {
                                    int32_t *S = A.out_S + 3 * pos;
                                    S[0] = ii.x + ds0;
                                    S[1] = ii.y + ds1;
                                    S[2] = ii.z + ds2;
                                }
                            }
                        }
                        if (lane == t) cursor += cnt;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (MODE == MODE_COUNT && lane < mg)
            A.counts[abase + A.sinfo[hb + grp0 + lane].w] = acc;
    }
}

/* Device-side parameter set-up (one thread): keeps the whole build free of
   host-to-device copies, so that it can be captured in a graph. */
__global__ void k_set_params(SysGeom *__restrict__ geo, SysGeom g, int32_t *__restrict__ atom_base,
                             int32_t *__restrict__ cell_base, int32_t nat, int32_t ncells) {
    if (threadIdx.x != 0) return;
    geo[0] = g;
    atom_base[0] = 0;
    atom_base[1] = nat;
    cell_base[0] = 0;
    cell_base[1] = ncells;
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
                            int32_t *ticket, int32_t *cell_count, hipStream_t stream) {
    if (nat <= 0) return hipSuccess;
    k_bin_dense<<<grid_for(nat), kBlock, 0, stream>>>(r, nat, geo, atom_base, nsys, raw4,
                                                      ticket, cell_count);
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
                              int4 *sinfo, int32_t *stype, hipStream_t stream) {
    if (nat <= 0) return hipSuccess;
    k_rank_gather<<<grid_for(nat), kBlock, 0, stream>>>(r, raw4, tmp, nat, cell_first,
                                                        cell_count, geo, atom_base, nsys,
                                                        per_atom, types, spos, sinfo, stype);
    return hipGetLastError();
}

int64_t scan_partial_size(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

template <typename TO>
static hipError_t scan_impl(const int32_t *in, int64_t n, TO *out, int64_t *partial,
                            int64_t *total, bool write_total, hipStream_t stream) {
    if (n <= 0) {
        if (write_total) {
            /* out[0] = 0 */
            if (sizeof(TO) == 8)
                k_fill_i64<<<1, kBlock, 0, stream>>>(reinterpret_cast<int64_t *>(out), 1, 0);
            else
                k_fill_i32<<<1, kBlock, 0, stream>>>(reinterpret_cast<int32_t *>(out), 1, 0);
        }
        if (total) k_fill_i64<<<1, kBlock, 0, stream>>>(total, 1, 0);
        return hipGetLastError();
    }
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    k_scan_reduce<int32_t><<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(in, n, partial);
    k_scan_spine<<<1, 1024, 0, stream>>>(partial, nb, total);
    k_scan_apply<int32_t, TO><<<static_cast<unsigned>(nb), kBlock, 0, stream>>>(
        in, n, partial, out, write_total ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_scan_i32_to_i32(const int32_t *in, int64_t n, int32_t *out,
                                  int64_t *partial, int64_t *total, hipStream_t stream) {
    return scan_impl<int32_t>(in, n, out, partial, total, false, stream);
}

hipError_t launch_scan_i32_to_i64(const int32_t *in, int64_t n, int64_t *out,
                                  int64_t *partial, int64_t *total, hipStream_t stream) {
    return scan_impl<int64_t>(in, n, out, partial, total, true, stream);
}

template <int MODE, int CUT, bool HASHED>
static void launch_pairs_t(const PairArgs &a, unsigned nblocks, hipStream_t stream) {
    k_pairs<MODE, CUT, HASHED><<<nblocks, kBlock, 0, stream>>>(a);
}

template <int MODE, int CUT>
static void launch_pairs_h(bool hashed, const PairArgs &a, unsigned nb, hipStream_t s) {
    if (hashed)
        launch_pairs_t<MODE, CUT, true>(a, nb, s);
    else
        launch_pairs_t<MODE, CUT, false>(a, nb, s);
}

template <int MODE>
static void launch_pairs_c(int cut, bool hashed, const PairArgs &a, unsigned nb, hipStream_t s) {
    if (cut == CUT_PER_ATOM)
        launch_pairs_h<MODE, CUT_PER_ATOM>(hashed, a, nb, s);
    else if (cut == CUT_PER_TYPE)
        launch_pairs_h<MODE, CUT_PER_TYPE>(hashed, a, nb, s);
    else
        launch_pairs_h<MODE, CUT_GLOBAL>(hashed, a, nb, s);
}

hipError_t launch_pairs(int mode, int cut, bool hashed, const PairArgs &args,
                        hipStream_t stream) {
    if (args.home_cells <= 0) return hipSuccess;
    const unsigned nb =
        static_cast<unsigned>((args.home_cells + kWavesPerBlock - 1) / kWavesPerBlock);
    PairArgs a = args;
    a.xcd_blocks = static_cast<int32_t>(nb);
    if (mode == MODE_COUNT)
        launch_pairs_c<MODE_COUNT>(cut, hashed, a, nb, stream);
    else if (mode == MODE_FILL64)
        launch_pairs_c<MODE_FILL64>(cut, hashed, a, nb, stream);
    else
        launch_pairs_c<MODE_FILL32>(cut, hashed, a, nb, stream);
    return hipGetLastError();
}

hipError_t launch_set_params(SysGeom *geo, const SysGeom &g, int32_t *atom_base,
                             int32_t *cell_base, int32_t nat, int32_t ncells,
                             hipStream_t stream) {
    k_set_params<<<1, kWave, 0, stream>>>(geo, g, atom_base, cell_base, nat, ncells);
    return hipGetLastError();
}

hipError_t launch_sqrt_probe(const double *x, double *y, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_sqrt_probe<<<grid_for(n), kBlock, 0, stream>>>(x, y, n);
    return hipGetLastError();
}

}  // namespace mnl

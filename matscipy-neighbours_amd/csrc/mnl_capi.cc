/*
 * mnl host core and C ABI (include/mnl.h).
 *
 * The host side does what the reference's build_device does on the host
 * (src/libneighbours/neighbour_list_gpu.cc:309-492): argument checks, the
 * grid geometry, the backend choice, workspace management and the launch
 * sequence; all per-atom and per-pair work is in mnl_kernels.hip.
 *
 * Differences from the reference's host side, by design:
 *   - workspaces live in a plan and are reused (no ~15 hipMalloc/hipFree per
 *     call), outputs may be caller-owned (plan API);
 *   - one host synchronisation per build (the pair total, to size the
 *     outputs), none inside the binning;
 *   - HIP errors are returned as MNL_ERROR with a message instead of
 *     abort() (device.hh:68-76);
 *   - the error message is thread-local.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "mnl.h"
#include "mnl_engine.hh"

using namespace mnl;

namespace {

thread_local std::string g_error;

int fail(const std::string &msg) {
    g_error = msg;
    return MNL_ERROR;
}

#define MNL_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return fail(std::string("HIP error in ") + #expr + ": " +              \
                        hipGetErrorString(e_));                                    \
    } while (0)

/* Switch to a device for the scope; restore the previous one on exit. */
struct DeviceScope {
    int prev = -1;
    bool ok = true;
    explicit DeviceScope(int dev) {
        if (dev < 0) return;
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) {
            ok = false;
            return;
        }
        if (cur != dev) {
            if (hipSetDevice(dev) != hipSuccess) {
                ok = false;
                return;
            }
            prev = cur;
        }
    }
    ~DeviceScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

/* Device buffer that only grows. */
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes + bytes / 8, 256);
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(ptr);
    }
};

int check_device_available() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail("no HIP device available: the mnl engine runs only on the GPU");
    return MNL_SUCCESS;
}

bool valid_quantities(int q) { return (q & ~31) == 0; }

/* CellOrder (src/libneighbours/cell_list.hh:25): both orders give each cell
   its atoms in ascending index, so the output is the same for both. */
bool valid_order(int order) { return order == MNL_ORDER_LINEAR || order == MNL_ORDER_MORTON; }

/* The search range of a system: refused when negative (a negative cutoff)
   or when the stencil has more than kMaxStencil neighbour-cell offsets. */
const char *stencil_error(const SysGeom &g) {
    const int64_t nseg = stencil_size(g.nxyz);
    if (nseg < 0) return "negative cutoff";
    if (nseg > kMaxStencil)
        return "cutoff spans more than 2^30 neighbour cells of the grid "
               "((2nx+1)(2ny+1)(2nz+1) with nx = ceil(cutoff n / width))";
    return nullptr;
}

}  // namespace

/* ======================================================================== */

struct mnl_plan {
    int device = 0;
    /* workspaces */
    DevBuf raw4, ticket, tmp, spos, sinfo, stype, cell_count, cell_first, scount, slot,
        counts, order_stats, home_list, home_items, item_base,
        offsets, partial, geo, atom_base, cell_base, pt, hkey, total, hpos, hper_atom,
        htypes, cell_partial, det_scratch, tri_m, tri_cnt, tri_off, tri_vidx, tri_vstart, tri_flag, lj,
        tune_out, scan_status;
    int64_t *host_total = nullptr; /* pinned */
    /* state of the last count */
    bool counted = false;
    int32_t nat = 0;
    int32_t nsys = 1;
    int backend = MNL_BACKEND_DENSE;
    int cut = CUT_GLOBAL;
    int32_t home_cells = 0;
    int32_t batch_items = 0; /* batches: waves of a pair-kernel launch */
    int64_t hmask = 0;
    int64_t npairs = 0;
    int32_t ncut = 0;
    int32_t fill_chunks = 4; /* fill window (chunks of 64 candidates) */
    /* launch shape set by mnl_plan_set_launch_shape or the autotuner while it
       times a candidate (0: the tuned or built-in choice) */
    int32_t force_split = 0, force_chunks = 0;
    hipStream_t stream = nullptr;
    /* host copy of the per-type table on the device */
    std::vector<double> pt_host;
    bool pt_symmetric = true; /* pt_host[a][b] == pt_host[b][a] for all a, b */
    /* triplets: rows and pairs of the last mnl_plan_triplets_count */
    int32_t tri_rows = 0;
    int64_t tri_total = -1;
    const int32_t *tri_first = nullptr;
    bool tri_all = false; /* no distance cutoff: every pair kept, no index table */
    /* batched geometry uploaded by mnl_plan_set_batch and still in place */
    bool batch_ready = false;
    bool batch_small = false; /* every system fits one binning block */
    /* CellOrder of the plan's builds (mnl_plan_set_cell_order); both orders
       give the same output */
    int order = MNL_ORDER_LINEAR;
    /* workspaces carved from a fixed device array (the allocation-free
       build): they never grow, a build that needs more is refused */
    bool fixed = false;
};

namespace {

/* Launch shapes chosen by the autotuner, per configuration class: the same
   systems, atoms, cells (or hash slots), grid backend and cutoff form. The
   shape changes how work is dealt to waves, never the outputs. */
struct TuneKey {
    int32_t nsys, nat, cells, backend, cut;
    bool operator<(const TuneKey &o) const {
        return std::tie(nsys, nat, cells, backend, cut) <
               std::tie(o.nsys, o.nat, o.cells, o.backend, o.cut);
    }
};
struct TuneShape {
    int32_t split, chunks;
};
std::mutex g_tune_mutex;
std::map<TuneKey, TuneShape> g_tune;

TuneKey tune_key(const mnl_plan *p) {
    return TuneKey{p->nsys, p->nat, p->home_cells, p->backend, p->cut};
}

/* The pair-kernel arguments of the plan's last count. */
PairArgs pair_args(const mnl_plan *p) {
    PairArgs a{};
    a.geo = p->geo.as<SysGeom>();
    a.cell_base = p->cell_base.as<int32_t>();
    a.nsys = p->nsys;
    a.home_cells = p->home_cells;
    a.cell_first = p->cell_first.as<int32_t>();
    a.cell_count = p->cell_count.as<int32_t>();
    a.hkey = p->backend == MNL_BACKEND_HASHED ? p->hkey.as<int64_t>() : nullptr;
    a.home_list = p->backend == MNL_BACKEND_HASHED ? p->home_list.as<int32_t>() : nullptr;
    a.home_items = p->backend == MNL_BACKEND_HASHED ? p->home_items.as<int64_t>() : nullptr;
    a.hmask = p->hmask;
    a.spos = p->spos.as<double4>();
    a.sinfo = p->sinfo.as<int4>();
    a.stype = p->cut == CUT_PER_TYPE ? p->stype.as<int32_t>() : nullptr;
    a.pt_sq = p->cut == CUT_PER_TYPE ? p->pt.as<double>() : nullptr;
    a.ncut = p->ncut;
    a.fill_chunks = p->fill_chunks;
    /* Systems with few home cells (or slots) share each cell's atoms out
       over several waves: about 4096 waves in all, at most one per atom
       on average (fcc 1M: 97k cells, one wave each; 1000 atoms in 64
       cells: 16 waves per cell). A hashed grid has at most one occupied
       slot per atom. */
    {
        const int64_t cells = std::max<int64_t>(
            p->backend == MNL_BACKEND_HASHED ? std::min(p->home_cells, p->nat) : p->home_cells, 1);
        const int64_t want = (4096 + cells - 1) / cells;
        const int64_t per_cell = (static_cast<int64_t>(p->nat) + cells - 1) / cells;
        a.split = static_cast<int32_t>(std::max<int64_t>(1, std::min(want, per_cell)));
    }
    /* a forced shape, else a tuned one for this configuration class */
    if (p->force_split > 0) {
        a.split = p->force_split;
        if (p->force_chunks > 0) a.fill_chunks = p->force_chunks;
    } else {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune.find(tune_key(p));
        if (it != g_tune.end()) {
            a.split = it->second.split;
            a.fill_chunks = it->second.chunks;
        }
    }
    if (p->batch_ready) {
        /* batches: per-system splits replace the uniform one */
        a.item_base = p->item_base.as<int32_t>();
        a.items = p->batch_items;
        a.split = 1;
    }
    a.atom_base = p->atom_base.as<int32_t>();
    a.nat = p->nat;
    /* the count visits half the offsets when the test is symmetric in i, j */
    a.half = (p->cut != CUT_PER_TYPE || p->ncut == 0 || p->pt_symmetric) ? 1 : 0;
    a.scount = p->scount.as<int32_t>();
    a.counts = p->counts.as<int32_t>();
    a.offsets = p->offsets.as<int64_t>();
    a.capacity = p->npairs;
    return a;
}

/*
 * Grow the plan's workspaces for nat atoms, `cells` dense cells or hash
 * slots and nsys systems. Growing frees old buffers, so pending work on the
 * stream must finish first; a stream under graph capture cannot wait, and
 * the caller runs one uncaptured build of the same size beforehand.
 */
int ensure_workspaces(mnl_plan *p, hipStream_t stream, int32_t nat, int64_t cells,
                      bool hashed, int cut, int32_t ncutoffs, int32_t nsys) {
    const size_t n = static_cast<size_t>(std::max(nat, 1));
    const int64_t nparts =
        std::max(scan_partial_size(cells), scan_partial_size(static_cast<int64_t>(n)));
    struct Need {
        DevBuf *b;
        size_t bytes;
    } needs[] = {
        {&p->raw4, n * sizeof(int4)},
        {&p->ticket, n * sizeof(int32_t)},
        {&p->tmp, n * sizeof(int32_t)},
        {&p->spos, n * sizeof(double4)},
        {&p->sinfo, n * sizeof(int4)},
        {&p->stype, cut == CUT_PER_TYPE ? n * sizeof(int32_t) : 0},
        {&p->pt, cut == CUT_PER_TYPE
                     ? sizeof(double) * std::max<int64_t>(1, int64_t(ncutoffs) * ncutoffs)
                     : 0},
        {&p->scount, n * sizeof(int32_t)},
        {&p->slot, n * sizeof(int32_t)},
        {&p->counts, n * sizeof(int32_t)},
        {&p->offsets, (n + 1) * sizeof(int64_t)},
        {&p->cell_count, static_cast<size_t>(std::max<int64_t>(cells, 1)) * sizeof(int32_t)},
        {&p->cell_first, static_cast<size_t>(std::max<int64_t>(cells, 1)) * sizeof(int32_t)},
        {&p->partial, static_cast<size_t>(nparts) * sizeof(int64_t)},
        {&p->geo, sizeof(SysGeom) * static_cast<size_t>(nsys)},
        {&p->atom_base, (static_cast<size_t>(nsys) + 1) * sizeof(int32_t)},
        {&p->cell_base, (static_cast<size_t>(nsys) + 1) * sizeof(int32_t)},
        {&p->total, sizeof(int64_t)},
        {&p->hkey, hashed ? static_cast<size_t>(cells) * sizeof(int64_t) : 0},
        {&p->order_stats, 3 * sizeof(int32_t)},
        {&p->home_list, hashed ? n * sizeof(int32_t) : 0},
        {&p->home_items, hashed ? sizeof(int64_t) : 0},
        /* the single-pass scans' status words (not in the fixed workspace:
           its scans take the two-kernel form) */
        {&p->scan_status, p->fixed ? 0 : static_cast<size_t>(scan_status_words()) * 8},
    };
    bool grow = false;
    for (const Need &nd : needs) grow = grow || nd.bytes > nd.b->cap;
    if (!grow) return MNL_SUCCESS;
    if (p->fixed)
        return fail("the build exceeds the allocation-free workspace (at most " +
                    std::to_string(kNoallocMaxAtoms) + " atoms and " +
                    std::to_string(kNoallocMaxCells) + " cells or hash slots)");
    hipStreamCaptureStatus cap_status = hipStreamCaptureStatusNone;
    MNL_HIP(hipStreamIsCapturing(stream, &cap_status));
    if (cap_status != hipStreamCaptureStatusNone)
        return fail("mnl workspace must grow during graph capture: run one "
                    "uncaptured build of this size first");
    MNL_HIP(hipStreamSynchronize(stream));
    for (const Need &nd : needs) MNL_HIP(nd.b->ensure(nd.bytes));
    p->pt_host.clear(); /* buffers may have moved: upload the table again */
    /* no atom order seen yet: the first build keeps the default layout */
    MNL_HIP(launch_fill_i32(p->order_stats.as<int32_t>(), 3, 0, stream));
    /* the scans keep their status words zero between scans from here on */
    if (p->scan_status.ptr)
        MNL_HIP(launch_fill_i64(p->scan_status.as<int64_t>(), scan_status_words(), 0, stream));
    return MNL_SUCCESS;
}

/*
 * Binning of a build whose geometry (geo, atom_base, cell_base) is already
 * on the device: clear (unless `cleared`) -> bin -> scan -> scatter ->
 * rank/gather. Records no memset and does not synchronise (capturable).
 */
int enqueue_binning(mnl_plan *p, hipStream_t stream, const double *positions,
                    const double *per_atom, const int32_t *types, bool cleared = false) {
    const bool hashed = p->backend == MNL_BACKEND_HASHED;
    const int64_t cells = p->home_cells;
    const int32_t nat = p->nat;
    if (p->batch_ready && p->batch_small) {
        /* one block per system writes every cell count (no clearing) */
        MNL_HIP(launch_bin_systems(positions, p->nsys, p->geo.as<SysGeom>(),
                                   p->cell_first.as<int32_t>(), p->cell_count.as<int32_t>(),
                                   p->spos.as<double4>(), p->sinfo.as<int4>(),
                                   p->scount.as<int32_t>(), p->slot.as<int32_t>(), stream));
        return MNL_SUCCESS;
    }
    if (!cleared) MNL_HIP(launch_fill_i32(p->cell_count.as<int32_t>(), cells, 0, stream));
    if (hashed) {
        if (!cleared) MNL_HIP(launch_fill_i64(p->hkey.as<int64_t>(), cells, -1, stream));
        MNL_HIP(launch_bin_hashed(positions, nat, p->geo.as<SysGeom>(), p->raw4.as<int4>(),
                                  p->ticket.as<int32_t>(), p->hkey.as<int64_t>(), p->hmask,
                                  p->cell_count.as<int32_t>(), stream));
    } else {
        /* one system: count the input's order for the next build's layout
           (batched geometry keeps the default layout) */
        MNL_HIP(launch_bin_dense(positions, nat, p->geo.as<SysGeom>(),
                                 p->atom_base.as<int32_t>(), p->nsys, p->raw4.as<int4>(),
                                 p->ticket.as<int32_t>(), p->cell_count.as<int32_t>(),
                                 p->batch_ready ? nullptr : p->order_stats.as<int32_t>(),
                                 stream));
    }
    MNL_HIP(launch_scan_i32_to_i32(p->cell_count.as<int32_t>(), cells,
                                   p->cell_first.as<int32_t>(), p->partial.as<int64_t>(),
                                   nullptr, stream, p->scan_status.as<uint64_t>()));
    MNL_HIP(launch_scatter(p->raw4.as<int4>(), p->ticket.as<int32_t>(), nat,
                           p->cell_first.as<int32_t>(), p->tmp.as<int32_t>(), stream));
    MNL_HIP(launch_rank_gather(positions, p->raw4.as<int4>(), p->tmp.as<int32_t>(), nat,
                               p->cell_first.as<int32_t>(), p->cell_count.as<int32_t>(),
                               p->geo.as<SysGeom>(), p->atom_base.as<int32_t>(), p->nsys,
                               per_atom, types, p->spos.as<double4>(), p->sinfo.as<int4>(),
                               p->cut == CUT_PER_TYPE ? p->stype.as<int32_t>() : nullptr,
                               p->scount.as<int32_t>(), p->slot.as<int32_t>(),
                               hashed ? p->ticket.as<int32_t>() : nullptr, stream));
    if (hashed) {
        /* the occupied slots in the order of their lowest atom: flags (in
           ticket, free after the scatter), their scan (in tmp, free after
           the rank/gather), the list and its length */
        MNL_HIP(launch_scan_i32_to_i32(p->ticket.as<int32_t>(), nat, p->tmp.as<int32_t>(),
                                       p->partial.as<int64_t>(), p->home_items.as<int64_t>(),
                                       stream, p->scan_status.as<uint64_t>()));
        MNL_HIP(launch_home_list(p->raw4.as<int4>(), p->ticket.as<int32_t>(),
                                 p->tmp.as<int32_t>(), nat, p->home_list.as<int32_t>(), stream));
    }
    return MNL_SUCCESS;
}

/*
 * binning -> count -> scan. Leaves the pair total in p->total (device) and,
 * if given, in *total_copy (device); does not synchronise and records no
 * memset, so it can be captured in a graph.
 */
int enqueue_count_scan(mnl_plan *p, hipStream_t stream, int64_t *total_copy) {
    const bool hashed = p->backend == MNL_BACKEND_HASHED;
    const int32_t nat = p->nat;
    PairArgs a = pair_args(p);
    MNL_HIP(launch_pairs(0, p->cut, hashed, a, stream));
    if (a.half)
        MNL_HIP(launch_scan_sorted_counts(p->scount.as<int32_t>(), p->slot.as<int32_t>(),
                                          p->counts.as<int32_t>(), nat,
                                          p->offsets.as<int64_t>(), p->partial.as<int64_t>(),
                                          p->total.as<int64_t>(), stream, total_copy,
                                          p->scan_status.as<uint64_t>()));
    else
        MNL_HIP(launch_scan_i32_to_i64(p->counts.as<int32_t>(), nat, p->offsets.as<int64_t>(),
                                       p->partial.as<int64_t>(), p->total.as<int64_t>(),
                                       stream, total_copy, p->scan_status.as<uint64_t>()));
    return MNL_SUCCESS;
}

int enqueue_build(mnl_plan *p, hipStream_t stream, const double *positions,
                  const double *per_atom, const int32_t *types,
                  int64_t *total_copy = nullptr, bool cleared = false) {
    if (enqueue_binning(p, stream, positions, per_atom, types, cleared) != MNL_SUCCESS)
        return MNL_ERROR;
    return enqueue_count_scan(p, stream, total_copy);
}

/*
 * Enqueue a whole count for one system (inputs are device pointers). Leaves
 * the pair total in p->total (device); does not synchronise.
 */
int enqueue_count(mnl_plan *p, hipStream_t stream, const double origin[3],
                  const double cell[9], const double inv_cell[9], const int32_t pbc[3],
                  int32_t nat, const double *positions, double cutoff,
                  const double *per_atom, const double *per_type_sq, int32_t ncutoffs,
                  const int32_t *types, int backend, bool with_count = true,
                  int64_t *total_copy = nullptr) {
    p->counted = false;
    p->batch_ready = false;
    if (nat < 0) return fail("negative atom count");
    if (!(cutoff == cutoff)) return fail("cutoff is NaN");

    SysGeom g{};
    if (!make_sys_geom(origin, cell, inv_cell, pbc, cutoff, 0, nat, 0, g))
        return fail("Zero cell volume.");
    if (const char *why = stencil_error(g)) return fail(why);
    const int64_t ncells = static_cast<int64_t>(g.n[0]) * g.n[1] * g.n[2];
    const int64_t budget = std::max<int64_t>(int64_t(1) << 20, int64_t(8) * nat);
    bool hashed;
    if (backend == MNL_BACKEND_HASHED)
        hashed = true;
    else if (backend == MNL_BACKEND_DENSE)
        hashed = false;
    else
        hashed = ncells > budget; /* cell_list.cc:193-196 */
    if (!hashed && ncells > INT32_MAX / 2)
        return fail("cell grid too large for the dense backend");

    p->nat = nat;
    p->nsys = 1;
    p->backend = hashed ? MNL_BACKEND_HASHED : MNL_BACKEND_DENSE;
    p->cut = per_atom ? CUT_PER_ATOM : ((per_type_sq && types) ? CUT_PER_TYPE : CUT_GLOBAL);
    p->ncut = p->cut == CUT_PER_TYPE ? ncutoffs : 0;
    p->stream = stream;

    int64_t cells_or_slots = ncells;
    if (hashed) {
        int64_t cap = 1;
        while (cap < 2 * static_cast<int64_t>(nat) + 1) cap <<= 1;
        p->hmask = cap - 1;
        cells_or_slots = cap;
    } else {
        p->hmask = 0;
    }
    p->home_cells = static_cast<int32_t>(cells_or_slots);
    /* a 5-chunk fill window when it covers the typical home cell's
       candidates (nseg cells at the mean density) in one pass */
    {
        const double cand = static_cast<double>(g.nseg) * nat /
                            static_cast<double>(std::max<int64_t>(ncells, 1));
        p->fill_chunks = (cand > 4 * 64 && cand <= 5 * 64) ? 5 : 4;
    }

    if (ensure_workspaces(p, stream, nat, cells_or_slots, hashed, p->cut, ncutoffs, 1) !=
        MNL_SUCCESS)
        return MNL_ERROR;

    /* a small dense system is binned by one block, parameters included */
    const bool small = !hashed && small_binning_fits(nat, ncells);
    /* parameters are written on the device by the kernel that clears the
       grid: no host-to-device copy, nothing read from host memory later */
    if (!small)
        MNL_HIP(launch_set_params(p->geo.as<SysGeom>(), g, p->atom_base.as<int32_t>(),
                                  p->cell_base.as<int32_t>(), nat,
                                  static_cast<int32_t>(hashed ? 0 : ncells),
                                  p->cell_count.as<int32_t>(), cells_or_slots,
                                  hashed ? p->hkey.as<int64_t>() : nullptr,
                                  hashed ? nullptr : p->order_stats.as<int32_t>(), stream));
    if (p->cut == CUT_PER_TYPE && ncutoffs > 0) {
        /* the table is uploaded only when it differs from the plan's copy, so
           steady-state builds with fixed cutoffs stay asynchronous */
        const size_t nt = size_t(ncutoffs) * size_t(ncutoffs);
        if (p->pt_host.size() != nt ||
            std::memcmp(p->pt_host.data(), per_type_sq, nt * sizeof(double)) != 0) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            MNL_HIP(hipStreamIsCapturing(stream, &cs));
            if (cs != hipStreamCaptureStatusNone)
                return fail("a new per-type cutoff table cannot be uploaded under graph "
                            "capture: run one uncaptured build with it first");
            p->pt_host.assign(per_type_sq, per_type_sq + nt);
            p->pt_symmetric = true;
            for (int32_t u = 0; u < ncutoffs; u++)
                for (int32_t v = 0; v < u; v++)
                    if (!(p->pt_host[size_t(u) * ncutoffs + v] ==
                          p->pt_host[size_t(v) * ncutoffs + u]))
                        p->pt_symmetric = false;
            MNL_HIP(hipMemcpyAsync(p->pt.ptr, p->pt_host.data(), nt * sizeof(double),
                                   hipMemcpyHostToDevice, stream));
            MNL_HIP(hipStreamSynchronize(stream));
        }
    }
    if (small) {
        MNL_HIP(launch_bin_small(positions, nat, g, static_cast<int32_t>(ncells),
                                 p->geo.as<SysGeom>(), p->atom_base.as<int32_t>(),
                                 p->cell_base.as<int32_t>(), p->order_stats.as<int32_t>(),
                                 per_atom, types, p->cell_first.as<int32_t>(),
                                 p->cell_count.as<int32_t>(), p->spos.as<double4>(),
                                 p->sinfo.as<int4>(),
                                 p->cut == CUT_PER_TYPE ? p->stype.as<int32_t>() : nullptr,
                                 p->scount.as<int32_t>(), p->slot.as<int32_t>(), stream));
        if (!with_count) return MNL_SUCCESS;
        return enqueue_count_scan(p, stream, total_copy);
    }
    if (!with_count) return enqueue_binning(p, stream, positions, per_atom, types, true);
    return enqueue_build(p, stream, positions, per_atom, types, total_copy, true);
}

int read_total(mnl_plan *p, hipStream_t stream, int64_t *npairs) {
    MNL_HIP(hipMemcpyAsync(p->host_total, p->total.ptr, sizeof(int64_t),
                           hipMemcpyDeviceToHost, stream));
    /* the single-pass scans' error flag (set only by a look-back that gave up
       waiting, never expected), read with the total */
    int32_t *flag = p->scan_status.ptr ? p->scan_status.as<int32_t>() +
                                             2 * scan_status_words() - 2
                                       : nullptr;
    if (flag)
        MNL_HIP(hipMemcpyAsync(p->host_total + 1, flag, sizeof(int32_t), hipMemcpyDeviceToHost,
                               stream));
    MNL_HIP(hipStreamSynchronize(stream));
    p->npairs = *p->host_total;
    if (flag && *reinterpret_cast<int32_t *>(p->host_total + 1) != 0) {
        p->counted = false;
        /* status words and flag back to zero, so that the next build can run */
        MNL_HIP(launch_fill_i64(p->scan_status.as<int64_t>(), scan_status_words(), 0, stream));
        return fail("internal error: a scan's look-back gave up waiting; the build's "
                    "offsets are wrong");
    }
    p->counted = true;
    if (npairs) *npairs = p->npairs;
    return MNL_SUCCESS;
}

int enqueue_fill(mnl_plan *p, hipStream_t stream, int mode, int32_t *first, int32_t *secnd,
                 void *distvec, void *absdist, int32_t *shift) {
    if (!p->counted) return fail("mnl_plan_fill called without a successful mnl_plan_count");
    if (p->npairs < 0) return fail("pair count unknown: use mnl_plan_fill_async with a capacity");
    if (p->npairs == 0 || p->nat == 0) return MNL_SUCCESS;
    PairArgs a = pair_args(p);
    a.out_i = first;
    a.out_j = secnd;
    a.out_D = distvec;
    a.out_d = absdist;
    a.out_S = shift;
    MNL_HIP(launch_pairs(mode, p->cut, p->backend == MNL_BACKEND_HASHED, a, stream));
    return MNL_SUCCESS;
}

/* One cached plan per device for the one-shot entry points. */
std::mutex g_plans_mutex;
std::vector<mnl_plan *> g_plans;

mnl_plan *shared_plan(int device) {
    if (device < 0) return nullptr;
    if (static_cast<size_t>(device) >= g_plans.size()) g_plans.resize(device + 1, nullptr);
    if (!g_plans[device]) {
        mnl_plan *p = nullptr;
        if (mnl_plan_create(device, &p) != MNL_SUCCESS) return nullptr;
        g_plans[device] = p;
    }
    return g_plans[device];
}

}  // namespace

/* ------------------------------------------------ allocation-free build */

namespace {

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

/* Bytes of every workspace at the allocation-free limits, in carve order
   (the order of ensure_workspaces' list). */
constexpr size_t kNoallocPartials = 1024; /* >= scan_partial_size(kNoallocMaxCells) */
constexpr size_t noalloc_sizes[] = {
    size_t(kNoallocMaxAtoms) * sizeof(int4),          /* raw4 */
    size_t(kNoallocMaxAtoms) * sizeof(int32_t),       /* ticket */
    size_t(kNoallocMaxAtoms) * sizeof(int32_t),       /* tmp */
    size_t(kNoallocMaxAtoms) * sizeof(double4),       /* spos */
    size_t(kNoallocMaxAtoms) * sizeof(int4),          /* sinfo */
    size_t(kNoallocMaxAtoms) * sizeof(int32_t),       /* scount */
    size_t(kNoallocMaxAtoms) * sizeof(int32_t),       /* slot */
    size_t(kNoallocMaxAtoms) * sizeof(int32_t),       /* counts */
    (size_t(kNoallocMaxAtoms) + 1) * sizeof(int64_t), /* offsets */
    size_t(kNoallocMaxCells) * sizeof(int32_t),       /* cell_count */
    size_t(kNoallocMaxCells) * sizeof(int32_t),       /* cell_first */
    kNoallocPartials * sizeof(int64_t),               /* partial */
    sizeof(SysGeom),                                  /* geo */
    2 * sizeof(int32_t),                              /* atom_base */
    2 * sizeof(int32_t),                              /* cell_base */
    sizeof(int64_t),                                  /* total */
    size_t(kNoallocMaxCells) * sizeof(int64_t),       /* hkey */
    3 * sizeof(int32_t),                              /* order_stats */
    size_t(kNoallocMaxAtoms) * sizeof(int32_t),       /* home_list */
    sizeof(int64_t),                                  /* home_items */
};

constexpr size_t noalloc_total() {
    size_t off = 0;
    for (size_t b : noalloc_sizes) off = align256(off) + b;
    return off;
}
static_assert(noalloc_total() <= kNoallocScratchBytes,
              "the allocation-free workspaces exceed the device array");

std::mutex g_fixed_mutex;
mnl_plan g_fixed_plans[kNoallocMaxDevices];

/* Point the plan's workspaces into scratch; they never grow. */
void carve_fixed_plan(mnl_plan *p, int device, void *scratch) {
    DevBuf *bufs[] = {&p->raw4,   &p->ticket,     &p->tmp,        &p->spos,    &p->sinfo,
                      &p->scount, &p->slot,       &p->counts,     &p->offsets, &p->cell_count,
                      &p->cell_first, &p->partial, &p->geo,       &p->atom_base,
                      &p->cell_base,  &p->total,   &p->hkey,      &p->order_stats,
                      &p->home_list,  &p->home_items};
    static_assert(sizeof(bufs) / sizeof(bufs[0]) ==
                      sizeof(noalloc_sizes) / sizeof(noalloc_sizes[0]),
                  "one size per carved workspace");
    char *base = static_cast<char *>(scratch);
    size_t off = 0;
    for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); k++) {
        off = align256(off);
        bufs[k]->ptr = base + off;
        bufs[k]->cap = noalloc_sizes[k];
        off += noalloc_sizes[k];
    }
    p->device = device;
    p->fixed = true;
}

}  // namespace

int mnl::set_error(const char *msg) { return fail(msg); }

int mnl::noalloc_build(int device, void *scratch, hipStream_t stream, const double origin[3],
                       const double cell[9], const double inv_cell[9], const int32_t pbc[3],
                       int32_t nat, const double *positions, double cutoff,
                       const double *per_atom_cutoff, int backend, int64_t capacity,
                       int32_t *first, int32_t *secnd, void *distvec, void *absdist,
                       int32_t *shift, int float32_output, int64_t *npairs_device) {
    if (device < 0 || device >= kNoallocMaxDevices)
        return fail("device index beyond the allocation-free build's table");
    if (!origin || !cell || !inv_cell || !pbc) return fail("geometry pointer is NULL");
    if (nat > 0 && !positions) return fail("positions is NULL");
    if (capacity < 0) return fail("negative capacity");
    if (nat > kNoallocMaxAtoms)
        return fail("the allocation-free build takes at most " +
                    std::to_string(kNoallocMaxAtoms) + " atoms");
    std::lock_guard<std::mutex> lock(g_fixed_mutex);
    mnl_plan *p = &g_fixed_plans[device];
    if (!p->fixed) carve_fixed_plan(p, device, scratch);
    if (enqueue_count(p, stream, origin, cell, inv_cell, pbc, nat, positions, cutoff,
                      per_atom_cutoff, nullptr, 0, nullptr, backend, true,
                      npairs_device) != MNL_SUCCESS)
        return MNL_ERROR;
    p->counted = true;
    p->npairs = capacity;
    const int st = enqueue_fill(p, stream, float32_output ? 2 : 1, first, secnd, distvec,
                                absdist, shift);
    p->npairs = -1;
    return st;
}

/* ======================================================================== */

struct mnl_graph {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct mnl_result {
    int device = 0;
    int32_t nat = 0;
    int64_t npairs = 0;
    int quantities = 0;
    void *buf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}; /* i j D d S c */
};

namespace {

int slot_of(char q) {
    switch (q) {
        case 'i': return 0;
        case 'j': return 1;
        case 'D': return 2;
        case 'd': return 3;
        case 'S': return 4;
        case 'c': return 5;
        default: return -1;
    }
}

size_t bytes_of(const mnl_result *r, int slot) {
    const size_t P = static_cast<size_t>(r->npairs);
    switch (slot) {
        case 0: case 1: return P * 4;
        case 2: return P * 24;
        case 3: return P * 8;
        case 4: return P * 12;
        case 5: return static_cast<size_t>(r->nat) * 4;
        default: return 0;
    }
}

/* Output buffers of one-shot results. Each is allocated at its size class
   (block_bytes: rounded up to 1/8 of its highest power of two, at least
   64 KiB), so lists whose lengths differ by a little, as an MD loop's do from
   step to step, share blocks. A freed buffer is kept, once the device has
   finished all work queued before the free (hipFree waits for it too), and
   handed to the next result whose buffer has the same class on the same
   device; at most kBlockCacheMax bytes are kept. The reference allocates and
   frees every output per call (neighbour_list_gpu.cc:264-305); with buffers
   of gigabytes that costs more than the build. */
struct CachedBlock {
    int device;
    size_t bytes;
    void *ptr;
};
/* Never destroyed: a DLPack deleter may run late in process exit. */
struct BlockCache {
    std::mutex mutex;
    std::vector<CachedBlock> blocks;
    size_t bytes = 0;
};
BlockCache &block_cache() {
    static BlockCache *c = new BlockCache();
    return *c;
}
constexpr size_t kBlockCacheMax = size_t(16) << 30;

/* Takes the kept blocks of `device` (all devices when device < 0) out of the
   cache and returns them to the device. */
void release_blocks(int device) {
    BlockCache &bc = block_cache();
    std::vector<CachedBlock> out;
    {
        std::lock_guard<std::mutex> lock(bc.mutex);
        for (size_t k = 0; k < bc.blocks.size();) {
            if (device < 0 || bc.blocks[k].device == device) {
                out.push_back(bc.blocks[k]);
                bc.bytes -= bc.blocks[k].bytes;
                bc.blocks.erase(bc.blocks.begin() + static_cast<std::ptrdiff_t>(k));
            } else {
                k++;
            }
        }
    }
    for (const CachedBlock &b : out) {
        DeviceScope ds(b.device);
        (void)hipFree(b.ptr);
    }
}

size_t block_bytes(size_t bytes) {
    size_t top = 1;
    while (top <= bytes / 2) top *= 2;
    const size_t g = std::max<size_t>(size_t(64) << 10, top / 8);
    return (bytes + g - 1) / g * g;
}

/* With the device current. When the device is out of memory, the kept
   blocks of this device go back to it and the allocation is tried once more. */
hipError_t result_malloc(int device, void **p, size_t need) {
    const size_t bytes = block_bytes(need);
    BlockCache &bc = block_cache();
    {
        std::lock_guard<std::mutex> lock(bc.mutex);
        for (size_t k = 0; k < bc.blocks.size(); k++) {
            if (bc.blocks[k].device == device && bc.blocks[k].bytes == bytes) {
                *p = bc.blocks[k].ptr;
                bc.bytes -= bytes;
                bc.blocks.erase(bc.blocks.begin() + static_cast<std::ptrdiff_t>(k));
                return hipSuccess;
            }
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        release_blocks(device);
        e = hipMalloc(p, bytes);
    }
    return e;
}

/* With the device current; `need` as passed to result_malloc. */
void result_free(int device, void *p, size_t need) {
    if (!p) return;
    const size_t bytes = need > 0 ? block_bytes(need) : 0;
    if (bytes > 0 && bytes <= kBlockCacheMax && hipDeviceSynchronize() == hipSuccess) {
        std::vector<void *> evict;
        BlockCache &bc = block_cache();
        {
            std::lock_guard<std::mutex> lock(bc.mutex);
            /* oldest blocks of this device make room first */
            for (size_t k = 0; k < bc.blocks.size() && bc.bytes + bytes > kBlockCacheMax;) {
                if (bc.blocks[k].device == device) {
                    evict.push_back(bc.blocks[k].ptr);
                    bc.bytes -= bc.blocks[k].bytes;
                    bc.blocks.erase(bc.blocks.begin() + static_cast<std::ptrdiff_t>(k));
                } else {
                    k++;
                }
            }
            if (bc.bytes + bytes <= kBlockCacheMax) {
                bc.blocks.push_back(CachedBlock{device, bytes, p});
                bc.bytes += bytes;
                p = nullptr;
            }
        }
        for (void *e : evict) (void)hipFree(e);
    }
    if (p) (void)hipFree(p);
}

void free_result_buffers(mnl_result *r) {
    DeviceScope ds(r->device);
    for (int s = 0; s < 6; s++) {
        result_free(r->device, r->buf[s], bytes_of(r, s));
        r->buf[s] = nullptr;
    }
}

int one_shot(int quantities, bool want_pairs, const double origin[3], const double cell[9],
             const double inv_cell[9], const int32_t pbc[3], int32_t nat,
             const double *positions, int positions_on_device, double cutoff,
             const double *per_atom, const double *per_type_sq, int32_t ncutoffs,
             const int32_t *types, int device_id, mnl_result **out,
             hipStream_t stream = nullptr, int order = MNL_ORDER_LINEAR,
             int backend = MNL_BACKEND_AUTO) {
    if (!out) return fail("out is NULL");
    *out = nullptr;
    if (!origin || !cell || !inv_cell || !pbc) return fail("geometry pointer is NULL");
    if (!valid_quantities(quantities)) return fail("Unsupported quantity specified.");
    if (!valid_order(order)) return fail("unknown cell order (MNL_ORDER_LINEAR or _MORTON)");
    if (backend < MNL_BACKEND_AUTO || backend > MNL_BACKEND_HASHED)
        return fail("unknown cell-grid backend");
    if (nat > 0 && !positions) return fail("positions is NULL");
    if (per_type_sq && ncutoffs < 0) return fail("negative ncutoffs");
    if (check_device_available() != MNL_SUCCESS) return MNL_ERROR;
    int dev = device_id;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return fail("hipGetDevice failed");
    DeviceScope ds(dev);
    if (!ds.ok) return fail("cannot switch to the requested device");

    std::lock_guard<std::mutex> lock(g_plans_mutex);
    mnl_plan *p = shared_plan(dev);
    if (!p) return MNL_ERROR;
    const size_t n = static_cast<size_t>(std::max(nat, 1));

    /* host inputs go through the plan's staging buffers, copied on the
       build's stream (growing a buffer waits for the stream first) */
    auto stage = [&](DevBuf &b, const void *src, size_t bytes) -> int {
        if (bytes > b.cap) {
            MNL_HIP(hipStreamSynchronize(stream));
            MNL_HIP(b.ensure(bytes));
        }
        MNL_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, stream));
        return MNL_SUCCESS;
    };
    const double *dpos = positions;
    if (!positions_on_device && nat > 0) {
        if (stage(p->hpos, positions, size_t(nat) * 3 * sizeof(double)) != MNL_SUCCESS)
            return MNL_ERROR;
        dpos = p->hpos.as<double>();
    }
    const double *dper = nullptr;
    if (per_atom && nat > 0) {
        if (stage(p->hper_atom, per_atom, size_t(nat) * sizeof(double)) != MNL_SUCCESS)
            return MNL_ERROR;
        dper = p->hper_atom.as<double>();
    }
    const int32_t *dtypes = nullptr;
    if (types && nat > 0) {
        if (stage(p->htypes, types, size_t(nat) * sizeof(int32_t)) != MNL_SUCCESS)
            return MNL_ERROR;
        dtypes = p->htypes.as<int32_t>();
    }
    (void)n;

    auto *res = new mnl_result();
    res->device = dev;
    res->nat = nat;
    res->quantities = want_pairs ? quantities : 0;

    int64_t np = 0;
    if (enqueue_count(p, stream, origin, cell, inv_cell, pbc, nat, dpos, cutoff, dper,
                      per_type_sq, ncutoffs, dtypes, backend) != MNL_SUCCESS ||
        read_total(p, stream, &np) != MNL_SUCCESS) {
        delete res;
        return MNL_ERROR;
    }
    res->npairs = np;
    auto alloc = [&](int slot) -> int {
        const size_t b = bytes_of(res, slot);
        if (b == 0) return MNL_SUCCESS;
        hipError_t e = result_malloc(dev, &res->buf[slot], b);
        if (e != hipSuccess) {
            res->buf[slot] = nullptr;
            return fail(std::string("hipMalloc of an output failed: ") + hipGetErrorString(e));
        }
        return MNL_SUCCESS;
    };
    int st = MNL_SUCCESS;
    if (st == MNL_SUCCESS) st = alloc(5);
    if (want_pairs) {
        for (int s = 0; s < 5 && st == MNL_SUCCESS; s++)
            if (quantities & (1 << s)) st = alloc(s);
    }
    if (st == MNL_SUCCESS && nat > 0)
        st = mnl_plan_counts(p, stream, static_cast<int32_t *>(res->buf[5]));
    if (st == MNL_SUCCESS && want_pairs)
        st = enqueue_fill(p, stream, 1, static_cast<int32_t *>(res->buf[0]),
                          static_cast<int32_t *>(res->buf[1]), res->buf[2], res->buf[3],
                          static_cast<int32_t *>(res->buf[4]));
    if (st == MNL_SUCCESS) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) st = fail(std::string("HIP error: ") + hipGetErrorString(e));
    }
    if (st != MNL_SUCCESS) {
        free_result_buffers(res);
        delete res;
        return MNL_ERROR;
    }
    *out = res;
    return MNL_SUCCESS;
}

struct DLCtx {
    void *ptr;
    size_t bytes;
    int device;
    int64_t shape[2];
};

void dl_deleter(DLManagedTensor *t) {
    if (!t) return;
    auto *c = static_cast<DLCtx *>(t->manager_ctx);
    if (c) {
        DeviceScope ds(c->device);
        result_free(c->device, c->ptr, c->bytes);
        delete c;
    }
    delete t;
}

}  // namespace

/* ======================================================================== */

extern "C" {

const char *mnl_last_error(void) { return g_error.c_str(); }

const char *mnl_version(void) { return "mnl 0.1.0 (gfx950)"; }

int mnl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mnl_plan_create(int device_id, mnl_plan **plan) {
    if (!plan) return fail("plan is NULL");
    *plan = nullptr;
    if (check_device_available() != MNL_SUCCESS) return MNL_ERROR;
    int dev = device_id;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return fail("hipGetDevice failed");
    DeviceScope ds(dev);
    if (!ds.ok) return fail("cannot switch to the requested device");
    auto *p = new mnl_plan();
    p->device = dev;
    void *h = nullptr;
    hipError_t e = hipHostMalloc(&h, 2 * sizeof(int64_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        delete p;
        return fail(std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    p->host_total = static_cast<int64_t *>(h);
    *plan = p;
    return MNL_SUCCESS;
}

void mnl_plan_destroy(mnl_plan *p) {
    if (!p) return;
    DeviceScope ds(p->device);
    DevBuf *bufs[] = {&p->raw4,      &p->ticket,    &p->tmp,    &p->spos,     &p->sinfo,
                      &p->stype,     &p->cell_count, &p->cell_first, &p->scount, &p->slot, &p->counts, &p->order_stats, &p->offsets,
                      &p->home_list, &p->home_items, &p->item_base,
                      &p->partial,   &p->geo,       &p->atom_base, &p->cell_base, &p->pt,
                      &p->hkey,      &p->total,     &p->hpos,   &p->hper_atom, &p->htypes,
                      &p->cell_partial, &p->det_scratch, &p->tri_m,   &p->tri_cnt,  &p->tri_off,
                      &p->tri_vidx,  &p->tri_vstart, &p->tri_flag, &p->lj, &p->tune_out,
                      &p->scan_status};
    for (DevBuf *b : bufs) b->release();
    if (p->host_total) (void)hipHostFree(p->host_total);
    delete p;
}

int mnl_plan_device(const mnl_plan *p) { return p ? p->device : -1; }

int mnl_plan_count(mnl_plan *p, void *stream, const double origin[3], const double cell[9],
                   const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                   const double *positions, double cutoff, const double *per_atom_cutoff,
                   const double *per_type_cutoff_sq, int32_t ncutoffs, const int32_t *types,
                   int backend, int64_t *npairs) {
    if (!p) return fail("plan is NULL");
    if (!origin || !cell || !inv_cell || !pbc) return fail("geometry pointer is NULL");
    if (nat > 0 && !positions) return fail("positions is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (enqueue_count(p, s, origin, cell, inv_cell, pbc, nat, positions, cutoff,
                      per_atom_cutoff, per_type_cutoff_sq, ncutoffs, types,
                      backend) != MNL_SUCCESS)
        return MNL_ERROR;
    return read_total(p, s, npairs);
}

int mnl_plan_count_async(mnl_plan *p, void *stream, const double origin[3],
                         const double cell[9], const double inv_cell[9], const int32_t pbc[3],
                         int32_t nat, const double *positions, double cutoff,
                         const double *per_atom_cutoff, const double *per_type_cutoff_sq,
                         int32_t ncutoffs, const int32_t *types, int backend,
                         int64_t *npairs_device) {
    if (!p) return fail("plan is NULL");
    if (!origin || !cell || !inv_cell || !pbc) return fail("geometry pointer is NULL");
    if (nat > 0 && !positions) return fail("positions is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (enqueue_count(p, s, origin, cell, inv_cell, pbc, nat, positions, cutoff,
                      per_atom_cutoff, per_type_cutoff_sq, ncutoffs, types, backend, true,
                      npairs_device) != MNL_SUCCESS)
        return MNL_ERROR;
    p->counted = true;
    p->npairs = -1; /* unknown on the host until read */
    return MNL_SUCCESS;
}

int mnl_plan_fill_async(mnl_plan *p, void *stream, int64_t capacity, int32_t *first,
                        int32_t *secnd, void *distvec, void *absdist, int32_t *shift,
                        int float32_output) {
    if (!p) return fail("plan is NULL");
    if (!p->counted) return fail("mnl_plan_fill_async called without a count");
    if (capacity < 0) return fail("negative capacity");
    DeviceScope ds(p->device);
    const int64_t saved = p->npairs;
    p->npairs = capacity;
    int st = enqueue_fill(p, static_cast<hipStream_t>(stream), float32_output ? 2 : 1, first,
                          secnd, distvec, absdist, shift);
    p->npairs = saved;
    return st;
}

int mnl_plan_fill(mnl_plan *p, void *stream, int32_t *first, int32_t *secnd, double *distvec,
                  double *absdist, int32_t *shift) {
    if (!p) return fail("plan is NULL");
    DeviceScope ds(p->device);
    return enqueue_fill(p, static_cast<hipStream_t>(stream), 1, first, secnd, distvec, absdist,
                        shift);
}

int mnl_plan_fill_f32(mnl_plan *p, void *stream, int32_t *first, int32_t *secnd,
                      float *distvec, float *absdist, int32_t *shift) {
    if (!p) return fail("plan is NULL");
    DeviceScope ds(p->device);
    return enqueue_fill(p, static_cast<hipStream_t>(stream), 2, first, secnd, distvec, absdist,
                        shift);
}

int mnl_plan_counts(mnl_plan *p, void *stream, int32_t *counts) {
    if (!p) return fail("plan is NULL");
    if (!p->counted) return fail("mnl_plan_counts called without a successful count");
    if (p->nat == 0) return MNL_SUCCESS;
    if (!counts) return fail("counts is NULL");
    DeviceScope ds(p->device);
    MNL_HIP(hipMemcpyAsync(counts, p->counts.ptr, sizeof(int32_t) * size_t(p->nat),
                           hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
    return MNL_SUCCESS;
}

const int64_t *mnl_plan_offsets(const mnl_plan *p) {
    return (p && p->counted) ? p->offsets.as<int64_t>() : nullptr;
}

int mnl_plan_backend(const mnl_plan *p) { return p ? p->backend : -1; }

int mnl_neighbour_list_device(int quantities, const double origin[3], const double cell[9],
                              const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                              const double *positions, int positions_on_device,
                              double cutoff, const double *per_atom_cutoff,
                              const double *per_type_cutoff_sq, int32_t ncutoffs,
                              const int32_t *types, int device_id, mnl_result **out) {
    return one_shot(quantities, true, origin, cell, inv_cell, pbc, nat, positions,
                    positions_on_device, cutoff, per_atom_cutoff, per_type_cutoff_sq,
                    ncutoffs, types, device_id, out);
}

int mnl_neighbour_count_device(const double origin[3], const double cell[9],
                               const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                               const double *positions, int positions_on_device,
                               double cutoff, const double *per_atom_cutoff,
                               const double *per_type_cutoff_sq, int32_t ncutoffs,
                               const int32_t *types, int device_id, mnl_result **out) {
    return one_shot(0, false, origin, cell, inv_cell, pbc, nat, positions,
                    positions_on_device, cutoff, per_atom_cutoff, per_type_cutoff_sq,
                    ncutoffs, types, device_id, out);
}

int mnl_neighbour_list_request(const mnl_request *rq, mnl_result **out) {
    if (!rq) return fail("request is NULL");
    return one_shot(rq->quantities, true, rq->cell_origin, rq->cell, rq->inv_cell, rq->pbc,
                    rq->nat, rq->positions, rq->positions_on_device, rq->cutoff,
                    rq->per_atom_cutoff, rq->per_type_cutoff_sq, rq->ncutoffs, rq->types,
                    rq->device_id, out, static_cast<hipStream_t>(rq->stream), rq->order,
                    rq->backend);
}

int mnl_neighbour_count_request(const mnl_request *rq, mnl_result **out) {
    if (!rq) return fail("request is NULL");
    return one_shot(0, false, rq->cell_origin, rq->cell, rq->inv_cell, rq->pbc, rq->nat,
                    rq->positions, rq->positions_on_device, rq->cutoff, rq->per_atom_cutoff,
                    rq->per_type_cutoff_sq, rq->ncutoffs, rq->types, rq->device_id, out,
                    static_cast<hipStream_t>(rq->stream), rq->order, rq->backend);
}

int mnl_plan_set_cell_order(mnl_plan *p, int order) {
    if (!p) return fail("plan is NULL");
    if (!valid_order(order)) return fail("unknown cell order (MNL_ORDER_LINEAR or _MORTON)");
    p->order = order;
    return MNL_SUCCESS;
}

int mnl_plan_cell_order(const mnl_plan *p) { return p ? p->order : -1; }

int64_t mnl_result_npairs(const mnl_result *r) { return r ? r->npairs : 0; }

int32_t mnl_result_nat(const mnl_result *r) { return r ? r->nat : 0; }

int mnl_result_device(const mnl_result *r) { return r ? r->device : -1; }

void *mnl_result_data(const mnl_result *r, char q) {
    const int s = slot_of(q);
    if (!r || s < 0) return nullptr;
    return r->buf[s];
}

int mnl_result_copy_to_host(const mnl_result *r, char q, void *dst) {
    if (!r) return fail("result is NULL");
    const int s = slot_of(q);
    if (s < 0) return fail("Unsupported quantity specified.");
    const size_t b = bytes_of(r, s);
    if (b == 0) return MNL_SUCCESS;
    if (!r->buf[s]) return fail(std::string("quantity '") + q + "' is not held by the result");
    if (!dst) return fail("dst is NULL");
    DeviceScope ds(r->device);
    MNL_HIP(hipMemcpy(dst, r->buf[s], b, hipMemcpyDeviceToHost));
    return MNL_SUCCESS;
}

DLManagedTensor *mnl_result_export_dlpack(mnl_result *r, char q) {
    const int s = slot_of(q);
    if (!r || s < 0) {
        fail("Unsupported quantity specified.");
        return nullptr;
    }
    const bool present = r->buf[s] != nullptr || bytes_of(r, s) == 0;
    const bool requested = s == 5 || (r->quantities & (1 << s));
    if (!present || !requested) {
        fail(std::string("quantity '") + q + "' is not held by the result");
        return nullptr;
    }
    auto *ctx = new DLCtx();
    ctx->ptr = r->buf[s];
    ctx->bytes = bytes_of(r, s);
    ctx->device = r->device;
    r->buf[s] = nullptr;
    if (s < 5) r->quantities &= ~(1 << s);
    auto *t = new DLManagedTensor();
    std::memset(t, 0, sizeof *t);
    const bool vec3 = (s == 2 || s == 4);
    ctx->shape[0] = s == 5 ? r->nat : r->npairs;
    ctx->shape[1] = 3;
    t->dl_tensor.data = ctx->ptr;
    t->dl_tensor.device.device_type = kDLROCM;
    t->dl_tensor.device.device_id = r->device;
    t->dl_tensor.ndim = vec3 ? 2 : 1;
    t->dl_tensor.dtype.code = (s == 2 || s == 3) ? kDLFloat : kDLInt;
    t->dl_tensor.dtype.bits = (s == 2 || s == 3) ? 64 : 32;
    t->dl_tensor.dtype.lanes = 1;
    t->dl_tensor.shape = ctx->shape;
    t->dl_tensor.strides = nullptr;
    t->dl_tensor.byte_offset = 0;
    t->manager_ctx = ctx;
    t->deleter = dl_deleter;
    return t;
}

void mnl_release_cached_memory(void) { release_blocks(-1); }

int64_t mnl_cached_bytes(int device) {
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lock(bc.mutex);
    size_t total = 0;
    for (const CachedBlock &b : bc.blocks)
        if (device < 0 || b.device == device) total += b.bytes;
    return static_cast<int64_t>(total);
}

void mnl_result_free(mnl_result *r) {
    if (!r) return;
    free_result_buffers(r);
    delete r;
}

namespace {
int matrix_enqueue(mnl_plan *p, hipStream_t st, const double origin[3], const double cell[9],
                   const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                   const double *positions, double cutoff, const double *per_atom,
                   const double *per_type_sq, int32_t ncutoffs, const int32_t *types,
                   int backend, int32_t max_neighbours, int32_t *idx, double *dist,
                   int32_t *count, int32_t *overflow) {
    if (!origin || !cell || !inv_cell || !pbc) return fail("geometry pointer is NULL");
    if (nat > 0 && !positions) return fail("positions is NULL");
    if (max_neighbours < 0) return fail("max_neighbours must be non-negative");
    const int64_t slots = int64_t(nat) * max_neighbours;
    if (nat > 0 && ((slots > 0 && (!idx || !dist)) || !count))
        return fail("matrix output pointer is NULL");
    if (enqueue_count(p, st, origin, cell, inv_cell, pbc, nat, positions, cutoff, per_atom,
                      per_type_sq, ncutoffs, types, backend, false) != MNL_SUCCESS)
        return MNL_ERROR;
    if (overflow) MNL_HIP(launch_fill_i32(overflow, 1, 0, st));
    /* no clearing of idx and dist: the traversal writes every slot, the
       unused ones (from each row's count on) as zeros */
    if (nat == 0) return MNL_SUCCESS;
    PairArgs a = pair_args(p);
    a.counts = count;
    a.out_j = idx;
    a.out_D = dist;
    a.max_nb = max_neighbours;
    a.overflow = overflow;
    MNL_HIP(launch_pairs(3, p->cut, p->backend == MNL_BACKEND_HASHED, a, st));
    p->counted = false; /* the plan holds no pair offsets after a matrix build */
    return MNL_SUCCESS;
}
}  // namespace

int mnl_plan_matrix_async(mnl_plan *p, void *stream, const double origin[3],
                          const double cell[9], const double inv_cell[9],
                          const int32_t pbc[3], int32_t nat, const double *positions,
                          double cutoff, const double *per_atom_cutoff,
                          const double *per_type_cutoff_sq, int32_t ncutoffs,
                          const int32_t *types, int backend, int32_t max_neighbours,
                          int32_t *idx, double *dist, int32_t *count,
                          int32_t *overflow_device) {
    if (!p) return fail("plan is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    return matrix_enqueue(p, static_cast<hipStream_t>(stream), origin, cell, inv_cell, pbc, nat,
                          positions, cutoff, per_atom_cutoff, per_type_cutoff_sq, ncutoffs,
                          types, backend, max_neighbours, idx, dist, count, overflow_device);
}

int mnl_plan_matrix(mnl_plan *p, void *stream, const double origin[3], const double cell[9],
                    const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                    const double *positions, double cutoff, const double *per_atom_cutoff,
                    const double *per_type_cutoff_sq, int32_t ncutoffs, const int32_t *types,
                    int backend, int32_t max_neighbours, int32_t *idx, double *dist,
                    int32_t *count, int32_t *overflow) {
    if (!p) return fail("plan is NULL");
    if (!overflow) return fail("overflow is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    /* the flag lives in the plan's total slot (an int64, used here as int32) */
    if (p->total.cap < sizeof(int64_t)) MNL_HIP(p->total.ensure(sizeof(int64_t)));
    int32_t *flag = p->total.as<int32_t>();
    if (matrix_enqueue(p, st, origin, cell, inv_cell, pbc, nat, positions, cutoff,
                       per_atom_cutoff, per_type_cutoff_sq, ncutoffs, types, backend,
                       max_neighbours, idx, dist, count, flag) != MNL_SUCCESS)
        return MNL_ERROR;
    MNL_HIP(hipMemcpyAsync(p->host_total, flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    MNL_HIP(hipStreamSynchronize(st));
    int32_t v = 0;
    std::memcpy(&v, p->host_total, sizeof v);
    *overflow = nat > 0 ? v : 0;
    return MNL_SUCCESS;
}

namespace {
/* candidate tests a batch wave aims at (fcc 1M: about 3,100 per home cell;
   on the 1024-system batch 8192 measured best of 1024 ... 32768) */
constexpr double kBatchWaveTests = 8192.0;
}  // namespace

int mnl_plan_set_batch(mnl_plan *p, void *stream, int32_t nsys, const int32_t *atom_offsets,
                       const double *origins, const double *cells, const double *inv_cells,
                       const int32_t *pbcs, const double *cutoffs) {
    if (!p) return fail("plan is NULL");
    if (nsys < 1) return fail("a batch needs at least one system");
    if (!atom_offsets || !cells || !inv_cells || !pbcs || !cutoffs)
        return fail("batch geometry pointer is NULL");
    if (atom_offsets[0] != 0) return fail("atom_offsets[0] must be 0");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    p->counted = false;
    p->batch_ready = false;
    std::vector<SysGeom> geo(static_cast<size_t>(nsys));
    std::vector<int32_t> abase(static_cast<size_t>(nsys) + 1), cbase(static_cast<size_t>(nsys) + 1);
    const double zero3[3] = {0.0, 0.0, 0.0};
    int64_t ncells = 0;
    for (int32_t s = 0; s < nsys; s++) {
        const int32_t a0 = atom_offsets[s], a1 = atom_offsets[s + 1];
        if (a1 < a0) return fail("atom_offsets must be non-decreasing");
        const double rc = cutoffs[s];
        if (!(rc == rc)) return fail("cutoff is NaN");
        SysGeom &g = geo[static_cast<size_t>(s)];
        if (!make_sys_geom(origins ? origins + 3 * int64_t(s) : zero3, cells + 9 * int64_t(s),
                           inv_cells + 9 * int64_t(s), pbcs + 3 * int64_t(s), rc, a0, a1 - a0,
                           static_cast<int32_t>(ncells), g))
            return fail("Zero cell volume (system " + std::to_string(s) + ").");
        if (const char *why = stencil_error(g))
            return fail(std::string(why) + " (system " + std::to_string(s) + ")");
        abase[static_cast<size_t>(s)] = a0;
        cbase[static_cast<size_t>(s)] = static_cast<int32_t>(ncells);
        ncells += static_cast<int64_t>(g.n[0]) * g.n[1] * g.n[2];
        if (ncells > INT32_MAX / 2) return fail("batched cell grids too large");
    }
    const int32_t nat = atom_offsets[nsys];
    abase[static_cast<size_t>(nsys)] = nat;
    cbase[static_cast<size_t>(nsys)] = static_cast<int32_t>(ncells);
    /* Waves per home cell, per system: a home cell costs about m * (nseg m)
       candidate tests at m atoms per cell. A system of one cell holds all
       its atoms there and tests 27 images of it: 50 atoms make 67,500 tests,
       twenty times a typical cell, so such cells are shared out over several
       waves (at most one per atom) to keep every wave under about
       kBatchWaveTests. */
    std::vector<int32_t> ibase(static_cast<size_t>(nsys) + 1);
    int64_t items = 0;
    int64_t max_atoms = 0, max_cells = 0;
    for (int32_t s = 0; s < nsys; s++) {
        SysGeom &g = geo[static_cast<size_t>(s)];
        const double cells_s = std::max<double>(1.0, double(g.ncells));
        const double m = double(g.natoms) / cells_s;
        const double tests = m * m * double(g.nseg);
        const int64_t want = static_cast<int64_t>(std::ceil(tests / kBatchWaveTests));
        const int64_t cap = std::max<int64_t>(1, static_cast<int64_t>(std::ceil(m)));
        g.split = static_cast<int32_t>(std::max<int64_t>(1, std::min(want, cap)));
        ibase[static_cast<size_t>(s)] = static_cast<int32_t>(items);
        max_atoms = std::max<int64_t>(max_atoms, g.natoms);
        max_cells = std::max<int64_t>(max_cells, g.ncells);
        items += static_cast<int64_t>(g.ncells) * g.split;
        if (items > INT32_MAX / 2) return fail("batched work items too many");
    }
    ibase[static_cast<size_t>(nsys)] = static_cast<int32_t>(items);
    if (ensure_workspaces(p, st, nat, ncells, false, CUT_GLOBAL, 0, nsys) != MNL_SUCCESS)
        return MNL_ERROR;
    MNL_HIP(p->item_base.ensure(sizeof(int32_t) * ibase.size()));
    MNL_HIP(hipMemcpyAsync(p->item_base.ptr, ibase.data(), sizeof(int32_t) * ibase.size(),
                           hipMemcpyHostToDevice, st));
    MNL_HIP(hipMemcpyAsync(p->geo.ptr, geo.data(), sizeof(SysGeom) * geo.size(),
                           hipMemcpyHostToDevice, st));
    MNL_HIP(hipMemcpyAsync(p->atom_base.ptr, abase.data(), sizeof(int32_t) * abase.size(),
                           hipMemcpyHostToDevice, st));
    MNL_HIP(hipMemcpyAsync(p->cell_base.ptr, cbase.data(), sizeof(int32_t) * cbase.size(),
                           hipMemcpyHostToDevice, st));
    MNL_HIP(hipStreamSynchronize(st));
    p->nat = nat;
    p->nsys = nsys;
    p->fill_chunks = 4;
    p->backend = MNL_BACKEND_DENSE;
    p->cut = CUT_GLOBAL;
    p->ncut = 0;
    p->hmask = 0;
    p->home_cells = static_cast<int32_t>(ncells);
    p->batch_items = static_cast<int32_t>(items);
    p->batch_small = batch_binning_fits(max_atoms, max_cells);
    p->stream = st;
    p->batch_ready = true;
    return MNL_SUCCESS;
}

namespace {
int batch_count_enqueue(mnl_plan *p, hipStream_t st, const double *positions,
                        int64_t *npairs_device) {
    if (!p->batch_ready)
        return fail("no batch geometry in the plan: call mnl_plan_set_batch first (a "
                    "single-system count on the plan replaces it)");
    if (p->nat > 0 && !positions) return fail("positions is NULL");
    p->counted = false;
    if (enqueue_build(p, st, positions, nullptr, nullptr, npairs_device) != MNL_SUCCESS)
        return MNL_ERROR;
    p->counted = true;
    p->npairs = -1;
    return MNL_SUCCESS;
}
}  // namespace

int mnl_plan_batch_count_async(mnl_plan *p, void *stream, const double *positions,
                               int64_t *npairs_device) {
    if (!p) return fail("plan is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    return batch_count_enqueue(p, static_cast<hipStream_t>(stream), positions, npairs_device);
}

int mnl_plan_batch_count(mnl_plan *p, void *stream, const double *positions,
                         int64_t *npairs) {
    if (!p) return fail("plan is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (batch_count_enqueue(p, st, positions, nullptr) != MNL_SUCCESS) return MNL_ERROR;
    return read_total(p, st, npairs);
}

int mnl_plan_system_offsets(mnl_plan *p, void *stream, int64_t *system_pair_offsets) {
    if (!p) return fail("plan is NULL");
    if (!p->counted) return fail("mnl_plan_system_offsets called without a count");
    if (!system_pair_offsets) return fail("system_pair_offsets is NULL");
    DeviceScope ds(p->device);
    MNL_HIP(launch_system_offsets(p->offsets.as<int64_t>(), p->atom_base.as<int32_t>(),
                                  p->nsys, system_pair_offsets,
                                  static_cast<hipStream_t>(stream)));
    return MNL_SUCCESS;
}

int32_t mnl_plan_nsys(const mnl_plan *p) { return p ? p->nsys : 0; }

int mnl_plan_batch_graph_create(mnl_plan *p, const double *positions, int64_t capacity,
                                int32_t *first, int32_t *secnd, void *distvec, void *absdist,
                                int32_t *shift, int float32_output,
                                int64_t *system_pair_offsets, int64_t *npairs_device,
                                mnl_graph **out) {
    if (!out) return fail("out is NULL");
    *out = nullptr;
    if (!p) return fail("plan is NULL");
    if (!p->batch_ready) return fail("no batch geometry in the plan: call mnl_plan_set_batch");
    if (capacity < 0) return fail("negative capacity");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t cs = nullptr;
    MNL_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fail(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    int st = batch_count_enqueue(p, cs, positions, npairs_device);
    if (st == MNL_SUCCESS) {
        const int64_t saved = p->npairs;
        p->npairs = capacity;
        st = enqueue_fill(p, cs, float32_output ? 2 : 1, first, secnd, distvec, absdist, shift);
        p->npairs = saved;
    }
    if (st == MNL_SUCCESS && system_pair_offsets) {
        e = launch_system_offsets(p->offsets.as<int64_t>(), p->atom_base.as<int32_t>(),
                                  p->nsys, system_pair_offsets, cs);
        if (e != hipSuccess) st = fail(std::string("HIP error: ") + hipGetErrorString(e));
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cs, &graph);
    (void)hipStreamDestroy(cs);
    if (st != MNL_SUCCESS) {
        if (graph) (void)hipGraphDestroy(graph);
        return MNL_ERROR;
    }
    if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    auto *g = new mnl_graph();
    g->device = p->device;
    g->graph = graph;
    g->exec = exec;
    *out = g;
    return MNL_SUCCESS;
}

int mnl_graph_launch(mnl_graph *g, void *stream) {
    if (!g) return fail("graph is NULL");
    DeviceScope ds(g->device);
    MNL_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    return MNL_SUCCESS;
}

void mnl_graph_destroy(mnl_graph *g) {
    if (!g) return;
    DeviceScope ds(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

int mnl_plan_pair_backward(mnl_plan *p, void *stream, int64_t npairs, int32_t nat,
                           const int32_t *first, const int32_t *secnd, const int32_t *shift,
                           const void *distvec, const void *absdist,
                           const void *grad_distvec, const void *grad_absdist,
                           int float32, const int32_t pbc[3], int deterministic,
                           double *grad_positions, double *grad_cell) {
    if (!p) return fail("plan is NULL");
    if (npairs < 0 || nat < 0) return fail("negative size");
    if (!grad_positions && nat > 0) return fail("grad_positions is NULL");
    if (!grad_cell) return fail("grad_cell is NULL");
    if (npairs > 0) {
        if (!first || !secnd || !shift) return fail("i, j and S are required");
        if (!grad_distvec && !grad_absdist) return fail("no incoming gradient");
        if (grad_absdist && (!distvec || !absdist))
            return fail("D and d are required for the gradient of d");
    }
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int pbc_mask = pbc ? (pbc[0] ? 1 : 0) | (pbc[1] ? 2 : 0) | (pbc[2] ? 4 : 0) : 7;
    DevBuf &ws = deterministic ? p->det_scratch : p->cell_partial;
    if (deterministic && npairs >= INT32_MAX)
        return fail("the deterministic backward takes fewer than 2^31 pairs");
    const size_t need = deterministic
                            ? backward_det_scratch_bytes(npairs, nat)
                            : sizeof(double) * static_cast<size_t>(backward_scratch_doubles(nat));
    if (need > ws.cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        MNL_HIP(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail("mnl workspace must grow during graph capture");
        MNL_HIP(hipStreamSynchronize(st));
        MNL_HIP(ws.ensure(need));
    }
    if (deterministic)
        MNL_HIP(launch_pair_backward_det(npairs, nat, first, secnd, shift, distvec, absdist,
                                         grad_distvec, grad_absdist, float32 != 0, pbc_mask,
                                         grad_positions, grad_cell, ws.ptr, st));
    else
        MNL_HIP(launch_pair_backward(npairs, nat, first, secnd, shift, distvec, absdist,
                                     grad_distvec, grad_absdist, float32 != 0, pbc_mask,
                                     grad_positions, grad_cell, p->cell_partial.as<double>(), st));
    return MNL_SUCCESS;
}

int mnl_plan_lj(mnl_plan *p, void *stream, const double origin[3], const double cell[9],
                const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                const double *positions, double cutoff, double epsilon, double sigma,
                double *forces, double *energy, double *virial) {
    if (!p) return fail("plan is NULL");
    if (nat < 0) return fail("negative atom count");
    if (nat > 0 && (!positions || !forces)) return fail("positions and forces are required");
    if (!(sigma == sigma) || !(epsilon == epsilon)) return fail("epsilon or sigma is NaN");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    /* the cell grid only (no count): the traversal below is the consumer */
    if (enqueue_count(p, st, origin, cell, inv_cell, pbc, nat, positions, cutoff, nullptr,
                      nullptr, 0, nullptr, -1, false) != MNL_SUCCESS)
        return MNL_ERROR;
    const PairArgs a = pair_args(p);
    const size_t need = sizeof(double) * static_cast<size_t>(lj_scratch_doubles(a));
    if (need > p->lj.cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        MNL_HIP(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail("mnl workspace must grow during graph capture");
        MNL_HIP(hipStreamSynchronize(st));
        MNL_HIP(p->lj.ensure(need));
    }
    MNL_HIP(launch_lj(a, p->backend == MNL_BACKEND_HASHED, epsilon, sigma,
                      p->slot.as<int32_t>(), p->lj.as<double>(), forces, energy, virial, st));
    return MNL_SUCCESS;
}

int mnl_plan_set_launch_shape(mnl_plan *p, int32_t split, int32_t chunks) {
    if (!p) return fail("plan is NULL");
    if (split < 0 || split > 64) return fail("split must be 0 (automatic) or 1..64");
    if (chunks != 0 && chunks != 4 && chunks != 5) return fail("chunks must be 0, 4 or 5");
    p->force_split = split;
    p->force_chunks = split > 0 ? chunks : 0;
    return MNL_SUCCESS;
}

int mnl_plan_autotune(mnl_plan *p, void *stream, int32_t *split, int32_t *chunks,
                      float *best_ms) {
    if (!p) return fail("plan is NULL");
    if (!p->counted) return fail("mnl_plan_autotune needs a count (or batch count) first");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MNL_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return fail("mnl_plan_autotune cannot run under capture");
    int64_t P = 0;
    if (read_total(p, st, &P) != MNL_SUCCESS) return MNL_ERROR;
    constexpr int64_t kTuneMaxPairs = int64_t(1) << 30;
    if (P > kTuneMaxPairs) return fail("too many pairs to autotune (limit 2^30)");
    /* scratch outputs i, j, D, d, S: 52 bytes per pair, 8-byte aligned parts */
    const size_t n = static_cast<size_t>(std::max<int64_t>(P, 1));
    MNL_HIP(hipStreamSynchronize(st));
    MNL_HIP(p->tune_out.ensure(52 * n));
    char *base = p->tune_out.as<char>();
    int32_t *oi = reinterpret_cast<int32_t *>(base);
    int32_t *oj = oi + n;
    double *oD = reinterpret_cast<double *>(base + 8 * n);
    double *od = oD + 3 * n;
    int32_t *oS = reinterpret_cast<int32_t *>(base + 40 * n);
    /* candidates: split over up to twice the mean atoms per cell */
    const int64_t cells = std::max<int64_t>(p->home_cells, 1);
    const int64_t per_cell = (static_cast<int64_t>(p->nat) + cells - 1) / cells;
    std::vector<TuneShape> cand;
    for (int32_t sp : {1, 2, 4, 8}) {
        if (sp > 1 && sp > 2 * per_cell) continue;
        if (sp > 1 && p->batch_ready) continue; /* batches split per system */
        for (int32_t ch : {4, 5}) cand.push_back(TuneShape{sp, ch});
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MNL_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail("hipEventCreate failed");
    }
    const int32_t saved_split = p->force_split, saved_chunks = p->force_chunks;
    TuneShape best{1, 4};
    float best_t = std::numeric_limits<float>::infinity();
    int status = MNL_SUCCESS;
    for (const TuneShape &c : cand) {
        p->force_split = c.split;
        p->force_chunks = c.chunks;
        float t = std::numeric_limits<float>::infinity();
        for (int rep = 0; rep < 4 && status == MNL_SUCCESS; rep++) {
            /* one warm-up, then the fastest of three */
            hipError_t e = hipEventRecord(e0, st);
            if (e == hipSuccess && p->nat > 0)
                e = launch_fill_i32(p->scount.as<int32_t>(), p->nat, 0, st);
            if (e != hipSuccess) {
                status = fail(std::string("autotune: ") + hipGetErrorString(e));
                break;
            }
            if (enqueue_count_scan(p, st, nullptr) != MNL_SUCCESS ||
                enqueue_fill(p, st, 1, oi, oj, oD, od, oS) != MNL_SUCCESS) {
                status = MNL_ERROR;
                break;
            }
            float ms = 0.0f;
            e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                status = fail(std::string("autotune: ") + hipGetErrorString(e));
                break;
            }
            if (rep > 0) t = std::min(t, ms);
        }
        if (status != MNL_SUCCESS) break;
        if (t < best_t) {
            best_t = t;
            best = c;
        }
    }
    p->force_split = saved_split;
    p->force_chunks = saved_chunks;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (status != MNL_SUCCESS) return status;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune[tune_key(p)] = best;
    }
    if (split) *split = best.split;
    if (chunks) *chunks = best.chunks;
    if (best_ms) *best_ms = best_t;
    return MNL_SUCCESS;
}

void mnl_autotune_clear(void) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune.clear();
}

int mnl_first_neighbours(void *stream, int32_t n, int64_t nn, const int32_t *i, int32_t *seed) {
    if (n < 0 || nn < 0) return fail("negative size");
    if (nn > INT32_MAX) return fail("first_neighbours: more than 2^31 - 1 pairs");
    if (!seed || (nn > 0 && !i)) return fail("first_neighbours: NULL pointer");
    if (check_device_available() != MNL_SUCCESS) return MNL_ERROR;
    MNL_HIP(launch_first_neighbours(i, nn, n, seed, static_cast<hipStream_t>(stream)));
    return MNL_SUCCESS;
}

int mnl_plan_count_distinct(mnl_plan *p, void *stream, int64_t nn, const int32_t *sorted,
                            int32_t *ndistinct) {
    if (!p) return fail("plan is NULL");
    if (!ndistinct) return fail("ndistinct is NULL");
    if (nn > 0 && !sorted) return fail("sorted is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (p->tri_flag.cap < sizeof(int64_t)) {
        MNL_HIP(hipStreamSynchronize(st));
        MNL_HIP(p->tri_flag.ensure(sizeof(int64_t)));
    }
    MNL_HIP(launch_count_jumps(sorted, nn, p->tri_flag.as<int32_t>(), st));
    MNL_HIP(hipMemcpyAsync(p->host_total, p->tri_flag.ptr, sizeof(int32_t),
                           hipMemcpyDeviceToHost, st));
    MNL_HIP(hipStreamSynchronize(st));
    int32_t jumps = 0;
    std::memcpy(&jumps, p->host_total, sizeof jumps);
    *ndistinct = jumps + 1; /* get_jump_indicies: jumps + 1, also for nn == 0 */
    return MNL_SUCCESS;
}

int mnl_plan_triplets_count(mnl_plan *p, void *stream, int32_t n_first, const int32_t *first_i,
                            int64_t nn, const double *absdist, double cutoff,
                            int64_t *ntriplets) {
    if (!p) return fail("plan is NULL");
    if (!ntriplets) return fail("ntriplets is NULL");
    if (n_first < 0 || nn < 0) return fail("negative size");
    if (n_first > 0 && !first_i) return fail("first_i is NULL");
    DeviceScope ds(p->device);
    if (!ds.ok) return fail("cannot switch to the plan's device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    p->tri_total = -1;
    const int32_t rows = n_first > 1 ? n_first - 1 : 0;
    const size_t r1 = static_cast<size_t>(rows) + 1;
    struct {
        DevBuf *b;
        size_t bytes;
    } needs[] = {{&p->tri_m, r1 * sizeof(int32_t)},
                 {&p->tri_cnt, r1 * sizeof(int64_t)},
                 {&p->tri_off, (r1 + 1) * sizeof(int64_t)},
                 {&p->tri_vidx, std::max<size_t>(1, static_cast<size_t>(nn)) * sizeof(int32_t)},
                 {&p->tri_vstart, r1 * sizeof(int32_t)},
                 {&p->tri_flag, sizeof(int64_t)},
                 {&p->partial, static_cast<size_t>(scan_partial_size(rows)) * sizeof(int64_t)}};
    bool grow = false;
    for (auto &nd : needs) grow = grow || nd.bytes > nd.b->cap;
    if (grow) {
        MNL_HIP(hipStreamSynchronize(st));
        for (auto &nd : needs) MNL_HIP(nd.b->ensure(nd.bytes));
    }
    MNL_HIP(launch_triplet_rows(first_i, rows, nn, absdist, cutoff, p->tri_m.as<int32_t>(),
                                p->tri_cnt.as<int64_t>(), p->tri_vstart.as<int32_t>(),
                                p->tri_vidx.as<int32_t>(), st));
    MNL_HIP(launch_scan_i64_to_i64(p->tri_cnt.as<int64_t>(), rows, p->tri_off.as<int64_t>(),
                                   p->partial.as<int64_t>(), p->tri_flag.as<int64_t>(), st));
    MNL_HIP(hipMemcpyAsync(p->host_total, p->tri_flag.ptr, sizeof(int64_t),
                           hipMemcpyDeviceToHost, st));
    MNL_HIP(hipStreamSynchronize(st));
    p->tri_total = *p->host_total;
    p->tri_rows = rows;
    p->tri_first = first_i;
    p->tri_all = absdist == nullptr;
    *ntriplets = p->tri_total;
    return MNL_SUCCESS;
}

int mnl_plan_triplets_fill(mnl_plan *p, void *stream, int32_t *ij, int32_t *ik) {
    if (!p) return fail("plan is NULL");
    if (p->tri_total < 0) return fail("mnl_plan_triplets_fill without mnl_plan_triplets_count");
    if (p->tri_total > 0 && (!ij || !ik)) return fail("triplet output is NULL");
    DeviceScope ds(p->device);
    /* without a distance cutoff every pair is kept: no index table */
    MNL_HIP(launch_triplet_fill(p->tri_all ? p->tri_first : p->tri_vstart.as<int32_t>(),
                                p->tri_rows, p->tri_m.as<int32_t>(),
                                p->tri_off.as<int64_t>(),
                                p->tri_all ? nullptr : p->tri_vidx.as<int32_t>(), ij, ik,
                                static_cast<hipStream_t>(stream)));
    return MNL_SUCCESS;
}

int mnl_debug_sqrt(const double *x_host, double *y_host, int64_t n) {
    if (n <= 0) return MNL_SUCCESS;
    if (check_device_available() != MNL_SUCCESS) return MNL_ERROR;
    void *dx = nullptr, *dy = nullptr;
    hipError_t e = hipMalloc(&dx, size_t(n) * 8);
    if (e == hipSuccess) e = hipMalloc(&dy, size_t(n) * 8);
    if (e == hipSuccess) e = hipMemcpy(dx, x_host, size_t(n) * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_sqrt_probe(static_cast<double *>(dx), static_cast<double *>(dy), n, nullptr);
    if (e == hipSuccess) e = hipMemcpy(y_host, dy, size_t(n) * 8, hipMemcpyDeviceToHost);
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return fail(std::string("mnl_debug_sqrt: ") + hipGetErrorString(e));
    return MNL_SUCCESS;
}

}  // extern "C"

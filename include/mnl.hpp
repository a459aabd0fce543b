/*
 * mnl C++ interface: the reference's C++ GPU API over the C ABI (mnl.h).
 *
 * Same names, fields, argument meaning and error behaviour as
 * src/libneighbours/neighbour_list_gpu.hh:26-136 (and the types of
 * types.hh:28-44, neighbour_list.hh:51-57, cell_list.hh:25), in namespace
 * mnl, so that C++ code written against matscipy's GPU entry points switches
 * by namespace (`namespace matscipy = mnl;` before the calls, or a using
 * declaration per name):
 *
 *   neighbour_list_gpu          host in, host out (NeighbourList)
 *   neighbour_list_gpu_device   result left on the device (NeighbourListDevice)
 *   neighbour_count_gpu_device  per-atom counts only, on the device
 *   neighbour_matrix_gpu_device fixed-capacity rows on the device
 *
 * Every function returns NL_SUCCESS or NL_ERROR; the message of the last
 * error on the calling thread is get_error_msg() (the reference keeps it in
 * error_string, error.hh:25-31). Device arrays are owned by the result
 * objects and freed with them (move-only). NeighbourListRequest carries two
 * fields the reference's lacks, with defaults that give its behaviour: the
 * stream the build runs on (NULL: the default stream) and the grid backend
 * (automatic: the reference's rule).
 *
 * Header-only, C++17, needs the HIP runtime headers; link libmnl.so and
 * amdhip64.
 */
#ifndef MNL_HPP
#define MNL_HPP

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "mnl.h"

namespace mnl {

/* types.hh:28-44 */
using index_t = int32_t;
using real_t = double;
using error_t = int;
constexpr error_t NL_SUCCESS = MNL_SUCCESS;
constexpr error_t NL_ERROR = MNL_ERROR;

enum Quantity : int {
    QUANTITY_FIRST = MNL_Q_FIRST,
    QUANTITY_SECOND = MNL_Q_SECOND,
    QUANTITY_DISTVEC = MNL_Q_DISTVEC,
    QUANTITY_ABSDIST = MNL_Q_ABSDIST,
    QUANTITY_SHIFT = MNL_Q_SHIFT,
};

/* cell_list.hh:25. Both orders give the same output here (mnl.h). */
enum class CellOrder { Linear, Morton };

inline const char *get_error_msg() { return mnl_last_error(); }

/* A typed view: pointer and element count (device or host memory). */
template <typename T>
struct Span {
    T *ptr = nullptr;
    size_t n = 0;
    T *data() const { return ptr; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
};

/* A device array owned by the result it belongs to (a view that the result
   object keeps valid until it is destroyed or reassigned). */
template <typename T>
struct DeviceArray {
    T *ptr = nullptr;
    size_t n = 0;
    T *data() const { return ptr; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
};

/* neighbour_list.hh:51-57: host result, requested quantities filled. */
struct NeighbourList {
    std::vector<index_t> first, secnd, shift;
    std::vector<real_t> distvec, absdist;
    index_t npairs = 0;
    Span<const index_t> first_view() const { return {first.data(), first.size()}; }
    Span<const index_t> secnd_view() const { return {secnd.data(), secnd.size()}; }
    Span<const real_t> distvec_view() const { return {distvec.data(), distvec.size()}; }
    Span<const real_t> absdist_view() const { return {absdist.data(), absdist.size()}; }
    Span<const index_t> shift_view() const { return {shift.data(), shift.size()}; }
};

/* neighbour_list_gpu.hh:49-65, plus stream and backend (defaults: the
   reference's behaviour). */
struct NeighbourListRequest {
    int quantities = 0;
    const real_t *cell_origin = nullptr;        /* [3] */
    const real_t *cell = nullptr;               /* [9] row-major */
    const real_t *inv_cell = nullptr;           /* [9] */
    const bool *pbc = nullptr;                  /* [3] */
    index_t nat = 0;
    const real_t *positions = nullptr;          /* [3*nat], host or device */
    bool positions_on_device = false;
    real_t cutoff = 0;
    const real_t *per_atom_cutoff = nullptr;    /* [nat] or null */
    const real_t *per_type_cutoff_sq = nullptr; /* [ncutoffs^2] or null */
    index_t ncutoffs = 0;
    const index_t *types = nullptr;             /* [nat] or null */
    CellOrder order = CellOrder::Linear;
    int device_id = -1;                         /* -1 = current device */
    void *stream = nullptr;                     /* hipStream_t; NULL = default */
    int backend = MNL_BACKEND_AUTO;
};

/* neighbour_list_gpu.hh:74-94: the quantities stay on the device, owned by
   this object; unrequested ones are empty. counts is filled by the count
   entry point only (the pair entry point leaves it empty). */
class NeighbourListDevice {
  public:
    DeviceArray<index_t> first, secnd, shift, counts;
    DeviceArray<real_t> distvec, absdist;
    index_t npairs = 0;

    NeighbourListDevice() = default;
    NeighbourListDevice(const NeighbourListDevice &) = delete;
    NeighbourListDevice &operator=(const NeighbourListDevice &) = delete;
    NeighbourListDevice(NeighbourListDevice &&o) noexcept { *this = std::move(o); }
    NeighbourListDevice &operator=(NeighbourListDevice &&o) noexcept {
        if (this != &o) {
            reset();
            first = o.first;
            secnd = o.secnd;
            shift = o.shift;
            counts = o.counts;
            distvec = o.distvec;
            absdist = o.absdist;
            npairs = o.npairs;
            res_ = o.res_;
            o.res_ = nullptr;
            o.clear_views();
        }
        return *this;
    }
    ~NeighbourListDevice() { reset(); }

    Span<const index_t> first_view() const { return {first.data(), first.size()}; }
    Span<const index_t> secnd_view() const { return {secnd.data(), secnd.size()}; }
    Span<const real_t> distvec_view() const { return {distvec.data(), distvec.size()}; }
    Span<const real_t> absdist_view() const { return {absdist.data(), absdist.size()}; }
    Span<const index_t> shift_view() const { return {shift.data(), shift.size()}; }

    /* Hand one quantity ('i', 'j', 'D', 'd', 'S' or 'c') to a DLPack
       consumer, which then owns its buffer (the array here becomes empty);
       NULL if the quantity is absent. */
    DLManagedTensor *export_dlpack(char quantity) {
        if (!res_) return nullptr;
        DLManagedTensor *t = mnl_result_export_dlpack(res_, quantity);
        if (t) view_of(quantity);
        return t;
    }

    /* The underlying C result (owned here). */
    const mnl_result *result() const { return res_; }

    /* Takes ownership of a C result and points the arrays at it. */
    void adopt(mnl_result *res) {
        reset();
        res_ = res;
        npairs = static_cast<index_t>(mnl_result_npairs(res));
        const size_t p = static_cast<size_t>(mnl_result_npairs(res));
        const size_t nat = static_cast<size_t>(mnl_result_nat(res));
        first = array<index_t>('i', p);
        secnd = array<index_t>('j', p);
        distvec = array<real_t>('D', 3 * p);
        absdist = array<real_t>('d', p);
        shift = array<index_t>('S', 3 * p);
        counts = array<index_t>('c', nat);
    }

  private:
    mnl_result *res_ = nullptr;

    template <typename T>
    DeviceArray<T> array(char q, size_t n) const {
        T *ptr = static_cast<T *>(mnl_result_data(res_, q));
        return ptr ? DeviceArray<T>{ptr, n} : DeviceArray<T>{};
    }
    void view_of(char q) {
        switch (q) {
        case 'i': first = {}; break;
        case 'j': secnd = {}; break;
        case 'D': distvec = {}; break;
        case 'd': absdist = {}; break;
        case 'S': shift = {}; break;
        case 'c': counts = {}; break;
        default: break;
        }
    }
    void clear_views() {
        first = secnd = shift = counts = {};
        distvec = absdist = {};
        npairs = 0;
    }
    void reset() {
        if (res_) mnl_result_free(res_);
        res_ = nullptr;
        clear_views();
    }
};

/* neighbour_list_gpu.hh:113-126: rows of max_neighbours slots on the device,
   owned by this object. */
class NeighbourMatrixDevice {
  public:
    index_t n = 0;
    index_t max_neighbours = 0;
    DeviceArray<index_t> idx;   /* [n*K] */
    DeviceArray<real_t> dist;   /* [n*K*3] */
    DeviceArray<index_t> count; /* [n] */
    bool overflow = false;

    NeighbourMatrixDevice() = default;
    NeighbourMatrixDevice(const NeighbourMatrixDevice &) = delete;
    NeighbourMatrixDevice &operator=(const NeighbourMatrixDevice &) = delete;
    NeighbourMatrixDevice(NeighbourMatrixDevice &&o) noexcept { *this = std::move(o); }
    NeighbourMatrixDevice &operator=(NeighbourMatrixDevice &&o) noexcept {
        if (this != &o) {
            release();
            n = o.n;
            max_neighbours = o.max_neighbours;
            idx = o.idx;
            dist = o.dist;
            count = o.count;
            overflow = o.overflow;
            o.idx = {};
            o.dist = {};
            o.count = {};
        }
        return *this;
    }
    ~NeighbourMatrixDevice() { release(); }

    /* Device buffers for n rows of k slots (called by the entry point). */
    error_t allocate(index_t rows, index_t k) {
        release();
        const size_t r = static_cast<size_t>(rows > 0 ? rows : 0);
        const size_t s = r * static_cast<size_t>(k > 0 ? k : 0);
        void *a = nullptr, *b = nullptr, *c = nullptr;
        if (hipMalloc(&a, 4 * (s ? s : 1)) != hipSuccess ||
            hipMalloc(&b, 24 * (s ? s : 1)) != hipSuccess ||
            hipMalloc(&c, 4 * (r ? r : 1)) != hipSuccess) {
            (void)hipFree(a);
            (void)hipFree(b);
            (void)hipFree(c);
            return NL_ERROR;
        }
        n = rows;
        max_neighbours = k;
        idx = {static_cast<index_t *>(a), s};
        dist = {static_cast<real_t *>(b), 3 * s};
        count = {static_cast<index_t *>(c), r};
        return NL_SUCCESS;
    }

  private:
    void release() {
        (void)hipFree(idx.ptr);
        (void)hipFree(dist.ptr);
        (void)hipFree(count.ptr);
        idx = {};
        dist = {};
        count = {};
    }
};

namespace detail {

inline void pbc_of(const bool *pbc, int32_t out[3]) {
    for (int k = 0; k < 3; k++) out[k] = (pbc && pbc[k]) ? 1 : 0;
}

inline mnl_request c_request(const NeighbourListRequest &rq, int32_t pbc[3]) {
    pbc_of(rq.pbc, pbc);
    mnl_request m{};
    m.quantities = rq.quantities;
    m.cell_origin = rq.cell_origin;
    m.cell = rq.cell;
    m.inv_cell = rq.inv_cell;
    m.pbc = rq.pbc ? pbc : nullptr;
    m.nat = rq.nat;
    m.positions = rq.positions;
    m.positions_on_device = rq.positions_on_device ? 1 : 0;
    m.cutoff = rq.cutoff;
    m.per_atom_cutoff = rq.per_atom_cutoff;
    m.per_type_cutoff_sq = rq.per_type_cutoff_sq;
    m.ncutoffs = rq.ncutoffs;
    m.types = rq.types;
    m.order = rq.order == CellOrder::Morton ? MNL_ORDER_MORTON : MNL_ORDER_LINEAR;
    m.device_id = rq.device_id;
    m.stream = rq.stream;
    m.backend = rq.backend;
    return m;
}

template <typename T>
inline error_t to_host(const mnl_result *res, char q, size_t n, std::vector<T> &v) {
    v.clear();
    if (!mnl_result_data(res, q)) return NL_SUCCESS;
    v.resize(n);
    return n ? mnl_result_copy_to_host(res, q, v.data()) : NL_SUCCESS;
}

}  // namespace detail

/* neighbour_list_gpu.hh:28-34: host pointers in, host vectors out. */
inline error_t neighbour_list_gpu(int quantities, const real_t cell_origin[3],
                                  const real_t cell[9], const real_t inv_cell[9],
                                  const bool pbc[3], index_t nat, const real_t *r,
                                  real_t cutoff, const real_t *per_atom_cutoff,
                                  const real_t *per_type_cutoff_sq, index_t ncutoffs,
                                  const index_t *types, NeighbourList &out,
                                  CellOrder order = CellOrder::Linear) {
    NeighbourListRequest rq;
    rq.quantities = quantities;
    rq.cell_origin = cell_origin;
    rq.cell = cell;
    rq.inv_cell = inv_cell;
    rq.pbc = pbc;
    rq.nat = nat;
    rq.positions = r;
    rq.cutoff = cutoff;
    rq.per_atom_cutoff = per_atom_cutoff;
    rq.per_type_cutoff_sq = per_type_cutoff_sq;
    rq.ncutoffs = ncutoffs;
    rq.types = types;
    rq.order = order;
    int32_t pb[3];
    const mnl_request m = detail::c_request(rq, pb);
    mnl_result *res = nullptr;
    if (mnl_neighbour_list_request(&m, &res) != MNL_SUCCESS) return NL_ERROR;
    const size_t p = static_cast<size_t>(mnl_result_npairs(res));
    error_t e = detail::to_host(res, 'i', p, out.first);
    if (e == NL_SUCCESS) e = detail::to_host(res, 'j', p, out.secnd);
    if (e == NL_SUCCESS) e = detail::to_host(res, 'D', 3 * p, out.distvec);
    if (e == NL_SUCCESS) e = detail::to_host(res, 'd', p, out.absdist);
    if (e == NL_SUCCESS) e = detail::to_host(res, 'S', 3 * p, out.shift);
    out.npairs = static_cast<index_t>(p);
    mnl_result_free(res);
    return e;
}

/* neighbour_list_gpu.hh:96-97 */
inline error_t neighbour_list_gpu_device(const NeighbourListRequest &req,
                                         NeighbourListDevice &out) {
    int32_t pb[3];
    const mnl_request m = detail::c_request(req, pb);
    mnl_result *res = nullptr;
    if (mnl_neighbour_list_request(&m, &res) != MNL_SUCCESS) return NL_ERROR;
    out.adopt(res);
    return NL_SUCCESS;
}

/* neighbour_list_gpu.hh:101-102 (req.quantities is ignored) */
inline error_t neighbour_count_gpu_device(const NeighbourListRequest &req,
                                          NeighbourListDevice &out) {
    int32_t pb[3];
    const mnl_request m = detail::c_request(req, pb);
    mnl_result *res = nullptr;
    if (mnl_neighbour_count_request(&m, &res) != MNL_SUCCESS) return NL_ERROR;
    out.adopt(res);
    return NL_SUCCESS;
}

/* neighbour_list_gpu.hh:128-130: rows clipped at max_neighbours, overflow set
   if any atom has more (the call still succeeds), unused slots 0. */
inline error_t neighbour_matrix_gpu_device(const NeighbourListRequest &req,
                                           index_t max_neighbours,
                                           NeighbourMatrixDevice &out) {
    int dev = req.device_id;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return NL_ERROR;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(dev) != hipSuccess) return NL_ERROR;
    error_t e = NL_SUCCESS;
    mnl_plan *plan = nullptr;
    double *pos = nullptr;
    const double *dpos = req.positions;
    const size_t nat = static_cast<size_t>(req.nat > 0 ? req.nat : 0);
    hipStream_t st = static_cast<hipStream_t>(req.stream);
    if (!req.positions_on_device && nat > 0 && req.positions) {
        if (hipMalloc(&pos, 24 * nat) != hipSuccess ||
            hipMemcpyAsync(pos, req.positions, 24 * nat, hipMemcpyHostToDevice, st) !=
                hipSuccess)
            e = NL_ERROR;
        dpos = pos;
    }
    if (e == NL_SUCCESS && mnl_plan_create(dev, &plan) != MNL_SUCCESS) e = NL_ERROR;
    if (e == NL_SUCCESS)
        e = mnl_plan_set_cell_order(plan, req.order == CellOrder::Morton ? MNL_ORDER_MORTON
                                                                          : MNL_ORDER_LINEAR);
    if (e == NL_SUCCESS && max_neighbours >= 0) e = out.allocate(req.nat, max_neighbours);
    int32_t ovf = 0;
    if (e == NL_SUCCESS) {
        int32_t pb[3];
        detail::pbc_of(req.pbc, pb);
        e = mnl_plan_matrix(plan, st, req.cell_origin, req.cell, req.inv_cell,
                            req.pbc ? pb : nullptr, req.nat, dpos, req.cutoff,
                            req.per_atom_cutoff, req.per_type_cutoff_sq, req.ncutoffs,
                            req.types, req.backend, max_neighbours, out.idx.data(),
                            out.dist.data(), out.count.data(), &ovf);
    }
    out.overflow = ovf != 0;
    mnl_plan_destroy(plan);
    if (pos) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(pos);
    }
    (void)hipSetDevice(prev);
    return e;
}

}  // namespace mnl

#endif

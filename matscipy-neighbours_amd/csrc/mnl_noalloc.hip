/*
 * The allocation-free build, mnl_neighbour_list_noalloc (include/mnl.h).
 *
 * Its workspaces (binning records, cell grid, hash keys, counts, offsets,
 * scan partials, parameters) live in one file-scope device array,
 * `scratch`, of kNoallocScratchBytes (24 MiB): every GPU the library runs on
 * holds its own copy until the process ends. Host code cannot use the
 * array's name as a device address, so the first call on a device, with
 * that device current, takes the array's address there with
 * hipGetSymbolAddress and keeps it per device; the kernels receive that
 * address, as they receive a plan's buffers. The call allocates nothing and
 * never waits for the device.
 *
 * Every call on one GPU shares that GPU's array, and the array keeps the
 * last call's contents (each build clears what it needs cleared, with
 * kernels). Calls that overlap on one GPU - on different streams, or from
 * different threads without ordering - race on it.
 */
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "mnl.h"
#include "mnl_engine.hh"

namespace {

static __device__ __align__(256) unsigned char scratch[mnl::kNoallocScratchBytes];

std::mutex g_scratch_mutex;
void *g_scratch_ptr[mnl::kNoallocMaxDevices] = {};

/* The array's address on `device` (the current device), taken once. */
void *scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

}  // namespace

extern "C" int mnl_neighbour_list_noalloc(void *stream, const double origin[3],
                                          const double cell[9], const double inv_cell[9],
                                          const int32_t pbc[3], int32_t nat,
                                          const double *positions, double cutoff,
                                          const double *per_atom_cutoff, int backend,
                                          int64_t capacity, int32_t *first, int32_t *secnd,
                                          void *distvec, void *absdist, int32_t *shift,
                                          int float32_output, int64_t *npairs_device) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return mnl::set_error("no HIP device available: the mnl engine runs only on the GPU");
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return mnl::set_error("hipGetDevice failed");
    if (device < 0 || device >= mnl::kNoallocMaxDevices)
        return mnl::set_error("device index beyond the allocation-free build's table");
    void *ptr = scratch_on(device);
    if (!ptr) return mnl::set_error("hipGetSymbolAddress of the allocation-free workspace failed");
    return mnl::noalloc_build(device, ptr, static_cast<hipStream_t>(stream), origin, cell,
                              inv_cell, pbc, nat, positions, cutoff, per_atom_cutoff, backend,
                              capacity, first, secnd, distvec, absdist, shift, float32_output,
                              npairs_device);
}

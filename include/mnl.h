/*
 * mnl — MI355X-native neighbour-list engine, C ABI.
 *
 * Drop-in for the device path of matscipy-neighbours'
 * `neighbour_list("ijdDS", ...)` and its src/libneighbours C++ API. All
 * compute runs in hand-written gfx950 HIP kernels; there is no CPU fallback:
 * every entry point returns MNL_ERROR with a message when no GPU is usable.
 *
 * Conventions (kept from the reference):
 *   - error codes: 0 on success, -1 on error; the message is read with
 *     mnl_last_error() (thread-local here, a GIL-protected global in the
 *     reference: src/libneighbours/error.hh:17-33);
 *   - quantities: bit flags i=1, j=2, D=4, d=8, S=16
 *     (src/libneighbours/types.hh:37-43);
 *   - cell is row-major with rows = lattice vectors; inv_cell is taken as
 *     given by the caller (fractional = inv_cell . (r - origin));
 *   - output pairs are sorted by i, and within one i in the reference's
 *     traversal order, bit-identical to the reference CPU path;
 *     D == r[j] - r[i] + S @ cell.
 *   - pbc is three int32 (0/1) instead of three C++ bools.
 *
 * Pointers marked [device] must be device memory on the plan's/result's
 * device; [host] must be host memory. Inputs are non-owning and need only
 * live for the duration of the call (plus, for asynchronous plan calls,
 * until the stream reaches the enqueued work).
 */
#ifndef MNL_H
#define MNL_H

#include <stdint.h>

#include "mnl_dlpack.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MNL_SUCCESS 0
#define MNL_ERROR (-1)

enum {
    MNL_Q_FIRST = 1,   /* i */
    MNL_Q_SECOND = 2,  /* j */
    MNL_Q_DISTVEC = 4, /* D */
    MNL_Q_ABSDIST = 8, /* d */
    MNL_Q_SHIFT = 16   /* S */
};

/* Cell-grid backend: AUTO applies the reference rule (cell_list.cc:193-196),
   hashed when ncells > max(2^20, 8*nat); DENSE and HASHED force one. The
   backend never changes the output. */
enum { MNL_BACKEND_AUTO = -1, MNL_BACKEND_DENSE = 0, MNL_BACKEND_HASHED = 1 };

/* Order of the atoms inside the dense cell grid (the reference's CellOrder,
   src/libneighbours/cell_list.hh:25; neighbour_list.hh:98-104,140,
   neighbour_list_gpu.hh:34,63). The reference's Morton build sorts the atoms
   by their cell's Z-curve key (cell_list.cc:81-102 on the CPU with an
   unstable std::sort, neighbour_list_gpu.cc:396-404 on the GPU with a stable
   radix sort over the atom indices), so its cells hold their atoms in an
   order the CPU leaves implementation-defined. Here every cell holds its
   atoms in ascending index under both orders, the order of the reference's
   stable GPU sort and of its Linear build: the pairs, their order and their
   bits are the same for MNL_ORDER_LINEAR and MNL_ORDER_MORTON. The locality
   the Morton curve is for comes from the engine's own cell layout, which
   follows the input's atom order. */
enum { MNL_ORDER_LINEAR = 0, MNL_ORDER_MORTON = 1 };

/* Message of the last error on this thread ("" if none). */
const char *mnl_last_error(void);

/* Library version string. */
const char *mnl_version(void);

/* Number of usable HIP devices (0 if none). Never fails. */
int mnl_device_count(void);

/* ------------------------------------------------------------------------
 * One-shot entry points (reference: neighbour_list_gpu.hh).
 * ---------------------------------------------------------------------- */

typedef struct mnl_result mnl_result;

/*
 * Build the neighbour list on the GPU and keep the requested quantities in
 * device memory, owned by the returned result.
 * Replaces matscipy::neighbour_list_gpu_device
 * (src/libneighbours/neighbour_list_gpu.hh:96, request fields :49-65).
 *   positions          [3*nat] f64, [device] if positions_on_device else [host]
 *   per_atom_cutoff    [nat] radii (pair cutoff = sum) [host] or NULL
 *   per_type_cutoff_sq [ncutoffs^2] squared cutoffs [host] or NULL
 *   types              [nat] int32 [host] or NULL
 *   cutoff             binning / global cutoff (max pair cutoff for arrays)
 *   device_id          device to run on; -1 = current device
 */
int mnl_neighbour_list_device(int quantities, const double origin[3],
                              const double cell[9], const double inv_cell[9],
                              const int32_t pbc[3], int32_t nat,
                              const double *positions, int positions_on_device,
                              double cutoff, const double *per_atom_cutoff,
                              const double *per_type_cutoff_sq, int32_t ncutoffs,
                              const int32_t *types, int device_id,
                              mnl_result **out);

/*
 * Per-atom neighbour counts only (coordination), never materialising pairs.
 * Replaces matscipy::neighbour_count_gpu_device
 * (src/libneighbours/neighbour_list_gpu.hh:101-102). The result holds
 * quantity 'c' (int32 [nat]) and npairs.
 */
int mnl_neighbour_count_device(const double origin[3], const double cell[9],
                               const double inv_cell[9], const int32_t pbc[3],
                               int32_t nat, const double *positions,
                               int positions_on_device, double cutoff,
                               const double *per_atom_cutoff,
                               const double *per_type_cutoff_sq,
                               int32_t ncutoffs, const int32_t *types,
                               int device_id, mnl_result **out);

/*
 * The request form of the two entry points above, field for field the
 * reference's NeighbourListRequest (src/libneighbours/neighbour_list_gpu.hh:49-65)
 * plus the stream and the grid backend. Replaces
 * neighbour_list_gpu_device(const NeighbourListRequest&, NeighbourListDevice&)
 * (:96) and neighbour_count_gpu_device (:101-102).
 *   stream   hipStream_t the build runs on (NULL = default stream); host
 *            inputs are copied on it, and the call returns once the result
 *            is complete on it
 *   order    MNL_ORDER_LINEAR or MNL_ORDER_MORTON (same output)
 *   backend  MNL_BACKEND_* (AUTO: the reference's rule)
 * Set unused fields to zero / NULL (backend AUTO is -1).
 */
typedef struct mnl_request {
    int quantities;
    const double *cell_origin;        /* [3] */
    const double *cell;               /* [9] row-major, rows = lattice vectors */
    const double *inv_cell;           /* [9] as given */
    const int32_t *pbc;               /* [3] */
    int32_t nat;
    const double *positions;          /* [3*nat], host or device */
    int positions_on_device;
    double cutoff;
    const double *per_atom_cutoff;    /* [nat] [host] or NULL */
    const double *per_type_cutoff_sq; /* [ncutoffs^2] [host] or NULL */
    int32_t ncutoffs;
    const int32_t *types;             /* [nat] [host] or NULL */
    int order;
    int device_id;                    /* -1 = current device */
    void *stream;
    int backend;
} mnl_request;

int mnl_neighbour_list_request(const mnl_request *request, mnl_result **out);
int mnl_neighbour_count_request(const mnl_request *request, mnl_result **out);

/* Number of pairs in a result. */
int64_t mnl_result_npairs(const mnl_result *res);

/* Atom count of a result. */
int32_t mnl_result_nat(const mnl_result *res);

/* Device the result lives on. */
int mnl_result_device(const mnl_result *res);

/* [device] pointer of one quantity: 'i','j' int32 [P]; 'D' f64 [P,3];
   'd' f64 [P]; 'S' int32 [P,3]; 'c' int32 [nat] counts. NULL if absent. */
void *mnl_result_data(const mnl_result *res, char quantity);

/* Copy one quantity to host memory `dst` (sized for the whole array).
   Together with a host-positions build this is the host-in/host-out call
   matscipy::neighbour_list_gpu (neighbour_list_gpu.hh:28-34). */
int mnl_result_copy_to_host(const mnl_result *res, char quantity, void *dst);

/*
 * Export one quantity as an unversioned DLPack tensor (device kROCM) that
 * owns its buffer; the result gives the buffer up (the quantity becomes
 * absent in `res`). The consumer calls tensor->deleter when done.
 * Replaces the reference's device_capsule export
 * (language_bindings/python/bind_py_dlpack.cc:66-112).
 */
DLManagedTensor *mnl_result_export_dlpack(mnl_result *res, char quantity);

/* Free a result and the device buffers it still owns. NULL is a no-op.
   Output buffers are allocated at size classes (rounded up to 1/8 of their
   highest power of two, at least 64 KiB). A freed buffer (here or by an
   exported tensor's DLPack deleter) is kept, once the device has finished
   the work queued before the free, for the next result whose buffer has its
   class on its device; at most 16 GiB is kept in all. An allocation that
   finds the device out of memory returns that device's kept buffers first
   and tries once more. */
void mnl_result_free(mnl_result *res);

/* Return every kept result buffer to the device (no counterpart in the
   reference, which keeps none). */
void mnl_release_cached_memory(void);

/* Bytes of result buffers kept for `device` (every device when negative);
   no counterpart in the reference. */
int64_t mnl_cached_bytes(int device);

/* ------------------------------------------------------------------------
 * Plan API: reusable device workspace, caller-owned outputs, explicit stream.
 * This is the zero-copy path used by the Python binding (outputs allocated by
 * the caller, e.g. from a caching allocator) and by MD loops.
 * ---------------------------------------------------------------------- */

typedef struct mnl_plan mnl_plan;

/* Create a plan on `device_id` (-1 = current). Allocates only small state;
   workspaces grow on first use and are reused across calls. */
int mnl_plan_create(int device_id, mnl_plan **plan);
void mnl_plan_destroy(mnl_plan *plan);

/* Device of a plan. */
int mnl_plan_device(const mnl_plan *plan);

/* CellOrder of the plan's later builds (MNL_ORDER_LINEAR, the default, or
   MNL_ORDER_MORTON; the output is the same for both, see above), the
   plan-side counterpart of the order argument of the reference's
   neighbour_list / neighbour_matrix (neighbour_list.hh:98-104,134-140). */
int mnl_plan_set_cell_order(mnl_plan *plan, int order);
int mnl_plan_cell_order(const mnl_plan *plan);

/*
 * Phase 1: bin the atoms, count every atom's neighbours and scan the counts
 * into pair offsets, on `stream` (a hipStream_t; NULL = default stream).
 * Waits for the stream and returns the pair count in *npairs.
 * Same arguments as mnl_neighbour_list_device except that positions,
 * per_atom_cutoff and types are [device] pointers; per_type_cutoff_sq is
 * [host]. `backend` is one of MNL_BACKEND_*.
 * Replaces the count half of count_and_fill
 * (src/libneighbours/neighbour_list_gpu.cc:264-285) and the binning of
 * build_device (:309-462).
 */
int mnl_plan_count(mnl_plan *plan, void *stream, const double origin[3],
                   const double cell[9], const double inv_cell[9],
                   const int32_t pbc[3], int32_t nat, const double *positions,
                   double cutoff, const double *per_atom_cutoff,
                   const double *per_type_cutoff_sq, int32_t ncutoffs,
                   const int32_t *types, int backend, int64_t *npairs);

/*
 * Phase 2: write the pairs of the last mnl_plan_count into caller [device]
 * buffers sized for npairs (NULL = quantity not wanted). Asynchronous on
 * `stream`. Replaces the fill half of count_and_fill
 * (neighbour_list_gpu.cc:286-305).
 */
int mnl_plan_fill(mnl_plan *plan, void *stream, int32_t *first, int32_t *secnd,
                  double *distvec, double *absdist, int32_t *shift);

/* As mnl_plan_fill, with D and d written as float32 (pairs still decided in
   float64). No counterpart in the reference: the FP32 output option. */
int mnl_plan_fill_f32(mnl_plan *plan, void *stream, int32_t *first,
                      int32_t *secnd, float *distvec, float *absdist,
                      int32_t *shift);

/*
 * Asynchronous phase 1 for steady-state loops and graph capture: as
 * mnl_plan_count but never waits for the stream and never copies from host
 * memory once the workspaces are large enough (the first build of a given
 * size, outside capture, sizes them; growing under capture is an error).
 * The pair count is written to *npairs_device ([device] int64, may be NULL).
 * A per-type cutoff table is uploaded (waiting for the stream) only when it
 * differs from the previous one, so only its first build is not capturable.
 */
int mnl_plan_count_async(mnl_plan *plan, void *stream, const double origin[3],
                         const double cell[9], const double inv_cell[9],
                         const int32_t pbc[3], int32_t nat, const double *positions,
                         double cutoff, const double *per_atom_cutoff,
                         const double *per_type_cutoff_sq, int32_t ncutoffs,
                         const int32_t *types, int backend, int64_t *npairs_device);

/*
 * Asynchronous phase 2 with a caller-chosen capacity: pairs at positions
 * >= capacity are not written (compare *npairs_device with the capacity to
 * detect overflow and retry larger). D and d are float32 when
 * float32_output is non-zero, float64 otherwise.
 */
int mnl_plan_fill_async(mnl_plan *plan, void *stream, int64_t capacity,
                        int32_t *first, int32_t *secnd, void *distvec,
                        void *absdist, int32_t *shift, int float32_output);

/* Copy the per-atom neighbour counts of the last count into [device] int32
   [nat] (coordination; asynchronous on `stream`). */
int mnl_plan_counts(mnl_plan *plan, void *stream, int32_t *counts);

/* [device] pointer to the plan's int64 [nat+1] pair offsets (CSR row
   pointers of the i-sorted list; valid until the next count). */
const int64_t *mnl_plan_offsets(const mnl_plan *plan);

/* Backend chosen by the last count (MNL_BACKEND_DENSE or _HASHED). */
int mnl_plan_backend(const mnl_plan *plan);

/*
 * Dense fixed-capacity neighbour matrix (the reference's neighbour_matrix,
 * src/libneighbours/neighbour_list.cc:206-254, and its device form
 * neighbour_list_gpu.cc:523-570): row i holds atom i's neighbours in the
 * pair list's order, idx [nat*K] int32 = j, dist [nat*K*3] float64 = D,
 * unused slots 0; count [nat] int32 = true neighbour count (rows keep only
 * the first K). Same geometry arguments as mnl_plan_count. One traversal,
 * no count pass. All outputs [device], caller-owned.
 * mnl_plan_matrix waits and sets *overflow [host] to 1 if some atom has
 * more than K neighbours (the reference raises then); the async form writes
 * that flag to *overflow_device ([device] int32, may be NULL).
 */
int mnl_plan_matrix(mnl_plan *plan, void *stream, const double origin[3],
                    const double cell[9], const double inv_cell[9], const int32_t pbc[3],
                    int32_t nat, const double *positions, double cutoff,
                    const double *per_atom_cutoff, const double *per_type_cutoff_sq,
                    int32_t ncutoffs, const int32_t *types, int backend,
                    int32_t max_neighbours, int32_t *idx, double *dist, int32_t *count,
                    int32_t *overflow);
int mnl_plan_matrix_async(mnl_plan *plan, void *stream, const double origin[3],
                          const double cell[9], const double inv_cell[9],
                          const int32_t pbc[3], int32_t nat, const double *positions,
                          double cutoff, const double *per_atom_cutoff,
                          const double *per_type_cutoff_sq, int32_t ncutoffs,
                          const int32_t *types, int backend, int32_t max_neighbours,
                          int32_t *idx, double *dist, int32_t *count,
                          int32_t *overflow_device);

/*
 * Allocation-free build: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). One call does what
 * mnl_plan_count_async followed by mnl_plan_fill_async does, on the current
 * device and on `stream`, but every workspace lives in a fixed 24 MiB device
 * array inside the library (each GPU the library runs on holds one until the
 * process ends), so it calls no hipMalloc or hipFree and never waits for the
 * device. Limits: nat <= 131072 and at most 524288 dense cells (or hash
 * slots); larger builds are refused with an error. Cutoffs: the scalar
 * cutoff, or per-atom radii per_atom_cutoff [device] (pair cutoff = sum of
 * radii, binning cutoff = `cutoff`); no per-type tables. Pairs at positions
 * >= capacity are not written; the pair total goes to *npairs_device
 * ([device] int64, may be NULL), as in mnl_plan_fill_async. D and d are
 * float32 when float32_output != 0. Output order and bits are those of the
 * plan API. All calls on one GPU share its array, which keeps the last
 * call's contents: calls that overlap on one GPU (other streams, other
 * threads) race; order them on one stream.
 */
int mnl_neighbour_list_noalloc(void *stream, const double origin[3], const double cell[9],
                               const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                               const double *positions, double cutoff,
                               const double *per_atom_cutoff, int backend, int64_t capacity,
                               int32_t *first, int32_t *secnd, void *distvec, void *absdist,
                               int32_t *shift, int float32_output, int64_t *npairs_device);

/* ---------------------------------------------------------------- batches
 *
 * Many independent systems in one build: atoms concatenated in system
 * order, each system with its own cell, pbc and scalar cutoff; every
 * system's pairs are exactly those of a single-system build of it (indices
 * i, j local to the system), concatenated in system order. There is no
 * counterpart in the reference (whose API builds one system per call); the
 * single-system semantics each system follows are those of
 * src/libneighbours/neighbour_list.hh:98-104.
 */

/*
 * Upload the geometry of a batch of nsys systems into the plan (host
 * arrays): atom_offsets [nsys+1] (first atom of each system, ascending,
 * atom_offsets[0] == 0), origins [nsys*3] (NULL = all zero), cells and
 * inv_cells [nsys*9] (row-major, lattice vectors as rows; inv_cell as given,
 * fractional = inv_cell . (r - origin)), pbcs [nsys*3] int32, cutoffs
 * [nsys]. Grows the workspaces and waits for `stream`. The geometry stays in
 * the plan for any number of batch counts until the plan is used for a
 * single-system count.
 */
int mnl_plan_set_batch(mnl_plan *plan, void *stream, int32_t nsys,
                       const int32_t *atom_offsets, const double *origins,
                       const double *cells, const double *inv_cells,
                       const int32_t *pbcs, const double *cutoffs);

/* Count of a batch (positions: [device] float64 [N][3], all systems
   concatenated): waits for the stream and returns the total pair count. */
int mnl_plan_batch_count(mnl_plan *plan, void *stream, const double *positions,
                         int64_t *npairs);

/* Asynchronous batch count: never waits and never copies from the host;
   the total goes to *npairs_device ([device] int64, may be NULL).
   Capturable. Fill with mnl_plan_fill_async / mnl_plan_fill. */
int mnl_plan_batch_count_async(mnl_plan *plan, void *stream,
                               const double *positions, int64_t *npairs_device);

/* [device] int64 [nsys+1]: the first pair of every system in the
   concatenated output, and the total at index nsys (asynchronous). Works
   for a single-system count too (nsys = 1). */
int mnl_plan_system_offsets(mnl_plan *plan, void *stream,
                            int64_t *system_pair_offsets);

/* Systems in the plan's last count or batch. */
int32_t mnl_plan_nsys(const mnl_plan *plan);

/* A batch build (count, fill into caller buffers of `capacity` pairs,
   per-system pair offsets, total pair count) captured once as a HIP graph,
   replayed by mnl_graph_launch with one launch. Replays read the positions
   buffer in place (update it between replays); the batch geometry and the
   output buffers are fixed at capture. */
typedef struct mnl_graph mnl_graph;
int mnl_plan_batch_graph_create(mnl_plan *plan, const double *positions,
                                int64_t capacity, int32_t *first, int32_t *secnd,
                                void *distvec, void *absdist, int32_t *shift,
                                int float32_output, int64_t *system_pair_offsets,
                                int64_t *npairs_device, mnl_graph **graph);
int mnl_graph_launch(mnl_graph *graph, void *stream);
void mnl_graph_destroy(mnl_graph *graph);

/* ---------------------------------------------------------------- gradients
 *
 * Backward of a pair list, no counterpart in the reference (which has no
 * derivatives): with D_p = r[j] - r[i] + S_p . cell and d_p = |D_p| over an
 * i-sorted list of npairs pairs (first/secnd/shift as written by the fill),
 * given dL/dD (grad_distvec [npairs*3], may be NULL) and dL/dd
 * (grad_absdist [npairs], may be NULL; then distvec and absdist are
 * required), writes dL/dr into grad_positions [nat*3] float64 and dL/dcell
 * (cell rows = lattice vectors, dL/dcell[a][b] = sum_p S_pa g_pb over the
 * periodic directions a) into grad_cell [9] float64, where
 * g_p = dL/dD_p + dL/dd_p D_p / d_p (the d term is dropped for d_p = 0).
 * pbc [3] int32 [host] (NULL = all periodic) is the list's periodicity:
 * along a non-periodic direction D = r_j - r_i does not involve the cell,
 * even where the reference's S is non-zero there (an atom a box length or
 * more outside). Both outputs are overwritten. distvec,
 * absdist and the incoming gradients are float32 when float32 != 0 (the FP32
 * output option), float64 otherwise. All pointers [device]; asynchronous.
 * The cell gradient is deterministic. With deterministic == 0,
 * grad_positions sums the j-side contributions with float64 atomics
 * (order-dependent in the last bits). With deterministic != 0 every sum
 * runs in a fixed order and the result is the same bits on every call: the
 * pairs are grouped by j with a stable radix sort, each atom adds its
 * j-side pairs in ascending pair order and its own row in pair order (first
 * must be sorted, as every list this library writes is; npairs < 2^31);
 * this takes about 40 bytes of workspace per pair and is slower.
 */
int mnl_plan_pair_backward(mnl_plan *plan, void *stream, int64_t npairs, int32_t nat,
                           const int32_t *first, const int32_t *secnd,
                           const int32_t *shift, const void *distvec,
                           const void *absdist, const void *grad_distvec,
                           const void *grad_absdist, int float32,
                           const int32_t pbc[3], int deterministic,
                           double *grad_positions, double *grad_cell);

/* ------------------------------------------------------------- autotuner
 *
 * Launch shapes of the pair kernels (no counterpart in the reference, which
 * has no GPU tuning): the home atoms of a cell are dealt to `split` waves
 * (1..64) and the fill holds `chunks` (4 or 5) chunks of 64 candidates in
 * registers per window. Any shape gives the same outputs, bit for bit.
 */

/*
 * Times the plan's pair kernels (count, scan, and the fill of all five
 * quantities into internal scratch) on the binned state of its last count or
 * batch count, for each shape: split 1, 2, 4 and 8 waves per cell (up to
 * twice the mean atoms per cell) with windows of 4 and 5 chunks; one
 * warm-up, then the fastest of three runs. The fastest shape is kept in a
 * process-wide cache keyed by the configuration class (systems, atoms,
 * cells or hash slots, grid backend, cutoff form), which every later build
 * of that class uses. Waits for the stream, allocates 52 bytes of scratch per
 * pair (at most 2^30 pairs) and is refused under graph capture. split,
 * chunks and best_ms (may be NULL) receive the choice and its time.
 */
int mnl_plan_autotune(mnl_plan *plan, void *stream, int32_t *split, int32_t *chunks,
                      float *best_ms);

/* Forget every tuned shape: builds return to the built-in choice. */
void mnl_autotune_clear(void);

/* Force the plan's launch shape (split 1..64, chunks 4 or 5, or 0 for the
   built-in choice; split 0 restores the tuned or built-in shape). */
int mnl_plan_set_launch_shape(mnl_plan *plan, int32_t split, int32_t chunks);

/* ------------------------------------------------------- fused consumer */

/*
 * Lennard-Jones forces, energy and virial of one system, computed straight
 * from the cell grid without writing a pair list. The reference has no
 * such entry point: its examples/lj_langevin/lj_langevin_cpu.cc:62-98
 * computes the same forces from an "ij" list (eps = sigma = 1, open
 * boundaries), and the build is for callers that want forces rather than
 * the list. The pairs are the neighbour list's with a uniform cutoff:
 * r < cutoff, D = r_j - r_i + S.cell, images included. With
 * phi(r) = 4 eps ((sigma/r)^12 - (sigma/r)^6), unshifted:
 *   forces [nat][3] float64: f_i = sum_j coef D_ij,
 *     coef = -24 eps (2 (sigma/r)^12 - (sigma/r)^6) / r^2;
 *   energy [1] float64 (may be NULL): sum over pairs of phi;
 *   virial [3][3] float64 (may be NULL): -1/2 sum over ordered pairs of
 *     coef D (x) D, i.e. sum over pairs of r_ij (x) f_ij.
 * All pointers [device]; asynchronous. Forces are float64 sums with
 * atomics (order-dependent in the last bits); energy and virial are
 * reduced in a fixed order. The plan's count state is cleared.
 */
int mnl_plan_lj(mnl_plan *plan, void *stream, const double origin[3], const double cell[9],
                const double inv_cell[9], const int32_t pbc[3], int32_t nat,
                const double *positions, double cutoff, double epsilon, double sigma,
                double *forces, double *energy, double *virial);

/* ---------------------------------------------------------------- lists
 * Post-processing of i-sorted pair lists, all on the device.
 */

/* Row starts of an i-sorted list: seed [n+1] int32 [device] from the first
   indices i [nn] int32 [device] (asynchronous on `stream`). Rows before the
   first atom with pairs get -1, rows after the last get nn, rows without
   pairs point at the next row's start; nn == 0 gives all 0. Replaces
   matscipy::first_neighbours (src/libneighbours/first_neighbours.cc:15-40). */
int mnl_first_neighbours(void *stream, int32_t n, int64_t nn, const int32_t *i,
                         int32_t *seed);

/* Number of distinct values of a sorted int32 array [device] (waits): the n
   that get_jump_indicies (first_neighbours.cc:42-55) passes to
   first_neighbours, 1 for an empty array. */
int mnl_plan_count_distinct(mnl_plan *plan, void *stream, int64_t nn, const int32_t *sorted,
                            int32_t *ndistinct);

/* Triplets (pairs of pairs sharing the first atom) of a first-neighbour
   array, as matscipy::triplet_list (src/libneighbours/triplet_list.cc:15-34):
   for each row r < n_first - 1 every ordered (ij, ik), ij != ik, of pair
   indices in [first_i[r], first_i[r+1]), ij outer and ik inner, both
   ascending; with absdist [nn] (may be NULL) only pairs with
   !(absdist >= cutoff). Phase 1 waits and returns the triplet count; phase 2
   writes ij, ik [ntriplets] int32 [device]. first_i [n_first] int32
   [device] must stay valid until phase 2. */
int mnl_plan_triplets_count(mnl_plan *plan, void *stream, int32_t n_first,
                            const int32_t *first_i, int64_t nn, const double *absdist,
                            double cutoff, int64_t *ntriplets);
int mnl_plan_triplets_fill(mnl_plan *plan, void *stream, int32_t *ij, int32_t *ik);

/* Diagnostic, no counterpart in the reference: y = sqrt(x) for n doubles
   (host buffers) computed by a device kernel, so that tests can pin that the
   device square root rounds exactly like the host's (d == sqrt(r2)). */
int mnl_debug_sqrt(const double *x, double *y, int64_t n);

#ifdef __cplusplus
}
#endif

#endif

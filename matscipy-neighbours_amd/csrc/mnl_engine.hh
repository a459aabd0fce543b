/*
 * mnl engine internals shared by the host core and the HIP kernels.
 *
 * Data layout in HBM (one build, N atoms, C cells):
 *   raw4   int4   [N]  raw (unwrapped) cell coords + global cell id, atom order
 *   ticket int32  [N]  arrival slot inside the cell (from the histogram atomic)
 *   tmp    int32  [N]  atoms scattered into cell slices, arbitrary order
 *   spos   double4[N]  cell-sorted: bin-relative position xyz + per-atom radius
 *   sinfo  int4   [N]  cell-sorted: per direction the raw cell (periodic) or
 *                      the image count (non-periodic), + atom index (system-local)
 *   stype  int32  [N]  cell-sorted atom type (per-type cutoffs only)
 *   cell_count / cell_first int32 [C] dense CSR (or hash slots when hashed)
 *   scount int32  [N]  neighbours per atom, cell-sorted order (half count)
 *   slot   int32  [N]  cell-sorted index of each atom
 *   counts int32  [N]  neighbours per atom, original order
 *   offsets int64 [N+1] exclusive scan of counts (CSR row pointers)
 * A "system" is one periodic cell with its atoms; the plain path is a batch
 * of one system, the batched path concatenates systems (atoms, cells and
 * pairs) with per-system base offsets.
 */
#ifndef MNL_ENGINE_HH
#define MNL_ENGINE_HH

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mnl {

/* Per-system grid description, device resident, read with scalar loads. */
struct SysGeom {
    double inv[9];      /* inv_cell as given: fractional = inv . (r - origin) */
    double origin[3];
    double bin[9];      /* bin[3*k + m] = cell[3*k + m] / n[k] (cell edge vectors) */
    double cutoff_sq;
    double rn[3];       /* 1.0 / n[k] (exact division of shift numerators) */
    int32_t n[3];       /* cells per lattice direction */
    int32_t nxyz[3];    /* neighbour cells scanned each way */
    int32_t pbc[3];
    int32_t atom_base;  /* first atom of the system in the concatenated arrays */
    int32_t natoms;
    int32_t cell_base;  /* first dense cell of the system */
    int32_t ncells;
    int32_t nseg;       /* (2nx+1)(2ny+1)(2nz+1) neighbour-cell offsets */
    int32_t split;      /* batches: waves per home cell of this system */
    /* dense cell layout: cell = c[f] + n[f] * (c[m] + n[m] * c[s]) with
       (f, m, s) = lay; {0, 1, 2} (x fastest) unless the binning of the
       previous build found the input's atom order running along another
       axis (k_set_params). Internal: no output depends on it. */
    int32_t lay[3];
    /* division by n[0], n[1], n[2], 2nx+1, 2ny+1 as a multiply and shift:
       x / d == (x * dmul) >> dshift for 0 <= x < 2^31 (make_div_magic) */
    uint32_t dmul[5];
    int32_t dshift[5];
};

/* Multiplier and shift with x / d == (uint64(x) * mul) >> shift for all
   0 <= x < 2^31 (d >= 1). */
void make_div_magic(int32_t d, uint32_t &mul, int32_t &shift);

/* Host geometry (src/libneighbours/cell_list.cc:153-182 and the search range
   of neighbour_list.cc:176-178), with the reference's operation order.
   Returns false for a zero-volume cell. */
bool grid_geometry(const double cell[9], double cutoff, int32_t n[3],
                   int32_t nxyz[3], double bin[9], double len[3]);

/* Neighbour-cell offsets one home cell visits, (2nx+1)(2ny+1)(2nz+1), or -1
   for a negative reach; values above kMaxStencil come back as
   kMaxStencil + 1. The pair kernels index the stencil with int32 and need
   it at most kMaxStencil (2^30, far beyond any system that finishes:
   each home atom then tests over 10^9 cells). */
constexpr int64_t kMaxStencil = int64_t(1) << 30;
int64_t stencil_size(const int32_t nxyz[3]);

/* Fill a SysGeom for one system. Returns false for a zero-volume cell. */
bool make_sys_geom(const double origin[3], const double cell[9],
                   const double inv_cell[9], const int32_t pbc[3], double cutoff,
                   int32_t atom_base, int32_t natoms, int32_t cell_base,
                   SysGeom &g);

enum CutMode { CUT_GLOBAL = 0, CUT_PER_ATOM = 1, CUT_PER_TYPE = 2 };

/* Arguments of the pair kernels (count and fill), passed by value. */
struct PairArgs {
    const SysGeom *geo;
    const int32_t *cell_base;  /* [nsys+1] */
    int32_t nsys;
    int32_t home_cells;        /* dense: total cells; hashed: slot capacity */
    /* hashed: the occupied slots in order of their lowest atom index
       (*home_items of them, on the device); NULL: cells 0..home_cells-1 */
    const int32_t *home_list;
    const int64_t *home_items;
    /* batches: first work item of each system ([nsys+1]; item = one of a
       home cell's geo[s].split waves), and their total; NULL: uniform
       split over all home cells */
    const int32_t *item_base;
    int32_t items;
    const int32_t *cell_first;
    const int32_t *cell_count;
    const int64_t *hkey;
    int64_t hmask;
    const double4 *spos;
    const int4 *sinfo;
    const int32_t *stype;
    const double *pt_sq;
    int32_t ncut;
    int32_t xcd_blocks;        /* grid size, for the XCD-aware block remap */
    int32_t fill_chunks;       /* fill: candidate chunks per register window (4 or 5) */
    int32_t split;             /* waves per home cell, each with a share of its atoms */
    const int32_t *atom_base;  /* [nsys+1] */
    int32_t nat;               /* atoms of all systems */
    int32_t half;              /* count: visit half the offsets (symmetric test) */
    int32_t *scount;           /* half count: [N] by cell-sorted index, zeroed */
    int32_t *counts;           /* count: [N] by global atom index */
    const int64_t *offsets;    /* fill: [N+1] */
    int64_t capacity;          /* fill: pairs beyond this are not written */
    int32_t max_nb;            /* matrix fill: slots per row */
    int32_t *overflow;         /* matrix fill: set to 1 if a row exceeds max_nb */
    int32_t *out_i;
    int32_t *out_j;
    void *out_D;
    void *out_d;
    int32_t *out_S;
};

/* Kernel launchers (mnl_kernels.hip). All asynchronous on `stream`. */
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t stream);
hipError_t launch_fill_i64(int64_t *p, int64_t n, int64_t v, hipStream_t stream);
/* order_stats (may be NULL): per axis, the number of consecutive atoms
   whose cells differ along it, added to (for the next build's layout). */
hipError_t launch_bin_dense(const double *r, int32_t nat, const SysGeom *geo,
                            const int32_t *atom_base, int32_t nsys, int4 *raw4,
                            int32_t *ticket, int32_t *cell_count, int32_t *order_stats,
                            hipStream_t stream);
hipError_t launch_bin_hashed(const double *r, int32_t nat, const SysGeom *geo,
                             int4 *raw4, int32_t *ticket, int64_t *hkey,
                             int64_t hmask, int32_t *hcount, hipStream_t stream);
hipError_t launch_scatter(const int4 *raw4, const int32_t *ticket, int32_t nat,
                          const int32_t *cell_first, int32_t *tmp,
                          hipStream_t stream);
hipError_t launch_rank_gather(const double *r, const int4 *raw4, const int32_t *tmp,
                              int32_t nat, const int32_t *cell_first,
                              const int32_t *cell_count, const SysGeom *geo,
                              const int32_t *atom_base, int32_t nsys,
                              const double *per_atom, const int32_t *types,
                              double4 *spos, int4 *sinfo, int32_t *stype,
                              int32_t *scount, int32_t *slot, int32_t *first_flag,
                              hipStream_t stream);
/* home_list[pos[a]] = raw4[a].w for every atom a with flag[a] (the lowest
   atom of its cell, flagged by the rank/gather; pos = exclusive scan of
   the flags). */
hipError_t launch_home_list(const int4 *raw4, const int32_t *flag, const int32_t *pos,
                            int32_t nat, int32_t *home_list, hipStream_t stream);
/* Waves a pair-kernel launch of these arguments takes. */
int64_t pair_waves(const PairArgs &a);
/* Exclusive scan of n int32 values; out gets n (+1 with the total) entries.
   partial needs scan_partial_size(n) int64 entries. total and total2
   (device, may be NULL) receive the sum. With status (scan_status_words()
   words, all zero before the first scan that uses them; each scan leaves
   them zero), scans of up to 1024 tiles of 4096 values run in one launch
   with a decoupled look-back instead of a reduce and an apply kernel; the
   same status words must not serve two scans at once. */
int64_t scan_partial_size(int64_t n);
int64_t scan_status_words();
hipError_t launch_scan_i32_to_i32(const int32_t *in, int64_t n, int32_t *out,
                                  int64_t *partial, int64_t *total,
                                  hipStream_t stream, uint64_t *status = nullptr);
hipError_t launch_scan_i32_to_i64(const int32_t *in, int64_t n, int64_t *out,
                                  int64_t *partial, int64_t *total,
                                  hipStream_t stream, int64_t *total2 = nullptr,
                                  uint64_t *status = nullptr);
/* The same scan over per-atom counts held in cell-sorted order (the half
   count): value of atom a = scount[slot[a]], also written to counts[a]. */
hipError_t launch_scan_sorted_counts(const int32_t *scount, const int32_t *slot,
                                     int32_t *counts, int64_t n, int64_t *out,
                                     int64_t *partial, int64_t *total, hipStream_t stream,
                                     int64_t *total2 = nullptr, uint64_t *status = nullptr);
hipError_t launch_scan_i64_to_i64(const int64_t *in, int64_t n, int64_t *out,
                                  int64_t *partial, int64_t *total,
                                  hipStream_t stream);

/* Post-processing of i-sorted pair lists (mnl_post.hip). */
hipError_t launch_first_neighbours(const int32_t *i, int64_t nn, int32_t n, int32_t *seed,
                                   hipStream_t stream);
hipError_t launch_count_jumps(const int32_t *sorted, int64_t nn, int32_t *count,
                              hipStream_t stream);
/* rows r < nrows (= n_first - 1): m[r] kept pairs, tcnt[r] = m (m - 1)
   (row ends clamped to nn). With absdist: vidx[vstart[r] + k] = k-th kept
   pair of the row; without, every pair is kept and vstart, vidx are not
   written. */
hipError_t launch_triplet_rows(const int32_t *first_i, int32_t nrows, int64_t nn,
                               const double *absdist, double cutoff, int32_t *m, int64_t *tcnt,
                               int32_t *vstart, int32_t *vidx, hipStream_t stream);
/* vbase: vstart (with vidx) or first_i (vidx NULL: every pair kept) */
hipError_t launch_triplet_fill(const int32_t *vbase, int32_t nrows, const int32_t *m,
                               const int64_t *toff, const int32_t *vidx, int32_t *ij,
                               int32_t *ik, hipStream_t stream);
/* Pair kernels (one wave per home cell). mode 0 = count, 1 = fill f64
   outputs, 2 = fill f32 outputs, 3 = dense matrix (rows of max_nb slots:
   out_j = neighbour index, out_D = D, counts = row counts; no count pass). */
hipError_t launch_pairs(int mode, int cut, bool hashed, const PairArgs &args,
                        hipStream_t stream);
/* One-system parameters into geo/atom_base/cell_base, and the grid cleared:
   cell_count[0..nslots) = 0 and, if hkey is given, hkey[0..nslots) = -1.
   order_stats (may be NULL): the previous binning's per-axis counts, which
   choose the dense layout, then cleared. */
/* Lennard-Jones forces [nat][3], energy and virial [3][3] (energy and
   virial may be NULL) from the binned grid of the plan's last binning,
   without a pair list; global cutoff only. scratch holds
   lj_scratch_doubles(args) doubles; slot maps atoms to cell-sorted
   indices. */
int64_t lj_scratch_doubles(const PairArgs &a);
hipError_t launch_lj(const PairArgs &args, bool hashed, double epsilon, double sigma,
                     const int32_t *slot, double *scratch, double *forces, double *energy,
                     double *virial, hipStream_t stream);

/* The whole binning of one small dense system (at most 4096 atoms and 4096
   cells: small_binning_fits) in a single block: writes the parameters (with
   the layout chosen from order_stats) and everything set_params, bin_dense,
   the cell scan, scatter and rank_gather write, except the scatter's
   scratch (raw4, ticket, tmp). */
bool small_binning_fits(int64_t nat, int64_t ncells);
/* The binning of a batch whose systems all fit one block each (at most
   2048 atoms and 2048 cells: batch_binning_fits), one block per system, from
   the uploaded geometry: writes what bin_dense, the cell scan, scatter and
   rank_gather write, except the scatter's scratch. */
bool batch_binning_fits(int64_t max_atoms, int64_t max_cells);
hipError_t launch_bin_systems(const double *r, int32_t nsys, const SysGeom *geo,
                              int32_t *cell_first, int32_t *cell_count, double4 *spos,
                              int4 *sinfo, int32_t *scount, int32_t *slot, hipStream_t stream);
hipError_t launch_bin_small(const double *r, int32_t nat, const SysGeom &g, int32_t ncells,
                            SysGeom *geo, int32_t *atom_base, int32_t *cell_base,
                            int32_t *order_stats, const double *per_atom, const int32_t *types,
                            int32_t *cell_first, int32_t *cell_count, double4 *spos,
                            int4 *sinfo, int32_t *stype, int32_t *scount, int32_t *slot,
                            hipStream_t stream);
hipError_t launch_set_params(SysGeom *geo, const SysGeom &g, int32_t *atom_base,
                             int32_t *cell_base, int32_t nat, int32_t ncells,
                             int32_t *cell_count, int64_t nslots, int64_t *hkey,
                             int32_t *order_stats, hipStream_t stream);
hipError_t launch_sqrt_probe(const double *x, double *y, int64_t n,
                             hipStream_t stream);
/* Backward of D = r[j] - r[i] + S.cell, d = |D| (see mnl_kernels.hip):
   overwrites grad_pos [nat*3] and grad_cell [9]; cell_partial is scratch of
   backward_scratch_doubles(nat) doubles. D, d, gD, gd are float when f32.
   pbc_mask: bit k set if direction k is periodic (S enters the cell
   gradient only there). */
constexpr int kBackwardMaxBlocks = 2048;
int backward_blocks(int64_t npairs);
inline int64_t backward_scratch_doubles(int32_t nat) {
    return 9 * int64_t(kBackwardMaxBlocks) + 3 * int64_t(nat);
}
hipError_t launch_pair_backward(int64_t npairs, int32_t nat, const int32_t *first,
                                const int32_t *secnd, const int32_t *shift, const void *D,
                                const void *d, const void *gD, const void *gd, bool f32,
                                int pbc_mask, double *grad_pos, double *grad_cell,
                                double *cell_partial, hipStream_t stream);
/* The allocation-free build (mnl_neighbour_list_noalloc): its limits and
   the size of the per-device array (mnl_noalloc.hip) that its workspaces
   are carved from. */
constexpr int32_t kNoallocMaxAtoms = 1 << 17;
constexpr int64_t kNoallocMaxCells = int64_t(1) << 19; /* dense cells or hash slots */
constexpr size_t kNoallocScratchBytes = size_t(24) << 20;
constexpr int kNoallocMaxDevices = 64;

/* Records msg as the calling thread's last error; returns -1 (MNL_ERROR). */
int set_error(const char *msg);

/* Count and fill of one system (mnl_plan_count_async followed by
   mnl_plan_fill_async, uniform or per-atom cutoffs) with every workspace
   inside `scratch` (device memory of kNoallocScratchBytes on `device`, the
   current device): no allocation, no host synchronisation. */
int noalloc_build(int device, void *scratch, hipStream_t stream, const double origin[3],
                  const double cell[9], const double inv_cell[9], const int32_t pbc[3],
                  int32_t nat, const double *positions, double cutoff,
                  const double *per_atom_cutoff, int backend, int64_t capacity,
                  int32_t *first, int32_t *secnd, void *distvec, void *absdist,
                  int32_t *shift, int float32_output, int64_t *npairs_device);

/* The same backward with every sum in a fixed order (mnl_backward_det.hip):
   the pairs grouped by j with a stable radix sort, one thread per atom
   adding its j-side pairs and its row in pair order, the cell gradient over
   fixed tiles. first must be sorted; npairs < 2^31. scratch holds
   backward_det_scratch_bytes(npairs, nat) bytes (about 40 per pair). */
size_t backward_det_scratch_bytes(int64_t npairs, int32_t nat);
hipError_t launch_pair_backward_det(int64_t npairs, int32_t nat, const int32_t *first,
                                    const int32_t *secnd, const int32_t *shift, const void *D,
                                    const void *d, const void *gD, const void *gd, bool f32,
                                    int pbc_mask, double *grad_pos, double *grad_cell,
                                    void *scratch, hipStream_t stream);

/* out[s] = offsets[atom_base[s]] for s = 0..nsys: each system's first pair. */
hipError_t launch_system_offsets(const int64_t *offsets, const int32_t *atom_base,
                                 int32_t nsys, int64_t *out, hipStream_t stream);

}  // namespace mnl

#endif

/*
 * mnl oracle: CPU restatement of the reference neighbour-list path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP engine is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never links or calls it.
 *
 * Every function restates the reference algorithm with the same floating-point
 * operation order and no FMA contraction (built with -ffp-contract=off), so its
 * outputs are bit-identical to the reference CPU path
 * (src/libneighbours/neighbour_list.cc:99-204 with CellOrder::Linear).
 */
#ifndef MNL_ORACLE_H
#define MNL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quantity bit flags; same values as the reference's Quantity enum
   (src/libneighbours/types.hh:37-43). */
enum {
    ORACLE_Q_FIRST = 1,
    ORACLE_Q_SECOND = 2,
    ORACLE_Q_DISTVEC = 4,
    ORACLE_Q_ABSDIST = 8,
    ORACLE_Q_SHIFT = 16
};

typedef struct oracle_result {
    int64_t npairs;
    int32_t *first;   /* [npairs] or NULL */
    int32_t *secnd;   /* [npairs] or NULL */
    double *distvec;  /* [3*npairs] or NULL */
    double *absdist;  /* [npairs] or NULL */
    int32_t *shift;   /* [3*npairs] or NULL */
    int32_t *counts;  /* [nat] per-atom neighbour count (always filled) */
    int sparse;       /* backend that was used: 0 dense, 1 hashed */
} oracle_result;

/*
 * Restates matscipy::neighbour_list (src/libneighbours/neighbour_list.hh:98-104).
 * force_backend: -1 = the reference rule (cell_list.cc:193-196), 0 = dense,
 * 1 = hashed. The backend never changes the output.
 * Returns 0 on success, -1 on error (message via oracle_last_error()).
 */
int oracle_neighbour_list(int quantities, const double origin[3],
                          const double cell[9], const double inv_cell[9],
                          const int32_t pbc[3], int32_t nat, const double *r,
                          double cutoff, const double *per_atom_cutoff,
                          const double *per_type_cutoff_sq, int32_t ncutoffs,
                          const int32_t *types, int force_backend,
                          oracle_result *out);

void oracle_free(oracle_result *res);
const char *oracle_last_error(void);

/* Grid geometry (cell_list.cc:153-182 and neighbour_list.cc:176-178):
   out_n[3] cells per direction, out_nxyz[3] search range, out_bins[9] cell
   edge vectors (rows), out_len[3] box widths. Returns -1 for zero volume. */
int oracle_grid_geometry(const double cell[9], double cutoff, int32_t out_n[3],
                         int32_t out_nxyz[3], double out_bins[9],
                         double out_len[3]);

/* Restates first_neighbours (src/libneighbours/first_neighbours.cc:15-40). */
void oracle_first_neighbours(int32_t n, int32_t nn, const int32_t *i_n,
                             int32_t *seed);

#ifdef __cplusplus
}
#endif

#endif

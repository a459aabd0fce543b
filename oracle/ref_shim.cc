/*
 * C-ABI harness around the compiled reference CPU core.
 *
 * TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
 * the reference's own CPU sources, read where they lie under the reference
 * checkout (error.cc, tools.cc, cell_list.cc, neighbour_list.cc,
 * first_neighbours.cc, triplet_list.cc), into oracle/_ref/libref_neighbours.so.
 * It is used to write the golden fixtures, to cross-check the oracle
 * restatement, and as bench.py's CPU baseline. The product never loads it.
 *
 * Entry points mirror matscipy::neighbour_list
 * (src/libneighbours/neighbour_list.hh:98-104) with plain C types, and
 * matscipy::neighbour_matrix / first_neighbours / triplet_list.
 */
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "error.hh"
#include "first_neighbours.hh"
#include "neighbour_list.hh"
#include "triplet_list.hh"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace matscipy;

template <typename T>
static T *dup(const std::vector<T> &v) {
    if (v.empty()) return nullptr;
    T *p = static_cast<T *>(std::malloc(v.size() * sizeof(T)));
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

extern "C" {

struct ref_result {
    int64_t npairs;
    int32_t *first, *secnd;
    double *distvec, *absdist;
    int32_t *shift;
};

int ref_neighbour_list(int quantities, const double *origin, const double *cell,
                       const double *inv_cell, const int32_t *pbc_in, int32_t nat,
                       const double *r, double cutoff, const double *per_atom,
                       const double *per_type_sq, int32_t ncutoffs,
                       const int32_t *types, ref_result *out) {
    const bool pbc[3] = {pbc_in[0] != 0, pbc_in[1] != 0, pbc_in[2] != 0};
    NeighbourList nl;
    std::memset(out, 0, sizeof *out);
    error_t e = neighbour_list(quantities, origin, cell, inv_cell, pbc, nat, r,
                               cutoff, per_atom, per_type_sq, ncutoffs, types, nl);
    if (e != NL_SUCCESS) return -1;
    out->npairs = nl.npairs;
    out->first = dup(nl.first);
    out->secnd = dup(nl.secnd);
    out->distvec = dup(nl.distvec);
    out->absdist = dup(nl.absdist);
    out->shift = dup(nl.shift);
    return 0;
}

/* Time n_rep calls (no copies out) and return the wall seconds of each; the
   pair count of the last call goes to *npairs. */
int ref_time_neighbour_list(int quantities, const double *origin, const double *cell,
                            const double *inv_cell, const int32_t *pbc_in,
                            int32_t nat, const double *r, double cutoff,
                            int32_t n_rep, double *seconds, int64_t *npairs) {
    const bool pbc[3] = {pbc_in[0] != 0, pbc_in[1] != 0, pbc_in[2] != 0};
    NeighbourList nl;
    for (int k = 0; k < n_rep; k++) {
#ifdef _OPENMP
        double t0 = omp_get_wtime();
#endif
        if (neighbour_list(quantities, origin, cell, inv_cell, pbc, nat, r, cutoff,
                           nullptr, nullptr, 0, nullptr, nl) != NL_SUCCESS)
            return -1;
#ifdef _OPENMP
        seconds[k] = omp_get_wtime() - t0;
#else
        seconds[k] = 0.0;
#endif
    }
    *npairs = nl.npairs;
    return 0;
}

int ref_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ref_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void ref_free(ref_result *res) {
    std::free(res->first);
    std::free(res->secnd);
    std::free(res->distvec);
    std::free(res->absdist);
    std::free(res->shift);
    std::memset(res, 0, sizeof *res);
}

const char *ref_last_error(void) { return error_string; }

void ref_first_neighbours(int32_t n, int32_t nn, const int32_t *i_n, int32_t *seed) {
    first_neighbours(n, nn, i_n, seed);
}

/* Writes the jump indices into `out` (capacity cap); returns their count. */
int32_t ref_get_jump_indicies(int32_t nn, const int32_t *sorted, int32_t *out,
                              int32_t cap) {
    std::vector<index_t> v = get_jump_indicies(nn, sorted);
    int32_t n = static_cast<int32_t>(v.size());
    for (int32_t k = 0; k < n && k < cap; k++) out[k] = v[k];
    return n;
}

/* triplet_list; returns the triplet count, writes up to cap of each. */
int64_t ref_triplet_list(int32_t n_first, const int32_t *first_i,
                         const double *absdist, double cutoff, int32_t *ij,
                         int32_t *ik, int64_t cap) {
    std::vector<index_t> ij_t, ik_t;
    triplet_list(n_first, first_i, absdist, cutoff, ij_t, ik_t);
    int64_t n = static_cast<int64_t>(ij_t.size());
    for (int64_t k = 0; k < n && k < cap; k++) {
        ij[k] = ij_t[k];
        ik[k] = ik_t[k];
    }
    return n;
}

int ref_neighbour_matrix(const double *origin, const double *cell,
                         const double *inv_cell, const int32_t *pbc_in,
                         int32_t nat, const double *r, double cutoff,
                         const double *per_atom, const double *per_type_sq,
                         int32_t ncutoffs, const int32_t *types, int32_t K,
                         int32_t *idx, double *dist, int32_t *count,
                         int32_t *overflow) {
    const bool pbc[3] = {pbc_in[0] != 0, pbc_in[1] != 0, pbc_in[2] != 0};
    NeighbourMatrix nm;
    if (neighbour_matrix(origin, cell, inv_cell, pbc, nat, r, cutoff, per_atom,
                         per_type_sq, ncutoffs, types, K, nm) != NL_SUCCESS)
        return -1;
    std::memcpy(idx, nm.idx.data(), nm.idx.size() * sizeof(int32_t));
    std::memcpy(dist, nm.dist.data(), nm.dist.size() * sizeof(double));
    std::memcpy(count, nm.count.data(), nm.count.size() * sizeof(int32_t));
    *overflow = nm.overflow ? 1 : 0;
    return 0;
}

}  // extern "C"

/*
 * mnl oracle: CPU restatement of the reference neighbour-list path.
 *
 * TEST INFRASTRUCTURE ONLY — the checker, never the thing measured or shipped.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * It follows, step by step and with the same floating-point operation order,
 *   geometry        src/libneighbours/cell_list.cc:153-182, tools.cc:17-25
 *   cell index      src/libneighbours/tools.hh:30-36,55-67
 *   binning         src/libneighbours/cell_list.cc:29-149,184-203
 *   gather          src/libneighbours/neighbour_list.cc:139-167
 *   context         src/libneighbours/neighbour_list.cc:169-192
 *   traversal       src/libneighbours/neighbour_visit.hh:51-120
 *   two passes      src/libneighbours/neighbour_list.cc:40-95
 * Built with -ffp-contract=off: no a*b+c is fused, as in the reference build
 * (x86-64 without -mfma has no FMA to contract into).
 *
 * The OpenMP structure is the reference's (static per-atom loops, serial
 * counting sort and prefix sums, dynamic(256) count and fill passes), so that
 * timing it is a fair stand-in for the reference CPU path. Pair offsets are
 * int64 here (the reference uses int); results agree wherever the reference
 * does not overflow.
 */
#include "mnl_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

static int set_err(const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return -1;
}

const char *oracle_last_error(void) { return g_err; }

/* tools.cc:17-21 */
static void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* tools.cc:23-25 */
static double norm3(const double *a) {
    return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

/* tools.hh:39-43 */
static int32_t wrap_idx(int32_t i, int32_t n) {
    while (i < 0) i += n;
    while (i >= n) i -= n;
    return i;
}

/* tools.hh:46-52 */
static int32_t clamp_idx(int32_t i, int32_t n) {
    if (i < 0) return 0;
    if (i >= n) return n - 1;
    return i;
}

typedef struct {
    int32_t n[3];
    int32_t nxyz[3];
    double bin[3][3]; /* bin[k] = cell row k / n[k] */
    double len[3];
} geom_t;

/* cell_list.cc:153-182 plus the search range of neighbour_list.cc:176-178. */
static int geometry(const double cell[9], double cutoff, geom_t *g) {
    const double *c1 = &cell[0], *c2 = &cell[3], *c3 = &cell[6];
    double nrm[3][3];
    cross3(c2, c3, nrm[0]);
    cross3(c3, c1, nrm[1]);
    cross3(c1, c2, nrm[2]);
    double volume =
        fabs(c3[0] * nrm[2][0] + c3[1] * nrm[2][1] + c3[2] * nrm[2][2]);
    if (volume < 1e-12) return -1;
    for (int k = 0; k < 3; k++) {
        g->len[k] = volume / norm3(nrm[k]);
        int32_t nk = (int32_t)floor(g->len[k] / cutoff);
        g->n[k] = nk > 1 ? nk : 1;
    }
    for (int m = 0; m < 3; m++) {
        g->bin[0][m] = c1[m] / g->n[0];
        g->bin[1][m] = c2[m] / g->n[1];
        g->bin[2][m] = c3[m] / g->n[2];
    }
    for (int k = 0; k < 3; k++)
        g->nxyz[k] = (int32_t)ceil(cutoff * g->n[k] / g->len[k]);
    return 0;
}

int oracle_grid_geometry(const double cell[9], double cutoff, int32_t out_n[3],
                         int32_t out_nxyz[3], double out_bins[9],
                         double out_len[3]) {
    geom_t g;
    if (geometry(cell, cutoff, &g) != 0) return set_err("Zero cell volume.");
    for (int k = 0; k < 3; k++) {
        out_n[k] = g.n[k];
        out_nxyz[k] = g.nxyz[k];
        out_len[k] = g.len[k];
        for (int m = 0; m < 3; m++) out_bins[3 * k + m] = g.bin[k][m];
    }
    return 0;
}

/* Cell slice lookup: dense (cell_list.hh:94-104) or hashed (:107-131). */
typedef struct {
    int sparse;
    int32_t n1, n2;
    const int32_t *cell_first, *cell_count;
    int64_t mask;
    const int64_t *hkey;
    const int32_t *hfirst, *hcount;
} query_t;

/* cell_list.hh:72-75 (Fibonacci hashing). */
static int64_t hash64(int64_t key) {
    uint64_t h = (uint64_t)key * 0x9E3779B97F4A7C15ull;
    return (int64_t)(h >> 1);
}

static void slice(const query_t *q, int32_t c1, int32_t c2, int32_t c3,
                  int32_t *b, int32_t *e) {
    if (!q->sparse) {
        int32_t c = c1 + q->n1 * (c2 + q->n2 * c3);
        *b = q->cell_first[c];
        *e = *b + q->cell_count[c];
        return;
    }
    int64_t key = (int64_t)c1 + (int64_t)q->n1 * ((int64_t)c2 + (int64_t)q->n2 * c3);
    int64_t h = hash64(key) & q->mask;
    while (q->hkey[h] != -1) {
        if (q->hkey[h] == key) {
            *b = q->hfirst[h];
            *e = *b + q->hcount[h];
            return;
        }
        h = (h + 1) & q->mask;
    }
    *b = 0;
    *e = 0;
}

typedef struct {
    int32_t n[3], nxyz[3];
    int pbc[3];
    double bin[3][3];
    double cutoff_sq;
    const double *per_type_sq;
    int32_t ncutoffs;
    const int32_t *sorted_atom, *raw_cell, *rel_cell;
    const double *rel_pos, *per_atom;
    const int32_t *types;
} ctx_t;

/* Sink for one accepted pair. */
typedef struct {
    int mode; /* 0 count, 1 fill */
    int64_t w;
    int32_t i;
    int quantities;
    oracle_result *out;
} sink_t;

/* neighbour_visit.hh:51-120, with the emit bodies of neighbour_list.cc:49-50
   (count) and :77-93 (fill). Returns the number of emitted pairs. */
static int64_t visit(const ctx_t *c, const query_t *q, int32_t si, sink_t *s) {
    const int32_t n1 = c->n[0], n2 = c->n[1], n3 = c->n[2];
    const int32_t *raw_i = &c->raw_cell[3 * (int64_t)si];
    const double *dri = &c->rel_pos[3 * (int64_t)si];
    const int32_t ci1 = c->pbc[0] ? wrap_idx(raw_i[0], n1) : clamp_idx(raw_i[0], n1);
    const int32_t ci2 = c->pbc[1] ? wrap_idx(raw_i[1], n2) : clamp_idx(raw_i[1], n2);
    const int32_t ci3 = c->pbc[2] ? wrap_idx(raw_i[2], n3) : clamp_idx(raw_i[2], n3);
    int64_t emitted = 0;

    for (int32_t z = -c->nxyz[2]; z <= c->nxyz[2]; z++) {
        int32_t cj3 = ci3 + z;
        if (c->pbc[2]) cj3 = wrap_idx(cj3, n3);
        if (cj3 < 0 || cj3 >= n3) continue;
        double off3[3] = {z * c->bin[2][0], z * c->bin[2][1], z * c->bin[2][2]};
        for (int32_t y = -c->nxyz[1]; y <= c->nxyz[1]; y++) {
            int32_t cj2 = ci2 + y;
            if (c->pbc[1]) cj2 = wrap_idx(cj2, n2);
            if (cj2 < 0 || cj2 >= n2) continue;
            double off2[3] = {off3[0] + y * c->bin[1][0], off3[1] + y * c->bin[1][1],
                              off3[2] + y * c->bin[1][2]};
            for (int32_t x = -c->nxyz[0]; x <= c->nxyz[0]; x++) {
                int32_t cj1 = ci1 + x;
                if (c->pbc[0]) cj1 = wrap_idx(cj1, n1);
                if (cj1 < 0 || cj1 >= n1) continue;
                double off[3] = {off2[0] + x * c->bin[0][0],
                                 off2[1] + x * c->bin[0][1],
                                 off2[2] + x * c->bin[0][2]};
                int32_t b, e;
                slice(q, cj1, cj2, cj3, &b, &e);
                for (int32_t sj = b; sj < e; sj++) {
                    if (sj == si && x == 0 && y == 0 && z == 0) continue;
                    const double *drj = &c->rel_pos[3 * (int64_t)sj];
                    double dr[3] = {drj[0] - dri[0] + off[0], drj[1] - dri[1] + off[1],
                                    drj[2] - dri[2] + off[2]};
                    double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                    if (r2 >= c->cutoff_sq) continue;
                    int inside = 1;
                    if (c->per_atom) {
                        double cc = c->per_atom[si] + c->per_atom[sj];
                        inside = r2 < cc * cc;
                    } else if (c->per_type_sq && c->types) {
                        int32_t ti = c->types[si], tj = c->types[sj];
                        if (ti >= 0 && ti < c->ncutoffs && tj >= 0 && tj < c->ncutoffs)
                            inside = r2 < c->per_type_sq[ti * c->ncutoffs + tj];
                    }
                    if (!inside) continue;
                    emitted++;
                    if (s->mode == 0) continue;
                    const int32_t *crj = &c->rel_cell[3 * (int64_t)sj];
                    int32_t sh[3] = {(raw_i[0] - crj[0] + x) / n1,
                                     (raw_i[1] - crj[1] + y) / n2,
                                     (raw_i[2] - crj[2] + z) / n3};
                    oracle_result *o = s->out;
                    int64_t w = s->w;
                    if (o->first) o->first[w] = s->i;
                    if (o->secnd) o->secnd[w] = c->sorted_atom[sj];
                    if (o->distvec) {
                        o->distvec[3 * w + 0] = dr[0];
                        o->distvec[3 * w + 1] = dr[1];
                        o->distvec[3 * w + 2] = dr[2];
                    }
                    if (o->absdist) o->absdist[w] = sqrt(r2);
                    if (o->shift) {
                        o->shift[3 * w + 0] = sh[0];
                        o->shift[3 * w + 1] = sh[1];
                        o->shift[3 * w + 2] = sh[2];
                    }
                    s->w = w + 1;
                }
            }
        }
    }
    return emitted;
}

void oracle_free(oracle_result *res) {
    if (!res) return;
    free(res->first);
    free(res->secnd);
    free(res->distvec);
    free(res->absdist);
    free(res->shift);
    free(res->counts);
    memset(res, 0, sizeof *res);
}

int oracle_neighbour_list(int quantities, const double origin[3],
                          const double cell[9], const double inv_cell[9],
                          const int32_t pbc_in[3], int32_t nat, const double *r,
                          double cutoff, const double *per_atom_cutoff,
                          const double *per_type_cutoff_sq, int32_t ncutoffs,
                          const int32_t *types, int force_backend,
                          oracle_result *out) {
    memset(out, 0, sizeof *out);
    g_err[0] = 0;
    geom_t g;
    if (geometry(cell, cutoff, &g) != 0) return set_err("Zero cell volume.");
    const int32_t n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
    const int pbc[3] = {pbc_in[0] != 0, pbc_in[1] != 0, pbc_in[2] != 0};
    out->counts = (int32_t *)calloc(nat > 0 ? nat : 1, sizeof(int32_t));
    if (nat <= 0) return 0;

    /* neighbour_list.cc:126-133 with tools.hh:55-67 */
    int32_t *raw = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)nat);
#pragma omp parallel for schedule(static)
    for (int32_t a = 0; a < nat; a++) {
        double dri[3], si[3];
        for (int k = 0; k < 3; k++) dri[k] = r[3 * (int64_t)a + k] - origin[k];
        for (int k = 0; k < 3; k++) {
            si[k] = 0.0;
            for (int m = 0; m < 3; m++) si[k] += inv_cell[3 * k + m] * dri[m];
        }
        raw[3 * (int64_t)a + 0] = (int32_t)floor(si[0] * n1);
        raw[3 * (int64_t)a + 1] = (int32_t)floor(si[1] * n2);
        raw[3 * (int64_t)a + 2] = (int32_t)floor(si[2] * n3);
    }

    /* cell_list.cc:29-41 folded coordinates */
    int32_t *cc = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)nat);
#pragma omp parallel for schedule(static)
    for (int32_t a = 0; a < nat; a++) {
        for (int k = 0; k < 3; k++) {
            int32_t v = raw[3 * (int64_t)a + k];
            cc[3 * (int64_t)a + k] = pbc[k] ? wrap_idx(v, g.n[k]) : clamp_idx(v, g.n[k]);
        }
    }

    /* cell_list.cc:193-196 backend rule */
    const int64_t ncells = (int64_t)n1 * n2 * n3;
    const int64_t budget = (int64_t)8 * nat > (1 << 20) ? (int64_t)8 * nat : (1 << 20);
    int sparse = force_backend >= 0 ? force_backend : (ncells > budget);
    out->sparse = sparse;

    int32_t *sorted_atom = (int32_t *)malloc(sizeof(int32_t) * (size_t)nat);
    int32_t *cell_first = NULL, *cell_count = NULL, *hfirst = NULL, *hcount = NULL;
    int64_t *hkey = NULL, hmask = 0;
    if (!sparse) {
        /* cell_list.cc:43-80 (Linear): histogram, serial prefix, serial scatter
           in atom order -> ascending original index within each cell. */
        cell_first = (int32_t *)calloc((size_t)ncells, sizeof(int32_t));
        cell_count = (int32_t *)calloc((size_t)ncells, sizeof(int32_t));
        int32_t *lin = (int32_t *)malloc(sizeof(int32_t) * (size_t)nat);
        for (int32_t a = 0; a < nat; a++) {
            lin[a] = cc[3 * (int64_t)a] + n1 * (cc[3 * (int64_t)a + 1] + n2 * cc[3 * (int64_t)a + 2]);
            cell_count[lin[a]]++;
        }
        int32_t acc = 0;
        for (int64_t c = 0; c < ncells; c++) {
            cell_first[c] = acc;
            acc += cell_count[c];
        }
        int32_t *cursor = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncells);
        memcpy(cursor, cell_first, sizeof(int32_t) * (size_t)ncells);
        for (int32_t a = 0; a < nat; a++) sorted_atom[cursor[lin[a]]++] = a;
        free(cursor);
        free(lin);
    } else {
        /* cell_list.cc:105-149: open addressing, capacity pow2 >= 2N+1,
           serial insert, prefix over slots, serial scatter. */
        int64_t cap = 1;
        while (cap < 2 * (int64_t)nat + 1) cap <<= 1;
        hmask = cap - 1;
        hkey = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
        for (int64_t h = 0; h < cap; h++) hkey[h] = -1;
        hfirst = (int32_t *)calloc((size_t)cap, sizeof(int32_t));
        hcount = (int32_t *)calloc((size_t)cap, sizeof(int32_t));
        int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (size_t)nat);
        for (int32_t a = 0; a < nat; a++) {
            int64_t key = (int64_t)cc[3 * (int64_t)a] +
                          (int64_t)n1 * ((int64_t)cc[3 * (int64_t)a + 1] +
                                         (int64_t)n2 * cc[3 * (int64_t)a + 2]);
            keys[a] = key;
            int64_t h = hash64(key) & hmask;
            while (hkey[h] != -1 && hkey[h] != key) h = (h + 1) & hmask;
            hkey[h] = key;
            hcount[h]++;
        }
        int32_t acc = 0;
        for (int64_t h = 0; h <= hmask; h++) {
            if (hkey[h] != -1) {
                hfirst[h] = acc;
                acc += hcount[h];
            }
        }
        int32_t *cursor = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        memcpy(cursor, hfirst, sizeof(int32_t) * (size_t)cap);
        for (int32_t a = 0; a < nat; a++) {
            int64_t h = hash64(keys[a]) & hmask;
            while (hkey[h] != keys[a]) h = (h + 1) & hmask;
            sorted_atom[cursor[h]++] = a;
        }
        free(cursor);
        free(keys);
    }
    free(cc);

    /* neighbour_list.cc:139-167 gather into cell-sorted order */
    int32_t *raw_s = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)nat);
    int32_t *rel_s = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)nat);
    double *pos_s = (double *)malloc(sizeof(double) * 3 * (size_t)nat);
    double *pa_s = per_atom_cutoff ? (double *)malloc(sizeof(double) * (size_t)nat) : NULL;
    int32_t *ty_s = types ? (int32_t *)malloc(sizeof(int32_t) * (size_t)nat) : NULL;
#pragma omp parallel for schedule(static)
    for (int32_t s = 0; s < nat; s++) {
        const int32_t a = sorted_atom[s];
        const int32_t c1 = raw[3 * (int64_t)a], c2 = raw[3 * (int64_t)a + 1],
                      c3 = raw[3 * (int64_t)a + 2];
        raw_s[3 * (int64_t)s + 0] = c1;
        raw_s[3 * (int64_t)s + 1] = c2;
        raw_s[3 * (int64_t)s + 2] = c3;
        const int32_t r1 = pbc[0] ? c1 : clamp_idx(c1, n1);
        const int32_t r2 = pbc[1] ? c2 : clamp_idx(c2, n2);
        const int32_t r3 = pbc[2] ? c3 : clamp_idx(c3, n3);
        rel_s[3 * (int64_t)s + 0] = r1;
        rel_s[3 * (int64_t)s + 1] = r2;
        rel_s[3 * (int64_t)s + 2] = r3;
        for (int k = 0; k < 3; k++)
            pos_s[3 * (int64_t)s + k] = r[3 * (int64_t)a + k] - r1 * g.bin[0][k] -
                                        r2 * g.bin[1][k] - r3 * g.bin[2][k];
        if (pa_s) pa_s[s] = per_atom_cutoff[a];
        if (ty_s) ty_s[s] = types[a];
    }
    free(raw);

    /* neighbour_list.cc:169-192 */
    ctx_t ctx;
    for (int k = 0; k < 3; k++) {
        ctx.n[k] = g.n[k];
        ctx.nxyz[k] = g.nxyz[k];
        ctx.pbc[k] = pbc[k];
        for (int m = 0; m < 3; m++) ctx.bin[k][m] = g.bin[k][m];
    }
    ctx.cutoff_sq = cutoff * cutoff;
    ctx.per_type_sq = per_type_cutoff_sq;
    ctx.ncutoffs = ncutoffs;
    ctx.sorted_atom = sorted_atom;
    ctx.raw_cell = raw_s;
    ctx.rel_cell = rel_s;
    ctx.rel_pos = pos_s;
    ctx.per_atom = pa_s;
    ctx.types = ty_s;

    query_t q;
    q.sparse = sparse;
    q.n1 = n1;
    q.n2 = n2;
    q.cell_first = cell_first;
    q.cell_count = cell_count;
    q.mask = hmask;
    q.hkey = hkey;
    q.hfirst = hfirst;
    q.hcount = hcount;

    /* neighbour_list.cc:40-56 count pass + prefix over original atom index */
    int64_t *offset = (int64_t *)calloc((size_t)nat + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t si = 0; si < nat; si++) {
        sink_t sk = {0, 0, 0, quantities, out};
        int64_t n = visit(&ctx, &q, si, &sk);
        offset[sorted_atom[si] + 1] = n;
        out->counts[sorted_atom[si]] = (int32_t)n;
    }
    for (int32_t i = 0; i < nat; i++) offset[i + 1] += offset[i];
    const int64_t np = offset[nat];
    out->npairs = np;

    /* neighbour_list.cc:58-94 exact-size outputs, fill pass */
    size_t n1b = (size_t)(np > 0 ? np : 1);
    if (quantities & ORACLE_Q_FIRST) out->first = (int32_t *)malloc(sizeof(int32_t) * n1b);
    if (quantities & ORACLE_Q_SECOND) out->secnd = (int32_t *)malloc(sizeof(int32_t) * n1b);
    if (quantities & ORACLE_Q_DISTVEC) out->distvec = (double *)malloc(sizeof(double) * 3 * n1b);
    if (quantities & ORACLE_Q_ABSDIST) out->absdist = (double *)malloc(sizeof(double) * n1b);
    if (quantities & ORACLE_Q_SHIFT) out->shift = (int32_t *)malloc(sizeof(int32_t) * 3 * n1b);
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t si = 0; si < nat; si++) {
        const int32_t i = sorted_atom[si];
        sink_t sk = {1, offset[i], i, quantities, out};
        visit(&ctx, &q, si, &sk);
    }

    free(offset);
    free(sorted_atom);
    free(cell_first);
    free(cell_count);
    free(hkey);
    free(hfirst);
    free(hcount);
    free(raw_s);
    free(rel_s);
    free(pos_s);
    free(pa_s);
    free(ty_s);
    return 0;
}

/* first_neighbours.cc:15-40 */
void oracle_first_neighbours(int32_t n, int32_t nn, const int32_t *i_n,
                             int32_t *seed) {
    for (int32_t k = 0; k <= n; k++) seed[k] = -1;
    if (nn <= 0) {
        for (int32_t k = 0; k <= n; k++) seed[k] = 0;
        return;
    }
    seed[i_n[0]] = 0;
    for (int32_t k = 1; k < nn; k++) {
        if (i_n[k] != i_n[k - 1]) {
            for (int32_t l = i_n[k - 1] + 1; l <= i_n[k]; l++) seed[l] = k;
        }
    }
    for (int32_t k = i_n[nn - 1] + 1; k <= n; k++) seed[k] = nn;
}

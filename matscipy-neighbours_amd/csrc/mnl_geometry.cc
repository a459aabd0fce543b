/*
 * Host geometry of the cell grid. Built with -ffp-contract=off: the cell
 * resolution, cell edge vectors and search range must come out bit-identical
 * to the reference's host arithmetic, since they feed every position the
 * kernels compute.
 */
#include <cmath>

#include "mnl_engine.hh"

#pragma clang fp contract(off)

namespace mnl {

namespace {

/* c = a x b (src/libneighbours/tools.cc:17-21). */
void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* |a| (src/libneighbours/tools.cc:23-25). */
double norm3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* The reference converts floor(width / cutoff) and ceil(cutoff n / width)
   with static_cast<index_t>; on x86-64 an out-of-range or NaN value gives
   INT_MIN (cvttsd2si), which its max(., 1) turns into one cell (cutoff 0, a
   width far above the cutoff). The same rule is spelled out here so that no
   conversion is undefined. */
int32_t to_index(double v) {
    return (v >= -2147483648.0 && v < 2147483648.0) ? static_cast<int32_t>(v) : INT32_MIN;
}

}  // namespace

/* src/libneighbours/cell_list.cc:153-182: volume from the triple product,
   widths = volume / |face normal|, n = max(floor(width / cutoff), 1),
   edge vectors = lattice vector / n; plus neighbour_list.cc:176-178:
   nx = ceil(cutoff * n / width). */
bool grid_geometry(const double cell[9], double cutoff, int32_t n[3],
                   int32_t nxyz[3], double bin[9], double len[3]) {
    const double *c1 = &cell[0], *c2 = &cell[3], *c3 = &cell[6];
    double nrm[3][3];
    cross3(c2, c3, nrm[0]);
    cross3(c3, c1, nrm[1]);
    cross3(c1, c2, nrm[2]);
    const double volume =
        std::fabs(c3[0] * nrm[2][0] + c3[1] * nrm[2][1] + c3[2] * nrm[2][2]);
    if (!(volume >= 1e-12)) return false;
    for (int k = 0; k < 3; k++) {
        len[k] = volume / norm3(nrm[k]);
        const int32_t nk = to_index(std::floor(len[k] / cutoff));
        n[k] = nk > 1 ? nk : 1;
    }
    for (int m = 0; m < 3; m++) {
        bin[0 + m] = c1[m] / n[0];
        bin[3 + m] = c2[m] / n[1];
        bin[6 + m] = c3[m] / n[2];
    }
    for (int k = 0; k < 3; k++) {
        /* a reach beyond int32 is refused by the caller (stencil_size) */
        const double reach = std::ceil(cutoff * n[k] / len[k]);
        nxyz[k] = reach >= 2147483647.0 ? INT32_MAX : to_index(reach);
    }
    return true;
}

bool make_sys_geom(const double origin[3], const double cell[9],
                   const double inv_cell[9], const int32_t pbc[3], double cutoff,
                   int32_t atom_base, int32_t natoms, int32_t cell_base,
                   SysGeom &g) {
    double len[3];
    if (!grid_geometry(cell, cutoff, g.n, g.nxyz, g.bin, len)) return false;
    for (int k = 0; k < 9; k++) g.inv[k] = inv_cell[k];
    for (int k = 0; k < 3; k++) {
        g.origin[k] = origin[k];
        g.pbc[k] = pbc[k] ? 1 : 0;
    }
    g.cutoff_sq = cutoff * cutoff;
    for (int k = 0; k < 3; k++) g.rn[k] = 1.0 / g.n[k];
    g.atom_base = atom_base;
    g.natoms = natoms;
    g.cell_base = cell_base;
    const int64_t nc = static_cast<int64_t>(g.n[0]) * g.n[1] * g.n[2];
    g.ncells = nc > INT32_MAX ? INT32_MAX : static_cast<int32_t>(nc);
    const int64_t nseg = stencil_size(g.nxyz);
    g.nseg = nseg > INT32_MAX ? INT32_MAX : static_cast<int32_t>(nseg);
    g.split = 1;
    for (int k = 0; k < 3; k++) g.lay[k] = k;
    int32_t divisors[5] = {g.n[0], g.n[1], g.n[2], 1, 1};
    if (nseg <= kMaxStencil) {
        divisors[3] = 2 * g.nxyz[0] + 1;
        divisors[4] = 2 * g.nxyz[1] + 1;
    }
    for (int k = 0; k < 5; k++) make_div_magic(divisors[k], g.dmul[k], g.dshift[k]);
    return true;
}

int64_t stencil_size(const int32_t nxyz[3]) {
    int64_t s = 1;
    for (int k = 0; k < 3; k++) {
        if (nxyz[k] < 0) return -1;
        s *= 2 * static_cast<int64_t>(nxyz[k]) + 1;
        if (s > kMaxStencil) return kMaxStencil + 1;
    }
    return s;
}

/* s = ceil(log2 d), mul = floor(2^(31+s) / d) + 1 < 2^32. For x < 2^31,
   x * mul / 2^(31+s) = x/d + x*e/2^(31+s) with 0 < e <= 1, and the excess
   is below 2^-s <= 1/d, which the fractional part of x/d (at most
   1 - 1/d) cannot carry past the next integer. d = 1: mul 1, shift 0. */
void make_div_magic(int32_t d, uint32_t &mul, int32_t &shift) {
    if (d <= 1) {
        mul = 1;
        shift = 0;
        return;
    }
    int s = 0;
    while ((int64_t(1) << s) < d) s++;
    mul = static_cast<uint32_t>((uint64_t(1) << (31 + s)) / uint64_t(d) + 1);
    shift = 31 + s;
}

}  // namespace mnl

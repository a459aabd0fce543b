/*
 * The C++ interface (include/mnl.hpp) on the scenarios of the reference's
 * GPU test (tests/test_neighbour_list_gpu.cc:95-233): the same sizes, boxes,
 * cutoffs and seeds, and the same entry points and types, called through
 * `namespace matscipy = mnl`. The reference compares its GPU build with its
 * CPU build as sorted multisets to 1e-12; here every array must equal the
 * oracle (the reference CPU path restated) bit for bit and in order. Test
 * infrastructure: links the oracle as the checker.
 *
 *   cpp_api_test --no-gpu   argument errors only (no device work)
 *   cpp_api_test            every scenario on the current device
 */
#include "mnl.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mnl_oracle.h"

namespace matscipy = mnl;
using matscipy::index_t;
using matscipy::real_t;

namespace {

int g_checks = 0, g_failed = 0;
std::string g_test;

#define EXPECT(cond, what)                                                              \
    do {                                                                                \
        if (cond) {                                                                     \
            ++g_checks;                                                                 \
        } else {                                                                        \
            ++g_failed;                                                                 \
            std::fprintf(stderr, "FAILED [%s] %s (%s:%d) [%s]\n", g_test.c_str(), what, \
                         __FILE__, __LINE__, matscipy::get_error_msg());                \
            return;                                                                     \
        }                                                                               \
    } while (0)

constexpr int kAll = matscipy::QUANTITY_FIRST | matscipy::QUANTITY_SECOND |
                     matscipy::QUANTITY_DISTVEC | matscipy::QUANTITY_ABSDIST |
                     matscipy::QUANTITY_SHIFT;

struct Box {
    real_t cell[9], inv[9], origin[3] = {0, 0, 0};
    bool pbc[3] = {true, true, true};
    int32_t ipbc[3] = {1, 1, 1};
    std::vector<real_t> r;
    int N;
};

/* The reference test's inputs: cubic box of edge L, pbc TTT, coordinates
   uniform in [0, L) (or [0, fill) for a cluster in vacuum) from mt19937. */
Box make_box(int N, double L, unsigned seed, double fill = -1.0) {
    Box b;
    const real_t c[9] = {L, 0, 0, 0, L, 0, 0, 0, L};
    const real_t v[9] = {1 / L, 0, 0, 0, 1 / L, 0, 0, 0, 1 / L};
    std::memcpy(b.cell, c, sizeof(c));
    std::memcpy(b.inv, v, sizeof(v));
    b.N = N;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<real_t> U(0.0, fill < 0 ? L : fill);
    b.r.resize(3 * static_cast<size_t>(N));
    for (auto &x : b.r) x = U(rng);
    return b;
}

bool oracle_list(const Box &b, double cutoff, oracle_result &out) {
    std::memset(&out, 0, sizeof(out));
    return oracle_neighbour_list(kAll, b.origin, b.cell, b.inv, b.ipbc, b.N, b.r.data(), cutoff,
                                 nullptr, nullptr, 0, nullptr, -1, &out) == 0;
}

template <typename T>
bool same_bits(const std::vector<T> &a, const T *b, size_t n) {
    return a.size() == n && (n == 0 || std::memcmp(a.data(), b, n * sizeof(T)) == 0);
}

template <typename T>
std::vector<T> to_host(const matscipy::DeviceArray<T> &a) {
    std::vector<T> v(a.size());
    if (!v.empty())
        (void)hipMemcpy(v.data(), a.data(), v.size() * sizeof(T), hipMemcpyDeviceToHost);
    return v;
}

/* neighbour_list_gpu (host in, host out) and neighbour_list_gpu_device
   (device positions, device result) against the oracle. */
void compare(int N, double L, double cutoff, unsigned seed,
             matscipy::CellOrder order = matscipy::CellOrder::Linear, double fill = -1.0) {
    const Box b = make_box(N, L, seed, fill);
    oracle_result ref;
    EXPECT(oracle_list(b, cutoff, ref), "oracle build");
    const size_t P = static_cast<size_t>(ref.npairs);

    matscipy::NeighbourList gpu;
    const bool built = matscipy::neighbour_list_gpu(kAll, b.origin, b.cell, b.inv, b.pbc, N,
                                                    b.r.data(), cutoff, nullptr, nullptr, 0,
                                                    nullptr, gpu, order) == matscipy::NL_SUCCESS;
    const bool host_ok = built && static_cast<size_t>(gpu.npairs) == P &&
                         same_bits(gpu.first, ref.first, P) &&
                         same_bits(gpu.secnd, ref.secnd, P) &&
                         same_bits(gpu.shift, ref.shift, 3 * P) &&
                         same_bits(gpu.distvec, ref.distvec, 3 * P) &&
                         same_bits(gpu.absdist, ref.absdist, P);

    double *dpos = nullptr;
    bool dev_ok = hipMalloc(&dpos, b.r.size() * sizeof(double)) == hipSuccess &&
                  hipMemcpy(dpos, b.r.data(), b.r.size() * sizeof(double),
                            hipMemcpyHostToDevice) == hipSuccess;
    matscipy::NeighbourListRequest req;
    req.quantities = kAll;
    req.cell_origin = b.origin;
    req.cell = b.cell;
    req.inv_cell = b.inv;
    req.pbc = b.pbc;
    req.nat = N;
    req.positions = dpos;
    req.positions_on_device = true;
    req.cutoff = cutoff;
    req.order = order;
    matscipy::NeighbourListDevice dev;
    dev_ok = dev_ok && matscipy::neighbour_list_gpu_device(req, dev) == matscipy::NL_SUCCESS;
    dev_ok = dev_ok && static_cast<size_t>(dev.npairs) == P && dev.first.size() == P &&
             same_bits(to_host(dev.first), ref.first, P) &&
             same_bits(to_host(dev.secnd), ref.secnd, P) &&
             same_bits(to_host(dev.shift), ref.shift, 3 * P) &&
             same_bits(to_host(dev.distvec), ref.distvec, 3 * P) &&
             same_bits(to_host(dev.absdist), ref.absdist, P);
    (void)hipFree(dpos);
    oracle_free(&ref);
    std::printf("  %s: N %d, L %.2f, rc %.2f, %zu pairs\n", g_test.c_str(), N, L, cutoff, P);
    EXPECT(host_ok, "neighbour_list_gpu equals the oracle bit for bit");
    EXPECT(dev_ok, "neighbour_list_gpu_device equals the oracle bit for bit");
}

/* TEST(NeighbourListGpu, MatrixMatchesCpu): rows of K = 80 slots. */
void matrix_matches_cpu() {
    const int N = 3000, K = 80;
    const double cutoff = 1.4;
    const Box b = make_box(N, 14.0, 4);
    oracle_result ref;
    EXPECT(oracle_list(b, cutoff, ref), "oracle build");
    /* the rows the reference's neighbour_matrix writes: atom i's pairs in
       list order, the first K of them, unused slots 0 */
    std::vector<index_t> widx(static_cast<size_t>(N) * K, 0), wcount(N, 0);
    std::vector<real_t> wdist(static_cast<size_t>(N) * K * 3, 0.0);
    bool wover = false;
    for (int64_t p = 0; p < ref.npairs; p++) {
        const index_t i = ref.first[p];
        const index_t s = wcount[i]++;
        if (s >= K) {
            wover = true;
            continue;
        }
        widx[static_cast<size_t>(i) * K + s] = ref.secnd[p];
        for (int k = 0; k < 3; k++)
            wdist[(static_cast<size_t>(i) * K + s) * 3 + k] = ref.distvec[3 * p + k];
    }
    oracle_free(&ref);
    matscipy::NeighbourListRequest req;
    req.cell_origin = b.origin;
    req.cell = b.cell;
    req.inv_cell = b.inv;
    req.pbc = b.pbc;
    req.nat = N;
    req.positions = b.r.data();
    req.cutoff = cutoff;
    matscipy::NeighbourMatrixDevice dev;
    EXPECT(matscipy::neighbour_matrix_gpu_device(req, K, dev) == matscipy::NL_SUCCESS,
           "neighbour_matrix_gpu_device");
    EXPECT(!wover && !dev.overflow, "no row overflows");
    EXPECT(dev.max_neighbours == K && dev.n == N, "shape");
    EXPECT(same_bits(widx, to_host(dev.idx).data(), widx.size()), "idx");
    EXPECT(same_bits(wcount, to_host(dev.count).data(), wcount.size()), "count");
    EXPECT(same_bits(wdist, to_host(dev.dist).data(), wdist.size()), "dist bits");
    /* a capacity below the longest row: rows clipped, overflow set */
    matscipy::NeighbourMatrixDevice small;
    EXPECT(matscipy::neighbour_matrix_gpu_device(req, 4, small) == matscipy::NL_SUCCESS,
           "clipped matrix");
    EXPECT(small.overflow, "overflow reported");
    const std::vector<index_t> cidx = to_host(small.idx);
    bool clipped_ok = true;
    for (int i = 0; i < N && clipped_ok; i++)
        for (int s = 0; s < 4; s++)
            clipped_ok = clipped_ok &&
                         cidx[static_cast<size_t>(i) * 4 + s] == widx[static_cast<size_t>(i) * K + s];
    EXPECT(clipped_ok, "clipped rows are the first slots of the full rows");
    std::printf("  %s: N %d, K %d\n", g_test.c_str(), N, K);
}

/* TEST(NeighbourListGpu, CoordinationMatchesCpu) */
void coordination_matches_cpu() {
    const int N = 3000;
    const double cutoff = 1.2;
    const Box b = make_box(N, 14.0, 7);
    oracle_result ref;
    EXPECT(oracle_list(b, cutoff, ref), "oracle build");
    std::vector<index_t> want(N, 0);
    for (int64_t p = 0; p < ref.npairs; p++) want[ref.first[p]]++;
    oracle_free(&ref);
    matscipy::NeighbourListRequest req;
    req.cell_origin = b.origin;
    req.cell = b.cell;
    req.inv_cell = b.inv;
    req.pbc = b.pbc;
    req.nat = N;
    req.positions = b.r.data();
    req.cutoff = cutoff;
    matscipy::NeighbourListDevice dev;
    EXPECT(matscipy::neighbour_count_gpu_device(req, dev) == matscipy::NL_SUCCESS,
           "neighbour_count_gpu_device");
    EXPECT(dev.counts.size() == static_cast<size_t>(N), "counts size");
    EXPECT(to_host(dev.counts) == want, "counts");
    std::printf("  %s: N %d\n", g_test.c_str(), N);
}

/* Errors come back as NL_ERROR with a message (no device work). */
void errors() {
    matscipy::NeighbourListRequest req;
    matscipy::NeighbourListDevice dev;
    EXPECT(matscipy::neighbour_list_gpu_device(req, dev) == matscipy::NL_ERROR,
           "missing geometry refused");
    EXPECT(std::strlen(matscipy::get_error_msg()) > 0, "message set");
    EXPECT(dev.npairs == 0 && dev.first.empty(), "result left empty");
}

void run(const char *name, void (*fn)()) {
    g_test = name;
    const int before = g_failed;
    fn();
    std::printf("%s %s\n", g_failed == before ? "[ OK ]" : "[FAIL]", name);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char **argv) {
    run("Errors", errors);
    if (argc > 1 && std::strcmp(argv[1], "--no-gpu") == 0) {
        std::printf("cpp api: %d checks passed, %d failed (no device calls)\n", g_checks,
                    g_failed);
        return g_failed ? 1 : 0;
    }
    if (mnl_device_count() < 1) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    run("MatchesCpuDense", [] { compare(2000, 12.0, 1.0, 1); });
    run("MatchesCpuFewLargeCells", [] { compare(500, 3.5, 1.0, 2); });
    run("MatchesCpuLargerCutoff", [] { compare(3000, 15.0, 2.0, 3); });
    run("MatchesCpuMortonOrder", [] {
        compare(2000, 12.0, 1.0, 1, matscipy::CellOrder::Morton);
        compare(3000, 15.0, 2.0, 3, matscipy::CellOrder::Morton);
    });
    run("MatchesCpuSparseVacuum", [] {
        const double Lc = std::cbrt(4000 / 12.0);
        compare(4000, Lc * 40.0, 1.0, 11, matscipy::CellOrder::Linear, Lc);
    });
    run("MatrixMatchesCpu", matrix_matches_cpu);
    run("CoordinationMatchesCpu", coordination_matches_cpu);
    std::printf("cpp api: %d checks passed, %d failed\n", g_checks, g_failed);
    return g_failed ? 1 : 0;
}

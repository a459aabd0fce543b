/*
 * A C++ consumer of the C ABI (include/mnl.h), built the way the reference's
 * C++ core would bind it: plain C calls on host and device pointers, no
 * Python. Test infrastructure: it links the oracle (the CPU restatement of
 * the reference path) as the checker.
 *
 *   capi_consumer --no-gpu   calls that need no device (version, device
 *                            count, argument errors); exit 0 if they behave
 *   capi_consumer            every check below on device 0
 *
 * Checks, each bit for bit against the oracle on the same inputs:
 *   - the one-shot entry point (host positions), all five quantities;
 *   - the request form with CellOrder Morton on a stream of its own;
 *   - the plan API, synchronous count + fill into device buffers;
 *   - the plan API, asynchronous count + fill with a short capacity
 *     (the first `capacity` pairs written, the full total reported);
 *   - a triclinic cell, pbc (F,T,F), atoms outside the cell.
 * The reference's tests of the same path: tests/test_neighbour_list_gpu.cc
 * (device build against the CPU build on the same inputs).
 */
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mnl.h"
#include "mnl_oracle.h"

extern "C" int capi_consumer_c_checks(void); /* capi_consumer_c.c */

namespace {

int g_checks = 0;

#define CHECK(cond, what)                                                          \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::fprintf(stderr, "FAILED: %s (%s:%d) [%s]\n", what, __FILE__,      \
                         __LINE__, mnl_last_error());                              \
            return false;                                                          \
        }                                                                          \
        ++g_checks;                                                                \
    } while (0)

#define HIPCHECK(call)                                                             \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                      \
            return false;                                                          \
        }                                                                          \
    } while (0)

constexpr int kAll = MNL_Q_FIRST | MNL_Q_SECOND | MNL_Q_DISTVEC | MNL_Q_ABSDIST | MNL_Q_SHIFT;

/* One test system: positions, cell (rows are lattice vectors), its inverse
   transposed as the reference's Python layer passes it (inv(cell.T)), pbc. */
struct System {
    std::vector<double> r;
    double cell[9], inv[9], origin[3] = {0.0, 0.0, 0.0};
    int32_t pbc[3];
    int32_t nat;
    double cutoff;
};

/* 64-bit LCG, uniform in [0, 1). */
struct Lcg {
    uint64_t s;
    double next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<double>(s >> 11) * (1.0 / 9007199254740992.0);
    }
};

/* inv = inverse of cell^T, by the adjugate. */
void inverse_of_transpose(const double c[9], double inv[9]) {
    double t[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) t[3 * a + b] = c[3 * b + a];
    const double det = t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
                       t[2] * (t[3] * t[7] - t[4] * t[6]);
    inv[0] = (t[4] * t[8] - t[5] * t[7]) / det;
    inv[1] = (t[2] * t[7] - t[1] * t[8]) / det;
    inv[2] = (t[1] * t[5] - t[2] * t[4]) / det;
    inv[3] = (t[5] * t[6] - t[3] * t[8]) / det;
    inv[4] = (t[0] * t[8] - t[2] * t[6]) / det;
    inv[5] = (t[2] * t[3] - t[0] * t[5]) / det;
    inv[6] = (t[3] * t[7] - t[4] * t[6]) / det;
    inv[7] = (t[1] * t[6] - t[0] * t[7]) / det;
    inv[8] = (t[0] * t[4] - t[1] * t[3]) / det;
}

/* Atoms at fractional coordinates uniform in [lo, hi) of the cell. */
System make_system(uint64_t seed, int32_t nat, const double cell[9], const int32_t pbc[3],
                   double cutoff, double lo, double hi) {
    System s;
    std::memcpy(s.cell, cell, sizeof(s.cell));
    std::memcpy(s.pbc, pbc, sizeof(s.pbc));
    inverse_of_transpose(cell, s.inv);
    s.nat = nat;
    s.cutoff = cutoff;
    s.r.resize(3 * static_cast<size_t>(nat));
    Lcg g{seed};
    for (int32_t a = 0; a < nat; a++) {
        double f[3];
        for (double &x : f) x = lo + (hi - lo) * g.next();
        for (int k = 0; k < 3; k++)
            s.r[3 * a + k] = f[0] * cell[k] + f[1] * cell[3 + k] + f[2] * cell[6 + k];
    }
    return s;
}

/* Host copy of one build's five quantities. */
struct Pairs {
    std::vector<int32_t> i, j, S;
    std::vector<double> D, d;
    int64_t n = 0;
    void resize(int64_t p) {
        n = p;
        i.resize(p);
        j.resize(p);
        S.resize(3 * p);
        D.resize(3 * p);
        d.resize(p);
    }
};

bool oracle_pairs(const System &s, Pairs &out) {
    oracle_result res{};
    CHECK(oracle_neighbour_list(kAll, s.origin, s.cell, s.inv, s.pbc, s.nat, s.r.data(),
                                s.cutoff, nullptr, nullptr, 0, nullptr, -1, &res) == 0,
          "oracle build");
    out.resize(res.npairs);
    std::memcpy(out.i.data(), res.first, 4 * res.npairs);
    std::memcpy(out.j.data(), res.secnd, 4 * res.npairs);
    std::memcpy(out.S.data(), res.shift, 12 * res.npairs);
    std::memcpy(out.D.data(), res.distvec, 24 * res.npairs);
    std::memcpy(out.d.data(), res.absdist, 8 * res.npairs);
    oracle_free(&res);
    return true;
}

/* The first n pairs of got equal those of want, bit for bit. */
bool same_prefix(const Pairs &got, const Pairs &want, int64_t n, const char *what) {
    CHECK(n <= got.n && n <= want.n, what);
    CHECK(std::memcmp(got.i.data(), want.i.data(), 4 * n) == 0, what);
    CHECK(std::memcmp(got.j.data(), want.j.data(), 4 * n) == 0, what);
    CHECK(std::memcmp(got.S.data(), want.S.data(), 12 * n) == 0, what);
    CHECK(std::memcmp(got.D.data(), want.D.data(), 24 * n) == 0, what);
    CHECK(std::memcmp(got.d.data(), want.d.data(), 8 * n) == 0, what);
    return true;
}

bool from_result(mnl_result *res, Pairs &out) {
    out.resize(mnl_result_npairs(res));
    CHECK(mnl_result_copy_to_host(res, 'i', out.i.data()) == MNL_SUCCESS, "copy i");
    CHECK(mnl_result_copy_to_host(res, 'j', out.j.data()) == MNL_SUCCESS, "copy j");
    CHECK(mnl_result_copy_to_host(res, 'S', out.S.data()) == MNL_SUCCESS, "copy S");
    CHECK(mnl_result_copy_to_host(res, 'D', out.D.data()) == MNL_SUCCESS, "copy D");
    CHECK(mnl_result_copy_to_host(res, 'd', out.d.data()) == MNL_SUCCESS, "copy d");
    return true;
}

bool check_one_shot(const System &s, const Pairs &want) {
    mnl_result *res = nullptr;
    CHECK(mnl_neighbour_list_device(kAll, s.origin, s.cell, s.inv, s.pbc, s.nat, s.r.data(), 0,
                                    s.cutoff, nullptr, nullptr, 0, nullptr, 0,
                                    &res) == MNL_SUCCESS,
          "one-shot build");
    CHECK(mnl_result_nat(res) == s.nat, "one-shot nat");
    Pairs got;
    const bool ok = from_result(res, got);
    mnl_result_free(res);
    if (!ok) return false;
    CHECK(got.n == want.n, "one-shot pair total");
    return same_prefix(got, want, want.n, "one-shot pairs");
}

bool check_request(const System &s, const Pairs &want) {
    hipStream_t st = nullptr;
    HIPCHECK(hipStreamCreate(&st));
    mnl_request rq;
    std::memset(&rq, 0, sizeof(rq));
    rq.quantities = kAll;
    rq.cell_origin = s.origin;
    rq.cell = s.cell;
    rq.inv_cell = s.inv;
    rq.pbc = s.pbc;
    rq.nat = s.nat;
    rq.positions = s.r.data();
    rq.cutoff = s.cutoff;
    rq.order = MNL_ORDER_MORTON;
    rq.device_id = 0;
    rq.stream = st;
    rq.backend = MNL_BACKEND_AUTO;
    mnl_result *res = nullptr;
    const int rc = mnl_neighbour_list_request(&rq, &res);
    Pairs got;
    const bool ok = rc == MNL_SUCCESS && from_result(res, got);
    if (res) mnl_result_free(res);
    HIPCHECK(hipStreamDestroy(st));
    CHECK(ok, "request build (Morton, own stream)");
    CHECK(got.n == want.n, "request pair total");
    return same_prefix(got, want, want.n, "request pairs");
}

/* Device buffers of one pair list. */
struct DevPairs {
    int32_t *i = nullptr, *j = nullptr, *S = nullptr;
    double *D = nullptr, *d = nullptr;
    ~DevPairs() {
        (void)hipFree(i);
        (void)hipFree(j);
        (void)hipFree(S);
        (void)hipFree(D);
        (void)hipFree(d);
    }
    bool alloc(int64_t n) {
        const size_t m = static_cast<size_t>(n > 0 ? n : 1);
        HIPCHECK(hipMalloc(&i, 4 * m));
        HIPCHECK(hipMalloc(&j, 4 * m));
        HIPCHECK(hipMalloc(&S, 12 * m));
        HIPCHECK(hipMalloc(&D, 24 * m));
        HIPCHECK(hipMalloc(&d, 8 * m));
        return true;
    }
    bool to_host(int64_t n, Pairs &out) const {
        out.resize(n);
        HIPCHECK(hipMemcpy(out.i.data(), i, 4 * n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(out.j.data(), j, 4 * n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(out.S.data(), S, 12 * n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(out.D.data(), D, 24 * n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(out.d.data(), d, 8 * n, hipMemcpyDeviceToHost));
        return true;
    }
};

bool check_plan(const System &s, const Pairs &want) {
    double *pos = nullptr;
    HIPCHECK(hipMalloc(&pos, 24 * static_cast<size_t>(s.nat)));
    HIPCHECK(hipMemcpy(pos, s.r.data(), 24 * static_cast<size_t>(s.nat), hipMemcpyHostToDevice));
    mnl_plan *plan = nullptr;
    bool ok = mnl_plan_create(0, &plan) == MNL_SUCCESS;
    int64_t np = -1;
    DevPairs out;
    Pairs got;
    ok = ok && mnl_plan_count(plan, nullptr, s.origin, s.cell, s.inv, s.pbc, s.nat, pos,
                              s.cutoff, nullptr, nullptr, 0, nullptr, MNL_BACKEND_AUTO,
                              &np) == MNL_SUCCESS;
    ok = ok && np == want.n && out.alloc(np);
    ok = ok && mnl_plan_fill(plan, nullptr, out.i, out.j, out.D, out.d, out.S) == MNL_SUCCESS;
    ok = ok && out.to_host(np, got);
    bool same = ok && same_prefix(got, want, want.n, "plan pairs");

    /* asynchronous form on a stream of its own, capacity short of the total */
    hipStream_t st = nullptr;
    int64_t *np_dev = nullptr;
    bool async_ok = ok && hipStreamCreate(&st) == hipSuccess &&
                    hipMalloc(&np_dev, sizeof(int64_t)) == hipSuccess;
    const int64_t cap = want.n > 10 ? want.n - 10 : want.n;
    DevPairs out2;
    async_ok = async_ok && out2.alloc(want.n);
    async_ok = async_ok && mnl_plan_count_async(plan, st, s.origin, s.cell, s.inv, s.pbc, s.nat,
                                                pos, s.cutoff, nullptr, nullptr, 0, nullptr,
                                                MNL_BACKEND_AUTO, np_dev) == MNL_SUCCESS;
    async_ok = async_ok && mnl_plan_fill_async(plan, st, cap, out2.i, out2.j, out2.D, out2.d,
                                               out2.S, 0) == MNL_SUCCESS;
    int64_t np_async = -1;
    async_ok = async_ok && hipMemcpyAsync(&np_async, np_dev, sizeof(int64_t),
                                          hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipStreamSynchronize(st) == hipSuccess;
    Pairs got2;
    async_ok = async_ok && out2.to_host(cap, got2);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(np_dev);
    mnl_plan_destroy(plan);
    (void)hipFree(pos);
    CHECK(ok, "plan count + fill");
    CHECK(same, "plan pairs");
    CHECK(async_ok, "plan async count + fill");
    CHECK(np_async == want.n, "async pair total beyond the capacity");
    return same_prefix(got2, want, cap, "async pairs within the capacity");
}

bool check_errors() {
    mnl_result *res = nullptr;
    const double cell[9] = {10, 0, 0, 0, 10, 0, 0, 0, 10};
    const double z3[3] = {0, 0, 0};
    const int32_t pbc[3] = {1, 1, 1};
    CHECK(mnl_neighbour_list_device(kAll, z3, cell, cell, pbc, 4, nullptr, 0, 2.5, nullptr,
                                    nullptr, 0, nullptr, 0, &res) == MNL_ERROR,
          "NULL positions refused");
    CHECK(std::strlen(mnl_last_error()) > 0, "error message set");
    CHECK(mnl_plan_fill(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) ==
              MNL_ERROR,
          "NULL plan refused");
    return true;
}

bool no_gpu_checks() {
    CHECK(mnl_version() && std::strlen(mnl_version()) > 0, "version string");
    CHECK(mnl_device_count() >= 0, "device count");
    CHECK(mnl_plan_fill(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) ==
              MNL_ERROR,
          "NULL plan refused");
    CHECK(std::string(mnl_last_error()).find("plan") != std::string::npos, "message names it");
    CHECK(capi_consumer_c_checks() == 0, "C translation unit");
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    const bool gpu = !(argc > 1 && std::strcmp(argv[1], "--no-gpu") == 0);
    if (!no_gpu_checks()) return 1;
    if (!gpu) {
        std::printf("capi consumer: %d checks passed (no device calls)\n", g_checks);
        return 0;
    }
    if (mnl_device_count() < 1) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    /* the 1000-atom uniform case: 10 A cube, pbc TTT, rc 2.5 */
    const double cube[9] = {10, 0, 0, 0, 10, 0, 0, 0, 10};
    const int32_t ttt[3] = {1, 1, 1};
    const System a = make_system(1, 1000, cube, ttt, 2.5, 0.0, 1.0);
    /* triclinic, pbc (F,T,F), atoms up to 20% outside the cell */
    const double tri[9] = {9.0, 0.0, 0.0, 1.5, 8.0, 0.0, -0.7, 0.9, 11.0};
    const int32_t ftf[3] = {0, 1, 0};
    const System b = make_system(2, 800, tri, ftf, 3.1, -0.2, 1.2);
    for (const System *s : {&a, &b}) {
        Pairs want;
        if (!oracle_pairs(*s, want)) return 1;
        std::printf("system of %d atoms: %lld pairs\n", s->nat, static_cast<long long>(want.n));
        std::fflush(stdout);
        if (!check_one_shot(*s, want) || !check_request(*s, want) || !check_plan(*s, want))
            return 1;
    }
    if (!check_errors()) return 1;
    std::printf("capi consumer: %d checks passed\n", g_checks);
    return 0;
}

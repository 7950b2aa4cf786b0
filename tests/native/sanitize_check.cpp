// Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer (the
// CPU build of tests/test_sanitize.py): the CPU oracle (plain, multi-rank,
// immersed boundary with the limiter, viscous, its LDL^T and stencil
// helpers), the library's host immersed-boundary set-up (cfs_circle_setup),
// its host helpers (configuration checks, immersed-boundary index tables,
// per-stage kernel arguments) and the batch arena allocator against a model.
// Needs no GPU.  Prints one line per part and exits 0 when every check holds;
// a sanitizer finding aborts the process.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <vector>

#include "../../2d_compressible_flow_solver_amd/csrc/arena.h"
#include "../../2d_compressible_flow_solver_amd/csrc/host.hpp"

extern "C" {
struct oracle_solver;
oracle_solver* oracle_create(const cfs_config* cfg, int nthreads);
void oracle_destroy(oracle_solver* s);
int oracle_set_ranks(oracle_solver* s, int n, int* dims_out);
int oracle_set_state(oracle_solver* s, const double* U);
int oracle_get_state(oracle_solver* s, double* U);
int oracle_set_circle(oracle_solver* s, double cx, double cy, double r, int npts, int layers);
int oracle_step(oracle_solver* s, double dt);
int oracle_stage(oracle_solver* s, int counter, double dt);
double oracle_cfl_smax(oracle_solver* s);
int oracle_checksums(oracle_solver* s, double out[4]);
void oracle_ldlt_factor(double* A16, int32_t* transp);
void oracle_ldlt_solve(const double* A16, const int32_t* transp, double* x);
void oracle_collect_stencil(int f, int N, int periodic, int out[6]);
}

static int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

static cfs_config base(int nx, int ny, double lx, double ly) {
    cfs_config c{};
    c.nx = nx;
    c.ny = ny;
    c.lx = lx;
    c.ly = ly;
    c.gamma = 1.4;
    c.rk = 3;
    c.precision = CFS_PREC_FP64;
    c.prandtl = 0.72;
    c.ma_ref = 1.0;
    return c;
}

static void cons(double rho, double u, double v, double p, double out[4]) {
    out[0] = rho;
    out[1] = rho * u;
    out[2] = rho * v;
    out[3] = p / 0.4 + 0.5 * rho * (u * u + v * v);
}

static std::vector<double> quadrants(const cfs_config& c) {
    const size_t n = (size_t)c.nx * c.ny;
    std::vector<double> U(4 * n);
    for (int i = 0; i < c.nx; ++i)
        for (int j = 0; j < c.ny; ++j) {
            const bool e = i >= c.nx / 2, nn = j >= c.ny / 2;
            double s[4];
            if (e && nn) cons(1.5, 0, 0, 1.5, s);
            else if (!e && nn) cons(0.5323, 1.206, 0, 0.3, s);
            else if (!e && !nn) cons(0.138, 1.206, 1.206, 0.029, s);
            else cons(0.5323, 0, 1.206, 0.3, s);
            for (int q = 0; q < 4; ++q) U[q * n + (size_t)i * c.ny + j] = s[q];
        }
    return U;
}

static bool finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

static void oracle_part() {
    // plain and multi-rank Riemann runs agree bit for bit
    cfs_config c = base(33, 29, 1.0, 1.0);
    std::vector<double> U = quadrants(c), A(U.size()), B(U.size());
    oracle_solver* a = oracle_create(&c, 2);
    oracle_solver* b = oracle_create(&c, 4);
    int dims[2];
    CHECK(oracle_set_ranks(b, 4, dims) == 0);
    oracle_set_state(a, U.data());
    oracle_set_state(b, U.data());
    for (int k = 0; k < 3; ++k) {
        const double dt = 0.5 / oracle_cfl_smax(a);
        oracle_step(a, dt);
        oracle_step(b, dt);
    }
    oracle_get_state(a, A.data());
    oracle_get_state(b, B.data());
    CHECK(finite(A) && std::memcmp(A.data(), B.data(), A.size() * sizeof(double)) == 0);
    double cs[4];
    oracle_checksums(a, cs);
    CHECK(std::isfinite(cs[0]) && cs[0] > 0);
    oracle_destroy(a);
    oracle_destroy(b);
    // immersed boundary (the oracle's own set-up) with the positivity limiter
    cfs_config cy = base(96, 48, 4.0, 2.0);
    cy.positivity = 2;
    cy.bc[CFS_SIDE_LEFT] = CFS_BC_DIRICHLET;
    cons(1.4, 3.0, 0.0, 1.0, cy.bc_state[CFS_SIDE_LEFT]);
    oracle_solver* o = oracle_create(&cy, 2);
    CHECK(oracle_set_circle(o, 1.0, 1.0, 0.25, 400, 3) == 0);
    const size_t n = 96 * 48;
    std::vector<double> V(4 * n);
    for (size_t k = 0; k < n; ++k) {
        double s[4];
        cons(1.4, 3.0, 0.0, 1.0, s);
        for (int q = 0; q < 4; ++q) V[q * n + k] = s[q];
    }
    oracle_set_state(o, V.data());
    // the oracle's own set-up is the reference's LDL^T interpolation, which
    // can turn this flow non-finite within a few steps: only the calls are
    // checked here (and everything they touch, by the sanitizers)
    for (int k = 0; k < 2; ++k) CHECK(oracle_step(o, 1e-4) == 0);
    oracle_get_state(o, V.data());
    oracle_destroy(o);
    // viscous terms, periodic, stage by stage
    cfs_config cv = base(24, 20, 1.0, 1.0);
    cv.viscous = 1;
    cv.mu = 1e-3;
    for (int s = 0; s < 4; ++s) cv.bc[s] = CFS_BC_PERIODIC;
    std::vector<double> W(4 * 24 * 20);
    for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 20; ++j) {
            double s[4];
            cons(1.0 + 0.1 * std::sin(0.3 * i), 0.2 * std::cos(0.5 * j), 0.1, 2.5, s);
            for (int q = 0; q < 4; ++q) W[q * 480 + i * 20 + j] = s[q];
        }
    o = oracle_create(&cv, 1);
    oracle_set_state(o, W.data());
    for (int i = 1; i <= 4; ++i) oracle_stage(o, i, 1e-3);
    oracle_get_state(o, W.data());
    CHECK(finite(W));
    oracle_destroy(o);
    // LDL^T of a donor matrix and its solve; face stencils at every face of N = 8
    double M[16] = {0.25, 0.5, 0.5, 1, 0.75, 1.5, 0.5, 1, 0.5, 0.5, 1, 1, 1.5, 1.5, 1, 1}, x[4] = {1, 2, 3, 4};
    int32_t tr[4];
    oracle_ldlt_factor(M, tr);
    oracle_ldlt_solve(M, tr, x);
    CHECK(std::isfinite(x[0]) && std::isfinite(x[3]));
    for (int f = 0; f <= 8; ++f) {
        int st[6];
        oracle_collect_stencil(f, 8, 0, st);
        oracle_collect_stencil(f, 8, 1, st);
    }
    std::printf("oracle ok\n");
}

static void host_part() {
    // host immersed-boundary set-up: count query, tables, both interpolations
    for (int interp : {CFS_IB_REFERENCE, CFS_IB_BILINEAR_POSITIVE}) {
        for (int g = 0; g < 3; ++g) {
            const int nx = 64 + 48 * g, ny = 32 + 24 * g;
            cfs_config c = base(nx, ny, 4.0, 2.0);
            std::vector<uint8_t> type((size_t)nx * ny);
            int32_t n_ig = 0;
            const double r = 0.2 + 0.1 * g;
            int rc = cfs_circle_setup(&c, 1.0, 1.0, r, 400, 3, interp, 0, type.data(), &n_ig, nullptr, nullptr,
                                      nullptr, nullptr, nullptr);
            CHECK(rc == CFS_E_ARG && n_ig > 0);
            std::vector<int32_t> ig(n_ig), donors(4 * (size_t)n_ig), transp(4 * (size_t)n_ig);
            std::vector<double> xy(2 * (size_t)n_ig), coef(10 * (size_t)n_ig);
            rc = cfs_circle_setup(&c, 1.0, 1.0, r, 400, 3, interp, n_ig, type.data(), &n_ig, ig.data(), donors.data(),
                                  xy.data(), coef.data(), transp.data());
            CHECK(rc == CFS_OK);
            // the library's index tables of these, for every tile shape
            std::vector<int> index;
            std::vector<uint8_t> tiles;
            for (int k = 0; k < cfs::NUM_SHAPES; ++k) {
                CHECK(cfs::build_ib_index(c, k, interp, type.data(), n_ig, ig.data(), donors.data(), transp.data(),
                                          index, tiles) == CFS_OK);
                CHECK((long long)tiles.size() == cfs::tiles_for(c, k));
            }
            // a table naming a fluid cell as a ghost is refused
            if (n_ig > 0) {
                std::vector<uint8_t> bad = type;
                bad[ig[0]] = CFS_CELL_FLUID;
                CHECK(cfs::build_ib_index(c, 0, interp, bad.data(), n_ig, ig.data(), donors.data(), transp.data(),
                                          index, tiles) == CFS_E_ARG);
            }
        }
    }
    // configuration checks and the stage arguments of every sub-stage
    cfs_config c = base(40, 40, 1.0, 1.0);
    CHECK(cfs::check_config(&c) == CFS_OK);
    cfs_config bad = c;
    bad.nx = 2;
    CHECK(cfs::check_config(&bad) == CFS_E_ARG);
    bad = c;
    bad.bc[0] = CFS_BC_PERIODIC;   // periodic on one side only
    CHECK(cfs::check_config(&bad) == CFS_E_ARG);
    bad = c;
    bad.viscous = 1;
    bad.positivity = 2;
    CHECK(cfs::check_config(&bad) == CFS_E_UNSUPPORTED);
    std::vector<double> scal(cfs::SCAL_DOUBLES), planes(12 * 1600);
    std::vector<int> exact(2);
    cfs::ProblemDev p;
    p.cfg = &c;
    p.plane = 1600;
    p.planes = planes.data();
    p.scal = scal.data();
    p.exact = exact.data();
    p.tiles_i = 3;
    p.tiles_j = 3;
    for (int rk : {2, 3}) {
        c.rk = rk;
        for (int sub = 0; sub < rk; ++sub) {
            cfs::StageArgs a;
            const int next = cfs::fill_stage_args(a, p, sub, sub & 1, sub & 1, cfs::cfl_stage_time(rk, sub, 0.5, 0.1));
            CHECK(next >= 0 && next < 3 && a.out != a.in && a.form == (sub == 0 ? 0 : sub));
        }
    }
    std::printf("host ok\n");
}

static void arena_part() {
    // random allocations and releases against a model of the occupied ranges
    const size_t size = 1 << 20;
    cfs::Arena ar;
    ar.reset(size);
    std::map<size_t, size_t> used;   // offset -> bytes
    std::mt19937_64 rng(7);
    size_t in_use = 0;
    for (int it = 0; it < 20000; ++it) {
        if (used.empty() || rng() % 3 != 0) {
            const size_t bytes = 256 * (1 + rng() % 64);
            size_t off = 0;
            if (ar.alloc(bytes, &off)) {
                CHECK(off + bytes <= size && off % 256 == 0);
                auto nx = used.lower_bound(off);
                if (nx != used.end()) CHECK(off + bytes <= nx->first);
                if (nx != used.begin()) {
                    auto pv = std::prev(nx);
                    CHECK(pv->first + pv->second <= off);
                }
                used[off] = bytes;
                in_use += bytes;
            }
        } else {
            auto it2 = used.begin();
            std::advance(it2, (long)(rng() % used.size()));
            ar.release(it2->first, it2->second);
            in_use -= it2->second;
            used.erase(it2);
        }
        CHECK(ar.free_bytes() == size - in_use);
    }
    for (auto& kv : used) ar.release(kv.first, kv.second);
    size_t off = 0;
    CHECK(ar.free_bytes() == size && ar.alloc(size, &off) && off == 0);   // fully coalesced
    std::printf("arena ok\n");
}

int main() {
    oracle_part();
    host_part();
    arena_part();
    std::printf("%d\n", failures);
    return failures ? 1 : 0;
}

// sanitize_driver.cpp -- host code under AddressSanitizer/UBSan.
//
// Built by tests/test_sanitizers.py with g++ -fsanitize=address,undefined
// together with the library's host physics (csrc/physics.cpp: model
// derivations, schedule generation, spec validation, grid parsing, lattice
// budgets) and the CPU oracle (oracle/oracle.cpp).  It drives both with
// ordinary and hostile inputs; any out-of-bounds access, leak or undefined
// operation aborts the run with the sanitizer's report.  No HIP code is
// involved: the library's device side is not built here.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bubblesim.h"

// physics.cpp reports errors through the library's thread-local error slot,
// which lives with the HIP-facing C API; a host-only stand-in suffices here
namespace bs {
int set_error(int rc, const std::string &) { return rc; }
void clear_error() {}
}  // namespace bs

extern "C" {
void *oracle_create(int precision, const float *params, const float *events, int nev);
void oracle_destroy(void *h);
int oracle_advance(void *h, long long nsteps);
void oracle_get_state(void *h, void *field, void *fluid);
}

static int failures = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                   \
        }                                                                 \
    } while (0)

static void grids() {
    uint32_t g[3];
    const char *ok[] = {"192", "256,128, 64", "8,8,8", " 32 , 33 , 34 "};
    for (const char *s : ok) CHECK(bs_parse_grid(s, g) == BS_OK);
    const char *bad[] = {"", "1,2", "abc", "1,2,3,4", "-5", ",,", "99999999999999999999", "8,,8",
                         "8,8,8,", "\t", "1e3", "0", "7"};
    for (const char *s : bad) (void)bs_parse_grid(s, g);  // must not crash; some are invalid
    std::string longs(10000, '9');
    (void)bs_parse_grid(longs.c_str(), g);
}

static void schedules() {
    bs_model m;
    bs_model_default(&m);
    CHECK(bs_model_validate(&m) == BS_OK);
    bs_derived d;
    CHECK(bs_model_derive(&m, &d) == BS_OK);
    const uint32_t grids3[][3] = {{8, 8, 8}, {32, 32, 32}, {192, 192, 192}, {17, 8, 33}};
    for (const auto &g : grids3) {
        for (int32_t mode = 0; mode < 3; ++mode) {
            for (uint32_t n : {1u, 2u, 12u, 200u, 1024u}) {
                std::vector<bs_event> ev(n);
                const int rc = bs_make_schedule(mode, n, 15.33f, g, 60.0f, 1, 200, 0x5EED, 0.2f, ev.data());
                if (rc == BS_OK) {
                    std::vector<uint64_t> steps(n);
                    CHECK(bs_event_steps(ev.data(), n, 0.2f, steps.data()) == BS_OK);
                }
            }
        }
    }
    std::vector<uint64_t> u(1000);
    std::vector<float> f(1000);
    bs_pcg64mcg_u64(42, 1000, u.data());
    bs_pcg64mcg_f32(7, 1000, f.data());
    for (float x : f) CHECK(x >= 0.0f && x < 1.0f);
}

static void specs() {
    bs_spec s;
    bs_spec_default(&s);
    CHECK(bs_spec_validate(&s) == BS_OK);
    bs_spec t = s;
    t.grid[1] = 7;
    CHECK(bs_spec_validate(&t) != BS_OK);
    t = s;
    t.grid[0] = 16384;
    t.grid[1] = 16384;
    CHECK(bs_spec_validate(&t) != BS_OK);
    t = s;
    t.model.cfl = 0.51f;
    CHECK(bs_spec_validate(&t) != BS_OK);
    t = s;
    t.model.eps_ratio = 1.0f;
    CHECK(bs_spec_validate(&t) != BS_OK);
    t = s;
    t.bubbles = 0;
    CHECK(bs_spec_validate(&t) != BS_OK);
    uint32_t side = 0;
    int32_t limited = 0;
    for (uint64_t budget : {0ull, 1ull << 20, 1ull << 30, 288ull << 30, ~0ull})
        for (int32_t prec : {BS_PRECISION_F32, BS_PRECISION_F64})
            for (int32_t kern : {BS_KERNEL_FUSED, BS_KERNEL_TWO_PASS})
                (void)bs_max_cubic_size(budget, prec, kern, &side, &limited);
    const uint32_t g[3] = {512, 512, 512};
    CHECK(bs_lattice_bytes(g, BS_PRECISION_F32, BS_KERNEL_FUSED) > 0);
    (void)bs_lattice_budget(1ull << 40, 1ull << 30);
}

static void oracle_run(int precision) {
    bs_model m;
    bs_model_default(&m);
    const uint32_t g[3] = {16, 12, 10};
    bs_sim_params p;
    CHECK(bs_sim_params_fill(&m, g, 0.0f, &p) == BS_OK);
    float words[24];
    static_assert(sizeof(bs_sim_params) == sizeof(words), "96-byte parameter block");
    std::memcpy(words, &p, sizeof words);
    const float events[2][5] = {{0.0f, 8.0f, 6.0f, 5.0f, 3.0f}, {0.4f, 1.0f, 11.0f, 9.0f, 3.0f}};
    void *h = oracle_create(precision, words, &events[0][0], 2);
    oracle_advance(h, 3);
    const size_t n = (size_t)g[0] * g[1] * g[2];
    const size_t r = precision ? 8 : 4;
    std::vector<unsigned char> field(n * 2 * r), fluid(n * 4 * r);
    oracle_get_state(h, field.data(), fluid.data());
    oracle_destroy(h);
    // the state is finite and the bubbles are there
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
        const double phi = precision ? ((const double *)field.data())[2 * i] : ((const float *)field.data())[2 * i];
        CHECK(std::isfinite(phi));
        any = any || phi != 0.0;
    }
    CHECK(any);
}

int main() {
    grids();
    schedules();
    specs();
    oracle_run(0);
    oracle_run(1);
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitized host run ok\n");
    return 0;
}

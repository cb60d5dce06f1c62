// physics.cpp -- host-side model derivations, validation, the seeded
// nucleation schedule and the parameter blocks.
//
// Everything here is f32 arithmetic in the reference's operation order and is
// compiled with -ffp-contract=off, so results are bit-identical to the Rust
// host (which does not contract and uses the platform libm for exp/ln).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "bubblesim.h"
#include "internal.h"

namespace bs {

// ---------------------------------------------------------------------------
//  Random numbers: rand_core 0.9 seed_from_u64 -> rand_pcg 0.9 Mcg128Xsl64
//  (Pcg64Mcg), rand 0.9 StandardUniform for f32.  physics.rs:256, 278-280.
// ---------------------------------------------------------------------------
typedef unsigned __int128 u128;
static const u128 kMcgMul = (((u128)0x2360ED051FC65DA4ULL) << 64) | (u128)0x4385DF649FCCF645ULL;

static inline uint64_t rotr64(uint64_t x, unsigned r) { return r ? (x >> r) | (x << (64 - r)) : x; }
static inline uint32_t rotr32(uint32_t x, unsigned r) { return r ? (x >> r) | (x << (32 - r)) : x; }

struct Mcg128 {
    u128 state;
    // Mcg128Xsl64::new: force the low bit
    explicit Mcg128(u128 s) : state(s | 1) {}
    // next_u64: multiply, then the XSL-RR output function
    uint64_t next_u64() {
        state = state * kMcgMul;
        const unsigned rot = (unsigned)(state >> 122);
        const uint64_t xsl = (uint64_t)(state >> 64) ^ (uint64_t)state;
        return rotr64(xsl, rot);
    }
    // next_u32: low half of next_u64
    uint32_t next_u32() { return (uint32_t)next_u64(); }
    // rand's StandardUniform f32: 24 most significant bits of a u32, times 2^-24
    float next_f32() {
        const uint32_t v = next_u32() >> 8;
        return (1.0f / 16777216.0f) * (float)v;
    }
};

// SeedableRng::seed_from_u64: four PCG32 (XSH-RR) outputs, advancing the
// 64-bit LCG before each, laid little-endian into the 16-byte seed; from_seed
// reads the seed as a little-endian u128.
static u128 seed_from_u64(uint64_t s) {
    uint8_t seed[16];
    for (int c = 0; c < 4; ++c) {
        s = s * 6364136223846793005ULL + 11634580027462260723ULL;
        const uint32_t xs = (uint32_t)(((s >> 18) ^ s) >> 27);
        const uint32_t x = rotr32(xs, (unsigned)(s >> 59));
        for (int b = 0; b < 4; ++b) seed[4 * c + b] = (uint8_t)(x >> (8 * b));
    }
    uint64_t lo = 0, hi = 0;
    for (int b = 0; b < 8; ++b) {
        lo |= (uint64_t)seed[b] << (8 * b);
        hi |= (uint64_t)seed[8 + b] << (8 * b);
    }
    return ((u128)hi << 64) | lo;
}

// powi by repeated squaring, as LLVM expands a constant-exponent powi
static float powi_f(float x, int n) {
    float res = 1.0f;
    bool have = false;
    float cur = x;
    while (n > 0) {
        if (n & 1) {
            res = have ? res * cur : cur;
            have = true;
        }
        n >>= 1;
        if (n) cur = cur * cur;
    }
    return have ? res : 1.0f;
}

// Rust f32::max: the larger, ignoring a NaN operand
static float rust_max(float a, float b) {
    if (std::isnan(a)) return b;
    if (std::isnan(b)) return a;
    return a < b ? b : a;
}

// ---------------------------------------------------------------------------
//  Model (physics.rs:44-218)
// ---------------------------------------------------------------------------
void model_default(bs_model *m) {
    m->lambda = 0.5f;
    m->eps_ratio = 0.3f;
    m->alpha = 0.1f;
    m->eta = 0.3f;
    m->phi_b = 1.0f;
    m->dx = 1.0f;
    m->cfl = 0.2f;
    m->seed_kind = BS_SEED_CRITICAL;
    m->seed_value = 1.15f;  // SeedSize::DEFAULT_FACTOR physics.rs:41
}

float model_eps(const bs_model &m) { return m.eps_ratio * m.lambda * powi_f(m.phi_b, 4) / 12.0f; }
float model_m2(const bs_model &m) {
    return 0.5f * m.lambda * powi_f(m.phi_b, 2) - 6.0f * model_eps(m) / powi_f(m.phi_b, 2);
}
float model_delta(const bs_model &m) {
    return 1.5f * m.lambda * m.phi_b - 6.0f * model_eps(m) / powi_f(m.phi_b, 3);
}
float model_wall_width(const bs_model &m) {
    const float sqrt2 = 1.41421356237309504880f;  // std::f32::consts::SQRT_2
    return 2.0f * sqrt2 / (m.phi_b * std::sqrt(m.lambda));
}
float model_surface_tension(const bs_model &m) { return std::sqrt(m.lambda / 2.0f) * powi_f(m.phi_b, 3) / 6.0f; }
float model_critical_radius(const bs_model &m) { return 2.0f * model_surface_tension(m) / model_eps(m); }
float model_e_ref(const bs_model &m) { return model_eps(m) / m.alpha; }
float model_t_ref(const bs_model &) { return 1.0f; }
float model_a_rad(const bs_model &m) { return model_e_ref(m) / powi_f(model_t_ref(m), 4); }
float model_dt(const bs_model &m) { return m.cfl * m.dx; }
float model_seed_radius(const bs_model &m) {
    const float r = m.seed_kind == BS_SEED_CELLS ? m.seed_value : m.seed_value * model_critical_radius(m);
    return rust_max(r, 0.25f * m.dx);
}

int model_validate(const bs_model &m, std::string &err) {
    if (!(m.eps_ratio >= 0.0f && m.eps_ratio < 1.0f) || m.eps_ratio <= 0.0f) {
        err = "eps_ratio must lie in (0, 1)";
        return BS_ERR_INVALID;
    }
    if (m.lambda <= 0.0f) { err = "lambda must be positive"; return BS_ERR_INVALID; }
    if (m.alpha <= 0.0f) { err = "alpha must be positive"; return BS_ERR_INVALID; }
    if (m.eta < 0.0f) { err = "eta must be non-negative"; return BS_ERR_INVALID; }
    if (m.cfl <= 0.0f || m.cfl > 0.5f) {
        err = "cfl must lie in (0, 0.5]; see docs/NUMERICS.md";
        return BS_ERR_INVALID;
    }
    if (m.seed_kind == BS_SEED_CRITICAL && m.seed_value <= 0.0f) {
        err = "seed factor must be positive";
        return BS_ERR_INVALID;
    }
    if (m.seed_kind == BS_SEED_CELLS && m.seed_value <= 0.0f) {
        err = "seed radius must be positive";
        return BS_ERR_INVALID;
    }
    if (m.seed_kind != BS_SEED_CRITICAL && m.seed_kind != BS_SEED_CELLS) {
        err = "unknown seed size kind";
        return BS_ERR_INVALID;
    }
    return BS_OK;
}

void model_derive(const bs_model &m, bs_derived *d) {
    d->eps = model_eps(m);
    d->m2 = model_m2(m);
    d->delta = model_delta(m);
    d->wall_width = model_wall_width(m);
    d->surface_tension = model_surface_tension(m);
    d->critical_radius = model_critical_radius(m);
    d->e_ref = model_e_ref(m);
    d->t_ref = model_t_ref(m);
    d->a_rad = model_a_rad(m);
    d->dt = model_dt(m);
    d->seed_radius = model_seed_radius(m);
    d->critical_to_wall_ratio = 1.0f / m.eps_ratio;
    {
        const float x = model_eps(m) / (rust_max(m.eta, 1e-6f) * model_surface_tension(m));
        d->estimated_wall_velocity = x / std::sqrt(1.0f + x * x);
    }
    d->sound_speed = std::sqrt(1.0f / 3.0f);
}

// ---------------------------------------------------------------------------
//  Schedule (physics.rs:248-308)
// ---------------------------------------------------------------------------
static float periodic_distance(const float a[3], const float b[3], const float n[3]) {
    float sum = 0.0f;
    for (int i = 0; i < 3; ++i) {
        float d = std::fabs(a[i] - b[i]);
        if (d > 0.5f * n[i]) d = n[i] - d;
        sum += d * d;
    }
    return std::sqrt(sum);
}

// f32 time after k steps of `time += dt` (sim.rs:913-915)
float accumulated_time(uint64_t k, float dt) {
    float t = 0.0f;
    for (uint64_t s = 0; s < k; ++s) t += dt;
    return t;
}

int make_schedule(int32_t mode, uint32_t count, float radius, const uint32_t grid[3], float duration, uint64_t seed,
                  uint32_t uniform_max_step, uint64_t time_seed, float dt, std::vector<bs_event> &out,
                  std::string &err) {
    if (mode != BS_NUC_SIMULTANEOUS && mode != BS_NUC_EXPONENTIAL && mode != BS_NUC_UNIFORM_STEPS) {
        err = "unknown nucleation mode";
        return BS_ERR_INVALID;
    }
    Mcg128 rng(seed_from_u64(seed));
    const float n[3] = {(float)grid[0], (float)grid[1], (float)grid[2]};
    const float min_sep = 2.5f * radius;
    out.clear();
    out.reserve(count);
    // separate generator for the uniform-steps extension, so positions are
    // the reference's stream untouched
    Mcg128 trng(seed_from_u64(time_seed));
    for (uint32_t k = 0; k < count; ++k) {
        float time = 0.0f;
        if (mode == BS_NUC_EXPONENTIAL) {
            const float growth = 4.0f;
            const float frac = ((float)k + 0.5f) / (float)count;
            time = duration / growth * std::log(1.0f + frac * (std::exp(growth) - 1.0f));
        } else if (mode == BS_NUC_UNIFORM_STEPS) {
            const uint64_t step = trng.next_u64() % ((uint64_t)uniform_max_step + 1u);
            time = accumulated_time(step, dt);
        }
        float pos[3] = {0.0f, 0.0f, 0.0f};
        for (int attempt = 0; attempt < 64; ++attempt) {
            float cand[3];
            cand[0] = rng.next_f32() * n[0];
            cand[1] = rng.next_f32() * n[1];
            cand[2] = rng.next_f32() * n[2];
            bool ok = true;
            for (const bs_event &e : out) {
                if (!(periodic_distance(cand, e.pos, n) > min_sep)) {
                    ok = false;
                    break;
                }
            }
            pos[0] = cand[0]; pos[1] = cand[1]; pos[2] = cand[2];
            if (ok) break;
        }
        bs_event ev;
        ev.time = time;
        ev.pos[0] = pos[0]; ev.pos[1] = pos[1]; ev.pos[2] = pos[2];
        ev.radius = radius;
        out.push_back(ev);
    }
    std::stable_sort(out.begin(), out.end(), [](const bs_event &a, const bs_event &b) { return a.time < b.time; });
    return BS_OK;
}

// The schedule Simulation::new builds for a spec (sim.rs:551-558), with the
// single-centre-bubble extension.
int spec_schedule(const bs_spec &s, std::vector<bs_event> &out, std::string &err) {
    const float radius = model_seed_radius(s.model);
    if (s.bubble_at_centre) {
        bs_event e;
        e.time = 0.0f;
        e.pos[0] = (float)(s.grid[0] / 2);
        e.pos[1] = (float)(s.grid[1] / 2);
        e.pos[2] = (float)(s.grid[2] / 2);
        e.radius = radius;
        out.assign(1, e);
        return BS_OK;
    }
    return make_schedule(s.nucleation, s.bubbles, radius, s.grid, s.nucleation_duration, s.seed, s.uniform_max_step,
                         s.time_seed, model_dt(s.model), out, err);
}

uint64_t event_step(float t, float dt) {
    // first s with accumulated time >= t; UINT64_MAX if the f32 clock stalls first
    float acc = 0.0f;
    uint64_t s = 0;
    while (!(t <= acc)) {
        const float nxt = acc + dt;
        if (!(nxt > acc)) return UINT64_MAX;
        acc = nxt;
        ++s;
    }
    return s;
}

// gpu_params (sim.rs:613-644)
void sim_params_fill(const bs_model &m, const uint32_t grid[3], float time, bs_sim_params *p) {
    std::memset(p, 0, sizeof(*p));
    const float e_ref = model_e_ref(m);
    p->nx = grid[0];
    p->ny = grid[1];
    p->nz = grid[2];
    p->dx = m.dx;
    p->inv_dx = 1.0f / m.dx;
    p->dt = model_dt(m);
    p->eta = m.eta;
    p->lambda = m.lambda;
    p->eps = model_eps(m);
    p->phi_b = m.phi_b;
    p->m2 = model_m2(m);
    p->delta = model_delta(m);
    p->a_rad = model_a_rad(m);
    p->wall_width = model_wall_width(m);
    p->p_floor = 1e-7f * e_ref;
    p->e_floor = 1e-7f * e_ref;
    p->z_max = 0.9999f;
    p->e_ref = e_ref;
    p->t_ref = model_t_ref(m);
    p->time = time;
    p->vis_gain = 1.0f;
}

// Config::validate (config.rs:151-178)
int spec_validate(const bs_spec &s, std::string &err) {
    int rc = model_validate(s.model, err);
    if (rc) return rc;
    if (s.bubbles == 0 && !s.bubble_at_centre) {
        err = "--bubbles must be at least 1";
        return BS_ERR_INVALID;
    }
    if (s.nucleation_duration <= 0.0f) {
        err = "--nucleation-duration must be positive";
        return BS_ERR_INVALID;
    }
    const float lw = model_wall_width(s.model);
    if (lw < 2.0f) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "lambda = %g gives a wall only %.2f cells thick, which the lattice cannot resolve; use "
                      "lambda <= %.2f",
                      (double)s.model.lambda, (double)lw, 2.0);
        err = buf;
        return BS_ERR_INVALID;
    }
    for (int a = 0; a < 3; ++a) {
        if (s.grid[a] < 8) {
            char buf[160];
            std::snprintf(buf, sizeof buf, "grid axis %d is %u; the 5-point stencil needs at least 8 cells", a,
                          s.grid[a]);
            err = buf;
            return BS_ERR_INVALID;
        }
    }
    if (s.precision != BS_PRECISION_F32 && s.precision != BS_PRECISION_F64) {
        err = "precision must be BS_PRECISION_F32 or BS_PRECISION_F64";
        return BS_ERR_INVALID;
    }
    {
        // the stage kernels address one x-y plane of fluid states (16 B per
        // cell in f32, 32 B in f64) with 32-bit byte offsets
        const uint64_t plane_cells = (uint64_t)s.grid[0] * s.grid[1];
        const uint64_t cap = s.precision == BS_PRECISION_F64 ? (1ull << 27) : (1ull << 28);
        if (plane_cells >= cap) {
            char buf[200];
            std::snprintf(buf, sizeof buf, "an x-y plane of %llu cells is too large; nx * ny must stay below %llu",
                          (unsigned long long)plane_cells, (unsigned long long)cap);
            err = buf;
            return BS_ERR_INVALID;
        }
    }
    if (s.math != BS_MATH_FAST && s.math != BS_MATH_EXACT) {
        err = "math must be BS_MATH_FAST or BS_MATH_EXACT";
        return BS_ERR_INVALID;
    }
    if (s.kernel != BS_KERNEL_FUSED && s.kernel != BS_KERNEL_TWO_PASS) {
        err = "kernel must be BS_KERNEL_FUSED or BS_KERNEL_TWO_PASS";
        return BS_ERR_INVALID;
    }
    if (s.nucleation != BS_NUC_SIMULTANEOUS && s.nucleation != BS_NUC_EXPONENTIAL &&
        s.nucleation != BS_NUC_UNIFORM_STEPS) {
        err = "unknown nucleation mode";
        return BS_ERR_INVALID;
    }
    if (s.eos != BS_EOS_BAG && s.eos != BS_EOS_GAMMA_LAW) {
        err = "eos must be BS_EOS_BAG or BS_EOS_GAMMA_LAW";
        return BS_ERR_INVALID;
    }
    if (s.eos == BS_EOS_GAMMA_LAW && !(s.eos_gamma > 1.0f && s.eos_gamma < 2.0f)) {
        char buf[160];
        std::snprintf(buf, sizeof buf,
                      "eos_gamma = %g is outside (1, 2): the sound speed squared, eos_gamma - 1, must lie in (0, 1)",
                      (double)s.eos_gamma);
        err = buf;
        return BS_ERR_INVALID;
    }
    if (!(s.dof_ratio > 0.0f) || !std::isfinite(s.dof_ratio)) {
        err = "dof_ratio must be positive";
        return BS_ERR_INVALID;
    }
    return BS_OK;
}

void spec_default(bs_spec *s) {
    std::memset(s, 0, sizeof(*s));
    model_default(&s->model);
    s->grid[0] = s->grid[1] = s->grid[2] = 192;
    s->nucleation = BS_NUC_EXPONENTIAL;
    s->bubbles = 12;
    s->nucleation_duration = 60.0f;
    s->seed = 1;
    s->precision = BS_PRECISION_F32;
    s->math = BS_MATH_FAST;
    s->kernel = BS_KERNEL_FUSED;
    s->bubble_at_centre = 0;
    s->uniform_max_step = 200;
    s->time_seed = 0x5eedULL;
    s->eos = BS_EOS_BAG;
    s->eos_gamma = 4.0f / 3.0f;
    s->dof_ratio = 1.0f;
}

int parse_grid(const char *s, uint32_t g[3], std::string &err) {
    std::vector<uint32_t> nums;
    std::string str(s ? s : "");
    size_t start = 0;
    while (true) {
        size_t comma = str.find(',', start);
        std::string part = str.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
        // trim
        size_t a = part.find_first_not_of(" \t"), b = part.find_last_not_of(" \t");
        part = a == std::string::npos ? std::string() : part.substr(a, b - a + 1);
        if (part.empty() || part.find_first_not_of("0123456789") != std::string::npos || part.size() > 10) {
            err = "`" + str + "` is not a lattice size: invalid digit found in string";
            return BS_ERR_INVALID;
        }
        unsigned long long v = std::stoull(part);
        if (v > 0xffffffffULL) {
            err = "`" + str + "` is not a lattice size: number too large to fit in target type";
            return BS_ERR_INVALID;
        }
        nums.push_back((uint32_t)v);
        if (comma == std::string::npos) break;
        start = comma + 1;
    }
    if (nums.size() == 1) {
        g[0] = g[1] = g[2] = nums[0];
        return BS_OK;
    }
    if (nums.size() == 3) {
        g[0] = nums[0]; g[1] = nums[1]; g[2] = nums[2];
        return BS_OK;
    }
    err = "expected 1 or 3 comma-separated sizes, got " + std::to_string(nums.size());
    return BS_ERR_INVALID;
}

uint64_t bytes_per_cell(int32_t precision, int32_t kernel) {
    const uint64_t real = precision == BS_PRECISION_F64 ? 8 : 4;
    uint64_t per = 3 * (2 + 4) * real;
    if (kernel == BS_KERNEL_TWO_PASS) per += 4 * real;
    return per;
}

}  // namespace bs

// ---------------------------------------------------------------------------
//  extern "C" surface for the host physics
// ---------------------------------------------------------------------------
using namespace bs;

extern "C" {

void bs_model_default(bs_model *out) {
    if (out) model_default(out);
}

int bs_model_validate(const bs_model *m) {
    if (!m) return set_error(BS_ERR_INVALID, "null model");
    std::string err;
    int rc = model_validate(*m, err);
    if (rc) return set_error(rc, err);
    return BS_OK;
}

int bs_model_derive(const bs_model *m, bs_derived *out) {
    if (!m || !out) return set_error(BS_ERR_INVALID, "null argument");
    model_derive(*m, out);
    return BS_OK;
}

int bs_make_schedule(int32_t mode, uint32_t count, float radius, const uint32_t grid[3], float duration,
                     uint64_t seed, uint32_t uniform_max_step, uint64_t time_seed, float dt, bs_event *out) {
    if (!grid || (!out && count)) return set_error(BS_ERR_INVALID, "null argument");
    std::vector<bs_event> ev;
    std::string err;
    int rc = make_schedule(mode, count, radius, grid, duration, seed, uniform_max_step, time_seed, dt, ev, err);
    if (rc) return set_error(rc, err);
    if (count) std::memcpy(out, ev.data(), sizeof(bs_event) * count);
    return BS_OK;
}

void bs_pcg64mcg_u64(uint64_t seed, uint32_t n, uint64_t *out) {
    Mcg128 r(seed_from_u64(seed));
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_u64();
}

void bs_pcg64mcg_f32(uint64_t seed, uint32_t n, float *out) {
    Mcg128 r(seed_from_u64(seed));
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_f32();
}

void bs_mcg128_from_state(uint64_t lo, uint64_t hi, uint32_t n, uint64_t *out) {
    Mcg128 r(((u128)hi << 64) | lo);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_u64();
}

int bs_sim_params_fill(const bs_model *m, const uint32_t grid[3], float time, bs_sim_params *out) {
    if (!m || !grid || !out) return set_error(BS_ERR_INVALID, "null argument");
    sim_params_fill(*m, grid, time, out);
    return BS_OK;
}

void bs_rk3_coefficients(float out[6]) {
    out[0] = 0.0f; out[1] = 1.0f;
    out[2] = 0.75f; out[3] = 0.25f;
    out[4] = 1.0f / 3.0f; out[5] = 2.0f / 3.0f;
}

void bs_stage_slots(uint32_t cur, uint32_t out[6]) {
    const uint32_t a = cur % 3, b = (cur + 1) % 3, c = (cur + 2) % 3;
    out[0] = a; out[1] = b;
    out[2] = b; out[3] = c;
    out[4] = c; out[5] = b;
}

int bs_event_steps(const bs_event *events, uint32_t n, float dt, uint64_t *steps_out) {
    if ((!events || !steps_out) && n) return set_error(BS_ERR_INVALID, "null argument");
    for (uint32_t i = 0; i < n; ++i) steps_out[i] = event_step(events[i].time, dt);
    return BS_OK;
}

int bs_spec_validate(const bs_spec *spec) {
    if (!spec) return set_error(BS_ERR_INVALID, "null spec");
    std::string err;
    int rc = spec_validate(*spec, err);
    if (rc) return set_error(rc, err);
    return BS_OK;
}

void bs_spec_default(bs_spec *out) {
    if (out) spec_default(out);
}

int bs_parse_grid(const char *s, uint32_t grid_out[3]) {
    std::string err;
    int rc = parse_grid(s, grid_out, err);
    if (rc) return set_error(rc, err);
    return BS_OK;
}

uint64_t bs_bytes_per_cell(int32_t precision, int32_t kernel) { return bytes_per_cell(precision, kernel); }

uint64_t bs_lattice_bytes(const uint32_t grid[3], int32_t precision, int32_t kernel) {
    return (uint64_t)grid[0] * grid[1] * grid[2] * bytes_per_cell(precision, kernel);
}

int bs_max_cubic_size(uint64_t budget_bytes, int32_t precision, int32_t kernel, uint32_t *side_out,
                      int32_t *limited_by_memory_out) {
    if (!side_out) return set_error(BS_ERR_INVALID, "null argument");
    const uint64_t per = bytes_per_cell(precision, kernel);
    const uint32_t by_budget = (uint32_t)std::floor(std::cbrt((double)(budget_bytes / per)));
    uint32_t side = std::min<uint32_t>(by_budget, 2048u);
    side = std::max<uint32_t>(side, 8u);
    *side_out = side;
    if (limited_by_memory_out) *limited_by_memory_out = by_budget < 2048u ? 1 : 0;
    return BS_OK;
}

uint64_t bs_lattice_budget(uint64_t available, uint64_t current_lattice) {
    uint64_t sum = available + current_lattice;
    if (sum < available) sum = UINT64_MAX;  // saturating_add
    return (uint64_t)((double)sum * 0.85);
}

}  // extern "C"

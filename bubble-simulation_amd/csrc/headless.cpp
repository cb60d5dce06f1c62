// headless.cpp -- `bubble-sim`, the command-line front end.
//
// Keeps the reference's flags (config.rs:13-115) and its headless report
// (headless.rs:23-116) byte for byte where the numbers agree, on top of the C
// ABI.  Extensions: --precision, --math, --kernel, --bubble-at-centre,
// --nucleation uniform-steps with --uniform-max-step/--time-seed, --dump,
// --check-f64 (the same run in f64 ExactMath beside it: field error and
// energy drift of the result against it).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bubblesim.h"

namespace {

// Rust's `{:.Ne}` (LowerExp): mantissa digits as printf, exponent without
// '+' and without leading zeros ("1.229e-2", "0.000e0").
std::string rust_exp(double v, int prec) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", prec, v);
    std::string s(buf);
    const size_t e = s.find('e');
    if (e == std::string::npos) return s;
    const int ex = std::atoi(s.c_str() + e + 1);
    return s.substr(0, e) + "e" + std::to_string(ex);
}

std::string rpad(const std::string &s, size_t w) { return s.size() >= w ? s : std::string(w - s.size(), ' ') + s; }

void usage(FILE *f) {
    std::fprintf(f,
                 "Live GPU lattice simulation of bubble nucleation and growth in a first-order cosmological phase "
                 "transition.\n\n"
                 "Usage: bubble-sim [OPTIONS]\n\n"
                 "Options:\n"
                 "      --grid <GRID>                  Lattice size: N or NX,NY,NZ [default: 192]\n"
                 "      --lambda <LAMBDA>              Quartic coupling [default: 0.5]\n"
                 "      --eps-ratio <EPS_RATIO>        Vacuum energy split [default: 0.3]\n"
                 "      --alpha <ALPHA>                Transition strength [default: 0.1]\n"
                 "      --eta <ETA>                    Wall friction [default: 0.3]\n"
                 "      --cfl <CFL>                    Courant number [default: 0.2]\n"
                 "      --seed-factor <SEED_FACTOR>    Seed radius in units of R_c [default: 1.15]\n"
                 "      --seed-radius <CELLS>          Seed radius in cells (overrides --seed-factor)\n"
                 "      --bubbles <BUBBLES>            Number of bubbles [default: 12]\n"
                 "      --nucleation <NUCLEATION>      simultaneous | exponential | uniform-steps [default: exponential]\n"
                 "      --nucleation-duration <T>      Time over which exponential nucleation completes [default: 60]\n"
                 "      --seed <SEED>                  Seed for bubble placement [default: 1]\n"
                 "      --steps-per-frame <N>          (interactive only) [default: 2]\n"
                 "      --field <FIELD>                Field the volume renderer shows: phi | speed | temperature | kinetic [default: kinetic]\n"
                 "      --no-vsync                     (interactive only)\n"
                 "      --headless <STEPS>             Run this many steps without a window, printing diagnostics\n"
                 "      --report-every <N>             Report interval in headless mode [default: steps/20]\n"
                 "      --screenshot <PATH>            Render one frame to this PNG and exit (after --headless N steps)\n"
                 "      --screenshot-size <WxH>        Resolution of --screenshot [default: 1280x960]\n"
                 "  extensions:\n"
                 "      --precision <f32|f64>          Lattice precision [default: f32]\n"
                 "      --math <fast|exact>            Arithmetic policy of the stage kernels [default: fast]\n"
                 "      --kernel <fused|two-pass>      Stage implementation [default: fused]\n"
                 "      --bubble-at-centre             One bubble at the box centre at t = 0\n"
                 "      --uniform-max-step <N>         uniform-steps nucleation: steps in [0, N] [default: 200]\n"
                 "      --time-seed <SEED>             uniform-steps nucleation time generator seed\n"
                 "      --eos <bag|gamma-law>          Equation of state [default: bag]\n"
                 "      --eos-gamma <G>                gamma-law adiabatic index, 1 < G < 2 [default: 4/3]\n"
                 "      --dof-ratio <R>                a_broken / a_symmetric of the temperature [default: 1]\n"
                 "      --dump <PATH>                  Write the final state (field then fluid, raw) to PATH\n"
                 "      --check-f64                    Step an f64 ExactMath copy of the run beside it; report the\n"
                 "                                     relative L2/Linf error of the result and both energy drifts\n"
                 "  -h, --help                         Print help\n"
                 "  -V, --version                      Print version\n");
}

bool parse_f32(const std::string &v, float &out) {
    char *end = nullptr;
    out = std::strtof(v.c_str(), &end);
    return end && *end == '\0' && !v.empty();
}
bool parse_u64(const std::string &v, uint64_t &out) {
    if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos) return false;
    out = std::strtoull(v.c_str(), nullptr, 10);
    return true;
}

// parse_size (config.rs:181-192): WIDTHxHEIGHT, either case of x, both
// positive
bool parse_size(const std::string &s, uint32_t &w, uint32_t &h, std::string &err) {
    const size_t k = s.find_first_of("xX");
    if (k == std::string::npos) {
        err = "`" + s + "` is not a size; expected WIDTHxHEIGHT";
        return false;
    }
    auto trim = [](std::string v) {
        const size_t a = v.find_first_not_of(" \t"), b = v.find_last_not_of(" \t");
        return a == std::string::npos ? std::string() : v.substr(a, b - a + 1);
    };
    const std::string parts[2] = {trim(s.substr(0, k)), trim(s.substr(k + 1))};
    const char *what[2] = {"width", "height"};
    uint32_t out[2];
    for (int i = 0; i < 2; ++i) {
        uint64_t u = 0;
        if (parts[i].empty()) {
            err = std::string(what[i]) + ": cannot parse integer from empty string";
            return false;
        }
        if (!parse_u64(parts[i], u) || u > 0xffffffffULL) {
            err = std::string(what[i]) + ": invalid digit found in string";
            return false;
        }
        if (u == 0) {
            err = std::string(what[i]) + " must be positive";
            return false;
        }
        out[i] = (uint32_t)u;
    }
    w = out[0];
    h = out[1];
    return true;
}

// Relative errors of a state block (k components per cell) against the f64
// run: sqrt(sum d^2 / sum r^2) and max|d| / max|r| over every component,
// accumulated in double.
struct BlockError {
    double l2 = 0.0, linf = 0.0;
};
BlockError block_error(const float *a, const double *r, size_t count) {
    double se = 0.0, sr = 0.0, me = 0.0, mr = 0.0;
    for (size_t i = 0; i < count; ++i) {
        const double d = (double)a[i] - r[i];
        se += d * d;
        sr += r[i] * r[i];
        me = std::fmax(me, std::fabs(d));
        mr = std::fmax(mr, std::fabs(r[i]));
    }
    BlockError e;
    e.l2 = sr > 0.0 ? std::sqrt(se / sr) : 0.0;
    e.linf = mr > 0.0 ? me / mr : 0.0;
    return e;
}

// --check-f64: the same run in f64 ExactMath, after the timed one, with the
// same report steps; prints the final state's error and the energy drift of
// both runs from the first report after the last nucleation (headless.rs:54-78)
int check_against_f64(const bs_spec &spec, bs_sim *sim, uint64_t steps, uint64_t every,
                      const std::vector<double> &energies, bool have_base, size_t base_idx, uint64_t base_step) {
    bs_spec s64 = spec;
    s64.precision = BS_PRECISION_F64;
    s64.math = BS_MATH_EXACT;
    bs_sim *ref = nullptr;
    if (bs_sim_create(&s64, &ref) != BS_OK) {
        std::fprintf(stderr, "Error: the f64 copy of the run: %s\n", bs_last_error());
        return 1;
    }
    std::vector<double> e64;
    uint64_t done = 0;
    int rc = 0;
    while (done < steps && !rc) {
        const uint64_t chunk = every < steps - done ? every : steps - done;
        bs_diagnostics diag;
        if (bs_sim_advance(ref, chunk) != BS_OK || bs_sim_diagnostics(ref, &diag) != BS_OK) {
            std::fprintf(stderr, "Error: the f64 copy of the run: %s\n", bs_last_error());
            rc = 1;
        } else {
            done += chunk;
            e64.push_back(diag.mean_energy);
        }
    }
    if (!rc) {
        const uint64_t n = bs_sim_cell_count(sim);
        std::vector<float> a(n * 6);
        std::vector<double> r(n * 6);
        if (bs_sim_download(sim, a.data(), a.data() + n * 2) != BS_OK ||
            bs_sim_download(ref, r.data(), r.data() + n * 2) != BS_OK) {
            std::fprintf(stderr, "Error: %s\n", bs_last_error());
            rc = 1;
        } else {
            const BlockError ef = block_error(a.data(), r.data(), n * 2);
            const BlockError eu = block_error(a.data() + n * 2, r.data() + n * 2, n * 4);
            std::printf("\nf64 check after %llu steps (f64 ExactMath run of the same spec):\n", (unsigned long long)steps);
            std::printf("  field (phi, pi): rel L2 %.3e, rel Linf %.3e\n", ef.l2, ef.linf);
            std::printf("  fluid (E, Z):    rel L2 %.3e, rel Linf %.3e\n", eu.l2, eu.linf);
            // the base report is the same in both runs: they share the schedule
            const size_t b = base_idx;
            if (have_base && b < energies.size() && energies.size() == e64.size()) {
                double worst = 0.0;
                for (size_t i = b; i < energies.size(); ++i)
                    worst = std::fmax(worst, std::fabs((energies[i] / energies[b] - 1.0) - (e64[i] / e64[b] - 1.0)));
                std::printf("  energy drift from step %llu: f32 %+.3e, f64 %+.3e (largest difference over the reports %.3e)\n",
                            (unsigned long long)base_step, energies.back() / energies[b] - 1.0,
                            e64.back() / e64[b] - 1.0, worst);
            } else {
                std::printf("  energy drift: no report after the last nucleation\n");
            }
        }
    }
    bs_sim_destroy(ref);
    return rc;
}

int fail_usage(const std::string &msg) {
    std::fprintf(stderr, "error: %s\n\nFor more information, try '--help'.\n", msg.c_str());
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    bs_spec spec;
    bs_spec_default(&spec);
    float seed_factor = 1.15f;
    bool have_seed_radius = false;
    float seed_radius = 0.0f;
    bool have_headless = false, have_screenshot = false;
    uint64_t headless = 0, report_every = 0;
    bool have_report = false;
    uint64_t steps_per_frame = 2;
    std::string dump_path;
    std::string screenshot_path;
    uint32_t shot_w = 1280, shot_h = 960;  // config.rs:112-114
    int32_t field_mode = 3;                // FieldMode::Kinetic (render.rs:26-28)
    bool check_f64 = false;

    std::vector<std::string> args(argv + 1, argv + argc);
    for (size_t i = 0; i < args.size(); ++i) {
        std::string a = args[i], val;
        bool has_inline = false;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            val = a.substr(eq + 1);
            a = a.substr(0, eq);
            has_inline = true;
        }
        auto need = [&](std::string &v) -> bool {
            if (has_inline) {
                v = val;
                return true;
            }
            if (i + 1 >= args.size()) return false;
            v = args[++i];
            return true;
        };
        std::string v;
        if (a == "-h" || a == "--help") {
            usage(stdout);
            return 0;
        } else if (a == "-V" || a == "--version") {
            std::printf("bubble-sim %s\n", bs_version());
            return 0;
        } else if (a == "--no-vsync") {
        } else if (a == "--bubble-at-centre") {
            spec.bubble_at_centre = 1;
            spec.bubbles = 1;
        } else if (a == "--check-f64") {
            check_f64 = true;
        } else if (!need(v)) {
            if (a.rfind("-", 0) == 0) return fail_usage("a value is required for '" + a + "' but none was supplied");
            return fail_usage("unexpected argument '" + a + "' found");
        } else if (a == "--grid") {
            if (bs_parse_grid(v.c_str(), spec.grid) != BS_OK)
                return fail_usage("invalid value '" + v + "' for '--grid <GRID>': " + bs_last_error());
        } else if (a == "--lambda" || a == "--eps-ratio" || a == "--alpha" || a == "--eta" || a == "--cfl" ||
                   a == "--seed-factor" || a == "--seed-radius" || a == "--nucleation-duration" ||
                   a == "--eos-gamma" || a == "--dof-ratio") {
            float f;
            if (!parse_f32(v, f)) return fail_usage("invalid value '" + v + "' for '" + a + "': invalid float literal");
            if (a == "--lambda") spec.model.lambda = f;
            else if (a == "--eps-ratio") spec.model.eps_ratio = f;
            else if (a == "--alpha") spec.model.alpha = f;
            else if (a == "--eta") spec.model.eta = f;
            else if (a == "--cfl") spec.model.cfl = f;
            else if (a == "--seed-factor") seed_factor = f;
            else if (a == "--seed-radius") { have_seed_radius = true; seed_radius = f; }
            else if (a == "--eos-gamma") spec.eos_gamma = f;
            else if (a == "--dof-ratio") spec.dof_ratio = f;
            else spec.nucleation_duration = f;
        } else if (a == "--bubbles" || a == "--seed" || a == "--steps-per-frame" || a == "--headless" ||
                   a == "--report-every" || a == "--uniform-max-step" || a == "--time-seed") {
            uint64_t u;
            if (!parse_u64(v, u)) return fail_usage("invalid value '" + v + "' for '" + a + "': invalid digit found in string");
            if (a == "--bubbles") spec.bubbles = (uint32_t)u;
            else if (a == "--seed") spec.seed = u;
            else if (a == "--steps-per-frame") steps_per_frame = u;
            else if (a == "--headless") { have_headless = true; headless = u; }
            else if (a == "--report-every") { have_report = true; report_every = u; }
            else if (a == "--uniform-max-step") spec.uniform_max_step = (uint32_t)u;
            else spec.time_seed = u;
        } else if (a == "--nucleation") {
            if (v == "simultaneous") spec.nucleation = BS_NUC_SIMULTANEOUS;
            else if (v == "exponential") spec.nucleation = BS_NUC_EXPONENTIAL;
            else if (v == "uniform-steps") spec.nucleation = BS_NUC_UNIFORM_STEPS;
            else return fail_usage("invalid value '" + v + "' for '--nucleation <NUCLEATION>'");
        } else if (a == "--field") {
            if (v == "phi") field_mode = 0;
            else if (v == "speed") field_mode = 1;
            else if (v == "temperature") field_mode = 2;
            else if (v == "kinetic") field_mode = 3;
            else return fail_usage("invalid value '" + v + "' for '--field <FIELD>'");
        } else if (a == "--screenshot") {
            have_screenshot = true;
            screenshot_path = v;
        } else if (a == "--screenshot-size") {
            std::string err;
            if (!parse_size(v, shot_w, shot_h, err))
                return fail_usage("invalid value '" + v + "' for '--screenshot-size <SCREENSHOT_SIZE>': " + err);
        } else if (a == "--precision") {
            if (v == "f32") spec.precision = BS_PRECISION_F32;
            else if (v == "f64") spec.precision = BS_PRECISION_F64;
            else return fail_usage("invalid value '" + v + "' for '--precision'");
        } else if (a == "--math") {
            if (v == "fast") spec.math = BS_MATH_FAST;
            else if (v == "exact") spec.math = BS_MATH_EXACT;
            else return fail_usage("invalid value '" + v + "' for '--math'");
        } else if (a == "--kernel") {
            if (v == "fused") spec.kernel = BS_KERNEL_FUSED;
            else if (v == "two-pass") spec.kernel = BS_KERNEL_TWO_PASS;
            else return fail_usage("invalid value '" + v + "' for '--kernel'");
        } else if (a == "--eos") {
            if (v == "bag") spec.eos = BS_EOS_BAG;
            else if (v == "gamma-law") spec.eos = BS_EOS_GAMMA_LAW;
            else return fail_usage("invalid value '" + v + "' for '--eos'");
        } else if (a == "--dump") {
            dump_path = v;
        } else {
            return fail_usage("unexpected argument '" + a + "' found");
        }
    }
    if (spec.precision == BS_PRECISION_F64) spec.math = BS_MATH_EXACT;
    if (check_f64 && spec.precision == BS_PRECISION_F64)
        return fail_usage("'--check-f64' compares an f32 run with its f64 copy; the run is already f64");
    spec.model.seed_kind = have_seed_radius ? BS_SEED_CELLS : BS_SEED_CRITICAL;
    spec.model.seed_value = have_seed_radius ? seed_radius : seed_factor;

    // Config::validate (config.rs:151-178)
    if (bs_spec_validate(&spec) != BS_OK) {
        std::fprintf(stderr, "Error: %s\n", bs_last_error());
        return 1;
    }
    if (steps_per_frame == 0) {
        std::fprintf(stderr, "Error: --steps-per-frame must be at least 1\n");
        return 1;
    }
    // a screenshot implies a windowless run (main.rs:29-33)
    if (!have_headless && have_screenshot) {
        have_headless = true;
        headless = 0;
    }
    if (!have_headless) {
        std::fprintf(stderr, "Error: the interactive window is not part of this build; run with --headless <STEPS>\n");
        return 1;
    }
    const uint64_t steps = headless;
    const uint64_t every = have_report ? report_every : (steps / 20 > 1 ? steps / 20 : 1);  // main.rs:32
    if (every == 0) {
        std::fprintf(stderr, "Error: --report-every must be at least 1\n");
        return 1;
    }

    char name[256] = "unknown";
    if (bs_device_name(name, sizeof name) != BS_OK) {
        std::fprintf(stderr, "Error: no suitable GPU adapter found: %s\n", bs_last_error());
        return 1;
    }
    std::printf("adapter: %s (HIP)\n", name);

    bs_sim *sim = nullptr;
    if (bs_sim_create(&spec, &sim) != BS_OK) {
        std::fprintf(stderr, "Error: failed to create the simulation\n\nCaused by:\n    %s\n", bs_last_error());
        return 1;
    }
    bs_derived d;
    bs_model_derive(&spec.model, &d);
    std::printf("lattice %ux%ux%u = %.2f M cells | %.2f GB | dt = %.3f\n", spec.grid[0], spec.grid[1], spec.grid[2],
                (double)bs_sim_cell_count(sim) / 1e6, (double)bs_sim_device_memory_bytes(sim) / 1e9, (double)d.dt);
    std::printf("wall %.2f cells | R_c %.1f cells | sigma %.4f | eps %.4f | e_ref %.4f\n", (double)d.wall_width,
                (double)d.critical_radius, (double)d.surface_tension, (double)d.eps, (double)d.e_ref);
    if (steps > 0) {
        std::printf("\n%10s %10s %10s %10s %11s %11s %10s\n", "step", "time", "broken", "R_eq", "max|v|", "K", "E/E0");
    }
    std::fflush(stdout);

    bool have_base = false;
    double base = 0.0;
    uint64_t base_step = 0;
    size_t base_idx = 0;
    std::vector<double> energies;  // mean energy at each report (for --check-f64)
    uint64_t done = 0;
    const auto start = std::chrono::steady_clock::now();
    int rc = 0;
    while (done < steps) {
        const uint64_t chunk = every < steps - done ? every : steps - done;
        if (bs_sim_advance(sim, chunk) != BS_OK) {
            std::fprintf(stderr, "Error: %s\n", bs_last_error());
            rc = 1;
            break;
        }
        bs_diagnostics diag;
        if (bs_sim_diagnostics(sim, &diag) != BS_OK) {
            std::fprintf(stderr, "Error: %s\n", bs_last_error());
            rc = 1;
            break;
        }
        done += chunk;
        if (diag.nonfinite || !std::isfinite(diag.mean_energy) || !std::isfinite(diag.max_velocity)) {
            std::fprintf(stderr,
                         "Error: the solution blew up at step %llu (non-finite diagnostics); reduce --cfl or --alpha\n",
                         (unsigned long long)done);
            rc = 1;
            break;
        }
        if (!have_base && bs_sim_remaining_bubbles(sim) == 0) {
            have_base = true;
            base = diag.mean_energy;
            base_step = done;
            base_idx = energies.size();
        }
        energies.push_back(diag.mean_energy);
        char ratio[64];
        if (have_base) std::snprintf(ratio, sizeof ratio, "%.6f", diag.mean_energy / base);
        else std::snprintf(ratio, sizeof ratio, "-");
        const float cells = (float)bs_sim_cell_count(sim);
        const float volume = (float)diag.broken_fraction * cells;
        const float r_eq = std::cbrt(volume * 3.0f / (4.0f * 3.14159265358979323846f));
        std::printf("%10llu %10.1f %9.2f%% %10.1f %11.4f %s %10s\n", (unsigned long long)done, (double)bs_sim_time(sim),
                    100.0 * diag.broken_fraction, (double)r_eq, diag.max_velocity,
                    rpad(rust_exp(diag.kinetic_fraction, 3), 11).c_str(), ratio);
        std::fflush(stdout);
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    if (!rc && steps > 0) {
        const double cell_steps = (double)steps * (double)bs_sim_cell_count(sim);
        std::printf("\n%llu steps in %.2f s  (%.1f steps/s, %.2f G cell-updates/s)\n", (unsigned long long)steps,
                    elapsed, (double)steps / elapsed, cell_steps / elapsed / 1e9);
    }
    if (!rc && check_f64 && steps > 0) rc = check_against_f64(spec, sim, steps, every, energies, have_base, base_idx, base_step);
    if (!rc && !dump_path.empty()) {
        const size_t real = spec.precision == BS_PRECISION_F64 ? 8 : 4;
        const uint64_t n = bs_sim_cell_count(sim);
        std::vector<unsigned char> buf(n * 6 * real);
        if (bs_sim_download(sim, buf.data(), buf.data() + n * 2 * real) != BS_OK) {
            std::fprintf(stderr, "Error: %s\n", bs_last_error());
            rc = 1;
        } else {
            FILE *f = std::fopen(dump_path.c_str(), "wb");
            if (!f || std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
                std::fprintf(stderr, "Error: cannot write %s\n", dump_path.c_str());
                rc = 1;
            }
            if (f) std::fclose(f);
        }
    }
    if (!rc && have_screenshot) {
        // capture::render_to_png with RenderSettings::with_field (headless.rs:102-113)
        bs_render_settings rs;
        if (bs_render_settings_default(field_mode, &rs) != BS_OK ||
            bs_sim_screenshot(sim, &rs, shot_w, shot_h, screenshot_path.c_str()) != BS_OK) {
            std::fprintf(stderr, "Error: %s\n", bs_last_error());
            rc = 1;
        } else {
            std::printf("wrote %s (%ux%u)\n", screenshot_path.c_str(), shot_w, shot_h);
        }
    }
    bs_sim_destroy(sim);
    return rc;
}

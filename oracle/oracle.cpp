// oracle.cpp -- CPU restatement of the reference lattice solver.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
// check and the CPU-baseline leg of the benchmark load it; the library never
// does.  It restates, operation by operation and in the reference's
// left-to-right expression order, the WGSL solver of AgenttiX/bubble-simulation:
//
//   shaders/common.wgsl     wrap1, cell_index, minmod3, limit_slope(4)
//   shaders/potential.wgsl  potential, dpotential
//   shaders/eos.wgsl        pressure_from_conserved, regularise, prim_from_cons,
//                           cons_from_prim, flux, wave_speeds
//   shaders/primitives.wgsl primitive pass
//   shaders/step.wgsl       sanitise, hll_flux, the stage (faces from both sides)
//   shaders/init.wgsl       initial condition
//   shaders/nucleate.wgsl   min_image, tanh profile, max-merge
//   shaders/reduce.wgsl     cell_contribution; naive f32 two-pass summation
//   shaders/vis.wgsl        visualisation channels (after the primitive pass)
//   shaders/volume.wgsl     the ray march of fs_main over those channels, with
//                           the trilinear clamp-to-edge filter of the sampler
//                           (gpu/render.rs:184-196) written out
//   src/gpu/sim.rs          RK3 (:135), stage_slots (:1046-1051), record_steps /
//                           record_nucleation / record_step (:829-916)
//
// Templated on the real type: double is the accuracy oracle, float is the
// reference's own arithmetic.  Compiled with -ffp-contract=off, IEEE '/'
// and sqrt.  The WGSL `tanh` has implementation-defined accuracy; this file
// uses expm1-based tanh (Cody-Waite reduction of 2|x| by ln 2, Taylor series
// of expm1 on the reduced argument), which the tests pin against libm's tanh.
#ifdef _OPENMP
#include <omp.h>
#endif

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <class R> struct P4 { R x, y, z, w; };

template <class R> struct Prm {
    int nx, ny, nz;
    R dx, inv_dx, dt, eta, lambda, eps, phi_b, m2, delta, a_rad, wall_width, p_floor, e_floor, z_max, e_ref, t_ref;
    int eos;       // 0: bag model (eos.wgsl); 1: gamma-law (docs/EOS.md:86-128)
    R eos_gamma;   // gamma-law adiabatic index
    R dof_ratio;   // temperature dof ratio across the wall (docs/EOS.md:65-84)
};

// WGSL min/max: min(e1,e2) = e2 if e2 < e1 else e1; max = e2 if e1 < e2 else e1
template <class R> inline R wmin(R a, R b) { return b < a ? b : a; }
template <class R> inline R wmax(R a, R b) { return a < b ? b : a; }
template <class R> inline R wsign(R a) { return a > R(0) ? R(1) : (a < R(0) ? R(-1) : R(0)); }
template <class R> inline R wdot(const P4<R> &a, const P4<R> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R tiny() { return R(1e-30); }
template <class R> inline R den_floor() { return R(1e-12); }
template <> inline float tiny<float>() { return 1e-30f; }
template <> inline float den_floor<float>() { return 1e-12f; }

// common.wgsl:80-95 (a full modulo; every offset used is within +-2 and the
// lattice is at least 8 cells along each axis)
inline int wrap1(int i, int n) {
    int r = i;
    if (r < 0) r += n;
    if (r < 0) r += n;
    if (r >= n) r -= n;
    if (r >= n) r -= n;
    return r;
}
template <class R> inline int64_t cell_index(const Prm<R> &P, int x, int y, int z) {
    return ((int64_t)wrap1(z, P.nz) * P.ny + wrap1(y, P.ny)) * P.nx + wrap1(x, P.nx);
}

// common.wgsl:114-136
template <class R> inline R minmod3(R a, R b, R c) {
    const R sa = wsign(a);
    if (sa != wsign(b) || sa != wsign(c)) return R(0);
    return sa * wmin(std::fabs(a), wmin(std::fabs(b), std::fabs(c)));
}
template <class R> inline R limit_slope(R dm, R dp) { return minmod3(R(0.5) * (dm + dp), R(2) * dm, R(2) * dp); }
template <class R> inline P4<R> limit_slope4(const P4<R> &dm, const P4<R> &dp) {
    return {limit_slope(dm.x, dp.x), limit_slope(dm.y, dp.y), limit_slope(dm.z, dp.z), limit_slope(dm.w, dp.w)};
}

// potential.wgsl:38-47
template <class R> inline R potential(const Prm<R> &P, R phi) {
    const R p2 = phi * phi;
    return R(0.5) * P.m2 * p2 - P.delta * p2 * phi / R(3) + R(0.25) * P.lambda * p2 * p2;
}
template <class R> inline R dpotential(const Prm<R> &P, R phi) {
    return phi * (P.m2 - P.delta * phi + P.lambda * phi * phi);
}

// eos.wgsl (bag model) and the gamma-law fluid p = (Gamma - 1) e of
// docs/EOS.md:86-128: c_s^2 = Gamma - 1, w = Gamma p / (Gamma - 1), inverted
// by a fixed count of Newton steps on f(p) = p - (Gamma - 1)(E - Z^2 / (E + p))
// from p0 = (Gamma - 1) E.  (The document's sketch writes de/dp as
// -(1 + Z^2/s^2) - 1; the derivative of its own e(p) = (s^2 - Z^2)/s - p is
// Z^2 / s^2, which is what is used here.)
template <class R> inline R eos_sound_speed2(const Prm<R> &P, R, R) {
    return P.eos == 1 ? P.eos_gamma - R(1) : R(1) / R(3);
}
template <class R> inline R eos_enthalpy_from_pressure(const Prm<R> &P, R p, R) {
    return P.eos == 1 ? P.eos_gamma * p / (P.eos_gamma - R(1)) : R(4) * p;
}
template <class R> inline R eos_pressure_from_conserved(const Prm<R> &P, R e_lab, R z2, R) {
    if (P.eos == 1) {
        const R gm1 = P.eos_gamma - R(1);
        R p = gm1 * e_lab;
        for (int it = 0; it < 8; ++it) {
            const R s = e_lab + p;
            const R rs = R(1) / s;
            const R zs = z2 * rs;
            const R f = p - gm1 * (e_lab - zs);
            const R df = R(1) - gm1 * (zs * rs);
            p = p - f / df;
        }
        return wmax(p, P.p_floor);
    }
    const R disc = wmax(R(4) * e_lab * e_lab - R(3) * z2, R(0));
    const R p = (e_lab * e_lab - z2) / (e_lab + std::sqrt(disc));
    return wmax(p, P.p_floor);
}
template <class R> inline P4<R> eos_regularise(const Prm<R> &P, const P4<R> &u) {
    const R e_lab = wmax(u.x, P.e_floor);
    P4<R> z{u.y, u.z, u.w, R(0)};
    const R zlen = std::sqrt(wdot(z, z));
    const R zcap = P.z_max * e_lab;
    P4<R> zc = z;
    if (zlen > zcap) {
        const R f = zcap / wmax(zlen, tiny<R>());
        zc = {z.x * f, z.y * f, z.z * f, R(0)};
    }
    return {e_lab, zc.x, zc.y, zc.z};
}
template <class R> inline P4<R> eos_prim_from_cons(const Prm<R> &P, const P4<R> &u_raw, R phi) {
    const P4<R> u = eos_regularise(P, u_raw);
    const R e_lab = u.x;
    const P4<R> z{u.y, u.z, u.w, R(0)};
    const R p = eos_pressure_from_conserved(P, e_lab, wdot(z, z), phi);
    const R s = e_lab + p;
    const R sm = wmax(s, tiny<R>());
    return {z.x / sm, z.y / sm, z.z / sm, p};
}
template <class R> inline P4<R> eos_cons_from_prim(const Prm<R> &P, const P4<R> &q, R phi) {
    const R v2 = wmin(wdot(q, q), P.z_max * P.z_max);
    const R w_gam2 = eos_enthalpy_from_pressure(P, q.w, phi) / (R(1) - v2);
    return {w_gam2 - q.w, w_gam2 * q.x, w_gam2 * q.y, w_gam2 * q.z};
}
template <class R> inline R axis_comp(const P4<R> &v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
template <class R> inline P4<R> eos_flux(const P4<R> &q, const P4<R> &u, int axis) {
    const R p = q.w;
    const R vd = axis_comp(q, axis);
    P4<R> f{(u.x + p) * vd, u.y * vd, u.z * vd, u.w * vd};
    if (axis == 0) f.y = f.y + p;
    if (axis == 1) f.z = f.z + p;
    if (axis == 2) f.w = f.w + p;
    return f;
}
template <class R> inline void eos_wave_speeds(const Prm<R> &P, const P4<R> &q, R phi, int axis, R &lo, R &hi) {
    const R v2 = wmin(wdot(q, q), P.z_max * P.z_max);
    const R vd = axis_comp(q, axis);
    const R cs2 = eos_sound_speed2(P, q.w, phi);
    const R cs = std::sqrt(cs2);
    const R denom = R(1) - v2 * cs2;
    const R rad = wmax((R(1) - v2) * (R(1) - vd * vd - cs2 * (v2 - vd * vd)), R(0));
    const R a = vd * (R(1) - cs2) / denom;
    const R b = cs * std::sqrt(rad) / denom;
    lo = a - b;
    hi = a + b;
}

// step.wgsl:63-90
template <class R> inline P4<R> sanitise(const Prm<R> &P, const P4<R> &q) {
    const R p = wmax(q.w, P.p_floor);
    P4<R> v{q.x, q.y, q.z, R(0)};
    const R v2 = wdot(v, v);
    const R vmax2 = P.z_max * P.z_max;
    if (v2 > vmax2) {
        const R f = std::sqrt(vmax2 / wmax(v2, tiny<R>()));
        v = {v.x * f, v.y * f, v.z * f, R(0)};
    }
    return {v.x, v.y, v.z, p};
}
template <class R>
inline P4<R> hll_flux(const Prm<R> &P, const P4<R> &q_l, R phi_l, const P4<R> &q_r, R phi_r, int axis) {
    const P4<R> u_l = eos_cons_from_prim(P, q_l, phi_l);
    const P4<R> u_r = eos_cons_from_prim(P, q_r, phi_r);
    const P4<R> f_l = eos_flux(q_l, u_l, axis);
    const P4<R> f_r = eos_flux(q_r, u_r, axis);
    R llo, lhi, rlo, rhi;
    eos_wave_speeds(P, q_l, phi_l, axis, llo, lhi);
    eos_wave_speeds(P, q_r, phi_r, axis, rlo, rhi);
    const R s_min = wmin(wmin(llo, rlo), R(0));
    const R s_max = wmax(wmax(lhi, rhi), R(0));
    const R den = wmax(s_max - s_min, den_floor<R>());
    const R smm = s_min * s_max;
    return {(s_max * f_l.x - s_min * f_r.x + smm * (u_r.x - u_l.x)) / den,
            (s_max * f_l.y - s_min * f_r.y + smm * (u_r.y - u_l.y)) / den,
            (s_max * f_l.z - s_min * f_r.z + smm * (u_r.z - u_l.z)) / den,
            (s_max * f_l.w - s_min * f_r.w + smm * (u_r.w - u_l.w)) / den};
}
template <class R> inline P4<R> sub(const P4<R> &a, const P4<R> &b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
template <class R> inline P4<R> addh(const P4<R> &a, const P4<R> &s) {
    return {a.x + R(0.5) * s.x, a.y + R(0.5) * s.y, a.z + R(0.5) * s.z, a.w + R(0.5) * s.w};
}
template <class R> inline P4<R> subh(const P4<R> &a, const P4<R> &s) {
    return {a.x - R(0.5) * s.x, a.y - R(0.5) * s.y, a.z - R(0.5) * s.z, a.w - R(0.5) * s.w};
}

// tanh (see header)
template <class R> struct TK;
template <> struct TK<float> {
    static constexpr int terms = 9;
    static constexpr float hi = 0.693145751953125f, lo = 1.428606765330187045e-06f, inv = 1.44269504088896341f, big = 9.0f;
};
template <> struct TK<double> {
    static constexpr int terms = 16;
    static constexpr double hi = 6.93147180369123816490e-01, lo = 1.90821492927058770002e-10,
                            inv = 1.44269504088896338700e+00, big = 19.5;
};
template <class R> R o_expm1_pos(R a) {
    const R kf = std::floor(a * TK<R>::inv + R(0.5));
    const int k = (int)kf;
    const R r = (a - kf * TK<R>::hi) - kf * TK<R>::lo;
    R t = R(1);
    for (int i = TK<R>::terms; i >= 2; --i) t = R(1) + t * r / R(i);
    const R em1r = r * t;
    if (k == 0) return em1r;
    R f = R(1);
    for (int i = 0; i < k; ++i) f = f * R(2);
    return (em1r + R(1)) * f - R(1);
}
template <class R> R o_tanh(R x) {
    const R ax = std::fabs(x);
    R t;
    if (ax >= TK<R>::big) t = R(1);
    else {
        const R e = o_expm1_pos(R(2) * ax);
        t = e / (e + R(2));
    }
    return x < R(0) ? -t : t;
}

// nucleate.wgsl:52-57
template <class R> inline R min_image(R d, R n) {
    R r = d;
    if (r > R(0.5) * n) r = r - n;
    if (r < R(-0.5) * n) r = r + n;
    return r;
}

// ---------------------------------------------------------------------------
//  The lattice
// ---------------------------------------------------------------------------
struct Event {
    float time, x, y, z, r0;
};

template <class R> struct Lattice {
    Prm<R> P;
    float dt32;  // the f32 time step of the reference clock
    int64_t n;
    std::vector<R> field[3];  // (phi, pi) interleaved
    std::vector<R> fluid[3];  // (E, Zx, Zy, Zz) interleaved
    std::vector<R> prim;      // (vx, vy, vz, p)
    int cur = 0;
    float time = 0.0f;
    uint64_t steps = 0;
    std::vector<Event> schedule;
    size_t next_event = 0;

    P4<R> fl(int s, int64_t i) const { return {fluid[s][4 * i], fluid[s][4 * i + 1], fluid[s][4 * i + 2], fluid[s][4 * i + 3]}; }
    P4<R> pr(int64_t i) const { return {prim[4 * i], prim[4 * i + 1], prim[4 * i + 2], prim[4 * i + 3]}; }

    // init.wgsl:18-26 on slot cur
    void init() {
        for (int64_t i = 0; i < n; ++i) {
            field[cur][2 * i] = R(0);
            field[cur][2 * i + 1] = R(0);
            fluid[cur][4 * i] = P.e_ref;
            fluid[cur][4 * i + 1] = R(0);
            fluid[cur][4 * i + 2] = R(0);
            fluid[cur][4 * i + 3] = R(0);
        }
    }

    // nucleate.wgsl:59-85 on slot cur
    void nucleate(const std::vector<Event> &batch) {
        const R nx = R(P.nx), ny = R(P.ny), nz = R(P.nz);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const int x = (int)(i % P.nx), y = (int)((i / P.nx) % P.ny), z = (int)(i / ((int64_t)P.nx * P.ny));
            const R px = R(x), py = R(y), pz = R(z);
            R phi = field[cur][2 * i];
            for (const Event &b : batch) {
                const P4<R> d{min_image(px - R(b.x), nx), min_image(py - R(b.y), ny), min_image(pz - R(b.z), nz), R(0)};
                const R r = std::sqrt(wdot(d, d)) * P.dx;
                const R profile = R(0.5) * P.phi_b * (R(1) - o_tanh((r - R(b.r0) * P.dx) / P.wall_width));
                phi = wmax(phi, profile);
            }
            field[cur][2 * i] = phi;
        }
    }

    // primitives.wgsl:16-23
    void primitives(int in) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const P4<R> q = eos_prim_from_cons(P, fl(in, i), field[in][2 * i]);
            prim[4 * i] = q.x;
            prim[4 * i + 1] = q.y;
            prim[4 * i + 2] = q.z;
            prim[4 * i + 3] = q.w;
        }
    }

    // step.wgsl:92-174
    void step_pass(int sn, int in, int out, R a0, R a1, R dt_stage) {
#pragma omp parallel for schedule(static)
        for (int64_t i0 = 0; i0 < n; ++i0) {
            const int cx = (int)(i0 % P.nx), cy = (int)((i0 / P.nx) % P.ny), cz = (int)(i0 / ((int64_t)P.nx * P.ny));
            const R phi0 = field[in][2 * i0], pi0 = field[in][2 * i0 + 1];
            const P4<R> q0 = sanitise(P, pr(i0));
            R lap_sum = R(0);
            R grad[3] = {R(0), R(0), R(0)};
            P4<R> flux_div{R(0), R(0), R(0), R(0)};
            for (int axis = 0; axis < 3; ++axis) {
                const int ex = axis == 0, ey = axis == 1, ez = axis == 2;
                const int64_t im1 = cell_index(P, cx - ex, cy - ey, cz - ez);
                const int64_t ip1 = cell_index(P, cx + ex, cy + ey, cz + ez);
                const int64_t im2 = cell_index(P, cx - 2 * ex, cy - 2 * ey, cz - 2 * ez);
                const int64_t ip2 = cell_index(P, cx + 2 * ex, cy + 2 * ey, cz + 2 * ez);
                const R phi_m = field[in][2 * im1];
                const R phi_p = field[in][2 * ip1];
                lap_sum = lap_sum + (phi_m + phi_p - R(2) * phi0);
                grad[axis] = R(0.5) * (phi_p - phi_m) * P.inv_dx;
                const P4<R> qm2 = sanitise(P, pr(im2));
                const P4<R> qm1 = sanitise(P, pr(im1));
                const P4<R> qp1 = sanitise(P, pr(ip1));
                const P4<R> qp2 = sanitise(P, pr(ip2));
                const P4<R> slope_m1 = limit_slope4(sub(qm1, qm2), sub(q0, qm1));
                const P4<R> slope_0 = limit_slope4(sub(q0, qm1), sub(qp1, q0));
                const P4<R> slope_p1 = limit_slope4(sub(qp1, q0), sub(qp2, qp1));
                const P4<R> f_minus =
                    hll_flux(P, sanitise(P, addh(qm1, slope_m1)), phi_m, sanitise(P, subh(q0, slope_0)), phi0, axis);
                const P4<R> f_plus =
                    hll_flux(P, sanitise(P, addh(q0, slope_0)), phi0, sanitise(P, subh(qp1, slope_p1)), phi_p, axis);
                flux_div = {flux_div.x + (f_plus.x - f_minus.x) * P.inv_dx, flux_div.y + (f_plus.y - f_minus.y) * P.inv_dx,
                            flux_div.z + (f_plus.z - f_minus.z) * P.inv_dx, flux_div.w + (f_plus.w - f_minus.w) * P.inv_dx};
            }
            const R laplacian = lap_sum * P.inv_dx * P.inv_dx;
            const P4<R> gv{grad[0], grad[1], grad[2], R(0)};
            const R gam = R(1) / std::sqrt(wmax(R(1) - wdot(q0, q0), den_floor<R>()));
            const R dphi_proper = gam * (pi0 + wdot(q0, gv));
            const R drag = P.eta * dphi_proper;
            const R d_phi = pi0;
            const R d_pi = laplacian - dpotential(P, phi0) - drag;
            P4<R> d_fluid{-flux_div.x, -flux_div.y, -flux_div.z, -flux_div.w};
            d_fluid.x = d_fluid.x + drag * pi0;
            d_fluid.y = d_fluid.y - drag * grad[0];
            d_fluid.z = d_fluid.z - drag * grad[1];
            d_fluid.w = d_fluid.w - drag * grad[2];
            const R fu_phi = phi0 + dt_stage * d_phi;
            const R fu_pi = pi0 + dt_stage * d_pi;
            const P4<R> uin = fl(in, i0);
            const P4<R> fu{uin.x + dt_stage * d_fluid.x, uin.y + dt_stage * d_fluid.y, uin.z + dt_stage * d_fluid.z,
                           uin.w + dt_stage * d_fluid.w};
            const R fn_phi = field[sn][2 * i0], fn_pi = field[sn][2 * i0 + 1];
            const P4<R> un = fl(sn, i0);
            field[out][2 * i0] = a0 * fn_phi + a1 * fu_phi;
            field[out][2 * i0 + 1] = a0 * fn_pi + a1 * fu_pi;
            fluid[out][4 * i0] = a0 * un.x + a1 * fu.x;
            fluid[out][4 * i0 + 1] = a0 * un.y + a1 * fu.y;
            fluid[out][4 * i0 + 2] = a0 * un.z + a1 * fu.z;
            fluid[out][4 * i0 + 3] = a0 * un.w + a1 * fu.w;
        }
    }

    // record_step (sim.rs:881-916) with RK3 (sim.rs:135) and stage_slots
    // (sim.rs:1046-1051)
    void record_step() {
        const float rk32[3][2] = {{0.0f, 1.0f}, {0.75f, 0.25f}, {1.0f / 3.0f, 2.0f / 3.0f}};
        const int a = cur, b = (cur + 1) % 3, c = (cur + 2) % 3;
        const int slots[3][2] = {{a, b}, {b, c}, {c, b}};
        for (int s = 0; s < 3; ++s) {
            const int in = slots[s][0], out = slots[s][1];
            primitives(in);
            step_pass(cur, in, out, R(rk32[s][0]), R(rk32[s][1]), P.dt);
        }
        cur = (cur + 1) % 3;
        time += dt32;
        steps += 1;
    }

    // record_nucleation (sim.rs:843-879)
    bool record_nucleation() {
        std::vector<Event> batch;
        while (next_event < schedule.size() && schedule[next_event].time <= time && batch.size() < 1024) {
            batch.push_back(schedule[next_event]);
            ++next_event;
        }
        if (batch.empty()) return false;
        nucleate(batch);
        return true;
    }
};

// reduce.wgsl:56-79 for cell i of slot s (the primitive cache holds slot s)
template <class R> inline void cell_contribution(const Lattice<R> &L, int s, int64_t i, R out[4], R &vabs) {
    const Prm<R> &P = L.P;
    const int x = (int)(i % P.nx), y = (int)((i / P.nx) % P.ny), z = (int)(i / ((int64_t)P.nx * P.ny));
    const R phi = L.field[s][2 * i], pi = L.field[s][2 * i + 1];
    R grad2 = R(0);
    for (int axis = 0; axis < 3; ++axis) {
        const int ex = axis == 0, ey = axis == 1, ez = axis == 2;
        const R d = R(0.5) * (L.field[s][2 * cell_index(P, x + ex, y + ey, z + ez)] -
                              L.field[s][2 * cell_index(P, x - ex, y - ey, z - ez)]) * P.inv_dx;
        grad2 = grad2 + d * d;
    }
    const R e_lab = wmax(L.fluid[s][4 * i], P.e_floor);
    const P4<R> q = L.pr(i);
    const R v2 = wdot(q, q);
    out[0] = R(0.5) * pi * pi + R(0.5) * grad2 + potential(P, phi) + e_lab;
    out[1] = phi > R(0.5) * P.phi_b ? R(1) : R(0);
    out[2] = (e_lab + q.w) * v2;
    out[3] = e_lab;
    vabs = std::sqrt(v2);
}

template <class R> Prm<R> make_prm(const float *p32) {
    // p32 holds the 96-byte SimParams block as 24 words (sim.rs:66-98)
    uint32_t u[4];
    std::memcpy(u, p32, 16);
    Prm<R> P;
    P.nx = (int)u[0]; P.ny = (int)u[1]; P.nz = (int)u[2];
    P.dx = p32[4]; P.inv_dx = p32[5]; P.dt = p32[6]; P.eta = p32[7];
    P.lambda = p32[8]; P.eps = p32[9]; P.phi_b = p32[10]; P.m2 = p32[11];
    P.delta = p32[12]; P.a_rad = p32[13]; P.wall_width = p32[14]; P.p_floor = p32[15];
    P.e_floor = p32[16]; P.z_max = p32[17]; P.e_ref = p32[18]; P.t_ref = p32[19];
    // the block's padding words carry the equation of state for the oracle:
    // word 3 the kind, words 22 and 23 gamma and the dof ratio (zero in the
    // reference's block: the bag model, dof ratio 1)
    P.eos = (int)u[3];
    P.eos_gamma = P.eos == 1 ? (R)p32[22] : (R)(4.0f / 3.0f);
    P.dof_ratio = p32[23] != 0.0f ? (R)p32[23] : R(1);
    return P;
}

struct Handle {
    int precision;  // 0 f32, 1 f64
    Lattice<float> *lf = nullptr;
    Lattice<double> *ld = nullptr;
};

template <class R> Lattice<R> *make_lattice(const float *params, const float *events, int nev) {
    Lattice<R> *L = new Lattice<R>();
    L->P = make_prm<R>(params);
    L->dt32 = params[6];
    L->n = (int64_t)L->P.nx * L->P.ny * L->P.nz;
    for (int s = 0; s < 3; ++s) {
        L->field[s].assign(2 * L->n, R(0));
        L->fluid[s].assign(4 * L->n, R(0));
    }
    L->prim.assign(4 * L->n, R(0));
    for (int k = 0; k < nev; ++k) {
        const float *e = events + 5 * k;
        L->schedule.push_back(Event{e[0], e[1], e[2], e[3], e[4]});
    }
    L->init();
    return L;
}

// eos_temperature (eos.wgsl:65-69): T = (3 max(p, p_floor) / a)^(1/4)
template <class R> inline R eos_temperature(const Prm<R> &P, R p, R phi) {
    R a = P.a_rad;
    if (P.dof_ratio != R(1)) {
        // docs/EOS.md:70-76: a(phi) = a (1 + (dof_ratio - 1) h^2 (3 - 2h))
        const R h = wmin(wmax(phi / P.phi_b, R(0)), R(1));
        const R smooth = h * h * (R(3) - R(2) * h);
        a = P.a_rad * (R(1) + (P.dof_ratio - R(1)) * smooth);
    }
    const R x = R(3) * wmax(p, P.p_floor) / a;
    return std::sqrt(std::sqrt(x));
}

// record_visualisation (sim.rs:920-942): primitives.wgsl on slot cur, then
// vis.wgsl:38-55 per cell; the four channels narrowed to f32 (f32(...)),
// before the texture's own conversion to half
template <class R> void vis_pack(const Lattice<R> &L, float *out) {
    const Prm<R> &P = L.P;
    for (int64_t i = 0; i < L.n; ++i) {
        const R phi = L.field[L.cur][2 * i];
        const P4<R> u = L.fl(L.cur, i);
        const P4<R> q = eos_prim_from_cons(P, u, phi);
        const R e_lab = wmax(u.x, P.e_floor);
        const R v2 = wdot(q, q);
        const R speed = std::sqrt(v2);
        const R temp = eos_temperature(P, q.w, phi);
        const R kinetic = (e_lab + q.w) * v2;
        out[4 * i] = (float)(phi / P.phi_b);
        out[4 * i + 1] = (float)speed;
        out[4 * i + 2] = (float)(temp / P.t_ref - R(1));
        out[4 * i + 3] = (float)(kinetic / P.e_ref);
    }
}

template <class R> void diag_exact(Lattice<R> &L, double *out) {
    // per-cell contributions in R; sums in long double
    L.primitives(L.cur);
    long double acc[4] = {0, 0, 0, 0};
    double vmax = 0.0;
    for (int64_t i = 0; i < L.n; ++i) {
        R c[4], va;
        cell_contribution(L, L.cur, i, c, va);
        for (int k = 0; k < 4; ++k) acc[k] += (long double)c[k];
        if ((double)va > vmax) vmax = (double)va;
    }
    for (int k = 0; k < 4; ++k) out[k] = (double)acc[k];
    out[4] = vmax;
}

// reduce.wgsl:81-140 with its exact f32 summation order: 1024 groups x 256
// threads grid-striding, shared-memory halving tree, one-group second pass
void diag_naive_f32(Lattice<float> &L, float *out4, uint32_t *vmax_bits) {
    L.primitives(L.cur);
    const uint32_t NP = 1024, WG = 256, stride = NP * WG;
    std::vector<float> partials(4 * NP);
    uint32_t maxbits = 0;
    std::vector<float> sh(4 * WG), shm(WG);
    for (uint32_t wid = 0; wid < NP; ++wid) {
        for (uint32_t lid = 0; lid < WG; ++lid) {
            float acc[4] = {0, 0, 0, 0};
            float vmax = 0.0f;
            for (uint64_t i = (uint64_t)wid * WG + lid; i < (uint64_t)L.n; i += stride) {
                float c[4], va;
                cell_contribution(L, L.cur, (int64_t)i, c, va);
                for (int k = 0; k < 4; ++k) acc[k] = acc[k] + c[k];
                vmax = wmax(vmax, va);
            }
            for (int k = 0; k < 4; ++k) sh[4 * lid + k] = acc[k];
            shm[lid] = vmax;
        }
        for (uint32_t s = WG / 2; s > 0; s >>= 1)
            for (uint32_t lid = 0; lid < s; ++lid) {
                for (int k = 0; k < 4; ++k) sh[4 * lid + k] = sh[4 * lid + k] + sh[4 * (lid + s) + k];
                shm[lid] = wmax(shm[lid], shm[lid + s]);
            }
        for (int k = 0; k < 4; ++k) partials[4 * wid + k] = sh[k];
        uint32_t b;
        std::memcpy(&b, &shm[0], 4);
        if (b > maxbits) maxbits = b;
    }
    for (uint32_t lid = 0; lid < WG; ++lid) {
        float acc[4] = {0, 0, 0, 0};
        for (uint32_t i = lid; i < NP; i += WG)
            for (int k = 0; k < 4; ++k) acc[k] = acc[k] + partials[4 * i + k];
        for (int k = 0; k < 4; ++k) sh[4 * lid + k] = acc[k];
    }
    for (uint32_t s = WG / 2; s > 0; s >>= 1)
        for (uint32_t lid = 0; lid < s; ++lid)
            for (int k = 0; k < 4; ++k) sh[4 * lid + k] = sh[4 * lid + k] + sh[4 * (lid + s) + k];
    for (int k = 0; k < 4; ++k) out4[k] = sh[k];
    *vmax_bits = maxbits;
}

}  // namespace

// ---------------------------------------------------------------------------
//  ctypes surface
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
//  volume.wgsl -- fs_main per pixel, in f32, in the shader's expression order
// ---------------------------------------------------------------------------
namespace vol {

struct V3 { float x, y, z; };
inline V3 v3(float a, float b, float c) { return V3{a, b, c}; }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 muls(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 divv(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline float dotv(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 normalize(V3 a) { return divs(a, std::sqrt(dotv(a, a))); }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }  // WGSL mix
inline V3 mixv(V3 a, V3 b, float t) { return v3(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
inline float fract(float x) { return x - std::floor(x); }
inline float smoothstep(float lo, float hi, float x) {
    const float t = clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

const float VIRIDIS[11][3] = {{0.267f, 0.005f, 0.329f}, {0.283f, 0.141f, 0.458f}, {0.254f, 0.265f, 0.530f},
                              {0.207f, 0.372f, 0.553f}, {0.164f, 0.471f, 0.558f}, {0.128f, 0.567f, 0.551f},
                              {0.135f, 0.659f, 0.518f}, {0.267f, 0.749f, 0.441f}, {0.478f, 0.821f, 0.318f},
                              {0.741f, 0.873f, 0.150f}, {0.993f, 0.906f, 0.144f}};
const float INFERNO[11][3] = {{0.001f, 0.000f, 0.014f}, {0.078f, 0.044f, 0.214f}, {0.225f, 0.036f, 0.388f},
                              {0.373f, 0.074f, 0.432f}, {0.522f, 0.128f, 0.420f}, {0.665f, 0.182f, 0.370f},
                              {0.798f, 0.280f, 0.270f}, {0.902f, 0.413f, 0.144f}, {0.969f, 0.579f, 0.032f},
                              {0.988f, 0.762f, 0.135f}, {0.988f, 0.998f, 0.645f}};
const float SIGNED[11][3] = {{0.30f, 0.55f, 1.00f}, {0.22f, 0.42f, 0.85f}, {0.15f, 0.30f, 0.66f},
                             {0.09f, 0.19f, 0.45f}, {0.04f, 0.09f, 0.24f}, {0.02f, 0.02f, 0.03f},
                             {0.26f, 0.09f, 0.04f}, {0.48f, 0.17f, 0.06f}, {0.72f, 0.29f, 0.08f},
                             {0.92f, 0.48f, 0.14f}, {1.00f, 0.72f, 0.32f}};

// ramp (volume.wgsl:81-92)
V3 ramp(float t, int which) {
    const float x = clampf(t, 0.0f, 1.0f) * 10.0f;
    const int i0 = (int)std::floor(x);
    const int i1 = std::min(i0 + 1, 10);
    const float f = x - std::floor(x);
    const float (*c)[3] = which == 0 ? VIRIDIS : (which == 1 ? INFERNO : SIGNED);
    return mixv(v3(c[i0][0], c[i0][1], c[i0][2]), v3(c[i1][0], c[i1][1], c[i1][2]), f);
}

// safe_inv (volume.wgsl:98-101)
float safe_inv(float x) {
    const float s = x < 0.0f ? -1.0f : 1.0f;
    return s / std::fmax(std::fabs(x), 1e-9f);
}

float half_to_float(uint16_t h) {
    const int sign = h >> 15, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((float)(m | 0x400), e - 25);
    return sign ? -v : v;
}

struct S4 { float r, g, b, a; };

struct Scene {
    const float *V;  // the 44-float view uniform
    const uint16_t *tex;
    int n[3];
    float vm(int block, int k) const { return V[16 + 4 * block + k]; }  // 0 cam, 1 half, 2 grid, 3..6 render0..3
};

S4 texel(const Scene &S, int x, int y, int z) {
    const uint16_t *t = S.tex + 4 * (((size_t)z * S.n[1] + y) * S.n[0] + x);
    return S4{half_to_float(t[0]), half_to_float(t[1]), half_to_float(t[2]), half_to_float(t[3])};
}
S4 lerp4(S4 a, S4 b, float t) { return S4{mixf(a.r, b.r, t), mixf(a.g, b.g, t), mixf(a.b, b.b, t), mixf(a.a, b.a, t)}; }

// linear filtering with clamp-to-edge addressing: i0 = floor(u n - 1/2),
// weight frac(u n - 1/2), both texels clamped into [0, n - 1]
S4 sample_state(const Scene &S, V3 p) {
    const V3 half = v3(S.vm(1, 0), S.vm(1, 1), S.vm(1, 2));
    const V3 tc = add(muls(divv(p, half), 0.5f), v3(0.5f, 0.5f, 0.5f));  // to_texcoord
    const float u[3] = {tc.x, tc.y, tc.z};
    int lo[3], hi[3];
    float w[3];
    for (int a = 0; a < 3; ++a) {
        const float c = u[a] * (float)S.n[a] - 0.5f;
        const float f = std::floor(c);
        w[a] = c - f;
        lo[a] = std::min(std::max((int)f, 0), S.n[a] - 1);
        hi[a] = std::min(std::max((int)f + 1, 0), S.n[a] - 1);
    }
    const S4 c00 = lerp4(texel(S, lo[0], lo[1], lo[2]), texel(S, hi[0], lo[1], lo[2]), w[0]);
    const S4 c10 = lerp4(texel(S, lo[0], hi[1], lo[2]), texel(S, hi[0], hi[1], lo[2]), w[0]);
    const S4 c01 = lerp4(texel(S, lo[0], lo[1], hi[2]), texel(S, hi[0], lo[1], hi[2]), w[0]);
    const S4 c11 = lerp4(texel(S, lo[0], hi[1], hi[2]), texel(S, hi[0], hi[1], hi[2]), w[0]);
    return lerp4(lerp4(c00, c10, w[1]), lerp4(c01, c11, w[1]), w[2]);
}

// grad_phi (volume.wgsl:131-141)
V3 grad_phi(const Scene &S, V3 p) {
    const float h = S.vm(1, 3);
    const V3 dx = v3(h, 0.0f, 0.0f), dy = v3(0.0f, h, 0.0f), dz = v3(0.0f, 0.0f, h);
    const V3 g = v3(sample_state(S, add(p, dx)).r - sample_state(S, sub(p, dx)).r,
                    sample_state(S, add(p, dy)).r - sample_state(S, sub(p, dy)).r,
                    sample_state(S, add(p, dz)).r - sample_state(S, sub(p, dz)).r);
    return divs(g, 2.0f * h);
}

// hash12 (volume.wgsl:143-147)
float hash12(float px, float py) {
    V3 p3 = v3(fract(px * 0.1031f), fract(py * 0.1031f), fract(px * 0.1031f));
    const float d = dotv(p3, add(v3(p3.y, p3.z, p3.x), v3(33.33f, 33.33f, 33.33f)));
    p3 = add(p3, v3(d, d, d));
    return fract((p3.x + p3.y) * p3.z);
}

inline uint8_t unorm8(float c) { return (uint8_t)(clampf(c, 0.0f, 1.0f) * 255.0f + 0.5f); }

void pixel(const Scene &S, int px, int py, int W, int H, uint8_t *out) {
    const float *M = S.V;
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const float nx = fx / (float)W * 2.0f - 1.0f, ny = 1.0f - fy / (float)H * 2.0f;
    const V3 ro = v3(S.vm(0, 0), S.vm(0, 1), S.vm(0, 2));
    // inv_view_proj * vec4(ndc, 1, 1), column-major
    const float fw[4] = {M[0] * nx + M[4] * ny + M[8] * 1.0f + M[12] * 1.0f,
                         M[1] * nx + M[5] * ny + M[9] * 1.0f + M[13] * 1.0f,
                         M[2] * nx + M[6] * ny + M[10] * 1.0f + M[14] * 1.0f,
                         M[3] * nx + M[7] * ny + M[11] * 1.0f + M[15] * 1.0f};
    const V3 rd = normalize(sub(divs(v3(fw[0], fw[1], fw[2]), fw[3]), ro));
    const float horizon = clampf(rd.y * 0.5f + 0.5f, 0.0f, 1.0f);
    V3 color = muls(mixv(v3(0.015f, 0.017f, 0.028f), v3(0.045f, 0.055f, 0.085f), horizon), S.vm(6, 2));
    auto emit_raw = [&](V3 c) {
        out[0] = unorm8(c.x); out[1] = unorm8(c.y); out[2] = unorm8(c.z); out[3] = 255;
    };
    const V3 half = v3(S.vm(1, 0), S.vm(1, 1), S.vm(1, 2));
    const V3 inv_rd = v3(safe_inv(rd.x), safe_inv(rd.y), safe_inv(rd.z));
    // intersect_box (volume.wgsl:104-113)
    const V3 t0 = mulv(sub(v3(-half.x, -half.y, -half.z), ro), inv_rd);
    const V3 t1 = mulv(sub(half, ro), inv_rd);
    const V3 tmin = v3(std::fmin(t0.x, t1.x), std::fmin(t0.y, t1.y), std::fmin(t0.z, t1.z));
    const V3 tmax = v3(std::fmax(t0.x, t1.x), std::fmax(t0.y, t1.y), std::fmax(t0.z, t1.z));
    float t_near = std::fmax(std::fmax(std::fmax(tmin.x, tmin.y), tmin.z), 0.0f);
    float t_far = std::fmin(std::fmin(tmax.x, tmax.y), tmax.z);
    if (t_near >= t_far) return emit_raw(color);
    const int clip_axis = (int)S.vm(4, 1);
    if (clip_axis >= 0) {
        const V3 n_axis = v3(clip_axis == 0 ? 1.0f : 0.0f, clip_axis == 1 ? 1.0f : 0.0f, clip_axis == 2 ? 1.0f : 0.0f);
        const float plane = S.vm(4, 2) * S.vm(1, clip_axis);
        const float denom = dotv(rd, n_axis);
        const float dist = plane - dotv(ro, n_axis);
        if (std::fabs(denom) < 1e-6f) {
            if (dist < 0.0f) return emit_raw(color);
        } else {
            const float t_plane = dist / denom;
            if (denom > 0.0f) t_far = std::fmin(t_far, t_plane);
            else t_near = std::fmax(t_near, t_plane);
            if (t_near >= t_far) return emit_raw(color);
        }
    }
    if (S.vm(6, 1) > 0.5f) {
        const float w = 0.004f * S.vm(1, 3) * S.vm(2, 0);
        float edge = 0.0f;
        for (int k = 0; k < 2; ++k) {
            const V3 p = add(ro, muls(rd, k == 0 ? t_near : t_far));
            const V3 d = sub(half, v3(std::fabs(p.x), std::fabs(p.y), std::fabs(p.z)));
            const float median = std::fmax(std::fmin(d.x, d.y), std::fmin(std::fmax(d.x, d.y), d.z));
            edge = std::fmax(edge, 1.0f - smoothstep(0.0f, w, median));
        }
        color = add(color, muls(v3(0.20f, 0.26f, 0.34f), edge));
    }
    const float cell = S.vm(1, 3);
    const float samples_per_cell = std::fmax(S.vm(3, 2), 0.25f);
    float ds = cell / samples_per_cell;
    const float span = t_far - t_near;
    if (span / ds > 2048.0f) ds = span / 2048.0f;
    const int n_steps = std::min((int)std::ceil(span / ds), 2048);
    const float jitter = hash12(fx + S.vm(6, 0), fy + S.vm(6, 0));
    const float iso = S.vm(3, 0), wall_opacity = S.vm(3, 1), exposure = S.vm(3, 3);
    const float absorption = S.vm(5, 1);
    const bool show_iso = S.vm(5, 2) > 0.5f, show_vol = S.vm(5, 3) > 0.5f;
    const int mode = (int)S.vm(4, 0);
    const float gain = S.vm(5, 0);
    V3 accum = v3(0.0f, 0.0f, 0.0f);
    float alpha = 0.0f;
    float t = t_near + jitter * ds;
    float prev_phi = sample_state(S, add(ro, muls(rd, t_near))).r;
    float prev_t = t_near;
    for (int i = 0; i < n_steps; ++i) {
        if (t > t_far || alpha > 0.995f) break;
        const V3 p = add(ro, muls(rd, t));
        const S4 s = sample_state(S, p);
        const float d_prev = prev_phi - iso, d_cur = s.r - iso;
        if (show_iso && d_prev != 0.0f && d_prev * d_cur <= 0.0f) {
            const float denom = s.r - prev_phi;
            const float frac = std::fabs(denom) > 1e-6f ? (iso - prev_phi) / denom : 0.5f;
            const float t_hit = mixf(prev_t, t, clampf(frac, 0.0f, 1.0f));
            const V3 p_hit = add(ro, muls(rd, t_hit));
            const V3 g = grad_phi(S, p_hit);
            const V3 n = normalize(sub(v3(-g.x, -g.y, -g.z), muls(rd, 1e-6f)));
            const S4 hs = sample_state(S, p_hit);
            const float drive = clampf(hs.g * 4.0f, 0.0f, 1.0f);
            const V3 base = mixv(v3(0.16f, 0.62f, 0.72f), v3(1.00f, 0.72f, 0.28f), drive);
            const V3 l1 = normalize(v3(0.55f, 0.72f, 0.42f));
            const V3 l2 = normalize(v3(-0.5f, -0.2f, -0.7f));
            const V3 view = v3(-rd.x, -rd.y, -rd.z);
            const float diff = std::fmax(dotv(n, l1), 0.0f) + 0.35f * std::fmax(dotv(n, l2), 0.0f);
            const V3 h = normalize(add(add(l1, view), v3(0.0f, 1e-5f, 0.0f)));
            const float spec = std::pow(std::fmax(dotv(n, h), 0.0f), 48.0f);
            const float fres = std::pow(1.0f - std::fabs(dotv(n, view)), 3.0f);
            const V3 shaded = add(add(muls(base, 0.22f + 0.85f * diff), muls(muls(v3(1.0f, 1.0f, 1.0f), spec), 0.35f)),
                                  muls(muls(base, fres), 0.9f));
            const float a = wall_opacity;
            accum = add(accum, muls(muls(shaded, 1.0f - alpha), a));
            alpha = alpha + (1.0f - alpha) * a;
        }
        if (show_vol) {
            // map_field (volume.wgsl:156-186)
            float mt, intensity;
            int ramp_id = (int)S.vm(4, 3);
            if (mode == 0) {
                const float x = clampf(s.r, 0.0f, 1.0f);
                mt = x; intensity = x * gain;
            } else if (mode == 1) {
                const float x = clampf(s.g * gain, 0.0f, 1.0f);
                mt = x; intensity = x;
            } else if (mode == 2) {
                const float x = clampf(s.b * gain, -1.0f, 1.0f);
                mt = 0.5f + 0.5f * x; intensity = std::fabs(x); ramp_id = 2;
            } else {
                const float x = clampf(s.a * gain, 0.0f, 1.0f);
                mt = x; intensity = x;
            }
            if (intensity > 1e-4f) {
                const float weight = intensity * intensity;
                const float a = 1.0f - std::exp(-absorption * weight * ds);
                const V3 emit = muls(ramp(mt, ramp_id), exposure);
                accum = add(accum, muls(muls(emit, 1.0f - alpha), a));
                alpha = alpha + (1.0f - alpha) * a;
            }
        }
        prev_phi = s.r;
        prev_t = t;
        t = t + ds;
    }
    color = add(muls(color, 1.0f - alpha), accum);
    color = divv(color, add(v3(1.0f, 1.0f, 1.0f), color));
    color = v3(std::pow(std::fmax(color.x, 0.0f), 1.0f / 2.2f), std::pow(std::fmax(color.y, 0.0f), 1.0f / 2.2f),
               std::pow(std::fmax(color.z, 0.0f), 1.0f / 2.2f));
    emit_raw(color);
}

}  // namespace vol

extern "C" {

// Worker threads of the OpenMP loops that follow (n >= 1); returns the count
// now in force.
int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;  // built without OpenMP: the loops run on the calling thread
#endif
}

// params: 24 words of the SimParams block; events: 5 floats each
// (time, x, y, z, r0), already sorted by time.
void *oracle_create(int precision, const float *params, const float *events, int nev) {
    Handle *h = new Handle();
    h->precision = precision;
    if (precision == 1) h->ld = make_lattice<double>(params, events, nev);
    else h->lf = make_lattice<float>(params, events, nev);
    return h;
}

void oracle_destroy(void *hp) {
    Handle *h = (Handle *)hp;
    delete h->lf;
    delete h->ld;
    delete h;
}

// headless::advance's stepping (headless.rs:126-131): n x (nucleate due
// events, one RK3 step).  Returns the number of steps that stamped bubbles.
int oracle_advance(void *hp, long long nsteps) {
    Handle *h = (Handle *)hp;
    int stamped = 0;
    for (long long k = 0; k < nsteps; ++k) {
        if (h->ld) {
            stamped += h->ld->record_nucleation();
            h->ld->record_step();
        } else {
            stamped += h->lf->record_nucleation();
            h->lf->record_step();
        }
    }
    return stamped;
}

// one record_steps(n) (sim.rs:829-841): nucleate once, then n steps
int oracle_record_steps(void *hp, long long nsteps) {
    Handle *h = (Handle *)hp;
    int stamped = h->ld ? h->ld->record_nucleation() : h->lf->record_nucleation();
    for (long long k = 0; k < nsteps; ++k) {
        if (h->ld) h->ld->record_step();
        else h->lf->record_step();
    }
    return stamped;
}

int oracle_remaining(void *hp) {
    Handle *h = (Handle *)hp;
    return h->ld ? (int)(h->ld->schedule.size() - h->ld->next_event) : (int)(h->lf->schedule.size() - h->lf->next_event);
}

float oracle_time(void *hp) {
    Handle *h = (Handle *)hp;
    return h->ld ? h->ld->time : h->lf->time;
}

// state of slot cur; field/fluid buffers in the lattice precision
void oracle_get_state(void *hp, void *field, void *fluid) {
    Handle *h = (Handle *)hp;
    if (h->ld) {
        std::memcpy(field, h->ld->field[h->ld->cur].data(), h->ld->field[0].size() * 8);
        std::memcpy(fluid, h->ld->fluid[h->ld->cur].data(), h->ld->fluid[0].size() * 8);
    } else {
        std::memcpy(field, h->lf->field[h->lf->cur].data(), h->lf->field[0].size() * 4);
        std::memcpy(fluid, h->lf->fluid[h->lf->cur].data(), h->lf->fluid[0].size() * 4);
    }
}

void oracle_set_state(void *hp, const void *field, const void *fluid) {
    Handle *h = (Handle *)hp;
    if (h->ld) {
        std::memcpy(h->ld->field[h->ld->cur].data(), field, h->ld->field[0].size() * 8);
        std::memcpy(h->ld->fluid[h->ld->cur].data(), fluid, h->ld->fluid[0].size() * 8);
    } else {
        std::memcpy(h->lf->field[h->lf->cur].data(), field, h->lf->field[0].size() * 4);
        std::memcpy(h->lf->fluid[h->lf->cur].data(), fluid, h->lf->fluid[0].size() * 4);
    }
}

// sums of (e_total, broken, kinetic, E) and max|v| -> out[5]; contributions
// in the lattice precision, summed in long double
void oracle_diagnostics(void *hp, double *out) {
    Handle *h = (Handle *)hp;
    if (h->ld) diag_exact(*h->ld, out);
    else diag_exact(*h->lf, out);
}

// visualisation channels of slot cur, 4 floats per cell
void oracle_vis_pack(void *hp, float *out) {
    Handle *h = (Handle *)hp;
    if (h->ld) vis_pack(*h->ld, out);
    else vis_pack(*h->lf, out);
}

// reduce.wgsl's own f32 result (f32 lattices only)
void oracle_diagnostics_naive_f32(void *hp, float *sums, uint32_t *vmax_bits) {
    Handle *h = (Handle *)hp;
    if (h->lf) diag_naive_f32(*h->lf, sums, vmax_bits);
}

// --- elementary functions, f64 and f32, for unit tests -------------------
// params: the 24-word SimParams block
void oracle_prim_from_cons(int precision, const float *params, const double *u, double phi, double *q) {
    if (precision == 1) {
        const Prm<double> P = make_prm<double>(params);
        const P4<double> r = eos_prim_from_cons(P, P4<double>{u[0], u[1], u[2], u[3]}, phi);
        q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
    } else {
        const Prm<float> P = make_prm<float>(params);
        const P4<float> r = eos_prim_from_cons(P, P4<float>{(float)u[0], (float)u[1], (float)u[2], (float)u[3]}, (float)phi);
        q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
    }
}

void oracle_cons_from_prim(const float *params, const double *q, double *u) {
    const Prm<double> P = make_prm<double>(params);
    const P4<double> r = eos_cons_from_prim(P, P4<double>{q[0], q[1], q[2], q[3]}, 0.0);
    u[0] = r.x; u[1] = r.y; u[2] = r.z; u[3] = r.w;
}

void oracle_wave_speeds(const float *params, const double *q, int axis, double *lohi) {
    const Prm<double> P = make_prm<double>(params);
    eos_wave_speeds(P, P4<double>{q[0], q[1], q[2], q[3]}, 0.0, axis, lohi[0], lohi[1]);
}

void oracle_hll_flux(const float *params, const double *ql, const double *qr, int axis, double *f) {
    const Prm<double> P = make_prm<double>(params);
    const P4<double> r = hll_flux(P, P4<double>{ql[0], ql[1], ql[2], ql[3]}, 0.0, P4<double>{qr[0], qr[1], qr[2], qr[3]}, 0.0, axis);
    f[0] = r.x; f[1] = r.y; f[2] = r.z; f[3] = r.w;
}

void oracle_eos_flux(const double *q, const double *u, int axis, double *f) {
    const P4<double> r = eos_flux(P4<double>{q[0], q[1], q[2], q[3]}, P4<double>{u[0], u[1], u[2], u[3]}, axis);
    f[0] = r.x; f[1] = r.y; f[2] = r.z; f[3] = r.w;
}

double oracle_limit_slope(double dm, double dp) { return limit_slope(dm, dp); }
float oracle_limit_slope_f32(float dm, float dp) { return limit_slope(dm, dp); }

double oracle_tanh_f64(double x) { return o_tanh(x); }
float oracle_tanh_f32(float x) { return o_tanh(x); }

// volume.wgsl fs_main for every pixel of a width x height frame (row 0 at
// the top), from the 44-float view uniform and 4 half texels per cell
void oracle_render(const float *view, const uint16_t *texels, int nx, int ny, int nz, int width, int height,
                   uint8_t *rgba) {
    const vol::Scene S{view, texels, {nx, ny, nz}};
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) vol::pixel(S, x, y, width, height, rgba + 4 * ((size_t)y * width + x));
}

double oracle_potential(const float *params, double phi) { return potential(make_prm<double>(params), phi); }
double oracle_dpotential(const float *params, double phi) { return dpotential(make_prm<double>(params), phi); }

}  // extern "C"

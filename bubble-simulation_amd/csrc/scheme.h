// scheme.h -- the lattice scheme of one SSP-RK3 stage, per cell and per face.
//
// __host__ __device__ templates over the real type R (float / double) and an
// arithmetic policy M (ExactMath / FastMathF32).  Every function names the
// reference WGSL it restates; the operation order inside each expression is
// the reference's left-to-right order, so ExactMath evaluations are a fixed,
// reproducible sequence of IEEE operations.
//
// Swap points (docs/EOS.md:10-34): the Eos* and potential functions take the
// field value phi even where the bag model ignores it, so a phase-dependent
// equation of state changes no call site.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace bs {

template <class R> struct alignas(4 * sizeof(R)) V4 { R x, y, z, w; };
template <class R> struct alignas(2 * sizeof(R)) V2 { R x, y; };

// Device-side parameter block: the solver-relevant subset of SimParams
// (common.wgsl:25-58) in the lattice precision.  Built on the host from the
// f32 bs_sim_params; for R = double every value is the exact widening of the
// f32 one, so the FP64 scheme runs the same physics in wider arithmetic.
template <class R> struct Params {
    int nx, ny, nz, eos;  // eos: kEosBag / kEosGammaLaw
    R dx, inv_dx, dt, eta;
    R lambda, eps, phi_b, m2;
    R delta, a_rad, wall_width, p_floor;
    R e_floor, z_max, e_ref, t_ref;
    R eos_gamma;  // adiabatic index of the gamma-law equation of state
    R dof_ratio;  // a_broken / a_symmetric for the temperature (docs/EOS.md:65-84)
};
constexpr int kEosBag = 0;
constexpr int kEosGammaLaw = 1;

template <class R> struct Consts;
template <> struct Consts<float> {
    static constexpr float tiny = 1e-30f;     // eos.wgsl:109,121, step.wgsl:69
    static constexpr float den_floor = 1e-12f; // step.wgsl:88, 154
};
template <> struct Consts<double> {
    static constexpr double tiny = 1e-30;
    static constexpr double den_floor = 1e-12;
};

// ---------------------------------------------------------------------------
//  Arithmetic policies
// ---------------------------------------------------------------------------

// IEEE correctly rounded '/', sqrt; WGSL min/max selection rule
// ("e2 if e2 < e1 else e1" / "e2 if e1 < e2 else e1"); the literal sign-test
// minmod of common.wgsl:114-120.
template <class R> struct ExactMath {
    static constexpr bool fast = false;
    __host__ __device__ static inline R div(R a, R b) { return a / b; }
    __host__ __device__ static inline R sqrt(R x) { return std::sqrt(x); }
    __host__ __device__ static inline R min(R a, R b) { return b < a ? b : a; }
    __host__ __device__ static inline R max(R a, R b) { return a < b ? b : a; }
    __host__ __device__ static inline R min3(R a, R b, R c) { return min(min(a, b), c); }
    __host__ __device__ static inline R max3(R a, R b, R c) { return max(max(a, b), c); }
    __host__ __device__ static inline R abs(R a) { return std::fabs(a); }
    __host__ __device__ static inline R inv_sqrt(R x) { return R(1) / std::sqrt(x); }
    __host__ __device__ static inline R sign(R a) { return a > R(0) ? R(1) : (a < R(0) ? R(-1) : R(0)); }
    // minmod3 (common.wgsl:114-120)
    __host__ __device__ static inline R minmod3(R a, R b, R c) {
        const R sa = sign(a);
        if (sa != sign(b) || sa != sign(c)) return R(0);
        return sa * min(abs(a), min(abs(b), abs(c)));
    }
    // limit_slope (common.wgsl:125-127)
    __host__ __device__ static inline R limit_slope(R dm, R dp) {
        return minmod3(R(0.5) * (dm + dp), R(2) * dm, R(2) * dp);
    }
    // 0.5 * limit_slope: the half slope added to q at a face (step.wgsl:136-143)
    __host__ __device__ static inline R half_slope(R dm, R dp) { return R(0.5) * limit_slope(dm, dp); }
};

// Hardware reciprocal / square root (1 ulp each, so a*rcp(b) stays inside
// the 2.5 ulp WGSL allows for '/'), FMA contraction by the compiler, and the
// minmod written as a median: minmod(a, b) = med3(a, b, 0), so
// minmod3(c, 2a, 2b) = med3(c, 2 med3(a, b, 0), 0) -- the same value as the
// sign-test form for every non-NaN input (up to the sign of a zero result).
// Two-operand min / max are medians against -FLT_MAX / +FLT_MAX (equal to
// min / max for finite operands): v_med3 takes its operands as they are,
// where v_min / v_max in IEEE mode first quiet every operand the compiler
// cannot prove canonical (loaded values, kernel arguments, values merged
// after a branch) with an extra instruction each.
struct FastMathF32 {
    static constexpr bool fast = true;
    __device__ static inline float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static inline float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static inline float min(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, -3.402823466e38f); }
    __device__ static inline float max(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 3.402823466e38f); }
    // three-operand forms on computed values: one v_min3 / v_max3
    __device__ static inline float min3(float a, float b, float c) { return fminf(fminf(a, b), c); }
    __device__ static inline float max3(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }
    __device__ static inline float abs(float a) { return fabsf(a); }
    __device__ static inline float inv_sqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    __device__ static inline float minmod3(float a, float b, float c) {
        return __builtin_amdgcn_fmed3f(a, __builtin_amdgcn_fmed3f(b, c, 0.0f), 0.0f);
    }
    // MC limiter: minmod3(c, 2 dm, 2 dp) with med3(2 dm, 2 dp, 0) written as
    // 2 med3(dm, dp, 0) -- exact, since scaling by 2 commutes with the median
    __device__ static inline float limit_slope(float dm, float dp) {
        return __builtin_amdgcn_fmed3f(0.5f * (dm + dp), 2.0f * __builtin_amdgcn_fmed3f(dm, dp, 0.0f), 0.0f);
    }
    // 0.5 * limit_slope with the halving moved inside the median (scaling by
    // a power of two commutes with it): med3(0.25 (dm + dp), med3(dm, dp, 0),
    // 0), four instructions, the same value bit for bit
    __device__ static inline float half_slope(float dm, float dp) {
        return __builtin_amdgcn_fmed3f(0.25f * (dm + dp), __builtin_amdgcn_fmed3f(dm, dp, 0.0f), 0.0f);
    }
};

// ---------------------------------------------------------------------------
//  Small vector helpers (fixed evaluation order)
// ---------------------------------------------------------------------------
template <class R> __host__ __device__ inline R dot3(R ax, R ay, R az, R bx, R by, R bz) {
    return ax * bx + ay * by + az * bz;   // ((ax bx + ay by) + az bz)
}
template <class R> __host__ __device__ inline R comp(const V4<R> &v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// ---------------------------------------------------------------------------
//  Indexing (common.wgsl:80-95)
// ---------------------------------------------------------------------------
// Periodic wrap for any offset (the reference's wrap1 covers reach <= 2; a
// full modulo also serves tiles wider than a small box).
__host__ __device__ inline int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------
//  Potential (potential.wgsl:38-47) -- swap point #1
// ---------------------------------------------------------------------------
template <class R> __host__ __device__ inline R potential(R phi, const Params<R> &P) {
    const R p2 = phi * phi;
    return R(0.5) * P.m2 * p2 - P.delta * p2 * phi / R(3) + R(0.25) * P.lambda * p2 * p2;
}
template <class R> __host__ __device__ inline R dpotential(R phi, const Params<R> &P) {
    return phi * (P.m2 - P.delta * phi + P.lambda * phi * phi);
}

// ---------------------------------------------------------------------------
//  Equation of state (eos.wgsl) -- swap point #2
//
//  An equation of state is a policy class with the scalar functions of
//  docs/EOS.md:12-23 (each takes phi, the hook for a phase-dependent EOS);
//  the vector functions below (prim_from_cons, cons_from_prim, flux, wave
//  speeds) are written once on top of it, from the two structural facts the
//  solver relies on: E = w W^2 - p, Z = w W^2 v.
// ---------------------------------------------------------------------------
// The bag model, e = 3p (eos.wgsl:50-98): closed-form inversion.
struct BagEos {
    static constexpr int kind = kEosBag;
    // eos_sound_speed2 (eos.wgsl:50-54): 1/3 for e = 3p.
    template <class R> __host__ __device__ static inline R sound_speed2(R /*p*/, R /*phi*/, const Params<R> &) {
        return R(1) / R(3);
    }
    // eos_enthalpy_from_pressure (eos.wgsl:56-59): w = 4p.
    template <class M, class R> __host__ __device__ static inline R enthalpy(R p, R /*phi*/, const Params<R> &) {
        return R(4) * p;
    }
    // w / p (the FastMath flux folds it into one loop-invariant factor)
    template <class M, class R> __host__ __device__ static inline R enthalpy_per_pressure(R /*phi*/, const Params<R> &) {
        return R(4);
    }
    // eos_energy_from_pressure (eos.wgsl:61-63): e = 3p.
    template <class R> __host__ __device__ static inline R energy(R p, R /*phi*/, const Params<R> &) {
        return R(3) * p;
    }
    // eos_pressure_from_conserved (eos.wgsl:94-98): conjugate root of
    // 3p^2 + 2Ep - (E^2 - Z^2) = 0.
    template <class M, class R>
    __host__ __device__ static inline R pressure_from_conserved(R e_lab, R z2, R /*phi*/, const Params<R> &P) {
        // disc >= E^2 (4 - 3 z_max^2) > 0 once |Z| <= z_max E; FastMath drops
        // the clamp
        const R disc = M::fast ? R(4) * e_lab * e_lab - R(3) * z2 : M::max(R(4) * e_lab * e_lab - R(3) * z2, R(0));
        const R p = M::div(e_lab * e_lab - z2, e_lab + M::sqrt(disc));
        return M::max(p, P.p_floor);
    }
};

// A conformal gamma-law fluid, p = (Gamma - 1) e (docs/EOS.md:86-128, worked
// example 2): c_s^2 = Gamma - 1, w = Gamma p / (Gamma - 1).  Gamma = 4/3 is
// the bag model.  The inversion iterates, as the document describes: with
// s = E + p the rest-frame energy is e(p) = E - Z^2 / s, and Newton's method
// solves f(p) = p - (Gamma - 1) e(p) = 0 with f'(p) = 1 - (Gamma - 1) Z^2 / s^2.
// f is increasing and convex on p > -E, so from p0 = (Gamma - 1) E (the root
// at Z = 0, above every root) the iterates decrease monotonically onto the
// root; the count is fixed (kIters), never a tolerance loop, so one bad cell
// cannot stall a wave.
struct GammaLawEos {
    static constexpr int kind = kEosGammaLaw;
    static constexpr int kIters = 8;
    template <class R> __host__ __device__ static inline R sound_speed2(R /*p*/, R /*phi*/, const Params<R> &P) {
        return P.eos_gamma - R(1);
    }
    // FastMath: p times Gamma / (Gamma - 1), a factor formed once per
    // launch (loop-invariant), instead of a division per call
    template <class M, class R> __host__ __device__ static inline R enthalpy(R p, R /*phi*/, const Params<R> &P) {
        if constexpr (M::fast) return p * M::div(P.eos_gamma, P.eos_gamma - R(1));
        else return P.eos_gamma * p / (P.eos_gamma - R(1));
    }
    template <class M, class R> __host__ __device__ static inline R enthalpy_per_pressure(R /*phi*/, const Params<R> &P) {
        return M::div(P.eos_gamma, P.eos_gamma - R(1));
    }
    template <class R> __host__ __device__ static inline R energy(R p, R /*phi*/, const Params<R> &P) {
        return p / (P.eos_gamma - R(1));
    }
    template <class M, class R>
    __host__ __device__ static inline R pressure_from_conserved(R e_lab, R z2, R /*phi*/, const Params<R> &P) {
        const R gm1 = P.eos_gamma - R(1);
        R p = gm1 * e_lab;
        if constexpr (M::fast) {
            // the same Newton step with f and f' scaled by s^2, one
            // reciprocal instead of two: f / f' = s (p s - (Gamma - 1)(E s -
            // Z^2)) / (s^2 - (Gamma - 1) Z^2), the denominator > 0 since
            // (Gamma - 1) Z^2 < s^2
            const R gz2 = gm1 * z2;
            for (int it = 0; it < kIters; ++it) {
                const R s = e_lab + p;
                const R fs = p * s - gm1 * (e_lab * s - z2);
                p = p - (s * fs) * M::div(R(1), s * s - gz2);
            }
        } else {
            for (int it = 0; it < kIters; ++it) {
                const R s = e_lab + p;
                const R rs = M::div(R(1), s);
                const R zs = z2 * rs;                        // Z^2 / s
                const R f = p - gm1 * (e_lab - zs);
                const R df = R(1) - gm1 * (zs * rs);         // 1 - (Gamma - 1) Z^2 / s^2
                p = p - M::div(f, df);
            }
        }
        return M::max(p, P.p_floor);
    }
};

// eos_temperature (eos.wgsl:65-69): T = (3 max(p, p_floor) / a)^(1/4), with
// the radiation constant interpolated across the wall when the two phases
// have different degrees of freedom (docs/EOS.md:65-84, worked example 1):
// a(phi) = a (1 + (dof_ratio - 1) h^2 (3 - 2h)), h = clamp(phi / phi_b, 0, 1).
// dof_ratio = 1 is the reference's expression, operation for operation.
template <class M, class R> __host__ __device__ inline R eos_temperature(R p, R phi, const Params<R> &P) {
    R a = P.a_rad;
    if (P.dof_ratio != R(1)) {
        const R h = M::min(M::max(phi / P.phi_b, R(0)), R(1));
        const R smooth = h * h * (R(3) - R(2) * h);
        a = P.a_rad * (R(1) + (P.dof_ratio - R(1)) * smooth);
    }
    const R x = R(3) * M::max(p, P.p_floor) / a;
    return M::sqrt(M::sqrt(x));
}

// eos_prim_from_cons with eos_regularise inlined (eos.wgsl:102-123):
// (E, Z) -> (v, p).
template <class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> eos_prim_from_cons(V4<R> u, R phi, const Params<R> &P) {
    const R e_lab = M::max(u.x, P.e_floor);
    R zx = u.y, zy = u.z, zz = u.w;
    const R z2raw = dot3(zx, zy, zz, zx, zy, zz);
    const R zcap = P.z_max * e_lab;
    if constexpr (M::fast) {
        // The cap without a square root or a branch (zcap > 0): the rescale
        // factor is min(1, zcap / |Z|) (1 for Z = 0, where the reciprocal
        // square root is +inf), and |Z|^2 after it is min(|Z|^2, zcap^2).
        // The factor joins the reciprocal of E + p, so v = Z (f / (E + p)):
        // one product per component.
        const R f = M::min(R(1), zcap * M::inv_sqrt(z2raw));
        const R z2 = M::min(z2raw, zcap * zcap);
        const R p = E::template pressure_from_conserved<M>(e_lab, z2, phi, P);
        const R rs = f * M::div(R(1), e_lab + p);  // E + p >= e_floor + p_floor > 0
        return V4<R>{zx * rs, zy * rs, zz * rs, p};
    } else {
        if (M::sqrt(z2raw) > zcap) {
            const R zlen = M::sqrt(z2raw);
            const R f = M::div(zcap, M::max(zlen, Consts<R>::tiny));
            zx = zx * f; zy = zy * f; zz = zz * f;
        }
        const R z2 = dot3(zx, zy, zz, zx, zy, zz);
        const R p = E::template pressure_from_conserved<M>(e_lab, z2, phi, P);
        const R sm = M::max(e_lab + p, Consts<R>::tiny);
        return V4<R>{M::div(zx, sm), M::div(zy, sm), M::div(zz, sm), p};
    }
}

// eos_cons_from_prim (eos.wgsl:129-135).
template <class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> eos_cons_from_prim(const V4<R> &q, R phi, const Params<R> &P) {
    const R v2 = M::min(dot3(q.x, q.y, q.z, q.x, q.y, q.z), P.z_max * P.z_max);
    const R w_gam2 = M::div(E::template enthalpy<M>(q.w, phi, P), R(1) - v2);
    return V4<R>{w_gam2 - q.w, w_gam2 * q.x, w_gam2 * q.y, w_gam2 * q.z};
}

// eos_flux (eos.wgsl:140-152) along compile-time axis A.
template <int A, class R>
__host__ __device__ inline V4<R> eos_flux(const V4<R> &q, const V4<R> &u) {
    const R p = q.w;
    const R vd = A == 0 ? q.x : (A == 1 ? q.y : q.z);
    V4<R> f{(u.x + p) * vd, u.y * vd, u.z * vd, u.w * vd};
    if (A == 0) f.y = f.y + p;
    if (A == 1) f.z = f.z + p;
    if (A == 2) f.w = f.w + p;
    return f;
}

// eos_wave_speeds (eos.wgsl:164-176): relativistic (lambda-, lambda+).
template <int A, class M, class R, class E = BagEos>
__host__ __device__ inline void eos_wave_speeds(const V4<R> &q, R phi, const Params<R> &P, R &lm, R &lp) {
    const R v2 = M::min(dot3(q.x, q.y, q.z, q.x, q.y, q.z), P.z_max * P.z_max);
    const R vd = A == 0 ? q.x : (A == 1 ? q.y : q.z);
    const R cs2 = E::sound_speed2(q.w, phi, P);
    const R cs = M::sqrt(cs2);
    const R denom = R(1) - v2 * cs2;
    const R rad = M::max((R(1) - v2) * (R(1) - vd * vd - cs2 * (v2 - vd * vd)), R(0));
    const R a = M::div(vd * (R(1) - cs2), denom);
    const R b = M::div(cs * M::sqrt(rad), denom);
    lm = a - b;
    lp = a + b;
}

// sanitise (step.wgsl:63-72): floor p, cap |v| at z_max.
template <class M, class R>
__host__ __device__ inline V4<R> sanitise(const V4<R> &q, const Params<R> &P) {
    const R p = M::max(q.w, P.p_floor);
    R vx = q.x, vy = q.y, vz = q.z;
    const R v2 = dot3(vx, vy, vz, vx, vy, vz);
    const R vmax2 = P.z_max * P.z_max;
    if (v2 > vmax2) {
        const R f = M::sqrt(M::div(vmax2, M::max(v2, Consts<R>::tiny)));
        vx = vx * f; vy = vy * f; vz = vz * f;
    }
    return V4<R>{vx, vy, vz, p};
}

// sanitise that also returns |v|^2 of its result: the dot product it already
// formed when nothing is clamped (the value a recomputation gives), else the
// recomputed one (ExactMath) or vmax^2 itself (FastMath).
template <class M, class R>
__host__ __device__ inline V4<R> sanitise_v2(const V4<R> &q, const Params<R> &P, R &v2out) {
    const R p = M::max(q.w, P.p_floor);
    R vx = q.x, vy = q.y, vz = q.z;
    const R v2 = dot3(vx, vy, vz, vx, vy, vz);
    const R vmax2 = P.z_max * P.z_max;
    v2out = v2;
    if (v2 > vmax2) {
        const R f = M::sqrt(M::div(vmax2, M::max(v2, Consts<R>::tiny)));
        vx = vx * f; vy = vy * f; vz = vz * f;
        if constexpr (M::fast) v2out = vmax2;
        else v2out = dot3(vx, vy, vz, vx, vy, vz);
    }
    return V4<R>{vx, vy, vz, p};
}

// sanitise_v2 of the two states of one face.  FastMath tests both against
// the cap in one branch, which no wave takes unless one of its states is
// clamped; inside it each state is scaled by its own factor (1 when it is
// not clamped), so every velocity is the one sanitise_v2 gives.  FastMath
// leaves out the pressure floor: the states are q +- h with the MC half
// slope |h| <= min(|q - q_-|, |q_+ - q|) of equal sign, so each lies between
// its cell and the neighbour across the face, both recovered with p >=
// p_floor; rounding can leave a few ulps below it, still positive, which is
// all the flux needs.
template <class M, class R>
__host__ __device__ inline void sanitise_pair_v2(V4<R> &a, R &v2a, V4<R> &b, R &v2b, const Params<R> &P) {
    if constexpr (M::fast) {
        const R vmax2 = P.z_max * P.z_max;
        v2a = dot3(a.x, a.y, a.z, a.x, a.y, a.z);
        v2b = dot3(b.x, b.y, b.z, b.x, b.y, b.z);
        const bool ca = v2a > vmax2, cb = v2b > vmax2;
        if (ca | cb) {
            const R fa = ca ? M::sqrt(M::div(vmax2, M::max(v2a, Consts<R>::tiny))) : R(1);
            const R fb = cb ? M::sqrt(M::div(vmax2, M::max(v2b, Consts<R>::tiny))) : R(1);
            a.x = a.x * fa; a.y = a.y * fa; a.z = a.z * fa;
            b.x = b.x * fb; b.y = b.y * fb; b.z = b.z * fb;
            if (ca) v2a = vmax2;
            if (cb) v2b = vmax2;
        }
    } else {
        a = sanitise_v2<M>(a, P, v2a);
        b = sanitise_v2<M>(b, P, v2b);
    }
}

// |v|^2 capped at z_max^2 (eos.wgsl:130, 166).  The FastMath sanitise_v2
// already returns a value <= z_max^2 (z_max^2 itself when it clamps), so the
// cap is the identity there and is left out.
template <class M, class R> __host__ __device__ inline R capped_v2(R v2raw, const Params<R> &P) {
    if constexpr (M::fast) return v2raw;
    else return M::min(v2raw, P.z_max * P.z_max);
}

// eos_cons_from_prim with |v|^2 of q given (from sanitise_v2)
template <class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> eos_cons_from_prim_v2(const V4<R> &q, R v2raw, R phi, const Params<R> &P,
                                                       R &w_gam2) {
    const R v2 = capped_v2<M>(v2raw, P);
    w_gam2 = M::div(E::template enthalpy<M>(q.w, phi, P), R(1) - v2);
    return V4<R>{w_gam2 - q.w, w_gam2 * q.x, w_gam2 * q.y, w_gam2 * q.z};
}

// eos_flux given w gamma^2 of the state as well: E + p is w gamma^2 itself,
// which FastMath uses in place of the reference's (E + p) sum
template <int A, class M, class R>
__host__ __device__ inline V4<R> eos_flux_w(const V4<R> &q, const V4<R> &u, R w_gam2) {
    V4<R> f = eos_flux<A>(q, u);
    if constexpr (M::fast) f.x = w_gam2 * (A == 0 ? q.x : (A == 1 ? q.y : q.z));
    return f;
}

// eos_wave_speeds with |v|^2 of q given (from sanitise_v2)
template <int A, class M, class R, class E = BagEos>
__host__ __device__ inline void eos_wave_speeds_v2(const V4<R> &q, R v2raw, R phi, const Params<R> &P, R &lm,
                                                   R &lp) {
    const R v2 = capped_v2<M>(v2raw, P);
    const R vd = A == 0 ? q.x : (A == 1 ? q.y : q.z);
    const R cs2 = E::sound_speed2(q.w, phi, P);
    const R denom = R(1) - v2 * cs2;
    const R cs = M::sqrt(cs2);
    const R rad = M::max((R(1) - v2) * (R(1) - vd * vd - cs2 * (v2 - vd * vd)), R(0));
    const R a = M::div(vd * (R(1) - cs2), denom);
    const R b = M::div(cs * M::sqrt(rad), denom);
    lm = a - b;
    lp = a + b;
}

// FastMath: what the HLL flux needs of one face state, given its velocity
// component v_d normal to the face -- w gamma^2 and the wave speeds (eos_cons_from_prim, eos_wave_speeds: eos.wgsl:129-135,
// 164-176) -- with the constants of the equation of state taken into the
// factors.  With g = 1 / (1 - c^2), o = g c^2 (1 - v^2) and
// D = g (1 - v^2 c^2) = g - g c^2 v^2, and since
// 1 - v_d^2 - c^2 (v^2 - v_d^2) = (1 - c^2) (D - v_d^2):
//   lambda+- = (v_d +- sqrt(o (D - v_d^2))) / D,
//   w gamma^2 = (w / p) p / (1 - v^2) = ((w / p) g c^2) p / o.
// D - v_d^2 >= g (1 - v^2) > 0 (v_d^2 <= |v|^2 <= z_max^2 < 1 after
// sanitise), so the radicand needs no clamp.  g, g c^2 and the enthalpy
// factor depend on the EOS constants only (1.5, 0.5 and 4 x 0.5 for the bag
// model), so they are formed once per launch.
template <class M, class R, class E = BagEos>
__host__ __device__ inline void flux_side_fast(const V4<R> &q, R vd, R v2, R phi, const Params<R> &P, R &w_gam2,
                                               R &lm, R &lp) {
    static_assert(M::fast, "the factored side terms are a FastMath form");
    const R cs2 = E::sound_speed2(q.w, phi, P);
    const R g = M::div(R(1), R(1) - cs2);
    const R gc2 = g * cs2;
    const R o = gc2 - gc2 * v2;
    const R d = g - gc2 * v2;
    const R rd = M::div(R(1), d);
    const R a = vd * rd;
    const R b = M::sqrt(o * (d - vd * vd)) * rd;
    lm = a - b;
    lp = a + b;
    const R hk = E::template enthalpy_per_pressure<M>(phi, P) * gc2;
    w_gam2 = (q.w * hk) * M::div(R(1), o);
}

// FastMath HLL flux along `axis` (0, 1, 2): a constant where the face
// direction is known at compile time, a per-lane value where one code path
// evaluates faces of two directions (the normal velocity and the pressure
// term then take one select each).  The flux is collected by term:
// F = cmax A - cmin B with al = cmax t_L, be = cmin t_R, and
// cmax s- = cmin s+ = s- s+ / (s+ - s-) multiplying both pressures.
template <class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> hll_flux_fast(const V4<R> &ql, R v2l, R phil, const V4<R> &qr, R v2r, R phir,
                                               const Params<R> &P, int axis) {
    static_assert(M::fast, "the factored flux is a FastMath form");
    const R vdl = axis == 0 ? ql.x : (axis == 1 ? ql.y : ql.z);
    const R vdr = axis == 0 ? qr.x : (axis == 1 ? qr.y : qr.z);
    R wl, llm, llp, wr, lrm, lrp;
    flux_side_fast<M, R, E>(ql, vdl, v2l, phil, P, wl, llm, llp);
    flux_side_fast<M, R, E>(qr, vdr, v2r, phir, P, wr, lrm, lrp);
    const R s_min = M::min3(llm, lrm, R(0));
    const R s_max = M::max3(llp, lrp, R(0));
    // s+ - s- >= lambda+ - lambda- of either state = 2 sqrt(o (D - v_d^2)) / D,
    // at least ~1e-7 for any sanitised state (|v| <= z_max, 0 < c^2 < 1),
    // so the reference's 1e-12 floor on the denominator never binds here
    const R rden = M::div(R(1), s_max - s_min);
    const R cmax = s_max * rden, cmin = s_min * rden;
    const R al = (wl * (vdl - s_min)) * cmax;
    const R be = (wr * (vdr - s_max)) * cmin;
    const R mu = s_min * cmax;
    const R kp = cmax * ql.w - cmin * qr.w;
    return V4<R>{(al - be) + mu * (ql.w - qr.w),
                 al * ql.x + ((axis == 0 ? kp : R(0)) - be * qr.x),
                 al * ql.y + ((axis == 1 ? kp : R(0)) - be * qr.y),
                 al * ql.z + ((axis == 2 ? kp : R(0)) - be * qr.z)};
}

// hll_flux (step.wgsl:77-90) along axis A, with |v|^2 of both states given
// (the reference recomputes dot(v, v) in eos_cons_from_prim and
// eos_wave_speeds; the value is the same).
//
// FastMath evaluates the same flux factored by side.  With w = w gamma^2,
// U = (w - p, w v), F = (w v_d, w v v_d + p e_d):
//   F_HLL = (s+ F_L - s- F_R + s- s+ (U_R - U_L)) / (s+ - s-)
//         = (s+ A - s- B) / (s+ - s-),   A = F_L - s- U_L,  B = F_R - s+ U_R,
// and with t_L = w_L (v_d,L - s-), t_R = w_R (v_d,R - s+):
//   A = (t_L + s- p_L, v_L t_L + p_L e_d),  B = (t_R + s+ p_R, v_R t_R + p_R e_d),
// one product per component and side instead of forming U and F.
template <int A, class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> hll_flux_v2(const V4<R> &ql, R v2l, R phil, const V4<R> &qr, R v2r, R phir,
                                             const Params<R> &P) {
    if constexpr (M::fast) {
        return hll_flux_fast<M, R, E>(ql, v2l, phil, qr, v2r, phir, P, A);
    } else {
        R wl, wr;
        const V4<R> ul = eos_cons_from_prim_v2<M, R, E>(ql, v2l, phil, P, wl);
        const V4<R> ur = eos_cons_from_prim_v2<M, R, E>(qr, v2r, phir, P, wr);
        const V4<R> fl = eos_flux_w<A, M>(ql, ul, wl);
        const V4<R> fr = eos_flux_w<A, M>(qr, ur, wr);
        R llm, llp, lrm, lrp;
        eos_wave_speeds_v2<A, M, R, E>(ql, v2l, phil, P, llm, llp);
        eos_wave_speeds_v2<A, M, R, E>(qr, v2r, phir, P, lrm, lrp);
        const R s_min = M::min3(llm, lrm, R(0));
        const R s_max = M::max3(llp, lrp, R(0));
        const R den = M::max(s_max - s_min, Consts<R>::den_floor);
        const R smm = s_min * s_max;
        return V4<R>{M::div(s_max * fl.x - s_min * fr.x + smm * (ur.x - ul.x), den),
                     M::div(s_max * fl.y - s_min * fr.y + smm * (ur.y - ul.y), den),
                     M::div(s_max * fl.z - s_min * fr.z + smm * (ur.z - ul.z), den),
                     M::div(s_max * fl.w - s_min * fr.w + smm * (ur.w - ul.w), den)};
    }
}

// hll_flux (step.wgsl:77-90) along axis A.
template <int A, class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> hll_flux(const V4<R> &ql, R phil, const V4<R> &qr, R phir,
                                          const Params<R> &P) {
    const V4<R> ul = eos_cons_from_prim<M, R, E>(ql, phil, P);
    const V4<R> ur = eos_cons_from_prim<M, R, E>(qr, phir, P);
    const V4<R> fl = eos_flux<A>(ql, ul);
    const V4<R> fr = eos_flux<A>(qr, ur);
    R llm, llp, lrm, lrp;
    eos_wave_speeds<A, M, R, E>(ql, phil, P, llm, llp);
    eos_wave_speeds<A, M, R, E>(qr, phir, P, lrm, lrp);
    const R s_min = M::min3(llm, lrm, R(0));
    const R s_max = M::max3(llp, lrp, R(0));
    const R den = M::max(s_max - s_min, Consts<R>::den_floor);
    const R smm = s_min * s_max;
    return V4<R>{M::div(s_max * fl.x - s_min * fr.x + smm * (ur.x - ul.x), den),
                 M::div(s_max * fl.y - s_min * fr.y + smm * (ur.y - ul.y), den),
                 M::div(s_max * fl.z - s_min * fr.z + smm * (ur.z - ul.z), den),
                 M::div(s_max * fl.w - s_min * fr.w + smm * (ur.w - ul.w), den)};
}

// limit_slope / limit_slope4 (common.wgsl:125-136): monotonised central.
template <class M, class R> __host__ __device__ inline R limit_slope(R dm, R dp) {
    return M::limit_slope(dm, dp);
}
template <class M, class R>
__host__ __device__ inline V4<R> limit_slope4(const V4<R> &dm, const V4<R> &dp) {
    return V4<R>{limit_slope<M>(dm.x, dp.x), limit_slope<M>(dm.y, dp.y),
                 limit_slope<M>(dm.z, dp.z), limit_slope<M>(dm.w, dp.w)};
}
// 0.5 * limit_slope4, per component
template <class M, class R>
__host__ __device__ inline V4<R> half_slope4(const V4<R> &dm, const V4<R> &dp) {
    return V4<R>{M::half_slope(dm.x, dp.x), M::half_slope(dm.y, dp.y), M::half_slope(dm.z, dp.z),
                 M::half_slope(dm.w, dp.w)};
}
template <class R> __host__ __device__ inline V4<R> sub4(const V4<R> &a, const V4<R> &b) {
    return V4<R>{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
}
// exchange of the x and y velocity slots of a primitive state, and of the
// x and y momentum slots of a flux (a y face evaluated as an x face)
template <class R> __host__ __device__ inline V4<R> swap_xy(const V4<R> &a) { return V4<R>{a.y, a.x, a.z, a.w}; }
template <class R> __host__ __device__ inline V4<R> swap_yz(const V4<R> &a) { return V4<R>{a.x, a.z, a.y, a.w}; }
// q + 0.5 s  and  q - 0.5 s  (step.wgsl:136-143), given h = 0.5 s (the
// product by 0.5 is exact, so q + h is the reference's q + 0.5 * s)
template <class R> __host__ __device__ inline V4<R> add4(const V4<R> &q, const V4<R> &h) {
    return V4<R>{q.x + h.x, q.y + h.y, q.z + h.z, q.w + h.w};
}

// The flux through the face between cells L and R along axis A, given the
// sanitised primitives of the four cells L-1, L, R, R+1 (step.wgsl:125-145:
// this is f_plus of cell L and f_minus of cell R; the reference evaluates it
// once from each side, with identical operands).
template <int A, class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> face_flux(const V4<R> &qlm, const V4<R> &ql, const V4<R> &qr,
                                           const V4<R> &qrp, R phil, R phir, const Params<R> &P) {
    const V4<R> d = sub4(qr, ql);
    const V4<R> hl = half_slope4<M>(sub4(ql, qlm), d);
    const V4<R> hr = half_slope4<M>(d, sub4(qrp, qr));
    R v2l, v2r;
    V4<R> a = add4(ql, hl), b = sub4(qr, hr);
    sanitise_pair_v2<M>(a, v2l, b, v2r, P);
    return hll_flux_v2<A, M, R, E>(a, v2l, phil, b, v2r, phir, P);
}
// face_flux under FastMath with the face direction as a per-lane value
// (0: x, 1: y), for a code path that evaluates faces of both directions
template <class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> face_flux_fast_xy(const V4<R> &qlm, const V4<R> &ql, const V4<R> &qr,
                                                   const V4<R> &qrp, R phil, R phir, const Params<R> &P, int axis) {
    const V4<R> d = sub4(qr, ql);
    const V4<R> hl = half_slope4<M>(sub4(ql, qlm), d);
    const V4<R> hr = half_slope4<M>(d, sub4(qrp, qr));
    R v2l, v2r;
    V4<R> a = add4(ql, hl), b = sub4(qr, hr);
    sanitise_pair_v2<M>(a, v2l, b, v2r, P);
    return hll_flux_fast<M, R, E>(a, v2l, phil, b, v2r, phir, P, axis);
}
// Same, with the two half slopes (0.5 * limit_slope) already known.
template <int A, class M, class R, class E = BagEos>
__host__ __device__ inline V4<R> face_flux_h(const V4<R> &ql, const V4<R> &hl, const V4<R> &qr,
                                             const V4<R> &hr, R phil, R phir, const Params<R> &P) {
    R v2l, v2r;
    V4<R> a = add4(ql, hl), b = sub4(qr, hr);
    sanitise_pair_v2<M>(a, v2l, b, v2r, P);
    return hll_flux_v2<A, M, R, E>(a, v2l, phil, b, v2r, phir, P);
}

// ---------------------------------------------------------------------------
//  Per-cell source terms and the RK stage combination (step.wgsl:150-173)
// ---------------------------------------------------------------------------
// Given phi0, pi0, the sanitised primitive q0, the Laplacian sum, the
// gradient, the flux divergence (already scaled by inv_dx), and U_in / U^n of
// the cell, produce U_out.
// S0: the first stage (a0 = 0, a1 = 1), where FastMath takes the combination
// as U_in + dt RHS itself; the exact policy keeps 0 * U^n + 1 * (...), whose
// zero signs follow the reference.
// RAW (FastMath only): flux_div holds the unscaled sum of the three flux
// differences and gx, gy, gz the unscaled central differences of phi (2 dx
// times the gradient); 1/dx joins dt in one factor, as 1/dx^2 does for the
// Laplacian, and 1/(2 dx) joins the drag: U_in + dt (drag terms) - (dt/dx)
// sum, each component two multiply-adds.
template <class M, class R, bool S0 = false, bool RAW = false>
__host__ __device__ inline void stage_update(R phi0, R pi0, const V4<R> &q0, R lap_sum, R gx, R gy, R gz,
                                             const V4<R> &flux_div, const V4<R> &fin,
                                             R phin, R pin, const V4<R> &fn,
                                             R a0, R a1, R dt, const Params<R> &P,
                                             R &phi_out, R &pi_out, V4<R> &f_out) {
    const R d_phi = pi0;
    R up_phi, up_pi;
    V4<R> up;
    if constexpr (RAW) {
        static_assert(M::fast, "the unscaled divergence is a FastMath form");
        // |v| < z_max for a recovered primitive (the recovery caps |Z| at
        // z_max E and divides by E + p > E), so 1 - v^2 >= 1 - z_max^2 needs
        // no floor here
        const R hdx = R(0.5) * P.inv_dx;
        const R gam = M::inv_sqrt(R(1) - dot3(q0.x, q0.y, q0.z, q0.x, q0.y, q0.z));
        const R dphi_proper = gam * (pi0 + hdx * dot3(q0.x, q0.y, q0.z, gx, gy, gz));
        const R drag = P.eta * dphi_proper;
        const R dtd = dt * drag;
        const R dtg = dtd * hdx;
        const R dtx = dt * P.inv_dx;
        up_phi = phi0 + dt * d_phi;
        up_pi = (pi0 - dt * (dpotential(phi0, P) + drag)) + (dt * (P.inv_dx * P.inv_dx)) * lap_sum;
        up = V4<R>{(fin.x + dtd * pi0) - dtx * flux_div.x, (fin.y - dtg * gx) - dtx * flux_div.y,
                   (fin.z - dtg * gy) - dtx * flux_div.z, (fin.w - dtg * gz) - dtx * flux_div.w};
    } else {
        const R laplacian = M::fast ? lap_sum * (P.inv_dx * P.inv_dx) : lap_sum * P.inv_dx * P.inv_dx;
        const R gam = M::inv_sqrt(M::max(R(1) - dot3(q0.x, q0.y, q0.z, q0.x, q0.y, q0.z), Consts<R>::den_floor));
        const R dphi_proper = gam * (pi0 + dot3(q0.x, q0.y, q0.z, gx, gy, gz));
        const R drag = P.eta * dphi_proper;
        const R d_pi = laplacian - dpotential(phi0, P) - drag;
        V4<R> d{-flux_div.x, -flux_div.y, -flux_div.z, -flux_div.w};
        d.x = d.x + drag * pi0;
        d.y = d.y - drag * gx;
        d.z = d.z - drag * gy;
        d.w = d.w - drag * gz;
        up_phi = phi0 + dt * d_phi;
        up_pi = pi0 + dt * d_pi;
        up = V4<R>{fin.x + dt * d.x, fin.y + dt * d.y, fin.z + dt * d.z, fin.w + dt * d.w};
    }
    if constexpr (M::fast && S0) {
        phi_out = up_phi;
        pi_out = up_pi;
        f_out = up;
    } else {
        // Two rounded products and a sum, as the reference writes the
        // combination, also under FastMath: fused into multiply-adds (either
        // product exact), or rewritten as U^n + a1 (up - U^n), the third
        // stage's f32 weights (1/3 and 2/3, rounded up, sum 1 + 3e-8) leave
        // a systematic energy error -- measured over 1500 steps at 128^3 on
        // three workloads: drift 1e-5 below f64's and 4x the reference's f32
        // field error with fusion, 4e-5 with the rewrite, against 2-3e-6 for
        // the reference's own order, which this keeps.
#pragma clang fp contract(off)
        phi_out = a0 * phin + a1 * up_phi;
        pi_out = a0 * pin + a1 * up_pi;
        f_out = V4<R>{a0 * fn.x + a1 * up.x, a0 * fn.y + a1 * up.y, a0 * fn.z + a1 * up.z, a0 * fn.w + a1 * up.w};
    }
}

// ---------------------------------------------------------------------------
//  Nucleation profile (nucleate.wgsl:52-85)
// ---------------------------------------------------------------------------
// tanh from expm1, both written out here so that every build (GPU f32/f64,
// host) evaluates the identical operation sequence:
//   expm1(a) for a >= 0: a = k ln2 + r, |r| <= ln2/2 (Cody-Waite split of
//   ln2), expm1(r) by its Taylor series (degree 9 for float, 16 for double),
//   expm1(a) = 2^k (expm1(r) + 1) - 1;
//   tanh(x) = sign(x) * e / (e + 2) with e = expm1(2|x|).
template <class R> struct TanhK;
template <> struct TanhK<float> {
    static constexpr int terms = 9;
    static constexpr float ln2_hi = 0.693145751953125f;        // 12-bit head of ln 2
    static constexpr float ln2_lo = 1.428606765330187045e-06f; // ln 2 - ln2_hi
    static constexpr float inv_ln2 = 1.44269504088896341f;
    static constexpr float big = 9.0f;                          // tanh(x) == 1 in f32 beyond this
};
template <> struct TanhK<double> {
    static constexpr int terms = 16;
    static constexpr double ln2_hi = 6.93147180369123816490e-01;
    static constexpr double ln2_lo = 1.90821492927058770002e-10;
    static constexpr double inv_ln2 = 1.44269504088896338700e+00;
    static constexpr double big = 19.5;
};

template <class R> __host__ __device__ inline R ldexp_int(R x, int k) {
    // multiply by 2^k with k in [-60, 60], as exact power-of-two factors
    R f = R(1);
    R two = k >= 0 ? R(2) : R(0.5);
    int n = k >= 0 ? k : -k;
    while (n > 0) { f = f * two; --n; }
    return x * f;
}

template <class R> __host__ __device__ inline R bs_expm1_pos(R a) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    // a in [0, 2*big]
    const R kf = std::floor(a * TanhK<R>::inv_ln2 + R(0.5));
    const int k = (int)kf;
    const R r = (a - kf * TanhK<R>::ln2_hi) - kf * TanhK<R>::ln2_lo;
    // Horner for r + r^2/2! + ... + r^n/n!
    R t = R(1);
    for (int i = TanhK<R>::terms; i >= 2; --i) t = R(1) + t * r / R(i);
    const R em1r = r * t;
    if (k == 0) return em1r;
    return ldexp_int(em1r + R(1), k) - R(1);
}

template <class R> __host__ __device__ inline R bs_tanh(R x) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    const R ax = std::fabs(x);
    R t;
    if (ax >= TanhK<R>::big) {
        t = R(1);
    } else {
        const R e = bs_expm1_pos(R(2) * ax);
        t = e / (e + R(2));
    }
    return x < R(0) ? -t : t;
}

// min_image (nucleate.wgsl:52-57)
template <class R> __host__ __device__ inline R min_image(R d, R n) {
    R r = d;
    if (r > R(0.5) * n) r = r - n;
    if (r < R(-0.5) * n) r = r + n;
    return r;
}

// One bubble's profile value at cell (x, y, z) (nucleate.wgsl:271-281).
template <class R>
__host__ __device__ inline R bubble_profile(int x, int y, int z, R cx, R cy, R cz, R r0, const Params<R> &P) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    const R dx_ = min_image(R(x) - cx, R(P.nx));
    const R dy_ = min_image(R(y) - cy, R(P.ny));
    const R dz_ = min_image(R(z) - cz, R(P.nz));
    const R r = std::sqrt(dot3(dx_, dy_, dz_, dx_, dy_, dz_)) * P.dx;
    return R(0.5) * P.phi_b * (R(1) - bs_tanh((r - r0 * P.dx) / P.wall_width));
}

// ---------------------------------------------------------------------------
//  Diagnostics contribution of one cell (reduce.wgsl:56-79, 96)
// ---------------------------------------------------------------------------
template <class M, class R> struct CellDiag {
    R e_total, broken, kinetic, e_lab, vabs;
};
template <class M, class R, class E = BagEos>
__host__ __device__ inline CellDiag<M, R> cell_contribution(R phi, R pi, R phixm, R phixp, R phiym, R phiyp,
                                                            R phizm, R phizp, const V4<R> &u,
                                                            const Params<R> &P) {
    R grad2 = R(0);
    {
        const R d = R(0.5) * (phixp - phixm) * P.inv_dx;
        grad2 = grad2 + d * d;
    }
    {
        const R d = R(0.5) * (phiyp - phiym) * P.inv_dx;
        grad2 = grad2 + d * d;
    }
    {
        const R d = R(0.5) * (phizp - phizm) * P.inv_dx;
        grad2 = grad2 + d * d;
    }
    const R e_lab = M::max(u.x, P.e_floor);
    const V4<R> q = eos_prim_from_cons<M, R, E>(u, phi, P);
    const R v2 = dot3(q.x, q.y, q.z, q.x, q.y, q.z);
    CellDiag<M, R> c;
    c.e_total = R(0.5) * pi * pi + R(0.5) * grad2 + potential(phi, P) + e_lab;
    c.broken = phi > R(0.5) * P.phi_b ? R(1) : R(0);
    c.kinetic = (e_lab + q.w) * v2;
    c.e_lab = e_lab;
    c.vabs = M::sqrt(v2);
    return c;
}

}  // namespace bs

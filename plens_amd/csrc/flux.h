// Device-side gas model and flux functions of the DGSEM hot path (gfx950).
//
// Each function computes the same quantity as the NavierStokes member cited beside it; the
// operation order is free where rounding is the only difference (e.g. the Chandrashekhar matrix
// dissipation is applied as K*(Lambda*(K^-1*du)) instead of forming K*Lambda*K^-1).
// Branch conditions (HLLC pressure estimates, log-mean series switch, AUSM Mach splits) follow the
// reference so that the same branch is taken.
//
// The flux functions are templates on the arithmetic type T: double for the plain FP64 path, float
// for the FP32-bulk path. For float, divisions use the hardware reciprocal (v_rcp_f32) and
// sqrt/log/pow the native single-precision instructions; for double every expression is the plain
// IEEE one.
#pragma once
#include <hip/hip_runtime.h>

namespace plens_dev {

struct Gas {
    double gma, R, cv, mu0, T0, S, Pr;
    double icv, iR, kfac;  // 1/c_v, 1/R, R gamma / ((gamma - 1) Pr)
    int inviscid;
};

// ---- arithmetic helpers (double: reciprocal-based division within one rounding of IEEE, IEEE
//      square root and logarithm; float: native single-precision instructions) -------------------
// 1/x from the hardware reciprocal refined by two Newton steps (full double precision; the quotient
// a*rcp(b) differs from a correctly rounded a/b by at most one extra rounding)
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double dv(double a, double b) { return a * rcp64(b); }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sq(double a) { return sqrt(a); }
__device__ __forceinline__ float sq(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double lg(double a) { return log(a); }
__device__ __forceinline__ float lg(float a) { return __logf(a); }
__device__ __forceinline__ double pw(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float pw(float a, float b) { return __powf(a, b); }
__device__ __forceinline__ double ab(double a) { return fabs(a); }
__device__ __forceinline__ float ab(float a) { return fabsf(a); }
__device__ __forceinline__ double mn(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float mn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double mx(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float mx(float a, float b) { return fmaxf(a, b); }

// internal energy e = (rhoE - |rho u|^2/(2 rho))/rho, NavierStokes::get_e (navier_stokes.cc:349-359)
template <typename T>
__device__ __forceinline__ T get_e(const T u[5]) {
    T ske = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    const T ri = dv(T(1), u[0]);
    ske *= T(0.5) * ri;
    return (u[4] - ske) * ri;
}
// pressure, NavierStokes::get_p (navier_stokes.cc:375-384)
template <typename T>
__device__ __forceinline__ T get_p(const T u[5], T gma) {
    T ske = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    ske *= dv(T(0.5), u[0]);
    return (gma - T(1)) * (u[4] - ske);
}
template <typename T>
__device__ __forceinline__ T get_a(const T u[5], T gma) {
    return sq(dv(gma * get_p(u, gma), u[0]));
}

// Sutherland viscosity and conductivity, navier_stokes.h:348-360
__device__ __forceinline__ double sutherland(const Gas& g, double T) {
    const double r = dv(T, g.T0);
    return dv(g.mu0 * (r * sqrt(r)) * (g.T0 + g.S), T + g.S);
}
__device__ __forceinline__ double conductivity(const Gas& g, double mu) { return mu * g.kfac; }

// exact Euler flux along dir (not necessarily unit), get_inv_flux (navier_stokes.cc:424-442)
template <typename T>
__device__ __forceinline__ void inv_flux(const T u[5], const T n[3], T gma, T f[5]) {
    const T ri = dv(T(1), u[0]);
    const T p = get_p(u, gma);
    const T vn = (u[1] * n[0] + u[2] * n[1] + u[3] * n[2]) * ri;
    f[0] = u[0] * vn;
    f[1] = vn * u[1] + p * n[0];
    f[2] = vn * u[2] + p * n[1];
    f[3] = vn * u[3] + p * n[2];
    f[4] = (u[4] + p) * vn;
}

// Ismail-Roe logarithmic mean with its series branch at u < 1e-2, log_avg (navier_stokes.h:494-501).
// Always FP64: the near-equal-state branch is among the parts kept in double.
__device__ __forceinline__ double log_avg(double xl, double xr) {
    const double xi = xl / xr, f = (xi - 1.0) / (xi + 1.0), u = f * f;
    const double F = (u < 1e-2) ? 1.0 + u * (1.0 / 3.0 + u * (1.0 / 5.0 + u * (1.0 / 7.0)))
                                : 0.5 * log(xi) / f;
    return 0.5 * (xl + xr) / F;
}

// HLLC x-flux with adaptive PVRS/TRRS/TSRS star pressure (navier_stokes.cc:936-1033)
template <typename T>
__device__ void hllc_xflux(const T l[5], const T r[5], T gma, T f[5]) {
    const T pl = get_p(l, gma), pr = get_p(r, gma);
    const T pmin = mn(pl, pr), pmax = mx(pl, pr);
    const T ul = dv(l[1], l[0]), ur = dv(r[1], r[0]);
    const T al = sq(dv(gma * pl, l[0])), ar = sq(dv(gma * pr, r[0]));
    const T p_pvrs = T(0.5) * (pl + pr) - T(0.5) * (ur - ul) * T(0.5) * (l[0] + r[0]) * T(0.5) * (al + ar);
    T ps;
    if (dv(pmax, pmin) < T(2) && pmin <= p_pvrs && p_pvrs <= pmax) {
        ps = p_pvrs;
    } else if (p_pvrs < pmin) {
        const T z = dv(gma - T(1), T(2) * gma);
        ps = pw(dv((al + ar) - dv(gma - T(1), T(2)) * (ur - ul), dv(al, pw(pl, z)) + dv(ar, pw(pr, z))), dv(T(1), z));
    } else {
        const T p0 = mx(T(0), p_pvrs);
        const T gl = sq(dv(dv(T(2), (gma + T(1)) * l[0]), p0 + dv(gma - T(1), gma + T(1)) * pl));
        const T gr = sq(dv(dv(T(2), (gma + T(1)) * r[0]), p0 + dv(gma - T(1), gma + T(1)) * pr));
        ps = dv(gl * pl + gr * pr - (ur - ul), gl + gr);
    }
    const T c1 = dv(gma + T(1), T(2) * gma);
    const T ql = (ps <= pl) ? T(1) : sq(T(1) + c1 * (dv(ps, pl) - T(1)));
    const T qr = (ps <= pr) ? T(1) : sq(T(1) + c1 * (dv(ps, pr) - T(1)));
    const T sl = ul - al * ql, sr = ur + ar * qr;
    const T s = dv(pr - pl + l[1] * (sl - ul) - r[1] * (sr - ur), l[0] * (sl - ul) - r[0] * (sr - ur));
    const T x[3] = {T(1), T(0), T(0)};
    if (sl > T(0)) {
        inv_flux(l, x, gma, f);
    } else if (s > T(0)) {
        const T t = dv(sl - ul, sl - s);
        T ff[5];
        inv_flux(l, x, gma, ff);
        const T ss[5] = {t * l[0], t * l[0] * s, t * l[2], t * l[3],
                         t * (l[4] + (s - ul) * (l[0] * s + dv(pl, sl - ul)))};
        for (int v = 0; v < 5; ++v) f[v] = ff[v] + sl * (ss[v] - l[v]);
    } else if (sr > T(0)) {
        const T t = dv(sr - ur, sr - s);
        T ff[5];
        inv_flux(r, x, gma, ff);
        const T ss[5] = {t * r[0], t * r[0] * s, t * r[2], t * r[3],
                         t * (r[4] + (s - ur) * (r[0] * s + dv(pr, sr - ur)))};
        for (int v = 0; v < 5; ++v) f[v] = ff[v] + sr * (ss[v] - r[v]);
    } else {
        inv_flux(r, x, gma, f);
    }
}

// HLLC flux along the unit normal n in the original frame: get_inv_surf_flux's rotation (momentum
// along n and two tangents), hllc_xflux and the rotation back, written out. The rotation is
// linear, so the rotated-back Euler x-flux is the normal flux F_n(u), the star momentum is
// t (rho s n + m - m_n n) with m_n = n . m, and mass and energy are those of hllc_xflux; the same
// star-pressure choice (PVRS / TRRS / TSRS) and wave speeds from the normal velocities. Neither the
// rotation nor the rotated states are formed (fewer registers); equal to the rotated form up to
// rounding. Like that rotation, a normal within 1e-3 of the x axis (|n_y, n_z| <= 1e-3) is taken
// as (sign n_x, 0, 0): the rotation is then sign(n_x) I, and the flux is that along this axis.
template <typename T>
__device__ void hllc_nflux(const T l[5], const T r[5], const T nin[3], T gma, T f[5]) {
    T n[3] = {nin[0], nin[1], nin[2]};
    if (!(sq(nin[2] * nin[2] + nin[1] * nin[1]) > T(1e-3))) {
        n[0] = nin[0] < T(0) ? T(-1) : T(1);
        n[1] = T(0);
        n[2] = T(0);
    }
    const T pl = get_p(l, gma), pr = get_p(r, gma);
    const T pmin = mn(pl, pr), pmax = mx(pl, pr);
    const T mnl = l[1] * n[0] + l[2] * n[1] + l[3] * n[2], mnr = r[1] * n[0] + r[2] * n[1] + r[3] * n[2];
    const T ul = dv(mnl, l[0]), ur = dv(mnr, r[0]);
    const T al = sq(dv(gma * pl, l[0])), ar = sq(dv(gma * pr, r[0]));
    const T p_pvrs = T(0.5) * (pl + pr) - T(0.5) * (ur - ul) * T(0.5) * (l[0] + r[0]) * T(0.5) * (al + ar);
    T ps;
    if (dv(pmax, pmin) < T(2) && pmin <= p_pvrs && p_pvrs <= pmax) {
        ps = p_pvrs;
    } else if (p_pvrs < pmin) {
        const T z = dv(gma - T(1), T(2) * gma);
        ps = pw(dv((al + ar) - dv(gma - T(1), T(2)) * (ur - ul), dv(al, pw(pl, z)) + dv(ar, pw(pr, z))), dv(T(1), z));
    } else {
        const T p0 = mx(T(0), p_pvrs);
        const T gl = sq(dv(dv(T(2), (gma + T(1)) * l[0]), p0 + dv(gma - T(1), gma + T(1)) * pl));
        const T gr = sq(dv(dv(T(2), (gma + T(1)) * r[0]), p0 + dv(gma - T(1), gma + T(1)) * pr));
        ps = dv(gl * pl + gr * pr - (ur - ul), gl + gr);
    }
    const T c1 = dv(gma + T(1), T(2) * gma);
    const T ql = (ps <= pl) ? T(1) : sq(T(1) + c1 * (dv(ps, pl) - T(1)));
    const T qr = (ps <= pr) ? T(1) : sq(T(1) + c1 * (dv(ps, pr) - T(1)));
    const T sl = ul - al * ql, sr = ur + ar * qr;
    const T s = dv(pr - pl + mnl * (sl - ul) - mnr * (sr - ur), l[0] * (sl - ul) - r[0] * (sr - ur));
    // the reference's four branches (sl > 0: F_L; s > 0: F_L*; sr > 0: F_R*; else F_R) as one
    // evaluation on the chosen side K, so that lanes of a wavefront on different sides share it
    const bool left = sl > T(0) || s > T(0);
    const bool star = !(sl > T(0)) && (s > T(0) || sr > T(0));
    T k[5];
    for (int v = 0; v < 5; ++v) k[v] = left ? l[v] : r[v];
    const T sk = left ? sl : sr, uk = left ? ul : ur, mnk = left ? mnl : mnr, pk = left ? pl : pr;
    inv_flux(k, n, gma, f);
    if (star) {
        const T t = dv(sk - uk, sk - s);
        f[0] += sk * (t * k[0] - k[0]);
        for (int a = 0; a < 3; ++a) f[1 + a] += sk * (t * (k[0] * s * n[a] + k[1 + a] - mnk * n[a]) - k[1 + a]);
        f[4] += sk * (t * (k[4] + (s - uk) * (k[0] * s + dv(pk, sk - uk))) - k[4]);
    }
}

// Rusanov x-flux (navier_stokes.cc:1044-1059)
template <typename T>
__device__ void rusanov_xflux(const T l[5], const T r[5], T gma, T f[5]) {
    const T x[3] = {T(1), T(0), T(0)};
    T lf[5], rf[5];
    inv_flux(l, x, gma, lf);
    inv_flux(r, x, gma, rf);
    const T S = mx(ab(dv(l[1], l[0])) + get_a(l, gma), ab(dv(r[1], r[0])) + get_a(r, gma));
    for (int v = 0; v < 5; ++v) f[v] = T(0.5) * (lf[v] + rf[v]) - T(0.5) * S * (r[v] - l[v]);
}

// AUSM+-up x-flux (navier_stokes.cc:1070-1121, modelling/ausm_functions.h). The reference's
// alpha = 3/16*(...) is an integer division and therefore 0; the pressure split uses alpha = 0.
template <typename T>
__device__ __forceinline__ T m2p(T M) { return T(0.25) * (M + T(1)) * (M + T(1)); }
template <typename T>
__device__ __forceinline__ T m2n(T M) { return T(-0.25) * (M - T(1)) * (M - T(1)); }
template <typename T>
__device__ void ausm_xflux(const T l[5], const T r[5], T gma, T f[5]) {
    const T pl = get_p(l, gma), pr = get_p(r, gma);
    const T vl0 = dv(l[1], l[0]), vr0 = dv(r[1], r[0]);
    const T Hl = dv(l[4] + pl, l[0]), Hr = dv(r[4] + pr, r[0]);
    const T a12 = T(0.5) * (sq(dv(gma * pl, l[0])) + sq(dv(gma * pr, r[0])));
    const T Ml = dv(vl0, a12), Mr = dv(vr0, a12);
    const T Mavgsq = T(0.5) * (Ml * Ml + Mr * Mr);
    const T Mo = sq(mn(T(1), mx(Mavgsq, T(0.3) * T(0.3))));
    const T fa = Mo * (T(2) - Mo);
    const T m4p = ab(Ml) >= T(1) ? T(0.5) * (Ml + ab(Ml)) : m2p(Ml) * (T(1) - T(2) * m2n(Ml));
    const T m4n = ab(Mr) >= T(1) ? T(0.5) * (Mr - ab(Mr)) : m2n(Mr) * (T(1) + T(2) * m2p(Mr));
    const T M12 = m4p + m4n - dv(dv(T(0.25), fa) * mx(T(1) - Mavgsq, T(0)) * T(2) * (pr - pl), a12 * a12 * (l[0] + r[0]));
    const T psp = ab(Ml) >= T(1) ? dv(T(0.5) * (Ml + ab(Ml)), Ml) : m2p(Ml) * (T(2) - Ml);
    const T psn = ab(Mr) >= T(1) ? dv(T(0.5) * (Mr - ab(Mr)), Mr) : m2n(Mr) * (T(-2) - Mr);
    const T p12 = psp * pl + psn * pr - T(0.75) * fa * psp * psn * (l[0] + r[0]) * a12 * (vr0 - vl0);
    const T* s = (M12 > T(0)) ? l : r;
    const T m12 = a12 * M12 * s[0];
    f[0] = m12;
    f[1] = m12 * dv(s[1], s[0]) + p12;
    f[2] = m12 * dv(s[2], s[0]);
    f[3] = m12 * dv(s[3], s[0]);
    f[4] = m12 * ((M12 > T(0)) ? Hl : Hr);
}

// log mean in the flux arithmetic with the series branch in FP64 (exact for T = double)
template <typename T>
__device__ __forceinline__ T log_avg_t(T xl, T xr) {
    return T(log_avg((double)xl, (double)xr));
}

// Chandrashekhar EC x-flux minus hybrid matrix dissipation 1/2 K |Lambda| K^-1 (u_r - u_l) with
// Roe-averaged K and the eigenvalue blend by the flux blender (navier_stokes.cc:1250-1318;
// K, K^-1 from get_xK / get_xKinv, navier_stokes.cc:641-678, 780-825)
template <typename T>
__device__ void chandrashekhar_xflux(const T l[5], const T r[5], T gma, T blender, T f[5]) {
    const T pl = get_p(l, gma), pr = get_p(r, gma);
    const T bl = dv(T(0.5) * l[0], pl), br = dv(T(0.5) * r[0], pr), bln = log_avg_t(bl, br);
    const T p_hat = dv(T(0.5) * (l[0] + r[0]), bl + br);
    const T rho_ln = log_avg_t(l[0], r[0]);
    const T ril = dv(T(1), l[0]), rir = dv(T(1), r[0]);
    T vl[3], vr[3], va[3], vsq = T(0);
    T H_hat = dv(T(0.5), (gma - T(1)) * bln) + dv(p_hat, rho_ln);
    for (int d = 0; d < 3; ++d) {
        vl[d] = l[1 + d] * ril;
        vr[d] = r[1 + d] * rir;
        va[d] = T(0.5) * (vl[d] + vr[d]);
        vsq = T(0.5) * (vl[d] * vl[d] + vr[d] * vr[d]);
        H_hat += va[d] * va[d] - T(0.5) * vsq;
    }
    f[0] = rho_ln * va[0];
    f[1] = rho_ln * va[0] * va[0] + p_hat;
    f[2] = rho_ln * va[0] * va[1];
    f[3] = rho_ln * va[0] * va[2];
    f[4] = rho_ln * va[0] * H_hat;
    // Roe average
    const T sql = sq(l[0]), sqr = sq(r[0]);
    const T inv = dv(T(1), sql + sqr);
    T vi[3];
    for (int d = 0; d < 3; ++d) vi[d] = inv * (sql * vl[d] + sqr * vr[d]);
    const T Hi = dv(dv(l[4] + pl, sql) + dv(r[4] + pr, sqr), sql + sqr);
    const T v2 = vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2];
    const T ai = sq((gma - T(1)) * (Hi - T(0.5) * v2));
    const T uabs = ab(vi[0]), lmax = uabs + ai, l2 = blender * lmax + (T(1) - blender) * uabs;
    T du[5];
    for (int v = 0; v < 5; ++v) du[v] = r[v] - l[v];
    // w = K^-1 du (rows of get_xKinv times 0.5 (gma-1)/a^2)
    const T t = dv(T(1), gma - T(1)), a2t = T(2) * ai * ai * t;
    const T fac = dv(T(0.5) * (gma - T(1)), ai * ai);
    T w[5];
    w[0] = fac * ((Hi + ai * (vi[0] - ai) * t) * du[0] + (-vi[0] - ai * t) * du[1] - vi[1] * du[2] - vi[2] * du[3] + du[4]);
    w[1] = fac * (T(2) * (T(2) * ai * ai * t - Hi) * du[0] + T(2) * vi[0] * du[1] + T(2) * vi[1] * du[2] +
                  T(2) * vi[2] * du[3] - T(2) * du[4]);
    w[2] = fac * (-vi[1] * a2t * du[0] + a2t * du[2]);
    w[3] = fac * (-vi[2] * a2t * du[0] + a2t * du[3]);
    w[4] = fac * ((Hi - ai * (vi[0] + ai) * t) * du[0] + (-vi[0] + ai * t) * du[1] - vi[1] * du[2] - vi[2] * du[3] + du[4]);
    w[0] *= lmax; w[1] *= l2; w[2] *= l2; w[3] *= l2; w[4] *= lmax;
    // K w
    T kw[5];
    kw[0] = w[0] + w[1] + w[4];
    kw[1] = (vi[0] - ai) * w[0] + vi[0] * w[1] + (vi[0] + ai) * w[4];
    kw[2] = vi[1] * w[0] + vi[1] * w[1] + w[2] + vi[1] * w[4];
    kw[3] = vi[2] * w[0] + vi[2] * w[1] + w[3] + vi[2] * w[4];
    kw[4] = (Hi - vi[0] * ai) * w[0] + T(0.5) * v2 * w[1] + vi[1] * w[2] + vi[2] * w[3] + (Hi + vi[0] * ai) * w[4];
    for (int v = 0; v < 5; ++v) f[v] -= T(0.5) * kw[v];
}

// K diag(lam) K^-1 x for the x-direction eigenvectors of velocity v, sound speed a and enthalpy H
// (get_xK / get_xKinv, navier_stokes.cc:641-678, 780-825; the rows as in chandrashekhar_xflux)
template <typename T>
__device__ __forceinline__ void kxlam_kinv(const T v[3], T a, T H, T gma, const T lam[5], const T x[5], T y[5]) {
    const T t = dv(T(1), gma - T(1)), a2t = T(2) * a * a * t;
    const T fac = dv(T(0.5) * (gma - T(1)), a * a);
    T w[5];
    w[0] = fac * ((H + a * (v[0] - a) * t) * x[0] + (-v[0] - a * t) * x[1] - v[1] * x[2] - v[2] * x[3] + x[4]);
    w[1] = fac * (T(2) * (T(2) * a * a * t - H) * x[0] + T(2) * v[0] * x[1] + T(2) * v[1] * x[2] +
                  T(2) * v[2] * x[3] - T(2) * x[4]);
    w[2] = fac * (-v[1] * a2t * x[0] + a2t * x[2]);
    w[3] = fac * (-v[2] * a2t * x[0] + a2t * x[3]);
    w[4] = fac * ((H - a * (v[0] + a) * t) * x[0] + (-v[0] + a * t) * x[1] - v[1] * x[2] - v[2] * x[3] + x[4]);
    for (int k = 0; k < 5; ++k) w[k] *= lam[k];
    const T v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    y[0] = w[0] + w[1] + w[4];
    y[1] = (v[0] - a) * w[0] + v[0] * w[1] + (v[0] + a) * w[4];
    y[2] = v[1] * w[0] + v[1] * w[1] + w[2] + v[1] * w[4];
    y[3] = v[2] * w[0] + v[2] * w[1] + w[3] + v[2] * w[4];
    y[4] = (H - v[0] * a) * w[0] + T(0.5) * v2 * w[1] + v[1] * w[2] + v[2] * w[3] + (H + v[0] * a) * w[4];
}

// modified Steger-Warming x-flux (modified_sw_xflux, navier_stokes.cc:1160-1235): pressure-switched
// states, smoothed split eigenvalues (eps 0.3), f = A+ u_l + A- u_r
template <typename T>
__device__ void modified_sw_xflux(const T l[5], const T r[5], T gma, T f[5]) {
    const T pl = get_p(l, gma), pr = get_p(r, gma);
    const T sw = dv(T(0.5) * (pr - pl), mn(pl, pr));
    const T om = dv(T(0.5), T(1) + sw * sw);
    const T ril = dv(T(1), l[0]), rir = dv(T(1), r[0]);
    T vp[3], vn[3];
    for (int d = 0; d < 3; ++d) {
        vp[d] = (T(1) - om) * (l[1 + d] * ril) + om * (r[1 + d] * rir);
        vn[d] = (T(1) - om) * (r[1 + d] * rir) + om * (l[1 + d] * ril);
    }
    const T rp = (T(1) - om) * l[0] + om * r[0], rn = (T(1) - om) * r[0] + om * l[0];
    const T pp = (T(1) - om) * pl + om * pr, pn = (T(1) - om) * pr + om * pl;
    const T ap = sq(dv(gma * pp, rp)), an = sq(dv(gma * pn, rn));
    const T Hp = dv(gma * pp, (gma - T(1)) * rp) + T(0.5) * (vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2]);
    const T Hn = dv(gma * pn, (gma - T(1)) * rn) + T(0.5) * (vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
    const T ep = T(0.3) * ap, en = T(0.3) * an;
    auto psm = [](T x, T e) { return T(0.5) * (x + sq(x * x + e * e)); };  // pos_smooth (navier_stokes.h:464)
    auto nsm = [](T x, T e) { return T(0.5) * (x - sq(x * x + e * e)); };  // neg_smooth (navier_stokes.h:473)
    const T lp[5] = {psm(vp[0] - ap, ep), psm(vp[0], ep), psm(vp[0], ep), psm(vp[0], ep), psm(vp[0] + ap, ep)};
    const T ln[5] = {nsm(vn[0] - an, en), nsm(vn[0], en), nsm(vn[0], en), nsm(vn[0], en), nsm(vn[0] + an, en)};
    T fp[5], fn[5];
    kxlam_kinv(vp, ap, Hp, gma, lp, l, fp);
    kxlam_kinv(vn, an, Hn, gma, ln, r, fn);
    for (int v = 0; v < 5; ++v) f[v] = fp[v] + fn[v];
}

// Kennedy-Gruber x-flux with the Rusanov-type stabilisation (kennedy_gruber_xflux,
// navier_stokes.cc:1322-1355)
template <typename T>
__device__ void kennedy_gruber_xflux(const T l[5], const T r[5], T gma, T f[5]) {
    const T ril = dv(T(1), l[0]), rir = dv(T(1), r[0]);
    T vl[3], vr[3], va[3];
    for (int d = 0; d < 3; ++d) {
        vl[d] = l[1 + d] * ril;
        vr[d] = r[1 + d] * rir;
        va[d] = T(0.5) * (vl[d] + vr[d]);
    }
    const T ra = T(0.5) * (l[0] + r[0]);
    const T el = get_e(l), er = get_e(r);
    const T Ea = T(0.5) * (el + er + T(0.5) * ((vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]) +
                                                (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2])));
    const T pa = (gma - T(1)) * T(0.5) * (l[0] * el + r[0] * er);
    f[0] = va[0] * ra;
    for (int d = 0; d < 3; ++d) f[1 + d] = ra * va[0] * va[d];
    f[1] += pa;
    f[4] = va[0] * (ra * Ea + pa);
    const T lm = mx(ab(vl[0]) + sq(gma * (gma - T(1)) * el), ab(vr[0]) + sq(gma * (gma - T(1)) * er));
    for (int v = 0; v < 5; ++v) f[v] -= T(0.5) * lm * (r[v] - l[v]);
}

// Ismail-Roe x-flux with the entropy-variable stabilisation (ismail_roe_xflux,
// navier_stokes.cc:1373-1442)
template <typename T>
__device__ void ismail_roe_xflux(const T l[5], const T r[5], T gma, T f[5]) {
    const T gs = sq(gma);
    const T al = get_a(l, gma), ar = get_a(r, gma);
    const T z1l = dv(gs, al), z1r = dv(gs, ar);
    const T z5l = dv(l[0] * al, gs), z5r = dv(r[0] * ar, gs);
    const T z5ln = log_avg_t(z5l, z5r);
    const T rh = T(0.5) * (z1l + z1r) * z5ln;
    const T p1 = dv(z5l + z5r, z1l + z1r);
    const T p2 = dv(T(0.5) * (dv((gma + T(1)) * z5ln, log_avg_t(z1l, z1r)) + (gma - T(1)) * p1), gma);
    T hh = dv(gma * p2, (gma - T(1)) * rh);
    T vl[3], vr[3], vh[3];
    for (int d = 0; d < 3; ++d) {
        vl[d] = dv(l[1 + d], l[0]);
        vr[d] = dv(r[1 + d], r[0]);
        vh[d] = dv(z1l * vl[d] + z1r * vr[d], z1l + z1r);
        hh += T(0.5) * (vh[d] * vh[d]);
    }
    f[0] = rh * vh[0];
    for (int d = 0; d < 3; ++d) f[1 + d] = rh * vh[0] * vh[d];
    f[1] += p1;
    f[4] = rh * hh * vh[0];
    const T ch[5] = {rh, rh * vh[0], rh * vh[1], rh * vh[2], rh * hh - p2};
    const T pl = get_p(l, gma), pr = get_p(r, gma);
    const T sl = lg(pl) - gma * lg(l[0]), sr = lg(pr) - gma * lg(r[0]);
    T dV[5];
    dV[0] = (dv(gma - sr, gma - T(1)) - dv(T(0.5) * r[0] * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]), pr)) -
            (dv(gma - sl, gma - T(1)) - dv(T(0.5) * l[0] * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]), pl));
    for (int d = 0; d < 3; ++d) dV[1 + d] = dv(r[1 + d], pr) - dv(l[1 + d], pl);
    dV[4] = -dv(r[0], pr) + dv(l[0], pl);
    T H[5][5];
    for (int i = 0; i < 5; ++i) H[i][0] = ch[i];
    for (int d = 0; d < 3; ++d) {
        const T e[3] = {T(d == 0), T(d == 1), T(d == 2)};
        T fl[5];
        inv_flux(ch, e, gma, fl);
        for (int i = 0; i < 5; ++i) H[i][1 + d] = fl[i];
    }
    for (int i = 0; i < 4; ++i) H[i][4] = H[4][i];
    H[4][4] = rh * hh * hh - dv(gma * p2 * p2, rh * (gma - T(1)));
    const T lm = ab(vh[0]) + sq(dv(gma * p2, rh));
    for (int i = 0; i < 5; ++i) {
        T s = T(0);
        for (int j = 0; j < 5; ++j) s += H[i][j] * dV[j];
        f[i] -= T(0.5) * lm * s;
    }
}

// two-point volume fluxes along an unnormalised direction a (linear in a): Kennedy-Gruber
// (kennedy_gruber_vol_flux, navier_stokes.cc:1551-1574) and Ismail-Roe (ismail_roe_vol_flux,
// 1589-1617)
template <typename T>
__device__ void kennedy_gruber_vol_flux(const T u1[5], const T u2[5], const T a[3], T gma, T f[5]) {
    const T r1 = dv(T(1), u1[0]), r2 = dv(T(1), u2[0]);
    T v1[3], v2[3], va[3];
    for (int d = 0; d < 3; ++d) {
        v1[d] = u1[1 + d] * r1;
        v2[d] = u2[1 + d] * r2;
        va[d] = T(0.5) * (v1[d] + v2[d]);
    }
    const T ra = T(0.5) * (u1[0] + u2[0]);
    const T e1 = get_e(u1), e2 = get_e(u2);
    const T Ea = T(0.5) * (e1 + e2 + T(0.5) * ((v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]) +
                                                (v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2])));
    const T pa = (gma - T(1)) * T(0.5) * (u1[0] * e1 + u2[0] * e2);
    const T vn = va[0] * a[0] + va[1] * a[1] + va[2] * a[2];
    f[0] = vn * ra;
    for (int d = 0; d < 3; ++d) f[1 + d] = ra * vn * va[d] + pa * a[d];
    f[4] = vn * (ra * Ea + pa);
}
template <typename T>
__device__ void ismail_roe_vol_flux(const T u1[5], const T u2[5], const T a[3], T gma, T f[5]) {
    const T gs = sq(gma);
    const T a1 = get_a(u1, gma), a2 = get_a(u2, gma);
    const T z11 = dv(gs, a1), z12 = dv(gs, a2);
    const T z51 = dv(u1[0] * a1, gs), z52 = dv(u2[0] * a2, gs);
    const T z5ln = log_avg_t(z51, z52);
    const T rh = T(0.5) * (z11 + z12) * z5ln;
    const T p1 = dv(z51 + z52, z11 + z12);
    const T p2 = dv(T(0.5) * (dv((gma + T(1)) * z5ln, log_avg_t(z11, z12)) + (gma - T(1)) * p1), gma);
    T hh = dv(gma * p2, (gma - T(1)) * rh);
    T vh[3];
    for (int d = 0; d < 3; ++d) {
        vh[d] = dv(z11 * dv(u1[1 + d], u1[0]) + z12 * dv(u2[1 + d], u2[0]), z11 + z12);
        hh += T(0.5) * (vh[d] * vh[d]);
    }
    const T vn = vh[0] * a[0] + vh[1] * a[1] + vh[2] * a[2];
    f[0] = rh * vn;
    for (int d = 0; d < 3; ++d) f[1 + d] = rh * vh[d] * vn + p1 * a[d];
    f[4] = rh * vn * hh;
}

// numerical flux through unit normal n with the normal-aligned rotation of
// get_inv_surf_flux (navier_stokes.cc:482-569)
// EXT = false compiles schemes 0-3 only (the kernels' common instantiation); EXT = true all ten
template <bool EXT, typename T>
__device__ void inv_surf_flux(int scheme, T blender, const T ul[5], const T ur[5], const T n[3], T gma, T f[5]) {
    // m = x cross n = (0, -n2, n1)
    T m1 = -n[2], m2 = n[1];
    const T Mm = sq(m1 * m1 + m2 * m2);
    T R[3][3];
    if (Mm > T(1e-3)) {
        m1 = dv(m1, Mm);
        m2 = dv(m2, Mm);
        const T c = n[0], t = T(1) - c, s = sq(T(1) - c * c);
        // Rodrigues rotation about (0, m1, m2)
        R[0][0] = c;             R[0][1] = -s * m2;            R[0][2] = s * m1;
        R[1][0] = s * m2;        R[1][1] = t * m1 * m1 + c;    R[1][2] = t * m1 * m2;
        R[2][0] = -s * m1;       R[2][1] = t * m1 * m2;        R[2][2] = t * m2 * m2 + c;
    } else {
        const T dg = n[0] < T(0) ? T(-1) : T(1);
        R[0][0] = dg; R[0][1] = T(0); R[0][2] = T(0);
        R[1][0] = T(0); R[1][1] = dg; R[1][2] = T(0);
        R[2][0] = T(0); R[2][1] = T(0); R[2][2] = dg;
    }
    T lr[5], rr[5], fr[5];
    lr[0] = ul[0]; rr[0] = ur[0]; lr[4] = ul[4]; rr[4] = ur[4];
    for (int a = 0; a < 3; ++a) {  // R^T * momentum
        lr[1 + a] = R[0][a] * ul[1] + R[1][a] * ul[2] + R[2][a] * ul[3];
        rr[1 + a] = R[0][a] * ur[1] + R[1][a] * ur[2] + R[2][a] * ur[3];
    }
    if constexpr (!EXT) {
        switch (scheme) {
            case 0: hllc_xflux(lr, rr, gma, fr); break;
            case 1: rusanov_xflux(lr, rr, gma, fr); break;
            case 2: ausm_xflux(lr, rr, gma, fr); break;
            default: chandrashekhar_xflux(lr, rr, gma, blender, fr); break;
        }
    } else {
        switch (scheme) {
            case 0: hllc_xflux(lr, rr, gma, fr); break;
            case 1: rusanov_xflux(lr, rr, gma, fr); break;
            case 2: ausm_xflux(lr, rr, gma, fr); break;
            case 3: chandrashekhar_xflux(lr, rr, gma, blender, fr); break;
            case 6: modified_sw_xflux(lr, rr, gma, fr); break;
            case 7: kennedy_gruber_xflux(lr, rr, gma, fr); break;
            case 9: ismail_roe_xflux(lr, rr, gma, fr); break;
            default: {  // Rusanov blends, weight = flux blender (navier_stokes.cc:1132-1160, 1364-1374)
                T fa[5], fb[5];
                rusanov_xflux(lr, rr, gma, fa);
                if (scheme == 4) hllc_xflux(lr, rr, gma, fb);
                else if (scheme == 5) ausm_xflux(lr, rr, gma, fb);
                else kennedy_gruber_xflux(lr, rr, gma, fb);
                for (int v = 0; v < 5; ++v) fr[v] = blender * fa[v] + (T(1) - blender) * fb[v];
                break;
            }
        }
    }
    f[0] = fr[0];
    f[4] = fr[4];
    for (int a = 0; a < 3; ++a) f[1 + a] = R[a][0] * fr[1] + R[a][1] * fr[2] + R[a][2] * fr[3];
}

// viscous flux along dir from (state, avars) with avars = (txx,txy,txz,tyy,tyz,tzz,qx,qy,qz)
// (get_dif_flux, navier_stokes.cc:599-624); linear in dir
template <typename T>
__device__ __forceinline__ void dif_flux(const T u[5], const T av[9], const T n[3], T f[5]) {
    const T m0 = av[0] * n[0] + av[1] * n[1] + av[2] * n[2];
    const T m1 = av[1] * n[0] + av[3] * n[1] + av[4] * n[2];
    const T m2 = av[2] * n[0] + av[4] * n[1] + av[5] * n[2];
    const T ri = dv(T(1), u[0]);
    f[0] = T(0);
    f[1] = m0;
    f[2] = m1;
    f[3] = m2;
    f[4] = -av[6] * n[0] + u[1] * ri * m0 - av[7] * n[1] + u[2] * ri * m1 - av[8] * n[2] + u[3] * ri * m2;
}

}  // namespace plens_dev

// Device-side gas model and flux functions of the DGSEM hot path (gfx950).
//
// Each function computes the same quantity as the NavierStokes member cited beside it; the
// operation order is free where rounding is the only difference (e.g. the Chandrashekhar matrix
// dissipation is applied as K*(Lambda*(K^-1*du)) instead of forming K*Lambda*K^-1).
// Branch conditions (HLLC pressure estimates, log-mean series switch, AUSM Mach splits) follow the
// reference so that the same branch is taken.
#pragma once
#include <hip/hip_runtime.h>

namespace plens_dev {

struct Gas {
    double gma, R, cv, mu0, T0, S, Pr;
    int inviscid;
};

// internal energy e = (rhoE - |rho u|^2/(2 rho))/rho, NavierStokes::get_e (navier_stokes.cc:349-359)
__device__ __forceinline__ double get_e(const double u[5]) {
    double ske = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    const double ri = 1.0 / u[0];
    ske *= 0.5 * ri;
    return (u[4] - ske) * ri;
}
// pressure, NavierStokes::get_p (navier_stokes.cc:375-384)
__device__ __forceinline__ double get_p(const double u[5], double gma) {
    double ske = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    ske *= 0.5 / u[0];
    return (gma - 1.0) * (u[4] - ske);
}
__device__ __forceinline__ double get_a(const double u[5], double gma) {
    return sqrt(gma * get_p(u, gma) / u[0]);
}

// Sutherland viscosity and conductivity, navier_stokes.h:348-360
__device__ __forceinline__ double sutherland(const Gas& g, double T) {
    const double r = T / g.T0;
    return g.mu0 * (r * sqrt(r)) * (g.T0 + g.S) / (T + g.S);
}
__device__ __forceinline__ double conductivity(const Gas& g, double mu) {
    return mu * g.R * g.gma / ((g.gma - 1.0) * g.Pr);
}

// exact Euler flux along dir (not necessarily unit), get_inv_flux (navier_stokes.cc:424-442)
__device__ __forceinline__ void inv_flux(const double u[5], const double n[3], double gma, double f[5]) {
    const double ri = 1.0 / u[0];
    const double p = get_p(u, gma);
    const double vn = (u[1] * n[0] + u[2] * n[1] + u[3] * n[2]) * ri;
    f[0] = u[0] * vn;
    f[1] = vn * u[1] + p * n[0];
    f[2] = vn * u[2] + p * n[1];
    f[3] = vn * u[3] + p * n[2];
    f[4] = (u[4] + p) * vn;
}

// Ismail-Roe logarithmic mean with its series branch at u < 1e-2, log_avg (navier_stokes.h:494-501)
__device__ __forceinline__ double log_avg(double xl, double xr) {
    const double xi = xl / xr, f = (xi - 1.0) / (xi + 1.0), u = f * f;
    const double F = (u < 1e-2) ? 1.0 + u * (1.0 / 3.0 + u * (1.0 / 5.0 + u * (1.0 / 7.0)))
                                : 0.5 * log(xi) / f;
    return 0.5 * (xl + xr) / F;
}

// HLLC x-flux with adaptive PVRS/TRRS/TSRS star pressure (navier_stokes.cc:936-1033)
__device__ void hllc_xflux(const double l[5], const double r[5], double gma, double f[5]) {
    const double pl = get_p(l, gma), pr = get_p(r, gma);
    const double pmin = fmin(pl, pr), pmax = fmax(pl, pr);
    const double ul = l[1] / l[0], ur = r[1] / r[0];
    const double al = sqrt(gma * pl / l[0]), ar = sqrt(gma * pr / r[0]);
    const double p_pvrs = 0.5 * (pl + pr) - 0.5 * (ur - ul) * 0.5 * (l[0] + r[0]) * 0.5 * (al + ar);
    double ps;
    if (pmax / pmin < 2.0 && pmin <= p_pvrs && p_pvrs <= pmax) {
        ps = p_pvrs;
    } else if (p_pvrs < pmin) {
        const double z = (gma - 1.0) / (2.0 * gma);
        ps = pow(((al + ar) - (gma - 1.0) / 2.0 * (ur - ul)) / (al / pow(pl, z) + ar / pow(pr, z)), 1.0 / z);
    } else {
        const double p0 = fmax(0.0, p_pvrs);
        const double gl = sqrt(2.0 / ((gma + 1.0) * l[0]) / (p0 + (gma - 1.0) / (gma + 1.0) * pl));
        const double gr = sqrt(2.0 / ((gma + 1.0) * r[0]) / (p0 + (gma - 1.0) / (gma + 1.0) * pr));
        ps = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
    }
    const double c1 = (gma + 1.0) / (2.0 * gma);
    const double ql = (ps <= pl) ? 1.0 : sqrt(1.0 + c1 * (ps / pl - 1.0));
    const double qr = (ps <= pr) ? 1.0 : sqrt(1.0 + c1 * (ps / pr - 1.0));
    const double sl = ul - al * ql, sr = ur + ar * qr;
    const double s = (pr - pl + l[1] * (sl - ul) - r[1] * (sr - ur)) / (l[0] * (sl - ul) - r[0] * (sr - ur));
    const double x[3] = {1.0, 0.0, 0.0};
    if (sl > 0.0) {
        inv_flux(l, x, gma, f);
    } else if (s > 0.0) {
        const double t = (sl - ul) / (sl - s);
        double ff[5];
        inv_flux(l, x, gma, ff);
        const double ss[5] = {t * l[0], t * l[0] * s, t * l[2], t * l[3],
                              t * (l[4] + (s - ul) * (l[0] * s + pl / (sl - ul)))};
        for (int v = 0; v < 5; ++v) f[v] = ff[v] + sl * (ss[v] - l[v]);
    } else if (sr > 0.0) {
        const double t = (sr - ur) / (sr - s);
        double ff[5];
        inv_flux(r, x, gma, ff);
        const double ss[5] = {t * r[0], t * r[0] * s, t * r[2], t * r[3],
                              t * (r[4] + (s - ur) * (r[0] * s + pr / (sr - ur)))};
        for (int v = 0; v < 5; ++v) f[v] = ff[v] + sr * (ss[v] - r[v]);
    } else {
        inv_flux(r, x, gma, f);
    }
}

// Rusanov x-flux (navier_stokes.cc:1044-1059)
__device__ void rusanov_xflux(const double l[5], const double r[5], double gma, double f[5]) {
    const double x[3] = {1.0, 0.0, 0.0};
    double lf[5], rf[5];
    inv_flux(l, x, gma, lf);
    inv_flux(r, x, gma, rf);
    const double S = fmax(fabs(l[1] / l[0]) + get_a(l, gma), fabs(r[1] / r[0]) + get_a(r, gma));
    for (int v = 0; v < 5; ++v) f[v] = 0.5 * (lf[v] + rf[v]) - 0.5 * S * (r[v] - l[v]);
}

// AUSM+-up x-flux (navier_stokes.cc:1070-1121, modelling/ausm_functions.h). The reference's
// alpha = 3/16*(...) is an integer division and therefore 0; the pressure split uses alpha = 0.
__device__ __forceinline__ double m2p(double M) { return 0.25 * (M + 1.0) * (M + 1.0); }
__device__ __forceinline__ double m2n(double M) { return -0.25 * (M - 1.0) * (M - 1.0); }
__device__ void ausm_xflux(const double l[5], const double r[5], double gma, double f[5]) {
    const double pl = get_p(l, gma), pr = get_p(r, gma);
    const double vl0 = l[1] / l[0], vr0 = r[1] / r[0];
    const double Hl = (l[4] + pl) / l[0], Hr = (r[4] + pr) / r[0];
    const double a12 = 0.5 * (sqrt(gma * pl / l[0]) + sqrt(gma * pr / r[0]));
    const double Ml = vl0 / a12, Mr = vr0 / a12;
    const double Mavgsq = 0.5 * (Ml * Ml + Mr * Mr);
    const double Mo = sqrt(fmin(1.0, fmax(Mavgsq, 0.3 * 0.3)));
    const double fa = Mo * (2.0 - Mo);
    const double m4p = fabs(Ml) >= 1.0 ? 0.5 * (Ml + fabs(Ml)) : m2p(Ml) * (1.0 - 2.0 * m2n(Ml));
    const double m4n = fabs(Mr) >= 1.0 ? 0.5 * (Mr - fabs(Mr)) : m2n(Mr) * (1.0 + 2.0 * m2p(Mr));
    const double M12 = m4p + m4n - 0.25 / fa * fmax(1.0 - Mavgsq, 0.0) * 2.0 * (pr - pl) / (a12 * a12 * (l[0] + r[0]));
    const double psp = fabs(Ml) >= 1.0 ? 0.5 * (Ml + fabs(Ml)) / Ml : m2p(Ml) * (2.0 - Ml);
    const double psn = fabs(Mr) >= 1.0 ? 0.5 * (Mr - fabs(Mr)) / Mr : m2n(Mr) * (-2.0 - Mr);
    const double p12 = psp * pl + psn * pr - 0.75 * fa * psp * psn * (l[0] + r[0]) * a12 * (vr0 - vl0);
    const double* s = (M12 > 0.0) ? l : r;
    const double m12 = a12 * M12 * s[0];
    f[0] = m12;
    f[1] = m12 * (s[1] / s[0]) + p12;
    f[2] = m12 * (s[2] / s[0]);
    f[3] = m12 * (s[3] / s[0]);
    f[4] = m12 * ((M12 > 0.0) ? Hl : Hr);
}

// Chandrashekhar EC x-flux minus hybrid matrix dissipation 1/2 K |Lambda| K^-1 (u_r - u_l) with
// Roe-averaged K and the eigenvalue blend by the flux blender (navier_stokes.cc:1250-1318;
// K, K^-1 from get_xK / get_xKinv, navier_stokes.cc:641-678, 780-825)
__device__ void chandrashekhar_xflux(const double l[5], const double r[5], double gma, double blender,
                                     double f[5]) {
    const double pl = get_p(l, gma), pr = get_p(r, gma);
    const double bl = 0.5 * l[0] / pl, br = 0.5 * r[0] / pr, bln = log_avg(bl, br);
    const double p_hat = 0.5 * (l[0] + r[0]) / (bl + br);
    const double rho_ln = log_avg(l[0], r[0]);
    const double ril = 1.0 / l[0], rir = 1.0 / r[0];
    double vl[3], vr[3], va[3], vsq = 0.0;
    double H_hat = 0.5 / ((gma - 1.0) * bln) + p_hat / rho_ln;
    for (int d = 0; d < 3; ++d) {
        vl[d] = l[1 + d] * ril;
        vr[d] = r[1 + d] * rir;
        va[d] = 0.5 * (vl[d] + vr[d]);
        vsq = 0.5 * (vl[d] * vl[d] + vr[d] * vr[d]);
        H_hat += va[d] * va[d] - 0.5 * vsq;
    }
    f[0] = rho_ln * va[0];
    f[1] = rho_ln * va[0] * va[0] + p_hat;
    f[2] = rho_ln * va[0] * va[1];
    f[3] = rho_ln * va[0] * va[2];
    f[4] = rho_ln * va[0] * H_hat;
    // Roe average
    const double sql = sqrt(l[0]), sqr = sqrt(r[0]);
    const double inv = 1.0 / (sql + sqr);
    double vi[3];
    for (int d = 0; d < 3; ++d) vi[d] = inv * (sql * vl[d] + sqr * vr[d]);
    const double Hi = ((l[4] + pl) / sql + (r[4] + pr) / sqr) / (sql + sqr);
    const double v2 = vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2];
    const double ai = sqrt((gma - 1.0) * (Hi - 0.5 * v2));
    const double uabs = fabs(vi[0]), lmax = uabs + ai, l2 = blender * lmax + (1.0 - blender) * uabs;
    double du[5];
    for (int v = 0; v < 5; ++v) du[v] = r[v] - l[v];
    // w = K^-1 du (rows of get_xKinv times 0.5 (gma-1)/a^2)
    const double t = 1.0 / (gma - 1.0), a2t = 2.0 * ai * ai * t;
    const double fac = 0.5 * (gma - 1.0) / (ai * ai);
    double w[5];
    w[0] = fac * ((Hi + ai * (vi[0] - ai) * t) * du[0] + (-vi[0] - ai * t) * du[1] - vi[1] * du[2] - vi[2] * du[3] + du[4]);
    w[1] = fac * (2.0 * (2.0 * ai * ai * t - Hi) * du[0] + 2.0 * vi[0] * du[1] + 2.0 * vi[1] * du[2] + 2.0 * vi[2] * du[3] - 2.0 * du[4]);
    w[2] = fac * (-vi[1] * a2t * du[0] + a2t * du[2]);
    w[3] = fac * (-vi[2] * a2t * du[0] + a2t * du[3]);
    w[4] = fac * ((Hi - ai * (vi[0] + ai) * t) * du[0] + (-vi[0] + ai * t) * du[1] - vi[1] * du[2] - vi[2] * du[3] + du[4]);
    w[0] *= lmax; w[1] *= l2; w[2] *= l2; w[3] *= l2; w[4] *= lmax;
    // K w
    double kw[5];
    kw[0] = w[0] + w[1] + w[4];
    kw[1] = (vi[0] - ai) * w[0] + vi[0] * w[1] + (vi[0] + ai) * w[4];
    kw[2] = vi[1] * w[0] + vi[1] * w[1] + w[2] + vi[1] * w[4];
    kw[3] = vi[2] * w[0] + vi[2] * w[1] + w[3] + vi[2] * w[4];
    kw[4] = (Hi - vi[0] * ai) * w[0] + 0.5 * v2 * w[1] + vi[1] * w[2] + vi[2] * w[3] + (Hi + vi[0] * ai) * w[4];
    for (int v = 0; v < 5; ++v) f[v] -= 0.5 * kw[v];
}

// numerical flux through unit normal n with the normal-aligned rotation of
// get_inv_surf_flux (navier_stokes.cc:482-569)
__device__ void inv_surf_flux(int scheme, double blender, const double ul[5], const double ur[5],
                              const double n[3], double gma, double f[5]) {
    // m = x cross n = (0, -n2, n1)
    double m1 = -n[2], m2 = n[1];
    const double Mm = sqrt(m1 * m1 + m2 * m2);
    double R[3][3];
    if (Mm > 1e-3) {
        m1 /= Mm; m2 /= Mm;
        const double c = n[0], t = 1.0 - c, s = sqrt(1.0 - c * c);
        // Rodrigues rotation about (0, m1, m2)
        R[0][0] = c;             R[0][1] = -s * m2;            R[0][2] = s * m1;
        R[1][0] = s * m2;        R[1][1] = t * m1 * m1 + c;    R[1][2] = t * m1 * m2;
        R[2][0] = -s * m1;       R[2][1] = t * m1 * m2;        R[2][2] = t * m2 * m2 + c;
    } else {
        const double dg = n[0] < 0.0 ? -1.0 : 1.0;
        R[0][0] = dg; R[0][1] = 0; R[0][2] = 0;
        R[1][0] = 0; R[1][1] = dg; R[1][2] = 0;
        R[2][0] = 0; R[2][1] = 0; R[2][2] = dg;
    }
    double lr[5], rr[5], fr[5];
    lr[0] = ul[0]; rr[0] = ur[0]; lr[4] = ul[4]; rr[4] = ur[4];
    for (int a = 0; a < 3; ++a) {  // R^T * momentum
        lr[1 + a] = R[0][a] * ul[1] + R[1][a] * ul[2] + R[2][a] * ul[3];
        rr[1 + a] = R[0][a] * ur[1] + R[1][a] * ur[2] + R[2][a] * ur[3];
    }
    switch (scheme) {
        case 0: hllc_xflux(lr, rr, gma, fr); break;
        case 1: rusanov_xflux(lr, rr, gma, fr); break;
        case 2: ausm_xflux(lr, rr, gma, fr); break;
        default: chandrashekhar_xflux(lr, rr, gma, blender, fr); break;
    }
    f[0] = fr[0];
    f[4] = fr[4];
    for (int a = 0; a < 3; ++a) f[1 + a] = R[a][0] * fr[1] + R[a][1] * fr[2] + R[a][2] * fr[3];
}

// Chandrashekhar symmetric two-point volume flux along dir (navier_stokes.cc:1510-1542)
__device__ __forceinline__ void vol_flux(const double u1[5], const double u2[5], const double dir[3], double gma,
                                         double f[5]) {
    const double p1 = get_p(u1, gma), p2 = get_p(u2, gma);
    const double b1 = 0.5 * u1[0] / p1, b2 = 0.5 * u2[0] / p2, bln = log_avg(b1, b2);
    const double p_hat = 0.5 * (u1[0] + u2[0]) / (b1 + b2);
    const double rho_ln = log_avg(u1[0], u2[0]);
    const double r1 = 1.0 / u1[0], r2 = 1.0 / u2[0];
    double va[3];
    double H_hat = 0.5 / ((gma - 1.0) * bln) + p_hat / rho_ln;
    for (int d = 0; d < 3; ++d) {
        const double v1 = u1[1 + d] * r1, v2 = u2[1 + d] * r2;
        va[d] = 0.5 * (v1 + v2);
        H_hat += va[d] * va[d] - 0.5 * (0.5 * (v1 * v1 + v2 * v2));
    }
    const double vn = va[0] * dir[0] + va[1] * dir[1] + va[2] * dir[2];
    const double rv = rho_ln * vn;
    f[0] = rv;
    f[1] = rv * va[0] + p_hat * dir[0];
    f[2] = rv * va[1] + p_hat * dir[1];
    f[3] = rv * va[2] + p_hat * dir[2];
    f[4] = rv * H_hat;
}

// viscous flux along dir from (state, avars) with avars = (txx,txy,txz,tyy,tyz,tzz,qx,qy,qz)
// (get_dif_flux, navier_stokes.cc:599-624)
__device__ __forceinline__ void dif_flux(const double u[5], const double av[9], const double n[3], double f[5]) {
    const double m0 = av[0] * n[0] + av[1] * n[1] + av[2] * n[2];
    const double m1 = av[1] * n[0] + av[3] * n[1] + av[4] * n[2];
    const double m2 = av[2] * n[0] + av[4] * n[1] + av[5] * n[2];
    const double ri = 1.0 / u[0];
    f[0] = 0.0;
    f[1] = m0;
    f[2] = m1;
    f[3] = m2;
    f[4] = -av[6] * n[0] + u[1] * ri * m0 - av[7] * n[1] + u[2] * ri * m1 - av[8] * n[2] + u[3] * ri * m2;
}

}  // namespace plens_dev

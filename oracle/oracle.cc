// TEST INFRASTRUCTURE ONLY: CPU restatement of the PLENS explicit DGSEM RHS + RK4 update.
// Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
// Each function cites the reference file:line it restates (paths relative to the reference repo).
#include "oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cctype>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_COUNT_FLOPS
#include "flopcount.h"  // Real counts its operations (liboracle_flops.so)
#define FLOP_SCOPE(b) const flopcount::Scope flop_scope_(b)
#define FLOP_BUCKET(b) (flopcount::g_bucket = (b))
#else
using Real = double;    // the oracle proper
#define FLOP_SCOPE(b)
#define FLOP_BUCKET(b)
#endif
// Flop buckets of the counting build: 0 positivity, 1 aux vars (lifting), 2 blender, 3 surface
// fluxes, 4 ho inviscid residual, 5 ho viscous residual, 6 lo inviscid residual, 7 blend of the
// residuals, 8 time step, 9 RK update; nested: 10 two-point volume fluxes (in 4), 11 dense change
// of basis (in 2), 12 stage-2 numerical fluxes (in 3), 13 viscous fluxes of the ho viscous volume
// term (in 5); 14 and 15 the surface parts of the ho inviscid and viscous residuals (after 4 and 5)

namespace {

// elementary functions called unqualified: the std ones for double, and (found by argument-dependent
// lookup) the counting ones of flopcount.h for the counting Real
using std::cos;
using std::exp;
using std::fabs;
using std::log;
using std::pow;
using std::sin;
using std::sqrt;
using std::tan;
using std::tanh;

using State = std::array<Real, 5>;
using Avars = std::array<Real, 9>;

// Planted mistakes (checker hook, oracle_set_mutation): each turns one operation of the time step,
// the inflow/outflow/isothermal-wall ghosts or the blender post-steps into a plausible misreading
// of the reference, so that the tests can show that their closed forms catch it. 0 = none.
enum Mutation {
    kMutNone = 0,
    kMutDtAbs = 1,          // max_d |(J^-T v)_d| instead of the signed maximum (plens.cc:3856-3858)
    kMutDtDegreeSq = 2,     // N^2 in place of N^1.5 (3860)
    kMutDtNoCap = 3,        // Courant >= 1 not capped to 0.9 (3818-3822)
    kMutDtJinv = 4,         // J^-1 v instead of J^-T v (3850-3855)
    kMutInflowS1 = 11,      // uniform inflow stage-1 ghost = prescribed state (uniform_inflow.cc:23-32)
    kMutOutflowSign = 12,   // outflow stage-2 energy correction with p_in - p_pr (outflow.cc:47-55)
    kMutOutflowInviscid = 13,  // outflow stage-2 correction also for inviscid flow (outflow.cc:48)
    kMutWallRho = 14,       // isothermal wall prescribed density from the inner temperature (uniform_temp_wall.cc:25-27)
    kMutWallAbsVel = 15,    // isothermal wall stage-2 reverses the absolute momentum (uniform_temp_wall.cc:67-68)
    kMutBlendLive = 21,     // diffusion reads the neighbours' diffused values (plens.cc:2926-2936 reads gh)
    kMutBlendPeriodic = 22, // diffusion across periodic faces (plens.cc:2928 skips every boundary face)
    kMutBlendOverrideLate = 23,  // wall override after the diffusion (plens.cc:2909-2920 comes first)
    kMutBlendNoCap = 24     // alpha_d = alpha / alpha_max without min(1, .) (plens.cc:2953-2957)
};
int g_mutation = kMutNone;
struct T1 {  // rank-1 tensor in 3D (deal.II Tensor<1,3>)
    Real v[3] = {0, 0, 0};
    Real& operator[](int i) { return v[i]; }
    Real operator[](int i) const { return v[i]; }
    Real norm() const { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
};
inline Real dot(const T1& a, const T1& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline T1 cross(const T1& a, const T1& b) {
    T1 c;
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    return c;
}

constexpr Real R0 = 8.314462618;  // modelling/navier_stokes.h:149

// ---------------------------------------------------------------------------------------------
// Expressions of the varying inflow BC (IBCs/varying_inflow.cc): deal.II's FunctionParser over
// muParser, variables x, y, z, t; evaluated directly while parsing. muParser's precedence, lowest
// first: ?: (right associative), ||, &&, comparisons, + -, * /, unary signs, ^ (right
// associative, above the signs: -2^2 = -4). Functions of muParser (sin ... abs, log = ln, min,
// max, sum, avg, atan2) and deal.II's additions (if, pow, floor, ceil, int, cot, csc, sec, erfc);
// constants _pi and _e.
// ---------------------------------------------------------------------------------------------
struct ExprEval {
    const std::string& s;
    size_t i;
    Real x, y, z, t;
    bool ok = true;
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    bool eat(const char* tok) {
        ws();
        const size_t n = std::strlen(tok);
        if (s.compare(i, n, tok) == 0) { i += n; return true; }
        return false;
    }
    Real expr() {  // ternary
        const Real c = lor();
        if (eat("?")) {
            const Real a = expr();
            if (!eat(":")) ok = false;
            const Real b = expr();
            return c != 0 ? a : b;
        }
        return c;
    }
    Real lor() { Real a = land(); while (eat("||")) { const Real b = land(); a = (a != 0 || b != 0); } return a; }
    Real land() { Real a = cmp(); while (eat("&&")) { const Real b = cmp(); a = (a != 0 && b != 0); } return a; }
    Real cmp() {
        Real a = sum();
        for (;;) {
            if (eat("<=")) { const Real b = sum(); a = a <= b; }
            else if (eat(">=")) { const Real b = sum(); a = a >= b; }
            else if (eat("==")) { const Real b = sum(); a = a == b; }
            else if (eat("!=")) { const Real b = sum(); a = a != b; }
            else if (eat("<")) { const Real b = sum(); a = a < b; }
            else if (eat(">")) { const Real b = sum(); a = a > b; }
            else return a;
        }
    }
    Real sum() {
        Real a = prod();
        for (;;) {
            if (eat("+")) a = a + prod();
            else if (eat("-")) a = a - prod();
            else return a;
        }
    }
    Real prod() {
        Real a = sign();
        for (;;) {
            if (eat("*")) a = a * sign();
            else if (eat("/")) a = a / sign();
            else return a;
        }
    }
    Real sign() {
        if (eat("-")) return -sign();
        if (eat("+")) return sign();
        const Real b = atom();
        if (eat("^")) return pow(b, sign());
        return b;
    }
    Real atom() {
        ws();
        if (eat("(")) { const Real v = expr(); if (!eat(")")) ok = false; return v; }
        if (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.')) {
            size_t used = 0;
            const Real v = std::stod(s.substr(i), &used);
            i += used;
            return v;
        }
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_')) ++i;
        const std::string id = s.substr(b, i - b);
        if (id.empty()) { ok = false; return 0; }
        if (eat("(")) {
            std::vector<Real> a;
            ws();
            if (!eat(")")) {
                do a.push_back(expr()); while (eat(","));
                if (!eat(")")) ok = false;
            }
            return call(id, a);
        }
        if (id == "x") return x;
        if (id == "y") return y;
        if (id == "z") return z;
        if (id == "t") return t;
        if (id == "_pi") return M_PI;
        if (id == "_e") return M_E;
        ok = false;
        return 0;
    }
    Real call(const std::string& f, const std::vector<Real>& a) {
        const size_t n = a.size();
        auto one = [&](Real v) { if (n != 1) ok = false; return v; };
        const Real a0 = n ? a[0] : Real(0);
        if (f == "sin") return one(sin(a0));
        if (f == "cos") return one(cos(a0));
        if (f == "tan") return one(tan(a0));
        if (f == "asin") return one(std::asin(a0));
        if (f == "acos") return one(std::acos(a0));
        if (f == "atan") return one(std::atan(a0));
        if (f == "sinh") return one(std::sinh(a0));
        if (f == "cosh") return one(std::cosh(a0));
        if (f == "tanh") return one(tanh(a0));
        if (f == "asinh") return one(std::asinh(a0));
        if (f == "acosh") return one(std::acosh(a0));
        if (f == "atanh") return one(std::atanh(a0));
        if (f == "log2") return one(std::log2(a0));
        if (f == "log10") return one(std::log10(a0));
        if (f == "log" || f == "ln") return one(log(a0));
        if (f == "exp") return one(exp(a0));
        if (f == "sqrt") return one(sqrt(a0));
        if (f == "sign") return one(a0 > 0 ? 1.0 : (a0 < 0 ? -1.0 : 0.0));
        if (f == "rint") return one(std::rint(a0));
        if (f == "abs") return one(fabs(a0));
        if (f == "floor") return one(std::floor(a0));
        if (f == "ceil") return one(std::ceil(a0));
        if (f == "int") return one((Real)(long long)a0);
        if (f == "cot") return one(1 / tan(a0));
        if (f == "csc") return one(1 / sin(a0));
        if (f == "sec") return one(1 / cos(a0));
        if (f == "erfc") return one(std::erfc(a0));
        if (f == "pow" && n == 2) return pow(a[0], a[1]);
        if (f == "atan2" && n == 2) return std::atan2(a[0], a[1]);
        if (f == "if" && n == 3) return a[0] != 0 ? a[1] : a[2];
        if (n >= 1 && (f == "min" || f == "max" || f == "sum" || f == "avg")) {
            Real r = a[0];
            for (size_t k = 1; k < n; ++k) {
                if (f == "min") r = std::min(r, a[k]);
                else if (f == "max") r = std::max(r, a[k]);
                else r += a[k];
            }
            return f == "avg" ? r / n : r;
        }
        ok = false;
        return 0;
    }
};

// value of expression `src` at (x, y, z, t); NaN on a syntax error
Real eval_expression(const std::string& src, Real x, Real y, Real z, Real t) {
    ExprEval e{src, 0, x, y, z, t};
    const Real v = e.expr();
    e.ws();
    return (e.ok && e.i == src.size()) ? v : Real(std::nan(""));
}

// components of "a; b; c"
std::vector<std::string> split_components(const char* s) {
    std::vector<std::string> out(1);
    for (; s && *s; ++s) {
        if (*s == ';') out.emplace_back();
        else out.back() += *s;
    }
    return out;
}

// ---------------------------------------------------------------------------------------------
// NavierStokes restatement (modelling/navier_stokes.{h,cc})
// ---------------------------------------------------------------------------------------------
struct NS {
    Real gma, M, Pr, mu0, T0, S;
    int scheme;
    Real blender = 1;  // flux_blender_value (ctor default 1, navier_stokes.cc:33)
    int vscheme = 0;     // inv_vol_flux_scheme: 0 Chandrashekhar, 1 Kennedy-Gruber, 2 Ismail-Roe

    Real R() const { return R0 / M; }                       // navier_stokes.h:337
    Real cv() const { return R() / (gma - 1); }             // navier_stokes.h:342
    Real mu(Real T) const {                               // navier_stokes.h:348-355
        return mu0 * pow(T / T0, 1.5) * (T0 + S) / (T + S);
    }
    Real kappa(Real m) const { return m * R() * gma / ((gma - 1) * Pr); }  // .h:360
    bool inviscid() const { return mu0 < 1e-16; }             // .h:438
    void set_blender(Real a) {                              // .h:443-448
        if (a <= 0) blender = 0;
        else if (a <= 1) blender = a;
        else blender = 1;
    }
    static Real get_e(const State& c) {                     // navier_stokes.cc:349-359
        Real ske = 0;
        for (int d = 0; d < 3; ++d) ske += c[1 + d] * c[1 + d];
        const Real ri = 1 / c[0];
        ske *= 0.5 * ri;
        return (c[4] - ske) * ri;
    }
    Real get_p(const State& c) const {                      // navier_stokes.cc:375-384
        Real ske = 0;
        for (int d = 0; d < 3; ++d) ske += c[1 + d] * c[1 + d];
        ske *= 0.5 / c[0];
        return (gma - 1) * (c[4] - ske);
    }
    Real get_a(const State& c) const { return sqrt(gma * get_p(c) / c[0]); }  // .cc:396-399
    void cons_to_entropy(const State& c, State& w) const {    // navier_stokes.h:404-430
        const Real p = get_p(c);
        const Real s = log(p) - gma * log(c[0]);
        const Real beta = 0.5 * c[0] / p;
        const Real rhoinv = 1 / c[0];
        Real v[3], vv = 0;
        for (int d = 0; d < 3; ++d) {
            v[d] = c[1 + d] * rhoinv;
            vv += v[d] * v[d];
        }
        w[0] = (gma - s) / (gma - 1) - beta * vv;
        for (int d = 0; d < 3; ++d) w[1 + d] = 2 * beta * v[d];
        w[4] = -2 * beta;
    }
    Real get_M(const State& c) const {                      // .cc:408-414
        Real s = 0;
        const Real ri = 1 / c[0];
        for (int d = 0; d < 3; ++d) s += (c[1 + d] * c[1 + d]) * (ri * ri);
        return sqrt(s) / get_a(c);
    }
    void prim_to_cons(Real rho, const T1& vel, Real p, State& c) const {  // .h:365-378
        c[0] = rho;
        c[4] = p / (gma - 1);
        for (int d = 0; d < 3; ++d) {
            c[1 + d] = rho * vel[d];
            c[4] += 0.5 * rho * vel[d] * vel[d];
        }
    }
    void get_inv_flux(const State& c, const T1& dir, State& f) const {  // .cc:424-442
        T1 vel;
        const Real ri = 1.0 / c[0];
        for (int d = 0; d < 3; ++d) vel[d] = c[1 + d] * ri;
        const Real p = get_p(c);
        const Real vn = dot(vel, dir);
        f[0] = c[0] * vn;
        for (int d = 0; d < 3; ++d) f[1 + d] = vn * c[1 + d] + p * dir[d];
        f[4] = (c[4] + p) * vn;
    }
    static Real log_avg(Real xl, Real xr) {             // navier_stokes.h:494-501
        const Real xi = xl / xr, f = (xi - 1) / (xi + 1), u = f * f,
                     F = (u < 1e-2 ? 1 + u / 3 + u * u / 5 + u * u * u / 7 : 0.5 * log(xi) / f);
        return 0.5 * (xl + xr) / F;
    }

    // HLLC, navier_stokes.cc:936-1033 (pressure-based wave speeds, Toro 10.6)
    void hllc_xflux(const State& l, const State& r, State& f) const {
        T1 x; x[0] = 1;
        const Real pl = get_p(l), pr = get_p(r);
        const Real pmin = std::min(pl, pr), pmax = std::max(pl, pr);
        const Real ul = l[1] / l[0], ur = r[1] / r[0];
        const Real al = sqrt(gma * pl / l[0]), ar = sqrt(gma * pr / r[0]);
        const Real p_pvrs = 0.5 * (pl + pr) - 0.5 * (ur - ul) * 0.5 * (l[0] + r[0]) * 0.5 * (al + ar);
        Real ps;
        if (pmax / pmin < 2 && pmin <= p_pvrs && p_pvrs <= pmax) {
            ps = p_pvrs;
        } else if (p_pvrs < pmin) {
            const Real z = (gma - 1) / (2 * gma);
            ps = pow(((al + ar) - (gma - 1) / 2 * (ur - ul)) /
                              (al / pow(pl, z) + ar / pow(pr, z)),
                          1 / z);
        } else {
            const Real p0 = std::max(Real(0.0), p_pvrs);
            const Real gl = sqrt(2 / ((gma + 1) * l[0]) / (p0 + (gma - 1) / (gma + 1) * pl)),
                         gr = sqrt(2 / ((gma + 1) * r[0]) / (p0 + (gma - 1) / (gma + 1) * pr));
            ps = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
        }
        const Real ql = (ps <= pl ? Real(1) : sqrt(1 + (gma + 1) / (2 * gma) * (ps / pl - 1))),
                     qr = (ps <= pr ? Real(1) : sqrt(1 + (gma + 1) / (2 * gma) * (ps / pr - 1)));
        const Real sl = ul - al * ql, sr = ur + ar * qr;
        const Real s = (pr - pl + l[1] * (sl - ul) - r[1] * (sr - ur)) / (l[0] * (sl - ul) - r[0] * (sr - ur));
        if (sl > 0) {
            get_inv_flux(l, x, f);
        } else if (s > 0) {
            const Real t = (sl - ul) / (sl - s);
            const State ss = {t * l[0], t * l[0] * s, t * l[2], t * l[3],
                              t * (l[4] + (s - ul) * (l[0] * s + pl / (sl - ul)))};
            State lf;
            get_inv_flux(l, x, lf);
            for (int v = 0; v < 5; ++v) f[v] = lf[v] + sl * (ss[v] - l[v]);
        } else if (sr > 0) {
            const Real t = (sr - ur) / (sr - s);
            const State ss = {t * r[0], t * r[0] * s, t * r[2], t * r[3],
                              t * (r[4] + (s - ur) * (r[0] * s + pr / (sr - ur)))};
            State rf;
            get_inv_flux(r, x, rf);
            for (int v = 0; v < 5; ++v) f[v] = rf[v] + sr * (ss[v] - r[v]);
        } else {
            get_inv_flux(r, x, f);
        }
    }
    // Rusanov, navier_stokes.cc:1044-1059
    void rusanov_xflux(const State& l, const State& r, State& f) const {
        T1 x; x[0] = 1;
        State lf, rf;
        const Real al = get_a(l), ar = get_a(r);
        const Real ul = l[1] / l[0], ur = r[1] / r[0];
        get_inv_flux(l, x, lf);
        get_inv_flux(r, x, rf);
        const Real S = std::max(fabs(ul) + al, fabs(ur) + ar);
        for (int v = 0; v < 5; ++v) f[v] = 0.5 * (lf[v] + rf[v]) - 0.5 * S * (r[v] - l[v]);
    }
    // AUSM+-up, navier_stokes.cc:1070-1121 with modelling/ausm_functions.h; note the
    // integer division 3/16 == 0 of the reference (navier_stokes.cc:1090) is kept
    static Real m1p(Real M) { return 0.5 * (M + fabs(M)); }
    static Real m1n(Real M) { return 0.5 * (M - fabs(M)); }
    static Real m2p(Real M) { return 0.25 * (M + 1) * (M + 1); }
    static Real m2n(Real M) { return -0.25 * (M - 1) * (M - 1); }
    static Real m4p(Real M) { return fabs(M) >= 1 ? m1p(M) : m2p(M) * (1 - 16 * 0.125 * m2n(M)); }
    static Real m4n(Real M) { return fabs(M) >= 1 ? m1n(M) : m2n(M) * (1 + 16 * 0.125 * m2p(M)); }
    static Real p5p(Real M, Real a) {
        return fabs(M) >= 1 ? m1p(M) / M : m2p(M) * ((2 - M) - 16 * a * M * m2n(M));
    }
    static Real p5n(Real M, Real a) {
        return fabs(M) >= 1 ? m1n(M) / M : m2n(M) * ((-2 - M) + 16 * a * M * m2p(M));
    }
    void ausm_xflux(const State& l, const State& r, State& f) const {
        const Real Minf = 0.3, Kp = 0.25, Ku = 0.75, sigma = 1.0;
        const Real pl = get_p(l), pr = get_p(r);
        T1 vl, vr;
        for (int d = 0; d < 3; ++d) { vl[d] = l[1 + d] / l[0]; vr[d] = r[1 + d] / r[0]; }
        const Real Hl = (l[4] + pl) / l[0], Hr = (r[4] + pr) / r[0];
        const Real a12 = 0.5 * (get_a(l) + get_a(r));
        const Real Ml = vl[0] / a12, Mr = vr[0] / a12;
        const Real Mavgsq = 0.5 * (Ml * Ml + Mr * Mr);
        const Real Mo = sqrt(std::min(Real(1.0), std::max(Mavgsq, Minf * Minf)));
        const Real fa = Mo * (2 - Mo), alpha = 3 / 16 * (-4 + 5 * fa * fa);
        const Real M12 = m4p(Ml) + m4n(Mr) -
                           Kp / fa * std::max(1 - sigma * Mavgsq, Real(0.0)) * 2 * (pr - pl) / (a12 * a12 * (l[0] + r[0]));
        const Real psp = p5p(Ml, alpha), psn = p5n(Mr, alpha);
        const Real p12 = psp * pl + psn * pr - Ku * fa * psp * psn * (l[0] + r[0]) * a12 * (vr[0] - vl[0]);
        if (M12 > 0) {
            const Real m12 = a12 * M12 * l[0];
            f[0] = m12; f[1] = m12 * vl[0] + p12; f[2] = m12 * vl[1]; f[3] = m12 * vl[2]; f[4] = m12 * Hl;
        } else {
            const Real m12 = a12 * M12 * r[0];
            f[0] = m12; f[1] = m12 * vr[0] + p12; f[2] = m12 * vr[1]; f[3] = m12 * vr[2]; f[4] = m12 * Hr;
        }
    }
    // x eigenvector matrices, navier_stokes.cc:641-678 and 780-825
    void get_xK(const T1& vel, Real a, Real H, Real K[5][5]) const {
        K[0][0] = 1;
        for (int d = 0; d < 3; ++d) K[1 + d][0] = vel[d];
        K[1][0] -= a;
        K[4][0] = H - vel[0] * a;
        K[0][1] = 1;
        for (int d = 0; d < 3; ++d) K[1 + d][1] = vel[d];
        K[4][1] = 0.5 * dot(vel, vel);
        K[0][2] = 0; K[1][2] = 0; K[2][2] = 1; K[3][2] = 0; K[4][2] = vel[1];
        K[0][3] = 0; K[1][3] = 0; K[2][3] = 0; K[3][3] = 1; K[4][3] = vel[2];
        K[0][4] = 1;
        for (int d = 0; d < 3; ++d) K[1 + d][4] = vel[d];
        K[1][4] += a;
        K[4][4] = H + vel[0] * a;
    }
    void get_xKinv(const T1& vel, Real a, Real H, Real Ki[5][5]) const {
        const Real t = 1 / (gma - 1);
        Ki[0][0] = H + a * (vel[0] - a) * t;
        for (int d = 0; d < 3; ++d) Ki[0][1 + d] = -vel[d];
        Ki[0][1] -= a * t;
        Ki[0][4] = 1;
        Ki[1][0] = 2 * (2 * a * a * t - H);
        for (int d = 0; d < 3; ++d) Ki[1][1 + d] = 2 * vel[d];
        Ki[1][4] = -2;
        Ki[2][0] = -2 * vel[1] * a * a * t; Ki[2][1] = 0; Ki[2][2] = 2 * a * a * t; Ki[2][3] = 0; Ki[2][4] = 0;
        Ki[3][0] = -2 * vel[2] * a * a * t; Ki[3][1] = 0; Ki[3][2] = 0; Ki[3][3] = 2 * a * a * t; Ki[3][4] = 0;
        Ki[4][0] = H - a * (vel[0] + a) * t;
        for (int d = 0; d < 3; ++d) Ki[4][1 + d] = -vel[d];
        Ki[4][1] += a * t;
        Ki[4][4] = 1;
        const Real fac = 0.5 * (gma - 1) / (a * a);
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) Ki[i][j] *= fac;
    }
    // Chandrashekhar EC flux with hybrid matrix dissipation, navier_stokes.cc:1250-1318
    void chandrashekhar_xflux(const State& l, const State& r, State& f) const {
        const Real pl = get_p(l), pr = get_p(r);
        const Real bl = 0.5 * l[0] / pl, br = 0.5 * r[0] / pr, bln = log_avg(bl, br);
        const Real p_hat = 0.5 * (l[0] + r[0]) / (bl + br);
        const Real rho_ln = log_avg(l[0], r[0]);
        T1 va, vsq, vl, vr;
        const Real ril = 1 / l[0], rir = 1 / r[0];
        for (int d = 0; d < 3; ++d) {
            vl[d] = l[1 + d] * ril;
            vr[d] = r[1 + d] * rir;
            va[d] = 0.5 * (vl[d] + vr[d]);
            vsq[d] = 0.5 * (vl[d] * vl[d] + vr[d] * vr[d]);
        }
        Real H_hat = 0.5 / ((gma - 1) * bln) + p_hat / rho_ln;
        for (int d = 0; d < 3; ++d) H_hat += va[d] * va[d] - 0.5 * vsq[d];
        f[0] = rho_ln * va[0];
        f[1] = rho_ln * va[0] * va[0] + p_hat;
        f[2] = rho_ln * va[0] * va[1];
        f[3] = rho_ln * va[0] * va[2];
        f[4] = rho_ln * va[0] * H_hat;
        const Real sql = sqrt(l[0]), sqr = sqrt(r[0]);
        T1 vi;
        const Real inv = 1 / (sql + sqr);
        for (int d = 0; d < 3; ++d) vi[d] = inv * (sql * vl[d] + sqr * vr[d]);
        const Real Hi = ((l[4] + pl) / sql + (r[4] + pr) / sqr) / (sql + sqr);
        const Real ai = sqrt((gma - 1) * (Hi - 0.5 * dot(vi, vi)));
        Real K[5][5], Ki[5][5], A[5][5];
        get_xK(vi, ai, Hi, K);
        get_xKinv(vi, ai, Hi, Ki);
        const Real uabs = fabs(vi[0]), lmax = uabs + ai,
                     l2 = blender * lmax + (1 - blender) * uabs;
        const Real eig[5] = {lmax, l2, l2, l2, lmax};
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
                A[i][j] = 0;
                for (int k = 0; k < 5; ++k) A[i][j] += K[i][k] * eig[k] * Ki[k][j];
            }
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) f[i] -= 0.5 * A[i][j] * (r[j] - l[j]);
    }
    // Rusanov blends: Rusanov weighted by flux_blender_value, navier_stokes.cc:1132-1160, 1364-1374
    void rusanov_blend_xflux(const State& l, const State& r, int other, State& f) const {
        State fr, fo;
        rusanov_xflux(l, r, fr);
        if (other == 0) hllc_xflux(l, r, fo);
        else if (other == 2) ausm_xflux(l, r, fo);
        else kennedy_gruber_xflux(l, r, fo);
        for (int v = 0; v < 5; ++v) f[v] = blender * fr[v] + (1 - blender) * fo[v];
    }
    // modified Steger-Warming, navier_stokes.cc:1174-1236 (pos_smooth/neg_smooth, .h:464-476)
    void modified_sw_xflux(const State& l, const State& r, State& f) const {
        const Real pl = get_p(l), pr = get_p(r);
        const Real q = 0.5 * (pr - pl) / std::min(pl, pr);
        const Real om = 0.5 / (1 + q * q);
        T1 vp, vn;
        for (int d = 0; d < 3; ++d) {
            vp[d] = (1 - om) * l[1 + d] / l[0] + om * r[1 + d] / r[0];
            vn[d] = (1 - om) * r[1 + d] / r[0] + om * l[1 + d] / l[0];
        }
        const Real rp = (1 - om) * l[0] + om * r[0], rn = (1 - om) * r[0] + om * l[0];
        const Real pp = (1 - om) * pl + om * pr, pn = (1 - om) * pr + om * pl;
        const Real ap = sqrt(gma * pp / rp), an = sqrt(gma * pn / rn);
        const Real Hp = gma * pp / ((gma - 1) * rp) + 0.5 * dot(vp, vp),
                     Hn = gma * pn / ((gma - 1) * rn) + 0.5 * dot(vn, vn);
        Real Kp[5][5], Kip[5][5], Kn[5][5], Kin[5][5];
        get_xK(vp, ap, Hp, Kp);
        get_xK(vn, an, Hn, Kn);
        get_xKinv(vp, ap, Hp, Kip);
        get_xKinv(vn, an, Hn, Kin);
        auto ps = [](Real x, Real e) { return 0.5 * (x + sqrt(x * x + e * e)); };
        auto ng = [](Real x, Real e) { return 0.5 * (x - sqrt(x * x + e * e)); };
        const Real ep = 0.3 * ap, en = 0.3 * an;
        const Real lp[5] = {ps(vp[0] - ap, ep), ps(vp[0], ep), ps(vp[0], ep), ps(vp[0], ep), ps(vp[0] + ap, ep)};
        const Real ln[5] = {ng(vn[0] - an, en), ng(vn[0], en), ng(vn[0], en), ng(vn[0], en), ng(vn[0] + an, en)};
        for (int i = 0; i < 5; ++i) {
            f[i] = 0;
            for (int j = 0; j < 5; ++j) {
                Real Ap = 0, An = 0;
                for (int k = 0; k < 5; ++k) {
                    Ap += Kp[i][k] * lp[k] * Kip[k][j];
                    An += Kn[i][k] * ln[k] * Kin[k][j];
                }
                f[i] += Ap * l[j] + An * r[j];
            }
        }
    }
    // Kennedy-Gruber with a local Lax-Friedrichs term, navier_stokes.cc:1325-1357
    void kennedy_gruber_xflux(const State& l, const State& r, State& f) const {
        T1 x; x[0] = 1;
        kennedy_gruber_vol_flux(l, r, x, f);
        const Real el = get_e(l), er = get_e(r);
        const Real lm = std::max(fabs(l[1] / l[0]) + sqrt(gma * (gma - 1) * el),
                                   fabs(r[1] / r[0]) + sqrt(gma * (gma - 1) * er));
        for (int v = 0; v < 5; ++v) f[v] -= 0.5 * lm * (r[v] - l[v]);
    }
    // Ismail-Roe with entropy-variable matrix dissipation, navier_stokes.cc:1381-1446
    void ismail_roe_xflux(const State& l, const State& r, State& f) const {
        T1 x; x[0] = 1;
        Real rh, vh[3], p2, hh;
        ismail_roe_means(l, r, rh, vh, p2, hh);
        ismail_roe_vol_flux(l, r, x, f);
        State ch;
        ch[0] = rh;
        for (int d = 0; d < 3; ++d) ch[1 + d] = rh * vh[d];
        ch[4] = rh * hh - p2;
        State Wl, Wr;
        cons_to_entropy(l, Wl);
        cons_to_entropy(r, Wr);
        Real H[5][5];
        for (int i = 0; i < 5; ++i) H[i][0] = ch[i];
        for (int d = 0; d < 3; ++d) {
            T1 e; e[d] = 1;
            State fl;
            get_inv_flux(ch, e, fl);
            for (int i = 0; i < 5; ++i) H[i][1 + d] = fl[i];
        }
        for (int i = 0; i < 4; ++i) H[i][4] = H[4][i];
        H[4][4] = rh * hh * hh - gma * p2 * p2 / (rh * (gma - 1));
        const Real lm = fabs(vh[0]) + sqrt(gma * p2 / rh);
        for (int i = 0; i < 5; ++i) {
            Real s = 0;
            for (int j = 0; j < 5; ++j) s += H[i][j] * (Wr[j] - Wl[j]);
            f[i] -= 0.5 * lm * s;
        }
    }
    void xflux(const State& l, const State& r, State& f) const {  // set_inv_surf_flux_scheme, .cc:140-197
        switch (scheme) {
            case 0: hllc_xflux(l, r, f); break;
            case 1: rusanov_xflux(l, r, f); break;
            case 2: ausm_xflux(l, r, f); break;
            case 3: chandrashekhar_xflux(l, r, f); break;
            case 4: rusanov_blend_xflux(l, r, 0, f); break;
            case 5: rusanov_blend_xflux(l, r, 2, f); break;
            case 6: modified_sw_xflux(l, r, f); break;
            case 7: kennedy_gruber_xflux(l, r, f); break;
            case 8: rusanov_blend_xflux(l, r, 7, f); break;
            default: ismail_roe_xflux(l, r, f); break;
        }
    }
    // rotation to the normal-aligned frame, navier_stokes.cc:482-569
    void get_inv_surf_flux(const State& o, const State& nb, const T1& n, State& f) const {
        T1 x; x[0] = 1;
        T1 m = cross(x, n);
        const Real Mm = m.norm();
        Real R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (Mm > 1e-3) {
            for (int d = 0; d < 3; ++d) m[d] /= Mm;
            const Real c = dot(x, n), t = 1 - c, s = sqrt(1 - c * c);
            R[0][0] = t * m[0] * m[0] + c;     R[0][1] = t * m[0] * m[1] - s * m[2]; R[0][2] = t * m[0] * m[2] + s * m[1];
            R[1][0] = t * m[0] * m[1] + s * m[2]; R[1][1] = t * m[1] * m[1] + c;     R[1][2] = t * m[1] * m[2] - s * m[0];
            R[2][0] = t * m[0] * m[2] - s * m[1]; R[2][1] = t * m[1] * m[2] + s * m[0]; R[2][2] = t * m[2] * m[2] + c;
        } else {
            const Real dg = n[0] < 0 ? -1 : 1;
            for (int d = 0; d < 3; ++d) R[d][d] = dg;
        }
        T1 om, nm, omr, nmr;
        for (int d = 0; d < 3; ++d) { om[d] = o[1 + d]; nm[d] = nb[1 + d]; }
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col < 3; ++col) {
                const Real RT = R[col][row];
                omr[row] += RT * om[col];
                nmr[row] += RT * nm[col];
            }
        State orr, nrr;
        orr[0] = o[0]; nrr[0] = nb[0];
        for (int d = 0; d < 3; ++d) { orr[1 + d] = omr[d]; nrr[1 + d] = nmr[d]; }
        orr[4] = o[4]; nrr[4] = nb[4];
        State fr;
        xflux(orr, nrr, fr);
        T1 mfr, mf;
        for (int d = 0; d < 3; ++d) mfr[d] = fr[1 + d];
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col < 3; ++col) mf[row] += R[row][col] * mfr[col];
        f[0] = fr[0];
        for (int d = 0; d < 3; ++d) f[1 + d] = mf[d];
        f[4] = fr[4];
    }
    void vol_flux(const State& c1, const State& c2, const T1& dir, State& f) const {  // .cc:203-228
        FLOP_SCOPE(10);
        if (vscheme == 1) kennedy_gruber_vol_flux(c1, c2, dir, f);
        else if (vscheme == 2) ismail_roe_vol_flux(c1, c2, dir, f);
        else chandrashekhar_vol_flux(c1, c2, dir, f);
    }
    // Kennedy-Gruber volume flux, navier_stokes.cc:1549-1575
    void kennedy_gruber_vol_flux(const State& c1, const State& c2, const T1& dir, State& f) const {
        T1 v1, v2, va;
        for (int d = 0; d < 3; ++d) {
            v1[d] = c1[1 + d] / c1[0];
            v2[d] = c2[1 + d] / c2[0];
            va[d] = 0.5 * (v1[d] + v2[d]);
        }
        const Real ra = 0.5 * (c1[0] + c2[0]);
        const Real e1 = get_e(c1), e2 = get_e(c2);
        const Real Ea = 0.5 * (e1 + e2 + 0.5 * (dot(v1, v1) + dot(v2, v2)));
        const Real pa = (gma - 1) * 0.5 * (c1[0] * e1 + c2[0] * e2);
        const Real vn = dot(va, dir);
        f[0] = vn * ra;
        for (int d = 0; d < 3; ++d) f[1 + d] = ra * vn * va[d] + pa * dir[d];
        f[4] = vn * (ra * Ea + pa);
    }
    // Ismail-Roe means through z = sqrt(rho/p) (1, v, p), navier_stokes.cc:1590-1610
    void ismail_roe_means(const State& c1, const State& c2, Real& rh, Real vh[3], Real& p2,
                          Real& hh) const {
        const Real a1 = get_a(c1), a2 = get_a(c2), g = sqrt(gma);
        const Real z11 = g / a1, z12 = g / a2, z51 = c1[0] * a1 / g, z52 = c2[0] * a2 / g;
        const Real z5ln = log_avg(z51, z52);
        rh = 0.5 * (z11 + z12) * z5ln;
        const Real p1 = (z51 + z52) / (z11 + z12);
        p2 = 0.5 * ((gma + 1) * z5ln / log_avg(z11, z12) + (gma - 1) * p1) / gma;
        hh = gma * p2 / ((gma - 1) * rh);
        for (int d = 0; d < 3; ++d) {
            vh[d] = (z11 * c1[1 + d] / c1[0] + z12 * c2[1 + d] / c2[0]) / (z11 + z12);
            hh += 0.5 * vh[d] * vh[d];
        }
    }
    // Ismail-Roe volume flux, navier_stokes.cc:1590-1622
    void ismail_roe_vol_flux(const State& c1, const State& c2, const T1& dir, State& f) const {
        const Real a1 = get_a(c1), a2 = get_a(c2), g = sqrt(gma);
        const Real p1 = (c1[0] * a1 / g + c2[0] * a2 / g) / (g / a1 + g / a2);
        Real rh, vh[3], p2, hh;
        ismail_roe_means(c1, c2, rh, vh, p2, hh);
        const Real vn = vh[0] * dir[0] + vh[1] * dir[1] + vh[2] * dir[2];
        f[0] = rh * vn;
        for (int d = 0; d < 3; ++d) f[1 + d] = rh * vh[d] * vn + p1 * dir[d];
        f[4] = rh * vn * hh;
    }
    // Chandrashekhar volume flux, navier_stokes.cc:1510-1542
    void chandrashekhar_vol_flux(const State& c1, const State& c2, const T1& dir, State& f) const {
        const Real p1 = get_p(c1), p2 = get_p(c2);
        const Real b1 = 0.5 * c1[0] / p1, b2 = 0.5 * c2[0] / p2, bln = log_avg(b1, b2);
        const Real p_hat = 0.5 * (c1[0] + c2[0]) / (b1 + b2);
        const Real rho_ln = log_avg(c1[0], c2[0]);
        T1 va, vsq;
        const Real r1 = 1 / c1[0], r2 = 1 / c2[0];
        for (int d = 0; d < 3; ++d) {
            const Real v1 = c1[1 + d] * r1, v2 = c2[1 + d] * r2;
            va[d] = 0.5 * (v1 + v2);
            vsq[d] = 0.5 * (v1 * v1 + v2 * v2);
        }
        Real H_hat = 0.5 / ((gma - 1) * bln) + p_hat / rho_ln;
        for (int d = 0; d < 3; ++d) H_hat += va[d] * va[d] - 0.5 * vsq[d];
        const Real vn = dot(va, dir);
        f[0] = vn * rho_ln;
        for (int d = 0; d < 3; ++d) f[1 + d] = rho_ln * vn * va[d] + p_hat * dir[d];
        f[4] = rho_ln * vn * H_hat;
    }
    // diffusive flux, navier_stokes.cc:599-624 (stress tensor from avars, .cc:578-589)
    static void dif_flux(const State& c, const Avars& av, const T1& dir, State& f) {
        Real st[3][3];
        int i = 0;
        for (int row = 0; row < 3; ++row)
            for (int col = row; col < 3; ++col) {
                st[row][col] = av[i];
                if (col != row) st[col][row] = av[i];
                ++i;
            }
        T1 mf;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) mf[a] += st[a][b] * dir[b];
        f[0] = 0;
        f[4] = 0;
        const Real ri = 1.0 / c[0];
        for (int d = 0; d < 3; ++d) {
            const Real v = c[1 + d] * ri;
            f[1 + d] = mf[d];
            f[4] += -av[6 + d] * dir[d] + v * mf[d];
        }
    }
};

// ---------------------------------------------------------------------------------------------
// 1D bases (QGaussLobatto / FE_DGQ<1> shape_grad / Legendre / QGauss), computed here in long
// Real by Newton on P_N' using the Legendre ODE for P_N''.
// ---------------------------------------------------------------------------------------------
void leg(int n, long double x, long double& P, long double& dP) {
    long double p0 = 1, p1 = x;
    if (n == 0) { P = 1; dP = 0; return; }
    for (int k = 2; k <= n; ++k) {
        const long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1; p1 = p2;
    }
    P = p1;
    dP = n * (x * p1 - p0) / (x * x - 1);
}
long double leg_val(int n, long double x) {
    long double p0 = 1, p1 = x;
    if (n == 0) return 1;
    for (int k = 2; k <= n; ++k) {
        const long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1; p1 = p2;
    }
    return p1;
}
void gll(int N, std::vector<Real>& x01, std::vector<Real>& w01) {
    std::vector<long double> x(N + 1);
    x[0] = -1; x[N] = 1;
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int j = 1; j < N; ++j) {
        long double t = -cosl(pi * j / N);
        for (int it = 0; it < 200; ++it) {  // Newton on P_N'(t) with P_N'' from the ODE
            long double P, dP;
            leg(N, t, P, dP);
            const long double d2P = (2 * t * dP - N * (N + 1) * P) / (1 - t * t);
            const long double dt = dP / d2P;
            t -= dt;
            if (fabsl(dt) < 1e-32L) break;
        }
        x[j] = t;
    }
    x01.resize(N + 1); w01.resize(N + 1);
    for (int j = 0; j <= N; ++j) {
        const long double P = leg_val(N, x[j]);
        x01[j] = (Real)((x[j] + 1) / 2);
        w01[j] = (Real)(1.0L / (N * (N + 1) * P * P));
    }
}
// FE_DGQ<1>::shape_grad(j, x_i): derivative of the Lagrange polynomial l_j at x_i
Real lagrange_deriv(const std::vector<Real>& xn, int j, long double x) {
    const int np = (int)xn.size();
    long double s = 0;
    for (int k = 0; k < np; ++k) {
        if (k == j) continue;
        long double p = 1.0L / ((long double)xn[j] - xn[k]);
        for (int m = 0; m < np; ++m)
            if (m != j && m != k) p *= (x - xn[m]) / ((long double)xn[j] - xn[m]);
        s += p;
    }
    return (Real)s;
}
long double lagrange_val(const std::vector<Real>& xn, int j, long double x) {
    long double p = 1;
    for (int m = 0; m < (int)xn.size(); ++m)
        if (m != j) p *= (x - xn[m]) / ((long double)xn[j] - xn[m]);
    return p;
}
void gauss(int m, std::vector<Real>& xg, std::vector<Real>& wg) {
    xg.resize(m); wg.resize(m);
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int j = 0; j < m; ++j) {
        long double t = -cosl(pi * (j + 0.75L) / (m + 0.5L)), P, dP;
        for (int it = 0; it < 200; ++it) {
            leg(m, t, P, dP);
            const long double dt = P / dP;
            t -= dt;
            if (fabsl(dt) < 1e-32L) break;
        }
        leg(m, t, P, dP);
        xg[j] = (Real)((t + 1) / 2);
        wg[j] = (Real)(1.0L / ((1 - t * t) * dP * dP));
    }
}
// ChangeOfBasisMatrix ctor (dgsem/change_of_basis_matrix.h:74-98): C_ij = sum_q w_q psi_i(p_q) phi_j(p_q)
// over QGauss<3>(N+1) with FE_DGQLegendre (orthonormal Legendre on [0,1]) and FE_DGQ (GLL Lagrange).
void cob_matrix(int N, std::vector<Real>& C) {
    std::vector<Real> xn, wn, xg, wg;
    gll(N, xn, wn);
    gauss(N + 1, xg, wg);
    const int np = N + 1, n = np * np * np;
    std::vector<Real> psi(np * np), phi(np * np);  // [poly][quad pt]
    for (int i = 0; i < np; ++i)
        for (int q = 0; q < np; ++q) {
            psi[i * np + q] = (Real)(sqrtl(2.0L * i + 1) * leg_val(i, 2.0L * xg[q] - 1));
            phi[i * np + q] = (Real)lagrange_val(xn, i, xg[q]);
        }
    C.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int i0 = i % np, i1 = (i / np) % np, i2 = i / (np * np);
        for (int j = 0; j < n; ++j) {
            const int j0 = j % np, j1 = (j / np) % np, j2 = j / (np * np);
            Real s = 0;
            for (int q = 0; q < n; ++q) {
                const int q0 = q % np, q1 = (q / np) % np, q2 = q / (np * np);
                const Real w = wg[q0] * wg[q1] * wg[q2];
                const Real ps = psi[i0 * np + q0] * psi[i1 * np + q1] * psi[i2 * np + q2];
                const Real ph = phi[j0 * np + q0] * phi[j1 * np + q1] * phi[j2 * np + q2];
                s += w * ps * ph;
            }
            C[(size_t)i * n + j] = s;
        }
    }
}

// FaceDoFInfo::form_maps (dgsem/face_dof_info.h:92-150)
int face_map(int N, int f, int q) {
    const int np = N + 1, nf = np * np;
    switch (f) {
        case 0: return np * q;
        case 1: return N + np * q;
        case 2: return (q / np) * nf + (q % np);
        case 3: return np * N + (q / np) * nf + (q % np);
        case 4: return q;
        default: return N * nf + q;
    }
}

// RK4Stage5Register3 ctor, dgsem/rk4_stage5_register3.cc:12-36
struct RKC {
    Real b[5], ao[5], ai[4];
    RKC() {
        b[0] = 314199625218.0 / 7198350928319; b[1] = 6410344372641.0 / 17000082738695;
        b[2] = 292278564125.0 / 5593752632744; b[3] = 5010207514426.0 / 21876007855139;
        b[4] = 5597675544274.0 / 18784428342765;
        ao[0] = 4745337637855.0 / 22386579876409; ao[1] = 6808157035527.0 / 13197844641179;
        ao[2] = 4367509502613.0 / 10454198590847; ao[3] = 1236962429870.0 / 3429868089329;
        ao[4] = b[4];
        ai[0] = 546509042554.0 / 9152262712923; ai[1] = 625707605167.0 / 5316659119056;
        ai[2] = 582400652113.0 / 7078426004906; ai[3] = b[3];
    }
};

Real blender_from_values(const NS& ns, const oracle_physics& ph, int N, const std::vector<Real>& C,
                           const Real* vals) {
    // BlenderCalculator::get_blender / get_trouble, dgsem/blender_calculator.cc:123-187, 277-303
    const int np = N + 1, n = np * np * np;
    std::vector<Real> modes(n);
    {
        FLOP_SCOPE(11);
        for (int row = 0; row < n; ++row) {
            modes[row] = 0;
            for (int col = 0; col < n; ++col) modes[row] += C[(size_t)row * n + col] * vals[col];
        }
    }
    Real etot = 0, e1 = 0, e2 = 0;
    for (Real m : modes) etot += m * m;
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const int id = i + j * np + k * np * np;
                e1 += modes[id] * modes[id];
            }
    if (N > 1)
        for (int k = 0; k < N - 1; ++k)
            for (int j = 0; j < N - 1; ++j)
                for (int i = 0; i < N - 1; ++i) {
                    const int id = i + j * np + k * np * np;
                    e2 += modes[id] * modes[id];
                }
    const Real trouble = (N > 2) ? std::max(1 - e1 / etot, 1 - e2 / e1) : 1 - e1 / etot;
    const Real thr = ph.threshold_factor * pow(10, -ph.threshold_exp_factor * (pow(N + 1, 0.25)));
    const Real at = 1 / (1 + exp(-ph.sharpness_factor * (trouble / thr - 1)));
    (void)ns;
    if (at < ph.alpha_min) return 0;
    else if (at < ph.alpha_max) return at;
    else return ph.alpha_max;
}

}  // namespace

// =============================================================================================
// PLENS restatement
// =============================================================================================
struct oracle {
    int N, np, n, nf, nc;
    size_t ndof;
    NS ns;
    oracle_physics ph;
    std::vector<oracle_bc> bcs;
    std::vector<int32_t> nbr, bid;
    std::vector<Real> w, x1, D, Qm;                   // w_1d, ref_D_1d, ref_Q_1d
    std::vector<Real> C;                              // change-of-basis matrix
    // metric terms: JxContra[c][node][dir] (T1), detJ[c][node], Jinv[c][node][3][3]
    std::vector<T1> Ja;
    std::vector<Real> detJ, Jinv, hmin, normals;     // normals [c][f][q][3]
    std::vector<T1> sub;                                // subcell normals [c][dir][(N+2)(N+1)^2]
    // solution vectors
    std::vector<Real> Q, avars, mu, kap, alpha, alpha_d, rhs, rhs1, rhs2, Qn, loc_dt;
    std::vector<Real> alpha_given, alpha_d_given;  // checker hook: blender values to use, when set
    std::vector<Real> raw_given;                    // checker hook: indicator values (pre-override)
    std::string x_courant, x_wall, x_region;          // expression-valued Courant number, wall blender
                                                      // limit, viscous blending region (empty: unset)
    int64_t blended_cell_stages = 0;                  // (cell, RHS evaluation) pairs with alpha > 0
    std::vector<Real> s1, s2, s3;                     // surface terms [var][c][f][q]
    std::vector<Real> ws1;                            // entropy variables of s1 (default build)
    std::vector<Real> verts;                          // cell vertices [c][8][3]
    Real bc_time = 0;                                 // BC::time_ = cur_time of the step
    std::vector<std::vector<std::string>> bexpr;       // varying inflow: p, T, u, v, w per bc
    int mq = 1;                                         // degree of the MappingQGeneric
    std::vector<Real> sup, tq;                        // support points [c][(mq+1)^3][3], GLL(mq)
    std::vector<Real> rhoE_old;                       // rho E at the start of the last step
    // local time stepping ("time integration" parameters, dgsem/plens.cc:540-575, 1481-1482)
    int lts_request = 0, lts_limit = 0;
    Real lts_start = 100, lts_nf = 1.05, lts_gf = 1.0;
    int64_t fail_dof = -1;
    int fail_kind = 0;
    double timers[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    RKC rk;

    int tl(int i, int j, int k) const { return i + j * np + k * np * np; }  // CellDoFInfo
    size_t gid(int c, int i) const { return (size_t)c * n + i; }
    size_t sidx(int var, int c, int f, int q) const { return (((size_t)var * nc + c) * 6 + f) * nf + q; }
    const T1& JaV(int c, int node, int dir) const { return Ja[((size_t)c * n + node) * 3 + dir]; }
    State cons(int c, int i) const {
        State s;
        for (int v = 0; v < 5; ++v) s[v] = Q[v * ndof + gid(c, i)];
        return s;
    }
    Avars av(int c, int i) const {
        Avars a;
        for (int v = 0; v < 9; ++v) a[v] = avars[v * ndof + gid(c, i)];
        return a;
    }
    T1 normal(int c, int f, int q) const {
        T1 t;
        for (int d = 0; d < 3; ++d) t[d] = normals[(((size_t)c * 6 + f) * nf + q) * 3 + d];
        return t;
    }

    // location of node `node` of cell c (the DoF location): the map at the GLL point
    T1 node_location(int c, int node) const {
        const Real xi[3] = {x1[node % np], x1[(node / np) % np], x1[node / (np * np)]};
        T1 X;
        if (mq == 1) {
            const Real* cv = &verts[(size_t)c * 24];
            for (int v = 0; v < 8; ++v) {
                Real w = 1;
                for (int a = 0; a < 3; ++a) w *= ((v >> a) & 1) ? xi[a] : 1.0 - xi[a];
                for (int d = 0; d < 3; ++d) X[d] += w * cv[3 * v + d];
            }
            return X;
        }
        const int qp = mq + 1;
        const Real* S = &sup[(size_t)c * qp * qp * qp * 3];
        for (int k = 0; k < qp; ++k)
            for (int j = 0; j < qp; ++j)
                for (int i = 0; i < qp; ++i) {
                    const Real l = (Real)(lagrange_val(tq, i, xi[0]) * lagrange_val(tq, j, xi[1]) *
                                              lagrange_val(tq, k, xi[2]));
                    for (int d = 0; d < 3; ++d) X[d] += l * S[3 * ((size_t)i + qp * ((size_t)j + (size_t)qp * k)) + d];
                }
        return X;
    }

    // Jacobian J[a][b] = dx_a/dxi_b of cell c's map at reference point xi: trilinear from the 8
    // vertices (MappingQGeneric(1)), else the degree-mq Lagrange interpolant of the support points
    void jacobian(int c, const Real xi[3], Real J[3][3]) const {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) J[a][b] = 0;
        if (mq == 1) {
            const Real* cv = &verts[(size_t)c * 24];
            for (int v = 0; v < 8; ++v) {
                const int bb[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
                for (int b = 0; b < 3; ++b) {
                    Real g = bb[b] ? 1.0 : -1.0;
                    for (int o = 0; o < 3; ++o)
                        if (o != b) g *= bb[o] ? xi[o] : 1.0 - xi[o];
                    for (int a = 0; a < 3; ++a) J[a][b] += g * cv[3 * v + a];
                }
            }
            return;
        }
        const int qp = mq + 1;
        std::vector<Real> L(3 * qp), dL(3 * qp);
        for (int d = 0; d < 3; ++d)
            for (int j = 0; j < qp; ++j) {
                L[d * qp + j] = (Real)lagrange_val(tq, j, xi[d]);
                dL[d * qp + j] = lagrange_deriv(tq, j, xi[d]);
            }
        const Real* S = &sup[(size_t)c * qp * qp * qp * 3];
        for (int k = 0; k < qp; ++k)
            for (int j = 0; j < qp; ++j)
                for (int i = 0; i < qp; ++i) {
                    const Real* sp = S + 3 * ((size_t)i + qp * ((size_t)j + (size_t)qp * k));
                    const Real g[3] = {dL[i] * L[qp + j] * L[2 * qp + k], L[i] * dL[qp + j] * L[2 * qp + k],
                                         L[i] * L[qp + j] * dL[2 * qp + k]};
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) J[a][b] += g[b] * sp[a];
                }
    }

    // ---- MetricTerms::reinit, dgsem/metric_terms.cc:67-142 ----------------------------------
    void metrics(const Real* V, int transpose_y) {
        Ja.assign((size_t)nc * n * 3, T1());
        detJ.assign((size_t)nc * n, 0);
        Jinv.assign((size_t)nc * n * 9, 0);
        hmin.assign(nc, 0);
        normals.assign((size_t)nc * 6 * nf * 3, 0);
        sub.assign((size_t)nc * 3 * (N + 2) * np * np, T1());
        for (int c = 0; c < nc; ++c) {
            const Real* cv = V + (size_t)c * 24;
            Real dmin = 1e300;
            for (int a = 0; a < 8; ++a)
                for (int b = a + 1; b < 8; ++b) {
                    Real d2 = 0;
                    for (int q = 0; q < 3; ++q) d2 += (cv[3 * a + q] - cv[3 * b + q]) * (cv[3 * a + q] - cv[3 * b + q]);
                    dmin = std::min(dmin, d2);
                }
            hmin[c] = sqrt(dmin);
            for (int k = 0; k < np; ++k)
                for (int j = 0; j < np; ++j)
                    for (int i = 0; i < np; ++i) {
                        const int node = tl(i, j, k);
                        const Real xi[3] = {x1[i], x1[j], x1[k]};
                        Real J[3][3];  // J[a][b] = dx_a/dxi_b
                        jacobian(c, xi, J);
                        const Real det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                        detJ[gid(c, node)] = det;
                        // inverse by cofactors
                        Real* Ji = &Jinv[gid(c, node) * 9];
                        Ji[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
                        Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
                        Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
                        Ji[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
                        Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
                        Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
                        Ji[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
                        Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
                        Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
                        T1 co[3];  // covariant vectors = rows of J^T
                        for (int d = 0; d < 3; ++d)
                            for (int a = 0; a < 3; ++a) co[d][a] = J[a][d];
                        for (int d = 0; d < 3; ++d)
                            Ja[((size_t)c * n + node) * 3 + d] = cross(co[(d + 1) % 3], co[(d + 2) % 3]);
                    }
            // subcell normals, metric_terms.cc:104-140
            for (int dir = 0; dir < 3; ++dir) {
                const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
                for (int id = 0; id <= N + 1; ++id)
                    for (int id1 = 0; id1 <= N; ++id1)
                        for (int id2 = 0; id2 <= N; ++id2) {
                            int t0[3];
                            t0[dir] = 0; t0[d1] = id1; t0[d2] = id2;
                            T1 s = JaV(c, tl(t0[0], t0[1], t0[2]), dir);
                            for (int l = 0; l <= id - 1; ++l)
                                for (int m = 0; m <= N; ++m) {
                                    int tm[3];
                                    tm[dir] = m; tm[d1] = id1; tm[d2] = id2;
                                    const T1& J = JaV(c, tl(tm[0], tm[1], tm[2]), dir);
                                    for (int a = 0; a < 3; ++a) s[a] += Qm[l * np + m] * J[a];
                                }
                            sub[sub_index(c, dir, id, id1, id2)] = s;
                        }
            }
            // FEFaceValues normals at face dofs (H1: faces 2/3 read the transposed point)
            for (int f = 0; f < 6; ++f) {
                const int dir = f / 2;
                const Real sg = (f % 2) ? 1.0 : -1.0;
                for (int q = 0; q < nf; ++q) {
                    int qq = q;
                    if ((f == 2 || f == 3) && transpose_y) qq = (q % np) * np + q / np;
                    const T1& J = JaV(c, face_map(N, f, qq), dir);
                    const Real nr = J.norm();
                    for (int d = 0; d < 3; ++d) normals[(((size_t)c * 6 + f) * nf + q) * 3 + d] = sg * J[d] / nr;
                }
            }
        }
    }
    size_t sub_index(int c, int dir, int id, int id1, int id2) const {
        return ((((size_t)c * 3 + dir) * (N + 2) + id) * np + id1) * np + id2;
    }
    // subcell_normals[dir](ti) with ti[dir] = id
    const T1& subn(int c, int dir, const int ti[3]) const {
        const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
        return sub[sub_index(c, dir, ti[dir], ti[d1], ti[d2])];
    }

    // ---- BC ghost getters (IBCs/*.cc) --------------------------------------------------------
    void ghost(int stage, int c, int f, int q, const State& in, const Avars& ain, const T1& nrm,
               State& g, Avars& ag) const {
        const oracle_bc& b = bcs[bid[(size_t)c * 6 + f]];
        ag = ain;
        switch (b.type) {
            case 0: {  // periodic: partner state/avars at the same face dof (periodic.cc:121-165)
                const int pc = nbr[(size_t)c * 6 + f];
                const int pf = (f % 2) ? f - 1 : f + 1;
                const int node = face_map(N, pf, q);
                g = cons(pc, node);
                if (stage == 3) ag = av(pc, node);
                break;
            }
            case 1: {  // uniform inflow (uniform_inflow.cc:23-77)
                State pr;
                T1 v; v[0] = b.vel[0]; v[1] = b.vel[1]; v[2] = b.vel[2];
                const Real rho = b.p / (ns.R() * b.T);  // set_BC, plens.cc:1268-1270
                ns.prim_to_cons(rho, v, b.p, pr);
                if (stage == 2 || (stage == 1 && g_mutation == kMutInflowS1)) g = pr;
                else for (int k = 0; k < 5; ++k) g[k] = 2 * pr[k] - in[k];
                break;
            }
            case 2: {  // outflow (outflow.cc:18-75)
                if (stage == 2) {
                    const Real M = ns.get_M(in);
                    if ((ns.inviscid() && g_mutation != kMutOutflowInviscid) || M >= 1) g = in;
                    else {
                        g = in;
                        const Real pin = ns.get_p(in);
                        if (g_mutation == kMutOutflowSign) g[4] += 2 * (pin - b.p) / (ns.gma - 1);
                        else g[4] += 2 * (b.p - pin) / (ns.gma - 1);
                    }
                } else g = in;
                break;
            }
            case 4:  // free (free.cc:10-60): inner state and avars at every stage
                g = in;
                break;
            case 5:    // symmetry (symmetry.cc:17-96)
            case 6: {  // empty (empty.cc): as symmetry, except the stage-2 ghost
                T1 vin, vt;
                for (int d = 0; d < 3; ++d) vin[d] = in[1 + d] / in[0];
                const Real vn = dot(vin, nrm);
                for (int d = 0; d < 3; ++d) vt[d] = vin[d] - vn * nrm[d];
                const Real pin = ns.get_p(in);
                State pr;
                ns.prim_to_cons(in[0], vt, pin, pr);
                if (stage != 2) {
                    for (int k = 0; k < 5; ++k) g[k] = 2 * pr[k] - in[k];
                } else if (b.type == 6) {
                    g = pr;
                } else {
                    T1 sm;
                    for (int d = 0; d < 3; ++d) sm[d] = in[1 + d];
                    const Real smn = dot(sm, nrm);
                    g[0] = in[0];
                    g[4] = in[4];
                    for (int d = 0; d < 3; ++d) g[1 + d] = sm[d] - 2 * smn * nrm[d];
                }
                break;
            }
            case 9: {  // varying inflow (varying_inflow.cc): uniform inflow's ghosts, state prescribed
                       // by the expressions at the DoF location and BC::time_
                const std::vector<std::string>& ex = bexpr[bid[(size_t)c * 6 + f]];
                const T1 X = node_location(c, face_map(N, f, q));
                Real val[5];
                for (int k = 0; k < 5; ++k) val[k] = eval_expression(ex[k], X[0], X[1], X[2], bc_time);
                T1 v; v[0] = val[2]; v[1] = val[3]; v[2] = val[4];
                State pr;
                ns.prim_to_cons(val[0] / (ns.R() * val[1]), v, val[0], pr);
                if (stage == 2) g = pr;
                else for (int k = 0; k < 5; ++k) g[k] = 2 * pr[k] - in[k];
                break;
            }
            case 8: {  // ZPG inflow (zpg_inflow.cc): prescribed velocity, inner density and pressure
                T1 vw; vw[0] = b.vel[0]; vw[1] = b.vel[1]; vw[2] = b.vel[2];
                State pr;
                ns.prim_to_cons(in[0], vw, ns.get_p(in), pr);
                if (stage == 2) g = pr;
                else for (int k = 0; k < 5; ++k) g[k] = 2 * pr[k] - in[k];
                break;
            }
            default: {  // uniform temperature wall (uniform_temp_wall.cc:17-104), insulated wall
                        // (insulated_wall.cc): the same stage-2 rule; stage 1/3 use the inner
                        // temperature for the insulated wall
                T1 vw; vw[0] = b.vel[0]; vw[1] = b.vel[1]; vw[2] = b.vel[2];
                if (stage == 2) {
                    g[0] = in[0];
                    g[4] = in[4];
                    if (ns.inviscid()) {
                        T1 sm, smr;
                        for (int d = 0; d < 3; ++d) { sm[d] = in[1 + d]; smr[d] = in[1 + d] - in[0] * vw[d]; }
                        const Real nsm = dot(smr, nrm);
                        for (int d = 0; d < 3; ++d) g[1 + d] = sm[d] - 2 * nsm * nrm[d];
                    } else if (g_mutation == kMutWallAbsVel) {
                        for (int d = 0; d < 3; ++d) g[1 + d] = -in[1 + d];
                    } else {
                        for (int d = 0; d < 3; ++d) g[1 + d] = 2 * in[0] * vw[d] - in[1 + d];
                    }
                } else {
                    State pr;
                    const Real pin = ns.get_p(in);
                    const bool inner_rho = b.type == 7 || g_mutation == kMutWallRho;
                    const Real rpr = inner_rho ? in[0] : pin / (ns.R() * b.T);
                    ns.prim_to_cons(rpr, vw, pin, pr);
                    for (int k = 0; k < 5; ++k) g[k] = 2 * pr[k] - in[k];
                    if (b.type == 7 && stage == 3) {
                        // heat flux with its normal component reversed, shear stresses kept
                        // (insulated_wall.cc:95-101 as documented there; the reference's loop writes
                        // heat_flux_in[6..8], out of the bounds of the 3-vector, at line 99)
                        Real qn = 0;
                        for (int d = 0; d < 3; ++d) qn += ain[6 + d] * nrm[d];
                        for (int d = 0; d < 3; ++d) ag[6 + d] = ain[6 + d] - 2 * qn * nrm[d];
                    }
                }
                break;
            }
        }
    }

    // surf_flux_wrappers[stage-1], navier_stokes.cc:273-319
    void surf_flux(int stage, const State& u, const Avars& a, const State& ug, const Avars& ag,
                   const T1& nrm, State& f, const NS& nsl) const {
        if (stage == 1) {
            for (int v = 0; v < 5; ++v) f[v] = 0.5 * (u[v] + ug[v]);
        } else if (stage == 2) {
            FLOP_SCOPE(12);
            nsl.get_inv_surf_flux(u, ug, nrm, f);
        } else {
            State f1, f2;
            NS::dif_flux(u, a, nrm, f1);
            NS::dif_flux(ug, ag, nrm, f2);
            for (int v = 0; v < 5; ++v) f[v] = 0.5 * (f1[v] + f2[v]);
        }
    }

    // ---- PLENS::calc_surf_flux, dgsem/plens.cc:1717-1913 --------------------------------------
    void calc_surf_flux(int stage, std::vector<Real>& sf) {
        sf.assign((size_t)5 * nc * 6 * nf, 0.0);
        const Real rsign = (stage == 1) ? 1.0f : -1.0f;
#pragma omp parallel for schedule(static)
        for (int c = 0; c < nc; ++c) {
            NS nsl = ns;  // per-thread copy: the reference mutates flux_blender_value
            for (int f = 0; f < 6; ++f) {
                const int b = bid[(size_t)c * 6 + f];
                if (b >= 0) {
                    nsl.set_blender(alpha[c]);
                    for (int q = 0; q < nf; ++q) {
                        const int node = face_map(N, f, q);
                        State u = cons(c, node), ug;
                        Avars a = av(c, node), ag;
                        const T1 nrm = normal(c, f, q);
                        ghost(stage, c, f, q, u, a, nrm, ug, ag);
                        State fl;
                        surf_flux(stage, u, a, ug, ag, nrm, fl, nsl);
                        for (int v = 0; v < 5; ++v) sf[sidx(v, c, f, q)] = fl[v];
                    }
                } else {
                    const int nb = nbr[(size_t)c * 6 + f];
                    if (!(c < nb)) continue;
                    const int fn = (f % 2) ? f - 1 : f + 1;  // neighbor_of_neighbor
                    nsl.set_blender(0.5 * (alpha[c] + alpha[nb]));
                    for (int q = 0; q < nf; ++q) {
                        const int qn = q;  // nei_face_matching_dofs: identity on these meshes
                        const State u = cons(c, face_map(N, f, q)), un = cons(nb, face_map(N, fn, qn));
                        const Avars a = av(c, face_map(N, f, q)), an = av(nb, face_map(N, fn, qn));
                        State fl;
                        surf_flux(stage, u, a, un, an, normal(c, f, q), fl, nsl);
                        for (int v = 0; v < 5; ++v) {
                            sf[sidx(v, c, f, q)] = fl[v];
                            sf[sidx(v, nb, fn, qn)] = rsign * fl[v];
                        }
                    }
                }
            }
        }
    }

    // ---- PLENS::calc_cell_cons_grad, dgsem/plens.cc:2174-2285 ---------------------------------
    void calc_cell_cons_grad(int c, const std::vector<Real>& s1f, std::vector<std::array<State, 3>>& g) const {
        std::vector<State> st(n);
        for (int i = 0; i < n; ++i) st[i] = cons(c, i);
        const Real ssign[2] = {1, -1};
        for (int d = 0; d < 3; ++d)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < 5; ++v) g[i][d][v] = 0;
        for (int gd = 0; gd < 3; ++gd) {
            for (int it = 0; it < n; ++it) {
                const State ct = st[it];
                const int ti[3] = {it % np, (it / np) % np, it / (np * np)};
                State fl;
                for (int m = 0; m <= N; ++m)
                    for (int md = 0; md < 3; ++md) {
                        int to[3] = {ti[0], ti[1], ti[2]};
                        to[md] = m;
                        const int io = tl(to[0], to[1], to[2]);
                        const State co = st[io];
                        for (int v = 0; v < 5; ++v) fl[v] = 0.5 * (ct[v] + co[v]);
                        const Real Jav = 0.5 * (JaV(c, it, md)[gd] + JaV(c, io, md)[gd]);
                        const Real Dm2 = 2 * D[ti[md] * np + m];
                        for (int v = 0; v < 5; ++v) g[it][gd][v] += Dm2 * fl[v] * Jav;
                    }
            }
            for (int sd = 0; sd < 3; ++sd)
                for (int lr = 0; lr <= 1; ++lr) {
                    const int f = 2 * sd + lr;
                    for (int q = 0; q < nf; ++q) {
                        const int node = face_map(N, f, q);
                        const State cs = st[node];
                        for (int v = 0; v < 5; ++v) {
                            const Real fin = cs[v] * JaV(c, node, sd)[gd];
                            const Real fsu = s1f[sidx(v, c, f, q)] * JaV(c, node, sd)[gd];
                            g[node][gd][v] -= ssign[lr] * (fsu - fin) / w[lr * N];
                        }
                    }
                }
            for (int i = 0; i < n; ++i) {
                const Real Ji = 1.0 / detJ[gid(c, i)];
                for (int v = 0; v < 5; ++v) g[i][gd][v] *= Ji;
            }
        }
    }

    // ---- PLENS::calc_cell_evar_grad, dgsem/plens.cc:2294-2405 (default build) ------------------
    // ws1: entropy variables of the stage-1 face values (calc_aux_vars, 2610-2637)
    void calc_cell_evar_grad(int c, const std::vector<Real>& ws1, std::vector<std::array<State, 3>>& g) const {
        std::vector<State> ev(n);
        for (int i = 0; i < n; ++i) ns.cons_to_entropy(cons(c, i), ev[i]);
        const Real ssign[2] = {1, -1};
        for (int d = 0; d < 3; ++d)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < 5; ++v) g[i][d][v] = 0;
        for (int gd = 0; gd < 3; ++gd) {
            for (int it = 0; it < n; ++it) {
                const int ti[3] = {it % np, (it / np) % np, it / (np * np)};
                for (int m = 0; m <= N; ++m)
                    for (int md = 0; md < 3; ++md) {
                        int to[3] = {ti[0], ti[1], ti[2]};
                        to[md] = m;
                        const int io = tl(to[0], to[1], to[2]);
                        const State eo = ev[io];
                        const Real Jc = JaV(c, io, md)[gd];
                        const Real Dv = D[ti[md] * np + m];
                        for (int v = 0; v < 5; ++v) g[it][gd][v] += Dv * eo[v] * Jc;
                    }
            }
            for (int sd = 0; sd < 3; ++sd)
                for (int lr = 0; lr <= 1; ++lr) {
                    const int f = 2 * sd + lr;
                    for (int q = 0; q < nf; ++q) {
                        const int node = face_map(N, f, q);
                        const State e = ev[node];
                        for (int v = 0; v < 5; ++v) {
                            const Real fin = e[v] * JaV(c, node, sd)[gd];
                            const Real fsu = ws1[sidx(v, c, f, q)] * JaV(c, node, sd)[gd];
                            g[node][gd][v] -= ssign[lr] * (fsu - fin) / w[lr * N];
                        }
                    }
                }
            for (int i = 0; i < n; ++i) {
                const Real Ji = 1.0 / detJ[gid(c, i)];
                for (int v = 0; v < 5; ++v) g[i][gd][v] *= Ji;
            }
        }
    }

    // ---- PLENS::assert_positivity, dgsem/plens.cc:2531-2546 -----------------------------------
    bool assert_positivity() {
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < n; ++i) {
                const State s = cons(c, i);
                if (!(s[0] > 0)) { fail_dof = (int64_t)gid(c, i); fail_kind = 1; return false; }
                if (!(NS::get_e(s) > 0)) { fail_dof = (int64_t)gid(c, i); fail_kind = 2; return false; }
            }
        return true;
    }

    // ---- PLENS::calc_aux_vars, dgsem/plens.cc:2593-2748 (both builds) --------------------------
    void calc_aux_vars() {
        const Real cvv = ns.cv();
        for (size_t i = 0; i < ndof; ++i) {
            State s;
            for (int v = 0; v < 5; ++v) s[v] = Q[v * ndof + i];
            const Real e = NS::get_e(s);
            mu[i] = ns.mu(e / cvv);
            kap[i] = ns.kappa(mu[i]);
        }
        calc_surf_flux(1, s1);
        if (ph.gradient_variables == 1) {
            // entropy variables of the stage-1 face values (dgsem/plens.cc:2610-2637)
            ws1.assign(s1.size(), 0.0);
            const size_t per_var = s1.size() / 5;
            for (size_t k = 0; k < per_var; ++k) {
                State cs, ws;
                for (int v = 0; v < 5; ++v) cs[v] = s1[v * per_var + k];
                ns.cons_to_entropy(cs, ws);
                for (int v = 0; v < 5; ++v) ws1[v * per_var + k] = ws[v];
            }
            calc_aux_vars_entropy();
            return;
        }
#pragma omp parallel
        {
            std::vector<std::array<State, 3>> g(n);
#pragma omp for schedule(static)
            for (int c = 0; c < nc; ++c) {
                calc_cell_cons_grad(c, s1, g);
                for (int i = 0; i < n; ++i) {
                    const State s = cons(c, i);
                    const Real vel[3] = {s[1] / s[0], s[2] / s[0], s[3] / s[0]};
                    Real vg[3][3];
                    for (int vd = 0; vd < 3; ++vd)
                        for (int gd = 0; gd < 3; ++gd)
                            vg[vd][gd] = (g[i][gd][1 + vd] - g[i][gd][0] * vel[vd]) / s[0];
                    int sid = 0;
                    Real tr = 0;
                    for (int d = 0; d < 3; ++d) tr += vg[d][d];
                    const size_t gi = gid(c, i);
                    for (int row = 0; row < 3; ++row)
                        for (int col = row; col < 3; ++col) {
                            if (col == row) avars[sid * ndof + gi] = vg[row][col] + vg[col][row] - 2.0 / 3 * tr;
                            else avars[sid * ndof + gi] = vg[row][col] + vg[col][row];
                            ++sid;
                        }
                    const Real e = NS::get_e(s);
                    Real eg[3];
                    for (int gd = 0; gd < 3; ++gd) {
                        eg[gd] = (g[i][gd][4] - g[i][gd][0] * e) / s[0];
                        for (int vd = 0; vd < 3; ++vd)
                            eg[gd] -= (0.5 * g[i][gd][0] * vel[vd] * vel[vd] + s[1 + vd] * vg[vd][gd]) / s[0];
                    }
                    for (int d = 0; d < 3; ++d) avars[(6 + d) * ndof + gi] = -eg[d] / cvv;
                }
            }
        }
        for (size_t i = 0; i < ndof; ++i) {
            for (int v = 0; v < 6; ++v) avars[v * ndof + i] *= mu[i];
            for (int v = 6; v < 9; ++v) avars[v * ndof + i] *= kap[i];
        }
    }

    // entropy-variable branch of calc_aux_vars (dgsem/plens.cc:2639-2746 without CVAR_GRAD)
    void calc_aux_vars_entropy() {
#pragma omp parallel
        {
            std::vector<std::array<State, 3>> g(n);
#pragma omp for schedule(static)
            for (int c = 0; c < nc; ++c) {
                calc_cell_evar_grad(c, ws1, g);
                for (int i = 0; i < n; ++i) {
                    State et;
                    ns.cons_to_entropy(cons(c, i), et);
                    Real vg[3][3];
                    for (int vd = 0; vd < 3; ++vd)
                        for (int gd = 0; gd < 3; ++gd)
                            vg[vd][gd] = (et[1 + vd] * g[i][gd][4] / et[4] - g[i][gd][1 + vd]) / et[4];
                    int sid = 0;
                    Real tr = 0;
                    for (int d = 0; d < 3; ++d) tr += vg[d][d];
                    const size_t gi = gid(c, i);
                    for (int row = 0; row < 3; ++row)
                        for (int col = row; col < 3; ++col) {
                            if (col == row) avars[sid * ndof + gi] = vg[row][col] + vg[col][row] - 2.0 / 3 * tr;
                            else avars[sid * ndof + gi] = vg[row][col] + vg[col][row];
                            ++sid;
                        }
                    for (int d = 0; d < 3; ++d)
                        avars[(6 + d) * ndof + gi] = -(g[i][d][4] / (et[4] * et[4] * ns.R()));
                }
            }
        }
        for (size_t i = 0; i < ndof; ++i) {
            for (int v = 0; v < 6; ++v) avars[v * ndof + i] *= mu[i];
            for (int v = 6; v < 9; ++v) avars[v * ndof + i] *= kap[i];
        }
    }

    // ---- PLENS::calc_blender, dgsem/plens.cc:2858-2962 ------------------------------------------
    void calc_blender() {
        std::vector<Real> raw(nc);
#pragma omp parallel
        {
            std::vector<Real> vals(n);
#pragma omp for schedule(static)
            for (int c = 0; c < nc; ++c) {
                for (int i = 0; i < n; ++i) {
                    const State s = cons(c, i);
                    vals[i] = ns.get_p(s) * s[0];  // 'pxrho'
                }
                raw[c] = blender_from_values(ns, ph, N, C, vals.data());
            }
        }
        if (!raw_given.empty()) raw = raw_given;  // checker hook: given indicator values
        // wall_blender_limit_function.value(Point<dim>()) at cur_time (dgsem/plens.cc:2884-2898)
        const Real wlim = x_wall.empty() ? Real(ph.wall_blender_limit) : eval_expression(x_wall, 0.0, 0.0, 0.0, bc_time);
        auto wall_override = [&]() {
            for (int c = 0; c < nc; ++c) {
                for (int f = 0; f < 6; ++f) {
                    const int b = bid[(size_t)c * 6 + f];
                    if (b >= 0 && (bcs[b].type == 3 || bcs[b].type == 7))  // uniform temp / insulated wall
                        alpha[c] = std::max(wlim, g_mutation == kMutBlendOverrideLate ? alpha[c] : raw[c]);
                }
            }
        };
        alpha = raw;
        if (g_mutation != kMutBlendOverrideLate) wall_override();
        const std::vector<Real> gh = alpha;  // gh_gcrk_alpha = gcrk_alpha
        const std::vector<Real>& nsrc = g_mutation == kMutBlendLive ? alpha : gh;
        for (int c = 0; c < nc; ++c)
            for (int f = 0; f < 6; ++f) {
                // face->at_boundary() (periodic included)
                const int b = bid[(size_t)c * 6 + f];
                if (b >= 0 && !(g_mutation == kMutBlendPeriodic && bcs[b].type == 0)) continue;
                const Real cur = alpha[c], nei = nsrc[nbr[(size_t)c * 6 + f]];
                alpha[c] = std::max(cur, 0.5 * nei);
            }
        if (g_mutation == kMutBlendOverrideLate) wall_override();
        for (int c = 0; c < nc; ++c) {  // vis_blending_region.value(cell->center()) > 0 (2950-2960)
            bool in_region = true;
            if (!x_region.empty()) {
                Real ctr[3] = {0, 0, 0};  // TriaAccessor::center(): the mean of the vertices
                for (int v = 0; v < 8; ++v)
                    for (int d = 0; d < 3; ++d) ctr[d] += verts[(size_t)c * 24 + 3 * v + d] / 8.0;
                in_region = eval_expression(x_region, ctr[0], ctr[1], ctr[2], bc_time) > 0;
            }
            alpha_d[c] = !in_region ? Real(0.0)
                         : g_mutation == kMutBlendNoCap ? alpha[c] / ph.alpha_max
                                                        : std::min(Real(1.0), alpha[c] / ph.alpha_max);
        }
    }

    // ---- PLENS::calc_cell_ho_inv_residual, dgsem/plens.cc:3153-3275 ----------------------------
    void ho_inv(int c, std::vector<State>& r) const {
        const Real ss = -1;
        std::vector<State> st(n);
        for (int i = 0; i < n; ++i) st[i] = cons(c, i);
        const Real sgn[2] = {1, -1};
        for (int i = 0; i < n; ++i) r[i] = State{0, 0, 0, 0, 0};
        for (int it = 0; it < n; ++it) {
            const State ct = st[it];
            const int ti[3] = {it % np, (it / np) % np, it / (np * np)};
            State fl;
            for (int m = 0; m <= N; ++m)
                for (int md = 0; md < 3; ++md) {
                    int to[3] = {ti[0], ti[1], ti[2]};
                    to[md] = m;
                    const int io = tl(to[0], to[1], to[2]);
                    T1 dir;
                    for (int a = 0; a < 3; ++a) dir[a] = 0.5 * (JaV(c, it, md)[a] + JaV(c, io, md)[a]);
                    const Real nr = dir.norm();
                    for (int a = 0; a < 3; ++a) dir[a] /= nr;
                    ns.vol_flux(ct, st[io], dir, fl);
                    const Real Dn = 2 * D[ti[md] * np + m] * nr;
                    for (int v = 0; v < 5; ++v) r[it][v] += Dn * fl[v];
                }
        }
        FLOP_BUCKET(14);
        for (int sd = 0; sd < 3; ++sd)
            for (int lr = 0; lr <= 1; ++lr) {
                const int f = 2 * sd + lr;
                for (int q = 0; q < nf; ++q) {
                    const int node = face_map(N, f, q);
                    T1 dir = JaV(c, node, sd);
                    const Real nr = dir.norm();
                    for (int a = 0; a < 3; ++a) dir[a] /= nr;
                    State fs, fi;
                    for (int v = 0; v < 5; ++v) fs[v] = -sgn[lr] * s2[sidx(v, c, f, q)];
                    ns.get_inv_flux(st[node], dir, fi);
                    for (int v = 0; v < 5; ++v) r[node][v] += -sgn[lr] * (fs[v] - fi[v]) * nr / w[lr * N];
                }
            }
        for (int i = 0; i < n; ++i) {
            const Real jm = 1 / (ss * detJ[gid(c, i)]);
            for (int v = 0; v < 5; ++v) r[i][v] *= jm;
        }
    }

    // ---- PLENS::calc_cell_ho_dif_residual, dgsem/plens.cc:3283-3413 ----------------------------
    void ho_dif(int c, std::vector<State>& r) const {
        const Real ss = 1;
        std::vector<State> st(n);
        std::vector<Avars> sa(n);
        for (int i = 0; i < n; ++i) { st[i] = cons(c, i); sa[i] = av(c, i); }
        const Real sgn[2] = {1, -1};
        for (int i = 0; i < n; ++i) r[i] = State{0, 0, 0, 0, 0};
        for (int it = 0; it < n; ++it) {
            const int ti[3] = {it % np, (it / np) % np, it / (np * np)};
            State fl;
            for (int m = 0; m <= N; ++m)
                for (int md = 0; md < 3; ++md) {
                    int to[3] = {ti[0], ti[1], ti[2]};
                    to[md] = m;
                    const int io = tl(to[0], to[1], to[2]);
                    T1 dir = JaV(c, io, md);
                    const Real nr = dir.norm();
                    for (int a = 0; a < 3; ++a) dir[a] /= nr;
                    {
                        FLOP_SCOPE(13);
                        NS::dif_flux(st[io], sa[io], dir, fl);
                    }
                    const Real Dn = D[ti[md] * np + m] * nr;
                    for (int v = 0; v < 5; ++v) r[it][v] += Dn * fl[v];
                }
        }
        FLOP_BUCKET(15);
        for (int sd = 0; sd < 3; ++sd)
            for (int lr = 0; lr <= 1; ++lr) {
                const int f = 2 * sd + lr;
                for (int q = 0; q < nf; ++q) {
                    const int node = face_map(N, f, q);
                    T1 dir = JaV(c, node, sd);
                    const Real nr = dir.norm();
                    for (int a = 0; a < 3; ++a) dir[a] /= nr;
                    State fs, fi;
                    for (int v = 0; v < 5; ++v) fs[v] = -sgn[lr] * s3[sidx(v, c, f, q)];
                    NS::dif_flux(st[node], sa[node], dir, fi);
                    for (int v = 0; v < 5; ++v) r[node][v] += -sgn[lr] * (fs[v] - fi[v]) * nr / w[lr * N];
                }
            }
        for (int i = 0; i < n; ++i) {
            const Real jm = 1 / (ss * detJ[gid(c, i)]);
            for (int v = 0; v < 5; ++v) r[i][v] *= jm;
        }
    }

    // ---- PLENS::calc_cell_lo_inv_residual, dgsem/plens.cc:3469-3583 ----------------------------
    void lo_inv(int c, std::vector<State>& r, NS& nsl) const {
        for (int i = 0; i < n; ++i) r[i] = State{0, 0, 0, 0, 0};
        nsl.set_blender(alpha[c]);
        for (int dir = 0; dir < 3; ++dir) {
            const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
            for (int id1 = 0; id1 <= N; ++id1)
                for (int id2 = 0; id2 <= N; ++id2)
                    for (int id = 1; id <= N; ++id) {
                        int tlf[3], trt[3];
                        tlf[dir] = id - 1; tlf[d1] = id1; tlf[d2] = id2;
                        trt[dir] = id; trt[d1] = id1; trt[d2] = id2;
                        const int il = tl(tlf[0], tlf[1], tlf[2]), ir = tl(trt[0], trt[1], trt[2]);
                        const State cl = cons(c, il), cr = cons(c, ir);
                        T1 nd = subn(c, dir, trt);
                        const Real nn = nd.norm();
                        for (int a = 0; a < 3; ++a) nd[a] /= nn;
                        State fl;
                        nsl.get_inv_surf_flux(cl, cr, nd, fl);
                        for (int v = 0; v < 5; ++v) {
                            r[il][v] += fl[v] * nn / w[id - 1];
                            r[ir][v] -= fl[v] * nn / w[id];
                        }
                    }
        }
        for (int dir = 0; dir < 3; ++dir)
            for (int lr = 0; lr <= 1; ++lr) {
                const int f = 2 * dir + lr;
                const float cs = lr ? 1.0f : -1.0f, fs = lr ? 1.0f : -1.0f;
                for (int q = 0; q < nf; ++q) {
                    const int node = face_map(N, f, q);
                    const Real nn = JaV(c, node, dir).norm();
                    for (int v = 0; v < 5; ++v) {
                        const Real fl = fs * s2[sidx(v, c, f, q)];
                        r[node][v] += cs * fl * nn / w[lr * N];
                    }
                }
            }
        for (int i = 0; i < n; ++i)
            for (int v = 0; v < 5; ++v) r[i][v] /= -detJ[gid(c, i)];
    }

    // ---- PLENS::calc_rhs, dgsem/plens.cc:3688-3779 ------------------------------------------------
    int calc_rhs() {
        auto t0 = std::chrono::steady_clock::now();
        auto lap = [&](int k) {
            auto t1 = std::chrono::steady_clock::now();
            timers[k] += std::chrono::duration<double>(t1 - t0).count();
            t0 = t1;
        };
        FLOP_BUCKET(0);
        if (!assert_positivity()) return -3;
        lap(0);
        FLOP_BUCKET(1);
        if (!ns.inviscid()) calc_aux_vars();
        lap(1);
        FLOP_BUCKET(2);
        if (!alpha_given.empty()) {
            alpha = alpha_given;
            alpha_d = alpha_d_given;
        } else {
            calc_blender();
        }
        for (int c = 0; c < nc; ++c) blended_cell_stages += alpha[c] > 0.0;
        lap(2);
        FLOP_BUCKET(3);
        calc_surf_flux(2, s2);
        if (!ns.inviscid()) calc_surf_flux(3, s3);
        lap(3);
        double tho = 0, tlo = 0;
#pragma omp parallel reduction(+ : tho, tlo)
        {
            std::vector<State> hi(n), hd(n), lo(n);
            NS nsl = ns;
#pragma omp for schedule(static)
            for (int c = 0; c < nc; ++c) {
                auto a0 = std::chrono::steady_clock::now();
                FLOP_BUCKET(4);
                ho_inv(c, hi);
                FLOP_BUCKET(5);
                if (!ns.inviscid()) ho_dif(c, hd);
                auto a1 = std::chrono::steady_clock::now();
                FLOP_BUCKET(6);
                lo_inv(c, lo, nsl);
                FLOP_BUCKET(7);
                auto a2 = std::chrono::steady_clock::now();
                tho += std::chrono::duration<double>(a1 - a0).count();
                tlo += std::chrono::duration<double>(a2 - a1).count();
                const Real al = alpha[c], ad = alpha_d[c];
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < 5; ++v) {
                        Real& o = rhs[v * ndof + gid(c, i)];
                        if (!ns.inviscid()) o = (1 - ad) * hd[i][v] + al * lo[i][v] + (1 - al) * hi[i][v];
                        else o = al * lo[i][v] + (1 - al) * hi[i][v];
                    }
            }
        }
        {  // the loop's wall time split between the ho and lo residuals by their summed thread times
            const auto t1 = std::chrono::steady_clock::now();
            const double wall = std::chrono::duration<double>(t1 - t0).count(), tot = tho + tlo;
            timers[4] += tot > 0 ? wall * tho / tot : wall;
            timers[5] += tot > 0 ? wall * tlo / tot : 0.0;
            t0 = t1;
        }
        return 0;
    }

    // ---- PLENS::calc_time_step, dgsem/plens.cc:3810-3912: per-cell steps into loc_dt, then global
    //      stepping (every cell gets the minimum) or, past the start time, local stepping
    Real calc_time_step(Real Co, Real cur_time = 0.0) {
        // courant_function.value(Point<dim>()) at cur_time (dgsem/plens.cc:3816-3817)
        if (!x_courant.empty()) Co = eval_expression(x_courant, 0.0, 0.0, 0.0, cur_time);
        if (Co >= 1 && g_mutation != kMutDtNoCap) Co = 0.9;
        Real proc = 1e6;
        loc_dt.assign(nc, 0.0);
        for (int c = 0; c < nc; ++c) {
            Real cs = 1e6;
            const Real h = hmin[c];
            for (int i = 0; i < n; ++i) {
                const State s = cons(c, i);
                const Real a = ns.get_a(s);
                const Real vel[3] = {s[1] / s[0], s[2] / s[0], s[3] / s[0]};
                const Real* Ji = &Jinv[gid(c, i) * 9];
                Real adv[3] = {0, 0, 0};
                for (int row = 0; row < 3; ++row)
                    for (int col = 0; col < 3; ++col)  // JinvT
                        adv[row] += (g_mutation == kMutDtJinv ? Ji[row * 3 + col] : Ji[col * 3 + row]) * vel[col];
                if (g_mutation == kMutDtAbs)
                    for (Real& x : adv) x = fabs(x);
                const Real madv = std::max({adv[0], adv[1], adv[2]});
                const Real pw = pow(N, g_mutation == kMutDtDegreeSq ? 2.0 : 1.5);
                const Real cur = Co / (pw * (madv + a / h + N * N * mu[gid(c, i)] / (h * h * s[0])));
                if (cur < cs) cs = cur;
            }
            loc_dt[c] = cs;
            if (cs < proc) proc = cs;
        }
        if (lts_request && cur_time > lts_start) {
            const std::vector<Real> gh = loc_dt;  // gh_loc_time_steps
            for (int c = 0; c < nc; ++c) {
                Real m = gh[c];
                if (lts_limit) {
                    for (int f = 0; f < 6; ++f)
                        if (bid[(size_t)c * 6 + f] < 0) m = std::min(m, lts_nf * gh[nbr[(size_t)c * 6 + f]]);
                    m = std::min(m, lts_gf * proc);
                }
                loc_dt[c] = m;
            }
        } else {
            for (int c = 0; c < nc; ++c) loc_dt[c] = proc;
        }
        return proc;
    }

    // ---- PLENS::calc_ss_error, dgsem/plens.cc:3985-4035: L2 norm of (rhoE - rhoE_old)/dt_cell over
    //      the L2 norm of rhoE, both by QGauss(N+1) through the mapping
    Real calc_ss_error() const {
        std::vector<Real> xg, wg;
        gauss(np, xg, wg);
        std::vector<Real> L((size_t)np * np);  // L[q][i] = l_i(xg_q), GLL Lagrange basis
        for (int q = 0; q < np; ++q)
            for (int i = 0; i < np; ++i) L[q * np + i] = (Real)lagrange_val(x1, i, xg[q]);
        Real num = 0, den = 0;
        for (int c = 0; c < nc; ++c) {
            for (int qk = 0; qk < np; ++qk)
                for (int qj = 0; qj < np; ++qj)
                    for (int qi = 0; qi < np; ++qi) {
                        const Real xi[3] = {xg[qi], xg[qj], xg[qk]};
                        Real J[3][3];
                        jacobian(c, xi, J);
                        const Real det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                        const Real jxw = fabs(det) * wg[qi] * wg[qj] * wg[qk];
                        Real e = 0, d = 0;
                        for (int k = 0; k < np; ++k)
                            for (int j = 0; j < np; ++j)
                                for (int i = 0; i < np; ++i) {
                                    const Real lw = L[qi * np + i] * L[qj * np + j] * L[qk * np + k];
                                    const size_t gi = gid(c, tl(i, j, k));
                                    const Real E = Q[4 * ndof + gi];
                                    e += lw * E;
                                    d += lw * (E - rhoE_old[gi]) / loc_dt[c];
                                }
                        den += e * e * jxw;
                        num += d * d * jxw;
                    }
        }
        return sqrt(num) / sqrt(den);
    }

    // ---- PLENS::update + update_rk4, dgsem/plens.cc:4334-4403, 4484-4518 --------------------------
    int rk4_step(Real Co, double* t, Real fixed_dt = 0) {
        if (t) bc_time = *t;  // BC::set_time(cur_time), dgsem/plens.cc:1798
        // gcrk_cvars = g_cvars (the state lives in Q throughout)
        int e = calc_rhs();
        if (e) return e;
        auto t0 = std::chrono::steady_clock::now();
        FLOP_BUCKET(8);
        const Real dt = fixed_dt > 0 ? fixed_dt : calc_time_step(Co, *t);
        if (fixed_dt > 0) loc_dt.assign(nc, fixed_dt);
        std::vector<Real> E0(Q.begin() + 4 * ndof, Q.begin() + 5 * ndof);  // g_cvars[4] at the start
        auto t1 = std::chrono::steady_clock::now();
        timers[6] += std::chrono::duration<double>(t1 - t0).count();
        for (int stage = 1; stage <= 5; ++stage) {
            if (stage > 1) {
                e = calc_rhs();
                if (e) return e;
            }
            auto u0 = std::chrono::steady_clock::now();
            FLOP_BUCKET(9);
            for (int v = 0; v < 5; ++v)  // multiply_time_step_to_rhs (per-cell steps)
                for (size_t i = 0; i < ndof; ++i) rhs[v * ndof + i] *= loc_dt[i / n];
            if (stage == 1) {
                for (size_t i = 0; i < 5 * ndof; ++i) {
                    Q[i] = Q[i] + rk.ao[0] * rhs[i];
                    rhs1[i] = rhs[i];
                }
            } else if (stage == 2) {
                for (size_t i = 0; i < 5 * ndof; ++i) {
                    Q[i] = Q[i] + ((rk.ai[0] - rk.ao[0]) * rhs1[i] + rk.ao[1] * rhs[i]);
                    rhs2[i] = rhs1[i];
                    rhs1[i] = rhs[i];
                }
            } else {
                const int s = stage;
                for (size_t i = 0; i < 5 * ndof; ++i) {
                    Q[i] = Q[i] + ((rk.b[s - 3] - rk.ai[s - 3]) * rhs2[i] +
                                   (rk.ai[s - 2] - rk.ao[s - 2]) * rhs1[i] + rk.ao[s - 1] * rhs[i]);
                    rhs2[i] = rhs1[i];
                    rhs1[i] = rhs[i];
                }
            }
            auto u1 = std::chrono::steady_clock::now();
            timers[7] += std::chrono::duration<double>(u1 - u0).count();
        }
        rhoE_old = E0;  // update(), dgsem/plens.cc:4507-4508
        *t += dt;
        return 0;
    }
};

// =============================================================================================
// C API
// =============================================================================================
extern "C" {

int oracle_create(int32_t n_cells, int32_t degree, const double* vertices, const int32_t* face_nbr,
                  const int32_t* face_bid, const oracle_physics* phys, const oracle_bc* bcs, int32_t n_bcs,
                  int32_t transpose_y, oracle** out) {
    if (n_cells < 1 || degree < 1 || degree > 8 || !vertices || !face_nbr || !face_bid || !phys || !out)
        return -1;
    oracle* o = new (std::nothrow) oracle;
    if (!o) return -4;
    o->N = degree; o->np = degree + 1; o->n = o->np * o->np * o->np; o->nf = o->np * o->np;
    o->nc = n_cells; o->ndof = (size_t)n_cells * o->n;
    o->ph = *phys;
    o->ns = NS{phys->gamma, phys->molecular_weight, phys->prandtl, phys->mu0, phys->T0, phys->S,
               phys->inv_surf_flux};
    o->ns.vscheme = phys->inv_vol_flux;
    o->bcs.assign(bcs, bcs + n_bcs);
    o->bexpr.assign(std::max(n_bcs, 0), {});
    for (int i = 0; i < n_bcs; ++i) {
        if (bcs[i].type != 9) continue;
        if (!bcs[i].p_expr || !bcs[i].T_expr || !bcs[i].vel_expr) { delete o; return -1; }
        std::vector<std::string> e = {bcs[i].p_expr, bcs[i].T_expr};
        for (const auto& v : split_components(bcs[i].vel_expr)) e.push_back(v);
        if (e.size() != 5) { delete o; return -1; }
        o->bexpr[i] = e;
    }
    for (size_t k = 0; k < o->bcs.size(); ++k) {  // the pointers are the caller's; keep copies only
        o->bcs[k].p_expr = o->bcs[k].T_expr = o->bcs[k].vel_expr = nullptr;
    }
    o->nbr.assign(face_nbr, face_nbr + (size_t)n_cells * 6);
    o->bid.assign(face_bid, face_bid + (size_t)n_cells * 6);
    gll(degree, o->x1, o->w);
    const int np = o->np;
    o->D.resize(np * np); o->Qm.resize(np * np);
    for (int r = 0; r < np; ++r)   // calc_metric_terms, plens.cc:1650-1663
        for (int c = 0; c < np; ++c) {
            o->D[r * np + c] = lagrange_deriv(o->x1, c, o->x1[r]);
            o->Qm[r * np + c] = o->w[r] * o->D[r * np + c];
        }
    cob_matrix(degree, o->C);
    o->verts.assign(vertices, vertices + (size_t)n_cells * 24);
    o->metrics(o->verts.data(), transpose_y);
    const size_t nd = o->ndof;
    o->Q.assign(5 * nd, 0); o->avars.assign(9 * nd, 0); o->mu.assign(nd, 0); o->kap.assign(nd, 0);
    o->alpha.assign(n_cells, 0); o->alpha_d.assign(n_cells, 0);
    o->rhs.assign(5 * nd, 0); o->rhs1.assign(5 * nd, 0); o->rhs2.assign(5 * nd, 0);
    *out = o;
    return 0;
}

int oracle_create_mapped(int32_t n_cells, int32_t degree, int32_t map_degree, const double* support,
                         const int32_t* face_nbr, const int32_t* face_bid, const oracle_physics* phys,
                         const oracle_bc* bcs, int32_t n_bcs, int32_t transpose_y, oracle** out) {
    if (map_degree < 1 || map_degree > 8 || !support || n_cells < 1) return -1;
    const int qp = map_degree + 1, nsp = qp * qp * qp;
    std::vector<double> corners((size_t)n_cells * 24);
    for (int c = 0; c < n_cells; ++c)
        for (int v = 0; v < 8; ++v) {
            const int s = (v & 1) * map_degree + qp * (((v >> 1) & 1) * map_degree + qp * (((v >> 2) & 1) * map_degree));
            for (int d = 0; d < 3; ++d)
                corners[(size_t)c * 24 + 3 * v + d] = support[((size_t)c * nsp + s) * 3 + d];
        }
    if (map_degree == 1)
        return oracle_create(n_cells, degree, corners.data(), face_nbr, face_bid, phys, bcs, n_bcs, transpose_y, out);
    // build on the corners, then redo the metric terms from the degree-q map
    int e = oracle_create(n_cells, degree, corners.data(), face_nbr, face_bid, phys, bcs, n_bcs, transpose_y, out);
    if (e) return e;
    oracle* o = *out;
    o->mq = map_degree;
    o->sup.assign(support, support + (size_t)n_cells * nsp * 3);
    std::vector<Real> wq;
    gll(map_degree, o->tq, wq);
    o->metrics(o->verts.data(), transpose_y);
    return 0;
}

void oracle_destroy(oracle* o) { delete o; }

int oracle_set_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int oracle_set_state(oracle* o, const double* Q) {
    std::memcpy(o->Q.data(), Q, sizeof(double) * 5 * o->ndof);
    return 0;
}
int oracle_get_state(const oracle* o, double* Q) {
    std::memcpy(Q, o->Q.data(), sizeof(double) * 5 * o->ndof);
    return 0;
}

int oracle_calc_rhs(oracle* o, double* rhs, double* alpha, double* alpha_d, double* avars) {
    const int e = o->calc_rhs();
    if (e) return e;
    if (rhs) std::memcpy(rhs, o->rhs.data(), sizeof(double) * 5 * o->ndof);
    if (alpha) std::memcpy(alpha, o->alpha.data(), sizeof(double) * o->nc);
    if (alpha_d) std::memcpy(alpha_d, o->alpha_d.data(), sizeof(double) * o->nc);
    if (avars) std::memcpy(avars, o->avars.data(), sizeof(double) * 9 * o->ndof);
    return 0;
}

int oracle_calc_dt(oracle* o, double courant, double* dt) {
    // mu from the state, as calc_aux_vars does before calc_time_step
    const double cvv = o->ns.cv();
    for (size_t i = 0; i < o->ndof; ++i) {
        State s;
        for (int v = 0; v < 5; ++v) s[v] = o->Q[v * o->ndof + i];
        o->mu[i] = o->ns.inviscid() ? Real(0.0) : o->ns.mu(NS::get_e(s) / cvv);
    }
    *dt = o->calc_time_step(courant, o->bc_time);
    return 0;
}

int oracle_rk4_step(oracle* o, double courant, double* t) { return o->rk4_step(courant, t); }

int oracle_node_location(const oracle* o, int32_t c, int32_t node, double* xyz) {
    if (c < 0 || c >= o->nc || node < 0 || node >= o->n) return -1;
    const T1 X = o->node_location(c, node);
    for (int d = 0; d < 3; ++d) xyz[d] = X[d];
    return 0;
}

int oracle_set_time(oracle* o, double t) {
    o->bc_time = t;
    return 0;
}

int oracle_set_alpha(oracle* o, const double* alpha, const double* alpha_d) {
    if (!alpha || !alpha_d) {
        o->alpha_given.clear();
        o->alpha_d_given.clear();
        return 0;
    }
    o->alpha_given.assign(alpha, alpha + o->nc);
    o->alpha_d_given.assign(alpha_d, alpha_d + o->nc);
    return 0;
}

int oracle_set_raw_alpha(oracle* o, const double* raw) {
    if (!raw) o->raw_given.clear();
    else o->raw_given.assign(raw, raw + o->nc);
    return 0;
}

int oracle_set_expressions(oracle* o, const char* courant, const char* wall_blender_limit,
                           const char* viscous_blending_region) {
    if (!o) return -1;
    o->x_courant = courant ? courant : "";
    o->x_wall = wall_blender_limit ? wall_blender_limit : "";
    o->x_region = viscous_blending_region ? viscous_blending_region : "";
    return 0;
}

int oracle_blended_cell_stages(const oracle* o, int64_t* count) {
    if (!o || !count) return -1;
    *count = o->blended_cell_stages;
    return 0;
}

int oracle_set_mutation(int32_t id) {
    g_mutation = id;
    return 0;
}

int oracle_bc_ghost(const oracle* o, int32_t stage, int32_t cell, int32_t face, int32_t q, const double* in,
                    const double* ain, const double* nrm, double* g, double* ag) {
    if (!o || !in || !ain || !nrm || !g || !ag || stage < 1 || stage > 3 || cell < 0 || cell >= o->nc ||
        face < 0 || face > 5 || q < 0 || q >= o->nf || o->bid[(size_t)cell * 6 + face] < 0)
        return -1;
    State s, gs;
    Avars a, ga;
    T1 n;
    for (int v = 0; v < 5; ++v) s[v] = in[v];
    for (int v = 0; v < 9; ++v) a[v] = ain[v];
    for (int d = 0; d < 3; ++d) n[d] = nrm[d];
    o->ghost(stage, cell, face, q, s, a, n, gs, ga);
    for (int v = 0; v < 5; ++v) g[v] = gs[v];
    for (int v = 0; v < 9; ++v) ag[v] = ga[v];
    return 0;
}

double oracle_eval_expression(const char* expr, double x, double y, double z, double t) {
    return eval_expression(expr ? expr : "", x, y, z, t);
}

// PLENS::write() fields (dgsem/plens.cc:4170-4322): write() follows calc_rhs(true) and
// calc_time_step() inside the first RK stage (4341-4350); post_process (3950-3963) gives p, T and
// the velocity at the DoFs; DataOut evaluates every field's FE interpolant at the equispaced
// (N+1)^3 points of each cell (build_patches with N subdivisions), here by direct summation over
// all nodal basis functions.
int oracle_output_fields(oracle* o, double courant, double t, double* nodal, double* cell) {
    o->bc_time = t;
    int e = o->calc_rhs();
    if (e) return e;
    o->calc_time_step(courant, t);
    const int np = o->np, n = o->n, N = o->N;
    const size_t nd = o->ndof;
    std::vector<double> L((size_t)np * np);  // L[e][i] = l_i(e/N)
    for (int ee = 0; ee < np; ++ee)
        for (int i = 0; i < np; ++i) L[ee * np + i] = (double)lagrange_val(o->x1, i, (long double)ee / N);
    const double cv = o->ns.cv();
    std::vector<double> f((size_t)15 * n);
    for (int c = 0; c < o->nc; ++c) {
        for (int i = 0; i < n; ++i) {
            const State s = o->cons(c, i);
            for (int a = 0; a < 9; ++a) f[(size_t)a * n + i] = o->avars[a * nd + o->gid(c, i)];
            f[(size_t)9 * n + i] = s[0];
            f[(size_t)10 * n + i] = o->ns.get_p(s);
            f[(size_t)11 * n + i] = NS::get_e(s) / cv;
            for (int d = 0; d < 3; ++d) f[(size_t)(12 + d) * n + i] = s[1 + d] / s[0];
        }
        for (int ek = 0; ek < np; ++ek)
            for (int ej = 0; ej < np; ++ej)
                for (int ei = 0; ei < np; ++ei) {
                    const int pt = ei + np * (ej + np * ek);
                    for (int q = 0; q < 15; ++q) {
                        double v = 0;
                        for (int k = 0; k < np; ++k)
                            for (int j = 0; j < np; ++j)
                                for (int i = 0; i < np; ++i)
                                    v += f[(size_t)q * n + o->tl(i, j, k)] * L[ei * np + i] * L[ej * np + j] * L[ek * np + k];
                        nodal[(size_t)q * nd + (size_t)c * n + pt] = v;
                    }
                }
        cell[c] = o->alpha[c];
        cell[o->nc + c] = o->alpha_d[c];
        cell[2 * (size_t)o->nc + c] = o->loc_dt[c];
    }
    return 0;
}

int oracle_rk4_step_fixed_dt(oracle* o, double dt, double* t) { return o->rk4_step(0.0, t, dt); }

int oracle_run(oracle* o, int32_t n_steps, double courant, double* t) {
    for (int s = 0; s < n_steps; ++s) {
        const int e = o->rk4_step(courant, t);
        if (e) return e;
    }
    return 0;
}

int oracle_last_failure(const oracle* o, int64_t* dof, int32_t* kind) {
    if (dof) *dof = o->fail_dof;
    if (kind) *kind = o->fail_kind;
    return 0;
}

int oracle_set_local_stepping(oracle* o, int32_t request, double start_time, int32_t limit,
                               double neighbor_factor, double global_factor) {
    o->lts_request = request;
    o->lts_start = start_time;
    o->lts_limit = limit;
    o->lts_nf = neighbor_factor;
    o->lts_gf = global_factor;
    return 0;
}

int oracle_loc_dt(const oracle* o, double* dt) {
    if ((int)o->loc_dt.size() != o->nc) return -1;
    for (int c = 0; c < o->nc; ++c) dt[c] = o->loc_dt[c];
    return 0;
}

int oracle_ss_error(const oracle* o, double* err) {
    if (o->rhoE_old.size() != o->ndof || (int)o->loc_dt.size() != o->nc) return -1;  // no step taken yet
    *err = o->calc_ss_error();
    return 0;
}

// flop counters of the counting build (liboracle_flops.so) per bucket and the number of entries
// into each nested bucket; -1 in the oracle proper
int oracle_flop_counts(uint64_t* flops, uint64_t* calls, int32_t n, int32_t reset) {
#ifdef ORACLE_COUNT_FLOPS
    for (int i = 0; i < n && i < flopcount::kBuckets; ++i) {
        if (flops) flops[i] = flopcount::g_flops[i];
        if (calls) calls[i] = flopcount::g_calls[i];
    }
    if (reset)
        for (int i = 0; i < flopcount::kBuckets; ++i) flopcount::g_flops[i] = flopcount::g_calls[i] = 0;
    return flopcount::kBuckets;
#else
    (void)flops; (void)calls; (void)n; (void)reset;
    return -1;
#endif
}

int oracle_section_times(const oracle* o, double* t, int32_t n) {
    for (int i = 0; i < n && i < 8; ++i) t[i] = o->timers[i];
    return 0;
}

int oracle_cell_metrics(const oracle* o, int32_t c, double* Ja, double* detJ, double* normals) {
    const int n = o->n, nf = o->nf;
    for (int node = 0; node < n; ++node) {
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < 3; ++a) Ja[(3 * d + a) * n + node] = o->JaV(c, node, d)[a];
        detJ[node] = o->detJ[o->gid(c, node)];
    }
    for (int f = 0; f < 6; ++f)
        for (int d = 0; d < 3; ++d)
            for (int q = 0; q < nf; ++q)
                normals[(f * 3 + d) * nf + q] = o->normals[(((size_t)c * 6 + f) * nf + q) * 3 + d];
    return 0;
}

int oracle_cell_subcell_normals(const oracle* o, int32_t c, double* sn) {
    const int N = o->N, np = o->np;
    size_t k = 0;
    for (int dir = 0; dir < 3; ++dir)
        for (int id = 0; id <= N + 1; ++id)
            for (int id1 = 0; id1 < np; ++id1)
                for (int id2 = 0; id2 < np; ++id2) {
                    const T1& s = o->sub[o->sub_index(c, dir, id, id1, id2)];
                    for (int a = 0; a < 3; ++a) sn[k++] = s[a];
                }
    return 0;
}

int oracle_basis(int32_t N, double* x, double* w, double* D) {
    std::vector<Real> xx, ww;
    gll(N, xx, ww);
    for (int i = 0; i <= N; ++i) { x[i] = xx[i]; w[i] = ww[i]; }
    if (D)
        for (int r = 0; r <= N; ++r)
            for (int c = 0; c <= N; ++c) D[r * (N + 1) + c] = lagrange_deriv(xx, c, xx[r]);
    return 0;
}

int oracle_cob_matrix(int32_t N, double* C) {
    std::vector<Real> c;
    cob_matrix(N, c);
    for (size_t i = 0; i < c.size(); ++i) C[i] = c[i];
    return 0;
}

double oracle_log_avg(double xl, double xr) { return NS::log_avg(xl, xr); }

static NS make_ns(const oracle_physics* p, int scheme) {
    NS ns{p->gamma, p->molecular_weight, p->prandtl, p->mu0, p->T0, p->S, scheme};
    ns.vscheme = p->inv_vol_flux;
    return ns;
}

void oracle_inv_surf_flux(const oracle_physics* p, int32_t scheme, double blender, const double* ul,
                          const double* ur, const double* n, double* f) {
    NS ns = make_ns(p, scheme);
    ns.set_blender(blender);
    State l, r, fl;
    for (int v = 0; v < 5; ++v) { l[v] = ul[v]; r[v] = ur[v]; }
    T1 nn; nn[0] = n[0]; nn[1] = n[1]; nn[2] = n[2];
    ns.get_inv_surf_flux(l, r, nn, fl);
    for (int v = 0; v < 5; ++v) f[v] = fl[v];
}

void oracle_inv_flux(const oracle_physics* p, const double* u, const double* dir, double* f) {
    NS ns = make_ns(p, 0);
    State s, fl;
    for (int v = 0; v < 5; ++v) s[v] = u[v];
    T1 d; d[0] = dir[0]; d[1] = dir[1]; d[2] = dir[2];
    ns.get_inv_flux(s, d, fl);
    for (int v = 0; v < 5; ++v) f[v] = fl[v];
}

void oracle_vol_flux(const oracle_physics* p, const double* u1, const double* u2, const double* dir, double* f) {
    NS ns = make_ns(p, 0);
    State a, b, fl;
    for (int v = 0; v < 5; ++v) { a[v] = u1[v]; b[v] = u2[v]; }
    T1 d; d[0] = dir[0]; d[1] = dir[1]; d[2] = dir[2];
    ns.vol_flux(a, b, d, fl);
    for (int v = 0; v < 5; ++v) f[v] = fl[v];
}

void oracle_dif_flux(const double* u, const double* av, const double* dir, double* f) {
    State s, fl;
    Avars a;
    for (int v = 0; v < 5; ++v) s[v] = u[v];
    for (int v = 0; v < 9; ++v) a[v] = av[v];
    T1 d; d[0] = dir[0]; d[1] = dir[1]; d[2] = dir[2];
    NS::dif_flux(s, a, d, fl);
    for (int v = 0; v < 5; ++v) f[v] = fl[v];
}

double oracle_blender_value(const oracle_physics* p, int32_t N, const double* var) {
    std::vector<Real> C;
    cob_matrix(N, C);
    NS ns = make_ns(p, 0);
    const std::vector<Real> vals(var, var + (size_t)(N + 1) * (N + 1) * (N + 1));
    return blender_from_values(ns, *p, N, C, vals.data());
}

void oracle_rk4_coeffs(double* ao, double* ai, double* b) {
    RKC r;
    for (int i = 0; i < 5; ++i) { ao[i] = r.ao[i]; b[i] = r.b[i]; }
    for (int i = 0; i < 4; ++i) ai[i] = r.ai[i];
}

}  // extern "C"

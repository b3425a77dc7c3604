// Host setup of the DGSEM hot path: structured hex mesh, GLL basis, metric terms, face normals,
// neighbour tables and the initial conditions of the workloads. deal.II-free restatement of the
// PLENS setup members listed in include/plens_host.h.
#include "plens_host.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <cstring>
#include <new>
#include <random>
#include <vector>

namespace {

constexpr double kR0 = 8.314462618;  // NavierStokes::R0, modelling/navier_stokes.h:149

// ---------------------------------------------------------------------------------------------
// 1D bases on [0,1]
// ---------------------------------------------------------------------------------------------

// Legendre P_0..P_n at x in [-1,1] (three-term recurrence), long double.
void legendre_all(int n, long double x, long double* P) {
    P[0] = 1;
    if (n >= 1) P[1] = x;
    for (int k = 2; k <= n; ++k) P[k] = ((2 * k - 1) * x * P[k - 1] - (k - 1) * P[k - 2]) / k;
}

// Gauss-Lobatto-Legendre nodes/weights with N+1 points on [0,1]; Newton iteration on
// (1-x^2) P_N'(x) = 0 written as x P_N - P_{N-1} = 0 (Chebyshev-Lobatto start), symmetrised.
void gll_01(int N, double* x01, double* w01) {
    std::vector<long double> x(N + 1), P(N + 2);
    for (int j = 0; j <= N; ++j) x[j] = -cosl(3.14159265358979323846264338327950288L * j / N);
    for (int it = 0; it < 100; ++it) {
        long double dmax = 0;
        for (int j = 0; j <= N; ++j) {
            legendre_all(N, x[j], P.data());
            long double dx = (x[j] * P[N] - P[N - 1]) / ((N + 1) * P[N]);
            x[j] -= dx;
            dmax = std::max(dmax, fabsl(dx));
        }
        if (dmax < 1e-30L) break;
    }
    std::vector<long double> w(N + 1);
    for (int j = 0; j <= N; ++j) {
        legendre_all(N, x[j], P.data());
        w[j] = 2.0L / (N * (N + 1) * P[N] * P[N]);
    }
    x[0] = -1; x[N] = 1;
    for (int j = 0; j <= N; ++j) {
        x01[j] = (double)((x[j] + 1) / 2);
        w01[j] = (double)(w[j] / 2);
    }
    for (int j = 0; j < (N + 1) / 2; ++j) {  // exact symmetry about 1/2
        x01[N - j] = 1.0 - x01[j];
        w01[N - j] = w01[j];
    }
    if (N % 2 == 0) x01[N / 2] = 0.5;
}

// Gauss-Legendre nodes/weights with m points on [0,1].
void gauss_01(int m, long double* xg, long double* wg) {
    std::vector<long double> P(m + 2);
    for (int j = 0; j < m; ++j) {
        long double x = -cosl(3.14159265358979323846264338327950288L * (j + 0.75L) / (m + 0.5L));
        for (int it = 0; it < 100; ++it) {
            legendre_all(m, x, P.data());
            long double dP = m * (x * P[m] - P[m - 1]) / (x * x - 1);
            long double dx = P[m] / dP;
            x -= dx;
            if (fabsl(dx) < 1e-30L) break;
        }
        legendre_all(m, x, P.data());
        long double dP = m * (x * P[m] - P[m - 1]) / (x * x - 1);
        xg[j] = (x + 1) / 2;
        wg[j] = 1.0L / ((1 - x * x) * dP * dP);  // 2/((1-x^2)P'^2) halved for [0,1]
    }
}

// Lagrange basis l_j(x) through nodes xn[0..N].
long double lagrange(int N, const double* xn, int j, long double x) {
    long double v = 1;
    for (int k = 0; k <= N; ++k)
        if (k != j) v *= (x - xn[k]) / ((long double)xn[j] - xn[k]);
    return v;
}

}  // namespace

extern "C" int plens_basis_1d(int32_t N, double* x, double* w, double* D, double* C1) {
    if (N < 1 || N > 8 || !x || !w) return PLENS_E_INVALID;
    gll_01(N, x, w);
    const int np = N + 1;
    if (D) {
        // barycentric differentiation matrix, diagonal by the negative-sum rule
        std::vector<long double> lam(np);
        for (int j = 0; j < np; ++j) {
            long double p = 1;
            for (int k = 0; k < np; ++k)
                if (k != j) p *= ((long double)x[j] - x[k]);
            lam[j] = 1 / p;
        }
        for (int i = 0; i < np; ++i) {
            long double s = 0;
            for (int j = 0; j < np; ++j) {
                if (j == i) continue;
                long double d = (lam[j] / lam[i]) / ((long double)x[i] - x[j]);
                D[i * np + j] = (double)d;
                s += d;
            }
            D[i * np + i] = (double)(-s);
        }
    }
    if (C1) {
        std::vector<long double> xg(np), wg(np), P(np + 1);
        gauss_01(np, xg.data(), wg.data());
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < np; ++j) {
                long double s = 0;
                for (int q = 0; q < np; ++q) {
                    legendre_all(N, 2 * xg[q] - 1, P.data());
                    s += wg[q] * sqrtl(2.0L * i + 1) * P[i] * lagrange(N, x, j, xg[q]);
                }
                C1[i * np + j] = (double)s;
            }
    }
    return PLENS_OK;
}

extern "C" void plens_prim_to_cons(double gamma, double rho, const double vel[3], double p,
                                   double cons[5]) {
    // same operation order as NavierStokes::prim_to_cons (modelling/navier_stokes.h:365-378)
    cons[0] = rho;
    cons[4] = p / (gamma - 1);
    for (int d = 0; d < 3; ++d) {
        cons[1 + d] = rho * vel[d];
        cons[4] += 0.5 * rho * vel[d] * vel[d];
    }
}

struct plens_mesh {
    int32_t N = 0, n = 0, nf = 0, n_cells = 0;
    int32_t nx = 0, ny = 0, nz = 0;
    int32_t q = 1;                 // mapping degree
    std::vector<double> vertices, nodes, Ja, detJ, normals, hmin;
    std::vector<double> support;   // [n_cells][(q+1)^3][3]
    std::vector<int32_t> face_nbr, face_bid;
    std::vector<double> x1d, w1d;
};

namespace {

// FaceDoFInfo::form_maps (dgsem/face_dof_info.h:92-150): cell-local node of face-local dof q.
inline int face_node(int N, int f, int q) {
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

struct Vec3 { double x, y, z; };
inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// trilinear Jacobian columns a_d = dx/dxi_d at reference point xi for the 8 deal.II-ordered vertices
void trilinear_jac(const double* V, const double xi[3], Vec3 a[3], double X[3]) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    X[0] = X[1] = X[2] = 0;
    for (int v = 0; v < 8; ++v) {
        const int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
        double f[3], df[3];
        for (int d = 0; d < 3; ++d) {
            f[d] = b[d] ? xi[d] : 1.0 - xi[d];
            df[d] = b[d] ? 1.0 : -1.0;
        }
        const double N = f[0] * f[1] * f[2];
        const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
        for (int i = 0; i < 3; ++i) {
            X[i] += N * V[3 * v + i];
            for (int d = 0; d < 3; ++d) J[i][d] += dN[d] * V[3 * v + i];
        }
    }
    for (int d = 0; d < 3; ++d) a[d] = {J[0][d], J[1][d], J[2][d]};
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double unit_uniform(uint64_t& s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" int plens_mesh_build_box(const plens_box_params* p, plens_mesh** out) {
    if (!p || !out) return PLENS_E_INVALID;
    if (p->degree < 1 || p->degree > 5 || p->nx < 1 || p->ny < 1 || p->nz < 1) return PLENS_E_INVALID;
    plens_mesh* m = new (std::nothrow) plens_mesh;
    if (!m) return PLENS_E_NOMEM;
    const int N = p->degree, np = N + 1, n = np * np * np, nf = np * np;
    m->N = N; m->n = n; m->nf = nf;
    m->nx = p->nx; m->ny = p->ny; m->nz = p->nz;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int nc = nx * ny * nz;
    m->n_cells = nc;
    m->x1d.resize(np); m->w1d.resize(np);
    plens_basis_1d(N, m->x1d.data(), m->w1d.data(), nullptr, nullptr);

    // ---- vertex lattice ------------------------------------------------------------------
    const int vx = nx + 1, vy = ny + 1, vz = nz + 1;
    std::vector<double> gx(vx), gy(vy), gz(vz), hx(nx), hy(ny), hz(nz);
    for (int i = 0; i < vx; ++i) gx[i] = p->lo[0] + (p->hi[0] - p->lo[0]) * i / nx;
    for (int k = 0; k < vz; ++k) gz[k] = p->lo[2] + (p->hi[2] - p->lo[2]) * k / nz;
    const double r = p->y_stretch_ratio > 0 ? p->y_stretch_ratio : 1.0;
    if (std::fabs(r - 1.0) < 1e-14) {
        for (int j = 0; j < vy; ++j) gy[j] = p->lo[1] + (p->hi[1] - p->lo[1]) * j / ny;
    } else {
        const double dy0 = (p->hi[1] - p->lo[1]) * (r - 1.0) / (std::pow(r, ny) - 1.0);
        for (int j = 0; j < vy; ++j) gy[j] = p->lo[1] + dy0 * (std::pow(r, j) - 1.0) / (r - 1.0);
        gy[ny] = p->hi[1];
    }
    for (int i = 0; i < nx; ++i) hx[i] = gx[i + 1] - gx[i];
    for (int j = 0; j < ny; ++j) hy[j] = gy[j + 1] - gy[j];
    for (int k = 0; k < nz; ++k) hz[k] = gz[k + 1] - gz[k];

    // smooth seeded displacement: per component, integer wave numbers in {1,2,3} and a phase
    uint64_t s = p->seed * 0x2545F4914F6CDD1Dull + 12345;
    int kw[3][3];
    double ph[3];
    for (int c = 0; c < 3; ++c) {
        for (int d = 0; d < 3; ++d) kw[c][d] = 1 + (int)(splitmix(s) % 3);
        ph[c] = 2.0 * M_PI * unit_uniform(s);
    }
    const int ncell_dir[3] = {nx, ny, nz};
    std::vector<double> V((size_t)vx * vy * vz * 3);
    for (int k = 0; k < vz; ++k)
        for (int j = 0; j < vy; ++j)
            for (int i = 0; i < vx; ++i) {
                const int I[3] = {i, j, k};
                double X[3] = {gx[i], gy[j], gz[k]};
                if (p->perturb_frac > 0) {
                    double xhat[3];
                    for (int d = 0; d < 3; ++d) {
                        int Id = p->periodic[d] ? I[d] % ncell_dir[d] : I[d];
                        xhat[d] = (double)Id / ncell_dir[d];
                    }
                    const double hloc[3] = {
                        std::min(hx[std::max(i - 1, 0)], hx[std::min(i, nx - 1)]),
                        std::min(hy[std::max(j - 1, 0)], hy[std::min(j, ny - 1)]),
                        std::min(hz[std::max(k - 1, 0)], hz[std::min(k, nz - 1)])};
                    for (int c = 0; c < 3; ++c) {
                        const bool on_bnd = !p->periodic[c] && (I[c] == 0 || I[c] == ncell_dir[c]);
                        if (on_bnd) continue;
                        const double arg = 2.0 * M_PI * (kw[c][0] * xhat[0] + kw[c][1] * xhat[1] +
                                                         kw[c][2] * xhat[2]) + ph[c];
                        X[c] += p->perturb_frac * hloc[c] * std::sin(arg);
                    }
                }
                double* dst = &V[3 * ((size_t)i + vx * ((size_t)j + (size_t)vy * k))];
                dst[0] = X[0]; dst[1] = X[1]; dst[2] = X[2];
            }

    // ---- per-cell arrays ----------------------------------------------------------------------
    m->vertices.resize((size_t)nc * 24);
    m->nodes.resize((size_t)nc * 3 * n);
    m->Ja.resize((size_t)nc * 9 * n);
    m->detJ.resize((size_t)nc * n);
    m->normals.resize((size_t)nc * 18 * nf);
    m->hmin.resize(nc);
    m->face_nbr.assign((size_t)nc * 6, -1);
    m->face_bid.assign((size_t)nc * 6, -1);
    const double* xg = m->x1d.data();
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int c = i + nx * (j + ny * k);
                double* cv = &m->vertices[(size_t)c * 24];
                for (int v = 0; v < 8; ++v) {
                    const int ii = i + (v & 1), jj = j + ((v >> 1) & 1), kk = k + ((v >> 2) & 1);
                    const double* src = &V[3 * ((size_t)ii + vx * ((size_t)jj + (size_t)vy * kk))];
                    cv[3 * v] = src[0]; cv[3 * v + 1] = src[1]; cv[3 * v + 2] = src[2];
                }
                // minimum vertex distance over all 28 pairs (TriaAccessor::minimum_vertex_distance)
                double d2min = 1e300;
                for (int a = 0; a < 8; ++a)
                    for (int b = a + 1; b < 8; ++b) {
                        double d2 = 0;
                        for (int q = 0; q < 3; ++q) {
                            const double t = cv[3 * a + q] - cv[3 * b + q];
                            d2 += t * t;
                        }
                        d2min = std::min(d2min, d2);
                    }
                m->hmin[c] = std::sqrt(d2min);
                // metric terms at GLL nodes (MetricTerms::reinit, dgsem/metric_terms.cc:82-102)
                double* Ja = &m->Ja[(size_t)c * 9 * n];
                double* dJ = &m->detJ[(size_t)c * n];
                double* XN = &m->nodes[(size_t)c * 3 * n];
                for (int nk = 0; nk < np; ++nk)
                    for (int nj = 0; nj < np; ++nj)
                        for (int ni = 0; ni < np; ++ni) {
                            const int node = ni + np * (nj + np * nk);
                            const double xi[3] = {xg[ni], xg[nj], xg[nk]};
                            Vec3 a[3];
                            double X[3];
                            trilinear_jac(cv, xi, a, X);
                            for (int d = 0; d < 3; ++d) XN[d * n + node] = X[d];
                            const Vec3 c0 = cross(a[1], a[2]), c1 = cross(a[2], a[0]),
                                       c2 = cross(a[0], a[1]);
                            const Vec3 cc[3] = {c0, c1, c2};
                            for (int d = 0; d < 3; ++d) {
                                Ja[(3 * d + 0) * n + node] = cc[d].x;
                                Ja[(3 * d + 1) * n + node] = cc[d].y;
                                Ja[(3 * d + 2) * n + node] = cc[d].z;
                            }
                            dJ[node] = a[0].x * c0.x + a[0].y * c0.y + a[0].z * c0.z;
                        }
                // outward unit face normals at face dofs, FEFaceValues::normal_vector(face_dof):
                // on faces 2/3 deal.II's face quadrature runs z-fastest, so face dof q reads the
                // normal of the transposed point when transpose_y_face_normals is set
                double* Nrm = &m->normals[(size_t)c * 18 * nf];
                for (int f = 0; f < 6; ++f) {
                    const int dir = f / 2;
                    const double sgn = (f % 2) ? 1.0 : -1.0;
                    for (int q = 0; q < nf; ++q) {
                        int qq = q;
                        if ((f == 2 || f == 3) && p->transpose_y_face_normals)
                            qq = (q % np) * np + q / np;
                        const int node = face_node(N, f, qq);
                        const double v0 = Ja[(3 * dir + 0) * n + node], v1 = Ja[(3 * dir + 1) * n + node],
                                     v2 = Ja[(3 * dir + 2) * n + node];
                        const double nrm = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                        Nrm[(f * 3 + 0) * nf + q] = sgn * v0 / nrm;
                        Nrm[(f * 3 + 1) * nf + q] = sgn * v1 / nrm;
                        Nrm[(f * 3 + 2) * nf + q] = sgn * v2 / nrm;
                    }
                }
                // neighbours
                const int I[3] = {i, j, k};
                for (int f = 0; f < 6; ++f) {
                    const int dir = f / 2, side = f % 2;
                    int J3[3] = {i, j, k};
                    J3[dir] += side ? 1 : -1;
                    const bool outside = J3[dir] < 0 || J3[dir] >= ncell_dir[dir];
                    if (!outside) {
                        m->face_nbr[(size_t)c * 6 + f] = J3[0] + nx * (J3[1] + ny * J3[2]);
                    } else if (p->periodic[dir]) {
                        J3[dir] = side ? 0 : ncell_dir[dir] - 1;
                        m->face_nbr[(size_t)c * 6 + f] = J3[0] + nx * (J3[1] + ny * J3[2]);
                        m->face_bid[(size_t)c * 6 + f] = p->periodic_bid[dir];
                    } else {
                        m->face_bid[(size_t)c * 6 + f] = p->side_bid[f];
                    }
                    (void)I;
                }
            }
    m->q = 1;
    m->support = m->vertices;
    *out = m;
    return PLENS_OK;
}

// ---------------------------------------------------------------------------------------------
// curved meshes: manifolds and MappingQGeneric support points
// ---------------------------------------------------------------------------------------------
namespace {

struct P3 {
    double v[3] = {0, 0, 0};
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};
inline P3 mk(double a, double b, double c) { P3 p; p[0] = a; p[1] = b; p[2] = c; return p; }
inline double dot3(const P3& a, const P3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline P3 cross3(const P3& a, const P3& b) {
    return mk(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline P3 axpy(double a, const P3& x, const P3& y) { return mk(a * x[0] + y[0], a * x[1] + y[1], a * x[2] + y[2]); }
inline P3 sub3(const P3& a, const P3& b) { return mk(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline double norm3(const P3& a) { return std::sqrt(dot3(a, a)); }

// One manifold: CylindricalManifold (chart r, phi, z about a unit axis; phi measured from a unit
// normal of the axis) or SphericalManifold (center). deal.II's grid/manifold_lib.cc documents both.
struct Manifold {
    int kind = 0;          // 1 cylinder, 2 sphere
    P3 dir, pt, nrm, dxn;  // cylinder
    P3 c;                  // sphere

    P3 pull_back(const P3& x) const {  // cylinder chart (r, phi, z)
        const P3 d = sub3(x, pt);
        const double lam = dot3(d, dir);
        const P3 pd = axpy(-lam, dir, d);
        const double phi = std::atan2(dot3(pd, dxn), dot3(pd, nrm));
        return mk(norm3(pd), phi, lam);
    }
    P3 push_forward(const P3& ch) const {
        const double cr = std::cos(ch[1]) * ch[0], sr = std::sin(ch[1]) * ch[0];
        P3 x = axpy(ch[2], dir, pt);
        x = axpy(cr, nrm, x);
        return axpy(sr, dxn, x);
    }
    // new point from surrounding points and weights (rows of a weight table share the points)
    P3 new_point(const std::vector<P3>& pts, const double* w) const {
        const size_t np = pts.size();
        if (kind == 1) {
            // ChartManifold: weighted mean in the chart, phi with period 2 pi handled as the
            // periodic FlatManifold does (shift by -2 pi what lies more than pi above the minimum)
            std::vector<P3> ch(np);
            double pmin = 1e300;
            for (size_t i = 0; i < np; ++i) {
                ch[i] = pull_back(pts[i]);
                pmin = std::min(pmin, ch[i][1]);
            }
            P3 a;
            for (size_t i = 0; i < np; ++i) {
                P3 q = ch[i];
                if (q[1] - pmin > M_PI) q[1] -= 2.0 * M_PI;
                for (int d = 0; d < 3; ++d) a[d] += w[i] * q[d];
            }
            return push_forward(a);
        }
        // sphere: radius = weighted mean distance; direction = normalised weighted mean of the unit
        // directions when all directions lie close together (squared chord < 2e-2), else the
        // stationary point of sum_i w_i theta_i^2 on the unit sphere (weighted spherical mean,
        // Buss & Fillmore 2001), iterated from that guess
        std::vector<P3> dirs(np);
        std::vector<double> dist(np);
        double maxd = 0;
        for (size_t i = 0; i < np; ++i) {
            dirs[i] = sub3(pts[i], c);
            dist[i] = norm3(dirs[i]);
            for (int d = 0; d < 3; ++d) dirs[i][d] /= dist[i];
            for (size_t k = 0; k < i; ++k) {
                const P3 t = sub3(dirs[i], dirs[k]);
                maxd = std::max(maxd, dot3(t, t));
            }
        }
        double rho = 0;
        P3 cand;
        for (size_t i = 0; i < np; ++i) {
            rho += w[i] * dist[i];
            for (int d = 0; d < 3; ++d) cand[d] += w[i] * dirs[i][d];
        }
        const double cn = norm3(cand);
        if (cn == 0.0) return c;
        for (int d = 0; d < 3; ++d) cand[d] /= cn;
        if (maxd >= 2e-2) {
            for (int it = 0; it < 100; ++it) {
                P3 g;  // sum_i w_i log_cand(d_i)
                for (size_t i = 0; i < np; ++i) {
                    if (w[i] == 0.0) continue;
                    const double ct = std::max(-1.0, std::min(1.0, dot3(cand, dirs[i])));
                    const P3 t = axpy(-ct, cand, dirs[i]);
                    const double tn = norm3(t);
                    if (tn < 1e-300) continue;
                    const double th = std::atan2(tn, ct);
                    for (int d = 0; d < 3; ++d) g[d] += w[i] * th * t[d] / tn;
                }
                const double gn = norm3(g);
                if (gn < 1e-16) break;
                P3 nc;
                for (int d = 0; d < 3; ++d) nc[d] = std::cos(gn) * cand[d] + std::sin(gn) * g[d] / gn;
                const double nn = norm3(nc);
                for (int d = 0; d < 3; ++d) cand[d] = nc[d] / nn;
            }
        }
        return axpy(rho, cand, c);
    }
};

Manifold make_cylinder(const double* axis, const double* point) {
    Manifold m;
    m.kind = 1;
    const double an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    m.dir = mk(axis[0] / an, axis[1] / an, axis[2] / an);
    m.pt = mk(point[0], point[1], point[2]);
    // unit normal of the axis: cross with the coordinate axis least aligned with it
    int k = 0;
    for (int d = 1; d < 3; ++d)
        if (std::fabs(m.dir[d]) < std::fabs(m.dir[k])) k = d;
    P3 e;
    e[k] = 1;
    m.nrm = cross3(m.dir, e);
    const double nn = norm3(m.nrm);
    for (int d = 0; d < 3; ++d) m.nrm[d] /= nn;
    m.dxn = cross3(m.dir, m.nrm);
    return m;
}
Manifold make_sphere(const double* center) {
    Manifold m;
    m.kind = 2;
    m.c = mk(center[0], center[1], center[2]);
    return m;
}

// Lagrange basis values and derivatives of the nodes xn[0..q] at x
void lagrange_all(int q, const double* xn, double x, double* L, double* dL) {
    for (int j = 0; j <= q; ++j) {
        double v = 1, dv = 0;
        for (int k = 0; k <= q; ++k) {
            if (k == j) continue;
            const double den = xn[j] - xn[k];
            dv = dv * (x - xn[k]) / den + v / den;
            v *= (x - xn[k]) / den;
        }
        L[j] = v;
        if (dL) dL[j] = dv;
    }
}

// Support points of one cell (lexicographic over the GLL(q) lattice t[0..q]).
void cell_support_points(int q, const double* t, const P3 V[8], const Manifold* line_m[12],
                         const Manifold* face_m[6], const Manifold& cell_m, std::vector<P3>& S) {
    const int qp = q + 1;
    S.assign((size_t)qp * qp * qp, P3());
    auto at = [&](int i, int j, int k) -> P3& { return S[(size_t)i + qp * ((size_t)j + (size_t)qp * k)]; };
    for (int v = 0; v < 8; ++v) at((v & 1) * q, ((v >> 1) & 1) * q, ((v >> 2) & 1) * q) = V[v];
    if (q == 1) return;
    // lines: direction d, the other two lattice indices at 0 or q; deal.II line order is
    // irrelevant here, each line is identified by (d, a, b)
    int li = 0;
    for (int d = 0; d < 3; ++d) {
        const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a, ++li) {
                int I0[3], I1[3];
                I0[d] = 0; I1[d] = q;
                I0[d1] = I1[d1] = a * q;
                I0[d2] = I1[d2] = b * q;
                const std::vector<P3> ends = {at(I0[0], I0[1], I0[2]), at(I1[0], I1[1], I1[2])};
                for (int m = 1; m < q; ++m) {
                    const double w[2] = {1.0 - t[m], t[m]};
                    int I[3] = {I0[0], I0[1], I0[2]};
                    I[d] = m;
                    at(I[0], I[1], I[2]) = line_m[li]->new_point(ends, w);
                }
            }
    }
    // faces: normal direction d at index e*q; face coordinates (u, v) = the other directions in
    // increasing order; perimeter = 4 vertices then the 4 lines' interior points; weights of the
    // 2D transfinite interpolation
    for (int f = 0; f < 6; ++f) {
        const int d = f / 2, e = f % 2;
        const int du = d == 0 ? 1 : 0, dvv = d == 2 ? 1 : 2;
        auto idx = [&](int u, int v, int* I) { I[d] = e * q; I[du] = u; I[dvv] = v; };
        std::vector<P3> per;
        std::vector<std::pair<int, int>> pid;  // lattice (u, v) of each perimeter point
        for (int v = 0; v < 2; ++v)
            for (int u = 0; u < 2; ++u) pid.push_back({u * q, v * q});
        for (int u = 0; u < 2; ++u)  // lines u = 0, u = q (along v)
            for (int m = 1; m < q; ++m) pid.push_back({u * q, m});
        for (int v = 0; v < 2; ++v)  // lines v = 0, v = q (along u)
            for (int m = 1; m < q; ++m) pid.push_back({m, v * q});
        for (auto& pr : pid) {
            int I[3];
            idx(pr.first, pr.second, I);
            per.push_back(at(I[0], I[1], I[2]));
        }
        std::vector<double> w(per.size());
        for (int mv = 1; mv < q; ++mv)
            for (int mu = 1; mu < q; ++mu) {
                const double su = t[mu], sv = t[mv];
                for (size_t i = 0; i < pid.size(); ++i) {
                    const int pu = pid[i].first, pv = pid[i].second;
                    const bool uc = pu == 0 || pu == q, vc = pv == 0 || pv == q;
                    double wi = 0;
                    if (uc && vc) {  // vertex: -bilinear
                        wi = -((pu ? su : 1 - su) * (pv ? sv : 1 - sv));
                    } else if (uc && pv == mv) {  // line u = const at the point's v
                        wi = pu ? su : 1 - su;
                    } else if (vc && pu == mu) {  // line v = const at the point's u
                        wi = pv ? sv : 1 - sv;
                    }
                    w[i] = wi;
                }
                int I[3];
                idx(mu, mv, I);
                at(I[0], I[1], I[2]) = face_m[f]->new_point(per, w.data());
            }
    }
    // interior: 3D transfinite interpolation, faces + , lines - , vertices +
    std::vector<P3> per;
    std::vector<std::array<int, 3>> pid;
    for (int k = 0; k <= q; ++k)
        for (int j = 0; j <= q; ++j)
            for (int i = 0; i <= q; ++i) {
                const int nb = (i == 0 || i == q) + (j == 0 || j == q) + (k == 0 || k == q);
                if (nb == 0) continue;
                pid.push_back({i, j, k});
                per.push_back(at(i, j, k));
            }
    std::vector<double> w(per.size());
    for (int k = 1; k < q; ++k)
        for (int j = 1; j < q; ++j)
            for (int i = 1; i < q; ++i) {
                const int me[3] = {i, j, k};
                for (size_t p = 0; p < pid.size(); ++p) {
                    double wi = 1;
                    int nb = 0;
                    bool match = true;
                    for (int d = 0; d < 3; ++d) {
                        const int c = pid[p][d];
                        if (c == 0 || c == q) {
                            ++nb;
                            wi *= c ? t[me[d]] : 1 - t[me[d]];
                        } else if (c != me[d]) {
                            match = false;
                        }
                    }
                    w[p] = match ? (nb == 2 ? -wi : wi) : 0.0;
                }
                at(i, j, k) = cell_m.new_point(per, w.data());
            }
}

}  // namespace

extern "C" int plens_mesh_build_curved(int32_t N, int32_t q, int32_t nc, int32_t nv, const double* X,
                                       const int32_t* cv, const int32_t* nbr, const int32_t* bid,
                                       const plens_manifold_desc* md, int32_t transpose_y, plens_mesh** out) {
    if (!out || !X || !cv || !nbr || !bid || !md) return PLENS_E_INVALID;
    if (N < 1 || N > 5 || q < 1 || q > 8 || nc < 1 || nv < 8) return PLENS_E_INVALID;
    if (md->type != PLENS_MANIFOLD_CYLINDER && md->type != PLENS_MANIFOLD_NOSE_CYLINDER) return PLENS_E_INVALID;
    for (size_t i = 0; i < (size_t)nc * 8; ++i)
        if (cv[i] < 0 || cv[i] >= nv) return PLENS_E_INVALID;
    for (size_t i = 0; i < (size_t)nc * 6; ++i)
        if (nbr[i] < -1 || nbr[i] >= nc) return PLENS_E_INVALID;
    // interior faces: a reciprocal neighbour other than the cell itself whose matching face f^1
    // has the same vertices in the standard orientation (vertex v of face f is vertex v ^ (1 << d)
    // of the neighbour), which is what the face-node maps of both sides assume
    // (form_neighbor_face_matchings, dgsem/plens.cc:1574-1627, for the orientation every mesh of
    // the reference's structured generators has)
    for (int c = 0; c < nc; ++c)
        for (int f = 0; f < 6; ++f) {
            if (bid[(size_t)c * 6 + f] >= 0) continue;
            const int nb = nbr[(size_t)c * 6 + f], d = f >> 1;
            if (nb < 0 || nb == c || nbr[(size_t)nb * 6 + (f ^ 1)] != c) return PLENS_E_INVALID;
            for (int v = 0; v < 8; ++v)
                if (((v >> d) & 1) == (f & 1) && cv[8 * (size_t)c + v] != cv[8 * (size_t)nb + (v ^ (1 << d))])
                    return PLENS_E_INVALID;
        }
    const double an = std::sqrt(md->axis_dir[0] * md->axis_dir[0] + md->axis_dir[1] * md->axis_dir[1] +
                                md->axis_dir[2] * md->axis_dir[2]);
    if (!(an > 0)) return PLENS_E_INVALID;
    const Manifold cyl = make_cylinder(md->axis_dir, md->axis_point);
    const Manifold sph = make_sphere(md->nose_center);
    // manifold of every cell (NoseCylinder::set, manifolds/nose_cylinder.cc:45-78)
    std::vector<int> cman(nc, 1);
    if (md->type == PLENS_MANIFOLD_NOSE_CYLINDER) {
        double sep = 0;
        for (int d = 0; d < 3; ++d) sep += md->axis_dir[d] * (md->nose_center[d] - md->axis_point[d]);
        if (!(sep < 0)) return PLENS_E_INVALID;  // nose center on the spherical side
        for (int c = 0; c < nc; ++c) {
            double dp = 0;
            for (int d = 0; d < 3; ++d) {
                double ctr = 0;
                for (int v = 0; v < 8; ++v) ctr += X[3 * (size_t)cv[8 * (size_t)c + v] + d];
                dp += md->axis_dir[d] * (ctr / 8.0 - md->axis_point[d]);
            }
            if (dp == 0.0) return PLENS_E_INVALID;
            cman[c] = dp > 0 ? 2 : 1;
        }
    }
    // a line or face takes the manifold of the last (highest-index) cell that contains it
    std::map<std::array<int, 2>, int> line_owner;
    std::map<std::array<int, 4>, int> face_owner;
    auto line_key = [&](int c, int d, int a, int b) {
        const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        int v0 = 0, v1 = 0;
        int I0[3], I1[3];
        I0[d] = 0; I1[d] = 1;
        I0[d1] = I1[d1] = a;
        I0[d2] = I1[d2] = b;
        v0 = cv[8 * (size_t)c + I0[0] + 2 * I0[1] + 4 * I0[2]];
        v1 = cv[8 * (size_t)c + I1[0] + 2 * I1[1] + 4 * I1[2]];
        return std::array<int, 2>{std::min(v0, v1), std::max(v0, v1)};
    };
    auto face_key = [&](int c, int f) {
        const int d = f / 2, e = f % 2;
        std::array<int, 4> k;
        int n = 0;
        for (int v = 0; v < 8; ++v) {
            const int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
            if (b[d] == e) k[n++] = cv[8 * (size_t)c + v];
        }
        std::sort(k.begin(), k.end());
        return k;
    };
    for (int c = 0; c < nc; ++c) {
        for (int d = 0; d < 3; ++d)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) line_owner[line_key(c, d, a, b)] = c;
        for (int f = 0; f < 6; ++f) face_owner[face_key(c, f)] = c;
    }
    plens_mesh* m = new (std::nothrow) plens_mesh;
    if (!m) return PLENS_E_NOMEM;
    const int np = N + 1, n = np * np * np, nf = np * np, qp = q + 1, nsp = qp * qp * qp;
    m->N = N; m->n = n; m->nf = nf; m->n_cells = nc; m->q = q;
    m->nx = nc; m->ny = 1; m->nz = 1;
    m->x1d.resize(np); m->w1d.resize(np);
    plens_basis_1d(N, m->x1d.data(), m->w1d.data(), nullptr, nullptr);
    std::vector<double> tq(qp), wq(qp);
    if (q == 1) { tq[0] = 0; tq[1] = 1; }
    else gll_01(q, tq.data(), wq.data());
    m->vertices.resize((size_t)nc * 24);
    m->support.resize((size_t)nc * nsp * 3);
    m->nodes.resize((size_t)nc * 3 * n);
    m->Ja.resize((size_t)nc * 9 * n);
    m->detJ.resize((size_t)nc * n);
    m->normals.resize((size_t)nc * 18 * nf);
    m->hmin.resize(nc);
    m->face_nbr.assign(nbr, nbr + (size_t)nc * 6);
    m->face_bid.assign(bid, bid + (size_t)nc * 6);
    // 1D Lagrange values/derivatives of the GLL(q) support lattice at the GLL(N) nodes
    std::vector<double> Lq((size_t)np * qp), dLq((size_t)np * qp);
    for (int i = 0; i < np; ++i) lagrange_all(q, tq.data(), m->x1d[i], &Lq[(size_t)i * qp], &dLq[(size_t)i * qp]);
    std::vector<P3> S;
    for (int c = 0; c < nc; ++c) {
        P3 V[8];
        for (int v = 0; v < 8; ++v) {
            const double* src = X + 3 * (size_t)cv[8 * (size_t)c + v];
            V[v] = mk(src[0], src[1], src[2]);
            for (int d = 0; d < 3; ++d) m->vertices[(size_t)c * 24 + 3 * v + d] = src[d];
        }
        const Manifold* lm[12];
        int li = 0;
        for (int d = 0; d < 3; ++d)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a, ++li) lm[li] = cman[line_owner[line_key(c, d, a, b)]] == 2 ? &sph : &cyl;
        const Manifold* fm[6];
        for (int f = 0; f < 6; ++f) fm[f] = cman[face_owner[face_key(c, f)]] == 2 ? &sph : &cyl;
        cell_support_points(q, tq.data(), V, lm, fm, cman[c] == 2 ? sph : cyl, S);
        for (int s = 0; s < nsp; ++s)
            for (int d = 0; d < 3; ++d) m->support[((size_t)c * nsp + s) * 3 + d] = S[s][d];
        // minimum vertex distance over the 28 vertex pairs
        double d2min = 1e300;
        for (int a = 0; a < 8; ++a)
            for (int b = a + 1; b < 8; ++b) {
                const P3 t = sub3(V[a], V[b]);
                d2min = std::min(d2min, dot3(t, t));
            }
        m->hmin[c] = std::sqrt(d2min);
        // metric terms at the GLL(N) nodes from the degree-q map (MetricTerms::reinit,
        // dgsem/metric_terms.cc:82-102): J = sum_s S_s grad L_s, Ja^d = a_{d+1} x a_{d+2}
        double* Ja = &m->Ja[(size_t)c * 9 * n];
        double* dJ = &m->detJ[(size_t)c * n];
        double* XN = &m->nodes[(size_t)c * 3 * n];
        for (int nk = 0; nk < np; ++nk)
            for (int nj = 0; nj < np; ++nj)
                for (int ni = 0; ni < np; ++ni) {
                    const int node = ni + np * (nj + np * nk);
                    const double* Li = &Lq[(size_t)ni * qp];
                    const double* Lj = &Lq[(size_t)nj * qp];
                    const double* Lk = &Lq[(size_t)nk * qp];
                    const double* Di = &dLq[(size_t)ni * qp];
                    const double* Dj = &dLq[(size_t)nj * qp];
                    const double* Dk = &dLq[(size_t)nk * qp];
                    Vec3 a[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    double Xn[3] = {0, 0, 0};
                    for (int k = 0; k < qp; ++k)
                        for (int j = 0; j < qp; ++j)
                            for (int i = 0; i < qp; ++i) {
                                const P3& sp = S[(size_t)i + qp * ((size_t)j + (size_t)qp * k)];
                                const double l = Li[i] * Lj[j] * Lk[k];
                                const double g[3] = {Di[i] * Lj[j] * Lk[k], Li[i] * Dj[j] * Lk[k], Li[i] * Lj[j] * Dk[k]};
                                for (int d = 0; d < 3; ++d) {
                                    a[d].x += g[d] * sp[0];
                                    a[d].y += g[d] * sp[1];
                                    a[d].z += g[d] * sp[2];
                                }
                                for (int d = 0; d < 3; ++d) Xn[d] += l * sp[d];
                            }
                    for (int d = 0; d < 3; ++d) XN[d * n + node] = Xn[d];
                    const Vec3 cc[3] = {cross(a[1], a[2]), cross(a[2], a[0]), cross(a[0], a[1])};
                    for (int d = 0; d < 3; ++d) {
                        Ja[(3 * d + 0) * n + node] = cc[d].x;
                        Ja[(3 * d + 1) * n + node] = cc[d].y;
                        Ja[(3 * d + 2) * n + node] = cc[d].z;
                    }
                    dJ[node] = a[0].x * cc[0].x + a[0].y * cc[0].y + a[0].z * cc[0].z;
                }
        double* Nrm = &m->normals[(size_t)c * 18 * nf];
        for (int f = 0; f < 6; ++f) {
            const int dir = f / 2;
            const double sgn = (f % 2) ? 1.0 : -1.0;
            for (int qq0 = 0; qq0 < nf; ++qq0) {
                int qq = qq0;
                if ((f == 2 || f == 3) && transpose_y) qq = (qq0 % np) * np + qq0 / np;
                const int node = face_node(N, f, qq);
                const double v0 = Ja[(3 * dir + 0) * n + node], v1 = Ja[(3 * dir + 1) * n + node],
                             v2 = Ja[(3 * dir + 2) * n + node];
                const double nrm = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                Nrm[(f * 3 + 0) * nf + qq0] = sgn * v0 / nrm;
                Nrm[(f * 3 + 1) * nf + qq0] = sgn * v1 / nrm;
                Nrm[(f * 3 + 2) * nf + qq0] = sgn * v2 / nrm;
            }
        }
    }
    *out = m;
    return PLENS_OK;
}

extern "C" int plens_mesh_output_points(const plens_mesh* m, double* xyz) {
    if (!m || !xyz) return PLENS_E_INVALID;
    const int N = m->N, np = N + 1, n = m->n, q = m->q, qp = q + 1, nsp = qp * qp * qp;
    std::vector<double> tq(qp), wq(qp), L((size_t)np * qp);
    if (q == 1) { tq[0] = 0; tq[1] = 1; }
    else gll_01(q, tq.data(), wq.data());
    for (int e = 0; e < np; ++e) lagrange_all(q, tq.data(), (double)e / N, &L[(size_t)e * qp], nullptr);
    for (int c = 0; c < m->n_cells; ++c) {
        const double* S = &m->support[(size_t)c * nsp * 3];
        for (int ek = 0; ek < np; ++ek)
            for (int ej = 0; ej < np; ++ej)
                for (int ei = 0; ei < np; ++ei) {
                    double X[3] = {0, 0, 0};
                    if (q == 1) {  // trilinear map of the vertices
                        const double xi[3] = {(double)ei / N, (double)ej / N, (double)ek / N};
                        Vec3 a[3];
                        trilinear_jac(S, xi, a, X);
                    } else {
                        for (int k = 0; k < qp; ++k)
                            for (int j = 0; j < qp; ++j)
                                for (int i = 0; i < qp; ++i) {
                                    const double l = L[(size_t)ei * qp + i] * L[(size_t)ej * qp + j] * L[(size_t)ek * qp + k];
                                    const double* sp = S + 3 * ((size_t)i + qp * ((size_t)j + (size_t)qp * k));
                                    for (int d = 0; d < 3; ++d) X[d] += l * sp[d];
                                }
                    }
                    double* dst = xyz + 3 * ((size_t)c * n + ei + np * (ej + np * ek));
                    for (int d = 0; d < 3; ++d) dst[d] = X[d];
                }
    }
    return PLENS_OK;
}

extern "C" int32_t plens_mesh_mapping_degree(const plens_mesh* m) { return m ? m->q : 0; }
extern "C" const double* plens_mesh_support_points(const plens_mesh* m) { return m->support.data(); }

extern "C" int32_t plens_face_node(int32_t degree, int32_t f, int32_t q) {
    if (degree < 1 || f < 0 || f > 5 || q < 0 || q >= (degree + 1) * (degree + 1)) return -1;
    return face_node(degree, f, q);
}

extern "C" void plens_mesh_destroy(plens_mesh* m) { delete m; }

extern "C" int plens_mesh_fill_desc(const plens_mesh* m, plens_mesh_desc* d) {
    if (!m || !d) return PLENS_E_INVALID;
    d->n_cells = m->n_cells;
    d->degree = m->N;
    d->Ja = m->Ja.data();
    d->detJ = m->detJ.data();
    d->face_normals = m->normals.data();
    d->min_vertex_dist = m->hmin.data();
    d->face_nbr = m->face_nbr.data();
    d->face_bid = m->face_bid.data();
    d->vertices = m->vertices.data();
    return PLENS_OK;
}

extern "C" int32_t plens_mesh_n_cells(const plens_mesh* m) { return m ? m->n_cells : 0; }
extern "C" int32_t plens_mesh_degree(const plens_mesh* m) { return m ? m->N : 0; }
extern "C" const double* plens_mesh_vertices(const plens_mesh* m) { return m->vertices.data(); }
extern "C" const double* plens_mesh_node_coords(const plens_mesh* m) { return m->nodes.data(); }
extern "C" const double* plens_mesh_Ja(const plens_mesh* m) { return m->Ja.data(); }
extern "C" const double* plens_mesh_detJ(const plens_mesh* m) { return m->detJ.data(); }
extern "C" const double* plens_mesh_face_normals(const plens_mesh* m) { return m->normals.data(); }
extern "C" const int32_t* plens_mesh_face_nbr(const plens_mesh* m) { return m->face_nbr.data(); }
extern "C" const int32_t* plens_mesh_face_bid(const plens_mesh* m) { return m->face_bid.data(); }
extern "C" const double* plens_mesh_min_vertex_dist(const plens_mesh* m) { return m->hmin.data(); }

// ---------------------------------------------------------------------------------------------
// initial conditions
// ---------------------------------------------------------------------------------------------

extern "C" int plens_ic_uniform(const plens_mesh* m, double gamma, double rho, const double vel[3],
                                double p, double* Q) {
    if (!m || !Q) return PLENS_E_INVALID;
    const size_t ndof = (size_t)m->n_cells * m->n;
    double c[5];
    plens_prim_to_cons(gamma, rho, vel, p, c);
    for (int v = 0; v < 5; ++v)
        for (size_t i = 0; i < ndof; ++i) Q[v * ndof + i] = c[v];
    return PLENS_OK;
}

extern "C" int plens_ic_density_wave(const plens_mesh* m, double gamma, double rho0,
                                     const double vel[3], double p, double amp, double kmax,
                                     uint64_t seed, int32_t k_out[3], double* Q) {
    if (!m || !Q) return PLENS_E_INVALID;
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 7;
    const int kr = (int)std::floor(kmax);
    int kv[3] = {0, 0, 0};
    for (int it = 0; it < 1000; ++it) {
        for (int d = 0; d < 3; ++d) kv[d] = (int)(splitmix(s) % (2 * kr + 1)) - kr;
        const double kn = std::sqrt((double)(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]));
        if (kn > 0 && kn <= kmax) break;
    }
    if (k_out) { k_out[0] = kv[0]; k_out[1] = kv[1]; k_out[2] = kv[2]; }
    const int n = m->n;
    const size_t ndof = (size_t)m->n_cells * n;
    for (int c = 0; c < m->n_cells; ++c)
        for (int i = 0; i < n; ++i) {
            const double* X = &m->nodes[(size_t)c * 3 * n];
            const double kx = kv[0] * X[i] + kv[1] * X[n + i] + kv[2] * X[2 * n + i];
            const double rho = rho0 * (1.0 + amp * std::sin(2.0 * M_PI * kx));
            double cons[5];
            plens_prim_to_cons(gamma, rho, vel, p, cons);
            for (int v = 0; v < 5; ++v) Q[v * ndof + (size_t)c * n + i] = cons[v];
        }
    return PLENS_OK;
}

extern "C" int plens_ic_normal_shock(const plens_mesh* m, double gamma, double rho1, double p1,
                                     double mach, double x_shock, double noise, uint64_t seed,
                                     double* Q) {
    if (!m || !Q || mach <= 1) return PLENS_E_INVALID;
    const double g = gamma, M2 = mach * mach;
    const double a1 = std::sqrt(g * p1 / rho1), u1 = mach * a1;
    const double rr = (g + 1) * M2 / ((g - 1) * M2 + 2);          // rho2/rho1
    const double pr = 1.0 + 2.0 * g / (g + 1) * (M2 - 1.0);       // p2/p1
    const double rho2 = rho1 * rr, p2 = p1 * pr, u2 = u1 / rr;
    uint64_t s = seed * 0xD1B54A32D192ED03ull + 3;
    const int n = m->n;
    const size_t ndof = (size_t)m->n_cells * n;
    for (int c = 0; c < m->n_cells; ++c)
        for (int i = 0; i < n; ++i) {
            const double x = m->nodes[(size_t)c * 3 * n + i];
            double cons[5];
            if (x < x_shock) {
                const double v[3] = {u1, 0, 0};
                plens_prim_to_cons(g, rho1, v, p1, cons);
            } else {
                const double v[3] = {u2, 0, 0};
                const double rho = rho2 * (1.0 + noise * (2.0 * unit_uniform(s) - 1.0));
                plens_prim_to_cons(g, rho, v, p2, cons);
            }
            for (int v = 0; v < 5; ++v) Q[v * ndof + (size_t)c * n + i] = cons[v];
        }
    return PLENS_OK;
}

extern "C" double plens_gas_R(double molecular_weight) { return kR0 / molecular_weight; }

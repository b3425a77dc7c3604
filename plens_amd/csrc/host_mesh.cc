// Host setup of the DGSEM hot path: structured hex mesh, GLL basis, metric terms, face normals,
// neighbour tables and the initial conditions of the workloads. deal.II-free restatement of the
// PLENS setup members listed in include/plens_host.h.
#include "plens_host.h"

#include <algorithm>
#include <cmath>
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
    std::vector<double> vertices, nodes, Ja, detJ, normals, hmin;
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
    *out = m;
    return PLENS_OK;
}

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

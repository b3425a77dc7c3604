/*
 * plens_host.h -- deal.II-free host setup for the PLENS DGSEM hot path.
 *
 * This is the host layer that plays the role of PLENS's setup members: mesh generation
 * (read_mesh, dgsem/plens.cc:699-803), DoF distribution and neighbour face matching
 * (set_dof_handler / form_neighbor_face_matchings, dgsem/plens.cc:918-1009, 1574-1627), the
 * 1D GLL basis (calc_metric_terms, dgsem/plens.cc:1646-1682), metric terms
 * (MetricTerms::reinit, dgsem/metric_terms.cc:67-142) and FEFaceValues normals
 * (dgsem/plens.cc:1751-1756). It fills the plens_mesh_desc consumed by plens_gpu_create().
 *
 * Structured, standard-orientation hexahedral boxes with straight-sided (trilinear, MappingQ1)
 * cells, and curved meshes of given hexahedra on the reference's manifold descriptions with a
 * MappingQGeneric of degree q.
 *
 * Conventions (identical to the reference):
 *   - reference cell [0,1]^3, GLL nodes, lexicographic node id i + j(N+1) + k(N+1)^2;
 *   - faces 0..5 = -x,+x,-y,+y,-z,+z; face-node maps as FaceDoFInfo::form_maps
 *     (dgsem/face_dof_info.h:92-150);
 *   - cell id = i + nx*(j + ny*k); deal.II vertex order inside a cell (x fastest).
 * All functions return 0 on success and a negative PLENS_E_* code on error.
 */
#ifndef PLENS_HOST_H
#define PLENS_HOST_H

#include <stdint.h>
#include "plens_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plens_mesh plens_mesh; /* opaque host mesh with all precomputed arrays */

typedef struct {
    int32_t degree;            /* polynomial degree N (1..5) */
    int32_t nx, ny, nz;        /* cells per direction */
    double lo[3], hi[3];       /* box corners */
    double y_stretch_ratio;    /* geometric growth of dy away from y=lo[1]; 1 = uniform */
    double perturb_frac;      /* smooth vertex displacement amplitude / local spacing; 0 = none */
    uint64_t seed;             /* seed of the displacement wave numbers and phases */
    int32_t periodic[3];       /* 1 = periodic in that direction */
    int32_t side_bid[6];       /* bc-table index of the 6 domain sides (used when not periodic) */
    int32_t periodic_bid[3];   /* bc-table index of the PERIODIC entry of each periodic direction */
    int32_t transpose_y_face_normals; /* 1 = deal.II face quadrature order on faces 2/3 */
} plens_box_params;

/* Build a box mesh and all its arrays. */
int plens_mesh_build_box(const plens_box_params* prm, plens_mesh** out);

/* Manifold descriptions of curved meshes ("curved subtype", dgsem/plens.cc:104-130, 737-797;
 * manifolds/cylinder.cc, manifolds/nose_cylinder.cc). */
enum {
    PLENS_MANIFOLD_CYLINDER = 1,       /* CylindricalManifold(axis_dir, axis_point) on everything */
    PLENS_MANIFOLD_NOSE_CYLINDER = 2   /* SphericalManifold(nose_center) on cells whose vertex mean
                                        * lies on the +axis_dir side of the plane through
                                        * axis_point (the separation point), CylindricalManifold(
                                        * axis_dir, axis_point) on the others; the nose center must
                                        * lie on the -axis_dir side, no cell center on the plane */
};
typedef struct {
    int32_t type;             /* PLENS_MANIFOLD_* */
    double axis_dir[3];       /* "axis direction" (any length) */
    double axis_point[3];     /* cylinder: "axis point"; nose cylinder: "separation point" */
    double nose_center[3];    /* nose cylinder: "nose center" */
} plens_manifold_desc;

/* Build a curved mesh: hexahedra in standard orientation given by their 8 global vertex ids
 * (deal.II vertex order), neighbours and boundary indices as for the box, a manifold description
 * and the degree q >= 1 of the MappingQGeneric (read_mesh, dgsem/plens.cc:699-803). The mapping's
 * support points (GLL(q) lattice of each cell) are placed as MappingQGeneric does: vertices, line
 * points from the line's manifold with weights (1-t, t), face and cell interior points by linear
 * transfinite (Gordon-Hall) interpolation of the perimeter through the face's / cell's manifold;
 * a line or face shared by cells of different manifolds takes the manifold of the cell with the
 * highest index (set_all_manifold_ids in cell order). Metric terms, node coordinates and face
 * normals follow from the degree-q map at the GLL(N) nodes exactly as for the box. */
int plens_mesh_build_curved(int32_t degree, int32_t mapping_degree, int32_t n_cells, int32_t n_vertices,
                            const double* vertex_coords, const int32_t* cell_vertices,
                            const int32_t* face_nbr, const int32_t* face_bid,
                            const plens_manifold_desc* manifold, int32_t transpose_y_face_normals,
                            plens_mesh** out);
/* Mapping degree q (1 for boxes) and the support points [n_cells][(q+1)^3][3], lexicographic over
 * the GLL(q) lattice of the reference cell (for q = 1: the 8 vertices). */
int32_t plens_mesh_mapping_degree(const plens_mesh* m);
/* Evaluate an expression of the varying inflow BC (the same compiler and evaluator the device
 * uses): n_components expressions separated by ';' at (x, y, z, t) into values; -1 on a syntax
 * error or another number of components. */
int plens_expr_eval(const char* expr, int32_t n_components, double x, double y, double z, double t,
                    double* values);

/* Coordinates of the equispaced (N+1)^3 output lattice of every cell through its map:
 * xyz[n_cells][(N+1)^3][3], lattice points lexicographic (x fastest). */
int plens_mesh_output_points(const plens_mesh* m, double* xyz);

/* Data output files of PLENS::write() (dgsem/plens.cc:4170-4322) from the fields of
 * plens_gpu_output_fields (nodal [15][n_cells][(N+1)^3], cell [3][n_cells]) for one process:
 *   <dir>/<base>_<counter, 6 digits>.0.vtu  one VTK Lagrange hexahedron (type 72, order N, VTK
 *       XML version 2.2 node order) per cell on the equispaced lattice (write_higher_order_cells,
 *       build_patches(mapping, N)); point data in the reference's order: txx txy txz tyy tyz tzz
 *       qx qy qz rho p T u v w subdomain alpha alpha_d loc_dt global_active_cell_index (the cell
 *       values constant over the cell's points); inline base64 Float64/Int64/UInt8 arrays with
 *       UInt64 byte-count headers;
 *   <dir>/<base>_<counter, 6 digits>.pvtu  the parallel record of that piece;
 *   <dir>/<base>.pvd  rewritten with one data set per output 0..n_records-1 at record_times[i]
 *       (the times_and_names of write_pvd_record);
 *   <dir>/<base>.times  "counter time wall_time" appended.
 * The directory must exist. Returns 0 or PLENS_E_INVALID (bad argument, file not writable). */
int plens_output_write(const plens_mesh* m, const char* directory, const char* base, int32_t counter,
                       double time, double wall_time, const double* nodal, const double* cell,
                       const double* record_times, int32_t n_records);
const double* plens_mesh_support_points(const plens_mesh* m);
void plens_mesh_destroy(plens_mesh* m);

/* Fill a descriptor that points into the mesh's arrays (valid while the mesh lives). */
int plens_mesh_fill_desc(const plens_mesh* m, plens_mesh_desc* desc);

/* Sizes and raw arrays (owned by the mesh). */
int32_t plens_mesh_n_cells(const plens_mesh* m);
int32_t plens_mesh_degree(const plens_mesh* m);
const double* plens_mesh_vertices(const plens_mesh* m);     /* [n_cells][8][3], deal.II order */
const double* plens_mesh_node_coords(const plens_mesh* m);  /* [n_cells][3][n] */
const double* plens_mesh_Ja(const plens_mesh* m);           /* [n_cells][9][n] */
const double* plens_mesh_detJ(const plens_mesh* m);         /* [n_cells][n] */
const double* plens_mesh_face_normals(const plens_mesh* m); /* [n_cells][6][3][n_f] */
const int32_t* plens_mesh_face_nbr(const plens_mesh* m);    /* [n_cells][6] */
const int32_t* plens_mesh_face_bid(const plens_mesh* m);    /* [n_cells][6] */
const double* plens_mesh_min_vertex_dist(const plens_mesh* m); /* [n_cells] */

/* 1D GLL basis on [0,1] of degree N: nodes x[N+1], weights w[N+1], D[(N+1)^2] (row-major,
 * D[i*(N+1)+j] = l_j'(x_i)) and the 1D nodal->Legendre matrix C1[(N+1)^2]
 * (C1[i*(N+1)+j] = int_0^1 psi_i phi_j, psi_i orthonormal Legendre). */
int plens_basis_1d(int32_t degree, double* x, double* w, double* D, double* C1);

/* Cell-local node of face-local dof q on face f (FaceDoFInfo::form_maps, dgsem/face_dof_info.h:92-150). */
int32_t plens_face_node(int32_t degree, int32_t face, int32_t q);

/* Primitive -> conservative with the reference's formula (modelling/navier_stokes.h:365-378). */
void plens_prim_to_cons(double gamma, double rho, const double vel[3], double p, double cons[5]);

/* Initial conditions written as Q[5][n_dof] (SoA, cell-major DoF order). */
/* rho = rho0*(1 + amp*sin(2 pi k.x)) with seeded integer k, 0<|k|<=kmax; uniform vel, p. */
int plens_ic_density_wave(const plens_mesh* m, double gamma, double rho0, const double vel[3],
                          double p, double amp, double kmax, uint64_t seed, int32_t k_out[3],
                          double* Q);
/* Stationary normal shock at x = x_shock from Rankine-Hugoniot with pre-shock (rho1, p1, Mach)
 * flowing in +x, plus seeded uniform nodal density noise of relative size `noise` downstream. */
int plens_ic_normal_shock(const plens_mesh* m, double gamma, double rho1, double p1, double mach,
                          double x_shock, double noise, uint64_t seed, double* Q);
/* Uniform state. */
int plens_ic_uniform(const plens_mesh* m, double gamma, double rho, const double vel[3], double p,
                     double* Q);

#ifdef __cplusplus
}
#endif
#endif

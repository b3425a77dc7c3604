/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY. CPU restatement of the PLENS explicit DGSEM right-hand
 * side and RK4(3)5[3R+]N update, used as the parity checker of the GPU path and as the CPU
 * baseline timed beside it. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product path never does.
 *
 * It follows the reference's loop nests and operation order in FP64:
 *   dgsem/plens.cc 1717-1913 (calc_surf_flux), 2174-2285 (calc_cell_cons_grad, CVAR_GRAD build),
 *   2531-2546 (assert_positivity), 2593-2748 (calc_aux_vars), 2858-2962 (calc_blender),
 *   3153-3583 (ho inviscid, ho diffusive, lo inviscid residuals), 3688-3779 (calc_rhs),
 *   3810-3940 (calc_time_step, multiply_time_step_to_rhs), 4334-4403 (update_rk4),
 *   4484-4518 (update);
 *   modelling/navier_stokes.{h,cc} (thermo, fluxes, log mean), modelling/ausm_functions.h;
 *   dgsem/blender_calculator.cc 24-87, 123-187, 277-303; dgsem/change_of_basis_matrix.h 74-98;
 *   dgsem/metric_terms.cc 67-142; dgsem/face_dof_info.h 92-150; dgsem/cell_dof_info.h 60-88;
 *   dgsem/rk4_stage5_register3.cc 12-36; IBCs/{periodic,uniform_inflow,outflow,uniform_temp_wall}.cc.
 * Its metric terms, basis and face normals are computed here from the cell vertices, independent
 * of the product's host layer.
 */
#ifndef PLENS_ORACLE_H
#define PLENS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

typedef struct {
    double gamma, molecular_weight, prandtl, mu0, T0, S;
    int32_t inv_surf_flux;  /* 0 HLLC, 1 Rusanov, 2 AUSM+-up, 3 Chandrashekhar, 4 Rusanov-HLLC,
                               5 Rusanov-AUSM+-up, 6 modified Steger-Warming, 7 Kennedy-Gruber,
                               8 Rusanov-Kennedy-Gruber, 9 Ismail-Roe */
    double threshold_factor, threshold_exp_factor, sharpness_factor, alpha_min, alpha_max,
        wall_blender_limit;
    int32_t gradient_variables;  /* 0 conservative (CVAR_GRAD build), 1 entropy (default build) */
    int32_t inv_vol_flux;        /* 0 Chandrashekhar, 1 Kennedy-Gruber, 2 Ismail-Roe */
} oracle_physics;

/* type as PLENS_BC_*; the expressions of a varying inflow (9) as plens_bc, copied at create */
typedef struct { int32_t type; double p, T, vel[3]; const char *p_expr, *T_expr, *vel_expr; } oracle_bc;

/* vertices [n_cells][8][3] (deal.II order), face_nbr/face_bid [n_cells][6] as plens_mesh_desc */
int oracle_create(int32_t n_cells, int32_t degree, const double* vertices, const int32_t* face_nbr,
                  const int32_t* face_bid, const oracle_physics* phys, const oracle_bc* bcs,
                  int32_t n_bcs, int32_t transpose_y_face_normals, oracle** out);
/* The same for a MappingQGeneric of degree map_degree given by its support points
 * [n_cells][(q+1)^3][3] (lexicographic over the GLL(q) lattice; the corners are the vertices). */
int oracle_create_mapped(int32_t n_cells, int32_t degree, int32_t map_degree, const double* support,
                         const int32_t* face_nbr, const int32_t* face_bid, const oracle_physics* phys,
                         const oracle_bc* bcs, int32_t n_bcs, int32_t transpose_y, oracle** out);
void oracle_destroy(oracle* o);
/* BC time (cur_time) of calc_rhs; steps set it from their time argument */
int oracle_set_time(oracle* o, double t);
/* Checker hook (no counterpart in the reference): with both arrays given ([n_cells] each), later RHS
 * evaluations use these blender values alpha and alpha_d in place of calc_blender's, which separates
 * the arithmetic of the DG operator from the conditioning of the blender; NULL restores calc_blender. */
int oracle_set_alpha(oracle* o, const double* alpha, const double* alpha_d);
/* Checker hook: with raw given ([n_cells]), calc_blender takes these indicator values in place of
 * get_blender's and applies its post-steps to them (wall override, diffusion, alpha_d); NULL clears. */
int oracle_set_raw_alpha(oracle* o, const double* raw);
/* Checker hook: the ghost state and auxiliary variables that boundary face `face` of `cell` gives at
 * face node q and stage 1..3 (the BC getters of IBCs/), for the inner state in[5], inner avars ain[9]
 * and outward unit normal nrm[3]; -1 for an interior face or bad indices. */
int oracle_bc_ghost(const oracle* o, int32_t stage, int32_t cell, int32_t face, int32_t q, const double* in,
                    const double* ain, const double* nrm, double* g, double* ag);
/* Checker hook: plant mistake `id` (process-wide; 0 = none) in the time step, the inflow/outflow/
 * isothermal-wall ghosts or the blender post-steps, so that tests can show their pins catch it:
 * 1 abs in the advection maximum, 2 N^2 for N^1.5, 3 no Courant cap, 4 J^-1 v for J^-T v;
 * 11 inflow stage-1 ghost = prescribed state, 12 outflow correction sign, 13 outflow correction
 * also when inviscid, 14 wall density from the inner temperature, 15 wall stage 2 reverses the
 * absolute momentum; 21 diffusion from diffused neighbours, 22 diffusion across periodic faces,
 * 23 wall override after diffusion, 24 alpha_d not capped at 1. */
int oracle_set_mutation(int32_t id);
/* expression-valued "Courant number" (origin, cur_time), "wall blender limit" (origin, cur_time) and
 * "viscous blending region" (cell center = vertex mean, cur_time; alpha_d = 0 where not > 0), each
 * nullable (NULL: the numeric value / every cell) */
int oracle_set_expressions(oracle* o, const char* courant, const char* wall_blender_limit,
                           const char* viscous_blending_region);
/* number of (cell, RHS evaluation) pairs with a blender value alpha > 0 since creation */
int oracle_blended_cell_stages(const oracle* o, int64_t* count);
/* location of node `node` of cell c (through the cell's map) */
int oracle_node_location(const oracle* o, int32_t c, int32_t node, double* xyz);
/* value of one expression of the varying inflow BC; NaN on a syntax error */
double oracle_eval_expression(const char* expr, double x, double y, double z, double t);
/* fields of PLENS::write() for the current state: nodal [15][n_cells][(N+1)^3] on the equispaced
 * lattice (txx..qz, rho, p, T, u, v, w), cell [3][n_cells] (alpha, alpha_d, loc_dt) */
int oracle_output_fields(oracle* o, double courant, double t, double* nodal, double* cell);
int oracle_set_threads(int32_t n);
int oracle_set_state(oracle* o, const double* Q);
int oracle_get_state(const oracle* o, double* Q);
/* calc_rhs on the current state; outputs nullable */
int oracle_calc_rhs(oracle* o, double* rhs, double* alpha, double* alpha_d, double* avars);
int oracle_calc_dt(oracle* o, double courant, double* dt);
/* one update() with RK4; *t += dt */
int oracle_rk4_step(oracle* o, double courant, double* t);
/* one update() with a prescribed dt (time-order checks of the RK registers) */
int oracle_rk4_step_fixed_dt(oracle* o, double dt, double* t);
int oracle_run(oracle* o, int32_t n_steps, double courant, double* t);
int oracle_last_failure(const oracle* o, int64_t* dof, int32_t* kind);
/* local time stepping parameters ("time integration": request local stepping, local stepping start
 * time, limit local time step, local step neighbor/global factor) */
int oracle_set_local_stepping(oracle* o, int32_t request, double start_time, int32_t limit,
                              double neighbor_factor, double global_factor);
/* per-cell time steps of the last step (loc_time_steps) */
int oracle_loc_dt(const oracle* o, double* dt);
/* steady-state error after at least one step (calc_ss_error) */
int oracle_ss_error(const oracle* o, double* err);
/* section timers (seconds, accumulated) named as plens_test::time(): [0] assert positivity,
 * [1] aux vars, [2] blender, [3] surface fluxes 2&3, [4] ho residual, [5] lo residual,
 * [6] time step, [7] updating cvars */
int oracle_section_times(const oracle* o, double* t, int32_t n);
/* flop counters of the counting build liboracle_flops.so (oracle/flopcount.h): per bucket
 * [0] positivity, [1] aux vars, [2] blender, [3] surface fluxes, [4] ho inviscid residual, [5] ho
 * viscous residual, [6] lo residual, [7] residual blend, [8] time step, [9] RK update, and the
 * nested [10] two-point volume fluxes, [11] dense change of basis, [12] stage-2 numerical fluxes,
 * [13] viscous fluxes of the ho viscous volume term, and [14], [15] the surface parts of the ho
 * inviscid and viscous residuals; calls[i] counts the entries into the nested buckets. Returns the number of buckets, or -1 in liboracle.so. reset != 0 zeroes the counters. */
int oracle_flop_counts(uint64_t* flops, uint64_t* calls, int32_t n, int32_t reset);

/* metric terms of the oracle for one cell: Ja [9][n], detJ [n], face normals [6][3][n_f] */
int oracle_cell_metrics(const oracle* o, int32_t cell, double* Ja, double* detJ, double* normals);
/* subcell normals of one cell: [3 dir][N+2 id][N+1 id1][N+1 id2][3] */
int oracle_cell_subcell_normals(const oracle* o, int32_t cell, double* sn);

/* ---- unit-level restatements (for known-answer and invariant tests) ---- */
int oracle_basis(int32_t degree, double* x, double* w, double* D);
/* dense (N+1)^3 x (N+1)^3 change-of-basis matrix, ChangeOfBasisMatrix ctor */
int oracle_cob_matrix(int32_t degree, double* C);
double oracle_log_avg(double xl, double xr);
/* inviscid numerical flux through unit normal n (get_inv_surf_flux with rotation) */
void oracle_inv_surf_flux(const oracle_physics* p, int32_t scheme, double blender, const double* ul,
                          const double* ur, const double* n, double* f);
void oracle_inv_flux(const oracle_physics* p, const double* u, const double* dir, double* f);
void oracle_vol_flux(const oracle_physics* p, const double* u1, const double* u2, const double* dir,
                     double* f);
void oracle_dif_flux(const double* u, const double* av, const double* dir, double* f);
/* blender of one cell from nodal values of the blender variable (get_blender) */
double oracle_blender_value(const oracle_physics* p, int32_t degree, const double* var);
void oracle_rk4_coeffs(double* a_outer5, double* a_inner4, double* b5);

#ifdef __cplusplus
}
#endif
#endif

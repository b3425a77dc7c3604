/*
 * plens_gpu.h -- C ABI of the MI355X (gfx950) DGSEM right-hand side and RK4(3)5[3R+]N update.
 *
 * Drop-in for the hot path of PLENS (vachan-potluri/plens):
 *   PLENS::calc_rhs()            dgsem/plens.cc:3688-3779   -> plens_gpu_calc_rhs()
 *   PLENS::calc_time_step()      dgsem/plens.cc:3810-3912   -> plens_gpu_calc_dt()   (global stepping)
 *   PLENS::update()+update_rk4() dgsem/plens.cc:4484-4518, 4334-4403 -> plens_gpu_rk4_step()
 *                                (or, with the step reduced by the caller, plens_gpu_rk4_step_dt())
 *   PLENS::run() time loop       dgsem/plens.cc:1547        -> plens_gpu_run()
 *   NavierStokes::assert_positivity's std::exit (dgsem/plens.cc:2540-2544)
 *                                                            -> PLENS_E_POSITIVITY + plens_gpu_last_failure()
 * The reference builds with -DCVAR_GRAD=ON (conservative-variable BR1 gradients,
 * calc_cell_cons_grad dgsem/plens.cc:2174-2285), Sutherland viscosity and the Hennemann
 * troubled-cell indicator (dgsem/blender_calculator.cc:123-187, 277-303).
 *
 * Arrays are plain host pointers; they are copied at create/set time. No C++ types or exceptions
 * cross this boundary. Every function returns 0 (PLENS_OK) or a negative PLENS_E_* code.
 * A handle owns its device memory and one HIP stream; calls on one handle must be serialised by
 * the caller, distinct handles may be used concurrently.
 */
#ifndef PLENS_GPU_H
#define PLENS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PLENS_OK = 0,
    PLENS_E_INVALID = -1,      /* bad argument / unsupported combination */
    PLENS_E_HIP = -2,          /* HIP runtime error */
    PLENS_E_POSITIVITY = -3,   /* rho<=0 or e<=0 at some DoF (replaces the reference's std::exit) */
    PLENS_E_NOMEM = -4,        /* device allocation failed */
    PLENS_E_NODEVICE = -5      /* no GPU visible */
};

/* Inviscid surface flux schemes ("inviscid surface flux scheme", dgsem/plens.cc:237-246; selection
 * dgsem/plens.cc:831-845, modelling/navier_stokes.cc:140-197). The prm value "SLAU2" never reaches
 * slau2_xflux in the reference (plens.cc:845 compares instead of assigning) and has no value here. */
enum {
    PLENS_FLUX_HLLC = 0,           /* "HLLC": hllc_xflux, modelling/navier_stokes.cc:936-1033 */
    PLENS_FLUX_RUSANOV = 1,        /* "Rusanov": rusanov_xflux, modelling/navier_stokes.cc:1044-1059 */
    PLENS_FLUX_AUSM_PLUS_UP = 2,   /* "AUSM+-up": ausm_plus_up_xflux, modelling/navier_stokes.cc:1070-1121 */
    PLENS_FLUX_CHANDRASHEKHAR = 3, /* "Chandrashekhar": chandrashekhar_xflux, navier_stokes.cc:1250-1318 */
    PLENS_FLUX_RUSANOV_HLLC = 4,   /* "Rusanov-HLLC": rusanov_hllc_blend_xflux, navier_stokes.cc:1132-1141 */
    PLENS_FLUX_RUSANOV_AUSM_PLUS_UP = 5, /* "Rusanov-AUSM+-up": navier_stokes.cc:1150-1159 */
    PLENS_FLUX_MODIFIED_SW = 6,    /* "Modified SW": modified_sw_xflux, navier_stokes.cc:1174-1236 */
    PLENS_FLUX_KG = 7,             /* "KG": kennedy_gruber_xflux, navier_stokes.cc:1325-1357 */
    PLENS_FLUX_RUSANOV_KG = 8,     /* "Rusanov-KG": rusanov_kennedy_gruber_blend_xflux, .cc:1364-1374 */
    PLENS_FLUX_IR = 9              /* "IR": ismail_roe_xflux, navier_stokes.cc:1381-1446 */
};

/* Inviscid volume (two-point) flux of the split form ("inviscid volume flux scheme",
 * dgsem/plens.cc:247-252, 847-850; modelling/navier_stokes.cc:203-228). */
enum {
    PLENS_VOL_FLUX_CHANDRASHEKHAR = 0, /* chandrashekhar_vol_flux, navier_stokes.cc:1510-1542 */
    PLENS_VOL_FLUX_KG = 1,             /* kennedy_gruber_vol_flux, navier_stokes.cc:1549-1575 */
    PLENS_VOL_FLUX_IR = 2              /* ismail_roe_vol_flux, navier_stokes.cc:1590-1622 */
};

/* Boundary condition types (IBCs/). */
enum {
    PLENS_BC_PERIODIC = 0,          /* IBCs/periodic.cc:121-165 */
    PLENS_BC_UNIFORM_INFLOW = 1,    /* IBCs/uniform_inflow.cc:23-77, state from (p, T, vel) */
    PLENS_BC_OUTFLOW = 2,           /* IBCs/outflow.cc:18-75, prescribed p */
    PLENS_BC_UNIFORM_TEMP_WALL = 3, /* IBCs/uniform_temp_wall.cc:17-104, (T, vel) */
    PLENS_BC_FREE = 4,              /* IBCs/free.cc, inner state and avars */
    PLENS_BC_SYMMETRY = 5,          /* IBCs/symmetry.cc */
    PLENS_BC_EMPTY = 6,             /* IBCs/empty.cc */
    PLENS_BC_INSULATED_WALL = 7,    /* IBCs/insulated_wall.cc, (vel); the stage-3 heat flux has its
                                     * normal component reversed as documented there (the
                                     * reference's copy loop at line 99 indexes past the 3-vector) */
    PLENS_BC_ZPG_INFLOW = 8,        /* IBCs/zpg_inflow.cc, (vel) with inner density and pressure */
    PLENS_BC_VARYING_INFLOW = 9     /* IBCs/varying_inflow.cc: p_expr, T_expr, vel_expr (deal.II
                                     * FunctionParser syntax in x, y, z, t; velocity components
                                     * separated by ';'), evaluated at the face DoF's location and
                                     * the solution time at the start of the step; needs
                                     * node_coords or vertices in the mesh; not in batches */
};

/* Arithmetic of the flux evaluations. FP64 registers/residual accumulation in both. */
enum {
    PLENS_PREC_FP64 = 0,
    PLENS_PREC_FP32_BULK = 1
};

/* Variables of the BR1 gradient (lifting) step of calc_aux_vars (dgsem/plens.cc:2593-2748):
 * conservative variables in the CVAR_GRAD build (calc_cell_cons_grad, 2174-2285), entropy
 * variables otherwise (calc_cell_evar_grad, 2294-2405; the reference's default, main.cc:76-81). */
enum {
    PLENS_GRAD_CONSERVATIVE = 0,
    PLENS_GRAD_ENTROPY = 1
};

typedef struct plens_gpu plens_gpu;

/* Mesh and metric terms of one partition (all owned cells), DoF order = cell-major,
 * lexicographic nodes inside a cell; n = (N+1)^3, n_f = (N+1)^2. */
typedef struct {
    int32_t n_cells;
    int32_t degree;                 /* N */
    const double* Ja;               /* [n_cells][9][n]: MetricTerms::JxContra_vecs, comp 3*dir+xyz (nullable, see vertices) */
    const double* detJ;             /* [n_cells][n]:    MetricTerms::detJ (nullable, see vertices) */
    const double* face_normals;     /* [n_cells][6][3][n_f]: FEFaceValues::normal_vector(face_dof) */
    const double* min_vertex_dist;  /* [n_cells]: cell->minimum_vertex_distance() */
    const int32_t* face_nbr;        /* [n_cells][6]: neighbour (interior) or periodic partner cell, else -1 */
    const int32_t* face_bid;        /* [n_cells][6]: -1 for interior faces, else index into the bc table */
    const double* vertices;         /* [n_cells][8][3] in deal.II vertex order, nullable. Given, the cells
                                     * are trilinear (MappingQ1): the kernels derive Ja and detJ at
                                     * every node from each cell's map instead of reading stored arrays
                                     * (Ja and detJ may then be NULL; if given, create() checks that the
                                     * map reproduces them to 1e-10 relative, else PLENS_E_INVALID). */
    const double* node_coords;      /* [n_cells][3][n]: DoF locations (dof_locations), nullable; read
                                     * only for the face nodes of varying-inflow faces (vertices, if
                                     * given, are used in their absence) */
} plens_mesh_desc;

/* Gas model ("Navier-Stokes" subsection, custom gas, dgsem/plens.cc:185-223) and blender
 * parameters ("blender parameters", dgsem/plens.cc:420-515). */
typedef struct {
    double gamma, molecular_weight, prandtl, mu0, T0, S;  /* R = 8.314462618/molecular_weight */
    int32_t inv_surf_flux;         /* PLENS_FLUX_* */
    int32_t precision;             /* PLENS_PREC_* */
    double threshold_factor;       /* default 0.5 */
    double threshold_exp_factor;   /* default 1.8 */
    double sharpness_factor;       /* default 9.21024 */
    double alpha_min;              /* default 1e-3 */
    double alpha_max;              /* default 0.5 */
    double wall_blender_limit;     /* default 0.5 */
    int32_t gradient_variables;    /* PLENS_GRAD_*: variables of the BR1 lifting */
    int32_t inv_vol_flux;          /* PLENS_VOL_FLUX_* */
} plens_physics;

typedef struct {
    int32_t type;      /* PLENS_BC_* */
    double p, T;       /* prescribed pressure / temperature where the type uses them */
    double vel[3];     /* prescribed (inflow) or wall velocity */
    const char* p_expr;    /* varying inflow: "prescribed p", "prescribed T", "prescribed velocity" */
    const char* T_expr;    /* (the prm strings; NULL for the other types) */
    const char* vel_expr;
} plens_bc;

/* Create a solver handle on the current HIP device. */
int plens_gpu_create(const plens_mesh_desc* mesh, const plens_physics* phys, const plens_bc* bcs,
                     int32_t n_bcs, plens_gpu** out);
void plens_gpu_destroy(plens_gpu* h);

/* State Q as [5][n_dof] (rho, rho u, rho v, rho w, rho E), the reference's g_cvars. */
int plens_gpu_set_state(plens_gpu* h, const double* Q);
int plens_gpu_get_state(plens_gpu* h, double* Q);

/* == PLENS::calc_rhs(): rhs_out [5][n_dof] (nullable), alpha_out [n_cells] (nullable, the
 * post-diffusion blender gcrk_alpha), alpha_d_out [n_cells] (nullable). */
int plens_gpu_calc_rhs(plens_gpu* h, double* rhs_out, double* alpha_out, double* alpha_d_out);

/* == PLENS::calc_time_step() (dgsem/plens.cc:3810-3912) with global stepping; needs mu, so it
 * evaluates mu from the state. dt_global = minimum over the cells of the per-node stable step
 * Co / (N^1.5 (max_d (J^-T v)_d + a/h + N^2 mu / (h^2 rho))), Co >= 1 capped to 0.9; dt_cell
 * [n_cells] (nullable) = each cell's own minimum (this_cell_step, 3844-3866), the values a caller
 * that owns several partitions reduces itself (3872-3875: MPI_Reduce + MPI_Bcast of the minimum). */
int plens_gpu_calc_dt(plens_gpu* h, double courant, double* dt_global, double* dt_cell);

/* One time step == PLENS::update() with RK order 4: dt from the stage-1 state (courant capped at
 * 0.9 when >= 1), 5 stages of calc_rhs + 3-register update. *t_inout += dt. */
int plens_gpu_rk4_step(plens_gpu* h, double courant, double* t_inout);

/* One time step with a step the caller supplies: update() + update_rk4() (4334-4403, 4484-4518)
 * with time_step = dt_global, e.g. the minimum of plens_gpu_calc_dt's dt_global over every
 * partition (the reference's MPI_Reduce/MPI_Bcast at 3872-3875), or that step clipped to
 * end_time - cur_time (4340). The five stages and the register update are those of
 * plens_gpu_rk4_step, so calc_dt followed by this call gives the same state bit for bit. t is the
 * solution time at the start of the step (the BCs' cur_time). Global stepping only: PLENS_E_INVALID
 * when local time stepping is requested, or for dt_global <= 0 or not finite. */
int plens_gpu_rk4_step_dt(plens_gpu* h, double t, double dt_global);

/* n_steps time steps replayed from a captured HIP graph; *t_inout += sum of dt. The state stays on
 * the device; positivity failures are reported at the end. */
int plens_gpu_run(plens_gpu* h, int32_t n_steps, double courant, double* t_inout);

/* Diagnostic with no counterpart in the reference: the number of (cell, stage) pairs with a blender
 * value alpha > 0 (cells that took the subcell FV blend of calc_rhs, dgsem/plens.cc:3737-3765) over
 * every RHS evaluation since the handle was created or last reset; reset != 0 clears it after
 * reading. */
int plens_gpu_blended_cell_stages(plens_gpu* h, int64_t* count, int32_t reset);

/* Device time (HIP events on the handle's stream) of the graph replays of the last plens_gpu_run. */
int plens_gpu_last_run_ms(const plens_gpu* h, double* ms);

/* Like plens_gpu_run but with the kernels launched one by one, each bracketed by HIP events;
 * afterwards plens_gpu_kernel_times() returns the average duration per kernel class. */
int plens_gpu_run_profiled(plens_gpu* h, int32_t n_steps, double courant, double* t_inout);

/* Last positivity failure: lowest failing DoF index and kind (1 = density, 2 = energy). */
int plens_gpu_last_failure(const plens_gpu* h, int64_t* dof, int32_t* kind);

/* Cells per workgroup of the cell kernels (autotuner hook); 0 = built-in default. */
int plens_gpu_set_tuning(plens_gpu* h, int32_t cells_per_block);

/* Average device time of each kernel class over the last plens_gpu_run (HIP events on the
 * handle's stream): times_ms[0..4] = gradient/blender, alpha, face flux, subcell FV, volume+RK. */
int plens_gpu_kernel_times(const plens_gpu* h, double* times_ms, int32_t n);

/* Local time stepping ("time integration" parameters request local stepping, local stepping start
 * time, limit local time step, local step neighbor/global factor; PLENS::calc_time_step,
 * dgsem/plens.cc:3877-3911): once the solution time at the start of a step exceeds start_time, every
 * cell advances with its own stable step, limited (limit != 0) to neighbor_factor x each interior
 * neighbour's step and global_factor x the global step; the returned time advances by the global
 * step. The solution time is taken from *t_inout of plens_gpu_run/rk4_step (NULL: continues). */
int plens_gpu_set_local_stepping(plens_gpu* h, int32_t request, double start_time, int32_t limit,
                                 double neighbor_factor, double global_factor);

/* Steady-state error after the last step (PLENS::calc_ss_error, dgsem/plens.cc:3985-4035):
 * ||(rhoE - rhoE_start)/dt_cell||_L2 / ||rhoE||_L2 with QGauss(N+1) through the mapping, rhoE_start
 * the energy at the start of that step and dt_cell its (local or global) step. PLENS_E_INVALID
 * before the first step. */
int plens_gpu_ss_error(plens_gpu* h, double* err);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): plens_gpu_ss_error with its
 * per-cell partial integrals in a fixed device array of the library (one per GPU, 2 x 2^18 + 2
 * doubles = 4 MiB, held until the process ends) instead of the handle's buffer; the call itself
 * allocates nothing. At most 2^18 cells (PLENS_E_INVALID beyond); the handle's device must be
 * current. Every call on one GPU uses that GPU's one array, so calls that overlap on a GPU (from
 * several threads or streams) race; the result equals plens_gpu_ss_error's bit for bit. */
int plens_gpu_ss_error_noalloc(plens_gpu* h, double* err);

/* Data output of PLENS::write() (dgsem/plens.cc:4170-4322), the fields only (files: plens_host.h).
 * write() runs inside the first RK stage (update_rk4, 4341-4350), after the stage-1 auxiliary
 * variables, blender and calc_time_step of the current state; this call computes those for the
 * current state (local steps switched on by the solution time t; rho E at the step start is not
 * recorded) and post_process (3950-3963), and returns, interpolated from the GLL nodes to the
 * equispaced (N+1)^3 lattice of every cell that DataOut::build_patches(mapping, N) evaluates
 * (lexicographic, x fastest):
 *   nodal[PLENS_OUT_NODAL][n_cells][(N+1)^3]: txx txy txz tyy tyz tzz qx qy qz rho p T u v w
 *   cell[PLENS_OUT_CELL][n_cells]: alpha, alpha_d, loc_dt (the cell's step).
 * The state is not changed; the time step and blender buffers hold the new values afterwards, so
 * a following plens_gpu_ss_error divides by the new steps as write() does. Single handles only.
 * The first call allocates the device buffer of the fields. */
/* Expression-valued parameters (deal.II FunctionParser syntax in x, y, z, t, as the prm strings):
 *   courant                 "Courant number" (dgsem/plens.cc:1484-1488), evaluated at the origin and
 *                           the step's start time (3816-3817), capped at 0.9 when >= 1; replaces the
 *                           courant argument of the run, step and calc_dt calls;
 *   wall_blender_limit      "wall blender limit" (2884-2898), at the origin and that time, in place of
 *                           plens_physics.wall_blender_limit;
 *   viscous_blending_region "viscous blending region" (2886-2889, 2953): alpha_d = min(1, alpha /
 *                           alpha_max) in cells whose center (the mean of the vertices, else of the
 *                           DoF locations) gives a value > 0, and 0 elsewhere (the default "1" is
 *                           every cell).
 * Each is nullable (NULL: the numeric parameter / every cell). PLENS_E_INVALID for a syntax error,
 * for a region without vertices or DoF locations in the mesh, or for a batch-merged handle; batches
 * reject handles with expressions. The solution time is the handle's (plens_gpu_set_time, and the
 * t of runs and steps). */
int plens_gpu_set_expressions(plens_gpu* h, const char* courant, const char* wall_blender_limit,
                              const char* viscous_blending_region);

/* Solution time of the handle (the BCs' cur_time for plens_gpu_calc_rhs; runs and steps set it
 * from their *t_inout and advance it). */
int plens_gpu_set_time(plens_gpu* h, double t);

#define PLENS_OUT_NODAL 15
#define PLENS_OUT_CELL 3
int plens_gpu_output_fields(plens_gpu* h, double courant, double t, double* nodal, double* cell);

/* Checkpoint and restart of the solution: the role of the ".ar" archive written by write()
 * (do_solution_transfer, 4117-4127, called at 4304-4308) and read by ICs::FromArchiveRestart
 * (IBCs/from_archive_restart.cc), whose p4est format is deal.II's own. Little-endian binary file:
 * header {char magic[8] = "PLENSGPU"; uint32 version = 1; int32 degree; int64 n_cells, n_dof;
 * double time; int64 step, output_counter; uint64 mesh_fingerprint, state_checksum} followed by
 * the state Q[5][n_dof] in FP64; checksum and fingerprint are FNV-1a 64 (of the state bytes; of the
 * degree, face_nbr, face_bid and min_vertex_dist of the mesh). Written through a temporary file
 * and renamed. Reading restores the state and returns time and counters (each nullable); a file
 * of another degree, size or mesh, a wrong checksum or trailing bytes give PLENS_E_INVALID and
 * leave the state as it was. As the reference's restart, only the conservative variables are
 * restored: the next step recomputes everything else. Single handles only. */
int plens_gpu_write_checkpoint(plens_gpu* h, const char* path, double time, int64_t step,
                               int64_t output_counter);
int plens_gpu_read_checkpoint(plens_gpu* h, const char* path, double* time, int64_t* step,
                              int64_t* output_counter);

/* Kernel timing with no counterpart in the reference: after one RHS evaluation of the current
 * state, launches kernel class `kernel` (0 gradient/blender, 2 face flux, 3 subcell FV residual of the
 * blended cells, 4 volume + RK update)
 * back to back on the handle's stream, `reps` times with the parameters of each of the 5 RK stages
 * in turn (new states go to a scratch buffer), between two HIP events, and returns the average
 * duration of one launch. The state is not modified; the residual registers are (they are dead
 * between steps). */
int plens_gpu_time_kernel(plens_gpu* h, int32_t kernel, int32_t reps, double* avg_ms);

/* Bandwidth probe with no counterpart in the reference: copies `bytes` between two device buffers
 * `iters` times with a streaming kernel; returns the average time per copy (2*bytes of HBM
 * traffic). Used to calibrate hardware byte counters and to state achievable bandwidth. */
int plens_gpu_stream_copy(int64_t bytes, int32_t iters, double* ms_per_copy);

const char* plens_gpu_strerror(int code);

/* ---- batched uneven solves ------------------------------------------------------------------
 * No counterpart in the reference, where each solve is its own PLENS::run (dgsem/plens.cc:1547)
 * driving update/update_rk4/calc_rhs (4484-4518, 4334-4403, 3688-3779). A batch takes the current
 * states of existing handles (any degrees and mesh sizes; per-handle boundary conditions, Sutherland
 * viscosity and time step) and advances all of them together: handles of one degree and precision
 * are merged into one group whose kernels run over the cells of all its instances, and one time
 * step of every group is captured into one HIP graph replayed per step. Every instance keeps its
 * own global time step (minimum over its own cells, as in calc_time_step 3872-3875). The
 * instances must agree on gamma, R, Pr, surface flux and blender parameters within a group.
 * The source handles are not modified. */
typedef struct plens_batch plens_batch;

/* tuning_db: path of a text file of autotuned launch shapes (cells per workgroup of the cell
 * kernels, workgroup size of the face kernel) keyed by degree, precision and log2 of the group's
 * cell count. Shapes found there are used; missing ones are timed on the group (RHS evaluations,
 * the state is untouched) and written back. NULL keeps the built-in defaults. */
int plens_gpu_batch_create(plens_gpu* const* instances, int32_t n, const char* tuning_db, plens_batch** out);
void plens_gpu_batch_destroy(plens_batch* b);

/* n_steps RK4 steps of every instance, instance i at Courant number courant[i]; t_inout[n]
 * (nullable) is advanced by each instance's own accumulated time. On PLENS_E_POSITIVITY the
 * failing instance and DoF are in plens_gpu_batch_last_failure. */
int plens_gpu_batch_run(plens_batch* b, int32_t n_steps, const double* courant, double* t_inout);
int plens_gpu_batch_get_state(const plens_batch* b, int32_t instance, double* Q);
int plens_gpu_batch_last_run_ms(const plens_batch* b, double* ms);
int plens_gpu_batch_last_failure(const plens_batch* b, int32_t* instance, int64_t* dof, int32_t* kind);
int plens_gpu_batch_n_groups(const plens_batch* b);
int plens_gpu_batch_group_info(const plens_batch* b, int32_t group, int32_t* degree, int32_t* n_instances,
                               int64_t* n_dof, int32_t* cells_per_workgroup, int32_t* face_workgroup);

#ifdef __cplusplus
}
#endif
#endif

/*
 * odis.h -- C ABI of geodesicodis_amd: the Laplace-Tidal-Equation (LTE) time
 * step of GeodesicODIS on the icosahedral-geodesic (SCVT) C-grid, run by
 * hand-written HIP kernels on one MI355X (gfx950).
 *
 * The reference has no FFI: its hot path is the pair of free functions
 *     int solveODIS(Globals *, Mesh *)    include/solver.h:10, src/solver.cpp:15-61
 *     int ab3Explicit(Globals *, Mesh *)  include/timeIntegrator.h:12, src/timeIntegrator.cpp:57-322
 * called from main (src/main.cpp:64).  This header is the seam beneath them:
 * plain pointers and sizes, no C++ types, int error codes (0 = OK, < 0 = error),
 * never exits or throws across the boundary.  The library owns every device
 * buffer; the caller owns host buffers, which are copied in and out and never
 * retained.
 *
 * Numbering: every array crossing this ABI uses the reference's numbering of
 * cells (grid-file node order), faces (first-visit order of AssignFaces,
 * src/mesh.cpp:524-666) and its face orientation (face_nodes[e][0] is the
 * "inner" node).  The device layout is renumbered internally for locality.
 */
#ifndef ODIS_H
#define ODIS_H

#ifdef __cplusplus
extern "C" {
#endif

#define ODIS_OK              0
#define ODIS_E_INVALID      -1   /* bad argument / shape                     */
#define ODIS_E_HIP          -2   /* HIP runtime error                        */
#define ODIS_E_NONFINITE    -3   /* non-finite dissipation (blow-up)         */
#define ODIS_E_UNSUPPORTED  -4   /* option with no live reference semantics  */
#define ODIS_E_NOMEM        -5   /* device or host allocation failed        */
#define ODIS_E_RANGE        -6   /* index / size outside what was created    */
#define ODIS_E_STOPPED      -7   /* odis_request_stop() honoured             */
#define ODIS_E_IO           -8   /* file not found / unreadable              */

/* Arithmetic of the stencil state.  ODIS_FP64 is the plain method (drop-in);
 * ODIS_FP32 runs the stencils and state in FP32 with FP64 dissipation sums. */
#define ODIS_FP64 0
#define ODIS_FP32 1

/* Potential / surface / friction codes: same ordinal values as the
 * reference enums (include/globals.h:46-82). */
enum odis_potential {
    ODIS_OBLIQ = 0, ODIS_OBLIQ_WEST, ODIS_OBLIQ_EAST, ODIS_ECC_RAD, ODIS_ECC_LIB, ODIS_ECC,
    ODIS_ECC_WEST, ODIS_ECC_EAST, ODIS_FULL, ODIS_FULL2, ODIS_TOTAL, ODIS_ECC_W3,
    ODIS_OBLIQ_W3, ODIS_PLANET, ODIS_PLANET_OBL, ODIS_GENERAL, ODIS_NONE
};
enum odis_surface { ODIS_FREE = 0, ODIS_FREE_LOADING, ODIS_LID_LOVE, ODIS_LID_MEMBR,
                    ODIS_LID_NUM, ODIS_LID_INF };
enum odis_friction { ODIS_LINEAR = 0, ODIS_QUADRATIC };

typedef struct odis_ctx odis_ctx;
typedef struct odis_grid odis_grid;
typedef struct odis_mesh odis_mesh;

/* ------------------------------------------------------------------------ */
/* Hot-path operands.  Replaces the Mesh members the step reads             */
/* (include/mesh.h:77-221: face_nodes, face_interp_friends/weights,         */
/* face_len, face_node_dist, face_centre_pos_sph, face_area, faces,         */
/* node_face_dir, control_volume_surf_area_map, operatorGradient,           */
/* operatorDivergence, operatorCoriolis, trig tables).                      */
/* ------------------------------------------------------------------------ */
typedef struct odis_mesh_desc {
    int n_cells;                /* N                                                 */
    int n_faces;                /* E = 3N - 6                                        */
    const int    *face_nodes;   /* [E][2] inner, outer cell                          */
    const double *grad_coef;    /* [E][2] G row (-m_in/d_e, m_out/d_e), src/mesh.cpp:3076-3080 */
    const int    *trisk_count;  /* [E]    9 or 10                                    */
    const int    *trisk_faces;  /* [E][10] face_interp_friends, -1 padded            */
    const double *trisk_coef;   /* [E][10] w_ee' l_e' / d_e (src/interpolation.cpp:43-46) */
    const double *face_lat;     /* [E]    face centre latitude (Coriolis)            */
    const double *face_area;    /* [E]    d_e l_e (src/mesh.cpp:1093)                */
    const int    *cell_faces;   /* [N][6] faces, -1 padded                           */
    const double *div_coef;     /* [N][6] -dir l_e / A_i (src/mesh.cpp:3242-3247)    */
    const double *cell_lat;     /* [N]    node latitude  (forcing fields)            */
    const double *cell_lon;     /* [N]    node longitude                             */
    /* Optional (may be null): with them the step uses the cell-local form of
     * the TRiSK operator (kite fractions per cell, derived from and checked
     * against trisk_coef); without them the tabulated trisk_faces/coef.     */
    const double *face_len;     /* [E]    l_e, Voronoi edge length (mesh.edge_len)  */
    const double *face_dist;    /* [E]    d_e, node-to-node distance (mesh.node_dists) */
    const double *cell_area;    /* [N]    A_i, control-volume area of div_coef      */
} odis_mesh_desc;

/* Per-instance physics, the Globals members the step reads
 * (include/globals.h:120-179) after applySurfaceBCs and period rounding. */
typedef struct odis_params {
    double g;               /* gravity at the ocean top                    */
    double h;               /* ocean thickness                             */
    double alpha;           /* linear drag (s^-1) or quadratic c_D         */
    double omega;           /* rotation rate after period rounding         */
    double radius;          /* ocean-top radius                            */
    double love_reduct;     /* forcing prefactor                           */
    double ecc;
    double obliquity;       /* radians                                     */
    double dt;              /* quantised time step (src/mesh.cpp:1566-1623)*/
    double shell_thickness; /* added back to the forcing radius for LID_*  */
    double semimajor;       /* PLANET forcing                              */
    int tide;               /* enum odis_potential                         */
    int surface;            /* enum odis_surface                           */
    int friction;           /* enum odis_friction                          */
    int init_load;          /* 1: AB3 from step 0 (INIT_LOAD semantics,
                               src/temporalOperators.cpp:36)               */
    int quad_drag;          /* 1: quadratic drag in the momentum tendency
                               (src/drag.cpp:19-38; never live in the
                               reference AB3 loop)                         */
    int reserved;
} odis_params;

/* ---- device context: the time loop (src/timeIntegrator.cpp:57-322) ------ */

/* Allocate device operands and n_inst instance states (all at rest, step 0).
 * All instances share the mesh; params holds n_inst entries. */
int odis_create(const odis_mesh_desc *mesh, const odis_params *params, int n_inst,
                int precision, odis_ctx **out);
int odis_destroy(odis_ctx *ctx);
const char *odis_last_error(const odis_ctx *ctx);

/* Replaces getInitialConditions / loadInitialConditions
 * (src/initialConditions.cpp:19-144, 279-313): v[E], dvdt[E][3], eta[N],
 * detadt[N][3] in the reference's Array2D row layout; iter = step counter.
 * Null pointers leave that array unchanged. */
int odis_set_state(odis_ctx *ctx, int inst, const double *v, const double *dvdt,
                   const double *eta, const double *detadt, long long iter);
int odis_get_state(odis_ctx *ctx, int inst, double *v, double *dvdt, double *eta,
                   double *detadt, long long *iter);

/* Advance every instance n_steps iterations of the ab3Explicit loop body
 * (src/timeIntegrator.cpp:207-277).  Stream-ordered; returns after enqueueing
 * unless ODIS_SYNC is set in flags. */
#define ODIS_SYNC 1
int odis_step(odis_ctx *ctx, int n_steps, int flags);
int odis_sync(odis_ctx *ctx);

/* FP64 dissipation of the current state of each instance (updateEnergy,
 * src/energy.cpp:13-62; W/m^2, multiply by 4 pi R^2 for W).  Returns
 * ODIS_E_NONFINITE if any instance is non-finite. */
int odis_get_dissipation(odis_ctx *ctx, double *per_inst);
/* Per-step dissipation series of instance inst: out[k] = dissipation of the
 * state at step first + k (ring of the last odis_series_capacity() steps). */
int odis_get_dissipation_series(odis_ctx *ctx, int inst, long long first, int count,
                                double *out);
int odis_series_capacity(const odis_ctx *ctx);
/* Orbit statistics for a run to converged dissipation.  The reference logs
 * the dissipation at its output cadence only (src/timeIntegrator.cpp:280-304,
 * src/energy.cpp:13-62); this is the per-orbit form of the same quantity.
 * For the n instances inst[j] whose orbit of steps_per_orbit[j] steps ended
 * at the current step: out[3j] = mean of the per-step FP64 dissipation over
 * those steps (W/m^2), out[3j+1] and out[3j+2] = the periodicity residuals
 * ||v - v'|| / ||v|| and ||eta - eta'|| / ||eta|| (FP64) against the state
 * v', eta' the previous call stored for that instance (NaN at its first
 * call); the call then stores the current state.  Orbits must lie within the
 * series held since the last odis_set_state (ODIS_E_RANGE otherwise). */
int odis_orbit_stats(odis_ctx *ctx, int n, const int *inst, const int *steps_per_orbit, double *out);
/* Retire instances from a batch: keep[0..n_keep) (strictly increasing)
 * become instances 0..n_keep-1, in that order, with their state, history,
 * parameters, dissipation series, orbit snapshots and shell factors; the
 * others are dropped and later steps cost only the kept instances.  No
 * counterpart in the reference (one instance per run). */
int odis_retire(odis_ctx *ctx, const int *keep, int n_keep);
/* Reconstructed (east, north) velocity per face, interpolateVelocity
 * (src/interpolation.cpp:26-62); uv[E][2]. */
int odis_get_velocity_en(odis_ctx *ctx, int inst, const double *face_normal, double *uv);
/* Geodesic -> lat-lon interpolation, interpolateGG2LLConservative
 * (src/interpolation.cpp:208-287): ll = W x with x the interleaved
 * [eta_i, r d(eta)/dlat_i, r d(eta)/dlon_i] of every cell i (3N columns in
 * reference cell numbering).  As in the reference, the gradient entries of x
 * are zero (its pressureGradient call is commented out, :236), so only the
 * eta columns contribute.  W is a CSR matrix of n_rows = (180/d)(360/d) rows
 * in the reference's row-major lat-lon order; it is copied to the device. */
int odis_interp_ll_set(odis_ctx *ctx, int n_rows, int n_cols, const int *row_ptr,
                       const int *col, const double *val);
int odis_interp_ll(odis_ctx *ctx, int inst, double *ll);
/* Cartesian velocity at cell centres by radial basis functions,
 * interpolateVelocityCartRBF (src/interpolation.cpp:64-174): xyz = R v with
 * R the reference's operatorRBFinterp (src/mesh.cpp:2263-2361, 360-430), a
 * CSR matrix of 3N rows (3i + x, y, z in reference cell numbering) and E
 * columns (reference face numbering) whose three rows of a cell hold the
 * same <= 6 faces in the same order (odis_mesh_get "rbf_row_ptr", "rbf_col",
 * "rbf_val" build it).  It is copied to the device as one 3 x 6 block per
 * cell; xyz[N][3] comes back in reference numbering. */
int odis_rbf_set(odis_ctx *ctx, int n_rows, int n_cols, const int *row_ptr, const int *col,
                 const double *val);
int odis_velocity_xyz(odis_ctx *ctx, int inst, double *xyz);

/* DATA/data.h5 in the reference's layout (OutFiles::CreateHDF5Framework,
 * DumpGridData, DumpData, src/outFiles.cpp:138-681): float32 datasets
 * [n_slices][n] per field, slice k - 1 holding output k, the face centres in
 * degrees as "face longitude" / "face latitude" (when face_lat / face_lon,
 * radians, are given).  Fields (a mask at open, one field per write):
 *   ODIS_H5_DISPLACEMENT  "displacement" [N]: data eta[N]
 *   ODIS_H5_VELOCITY      "east velocity", "north velocity" [E]: data uv[E][2]
 *   ODIS_H5_VELOCITY_XYZ  "x velocity", "y velocity", "z velocity" [N]: data xyz[N][3]
 *   ODIS_H5_DISSIPATION   "dissipated energy" [E]: data e[E]
 *   ODIS_H5_DISS_AVG      "dissipation avg output" (1-D): data[0]
 * The HDF5 C library is loaded at the first open (ODIS_E_UNSUPPORTED when
 * it cannot be; odis_h5_last_error says why).  Host-only. */
typedef struct odis_h5 odis_h5;
#define ODIS_H5_DISPLACEMENT 1
#define ODIS_H5_VELOCITY     2
#define ODIS_H5_VELOCITY_XYZ 4
#define ODIS_H5_DISSIPATION  8
#define ODIS_H5_DISS_AVG     16
int odis_h5_open(const char *path, int n_cells, int n_faces, int n_slices, int fields,
                 const double *face_lat, const double *face_lon, odis_h5 **out);
int odis_h5_write(odis_h5 *h, int field, int slice, const double *data);
int odis_h5_close(odis_h5 *h);
const char *odis_h5_last_error(void);

/* Self-gravity / ice-shell coupling through a spherical-harmonic projection
 * of eta (the reference's commented design: CalcLegendreFuncs,
 * src/mesh.cpp:2154-2259, and pressureGradientSH, src/spatialOperators.cpp:
 * 387-460).  Each step the pressure gradient acts on
 *     eta_eff = eta + sum_{l=2..L} sum_m (beta_l - 1) c_lm Y_lm,
 * c = least-squares coefficients of eta on the nodes up to degree L
 * (as SHExpandLSQ, src/extractSHCoeffGG.f95:29), Y_lm real 4pi-normalised
 * harmonics, index k = l*l + l + m.  beta holds n_inst rows of (l_max + 1)
 * per-degree factors (e.g. membraneNuBeta's beta_l for LID_MEMBR).
 * mode: ODIS_SH_OFF; ODIS_SH_FP64 (basis stored in FP64, FP64 sums);
 * ODIS_SH_SPLIT_BF16 (basis stored as bf16 hi + bf16 lo, FP32 accumulation);
 * ODIS_SH_RECUR (basis evaluated on the fly by the Legendre recurrence in
 * the context's precision; nothing of size N x K is kept); ODIS_SH_MFMA
 * (the analysis and synthesis of all instances as matrix products on the
 * bf16 matrix cores: basis and operands split into bf16 hi + lo, three
 * products per term, FP32 accumulation, FP64 partial sums; the basis is
 * read once per step for the whole batch, so it suits batches of
 * instances on one grid).  The Gram matrix, its inverse and the degree
 * factors are FP64 in every mode. */
#define ODIS_SH_OFF 0
#define ODIS_SH_FP64 1
#define ODIS_SH_SPLIT_BF16 2
#define ODIS_SH_RECUR 3
#define ODIS_SH_MFMA 4
int odis_sh_set_basis(odis_ctx *ctx, int l_max, const double *beta, int mode);
/* Coefficients (K = (l_max+1)^2) of a cell field in reference numbering, and
 * the synthesis of a field from coefficients, with the context's basis. */
int odis_sh_project(odis_ctx *ctx, const double *field, double *coeffs);
int odis_sh_synthesize(odis_ctx *ctx, const double *coeffs, double *field);
/* odis_sh_project without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as odis_sh_project (bit for bit);
 * its device scratch (the field twice and the K coefficients) is a
 * file-scope array of 2^20 doubles (8 MiB per GPU, held until the process
 * ends), so it returns ODIS_E_RANGE when 2 N + K exceeds 2^20.  Calls that
 * overlap on one GPU share that array and race. */
int odis_sh_project_noalloc(odis_ctx *ctx, const double *field, double *coeffs);

/* Async-signal-safe stop request (replaces CatchExit, src/timeIntegrator.cpp:34-37):
 * odis_step returns ODIS_E_STOPPED at the next step boundary. */
int odis_request_stop(odis_ctx *ctx);

/* Launch options (no counterpart in the reference):
 *  ODIS_OPT_GRAPHS   1: replay chunks of 64 steps from captured HIP graphs
 *                    (default 0: direct launches measured faster on MI355X;
 *                    results are bit for bit the same).
 *  ODIS_OPT_COMPACT  1: cell-local TRiSK form (default when the descriptor
 *                    carries face_len / face_dist / cell_area and the form
 *                    reproduces trisk_coef); 0: tabulated trisk_faces/coef,
 *                    the reference's operator coefficient by coefficient.
 *                    The two differ by rounding only.  Setting 1 on a context
 *                    without the compact form returns ODIS_E_UNSUPPORTED.
 *  ODIS_OPT_ONE_KERNEL 1: with the compact form, one kernel per step over
 *                    patches of cells (default; with an SH basis it follows
 *                    the SH kernels of the step); 0: a face and a cell
 *                    kernel per step.  The state is the same bit
 *                    for bit; dissipation sums differ by rounding (another
 *                    grouping of the per-face terms).
 *  ODIS_OPT_PATCH_CELLS at most 64 (default), 128 or 256 cells per patch of
 *                    the one-kernel step.
 *  ODIS_OPT_STEP_BLOCKS workgroups of the one-kernel step; each runs a
 *                    contiguous run of patches (0, default: one per patch;
 *                    reading it returns the count in use).  The dissipation
 *                    sums depend on it by rounding only.
 *  ODIS_OPT_STEP_LANES lanes per cell of the one-kernel step: 4 (default,
 *                    about one face per lane) or 2 (about two).
 *  ODIS_OPT_SH_CELLS cells per lane of the on-the-fly SH analysis: 0
 *                    (default: for degrees 8, 16 and 32, which have
 *                    compiled kernels, the fewest even count >= 4 whose grid
 *                    fits the workgroups the device holds at once; 8 for the
 *                    others), 4, 8 or 16; the coefficients depend on it by
 *                    rounding only.
 *  ODIS_OPT_INTERLEAVE a batch steps with its state held interleaved
 *                    (groups of 16 instances, instances innermost), one
 *                    workgroup per patch and 16 instances, so that the
 *                    operators are read once per 16 instances and each state
 *                    access is 16 consecutive values: 0 off, 1 on, 2
 *                    (default) on for 16 instances or more on grids of 20000
 *                    cells or more.  It applies to the one-kernel step with
 *                    64-cell patches, without an SH basis or with
 *                    ODIS_SH_MFMA; the instance-major arrays are brought up
 *                    to date when the state is read or set.  The state is the
 *                    same bit for bit; dissipation sums differ by rounding
 *                    (another grouping of the per-face terms).
 *  ODIS_OPT_INTERLEAVE_ACTIVE (read only) 1 when the next step would run
 *                    interleaved.
 *  ODIS_OPT_IL_UNROLL faces and (cell, instance) pairs each lane of the
 *                    interleaved step keeps in flight: 0 (default: 4 and 2
 *                    in FP64, 2 and 1 in FP32), 1 (1, 1), 2 (2, 1), 3 (4, 2).
 *                    The state is the same bit for bit.
 *  ODIS_OPT_IL_ORDER  workgroup order of the interleaved step: 0 (default)
 *                    each group of 16 instances has its patches spread over
 *                    the XCDs as contiguous ranges; 1 the groups of a patch
 *                    run next to each other on one XCD (its operands read
 *                    from L2 once for all groups); 2, 4 or 8 the same for
 *                    bands of that many groups, band after band (each XCD
 *                    streams through the state of fewer groups at a time).
 *                    Same state bit for bit.
 *  ODIS_OPT_IL_RECORDS operands of the interleaved step: 1 (default) from
 *                    per-face and per-cell records that each group of 16
 *                    lanes loads once and shares by row broadcasts (tides
 *                    with at most 4 forcing fields), 0 every lane loads
 *                    every operand.  Same state bit for bit. */
#define ODIS_OPT_GRAPHS     1
#define ODIS_OPT_COMPACT    2
#define ODIS_OPT_ONE_KERNEL 3
#define ODIS_OPT_PATCH_CELLS 4
#define ODIS_OPT_STEP_BLOCKS 5
#define ODIS_OPT_STEP_LANES  6
#define ODIS_OPT_SH_CELLS    7
#define ODIS_OPT_INTERLEAVE  8
#define ODIS_OPT_INTERLEAVE_ACTIVE 9
#define ODIS_OPT_IL_UNROLL 10
#define ODIS_OPT_IL_ORDER 11
#define ODIS_OPT_IL_RECORDS 12
int odis_set_option(odis_ctx *ctx, int option, int value);
int odis_get_option(odis_ctx *ctx, int option, int *value);
/* Timing for benchmarks: with stride T > 0, the kernels of every step n of
 * the next odis_step calls with n % T == 0 are launched with dispatch-stamped
 * start/stop HIP events on the context stream; odis_get_timing returns the
 * average duration (ms) of the face kernel, of the cell kernel and of a whole
 * sampled step (face start to cell end), over the sampled steps of the last
 * odis_step call.  T = 0 disables.  odis_enable_timing creates the events a
 * chunk of steps needs; the events are read (waiting for the stream) when
 * odis_get_timing or odis_get_kernel_times asks for them, or when the next
 * odis_step starts, so a timed odis_step adds no host work of its own. */
int odis_enable_timing(odis_ctx *ctx, int stride);
int odis_get_timing(odis_ctx *ctx, double *face_ms, double *cell_ms, double *step_ms);
/* The same averages and those of the SH kernels of sampled steps, indexed
 * by ODIS_T_*; fills min(n, ODIS_T_COUNT) entries. */
#define ODIS_T_MOMENTUM    0
#define ODIS_T_CONTINUITY  1
#define ODIS_T_STEP        2
#define ODIS_T_SH_ANALYSIS 3
#define ODIS_T_SH_COEF     4
#define ODIS_T_SH_SYNTH    5
#define ODIS_T_COUNT       6
int odis_get_kernel_times(odis_ctx *ctx, double *ms, int n);

/* ---- learned warm start: the network (no counterpart in the reference) --- */

/* A fully connected network n_in -> width x depth (tanh) -> n_out (linear)
 * on the GPU, FP32 weights and products (matrix cores), for the warm-start
 * prediction of periodic states.  Parameters: per layer l, W_l [out][in]
 * row-major then b_l [out]; Glorot-uniform weights and zero biases from
 * seed at creation.  Loss: sum (y - t)^2 / (2 n n_out) over the rows. */
typedef struct odis_mlp odis_mlp;
int odis_mlp_create(int n_in, int width, int depth, int n_out, unsigned long long seed, odis_mlp **out);
int odis_mlp_destroy(odis_mlp *m);
long long odis_mlp_param_count(const odis_mlp *m);
int odis_mlp_get_params(odis_mlp *m, float *params);
/* also resets the Adam moments */
int odis_mlp_set_params(odis_mlp *m, const float *params);
/* y[n][n_out] = network(x[n][n_in]) */
int odis_mlp_forward(odis_mlp *m, int n, const float *x, float *y);
/* the training set, copied to the device: x[n][n_in], y[n][n_out] */
int odis_mlp_set_data(odis_mlp *m, int n, const float *x, const float *y);
/* loss and gradient (param_count floats, nullable) over rows idx[0..n) of the
 * training set; the parameters do not change */
int odis_mlp_loss_grad(odis_mlp *m, int n, const int *idx, double *loss, float *grad);
/* Adam (beta 0.9, 0.999, eps 1e-8) over `epochs` passes of the training set
 * in shuffled minibatches of `batch` rows (order from seed), learning rate
 * from lr_start to lr_end by cosine decay; epoch_loss (nullable, epochs
 * entries) receives each epoch's mean minibatch loss.  Deterministic for a
 * given seed. */
int odis_mlp_train(odis_mlp *m, int epochs, int batch, float lr_start, float lr_end, unsigned long long seed,
                   double *epoch_loss);
const char *odis_mlp_last_error(void);

/* ---- host: grid and mesh (src/mesh.cpp:32-197, 436-1154, 2808-3287) ---- */

/* Check of the cell-local TRiSK form on the host (odis_create runs the same
 * check): ODIS_OK when the form reproduces every trisk_coef and div_coef to
 * 1e-12 relative; the largest relative deviations found are returned (any
 * pointer may be null).  ODIS_E_UNSUPPORTED with the reason in
 * odis_host_last_error otherwise. */
int odis_mesh_compact_check(const odis_mesh_desc *mesh, double *coef_err, double *div_err,
                            double *area_err);

/* Read a reference-format grid file (src/mesh.cpp:4016-4101). */
int odis_grid_read(const char *path, odis_grid **out);
/* Icosahedral bisection grid at reference level `level` (N = 10*4^(level-1)+2)
 * with `lloyd_iters` fixed-topology Lloyd (centroidal Voronoi) sweeps. */
int odis_grid_generate(int level, int lloyd_iters, odis_grid **out);
int odis_grid_write(const odis_grid *g, const char *path);
int odis_grid_size(const odis_grid *g);
/* lat, lon [N] radians; friends [N][6] (-1 pad); clat, clon [N][6] Voronoi vertices. */
int odis_grid_arrays(const odis_grid *g, double *lat, double *lon, int *friends,
                     double *clat, double *clon);
int odis_grid_from_arrays(int n, const double *lat, const double *lon, const int *friends,
                          const double *clat, const double *clon, odis_grid **out);
int odis_grid_free(odis_grid *g);

/* Assemble the TRiSK C-grid operands on the host (Mesh::Mesh). */
int odis_mesh_build(const odis_grid *g, double radius, odis_mesh **out);
int odis_mesh_sizes(const odis_mesh *m, int *n_cells, int *n_faces, int *n_vertices);
/* Fill a descriptor whose pointers stay valid while m lives. */
int odis_mesh_describe(const odis_mesh *m, odis_mesh_desc *desc);
/* Copy a named array (same names as the Mesh members, e.g. "face_len",
 * "face_node_dist", "face_normal", "interp_weights", "cv_area"; also
 * "face_normal_xyz" [E][3] and the RBF velocity operator "rbf_row_ptr"
 * [3N + 1], "rbf_col", "rbf_val"); returns the element count, or < 0.  dst
 * may be null to query the count. */
int odis_mesh_get(const odis_mesh *m, const char *name, void *dst);
int odis_mesh_free(odis_mesh *m);

/* ---- host: configuration (Globals, src/globals.cpp) ---------------------- */

/* Period rounded to an even number of seconds and the matching rate,
 * src/globals.cpp:209-214. */
int odis_round_period(double ang_vel, double *period, double *omega);
/* CalcMaxTimeStep, src/mesh.cpp:1601-1618: dt = P / n_steps, n_steps the
 * smallest multiple of 100 with dt <= target. */
int odis_quantise_dt(double period, double target_dt, double *dt, int *n_steps);
/* Read case_dir/input.in with the reference semantics, apply applySurfaceBCs
 * and the dt quantisation, and fill the step parameters.  Any of the
 * optional outputs may be null. */
int odis_config_read(const char *case_dir, odis_params *params, int *steps_per_orbit,
                     double *end_time_orbits, int *grid_level, int *output_time);
/* Per-degree pressure factors beta_l, l = 0..l_max, of a surface type:
 * LID_MEMBR -> membraneNuBeta (src/membraneConstants.cpp:8-144) with the
 * top radius, shell and ocean thicknesses; FREE_LOADING -> 1 - loading factor
 * (src/boundaryConditions.cpp:29-77) with gravity g; otherwise 1. */
int odis_surface_beta(int surface, int l_max, double top_radius, double g, double shell_thickness,
                      double ocean_thickness, double *beta);
/* analyticalLTE, the live OBLIQ_WEST case (src/analyticalLTE.cpp:47-228,
 * 133-180): the forced, Rayleigh-damped degree-2 obliquity-west tide of the
 * ocean p->radius, p->omega, p->g, p->h, p->obliquity (rad), p->alpha;
 * out[6] = (U, V, dU/dt, dV/dt, eta, deta/dt) at (lat, lon) radians and time
 * t, with U, V the components whose projection on a face normal (nx, ny)
 * the reference's analyticalInitialConditions takes as the face velocity
 * (src/initialConditions.cpp:146-206).  The INIT_ANALYTICAL start of
 * odis_run_case uses it. */
int odis_analytical_obliq_west(const odis_params *p, double lat, double lon, double t, double *out);
/* Real 4pi-normalised spherical harmonics to degree l_max at n points
 * (radians), the basis of the SH coupling: Y[i][l*l + l + m], m >= 0 the
 * cos(m lon) and m < 0 the sin(|m| lon) harmonic, no Condon-Shortley phase.
 * Host-only (the reference's commented CalcLegendreFuncs, src/mesh.cpp:
 * 2154-2259, tabulated the same functions). */
int odis_sh_basis_eval(int l_max, int n, const double *lat, const double *lon, double *Y);
/* Last error of a host entry point (grid, mesh, config). */
const char *odis_host_last_error(void);

/* ---- host: drivers (solveODIS / ab3Explicit) ----------------------------- */

/* Run the reference's program flow from a case directory holding input.in
 * (src/main.cpp:46-68): Globals, Mesh, solveODIS.  Writes DATA/OUTPUT.txt with
 * the reference's log lines and InitialConditions/{vel,pres}_init.txt.
 * grid_dir (nullable) overrides where grid_l<g>.txt is looked up; when the
 * file is absent the grid is generated. */
int odis_run_case(const char *case_dir, const char *grid_dir, int precision);

#ifdef __cplusplus
}
#endif
#endif /* ODIS_H */

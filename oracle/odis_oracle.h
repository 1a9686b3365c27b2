/*
 * odis_oracle.h -- CPU FP64 restatement of GeodesicODIS's LTE time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product (geodesicodis_amd) never links it.
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each part is listed in the header of odis_oracle.c.
 */
#ifndef ODIS_ORACLE_H
#define ODIS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Potential enum: same ordinal values as the reference's `enum Potential`
 * (include/globals.h:66-82). */
enum orc_potential {
    ORC_OBLIQ = 0, ORC_OBLIQ_WEST, ORC_OBLIQ_EAST, ORC_ECC_RAD, ORC_ECC_LIB,
    ORC_ECC, ORC_ECC_WEST, ORC_ECC_EAST, ORC_FULL, ORC_FULL2, ORC_TOTAL,
    ORC_ECC_W3, ORC_OBLIQ_W3, ORC_PLANET, ORC_PLANET_OBL, ORC_GENERAL, ORC_NONE
};
/* Surface enum (include/globals.h:48-53). */
enum orc_surface { ORC_FREE = 0, ORC_FREE_LOADING, ORC_LID_LOVE, ORC_LID_MEMBR,
                   ORC_LID_NUM, ORC_LID_INF };
/* Friction enum (include/globals.h:46-47). */
enum orc_friction { ORC_LINEAR = 0, ORC_QUADRATIC };

typedef struct orc_params {
    double g;               /* surface gravity (after BCs)          */
    double h;               /* ocean thickness                       */
    double alpha;           /* linear drag / quadratic c_D           */
    double omega;           /* rotation rate after period rounding   */
    double radius;          /* ocean-top radius (after BCs)          */
    double love_reduct;     /* forcing prefactor                     */
    double ecc;             /* eccentricity                          */
    double obliquity;       /* radians                               */
    double dt;              /* quantised step                        */
    double shell_thickness; /* added back to the forcing radius for LID_* */
    double semimajor;       /* PLANET forcing only                   */
    int tide;               /* enum orc_potential                    */
    int surface;            /* enum orc_surface                      */
    int friction;           /* enum orc_friction                     */
    int init_load;          /* 1: AB3 from step 0 (INIT_LOAD)        */
    int quad_drag_momentum; /* 1: apply quadraticDrag in momentum (config-4 semantics) */
    int pad_;
} orc_params;

/* ---- mesh ---------------------------------------------------------------- */
void *orc_mesh_build(int n_nodes, const double *node_lat, const double *node_lon,
                     const int *friends, const double *cen_lat, const double *cen_lon,
                     double radius, int *err);
void  orc_mesh_free(void *mesh);
void  orc_mesh_sizes(void *mesh, int *n_nodes, int *n_faces, int *n_vertices);
/* Copy a named per-node/per-face array into dst; returns element count or -1. */
int   orc_mesh_get(void *mesh, const char *name, void *dst);
/* CSR operators (Eigen RowMajor semantics: columns sorted ascending). */
int   orc_mesh_grad_csr(void *mesh, int *rowptr, int *col, double *val);
int   orc_mesh_div_csr(void *mesh, int *rowptr, int *col, double *val);
int   orc_mesh_coriolis_nnz(void *mesh);
int   orc_mesh_coriolis_csr(void *mesh, double omega, int *rowptr, int *col, double *val);

/* ---- time stepping ------------------------------------------------------- */
void *orc_sim_create(void *mesh, const orc_params *p);
void  orc_sim_free(void *sim);
void  orc_sim_set_state(void *sim, const double *v, const double *dvdt3,
                        const double *eta, const double *detadt3, int iter);
void  orc_sim_get_state(void *sim, double *v, double *dvdt3, double *eta,
                        double *detadt3, int *iter);
/* Advance nsteps; diss (nullable) receives the dissipation after each step. */
int   orc_sim_step(void *sim, int nsteps, double *diss);
double orc_sim_dissipation(void *sim);
void  orc_sim_forcing(void *sim, double time, double *potential);
void  orc_sim_interp_velocity(void *sim, const double *v, double *uv);
void  orc_sim_momentum(void *sim, const double *v, const double *eta, double *dvdt);
/* pressure term acts on eta + Pb (Pc eta); Pb [N][K], Pc [K][N]; K = 0 clears */
void  orc_sim_set_projector(void *sim, int K, const double *Pb, const double *Pc);

/* ---- scalar helpers ------------------------------------------------------ */
void  orc_round_period(double ang_vel, double *period, double *omega);
void  orc_quantise_dt(double period, double target_dt, double *dt, int *n_steps);
void  orc_analytical_obliq_west(const orc_params *p, double lat, double lon,
                                double t, double *out6);
/* OpenMP threads of the row loops (n >= 1), and the count in use; the
 * results do not depend on it. */
void  orc_set_threads(int n);
int   orc_get_threads(void);
void  orc_membrane_nu_beta(double radius, double shell_thickness, double ocean_thickness,
                           int l_max, double *nu, double *beta,
                           double *radius_ocean, double *grav_ocean);
/* The same with the body mass and the core / shell rigidities as arguments. */
void  orc_membrane_nu_beta_gen(double radius, double shell_thickness, double ocean_thickness,
                               int l_max, double mass_total, double rig_core, double rig_shell,
                               double *nu, double *beta, double *radius_ocean, double *grav_ocean);

/* ---- warm-start network (FP64 checker of the GPU network) --------------- */
void  orc_mlp_forward(int n_in, int width, int depth, int n_out, const double *params, int n,
                      const double *x, double *y);
double orc_mlp_loss_grad(int n_in, int width, int depth, int n_out, const double *params, int n,
                         const double *x, const double *t, double *grad);

#ifdef __cplusplus
}
#endif
#endif

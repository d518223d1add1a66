/*
 * FP64 CPU oracle for the soap-film time step (plain, non-BIMOCQ form).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product (soapbubble_amd/libsoapfilm.so) never links, loads or calls it.
 *
 * Every array is dense, row-major by theta with phi contiguous (the
 * reference's index convention, header.cuh:24-35, solver.cu:96-103):
 *   u_theta : (nTheta-1) x nPhi   (theta = (j+1) h,   phi = i h)
 *   u_phi   :  nTheta    x nPhi   (theta = (j+1/2) h, phi = (i-1/2) h)
 *   eta,gam :  nTheta    x nPhi   (theta = (j+1/2) h, phi = i h)
 * with nPhi = 2 nTheta and h = pi / nTheta (bubble.cu:2576-2577).
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t n_theta;       /* nTheta; nPhi = 2 nTheta                          */
    int32_t gravity;       /* header.cuh:94 `gravity`                          */
    int32_t evaporation;   /* header.cuh:95 `evaporation`                      */
    int32_t airflow;       /* header.cuh:96 `airflow`                          */
    int32_t floorf_mode;   /* 1: round sample indices through float like the
                              reference's floorf on doubles (bubble.cu:183)    */
    int32_t tol_absolute;  /* 0: stop on ||r||/||b|| <= tol; 1: on ||r|| <= tol
                              (AmgX "convergence": "ABSOLUTE", PCGF_F.json)    */
    int32_t max_iters;     /* solver iteration cap                             */
    int32_t viscous;       /* header.cuh:99 `VISCOUS` (off in the reference)   */
    double dt;             /* non-dimensional step, bubble.cu:2578             */
    double M, Cr, gs, gamma_max, eva, radius, U;
    double time_s;         /* currentTimeGlobal: air-flow time in seconds      */
    double tol;            /* solver tolerance                                  */
    double re;             /* Re^-1 = nu / (U radius), bubble.cu:2575          */
} sfo_params;

typedef struct {
    int32_t iters;
    int32_t converged;
    int32_t broken;        /* 1 if no cell has eta > 0 (solver.cu:568-574)     */
    int32_t pad_;
    double rel_residual;   /* ||b - A x|| / ||b||, recomputed in FP64           */
    double abs_residual;
    double volume;         /* sum eta_ij sin(theta_j) h^2                      */
} sfo_stats;

/* Parameter derivation of Bubble::Bubble and Bubble::run (bubble.cu:2573-2580,
 * 2627-2638). Physical inputs as in config.txt. */
void sfo_derive(double radius, double H, double U, double Gamma_m, double T,
                double dt_seconds, int gravity, int evaporation, sfo_params* out);

/* Samplers (bubble.cu:178-376). kind: 0 centred, 1 u_phi, 2 u_theta. */
double sfo_sample(const sfo_params* p, int kind, const double* field,
                  double theta_id, double phi_id);
/* Great-circle back-trace of a point (bubble.cu:416-506); out = (theta,phi)/h. */
void sfo_trace(const sfo_params* p, const double* uth, const double* uph,
               double theta_id, double phi_id, double* out2);

/* Solver::advection, plain branch (bubble.cu:1171-1214): new fields. */
void sfo_advect(const sfo_params* p, const double* uth, const double* uph,
                const double* eta, const double* gam, double* uth_out,
                double* uph_out, double* eta_out, double* gam_out);
/* airFlowU / airFlowV (bubble.cu:2301-2380) at p->time_s. */
void sfo_airflow(const sfo_params* p, const double* uair0, const double* vair0,
                 double* uair, double* vair);
/* CSR pattern of initLinearSystem (solver.cu:25-60): col[5N]. */
void sfo_col_ind(const sfo_params* p, int32_t* col);
/* concentrationLinearSystemKernel (bubble.cu:1301-1473). gam is updated in
 * place for broken cells, as the reference does. val[5N], rhs[N]. */
void sfo_assemble(const sfo_params* p, const double* uth, const double* uph,
                  double* gam, const double* eta, const double* uair,
                  const double* vair, double* val, double* rhs);
/* The VISCOUS stages: divergenceKernel (bubble.cu:1217-1269), div[N] at the
 * centres from u; computeforcevelthetaKernel (bubble.cu:1511-1646), fv on
 * the u_theta faces, and computeforcevelphiKernel_viscous (bubble.cu:
 * 1649-1816), fu on the u_phi faces, from u, eta and that div. With
 * p->viscous, sfo_assemble (with gravity) and sfo_update compute them from
 * their velocity and thickness inputs. */
void sfo_visc_div(const sfo_params* p, const double* uth, const double* uph, double* div);
void sfo_visc_forces(const sfo_params* p, const double* uth, const double* uph,
                     const double* eta, const double* div, double* fu, double* fv);
/* Solves the assembled system (replacing AmgX PCGF, bubble.cu:1479-1508).
 * x holds x0 on entry and the solution on exit. Returns 0 if converged. */
int sfo_solve(const sfo_params* p, const double* val, const double* rhs,
              double* x, sfo_stats* st);
/* y = A x for the assembled CSR system. */
void sfo_spmv(const sfo_params* p, const double* val, const double* x, double* y);
/* divergenceKernel_fromGamma + applyforceveltheta/phi + applyforceThickness
 * (bubble.cu:1276-1295, 1819-1935, 2041-2063). gam_next is clamped in place. */
void sfo_update(const sfo_params* p, const double* uth, const double* uph,
                const double* eta, const double* gam, double* gam_next,
                const double* uair, const double* vair, double* uth_out,
                double* uph_out, double* eta_out);
/* One Solver::stepForward (solver.cu:260-292): advection then bodyforce.
 * State arrays are updated in place. Returns 0 on success, 1 if the solve
 * did not converge (the reference then sets `broken`, bubble.cu:1500-1507). */
int sfo_step(const sfo_params* p, double* uth, double* uph, double* eta,
             double* gam, const double* uair0, const double* vair0, sfo_stats* st);
/* Film volume sum eta_ij sin(theta_j) h^2. */
double sfo_volume(const sfo_params* p, const double* eta);
/* Number of OpenMP threads the oracle uses. */
int sfo_num_threads(void);

/* ---- BIMOCQ scalar advection (header.cuh:97-98 BIMOCQ, MULTISAMPLE) ----
 * Caller-owned arrays of N = nTheta nPhi: forward map fwd (initial ->
 * current positions), backward maps bwd and bwdprev (current -> initial, and
 * the map of the previous reinitialisation interval), and for eta and Gamma
 * the initial fields and accumulated changes of the current and previous
 * interval (Quantity::getGPUInit/Delta/InitLast/DeltaLast). */
typedef struct {
    double *fwd_t, *fwd_p, *bwd_t, *bwd_p, *bwdprev_t, *bwdprev_p;
    double *eta_init, *eta_initlast, *eta_delta, *eta_deltalast;
    double *gam_init, *gam_initlast, *gam_delta, *gam_deltalast;
    double blend;            /* blendCoeff (config.txt)                     */
    double threshold;        /* reinitialisation distortion, cells (4)     */
    int32_t multisample;     /* 1: five samples per point                  */
    int32_t reinit_count;    /* out: reinitialisations so far              */
    double last_distortion;  /* out: distortion after the last step        */
} sfo_bimocq;

void sfo_lerp_coords(const sfo_params* p, const double* from2, const double* to2, double alpha,
                     double* out2);
void sfo_sample_mapping(const sfo_params* p, const double* map_t, const double* map_p,
                        double rx, double ry, double* out2);
void sfo_map_update(const sfo_params* p, const double* uth, const double* uph, double dt,
                    const double* map_t, const double* map_p, double* out_t, double* out_p);
void sfo_bimocq_init(const sfo_params* p, const double* eta, const double* gam, sfo_bimocq* b);
void sfo_advect_bimocq(const sfo_params* p, const double* eta, const sfo_bimocq* b,
                       double* eta_out);
void sfo_accumulate(const sfo_params* p, const double* fwd_t, const double* fwd_p,
                    const double* change, double* delta, int multisample);
double sfo_distortion(const sfo_params* p, const sfo_bimocq* b);
void sfo_bimocq_reinit(const sfo_params* p, double* eta, double* gam, sfo_bimocq* b);
int sfo_bimocq_step(const sfo_params* p, double* uth, double* uph, double* eta, double* gam,
                    const double* uair0, const double* vair0, sfo_bimocq* b, sfo_stats* st);

#ifdef __cplusplus
}
#endif
#endif

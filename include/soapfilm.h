/*
 * soapfilm.h - C ABI of the MI355X soap-film step (libsoapfilm.so).
 *
 * Drop-in for the time step of RiverIntheSky/SoapBubble: the Solver surface
 * (include/solver.cuh:107-150) and the AmgX solve behind it (bubble.cu:
 * 1479-1508). The reference has no FFI of its own; these entry points are
 * what its C++ Solver would bind to in place of its private CUDA stages and
 * the AmgX C API. Every function returns SF_OK (0) or an sf_status code;
 * nothing in the library exits the process (the reference's checkCudaErrors
 * exits, helper_cuda.h:981-995).
 *
 * Fields cross the boundary as FP64 host arrays in the reference's index
 * convention: row-major by theta, phi contiguous, nPhi = 2 nTheta,
 *   SF_FIELD_UTHETA : (nTheta-1) x nPhi   at theta = (j+1)h,   phi = i h
 *   SF_FIELD_UPHI   :  nTheta    x nPhi   at theta = (j+1/2)h, phi = (i-1/2)h
 *   SF_FIELD_ETA, SF_FIELD_GAMMA : nTheta x nPhi at theta = (j+1/2)h, phi = i h
 * (header.cuh:24-35, 72-74; solver.cu:96-103). The caller owns the host
 * buffers; the library copies.
 */
#ifndef SOAPFILM_H
#define SOAPFILM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sf_ctx sf_ctx;

typedef enum {
    SF_OK = 0,
    SF_ERR_INVALID = 1,       /* bad argument                                 */
    SF_ERR_HIP = 2,           /* HIP runtime error (message in sf_last_error)  */
    SF_ERR_UNSUPPORTED = 3,   /* configuration outside what the build supports */
    SF_ERR_NOMEM = 4,         /* device allocation failed                      */
    SF_ERR_NOT_CONVERGED = 5, /* a solve missed its tolerance (bubble.cu:1500) */
    SF_ERR_ORDER = 6          /* stage called out of order                     */
} sf_status;

typedef enum {
    SF_PREC_FP64 = 0,     /* FP64 fields, FP64 PCG: the plain path             */
    SF_PREC_MIXED = 1,    /* FP32 fields, FP32 PCG inside FP64 refinement      */
    SF_PREC_FP32_ADV = 2  /* FP32 fields, FP64 PCG inside FP64 refinement      */
} sf_precision;

typedef enum {
    SF_FIELD_UTHETA = 0,
    SF_FIELD_UPHI = 1,
    SF_FIELD_ETA = 2,
    SF_FIELD_GAMMA = 3,
    SF_FIELD_UAIR0 = 4,       /* uair_init on u_phi faces (solver.cu:82)          */
    SF_FIELD_VAIR0 = 5,       /* vair_init on u_theta faces (solver.cu:83)        */
    /* intermediates of the current step (read-only): */
    SF_FIELD_UTHETA_STAR = 6, /* after advection                                 */
    SF_FIELD_UPHI_STAR = 7,
    SF_FIELD_ETA_STAR = 8,
    SF_FIELD_GAMMA_STAR = 9,  /* after advection and the broken-cell override    */
    SF_FIELD_RHS = 10,        /* assembled right-hand side                       */
    SF_FIELD_DIAG = 11,       /* diagonal of the assembled operator              */
    SF_FIELD_WNORTH = 12,     /* -A(j,i; j-1,i): north-face weight               */
    SF_FIELD_WWEST = 13,      /* -A(j,i; j,i-1): west-face weight                */
    SF_FIELD_SOLUTION = 14    /* solution of the implicit solve (Gamma n+1)      */
} sf_field;

/* One bubble: the config.txt inputs of Bubble::Bubble (bubble.cu:2568-2580,
 * main.cu:20-36) that reach the step, plus the switches of header.cuh:86-98
 * as run-time flags. */
typedef struct {
    int32_t n_theta;      /* numberOfGridsInTheta; nPhi = 2 n_theta; 4..8192, the row
                             splitting into at most 1024 equal parts of 2, 4, 8
                             or 16 cells (every n_theta up to 1024)          */
    int32_t gravity;      /* header.cuh:94 `gravity`                            */
    int32_t evaporation;  /* header.cuh:95 `evaporation`                        */
    int32_t airflow;      /* header.cuh:96 `airflow`                            */
    double radius;        /* m                                                  */
    double H;             /* mean half film thickness, m                        */
    double U;             /* characteristic velocity, m/s (the reference assumes 1) */
    double Gamma_m;       /* surface mean concentration, mol/m^2               */
    double T;             /* temperature, K                                     */
    double dt;            /* timePerStep, s                                     */
    double M_scale;       /* multiplies the derived Marangoni number M (1 = reference) */
} sf_bubble_desc;

/* Derived non-dimensional parameters (bubble.cu:2573-2580, 2627-2638). */
typedef struct {
    double M, Cr, gs, gamma_max, eva, dt, radius, U, h;
    double re;            /* Re^-1 = nu / (U radius) (bubble.cu:2575), VISCOUS only */
} sf_derived;

typedef struct {
    int32_t precision;    /* sf_precision                                        */
    int32_t device;       /* HIP device ordinal (Solver(..., device), solver.cu:72) */
    int32_t use_graph;    /* capture sf_step in a hipGraph (mixed paths)          */
    int32_t tol_absolute; /* 1: stop on |r| <= tol (PCGF_F.json "ABSOLUTE")       */
    double tol;           /* solver tolerance; 0 -> 1e-10 relative                */
    int32_t max_outer;    /* refinement iterations; 0 -> automatic                */
    int32_t max_inner;    /* inner PCG iterations per refinement; 0 -> automatic  */
    double inner_rho;     /* inner residual reduction; 0 -> automatic             */
    int32_t advance_time; /* 1: air-flow time advances by dt every step           */
    int32_t inner_method; /* inner solve: 0 automatic, 1 PCG with the phi-line
                             Jacobi preconditioner, 2 Chebyshev, 3 PCG with the
                             geometric-multigrid preconditioner (phi-line
                             zebra relaxation, semi-coarsening in theta)     */
    int32_t scalar_advection; /* 0 plain (advectionAllCentered), 1 BIMOCQ for eta
                                 (header.cuh:97 BIMOCQ, bubble.cu:1188-1205)      */
    int32_t multisample;  /* BIMOCQ: 1 five samples per point (header.cuh:98)     */
    double blend;         /* BIMOCQ: blendCoeff of config.txt (bubble.cu:1008-1011) */
    double reinit_threshold; /* BIMOCQ: reinitialise the maps when their
                                distortion exceeds this many cells; 0 -> 4
                                (solver.cu:279-286)                          */
    int32_t initial_guess; /* first iterate of the Gamma solve: 0 automatic
                              (2 for the mixed paths, 1 for FP64), 1 Gamma*
                              as the reference, 2 Gamma* plus the previous
                              two solves' corrections extrapolated linearly */
    int32_t viscous;      /* 1: the VISCOUS stress path (header.cuh:99, off
                             in the reference): divergenceKernel and the
                             viscous forces enter the assembly (with
                             gravity) and the velocity updates
                             (bubble.cu:1217-1269, 1458-1464, 1511-2038)  */
    int32_t halt_on_failure; /* 1: a bubble whose solve misses its tolerance
                             takes no further step (its kernels return at
                             once; the reference sets `broken` and its frame
                             loop stops, bubble.cu:1500-1507, 2669), so a
                             run of several steps, captured or not, ends
                             where the reference's would; 0: keep stepping,
                             counting the failure                         */
    int32_t autotune;     /* 1: at the start of the first sf_step call of a
                             captured context after 16 steps (the flow has
                             developed), time the advection's cell tiles (256, 128 and 64
                             columns of 256 threads) and the Chebyshev launch
                             at C, 2C and 4C cells a thread per row width (C
                             the assembly's) on the context's own data, and
                             keep the fastest (sf_get_tuning reports them);
                             0: the fixed choices (256 columns; 8 cells a
                             thread for rows of >= 2048 cells, else C)     */
} sf_options;

typedef struct {
    int32_t iterations;   /* inner PCG iterations of the last solve              */
    int32_t outer;        /* FP64 residual evaluations of the last solve         */
    int32_t converged;    /* 1 if the last solve met the tolerance               */
    int32_t failed_solves;/* solves that missed the tolerance since creation     */
    double residual;      /* |b - A x| / |b| (or |b - A x|) in FP64, last solve   */
    double max_residual;  /* largest final residual since creation               */
    double volume;        /* film volume sum eta sin(theta) h^2                  */
    int32_t broken;       /* 1 if no cell has eta > 0, or a solve failed         */
    int32_t steps;        /* steps taken                                         */
    double time;          /* air-flow time, s                                    */
    double distortion;    /* BIMOCQ: map distortion after the last step, cells   */
    int32_t reinits;      /* BIMOCQ: map reinitialisations so far                */
    int32_t no_film;      /* 1 if no cell has eta > 0 (the broken test of
                             write_thickness_img, solver.cu:568-575)          */
} sf_bubble_stats;

/* The autotuner's choices (sf_options.autotune). */
typedef struct {
    int32_t tuned;             /* 1 once the autotuner has run                    */
    int32_t advect_tile_width; /* columns of an advection tile (256 threads)      */
    int32_t nwidths;           /* row widths of the context (at most 8 reported)  */
    int32_t pad_;
    int32_t row_width[8];      /* nPhi of each row width                          */
    int32_t inner_cells[8];    /* cells a thread of its Chebyshev launch          */
    double advect_ms[3];       /* advection at tile widths 256, 128, 64 (0: not timed) */
    double inner_ms[8][3];     /* Chebyshev launch at C, 2C, 4C cells a thread    */
} sf_tuning;

typedef struct {
    double advect_ms, assemble_ms, solve_ms, update_ms, step_ms;
    int32_t steps;        /* steps the timings accumulate over                   */
    int32_t pad_;
} sf_timing;

/* Replaces Solver::Solver (solver.cu:63-160): allocates the packed state of
 * nbubbles bubbles on the device. Fields start as the reference's
 * initWithConst values (u = 0, eta = 0.5, Gamma = 1, air = 0). */
int sf_create(const sf_options* opt, int32_t nbubbles, const sf_bubble_desc* bubbles,
              sf_ctx** out);
/* Replaces Solver::~Solver (solver.cu:163-216). */
int sf_destroy(sf_ctx* ctx);

/* Host <-> device copies of one bubble's field; replace Quantity::copyToGPU /
 * copyBackToCPU and Solver::copyToCPU (quantity.cu:3-17, solver.cu:243-249). */
int sf_upload_field(sf_ctx* ctx, int32_t bubble, int32_t field, const double* host);
int sf_download_field(sf_ctx* ctx, int32_t bubble, int32_t field, double* host);

/* currentTimeGlobal (bubble.cu:12, 2662): the air-flow time in seconds;
 * bubble = -1 sets every bubble. */
int sf_set_time(sf_ctx* ctx, int32_t bubble, double t_seconds);
/* Overrides the step of one bubble (or -1: all) like Solver::stepForward(dt)
 * with the timeStepGlobal upload (solver.cu:252-257, bubble.cu:2673-2677);
 * dt in seconds. */
int sf_set_dt(sf_ctx* ctx, int32_t bubble, double dt_seconds);

/* Solver::stepForward (solver.cu:260-292), plain branch: advection then
 * bodyforce, nsteps times, on the context's stream. */
int sf_step(sf_ctx* ctx, int32_t nsteps);

/* Stage-level entry points, in this order, for one step:
 * Solver::advection (bubble.cu:1171-1214); the air flow and assembly of
 * Solver::bodyforce (bubble.cu:2389-2432); the AmgX solve
 * (AlgebraicMultiGridCG, bubble.cu:1479-1508); the post-solve updates
 * (bubble.cu:2444-2506). */
int sf_advect(sf_ctx* ctx);
int sf_assemble(sf_ctx* ctx);
int sf_solve(sf_ctx* ctx);
int sf_update(sf_ctx* ctx);

int sf_synchronize(sf_ctx* ctx);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): nsteps steps of Solver::stepForward (plain scalar
 * advection) of one bubble whose state already lives on the device, like
 * sf_step on a one-bubble SF_PREC_MIXED context, without a context.
 * Allocates nothing: every work array of the step is carved from a fixed
 * 48 MiB device array inside the library (held by each GPU the library runs
 * on until the process ends), so the bubble has at most 2^18 cells
 * (n_theta <= 256 for the power-of-two grids); larger ones return
 * SF_ERR_UNSUPPORTED. Fields are device arrays of float in the layout above
 * (u_theta (nTheta-1) x nPhi, the others nTheta x nPhi; uair0 / vair0 the
 * air flow at t = 0, as SF_FIELD_UAIR0 / SF_FIELD_VAIR0); outputs may alias
 * the inputs. opt->precision must be SF_PREC_MIXED and scalar_advection 0;
 * the solve runs a fixed schedule of opt->max_outer refinements (0: 2) of
 * opt->max_inner inner iterations (0: 16) plus one spare refinement, and a
 * bubble that has converged skips the rest; solves that miss the tolerance
 * are counted in stats->failed_solves (stats may be NULL). t_seconds is the
 * air-flow time at the first step. Work is queued on `stream` (a
 * hipStream_t; NULL: the default stream) after what the caller queued there;
 * the call returns when it has finished. The array is shared by every call
 * on a GPU and keeps the last call's contents: calls are serialised by a
 * lock inside the library, and nothing else may use the array. */
int sf_step_noalloc(const sf_options* opt, const sf_bubble_desc* bubble, double t_seconds,
                    int32_t nsteps, void* stream, const float* u_theta, const float* u_phi,
                    const float* eta, const float* gamma, const float* uair0, const float* vair0,
                    float* u_theta_out, float* u_phi_out, float* eta_out, float* gamma_out,
                    sf_bubble_stats* stats);

/* Average device time of one advection launch over nlaunch back-to-back
 * launches between two events on the context's stream (the launches
 * recompute the advected fields from the unchanged state; call at the start
 * of a step). A measurement aid with no counterpart in the reference. */
int sf_time_advection(sf_ctx* ctx, int32_t nlaunch, double* avg_ms);
/* Average device time of one inner (Chebyshev) iteration launch over
 * nlaunch back-to-back launches between two events on the context's stream.
 * Call between sf_assemble and sf_solve; the solve that follows is
 * unchanged. A measurement aid with no counterpart in the reference. */
int sf_time_inner(sf_ctx* ctx, int32_t nlaunch, double* avg_ms);
/* Solver statistics and Solver::isBroken (solver.cu:296-298, 568-574). */
int sf_get_stats(sf_ctx* ctx, int32_t bubble, sf_bubble_stats* out);
/* The autotuner's choices and timings (a measurement aid with no
 * counterpart in the reference). */
int sf_get_tuning(sf_ctx* ctx, sf_tuning* out);
/* Device time per stage, accumulated since creation or the last reset. */
int sf_get_timing(sf_ctx* ctx, sf_timing* out);
int sf_reset_timing(sf_ctx* ctx);
/* Message of the last failure on ctx (or of the last sf_create when ctx is
 * NULL). */
const char* sf_last_error(const sf_ctx* ctx);

/* Parameter derivation of Bubble::Bubble (host only, no device needed). */
int sf_derive_params(const sf_bubble_desc* desc, sf_derived* out);

/* Deterministic synthetic initial fields (host only): eta = 1 + 0.1 f_s,
 * Gamma = 1 + 0.05 f_{s+1}, f = sum of 16 seeded real spherical harmonics of
 * degree <= l_max normalised to max |f| = 1, plus noise * U[-1,1) on eta;
 * air velocity from the stream function f_{s+2} (degree <= l_air), scaled to
 * max speed air_amp. Replaces the texture init of initThicknessfromPic /
 * initAirflowfromPic (initializer.cu:50-135). Arrays sized as the fields. */
int sf_synth_fields(int32_t n_theta, uint64_t seed, int32_t l_max, int32_t l_air, double noise,
                    double air_amp, double* eta, double* gamma, double* uair0, double* vair0);
/* initThicknessfromPic / initAirflowfromPic (initializer.cu:50-135), host
 * only: an uncompressed HALF/FLOAT EXR texture (channels B, G, R) resized to
 * nPhi x nTheta by bilinear interpolation with aligned pixel centres and
 * replicated borders (OpenCV INTER_LINEAR semantics); eta = (B - 0.5) 0.3 +
 * 0.5; uair0 = (G_i + G_{i-1}) / 2 * radius; vair0 = -(R_j + R_{j+1}) / 2 *
 * radius. As in the reference, a missing or unreadable file leaves the
 * output as it was; *loaded gets bit 0 for the thickness and bit 1 for the
 * air texture actually read. Any output pointer may be NULL. */
int sf_texture_fields(const char* thickness_exr, const char* air_exr, int32_t n_theta,
                      double radius, double* eta, double* uair0, double* vair0, int32_t* loaded);
/* Real spherical harmonic Y_lm(theta, phi), orthonormal on the unit sphere. */
double sf_real_ylm(int32_t l, int32_t m, double theta, double phi);

/* write_thickness_img (solver.cu:541-593): eta * H * 2e5 as a HALF RGB EXR
 * (channels B, G, R, uncompressed). */
int sf_write_thickness_exr(sf_ctx* ctx, int32_t bubble, const char* path);
/* The pixels of write_thickness_img's frame (solver.cu:560-582) as binary16
 * values: eta H 2e5 rounded to binary32, then to binary16 as the reference's
 * EXR writer rounds (nearest, ties away from zero), converted on the device;
 * host holds n_theta x nPhi values (pixel (x = i, y = j) of the nPhi x
 * n_theta image at host[j nPhi + i]). With sf_write_half_exr it writes the
 * same file as sf_write_thickness_exr, and lets a caller write the file on
 * another thread while the next steps run (the reference writes it
 * synchronously). */
int sf_thickness_half(sf_ctx* ctx, int32_t bubble, uint16_t* host);
/* sf_write_image_exr of a gray image given as its binary16 values (row-major,
 * width x height; host only): the bytes sf_write_image_exr writes when its
 * three planes are the gray plane whose halves these are. */
int sf_write_half_exr(const char* path, int32_t width, int32_t height, const uint16_t* gray);
/* Solver::write_image (solver.cu:338-385), host only: width x height float
 * planes r, g, b (row-major) written as an uncompressed scanline OpenEXR with
 * HALF channels B, G, R, byte for byte as the reference's tinyexr call
 * writes it (its header attributes, and its float-to-half rounding: nearest,
 * ties away from zero). */
int sf_write_image_exr(const char* path, int32_t width, int32_t height, const float* r,
                       const float* g, const float* b);

#ifdef __cplusplus
}
#endif
#endif

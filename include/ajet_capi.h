/*
 * ajet_capi.h -- C ABI of the MI355X (gfx950) RK4 time step of the axisymmetric jet solver.
 *
 * This is the drop-in boundary for the reference's hot path. The reference drives that path
 * from PROGRAM ModSolver with two argument-less module procedures over global module state:
 *     DO iteration = ...; CALL ComputeTimeStep; CALL RK4; ENDDO     (ModSolver.f90:122-125)
 * The Fortran host of this project keeps those entry points (ComputeTimeStep, RK4) and
 * implements their bodies with the calls below through ISO_C_BINDING.  Nothing here uses
 * torch types: plain pointers, sizes and POD structs only.
 *
 * Conventions
 *   - Host arrays are Fortran column-major.  A "field" is (nz, nr) with z (index i) fastest;
 *     a "state" is (nz, nr, 5) holding the primitives rho, Vr, Vtheta, Vz, T in that order
 *     (the pv(:,:,1,1:5) array of ModuleVariables.f90:39).
 *   - Return value: 0 on success, a negative AJET_E* code for a library error, a positive
 *     value for a hipError_t.  ajet_last_error() gives the message.
 *   - One context owns its device buffers and one HIP stream.  Calls on one context are not
 *     thread-safe.  Work is stream ordered; only ajet_get_* and ajet_compute_dt synchronise.
 */
#ifndef AJET_CAPI_H
#define AJET_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  AJET_OK = 0,
  AJET_EINVAL = -1,        /* bad argument (null pointer, size, flag) */
  AJET_EGRID = -2,         /* grid is not a tensor-product (z=z(i), r=r(j)) grid */
  AJET_EPATCH = -3,        /* a bc line gives a patch outside the grid or an unusable layout */
  AJET_ENODEV = -4,        /* no usable gfx950 device */
  AJET_EBLOWUP = -5,       /* time step blow-up: every point blanked (ModInputOutput.f90:424-428) */
  AJET_ENONFINITE = -6,    /* non-finite value met in the time-step reduction */
  AJET_ETILE = -7,         /* a boundary closure does not fit the kernel tiling */
  AJET_EUNSUPPORTED = -8   /* feature not available in this build */
};

/* Boundary-condition codes of the bc file (ModuleVariables.f90:112-117). */
enum {
  AJET_BC_AXIS = 90,
  AJET_BC_SAT_FARFIELD = 91,
  AJET_BC_NSCBC_NONREFLECTING = 92,
  AJET_BC_SAT_NOSLIP_WALL = 93,
  AJET_BC_NSCBC_NOSLIP_WALL = 94,
  AJET_BC_SPONGE = 95
};

/* Solver constants.  The reference keeps these as module variables and parameters
 * (ModuleVariables.f90:12-13, 66, 71, 73, 152, 164, 177); ajet_params_default() fills the
 * reference's values. */
typedef struct ajet_params {
  int32_t nz, nr;            /* grid points along z (i) and r (j) */
  int32_t viscous;           /* do_viscous_terms */
  int32_t shock_capturing;   /* ShockCapturing: localized artificial diffusivity */
  int32_t forcing;           /* ControlForcing: 0 = off (the reference default), else AJET_FORCING_* bits */
  int32_t reserved;
  double gam, Re, Pr, CFL;
  double C_mu, C_beta, C_kappa;
  double sponge_start_top, sponge_A_top, sponge_end_left, sponge_A_left;
} ajet_params;

/* One line of the bc file as ModPatch.f90:49 reads it: grid, bc, bc_dir, is, ie, js, je
 * (1-based; -1 stands for the last grid index).  The trailing ks, ke columns are unused. */
typedef struct ajet_bc {
  int64_t grid, bc, bc_dir, i_start, i_end, j_start, j_end;
} ajet_bc;

/* A 4-deep closure patch as ModPatch.f90:264-309 builds it (patch_type,
 * ModuleVariables.f90:102-106), 1-based inclusive. */
typedef struct ajet_patch {
  int64_t i_start, i_end, j_start, j_end, bc, bc_dir;
} ajet_patch;

/* Forcing terms (ajet_params.forcing bits).
 *   AJET_FORCING_FEEDBACK   ControlForcingTerms (ModControlForcing.f90:6-65):
 *       rhs(e) += alpha * F(e,k) * G(z,r) * (q_k - q_target_k), k = 1..5, alpha = 6
 *       (ModuleVariables.f90:184), at every point, after the viscous terms and before the
 *       boundary conditions (ModRHS.f90:223-228);
 *   AJET_FORCING_EXCITATION inflow excitation, an extension of this library (no counterpart in
 *       the reference): rhs(Vz) += amp * Uj * sin(2 pi St Uj (t0 + t)) * G(z,r), with t the
 *       time of the RK stage (time at the step start + c dt, c = 0, 1/2, 1/2, 1) and the
 *       jet diameter as the length unit.
 * G(z,r) = exp(-(z - z0)^2 / lz^2 - (r - r0)^2 / lr^2) in both. */
enum { AJET_FORCING_FEEDBACK = 1, AJET_FORCING_EXCITATION = 2 };

typedef struct ajet_forcing {
  double F[5][5];            /* forcing_mat(e,k) as F[e-1][k-1] (read from Vz_Control_Vr_Feedback.dat) */
  double z0, r0, lz, lr;     /* Gaussian centre and widths */
  double amp, strouhal, uj;  /* excitation: amplitude (fraction of Uj), Strouhal number, jet velocity */
  double t0;                 /* excitation: time at ajet_set_state */
} ajet_forcing;

typedef struct ajet_ctx ajet_ctx;

/* Fill *p with the reference constants for an nz x nr grid. */
void ajet_params_default(ajet_params* p, int32_t nz, int32_t nr);

/* Build the closure patches of a bc file for a single-domain nz x nr grid
 * (CreateAllPatches/PatchCreation, ModPatch.f90:117-309).  Writes at most max_out patches and
 * returns the number built (>= 0) or a negative error code. */
int ajet_build_patches(int32_t nz, int32_t nr, const ajet_bc* lines, int32_t nlines,
                       ajet_patch* out, int32_t max_out);

/* Create a context: grid coordinates z, r and IBLANK are (nz, nr) fields
 * (AssignGridCoordinates, ModInputOutput.f90:69-89); the bc lines replace the bc file.
 * Computes the metrics (ModSolver.f90:57-76, ComputeMetric ModDeriv.f90:241-382) and the
 * per-point patch maps (ModPatch.f90:214-242).  iblank may be NULL (all points active). */
int ajet_create(const ajet_params* p, const double* z, const double* r, const int32_t* iblank,
                const ajet_bc* lines, int32_t nlines, ajet_ctx** out);

int ajet_destroy(ajet_ctx* ctx);
const char* ajet_last_error(const ajet_ctx* ctx);

/* Load the state and the target (SolnRead, ModInputOutput.f90:91-148), both (nz, nr, 5).
 * Resets the stale time-step temperature to 0 as a fresh start or restart does. */
int ajet_set_state(ajet_ctx* ctx, const double* pv, const double* pv_target);
/* Copy the state back (the pv array WriteOutput gathers, ModInputOutput.f90:201-263). */
int ajet_get_state(ajet_ctx* ctx, double* pv);

/* ComputeTimeStep (ModInputOutput.f90:389-438): the CFL time step of the current state,
 * kept on the device for the next ajet_rk4_step and copied to *dt (may be NULL to skip the
 * copy and the synchronisation). */
int ajet_compute_dt(ajet_ctx* ctx, double* dt);

/* RK4 (ModRungeKutta.f90:5-171) without its output: advances the state by the device-side
 * time step of the last ajet_compute_dt (or of the previous step's fused reduction). */
int ajet_rk4_step(ajet_ctx* ctx);

/* One of the four RK stages of ajet_rk4_step (stage = 1..4), launched alone so that a caller
 * can time the stage kernel; calling stages 1, 2, 3, 4 in order is one ajet_rk4_step. */
int ajet_rk4_stage(ajet_ctx* ctx, int32_t stage);

/* nsteps iterations of {ComputeTimeStep; RK4} (ModSolver.f90:122-125).  The time step of
 * each iteration is reduced on the device inside the previous step's last stage; the
 * sequence is appended to the device time-step log (see ajet_get_dt_log). */
int ajet_run(ajet_ctx* ctx, int32_t nsteps);

/* Time steps and accumulated time since ajet_set_state (ModRungeKutta.f90:152-155).
 * Copies min(n, logged, AJET_DTLOG_CAP) entries to dts (may be NULL) and returns the number of
 * steps logged; the device log holds the first AJET_DTLOG_CAP of them (the time keeps adding up). */
enum { AJET_DTLOG_CAP = 1 << 20 };
int ajet_get_dt_log(ajet_ctx* ctx, double* dts, int32_t n, double* time);

/* Test hook: one ComputeRHS (ModRHS.f90:5-258) of the state q (nz, nr, 5), with AssignValues
 * (ModAssign.f90:5-29) before it; writes rhs (nz, nr, 5).  With shock capturing it takes the
 * reference's LAD coefficients of q, in the learned mode too.  Does not change the context state. */
int ajet_rhs(ajet_ctx* ctx, const double* q, double* rhs);

/* Test hook: the artificial transport coefficients mu*, beta*, kappa* that
 * ComputeArtTransCoeffs (ModShockCapturing.f90:5-106) adds for the state q; each (nz, nr). */
int ajet_lad_coeffs(ajet_ctx* ctx, const double* q, double* mustar, double* betastar,
                    double* kappastar);

/* Forcing parameters of a context created with ajet_params.forcing != 0 (the values
 * ControlForcingTerms reads from Vz_Control_Vr_Feedback.dat, ModControlForcing.f90:13-31, and
 * the excitation's).  Required before the first step of such a context. */
int ajet_set_forcing(ajet_ctx* ctx, const ajet_forcing* f);

/* Ensemble: several contexts advanced together -- an extension with no counterpart in the
 * reference, which runs one jet per job.  Every kernel of an RK stage is one launch over the
 * tiles of all runs, and the steps after the first replay a captured HIP graph.  Each run keeps
 * its own grid, state, time step, time, forcing and time-step log, and ends exactly as
 * ajet_run on that context alone would leave it (bit for bit).  All contexts on one device,
 * the same do_viscous_terms and shock capturing, states loaded and, with forcing, its
 * parameters set; all must have taken the same number of steps modulo 2.  With shock capturing
 * (non-zero C_mu, C_beta, C_kappa) each run's artificial coefficients -- the reference's LAD or
 * the learned ones of ajet_set_learned_lad -- are its own launches before each batched stage,
 * and steps are launched directly rather than replayed.  The contexts must outlive the batch;
 * between ajet_batch_run calls they can be used on their own. */
typedef struct ajet_batch ajet_batch;
int ajet_batch_create(ajet_ctx* const* runs, int32_t nruns, ajet_batch** out);
int ajet_batch_run(ajet_batch* b, int32_t nsteps);
int ajet_batch_destroy(ajet_batch* b);
const char* ajet_batch_last_error(const ajet_batch* b);

/* ---- Learned LAD coefficient (an extension with no counterpart in the reference).
 * A multilayer perceptron stands in for the costly part of ComputeArtTransCoeffs
 * (ModShockCapturing.f90:5-106): from a 5 x 5 stencil of rho, Vr, Vz, T around a point (the
 * centre as its value, the 24 neighbours as differences from it, indices clamped to the grid)
 * and dz(i), dr(j), ramp(i), r(j) -- 104 features -- it predicts
 *     beta* = ramp * out_scale[0] * max(y0, 0),  kappa* = ramp * out_scale[1] * max(y1, 0),
 * mu* = 0.  Layers 104 -> 128 -> 128 -> 128 -> 2, GELU (tanh form) after the first three,
 * weights and activations in BF16 with FP32 accumulation on the matrix cores; inputs are
 * normalised as (x - shift) * scale in FP32 first. */
enum { AJET_MLP_NIN = 104, AJET_MLP_HID = 128, AJET_MLP_NOUT = 2 };

typedef struct ajet_mlp {
  float w1[AJET_MLP_HID][AJET_MLP_NIN], b1[AJET_MLP_HID];
  float w2[AJET_MLP_HID][AJET_MLP_HID], b2[AJET_MLP_HID];
  float w3[AJET_MLP_HID][AJET_MLP_HID], b3[AJET_MLP_HID];
  float w4[AJET_MLP_NOUT][AJET_MLP_HID], b4[AJET_MLP_NOUT];
  float shift[AJET_MLP_NIN], scale[AJET_MLP_NIN];
  float out_scale[AJET_MLP_NOUT];
} ajet_mlp;

/* Learned mode of a context with shock capturing on (model NULL: back to the reference's LAD).
 * At the first RK stage of every refresh-th step (refresh >= 1), the network gives beta*,
 * kappa* (mu* = 0) at every point of each 16-point group along z whose ramp reaches ramp_min
 * (zero elsewhere); they are kept until the next refresh.  At the first stage of every step a
 * dilatation sensor s = max(0, -dil) max(dz, dr) / sqrt(T) is taken, and every z/r filter block
 * (a z segment times an r line group) holding a point with s > sensor_threshold and
 * ramp > ramp_min gets the reference's coefficients at each of the step's four stages instead (and
 * the network's values again in the next step whose sensor does not flag it):
 * the LAD kernels run only for those blocks and the tiles that feed them, and not at all in a
 * step that flags none.  The sensor of a state is taken after the step that produced it (same
 * state, same flags) and publishes whether it flagged a block to pinned host memory, so the host
 * knows before the next step's first launch: each step then waits once (polling) for the
 * previous one.  sensor_threshold < 0 reproduces the reference's LAD exactly.  A flagged block's
 * z filter walks its segment in parts, each started 48 columns early from the unfiltered values
 * (within 1.3e-16 of the whole-segment walk): at most AJET_FALLBACK_PARTS (environment, read
 * here: 1, 2, 4 or 8; unset: 8) parts, 1 for sensor_threshold < 0. */
int ajet_set_learned_lad(ajet_ctx* ctx, const ajet_mlp* model, double sensor_threshold, double ramp_min,
                         int32_t refresh);

/* Test and training hooks: the raw features (n x 104, row major) of the points (i, j), 0-based
 * (points holds 2 n int32), of the state q (nz, nr, 5); the network's mu*, beta*, kappa* of
 * the state q (no fallback), each (nz, nr); the fraction of filter blocks the sensor flags
 * for the current state (taken after each step, for the next one). */
int ajet_lad_features(ajet_ctx* ctx, const double* q, const int32_t* points, int32_t n, float* x);
int ajet_lad_learned_coeffs(ajet_ctx* ctx, const double* q, double* mustar, double* betastar,
                            double* kappastar);
int ajet_lad_fallback_fraction(ajet_ctx* ctx, double* fraction);

/* Training on the GPU with the library's own forward and backward kernels: minibatches of
 * `batch` rows (a multiple of 32) drawn from the samples loaded, mean-square loss
 * sum((y - t)^2) / (2 batch), gradients through every layer with BF16 operands and FP32
 * accumulation, torch.optim.Adam's update (beta 0.9, 0.999, eps 1e-8) on FP32 masters. */
typedef struct ajet_trainer ajet_trainer;
/* On failure *out may still hold a trainer (for ajet_trainer_last_error): destroy it. */
int ajet_trainer_create(int32_t batch, const ajet_mlp* init, ajet_trainer** out);
/* samples: x (n x 104, normalised as the model's shift and scale say) and targets t (n x 2) */
int ajet_trainer_load(ajet_trainer* tr, const float* x, const float* t, int64_t n);
/* steps Adam steps at learning rate lr; losses (may be NULL) receives each step's loss.  With
 * exactly `batch` samples loaded every step takes them all (full batch, no sampling). */
int ajet_trainer_run(ajet_trainer* tr, int32_t steps, float lr, uint64_t seed, double* losses);
/* test hooks on n rows (n <= batch, a multiple of 32) of normalised x: the outputs y (n x 2),
 * and the loss with its gradients (written into the w and b members of *grads) */
int ajet_trainer_forward(ajet_trainer* tr, const float* x, int32_t n, float* y);
int ajet_trainer_grads(ajet_trainer* tr, const float* x, const float* t, int32_t n, ajet_mlp* grads,
                       double* loss);
/* the current weights (shift, scale and out_scale as given at creation) */
int ajet_trainer_get(ajet_trainer* tr, ajet_mlp* model);
int ajet_trainer_destroy(ajet_trainer* tr);
const char* ajet_trainer_last_error(const ajet_trainer* tr);

/* Probes (an extension: the reference writes whole solution files only): the points (i, j),
 * 0-based, whose primitives ajet_probe copies out -- values[5 k + v] = pv(i_k, j_k, v) of the
 * current state -- e.g. pressure histories for far-field acoustics. */
int ajet_set_probes(ajet_ctx* ctx, const int32_t* points, int32_t n);
int ajet_probe(ajet_ctx* ctx, double* values);

/* Probes without device-memory allocation -- no counterpart in the reference; for callers that
 * cannot allocate device memory (tight sampling loops, harnesses that forbid hipMalloc): the
 * values ajet_set_probes(points, n) followed by ajet_probe would give, values[5 k + v] =
 * pv(i_k, j_k, v) of the current state, with the n points (0-based (i, j) pairs) and the values
 * passed through a fixed scratch area of AJET_PROBE_AT_MAX points that the library holds on each
 * GPU for the life of the process.  n > AJET_PROBE_AT_MAX is AJET_EINVAL.  Every call on one GPU
 * uses the same scratch area: calls that overlap on one GPU (contexts on other streams, other
 * host threads) race.  Leaves the points of ajet_set_probes as they are.  Synchronises. */
enum { AJET_PROBE_AT_MAX = 4096 };
int ajet_probe_at(ajet_ctx* ctx, const int32_t* points, int32_t n, double* values);

/* The HIP stream the context launches on (a hipStream_t), for callers that time it. */
void* ajet_stream(ajet_ctx* ctx);

/* Library version string. */
const char* ajet_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AJET_CAPI_H */

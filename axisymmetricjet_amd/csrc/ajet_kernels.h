// Kernel argument blocks and launchers (ajet_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "ajet_internal.h"
#include "ajet_physics.h"

namespace ajet {

// Forcing terms (ControlForcingTerms ModControlForcing.f90:6-65, and the inflow excitation).
struct ForcingDev {
  double aF[5][5];         // alpha * forcing_mat(e,k), alpha = 6 (ModuleVariables.f90:184)
  double z0, r0, lz2, lr2; // Gaussian centre, lz*lz, lr*lr
  double au, w, t0;        // excitation: amp*Uj, 2 pi St Uj, time offset
};

struct StageArgs {
  const double* qin;       // stage input state (5 fields), read with a 2-point halo
  const double* qn;        // q^n (stages 2, 3)
  double* qout;            // stage output state
  double* acc_a;           // A = q^n + k1/6
  double* acc_s;           // S = k2 (+ k3)
  double* rhs_out;         // stage 0 only: the right-hand side
  const double* tgt;       // target state
  const uint32_t* meta;    // per-point metadata word
  const double* zc;        // z(i)
  const double* dxidz;     // (nz)
  const double* ddxidxidz;
  const double* rc;        // r(j)
  const double* rinv;      // 1/r(j) (0 on the axis)
  const double* detadr;    // (nr)
  const double* ddetadetadr;
  const double* lad;       // mu*, beta*, kappa* (3 fields) when shock capturing is on
  const PatchDev* patches;
  const uint8_t* seq;
  const int* zcuts;        // tile boundaries along z (ntz + 1 entries)
  const int* rcuts;        // tile boundaries along r (ntr + 1 entries)
  const int* tiles;        // tiles of this launch (id = tri * ntz + tzi)
  const int4* boxes;       // their owned columns [x, y) and rows [z, w) (single-run launches)
  int ntiles;
  StepScalars* sc;
  Geometry g;
  PhysConst pc;
  double* dtlog;           // time-step log (stage 1 appends when dt_from_dmax)
  long long logcap;
  double CFL;
  int dt_from_dmax;        // stage 1: dt = CFL / (max denominator of the last stage 4)
  int forcing;             // AJET_FORCING_* bits
  int parity;              // step parity: the stage time is sc->tpar[parity] + c dt
  ForcingDev fc;
};

struct CflArgs {
  const double* q;
  const uint32_t* meta;
  const double* dxidz;
  const double* detadr;
  StepScalars* sc;
  Geometry g;
  double Re, Pr;
};

// Artificial transport coefficients (ComputeArtTransCoeffs, ModShockCapturing.f90:5-106) of
// the state q, in four launches: strain rate / dilatation / switch, the filtered-quantity
// sources, the in-place z filter, the r filter with the C Re scaling and the ramp.
struct LadArgs {
  const double* q;         // state whose coefficients are wanted (5 fields)
  const uint32_t* meta;
  const uint8_t* codes;    // z row code | r row code << 4, per point (the low byte of meta)
  const double* dxidz;     // (nz)
  const double* detadr;    // (nr)
  const double* rc;        // r(j)
  const double* dzl;       // grid spacing of ComputeTobeFilteredQuantities along z (nz)
  const double* drl;       // ... along r (nr)
  const double* ramp;      // (1 + tanh(5 (z - 10))) / 2 (nz)
  double* A;               // 3 fields: S, dil, f_sw
  double* B;               // 3 fields: the quantities before filtering
  double* out;             // 3 fields: mu*, beta*, kappa*
  const int* zcuts;
  const int* rcuts;
  const int* tiles;        // source tiles: t fused (takes |S| and dil from the velocity), ~t
  int ntiles;              // reads them from the strain kernel (classify_lad_source), in the
                           // order of mixed_tile_list
  const int* str_tiles;    // strain tiles, those the ~t source tiles read: t interior (all
  int nstr;                // points and stencil rows interior), ~t boundary
  const int* rgroups;      // line groups of the fused z/r filter: (wb, out_lo, out_hi) each
  int nrgroups;
  Geometry g;
  double gam;
  double cmu, cbeta, ckappa;  // C_mu Re, C_beta Re, C_kappa Re Pr
  // learned mode: only the flagged z/r filter blocks (segment + nseg * group) and the tiles
  // that feed them run (all run when these are null)
  const uint32_t* zr_flags;
  const uint8_t* src_mask;  // per tile: the sources are needed
  const uint8_t* str_mask;  // per tile: the strain rate and dilatation are needed
  int nseg;                 // filter segments per line: ceil(nz / zseg)
  int zseg;                 // filter segment length in columns, a multiple of 16, at least 64
  uint32_t gen;             // the step's stamp: a block is flagged when its flag equals it
  // where the ramp is exactly zero the coefficients are +-0 and only ever added, so that work is
  // skipped: filter segments below seg_first, source tiles ending at or before column src_imin,
  // strain tiles ending at or before str_imin (the buffer holds zeros there)
  int seg_first, src_imin, str_imin;
  // the z/r filter walks each segment as zsub sub-segments of zseg / zsub columns, each with its
  // own warm-up (1: whole segments).  The learned mode's fallback takes lad_fallback_zsub: its few
  // flagged blocks then finish in a fraction of the time, within 1.3e-16 of the whole-segment walk.
  int zsub;
};

hipError_t upload_tables(const double sbp1[9][6], const double sbp2[9][6], const double tg[9][11], hipStream_t s);
hipError_t upload_lad_tables(const double sbp1[9][6], const double sbp2[9][6], const double tg[9][11], hipStream_t s);
// The three LAD launches (strain, source, fused z/r filter) on stream s.
hipError_t launch_lad(const LadArgs& a, hipStream_t s);
// Resident blocks of the fused z/r filter on device (blocks per CU x CUs); 0 when unknown.
int lad_zrfilter_slots(int device);
// The filter segment length for a grid of nz columns, ramp[0..nz), ngroups line groups and
// slots resident blocks (see ajet_lad.hip).
int lad_zfilter_segment(const double* ramp, int nz, int ngroups, int slots);
// Sub-segments per segment of the learned mode's fallback filter launch: parts (a power of two;
// 0: eight), or the most below it whose length stays a multiple of 16 columns.
int lad_fallback_zsub(int seg, int parts);
void lad_skip_bounds(const double* ramp, int nz, int seg, int& seg_first, int& src_imin, int& str_imin);
// the dilatation sensor of the learned mode: flags[group(j) * nseg + i / seg] = gen where
// max(0, -dil) max(dz, dr) / sqrt(T) > thr and ramp > ramp_min (columns ifirst on, where the ramp
// exceeds ramp_min), and *any = gen when it flags a block; then the publication of its result
// to host-visible memory, host[1] = (*any == gen) and host[0] = gen (stored last, with
// system-scope release); the tile masks of the flagged blocks
hipError_t launch_lad_sensor(const LadArgs& a, uint32_t* flags, uint32_t* any, const int* line_group, double thr,
                             double ramp_min, int ifirst, uint32_t* host, hipStream_t s);
hipError_t launch_lad_masks(const LadArgs& a, const uint32_t* flags, uint8_t* src_mask, uint8_t* str_mask,
                            hipStream_t s);
// dst = src (3 fields) over the output points of every z/r filter block whose flag equals gen
hipError_t launch_lad_restore(const LadArgs& a, const uint32_t* flags, uint32_t gen, const double* src, double* dst,
                              hipStream_t s);

// ---- learned LAD coefficient (ajet_mlp.hip): 104 features -> 128 -> 128 -> 128 -> 2.
constexpr int MLP_NIN = 104;       // features
constexpr int MLP_K = 128;         // padded input width of every layer
constexpr int MLP_H = 128;         // hidden width
constexpr int MLP_NOUT = 2;        // beta*, kappa*
constexpr int MLP_NOUT_PAD = 32;   // output rows as the training GEMM tiles them
enum { MLP_EPI_FWD_GELU = 0, MLP_EPI_FWD_LAST = 1, MLP_EPI_BWD_GELU = 2, MLP_EPI_PARTIAL = 3 };

// C[M x N] = A[M x K] B[N x K]^T for K-slices s = 0 .. splits-1 (A and B advance by s K).
//   FWD_GELU: Z = C + bias (FP32), O = bf16(gelu(Z)) as M x N and OT as N x M
//   FWD_LAST: Z = C + bias
//   BWD_GELU: O = bf16(C gelu'(Zin)) and OT
//   PARTIAL:  Z[s] = C (FP32 split-K partials, M x N each)
struct MlpGemm {
  const __bf16* A;
  int lda;
  const __bf16* B;
  int ldb;
  int M, N, K, splits;
  const float* bias;
  float* Z;
  const float* Zin;
  __bf16* O;
  __bf16* OT;
};

// Adam over one layer's parameters p = [W (N x K), b (N)] (FP32 masters).
struct MlpAdam {
  float* p;
  float* m;
  float* v;
  const float* partial;  // splits x N x K
  const float* db;       // N
  int splits, N, K;
  __bf16* W;             // N x K
  __bf16* WT;            // K x N
  float* b;              // N, the bias the GEMM reads
  float beta2, one_m_beta1, one_m_beta2, step, bc2_sqrt, eps;
  int grad_only;
  float* grad;           // grad_only: the reduced gradient, N x K + N
};

// Inference over the grid (and the feature hook).
struct MlpInfer {
  const double* q;
  Geometry g;
  const double* dzl;
  const double* drl;
  const double* ramp;
  const double* rc;
  const __bf16* W;        // (3 x 128 + 16) x 128: W1 (K padded), W2, W3, W4 (rows >= 2 zero)
  const float* bias;      // b1, b2, b3 (128 each), b4 (2)
  const float* shift;     // 104
  const float* scale;     // 104
  float out_scale[2];
  double ramp_min;        // groups whose ramp stays below this get zero coefficients
  double* out;            // 3 fields: mu* (zero), beta*, kappa*
  double* out2;           // a second copy of out (may be null)
};

hipError_t launch_mlp_gemm(int epi, const MlpGemm& g, hipStream_t s);
hipError_t launch_mlp_loss(const float* y, const float* t, int n, __bf16* dz, __bf16* dzT, double* loss, hipStream_t s);
hipError_t launch_mlp_bias_grad(const __bf16* dzT, int n, int ncols, float* db, hipStream_t s);
hipError_t launch_mlp_adam(const MlpAdam& a, hipStream_t s);
hipError_t launch_mlp_rows(const float* x, const float* t, const long long* idx, int n, __bf16* a0, __bf16* a0T,
                           float* tb, hipStream_t s);
hipError_t launch_mlp_sample(unsigned long long seed, long long step, long long nsamples, int n, long long* idx,
                             hipStream_t s);
hipError_t launch_mlp_features(const MlpInfer& a, const int* pts, int n, float* out, hipStream_t s);
hipError_t launch_mlp_infer(const MlpInfer& a, int nblocks, hipStream_t s);

// Physics variants of the stage kernel.
enum { MODE_PLAIN = 0, MODE_SC = 1, MODE_LAD = 2, MODE_INVISCID = 3 };

hipError_t launch_stage(int stage, int mode, bool interior, const StageArgs& a, hipStream_t s);
// A stage with the interior and boundary tiles in one launch: a.tiles holds the list of
// mixed_tile_list (t interior, ~t boundary, a boundary tile first).  Stages 1-3 of the plain
// mode (three blocks per CU) and of the LAD mode (two).
bool stage_mixed_supported(int stage, int mode);
hipError_t launch_stage_mixed(int stage, int mode, const StageArgs& a, hipStream_t s);
// Batched stages (MODE_PLAIN and MODE_INVISCID): `runs` and `list` in device memory, list
// entries = tile | run << 20 | writer << 31.
hipError_t launch_stage_batch(int stage, int mode, bool interior, bool force, const StageArgs* runs,
                              const uint32_t* list, int n, int dt_from_dmax, int parity, hipStream_t s);
hipError_t launch_cfl(const CflArgs& a, double CFL, int any_blank, double* log, long long cap, hipStream_t s);
hipError_t launch_finalize(StepScalars* sc, double CFL, int any_blank, double* log, long long cap, hipStream_t s);
hipError_t launch_probe(const double* q, const Geometry& g, const int* pts, int n, double* out, hipStream_t s);

// What ajet_probe_at (ajet_noalloc.cpp) needs of a context: its device, stream, current state
// and geometry.  False for a context without a stream.
struct ProbeView {
  int device;
  hipStream_t stream;
  const double* q;
  Geometry g;
};
bool ctx_probe_view(ajet_ctx* c, ProbeView* v);
void ctx_set_error(ajet_ctx* c, const std::string& msg);

}  // namespace ajet

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
  const double* detadr;    // (nr)
  const double* ddetadetadr;
  const double* lad;       // mu*, beta*, kappa* (3 fields) when shock capturing is on
  const PatchDev* patches;
  const uint8_t* seq;
  const int* zcuts;        // tile boundaries along z (ntz + 1 entries)
  const int* rcuts;        // tile boundaries along r (ntr + 1 entries)
  const int* tiles;        // tiles of this launch (id = tri * ntz + tzi)
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
  const int* int_tiles;    // tiles whose points and stencil rows are all interior
  const int* bnd_tiles;    // the others
  int n_int, n_bnd;
  const int* rgroups;      // line groups of the fused z/r filter: (wb, out_lo, out_hi) each
  int nrgroups;
  Geometry g;
  double gam;
  double cmu, cbeta, ckappa;  // C_mu Re, C_beta Re, C_kappa Re Pr
};

hipError_t upload_tables(const double sbp1[9][6], const double sbp2[9][6], const double tg[9][11], hipStream_t s);
hipError_t upload_lad_tables(const double sbp1[9][6], const double sbp2[9][6], const double tg[9][11], hipStream_t s);
hipError_t launch_lad(const LadArgs& a, hipStream_t s);

// Physics variants of the stage kernel.
enum { MODE_PLAIN = 0, MODE_SC = 1, MODE_LAD = 2, MODE_INVISCID = 3 };

hipError_t launch_stage(int stage, int mode, bool interior, const StageArgs& a, hipStream_t s);
// Batched stages (MODE_PLAIN and MODE_INVISCID): `runs` and `list` in device memory, list
// entries = tile | run << 20 | writer << 31.
hipError_t launch_stage_batch(int stage, int mode, bool interior, bool force, const StageArgs* runs,
                              const uint32_t* list, int n, int dt_from_dmax, int parity, hipStream_t s);
hipError_t launch_cfl(const CflArgs& a, double CFL, int any_blank, double* log, long long cap, hipStream_t s);
hipError_t launch_finalize(StepScalars* sc, double CFL, int any_blank, double* log, long long cap, hipStream_t s);

}  // namespace ajet

/*
 * ORACLE -- test infrastructure only.  Nothing in the product links, loads or calls this
 * library: it is the CPU checker that tests/, the smoke check and the benchmark's CPU-baseline
 * leg compare the HIP path against.
 *
 * A single-domain CPU restatement, in plain C, of the reference solver's RK4 time step
 * (ModRungeKutta.f90, ModRHS.f90, ModDeriv.f90, ModOperator.f90, ModAssign.f90, ModAxis.f90,
 * ModBC.f90, ModShockCapturing.f90, ModTGfilter.f90, ModPatch.f90 and ComputeTimeStep in
 * ModInputOutput.f90).  Same loops, same halo'd per-variable arrays, same left-to-right
 * association of every expression, built with -ffp-contract=off so no multiply-add is fused.
 */
#ifndef AJET_ORACLE_H
#define AJET_ORACLE_H

#include "../include/ajet_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc* orc_create(const ajet_params* p, const double* z, const double* r, const int32_t* iblank,
                const ajet_bc* lines, int32_t nlines);
void orc_destroy(orc* o);
int orc_npatch(const orc* o);
int orc_patches(const orc* o, ajet_patch* out, int max_out);
void orc_tables(const orc* o, double* sbp1, double* sbp2, double* tg);
void orc_metrics(const orc* o, double* dxidz, double* detadr, double* ddxidxidz,
                 double* ddetadetadr);
void orc_set_state(orc* o, const double* pv, const double* pv_target);
void orc_get_state(const orc* o, double* pv);
int orc_rhs(orc* o, const double* q, double* rhs);
int orc_compute_dt(orc* o, double* dt);
void orc_set_dt(orc* o, double dt);
int orc_rk4(orc* o);
int orc_run(orc* o, int nsteps, double* dts);
double orc_time(const orc* o);
double orc_stale_T(const orc* o);
void orc_set_stale_T(orc* o, double T);
int orc_lad(orc* o, const double* q, double* mustar, double* betastar, double* kappastar);
int orc_lad_stages(orc* o, const double* q, double* quantities, double* zfiltered,
                   double* rfiltered);
void orc_d1(orc* o, const double* field, int dirn, double* out);
void orc_d2(orc* o, const double* field, int dirn, const double* first, double* out);
void orc_d2comp(orc* o, const double* field, int dirn, double* out);
void orc_tgfilter(orc* o, const double* field, int dirn, int inplace, double* out);
void orc_set_threads(int n);
/* forcing: mode = AJET_FORCING_* bits, F = forcing_mat row-major (e, k) */
void orc_set_forcing(orc* o, int mode, const double* F, double z0, double r0, double lz, double lr, double amp,
                     double st, double uj, double t0);

#ifdef __cplusplus
}
#endif
#endif

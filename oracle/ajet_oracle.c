/*
 * ORACLE -- test infrastructure only (see ajet_oracle.h).  Only tests/, the smoke check and
 * the benchmark's CPU-baseline leg use it, as the checker; the product never does.
 *
 * Single-domain CPU restatement of the reference's RK4 time step.  Every function cites the
 * reference lines it follows.  Arrays mirror ModAllocate.f90: "halo" fields are
 * (-3:Nz+4, -3:Nr+4) column-major, "grid" fields (Nz, Nr), the state (Nz, Nr, 5).
 * Build with -ffp-contract=off: the reference is compiled without optimisation
 * (run_compile.sh:5-7), so no multiply-add is ever fused, and every sum below keeps the
 * Fortran left-to-right association.  Loop nests may run j-outer where each point is
 * independent; the two places where order matters (the in-place z filter and the last point
 * that sets the stale scalar T) keep the reference's order.
 */
#include "ajet_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc {
  int nz, nr, nzh, nrh;
  ajet_params p;
  double SBP1[9][6], SBP2[9][6], TG[9][11];
  ajet_patch* patch;
  int npatch;
  int64_t *is_i, *is_j, *which_i, *which_j; /* (Nz, Nr) */
  double* ibval;                             /* (Nz, Nr) */
  /* halo fields */
  double *zvec, *rvec, *dzdxi, *drdeta, *dxidz, *detadr, *ddxidxidz, *ddetadetadr;
  double *rho, *Vr, *Vt, *Vz, *T, *pr, *mu, *kappa, *lambda, *beta;
  double *drhodz, *dVrdz, *dVtdz, *dVzdz, *dTdz, *dpdz, *dmudz, *dkappadz;
  double *drhodr, *dVrdr, *dVtdr, *dVzdr, *dTdr, *dpdr, *dmudr, *dkappadr;
  double *d2Vrdz2, *d2Vtdz2, *d2Vzdz2, *d2Tdz2, *d2Vrdr2, *d2Vtdr2, *d2Vzdr2, *d2Tdr2;
  double *d2Vrdzdr, *d2Vzdzdr;
  /* LAD halo fields (ModAllocate.f90:84-112) */
  double *S, *d2Sdxi2, *d2Sdeta2, *d4Sdxi4, *d4Sdeta4;
  double *d2Tdxi2, *d2Tdeta2, *d4Tdxi4, *d4Tdeta4;
  double *dil, *d2dildxi2, *d2dildeta2, *d4dildxi4, *d4dildeta4;
  double *muQ, *mustar, *betaQ, *betastar, *kappaQ, *kappastar;
  /* state arrays */
  double *pv, *pv_old, *pv_target, *rhs, *k1, *k2, *k3, *k4, *deltvec;
  double T_scalar; /* the module scalar T of ModuleVariables.f90:53 */
  double delt, time;
  /* forcing (ajet_capi.h AJET_FORCING_*): ControlForcingTerms and the inflow excitation */
  int forcing;
  double fmat[5][5], fz0, fr0, flz, flr, famp, fst, fuj, ft0;
  double t_stage; /* time of the RHS being evaluated */
  int T_set;
};

#define HX(o, i, j) ((size_t)((i) + 3) + (size_t)(o)->nzh * (size_t)((j) + 3))
#define GX(o, i, j) ((size_t)((i)-1) + (size_t)(o)->nz * (size_t)((j)-1))
#define SX(o, i, j, e) (GX(o, i, j) + (size_t)(o)->nz * (size_t)(o)->nr * (size_t)((e)-1))

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static double* halo_alloc(const orc* o) {
  return (double*)calloc((size_t)o->nzh * (size_t)o->nrh, sizeof(double));
}

/* SBP_operator, ModOperator.f90:5-121.  Rows 1-3 of the TG filter are default-REAL literals
 * in the reference (ModOperator.f90:94-97), so they are float values widened to double. */
static void sbp_tables(orc* o) {
  memset(o->SBP1, 0, sizeof o->SBP1);
  memset(o->SBP2, 0, sizeof o->SBP2);
  memset(o->TG, 0, sizeof o->TG);
  double (*A)[6] = o->SBP1;
  A[0][0] = -(24.0 / 17.0); A[0][1] = 59.0 / 34.0; A[0][2] = -(4.0 / 17.0); A[0][3] = -(3.0 / 34.0);
  A[1][0] = -1.0 / 2.0; A[1][1] = 0.0; A[1][2] = 1.0 / 2.0;
  A[2][0] = 4.0 / 43.0; A[2][1] = -(59.0 / 86.0); A[2][2] = 0.0; A[2][3] = 59.0 / 86.0; A[2][4] = -(4.0 / 43.0);
  A[3][0] = 3.0 / 98.0; A[3][1] = 0.0; A[3][2] = -(59.0 / 98.0); A[3][3] = 0.0; A[3][4] = 32.0 / 49.0; A[3][5] = -(4.0 / 49.0);
  for (int r = 5; r <= 8; ++r)
    for (int c = 0; c <= 5; ++c) A[r][c] = -1.0 * A[8 - r][5 - c];
  A[4][0] = 1.0 / 12.0; A[4][1] = -(2.0 / 3.0); A[4][2] = 0.0; A[4][3] = 2.0 / 3.0; A[4][4] = -(1.0 / 12.0);

  double (*B)[6] = o->SBP2;
  B[0][0] = 2.0; B[0][1] = -5.0; B[0][2] = 4.0; B[0][3] = -1.0;
  B[1][0] = 1.0; B[1][1] = -2.0; B[1][2] = 1.0;
  B[2][0] = -(4.0 / 43.0); B[2][1] = 59.0 / 43.0; B[2][2] = -(110.0 / 43.0); B[2][3] = 59.0 / 43.0; B[2][4] = -(4.0 / 43.0);
  B[3][0] = -(1.0 / 49.0); B[3][1] = 0.0; B[3][2] = 59.0 / 49.0; B[3][3] = -(118.0 / 49.0); B[3][4] = 64.0 / 49.0; B[3][5] = -(4.0 / 49.0);
  for (int r = 5; r <= 8; ++r)
    for (int c = 0; c <= 5; ++c) B[r][c] = 1.0 * B[8 - r][5 - c];
  B[4][0] = -(1.0 / 12.0); B[4][1] = 4.0 / 3.0; B[4][2] = -(5.0 / 2.0); B[4][3] = 4.0 / 3.0; B[4][4] = -(1.0 / 12.0);

  double (*F)[11] = o->TG;
  F[0][0] = 1.0;
  static const float r1[7] = {0.085777408970000f, 0.722371828476000f, 0.356848072173000f, -0.223119093072000f,
                              0.057347064865000f, 0.000747264596000f, 0.000027453993000f};
  static const float r2[7] = {-0.032649010764000f, 0.143339502575000f, 0.726678822020000f, 0.294622121167000f,
                              -0.186711738069000f, 0.062038376258000f, -0.007318073189000f};
  static const float r3[11] = {0.000054596010000f, -0.042124772446000f, 0.173103107841000f, 0.700384128648000f,
                               0.276543612935000f, -0.131223506571000f, 0.023424966418000f, -0.013937561779000f,
                               0.024565095706000f, -0.013098287852000f, 0.002308621090000f};
  for (int c = 0; c < 7; ++c) { F[1][c] = (double)r1[c]; F[2][c] = (double)r2[c]; }
  for (int c = 0; c < 11; ++c) F[3][c] = (double)r3[c];
  for (int r = 5; r <= 8; ++r)
    for (int c = 0; c <= 10; ++c) F[r][c] = 1.0 * F[8 - r][10 - c];
  const double tgi[9] = {107.0 / 103680.0, 149.0 / 12960.0, 1997.0 / 25920.0, 3091.0 / 12960.0, 3565.0 / 10368.0,
                         3091.0 / 12960.0, 1997.0 / 25920.0, 149.0 / 12960.0, 107.0 / 103680.0};
  for (int c = 0; c < 9; ++c) F[4][c] = tgi[c];
}

/* DOT_PRODUCT, summed sequentially from the first element (no fused multiply-add). */
static inline double dotn(const double* a, const double* v, size_t stride, int n) {
  double s = 0.0;
  for (int m = 0; m < n; ++m) s = s + a[m] * v[(size_t)m * stride];
  return s;
}

/* CreateAllPatches + PatchCreation for one domain (ModPatch.f90:117-309): each bc line becomes
 * a 4-deep strip clipped to the grid; the last patch in file order owns a point in the maps
 * (ModPatch.f90:214-242).  Returns the number of patches or a negative error. */
static int build_patches(int nz, int nr, const ajet_bc* lines, int nlines, ajet_patch* out, int max_out) {
  int np = 0;
  for (int l = 0; l < nlines; ++l) {
    int64_t is = lines[l].i_start, ie = lines[l].i_end, js = lines[l].j_start, je = lines[l].j_end;
    int64_t dir = lines[l].bc_dir;
    if (is == -1) is = nz;
    if (js == -1) js = nr;
    if (ie == -1) ie = nz;
    if (je == -1) je = nr;
    int64_t pis, pie, pjs, pje;
    if (dir == 1) { pis = is; pie = is + 3; pjs = js; pje = je; }
    else if (dir == -1) { pis = ie - 3; pie = ie; pjs = js; pje = je; }
    else if (dir == 2) { pis = is; pie = ie; pjs = js; pje = js + 3; }
    else if (dir == -2) { pis = is; pie = ie; pjs = je - 3; pje = je; }
    else return AJET_EPATCH;
    /* first and last grid point (i outer, j inner) inside the strip */
    int found = 0;
    int64_t fi = 0, fj = 0, li = 0, lj = 0;
    for (int64_t i = 1; i <= nz; ++i)
      for (int64_t j = 1; j <= nr; ++j) {
        int64_t v1 = (i - pis) * (i - pie), v2 = (j - pjs) * (j - pje);
        if (v1 <= 0 && v2 <= 0) {
          if (!found) { fi = i; fj = j; found = 1; }
          else { li = i; lj = j; }
        }
      }
    if (!found) continue;
    ajet_patch P;
    P.bc = lines[l].bc;
    P.bc_dir = dir;
    if (dir == 1) { P.i_start = fi; P.i_end = fi + 3; P.j_start = fj; P.j_end = lj; }
    else if (dir == -1) { P.i_start = li - 3; P.i_end = li; P.j_start = fj; P.j_end = lj; }
    else if (dir == 2) { P.i_start = fi; P.i_end = li; P.j_start = fj; P.j_end = fj + 3; }
    else { P.i_start = fi; P.i_end = li; P.j_start = lj - 3; P.j_end = lj; }
    if (np < max_out) out[np] = P;
    ++np;
  }
  return np;
}

/* Compute1stDerivative (ModDeriv.f90:5-120): SBP D1 with patch closures, zero at blanked
 * points, then the metric multiply in a second pass. */
static void first_deriv(orc* o, const double* var, int dirn, double* out) {
  memset(out, 0, sizeof(double) * (size_t)o->nzh * o->nrh);
  const int nz = o->nz, nr = o->nr;
  const size_t sj = (size_t)o->nzh;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) {
      size_t g = GX(o, i, j);
      if (o->ibval[g] == 0.0) continue;
      double s;
      if (dirn == 1) {
        if (o->is_i[g] == 1) {
          const ajet_patch* P = &o->patch[o->which_i[g] - 1];
          if (P->bc_dir == 1) s = dotn(o->SBP1[i - P->i_start], &var[HX(o, P->i_start, j)], 1, 6);
          else if (P->bc_dir == -1) s = dotn(o->SBP1[8 + i - P->i_end], &var[HX(o, P->i_end - 5, j)], 1, 6);
          else continue;
        } else s = dotn(o->SBP1[4], &var[HX(o, i - 2, j)], 1, 5);
      } else {
        if (o->is_j[g] == 1) {
          const ajet_patch* P = &o->patch[o->which_j[g] - 1];
          if (P->bc_dir == 2) s = dotn(o->SBP1[j - P->j_start], &var[HX(o, i, P->j_start)], sj, 6);
          else if (P->bc_dir == -2) s = dotn(o->SBP1[8 + j - P->j_end], &var[HX(o, i, P->j_end - 5)], sj, 6);
          else continue;
        } else s = dotn(o->SBP1[4], &var[HX(o, i, j - 2)], sj, 5);
      }
      out[HX(o, i, j)] = s;
    }
  const double* met = dirn == 1 ? o->dxidz : o->detadr;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) out[HX(o, i, j)] = out[HX(o, i, j)] * met[HX(o, i, j)];
}

/* The SBP2 stencil pass shared by Compute2ndDerivative (ModDeriv.f90:123-238) and
 * Compute2ndDerivativeComputational (ModDeriv.f90:387-481). */
static void second_stencil(orc* o, const double* var, int dirn, double* out) {
  memset(out, 0, sizeof(double) * (size_t)o->nzh * o->nrh);
  const int nz = o->nz, nr = o->nr;
  const size_t sj = (size_t)o->nzh;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) {
      size_t g = GX(o, i, j);
      if (o->ibval[g] == 0.0) continue;
      double s;
      if (dirn == 1) {
        if (o->is_i[g] == 1) {
          const ajet_patch* P = &o->patch[o->which_i[g] - 1];
          if (P->bc_dir == 1) s = dotn(o->SBP2[i - P->i_start], &var[HX(o, P->i_start, j)], 1, 6);
          else if (P->bc_dir == -1) s = dotn(o->SBP2[8 + i - P->i_end], &var[HX(o, P->i_end - 5, j)], 1, 6);
          else continue;
        } else s = dotn(o->SBP2[4], &var[HX(o, i - 2, j)], 1, 5);
      } else {
        if (o->is_j[g] == 1) {
          const ajet_patch* P = &o->patch[o->which_j[g] - 1];
          if (P->bc_dir == 2) s = dotn(o->SBP2[j - P->j_start], &var[HX(o, i, P->j_start)], sj, 6);
          else if (P->bc_dir == -2) s = dotn(o->SBP2[8 + j - P->j_end], &var[HX(o, i, P->j_end - 5)], sj, 6);
          else continue;
        } else s = dotn(o->SBP2[4], &var[HX(o, i, j - 2)], sj, 5);
      }
      out[HX(o, i, j)] = s;
    }
}

static void second_deriv(orc* o, const double* var, int dirn, const double* first, double* out) {
  second_stencil(o, var, dirn, out);
  const double* met = dirn == 1 ? o->dxidz : o->detadr;
  const double* dmet = dirn == 1 ? o->ddxidxidz : o->ddetadetadr;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      out[h] = out[h] * (met[h] * met[h]) + first[h] * dmet[h];
    }
}

/* ComputeMetric (ModDeriv.f90:241-382), single domain: every edge is a physical boundary. */
static void metric(orc* o, const double* var, int dirn, double* out) {
  memset(out, 0, sizeof(double) * (size_t)o->nzh * o->nrh);
  const int nz = o->nz, nr = o->nr;
  const size_t sj = (size_t)o->nzh;
  if (dirn == 1) {
    for (int i = 5; i <= nz - 4; ++i)
      for (int j = 1; j <= nr; ++j) out[HX(o, i, j)] = dotn(o->SBP1[4], &var[HX(o, i - 2, j)], 1, 5);
    for (int i = 1; i <= 4; ++i)
      for (int j = 1; j <= nr; ++j) out[HX(o, i, j)] = dotn(o->SBP1[i - 1], &var[HX(o, 1, j)], 1, 6);
    for (int i = nz - 3; i <= nz; ++i)
      for (int j = 1; j <= nr; ++j) out[HX(o, i, j)] = dotn(o->SBP1[8 - nz + i], &var[HX(o, nz - 5, j)], 1, 6);
  } else {
    for (int i = 1; i <= nz; ++i)
      for (int j = 5; j <= nr - 4; ++j) out[HX(o, i, j)] = dotn(o->SBP1[4], &var[HX(o, i, j - 2)], sj, 5);
    for (int i = 1; i <= nz; ++i)
      for (int j = 1; j <= 4; ++j) out[HX(o, i, j)] = dotn(o->SBP1[j - 1], &var[HX(o, i, 1)], sj, 6);
    for (int i = 1; i <= nz; ++i)
      for (int j = nr - 3; j <= nr; ++j) out[HX(o, i, j)] = dotn(o->SBP1[8 - nr + j], &var[HX(o, i, nr - 5)], sj, 6);
  }
}

/* AssignValues (ModAssign.f90:5-29). */
static void assign_values(orc* o) {
  const double c35 = 3.0 / 5.0, c23 = 2.0 / 3.0;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      o->rho[h] = o->pv[SX(o, i, j, 1)];
      o->Vr[h] = o->pv[SX(o, i, j, 2)];
      o->Vt[h] = o->pv[SX(o, i, j, 3)];
      o->Vz[h] = o->pv[SX(o, i, j, 4)];
      o->T[h] = o->pv[SX(o, i, j, 5)];
      o->pr[h] = o->rho[h] * o->T[h];
      o->mu[h] = pow(o->T[h], 0.666);
      o->kappa[h] = o->mu[h];
      o->beta[h] = c35 * o->mu[h];
      o->lambda[h] = o->beta[h] - c23 * o->mu[h];
    }
}

/* axis_conditions (ModAxis.f90:5-35): rho, Vz, T on j=1 from the primitives of the last
 * AssignValues (the stage input), copied into pv. */
static void axis_conditions(orc* o) {
  const double a00 = o->SBP1[0][0];
  const size_t sj = (size_t)o->nzh;
  for (int i = 1; i <= o->nz; ++i) {
    double* f[3] = {o->rho, o->Vz, o->T};
    for (int q = 0; q < 3; ++q) f[q][HX(o, i, 1)] = -1.0 * dotn(&o->SBP1[0][1], &f[q][HX(o, i, 2)], sj, 5) / a00;
  }
  for (int i = 1; i <= o->nz; ++i) {
    o->pv[SX(o, i, 1, 1)] = o->rho[HX(o, i, 1)];
    o->pv[SX(o, i, 1, 4)] = o->Vz[HX(o, i, 1)];
    o->pv[SX(o, i, 1, 5)] = o->T[HX(o, i, 1)];
  }
}

/* ---------------------------------------------------------------- LAD (ModShockCapturing.f90) */

/* PerformTGfilter (ModTGfilter.f90:5-95).  out is not cleared (ModTGfilter.f90:14).  Called
 * with var == out for z (ModShockCapturing.f90:63,69,75), it becomes a causal recursion
 * along i; each r-line is independent, so lines run in parallel and i stays sequential. */
static void tg_filter(orc* o, const double* var, int dirn, double* out) {
  const int nz = o->nz, nr = o->nr;
  const size_t sj = (size_t)o->nzh;
  if (dirn == 1) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nr; ++j)
      for (int i = 1; i <= nz; ++i) {
        size_t g = GX(o, i, j);
        if (o->ibval[g] == 0.0) continue;
        double s;
        if (o->is_i[g] == 1) {
          const ajet_patch* P = &o->patch[o->which_i[g] - 1];
          if (P->bc_dir == 1) s = dotn(o->TG[i - P->i_start], &var[HX(o, P->i_start, j)], 1, 11);
          else if (P->bc_dir == -1) s = dotn(o->TG[8 + i - P->i_end], &var[HX(o, P->i_end - 10, j)], 1, 11);
          else continue;
        } else s = dotn(o->TG[4], &var[HX(o, i - 4, j)], 1, 9);
        out[HX(o, i, j)] = s;
      }
  } else {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nr; ++j)
      for (int i = 1; i <= nz; ++i) {
        size_t g = GX(o, i, j);
        if (o->ibval[g] == 0.0) continue;
        double s;
        if (o->is_j[g] == 1) {
          const ajet_patch* P = &o->patch[o->which_j[g] - 1];
          if (P->bc_dir == 2) s = dotn(o->TG[j - P->j_start], &var[HX(o, i, P->j_start)], sj, 11);
          else if (P->bc_dir == -2) s = dotn(o->TG[8 + j - P->j_end], &var[HX(o, i, P->j_end - 10)], sj, 11);
          else continue;
        } else s = dotn(o->TG[4], &var[HX(o, i, j - 4)], sj, 9);
        out[HX(o, i, j)] = s;
      }
  }
}

/* ComputeDilatation and ComputeMgnStrRate (ModShockCapturing.f90:108-159). */
static void dilatation_strain(orc* o) {
  const double c13 = 1.0 / 3.0;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      if (j == 1) o->dil[h] = 2.0 * o->dVrdr[h] + o->dVzdz[h];
      else o->dil[h] = o->dVrdr[h] + o->Vr[h] / o->rvec[h] + o->dVzdz[h];
    }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      double S11 = o->dVrdr[h] - c13 * o->dil[h];
      double S12 = 0.0;
      double S13 = 0.5 * (o->dVrdz[h] + o->dVzdr[h]);
      double S21 = 0.0, S23 = 0.0, S31 = S13, S32 = 0.0;
      double S33 = o->dVzdz[h] - c13 * o->dil[h];
      double S22 = (j == 1) ? o->dVrdr[h] - c13 * o->dil[h] : o->Vr[h] / o->rvec[h] - c13 * o->dil[h];
      o->S[h] = sqrt(S11 * S11 + S12 * S12 + S13 * S13 + S21 * S21 + S22 * S22 + S23 * S23 + S31 * S31 +
                     S32 * S32 + S33 * S33);
    }
}

/* ComputeTobeFilteredQuantities (ModShockCapturing.f90:162-222).  1e-15 is a default-REAL
 * literal there, so it enters as a float widened to double. */
static void lad_quantities(orc* o) {
  const int nz = o->nz, nr = o->nr;
  const double eps = (double)1e-15f, gam = o->p.gam;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) {
      size_t h = HX(o, i, j);
      double dz = 0.0, dr = 0.0;
      if (i >= 2 && i <= nz - 1) dz = (o->zvec[HX(o, i + 1, j)] - o->zvec[HX(o, i - 1, j)]) / 2.0;
      if (i == 1) dz = o->zvec[HX(o, 2, j)] - o->zvec[HX(o, 1, j)];
      if (i == nz) dz = o->zvec[HX(o, nz, j)] - o->zvec[HX(o, nz - 1, j)];
      if (j >= 2 && j <= nr - 1) dr = (o->rvec[HX(o, i, j + 1)] - o->rvec[HX(o, i, j - 1)]) / 2.0;
      if (j == 1) dr = o->rvec[HX(o, i, 2)] - o->rvec[HX(o, i, 1)];
      if (j == nr) dr = o->rvec[HX(o, i, nr)] - o->rvec[HX(o, i, nr - 1)];
      double d;
      if (j == 1) d = 2.0 * o->dVrdr[h] + o->dVzdz[h];
      else d = o->dVrdr[h] + o->Vr[h] / o->rvec[h] + o->dVzdz[h];
      double vort = fabs(o->dVrdz[h] - o->dVzdr[h]);
      double md = -1.0 * d;
      double H = (md < 0.0) ? 0.0 : 1.0;
      double fsw = H * (d * d) / (d * d + vort * vort + eps);
      o->muQ[h] = o->rho[h] * fabs(o->d4Sdxi4[h] * (dz * dz) + o->d4Sdeta4[h] * (dr * dr));
      o->betaQ[h] = o->rho[h] * fsw * fabs(o->d4dildxi4[h] * (dz * dz) + o->d4dildeta4[h] * (dr * dr));
      o->kappaQ[h] = o->rho[h] / (gam * sqrt(o->T[h])) * fabs(o->d4Tdxi4[h] * dz + o->d4Tdeta4[h] * dr);
    }
}

/* The quantities and their z (in place) and r filters, ModShockCapturing.f90:18-77, without
 * the C*Re scaling. */
static void lad_filtered(orc* o) {
  dilatation_strain(o);
  second_stencil(o, o->S, 1, o->d2Sdxi2);
  second_stencil(o, o->T, 1, o->d2Tdxi2);
  second_stencil(o, o->dil, 1, o->d2dildxi2);
  second_stencil(o, o->S, 2, o->d2Sdeta2);
  second_stencil(o, o->T, 2, o->d2Tdeta2);
  second_stencil(o, o->dil, 2, o->d2dildeta2);
  second_stencil(o, o->d2Sdxi2, 1, o->d4Sdxi4);
  second_stencil(o, o->d2Tdxi2, 1, o->d4Tdxi4);
  second_stencil(o, o->d2dildxi2, 1, o->d4dildxi4);
  second_stencil(o, o->d2Sdeta2, 2, o->d4Sdeta4);
  second_stencil(o, o->d2Tdeta2, 2, o->d4Tdeta4);
  second_stencil(o, o->d2dildeta2, 2, o->d4dildeta4);
  lad_quantities(o);
  tg_filter(o, o->muQ, 1, o->muQ);
  tg_filter(o, o->muQ, 2, o->mustar);
  tg_filter(o, o->betaQ, 1, o->betaQ);
  tg_filter(o, o->betaQ, 2, o->betastar);
  tg_filter(o, o->kappaQ, 1, o->kappaQ);
  tg_filter(o, o->kappaQ, 2, o->kappastar);
}

/* ComputeArtTransCoeffs (ModShockCapturing.f90:5-106) including UpdateTransportCoeffs
 * (ModShockCapturing.f90:238-258). */
static void art_trans_coeffs(orc* o) {
  lad_filtered(o);
  const double Re = o->p.Re, Pr = o->p.Pr, c23 = 2.0 / 3.0;
  const double cm = o->p.C_mu * Re, cb = o->p.C_beta * Re, ck = o->p.C_kappa * Re * Pr;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      o->mustar[h] = cm * o->mustar[h];
      o->betastar[h] = cb * o->betastar[h];
      o->kappastar[h] = ck * o->kappastar[h];
    }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      double tmp = (1.0 + tanh(5.0 * (o->zvec[h] - 10.0))) / 2.0;
      o->mustar[h] = o->mustar[h] * tmp;
      o->betastar[h] = o->betastar[h] * tmp;
      o->kappastar[h] = o->kappastar[h] * tmp;
    }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      o->mu[h] = o->mu[h] + o->mustar[h];
      o->beta[h] = o->beta[h] + o->betastar[h];
      o->lambda[h] = o->beta[h] - c23 * o->mu[h];
      o->kappa[h] = o->kappa[h] + o->kappastar[h];
    }
}

/* ControlForcingTerms (ModControlForcing.f90:46-62): at every point,
 * rhs(e) += alpha*forcing_mat(e,k)*GaussianFactor*(pv(k) - pv_target(k)), k = 1..5 in order,
 * alpha = 6 (ModuleVariables.f90:184), GaussianFactor = exp(-(z-z0)**2/lz**2 - (r-r0)**2/lr**2).
 * Then the inflow excitation of this project (no counterpart in the reference; ajet_capi.h):
 * rhs(Vz) += amp*Uj*sin(2 pi St Uj (t0 + t))*GaussianFactor. */
static void control_forcing(orc* o) {
  const double alpha = 6.0;
  const double w = 2.0 * 3.14159265358979323846 * o->fst * o->fuj;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      size_t h = HX(o, i, j);
      double dz = o->zvec[h] - o->fz0, dr = o->rvec[h] - o->fr0;
      double G = exp(-1.0 * ((dz * dz) / (o->flz * o->flz)) - (dr * dr) / (o->flr * o->flr));
      if (o->forcing & AJET_FORCING_FEEDBACK)
        for (int e = 1; e <= 5; ++e)
          for (int k = 1; k <= 5; ++k) {
            size_t se = SX(o, i, j, e), sk = SX(o, i, j, k);
            o->rhs[se] = o->rhs[se] + alpha * o->fmat[e - 1][k - 1] * G * (o->pv[sk] - o->pv_target[sk]);
          }
      if (o->forcing & AJET_FORCING_EXCITATION) {
        size_t s4 = SX(o, i, j, 4);
        o->rhs[s4] = o->rhs[s4] + o->famp * o->fuj * sin(w * (o->ft0 + o->t_stage)) * G;
      }
    }
}

void orc_set_forcing(orc* o, int mode, const double* F, double z0, double r0, double lz, double lr, double amp,
                     double st, double uj, double t0) {
  o->forcing = mode;
  for (int e = 0; e < 5; ++e)
    for (int k = 0; k < 5; ++k) o->fmat[e][k] = F ? F[5 * e + k] : 0.0;
  o->fz0 = z0; o->fr0 = r0; o->flz = lz; o->flr = lr;
  o->famp = amp; o->fst = st; o->fuj = uj; o->ft0 = t0;
}

/* ---------------------------------------------------------------- boundary conditions (ModBC.f90) */

typedef struct {
  double lam[5]; /* diagonal of lam */
  double Tm[5][5], Ti[5][5];
} eigsys;

/* eigen_matrices_in_z_direction / _in_r_direction (ModBC.f90:280-437); vn is Vz or Vr. */
static void eigen(double rho, double vn, double T, double gam, int rdir, eigsys* e) {
  memset(e, 0, sizeof *e);
  double sT = sqrt(T);
  e->lam[0] = vn; e->lam[1] = vn; e->lam[2] = vn;
  e->lam[3] = vn - sT; e->lam[4] = vn + sT;
  e->Tm[0][0] = -1.0 * rho / T;
  e->Tm[0][3] = rho / ((gam - 1.0) * T);
  e->Tm[0][4] = rho / ((gam - 1.0) * T);
  e->Ti[0][0] = -1.0 * (gam - 1.0) * T / (gam * rho);
  e->Ti[0][4] = 1.0 / gam;
  e->Ti[3][0] = (gam - 1.0) * T / (2.0 * gam * rho);
  e->Ti[3][4] = (gam - 1.0) / (2.0 * gam);
  e->Ti[4][0] = (gam - 1.0) * T / (2.0 * gam * rho);
  e->Ti[4][4] = (gam - 1.0) / (2.0 * gam);
  e->Tm[4][0] = 1.0; e->Tm[4][3] = 1.0; e->Tm[4][4] = 1.0;
  if (!rdir) {
    e->Tm[1][2] = 1.0; e->Tm[2][1] = 1.0;
    e->Tm[3][3] = -1.0 / ((gam - 1.0) * sT);
    e->Tm[3][4] = 1.0 / ((gam - 1.0) * sT);
    e->Ti[1][2] = 1.0; e->Ti[2][1] = 1.0;
    e->Ti[3][3] = -1.0 * (gam - 1.0) * sT / 2.0;
    e->Ti[4][3] = (gam - 1.0) * sT / 2.0;
  } else {
    e->Tm[1][3] = -1.0 / ((gam - 1.0) * sT);
    e->Tm[1][4] = 1.0 / ((gam - 1.0) * sT);
    e->Tm[2][2] = 1.0; e->Tm[3][1] = 1.0;
    e->Ti[1][3] = 1.0; e->Ti[2][2] = 1.0;
    e->Ti[3][1] = -1.0 * (gam - 1.0) * sT / 2.0;
    e->Ti[4][1] = (gam - 1.0) * sT / 2.0;
  }
}

/* MATMUL(A, v) for 5x5 by 5, each element summed sequentially from the first column. */
static void matvec5(const double A[5][5], const double* v, double* out) {
  for (int r = 0; r < 5; ++r) {
    double s = 0.0;
    for (int c = 0; c < 5; ++c) s = s + A[r][c] * v[c];
    out[r] = s;
  }
}

/* MATMUL(MATMUL(T_mat, lam), w) with a diagonal lam (ModBC.f90:466, 543). */
static void tlam_w(const double Tm[5][5], const double* lam, const double* w, double* out) {
  double M[5][5];
  for (int r = 0; r < 5; ++r)
    for (int c = 0; c < 5; ++c) {
      double s = 0.0;
      for (int k = 0; k < 5; ++k) s = s + Tm[r][k] * (k == c ? lam[c] : 0.0);
      M[r][c] = s;
    }
  matvec5(M, w, out);
}

/* BoundaryConditions (ModBC.f90:16-277) for the patch list in file order, then sponge_bcs
 * (ModBC.f90:548-617).  Assignments of the module scalar T are mirrored into o->T_scalar. */
static void boundary_conditions(orc* o) {
  const double gam = o->p.gam, sigma_1 = -2.0, sigma_2 = -2.0, h44 = 48.0 / 49.0;
  const int nz = o->nz, nr = o->nr;
  for (int pn = 0; pn < o->npatch; ++pn) {
    const ajet_patch* P = &o->patch[pn];
    if (P->bc == AJET_BC_AXIS) continue;
    int64_t is, ie, js, je;
    if (P->bc_dir == 1) { is = P->i_start; ie = P->i_start; js = P->j_start; je = P->j_end; }
    else if (P->bc_dir == -1) { is = P->i_end; ie = P->i_end; js = P->j_start; je = P->j_end; }
    else if (P->bc_dir == 2) { is = P->i_start; ie = P->i_end; js = P->j_start; je = P->j_start; }
    else { is = P->i_start; ie = P->i_end; js = P->j_end; je = P->j_end; }
    const int adir = P->bc_dir < 0 ? (int)-P->bc_dir : (int)P->bc_dir;
    for (int64_t i = is; i <= ie; ++i)
      for (int64_t j = js; j <= je; ++j) {
        size_t h = HX(o, i, j);
        double rhs[5];
        for (int e = 0; e < 5; ++e) rhs[e] = o->rhs[SX(o, i, j, e + 1)];
        double q[5], qt[5];
        for (int e = 0; e < 5; ++e) { q[e] = o->pv[SX(o, i, j, e + 1)]; qt[e] = o->pv_target[SX(o, i, j, e + 1)]; }
        eigsys E;
        double w[5], wt[5], out[5];
        if (P->bc == AJET_BC_SAT_FARFIELD) {
          if (j == 1) continue;
          if (adir == 1) {
            eigen(o->rho[h], o->Vz[h], o->T[h], gam, 0, &E);
            o->T_scalar = o->T[h];
            double lam[5];
            for (int c = 0; c < 5; ++c)
              lam[c] = P->bc_dir > 0 ? (fabs(E.lam[c]) + E.lam[c]) / 2.0 : (fabs(E.lam[c]) - E.lam[c]) / 2.0;
            for (int c = 0; c < 5; ++c) w[c] = q[c] - qt[c];
            matvec5(E.Ti, w, wt);
            tlam_w(E.Tm, lam, wt, out);
            for (int e = 0; e < 5; ++e) rhs[e] = rhs[e] + sigma_1 * o->dxidz[h] * h44 * out[e];
          }
        } else if (P->bc == AJET_BC_SAT_NOSLIP_WALL) {
          double qI[5], qV[5], lam[5];
          const double* met = adir == 1 ? o->dxidz : o->detadr;
          for (int c = 0; c < 5; ++c) qI[c] = q[c];
          if (adir == 1) qI[3] = 0.0;
          else qI[1] = 0.0;
          eigen(o->rho[h], adir == 1 ? o->Vz[h] : o->Vr[h], o->T[h], gam, adir == 2, &E);
          o->T_scalar = o->T[h];
          for (int c = 0; c < 5; ++c)
            lam[c] = P->bc_dir > 0 ? (fabs(E.lam[c]) + E.lam[c]) / 2.0 : (fabs(E.lam[c]) - E.lam[c]) / 2.0;
          for (int c = 0; c < 5; ++c) w[c] = q[c] - qI[c];
          matvec5(E.Ti, w, wt);
          tlam_w(E.Tm, lam, wt, out);
          for (int e = 0; e < 5; ++e) rhs[e] = rhs[e] + sigma_1 * met[h] * h44 * out[e];
          qV[0] = q[0]; qV[1] = 0.0; qV[2] = 0.0; qV[3] = 0.0; qV[4] = qt[4];
          for (int e = 0; e < 5; ++e) rhs[e] = rhs[e] + sigma_2 * met[h] * h44 * (q[e] - qV[e]);
        } else if (P->bc == AJET_BC_NSCBC_NONREFLECTING) {
          if (j == 1) continue;
          double rho = o->rho[h], Vr = o->Vr[h], Vz = o->Vz[h], T = o->T[h];
          o->T_scalar = T;
          double dv[5] = {o->drhodz[h], o->dVrdz[h], o->dVtdz[h], o->dVzdz[h], o->dTdz[h]};
          if (adir == 1) {
            rhs[0] = rhs[0] + (Vz * o->drhodz[h] + rho * o->dVzdz[h]);
            rhs[1] = rhs[1] + Vz * o->dVrdz[h];
            rhs[2] = rhs[2] + Vz * o->dVtdz[h];
            rhs[3] = rhs[3] + (Vz * o->dVzdz[h] + 1.0 / (gam * rho) * o->dpdz[h]);
            rhs[4] = rhs[4] + (Vz * o->dTdz[h] + (gam - 1.0) * T * o->dVzdz[h]);
            eigen(rho, Vz, T, gam, 0, &E);
          } else {
            rhs[0] = rhs[0] + (Vr * o->drhodr[h] + rho * o->dVrdr[h]);
            rhs[1] = rhs[1] + (Vr * o->dVrdr[h] + 1.0 / (gam * rho) * o->dpdr[h]);
            rhs[2] = rhs[2] + Vr * o->dVtdr[h];
            rhs[3] = rhs[3] + Vr * o->dVzdr[h];
            rhs[4] = rhs[4] + (Vr * o->dTdr[h] + (gam - 1.0) * T * o->dVrdr[h]);
            eigen(rho, Vr, T, gam, 1, &E);
          }
          matvec5(E.Ti, dv, wt);
          for (int c = 0; c < 5; ++c) {
            if (P->bc_dir > 0 && E.lam[c] > 0.0) wt[c] = 0.0;
            if (P->bc_dir < 0 && E.lam[c] < 0.0) wt[c] = 0.0;
          }
          tlam_w(E.Tm, E.lam, wt, out);
          for (int e = 0; e < 5; ++e) rhs[e] = rhs[e] - out[e];
        } else if (P->bc == AJET_BC_NSCBC_NOSLIP_WALL) {
          double rho = o->rho[h], Vr = o->Vr[h], Vz = o->Vz[h];
          double dv[5];
          if (adir == 1) {
            eigen(rho, Vz, o->T[h], gam, 0, &E);
            dv[0] = o->drhodz[h]; dv[1] = o->dVrdz[h]; dv[2] = o->dVtdz[h]; dv[3] = o->dVzdz[h]; dv[4] = o->dTdz[h];
          } else {
            eigen(rho, Vr, o->T[h], gam, 1, &E);
            dv[0] = o->drhodr[h]; dv[1] = o->dVrdr[h]; dv[2] = o->dVtdr[h]; dv[3] = o->dVzdr[h]; dv[4] = o->dTdr[h];
          }
          o->T_scalar = o->T[h];
          matvec5(E.Ti, dv, wt);
          if (P->bc_dir < 0) wt[3] = -1.0 * wt[4];
          else if (P->bc_dir > 0) wt[4] = -1.0 * wt[3];
          tlam_w(E.Tm, E.lam, wt, out);
          if (adir == 1) rhs[0] = rhs[0] + (Vz * o->drhodz[h] + rho * o->dVzdz[h]);
          else rhs[0] = rhs[0] + (Vr * o->drhodr[h] + rho * o->dVrdr[h]);
          for (int e = 0; e < 5; ++e) rhs[e] = rhs[e] - out[e];
          for (int e = 1; e < 5; ++e) rhs[e] = 0.0;
        } else {
          continue;
        }
        for (int e = 0; e < 5; ++e) o->rhs[SX(o, i, j, e + 1)] = rhs[e];
      }
  }

  /* sponge_bcs: top, then inflow */
  const double st = o->p.sponge_start_top, At = o->p.sponge_A_top;
  const double se = o->p.sponge_end_left, Al = o->p.sponge_A_left;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) {
      size_t h = HX(o, i, j);
      if (o->rvec[h] >= st) {
        for (int e = 1; e <= 5; ++e) {
          double zeta = (o->rvec[h] - st) / (o->rvec[HX(o, i, nr)] - st);
          size_t s = SX(o, i, j, e);
          o->rhs[s] = o->rhs[s] - At * (zeta * zeta) * (o->pv[s] - o->pv_target[s]);
        }
      }
    }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) {
      size_t h = HX(o, i, j);
      if (o->zvec[h] <= se) {
        for (int e = 1; e <= 5; ++e) {
          double zeta = (se - o->zvec[h]) / se;
          size_t s = SX(o, i, j, e);
          o->rhs[s] = o->rhs[s] - Al * (zeta * zeta) * (o->pv[s] - o->pv_target[s]);
        }
      }
    }
}

/* ComputeRHS (ModRHS.f90:5-258) of the primitives in the halo fields and the state pv. */
static void compute_rhs(orc* o) {
  const int nz = o->nz, nr = o->nr;
  const double gam = o->p.gam, Re = o->p.Re, Pr = o->p.Pr;
  first_deriv(o, o->rho, 1, o->drhodz);
  first_deriv(o, o->Vr, 1, o->dVrdz);
  first_deriv(o, o->Vt, 1, o->dVtdz);
  first_deriv(o, o->Vz, 1, o->dVzdz);
  first_deriv(o, o->T, 1, o->dTdz);
  first_deriv(o, o->pr, 1, o->dpdz);
  first_deriv(o, o->rho, 2, o->drhodr);
  first_deriv(o, o->Vr, 2, o->dVrdr);
  first_deriv(o, o->Vt, 2, o->dVtdr);
  first_deriv(o, o->Vz, 2, o->dVzdr);
  first_deriv(o, o->T, 2, o->dTdr);
  first_deriv(o, o->pr, 2, o->dpdr);
  if (o->p.shock_capturing == 1) art_trans_coeffs(o);

  memset(o->rhs, 0, sizeof(double) * (size_t)nz * nr * 5);
  const int j0 = 2; /* the single domain owns the axis: nbrs(2) < 0 */
#pragma omp parallel for schedule(static)
  for (int j = j0; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) {
      size_t h = HX(o, i, j);
      double r = o->rvec[h], rho = o->rho[h], Vr = o->Vr[h], Vt = o->Vt[h], Vz = o->Vz[h], T = o->T[h];
      double drhodr = o->drhodr[h], dVrdr = o->dVrdr[h], dVtdr = o->dVtdr[h], dVzdr = o->dVzdr[h];
      double dTdr = o->dTdr[h], dpdr = o->dpdr[h];
      double drhodz = o->drhodz[h], dVrdz = o->dVrdz[h], dVtdz = o->dVtdz[h], dVzdz = o->dVzdz[h];
      double dTdz = o->dTdz[h], dpdz = o->dpdz[h];
      o->rhs[SX(o, i, j, 1)] = Vr * drhodr + Vz * drhodz + rho * (dVrdr + Vr / r + dVzdz);
      o->rhs[SX(o, i, j, 2)] = Vr * dVrdr + Vz * dVrdz + 1.0 / (gam * rho) * dpdr;
      o->rhs[SX(o, i, j, 3)] = Vr * dVtdr + Vr * Vt / r + Vz * dVtdz;
      o->rhs[SX(o, i, j, 4)] = Vr * dVzdr + Vz * dVzdz + 1.0 / (gam * rho) * dpdz;
      o->rhs[SX(o, i, j, 5)] = Vr * dTdr + Vz * dTdz + (gam - 1.0) * T * (dVrdr + Vr / r + dVzdz);
    }
  for (size_t s = 0; s < (size_t)nz * nr * 5; ++s) o->rhs[s] = -1.0 * o->rhs[s];
  if (nr >= j0) o->T_scalar = o->T[HX(o, nz, nr)];

  if (o->p.viscous == 1) {
    second_deriv(o, o->Vr, 1, o->dVrdz, o->d2Vrdz2);
    second_deriv(o, o->Vt, 1, o->dVtdz, o->d2Vtdz2);
    second_deriv(o, o->Vz, 1, o->dVzdz, o->d2Vzdz2);
    second_deriv(o, o->T, 1, o->dTdz, o->d2Tdz2);
    first_deriv(o, o->mu, 1, o->dmudz);
    first_deriv(o, o->kappa, 1, o->dkappadz);
    second_deriv(o, o->Vr, 2, o->dVrdr, o->d2Vrdr2);
    second_deriv(o, o->Vt, 2, o->dVtdr, o->d2Vtdr2);
    second_deriv(o, o->Vz, 2, o->dVzdr, o->d2Vzdr2);
    second_deriv(o, o->T, 2, o->dTdr, o->d2Tdr2);
    first_deriv(o, o->mu, 2, o->dmudr);
    first_deriv(o, o->kappa, 2, o->dkappadr);
    first_deriv(o, o->dVrdz, 2, o->d2Vrdzdr);
    first_deriv(o, o->dVzdz, 2, o->d2Vzdzdr);
    const int sc = o->p.shock_capturing == 1;
#pragma omp parallel for schedule(static)
    for (int j = j0; j <= nr; ++j)
      for (int i = 1; i <= nz; ++i) {
        size_t h = HX(o, i, j);
        double r = o->rvec[h];
        double onebyr = 1.0 / r;
        double rho = o->rho[h], Vr = o->Vr[h], Vt = o->Vt[h], Vz = o->Vz[h];
        double mu = o->mu[h], kappa = o->kappa[h];
        double dVrdr = o->dVrdr[h], dVtdr = o->dVtdr[h], dVzdr = o->dVzdr[h], dTdr = o->dTdr[h];
        double dmudr = o->dmudr[h], dkappadr = o->dkappadr[h];
        double dVrdz = o->dVrdz[h], dVtdz = o->dVtdz[h], dVzdz = o->dVzdz[h], dTdz = o->dTdz[h];
        double dmudz = o->dmudz[h], dkappadz = o->dkappadz[h];
        double d2Vrdz2 = o->d2Vrdz2[h], d2Vtdz2 = o->d2Vtdz2[h], d2Vzdz2 = o->d2Vzdz2[h], d2Tdz2 = o->d2Tdz2[h];
        double d2Vrdr2 = o->d2Vrdr2[h], d2Vtdr2 = o->d2Vtdr2[h], d2Vzdr2 = o->d2Vzdr2[h], d2Tdr2 = o->d2Tdr2[h];
        double d2Vrdzdr = o->d2Vrdzdr[h], d2Vzdzdr = o->d2Vzdzdr[h];
        double fourby3 = 29.0 / 15.0, oneby3 = 14.0 / 15.0, twoby3 = 1.0 / 15.0;
        if (sc) {
          double tmp = o->lambda[h] / o->mu[h];
          fourby3 = 2.0 + tmp;
          oneby3 = 1.0 + tmp;
          twoby3 = -1.0 * tmp;
        }
        double ob2 = onebyr * onebyr;
        size_t s2 = SX(o, i, j, 2), s3 = SX(o, i, j, 3), s4 = SX(o, i, j, 4), s5 = SX(o, i, j, 5);
        o->rhs[s2] = o->rhs[s2] +
                     1.0 / (Re * rho) *
                         (mu * (fourby3 * d2Vrdr2 + fourby3 * onebyr * dVrdr + oneby3 * d2Vzdzdr + d2Vrdz2 -
                                fourby3 * Vr * ob2) +
                          dmudr * (fourby3 * dVrdr - twoby3 * Vr * onebyr - twoby3 * dVzdz) +
                          dmudz * (dVrdz + dVzdr));
        o->rhs[s3] = o->rhs[s3] + 1.0 / (Re * rho) *
                                      (mu * (d2Vtdr2 + d2Vtdz2 + onebyr * dVtdr - Vt * ob2) +
                                       dmudr * (dVtdr - Vt * onebyr) + dmudz * dVtdz);
        o->rhs[s4] = o->rhs[s4] +
                     1.0 / (Re * rho) *
                         (mu * (d2Vzdr2 + fourby3 * d2Vzdz2 + oneby3 * d2Vrdzdr + onebyr * dVzdr +
                                oneby3 * onebyr * dVrdz) +
                          dmudr * (dVzdr + dVrdz) + dmudz * (fourby3 * dVzdz - twoby3 * dVrdr - twoby3 * Vr * onebyr));
        o->rhs[s5] = o->rhs[s5] +
                     gam / (Re * Pr * rho) *
                         (kappa * (d2Tdr2 + onebyr * dTdr + d2Tdz2) + dkappadr * dTdr + dkappadz * dTdz) +
                     gam * (gam - 1.0) / Re * mu / rho *
                         (dVrdr * (fourby3 * dVrdr - twoby3 * Vr * onebyr - twoby3 * dVzdz) +
                          (dVtdr - Vt * onebyr) * (dVtdr - Vr * onebyr) +
                          Vr * onebyr * (fourby3 * Vr * onebyr - twoby3 * dVrdr - twoby3 * dVzdz) + dVtdz * dVtdz +
                          dVzdz * (fourby3 * dVzdz - twoby3 * dVrdr - twoby3 * Vr * onebyr) +
                          (dVrdz + dVzdr) * (dVzdr + dVrdz));
      }
    if (nr >= j0) o->T_scalar = o->T[HX(o, nz, nr)];
  }
  if (o->forcing) control_forcing(o); /* ModRHS.f90:223-225 */
  boundary_conditions(o);
}

/* ---------------------------------------------------------------- public API */

orc* orc_create(const ajet_params* p, const double* z, const double* r, const int32_t* iblank,
                const ajet_bc* lines, int32_t nlines) {
  if (!p || !z || !r || p->nz < 8 || p->nr < 8) return NULL;
  orc* o = (orc*)calloc(1, sizeof(orc));
  o->p = *p;
  o->nz = p->nz;
  o->nr = p->nr;
  o->nzh = p->nz + 8;
  o->nrh = p->nr + 8;
  sbp_tables(o);
  int np = build_patches(o->nz, o->nr, lines, nlines, NULL, 0);
  if (np < 0) { free(o); return NULL; }
  o->patch = (ajet_patch*)calloc((size_t)(np > 0 ? np : 1), sizeof(ajet_patch));
  o->npatch = build_patches(o->nz, o->nr, lines, nlines, o->patch, np);
  size_t ng = (size_t)o->nz * o->nr;
  o->is_i = (int64_t*)calloc(ng, sizeof(int64_t));
  o->is_j = (int64_t*)calloc(ng, sizeof(int64_t));
  o->which_i = (int64_t*)calloc(ng, sizeof(int64_t));
  o->which_j = (int64_t*)calloc(ng, sizeof(int64_t));
  for (int pn = 0; pn < o->npatch; ++pn) {
    const ajet_patch* P = &o->patch[pn];
    for (int64_t i = P->i_start; i <= P->i_end; ++i)
      for (int64_t j = P->j_start; j <= P->j_end; ++j) {
        if (i < 1 || i > o->nz || j < 1 || j > o->nr) { orc_destroy(o); return NULL; }
        size_t g = GX(o, i, j);
        if (P->bc_dir == 1 || P->bc_dir == -1) { o->is_i[g] = 1; o->which_i[g] = pn + 1; }
        if (P->bc_dir == 2 || P->bc_dir == -2) { o->is_j[g] = 1; o->which_j[g] = pn + 1; }
      }
  }
  o->ibval = (double*)malloc(ng * sizeof(double));
  for (size_t g = 0; g < ng; ++g) o->ibval[g] = iblank ? (double)iblank[g] : 1.0;
  double** hf[] = {&o->zvec, &o->rvec, &o->dzdxi, &o->drdeta, &o->dxidz, &o->detadr, &o->ddxidxidz,
                   &o->ddetadetadr, &o->rho, &o->Vr, &o->Vt, &o->Vz, &o->T, &o->pr, &o->mu, &o->kappa,
                   &o->lambda, &o->beta, &o->drhodz, &o->dVrdz, &o->dVtdz, &o->dVzdz, &o->dTdz, &o->dpdz,
                   &o->dmudz, &o->dkappadz, &o->drhodr, &o->dVrdr, &o->dVtdr, &o->dVzdr, &o->dTdr, &o->dpdr,
                   &o->dmudr, &o->dkappadr, &o->d2Vrdz2, &o->d2Vtdz2, &o->d2Vzdz2, &o->d2Tdz2, &o->d2Vrdr2,
                   &o->d2Vtdr2, &o->d2Vzdr2, &o->d2Tdr2, &o->d2Vrdzdr, &o->d2Vzdzdr, &o->S, &o->d2Sdxi2,
                   &o->d2Sdeta2, &o->d4Sdxi4, &o->d4Sdeta4, &o->d2Tdxi2, &o->d2Tdeta2, &o->d4Tdxi4,
                   &o->d4Tdeta4, &o->dil, &o->d2dildxi2, &o->d2dildeta2, &o->d4dildxi4, &o->d4dildeta4,
                   &o->muQ, &o->mustar, &o->betaQ, &o->betastar, &o->kappaQ, &o->kappastar};
  for (size_t k = 0; k < sizeof hf / sizeof hf[0]; ++k) *hf[k] = halo_alloc(o);
  double** sf[] = {&o->pv, &o->pv_old, &o->pv_target, &o->rhs, &o->k1, &o->k2, &o->k3, &o->k4};
  for (size_t k = 0; k < sizeof sf / sizeof sf[0]; ++k) *sf[k] = (double*)calloc(ng * 5, sizeof(double));
  o->deltvec = (double*)calloc(ng, sizeof(double));
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      o->zvec[HX(o, i, j)] = z[GX(o, i, j)];
      o->rvec[HX(o, i, j)] = r[GX(o, i, j)];
    }
  /* metrics, ModSolver.f90:57-76 */
  metric(o, o->zvec, 1, o->dzdxi);
  metric(o, o->rvec, 2, o->drdeta);
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) {
      o->dxidz[HX(o, i, j)] = 1.0 / o->dzdxi[HX(o, i, j)];
      o->detadr[HX(o, i, j)] = 1.0 / o->drdeta[HX(o, i, j)];
    }
  metric(o, o->dxidz, 1, o->ddxidxidz);
  metric(o, o->detadr, 2, o->ddetadetadr);
  o->delt = 0.00025; /* ModuleVariables.f90:67 */
  return o;
}

void orc_destroy(orc* o) {
  if (!o) return;
  double* hf[] = {o->zvec, o->rvec, o->dzdxi, o->drdeta, o->dxidz, o->detadr, o->ddxidxidz, o->ddetadetadr,
                  o->rho, o->Vr, o->Vt, o->Vz, o->T, o->pr, o->mu, o->kappa, o->lambda, o->beta,
                  o->drhodz, o->dVrdz, o->dVtdz, o->dVzdz, o->dTdz, o->dpdz, o->dmudz, o->dkappadz,
                  o->drhodr, o->dVrdr, o->dVtdr, o->dVzdr, o->dTdr, o->dpdr, o->dmudr, o->dkappadr,
                  o->d2Vrdz2, o->d2Vtdz2, o->d2Vzdz2, o->d2Tdz2, o->d2Vrdr2, o->d2Vtdr2, o->d2Vzdr2,
                  o->d2Tdr2, o->d2Vrdzdr, o->d2Vzdzdr, o->S, o->d2Sdxi2, o->d2Sdeta2, o->d4Sdxi4,
                  o->d4Sdeta4, o->d2Tdxi2, o->d2Tdeta2, o->d4Tdxi4, o->d4Tdeta4, o->dil, o->d2dildxi2,
                  o->d2dildeta2, o->d4dildxi4, o->d4dildeta4, o->muQ, o->mustar, o->betaQ, o->betastar,
                  o->kappaQ, o->kappastar, o->pv, o->pv_old, o->pv_target, o->rhs, o->k1, o->k2, o->k3,
                  o->k4, o->deltvec, o->ibval};
  for (size_t k = 0; k < sizeof hf / sizeof hf[0]; ++k) free(hf[k]);
  free(o->is_i); free(o->is_j); free(o->which_i); free(o->which_j); free(o->patch);
  free(o);
}

int orc_npatch(const orc* o) { return o->npatch; }

int orc_patches(const orc* o, ajet_patch* out, int max_out) {
  for (int k = 0; k < o->npatch && k < max_out; ++k) out[k] = o->patch[k];
  return o->npatch;
}

void orc_tables(const orc* o, double* sbp1, double* sbp2, double* tg) {
  /* column-major (0:8, 0:5) and (0:8, 0:10) as the reference allocates them */
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 9; ++r) { sbp1[r + 9 * c] = o->SBP1[r][c]; sbp2[r + 9 * c] = o->SBP2[r][c]; }
  for (int c = 0; c < 11; ++c)
    for (int r = 0; r < 9; ++r) tg[r + 9 * c] = o->TG[r][c];
}

static void copy_out(const orc* o, const double* h, double* g) {
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) g[GX(o, i, j)] = h[HX(o, i, j)];
}

static void copy_in(orc* o, const double* g, double* h) {
  for (int j = 1; j <= o->nr; ++j)
    for (int i = 1; i <= o->nz; ++i) h[HX(o, i, j)] = g[GX(o, i, j)];
}

void orc_metrics(const orc* o, double* dxidz, double* detadr, double* ddxidxidz, double* ddetadetadr) {
  copy_out(o, o->dxidz, dxidz);
  copy_out(o, o->detadr, detadr);
  copy_out(o, o->ddxidxidz, ddxidxidz);
  copy_out(o, o->ddetadetadr, ddetadetadr);
}

void orc_set_state(orc* o, const double* pv, const double* pv_target) {
  size_t n = (size_t)o->nz * o->nr * 5;
  memcpy(o->pv, pv, n * sizeof(double));
  if (pv_target) memcpy(o->pv_target, pv_target, n * sizeof(double));
  assign_values(o);
  o->T_scalar = 0.0;
  o->time = 0.0;
}

void orc_get_state(const orc* o, double* pv) { memcpy(pv, o->pv, (size_t)o->nz * o->nr * 5 * sizeof(double)); }

int orc_rhs(orc* o, const double* q, double* rhs) {
  size_t n = (size_t)o->nz * o->nr * 5;
  double* save = (double*)malloc(n * sizeof(double));
  memcpy(save, o->pv, n * sizeof(double));
  double Tsave = o->T_scalar;
  memcpy(o->pv, q, n * sizeof(double));
  assign_values(o);
  o->t_stage = o->time;
  compute_rhs(o);
  memcpy(rhs, o->rhs, n * sizeof(double));
  memcpy(o->pv, save, n * sizeof(double));
  free(save);
  assign_values(o);
  o->T_scalar = Tsave;
  return 0;
}

/* ComputeTimeStep (ModInputOutput.f90:389-438) with the module scalar T. */
int orc_compute_dt(orc* o, double* dt) {
  const double Re = o->p.Re, Pr = o->p.Pr, CFL = o->p.CFL;
  const double sT = sqrt(o->T_scalar);
  double mn = INFINITY;
  for (int i = 1; i <= o->nz; ++i)
    for (int j = 1; j <= o->nr; ++j) {
      size_t g = GX(o, i, j), h = HX(o, i, j);
      if (o->ibval[g] == 0.0) { o->deltvec[g] = 100.0; }
      else {
        double a = o->dxidz[h], b = o->detadr[h];
        double inv = fabs(o->Vz[h] * a + o->Vr[h] * b) + sT * sqrt(a * a + b * b);
        double m1 = o->mu[h] / Re, m2 = o->lambda[h] / Re, m3 = o->kappa[h] / (Re * Pr);
        double mx = m1;
        if (m2 > mx) mx = m2;
        if (m3 > mx) mx = m3;
        double s = fabs(a) + fabs(b);
        double visc = 2.0 * mx * (s * s);
        o->deltvec[g] = CFL / (inv + visc);
      }
      if (o->deltvec[g] < mn) mn = o->deltvec[g];
    }
  o->delt = mn;
  if (dt) *dt = mn;
  if (mn == 100.0) return AJET_EBLOWUP;
  return 0;
}

void orc_set_dt(orc* o, double dt) { o->delt = dt; }

/* RK4 (ModRungeKutta.f90:5-171) without output. */
int orc_rk4(orc* o) {
  const size_t n = (size_t)o->nz * o->nr * 5;
  const double dt = o->delt;
  const double c16 = 1.0 / 6.0, c13 = 1.0 / 3.0;
  memcpy(o->pv_old, o->pv, n * sizeof(double));
  double* ks[4] = {o->k1, o->k2, o->k3, o->k4};
  const double cs[3] = {0.5, 0.5, 1.0};
  const double ct[4] = {0.0, 0.5, 0.5, 1.0};
  for (int s = 0; s < 4; ++s) {
    assign_values(o);
    o->t_stage = o->time + ct[s] * dt;
    compute_rhs(o);
    double* k = ks[s];
    for (size_t q = 0; q < n; ++q) k[q] = dt * o->rhs[q];
    if (s < 3) {
      if (s < 2)
        for (size_t q = 0; q < n; ++q) o->pv[q] = o->pv_old[q] + cs[s] * k[q];
      else
        for (size_t q = 0; q < n; ++q) o->pv[q] = o->pv_old[q] + k[q];
      axis_conditions(o);
    }
  }
  for (int e = 1; e <= 5; ++e)
    for (int j = 1; j <= o->nr; ++j)
      for (int i = 1; i <= o->nz; ++i) {
        if (o->ibval[GX(o, i, j)] == 0.0) continue;
        size_t q = SX(o, i, j, e);
        o->pv[q] = o->pv_old[q] + c16 * o->k1[q] + c13 * (o->k2[q] + o->k3[q]) + c16 * o->k4[q];
      }
  axis_conditions(o);
  assign_values(o);
  o->time = o->time + dt;
  return 0;
}

int orc_run(orc* o, int nsteps, double* dts) {
  for (int s = 0; s < nsteps; ++s) {
    double dt;
    int rc = orc_compute_dt(o, &dt);
    if (rc) return rc;
    if (dts) dts[s] = dt;
    orc_rk4(o);
  }
  return 0;
}

double orc_time(const orc* o) { return o->time; }
double orc_stale_T(const orc* o) { return o->T_scalar; }
void orc_set_stale_T(orc* o, double T) { o->T_scalar = T; }

static void with_state(orc* o, const double* q, double** save, double* Tsave) {
  size_t n = (size_t)o->nz * o->nr * 5;
  *save = (double*)malloc(n * sizeof(double));
  memcpy(*save, o->pv, n * sizeof(double));
  *Tsave = o->T_scalar;
  memcpy(o->pv, q, n * sizeof(double));
  assign_values(o);
  first_deriv(o, o->Vr, 1, o->dVrdz);
  first_deriv(o, o->Vz, 1, o->dVzdz);
  first_deriv(o, o->Vr, 2, o->dVrdr);
  first_deriv(o, o->Vz, 2, o->dVzdr);
}

static void restore_state(orc* o, double* save, double Tsave) {
  memcpy(o->pv, save, (size_t)o->nz * o->nr * 5 * sizeof(double));
  free(save);
  assign_values(o);
  o->T_scalar = Tsave;
}

int orc_lad(orc* o, const double* q, double* mustar, double* betastar, double* kappastar) {
  double* save;
  double Tsave;
  with_state(o, q, &save, &Tsave);
  art_trans_coeffs(o);
  copy_out(o, o->mustar, mustar);
  copy_out(o, o->betastar, betastar);
  copy_out(o, o->kappastar, kappastar);
  restore_state(o, save, Tsave);
  return 0;
}

/* quantities / zfiltered / rfiltered: 3 (nz, nr) fields each (mu, beta, kappa). */
int orc_lad_stages(orc* o, const double* q, double* quantities, double* zfiltered, double* rfiltered) {
  double* save;
  double Tsave;
  with_state(o, q, &save, &Tsave);
  size_t ng = (size_t)o->nz * o->nr;
  dilatation_strain(o);
  second_stencil(o, o->S, 1, o->d2Sdxi2);
  second_stencil(o, o->T, 1, o->d2Tdxi2);
  second_stencil(o, o->dil, 1, o->d2dildxi2);
  second_stencil(o, o->S, 2, o->d2Sdeta2);
  second_stencil(o, o->T, 2, o->d2Tdeta2);
  second_stencil(o, o->dil, 2, o->d2dildeta2);
  second_stencil(o, o->d2Sdxi2, 1, o->d4Sdxi4);
  second_stencil(o, o->d2Tdxi2, 1, o->d4Tdxi4);
  second_stencil(o, o->d2dildxi2, 1, o->d4dildxi4);
  second_stencil(o, o->d2Sdeta2, 2, o->d4Sdeta4);
  second_stencil(o, o->d2Tdeta2, 2, o->d4Tdeta4);
  second_stencil(o, o->d2dildeta2, 2, o->d4dildeta4);
  lad_quantities(o);
  copy_out(o, o->muQ, quantities);
  copy_out(o, o->betaQ, quantities + ng);
  copy_out(o, o->kappaQ, quantities + 2 * ng);
  tg_filter(o, o->muQ, 1, o->muQ);
  tg_filter(o, o->betaQ, 1, o->betaQ);
  tg_filter(o, o->kappaQ, 1, o->kappaQ);
  copy_out(o, o->muQ, zfiltered);
  copy_out(o, o->betaQ, zfiltered + ng);
  copy_out(o, o->kappaQ, zfiltered + 2 * ng);
  tg_filter(o, o->muQ, 2, o->mustar);
  tg_filter(o, o->betaQ, 2, o->betastar);
  tg_filter(o, o->kappaQ, 2, o->kappastar);
  copy_out(o, o->mustar, rfiltered);
  copy_out(o, o->betastar, rfiltered + ng);
  copy_out(o, o->kappastar, rfiltered + 2 * ng);
  restore_state(o, save, Tsave);
  return 0;
}

/* Single-operator hooks for the stencil known-answer tests (fields are (nz, nr)). */
void orc_d1(orc* o, const double* field, int dirn, double* out) {
  double* a = halo_alloc(o);
  double* b = halo_alloc(o);
  copy_in(o, field, a);
  first_deriv(o, a, dirn, b);
  copy_out(o, b, out);
  free(a);
  free(b);
}

void orc_d2(orc* o, const double* field, int dirn, const double* first, double* out) {
  double* a = halo_alloc(o);
  double* f = halo_alloc(o);
  double* b = halo_alloc(o);
  copy_in(o, field, a);
  copy_in(o, first, f);
  second_deriv(o, a, dirn, f, b);
  copy_out(o, b, out);
  free(a);
  free(f);
  free(b);
}

void orc_d2comp(orc* o, const double* field, int dirn, double* out) {
  double* a = halo_alloc(o);
  double* b = halo_alloc(o);
  copy_in(o, field, a);
  second_stencil(o, a, dirn, b);
  copy_out(o, b, out);
  free(a);
  free(b);
}

void orc_tgfilter(orc* o, const double* field, int dirn, int inplace, double* out) {
  double* a = halo_alloc(o);
  copy_in(o, field, a);
  if (inplace) {
    tg_filter(o, a, dirn, a);
    copy_out(o, a, out);
  } else {
    double* b = halo_alloc(o);
    tg_filter(o, a, dirn, b);
    copy_out(o, b, out);
    free(b);
  }
  free(a);
}

// Pointwise physics of ComputeRHS (ModRHS.f90:5-258) and the boundary conditions
// (ModBC.f90:16-617) as device functions.  inviscid, viscous and the boundary conditions keep
// the reference's left-to-right association; the library is built with -ffp-contract=off, so
// each product and sum is rounded on its own exactly as in the reference's unoptimised build.
// inviscid_fast and viscous_fast are the interior tiles' fused forms.
#pragma once

#include <hip/hip_runtime.h>

#include "ajet_internal.h"
#include "ajet_fastmath.h"

namespace ajet {

struct PointDerivs {
  // first derivatives (already multiplied by the metric)
  double drhodz, dVrdz, dVtdz, dVzdz, dTdz, dpdz, dmudz, dkappadz;
  double drhodr, dVrdr, dVtdr, dVzdr, dTdr, dpdr, dmudr, dkappadr;
  // second and cross derivatives
  double d2Vrdz2, d2Vtdz2, d2Vzdz2, d2Tdz2, d2Vrdr2, d2Vtdr2, d2Vzdr2, d2Tdr2;
  double d2Vrdzdr, d2Vzdzdr;
};

struct PhysConst {
  double gam, Re, Pr;
  double sponge_start_top, sponge_A_top, sponge_end_left, sponge_A_left;
  double r_top;  // rvec(i, Nr) of the tensor-product grid
  // products of the constants, for the interior-tile forms (viscous_fast and the CFL rule)
  double inv_gam, gm1, inv_Re, g_RePr, g_gm1_Re;  // 1/gam, gam - 1, 1/Re, gam/(Re Pr), gam (gam - 1)/Re
  double cfl_visc;                                 // max(1/Re, 1/(Re Pr))
};

// Inviscid terms, ModRHS.f90:92-96, already negated (ModRHS.f90:103-113).
__device__ __forceinline__ void inviscid(const double q[5], double r, const PointDerivs& d, double gam,
                                         double rhs[5]) {
  const double rho = q[0], Vr = q[1], Vt = q[2], Vz = q[3], T = q[4];
  rhs[0] = -1.0 * (Vr * d.drhodr + Vz * d.drhodz + rho * (d.dVrdr + Vr / r + d.dVzdz));
  rhs[1] = -1.0 * (Vr * d.dVrdr + Vz * d.dVrdz + 1.0 / (gam * rho) * d.dpdr);
  rhs[2] = -1.0 * (Vr * d.dVtdr + Vr * Vt / r + Vz * d.dVtdz);
  rhs[3] = -1.0 * (Vr * d.dVzdr + Vz * d.dVzdz + 1.0 / (gam * rho) * d.dpdz);
  rhs[4] = -1.0 * (Vr * d.dTdr + Vz * d.dTdz + (gam - 1.0) * T * (d.dVrdr + Vr / r + d.dVzdz));
}

// Viscous terms, ModRHS.f90:180-205; sc selects the per-point coefficients of the
// shock-capturing build (ModRHS.f90:180-188).
__device__ __forceinline__ void viscous(const double q[5], double r, double mu, double kappa, double lambda,
                                        bool sc, const PointDerivs& d, const PhysConst& c, double rhs[5]) {
  const double gam = c.gam, Re = c.Re, Pr = c.Pr;
  const double onebyr = 1.0 / r;
  const double rho = q[0], Vr = q[1], Vt = q[2];
  double fourby3 = 29.0 / 15.0, oneby3 = 14.0 / 15.0, twoby3 = 1.0 / 15.0;
  if (sc) {
    const double tmp = lambda / mu;
    fourby3 = 2.0 + tmp;
    oneby3 = 1.0 + tmp;
    twoby3 = -1.0 * tmp;
  }
  const double ob2 = onebyr * onebyr;
  rhs[1] = rhs[1] + 1.0 / (Re * rho) *
                        (mu * (fourby3 * d.d2Vrdr2 + fourby3 * onebyr * d.dVrdr + oneby3 * d.d2Vzdzdr + d.d2Vrdz2 -
                               fourby3 * Vr * ob2) +
                         d.dmudr * (fourby3 * d.dVrdr - twoby3 * Vr * onebyr - twoby3 * d.dVzdz) +
                         d.dmudz * (d.dVrdz + d.dVzdr));
  rhs[2] = rhs[2] + 1.0 / (Re * rho) *
                        (mu * (d.d2Vtdr2 + d.d2Vtdz2 + onebyr * d.dVtdr - Vt * ob2) +
                         d.dmudr * (d.dVtdr - Vt * onebyr) + d.dmudz * d.dVtdz);
  rhs[3] = rhs[3] + 1.0 / (Re * rho) *
                        (mu * (d.d2Vzdr2 + fourby3 * d.d2Vzdz2 + oneby3 * d.d2Vrdzdr + onebyr * d.dVzdr +
                               oneby3 * onebyr * d.dVrdz) +
                         d.dmudr * (d.dVzdr + d.dVrdz) +
                         d.dmudz * (fourby3 * d.dVzdz - twoby3 * d.dVrdr - twoby3 * Vr * onebyr));
  rhs[4] = rhs[4] +
           gam / (Re * Pr * rho) *
               (kappa * (d.d2Tdr2 + onebyr * d.dTdr + d.d2Tdz2) + d.dkappadr * d.dTdr + d.dkappadz * d.dTdz) +
           gam * (gam - 1.0) / Re * mu / rho *
               (d.dVrdr * (fourby3 * d.dVrdr - twoby3 * Vr * onebyr - twoby3 * d.dVzdz) +
                (d.dVtdr - Vt * onebyr) * (d.dVtdr - Vr * onebyr) +
                Vr * onebyr * (fourby3 * Vr * onebyr - twoby3 * d.dVrdr - twoby3 * d.dVzdz) + d.dVtdz * d.dVtdz +
                d.dVzdz * (fourby3 * d.dVzdz - twoby3 * d.dVrdr - twoby3 * Vr * onebyr) +
                (d.dVrdz + d.dVzdr) * (d.dVzdr + d.dVrdz));
}

// The same inviscid and viscous terms for interior tiles, factored around fused multiply-adds
// with one reciprocal of rho and the row's 1/r (ir) from a table; the association differs from
// the reference's, so the results differ by a few ulp of the terms.
__device__ __forceinline__ void inviscid_fast(const double q[5], double ir, double irho, const PointDerivs& d,
                                              const PhysConst& c, double rhs[5]) {
  const double rho = q[0], Vr = q[1], Vt = q[2], Vz = q[3], T = q[4];
  const double div = d.dVrdr + fma(Vr, ir, d.dVzdz);  // dVr/dr + Vr/r + dVz/dz
  const double igr = irho * c.inv_gam;                 // 1/(gam rho)
  rhs[0] = -fma(Vr, d.drhodr, fma(Vz, d.drhodz, rho * div));
  rhs[1] = -fma(Vr, d.dVrdr, fma(Vz, d.dVrdz, igr * d.dpdr));
  rhs[2] = -fma(Vr, fma(Vt, ir, d.dVtdr), Vz * d.dVtdz);
  rhs[3] = -fma(Vr, d.dVzdr, fma(Vz, d.dVzdz, igr * d.dpdz));
  rhs[4] = -fma(Vr, d.dTdr, fma(Vz, d.dTdz, (c.gm1 * T) * div));
}

// ModRHS.f90:180-205 with vri = Vr/r and the three strain combinations shared between the
// momentum and the dissipation terms.
__device__ __forceinline__ void viscous_fast(const double q[5], double ir, double irho, double mu, double kappa,
                                             double lambda, bool sc, const PointDerivs& d, const PhysConst& c,
                                             double rhs[5]) {
  const double Vr = q[1], Vt = q[2];
  double f43 = 29.0 / 15.0, f13 = 14.0 / 15.0, f23 = 1.0 / 15.0;
  if (sc) {
    const double tmp = lambda * rcp(mu);
    f43 = 2.0 + tmp;
    f13 = 1.0 + tmp;
    f23 = -tmp;
  }
  const double vri = Vr * ir;
  const double sA = fma(f43, d.dVrdr, -f23 * (vri + d.dVzdz));   // f43 Vr_r - f23 Vr/r - f23 Vz_z
  const double sB = fma(f43, vri, -f23 * (d.dVrdr + d.dVzdz));   // f43 Vr/r - f23 Vr_r - f23 Vz_z
  const double sC = fma(f43, d.dVzdz, -f23 * (d.dVrdr + vri));   // f43 Vz_z - f23 Vr_r - f23 Vr/r
  const double shear = d.dVrdz + d.dVzdr;
  const double w = fma(-Vt, ir, d.dVtdr);                         // Vt_r - Vt/r
  const double a = irho * c.inv_Re;                               // 1/(Re rho)
  rhs[1] = fma(a, fma(mu, fma(f43, fma(ir, d.dVrdr - vri, d.d2Vrdr2), fma(f13, d.d2Vzdzdr, d.d2Vrdz2)),
                      fma(d.dmudr, sA, d.dmudz * shear)), rhs[1]);
  rhs[2] = fma(a, fma(mu, fma(ir, w, d.d2Vtdr2 + d.d2Vtdz2), fma(d.dmudr, w, d.dmudz * d.dVtdz)), rhs[2]);
  rhs[3] = fma(a, fma(mu, d.d2Vzdr2 + fma(f43, d.d2Vzdz2, fma(f13, fma(ir, d.dVrdz, d.d2Vrdzdr), ir * d.dVzdr)),
                      fma(d.dmudr, shear, d.dmudz * sC)), rhs[3]);
  const double heat = fma(kappa, d.d2Tdr2 + fma(ir, d.dTdr, d.d2Tdz2), fma(d.dkappadr, d.dTdr, d.dkappadz * d.dTdz));
  const double phi = fma(d.dVrdr, sA, fma(w, d.dVtdr - vri, fma(vri, sB, fma(d.dVtdz, d.dVtdz,
                     fma(d.dVzdz, sC, shear * shear)))));
  rhs[4] = fma(c.g_RePr * irho, heat, fma(c.g_gm1_Re * mu * irho, phi, rhs[4]));
}

// Characteristic matrices of eigen_matrices_in_z_direction / _in_r_direction
// (ModBC.f90:280-437), kept as their non-zero entries.
struct Eig {
  double lam[5];
  double t11, t14, t15;          // T_mat row 1 (columns 1, 4, 5)
  double tA, tB;                 // -1/((g-1)sqrtT), 1/((g-1)sqrtT)
  double i11, i15, i41, i45;     // T_mat_inv entries shared by both directions
  double i4s, i5s;               // -(g-1)sqrtT/2, (g-1)sqrtT/2
};

__device__ __forceinline__ void eigen(double rho, double vn, double T, double gam, Eig& e) {
  const double sT = sqrt(T);
  e.lam[0] = vn; e.lam[1] = vn; e.lam[2] = vn;
  e.lam[3] = vn - sT;
  e.lam[4] = vn + sT;
  e.t11 = -1.0 * rho / T;
  e.t14 = rho / ((gam - 1.0) * T);
  e.t15 = rho / ((gam - 1.0) * T);
  e.tA = -1.0 / ((gam - 1.0) * sT);
  e.tB = 1.0 / ((gam - 1.0) * sT);
  e.i11 = -1.0 * (gam - 1.0) * T / (gam * rho);
  e.i15 = 1.0 / gam;
  e.i41 = (gam - 1.0) * T / (2.0 * gam * rho);
  e.i45 = (gam - 1.0) / (2.0 * gam);
  e.i4s = -1.0 * (gam - 1.0) * sT / 2.0;
  e.i5s = (gam - 1.0) * sT / 2.0;
}

// waveamp = MATMUL(T_mat_inv, w): each row summed from its first column; the structural
// zeros of the matrix add exact zeros and are left out.
__device__ __forceinline__ void tinv_w(const Eig& e, bool rdir, const double w[5], double o[5]) {
  o[0] = e.i11 * w[0] + e.i15 * w[4];
  if (!rdir) {
    o[1] = 1.0 * w[2];
    o[2] = 1.0 * w[1];
    o[3] = e.i41 * w[0] + e.i4s * w[3] + e.i45 * w[4];
    o[4] = e.i41 * w[0] + e.i5s * w[3] + e.i45 * w[4];
  } else {
    o[1] = 1.0 * w[3];
    o[2] = 1.0 * w[2];
    o[3] = e.i41 * w[0] + e.i4s * w[1] + e.i45 * w[4];
    o[4] = e.i41 * w[0] + e.i5s * w[1] + e.i45 * w[4];
  }
}

// MATMUL(MATMUL(T_mat, lam), w) with diagonal lam: entry (r, c) of T_mat*lam is
// T_mat(r, c) * lam(c).
__device__ __forceinline__ void tlam_w(const Eig& e, bool rdir, const double lam[5], const double w[5], double o[5]) {
  o[0] = e.t11 * lam[0] * w[0] + e.t14 * lam[3] * w[3] + e.t15 * lam[4] * w[4];
  if (!rdir) {
    o[1] = 1.0 * lam[2] * w[2];
    o[2] = 1.0 * lam[1] * w[1];
    o[3] = e.tA * lam[3] * w[3] + e.tB * lam[4] * w[4];
  } else {
    o[1] = e.tA * lam[3] * w[3] + e.tB * lam[4] * w[4];
    o[2] = 1.0 * lam[2] * w[2];
    o[3] = 1.0 * lam[1] * w[1];
  }
  o[4] = 1.0 * lam[0] * w[0] + 1.0 * lam[3] * w[3] + 1.0 * lam[4] * w[4];
}

// One patch of BoundaryConditions (ModBC.f90:54-267) at one point.  q is the stage state,
// qt the target; met is dxidz for |dir| = 1 and detadr for |dir| = 2.
__device__ __forceinline__ void apply_patch(int bc, int dir, const double q[5], const double qt[5],
                                         const PointDerivs& d, double dxidz, double detadr, double gam,
                                         double rhs[5]) {
  const double sigma_1 = -2.0, sigma_2 = -2.0, h44 = 48.0 / 49.0;
  const int adir = dir < 0 ? -dir : dir;
  const bool rdir = adir == 2;
  const double rho = q[0], Vr = q[1], Vz = q[3], T = q[4];
  Eig e;
  double w[5], wt[5], out[5];
  if (bc == AJET_BC_SAT_FARFIELD) {
    // only faces normal to z carry the far-field penalty (ModBC.f90:64)
    eigen(rho, Vz, T, gam, e);
    double lam[5];
    for (int c = 0; c < 5; ++c) lam[c] = dir > 0 ? (fabs(e.lam[c]) + e.lam[c]) / 2.0 : (fabs(e.lam[c]) - e.lam[c]) / 2.0;
    for (int c = 0; c < 5; ++c) w[c] = q[c] - qt[c];
    tinv_w(e, false, w, wt);
    tlam_w(e, false, lam, wt, out);
    for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] + sigma_1 * dxidz * h44 * out[k];
  } else if (bc == AJET_BC_SAT_NOSLIP_WALL) {
    const double met = rdir ? detadr : dxidz;
    double qI[5];
    for (int c = 0; c < 5; ++c) qI[c] = q[c];
    if (!rdir) qI[3] = 0.0;
    else qI[1] = 0.0;
    eigen(rho, rdir ? Vr : Vz, T, gam, e);
    double lam[5];
    for (int c = 0; c < 5; ++c) lam[c] = dir > 0 ? (fabs(e.lam[c]) + e.lam[c]) / 2.0 : (fabs(e.lam[c]) - e.lam[c]) / 2.0;
    for (int c = 0; c < 5; ++c) w[c] = q[c] - qI[c];
    tinv_w(e, rdir, w, wt);
    tlam_w(e, rdir, lam, wt, out);
    for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] + sigma_1 * met * h44 * out[k];
    const double qV[5] = {q[0], 0.0, 0.0, 0.0, qt[4]};
    for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] + sigma_2 * met * h44 * (q[k] - qV[k]);
  } else if (bc == AJET_BC_NSCBC_NONREFLECTING) {
    // wave amplitudes from z derivatives on every face (ModBC.f90:487-495)
    const double dv[5] = {d.drhodz, d.dVrdz, d.dVtdz, d.dVzdz, d.dTdz};
    if (!rdir) {
      rhs[0] = rhs[0] + (Vz * d.drhodz + rho * d.dVzdz);
      rhs[1] = rhs[1] + Vz * d.dVrdz;
      rhs[2] = rhs[2] + Vz * d.dVtdz;
      rhs[3] = rhs[3] + (Vz * d.dVzdz + 1.0 / (gam * rho) * d.dpdz);
      rhs[4] = rhs[4] + (Vz * d.dTdz + (gam - 1.0) * T * d.dVzdz);
      eigen(rho, Vz, T, gam, e);
    } else {
      rhs[0] = rhs[0] + (Vr * d.drhodr + rho * d.dVrdr);
      rhs[1] = rhs[1] + (Vr * d.dVrdr + 1.0 / (gam * rho) * d.dpdr);
      rhs[2] = rhs[2] + Vr * d.dVtdr;
      rhs[3] = rhs[3] + Vr * d.dVzdr;
      rhs[4] = rhs[4] + (Vr * d.dTdr + (gam - 1.0) * T * d.dVrdr);
      eigen(rho, Vr, T, gam, e);
    }
    tinv_w(e, rdir, dv, wt);
    for (int c = 0; c < 5; ++c) {
      if (dir > 0 && e.lam[c] > 0.0) wt[c] = 0.0;
      if (dir < 0 && e.lam[c] < 0.0) wt[c] = 0.0;
    }
    tlam_w(e, rdir, e.lam, wt, out);
    for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] - out[k];
  } else if (bc == AJET_BC_NSCBC_NOSLIP_WALL) {
    double dv[5];
    if (!rdir) {
      eigen(rho, Vz, T, gam, e);
      dv[0] = d.drhodz; dv[1] = d.dVrdz; dv[2] = d.dVtdz; dv[3] = d.dVzdz; dv[4] = d.dTdz;
    } else {
      eigen(rho, Vr, T, gam, e);
      dv[0] = d.drhodr; dv[1] = d.dVrdr; dv[2] = d.dVtdr; dv[3] = d.dVzdr; dv[4] = d.dTdr;
    }
    tinv_w(e, rdir, dv, wt);
    if (dir < 0) wt[3] = -1.0 * wt[4];
    else if (dir > 0) wt[4] = -1.0 * wt[3];
    tlam_w(e, rdir, e.lam, wt, out);
    if (!rdir) rhs[0] = rhs[0] + (Vz * d.drhodz + rho * d.dVzdz);
    else rhs[0] = rhs[0] + (Vr * d.drhodr + rho * d.dVrdr);
    for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] - out[k];
    for (int k = 1; k < 5; ++k) rhs[k] = 0.0;
  }
}

}  // namespace ajet

// gfx950 kernels of the RK4 time step: the fused stage kernel (ComputeRHS + stage update +
// axis fix, with the CFL reduction of the next step fused into the last stage), the
// stand-alone CFL reduction, and its finalisation.  Build with -ffp-contract=off: the only fused
// multiply-adds are the explicit fma() calls of the interior-tile forms.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ajet_kernels.h"
#include "ajet_physics.h"
#include "ajet_stencil.h"

namespace ajet {

// Waves per SIMD the stage kernels are compiled for.  Two (up to 256 VGPRs) lets every
// per-point input load issue at the top of the point's work without spilling; three blocks'
// LDS would fit a CU, but at 168 VGPRs the point code spills (measured slower).
#ifndef AJET_STAGE_WAVES
#define AJET_STAGE_WAVES 2
#endif
constexpr int STAGE_WAVES_PER_EU = AJET_STAGE_WAVES;

#define AJET_FENCE() do { } while (0)

__device__ __forceinline__ double warp_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Denominator of ComputeTimeStep (ModInputOutput.f90:404-412) at one point.
__device__ __forceinline__ double cfl_denominator(double Vr, double Vz, double T, double sTst, double a, double b,
                                                  double Re, double Pr) {
  const double c35 = 3.0 / 5.0, c23 = 2.0 / 3.0;
  const double mu = pow(T, 0.666);
  const double beta = c35 * mu;
  const double lambda = beta - c23 * mu;
  const double inv = fabs(Vz * a + Vr * b) + sTst * sqrt(a * a + b * b);
  const double m1 = mu / Re, m2 = lambda / Re, m3 = mu / (Re * Pr);
  double mx = m1;
  if (m2 > mx) mx = m2;
  if (m3 > mx) mx = m3;
  const double s = fabs(a) + fabs(b);
  return inv + 2.0 * mx * (s * s);
}

// dt = MINVAL(CFL / denominator) = CFL / MAXVAL(denominator), since a correctly rounded
// division by a positive number is monotone; blanked points contribute 100
// (ModInputOutput.f90:399-402).
__device__ __forceinline__ double dt_of_dmax(unsigned long long b, double CFL, int any_blank) {
  double dt = b == 0ull ? 100.0 : CFL / __longlong_as_double((long long)b);
  if (any_blank && dt > 100.0) dt = 100.0;
  return dt;
}

__device__ __forceinline__ void block_max_atomic(double v, unsigned long long* dst, double* red) {
  v = warp_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    if (m > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// NF staged fields: rho, Vr, Vt, Vz, T, p, mu, kappa, T**0.666 (the last two only with LAD:
// without it kappa = mu = T**0.666, AssignValues ModAssign.f90:19-22).
// INT: the tile list holds interior tiles only -- the window lies inside the grid, every owned
// point and every window row of the owned columns takes the interior stencils, and no owned
// point is on the axis or on a boundary patch -- so the metadata word, the closures, the
// patches, the axis fix and the window's bounds tests are compiled out, and the arithmetic takes
// the fused forms (ajet_fastmath.h, inviscid_fast, viscous_fast).  Boundary tiles keep the
// reference's operation order.
// MODE: MODE_PLAIN (viscous), MODE_SC (viscous with the per-point coefficients of
// ShockCapturing = 1 and zero artificial terms), MODE_LAD (with mu*, beta*, kappa*),
// MODE_INVISCID (do_viscous_terms = 0).
// FORCE: the forcing terms are compiled in (a.forcing != 0).
// The body works on tile t of the run described by `a`; `writer` marks the one boundary tile of
// the run whose thread 0 updates the run's step scalars.  Its LDS is the kernel's, so that a
// kernel holding both tile classes allocates it once.
template <int MODE>
constexpr int stage_nf() { return MODE == MODE_LAD ? 9 : 7; }
#define AJET_STAGE_LDS(MODE)                                  \
  __shared__ double sF[stage_nf<MODE>()][WR][WZ];             \
  __shared__ double sD[2][WR][TZ];                            \
  __shared__ double red[STAGE_THREADS / 64]

// What a boundary tile keeps in LDS besides its window, loaded with the window so that no
// per-point step waits on a dependent load: the window's metadata words, the ordered patch lists
// and the patches.  Interior tiles need none of it.
struct BndLds {
  uint16_t meta[WR][WZ];                // the metadata words (16 bits carry them)
  uint32_t seq[MAX_SEQ * SEQ_LEN / 4];  // the bytes of the patch lists, four to a word
  int8_t bcdir[MAX_PATCH][2];           // each patch's bc type and direction
};
struct NoLds {
  char unused;
};
template <bool INT>
using StageBndLds = std::conditional_t<INT, NoLds, BndLds>;
static_assert(MAX_SEQ * SEQ_LEN % 4 == 0, "the patch lists fill whole words");
__device__ __forceinline__ uint32_t bl_meta(const BndLds& b, int wr, int wz) { return b.meta[wr][wz]; }
__device__ __forceinline__ uint32_t bl_meta(const NoLds&, int, int) { return ROW_INTERIOR | (ROW_INTERIOR << 4); }
__device__ __forceinline__ int bl_seq(const BndLds& b, int k) { return reinterpret_cast<const uint8_t*>(b.seq)[k]; }
__device__ __forceinline__ int bl_seq(const NoLds&, int) { return 0; }
__device__ __forceinline__ PatchDev bl_patch(const BndLds& b, int k) {
  PatchDev P{};
  P.bc = b.bcdir[k][0];
  P.dir = b.bcdir[k][1];
  return P;
}
__device__ __forceinline__ PatchDev bl_patch(const NoLds&, int) { return PatchDev{}; }

// Interior tiles address fields through raw buffer descriptors: one 32-bit byte offset per
// point (voffset) shared by every field, the field's byte offset f * fstride * 8 in a scalar
// register (soffset), so no access needs vector address arithmetic.  ajet_create keeps a state's
// five fields below 2 GiB.  The descriptor's fourth word is the raw-buffer format of the gfx9
// family (gfx950 included).
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc rsrc(const double* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ double bld(Rsrc r, uint32_t voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t voff, int soff, double v) {
  using W = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(W, v), r, voff, soff, 0);
}

// Interior-tile first and second derivatives on an LDS line (stride s): the factored forms of
// ajet_fastmath.h.
__device__ __forceinline__ double d1i(const double* c, int s) { return d1_int(c[-2 * s], c[-s], c[s], c[2 * s]); }
__device__ __forceinline__ double d2i(const double* c, int s) {
  return d2_int(c[-2 * s], c[-s], c[0], c[s], c[2 * s]);
}

// Pointwise inputs of an interior point, loaded before the stencil work so that their latency
// hides behind it: the stage update's inputs (stages 2, 3: q^n, and S; stage 4: A and S), the
// row's eta metrics, r and 1/r, and the target where a sponge or the feedback forcing needs it.
struct PointIn {
  double pin0[5], pin1[5];
  double qt[5];
  double mr, mr2, r, ir;
  bool top, inflow;
};

template <int STAGE, bool FORCE>
__device__ __forceinline__ void int_inputs(const StageArgs& a, uint32_t boff, int j, double zc, PointIn& in) {
  const int fsb = (int)(8 * a.g.fstride);
  in.mr = a.detadr[j];
  in.mr2 = a.ddetadetadr[j];
  in.r = a.rc[j];
  in.ir = a.rinv[j];
  if (STAGE >= 2) {
    const Rsrc r0 = rsrc(STAGE == 4 ? a.acc_a : a.qn);
#pragma unroll
    for (int k = 0; k < 5; ++k) in.pin0[k] = bld(r0, boff, k * fsb);
  }
  if (STAGE >= 3) {
    const Rsrc r1 = rsrc(a.acc_s);
#pragma unroll
    for (int k = 0; k < 5; ++k) in.pin1[k] = bld(r1, boff, k * fsb);
  }
  in.top = in.r >= a.pc.sponge_start_top;
  in.inflow = j >= 1 && zc <= a.pc.sponge_end_left;
#pragma unroll
  for (int f = 0; f < 5; ++f) in.qt[f] = 0.0;
  if (in.top || in.inflow || (FORCE && (a.forcing & AJET_FORCING_FEEDBACK))) {
    const Rsrc rt = rsrc(a.tgt);
#pragma unroll
    for (int f = 0; f < 5; ++f) in.qt[f] = bld(rt, boff, f * fsb);
  }
}

// Window accessors of int_point: fields at the point and its SBP interior derivatives (the
// factored forms of ajet_fastmath.h, before the metric factors), and the staged dVr/dz, dVz/dz.
// TileWin: the window rows at fixed positions (row stride WZ); c points at field 0 of the point,
// d at dVr/dz of the point.
template <int FS, int DS>
struct TileWin {
  const double* c;
  const double* d;
  __device__ __forceinline__ double at(int f) const { return c[f * FS]; }
  __device__ __forceinline__ double z1(int f) const { return d1i(c + f * FS, 1); }
  __device__ __forceinline__ double r1(int f) const { return d1i(c + f * FS, WZ); }
  __device__ __forceinline__ double z2(int f) const { return d2i(c + f * FS, 1); }
  __device__ __forceinline__ double r2(int f) const { return d2i(c + f * FS, WZ); }
  __device__ __forceinline__ double dz(int k) const { return d[k * DS]; }
  __device__ __forceinline__ double dzr(int k) const { return d1i(d + k * DS, TZ); }
};

// lambda = beta - (2/3) mu with beta = (3/5) T**0.666 (+ beta*) (ModAssign.f90:21-23,
// ModShockCapturing.f90:249-250) in the interior tiles' fused form; mu_phys = T**0.666.
template <bool LAD>
__device__ __forceinline__ double int_lambda(double mu, double mu_phys, double beta_star) {
  double beta = 0.6 * mu_phys;
  if (LAD) beta = beta + beta_star;
  return fma(-2.0 / 3.0, mu, beta);
}

// The stage update's pointwise inputs held in registers (PointIn).
struct PinRegs {
  const PointIn& in;
  __device__ __forceinline__ double p0(int k) const { return in.pin0[k]; }
  __device__ __forceinline__ double p1(int k) const { return in.pin1[k]; }
};
// One interior point of a stage with the fused forms (ajet_fastmath.h, inviscid_fast,
// viscous_fast): derivatives from the window W, forcing, sponges, the stage update (its inputs
// from P) and its stores at byte offset boff; stage 4 folds the point's CFL denominator into
// dmax.  mz, mz2, zc: the column's xi metrics and z; in: the row's metrics and the target.
template <int STAGE, int MODE, bool FORCE, class Win, class Pins>
__device__ __forceinline__ void int_point(const StageArgs& a, uint32_t boff, const Win& W, double mz, double mz2,
                                          double zc, double dt, double sTst, double tst, const PointIn& in,
                                          const Pins& P, double& dmax) {
  constexpr bool LAD = MODE == MODE_LAD;
  constexpr bool VISC = MODE != MODE_INVISCID;
  constexpr bool SC = MODE == MODE_SC || MODE == MODE_LAD;
  constexpr int FK = 7, FM = 8;  // LAD: kappa and the physical mu
  const int fsb = (int)(8 * a.g.fstride);
  const double mr = in.mr, mr2 = in.mr2, r = in.r, ir = in.ir;

  double q[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) q[f] = W.at(f);

  PointDerivs d;
  double rhs[5];
  const double mu = W.at(6);
  const double irho = rcp(q[0]);
  d.drhodz = W.z1(0) * mz;
  d.dVrdz = W.dz(0);
  d.dVtdz = W.z1(2) * mz;
  d.dVzdz = W.dz(1);
  d.dTdz = W.z1(4) * mz;
  d.dpdz = W.z1(5) * mz;
  d.drhodr = W.r1(0) * mr;
  d.dVrdr = W.r1(1) * mr;
  d.dVtdr = W.r1(2) * mr;
  d.dVzdr = W.r1(3) * mr;
  d.dTdr = W.r1(4) * mr;
  d.dpdr = W.r1(5) * mr;
  inviscid_fast(q, ir, irho, d, a.pc, rhs);
  if (VISC) {
    d.dmudz = W.z1(6) * mz;
    d.dmudr = W.r1(6) * mr;
    double kappa = mu;
    if constexpr (LAD) {
      kappa = W.at(FK);
      d.dkappadz = W.z1(FK) * mz;
      d.dkappadr = W.r1(FK) * mr;
    } else {
      d.dkappadz = d.dmudz;
      d.dkappadr = d.dmudr;
    }
    const double z2 = mz * mz, r2 = mr * mr;
    d.d2Vrdz2 = fma(W.z2(1), z2, d.dVrdz * mz2);
    d.d2Vtdz2 = fma(W.z2(2), z2, d.dVtdz * mz2);
    d.d2Vzdz2 = fma(W.z2(3), z2, d.dVzdz * mz2);
    d.d2Tdz2 = fma(W.z2(4), z2, d.dTdz * mz2);
    d.d2Vrdr2 = fma(W.r2(1), r2, d.dVrdr * mr2);
    d.d2Vtdr2 = fma(W.r2(2), r2, d.dVtdr * mr2);
    d.d2Vzdr2 = fma(W.r2(3), r2, d.dVzdr * mr2);
    d.d2Tdr2 = fma(W.r2(4), r2, d.dTdr * mr2);
    d.d2Vrdzdr = W.dzr(0) * mr;
    d.d2Vzdzdr = W.dzr(1) * mr;
    double lambda = 0.0;
    if (SC) {
      lambda = int_lambda<LAD>(mu, LAD ? W.at(FM) : mu, LAD ? bld(rsrc(a.lad), boff, fsb) : 0.0);
    }
    viscous_fast(q, ir, irho, mu, kappa, lambda, SC, d, a.pc, rhs);
  }

  // forcing (ModRHS.f90:223-225), at every point, before the boundary conditions
  if (FORCE) {
    const double dzf = zc - a.fc.z0, drf = r - a.fc.r0;
    const double G = exp(-1.0 * ((dzf * dzf) / a.fc.lz2) - (drf * drf) / a.fc.lr2);
    if (a.forcing & AJET_FORCING_FEEDBACK) {
      // rhs(e) += alpha*forcing_mat(e,k)*GaussianFactor*(pv(k) - pv_target(k)), k in order
#pragma unroll
      for (int e = 0; e < 5; ++e)
#pragma unroll
        for (int k = 0; k < 5; ++k) rhs[e] = rhs[e] + (a.fc.aF[e][k] * G) * (q[k] - in.qt[k]);
    }
    if (a.forcing & AJET_FORCING_EXCITATION) rhs[3] = rhs[3] + (a.fc.au * sin(a.fc.w * (a.fc.t0 + tst))) * G;
  }

  // the sponges (ModBC.f90:548-617); an interior point carries no boundary patch
  if (in.top) {
    const double zeta = (r - a.pc.sponge_start_top) / (a.pc.r_top - a.pc.sponge_start_top);
#pragma unroll
    for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] - a.pc.sponge_A_top * (zeta * zeta) * (q[k] - in.qt[k]);
  }
  if (in.inflow) {
    const double zeta = (a.pc.sponge_end_left - zc) / a.pc.sponge_end_left;
#pragma unroll
    for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] - a.pc.sponge_A_left * (zeta * zeta) * (q[k] - in.qt[k]);
  }

  if (STAGE == 0) {
#pragma unroll
    for (int k = 0; k < 5; ++k) bst(rsrc(a.rhs_out), boff, k * fsb, rhs[k]);
    return;
  }

  // ---- stage update (ModRungeKutta.f90:32-146), k = dt rhs folded into one fused multiply-add
  // per value
  const double c16 = 1.0 / 6.0, c13 = 1.0 / 3.0;
  double out[5], acc[5];
  const double hdt = 0.5 * dt, dt6 = c16 * dt;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    if (STAGE == 1) {
      out[k] = fma(hdt, rhs[k], q[k]);
      acc[k] = fma(dt6, rhs[k], q[k]);
    } else if (STAGE == 2) {
      out[k] = fma(hdt, rhs[k], P.p0(k));
      acc[k] = dt * rhs[k];
    } else if (STAGE == 3) {
      out[k] = fma(dt, rhs[k], P.p0(k));
      acc[k] = fma(dt, rhs[k], P.p1(k));
    } else {
      out[k] = fma(dt6, rhs[k], fma(c13, P.p1(k), P.p0(k)));
    }
  }
  if (STAGE == 1) {
#pragma unroll
    for (int k = 0; k < 5; ++k) bst(rsrc(a.acc_a), boff, k * fsb, acc[k]);
  } else if (STAGE == 2 || STAGE == 3) {
#pragma unroll
    for (int k = 0; k < 5; ++k) bst(rsrc(a.acc_s), boff, k * fsb, acc[k]);
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) bst(rsrc(a.qout), boff, k * fsb, out[k]);
  if (STAGE == 4) {
    // ComputeTimeStep's denominator of the new state (lambda / Re never wins: lambda < 0)
    const double T = out[4];
    const double mun = pow0666(T);
    const double inv = fabs(fma(out[3], mz, out[1] * mr)) + sTst * sqrt(fma(mz, mz, mr * mr));
    const double sm = fabs(mz) + fabs(mr);
    const double den = fma(2.0 * (mun * a.pc.cfl_visc), sm * sm, inv);
    if (!(den <= 1.0e300)) a.sc->nonfinite = 1;
    dmax = fmax(dmax, den);
  }
}

// The stage on one tile: owned columns [box.x, box.y), rows [box.z, box.w).
template <int STAGE, int MODE, bool INT, bool FORCE>
__device__ __forceinline__ void stage_body(const StageArgs& a, int4 box, bool writer, int dt_from_dmax, int parity,
                                           double (&sF)[stage_nf<MODE>()][WR][WZ], double (&sD)[2][WR][TZ],
                                           double (&red)[STAGE_THREADS / 64], StageBndLds<INT>& bl) {
  constexpr bool LAD = MODE == MODE_LAD;
  constexpr bool VISC = MODE != MODE_INVISCID;
  constexpr bool SC = MODE == MODE_SC || MODE == MODE_LAD;
  constexpr int FK = 7, FM = 8;  // LAD: kappa and the physical mu

  const Geometry& g = a.g;
  const int nz = g.nz, nr = g.nr, pitch = g.pitch;
  const long long fs = g.fstride;
  const int fsb = (int)(8 * fs);  // a field's byte stride (interior tiles' buffer offsets)
  const int i0 = box.x, oi_lo = box.x, oi_hi = box.y, j0 = box.z, oj_lo = box.z, oj_hi = box.w;

  // ---- stage the window: primitives, p = rho T, mu = T**0.666 (AssignValues).  Every load
  // of the thread's window points (with LAD also mu* and kappa*) is issued before any of them
  // is used, so the block waits for memory once.  An interior tile's window lies inside the
  // grid (classify_tiles), so it loads without bounds tests.
  constexpr int NWIN = (WZ * WR + STAGE_THREADS - 1) / STAGE_THREADS;
  double win[NWIN][5];
  double wlad[NWIN][2];  // mu*, kappa* (LAD)
  bool inside[NWIN];
  // the xi metric of this thread's column in the pass below (the column is threadIdx.x mod TZ
  // in every iteration), loaded with the window
  static_assert(STAGE_THREADS % TZ == 0, "a thread keeps its column across the window rows");
  const int icol = min(i0 + (int)threadIdx.x % TZ, nz - 1);
  const double mcol = a.dxidz[icol], mcol2 = a.ddxidxidz[icol], zcol = a.zc[icol];
#pragma unroll
  for (int u = 0; u < NWIN; ++u) {
    const int k = threadIdx.x + u * STAGE_THREADS;
    const int wz = k % WZ, wr = k / WZ;
    const int i = i0 - HALO + wz, j = j0 - HALO + wr;
    if constexpr (INT) {
      inside[u] = k < WZ * WR;
      // the tail of the last pass loads the window's first point again
      const uint32_t boff = inside[u] ? 8u * (uint32_t)(i + pitch * j) : 8u * (uint32_t)(i0 - HALO + pitch * (j0 - HALO));
      const Rsrc rq = rsrc(a.qin);
#pragma unroll
      for (int f = 0; f < 5; ++f) win[u][f] = bld(rq, boff, f * fsb);
      if constexpr (LAD) {
        const Rsrc rl = rsrc(a.lad);
        wlad[u][0] = bld(rl, boff, 0);
        wlad[u][1] = bld(rl, boff, 2 * fsb);
      }
    } else {
      inside[u] = k < WZ * WR && i >= 0 && i < nz && j >= 0 && j < nr;
      // points outside the grid load point 0 (a valid address) and are zeroed after the load,
      // so every load issues without a branch
      const long long gi = inside[u] ? (long long)i + (long long)pitch * j : 0;
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        const double x = a.qin[gi + f * fs];
        win[u][f] = inside[u] ? x : 0.0;
      }
      if constexpr (LAD) {
        wlad[u][0] = a.lad[gi];
        wlad[u][1] = a.lad[gi + 2 * fs];
      }
      if (k < WZ * WR) bl.meta[wr][wz] = (uint16_t)(inside[u] ? a.meta[gi] : (ROW_BLANK | (ROW_BLANK << 4)));
    }
  }
  if constexpr (!INT) {
    // the patch lists and each patch's type and direction (every list, every patch slot)
    constexpr int NSW = MAX_SEQ * SEQ_LEN / 4;
    static_assert(NSW + MAX_PATCH <= 2 * STAGE_THREADS, "two entries per thread");
    static_assert(AJET_BC_SPONGE < 128, "bc types fit a byte");
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int k = threadIdx.x + u * STAGE_THREADS;
      if (k < NSW) {
        bl.seq[k] = reinterpret_cast<const uint32_t*>(a.seq)[k];
      } else if (k < NSW + MAX_PATCH) {
        const PatchDev P = a.patches[k - NSW];
        bl.bcdir[k - NSW][0] = (int8_t)P.bc;
        bl.bcdir[k - NSW][1] = (int8_t)P.dir;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NWIN; ++u) {
    const int k = threadIdx.x + u * STAGE_THREADS;
    if (k >= WZ * WR) break;
    const int wz = k % WZ, wr = k / WZ;
    double p = 0.0, mu = 0.0, kap = 0.0, mphys = 0.0;
    if (INT || inside[u]) {
      p = win[u][0] * win[u][4];
      mu = INT ? pow0666(win[u][4]) : pow(win[u][4], 0.666);
      mphys = mu;
      kap = mu;
      if constexpr (LAD) {
        mu = mu + wlad[u][0];
        kap = kap + wlad[u][1];
      }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) sF[f][wr][wz] = win[u][f];
    sF[5][wr][wz] = p;
    sF[6][wr][wz] = mu;
    if constexpr (LAD) {
      sF[FK][wr][wz] = kap;
      sF[FM][wr][wz] = mphys;
    }
  }
  __syncthreads();

  // ---- dVr/dz and dVz/dz on every window row of the owned columns (for the cross terms)
#pragma unroll 1
  for (int k = threadIdx.x; k < TZ * WR; k += STAGE_THREADS) {
    const int tz = k % TZ, wr = k / TZ;
    const int i = i0 + tz, j = j0 - HALO + wr;
    double dvr = 0.0, dvz = 0.0;
    if (INT) {
      if (i < oi_hi) {
        dvr = d1i(&sF[1][wr][tz + HALO], 1) * mcol;
        dvz = d1i(&sF[3][wr][tz + HALO], 1) * mcol;
      }
    } else if (i >= oi_lo && i < oi_hi && j >= 0 && j < nr) {
      const uint32_t zr = bl_meta(bl, wr, tz + HALO) & 15u;
      dvr = d1_sum(zr, &sF[1][wr][tz + HALO], 1) * mcol;
      dvz = d1_sum(zr, &sF[3][wr][tz + HALO], 1) * mcol;
    }
    sD[0][wr][tz] = dvr;
    sD[1][wr][tz] = dvz;
  }
  __syncthreads();

  // The time step: stage 1 finishes the reduction of the previous step's last stage itself
  // when asked to (every thread derives the same value; one writes it and logs it).  The
  // reduction slot is cleared in stage 2, after both stage-1 launches have read it.
  double dt = (STAGE > 0) ? a.sc->dt : 0.0;
  if (STAGE == 1 && dt_from_dmax) {
    dt = dt_of_dmax(a.sc->dmax_bits, a.CFL, g.any_blank);
    if (!INT && writer && threadIdx.x == 0) {
      a.sc->dt = dt;
      const long long n = a.sc->nlog;
      if (n < a.logcap) a.dtlog[n] = dt;
      a.sc->nlog = n + 1;
    }
  }
  if (STAGE == 2 && !INT && writer && threadIdx.x == 0) a.sc->dmax_bits = 0ull;
  double dmax = 0.0;
  double sTst = 0.0;
  // time of this stage's RHS (used by the inflow excitation): step start + c dt
  const double tstep = a.sc->tpar[parity];
  const double tst = STAGE <= 1 ? tstep : (STAGE == 4 ? tstep + dt : tstep + 0.5 * dt);
  if (STAGE == 4) {
    const double Tst = a.qin[(long long)g.iT + (long long)pitch * g.jT + 4 * fs];
    sTst = sqrt(Tst);
    if (!INT && writer && threadIdx.x == 0) {
      a.sc->T_stale = Tst;
      a.sc->time = a.sc->time + dt;
      a.sc->tpar[parity ^ 1] = tstep + dt;
    }
  }

  const int tz = threadIdx.x % TZ;
#ifndef AJET_POINT_UNROLL
#define AJET_POINT_UNROLL 1
#endif
#pragma unroll AJET_POINT_UNROLL
  for (int ty = threadIdx.x / TZ; ty < TR; ty += STAGE_THREADS / TZ) {
    const int i = i0 + tz, j = j0 + ty;
    if (i < oi_lo || i >= oi_hi || j < oj_lo || j >= oj_hi) continue;
    const int wz = tz + HALO, wr = ty + HALO;
    if constexpr (INT) {
      const uint32_t boff = 8u * (uint32_t)(i + pitch * j);
      PointIn in;
      int_inputs<STAGE, FORCE>(a, boff, j, zcol, in);
      const TileWin<WR * WZ, WR * TZ> W{&sF[0][wr][wz], &sD[0][wr][tz]};
      int_point<STAGE, MODE, FORCE>(a, boff, W, mcol, mcol2, zcol, dt, sTst, tst, in, PinRegs{in}, dmax);
      continue;
    }
    const long long gi = (long long)i + (long long)pitch * j;
    const uint32_t meta = bl_meta(bl, wr, wz);
    const uint32_t zr = meta & 15u, rr = (meta >> 4) & 15u;
    const double mz = mcol, mz2 = mcol2;  // the column's metrics, loaded with the window
    const double mr = a.detadr[j], mr2 = a.ddetadetadr[j];
    const double r = a.rc[j];
    const double zc = zcol;
    // pointwise inputs of the stage update, loaded before the stencil work so that their
    // latency hides behind it: stage 2, 3: q^n (and S); stage 4: A and S
    double pin0[5], pin1[5];
    if (STAGE >= 2) {
      const double* __restrict__ src0 = STAGE == 4 ? a.acc_a : a.qn;
#pragma unroll
      for (int k = 0; k < 5; ++k) pin0[k] = src0[gi + k * fs];
    }
    if (STAGE >= 3) {
      const double* __restrict__ src1 = a.acc_s;
#pragma unroll
      for (int k = 0; k < 5; ++k) pin1[k] = src1[gi + k * fs];
    }
    // the target where a boundary patch or a sponge needs it (rows from g.jtop on may be in the
    // top sponge: the test needs no load of r)
    const uint32_t sid = (meta >> 8) & 255u;
    const bool top = r >= a.pc.sponge_start_top;
    const bool inflow = j >= 1 && zc <= a.pc.sponge_end_left;
    double qt[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (sid != 0 || j >= g.jtop || inflow || (FORCE && (a.forcing & AJET_FORCING_FEEDBACK))) {
#pragma unroll
      for (int f = 0; f < 5; ++f) qt[f] = a.tgt[gi + f * fs];
    }

    double q[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) q[f] = sF[f][wr][wz];

    PointDerivs d;
    double rhs[5];
    double mu = sF[6][wr][wz];
    d.drhodz = d1_sum(zr, &sF[0][wr][wz], 1) * mz;
    d.dVrdz = sD[0][wr][tz];
    d.dVtdz = d1_sum(zr, &sF[2][wr][wz], 1) * mz;
    d.dVzdz = sD[1][wr][tz];
    d.dTdz = d1_sum(zr, &sF[4][wr][wz], 1) * mz;
    d.dpdz = d1_sum(zr, &sF[5][wr][wz], 1) * mz;
    AJET_FENCE();
    d.drhodr = d1_sum(rr, &sF[0][wr][wz], WZ) * mr;
    d.dVrdr = d1_sum(rr, &sF[1][wr][wz], WZ) * mr;
    d.dVtdr = d1_sum(rr, &sF[2][wr][wz], WZ) * mr;
    d.dVzdr = d1_sum(rr, &sF[3][wr][wz], WZ) * mr;
    d.dTdr = d1_sum(rr, &sF[4][wr][wz], WZ) * mr;
    d.dpdr = d1_sum(rr, &sF[5][wr][wz], WZ) * mr;
    AJET_FENCE();
    if (j == 0) {
      // the axis row is skipped by the interior loops (ModRHS.f90:59-67): rhs = -0
#pragma unroll
      for (int k = 0; k < 5; ++k) rhs[k] = -1.0 * 0.0;
    } else {
      inviscid(q, r, d, a.pc.gam, rhs);
    }
    AJET_FENCE();
    if (VISC) {
      d.dmudz = d1_sum(zr, &sF[6][wr][wz], 1) * mz;
      d.dmudr = d1_sum(rr, &sF[6][wr][wz], WZ) * mr;
      double kappa = mu;
      if constexpr (LAD) {
        kappa = sF[FK][wr][wz];
        d.dkappadz = d1_sum(zr, &sF[FK][wr][wz], 1) * mz;
        d.dkappadr = d1_sum(rr, &sF[FK][wr][wz], WZ) * mr;
      } else {
        d.dkappadz = d.dmudz;
        d.dkappadr = d.dmudr;
      }
      const double z2 = mz * mz, r2 = mr * mr;
      AJET_FENCE();
      d.d2Vrdz2 = d2_sum(zr, &sF[1][wr][wz], 1) * z2 + d.dVrdz * mz2;
      d.d2Vtdz2 = d2_sum(zr, &sF[2][wr][wz], 1) * z2 + d.dVtdz * mz2;
      d.d2Vzdz2 = d2_sum(zr, &sF[3][wr][wz], 1) * z2 + d.dVzdz * mz2;
      d.d2Tdz2 = d2_sum(zr, &sF[4][wr][wz], 1) * z2 + d.dTdz * mz2;
      AJET_FENCE();
      d.d2Vrdr2 = d2_sum(rr, &sF[1][wr][wz], WZ) * r2 + d.dVrdr * mr2;
      d.d2Vtdr2 = d2_sum(rr, &sF[2][wr][wz], WZ) * r2 + d.dVtdr * mr2;
      d.d2Vzdr2 = d2_sum(rr, &sF[3][wr][wz], WZ) * r2 + d.dVzdr * mr2;
      d.d2Tdr2 = d2_sum(rr, &sF[4][wr][wz], WZ) * r2 + d.dTdr * mr2;
      AJET_FENCE();
      d.d2Vrdzdr = d1_sum(rr, &sD[0][wr][tz], TZ) * mr;
      d.d2Vzdzdr = d1_sum(rr, &sD[1][wr][tz], TZ) * mr;
      AJET_FENCE();
      if (j != 0) {
        double lambda = 0.0;
        if (SC) {
          // beta = (3/5) T**0.666 (+ beta*), lambda = beta - (2/3) mu (ModAssign.f90:21-23,
          // ModShockCapturing.f90:249-250); without the artificial terms mu is T**0.666 itself
          const double mu_phys = LAD ? sF[FM][wr][wz] : mu;
          double beta = (3.0 / 5.0) * mu_phys;
          if (LAD) beta = beta + a.lad[gi + fs];
          lambda = beta - (2.0 / 3.0) * mu;
        }
        viscous(q, r, mu, kappa, lambda, SC, d, a.pc, rhs);
      }
    }

    // forcing (ModRHS.f90:223-225), at every point, before the boundary conditions
    if (FORCE) {
      const double dzf = zc - a.fc.z0, drf = r - a.fc.r0;
      const double G = exp(-1.0 * ((dzf * dzf) / a.fc.lz2) - (drf * drf) / a.fc.lr2);
      if (a.forcing & AJET_FORCING_FEEDBACK) {
        // rhs(e) += alpha*forcing_mat(e,k)*GaussianFactor*(pv(k) - pv_target(k)), k in order
#pragma unroll
        for (int e = 0; e < 5; ++e)
#pragma unroll
          for (int k = 0; k < 5; ++k) rhs[e] = rhs[e] + (a.fc.aF[e][k] * G) * (q[k] - qt[k]);
      }
      if (a.forcing & AJET_FORCING_EXCITATION) rhs[3] = rhs[3] + (a.fc.au * sin(a.fc.w * (a.fc.t0 + tst))) * G;
    }

    // boundary conditions (ModBC.f90:16-277), patches in bc-file order, then the sponges
    if (sid != 0 || top || inflow) {
      if (sid != 0) {
        for (int s = 0; s < SEQ_LEN; ++s) {
          const int pidx = bl_seq(bl, sid * SEQ_LEN + s);
          if (pidx == 0) break;
          const PatchDev P = bl_patch(bl, pidx - 1);
          apply_patch(P.bc, P.dir, q, qt, d, mz, mr, a.pc.gam, rhs);
        }
      }
      if (top) {
        const double zeta = (r - a.pc.sponge_start_top) / (a.pc.r_top - a.pc.sponge_start_top);
#pragma unroll
        for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] - a.pc.sponge_A_top * (zeta * zeta) * (q[k] - qt[k]);
      }
      if (inflow) {
        const double zeta = (a.pc.sponge_end_left - zc) / a.pc.sponge_end_left;
#pragma unroll
        for (int k = 0; k < 5; ++k) rhs[k] = rhs[k] - a.pc.sponge_A_left * (zeta * zeta) * (q[k] - qt[k]);
      }
    }

    if (STAGE == 0) {
#pragma unroll
      for (int k = 0; k < 5; ++k) a.rhs_out[gi + k * fs] = rhs[k];
      continue;
    }

    // ---- stage update (ModRungeKutta.f90:32-146), from the inputs loaded at the top
    double out[5], acc[5];
    const double c16 = 1.0 / 6.0, c13 = 1.0 / 3.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double kk = dt * rhs[k];
      if (STAGE == 1) {
        out[k] = q[k] + 0.5 * kk;
        acc[k] = q[k] + c16 * kk;
      } else if (STAGE == 2) {
        out[k] = pin0[k] + 0.5 * kk;
        acc[k] = kk;
      } else if (STAGE == 3) {
        out[k] = pin0[k] + kk;
        acc[k] = pin1[k] + kk;
      } else {
        out[k] = (pin0[k] + c13 * pin1[k]) + c16 * kk;
      }
    }
    if (STAGE == 1) {
      double* __restrict__ accA = a.acc_a;
#pragma unroll
      for (int k = 0; k < 5; ++k) accA[gi + k * fs] = acc[k];
    } else if (STAGE == 2 || STAGE == 3) {
      double* __restrict__ accS = a.acc_s;
#pragma unroll
      for (int k = 0; k < 5; ++k) accS[gi + k * fs] = acc[k];
    }
    if (STAGE == 4 && zr == ROW_BLANK) {
      // blanked points skip the final update and keep the last stage value (ModRungeKutta.f90:139)
#pragma unroll
      for (int k = 0; k < 5; ++k) out[k] = q[k];
    }
    if (j == 0) {
      // axis_conditions (ModAxis.f90:14-16) from the stage input rows 2..6
      const double a00 = c_sbp1[0][0];
      const int fl[3] = {0, 3, 4};
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int f = fl[u];
        const double s = (c_sbp1[0][1] * sF[f][wr + 1][wz] + c_sbp1[0][2] * sF[f][wr + 2][wz]) +
                         c_sbp1[0][3] * sF[f][wr + 3][wz];
        out[f] = -1.0 * s / a00;
      }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) a.qout[gi + k * fs] = out[k];

    if (STAGE == 4 && zr != ROW_BLANK) {
      const double den = cfl_denominator(out[1], out[3], out[4], sTst, mz, mr, a.pc.Re, a.pc.Pr);
      if (!(den <= 1.0e300)) a.sc->nonfinite = 1;
      dmax = fmax(dmax, den);
    }
  }
  if (STAGE == 4) block_max_atomic(dmax, &a.sc->dmax_bits, red);
}

// Waves per SIMD the compiler must allow: the interior tiles of stages 0-2 without LAD and
// forcing fit 3 (LDS 50.5 KB per block, at most 168 VGPRs, which the memory-clause scheduler
// would otherwise exceed by a few); the others are held to 2 by LDS or registers.
template <int STAGE, int MODE, bool INT, bool FORCE>
constexpr int stage_waves() {
  return INT ? (MODE != MODE_LAD && !FORCE && STAGE <= 2 ? 3 : STAGE_WAVES_PER_EU) : 2;
}

template <int STAGE, int MODE, bool INT, bool FORCE>
__global__ __launch_bounds__(STAGE_THREADS, (stage_waves<STAGE, MODE, INT, FORCE>())) void stage_kernel(StageArgs a) {
  AJET_STAGE_LDS(MODE);
  __shared__ StageBndLds<INT> bl;
  const int li = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (li >= a.ntiles) return;
  stage_body<STAGE, MODE, INT, FORCE>(a, a.boxes[li], li == 0, a.dt_from_dmax, a.parity, sF, sD, red, bl);
}

// Interior and boundary tiles in one launch (list entry t: interior tile t, ~t: boundary tile
// t; entry 0 is a boundary tile and writes the step scalars), so that the boundary tiles, which
// alone leave most of the GPU idle for ~30 us, run beside the interior tiles.  Each XCD's range
// starts with its boundary tiles (mixed_tile_list).
template <int STAGE, int MODE, bool FORCE>
__device__ __forceinline__ void mixed_body() {
  AJET_STAGE_LDS(MODE);
  // The arguments are read through the kernel-argument segment where each class uses them:
  // used as a by-value parameter they would all be loaded at the kernel's entry and, live across
  // the branch, spill scalar registers.
  const StageArgs& a = *(const StageArgs*)__builtin_amdgcn_kernarg_segment_ptr();
  const int li = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (li >= a.ntiles) return;
  __shared__ BndLds bl;
  NoLds nl;
  const int e = a.tiles[li];
  const int4 box = a.boxes[li];
  if (e >= 0) stage_body<STAGE, MODE, true, FORCE>(a, box, false, a.dt_from_dmax, a.parity, sF, sD, red, nl);
  else stage_body<STAGE, MODE, false, FORCE>(a, box, li == 0, a.dt_from_dmax, a.parity, sF, sD, red, bl);
}

// The LAD mode's stages 1-3: its 62 KB window allows two blocks per CU, at which both classes fit
// their registers.
template <int STAGE, int MODE, bool FORCE>
__global__ __launch_bounds__(STAGE_THREADS, 2) void stage_mixed_kernel(StageArgs args) {
  (void)args;
  mixed_body<STAGE, MODE, FORCE>();
}

// The plain mode's stages 1-3 at three blocks per CU (50.5 + 2.2 KB of LDS each): the interior
// tiles fit 168 VGPRs, the boundary tiles spill, which the interior tiles around them hide.  The
// VGPR count is set explicitly: from the launch bound alone the compiler takes the boundary
// class's 200-240 VGPRs and two blocks per CU.  (Stage 4 in this form ran 2.5x slower than its
// two classes in two launches, for a reason not found; it keeps them apart.)
template <int STAGE, int MODE, bool FORCE>
__global__ __launch_bounds__(STAGE_THREADS, 3) __attribute__((amdgpu_num_vgpr(168))) void stage_mixed3_kernel(
    StageArgs args) {
  (void)args;
  mixed_body<STAGE, MODE, FORCE>();
}

// The batched form: the work list holds (run, tile) entries of many runs, each run described by
// its own StageArgs in `runs`; entry bits 0-19 tile, 20-30 run, 31 writer tile of the run.
template <int STAGE, int MODE, bool INT, bool FORCE>
__global__ __launch_bounds__(STAGE_THREADS, INT ? STAGE_WAVES_PER_EU : 2) void stage_batch_kernel(
    const StageArgs* __restrict__ runs, const uint32_t* __restrict__ list, int n, int dt_from_dmax, int parity) {
  AJET_STAGE_LDS(MODE);
  __shared__ StageBndLds<INT> bl;
  const int li = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (li >= n) return;
  const uint32_t e = list[li];
  const StageArgs& a = runs[(e >> 20) & 0x7ffu];
  const int t = (int)(e & 0xfffffu), tzi = t % a.g.ntz, tri = t / a.g.ntz;
  const int4 box = make_int4(a.zcuts[tzi], a.zcuts[tzi + 1], a.rcuts[tri], a.rcuts[tri + 1]);
  stage_body<STAGE, MODE, INT, FORCE>(a, box, (e >> 31) != 0u, dt_from_dmax, parity, sF, sD, red, bl);
}

// ComputeTimeStep's sweep (ModInputOutput.f90:395-416) over the current state with the
// stale scalar T kept on the device.
__global__ __launch_bounds__(256) void cfl_kernel(CflArgs a) {
  __shared__ double red[4];
  const Geometry& g = a.g;
  const double sTst = sqrt(a.sc->T_stale);
  double dmax = 0.0;
  const long long n = (long long)g.nz * g.nr;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(p % g.nz), j = (int)(p / g.nz);
    const long long gi = (long long)i + (long long)g.pitch * j;
    if ((a.meta[gi] & 15u) == ROW_BLANK) continue;
    const double den = cfl_denominator(a.q[gi + g.fstride], a.q[gi + 3 * g.fstride], a.q[gi + 4 * g.fstride], sTst,
                                       a.dxidz[i], a.detadr[j], a.Re, a.Pr);
    if (!(den <= 1.0e300)) a.sc->nonfinite = 1;
    dmax = fmax(dmax, den);
  }
  block_max_atomic(dmax, &a.sc->dmax_bits, red);
}

__global__ void scalar_reset_kernel(StepScalars* sc) {
  if (threadIdx.x == 0) sc->dmax_bits = 0ull;
}

// dt = MINVAL(CFL / denominator) = CFL / MAXVAL(denominator), since a correctly rounded
// division by a positive number is monotone; blanked points contribute 100
// (ModInputOutput.f90:399-402).
__global__ void cfl_finalize_kernel(StepScalars* sc, double* log, long long log_cap, double CFL, int any_blank) {
  if (threadIdx.x != 0) return;
  const double dt = dt_of_dmax(sc->dmax_bits, CFL, any_blank);
  sc->dt = dt;
  if (sc->nlog < log_cap) log[sc->nlog] = dt;
  sc->nlog = sc->nlog + 1;
}

hipError_t upload_tables(const double sbp1[9][6], const double sbp2[9][6], const double tg[9][11], hipStream_t s) {
  hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sbp1), sbp1, sizeof(double) * 54, 0, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sbp2), sbp2, sizeof(double) * 54, 0, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tg), tg, sizeof(double) * 99, 0, hipMemcpyHostToDevice, s);
  return e;
}

static int pad8(int n) { return (n + 7) / 8 * 8; }

template <int MODE, bool INT, bool FORCE>
static void launch_stage_t(int stage, dim3 grid, dim3 block, const StageArgs& a, hipStream_t s) {
  switch (stage) {
    case 0: hipLaunchKernelGGL((stage_kernel<0, MODE, INT, FORCE>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((stage_kernel<1, MODE, INT, FORCE>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((stage_kernel<2, MODE, INT, FORCE>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((stage_kernel<3, MODE, INT, FORCE>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((stage_kernel<4, MODE, INT, FORCE>), grid, block, 0, s, a); break;
  }
}

template <int MODE, bool INT>
static void launch_mode(int stage, dim3 grid, dim3 block, const StageArgs& a, hipStream_t s) {
  if (a.forcing) launch_stage_t<MODE, INT, true>(stage, grid, block, a, s);
  else launch_stage_t<MODE, INT, false>(stage, grid, block, a, s);
}

hipError_t launch_stage(int stage, int mode, bool interior, const StageArgs& a, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  const int nblk = pad8(a.ntiles);
  dim3 grid(nblk), block(STAGE_THREADS);
  switch (mode) {
    case MODE_PLAIN:
      interior ? launch_mode<MODE_PLAIN, true>(stage, grid, block, a, s)
               : launch_mode<MODE_PLAIN, false>(stage, grid, block, a, s);
      break;
    case MODE_SC:
      interior ? launch_mode<MODE_SC, true>(stage, grid, block, a, s)
               : launch_mode<MODE_SC, false>(stage, grid, block, a, s);
      break;
    case MODE_LAD:
      interior ? launch_mode<MODE_LAD, true>(stage, grid, block, a, s)
               : launch_mode<MODE_LAD, false>(stage, grid, block, a, s);
      break;
    default:
      interior ? launch_mode<MODE_INVISCID, true>(stage, grid, block, a, s)
               : launch_mode<MODE_INVISCID, false>(stage, grid, block, a, s);
      break;
  }
  return hipGetLastError();
}

template <bool FORCE>
static void launch_mixed_t(int stage, int mode, dim3 grid, const StageArgs& a, hipStream_t s) {
  const dim3 block(STAGE_THREADS);
  if (mode == MODE_PLAIN) {
    switch (stage) {
      case 1: hipLaunchKernelGGL((stage_mixed3_kernel<1, MODE_PLAIN, FORCE>), grid, block, 0, s, a); break;
      case 2: hipLaunchKernelGGL((stage_mixed3_kernel<2, MODE_PLAIN, FORCE>), grid, block, 0, s, a); break;
      default: hipLaunchKernelGGL((stage_mixed3_kernel<3, MODE_PLAIN, FORCE>), grid, block, 0, s, a); break;
    }
    return;
  }
  switch (stage) {
    case 1: hipLaunchKernelGGL((stage_mixed_kernel<1, MODE_LAD, FORCE>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((stage_mixed_kernel<2, MODE_LAD, FORCE>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((stage_mixed_kernel<3, MODE_LAD, FORCE>), grid, block, 0, s, a); break;
  }
}

bool stage_mixed_supported(int stage, int mode) {
  return (mode == MODE_LAD || mode == MODE_PLAIN) && stage >= 1 && stage <= 3;
}

hipError_t launch_stage_mixed(int stage, int mode, const StageArgs& a, hipStream_t s) {
  if (a.ntiles <= 0 || !stage_mixed_supported(stage, mode)) return hipErrorInvalidValue;
  const dim3 grid(pad8(a.ntiles));
  if (a.forcing) launch_mixed_t<true>(stage, mode, grid, a, s);
  else launch_mixed_t<false>(stage, mode, grid, a, s);
  return hipGetLastError();
}

template <int MODE, bool INT, bool FORCE>
static void launch_batch_t(int stage, dim3 grid, const StageArgs* runs, const uint32_t* list, int n, int dtm, int par,
                           hipStream_t s) {
  const dim3 block(STAGE_THREADS);
  switch (stage) {
    case 1: hipLaunchKernelGGL((stage_batch_kernel<1, MODE, INT, FORCE>), grid, block, 0, s, runs, list, n, dtm, par); break;
    case 2: hipLaunchKernelGGL((stage_batch_kernel<2, MODE, INT, FORCE>), grid, block, 0, s, runs, list, n, dtm, par); break;
    case 3: hipLaunchKernelGGL((stage_batch_kernel<3, MODE, INT, FORCE>), grid, block, 0, s, runs, list, n, dtm, par); break;
    default: hipLaunchKernelGGL((stage_batch_kernel<4, MODE, INT, FORCE>), grid, block, 0, s, runs, list, n, dtm, par); break;
  }
}

hipError_t launch_stage_batch(int stage, int mode, bool interior, bool force, const StageArgs* runs,
                              const uint32_t* list, int n, int dt_from_dmax, int parity, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 grid(pad8(n));
  if (mode == MODE_PLAIN) {
    if (interior) force ? launch_batch_t<MODE_PLAIN, true, true>(stage, grid, runs, list, n, dt_from_dmax, parity, s)
                        : launch_batch_t<MODE_PLAIN, true, false>(stage, grid, runs, list, n, dt_from_dmax, parity, s);
    else force ? launch_batch_t<MODE_PLAIN, false, true>(stage, grid, runs, list, n, dt_from_dmax, parity, s)
               : launch_batch_t<MODE_PLAIN, false, false>(stage, grid, runs, list, n, dt_from_dmax, parity, s);
  } else if (mode == MODE_INVISCID) {
    if (interior) force ? launch_batch_t<MODE_INVISCID, true, true>(stage, grid, runs, list, n, dt_from_dmax, parity, s)
                        : launch_batch_t<MODE_INVISCID, true, false>(stage, grid, runs, list, n, dt_from_dmax, parity, s);
    else force ? launch_batch_t<MODE_INVISCID, false, true>(stage, grid, runs, list, n, dt_from_dmax, parity, s)
               : launch_batch_t<MODE_INVISCID, false, false>(stage, grid, runs, list, n, dt_from_dmax, parity, s);
  } else if (mode == MODE_LAD) {
    if (interior) force ? launch_batch_t<MODE_LAD, true, true>(stage, grid, runs, list, n, dt_from_dmax, parity, s)
                        : launch_batch_t<MODE_LAD, true, false>(stage, grid, runs, list, n, dt_from_dmax, parity, s);
    else force ? launch_batch_t<MODE_LAD, false, true>(stage, grid, runs, list, n, dt_from_dmax, parity, s)
               : launch_batch_t<MODE_LAD, false, false>(stage, grid, runs, list, n, dt_from_dmax, parity, s);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cfl(const CflArgs& a, double CFL, int any_blank, double* log, long long cap, hipStream_t s) {
  hipLaunchKernelGGL(scalar_reset_kernel, dim3(1), dim3(64), 0, s, a.sc);
  const long long n = (long long)a.g.nz * a.g.nr;
  int nb = (int)((n + 255) / 256);
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(cfl_kernel, dim3(nb), dim3(256), 0, s, a);
  hipLaunchKernelGGL(cfl_finalize_kernel, dim3(1), dim3(64), 0, s, a.sc, log, cap, CFL, any_blank);
  return hipGetLastError();
}

hipError_t launch_finalize(StepScalars* sc, double CFL, int any_blank, double* log, long long cap, hipStream_t s) {
  hipLaunchKernelGGL(cfl_finalize_kernel, dim3(1), dim3(64), 0, s, sc, log, cap, CFL, any_blank);
  return hipGetLastError();
}

// The five primitives at n points (i, j), 0-based: out[5 k + v].
__global__ __launch_bounds__(256) void probe_kernel(const double* __restrict__ q, Geometry g, const int* __restrict__ pts,
                                                    int n, double* __restrict__ out) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= 5 * n) return;
  const int p = k / 5, v = k - 5 * p;
  out[k] = q[(long long)v * g.fstride + pts[2 * p] + (long long)g.pitch * pts[2 * p + 1]];
}

hipError_t launch_probe(const double* q, const Geometry& g, const int* pts, int n, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(probe_kernel, dim3((unsigned)((5 * n + 255) / 256)), dim3(256), 0, s, q, g, pts, n, out);
  return hipGetLastError();
}

}  // namespace ajet

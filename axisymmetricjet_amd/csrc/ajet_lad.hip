// gfx950 kernels of the localized artificial diffusivity (ComputeArtTransCoeffs,
// ModShockCapturing.f90:5-106 with PerformTGfilter, ModTGfilter.f90:5-95):
//
//   lad_strain_kernel   |S|, dilatation and the Ducros-type switch f_sw, on the tiles whose
//                       values a source tile reads from memory (near closures and the axis)
//   lad_source_kernel   D2(D2(.)) of |S|, dil and T in xi and in eta; the three quantities.
//                       Interior tiles compute |S|, dil and f_sw themselves (fused tiles)
//   lad_zrfilter_kernel the z filter, applied in place by the reference, i.e. the causal
//                       recursion y_i = sum_{m<0} c_m y_{i+m} + sum_{m>=0} c_m x_{i+m}, then
//                       the r filter, C Re (Pr) scaling and the tanh ramp -> mu*, beta*, kappa*
//
// Build with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ajet_kernels.h"
#include "ajet_stencil.h"

namespace ajet {

// ---- |S|, dil, f_sw (ComputeDilatation, ComputeMgnStrRate ModShockCapturing.f90:108-159, and
// the switch of ComputeTobeFilteredQuantities :196-202) at one point from its four velocity
// derivatives.  On the axis row (axis) Vr/r is replaced by dVr/dr (ModShockCapturing.f90:116-120,
// 146-150).  The strain kernel and the fused source tiles both take this arithmetic.
__device__ __forceinline__ void strain_point(double dVrdz, double dVzdz, double dVrdr, double dVzdr, double Vr,
                                             double rcj, bool axis, double& S, double& dil, double& fsw) {
  const double c13 = 1.0 / 3.0;
  const double eps = (double)1e-15f;  // a default-REAL literal (ModShockCapturing.f90:202)
  double S22;
  if (axis) {
    dil = 2.0 * dVrdr + dVzdz;
    S22 = dVrdr - c13 * dil;
  } else {
    const double vbr = Vr / rcj;
    dil = (dVrdr + vbr) + dVzdz;
    S22 = vbr - c13 * dil;
  }
  const double S11 = dVrdr - c13 * dil;
  const double S13 = 0.5 * (dVrdz + dVzdr);
  const double S33 = dVzdz - c13 * dil;
  // the nine squares summed in order; S12, S21, S23, S32 are zero and add exact zeros
  S = sqrt((((S11 * S11 + S13 * S13) + S22 * S22) + S13 * S13) + S33 * S33);
  const double vort = fabs(dVrdz - dVzdr);
  const double md = -1.0 * dil;
  const double H = md < 0.0 ? 0.0 : 1.0;
  fsw = H * (dil * dil) / ((dil * dil + vort * vort) + eps);
}

// The strain kernel: the same tiles and 2-point window as the stage kernel; the first
// derivatives take the patch closures and are zero at blanked points.  It covers the tiles
// whose values the non-fused source tiles read (classify_lad_source).  INT: tile t's points
// and stencil rows are all interior (classify_lad_tiles), so no metadata is read.
template <bool INT>
__device__ __forceinline__ void strain_tile(const LadArgs& a, int t, double (&sVr)[WR][WZ], double (&sVz)[WR][WZ]) {
  const Geometry& g = a.g;
  const int nz = g.nz, nr = g.nr, pitch = g.pitch;
  const long long fs = g.fstride;
  if (a.str_mask && !a.str_mask[t]) return;
  const int tzi = t % g.ntz, tri = t / g.ntz;
  const int i0 = a.zcuts[tzi], i1 = a.zcuts[tzi + 1], j0 = a.rcuts[tri], j1 = a.rcuts[tri + 1];
  if (i1 <= a.str_imin) return;  // only feeds zero-ramp columns
  // all of this thread's window loads, and the metrics of its points, are issued before the
  // first store (one memory round trip)
  constexpr int NP = (TR * TZ + 255) / 256;  // owned points per thread, rows 256 / TZ apart
  static_assert(256 % TZ == 0, "a thread keeps its column across the rows");
  const double mz = a.dxidz[min(i0 + (int)threadIdx.x % TZ, nz - 1)];
  double mr[NP], rcj[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int j = min(j0 + (int)threadIdx.x / TZ + p * (256 / TZ), nr - 1);
    mr[p] = a.detadr[j];
    rcj[p] = a.rc[j];
  }
  constexpr int NW = (WZ * WR + 255) / 256;
  double vr[NW], vz[NW];
  bool vin[NW];
#pragma unroll
  for (int u = 0; u < NW; ++u) {
    const int k = (int)threadIdx.x + 256 * u;
    const int wz = k % WZ, wr = k / WZ;
    const int i = i0 - HALO + wz, j = j0 - HALO + wr;
    // outside the grid: load point 0 (a valid address) and zero it, so no load is branched
    vin[u] = k < WZ * WR && i >= 0 && i < nz && j >= 0 && j < nr;
    const long long gi = vin[u] ? (long long)i + (long long)pitch * j : 0;
    vr[u] = a.q[gi + fs];
    vz[u] = a.q[gi + 3 * fs];
  }
#pragma unroll
  for (int u = 0; u < NW; ++u) {
    const int k = (int)threadIdx.x + 256 * u;
    if (k < WZ * WR) {
      (&sVr[0][0])[k] = vin[u] ? vr[u] : 0.0;
      (&sVz[0][0])[k] = vin[u] ? vz[u] : 0.0;
    }
  }
  __syncthreads();
  const int tz = threadIdx.x % TZ;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int ty = (int)threadIdx.x / TZ + p * (256 / TZ);
    const int i = i0 + tz, j = j0 + ty;
    if (ty >= TR || i >= i1 || j >= j1) continue;
    const long long gi = (long long)i + (long long)pitch * j;
    const uint32_t meta = INT ? (ROW_INTERIOR | (ROW_INTERIOR << 4)) : a.meta[gi];
    const uint32_t zr = meta & 15u, rr = (meta >> 4) & 15u;
    const int wz = tz + HALO, wr = ty + HALO;
    const double dVrdz = d1_sum(zr, &sVr[wr][wz], 1) * mz;
    const double dVzdz = d1_sum(zr, &sVz[wr][wz], 1) * mz;
    const double dVrdr = d1_sum(rr, &sVr[wr][wz], WZ) * mr[p];
    const double dVzdr = d1_sum(rr, &sVz[wr][wz], WZ) * mr[p];
    double S, dil, fsw;
    strain_point(dVrdz, dVzdz, dVrdr, dVzdr, sVr[wr][wz], rcj[p], !INT && j == 0, S, dil, fsw);
    a.A[gi] = S;
    a.A[gi + fs] = dil;
    a.A[gi + 2 * fs] = fsw;
  }
  __syncthreads();  // the window is reused by the next tile
}

// Interior and boundary tiles in one launch: list entry t >= 0 is interior tile t, ~t boundary
// tile t.  One tile per block, or (a grid smaller than the list) several in turn.
__global__ __launch_bounds__(256) void lad_strain_kernel(LadArgs a) {
  __shared__ double sVr[WR][WZ];
  __shared__ double sVz[WR][WZ];
  for (int li = xcd_tile((int)blockIdx.x, (int)gridDim.x); li < a.nstr; li += (int)gridDim.x) {
    const int e = a.str_tiles[li];
    if (e >= 0) strain_tile<true>(a, e, sVr, sVz);
    else strain_tile<false>(a, ~e, sVr, sVz);
  }
}

// ---- D2(D2(.)) along xi and eta of |S|, dil and T, and the three quantities
// (ComputeTobeFilteredQuantities, ModShockCapturing.f90:162-222).  The second stencil of each
// pair reads the first one at up to 2 points beyond the tile (closure strips never straddle a
// tile edge), and that one reads its argument at up to 5 beyond, so |S|, dil and T are staged
// with a 7-point halo.  A second derivative at a blanked point or outside the grid is zero.
constexpr int LH = 7;
constexpr int LWZ = TZ + 2 * LH;
constexpr int LWR = TR + 2 * LH;

// The three fields pass through one LDS window in turn (|S|, dil, then T): 22 KB per block
// instead of 62 KB for all three at once, so about seven blocks share a CU instead of two; each
// point keeps its three pairs of fourth derivatives in registers across the passes.
struct SourceLds {
  double sB[LWR][LWZ];     // the field of this pass
  double sZ[TR][TZ + 4];   // its D2 along xi, columns i0-2 .. i0+TZ+1
  double sR[TR + 4][TZ];   // its D2 along eta, rows j0-2 .. j0+TR+1
  uint8_t cZ[TR][TZ + 4];  // z row codes where the D2 along xi is taken
  uint8_t cR[TR + 4][TZ];  // r row codes where the D2 along eta is taken
};

__device__ __forceinline__ void source_tile(const LadArgs& a, int t, SourceLds& sh) {
  auto& sB = sh.sB;
  auto& sZ = sh.sZ;
  auto& sR = sh.sR;
  auto& cZ = sh.cZ;
  auto& cR = sh.cR;
  constexpr int ROWS = 256 / TZ;                 // rows of owned points per pass of the block
  constexpr int NP = (TR + ROWS - 1) / ROWS;     // owned points per thread
  const Geometry& g = a.g;
  const int nz = g.nz, nr = g.nr, pitch = g.pitch;
  const long long fs = g.fstride;
  if (a.src_mask && !a.src_mask[t]) return;
  const int tzi = t % g.ntz, tri = t / g.ntz;
  const int i0 = a.zcuts[tzi], i1 = a.zcuts[tzi + 1], j0 = a.rcuts[tri], j1 = a.rcuts[tri + 1];
  if (i1 <= a.src_imin) return;  // only feeds zero-ramp columns
  {
    // codes of the points where a second derivative is taken; outside the grid it is zero,
    // which the blank code gives
    for (int k = threadIdx.x; k < TR * (TZ + 4); k += blockDim.x) {
      const int c = k % (TZ + 4), ty = k / (TZ + 4);
      const int i = i0 - 2 + c, j = j0 + ty;
      const bool in = j < j1 && i >= 0 && i < nz;
      const uint32_t m = a.meta[in ? (long long)i + (long long)pitch * j : 0];
      cZ[ty][c] = (uint8_t)(in ? (m & 15u) : ROW_BLANK);
    }
    for (int k = threadIdx.x; k < (TR + 4) * TZ; k += blockDim.x) {
      const int c = k % TZ, w = k / TZ;
      const int i = i0 + c, j = j0 - 2 + w;
      const bool in = i < i1 && j >= 0 && j < nr;
      const uint32_t m = a.meta[in ? (long long)i + (long long)pitch * j : 0];
      cR[w][c] = (uint8_t)(in ? ((m >> 4) & 15u) : ROW_BLANK);
    }
  }
  const int tz = threadIdx.x % TZ;
  // The window element k = threadIdx.x + 256 u of this thread sits at (wz, wr) = (k mod LWZ,
  // k / LWZ), stepped from u = 0 by the constants 256 mod LWZ and 256 / LWZ.  The arrays are
  // addressed flat: element k of sB, sZ, cZ, sR, cR is its k-th double or byte.
  double* const fB = &sB[0][0];
  double* const fZ = &sZ[0][0];
  double* const fR = &sR[0][0];
  const uint8_t* const fcZ = &cZ[0][0];
  const uint8_t* const fcR = &cR[0][0];
  const int wz0 = (int)threadIdx.x % LWZ, wr0 = (int)threadIdx.x / LWZ;
  double dz4[3][NP], dr4[3][NP], Tc[NP];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (f > 0) __syncthreads();  // the previous field's stencils are read
    const double* __restrict__ src = f == 2 ? a.q + 4 * fs : a.A + f * fs;
    // the thread's window elements in groups of SW whose loads are all issued before the first
    // store, so it waits on a memory round trip per group instead of one per element (groups,
    // not all at once: the kernel's 72 VGPRs hold three loads in flight without spilling)
    constexpr int SW = 3;
    constexpr int NG = ((LWZ * LWR + 255) / 256 + SW - 1) / SW;
    int wz = wz0, wr = wr0;
#pragma unroll 1
    for (int gq = 0; gq < NG; ++gq) {
      double v[SW];
      bool vin[SW];
#pragma unroll
      for (int u = 0; u < SW; ++u) {
        const int i = i0 - LH + wz, j = j0 - LH + wr;
        vin[u] = (int)threadIdx.x + 256 * (SW * gq + u) < LWZ * LWR && (unsigned)i < (unsigned)nz &&
                 (unsigned)j < (unsigned)nr;
        v[u] = src[vin[u] ? (long long)i + (long long)pitch * j : 0];
        wz += 256 % LWZ;
        wr += 256 / LWZ;
        if (wz >= LWZ) {
          wz -= LWZ;
          ++wr;
        }
      }
#pragma unroll
      for (int u = 0; u < SW; ++u) {
        const int k = (int)threadIdx.x + 256 * (SW * gq + u);
        if (k < LWZ * LWR) fB[k] = vin[u] ? v[u] : 0.0;
      }
    }
    __syncthreads();
    // sZ[ty][c] and cZ[ty][c] are element k = 36 ty + c; its centre sB[ty + LH][c - 2 + LH]
    // is element k + 10 ty + LH LWZ + LH - 2 of the window
#pragma unroll 1
    for (int k = threadIdx.x; k < TR * (TZ + 4); k += 256) {
      const int ty = k / (TZ + 4);
      const uint32_t zr = fcZ[k];
      fZ[k] = d2_sum(zr, fB + k + (LWZ - TZ - 4) * ty + LH * LWZ + LH - 2, 1);
    }
    // sR[w][c] and cR[w][c] are element k = 32 w + c; its centre sB[w - 2 + LH][c + LH]
#pragma unroll 1
    for (int k = threadIdx.x; k < (TR + 4) * TZ; k += 256) {
      const int w = k / TZ;
      const uint32_t rr = fcR[k];
      fR[k] = d2_sum(rr, fB + k + (LWZ - TZ) * w + (LH - 2) * LWZ + LH, LWZ);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int ty = (int)threadIdx.x / TZ + p * ROWS;
      if (ty >= TR) continue;
      const uint32_t zr = cZ[ty][tz + 2], rr = cR[ty + 2][tz];
      dz4[f][p] = d2_sum(zr, &sZ[ty][tz + 2], 1);
      dr4[f][p] = d2_sum(rr, &sR[ty + 2][tz], TZ);
      if (f == 2) Tc[p] = sB[ty + LH][tz + LH];
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int ty = (int)threadIdx.x / TZ + p * ROWS;
    const int i = i0 + tz, j = j0 + ty;
    if (ty >= TR || i >= i1 || j >= j1) continue;
    const long long gi = (long long)i + (long long)pitch * j;
    const double rho = a.q[gi];
    const double fsw = a.A[gi + 2 * fs];
    const double T = Tc[p];
    const double dz = a.dzl[i], dr = a.drl[j];
    const double dz2 = dz * dz, dr2 = dr * dr;
    a.B[gi] = rho * fabs(dz4[0][p] * dz2 + dr4[0][p] * dr2);
    a.B[gi + fs] = rho * fsw * fabs(dz4[1][p] * dz2 + dr4[1][p] * dr2);
    a.B[gi + 2 * fs] = rho / (a.gam * sqrt(T)) * fabs(dz4[2][p] * dz + dr4[2][p] * dr);
  }
  __syncthreads();  // the window and codes are reused by the next tile
}

// ---- fused source tile: |S|, dil and f_sw from the velocity, then the three D2(D2(.)) pairs.
// For a tile whose |S| points are interior, off the axis and inside the grid
// (classify_lad_source), the strain kernel's interior arithmetic (strain_point) runs here, over
// the tile widened by FH = 4 points on every side: the pairs read |S| and dil over the owned
// rows at the owned columns +-4 and over the owned columns at the owned rows +-4 (the corners
// are computed and not read).  The values equal the strain kernel's bit for bit and never pass
// through memory.  The velocity window, 2 points wider still, is loaded in one round trip; T's
// window is loaded during the strain arithmetic.
constexpr int FH = 4;
constexpr int FWZ = TZ + 2 * FH, FWR = TR + 2 * FH;  // the |S|, dil and T frame
constexpr int VWZ = FWZ + 4, VWR = FWR + 4;          // the velocity window
static_assert(FWZ + 2 * FWR <= 256, "one metric per thread");
// Four blocks per CU: LDS would allow six, but the fused tile keeps 115 VGPRs (at the 96 that
// five need it spills loop-invariant registers and runs 82 us per launch instead of 66).
#ifndef AJET_SRC_BLOCKS
#define AJET_SRC_BLOCKS 4
#endif

// The velocity is read (into registers) before |S| and dil are stored over it: 26 KB per block.
struct FusedLds {
  union {
    struct {
      double vr[VWR][VWZ];
      double vz[VWR][VWZ];
    } v;  // the strain arithmetic's velocity
    struct {
      double f[FWR][FWZ];    // |S|, then T
      double dl[FWR][FWZ];   // dil
      double z[TR][TZ + 4];  // a field's D2 along xi, columns i0-2 .. i0+TZ+1
      double r[TR + 4][TZ];  // its D2 along eta, rows j0-2 .. j0+TR+1
    } d;
  } u;
  double mz[FWZ], mr[FWR], rc[FWR];
};

__device__ __forceinline__ void source_tile_fused(const LadArgs& a, int t, FusedLds& sh) {
  constexpr int ROWS = 256 / TZ;              // rows of owned points per pass of the block
  constexpr int NP = (TR + ROWS - 1) / ROWS;  // owned points per thread
  const Geometry& g = a.g;
  const int nz = g.nz, nr = g.nr, pitch = g.pitch;
  const long long fs = g.fstride;
  if (a.src_mask && !a.src_mask[t]) return;
  const int tzi = t % g.ntz, tri = t / g.ntz;
  const int i0 = a.zcuts[tzi], i1 = a.zcuts[tzi + 1], j0 = a.rcuts[tri], j1 = a.rcuts[tri + 1];
  if (i1 <= a.src_imin) return;  // only feeds zero-ramp columns
  const int tid = (int)threadIdx.x, tz = tid % TZ;
  // frame position (sz, sr) is the point (i0 - FH + sz, j0 - FH + sr); positions a narrow tile
  // does not need are clamped into the grid and go unused
  auto gidx = [&](int i, int j) { return (long long)min(max(i, 0), nz - 1) + (long long)pitch * min(max(j, 0), nr - 1); };
  constexpr int NV = (VWZ * VWR + 255) / 256;
  double vr[NV], vz[NV];
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int k = tid + 256 * u;
    const long long gi = k < VWZ * VWR ? gidx(i0 - FH - 2 + k % VWZ, j0 - FH - 2 + k / VWZ) : gidx(i0, j0);
    vr[u] = a.q[gi + fs];
    vz[u] = a.q[gi + 3 * fs];
  }
  double m = 0.0;
  if (tid < FWZ) m = a.dxidz[min(i0 - FH + tid, nz - 1)];
  else if (tid < FWZ + FWR) m = a.detadr[min(j0 - FH + tid - FWZ, nr - 1)];
  else if (tid < FWZ + 2 * FWR) m = a.rc[min(j0 - FH + tid - FWZ - FWR, nr - 1)];
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int k = tid + 256 * u;
    if (k < VWZ * VWR) {
      (&sh.u.v.vr[0][0])[k] = vr[u];
      (&sh.u.v.vz[0][0])[k] = vz[u];
    }
  }
  if (tid < FWZ) sh.mz[tid] = m;
  else if (tid < FWZ + FWR) sh.mr[tid - FWZ] = m;
  else if (tid < FWZ + 2 * FWR) sh.rc[tid - FWZ - FWR] = m;
  __syncthreads();
  auto strain_at = [&](int sz, int sr, double& S, double& dil, double& fsw) {
    const double* pr = &sh.u.v.vr[sr + 2][sz + 2];
    const double* pz = &sh.u.v.vz[sr + 2][sz + 2];
    const double mz = sh.mz[sz], mr = sh.mr[sr];
    const double dVrdz = d1_sum(ROW_INTERIOR, pr, 1) * mz;
    const double dVzdz = d1_sum(ROW_INTERIOR, pz, 1) * mz;
    const double dVrdr = d1_sum(ROW_INTERIOR, pr, VWZ) * mr;
    const double dVzdr = d1_sum(ROW_INTERIOR, pz, VWZ) * mr;
    strain_point(dVrdz, dVzdz, dVrdr, dVzdr, *pr, sh.rc[sr], false, S, dil, fsw);
  };
  // the owned points, then the ring around them: FH rows above and below the owned columns,
  // FH columns left and right of the owned rows
  constexpr int NRING = 2 * FH * TZ + 2 * FH * TR, NU = (NRING + 255) / 256;
  double fsw[NP], Sv[NP + NU], Dv[NP + NU];
  int at[NP + NU];  // frame element, -1: none
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int sz = tz + FH, sr = tid / TZ + p * ROWS + FH;
    fsw[p] = 0.0;
    at[p] = -1;
    if (sr < TR + FH) {
      strain_at(sz, sr, Sv[p], Dv[p], fsw[p]);
      at[p] = sr * FWZ + sz;
    }
  }
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int k = tid + 256 * u;
    at[NP + u] = -1;
    if (k >= NRING) continue;
    int sz, sr;
    if (k < 2 * FH * TZ) {
      const int kk = k < FH * TZ ? k : k - FH * TZ;
      sr = kk / TZ + (k < FH * TZ ? 0 : FH + TR);
      sz = FH + kk % TZ;
    } else {
      const int kk = k - 2 * FH * TZ, side = kk < FH * TR ? 0 : 1, k2 = kk - side * FH * TR;
      sr = FH + k2 / FH;
      sz = k2 % FH + side * (FH + TZ);
    }
    double unused;
    strain_at(sz, sr, Sv[NP + u], Dv[NP + u], unused);
    at[NP + u] = sr * FWZ + sz;
  }
  __syncthreads();  // the velocity is read: |S| and dil take its place
  // T's window and the owned points' pointwise inputs, in flight during the first pass
  constexpr int NT = (FWZ * FWR + 255) / 256;
  double tv[NT], rho[NP], dr[NP];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int k = tid + 256 * u;
    tv[u] = a.q[4 * fs + (k < FWZ * FWR ? gidx(i0 - FH + k % FWZ, j0 - FH + k / FWZ) : gidx(i0, j0))];
  }
  const double dz = a.dzl[min(i0 + tz, nz - 1)], dz2 = dz * dz;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int ty = tid / TZ + p * ROWS;
    rho[p] = a.q[gidx(i0 + tz, j0 + ty)];
    dr[p] = a.drl[min(j0 + ty, nr - 1)];
  }
#pragma unroll
  for (int u = 0; u < NP + NU; ++u)
    if (at[u] >= 0) {
      (&sh.u.d.f[0][0])[at[u]] = Sv[u];
      (&sh.u.d.dl[0][0])[at[u]] = Dv[u];
    }
  __syncthreads();
  // one field's D2 along xi and eta, then its second pair at the owned points, weighted by the
  // grid spacings (squared for |S| and dil) and summed: the magnitudes ComputeTobeFilteredQuantities
  // scales (ModShockCapturing.f90:204-222)
  double mag[3][NP];
  auto pass = [&](const double (*F)[FWZ], int fi) {
    for (int k = tid; k < TR * (TZ + 4); k += 256) {
      const int ty = k / (TZ + 4), c = k % (TZ + 4);
      sh.u.d.z[ty][c] = d2_sum(ROW_INTERIOR, &F[ty + FH][c + 2], 1);
    }
    for (int k = tid; k < (TR + 4) * TZ; k += 256) {
      const int w = k / TZ, c = k % TZ;
      sh.u.d.r[w][c] = d2_sum(ROW_INTERIOR, &F[w + 2][c + FH], FWZ);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int ty = tid / TZ + p * ROWS;
      if (ty >= TR) continue;
      const double dz4 = d2_sum(ROW_INTERIOR, &sh.u.d.z[ty][tz + 2], 1);
      const double dr4 = d2_sum(ROW_INTERIOR, &sh.u.d.r[ty + 2][tz], TZ);
      mag[fi][p] = fi < 2 ? fabs(dz4 * dz2 + dr4 * (dr[p] * dr[p])) : fabs(dz4 * dz + dr4 * dr[p]);
    }
  };
  pass(sh.u.d.f, 0);
  // |S| has been read: T takes its place
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int k = tid + 256 * u;
    if (k < FWZ * FWR) (&sh.u.d.f[0][0])[k] = tv[u];
  }
  __syncthreads();
  pass(sh.u.d.dl, 1);
  __syncthreads();
  pass(sh.u.d.f, 2);
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int ty = tid / TZ + p * ROWS;
    const int i = i0 + tz, j = j0 + ty;
    if (ty >= TR || i >= i1 || j >= j1) continue;
    const long long gi = (long long)i + (long long)pitch * j;
    const double T = sh.u.d.f[ty + FH][tz + FH];
    a.B[gi] = rho[p] * mag[0][p];
    a.B[gi + fs] = rho[p] * fsw[p] * mag[1][p];
    a.B[gi + 2 * fs] = rho[p] / (a.gam * sqrt(T)) * mag[2][p];
  }
  __syncthreads();  // the frame is reused by the next tile
}

// Fused and other tiles in one launch: list entry t >= 0 is fused tile t, ~t a tile that reads
// the strain kernel's values.
// The arguments are read through the kernel-argument segment where the tiles use them: taken by
// value they are all loaded at the kernel's entry and, live across the tile loop, spill 42 scalar
// registers into vector-register lanes (the launch took 67.2 us against 62.3).
__global__ __launch_bounds__(256, AJET_SRC_BLOCKS) void lad_source_kernel(LadArgs args) {
  (void)args;
  const LadArgs& a = *(const LadArgs*)__builtin_amdgcn_kernarg_segment_ptr();
  __shared__ union {
    SourceLds general;
    FusedLds fused;
  } sh;
  for (int li = xcd_tile((int)blockIdx.x, (int)gridDim.x); li < a.ntiles; li += (int)gridDim.x) {
    const int e = a.tiles[li];
    if (e >= 0) source_tile_fused(a, e, sh.fused);
    else source_tile(a, ~e, sh.general);
  }
}

// ---- the z filter, which the reference applies in place (PerformTGfilter with var = out,
// ModShockCapturing.f90:63, 69, 75): walking i upwards, each point reads the already filtered
// values before it and the unfiltered ones from itself on.  The recursion forgets its start at
// the rate of the largest root of z^4 - c3 z^3 - c2 z^2 - c1 z - c0 (c = interior weights
// before the centre), 0.467, so a line is cut into segments that each start ZWARM points early
// from the unfiltered values: the start error is damped by 0.467^48 < 1.3e-16 before the first
// point a segment writes.  A segment within ZWARM of the line start runs the exact recursion
// from i = 1.  The segment length is chosen per grid (lad_zfilter_segment).
//
// One wave walks 64 lines (one lane each) through its segment.  The columns stream through an
// LDS ring of ZRING columns in chunks of ZCH; a second wave takes the r filter beside the walk.
// The walk keeps the last four filtered values and the next nine inputs in registers (and
// reads the ring 9 points ahead), so an interior step waits only on its own arithmetic; closure rows read their 11 points from the ring (up to 10 behind or 10 ahead).
// While walk chunk k runs, the ring holds ring chunks k .. k + 2 (columns ib - 16 + 16 k ..
// ib + 31 + 16 k), which covers every read of its steps i = ib + 16 k .. ib + 16 k + 15
// (columns i - 10 .. i + 10).
constexpr int ZWARM = 48;
constexpr int ZLINES = LAD_LINES;
constexpr int ZCH = 16;
constexpr int ZRING = 64;
// chunks in flight in the walker's registers: a chunk's loads are issued ZPRE iterations
// before it enters the ring (one measured fastest: the walk of a chunk covers the loads, and a
// second register set costs the walker 40 VGPRs)
#ifndef AJET_ZPRE
#define AJET_ZPRE 1
#endif
constexpr int ZPRE = AJET_ZPRE;


// Walk state of one lane: the last four filtered values and the next nine unfiltered ones.
struct ZWalk {
  double y1, y2, y3, y4;
  double X[9];
};

__device__ __forceinline__ double tg_interior(const ZWalk& st) {
  return (((((((TG0 * st.y4 + TG1 * st.y3) + TG2 * st.y2) + TG3 * st.y1) + TG4 * st.X[0]) + TG3 * st.X[1]) +
           TG2 * st.X[2]) + TG1 * st.X[3]) + TG0 * st.X[4];
}

__device__ __forceinline__ void zwalk_advance(ZWalk& st, double w, double xn) {
  st.y4 = st.y3;
  st.y3 = st.y2;
  st.y2 = st.y1;
  st.y1 = w;
#pragma unroll
  for (int m = 0; m < 8; ++m) st.X[m] = st.X[m + 1];
  st.X[8] = xn;
}

// The interior filter in fused form for the chunks whose points all take it: the symmetric
// weights pair the values either side, and the last filtered value y1 enters last, so a step
// waits on one add and one fused multiply-add of the step before it (the reference's
// left-to-right sum makes that six dependent operations).  It rounds differently from
// tg_interior by a few ulp of the filtered values.
__device__ __forceinline__ double tg_interior_fast(const ZWalk& st) {
  const double t = fma(TG0, st.y4 + st.X[4], fma(TG1, st.y3 + st.X[3], fma(TG2, st.y2 + st.X[2], TG4 * st.X[0])));
  return fma(TG3, st.y1 + st.X[1], t);
}

// 16 steps of a chunk whose points all take the interior filter.  Walk chunk k fills ring
// chunk k + 1, slots 16 ((k + 1) mod 4) .., so KM = (k + 1) mod 4 fixes every slot at compile
// time.
template <int KM>
__device__ __forceinline__ void zwalk_interior(ZWalk& st, double (*ring)[ZLINES + 1], int lane) {
#pragma unroll
  for (int t = 0; t < ZCH; ++t) {
    const int sl = (ZCH * KM + t) & (ZRING - 1);
    const double xn = ring[(sl + 9) & (ZRING - 1)][lane];
    const double w = tg_interior_fast(st);
    ring[sl][lane] = w;
    zwalk_advance(st, w, xn);
  }
}

// The fused z and r filters.  Block (segment, line group, field), two waves: the walker (wave 0)
// takes the group's 64 r-lines wb..wb+63 through the z recursion, one lane a line, and moves the
// chunks from memory into the ring; the writer (wave 1) takes the r filter of the chunk walked
// before (PerformTGfilter along eta, ModShockCapturing.f90:64, 70, 76) over the group's output
// lines [out_lo, out_hi), scales it by C Re (C_kappa Re Pr) and the ramp
// (ModShockCapturing.f90:66-92) and writes mu*, beta* or kappa*.  All loads are the walker's and
// all stores the writer's, so neither waits on the other's memory traffic.  One barrier per
// chunk: in iteration k the walker walks chunk k (ring chunk k + 1, reading ring chunks k ..
// k + 2) and then overwrites ring chunk k - 1 with ring chunk k + 3, whose loads it issued
// ZPRE iterations before; the writer reads ring chunk k (walk chunk k - 1).  Blanked points
// are left out by the filter and keep zero.  Ring codes: z row in the low nibble, r row in the
// high one.
__global__ __launch_bounds__(2 * ZLINES, 2) void lad_zrfilter_kernel(LadArgs a) {
  __shared__ double ring[ZRING][ZLINES + 1];
  // row codes, one byte per [line][slot], four slots a word (written as words, read as bytes)
  __shared__ uint32_t code_w[ZLINES][(ZRING + 4) / 4];
  __shared__ double ramp_s[ZRING / ZCH][ZCH];  // the ramp of walk chunk m's columns (slot m mod 4)
  __shared__ double tg_s[9][11];               // the TG closure rows (c_tg), read by lanes that differ in row
  __shared__ uint32_t rint_s[ZRING / ZCH];     // ring chunk m (slot m mod 4): every point takes the interior r row
  __shared__ unsigned long long colmask_s[ZRING];  // slot s, where rint_s is 0: lines whose r row is not interior
  const Geometry& g = a.g;
  const int nz = g.nz, nr = g.nr, pitch = g.pitch;
  const long long fs = g.fstride;
  // segments before seg_first have zero ramps; block x walks sub-segment x mod zsub of segment
  // seg_first + x / zsub
  const int segi = (int)blockIdx.x / a.zsub + a.seg_first;
  if (a.zr_flags && a.zr_flags[segi + a.nseg * (int)blockIdx.y] != a.gen) return;
  const int seg = a.zseg / a.zsub;
  const int S0 = segi * a.zseg + ((int)blockIdx.x % a.zsub) * seg;
  if (S0 >= nz) return;  // past the end of a shorter last segment
  const int S1 = min(nz, S0 + seg);
  const int jb = a.rgroups[3 * blockIdx.y];
  const int out_lo = a.rgroups[3 * blockIdx.y + 1], out_hi = a.rgroups[3 * blockIdx.y + 2];
  const int f = (int)blockIdx.z;
  const double* __restrict__ x = a.B + f * fs;
  double* __restrict__ out = a.out + f * fs;
  const double C = f == 0 ? a.cmu : (f == 1 ? a.cbeta : a.ckappa);
  const bool walker = threadIdx.x < ZLINES;
  const int lane = (int)threadIdx.x & (ZLINES - 1);
  const int c16 = lane & (ZCH - 1), l4 = lane / ZCH;  // r filter: column c16, lines quarter l4
  const int c8 = lane & 7, l8 = lane / 8;              // chunk transfers: 8 lines x 16 columns per instruction
  const int ib = max(0, S0 - ZWARM);                   // first point the walk filters
  // column of ring slot 0: 16 before ib (closure rows read up to 10 back); ib, the segment
  // starts and the row pitch are multiples of 16, so every chunk is one 128-byte line per line
  const int P0 = ib - ZCH;
  // ring chunk m holds columns P0 + 16 m .. P0 + 16 m + 15; walk chunk k (points ib + 16 k ..)
  // is ring chunk k + 1.  Chunks past the last column the walk reads (S1 + 9) load the last
  // chunk it reads again instead (cache hits): loads stay unconditional, so none waits on a
  // branch.
  const int mlast = (S1 + 9 - P0) / ZCH;
  // a chunk in the walker's registers: lane (c8, l8) holds columns 2 c8, 2 c8 + 1 of lines
  // 8 r + l8, lane l the row codes of line l (16 bytes), lane c16 the ramp of column c16
  struct Chunk {
    double2 v[ZLINES / 8];
    uint4 cz;
    double ramp;
  };
  auto col = [&](int m) { return P0 + min(m, mlast) * ZCH; };  // first column ring chunk m loads
  auto code = [&](int l, int sl) { return (uint32_t)reinterpret_cast<const uint8_t*>(&code_w[l][0])[sl]; };
  unsigned long long interior = 0ull;  // walker: bit (m mod 64) -- every z code of ring chunk m is interior
  // Loads read clamped points (inside the row's pitch) as they are; the store replaces the
  // points outside the grid by zero and the blank code.  No select between the two keeps a load
  // in flight across the loop's back edge.
  auto load_chunk = [&](Chunk& c, int m) {
    const int q = col(m);
#pragma unroll
    for (int r = 0; r < ZLINES / 8; ++r) {
      const int j = min(jb + 8 * r + l8, nr - 1), i = min(max(q + 2 * c8, 0), pitch - 2);
      c.v[r] = *reinterpret_cast<const double2*>(x + ((long long)i + (long long)pitch * j));
    }
    // the codes of columns q .. q + 15 (a multiple of 16, like the pitch: inside one row)
    const int cq = min(max(q, 0), pitch - ZCH), cj = min(jb + lane, nr - 1);
    c.cz = *reinterpret_cast<const uint4*>(a.codes + ((long long)cq + (long long)pitch * cj));
    c.ramp = a.ramp[min(max(q + c16, 0), nz - 1)];
  };
  auto store_chunk = [&](const Chunk& c, int m) {
    const int q = col(m);
    const int slot = (m * ZCH + 2 * c8) & (ZRING - 1);
    const int i = q + 2 * c8;
    const bool in0 = i >= 0 && i < nz, in1 = i + 1 >= 0 && i + 1 < nz;
#pragma unroll
    for (int r = 0; r < ZLINES / 8; ++r) {
      const int l = 8 * r + l8;
      const bool jin = jb + l < nr;
      ring[slot][l] = (jin && in0) ? c.v[r].x : 0.0;
      ring[slot + 1][l] = (jin && in1) ? c.v[r].y : 0.0;
    }
    // codes: a chunk wholly outside the grid's columns (q < 0 or q >= pitch) or lines takes the
    // blank code; the columns nz .. pitch - 1 hold it already
    const bool cin = q >= 0 && q < pitch && jb + lane < nr;
    constexpr uint32_t BLANK4 = 0x01010101u * (ROW_BLANK | (ROW_BLANK << 4));
    const uint32_t w0 = cin ? c.cz.x : BLANK4, w1 = cin ? c.cz.y : BLANK4, w2 = cin ? c.cz.z : BLANK4,
                   w3 = cin ? c.cz.w : BLANK4;
    uint32_t* const cw = &code_w[lane][((m * ZCH) & (ZRING - 1)) / 4];
    cw[0] = w0;
    cw[1] = w1;
    cw[2] = w2;
    cw[3] = w3;
    if (lane < ZCH) ramp_s[(m + 3) & 3][lane] = c.ramp;  // ring chunk m is walk chunk m - 1
    // every code interior in z (low nibbles 4), and in r (high nibbles 4)
    constexpr uint32_t ZMASK = 0x0F0F0F0Fu, ZINT = 0x01010101u * ROW_INTERIOR;
    constexpr uint32_t RMASK = 0xF0F0F0F0u, RINT = 0x01010101u * (ROW_INTERIOR << 4);
    const bool zin = (w0 & ZMASK) == ZINT && (w1 & ZMASK) == ZINT && (w2 & ZMASK) == ZINT && (w3 & ZMASK) == ZINT;
    const bool rin = (w0 & RMASK) == RINT && (w1 & RMASK) == RINT && (w2 & RMASK) == RINT && (w3 & RMASK) == RINT;
    const bool rall = __all(rin);
    if (lane == 0) rint_s[m & 3] = rall ? 1u : 0u;
    if (!rall) {
      // column c of the chunk: bit l set where line l takes a closure or blanked r row
      unsigned long long mine = 0ull;
#pragma unroll
      for (int cc = 0; cc < ZCH; ++cc) {
        const uint32_t wv = cc < 4 ? w0 : (cc < 8 ? w1 : (cc < 12 ? w2 : w3));
        const unsigned long long bm = __ballot(((wv >> (8 * (cc & 3) + 4)) & 15u) != ROW_INTERIOR);
        mine = lane == cc ? bm : mine;
      }
      if (lane < ZCH) colmask_s[((m * ZCH) & (ZRING - 1)) + lane] = mine;
    }
    const unsigned long long bit = 1ull << (m & 63);
    interior = __all(zin) ? (interior | bit) : (interior & ~bit);
  };
  for (int k = (int)threadIdx.x; k < 99; k += 2 * ZLINES) (&tg_s[0][0])[k] = (&c_tg[0][0])[k];
  Chunk ch[ZPRE];
  ZWalk st;
  if (walker) {
    Chunk c2;
    load_chunk(ch[0], 0);
    load_chunk(c2, 1);
    store_chunk(ch[0], 0);
    load_chunk(ch[0], 2);
    store_chunk(c2, 1);
    store_chunk(ch[0], 2);
#pragma unroll
    for (int u = 0; u < ZPRE; ++u) load_chunk(ch[u], 3 + u);  // stored in iteration u
    // the walk's start from the ring this wave has just written
    const int sl = ZCH;  // slot of ib
    st.y1 = ring[sl - 1][lane];
    st.y2 = ring[sl - 2][lane];
    st.y3 = ring[sl - 3][lane];
    st.y4 = ring[sl - 4][lane];
#pragma unroll
    for (int m = 0; m < 9; ++m) st.X[m] = ring[sl + m][lane];
  }
  // general step: closure rows, blanked points and the interior filter chosen per point
  auto general_step = [&](int i) {
    const int sl = (i - P0) & (ZRING - 1);
    const double xn = ring[(sl + 9) & (ZRING - 1)][lane];
    const uint32_t zr = code(lane, sl) & 15u;
    double w = tg_interior(st);
    if (zr != ROW_INTERIOR) {
      if (zr == ROW_BLANK) {
        w = st.X[0];  // blanked points keep their value (ModTGfilter.f90:17)
      } else {
        // closure row: 11 ring columns from base = -row (rows 0-3) or -row-2 (rows 5-8)
        const int base = zr < 4 ? -(int)zr : -(int)zr - 2;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 11; ++m) acc = acc + tg_s[zr][m] * ring[(sl + base + m) & (ZRING - 1)][lane];
        w = acc;
      }
    }
    ring[sl][lane] = w;
    zwalk_advance(st, w, xn);
  };
  // the r filter of one column: lane (c16, l4) takes lines j0 = out_lo + l4 * per .. j1 - 1 of
  // column c16.  Interior rows take the interior weights on a sliding window of nine lines (the
  // next line read one step ahead).  Where the chunk holds a closure or blanked row (rint_s,
  // colmask_s), the four lanes of a column then share the column's such rows in turn (the first
  // to lane l4 = 0, ...) with their closure weights, or zero, so the rows at the grid's r ends
  // take one pass of the wave; the interior pass leaves those points alone.
  const int nout = out_hi - out_lo;
  const int per = (nout + 3) / 4;
  auto rfilter_chunk = [&](int c0, int n, double ramp, bool allint) {
    const int ii = c0 + c16;
    if (c16 >= n || ii < S0 || ii >= S1) return;  // the same for the four lanes of a column
    const int sl = (ii - P0) & (ZRING - 1);
    const int j0 = out_lo + l4 * per;
    const int j1 = min(out_hi, j0 + per);
    const unsigned long long cm = allint ? 0ull : colmask_s[sl];  // bit j - jb: line j is left to the fix-up
    // An interior row's nine lines lie inside the grid and the group (the groups are built so),
    // so the clamped reads give every value an interior row uses unclamped.
    auto win = [&](int j) { return ring[sl][min(max(j - jb, 0), ZLINES - 1)]; };
    // this lane's rows left to the fix-up (bit t: line j0 + t; per <= 16): those at the two ends
    // of its range are cut off the interior pass, which then runs without a test per row; rows
    // between interior ones (walls inside the group) take the pass that tests every row
    const int span = max(0, j1 - j0);
    const uint32_t sk = (uint32_t)(cm >> min(j0 - jb, 63)) & ((1u << span) - 1u);
    const int lead = min(span, __builtin_ctz(~sk));
    const int trail = span > 0 ? min(span - lead, (int)__builtin_clz(~(sk << (32 - span)))) : 0;
    const int mid = span - lead - trail;
    const bool clean = ((sk >> lead) & ((1u << mid) - 1u)) == 0u;
    auto pass = [&](int ja, int jz, auto skipping) {
      double w[9];
#pragma unroll
      for (int m = 0; m < 8; ++m) w[m + 1] = win(ja - 4 + m);
      double nx = win(ja + 4);
      double* op = out + ((long long)ii + (long long)pitch * ja);
      uint32_t skip = sk;
#pragma unroll 2
      for (int j = ja; j < jz; ++j) {
#pragma unroll
        for (int m = 0; m < 8; ++m) w[m] = w[m + 1];
        w[8] = nx;
        nx = win(j + 5);
        // symmetric weights paired, fused (a few ulp from the reference's left-to-right sum)
        double v = fma(TG0, w[0] + w[8], fma(TG1, w[1] + w[7], fma(TG2, w[2] + w[6], fma(TG3, w[3] + w[5], TG4 * w[4]))));
        v = C * v;
        v = v * ramp;
        if (!decltype(skipping)::value || (skip & 1u) == 0u) *op = v;
        skip >>= 1;
        op += pitch;
      }
    };
    if (clean) pass(j0 + lead, j1 - trail, std::false_type{});
    else pass(j0, j1, std::true_type{});
    if (!allint) {
      auto val = [&](int j) {  // z-filtered value of line j (zero outside the grid)
        const int l = j - jb;
        return (j >= 0 && j < nr && l >= 0 && l < ZLINES) ? ring[sl][l] : 0.0;
      };
      // the column's such rows among the group's outputs, every fourth from the l4-th
      const int lo = out_lo - jb, hi = out_hi - jb;  // 0 <= lo < hi <= 64
      const unsigned long long outm = (hi >= 64 ? ~0ull : ((1ull << hi) - 1ull)) & ~((1ull << lo) - 1ull);
      unsigned long long m = cm & outm;
      for (int q = 0; q < l4; ++q) m &= m - 1ull;
      while (m != 0ull) {
        const int j = jb + __builtin_ctzll(m);
        const uint32_t rr = code(j - jb, sl) >> 4;
        double v = 0.0;
        if (rr != ROW_BLANK) {
          const int base = rr < 4 ? -(int)rr : -(int)rr - 2;
#pragma unroll
          for (int q = 0; q < 11; ++q) v = v + tg_s[rr][q] * val(j + base + q);
          v = C * v;
          v = v * ramp;
        }
        out[(long long)ii + (long long)pitch * j] = v;
#pragma unroll
        for (int q = 0; q < 4; ++q) m &= m - 1ull;  // this lane's next row of the column
      }
    }
  };
  __syncthreads();
  const int nchunks = (S1 - ib + ZCH - 1) / ZCH;
  // iteration k: the walker walks chunk k (k < nchunks) and moves a chunk in, the writer
  // filters chunk k - 1 in r (k >= 1); both take every iteration's barrier.  The iterations run
  // in groups of ZPRE, so that the register set each moves is fixed at compile time; those past
  // k = nchunks only move chunks past the last one read.
  for (int k0 = 0; k0 <= nchunks; k0 += ZPRE)
#pragma unroll
  for (int u = 0; u < ZPRE; ++u) {
    const int k = k0 + u;
    if (walker) {
      if (k < nchunks) {
        const int c0 = ib + k * ZCH;  // first point of this chunk
        const int n = min(ZCH, S1 - c0);
        const bool fast = n == ZCH && ((interior >> ((k + 1) & 63)) & 1ull);
        if (fast) {
          switch ((k + 1) & 3) {
            case 0: zwalk_interior<0>(st, ring, lane); break;
            case 1: zwalk_interior<1>(st, ring, lane); break;
            case 2: zwalk_interior<2>(st, ring, lane); break;
            default: zwalk_interior<3>(st, ring, lane); break;
          }
        } else {
          for (int t = 0; t < n; ++t) general_step(c0 + t);
        }
      }
      store_chunk(ch[u], k + 3);  // over ring chunk k - 1, filtered in r in the iteration before
      load_chunk(ch[u], k + 3 + ZPRE);
    } else if (k >= 1) {
      const int c0 = ib + (k - 1) * ZCH;  // first point of walk chunk k - 1
      rfilter_chunk(c0, min(ZCH, S1 - c0), ramp_s[(k - 1) & 3][c16], rint_s[k & 3] != 0u);
    }
    __syncthreads();
  }
}

// ---- learned mode: the dilatation sensor and the tile masks of the flagged filter blocks.
// The sensor covers columns [ifirst, nz), where the ramp exceeds ramp_min (it rises with z),
// one point a thread.
__global__ __launch_bounds__(256) void lad_sensor_kernel(LadArgs a, uint32_t* __restrict__ flags,
                                                         uint32_t* __restrict__ any,
                                                         const int* __restrict__ line_group, double thr,
                                                         double ramp_min, int ifirst) {
  const Geometry& g = a.g;
  const int nz = g.nz, nr = g.nr, pitch = g.pitch, w = nz - ifirst;
  const long long fs = g.fstride;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)w * nr) return;
  const int i = ifirst + (int)(k % w), j = (int)(k / w);
  if (!(a.ramp[i] > ramp_min)) return;
  const int im = max(i - 1, 0), ip = min(i + 1, nz - 1), jm = max(j - 1, 0), jp = min(j + 1, nr - 1);
  auto at = [&](int v, int ii, int jj) { return a.q[(long long)v * fs + ii + (long long)pitch * jj]; };
  const double dVzdz = 0.5 * (at(3, ip, j) - at(3, im, j)) * a.dxidz[i];
  const double dVrdr = 0.5 * (at(1, i, jp) - at(1, i, jm)) * a.detadr[j];
  const double r = a.rc[j];
  const double dil = dVzdz + dVrdr + (r > 0.0 ? at(1, i, j) / r : dVrdr);  // 2 dVr/dr on the axis
  const double s = fmax(0.0, -dil) * fmax(a.dzl[i], a.drl[j]) / sqrt(at(4, i, j));
  if (s > thr) {
    // the step's stamp: every writer stores the same value, and older stamps need no clearing
    uint32_t* f = &flags[line_group[j] * a.nseg + i / a.zseg];
    if (*f != a.gen) *f = a.gen;
    if (*any != a.gen) *any = a.gen;
  }
}

// One thread: whether the sensor flagged any block, then the stamp, into host memory that the
// host polls (fine-grained, so the stores reach it without a cache flush).  A kernel of its own:
// the sensor's blocks publishing through a last-block counter need a device-scope release in
// every block, an L2 write-back each, which took the learned step from 0.70 to 1.39 ms.
__global__ void lad_publish_kernel(const uint32_t* __restrict__ any, uint32_t gen, uint32_t* host) {
  if (threadIdx.x != 0) return;
  __hip_atomic_store(&host[1], *any == gen ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&host[0], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A flagged block (segment sg, group gr) reads the sources over columns [ib - 16, S1 + 25)
// (ib = S0 - warm-up) and lines [wb, wb + 64); a source tile reads the strain rate and
// dilatation up to 7 points beyond itself, and a tile spans at most TZ x TR points.
__global__ __launch_bounds__(256) void lad_mask_kernel(LadArgs a, const uint32_t* __restrict__ flags,
                                                       uint8_t* __restrict__ src_mask, uint8_t* __restrict__ str_mask) {
  const Geometry& g = a.g;
  const int seg = a.zseg;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= g.ntz * g.ntr) return;
  const int tzi = t % g.ntz, tri = t / g.ntz;
  const int i0 = a.zcuts[tzi], i1 = a.zcuts[tzi + 1], j0 = a.rcuts[tri], j1 = a.rcuts[tri + 1];
  constexpr int DZ = TZ + 8, DR = TR + 8;
  bool src = false, str = false;
  const int sg_lo = max(0, (i0 - DZ - 32) / seg - 1), sg_hi = min(a.nseg - 1, (i1 + DZ + 16 + ZWARM) / seg + 1);
  for (int gr = 0; gr < a.nrgroups; ++gr) {
    const int wb = a.rgroups[3 * gr];
    const bool rsrc = j0 < wb + ZLINES && j1 > wb, rstr = j0 < wb + ZLINES + DR && j1 > wb - DR;
    if (!rstr) continue;
    for (int sg = sg_lo; sg <= sg_hi; ++sg) {
      if (flags[sg + a.nseg * gr] != a.gen) continue;
      const int S0 = sg * seg, S1 = min(g.nz, S0 + seg), ib = max(0, S0 - ZWARM);
      const int zlo = ib - 16, zhi = S1 + 32;
      src = src || (rsrc && i0 < zhi && i1 > zlo);
      str = str || (i0 < zhi + DZ && i1 > zlo - DZ);
    }
  }
  src_mask[t] = src ? 1 : 0;
  str_mask[t] = str ? 1 : 0;
}

// Learned mode: the blocks the reference's LAD wrote at stamp gen get the network's values
// back.  A z/r filter block (segment sg, group gr) writes columns [sg seg, (sg + 1) seg) of the
// group's output lines and nothing else.
__global__ __launch_bounds__(256) void lad_restore_kernel(LadArgs a, const uint32_t* __restrict__ flags, uint32_t gen,
                                                          const double* __restrict__ src, double* __restrict__ dst) {
  const int sg = (int)blockIdx.x, gr = (int)blockIdx.y;
  if (flags[sg + a.nseg * gr] != gen) return;
  const Geometry& g = a.g;
  const int S0 = sg * a.zseg, S1 = min(g.nz, S0 + a.zseg);
  const int lo = a.rgroups[3 * gr + 1], hi = a.rgroups[3 * gr + 2];
  const int w = S1 - S0, n = w * (hi - lo);
  for (int f = 0; f < 3; ++f)
    for (int k = (int)threadIdx.x; k < n; k += (int)blockDim.x) {
      const long long gi = (long long)f * g.fstride + (S0 + k % w) + (long long)g.pitch * (lo + k / w);
      dst[gi] = src[gi];
    }
}

hipError_t launch_lad_restore(const LadArgs& a, const uint32_t* flags, uint32_t gen, const double* src, double* dst,
                              hipStream_t s) {
  if (a.nseg <= 0 || a.nrgroups <= 0) return hipSuccess;
  hipLaunchKernelGGL(lad_restore_kernel, dim3((unsigned)a.nseg, (unsigned)a.nrgroups), dim3(256), 0, s, a, flags, gen,
                     src, dst);
  return hipGetLastError();
}

int lad_zrfilter_slots(int device) {
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lad_zrfilter_kernel, 2 * ZLINES, 0) != hipSuccess)
    return 0;
  return cus * per_cu;
}

// Filter blocks from segment seg_first on (the ramp is zero before it).
static int zfilter_first_segment(const double* ramp, int nz, int seg) {
  int ifirst = nz;
  for (int i = 0; i < nz; ++i)
    if (ramp[i] != 0.0) {
      ifirst = i;
      break;
    }
  const int nseg = (nz + seg - 1) / seg;
  int seg_first = 0;
  while (seg_first < nseg && min(nz, (seg_first + 1) * seg) <= ifirst) ++seg_first;
  return seg_first;
}

// The segment length: a multiple of ZCH and at least 64.  A block (one wave, held to a few per
// CU by its LDS ring) walks its segment and ZWARM warm-up columns, and the blocks beyond the
// slots that can be resident at once run in a further round; so the length that minimizes
// rounds x (length + ZWARM) over the blocks the ramp leaves to run is taken, the shorter on a tie.
// With slots <= 0 (unknown) every length takes one round and the shortest is taken.
int lad_zfilter_segment(const double* ramp, int nz, int ngroups, int slots) {
  int best = 64;
  long long best_cost = -1;
  const int top = max(64, (nz + ZCH - 1) / ZCH * ZCH);
  for (int seg = 64; seg <= top; seg += ZCH) {
    const long long blocks = 3ll * ngroups * ((nz + seg - 1) / seg - zfilter_first_segment(ramp, nz, seg));
    const long long rounds = slots > 0 ? max(1ll, (blocks + slots - 1) / slots) : 1;
    const long long cost = rounds * (min(seg, nz) + ZWARM);
    if (best_cost < 0 || cost < best_cost) {
      best = seg;
      best_cost = cost;
    }
  }
  return best;
}

// (eight: a flagged block then walks 16 + 48 columns instead of 128 + 48; on the npr 1.5 jet,
// with 0.7 % of the blocks flagged, the learned step took 0.950, 0.900 and 0.885 ms with one,
// four and eight)
#ifndef AJET_FALLBACK_ZSUB
#define AJET_FALLBACK_ZSUB 8
#endif
int lad_fallback_zsub(int seg, int parts) {
  for (int n = parts > 0 ? parts : AJET_FALLBACK_ZSUB; n > 1; n /= 2)
    if (seg % (ZCH * n) == 0) return n;
  return 1;
}

// The work the exactly-zero ramp makes unnecessary (the ramp rises with z).  A filter segment
// whose columns all have ramp 0 writes only +-0; the first segment that does not starts its
// walk ZWARM columns early and reads the sources from 16 columns before that; a source tile
// reads the strain rate up to LH = 7 columns beyond itself and spans at most TZ columns.
void lad_skip_bounds(const double* ramp, int nz, int seg, int& seg_first, int& src_imin, int& str_imin) {
  seg_first = zfilter_first_segment(ramp, nz, seg);
  const int ib = max(0, seg_first * seg - ZWARM);
  src_imin = ib - ZCH;
  str_imin = src_imin - TZ - LH - 1;
}

hipError_t launch_lad_sensor(const LadArgs& a, uint32_t* flags, uint32_t* any, const int* line_group, double thr,
                             double ramp_min, int ifirst, uint32_t* host, hipStream_t s) {
  const long long n = (long long)(a.g.nz - ifirst) * a.g.nr;
  if (n > 0)
    hipLaunchKernelGGL(lad_sensor_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, flags, any,
                       line_group, thr, ramp_min, ifirst);
  hipLaunchKernelGGL(lad_publish_kernel, dim3(1), dim3(64), 0, s, any, a.gen, host);
  return hipGetLastError();
}

hipError_t launch_lad_masks(const LadArgs& a, const uint32_t* flags, uint8_t* src_mask, uint8_t* str_mask,
                            hipStream_t s) {
  const int nt = a.g.ntz * a.g.ntr;
  hipLaunchKernelGGL(lad_mask_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a, flags, src_mask,
                     str_mask);
  return hipGetLastError();
}

hipError_t upload_lad_tables(const double sbp1[9][6], const double sbp2[9][6], const double tg[9][11], hipStream_t s) {
  hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sbp1), sbp1, sizeof(double) * 54, 0, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sbp2), sbp2, sizeof(double) * 54, 0, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tg), tg, sizeof(double) * 99, 0, hipMemcpyHostToDevice, s);
  return e;
}

hipError_t launch_lad(const LadArgs& a, hipStream_t s) {
  const Geometry& g = a.g;
  // masked (learned mode): most tiles return at once, so a grid of at most MASKED_GRID blocks
  // walks the list instead of one block per tile
  constexpr int MASKED_GRID = 1024;
  const bool masked = a.src_mask != nullptr;
  auto grid = [&](int n) { return dim3((unsigned)(((masked && n > MASKED_GRID ? MASKED_GRID : n) + 7) / 8 * 8)); };
  // the interior and the boundary tiles of each kernel in one launch: a boundary tile's launch
  // of its own took 10-25 us, most of it latency with a few hundred blocks on 256 CUs
  if (a.nstr > 0) hipLaunchKernelGGL(lad_strain_kernel, grid(a.nstr), dim3(256), 0, s, a);
  if (a.ntiles > 0) hipLaunchKernelGGL(lad_source_kernel, grid(a.ntiles), dim3(256), 0, s, a);
  const int nseg = (g.nz + a.zseg - 1) / a.zseg;
  if (nseg > a.seg_first)
    hipLaunchKernelGGL(lad_zrfilter_kernel, dim3((nseg - a.seg_first) * a.zsub, a.nrgroups, 3), dim3(2 * ZLINES), 0, s,
                       a);
  return hipGetLastError();
}

}  // namespace ajet

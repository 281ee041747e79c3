// Host-side setup of the time-step context: SBP tables, closure patches from the bc lines,
// 1-D metrics of a tensor-product grid, and the per-point metadata word the stage kernel
// reads.  Pure CPU; compiled with -ffp-contract=off so the metrics round exactly as the
// reference's unoptimised build computes them.
#include <cmath>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "ajet_internal.h"

namespace ajet {

// SBP_operator (ModOperator.f90:5-121): D1 and D2 with 4-row closures, and the truncated
// Gaussian filter whose closure rows 1-3 are single-precision literals in the reference
// (ModOperator.f90:94-97), i.e. float values widened to double.
void sbp_tables(double A[9][6], double B[9][6], double F[9][11]) {
  std::memset(A, 0, sizeof(double) * 54);
  std::memset(B, 0, sizeof(double) * 54);
  std::memset(F, 0, sizeof(double) * 99);
  const double a[4][6] = {{-(24.0 / 17.0), 59.0 / 34.0, -(4.0 / 17.0), -(3.0 / 34.0), 0.0, 0.0},
                          {-1.0 / 2.0, 0.0, 1.0 / 2.0, 0.0, 0.0, 0.0},
                          {4.0 / 43.0, -(59.0 / 86.0), 0.0, 59.0 / 86.0, -(4.0 / 43.0), 0.0},
                          {3.0 / 98.0, 0.0, -(59.0 / 98.0), 0.0, 32.0 / 49.0, -(4.0 / 49.0)}};
  const double b[4][6] = {{2.0, -5.0, 4.0, -1.0, 0.0, 0.0},
                          {1.0, -2.0, 1.0, 0.0, 0.0, 0.0},
                          {-(4.0 / 43.0), 59.0 / 43.0, -(110.0 / 43.0), 59.0 / 43.0, -(4.0 / 43.0), 0.0},
                          {-(1.0 / 49.0), 0.0, 59.0 / 49.0, -(118.0 / 49.0), 64.0 / 49.0, -(4.0 / 49.0)}};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 6; ++c) { A[r][c] = a[r][c]; B[r][c] = b[r][c]; }
  for (int r = 5; r <= 8; ++r)
    for (int c = 0; c <= 5; ++c) { A[r][c] = -1.0 * A[8 - r][5 - c]; B[r][c] = 1.0 * B[8 - r][5 - c]; }
  const double ai[5] = {1.0 / 12.0, -(2.0 / 3.0), 0.0, 2.0 / 3.0, -(1.0 / 12.0)};
  const double bi[5] = {-(1.0 / 12.0), 4.0 / 3.0, -(5.0 / 2.0), 4.0 / 3.0, -(1.0 / 12.0)};
  for (int c = 0; c < 5; ++c) { A[4][c] = ai[c]; B[4][c] = bi[c]; }

  F[0][0] = 1.0;
  const float f1[7] = {0.085777408970000f, 0.722371828476000f, 0.356848072173000f, -0.223119093072000f,
                       0.057347064865000f, 0.000747264596000f, 0.000027453993000f};
  const float f2[7] = {-0.032649010764000f, 0.143339502575000f, 0.726678822020000f, 0.294622121167000f,
                       -0.186711738069000f, 0.062038376258000f, -0.007318073189000f};
  const float f3[11] = {0.000054596010000f, -0.042124772446000f, 0.173103107841000f, 0.700384128648000f,
                        0.276543612935000f, -0.131223506571000f, 0.023424966418000f, -0.013937561779000f,
                        0.024565095706000f, -0.013098287852000f, 0.002308621090000f};
  for (int c = 0; c < 7; ++c) { F[1][c] = (double)f1[c]; F[2][c] = (double)f2[c]; }
  for (int c = 0; c < 11; ++c) F[3][c] = (double)f3[c];
  for (int r = 5; r <= 8; ++r)
    for (int c = 0; c <= 10; ++c) F[r][c] = 1.0 * F[8 - r][10 - c];
  const double fi[9] = {107.0 / 103680.0, 149.0 / 12960.0, 1997.0 / 25920.0, 3091.0 / 12960.0, 3565.0 / 10368.0,
                        3091.0 / 12960.0, 1997.0 / 25920.0, 149.0 / 12960.0, 107.0 / 103680.0};
  for (int c = 0; c < 9; ++c) F[4][c] = fi[c];
}

// DOT_PRODUCT as the reference evaluates it: one rounding per product and per partial sum,
// summed from the first element.
static double dot_seq(const double* a, const double* v, int n) {
  double s = 0.0;
  for (int m = 0; m < n; ++m) s = s + a[m] * v[m];
  return s;
}

// ComputeMetric (ModDeriv.f90:241-382) along one grid line, every end a physical boundary.
static void metric_line(const double A[9][6], const std::vector<double>& x, std::vector<double>& out) {
  const int n = (int)x.size();
  out.assign(n, 0.0);
  for (int i = 4; i <= n - 5; ++i) out[i] = dot_seq(A[4], &x[i - 2], 5);
  for (int i = 0; i < 4; ++i) out[i] = dot_seq(A[i], &x[0], 6);
  for (int i = n - 4; i < n; ++i) out[i] = dot_seq(A[8 - (n - 1) + i], &x[n - 6], 6);
}

// CreateAllPatches + PatchCreation for a single domain (ModPatch.f90:117-309).
int build_patches(int nz, int nr, const ajet_bc* lines, int nlines, std::vector<ajet_patch>& out) {
  out.clear();
  for (int l = 0; l < nlines; ++l) {
    int64_t is = lines[l].i_start, ie = lines[l].i_end, js = lines[l].j_start, je = lines[l].j_end;
    const int64_t dir = lines[l].bc_dir;
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
    // the first and last point (i outer, j inner) of the strip inside the grid
    bool found = false;
    int64_t fi = 0, fj = 0, li = 0, lj = 0;
    for (int64_t i = 1; i <= nz; ++i) {
      if ((i - pis) * (i - pie) > 0) continue;
      for (int64_t j = 1; j <= nr; ++j) {
        if ((j - pjs) * (j - pje) > 0) continue;
        if (!found) { fi = i; fj = j; found = true; }
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
    out.push_back(P);
  }
  return (int)out.size();
}

// Boundary line of a patch as BoundaryConditions walks it (ModBC.f90:29-52).
static void bc_line(const ajet_patch& P, int64_t& is, int64_t& ie, int64_t& js, int64_t& je) {
  if (P.bc_dir == 1) { is = P.i_start; ie = P.i_start; js = P.j_start; je = P.j_end; }
  else if (P.bc_dir == -1) { is = P.i_end; ie = P.i_end; js = P.j_start; je = P.j_end; }
  else if (P.bc_dir == 2) { is = P.i_start; ie = P.i_end; js = P.j_start; je = P.j_start; }
  else { is = P.i_start; ie = P.i_end; js = P.j_end; je = P.j_end; }
}

// Whether patch P acts at point (i, j) (1-based) in BoundaryConditions: the types with a
// branch there, skipping what the reference skips on the axis row.
static bool bc_acts(const ajet_patch& P, int64_t j) {
  const int64_t adir = P.bc_dir < 0 ? -P.bc_dir : P.bc_dir;
  switch (P.bc) {
    case AJET_BC_SAT_FARFIELD: return j != 1 && adir == 1;
    case AJET_BC_NSCBC_NONREFLECTING: return j != 1;
    case AJET_BC_SAT_NOSLIP_WALL:
    case AJET_BC_NSCBC_NOSLIP_WALL: return adir == 1 || adir == 2;
    default: return false;
  }
}

int build_setup(const ajet_params& p, const double* z, const double* r, const int32_t* iblank,
                const ajet_bc* lines, int nlines, Setup& s, std::string& err) {
  const int nz = p.nz, nr = p.nr;
  if (nz < 8 || nr < 8) { err = "grid must be at least 8 x 8"; return AJET_EINVAL; }
  const size_t ng = (size_t)nz * nr;
  // tensor-product grid: z = z(i), r = r(j), bit for bit
  s.zc.assign(z, z + nz);
  s.rc.resize(nr);
  for (int j = 0; j < nr; ++j) s.rc[j] = r[(size_t)nz * j];
  for (int j = 0; j < nr; ++j)
    for (int i = 0; i < nz; ++i) {
      const size_t g = (size_t)i + (size_t)nz * j;
      if (z[g] != s.zc[i] || r[g] != s.rc[j]) {
        err = "grid is not a tensor-product grid (z must depend on i only and r on j only)";
        return AJET_EGRID;
      }
    }
  double A[9][6], B[9][6], F[9][11];
  sbp_tables(A, B, F);
  // metrics, ModSolver.f90:57-76
  std::vector<double> dzdxi, drdeta;
  metric_line(A, s.zc, dzdxi);
  metric_line(A, s.rc, drdeta);
  s.dxidz.resize(nz);
  s.detadr.resize(nr);
  for (int i = 0; i < nz; ++i) s.dxidz[i] = 1.0 / dzdxi[i];
  for (int j = 0; j < nr; ++j) s.detadr[j] = 1.0 / drdeta[j];
  metric_line(A, s.dxidz, s.ddxidxidz);
  metric_line(A, s.detadr, s.ddetadetadr);

  const int np = build_patches(nz, nr, lines, nlines, s.patches);
  if (np < 0) { err = "bc line with a bc_dir other than +-1, +-2"; return AJET_EPATCH; }
  if (np > MAX_PATCH) { err = "too many boundary patches"; return AJET_EPATCH; }
  // patch maps, last patch in file order wins (ModPatch.f90:214-242)
  std::vector<int> which_i(ng, 0), which_j(ng, 0);
  for (int k = 0; k < np; ++k) {
    const ajet_patch& P = s.patches[k];
    if (P.i_start < 1 || P.i_end > nz || P.j_start < 1 || P.j_end > nr) {
      err = "a bc line gives a closure patch outside the grid";
      return AJET_EPATCH;
    }
    for (int64_t i = P.i_start; i <= P.i_end; ++i)
      for (int64_t j = P.j_start; j <= P.j_end; ++j) {
        const size_t g = (size_t)(i - 1) + (size_t)nz * (j - 1);
        if (P.bc_dir == 1 || P.bc_dir == -1) which_i[g] = k + 1;
        if (P.bc_dir == 2 || P.bc_dir == -2) which_j[g] = k + 1;
      }
  }
  // ordered lists of acting patches per point
  std::vector<std::vector<uint8_t>> seqpt(ng);
  for (int k = 0; k < np; ++k) {
    const ajet_patch& P = s.patches[k];
    if (P.bc == AJET_BC_AXIS) continue;
    int64_t is, ie, js, je;
    bc_line(P, is, ie, js, je);
    for (int64_t i = is; i <= ie; ++i)
      for (int64_t j = js; j <= je; ++j) {
        if (!bc_acts(P, j)) continue;
        // the stale scalar T: last point (patch order, i outer, j inner) that sets it
        s.iT = (int)(i - 1);
        s.jT = (int)(j - 1);
        seqpt[(size_t)(i - 1) + (size_t)nz * (j - 1)].push_back((uint8_t)(k + 1));
      }
  }
  if (s.iT < 0 && nr >= 2) { s.iT = nz - 1; s.jT = nr - 1; }  // set by the interior loops instead

  std::map<std::vector<uint8_t>, int> ids;
  s.seq.assign((size_t)MAX_SEQ * SEQ_LEN, 0);
  s.nseq = 1;
  s.meta.assign(ng, 0);
  s.any_blank = 0;
  for (int j = 1; j <= nr; ++j)
    for (int i = 1; i <= nz; ++i) {
      const size_t g = (size_t)(i - 1) + (size_t)nz * (j - 1);
      const bool blank = iblank && iblank[g] == 0;
      s.any_blank |= blank;
      uint32_t zr = ROW_INTERIOR, rr = ROW_INTERIOR;
      if (blank) { zr = ROW_BLANK; rr = ROW_BLANK; }
      else {
        if (which_i[g]) {
          const ajet_patch& P = s.patches[which_i[g] - 1];
          zr = P.bc_dir == 1 ? (uint32_t)(i - P.i_start) : (uint32_t)(8 + i - P.i_end);
        }
        if (which_j[g]) {
          const ajet_patch& P = s.patches[which_j[g] - 1];
          rr = P.bc_dir == 2 ? (uint32_t)(j - P.j_start) : (uint32_t)(8 + j - P.j_end);
        }
      }
      uint32_t sid = 0;
      if (!seqpt[g].empty()) {
        if (seqpt[g].size() > (size_t)SEQ_LEN) { err = "too many boundary patches meet at one point"; return AJET_EPATCH; }
        auto it = ids.find(seqpt[g]);
        if (it == ids.end()) {
          if (s.nseq >= MAX_SEQ) { err = "too many distinct boundary-patch combinations"; return AJET_EPATCH; }
          for (size_t q = 0; q < seqpt[g].size(); ++q) s.seq[(size_t)s.nseq * SEQ_LEN + q] = seqpt[g][q];
          it = ids.emplace(seqpt[g], s.nseq++).first;
        }
        sid = (uint32_t)it->second;
      }
      s.meta[g] = zr | (rr << 4) | (sid << 8);
    }
  return AJET_OK;
}

// The stage kernel stages a tile plus a 2-point halo; every closure stencil it evaluates
// must read inside that window.  Cut one direction into segments of at most T points so
// that no closure of an owned point reads outside [start - 2, end + 2): a closure reads six
// points from its strip edge, so a cut may have to move back by up to four points.
// lo[k], hi[k]: the union over the other direction of the closure read range of index k
// (lo = INT_MAX where no closure).
static int cut_segments(int n, int T, const std::vector<int>& lo, const std::vector<int>& hi, std::vector<int>& cuts) {
  cuts.assign(1, 0);
  int s = 0;
  while (s < n) {
    int e = s + T < n ? s + T : n;
    for (;;) {
      bool bad = false;
      const int top = e + 6 < n ? e + 6 : n;
      for (int k = s; k < top && !bad; ++k) {
        if (lo[k] > hi[k]) continue;
        if (k < e && (lo[k] < s - HALO || hi[k] > e + HALO - 1)) bad = true;
        if (k >= e && lo[k] < e - HALO) bad = true;
      }
      if (!bad) break;
      if (--e <= s) return AJET_ETILE;
    }
    cuts.push_back(e);
    s = e;
  }
  return AJET_OK;
}

// A tile is interior when no owned point is on the axis row, on a boundary patch or off the
// interior r stencil, and every window row of its owned columns takes the interior z
// stencil (dVr/dz and dVz/dz are staged on those rows for the cross derivatives).
void classify_tiles(const Setup& s, int nz, int nr, std::vector<int>& interior, std::vector<int>& boundary) {
  interior.clear();
  boundary.clear();
  const int ntz = (int)s.zcuts.size() - 1, ntr = (int)s.rcuts.size() - 1;
  for (int tr = 0; tr < ntr; ++tr)
    for (int tz = 0; tz < ntz; ++tz) {
      const int i0 = s.zcuts[tz], i1 = s.zcuts[tz + 1], j0 = s.rcuts[tr], j1 = s.rcuts[tr + 1];
      // the window (HALO points on every side) lies inside the grid
      bool ok = j0 - HALO >= 0 && j1 + HALO <= nr && i0 - HALO >= 0 && i1 + HALO <= nz;
      for (int j = std::max(0, j0 - HALO); ok && j < std::min(nr, j1 + HALO); ++j)
        for (int i = i0; ok && i < i1; ++i) {
          const uint32_t m = s.meta[(size_t)i + (size_t)nz * j];
          if ((m & 15u) != ROW_INTERIOR) ok = false;
          if (j >= j0 && j < j1 && (((m >> 4) & 15u) != ROW_INTERIOR || (m >> 8) != 0u)) ok = false;
        }
      (ok ? interior : boundary).push_back(tr * ntz + tz);
    }
}

// Tiles of the artificial-diffusivity kernels that need no metadata: every point where the
// tile takes a second derivative -- the owned rows over the owned columns +-2 (along xi), the
// owned columns over the owned rows +-2 (along eta) -- is inside the grid and takes the
// interior stencil in that direction.  That also makes every owned point interior in z and r
// and keeps the axis row out.
void classify_lad_tiles(const Setup& s, int nz, int nr, std::vector<int>& interior, std::vector<int>& boundary) {
  interior.clear();
  boundary.clear();
  const int ntz = (int)s.zcuts.size() - 1, ntr = (int)s.rcuts.size() - 1;
  for (int tr = 0; tr < ntr; ++tr)
    for (int tz = 0; tz < ntz; ++tz) {
      const int i0 = s.zcuts[tz], i1 = s.zcuts[tz + 1], j0 = s.rcuts[tr], j1 = s.rcuts[tr + 1];
      bool ok = i0 - 2 >= 0 && i1 + 2 <= nz && j0 - 2 >= 0 && j1 + 2 <= nr;
      for (int j = j0; ok && j < j1; ++j)
        for (int i = i0 - 2; ok && i < i1 + 2; ++i)
          if ((s.meta[(size_t)i + (size_t)nz * j] & 15u) != ROW_INTERIOR) ok = false;
      for (int j = j0 - 2; ok && j < j1 + 2; ++j)
        for (int i = i0; ok && i < i1; ++i)
          if (((s.meta[(size_t)i + (size_t)nz * j] >> 4) & 15u) != ROW_INTERIOR) ok = false;
      (ok ? interior : boundary).push_back(tr * ntz + tz);
    }
}

// Source tiles of the artificial-diffusivity kernels that take |S| and the dilatation from the
// velocity themselves (fused), and the others, which read them from the strain kernel.  A tile
// is fused when it is one of classify_lad_tiles' interior tiles and every point where its
// second-derivative pairs read |S| and dil -- the owned rows over the owned columns +-4, the
// owned columns over the owned rows +-4 -- takes the interior first derivative along z and
// along r, lies off the axis, and has the velocity it differentiates (2 more points along z
// and r) inside the grid.  strain: the tiles whose |S|, dil and f_sw a non-fused tile reads,
// i.e. every tile within LAD_REACH points of one (its own included), split as
// classify_lad_tiles splits them.
void classify_lad_source(const Setup& s, int nz, int nr, std::vector<int>& fused, std::vector<int>& other,
                         std::vector<int>& str_interior, std::vector<int>& str_boundary) {
  std::vector<int> it, bt;
  classify_lad_tiles(s, nz, nr, it, bt);
  const int ntz = (int)s.zcuts.size() - 1, ntr = (int)s.rcuts.size() - 1;
  std::vector<uint8_t> is_int((size_t)ntz * ntr, 0);
  for (int t : it) is_int[(size_t)t] = 1;
  auto interior_point = [&](int i, int j) {
    const uint32_t m = s.meta[(size_t)i + (size_t)nz * j];
    return (m & 15u) == ROW_INTERIOR && ((m >> 4) & 15u) == ROW_INTERIOR;
  };
  fused.clear();
  other.clear();
  for (int tr = 0; tr < ntr; ++tr)
    for (int tz = 0; tz < ntz; ++tz) {
      const int t = tr * ntz + tz;
      const int i0 = s.zcuts[tz], i1 = s.zcuts[tz + 1], j0 = s.rcuts[tr], j1 = s.rcuts[tr + 1];
      bool ok = is_int[(size_t)t] && i0 - 6 >= 0 && i1 + 6 <= nz && j0 - 6 >= 1 && j1 + 6 <= nr;
      for (int j = j0; ok && j < j1; ++j)
        for (int i = i0 - 4; ok && i < i1 + 4; ++i) ok = interior_point(i, j);
      for (int j = j0 - 4; ok && j < j1 + 4; ++j)
        for (int i = i0; ok && i < i1; ++i) ok = interior_point(i, j);
      (ok ? fused : other).push_back(t);
    }
  // the tiles any non-fused tile's window (LAD_REACH beyond it) touches
  std::vector<uint8_t> need((size_t)ntz * ntr, 0);
  for (int t : other) {
    const int tz = t % ntz, tr = t / ntz;
    const int lo_i = s.zcuts[tz] - LAD_REACH, hi_i = s.zcuts[tz + 1] + LAD_REACH;
    const int lo_j = s.rcuts[tr] - LAD_REACH, hi_j = s.rcuts[tr + 1] + LAD_REACH;
    for (int ur = 0; ur < ntr; ++ur) {
      if (s.rcuts[ur + 1] <= lo_j || s.rcuts[ur] >= hi_j) continue;
      for (int uz = 0; uz < ntz; ++uz)
        if (s.zcuts[uz + 1] > lo_i && s.zcuts[uz] < hi_i) need[(size_t)ur * ntz + uz] = 1;
    }
  }
  str_interior.clear();
  str_boundary.clear();
  for (int t : it)
    if (need[(size_t)t]) str_interior.push_back(t);
  for (int t : bt)
    if (need[(size_t)t]) str_boundary.push_back(t);
}

// One launch list for interior and boundary tiles: interior tile t as t, boundary tile t as ~t.
// A launch of pad8(n) blocks gives XCD x (blocks x, x + 8, ...) the list positions
// [x C, (x + 1) C), C = pad8(n) / 8, in order.  Each XCD's range starts with its share of the
// boundary tiles (the slower ones, first) and goes on with a contiguous run of interior tiles,
// so the XCDs hold about the same work and neighbouring interior tiles share an L2.  Position 0
// holds a boundary tile when there is one.
void mixed_tile_list(const std::vector<int>& interior, const std::vector<int>& boundary, std::vector<int>& out) {
  const int nb = (int)boundary.size(), ni = (int)interior.size(), n = nb + ni;
  const int C = (n + 7) / 8;
  out.clear();
  out.reserve(n);
  int bi = 0, ii = 0;
  for (int x = 0; x < 8; ++x) {
    const int size = std::max(0, std::min(C, n - x * C));
    const int share = (nb + 7 - x) / 8;  // the eight shares sum to nb, the first ones larger
    const int kb = std::min(std::min(share, size), nb - bi);
    for (int k = 0; k < kb; ++k) out.push_back(~boundary[bi++]);
    const int ki = std::min(size - kb, ni - ii);
    for (int k = 0; k < ki; ++k) out.push_back(interior[ii++]);
    for (int k = kb + ki; k < size; ++k) out.push_back(~boundary[bi++]);  // interior ran out
  }
}

// Line groups of the fused z/r filter: each group walks `lines` consecutive r-lines from wb
// through the z recursion and writes the r-filtered output of lines [out_lo, out_hi).  The r
// filter of line j reads lines j-4..j+4 (interior rows) or its closure row's 11 lines, clipped
// to the grid; a group holds every line those windows touch.  Groups are triples (wb, out_lo,
// out_hi); the last ones are pulled down so that all their walked lines are inside the grid.
void build_lad_line_groups(const Setup& s, int nz, int nr, int lines, std::vector<int>& groups) {
  std::vector<int> lo(nr), hi(nr);
  for (int j = 0; j < nr; ++j) {
    lo[j] = j;
    hi[j] = j;
    for (int i = 0; i < nz; ++i) {
      const int rr = (int)((s.meta[(size_t)i + (size_t)nz * j] >> 4) & 15u);
      int b, e;
      if (rr == (int)ROW_BLANK) continue;
      if (rr == (int)ROW_INTERIOR) { b = j - 4; e = j + 4; }
      else { b = rr < 4 ? j - rr : j - rr - 2; e = b + 10; }
      lo[j] = std::min(lo[j], std::max(b, 0));
      hi[j] = std::max(hi[j], std::min(e, nr - 1));
    }
  }
  groups.clear();
  int j = 0;
  while (j < nr) {
    int mlo = lo[j], mhi = hi[j];
    int k = j + 1;
    while (k < nr && std::max(mhi, hi[k]) - std::min(mlo, lo[k]) + 1 <= lines) {
      mlo = std::min(mlo, lo[k]);
      mhi = std::max(mhi, hi[k]);
      ++k;
    }
    int wb = mlo;
    if (wb + lines > nr) wb = std::max(0, std::min(mlo, nr - lines));
    groups.push_back(wb);
    groups.push_back(j);
    groups.push_back(k);
    j = k;
  }
}

int build_tiling(Setup& s, int nz, int nr, std::string& err) {
  const int BIG = 1 << 30;
  std::vector<int> zlo(nz, BIG), zhi(nz, -BIG), rlo(nr, BIG), rhi(nr, -BIG);
  for (int j = 0; j < nr; ++j)
    for (int i = 0; i < nz; ++i) {
      const uint32_t m = s.meta[(size_t)i + (size_t)nz * j];
      const int zr = (int)(m & 15u), rr = (int)((m >> 4) & 15u);
      if (zr != (int)ROW_INTERIOR && zr != (int)ROW_BLANK) {
        const int base = zr < 4 ? i - zr : i - zr + 3;
        zlo[i] = std::min(zlo[i], base);
        zhi[i] = std::max(zhi[i], base + 5);
      }
      if (rr != (int)ROW_INTERIOR && rr != (int)ROW_BLANK) {
        const int base = rr < 4 ? j - rr : j - rr + 3;
        rlo[j] = std::min(rlo[j], base);
        rhi[j] = std::max(rhi[j], base + 5);
      }
    }
  if (cut_segments(nz, TZ, zlo, zhi, s.zcuts) != AJET_OK) {
    err = "the z closures cannot be tiled";
    return AJET_ETILE;
  }
  if (cut_segments(nr, TR, rlo, rhi, s.rcuts) != AJET_OK) {
    err = "the r closures cannot be tiled";
    return AJET_ETILE;
  }
  return AJET_OK;
}

// Grid spacings of ComputeTobeFilteredQuantities (ModShockCapturing.f90:169-187), centred
// inside and one-sided at the two ends, and the ramp (1 + tanh(5 (z - 10))) / 2 of
// ComputeArtTransCoeffs (ModShockCapturing.f90:85).  The grid is a tensor product, so all
// three are 1-D; they are computed here, on the host, with the host's tanh.
void lad_tables(const Setup& s, int nz, int nr, std::vector<double>& dz, std::vector<double>& dr,
                std::vector<double>& ramp) {
  dz.assign((size_t)nz, 0.0);
  dr.assign((size_t)nr, 0.0);
  ramp.assign((size_t)nz, 0.0);
  for (int i = 0; i < nz; ++i) {
    double d = 0.0;
    if (i >= 1 && i <= nz - 2) d = (s.zc[i + 1] - s.zc[i - 1]) / 2.0;
    if (i == 0 && nz > 1) d = s.zc[1] - s.zc[0];
    if (i == nz - 1 && nz > 1) d = s.zc[nz - 1] - s.zc[nz - 2];
    dz[i] = d;
    ramp[i] = (1.0 + std::tanh(5.0 * (s.zc[i] - 10.0))) / 2.0;
  }
  for (int j = 0; j < nr; ++j) {
    double d = 0.0;
    if (j >= 1 && j <= nr - 2) d = (s.rc[j + 1] - s.rc[j - 1]) / 2.0;
    if (j == 0 && nr > 1) d = s.rc[1] - s.rc[0];
    if (j == nr - 1 && nr > 1) d = s.rc[nr - 1] - s.rc[nr - 2];
    dr[j] = d;
  }
}

}  // namespace ajet

// Checks classify_lad_source (which LAD source tiles take |S| and dil from the velocity
// themselves, and which tiles the strain kernel still covers) on a case read from a file:
//   int32 nz, nr, nlines; float64 z[nz nr], r[nz nr] (Fortran order); int32 iblank[nz nr];
//   int64 lines[nlines][7]
// Against the point metadata, by brute force: every tile is a source tile once; a fused tile's
// |S| points (owned rows over owned columns +-4, owned columns over owned rows +-4) are
// interior both ways, off the axis, with the velocity 2 beyond them in the grid, and its second
// derivatives are interior; every point within LAD_REACH of a non-fused tile lies in a strain
// tile; strain tiles split as classify_lad_tiles splits them.  Prints "ok fused F other O
// strain S of T" or the first failure.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../axisymmetricjet_amd/csrc/ajet_internal.h"

using namespace ajet;

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[3];
  if (std::fread(hdr, sizeof hdr, 1, f) != 1) return 2;
  const int nz = hdr[0], nr = hdr[1], nl = hdr[2];
  const size_t ng = (size_t)nz * nr;
  std::vector<double> z(ng), r(ng);
  std::vector<int32_t> ib(ng);
  std::vector<int64_t> raw((size_t)nl * 7);
  if (std::fread(z.data(), 8, ng, f) != ng || std::fread(r.data(), 8, ng, f) != ng ||
      std::fread(ib.data(), 4, ng, f) != ng || std::fread(raw.data(), 8, raw.size(), f) != raw.size())
    return 2;
  std::fclose(f);
  std::vector<ajet_bc> lines((size_t)nl);
  for (int k = 0; k < nl; ++k)
    lines[k] = {raw[7 * k], raw[7 * k + 1], raw[7 * k + 2], raw[7 * k + 3], raw[7 * k + 4], raw[7 * k + 5],
                raw[7 * k + 6]};
  ajet_params p{};
  p.nz = nz;
  p.nr = nr;
  p.gam = 1.4;
  p.Re = 1e4;
  p.Pr = 0.7;
  p.CFL = 0.5;
  p.viscous = 1;
  Setup s;
  std::string err;
  if (build_setup(p, z.data(), r.data(), ib.data(), lines.data(), nl, s, err) != AJET_OK ||
      build_tiling(s, nz, nr, err) != AJET_OK) {
    std::printf("setup failed: %s\n", err.c_str());
    return 1;
  }
  std::vector<int> fused, other, si, sb, it, bt;
  classify_lad_source(s, nz, nr, fused, other, si, sb);
  classify_lad_tiles(s, nz, nr, it, bt);
  const int ntz = (int)s.zcuts.size() - 1, ntr = (int)s.rcuts.size() - 1, nt = ntz * ntr;
  auto fail = [](const char* what, int t) {
    std::printf("fail: %s (tile %d)\n", what, t);
    return 1;
  };
  std::vector<int> seen(nt, 0), cls(nt, -1);
  for (int t : fused) ++seen[t];
  for (int t : other) ++seen[t];
  for (int t = 0; t < nt; ++t)
    if (seen[t] != 1) return fail("source tile listed other than once", t);
  for (int t : it) cls[t] = 1;
  for (int t : bt) cls[t] = 0;
  auto meta = [&](int i, int j) { return s.meta[(size_t)i + (size_t)nz * j]; };
  auto row_int = [&](int i, int j, int shift) { return ((meta(i, j) >> shift) & 15u) == ROW_INTERIOR; };
  for (int t : fused) {
    const int tz = t % ntz, tr = t / ntz;
    const int i0 = s.zcuts[tz], i1 = s.zcuts[tz + 1], j0 = s.rcuts[tr], j1 = s.rcuts[tr + 1];
    if (cls[t] != 1) return fail("fused tile is not an interior tile", t);
    for (int j = j0 - 4; j < j1 + 4; ++j)
      for (int i = i0 - 4; i < i1 + 4; ++i) {
        const bool zarm = j >= j0 && j < j1, rarm = i >= i0 && i < i1;
        if (!zarm && !rarm) continue;
        if (i - 2 < 0 || i + 2 >= nz || j - 2 < 1 || j + 2 >= nr) return fail("velocity window leaves the grid", t);
        if (!row_int(i, j, 0) || !row_int(i, j, 4)) return fail("|S| point is not interior", t);
      }
  }
  std::vector<uint8_t> cover(ng, 0);
  std::vector<int> sseen(nt, 0);
  for (int pass = 0; pass < 2; ++pass)
    for (int t : pass ? sb : si) {
      if (cls[t] != (pass ? 0 : 1)) return fail("strain tile in the wrong class", t);
      ++sseen[t];
      const int tz = t % ntz, tr = t / ntz;
      for (int j = s.rcuts[tr]; j < s.rcuts[tr + 1]; ++j)
        for (int i = s.zcuts[tz]; i < s.zcuts[tz + 1]; ++i) cover[(size_t)i + (size_t)nz * j] = 1;
    }
  for (int t = 0; t < nt; ++t)
    if (sseen[t] > 1) return fail("strain tile listed twice", t);
  for (int t : other) {
    const int tz = t % ntz, tr = t / ntz;
    for (int j = s.rcuts[tr] - LAD_REACH; j < s.rcuts[tr + 1] + LAD_REACH; ++j)
      for (int i = s.zcuts[tz] - LAD_REACH; i < s.zcuts[tz + 1] + LAD_REACH; ++i)
        if (i >= 0 && i < nz && j >= 0 && j < nr && !cover[(size_t)i + (size_t)nz * j])
          return fail("a point a non-fused tile reads is in no strain tile", t);
  }
  std::printf("ok fused %zu other %zu strain %zu of %d\n", fused.size(), other.size(), si.size() + sb.size(), nt);
  return 0;
}

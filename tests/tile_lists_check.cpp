// Checks mixed_tile_list (the launch list of interior and boundary tiles) on synthetic lists:
// every tile appears once, boundary tiles as ~t, a boundary tile first, and each of the eight
// XCD ranges of a pad8(n)-block launch holds its share of the boundary tiles at its start.
// Prints "ok" or the first failing case.
#include <cstdio>
#include <vector>

#include "../axisymmetricjet_amd/csrc/ajet_internal.h"

int main() {
  const int cases[][2] = {{0, 0}, {1, 0}, {0, 5}, {1, 1}, {3, 4}, {7, 1}, {8, 0}, {9, 100}, {356, 5164},
                          {50, 3}, {17, 17}, {100, 900}, {5, 3000}};
  for (const auto& cs : cases) {
    const int nb = cs[0], ni = cs[1], n = nb + ni;
    std::vector<int> it, bt, out;
    for (int k = 0; k < n; ++k) ((k * 7) % 3 == 0 && (int)bt.size() < nb ? bt : (int)it.size() < ni ? it : bt).push_back(k);
    ajet::mixed_tile_list(it, bt, out);
    std::vector<int> seen(n, 0);
    bool ok = (int)out.size() == n;
    for (int e : out) {
      const int t = e >= 0 ? e : ~e;
      if (t < 0 || t >= n) { ok = false; break; }
      ++seen[t];
      bool isb = false;
      for (int b : bt) isb |= b == t;
      if (isb != (e < 0)) ok = false;
    }
    for (int s : seen) ok &= s == 1;
    if (nb > 0 && !out.empty()) ok &= out[0] < 0;
    // interior entries keep their order
    int last = -1;
    for (int e : out)
      if (e >= 0) { ok &= e > last; last = e; }
    // per XCD range: boundary entries first; shares differ by at most one where the range allows
    const int C = (n + 7) / 8;
    int mn = 1 << 30, mx = -1;
    for (int x = 0; x < 8; ++x) {
      int nbx = 0;
      bool interior_seen = false;
      for (int p = x * C; p < (x + 1) * C && p < n; ++p) {
        if (out[p] >= 0) interior_seen = true;
        else { ++nbx; if (interior_seen && ni > 0 && (x + 1) * C <= n) ok = false; }
      }
      if ((x + 1) * C <= n && C > (nb + 7) / 8) { mn = nbx < mn ? nbx : mn; mx = nbx > mx ? nbx : mx; }
    }
    if (mx >= 0) ok &= mx - mn <= 1;
    if (!ok) {
      std::printf("fail nb=%d ni=%d\n", nb, ni);
      return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}

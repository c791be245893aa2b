// Random cases of the persistent batch's overlap rule (convfpga_amd/csrc/
// batch_spans.h): the sweep must find a conflict exactly when a brute-force
// check over every pair of instances does, and a pair it reports must break
// the rule.  Prints "ok" and exits 0 when every case agrees.
#include <cstdio>
#include <random>
#include <vector>

#include "batch_spans.h"

using cfspan::Span;

static bool meet(const Span& p, const Span& q) { return p.a < q.b && q.a < p.b; }

// the rule, pair by pair: an output of one instance meets any span of another
static bool brute(const std::vector<Span>& s) {
  for (const Span& p : s)
    for (const Span& q : s)
      if (p.owner != q.owner && p.out && meet(p, q)) return true;
  return false;
}

static bool pair_breaks(const std::vector<Span>& s, unsigned x, unsigned y) {
  if (x == y) return false;
  for (const Span& p : s)
    for (const Span& q : s)
      if (p.owner == x && q.owner == y && p.out && meet(p, q)) return true;
  return false;
}

int main() {
  std::mt19937_64 rng(20261203);
  int bad = 0, conflicts = 0;
  for (int c = 0; c < 20000; ++c) {
    const unsigned n = 1 + rng() % 40;
    const uintptr_t range = 64 + rng() % 4000, len = 1 + rng() % 64;
    std::vector<Span> s;
    for (unsigned i = 0; i < n; ++i) {
      const uintptr_t o = 1 + rng() % range;
      s.push_back({o, o + len, i, true});
      if (rng() % 3) {  // a separate signal (else in place); never a partial overlap with its own output
        uintptr_t g = 1 + rng() % range;
        while (g != o && g < o + len && o < g + len) g = 1 + rng() % range;
        if (g != o) s.push_back({g, g + len, i, false});
      }
    }
    unsigned x = 0, y = 0;
    const bool found = cfspan::find_conflict(s, &x, &y), expect = brute(s);
    conflicts += expect;
    if (found != expect || (found && !pair_breaks(s, x, y))) {
      if (bad < 5) std::printf("case %d: sweep %d brute %d pair %u %u\n", c, found, expect, x, y);
      ++bad;
    }
  }
  std::printf("%s (%d mismatches; %d of 20000 cases with a conflict)\n", bad ? "FAILED" : "ok", bad, conflicts);
  return bad ? 1 : 0;
}

// batch_spans.h -- the overlap rule of a persistent batch, whose instances run
// concurrently: no instance's output may meet another instance's signal or
// output (signals may be shared; an instance's own signal and output are the
// caller's business).  One sweep over the spans sorted by start, O(n log n),
// so that batches of many thousands of instances are checked at once.
// Plain C++ (the host runtime and a CPU test include it).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace cfspan {

struct Span {
  uintptr_t a, b;  // [a, b), a < b
  unsigned owner;  // instance
  bool out;        // the instance's output (else its signal)
};

// The largest end among the spans seen so far, and the largest among those of
// every other owner: enough to answer "largest end of a span not owned by o".
struct TopEnds {
  uintptr_t end1 = 0, end2 = 0;
  unsigned owner1 = ~0u, owner2 = ~0u;
  void push(uintptr_t e, unsigned o) {
    if (o == owner1) {
      end1 = std::max(end1, e);
    } else if (e > end1) {
      end2 = end1;
      owner2 = owner1;
      end1 = e;
      owner1 = o;
    } else if (e > end2) {
      end2 = e;
      owner2 = o;
    }
  }
  // largest end of a span owned by another instance than o, and its owner
  uintptr_t max_not(unsigned o, unsigned* who) const {
    *who = o == owner1 ? owner2 : owner1;
    return o == owner1 ? end2 : end1;
  }
};

// Finds a pair of instances that break the rule: returns true with *x the
// owner of an output and *y the other instance whose span it meets.
inline bool find_conflict(std::vector<Span> s, unsigned* x, unsigned* y) {
  std::sort(s.begin(), s.end(), [](const Span& p, const Span& q) { return p.a < q.a; });
  TopEnds any, outs;  // over the spans before the current one: all of them, the outputs
  for (const Span& c : s) {
    // a span that started earlier meets c iff it ends after c starts
    unsigned who = 0;
    if (c.out && any.max_not(c.owner, &who) > c.a) {
      *x = c.owner;
      *y = who;
      return true;
    }
    if (!c.out && outs.max_not(c.owner, &who) > c.a) {
      *x = who;
      *y = c.owner;
      return true;
    }
    any.push(c.b, c.owner);
    if (c.out) outs.push(c.b, c.owner);
  }
  return false;
}

}  // namespace cfspan

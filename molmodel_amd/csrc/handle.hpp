// handle.hpp -- the opaque mm_engine handle of the C ABI and the internal
// entry points that host-side modules built on it share.
#pragma once
#include "engine.hpp"

struct mm_engine {
  mmamd::Engine e;
  int pos_dirty = 1;
  int want_virial_last = 0;
};

namespace mmamd {
// one full force evaluation (neighbour list, pair, reciprocal space,
// finish) at the current positions, with overflow retries; synchronises
int handle_compute(mm_engine *m);
}  // namespace mmamd

// Per-shape-class autotuner shared by the operator families.
#pragma once
#include <functional>

#include "tm_common.h"

namespace tmh {

// Returns the fastest of `n` launch variants for the shape class (op, key).
// The first call for a class times every variant with HIP events on `st`
// (each variant must produce identical outputs, so the last timed run leaves
// the caller's outputs complete); later calls reuse the cached choice.  While
// `st` is capturing a graph, or when tuning is disabled, `dflt` is returned
// and run by the caller.  `ran` reports whether the outputs were produced.
int tune_pick(const char* op, const int64_t* key, int nkey, int n, int dflt,
              const std::function<int(int)>& run, hipStream_t st, bool* ran);

// While one of these lives on a host thread, tune_pick on that thread times
// nothing: it returns the cached variant of the class, or `dflt`.  For entry
// points that must not create HIP objects (the timing events) or run extra
// launches.
struct TuneSuspend {
  TuneSuspend();
  ~TuneSuspend();
};

}  // namespace tmh

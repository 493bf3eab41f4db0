// Autotuner of launch choices: for a launch key (pass, precision, lengths, sizes) the kernel
// variant and the resident work-groups per CU of a persistent grid are chosen by timing the
// candidates once, and kept for the context's lifetime.
#pragma once
#include <functional>
#include <string>

#include "vpmx_ctx.h"

namespace vpmx {

// Candidates: variant fi < nf at each resident work-group count per CU from its occupancy
// down (at most four counts).  The first time `key` is seen -- with tuning on and no graph
// being captured -- every candidate is run three times through trial(fi, per_cu), which queues
// one launch, timed with events on the context's stream; the fastest is kept.  A trial either
// writes scratch arrays nothing reads, or, where every variant writes the same bits, the real
// output.  Otherwise variant 0 at its occupancy.
// persistent = false: the launches are not persistent grids, so each variant is timed once
// (three runs) and per_cu is not a choice.
int tune(Ctx *c, const std::string &key, int nf, const void *const *fns, const char *const *names, size_t lds,
         int threads, const std::function<int(int, int)> &trial, Ctx::Tuned *out, bool persistent = true);
// The same with each variant's own dynamic LDS bytes and work-group size (for its occupancy).
int tune_v(Ctx *c, const std::string &key, int nf, const void *const *fns, const char *const *names,
           const size_t *lds, const int *threads, const std::function<int(int, int)> &trial, Ctx::Tuned *out,
           bool persistent = true);

}  // namespace vpmx

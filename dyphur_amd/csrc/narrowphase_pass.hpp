// narrowphase_pass.hpp — the one-pass narrowphase (narrowphase.hip) as the
// entry points enqueue it.
#pragma once
#include "dy_internal.hpp"

namespace dy {

// counts[]: [0] stored (clamped to capacity), [1] true count, [2] an
// unsupported shape was met, [3] a wait for an earlier tile's count gave up
// (internal error), [4] the tile ticket.
constexpr uint32_t kNpCounts = DY_NARROWPHASE_COUNT_WORDS, kNpTicket = 4;
static_assert(kNpCounts == 5, "counts layout");
constexpr uint32_t kNpTilePairs = 256;  // pairs per tile (one per lane of a workgroup)

struct NpOut {
    dy_contact_view cv;  // the store: arrays and capacity (cv.n is not read)
    uint32_t *counts;    // kNpCounts words, cleared by np_enqueue
    uint32_t *feat;      // per-contact feature word, may be null
};

// Enqueues one pass over at most n_bound pairs (the first min(*d_n, n_bound)
// when d_n is not null): clears counts, then tests the pairs and writes the
// contacts in pair order.  status holds ceil(n_bound / 256) tile words, each
// zero or tagged with an epoch other than this run's (epoch != 0).
int np_enqueue(hipStream_t st, const dy_contact_pair *pairs, uint32_t n_bound, const uint32_t *d_n,
               const dy_body_view &bv, const dy_shape_view &sv, uint64_t *status, uint32_t epoch, NpOut out);

}  // namespace dy

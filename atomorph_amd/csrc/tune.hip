// Autotuner of the batched refinement (no counterpart in the reference,
// whose descent has no tunable schedule).
//
// The batched descent has two free parameters: the kernel variant (LDS tile
// capacity and threads per block, amh_refine_tiles) and the number of
// conflict-free rounds each work item runs per pass.  Larger tiles keep
// more of a long chain in one block; more rounds per pass amortise the
// row load/store against more swap attempts but spend them on a row whose
// easy gains are already taken.  Which wins depends on the chain widths and
// on how far the chains are from their optimum, so the tuner measures it on
// the caller's own chains: every candidate runs from the same starting
// assignment (a scratch copy) for a fixed budget of device time, and the
// one with the largest exact cost decrease at the budget wins.
#include "common.h"
#include "../../include/atomorph_hip.h"

#include <algorithm>
#include <vector>

using namespace amh;

namespace {

constexpr int32_t kRoundChoices[] = {4, 8, 16, 32, 64};
constexpr int kMaxPassesPerCandidate = 512;

struct Event {
    hipEvent_t e = nullptr;
    ~Event() {
        if (e) (void)hipEventDestroy(e);
    }
};

struct DevMem {
    void *p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int amh_refine_autotune(const uint64_t *pts_dev, uint64_t *work_dev, int32_t nkeys, int64_t wtot,
                                   const int64_t *chain_off_dev, const int32_t *chain_w_dev, const int32_t *chain_w,
                                   int32_t nchains, double budget_ms, uint64_t seed, int32_t rounds_fixed,
                                   int32_t *tile_out, int32_t *rounds_out, amh_tune_entry *report,
                                   int32_t max_report, int32_t *nreport, void *stream) {
    if (!pts_dev || !work_dev || !chain_off_dev || !chain_w_dev || !chain_w || !tile_out || !rounds_out ||
        nkeys < 2 || wtot <= 0 || nchains <= 0 || pts_dev == work_dev || max_report < 0 || rounds_fixed < 0 ||
        (max_report > 0 && !report)) {
        set_error("amh_refine_autotune: invalid argument");
        return AMH_E_ARG;
    }
    if (!(budget_ms > 0.0)) budget_ms = 3.0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t pts_bytes = sizeof(uint64_t) * size_t(wtot) * size_t(nkeys);

    int32_t ntiles = 0;
    amh_refine_tiles(&ntiles, nullptr, nullptr, 0);
    int32_t max_items = 0;
    for (int32_t t = 0; t < ntiles; ++t) {
        int32_t n = 0;
        if (int r = amh_refine_items(chain_w, nchains, t, nullptr, 0, &n)) return r;
        max_items = std::max(max_items, n);
    }
    if (max_items == 0) {
        set_error("amh_refine_autotune: no chain has two slots");
        return AMH_E_ARG;
    }
    // queue (16 B) + stats (32 B) + items, one allocation
    DevMem mem;
    const size_t items_off = 64, bytes = items_off + sizeof(int32_t) * 2 * size_t(max_items);
    if (int r = check_hip(hipMalloc(&mem.p, bytes), "amh_refine_autotune alloc")) return r;
    uint32_t *queue = static_cast<uint32_t *>(mem.p);
    amh_refine_stats *stats = reinterpret_cast<amh_refine_stats *>(static_cast<char *>(mem.p) + 16);
    int32_t *items_dev = reinterpret_cast<int32_t *>(static_cast<char *>(mem.p) + items_off);
    Event e0, e1;
    if (int r = check_hip(hipEventCreate(&e0.e), "amh_refine_autotune event")) return r;
    if (int r = check_hip(hipEventCreate(&e1.e), "amh_refine_autotune event")) return r;

    std::vector<int32_t> items(2 * size_t(max_items));
    int32_t best_tile = -1, best_rounds = 0, n_rep = 0;
    double best_gain = -1.0;
    for (int32_t t = 0; t < ntiles; ++t) {
        int32_t nitems = 0;
        if (int r = amh_refine_items(chain_w, nchains, t, items.data(), max_items, &nitems)) return r;
        if (int r = check_hip(hipMemcpyAsync(items_dev, items.data(), sizeof(int32_t) * 2 * size_t(nitems),
                                             hipMemcpyHostToDevice, s),
                              "amh_refine_autotune items"))
            return r;
        for (int32_t rounds : kRoundChoices) {
            if (rounds_fixed > 0 && rounds != kRoundChoices[0]) continue;  // one candidate per tile
            if (rounds_fixed > 0) rounds = rounds_fixed;
            // same start for every candidate
            if (int r = check_hip(hipMemcpyAsync(work_dev, pts_dev, pts_bytes, hipMemcpyDeviceToDevice, s),
                                  "amh_refine_autotune copy"))
                return r;
            if (int r = check_hip(hipMemsetAsync(mem.p, 0, items_off, s), "amh_refine_autotune clear")) return r;
            double ms = 0.0, prev_ms = 0.0, gain_at = 0.0;
            int64_t prev_gain = 0;
            int passes = 0;
            while (passes < kMaxPassesPerCandidate) {
                if (int r = check_hip(hipEventRecord(e0.e, s), "amh_refine_autotune record")) return r;
                if (int r = amh_refine_batched(work_dev, nkeys, wtot, chain_off_dev, chain_w_dev, items_dev, nitems,
                                               t, uint64_t(passes), rounds, seed, queue, stats, s))
                    return r;
                if (int r = check_hip(hipEventRecord(e1.e, s), "amh_refine_autotune record")) return r;
                amh_refine_stats st;
                if (int r = check_hip(hipMemcpyAsync(&st, stats, sizeof(st), hipMemcpyDeviceToHost, s),
                                      "amh_refine_autotune stats"))
                    return r;
                if (int r = check_hip(hipStreamSynchronize(s), "amh_refine_autotune sync")) return r;
                float dt = 0.0f;
                if (int r = check_hip(hipEventElapsedTime(&dt, e0.e, e1.e), "amh_refine_autotune elapsed")) return r;
                ++passes;
                ms += double(dt);
                if (ms >= budget_ms) {
                    // gain at exactly the budget, linear within the last pass
                    const double f = (budget_ms - prev_ms) / std::max(ms - prev_ms, 1e-9);
                    gain_at = double(prev_gain) + f * double(st.gain - prev_gain);
                    prev_gain = st.gain;
                    break;
                }
                prev_ms = ms;
                prev_gain = st.gain;
                gain_at = double(st.gain);
            }
            if (n_rep < max_report) {
                amh_tune_entry &e = report[n_rep];
                e.tile = t;
                e.rounds = rounds;
                e.passes = passes;
                e.reserved = 0;
                e.ms = ms;
                e.gain_at_budget = gain_at;
                e.gain = prev_gain;
            }
            ++n_rep;
            if (gain_at > best_gain) {
                best_gain = gain_at;
                best_tile = t;
                best_rounds = rounds;
            }
        }
    }
    *tile_out = best_tile;
    *rounds_out = best_rounds;
    if (nreport) *nreport = n_rep;
    return AMH_OK;
}

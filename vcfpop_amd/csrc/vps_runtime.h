// vps_runtime.h — host-side chain handle and the sweep plan shared by the
// plain, batched and served drivers.
#pragma once
#include <map>
#include "vps_chain.h"

namespace vps {

int launch_set_mi_binary(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_arrange(const DeviceDataset& ds, const ChainDev& c, int nr, hipStream_t st);
uint32_t hash_ulong_host(uint64_t v);

// UpdateQ variants of one sweep
constexpr uint8_t kQAdmix = 0;    // Dirichlet Q (admixture, or no-admixture burn-in)
constexpr uint8_t kQMetro = 1;    // Metropolis Q (every metrofreq sweeps)
constexpr uint8_t kQNoAdmix = 2;  // PolyLog Z, binary Q

// what sweep m launches; equal plans launch identical kernel sequences
struct SweepPlan {
    uint8_t q, record, post, trace;
    uint32_t code() const { return (uint32_t)q | (uint32_t)record << 2 | (uint32_t)post << 3 | (uint32_t)trace << 4; }
};

}  // namespace vps

struct vps_chain {
    vps_dataset* ds = nullptr;
    vps_params par{};
    vps::ChainDev c{};
    int K = 0, id = 0, m = 0, nr = 0, nlnl = 0, binaryq = 0, singlez = 0, finished = 0;
    int nburnin = 0, nreps = 0, nthinning = 1, metrofreq = 0, writelnl = 0, nadmburnin = 0;
    bool lnl_groups = false;   // automatic chunk sizes of whole lnL groups (Record can sum in the Z pass)
    int64_t lnl_cap = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;                  // parameter updates beside the Z pass
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool side_pending = false;                  // the second stream's updates are not joined yet
    bool capturing = false;                     // launch_sweep is being captured into a graph
    std::vector<void*> allocs;
    int64_t device_bytes = 0;
    bool timing = false;
    int timing_every = 1;          // time every n-th Z pass
    int64_t timed_calls = 0;
    std::vector<hipEvent_t> evs;   // pairs bracketing each timed Z pass
    size_t ev_used = 0;
    std::map<uint32_t, hipGraphExec_t> graphs;   // one captured sweep per plan
    bool graph_off = false;                       // a capture failed: launch kernel by kernel
};

namespace vps {

SweepPlan plan_sweep(const vps_chain* ch);
// A chain's chunking of individuals in the Z pass (the draws and counts do
// not depend on it): set_chunks(ch, G) takes G chunks (at most the chunk
// count the chain's buffers were sized for; its own single-chain graphs are
// dropped), batch_chunk_count gives the count a chain takes in a batch
int set_chunks(vps_chain* ch, int G);
int batch_chunk_count(const vps_chain* ch);
int launch_sweep(vps_chain* ch, const SweepPlan& p);   // every launch of the sweep, on ch->stream
int settle_side(vps_chain* ch);                        // ch->stream waits for the second stream's updates
void commit_sweep(vps_chain* ch, const SweepPlan& p);  // host bookkeeping after the launch

}  // namespace vps

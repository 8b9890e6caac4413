// vps_batch.cpp — many chains on one device advanced together, one graph
// launch per sweep.
//
// The reference runs its chains (K range x replicates) as a task pool of
// host threads, each leasing a device slot and driving one chain with
// synchronous per-pass calls (CalcBayesian / BayesianThread / ManageCUDA,
// src/structure.cpp:2397-2501, 711-756; one chain per device at a time,
// structure_nsimult = 1).  Here a batch holds any number of chains that
// share a device.  A sweep of every chain is captured once as a HIP graph:
// each chain's kernels form one branch, forked from and joined to the batch
// stream, so independent chains overlap on the device and the host issues
// one launch per sweep.
//
// Kernels read the sweep index from device memory and the sweep's last
// kernel advances it, so a captured sweep is valid at every m with the same
// plan (UpdateQ variant, Record or not, burn-in or not).  Graphs are cached
// per vector of per-chain plans; a batch of chains with one schedule needs
// at most a handful.
#include <map>
#include <vector>
#include "vps_runtime.h"

using namespace vps;

struct vps_batch {
    std::vector<vps_chain*> chains;
    std::vector<int> own_chunks;   // each chain's chunk count before the batch
    int device = 0;
    hipStream_t root = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> join, ready;
    std::map<std::vector<uint32_t>, hipGraphExec_t> graphs;
    int64_t launches = 0, captures = 0, eager = 0;
};

namespace {

constexpr size_t kMaxGraphs = 32;   // beyond this, sweeps launch stream by stream

void destroy_batch(vps_batch* b)
{
    // the chains continue alone with their own chunking
    for (size_t i = 0; i < b->own_chunks.size() && i < b->chains.size(); ++i) (void)set_chunks(b->chains[i], b->own_chunks[i]);
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    for (hipEvent_t e : b->join) if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->ready) if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->root) (void)hipStreamDestroy(b->root);
    delete b;
}

// capture one sweep of every chain with the given plans
int capture(vps_batch* b, const std::vector<SweepPlan>& plans, hipGraphExec_t* out)
{
    VPS_HIP(hipStreamBeginCapture(b->root, hipStreamCaptureModeRelaxed));
    int rc = VPS_OK;
    hipError_t e = hipEventRecord(b->fork, b->root);
    if (e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "capture fork: %s", hipGetErrorString(e));
    for (size_t i = 0; i < b->chains.size() && !rc; ++i) {
        vps_chain* ch = b->chains[i];
        e = hipStreamWaitEvent(ch->stream, b->fork, 0);
        if (e != hipSuccess) { rc = fail(VPS_ERR_DEVICE, "capture wait: %s", hipGetErrorString(e)); break; }
        const bool timing = ch->timing;
        ch->timing = false;   // no event records inside a graph
        ch->capturing = true;
        rc = launch_sweep(ch, plans[i]);
        ch->capturing = false;
        ch->timing = timing;
        if (rc) break;
        e = hipEventRecord(b->join[i], ch->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->root, b->join[i], 0);
        if (e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "capture join: %s", hipGetErrorString(e));
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->root, &g);
    if (!rc && e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "end capture: %s", hipGetErrorString(e));
    if (!rc) {
        e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e));
    }
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

}  // namespace

extern "C" int vps_batch_create(vps_chain* const* chains, int n, vps_batch** out)
{
    if (!chains || !out || n < 1) return fail(VPS_ERR_INVALID, "null argument or empty batch");
    *out = nullptr;
    vps_batch* b = new vps_batch();
    b->device = chains[0]->ds->device;
    for (int i = 0; i < n; ++i) {
        vps_chain* ch = chains[i];
        if (!ch) { destroy_batch(b); return fail(VPS_ERR_INVALID, "null chain %d", i); }
        if (ch->ds->device != b->device) { destroy_batch(b); return fail(VPS_ERR_INVALID, "chains on different devices"); }
        if (ch->finished) { destroy_batch(b); return fail(VPS_ERR_STATE, "chain %d already finished", i); }
        for (int j = 0; j < i; ++j)
            if (chains[j] == ch) { destroy_batch(b); return fail(VPS_ERR_INVALID, "chain %d listed twice", i); }
        b->chains.push_back(ch);
    }
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->root, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    b->join.assign(n, nullptr);
    b->ready.assign(n, nullptr);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ready[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) { destroy_batch(b); return fail(VPS_ERR_DEVICE, "batch streams: %s", hipGetErrorString(e)); }
    // the chains' chunking for a batch (results do not depend on it)
    for (vps_chain* ch : b->chains) {
        b->own_chunks.push_back(ch->c.G);
        const int rc = set_chunks(ch, batch_chunk_count(ch));
        if (rc) { destroy_batch(b); return rc; }
    }
    *out = b;
    return VPS_OK;
}

extern "C" int vps_batch_destroy(vps_batch* b)
{
    if (!b) return VPS_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->root);
    destroy_batch(b);
    return VPS_OK;
}

extern "C" int vps_batch_step(vps_batch* b, int n_sweeps)
{
    if (!b) return fail(VPS_ERR_INVALID, "null batch");
    if (n_sweeps < 0) return fail(VPS_ERR_INVALID, "n_sweeps %d", n_sweeps);
    for (vps_chain* ch : b->chains)
        if (ch->finished) return fail(VPS_ERR_STATE, "a chain of the batch is finished");
    VPS_HIP(hipSetDevice(b->device));
    const size_t n = b->chains.size();
    // the batch stream starts after everything already queued on the chains
    for (size_t i = 0; i < n; ++i) {
        VPS_HIP(hipEventRecord(b->ready[i], b->chains[i]->stream));
        VPS_HIP(hipStreamWaitEvent(b->root, b->ready[i], 0));
    }
    std::vector<SweepPlan> plans(n);
    std::vector<uint32_t> key(n);
    int rc = VPS_OK;
    for (int s = 0; s < n_sweeps && !rc; ++s) {
        for (size_t i = 0; i < n; ++i) {
            plans[i] = plan_sweep(b->chains[i]);
            key[i] = plans[i].code();
        }
        auto it = b->graphs.find(key);
        hipGraphExec_t exec = nullptr;
        if (it != b->graphs.end()) {
            exec = it->second;
        } else if (b->graphs.size() < kMaxGraphs) {
            rc = capture(b, plans, &exec);
            if (rc) break;
            b->graphs.emplace(key, exec);
            b->captures++;
        }
        if (exec) {
            hipError_t e = hipGraphLaunch(exec, b->root);
            if (e != hipSuccess) { rc = fail(VPS_ERR_DEVICE, "graph launch: %s", hipGetErrorString(e)); break; }
            b->launches++;
        } else {
            // too many distinct plans: the same launches, stream by stream
            VPS_HIP(hipEventRecord(b->fork, b->root));
            for (size_t i = 0; i < n && !rc; ++i) {
                vps_chain* ch = b->chains[i];
                VPS_HIP(hipStreamWaitEvent(ch->stream, b->fork, 0));
                rc = launch_sweep(ch, plans[i]);
                if (!rc) rc = settle_side(ch);
                if (!rc) {
                    VPS_HIP(hipEventRecord(b->join[i], ch->stream));
                    VPS_HIP(hipStreamWaitEvent(b->root, b->join[i], 0));
                }
            }
            b->eager++;
        }
        if (!rc)
            for (size_t i = 0; i < n; ++i) commit_sweep(b->chains[i], plans[i]);
    }
    // later work on any chain's own stream follows the batch's sweeps
    VPS_HIP(hipEventRecord(b->done, b->root));
    for (vps_chain* ch : b->chains) VPS_HIP(hipStreamWaitEvent(ch->stream, b->done, 0));
    return rc;
}

extern "C" int vps_batch_sync(vps_batch* b)
{
    if (!b) return fail(VPS_ERR_INVALID, "null batch");
    VPS_HIP(hipSetDevice(b->device));
    VPS_HIP(hipStreamSynchronize(b->root));
    return VPS_OK;
}

extern "C" int vps_batch_stats(const vps_batch* b, int64_t* graph_launches, int64_t* captures, int64_t* eager_sweeps)
{
    if (!b || !graph_launches || !captures || !eager_sweeps) return fail(VPS_ERR_INVALID, "null argument");
    *graph_launches = b->launches;
    *captures = b->captures;
    *eager_sweeps = b->eager;
    return VPS_OK;
}

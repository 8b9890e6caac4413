// vps_serve.cpp — a clustering-job server on one device.
//
// The reference's analogue is its run task pool with device-slot leasing
// (CalcBayesian / BayesianThread / ManageCUDA, src/structure.cpp:2397-2501,
// 711-756): runs queue for a free device slot and hold it for their whole
// chain, one chain per device.  Here jobs (a genotype image, options, K, a
// run id and a sweep count) arrive at any time and are admitted while the
// device memory they need fits an HBM budget and an active-job limit.
// Admitted jobs build their device dataset and chain on a loader thread,
// then run concurrently, each on its own stream, in slices of sweeps: a
// scheduler thread issues a job's next slice as soon as its previous one
// completes, so the device always holds work from every active job
// (continuous batching).  A finished job runs Arrange, copies its results to
// host memory and frees its device memory, which admits the next jobs.
// A job's kernels read only its own state, and every reduction has a fixed
// order, so a job's results do not depend on what else runs beside it.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <string.h>
#include <string>
#include <thread>
#include <vector>
#include "vps_runtime.h"

using namespace vps;

namespace {

using Clock = std::chrono::steady_clock;

// owned copy of a genotype image
struct HostImage {
    int32_t N = 0, S = 0;
    int64_t L = 0;
    std::vector<int32_t> ind_pop, n_allele;
    std::vector<uint8_t> geno, bitsize, gt_ploidy;
    std::vector<uint64_t> byte_off;
    std::vector<int64_t> gtab_offset, gt_allele_offset;
    std::vector<uint16_t> pool;
    vps_dataset_desc desc() const
    {
        vps_dataset_desc d{};
        d.n_ind = N; d.n_loc = L; d.n_pop = S;
        d.ind_pop = S > 0 ? ind_pop.data() : nullptr;
        d.geno_bits = geno.data(); d.geno_byte_offset = byte_off.data();
        d.bitsize = bitsize.data(); d.n_allele = n_allele.data();
        d.gtab_offset = gtab_offset.data(); d.gt_ploidy = gt_ploidy.data();
        d.gt_allele_offset = gt_allele_offset.data(); d.allele_pool = pool.data();
        return d;
    }
};

int copy_image(const vps_dataset_desc* in, HostImage& h)
{
    if (in->n_ind < 1 || in->n_loc < 1 || !in->geno_bits || !in->geno_byte_offset || !in->bitsize ||
        !in->n_allele || !in->gtab_offset || !in->gt_ploidy || !in->gt_allele_offset || !in->allele_pool)
        return fail(VPS_ERR_INVALID, "incomplete dataset description");
    h.N = in->n_ind; h.L = in->n_loc; h.S = in->n_pop;
    const int64_t L = h.L;
    uint64_t gbytes = 0;
    for (int64_t l = 0; l < L; ++l) {
        const uint64_t e = in->geno_byte_offset[l] + (((uint64_t)in->bitsize[l] * (uint64_t)h.N + 7) >> 3);
        gbytes = std::max(gbytes, e);
    }
    const int64_t G = in->gtab_offset[L];
    if (G < 0) return fail(VPS_ERR_INVALID, "genotype table size");
    int64_t psize = 0;
    for (int64_t g = 0; g < G; ++g) psize = std::max<int64_t>(psize, in->gt_allele_offset[g] + in->gt_ploidy[g]);
    h.geno.assign(in->geno_bits, in->geno_bits + gbytes);
    h.geno.resize(gbytes + 8, 0);   // readers may fetch whole words past the last id
    h.byte_off.assign(in->geno_byte_offset, in->geno_byte_offset + L);
    h.bitsize.assign(in->bitsize, in->bitsize + L);
    h.n_allele.assign(in->n_allele, in->n_allele + L);
    h.gtab_offset.assign(in->gtab_offset, in->gtab_offset + L + 1);
    h.gt_ploidy.assign(in->gt_ploidy, in->gt_ploidy + G);
    h.gt_allele_offset.assign(in->gt_allele_offset, in->gt_allele_offset + G);
    h.pool.assign(in->allele_pool, in->allele_pool + psize);
    if (h.S > 0) {
        if (!in->ind_pop) return fail(VPS_ERR_INVALID, "n_pop > 0 without ind_pop");
        h.ind_pop.assign(in->ind_pop, in->ind_pop + h.N);
    }
    return VPS_OK;
}

// Device bytes of a job, from its sizes: the dataset (transposed ids padded
// to 64-locus tiles, per-locus tables, genotype descriptors, one chunk
// jump-table set) and the chain (P, Ni, FreqM, FreqA, N x K state,
// likelihood partials, Ni and Record partials for kEstChunkParts chunks).
// The reservation is corrected to the measured bytes once the job is loaded.
constexpr int64_t kEstChunkParts = 16;

int64_t estimate_bytes(const HostImage& h, int K, const vps_params& par)
{
    const int64_t N = h.N, L = h.L;
    int64_t KT = 0;
    for (int64_t l = 0; l < L; ++l) KT += h.n_allele[l];
    // tiles: per id width, its loci split by kind, so at most one partial
    // tile per width and kind
    int64_t per_bits[33] = {0};
    for (int64_t l = 0; l < L; ++l) per_bits[std::min<int>(std::max<int>(1, h.bitsize[l]), 32)]++;
    int64_t ntiles = 0, words = 0;
    for (int b = 1; b <= 32; ++b) {
        if (!per_bits[b]) continue;
        const int64_t t = (per_bits[b] + 63) / 64 + 1;
        const int64_t ipw = std::max(1, 32 / b);
        ntiles += t;
        words += t * ((N + ipw - 1) / ipw) * 64;
    }
    int64_t btab = 0;   // biallelic genotype bytes, 2^bits per locus
    for (int64_t l = 0; l < L; ++l) btab += (int64_t)1 << std::min<int>(std::max<int>(1, h.bitsize[l]), 6);
    const int64_t dataset = words * 4 + L * (4 + 4 + 8 + 8 + 4) + (int64_t)h.gt_ploidy.size() * 8 + N * 8 +
                            ntiles * (56 + 64 * 4) + btab + ((N + 15) / 16) * 8192 + (64 << 10);
    const int64_t parts = std::min<int64_t>(kEstChunkParts, (N + 15) / 16);
    const int64_t chain = K * KT * (8 + 4 + 8) + parts * (K * KT * 4 + (ntiles / 4 + ntiles) * 8) + 3 * KT * 8 +
                          N * K * (8 + 4 + 4 + 8 + 8 + 8) +
                          N * 64 + (ntiles / 4 + 2) * N * std::max<int64_t>(K, 2) * 8 + (64 << 10);
    // the stream graph of the Metropolis proposals (at most N typed)
    int64_t pg = 0;
    if (par.admix_val == 1 && par.locpriori_val != 1 && par.rng_topology == VPS_RNG_REFERENCE && par.metrofreq > 0) {
        int64_t W = 0, S = 0;
        pg_sizes((int)N, K, &W, &S);
        pg = W * 8 + (W / 8) * 16 + 2 * S * 4 + N * K * 4;
    }
    // the reference streams' generators prepared one sweep ahead
    const int64_t prep = par.rng_topology == VPS_RNG_REFERENCE ? 2 * kSppStreams * (1024 * 16 + 4) : 0;
    return dataset + chain + pg + prep;
}

struct Job {
    int64_t id = 0;
    int K = 0, run_id = 0;
    int64_t sweeps = 0, inflight = 0;
    std::atomic<int64_t> done{0};   // advanced by the scheduler, read by job_info
    vps_params par{};
    HostImage img;
    int64_t est = 0, bytes = 0;
    int state = VPS_JOB_PENDING, status = VPS_OK;
    std::string err;
    double t_submit = 0, t_start = 0, t_end = 0;
    vps_dataset* ds = nullptr;
    vps_chain* ch = nullptr;
    hipEvent_t ev = nullptr;
    // results
    std::vector<double> Q, ancestry, rout, lnltrace;
    double lnl_last = 0;
};

}  // namespace

struct vps_server {
    vps_server_config cfg{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_done, cv_work, cv_load;
    std::map<int64_t, std::shared_ptr<Job>> jobs;
    std::deque<std::shared_ptr<Job>> pending, to_load;
    std::vector<std::shared_ptr<Job>> active;   // loaded and running
    int64_t next_id = 1, mem_reserved = 0, mem_peak = 0;
    int64_t n_loading = 0;
    int64_t completed = 0, failed = 0;
    bool stop = false;
    std::thread scheduler;
    std::vector<std::thread> loaders;
};

namespace {

double now_s(const vps_server* s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

// pending -> loading while the budget and the active-job limit allow (FIFO:
// a job that does not fit yet blocks the ones behind it, so none starves)
void admit_locked(vps_server* s)
{
    while (!s->pending.empty()) {
        auto j = s->pending.front();
        const int64_t running = (int64_t)s->active.size() + s->n_loading;
        if (s->cfg.max_active > 0 && running >= s->cfg.max_active) break;
        if (s->mem_reserved + j->est > s->cfg.hbm_budget_bytes) break;
        s->pending.pop_front();
        s->mem_reserved += j->est;
        s->mem_peak = std::max(s->mem_peak, s->mem_reserved);
        j->state = VPS_JOB_LOADING;
        j->t_start = now_s(s);
        s->n_loading++;
        s->to_load.push_back(j);
        s->cv_load.notify_one();
    }
}

void release_device(Job& j)
{
    if (j.ch) { vps_chain_destroy(j.ch); j.ch = nullptr; }
    if (j.ds) { vps_dataset_destroy(j.ds); j.ds = nullptr; }
    if (j.ev) { (void)hipEventDestroy(j.ev); j.ev = nullptr; }
}

void loader_main(vps_server* s)
{
    (void)hipSetDevice(s->cfg.device);
    for (;;) {
        std::shared_ptr<Job> j;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_load.wait(lk, [&] { return s->stop || !s->to_load.empty(); });
            if (s->stop) return;
            j = s->to_load.front();
            s->to_load.pop_front();
        }
        const vps_dataset_desc d = j->img.desc();
        int rc = vps_dataset_create(&d, s->cfg.device, &j->ds);
        if (!rc) rc = vps_chain_create(j->ds, &j->par, j->K, j->run_id, &j->ch);
        if (!rc && hipEventCreateWithFlags(&j->ev, hipEventDisableTiming) != hipSuccess)
            rc = fail(VPS_ERR_DEVICE, "job event");
        std::string msg = rc ? vps_last_error() : "";
        if (!rc) {
            vps_dataset_info inf;
            vps_dataset_get_info(j->ds, &inf);
            j->bytes = inf.device_bytes + j->ch->device_bytes;
        }
        if (rc) release_device(*j);
        std::lock_guard<std::mutex> lk(s->mu);
        s->n_loading--;
        if (rc) {
            j->state = VPS_JOB_FAILED;
            j->status = rc;
            j->err = msg;
            j->t_end = now_s(s);
            s->mem_reserved -= j->est;
            s->failed++;
            { std::vector<uint8_t>().swap(j->img.geno); }
            admit_locked(s);
            s->cv_done.notify_all();
        } else {
            s->mem_reserved += j->bytes - j->est;   // measured, not estimated
            j->est = j->bytes;
            s->mem_peak = std::max(s->mem_peak, s->mem_reserved);
            j->state = VPS_JOB_RUNNING;
            s->active.push_back(j);
            s->cv_work.notify_one();
        }
    }
}

// A finished job's results, copied to host memory by the scheduler thread
// without the server lock and moved into the job under it.
struct JobResults {
    std::vector<double> Q, ancestry, rout, lnltrace;
    double lnl_last = 0;
};

int collect_results(vps_chain* ch, JobResults& r)
{
    const int64_t N = ch->c.N, K = ch->K;
    int rc = vps_chain_finish(ch);
    r.Q.resize(N * K);
    r.ancestry.resize(N * K);
    int64_t rlen = 0, nl = 0;
    if (!rc) rc = vps_chain_get_sizes(ch, &rlen, &nl);
    r.rout.resize(std::max<int64_t>(rlen, 1));
    r.lnltrace.resize(std::max<int64_t>(nl, 1));
    if (!rc) rc = vps_chain_get_state(ch, VPS_STATE_Q, r.Q.data(), r.Q.size() * 8);
    if (!rc) rc = vps_chain_get_state(ch, VPS_STATE_MISUM, r.ancestry.data(), r.ancestry.size() * 8);
    if (!rc && rlen) rc = vps_chain_get_state(ch, VPS_STATE_ROUT, r.rout.data(), rlen * 8);
    if (!rc && nl) rc = vps_chain_get_state(ch, VPS_STATE_LNLTRACE, r.lnltrace.data(), nl * 8);
    if (!rc) rc = vps_chain_get_state(ch, VPS_STATE_LNL, &r.lnl_last, 8);
    r.rout.resize(rlen);
    r.lnltrace.resize(nl);
    return rc;
}

void scheduler_main(vps_server* s)
{
    (void)hipSetDevice(s->cfg.device);
    const int64_t slice = std::max(1, s->cfg.slice_sweeps);
    for (;;) {
        std::vector<std::shared_ptr<Job>> act;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait_for(lk, std::chrono::milliseconds(20), [&] { return s->stop || !s->active.empty(); });
            if (s->stop) return;
            act = s->active;
        }
        bool progressed = false;
        for (auto& jp : act) {
            Job& j = *jp;
            int rc = VPS_OK;
            if (j.inflight > 0) {
                const hipError_t q = hipEventQuery(j.ev);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) rc = fail(VPS_ERR_DEVICE, "job %lld: %s", (long long)j.id, hipGetErrorString(q));
                j.done += j.inflight;
                j.inflight = 0;
                progressed = true;
            }
            if (!rc && j.done < j.sweeps) {
                const int64_t n = std::min(slice, j.sweeps - j.done);
                rc = vps_chain_step(j.ch, (int)n);
                if (!rc && hipEventRecord(j.ev, j.ch->stream) != hipSuccess) rc = fail(VPS_ERR_DEVICE, "job event record");
                if (!rc) j.inflight = n;
                progressed = true;
                if (!rc) continue;
            }
            // finished (or failed): results to the host, device memory back
            std::string msg;
            JobResults res;
            if (!rc) rc = collect_results(j.ch, res);
            if (rc) msg = vps_last_error();
            release_device(j);
            std::lock_guard<std::mutex> lk(s->mu);
            if (!rc) {
                j.Q = std::move(res.Q);
                j.ancestry = std::move(res.ancestry);
                j.rout = std::move(res.rout);
                j.lnltrace = std::move(res.lnltrace);
                j.lnl_last = res.lnl_last;
            }
            j.state = rc ? VPS_JOB_FAILED : VPS_JOB_DONE;
            j.status = rc;
            j.err = msg;
            j.t_end = now_s(s);
            s->mem_reserved -= j.est;
            if (rc) s->failed++; else s->completed++;
            { std::vector<uint8_t>().swap(j.img.geno); }
            for (size_t x = 0; x < s->active.size(); ++x)
                if (s->active[x].get() == &j) { s->active.erase(s->active.begin() + x); break; }
            admit_locked(s);
            s->cv_done.notify_all();
            progressed = true;
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
}

}  // namespace

extern "C" void vps_server_config_default(vps_server_config* c)
{
    memset(c, 0, sizeof(*c));
    c->device = 0;
    c->hbm_budget_bytes = (int64_t)64 << 30;
    c->max_active = 64;
    c->slice_sweeps = 16;
    c->loaders = 2;
}

extern "C" int vps_server_create(const vps_server_config* cfg, vps_server** out)
{
    if (!cfg || !out) return fail(VPS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->hbm_budget_bytes <= 0 || cfg->slice_sweeps < 1 || cfg->max_active < 0 || cfg->loaders < 0 || cfg->loaders > 16)
        return fail(VPS_ERR_INVALID, "server configuration");
    int n = 0;
    int rc = vps_device_count(&n);
    if (rc) return rc;
    if (cfg->device < 0 || cfg->device >= n) return fail(VPS_ERR_INVALID, "device %d of %d", cfg->device, n);
    vps_server* s = new vps_server();
    s->cfg = *cfg;
    s->t0 = Clock::now();
    for (int t = 0; t < std::max(1, cfg->loaders); ++t) s->loaders.emplace_back(loader_main, s);
    s->scheduler = std::thread(scheduler_main, s);
    *out = s;
    return VPS_OK;
}

extern "C" int vps_server_submit(vps_server* s, const vps_dataset_desc* desc, const vps_params* par, int K,
                                 int run_id, int64_t n_sweeps, int64_t* job_id)
{
    if (!s || !desc || !par || !job_id) return fail(VPS_ERR_INVALID, "null argument");
    if (K < 1 || K > VPS_MAX_K) return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", K, VPS_MAX_K);
    // a run is nburnin + nreps sweeps, plus the admixture burn-in of the
    // no-admixture model without a location prior (ReadPar, src/structure.cpp:460-462)
    const int64_t run_len = (int64_t)par->nburnin + par->nreps +
                            ((par->admix_val != 1 && par->locpriori_val != 1) ? par->nadmburnin : 0);
    if (par->nburnin < 0 || par->nreps < 1 || par->nadmburnin < 0)
        return fail(VPS_ERR_INVALID, "nburnin %d, nreps %d, nadmburnin %d", par->nburnin, par->nreps, par->nadmburnin);
    if (n_sweeps == 0) n_sweeps = run_len;
    if (n_sweeps != run_len)
        return fail(VPS_ERR_INVALID, "n_sweeps %lld differs from the run length %lld the parameters give",
                    (long long)n_sweeps, (long long)run_len);
    auto j = std::make_shared<Job>();
    int rc = copy_image(desc, j->img);
    if (rc) return rc;
    j->K = K;
    j->run_id = run_id;
    j->par = *par;
    j->par.sweep_graphs = 0;   // loader threads make HIP calls while jobs run
    if (j->par.z_chunk == 0) j->par.z_chunk = -1;   // no timing trials beside running jobs
    j->sweeps = n_sweeps;
    j->est = estimate_bytes(j->img, K, j->par);
    if (j->est > s->cfg.hbm_budget_bytes)
        return fail(VPS_ERR_BUDGET, "job needs ~%lld bytes, budget %lld", (long long)j->est,
                    (long long)s->cfg.hbm_budget_bytes);
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = s->next_id++;
    j->t_submit = now_s(s);
    s->jobs[j->id] = j;
    s->pending.push_back(j);
    admit_locked(s);
    *job_id = j->id;
    return VPS_OK;
}

extern "C" int vps_server_wait(vps_server* s, int64_t job_id, double timeout_s)
{
    if (!s) return fail(VPS_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(VPS_ERR_NOTFOUND, "job %lld", (long long)job_id);
    auto j = it->second;
    auto finished = [&] { return j->state == VPS_JOB_DONE || j->state == VPS_JOB_FAILED; };
    if (timeout_s < 0) s->cv_done.wait(lk, finished);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished))
        return fail(VPS_ERR_STATE, "job %lld not finished", (long long)job_id);
    if (j->state == VPS_JOB_FAILED) return fail(j->status, "job %lld failed: %s", (long long)job_id, j->err.c_str());
    return VPS_OK;
}

extern "C" int vps_server_job_info(vps_server* s, int64_t job_id, vps_job_info* info)
{
    if (!s || !info) return fail(VPS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(VPS_ERR_NOTFOUND, "job %lld", (long long)job_id);
    const Job& j = *it->second;
    memset(info, 0, sizeof(*info));
    info->job_id = j.id;
    info->state = j.state;
    info->status = j.status;
    info->n_ind = j.img.N;
    info->K = j.K;
    info->n_loc = j.img.L;
    info->sweeps_done = j.done.load();
    info->sweeps_total = j.sweeps;
    info->device_bytes = j.bytes;
    info->estimated_bytes = j.est;
    info->t_submit = j.t_submit;
    info->t_start = j.t_start;
    info->t_end = j.t_end;
    info->rout_len = (int64_t)j.rout.size();
    info->lnltrace_len = (int64_t)j.lnltrace.size();
    return VPS_OK;
}

extern "C" int vps_server_job_result(vps_server* s, int64_t job_id, int which, void* host, size_t bytes)
{
    if (!s || !host) return fail(VPS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(VPS_ERR_NOTFOUND, "job %lld", (long long)job_id);
    const Job& j = *it->second;
    if (j.state != VPS_JOB_DONE) return fail(VPS_ERR_STATE, "job %lld has no results", (long long)job_id);
    const std::vector<double>* v = nullptr;
    switch (which) {
    case VPS_STATE_Q: v = &j.Q; break;
    case VPS_STATE_MISUM: v = &j.ancestry; break;
    case VPS_STATE_ROUT: v = &j.rout; break;
    case VPS_STATE_LNLTRACE: v = &j.lnltrace; break;
    case VPS_STATE_LNL:
        if (bytes != 8) return fail(VPS_ERR_INVALID, "lnL is 8 bytes");
        memcpy(host, &j.lnl_last, 8);
        return VPS_OK;
    default: return fail(VPS_ERR_INVALID, "result %d not kept by the server", which);
    }
    if (bytes != v->size() * 8) return fail(VPS_ERR_INVALID, "result %d: %zu bytes, expected %zu", which, bytes, v->size() * 8);
    memcpy(host, v->data(), bytes);
    return VPS_OK;
}

extern "C" int vps_server_release(vps_server* s, int64_t job_id)
{
    if (!s) return fail(VPS_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(VPS_ERR_NOTFOUND, "job %lld", (long long)job_id);
    if (it->second->state != VPS_JOB_DONE && it->second->state != VPS_JOB_FAILED)
        return fail(VPS_ERR_STATE, "job %lld still queued or running", (long long)job_id);
    s->jobs.erase(it);
    return VPS_OK;
}

extern "C" int vps_server_stats(vps_server* s, vps_server_stats_t* st)
{
    if (!s || !st) return fail(VPS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof(*st));
    st->pending = (int64_t)s->pending.size();
    st->loading = s->n_loading;
    st->running = (int64_t)s->active.size();
    st->completed = s->completed;
    st->failed = s->failed;
    st->reserved_bytes = s->mem_reserved;
    st->peak_reserved_bytes = s->mem_peak;
    st->uptime_s = now_s(s);
    return VPS_OK;
}

extern "C" int vps_server_destroy(vps_server* s)
{
    if (!s) return VPS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_load.notify_all();
    s->cv_done.notify_all();
    if (s->scheduler.joinable()) s->scheduler.join();
    for (auto& t : s->loaders)
        if (t.joinable()) t.join();
    (void)hipSetDevice(s->cfg.device);
    for (auto& jp : s->active) {
        if (jp->ch) (void)hipStreamSynchronize(jp->ch->stream);
        release_device(*jp);
    }
    for (auto& jp : s->to_load) release_device(*jp);
    delete s;
    return VPS_OK;
}

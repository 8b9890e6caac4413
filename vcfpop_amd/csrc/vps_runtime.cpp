// vps_runtime.cpp — chain lifecycle and the sweep driver (BAYESIAN<double>::
// MCMC, src/structure.cpp:2112-2166) over the device kernels, plus the C ABI.
#include <algorithm>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include "vps_runtime.h"

using namespace vps;

namespace vps {

static thread_local char g_err[1024] = "";

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// CRC32C (reflected 0x82F63B78) of one 64-bit word, the reference's
// _mm_crc32_u64 (src/hash.cpp:9-16); HashULong = crc(0xB385FDA9, v)
uint32_t hash_ulong_host(uint64_t v)
{
    static uint32_t tab[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t n = 0; n < 256; ++n) {
            uint32_t c = n;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : c >> 1;
            tab[n] = c;
        }
        ready = true;
    }
    uint32_t c = 0xB385FDA9u;
    for (int b = 0; b < 8; ++b) c = (c >> 8) ^ tab[(c ^ (uint32_t)(v >> (8 * b))) & 0xFF];
    return c;
}

}  // namespace vps

extern "C" void vps_params_default(vps_params* p)
{
    memset(p, 0, sizeof(*p));
    p->admix_val = 2; p->locpriori_val = 2; p->f_val = 2;
    p->nburnin = 1000; p->nreps = 10000; p->nthinning = 10; p->nruns = 1; p->nadmburnin = 500;
    p->lambda = 1; p->stdlambda = 0.3; p->maxlambda = 10;
    p->inferlambda_val = 2; p->difflambda_val = 2; p->diversity_val = 2;
    p->alpha = 1; p->stdalpha = 0.025; p->maxalpha = 10; p->alphapriora = 0.05; p->alphapriorb = 0.001;
    p->inferalpha_val = 1; p->diffalpha_val = 0; p->uniformalpha_val = 1; p->metrofreq = 10;
    p->r = 1; p->maxr = 20; p->stdr = 0.1; p->epseta = 0.025; p->epsgamma = 0.025;
    p->pmeanf = 0.01; p->pstdf = 0.05; p->stdf = 0.05; p->singlef_val = 2;
    p->writelnl_val = 2; p->krange_min = 1; p->krange_max = 5;
    p->g_seed = 0;
    p->rng_topology = VPS_RNG_REFERENCE;
    p->precision = VPS_PREC_FP64;
    p->z_chunk = 0;
    p->sweep_graphs = 0;
}

extern "C" int vps_abi_version(void) { return VPS_ABI_VERSION; }
extern "C" const char* vps_last_error(void) { return vps::g_err; }

extern "C" int vps_device_count(int* n)
{
    if (!n) return fail(VPS_ERR_INVALID, "null argument");
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) { *n = 0; return fail(VPS_ERR_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e)); }
    return VPS_OK;
}

namespace {

template <typename T>
int chain_alloc(vps_chain* ch, T** p, int64_t count)
{
    return dev_alloc((void**)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T), ch->allocs, &ch->device_bytes, ch->stream);
}

int fill_f64(vps_chain* ch, double* p, int64_t n, double v)
{
    std::vector<double> h(std::max<int64_t>(n, 1), v);
    VPS_HIP(hipMemcpyAsync(p, h.data(), n * sizeof(double), hipMemcpyHostToDevice, ch->stream));
    VPS_HIP(hipStreamSynchronize(ch->stream));
    return VPS_OK;
}

void chain_free(vps_chain* ch)
{
    if (!ch) return;
    if (ch->ds) (void)hipSetDevice(ch->ds->device);
    if (ch->aux) (void)hipStreamSynchronize(ch->aux);
    if (ch->stream) {
        (void)hipStreamSynchronize(ch->stream);
        dev_free_all(ch->allocs, ch->stream);
    }
    for (hipEvent_t e : ch->evs) (void)hipEventDestroy(e);
    for (auto& kv : ch->graphs) (void)hipGraphExecDestroy(kv.second);
    if (ch->aux) (void)hipStreamDestroy(ch->aux);
    if (ch->ev_fork) (void)hipEventDestroy(ch->ev_fork);
    if (ch->ev_join) (void)hipEventDestroy(ch->ev_join);
    if (ch->stream) (void)hipStreamDestroy(ch->stream);
    delete ch;
}

// Chunk-size autotuner: time the whole Z pass (count kernels and chunk
// reduction; the fastest of 3 passes after a warm-up) for the model's
// chunking and a ladder of chunk sizes on the
// freshly initialised chain, and keep the fastest.  Chunking changes no draw
// and no count, so the choice only moves time.  The trial passes write a
// scratch Mi (the Metropolis buffer, unused until the first sweep), scratch
// partials and a scratch Ni, leaving the initial state untouched.  At most
// kTuneMaxChunks chunks are tried.
constexpr int kTuneMaxChunks = 32;

int tune_chunks(vps_chain* ch)
{
    ChainDev& c = ch->c;
    const DeviceDataset& d = ch->ds->d;
    const int N = c.N;
    std::vector<int> cand;
    // candidates by chunk count (chunk sizes of whole lnL groups when
    // Record can sum in the Z pass)
    const bool groups = ch->lnl_groups;
    auto add = [&](int nc) {
        nc = std::max(1, std::min(nc, std::min(N, (int)kMaxChunk)));
        int G = (N + nc - 1) / nc;
        G = (N + chunk_size(N, G, groups) - 1) / chunk_size(N, G, groups);
        if (G > kTuneMaxChunks && G != c.G) return;
        if (std::find(cand.begin(), cand.end(), G) == cand.end()) cand.push_back(G);
    };
    add(c.Nc);
    for (int nc : {32, 64, 96, 128, 160, 192, 256, 320, 384, 512}) add(nc);
    if (cand.size() < 2) return VPS_OK;
    const int Gmax = *std::max_element(cand.begin(), cand.end());
    const size_t part_bytes = (size_t)(Gmax + 1) * c.K * c.KT * sizeof(int32_t);   // partials + a scratch Ni
    int32_t* tmp = nullptr;
    VPS_HIP(hipMallocAsync((void**)&tmp, std::max<size_t>(part_bytes, 16), ch->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = VPS_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(VPS_ERR_DEVICE, "tuner events");
    float best_ms = 0;
    int bestG = c.G;
    for (size_t x = 0; x < cand.size() && !rc; ++x) {
        ChainDev t = c;
        t.Nc = chunk_size(N, cand[x], groups);
        t.G = (N + t.Nc - 1) / t.Nc;
        rc = dataset_jump(ch->ds, t.Nc, c.precision == VPS_PREC_FAST32, &t.jump);
        if (rc) break;
        t.Ni_part = tmp;
        t.Ni = tmp + (size_t)Gmax * c.K * c.KT;
        t.Mi = reinterpret_cast<int32_t*>(c.bufNK1);   // N*K doubles >= N*K ints
        rc = launch_z_admix_counts(d, t, ch->stream);  // warm-up
        // the fastest of 3 timed passes: robust to a pass slowed by other work
        float ms = 0;
        hipError_t e = hipSuccess;
        for (int r = 0; r < 3 && !rc && e == hipSuccess; ++r) {
            e = hipEventRecord(e0, ch->stream);
            if (e == hipSuccess) rc = launch_z_admix_counts(d, t, ch->stream);
            if (!rc && e == hipSuccess) e = hipEventRecord(e1, ch->stream);
            if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
            float one = 0;
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&one, e0, e1);
            if (!rc && e == hipSuccess && (r == 0 || one < ms)) ms = one;
        }
        if (!rc && e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "tuner timing: %s", hipGetErrorString(e));
        if (!rc && (x == 0 || ms < best_ms * 0.99f)) { best_ms = ms; bestG = t.G; }
    }
    (void)hipFreeAsync(tmp, ch->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!rc) rc = fill_f64(ch, c.bufNK1, (int64_t)N * c.K, 0.0);   // the scratch was the Metropolis buffer
    if (rc) return rc;
    VPS_HIP(hipStreamSynchronize(ch->stream));
    if (bestG == c.G) return VPS_OK;
    // swap the per-chunk buffers for the chosen chunking
    auto drop = [&](void* p, int64_t bytes) -> int {
        auto it = std::find(ch->allocs.begin(), ch->allocs.end(), p);
        if (it == ch->allocs.end()) return VPS_OK;
        ch->device_bytes -= bytes;
        ch->allocs.erase(it);
        VPS_HIP(hipFreeAsync(p, ch->stream));
        return VPS_OK;
    };
    rc = drop(c.Ni_part, (int64_t)c.G * c.K * c.KT * sizeof(int32_t));
    if (rc) return rc;
    c.Nc = chunk_size(N, bestG, groups);
    c.G = (N + c.Nc - 1) / c.Nc;
    rc = dataset_jump(ch->ds, c.Nc, c.precision == VPS_PREC_FAST32, &c.jump);
    if (!rc) rc = chain_alloc(ch, &c.Ni_part, (int64_t)c.G * c.K * c.KT);
    c.Gcap = c.G;
    return rc;
}

}  // namespace

extern "C" int vps_chain_create(vps_dataset* ds, const vps_params* p, int K, int run_id, vps_chain** out)
{
    if (!ds || !p || !out) return fail(VPS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (K < 1 || K > VPS_MAX_K) return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", K, VPS_MAX_K);
    if (p->nthinning < 1) return fail(VPS_ERR_INVALID, "nthinning must be >= 1");
    if (p->rng_topology != VPS_RNG_REFERENCE && p->rng_topology != VPS_RNG_PARALLEL)
        return fail(VPS_ERR_INVALID, "rng_topology %d", p->rng_topology);
    if (p->precision < VPS_PREC_FP64 || p->precision > VPS_PREC_FAST32)
        return fail(VPS_ERR_INVALID, "precision %d", p->precision);
    if (p->precision >= VPS_PREC_REAL32 && p->f_val == 1)
        return fail(VPS_ERR_UNSUPPORTED, "the F model runs with REAL = double only (precision FP64 or MIXED)");
    const DeviceDataset& d = ds->d;
    VPS_HIP(hipSetDevice(ds->device));

    vps_chain* ch = new vps_chain();
    ch->ds = ds; ch->par = *p; ch->K = K; ch->id = run_id;
    // non-blocking: other threads' null-stream copies (dataset uploads of a
    // server) do not serialise with this chain's kernels
    hipError_t e = hipStreamCreateWithFlags(&ch->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ch->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ch->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ch->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) { chain_free(ch); return fail(VPS_ERR_DEVICE, "chain streams: %s", hipGetErrorString(e)); }

    // ReadPar, src/structure.cpp:405-464
    ChainDev& c = ch->c;
    c.K = K; c.N = d.N; c.S = d.S; c.L = d.L; c.KT = d.KT; c.maxploidy = d.maxploidy;
    c.admix = p->admix_val == 1; c.locpriori = p->locpriori_val == 1; c.fmodel = p->f_val == 1;
    c.topology = p->rng_topology; c.precision = p->precision;
    c.inferalpha = p->inferalpha_val == 1; c.diffalpha = p->diffalpha_val == 1; c.uniformalpha = p->uniformalpha_val == 1;
    c.inferlambda = p->inferlambda_val == 0; c.difflambda = p->difflambda_val == 1; c.fsame = p->singlef_val == 1;
    c.stdalpha = p->stdalpha; c.maxalpha = p->maxalpha; c.alphapriora = p->alphapriora; c.alphapriorb = p->alphapriorb;
    c.stdlambda = p->stdlambda; c.maxlambda = p->maxlambda;
    c.maxr = p->maxr; c.epsr = p->stdr; c.epseta = p->epseta; c.epsgamma = p->epsgamma;
    c.pmeanf = p->pmeanf; c.pstdf = p->pstdf; c.stdf = p->stdf;
    ch->nadmburnin = (!c.locpriori && !c.admix) ? p->nadmburnin : 0;
    c.nadmburnin = ch->nadmburnin;
    ch->nburnin = p->nburnin + ch->nadmburnin;
    ch->nreps = p->nreps; ch->nthinning = p->nthinning; ch->metrofreq = p->metrofreq;
    ch->writelnl = p->writelnl_val == 1;
    // Record sums its lnL inside the Z pass on FP64 admixture sweeps without a
    // Metropolis UpdateQ.  Every Record sweep is a Metropolis sweep when the
    // thinning and the burn-in are multiples of metrofreq; otherwise the
    // automatic chunkings keep the lnL groups whole, so the reported lnL does
    // not depend on the chunk count chosen.
    ch->lnl_groups = c.admix && p->precision == VPS_PREC_FP64 &&
                     !(ch->metrofreq > 0 && ch->nthinning % ch->metrofreq == 0 && ch->nburnin % ch->metrofreq == 0);
    c.seed = hash_ulong_host(p->g_seed + (uint64_t)run_id);
    const int nf = c.fmodel ? (c.fsame ? 1 : K) : 0;
    const int na = c.admix ? (c.diffalpha ? K : 1) : 0;
    const int nl = c.difflambda ? K : 1;
    c.rlen = c.locpriori ? 4 + nl + 1 + K + K * d.S + nf : 4 + nl + na + nf;
    ch->lnl_cap = (ch->nburnin + ch->nthinning - 1) / ch->nthinning + (ch->nreps + ch->nthinning - 1) / ch->nthinning + 1;
    c.nburnin = ch->nburnin; c.nthinning = ch->nthinning; c.writelnl = ch->writelnl; c.lnl_cap = (int32_t)ch->lnl_cap;

    const int64_t N = d.N, KT = d.KT, S = d.S;
    int rc = 0;
    rc = choose_lik_tpr(d, K, ds->device, c.lik_tpr);
    const int nparts = rc ? 0 : std::max({lik_part_count(d, 0, K, c.lik_tpr), lik_part_count(d, 2, K, c.lik_tpr),
                                          lik_part_count(d, 3, K, c.lik_tpr)});
    if (!rc) rc = chain_alloc(ch, &c.Freq, K * KT + 2);   // +2: paired allele loads past the last locus
    if (!rc) rc = chain_alloc(ch, &c.Q, N * K);
    if (!rc) rc = chain_alloc(ch, &c.Qf, N * K);
    if (!rc) rc = chain_alloc(ch, &c.Z, N + 32);
    if (!rc) rc = chain_alloc(ch, &c.Ni, K * KT);
    // individual chunking of the genotype passes (model choice here; z_chunk
    // = 0 then measures candidates once the chain is initialised)
    if (p->z_chunk < -1 || p->z_chunk > kMaxChunk)
        rc = fail(VPS_ERR_INVALID, "z_chunk %d outside -1..%d", p->z_chunk, (int)kMaxChunk);
    if (!rc) {
        int G = 0;
        // an explicit z_chunk fixes the chunk count, with balanced chunks;
        // otherwise chunks are whole lnL groups when Record can sum in the
        // Z pass (chunk_size)
        if (p->z_chunk > 0) G = (int)((N + p->z_chunk - 1) / p->z_chunk);
        else rc = choose_chunks(d, K, p->precision, ds->device, ch->lnl_groups, &G);
        if (!rc) {
            c.Nc = p->z_chunk > 0 ? (int32_t)((N + G - 1) / G) : (int32_t)chunk_size(N, G, ch->lnl_groups);
            c.G = (int32_t)((N + c.Nc - 1) / c.Nc);
            rc = dataset_jump(ds, c.Nc, p->precision == VPS_PREC_FAST32, &c.jump);
        }
    }
    if (!rc) rc = chain_alloc(ch, &c.Ni_part, (int64_t)c.G * K * KT);
    c.Gcap = c.G;
    if (!rc) rc = chain_alloc(ch, &c.zlik, z_lik_parts(d) + kLnlStage1);
    if (!rc) rc = chain_alloc(ch, &c.Mi, N * K);
    if (!rc) rc = chain_alloc(ch, &c.Alpha, K);
    if (!rc) rc = chain_alloc(ch, &c.Lambda, K);
    if (!rc) rc = chain_alloc(ch, &c.bufNK1, N * K);
    if (!rc) rc = chain_alloc(ch, &c.bufN1, N);
    if (!rc) rc = chain_alloc(ch, &c.bufN2, N);
    if (!rc) rc = chain_alloc(ch, &c.part, nparts);
    if (!rc) rc = chain_alloc(ch, &c.MiSum, N * K);
    if (!rc) rc = chain_alloc(ch, &c.FreqM, K * KT);
    if (!rc) rc = chain_alloc(ch, &c.FreqA, 3 * KT);
    if (!rc) rc = chain_alloc(ch, &c.Fst, 2 * K);
    if (!rc) rc = chain_alloc(ch, &c.frng, 1);
    if (!rc) rc = chain_alloc(ch, &c.fctl, 4);
    if (!rc) rc = chain_alloc(ch, &c.fpart, 256);
    if (!rc) rc = chain_alloc(ch, &c.fcnt, 1);
    if (!rc) rc = chain_alloc(ch, &c.lpart, 2 * std::max<int64_t>(kLpCap, (int64_t)S * K));
    if (!rc) rc = chain_alloc(ch, &c.lcnt, 1);
    if (!rc) {
        const uint32_t zero = 0;
        VPS_HIP(hipMemcpyAsync(c.fcnt, &zero, sizeof(zero), hipMemcpyHostToDevice, ch->stream));
        VPS_HIP(hipMemcpyAsync(c.lcnt, &zero, sizeof(zero), hipMemcpyHostToDevice, ch->stream));
        VPS_HIP(hipStreamSynchronize(ch->stream));
    }
    // the stream-graph parse of the Metropolis proposals (reference draw
    // streams, admixture with a global alpha)
    if (!rc && c.admix && !c.locpriori && c.topology == VPS_RNG_REFERENCE && p->metrofreq > 0) {
        int64_t W = 0, Sg = 0;
        pg_sizes(d.ntyped, K, &W, &Sg);
        if (W > 0) {
            c.pg_W = W;
            c.pg_S = Sg;
            rc = chain_alloc(ch, &c.pg_words, W);
            if (!rc) rc = chain_alloc(ch, &c.pg_st8, W / 8);
            if (!rc) rc = chain_alloc(ch, &c.pg_next, 2 * Sg);
            if (!rc) rc = chain_alloc(ch, &c.pg_path, (int64_t)d.ntyped * K);
        }
    }
    if (!rc && c.topology == VPS_RNG_REFERENCE) {
        rc = chain_alloc(ch, &c.spp_prep, 2 * (int64_t)kSppStreams * 1024);
        if (!rc) rc = chain_alloc(ch, &c.spp_tag, 2 * kSppStreams);
        if (!rc && hipMemsetAsync(c.spp_tag, 0xFF, 2 * kSppStreams * sizeof(int32_t), ch->stream) != hipSuccess)
            rc = fail(VPS_ERR_DEVICE, "clear stream preparation tags");
    }
    if (!rc) rc = chain_alloc(ch, &c.Eta, K);
    if (!rc) rc = chain_alloc(ch, &c.Gamma, S * K);
    if (!rc) rc = chain_alloc(ch, &c.Di, S * K);
    if (!rc) rc = chain_alloc(ch, &c.AlphaLocal, S * K);
    if (!rc) rc = chain_alloc(ch, &c.SumAlpha, S);
    if (!rc) rc = chain_alloc(ch, &c.scal, 8);
    if (!rc) rc = chain_alloc(ch, &c.rout, c.rlen);
    if (!rc) rc = chain_alloc(ch, &c.lnltrace, ch->lnl_cap);
    if (!rc) rc = chain_alloc(ch, &c.rngbuf, N);
    if (!rc) rc = chain_alloc(ch, &c.mdev, 1);
    if (!rc) rc = chain_alloc(ch, &c.mside, 1);
    if (!rc && hipMemsetAsync(c.mdev, 0, sizeof(int32_t), ch->stream) != hipSuccess) rc = fail(VPS_ERR_DEVICE, "clear sweep index");
    // the padding past P: read by paired allele loads, never charged
    if (!rc && hipMemsetAsync(c.Freq + (int64_t)K * KT, 0, 2 * sizeof(double), ch->stream) != hipSuccess)
        rc = fail(VPS_ERR_DEVICE, "clear P padding");
    if (rc) { chain_free(ch); return rc; }

    // InitFix (src/structure.cpp:468-553): P = 0.5 (1/k2 off biallelic), Q = 1/K,
    // Alpha = alpha, Lambda = lambda, r = 1, accumulators 0
    {
        std::vector<double> fr((size_t)K * KT, 0.5);
        for (int64_t l = 0; l < d.L; ++l)
            if (ds->k2[l] != 2)
                for (int k = 0; k < K; ++k)
                    for (int a = 0; a < ds->k2[l]; ++a)
                        fr[(size_t)k * KT + ds->afo[l] + a] = real_store(1.0 / ds->k2[l], p->precision >= VPS_PREC_REAL32);
        // InitFmodel (src/structure.cpp:679-708) sets P = Lambda; UpdateP
        // overwrites P before any use, so the value is immaterial
        e = hipMemcpy(c.Freq, fr.data(), fr.size() * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) { chain_free(ch); return fail(VPS_ERR_DEVICE, "upload P: %s", hipGetErrorString(e)); }
    }
    rc = fill_f64(ch, c.Q, N * K, real_store(1.0 / K, p->precision >= VPS_PREC_REAL32));
    if (!rc) rc = fill_f64(ch, c.Alpha, K, p->alpha);
    if (!rc) rc = fill_f64(ch, c.Lambda, K, p->lambda);
    if (!rc) rc = fill_f64(ch, c.FreqM, K * KT, 0.0);
    if (!rc) rc = fill_f64(ch, c.rout, c.rlen, 0.0);
    if (!rc) rc = fill_f64(ch, c.bufNK1, N * K, 0.0);
    if (!rc) rc = fill_f64(ch, c.bufN1, N, 0.0);
    if (!rc) rc = fill_f64(ch, c.bufN2, N, 0.0);
    if (!rc) rc = fill_f64(ch, c.scal, 8, 0.0);
    if (!rc) rc = fill_f64(ch, c.scal, 1, 1.0);   // r = 1 (InitFix)
    if (!rc) rc = fill_f64(ch, c.FreqA, 3 * KT, 0.0);
    if (rc) { chain_free(ch); return rc; }
    if (hipMemsetAsync(c.MiSum, 0, N * K * 8, ch->stream) != hipSuccess ||
        hipMemsetAsync(c.Z, 0, (N + 32) * 2, ch->stream) != hipSuccess ||
        hipMemsetAsync(c.Ni, 0, K * KT * 4, ch->stream) != hipSuccess) {
        chain_free(ch);
        return fail(VPS_ERR_DEVICE, "clear chain state");
    }
    // InitLocPriori (src/structure.cpp:653-675)
    rc = fill_f64(ch, c.AlphaLocal, S * K, p->alpha);
    if (!rc) rc = fill_f64(ch, c.SumAlpha, S, 0.0);
    if (!rc) rc = fill_f64(ch, c.Eta, K, 1.0 / K);
    if (!rc) rc = fill_f64(ch, c.Gamma, S * K, 1.0 / K);
    if (!rc) rc = fill_f64(ch, c.Di, S * K, 0.0);
    // InitFmodel: F = pmeanf, f = (1-F)/F, ancestral frequencies from the
    // total-population frequencies
    if (!rc) {
        std::vector<double> fst(2 * K);
        for (int k = 0; k < K; ++k) { fst[k] = p->pmeanf; fst[K + k] = (1 - fst[k]) / fst[k]; }
        e = hipMemcpy(c.Fst, fst.data(), fst.size() * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "upload F");
    }
    if (!rc && c.fmodel) {
        std::vector<double> tot(KT), nh(d.L);
        std::vector<int32_t> nhi(d.L);
        e = hipMemcpy(tot.data(), d.totfreq, KT * 8, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(nhi.data(), d.nhaplo, d.L * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "read frequencies");
        std::vector<double> fa(3 * KT, 0.0);
        for (int64_t l = 0; !rc && l < d.L; ++l) {
            const int k2 = ds->k2[l];
            for (int a = 0; a < k2; ++a)
                fa[ds->afo[l] + a] = (p->lambda + tot[ds->afo[l] + a] * nhi[l]) / (k2 * p->lambda + nhi[l]);
        }
        if (!rc && hipMemcpy(c.FreqA, fa.data(), fa.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(VPS_ERR_DEVICE, "upload ancestral frequencies");
    }
    // InitAdmix (src/structure.cpp:557-649)
    ch->singlez = ch->nadmburnin == 0 && !c.admix;
    if (!rc) rc = launch_init_admix(d, c, ch->singlez, ch->stream);
    if (!rc) { e = hipStreamSynchronize(ch->stream); if (e != hipSuccess) rc = fail(VPS_ERR_DEVICE, "init: %s", hipGetErrorString(e)); }
    if (rc) { chain_free(ch); return rc; }
    if (p->z_chunk == 0) {
        rc = tune_chunks(ch);
        if (rc) { chain_free(ch); return rc; }
    }
    *out = ch;
    return VPS_OK;
}

extern "C" int vps_chain_destroy(vps_chain* ch)
{
    chain_free(ch);
    return VPS_OK;
}

namespace {

int do_update_p(vps_chain* ch) { return launch_update_p(ch->ds->d, ch->c, ch->stream); }

int do_q_noadmix(vps_chain* ch)
{
    int rc = launch_lik_pass(ch->ds->d, ch->c, 3, ch->stream);
    if (!rc) rc = launch_q_noadmix_draw(ch->ds->d, ch->c, ch->stream);
    return rc;
}

int do_q_admix(vps_chain* ch) { return launch_update_q_admix(ch->ds->d, ch->c, ch->stream); }

int do_q_metro(vps_chain* ch)
{
    int rc = launch_q_metro_propose(ch->ds->d, ch->c, ch->stream);
    if (!rc) rc = launch_lik_pass(ch->ds->d, ch->c, 2, ch->stream);
    if (!rc) rc = launch_q_metro_accept(ch->ds->d, ch->c, ch->stream);
    return rc;
}

int do_z_admix(vps_chain* ch, bool clear_mi = true, bool rec = false, bool advance = false, bool qf_fresh = false)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // timing samples the plain pass: a Record sweep's pass (which also sums
    // lnL) neither is timed nor counts
    if (ch->timing && !rec && (ch->timing_every <= 1 || ch->timed_calls++ % ch->timing_every == 0)) {
        if (ch->ev_used + 2 > ch->evs.size()) {
            for (int x = 0; x < 64; ++x) {
                hipEvent_t e;
                VPS_HIP(hipEventCreate(&e));
                ch->evs.push_back(e);
            }
        }
        e0 = ch->evs[ch->ev_used];
        e1 = ch->evs[ch->ev_used + 1];
        ch->ev_used += 2;
    }
    return launch_z_admix(ch->ds->d, ch->c, ch->stream, e0, e1, clear_mi, rec, advance, qf_fresh);
}

int do_z_noadmix(vps_chain* ch)
{
    int rc = launch_set_mi_binary(ch->ds->d, ch->c, ch->stream);
    if (!rc) rc = launch_z_noadmix(ch->ds->d, ch->c, ch->stream);
    return rc;
}

}  // namespace

namespace vps {

int set_chunks(vps_chain* ch, int G)
{
    ChainDev& c = ch->c;
    const int N = c.N;
    const int Nc = chunk_size(N, std::max(1, G), ch->lnl_groups);
    const int Gr = (N + Nc - 1) / Nc;
    if (Gr > c.Gcap) return fail(VPS_ERR_INVALID, "%d chunks: the chain's partial counts hold %d", Gr, c.Gcap);
    if (Gr == c.G) return VPS_OK;
    const uint32_t* jump = nullptr;
    int rc = dataset_jump(ch->ds, Nc, c.precision == VPS_PREC_FAST32, &jump);
    if (rc) return rc;
    VPS_HIP(hipStreamSynchronize(ch->stream));
    for (auto& kv : ch->graphs) (void)hipGraphExecDestroy(kv.second);
    ch->graphs.clear();
    c.Nc = Nc;
    c.G = Gr;
    c.jump = jump;
    return VPS_OK;
}

// In a batch the other chains' blocks keep the device busy, so a chain takes
// chunks of about kBatchChunk individuals (fewer jump-aheads, partial counts
// and Mi flushes per pass) unless its own chunks are coarser or were fixed by
// z_chunk > 0.  Measured on 64 chains at N = 1,000, L = 50,000: 10.68 ->
// 10.50 ms per sweep of the 64 in FP64, 7.47 -> 6.91 ms with FP32 weights
// (chunks of 125 to 512 tried; 192 to 320 best).
constexpr int kBatchChunk = 256;

int batch_chunk_count(const vps_chain* ch)
{
    if (ch->par.z_chunk > 0) return ch->c.G;
    const int N = ch->c.N;
    const int G = (N + kBatchChunk - 1) / kBatchChunk;
    return std::min(ch->c.G, std::max(1, G));
}

// What sweep m does, decided on the host from m and the run schedule
// (MCMC loop, src/structure.cpp:2122-2158).
SweepPlan plan_sweep(const vps_chain* ch)
{
    const ChainDev& c = ch->c;
    const int m = ch->m;
    SweepPlan p{};
    // UpdateQ, src/structure.cpp:1140-1166
    if (!c.admix) p.q = m >= ch->nadmburnin ? kQNoAdmix : kQAdmix;
    else p.q = (ch->metrofreq > 0 && m % ch->metrofreq == 0) ? kQMetro : kQAdmix;
    // Record, src/structure.cpp:1943-1950
    p.record = (ch->writelnl && m < ch->nburnin && m % ch->nthinning == 0) ||
               (m >= ch->nburnin && (m - ch->nburnin) % ch->nthinning == 0);
    p.post = m >= ch->nburnin;
    p.trace = p.record && ch->writelnl && ch->nlnl < ch->lnl_cap;
    return p;
}

// Every launch of sweep m, all on the chain's stream.  Kernels read m from
// the device, and the last launch advances it.
int settle_side(vps_chain* ch)
{
    if (!ch->side_pending) return VPS_OK;
    ch->side_pending = false;
    VPS_HIP(hipStreamWaitEvent(ch->stream, ch->ev_join, 0));
    return VPS_OK;
}

int launch_sweep(vps_chain* ch, const SweepPlan& p)
{
    const DeviceDataset& d = ch->ds->d;
    const ChainDev& c = ch->c;
    // the previous sweep's alpha/lambda/A/F: read by this sweep's UpdateQ and UpdateP
    int rc = settle_side(ch);
    if (rc) return rc;
    // Sweeps with a Dirichlet UpdateQ: UpdateP, UpdateQ and the Mi clear in
    // one step (one launch in the parallel topology; the parallel parse of the
    // reference's streams and its normalisation in the reference topology).
    // Uncaptured, the parameter updates also run beside the Z pass, joined at
    // the start of the next sweep or before Record (they read the sweep index
    // from a copy, so the Z pass's last kernel can advance it); captured, they
    // run in series.
    const bool joint = p.q == kQAdmix;
    const bool fast = joint && !ch->capturing;
    if (joint) {
        rc = c.topology == VPS_RNG_PARALLEL ? launch_update_pq(d, c, ch->stream)
                                            : launch_spp(d, c, true, kSppQAdmix, true, ch->stream);
    } else if (p.q == kQMetro && c.topology == VPS_RNG_REFERENCE) {
        // UpdateP's streams, and the proposals' stream as a stream graph (or
        // in the same window parse)
        if (pg_applies(c) && !ch->capturing) {
            // the proposals' graph (UpdateQ's stream) and UpdateP's streams
            // read and write disjoint state: the graph runs on the second
            // stream beside the P streams' parse, joined before the ratio pass
            VPS_HIP(hipEventRecord(ch->ev_fork, ch->stream));
            VPS_HIP(hipStreamWaitEvent(ch->aux, ch->ev_fork, 0));
            rc = launch_pg_proposals(d, c, ch->aux);
            if (!rc) VPS_HIP(hipEventRecord(ch->ev_join, ch->aux));
            if (!rc) rc = launch_spp(d, c, true, kSppQNone, false, ch->stream);
            if (!rc) VPS_HIP(hipStreamWaitEvent(ch->stream, ch->ev_join, 0));
        } else if (pg_applies(c)) {
            rc = launch_spp(d, c, true, kSppQNone, false, ch->stream);
            if (!rc) rc = launch_pg_proposals(d, c, ch->stream);
        } else {
            rc = launch_spp(d, c, true, kSppQProposals, false, ch->stream);
        }
        if (!rc) rc = launch_lik_pass(d, c, 2, ch->stream);
        if (!rc) rc = launch_q_metro_accept(d, c, ch->stream);
    } else {
        rc = do_update_p(ch);
        if (!rc) rc = p.q == kQNoAdmix ? do_q_noadmix(ch) : p.q == kQMetro ? do_q_metro(ch) : do_q_admix(ch);
    }
    if (!rc && !c.admix && c.locpriori) rc = launch_update_di(d, c, ch->stream);
    if (!rc) rc = launch_update_locpriori(d, c, ch->stream);
    // UpdateAlpha, UpdateLambda, UpdateA and UpdateF read Q and P and write
    // only their own parameters, none of which the Z pass reads or writes:
    // they run beside it on the chain's second stream (captured sweeps keep
    // one stream: graph branches cost more than they hide)
    hipStream_t side = ch->capturing ? ch->stream : ch->aux;
    ChainDev cs = c;
    if (fast) cs.mdev = c.mside;
    if (!rc && side != ch->stream) {
        VPS_HIP(hipEventRecord(ch->ev_fork, ch->stream));
        VPS_HIP(hipStreamWaitEvent(side, ch->ev_fork, 0));
    }
    if (!rc) rc = launch_update_alpha(d, cs, side);
    if (!rc) rc = launch_update_lambda(d, cs, side);
    if (!rc) rc = launch_update_a(d, cs, side);
    if (!rc) rc = launch_update_f(d, cs, side);
    // Record's admixture lnL: after a Metropolis UpdateQ, the sum of the
    // per-individual log likelihoods that pass left under the kept Q (the
    // same P and Q Record reads); otherwise from the FP64 Z pass itself, whose
    // weights start from the same products P[k][l][x] Q[i][k].  With
    // single-precision ratios (FAST32) the kept sum would carry their
    // rounding, so Record takes its own pass.
    const bool mrec = p.record && p.q == kQMetro && c.precision != VPS_PREC_FAST32;
    const bool zrec = p.record && !mrec && p.q != kQNoAdmix && c.precision != VPS_PREC_MIXED &&
                      c.precision != VPS_PREC_FAST32 && ch->lnl_groups && z_lnl_aligned(c);
    // without Record, the Z pass's last kernel advances the sweep index when
    // nothing beside it still reads the index
    const bool zadv = !p.record && p.q != kQNoAdmix && (fast || side == ch->stream);
    // (the joint P/Q step writes Q's single-precision copy beside Q)
    if (!rc) rc = p.q == kQNoAdmix ? do_z_noadmix(ch) : do_z_admix(ch, !joint, zrec, zadv, joint);
    // Record reads alpha, lambda and F: a Record sweep joins here
    if (!rc && side != ch->stream) {
        VPS_HIP(hipEventRecord(ch->ev_join, side));
        if (fast && !p.record) ch->side_pending = true;
        else VPS_HIP(hipStreamWaitEvent(ch->stream, ch->ev_join, 0));
    }
    if (!rc && p.record) {
        if (mrec) rc = launch_reduce_lnl(c.bufN2, d.N, c.scal + 1, ch->stream);
        else if (!zrec) rc = launch_lik_pass(d, c, p.q == kQNoAdmix ? 1 : 0, ch->stream);
        if (!rc) rc = launch_record_accumulate(d, c, p.post, p.trace, 1, ch->stream);
    }
    if (!rc && !zadv) rc = launch_advance_m(c, ch->stream);
    return rc;
}

// host bookkeeping after sweep m was launched
void commit_sweep(vps_chain* ch, const SweepPlan& p)
{
    ch->binaryq = p.q == kQNoAdmix;
    if (p.trace) ch->nlnl++;
    if (p.record && p.post) ch->nr++;
    ch->m++;
}

}  // namespace vps

namespace {

// One sweep as a cached graph of its plan, captured on first use.  Relaxed
// capture mode: other host threads (a server's loaders) may allocate, copy
// and synchronise while this thread captures.  If a capture still fails,
// the chain falls back to kernel-by-kernel launches for good; a failed
// capture has run nothing, so the sweep is then simply launched directly.
int graph_sweep(vps_chain* ch, const SweepPlan& p, bool* launched)
{
    *launched = false;
    auto it = ch->graphs.find(p.code());
    if (it == ch->graphs.end()) {
        if (hipStreamBeginCapture(ch->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            ch->graph_off = true;
            return VPS_OK;
        }
        ch->capturing = true;
        int rc = launch_sweep(ch, p);
        ch->capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ch->stream, &g);
        hipGraphExec_t exec = nullptr;
        if (!rc && e == hipSuccess && g) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (rc || e != hipSuccess || !exec) {
            (void)hipGetLastError();
            if (exec) (void)hipGraphExecDestroy(exec);
            ch->graph_off = true;
            return VPS_OK;
        }
        it = ch->graphs.emplace(p.code(), exec).first;
    }
    VPS_HIP(hipGraphLaunch(it->second, ch->stream));
    *launched = true;
    return VPS_OK;
}

int do_sweep(vps_chain* ch)
{
    const SweepPlan p = plan_sweep(ch);
    // timed sweeps record events around the Z pass, so they launch kernel by kernel
    bool launched = false;
    int rc = VPS_OK;
    if (ch->par.sweep_graphs && !ch->timing && !ch->graph_off) rc = graph_sweep(ch, p, &launched);
    if (!rc && !launched) rc = launch_sweep(ch, p);
    if (!rc) commit_sweep(ch, p);
    return rc;
}

}  // namespace

extern "C" int vps_chain_step(vps_chain* ch, int n)
{
    if (!ch) return fail(VPS_ERR_INVALID, "null chain");
    if (ch->finished) return fail(VPS_ERR_STATE, "chain already finished");
    VPS_HIP(hipSetDevice(ch->ds->device));
    for (int s = 0; s < n; ++s) {
        int rc = do_sweep(ch);
        if (rc) return rc;
    }
    return settle_side(ch);   // the chain's stream is again the point of completion
}

extern "C" int vps_chain_sync(vps_chain* ch)
{
    if (!ch) return fail(VPS_ERR_INVALID, "null chain");
    VPS_HIP(hipStreamSynchronize(ch->stream));
    return VPS_OK;
}

extern "C" int vps_chain_finish(vps_chain* ch)
{
    if (!ch) return fail(VPS_ERR_INVALID, "null chain");
    if (ch->finished) return VPS_OK;
    // Arrange divides the accumulators by the number of recorded sampling
    // sweeps; a run stopped before its first one has no result
    if (ch->nr == 0)
        return fail(VPS_ERR_STATE, "no sampling sweep recorded yet (%d sweeps run, burn-in %d)", ch->m, ch->nburnin);
    int rc = launch_arrange(ch->ds->d, ch->c, ch->nr, ch->stream);
    if (rc) return rc;
    VPS_HIP(hipStreamSynchronize(ch->stream));
    ch->finished = 1;
    return VPS_OK;
}

extern "C" int vps_chain_get_sizes(vps_chain* ch, int64_t* rlen, int64_t* nlnl)
{
    if (!ch) return fail(VPS_ERR_INVALID, "null chain");
    if (rlen) *rlen = ch->c.rlen;
    if (nlnl) *nlnl = ch->nlnl;
    return VPS_OK;
}

// single passes at the current sweep index (the BAYESIAN::*CUDA seam)
#define VPS_PASS_PROLOGUE(ch)                                                   \
    if (!(ch)) return fail(VPS_ERR_INVALID, "null chain");                       \
    if ((ch)->finished) return fail(VPS_ERR_STATE, "chain already finished");    \
    VPS_HIP(hipSetDevice((ch)->ds->device));

extern "C" int vps_pass_update_z_admix(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = do_z_admix(ch);
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

extern "C" int vps_pass_update_z_noadmix(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = do_z_noadmix(ch);
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

extern "C" int vps_pass_update_q_noadmix(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = do_q_noadmix(ch);
    ch->binaryq = 1;
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

extern "C" int vps_pass_update_q_metro(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = do_q_metro(ch);
    ch->binaryq = 0;
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

extern "C" int vps_pass_update_q_admix(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = do_q_admix(ch);
    ch->binaryq = 0;
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

extern "C" int vps_pass_update_p(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = do_update_p(ch);
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

extern "C" int vps_pass_update_alpha(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = launch_update_alpha(ch->ds->d, ch->c, ch->stream);
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

// the genotype pass of Record alone (lnL into VPS_STATE_LNL; MiSum += Mi)
extern "C" int vps_pass_record(vps_chain* ch)
{
    VPS_PASS_PROLOGUE(ch);
    int rc = launch_lik_pass(ch->ds->d, ch->c, ch->binaryq ? 1 : 0, ch->stream);
    if (!rc) rc = launch_record_accumulate(ch->ds->d, ch->c, 0, 0, 1, ch->stream);
    if (!rc) VPS_HIP(hipStreamSynchronize(ch->stream));
    return rc;
}

// ---------------------------------------------------------------------------
// state transfer
// ---------------------------------------------------------------------------
namespace {

int copy_dev(vps_chain* ch, void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    VPS_HIP(hipStreamSynchronize(ch->stream));
    VPS_HIP(hipMemcpy(dst, src, bytes, kind));
    return VPS_OK;
}

}  // namespace

extern "C" int vps_chain_get_state(vps_chain* ch, int which, void* host, size_t bytes)
{
    if (!ch || !host) return fail(VPS_ERR_INVALID, "null argument");
    VPS_HIP(hipSetDevice(ch->ds->device));
    const ChainDev& c = ch->c;
    const int64_t N = c.N, K = c.K, KT = c.KT, S = c.S;
    auto need = [&](size_t b) { return bytes == b ? VPS_OK : fail(VPS_ERR_INVALID, "state %d: %zu bytes given, %zu needed", which, bytes, b); };
    int rc;
    switch (which) {
    case VPS_STATE_FREQ: if ((rc = need(K * KT * 8))) return rc; return copy_dev(ch, host, c.Freq, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_Q: if ((rc = need(N * K * 8))) return rc; return copy_dev(ch, host, c.Q, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_Z: if ((rc = need(N * 2))) return rc; return copy_dev(ch, host, c.Z, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_NI: if ((rc = need(K * KT * 4))) return rc; return copy_dev(ch, host, c.Ni, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_MI: {
        if ((rc = need(N * K * 8))) return rc;
        std::vector<int32_t> t(N * K);
        if ((rc = copy_dev(ch, t.data(), c.Mi, N * K * 4, hipMemcpyDeviceToHost))) return rc;
        for (int64_t x = 0; x < N * K; ++x) ((int64_t*)host)[x] = t[x];
        return VPS_OK;
    }
    case VPS_STATE_ALPHA: if ((rc = need(K * 8))) return rc; return copy_dev(ch, host, c.Alpha, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_LAMBDA: if ((rc = need(K * 8))) return rc; return copy_dev(ch, host, c.Lambda, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_ITER: if ((rc = need(4))) return rc; *(int32_t*)host = ch->m; return VPS_OK;
    case VPS_STATE_BUFN1: if ((rc = need(N * 8))) return rc; return copy_dev(ch, host, c.bufN1, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_BUFNK1: if ((rc = need(N * K * 8))) return rc; return copy_dev(ch, host, c.bufNK1, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_LNL: if ((rc = need(8))) return rc; return copy_dev(ch, host, c.scal + 1, 8, hipMemcpyDeviceToHost);
    case VPS_STATE_MISUM: if ((rc = need(N * K * 8))) return rc; return copy_dev(ch, host, c.MiSum, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_FREQM: if ((rc = need(K * KT * 8))) return rc; return copy_dev(ch, host, c.FreqM, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_ROUT: if ((rc = need((size_t)c.rlen * 8))) return rc; return copy_dev(ch, host, c.rout, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_FREQA: if ((rc = need(3 * KT * 8))) return rc; return copy_dev(ch, host, c.FreqA, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_FST: if ((rc = need(2 * K * 8))) return rc; return copy_dev(ch, host, c.Fst, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_BINARYQ: if ((rc = need(4))) return rc; *(int32_t*)host = ch->binaryq; return VPS_OK;
    case VPS_STATE_LNLTRACE: if ((rc = need((size_t)ch->nlnl * 8))) return rc;
        return ch->nlnl ? copy_dev(ch, host, c.lnltrace, bytes, hipMemcpyDeviceToHost) : VPS_OK;
    case VPS_STATE_QF: if ((rc = need(N * K * 4))) return rc; return copy_dev(ch, host, c.Qf, bytes, hipMemcpyDeviceToHost);
    case VPS_STATE_LOC: {
        if ((rc = need((size_t)(1 + K + 3 * S * K + S) * 8))) return rc;
        double* h = (double*)host;
        if ((rc = copy_dev(ch, h, c.scal, 8, hipMemcpyDeviceToHost))) return rc;
        if ((rc = copy_dev(ch, h + 1, c.Eta, K * 8, hipMemcpyDeviceToHost))) return rc;
        if ((rc = copy_dev(ch, h + 1 + K, c.Gamma, S * K * 8, hipMemcpyDeviceToHost))) return rc;
        if ((rc = copy_dev(ch, h + 1 + K + S * K, c.AlphaLocal, S * K * 8, hipMemcpyDeviceToHost))) return rc;
        if ((rc = copy_dev(ch, h + 1 + K + 2 * S * K, c.Di, S * K * 8, hipMemcpyDeviceToHost))) return rc;
        return copy_dev(ch, h + 1 + K + 3 * S * K, c.SumAlpha, S * 8, hipMemcpyDeviceToHost);
    }
    default: return fail(VPS_ERR_INVALID, "unknown state %d", which);
    }
}

extern "C" int vps_chain_set_state(vps_chain* ch, int which, const void* host, size_t bytes)
{
    if (!ch || !host) return fail(VPS_ERR_INVALID, "null argument");
    VPS_HIP(hipSetDevice(ch->ds->device));
    ChainDev& c = ch->c;
    const int64_t N = c.N, K = c.K, KT = c.KT, S = c.S;
    auto need = [&](size_t b) { return bytes == b ? VPS_OK : fail(VPS_ERR_INVALID, "state %d: %zu bytes given, %zu needed", which, bytes, b); };
    int rc;
    switch (which) {
    case VPS_STATE_FREQ: if ((rc = need(K * KT * 8))) return rc; return copy_dev(ch, c.Freq, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_Q: if ((rc = need(N * K * 8))) return rc; return copy_dev(ch, c.Q, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_Z: {
        if ((rc = need(N * 2))) return rc;
        for (int64_t i = 0; i < N; ++i)
            if (((const uint16_t*)host)[i] >= K) return fail(VPS_ERR_INVALID, "Z[%lld] >= K", (long long)i);
        return copy_dev(ch, c.Z, host, bytes, hipMemcpyHostToDevice);
    }
    case VPS_STATE_NI: if ((rc = need(K * KT * 4))) return rc; return copy_dev(ch, c.Ni, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_MI: {
        if ((rc = need(N * K * 8))) return rc;
        std::vector<int32_t> t(N * K);
        for (int64_t x = 0; x < N * K; ++x) t[x] = (int32_t)((const int64_t*)host)[x];
        return copy_dev(ch, c.Mi, t.data(), N * K * 4, hipMemcpyHostToDevice);
    }
    case VPS_STATE_ALPHA: if ((rc = need(K * 8))) return rc; return copy_dev(ch, c.Alpha, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_LAMBDA: if ((rc = need(K * 8))) return rc; return copy_dev(ch, c.Lambda, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_ITER:
        if ((rc = need(4))) return rc;
        ch->m = *(const int32_t*)host;
        return copy_dev(ch, ch->c.mdev, host, 4, hipMemcpyHostToDevice);
    case VPS_STATE_BUFN1: if ((rc = need(N * 8))) return rc; return copy_dev(ch, c.bufN1, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_BUFNK1: if ((rc = need(N * K * 8))) return rc; return copy_dev(ch, c.bufNK1, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_MISUM: if ((rc = need(N * K * 8))) return rc; return copy_dev(ch, c.MiSum, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_FREQM: if ((rc = need(K * KT * 8))) return rc; return copy_dev(ch, c.FreqM, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_ROUT: if ((rc = need((size_t)c.rlen * 8))) return rc; return copy_dev(ch, c.rout, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_FREQA: if ((rc = need(3 * KT * 8))) return rc; return copy_dev(ch, c.FreqA, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_FST: if ((rc = need(2 * K * 8))) return rc; return copy_dev(ch, c.Fst, host, bytes, hipMemcpyHostToDevice);
    case VPS_STATE_BINARYQ: if ((rc = need(4))) return rc; ch->binaryq = *(const int32_t*)host ? 1 : 0; return VPS_OK;
    case VPS_STATE_LOC: {
        if ((rc = need((size_t)(1 + K + 3 * S * K + S) * 8))) return rc;
        const double* h = (const double*)host;
        if ((rc = copy_dev(ch, c.scal, h, 8, hipMemcpyHostToDevice))) return rc;
        if ((rc = copy_dev(ch, c.Eta, h + 1, K * 8, hipMemcpyHostToDevice))) return rc;
        if ((rc = copy_dev(ch, c.Gamma, h + 1 + K, S * K * 8, hipMemcpyHostToDevice))) return rc;
        if ((rc = copy_dev(ch, c.AlphaLocal, h + 1 + K + S * K, S * K * 8, hipMemcpyHostToDevice))) return rc;
        if ((rc = copy_dev(ch, c.Di, h + 1 + K + 2 * S * K, S * K * 8, hipMemcpyHostToDevice))) return rc;
        return copy_dev(ch, c.SumAlpha, h + 1 + K + 3 * S * K, S * 8, hipMemcpyHostToDevice);
    }
    default: return fail(VPS_ERR_INVALID, "state %d is read-only or unknown", which);
    }
}

// Z-pass timing: a timed pass is bracketed by its own pair of events on the
// chain's stream, so timing adds no host synchronisation to the sweep.
// `every` = n times every n-th Z pass (1 = all), which keeps the cost of the
// event records out of most sweeps.
extern "C" int vps_chain_enable_timing(vps_chain* ch, int every)
{
    if (!ch) return fail(VPS_ERR_INVALID, "null chain");
    if (every < 0) return fail(VPS_ERR_INVALID, "timing period %d", every);
    VPS_HIP(hipSetDevice(ch->ds->device));
    VPS_HIP(hipStreamSynchronize(ch->stream));
    ch->timing = every != 0;
    ch->timing_every = every;
    ch->timed_calls = 0;
    ch->ev_used = 0;
    return VPS_OK;
}

extern "C" int vps_chain_kernel_time(vps_chain* ch, int kernel, double* total_ms, int64_t* launches)
{
    if (!ch || !total_ms || !launches) return fail(VPS_ERR_INVALID, "null argument");
    if (kernel != 0) return fail(VPS_ERR_INVALID, "kernel id %d", kernel);
    VPS_HIP(hipStreamSynchronize(ch->stream));
    double tot = 0;
    for (size_t x = 0; x + 1 < ch->ev_used; x += 2) {
        float ms = 0;
        VPS_HIP(hipEventElapsedTime(&ms, ch->evs[x], ch->evs[x + 1]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int64_t)(ch->ev_used / 2);
    return VPS_OK;
}

extern "C" int vps_debug_chunking(const vps_chain* ch, int32_t* G, int32_t* Nc)
{
    if (!ch || !G || !Nc) return fail(VPS_ERR_INVALID, "null argument");
    *G = ch->c.G;
    *Nc = ch->c.Nc;
    return VPS_OK;
}

extern "C" int vps_debug_gen_waves(const vps_chain* ch, int32_t* waves)
{
    if (!ch || !waves) return fail(VPS_ERR_INVALID, "null argument");
    int w = 0;
    const int rc = z_gen_block_waves(ch->ds->d, ch->K, ch->c.precision, ch->c.Nc, &w);
    *waves = w;
    return rc;
}

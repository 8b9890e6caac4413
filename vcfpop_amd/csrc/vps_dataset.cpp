// vps_dataset.cpp — builds the device genotype image from the vcfpop-style
// host description (vps_dataset_desc) and performs CalcBayesian's prologue:
// locus sort by (bitsize, original index), allele_freq_offset and KT
// (src/structure.cpp:2350-2381), individual vt / ploidy range
// (src/function.cpp:4904-4915) and total-population frequencies for the F
// model (POP::CalcFreqGcount, src/function.cpp:3974-4017).
#include <algorithm>
#include <numeric>
#include <string.h>
#include <thread>
#include <mutex>
#include "vps_internal.h"

using namespace vps;

namespace vps {

// Stream-ordered allocations from the device's default pool, which keeps
// freed blocks: releasing a finished chain or dataset then neither returns
// memory to the driver nor synchronises the device, so jobs that still run
// beside it are not stalled.
int dev_alloc(void** p, size_t bytes, std::vector<void*>& list, int64_t* counter, hipStream_t st)
{
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    static std::once_flag pool_once[64];
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64)
        std::call_once(pool_once[dev], [dev] {
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t keep = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
        });
    hipError_t e = hipMallocAsync(p, bytes, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(VPS_ERR_NOMEM, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    list.push_back(*p);
    if (counter) *counter += (int64_t)bytes;
    return VPS_OK;
}

void dev_free_all(std::vector<void*>& list, hipStream_t st)
{
    for (void* p : list) (void)hipFreeAsync(p, st);
    list.clear();
    (void)hipStreamSynchronize(st);
}

}  // namespace vps

namespace {

template <typename T>
int upload(T** dst, const std::vector<T>& src, vps_dataset* ds)
{
    int rc = dev_alloc((void**)dst, src.size() * sizeof(T), ds->allocations, &ds->device_bytes, ds->stream);
    if (rc) return rc;
    if (!src.empty()) VPS_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return VPS_OK;
}

inline uint32_t read_id(const uint8_t* bits, uint64_t byteoff, int b, int64_t i, uint64_t nbytes)
{
    uint64_t bit = (uint64_t)i * b;
    uint64_t pos = byteoff + (bit >> 3);
    uint64_t w = 0;
    int nb = (int)std::min<uint64_t>(8, nbytes - pos);
    memcpy(&w, bits + pos, nb);
    return (uint32_t)((w >> (bit & 7)) & ((1ull << b) - 1ull));
}

template <typename F>
void parallel_for(int64_t n, F&& f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64 || nt == 1) { for (int64_t i = 0; i < n; ++i) f(i); return; }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() { for (int64_t i = t; i < n; i += nt) f(i); });
    for (auto& x : th) x.join();
}

}  // namespace

namespace {

// GF(2) jump tables of the parallel stream parse: tables l < 64 advance an
// xorshift128+ state by 8*l draws, tables 64 + w by 512*w draws, so thread
// 64w + l of a 1024-thread block reaches its ring offset 8*(64w + l) in two
// jumps.  The same for every dataset; built once.
const std::vector<uint32_t>& ring_jump_tables()
{
    static std::once_flag once;
    static std::vector<uint32_t> tabs;
    std::call_once(once, [] {
        tabs.resize((size_t)kRingJumps * kJumpWords);
        for (int l = 0; l < kRingLaneTabs; ++l) jump_table((uint64_t)8 * l, tabs.data() + (size_t)l * kJumpWords);
        for (int w = 0; w < kRingWaveTabs; ++w)
            jump_table((uint64_t)512 * w, tabs.data() + (size_t)(kRingLaneTabs + w) * kJumpWords);
        for (int j = 0; j < kRingPow2Tabs; ++j)
            jump_table((uint64_t)4096 << j, tabs.data() + (size_t)(kRingPow2Base + j) * kJumpWords);
    });
    return tabs;
}

}  // namespace

extern "C" int vps_dataset_create(const vps_dataset_desc* in, int device, vps_dataset** out)
{
    if (!in || !out) return fail(VPS_ERR_INVALID, "null argument");
    *out = nullptr;
    const int64_t L = in->n_loc;
    const int N = in->n_ind;
    if (N <= 0 || L <= 0) return fail(VPS_ERR_INVALID, "empty dataset (N=%d, L=%lld)", N, (long long)L);
    if (N > 1000000) return fail(VPS_ERR_UNSUPPORTED, "N=%d exceeds 1e6", N);
    if (!in->geno_bits || !in->geno_byte_offset || !in->bitsize || !in->n_allele || !in->gtab_offset ||
        !in->gt_ploidy || !in->gt_allele_offset || !in->allele_pool)
        return fail(VPS_ERR_INVALID, "missing dataset array");
    int S = in->n_pop > 0 ? in->n_pop : 1;

    const bool host_only = device < 0;
    if (!host_only) VPS_HIP(hipSetDevice(device));
    vps_dataset* ds = new vps_dataset();
    ds->device = device;
    ds->host_only = host_only;
    DeviceDataset& d = ds->d;
    d.N = N; d.L = L; d.S = S;

    // byte extent of the packed ids
    uint64_t nbytes = 0;
    for (int64_t l = 0; l < L; ++l) {
        int b = in->bitsize[l];
        if (b < 1 || b > 24) { delete ds; return fail(VPS_ERR_INVALID, "locus %lld: bitsize %d out of range", (long long)l, b); }
        nbytes = std::max<uint64_t>(nbytes, in->geno_byte_offset[l] + ((uint64_t)b * N + 7) / 8);
        int64_t ng = in->gtab_offset[l + 1] - in->gtab_offset[l];
        if (ng <= 0 || ng > (1ll << b)) { delete ds; return fail(VPS_ERR_INVALID, "locus %lld: %lld genotypes for %d bits", (long long)l, (long long)ng, b); }
        int k2 = in->n_allele[l];
        if (k2 < 1) { delete ds; return fail(VPS_ERR_INVALID, "locus %lld: %d alleles", (long long)l, k2); }
        if (k2 > VPS_MAX_ALLELES) { delete ds; return fail(VPS_ERR_UNSUPPORTED, "locus %lld has %d alleles; the device path supports at most %d", (long long)l, k2, VPS_MAX_ALLELES); }
    }

    // locus order (QSLocus2, src/structure.cpp:2202-2273): (bitsize, original index)
    ds->orig.resize(L);
    std::iota(ds->orig.begin(), ds->orig.end(), 0);
    std::stable_sort(ds->orig.begin(), ds->orig.end(),
                     [&](int64_t a, int64_t b) { return in->bitsize[a] < in->bitsize[b]; });
    ds->afo.resize(L); ds->k2.resize(L); ds->bitsize.resize(L);
    int64_t KT = 0; int maxK = 0;
    for (int64_t l = 0; l < L; ++l) {
        int64_t o = ds->orig[l];
        ds->afo[l] = KT;
        ds->k2[l] = in->n_allele[o];
        ds->bitsize[l] = in->bitsize[o];
        KT += ds->k2[l];
        maxK = std::max(maxK, ds->k2[l]);
    }
    if (KT >= (1ll << 31)) { delete ds; return fail(VPS_ERR_UNSUPPORTED, "KT=%lld exceeds 2^31", (long long)KT); }
    d.KT = KT; d.maxK = maxK;
    ds->kdis.assign(maxK + 1, 0);
    for (int64_t l = 0; l < L; ++l) ds->kdis[ds->k2[l]]++;

    // validate genotype tables and ids; per-individual vt and ploidy range
    std::vector<int64_t> vt(N, 0);
    std::vector<int> vmin(N, 100), vmax(N, 0);
    std::vector<double> totfreq(KT, 0.0);
    std::vector<int32_t> ncopies(KT, 0);   // typed copies of each allele over all individuals
    std::vector<int32_t> nhaplo(L, 0);
    std::vector<int> bad(L, 0);
    {
        // per-locus pass (thread-parallel over loci; per-thread ind stats merged after)
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::vector<int64_t>> tvt(nt, std::vector<int64_t>(N, 0));
        std::vector<std::vector<int>> tmin(nt, std::vector<int>(N, 100)), tmax(nt, std::vector<int>(N, 0));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                for (int64_t l = t; l < L; l += nt) {
                    int64_t o = ds->orig[l];
                    int b = in->bitsize[o];
                    int64_t g0 = in->gtab_offset[o], ng = in->gtab_offset[o + 1] - g0;
                    int k2 = in->n_allele[o];
                    for (int64_t g = 0; g < ng; ++g) {
                        int v = in->gt_ploidy[g0 + g];
                        if (v > VPS_MAX_PLOIDY) { bad[l] = 3; break; }
                        for (int a = 0; a < v; ++a)
                            if (in->allele_pool[in->gt_allele_offset[g0 + g] + a] >= k2) { bad[l] = 2; break; }
                    }
                    if (bad[l]) continue;
                    double* p = totfreq.data() + ds->afo[l];
                    int nh = 0;
                    for (int i = 0; i < N; ++i) {
                        uint32_t id = read_id(in->geno_bits, in->geno_byte_offset[o], b, i, nbytes);
                        if (id >= (uint64_t)ng) { bad[l] = 1; break; }
                        int v = in->gt_ploidy[g0 + id];
                        if (!v) continue;
                        tvt[t][i] += v;
                        tmin[t][i] = std::min(tmin[t][i], v);
                        tmax[t][i] = std::max(tmax[t][i], v);
                        nh += v;
                        const uint16_t* als = in->allele_pool + in->gt_allele_offset[g0 + id];
                        for (int a = 0; a < v; ++a) p[als[a]] += 1;
                    }
                    nhaplo[l] = nh;
                    for (int a = 0; a < k2; ++a) ncopies[ds->afo[l] + a] = (int32_t)p[a];
                    unify_row(p, k2);   // CalcFreqGcount's Unify (k2 < 32 path)
                }
            });
        for (auto& x : th) x.join();
        for (int64_t l = 0; l < L; ++l)
            if (bad[l]) {
                int code = bad[l];
                int64_t o = ds->orig[l];
                delete ds;
                if (code == 1) return fail(VPS_ERR_INVALID, "locus %lld: genotype id beyond its table", (long long)o);
                if (code == 2) return fail(VPS_ERR_INVALID, "locus %lld: allele index beyond n_allele", (long long)o);
                return fail(VPS_ERR_UNSUPPORTED, "locus %lld: ploidy above %d", (long long)o, VPS_MAX_PLOIDY);
            }
        for (unsigned t = 0; t < nt; ++t)
            for (int i = 0; i < N; ++i) {
                vt[i] += tvt[t][i];
                vmin[i] = std::min(vmin[i], tmin[t][i]);
                vmax[i] = std::max(vmax[i], tmax[t][i]);
            }
    }
    int maxploidy = 0, minploidy = 100;
    for (int i = 0; i < N; ++i) {
        minploidy = std::min(minploidy, vmin[i] == 100 ? 0 : vmin[i]);
        maxploidy = std::max(maxploidy, vmax[i]);
    }
    if (maxploidy == 0) { delete ds; return fail(VPS_ERR_INVALID, "no typed genotype in the dataset"); }
    d.maxploidy = maxploidy; d.minploidy = minploidy;
    ds->vt = vt;

    // populations (sampling locations): individuals of one population are contiguous
    ds->popid.resize(N);
    ds->pop_ind0.assign(S, -1); ds->pop_nind.assign(S, 0);
    for (int i = 0; i < N; ++i) {
        int s = in->ind_pop ? in->ind_pop[i] : 0;
        if (s < 0 || s >= S) { delete ds; return fail(VPS_ERR_INVALID, "individual %d: population %d out of range", i, s); }
        ds->popid[i] = s;
        if (ds->pop_ind0[s] < 0) ds->pop_ind0[s] = i;
        else if (ds->popid[i - 1] != s) { delete ds; return fail(VPS_ERR_INVALID, "individuals of population %d are not contiguous", s); }
        ds->pop_nind[s]++;
    }
    for (int s = 0; s < S; ++s) if (ds->pop_ind0[s] < 0) ds->pop_ind0[s] = 0;

    // tiles: each run of equal bitsize split into BI and GEN loci, each part
    // cut in sorted order into tiles of up to 64 loci
    std::vector<TileDesc>& tiles = ds->tiles;
    std::vector<int32_t>& tloc = ds->tloc;
    ds->lpos.assign(L, 0);
    auto bi_locus = [&](int64_t l) {
        if (ds->k2[l] > 2 || ds->bitsize[l] > kBiMaxBits) return false;
        const int64_t o = ds->orig[l];
        // a BI entry needs the allele copies in ascending order
        for (int64_t g = in->gtab_offset[o]; g < in->gtab_offset[o + 1]; ++g)
            for (int a = 1; a < in->gt_ploidy[g]; ++a)
                if (in->allele_pool[in->gt_allele_offset[g] + a] < in->allele_pool[in->gt_allele_offset[g] + a - 1]) return false;
        return true;
    };
    {
        int64_t l = 0, wbase = 0, dbase = 0;
        std::vector<int64_t> part[2];
        while (l < L) {
            const int b = ds->bitsize[l];
            int64_t e = l;
            part[0].clear(); part[1].clear();
            for (; e < L && ds->bitsize[e] == b; ++e) part[bi_locus(e) ? 0 : 1].push_back(e);
            for (int kind = 0; kind < 2; ++kind)
                for (size_t s0 = 0; s0 < part[kind].size(); s0 += kTileLoci) {
                    const int n = (int)std::min<size_t>(kTileLoci, part[kind].size() - s0);
                    TileDesc t{};
                    t.lmap = (int64_t)tiles.size() * kTileLoci;
                    t.nloc = n; t.bits = b; t.ipw = 32 / b;
                    t.wbase = wbase;
                    wbase += (int64_t)((N + t.ipw - 1) / t.ipw) * kTileLoci;
                    t.kind = kind == 0 ? kKindBi : kKindGen;
                    t.dbase = -1;
                    if (t.kind == kKindBi) { t.dbase = dbase; dbase += (int64_t)kTileLoci << b; }
                    int mk = 0;
                    for (int j = 0; j < kTileLoci; ++j) {
                        const int64_t lj = part[kind][s0 + (j < n ? j : 0)];
                        tloc.push_back((int32_t)lj);
                        if (j < n) { ds->lpos[lj] = t.lmap + j; mk = std::max(mk, ds->k2[lj]); }
                    }
                    t.maxk2 = mk;
                    tiles.push_back(t);
                }
            l = e;
        }
    }
    d.ntiles = (int64_t)tiles.size();
    std::vector<int32_t> bi_list, gen_list;
    for (int64_t t = 0; t < d.ntiles; ++t) (tiles[t].kind == kKindBi ? bi_list : gen_list).push_back((int32_t)t);
    d.nbi = (int64_t)bi_list.size(); d.ngen = (int64_t)gen_list.size();

    // transposed words
    int64_t total_words = 0;
    for (auto& t : tiles) total_words = std::max<int64_t>(total_words, t.wbase + (int64_t)((N + t.ipw - 1) / t.ipw) * kTileLoci);
    std::vector<uint32_t> words(total_words, 0);
    parallel_for(d.ntiles, [&](int64_t ti) {
        const TileDesc& t = tiles[ti];
        int64_t nwords = (N + t.ipw - 1) / t.ipw;
        for (int j = 0; j < t.nloc; ++j) {
            int64_t o = ds->orig[tloc[t.lmap + j]];
            for (int64_t w = 0; w < nwords; ++w) {
                uint32_t word = 0;
                for (int s = 0; s < t.ipw; ++s) {
                    int64_t i = w * t.ipw + s;
                    if (i >= N) break;
                    word |= read_id(in->geno_bits, in->geno_byte_offset[o], t.bits, i, nbytes) << (s * t.bits);
                }
                words[t.wbase + w * kTileLoci + j] = word;
            }
        }
    });

    // decode tables
    std::vector<uint64_t> bitab(L, 0);
    std::vector<int64_t> gbase(L, 0);
    std::vector<uint64_t> gdesc;
    int64_t btab_bytes = 0;
    int bi_rows = 0;
    for (auto& t : tiles)
        if (t.kind == kKindBi) {
            btab_bytes = std::max<int64_t>(btab_bytes, t.dbase + ((int64_t)kTileLoci << t.bits));
            if (t.bits > kBiRegBits) bi_rows = std::max(bi_rows, 1 << t.bits);
        }
    std::vector<uint8_t> btab(std::max<int64_t>(btab_bytes, 1), 0);
    d.bi_tab_rows = bi_rows;
    for (auto& t : tiles) {
        for (int j = 0; j < t.nloc; ++j) {
            int64_t l = tloc[t.lmap + j], o = ds->orig[l];
            int64_t g0 = in->gtab_offset[o], ng = in->gtab_offset[o + 1] - g0;
            if (t.kind == kKindBi) {
                uint64_t tab = 0;
                for (int64_t g = 0; g < ng; ++g) {
                    int v = in->gt_ploidy[g0 + g];
                    int c = 0;
                    for (int a = 0; a < v; ++a) c += in->allele_pool[in->gt_allele_offset[g0 + g] + a] == 1;
                    const uint8_t e = (uint8_t)(v | (c << 4));
                    btab[t.dbase + g * kTileLoci + j] = e;
                    if (g < 8) tab |= (uint64_t)e << (8 * g);
                }
                bitab[l] = tab;
            } else {
                gbase[l] = (int64_t)gdesc.size();
                for (int64_t g = 0; g < ng; ++g) {
                    int v = in->gt_ploidy[g0 + g];
                    uint64_t e = (uint64_t)v;
                    for (int a = 0; a < v; ++a)
                        e |= (uint64_t)(in->allele_pool[in->gt_allele_offset[g0 + g] + a] & 63) << (4 + 6 * a);
                    gdesc.push_back(e);
                }
            }
        }
    }

    if (in->ind_name) for (int i = 0; i < N; ++i) ds->ind_names.emplace_back(in->ind_name[i] ? in->ind_name[i] : "");
    if (in->pop_name) for (int s = 0; s < S; ++s) ds->pop_names.emplace_back(in->pop_name[s] ? in->pop_name[s] : "");
    if (in->pop_region) for (int s = 0; s < S; ++s) ds->pop_regions.emplace_back(in->pop_region[s] ? in->pop_region[s] : "");
    if (in->loc_name) for (int64_t l = 0; l < L; ++l) ds->loc_names.emplace_back(in->loc_name[ds->orig[l]] ? in->loc_name[ds->orig[l]] : "");

    if (host_only) {
        ds->h_words.swap(words);
        ds->h_bitab.swap(bitab);
        ds->h_btab.swap(btab);
        ds->h_gdesc.swap(gdesc);
        ds->h_gbase.swap(gbase);
        *out = ds;
        return VPS_OK;
    }

    // upload
    {
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&ds->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ds;
            return fail(VPS_ERR_DEVICE, "dataset stream on device %d: %s", device, hipGetErrorString(e));
        }
    }
    std::vector<int32_t> afo32(L), vt32(N);
    for (int64_t l = 0; l < L; ++l) afo32[l] = (int32_t)ds->afo[l];
    for (int i = 0; i < N; ++i) vt32[i] = (int32_t)vt[i];
    int rc = 0;
    if (!rc) rc = upload(&d.tiles, tiles, ds);
    if (!rc) rc = upload(&d.bi_list, bi_list, ds);
    if (!rc) rc = upload(&d.gen_list, gen_list, ds);
    if (!rc) rc = upload(&d.tloc, tloc, ds);
    if (!rc) rc = upload(&d.btab, btab, ds);
    if (!rc) rc = upload(&d.words, words, ds);
    if (!rc) rc = upload(&d.afo, afo32, ds);
    if (!rc) rc = upload(&d.k2, ds->k2, ds);
    if (!rc) rc = upload(&d.bitab, bitab, ds);
    if (!rc) rc = upload(&d.gbase, gbase, ds);
    if (!rc) rc = upload(&d.gdesc, gdesc, ds);
    if (!rc) rc = upload(&d.vt, vt32, ds);
    if (!rc) rc = upload(&d.popid, ds->popid, ds);
    if (!rc) rc = upload(&d.pop_ind0, ds->pop_ind0, ds);
    if (!rc) rc = upload(&d.pop_nind, ds->pop_nind, ds);
    if (!rc) rc = upload(&d.totfreq, totfreq, ds);
    if (!rc) rc = upload(&d.nhaplo, nhaplo, ds);
    if (!rc) rc = upload(&d.ncopies, ncopies, ds);
    if (!rc) rc = upload(&d.kdis, ds->kdis, ds);
    {
        std::vector<int32_t> typed;
        for (int i = 0; i < N; ++i)
            if (vt[i]) typed.push_back(i);
        d.ntyped = (int32_t)typed.size();
        if (!rc) rc = upload(&d.typed, typed, ds);
    }
    if (!rc) rc = upload(&d.ring_jump, ring_jump_tables(), ds);
    if (rc) { vps_dataset_destroy(ds); return rc; }
    *out = ds;
    return VPS_OK;
}

extern "C" int vps_dataset_destroy(vps_dataset* ds)
{
    if (!ds) return VPS_OK;
    if (!ds->host_only) {
        (void)hipSetDevice(ds->device);
        dev_free_all(ds->allocations, ds->stream);
        if (ds->stream) (void)hipStreamDestroy(ds->stream);
    }
    delete ds;
    return VPS_OK;
}

extern "C" int vps_dataset_get_info(const vps_dataset* ds, vps_dataset_info* info)
{
    if (!ds || !info) return fail(VPS_ERR_INVALID, "null argument");
    info->n_ind = ds->d.N;
    info->n_loc = ds->d.L;
    info->kt = ds->d.KT;
    info->max_ploidy = ds->d.maxploidy;
    info->min_ploidy = ds->d.minploidy;
    info->max_alleles = ds->d.maxK;
    info->n_pop = ds->d.S;
    info->n_tiles = ds->d.ntiles;
    info->device_bytes = ds->device_bytes;
    return VPS_OK;
}

extern "C" int vps_dataset_get_order(const vps_dataset* ds, int64_t* orig, int64_t* afo, int64_t* vt)
{
    if (!ds) return fail(VPS_ERR_INVALID, "null dataset");
    if (orig) memcpy(orig, ds->orig.data(), ds->orig.size() * 8);
    if (afo) memcpy(afo, ds->afo.data(), ds->afo.size() * 8);
    if (vt) memcpy(vt, ds->vt.data(), ds->vt.size() * 8);
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// test hooks (host-only datasets)
// ---------------------------------------------------------------------------
extern "C" int vps_debug_decode(const vps_dataset* ds, int64_t l, int32_t i, int32_t* ploidy, uint16_t* alleles)
{
    if (!ds || !ploidy || !alleles) return fail(VPS_ERR_INVALID, "null argument");
    if (!ds->host_only) return fail(VPS_ERR_STATE, "decode needs a host-only dataset (device = -1)");
    if (l < 0 || l >= ds->d.L || i < 0 || i >= ds->d.N) return fail(VPS_ERR_INVALID, "index out of range");
    // the tile holding sorted locus l
    const int64_t pos = ds->lpos[l];
    const TileDesc& t = ds->tiles[pos / kTileLoci];
    const int j = (int)(pos % kTileLoci);
    const uint32_t word = ds->h_words[t.wbase + (int64_t)(i / t.ipw) * kTileLoci + j];
    const uint32_t gid = (word >> ((i % t.ipw) * t.bits)) & ((1u << t.bits) - 1u);
    if (t.kind == kKindBi) {
        const uint32_t desc = t.bits <= kBiRegBits ? (uint32_t)(ds->h_bitab[l] >> (gid * 8)) & 0xFFu
                                                   : ds->h_btab[t.dbase + (int64_t)gid * kTileLoci + j];
        const int v = desc & 15, c = desc >> 4;
        *ploidy = v;
        for (int a = 0; a < v; ++a) alleles[a] = a >= v - c ? 1 : 0;
    } else {
        const uint64_t d = ds->h_gdesc[ds->h_gbase[l] + gid];
        const int v = (int)(d & 15u);
        *ploidy = v;
        for (int a = 0; a < v; ++a) alleles[a] = (uint16_t)((d >> (4 + 6 * a)) & 63u);
    }
    return VPS_OK;
}

extern "C" int vps_debug_jump(uint64_t D, uint64_t* x, uint64_t* y)
{
    if (!x || !y) return fail(VPS_ERR_INVALID, "null argument");
    std::vector<uint32_t> tab(kJumpWords);
    jump_table(D, tab.data());
    jump_apply_host(tab.data(), *x, *y);
    return VPS_OK;
}

extern "C" int vps_debug_mt_accept(const double* u, const double* t, const double* v, const double* d, int64_t n,
                                   uint8_t* accept, int64_t* n_exact)
{
    if (!u || !t || !v || !d || !accept || n < 0) return fail(VPS_ERR_INVALID, "null argument");
    int64_t ex = 0;
    for (int64_t x = 0; x < n; ++x) {
        const MtDecision r = mt_log_decide(u[x], t[x], v[x], d[x]);
        accept[x] = r == kMtAccept ? 1 : r == kMtReject ? 0 : (mt_log_exact(u[x], t[x], v[x], d[x]) ? 1 : 0);
        ex += r == kMtExact;
    }
    if (n_exact) *n_exact = ex;
    return VPS_OK;
}

extern "C" int vps_debug_jump96(uint64_t D, uint32_t* x, uint32_t* y, uint32_t* z)
{
    if (!x || !y || !z) return fail(VPS_ERR_INVALID, "null argument");
    // tables for chunks 0 (identity) and 1 (D steps); apply the second
    std::vector<uint32_t> tab(2 * (size_t)kJumpWords);
    jump_tables96(D, 2, tab.data());
    const uint32_t st[3] = {*x, *y, *z};
    uint32_t r[3] = {0, 0, 0};
    const uint32_t* t = tab.data() + kJumpWords;
    for (int p = 0; p < 24; ++p) {
        const uint32_t nib = (st[p >> 3] >> ((p & 7) * 4)) & 15u;
        for (int w = 0; w < 3; ++w) r[w] ^= t[(p * 16 + nib) * 4 + w];
    }
    *x = r[0]; *y = r[1]; *z = r[2];
    return VPS_OK;
}

namespace vps {

int dataset_jump(vps_dataset* ds, int Nc, bool x96, const uint32_t** out)
{
    std::lock_guard<std::mutex> lock(ds->jump_mu);
    const int key = 2 * Nc + (x96 ? 1 : 0);
    auto it = ds->jump_cache.find(key);
    if (it != ds->jump_cache.end()) { *out = it->second; return VPS_OK; }
    const int G = (ds->d.N + Nc - 1) / Nc;
    std::vector<uint32_t> tab((size_t)G * kJumpWords);
    if (x96) jump_tables96((uint64_t)Nc * ds->d.maxploidy, G, tab.data());
    else jump_tables((uint64_t)Nc * ds->d.maxploidy, G, tab.data());
    uint32_t* dev = nullptr;
    int rc = upload(&dev, tab, ds);
    if (rc) return rc;
    ds->jump_cache[key] = dev;
    *out = dev;
    return VPS_OK;
}

}  // namespace vps

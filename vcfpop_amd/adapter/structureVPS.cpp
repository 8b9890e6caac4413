/*
 * structureVPS.cpp — vcfpop's GPU seam on the MI355X engine (libvps.so).
 *
 * A vcfpop build links this file in place of src/structureCUDA.cpp and adds
 * -lvps.  It defines:
 *   - the extern "C" runtime seam of src/structureCUDA.h:6-52 (WaitCUDA ...
 *     ResetDeviceCUDA), with the reference's own stubs for Svd, Eig and
 *     MatrixMul, which are not part of this path;
 *   - the BAYESIAN<REAL> members behind `usecuda` (src/structure.h:186-306):
 *     InitCUDA, UninitCUDA, UpdatePCUDA, UpdateQNoAdmixCUDA, UpdateQMetroCUDA,
 *     UpdateZAdmixCUDA, UpdateZNoAdmixCUDA, RecordCUDA, each on the engine's
 *     per-pass entry points with the reference's host-owned state copied in
 *     and out;
 *   - VpsMCMC (structureVPS.h), which hands a whole run to the device.
 * It touches vcfpop only through names vcfpop declares (src/vcfpop.h), and
 * the engine only through include/vps.h.
 *
 * The reference compiles its members for REAL = double and float
 * (src/structureCUDA.cpp:6-19); the engine's state is FP64, so a float run's
 * buffers are converted at the seam.
 *
 * Errors: every engine call returns a status; a failure ends in vcfpop's
 * Exit with the engine's message, where the reference's checkErrors printed
 * and exited (src/structureCUDA.cpp:183-196).
 */
#include <string.h>
#include <type_traits>
#include <vector>
#include "vcfpop.h"
#include "../../include/vps.h"
#include "structureVPS.h"

namespace {

constexpr int kMaxDevices = 64;

vps_dataset* g_dataset[kMaxDevices];     // one genotype image per device (CopyStructureMemory)
thread_local int t_device = 0;            // the device this host thread leased (CreateStreamCUDA)
thread_local vps_chain* t_chain = nullptr; // the run this host thread drives on it (InitCUDA)

void check(int rc, const char* what)
{
    if (rc) Exit("\nError: GPU engine, %s: %s\n", what, vps_last_error());
}

// REAL buffers to and from the engine's FP64 state
template <typename REAL>
void put(int which, const REAL* p, int64 n)
{
    if (std::is_same<REAL, double>::value) {
        check(vps_chain_set_state(t_chain, which, p, (size_t)n * 8), "upload");
    } else {
        std::vector<double> t(p, p + n);
        check(vps_chain_set_state(t_chain, which, t.data(), (size_t)n * 8), "upload");
    }
}

template <typename REAL>
void get(int which, REAL* p, int64 n)
{
    if (std::is_same<REAL, double>::value) {
        check(vps_chain_get_state(t_chain, which, p, (size_t)n * 8), "download");
    } else {
        std::vector<double> t(n);
        check(vps_chain_get_state(t_chain, which, t.data(), (size_t)n * 8), "download");
        for (int64 x = 0; x < n; ++x) p[x] = (REAL)t[x];
    }
}

void put_raw(int which, const void* p, size_t bytes) { check(vps_chain_set_state(t_chain, which, p, bytes), "upload"); }
void get_raw(int which, void* p, size_t bytes) { check(vps_chain_get_state(t_chain, which, p, bytes), "download"); }

// the -structure options of this process (ReadPar, src/structure.cpp:405-464,
// reads the same globals) in the engine's parameter block
// real32: the run is a BAYESIAN<float> (-g_float=single); -g_fastsingle
// then selects the single-precision genotype passes (src/parameters.cpp:
// 443-444, 795-798)
vps_params params_from_options(bool real32)
{
    vps_params p;
    vps_params_default(&p);
    p.admix_val = structure_admix_val;
    p.locpriori_val = structure_locpriori_val;
    p.f_val = structure_f_val;
    p.nburnin = structure_nburnin_val;
    p.nreps = structure_nreps_val;
    p.nthinning = structure_nthinning_val;
    p.nruns = structure_nruns_val;
    p.nadmburnin = structure_nadmburnin_val;
    p.lambda = structure_lambda_val;
    p.stdlambda = structure_stdlambda_val;
    p.maxlambda = structure_maxlambda_val;
    p.inferlambda_val = structure_inferlambda_val;
    p.difflambda_val = structure_difflambda_val;
    p.diversity_val = structure_diversity_val;
    p.alpha = structure_alpha_val;
    p.stdalpha = structure_stdalpha_val;
    p.maxalpha = structure_maxalpha_val;
    p.alphapriora = structure_alphapriora_val;
    p.alphapriorb = structure_alphapriorb_val;
    p.inferalpha_val = structure_inferalpha_val;
    p.diffalpha_val = structure_diffalpha_val;
    p.uniformalpha_val = structure_uniformalpha_val;
    p.metrofreq = structure_metrofreq_val;
    p.r = structure_r_val;
    p.maxr = structure_maxr_val;
    p.stdr = structure_stdr_val;
    p.epseta = structure_epseta_val;
    p.epsgamma = structure_epsgamma_val;
    p.pmeanf = structure_pmeanf_val;
    p.pstdf = structure_pstdf_val;
    p.stdf = structure_stdf_val;
    p.singlef_val = structure_singlef_val;
    p.writelnl_val = structure_writelnl_val;
    p.krange_min = structure_krange_min;
    p.krange_max = structure_krange_max;
    p.g_seed = (uint64_t)(int64_t)g_seed_val;
    p.rng_topology = VPS_RNG_REFERENCE;   // the reference's draw streams
    p.precision = !real32 ? VPS_PREC_FP64 : g_fastsingle_val == 1 ? VPS_PREC_FAST32 : VPS_PREC_REAL32;
    return p;
}

// the location-prior state block: r, Eta[K], Gamma[S*K], AlphaLocal[S*K], Di[S*K], SumAlpha[S]
template <typename REAL>
std::vector<double> loc_block(const BAYESIAN<REAL>& b)
{
    const int K = b.K, S = b.S;
    std::vector<double> v(1 + K + 3 * (size_t)S * K + S);
    double* o = v.data();
    *o++ = b.r;
    for (int x = 0; x < K; ++x) *o++ = b.Eta[x];
    for (int x = 0; x < S * K; ++x) *o++ = b.Gamma[x];
    for (int x = 0; x < S * K; ++x) *o++ = b.AlphaLocal[x];
    for (int x = 0; x < S * K; ++x) *o++ = (double)b.Di[x];
    for (int x = 0; x < S; ++x) *o++ = b.SumAlpha[x];
    return v;
}

template <typename REAL>
void put_loc(const BAYESIAN<REAL>& b)
{
    const std::vector<double> v = loc_block(b);
    put_raw(VPS_STATE_LOC, v.data(), v.size() * 8);
}

template <typename REAL>
void get_loc(BAYESIAN<REAL>& b)
{
    std::vector<double> v = loc_block(b);
    get_raw(VPS_STATE_LOC, v.data(), v.size() * 8);
    const int K = b.K, S = b.S;
    const double* o = v.data();
    b.r = *o++;
    for (int x = 0; x < K; ++x) b.Eta[x] = *o++;
    for (int x = 0; x < S * K; ++x) b.Gamma[x] = *o++;
    for (int x = 0; x < S * K; ++x) b.AlphaLocal[x] = *o++;
    for (int x = 0; x < S * K; ++x) b.Di[x] = (REAL)*o++;
    for (int x = 0; x < S; ++x) b.SumAlpha[x] = *o++;
}

template <typename REAL>
void put_iter(const BAYESIAN<REAL>& b)
{
    const int32_t m = b.m;
    put_raw(VPS_STATE_ITER, &m, 4);
}

}  // namespace

/* ------------------------------------------------------------------------- */
/* Runtime seam, src/structureCUDA.h:6-52                                     */
/* ------------------------------------------------------------------------- */
extern "C" void WaitCUDA()
{
    if (t_chain) check(vps_chain_sync(t_chain), "WaitCUDA");
}

extern "C" void* Malloc(uint64 size)
{
    void* re = malloc(size ? size : 1);
    if (!re) Exit("\nError: fail to allocate %llu bytes of memory.\n", (unsigned long long)size);
    return re;
}

// the engine copies host buffers itself, so pinned memory buys nothing
extern "C" void* MallocHostCUDA(uint64 size) { return Malloc(size); }
extern "C" void FreeHostCUDA(void* addr) { free(addr); }

extern "C" void* MallocDeviceCUDA(uint64 size)
{
    void* p = nullptr;
    check(vps_device_alloc(t_device, size, &p), "MallocDeviceCUDA");
    return p;
}

extern "C" void FreeCUDA(void* addr) { check(vps_device_free(addr), "FreeCUDA"); }

extern "C" void MemsetCUDA(void* addr, int val, uint64 size) { check(vps_device_memset(addr, val, size), "MemsetCUDA"); }

extern "C" void MemcpyCUDA(void* dst, void* src, uint64 size, bool todevice)
{
    check(vps_device_memcpy(dst, src, size, todevice ? 1 : 0), "MemcpyCUDA");
}

// The genotype image (src/structureCUDA.cpp:969-1094): per locus the packed
// ids (geno_bucket, CeilLog2(ngeno) bits each) and the genotype table, whose
// entries give the ploidy and the allele copies in stored order.  The engine
// sorts loci as CalcBayesian does (src/structure.cpp:2350-2395).
extern "C" void CopyStructureMemory(int devID)
{
    if (devID < 0 || devID >= kMaxDevices) Exit("\nError: device %d beyond %d.\n", devID, kMaxDevices);
    const int64 L = nloc;
    std::vector<int64_t> gtab_offset(L + 1), gt_allele_offset;
    std::vector<uint8_t> gt_ploidy, bitsize(L);
    std::vector<uint16_t> pool;
    std::vector<uint64_t> byte_off(L);
    std::vector<int32_t> nallele(L), ind_pop(nind);
    for (int64 l = 0; l < L; ++l) {
        SLOCUS& loc = GetLoc(l);
        GENOTYPE* gtab = loc.GetGtab();
        gtab_offset[l] = (int64_t)gt_ploidy.size();
        for (uint g = 0; g < loc.ngeno; ++g) {
            GENOTYPE& gt = gtab[g];
            const int v = gt.Nalleles() ? gt.Ploidy() : 0;
            gt_ploidy.push_back((uint8_t)v);
            gt_allele_offset.push_back((int64_t)pool.size());
            const ushort* als = gt.GetAlleleArray();
            for (int a = 0; a < v; ++a) pool.push_back(als[a]);
        }
        byte_off[l] = geno_bucket.offset[l].offset;
        bitsize[l] = (uint8_t)geno_bucket.offset[l].size;
        nallele[l] = (int32_t)loc.k;
    }
    gtab_offset[L] = (int64_t)gt_ploidy.size();
    // (vcfpop fills the individuals of its REAL type)
    for (int i = 0; i < nind; ++i) ind_pop[i] = ainds<double> ? ainds<double>[i]->popid : ainds<float>[i]->popid;
    vps_dataset_desc d;
    memset(&d, 0, sizeof(d));
    d.n_ind = nind;
    d.n_loc = L;
    d.n_pop = npop;
    d.ind_pop = ind_pop.data();
    d.geno_bits = geno_bucket.base_addr;
    d.geno_byte_offset = byte_off.data();
    d.bitsize = bitsize.data();
    d.n_allele = nallele.data();
    d.gtab_offset = gtab_offset.data();
    d.gt_ploidy = gt_ploidy.data();
    d.gt_allele_offset = gt_allele_offset.data();
    d.allele_pool = pool.data();
    if (g_dataset[devID]) check(vps_dataset_destroy(g_dataset[devID]), "FreeStructureMemory");
    g_dataset[devID] = nullptr;
    check(vps_dataset_create(&d, devID, &g_dataset[devID]), "CopyStructureMemory");
}

extern "C" void FreeStructureMemory(int devID)
{
    if (devID < 0 || devID >= kMaxDevices || !g_dataset[devID]) return;
    check(vps_dataset_destroy(g_dataset[devID]), "FreeStructureMemory");
    g_dataset[devID] = nullptr;
}

extern "C" int GetDeviceCountCUDA()
{
    int n = 0;
    return vps_device_count(&n) == VPS_OK ? n : 0;
}

// the reference allocates per-device bookkeeping here; the engine keeps its own
extern "C" void AllocMemoryCUDA() {}
extern "C" void FreeMemoryCUDA() {}

// a run's device: the engine gives every chain streams of its own
extern "C" void CreateStreamCUDA(int devID) { t_device = devID; }
extern "C" void DestroyStreamCUDA() {}

extern "C" void ShowDevicesCUDA()
{
    const int n = GetDeviceCountCUDA();
    printf("List of GPU devices:");
    for (int dev = 0; dev < n; ++dev) {
        char name[256];
        if (vps_device_name(dev, name, sizeof(name)) == VPS_OK) printf("\n  %d: %s", dev, name);
    }
    printf("\n");
}

extern "C" void ResetDeviceCUDA()
{
    for (int dev = 0; dev < GetDeviceCountCUDA() && dev < kMaxDevices; ++dev) {
        FreeStructureMemory(dev);
        check(vps_device_reset(dev), "ResetDeviceCUDA");
    }
}

// not part of this path: the reference's own no-GPU stubs (src/structureCUDA.cpp:2446-2456)
extern "C" void Svd64CUDA(double*, double*, double*, double*, int64, int64) {}
extern "C" void Svd32CUDA(float*, float*, float*, float*, int64, int64) {}
extern "C" void Eig64CUDA(double*, double*, double*, int64) {}
extern "C" void Eig32CUDA(float*, float*, float*, int64) {}
extern "C" void MatrixMul64CUDA(double*, double*, double*, int64, int64, int64, bool, bool) {}
extern "C" void MatrixMul32CUDA(float*, float*, float*, int64, int64, int64, bool, bool) {}

/* ------------------------------------------------------------------------- */
/* BAYESIAN<REAL> members, src/structure.h:186-306                           */
/* ------------------------------------------------------------------------- */

// The run leased device t_device (ManageCUDA, src/structure.cpp:711-756):
// a chain there, holding the host's state as of now (src/structure.cpp:
// 760-799 copies the same fields up).
template <typename REAL>
void BAYESIAN<REAL>::InitCUDA()
{
    if (t_chain) check(vps_chain_destroy(t_chain), "UninitCUDA");
    t_chain = nullptr;
    if (!g_dataset[t_device]) Exit("\nError: no genotype image on device %d (CopyStructureMemory).\n", t_device);
    const vps_params p = params_from_options(std::is_same<REAL, float>::value);
    check(vps_chain_create(g_dataset[t_device], &p, K, id, &t_chain), "InitCUDA");
    const int64 NK = (int64)N * K, KKT = (int64)K * KT;
    put(VPS_STATE_FREQ, Freq, KKT);
    put(VPS_STATE_Q, Q, NK);
    put_raw(VPS_STATE_Z, Z, (size_t)N * 2);
    put_raw(VPS_STATE_NI, Ni, (size_t)KKT * 4);
    put_raw(VPS_STATE_MI, Mi, (size_t)NK * 8);
    put_raw(VPS_STATE_ALPHA, Alpha, (size_t)K * 8);
    put_raw(VPS_STATE_LAMBDA, Lambda, (size_t)K * 8);
    if (locpriori) put_loc(*this);
    if (fmodel) {
        put(VPS_STATE_FREQA, FreqA, 3 * (int64)KT);
        std::vector<double> fst(2 * (size_t)K);
        for (int k = 0; k < K; ++k) { fst[k] = F[k]; fst[K + k] = f[k]; }
        put_raw(VPS_STATE_FST, fst.data(), fst.size() * 8);
    }
    const int32_t bq = binaryq ? 1 : 0;
    put_raw(VPS_STATE_BINARYQ, &bq, 4);
    put_iter(*this);
}

template <typename REAL>
void BAYESIAN<REAL>::UninitCUDA()
{
    if (t_chain) check(vps_chain_destroy(t_chain), "UninitCUDA");
    t_chain = nullptr;
}

// UpdateP (src/structure.cpp:812-862): P from Ni and the prior
template <typename REAL>
void BAYESIAN<REAL>::UpdatePCUDA()
{
    const int64 KKT = (int64)K * KT;
    put_iter(*this);
    put_raw(VPS_STATE_NI, Ni, (size_t)KKT * 4);
    put_raw(VPS_STATE_LAMBDA, Lambda, (size_t)K * 8);
    if (fmodel) {
        put(VPS_STATE_FREQA, FreqA, 3 * (int64)KT);
        std::vector<double> fst(2 * (size_t)K);
        for (int k = 0; k < K; ++k) { fst[k] = F[k]; fst[K + k] = f[k]; }
        put_raw(VPS_STATE_FST, fst.data(), fst.size() * 8);
    }
    check(vps_pass_update_p(t_chain), "UpdatePCUDA");
    get(VPS_STATE_FREQ, Freq, KKT);
}

// src/structureCUDA.cpp:2175-2214: per-(i,k) log likelihoods (with the
// location prior), PolyLog draws of Z, binary Q
template <typename REAL>
void BAYESIAN<REAL>::UpdateQNoAdmixCUDA()
{
    const int64 NK = (int64)N * K;
    put_iter(*this);
    put(VPS_STATE_FREQ, Freq, (int64)K * KT);
    if (locpriori) put_loc(*this);
    check(vps_pass_update_q_noadmix(t_chain), "UpdateQNoAdmixCUDA");
    get(VPS_STATE_Q, Q, NK);
    get_raw(VPS_STATE_Z, Z, (size_t)N * 2);
    get_raw(VPS_STATE_BUFNK1, bufNK1, (size_t)NK * 8);
    binaryq = true;
}

// src/structureCUDA.cpp:2269-2312: proposals, the log-ratio pass, acceptance
template <typename REAL>
void BAYESIAN<REAL>::UpdateQMetroCUDA()
{
    const int64 NK = (int64)N * K;
    put_iter(*this);
    put(VPS_STATE_FREQ, Freq, (int64)K * KT);
    put(VPS_STATE_Q, Q, NK);
    put_raw(VPS_STATE_ALPHA, Alpha, (size_t)K * 8);
    if (locpriori) put_loc(*this);
    check(vps_pass_update_q_metro(t_chain), "UpdateQMetroCUDA");
    get(VPS_STATE_Q, Q, NK);
    get_raw(VPS_STATE_BUFNK1, bufNK1, (size_t)NK * 8);
    get_raw(VPS_STATE_BUFN1, bufN1, (size_t)N * 8);
    binaryq = false;
}

// src/structureCUDA.cpp:2092-2119: the Z pass; Ni and Mi back to the host
template <typename REAL>
void BAYESIAN<REAL>::UpdateZAdmixCUDA()
{
    put_iter(*this);
    put(VPS_STATE_FREQ, Freq, (int64)K * KT);
    put(VPS_STATE_Q, Q, (int64)N * K);
    check(vps_pass_update_z_admix(t_chain), "UpdateZAdmixCUDA");
    get_raw(VPS_STATE_NI, Ni, (size_t)K * KT * 4);
    get_raw(VPS_STATE_MI, Mi, (size_t)N * K * 8);
}

// src/structureCUDA.cpp:1987-2009: counts under binary Q
template <typename REAL>
void BAYESIAN<REAL>::UpdateZNoAdmixCUDA()
{
    put_raw(VPS_STATE_Z, Z, (size_t)N * 2);
    check(vps_pass_update_z_noadmix(t_chain), "UpdateZNoAdmixCUDA");
    get_raw(VPS_STATE_NI, Ni, (size_t)K * KT * 4);
    get_raw(VPS_STATE_MI, Mi, (size_t)N * K * 8);
}

// src/structureCUDA.cpp:2381-2410: the likelihood pass; the lnL is left in
// bufNK1[0] for Record, which accumulates rout, FreqM and MiSum on the host
template <typename REAL>
void BAYESIAN<REAL>::RecordCUDA()
{
    put(VPS_STATE_FREQ, Freq, (int64)K * KT);
    const int32_t bq = binaryq ? 1 : 0;
    put_raw(VPS_STATE_BINARYQ, &bq, 4);
    if (binaryq) put_raw(VPS_STATE_Z, Z, (size_t)N * 2);
    else put(VPS_STATE_Q, Q, (int64)N * K);
    check(vps_pass_record(t_chain), "RecordCUDA");
    double lnl = 0;
    get_raw(VPS_STATE_LNL, &lnl, 8);
    bufNK1[0] = lnl;
}

// The whole run on the device (structureVPS.h): sweeps m = 0 .. nburnin +
// nreps - 1 (nburnin includes nadmburnin, ReadPar) and Arrange; the results
// land in the host fields WriteStructure reads.
template <typename REAL>
void VpsMCMC(BAYESIAN<REAL>& b)
{
    if (!t_chain) b.InitCUDA();
    check(vps_chain_step(t_chain, b.nburnin + b.nreps), "MCMC");
    check(vps_chain_finish(t_chain), "Arrange");
    const int64 NK = (int64)b.N * b.K, KKT = (int64)b.K * KT;
    get(VPS_STATE_Q, b.Q, NK);
    get(VPS_STATE_FREQ, b.Freq, KKT);
    get_raw(VPS_STATE_MISUM, b.MiSum, (size_t)NK * 8);   // normalised in place, as Arrange leaves it
    get(VPS_STATE_FREQM, b.FreqM, KKT);
    get_raw(VPS_STATE_ALPHA, b.Alpha, (size_t)b.K * 8);
    get_raw(VPS_STATE_LAMBDA, b.Lambda, (size_t)b.K * 8);
    if (b.locpriori) get_loc(b);
    int64_t rlen = 0, nl = 0;
    check(vps_chain_get_sizes(t_chain, &rlen, &nl), "MCMC");
    if (rlen != b.rlen) Exit("\nError: GPU engine: run summary of %lld values, vcfpop holds %d.\n", (long long)rlen, b.rlen);
    get_raw(VPS_STATE_ROUT, b.rout, (size_t)rlen * 8);
    if (nl && b.lnL) get_raw(VPS_STATE_LNLTRACE, b.lnL, (size_t)nl * 8);
    b.nlnL = (int)nl;
    int32_t m = 0;
    get_raw(VPS_STATE_ITER, &m, 4);
    b.m = m;
    b.nr = (b.nreps + b.nthinning - 1) / b.nthinning;
}

template void BAYESIAN<double>::InitCUDA();
template void BAYESIAN<double>::UninitCUDA();
template void BAYESIAN<double>::UpdatePCUDA();
template void BAYESIAN<double>::UpdateQNoAdmixCUDA();
template void BAYESIAN<double>::UpdateQMetroCUDA();
template void BAYESIAN<double>::UpdateZAdmixCUDA();
template void BAYESIAN<double>::UpdateZNoAdmixCUDA();
template void BAYESIAN<double>::RecordCUDA();
template void VpsMCMC<double>(BAYESIAN<double>&);
template void BAYESIAN<float>::InitCUDA();
template void BAYESIAN<float>::UninitCUDA();
template void BAYESIAN<float>::UpdatePCUDA();
template void BAYESIAN<float>::UpdateQNoAdmixCUDA();
template void BAYESIAN<float>::UpdateQMetroCUDA();
template void BAYESIAN<float>::UpdateZAdmixCUDA();
template void BAYESIAN<float>::UpdateZNoAdmixCUDA();
template void BAYESIAN<float>::RecordCUDA();
template void VpsMCMC<float>(BAYESIAN<float>&);

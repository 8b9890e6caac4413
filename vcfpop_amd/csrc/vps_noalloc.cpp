// vps_noalloc.cpp — the Record likelihood pass on caller-owned device
// buffers, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops).
//
// The pass is the genotype pass of vps_pass_record (RecordCUDA's
// counterpart): one partial per (locus range, block of 256 individuals),
// then a fixed-order sum.  Here the partials live in `scratch`, a device
// array of this translation unit, instead of a chain's allocation, so the
// entry point makes no allocation of any kind.
//
// Every call on a device shares that device's array: calls that overlap on
// one device (two streams, two threads) race on it.  Each call writes every
// partial it later reads, so nothing needs clearing between calls.
#include <mutex>
#include "vps_runtime.h"

namespace {

// 4 MiB of partials: (locus ranges) x ceil(N / 256) doubles; N = 100,000
// individuals and L = 1,000,000 biallelic loci need about 3 MiB
constexpr int64_t kScratchDoubles = (int64_t)1 << 19;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};   // per device: the device address of `scratch`

// Device address of `scratch` on `dev` (current), looked up once per device.
// The array's name is only a host-side stand-in; kernels get this address.
int scratch_on(int dev, double** out)
{
    if (dev < 0 || dev >= kMaxDevices) return vps::fail(VPS_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return vps::fail(VPS_ERR_DEVICE, "scratch address on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return VPS_OK;
}

}  // namespace

extern "C" int vps_record_lnl_noalloc(const vps_dataset* ds, int K, const double* freq, const double* q,
                                      const uint16_t* z, double* lnl, void* stream)
{
    using namespace vps;
    if (!ds || !freq || !lnl || (!q && !z)) return fail(VPS_ERR_INVALID, "null argument");
    if (ds->host_only) return fail(VPS_ERR_STATE, "dataset has no device image");
    if (K < 1 || K > VPS_MAX_K) return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", K, VPS_MAX_K);
    VPS_HIP(hipSetDevice(ds->device));
    const DeviceDataset& d = ds->d;
    const int mode = z ? 1 : 0;
    ChainDev c{};
    int rc = choose_lik_tpr(d, K, ds->device, c.lik_tpr);
    if (rc) return rc;
    const int64_t need = lik_part_count(d, mode, K, c.lik_tpr);
    if (need > kScratchDoubles)
        return fail(VPS_ERR_UNSUPPORTED, "the pass needs %lld partials, the scratch holds %lld (N=%d, L=%lld)",
                    (long long)need, (long long)kScratchDoubles, d.N, (long long)d.L);
    double* sp = nullptr;
    rc = scratch_on(ds->device, &sp);
    if (rc) return rc;
    c.K = K; c.N = d.N; c.S = d.S; c.L = d.L; c.KT = d.KT; c.maxploidy = d.maxploidy;
    c.precision = VPS_PREC_FP64;
    c.Freq = const_cast<double*>(freq);
    c.Q = const_cast<double*>(q);
    c.Z = const_cast<uint16_t*>(z);
    c.part = sp;
    return launch_lik_pass(d, c, mode, static_cast<hipStream_t>(stream), lnl);
}

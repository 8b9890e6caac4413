// vps_lik.hip — the genotype passes that reduce log likelihoods:
//   mode 0  Record, admixture        lnL = sum log(sum_k Q[i][k] P[k][l][x])
//   mode 1  Record, binary Q          lnL = sum log P[Z_i][l][x]
//           (RecordAVX, src/structureAVX.cpp:986-1187)
//   mode 2  Metropolis log ratio     r_i = sum log(sum_k q'_ik p / sum_k q_ik p)
//           (UpdateQMetroAVX, src/structureAVX.cpp:557-783)
//   mode 3  no-admix log likelihood  b_ik = sum log P[k][l][x]
//           (UpdateQNoAdmixAVX, src/structureAVX.cpp:254-428)
// sums run over typed allele copies x of individual i at locus l.
//
// These passes draw no random numbers, so lanes map to individuals: a lane
// keeps its individual's Q row in registers and its own exponent-tracking
// log accumulator (src/math2.h:716-902) across a range of loci, with no
// cross-lane reduction inside the loop.  Each (locus range, individual)
// writes one closed partial; a fixed-order reduction adds the partials.
// The reference multiplies 64-lane products into one accumulator instead;
// the two orders agree to rounding (a few ulp per partial).
#include "vps_chain.h"

namespace vps {

constexpr int kLikThreads = 256;
constexpr int kLikTilesPerRange = 8;

template <int K, int MODE>
__global__ __launch_bounds__(kLikThreads) void k_lik(const DeviceDataset ds, const ChainDev c, int tiles_per_range)
{
    const int i = blockIdx.y * kLikThreads + threadIdx.x;
    const bool ok = i < ds.N;
    const int ii = ok ? i : 0;
    const int64_t KT = ds.KT;
    double q[K], qp[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        q[k] = c.Q[(int64_t)ii * K + k];
        qp[k] = MODE == 2 ? c.bufNK1[(int64_t)ii * K + k] : 0.0;
    }
    const int zi = MODE == 1 ? c.Z[ii] : 0;
    constexpr int NA = MODE == 3 ? K : 1;
    LogAcc acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k].open();

    const int64_t t0 = (int64_t)blockIdx.x * tiles_per_range;
    const int64_t t1 = min(ds.ntiles, t0 + tiles_per_range);
    for (int64_t t = t0; t < t1; ++t) {
        const TileDesc td = ds.tiles[t];
        const int w = ii / td.ipw;
        const int sh = (ii - w * td.ipw) * td.bits;
        const uint32_t mask = (1u << td.bits) - 1u;
        const uint32_t* wp = ds.words + td.wbase + (int64_t)w * kTileLoci;
        for (int j = 0; j < td.nloc; ++j) {
            const int64_t l = td.l0 + j;
            const uint32_t gid = (wp[j] >> sh) & mask;
            const double* pl = c.Freq + ds.afo[l];
            int v;
            uint64_t d64 = 0;
            int h = 0;
            if (td.kind == kKindBi) {
                const uint32_t desc = (uint32_t)(ds.bitab[l] >> (gid * 8)) & 0xFFu;
                v = (int)(desc & 15u);
                h = v - (int)(desc >> 4);
            } else {
                d64 = ds.gdesc[ds.gbase[l] + gid];
                v = (int)(d64 & 15u);
            }
            if (!ok) v = 0;
            for (int a = 0; a < v; ++a) {
                const int x = td.kind == kKindBi ? (a >= h ? 1 : 0) : (int)((d64 >> (4 + 6 * a)) & 63u);
                const double* px = pl + x;
                if (MODE == 0) {
                    double f = 0;
#pragma unroll
                    for (int k = 0; k < K; ++k) f = f + px[k * KT] * q[k];
                    acc[0].charge(f);
                } else if (MODE == 1) {
                    acc[0].charge(px[(int64_t)zi * KT]);
                } else if (MODE == 2) {
                    double f1 = 0, f2 = 0;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const double p = px[k * KT];
                        f1 = f1 + p * qp[k];
                        f2 = f2 + p * q[k];
                    }
                    acc[0].charge(f1 / f2);
                } else {
#pragma unroll
                    for (int k = 0; k < K; ++k) acc[k < NA ? k : 0].charge(px[k * KT]);
                }
            }
        }
    }
    if (!ok) return;
    double* out = c.part + (int64_t)blockIdx.x * ds.N * NA + (int64_t)i * NA;
#pragma unroll
    for (int k = 0; k < NA; ++k) out[k] = acc[k].close();
}

// lnL = sum of all partials, one block, fixed order
__global__ __launch_bounds__(256) void k_reduce_lnl(const double* __restrict__ part, int64_t n, double* __restrict__ out)
{
    __shared__ double sh[256];
    double s = 0;
    for (int64_t x = threadIdx.x; x < n; x += 256) s += part[x];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// per-individual (x NA) sums over locus ranges; mode 3 adds the location
// prior log Gamma[pop][k] for typed individuals (charged first in the
// reference, src/structureAVX.cpp:260-266)
__global__ void k_reduce_rows(const DeviceDataset ds, const ChainDev c, int nranges, int NA, int add_prior, double* __restrict__ dst)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)ds.N * NA;
    if (x >= n) return;
    double s = 0;
    if (add_prior) {
        const int i = (int)(x / NA), k = (int)(x % NA);
        if (ds.vt[i]) {
            LogAcc a;
            a.open();
            a.charge(c.Gamma[(int64_t)ds.popid[i] * c.K + k]);
            s = a.close();
        }
    }
    for (int r = 0; r < nranges; ++r) s += c.part[(int64_t)r * n + x];
    dst[x] = s;
}

namespace {

template <int MODE>
int launch_lik_mode(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, dim3 grid)
{
    switch (c.K) {
#define VPS_CASE(k) case k: hipLaunchKernelGGL((k_lik<k, MODE>), grid, dim3(kLikThreads), 0, st, ds, c, kLikTilesPerRange); break;
        VPS_CASE(1) VPS_CASE(2) VPS_CASE(3) VPS_CASE(4) VPS_CASE(5) VPS_CASE(6) VPS_CASE(7) VPS_CASE(8)
        VPS_CASE(9) VPS_CASE(10) VPS_CASE(11) VPS_CASE(12) VPS_CASE(13) VPS_CASE(14) VPS_CASE(15) VPS_CASE(16)
#undef VPS_CASE
    default: return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", c.K, VPS_MAX_K);
    }
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

}  // namespace

int lik_part_count(const DeviceDataset& ds, int mode, int K)
{
    const int64_t nranges = (ds.ntiles + kLikTilesPerRange - 1) / kLikTilesPerRange;
    return (int)(nranges * ds.N * (mode == 3 ? K : 1));
}

int launch_lik_pass(const DeviceDataset& ds, const ChainDev& c, int mode, hipStream_t st)
{
    const int nranges = (int)((ds.ntiles + kLikTilesPerRange - 1) / kLikTilesPerRange);
    dim3 grid((unsigned)nranges, (unsigned)((ds.N + kLikThreads - 1) / kLikThreads));
    int rc;
    switch (mode) {
    case 0: rc = launch_lik_mode<0>(ds, c, st, grid); break;
    case 1: rc = launch_lik_mode<1>(ds, c, st, grid); break;
    case 2: rc = launch_lik_mode<2>(ds, c, st, grid); break;
    case 3: rc = launch_lik_mode<3>(ds, c, st, grid); break;
    default: return fail(VPS_ERR_INVALID, "likelihood mode %d", mode);
    }
    if (rc) return rc;
    if (mode <= 1) {
        hipLaunchKernelGGL(k_reduce_lnl, dim3(1), dim3(256), 0, st, c.part, (int64_t)nranges * ds.N, c.scal + 1);
    } else {
        const int NA = mode == 3 ? c.K : 1;
        const int64_t n = (int64_t)ds.N * NA;
        double* dst = mode == 2 ? c.bufN1 : c.bufNK1;
        hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ds, c, nranges, NA,
                           (mode == 3 && c.locpriori) ? 1 : 0, dst);
    }
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

}  // namespace vps

// vps_updates.hip — the non-genotype steps of a sweep on the device:
// Dirichlet draws for P (UpdateP) and Q (UpdateQAdmix, Metropolis
// proposals/acceptance, PolyLog for binary Q), the scalar Metropolis updates
// (alpha, lambda, location prior, F model), Record's accumulators and the
// initial cluster assignment (InitAdmix).
//
// Two draw topologies (vps_params::rng_topology):
//   VPS_RNG_REFERENCE  the reference's streams: P in 128 sequential blocks
//                      RNG(seed + m*128 + b, UPDATEP) over kl = k*L + l
//                      (src/structure.cpp:843-861); Q in one stream
//                      RNG(seed + m, UPDATEQ) over individuals
//                      (src/structure.cpp:990-1003); A in 32 blocks.
//                      The P and Q streams are parsed in parallel
//                      (vps_spp.hip); A runs 32 device lanes.
//   VPS_RNG_PARALLEL   one stream per (k, l): RNG(seed + (m*K + k)*L + l,
//                      UPDATEP); per individual RNG(seed + m*N + i, UPDATEQ);
//                      per locus RNG(seed + m*L + l, UPDATEA).  Same
//                      distributions with full device parallelism.
// Every other stream (Z, alpha, lambda, location prior, F) is the
// reference's in both modes.
#include <algorithm>
#include "vps_chain.h"

namespace vps {

// ---------------------------------------------------------------------------
// UpdateP, src/structure.cpp:812-862
// ---------------------------------------------------------------------------
__device__ __forceinline__ void draw_p_one(const DeviceDataset& ds, const ChainDev& c, Rng& r, int64_t k, int64_t l)
{
    const int64_t KT = ds.KT;
    const int afo = ds.afo[l];
    const int k2 = ds.k2[l];
    double* p = c.Freq + k * KT + afo;
    const int32_t* ni = c.Ni + k * KT + afo;
    const double fk = c.fmodel ? c.Fst[c.K + (c.fsame ? 0 : k)] : 0.0;
    // the prior is first stored in P (SetVal(ClusterFreq(k), (REAL)Lambda[k]))
    const double lam = real_store(c.Lambda[k], real32(c));
    double s = 0;
    for (int a = 0; a < k2; ++a) {
        const double prior = c.fmodel ? c.FreqA[afo + a] * fk : lam;
        const double v = r.gamma(prior + (double)ni[a]);
        p[a] = v;
        s += v;
    }
    dirichlet_scale(p, s, k2, real32(c));
}

__global__ void k_update_p_par(const DeviceDataset ds, const ChainDev c)
{
    const int m = *c.mdev;
    const int64_t kl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (kl >= (int64_t)c.K * ds.L) return;
    const int64_t k = kl / ds.L, l = kl % ds.L;
    Rng r;
    r.init(c.seed + ((uint64_t)m * c.K + k) * (uint64_t)ds.L + l, kSaltUpdateP);
    draw_p_one(ds, c, r, k, l);
}

int launch_update_p(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (c.topology == VPS_RNG_REFERENCE) {
        return launch_spp(ds, c, true, kSppQNone, false, st);
    } else {
        const int64_t n = (int64_t)c.K * ds.L;
        hipLaunchKernelGGL(k_update_p_par, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ds, c);
    }
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// UpdateQAdmix, src/structure.cpp:986-1004
// ---------------------------------------------------------------------------
__device__ __forceinline__ const double* alpha_row(const DeviceDataset& ds, const ChainDev& c, int i)
{
    return c.locpriori ? c.AlphaLocal + (int64_t)ds.popid[i] * c.K : c.Alpha;
}

__global__ void k_q_admix(const DeviceDataset ds, const ChainDev c)
{
    const int m = *c.mdev;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ds.N || !ds.vt[i]) return;
    Rng r;
    r.init(c.seed + (uint64_t)m * ds.N + i, kSaltUpdateQ);
    dirichlet_ab(r, c.Q + (int64_t)i * c.K, alpha_row(ds, c, i), c.Mi + (int64_t)i * c.K, c.K, real32(c));
}

int launch_update_q_admix(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (c.topology == VPS_RNG_REFERENCE) return launch_spp(ds, c, false, kSppQAdmix, false, st);
    const unsigned nb = (unsigned)((ds.N + 255) / 256);
    hipLaunchKernelGGL(k_q_admix, dim3(nb), dim3(256), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// UpdateP and the Dirichlet UpdateQ of the parallel topology in one launch
// (the two read disjoint state): blocks [0, nbq) draw Q, one individual per
// thread, and then clear the individual's Mi row for the Z pass; the rest
// draw P, one (k, l) per thread.  Thread 0 also copies the sweep index for
// the kernels that run beside the Z pass.
__global__ void k_update_pq(const DeviceDataset ds, const ChainDev c, unsigned nbq)
{
    const int m = *c.mdev;
    if (blockIdx.x == 0 && threadIdx.x == 0) c.mside[0] = m;
    if (blockIdx.x < nbq) {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= ds.N) return;
        int32_t* mi = c.Mi + (int64_t)i * c.K;
        if (ds.vt[i]) {
            Rng r;
            r.init(c.seed + (uint64_t)m * ds.N + i, kSaltUpdateQ);
            double* q = c.Q + (int64_t)i * c.K;
            dirichlet_ab(r, q, alpha_row(ds, c, i), mi, c.K, real32(c));
            // the single-precision copy the FP32-weight Z passes read
            if (c.precision == VPS_PREC_MIXED || c.precision == VPS_PREC_FAST32)
                for (int k = 0; k < c.K; ++k) c.Qf[(int64_t)i * c.K + k] = (float)q[k];
        }
        for (int k = 0; k < c.K; ++k) mi[k] = 0;
        return;
    }
    const int64_t kl = (int64_t)(blockIdx.x - nbq) * blockDim.x + threadIdx.x;
    if (kl >= (int64_t)c.K * ds.L) return;
    const int64_t k = kl / ds.L, l = kl % ds.L;
    Rng r;
    r.init(c.seed + ((uint64_t)m * c.K + k) * (uint64_t)ds.L + l, kSaltUpdateP);
    draw_p_one(ds, c, r, k, l);
}

int launch_update_pq(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (c.topology != VPS_RNG_PARALLEL) return fail(VPS_ERR_INVALID, "the joint P/Q launch draws the parallel streams");
    const unsigned nbq = (unsigned)((ds.N + 255) / 256);
    const unsigned nbp = (unsigned)(((int64_t)c.K * ds.L + 255) / 256);
    hipLaunchKernelGGL(k_update_pq, dim3(nbq + nbp), dim3(256), 0, st, ds, c, nbq);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// UpdateQMetro, src/structure.cpp:1090-1136: proposals q'_i ~ Dir(alpha),
// the log-ratio pass, then acceptance with conditional uniform draws from
// the same stream(s).
// ---------------------------------------------------------------------------
__global__ void k_q_metro_propose(const DeviceDataset ds, const ChainDev c)
{
    const int m = *c.mdev;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ds.N) return;
    Rng r;
    r.init(c.seed + (uint64_t)m * ds.N + i, kSaltUpdateQ);
    if (ds.vt[i]) dirichlet_a(r, c.bufNK1 + (int64_t)i * c.K, alpha_row(ds, c, i), c.K, real32(c));
    c.rngbuf[i] = r;
}

// Acceptance (src/structure.cpp:1039): a proposal is kept when its log
// ratio is >= -1e-10 or, only then drawn, a uniform falls below exp(ratio).
// bufN2 arrives as each individual's log likelihood under its current Q and
// leaves as the one under the Q it keeps (current + ratio when accepted):
// a Record in the same sweep sums it instead of another genotype pass.
__global__ void k_q_metro_accept(const DeviceDataset ds, const ChainDev c)
{
    const int K = c.K;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ds.N || !ds.vt[i]) return;
    Rng r = c.rngbuf[i];
    if (c.bufN1[i] >= kNZero || r.uniform() < exp(c.bufN1[i])) {
        for (int k = 0; k < K; ++k) c.Q[(int64_t)i * K + k] = c.bufNK1[(int64_t)i * K + k];
        c.bufN2[i] += c.bufN1[i];   // the log likelihood under the Q it keeps
    }
}

int launch_q_metro_propose(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (c.topology == VPS_RNG_REFERENCE)
        return pg_applies(c) ? launch_pg_proposals(ds, c, st) : launch_spp(ds, c, false, kSppQProposals, false, st);
    const unsigned nb = (unsigned)((ds.N + 255) / 256);
    hipLaunchKernelGGL(k_q_metro_propose, dim3(nb), dim3(256), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

int launch_q_metro_accept(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (c.topology == VPS_RNG_REFERENCE) return launch_spp_metro_accept(ds, c, st);
    const unsigned nb = (unsigned)((ds.N + 255) / 256);
    hipLaunchKernelGGL(k_q_metro_accept, dim3(nb), dim3(256), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// UpdateQNoAdmix's draw: Q = 0, then per typed individual PolyLog over the
// K log likelihoods picks Z_i and sets Q[i][Z_i] = 1 (src/structureAVX.cpp:
// 279-287).  bufNK1 rows are overwritten by PolyLog, as in the reference.
// ---------------------------------------------------------------------------
__global__ void k_q_noadmix_draw(const DeviceDataset ds, const ChainDev c)
{
    const int m = *c.mdev;
    const int K = c.K;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ds.N) return;
    for (int k = 0; k < K; ++k) c.Q[(int64_t)i * K + k] = 0;
    if (!ds.vt[i]) return;
    Rng r;
    r.init(c.seed + (uint64_t)m * ds.N + i, kSaltUpdateQ);
    const int k2 = polylog(r, c.bufNK1 + (int64_t)i * K, K);
    c.Q[(int64_t)i * K + k2] = 1;
    c.Z[i] = (uint16_t)k2;
}

int launch_q_noadmix_draw(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (c.topology == VPS_RNG_REFERENCE) return launch_spp_polylog(ds, c, st);
    const unsigned nb = (unsigned)((ds.N + 255) / 256);
    hipLaunchKernelGGL(k_q_noadmix_draw, dim3(nb), dim3(256), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// Di[s][k] = sum of Q rows of typed individuals of location s
// (UpdateQ, src/structure.cpp:1148-1155); Mi[i][Z_i] = vt_i for binary Q
// (UpdateZNoAdmix, src/structure.cpp:1268-1269)
__global__ void k_update_di(const DeviceDataset ds, const ChainDev c)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= c.S * c.K) return;
    const int s = x / c.K, k = x % c.K;
    double acc = 0;
    for (int i = ds.pop_ind0[s]; i < ds.pop_ind0[s] + ds.pop_nind[s]; ++i)
        if (ds.vt[i]) acc += c.Q[(int64_t)i * c.K + k];
    c.Di[x] = acc;
}

int launch_update_di(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    hipLaunchKernelGGL(k_update_di, dim3((unsigned)((c.S * c.K + 63) / 64)), dim3(64), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

constexpr int kScalarThreads = 256;

// ---------------------------------------------------------------------------
// Grid-wide LogProd for UpdateAlpha and UpdateLambda.  Their log ratios need
// LogProd over columns of Q or P (up to N*K or K*KT values), and the
// Metropolis chain itself is a few scalars: the grid is ncol x B blocks,
// block (col, b) reduces slice b of column col to an exponent and a mantissa,
// and the last block to finish (a counter the last block resets) combines
// each column's B partials in block order, then runs the chain on thread 0.
// ---------------------------------------------------------------------------
struct LpCol {
    const double* a;
    int64_t n, sep;   // values a[x * sep], x < n
};

// one block's share of column `col`, written as (exponent bits, mantissa)
__device__ void lp_block(const LpCol col, int b, int B, int64_t* sh_s, double* sh_p, double* out)
{
    const int64_t x1 = col.n * (b + 1) / B, st = blockDim.x;
    LogAcc acc;
    acc.open();
    int64_t x = col.n * b / B + threadIdx.x;
    for (; x + 3 * st < x1; x += 4 * st) {   // four loads in flight per thread
        const double v0 = col.a[x * col.sep], v1 = col.a[(x + st) * col.sep];
        const double v2 = col.a[(x + 2 * st) * col.sep], v3 = col.a[(x + 3 * st) * col.sep];
        acc.charge(v0);
        acc.charge(v1);
        acc.charge(v2);
        acc.charge(v3);
    }
    for (; x < x1; x += st) acc.charge(col.a[x * col.sep]);
    LogAcc::strip(acc.slog, acc.prod);
    sh_s[threadIdx.x] = acc.slog;
    sh_p[threadIdx.x] = acc.prod;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            int64_t s = sh_s[threadIdx.x] + sh_s[threadIdx.x + w];
            double p = sh_p[threadIdx.x] * sh_p[threadIdx.x + w];
            LogAcc::strip(s, p);
            sh_s[threadIdx.x] = s;
            sh_p[threadIdx.x] = p;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = as_f64((uint64_t)sh_s[0]);
        out[1] = sh_p[0];
    }
}

// true in every thread of the last block of the grid to arrive
__device__ bool lp_arrive(uint32_t* cnt, uint32_t nblocks)
{
    __shared__ uint32_t ticket;
    if (threadIdx.x == 0) {
        __threadfence();
        ticket = atomicAdd(cnt, 1u);
    }
    __syncthreads();
    if (ticket != nblocks - 1) return false;
    __threadfence();
    if (threadIdx.x == 0) *cnt = 0;
    return true;
}

// the last block: column col's B partials combined in block order; the
// result replaces the mantissa of the column's first partial
__device__ void lp_combine_all(double* part, int ncol, int B)
{
    for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
        double* q = part + (int64_t)2 * col * B;
        int64_t s = 0;
        double p = 1;
        for (int b = 0; b < B; ++b) {
            s += (int64_t)as_u64(q[2 * b]);
            p *= q[2 * b + 1];
            LogAcc::strip(s, p);
        }
        q[1] = ((double)s + log2(p)) * kLn2Trunc;
    }
    __syncthreads();
}

constexpr int kLpThreads = 256;
constexpr int kLpItems = 8;       // values per thread a column slice aims at

// blocks per column: about kLpItems values per thread, within the partials' capacity
int lp_blocks(int ncol, int64_t n)
{
    const int64_t want = (n + (int64_t)kLpThreads * kLpItems - 1) / ((int64_t)kLpThreads * kLpItems);
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, std::max(1, kLpCap / ncol)));
}

// ---------------------------------------------------------------------------
// UpdateAlpha, src/structure.cpp:1501-1594
// ---------------------------------------------------------------------------
enum { kAlphaOne = 0, kAlphaDiff = 1, kAlphaLoc = 2 };

__device__ LpCol alpha_col(const DeviceDataset& ds, const ChainDev& c, int mode, int col)
{
    const int K = c.K;
    if (mode == kAlphaLoc) {
        const int s = col / K, k = col % K;
        return {c.Q + (int64_t)ds.pop_ind0[s] * K + k, ds.pop_nind[s], K};
    }
    if (mode == kAlphaDiff) return {c.Q + col, ds.N, K};
    return {c.Q, (int64_t)ds.N * K, 1};
}

__global__ __launch_bounds__(kLpThreads) void k_update_alpha(const DeviceDataset ds, const ChainDev c, int mode, int B)
{
    const int m = *c.mdev;
    __shared__ int64_t sh_s[kLpThreads];
    __shared__ double sh_p[kLpThreads];
    const int K = c.K, N = ds.N, S = c.S;
    if (mode != kAlphaLoc && !c.admix && m >= c.nadmburnin) return;
    const int col = blockIdx.x / B, ncol = gridDim.x / B;
    lp_block(alpha_col(ds, c, mode, col), blockIdx.x % B, B, sh_s, sh_p, c.lpart + 2 * (int64_t)blockIdx.x);
    if (!lp_arrive(c.lcnt, gridDim.x)) return;
    lp_combine_all(c.lpart, ncol, B);
    if (threadIdx.x != 0) return;
    const double* lp = c.lpart + 1;   // LogProd of column j at lp[2 * j * B]
    const int64_t ls = 2 * (int64_t)B;
    Rng r;
    r.init(c.seed + (uint64_t)m, kSaltUpdateAlpha);
    if (mode == kAlphaLoc) {
        const double rr = c.scal[0];
        for (int k = 0; k < K; ++k) {
            const double am = r.normal(c.Alpha[k], c.stdalpha);
            if (am > 0 && am < c.maxalpha) {
                const double d = am - c.Alpha[k];
                double dlnL = S * (d * rr * my_log(rr) + lgamma(rr * c.Alpha[k]) - lgamma(rr * am));
                dlnL += rr * d * logprod_sep(c.AlphaLocal + k, S, K);
                if (dlnL >= kNZero || r.uniform() < exp(dlnL)) c.Alpha[k] = am;
            }
        }
        if (m % 10 == 0)
            for (int s = 0; s < S; ++s) {
                double sm = 0;
                for (int k = 0; k < K; ++k) sm += c.AlphaLocal[s * K + k];
                c.SumAlpha[s] = sm;
            }
        for (int s = 0; s < S; ++s) {
            for (int k = 0; k < K; ++k) {
                const double ao = c.AlphaLocal[s * K + k], am = r.normal(ao, c.stdalpha);
                if (am > 0 && am < c.maxalpha) {
                    const double d = am - ao;
                    double dlnL = (rr * c.Alpha[k] - 1) * my_log(am / ao) - d * rr +
                                  ds.pop_nind[s] * (lgamma(c.SumAlpha[s] + d) - lgamma(am) - lgamma(c.SumAlpha[s]) + lgamma(ao));
                    dlnL += d * lp[(s * K + k) * ls];
                    if (dlnL >= kNZero || r.uniform() < exp(dlnL)) {
                        c.AlphaLocal[s * K + k] = am;
                        c.SumAlpha[s] += d;
                    }
                }
            }
        }
        return;
    }
    if (mode == kAlphaDiff) {
        double sumalpha = 0;
        for (int k = 0; k < K; ++k) sumalpha += c.Alpha[k];
        for (int k = 0; k < K; ++k) {
            const double am = r.normal(c.Alpha[k], c.stdalpha);
            if (am <= 0 || (am >= c.maxalpha && c.uniformalpha)) continue;
            const double ao = c.Alpha[k];
            const double d = am - ao;
            double dlnL = c.uniformalpha ? 0 : ((c.alphapriora - 1) * my_log(am / ao) + (ao - am) / c.alphapriorb);
            dlnL += (am - ao) * lp[k * ls] + (lgamma(sumalpha + d) - lgamma(sumalpha) - lgamma(am) + lgamma(ao)) * N;
            if (dlnL >= kNZero || r.uniform() < exp(dlnL)) c.Alpha[k] = am;
        }
        return;
    }
    const double am = r.normal(c.Alpha[0], c.stdalpha);
    if (am <= 0 || (am >= c.maxalpha && c.uniformalpha)) return;
    const double ao = c.Alpha[0];
    double dlnL = c.uniformalpha ? 0 : ((c.alphapriora - 1) * my_log(am / ao) + (ao - am) / c.alphapriorb);
    dlnL += (am - ao) * lp[0] + (lgamma(K * am) - lgamma(K * ao) - K * lgamma(am) + K * lgamma(ao)) * N;
    if (dlnL >= kNZero || r.uniform() < exp(dlnL))
        for (int k = 0; k < K; ++k) c.Alpha[k] = am;
}

int launch_update_alpha(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    int mode, ncol;
    int64_t n;
    if (c.locpriori && c.admix) {
        mode = kAlphaLoc;
        ncol = c.S * c.K;
        n = (ds.N + c.S - 1) / c.S;
    } else if (!c.inferalpha) {
        return VPS_OK;
    } else if (c.diffalpha) {
        mode = kAlphaDiff;
        ncol = c.K;
        n = ds.N;
    } else {
        mode = kAlphaOne;
        ncol = 1;
        n = (int64_t)ds.N * c.K;
    }
    const int B = lp_blocks(ncol, n);
    hipLaunchKernelGGL(k_update_alpha, dim3((unsigned)(ncol * B)), dim3(kLpThreads), 0, st, ds, c, mode, B);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// UpdateLambda, src/structure.cpp:1598-1635 (runs only when inferlambda):
// LogProd of each cluster's frequencies (difflambda), else of all of them
// (of the ancestral frequencies under the F model)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kLpThreads) void k_update_lambda(const DeviceDataset ds, const ChainDev c, int B)
{
    const int m = *c.mdev;
    __shared__ int64_t sh_s[kLpThreads];
    __shared__ double sh_p[kLpThreads];
    const int K = c.K;
    const int64_t KT = ds.KT;
    const int col = blockIdx.x / B, ncol = gridDim.x / B;
    const LpCol lc = c.difflambda ? LpCol{c.Freq + col * KT, KT, 1}
                                  : c.fmodel ? LpCol{c.FreqA, KT, 1} : LpCol{c.Freq, K * KT, 1};
    lp_block(lc, blockIdx.x % B, B, sh_s, sh_p, c.lpart + 2 * (int64_t)blockIdx.x);
    if (!lp_arrive(c.lcnt, gridDim.x)) return;
    lp_combine_all(c.lpart, ncol, B);
    if (threadIdx.x != 0) return;
    const double* lp = c.lpart + 1;
    const int64_t ls = 2 * (int64_t)B;
    Rng r;
    r.init(c.seed + (uint64_t)m, kSaltUpdateLambda);
    if (c.difflambda) {
        for (int k = 0; k < K; ++k) {
            const double lo = c.Lambda[k], lm = r.normal(lo, c.stdlambda);
            if (lm <= 0 || lm >= c.maxlambda) continue;
            double dlnL = 0, gd = lgamma(lo) - lgamma(lm);
            for (int i = 2; i <= ds.maxK; ++i) dlnL += ds.kdis[i] * (lgamma(i * lm) - lgamma(i * lo) + i * gd);
            dlnL += (lm - lo) * lp[k * ls];
            if (dlnL >= kNZero || r.uniform() < exp(dlnL)) c.Lambda[k] = lm;
        }
    } else {
        const int np = c.fmodel ? 1 : K;
        const double lo = c.Lambda[0], lm = r.normal(lo, c.stdlambda);
        if (lm <= 0 || lm >= c.maxlambda) return;
        double dlnL = 0, gd = lgamma(lo) - lgamma(lm);
        for (int i = 2; i <= ds.maxK; ++i) dlnL += ds.kdis[i] * np * (lgamma(i * lm) - lgamma(i * lo) + i * gd);
        dlnL += (lm - lo) * lp[0];
        if (dlnL >= kNZero || r.uniform() < exp(dlnL))
            for (int k = 0; k < K; ++k) c.Lambda[k] = lm;
    }
}

int launch_update_lambda(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (!c.inferlambda) return VPS_OK;
    const int ncol = c.difflambda ? c.K : 1;
    const int64_t n = c.difflambda || c.fmodel ? ds.KT : c.K * ds.KT;
    const int B = lp_blocks(ncol, n);
    hipLaunchKernelGGL(k_update_lambda, dim3((unsigned)(ncol * B)), dim3(kLpThreads), 0, st, ds, c, B);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// UpdateLocPriori, src/structure.cpp:1170-1255 (S x K scalars, thread 0)
// ---------------------------------------------------------------------------
__global__ void k_update_locpriori(const DeviceDataset ds, const ChainDev c)
{
    const int m = *c.mdev;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int K = c.K, S = c.S;
    Rng rng;
    rng.init(c.seed + (uint64_t)m, kSaltUpdateLoc);
    double r = c.scal[0];
    if (c.admix) {
        const double rm = rng.uniform(r - c.epsr, r + c.epsr);
        if (rm > 0 && rm < c.maxr) {
            double dlnL = 0, d = rm - r, dt1 = rm * my_log(rm) - r * my_log(r);
            for (int k = 0; k < K; ++k) {
                dlnL += S * (c.Alpha[k] * dt1 - lgamma(rm * c.Alpha[k]) + lgamma(r * c.Alpha[k]));
                double sa = 0;
                for (int s = 0; s < S; ++s) sa += c.AlphaLocal[s * K + k];
                dlnL += c.Alpha[k] * d * logprod_sep(c.AlphaLocal + k, S, K) - sa * d;
            }
            if (dlnL >= kNZero || rng.uniform() < exp(dlnL)) c.scal[0] = rm;
        }
        return;
    }
    {
        const double rm = rng.uniform(r - c.epsr, r + c.epsr);
        if (rm > 0 && rm < c.maxr) {
            double dlnL = lgamma(rm) - lgamma(r);
            for (int k = 0; k < K; ++k) dlnL += lgamma(r * c.Eta[k]) - lgamma(rm * c.Eta[k]);
            dlnL *= S;
            for (int k = 0; k < K; ++k) dlnL += (rm - r) * c.Eta[k] * logprod_sep(c.Gamma + k, S, K);
            if (dlnL >= kNZero || rng.uniform() < exp(dlnL)) c.scal[0] = rm;
        }
    }
    r = c.scal[0];
    if (K >= 2) {
        const int64_t i1 = (int64_t)rng.next(K), j1 = (int64_t)rng.next_avoid(K, i1);
        const double delta = rng.uniform(0, c.epseta);
        const double e1 = c.Eta[i1] + delta, e2 = c.Eta[j1] - delta;
        if (e1 < 1 && e2 > 0) {
            double dlnL = S * (lgamma(r * c.Eta[i1]) + lgamma(r * c.Eta[j1]) - lgamma(r * e1) - lgamma(r * e2));
            LogAcc a1, a2;
            a1.open(); a2.open();
            for (int s = 0; s < S; ++s) { a1.charge(c.Gamma[s * K + i1]); a2.charge(c.Gamma[s * K + j1]); }
            dlnL += r * delta * (a1.close() - a2.close());
            if (dlnL >= kNZero || rng.uniform() < exp(dlnL)) {
                c.Eta[i1] += delta;
                c.Eta[j1] -= delta;
            }
        }
        for (int s = 0; s < S; ++s) {
            const int64_t a = (int64_t)rng.next(K), b = (int64_t)rng.next_avoid(K, a);
            const double dl = rng.uniform(0, c.epsgamma);
            double* G = c.Gamma + (int64_t)s * K;
            if (G[a] + dl < 1 && G[b] - dl > 0) {
                const double dlnL = (r * c.Eta[a] - 1.0 + c.Di[s * K + a]) * my_log(1 + dl / G[a]) +
                                    (r * c.Eta[b] - 1.0 + c.Di[s * K + b]) * my_log(1 - dl / G[b]);
                if (dlnL >= kNZero || rng.uniform() < exp(dlnL)) {
                    G[a] += dl;
                    G[b] -= dl;
                }
            }
        }
    }
    if (m % 10 == 0) {
        unify_row(c.Eta, K);
        for (int s = 0; s < S; ++s) unify_row(c.Gamma + (int64_t)s * K, K);
    }
}

int launch_update_locpriori(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (!c.locpriori) return VPS_OK;
    hipLaunchKernelGGL(k_update_locpriori, dim3(1), dim3(64), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// UpdateA (A1: Dirichlet proposal / A2: pairwise shift), src/structure.cpp:
// 1639-1804, and UpdateF, src/structure.cpp:1808-1858
// ---------------------------------------------------------------------------
__device__ void update_a1_locus(const DeviceDataset& ds, const ChainDev& c, Rng& r, int64_t l)
{
    const int K = c.K;
    const int64_t KT = ds.KT;
    const double lam = c.Lambda[0];
    const int k2 = ds.k2[l];
    const int afo = ds.afo[l];
    double* pa = c.FreqA + afo;
    double pb[VPS_MAX_ALLELES], pc[VPS_MAX_ALLELES];
    for (int a = 0; a < k2; ++a) pb[a] = lam;
    for (int k = 0; k < K; ++k)
        for (int a = 0; a < k2; ++a) pb[a] += c.Freq[k * KT + afo + a] * c.Fst[K + k];
    dirichlet_a(r, pc, pb, k2, false);
    double dlnL = 0;
    for (int a = 0; a < k2; ++a) dlnL += (pb[a] - lam) * (my_log(pa[a] / pc[a]));
    for (int k = 0; k < K; ++k) {
        const double* pk = c.Freq + k * KT + afo;
        const double fr = c.Fst[K + k];
        for (int a = 0; a < k2; ++a)
            dlnL += lgamma(fr * pa[a]) - lgamma(fr * pc[a]) + fr * (pc[a] - pa[a]) * my_log(pk[a]);
    }
    if (dlnL >= kNZero || r.uniform() < exp(dlnL))
        for (int a = 0; a < k2; ++a) pa[a] = pc[a];
}

__device__ void update_a2_locus(const DeviceDataset& ds, const ChainDev& c, Rng& r, int64_t l)
{
    const int K = c.K;
    const int64_t KT = ds.KT;
    const double lam = c.Lambda[0];
    const int k2 = ds.k2[l];
    const int afo = ds.afo[l];
    double* pa = c.FreqA + afo;
    const double* p = c.Freq + afo;
    const int m1 = (int)r.next(k2), n1 = (int)r.next_avoid(k2, m1);
    const double delta = r.uniform(0, pow((double)ds.N, -0.5));
    const double pm0 = pa[m1], pm1 = pm0 + delta, pn0 = pa[n1], pn1 = pn0 - delta;
    if (pm1 >= 1 || pn1 <= 0) return;
    double dlnL = (lam - 1) * my_log(pm1 * pn1 / pm0 / pn0);
    for (int k = 0; k < K; ++k) {
        const double fr = c.Fst[K + k];
        dlnL += lgamma(fr * pm0) + lgamma(fr * pn0) - lgamma(fr * pm1) - lgamma(fr * pn1) +
                (fr * delta) * my_log(p[k * KT + m1] / p[k * KT + n1]);
    }
    if (dlnL >= kNZero || r.uniform() < exp(dlnL)) { pa[m1] = pm1; pa[n1] = pn1; }
}

// The reference's 32 serial streams (one per block of L/32 loci): one
// wavefront per stream.  Every lane runs the stream's draws (the same
// values on every lane), so the proposal and the decision are wave-uniform.
// The lanes evaluate the log ratio's lgamma calls side by side (one per
// lane), each term is assembled in the reference's expression order, and the
// terms are added one by one in the reference's order, so dlnL has the
// reference's rounding.  A locus's fields come from a batch of 64 loaded one
// per lane.  Lane 0 writes the kept frequencies.
constexpr int kAStreams = 32;

__device__ __forceinline__ double lane_f64(double v, int l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// UpdateA1 at one locus (src/structure.cpp:1672-1700): Dirichlet proposal
// from lam + sum_k f_k P_k, accepted by its log ratio
__device__ void update_a1_locus_wave(const DeviceDataset& ds, const ChainDev& c, Rng& r, int k2, int afo, int lane)
{
    const int K = c.K;
    const int64_t KT = ds.KT;
    const double lam = c.Lambda[0];
    double* pa = c.FreqA + afo;
    // lane a < k2: pb[a] = lam, then + P_k[a] * f_k in cluster order, and pa[a]
    double pb = lam, pal = 0;
    if (lane < k2) {
        for (int k = 0; k < K; ++k) pb += c.Freq[k * KT + afo + lane] * c.Fst[K + k];
        pal = pa[lane];
    }
    // the Dirichlet draw: the stream's k2 Gamma draws in allele order (every
    // lane), their sum in that order, each lane keeping its allele's draw
    double g = 0, sum = 0;
    for (int a = 0; a < k2; ++a) {
        const double v = r.gamma(lane_f64(pb, a) * 1.0);
        if (lane == a) g = v;
        sum += v;
    }
    const double pcl = g * (1.0 / sum);
    // sum_a (pb[a] - lam) log(pa[a] / pc[a]), in allele order
    const double u = lane < k2 ? (pb - lam) * (my_log(pal / pcl)) : 0.0;
    double dlnL = 0;
    for (int a = 0; a < k2; ++a) dlnL += lane_f64(u, a);
    // terms (k, a), k-major as the reference adds them, 64 at a time
    const int T = K * k2;
    for (int t0 = 0; t0 < T; t0 += 64) {
        const int t = t0 + lane;
        const int k = t / k2, a = t - k * k2;
        const double pat = __shfl(pal, a, 64), pct = __shfl(pcl, a, 64);
        double term = 0;
        if (t < T) {
            const double fr = c.Fst[K + k];
            term = lgamma(fr * pat) - lgamma(fr * pct) + fr * (pct - pat) * my_log(c.Freq[k * KT + afo + a]);
        }
        const int n = min(64, T - t0);
        for (int x = 0; x < n; ++x) dlnL += lane_f64(term, x);
    }
    if (dlnL >= kNZero || r.uniform() < exp(dlnL)) {
        // every lane has read pa: the block is this one wavefront
        for (int a = 0; a < k2; ++a) {
            const double v = lane_f64(pcl, a);
            if (lane == 0) pa[a] = v;
        }
    }
}

// UpdateA2 at one locus (src/structure.cpp:1760-1800): a shift between two
// alleles, accepted by its log ratio; lane 4k + j evaluates the j-th lgamma
// of cluster k
__device__ void update_a2_locus_wave(const DeviceDataset& ds, const ChainDev& c, Rng& r, int k2, int afo, int lane,
                                     double sdn)
{
    const int K = c.K;
    const int64_t KT = ds.KT;
    const double lam = c.Lambda[0];
    double* pa = c.FreqA + afo;
    const double* p = c.Freq + afo;
    const int m1 = (int)r.next(k2), n1 = (int)r.next_avoid(k2, m1);
    const double delta = r.uniform(0, sdn);
    const double pm0 = pa[m1], pm1 = pm0 + delta, pn0 = pa[n1], pn1 = pn0 - delta;
    if (pm1 >= 1 || pn1 <= 0) return;
    double dlnL = (lam - 1) * my_log(pm1 * pn1 / pm0 / pn0);
    const int k = lane >> 2, j = lane & 3;
    const double fr = k < K ? c.Fst[K + k] : 1.0;
    const double lg = lgamma(fr * (j == 0 ? pm0 : j == 1 ? pn0 : j == 2 ? pm1 : pn1));
    const double lg1 = __shfl(lg, lane + 1, 64), lg2 = __shfl(lg, lane + 2, 64), lg3 = __shfl(lg, lane + 3, 64);
    double term = 0;
    if (j == 0 && k < K) term = lg + lg1 - lg2 - lg3 + (fr * delta) * my_log(p[k * KT + m1] / p[k * KT + n1]);
    for (int x = 0; x < K; ++x) dlnL += lane_f64(term, 4 * x);
    if (dlnL >= kNZero || r.uniform() < exp(dlnL)) {
        if (lane == 0) { pa[m1] = pm1; pa[n1] = pn1; }
    }
}

__global__ void k_update_a(const DeviceDataset ds, const ChainDev c)
{
    const int m = *c.mdev;
    // UpdateA picks A1 or A2 by HashULong(seed + m) & 1, src/structure.cpp:1639-1647
    const int use_a1 = hash_ulong(c.seed + (uint64_t)m) & 1;
    if (c.topology == VPS_RNG_REFERENCE) {
        // one block of 64 threads (one wavefront) per stream
        const int b = blockIdx.x, lane = threadIdx.x;
        const int64_t l0 = (int64_t)b * ds.L / kAStreams, l1 = (int64_t)(b + 1) * ds.L / kAStreams;
        Rng r;
        r.init(c.seed + (uint64_t)m * kAStreams + b, kSaltUpdateA);
        const double sdn = pow((double)ds.N, -0.5);   // A2's step bound, the same at every locus
        for (int64_t lb = l0; lb < l1; lb += 64) {
            // the next 64 loci's allele counts and offsets, one per lane
            const int64_t ll = min(lb + lane, l1 - 1);
            const int k2v = ds.k2[ll], afov = ds.afo[ll];
            const int nl = (int)min<int64_t>(64, l1 - lb);
            for (int x = 0; x < nl; ++x) {
                const int k2 = __builtin_amdgcn_readlane(k2v, x), afo = __builtin_amdgcn_readlane(afov, x);
                if (use_a1) update_a1_locus_wave(ds, c, r, k2, afo, lane);
                else update_a2_locus_wave(ds, c, r, k2, afo, lane, sdn);
            }
        }
        return;
    }
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ds.L) return;
    Rng r;
    r.init(c.seed + (uint64_t)m * ds.L + x, kSaltUpdateA);
    use_a1 ? update_a1_locus(ds, c, r, x) : update_a2_locus(ds, c, r, x);
}

int launch_update_a(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (!c.fmodel) return VPS_OK;
    if (c.topology == VPS_RNG_REFERENCE)
        hipLaunchKernelGGL(k_update_a, dim3(kAStreams), dim3(64), 0, st, ds, c);
    else
        hipLaunchKernelGGL(k_update_a, dim3((unsigned)((ds.L + 255) / 256)), dim3(256), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// UpdateF runs as K steps of a grid-wide sum.  The per-cluster log ratio
// sums a term over every allele of the cluster (all clusters with a single
// F), lgamma twice per term: spread over many blocks, with the proposal and
// the decision on one thread.  k_update_f_start draws the first proposal;
// step k sums its terms per block, and the last block to finish (a counter
// in device memory) adds the block sums in block order, decides, and draws
// the next proposal from the same stream.  The stream's draws (a normal per
// cluster, a uniform only when the log ratio is below -1e-10) are the
// reference's, in its order.
constexpr int kFBlocks = 256;

__device__ void f_propose(const ChainDev& c, Rng& r, int k)
{
    const double Fo = c.Fst[k];
    const double Fm = r.normal(Fo, c.stdf);
    c.fctl[0] = Fo;
    c.fctl[1] = Fm;
    c.fctl[2] = (Fm < 0 || Fm > 1) ? 1.0 : 0.0;
}

__global__ void k_update_f_start(const ChainDev c)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    Rng r;
    r.init(c.seed + (uint64_t)*c.mdev, kSaltUpdateF);
    f_propose(c, r, 0);
    c.fctl[3] = 0.0;
    c.frng[0] = r;
}

__global__ __launch_bounds__(kScalarThreads) void k_update_f_step(const DeviceDataset ds, const ChainDev c, int k)
{
    __shared__ double sh[kScalarThreads];
    __shared__ uint32_t ticket;
    const int K = c.K;
    const int64_t KT = ds.KT;
    if (c.fctl[3] != 0.0) return;   // done: a single F decided, or all clusters
    const bool skip = c.fctl[2] != 0.0;
    const double Fo = c.fctl[0], Fm = c.fctl[1];
    const double fo = c.Fst[K + k], fm = (1 - Fm) / Fm;
    double s = 0;
    if (!skip) {
        // terms: cluster k (clusters k..K-1 with a single F) x every allele
        const int64_t n = (c.fsame ? (int64_t)(K - k) : 1) * KT;
        for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
            const int64_t k2 = k + x / KT, a = x % KT;
            const double pa2 = c.FreqA[a];
            s += pa2 * (fm - fo) * my_log(c.Freq[k2 * KT + a]) + lgamma(fo * pa2) - lgamma(fm * pa2);
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        c.fpart[blockIdx.x] = sh[0];
        __threadfence();
        ticket = atomicAdd(c.fcnt, 1u);
    }
    __syncthreads();
    if (ticket != gridDim.x - 1 || threadIdx.x != 0) return;
    // the last block: every partial is visible
    __threadfence();
    *c.fcnt = 0;
    Rng r = c.frng[0];
    if (!skip) {
        double total = 0;
        for (unsigned b = 0; b < gridDim.x; ++b) total += c.fpart[b];
        double dlnL = c.pmeanf * (Fo - Fm) / (c.pstdf * c.pstdf) +
                      (c.pmeanf * c.pmeanf / (c.pstdf * c.pstdf) - 1) * my_log(Fm / Fo) +
                      (double)((int64_t)(c.fsame ? K : 1) * ds.L) * (lgamma(fm) - lgamma(fo));
        dlnL += total;
        if (dlnL >= kNZero || r.uniform() < exp(dlnL)) {
            if (c.fsame) for (int j = 0; j < K; ++j) { c.Fst[j] = Fm; c.Fst[K + j] = fm; }
            else { c.Fst[k] = Fm; c.Fst[K + k] = fm; }
        }
    }
    if (c.fsame || k + 1 >= K) c.fctl[3] = 1.0;
    else f_propose(c, r, k + 1);
    c.frng[0] = r;
}

int launch_update_f(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    if (!c.fmodel) return VPS_OK;
    hipLaunchKernelGGL(k_update_f_start, dim3(1), dim3(64), 0, st, c);
    // a single F decides at k = 0, over every cluster's terms
    const int steps = c.fsame ? 1 : c.K;
    for (int k = 0; k < steps; ++k) {
        const int64_t n = (c.fsame ? (int64_t)c.K : 1) * ds.KT;
        const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kFBlocks, (n + 4 * kScalarThreads - 1) / (4 * kScalarThreads)));
        hipLaunchKernelGGL(k_update_f_step, dim3(nb), dim3(kScalarThreads), 0, st, ds, c, k);
    }
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// Record's accumulators, src/structure.cpp:1877-1878 and 2017-2041
// ---------------------------------------------------------------------------
// writelnl trace slot of the Record at sweep m: Records happen at burn-in
// sweeps m % nthinning == 0 (writelnl only) and at post-burn-in sweeps
// (m - nburnin) % nthinning == 0 (src/structure.cpp:1943-1950); the slot is
// the number of Records before m.
__device__ int record_trace_slot(const ChainDev& c, int m)
{
    if (!c.writelnl) return -1;
    const int burn = min(m, c.nburnin);
    const int pre = (burn + c.nthinning - 1) / c.nthinning;
    const int post = m > c.nburnin ? (m - c.nburnin + c.nthinning - 1) / c.nthinning : 0;
    const int slot = pre + post;
    return slot < c.lnl_cap ? slot : -1;
}

__global__ void k_record_acc(const DeviceDataset ds, const ChainDev c, int post, int trace, int acc_misum)
{
    const int trace_slot = trace ? record_trace_slot(c, *c.mdev) : -1;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int K = c.K;
    const int64_t nmi = (int64_t)ds.N * K, nfr = (int64_t)K * ds.KT;
    if (acc_misum && x < nmi) c.MiSum[x] += c.Mi[x];
    if (post && x < nfr) {
        // Add(FreqM, Freq): on float arrays when REAL = float
        if (real32(c)) c.FreqM[x] = (double)((float)c.FreqM[x] + (float)c.Freq[x]);
        else c.FreqM[x] += c.Freq[x];
    }
    if (x != 0) return;
    const double lik = c.scal[1];
    if (trace_slot >= 0) c.lnltrace[trace_slot] = lik;
    if (!post) return;
    c.rout[0] += lik;
    c.rout[1] += lik * lik;
    const int nf = c.fmodel ? (c.fsame ? 1 : K) : 0;
    const int na = c.admix ? (c.diffalpha ? K : 1) : 0;
    const int nl = c.difflambda ? K : 1;
    for (int j = 0; j < nl; ++j) c.rout[4 + j] += c.Lambda[j];
    if (c.locpriori) {
        c.rout[4 + nl] += c.scal[0];
        for (int k = 0; k < K; ++k) c.rout[4 + nl + 1 + k] += c.admix ? c.Alpha[k] : c.Eta[k];
        for (int j = 0; j < c.S * K; ++j) c.rout[4 + nl + 1 + K + j] += c.admix ? c.AlphaLocal[j] : c.Gamma[j];
    } else if (c.admix) {
        for (int j = 0; j < na; ++j) c.rout[4 + nl + j] += c.Alpha[j];
    }
    for (int j = 0; j < nf; ++j) c.rout[c.rlen - nf + j] += c.Fst[j];
}

__global__ void k_advance_m(int32_t* mdev) { mdev[0] += 1; }

int launch_advance_m(const ChainDev& c, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance_m, dim3(1), dim3(1), 0, st, c.mdev);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

int launch_record_accumulate(const DeviceDataset& ds, const ChainDev& c, int post, int trace, int acc_misum, hipStream_t st)
{
    const int64_t n = std::max<int64_t>((int64_t)ds.N * c.K, (int64_t)c.K * ds.KT);
    hipLaunchKernelGGL(k_record_acc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ds, c, post, trace, acc_misum);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// InitAdmix, src/structure.cpp:557-649.  Non-single-Z start: per locus one
// stream RNG(seed + l, INITADM), Next(K) per typed allele copy (no draws for
// missing copies).  Lanes = loci; allele counts in lane-owned LDS columns;
// per individual the wavefront ballots its K cluster counts into Mi.
// Single-Z start: Z_i = Next(K) from RNG(seed - 1, INITADM), Mi[i][Z_i] = vt_i.
// ---------------------------------------------------------------------------
// (in_lds = 0: the K x alleles table does not fit in LDS; counts go to Ni
// by global atomics, Ni cleared first)
__global__ __launch_bounds__(64) void k_init_admix(const DeviceDataset ds, const ChainDev c, int A, int list_kind,
                                                  int in_lds)
{
    extern __shared__ int32_t ni_lds[];
    const int lane = threadIdx.x;
    const int K = c.K;
    if (in_lds)
        for (int x = lane; x < K * A * kTileLoci; x += 64) ni_lds[x] = 0;
    __syncthreads();
    const int64_t tix = blockIdx.x;
    const int64_t ntl = list_kind == kKindBi ? ds.nbi : ds.ngen;
    if (tix >= ntl) return;
    const TileDesc t = ds.tiles[list_kind == kKindBi ? ds.bi_list[tix] : ds.gen_list[tix]];
    const bool valid = lane < t.nloc;
    const int64_t l = ds.tloc[t.lmap + lane];
    const int afo = ds.afo[l];
    const int k2 = ds.k2[l];
    const uint64_t tab = valid ? ds.bitab[l] : 0ull;
    const uint8_t* btab = ds.btab + t.dbase + lane;   // BI tiles above 3 bits
    const uint64_t* gd = ds.gdesc + ds.gbase[l];
    Rng r;
    r.init(c.seed + (uint64_t)l, kSaltInitAdm);
    const int bits = t.bits, ipw = t.ipw;
    const uint32_t mask = (1u << bits) - 1u;
    const uint32_t* wp = ds.words + t.wbase + lane;
    uint32_t word = wp[0];
    int j = 0, w = 0;
    for (int i = 0; i < ds.N; ++i) {
        const uint32_t gid = word & mask;
        word >>= bits;
        if (++j == ipw) {
            j = 0; ++w;
            if (i + 1 < ds.N) word = wp[(int64_t)w * kTileLoci];
        }
        int v, h = 0;
        uint64_t d64 = 0;
        if (t.kind == kKindBi) {
            const uint32_t desc = bits > kBiRegBits ? (uint32_t)btab[gid * kTileLoci] : (uint32_t)(tab >> (gid * 8)) & 0xFFu;
            v = (int)(desc & 15u);
            h = v - (int)(desc >> 4);
        } else {
            d64 = valid ? gd[gid] : 0ull;
            v = (int)(d64 & 15u);
        }
        const int MP = ds.maxploidy;
        for (int a = 0; a < MP; ++a) {
            const bool typed = a < v;
            int kk = 0;
            if (typed) {
                kk = (int)r.next((uint64_t)K);
                const int x = t.kind == kKindBi ? (a >= h ? 1 : 0) : (int)((d64 >> (4 + 6 * a)) & 63u);
                if (in_lds) ni_lds[(kk * A + x) * kTileLoci + lane] += 1;
                else atomicAdd(c.Ni + (int64_t)kk * ds.KT + afo + x, 1);
            }
            for (int k = 0; k < K; ++k) {
                const int cnt = __builtin_popcountll(__builtin_amdgcn_ballot_w64(typed && kk == k));
                if (lane == 0 && cnt) atomicAdd(c.Mi + (int64_t)i * K + k, cnt);
            }
        }
    }
    if (valid && in_lds)
        for (int k = 0; k < K; ++k)
            for (int a = 0; a < k2; ++a) c.Ni[(int64_t)k * ds.KT + afo + a] = ni_lds[(k * A + a) * kTileLoci + lane];
}

__global__ void k_init_singlez(const DeviceDataset ds, const ChainDev c)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Rng r;
    r.init(c.seed - 1, kSaltInitAdm);
    for (int i = 0; i < ds.N; ++i) {
        const int z = (int)r.next((uint64_t)c.K);
        c.Z[i] = (uint16_t)z;
        c.Mi[(int64_t)i * c.K + z] = ds.vt[i];
    }
}

__global__ void k_set_mi_binary(const DeviceDataset ds, const ChainDev c)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)ds.N * c.K) return;
    const int i = (int)(x / c.K), k = (int)(x % c.K);
    c.Mi[x] = c.Z[i] == k ? ds.vt[i] : 0;
}

int launch_set_mi_binary(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    const int64_t n = (int64_t)ds.N * c.K;
    hipLaunchKernelGGL(k_set_mi_binary, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// Arrange, src/structure.cpp:1862-1869: MiSum rows -> normalised doubles with
// MIN_FREQ (UnifyInt64ToDouble), rout and FreqM scaled by 1/nr, variance and
// lnP(D) = mean - var/2.
// ---------------------------------------------------------------------------
__global__ void k_arrange(const DeviceDataset ds, const ChainDev c, int nr)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int K = c.K;
    if (x < ds.N) {
        int64_t* row = c.MiSum + x * K;
        double* rd = reinterpret_cast<double*>(row);
        double tmp[VPS_MAX_K];
        for (int k = 0; k < K; ++k) tmp[k] = (double)row[k];
        unify_row(tmp, K);
        for (int k = 0; k < K; ++k) rd[k] = tmp[k];
    }
    const double inv = 1.0 / nr;
    if (x < (int64_t)K * ds.KT) {
        // Mul(FreqM, (REAL)(1.0 / nr))
        if (real32(c)) c.FreqM[x] = (double)((float)c.FreqM[x] * (float)inv);
        else c.FreqM[x] *= inv;
    }
    if (x == 0) {
        for (int j = 0; j < c.rlen; ++j) c.rout[j] *= inv;
        c.rout[2] = c.rout[1] - c.rout[0] * c.rout[0];
        c.rout[3] = c.rout[0] - c.rout[2] / 2;
    }
}

int launch_arrange(const DeviceDataset& ds, const ChainDev& c, int nr, hipStream_t st)
{
    const int64_t n = std::max<int64_t>((int64_t)ds.N, (int64_t)c.K * ds.KT);
    hipLaunchKernelGGL(k_arrange, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ds, c, nr);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

int launch_init_admix(const DeviceDataset& ds, const ChainDev& c, int singlez, hipStream_t st)
{
    int rc = launch_fill_i32(c.Mi, (int64_t)ds.N * c.K, 0, st);
    if (rc) return rc;
    if (singlez) {
        hipLaunchKernelGGL(k_init_singlez, dim3(1), dim3(64), 0, st, ds, c);
        VPS_HIP(hipGetLastError());
        return launch_z_noadmix(ds, c, st);
    }
    const int A = ds.maxK > 2 ? ds.maxK : 2;
    size_t shm = (size_t)c.K * A * kTileLoci * sizeof(int32_t);
    const int in_lds = shm <= 64 * 1024;
    if (!in_lds) {
        rc = launch_fill_i32(c.Ni, (int64_t)c.K * ds.KT, 0, st);
        if (rc) return rc;
        shm = 0;
    }
    if (ds.nbi > 0) hipLaunchKernelGGL(k_init_admix, dim3((unsigned)ds.nbi), dim3(64), shm, st, ds, c, A, kKindBi, in_lds);
    if (ds.ngen > 0) hipLaunchKernelGGL(k_init_admix, dim3((unsigned)ds.ngen), dim3(64), shm, st, ds, c, A, kKindGen, in_lds);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

}  // namespace vps

// vps_lik.hip — the genotype passes that reduce log likelihoods:
//   mode 0  Record, admixture        lnL = sum log(sum_k Q[i][k] P[k][l][x])
//   mode 1  Record, binary Q          lnL = sum log P[Z_i][l][x]
//           (RecordAVX, src/structureAVX.cpp:986-1187)
//   mode 2  Metropolis log ratio     r_i = sum log(sum_k q'_ik p / sum_k q_ik p)
//           (UpdateQMetroAVX, src/structureAVX.cpp:557-783): each copy's
//           ratio is formed as in SumProdDiv (src/math2.cpp:1557-1580),
//           products and sums rounded separately, then divided, and the
//           ratios are charged into one accumulator (FP64 mode)
//   mode 4  the same log ratio as the difference of two accumulated
//           mixtures formed with fused multiply-adds (mixed-precision mode:
//           half the FP64 instructions, no division)
//   mode 3  no-admix log likelihood  b_ik = sum log P[k][l][x]
//           (UpdateQNoAdmixAVX, src/structureAVX.cpp:254-428)
//   mode 5  mode 0 with the mixture formed in single precision, and
//   mode 6  mode 2 with both mixtures and their ratio in single precision:
//           REAL = float with -g_fastsingle (src/structureAVX.cpp:725-760,
//           1148-1168); the factors are then charged as doubles
// sums run over typed allele copies x of individual i at locus l.
//
// These passes draw no random numbers, so lanes map to individuals: a lane
// keeps its individual's Q row in registers and its own log accumulator
// across a range of tiles, with no cross-lane reduction inside the loop.
// Within one tile every lane reads the same locus, so on biallelic tiles the
// 2K frequencies of a locus are wave-uniform scalar loads and a lane forms
// both allele mixtures f_a = sum_k q_k P[k][l][a] once per locus, then
// charges one of them per allele copy.  Multiallelic tiles gather P per copy.
//
// The accumulator keeps the product as mantissa in [0.5, 1) times 2^e
// (frexp after each factor), which is the reference's exponent-stripping
// log accumulator (src/math2.h:716-902) with the strip done every time:
// scaling by powers of two is exact, so only the product order differs from
// the reference's 64-lane products, and the sums agree to rounding.
// Each (locus range, individual) closes one partial; record passes sum a
// block's partials in LDS in a fixed tree, and a fixed-order pass adds the
// block sums (or, per individual, the range partials).
#include <type_traits>
#include "vps_chain.h"

namespace vps {

constexpr int kLikThreads = 256;

// modes 0/1: sum the block's 256 partials in a fixed tree
__device__ void block_store_sum(double v, double* dst)
{
    __shared__ double sh[kLikThreads];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kLikThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *dst = sh[0];
}

template <int K, int MODE>
__device__ void lik_load_q(const ChainDev& c, int i, double* q, double* qp)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (MODE == 1) q[k] = c.Z[i] == k ? 1.0 : 0.0;   // one-hot: sum_k q_k p_k == p_Z exactly
        else if (MODE != 3) q[k] = c.Q[(int64_t)i * K + k];
        else q[k] = 0.0;
        qp[k] = (MODE == 2 || MODE == 4 || MODE == 6) ? c.bufNK1[(int64_t)i * K + k] : 0.0;
    }
}

template <int K, int MODE>
__device__ void lik_write(const DeviceDataset& ds, const ChainDev& c, FrexpAcc* acc, bool ok, int i, int64_t range,
                          int64_t block_slot)
{
    if (MODE <= 1 || MODE == 5) {
        block_store_sum(ok ? acc[0].close() : 0.0, c.part + block_slot);
    } else if (MODE == 2 || MODE == 4 || MODE == 6) {
        if (ok) {
            const double cur = acc[1].close();
            c.part[range * ds.N + i] = MODE != 4 ? acc[0].close() : acc[0].close() - cur;
            c.part_cur[range * ds.N + i] = cur;
        }
    } else {
        if (ok) {
            double* out = c.part + range * ds.N * K + (int64_t)i * K;
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] = acc[k].close();
        }
    }
}

// number of accumulators per lane
template <int K, int MODE>
struct LikNA { static constexpr int value = MODE == 3 ? K : ((MODE == 2 || MODE == 4 || MODE == 6) ? 2 : 1); };

// biallelic tiles (at most two alleles; genotype bytes in ds.bitab, or in
// the tile's ds.btab columns above 3 bits per id).  MPT is
// the maximum ploidy when it is 2 or 4 (copies loop fully unrolled), else 0.
// Resident wavefronts per SIMD the BI likelihood kernel asks the register
// allocator for: the FP64 Metropolis ratio (mode 2) at K = 5 takes 106
// registers on its own (4 wavefronts) and fits 5 in 94 with no spills (its
// pass 0.69 -> 0.65 ms at N = 2,000, L = 100,000); at K <= 4 every mode
// already fits 5, and from K = 6 a cap of 5 spills.
template <int K, int MODE>
constexpr int lik_bi_waves() { return MODE == 2 && K == 5 ? 5 : 1; }

template <int K, int MODE, int MPT>
__global__ __launch_bounds__(kLikThreads, (lik_bi_waves<K, MODE>())) void k_lik_bi(const DeviceDataset ds, const ChainDev c,
                                                                               int tiles_per_range)
{
    const int i = blockIdx.y * kLikThreads + threadIdx.x;
    const bool ok = i < ds.N;
    const int ii = ok ? i : 0;
    const int64_t KT = ds.KT;
    double q[K], qp[K];
    lik_load_q<K, MODE>(c, ii, q, qp);
    constexpr int NA = LikNA<K, MODE>::value;
    FrexpAcc acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a].open();
    const int ncopies = MPT > 0 ? MPT : ds.maxploidy;

    const int lane = threadIdx.x & 63;
    const int64_t r0 = (int64_t)blockIdx.x * tiles_per_range;
    const int64_t r1 = min(ds.nbi, r0 + tiles_per_range);
    for (int64_t r = r0; r < r1; ++r) {
        const TileDesc td = ds.tiles[ds.bi_list[r]];
        const int w = ii / td.ipw;
        const int sh = (ii - w * td.ipw) * td.bits;
        const uint32_t mask = (1u << td.bits) - 1u;
        const uint32_t* wp = ds.words + td.wbase + (int64_t)w * kTileLoci;
        const bool big = td.bits > kBiRegBits;
        const uint8_t* btab = ds.btab + td.dbase;
        // the tile's per-locus fields, lane j holding locus j (one vector
        // load each instead of a chain of scalar loads per locus)
        const int64_t lj = ds.tloc[td.lmap + lane];
        const int afo_v = ds.afo[lj];
        const uint64_t tab_v = ds.bitab[lj];
        // one copy of the loop per table kind: the register-table loop then
        // has no vector load besides the id prefetch to wait for
        auto tile_loop = [&](auto big_tag) {
            constexpr bool BIG = decltype(big_tag)::value;
            constexpr int kPairs = MPT > 0 ? (MPT + 1) / 2 : 1;   // copy pairs of the unrolled loop
            // ids of 4 loci per 16-byte load (a tile's word rows hold 64
            // entries, so a group never leaves its row); the next group's load
            // is issued before this group's work
            uint4 wn = *reinterpret_cast<const uint4*>(wp);
            // a locus's 2K frequencies (both alleles in one 16-byte scalar
            // load; P is padded past KT), loaded one locus ahead: they are
            // first used a locus later, so the wait for them covers loads
            // issued a whole locus of work earlier
            // (two buffers, alternating by locus parity, so no copies)
            double xa0[K], xa1[K], xb0[K], xb1[K];
            auto load_p = [&](int j, double* d0, double* d1) {
                const double* pl = c.Freq + __builtin_amdgcn_readlane(afo_v, j);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    d0[k] = pl[k * KT];
                    d1[k] = pl[k * KT + 1];
                }
            };
            load_p(0, xa0, xa1);
            for (int j0 = 0; j0 < td.nloc; j0 += 4) {
                const uint4 w4 = wn;
                if (j0 + 4 < td.nloc) wn = *reinterpret_cast<const uint4*>(wp + j0 + 4);
                const uint32_t wq[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = j0 + u;
                    if (j >= td.nloc) break;
                    // (on a monomorphic locus no copy carries allele 1, so the
                    // second value, the next locus's, is never charged)
                    const double* p0 = (u & 1) ? xb0 : xa0;
                    const double* p1 = (u & 1) ? xb1 : xa1;
                    if (j + 1 < td.nloc) {
                        if (u & 1) load_p(j + 1, xa0, xa1);
                        else load_p(j + 1, xb0, xb1);
                    }
                    const uint64_t tab = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(tab_v >> 32), j) << 32) |
                                         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)tab_v, j);
                    const uint32_t gid = (wq[u] >> sh) & mask;
                    const uint32_t desc = BIG ? (uint32_t)btab[gid * kTileLoci + j] : (uint32_t)(tab >> (gid * 8)) & 0xFFu;
                    const int v = ok ? (int)(desc & 15u) : 0;
                    const int h = v - (int)(desc >> 4);    // copies a >= h carry allele 1
                    // Copies are charged in pairs: the product of two factors is
                    // formed first.  Every factor is a frequency or a mixture of
                    // frequencies, above 1e-154 in any chain, so a pair cannot
                    // underflow; only the product order differs from per-copy
                    // charging.
                    if (MODE == 3) {
#pragma unroll kPairs
                        for (int a = 0; a < ncopies; a += 2) {
                            const bool on0 = a < v, hi0 = a >= h, on1 = a + 1 < v, hi1 = a + 1 >= h;
#pragma unroll
                            for (int k = 0; k < K; ++k)
                                acc[k].charge((on0 ? (hi0 ? p1[k] : p0[k]) : 1.0) * (on1 ? (hi1 ? p1[k] : p0[k]) : 1.0));
                        }
                    } else if (MODE == 5 || MODE == 6) {
                        // single-precision mixtures of single-precision P and
                        // Q, and (mode 6) their ratio in single precision
                        float f0 = (float)p0[0] * (float)q[0], f1 = (float)p1[0] * (float)q[0];
                        float g0 = MODE == 6 ? (float)p0[0] * (float)qp[0] : 0.f;
                        float g1 = MODE == 6 ? (float)p1[0] * (float)qp[0] : 0.f;
#pragma unroll
                        for (int k = 1; k < K; ++k) {
                            f0 = f0 + (float)p0[k] * (float)q[k];
                            f1 = f1 + (float)p1[k] * (float)q[k];
                            if (MODE == 6) {
                                g0 = g0 + (float)p0[k] * (float)qp[k];
                                g1 = g1 + (float)p1[k] * (float)qp[k];
                            }
                        }
                        const double d0 = (double)f0, d1 = (double)f1;
                        const double r0 = MODE == 6 ? (double)(g0 / f0) : 0.0, r1 = MODE == 6 ? (double)(g1 / f1) : 0.0;
#pragma unroll kPairs
                        for (int a = 0; a < ncopies; a += 2) {
                            const bool on0 = a < v, hi0 = a >= h, on1 = a + 1 < v, hi1 = a + 1 >= h;
                            if (MODE == 6)
                                acc[0].charge((on0 ? (hi0 ? r1 : r0) : 1.0) * (on1 ? (hi1 ? r1 : r0) : 1.0));
                            acc[MODE == 6 ? 1 : 0].charge((on0 ? (hi0 ? d1 : d0) : 1.0) * (on1 ? (hi1 ? d1 : d0) : 1.0));
                        }
                    } else {
                        // sums start from the k = 0 term (0 + x == x for the
                        // non-negative terms here).  Record and the FP64
                        // Metropolis ratio keep the product-then-sum rounding
                        // of SumProdDiv; mode 4 fuses the multiply-adds (one
                        // rounding per term instead of two).
                        constexpr bool MET = MODE == 2 || MODE == 4;
                        double f0 = p0[0] * q[0], f1 = p1[0] * q[0];
                        double g0 = MET ? p0[0] * qp[0] : 0.0, g1 = MET ? p1[0] * qp[0] : 0.0;
#pragma unroll
                        for (int k = 1; k < K; ++k) {
                            if (MODE == 4) {
                                f0 = __builtin_fma(p0[k], q[k], f0);
                                f1 = __builtin_fma(p1[k], q[k], f1);
                                g0 = __builtin_fma(p0[k], qp[k], g0);
                                g1 = __builtin_fma(p1[k], qp[k], g1);
                            } else {
                                f0 = f0 + p0[k] * q[k];
                                f1 = f1 + p1[k] * q[k];
                                if (MODE == 2) {
                                    g0 = g0 + p0[k] * qp[k];
                                    g1 = g1 + p1[k] * qp[k];
                                }
                            }
                        }
                        // the per-copy ratios (correctly rounded divisions);
                        // a ratio of an allele no copy carries is never charged
                        const double r0 = MODE == 2 ? g0 / f0 : 0.0, r1 = MODE == 2 ? g1 / f1 : 0.0;
#pragma unroll kPairs
                        for (int a = 0; a < ncopies; a += 2) {
                            const bool on0 = a < v, hi0 = a >= h, on1 = a + 1 < v, hi1 = a + 1 >= h;
                            if (MODE == 2)
                                acc[0].charge((on0 ? (hi0 ? r1 : r0) : 1.0) * (on1 ? (hi1 ? r1 : r0) : 1.0));
                            else if (MODE == 4)
                                acc[0].charge((on0 ? (hi0 ? g1 : g0) : 1.0) * (on1 ? (hi1 ? g1 : g0) : 1.0));
                            acc[MET ? 1 : 0].charge((on0 ? (hi0 ? f1 : f0) : 1.0) * (on1 ? (hi1 ? f1 : f0) : 1.0));
                        }
                    }
                }
            }
        };
        if (big) tile_loop(std::true_type{});
        else tile_loop(std::false_type{});
    }
    lik_write<K, MODE>(ds, c, acc, ok, i, (int64_t)blockIdx.x, (int64_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// multiallelic tiles: genotype descriptors in ds.gdesc, P gathered per copy
template <int K, int MODE>
__global__ __launch_bounds__(kLikThreads) void k_lik_gen(const DeviceDataset ds, const ChainDev c, int tiles_per_range,
                                                        int range_base, int64_t slot_base)
{
    const int i = blockIdx.y * kLikThreads + threadIdx.x;
    const bool ok = i < ds.N;
    const int ii = ok ? i : 0;
    const int64_t KT = ds.KT;
    double q[K], qp[K];
    lik_load_q<K, MODE>(c, ii, q, qp);
    constexpr int NA = LikNA<K, MODE>::value;
    FrexpAcc acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a].open();

    const int lane = threadIdx.x & 63;
    const int64_t r0 = (int64_t)blockIdx.x * tiles_per_range;
    const int64_t r1 = min(ds.ngen, r0 + tiles_per_range);
    for (int64_t r = r0; r < r1; ++r) {
        const TileDesc td = ds.tiles[ds.gen_list[r]];
        const int w = ii / td.ipw;
        const int sh = (ii - w * td.ipw) * td.bits;
        const uint32_t mask = (1u << td.bits) - 1u;
        const uint32_t* wp = ds.words + td.wbase + (int64_t)w * kTileLoci;
        // the tile's per-locus fields, lane j holding locus j
        const int64_t lj = ds.tloc[td.lmap + lane];
        const int afo_v = ds.afo[lj];
        const int64_t gb_v = ds.gbase[lj];
        // ids of 4 loci per 16-byte load, the next group's issued first;
        // each locus's genotype descriptor is loaded one locus ahead
        auto desc_of = [&](int j, uint32_t word) -> uint64_t {
            const int64_t gb = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(gb_v >> 32), j) << 32) |
                                         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)gb_v, j));
            return ds.gdesc[gb + ((word >> sh) & mask)];
        };
        uint4 wn = *reinterpret_cast<const uint4*>(wp);
        uint64_t dn = desc_of(0, wn.x);
        for (int j0 = 0; j0 < td.nloc; j0 += 4) {
            const uint4 w4 = wn;
            if (j0 + 4 < td.nloc) wn = *reinterpret_cast<const uint4*>(wp + j0 + 4);
            const uint32_t wq[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u;
                if (j >= td.nloc) break;
                const uint64_t d64 = dn;
                if (j + 1 < td.nloc) dn = desc_of(j + 1, u < 3 ? wq[u + 1] : wn.x);
                const int v = ok ? (int)(d64 & 15u) : 0;
                const double* pl = c.Freq + __builtin_amdgcn_readlane(afo_v, j);
                for (int a = 0; a < v; ++a) {
                    const double* px = pl + (int)((d64 >> (4 + 6 * a)) & 63u);
                    if (MODE == 3) {
#pragma unroll
                        for (int k = 0; k < K; ++k) acc[k].charge(px[k * KT]);
                    } else if (MODE == 5 || MODE == 6) {
                        float f = 0.f, g = 0.f;
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const float p = (float)px[k * KT];
                            f = k ? f + p * (float)q[k] : p * (float)q[k];
                            if (MODE == 6) g = k ? g + p * (float)qp[k] : p * (float)qp[k];
                        }
                        if (MODE == 6) {
                            acc[0].charge((double)(g / f));
                            acc[1].charge((double)f);
                        } else {
                            acc[0].charge((double)f);
                        }
                    } else {
                        // (rounding as on BI tiles: separate products and sums
                        // and a per-copy ratio in mode 2, fused in mode 4)
                        double f = 0, g = 0;
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const double p = px[k * KT];
                            if (MODE == 4) {
                                f = __builtin_fma(p, q[k], f);
                                g = __builtin_fma(p, qp[k], g);
                            } else {
                                f = f + p * q[k];
                                if (MODE == 2) g = g + p * qp[k];
                            }
                        }
                        if (MODE == 2) {
                            acc[0].charge(g / f);
                            acc[1].charge(f);
                        } else if (MODE == 4) {
                            acc[0].charge(g);
                            acc[1].charge(f);
                        } else {
                            acc[0].charge(f);
                        }
                    }
                }
            }
        }
    }
    lik_write<K, MODE>(ds, c, acc, ok, i, range_base + (int64_t)blockIdx.x,
                       slot_base + (int64_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// lnL = sum of the block partials, one block, fixed order
__global__ __launch_bounds__(256) void k_reduce_lnl(const double* __restrict__ part, int64_t n, double* __restrict__ out)
{
    __shared__ double sh[256];
    double s = 0;
    for (int64_t x = threadIdx.x; x < n; x += 256) s += part[x];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// per-individual (x NA) sums over locus ranges; mode 3 adds the location
// prior log Gamma[pop][k] for typed individuals (charged first in the
// reference, src/structureAVX.cpp:260-266).  A block covers 64 outputs with
// 4 threads each; thread s sums ranges s, s+4, ... and the four partials are
// added in a fixed order.
constexpr int kRowSplit = 4;
// rows_sum: one split's share of the range partials of entry x, in a fixed order
__device__ __forceinline__ double rows_sum(const double* __restrict__ p, int64_t n, int nranges, int split)
{
    int r = split;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (; r + 3 * kRowSplit < nranges; r += 4 * kRowSplit) {
        a0 += p[(int64_t)r * n];
        a1 += p[(int64_t)(r + kRowSplit) * n];
        a2 += p[(int64_t)(r + 2 * kRowSplit) * n];
        a3 += p[(int64_t)(r + 3 * kRowSplit) * n];
    }
    for (; r < nranges; r += kRowSplit) a0 += p[(int64_t)r * n];
    return (a0 + a1) + (a2 + a3);
}

// cur/dst_cur (the Metropolis pass): the current-Q partials, summed the same way
__global__ __launch_bounds__(256) void k_reduce_rows(const DeviceDataset ds, const ChainDev c, int nranges, int NA,
                                                     int add_prior, double* __restrict__ dst,
                                                     const double* __restrict__ cur, double* __restrict__ dst_cur)
{
    __shared__ double sh[kRowSplit][64];
    __shared__ double sc[kRowSplit][64];
    const int lane = threadIdx.x & 63, split = threadIdx.x >> 6;
    const int64_t x = (int64_t)blockIdx.x * 64 + lane;
    const int64_t n = (int64_t)ds.N * NA;
    double s = 0, sk = 0;
    if (x < n) {
        s = rows_sum(c.part + x, n, nranges, split);
        if (cur) sk = rows_sum(cur + x, n, nranges, split);
    }
    sh[split][lane] = s;
    sc[split][lane] = sk;
    __syncthreads();
    if (split != 0 || x >= n) return;
    if (cur) dst_cur[x] = (sc[0][lane] + sc[1][lane]) + (sc[2][lane] + sc[3][lane]);
    double t = 0;
    if (add_prior) {
        const int i = (int)(x / NA), k = (int)(x % NA);
        if (ds.vt[i]) {
            LogAcc a;
            a.open();
            a.charge(c.Gamma[(int64_t)ds.popid[i] * c.K + k]);
            t = a.close();
        }
    }
    dst[x] = t + ((sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]));
}

namespace {

struct LikGrid {
    int nr_bi, nr_gen, gy;
};

LikGrid lik_grid(const DeviceDataset& ds, const int32_t* tpr)
{
    LikGrid g;
    g.nr_bi = (int)((ds.nbi + tpr[0] - 1) / tpr[0]);
    g.nr_gen = (int)((ds.ngen + tpr[1] - 1) / tpr[1]);
    g.gy = (ds.N + kLikThreads - 1) / kLikThreads;
    return g;
}

template <int K, int MODE>
int launch_lik_k(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, const LikGrid& g)
{
    if (g.nr_bi > 0) {
        const dim3 grid((unsigned)g.nr_bi, (unsigned)g.gy);
        if (ds.maxploidy == 2)
            hipLaunchKernelGGL((k_lik_bi<K, MODE, 2>), grid, dim3(kLikThreads), 0, st, ds, c, c.lik_tpr[0]);
        else if (ds.maxploidy == 4)
            hipLaunchKernelGGL((k_lik_bi<K, MODE, 4>), grid, dim3(kLikThreads), 0, st, ds, c, c.lik_tpr[0]);
        else
            hipLaunchKernelGGL((k_lik_bi<K, MODE, 0>), grid, dim3(kLikThreads), 0, st, ds, c, c.lik_tpr[0]);
        VPS_HIP(hipGetLastError());
    }
    if (g.nr_gen > 0) {
        const dim3 grid((unsigned)g.nr_gen, (unsigned)g.gy);
        hipLaunchKernelGGL((k_lik_gen<K, MODE>), grid, dim3(kLikThreads), 0, st, ds, c, c.lik_tpr[1], g.nr_bi,
                           (int64_t)g.nr_bi * g.gy);
        VPS_HIP(hipGetLastError());
    }
    return VPS_OK;
}

template <int MODE>
int launch_lik_mode(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, const LikGrid& g)
{
    switch (c.K) {
#define VPS_CASE(k) case k: return launch_lik_k<k, MODE>(ds, c, st, g);
        VPS_CASE(1) VPS_CASE(2) VPS_CASE(3) VPS_CASE(4) VPS_CASE(5) VPS_CASE(6) VPS_CASE(7) VPS_CASE(8)
        VPS_CASE(9) VPS_CASE(10) VPS_CASE(11) VPS_CASE(12) VPS_CASE(13) VPS_CASE(14) VPS_CASE(15) VPS_CASE(16)
#undef VPS_CASE
    default: return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", c.K, VPS_MAX_K);
    }
}

}  // namespace

// stage 1 of a fixed-order sum of n partials: block b adds the range
// [b n / nb, (b+1) n / nb) (the ranges depend only on n and nb)
__global__ __launch_bounds__(256) void k_reduce_lnl_stage(const double* __restrict__ part, int64_t n, double* __restrict__ out)
{
    __shared__ double sh[256];
    const int64_t b0 = (int64_t)blockIdx.x * n / gridDim.x, b1 = (int64_t)(blockIdx.x + 1) * n / gridDim.x;
    double s = 0;
    for (int64_t x = b0 + threadIdx.x; x < b1; x += 256) s += part[x];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

int launch_reduce_lnl_parts(const double* part, int64_t n, double* stage, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_lnl_stage, dim3(kLnlStage1), dim3(256), 0, st, part, n, stage);
    VPS_HIP(hipGetLastError());
    return launch_reduce_lnl(stage, kLnlStage1, out, st);
}

int launch_reduce_lnl(const double* part, int64_t n, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_lnl, dim3(1), dim3(256), 0, st, part, n, out);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

namespace {

template <int K>
int lik_blocks_per_cu(const DeviceDataset& ds, int* bi, int* gen)
{
    const void* f = ds.maxploidy == 2 ? (const void*)k_lik_bi<K, 2, 2>
                    : ds.maxploidy == 4 ? (const void*)k_lik_bi<K, 2, 4> : (const void*)k_lik_bi<K, 2, 0>;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(bi, f, kLikThreads, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(gen, (const void*)k_lik_gen<K, 2>, kLikThreads, 0);
    if (e != hipSuccess) return fail(VPS_ERR_DEVICE, "occupancy query: %s", hipGetErrorString(e));
    *bi = std::max(*bi, 1);
    *gen = std::max(*gen, 1);
    return VPS_OK;
}

// tiles per range for one launch of `ntiles` equal tiles: the fewest, from 4
// to 16, whose blocks all fit in one round of resident blocks (a block that
// has to wait for a second round then runs alone, latency-bound); when no
// such count exists, 8 (measured: many rounds of short blocks balance better
// than one round of long ones)
int tiles_per_range(int64_t ntiles, int64_t gy, int64_t cap)
{
    if (ntiles <= 0) return 1;
    for (int t = 4; t <= 16; ++t)
        if ((ntiles + t - 1) / t * gy <= cap) return (int)std::min<int64_t>(t, ntiles);
    return (int)std::min<int64_t>(8, ntiles);
}

}  // namespace

// Tiles per locus range of the likelihood passes, for the biallelic and the
// multiallelic launch (blocks = one range x 256 individuals).  Resident
// block counts are those of the Metropolis pass, the heaviest.
int choose_lik_tpr(const DeviceDataset& ds, int K, int device, int32_t* tpr)
{
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
    int bpc = 1, gpc = 1, rc = VPS_OK;
    switch (K) {
#define VPS_CASE(k) case k: rc = lik_blocks_per_cu<k>(ds, &bpc, &gpc); break;
        VPS_CASE(1) VPS_CASE(2) VPS_CASE(3) VPS_CASE(4) VPS_CASE(5) VPS_CASE(6) VPS_CASE(7) VPS_CASE(8)
        VPS_CASE(9) VPS_CASE(10) VPS_CASE(11) VPS_CASE(12) VPS_CASE(13) VPS_CASE(14) VPS_CASE(15) VPS_CASE(16)
#undef VPS_CASE
    default: return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", K, VPS_MAX_K);
    }
    if (rc) return rc;
    const int64_t gy = (ds.N + kLikThreads - 1) / kLikThreads;
    tpr[0] = tiles_per_range(ds.nbi, gy, (int64_t)cus * bpc);
    tpr[1] = tiles_per_range(ds.ngen, gy, (int64_t)cus * gpc);
    return VPS_OK;
}

int lik_part_count(const DeviceDataset& ds, int mode, int K, const int32_t* tpr)
{
    const LikGrid g = lik_grid(ds, tpr);
    const int64_t nr = (int64_t)g.nr_bi + g.nr_gen;
    if (mode <= 1 || mode == 5) return (int)(nr * g.gy);
    return (int)(nr * ds.N * (mode == 3 ? K : 2));   // modes 2, 4, 6: ratios, then current-Q partials
}

int launch_lik_pass(const DeviceDataset& ds, const ChainDev& c0, int mode, hipStream_t st, double* lnl_out)
{
    const LikGrid g = lik_grid(ds, c0.lik_tpr);
    const int nranges = g.nr_bi + g.nr_gen;
    // the Metropolis pass also keeps its current-Q partials, after the ratios
    ChainDev c = c0;
    c.part_cur = c0.part + (int64_t)nranges * ds.N;
    int rc;
    switch (mode) {
    case 0: rc = c.precision == VPS_PREC_FAST32 ? launch_lik_mode<5>(ds, c, st, g) : launch_lik_mode<0>(ds, c, st, g); break;
    case 1: rc = launch_lik_mode<1>(ds, c, st, g); break;
    case 2:
        rc = c.precision == VPS_PREC_MIXED    ? launch_lik_mode<4>(ds, c, st, g)
             : c.precision == VPS_PREC_FAST32 ? launch_lik_mode<6>(ds, c, st, g)
                                              : launch_lik_mode<2>(ds, c, st, g);
        break;
    case 3: rc = launch_lik_mode<3>(ds, c, st, g); break;
    default: return fail(VPS_ERR_INVALID, "likelihood mode %d", mode);
    }
    if (rc) return rc;
    if (mode <= 1) {
        hipLaunchKernelGGL(k_reduce_lnl, dim3(1), dim3(256), 0, st, c.part, (int64_t)nranges * g.gy,
                           lnl_out ? lnl_out : c.scal + 1);
    } else {
        const int NA = mode == 3 ? c.K : 1;
        const int64_t n = (int64_t)ds.N * NA;
        double* dst = mode == 2 ? c.bufN1 : c.bufNK1;
        hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)((n + 63) / 64)), dim3(64 * kRowSplit), 0, st, ds, c, nranges, NA,
                           (mode == 3 && c.locpriori) ? 1 : 0, dst, mode == 2 ? (const double*)c.part_cur : nullptr,
                           mode == 2 ? c.bufN2 : nullptr);
    }
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

}  // namespace vps

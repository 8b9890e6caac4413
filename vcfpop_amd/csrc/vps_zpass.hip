// vps_zpass.hip — allele-copy origin sampling (UpdateZAdmix) and the
// binary-Q count pass (UpdateZNoAdmix) on gfx950.
//
// Reference semantics (src/structureAVX.cpp:788-981, the default x86 path):
// for every sorted locus l one stream RNG(seed + m*L + l, UPDATEZ) draws
// exactly one number per (individual i ascending, slot a < maxploidy),
// including slots that are missing or beyond the genotype's ploidy.  For a
// typed slot with allele x the weights are w_k = P[k][l][x] * Q[i][k] +
// MIN_FREQ, s = w_0 + ... + w_{K-1}, t = (U - 1) * s with U in [1,2) built
// from the draw's 52 low bits, and the cluster is the first k with t < w_k
// after subtracting w_0..w_{k-1} (K-1 if none), src/mathAVX.cpp:73-122.
//
// Mapping: one wavefront = one 64-locus tile x one chunk of individuals.
// Lane j owns the tile's j-th locus and its stream; a wavefront that starts at
// individual i0 jumps its stream ahead by i0 * maxploidy steps with the
// GF(2) table of its chunk, so a locus is split over G wavefronts without
// changing any draw.  Per slot:
//   * "first k with t_k < w_k" is evaluated as sign tests of the running
//     remainder: t_k < w_k  <=>  fl(t_k - w_k) < 0 for finite doubles, and
//     once negative the remainder stays negative, so
//     C_k = [t_{k+1} < 0] = [k* <= k]  (cumulative indicator, k < K-1).
//   * Ni: each lane accumulates C_k per allele in registers;
//     Ni[k] = C_k - C_{k-1} after the chunk (linear in the counts), written
//     to the chunk's partial table, which k_reduce_ni sums.
//   * Mi: per k the wavefront ballots C_k over typed lanes and popcounts
//     it on the scalar unit; per individual the counts go to the
//     wavefront's LDS row, and the block adds its wavefronts' rows (same
//     individuals, other tiles) to Mi once at the end.
//   * Only clusters k < K-1 are counted: every typed copy lands in one
//     cluster, so after the pass Ni[K-1] = copies of the allele - the rest
//     and Mi[i][K-1] = vt_i - the rest (k_reduce_ni).
#include <algorithm>
#include <type_traits>
#include "vps_chain.h"

namespace vps {

// U - 1 for U in [1, 2) built from a draw.  The exponent word is set with
// one bitfield insert whose mask and exponent bits sit in registers (a
// 3-operand VALU instruction takes no literal here); the registers are set
// once per kernel, outside the loops.
struct UnitBits {
    uint32_t m52, one64, m23, one32;
    __device__ __forceinline__ void init()
    {
        asm("s_mov_b32 %0, 0xfffff" : "=s"(m52));
        asm("v_mov_b32 %0, 0x3ff00000" : "=v"(one64));
        asm("s_mov_b32 %0, 0x7fffff" : "=s"(m23));
        asm("v_mov_b32 %0, 0x3f800000" : "=v"(one32));
    }
};
__device__ __forceinline__ uint32_t bfi(uint32_t mask, uint32_t a, uint32_t b)
{
    uint32_t d;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(d) : "s"(mask), "v"(a), "v"(b));
    return d;
}

template <typename Real> __device__ __forceinline__ Real unit_minus1(uint64_t r, const UnitBits& ub);
// FP64 weights: U keeps the 52 low bits of the draw
template <> __device__ __forceinline__ double unit_minus1<double>(uint64_t r, const UnitBits& ub)
{
    const uint32_t hi = bfi(ub.m52, (uint32_t)(r >> 32), ub.one64);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint32_t)r) - 1.0;
}
// FP32 weights: U keeps the 23 high bits of the 52-bit mantissa
template <> __device__ __forceinline__ float unit_minus1<float>(uint64_t r, const UnitBits& ub)
{
    return __builtin_bit_cast(float, bfi(ub.m23, (uint32_t)(r >> 29), ub.one32)) - 1.0f;
}
// FP32 weights, and the slot's lane mask folded into the draw: an untyped
// slot's U takes the exponent of 2^127 instead of 2^0, so its scaled draw
// lies above every running sum of the weights (or is +inf) and no cluster
// test passes on that lane; the counts then need no typed-lane mask
__device__ __forceinline__ float unit_minus1_typed(uint64_t r, const UnitBits& ub, bool typed)
{
    return __builtin_bit_cast(float, bfi(ub.m23, (uint32_t)(r >> 29), typed ? ub.one32 : 0x7f000000u)) - 1.0f;
}
// REAL = float with -g_fastsingle: U from the low 23 bits of an xorshift96
// output (RNGAVX<float>::Poly, src/mathAVX.cpp:357-365)
template <typename Real> __device__ __forceinline__ Real unit_minus1(uint32_t r, const UnitBits& ub);
template <> __device__ __forceinline__ float unit_minus1<float>(uint32_t r, const UnitBits& ub)
{
    return __builtin_bit_cast(float, bfi(ub.m23, r, ub.one32)) - 1.0f;
}

// the mixed-precision mode (FP32 weights, RNG<double> streams): this build's
// own mode, free of the reference's rounding; its cluster test compares the
// scaled draw with the running sums of the weights (t < w_0 + ... + w_k, the
// reference's "first k with t - w_0 - ... - w_{k-1} < w_k" in exact
// arithmetic), one compare per cluster instead of a subtraction and a compare
template <typename Real, bool X96>
constexpr bool z_mixed() { return std::is_same<Real, float>::value && !X96; }

// a slot's cluster weight w = P Q + eps: a product and a sum rounded
// separately, as the reference forms it, except with FP32 weights in the
// mixed-precision mode (this build's own mode, with no reference rounding to
// keep), where one fused multiply-add forms it
template <typename Real, bool X96>
__device__ __forceinline__ Real z_weight(Real p, Real q, Real eps)
{
    if constexpr (z_mixed<Real, X96>()) return __builtin_fmaf((float)p, (float)q, (float)eps);
    else return p * q + eps;
}

// the Z pass's per-locus streams: RNG<double> (xorshift128+), or RNG<float>
// (xorshift96) for REAL = float with -g_fastsingle
template <bool X96> struct ZStream { using T = Xs128; };
template <> struct ZStream<true> { using T = Xs96; };

__device__ __forceinline__ uint64_t ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }

// v_writelane with a lane index in a register: the compiler's intrinsic
// (it routes the index through M0), which this HIP version declares no
// builtin for
__device__ int writelane_i32(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

constexpr int kWavesMax = 8;   // wavefronts per block of the Z kernels (BI: 4; GEN: up to 8)
constexpr int kXcds = 8;       // XCDs of the device; blocks are dealt to them round-robin

// Flush the K per-individual cluster counts of one wavefront (cumulative
// form, wave-uniform) into the block's LDS row, two counts per 32-bit word
// (cluster 2j low half, 2j+1 high half; a block adds at most 4 waves x 64
// loci x 10 slots < 65536 per count).  `val` keeps its other lanes between
// calls: only lanes < (K+1)/2 are written and added.
template <int K>
__device__ __forceinline__ void flush_mi(uint32_t* row, const int (&cum)[K], int lane, uint32_t& val)
{
#pragma unroll
    for (int k = 0; k < K; k += 2) {
        const uint32_t lo = (uint32_t)(cum[k] - (k ? cum[k - 1] : 0));
        const uint32_t hi = k + 1 < K ? (uint32_t)(cum[k + 1] - cum[k]) : 0u;
        const uint32_t pk = lo | (hi << 16);
        asm volatile("v_writelane_b32 %0, %1, %2" : "+v"(val) : "s"(pk), "i"(k / 2));
    }
    if (lane < (K + 1) / 2) atomicAdd(row + lane, val);
}

// add a block's packed per-individual rows to Mi
template <int K>
__device__ __forceinline__ void flush_mi_rows(const uint32_t* mi_lds, int nci, int32_t* __restrict__ Mi_blk)
{
    constexpr int KH = (K + 1) / 2;
    for (int x = threadIdx.x; x < nci * K; x += blockDim.x) {
        const int i = x / K, k = x - i * K;
        const int v = (int)((mi_lds[i * KH + (k >> 1)] >> ((k & 1) * 16)) & 0xFFFFu);
        if (v) atomicAdd(Mi_blk + x, v);
    }
}

// ---------------------------------------------------------------------------
// BI tiles: <= 2 alleles, <= 64 genotypes per locus.  A genotype decodes to
// one byte (typed ploidy | dosage of allele 1 << 4; slot a holds allele
// (a >= ploidy - dosage)): from one 64-bit register table per lane when the
// tile's ids have <= 3 bits, else from the lane's column of the tile's byte
// table, copied into LDS.  The tile's P values sit in LDS as
// [k][allele][lane] values, so selecting a slot's allele is one LDS read
// per cluster instead of two register selects.  Per-lane cumulative cluster
// counts are packed: low 16 bits all typed copies, high 16 bits the copies
// of allele 1 (a chunk has <= 512 individuals x 10 slots < 65536).
// MPT > 0 fixes maxploidy at compile time (unrolled slot loop).
// ---------------------------------------------------------------------------

// a genotype byte recoded as (allele-1 flag of slot a) << a | ploidy << 4,
// a < 4, so a slot's allele is one bit extract instead of a compare and
// select, and slot a is typed iff the byte is >= (a + 1) * 16 (tables of
// maximum ploidy 2 and 4)
__device__ __forceinline__ uint32_t slot_flags(uint32_t e)
{
    const int v = (int)(e & 15u), h = v - (int)((e >> 4) & 15u);
    uint32_t f = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) f |= (a >= h ? 1u : 0u) << a;
    return f | ((uint32_t)v << 4);
}

// Per-individual Mi rows of the BI kernel: one row per (wavefront,
// individual) holding the counts of clusters k < K-1 (the last cluster's
// count is vt_i minus the others, set after the pass), FB-bit fields, FPW
// per 32-bit word.  A wavefront adds at most 64 loci x MPT slots per count:
// 128 < 2^8 for diploid tables, 256 < 2^9 for tetraploid ones, 640 < 2^10
// otherwise.
template <int K, int MPT>
struct MiPack {
    static constexpr int KC = K > 1 ? K - 1 : 1;     // counted clusters
    static constexpr int FB = MPT == 2 ? 8 : MPT == 4 ? 9 : 10;
    static constexpr int FPW = 32 / FB;
    static constexpr int MW = (KC + FPW - 1) / FPW;   // words per row
};
template <int K, int MPT>
__host__ __device__ constexpr int mi_row_words() { return MiPack<K, MPT>::MW; }

// add a BI block's packed per-individual rows (nwv wavefronts' rows of nci
// individuals each) to Mi
template <int K, int MPT>
__device__ __forceinline__ void flush_bi_rows(const uint32_t* mi_lds, int nci, int nwv, int32_t* __restrict__ Mi_blk)
{
    using MP = MiPack<K, MPT>;
    if (K < 2) return;
    for (int x = threadIdx.x; x < nci * (K - 1); x += blockDim.x) {
        const int i = x / (K - 1), k = x - i * (K - 1);
        int v = 0;
        for (int w = 0; w < nwv; ++w)
            v += (int)((mi_lds[((size_t)w * nci + i) * MP::MW + k / MP::FPW] >> ((k % MP::FPW) * MP::FB)) &
                       ((1u << MP::FB) - 1u));
        if (v) atomicAdd(Mi_blk + i * K + k, v);
    }
}

// one wavefront's chunk of one BI tile; BIG: ids of more than 3 bits, the
// genotype bytes come from the lane's LDS column `tab_lds`
// fixed-order sum over the wavefront (every lane ends with the sum)
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// REC: after individual i, close the lanes' lnL accumulators at the end of
// each group of kLnlGroup individuals (or of the chunk) into the partial of
// (tile, group)
__device__ __forceinline__ void lnl_group_end(FrexpAcc& acc, int i, int i1, int N, int64_t tile, int lane,
                                              double* __restrict__ zlik)
{
    if ((i + 1) % kLnlGroup == 0 || i + 1 == i1) {
        const double v = wave_sum(acc.close());
        if (lane == 0) zlik[tile * ((N + kLnlGroup - 1) / kLnlGroup) + i / kLnlGroup] = v;
        acc.open();
    }
}

template <int K, typename Real, int MPT, bool BIG, bool REC, bool X96>
__device__ __forceinline__ void z_bi_tile(const DeviceDataset& ds, const TileDesc& t, const Real* __restrict__ Q,
                                          uint64_t seed, int m, int64_t l, bool valid, int afo, int k2, int lane,
                                          int c, int i0, int i1, const Real* pl0, uint8_t* tab_lds,
                                          uint32_t* mi_lds, int32_t* __restrict__ Ni_part,
                                          const uint32_t* __restrict__ jump, int64_t tile, double* __restrict__ zlik)
{
    const int64_t KT = ds.KT;
    uint64_t tab = 0;
    if (BIG) {
        // each lane copies (and reads back) only its own column
        const int rows = 1 << t.bits;
        const uint8_t* src = ds.btab + t.dbase + lane;
        for (int g = 0; g < rows; ++g) {
            uint32_t e = src[(int64_t)g * kTileLoci];
            if (MPT == 2 || MPT == 4) e = slot_flags(e);
            tab_lds[g * kTileLoci + lane] = (uint8_t)e;
        }
    } else {
        tab = valid ? ds.bitab[l] : 0ull;   // invalid lane: every id is untyped
        if (MPT == 2 || MPT == 4) {
            uint64_t st = 0;
#pragma unroll
            for (int g = 0; g < 8; ++g) st |= (uint64_t)slot_flags((uint32_t)(tab >> (8 * g)) & 0xFFu) << (8 * g);
            tab = st;
        }
    }
    typename ZStream<X96>::T rng;
    rng.seed(seed + (uint64_t)m * (uint64_t)ds.L + (uint64_t)l, kSaltUpdateZ);
    if (c > 0) jump_apply(jump + (size_t)c * kJumpWords, rng);

    // cumulative counts C_k, k < K-1 (C_{K-1} = all typed copies is
    // fixed by the data: the last cluster's counts come from ncopies)
    uint32_t nc[K > 1 ? K - 1 : 1];
#pragma unroll
    for (int k = 0; k < K - 1; ++k) nc[k] = 0;
    using MPK = MiPack<K, MPT>;
    // lane x of mrow[] holds the packed row of individual i0 + 64 g + x of
    // the current group g of 64 individuals; a group's rows go to LDS at once
    uint32_t mrow[MPK::MW];
#pragma unroll
    for (int x = 0; x < MPK::MW; ++x) mrow[x] = 0;
    const Real eps = (Real)kMinFreq;
    const int bits = t.bits, ipw = t.ipw;
    // ids: word by word, each id a bitfield extract at a uniform offset; the
    // next word's load is issued before the current word's ids (clamped to
    // the chunk's last word, so every load is in bounds)
    const int w_first = i0 / ipw, w_last = (i1 - 1) / ipw;
    const uint32_t* wp = ds.words + t.wbase + lane;
    uint32_t nxt = wp[(int64_t)w_first * kTileLoci];
    // the extract's width in a vector register, set once (the extract
    // takes its offset from a scalar register, and one scalar operand is
    // the limit)
    uint32_t vbits;
    asm("v_mov_b32 %0, %1" : "=v"(vbits) : "s"(bits));
    const int MP = MPT > 0 ? MPT : ds.maxploidy;
    UnitBits ub;
    ub.init();
    // REC: the lane's Record lnL, log f over its typed copies with
    // f = sum_k P[k][l][x] Q[i][k] (the products the weights start from)
    FrexpAcc acc;
    acc.open();

    for (int w = w_first; w <= w_last; ++w) {
        const uint32_t word = nxt;
        nxt = wp[(int64_t)min(w + 1, w_last) * kTileLoci];
        const int ia = max(i0, w * ipw), ib = min(i1, (w + 1) * ipw);
        for (int i = ia, sh = (ia - w * ipw) * bits; i < ib; ++i, sh += bits) {
            const uint32_t gid = __builtin_amdgcn_ubfe(word, (uint32_t)sh, vbits);
            // register table: a byte permute picks byte gid of the 8-byte table
            // (only the low byte of desc is read)
            const uint32_t desc = BIG ? (uint32_t)tab_lds[gid * kTileLoci + lane]
                                      : __builtin_amdgcn_perm((uint32_t)(tab >> 32), (uint32_t)tab, gid);
            const uint32_t db = desc & 0xFFu;
            const int v = (int)(desc & 15u);
            const int h = v - (int)((desc >> 4) & 15u);   // the generic path's dosage form
            Real q[K];
#pragma unroll
            for (int k = 0; k < K; ++k) q[k] = Q[(int64_t)i * K + k];
            int cum[K];
#pragma unroll
            for (int k = 0; k < K; ++k) cum[k] = 0;

            double pr = 1.0;   // REC, diploid: the individual's two factors, charged as one pair
            constexpr int kSlotUnroll = MPT > 0 ? MPT : 1;
#pragma unroll kSlotUnroll
            for (int a = 0; a < MP; ++a) {
                const auto r = rng.next();
                const bool typed = (MPT == 2 || MPT == 4) ? db >= (uint32_t)(a + 1) * 16u : a < v;
                const uint64_t tmask = ballot(typed);
                // a slot no lane types (an individual below maxploidy) only
                // consumes its draw; diploid-only tiles keep straight-line code
                if (MPT != 2 && tmask == 0) continue;
                // allele of slot a: a flag bit on the recoded tables, a >= h otherwise
                const uint32_t al1b = (MPT == 2 || MPT == 4) ? ((desc >> a) & 1u) : (a >= h ? 1u : 0u);
                const Real* pa = pl0 + al1b * kTileLoci;
                Real wk[K];
                Real f = 0;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (REC) {
                        const Real pq = pa[k * 2 * kTileLoci] * q[k];
                        wk[k] = pq + eps;
                        f = k ? f + pq : pq;
                    } else {
                        wk[k] = z_weight<Real, X96>(pa[k * 2 * kTileLoci], q[k], eps);
                    }
                }
                if (REC) {
                    const double fc = typed ? (double)f : 1.0;
                    if (MPT == 2) pr = pr * fc;
                    else acc.charge(fc);
                }
                // the running sums of the weights; s is the last
                Real cs[K];
                cs[0] = wk[0];
#pragma unroll
                for (int k = 1; k < K; ++k) cs[k] = cs[k - 1] + wk[k];
                const Real s = cs[K - 1];
                // mixed precision: untyped lanes fail every cluster test
                constexpr bool TYPED_U = z_mixed<Real, X96>() && !REC;
                Real tt;
                if constexpr (TYPED_U) tt = unit_minus1_typed(r, ub, typed) * s;
                else tt = unit_minus1<Real>(r, ub) * s;
                const uint64_t cmask = TYPED_U ? ~0ull : tmask;
                const uint32_t inc = (al1b << 16) | 1u;
#pragma unroll
                for (int k = 0; k < K - 1; ++k) {
                    bool neg;
                    if constexpr (TYPED_U) {
                        neg = tt < cs[k];
                    } else {
                        tt = tt - wk[k];
                        neg = tt < (Real)0;
                    }
                    cum[k] += __builtin_popcountll(ballot(neg) & cmask);
                    nc[k] += (neg && (TYPED_U || typed)) ? inc : 0u;
                }
            }
            if (REC && MPT == 2) acc.charge(pr);
            if (REC) lnl_group_end(acc, i, i1, ds.N, tile, lane, zlik);
            // the individual's counts of clusters k < K-1, packed, into lane
            // (i - i0) % 64 of mrow
            const int rr = i - i0;
#pragma unroll
            for (int x = 0; x < MPK::MW; ++x) {
                uint32_t pk = 0;
#pragma unroll
                for (int f = 0; f < MPK::FPW; ++f) {
                    const int k = x * MPK::FPW + f;
                    if (k < K - 1) pk |= (uint32_t)(cum[k] - (k ? cum[k - 1] : 0)) << (f * MPK::FB);
                }
                mrow[x] = (uint32_t)writelane_i32((int)pk, rr & 63, (int)mrow[x]);
            }
            if ((rr & 63) == 63) {
#pragma unroll
                for (int x = 0; x < MPK::MW; ++x) mi_lds[((rr & ~63) + lane) * MPK::MW + x] = mrow[x];
            }
        }
    }
    // the last group, when partial: its rows are lanes 0 .. (i1 - i0 - 1) % 64
    {
        const int rr = i1 - 1 - i0;
        if ((rr & 63) != 63 && lane <= (rr & 63)) {
#pragma unroll
            for (int x = 0; x < MPK::MW; ++x) mi_lds[((rr & ~63) + lane) * MPK::MW + x] = mrow[x];
        }
    }

    // the lane's counts of clusters k < K-1, into the chunk's partial
    // (plain stores: device-wide atomics would each reach memory); the last
    // cluster's follow from the allele totals after the pass
    if (valid) {
        int32_t* dst = Ni_part + (int64_t)c * K * KT + afo;
        uint32_t prev = 0;
#pragma unroll
        for (int k = 0; k < K - 1; ++k) {
            const uint32_t d = nc[k] - prev;   // per-half differences never borrow (cumulative counts)
            prev = nc[k];
            const int all = (int)(d & 0xFFFFu), one = (int)(d >> 16);
            dst[k * KT] = all - one;
            if (k2 > 1) dst[k * KT + 1] = one;
        }
    }
}

// LDS of the BI kernel: [wave][k][allele][lane] P values, then
// [wave][genotype][lane] table bytes (bi_tab_rows rows; none when every BI
// tile decodes from registers), then the packed Mi rows [wave][individual]
template <int K, typename Real>
__host__ __device__ constexpr size_t bi_p_bytes(int nwv) { return (size_t)nwv * K * 2 * kTileLoci * sizeof(Real); }
__host__ __device__ inline size_t bi_tab_bytes(int nwv, int rows) { return (size_t)nwv * rows * kTileLoci; }
// words per Mi row of the BI kernel launched for this dataset (its MPT)
template <int K>
inline int bi_row_words(const DeviceDataset& ds)
{
    return ds.maxploidy == 2 ? mi_row_words<K, 2>() : ds.maxploidy == 4 ? mi_row_words<K, 4>() : mi_row_words<K, 0>();
}

// Work items are (tile group of nwv tiles, chunk) pairs, w = group * G +
// chunk.  Blocks are dealt to the XCDs round-robin by block index, so block
// b runs item (b % kXcds) * per + b / kXcds (gridDim.x = kXcds * per): each
// XCD takes a contiguous run of items, and the chunks of one tile group run
// back to back on it and read the group's P from its L2 rather than HBM.
// REC: each wavefront also writes the Record lnL partials of its tile's
// groups of individuals in its chunk (lnl_group_end).
template <int K, typename Real, int MPT, bool REC, bool X96>
__global__ __launch_bounds__(256) void k_z_admix_bi(const DeviceDataset ds, const double* __restrict__ Freq,
                                                   const Real* __restrict__ Q, uint64_t seed,
                                                   const int32_t* __restrict__ mp,
                                                   int32_t* __restrict__ Ni_part, int32_t* __restrict__ Mi,
                                                   int Nc, int G, const uint32_t* __restrict__ jump,
                                                   double* __restrict__ zlik)
{
    const int64_t per = gridDim.x / kXcds;
    const int64_t w = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    const int64_t ngroups = ((int64_t)ds.nbi + (blockDim.x >> 6) - 1) / (blockDim.x >> 6);
    if (w >= ngroups * G) return;   // the whole block: grid padding
    const int m = *mp;
    extern __shared__ int32_t lds_i32[];
    const int lane = threadIdx.x & 63;
    // the wavefront index is uniform: tile fields and loop counters then
    // live in scalar registers
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwv = blockDim.x >> 6;
    const int c = (int)(w % G);
    const int i0 = c * Nc;
    const int i1 = min(ds.N, i0 + Nc);
    const int nci = i1 - i0;
    constexpr int MW = mi_row_words<K, MPT>();
    uint8_t* lds_u8 = reinterpret_cast<uint8_t*>(lds_i32);
    Real* p_lds = reinterpret_cast<Real*>(lds_u8) + (size_t)wv * (K * 2 * kTileLoci);
    uint8_t* tab_lds = lds_u8 + bi_p_bytes<K, Real>(nwv) + (size_t)wv * ds.bi_tab_rows * kTileLoci;
    uint32_t* mi_all = reinterpret_cast<uint32_t*>(lds_u8 + bi_p_bytes<K, Real>(nwv) + bi_tab_bytes(nwv, ds.bi_tab_rows));
    for (int x = threadIdx.x; x < nwv * nci * MW; x += blockDim.x) mi_all[x] = 0;
    uint32_t* mi_lds = mi_all + (size_t)wv * nci * MW;   // this wavefront's rows

    const int64_t tix = (w / G) * nwv + wv;
    const bool has_tile = tix < ds.nbi;
    TileDesc t{};
    int afo = 0, k2 = 1;
    bool valid = false;
    int64_t l = 0;
    if (has_tile) {
        t = ds.tiles[ds.bi_list[tix]];
        valid = lane < t.nloc;
        l = ds.tloc[t.lmap + lane];
        afo = ds.afo[l];
        k2 = ds.k2[l];
        const int64_t KT = ds.KT;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double a0 = Freq[k * KT + afo];
            p_lds[(k * 2 + 0) * kTileLoci + lane] = (Real)a0;
            p_lds[(k * 2 + 1) * kTileLoci + lane] = (Real)(k2 > 1 ? Freq[k * KT + afo + 1] : a0);
        }
    }
    __syncthreads();

    if (has_tile) {
        const int64_t tile = ds.bi_list[tix];
        if (t.bits > kBiRegBits)
            z_bi_tile<K, Real, MPT, true, REC, X96>(ds, t, Q, seed, m, l, valid, afo, k2, lane, c, i0, i1, p_lds + lane,
                                                    tab_lds, mi_lds, Ni_part, jump, tile, zlik);
        else
            z_bi_tile<K, Real, MPT, false, REC, X96>(ds, t, Q, seed, m, l, valid, afo, k2, lane, c, i0, i1, p_lds + lane,
                                                     tab_lds, mi_lds, Ni_part, jump, tile, zlik);
    }
    __syncthreads();
    flush_bi_rows<K, MPT>(mi_all, nci, nwv, Mi + (int64_t)i0 * K);
}

// ---------------------------------------------------------------------------
// GEN tiles: any allele count (< 64) and ploidy mix.  A block holds one tile
// and up to 8 consecutive chunks of individuals, one wavefront per chunk, so
// the wavefronts share the tile's tables: with PLDS its P values sit in LDS
// as [k][allele][lane] values (A alleles per lane column), so a slot's K
// weights are K conflict-free LDS reads rather than K scattered HBM/L2
// gathers (without PLDS, P is gathered from Freq); the allele counts of
// clusters k < K-1 are one LDS table [k][allele][lane] that the wavefronts
// add into (or global atomics into the wavefront's chunk partial when the
// table does not fit).  A genotype decodes from a 64-bit descriptor, loaded
// one individual ahead.
//
// LDS: [PLDS] P values, then [ni_in_lds] counts, then packed Mi rows of the
// block's individuals.  The block's counts go to the partial of its first
// chunk; the partials of its other chunks are written as zeros.
// ---------------------------------------------------------------------------
//
// REC: each wavefront also writes the Record lnL partials of the tile's
// groups of individuals in its chunk (lnl_group_end).
template <int K, typename Real, bool PLDS, bool REC, bool X96>
__global__ __launch_bounds__(64 * kWavesMax) void k_z_admix_gen(const DeviceDataset ds, const double* __restrict__ Freq,
                                                    const Real* __restrict__ Q, uint64_t seed,
                                                    const int32_t* __restrict__ mp,
                                                    int32_t* __restrict__ Ni_part, int32_t* __restrict__ Mi,
                                                    int A, int ni_in_lds, int Nc, int G,
                                                    const uint32_t* __restrict__ jump, double* __restrict__ zlik)
{
    const int m = *mp;
    extern __shared__ int32_t lds[];
    const int lane = threadIdx.x & 63;
    // the wavefront index is uniform: tile fields and loop counters then
    // live in scalar registers
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwv = blockDim.x >> 6;
    const int c0 = blockIdx.y * nwv;            // the block's first chunk
    const int c = c0 + wv;                      // this wavefront's chunk
    const int gi0 = c0 * Nc;
    const int gi1 = min(ds.N, (c0 + nwv) * Nc);
    const int i0 = min(ds.N, c * Nc);
    const int i1 = min(ds.N, i0 + Nc);
    constexpr int KH = (K + 1) / 2;
    constexpr int KC = K > 1 ? K - 1 : 1;       // count rows
    const int nP = PLDS ? K * A * kTileLoci : 0;
    const int nni = ni_in_lds ? KC * A * kTileLoci : 0;
    Real* p_lds = reinterpret_cast<Real*>(lds);
    int32_t* ni_lds = reinterpret_cast<int32_t*>(p_lds + nP);
    uint32_t* mi_lds = reinterpret_cast<uint32_t*>(ni_lds + nni);
    for (int x = threadIdx.x; x < nni + (gi1 - gi0) * KH; x += blockDim.x) ni_lds[x] = 0;   // counts, Mi rows

    const TileDesc t = ds.tiles[ds.gen_list[blockIdx.x]];
    const bool valid = lane < t.nloc;
    const int64_t l = ds.tloc[t.lmap + lane];
    const int64_t KT = ds.KT;
    const int afo = ds.afo[l];
    const int k2 = ds.k2[l];
    const double* pl = Freq + afo;
    if (PLDS) {
        // wavefront w fills cluster rows w, w + nwv, ... of every lane column
        for (int k = wv; k < K; k += nwv)
            for (int x = 0; x < A; ++x)
                p_lds[(k * A + x) * kTileLoci + lane] = x < k2 ? (Real)pl[(int64_t)k * KT + x] : (Real)0;
    }
    __syncthreads();

    if (i0 < i1) {
        const uint64_t* gd = ds.gdesc + ds.gbase[l];
        typename ZStream<X96>::T rng;
        rng.seed(seed + (uint64_t)m * (uint64_t)ds.L + (uint64_t)l, kSaltUpdateZ);
        if (c > 0) jump_apply(jump + (size_t)c * kJumpWords, rng);

        const Real eps = (Real)kMinFreq;
        const int bits = t.bits, ipw = t.ipw;
        const uint32_t mask = (1u << bits) - 1u;
        int w = i0 / ipw, j = i0 - w * ipw;
        const uint32_t* wp = ds.words + t.wbase + lane;
        uint32_t word = wp[(int64_t)w * kTileLoci] >> (j * bits);
        // id of individual i, taken in order (the next word is read only
        // while the chunk has individuals left)
        auto take = [&](int i) -> uint32_t {
            const uint32_t g = word & mask;
            word >>= bits;
            if (++j == ipw) {
                j = 0; ++w;
                if (i + 1 < i1) word = wp[(int64_t)w * kTileLoci];
            }
            return g;
        };
        const int MP = ds.maxploidy;
        int32_t* nig = Ni_part + (int64_t)c * K * KT + afo;
        uint32_t mval = 0;

        UnitBits ub;
        ub.init();
        FrexpAcc acc;
        acc.open();
        uint64_t desc = valid ? gd[take(i0)] : 0ull;
        for (int i = i0; i < i1; ++i) {
            const uint64_t cur = desc;
            if (i + 1 < i1) {
                const uint32_t g = take(i + 1);
                desc = valid ? gd[g] : 0ull;
            }
            const int v = (int)(cur & 15u);
            Real q[K];
#pragma unroll
            for (int k = 0; k < K; ++k) q[k] = Q[(int64_t)i * K + k];
            int cum[K];
#pragma unroll
            for (int k = 0; k < K; ++k) cum[k] = 0;

            for (int a = 0; a < MP; ++a) {
                const auto r = rng.next();
                const bool typed = a < v;
                const uint64_t tmask = ballot(typed);
                if (tmask == 0) continue;   // no lane types this slot: the draw only
                const int al = a < VPS_MAX_PLOIDY ? (int)((cur >> (4 + 6 * a)) & 63u) : 0;
                const int x = typed ? al : 0;
                Real wk[K];
                Real f = 0;
                const Real* px = p_lds + x * kTileLoci + lane;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const Real pk = PLDS ? px[k * A * kTileLoci] : (Real)pl[k * KT + x];
                    if (REC) {
                        const Real pq = pk * q[k];
                        wk[k] = pq + eps;
                        f = k ? f + pq : pq;
                    } else {
                        wk[k] = z_weight<Real, X96>(pk, q[k], eps);
                    }
                }
                if (REC) acc.charge(typed ? (double)f : 1.0);
                Real cs[K];
                cs[0] = wk[0];
#pragma unroll
                for (int k = 1; k < K; ++k) cs[k] = cs[k - 1] + wk[k];
                const Real s = cs[K - 1];
                // mixed precision: untyped lanes fail every cluster test
                constexpr bool TYPED_U = z_mixed<Real, X96>() && !REC;
                Real tt;
                if constexpr (TYPED_U) tt = unit_minus1_typed(r, ub, typed) * s;
                else tt = unit_minus1<Real>(r, ub) * s;
                const uint64_t cmask = TYPED_U ? ~0ull : tmask;
                int nneg = 0;
#pragma unroll
                for (int k = 0; k < K - 1; ++k) {
                    bool neg;
                    if constexpr (TYPED_U) {
                        neg = tt < cs[k];
                    } else {
                        tt = tt - wk[k];
                        neg = tt < (Real)0;
                    }
                    cum[k] += __builtin_popcountll(ballot(neg) & cmask);
                    nneg += neg ? 1 : 0;
                }
                cum[K - 1] += __builtin_popcountll(tmask);
                // the last cluster's counts follow from the typed copies
                const int ks = K - 1 - nneg;
                if (typed && ks < K - 1) {
                    if (ni_in_lds) atomicAdd(ni_lds + (ks * A + x) * kTileLoci + lane, 1);
                    else atomicAdd(nig + (int64_t)ks * KT + x, 1);
                }
            }
            flush_mi<K>(mi_lds + (i - gi0) * KH, cum, lane, mval);
            if (REC) lnl_group_end(acc, i, i1, ds.N, ds.gen_list[blockIdx.x], lane, zlik);
        }
    }
    __syncthreads();
    if (ni_in_lds && valid) {
        for (int k = wv; k < K - 1; k += nwv) {
            int32_t* dst = Ni_part + ((int64_t)c0 * K + k) * KT + afo;
            for (int x = 0; x < k2; ++x) dst[x] = ni_lds[(k * A + x) * kTileLoci + lane];
            for (int cc = 1; cc < nwv && c0 + cc < G; ++cc)
                for (int x = 0; x < k2; ++x) dst[(int64_t)cc * K * KT + x] = 0;
        }
    }
    flush_mi_rows<K>(mi_lds, gi1 - gi0, Mi + (int64_t)gi0 * K);
}

// ---------------------------------------------------------------------------
// UpdateZNoAdmix (src/structureAVX.cpp:432-552): Ni[Z_i][l][x] += 1 per
// typed copy; Mi[i][Z_i] = vt_i.  Lanes = loci, counts in lane-owned LDS
// columns, each chunk writes its own Ni partial (in_lds = 0: the table does
// not fit in LDS; counts go to the partials by global atomics, cleared first).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_z_noadmix(const DeviceDataset ds, const uint16_t* __restrict__ Z, int K,
                                                 int32_t* __restrict__ Ni_part, int A, int list_kind, int Nc,
                                                 int in_lds)
{
    extern __shared__ int32_t ni_lds[];   // [K][A][64]
    const int lane = threadIdx.x;
    const int c = blockIdx.y;
    const int i0 = c * Nc;
    const int i1 = min(ds.N, i0 + Nc);
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
    const int bits = t.bits, ipw = t.ipw;
    const uint32_t mask = (1u << bits) - 1u;
    int w = i0 / ipw, j = i0 - w * ipw;
    const uint32_t* wp = ds.words + t.wbase + lane;
    uint32_t word = wp[(int64_t)w * kTileLoci] >> (j * bits);
    for (int i = i0; i < i1; ++i) {
        const uint32_t gid = word & mask;
        word >>= bits;
        if (++j == ipw) {
            j = 0; ++w;
            if (i + 1 < i1) word = wp[(int64_t)w * kTileLoci];
        }
        const int zk = Z[i];
        int32_t* col = ni_lds + (zk * A) * kTileLoci + lane;
        int32_t* gcol = Ni_part + ((int64_t)c * K + zk) * ds.KT + afo;
        if (t.kind == kKindBi) {
            const uint32_t desc = bits > kBiRegBits ? (uint32_t)btab[gid * kTileLoci] : (uint32_t)(tab >> (gid * 8)) & 0xFFu;
            const int v = (int)(desc & 15u), cc = (int)(desc >> 4);
            if (in_lds) {
                col[0] += v - cc;
                col[kTileLoci] += cc;
            } else if (v) {
                atomicAdd(gcol, v - cc);
                if (cc) atomicAdd(gcol + 1, cc);
            }
        } else {
            const uint64_t desc = valid ? gd[gid] : 0ull;
            const int v = (int)(desc & 15u);
            for (int a = 0; a < v; ++a) {
                const int x = (int)((desc >> (4 + 6 * a)) & 63u);
                if (in_lds) col[x * kTileLoci] += 1;
                else atomicAdd(gcol + x, 1);
            }
        }
    }
    if (valid && in_lds) {
        int32_t* dst = Ni_part + (int64_t)c * K * ds.KT + afo;
        for (int k = 0; k < K; ++k)
            for (int a = 0; a < k2; ++a) dst[(int64_t)k * ds.KT + a] = ni_lds[(k * A + a) * kTileLoci + lane];
    }
}

// Ni = sum over chunks of Ni_part
// Ni[k][a] = sum over chunks of the partial counts for k < K-1; every typed
// copy lands in exactly one cluster, so Ni[K-1][a] = ncopies[a] - the rest
// (the Z passes leave the last cluster's partials unwritten).  A block covers
// 64 alleles x 4 cluster rows; the row totals meet in LDS in a fixed order.
constexpr int kReduceRows = 4;
// With vt (after the admixture Z pass, which counts clusters k < K-1 only)
// it also sets Mi[i][K-1] = vt_i - the individual's other counts, one thread
// per individual.
// (advance, when not null: the sweep index, advanced once by the first thread)
__global__ __launch_bounds__(64 * kReduceRows) void k_reduce_ni(const int32_t* __restrict__ Ni_part,
                                                               const int32_t* __restrict__ ncopies,
                                                               int32_t* __restrict__ Ni, int64_t KT, int K, int G,
                                                               int32_t* __restrict__ advance,
                                                               const int32_t* __restrict__ vt, int32_t* __restrict__ Mi,
                                                               int N)
{
    if (advance && blockIdx.x == 0 && threadIdx.x == 0) advance[0] += 1;
    if (vt) {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i < N) {
            int32_t* row = Mi + i * K;
            int32_t t = 0;
            for (int k = 0; k < K - 1; ++k) t += row[k];
            row[K - 1] = vt[i] - t;
        }
    }
    __shared__ int32_t tot[kReduceRows][64];
    const int al = threadIdx.x & 63, kk = threadIdx.x >> 6;
    const int64_t a = (int64_t)blockIdx.x * 64 + al;
    const int64_t n = (int64_t)K * KT;
    int32_t mine = 0;
    if (a < KT) {
        for (int k = kk; k < K - 1; k += kReduceRows) {
            const int32_t* p = Ni_part + k * KT + a;
            int32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            int c = 0;
            for (; c + 4 <= G; c += 4) {
                s0 += p[(int64_t)c * n];
                s1 += p[(int64_t)(c + 1) * n];
                s2 += p[(int64_t)(c + 2) * n];
                s3 += p[(int64_t)(c + 3) * n];
            }
            for (; c < G; ++c) s0 += p[(int64_t)c * n];
            const int32_t s = (s0 + s1) + (s2 + s3);
            Ni[k * KT + a] = s;
            mine += s;
        }
    }
    tot[kk][al] = mine;
    __syncthreads();
    if (kk == 0 && a < KT) {
        int32_t t = 0;
        for (int r = 0; r < kReduceRows; ++r) t += tot[r][al];
        Ni[(int64_t)(K - 1) * KT + a] = ncopies[a] - t;
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) p[x] = v;
}

// Q rounded to single precision once per pass, so that the single-precision
// Z passes read their Q values ready to use (scalar loads, no conversion per
// individual and tile)
__global__ void k_q_to_float(const double* __restrict__ q, float* __restrict__ qf, int64_t n)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) qf[x] = (float)q[x];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {

constexpr int kWavesPerBlock = 4;

// LDS available to one block (gfx950: the CU's 160 KiB)
int max_block_lds()
{
    static int bytes = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        return v;
    }();
    return bytes;
}

// Launch shape of the GEN kernel: one tile x `waves` chunks per block; P
// staged when P, the count table and the Mi rows fit one block's LDS, else
// counts alone in LDS, else counts by global atomics.  The tile's P and
// count tables are shared by the block's waves, so when LDS limits the
// resident blocks, 8 chunks per block keep more waves resident than 4 (the
// shape with more resident waves is taken; fewer when there are fewer chunks).
struct GenPlan {
    int waves;
    bool plds;
    bool ni_in_lds;
    size_t shm;
};

template <int K, typename Real>
GenPlan gen_plan_waves(const DeviceDataset& ds, int Nc, int wv)
{
    const size_t A = (size_t)std::max(ds.maxK, 1);
    const size_t kc = K > 1 ? K - 1 : 1;
    const size_t mi = (size_t)wv * Nc * ((K + 1) / 2) * sizeof(uint32_t);
    const size_t ni = kc * A * kTileLoci * sizeof(int32_t);
    const size_t pb = (size_t)K * A * kTileLoci * sizeof(Real);
    const size_t cap = (size_t)max_block_lds();
    if (pb + ni + mi <= cap) return {wv, true, true, pb + ni + mi};
    if (ni + mi <= cap) return {wv, false, true, ni + mi};
    return {wv, false, false, mi};
}

template <int K, typename Real>
GenPlan gen_plan(const DeviceDataset& ds, int Nc, int G)
{
    const GenPlan g4 = gen_plan_waves<K, Real>(ds, Nc, std::max(1, std::min(kWavesPerBlock, G)));
    if (G < kWavesMax) return g4;
    const GenPlan g8 = gen_plan_waves<K, Real>(ds, Nc, kWavesMax);
    if (g8.plds != g4.plds || g8.ni_in_lds != g4.ni_in_lds) return g4;
    // resident waves per CU as LDS allows (160 KiB per CU), capped by 32 waves
    const size_t cu_lds = 160 * 1024;
    auto resident = [&](const GenPlan& g) {
        const size_t blocks = g.shm ? cu_lds / g.shm : 64;
        return std::min<size_t>(blocks * g.waves, 32);
    };
    return resident(g8) > resident(g4) ? g8 : g4;
}

// the Q values a Z pass of this precision reads
template <typename Real>
const Real* z_q(const ChainDev& c);
template <>
const double* z_q<double>(const ChainDev& c) { return c.Q; }
template <>
const float* z_q<float>(const ChainDev& c) { return c.Qf; }

// nparts (may be null): the number of Record partials written
template <int K, typename Real, bool REC, bool X96>
int launch_z_admix_k(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, int64_t* nparts)
{
    const int nthr = 64 * kWavesPerBlock;
    const int A = std::max(ds.maxK, 1);
    const size_t mi_packed = (size_t)kWavesPerBlock * c.Nc * bi_row_words<K>(ds) * sizeof(uint32_t);
    const GenPlan g = gen_plan<K, Real>(ds, c.Nc, c.G);
    // GEN counts in global atomics add into the chunk partials, which must
    // then be cleared before any tile (of either kind) writes its counts
    if (ds.ngen > 0 && !g.ni_in_lds) {
        int rc = launch_fill_i32(c.Ni_part, (int64_t)c.G * K * ds.KT, 0, st);
        if (rc) return rc;
    }
    if (ds.nbi > 0) {
        // one block per (group of 4 tiles, chunk) item, dealt round-robin to
        // the XCDs; padded to a multiple of kXcds (padding blocks exit)
        const int64_t items = (ds.nbi + kWavesPerBlock - 1) / kWavesPerBlock * (int64_t)c.G;
        const int64_t per = (items + kXcds - 1) / kXcds;
        if (per * kXcds > 0x7FFFFFFF) return fail(VPS_ERR_UNSUPPORTED, "Z pass grid of %lld blocks", (long long)(per * kXcds));
        const dim3 grid((unsigned)(per * kXcds));
        const size_t shm = bi_p_bytes<K, Real>(kWavesPerBlock) + bi_tab_bytes(kWavesPerBlock, ds.bi_tab_rows) + mi_packed;
        if (ds.maxploidy == 2)
            hipLaunchKernelGGL((k_z_admix_bi<K, Real, 2, REC, X96>), grid, dim3(nthr), shm, st, ds, c.Freq, z_q<Real>(c), c.seed, c.mdev,
                               c.Ni_part, c.Mi, c.Nc, c.G, c.jump, c.zlik);
        else if (ds.maxploidy == 4)
            hipLaunchKernelGGL((k_z_admix_bi<K, Real, 4, REC, X96>), grid, dim3(nthr), shm, st, ds, c.Freq, z_q<Real>(c), c.seed, c.mdev,
                               c.Ni_part, c.Mi, c.Nc, c.G, c.jump, c.zlik);
        else
            hipLaunchKernelGGL((k_z_admix_bi<K, Real, 0, REC, X96>), grid, dim3(nthr), shm, st, ds, c.Freq, z_q<Real>(c), c.seed, c.mdev,
                               c.Ni_part, c.Mi, c.Nc, c.G, c.jump, c.zlik);
        VPS_HIP(hipGetLastError());
    }
    if (ds.ngen > 0) {
        dim3 grid((unsigned)ds.ngen, (unsigned)((c.G + g.waves - 1) / g.waves));
        if (g.plds)
            hipLaunchKernelGGL((k_z_admix_gen<K, Real, true, REC, X96>), grid, dim3(64 * g.waves), g.shm, st, ds, c.Freq, z_q<Real>(c),
                               c.seed, c.mdev, c.Ni_part, c.Mi, A, 1, c.Nc, c.G, c.jump, c.zlik);
        else
            hipLaunchKernelGGL((k_z_admix_gen<K, Real, false, REC, X96>), grid, dim3(64 * g.waves), g.shm, st, ds, c.Freq, z_q<Real>(c),
                               c.seed, c.mdev, c.Ni_part, c.Mi, A, g.ni_in_lds ? 1 : 0, c.Nc, c.G, c.jump, c.zlik);
        VPS_HIP(hipGetLastError());
    }
    if (nparts) *nparts = z_lik_parts(ds);
    return VPS_OK;
}

// resident blocks per CU of the BI kernel, and resident GEN wavefronts per
// SIMD (GEN blocks may hold fewer than 4 wavefronts), for chunks of Nc
template <int K, typename Real, bool X96>
int z_occupancy(const DeviceDataset& ds, int Nc, int* bi_blocks, int* gen_waves, int* gen_block_waves)
{
    const size_t mi_packed = (size_t)kWavesPerBlock * Nc * bi_row_words<K>(ds) * sizeof(uint32_t);
    const size_t p_bytes = bi_p_bytes<K, Real>(kWavesPerBlock) + bi_tab_bytes(kWavesPerBlock, ds.bi_tab_rows);
    *bi_blocks = 1;
    hipError_t e = hipSuccess;
    if (ds.maxploidy == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(bi_blocks, (const void*)k_z_admix_bi<K, Real, 2, false, X96>, 256, p_bytes + mi_packed);
    else if (ds.maxploidy == 4)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(bi_blocks, (const void*)k_z_admix_bi<K, Real, 4, false, X96>, 256, p_bytes + mi_packed);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(bi_blocks, (const void*)k_z_admix_bi<K, Real, 0, false, X96>, 256, p_bytes + mi_packed);
    if (e != hipSuccess) return fail(VPS_ERR_DEVICE, "occupancy query: %s", hipGetErrorString(e));
    const GenPlan g = gen_plan<K, Real>(ds, Nc, (ds.N + Nc - 1) / Nc);
    int gb = 1;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&gb, g.plds ? (const void*)k_z_admix_gen<K, Real, true, false, X96>
                                                                 : (const void*)k_z_admix_gen<K, Real, false, false, X96>,
                                                     64 * g.waves, g.shm);
    if (e != hipSuccess) return fail(VPS_ERR_DEVICE, "occupancy query: %s", hipGetErrorString(e));
    *bi_blocks = std::max(*bi_blocks, 1);
    *gen_waves = std::max(1, std::max(gb, 1) * g.waves / kWavesPerBlock);
    *gen_block_waves = g.waves;
    return VPS_OK;
}

template <typename Real, bool X96>
int z_occupancy_k(const DeviceDataset& ds, int K, int Nc, int* b, int* g, int* gw)
{
    switch (K) {
#define VPS_CASE(k) case k: return z_occupancy<k, Real, X96>(ds, Nc, b, g, gw);
        VPS_CASE(1) VPS_CASE(2) VPS_CASE(3) VPS_CASE(4) VPS_CASE(5) VPS_CASE(6) VPS_CASE(7) VPS_CASE(8)
        VPS_CASE(9) VPS_CASE(10) VPS_CASE(11) VPS_CASE(12) VPS_CASE(13) VPS_CASE(14) VPS_CASE(15) VPS_CASE(16)
#undef VPS_CASE
    default: return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", K, VPS_MAX_K);
    }
}

template <typename Real, bool REC, bool X96 = false>
int dispatch_k(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, int64_t* nparts = nullptr)
{
    switch (c.K) {
#define VPS_CASE(k) case k: return launch_z_admix_k<k, Real, REC, X96>(ds, c, st, nparts);
        VPS_CASE(1) VPS_CASE(2) VPS_CASE(3) VPS_CASE(4) VPS_CASE(5) VPS_CASE(6) VPS_CASE(7) VPS_CASE(8)
        VPS_CASE(9) VPS_CASE(10) VPS_CASE(11) VPS_CASE(12) VPS_CASE(13) VPS_CASE(14) VPS_CASE(15) VPS_CASE(16)
#undef VPS_CASE
    default: return fail(VPS_ERR_UNSUPPORTED, "K=%d outside 1..%d", c.K, VPS_MAX_K);
    }
}

// occupancy of the Z kernels a precision mode launches
int z_occupancy_prec(const DeviceDataset& ds, int K, int precision, int Nc, int* b, int* g, int* gw)
{
    if (precision == VPS_PREC_MIXED) return z_occupancy_k<float, false>(ds, K, Nc, b, g, gw);
    if (precision == VPS_PREC_FAST32) return z_occupancy_k<float, true>(ds, K, Nc, b, g, gw);
    return z_occupancy_k<double, false>(ds, K, Nc, b, g, gw);
}

// the count-only Z pass of a precision mode: FP64 weights (FP64, REAL32),
// FP32 weights from the xorshift128+ draws (MIXED) or from xorshift96
// streams (FAST32).  The FP32-weight passes read Q's single-precision copy,
// rounded here first unless qf_fresh (the sweep's UpdateQ wrote both)
int dispatch_counts(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, bool qf_fresh = false)
{
    if (c.precision == VPS_PREC_MIXED || c.precision == VPS_PREC_FAST32) {
        if (!qf_fresh) {
            const int64_t n = (int64_t)ds.N * c.K;
            hipLaunchKernelGGL(k_q_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c.Q, c.Qf, n);
            VPS_HIP(hipGetLastError());
        }
        return c.precision == VPS_PREC_MIXED ? dispatch_k<float, false>(ds, c, st) : dispatch_k<float, false, true>(ds, c, st);
    }
    return dispatch_k<double, false>(ds, c, st);
}

}  // namespace

// Chunking model.  A group of 4 wavefronts (one per SIMD) costs
// (Nc * maxploidy + 8) slot-steps (jump-ahead and set-up ~ 8 slots); a CU
// streams its blocks back to back, so the pass takes ceil(blocks / CUs)
// block-times, stretched when fewer than 3 wavefronts per SIMD are resident
// to hide FP64 latency.  Chunk partial counts add 8 bytes per (chunk, cluster,
// allele) of HBM traffic (written, then reduced).  A slot-step is ~(12 + 10K)
// VALU instructions of 4 cycles at 2.4 GHz.  The chunk count with the
// smallest modelled time wins, among chunks of 16..kMaxChunk individuals.
int choose_chunks(const DeviceDataset& ds, int K, int precision, int device, bool groups, int* Gout)
{
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
    const int N = ds.N;
    const double slot_us = (12.0 + 10.0 * K) * 4.0 / 2400.0;
    const double hbm_bytes_per_us = 5.0e6;
    double best = 1e300;
    int bestG = (N + kMaxChunk - 1) / kMaxChunk;
    const int gmin = std::max(1, (N + kMaxChunk - 1) / kMaxChunk);
    const int gmax = std::max(gmin, std::min(256, (N + 15) / 16));
    int lastNc = -1;
    auto eff = [](int64_t waves_per_simd) { return waves_per_simd >= 3 ? 1.0 : waves_per_simd == 2 ? 0.8 : 0.55; };
    for (int G = gmin; G <= gmax; ++G) {
        const int Nc = chunk_size(N, G, groups);
        if (Nc == lastNc) continue;
        lastNc = Nc;
        const int Gr = (N + Nc - 1) / Nc;
        // occupancy as wavefronts per SIMD (a 4-wave BI block puts one on each)
        int bb = 1, gb = 1, gw = kWavesPerBlock;
        int rc = z_occupancy_prec(ds, K, precision, Nc, &bb, &gb, &gw);
        if (rc) return rc;
        const double work = ((double)Nc * ds.maxploidy + 8.0) * slot_us;
        double t = (double)Gr * K * ds.KT * 8.0 / hbm_bytes_per_us;
        const int64_t nb[2] = {(ds.nbi + kWavesPerBlock - 1) / kWavesPerBlock * (int64_t)Gr,
                               (ds.ngen + kWavesPerBlock - 1) / kWavesPerBlock * (int64_t)Gr};
        const int occ[2] = {bb, gb};
        for (int x = 0; x < 2; ++x) {
            if (!nb[x]) continue;
            const int64_t per_cu = (nb[x] + cus - 1) / cus;
            t += (double)per_cu * work / eff(std::min<int64_t>(per_cu, occ[x]));
        }
        if (t < best * 0.999) { best = t; bestG = Gr; }
    }
    *Gout = bestG;
    return VPS_OK;
}

int launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st)
{
    if (n <= 0) return VPS_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

// the Z pass into the buffers `c` names (the chunk-size tuner passes
// scratch Mi, partials and Ni)
int launch_z_admix_counts(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    int rc = launch_fill_i32(c.Mi, (int64_t)ds.N * c.K, 0, st);
    if (!rc) rc = dispatch_counts(ds, c, st);
    if (!rc) rc = launch_reduce_ni(ds, c, st, false, true);
    return rc;
}

// wavefronts per block of the GEN Z kernel for a chain's chunking
int z_gen_block_waves(const DeviceDataset& ds, int K, int precision, int Nc, int* waves)
{
    int bb = 1, gb = 1;
    return z_occupancy_prec(ds, K, precision, Nc, &bb, &gb, waves);
}

int64_t z_lik_parts(const DeviceDataset& ds)
{
    return ds.ntiles * (((int64_t)ds.N + kLnlGroup - 1) / kLnlGroup);
}

int launch_z_admix(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, bool clear_mi,
                   bool rec, bool advance, bool qf_fresh)
{
    if (rec && advance) return fail(VPS_ERR_INVALID, "Record reads the sweep index after the Z pass");
    if (rec && (c.precision == VPS_PREC_MIXED || c.precision == VPS_PREC_FAST32))
        return fail(VPS_ERR_INVALID, "the Record lnL of the Z pass needs FP64 weights");
    if (rec && !z_lnl_aligned(c))
        return fail(VPS_ERR_INVALID, "the Record lnL of the Z pass needs chunks of whole lnL groups");
    int rc = clear_mi ? launch_fill_i32(c.Mi, (int64_t)ds.N * c.K, 0, st) : VPS_OK;
    if (rc) return rc;
    if (ev0) VPS_HIP(hipEventRecord(ev0, st));
    int64_t nparts = 0;
    rc = rec ? dispatch_k<double, true>(ds, c, st, &nparts) : dispatch_counts(ds, c, st, qf_fresh);
    if (ev1) VPS_HIP(hipEventRecord(ev1, st));
    if (rc) return rc;
    rc = launch_reduce_ni(ds, c, st, advance, true);
    if (!rc && rec) rc = launch_reduce_lnl_parts(c.zlik, nparts, c.zlik + nparts, c.scal + 1, st);
    return rc;
}

int launch_reduce_ni(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, bool advance, bool mi_last)
{
    const int64_t nb = std::max<int64_t>((ds.KT + 63) / 64, mi_last ? (ds.N + 64 * kReduceRows - 1) / (64 * kReduceRows) : 0);
    hipLaunchKernelGGL(k_reduce_ni, dim3((unsigned)nb), dim3(64 * kReduceRows), 0, st, c.Ni_part, ds.ncopies, c.Ni, ds.KT,
                       c.K, c.G, advance ? c.mdev : nullptr, mi_last ? ds.vt : nullptr, c.Mi, ds.N);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

int launch_z_noadmix(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    const int A = ds.maxK > 2 ? ds.maxK : 2;
    size_t shm = (size_t)c.K * A * kTileLoci * sizeof(int32_t);
    const int in_lds = shm <= 64 * 1024;
    if (!in_lds) {
        int rc = launch_fill_i32(c.Ni_part, (int64_t)c.G * c.K * ds.KT, 0, st);
        if (rc) return rc;
        shm = 0;
    }
    if (ds.nbi > 0)
        hipLaunchKernelGGL(k_z_noadmix, dim3((unsigned)ds.nbi, (unsigned)c.G), dim3(64), shm, st, ds, c.Z, c.K, c.Ni_part, A, kKindBi, c.Nc, in_lds);
    if (ds.ngen > 0)
        hipLaunchKernelGGL(k_z_noadmix, dim3((unsigned)ds.ngen, (unsigned)c.G), dim3(64), shm, st, ds, c.Z, c.K, c.Ni_part, A, kKindGen, c.Nc, in_lds);
    VPS_HIP(hipGetLastError());
    return launch_reduce_ni(ds, c, st);
}

}  // namespace vps

namespace vps {

// ---------------------------------------------------------------------------
// Test hook: a plain device copy of known size, used to scale hardware
// traffic counters (bytes moved per counted unit) for the roofline report.
// ---------------------------------------------------------------------------
__global__ void k_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n)
{
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
        dst[x] = src[x];
}

}  // namespace vps

extern "C" int vps_debug_copy(int device, size_t bytes, int iters, double* ms_per_copy)
{
    using namespace vps;
    if (bytes < 16 || iters < 1 || !ms_per_copy) return fail(VPS_ERR_INVALID, "copy arguments");
    VPS_HIP(hipSetDevice(device));
    void *a = nullptr, *b = nullptr;
    VPS_HIP(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) { (void)hipFree(a); return fail(VPS_ERR_NOMEM, "hipMalloc copy buffer"); }
    const int64_t n = (int64_t)(bytes / 16);
    hipEvent_t e0, e1;
    int rc = VPS_OK;
    if (hipMemset(a, 1, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(VPS_ERR_DEVICE, "copy setup");
    if (!rc) {
        hipLaunchKernelGGL(k_copy_u4, dim3(4096), dim3(256), 0, 0, (const uint4*)a, (uint4*)b, n);
        (void)hipEventRecord(e0, 0);
        for (int it = 0; it < iters; ++it)
            hipLaunchKernelGGL(k_copy_u4, dim3(4096), dim3(256), 0, 0, (const uint4*)a, (uint4*)b, n);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail(VPS_ERR_DEVICE, "copy run");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *ms_per_copy = ms / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}


// vps_chain.h — device view of one MCMC chain and kernel entry points.
#pragma once
#include <algorithm>
#include "vps_internal.h"

namespace vps {

// A log accumulator that keeps a product as mantissa in [0.5, 1) times 2^e
// (frexp after each factor): the reference's exponent-stripping log
// accumulator (src/math2.h:716-902) with the strip done every time.  Scaling
// by powers of two is exact, so only the product order differs from the
// reference's products, and sums agree to rounding.
struct FrexpAcc {
    double m;
    int e;
    __device__ void open() { m = 1.0; e = 0; }
    __device__ void charge(double x)
    {
        const double y = m * x;
        e += __builtin_amdgcn_frexp_exp(y);
        m = __builtin_amdgcn_frexp_mant(y);
    }
    __device__ double close() const { return ((double)e + log2(m)) * kLn2Trunc; }
};

// apply a jump table (32 nibbles x 16 entries x 128 bits) to a stream.
// One state word (8 nibbles) per outer step: at most 8 table loads are in
// flight, so the prologue needs no more registers than the sampling loop
// (the loop, not the jump, then sets the kernel's occupancy).
__device__ __forceinline__ void jump_apply(const uint32_t* __restrict__ tab, Xs128& r)
{
    uint32_t w0 = (uint32_t)r.x, w1 = (uint32_t)(r.x >> 32), w2 = (uint32_t)r.y, w3 = (uint32_t)(r.y >> 32);
    uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    const uint4* t4 = reinterpret_cast<const uint4*>(tab);
#pragma unroll 1
    for (int q = 0; q < 4; ++q, t4 += 8 * 16) {
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const uint4 e = t4[p * 16 + ((w0 >> (p * 4)) & 15u)];
            a0 ^= e.x; a1 ^= e.y; a2 ^= e.z; a3 ^= e.w;
        }
        w0 = w1; w1 = w2; w2 = w3;
    }
    r.x = (uint64_t)a0 | ((uint64_t)a1 << 32);
    r.y = (uint64_t)a2 | ((uint64_t)a3 << 32);
}

// the same for an xorshift96 stream (tables of jump_tables96: words 0..2)
__device__ __forceinline__ void jump_apply(const uint32_t* __restrict__ tab, Xs96& r)
{
    uint32_t w0 = r.x, w1 = r.y, w2 = r.z;
    uint32_t a0 = 0, a1 = 0, a2 = 0;
    const uint4* t4 = reinterpret_cast<const uint4*>(tab);
#pragma unroll 1
    for (int q = 0; q < 3; ++q, t4 += 8 * 16) {
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const uint4 e = t4[p * 16 + ((w0 >> (p * 4)) & 15u)];
            a0 ^= e.x; a1 ^= e.y; a2 ^= e.z;
        }
        w0 = w1; w1 = w2;
    }
    r.x = a0; r.y = a1; r.z = a2;
}

// Pointers and scalars of one chain, passed by value to every kernel.
struct ChainDev {
    int32_t K, N, S, maxploidy;
    int64_t L, KT;
    uint64_t seed;
    int32_t admix, locpriori, fmodel, topology, precision;
    int32_t inferalpha, diffalpha, uniformalpha, inferlambda, difflambda, fsame;
    int32_t nadmburnin, rlen;
    int32_t nburnin, nthinning, writelnl, lnl_cap;   // Record schedule (nburnin includes nadmburnin)
    double stdalpha, maxalpha, alphapriora, alphapriorb;
    double stdlambda, maxlambda;
    double maxr, epsr, epseta, epsgamma;
    double pmeanf, pstdf, stdf;
    // state
    double*   Freq;      // [K*KT]
    double*   Q;         // [N*K]
    float*    Qf;        // [N*K] Q rounded to single precision, the single-precision Z passes' copy
    uint16_t* Z;         // [N]
    int32_t*  Ni;        // [K*KT]
    int32_t*  Ni_part;   // [G*K*KT]
    int32_t*  Mi;        // [N*K]
    double*   Alpha;     // [K]
    double*   Lambda;    // [K]
    double*   bufNK1;    // [N*K] Metropolis proposals / no-admix log likelihoods
    double*   bufN1;     // [N]   Metropolis log ratios
    double*   bufN2;     // [N]   Metropolis pass: each individual's log likelihood under its current Q,
                         //       then (after acceptance) under the Q it keeps
    double*   part;      // partial sums of the likelihood passes
    double*   part_cur;  // Metropolis pass only: the current-Q partials (set by the launch)
    double*   zlik;      // [z_lik_parts(G)] Record lnL partials of the FP64 Z pass
    int64_t*  MiSum;     // [N*K]
    double*   FreqM;     // [K*KT]
    double*   FreqA;     // [3*KT]
    double*   Fst;       // [2*K]: F[k], then f[k] = (1-F)/F
    double*   Eta;       // [K]
    double*   Gamma;     // [S*K]
    double*   Di;        // [S*K]
    double*   AlphaLocal;// [S*K]
    double*   SumAlpha;  // [S]
    double*   scal;      // [8]: 0 r, 1 lnL of the last Record pass
    double*   rout;      // [rlen]
    double*   lnltrace;  // [nlnl capacity]
    Rng*      rngbuf;    // [N] streams kept between a draw kernel and its acceptance kernel
    Rng*      frng;      // [1] the UpdateF stream between the kernels of one UpdateF
    double*   fctl;      // [4] UpdateF step: Fo, Fm, out of range, done
    double*   fpart;     // [256] UpdateF block partial sums
    uint32_t* fcnt;      // [1] UpdateF blocks finished (the last one decides; reset by it)
    double*   lpart;     // [2 * max(kLpCap, S*K)] UpdateAlpha/UpdateLambda LogProd partials
    uint32_t* lcnt;      // [1] their blocks finished (reset by the last)
    int32_t*  mdev;      // [1] sweep index m; kernels read it, the sweep's last kernel advances it,
                         // so one captured sweep replays at any m
    int32_t*  mside;     // [1] copy of m for the kernels beside the Z pass (written by UpdateP/Q)
    // individual chunking of the genotype passes: G chunks of Nc individuals;
    // jump[c] advances a locus stream to individual c*Nc
    int32_t   G, Nc;
    int32_t   Gcap;      // the chunk count Ni_part holds
    const uint32_t* jump;
    int32_t   lik_tpr[2];   // tiles per locus range of the likelihood passes (BI, GEN tiles)
    // Metropolis proposals parsed as a stream graph (reference draw streams,
    // global alpha): the stream's first pg_W words, the generator at every
    // 8th word, two successor tables of pg_S states, the path of draw states
    uint64_t* pg_words;
    Xs128*    pg_st8;
    int32_t*  pg_next;
    int32_t*  pg_path;
    int64_t   pg_W, pg_S;
    // reference draw streams: the generator of every thread of the parallel
    // stream parse at its ring offset, prepared one sweep ahead for UpdateP's
    // 128 streams and UpdateQ's stream, [2][kSppStreams][1024] by sweep
    // parity, with the sweep each stream's entry was prepared for
    Xs128*    spp_prep;
    int32_t*  spp_tag;   // [2][kSppStreams]
};
constexpr int kSppStreams = 129;    // UpdateP's 128 streams, then UpdateQ's

// REAL = float (BAYESIAN<float>): P, Q, the proposals and FreqM hold
// single-precision values
__host__ __device__ inline bool real32(const ChainDev& c) { return c.precision >= VPS_PREC_REAL32; }

// partials of the grid-wide LogProds of UpdateAlpha and UpdateLambda (more
// when S*K columns exceed it: one block per column then)
constexpr int kLpCap = 1024;

// likelihood-pass partials: lanes = individuals, one partial per locus range
constexpr int kLikLociPerWave = 1024;

// launchers (vps_kernels.hip / vps_updates.hip); all asynchronous on `st`
// clear_mi = false: the caller has already zeroed Mi on an ordered stream
// rec = true: the FP64 Z pass also sums the Record lnL (mode 0) into c.scal[1]
// advance = true: the pass's last kernel also advances the sweep index
// qf_fresh = true: Qf already holds Q rounded to single precision (the
// sweep's UpdateQ wrote both), so the FP32-weight passes skip the rounding
int launch_z_admix(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1,
                   bool clear_mi = true, bool rec = false, bool advance = false, bool qf_fresh = false);
// number of Record partials the Z pass writes with G chunks
// Record's lnL summed inside the FP64 Z pass: one partial per (tile, group of
// kLnlGroup individuals), each the wavefront sum of its lanes' closed
// accumulators, so the total does not depend on the individual chunking as
// long as chunk boundaries fall on group boundaries (every automatic chunking
// of a chain whose Record sweeps can take this path does: chunk_size); then
// kLnlStage1 fixed ranges and their sum
constexpr int kLnlGroup = 32;
constexpr int kLnlStage1 = 64;
int64_t z_lik_parts(const DeviceDataset& ds);
// individuals per chunk for G chunks of N: balanced, and with `groups`
// rounded up to whole lnL groups when there is more than one chunk
inline int chunk_size(int N, int G, bool groups)
{
    int nc = (N + G - 1) / G;
    if (groups && G > 1) nc = (nc + kLnlGroup - 1) / kLnlGroup * kLnlGroup;
    return std::min(nc, std::max(N, 1));
}
// whether the chain's chunking keeps the Z pass's Record lnL partials whole
inline bool z_lnl_aligned(const ChainDev& c) { return c.G == 1 || c.Nc % kLnlGroup == 0; }
int launch_reduce_lnl_parts(const double* part, int64_t n, double* stage, double* out, hipStream_t st);
// out = the fixed-order sum of n partials (one block)
int launch_reduce_lnl(const double* part, int64_t n, double* out, hipStream_t st);
int launch_z_noadmix(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_z_admix_counts(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
// Ni from the chunk partials (and, mi_last: the Mi column of the last
// cluster, after the admixture Z pass)
int launch_reduce_ni(const DeviceDataset& ds, const ChainDev& c, hipStream_t st, bool advance = false, bool mi_last = false);
// likelihood passes: mode 0 = Record lnL (admix), 1 = Record lnL (no admix),
// 2 = Metropolis log ratio per individual, 3 = no-admix per-(i,k) log likelihood.
// Modes 0/1 write the lnL to lnl_out, or to c.scal[1] when it is null.
int launch_lik_pass(const DeviceDataset& ds, const ChainDev& c, int mode, hipStream_t st, double* lnl_out = nullptr);
int lik_part_count(const DeviceDataset& ds, int mode, int K, const int32_t* tpr);
// tiles per locus range of the likelihood passes for this device: [0]
// biallelic tiles, [1] multiallelic tiles
int choose_lik_tpr(const DeviceDataset& ds, int K, int device, int32_t* tpr);

// The reference topology's sequential Gamma streams, parsed in parallel
// (vps_spp.hip): UpdateP's 128 streams when do_p, and UpdateQ's stream with
// draws of kind qkind (kSppQ*); then the rows are normalised, Mi is cleared
// when clear_mi, and the sweep index is copied for the side-stream kernels.
constexpr int kSppQNone = -1, kSppQAdmix = 1, kSppQProposals = 2;
int launch_spp(const DeviceDataset& ds, const ChainDev& c, bool do_p, int qkind, bool clear_mi, hipStream_t st);
int launch_spp_metro_accept(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_spp_polylog(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
// the stream-graph parse of the Metropolis proposals (then their Dirichlet
// rows), when the chain has its buffers (pg_words); see vps_spp.hip
bool pg_applies(const ChainDev& c);
int launch_pg_proposals(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
// the stream graph's sizes for a chain: words W and states S (0 = not used)
void pg_sizes(int ntyped, int K, int64_t* W, int64_t* S);

int launch_update_p(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
// UpdateP + Dirichlet UpdateQ + Mi clear in one launch (parallel topology)
int launch_update_pq(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_update_q_admix(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_q_metro_propose(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_q_metro_accept(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_q_noadmix_draw(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_update_di(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_update_locpriori(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_update_alpha(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_update_lambda(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_update_a(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
int launch_update_f(const DeviceDataset& ds, const ChainDev& c, hipStream_t st);
// trace: 1 = store lnL in the writelnl trace at the slot that sweep m fills
int launch_record_accumulate(const DeviceDataset& ds, const ChainDev& c, int post_burnin, int trace,
                             int accumulate_misum, hipStream_t st);
// m += 1 on the device (the sweep's last launch)
int launch_advance_m(const ChainDev& c, hipStream_t st);
int launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
// wavefronts per block of the GEN Z kernel (4 or 8) for chunks of Nc
int z_gen_block_waves(const DeviceDataset& ds, int K, int precision, int Nc, int* waves);
// chunk count minimising (grid rounds x per-block work) at the kernel's occupancy
int choose_chunks(const DeviceDataset& ds, int K, int precision, int device, bool groups, int* G);
int launch_init_admix(const DeviceDataset& ds, const ChainDev& c, int singlez, hipStream_t st);

}  // namespace vps

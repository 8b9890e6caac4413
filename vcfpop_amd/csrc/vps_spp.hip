// vps_spp.hip — the reference's sequential Gamma streams drawn in parallel.
//
// In the reference's draw topology UpdateP runs 128 streams
// RNG(seed + m*128 + b, UPDATEP), each drawing the Dirichlet rows of its
// slice kl in [b*KL/128, (b+1)*KL/128) one after another
// (src/structure.cpp:843-861), and UpdateQ runs one stream
// RNG(seed + m, UPDATEQ) over all typed individuals (src/structure.cpp:
// 990-1003; the Metropolis proposals, :1012-1020).  Each Gamma draw
// (Marsaglia–Tsang, src/statistics.cpp:212-242) consumes a data-dependent
// number of words: a Box–Muller pair when no second normal is cached, one
// uniform for the acceptance test, one more for a shape below 1, and more for
// every rejection.  A stream therefore has to be parsed in order.
//
// Parsing in parallel.  One workgroup of 1024 lanes owns a stream; lane j
// takes draw front + j (and, in long windows, front + 1024 + j; see the
// thread layout below).  Almost every draw takes its first candidate, so the
// lane assumes every earlier draw of the window did: draws then alternate
// between "pair + uniform" (3 words, leaves a normal cached) and "cached
// normal + uniform" (1 word), plus one word per shape below 1 (a prefix
// count).  Each lane evaluates its draw exactly, from that speculated start,
// and checks whether its own consumption was the typical one.  The first lane
// that deviates (a rejection, v <= 0, or a Box–Muller value whose low word
// is the cache's empty marker) still started in the right place, so its draw
// is correct; the window commits lanes 0..f and the next window starts from
// lane f's true end state.  Rejections are rare (about 0.3% of draws at shape
// 10, under 0.03% at shape 1000), so a window commits ~1000 draws.
//
// Words come from a circular LDS ring of 8,192 consecutive stream outputs:
// thread t writes the 8 words at indices [8t, 8t + 8); the ring moves on in
// steps of 4,096 words, the 512 threads of the stale half each taking one
// GF(2) jump of a ring cycle.  Every lane reads words at known positions, so
// no lane steps a generator serially.
//
// The second acceptance test of each Gamma draw is decided from
// single-precision estimates with an error bound (mt_log_accept), which keeps
// the outcome of the reference's double expression.
//
// The values equal the serial parse exactly: each committed draw is the
// reference's arithmetic (vps_numeric.h) on the words the serial parse would
// give it.  A draw that would need more than the ring holds (impossible for a
// finite shape; a NaN shape spins) ends the parse's wait: it is stored as NaN
// and its stream moves to the end of the ring.
#include <climits>
#include <utility>
#include "vps_chain.h"

namespace vps {

constexpr int kSppThreads = 1024;
constexpr int kRingWords = 8 * kSppThreads;
constexpr int64_t kRingStep = kRingWords / 2;
constexpr int kPBlocks = 128;       // the reference's UpdateP streams (src/structure.cpp:843)

// A stream's parse state between two draws: words consumed, and the
// Box–Muller cache (the pair drawn at `bm`, values U1 and U2).
struct SppFront {
    int64_t d;        // next draw
    int64_t pos;      // words consumed
    int64_t bm;       // position of the cached pair
    double U1, U2;
    int32_t cached;
    int32_t pad;
    int64_t qa, qb;   // d as (cluster, allele entry) k*KT + x for P, (typed individual, cluster) t*K + k for Q
};

// the draw counts a block keeps in LDS: P's Ni entries of its slice, or Q's
// Mi entries of the typed individuals in draw order and then their indices
constexpr int kStageInts = 20480;

// this thread's generator: the state `8 * threadIdx.x` draws into the ring
// (two jumps: 512 per wave, then 8 per lane)
__device__ __forceinline__ void ring_seed(Xs128& st, const uint32_t* __restrict__ J)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (w) jump_apply(J + (size_t)(kRingLaneTabs + w) * kJumpWords, st);
    if (l) jump_apply(J + (size_t)l * kJumpWords, st);
}

// the seed and salt of stream s of sweep m: UpdateP's streams s < 128
// (src/structure.cpp:843-848), UpdateQ's stream s = 128 (src/structure.cpp:990)
__device__ __forceinline__ void spp_stream_seed(const ChainDev& c, int s, int m, uint64_t* seed, uint64_t* salt)
{
    if (s < kPBlocks) {
        *seed = c.seed + (uint64_t)m * kPBlocks + s;
        *salt = kSaltUpdateP;
    } else {
        *seed = c.seed + (uint64_t)m;
        *salt = kSaltUpdateQ;
    }
}

__device__ __forceinline__ void ring_fill(const Xs128& st, uint64_t* ring)
{
    Xs128 s = st;
    uint64_t* o = ring + 8 * threadIdx.x;
#pragma unroll
    for (int u = 0; u < 8; ++u) o[u] = s.next();
}

// A workgroup barrier that orders LDS only: the parse's global stores are
// each written once and read by later kernels, so no wave waits for them.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// the ring's word at stream position p (held while ws <= p < ws + kRingWords)
__device__ __forceinline__ uint64_t ring_word(const uint64_t* ring, int64_t p) { return ring[p & (kRingWords - 1)]; }

// the 8,192-draw jump table (one ring cycle), copied to LDS once per block
__device__ __forceinline__ void ring_load_advance(const uint32_t* __restrict__ J, uint32_t* tab)
{
    const uint32_t* src = J + (size_t)(kRingPow2Base + 1) * kJumpWords;
    for (int x = threadIdx.x; x < kJumpWords; x += kSppThreads) tab[x] = src[x];
}

// The ring is circular: position p sits at index p mod kRingWords, and
// thread t owns indices [8t, 8t + 8), i.e. positions 8t + kRingWords * r.
// `st` is the thread's generator at the start of its current block.  While
// `pos` lies past the ring's first half, the half holding [ws, ws + kRingStep)
// is refilled with the next cycle's words by the 512 threads that own it (one
// cycle jump each) and the window moves on by kRingStep (a block-uniform
// condition; every thread calls it).
__device__ __forceinline__ void ring_follow(int64_t pos, int64_t& ws, Xs128& st, uint64_t* ring, const uint32_t* tab)
{
    if (pos - ws < kRingStep) return;
    const int64_t n = (pos - ws) / kRingStep;
    lds_barrier();   // (the ring's last readers are done)
#pragma unroll 1
    for (int64_t x = 0; x < n; ++x) {
        const int half = (int)((ws / kRingStep) & 1);
        if ((int)(threadIdx.x >> 9) == half) {
            jump_apply(tab, st);
            ring_fill(st, ring);
        }
        ws += kRingStep;
    }
    lds_barrier();
}

// exclusive prefix count of `flag` over the block's lanes; *total = all
__device__ __forceinline__ int block_prefix(bool flag, int* s_cnt, int* total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t bal = __ballot(flag);
    const int in = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_cnt[w] = __popcll(bal);
    __syncthreads();
    int before = 0, all = 0;
    for (int x = 0; x < kSppThreads / 64; ++x) {
        if (x < w) before += s_cnt[x];
        all += s_cnt[x];
    }
    *total = all;
    return before + in;
}

// the draw kinds of one parse
constexpr int kDrawP = 0;       // Freq[g] <- Gamma(prior + Ni[g]), g = k*KT + allele
constexpr int kDrawQ = 1;       // Q[i][k] <- Gamma(alpha_row(i)[k] + Mi[i][k]), typed i
constexpr int kDrawProp = 2;    // bufNK1[i][k] <- Gamma(alpha_row(i)[k]), typed i

// (qa, qb) of draw d (see SppFront)
__device__ __forceinline__ void draw_coords(int kind, int64_t d, int64_t KT, int K, int64_t* qa, int64_t* qb)
{
    const int64_t n = kind == kDrawP ? KT : K;
    *qa = d / n;
    *qb = d - *qa * n;
}

// draw j of a window whose first draw is the front f: its shape and where its
// value goes.  The coordinates follow from the front's without a division
// (the window holds at most kSppD * 1024 draws); counts come from the LDS stage
// (indexed from the block's first draw d0) when the block has one.
__device__ __forceinline__ double* draw_target(const DeviceDataset& ds, const ChainDev& c, int kind, const SppFront& f,
                                               int j, const int32_t* stage, int64_t d0, int64_t nd,
                                               const double* prior_k, double* alpha)
{
    const int64_t g = f.d + j;
    if (kind == kDrawP) {
        int64_t k = f.qa, x = f.qb + j;
        if (x >= ds.KT) {
            k += x / ds.KT;
            x %= ds.KT;
        }
        const double prior = c.fmodel ? c.FreqA[x] * c.Fst[c.K + (c.fsame ? 0 : k)]
                                      : prior_k ? prior_k[k] : real_store(c.Lambda[k], real32(c));
        *alpha = prior + (double)(stage ? stage[g - d0] : c.Ni[g]);
        return c.Freq + g;
    }
    const int K = c.K;
    int k = (int)f.qb + j % K;
    int64_t t = f.qa + j / K;
    if (k >= K) {
        k -= K;
        ++t;
    }
    const int i = stage ? stage[nd + t] : ds.typed[t];
    const double* ar = c.locpriori ? c.AlphaLocal + (int64_t)ds.popid[i] * K : prior_k ? prior_k : c.Alpha;
    if (kind == kDrawQ) {
        *alpha = ar[k] + (double)(stage ? stage[g] : c.Mi[(int64_t)i * K + k]);
        return c.Q + (int64_t)i * K + k;
    }
    *alpha = ar[k] * 1.0;
    return c.bufNK1 + (int64_t)i * K + k;
}

// Slots.  Seen from the front of a window, the stream is a sequence of
// slots, each one normal (a fresh Box-Muller pair, 2 words, when none is
// cached; else the cached one, 0 words) and one acceptance uniform; the
// pairs alternate with the cached normals.  A draw that takes its first
// candidate fills one slot, and each rejection fills one more, so a draw
// after r rejections in the window starts r slots later than it would
// otherwise.  Shapes below 1 add one uniform after their draw.  Slot s of a
// draw with A such uniforms before it starts at word
//     front + s + 2 * (pairs among slots [0, s)) + A.
// A window evaluates each of its draws under H hypotheses h = 0..H-1 ("h
// rejections before me in this window"); the true hypothesis of draw j+1 is
// h_j + r_j, so a scan over the draws that deviate picks each draw's true
// evaluation.  A v <= 0 candidate uses a normal without a uniform and leaves
// the slot pattern: such a draw ends the window (its own value is right).
__device__ __forceinline__ int64_t slot_words(int64_t s, bool front_cached)
{
    return s + 2 * ((s + (front_cached ? 0 : 1)) >> 1);
}

// The parse of one stream per block.  Blocks [0, np) are UpdateP's 128
// streams (np = 0 or 128); block np, when qkind >= 0, is UpdateQ's stream with
// draws of kind qkind.  Each block turns its rows of Gamma draws into their
// Dirichlet draws once its stream is parsed.  The Metropolis proposals' stream
// state after the parse goes to rngbuf[0] for the acceptance kernel.
//
// Thread layout of a window with H hypotheses (H in 1, 2, 4, 8, 16) and D
// draws per thread (D in 1, 2): thread evaluates D draws, as the virtual
// waves w and w + 16 (w its wave); virtual wave v takes hypothesis v % H for
// draws 64 * (v / H) + lane, so a window holds D * 1024 / H draws.  H and D
// adapt to the deviation rate seen so far: a window commits about
// min(D * 1024 / H, H / rate) draws at a cost of about D + kSppWindowCost
// 1024-draw evaluations.  With one hypothesis the window ends at its first
// deviating draw, which one wave finds by ballots; with more, one thread
// follows the path through the hypotheses over the deviating draws.
constexpr int kSppD = 2;                        // most draws per thread per window
constexpr int kSppVT = kSppD * kSppThreads;     // draws (x hypotheses) per window
constexpr int kSppVW = kSppVT / 64;             // virtual waves
constexpr float kSppWindowCost = 0.5f;          // a window's fixed cost, in 1024-draw evaluations

// one draw of a window, evaluated from its speculated start
struct SppDraw {
    double* out;
    double val, U1, U2;
    int64_t pos, bm;
    bool cache, live;
};

__global__ __launch_bounds__(kSppThreads) void k_spp_gamma(const DeviceDataset ds, const ChainDev c, int np, int qkind,
                                                            int clear_mi)
{
    __shared__ uint64_t ring[kRingWords];
    __shared__ uint32_t s_adv[kJumpWords];
    __shared__ int s_cnt[kSppVW];                // per draw group: shapes below 1
    __shared__ uint8_t s_lt1[kSppVT];            // per draw of the window
    __shared__ uint8_t s_next[kSppVT];           // per virtual thread: next hypothesis, or kNextStop / kNextOvf
    __shared__ uint64_t s_dev[kSppVW];           // per virtual wave: draws not typical under its hypothesis
    __shared__ int s_seg_lane[17];
    __shared__ uint8_t s_seg_h[17];
    __shared__ int s_nseg, s_commit, s_H, s_D, s_lt1_tot;
    __shared__ float s_rate;
    __shared__ SppFront s_front;
    __shared__ double s_prior[VPS_MAX_K];        // per-cluster prior of the draws, when it has no other index
    __shared__ int32_t s_stage[kStageInts];
    constexpr uint8_t kNextStop = 0xFF, kNextOvf = 0xFE;

    const int m = *c.mdev;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int kind;
    uint64_t sseed, salt;
    int64_t d0 = 0, d1, kl0 = 0, kl1 = 0;
    // the sweep index for the kernels that run beside the Z pass
    if (blockIdx.x == 0 && tid == 0) c.mside[0] = m;
    if ((int)blockIdx.x < np) {
        kind = kDrawP;
        const int b = blockIdx.x;
        const int64_t KL = (int64_t)c.K * ds.L;
        kl0 = (int64_t)b * KL / kPBlocks;
        kl1 = (int64_t)(b + 1) * KL / kPBlocks;
        // a slice of kl is a contiguous range of Freq: g = k*KT + afo[l] + allele
        auto first_g = [&](int64_t kl) -> int64_t {
            if (kl >= KL) return (int64_t)c.K * ds.KT;
            const int64_t k = kl / ds.L;
            return k * ds.KT + ds.afo[kl - k * ds.L];
        };
        d0 = first_g(kl0);
        d1 = first_g(kl1);
    } else if ((int)blockIdx.x == np && qkind >= 0) {
        kind = qkind;
        d1 = (int64_t)ds.ntyped * c.K;
    } else {
        // the blocks past the streams prepare the next sweep's stream
        // generators at every thread's ring offset (two jumps each), on CUs
        // the streams leave free
        const int pb = (int)blockIdx.x - np - (qkind >= 0 ? 1 : 0);
        for (int s = 2 * pb; s < min(2 * pb + 2, kSppStreams); ++s) {
            uint64_t sd, sa;
            spp_stream_seed(c, s, m + 1, &sd, &sa);
            Xs128 g;
            g.seed(sd, sa);
            ring_seed(g, ds.ring_jump);
            c.spp_prep[((int64_t)((m + 1) & 1) * kSppStreams + s) * kSppThreads + tid] = g;
            __syncthreads();
            if (tid == 0) {
                __threadfence();
                c.spp_tag[((m + 1) & 1) * kSppStreams + s] = m + 1;
            }
        }
        return;
    }
    const int stream = kind == kDrawP ? (int)blockIdx.x : kPBlocks;
    spp_stream_seed(c, stream, m, &sseed, &salt);
    // the counts the draws read, staged in LDS when they fit
    const int64_t nd = d1 - d0;
    const int32_t* stage = nullptr;
    if (kind == kDrawP && nd <= kStageInts) {
        for (int64_t x = tid; x < nd; x += kSppThreads) s_stage[x] = c.Ni[d0 + x];
        stage = s_stage;
    } else if (kind == kDrawQ && nd + ds.ntyped <= kStageInts) {
        const int K = c.K;
        for (int t = tid; t < ds.ntyped; t += kSppThreads) {
            const int i = ds.typed[t];
            for (int k = 0; k < K; ++k) s_stage[t * K + k] = c.Mi[(int64_t)i * K + k];
            s_stage[nd + t] = i;
        }
        stage = s_stage;
    }
    // the per-cluster priors: lambda for P (no F model), the global alpha
    // for Q and the proposals (no location prior)
    const bool prior_lds = kind == kDrawP ? !c.fmodel : !c.locpriori;
    if (tid < c.K && prior_lds) s_prior[tid] = kind == kDrawP ? real_store(c.Lambda[tid], real32(c)) : c.Alpha[tid];
    const uint32_t* J = ds.ring_jump;
    ring_load_advance(J, s_adv);
    Xs128 st;
    // the generator at this thread's ring offset: prepared by the previous
    // sweep's launch when it ran one, else two jumps here
    const int64_t prep = ((int64_t)(m & 1) * kSppStreams + stream);
    if (c.spp_prep && c.spp_tag[prep] == m) {
        st = c.spp_prep[prep * kSppThreads + tid];
    } else {
        st.seed(sseed, salt);
        ring_seed(st, J);
    }
    int64_t ws = 0;
    ring_fill(st, ring);
    if (tid == 0) {
        s_front.d = d0;
        draw_coords(kind, d0, ds.KT, c.K, &s_front.qa, &s_front.qb);
        s_front.pos = 0;
        s_front.bm = 0;
        s_front.U1 = as_f64(0xFFFFFFFFull);
        s_front.U2 = 0;
        s_front.cached = 0;
        s_H = 1;
        s_D = kSppD;
        s_rate = 0.f;
        s_lt1_tot = 0;
    }
    __syncthreads();

    for (;;) {
        if (s_front.d >= d1) break;
        ring_follow(s_front.pos, ws, st, ring, s_adv);
        const SppFront f = s_front;
        const int H = s_H, Dw = s_D;
        const int G = 16 * Dw / H;                     // draw groups of 64
        const int h = wave % H;                        // the same in both virtual waves (16 % H == 0)
        const int64_t nact = min((int64_t)(64 * G), d1 - f.d);
        int jd[kSppD];
        double alpha[kSppD];
        SppDraw r[kSppD];
#pragma unroll
        for (int x = 0; x < kSppD; ++x) {
            const int grp = (wave + 16 * x) / H;
            jd[x] = grp * 64 + lane;                   // this draw's place in the window
            const bool act = x < Dw && jd[x] < nact;
            alpha[x] = 2.0;
            r[x].out = act ? draw_target(ds, c, kind, f, jd[x], stage, d0, nd, prior_lds ? s_prior : nullptr, &alpha[x])
                           : nullptr;
            r[x].live = act && (jd[x] > 0 || h == 0);  // draw 0 starts at the front
            // count of shapes below 1 per group of 64 draws (their ballot is
            // the same in every hypothesis's wave of a group)
            const bool lt1 = act && alpha[x] < 1;
            const uint64_t bal = __ballot(lt1);
            if (h == 0 && x < Dw) {
                s_lt1[jd[x]] = lt1 ? 1 : 0;
                if (lane == 0) {
                    s_cnt[grp] = __popcll(bal);
                    if (bal) atomicAdd(&s_lt1_tot, __popcll(bal));
                }
            }
        }
        lds_barrier();
        const bool any_lt1 = s_lt1_tot != 0;

        bool dev[kSppD];
#pragma unroll
        for (int x = 0; x < kSppD; ++x) {
            dev[x] = false;
            if (x >= Dw) continue;
            const int grp = jd[x] >> 6;
            const bool lt1 = r[x].out && alpha[x] < 1;
            // shapes below 1 before this draw in the window
            int A = 0;
            if (any_lt1) {
                A = __popcll(__ballot(lt1) & ((1ull << lane) - 1ull));
                for (int g = 0; g < grp; ++g) A += s_cnt[g];
            }
            // start of this draw under hypothesis h
            const int64_t sl = jd[x] + h;
            const bool fc = f.cached != 0;
            const bool cj = fc ^ (sl & 1);
            const int64_t pj = f.pos + slot_words(sl, fc) + A;
            const bool live = r[x].live;
            const int64_t bmj = cj ? (sl == 0 ? f.bm : pj - 3 - (jd[x] > 0 ? s_lt1[jd[x] - 1] : 0)) : pj;

            bool ovf = false;
            auto u01 = [&](int64_t p) -> double {
                const int64_t o = p - ws;
                if (o < 0 || o >= kRingWords) { ovf = true; return 0.5; }
                return unit12(ring_word(ring, p)) - 1.0;
            };
            // the draw's first normal: the cached pair's sine, or a new pair's
            // cosine (one form for both, so the wave stays convergent).  At the
            // front a cached pair comes from the front's values (its words may
            // have left the ring); the ring is read in range only to keep the form.
            const int64_t bsrc = (sl == 0 && cj) || !live ? f.pos : bmj;
            double U1 = my_sqrt(-2.0 * log(u01(bsrc)));
            double U2 = 2.0 * kPi * u01(bsrc + 1);
            if (sl == 0 && cj) { U1 = f.U1; U2 = f.U2; }
            double t = U1 * my_sin(cj ? U2 : U2 + (kPi / 2));
            int64_t pos = pj, bm = f.bm;
            bool cache;
            if (cj) {
                cache = false;
            } else {
                pos += 2;
                bm = pj;
                cache = (uint32_t)as_u64(U1) != 0xFFFFFFFFu;
            }
            bool off_slot = !cj && !cache;   // a pair whose U1 reads as the empty marker
            auto normal = [&]() -> double {
                if (cache) {
                    cache = false;
                    return U1 * my_sin(U2);
                }
                const double v1 = u01(pos), v2 = u01(pos + 1);
                bm = pos;
                pos += 2;
                U1 = my_sqrt(-2.0 * log(v1));
                U2 = 2.0 * kPi * v2;
                cache = (uint32_t)as_u64(U1) != 0xFFFFFFFFu;
                if (!cache) off_slot = true;
                return U1 * my_cos(U2);
            };
            // Gamma(alpha) of RNG<double>, src/statistics.cpp:212-242: shapes
            // below 1 draw Gamma(1 + alpha) and a boost
            const double a = alpha[x];
            const double ag = a < 1 ? 1.0 + a : a;
            const double dd = ag - 0.333333333333333;
            const double cc = 1.0 / (sqrt(dd) * 3.0);
            double v = 1.0, u;
            bool first = true;
            int slots = 0;
            for (int guard = 0; guard < 4096 && !ovf; ++guard) {
                int normals = 0;
                do {
                    if (!first) t = normal();
                    first = false;
                    ++normals;
                    const double v1 = cc * t;
                    v = 1.0 + v1;
                } while (v <= 0 && !ovf);
                if (normals > 1) off_slot = true;
                v = v * v * v;
                u = u01(pos++);
                ++slots;
                if (u < 1.0 - 0.0331 * t * t * t * t) break;
                if (mt_log_accept(u, t, v, dd)) break;
            }
            double val = dd * v;
            if (a < 1) {
                const double boost = pow(u01(pos++), 1.0 / a);
                val = val * boost;
            }
            // where the next draw starts if this one filled `slots` slots: slot
            // sl + slots, behind this draw's own uniform for a shape below 1
            const int64_t nsl = sl + slots;
            const bool ncached = fc ^ (nsl & 1);
            const int64_t npos = f.pos + slot_words(nsl, fc) + A + (lt1 ? 1 : 0);
            const bool aligned = !ovf && !off_slot && pos == npos && cache == ncached &&
                                 (!cache || bm == npos - 3 - (lt1 ? 1 : 0));
            const int nh = h + slots - 1;
            uint8_t nx;
            if (!live) nx = kNextStop;
            else if (ovf) nx = kNextOvf;
            else if (!aligned || nh >= H) nx = kNextStop;
            else nx = (uint8_t)nh;
            s_next[(wave + 16 * x) * 64 + lane] = nx;
            dev[x] = live && nx != (uint8_t)h;
            r[x].val = val;
            r[x].U1 = U1;
            r[x].U2 = U2;
            r[x].pos = pos;
            r[x].bm = bm;
            r[x].cache = cache;
        }
#pragma unroll
        for (int x = 0; x < kSppD; ++x) {
            const uint64_t devb = __ballot(dev[x]);
            if (lane == 0) s_dev[wave + 16 * x] = devb;
        }
        lds_barrier();

        // the window's end and the path through the hypotheses
        if (wave == 0) {
            int commit = (int)nact, events = 0;
            if (H == 1) {
                // the first deviating draw ends the window (its next
                // hypothesis would be 1 or more, or it ran out of ring)
                const uint64_t dv = lane < G ? s_dev[lane] : 0ull;
                const uint64_t nz = __ballot(dv != 0ull);
                if (nz) {
                    const int g0 = __ffsll((unsigned long long)nz) - 1;
                    const uint64_t dg = __shfl(dv, g0, 64);
                    const int draw = g0 * 64 + __ffsll((unsigned long long)dg) - 1;
                    if (draw < (int)nact) {
                        const uint8_t e = s_next[draw];
                        commit = e == kNextOvf ? draw : draw + 1;
                        events = 1;
                    }
                }
                if (lane == 0) {
                    s_seg_lane[0] = 0;
                    s_seg_h[0] = 0;
                    s_nseg = 1;
                }
            } else if (lane == 0) {
                int hh = 0, nseg = 1;
                s_seg_lane[0] = 0;
                s_seg_h[0] = 0;
                for (int g = 0; g < G && commit == (int)nact; ++g) {
                    int j = 0;
                    while (j < 64) {
                        const uint64_t msk = s_dev[g * H + hh] & (~0ull << j);
                        if (!msk) break;
                        const int jj = __ffsll((unsigned long long)msk) - 1;
                        const int draw = g * 64 + jj;
                        if (draw >= (int)nact) break;
                        const uint8_t e = s_next[(g * H + hh) * 64 + jj];
                        if (e == kNextStop || e == kNextOvf) {
                            commit = e == kNextOvf ? draw : draw + 1;
                            ++events;
                            break;
                        }
                        events += e - hh;
                        hh = e;
                        if (nseg < 17) {
                            s_seg_lane[nseg] = draw + 1;
                            s_seg_h[nseg] = (uint8_t)hh;
                            ++nseg;
                        }
                        j = jj + 1;
                    }
                }
                s_nseg = nseg;
            }
            if (lane == 0) {
                s_commit = commit;
                s_lt1_tot = 0;
                // the deviation rate so far
                const int done = max(commit, 1);
                const float rate = 0.5f * s_rate + 0.5f * (float)events / (float)done;
                s_rate = rate;
                // the next window's hypotheses H and draws per thread D: the
                // most draws expected per unit of time, a window of D * 1024
                // evaluations costing D plus a fixed kSppWindowCost
                int bestH = 1, bestD = kSppD;
                float best = 0.f;
                for (int DD = 1; DD <= kSppD; ++DD)
                    for (int HH = 1; HH <= 16; HH *= 2) {
                        const float expect = fminf((float)(DD * kSppThreads / HH), (float)HH / fmaxf(rate, 1e-5f)) /
                                             ((float)DD + kSppWindowCost);
                        if (expect > best * 1.05f) { best = expect; bestH = HH; bestD = DD; }
                    }
                s_H = bestH;
                s_D = bestD;
            }
        }
        lds_barrier();
        const int commit = s_commit;
        const int nseg = s_nseg;
#pragma unroll
        for (int x = 0; x < kSppD; ++x) {
            int hj = 0;
            for (int y = 0; y < nseg; ++y)
                if (s_seg_lane[y] <= jd[x]) hj = s_seg_h[y];
            const bool mine = r[x].live && h == hj;
            if (commit == 0) {
                // the front draw would need more words than the ring holds
                if (jd[x] == 0 && h == 0) {
                    *r[x].out = __builtin_nan("");
                    SppFront nf;
                    nf.d = f.d + 1;
                    nf.pos = ws + kRingWords;
                    nf.cached = 0;
                    nf.bm = 0;
                    nf.U1 = as_f64(0xFFFFFFFFull);
                    nf.U2 = 0;
                    nf.pad = 0;
                    draw_coords(kind, nf.d, ds.KT, c.K, &nf.qa, &nf.qb);
                    s_front = nf;
                }
            } else if (mine && jd[x] < commit) {
                *r[x].out = r[x].val;
                if (jd[x] == commit - 1) {
                    SppFront nf;
                    nf.d = f.d + commit;
                    nf.pos = r[x].pos;
                    nf.cached = r[x].cache ? 1 : 0;
                    nf.bm = r[x].bm;
                    nf.U1 = r[x].cache ? r[x].U1 : as_f64(0xFFFFFFFFull);
                    nf.U2 = r[x].U2;
                    nf.pad = 0;
                    draw_coords(kind, nf.d, ds.KT, c.K, &nf.qa, &nf.qb);
                    s_front = nf;
                }
            }
        }
        lds_barrier();
    }

    // the stream's Dirichlet rows, each scaled by 1/sum (the reference's
    // Dirichlet, src/statistics.cpp:269-296): P rows (k, l) of the slice, or
    // each typed individual's Q (or proposal) row; then the Mi clear for the
    // Z pass (Mi is read only by UpdateQ's stream)
    __syncthreads();   // (the block's stores of the values are visible to all its threads)
    if (kind == kDrawP) {
        for (int64_t kl = kl0 + tid; kl < kl1; kl += kSppThreads) {
            const int64_t k = kl / ds.L, l = kl - k * ds.L;
            double* p = c.Freq + k * ds.KT + ds.afo[l];
            const int k2 = ds.k2[l];
            double sum = 0;
            for (int a = 0; a < k2; ++a) sum += p[a];
            dirichlet_scale(p, sum, k2, real32(c));
        }
    } else {
        const int K = c.K;
        double* base = kind == kDrawQ ? c.Q : c.bufNK1;
        const bool qf = kind == kDrawQ && (c.precision == VPS_PREC_MIXED || c.precision == VPS_PREC_FAST32);
        for (int t = tid; t < ds.ntyped; t += kSppThreads) {
            double* r = base + (int64_t)ds.typed[t] * K;
            double sum = 0;
            for (int k = 0; k < K; ++k) sum += r[k];
            dirichlet_scale(r, sum, K, real32(c));
            // the single-precision copy the FP32-weight Z passes read
            if (qf)
                for (int k = 0; k < K; ++k) c.Qf[(int64_t)ds.typed[t] * K + k] = (float)r[k];
        }
        if (clear_mi) {
            __syncthreads();   // (the stage may have been read from Mi)
            for (int64_t x = tid; x < (int64_t)ds.N * K; x += kSppThreads) c.Mi[x] = 0;
        }
    }

    // the proposals' stream continues in the acceptance kernel: its state
    // after the last draw (generator at `pos`, and the cache)
    if (kind == kDrawProp) {
        ring_follow(s_front.pos, ws, st, ring, s_adv);
        const int64_t p = s_front.pos;
        if (tid == (int)((p & (kRingWords - 1)) >> 3)) {
            Xs128 s = st;
            for (int x = 0; x < (int)(p & 7); ++x) s.next();
            Rng r;
            r.s = s;
            r.U1 = s_front.cached ? s_front.U1 : as_f64(0xFFFFFFFFull);
            r.U2 = s_front.U2;
            c.rngbuf[0] = r;
        }
    }
}

// Dirichlet rows from Gamma draws, for the launches whose rows no stream
// block normalises (the proposals' stream graph; a launch with no stream):
// each P row (k, l) and each typed individual's Q (or proposal) row is scaled
// by 1/sum (the reference's Dirichlet, src/statistics.cpp:269-296).  clear_mi
// zeroes every Mi row for the Z pass; thread 0 copies the sweep index for the
// kernels that run beside the Z pass.
__global__ void k_spp_normalize(const DeviceDataset ds, const ChainDev c, int do_p, int qkind, int clear_mi)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) c.mside[0] = *c.mdev;
    const int64_t np = do_p ? (int64_t)c.K * ds.L : 0;
    if (x < np) {
        const int64_t k = x / ds.L, l = x - k * ds.L;
        double* p = c.Freq + k * ds.KT + ds.afo[l];
        const int k2 = ds.k2[l];
        double s = 0;
        for (int a = 0; a < k2; ++a) s += p[a];
        dirichlet_scale(p, s, k2, real32(c));
        return;
    }
    const int64_t i = x - np;
    if (i >= ds.N) return;
    const int K = c.K;
    if (qkind >= 0 && ds.vt[i]) {
        double* r = (qkind == kDrawQ ? c.Q : c.bufNK1) + i * K;
        double s = 0;
        for (int k = 0; k < K; ++k) s += r[k];
        dirichlet_scale(r, s, K, real32(c));
    }
    if (clear_mi)
        for (int k = 0; k < K; ++k) c.Mi[i * K + k] = 0;
}

// Metropolis acceptance on the proposals' stream (src/structure.cpp:1034-
// 1042): individuals in order, a uniform drawn only for a log ratio below
// NZERO.  The position of each such uniform is a prefix count, so one block
// decides 1024 individuals at a time.  bufN2 arrives as each individual's log
// likelihood under its current Q and leaves as the one under the Q it keeps.
__global__ __launch_bounds__(kSppThreads) void k_spp_metro_accept(const DeviceDataset ds, const ChainDev c)
{
    __shared__ uint64_t ring[kRingWords];
    __shared__ uint32_t s_adv[kJumpWords];
    __shared__ int s_cnt[kSppThreads / 64];
    const uint32_t* J = ds.ring_jump;
    ring_load_advance(J, s_adv);
    Xs128 st = c.rngbuf[0].s;
    ring_seed(st, J);
    int64_t ws = 0, used = 0;
    ring_fill(st, ring);
    __syncthreads();
    const int K = c.K;
    for (int base = 0; base < ds.N; base += kSppThreads) {
        ring_follow(used, ws, st, ring, s_adv);
        const int i = base + threadIdx.x;
        const bool typed = i < ds.N && ds.vt[i];
        const double r = typed ? c.bufN1[i] : 0.0;
        const bool need = typed && !(r >= kNZero);
        int total;
        const int off = block_prefix(need, s_cnt, &total);
        bool acc = typed && r >= kNZero;
        if (need) acc = unit12(ring_word(ring, used + off)) - 1.0 < exp(r);
        if (acc) {
            for (int k = 0; k < K; ++k) c.Q[(int64_t)i * K + k] = c.bufNK1[(int64_t)i * K + k];
            c.bufN2[i] += r;
        }
        used += total;
        __syncthreads();   // (s_cnt and the ring are reused)
    }
}

// UpdateQNoAdmix's PolyLog draws (src/structureAVX.cpp:279-287): typed
// individual number t takes the stream's t-th word; Q rows are cleared and
// Q[i][Z_i] = 1.
__global__ __launch_bounds__(kSppThreads) void k_spp_polylog(const DeviceDataset ds, const ChainDev c)
{
    __shared__ uint64_t ring[kRingWords];
    __shared__ uint32_t s_adv[kJumpWords];
    __shared__ int s_cnt[kSppThreads / 64];
    const int m = *c.mdev;
    const uint32_t* J = ds.ring_jump;
    ring_load_advance(J, s_adv);
    Xs128 st;
    st.seed(c.seed + (uint64_t)m, kSaltUpdateQ);
    ring_seed(st, J);
    int64_t ws = 0, used = 0;
    ring_fill(st, ring);
    __syncthreads();
    const int K = c.K;
    for (int base = 0; base < ds.N; base += kSppThreads) {
        ring_follow(used, ws, st, ring, s_adv);
        const int i = base + threadIdx.x;
        const bool typed = i < ds.N && ds.vt[i];
        int total;
        const int off = block_prefix(typed, s_cnt, &total);
        if (i < ds.N) {
            for (int k = 0; k < K; ++k) c.Q[(int64_t)i * K + k] = 0;
            if (typed) {
                const int k2 = polylog_word(ring_word(ring, used + off), c.bufNK1 + (int64_t)i * K, K);
                c.Q[(int64_t)i * K + k2] = 1;
                c.Z[i] = (uint16_t)k2;
            }
        }
        used += total;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The Metropolis proposals as a stream graph.
//
// With a global alpha (no location prior) the proposal draws of UpdateQMetro
// (src/structure.cpp:1012-1020) are Gamma(Alpha[k]) for k = 0..K-1 of each
// typed individual in turn: the shape depends only on the draw's position
// modulo K.  A draw's start is then fully described by a state
// (p, c, k): the next word p of the stream, whether a Box-Muller normal is
// cached, and k.  A cached normal always comes from the pair 3 words back (4
// when the previous draw's shape was below 1, for its boost uniform), so
// the state fixes its value.  Each state has one successor, the state after
// its draw.  Rather than parsing the stream in order, every state of the
// stream's first W words is evaluated at once (one thread each), and the
// path from the start (0, empty, 0) is found by pointer doubling: with
// path[0..n) known and T = succ^n, path[n + i] = T[path[i]], then T = T o T.
// The values of the D draws on the path are then evaluated again, in
// parallel.  W = 4D + 8192 words covers the path (a draw takes 2 to 3 words
// on average, one more per rejection); a path that would run past W makes
// the draws from there NaN, as the window parse does for a draw longer than
// its ring.  The draws equal the serial parse's exactly: each is the
// reference's Gamma on the words the serial parse gives it.
// ---------------------------------------------------------------------------
namespace {

struct WordSrc {
    const uint64_t* w;
    int64_t pos, lim;
    bool ovf;
    __device__ uint64_t next()
    {
        if (pos >= lim) {
            ovf = true;
            return 0x3FF8000000000000ull;   // U = 1.5: ends any loop
        }
        return w[pos++];
    }
};

struct PgStep {
    double val;
    int64_t p;
    bool cached, ok;
};

// the proposal draw starting at state (p, cf, k)
__device__ PgStep pg_draw(const uint64_t* w, int64_t W, const double* alpha, int K, int64_t p, int cf, int k)
{
    PgStep out{0.0, 0, false, false};
    RngT<WordSrc> r;
    r.s = WordSrc{w, p, W, false};
    r.U1 = as_f64(0xFFFFFFFFull);
    r.U2 = 0;
    if (cf) {
        const int kp = k == 0 ? K - 1 : k - 1;
        const int64_t bm = p - 3 - (alpha[kp] * 1.0 < 1 ? 1 : 0);
        if (bm < 0) return out;
        r.U1 = my_sqrt(-2.0 * log(unit12(w[bm]) - 1.0));
        r.U2 = 2.0 * kPi * (unit12(w[bm + 1]) - 1.0);
        if (!r.cached()) return out;   // (no draw leaves such a state)
    }
    const double a = alpha[k] * 1.0;
    out.val = r.gamma(a);
    out.p = r.s.pos;
    out.cached = r.cached();
    out.ok = !r.s.ovf;
    if (out.ok && out.cached) {
        // the cache must be the pair 3 (+1) words back, as the states assume
        const int64_t bm = out.p - 3 - (a < 1 ? 1 : 0);
        const bool same = bm >= 0 && as_u64(r.U1) == as_u64(my_sqrt(-2.0 * log(unit12(w[bm]) - 1.0))) &&
                          as_u64(r.U2) == as_u64(2.0 * kPi * (unit12(w[bm + 1]) - 1.0));
        if (!same) out.ok = false;
    }
    return out;
}

__device__ __forceinline__ int64_t pg_state(int64_t p, int k, int K, bool cf) { return ((p * K + k) << 1) | (cf ? 1 : 0); }

// the stream's words, 4,096 per block (each thread 8), the generator at each
// 8th word, and path[0] = the start state
__global__ __launch_bounds__(512) void k_pg_words(const DeviceDataset ds, const ChainDev c)
{
    const int m = *c.mdev;
    const int64_t t = (int64_t)blockIdx.x * 512 + threadIdx.x;
    Xs128 s;
    s.seed(c.seed + (uint64_t)m, kSaltUpdateQ);
    const uint32_t* J = ds.ring_jump;
    const uint64_t b = blockIdx.x;
    for (int j = 0; j < kRingPow2Tabs && (b >> j); ++j)
        if ((b >> j) & 1) jump_apply(J + (size_t)(kRingPow2Base + j) * kJumpWords, s);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (wv) jump_apply(J + (size_t)(kRingLaneTabs + wv) * kJumpWords, s);
    if (l) jump_apply(J + (size_t)l * kJumpWords, s);
    c.pg_st8[t] = s;
#pragma unroll
    for (int u = 0; u < 8; ++u) c.pg_words[t * 8 + u] = s.next();
    if (t == 0) c.pg_path[0] = 0;
}

// successor of every state
__global__ void k_pg_next(const ChainDev c)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= c.pg_S) return;
    const int K = c.K;
    const int cf = (int)(x & 1);
    const int64_t pk = x >> 1;
    const int k = (int)(pk % K);
    const int64_t p = pk / K;
    const PgStep st = pg_draw(c.pg_words, c.pg_W, c.Alpha, K, p, cf, k);
    c.pg_next[x] = (st.ok && st.p < c.pg_W) ? (int32_t)pg_state(st.p, (k + 1) % K, K, st.cached) : -1;
}

// one doubling step: with path[0..n) known and T = succ^n, path[n + i] =
// T[path[i]] for i < n, and (square) T2 = T o T; both only read T
__global__ void k_pg_double(int32_t* path, const int32_t* T, int32_t* T2, int64_t n, int64_t D, int64_t S, int square)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n && n + x < D) {
        const int32_t s = path[x];
        path[n + x] = s < 0 ? -1 : T[s];
    }
    if (square && x < S) {
        const int32_t t = T[x];
        T2[x] = t < 0 ? -1 : T[t];
    }
}

// the draws on the path into bufNK1 (typed individual d / K, cluster d % K);
// the last one also leaves the stream's state for the acceptance uniforms
__global__ void k_pg_values(const DeviceDataset ds, const ChainDev c, int64_t D)
{
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    const int K = c.K;
    const int k = (int)(d % K);
    const int i = ds.typed[d / K];
    const int32_t s = c.pg_path[d];
    double* out = c.bufNK1 + (int64_t)i * K + k;
    const PgStep st = s < 0 ? PgStep{0.0, 0, false, false}
                            : pg_draw(c.pg_words, c.pg_W, c.Alpha, K, (int64_t)(s >> 1) / K, s & 1, k);
    *out = st.ok ? st.val : __builtin_nan("");
    if (d != D - 1) return;
    if (!st.ok || st.p >= c.pg_W) {
        // (a path past the words: its draws are NaN, so no proposal is
        // accepted whatever the acceptance uniforms)
        Rng r;
        r.init(c.seed + (uint64_t)*c.mdev, kSaltUpdateQ);
        c.rngbuf[0] = r;
    } else {
        Rng r;
        r.s = c.pg_st8[st.p >> 3];
        for (int x = 0; x < (int)(st.p & 7); ++x) r.s.next();
        if (st.cached) {
            const int64_t bm = st.p - 3 - (c.Alpha[k] * 1.0 < 1 ? 1 : 0);
            r.U1 = my_sqrt(-2.0 * log(unit12(c.pg_words[bm]) - 1.0));
            r.U2 = 2.0 * kPi * (unit12(c.pg_words[bm + 1]) - 1.0);
        } else {
            r.U1 = as_f64(0xFFFFFFFFull);
            r.U2 = 0;
        }
        c.rngbuf[0] = r;
    }
}

}  // namespace

void pg_sizes(int ntyped, int K, int64_t* W, int64_t* S)
{
    const int64_t D = (int64_t)ntyped * K;
    int64_t w = 4 * D + 8192;
    w = (w + 4095) / 4096 * 4096;
    const int64_t s = 2 * (int64_t)K * w;
    // state indices are int32; the pow2 jump tables reach 4096 * 2^24 words
    if (D <= 0 || s >= ((int64_t)1 << 30) || w / 4096 >= ((int64_t)1 << kRingPow2Tabs)) {
        *W = *S = 0;
        return;
    }
    *W = w;
    *S = s;
}

bool pg_applies(const ChainDev& c)
{
    return c.pg_words != nullptr && c.topology == VPS_RNG_REFERENCE && !c.locpriori;
}

int launch_pg_proposals(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    const int64_t D = (int64_t)ds.ntyped * c.K;
    const int64_t S = c.pg_S;
    hipLaunchKernelGGL(k_pg_words, dim3((unsigned)(c.pg_W / 4096)), dim3(512), 0, st, ds, c);
    hipLaunchKernelGGL(k_pg_next, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st, c);
    int32_t* T = c.pg_next;
    int32_t* T2 = c.pg_next + S;
    for (int64_t n = 1; n < D; n *= 2) {
        const int square = 2 * n < D ? 1 : 0;
        const int64_t m = std::max(std::min(n, D - n), square ? S : (int64_t)0);
        hipLaunchKernelGGL(k_pg_double, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, c.pg_path, T, T2, n, D, S,
                           square);
        if (square) std::swap(T, T2);
    }
    hipLaunchKernelGGL(k_pg_values, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, st, ds, c, D);
    VPS_HIP(hipGetLastError());
    // the proposals' Dirichlet rows
    hipLaunchKernelGGL(k_spp_normalize, dim3((unsigned)((ds.N + 255) / 256)), dim3(256), 0, st, ds, c, 0, kDrawProp, 0);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

int launch_spp(const DeviceDataset& ds, const ChainDev& c, bool do_p, int qkind, bool clear_mi, hipStream_t st)
{
    const int np = do_p ? kPBlocks : 0;
    // with UpdateP's streams, blocks that prepare the next sweep's streams
    const int nprep = do_p && c.spp_prep ? (kSppStreams + 1) / 2 : 0;
    const int nb = np + (qkind >= 0 ? 1 : 0) + nprep;
    // the stream blocks normalise their own Dirichlet rows; UpdateQ's block
    // clears Mi after reading it
    if (nb > 0) {
        hipLaunchKernelGGL(k_spp_gamma, dim3(nb), dim3(kSppThreads), 0, st, ds, c, np, qkind,
                           clear_mi && qkind >= 0 ? 1 : 0);
        VPS_HIP(hipGetLastError());
    }
    if (nb == 0 || (clear_mi && qkind < 0)) {
        const int64_t n = ds.N;
        hipLaunchKernelGGL(k_spp_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ds, c, 0, -1,
                           clear_mi ? 1 : 0);
        VPS_HIP(hipGetLastError());
    }
    return VPS_OK;
}

int launch_spp_metro_accept(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    hipLaunchKernelGGL(k_spp_metro_accept, dim3(1), dim3(kSppThreads), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

int launch_spp_polylog(const DeviceDataset& ds, const ChainDev& c, hipStream_t st)
{
    hipLaunchKernelGGL(k_spp_polylog, dim3(1), dim3(kSppThreads), 0, st, ds, c);
    VPS_HIP(hipGetLastError());
    return VPS_OK;
}

}  // namespace vps

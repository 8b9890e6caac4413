// vps_numeric.h — numerics shared by host and device code of the engine:
// seed hashing, the xorshift128+ generator and its distributions, the
// exponent-tracking log accumulator.  Every function is a restatement of
// the algorithm the reference uses, so the engine consumes the reference's
// random-draw streams (the cited lines define the algorithm).
//
// Compile with -ffp-contract=off: each product and sum must round
// separately, as in the reference build (-mno-fma -ffp-contract=off,
// src/makefile.linux.gcc:9).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VPS_HD __host__ __device__ __forceinline__

namespace vps {

constexpr double kMinFreq   = 1e-10;     // MIN_FREQ, src/global.h:60
constexpr double kNZero     = -1e-10;    // NZERO, src/global.h:50
constexpr double kUnderflow = 1e-100;    // DOUBLE_UNDERFLOW, src/global.h:56
constexpr double kOverflow  = 1e100;     // DOUBLE_OVERFLOW, src/global.h:57
constexpr double kPi        = 3.14159265358979323846264338328;  // M_PI, src/global.h:37
constexpr double kLn2Trunc  = 0.693147180559945;                // CloseLog, src/math2.h:894

// stream salts, src/global.h:70-80
constexpr uint64_t kSaltInitAdm  = 0xDA310D821043E1D3ull;
constexpr uint64_t kSaltUpdateQ  = 0xE7853EB09C92935Aull;
constexpr uint64_t kSaltUpdateZ  = 0x9B83516C63E249C3ull;
constexpr uint64_t kSaltUpdateP  = 0xD4EA1DB746DF14E3ull;
constexpr uint64_t kSaltUpdateA  = 0x49913AD9EDBD2EB8ull;
constexpr uint64_t kSaltUpdateLoc   = 0x95EE7991226E2D14ull;
constexpr uint64_t kSaltUpdateAlpha = 0x8A26E2C4D1BC5E54ull;
constexpr uint64_t kSaltUpdateLambda = 0x67099E2FE2510092ull;
constexpr uint64_t kSaltUpdateF  = 0xF82104A5ADDAAF9Full;

VPS_HD uint64_t as_u64(double v) { return __builtin_bit_cast(uint64_t, v); }
VPS_HD double as_f64(uint64_t u) { return __builtin_bit_cast(double, u); }

// MurmurHash2 of one 32-bit word (Appleby's 4-byte case), src/hash.cpp:429-447
VPS_HD uint32_t murmur2(uint32_t data, uint32_t seed)
{
    const uint32_t m = 0x5bd1e995u;
    uint32_t s = (seed ^ 4u) * m;
    uint32_t a = data * m;
    a ^= a >> 24;
    a *= m;
    a ^= s;
    a ^= a >> 13;
    a *= m;
    a ^= a >> 15;
    return a;
}

// MurmurHash64: the high word is first mixed with the complement of the low
// word, src/hash.cpp:450-460
VPS_HD uint64_t murmur64(uint64_t data, uint64_t seed)
{
    uint32_t d0 = (uint32_t)data, s0 = (uint32_t)seed;
    uint32_t d1 = ~d0 ^ (uint32_t)(data >> 32);
    uint32_t s1 = ~s0 ^ (uint32_t)(seed >> 32);
    return ((uint64_t)murmur2(d1, s1) << 32) | murmur2(d0, s0);
}

// HashULong: CRC32C (reflected polynomial 0x82F63B78) of the 8 bytes of v,
// little-endian, from the initial value 0xB385FDA9 (src/hash.cpp:219-223).
// Bitwise form: eight shift/xor steps per byte.
VPS_HD uint32_t hash_ulong(uint64_t v)
{
    uint32_t c = 0xB385FDA9u;
    for (int b = 0; b < 8; ++b) {
        c ^= (uint32_t)(v >> (8 * b)) & 0xFFu;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
    }
    return c;
}

// xorshift128+ state of one stream, seeded as RNG<double>(s, salt),
// src/statistics.cpp:62-70
struct Xs128 {
    uint64_t x, y;
    VPS_HD void seed(uint64_t s, uint64_t salt)
    {
        uint64_t h = murmur64(s, salt);
        x = 0x159A55E5075BCD15ull ^ h;
        y = 0x054913331F123BB5ull ^ (h << 6);
    }
    // one step with shifts (23, 18, 5); returns new_x + new_y,
    // src/statistics.cpp:73-86.  On gfx950 the two right shifts are single
    // 64-bit v_lshrrev_b64 (the compiler splits them into two 32-bit
    // funnel shifts each) and the four-way XOR a1 ^ (a1 >> 18) ^ b ^ (b >> 5)
    // uses the 3-input v_bitop3_b32.
    VPS_HD uint64_t next()
    {
        uint64_t a = x, b = y;
        x = b;
        a ^= a << 23;
#if defined(__HIP_DEVICE_COMPILE__) && defined(__gfx950__)
        uint64_t s, u;
        asm("v_lshrrev_b64 %0, 18, %1" : "=v"(s) : "v"(a));
        asm("v_lshrrev_b64 %0, 5, %1" : "=v"(u) : "v"(b));
        const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)s, (uint32_t)b, 0x96) ^ (uint32_t)u;
        const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(s >> 32), (uint32_t)(b >> 32), 0x96) ^
                            (uint32_t)(u >> 32);
        a = (uint64_t)lo | ((uint64_t)hi << 32);
#else
        a ^= a >> 18;
        a ^= b;
        a ^= b >> 5;
#endif
        y = a;
        return a + b;
    }
};

// MurmurHash32, src/hash.cpp:462-469
VPS_HD uint32_t murmur32(uint64_t data, uint64_t seed)
{
    const uint32_t a = (~(uint32_t)data) ^ (uint32_t)(data >> 32);
    const uint32_t b = (~(uint32_t)seed) ^ (uint32_t)(seed >> 32);
    return murmur2(a, b);
}

// RNG<float>: xorshift96 with shifts (16, 5, 1), seeded by MurmurHash32
// (src/statistics.cpp:330-354); the Z pass's per-locus streams when
// REAL = float with -g_fastsingle
struct Xs96 {
    uint32_t x, y, z;
    VPS_HD void seed(uint64_t s, uint64_t salt)
    {
        const uint32_t h = murmur32(s, salt);
        x = 0x075BCD15u ^ h;
        y = 0x159A55E5u ^ (h << 3);
        z = 0x1F123BB5u ^ (h << 6);
    }
    VPS_HD uint32_t next()
    {
        uint32_t t = x;
        t ^= t << 16;
        t ^= t >> 5;
        t ^= t << 1;
        x = y;
        y = z;
        z = t ^ x ^ y;
        return z;
    }
};

// U in [1, 2) from the 52 low bits of a draw
VPS_HD double unit12(uint64_t r) { return as_f64((r & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull); }

// MyLog, src/math2.h:608-615
VPS_HD double my_log(double v)
{
    if (v < 1e-300) return -6.90775527898214000E+02;
    if (v > 1e300) return 6.90775527898214000E+02;
    return log(v);
}

// MySqrt, src/math2.h:294-302
VPS_HD double my_sqrt(double v)
{
    if (v > 0) return sqrt(v);
    if (v > -kMinFreq) return 0;
    return __builtin_nan("");
}

// Sine/cosine used by the reference's Box–Muller transform: quadrant
// reduction and the minimax polynomials of src/math2.h:1087-1142.  The
// normal variates' bit patterns depend on these exact coefficients.
VPS_HD double cos_poly(double x)
{
    double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
    return (-2.7236370439787708e-7 * x2 + 2.4799852696610628e-5) * x8 +
           (-1.3888885054799695e-3 * x2 + 4.1666666636943683e-2) * x4 +
           (-4.9999999999963024e-1 * x2 + 1.0000000000000000e+0);
}
VPS_HD double sin_poly(double x)
{
    double x2 = x * x, x4 = x2 * x2;
    return ((2.7181216275479732e-6 * x2 - 1.9839312269456257e-4) * x4 +
            (8.3333293048425631e-3 * x2 - 1.6666666640797048e-1)) * x2 * x + x;
}
VPS_HD double my_sin(double x)
{
    double y = x * 6.3661977236758138e-1;
    double t0 = floor(fabs(y) + 0.5);
    double q = y < 0.0 ? -t0 : t0;
    int quadrant = (int)q;
    double t = x - q * 1.5707963267923333e+00;
    t = t - q * 2.5633441515945189e-12;
    t = (quadrant & 1) ? cos_poly(t) : sin_poly(t);
    return (quadrant & 2) ? -t : t;
}
VPS_HD double my_cos(double x) { return my_sin(x + (kPi / 2)); }

// Marsaglia–Tsang's second acceptance test as the reference evaluates it,
//     log(u) < 0.5 * t * t + d * (1.0 - v + log(v))      (src/statistics.cpp:239)
// in double precision.  Two double-precision logarithms cost about 200
// instructions, so the two sides are first estimated cheaply:
//   * log(u) from the single-precision logarithm of (float)u: input
//     rounding moves it by at most 2^-24 and the single-precision result is
//     within a few ulps, so it is within 3e-7 * (1 + |log u|);
//   * g = 1 - v + log(v) = log1p(w) - w with w = v - 1 (exact for v in
//     [0.5, 2]): for |w| <= 1/8 the alternating series -w^2/2 + w^3/3 - ...
//     to w^17, whose tail is below |w|^18 / 18 / (1 - |w|) < 4e-18; otherwise
//     1 - v plus the single-precision log(v), within 3e-7 * (1 + |log v|);
//   * the reference's own rounding of the double expression is a few ulps
//     of its largest intermediate, bounded here by 1e-15 of their sum.
// When the estimates differ by more than the bound the exact comparison has
// the same outcome; otherwise it is evaluated as the reference writes it.
// Inputs outside the ranges the bounds cover (u below 2^-100, v outside
// [2^-100, 2^100], NaN) take the exact form too.
enum MtDecision { kMtReject = 0, kMtAccept = 1, kMtExact = 2 };
VPS_HD bool mt_log_exact(double u, double t, double v, double d)
{
    return log(u) < 0.5 * t * t + d * (1.0 - v + log(v));
}
VPS_HD MtDecision mt_log_decide(double u, double t, double v, double d)
{
    const double tt = 0.5 * t * t;
    if (u >= 0x1p-100 && v >= 0x1p-100 && v <= 0x1p100 && d >= 0) {
        const double lu = (double)logf((float)u);
        const double w = v - 1.0;
        double g, eg;
        if (w >= -0.125 && w <= 0.125) {
            // log1p(w) - w = w^2 * S, S = sum_{n>=2} (-1)^(n+1) w^(n-2) / n,
            // by Horner from n = 17
            double s = 1.0 / 17;
#pragma unroll
            for (int n = 16; n >= 2; --n) s = ((n & 1) ? 1.0 : -1.0) / n + w * s;
            g = (w * w) * s;
            eg = 4e-18 + 4e-15 * (w * w);
        } else {
            const double lv = (double)logf((float)v);
            g = -w + lv;
            eg = 3e-7 * (1.0 + fabs(lv)) + 1e-15 * fabs(w);
        }
        const double rhs = tt + d * g;
        const double bound = 3e-7 * (1.0 + fabs(lu)) + d * eg +
                             1e-15 * (tt + d * (fabs(w) + fabs(g) + 1.0) + fabs(lu) + 1.0);
        const double diff = lu - rhs;
        if (diff < -bound) return kMtAccept;
        if (diff > bound) return kMtReject;
    }
    return kMtExact;
}
VPS_HD bool mt_log_accept(double u, double t, double v, double d)
{
    const MtDecision r = mt_log_decide(u, t, v, d);
    return r == kMtExact ? mt_log_exact(u, t, v, d) : r == kMtAccept;
}

// RNG<double> with the Box–Muller cache (U1, U2), src/statistics.h:15-87.
// The cache is empty when U1's low 32 bits are all ones.  Src supplies the
// raw 64-bit draws: the xorshift128+ generator, or words already drawn.
template <typename Src>
struct RngT {
    Src s;
    double U1, U2;
    VPS_HD void init(uint64_t seed, uint64_t salt)
    {
        s.seed(seed, salt);
        U1 = as_f64(0xFFFFFFFFull);
        U2 = 0;
    }
    VPS_HD uint64_t xorshift() { return s.next(); }
    // Uniform(min, max), src/statistics.cpp:168-174
    VPS_HD double uniform(double mn = 0, double mx = 1) { return (unit12(s.next()) - 1.0) * (mx - mn) + mn; }
    VPS_HD bool cached() const { return (uint32_t)as_u64(U1) != 0xFFFFFFFFu; }
    VPS_HD void clear() { U1 = as_f64((as_u64(U1) & 0xFFFFFFFF00000000ull) | 0xFFFFFFFFull); }
    // Normal(mean, std), src/statistics.cpp:185-201
    VPS_HD double normal(double mean = 0, double sd = 1)
    {
        if (cached()) {
            double re = U1 * my_sin(U2);
            clear();
            return re * sd + mean;
        }
        double v1 = uniform();
        double v2 = uniform();
        U1 = my_sqrt(-2.0 * log(v1));
        U2 = 2.0 * kPi * v2;
        return U1 * my_cos(U2) * sd + mean;
    }
    // Gamma(alpha, beta): Marsaglia–Tsang, alpha < 1 boosted,
    // src/statistics.cpp:212-242 (evaluation order kept)
    VPS_HD double gamma(double alpha, double beta = 1)
    {
        double boost = 1.0;
        if (alpha < 1) {
            double g = gamma_ge1(1.0 + alpha, beta);
            boost = pow(uniform(), 1.0 / alpha);
            return g * boost;
        }
        return gamma_ge1(alpha, beta);
    }
    // The rejection loop is capped at 4096 attempts: for a finite shape the
    // acceptance rate is above 0.95, so the cap never binds (probability
    // < 1e-5000); for a NaN shape it ends the loop instead of spinning.
    VPS_HD double gamma_ge1(double alpha, double beta)
    {
        double t, v, u;
        double d = alpha - 0.333333333333333;
        double c = 1.0 / (sqrt(d) * 3.0);
        for (int guard = 0; guard < 4096; ++guard) {
            do {
                t = normal();
                double v1 = c * t;
                v = 1.0 + v1;
            } while (v <= 0);
            v = v * v * v;
            u = uniform();
            if (u < 1.0 - 0.0331 * t * t * t * t) break;
            if (mt_log_accept(u, t, v, d)) break;
        }
        return beta * d * v;
    }
    // Next(max), NextAvoid(max, av), src/statistics.cpp:301-320
    VPS_HD uint64_t next(uint64_t mx) { return s.next() % mx; }
    VPS_HD uint64_t next_avoid(uint64_t mx, uint64_t av)
    {
        uint64_t a = next(mx - 1);
        if (a >= av) a++;
        return a;
    }
};
using Rng = RngT<Xs128>;

// Dirichlet(res, a, b, n): Gamma(a[i] + b[i]) then scale by 1/sum,
// src/statistics.cpp:284-296
// Scales a row of n Gamma draws v (sum s, added in row order) into its
// Dirichlet draw.  real32: the row is a REAL = float array of BAYESIAN<float>:
// v is stored as float and multiplied by (float)(1/s) in single precision
// (Mul(float*, float, n), src/math2.cpp:2088-2102); the values stay doubles
// holding single-precision numbers.
VPS_HD void dirichlet_scale(double* res, double s, int n, bool real32)
{
    if (real32) {
        const float inv = (float)(1.0 / s);
        for (int i = 0; i < n; ++i) res[i] = (double)((float)res[i] * inv);
    } else {
        const double inv = 1.0 / s;
        for (int i = 0; i < n; ++i) res[i] *= inv;
    }
}

// a value stored into a REAL array (real32: rounded to float)
VPS_HD double real_store(double v, bool real32) { return real32 ? (double)(float)v : v; }

template <typename TB>
VPS_HD void dirichlet_ab(Rng& r, double* res, const double* a, const TB* b, int n, bool real32)
{
    double s = 0;
    for (int i = 0; i < n; ++i) {
        double v = r.gamma(a[i] + (double)b[i]);
        res[i] = v;
        s += v;
    }
    dirichlet_scale(res, s, n, real32);
}

// Dirichlet(res, a, n, f = 1): Gamma(a[i] * f), src/statistics.cpp:269-281
VPS_HD void dirichlet_a(Rng& r, double* res, const double* a, int n, bool real32)
{
    double s = 0;
    for (int i = 0; i < n; ++i) {
        double v = r.gamma(a[i] * 1.0);
        res[i] = v;
        s += v;
    }
    dirichlet_scale(res, s, n, real32);
}

// PolyLog(a, n), src/statistics.cpp:116-155 (overwrites a), with its one
// draw given: `word` is the stream's next xorshift output
VPS_HD int polylog_word(uint64_t word, double* a, int n)
{
    double mx = a[0];
    for (int i = 1; i < n; ++i) mx = a[i] > mx ? a[i] : mx;
    double s = kMinFreq * n;
    for (int i = 0; i < n; ++i) {
        double diff = a[i] - mx;
        a[i] = diff < -23 ? kMinFreq : exp(diff);
        s += a[i];
    }
    double t = (unit12(word) - 1.0) * (s - 0) + 0;   // Uniform(0, s)
    for (int i = 0; i < n; ++i) {
        if (t < a[i]) return i;
        t -= a[i];
    }
    return n - 1;
}
VPS_HD int polylog(Rng& r, double* a, int n) { return polylog_word(r.xorshift(), a, n); }

// Exponent-tracking product accumulator (OpenLog/ChargeLog/CloseLog,
// src/math2.h:716-902): the binary exponent of a factor or of the running
// product moves into an integer when the value leaves [1e-100, 1e100].
struct LogAcc {
    int64_t slog;
    double prod;
    VPS_HD void open() { slog = 0; prod = 1; }
    VPS_HD static void strip(int64_t& sl, double& v)
    {
        uint64_t vv = as_u64(v);
        sl += (int64_t)((vv & 0x7FF0000000000000ull) >> 52) - 1023;
        v = as_f64((vv & 0x800FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    }
    VPS_HD void charge(double val)
    {
        if (val < kUnderflow || val > kOverflow) strip(slog, val);
        prod *= val;
        if (prod < kUnderflow || prod > kOverflow) strip(slog, prod);
    }
    VPS_HD double close() const { return ((double)slog + log2(prod)) * kLn2Trunc; }
};

// LogProd(A, n, sep) in sequential order (the strided form of the AVX path
// is a scalar loop, src/mathAVX.cpp:1115-1125)
VPS_HD double logprod_sep(const double* a, int64_t n, int64_t sep)
{
    LogAcc acc;
    acc.open();
    for (int64_t i = 0; i < n; ++i) acc.charge(a[i * sep]);
    return acc.close();
}

// Row normalisation with MIN_FREQ (UnifyAVX, src/mathAVX.cpp:2879-2900):
// groups of four use x*inv + MIN_FREQ*inv, the tail (x + MIN_FREQ)*inv.
VPS_HD void unify_row(double* a, int n)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i];
    double inv = 1.0 / (s + (double)n * kMinFreq);
    int i = 0;
    if (n >= 4) {
        double c = kMinFreq * inv;
        for (; i + 4 <= n; i += 4)
            for (int j = 0; j < 4; ++j) a[i + j] = a[i + j] * inv + c;
    }
    for (; i < n; ++i) a[i] = (a[i] + kMinFreq) * inv;
}

}  // namespace vps

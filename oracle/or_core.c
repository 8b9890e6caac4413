/*
 * or_core.c — oracle numeric core: seed hashing, RNG<double>/RNG<float>,
 * distributions, the exponent-tracking log accumulator and the SIMD-order
 * reductions of the AVX path.  Test infrastructure only (see vps_oracle.h).
 *
 * Compiled with -ffp-contract=off: the reference is built with
 * -mno-fma -ffp-contract=off (src/makefile.linux.gcc:9), so every product
 * and sum below is a separately rounded IEEE operation.
 */
#include <math.h>
#include <string.h>
#include "vps_oracle.h"

/* ------------------------------------------------------------------ */
/* CRC32C (Castagnoli, reflected polynomial 0x82F63B78).  The reference   */
/* uses the SSE4.2 instruction _mm_crc32_u64 (src/hash.cpp:9-16); this is */
/* the bitwise software definition of the same function.                  */
/* ------------------------------------------------------------------ */
static uint32_t crc_table[8][256];
static int crc_ready = 0;

static void crc_init(void)
{
    for (uint32_t n = 0; n < 256; ++n) {
        uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : c >> 1;
        crc_table[0][n] = c;
    }
    for (uint32_t n = 0; n < 256; ++n)
        for (int t = 1; t < 8; ++t)
            crc_table[t][n] = (crc_table[t - 1][n] >> 8) ^ crc_table[0][crc_table[t - 1][n] & 0xFF];
    crc_ready = 1;
}

uint32_t or_crc32c_u64(uint32_t prev, uint64_t v)
{
    if (!crc_ready) crc_init();
    uint32_t c = prev;
    for (int b = 0; b < 8; ++b) {
        c = (c >> 8) ^ crc_table[0][(c ^ (uint32_t)(v >> (8 * b))) & 0xFF];
    }
    return c;
}

/* HashULong, src/hash.cpp:219-223 (HASH = uint, src/vcfpop.h:546-549) */
uint32_t or_hash_ulong(uint64_t v) { return or_crc32c_u64(0xB385FDA9u, v); }

/* HashString, src/hash.cpp:182-214: 8-byte little-endian words, then a tail word */
uint32_t or_hash_string(const void* s, int len)
{
    const unsigned char* p = (const unsigned char*)s;
    uint32_t crc = 0xB385FDA9u;
    for (; len >= 8; len -= 8, p += 8) {
        uint64_t w; memcpy(&w, p, 8);
        crc = or_crc32c_u64(crc, w);
    }
    if (len > 0) {
        uint64_t w = 0;
        for (int i = 0; i < len; ++i) w |= (uint64_t)p[i] << (8 * i);
        crc = or_crc32c_u64(crc, w);
    }
    return crc;
}

/* MurmurHash2 of one 32-bit word, src/hash.cpp:429-447 */
uint32_t or_murmur2(uint32_t data, uint32_t seed)
{
    const uint32_t m = 0x5bd1e995u;
    uint32_t s = seed ^ 4u;
    uint32_t a = data;
    a *= m;
    a ^= a >> 24;
    a *= m;
    s *= m;
    a ^= s;
    a ^= a >> 13;
    a *= m;
    a ^= a >> 15;
    return a;
}

/* MurmurHash64, src/hash.cpp:450-460: high halves are mixed with ~low */
uint64_t or_murmur64(uint64_t data, uint64_t seed)
{
    uint32_t d0 = (uint32_t)data, d1 = (uint32_t)(data >> 32);
    uint32_t s0 = (uint32_t)seed, s1 = (uint32_t)(seed >> 32);
    d1 = (~d0) ^ d1;
    s1 = (~s0) ^ s1;
    return ((uint64_t)or_murmur2(d1, s1) << 32) | (uint64_t)or_murmur2(d0, s0);
}

/* MurmurHash32, src/hash.cpp:462-469 */
uint32_t or_murmur32(uint64_t data, uint64_t seed)
{
    uint32_t a32 = (~(uint32_t)data) ^ (uint32_t)(data >> 32);
    uint32_t s32 = (~(uint32_t)seed) ^ (uint32_t)(seed >> 32);
    return or_murmur2(a32, s32);
}

/* ------------------------------------------------------------------ */
/* math helpers                                                         */
/* ------------------------------------------------------------------ */
static inline uint64_t d2u(double v) { uint64_t u; memcpy(&u, &v, 8); return u; }
static inline double u2d(uint64_t u) { double v; memcpy(&v, &u, 8); return v; }

/* MyLog, src/math2.h:608-615 */
double or_mylog(double v)
{
    if (v < 1e-300) return -6.90775527898214000E+02;
    else if (v > 1e300) return +6.90775527898214000E+02;
    return log(v);
}

/* MySqrt, src/math2.h:294-302 */
double or_mysqrt(double v)
{
    if (v > 0) return sqrt(v);
    else if (v > -OR_MIN_FREQ) return 0;
    return NAN;
}

/* Polynomial kernels of MySin/MyCos, src/math2.h:1087-1105.  The
   coefficients are those of the public minimax fit the reference cites
   (src/math2.h:1119, CyberGeeks math.c); they define the Box–Muller output
   bit pattern and must be identical for parity. */
static double cos_core(double x)
{
    double x2 = x * x;
    double x4 = x2 * x2;
    double x8 = x4 * x4;
    return (-2.7236370439787708e-7 * x2 + 2.4799852696610628e-5) * x8 +
           (-1.3888885054799695e-3 * x2 + 4.1666666636943683e-2) * x4 +
           (-4.9999999999963024e-1 * x2 + 1.0000000000000000e+0);
}

static double sin_core(double x)
{
    double x2 = x * x;
    double x4 = x2 * x2;
    return ((2.7181216275479732e-6 * x2 - 1.9839312269456257e-4) * x4 +
            (8.3333293048425631e-3 * x2 - 1.6666666640797048e-1)) * x2 * x + x;
}

/* MyRInt + MySin, src/math2.h:1119-1136 */
double or_mysin(double x)
{
    double t0 = floor(fabs(x * 6.3661977236758138e-1) + 0.5);
    double q = (x * 6.3661977236758138e-1 < 0.0) ? -t0 : t0;
    int quadrant = (int)q;
    double t = x - q * 1.5707963267923333e+00;
    t = t - q * 2.5633441515945189e-12;
    t = (quadrant & 1) ? cos_core(t) : sin_core(t);
    return (quadrant & 2) ? -t : t;
}

/* MyCos, src/math2.h:1139-1142; M_PI as src/global.h:37 */
double or_mycos(double x) { return or_mysin(x + (3.14159265358979323846264338328 / 2)); }

/* AddExponent, src/math2.h:736-746: move the binary exponent into slog */
static inline void add_exponent(int64_t* slog, double* val)
{
    uint64_t vv = d2u(*val);
    *slog += (int64_t)((vv & 0x7FF0000000000000ull) >> 52) - 1023;
    *val = u2d((vv & 0x800FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
}

/* ChargeLog(slog, prod, double), src/math2.h:760-769 */
void or_charge_log(int64_t* slog, double* prod, double val)
{
    if (val < OR_DOUBLE_UNDERFLOW || val > OR_DOUBLE_OVERFLOW) add_exponent(slog, &val);
    *prod *= val;
    if (*prod < OR_DOUBLE_UNDERFLOW || *prod > OR_DOUBLE_OVERFLOW) add_exponent(slog, prod);
}

/* CloseLog, src/math2.h:891-895 (ln 2 truncated to 0.693147180559945) */
double or_close_log(int64_t slog, double prod)
{
    return ((double)slog + log2(prod)) * 0.693147180559945;
}

/* LogProdAVX(double*, n), src/mathAVX.cpp:1053-1082: for n >= 32 eight
   lanes take elements i = j (mod 8) in order, exponents are stripped after
   every 32 elements (AddExponentAVX, src/mathAVX.cpp:1001-1014), then the
   lanes are folded pairwise (ChargeLogSSE, src/mathSSE.cpp:1020-1028) and
   the tail is charged one by one. */
double or_logprod_avx(const double* a, int64_t n)
{
    int64_t i = 0, slog = 0;
    double prod = 1;
    if (n >= 32) {
        double pd[8];
        for (int j = 0; j < 8; ++j) pd[j] = 1.0;
        for (int64_t l1 = n - 32; i <= l1; i += 32) {
            for (int u = 0; u < 4; ++u)
                for (int j = 0; j < 8; ++j) pd[j] *= a[i + u * 8 + j];
            int tsum = 0;
            for (int j = 0; j < 8; ++j) {
                uint64_t vv = d2u(pd[j]);
                tsum += (int)(((vv & 0x7FF0000000000000ull) >> 52) - 1023);
                pd[j] = u2d((vv & 0x800FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
            }
            slog += tsum;
        }
        for (int p = 0; p < 4; ++p) {
            double v0 = pd[2 * p], v1 = pd[2 * p + 1];
            uint64_t a0 = d2u(v0), a1 = d2u(v1);
            int t = (int)(((a0 & 0x7FF0000000000000ull) >> 52) - 1023) +
                    (int)(((a1 & 0x7FF0000000000000ull) >> 52) - 1023);
            slog += t;
            v0 = u2d((a0 & 0x800FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
            v1 = u2d((a1 & 0x800FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
            prod *= v0 * v1;
            if (prod < OR_DOUBLE_UNDERFLOW || prod > OR_DOUBLE_OVERFLOW) add_exponent(&slog, &prod);
        }
    }
    for (; i < n; ++i) or_charge_log(&slog, &prod, a[i]);
    return or_close_log(slog, prod);
}

/* LogProdAVX(double*, n, sep), src/mathAVX.cpp:1115-1125 (scalar loop) */
double or_logprod_sep(const double* a, int64_t n, int64_t sep)
{
    int64_t slog = 0;
    double prod = 1;
    for (int64_t i = 0; i < n; ++i, a += sep) or_charge_log(&slog, &prod, *a);
    return or_close_log(slog, prod);
}

/* SumAVX(double*, n), src/mathAVX.cpp:1236-1264 */
double or_sum_avx(const double* a, int64_t n)
{
    int64_t i = 0;
    volatile double re = 0;
    if (n >= 32) {
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t l1 = n - 32; i <= l1; i += 32)
            for (int u = 0; u < 4; ++u)
                for (int j = 0; j < 8; ++j) s[j] += a[i + u * 8 + j];
        double l[4];
        for (int j = 0; j < 4; ++j) l[j] = s[j] + s[j + 4];
        re = (l[0] + l[2]) + (l[1] + l[3]);
    }
    for (; i < n; ++i) { volatile double v1 = a[i]; re += v1; }
    return re;
}

/* UnifyAVX(double*, n), src/mathAVX.cpp:2879-2900 (fmaddx is mul then add
   without FMA, src/vcfpop.h:237-240) */
void or_unify_avx(double* a, int64_t n)
{
    int64_t i = 0;
    double invsum = 1.0 / (or_sum_avx(a, n) + (double)n * OR_MIN_FREQ);
    if (n >= 4) {
        double b = invsum, c = OR_MIN_FREQ * invsum;
        for (int64_t l1 = n - 4; i <= l1; i += 4)
            for (int j = 0; j < 4; ++j) a[i + j] = a[i + j] * b + c;
    }
    for (; i < n; ++i) { volatile double v1 = a[i] + OR_MIN_FREQ; a[i] = v1 * invsum; }
}

/* ------------------------------------------------------------------ */
/* RNG<double>, src/statistics.cpp:57-321                              */
/* ------------------------------------------------------------------ */
static inline int u1_cached(const or_rng* r) { return (uint32_t)d2u(r->U1) != 0xFFFFFFFFu; }
static inline void u1_clear(or_rng* r)
{
    uint64_t u = d2u(r->U1);
    r->U1 = u2d((u & 0xFFFFFFFF00000000ull) | 0xFFFFFFFFull);
}

/* RNG(s, salt), src/statistics.cpp:62-70 */
void or_rng_init(or_rng* r, uint64_t s, uint64_t salt)
{
    s = or_murmur64(s, salt);
    r->x = 0x159A55E5075BCD15ull ^ s;
    r->y = 0x054913331F123BB5ull ^ (s << 6);
    r->U1 = 0; r->U2 = 0;
    u1_clear(r);
}

/* XorShift, src/statistics.cpp:73-86: xorshift128+ with shifts (23,18,5) */
uint64_t or_rng_xorshift(or_rng* r)
{
    uint64_t a = r->x, b = r->y;
    r->x = b;
    a ^= a << 23;
    a ^= a >> 18;
    a ^= b;
    a ^= b >> 5;
    r->y = a;
    return a + b;
}

/* Uniform(min, max), src/statistics.cpp:168-174 */
double or_rng_uniform(or_rng* r, double mn, double mx)
{
    uint64_t u = or_rng_xorshift(r), b = 0x3FF0000000000000ull;
    b |= u & 0x000FFFFFFFFFFFFFull;
    return (u2d(b) - 1.0) * (mx - mn) + mn;
}

/* Poly(a, n) scalar form, src/statistics.cpp:89-113 */
int or_rng_poly_scalar(or_rng* r, const double* a, int n)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i];
    s = s + OR_MIN_FREQ * n;
    double t = or_rng_uniform(r, 0, s);
    for (int i = 0; i < n; ++i) {
        if (t < a[i] + OR_MIN_FREQ) return i;
        t -= a[i] + OR_MIN_FREQ;
    }
    return n - 1;
}

/* RNGAVX<double>::Poly one lane, src/mathAVX.cpp:73-122: w[k] already holds
   q*p + MIN_FREQ (src/structureAVX.cpp:935-940); s = 0 + w0 + w1 + ...;
   t = (U - 1) * s with U from the next draw; first k with t < w[k] after
   subtracting the previous weights, else n - 1. */
int or_rng_poly_avx(or_rng* r, const double* w, int n)
{
    double s = 0;
    for (int k = 0; k < n; ++k) s += w[k];
    uint64_t u = or_rng_xorshift(r);
    double t = u2d((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    t = t - 1.0;
    t = t * s;
    int idx = n - 1, found = 0;
    for (int k = 0; k < n; ++k) {
        int b = t < w[k];
        t = t - w[k];
        if (b && !found) { idx = k; found = 1; }
    }
    return idx;
}

/* PolyLog(a, n), src/statistics.cpp:116-155 (overwrites a) */
int or_rng_polylog(or_rng* r, double* a, int n)
{
    double maxval = a[0];
    for (int i = 1; i < n; ++i) if (a[i] > maxval) maxval = a[i];
    double s = OR_MIN_FREQ * n;
    for (int i = 0; i < n; ++i) {
        double diff = a[i] - maxval;
        a[i] = diff < -23 ? OR_MIN_FREQ : exp(diff);
        s += a[i];
    }
    double t = or_rng_uniform(r, 0, s);
    for (int i = 0; i < n; ++i) {
        if (t < a[i]) return i;
        t -= a[i];
    }
    return n - 1;
}

/* Normal(mean, std): Box–Muller with the second variate cached in U1/U2;
   U1 == 0xFFFFFFFF in its low 32 bits means empty, src/statistics.cpp:185-201 */
double or_rng_normal(or_rng* r, double mean, double sd)
{
    if (u1_cached(r)) {
        double re = r->U1 * or_mysin(r->U2);
        u1_clear(r);
        return re * sd + mean;
    }
    volatile double v1 = or_rng_uniform(r, 0, 1);
    volatile double v2 = or_rng_uniform(r, 0, 1);
    r->U1 = or_mysqrt(-2.0 * log(v1));
    r->U2 = 2.0 * 3.14159265358979323846264338328 * v2;
    return r->U1 * or_mycos(r->U2) * sd + mean;
}

/* Gamma(alpha, beta): Marsaglia–Tsang with the alpha < 1 boost,
   src/statistics.cpp:212-242 */
double or_rng_gamma(or_rng* r, double alpha, double beta)
{
    if (alpha < 1) {
        volatile double v1 = or_rng_gamma(r, 1.0 + alpha, beta);
        volatile double v2 = pow(or_rng_uniform(r, 0, 1), 1.0 / alpha);
        return v1 * v2;
    }
    double t, v, u;
    double d = alpha - 0.333333333333333;
    double c = 1.0 / (sqrt(d) * 3.0);
    for (;;) {
        do {
            t = or_rng_normal(r, 0, 1);
            volatile double v1 = c * t;
            v = 1.0 + v1;
        } while (v <= 0);
        v = v * v * v;
        u = or_rng_uniform(r, 0, 1);
        if (u < 1.0 - 0.0331 * t * t * t * t) break;
        if (log(u) < 0.5 * t * t + d * (1.0 - v + log(v))) break;
    }
    return beta * d * v;
}

/* Next(max), NextAvoid(max, av), src/statistics.cpp:301-320 */
uint64_t or_rng_next(or_rng* r, uint64_t mx) { return or_rng_xorshift(r) % mx; }
uint64_t or_rng_next_avoid(or_rng* r, uint64_t mx, uint64_t av)
{
    uint64_t a = or_rng_next(r, mx - 1);
    if (a >= av) a++;
    return a;
}

/* Dirichlet(res, a, b, n): Gamma(a[i] + b[i]) then scale by 1/sum,
   src/statistics.cpp:284-296.  b is int32 (Ni) or int64 (Mi). */
void or_rng_dirichlet_ab(or_rng* r, double* res, const double* a, const void* b, int b_is_int64, int n)
{
    double s = 0;
    for (int i = 0; i < n; ++i) {
        double bi = b_is_int64 ? (double)((const int64_t*)b)[i] : (double)((const int32_t*)b)[i];
        double v = or_rng_gamma(r, a[i] + bi, 1);
        res[i] = v;
        s += v;
    }
    double inv = 1.0 / s;
    for (int i = 0; i < n; ++i) res[i] *= inv;
}

/* Dirichlet(res, a, n, f): Gamma(a[i] * f), src/statistics.cpp:269-281 */
void or_rng_dirichlet_af(or_rng* r, double* res, const double* a, int n, double f)
{
    double s = 0;
    for (int i = 0; i < n; ++i) {
        double v = or_rng_gamma(r, a[i] * f, 1);
        res[i] = v;
        s += v;
    }
    double inv = 1.0 / s;
    for (int i = 0; i < n; ++i) res[i] *= inv;
}

/* The same Dirichlet draws stored as REAL = float (BAYESIAN<float>):
   res[i] = (float)v while s adds the unrounded v, then res[i] *= (float)(1/s)
   in single precision (Mul(float*, float, n), src/math2.cpp:2088-2102). */
void or_rng_dirichlet_ab32(or_rng* r, double* res, const double* a, const void* b, int b_is_int64, int n)
{
    double s = 0;
    float v32[VPS_MAX_K * 8];
    for (int i = 0; i < n; ++i) {
        double bi = b_is_int64 ? (double)((const int64_t*)b)[i] : (double)((const int32_t*)b)[i];
        double v = or_rng_gamma(r, a[i] + bi, 1);
        v32[i] = (float)v;
        s += v;
    }
    float inv = (float)(1.0 / s);
    for (int i = 0; i < n; ++i) res[i] = (double)(v32[i] * inv);
}

void or_rng_dirichlet_af32(or_rng* r, double* res, const double* a, int n, double f)
{
    double s = 0;
    float v32[VPS_MAX_K * 8];
    for (int i = 0; i < n; ++i) {
        double v = or_rng_gamma(r, a[i] * f, 1);
        v32[i] = (float)v;
        s += v;
    }
    float inv = (float)(1.0 / s);
    for (int i = 0; i < n; ++i) res[i] = (double)(v32[i] * inv);
}

/* ------------------------------------------------------------------ */
/* RNG<float> (xorshift96), src/statistics.cpp:330-391                  */
/* ------------------------------------------------------------------ */
void or_rngf_init(or_rngf* r, uint64_t s, uint64_t salt)
{
    uint32_t h = or_murmur32(s, salt);
    r->x = 0x075BCD15u ^ h;
    r->y = 0x159A55E5u ^ (h << 3);
    r->z = 0x1F123BB5u ^ (h << 6);
}

uint32_t or_rngf_xorshift(or_rngf* r)
{
    uint32_t t;
    r->x ^= r->x << 16;
    r->x ^= r->x >> 5;
    r->x ^= r->x << 1;
    t = r->x;
    r->x = r->y;
    r->y = r->z;
    r->z = t ^ r->x ^ r->y;
    return r->z;
}

float or_rngf_uniform(or_rngf* r, float mn, float mx)
{
    uint32_t u = or_rngf_xorshift(r), b = 0x3F800000u;
    b |= u & 0x007FFFFFu;
    float re; memcpy(&re, &b, 4);
    return (re - 1.0f) * (mx - mn) + mn;
}

/* RNGAVX<float>::Poly, src/mathAVX.cpp:340-386: w[k] are the fast
   single-precision weights q*p + (float)MIN_FREQ (src/structureAVX.cpp:
   942-947); s = 0 + w0 + w1 + ... in float; U from the low 23 bits of the
   next xorshift96 output; t = (U - 1) * s; the first k with t < w[k] after
   subtracting the previous weights, else n - 1. */
int or_rngf_poly_avx(or_rngf* r, const float* w, int n)
{
    float s = 0.0f;
    for (int k = 0; k < n; ++k) s = s + w[k];
    uint32_t u = or_rngf_xorshift(r), b = 0x3F800000u | (u & 0x007FFFFFu);
    float t;
    memcpy(&t, &b, 4);
    t = t - 1.0f;
    t = t * s;
    int idx = n - 1, found = 0;
    for (int k = 0; k < n; ++k) {
        int lt = t < w[k];
        t = t - w[k];
        if (lt && !found) { idx = k; found = 1; }
    }
    return idx;
}

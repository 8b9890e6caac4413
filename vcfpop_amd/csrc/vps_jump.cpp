// vps_jump.cpp — jump-ahead for the xorshift128+ and xorshift96 streams.
//
// The per-locus stream of the Z pass draws one number per (individual,
// slot) in individual order (src/structure.cpp:1386-1413).  To split one
// locus over several wavefronts, a wavefront that starts at individual i0
// advances its stream by D = i0 * maxploidy steps.  One xorshift128+ step is
// linear over GF(2) on the 128-bit state (x, y), so D steps are one 128x128
// bit-matrix product M = T^D, built here by square-and-multiply and stored
// as 32 nibble tables of 16 entries (a device lookup is 32 table reads and
// XORs instead of D steps).
#include <string.h>
#include "vps_internal.h"

namespace vps {
namespace {

struct V128 { uint64_t lo, hi; };   // lo = x, hi = y
struct M128 { V128 col[128]; };     // col[j] = image of basis vector j

inline V128 step(V128 s)
{
    Xs128 r{s.lo, s.hi};
    r.next();
    return V128{r.x, r.y};
}

inline V128 mulv(const M128& A, V128 v)
{
    V128 r{0, 0};
    for (int b = 0; b < 64; ++b)
        if ((v.lo >> b) & 1) { r.lo ^= A.col[b].lo; r.hi ^= A.col[b].hi; }
    for (int b = 0; b < 64; ++b)
        if ((v.hi >> b) & 1) { r.lo ^= A.col[64 + b].lo; r.hi ^= A.col[64 + b].hi; }
    return r;
}

inline M128 mulm(const M128& A, const M128& B)
{
    M128 C;
    for (int j = 0; j < 128; ++j) C.col[j] = mulv(A, B.col[j]);
    return C;
}

M128 identity()
{
    M128 I;
    for (int j = 0; j < 128; ++j) {
        I.col[j].lo = j < 64 ? (1ull << j) : 0;
        I.col[j].hi = j < 64 ? 0 : (1ull << (j - 64));
    }
    return I;
}

M128 step_matrix()
{
    M128 T = identity();
    for (int j = 0; j < 128; ++j) T.col[j] = step(T.col[j]);
    return T;
}

}  // namespace

void jump_table(uint64_t D, uint32_t* out)
{
    M128 R = identity(), P = step_matrix();
    for (uint64_t d = D; d; d >>= 1) {
        if (d & 1) R = mulm(P, R);
        P = mulm(P, P);
    }
    for (int p = 0; p < 32; ++p)
        for (int v = 0; v < 16; ++v) {
            V128 acc{0, 0};
            for (int b = 0; b < 4; ++b)
                if ((v >> b) & 1) { acc.lo ^= R.col[4 * p + b].lo; acc.hi ^= R.col[4 * p + b].hi; }
            uint32_t* e = out + (p * 16 + v) * 4;
            e[0] = (uint32_t)acc.lo; e[1] = (uint32_t)(acc.lo >> 32);
            e[2] = (uint32_t)acc.hi; e[3] = (uint32_t)(acc.hi >> 32);
        }
}

void jump_tables(uint64_t D, int count, uint32_t* out)
{
    M128 P = step_matrix(), M = identity();
    for (uint64_t d = D; d; d >>= 1) {
        if (d & 1) M = mulm(P, M);
        P = mulm(P, P);
    }
    M128 R = identity();
    for (int c = 0; c < count; ++c) {
        uint32_t* o = out + (size_t)c * kJumpWords;
        for (int p = 0; p < 32; ++p)
            for (int v = 0; v < 16; ++v) {
                V128 acc{0, 0};
                for (int b = 0; b < 4; ++b)
                    if ((v >> b) & 1) { acc.lo ^= R.col[4 * p + b].lo; acc.hi ^= R.col[4 * p + b].hi; }
                uint32_t* e = o + (p * 16 + v) * 4;
                e[0] = (uint32_t)acc.lo; e[1] = (uint32_t)(acc.lo >> 32);
                e[2] = (uint32_t)acc.hi; e[3] = (uint32_t)(acc.hi >> 32);
            }
        R = mulm(M, R);
    }
}

// xorshift96 (RNG<float>): the 96-bit state (x, y, z) in the same table
// layout, state words 0..2 (nibbles 0..23; the entries of nibbles 24..31 are
// zero)
namespace {

struct M96 { uint32_t col[96][3]; };

M96 identity96()
{
    M96 I;
    memset(&I, 0, sizeof I);
    for (int j = 0; j < 96; ++j) I.col[j][j >> 5] = 1u << (j & 31);
    return I;
}

void mulv96(const M96& A, const uint32_t* v, uint32_t* r)
{
    r[0] = r[1] = r[2] = 0;
    for (int b = 0; b < 96; ++b)
        if ((v[b >> 5] >> (b & 31)) & 1) { r[0] ^= A.col[b][0]; r[1] ^= A.col[b][1]; r[2] ^= A.col[b][2]; }
}

M96 mulm96(const M96& A, const M96& B)
{
    M96 C;
    for (int j = 0; j < 96; ++j) mulv96(A, B.col[j], C.col[j]);
    return C;
}

M96 step_matrix96()
{
    M96 T = identity96();
    for (int j = 0; j < 96; ++j) {
        Xs96 r{T.col[j][0], T.col[j][1], T.col[j][2]};
        r.next();
        T.col[j][0] = r.x; T.col[j][1] = r.y; T.col[j][2] = r.z;
    }
    return T;
}

}  // namespace

void jump_tables96(uint64_t D, int count, uint32_t* out)
{
    M96 P = step_matrix96(), M = identity96();
    for (uint64_t d = D; d; d >>= 1) {
        if (d & 1) M = mulm96(P, M);
        P = mulm96(P, P);
    }
    M96 R = identity96();
    for (int c = 0; c < count; ++c) {
        uint32_t* o = out + (size_t)c * kJumpWords;
        memset(o, 0, kJumpWords * sizeof(uint32_t));
        for (int p = 0; p < 24; ++p)
            for (int v = 0; v < 16; ++v) {
                uint32_t acc[3] = {0, 0, 0};
                for (int b = 0; b < 4; ++b)
                    if ((v >> b) & 1)
                        for (int w = 0; w < 3; ++w) acc[w] ^= R.col[4 * p + b][w];
                uint32_t* e = o + (p * 16 + v) * 4;
                e[0] = acc[0]; e[1] = acc[1]; e[2] = acc[2]; e[3] = 0;
            }
        R = mulm96(M, R);
    }
}

void jump_apply_host(const uint32_t* tab, uint64_t& x, uint64_t& y)
{
    uint32_t st[4] = {(uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32)};
    uint32_t r[4] = {0, 0, 0, 0};
    for (int p = 0; p < 32; ++p) {
        uint32_t nib = (st[p >> 3] >> ((p & 7) * 4)) & 15u;
        const uint32_t* e = tab + (p * 16 + nib) * 4;
        for (int w = 0; w < 4; ++w) r[w] ^= e[w];
    }
    x = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
    y = (uint64_t)r[2] | ((uint64_t)r[3] << 32);
}

}  // namespace vps

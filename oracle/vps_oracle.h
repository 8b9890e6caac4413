/*
 * vps_oracle.h — CPU restatement of vcfpop's Bayesian clustering path
 * (`-structure`), used ONLY as test infrastructure: by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker.
 * Nothing in the product (vcfpop_amd/) links, loads or calls it.
 *
 * Semantics follow the reference's default x86 path (AVX, REAL = double):
 *   genotype passes  src/structureAVX.cpp:254-1187
 *   lane RNG         src/mathAVX.cpp:24-144  (== per-locus scalar RNG)
 *   scalar updates   src/structure.cpp:405-2045
 *   RNG/distributions src/statistics.cpp:57-321
 *   hashing          src/hash.cpp:9-16, 219-223, 429-469
 *   log accumulator  src/math2.h:716-902
 * Accumulation orders of the log-likelihood passes are those of one
 * subthread (structure_nsubthread = 1); the Z pass and integer counts are
 * thread-count invariant in the reference and here.
 *
 * Each function is pinned by tests/test_oracle_*.py (closed forms, invariants,
 * distribution tests); functions without an independent check: the locprior
 * r/eta/gamma and F-model Metropolis updates, per-cluster and local alpha,
 * lambda and Arrange -- parity unpinned.
 */
#ifndef VPS_ORACLE_H
#define VPS_ORACLE_H

#include <stdint.h>
#include "../include/vps.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants, src/global.h:50-80 ---- */
#define OR_NZERO            (-1e-10)
#define OR_MIN_FREQ         (1e-10)
#define OR_DOUBLE_UNDERFLOW (1e-100)
#define OR_DOUBLE_OVERFLOW  (1e100)
#define OR_SALT_INITADM     0xDA310D821043E1D3ull
#define OR_SALT_UPDATEQ     0xE7853EB09C92935Aull
#define OR_SALT_UPDATEZ     0x9B83516C63E249C3ull
#define OR_SALT_UPDATEP     0xD4EA1DB746DF14E3ull
#define OR_SALT_UPDATEA     0x49913AD9EDBD2EB8ull
#define OR_SALT_UPDATELOC   0x95EE7991226E2D14ull
#define OR_SALT_UPDATEALPHA 0x8A26E2C4D1BC5E54ull
#define OR_SALT_UPDATELAMBDA 0x67099E2FE2510092ull
#define OR_SALT_UPDATEF     0xF82104A5ADDAAF9Full

/* ---- hashing ---- */
uint32_t or_crc32c_u64(uint32_t prev, uint64_t v);
uint32_t or_hash_ulong(uint64_t v);
uint32_t or_hash_string(const void* s, int len);
uint32_t or_murmur2(uint32_t data, uint32_t seed);
uint64_t or_murmur64(uint64_t data, uint64_t seed);
uint32_t or_murmur32(uint64_t data, uint64_t seed);

/* ---- RNG<double>, src/statistics.h:15-87 ---- */
typedef struct or_rng { uint64_t x, y; double U1, U2; } or_rng;
void     or_rng_init(or_rng* r, uint64_t s, uint64_t salt);
uint64_t or_rng_xorshift(or_rng* r);
double   or_rng_uniform(or_rng* r, double mn, double mx);
double   or_rng_normal(or_rng* r, double mean, double sd);
double   or_rng_gamma(or_rng* r, double alpha, double beta);
uint64_t or_rng_next(or_rng* r, uint64_t mx);
uint64_t or_rng_next_avoid(or_rng* r, uint64_t mx, uint64_t av);
int      or_rng_poly_scalar(or_rng* r, const double* a, int n);   /* RNG::Poly, Σa + nε */
int      or_rng_poly_avx(or_rng* r, const double* w, int n);      /* RNGAVX::Poly, w already + ε */
int      or_rng_polylog(or_rng* r, double* a, int n);
void     or_rng_dirichlet_ab(or_rng* r, double* res, const double* a, const void* b, int b_is_int64, int n);
void     or_rng_dirichlet_af(or_rng* r, double* res, const double* a, int n, double f);

/* ---- RNG<float>, src/statistics.cpp:325-393 ---- */
typedef struct or_rngf { uint32_t x, y, z; } or_rngf;
void     or_rngf_init(or_rngf* r, uint64_t s, uint64_t salt);
uint32_t or_rngf_xorshift(or_rngf* r);
float    or_rngf_uniform(or_rngf* r, float mn, float mx);
int      or_rngf_poly_avx(or_rngf* r, const float* w, int n);
void     or_rng_dirichlet_ab32(or_rng* r, double* res, const double* a, const void* b, int b_is_int64, int n);
void     or_rng_dirichlet_af32(or_rng* r, double* res, const double* a, int n, double f);

/* ---- math ---- */
double or_mylog(double v);
double or_mysin(double x);
double or_mycos(double x);
double or_mysqrt(double v);
void   or_charge_log(int64_t* slog, double* prod, double val);
double or_close_log(int64_t slog, double prod);
double or_logprod_avx(const double* a, int64_t n);             /* LogProd(A,n), AVX path */
double or_logprod_sep(const double* a, int64_t n, int64_t sep); /* LogProd(A,n,sep) */
double or_sum_avx(const double* a, int64_t n);                 /* Sum(A,n), AVX path */
void   or_unify_avx(double* a, int64_t n);                     /* Unify(A,n), AVX path */

/* ---- dataset and chain ---- */
typedef struct or_data  or_data;
typedef struct or_chain or_chain;

or_data* or_data_create(const vps_dataset_desc* d);
void     or_data_destroy(or_data* d);
int64_t  or_data_kt(const or_data* d);
int      or_data_maxploidy(const or_data* d);
void     or_data_order(const or_data* d, int64_t* orig, int64_t* afo, int64_t* vt);
/* genotype id of individual i at sorted locus l */
int      or_data_gid(const or_data* d, int64_t l, int i);

void     or_set_threads(int n);

or_chain* or_chain_create(or_data* d, const vps_params* p, int K, int run_id);
void      or_chain_destroy(or_chain* c);
int       or_chain_get_state(or_chain* c, int which, void* host, size_t bytes);
int       or_chain_set_state(or_chain* c, int which, const void* host, size_t bytes);
void      or_chain_sizes(or_chain* c, int64_t* rlen, int64_t* nlnl);
uint64_t  or_chain_seed(or_chain* c);

/* sweep pieces, src/structure.cpp:2122-2158 */
void or_update_p(or_chain* c);
void or_update_q(or_chain* c);
void or_update_q_admix(or_chain* c);
void or_update_q_metro(or_chain* c);
void or_update_q_noadmix(or_chain* c);
void or_update_locpriori(or_chain* c);
void or_update_z(or_chain* c);
void or_update_z_admix(or_chain* c);
void or_update_z_noadmix(or_chain* c);
void or_update_alpha(or_chain* c);
void or_update_lambda(or_chain* c);
void or_update_a(or_chain* c);
void or_update_f(or_chain* c);
void or_record(or_chain* c);            /* Record: condition check + pass + accumulators */
double or_record_pass(or_chain* c);     /* the genotype pass alone (lnL), also MiSum += Mi */
void or_arrange(or_chain* c);
void or_sweep(or_chain* c, int n);      /* n full MCMC iterations from the current m */

/* the Metropolis log-ratio pass alone, for given proposals (bufNK1 rows) */
void or_q_metro_logratio(or_chain* c, const double* qprop, double* out_logratio);
/* the no-admix per-(i,k) log-likelihood pass alone (before PolyLog) */
void or_q_noadmix_loglik(or_chain* c, double* out_nk);

#ifdef __cplusplus
}
#endif
#endif

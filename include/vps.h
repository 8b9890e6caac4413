/*
 * vps.h — C ABI of the MI355X (gfx950) Bayesian-clustering engine that stands
 * behind vcfpop's `-structure` GPU seam.
 *
 * The reference exposes its GPU path in two layers:
 *   - an extern "C" runtime seam, src/structureCUDA.h:6-52 (WaitCUDA, MallocDeviceCUDA,
 *     CopyStructureMemory, CreateStreamCUDA, ...), and
 *   - five C++ members of BAYESIAN<REAL> selected when `usecuda` is set
 *     (src/structure.h:259-306): UpdateZAdmixCUDA, UpdateZNoAdmixCUDA,
 *     UpdateQNoAdmixCUDA, UpdateQMetroCUDA, RecordCUDA.
 * This header replaces both with one plain C interface: opaque handles,
 * plain pointers and sizes, int status codes (no exit() on error, unlike
 * checkErrors at src/structureCUDA.cpp:183-196).  Every entry point below names
 * the reference interface it replaces.
 *
 * Conventions
 *   - All functions return VPS_OK (0) or a positive VPS_ERR_* code; the
 *     message of the last failure on the calling thread is vps_last_error().
 *   - Host buffers are plain host memory; the library copies them.
 *   - A chain handle is used by one host thread at a time.  Different chains
 *     may be driven from different threads.
 *   - Locus index l everywhere is the index in the reference's sorted order
 *     (ascending genotype bitsize, ties by original index; QSLocus2,
 *     src/structure.cpp:2202-2273).  vps_dataset_create performs that sort.
 */
#ifndef VPS_H
#define VPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPS_ABI_VERSION 1

/* status codes */
#define VPS_OK               0
#define VPS_ERR_INVALID      1   /* bad argument or inconsistent sizes */
#define VPS_ERR_UNSUPPORTED  2   /* valid input outside what the engine supports */
#define VPS_ERR_DEVICE       3   /* HIP runtime error */
#define VPS_ERR_NOMEM        4   /* device or host allocation failed */
#define VPS_ERR_STATE        5   /* call not valid in the handle's current state */
#define VPS_ERR_BUDGET       6   /* serving: job does not fit the HBM budget */
#define VPS_ERR_NOTFOUND     7   /* serving: unknown job id */

/* engine limits */
#define VPS_MAX_PLOIDY       10  /* N_MAX_PLOIDY, src/global.h:27 */
#define VPS_MAX_K            16  /* clusters per chain on the device path */
#define VPS_MAX_ALLELES      64  /* alleles per locus on the device path */

/* ------------------------------------------------------------------------- */
/* Genotype image (host side).  Mirrors vcfpop's in-memory layout:            */
/*   per locus, N genotype ids of CeilLog2(ngeno) bits, LSB-first, starting   */
/*   on a byte boundary (GENO_READER, src/function.cpp:3499-3569;             */
/*   AddOffsetGT, src/misc.cpp:1624-1636), and a per-locus genotype table     */
/*   whose entries carry ploidy and the allele copies in stored order          */
/*   (GENOTYPE, src/function.h:53-149; GetAlleleArray, src/function.cpp:514).  */
/* ------------------------------------------------------------------------- */
typedef struct vps_dataset_desc {
    int32_t         n_ind;             /* N individuals                                   */
    int64_t         n_loc;             /* L loci, in the caller's (original) order         */
    int32_t         n_pop;             /* S sampling locations (populations)               */
    const int32_t*  ind_pop;           /* [N] population of each individual; individuals
                                          of one population are contiguous (IND::popid)    */
    const uint8_t*  geno_bits;         /* packed genotype ids                              */
    const uint64_t* geno_byte_offset;  /* [L] first byte of locus l's ids in geno_bits    */
    const uint8_t*  bitsize;           /* [L] bits per id = CeilLog2(ngeno) (>= 1)         */
    const int32_t*  n_allele;          /* [L] number of alleles k2 (SLOCUS::k)             */
    const int64_t*  gtab_offset;       /* [L+1] genotype-table range of locus l            */
    const uint8_t*  gt_ploidy;         /* [G] ploidy of each genotype; 0 = missing         */
    const int64_t*  gt_allele_offset;  /* [G] first allele copy in allele_pool             */
    const uint16_t* allele_pool;       /* allele copies, ascending, gt_ploidy[g] per entry */
    const char* const* ind_name;       /* [N] optional (NULL): names for output tables     */
    const char* const* pop_name;       /* [S] optional (NULL)                               */
    const char* const* loc_name;       /* [L] optional (NULL)                               */
    const char* const* pop_region;     /* [S] optional (NULL): the level-1 region of each
                                          population (the ancestry table's RegL1 column,
                                          src/structure.cpp:283); NULL = "Total"           */
} vps_dataset_desc;

/* ------------------------------------------------------------------------- */
/* Run parameters: the `-structure_*` option surface (src/parameters.cpp:    */
/* 584-627 defaults, :1230-1315 parser) plus engine options.                  */
/* Enumerated options keep the reference's 1-based encoding: 1 = yes, 2 = no */
/* (GetParString, src/string2.cpp:1107-1124), and the same derived booleans  */
/* (ReadPar, src/structure.cpp:405-464).                                     */
/* ------------------------------------------------------------------------- */
typedef struct vps_params {
    /* model */
    int32_t admix_val;        /* -structure_admix      (default 2 = no)   */
    int32_t locpriori_val;    /* -structure_locpriori  (default 2)        */
    int32_t f_val;            /* -structure_f          (default 2)        */
    /* MCMC */
    int32_t nburnin;          /* default 1000  */
    int32_t nreps;            /* default 10000 */
    int32_t nthinning;        /* default 10    */
    int32_t nruns;            /* default 1     */
    int32_t nadmburnin;       /* default 500   */
    /* lambda */
    double  lambda, stdlambda, maxlambda;      /* 1, 0.3, 10 */
    int32_t inferlambda_val;  /* default 2; the reference derives inferlambda = (val == 0) */
    int32_t difflambda_val;   /* default 2 */
    int32_t diversity_val;    /* default 2 */
    /* alpha */
    double  alpha, stdalpha, maxalpha, alphapriora, alphapriorb;  /* 1, 0.025, 10, 0.05, 0.001 */
    int32_t inferalpha_val;   /* default 1 */
    int32_t diffalpha_val;    /* default 0 (derived diffalpha = (val == 1)) */
    int32_t uniformalpha_val; /* default 1 */
    int32_t metrofreq;        /* default 10 */
    /* locprior */
    double  r, maxr, stdr, epseta, epsgamma;   /* 1, 20, 0.1, 0.025, 0.025 */
    /* F model */
    double  pmeanf, pstdf, stdf;               /* 0.01, 0.05, 0.05 */
    int32_t singlef_val;      /* default 2 */
    /* output / run control */
    int32_t writelnl_val;     /* default 2 */
    int32_t krange_min, krange_max;            /* 1, 5 */
    uint64_t g_seed;          /* -g_seed; run seed = HashULong(g_seed + id) */
    /* engine options (no counterpart in the reference) */
    int32_t rng_topology;     /* VPS_RNG_REFERENCE or VPS_RNG_PARALLEL           */
    int32_t precision;        /* VPS_PREC_FP64, _MIXED, _REAL32 or _FAST32       */
    int32_t z_chunk;          /* individuals per work chunk of the genotype passes:
                                 > 0 fixed; 0 (default) = timed candidates at chain
                                 creation; -1 = occupancy model only.  Results do
                                 not depend on it.                              */
    int32_t sweep_graphs;     /* 1: vps_chain_step launches each sweep as a cached
                                 HIP graph (one launch per sweep); 0 (default):
                                 kernel by kernel.  Results do not depend on it.
                                 Capture needs no other host thread to make HIP
                                 calls meanwhile; the server always uses 0.      */
} vps_params;

#define VPS_RNG_REFERENCE 0  /* P: 128 streams RNG(seed+m*128+b); Q: one stream RNG(seed+m) —
                                the reference's draw topology (src/structure.cpp:843-861, 990-1003) */
#define VPS_RNG_PARALLEL  1  /* P: one stream per (k,l); Q: one stream per individual —
                                same distributions, different draws; Z pass unchanged */

#define VPS_PREC_FP64  0     /* every quantity in FP64 (reference REAL=double)                     */
#define VPS_PREC_MIXED 1     /* FP32 per-genotype weights; FP64 lnL, Dirichlet, Metropolis          */
/* The reference's BAYESIAN<float> (-g_float=single): P, Q, the Metropolis
   proposals and the mean frequencies are single-precision values (each
   Dirichlet row is v_i rounded to float times (float)(1/sum v) in float;
   priors and initial values rounded to float; FreqM accumulated in float);
   the genotype passes compute in FP64 from them (-g_fastsingle=no). */
#define VPS_PREC_REAL32 2
/* REAL32 with -g_fastsingle=yes (vcfpop's default for -g_float=single):
   the Z pass draws from RNG<float> (xorshift96) per-locus streams with
   single-precision weights, and the Record mixture and Metropolis ratios are
   formed in single precision (src/structureAVX.cpp:725-760, 942-947,
   1148-1168).  The F model is not available with REAL32 or FAST32. */
#define VPS_PREC_FAST32 3

/* fills *p with the reference defaults (src/parameters.cpp:584-627, :432-445) */
void vps_params_default(vps_params* p);

/* ------------------------------------------------------------------------- */
/* Library                                                                    */
/* ------------------------------------------------------------------------- */
int         vps_abi_version(void);
const char* vps_last_error(void);
/* GetDeviceCountCUDA, src/structureCUDA.h:30 */
int         vps_device_count(int* n);

/* Plain device memory and device control, for callers that keep device
   buffers of their own: MallocDeviceCUDA, FreeCUDA, MemsetCUDA, MemcpyCUDA
   (src/structureCUDA.h:16-22), WaitCUDA (:8), ResetDeviceCUDA (:42),
   ShowDevicesCUDA (:40).  Status codes instead of exit(). */
int vps_device_alloc(int device, size_t bytes, void** out);
int vps_device_free(void* p);
int vps_device_memset(void* p, int value, size_t bytes);
/* to_device = 1: host -> device, 0: device -> host (MemcpyCUDA's todevice) */
int vps_device_memcpy(void* dst, const void* src, size_t bytes, int to_device);
int vps_device_synchronize(int device);
int vps_device_reset(int device);
int vps_device_name(int device, char* buf, size_t n);

/* ------------------------------------------------------------------------- */
/* Dataset: device-resident genotype image shared by every chain on a device. */
/* Replaces AllocMemoryCUDA + CopyStructureMemory (src/structureCUDA.h:24-32, */
/* src/structureCUDA.cpp:969-1094) and performs CalcBayesian's prologue:     */
/* locus sort and allele_freq_offset / KT (src/structure.cpp:2350-2395).      */
/* ------------------------------------------------------------------------- */
typedef struct vps_dataset vps_dataset;

int vps_dataset_create(const vps_dataset_desc* desc, int device, vps_dataset** out);
/* FreeStructureMemory + FreeMemoryCUDA, src/structureCUDA.h:22,34 */
int vps_dataset_destroy(vps_dataset* ds);

typedef struct vps_dataset_info {
    int32_t n_ind;
    int64_t n_loc;
    int64_t kt;             /* KT = total alleles over loci                      */
    int32_t max_ploidy;     /* maxploidy, src/global.h:189                        */
    int32_t min_ploidy;
    int32_t max_alleles;    /* maxK                                               */
    int32_t n_pop;
    int64_t n_tiles;        /* 64-locus tiles (one per wavefront)                 */
    int64_t device_bytes;   /* HBM held by the genotype image                     */
} vps_dataset_info;
int vps_dataset_get_info(const vps_dataset* ds, vps_dataset_info* info);
/* sorted-order tables: original index of sorted locus l, allele_freq_offset[l], vt[i] */
int vps_dataset_get_order(const vps_dataset* ds, int64_t* orig_index /*[L]*/,
                          int64_t* allele_freq_offset /*[L]*/, int64_t* vt /*[N]*/);

/* ------------------------------------------------------------------------- */
/* Chain: one MCMC run (BAYESIAN<double>, src/structure.h:50-307).           */
/* ------------------------------------------------------------------------- */
typedef struct vps_chain vps_chain;

/* ReadPar + InitFix + InitAdmix + InitLocPriori + InitFmodel
   (src/structure.cpp:405-708).  run_id is the reference's 1-based task id
   (src/structure.cpp:2402-2413), K the number of clusters. */
int vps_chain_create(vps_dataset* ds, const vps_params* par, int K, int run_id, vps_chain** out);
int vps_chain_destroy(vps_chain* ch);

/* MCMC loop body for sweeps m .. m+n-1 (src/structure.cpp:2122-2158):
   UpdateP, UpdateQ, UpdateLocPriori, UpdateZ, UpdateAlpha, UpdateLambda,
   UpdateA, UpdateF, Record.  Asynchronous on the chain's stream. */
int vps_chain_step(vps_chain* ch, int n_sweeps);
/* Arrange (src/structure.cpp:1862-1869) after the last sweep;
   VPS_ERR_STATE while no sampling sweep has been recorded (nr = 0) */
int vps_chain_finish(vps_chain* ch);
/* WaitCUDA (src/structureCUDA.h:8) */
int vps_chain_sync(vps_chain* ch);

/* chain state, copied to / from host buffers */
enum vps_state {
    VPS_STATE_FREQ   = 1,   /* double[K*KT]  current allele frequencies P     */
    VPS_STATE_Q      = 2,   /* double[N*K]   admixture proportions           */
    VPS_STATE_Z      = 3,   /* uint16[N]     cluster of each individual       */
    VPS_STATE_NI     = 4,   /* int32[K*KT]   allele copies per cluster        */
    VPS_STATE_MI     = 5,   /* int64[N*K]    allele copies per individual    */
    VPS_STATE_ALPHA  = 6,   /* double[K]                                       */
    VPS_STATE_LAMBDA = 7,   /* double[K]                                       */
    VPS_STATE_ITER   = 8,   /* int32[1]      next sweep index m              */
    VPS_STATE_BUFN1  = 9,   /* double[N]     Metropolis log ratio of the last QMetro */
    VPS_STATE_BUFNK1 = 10,  /* double[N*K]   no-admix log likelihoods of the last QNoAdmix */
    VPS_STATE_LNL    = 11,  /* double[1]     lnL of the last Record pass     */
    VPS_STATE_MISUM  = 12,  /* int64[N*K] before finish, double[N*K] after    */
    VPS_STATE_FREQM  = 13,  /* double[K*KT]                                    */
    VPS_STATE_ROUT   = 14,  /* double[rlen]  run summary accumulators          */
    VPS_STATE_FREQA  = 15,  /* double[3*KT]  F model ancestral frequencies    */
    VPS_STATE_FST    = 16,  /* double[2*K]   F[k] then f[k] = (1-F)/F          */
    VPS_STATE_LOC    = 17,  /* double[2 + K + 3*S*K]: r, nr-unused, Eta[K], Gamma[S*K], AlphaLocal[S*K], Di[S*K] */
    VPS_STATE_BINARYQ= 18,  /* int32[1]                                        */
    VPS_STATE_LNLTRACE=19,  /* double[nlnL]  per-record lnL trace (writelnl)   */
    VPS_STATE_QF     = 20   /* float[N*K]    read-only: Q rounded to single precision, the copy the
                               FP32-weight Z passes read (current after each sweep's Z pass) */
};
int vps_chain_get_state(vps_chain* ch, int which, void* host, size_t bytes);
int vps_chain_set_state(vps_chain* ch, int which, const void* host, size_t bytes);
/* rlen of VPS_STATE_ROUT and count of VPS_STATE_LNLTRACE */
int vps_chain_get_sizes(vps_chain* ch, int64_t* rlen, int64_t* nlnl);

/* Single passes at the chain's current sweep index m.  These are the
   drop-in counterparts of the BAYESIAN<REAL>::*CUDA members. */
int vps_pass_update_z_admix(vps_chain* ch);    /* UpdateZAdmixCUDA,   src/structureCUDA.cpp:2092-2119 */
int vps_pass_update_z_noadmix(vps_chain* ch);  /* UpdateZNoAdmixCUDA, src/structureCUDA.cpp:1987-2009 */
int vps_pass_update_q_noadmix(vps_chain* ch);  /* UpdateQNoAdmixCUDA, src/structureCUDA.cpp:2175-2214 */
int vps_pass_update_q_metro(vps_chain* ch);    /* UpdateQMetroCUDA,   src/structureCUDA.cpp:2269-2312 */
int vps_pass_record(vps_chain* ch);            /* RecordCUDA,         src/structureCUDA.cpp:2381-2410 */
int vps_pass_update_p(vps_chain* ch);          /* UpdatePCUDA (dead in the reference), src/structure.cpp:812-862 */
int vps_pass_update_q_admix(vps_chain* ch);    /* UpdateQAdmix,       src/structure.cpp:986-1004 */
int vps_pass_update_alpha(vps_chain* ch);      /* UpdateAlpha,        src/structure.cpp:1501-1594 */

/* Record's likelihood pass with no device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops).  It runs the genotype pass of
   vps_pass_record on caller-owned device buffers and writes the lnL to the
   device double *lnl:
     freq  [K*KT + 2] doubles, P in sorted locus order (the 2 trailing values
           are read by paired loads and never charged);
     q     [N*K] doubles (admixture form), or
     z     [N] uint16 cluster of each individual (binary-Q form; q unused).
   The pass's block partials live in one static device array per GPU (4 MiB,
   held until the process ends): a dataset needing more partials than it holds
   (about N = 100,000 with L = 1,000,000) is refused with VPS_ERR_UNSUPPORTED.
   Calls that overlap on one GPU race on that array.  Asynchronous on `stream`
   (a hipStream_t; NULL = the null stream). */
int vps_record_lnl_noalloc(const vps_dataset* ds, int K, const double* freq, const double* q,
                           const uint16_t* z, double* lnl, void* stream);

/* ------------------------------------------------------------------------- */
/* Result tables (WriteStructure / WriteStructureSummary,                     */
/* src/structure.cpp:88-401): the run file "<prefix>.structure.k=K_rep=R_id=I */
/* .txt" (and ".lnl.txt" with writelnl) with the reference's sections and     */
/* number format.  The RegL1 column reads the dataset's pop_region names, or  */
/* "Total" (the reference's value for a population without a region).         */
/* ------------------------------------------------------------------------- */
typedef struct vps_output_options {
    int32_t     decimal;      /* digits after the point (-decimal, default 5)          */
    int32_t     scientific;   /* 1 = %e (-scientific, default 2)                       */
    char        delimiter;    /* '\t' (default) or ',' (-delimiter)                    */
    char        pad[7];
    const char* linebreak;    /* "\r\n" (default) or "\n" (-linebreak)                 */
    int32_t     diversity;    /* 1 = per-locus diversity table (-structure_diversity)  */
    int32_t     rep;          /* replicate number of the run                           */
} vps_output_options;

typedef struct vps_run_summary {
    int32_t id, k, rep, pad;
    double  mean_lnl, var_lnl, lnpd;
} vps_run_summary;

/* a finished run's host-side state, for writing tables without a device */
typedef struct vps_run_state {
    const double* Q;          /* [N*K] final admixture proportions              */
    const double* ancestry;   /* [N*K] normalised MiSum (after finish)          */
    const double* FreqM;      /* [K*KT] mean allele frequencies                 */
    const double* rout;       /* [rout_len] arranged run accumulators           */
    int64_t       rout_len;
    const double* lnltrace;   /* [lnltrace_len] or NULL                         */
    int64_t       lnltrace_len;
    uint64_t      seed;       /* run seed, HashULong(g_seed + id)               */
} vps_run_state;

void vps_output_options_default(vps_output_options* o);
/* after vps_chain_finish; fills *summary (may be NULL) for the summary table */
int  vps_chain_write_tables(vps_chain* ch, const char* prefix, const vps_output_options* opt, vps_run_summary* summary);
/* the same tables from host arrays (any dataset, including device = -1) */
int  vps_write_tables_host(const vps_dataset* ds, const vps_params* par, int K, int run_id, const vps_run_state* st,
                           const char* prefix, const vps_output_options* opt, vps_run_summary* summary);
/* The head of a result file (OpenResFile, src/file.cpp:269-287, and
   WriteParameters, src/parameters.cpp:1408-1418):
     ##Bayesian clustering calculated by vcfpop v1.08b
     #  Input: <input>
     #  Output: <output>
     #  Time: YYYY-MM-DD HH:MM:SS        (local time)
     #  Parameter: <argv[0]>
     #    <argv[i]>                      for each later argument that starts
                                         (case-insensitively) with -g_ (but not
                                         -g_indtext or -g_indtab), with
                                         -structure, or with -f_ when filters on
   argv is vcfpop's argument list: the command line, or for a parameter file
   (-p=) the file's blank-separated tokens with their quotes trimmed
   (SetParameters, src/parameters.cpp:1345-1362). */
#define VPS_VCFPOP_VERSION "1.08b"     /* VERSION, src/global.h:10 */
typedef struct vps_result_header {
    const char*        input;      /* -g_input value                          */
    const char*        output;     /* -g_output value                         */
    const char* const* argv;       /* [argc] vcfpop's arguments               */
    int32_t            argc;
    int32_t            filters;    /* 1 = -f given: -f_* arguments are listed */
    int64_t            time;       /* seconds since the epoch; 0 = now        */
} vps_result_header;

/* WriteStructureSummary (src/structure.cpp:380-401): the header (hdr NULL:
   the title line only), then one row per run (ID, k, rep, mean lnL, var lnL,
   lnP(D)); like the reference, no line break after the last row. */
int  vps_write_summary(const char* path, const vps_run_summary* runs, int n, const vps_output_options* opt,
                       const vps_result_header* hdr);

/* ------------------------------------------------------------------------- */
/* Batched chains.  The reference runs K values x replicates as a task pool of */
/* host threads, each leasing a device slot for one chain and driving it with */
/* synchronous per-pass calls (CalcBayesian, BayesianThread, ManageCUDA;      */
/* src/structure.cpp:2397-2501, 711-756).  A batch advances any number of      */
/* chains on one device together: one HIP-graph launch per sweep, with every   */
/* chain a parallel branch.  Results equal those of vps_chain_step on each     */
/* chain alone, bit for bit.  Chains stay owned by the caller and must outlive */
/* the batch; chain calls may be mixed with batch steps.                       */
/* ------------------------------------------------------------------------- */
typedef struct vps_batch vps_batch;
int vps_batch_create(vps_chain* const* chains, int n, vps_batch** out);
/* n_sweeps sweeps of every chain; asynchronous */
int vps_batch_step(vps_batch* b, int n_sweeps);
int vps_batch_sync(vps_batch* b);
int vps_batch_destroy(vps_batch* b);
/* graph launches, graphs captured, sweeps launched without a graph */
int vps_batch_stats(const vps_batch* b, int64_t* graph_launches, int64_t* captures, int64_t* eager_sweeps);

/* ------------------------------------------------------------------------- */
/* Serving: a job server on one device.  The reference queues runs for a free */
/* device slot and holds the slot for a whole chain (ManageCUDA, BayesianThread,*/
/* src/structure.cpp:711-756, 2468-2501).  The server admits jobs while their  */
/* device memory fits an HBM budget and an active-job limit, builds each job's */
/* dataset and chain on a loader thread, and runs all admitted jobs at once,   */
/* each on its own stream in slices of sweeps (continuous batching).  A job's  */
/* results are bit-identical to the same job run alone.                        */
/* ------------------------------------------------------------------------- */
typedef struct vps_server vps_server;

typedef struct vps_server_config {
    int32_t device;
    int32_t max_active;         /* jobs loaded or running at once (0 = no limit) */
    int64_t hbm_budget_bytes;   /* device bytes of all admitted jobs              */
    int32_t slice_sweeps;       /* sweeps issued per job per scheduling step      */
    int32_t loaders;            /* threads building admitted jobs' datasets and chains */
} vps_server_config;

#define VPS_JOB_PENDING  0
#define VPS_JOB_LOADING  1
#define VPS_JOB_RUNNING  2
#define VPS_JOB_DONE     3
#define VPS_JOB_FAILED   4

typedef struct vps_job_info {
    int64_t job_id;
    int32_t state;              /* VPS_JOB_*                                      */
    int32_t status;             /* VPS_OK, or the failure code                    */
    int32_t n_ind, K;
    int64_t n_loc;
    int64_t sweeps_done, sweeps_total;
    int64_t device_bytes;       /* measured once loaded                           */
    int64_t estimated_bytes;    /* reservation held against the budget            */
    double  t_submit, t_start, t_end;   /* seconds on the server clock; t_start = admission */
    int64_t rout_len, lnltrace_len;     /* result lengths once done                */
} vps_job_info;

typedef struct vps_server_stats_t {
    int64_t pending, loading, running, completed, failed;
    int64_t reserved_bytes, peak_reserved_bytes;
    double  uptime_s;
} vps_server_stats_t;

void vps_server_config_default(vps_server_config* c);
int  vps_server_create(const vps_server_config* cfg, vps_server** out);
/* copies the genotype image; VPS_ERR_BUDGET if the job can never fit.
   n_sweeps is the run length nburnin + nreps (+ nadmburnin for the
   no-admixture model without a location prior); 0 takes that length, and
   any other value is VPS_ERR_INVALID */
int  vps_server_submit(vps_server* s, const vps_dataset_desc* desc, const vps_params* par, int K, int run_id,
                       int64_t n_sweeps, int64_t* job_id);
/* blocks until the job is done (timeout_s < 0: no limit); VPS_ERR_STATE on timeout */
int  vps_server_wait(vps_server* s, int64_t job_id, double timeout_s);
int  vps_server_job_info(vps_server* s, int64_t job_id, vps_job_info* info);
/* results of a finished job: VPS_STATE_Q, VPS_STATE_MISUM (normalised ancestry),
   VPS_STATE_ROUT, VPS_STATE_LNLTRACE, VPS_STATE_LNL */
int  vps_server_job_result(vps_server* s, int64_t job_id, int which, void* host, size_t bytes);
/* forget a finished job's results */
int  vps_server_release(vps_server* s, int64_t job_id);
int  vps_server_stats(vps_server* s, vps_server_stats_t* st);
/* stops the threads; running jobs are abandoned */
int  vps_server_destroy(vps_server* s);

/* ------------------------------------------------------------------------- */
/* Timing hook for measurement: average device time of the last launches of  */
/* a named kernel family, measured with HIP events on the chain's stream.     */
/* ------------------------------------------------------------------------- */
/* every = n: time every n-th Z pass (1 = all, 0 = off); the passes of Record
   sweeps, which also sum the lnL, are neither timed nor counted */
int vps_chain_enable_timing(vps_chain* ch, int every);
/* kernel: 0 = Z admix pass; returns total ms and number of timed launches */
int vps_chain_kernel_time(vps_chain* ch, int kernel, double* total_ms, int64_t* launches);

/* ------------------------------------------------------------------------- */
/* Test hooks (no counterpart in the reference).  A dataset created with     */
/* device = -1 is built on the host only (no HIP call) and keeps its packed   */
/* tile layout, so the layout can be checked without a GPU.                   */
/* ------------------------------------------------------------------------- */
/* decode genotype (sorted locus l, individual i) from the tile layout */
int vps_debug_decode(const vps_dataset* ds, int64_t l, int32_t i, int32_t* ploidy, uint16_t* alleles /*[VPS_MAX_PLOIDY]*/);
/* advance a xorshift128+ state (x, y) by D steps with the GF(2) jump table */
int vps_debug_jump(uint64_t D, uint64_t* x, uint64_t* y);
/* advance a xorshift96 state (x, y, z) (RNG<float>) by D steps with its table */
int vps_debug_jump96(uint64_t D, uint32_t* x, uint32_t* y, uint32_t* z);
/* a chain's individual chunking of the genotype passes: G chunks of Nc individuals */
int vps_debug_chunking(const vps_chain* ch, int32_t* G, int32_t* Nc);
/* wavefronts per block of the multiallelic (GEN) Z-pass kernel for the chain's
   chunking: 4, or 8 when LDS limits the resident blocks and G >= 8 */
int vps_debug_gen_waves(const vps_chain* ch, int32_t* waves);
/* Marsaglia–Tsang's second acceptance test (src/statistics.cpp:239) as the
   device's Gamma draws decide it, from single-precision estimates with an
   error bound and the exact double form inside the bound: accept = 1/0 for n
   values, and *n_exact = how many needed the exact form */
int vps_debug_mt_accept(const double* u, const double* t, const double* v, const double* d, int64_t n,
                        uint8_t* accept, int64_t* n_exact);
/* one warm-up plus `iters` launches of a plain device copy of `bytes` bytes
   (scales hardware traffic counters); average ms per copy */
int vps_debug_copy(int device, size_t bytes, int iters, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* VPS_H */

/*
 * vps_synth.h — deterministic synthetic genotype data in vcfpop's packed
 * layout, for benchmark workloads and the serving job stream.  No
 * counterpart in the reference (which reads VCF/BCF/genepop files).
 */
#ifndef VPS_SYNTH_H
#define VPS_SYNTH_H
#include <stdint.h>
#include "vps.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vps_synth_config {
    int32_t n_ind;                 /* N                                                    */
    int64_t n_loc;                 /* L                                                    */
    int32_t k_true;                /* generating clusters                                  */
    double  alpha_q;               /* Q_i ~ Dirichlet(alpha_q)                             */
    double  fst;                   /* 0: P ~ Dir(1,1); >0: Beta drift from p0 ~ U(.05,.95) */
    double  missing;               /* probability a genotype is missing                    */
    double  frac_multi;            /* fraction of multiallelic loci                        */
    int32_t min_alleles, max_alleles;  /* allele count range of multiallelic loci          */
    double  multi_dir;             /* multiallelic P ~ Dirichlet(multi_dir)                */
    const int32_t* ploidy_levels;  /* [n_ploidy] ploidy mix (NULL: all diploid)            */
    const double*  ploidy_frac;    /* [n_ploidy]                                           */
    int32_t n_ploidy;
    int32_t n_pop;                 /* sampling locations (individuals split evenly)        */
    uint64_t seed;
} vps_synth_config;

typedef struct vps_synth vps_synth;

void vps_synth_default(vps_synth_config* c);
int  vps_synth_create(const vps_synth_config* cfg, vps_synth** out);
/* a view of the generated arrays, valid while the generator lives */
int  vps_synth_desc(const vps_synth* s, vps_dataset_desc* desc);
/* generating admixture proportions, [N * k_true] */
int  vps_synth_truth(const vps_synth* s, double* q);
int  vps_synth_sizes(const vps_synth* s, int64_t* packed_bytes, int64_t* n_genotypes, int64_t* n_alleles);
int  vps_synth_destroy(vps_synth* s);

#ifdef __cplusplus
}
#endif
#endif

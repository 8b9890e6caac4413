/*
 * vps_ingest.h — genotype text files read into the packed genotype image
 * the engine takes (vps_dataset_desc).  Replaces the part of vcfpop's input
 * path that the Bayesian clustering reads: the individual reader
 * (IND::polygene, src/load.cpp:3047-3120), the per-locus genotype table in
 * order of first appearance (src/load.cpp:3093-3110), the allele indexing
 * (IndexAlleleLength, src/load.cpp:1535-1590), the default locus filter
 * (fewer than two alleles, src/load.cpp:1575; RemoveLocus,
 * src/filter.cpp:137-160) and the CeilLog2 id packing
 * (src/math2.h:503-512, AddOffsetGT src/misc.cpp:1624-1636).
 */
#ifndef VPS_INGEST_H
#define VPS_INGEST_H
#include <stdint.h>
#include "vps.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vps_image vps_image;

/* Read a polygene-format file (tab-separated: ID, population, ploidy, then
   one field per locus of comma-separated integer allele identifiers; an
   empty field is a missing genotype).  Errors (unreadable file, a field
   with a number of alleles other than the ploidy, ploidy outside 1..10,
   allele identifiers outside 0..65534) return VPS_ERR_INVALID with a
   message naming the line. */
int vps_image_read_polygene(const char* path, vps_image** out);
/* a view of the image's arrays and names, valid while the image lives */
int vps_image_desc(const vps_image* im, vps_dataset_desc* desc);
/* loci in the file, loci kept (the others have fewer than two alleles),
   bytes of packed ids, genotype-table entries and allele copies */
int vps_image_sizes(const vps_image* im, int64_t* loci_read, int64_t* loci_kept, int64_t* packed_bytes,
                    int64_t* n_genotypes, int64_t* n_alleles);
int vps_image_destroy(vps_image* im);

#ifdef __cplusplus
}
#endif
#endif

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
/* Read a VCF file, plain or gzip/BGZF-compressed (LoadVCF, src/load.cpp:
   2111-2330; IND::AddVCFGenotype, :3495-3560; the VCF LOCUS,
   src/function.cpp:3166-3240).  Samples are the #CHROM line's columns after
   FORMAT; a variant has k = REF + ALT entries alleles (an ALT of "." counts
   as one) and its GT numbers are the allele indices; the locus name is ID,
   or POS_REF_ALT without one.  A GT of ploidy v (1 + its '/' and '|') that
   starts with '.' or holds a '.' allele is the missing genotype of ploidy v.
   Genotype ids follow first appearance over samples (each sample's missing
   genotype of its ploidy first); loci whose typed genotypes carry fewer than
   two distinct alleles are dropped (src/filter.cpp:100-133).  Every sample
   is in one population, "Total", until vps_image_assign_indtext.  Errors (no
   #CHROM line, a FORMAT without GT, a line with another number of fields, an
   empty or malformed GT, an allele index >= k, ploidy outside 1..10) return
   VPS_ERR_INVALID with a message naming the line and sample; more than
   VPS_MAX_ALLELES alleles at a kept locus is VPS_ERR_UNSUPPORTED.  The
   -f_* filters, -ad and phased (-haplotype) reading are not applied. */
int vps_image_read_vcf(const char* path, vps_image** out);
/* Populations and regions from vcfpop's -g_indtext option, replacing those
   of the file's population column (which vcfpop does not read).  NULL or an
   empty text puts every individual in one population, "Total", as vcfpop
   does without -g_indtext.  The text is blank-separated tokens: before the
   first "#REG", each "name:m1,m2,..." is a population of individuals named,
   given as "#i" (1-based; a name is tried first) or as a range "#i-#j";
   after each "#REG", "name:m1,..." is a region of the level below (its
   populations, or regions), named or given as "#i" / "#i-#j" (1-based).
   Individuals not listed form "DefPop"; populations not in a region form
   "DefRegL1".  As vcfpop does after loading (AssignInds, src/load.cpp:88-160;
   AssignVInds and SetVReg, src/filter.cpp:166-398): empty populations are
   dropped, populations are ordered by a walk of the region tree, and the
   individuals move to population order (their order within a population
   kept; the genotype tables are unchanged, the packed ids move with them).
   With a single population left it is named "Total".  Each population's
   level-1 region becomes its pop_region ("Total" when the populations fall
   in fewer than two level-1 regions).  A duplicate name, an individual or
   population listed twice, an unknown name or an index out of range is
   VPS_ERR_INVALID with a message naming it.  The image's descriptor must be
   fetched again afterwards. */
int vps_image_assign_indtext(vps_image* im, const char* indtext);
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

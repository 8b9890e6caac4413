// vps_ingest.cpp — genotype text files read into the packed genotype image
// (vps_dataset_desc): the polygene format.
//
// The reference's reader, restated (src/load.cpp):
//   IND::polygene (:3047-3120): a line per individual; name, population and
//     ploidy, then per locus the allele identifiers separated by commas (an
//     empty field is missing).  Alleles are sorted: genotypes are unphased.
//   Genotype ids per locus in order of first appearance over individuals;
//     an individual registers the missing genotype of its ploidy before its
//     own genotype (:3093-3110), so every ploidy present has a missing entry.
//   IndexAlleleLength (:1535-1590): allele indices in order of first
//     appearance over the genotype table (copies in stored order, missing
//     entries skipped); each genotype's copies are then re-sorted by index.
//     A locus with fewer than two alleles fails the locus filter (:1575) and
//     is removed before any analysis (RemoveLocus, src/filter.cpp:137-160).
//   Ids are packed at CeilLog2(#genotypes) bits, LSB-first from a byte-aligned
//     start (src/math2.h:503-512, src/misc.cpp:1624-1636).
// One deviation: a single allele in a field of ploidy > 1 is an error here;
// the reference pads such a field with missing copies.
//
// The VCF format (plain or gzip/BGZF-compressed text), restated from the
// reference's loader (src/load.cpp:2111-2330, LoadVCF; :3495-3560,
// IND::AddVCFGenotype; src/function.cpp:3166-3240, the LOCUS of a VCF line;
// src/string2.cpp:41-60, ReadVCFGenoString; src/load.h:86-93, CountPloidy):
//   Samples are the #CHROM line's columns after FORMAT.  A variant's alleles
//     are REF and the comma-separated ALT entries: k = commas in ALT + 2
//     (an ALT of "." counts as one allele), and allele indices are the GT
//     numbers themselves.  The locus name is ID, or "POS_REF_ALT" when ID is
//     "." or empty.
//   A sample's genotype is its GT subfield: ploidy = 1 + the '/' and '|'
//     after the first character (1..10, else an error); a genotype that
//     starts with '.' or has any '.' allele is the missing genotype of that
//     ploidy; alleles are sorted (unphased); an allele index >= k is an
//     error.  Genotype ids follow first appearance over samples, each sample
//     registering the missing genotype of its ploidy before its own.
//   ApplyFilter (src/filter.cpp:100-133, run on every input) then drops the
//     loci whose typed genotypes carry fewer than two distinct alleles; k
//     stays the REF/ALT count, so unobserved alleles keep their P entries.
//   The options that filter or rewrite genotypes (-f_*, -ad, -haplotype
//     phasing) are not applied.
#include <zlib.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>
#include "../../include/vps_ingest.h"
#include "vps_internal.h"

struct vps_image {
    int32_t N = 0;
    int64_t L = 0, L_read = 0;
    int32_t S = 0;
    std::vector<uint8_t> bits;
    std::vector<uint64_t> byteoff;
    std::vector<uint8_t> bitsize;
    std::vector<int32_t> nallele;
    std::vector<int64_t> gtoff;
    std::vector<uint8_t> gtploidy;
    std::vector<int64_t> gtaoff;
    std::vector<uint16_t> pool;
    std::vector<int32_t> pop;
    std::vector<std::string> ind_names, pop_names, loc_names, reg_names;
    std::vector<const char*> ind_ptr, pop_ptr, loc_ptr, reg_ptr;
};

using namespace vps;

namespace {

int ceil_log2(int64_t v)
{
    if (v <= 1) return 1;
    int b = 0;
    uint64_t x = (uint64_t)(v - 1);
    while (x) { ++b; x >>= 1; }
    return b;
}

std::vector<std::string> split(const std::string& s, char sep)
{
    std::vector<std::string> out;
    size_t a = 0;
    for (;;) {
        const size_t b = s.find(sep, a);
        out.emplace_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return out;
}

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

bool parse_int(const std::string& s, int64_t lo, int64_t hi, int64_t* v)
{
    if (s.empty() || s.size() > 18) return false;
    int64_t x = 0;
    for (char ch : s) {
        if (ch < '0' || ch > '9') return false;
        x = x * 10 + (ch - '0');
    }
    if (x < lo || x > hi) return false;
    *v = x;
    return true;
}

// a genotype of one locus: ploidy and sorted allele identifiers (none if missing)
struct Geno {
    int v = 0;
    bool missing = false;
    std::vector<uint16_t> al;
};

std::string geno_key(const Geno& g)
{
    std::string k(1, (char)g.v);
    k.push_back(g.missing ? 'm' : 't');
    for (uint16_t a : g.al) { k.push_back((char)(a & 0xFF)); k.push_back((char)(a >> 8)); }
    return k;
}

}  // namespace

extern "C" int vps_image_read_polygene(const char* path, vps_image** out)
{
    if (!path || !out) return fail(VPS_ERR_INVALID, "null argument");
    *out = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(VPS_ERR_INVALID, "cannot open %s", path);
    std::string text;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    fclose(f);

    std::vector<std::string> lines = split(text, '\n');
    size_t li = 0;
    while (li < lines.size() && trim(lines[li]).empty()) ++li;
    if (li == lines.size()) return fail(VPS_ERR_INVALID, "%s: empty file", path);
    const std::vector<std::string> head = split(trim(lines[li]), '\t');
    if (head.size() < 4) return fail(VPS_ERR_INVALID, "%s: the header needs ID, Pop, Ploidy and at least one locus", path);
    const int64_t L = (int64_t)head.size() - 3;
    ++li;

    vps_image* im = new vps_image();
    im->L_read = L;
    // per locus: genotype table (first appearance) and each individual's id
    std::vector<std::unordered_map<std::string, uint32_t>> gid(L);
    std::vector<std::vector<Geno>> gtab(L);
    std::vector<std::vector<uint32_t>> ids(L);
    std::map<std::string, int32_t> pop_index;
    for (; li < lines.size(); ++li) {
        const std::string line = trim(lines[li]);
        if (line.empty()) continue;
        const std::vector<std::string> fld = split(line, '\t');
        const int lineno = (int)li + 1;
        if ((int64_t)fld.size() < 3 + L) {
            delete im;
            return fail(VPS_ERR_INVALID, "%s line %d: %zu fields, expected %lld", path, lineno, fld.size(), (long long)(3 + L));
        }
        for (size_t x = 3 + L; x < fld.size(); ++x)
            if (!trim(fld[x]).empty()) {
                delete im;
                return fail(VPS_ERR_INVALID, "%s line %d: more fields than loci", path, lineno);
            }
        int64_t v = 0;
        if (!parse_int(trim(fld[2]), 1, VPS_MAX_PLOIDY, &v)) {
            delete im;
            return fail(VPS_ERR_INVALID, "%s line %d: ploidy '%s' outside 1..%d", path, lineno, fld[2].c_str(), VPS_MAX_PLOIDY);
        }
        im->ind_names.push_back(trim(fld[0]));
        const std::string pname = trim(fld[1]);
        auto pit = pop_index.find(pname);
        if (pit == pop_index.end()) {
            pit = pop_index.emplace(pname, (int32_t)im->pop_names.size()).first;
            im->pop_names.push_back(pname);
        }
        im->pop.push_back(pit->second);
        for (int64_t l = 0; l < L; ++l) {
            const std::string gs = trim(fld[3 + l]);
            Geno g;
            g.v = (int)v;
            if (gs.empty()) {
                g.missing = true;
            } else {
                const std::vector<std::string> al = split(gs, ',');
                if ((int64_t)al.size() != v) {
                    delete im;
                    return fail(VPS_ERR_INVALID, "%s line %d locus %lld: %zu alleles for ploidy %lld (allelic phenotypes are not supported)",
                                path, lineno, (long long)(l + 1), al.size(), (long long)v);
                }
                for (const std::string& a : al) {
                    int64_t x = 0;
                    if (!parse_int(trim(a), 0, 0xFFFE, &x)) {
                        delete im;
                        return fail(VPS_ERR_INVALID, "%s line %d locus %lld: allele '%s' is not an integer in 0..65534",
                                    path, lineno, (long long)(l + 1), a.c_str());
                    }
                    g.al.push_back((uint16_t)x);
                }
                std::sort(g.al.begin(), g.al.end());   // unphased
            }
            // the missing genotype of this ploidy first, then the genotype
            Geno miss;
            miss.v = (int)v;
            miss.missing = true;
            uint32_t id = 0;
            for (const Geno* e : {&miss, &g}) {
                const std::string k = geno_key(*e);
                auto it = gid[l].find(k);
                if (it == gid[l].end()) {
                    it = gid[l].emplace(k, (uint32_t)gtab[l].size()).first;
                    gtab[l].push_back(*e);
                }
                id = it->second;
                if (g.missing) break;
            }
            ids[l].push_back(id);
        }
    }
    const int64_t N = (int64_t)im->ind_names.size();
    if (N == 0) { delete im; return fail(VPS_ERR_INVALID, "%s: no individuals", path); }
    if (N > 0x7FFFFFFF) { delete im; return fail(VPS_ERR_UNSUPPORTED, "%s: too many individuals", path); }
    im->N = (int32_t)N;
    im->S = (int32_t)im->pop_names.size();

    // allele indices by first appearance over the genotype table, copies
    // re-sorted by index; loci with fewer than two alleles are dropped
    im->gtoff.push_back(0);
    uint64_t bytes = 0;
    for (int64_t l = 0; l < L; ++l) {
        std::unordered_map<uint16_t, uint16_t> aix;
        for (const Geno& g : gtab[l])
            for (uint16_t a : g.al)
                if (!aix.count(a)) aix.emplace(a, (uint16_t)aix.size());
        const int k = (int)aix.size();
        if (k < 2) continue;
        if (k > VPS_MAX_ALLELES) {
            delete im;
            return fail(VPS_ERR_UNSUPPORTED, "%s locus %s: %d alleles, the device path takes at most %d", path,
                        head[3 + l].c_str(), k, VPS_MAX_ALLELES);
        }
        const int64_t ng = (int64_t)gtab[l].size();
        const int b = ceil_log2(ng);
        if (b > 24) { delete im; return fail(VPS_ERR_UNSUPPORTED, "%s locus %s: %lld genotypes", path, head[3 + l].c_str(), (long long)ng); }
        im->loc_names.push_back(head[3 + l]);
        im->nallele.push_back(k);
        im->bitsize.push_back((uint8_t)b);
        im->byteoff.push_back(bytes);
        for (const Geno& g : gtab[l]) {
            im->gtploidy.push_back(g.missing ? 0 : (uint8_t)g.v);
            im->gtaoff.push_back((int64_t)im->pool.size());
            std::vector<uint16_t> m;
            for (uint16_t a : g.al) m.push_back(aix[a]);
            std::sort(m.begin(), m.end());
            im->pool.insert(im->pool.end(), m.begin(), m.end());
        }
        im->gtoff.push_back((int64_t)im->gtploidy.size());
        const uint64_t nb = ((uint64_t)b * N + 7) / 8;
        im->bits.resize(bytes + nb, 0);
        for (int64_t i = 0; i < N; ++i) {
            const uint64_t v = ids[l][i];
            const uint64_t bit = (uint64_t)i * b;
            for (int x = 0; x < b; ++x)
                if (v >> x & 1) im->bits[bytes + ((bit + x) >> 3)] |= (uint8_t)(1u << ((bit + x) & 7));
        }
        bytes += nb;
    }
    im->L = (int64_t)im->nallele.size();
    if (im->L == 0) { delete im; return fail(VPS_ERR_INVALID, "%s: every locus has fewer than two alleles", path); }
    if (im->pool.empty()) im->pool.push_back(0);
    for (auto& s : im->ind_names) im->ind_ptr.push_back(s.c_str());
    for (auto& s : im->pop_names) im->pop_ptr.push_back(s.c_str());
    for (auto& s : im->loc_names) im->loc_ptr.push_back(s.c_str());
    *out = im;
    return VPS_OK;
}

namespace {

// the image's packed ids, tables and names from per-locus genotype tables:
// locus l keeps its alleles' indices (k alleles) and ids[l][i] per individual
int pack_loci(vps_image* im, const char* path, std::vector<std::vector<Geno>>& gtab,
              std::vector<std::vector<uint32_t>>& ids, const std::vector<int>& kk, const std::vector<std::string>& names,
              const std::vector<bool>& keep)
{
    const int64_t N = im->N;
    im->gtoff.assign(1, 0);
    uint64_t bytes = 0;
    for (size_t l = 0; l < gtab.size(); ++l) {
        if (!keep[l]) continue;
        const int k = kk[l];
        if (k > VPS_MAX_ALLELES)
            return fail(VPS_ERR_UNSUPPORTED, "%s locus %s: %d alleles, the device path takes at most %d", path,
                        names[l].c_str(), k, VPS_MAX_ALLELES);
        const int64_t ng = (int64_t)gtab[l].size();
        const int b = ceil_log2(ng);
        if (b > 24) return fail(VPS_ERR_UNSUPPORTED, "%s locus %s: %lld genotypes", path, names[l].c_str(), (long long)ng);
        im->loc_names.push_back(names[l]);
        im->nallele.push_back(k);
        im->bitsize.push_back((uint8_t)b);
        im->byteoff.push_back(bytes);
        for (const Geno& g : gtab[l]) {
            im->gtploidy.push_back(g.missing ? 0 : (uint8_t)g.v);
            im->gtaoff.push_back((int64_t)im->pool.size());
            im->pool.insert(im->pool.end(), g.al.begin(), g.al.end());
        }
        im->gtoff.push_back((int64_t)im->gtploidy.size());
        const uint64_t nb = ((uint64_t)b * N + 7) / 8;
        im->bits.resize(bytes + nb, 0);
        for (int64_t i = 0; i < N; ++i) {
            const uint64_t v = ids[l][i];
            const uint64_t bit = (uint64_t)i * b;
            for (int x = 0; x < b; ++x)
                if (v >> x & 1) im->bits[bytes + ((bit + x) >> 3)] |= (uint8_t)(1u << ((bit + x) & 7));
        }
        bytes += nb;
    }
    im->L = (int64_t)im->nallele.size();
    if (im->L == 0) return fail(VPS_ERR_INVALID, "%s: no locus carries two alleles", path);
    if (im->pool.empty()) im->pool.push_back(0);
    return VPS_OK;
}

// one VCF genotype (the GT subfield) of ploidy v into g; false on a malformed
// allele or an index >= k
bool vcf_genotype(const std::string& gt, int k, Geno* g)
{
    int v = 1;
    for (size_t x = 1; x < gt.size(); ++x) v += (gt[x] == '/' || gt[x] == '|') ? 1 : 0;
    g->v = v;
    g->al.clear();
    g->missing = gt.empty() || gt[0] == '.';
    if (g->missing) return true;
    size_t p = 0;
    for (int a = 0; a < v; ++a) {
        if (p < gt.size() && gt[p] == '.') {
            g->missing = true;
            g->al.clear();
            return true;
        }
        int64_t x = 0;
        size_t q = p;
        while (q < gt.size() && gt[q] >= '0' && gt[q] <= '9') x = x * 10 + (gt[q++] - '0');
        if (q == p || x >= k || x > 0xFFFE) return false;
        g->al.push_back((uint16_t)x);
        if (q < gt.size() && gt[q] != '/' && gt[q] != '|') return false;
        p = q + 1;
    }
    std::sort(g->al.begin(), g->al.end());   // unphased
    return true;
}

}  // namespace

extern "C" int vps_image_read_vcf(const char* path, vps_image** out)
{
    if (!path || !out) return fail(VPS_ERR_INVALID, "null argument");
    *out = nullptr;
    // gzread reads plain text too, and every member of a BGZF file
    gzFile f = gzopen(path, "rb");
    if (!f) return fail(VPS_ERR_INVALID, "cannot open %s", path);
    std::string text;
    std::vector<char> buf(1 << 20);
    int n;
    while ((n = gzread(f, buf.data(), (unsigned)buf.size())) > 0) text.append(buf.data(), (size_t)n);
    const bool err = n < 0;
    gzclose(f);
    if (err) return fail(VPS_ERR_INVALID, "%s: read or decompression error", path);

    std::vector<std::string> lines = split(text, '\n');
    std::string().swap(text);
    vps_image* im = new vps_image();
    auto bad = [&](int code, const char* fmt, auto... a) {
        delete im;
        return fail(code, fmt, a...);
    };
    size_t li = 0;
    std::vector<std::string> samples;
    for (; li < lines.size(); ++li) {
        std::string& line = lines[li];
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.compare(0, 2, "##") == 0 || line.empty()) continue;
        if (line.compare(0, 6, "#CHROM") != 0) return bad(VPS_ERR_INVALID, "%s line %zu: expected the #CHROM header line", path, li + 1);
        const std::vector<std::string> head = split(line, '\t');
        if (head.size() < 10) return bad(VPS_ERR_INVALID, "%s: the #CHROM line names no samples", path);
        samples.assign(head.begin() + 9, head.end());
        ++li;
        break;
    }
    if (samples.empty()) return bad(VPS_ERR_INVALID, "%s: no #CHROM header line", path);
    const int64_t N = (int64_t)samples.size();
    if (N > 0x7FFFFFFF) return bad(VPS_ERR_UNSUPPORTED, "%s: too many samples", path);
    im->N = (int32_t)N;
    im->ind_names = samples;
    im->pop.assign((size_t)N, 0);
    im->pop_names.assign(1, "Total");
    im->S = 1;

    std::vector<std::vector<Geno>> gtab;
    std::vector<std::vector<uint32_t>> ids;
    std::vector<int> kk;
    std::vector<std::string> names;
    std::vector<bool> keep;
    for (; li < lines.size(); ++li) {
        std::string& line = lines[li];
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        const std::vector<std::string> fld = split(line, '\t');
        const size_t lineno = li + 1;
        if ((int64_t)fld.size() != 9 + N) return bad(VPS_ERR_INVALID, "%s line %zu: %zu fields, expected %lld", path, lineno, fld.size(), (long long)(9 + N));
        const std::string &pos = fld[1], &id = fld[2], &ref = fld[3], &alt = fld[4];
        const int k = (int)std::count(alt.begin(), alt.end(), ',') + 2;
        const std::vector<std::string> fmt = split(fld[8], ':');
        int gti = -1;
        for (size_t x = 0; x < fmt.size(); ++x)
            if (fmt[x] == "GT") { gti = (int)x; break; }
        if (gti < 0) return bad(VPS_ERR_INVALID, "%s line %zu: FORMAT has no GT", path, lineno);
        names.push_back(id.empty() || id == "." ? pos + "_" + ref + "_" + alt : id);
        kk.push_back(k);
        gtab.emplace_back();
        ids.emplace_back();
        std::unordered_map<std::string, uint32_t> gid;
        std::vector<uint32_t>& row = ids.back();
        row.reserve((size_t)N);
        std::vector<bool> seen(k, false);
        int distinct = 0;
        bool multi = false;   // a typed genotype with two or more distinct alleles
        for (int64_t i = 0; i < N; ++i) {
            const std::string& sf = fld[9 + i];
            // the GT subfield (trailing subfields may be left out)
            size_t a = 0;
            for (int x = 0; x < gti && a != std::string::npos; ++x) {
                a = sf.find(':', a);
                if (a != std::string::npos) ++a;
            }
            std::string gt;
            if (a != std::string::npos) gt = sf.substr(a, sf.find(':', a) == std::string::npos ? std::string::npos : sf.find(':', a) - a);
            Geno g;
            if (gt.empty()) return bad(VPS_ERR_INVALID, "%s line %zu sample %s: no GT value", path, lineno, samples[i].c_str());
            if (!vcf_genotype(gt, k, &g))
                return bad(VPS_ERR_INVALID, "%s line %zu sample %s: genotype '%s' (alleles 0..%d)", path, lineno,
                           samples[i].c_str(), gt.c_str(), k - 1);
            if (g.v > VPS_MAX_PLOIDY)
                return bad(VPS_ERR_INVALID, "%s line %zu sample %s: ploidy %d outside 1..%d", path, lineno, samples[i].c_str(), g.v,
                           VPS_MAX_PLOIDY);
            if (!g.missing) {
                int nd = 0;
                for (size_t x = 0; x < g.al.size(); ++x) {
                    if (x == 0 || g.al[x] != g.al[x - 1]) ++nd;
                    if (!seen[g.al[x]]) { seen[g.al[x]] = true; ++distinct; }
                }
                multi = multi || nd >= 2;
            }
            Geno miss;
            miss.v = g.v;
            miss.missing = true;
            uint32_t v = 0;
            for (const Geno* e : {&miss, &g}) {
                const std::string key = geno_key(*e);
                auto it = gid.find(key);
                if (it == gid.end()) {
                    it = gid.emplace(key, (uint32_t)gtab.back().size()).first;
                    gtab.back().push_back(*e);
                }
                v = it->second;
                if (g.missing) break;
            }
            row.push_back(v);
        }
        keep.push_back(multi || distinct >= 2);
        std::string().swap(line);
    }
    im->L_read = (int64_t)gtab.size();
    const int rc = pack_loci(im, path, gtab, ids, kk, names, keep);
    if (rc) { delete im; return rc; }
    for (auto& s2 : im->ind_names) im->ind_ptr.push_back(s2.c_str());
    for (auto& s2 : im->pop_names) im->pop_ptr.push_back(s2.c_str());
    for (auto& s2 : im->loc_names) im->loc_ptr.push_back(s2.c_str());
    *out = im;
    return VPS_OK;
}

// ---------------------------------------------------------------------------
// Populations and regions from -g_indtext (vcfpop's Initialize,
// src/function.cpp:4352-4500; AssignInds, src/load.cpp:88-160; AssignVInds
// and SetVReg, src/filter.cpp:166-398).  Each level's entry 0 is its default
// (DefPop, DefRegL1, DefRegL2, ...).  The region walk keeps each unit's
// parent as parsed and numbers the kept units separately; vcfpop renumbers
// the parent field in place, which gives the same order wherever no renumbered
// value meets a later unit's parsed index (always with one region level).

namespace {

struct Unit {
    std::string name;
    std::vector<std::string> members;
    int parent = 0;      // index of the enclosing region one level up
    int64_t nind = 0;
    int npop = 0;
};

// ParseTwoNumber (src/string2.cpp:189-194)
void two_numbers(const std::string& s, int& a, int& b)
{
    a = b = -1;
    if (!s.empty() && s[0] == '#') sscanf(s.c_str(), "#%d-#%d", &a, &b);
}

bool starts_lower(const std::string& s, const char* prefix)
{
    size_t i = 0;
    for (; prefix[i]; ++i) {
        if (i >= s.size()) return false;
        char c = s[i];
        if (c >= 'A' && c <= 'Z') c = (char)(c + 32);
        if (c != prefix[i]) return false;
    }
    return true;
}

struct Assignment {
    std::vector<int32_t> order;            // new position -> old individual
    std::vector<int32_t> pop;              // new position -> population
    std::vector<std::string> pop_names, reg_names;
};

int assign(const std::vector<std::string>& ind_names, const char* indtext, Assignment* out)
{
    const int N = (int)ind_names.size();
    std::vector<Unit> pops(1);
    pops[0].name = "DefPop";
    std::vector<std::vector<Unit>> regs(1, std::vector<Unit>(1));
    regs[0][0].name = "DefRegL1";
    int rl = -1;
    std::vector<std::string> toks;
    if (indtext) {
        std::string t;
        for (const char* c = indtext;; ++c) {
            if (!*c || *c == ' ' || *c == '\t' || *c == '\r' || *c == '\n') {
                if (!t.empty()) toks.push_back(t);
                t.clear();
                if (!*c) break;
            } else {
                t.push_back(*c);
            }
        }
    }
    for (const std::string& tok : toks) {
        if (starts_lower(tok, "#reg")) {
            ++rl;
            regs.emplace_back(1);
            regs.back()[0].name = "DefRegL" + std::to_string(rl + 2);
            continue;
        }
        const size_t colon = tok.find(':');
        if (colon == std::string::npos) return fail(VPS_ERR_INVALID, "-g_indtext: '%s' has no ':'", tok.c_str());
        Unit u;
        u.name = tok.substr(0, colon);
        for (const std::string& m : split(tok.substr(colon + 1), ','))
            if (!m.empty()) u.members.push_back(m);
        if (rl == -1) {
            for (const Unit& p : pops)
                if (p.name == u.name) return fail(VPS_ERR_INVALID, "-g_indtext: two populations have the same name: %s", u.name.c_str());
            pops.push_back(u);
            continue;
        }
        for (int r = 0; r <= rl; ++r)
            for (const Unit& x : regs[r])
                if (x.name == u.name) return fail(VPS_ERR_INVALID, "-g_indtext: repeat population/region name: %s", u.name.c_str());
        for (const Unit& p : pops)
            if (p.name == u.name) return fail(VPS_ERR_INVALID, "-g_indtext: repeat population/region name: %s", u.name.c_str());
        std::vector<Unit>& below = rl == 0 ? pops : regs[rl - 1];
        const int id = (int)regs[rl].size();
        auto take = [&](int i) {
            if (below[i].parent == id)
                return fail(VPS_ERR_INVALID, "-g_indtext: %s appears twice in region %s", below[i].name.c_str(), u.name.c_str());
            if (below[i].parent != 0)
                return fail(VPS_ERR_INVALID, "-g_indtext: regions %s and %s have the same member %s",
                            regs[rl][below[i].parent].name.c_str(), u.name.c_str(), below[i].name.c_str());
            below[i].parent = id;
            return (int)VPS_OK;
        };
        for (const std::string& m : u.members) {
            int v1, v2;
            two_numbers(m, v1, v2);
            if (v1 != -1 && v2 != -1) {
                if (v1 <= 0 || v1 >= (int)below.size() || v2 <= 0 || v2 >= (int)below.size())
                    return fail(VPS_ERR_INVALID, "-g_indtext: %s of region %s exceeds range", m.c_str(), u.name.c_str());
                for (int i = std::min(v1, v2); i <= std::max(v1, v2); ++i)
                    if (int e = take(i)) return e;
            } else {
                int found = -1;
                for (int i = 0; i < (int)below.size() && found < 0; ++i)
                    if (v1 == i || below[i].name == m) found = i;
                if (found < 0) return fail(VPS_ERR_INVALID, "-g_indtext: cannot find %s of region %s", m.c_str(), u.name.c_str());
                if (int e = take(found)) return e;
            }
        }
        regs[rl].push_back(u);
    }

    // individuals (AssignInds)
    std::unordered_map<std::string, int> name2id;
    for (int i = 0; i < N; ++i)
        if (!name2id.emplace(ind_names[i], i).second)
            return fail(VPS_ERR_INVALID, "-g_indtext: individuals #%d and #%d have the same name %s", name2id[ind_names[i]] + 1,
                        i + 1, ind_names[i].c_str());
    std::vector<int> popid(N, 0);
    for (int pp = (int)pops.size() - 1; pp >= 0; --pp) {
        auto put = [&](int i) {
            if (popid[i] != 0)
                return fail(VPS_ERR_INVALID, "-g_indtext: populations %s and %s have the same individual %s",
                            pops[popid[i]].name.c_str(), pops[pp].name.c_str(), ind_names[i].c_str());
            popid[i] = pp;
            return (int)VPS_OK;
        };
        for (const std::string& m : pops[pp].members) {
            int v1, v2;
            two_numbers(m, v1, v2);
            if (v1 == -1 && v2 == -1) {
                auto it = name2id.find(m);
                if (it == name2id.end())
                    return fail(VPS_ERR_INVALID, "-g_indtext: individual %s of population %s is not in the genotype data",
                                m.c_str(), pops[pp].name.c_str());
                if (int e = put(it->second)) return e;
            } else if (v2 == -1) {
                auto it = name2id.find(m);
                const int i = it != name2id.end() ? it->second : v1 - 1;
                if (i < 0 || i >= N)
                    return fail(VPS_ERR_INVALID, "-g_indtext: cannot find individual %s of population %s", m.c_str(),
                                pops[pp].name.c_str());
                if (int e = put(i)) return e;
            } else {
                if (v1 <= 0 || v1 > N || v2 <= 0 || v2 > N)
                    return fail(VPS_ERR_INVALID, "-g_indtext: %s of population %s exceeds range", m.c_str(), pops[pp].name.c_str());
                for (int i = std::min(v1, v2) - 1; i <= std::max(v1, v2) - 1; ++i)
                    if (int e = put(i)) return e;
            }
        }
    }

    // counts up the region tree; top levels holding a single region dropped
    for (int i = 0; i < N; ++i) pops[popid[i]].nind++;
    const int levels = (int)regs.size();
    std::vector<int> nreg(levels, 0);
    for (const Unit& p : pops)
        if (p.nind > 0) { regs[0][p.parent].npop++; regs[0][p.parent].nind += p.nind; }
    for (int r = 0; r < levels; ++r)
        for (const Unit& x : regs[r])
            if (x.npop > 0) {
                nreg[r]++;
                if (r + 1 < levels) { regs[r + 1][x.parent].npop++; regs[r + 1][x.parent].nind += x.nind; }
            }
    int lreg = levels;
    while (lreg && nreg[lreg - 1] == 1) {
        --lreg;
        if (lreg == 0) break;
    }
    int npop = 0;
    for (const Unit& p : pops) npop += p.nind > 0;

    // population order: the walk of SetVReg from the single top region
    std::vector<int> apops;
    if (lreg == 0 && npop == 1) {
        lreg = -1;
        for (int j = 0; j < (int)pops.size(); ++j)
            if (pops[j].nind > 0) apops.push_back(j);
    } else {
        int top = -1;
        for (int i = 0; i < (int)regs[lreg].size(); ++i)
            if (regs[lreg][i].nind > 0) top = i;
        std::vector<int> kept_l0;     // level-0 region of each population in apops
        std::function<void(int, int)> walk = [&](int r, int i) {
            const Unit& x = regs[r][i];
            if (x.npop == 0 || x.nind == 0) return;
            if (r > 0) {
                for (int j = 0; j < (int)regs[r - 1].size(); ++j)
                    if (regs[r - 1][j].parent == i) walk(r - 1, j);
            } else {
                for (int j = 0; j < (int)pops.size(); ++j)
                    if (pops[j].parent == i && pops[j].nind > 0) { apops.push_back(j); kept_l0.push_back(i); }
            }
        };
        walk(lreg, top);
        for (size_t a = 0; a < apops.size(); ++a)
            out->reg_names.push_back(lreg >= 1 ? regs[0][kept_l0[a]].name : std::string("Total"));
    }
    if ((int)apops.size() != npop) return fail(VPS_ERR_INVALID, "-g_indtext: a population outside the region tree");
    std::vector<int> newpop(pops.size(), -1);
    for (int a = 0; a < npop; ++a) {
        newpop[apops[a]] = a;
        out->pop_names.push_back(lreg == -1 ? std::string("Total") : pops[apops[a]].name);
    }
    if (lreg == -1) out->reg_names.assign(1, "Total");
    for (int a = 0; a < npop; ++a)
        for (int i = 0; i < N; ++i)
            if (popid[i] == apops[a]) { out->order.push_back(i); out->pop.push_back(a); }
    return VPS_OK;
}

}  // namespace

extern "C" int vps_image_assign_indtext(vps_image* im, const char* indtext)
{
    if (!im) return fail(VPS_ERR_INVALID, "null image");
    Assignment as;
    if (int e = assign(im->ind_names, indtext, &as)) return e;
    const int64_t N = im->N;
    // the packed ids of each locus in the new individual order
    std::vector<uint8_t> bits(im->bits.size(), 0);
    for (int64_t l = 0; l < im->L; ++l) {
        const uint64_t base = im->byteoff[l] * 8;
        const int b = im->bitsize[l];
        for (int64_t i = 0; i < N; ++i) {
            const uint64_t src = base + (uint64_t)as.order[i] * b, dst = base + (uint64_t)i * b;
            for (int x = 0; x < b; ++x)
                if (im->bits[(src + x) >> 3] >> ((src + x) & 7) & 1) bits[(dst + x) >> 3] |= (uint8_t)(1u << ((dst + x) & 7));
        }
    }
    std::vector<std::string> names(N);
    for (int64_t i = 0; i < N; ++i) names[i] = im->ind_names[as.order[i]];
    im->bits.swap(bits);
    im->ind_names.swap(names);
    im->pop = as.pop;
    im->pop_names = as.pop_names;
    im->reg_names = as.reg_names;
    im->S = (int32_t)im->pop_names.size();
    im->ind_ptr.clear();
    im->pop_ptr.clear();
    im->reg_ptr.clear();
    for (auto& x : im->ind_names) im->ind_ptr.push_back(x.c_str());
    for (auto& x : im->pop_names) im->pop_ptr.push_back(x.c_str());
    for (auto& x : im->reg_names) im->reg_ptr.push_back(x.c_str());
    return VPS_OK;
}

extern "C" int vps_image_desc(const vps_image* im, vps_dataset_desc* d)
{
    if (!im || !d) return fail(VPS_ERR_INVALID, "null argument");
    memset(d, 0, sizeof *d);
    d->n_ind = im->N;
    d->n_loc = im->L;
    d->n_pop = im->S;
    d->ind_pop = im->pop.data();
    d->geno_bits = im->bits.data();
    d->geno_byte_offset = im->byteoff.data();
    d->bitsize = im->bitsize.data();
    d->n_allele = im->nallele.data();
    d->gtab_offset = im->gtoff.data();
    d->gt_ploidy = im->gtploidy.data();
    d->gt_allele_offset = im->gtaoff.data();
    d->allele_pool = im->pool.data();
    d->ind_name = im->ind_ptr.data();
    d->pop_name = im->pop_ptr.data();
    d->loc_name = im->loc_ptr.data();
    d->pop_region = im->reg_ptr.empty() ? nullptr : im->reg_ptr.data();
    return VPS_OK;
}

extern "C" int vps_image_sizes(const vps_image* im, int64_t* loci_read, int64_t* loci_kept, int64_t* packed_bytes,
                               int64_t* n_genotypes, int64_t* n_alleles)
{
    if (!im) return fail(VPS_ERR_INVALID, "null image");
    if (loci_read) *loci_read = im->L_read;
    if (loci_kept) *loci_kept = im->L;
    if (packed_bytes) *packed_bytes = (int64_t)im->bits.size();
    if (n_genotypes) *n_genotypes = (int64_t)im->gtploidy.size();
    if (n_alleles) *n_alleles = (int64_t)im->pool.size();
    return VPS_OK;
}

extern "C" int vps_image_destroy(vps_image* im)
{
    delete im;
    return VPS_OK;
}

// vps_internal.h — device-side data layout and host handles of the engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <map>
#include <mutex>
#include <vector>
#include "../../include/vps.h"
#include "vps_numeric.h"

namespace vps {

// ---------------------------------------------------------------------------
// Genotype image in HBM.
//
// Loci are kept in the reference's sorted order (bitsize, then original
// index; src/structure.cpp:2202-2273).  Each run of equal bitsize is split
// into the loci the biallelic kernels decode (kind BI) and the rest (kind
// GEN); each part is cut, in sorted order, into tiles of up to 64 loci.  One
// wavefront owns one tile: lane j holds sorted locus tloc[lmap + j] (lanes
// past nloc repeat the tile's first locus and decode every id as untyped).
// The genotype ids of a tile are re-packed so that one 32-bit load per lane
// is one coalesced 256-byte wavefront load:
//     words[wbase + w*64 + j]  = ids of individuals [w*ipw, (w+1)*ipw) of
//                                lane j's locus, id i at bits (i % ipw) * bits
// with ipw = floor(32 / bits) ids per word (no id straddles two words).
//
// Genotype ids decode through one of these tables:
//   kind BI  (every locus of the tile has <= 2 alleles, <= 64 genotypes and
//            its allele copies in ascending order): one byte per genotype,
//            g -> v | c << 4 where v is the typed ploidy (0 = missing) and c
//            the dosage of allele 1; slot a holds allele (a >= v - c).
//            btab[dbase + g*64 + j] holds the byte of lane j for every
//            g < 2^bits (zero past the locus' table and on idle lanes); for
//            bits <= 3 the 8 bytes of locus l are also packed in bitab[l],
//            one 64-bit register per lane.
//   kind GEN (any other tile): gdesc[gbase[l] + g] = v | allele(a) << (4+6a)
//            for a < v (alleles < 64, ploidy <= 10).
// ---------------------------------------------------------------------------
constexpr int kTileLoci = 64;
constexpr int kKindBi = 0;
constexpr int kKindGen = 1;
constexpr int kBiRegBits = 3;    // BI tiles up to this bitsize decode from bitab
constexpr int kBiMaxBits = 6;    // BI tiles up to this bitsize (64 genotypes)

struct TileDesc {
    int64_t lmap;     // first entry of the tile in `tloc` (tile index * 64)
    int64_t wbase;    // first word of the tile in `words`
    int64_t dbase;    // first byte of the tile in `btab` (BI tiles)
    int32_t nloc;     // loci in the tile (<= 64)
    int32_t bits;     // genotype-id bitsize of every locus in the tile
    int32_t ipw;      // ids per 32-bit word
    int32_t kind;     // kKindBi or kKindGen
    int32_t maxk2;    // largest allele count in the tile
    int32_t pad;
};

// Jump-ahead table of xorshift128+: s' = M s over GF(2), split into 32
// nibbles of the 128-bit state; entry [p][v] is the 128-bit image of nibble
// p holding v (4 dwords: x_lo, x_hi, y_lo, y_hi).
constexpr int kJumpWords = 32 * 16 * 4;
// jump tables of the parallel stream parse: [0, 64) advance 8*l draws (l =
// lane), [64, 80) advance 512*w draws (w = wave)
constexpr int kRingLaneTabs = 64, kRingWaveTabs = 16;
// then tables by 4096 * 2^j draws, j < kRingPow2Tabs (any multiple of 4096
// in at most kRingPow2Tabs jumps; table j = 1 is one 8,192-word ring cycle)
constexpr int kRingPow2Base = kRingLaneTabs + kRingWaveTabs, kRingPow2Tabs = 24;
constexpr int kRingJumps = kRingPow2Base + kRingPow2Tabs;

struct DeviceDataset {
    // sizes
    int32_t N;
    int32_t S;
    int64_t L;
    int64_t KT;
    int32_t maxploidy;
    int32_t minploidy;
    int32_t maxK;
    int64_t ntiles;
    int64_t nbi, ngen;         // tiles of each kind
    // device arrays
    TileDesc* tiles;           // [ntiles]
    int32_t*  bi_list;         // tile indices of kind BI
    int32_t*  gen_list;        // tile indices of kind GEN
    int32_t*  tloc;            // [ntiles * 64] sorted locus of each tile lane
    uint8_t*  btab;            // BI genotype bytes, tile-major (see above)
    int32_t   bi_tab_rows;     // largest 2^bits over BI tiles above kBiRegBits (0: none)
    uint32_t* words;           // transposed genotype ids
    int32_t*  afo;             // [L] allele_freq_offset
    int32_t*  k2;              // [L]
    uint64_t* bitab;           // [L]
    int64_t*  gbase;           // [L]
    uint64_t* gdesc;           // [sum ngeno of GEN loci]
    int32_t*  vt;              // [N]
    int32_t*  popid;           // [N]
    int32_t*  pop_ind0;        // [S]
    int32_t*  pop_nind;        // [S]
    double*   totfreq;         // [KT] (F model)
    int32_t*  nhaplo;          // [L]
    int32_t*  ncopies;         // [KT] typed copies of each allele (sum over clusters of Ni)
    int64_t*  kdis;            // [maxK+1] loci per allele count
    int32_t*  typed;           // [ntyped] individuals with vt > 0, ascending
    int32_t   ntyped;
    uint32_t* ring_jump;       // [kRingJumps * kJumpWords] (see kRingLaneTabs)
};

}  // namespace vps

struct vps_dataset {
    int device = 0;
    vps::DeviceDataset d{};
    // host copies kept for state conversion, output tables and tests
    std::vector<int64_t> orig;         // sorted -> original locus index
    std::vector<int64_t> afo;
    std::vector<int32_t> k2;
    std::vector<int64_t> vt;
    std::vector<int32_t> popid;
    std::vector<int32_t> pop_ind0, pop_nind;
    std::vector<vps::TileDesc> tiles;
    std::vector<int32_t> tloc;         // as DeviceDataset::tloc
    std::vector<int64_t> lpos;         // sorted locus -> its entry in tloc
    std::vector<uint8_t> bitsize;
    std::vector<std::string> ind_names, pop_names, loc_names;
    std::vector<std::string> pop_regions;   // RegL1 name of each population (empty = "Total")
    std::vector<int64_t> kdis;
    int64_t device_bytes = 0;
    std::vector<void*> allocations;
    hipStream_t stream = nullptr;      // allocation and upload stream (non-blocking)
    // host-only build (device < 0): the packed layout stays on the host
    bool host_only = false;
    std::vector<uint32_t> h_words;
    std::vector<uint64_t> h_bitab, h_gdesc;
    std::vector<uint8_t> h_btab;
    std::vector<int64_t> h_gbase;
    // device jump tables per chunk size Nc: [ceil(N/Nc)][kJumpWords], entry c
    // advances a locus stream by c*Nc*maxploidy draws
    std::mutex jump_mu;
    std::map<int, uint32_t*> jump_cache;
};

namespace vps {

// error reporting (thread local), vps_runtime.cpp
int fail(int code, const char* fmt, ...);
#define VPS_HIP(call)                                                           \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return ::vps::fail(VPS_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,  \
                               hipGetErrorString(e_), __FILE__, __LINE__);      \
    } while (0)

// GF(2) jump table for D steps of xorshift128+ (vps_jump.cpp)
void jump_table(uint64_t D, uint32_t* out /* kJumpWords */);
// tables for 0, D, 2D, ..., (count-1)D steps
void jump_tables(uint64_t D, int count, uint32_t* out /* count * kJumpWords */);
// the same for xorshift96 (RNG<float>), state words x, y, z in entries 0..2
void jump_tables96(uint64_t D, int count, uint32_t* out /* count * kJumpWords */);
// largest individual chunk of the genotype passes (a block keeps
// kMaxChunk x VPS_MAX_K cluster counts in 32 KB of LDS)
constexpr int kMaxChunk = 512;
// device jump tables of a chunk size, built once per dataset (vps_dataset.cpp)
// the chunk jump tables of the Z pass streams for chunks of Nc individuals
// (x96: the xorshift96 streams of REAL = float with -g_fastsingle), cached
int dataset_jump(vps_dataset* ds, int Nc, bool x96, const uint32_t** out);
// apply a jump table to a state on the host (tests)
void jump_apply_host(const uint32_t* tab, uint64_t& x, uint64_t& y);

// device allocation helper that records the block for release
int dev_alloc(void** p, size_t bytes, std::vector<void*>& list, int64_t* counter, hipStream_t st);
void dev_free_all(std::vector<void*>& list, hipStream_t st);

}  // namespace vps

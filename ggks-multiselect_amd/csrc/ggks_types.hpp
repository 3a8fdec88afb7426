// ggks_types.hpp -- device-side tables shared by the selection kernels.
//
// The engine works on a table of top-level segments (one for a single-vector
// select, thousands for a batch).  A segment is split into chunks; every chunk
// is one work-group in the two streaming passes (count, compact), so the
// per-chunk histogram rows written by the count pass give every work-group its
// private write offsets in the compaction pass.  Each bucket that holds a
// requested rank becomes a child; children are finished in LDS.
#pragma once
#include <stdint.h>

namespace ggks {

constexpr int kThreads = 1024;             // work-group size of every kernel
constexpr uint32_t kMaxBuckets = 32768;    // histogram bins per segment (LDS)
constexpr uint32_t kMaxPrefixBits = 12;    // LUT prefixes per segment <= 4096
constexpr uint32_t kCursorCap = 4096;      // LDS write cursors in compaction
constexpr uint32_t kFinishLdsBytes = 65536;// keys kept in LDS by the finisher
constexpr uint32_t kQCap = 1024;           // queries per finisher batch
constexpr uint32_t kGroupMax = 32;         // prefix groups per histogram pass
constexpr uint32_t kReduceBins = 64;       // bins per reduce work item
constexpr uint32_t kSortRanks = 32;        // finisher sorts a child above this
constexpr uint32_t kMaxChunks = 512;       // chunks per segment (reduce registers)

// Where a segment's queries come from.
enum QuerySource : uint32_t {
  QSRC_QUERIES = 0,  // explicit Query records at q_off
  QSRC_K = 1,        // raw k values (k-th largest) at q_off; output slot out_off + i
  QSRC_TOPK = 2      // the q_cnt largest: rank len - i into output slot out_off + i
};

// Bucket-function modes of a segment.
enum SegMode : uint32_t {
  MODE_SAMPLE = 0,  // prefix LUT sized from a random sample of the segment
  MODE_RADIX = 1,   // the top radix_bits bits of the key (MSD radix digit)
  MODE_RANGE = 2    // equal-width buckets over the known key range [lo, hi]
};

// A requested order statistic: 1-based ascending rank inside its segment or
// child, and the index of the output slot it fills.
struct Query {
  uint32_t rank;
  uint32_t out;
};

struct SegDesc {          // planned on the host, read-only on the device
  uint64_t data_off;      // element offset of the segment in the source
  uint64_t len;           // elements
  uint64_t row_base;      // u32 offset of the first chunk's histogram row
  uint64_t red_base;      // element offset of its children in the reduced buffer
  uint64_t lo, hi;        // MODE_RANGE: key range of every element
  uint32_t bmax;          // bucket budget including the two tail buckets
  uint32_t pbits;         // LUT prefix bits (P <= 2^pbits)
  uint32_t nchunks;
  uint32_t sample_off;    // first sample slot (key units)
  uint32_t nsample;
  uint32_t lut_off;       // first LUT entry (u32)
  uint32_t tot_off;       // first total (u32), bmax entries
  uint32_t bits_off;      // first bitmap word (u32), bmax/32 words
  uint32_t q_off, q_cnt;  // its queries; child slots use the same range
  uint32_t mode;          // SegMode
  uint32_t radix_bits;    // MODE_RADIX digit width
  uint32_t qsrc;          // QuerySource
  uint32_t out_off;       // first output slot (QSRC_K, QSRC_TOPK)
  uint32_t flags;         // SEG_LEVEL2: its heavy children go to the second level
};

enum SegFlags : uint32_t { SEG_LEVEL2 = 1u };

// The bucket function of a segment (BucketFn in ggks_kernels.hpp):
//   low = min(max(key, kb) - kb, lmax)      keys outside the sampled range
//                                           fall into the edge buckets
//   p   = (low >> poff) & (2^pwid - 1)      prefix, looked up in the LUT
//   t   = ((low >> rsh) & (2^twid - 1)) << lsh   the low s bits of low,
//                                           normalised to 24 bits
//   b   = base[p] + (t * nb[p]) >> 24       (one 24-bit multiply-high)
//
// A heavy value hv (a key covering a large share of the sample) gets a
// bucket of its own: with L the function above, the bucket is
//   F(key) = L(key) + [key >= hv] + [key > hv]
// which is monotone (keys below hv keep L <= L(hv), hv maps to L(hv) + 1,
// keys above to L + 2 >= L(hv) + 2), and bucket L(hv) + 1 holds hv alone.
struct SegState {         // written on the device by the LUT kernel
  uint64_t kb;            // base key
  uint64_t kmax;          // largest sampled key (information)
  uint64_t lmax;          // clamp of key - kb
  uint64_t hv;            // heavy value (when hvon)
  uint32_t s;             // prefix shift
  uint32_t P;             // number of prefixes
  uint32_t B;             // number of buckets (F's when hvon)
  uint32_t nact;          // active buckets (children)
  uint32_t poff, pwid;    // prefix bit field
  uint32_t rsh, twid, lsh;// sub-bucket bit field and normalising shift
  uint32_t hvon;          // 1: the heavy-value bucket is in use
  uint32_t hb;            // F's bucket of hv (L(hv) + 1) when hvon
  uint32_t nfree;         // next free child slot (second-level children)
  uint32_t arrive;        // LUT parts done (k_lut; zeroed by k_sample)
  uint32_t nred;          // keys of the active buckets (k_findk)
};

// A segment of at most kDenseChildren children whose active buckets hold at
// least len / 2^kDenseLog2 keys (dense) is compacted by sorting each round's
// keys by child in LDS and storing runs (the sort's tables hold
// kDenseChildren).
constexpr uint32_t kDenseLog2 = 2;
constexpr uint32_t kDenseChildren = 1024;

// Work-groups that build one segment's LUT: a sampled segment whose sample
// is more than kLutPartBytes splits its sample histogram over parts of
// kLutPartBytes (each histograms its share and adds it into the segment's
// LUT region, which k_sample zeroes; the last to arrive builds the LUT from
// the sums), so no single CU reads the whole sample.  At 64 Ki samples the
// split pays for 8-byte keys (25 -> 20 us) and not for 4-byte keys (12.5 ->
// 17 us with 16 parts: the parts' device-scope atomics on the same few
// thousand counters cost more than the reads they spread).
constexpr uint32_t kLutPartBytes = 128u << 10;
constexpr uint32_t kLutSplitBytes = 256u << 10;
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t lut_parts(uint32_t mode, uint32_t nsample, uint32_t key_bytes) {
  const uint64_t bytes = (uint64_t)nsample * key_bytes;
  return (mode == MODE_SAMPLE && bytes > kLutSplitBytes) ? (uint32_t)(bytes / kLutPartBytes) : 1u;
}

struct Chunk {            // one work-group of the streaming passes
  uint32_t seg;
  uint32_t idx;           // chunk index inside the segment (row index)
  uint64_t start;         // element offset inside the segment
  uint64_t len;
};

struct SampleBlock {      // one work-group of the sample kernel: kThreads samples
  uint32_t seg;
  uint32_t blk;           // which kThreads-sample block of the segment
};

struct ReduceItem {       // 2^gi_log2 groups of kReduceBins consecutive bins of one segment
  uint32_t seg;
  uint16_t group;         // first group
  uint16_t gi_log2;       // groups in the item, log2: the chunk range splits 16 >> gi_log2 ways
};

// log2 of the groups per reduce item for a segment of g chunks: the 16
// wavefronts of a work-group split into parts of the chunk range times
// groups, with as few parts as keep a lane at <= 16 chunks (a segment of
// <= 16 chunks covers 16 groups, 1024 buckets, per work-group; 256 chunks
// split 16 ways).
inline uint32_t reduce_groups_log2(uint64_t g) {
  uint32_t parts_log2 = 0;
  while (parts_log2 < 4 && (16ull << parts_log2) < g) ++parts_log2;
  return 4 - parts_log2;
}

struct Child {            // one bucket to finish, or one small whole segment
  uint64_t start;         // element offset in its source buffer
  uint64_t len;
  uint64_t lo, hi;        // key range of its elements (unless need_range)
  uint32_t q_off, q_cnt;  // its queries (ranks inside the child)
  uint32_t src;           // 0: the input buffer, 1: the reduced buffer
  uint32_t need_range;    // 1: lo/hi unknown, computed by the finisher
  uint32_t bucket;        // bucket of the parent segment it came from
  uint32_t seg;           // parent segment
};

// Second level.  A child of more than kHeavyKeys keys found by k_findk is
// re-bucketed by the whole GPU instead of being finished by one work-group:
// k_count2 histograms its keys in chunks of kChunk2 keys over kBuckets2
// buckets (the exact digit (key - lo) >> shift of its key range),
// k_resolve2 resolves its ranks (single-key buckets are answered) and lays
// out its second-level children, and k_compact2 copies their keys into the
// second reduced buffer at the child's own offsets.  Its children take the
// heavy child's slot and free slots of its segment's child table.
constexpr uint64_t kHeavyKeys = 1ull << 16;
constexpr uint64_t kChunk2 = 1ull << 18;
constexpr uint32_t kBuckets2 = 4096;
constexpr uint32_t kBits2 = 12;  // log2 kBuckets2

struct Heavy {             // written by k_findk
  uint64_t start, len;     // its keys in the first reduced buffer
  uint64_t lo, hi;         // key range
  uint32_t seg;            // first-level segment
  uint32_t slot;           // its child slot
  uint32_t q_off, q_cnt;   // its ranks (Query records of the child)
  uint32_t chunk_base;     // first second-level chunk
  uint32_t nchunks;
  uint32_t shift;          // bucket = (key - lo) >> shift
  uint32_t done;           // unused (zero)
  uint32_t nact;           // second-level children (k_compact2 skips 0)
  uint32_t slot_base;      // child slots of children 1.. (child 0 takes `slot`)
  uint32_t pad[2];
};

struct Chunk2 {
  uint32_t h;              // heavy child
  uint32_t idx;            // chunk index inside it
  uint64_t start, len;     // inside the heavy child
};

// Per-context status bits written by kernels.
enum StatusBits : uint32_t {
  STATUS_BAD_K = 1u,      // a device k list held k outside [1, n]
  STATUS_OVERFLOW = 2u    // internal capacity exceeded (should not happen)
};

}  // namespace ggks

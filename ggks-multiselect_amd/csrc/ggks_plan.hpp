// ggks_plan.hpp -- host-side planning of one select: the segment / chunk /
// reduce tables and the layout of the context arena they and the kernels'
// scratch live in.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ggks.h"
#include "ggks_keys.hpp"
#include "ggks_types.hpp"

namespace ggks {

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// Tuned defaults per length band (bucket budget and chunk count).  The bucket
// budget trades the reduced-set size (~ ranks * n / buckets) against the
// per-chunk histogram rows (chunks * buckets * 4 bytes).
struct Tuning {
  uint32_t buckets_log2;
  uint32_t chunks;
};

inline Tuning default_tuning(uint64_t n) {
  if (n >= (1ull << 26)) return {15, 256};
  if (n >= (1ull << 24)) return {13, 256};
  if (n >= (1ull << 22)) return {12, 128};
  if (n >= (1ull << 20)) return {11, 64};
  return {10, 32};
}

// Single-vector tuning per (dtype, floor(log2 n)): the defaults above until
// the autotuner sets a band (ggks_tuning_single_set).
constexpr int kSingleBands = 33;
struct SingleTuningTable {
  Tuning t[4][kSingleBands];
  SingleTuningTable() {
    for (int d = 0; d < 4; ++d)
      for (int b = 0; b < kSingleBands; ++b) t[d][b] = default_tuning(1ull << b);
  }
};
inline SingleTuningTable& single_tuning_table() {
  static SingleTuningTable tab;
  return tab;
}
inline Tuning single_tuning(int dtype, uint64_t n) {
  int b = 0;
  while (b < kSingleBands - 1 && (n >> (b + 1))) ++b;
  return single_tuning_table().t[dtype][b];
}

// Bucket budget (log2) of a radix segment with digits of rb bits: room for
// the 2^rb digits plus the edge buckets, and never below 2^8, so that every
// segment's bucket tables are whole 16-byte words of group sums (the bucket
// prefix kernel loads bmax / 256 of them) and every segment's table offsets
// in a batch stay multiples of 256 buckets.
inline uint32_t radix_budget_log2(uint32_t rb) { return std::max<uint32_t>(rb + 1, 8); }

// LUT prefix bits of a radix segment: up to 2^min(rb, 11) digit prefixes and
// two guard prefixes (keys below / above the sampled digit window).
inline uint32_t radix_lut_bits(uint32_t rb) { return std::min<uint32_t>(kMaxPrefixBits, rb + 1); }

// Batches of at most this many segments split large segments' LUT builds.
constexpr size_t kLutSplitSegs = 16;

struct Plan {
  std::vector<SegDesc> segs;
  std::vector<Chunk> chunks;
  std::vector<ReduceItem> items;
  std::vector<Child> direct_children;
  bool direct = false;  // single-vector plan answered by the finisher alone
  uint32_t nq = 0;
  uint32_t qmax = 0;  // most ranks of one engine segment
  uint32_t bmax_max = 0;
  uint32_t pbits_max = 0;
  // table sizes (equal to the vectors' sizes when the tables are built)
  uint32_t nsegs = 0, nchunks = 0, nitems = 0, nsblocks = 0, ndirect = 0;
  uint32_t lut_parts = 1;  // most LUT work-groups of a segment (lut_parts())
  // longest direct child, and the largest mean bucket size of an engine
  // segment (a radix segment's whole length): picks the finisher variant
  uint64_t dmax_keys = 0, emax_keys = 0;
  // second level: enabled, capacity (heavy children, chunks) and regions
  bool level2 = false;
  uint32_t hmax = 0, cmax = 0;
  size_t off_hcount = 0, off_heavy = 0, off_chunks2 = 0, off_rows2 = 0, off_bits2 = 0, off_wb2 = 0;
  size_t off_cq2 = 0, off_red2 = 0;
  uint64_t seed = 0;
  // arena offsets (bytes)
  size_t off_tables = 0, off_segs = 0, off_chunks = 0, off_items = 0, off_sblocks = 0, tables_end = 0;
  size_t off_states = 0, off_samples = 0, off_minmax = 0, off_lut = 0, off_rows = 0, off_tot = 0;
  size_t off_bits = 0, off_wbase = 0, off_q = 0, off_qb = 0, off_qr = 0;
  size_t off_gsum = 0;  // per-group bucket sums (k_reduce -> k_findk)
  size_t off_big = 0;   // children k_finish takes (k_findk counts, k_finish reads)
  size_t off_cinfo = 0;  // per child slot: slice start in its segment's part, bucket
  size_t off_children = 0, off_cq = 0, off_red = 0, total = 0;
  size_t off_kvals = 0, off_dchildren = 0, off_dqueries = 0;  // batch / direct tables
  std::vector<uint32_t> kvals;         // k lists of the engine segments, by q_off
  std::vector<Query> direct_queries;   // explicit queries of the direct children
  std::vector<SampleBlock> sblocks;    // the sample kernel's work-groups, densely

  size_t table_bytes() const { return tables_end - off_tables; }

  void write_tables(char* dst) const {
    if (!segs.empty()) memcpy(dst + (off_segs - off_tables), segs.data(), segs.size() * sizeof(SegDesc));
    if (!chunks.empty())
      memcpy(dst + (off_chunks - off_tables), chunks.data(), chunks.size() * sizeof(Chunk));
    if (!items.empty())
      memcpy(dst + (off_items - off_tables), items.data(), items.size() * sizeof(ReduceItem));
    if (!sblocks.empty())
      memcpy(dst + (off_sblocks - off_tables), sblocks.data(), sblocks.size() * sizeof(SampleBlock));
    if (!kvals.empty())
      memcpy(dst + (off_kvals - off_tables), kvals.data(), kvals.size() * 4);
    if (!direct_children.empty())
      memcpy(dst + (off_dchildren - off_tables), direct_children.data(),
             direct_children.size() * sizeof(Child));
    if (!direct_queries.empty())
      memcpy(dst + (off_dqueries - off_tables), direct_queries.data(),
             direct_queries.size() * sizeof(Query));
  }
};

// Whether a segment's heavy children go to the second level (ggks_opts
// level2: 0 = radix segments only, 1 = every segment, -1 = none), and the
// most heavy children it can have (each holds a rank and more than
// kHeavyKeys keys).
inline bool seg_level2(const ggks_opts& o, uint32_t mode) {
  return o.level2 > 0 || (o.level2 == 0 && mode == MODE_RADIX);
}
inline uint64_t seg_heavy_max(uint64_t len, uint32_t nq) {
  return std::min<uint64_t>(nq, len / (kHeavyKeys + 1));
}

// Second-level regions at the end of an arena layout (tables and scratch
// sized by the capacity: hmax heavy children, cmax chunks of kChunk2 keys,
// red elements in the second reduced buffer, nq query slots).
template <typename Take>
inline void layout_level2(Plan& p, Take&& take, uint64_t red, uint64_t nq, int eb) {
  if (!p.level2) return;
  p.off_hcount = take(256);
  p.off_heavy = take((size_t)p.hmax * sizeof(Heavy));
  p.off_chunks2 = take((size_t)p.cmax * sizeof(Chunk2));
  p.off_rows2 = take((size_t)p.cmax * kBuckets2 * 4);
  p.off_bits2 = take((size_t)p.hmax * (kBuckets2 / 32) * 4);
  p.off_wb2 = take((size_t)p.hmax * (kBuckets2 / 32) * 4);
  p.off_cq2 = take((size_t)nq * sizeof(Query));
  p.off_red2 = take((size_t)red * eb);
}

// The shape of a single-vector plan: its route, bucket function parameters,
// chunking and sample size, computed from scalars alone (no tables).
struct SingleShape {
  bool direct;        // the whole vector fits the finisher's LDS
  uint32_t mode, rb;  // SegMode and radix digit width
  uint32_t bmax, pb;  // bucket budget, LUT prefix bits
  uint32_t ns;        // samples
  uint32_t nitems;    // reduce items
  uint32_t gl;        // log2 groups per reduce item
  uint64_t G, csz;    // chunks and chunk length
};

inline SingleShape single_shape(int kind, int dtype, uint64_t n, const ggks_opts& o) {
  SingleShape sh = {};
  const int eb = dtype_bytes(dtype);
  const uint64_t cap = kFinishLdsBytes / (uint32_t)eb;
  if (n <= cap) {
    sh.direct = true;
    return sh;
  }
  Tuning t = single_tuning(dtype, n);
  uint32_t bl;
  sh.mode = MODE_SAMPLE;
  if (kind == GGKS_KIND_RADIX) {
    sh.mode = MODE_RADIX;
    uint32_t rb = o.radix_bits ? (uint32_t)o.radix_bits : 14u;
    sh.rb = std::min<uint32_t>(14, std::max<uint32_t>(4, rb));
    bl = radix_budget_log2(sh.rb);
    sh.pb = radix_lut_bits(sh.rb);
  } else {
    bl = o.buckets_log2 ? (uint32_t)o.buckets_log2 : t.buckets_log2;
    bl = std::min<uint32_t>(15, std::max<uint32_t>(8, bl));
    sh.pb = std::min<uint32_t>(kMaxPrefixBits, bl - 2);
  }
  sh.bmax = 1u << bl;
  uint64_t G = o.blocks > 0 ? (uint64_t)o.blocks : t.chunks;
  G = std::max<uint64_t>(1, std::min<uint64_t>(G, (n + 4095) / 4096));
  G = std::min<uint64_t>(G, kMaxChunks);
  uint64_t csz = (n + G - 1) / G;
  csz = (csz + 63) / 64 * 64;
  sh.G = (n + csz - 1) / csz;
  sh.csz = csz;
  {  // both modes sample: the LUT (bucket) or the digit window (radix)
    uint64_t want = (n / 256 + 1023) / 1024 * 1024;
    sh.ns = (uint32_t)std::min<uint64_t>(65536, std::max<uint64_t>(4096, want));
  }
  sh.gl = reduce_groups_log2(sh.G);
  const uint32_t ng = (sh.bmax + kReduceBins - 1) / kReduceBins;
  sh.nitems = (ng + (1u << sh.gl) - 1) >> sh.gl;
  return sh;
}

// The descriptor of a single-vector plan's one engine segment.
inline SegDesc single_segment(const SingleShape& sh, int kind, uint64_t n, uint32_t nq) {
  SegDesc s = {};
  s.data_off = 0;
  s.len = n;
  s.row_base = 0;
  s.red_base = 0;
  s.bmax = sh.bmax;
  s.pbits = sh.pb;
  s.nchunks = (uint32_t)sh.G;
  s.sample_off = 0;
  s.nsample = sh.ns;
  s.lut_off = 0;
  s.tot_off = 0;
  s.bits_off = 0;
  s.q_off = 0;
  s.q_cnt = nq;
  s.mode = sh.mode;
  s.radix_bits = sh.rb;
  s.qsrc = kind == GGKS_KIND_TOPK ? QSRC_TOPK : QSRC_K;
  s.out_off = 0;
  s.flags = 0;
  return s;
}

// The direct plan's one child: the whole vector, range found by the finisher.
inline Child single_direct_child(uint64_t n, uint32_t nq) {
  Child c = {};
  c.start = 0;
  c.len = n;
  c.q_off = 0;
  c.q_cnt = nq;
  c.src = 0;
  c.need_range = 1;
  return c;
}

// Plan a single-vector select of nq ranks over n keys.  The engine's single
// segment reads its queries straight from the k list (or as ranks of a
// top-K); the small-n direct path takes explicit Query records.  With
// tables = false only the arena layout is computed (the table vectors stay
// empty and allocate nothing): the caller materialises the tables itself.
inline void plan_single(Plan& p, int kind, int dtype, uint64_t n, uint32_t nq,
                        const ggks_opts& o, bool tables = true) {
  const int eb = dtype_bytes(dtype);
  const SingleShape sh = single_shape(kind, dtype, n, o);
  p.nq = nq;
  p.qmax = nq;
  p.seed = o.seed ? o.seed : 0x6A09E667F3BCC909ull;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off = align_up(off + bytes);
    return at;
  };
  if (sh.direct) {
    p.direct = true;
    p.ndirect = 1;
    p.dmax_keys = n;
    if (tables) p.direct_children.push_back(single_direct_child(n, nq));
    p.off_tables = off;
    p.off_dchildren = take(sizeof(Child));
    p.tables_end = off;
    p.off_dqueries = take((size_t)nq * sizeof(Query));  // written by a query kernel
    p.total = off;
    return;
  }
  const uint32_t bmax = sh.bmax, pb = sh.pb, ns = sh.ns;
  const uint64_t G = sh.G;
  if (tables) {
    p.segs.push_back(single_segment(sh, kind, n, nq));
    for (uint64_t i = 0; i < G; ++i) {
      Chunk c = {};
      c.seg = 0;
      c.idx = (uint32_t)i;
      c.start = i * sh.csz;
      c.len = std::min<uint64_t>(sh.csz, n - i * sh.csz);
      p.chunks.push_back(c);
    }
    for (uint32_t i = 0; i < sh.nitems; ++i)
      p.items.push_back({0u, (uint16_t)(i << sh.gl), (uint16_t)sh.gl});
    for (uint32_t y = 0; y < ns / kThreads; ++y) p.sblocks.push_back({0u, y});
  }
  if (seg_level2(o, sh.mode)) {
    const uint64_t hm = seg_heavy_max(n, nq);
    if (hm > 0) {
      p.level2 = true;
      p.hmax = (uint32_t)hm;
      p.cmax = (uint32_t)(n / kChunk2 + hm);
    }
  }
  if (tables && p.level2) p.segs[0].flags |= SEG_LEVEL2;
  p.bmax_max = bmax;
  p.pbits_max = pb;
  p.nsegs = 1;
  p.nchunks = (uint32_t)G;
  p.nitems = sh.nitems;
  p.nsblocks = ns / kThreads;
  p.lut_parts = ggks::lut_parts(sh.mode, ns, (uint32_t)eb);
  p.emax_keys = sh.mode == MODE_RADIX ? n : (n + bmax - 1) / bmax;

  p.off_tables = off;
  p.off_segs = take(sizeof(SegDesc));
  p.off_chunks = take((size_t)G * sizeof(Chunk));
  p.off_items = take((size_t)sh.nitems * sizeof(ReduceItem));
  p.off_sblocks = take((size_t)(ns / kThreads) * sizeof(SampleBlock));
  p.tables_end = off;
  p.off_states = take(sizeof(SegState));
  p.off_samples = take((size_t)std::max<uint32_t>(ns, 1) * eb);
  p.off_minmax = take((size_t)std::max<uint32_t>(ns / 1024, 1) * 4 * eb);
  p.off_lut = take((size_t)(1u << pb) * 4);
  p.off_rows = take((size_t)G * bmax * 4);
  p.off_tot = take((size_t)bmax * 4);
  p.off_gsum = take((size_t)(bmax / kReduceBins + 1) * 4);
  p.off_bits = take((size_t)((bmax + 31) / 32) * 4);
  p.off_wbase = take((size_t)((bmax + 31) / 32) * 4);
  p.off_q = take((size_t)nq * sizeof(Query));
  p.off_qb = take((size_t)nq * 4);
  p.off_qr = take((size_t)nq * 4);
  p.off_children = take((size_t)nq * sizeof(Child));
  p.off_cinfo = take((size_t)nq * 8);
  p.off_cq = take((size_t)nq * sizeof(Query));
  p.off_red = take((size_t)n * eb);
  p.off_big = take(4);
  layout_level2(p, take, n, nq, eb);
  p.total = off;
}

// ------------------------------------------------------------------ batches
// Per (dtype, floor(log2 length)) band: the bucket budget, the chunk size of
// the streaming passes and whether a segment of the band goes straight to
// the finisher (whole segment in LDS, no bucketing).  Defaults below; the
// autotuner overwrites entries through ggks_tuning_set.
struct BandTuning {
  uint8_t buckets_log2;
  uint8_t chunk_log2;  // elements per chunk
  uint8_t direct;
  uint8_t pad;
};
constexpr int kBands = 33;  // floor(log2 n) = 0 .. 32

inline BandTuning band_default(int dtype, int band) {
  const uint32_t cap = kFinishLdsBytes / (uint32_t)dtype_bytes(dtype);
  BandTuning t = {};
  if ((1ull << band) * 2 - 1 <= cap) {  // every length of the band fits in LDS
    t.direct = 1;
    return t;
  }
  const int bl = band <= 15 ? 8 : band <= 17 ? 10 : band <= 19 ? 11 : band <= 21 ? 12
               : band <= 23 ? 13 : 15;
  const int cl = band <= 15 ? 14 : band <= 17 ? 15 : band <= 19 ? 16 : band <= 21 ? 17
               : band <= 23 ? 18 : 20;
  t.buckets_log2 = (uint8_t)bl;
  t.chunk_log2 = (uint8_t)cl;
  return t;
}

struct TuningTable {
  BandTuning t[4][kBands];
  TuningTable() {
    for (int d = 0; d < 4; ++d)
      for (int b = 0; b < kBands; ++b) t[d][b] = band_default(d, b);
  }
};

inline TuningTable& tuning_table() {
  static TuningTable tab;
  return tab;
}

inline int band_of(uint64_t n) {
  int b = 0;
  while (b < 32 && (n >> (b + 1))) ++b;
  return b;
}

// One segment of a batch as the planner sees it (validated by the caller).
struct BatchSeg {
  uint64_t offset, length;
  const uint32_t* k;  // host k values (multiselect), null for top-K
  uint32_t num_k;     // k count, or K for a top-K
  uint32_t kind;      // GGKS_KIND_*
  uint64_t out_offset;
};

// Largest value of the 32-bit table offsets a batch plan may reach (segment
// descriptors hold sample, LUT, bucket-table, bitmap and query offsets as
// uint32; kernels index chunks, items and sample blocks by 32-bit grid
// indices).  Element offsets (segment data, reduced buffer, rows) are 64-bit
// and unbounded here.
constexpr uint64_t kPlanTableLimit = 0xFFFFFFFFull;
constexpr uint64_t kPlanGridLimit = 0x7FFFFFFFull;

// Plan a batch: segments whose band is marked direct (and whose length fits
// in the finisher's LDS) become direct children with explicit queries; the
// rest become engine segments with their own chunk rows, LUT, bitmaps and a
// slice of the reduced buffer.  Returns false (and leaves p unusable) when
// the batch is too large for the 32-bit table offsets: the caller refuses it
// or splits it.
inline bool plan_batch(Plan& p, int dtype, const std::vector<BatchSeg>& in, const ggks_opts& o) {
  const int eb = dtype_bytes(dtype);
  const uint64_t cap = kFinishLdsBytes / (uint32_t)eb;
  const TuningTable& tab = tuning_table();
  p.seed = o.seed ? o.seed : 0x6A09E667F3BCC909ull;
  uint64_t red = 0, rowc = 0;
  // 64-bit running totals of the 32-bit offsets, checked before each use
  uint64_t qoff = 0, sam = 0, lutc = 0, totc = 0, bitc = 0, dq = 0;
  uint64_t hmax = 0, cmax = 0;  // second-level capacity
  for (size_t i = 0; i < in.size(); ++i) {
    const BatchSeg& g = in[i];
    if (g.num_k == 0) continue;
    const uint64_t n = g.length;
    const BandTuning bt = tab.t[dtype][band_of(n)];
    const bool topk = g.kind == GGKS_KIND_TOPK;
    if (n <= cap && (bt.direct || n <= 64)) {
      if (dq + g.num_k > kPlanTableLimit || p.direct_children.size() >= kPlanGridLimit)
        return false;
      Child c = {};
      c.start = g.offset;
      c.len = n;
      c.q_off = (uint32_t)dq;
      c.q_cnt = g.num_k;
      c.src = 0;
      c.need_range = 1;
      c.seg = (uint32_t)i;
      p.direct_children.push_back(c);
      for (uint32_t j = 0; j < g.num_k; ++j) {
        Query q;
        q.rank = topk ? (uint32_t)(n - j) : (uint32_t)(n - g.k[j] + 1);
        q.out = (uint32_t)(g.out_offset + j);
        p.direct_queries.push_back(q);
      }
      dq += g.num_k;
      continue;
    }
    uint32_t mode = MODE_SAMPLE, rb = 0, bl, pb;
    if (g.kind == GGKS_KIND_RADIX) {
      mode = MODE_RADIX;
      rb = o.radix_bits ? (uint32_t)o.radix_bits : 14u;
      rb = std::min<uint32_t>(14, std::max<uint32_t>(4, rb));
      bl = radix_budget_log2(rb);
      pb = radix_lut_bits(rb);
    } else {
      bl = o.buckets_log2 ? (uint32_t)o.buckets_log2 : (bt.buckets_log2 ? bt.buckets_log2 : 12u);
      bl = std::min<uint32_t>(15, std::max<uint32_t>(8, bl));
      pb = std::min<uint32_t>(kMaxPrefixBits, bl - 2);
    }
    const uint32_t bmax = 1u << bl;
    uint64_t csz = 1ull << (bt.chunk_log2 ? bt.chunk_log2 : 16);
    uint64_t G = (n + csz - 1) / csz;
    G = std::max<uint64_t>(1, std::min<uint64_t>(G, kMaxChunks));
    csz = (n + G - 1) / G;
    csz = (csz + 63) / 64 * 64;
    G = (n + csz - 1) / csz;
    uint32_t ns = 0;
    // one sample per 1024 keys, 1024 to 65536: every sample is a random
    // gather that fetches a whole line, and a batch of many segments pays
    // for them in HBM bandwidth; sparser samples leave buckets a little less
    // even, which costs less than the gathers save
    {  // both modes sample: the LUT (bucket) or the digit window (radix)
      uint64_t want = (n / 1024 + 1023) / 1024 * 1024;
      ns = (uint32_t)std::min<uint64_t>(65536, std::max<uint64_t>(1024, want));
    }
    // every 32-bit offset this segment would take, and the grids, must fit
    const uint32_t nitems = (((bmax + kReduceBins - 1) / kReduceBins) >> reduce_groups_log2(G)) + 1;
    if (qoff + g.num_k > kPlanTableLimit || sam + ns > kPlanTableLimit ||
        lutc + (1u << pb) > kPlanTableLimit || totc + bmax > kPlanTableLimit ||
        bitc + bmax / 32 > kPlanTableLimit || p.segs.size() + 1 > kPlanGridLimit ||
        p.chunks.size() + G > kPlanGridLimit || p.items.size() + nitems > kPlanGridLimit ||
        p.sblocks.size() + ns / kThreads > kPlanGridLimit)
      return false;
    for (uint32_t y = 0; y < ns / kThreads; ++y) p.sblocks.push_back({(uint32_t)p.segs.size(), y});
    SegDesc s = {};
    s.data_off = g.offset;
    s.len = n;
    s.row_base = rowc;
    s.red_base = red;
    s.bmax = bmax;
    s.pbits = pb;
    s.nchunks = (uint32_t)G;
    s.sample_off = (uint32_t)sam;
    s.nsample = ns;
    s.lut_off = (uint32_t)lutc;
    s.tot_off = (uint32_t)totc;
    s.bits_off = (uint32_t)bitc;
    s.q_off = (uint32_t)qoff;
    s.q_cnt = g.num_k;
    p.qmax = std::max(p.qmax, s.q_cnt);
    s.mode = mode;
    s.radix_bits = rb;
    s.qsrc = topk ? QSRC_TOPK : QSRC_K;
    s.out_off = (uint32_t)g.out_offset;
    s.flags = 0;
    if (seg_level2(o, mode)) {
      const uint64_t hm = seg_heavy_max(n, g.num_k);
      if (hm > 0) {
        s.flags |= SEG_LEVEL2;
        hmax += hm;
        cmax += n / kChunk2 + hm;
      }
    }
    const uint32_t si = (uint32_t)p.segs.size();
    p.segs.push_back(s);
    for (uint64_t c = 0; c < G; ++c) {
      Chunk ch = {};
      ch.seg = si;
      ch.idx = (uint32_t)c;
      ch.start = c * csz;
      ch.len = std::min<uint64_t>(csz, n - c * csz);
      p.chunks.push_back(ch);
    }
    {
      const uint32_t gl = reduce_groups_log2(G), ng = (bmax + kReduceBins - 1) / kReduceBins;
      for (uint32_t gi = 0; gi < ng; gi += 1u << gl) p.items.push_back({si, (uint16_t)gi, (uint16_t)gl});
    }
    p.kvals.resize(qoff + g.num_k, 0u);
    if (!topk) memcpy(p.kvals.data() + qoff, g.k, (size_t)g.num_k * 4);
    p.bmax_max = std::max(p.bmax_max, bmax);
    p.pbits_max = std::max(p.pbits_max, pb);
    red += n;
    rowc += G * bmax;
    qoff += g.num_k;
    sam += ns;
    lutc += 1u << pb;
    totc += bmax;
    bitc += (bmax + 31) / 32;
  }
  if (hmax > kPlanTableLimit || cmax > kPlanGridLimit) return false;
  p.level2 = hmax > 0;
  p.hmax = (uint32_t)hmax;
  p.cmax = (uint32_t)cmax;
  p.nq = (uint32_t)qoff;
  p.nsegs = (uint32_t)p.segs.size();
  p.nchunks = (uint32_t)p.chunks.size();
  p.nitems = (uint32_t)p.items.size();
  p.nsblocks = (uint32_t)p.sblocks.size();
  // a batch of many segments builds its LUTs side by side, one work-group
  // each: split builds would only add work-groups that exit at once
  if (p.segs.size() <= kLutSplitSegs)
    for (const SegDesc& sd : p.segs)
      p.lut_parts = std::max(p.lut_parts, ggks::lut_parts(sd.mode, sd.nsample, (uint32_t)eb));
  p.ndirect = (uint32_t)p.direct_children.size();
  for (const Child& c : p.direct_children) p.dmax_keys = std::max<uint64_t>(p.dmax_keys, c.len);
  for (const SegDesc& sd : p.segs)
    p.emax_keys = std::max<uint64_t>(
        p.emax_keys, sd.mode == MODE_RADIX ? sd.len : (sd.len + sd.bmax - 1) / sd.bmax);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off = align_up(off + bytes);
    return at;
  };
  const size_t nseg = p.segs.size();
  p.off_tables = off;
  p.off_segs = take(nseg * sizeof(SegDesc));
  p.off_chunks = take(p.chunks.size() * sizeof(Chunk));
  p.off_items = take(p.items.size() * sizeof(ReduceItem));
  p.off_sblocks = take(p.sblocks.size() * sizeof(SampleBlock));
  p.off_kvals = take(p.kvals.size() * 4);
  p.off_dchildren = take(p.direct_children.size() * sizeof(Child));
  p.off_dqueries = take(p.direct_queries.size() * sizeof(Query));
  p.tables_end = off;
  p.off_states = take(nseg * sizeof(SegState));
  p.off_samples = take((size_t)std::max<uint64_t>(sam, 1) * eb);
  p.off_minmax = take((size_t)std::max<uint64_t>(sam / 1024, 1) * 4 * eb);
  p.off_lut = take((size_t)lutc * 4);
  p.off_rows = take((size_t)rowc * 4);
  p.off_tot = take((size_t)totc * 4);
  p.off_gsum = take((size_t)(totc / kReduceBins + 1) * 4);
  p.off_bits = take((size_t)bitc * 4);
  p.off_wbase = take((size_t)bitc * 4);
  p.off_q = off;  // engine segments read k lists or top-K ranks, no Query records
  p.off_qb = take((size_t)qoff * 4);
  p.off_qr = take((size_t)qoff * 4);
  p.off_children = take((size_t)qoff * sizeof(Child));
  p.off_cinfo = take((size_t)qoff * 8);
  p.off_cq = take((size_t)qoff * sizeof(Query));
  p.off_red = take((size_t)red * eb);
  p.off_big = take(4);
  layout_level2(p, take, red, qoff, eb);
  p.total = off;
  return true;
}

}  // namespace ggks

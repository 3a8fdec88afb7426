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

inline Tuning default_tuning(uint64_t n, int eb) {
  (void)eb;
  if (n >= (1ull << 26)) return {15, 256};
  if (n >= (1ull << 24)) return {13, 256};
  if (n >= (1ull << 22)) return {12, 128};
  if (n >= (1ull << 20)) return {11, 64};
  return {10, 32};
}

struct Plan {
  std::vector<SegDesc> segs;
  std::vector<Chunk> chunks;
  std::vector<ReduceItem> items;
  std::vector<Child> direct_children;
  bool direct = false;
  bool any_sample = false;
  uint32_t nq = 0;
  uint32_t max_nsample = 0;
  uint32_t bmax_max = 0;
  uint32_t pbits_max = 0;
  uint64_t seed = 0;
  // arena offsets (bytes)
  size_t off_tables = 0, off_segs = 0, off_chunks = 0, off_items = 0, tables_end = 0;
  size_t off_states = 0, off_samples = 0, off_lut = 0, off_rows = 0, off_tot = 0;
  size_t off_bits = 0, off_wbase = 0, off_cand = 0, off_q = 0, off_qb = 0, off_qr = 0;
  size_t off_children = 0, off_cq = 0, off_red = 0, total = 0;

  size_t table_bytes() const { return tables_end - off_tables; }

  void write_tables(char* dst) const {
    if (!segs.empty()) memcpy(dst + (off_segs - off_tables), segs.data(), segs.size() * sizeof(SegDesc));
    if (!chunks.empty())
      memcpy(dst + (off_chunks - off_tables), chunks.data(), chunks.size() * sizeof(Chunk));
    if (!items.empty())
      memcpy(dst + (off_items - off_tables), items.data(), items.size() * sizeof(ReduceItem));
    if (direct)
      memcpy(dst + (off_children - off_tables), direct_children.data(),
             direct_children.size() * sizeof(Child));
  }
};

// Plan a single-vector select of nq ranks over n keys.  The engine's single
// segment reads its queries straight from the k list (or as ranks of a
// top-K); the small-n direct path takes explicit Query records.
inline void plan_single(Plan& p, int kind, int dtype, uint64_t n, uint32_t nq,
                        const ggks_opts& o) {
  const int eb = dtype_bytes(dtype);
  const uint64_t cap = kFinishLdsBytes / (uint32_t)eb;
  p.nq = nq;
  p.seed = o.seed ? o.seed : 0x6A09E667F3BCC909ull;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off = align_up(off + bytes);
    return at;
  };
  if (n <= cap) {
    p.direct = true;
    Child c = {};
    c.start = 0;
    c.len = n;
    c.q_off = 0;
    c.q_cnt = nq;
    c.src = 0;
    c.need_range = 1;
    p.direct_children.push_back(c);
    p.off_tables = off;
    p.off_children = take(sizeof(Child));
    p.tables_end = off;
    p.off_q = take((size_t)nq * sizeof(Query));
    p.total = off;
    return;
  }
  Tuning t = default_tuning(n, eb);
  uint32_t mode = MODE_SAMPLE, rb = 0, bl, pb;
  if (kind == GGKS_KIND_RADIX) {
    mode = MODE_RADIX;
    rb = o.radix_bits ? (uint32_t)o.radix_bits : 14u;
    rb = std::min<uint32_t>(14, std::max<uint32_t>(4, rb));
    bl = rb + 1;
    pb = std::min<uint32_t>(kMaxPrefixBits, rb);
  } else {
    bl = o.buckets_log2 ? (uint32_t)o.buckets_log2 : t.buckets_log2;
    bl = std::min<uint32_t>(15, std::max<uint32_t>(8, bl));
    pb = std::min<uint32_t>(kMaxPrefixBits, bl - 2);
  }
  const uint32_t bmax = 1u << bl;
  uint64_t G = o.blocks > 0 ? (uint64_t)o.blocks : t.chunks;
  G = std::max<uint64_t>(1, std::min<uint64_t>(G, (n + 4095) / 4096));
  G = std::min<uint64_t>(G, kMaxChunks);
  uint64_t csz = (n + G - 1) / G;
  csz = (csz + 63) / 64 * 64;
  G = (n + csz - 1) / csz;
  uint32_t ns = 0;
  if (mode == MODE_SAMPLE) {
    uint64_t want = (n / 256 + 1023) / 1024 * 1024;
    ns = (uint32_t)std::min<uint64_t>(65536, std::max<uint64_t>(4096, want));
  }
  SegDesc s = {};
  s.data_off = 0;
  s.len = n;
  s.row_base = 0;
  s.red_base = 0;
  s.bmax = bmax;
  s.pbits = pb;
  s.nchunks = (uint32_t)G;
  s.sample_off = 0;
  s.nsample = ns;
  s.lut_off = 0;
  s.tot_off = 0;
  s.bits_off = 0;
  s.cand_off = 0;
  s.q_off = 0;
  s.q_cnt = nq;
  s.mode = mode;
  s.radix_bits = rb;
  s.qsrc = kind == GGKS_KIND_TOPK ? QSRC_TOPK : QSRC_K;
  s.out_off = 0;
  p.segs.push_back(s);
  for (uint64_t i = 0; i < G; ++i) {
    Chunk c = {};
    c.seg = 0;
    c.idx = (uint32_t)i;
    c.start = i * csz;
    c.len = std::min<uint64_t>(csz, n - i * csz);
    p.chunks.push_back(c);
  }
  for (uint32_t g = 0; g < (bmax + kReduceBins - 1) / kReduceBins; ++g) p.items.push_back({0u, g});
  p.any_sample = mode == MODE_SAMPLE;
  p.max_nsample = ns;
  p.bmax_max = bmax;
  p.pbits_max = pb;

  p.off_tables = off;
  p.off_segs = take(p.segs.size() * sizeof(SegDesc));
  p.off_chunks = take(p.chunks.size() * sizeof(Chunk));
  p.off_items = take(p.items.size() * sizeof(ReduceItem));
  p.tables_end = off;
  p.off_states = take(p.segs.size() * sizeof(SegState));
  p.off_samples = take((size_t)std::max<uint32_t>(ns, 1) * eb);
  p.off_lut = take((size_t)(1u << pb) * 4);
  p.off_rows = take((size_t)G * bmax * 4);
  p.off_tot = take((size_t)bmax * 4);
  p.off_bits = take((size_t)((bmax + 31) / 32) * 4);
  p.off_wbase = take((size_t)((bmax + 31) / 32) * 4);
  p.off_cand = take((size_t)(1u << pb) << (kCellBits - 3));
  p.off_q = take((size_t)nq * sizeof(Query));
  p.off_qb = take((size_t)nq * 4);
  p.off_qr = take((size_t)nq * 4);
  p.off_children = take((size_t)nq * sizeof(Child));
  p.off_cq = take((size_t)nq * sizeof(Query));
  p.off_red = take((size_t)n * eb);
  p.total = off;
}

}  // namespace ggks

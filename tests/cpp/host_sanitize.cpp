// Host-logic harness, built with g++ -fsanitize=address,undefined (no GPU):
// drives the library's own planner (plan_single, plan_batch), the server's
// scratch allocator (RangeAlloc) and its admission logic (serve_form_group,
// serve_admit, serve_remove_launched) through randomised and edge-case
// inputs, and checks their invariants.  Any sanitizer report aborts the run;
// a broken invariant prints FAIL and exits non-zero.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "ggks_host.hpp"

using namespace ggks;

static int g_bad = 0;
#define CHECK(c)                                                              \
  do {                                                                        \
    if (!(c)) {                                                               \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);               \
      if (++g_bad > 20) std::exit(1);                                         \
    }                                                                         \
  } while (0)

// Every arena region of a plan lies inside [0, total), 256-byte aligned, in
// layout order; the tables serialise into exactly table_bytes().
static void check_layout(const Plan& p) {
  const size_t offs[] = {p.off_tables, p.off_segs,  p.off_chunks,   p.off_items,
                         p.off_sblocks, p.tables_end, p.off_states,  p.off_samples,
                         p.off_minmax,  p.off_lut,    p.off_rows,    p.off_tot,
                         p.off_gsum,    p.off_bits,   p.off_wbase,   p.off_qb,
                         p.off_qr,      p.off_children, p.off_cq,    p.off_red};
  for (size_t o : offs) {
    CHECK(o % 256 == 0);
    CHECK(o <= p.total);
  }
  CHECK(p.tables_end >= p.off_tables);
  std::vector<char> tab(p.table_bytes() + 1, 0);
  p.write_tables(tab.data());  // the sanitizer checks the bounds
  for (const SegDesc& s : p.segs) {
    CHECK(s.bmax >= 256 && (s.bmax & (s.bmax - 1)) == 0);
    CHECK(s.tot_off % 256 == 0);  // bucket tables in whole 16-byte group-sum words
    CHECK(s.bits_off % 8 == 0);
    CHECK(s.nchunks >= 1 && s.nchunks <= kMaxChunks);
    CHECK((1u << s.pbits) <= (1u << kMaxPrefixBits));
    if (s.mode == MODE_RADIX) CHECK((1u << s.radix_bits) < s.bmax);
  }
  // chunks of every segment tile [0, len) in order
  std::vector<uint64_t> next(p.segs.size(), 0);
  std::vector<uint32_t> idx(p.segs.size(), 0);
  for (const Chunk& c : p.chunks) {
    CHECK(c.seg < p.segs.size());
    if (c.seg >= p.segs.size()) continue;
    CHECK(c.start == next[c.seg]);
    CHECK(c.idx == idx[c.seg]);
    CHECK(c.len > 0);
    next[c.seg] = c.start + c.len;
    idx[c.seg]++;
  }
  for (size_t i = 0; i < p.segs.size(); ++i) {
    CHECK(next[i] == p.segs[i].len);
    CHECK(idx[i] == p.segs[i].nchunks);
  }
  // reduce items cover every 64-bucket group of every segment once
  std::vector<uint32_t> groups(p.segs.size(), 0);
  for (const ReduceItem& it : p.items) {
    CHECK(it.seg < p.segs.size());
    if (it.seg >= p.segs.size()) continue;
    CHECK(it.group == groups[it.seg]);
    groups[it.seg] += 1u << it.gi_log2;
  }
  for (size_t i = 0; i < p.segs.size(); ++i)
    CHECK(groups[i] >= p.segs[i].bmax / kReduceBins);
}

static void test_plan_single() {
  const uint64_t ns[] = {1,        2,          63,         64,        16383,    16384,
                         16385,    8191,       8192,       8193,      65537,    1000003,
                         1u << 20, (1u << 22) + 5, 1u << 24, 1ull << 28, 0xFFFFFFFFull};
  for (int kind = 0; kind < 3; ++kind)
    for (int dt = 0; dt < 4; ++dt)
      for (uint64_t n : ns)
        for (uint32_t nq : {1u, 101u, 1024u, 5000u})
          for (int rb : {0, 4, 5, 6, 7, 8, 14})
            for (int bl : {0, 8, 15}) {
              if (kind == GGKS_KIND_TOPK && nq > n) continue;
              ggks_opts o = {};
              o.radix_bits = rb;
              o.buckets_log2 = bl;
              o.blocks = (int)(n % 7 == 0 ? 3 : 0);
              Plan p;
              plan_single(p, kind, dt, n, nq, o);
              check_layout(p);
              if (!p.direct) {
                CHECK(p.off_red + n * dtype_bytes(dt) <= p.total);
                CHECK(p.segs.size() == 1);
              }
            }
}

static void test_plan_batch(std::mt19937_64& rng) {
  for (int trial = 0; trial < 40; ++trial) {
    const int dt = (int)(rng() % 4);
    const size_t nseg = 1 + rng() % 300;
    std::vector<std::vector<uint32_t>> ks(nseg);
    std::vector<BatchSeg> in;
    uint64_t off = 0, out = 0;
    for (size_t i = 0; i < nseg; ++i) {
      BatchSeg g = {};
      g.length = (uint64_t)std::exp2(std::uniform_real_distribution<double>(0, 23)(rng));
      if (g.length < 1) g.length = 1;
      g.offset = off;
      off += g.length;
      g.kind = (uint32_t)(rng() % 3);
      g.num_k = (uint32_t)(rng() % 65);
      if (g.kind == GGKS_KIND_TOPK) {
        g.num_k = (uint32_t)std::min<uint64_t>(g.num_k, g.length);
      } else {
        for (uint32_t j = 0; j < g.num_k; ++j) ks[i].push_back(1 + (uint32_t)(rng() % g.length));
        g.k = ks[i].data();
      }
      g.out_offset = out;
      out += g.num_k;
      in.push_back(g);
    }
    ggks_opts o = {};
    o.radix_bits = (int)(rng() % 12);
    Plan p;
    CHECK(plan_batch(p, dt, in, o));
    check_layout(p);
    uint64_t red = 0;
    for (const SegDesc& s : p.segs) {
      CHECK(s.red_base == red);
      red += s.len;
    }
    CHECK(p.off_red + red * dtype_bytes(dt) <= p.total);
  }
  // engine segments summing to more than 2^32 keys: element offsets are
  // 64-bit, so this plans (and the reduced-buffer bases pass 2^32)
  {
    std::vector<BatchSeg> in;
    std::vector<uint32_t> k = {1, 7, 1u << 29};
    for (int i = 0; i < 5; ++i) {
      BatchSeg g = {};
      g.offset = (uint64_t)i << 27;
      g.length = 1ull << 30;
      g.k = k.data();
      g.num_k = 3;
      g.kind = GGKS_KIND_BUCKET;
      g.out_offset = 3ull * i;
      in.push_back(g);
    }
    Plan p;
    ggks_opts o = {};
    CHECK(plan_batch(p, GGKS_U32, in, o));
    CHECK(p.segs.size() == 5 && p.segs[4].red_base == 4ull << 30);
    check_layout(p);
  }
  // a table whose 32-bit bucket-table offsets would overflow is refused
  {
    const size_t nseg = 131073;  // x 2^15 buckets > 2^32
    std::vector<BatchSeg> in;
    uint32_t one = 1;
    for (size_t i = 0; i < nseg; ++i) {
      BatchSeg g = {};
      g.offset = 0;
      g.length = 20000;
      g.k = &one;
      g.num_k = 1;
      g.kind = GGKS_KIND_BUCKET;
      g.out_offset = i;
      in.push_back(g);
    }
    ggks_opts o = {};
    o.buckets_log2 = 15;
    Plan p;
    CHECK(!plan_batch(p, GGKS_F32, in, o));
    in.resize(131071);
    Plan q;
    CHECK(plan_batch(q, GGKS_F32, in, o));
  }
}

static void test_range_alloc(std::mt19937_64& rng) {
  RangeAlloc a;
  const size_t size = 1 << 24;
  a.reset(size);
  struct R { size_t off, bytes; };
  std::vector<R> live;
  size_t off = 0;
  CHECK(!a.alloc(0, off));
  CHECK(!a.alloc(SIZE_MAX, off));  // rounding must not wrap
  CHECK(!a.alloc(size + 1, off));
  for (int step = 0; step < 20000; ++step) {
    if (live.empty() || rng() % 3) {
      const size_t bytes = 1 + rng() % (size / 64);
      if (a.alloc(bytes, off)) {
        CHECK(off % 256 == 0 && off + align_up(bytes) <= size);
        for (const R& r : live)  // no overlap with a live range
          CHECK(off + align_up(bytes) <= r.off || r.off + align_up(r.bytes) <= off);
        live.push_back({off, bytes});
      }
    } else {
      const size_t i = rng() % live.size();
      a.release(live[i].off, live[i].bytes);
      live[i] = live.back();
      live.pop_back();
    }
    size_t used = 0;
    for (const R& r : live) used += align_up(r.bytes);
    CHECK(a.used() == used);
  }
  for (const R& r : live) a.release(r.off, r.bytes);
  CHECK(a.used() == 0);
  CHECK(a.free_ranges().size() == 1 && a.free_ranges().begin()->second == size);
}

// The server's admission loop without the device: requests of three dtypes
// arrive, batches are formed and admitted into a small budget, a random
// in-flight batch completes and frees its range, until every request ran
// exactly once, in queue order within its dtype.
static void test_admission(std::mt19937_64& rng) {
  std::vector<ServeReq> reqs(3000);
  for (size_t i = 0; i < reqs.size(); ++i) {
    ServeReq& q = reqs[i];
    q.r = {};
    q.r.dtype = (int)(rng() % 3);
    q.r.length = (uint64_t)std::exp2(std::uniform_real_distribution<double>(12, 22)(rng));
    q.r.offset = 0;
    q.r.kind = (int)(rng() % 3);
    q.r.num_k = 1 + (uint32_t)(rng() % 64);
    if (q.r.kind == GGKS_KIND_TOPK) {
      q.r.num_k = (uint32_t)std::min<uint64_t>(q.r.num_k, q.r.length);
    } else {
      for (uint32_t j = 0; j < q.r.num_k; ++j) q.k.push_back(1 + (uint32_t)(rng() % q.r.length));
    }
  }
  RangeAlloc a;
  a.reset(size_t(256) << 20);  // 256 MB: forces halving and waiting
  std::deque<size_t> pending;
  struct Flight { size_t off, foot; std::vector<size_t> ids; };
  std::vector<Flight> flying;
  std::vector<int> ran(reqs.size(), 0);
  std::vector<int64_t> last_of(4, -1);
  size_t next = 0;
  for (int step = 0; step < 200000 && (next < reqs.size() || !pending.empty() || !flying.empty());
       ++step) {
    for (int j = 0; j < 3 && next < reqs.size(); ++j) pending.push_back(next++);
    if (flying.size() < 4 && !pending.empty()) {
      const std::vector<size_t> group = serve_form_group(reqs, pending, 64);
      CHECK(!group.empty() && group[0] == pending.front());
      for (size_t id : group) CHECK(reqs[id].r.dtype == reqs[group[0]].r.dtype);
      Plan p;
      size_t foot = 0, off = 0;
      const size_t take = serve_admit(a, reqs, group, p, foot, off);
      if (take) {
        CHECK(foot >= p.total);
        std::vector<size_t> ids(group.begin(), group.begin() + take);
        for (size_t id : ids) {
          CHECK(ran[id] == 0);
          ran[id] = 1;
          const int d = reqs[id].r.dtype;
          CHECK((int64_t)id > last_of[d]);  // FIFO within a dtype
          last_of[d] = (int64_t)id;
        }
        serve_remove_launched(pending, ids);
        for (size_t id : pending)
          for (size_t x : ids) CHECK(id != x);
        flying.push_back({off, foot, ids});
      }
    }
    if (!flying.empty() && (rng() % 2 || flying.size() >= 4 || next >= reqs.size())) {
      const size_t i = rng() % flying.size();
      a.release(flying[i].off, flying[i].foot);
      flying[i] = flying.back();
      flying.pop_back();
    }
  }
  for (int r : ran) CHECK(r == 1);
  CHECK(a.used() == 0);
}

int main() {
  std::mt19937_64 rng(12345);
  test_plan_single();
  std::printf("plan_single ok\n");
  test_plan_batch(rng);
  std::printf("plan_batch ok\n");
  test_range_alloc(rng);
  std::printf("range allocator ok\n");
  test_admission(rng);
  std::printf("admission ok\n");
  std::printf("%s\n", g_bad ? "FAILED" : "ALL OK");
  return g_bad ? 1 : 0;
}

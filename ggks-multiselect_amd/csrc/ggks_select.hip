// ggks_select.hip -- host orchestration and the C ABI of libggks.so.
//
// One call = one plan: a table of top-level segments (a single vector is one
// segment), their chunks and reduce items, laid out in the context's arena.
// The plan tables go to the device with one async copy from pinned memory,
// then the level kernels run back to back on the context's stream.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ggks.h"
#include "ggks_kernels.hpp"
#include "ggks_plan.hpp"
#include "ggks_host.hpp"
#include "ggks_gen.hpp"

namespace ggks {

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(GGKS_EDEVICE, "%s failed: %s (%s:%d)", #expr,                \
                  hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)

// ------------------------------------------------------------------ context
// Host staging for the plan tables, k lists and outputs: a ring of pinned
// buffers, each with an event recorded after the async copies that read it,
// so a call waits only for the copies issued kPinSlots calls earlier
// instead of draining the stream (back-to-back async selects with changing
// tables keep the host ahead of the device).
constexpr int kPinSlots = 4;
struct PinSlot {
  char* buf = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool busy = false;  // a copy reading buf may be in flight until ev completes
};
}  // namespace ggks

struct ggks_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  char* arena = nullptr;
  size_t arena_cap = 0;
  ggks::PinSlot pin[ggks::kPinSlots];
  int pin_next = 0;                  // slot the next pin_acquire takes
  ggks::PinSlot* pin_cur = nullptr;  // slot of the call in progress
  uint32_t* status = nullptr;  // device word
  uint32_t* hgram = nullptr;   // sample histogram (Exec::hgram), zero between selects
  std::mutex mu;
  // stage timing (ggks_ctx_profile_arm)
  std::vector<hipEvent_t> prof;  // nslots * (GGKS_NSTAGES + 1)
  int prof_slots = 0;
  uint64_t prof_calls = 0;
  int prof_cur = -1;  // slot of the call in flight, -1 when not armed
  // events to record (ggks_ctx_profile_marks; default: every stage's, as
  // markers), and the mask latched by the last arm: recorded and read
  uint32_t prof_mask = (1u << (GGKS_NSTAGES + 1)) - 1u;
  uint32_t prof_armed = (1u << (GGKS_NSTAGES + 1)) - 1u;
  // plan-table cache: the tables in the arena match this signature
  uint64_t table_sig = 0;
  ggks_debug_regions dbg = {};  // the last select's tables (ggks_ctx_debug_regions)
};

namespace ggks {

static int ensure_arena(ggks_ctx* c, size_t bytes) {
  if (bytes <= c->arena_cap) return GGKS_OK;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->arena) HIP_TRY(hipFree(c->arena));
  c->arena = nullptr;
  c->arena_cap = 0;
  size_t cap = bytes + bytes / 8;
  if (hipMalloc(&c->arena, cap) != hipSuccess) {
    c->arena = nullptr;
    (void)hipGetLastError();
    return fail(GGKS_ENOMEM, "cannot allocate %zu bytes of device scratch", cap);
  }
  c->arena_cap = cap;
  c->table_sig = 0;
  return GGKS_OK;
}

// Wait until no copy reads the slot, then make it at least `bytes` long.
static int pin_ready(PinSlot& s, size_t bytes) {
  if (s.busy) {
    HIP_TRY(hipEventSynchronize(s.ev));
    s.busy = false;
  }
  // the event only tells the host that the copies reading the buffer are
  // done, so it needs neither timing nor a system-scope release
  if (!s.ev) HIP_TRY(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming | hipEventDisableSystemFence));
  if (bytes <= s.cap) return GGKS_OK;
  if (s.buf) HIP_TRY(hipHostFree(s.buf));
  s.buf = nullptr;
  s.cap = 0;
  const size_t cap = std::max<size_t>(bytes + bytes / 8, 1 << 16);
  if (hipHostMalloc(&s.buf, cap, hipHostMallocDefault) != hipSuccess) {
    s.buf = nullptr;
    (void)hipGetLastError();
    return fail(GGKS_ENOMEM, "cannot allocate %zu bytes of pinned memory", cap);
  }
  s.cap = cap;
  return GGKS_OK;
}

// The next staging buffer of the ring, at least `bytes` long.
static int pin_acquire(ggks_ctx* c, size_t bytes, char** out) {
  PinSlot& s = c->pin[c->pin_next];
  c->pin_next = (c->pin_next + 1) % kPinSlots;
  int rc = pin_ready(s, bytes);
  if (rc) return rc;
  c->pin_cur = &s;
  *out = s.buf;
  return GGKS_OK;
}

// The async copies reading the acquired buffer are enqueued: mark their end.
static int pin_mark(ggks_ctx* c) {
  HIP_TRY(hipEventRecord(c->pin_cur->ev, c->stream));
  c->pin_cur->busy = true;
  return GGKS_OK;
}

// The stream has drained: no staging buffer is read any more.
static void pin_all_free(ggks_ctx* c) {
  for (PinSlot& s : c->pin) s.busy = false;
}

static int ctx_init(ggks_ctx* c, int device, hipStream_t stream) {
  c->device = device;
  c->stream = stream;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipMalloc(&c->status, 256));
  HIP_TRY(hipMemset(c->status, 0, 256));
  HIP_TRY(hipMalloc(&c->hgram, (size_t)4 << kMaxPrefixBits));
  HIP_TRY(hipMemset(c->hgram, 0, (size_t)4 << kMaxPrefixBits));
  return GGKS_OK;
}

static std::mutex g_default_mu;
static ggks_ctx* g_default[64];

static int default_ctx(ggks_ctx** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(GGKS_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default[dev]) {
    ggks_ctx* c = new ggks_ctx();
    int rc = ctx_init(c, dev, nullptr);
    if (rc) { delete c; return rc; }
    g_default[dev] = c;
  }
  *out = g_default[dev];
  return GGKS_OK;
}

// Record the event that opens `stage` (GGKS_NSTAGES closes the call).  Stages
// a call skips are recorded at the same point, so they measure zero.
static void mark(ggks_ctx* c, int stage) {
  if (c->prof_cur < 0 || !((c->prof_armed >> stage) & 1u)) return;
  if (c->prof_armed & GGKS_PROFILE_KERNEL_EVENTS) return;  // carried by the kernel launches
  const int per = GGKS_NSTAGES + 1;
  (void)hipEventRecord(c->prof[(size_t)c->prof_cur * per + stage], c->stream);
}

// Stages whose kernel launch can carry their events (one kernel each).
static bool is_kernel_stage(int stage) { return stage == 3 || stage == 6; }

// The events of single-kernel stage `stage` (its own and the next stage's)
// when they are armed as kernel events, else null: the launch of that
// stage's kernel then carries them as its start and stop events.
static void kernel_events(ggks_ctx* c, int stage, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (!c || c->prof_cur < 0 || !(c->prof_armed & GGKS_PROFILE_KERNEL_EVENTS)) return;
  if (!is_kernel_stage(stage)) return;
  if (!((c->prof_armed >> stage) & 1u) || !((c->prof_armed >> (stage + 1)) & 1u)) return;
  const int per = GGKS_NSTAGES + 1;
  *e0 = c->prof[(size_t)c->prof_cur * per + stage];
  *e1 = c->prof[(size_t)c->prof_cur * per + stage + 1];
}

// Launch `kernel` on `s`; with stage `stage`'s events armed as kernel
// events, as their start and stop events.
template <typename K, typename... Args>
static void launch_stage(ggks_ctx* prof, int stage, K kernel, dim3 grid, dim3 block, uint32_t lds,
                         hipStream_t s, Args... args) {
  hipEvent_t e0, e1;
  kernel_events(prof, stage, &e0, &e1);
  if (e0)
    hipExtLaunchKernelGGL(kernel, grid, block, lds, s, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

// ----------------------------------------------------------- kernel launch
template <typename K>
static void set_lds_limit(K kernel, size_t bytes) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Where one launch sequence runs: its arena (plan tables + scratch), stream,
// status word and, when stage timing is armed, the context recording it.
struct Exec {
  char* arena;
  hipStream_t stream;
  uint32_t* status;
  ggks_ctx* prof;  // null: no stage events (also during graph capture)
  // optional: a second stream (and fork/join events) on which the direct
  // children are finished beside the engine pipeline
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // optional: a zeroed histogram of 2^kMaxPrefixBits words for a plan of one
  // sampled segment (k_sample counts the sample into it, k_lut reads it and
  // clears it again), owned by the context
  uint32_t* hgram = nullptr;
};

static void stage(const Exec& x, int st) {
  if (x.prof) mark(x.prof, st);
}

// Dynamic-LDS limits of the level kernels: the largest size any plan can
// request, set once per instantiation (the limit is a per-function attribute
// shared by every plan and graph; occupancy follows the size each launch
// actually asks for).
template <typename KT, int C>
static void level_attrs(const Plan&) {
  static const bool once = [] {
    set_lds_limit(k_count<KT, C>, ((size_t)kMaxBuckets + kLaneSlots + (1u << kMaxPrefixBits)) * 4);
    set_lds_limit(k_findk<KT, C>, ((size_t)kMaxBuckets + (1u << kMaxPrefixBits) +
                                   3 * ((kMaxBuckets + 31) / 32) + 32) * 4);
    set_lds_limit(k_findk_small<KT, C>, ((size_t)kFindkSmallBuckets + (1u << kMaxPrefixBits) +
                                         3 * ((kFindkSmallBuckets + 31) / 32) + 32) * 4);
    set_lds_limit(k_compact<KT, C>, compact_lds_bytes<false>());
    set_lds_limit(k_compact_win<KT, C>, compact_lds_bytes<false>());
    set_lds_limit(k_compact_small<KT, C>, compact_lds_bytes<true>());
    set_lds_limit(k_finish<KT, C>, kFinishLdsTotal);
    return true;
  }();
  (void)once;
}

// Both finishers over a child list: a 256-thread work-group per child with
// few ranks (or many ranks and at most 1024 keys), a 1024-thread work-group
// per child otherwise; each kernel screens out the other's children.  The
// 1024-thread finisher's LDS allows one group per CU, so a grid of
// max_groups (256: one per CU) loops over the list without losing
// concurrency.
template <typename KT, int C>
static void finish(hipStream_t s, const KT* d_src, const KT* red, const KT* red2,
                   const Child* kids, uint32_t n, const Query* cq, const Query* cq2, KT* d_out,
                   uint32_t max_groups, bool lite, bool big, const uint32_t* bigw) {
  bool launched = false;
  if constexpr (sizeof(KT) == 4) {
    if (lite) {
      hipLaunchKernelGGL((k_finish_small_lite<KT, C>), dim3(std::min<uint32_t>(n, 65536)),
                         dim3(kFThreads), 0, s, d_src, red, red2, kids, n, cq, cq2, d_out);
      launched = true;
    }
  }
  if (!launched)
    hipLaunchKernelGGL((k_finish_small<KT, C>), dim3(std::min<uint32_t>(n, 65536)),
                       dim3(kFThreads), 0, s, d_src, red, red2, kids, n, cq, cq2, d_out);
  // big = false: no child for k_finish (known on the host); bigw: a device
  // count of them, read by k_finish before anything else
  if (big)
    hipLaunchKernelGGL((k_finish<KT, C>), dim3(std::min<uint32_t>(n, max_groups)), dim3(kThreads),
                       kFinishLdsTotal, s, d_src, red, red2, kids, n, cq, cq2, d_out, bigw);
}

// Grid of the second-level kernels (they walk the chunk table with a grid
// stride): two 1024-thread work-groups per CU.
constexpr uint32_t kLevel2Groups = 512;

// The level pipeline over every engine segment of the plan, and the
// finisher over the plan's direct children (whole small segments): after
// the pipeline, or, with a side stream, beside it from the streaming count
// pass on (the two read the input and write disjoint outputs; the side
// stream is joined back before returning, also under graph capture, where
// it forms a parallel branch).  Forked any earlier, the direct finishers
// hold CUs that the many one-work-group-per-segment LUT builds wait for.
template <typename KT, int C>
static int enqueue_level(const Exec& x, const Plan& p, const KT* d_src, KT* d_out,
                         const uint32_t* d_k) {
  char* A = x.arena;
  hipStream_t s = x.stream;
  const uint32_t ndirect = p.ndirect;
  const bool beside = ndirect > 0 && p.nsegs > 0 && x.side;
  // the register-light 256-thread finisher when every child a plan can
  // make is expected to fit its registers twice over (a longer child is
  // still finished exactly, from memory): direct children by their length,
  // an engine segment's by its mean bucket size
  const uint64_t lite_keys = finish_small_reg_keys<KT>(true);
  const bool dlite = p.dmax_keys <= lite_keys, elite = 2 * p.emax_keys <= lite_keys;
  // whether a direct child (planned on the host) is one for k_finish; a plan
  // without its host tables (the no-allocation path) may have one
  bool dbig = ndirect > 0 && p.direct_children.empty();
  for (const Child& c : p.direct_children) dbig = dbig || big_child(c.len, c.q_cnt);
  auto finish_direct = [&](hipStream_t ds) {
    const Child* dk = reinterpret_cast<const Child*>(A + p.off_dchildren);
    const Query* dq = reinterpret_cast<const Query*>(A + p.off_dqueries);
    finish<KT, C>(ds, d_src, (const KT*)nullptr, (const KT*)nullptr, dk, ndirect, dq,
                  (const Query*)nullptr, d_out, 256, dlite, dbig, (const uint32_t*)nullptr);
  };
  const uint32_t nseg = p.nsegs;
  if (nseg > 0) {
    const SegDesc* segs = reinterpret_cast<const SegDesc*>(A + p.off_segs);
    SegState* states = reinterpret_cast<SegState*>(A + p.off_states);
    const Chunk* chunks = reinterpret_cast<const Chunk*>(A + p.off_chunks);
    const ReduceItem* items = reinterpret_cast<const ReduceItem*>(A + p.off_items);
    KT* samples = reinterpret_cast<KT*>(A + p.off_samples);
    KT* minmax = reinterpret_cast<KT*>(A + p.off_minmax);
    uint32_t* lut = reinterpret_cast<uint32_t*>(A + p.off_lut);
    uint32_t* rows = reinterpret_cast<uint32_t*>(A + p.off_rows);
    uint32_t* tot = reinterpret_cast<uint32_t*>(A + p.off_tot);
    uint32_t* gsum = reinterpret_cast<uint32_t*>(A + p.off_gsum);
    uint32_t* bits = reinterpret_cast<uint32_t*>(A + p.off_bits);
    uint32_t* wbase = reinterpret_cast<uint32_t*>(A + p.off_wbase);
    const Query* q = reinterpret_cast<const Query*>(A + p.off_q);
    uint32_t* qb = reinterpret_cast<uint32_t*>(A + p.off_qb);
    uint32_t* qr = reinterpret_cast<uint32_t*>(A + p.off_qr);
    Child* kids = reinterpret_cast<Child*>(A + p.off_children);
    Query* cq = reinterpret_cast<Query*>(A + p.off_cq);
    KT* red = reinterpret_cast<KT*>(A + p.off_red);
    const uint32_t nchunks = p.nchunks;
    // second level (null pointers when the plan has none)
    uint32_t* l2count = p.level2 ? reinterpret_cast<uint32_t*>(A + p.off_hcount) : nullptr;
    Heavy* heavy = p.level2 ? reinterpret_cast<Heavy*>(A + p.off_heavy) : nullptr;
    Chunk2* chunks2 = p.level2 ? reinterpret_cast<Chunk2*>(A + p.off_chunks2) : nullptr;
    uint32_t* rows2 = p.level2 ? reinterpret_cast<uint32_t*>(A + p.off_rows2) : nullptr;
    uint32_t* bits2 = p.level2 ? reinterpret_cast<uint32_t*>(A + p.off_bits2) : nullptr;
    uint32_t* wb2 = p.level2 ? reinterpret_cast<uint32_t*>(A + p.off_wb2) : nullptr;
    Query* cq2 = p.level2 ? reinterpret_cast<Query*>(A + p.off_cq2) : nullptr;
    KT* red2 = p.level2 ? reinterpret_cast<KT*>(A + p.off_red2) : nullptr;

    // the context's sample histogram: one segment, and k_sample runs
    uint32_t* hg = (nseg == 1 && p.nsblocks > 0) ? x.hgram : nullptr;
    stage(x, 1);
    if (p.nsblocks > 0) {
      const SampleBlock* sblocks = reinterpret_cast<const SampleBlock*>(A + p.off_sblocks);
      hipLaunchKernelGGL((k_sample<KT, C>), dim3(p.nsblocks), dim3(kSampleThreads), 0, s,
                         d_src, segs, sblocks, samples, minmax, p.seed, states, lut, hg);
    }
    stage(x, 2);
    uint32_t* bigw = reinterpret_cast<uint32_t*>(A + p.off_big);
    uint2* cinfo = reinterpret_cast<uint2*>(A + p.off_cinfo);
    hipLaunchKernelGGL((k_lut<KT>), dim3(nseg, p.lut_parts), dim3(kThreads), 0, s, segs, samples,
                       minmax, states,
                       lut, l2count, bigw, hg);
    stage(x, 3);
    if (beside) {  // nseg > 0
      HIP_TRY(hipEventRecord(x.fork, s));
      HIP_TRY(hipStreamWaitEvent(x.side, x.fork, 0));
      finish_direct(x.side);
      HIP_TRY(hipEventRecord(x.join, x.side));
    }
    launch_stage(x.prof, 3, k_count<KT, C>, dim3(nchunks), dim3(kThreads),
                 (uint32_t)(((size_t)p.bmax_max + kLaneSlots + (1u << p.pbits_max)) * 4), s, d_src, segs, states,
                 chunks, lut, rows);
    stage(x, 4);
    hipLaunchKernelGGL(k_reduce, dim3(p.nitems), dim3(kThreads), 0, s, segs,
                       states, items, rows, tot, gsum);
    stage(x, 5);
    if (p.bmax_max <= kFindkSmallBuckets)
      hipLaunchKernelGGL((k_findk_small<KT, C>), dim3(nseg), dim3(kThreads),
                         ((size_t)p.bmax_max + (1u << p.pbits_max) + 3 * ((p.bmax_max + 31) / 32) +
                          32) * 4,
                         s, segs, states, lut, tot, gsum, q, d_k, x.status, qb, qr, kids, cq, bits, wbase,
                         d_out, l2count, heavy, chunks2, p.hmax, bigw, cinfo);
    else
      hipLaunchKernelGGL((k_findk<KT, C>), dim3(nseg), dim3(kThreads),
                         ((size_t)p.bmax_max + (1u << p.pbits_max) + 3 * ((p.bmax_max + 31) / 32) +
                          32) * 4,
                         s, segs, states, lut, tot, gsum, q, d_k, x.status, qb, qr, kids, cq, bits, wbase,
                         d_out, l2count, heavy, chunks2, p.hmax, bigw, cinfo);
    stage(x, 6);
    if (compact_small(p.pbits_max, p.bmax_max))
      launch_stage(x.prof, 6, k_compact_small<KT, C>, dim3(nchunks), dim3(kThreads),
                   (uint32_t)compact_lds_bytes<true>(), s, d_src, segs, states, chunks, lut, rows,
                   bits, wbase, cinfo, red);
    else if (p.qmax > CompactCfg<false>::template cursors<KT>)  // cursor windows possible
      launch_stage(x.prof, 6, k_compact_win<KT, C>, dim3(nchunks), dim3(kThreads),
                   (uint32_t)compact_lds_bytes<false>(), s, d_src, segs, states, chunks, lut, rows,
                   bits, wbase, cinfo, red);
    else
      launch_stage(x.prof, 6, k_compact<KT, C>, dim3(nchunks), dim3(kThreads),
                   (uint32_t)compact_lds_bytes<false>(), s, d_src, segs, states, chunks, lut, rows,
                   bits, wbase, cinfo, red);
    stage(x, 7);
    if (p.level2) {  // the heavy children's second level (part of the finish stage)
      const uint32_t g2 = std::max<uint32_t>(1, std::min<uint32_t>(p.cmax, kLevel2Groups));
      hipLaunchKernelGGL((k_count2<KT, C>), dim3(g2), dim3(kThreads), 0, s, red, heavy, chunks2,
                         l2count, p.cmax, rows2);
      const uint32_t gr = std::max<uint32_t>(1, std::min<uint32_t>(p.hmax, kLevel2Groups));
      hipLaunchKernelGGL((k_resolve2<KT, C>), dim3(gr), dim3(kThreads), 0, s, heavy, l2count, p.hmax,
                         rows2, cq, cq2, qb, qr, kids, states, segs, bits2, wb2, d_out);
      hipLaunchKernelGGL((k_compact2<KT, C>), dim3(g2), dim3(kThreads), 0, s, red, heavy, chunks2,
                         l2count, p.cmax, rows2, bits2, wb2, red2);
    }
    // the second level's children are not counted: k_finish screens them
    finish<KT, C>(s, d_src, red, red2, kids, p.nq, cq, cq2, d_out, 256, elite, true,
                  p.level2 ? (const uint32_t*)nullptr : bigw);
  } else {
    for (int st = 1; st < GGKS_NSTAGES; ++st) stage(x, st);
  }
  if (beside) {
    HIP_TRY(hipStreamWaitEvent(s, x.join, 0));
  } else if (ndirect > 0) {
    finish_direct(s);
  }
  HIP_TRY(hipGetLastError());
  return GGKS_OK;
}

template <typename KT, int C>
static int launch_typed(const Exec& x, const Plan& p, const void* d_src, void* d_out,
                        const uint32_t* d_k) {
  level_attrs<KT, C>(p);
  return enqueue_level<KT, C>(x, p, (const KT*)d_src, (KT*)d_out, d_k);
}

static int launch(const Exec& x, int dtype, const Plan& p, const void* d_src, void* d_out,
                  const uint32_t* d_k) {
  switch (dtype) {
    case GGKS_F32: return launch_typed<uint32_t, CONV_F>(x, p, d_src, d_out, d_k);
    case GGKS_F64: return launch_typed<uint64_t, CONV_F>(x, p, d_src, d_out, d_k);
    case GGKS_U32: return launch_typed<uint32_t, CONV_U>(x, p, d_src, d_out, d_k);
    case GGKS_I32: return launch_typed<uint32_t, CONV_I>(x, p, d_src, d_out, d_k);
  }
  return fail(GGKS_EINVAL, "bad dtype %d", dtype);
}

// FNV-1a over the plan tables (decides whether the device copy is current).
static uint64_t table_signature(const char* p, size_t n, const void* arena) {
  uint64_t h = 1469598103934665603ull ^ (uint64_t)(uintptr_t)arena;
  const uint64_t* w = reinterpret_cast<const uint64_t*>(p);
  for (size_t i = 0; i < n / 8; ++i) h = (h ^ w[i]) * 1099511628211ull;
  for (size_t i = n / 8 * 8; i < n; ++i) h = (h ^ (uint8_t)p[i]) * 1099511628211ull;
  return h;
}

// -------------------------------------------------------------- one select
enum class KSource { HOST, DEVICE, TOPK };

static int check_common(int dtype, const void* d_vec, int64_t n, int64_t num_k) {
  if (dtype < 0 || dtype > 3) return fail(GGKS_EINVAL, "bad dtype %d", dtype);
  if (!d_vec) return fail(GGKS_EINVAL, "null vector");
  if (n < 1 || n > (int64_t)0xFFFFFFFFll) return fail(GGKS_EINVAL, "n=%lld outside [1, 2^32)", (long long)n);
  if (num_k < 0 || num_k > (int64_t)0x7FFFFFFFll) return fail(GGKS_EINVAL, "num_k=%lld", (long long)num_k);
  return GGKS_OK;
}

static int select_one(ggks_ctx* c, int kind, int dtype, const void* d_vec, int64_t n,
                      const uint32_t* k, int64_t num_k, KSource ksrc, void* out,
                      bool host_out, const ggks_opts* opt) {
  int rc = check_common(dtype, d_vec, n, num_k);
  if (rc) return rc;
  if (ksrc == KSource::HOST) {
    if (num_k > 0 && (!k || !out)) return fail(GGKS_EINVAL, "null k or output array");
    for (int64_t i = 0; i < num_k; ++i)
      if (k[i] < 1 || (int64_t)k[i] > n)
        return fail(GGKS_EINVAL, "k[%lld]=%u outside [1, %lld]", (long long)i, k[i], (long long)n);
  } else if (ksrc == KSource::DEVICE) {
    if (num_k > 0 && (!k || !out)) return fail(GGKS_EINVAL, "null k or output array");
  } else {
    if (num_k < 1 || num_k > n) return fail(GGKS_EINVAL, "K=%lld outside [1, n]", (long long)num_k);
    if (!out) return fail(GGKS_EINVAL, "null output array");
  }
  if (num_k == 0) return GGKS_OK;
  if (!c && (rc = default_ctx(&c))) return rc;
  ggks_opts o = {};
  if (opt) o = *opt;
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));

  Plan p;
  plan_single(p, kind, dtype, (uint64_t)n, (uint32_t)num_k, o);
  const size_t eb = (size_t)dtype_bytes(dtype);
  // device output staging for host outputs, device k copy for host k lists
  const size_t off_dout = p.total;
  const size_t off_dk = align_up(off_dout + (host_out ? (size_t)num_k * eb : 0));
  const size_t total = align_up(off_dk + (ksrc == KSource::HOST ? (size_t)num_k * 4 : 0));
  if ((rc = ensure_arena(c, total))) return rc;
  const size_t tbytes = p.table_bytes();
  const size_t pin_k = align_up(tbytes);
  const size_t pin_out = align_up(pin_k + (ksrc == KSource::HOST ? (size_t)num_k * 4 : 0));
  const size_t pin_total = pin_out + (host_out ? (size_t)num_k * eb : 0);

  c->prof_cur = c->prof_slots > 0 ? (int)(c->prof_calls++ % (uint64_t)c->prof_slots) : -1;
  mark(c, 0);
  // plan tables: upload only when they differ from what the arena holds
  std::vector<char> tab(tbytes);
  p.write_tables(tab.data());
  const uint64_t sig = table_signature(tab.data(), tbytes, c->arena);
  const bool up_tab = sig != c->table_sig, up_k = ksrc == KSource::HOST;
  char* pin = nullptr;
  if (up_tab || up_k || host_out) {
    if ((rc = pin_acquire(c, pin_total, &pin))) return rc;
  }
  if (up_tab) {
    memcpy(pin, tab.data(), tbytes);
    HIP_TRY(hipMemcpyAsync(c->arena + p.off_tables, pin, tbytes, hipMemcpyHostToDevice,
                           c->stream));
    c->table_sig = sig;
  }
  const uint32_t nk = (uint32_t)num_k;
  const uint32_t qgrid = (nk + 255) / 256;
  const uint32_t* d_k = k;
  if (up_k) {
    memcpy(pin + pin_k, k, (size_t)num_k * 4);
    HIP_TRY(hipMemcpyAsync(c->arena + off_dk, pin + pin_k, (size_t)num_k * 4,
                           hipMemcpyHostToDevice, c->stream));
    d_k = reinterpret_cast<const uint32_t*>(c->arena + off_dk);
  }
  if ((up_tab || up_k) && (rc = pin_mark(c))) return rc;
  if (p.direct) {  // the direct path takes explicit Query records
    Query* d_q = reinterpret_cast<Query*>(c->arena + p.off_dqueries);
    if (ksrc == KSource::TOPK)
      hipLaunchKernelGGL(k_topk_queries, dim3(qgrid), dim3(256), 0, c->stream, nk, (uint64_t)n,
                         d_q);
    else
      hipLaunchKernelGGL(k_queries, dim3(qgrid), dim3(256), 0, c->stream, d_k, nk, (uint64_t)n,
                         d_q, c->status);
  }
  void* d_out = host_out ? (void*)(c->arena + off_dout) : out;
  Exec x = {c->arena, c->stream, c->status, c->prof_cur >= 0 ? c : nullptr};
  x.hgram = c->hgram;
  rc = launch(x, dtype, p, d_vec, d_out, d_k);
  {
    ggks_debug_regions& g = c->dbg;
    g = ggks_debug_regions{};
    g.n = (uint64_t)n;
    g.dtype = dtype;
    g.direct = p.direct ? 1 : 0;
    g.nq = p.nq;
    if (!p.direct) {
      char* A = c->arena;
      g.segs = A + p.off_segs;
      g.states = A + p.off_states;
      g.lut = A + p.off_lut;
      g.chunks = A + p.off_chunks;
      g.rows = A + p.off_rows;
      g.totals = A + p.off_tot;
      g.gsums = A + p.off_gsum;
      g.bits = A + p.off_bits;
      g.children = A + p.off_children;
      g.cq = A + p.off_cq;
      g.red = A + p.off_red;
      g.nchunks = p.nchunks;
      g.bmax = p.bmax_max;
      g.pbits = p.pbits_max;
    }
  }
  mark(c, GGKS_NSTAGES);
  c->prof_cur = -1;
  if (rc) return rc;
  if (host_out) {
    // pin_out is disjoint from every host-written staging region
    HIP_TRY(hipMemcpyAsync(pin + pin_out, d_out, (size_t)num_k * eb, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    pin_all_free(c);
    memcpy(out, pin + pin_out, (size_t)num_k * eb);
  }
  return GGKS_OK;
}

#include "ggks_batch.hpp"
#include "ggks_serve.hpp"

// ---------------------------------------------- sort-and-choose (baseline)
// The reference's sort-and-choose (multiselectTimingFunctions.cu:41-95) on
// the device: a full radix sort of a copy of the vector (ggks_sort.hip),
// then out[i] = sorted[n - k_i].  The baseline the crossover driver races
// bucket multiselect against (analyzeMultiselect.cu:209-274).
size_t device_sort_temp_bytes(int dtype, uint64_t n);
hipError_t device_sort(int dtype, const void* in, void* out, uint64_t n, void* temp,
                       size_t temp_bytes, hipStream_t s);

template <typename T>
__global__ __launch_bounds__(256) void k_pick_sorted(const T* __restrict__ sorted, uint64_t n,
                                                     const uint32_t* __restrict__ k, uint32_t nk,
                                                     T* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nk) out[i] = sorted[n - k[i]];
}

static int sort_choose(ggks_ctx* c, int dtype, const void* d_vec, int64_t n, const uint32_t* k,
                       int64_t num_k, void* out) {
  int rc = check_common(dtype, d_vec, n, num_k);
  if (rc) return rc;
  if (num_k > 0 && (!k || !out)) return fail(GGKS_EINVAL, "null k or output array");
  for (int64_t i = 0; i < num_k; ++i)
    if (k[i] < 1 || (int64_t)k[i] > n)
      return fail(GGKS_EINVAL, "k[%lld]=%u outside [1, %lld]", (long long)i, k[i], (long long)n);
  if (num_k == 0) return GGKS_OK;
  if (!c && (rc = default_ctx(&c))) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  const size_t eb = (size_t)dtype_bytes(dtype);
  const size_t tb = device_sort_temp_bytes(dtype, (uint64_t)n);
  if (!tb) return fail(GGKS_EDEVICE, "rocPRIM sort size query failed");
  const size_t off_sorted = 0, off_temp = align_up((size_t)n * eb);
  const size_t off_k = align_up(off_temp + tb), off_out = align_up(off_k + (size_t)num_k * 4);
  if ((rc = ensure_arena(c, off_out + (size_t)num_k * eb))) return rc;
  c->table_sig = 0;  // the plan tables' place in the arena is overwritten
  const size_t pin_out = align_up((size_t)num_k * 4);
  char* pin = nullptr;
  if ((rc = pin_acquire(c, pin_out + (size_t)num_k * eb, &pin))) return rc;
  memcpy(pin, k, (size_t)num_k * 4);
  char* A = c->arena;
  HIP_TRY(hipMemcpyAsync(A + off_k, pin, (size_t)num_k * 4, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(device_sort(dtype, d_vec, A + off_sorted, (uint64_t)n, A + off_temp, tb, c->stream));
  const uint32_t nk = (uint32_t)num_k, grid = (nk + 255) / 256;
  const uint32_t* dk = reinterpret_cast<const uint32_t*>(A + off_k);
  if (eb == 8)
    hipLaunchKernelGGL(k_pick_sorted<uint64_t>, dim3(grid), dim3(256), 0, c->stream,
                       reinterpret_cast<const uint64_t*>(A + off_sorted), (uint64_t)n, dk, nk,
                       reinterpret_cast<uint64_t*>(A + off_out));
  else
    hipLaunchKernelGGL(k_pick_sorted<uint32_t>, dim3(grid), dim3(256), 0, c->stream,
                       reinterpret_cast<const uint32_t*>(A + off_sorted), (uint64_t)n, dk, nk,
                       reinterpret_cast<uint32_t*>(A + off_out));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(pin + pin_out, A + off_out, (size_t)num_k * eb, hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  pin_all_free(c);
  memcpy(out, pin + pin_out, (size_t)num_k * eb);
  return GGKS_OK;
}

// ------------------------------------------------- no-allocation entry point
// ggks_bucket_multiselect_noalloc runs the single-vector engine over a fixed
// device array instead of a context arena: no hipMalloc, hipFree, pinned
// buffer or host staging on any call.  Every GPU the library runs on holds
// its own copy of `scratch` (kNoAllocScratchBytes) until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (the array's name is a host-side stand-in).  All
// calls on one GPU share that array, so calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = GGKS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[64];

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
static int noalloc_base(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(GGKS_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(GGKS_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                  hipGetErrorString(e));
    g_noalloc_ptr[dev] = ptr;
  }
  *out = static_cast<char*>(g_noalloc_ptr[dev]);
  return GGKS_OK;
}

// Everything the plan kernel writes, by value (byte offsets into the scratch).
struct SinglePlanArgs {
  SegDesc seg;
  Child dchild;
  uint64_t n, csz;
  uint32_t G, nitems, gl, nsb, direct;
  uint64_t off_segs, off_chunks, off_items, off_sblocks, off_dchildren, off_status;
};

// Materialise a single-vector plan's tables in device memory (what the
// context path uploads from pinned memory) and clear the status block.
// Thread i writes chunk i, reduce item i and sample block i.
__global__ __launch_bounds__(256) void k_plan_single(char* __restrict__ base, SinglePlanArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 64) reinterpret_cast<uint32_t*>(base + a.off_status)[i] = 0u;
  if (a.direct) {
    if (i == 0) *reinterpret_cast<Child*>(base + a.off_dchildren) = a.dchild;
    return;
  }
  if (i == 0) *reinterpret_cast<SegDesc*>(base + a.off_segs) = a.seg;
  if (i < a.G) {
    Chunk c;
    c.seg = 0;
    c.idx = i;
    c.start = (uint64_t)i * a.csz;
    c.len = min(a.csz, a.n - c.start);
    reinterpret_cast<Chunk*>(base + a.off_chunks)[i] = c;
  }
  if (i < a.nitems) {
    ReduceItem it;
    it.seg = 0;
    it.group = (uint16_t)(i << a.gl);
    it.gi_log2 = (uint16_t)a.gl;
    reinterpret_cast<ReduceItem*>(base + a.off_items)[i] = it;
  }
  if (i < a.nsb) {
    SampleBlock sb;
    sb.seg = 0;
    sb.blk = i;
    reinterpret_cast<SampleBlock*>(base + a.off_sblocks)[i] = sb;
  }
}

static int select_noalloc(int dtype, const void* d_vec, int64_t n, const uint32_t* d_k,
                          int64_t num_k, void* d_out, const ggks_opts* opt, hipStream_t stream) {
  int rc = check_common(dtype, d_vec, n, num_k);
  if (rc) return rc;
  if (num_k > 0 && (!d_k || !d_out)) return fail(GGKS_EINVAL, "null k or output array");
  if (num_k == 0) return GGKS_OK;
  ggks_opts o = {};
  if (opt) o = *opt;
  Plan p;  // layout only: its table vectors stay empty
  plan_single(p, GGKS_KIND_BUCKET, dtype, (uint64_t)n, (uint32_t)num_k, o, false);
  const size_t off_status = align_up(p.total);
  const size_t need = off_status + 256;
  if (need > kNoAllocScratchBytes)
    return fail(GGKS_ENOMEM,
                "n=%lld %s with %lld k needs %zu bytes of scratch; the no-allocation entry "
                "point has %zu (use the context entry points for larger selects)",
                (long long)n, dtype == GGKS_F64 ? "f64" : "4-byte keys", (long long)num_k, need,
                kNoAllocScratchBytes);
  char* base = nullptr;
  if ((rc = noalloc_base(&base))) return rc;
  const SingleShape sh = single_shape(GGKS_KIND_BUCKET, dtype, (uint64_t)n, o);
  SinglePlanArgs a = {};
  a.n = (uint64_t)n;
  a.direct = p.direct ? 1u : 0u;
  if (p.direct) {
    a.dchild = single_direct_child((uint64_t)n, (uint32_t)num_k);
  } else {
    a.seg = single_segment(sh, GGKS_KIND_BUCKET, (uint64_t)n, (uint32_t)num_k);
    if (p.level2) a.seg.flags |= SEG_LEVEL2;
    a.csz = sh.csz;
    a.G = (uint32_t)sh.G;
    a.nitems = sh.nitems;
    a.gl = sh.gl;
    a.nsb = sh.ns / kThreads;
  }
  a.off_segs = p.off_segs;
  a.off_chunks = p.off_chunks;
  a.off_items = p.off_items;
  a.off_sblocks = p.off_sblocks;
  a.off_dchildren = p.off_dchildren;
  a.off_status = off_status;
  const uint32_t nthreads = std::max<uint32_t>(64, std::max(a.G, std::max(a.nitems, a.nsb)));
  hipLaunchKernelGGL(k_plan_single, dim3((nthreads + 255) / 256), dim3(256), 0, stream, base, a);
  uint32_t* status = reinterpret_cast<uint32_t*>(base + off_status);
  if (p.direct) {
    const uint32_t nk = (uint32_t)num_k;
    hipLaunchKernelGGL(k_queries, dim3((nk + 255) / 256), dim3(256), 0, stream, d_k, nk,
                       (uint64_t)n, reinterpret_cast<Query*>(base + p.off_dqueries), status);
  }
  const Exec x = {base, stream, status, nullptr};
  return launch(x, dtype, p, d_vec, d_out, d_k);
}

}  // namespace ggks

using namespace ggks;

// ================================================================== C ABI
extern "C" {

const char* ggks_last_error(void) { return g_err.c_str(); }
const char* ggks_version(void) { return "ggks-mi355x 0.1 (gfx950)"; }

int ggks_ctx_create(ggks_ctx** ctx, int device, void* stream) {
  if (!ctx) return fail(GGKS_EINVAL, "null ctx pointer");
  ggks_ctx* c = new ggks_ctx();
  int rc = ctx_init(c, device, (hipStream_t)stream);
  if (rc) { delete c; return rc; }
  *ctx = c;
  return GGKS_OK;
}

int ggks_ctx_destroy(ggks_ctx* c) {
  if (!c) return GGKS_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->arena) (void)hipFree(c->arena);
  for (PinSlot& s : c->pin) {
    if (s.buf) (void)hipHostFree(s.buf);
    if (s.ev) (void)hipEventDestroy(s.ev);
  }
  if (c->status) (void)hipFree(c->status);
  if (c->hgram) (void)hipFree(c->hgram);
  delete c;
  return GGKS_OK;
}

void* ggks_ctx_stream(ggks_ctx* c) { return c ? (void*)c->stream : nullptr; }

int ggks_ctx_debug_regions(ggks_ctx* c, ggks_debug_regions* out) {
  if (!c || !out) return fail(GGKS_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  *out = c->dbg;
  return GGKS_OK;
}

int ggks_ctx_profile_arm(ggks_ctx* c, int nslots) {
  if (!c || nslots < 0) return fail(GGKS_EINVAL, "bad profile arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (hipEvent_t e : c->prof) (void)hipEventDestroy(e);
  c->prof.clear();
  c->prof_slots = 0;
  c->prof_calls = 0;
  // timing-only events: no system-scope fence (cache writeback and
  // invalidate) when they complete, which would cost device time between the
  // kernels they separate; ggks_ctx_profile_read synchronises the stream
  // before reading them
  const size_t n = (size_t)nslots * (GGKS_NSTAGES + 1);
  for (size_t i = 0; i < n; ++i) {
    hipEvent_t e;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    c->prof.push_back(e);
  }
  c->prof_slots = nslots;
  c->prof_armed = c->prof_mask;
  return GGKS_OK;
}

int ggks_ctx_profile_read(ggks_ctx* c, float* ms, int nslots) {
  if (!c || !ms || nslots < 1 || nslots > c->prof_slots)
    return fail(GGKS_EINVAL, "bad profile arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int per = GGKS_NSTAGES + 1;
  const uint32_t mask = c->prof_armed;
  const bool kev = (mask & GGKS_PROFILE_KERNEL_EVENTS) != 0;
  auto armed = [&](int m) { return m >= 0 && m <= GGKS_NSTAGES && ((mask >> m) & 1u) != 0; };
  // with kernel events only the pairs of the kernel-timed stages exist
  auto kpair = [&](int st) { return is_kernel_stage(st) && armed(st) && armed(st + 1); };
  auto recorded = [&](int m) { return kev ? (kpair(m) || kpair(m - 1)) : armed(m); };
  int first = -1, last = -1;
  for (int m = 0; m <= GGKS_NSTAGES; ++m)
    if (recorded(m)) {
      if (first < 0) first = m;
      last = m;
    }
  for (int sl = 0; sl < nslots; ++sl) {
    hipEvent_t* e = &c->prof[(size_t)sl * per];
    float* row = ms + (size_t)sl * per;
    for (int st = 0; st <= GGKS_NSTAGES; ++st) row[st] = 0.f;
    for (int st = 0; st < GGKS_NSTAGES; ++st) {
      if (kev) {  // a kernel-timed stage: its launch's own two events
        if (kpair(st)) HIP_TRY(hipEventElapsedTime(&row[st], e[st], e[st + 1]));
        continue;
      }
      if (!recorded(st)) continue;
      int nx = st + 1;
      while (nx <= GGKS_NSTAGES && !recorded(nx)) ++nx;
      if (nx <= GGKS_NSTAGES) HIP_TRY(hipEventElapsedTime(&row[st], e[st], e[nx]));
    }
    if (first >= 0 && last > first)
      HIP_TRY(hipEventElapsedTime(&row[GGKS_NSTAGES], e[first], e[last]));
  }
  return GGKS_OK;
}

int ggks_ctx_profile_marks(ggks_ctx* c, uint32_t mask) {
  if (!c) return fail(GGKS_EINVAL, "null ctx");
  std::lock_guard<std::mutex> lk(c->mu);
  c->prof_mask = mask & (((1u << (GGKS_NSTAGES + 1)) - 1u) | GGKS_PROFILE_KERNEL_EVENTS);
  return GGKS_OK;
}

size_t ggks_scratch_bytes(int kind, int dtype, int64_t n, int64_t num_k) {
  if (dtype < 0 || dtype > 3 || n < 1 || num_k < 0) return 0;
  Plan p;
  ggks_opts o = {};
  plan_single(p, kind, dtype, (uint64_t)n, (uint32_t)std::max<int64_t>(num_k, 1), o);
  return p.total + align_up((size_t)num_k * (dtype_bytes(dtype) + 4));
}

int ggks_ctx_reserve(ggks_ctx* c, int kind, int dtype, int64_t n, int64_t num_k) {
  if (!c) return fail(GGKS_EINVAL, "null ctx");
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  int rc = ensure_arena(c, ggks_scratch_bytes(kind, dtype, n, num_k));
  if (rc) return rc;
  Plan p;
  ggks_opts o = {};
  plan_single(p, kind, dtype, (uint64_t)n, (uint32_t)std::max<int64_t>(num_k, 1), o);
  // every staging slot holds the tables and a 1024-entry k list
  for (PinSlot& s : c->pin)
    if ((rc = pin_ready(s, p.table_bytes() + 4096))) return rc;
  return GGKS_OK;
}

int ggks_ctx_check(ggks_ctx* c) {
  if (!c) return fail(GGKS_EINVAL, "null ctx");
  uint32_t st = 0;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(&st, c->status, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (st) {
    HIP_TRY(hipMemsetAsync(c->status, 0, 4, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (st & STATUS_BAD_K) return fail(GGKS_EINVAL, "a device k list held k outside [1, n]");
    return fail(GGKS_EDEVICE, "internal status 0x%x", st);
  }
  return GGKS_OK;
}

int ggks_bucket_multiselect(ggks_ctx* c, int dtype, const void* d_vec, int64_t n,
                            const uint32_t* h_k, int64_t num_k, void* h_out,
                            const ggks_opts* opt) {
  return select_one(c, GGKS_KIND_BUCKET, dtype, d_vec, n, h_k, num_k, KSource::HOST, h_out, true,
                    opt);
}
int ggks_sort_choose(ggks_ctx* c, int dtype, const void* d_vec, int64_t n, const uint32_t* h_k,
                     int64_t num_k, void* h_out) {
  return sort_choose(c, dtype, d_vec, n, h_k, num_k, h_out);
}
int ggks_radix_multiselect(ggks_ctx* c, int dtype, const void* d_vec, int64_t n,
                           const uint32_t* h_k, int64_t num_k, void* h_out,
                           const ggks_opts* opt) {
  return select_one(c, GGKS_KIND_RADIX, dtype, d_vec, n, h_k, num_k, KSource::HOST, h_out, true,
                    opt);
}
int ggks_topk(ggks_ctx* c, int dtype, const void* d_vec, int64_t n, int64_t K, void* h_out,
              const ggks_opts* opt) {
  return select_one(c, GGKS_KIND_TOPK, dtype, d_vec, n, nullptr, K, KSource::TOPK, h_out, true,
                    opt);
}
int ggks_bucket_multiselect_async(ggks_ctx* c, int dtype, const void* d_vec, int64_t n,
                                  const uint32_t* d_k, int64_t num_k, void* d_out,
                                  const ggks_opts* opt) {
  return select_one(c, GGKS_KIND_BUCKET, dtype, d_vec, n, d_k, num_k, KSource::DEVICE, d_out,
                    false, opt);
}
int ggks_radix_multiselect_async(ggks_ctx* c, int dtype, const void* d_vec, int64_t n,
                                 const uint32_t* d_k, int64_t num_k, void* d_out,
                                 const ggks_opts* opt) {
  return select_one(c, GGKS_KIND_RADIX, dtype, d_vec, n, d_k, num_k, KSource::DEVICE, d_out,
                    false, opt);
}
int ggks_topk_async(ggks_ctx* c, int dtype, const void* d_vec, int64_t n, int64_t K, void* d_out,
                    const ggks_opts* opt) {
  return select_one(c, GGKS_KIND_TOPK, dtype, d_vec, n, nullptr, K, KSource::TOPK, d_out, false,
                    opt);
}

#define GGKS_TYPED(NAME, T, DT)                                                          \
  int NAME(const T* d_vec, int64_t n, const uint32_t* h_k, int32_t num_k, T* h_out,      \
           const ggks_opts* opt) {                                                       \
    return ggks_bucket_multiselect(nullptr, DT, d_vec, n, h_k, num_k, h_out, opt);       \
  }
GGKS_TYPED(ggks_bucket_multiselect_f32, float, GGKS_F32)
GGKS_TYPED(ggks_bucket_multiselect_f64, double, GGKS_F64)
GGKS_TYPED(ggks_bucket_multiselect_u32, uint32_t, GGKS_U32)
GGKS_TYPED(ggks_bucket_multiselect_i32, int32_t, GGKS_I32)
#undef GGKS_TYPED

#define GGKS_TYPED_RADIX(NAME, T, DT)                                                    \
  int NAME(const T* d_vec, int64_t n, const uint32_t* h_k, int32_t num_k, T* h_out,      \
           const ggks_opts* opt) {                                                       \
    return ggks_radix_multiselect(nullptr, DT, d_vec, n, h_k, num_k, h_out, opt);        \
  }
GGKS_TYPED_RADIX(ggks_radix_multiselect_f32, float, GGKS_F32)
GGKS_TYPED_RADIX(ggks_radix_multiselect_f64, double, GGKS_F64)
GGKS_TYPED_RADIX(ggks_radix_multiselect_u32, uint32_t, GGKS_U32)
GGKS_TYPED_RADIX(ggks_radix_multiselect_i32, int32_t, GGKS_I32)
#undef GGKS_TYPED_RADIX

#define GGKS_TYPED_TOPK(NAME, T, DT)                                                      \
  int NAME(const T* d_vec, int64_t n, int32_t K, T* h_out, const ggks_opts* opt) {       \
    return ggks_topk(nullptr, DT, d_vec, n, K, h_out, opt);                               \
  }
GGKS_TYPED_TOPK(ggks_topk_f32, float, GGKS_F32)
GGKS_TYPED_TOPK(ggks_topk_f64, double, GGKS_F64)
GGKS_TYPED_TOPK(ggks_topk_u32, uint32_t, GGKS_U32)
GGKS_TYPED_TOPK(ggks_topk_i32, int32_t, GGKS_I32)
#undef GGKS_TYPED_TOPK

int ggks_bucket_multiselect_noalloc(int dtype, const void* d_vec, int64_t n, const uint32_t* d_k,
                                    int64_t num_k, void* d_out, const ggks_opts* opt,
                                    void* stream) {
  return select_noalloc(dtype, d_vec, n, d_k, num_k, d_out, opt, (hipStream_t)stream);
}


// ---------------------------------------------------------- plumbing
int ggks_dev_alloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(GGKS_EINVAL, "null pointer");
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(GGKS_ENOMEM, "hipMalloc(%zu) failed", bytes);
  }
  return GGKS_OK;
}
int ggks_dev_free(void* ptr) {
  HIP_TRY(hipFree(ptr));
  return GGKS_OK;
}
int ggks_h2d(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GGKS_OK;
}
int ggks_d2h(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GGKS_OK;
}
int ggks_d2d(void* dst, const void* src, size_t bytes) {
  // 16-byte aligned whole-vector copies run the library's own streaming copy
  // kernel (one read and one write of every byte: a known-size reference for
  // memory-traffic counters); anything else goes through the runtime.
  if (((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0 && bytes > 0) {
    const uint64_t nv = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((nv + kThreads - 1) / kThreads, 4096);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(kThreads), 0, nullptr,
                       reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), nv);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return GGKS_OK;
  }
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return GGKS_OK;
}
int ggks_sync(void) {
  HIP_TRY(hipDeviceSynchronize());
  return GGKS_OK;
}
int ggks_device_count(int* count) {
  HIP_TRY(hipGetDeviceCount(count));
  return GGKS_OK;
}
int ggks_set_device(int device) {
  HIP_TRY(hipSetDevice(device));
  return GGKS_OK;
}
int ggks_dev_fill(void* ptr, int value, size_t bytes) {
  HIP_TRY(hipMemset(ptr, value, bytes));
  return GGKS_OK;
}


// ---------------------------------------------------------------- batches
int ggks_batch_create(ggks_batch** batch, int device, int dtype, const ggks_segment* segs,
                      int64_t num_segments, const uint32_t* h_k, int64_t total_k,
                      const ggks_opts* opt, void* d_scratch, size_t scratch_bytes,
                      void* stream) {
  return batch_create(batch, device, dtype, segs, num_segments, h_k, total_k, opt, d_scratch,
                      scratch_bytes, stream);
}

size_t ggks_batch_scratch_bytes(int dtype, const ggks_segment* segs, int64_t num_segments,
                                const uint32_t* h_k, int64_t total_k, const ggks_opts* opt) {
  std::vector<BatchSeg> in;
  uint64_t elements = 0;
  if (batch_segments(dtype, segs, num_segments, h_k, total_k, in, elements)) return 0;
  ggks_opts o = {};
  if (opt) o = *opt;
  Plan p;
  if (!plan_batch(p, dtype, in, o)) {
    fail(GGKS_EINVAL, "batch too large for one plan (32-bit table offsets); split it");
    return 0;
  }
  return batch_arena_bytes(p);
}

int ggks_batch_run(ggks_batch* b, const void* d_vec, void* d_out, int flags) {
  return batch_run(b, d_vec, d_out, flags);
}

int ggks_batch_sync(ggks_batch* b) { return batch_sync(b); }

void* ggks_batch_stream(ggks_batch* b) { return b ? (void*)b->stream : nullptr; }

int ggks_batch_info(ggks_batch* b, ggks_batch_stats* info) {
  if (!b || !info) return fail(GGKS_EINVAL, "null argument");
  info->elements = b->elements;
  info->queries = b->plan.nq + b->plan.direct_queries.size();
  info->scratch_bytes = b->arena_bytes;
  info->engine_segments = b->engine;
  info->direct_segments = b->direct;
  info->chunks = (uint32_t)b->plan.chunks.size();
  info->graph = b->exec ? 1 : 0;
  return GGKS_OK;
}

int ggks_batch_destroy(ggks_batch* b) { return batch_destroy(b); }

int ggks_tuning_set(int dtype, int log2_len, int buckets_log2, int chunk_log2, int direct) {
  if (dtype < 0 || dtype > 3 || log2_len < 0 || log2_len >= kBands)
    return fail(GGKS_EINVAL, "bad tuning band (dtype %d, log2 length %d)", dtype, log2_len);
  if (buckets_log2 < 8 || buckets_log2 > 15 || chunk_log2 < 10 || chunk_log2 > 26)
    return fail(GGKS_EINVAL, "bad tuning values (buckets 2^%d, chunk 2^%d)", buckets_log2,
                chunk_log2);
  BandTuning t = {};
  t.buckets_log2 = (uint8_t)buckets_log2;
  t.chunk_log2 = (uint8_t)chunk_log2;
  t.direct = direct ? 1 : 0;
  tuning_table().t[dtype][log2_len] = t;
  return GGKS_OK;
}

int ggks_tuning_get(int dtype, int log2_len, int* buckets_log2, int* chunk_log2, int* direct) {
  if (dtype < 0 || dtype > 3 || log2_len < 0 || log2_len >= kBands)
    return fail(GGKS_EINVAL, "bad tuning band (dtype %d, log2 length %d)", dtype, log2_len);
  const BandTuning t = tuning_table().t[dtype][log2_len];
  if (buckets_log2) *buckets_log2 = t.buckets_log2;
  if (chunk_log2) *chunk_log2 = t.chunk_log2;
  if (direct) *direct = t.direct;
  return GGKS_OK;
}

int ggks_tuning_reset(void) {
  tuning_table() = TuningTable();
  single_tuning_table() = SingleTuningTable();
  return GGKS_OK;
}

int ggks_tuning_single_set(int dtype, int log2_len, int buckets_log2, int chunks) {
  if (dtype < 0 || dtype > 3 || log2_len < 0 || log2_len >= kSingleBands)
    return fail(GGKS_EINVAL, "bad tuning band (dtype %d, log2 length %d)", dtype, log2_len);
  if (buckets_log2 < 8 || buckets_log2 > 15 || chunks < 1 || chunks > (int)kMaxChunks)
    return fail(GGKS_EINVAL, "bad single-vector tuning (buckets 2^%d, %d chunks)", buckets_log2,
                chunks);
  single_tuning_table().t[dtype][log2_len] = {(uint32_t)buckets_log2, (uint32_t)chunks};
  return GGKS_OK;
}

int ggks_tuning_single_get(int dtype, int log2_len, int* buckets_log2, int* chunks) {
  if (dtype < 0 || dtype > 3 || log2_len < 0 || log2_len >= kSingleBands)
    return fail(GGKS_EINVAL, "bad tuning band (dtype %d, log2 length %d)", dtype, log2_len);
  const Tuning t = single_tuning_table().t[dtype][log2_len];
  if (buckets_log2) *buckets_log2 = (int)t.buckets_log2;
  if (chunks) *chunks = (int)t.chunks;
  return GGKS_OK;
}

// ---------------------------------------------------------- device inputs
int ggks_generate(void* d_ptr, int dtype, int family, uint64_t seed, uint64_t start, uint64_t n,
                  int64_t lo, int64_t hi) {
  if (!d_ptr && n) return fail(GGKS_EINVAL, "null pointer");
  if (dtype < 0 || dtype > 3 || family < 0 || family >= GEN_FAMILIES)
    return fail(GGKS_EINVAL, "bad dtype %d or family %d", dtype, family);
  const bool integer = dtype == GGKS_U32 || dtype == GGKS_I32;
  if (integer && family != GEN_UNIFORM && family != GEN_LEVELS && family != GEN_SORTED &&
      family != GEN_ONES && family != GEN_ONES_TWOS)
    return fail(GGKS_EINVAL, "family %d is for float dtypes", family);
  if (family == GEN_SORTED) {
    if (hi < 1 || start > (uint64_t)hi || n > (uint64_t)hi - start)
      return fail(GGKS_EINVAL, "sorted: positions [%llu, +%llu) outside the length %lld",
                  (unsigned long long)start, (unsigned long long)n, (long long)hi);
  } else if (family == GEN_LEVELS && !integer && (hi < lo || lo < 0 || hi > 0xFFFFFFFFll)) {
    return fail(GGKS_EINVAL, "bad levels range [%lld, %lld]", (long long)lo, (long long)hi);
  } else if (integer && (family == GEN_UNIFORM || family == GEN_LEVELS) &&
             (hi < lo || (dtype == GGKS_U32 && (lo < 0 || hi > 0xFFFFFFFFll)) ||
              (dtype == GGKS_I32 && (lo < INT32_MIN || hi > INT32_MAX)))) {
    return fail(GGKS_EINVAL, "bad integer range [%lld, %lld]", (long long)lo, (long long)hi);
  }
  if (n == 0) return GGKS_OK;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_generate, dim3(grid), dim3(256), 0, nullptr, d_ptr, dtype, family, seed,
                     start, n, lo, hi);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return GGKS_OK;
}

// ---------------------------------------------------------------- serving
int ggks_server_create(ggks_server** out, int device, const void* const* resident,
                       const uint64_t* resident_len, size_t scratch_budget, int max_inflight,
                       int max_batch) {
  if (!out || !resident || !resident_len) return fail(GGKS_EINVAL, "null argument");
  if (scratch_budget < (1u << 20) || max_inflight < 1 || max_inflight > 16 || max_batch < 1)
    return fail(GGKS_EINVAL, "bad server limits");
  HIP_TRY(hipSetDevice(device));
  ggks_server* s = new ggks_server();
  s->device = device;
  for (int d = 0; d < 4; ++d) {
    s->resident[d] = static_cast<const char*>(resident[d]);
    s->resident_len[d] = resident[d] ? resident_len[d] : 0;
  }
  s->budget = align_up(scratch_budget);
  s->max_batch = (uint32_t)max_batch;
  if (hipMalloc(&s->pool, s->budget) != hipSuccess) {
    (void)hipGetLastError();
    s->pool = nullptr;
    serve_destroy(s);
    return fail(GGKS_ENOMEM, "cannot allocate the %zu-byte scratch budget", scratch_budget);
  }
  s->alloc.reset(s->budget);
  s->st.scratch_budget = s->budget;
  s->slots.resize((size_t)max_inflight);
  for (ServeSlot& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      const hipError_t e = hipGetLastError();
      serve_destroy(s);
      return fail(GGKS_EDEVICE, "cannot create server streams: %s", hipGetErrorString(e));
    }
  }
  *out = s;
  return GGKS_OK;
}

int ggks_server_submit(ggks_server* s, const ggks_request* r) {
  if (!s) return fail(GGKS_EINVAL, "null server");
  HIP_TRY(hipSetDevice(s->device));
  int rc = serve_submit(s, r, now_s());
  if (rc) return rc;
  size_t nc = 0;
  return serve_step(s, nc);
}

int ggks_server_poll(ggks_server* s, uint64_t* tags, double* latency_s, int max_done) {
  if (!s || max_done < 0 || (max_done > 0 && (!tags || !latency_s)))
    return fail(GGKS_EINVAL, "bad poll arguments");
  HIP_TRY(hipSetDevice(s->device));
  size_t nc = 0;
  int rc = serve_step(s, nc);
  if (rc) return rc;
  int k = 0;
  while (k < max_done && !s->finished.empty()) {
    tags[k] = s->finished.front().first;
    latency_s[k] = s->finished.front().second;
    s->finished.pop_front();
    ++k;
  }
  return k;
}

int ggks_server_drain(ggks_server* s) {
  if (!s) return fail(GGKS_EINVAL, "null server");
  HIP_TRY(hipSetDevice(s->device));
  size_t nc = 0;
  while (!serve_idle(s)) {
    int rc = serve_step(s, nc);
    if (rc) return rc;
  }
  return GGKS_OK;
}

int ggks_server_run_trace(ggks_server* s, const ggks_request* reqs, const double* arrival_s,
                          int64_t n, double* latency_s) {
  if (!s || n < 0 || (n > 0 && (!reqs || !arrival_s || !latency_s)))
    return fail(GGKS_EINVAL, "bad trace arguments");
  HIP_TRY(hipSetDevice(s->device));
  return serve_run_trace(s, reqs, arrival_s, n, latency_s);
}

int ggks_server_stats_get(ggks_server* s, ggks_server_stats* out) {
  if (!s || !out) return fail(GGKS_EINVAL, "null argument");
  *out = s->st;
  out->pending = s->pending.size();
  out->scratch_in_use = s->alloc.used();
  return GGKS_OK;
}

int ggks_server_destroy(ggks_server* s) { return serve_destroy(s); }

}  // extern "C"

// ggks_batch.hpp -- batched selection over a segment table.
//
// A batch is planned once on the host (plan_batch): its tables go to the
// device at creation and stay there, its scratch is allocated once, and every
// run enqueues the same launch sequence -- one launch per level kernel for
// all engine segments together, plus one finisher launch for every small
// segment -- either directly or as a replay of a captured hipGraph.
// Included by ggks_select.hip inside namespace ggks (uses Exec, launch,
// fail, HIP_TRY from there).
#pragma once

}  // namespace ggks

struct ggks_batch {
  int device = 0;
  int dtype = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipStream_t side = nullptr;  // direct children finish here, beside the engine pipeline
  hipEvent_t fork = nullptr, join = nullptr;
  ggks::Plan plan;
  char* arena = nullptr;  // plan tables + scratch + status word
  bool own_arena = false;
  size_t arena_bytes = 0;
  uint32_t* status = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* g_src = nullptr;  // pointers the captured graph was built for
  void* g_out = nullptr;
  uint64_t elements = 0;        // sum of segment lengths with queries
  uint32_t engine = 0, direct = 0;
  std::mutex mu;
};

namespace ggks {

// Validate a segment table and turn it into planner segments.
static int batch_segments(int dtype, const ggks_segment* segs, int64_t nseg, const uint32_t* h_k,
                          int64_t total_k, std::vector<BatchSeg>& out, uint64_t& elements) {
  if (dtype < 0 || dtype > 3) return fail(GGKS_EINVAL, "bad dtype %d", dtype);
  if (nseg < 0 || (nseg > 0 && !segs)) return fail(GGKS_EINVAL, "bad segment table");
  if (total_k < 0 || (total_k > 0 && !h_k)) return fail(GGKS_EINVAL, "bad k list");
  out.clear();
  out.reserve((size_t)nseg);
  elements = 0;
  for (int64_t i = 0; i < nseg; ++i) {
    const ggks_segment& s = segs[i];
    if (s.length < 1 || s.length > 0xFFFFFFFFull)
      return fail(GGKS_EINVAL, "segment %lld: length %llu outside [1, 2^32)", (long long)i,
                  (unsigned long long)s.length);
    if (s.kind != GGKS_KIND_BUCKET && s.kind != GGKS_KIND_RADIX && s.kind != GGKS_KIND_TOPK)
      return fail(GGKS_EINVAL, "segment %lld: bad kind %u", (long long)i, s.kind);
    if (s.out_offset + s.num_k > 0x100000000ull)
      return fail(GGKS_EINVAL, "segment %lld: outputs beyond 2^32 slots", (long long)i);
    BatchSeg g = {};
    g.offset = s.offset;
    g.length = s.length;
    g.num_k = s.num_k;
    g.kind = s.kind;
    g.out_offset = s.out_offset;
    if (s.kind == GGKS_KIND_TOPK) {
      if (s.num_k > s.length)
        return fail(GGKS_EINVAL, "segment %lld: K=%u > length", (long long)i, s.num_k);
    } else if (s.num_k > 0) {
      if (s.k_offset + s.num_k > (uint64_t)total_k)
        return fail(GGKS_EINVAL, "segment %lld: k range beyond the k list", (long long)i);
      g.k = h_k + s.k_offset;
      for (uint32_t j = 0; j < s.num_k; ++j)
        if (g.k[j] < 1 || g.k[j] > s.length)
          return fail(GGKS_EINVAL, "segment %lld: k=%u outside [1, %llu]", (long long)i, g.k[j],
                      (unsigned long long)s.length);
    }
    if (s.num_k > 0) elements += s.length;
    out.push_back(g);
  }
  return GGKS_OK;
}

template <typename KT, int C>
static void batch_attrs_typed(const Plan& p) {
  level_attrs<KT, C>(p);
}

static void batch_attrs(int dtype, const Plan& p) {
  switch (dtype) {
    case GGKS_F32: batch_attrs_typed<uint32_t, CONV_F>(p); break;
    case GGKS_F64: batch_attrs_typed<uint64_t, CONV_F>(p); break;
    case GGKS_U32: batch_attrs_typed<uint32_t, CONV_U>(p); break;
    default: batch_attrs_typed<uint32_t, CONV_I>(p); break;
  }
}

template <typename KT, int C>
static int batch_enqueue_typed(const Exec& x, const Plan& p, const void* src, void* out,
                               const uint32_t* d_k) {
  return enqueue_level<KT, C>(x, p, (const KT*)src, (KT*)out, d_k);
}

static int batch_enqueue(ggks_batch* b, const void* src, void* out) {
  const Exec x = {b->arena, b->stream, b->status, nullptr, b->side, b->fork, b->join};
  const uint32_t* d_k = reinterpret_cast<const uint32_t*>(b->arena + b->plan.off_kvals);
  switch (b->dtype) {
    case GGKS_F32: return batch_enqueue_typed<uint32_t, CONV_F>(x, b->plan, src, out, d_k);
    case GGKS_F64: return batch_enqueue_typed<uint64_t, CONV_F>(x, b->plan, src, out, d_k);
    case GGKS_U32: return batch_enqueue_typed<uint32_t, CONV_U>(x, b->plan, src, out, d_k);
    default: return batch_enqueue_typed<uint32_t, CONV_I>(x, b->plan, src, out, d_k);
  }
}

static void batch_drop_graph(ggks_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->g_src = nullptr;
  b->g_out = nullptr;
}

// Bytes of device memory a batch's arena takes (tables, scratch, status).
static size_t batch_arena_bytes(const Plan& p) { return align_up(p.total) + 256; }

// Plan, place and upload.  arena == null: the batch allocates its own.
static int batch_create(ggks_batch** out, int device, int dtype, const ggks_segment* segs,
                        int64_t nseg, const uint32_t* h_k, int64_t total_k, const ggks_opts* opt,
                        void* arena, size_t arena_bytes, void* stream) {
  if (!out) return fail(GGKS_EINVAL, "null batch pointer");
  std::vector<BatchSeg> in;
  uint64_t elements = 0;
  int rc = batch_segments(dtype, segs, nseg, h_k, total_k, in, elements);
  if (rc) return rc;
  ggks_opts o = {};
  if (opt) o = *opt;
  {
    Plan probe;
    if (!plan_batch(probe, dtype, in, o))
      return fail(GGKS_EINVAL, "batch too large for one plan (32-bit table offsets); split it");
  }
  HIP_TRY(hipSetDevice(device));
  ggks_batch* b = new ggks_batch();
  b->device = device;
  b->dtype = dtype;
  b->elements = elements;
  plan_batch(b->plan, dtype, in, o);
  b->engine = (uint32_t)b->plan.segs.size();
  b->direct = (uint32_t)b->plan.direct_children.size();
  const size_t need = batch_arena_bytes(b->plan);
  auto bail = [&](int code) {
    if (b->own_arena && b->arena) (void)hipFree(b->arena);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    delete b;
    return code;
  };
  if (arena) {
    if (arena_bytes < need)
      return bail(fail(GGKS_ENOMEM, "batch needs %zu bytes of scratch, %zu given", need,
                       arena_bytes));
    b->arena = static_cast<char*>(arena);
  } else {
    if (hipMalloc(&b->arena, need) != hipSuccess) {
      (void)hipGetLastError();
      b->arena = nullptr;
      return bail(fail(GGKS_ENOMEM, "cannot allocate %zu bytes of batch scratch", need));
    }
    b->own_arena = true;
  }
  b->arena_bytes = need;
  b->status = reinterpret_cast<uint32_t*>(b->arena + align_up(b->plan.total));
  if (stream) {
    b->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      b->stream = nullptr;
      return bail(fail(GGKS_EDEVICE, "cannot create a stream"));
    }
    b->own_stream = true;
  }
  if (b->engine > 0 && b->direct > 0) {  // default priority, like every stream here
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      return bail(fail(GGKS_EDEVICE, "cannot create the batch's side stream"));
    }
  }
  const size_t tb = b->plan.table_bytes();
  std::vector<char> tab(tb + 256, 0);
  b->plan.write_tables(tab.data());
  if (hipMemcpyAsync(b->arena + b->plan.off_tables, tab.data(), tb, hipMemcpyHostToDevice,
                     b->stream) != hipSuccess ||
      hipMemsetAsync(b->status, 0, 256, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    const hipError_t e = hipGetLastError();
    return bail(fail(GGKS_EDEVICE, "batch table upload failed: %s", hipGetErrorString(e)));
  }
  batch_attrs(dtype, b->plan);
  *out = b;
  return GGKS_OK;
}

static int batch_run(ggks_batch* b, const void* d_vec, void* d_out, int flags) {
  if (!b) return fail(GGKS_EINVAL, "null batch");
  if ((b->engine || b->direct) && (!d_vec || !d_out))
    return fail(GGKS_EINVAL, "null vector or output");
  std::lock_guard<std::mutex> lk(b->mu);
  HIP_TRY(hipSetDevice(b->device));
  if (!(flags & GGKS_BATCH_GRAPH)) return batch_enqueue(b, d_vec, d_out);
  if (b->exec && (b->g_src != d_vec || b->g_out != d_out)) batch_drop_graph(b);
  if (!b->exec) {
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, d_vec, d_out);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return fail(GGKS_EDEVICE, "graph capture failed: %s", hipGetErrorString(ec));
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
      const hipError_t e = hipGetLastError();
      batch_drop_graph(b);
      return fail(GGKS_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    b->g_src = d_vec;
    b->g_out = d_out;
  }
  HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  return GGKS_OK;
}

static int batch_sync(ggks_batch* b) {
  if (!b) return fail(GGKS_EINVAL, "null batch");
  std::lock_guard<std::mutex> lk(b->mu);
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  uint32_t st = 0;
  HIP_TRY(hipMemcpy(&st, b->status, 4, hipMemcpyDeviceToHost));
  if (st) {
    const uint32_t zero = 0;
    HIP_TRY(hipMemcpy(b->status, &zero, 4, hipMemcpyHostToDevice));
    return fail(GGKS_EINVAL, "batch kernels reported status 0x%x", st);
  }
  return GGKS_OK;
}

static int batch_destroy(ggks_batch* b) {
  if (!b) return GGKS_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->own_arena && b->arena) (void)hipFree(b->arena);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  delete b;
  return GGKS_OK;
}

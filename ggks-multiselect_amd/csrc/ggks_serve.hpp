// ggks_serve.hpp -- the serving loop: continuous batching of select requests
// over vectors resident in HBM.
//
// Requests (a multiselect k list or a top-K over a vector in one of the
// per-dtype resident buffers) wait in a FIFO.  Whenever an in-flight slot is
// free, the scheduler takes the oldest request and every later pending
// request of the same dtype (up to max_batch), plans them as one batch
// (plan_batch), and admits the batch if its scratch (tables, bucket rows,
// reduced buffer, outputs) fits in the free part of the scratch budget --
// halving the group until it does.  The batch's tables go up from the slot's
// pinned buffer, the level kernels run on the slot's stream, the outputs come
// back to the same pinned buffer and an event marks completion.  Up to
// max_inflight batches run at once on separate streams, so requests that
// arrive while others are on the GPU are admitted as slots and memory free
// up.  Polling harvests finished slots (outputs copied to each request's
// host buffer, latency = completion seen - submission) and schedules again.
// One host thread drives a server.  Included by ggks_select.hip inside
// namespace ggks.
#pragma once

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct ServeSlot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  size_t off = 0, bytes = 0;  // scratch sub-allocation
  char* pinned = nullptr;
  size_t pinned_cap = 0;
  size_t pin_stat = 0;  // the batch's status block inside the pinned buffer
  size_t pin_out = 0;   // outputs inside the pinned buffer
  std::vector<size_t> ids;
  std::vector<uint64_t> out_at;  // per request: first output slot in the batch
  int dtype = 0;
};

}  // namespace ggks

struct ggks_server {
  int device = 0;
  const char* resident[4] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t resident_len[4] = {0, 0, 0, 0};  // elements
  char* pool = nullptr;
  size_t budget = 0;
  ggks::RangeAlloc alloc;
  std::vector<ggks::ServeSlot> slots;
  std::deque<size_t> pending;
  std::vector<ggks::ServeReq> reqs;  // by internal id (slots are recycled)
  std::vector<size_t> free_ids;      // request slots free for reuse
  std::deque<std::pair<uint64_t, double>> finished;  // (tag, latency; -1: failed)
  uint32_t max_batch = 256;
  ggks_server_stats st = {};
  double* trace_latency = nullptr;  // run_trace: latency per request id
  size_t last_id = 0;               // slot of the last submitted request
};

namespace ggks {

static int serve_validate(const ggks_server* s, const ggks_request* r) {
  if (!r) return fail(GGKS_EINVAL, "null request");
  if (r->dtype < 0 || r->dtype > 3 || !s->resident[r->dtype])
    return fail(GGKS_EINVAL, "dtype %d has no resident buffer", r->dtype);
  if (r->kind != GGKS_KIND_BUCKET && r->kind != GGKS_KIND_RADIX && r->kind != GGKS_KIND_TOPK)
    return fail(GGKS_EINVAL, "bad kind %d", r->kind);
  if (r->length < 1 || r->length > 0xFFFFFFFFull)
    return fail(GGKS_EINVAL, "length %llu outside [1, 2^32)", (unsigned long long)r->length);
  if (r->offset > s->resident_len[r->dtype] || r->length > s->resident_len[r->dtype] - r->offset)
    return fail(GGKS_EINVAL, "vector [%llu, +%llu) outside the resident buffer",
                (unsigned long long)r->offset, (unsigned long long)r->length);
  if (r->num_k < 1 || !r->h_out) return fail(GGKS_EINVAL, "no outputs requested");
  if (r->kind == GGKS_KIND_TOPK) {
    if (r->num_k > r->length) return fail(GGKS_EINVAL, "K=%u > length", r->num_k);
  } else {
    if (!r->h_k) return fail(GGKS_EINVAL, "null k list");
    for (uint32_t j = 0; j < r->num_k; ++j)
      if (r->h_k[j] < 1 || r->h_k[j] > r->length)
        return fail(GGKS_EINVAL, "k=%u outside [1, %llu]", r->h_k[j],
                    (unsigned long long)r->length);
  }
  return GGKS_OK;
}

template <typename KT, int C>
static int serve_enqueue_typed(const Exec& x, const Plan& p, const void* src, void* out,
                               const uint32_t* d_k) {
  level_attrs<KT, C>(p);
  return enqueue_level<KT, C>(x, p, (const KT*)src, (KT*)out, d_k);
}

// Launch one batch in `slot` for the requests `ids` (same dtype).
static int serve_launch(ggks_server* s, ServeSlot& sl, const std::vector<size_t>& ids,
                        const Plan& p, size_t foot, size_t off) {
  const int dtype = s->reqs[ids[0]].r.dtype;
  const size_t eb = (size_t)dtype_bytes(dtype);
  uint64_t outs = 0;
  sl.out_at.clear();
  for (size_t id : ids) {
    sl.out_at.push_back(outs);
    outs += s->reqs[id].r.num_k;
  }
  const size_t tb = p.table_bytes();
  // pinned layout: tables | status block (256 B) | outputs; the status block
  // and the outputs come back in one copy (they are adjacent on the device)
  const size_t pin_stat = align_up(tb);
  const size_t pin_out = pin_stat + 256;
  const size_t pin_need = pin_out + (size_t)outs * eb;
  if (pin_need > sl.pinned_cap) {
    if (sl.pinned) HIP_TRY(hipHostFree(sl.pinned));
    sl.pinned = nullptr;
    sl.pinned_cap = 0;
    const size_t cap = std::max<size_t>(pin_need + pin_need / 4, 1 << 20);
    if (hipHostMalloc(&sl.pinned, cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      sl.pinned = nullptr;
      return fail(GGKS_ENOMEM, "cannot allocate %zu bytes of pinned memory", cap);
    }
    sl.pinned_cap = cap;
  }
  p.write_tables(sl.pinned);
  char* arena = s->pool + off;
  uint32_t* status = reinterpret_cast<uint32_t*>(arena + align_up(p.total));
  void* d_out = arena + align_up(p.total) + 256;
  HIP_TRY(hipMemcpyAsync(arena + p.off_tables, sl.pinned, tb, hipMemcpyHostToDevice, sl.stream));
  HIP_TRY(hipMemsetAsync(status, 0, 4, sl.stream));
  const Exec x = {arena, sl.stream, status, nullptr};
  const uint32_t* d_k = reinterpret_cast<const uint32_t*>(arena + p.off_kvals);
  const void* src = s->resident[dtype];
  int rc;
  switch (dtype) {
    case GGKS_F32: rc = serve_enqueue_typed<uint32_t, CONV_F>(x, p, src, d_out, d_k); break;
    case GGKS_F64: rc = serve_enqueue_typed<uint64_t, CONV_F>(x, p, src, d_out, d_k); break;
    case GGKS_U32: rc = serve_enqueue_typed<uint32_t, CONV_U>(x, p, src, d_out, d_k); break;
    default: rc = serve_enqueue_typed<uint32_t, CONV_I>(x, p, src, d_out, d_k); break;
  }
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(sl.pinned + pin_stat, status, 256 + (size_t)outs * eb,
                         hipMemcpyDeviceToHost, sl.stream));
  HIP_TRY(hipEventRecord(sl.done, sl.stream));
  sl.busy = true;
  sl.off = off;
  sl.bytes = foot;
  sl.pin_stat = pin_stat;
  sl.pin_out = pin_out;
  sl.ids = ids;
  sl.dtype = dtype;
  s->st.batches++;
  s->st.peak_scratch_bytes = std::max<uint64_t>(s->st.peak_scratch_bytes, s->alloc.used());
  return GGKS_OK;
}

// Form and launch batches while slots and scratch allow (admission:
// serve_form_group / serve_admit).
static int serve_schedule(ggks_server* s) {
  for (ServeSlot& sl : s->slots) {
    if (sl.busy) continue;
    if (s->pending.empty()) return GGKS_OK;
    const std::vector<size_t> group = serve_form_group(s->reqs, s->pending, s->max_batch);
    Plan p;
    size_t foot = 0, off = 0;
    const size_t take = serve_admit(s->alloc, s->reqs, group, p, foot, off);
    if (take == 0) return GGKS_OK;  // wait for memory to free up
    std::vector<size_t> ids(group.begin(), group.begin() + take);
    int rc = serve_launch(s, sl, ids, p, foot, off);
    if (rc) {
      s->alloc.release(off, foot);
      return rc;
    }
    serve_remove_launched(s->pending, ids);
  }
  return GGKS_OK;
}

// Harvest finished slots; returns the number of requests completed.
static int serve_harvest(ggks_server* s, size_t& ncompleted) {
  ncompleted = 0;
  for (ServeSlot& sl : s->slots) {
    if (!sl.busy) continue;
    const hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return fail(GGKS_EDEVICE, "batch failed: %s", hipGetErrorString(q));
    const double t = now_s();
    const size_t eb = (size_t)dtype_bytes(sl.dtype);
    // a status bit set by a kernel of this batch (requests are validated on
    // submission, so only an internal capacity error can set one): its
    // outputs are not trusted, each request is reported failed (latency -1,
    // output untouched) and the poll / drain / trace call returns an error
    uint32_t status = 0;
    memcpy(&status, sl.pinned + sl.pin_stat, 4);
    for (size_t i = 0; i < sl.ids.size(); ++i) {
      ServeReq& rq = s->reqs[sl.ids[i]];
      double lat = -1.0;
      if (!status) {
        memcpy(rq.r.h_out, sl.pinned + sl.pin_out + sl.out_at[i] * eb, (size_t)rq.r.num_k * eb);
        lat = t - rq.submitted;
        s->st.completed++;
        s->st.elements += rq.r.length;
      } else {
        s->st.failed++;
      }
      if (s->trace_latency && rq.trace_idx >= 0) s->trace_latency[rq.trace_idx] = lat;
      else s->finished.emplace_back(rq.r.tag, lat);
      std::vector<uint32_t>().swap(rq.k);
      rq.trace_idx = -1;
      s->free_ids.push_back(sl.ids[i]);  // the request slot is reused
    }
    ncompleted += sl.ids.size();
    s->alloc.release(sl.off, sl.bytes);
    sl.busy = false;
    sl.ids.clear();
    if (status)
      return fail(GGKS_EDEVICE, "a served batch reported internal status 0x%x", status);
  }
  return GGKS_OK;
}

static int serve_submit(ggks_server* s, const ggks_request* r, double at) {
  int rc = serve_validate(s, r);
  if (rc) {
    s->st.rejected++;
    return rc;
  }
  ServeReq q;
  q.r = *r;
  if (r->kind != GGKS_KIND_TOPK) q.k.assign(r->h_k, r->h_k + r->num_k);
  q.submitted = at;
  // a request that could never fit is refused now rather than left waiting
  Plan p;
  std::vector<BatchSeg> one = {serve_seg(q, 0)};
  if (serve_footprint(r->dtype, one, p) > s->budget) {
    s->st.rejected++;
    return fail(GGKS_ENOMEM, "request needs more scratch than the %zu-byte budget", s->budget);
  }
  size_t id;
  if (!s->free_ids.empty()) {  // reuse a finished request's slot
    id = s->free_ids.back();
    s->free_ids.pop_back();
    s->reqs[id] = std::move(q);
  } else {
    id = s->reqs.size();
    s->reqs.push_back(std::move(q));
  }
  s->pending.push_back(id);
  s->last_id = id;
  s->st.submitted++;
  s->st.max_pending = std::max<uint64_t>(s->st.max_pending, s->pending.size());
  return GGKS_OK;
}

static int serve_step(ggks_server* s, size_t& ncompleted) {
  int rc = serve_harvest(s, ncompleted);
  if (rc) return rc;
  return serve_schedule(s);
}

static bool serve_idle(const ggks_server* s) {
  if (!s->pending.empty()) return false;
  for (const ServeSlot& sl : s->slots)
    if (sl.busy) return false;
  return true;
}

static int serve_destroy(ggks_server* s) {
  if (!s) return GGKS_OK;
  (void)hipSetDevice(s->device);
  for (ServeSlot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.pinned) (void)hipHostFree(sl.pinned);
  }
  if (s->pool) (void)hipFree(s->pool);
  delete s;
  return GGKS_OK;
}

// Open-loop replay of an arrival trace (arrival_s ascending, seconds from the
// start): each request is submitted when its time comes, the scheduler runs
// in between, and latency[i] = completion seen - arrival (-1 if refused).
static int serve_run_trace(ggks_server* s, const ggks_request* reqs, const double* arrival,
                           int64_t n, double* latency) {
  for (int64_t i = 1; i < n; ++i)
    if (arrival[i] < arrival[i - 1]) return fail(GGKS_EINVAL, "arrivals not ascending");
  s->trace_latency = latency;
  const double t0 = now_s();
  int64_t i = 0;
  int rc = GGKS_OK;
  size_t nc = 0;
  while (i < n || !serve_idle(s)) {
    const double t = now_s() - t0;
    while (i < n && arrival[i] <= t) {
      if (serve_submit(s, &reqs[i], t0 + arrival[i]) == GGKS_OK) s->reqs[s->last_id].trace_idx = i;
      else latency[i] = -1.0;
      ++i;
    }
    if ((rc = serve_step(s, nc))) break;
  }
  s->trace_latency = nullptr;
  return rc;
}

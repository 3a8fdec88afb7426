// ggks_host.hpp -- host-only logic of the serving loop: the scratch range
// allocator, the request record, the batch footprint and the admission
// decision (which pending requests form the next batch, and how many of them
// fit in the free scratch).  No HIP types or calls: ggks_serve.hpp drives
// these around its device work, and a CPU harness built with address and
// undefined-behaviour sanitizers drives the same code on the host.
#pragma once
#include <stdint.h>

#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "ggks.h"
#include "ggks_plan.hpp"

namespace ggks {

// First-fit allocator over [0, size) with 256-byte granules.
class RangeAlloc {
 public:
  void reset(size_t size) {
    free_.clear();
    size_ = size / 256 * 256;
    if (size_) free_[0] = size_;
    used_ = 0;
  }
  // false when no free range holds `bytes` (also for requests larger than
  // the whole pool, whose rounding would overflow)
  bool alloc(size_t bytes, size_t& off) {
    if (bytes == 0 || bytes > size_) return false;
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      off = it->first;
      const size_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }
  const std::map<size_t, size_t>& free_ranges() const { return free_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0;
  size_t used_ = 0;
};

struct ServeReq {
  ggks_request r;
  std::vector<uint32_t> k;
  double submitted = 0;
  int64_t trace_idx = -1;  // position in a run_trace request list
};

inline BatchSeg serve_seg(const ServeReq& q, uint64_t out_at) {
  BatchSeg g = {};
  g.offset = q.r.offset;
  g.length = q.r.length;
  g.k = q.r.kind == GGKS_KIND_TOPK ? nullptr : q.k.data();
  g.num_k = q.r.num_k;
  g.kind = (uint32_t)q.r.kind;
  g.out_offset = out_at;
  return g;
}

// Device bytes one batch of `segs` takes: its arena (tables and scratch), a
// 256-byte status block and its outputs.  SIZE_MAX when the group cannot be
// planned as one batch (32-bit table offsets): the scheduler then splits it.
inline size_t serve_footprint(int dtype, const std::vector<BatchSeg>& segs, Plan& p) {
  ggks_opts o = {};
  uint64_t outs = 0;
  for (const BatchSeg& g : segs) outs += g.num_k;
  if (outs > kPlanTableLimit || !plan_batch(p, dtype, segs, o)) return SIZE_MAX;
  return align_up(p.total) + 256 + align_up((size_t)outs * dtype_bytes(dtype));
}

// The next batch's candidates: the oldest pending request and the later
// pending requests of its dtype, in queue order, at most max_batch.
inline std::vector<size_t> serve_form_group(const std::vector<ServeReq>& reqs,
                                            const std::deque<size_t>& pending, size_t max_batch) {
  std::vector<size_t> group;
  if (pending.empty()) return group;
  const int dtype = reqs[pending.front()].r.dtype;
  for (size_t id : pending) {
    if (reqs[id].r.dtype != dtype) continue;
    group.push_back(id);
    if (group.size() >= max_batch) break;
  }
  return group;
}

// Admission: the largest prefix of `group` (halving from the whole group)
// whose footprint fits a free range of `alloc`.  On success the range is
// taken, p holds the prefix's plan, and the prefix length is returned; 0
// means not even the oldest request fits now (it waits for memory).
inline size_t serve_admit(RangeAlloc& alloc, const std::vector<ServeReq>& reqs,
                          const std::vector<size_t>& group, Plan& p, size_t& foot, size_t& off) {
  if (group.empty()) return 0;
  const int dtype = reqs[group[0]].r.dtype;
  size_t take = group.size();
  for (;;) {
    std::vector<BatchSeg> segs;
    uint64_t at = 0;
    for (size_t i = 0; i < take; ++i) {
      segs.push_back(serve_seg(reqs[group[i]], at));
      at += reqs[group[i]].r.num_k;
    }
    p = Plan();
    foot = serve_footprint(dtype, segs, p);
    if (foot != SIZE_MAX && alloc.alloc(foot, off)) return take;
    if (take == 1) return 0;
    take = (take + 1) / 2;
  }
}

// Drop the launched requests `ids` (a subsequence of `pending`, in its
// order) from the pending queue, keeping the order of the rest.
inline void serve_remove_launched(std::deque<size_t>& pending, const std::vector<size_t>& ids) {
  std::deque<size_t> rest;
  size_t j = 0;
  for (size_t id : pending) {
    if (j < ids.size() && id == ids[j]) {
      ++j;
      continue;
    }
    rest.push_back(id);
  }
  pending.swap(rest);
}

}  // namespace ggks

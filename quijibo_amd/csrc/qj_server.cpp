// Request loop: continuous batching of field solves on one GPU.
//
// Requests (any grid size, any root count up to max_roots) are copied on
// submit and wait in a pending list; a submission is refused (QJ_ERR_BUSY)
// when the work queued ahead of it in shortest-first order (the outstanding
// requests no larger than it, its own included) exceeds max_pending_work, so
// under overload the large requests are the ones turned away. A scheduler
// thread admits requests while their HBM footprint (field + root records)
// fits a preallocated arena, smallest first, splits admitted requests into
// fixed-size voxel chunks and packs chunks of several requests into one
// launch, shortest remaining work first. A request older than aging_s is
// admitted before all others (oldest first; while it does not fit, requests
// of a quarter of its footprint or more wait behind it) and takes up to a
// quarter of each launch, so large requests are not starved and short ones
// keep flowing.
// Each launch is a hipGraph captured once per in-flight slot and set of FP32
// variants present ([descriptor fetch, FP32 bulk per variant, FP64 band] or
// [descriptor fetch, FP64]) and replayed with new descriptors, which the
// graph's first kernel reads from coherent pinned host memory; requests that
// arrive while launches are on the GPU join the next launch. A finished
// field gets its sampled FP64 cross-check (check_samples) on the server's own
// stream before its completion is posted; with escalate_abs_dF set, an FP32
// result whose sample shows a sign flip, a class mismatch or a larger error
// is computed again with every voxel in the FP64 band kernels. Fields stay resident in the arena
// until completion (or until released when the caller asked to keep them).
// The reference has no serving path; its analogue is one
// OPTIMIZE_3D::computeLogScaledPowerRationalMap call per solve
// (OPTIMIZE_3D.cpp:1272-1370).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <set>
#include <thread>
#include <unordered_map>

#include "qj_internal.h"

namespace qj {

namespace {

// First-fit allocator over one device allocation (offsets, 256 B aligned).
class Arena {
 public:
  bool init(size_t bytes) {
    if (hipMalloc(&base_, bytes) != hipSuccess) return false;
    size_ = bytes;
    free_[0] = bytes;
    return true;
  }
  void release() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
  }
  bool alloc(size_t n, size_t* off) {
    n = (n + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n;
        const size_t o = it->first;
        free_.erase(it);
        if (rest) free_[o + n] = rest;
        used_ += n;
        return true;
      }
    }
    return false;
  }
  void free(size_t off, size_t n) {
    n = (n + 255) & ~(size_t)255;
    used_ -= n;
    auto it = free_.emplace(off, n).first;
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
  char* base() const { return (char*)base_; }
  size_t size() const { return size_; }

 private:
  void* base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

struct Req {
  uint64_t id = 0;
  qj_grid grid{};
  qj_map_params params{};
  int keep = 0;
  double arrival = 0, admit = 0, done = 0;
  std::vector<RootRec64> h64;
  std::vector<RootRec32> h32;
  std::vector<uint8_t> blk32;  // FP32 root block (roots32_view)
  int nt = 0, nb = 0;
  double sct = 1, scb = 1;
  int64_t voxels = 0, next = 0, chunks = 0, chunks_done = 0;
  double work = 0;       // root-term iterations: voxels x max_iter x roots
  double submitted = 0;  // server clock at submission (aging)
  int check_slot = -1;   // per-result cross-check in flight
  bool escalated = false;  // recomputed in FP64 after its FP32 cross-check
  unsigned long long chk[CNT_N] = {};
  size_t off = 0, bytes = 0, off64 = 0, off32 = 0, off_field = 0;
  bool admitted = false;
  qj_status status = QJ_OK;
  MapArgs args{};
  Consts32 k32{};
};

struct Slot {
  ChunkDesc* h_desc = nullptr;
  ChunkDesc* d_desc = nullptr;
  uint32_t* d_band = nullptr;
  uint32_t* d_band_count = nullptr;  // batch counters: band counts and work queues (8 x uint32)
  // one graph per set of FP32 variants present in a batch (bit v = variant v);
  // an FP64 server uses exec[0] only
  hipGraphExec_t exec[16] = {};
  unsigned variants = 0;
  hipEvent_t done = nullptr;
  hipStream_t stream = nullptr;  // one stream per slot: one launch's FP64 band pass overlaps the next's FP32 bulk
  bool busy = false;
  std::vector<std::pair<uint64_t, int>> work;  // (request id, chunks in this launch)
};

}  // namespace

}  // namespace qj

struct qj_server {
  qj_server_config cfg{};
  hipStream_t stream = nullptr;
  qj::Arena arena;
  std::vector<qj::Slot> slots;
  std::vector<qj::ChunkDesc> batch_tmp;  // build_batch scratch
  std::chrono::steady_clock::time_point t0;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<std::unique_ptr<qj::Req>> incoming;
  std::deque<qj_completion> completions;
  std::atomic<bool> stop{false};
  std::atomic<uint64_t> next_id{1};
  // owned by the worker
  std::vector<std::unique_ptr<qj::Req>> pending;              // not admitted yet
  std::unordered_map<uint64_t, std::unique_ptr<qj::Req>> live; // admitted, running
  std::unordered_map<uint64_t, std::unique_ptr<qj::Req>> kept; // finished, field kept
  std::unordered_map<uint64_t, std::unique_ptr<qj::Req>> checking; // finished, cross-check in flight
  // per-result cross-checks: kCheckSlots blocks of CNT_N counters on the
  // device and in pinned host memory, one event each
  static constexpr int kCheckSlots = 64;
  unsigned long long* d_check = nullptr;
  unsigned long long* h_check = nullptr;
  hipEvent_t check_ev[kCheckSlots] = {};
  std::vector<int> check_free;
  double pending_work = 0;  // guarded by mu: work submitted and not yet completed
  std::multiset<double> outstanding;  // guarded by mu: the work of each submitted, not completed request
  std::vector<uint64_t> release_q;
  size_t slot_next = 0;
  qj_status fatal = QJ_OK;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace qj {

static void finish(qj_server* s, Req* r, qj_status st) {
  r->done = s->now();
  qj_completion c{};
  c.id = r->id;
  c.status = st;
  c.arrival_time = r->arrival;
  c.admit_time = r->admit;
  c.done_time = r->done;
  c.voxels = r->voxels;
  c.d_field = (r->keep && st == QJ_OK && r->admitted) ? (const double*)(s->arena.base() + r->off_field) : nullptr;
  c.escalated = r->escalated ? 1 : 0;
  c.checked = (int64_t)r->chk[CNT_CHECKED];
  c.check_sign_flips = (int64_t)r->chk[CNT_CHECK_FLIPS];
  c.check_class_mismatch = (int64_t)r->chk[CNT_CHECK_CLASS];
  memcpy(&c.check_max_abs_dF, &r->chk[CNT_CHECK_MAXABS], sizeof(double));
  memcpy(&c.check_max_rel_dF, &r->chk[CNT_CHECK_MAXREL], sizeof(double));
  std::lock_guard<std::mutex> lk(s->mu);
  s->pending_work -= r->work;
  auto ow = s->outstanding.find(r->work);
  if (ow != s->outstanding.end()) s->outstanding.erase(ow);
  if (s->pending_work < 0.5) s->pending_work = 0.0;  // rounding of the running sum
  s->completions.push_back(c);
  s->cv_done.notify_all();
}

static bool aged(const qj_server* s, const Req* r, double now) { return now - r->submitted > s->cfg.aging_s; }

// Posts a finished request: its completion, then the arena space freed or
// the field kept for the caller.
static void retire(qj_server* s, std::unique_ptr<Req> r) {
  finish(s, r.get(), r->status);
  if (r->keep && r->status == QJ_OK) {
    const uint64_t id = r->id;
    s->kept[id] = std::move(r);
  } else {
    s->arena.free(r->off, r->bytes);
  }
}

static bool try_admit(qj_server* s, Req* r) {
  size_t off;
  if (!s->arena.alloc(r->bytes, &off)) return false;
  r->off = off;
  r->off_field = off;
  r->off64 = off + (((size_t)r->voxels * sizeof(double) + 255) & ~(size_t)255);
  r->off32 = r->off64 + ((r->h64.size() * sizeof(RootRec64) + 255) & ~(size_t)255);
  char* b = s->arena.base();
  // synchronous (a few KB): the launches run on the slots' own streams
  if (hipMemcpy(b + r->off64, r->h64.data(), r->h64.size() * sizeof(RootRec64), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b + r->off32, r->blk32.data(), r->blk32.size(), hipMemcpyHostToDevice) !=
          hipSuccess) {
    s->arena.free(off, r->bytes);
    return false;
  }
  r->admitted = true;
  r->admit = s->now();
  return true;
}

// Fills slot descriptors with up to chunks_per_launch chunks until the launch
// holds about launch_work root-term iterations: requests older than aging_s
// (oldest first) take up to kAgedShare of it while younger ones wait, the
// rest goes shortest remaining work first, and aged requests take whatever
// room is left. An aged request thus progresses at a quarter of the device
// or more without stalling the short requests behind it.
constexpr double kAgedShare = 0.25;

static int build_batch(qj_server* s, Slot& sl) {
  std::vector<Req*> aged_q, young;
  const double now = s->now();
  for (auto& kv : s->live) {
    Req* r = kv.second.get();
    if (r->next < r->voxels) (aged(s, r, now) ? aged_q : young).push_back(r);
  }
  std::sort(aged_q.begin(), aged_q.end(), [](const Req* a, const Req* b) { return a->submitted < b->submitted; });
  std::sort(young.begin(), young.end(), [](const Req* a, const Req* b) {
    const int64_t ra = a->voxels - a->next, rb = b->voxels - b->next;
    return ra != rb ? ra < rb : a->arrival < b->arrival;
  });
  const int maxc = s->cfg.chunks_per_launch, cv = s->cfg.chunk_voxels;
  const double full = s->cfg.launch_work;
  int n = 0;
  double work = 0.0;
  sl.work.clear();
  char* b = s->arena.base();
  std::vector<ChunkDesc>& tmp = s->batch_tmp;
  tmp.resize(maxc);
  // chunks of r while the launch has room below `limit`
  auto take = [&](Req* r, double limit) {
    int taken = 0;
    const double per_voxel = (double)r->params.max_iter * (double)(r->nt + r->nb);
    while (n < maxc && r->next < r->voxels && (n == 0 || work < limit)) {
      ChunkDesc& d = tmp[n++];
      d.a = r->args;
      d.r64 = (const RootRec64*)(b + r->off64);
      d.r32 = (const F4*)(b + r->off32);
      d.out = (double*)(b + r->off_field);
      d.start = r->next;
      const int64_t left = r->voxels - r->next;
      d.count = (int32_t)(left < cv ? left : cv);
      d.k = r->k32;
      r->next += d.count;
      work += per_voxel * d.count;
      taken++;
    }
    if (taken) sl.work.emplace_back(r->id, taken);
  };
  const double aged_limit = young.empty() ? full : full * kAgedShare;
  for (Req* r : aged_q) {
    if (n == maxc || work >= aged_limit) break;
    take(r, aged_limit);
  }
  for (Req* r : young) {
    if (n == maxc || work >= full) break;
    take(r, full);
  }
  for (Req* r : aged_q) {
    if (n == maxc || work >= full) break;
    take(r, full);
  }
  // group by FP32 variant (the kernels take one group each); order inside a
  // group keeps the shortest-remaining-work order
  int g[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < n; i++) g[batch_variant(tmp[i].a) + 1]++;
  int pos[4] = {0, g[1], g[1] + g[2], g[1] + g[2] + g[3]};
  for (int i = 0; i < n; i++) sl.h_desc[pos[batch_variant(tmp[i].a)]++] = tmp[i];
  for (int i = n; i < maxc; i++) sl.h_desc[i].count = 0;
  sl.h_desc[0].total = n;
  sl.h_desc[0].g1 = g[1];
  sl.h_desc[0].g2 = g[1] + g[2];
  sl.h_desc[0].g3 = g[1] + g[2] + g[3];
  sl.variants = s->cfg.precision == QJ_FP64
                    ? 0u
                    : (g[1] ? 1u : 0u) | (g[2] ? 2u : 0u) | (g[3] ? 4u : 0u) | (g[4] ? 8u : 0u);
  return n;
}

static void complete_slot(qj_server* s, Slot& sl, qj_status st) {
  for (auto& w : sl.work) {
    auto it = s->live.find(w.first);
    if (it == s->live.end()) continue;
    Req* r = it->second.get();
    r->chunks_done += w.second;
    if (st != QJ_OK) r->status = st;
    if (r->chunks_done >= r->chunks) {
      std::unique_ptr<Req> own = std::move(it->second);
      s->live.erase(it);
      if (r->status == QJ_OK && s->cfg.check_samples > 0 && !s->check_free.empty()) {
        // the field is complete (every launch that held its chunks has
        // finished): its cross-check runs on the server's stream
        const int slot = s->check_free.back();
        unsigned long long* d = s->d_check + (size_t)slot * CNT_N;
        char* b = s->arena.base();
        qj_status cs = launch_check_one(r->args, (const RootRec64*)(b + r->off64), (const double*)(b + r->off_field),
                                        s->cfg.check_samples, d, s->stream);
        if (cs == QJ_OK &&
            hipMemcpyAsync(s->h_check + (size_t)slot * CNT_N, d, sizeof(unsigned long long) * CNT_N,
                           hipMemcpyDeviceToHost, s->stream) == hipSuccess &&
            hipEventRecord(s->check_ev[slot], s->stream) == hipSuccess) {
          s->check_free.pop_back();
          r->check_slot = slot;
          s->checking[r->id] = std::move(own);
          continue;
        }
      }
      retire(s, std::move(own));
    }
  }
  sl.work.clear();
  sl.busy = false;
}

// Launches a batch into every free slot while live requests have chunks left;
// returns whether it launched anything.
static bool launch_free_slots(qj_server* s) {
  bool launched = false;
  const size_t first = s->slot_next;
  for (size_t k = 0; k < s->slots.size(); k++) {
    Slot& sl = s->slots[(first + k) % s->slots.size()];
    if (sl.busy) continue;
    bool has_work = false;
    for (auto& kv : s->live)
      if (kv.second->next < kv.second->voxels) {
        has_work = true;
        break;
      }
    if (!has_work || build_batch(s, sl) == 0) break;
    hipError_t e = hipGraphLaunch(sl.exec[sl.variants], sl.stream);
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
    sl.busy = true;
    if (e != hipSuccess) complete_slot(s, sl, QJ_ERR_HIP);
    s->slot_next = (first + k + 1) % s->slots.size();
    launched = true;
  }
  return launched;
}

// The worker: each pass retires finished launches and refills their slots
// first (so the GPU waits for the host as little as possible), then admits
// requests and retires cross-checked results.
static void worker_main(qj_server* s) {
  hipSetDevice(s->cfg.device);
  while (true) {
    // 1. take submissions and releases
    {
      std::unique_lock<std::mutex> lk(s->mu);
      while (!s->incoming.empty()) {
        s->pending.push_back(std::move(s->incoming.front()));
        s->incoming.pop_front();
      }
      for (uint64_t id : s->release_q) {
        auto it = s->kept.find(id);
        if (it != s->kept.end()) {
          s->arena.free(it->second->off, it->second->bytes);
          s->kept.erase(it);
        }
      }
      s->release_q.clear();
    }
    // 2. retire finished launches and refill their slots
    for (Slot& sl : s->slots)
      if (sl.busy && hipEventQuery(sl.done) == hipSuccess) complete_slot(s, sl, QJ_OK);
    bool launched = launch_free_slots(s);
    // 3. admit while the arena has room: requests older than aging_s first,
    //    oldest first (and nothing after an aged request that does not fit, so
    //    the space it needs frees up), then smallest footprint first
    bool admitted = false;
    if (!s->pending.empty()) {
      const double now = s->now();
      std::stable_sort(s->pending.begin(), s->pending.end(),
                       [s, now](const std::unique_ptr<Req>& a, const std::unique_ptr<Req>& b) {
                         const bool oa = aged(s, a.get(), now), ob = aged(s, b.get(), now);
                         if (oa != ob) return oa;
                         if (oa) return a->submitted < b->submitted;
                         return a->bytes < b->bytes;
                       });
      // An aged request that does not fit holds back the requests behind it
      // that would compete for its memory (a quarter of its footprint or
      // more); smaller ones still go, since shortest-first packing finishes
      // them, and frees their memory, long before it.
      std::vector<std::unique_ptr<Req>> keep;
      size_t blocked_bytes = SIZE_MAX;
      for (auto& r : s->pending) {
        if (r->bytes > s->arena.size()) {
          finish(s, r.get(), QJ_ERR_CAPACITY);
          continue;
        }
        const bool held = blocked_bytes != SIZE_MAX && r->bytes >= blocked_bytes / 4;
        if (!held && try_admit(s, r.get())) {
          const uint64_t id = r->id;
          s->live[id] = std::move(r);
          admitted = true;
        } else {
          if (!held && aged(s, r.get(), now) && r->bytes < blocked_bytes) blocked_bytes = r->bytes;
          keep.push_back(std::move(r));
        }
      }
      s->pending.swap(keep);
    }
    // 4. retire fields whose cross-check is done
    for (auto it = s->checking.begin(); it != s->checking.end();) {
      Req* r = it->second.get();
      if (hipEventQuery(s->check_ev[r->check_slot]) != hipSuccess) {
        ++it;
        continue;
      }
      memcpy(r->chk, s->h_check + (size_t)r->check_slot * CNT_N, sizeof(r->chk));
      s->check_free.push_back(r->check_slot);
      std::unique_ptr<Req> own = std::move(it->second);
      it = s->checking.erase(it);
      double max_abs;
      memcpy(&max_abs, &own->chk[CNT_CHECK_MAXABS], sizeof max_abs);
      if (s->cfg.escalate_abs_dF > 0 && s->cfg.precision == QJ_FP32_BAND && !own->escalated &&
          (own->chk[CNT_CHECK_FLIPS] > 0 || own->chk[CNT_CHECK_CLASS] > 0 || !(max_abs <= s->cfg.escalate_abs_dF))) {
        // the FP32 bulk missed the tolerance on the sample: every voxel of the
        // request goes again, this time to the FP64 band kernels (a band of
        // infinite width), keeping its arena space and its place in the queue
        own->escalated = true;
        own->next = 0;
        own->chunks_done = 0;
        own->k32.band = __builtin_inff();
        const uint64_t id = own->id;
        s->live[id] = std::move(own);
        admitted = true;
        continue;
      }
      retire(s, std::move(own));
    }
    // new work and a free slot: go round again at once
    bool free_slot = false;
    for (Slot& sl : s->slots) free_slot |= !sl.busy;
    if (admitted && free_slot) continue;
    if (launched) continue;
    bool any_busy = !s->checking.empty();
    for (Slot& sl : s->slots) any_busy |= sl.busy;
    if (s->stop.load() && !any_busy && s->live.empty()) {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->incoming.empty()) break;
    }
    if (any_busy) {
      // all slots busy (or no new work): wait briefly for a launch to finish or a submission
      std::this_thread::sleep_for(std::chrono::microseconds(10));
    } else {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_submit.wait_for(lk, std::chrono::milliseconds(1),
                            [s] { return !s->incoming.empty() || !s->release_q.empty() || s->stop.load(); });
    }
  }
  // fail whatever never got admitted
  for (auto& r : s->pending) finish(s, r.get(), QJ_ERR_SHUTDOWN);
  s->pending.clear();
}

static qj_status capture_slot(qj_server* s, Slot& sl) {
  const int nch = s->cfg.chunks_per_launch;
  // coherent: the graph's first kernel reads the descriptors over the bus
  QJ_HIP(hipHostMalloc((void**)&sl.h_desc, sizeof(ChunkDesc) * nch, hipHostMallocCoherent));
  memset(sl.h_desc, 0, sizeof(ChunkDesc) * nch);
  QJ_HIP(hipMalloc(&sl.d_desc, sizeof(ChunkDesc) * nch));
  QJ_HIP(hipMemset(sl.d_desc, 0, sizeof(ChunkDesc) * nch));
  if (s->cfg.precision == QJ_FP64) {
    QJ_HIP(hipMalloc(&sl.d_band_count, 8 * sizeof(uint32_t)));  // work-queue counters
  } else {
    QJ_HIP(hipMalloc(&sl.d_band, sizeof(uint32_t) * (size_t)nch * s->cfg.chunk_voxels));
    QJ_HIP(hipMalloc(&sl.d_band_count, 8 * sizeof(uint32_t)));
  }
  QJ_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  QJ_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
  const unsigned first = s->cfg.precision == QJ_FP64 ? 0u : 1u, last = s->cfg.precision == QJ_FP64 ? 0u : 15u;
  for (unsigned v = first; v <= last; v++) {
    hipGraph_t g;
    QJ_HIP(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
    qj_status st = launch_desc_fetch(sl.h_desc, sl.d_desc, nch, sl.stream);
    if (st == QJ_OK)
      st = launch_batch(sl.d_desc, nch, s->cfg.chunk_voxels, (qj_precision)s->cfg.precision, sl.d_band,
                        sl.d_band_count, v, sl.stream);
    hipError_t e2 = hipStreamEndCapture(sl.stream, &g);
    if (st) return st;
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(&sl.exec[v], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  }
  return QJ_OK;
}

static void destroy_server(qj_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    for (hipGraphExec_t x : sl.exec)
      if (x) (void)hipGraphExecDestroy(x);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.h_desc) (void)hipHostFree(sl.h_desc);
    if (sl.d_desc) (void)hipFree(sl.d_desc);
    if (sl.d_band) (void)hipFree(sl.d_band);
    if (sl.d_band_count) (void)hipFree(sl.d_band_count);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  s->arena.release();
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (hipEvent_t e : s->check_ev)
    if (e) (void)hipEventDestroy(e);
  if (s->d_check) (void)hipFree(s->d_check);
  if (s->h_check) (void)hipHostFree(s->h_check);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // namespace qj

using namespace qj;

extern "C" {

qj_status qj_server_create(const qj_server_config* cfg, qj_server** out) {
  if (!out) return QJ_ERR_INVALID;
  *out = nullptr;
  qj_server* s = new qj_server();
  if (cfg) s->cfg = *cfg;
  qj_server_config& c = s->cfg;
  if (c.hbm_budget_bytes <= 0) c.hbm_budget_bytes = (int64_t)8 << 30;
  if (c.chunk_voxels <= 0) c.chunk_voxels = 65536;
  if (c.chunks_per_launch <= 0) c.chunks_per_launch = 512;
  if (c.max_roots <= 0) c.max_roots = 256;
  if (c.inflight_launches <= 0) c.inflight_launches = 2;
  if (c.launch_work <= 0) c.launch_work = 3e9;
  if (c.band <= 0) c.band = 0.05;
  if (c.aging_s <= 0) c.aging_s = 0.25;
  if (c.max_pending_work < 0 || c.check_samples < 0 || c.check_samples > (1 << 24) || !(c.escalate_abs_dF >= 0) ||
      (c.escalate_abs_dF > 0 && c.check_samples == 0)) {
    delete s;
    set_error("qj_server_create: max_pending_work must be >= 0, check_samples in [0, 2^24], escalate_abs_dF >= 0 "
              "(and > 0 only with check_samples)");
    return QJ_ERR_INVALID;
  }
  if (c.precision != QJ_FP64 && c.precision != QJ_FP32_BAND) {
    delete s;
    set_error("qj_server_create: unknown precision");
    return QJ_ERR_INVALID;
  }
  if ((int64_t)c.chunk_voxels * c.chunks_per_launch > (int64_t)UINT32_MAX) {
    delete s;
    set_error("qj_server_create: chunk_voxels * chunks_per_launch must stay below 2^32");
    return QJ_ERR_INVALID;
  }
  hipError_t e = hipSetDevice(c.device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "qj_server_create");
  }
  if (!s->arena.init((size_t)c.hbm_budget_bytes)) {
    destroy_server(s);
    set_error("qj_server_create: cannot allocate %lld bytes of HBM", (long long)c.hbm_budget_bytes);
    return QJ_ERR_NOMEM;
  }
  if (hipMalloc(&s->d_check, sizeof(unsigned long long) * CNT_N * qj_server::kCheckSlots) != hipSuccess ||
      hipHostMalloc((void**)&s->h_check, sizeof(unsigned long long) * CNT_N * qj_server::kCheckSlots,
                    hipHostMallocDefault) != hipSuccess) {
    destroy_server(s);
    set_error("qj_server_create: cannot allocate the cross-check buffers");
    return QJ_ERR_NOMEM;
  }
  for (int k = 0; k < qj_server::kCheckSlots; k++) {
    if (hipEventCreateWithFlags(&s->check_ev[k], hipEventDisableTiming) != hipSuccess) {
      destroy_server(s);
      return hip_fail(hipErrorOutOfMemory, "qj_server_create: check events");
    }
    s->check_free.push_back(k);
  }
  s->slots.resize(c.inflight_launches);
  for (Slot& sl : s->slots) {
    qj_status st = capture_slot(s, sl);
    if (st) {
      destroy_server(s);
      return st;
    }
  }
  s->t0 = std::chrono::steady_clock::now();
  s->worker = std::thread(worker_main, s);
  *out = s;
  return QJ_OK;
}

double qj_server_now(qj_server* s) { return s ? s->now() : 0.0; }

qj_status qj_server_submit(qj_server* s, const qj_request* req, uint64_t* id) {
  if (!s || !req) return QJ_ERR_INVALID;
  if (s->stop.load()) return QJ_ERR_SHUTDOWN;
  const qj_poly& t = req->top;
  const qj_poly& b = req->bottom;
  if (t.n_roots < 1 || b.n_roots < 0 || t.n_roots > s->cfg.max_roots || b.n_roots > s->cfg.max_roots ||
      !t.roots || !t.powers || (b.n_roots && (!b.roots || !b.powers))) {
    set_error("qj_server_submit: root counts must be 1..%d (top) and 0..%d (bottom)", s->cfg.max_roots,
              s->cfg.max_roots);
    return QJ_ERR_INVALID;
  }
  if (req->params.flags & QJ_MAP_ADAPTIVE) {
    set_error("qj_server_submit: QJ_MAP_ADAPTIVE is not served by the request loop");
    return QJ_ERR_INVALID;
  }
  if ((req->params.flags & QJ_MAP_LAGRANGE) && s->cfg.precision != QJ_FP64) {
    set_error("qj_server_submit: QJ_MAP_LAGRANGE needs a QJ_FP64 server");
    return QJ_ERR_INVALID;
  }
  auto r = std::make_unique<Req>();
  r->grid = req->grid;
  r->params = req->params;
  r->keep = req->keep_field;
  r->arrival = req->arrival_time;
  // the same root records qj_rational_create builds
  for (int k = 0; k < 2; k++) {
    const qj_poly& p = k ? b : t;
    for (int i = 0; i < p.n_roots; i++) {
      RootRec64 x;
      x.w = p.roots[4 * i];
      x.x = p.roots[4 * i + 1];
      x.y = p.roots[4 * i + 2];
      x.z = p.roots[4 * i + 3];
      x.p = p.powers[i];
      x.e = p.power_scalar * p.powers[i];
      r->h64.push_back(x);
      RootRec32 y;
      y.w = (float)x.w;
      y.x = (float)x.x;
      y.y = (float)x.y;
      y.z = (float)x.z;
      y.e = (float)x.e;
      y.kind = (x.e == 1.0) ? 1 : (x.e == 2.0 ? 2 : 0);
      r->h32.push_back(y);
    }
  }
  r->nt = t.n_roots;
  r->nb = b.n_roots;
  pack_roots32(r->h32.data(), r->nt, r->nb, r->blk32);
  qj_rational f;
  f.nt = r->nt;
  f.nb = r->nb;
  f.sc_top = t.power_scalar;
  f.sc_bottom = b.power_scalar;
  f.h64 = r->h64;
  qj_status st = make_map_args(&r->grid, &f, &r->params, &r->args);
  if (st) return st;
  r->k32 = consts32(r->args, s->cfg.band);
  r->voxels = (int64_t)r->grid.res[0] * r->grid.res[1] * r->grid.res[2];
  r->chunks = (r->voxels + s->cfg.chunk_voxels - 1) / s->cfg.chunk_voxels;
  r->bytes = (((size_t)r->voxels * sizeof(double) + 255) & ~(size_t)255) +
             ((r->h64.size() * sizeof(RootRec64) + 255) & ~(size_t)255) + r->blk32.size();
  r->work = (double)r->voxels * (double)r->params.max_iter * (double)(r->nt + r->nb);
  r->submitted = s->now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    // work queued ahead of this request in shortest-first order, its own
    // included: outstanding requests no larger than it
    double ahead = r->work;
    for (auto it = s->outstanding.begin(); it != s->outstanding.end() && *it <= r->work; ++it) ahead += *it;
    if (s->cfg.max_pending_work > 0 && !s->outstanding.empty() && ahead > s->cfg.max_pending_work) {
      set_error("qj_server_submit: refused, %.3g root-term iterations queued ahead of it in shortest-first order "
                "(limit %.3g)", ahead, s->cfg.max_pending_work);
      return QJ_ERR_BUSY;
    }
    s->pending_work += r->work;
    s->outstanding.insert(r->work);
    r->id = s->next_id.fetch_add(1);
    if (id) *id = r->id;
    s->incoming.push_back(std::move(r));
  }
  s->cv_submit.notify_one();
  return QJ_OK;
}

qj_status qj_server_poll(qj_server* s, qj_completion* out, int32_t max, int32_t* n_out, double timeout_s) {
  if (!s || !n_out || (max > 0 && !out)) return QJ_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completions.empty() && timeout_s > 0)
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [s] { return !s->completions.empty(); });
  int n = 0;
  while (n < max && !s->completions.empty()) {
    out[n++] = s->completions.front();
    s->completions.pop_front();
  }
  *n_out = n;
  return QJ_OK;
}

qj_status qj_server_fetch(qj_server* s, const qj_completion* c, double* dst, void* stream) {
  if (!s || !c || !dst || !c->d_field) {
    set_error("qj_server_fetch: no kept field in this completion");
    return QJ_ERR_INVALID;
  }
  QJ_HIP(hipMemcpyAsync(dst, c->d_field, sizeof(double) * (size_t)c->voxels, hipMemcpyDefault,
                        (hipStream_t)stream));
  QJ_HIP(hipStreamSynchronize((hipStream_t)stream));
  return QJ_OK;
}

qj_status qj_server_release(qj_server* s, uint64_t id) {
  if (!s) return QJ_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->release_q.push_back(id);
  }
  s->cv_submit.notify_one();
  return QJ_OK;
}

qj_status qj_server_destroy(qj_server* s) {
  if (!s) return QJ_OK;
  s->stop.store(true);
  s->cv_submit.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipStreamSynchronize(s->stream);
  destroy_server(s);
  return QJ_OK;
}

}  // extern "C"

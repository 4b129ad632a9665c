// sdv_server.hip -- request engine for many concurrent propagations.
//
// Requests (Coulomb pair potentials of one orbital set, or the kinetic apply
// of one orbital set) from many propagations are served with continuous
// batching:
//   * admission by HBM footprint: each request's output lives in one device
//     arena of a fixed budget; a queued request is admitted (FIFO) as soon as
//     its output block fits;
//   * ticks: the pairs of every admitted request still to be convolved are
//     cut into ticks of a fixed work-buffer size; one tick mixes pairs of
//     several requests (and requests of different grids) in one grouped batch
//     launch, so requests that arrive while others run join the next tick and
//     their work overlaps on the device;
//   * two ticks in flight, each on its own stream with its own work buffer
//     and index tables: the host builds tick t+1 while the device runs tick
//     t, and the device starts tick t+1 while tick t's last workgroups drain;
//   * a request is done when every tick holding its work has completed;
//   * the HBM budget covers the output arena and the two tick work buffers.
// Each pair is computed by the same kernels whatever else is in the tick, so
// a request's result does not depend on the mix.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "sdv_ctx.h"

using namespace sdv;

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// First-fit sub-allocator over one device allocation (offsets only).
struct Arena {
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  size_t used = 0, peak = 0;
  void init(size_t bytes) {
    size = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align_up(bytes, 256);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t roff = it->first + bytes;
        free_.erase(it);
        if (rest) free_[roff] = rest;
        used += bytes;
        peak = std::max(peak, used);
        return true;
      }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes, 256);
    used -= bytes;
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
};

enum { ST_QUEUED = 0, ST_ADMITTED = 1, ST_DONE = 2 };

struct Req {
  long long id = 0;
  sdv_ctx* ctx = nullptr;
  int kind = SDV_REQ_COULOMB;
  const void* in = nullptr;
  size_t out_off = 0, out_bytes = 0;
  const std::vector<PairItem>* items = nullptr;  // Coulomb: pair list of the context (1..N, symmetric)
  size_t next = 0;                                // next pair to schedule
  int inflight = 0;                               // launched ticks holding its work, not yet retired
  hipEvent_t copy_ev = nullptr;                   // after the last sdv_server_copy_result of the output
  int state = ST_QUEUED;
  double t_submit = 0, t_admit = 0, t_done = 0;
};

// One tick slot: its stream (plus the forked streams of the grouped launch),
// work buffer, device tables and pinned staging.
struct Slot {
  hipStream_t st = nullptr;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> fork_ev;
  void* X = nullptr;  // work buffer (tick_bytes)
  void* dtab = nullptr;
  size_t dtab_cap = 0;
  char* htab = nullptr;  // pinned staging
  size_t htab_cap = 0;
  hipEvent_t done = nullptr;
  long long tick = -1;  // tick currently using the slot (-1: free)
  std::vector<long long> members;  // requests with work in this tick
};

}  // namespace

struct sdv_server {
  int precision = 1;
  size_t budget = 0, tick_bytes = 0;
  void* arena_mem = nullptr;
  Arena arena;
  Slot slots[2];
  long long next_id = 1, next_tick = 0;
  std::map<long long, Req> reqs;
  std::deque<long long> queue;    // submitted, not admitted (FIFO)
  std::deque<long long> active;   // admitted Coulomb requests with pairs left to schedule
  std::deque<long long> ke_ready; // admitted kinetic requests not yet launched
  // Released outputs whose copies (sdv_server_copy_result, on the caller's
  // stream) may still be reading them: their blocks return to the arena once
  // the copy's event has completed, so a later request cannot overwrite them.
  struct Deferred {
    size_t off, bytes;
    hipEvent_t ev;
  };
  std::vector<Deferred> deferred;
  std::vector<hipEvent_t> ev_pool;
  int ncu = 256;
  long long completed = 0;
};

static size_t pair_work_bytes(const Geom& g, int precision) {
  return (size_t)2 * g.m * plane_stride(g.n) * (precision == 0 ? sizeof(z64) : sizeof(c32));
}

static int grow_slot_tables(Slot& s, size_t bytes) {
  if (s.dtab_cap < bytes) {
    SDV_HIP_CHECK(hipEventSynchronize(s.done));
    if (s.dtab) SDV_HIP_CHECK(hipFree(s.dtab));
    s.dtab = nullptr;
    s.dtab_cap = 0;
    const size_t cap = align_up(bytes * 2, 1 << 20);
    SDV_REQUIRE(hipMalloc(&s.dtab, cap) == hipSuccess, SDV_ERR_NOMEM, "device allocation of tick tables failed");
    s.dtab_cap = cap;
  }
  if (s.htab_cap < bytes) {
    SDV_HIP_CHECK(hipEventSynchronize(s.done));
    if (s.htab) SDV_HIP_CHECK(hipHostFree(s.htab));
    s.htab = nullptr;
    s.htab_cap = 0;
    const size_t cap = align_up(bytes * 2, 1 << 20);
    SDV_REQUIRE(hipHostMalloc((void**)&s.htab, cap, hipHostMallocDefault) == hipSuccess, SDV_ERR_NOMEM,
                "pinned allocation of tick tables failed");
    s.htab_cap = cap;
  }
  return SDV_OK;
}

// Retire completed ticks; a request is done once every tick holding its
// work has completed and none of its work is left to schedule.
static void retire(sdv_server* v, bool wait_all) {
  for (Slot& s : v->slots) {
    if (s.tick < 0) continue;
    if (!wait_all && hipEventQuery(s.done) != hipSuccess) continue;
    if (wait_all) (void)hipEventSynchronize(s.done);
    const double t = now_ms();
    for (long long id : s.members) {
      auto it = v->reqs.find(id);
      if (it == v->reqs.end()) continue;
      Req& r = it->second;
      r.inflight--;
      if (r.inflight == 0 && r.state == ST_ADMITTED && (r.kind != SDV_REQ_COULOMB || r.next >= r.items->size())) {
        r.state = ST_DONE;
        r.t_done = t;
        v->completed++;
      }
    }
    s.members.clear();
    s.tick = -1;
  }
}

// Return to the arena the released blocks whose copies have finished (wait:
// block until they have).
static void reclaim(sdv_server* v, bool wait) {
  size_t k = 0;
  for (auto& d : v->deferred) {
    if (wait) (void)hipEventSynchronize(d.ev);
    if (wait || hipEventQuery(d.ev) == hipSuccess) {
      v->arena.release(d.off, d.bytes);
      v->ev_pool.push_back(d.ev);
    } else {
      v->deferred[k++] = d;
    }
  }
  v->deferred.resize(k);
}

static void admit(sdv_server* v) {
  reclaim(v, false);
  while (!v->queue.empty()) {
    Req& r = v->reqs[v->queue.front()];
    size_t off = 0;
    if (!v->arena.alloc(r.out_bytes, &off)) break;  // FIFO: wait for memory
    r.out_off = off;
    r.state = ST_ADMITTED;
    r.t_admit = now_ms();
    v->queue.pop_front();
    if (r.kind == SDV_REQ_COULOMB)
      v->active.push_back(r.id);
    else
      v->ke_ready.push_back(r.id);
  }
}

// Build and launch one tick in a free slot; *launched = 0 when there was
// nothing to do or no free slot.  Request and slot state change only after
// every launch of the tick has been enqueued: on an error nothing is marked,
// and the same work is scheduled again by the next poll.
static int launch_tick(sdv_server* v, int* launched) {
  *launched = 0;
  if (v->active.empty() && v->ke_ready.empty()) return SDV_OK;
  Slot* slot = nullptr;
  for (Slot& s : v->slots)
    if (s.tick < 0) {
      slot = &s;
      break;
    }
  if (!slot) return SDV_OK;
  char* base = (char*)v->arena_mem;

  // Coulomb pairs, FIFO over admitted requests, up to the tick's work buffer
  // (planned without touching the requests)
  std::vector<std::vector<PairItem>> slices;
  slices.reserve(v->active.size());  // no reallocation: the sets keep pointers to the slices
  std::vector<std::pair<long long, size_t>> takes;  // (request, pairs taken)
  std::vector<BatchSetHost> sets;
  size_t used = 0;
  for (long long id : v->active) {
    Req& r = v->reqs[id];
    const size_t per = pair_work_bytes(r.ctx->spec->g, v->precision);
    const size_t left = r.items->size() - r.next;
    const size_t take = std::min(left, (v->tick_bytes - used) / per);
    if (take == 0) break;  // tick full (a single pair always fits: checked at submit)
    slices.emplace_back(r.items->begin() + r.next, r.items->begin() + r.next + take);
    takes.emplace_back(id, take);
    used += take * per;
    BatchSetHost h;
    h.spec = r.ctx->spec;
    h.items = &slices.back();
    h.phi = r.in;
    h.out = base + r.out_off;
    h.cnorm = r.ctx->cfg.cnormflag;
    sets.push_back(h);
    if (take < left) break;  // the tick is full
  }

  // kinetic requests: all that are ready
  for (long long id : v->ke_ready) {
    Req& r = v->reqs[id];
    sdv_ctx* c = r.ctx;
    KeArgsHost a;
    a.n = c->n;
    a.howmany = c->N;
    a.ke = c->ke;
    a.invj = c->invj;
    a.sdiag = c->sdiag;
    a.mask = c->mask;
    a.in = r.in;
    a.out = base + r.out_off;
    SDV_TRY(ke_launch(a, v->precision, slot->st, nullptr));
  }
  if (!sets.empty()) {
    BatchPlan bp = batch_plan(sets, v->precision, BATCH_GROUPED, 1, v->ncu);
    SDV_REQUIRE(bp.grid_x > 0, SDV_ERR_UNSUPPORTED, "request holds an unsupported grid");
    SDV_REQUIRE(bp.x_bytes <= v->tick_bytes, SDV_ERR_ARG, "tick work exceeds its buffer");
    const size_t sb = batch_set_bytes(v->precision);
    const size_t o_sets = 0;
    const size_t o_items = align_up(o_sets + sb * sets.size(), 256);
    const size_t o_wx = align_up(o_items + bp.items.size() * batch_item_bytes(), 256);
    const size_t o_wy = align_up(o_wx + bp.wx.size() * batch_work_bytes(), 256);
    const size_t total = align_up(o_wy + bp.wy.size() * batch_work_bytes(), 256);
    SDV_TRY(grow_slot_tables(*slot, total));
    // the slot's previous tick has completed (retired), so its staging is free
    for (size_t i = 0; i < sets.size(); ++i) batch_fill_set(v->precision, sets[i], slot->htab + o_sets + i * sb);
    std::memcpy(slot->htab + o_items, bp.items.data(), bp.items.size() * batch_item_bytes());
    std::memcpy(slot->htab + o_wx, bp.wx.data(), bp.wx.size() * batch_work_bytes());
    std::memcpy(slot->htab + o_wy, bp.wy.data(), bp.wy.size() * batch_work_bytes());
    SDV_HIP_CHECK(hipMemcpyAsync(slot->dtab, slot->htab, total, hipMemcpyHostToDevice, slot->st));
    BatchDev d;
    char* dt = (char*)slot->dtab;
    d.sets = dt + o_sets;
    d.items = (const BItemDev*)(dt + o_items);
    d.wx = (const BWork*)(dt + o_wx);
    d.wy = (const BWork*)(dt + o_wy);
    d.X = slot->X;
    d.nside = (int)slot->side.size();
    d.side = d.nside ? slot->side.data() : nullptr;
    d.ev = d.nside ? slot->fork_ev.data() : nullptr;
    SDV_TRY(batch_launch(bp, d, slot->st, nullptr));
  }
  SDV_HIP_CHECK(hipEventRecord(slot->done, slot->st));

  // commit
  slot->tick = v->next_tick++;
  slot->members.clear();
  for (long long id : v->ke_ready) {
    v->reqs[id].inflight++;
    slot->members.push_back(id);
  }
  v->ke_ready.clear();
  for (const auto& tk : takes) {
    Req& r = v->reqs[tk.first];
    r.next += tk.second;
    r.inflight++;
    slot->members.push_back(r.id);
    if (r.next >= r.items->size()) {
      // only the front request can be finished here: takes run over the FIFO in order
      v->active.pop_front();
    }
  }
  *launched = 1;
  return SDV_OK;
}

static void server_release(sdv_server* v) {
  for (Slot& s : v->slots)
    if (s.st) (void)hipStreamSynchronize(s.st);
  for (Slot& s : v->slots) {
    (void)hipFree(s.X);
    (void)hipFree(s.dtab);
    if (s.htab) (void)hipHostFree(s.htab);
    if (s.done) (void)hipEventDestroy(s.done);
    for (hipStream_t st : s.side) (void)hipStreamDestroy(st);
    for (hipEvent_t e : s.fork_ev) (void)hipEventDestroy(e);
    if (s.st) (void)hipStreamDestroy(s.st);
  }
  for (auto& d : v->deferred) (void)hipEventSynchronize(d.ev);
  for (auto& d : v->deferred) (void)hipEventDestroy(d.ev);
  for (hipEvent_t e : v->ev_pool) (void)hipEventDestroy(e);
  for (auto& kv : v->reqs)
    if (kv.second.copy_ev) {
      (void)hipEventSynchronize(kv.second.copy_ev);
      (void)hipEventDestroy(kv.second.copy_ev);
    }
  (void)hipFree(v->arena_mem);
}

extern "C" {

int sdv_server_create(long long hbm_budget, long long tick_bytes, int precision, sdv_server** out) {
  SDV_REQUIRE(out && hbm_budget > 0 && tick_bytes > 0, SDV_ERR_ARG, "bad server arguments");
  SDV_REQUIRE(precision == 0 || precision == 1, SDV_ERR_ARG, "precision must be 0 or 1");
  sdv_server* v = new sdv_server();
  v->precision = precision;
  v->budget = (size_t)hbm_budget;
  v->tick_bytes = (size_t)tick_bytes;
  // the budget holds the two tick work buffers; the rest is the output arena
  const size_t work = 2 * align_up(v->tick_bytes + 256, 256);
  if (v->budget <= work) {
    delete v;
    set_error("HBM budget does not exceed the two tick work buffers");
    return SDV_ERR_ARG;
  }
  const size_t arena_bytes = v->budget - work;
  int dev = 0;
  int rc = SDV_OK;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    set_error("HIP set-up of the server failed");
    rc = SDV_ERR_HIP;
  }
  for (Slot& s : v->slots) {
    if (rc == SDV_OK && hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess) rc = SDV_ERR_HIP;
    for (int i = 0; i < 4 && rc == SDV_OK; ++i) {
      hipStream_t st = nullptr;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = SDV_ERR_HIP;
      else s.side.push_back(st);
    }
    for (int i = 0; i < 5 && rc == SDV_OK; ++i) {
      hipEvent_t e = nullptr;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) rc = SDV_ERR_HIP;
      else s.fork_ev.push_back(e);
    }
    if (rc == SDV_OK && hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) rc = SDV_ERR_HIP;
    if (rc == SDV_OK && hipMalloc(&s.X, v->tick_bytes + 256) != hipSuccess) rc = SDV_ERR_NOMEM;
  }
  if (rc == SDV_OK && hipMalloc(&v->arena_mem, arena_bytes) != hipSuccess) rc = SDV_ERR_NOMEM;
  if (rc != SDV_OK) {
    if (rc == SDV_ERR_NOMEM) set_error("device allocation of the server arena or work buffers failed");
    server_release(v);
    delete v;
    return rc;
  }
  v->arena.init(arena_bytes);
  *out = v;
  return SDV_OK;
}

int sdv_server_submit(sdv_server* v, sdv_ctx* c, int kind, const void* inspfs, long long* request_id) {
  SDV_REQUIRE(v && c && inspfs && request_id, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(kind == SDV_REQ_COULOMB || kind == SDV_REQ_KE, SDV_ERR_ARG, "unknown request kind");
  Req r;
  r.id = v->next_id++;
  r.ctx = c;
  r.kind = kind;
  r.in = inspfs;
  const size_t col = (size_t)c->n3 * sizeof(z64);
  if (kind == SDV_REQ_COULOMB) {
    IdxPlan* pl = nullptr;
    SDV_TRY(twoe_plan(c, 1, c->N, true, &pl));
    r.items = &pl->items;
    r.out_bytes = col * c->N * c->N;
    SDV_REQUIRE(pair_work_bytes(c->spec->g, v->precision) <= v->tick_bytes, SDV_ERR_ARG,
                "one pair's work exceeds the tick buffer");
  } else {
    r.out_bytes = col * c->N;
  }
  SDV_REQUIRE(align_up(r.out_bytes, 256) <= v->arena.size, SDV_ERR_ARG,
              "request output exceeds the HBM budget left beside the tick work buffers");
  r.t_submit = now_ms();
  v->reqs[r.id] = r;
  v->queue.push_back(r.id);
  *request_id = r.id;
  return SDV_OK;
}

int sdv_server_poll(sdv_server* v, int wait, int* ncompleted) {
  SDV_REQUIRE(v, SDV_ERR_ARG, "null server");
  const long long before = v->completed;
  retire(v, false);
  admit(v);
  for (;;) {
    int launched = 0;
    SDV_TRY(launch_tick(v, &launched));
    if (!launched) break;
    admit(v);
  }
  if (wait) {  // block until every request submitted so far is done
    while (!v->queue.empty() || !v->active.empty() || !v->ke_ready.empty() || v->slots[0].tick >= 0 ||
           v->slots[1].tick >= 0) {
      retire(v, true);
      admit(v);
      int launched = 1;
      while (launched) SDV_TRY(launch_tick(v, &launched));
      if (!v->queue.empty() && v->active.empty() && v->ke_ready.empty() && v->slots[0].tick < 0 &&
          v->slots[1].tick < 0) {
        if (v->deferred.empty()) break;  // queued requests wait for outputs still to be released
        reclaim(v, true);                // or for copies of released outputs to finish
      }
    }
  }
  if (ncompleted) *ncompleted = (int)(v->completed - before);
  return SDV_OK;
}

int sdv_server_result(sdv_server* v, long long id, void** out, int* state, double* latency_ms, double* queue_ms) {
  SDV_REQUIRE(v, SDV_ERR_ARG, "null server");
  auto it = v->reqs.find(id);
  SDV_REQUIRE(it != v->reqs.end(), SDV_ERR_ARG, "unknown or released request");
  const Req& r = it->second;
  if (state) *state = r.state;
  if (out) *out = r.state == ST_DONE ? (char*)v->arena_mem + r.out_off : nullptr;
  if (latency_ms) *latency_ms = r.state == ST_DONE ? r.t_done - r.t_submit : -1.0;
  if (queue_ms) *queue_ms = r.state != ST_QUEUED ? r.t_admit - r.t_submit : -1.0;
  return SDV_OK;
}

int sdv_server_copy_result(sdv_server* v, long long id, void* dst, void* stream) {
  SDV_REQUIRE(v && dst, SDV_ERR_ARG, "null argument");
  auto it = v->reqs.find(id);
  SDV_REQUIRE(it != v->reqs.end(), SDV_ERR_ARG, "unknown or released request");
  SDV_REQUIRE(it->second.state == ST_DONE, SDV_ERR_ARG, "request not done");
  Req& r = it->second;
  SDV_HIP_CHECK(hipMemcpyAsync(dst, (char*)v->arena_mem + r.out_off, r.out_bytes, hipMemcpyDeviceToDevice,
                               (hipStream_t)stream));
  if (!r.copy_ev) {
    if (!v->ev_pool.empty()) {
      r.copy_ev = v->ev_pool.back();
      v->ev_pool.pop_back();
    } else {
      SDV_HIP_CHECK(hipEventCreateWithFlags(&r.copy_ev, hipEventDisableTiming));
    }
  }
  SDV_HIP_CHECK(hipEventRecord(r.copy_ev, (hipStream_t)stream));
  return SDV_OK;
}

int sdv_server_release(sdv_server* v, long long id) {
  SDV_REQUIRE(v, SDV_ERR_ARG, "null server");
  auto it = v->reqs.find(id);
  SDV_REQUIRE(it != v->reqs.end(), SDV_ERR_ARG, "unknown or released request");
  SDV_REQUIRE(it->second.state == ST_DONE, SDV_ERR_ARG, "request not done");
  const Req& r = it->second;
  if (r.copy_ev)
    v->deferred.push_back(sdv_server::Deferred{r.out_off, r.out_bytes, r.copy_ev});
  else
    v->arena.release(r.out_off, r.out_bytes);
  v->reqs.erase(it);
  return SDV_OK;
}

int sdv_server_stats(sdv_server* v, long long* queued, long long* in_flight, long long* arena_used,
                     long long* arena_peak, long long* ticks) {
  SDV_REQUIRE(v, SDV_ERR_ARG, "null server");
  reclaim(v, false);
  long long q = (long long)v->queue.size(), f = 0;
  for (auto& kv : v->reqs)
    if (kv.second.state == ST_ADMITTED) ++f;
  if (queued) *queued = q;
  if (in_flight) *in_flight = f;
  if (arena_used) *arena_used = (long long)v->arena.used;
  if (arena_peak) *arena_peak = (long long)v->arena.peak;
  if (ticks) *ticks = v->next_tick;
  return SDV_OK;
}

int sdv_server_free(sdv_server* v) {
  if (!v) return SDV_OK;
  server_release(v);
  delete v;
  return SDV_OK;
}

}  // extern "C"

// Batched and served solves on one GPU (include/ecogen_pool.h).
//
// A pool owns an HBM arena (one hipMalloc of the budget), a table of device
// control blocks (one per case handle), and one kernel class per (model,
// dimension, order, precision, limiter form).  Each class keeps a device
// argument table (four StageArgs per case: launch-slot parity x first/second
// kernel of the step) and a tile table listing every tile of its running
// cases.  A chunk of K steps (K a multiple of 6, so that slot-indexed
// ping-pong and ring entries repeat) launches, per step, the persistent
// first-stage kernel of every class and then the second-stage kernel of every
// class; the chunk is captured once as a hipGraph and replayed.  Table
// contents change between replays by stream-ordered copies; the graph is
// recaptured only when the set of non-empty classes changes.
//
// After each chunk the control blocks are copied to a pinned snapshot and an
// event is recorded; polling reads the snapshots of completed chunks, so the
// host never waits for the chunk that is running.  Memory of a released case
// is returned to the arena once no enqueued chunk refers to it.
//
// Admissions upload and initialise on a second stream (default priority):
// FP64 initial state -> staging -> total energy (k_init) -> FP32 storage
// conversion when needed; the compute stream waits on the admission event
// before the next chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "../../include/ecogen_pool.h"
#include "eco_host.h"
#include "eco_launch.h"

using namespace eco;

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation, coalescing on free.
class Arena {
 public:
  bool init(size_t bytes) {
    size_ = bytes & ~(kAlign - 1);
    if (hipMalloc(&base_, size_) != hipSuccess) { base_ = nullptr; return false; }
    free_[0] = size_;
    return true;
  }
  ~Arena() { if (base_) (void)hipFree(base_); }
  // returns the byte offset or SIZE_MAX
  size_t alloc(size_t bytes) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_ += bytes;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
    }
  }
  size_t largest() const {
    size_t m = 0;
    for (auto &f : free_) m = std::max(m, f.second);
    return m;
  }
  size_t used() const { return used_; }
  char *ptr(size_t off) const { return (char *)base_ + off; }

 private:
  void *base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

struct ClassKey {
  int model, dim, order, prec, same, nph;
  bool operator==(const ClassKey &o) const {
    return model == o.model && dim == o.dim && order == o.order && prec == o.prec && same == o.same && nph == o.nph;
  }
};

struct Case {
  int id = -1;
  int cls = -1;
  int argslot = -1;         // slot in the class's argument table
  eco_config cfg;
  Lowered low;
  size_t off = 0, bytes = 0;  // arena block
  double *buf[2] = {nullptr, nullptr};
  double *bstar = nullptr;
  double *tables = nullptr;
  long long g0 = 0;         // pool slot of the case's first step
  int ntiles = 0;
  int running = 0;          // in the tile table
  int finished = 0;         // reported by poll
  long long first_chunk = -1;  // first chunk that included the case
  long long ref_chunk = -1; // last chunk that may refer to the case
  int released = 0;
};

struct Class {
  ClassKey key;
  int grid = 0;
  StageArgs *d_args = nullptr;      // max_cases * 4
  StageArgs *h_args = nullptr;      // pinned mirror
  std::vector<int> free_argslots;
  TileRef *d_tiles = nullptr;
  size_t tiles_cap = 0;
  BatchDesc *d_desc = nullptr;
  std::vector<int> running;         // case ids
  bool dirty = true;
  int ntiles = 0;
};

struct PinnedSlot {
  void *p = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool used = false;
};

struct Chunk {
  long long seq;
  hipEvent_t ev;
  Control *snap;  // pinned snapshot of the control table after the chunk
  int ncases;     // control entries copied
};

}  // namespace

struct eco_pool {
  eco_pool_config pc;
  hipStream_t cs = nullptr;  // compute stream
  hipStream_t us = nullptr;  // upload / download stream
  int cus = 0;
  Arena arena;
  int max_cases = 0;
  Control *d_ctl = nullptr;
  std::vector<Case> cases;       // indexed by handle
  std::vector<int> free_ids;
  int hi_id = 0;                 // 1 + highest handle ever used
  std::vector<Class> classes;
  int K = 24;
  long long slot = 0;            // pool slot of the next step to enqueue
  long long chunk_seq = 0;       // chunks enqueued
  long long chunk_done = 0;      // chunks whose snapshot has been read
  std::deque<Chunk> inflight;
  std::vector<Control *> snap_free;
  std::vector<PinnedSlot> pinned;  // staging of table uploads
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<int> graph_classes;  // classes in the captured graph
  bool admit_pending = false;
  hipEvent_t admit_ev = nullptr;
  std::vector<std::pair<size_t, size_t>> staging_to_free;  // arena blocks freed after admit_ev
  long long steps_launched = 0, tiles_launched = 0;
  int captures = 0;
  std::vector<int> ready;          // finished, not yet reported
  double *fetch_tmp = nullptr;     // FP64 staging of downloads (outside the arena)
  size_t fetch_cap = 0;
};

static int hipc(hipError_t e) { return e == hipSuccess ? ECO_OK : ECO_ERR_HIP; }

static size_t case_bytes(const eco_config &c, const Lowered &L) {
  const size_t nc = (size_t)L.geo.ncell;
  size_t b = 2 * align_up(nc * L.nlevel * L.es);
  if (c.order == 2) b += align_up(nc * L.nprim * L.es);
  b += align_up(L.tab.size() * sizeof(double));
  return b;
}

static size_t staging_bytes(const eco_config &c, const Lowered &L) {
  // FP64 device-layout staging of the initial state (FP32 storage converts from it)
  return (c.precision == ECO_FP32S) ? align_up((size_t)L.geo.ncell * (L.nprim + 1) * sizeof(double)) : 0;
}

extern "C" unsigned long long eco_pool_case_bytes(const eco_config *cfg) {
  if (validate_config(cfg) != ECO_OK) return 0;
  Lowered L;
  lower_config(*cfg, L);
  return case_bytes(*cfg, L);
}

static void wait_event(hipEvent_t e) { (void)hipEventSynchronize(e); }

// pinned staging buffer for a host->device table copy enqueued on the compute stream
static void *pinned_get(eco_pool *p, size_t bytes) {
  for (auto &s : p->pinned) {
    if (s.used && hipEventQuery(s.ev) != hipSuccess) continue;
    if (s.cap < bytes) {
      if (s.p) (void)hipHostFree(s.p);
      s.cap = std::max(bytes, (size_t)1 << 16);
      if (hipHostMalloc(&s.p, s.cap, 0) != hipSuccess) { s.p = nullptr; s.cap = 0; return nullptr; }
    }
    s.used = true;
    return s.p;
  }
  PinnedSlot s;
  s.cap = std::max(bytes, (size_t)1 << 16);
  if (hipHostMalloc(&s.p, s.cap, 0) != hipSuccess) return nullptr;
  if (hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess) return nullptr;
  s.used = true;
  p->pinned.push_back(s);
  return p->pinned.back().p;
}

static void pinned_mark(eco_pool *p, void *ptr) {
  for (auto &s : p->pinned)
    if (s.p == ptr) { (void)hipEventRecord(s.ev, p->cs); return; }
}

extern "C" int eco_pool_create(const eco_pool_config *pc, eco_pool **out) {
  if (!pc || !out) return ECO_ERR_INVALID;
  *out = nullptr;
  if (pc->hbm_budget < (1ull << 20)) return ECO_ERR_INVALID;
  std::unique_ptr<eco_pool> p(new (std::nothrow) eco_pool());
  if (!p) return ECO_ERR_NOMEM;
  p->pc = *pc;
  p->max_cases = pc->max_cases > 0 ? pc->max_cases : 8192;
  int k = pc->steps_per_chunk > 0 ? pc->steps_per_chunk : 24;
  p->K = ((k + 5) / 6) * 6;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return ECO_ERR_HIP;
  p->cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&p->cs, hipStreamNonBlocking) != hipSuccess) return ECO_ERR_HIP;
  if (hipStreamCreateWithFlags(&p->us, hipStreamNonBlocking) != hipSuccess) return ECO_ERR_HIP;
  if (hipEventCreateWithFlags(&p->admit_ev, hipEventDisableTiming) != hipSuccess) return ECO_ERR_HIP;
  if (!p->arena.init(pc->hbm_budget)) return ECO_ERR_NOMEM;
  if (hipMalloc(&p->d_ctl, sizeof(Control) * (size_t)p->max_cases) != hipSuccess) return ECO_ERR_NOMEM;
  p->cases.resize(p->max_cases);
  for (int i = p->max_cases - 1; i >= 0; i--) p->free_ids.push_back(i);
  *out = p.release();
  return ECO_OK;
}

static int class_for(eco_pool *p, const ClassKey &key) {
  for (size_t i = 0; i < p->classes.size(); i++)
    if (p->classes[i].key == key) return (int)i;
  Class c;
  c.key = key;
  if (hipMalloc(&c.d_args, sizeof(StageArgs) * 4 * (size_t)p->max_cases) != hipSuccess ||
      hipHostMalloc(&c.h_args, sizeof(StageArgs) * 4 * (size_t)p->max_cases, 0) != hipSuccess ||
      hipMalloc(&c.d_desc, sizeof(BatchDesc)) != hipSuccess) {
    if (c.d_args) (void)hipFree(c.d_args);
    if (c.h_args) (void)hipHostFree(c.h_args);
    if (c.d_desc) (void)hipFree(c.d_desc);
    return -1;
  }
  for (int i = p->max_cases - 1; i >= 0; i--) c.free_argslots.push_back(i);
  // persistent grid: resident workgroups of the heavier (second) kernel of the step
  int occ = 0;
  BatchLaunch L{key.dim, key.order == 1 ? 0 : 2, key.same, nullptr, nullptr, nullptr, 0, 0, &occ};
  if (launch_batch(key.model, key.prec, key.nph, L, p->cs) != 0 || occ <= 0) occ = 1;
  c.grid = p->pc.grid_blocks > 0 ? p->pc.grid_blocks : occ * p->cus;
  p->classes.push_back(std::move(c));
  return (int)p->classes.size() - 1;
}

// StageArgs of a case for slot parity `par`, kernel `kidx` (0 first, 1 second)
static StageArgs case_args(const Case &cs, Control *ctl, int par, int kidx) {
  StageArgs A;
  A.geo = cs.low.geo;
  A.sch = cs.low.sch;
  A.ctl = ctl;
  A.step = 0;
  A.wn = cs.buf[par];
  A.wsrc = A.wn;
  A.wout = cs.buf[par ^ 1];
  if (cs.cfg.order == 2) {
    if (kidx == 0) A.wout = cs.bstar;
    else A.wsrc = cs.bstar;
  }
  return A;
}

extern "C" int eco_pool_admit(eco_pool *p, const eco_config *cfg, const double *host_state, int nsteps,
                              double t_final, int *case_id) {
  if (!p || !cfg || !host_state || !case_id) return ECO_ERR_INVALID;
  int rc = validate_config(cfg);
  if (rc) return rc;
  if (cfg->precision == ECO_FP32S && cfg->order != 2) return ECO_ERR_UNSUPPORTED;
  // transported scalars run in single-solve contexts (their own pass per stage)
  if (cfg->ntransport > 0) return ECO_ERR_UNSUPPORTED;
  // ThermalEq, EulerHomogeneous and RelaxationPTMu run in single-solve contexts
  if (cfg->model == ECO_MODEL_THERMALEQ || cfg->model == ECO_MODEL_EULERHOM || cfg->relax_ptmu) return ECO_ERR_UNSUPPORTED;
  if (p->free_ids.empty()) return ECO_ERR_FULL;
  Lowered L;
  lower_config(*cfg, L);
  const size_t need = case_bytes(*cfg, L), stg = staging_bytes(*cfg, L);
  size_t off = p->arena.alloc(need);
  if (off == SIZE_MAX) return ECO_ERR_BUDGET;
  size_t soff = SIZE_MAX;
  if (stg) {
    soff = p->arena.alloc(stg);
    if (soff == SIZE_MAX) { p->arena.release(off, need); return ECO_ERR_BUDGET; }
  }
  ClassKey key{cfg->model, L.geo.dim, cfg->order, cfg->precision, L.sch.same_lim ? 1 : 0,
               cfg->model == ECO_MODEL_EULER ? 1 : cfg->nphase};
  int ci = class_for(p, key);
  if (ci < 0) {
    p->arena.release(off, need);
    if (stg) p->arena.release(soff, stg);
    return ECO_ERR_NOMEM;
  }
  Class &cl = p->classes[ci];
  const int id = p->free_ids.back();
  p->free_ids.pop_back();
  p->hi_id = std::max(p->hi_id, id + 1);
  Case &cs = p->cases[id];
  cs = Case();
  cs.id = id;
  cs.cls = ci;
  cs.cfg = *cfg;
  cs.low = L;
  cs.off = off;
  cs.bytes = need;
  const size_t nc = (size_t)L.geo.ncell;
  char *b = p->arena.ptr(off);
  cs.buf[0] = (double *)b; b += align_up(nc * L.nlevel * L.es);
  cs.buf[1] = (double *)b; b += align_up(nc * L.nlevel * L.es);
  if (cfg->order == 2) { cs.bstar = (double *)b; b += align_up(nc * L.nprim * L.es); }
  cs.tables = (double *)b;
  bind_tables(*cfg, cs.low, cs.tables);
  int tt[3];
  tile_shape(L.geo.dim, tt);
  cs.ntiles = 1;
  for (int d = 0; d < 3; d++) cs.ntiles *= (L.geo.n[d] + tt[d] - 1) / tt[d];
  cs.g0 = p->slot;  // joins at the next chunk
  cs.argslot = cl.free_argslots.back();
  cl.free_argslots.pop_back();
  // a failure from here on gives back what the admission took: the argument
  // slot, the handle and the arena blocks (after the upload stream has
  // drained, so no copy still writes into them), and resets the case
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(p->us);
    cl.free_argslots.push_back(cs.argslot);
    p->arena.release(off, need);
    if (stg) p->arena.release(soff, stg);
    cs = Case();
    p->free_ids.push_back(id);
    return code;
  };

  // --- upload and initialise on the upload stream ---
  hipError_t e = hipMemcpyAsync(cs.tables, cs.low.tab.data(), cs.low.tab.size() * sizeof(double),
                                hipMemcpyHostToDevice, p->us);
  double *dst = (cfg->precision == ECO_FP32S) ? (double *)p->arena.ptr(soff) : cs.buf[cs.g0 & 1];
  std::vector<int> map;
  host_field_map(*cfg, L.geo.dim, map);
  for (size_t f = 0; f < map.size() && e == hipSuccess; f++)
    e = hipMemcpyAsync(dst + f * nc, host_state + (size_t)map[f] * nc, nc * sizeof(double), hipMemcpyHostToDevice,
                       p->us);
  if (e != hipSuccess) return fail(ECO_ERR_HIP);
  StageArgs A = case_args(cs, p->d_ctl + id, (int)(cs.g0 & 1), 0);
  A.wsrc = A.wn = dst;
  A.wout = dst;
  rc = launch_stage(A, 3, p->us);  // total energy (Cell::completeFulfillState)
  if (!rc && cfg->precision == ECO_FP32S) {
    A.wout = cs.buf[cs.g0 & 1];
    rc = launch_convert(A, 1, dst, p->us);
  }
  if (rc) return fail(rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP);
  // control block: zero, target
  Control h;
  std::memset(&h, 0, sizeof(h));
  h.t_final = (nsteps > 0) ? -1. : t_final;
  h.max_steps = (nsteps > 0) ? nsteps : 0;
  // the control entry may still be read by no-op tiles of an earlier case with
  // this handle only if that case's chunks are still in flight; released
  // handles return to free_ids only after those chunks completed
  e = hipMemcpyAsync(p->d_ctl + id, &h, sizeof(Control), hipMemcpyHostToDevice, p->us);
  // argument blocks
  for (int par = 0; par < 2 && e == hipSuccess; par++)
    for (int k = 0; k < 2; k++) cl.h_args[(size_t)cs.argslot * 4 + par * 2 + k] = case_args(cs, p->d_ctl + id, par, k);
  if (e == hipSuccess)
    e = hipMemcpyAsync(cl.d_args + (size_t)cs.argslot * 4, cl.h_args + (size_t)cs.argslot * 4, 4 * sizeof(StageArgs),
                       hipMemcpyHostToDevice, p->us);
  if (e == hipSuccess) e = hipEventRecord(p->admit_ev, p->us);
  if (e != hipSuccess) return fail(ECO_ERR_HIP);
  // the upload stream is synchronised here only for the host arrays it reads
  // (pageable copies complete before hipMemcpyAsync returns for pageable
  // sources; the control block and h is on the stack)
  e = hipStreamSynchronize(p->us);
  if (e != hipSuccess) return fail(ECO_ERR_HIP);
  if (stg) p->arena.release(soff, stg);
  p->admit_pending = true;
  cs.running = 1;
  cl.running.push_back(id);
  cl.dirty = true;
  *case_id = id;
  return ECO_OK;
}

static int upload_tables(eco_pool *p) {
  for (auto &cl : p->classes) {
    if (!cl.dirty) continue;
    size_t nt = 0;
    for (int id : cl.running) nt += (size_t)p->cases[id].ntiles;
    if (nt > cl.tiles_cap) {
      // grow: no enqueued work may still read the old table
      hipError_t e = hipStreamSynchronize(p->cs);
      if (e != hipSuccess) return ECO_ERR_HIP;
      if (cl.d_tiles) (void)hipFree(cl.d_tiles);
      cl.tiles_cap = std::max(nt, cl.tiles_cap * 2);
      if (hipMalloc(&cl.d_tiles, cl.tiles_cap * sizeof(TileRef)) != hipSuccess) { cl.d_tiles = nullptr; return ECO_ERR_NOMEM; }
      // the captured launches hold the table pointer: capture again
      if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
      if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
    }
    const size_t bytes = nt * sizeof(TileRef) + sizeof(BatchDesc);
    char *h = (char *)pinned_get(p, bytes);
    if (!h) return ECO_ERR_NOMEM;
    TileRef *ht = (TileRef *)(h + sizeof(BatchDesc));
    size_t w = 0;
    for (int id : cl.running) {
      const Case &cs = p->cases[id];
      for (int t = 0; t < cs.ntiles; t++) ht[w++] = TileRef{cs.argslot, t};
    }
    BatchDesc *hd = (BatchDesc *)h;
    hd->args = cl.d_args;
    hd->tiles = cl.d_tiles;
    hd->ntiles = (int)nt;
    hipError_t e = hipSuccess;
    if (nt) e = hipMemcpyAsync(cl.d_tiles, ht, nt * sizeof(TileRef), hipMemcpyHostToDevice, p->cs);
    if (e == hipSuccess) e = hipMemcpyAsync(cl.d_desc, hd, sizeof(BatchDesc), hipMemcpyHostToDevice, p->cs);
    if (e != hipSuccess) return ECO_ERR_HIP;
    pinned_mark(p, h);
    cl.ntiles = (int)nt;
    cl.dirty = false;
  }
  return ECO_OK;
}

// launches of one chunk (K steps) for the classes in `cls`, slots base..base+K-1
static int enqueue_steps(eco_pool *p, const std::vector<int> &cls, long long base) {
  for (int i = 0; i < p->K; i++) {
    const int sl = (int)((base + i) % 6);
    for (int pass = 0; pass < 2; pass++) {
      for (int c : cls) {
        Class &cl = p->classes[c];
        const int stage = (cl.key.order == 1) ? (pass == 0 ? 0 : -1) : pass + 1;
        if (stage < 0) continue;
        BatchLaunch L{cl.key.dim, stage, cl.key.same, cl.d_args, cl.d_tiles, &cl.d_desc->ntiles, sl, cl.grid,
                      nullptr};
        int rc = launch_batch(cl.key.model, cl.key.prec, cl.key.nph, L, p->cs);
        if (rc) return rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;
      }
    }
  }
  return ECO_OK;
}

extern "C" int eco_pool_advance(eco_pool *p) {
  if (!p) return ECO_ERR_INVALID;
  int rc = upload_tables(p);
  if (rc) return rc;
  std::vector<int> cls;
  for (size_t i = 0; i < p->classes.size(); i++)
    if (p->classes[i].ntiles > 0) cls.push_back((int)i);
  if (p->admit_pending) {
    if (hipStreamWaitEvent(p->cs, p->admit_ev, 0) != hipSuccess) return ECO_ERR_HIP;
    p->admit_pending = false;
  }
  if (!cls.empty()) {
    if (p->pc.use_graph) {
      if (!p->gexec || cls != p->graph_classes) {
        if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
        if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
        if (hipStreamBeginCapture(p->cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return ECO_ERR_HIP;
        // slots 0..K-1 are baked in; K % 6 == 0 keeps every replay consistent
        int rc2 = enqueue_steps(p, cls, 0);
        hipError_t e = hipStreamEndCapture(p->cs, &p->graph);
        if (rc2) return rc2;
        if (e != hipSuccess) return ECO_ERR_HIP;
        if (hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0) != hipSuccess) return ECO_ERR_HIP;
        p->graph_classes = cls;
        p->captures++;
      }
      if (hipGraphLaunch(p->gexec, p->cs) != hipSuccess) return ECO_ERR_HIP;
    } else {
      rc = enqueue_steps(p, cls, p->slot);
      if (rc) return rc;
    }
  }
  for (int c : cls) p->tiles_launched += (long long)p->classes[c].ntiles * p->K * (p->classes[c].key.order == 1 ? 1 : 2);
  p->slot += p->K;
  p->steps_launched += p->K;
  // control snapshot of this chunk
  Chunk ch;
  ch.seq = p->chunk_seq++;
  ch.ncases = p->hi_id;
  if (!p->snap_free.empty()) { ch.snap = p->snap_free.back(); p->snap_free.pop_back(); }
  else if (hipHostMalloc(&ch.snap, sizeof(Control) * (size_t)p->max_cases, 0) != hipSuccess) return ECO_ERR_NOMEM;
  if (hipEventCreateWithFlags(&ch.ev, hipEventDisableTiming) != hipSuccess) return ECO_ERR_HIP;
  hipError_t e = hipSuccess;
  if (ch.ncases) e = hipMemcpyAsync(ch.snap, p->d_ctl, sizeof(Control) * (size_t)ch.ncases, hipMemcpyDeviceToHost, p->cs);
  if (e == hipSuccess) e = hipEventRecord(ch.ev, p->cs);
  if (e != hipSuccess) return ECO_ERR_HIP;
  for (int c : cls)
    for (int id : p->classes[c].running) {
      Case &cs = p->cases[id];
      if (cs.first_chunk < 0) cs.first_chunk = ch.seq;
      cs.ref_chunk = ch.seq;
    }
  p->inflight.push_back(ch);
  return ECO_OK;
}

static void reclaim(eco_pool *p) {
  // cases released whose last referring chunk has completed
  for (int id = 0; id < p->hi_id; id++) {
    Case &cs = p->cases[id];
    if (cs.id < 0 || !cs.released) continue;
    if (cs.ref_chunk >= p->chunk_done) continue;
    p->arena.release(cs.off, cs.bytes);
    p->classes[cs.cls].free_argslots.push_back(cs.argslot);
    cs = Case();
    p->free_ids.push_back(id);
  }
}

extern "C" int eco_pool_poll(eco_pool *p, int keep, int *done_ids, int max_ids, int *ndone) {
  if (!p || keep < 0) return ECO_ERR_INVALID;
  while ((int)p->inflight.size() > keep) {
    Chunk ch = p->inflight.front();
    p->inflight.pop_front();
    if (hipEventSynchronize(ch.ev) != hipSuccess) return ECO_ERR_HIP;
    (void)hipEventDestroy(ch.ev);
    for (int id = 0; id < ch.ncases; id++) {
      Case &cs = p->cases[id];
      // only snapshots of chunks that ran the case describe it (a reused
      // handle's entry may still show the previous case in older ones)
      if (cs.id < 0 || !cs.running || cs.first_chunk < 0 || cs.first_chunk > ch.seq) continue;
      if (!ch.snap[id].finished) continue;
      cs.running = 0;
      cs.finished = 1;
      Class &cl = p->classes[cs.cls];
      cl.running.erase(std::remove(cl.running.begin(), cl.running.end(), id), cl.running.end());
      cl.dirty = true;
      p->ready.push_back(id);
    }
    p->snap_free.push_back(ch.snap);
    p->chunk_done = ch.seq + 1;
  }
  reclaim(p);
  int n = 0;
  while (!p->ready.empty() && n < max_ids) {
    if (done_ids) done_ids[n] = p->ready.front();
    p->ready.erase(p->ready.begin());
    n++;
  }
  if (ndone) *ndone = n;
  return ECO_OK;
}

// download a case's state on the upload stream (the case is not written by
// enqueued work once finished; for a running case the pool is synchronised)
static int fetch(eco_pool *p, Case &cs, double *host_state, eco_status *st) {
  if (!cs.finished) {
    if (hipStreamSynchronize(p->cs) != hipSuccess) return ECO_ERR_HIP;
  }
  Control h;
  hipError_t e = hipMemcpyAsync(&h, p->d_ctl + cs.id, sizeof(Control), hipMemcpyDeviceToHost, p->us);
  if (e == hipSuccess) e = hipStreamSynchronize(p->us);
  if (e != hipSuccess) return ECO_ERR_HIP;
  const eco_config &c = cs.cfg;
  const size_t nc = (size_t)cs.low.geo.ncell;
  const int par = (int)((cs.g0 + h.steps) & 1);  // buffer written by the case's last step
  if (st) {
    st->steps = h.steps;
    st->t = h.steps > 0 ? h.t_now : 0.;
    st->dt_last = h.dt_last;
    double dtn = 1.e-15;
    if (h.steps > 0) {
      unsigned long long bb = h.smax[(cs.g0 + h.steps - 1) % 3];
      double sm;
      std::memcpy(&sm, &bb, sizeof(sm));
      dtn = c.cfl * ((sm > 0.) ? std::min(1.e10, cs.low.geo.lcfl / sm) : 1.e10);
    }
    st->dt_next = dtn;
    st->flags = h.err;
    st->done = h.finished;
  }
  if (!host_state) return ECO_OK;
  const double *src = cs.buf[par];
  if (c.precision == ECO_FP32S) {
    const size_t stg = staging_bytes(c, cs.low);
    if (stg > p->fetch_cap) {
      if (p->fetch_tmp) (void)hipFree(p->fetch_tmp);
      p->fetch_tmp = nullptr;
      p->fetch_cap = 0;
      if (hipMalloc(&p->fetch_tmp, stg) != hipSuccess) return ECO_ERR_NOMEM;
      p->fetch_cap = stg;
    }
    StageArgs A = case_args(cs, p->d_ctl + cs.id, par, 0);
    A.wn = cs.buf[par];
    int rc = launch_convert(A, 0, p->fetch_tmp, p->us);
    if (rc) return rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;
    src = p->fetch_tmp;
  }
  std::vector<int> map;
  host_field_map(c, cs.low.geo.dim, map);
  const int nf = ECO_NFIELDS(c.nphase);
  std::memset(host_state, 0, sizeof(double) * nf * nc);
  for (size_t f = 0; f < map.size() && e == hipSuccess; f++)
    e = hipMemcpyAsync(host_state + (size_t)map[f] * nc, src + f * nc, nc * sizeof(double), hipMemcpyDeviceToHost, p->us);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_state + (size_t)(nf - 1) * nc, src + (size_t)cs.low.nprim * nc, nc * sizeof(double),
                       hipMemcpyDeviceToHost, p->us);
  if (e == hipSuccess) e = hipStreamSynchronize(p->us);
  if (e != hipSuccess) return ECO_ERR_HIP;
  if (c.model == ECO_MODEL_EULER)
    for (size_t i = 0; i < nc; i++) host_state[i] = 1.;
  return ECO_OK;
}

extern "C" int eco_pool_result(eco_pool *p, int id, double *host_state, eco_status *st) {
  if (!p || id < 0 || id >= p->max_cases || p->cases[id].id != id || p->cases[id].released) return ECO_ERR_INVALID;
  return fetch(p, p->cases[id], host_state, st);
}

extern "C" int eco_pool_release(eco_pool *p, int id) {
  if (!p || id < 0 || id >= p->max_cases || p->cases[id].id != id || p->cases[id].released) return ECO_ERR_INVALID;
  Case &cs = p->cases[id];
  if (cs.running) {
    // stop a running case: drop it from the tables
    Class &cl = p->classes[cs.cls];
    cl.running.erase(std::remove(cl.running.begin(), cl.running.end(), id), cl.running.end());
    cl.dirty = true;
    cs.running = 0;
  }
  cs.released = 1;
  reclaim(p);
  return ECO_OK;
}

extern "C" int eco_pool_stats_get(eco_pool *p, eco_pool_stats *s) {
  if (!p || !s) return ECO_ERR_INVALID;
  std::memset(s, 0, sizeof(*s));
  for (int id = 0; id < p->hi_id; id++) {
    const Case &cs = p->cases[id];
    if (cs.id < 0 || cs.released) continue;
    s->active++;
    s->running += cs.running;
  }
  s->bytes_used = p->arena.used();
  s->bytes_free = p->arena.largest();
  s->steps_launched = p->steps_launched;
  s->tiles_launched = p->tiles_launched;
  s->graph_captures = p->captures;
  return ECO_OK;
}

extern "C" void eco_pool_destroy(eco_pool *p) {
  if (!p) return;
  if (p->cs) (void)hipStreamSynchronize(p->cs);
  if (p->us) (void)hipStreamSynchronize(p->us);
  for (auto &ch : p->inflight) { (void)hipEventDestroy(ch.ev); p->snap_free.push_back(ch.snap); }
  for (auto *s : p->snap_free) (void)hipHostFree(s);
  for (auto &s : p->pinned) { if (s.p) (void)hipHostFree(s.p); if (s.ev) (void)hipEventDestroy(s.ev); }
  for (auto &cl : p->classes) {
    if (cl.d_args) (void)hipFree(cl.d_args);
    if (cl.h_args) (void)hipHostFree(cl.h_args);
    if (cl.d_tiles) (void)hipFree(cl.d_tiles);
    if (cl.d_desc) (void)hipFree(cl.d_desc);
  }
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_ctl) (void)hipFree(p->d_ctl);
  if (p->fetch_tmp) (void)hipFree(p->fetch_tmp);
  if (p->admit_ev) (void)hipEventDestroy(p->admit_ev);
  if (p->cs) (void)hipStreamDestroy(p->cs);
  if (p->us) (void)hipStreamDestroy(p->us);
  delete p;
}

// ---------------------------------------------------------------------------
// serving loop
// ---------------------------------------------------------------------------
extern "C" int eco_serve(eco_pool *p, const eco_request *req, int nreq, eco_request_result *res) {
  if (!p || (nreq > 0 && (!req || !res))) return ECO_ERR_INVALID;
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
  std::vector<int> order(nreq);
  for (int i = 0; i < nreq; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
  std::vector<int> owner(p->max_cases, -1);  // case handle -> request
  std::deque<int> waiting;
  size_t next = 0;
  int finished = 0, running = 0;
  std::vector<int> done(256);
  for (int i = 0; i < nreq; i++) { res[i] = eco_request_result{-1., -1., 0, 0., 0u}; }
  while (finished < nreq) {
    const double tn = now();
    while (next < order.size() && req[order[next]].arrival_s <= tn) waiting.push_back(order[next++]);
    // admission in arrival order while the budget allows (continuous batching)
    while (!waiting.empty()) {
      const int r = waiting.front();
      int id = -1;
      int rc = eco_pool_admit(p, req[r].cfg, req[r].state, req[r].nsteps, req[r].t_final, &id);
      if (rc == ECO_ERR_BUDGET || rc == ECO_ERR_FULL) break;
      if (rc) return rc;
      waiting.pop_front();
      owner[id] = r;
      res[r].admit_s = now();
      running++;
    }
    if (running == 0 && p->inflight.empty()) {
      if (!waiting.empty()) return ECO_ERR_BUDGET;  // a request larger than the whole budget
      if (next < order.size()) {
        const double wait = req[order[next]].arrival_s - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.002)));
      }
      continue;
    }
    if (running > 0) {
      int rc = eco_pool_advance(p);
      if (rc) return rc;
    }
    int nd = 0;
    int rc = eco_pool_poll(p, running > 0 ? 1 : 0, done.data(), (int)done.size(), &nd);
    if (rc) return rc;
    for (int j = 0; j < nd; j++) {
      const int id = done[j], r = owner[id];
      eco_status st;
      rc = fetch(p, p->cases[id], req[r].out_state, &st);
      if (rc) return rc;
      res[r].finish_s = now();
      res[r].steps = st.steps;
      res[r].t = st.t;
      res[r].flags = st.flags;
      owner[id] = -1;
      rc = eco_pool_release(p, id);
      if (rc) return rc;
      finished++;
      running--;
    }
  }
  // drain
  int nd = 0;
  return eco_pool_poll(p, 0, nullptr, 0, &nd);
}

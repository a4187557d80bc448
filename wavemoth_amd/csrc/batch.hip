// Batched synthesis: many independent a_lm -> map transforms of uneven
// (Nside, lmax, mmax, nmaps) in one execute.
//
// The reference runs one transform per plan (src/wavemoth.c:1271-1283) and
// spreads one transform's m values over a thread pool with a dynamic queue
// (legendre_transforms_thread, src/wavemoth.c:888-931).  Here the queue
// spans all requests: one persistent sweep kernel whose waves pull
// (request, m, ring chunk, map block) items, then one persistent ring
// kernel per LDS class pulling (request, ring pair) items.  Requests of
// equal (Nside, lmax, mmax) share their seeds, coefficient and ring tables.
// The launches are captured once into a HIP graph and replayed per execute;
// the per-execute input/output pointers live in a device array that the
// kernels read, so the graph stays valid when the caller's buffers change.
//
// Sweep shape per (Nside, lmax, mmax, nmaps): rings per lane RL and maps per
// pass KB.  Heuristic: KB = nmaps (<= 4) or 8, RL = 2 for nmaps <= 2 else 1.
// With WAVEMOTH_MEASURE each candidate shape the batched kernel supports is
// timed on the device and the fastest kept (process-wide wisdom table).
#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <tuple>

#include "../../include/wavemoth.h"
#include "wm_host.h"

using namespace wm;

namespace {

struct Geometry {
  int N = 0, L = 0, M = 0;
  LegTables base;  // tables of (N, L, M); no wave bounds, no per-map buffers
  LegCounts cnt;
  // ring tables for few maps ([0]) and for kRingWideMaps or more ([1])
  std::array<RingTables, 2> rt;
  std::array<RingHost, 2> rhost;
  std::array<bool, 2> have_rt{};
  std::array<LegTables, 5> waves;  // base + wave bounds, indexed by RL
  std::array<int64_t, 5> T_exec{};
  std::array<bool, 5> have{};
};

struct Shape {
  int RL = 1, KB = 1;
  // ring tables: 1 = wrapping rings in the register-FFT kernels (the
  // single plan's choice from kRingWideMaps maps up), 0 = in the general one
  int RW = 0;
};

std::mutex g_wisdom_mu;
std::map<std::tuple<int, int, int, int>, Shape> g_wisdom;  // (N, L, M, K) -> shape

int heuristic_kb(int K) { return legendre_map_block(K); }
int heuristic_rl(int K) { return legendre_default_rl(K); }
int heuristic_rw(int K) { return K >= kRingWideMaps ? 1 : 0; }

std::vector<Shape> candidate_shapes(int K) {
  std::vector<Shape> out;
  std::vector<int> kbs = {heuristic_kb(K)};
  if (K > 4) kbs.push_back(8);
  for (int KB : kbs)
    for (int RL : {1, 2, 4})
      if (batch_shape_supported(RL, KB)) out.push_back({RL, KB});
  return out;
}

int64_t nalm_of(int L, int M) {
  int64_t n = 0;
  for (int m = 0; m <= M; ++m) n += L - m + 1;
  return n;
}

void sweep_items(const LegTables &t, int req, std::vector<BatchLegItem> &out) {
  const int nkb = (t.K + t.KB - 1) / t.KB;
  for (int m = 0; m <= t.M; ++m)
    for (int kb = 0; kb < nkb; ++kb)
      for (int w = 0; w < t.W; ++w) out.push_back({req, m, w, kb});
}

// Register-FFT ring class (kind, FFT size P): r16items[first .. first + count)
struct R16Class {
  int kind, P, first, count;
};

// Ring-stage queues of a set of requests: the general kernel's (ring pair,
// map) items in classes c = 4 big + log2(size class / 1024) (big: LDS need
// above 80 KiB, one workgroup per CU; size class: ring_class, 1024 .. 8192),
// ritems[first[c] .. first[c] + count[c]) with slots[c] complex LDS slots,
// and the register-FFT launch blocks in one class per (kind, FFT size); one
// workgroup per item in both.
constexpr int kGenQueues = 8;
struct RingQueues {
  int first[kGenQueues] = {0}, count[kGenQueues] = {0}, slots[kGenQueues] = {0};
  std::vector<R16Class> r16;
  std::vector<BatchRingItem> ritems;
  std::vector<BatchRing16Item> r16items;
};

int gen_queue_maxp(int c) { return 1024 << (c & 3); }

}  // namespace

struct _wavemoth_batch {
  int nreq = 0;
  unsigned flags = 0;
  int device = 0;
  int cus = 0;
  hipStream_t stream = nullptr;
  std::vector<int> N, L, M, K, geom, measured;
  std::vector<Shape> shape;
  std::vector<std::unique_ptr<Geometry>> geoms;
  std::vector<LegTables> tabs;
  std::vector<int64_t> nalm, npix, exec_slots;
  DevArena arena;
  BatchDev dev;
  void **d_ptrs = nullptr;  // [2 nreq] inputs then outputs
  void **h_ptrs = nullptr;  // pinned staging of d_ptrs
  std::vector<const void *> cur_ptrs;
  hipEvent_t ptr_ev = nullptr;
  bool ptr_ev_live = false;
  // sweep classes (one per shape): litems[first .. first + count), costliest
  // class first; leg_blocks: the most blocks of any class's launch
  struct SweepClass {
    int RL, KB, first, count, blocks;
  };
  std::vector<SweepClass> sweep;
  int leg_blocks = 0;
  int nlitems = 0;
  RingQueues rq;  // ring-stage classes (their item lists live on the device)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // device staging of wavemoth_execute_batch, allocated on first use
  std::vector<double2 *> st_alm;
  std::vector<double *> st_map;
  int64_t sum_eps = 0, sum_dense = 0, sum_flops = 0, sum_exec = 0, maps = 0;
  // stage timing
  bool profiling = false;
  std::vector<hipEvent_t> ev_free;
  std::vector<std::array<hipEvent_t, 4>> ev_used;
};

namespace {

Geometry *get_geometry(wavemoth_batch b, int N, int L, int M) {
  for (auto &g : b->geoms)
    if (g->N == N && g->L == L && g->M == M) return g.get();
  auto g = std::make_unique<Geometry>();
  g->N = N;
  g->L = L;
  g->M = M;
  if (!build_leg_geometry(b->arena, g->base, N, L, M, b->stream, &g->cnt)) return nullptr;
  b->geoms.push_back(std::move(g));
  return b->geoms.back().get();
}

// Ring tables of a geometry in variant w (Shape::RW), built on first use.
bool ensure_rings(wavemoth_batch b, Geometry *g, int w) {
  if (g->have_rt[w]) return true;
  if (!build_ring_geometry(b->arena, g->rt[w], g->rhost[w], g->N, g->M, true, w == 1,
                           b->stream))
    return false;
  if (!g->rhost[w].lr.groups.empty()) {  // none up to Nside kMaxNside
    set_err("batch: rings longer than the register-FFT kernels hold");
    return false;
  }
  g->have_rt[w] = true;
  return true;
}

// Ring queues of requests r with ring tables rts[r] / hosts[r] and K[r]
// maps.  General-kernel pairs needing more than 80 KiB of LDS (one
// workgroup per CU) form class 0, the rest class 1, costliest first within
// a class; register-FFT blocks go one launch per (kind, FFT size), the
// largest requests first.
bool ring_queues(int cus, const std::vector<const RingTables *> &rts,
                 const std::vector<const RingHost *> &hosts, const std::vector<int> &K,
                 const std::vector<int> &Ms, RingQueues *q) {
  const int nreq = (int)rts.size();
  const int big_slots = (80 * 1024) / (int)sizeof(double2);
  struct RI {
    BatchRingItem it;
    double c;
    int need;
  };
  std::vector<RI> cls[kGenQueues];
  for (int r = 0; r < nreq; ++r) {
    const RingHost &h = *hosts[r];
    for (int p = 0; p < (int)h.pairs.size(); ++p) {
      if (h.kind[p] != kRingGeneral) continue;
      const int mp = ring_class(h.pairs[p].n, h.pairs[p].P);
      const int need = ring_lds_slots(h.pairs[p].n, h.pairs[p].P, Ms[r], mp);
      int lg = 0;
      while ((1024 << lg) < mp) ++lg;
      const int c = (need > big_slots ? 0 : 4) + lg;
      for (int k = 0; k < K[r]; ++k) cls[c].push_back({{r, p, k}, h.cost[p], need});
    }
  }
  std::map<std::pair<int, int>, std::vector<BatchRing16Item>> q16;
  std::vector<int> rord(nreq);
  std::iota(rord.begin(), rord.end(), 0);
  std::stable_sort(rord.begin(), rord.end(), [&](int x, int y) {
    return hosts[x]->npix * K[x] > hosts[y]->npix * K[y];
  });
  for (int r : rord) {
    const RingTables &rt = *rts[r];
    for (int g = 0; g < rt.nlaunch; ++g) {
      if (rt.kind[g] == kRingGeneral || rt.count[g] == 0) continue;
      const int64_t nb = ring16_blocks(rt.kind[g], rt.fftP[g], K[r], rt.count[g]);
      auto &v = q16[{rt.kind[g], rt.fftP[g]}];
      for (int64_t x = 0; x < nb; ++x) v.push_back({r, g, (int)x});
    }
  }
  q->r16.clear();
  q->r16items.clear();
  for (auto &kv : q16) {
    R16Class c;
    c.kind = kv.first.first;
    c.P = kv.first.second;
    c.first = (int)q->r16items.size();
    c.count = (int)kv.second.size();
    q->r16items.insert(q->r16items.end(), kv.second.begin(), kv.second.end());
    q->r16.push_back(c);
  }
  q->ritems.clear();
  for (int c = 0; c < kGenQueues; ++c) {
    std::stable_sort(cls[c].begin(), cls[c].end(),
                     [](const RI &x, const RI &y) { return x.c > y.c; });
    q->first[c] = (int)q->ritems.size();
    q->count[c] = (int)cls[c].size();
    q->slots[c] = 0;
    for (const RI &x : cls[c]) {
      q->ritems.push_back(x.it);
      q->slots[c] = std::max(q->slots[c], x.need);
    }
    if (q->count[c] > 0) {
      WM_TRY(batch_rings_set_lds_limit(q->slots[c]), "batch ring LDS limit");
    }
  }
  return true;
}

hipError_t launch_ring_queues(const BatchDev &d, const RingQueues &q, hipStream_t s) {
  for (int c = 0; c < kGenQueues; ++c) {
    const hipError_t e =
        launch_batch_rings(d, gen_queue_maxp(c), q.first[c], q.count[c], q.slots[c], s);
    if (e != hipSuccess) return e;
  }
  for (size_t c = 0; c < q.r16.size(); ++c) {
    const R16Class &x = q.r16[c];
    const hipError_t e = launch_batch_ring16(x.kind, x.P, d, x.first, x.count, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

bool ensure_waves(DevArena &a, Geometry *g, int RL, hipStream_t s) {
  if (g->have[RL]) return true;
  g->waves[RL] = g->base;
  if (!build_leg_waves(a, g->waves[RL], RL, s, &g->T_exec[RL])) return false;
  g->have[RL] = true;
  return true;
}

// Request tables with shape `sh`: the geometry's waves for sh.RL (kept in
// the batch arena `keep`, since the geometry caches them) plus per-request
// atil / qbuf allocated in `a`.
bool request_tables(DevArena &keep, DevArena &a, Geometry *g, int K, Shape sh, hipStream_t s,
                    LegTables *out) {
  if (!ensure_waves(keep, g, sh.RL, s)) return false;
  LegTables t = g->waves[sh.RL];
  t.K = K;
  t.KB = sh.KB;
  const int nkb = (K + sh.KB - 1) / sh.KB;
  if (!dalloc(a, &t.atil, (size_t)t.nlpad * nkb * sh.KB, "batch atil")) return false;
  if (!dalloc(a, &t.qbuf, (size_t)q_alloc_elems(t.R, K, t.M + 1), "batch qbuf")) return false;
  t.npix = 12 * (int64_t)g->N * g->N;
  *out = t;
  return true;
}

// Device arrays of a BatchDev over `tabs` with the given work lists; the
// pointer array (inputs then outputs) is allocated but not filled.
bool make_dev(DevArena &a, const std::vector<LegTables> &tabs,
              const std::vector<RingTables> &rtabs, const std::vector<BatchLegItem> &litems,
              const std::vector<BatchRingItem> &ritems,
              const std::vector<BatchRing16Item> &r16items, const double2 *twiddle, BatchDev *d,
              void ***d_ptrs) {
  const int nreq = (int)tabs.size();
  d->nreq = nreq;
  LegTables *dt = nullptr;
  RingTables *dr = nullptr;
  if (!upload(a, &dt, tabs, "batch tables")) return false;
  if (!upload(a, &dr, rtabs, "batch ring tables")) return false;
  d->tabs = dt;
  d->rtabs = dr;
  std::vector<int64_t> off(nreq + 1, 0);
  // prescale entries: (padded l-entry, map slot) pairs of each request
  for (int r = 0; r < nreq; ++r)
    off[r + 1] = off[r] + tabs[r].nlpad * (int64_t)((tabs[r].K + tabs[r].KB - 1) / tabs[r].KB *
                                                   tabs[r].KB);
  int64_t *doff = nullptr;
  if (!upload(a, &doff, off, "batch prep offsets")) return false;
  d->prep_off = doff;
  d->prep_total = off[nreq];
  BatchLegItem *dl = nullptr;
  if (!upload(a, &dl, litems, "sweep items")) return false;
  d->litems = dl;
  d->nlitems = (int)litems.size();
  BatchRingItem *dri = nullptr;
  if (!upload(a, &dri, ritems, "ring items")) return false;
  d->ritems = dri;
  BatchRing16Item *dr16 = nullptr;
  if (!upload(a, &dr16, r16items, "register-FFT ring items")) return false;
  d->r16items = dr16;
  d->twiddle = twiddle;
  if (!dalloc(a, &d->counters, kBatchCounters, "queue heads")) return false;
  void **p = nullptr;
  if (!dalloc(a, &p, (size_t)2 * nreq, "pointer array")) return false;
  d->alm = (const double2 *const *)p;
  d->map = (double *const *)(p + nreq);
  *d_ptrs = p;
  return true;
}

int sweep_blocks(int cus, int RL, int KB, int nitems) {
  int per_cu = 0;
  if (batch_legendre_occupancy(RL, KB, &per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
  const int need = (nitems + 3) / 4;
  return std::max(1, std::min(per_cu * cus, need));
}

// Time the sweep of one request of (g, K) with shape sh (prescale + sweep,
// best of 3 after a warm-up).  Returns milliseconds, or -1 on error.
double time_shape(wavemoth_batch b, Geometry *g, int K, Shape sh) {
  DevArena a;
  double best = -1.0;
  do {
    LegTables t;
    if (!request_tables(b->arena, a, g, K, sh, b->stream, &t)) break;
    double2 *alm = nullptr;
    const int64_t na = nalm_of(g->L, g->M) * K;
    if (!dalloc(a, &alm, (size_t)na, "tuning input")) break;
    if (hipMemsetAsync(alm, 0, (size_t)na * sizeof(double2), b->stream) != hipSuccess) break;
    std::vector<BatchLegItem> items;
    sweep_items(t, 0, items);
    BatchDev d;
    void **dp = nullptr;
    const int w = heuristic_rw(K);
    if (!ensure_rings(b, g, w)) break;
    if (!make_dev(a, {t}, {g->rt[w]}, items, {}, {}, g->rt[w].twiddle, &d, &dp)) break;
    std::vector<void *> ptrs = {(void *)alm, nullptr};
    if (hipMemcpy(dp, ptrs.data(), 2 * sizeof(void *), hipMemcpyHostToDevice) != hipSuccess)
      break;
    const int nb = sweep_blocks(b->cus, sh.RL, sh.KB, d.nlitems);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) break;
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      break;
    }
    bool ok = true;
    for (int rep = 0; rep < 4 && ok; ++rep) {
      ok = launch_batch_reset(d, b->stream) == hipSuccess &&
           launch_batch_prep(d, b->stream) == hipSuccess &&
           hipEventRecord(e0, b->stream) == hipSuccess &&
           launch_batch_legendre(d, sh.RL, sh.KB, 0, d.nlitems, nb, b->stream) == hipSuccess &&
           hipEventRecord(e1, b->stream) == hipSuccess &&
           hipEventSynchronize(e1) == hipSuccess;
      float ms = 0.f;
      if (ok && rep > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
        best = (best < 0 || ms < best) ? ms : best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) best = -1.0;
  } while (false);
  (void)hipStreamSynchronize(b->stream);
  a.release();
  return best;
}

// Time the ring stage of one request of (g, K) in ring variant w (q
// zeroed; best of 3 after a warm-up).  Returns milliseconds, or -1 on error.
double time_rings(wavemoth_batch b, Geometry *g, int K, int w) {
  DevArena a;
  double best = -1.0;
  do {
    if (!ensure_rings(b, g, w)) break;
    LegTables t;
    if (!request_tables(b->arena, a, g, K, {heuristic_rl(K), heuristic_kb(K), w}, b->stream,
                        &t))
      break;
    if (hipMemsetAsync(t.qbuf, 0, (size_t)q_alloc_elems(t.R, K, t.M + 1) * sizeof(double2),
                       b->stream) != hipSuccess)
      break;
    double *map = nullptr;
    if (!dalloc(a, &map, (size_t)g->rhost[w].npix * K, "tuning map")) break;
    RingQueues q;
    if (!ring_queues(b->cus, {&g->rt[w]}, {&g->rhost[w]}, {K}, {g->M}, &q)) break;
    BatchDev d;
    void **dp = nullptr;
    if (!make_dev(a, {t}, {g->rt[w]}, {}, q.ritems, q.r16items, g->rt[w].twiddle, &d, &dp))
      break;
    std::vector<void *> ptrs = {nullptr, (void *)map};
    if (hipMemcpy(dp, ptrs.data(), 2 * sizeof(void *), hipMemcpyHostToDevice) != hipSuccess)
      break;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) break;
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      break;
    }
    bool ok = true;
    for (int rep = 0; rep < 4 && ok; ++rep) {
      ok = launch_batch_reset(d, b->stream) == hipSuccess &&
           hipEventRecord(e0, b->stream) == hipSuccess &&
           launch_ring_queues(d, q, b->stream) == hipSuccess &&
           hipEventRecord(e1, b->stream) == hipSuccess &&
           hipEventSynchronize(e1) == hipSuccess;
      float ms = 0.f;
      if (ok && rep > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
        best = (best < 0 || ms < best) ? ms : best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) best = -1.0;
  } while (false);
  (void)hipStreamSynchronize(b->stream);
  a.release();
  return best;
}

// Whether the two ring variants of a geometry differ (some wrapping ring
// changes kernel); *ok false, with the error set, when building one failed.
bool ring_variants_differ(wavemoth_batch b, Geometry *g, bool *ok) {
  *ok = ensure_rings(b, g, 0) && ensure_rings(b, g, 1);
  return *ok && g->rhost[0].kind != g->rhost[1].kind;
}

bool choose_shape(wavemoth_batch b, Geometry *g, int K, Shape *out, int *measured) {
  const Shape h = {heuristic_rl(K), heuristic_kb(K), heuristic_rw(K)};
  *out = h;
  *measured = 0;
  if (!(b->flags & WAVEMOTH_MEASURE)) return true;
  const auto key = std::make_tuple(g->N, g->L, g->M, K);
  {
    std::lock_guard<std::mutex> lock(g_wisdom_mu);
    auto it = g_wisdom.find(key);
    if (it != g_wisdom.end()) {
      *out = it->second;
      *measured = 1;
      return true;
    }
  }
  double best = -1.0;
  for (Shape sh : candidate_shapes(K)) {
    const double ms = time_shape(b, g, K, sh);
    if (ms < 0) {
      set_err(std::string("autotuner: timing a sweep shape failed: ") + last_err());
      return false;
    }
    if (best < 0 || ms < best) {
      best = ms;
      *out = sh;
    }
  }
  // the ring stage: wrapping rings in the general or the register kernels
  out->RW = h.RW;
  bool built = true;
  if (ring_variants_differ(b, g, &built)) {
    double t[2];
    for (int w = 0; w < 2; ++w) {
      t[w] = time_rings(b, g, K, w);
      if (t[w] < 0) {
        set_err(std::string("autotuner: timing a ring stage failed: ") + last_err());
        return false;
      }
    }
    out->RW = t[1] < t[0] ? 1 : 0;
  }
  if (!built) return false;
  *measured = 1;
  std::lock_guard<std::mutex> lock(g_wisdom_mu);
  g_wisdom[key] = *out;
  return true;
}

hipEvent_t take_event(wavemoth_batch b) {
  if (!b->ev_free.empty()) {
    hipEvent_t e = b->ev_free.back();
    b->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

bool mark(wavemoth_batch b, int i, hipStream_t s) {
  if (!b->profiling) return true;
  if (i == 0) b->ev_used.push_back({nullptr, nullptr, nullptr, nullptr});
  hipEvent_t e = take_event(b);
  if (!e) {
    set_err("hipEventCreate failed");
    return false;
  }
  b->ev_used.back()[i] = e;
  WM_TRY(hipEventRecord(e, s), "hipEventRecord");
  return true;
}

void release_events(wavemoth_batch b) {
  for (auto &a : b->ev_used)
    for (hipEvent_t e : a)
      if (e) b->ev_free.push_back(e);
  b->ev_used.clear();
}

// The launches of one execute (also what the graph holds).
bool enqueue(wavemoth_batch b, hipStream_t s) {
  const BatchDev &d = b->dev;
  if (!mark(b, 0, s)) return false;
  WM_TRY(launch_batch_reset(d, s), "batch queue reset");
  WM_TRY(launch_batch_prep(d, s), "batch prescale");
  if (!mark(b, 1, s)) return false;
  for (const auto &c : b->sweep)
    WM_TRY(launch_batch_legendre(d, c.RL, c.KB, c.first, c.count, c.blocks, s), "batch sweep");
  if (!mark(b, 2, s)) return false;
  WM_TRY(launch_ring_queues(d, b->rq, s), "batch ring stage");
  if (!mark(b, 3, s)) return false;
  return true;
}

void free_batch(wavemoth_batch b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  release_events(b);
  for (hipEvent_t e : b->ev_free) (void)hipEventDestroy(e);
  if (b->ptr_ev) (void)hipEventDestroy(b->ptr_ev);
  if (b->h_ptrs) (void)hipHostFree(b->h_ptrs);
  b->arena.release();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

bool build_batch(wavemoth_batch b) {
  const int nreq = b->nreq;
  WM_TRY(hipGetDevice(&b->device), "hipGetDevice");
  WM_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device),
         "CU count");
  WM_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  b->geom.resize(nreq);
  b->shape.resize(nreq);
  b->measured.resize(nreq);
  b->tabs.resize(nreq);
  b->nalm.resize(nreq);
  b->npix.resize(nreq);
  b->exec_slots.resize(nreq);
  for (int r = 0; r < nreq; ++r) {
    Geometry *g = get_geometry(b, b->N[r], b->L[r], b->M[r]);
    if (!g) return false;
    b->geom[r] = -1;
    for (size_t i = 0; i < b->geoms.size(); ++i)
      if (b->geoms[i].get() == g) b->geom[r] = (int)i;
    b->nalm[r] = nalm_of(b->L[r], b->M[r]);
    b->npix[r] = 12 * (int64_t)b->N[r] * b->N[r];
  }
  for (int r = 0; r < nreq; ++r) {
    Geometry *g = b->geoms[b->geom[r]].get();
    if (!choose_shape(b, g, b->K[r], &b->shape[r], &b->measured[r])) return false;
    if (!ensure_rings(b, g, b->shape[r].RW)) return false;
    if (!request_tables(b->arena, b->arena, g, b->K[r], b->shape[r], b->stream, &b->tabs[r]))
      return false;
    const int nkb = (b->K[r] + b->shape[r].KB - 1) / b->shape[r].KB;
    b->exec_slots[r] = g->T_exec[b->shape[r].RL] * nkb;
    b->sum_eps += g->cnt.T_eps;
    b->sum_dense += g->cnt.T_dense;
    b->sum_flops += (5 + 4 * (int64_t)b->K[r]) * g->cnt.T_eps;
    b->sum_exec += b->exec_slots[r];
    b->maps += b->K[r];
  }

  // Sweep queue: requests by decreasing cost (executed slots times the
  // per-slot work), each request's items m-major.
  std::vector<int> ord(nreq);
  std::iota(ord.begin(), ord.end(), 0);
  auto cost = [&](int r) {
    return (double)b->exec_slots[r] * (5.0 + 4.0 * b->shape[r].KB);
  };
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost(x) > cost(y); });
  // one queue per shape, costliest shape first, requests in cost order
  // within each
  std::map<int, double> shape_cost;
  for (int r : ord) shape_cost[batch_shape_index(b->shape[r].RL, b->shape[r].KB)] += cost(r);
  std::vector<int> shapes;
  for (auto &kv : shape_cost) shapes.push_back(kv.first);
  std::stable_sort(shapes.begin(), shapes.end(),
                   [&](int x, int y) { return shape_cost[x] > shape_cost[y]; });
  std::vector<BatchLegItem> litems;
  b->sweep.clear();
  for (int si : shapes) {
    _wavemoth_batch::SweepClass c{0, 0, (int)std::min<size_t>(litems.size(), 0x7fffffff), 0, 0};
    for (int r : ord) {
      if (batch_shape_index(b->shape[r].RL, b->shape[r].KB) != si) continue;
      c.RL = b->shape[r].RL;
      c.KB = b->shape[r].KB;
      sweep_items(b->tabs[r], r, litems);
    }
    if (litems.size() > (size_t)0x7fffffff) {
      set_err("batch too large: more than 2^31 sweep items");
      return false;
    }
    c.count = (int)litems.size() - c.first;
    b->sweep.push_back(c);
  }
  // Ring queues over every request's ring tables
  std::vector<const RingTables *> rts(nreq);
  std::vector<const RingHost *> hosts(nreq);
  for (int r = 0; r < nreq; ++r) {
    const Geometry *g = b->geoms[b->geom[r]].get();
    rts[r] = &g->rt[b->shape[r].RW];
    hosts[r] = &g->rhost[b->shape[r].RW];
  }
  if (!ring_queues(b->cus, rts, hosts, b->K, b->M, &b->rq)) return false;
  b->nlitems = (int)litems.size();
  std::vector<RingTables> rtabs(nreq);
  for (int r = 0; r < nreq; ++r) rtabs[r] = *rts[r];
  if (!make_dev(b->arena, b->tabs, rtabs, litems, b->rq.ritems, b->rq.r16items,
                rtabs[0].twiddle, &b->dev, &b->d_ptrs))
    return false;
  b->leg_blocks = 0;
  for (auto &c : b->sweep) {
    c.blocks = sweep_blocks(b->cus, c.RL, c.KB, c.count);
    b->leg_blocks = std::max(b->leg_blocks, c.blocks);
  }
  WM_TRY(hipHostMalloc((void **)&b->h_ptrs, (size_t)2 * nreq * sizeof(void *),
                       hipHostMallocDefault),
         "pinned pointer staging");
  WM_TRY(hipEventCreateWithFlags(&b->ptr_ev, hipEventDisableTiming), "hipEventCreate");
  b->cur_ptrs.assign((size_t)2 * nreq, nullptr);
  WM_TRY(hipStreamSynchronize(b->stream), "batch tables");

  if (!(b->flags & WAVEMOTH_BATCH_NO_GRAPH)) {
    WM_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
           "graph capture begin");
    const bool ok = enqueue(b, b->stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (!ok) {
      if (g) (void)hipGraphDestroy(g);
      return false;
    }
    WM_TRY(e, "graph capture end");
    b->graph = g;
    WM_TRY(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0), "graph instantiate");
  }
  return true;
}

bool set_pointers(wavemoth_batch b, const void *const *in, const void *const *out,
                  hipStream_t s) {
  const int nreq = b->nreq;
  bool same = true;
  for (int r = 0; r < nreq && same; ++r)
    same = b->cur_ptrs[r] == in[r] && b->cur_ptrs[nreq + r] == out[r];
  if (same) return true;
  if (b->ptr_ev_live) WM_TRY(hipEventSynchronize(b->ptr_ev), "pointer staging wait");
  for (int r = 0; r < nreq; ++r) {
    b->h_ptrs[r] = const_cast<void *>(in[r]);
    b->h_ptrs[nreq + r] = const_cast<void *>(out[r]);
    b->cur_ptrs[r] = in[r];
    b->cur_ptrs[nreq + r] = out[r];
  }
  WM_TRY(hipMemcpyAsync(b->d_ptrs, b->h_ptrs, (size_t)2 * nreq * sizeof(void *),
                        hipMemcpyHostToDevice, s),
         "pointer upload");
  WM_TRY(hipEventRecord(b->ptr_ev, s), "hipEventRecord");
  b->ptr_ev_live = true;
  return true;
}

bool execute_device(wavemoth_batch b, const double *const *in, double *const *out,
                    hipStream_t s) {
  for (int r = 0; r < b->nreq; ++r) {
    if (!in[r] || !out[r]) {
      set_err("NULL buffer of request " + std::to_string(r));
      return false;
    }
    if (((uintptr_t)in[r] | (uintptr_t)out[r]) & 15) {
      set_err("device buffers must be 16-byte aligned (request " + std::to_string(r) + ")");
      return false;
    }
  }
  if (!set_pointers(b, (const void *const *)in, (const void *const *)out, s)) return false;
  if (b->gexec && !b->profiling) {
    WM_TRY(hipGraphLaunch(b->gexec, s), "graph launch");
    return true;
  }
  return enqueue(b, s);
}

}  // namespace

extern "C" {

wavemoth_batch wavemoth_plan_batch(int nreq, const int *Nside, const int *lmax,
                                   const int *mmax, const int *nmaps, unsigned flags) {
  clear_err();
  if (nreq < 1 || !Nside || !lmax || !mmax || !nmaps) {
    set_err("need nreq >= 1 and non-NULL request arrays");
    return nullptr;
  }
  for (int r = 0; r < nreq; ++r) {
    const std::string at = " (request " + std::to_string(r) + ")";
    if (Nside[r] < 1 || Nside[r] > kMaxNside) {
      set_err("Nside must be in [1, " + std::to_string(kMaxNside) + "]" + at);
      return nullptr;
    }
    if (lmax[r] < 0 || mmax[r] < 0 || mmax[r] > lmax[r]) {
      set_err("require 0 <= mmax <= lmax" + at);
      return nullptr;
    }
    if (nmaps[r] < 1) {
      set_err("nmaps must be >= 1" + at);
      return nullptr;
    }
    if ((int64_t)lmax[r] + kChunk > (1 << 28)) {
      set_err("lmax too large" + at);
      return nullptr;
    }
  }
  wavemoth_batch b = new _wavemoth_batch();
  b->nreq = nreq;
  b->flags = flags;
  b->N.assign(Nside, Nside + nreq);
  b->L.assign(lmax, lmax + nreq);
  b->M.assign(mmax, mmax + nreq);
  b->K.assign(nmaps, nmaps + nreq);
  if (!build_batch(b)) {
    const std::string e = last_err();
    free_batch(b);
    set_err(e);
    return nullptr;
  }
  return b;
}

void wavemoth_destroy_batch(wavemoth_batch b) { free_batch(b); }

int wavemoth_execute_batch_device(wavemoth_batch b, const double *const *d_inputs,
                                  double *const *d_outputs, void *stream) {
  clear_err();
  if (!b || !d_inputs || !d_outputs) {
    set_err("NULL argument");
    return -1;
  }
  return execute_device(b, d_inputs, d_outputs, (hipStream_t)stream) ? 0 : -1;
}

int wavemoth_execute_batch(wavemoth_batch b, const double *const *inputs,
                           double *const *outputs) {
  clear_err();
  if (!b || !inputs || !outputs) {
    set_err("NULL argument");
    return -1;
  }
  const int nreq = b->nreq;
  if (b->st_alm.empty()) {
    b->st_alm.assign(nreq, nullptr);
    b->st_map.assign(nreq, nullptr);
    for (int r = 0; r < nreq; ++r) {
      if (!dalloc(b->arena, &b->st_alm[r], (size_t)b->nalm[r] * b->K[r], "staging input") ||
          !dalloc(b->arena, &b->st_map[r], (size_t)b->npix[r] * b->K[r], "staging output")) {
        b->st_alm.clear();
        b->st_map.clear();
        return -1;
      }
    }
  }
  hipStream_t s = b->stream;
  for (int r = 0; r < nreq; ++r) {
    if (!inputs[r] || !outputs[r]) {
      set_err("NULL buffer of request " + std::to_string(r));
      return -1;
    }
    if (hipMemcpyAsync(b->st_alm[r], inputs[r], (size_t)b->nalm[r] * b->K[r] * sizeof(double2),
                       hipMemcpyHostToDevice, s) != hipSuccess) {
      set_err("input upload failed");
      return -1;
    }
  }
  std::vector<const double *> din(nreq);
  std::vector<double *> dout(nreq);
  for (int r = 0; r < nreq; ++r) {
    din[r] = (const double *)b->st_alm[r];
    dout[r] = b->st_map[r];
  }
  if (!execute_device(b, din.data(), dout.data(), s)) return -1;
  for (int r = 0; r < nreq; ++r)
    if (hipMemcpyAsync(outputs[r], b->st_map[r], (size_t)b->npix[r] * b->K[r] * sizeof(double),
                       hipMemcpyDeviceToHost, s) != hipSuccess) {
      set_err("output download failed");
      return -1;
    }
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_err("batch execute failed");
    return -1;
  }
  return 0;
}

int wavemoth_batch_set_profiling(wavemoth_batch b, int on) {
  clear_err();
  if (!b) return -1;
  (void)hipStreamSynchronize(b->stream);
  release_events(b);
  b->profiling = on != 0;
  return 0;
}

int wavemoth_batch_stage_times(wavemoth_batch b, double *avg_ms, int *count) {
  clear_err();
  if (!b || !avg_ms || !count) return -1;
  double acc[3] = {0, 0, 0};
  int n = 0;
  for (auto &a : b->ev_used) {
    if (!a[0] || !a[1] || !a[2] || !a[3]) continue;
    if (hipEventSynchronize(a[3]) != hipSuccess) {
      set_err("hipEventSynchronize failed");
      return -1;
    }
    for (int i = 0; i < 3; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, a[i], a[i + 1]) != hipSuccess) {
        set_err("hipEventElapsedTime failed");
        return -1;
      }
      acc[i] += ms;
    }
    ++n;
  }
  for (int i = 0; i < 3; ++i) avg_ms[i] = n ? acc[i] / n : 0.0;
  *count = n;
  release_events(b);
  return 0;
}

int64_t wavemoth_batch_info(wavemoth_batch b, int req, int key) {
  if (!b) return -1;
  if (req < 0) {
    switch (key) {
      case 0: return b->nreq;
      case 1: return b->maps;
      case 2: return b->sum_eps;
      case 3: return b->sum_dense;
      case 4: return b->sum_flops;
      case 5: return (int64_t)b->arena.bytes;
      case 6: return b->nlitems;
      case 7: return (int64_t)b->rq.ritems.size() + (int64_t)b->rq.r16items.size();
      case 8: return b->gexec ? 1 : 0;
      case 9: return b->sum_exec;
      case 10: return b->leg_blocks;
      default: return -1;
    }
  }
  if (req >= b->nreq) return -1;
  const Geometry *g = b->geoms[b->geom[req]].get();
  switch (key) {
    case 0: return b->N[req];
    case 1: return b->L[req];
    case 2: return b->M[req];
    case 3: return b->K[req];
    case 4: return b->shape[req].RL;
    case 5: return b->shape[req].KB;
    case 6: return g->cnt.T_eps;
    case 7: return g->cnt.T_dense;
    case 8: return b->npix[req];
    case 9: return b->exec_slots[req];
    case 10: return b->measured[req];
    case 11: return b->shape[req].RW;
    default: return -1;
  }
}

}  // extern "C"

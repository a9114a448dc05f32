// hyb_server.cpp -- request loop over resident domains: continuous batching
// of hybgen steps (hycom_hybgen_server_* in include/hycom_hybgen.h).
//
// One scheduler thread per server owns the device work.  An iteration
// steps every active request once: the descriptors of the active domains
// (KParams, View, Scratch, and the block layout of each launch) are written
// to a pinned staging buffer, copied to the device in one transfer and
// consumed by the batched kernels (launch_batch), all on the server's
// stream.  Two iterations are kept in flight so that the host prepares the
// next one while the device runs the current one; a request completes when
// the event after its last iteration has fired.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/hycom_hybgen.h"
#include "hyb_host.h"
#include "hyb_kernels.h"
#include "hyb_params_conv.h"

using namespace hyb;

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// byte offsets of the fields of one resident domain in its device block
struct Layout {
  size_t dp, dpu, dpv, temp, saln, th3d, u, v, tracer, p, pu, pv, theta;
  size_t depths, topiso, depthu, depthv, ubavg, vbavg, ip, iu, iv;
  size_t state;  // bytes of the fields: what an eviction saves
  size_t pres, lcm, kp, uvf;
  size_t total;
  size_t n2, n3, n3p, n4, n5;
};

Layout layout_of(const hyb_params& P) {
  Layout L;
  const size_t NI = P.idm + 2 * P.nbdy, NJ = P.jdm + 2 * P.nbdy, K = P.kdm;
  L.n2 = NI * NJ;
  L.n3 = L.n2 * K;
  L.n3p = L.n2 * (K + 1);
  L.n4 = 2 * L.n3;
  L.n5 = P.ntracr > 0 ? L.n4 * P.ntracr : 0;
  size_t o = 0;
  auto put = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes);
    return at;
  };
  L.dp = put(8 * L.n4);
  L.dpu = put(8 * L.n4);
  L.dpv = put(8 * L.n4);
  L.temp = put(8 * L.n4);
  L.saln = put(8 * L.n4);
  L.th3d = put(8 * L.n4);
  L.u = put(8 * L.n4);
  L.v = put(8 * L.n4);
  L.tracer = put(8 * L.n5);
  L.p = put(8 * L.n3p);
  L.pu = put(8 * L.n3p);
  L.pv = put(8 * L.n3p);
  L.theta = put(8 * L.n3);
  L.depths = put(8 * L.n2);
  L.topiso = put(8 * L.n2);
  L.depthu = put(8 * L.n2);
  L.depthv = put(8 * L.n2);
  L.ubavg = put(24 * L.n2);
  L.vbavg = put(24 * L.n2);
  L.ip = put(4 * L.n2);
  L.iu = put(4 * L.n2);
  L.iv = put(4 * L.n2);
  L.state = o;
  L.pres = put(8 * L.n3p);
  L.lcm = put(8 * L.n2);
  L.kp = put(4 * L.n2);
  L.uvf = put(8 * L.n2);
  L.total = o;
  return L;
}

// the fields of a block laid out by L (device or host memory)
hyb_fields fields_at(unsigned char* b, const Layout& L, int ntracr) {
  hyb_fields F;
  F.dp = (double*)(b + L.dp);
  F.dpo = nullptr;
  F.dpu = (double*)(b + L.dpu);
  F.dpv = (double*)(b + L.dpv);
  F.temp = (double*)(b + L.temp);
  F.saln = (double*)(b + L.saln);
  F.th3d = (double*)(b + L.th3d);
  F.u = (double*)(b + L.u);
  F.v = (double*)(b + L.v);
  F.tracer = ntracr > 0 ? (double*)(b + L.tracer) : nullptr;
  F.p = (double*)(b + L.p);
  F.pu = (double*)(b + L.pu);
  F.pv = (double*)(b + L.pv);
  F.theta = (const double*)(b + L.theta);
  F.depths = (const double*)(b + L.depths);
  F.topiso = (const double*)(b + L.topiso);
  F.depthu = (const double*)(b + L.depthu);
  F.depthv = (const double*)(b + L.depthv);
  F.ubavg = (const double*)(b + L.ubavg);
  F.vbavg = (const double*)(b + L.vbavg);
  F.ip = (const int32_t*)(b + L.ip);
  F.iu = (const int32_t*)(b + L.iu);
  F.iv = (const int32_t*)(b + L.iv);
  return F;
}

// (offset of the pointer in hyb_fields, offset in the block, bytes) of
// every field; every member of hyb_fields is a data pointer
struct FieldRef {
  size_t member, off, bytes;
};

std::vector<FieldRef> field_refs(const Layout& L) {
  const size_t b4 = 8 * L.n4, bp = 8 * L.n3p, b2 = 8 * L.n2, i2 = 4 * L.n2;
#define HYB_FR(name, bytes) FieldRef{offsetof(hyb_fields, name), L.name, bytes}
  return {HYB_FR(dp, b4),         HYB_FR(dpu, b4),         HYB_FR(dpv, b4),
          HYB_FR(temp, b4),       HYB_FR(saln, b4),        HYB_FR(th3d, b4),
          HYB_FR(u, b4),          HYB_FR(v, b4),           HYB_FR(tracer, 8 * L.n5),
          HYB_FR(p, bp),          HYB_FR(pu, bp),          HYB_FR(pv, bp),
          HYB_FR(theta, 8 * L.n3), HYB_FR(depths, b2),     HYB_FR(topiso, b2),
          HYB_FR(depthu, b2),     HYB_FR(depthv, b2),      HYB_FR(ubavg, 3 * b2),
          HYB_FR(vbavg, 3 * b2),  HYB_FR(ip, i2),          HYB_FR(iu, i2),
          HYB_FR(iv, i2)};
#undef HYB_FR
}

void* field_ptr(const hyb_fields* F, const FieldRef& r) {
  void* q;
  memcpy(&q, (const unsigned char*)F + r.member, sizeof q);
  return q;
}

struct Domain {
  hyb_params P;
  KParams k[2];  // per precision
  Layout L;
  hyb_synth S;
  bool synth = false;
  std::vector<unsigned char> host;  // state while not resident (L.state bytes)
  bool host_valid = false;
  unsigned char* dev = nullptr;
  View V;
  Scratch W;
  int64_t sea = -1;  // sea columns, known after the first open
  bool resident = false;
  bool active = false;    // a request of this domain is in the active set
  bool pinned = false;    // being read by the caller (not evictable)
  bool evicting = false;  // being evicted by the loop
  int queued = 0;         // requests submitted and not completed
  uint64_t lru = 0;
};

struct Req {
  Domain* d;
  int64_t tag;
  int32_t dom, nsteps, prec, left;
  double t_submit, t_start;
};

struct Inflight {
  int slot;  // staging buffer and event
  std::vector<Req> finished;
};

constexpr int kRing = 3;
constexpr int kMaxActive = 256;

}  // namespace

struct hyb_server {
  hyb_server_config cfg{};
  int device = 0;
  std::chrono::steady_clock::time_point t0;
  // shared with the caller's threads (guarded by mu)
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::vector<std::unique_ptr<Domain>> doms;
  std::deque<Req> incoming;
  std::deque<hyb_completion> done;
  bool stop = false;
  int err = 0;
  char errmsg[512] = "";
  hyb_server_stats st{};
  // owned by the scheduler thread
  std::thread th;
  std::deque<Req> waiting;
  std::vector<Req> active;
  std::deque<Inflight> inflight;
  hipStream_t stream = nullptr;
  unsigned char* stage[kRing] = {};
  hipEvent_t ev[kRing] = {};
  bool ev_used[kRing] = {};
  unsigned char* ddesc = nullptr;
  size_t desc_cap = 0;
  int64_t iter = 0;
  uint64_t clock = 0;
  size_t used = 0, budget = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

size_t desc_capacity(int na) {
  return align_up(na * sizeof(KParams)) + align_up(na * sizeof(View)) +
         align_up(na * sizeof(Scratch)) + sizeof(int32_t) * (16 * (size_t)na + 64);
}

// records a loop failure once, for every later call to report
void set_error(hyb_server* s, int code) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->err) {
    s->err = code;
    snprintf(s->errmsg, sizeof s->errmsg, "%s", last_error());
  }
}

void complete(hyb_server* s, const Req& r, int status, double t_done) {
  Domain& d = *r.d;
  hyb_completion c;
  c.tag = r.tag;
  c.domain = r.dom;
  c.status = status;
  c.nsteps = r.nsteps;
  c.precision = r.prec;
  c.t_submit = r.t_submit;
  c.t_start = r.t_start;
  c.t_done = t_done;
  c.column_layers = status == 0 ? (int64_t)r.nsteps * d.sea * d.P.kdm : 0;
  std::lock_guard<std::mutex> lk(s->mu);
  d.queued--;
  if (status == 0) {
    s->st.completed++;
    s->st.column_layers += c.column_layers;
  }
  s->done.push_back(c);
  s->cv_done.notify_all();
}

int evict(hyb_server* s, Domain& e) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (e.pinned) return 1;
    e.evicting = true;
  }
  e.host.resize(e.L.state);
  int rc = 0;
  hipError_t h = hipMemcpyAsync(e.host.data(), e.dev, e.L.state, hipMemcpyDeviceToHost, s->stream);
  if (h == hipSuccess) h = hipStreamSynchronize(s->stream);
  if (h == hipSuccess) h = hipFree(e.dev);
  if (h != hipSuccess) rc = fail(HYB_EDEVICE, "evicting a domain: %s", hipGetErrorString(h));
  std::lock_guard<std::mutex> lk(s->mu);
  e.evicting = false;
  if (rc) return rc;
  e.dev = nullptr;
  e.resident = false;
  e.host_valid = true;
  s->used -= e.L.total;
  s->st.evictions++;
  s->st.resident_domains--;
  s->st.resident_bytes = (int64_t)s->used;
  return 0;
}

// 0 opened, 1 not now (memory held by running work), <0 error
int open_domain(hyb_server* s, Domain& d) {
  const size_t need = d.L.total;
  if (need > s->budget)
    return fail(HYB_ENOMEM, "domain needs %zu bytes, over the budget of %zu", need, s->budget);
  while (s->used + need > s->budget) {
    Domain* victim = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    for (auto& up : s->doms) {
      Domain& e = *up;
      if (!e.resident || e.active || e.pinned || &e == &d) continue;
      // idle first (no queued request), then least recently used
      if (!victim || (e.queued == 0 && victim->queued > 0) ||
          ((e.queued == 0) == (victim->queued == 0) && e.lru < victim->lru))
        victim = &e;
    }
    lk.unlock();
    if (!victim) return 1;
    int rc = evict(s, *victim);
    if (rc < 0) return rc;
    if (rc > 0) return 1;
  }
  hipError_t h = hipMalloc(&d.dev, need);
  if (h == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    return 1;
  }
  HIPCHK(h);
  hyb_fields F = fields_at(d.dev, d.L, d.P.ntracr);
  if (d.host_valid) {
    HIPCHK(hipMemcpyAsync(d.dev, d.host.data(), d.L.state, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    std::vector<unsigned char>().swap(d.host);
    d.host_valid = false;
  } else {
    int rc = hycom_hybgen_synth_device(&d.P, &d.S, &F, 1, 2, s->stream);
    if (rc) return rc;
  }
  if (d.sea < 0) {
    std::vector<int32_t> ip(d.L.n2);
    HIPCHK(hipMemcpyAsync(ip.data(), F.ip, 4 * d.L.n2, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int64_t NI = d.P.idm + 2 * d.P.nbdy;
    int64_t n = 0;
    for (int j = 0; j < d.P.jj; j++)
      for (int i = 0; i < d.P.ii; i++) n += ip[(j + d.P.nbdy) * NI + i + d.P.nbdy] > 0;
    d.sea = n;
  }
  d.V = view_of(&d.P, &F);
  d.W = Scratch{(double*)(d.dev + d.L.pres), (int32_t*)(d.dev + d.L.kp),
                (uint64_t*)(d.dev + d.L.lcm), (int32_t*)(d.dev + d.L.uvf)};
  std::lock_guard<std::mutex> lk(s->mu);
  d.resident = true;
  s->used += need;
  s->st.opens++;
  s->st.resident_domains++;
  s->st.resident_bytes = (int64_t)s->used;
  return 0;
}

// moves waiting requests into the active set while slots and memory allow
int admit(hyb_server* s) {
  for (auto it = s->waiting.begin(); it != s->waiting.end();) {
    if ((int)s->active.size() >= s->cfg.max_active) break;
    Domain& d = *it->d;
    bool blocked = d.active;
    if (!blocked) {
      std::lock_guard<std::mutex> lk(s->mu);
      blocked = d.pinned;
    }
    if (blocked) {
      ++it;
      continue;
    }
    if (!d.resident) {
      const int rc = open_domain(s, d);
      if (rc < 0) {
        if (rc == HYB_ENOMEM && d.L.total > s->budget) {  // can never fit: fail the request
          complete(s, *it, rc, s->now());
          it = s->waiting.erase(it);
          continue;
        }
        return rc;
      }
      if (rc > 0) {
        ++it;
        continue;
      }
    }
    d.active = true;
    d.lru = ++s->clock;
    Req r = *it;
    r.t_start = s->now();
    s->active.push_back(r);
    it = s->waiting.erase(it);
  }
  return 0;
}

int launch_iteration(hyb_server* s) {
  const int na = (int)s->active.size();
  const int slot = (int)(s->iter % kRing);
  if (s->ev_used[slot]) HIPCHK(hipEventSynchronize(s->ev[slot]));
  unsigned char* h = s->stage[slot];
  const size_t oV = align_up(na * sizeof(KParams));
  const size_t oW = oV + align_up(na * sizeof(View));
  const size_t oI = oW + align_up(na * sizeof(Scratch));
  KParams* hp = (KParams*)h;
  View* hv = (View*)(h + oV);
  Scratch* hw = (Scratch*)(h + oW);
  for (int q = 0; q < na; q++) {
    const Domain& d = *s->active[q].d;
    hp[q] = d.k[s->active[q].prec];
    hv[q] = d.V;
    hw[q] = d.W;
  }
  BatchLaunch L{};
  L.remap_cols = s->cfg.remap_cols;
  L.regrid_cols = s->cfg.regrid_cols;
  std::vector<int32_t> I;
  batch_layout(hp, na, L, I);
  const size_t bytes = oI + sizeof(int32_t) * I.size();
  if (bytes > s->desc_cap) return fail(HYB_EDEVICE, "iteration descriptor of %zu bytes", bytes);
  memcpy(h + oI, I.data(), sizeof(int32_t) * I.size());
  HIPCHK(hipMemcpyAsync(s->ddesc, h, bytes, hipMemcpyHostToDevice, s->stream));
  batch_bind(L, (const int32_t*)(s->ddesc + oI));
  const BatchDevice B{(const KParams*)s->ddesc, (const View*)(s->ddesc + oV),
                      (const Scratch*)(s->ddesc + oW)};
  HIPCHK(launch_batch(B, L, 2, s->stream));
  HIPCHK(hipEventRecord(s->ev[slot], s->stream));
  s->ev_used[slot] = true;
  Inflight f;
  f.slot = slot;
  std::vector<Req> keep;
  keep.reserve(na);
  for (Req& r : s->active) {
    Domain& d = *r.d;
    d.lru = ++s->clock;
    if (--r.left == 0) {
      f.finished.push_back(r);
      d.active = false;  // its next request may join the next iteration
    } else {
      keep.push_back(r);
    }
  }
  s->active.swap(keep);
  s->inflight.push_back(std::move(f));
  s->iter++;
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.iterations++;
  s->st.active_sum += na;
  if (na > s->st.max_active_seen) s->st.max_active_seen = na;
  return 0;
}

int complete_oldest(hyb_server* s) {
  Inflight f = std::move(s->inflight.front());
  s->inflight.pop_front();
  HIPCHK(hipEventSynchronize(s->ev[f.slot]));
  const double t = s->now();
  for (const Req& r : f.finished) complete(s, r, 0, t);
  return 0;
}

void cancel_all(hyb_server* s, int status) {
  const double t = s->now();
  for (Inflight& f : s->inflight)
    for (const Req& r : f.finished) complete(s, r, status, t);
  s->inflight.clear();
  for (const Req& r : s->active) {
    r.d->active = false;
    complete(s, r, status, t);
  }
  s->active.clear();
  for (const Req& r : s->waiting) complete(s, r, status, t);
  s->waiting.clear();
}

void loop(hyb_server* s) {
  (void)hipSetDevice(s->device);
  int rc = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->active.empty() && s->inflight.empty())
        s->cv.wait(lk, [&] { return s->stop || !s->incoming.empty() || !s->waiting.empty(); });
      if (s->stop) break;
      while (!s->incoming.empty()) {
        s->waiting.push_back(s->incoming.front());
        s->incoming.pop_front();
      }
    }
    if ((rc = admit(s))) break;
    if (!s->active.empty()) {
      if ((rc = launch_iteration(s))) break;
    } else if (s->inflight.empty() && !s->waiting.empty()) {
      // nothing runs and nothing could be admitted: the memory is held by
      // the caller (pinned domains) or by other processes; retry shortly
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait_for(lk, std::chrono::milliseconds(1));
      continue;
    }
    while (!s->inflight.empty() && (s->inflight.size() >= 2 || s->active.empty())) {
      if ((rc = complete_oldest(s))) break;
      // admit arrivals before draining further
      if (!s->active.empty()) break;
      std::lock_guard<std::mutex> lk(s->mu);
      if (!s->incoming.empty()) break;
    }
    if (rc) break;
  }
  if (rc) set_error(s, rc);
  (void)hipStreamSynchronize(s->stream);
  cancel_all(s, rc ? rc : HYB_EBUSY);
  std::lock_guard<std::mutex> lk(s->mu);
  for (const Req& r : s->incoming) {
    hyb_completion c{};
    c.tag = r.tag;
    c.domain = r.dom;
    c.status = rc ? rc : HYB_EBUSY;
    r.d->queued--;
    s->done.push_back(c);
  }
  s->incoming.clear();
  s->cv_done.notify_all();
}

int server_error(hyb_server* s) {
  // with s->mu held
  return s->err ? fail(s->err, "server failed: %s", s->errmsg) : 0;
}

}  // namespace

extern "C" {

int hycom_hybgen_server_create(const hyb_server_config* C, hyb_server** out) {
  if (!out) return fail(HYB_ENULL, "out is NULL");
  *out = nullptr;
  hyb_server_config c{};
  if (C) c = *C;
  if (c.max_active == 0) c.max_active = 64;
  if (c.remap_cols == 0) c.remap_cols = 64;
  if (c.regrid_cols == 0) c.regrid_cols = 256;
  if (c.max_active < 1 || c.max_active > kMaxActive)
    return fail(HYB_EINVAL, "max_active=%d outside 1..%d", c.max_active, kMaxActive);
  auto width_ok = [](int w) { return w == 64 || w == 128 || w == 256; };
  if (!width_ok(c.remap_cols) || !width_ok(c.regrid_cols))
    return fail(HYB_EINVAL, "workgroup widths must be 64, 128 or 256");
  if (c.hbm_budget < 0) return fail(HYB_EINVAL, "hbm_budget < 0");
  int rc;
  if ((rc = device_ready())) return rc;
  std::unique_ptr<hyb_server> s(new hyb_server);
  HIPCHK(hipGetDevice(&s->device));
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  s->budget = c.hbm_budget ? (size_t)c.hbm_budget : (size_t)(0.9 * (double)fr);
  s->cfg = c;
  s->st.budget_bytes = (int64_t)s->budget;
  HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  s->desc_cap = desc_capacity(c.max_active);
  for (int q = 0; q < kRing; q++) {
    HIPCHK(hipHostMalloc((void**)&s->stage[q], s->desc_cap, hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&s->ev[q], hipEventDisableTiming));
  }
  HIPCHK(hipMalloc(&s->ddesc, s->desc_cap));
  s->t0 = std::chrono::steady_clock::now();
  s->th = std::thread(loop, s.get());
  *out = s.release();
  return 0;
}

int hycom_hybgen_server_destroy(hyb_server* s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  (void)hipStreamSynchronize(s->stream);
  for (auto& d : s->doms)
    if (d->dev) (void)hipFree(d->dev);
  for (int q = 0; q < kRing; q++) {
    if (s->stage[q]) (void)hipHostFree(s->stage[q]);
    if (s->ev[q]) (void)hipEventDestroy(s->ev[q]);
  }
  if (s->ddesc) (void)hipFree(s->ddesc);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

int hycom_hybgen_server_add_domain(hyb_server* s, const hyb_params* P, const hyb_synth* S,
                                   const hyb_fields* F, int32_t* domain_id) {
  if (!s || !P || !domain_id) return fail(HYB_ENULL, "NULL argument");
  if (!S && !F) return fail(HYB_ENULL, "a domain needs a synthetic description or host fields");
  int rc;
  if ((rc = hycom_hybgen_check_params(P))) return rc;
  if (F && (rc = check_fields(P, F, 0))) return rc;
  std::unique_ptr<Domain> d(new Domain);
  d->P = *P;
  d->P.precision = HYB_PREC_FP64;
  hyb_params pm = d->P;
  pm.precision = HYB_PREC_MIXED;
  d->k[0] = kparams_of(&d->P);
  d->k[1] = kparams_of(&pm);
  if (!stream_path_ok(d->k[0]))
    return fail(HYB_EUNSUPPORTED, "served domains need nhybrd==kdm");
  if (P->itracr801 || P->itracr802 || P->itracr803)
    return fail(HYB_EUNSUPPORTED, "no itracr diagnostics in served domains");
  d->L = layout_of(d->P);
  if (S) {
    d->S = *S;
    d->synth = true;
  } else {
    d->host.assign(d->L.state, 0);
    for (const FieldRef& r : field_refs(d->L)) {
      const void* src = field_ptr(F, r);
      if (src && r.bytes) memcpy(d->host.data() + r.off, src, r.bytes);
    }
    d->host_valid = true;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if ((rc = server_error(s))) return rc;
  *domain_id = (int32_t)s->doms.size();
  s->doms.push_back(std::move(d));
  return 0;
}

int hycom_hybgen_server_submit(hyb_server* s, int32_t domain_id, int32_t nsteps,
                               int32_t precision, int64_t tag) {
  if (!s) return fail(HYB_ENULL, "server is NULL");
  if (nsteps < 1 || nsteps > 100000) return fail(HYB_EINVAL, "nsteps=%d", nsteps);
  if (precision != HYB_PREC_FP64 && precision != HYB_PREC_MIXED)
    return fail(HYB_EINVAL, "precision=%d", precision);
  std::lock_guard<std::mutex> lk(s->mu);
  int rc;
  if ((rc = server_error(s))) return rc;
  if (s->stop) return fail(HYB_EBUSY, "server is stopping");
  if (domain_id < 0 || domain_id >= (int32_t)s->doms.size())
    return fail(HYB_EINVAL, "no domain %d", domain_id);
  Req r;
  r.d = s->doms[domain_id].get();
  r.tag = tag;
  r.dom = domain_id;
  r.nsteps = nsteps;
  r.prec = precision;
  r.left = nsteps;
  r.t_submit = s->now();
  r.t_start = 0.0;
  s->doms[domain_id]->queued++;
  s->incoming.push_back(r);
  s->cv.notify_all();
  return 0;
}

int hycom_hybgen_server_poll(hyb_server* s, hyb_completion* out, int32_t max, int32_t wait_ms,
                             int32_t* n) {
  if (!s || !n || (max > 0 && !out)) return fail(HYB_ENULL, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms),
                        [&] { return !s->done.empty() || s->err; });
  int32_t k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n = k;
  return server_error(s);
}

int hycom_hybgen_server_read_domain(hyb_server* s, int32_t domain_id, const hyb_fields* F) {
  if (!s || !F) return fail(HYB_ENULL, "NULL argument");
  Domain* d;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (domain_id < 0 || domain_id >= (int32_t)s->doms.size())
      return fail(HYB_EINVAL, "no domain %d", domain_id);
    d = s->doms[domain_id].get();
    if (d->queued > 0 || d->evicting)
      return fail(HYB_EBUSY, "domain %d has requests queued or running", domain_id);
    if (!d->resident && !d->host_valid)
      return fail(HYB_EINVAL, "domain %d has not been stepped yet", domain_id);
    d->pinned = true;
  }
  int rc = 0;
  for (const FieldRef& r : field_refs(d->L)) {
    void* dst = field_ptr(F, r);
    if (!dst || !r.bytes) continue;
    if (d->resident) {
      hipError_t h = hipMemcpy(dst, d->dev + r.off, r.bytes, hipMemcpyDeviceToHost);
      if (h != hipSuccess) {
        rc = fail(HYB_EDEVICE, "reading domain %d: %s", domain_id, hipGetErrorString(h));
        break;
      }
    } else {
      memcpy(dst, d->host.data() + r.off, r.bytes);
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  d->pinned = false;
  s->cv.notify_all();
  return rc;
}

int hycom_hybgen_server_stats(hyb_server* s, hyb_server_stats* st) {
  if (!s || !st) return fail(HYB_ENULL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->st;
  return server_error(s);
}

}  // extern "C"

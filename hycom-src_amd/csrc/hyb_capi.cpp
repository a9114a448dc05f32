// hyb_capi.cpp -- the C ABI of include/hycom_hybgen.h: parameter checks and
// derivation (host), the copy-mode and device-resident hybgen calls, and
// the synthetic-input generator entry points.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <omp.h>

#include <mutex>
#include <vector>

#include "../../include/hycom_hybgen.h"
#include "hyb_host.h"
#include "hyb_kernels.h"
#include "hyb_params_conv.h"

using namespace hyb;

namespace hyb {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char* last_error() { return g_err; }

struct Shape {
  int64_t NI = 0, NJ = 0, NIJ = 0;
  int K = 0, ntr = 0;
  bool operator==(const Shape& o) const {
    return NI == o.NI && NJ == o.NJ && K == o.K && ntr == o.ntr;
  }
};

// Tracers the array holds: the ntracr the call remaps, and the diagnostic
// tracers 801-803 wherever their indices put them.  The model keeps those
// after the remapped ones (its mtracr, blkdat.F90:1307-1330), and hybgen
// writes them without remapping them (hybgen.F90:44-164 beside the
// do ktr=1,ntracr loops of hybgenaj).
int tracer_extent(const hyb_params* P) {
  int e = P->ntracr;
  const int it[3] = {P->itracr801, P->itracr802, P->itracr803};
  for (int a = 0; a < 3; a++) e = it[a] > e ? it[a] : e;
  return e;
}

Shape shape_of(const hyb_params* P) {
  Shape s;
  s.NI = P->idm + 2 * P->nbdy;
  s.NJ = P->jdm + 2 * P->nbdy;
  s.NIJ = s.NI * s.NJ;
  s.K = P->kdm;
  s.ntr = tracer_extent(P);
  return s;
}

int device_ready() {
  static int state = 0;  // 0 unknown, 1 ok, <0 error code
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (state == 1) return 0;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1)
    return fail(HYB_ENODEVICE, "no HIP device visible (%s)", hipGetErrorString(e));
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(HYB_ENODEVICE, "device %d is %s; this library is built for gfx950",
                dev, prop.gcnArchName);
  state = 1;
  return 0;
}

int check_fields(const hyb_params* P, const hyb_fields* F, int raflag) {
  if (!F) return fail(HYB_ENULL, "fields pointer is NULL");
  const void* need[] = {F->dp, F->dpu, F->dpv, F->temp, F->saln, F->th3d, F->u, F->v,
                        F->p, F->pu, F->pv, F->theta, F->depths, F->topiso,
                        F->depthu, F->depthv, F->ubavg, F->vbavg, F->ip, F->iu, F->iv};
  const char* names[] = {"dp", "dpu", "dpv", "temp", "saln", "th3d", "u", "v", "p", "pu",
                         "pv", "theta", "depths", "topiso", "depthu", "depthv", "ubavg",
                         "vbavg", "ip", "iu", "iv"};
  for (size_t q = 0; q < sizeof need / sizeof need[0]; q++)
    if (!need[q]) return fail(HYB_ENULL, "field %s is NULL", names[q]);
  if (tracer_extent(P) > 0 && !F->tracer)
    return fail(HYB_ENULL, "field tracer is NULL (ntracr or itracr > 0)");
  if (raflag && !F->dpo) return fail(HYB_ENULL, "field dpo is NULL (hybgen_raflag set)");
  return 0;
}

int batch_domain_ok(const hyb_params* P, const hyb_fields* F, int d) {
  if (!P || !F) return fail(HYB_ENULL, "domain %d: NULL params or fields", d);
  int rc;
  if ((rc = hycom_hybgen_check_params(P))) return rc;
  if ((rc = check_fields(P, F, 0))) return rc;
  if (!stream_path_ok(kparams_of(P)))
    return fail(HYB_EUNSUPPORTED, "domain %d: batches need nhybrd==kdm", d);
  if (P->itracr801 || P->itracr802 || P->itracr803)
    return fail(HYB_EUNSUPPORTED, "domain %d: no itracr diagnostics in batches", d);
  return 0;
}

void batch_layout(const KParams* hP, int nd, BatchLaunch& L, std::vector<int32_t>& I) {
  auto group = [&](const std::vector<int>& doms, auto blocks_of, int key, BatchGroup* G) {
    G->ndom = (int32_t)doms.size();
    G->key = key;
    const size_t od = I.size();
    for (int d : doms) I.push_back(d);
    const size_t op = I.size();
    int64_t acc = 0;
    I.push_back(0);
    for (int d : doms) {
      acc += blocks_of(hP[d]);
      I.push_back((int32_t)acc);
    }
    G->nblocks = (int32_t)acc;
    G->dom = (const int32_t*)(uintptr_t)od;
    G->prefix = (const int32_t*)(uintptr_t)op;
  };
  // row kernels map rows as i_aligned0 does: kAlignPad more columns a row
  auto rows = [](int cols) {
    return [cols](const KParams& P) {
      return (int64_t)P.jj * ((P.ii + kAlignPad + cols - 1) / cols);
    };
  };
  auto uvblocks = [](const KParams& P) {
    return 2 * (int64_t)P.jj * ((P.ii + kAlignPad + 63) / 64);
  };
  std::vector<int> all(nd);
  for (int d = 0; d < nd; d++) all[d] = d;
  L.kmax = 0;
  for (int d = 0; d < nd; d++) L.kmax = hP[d].kk > L.kmax ? hP[d].kk : L.kmax;
  group(all, rows(L.regrid_cols), 0, &L.regrid);
  L.nremap = 0;
  for (int key = 0; key < kMaxGroups; key++) {
    std::vector<int> ds;
    for (int d = 0; d < nd; d++)
      if (hP[d].hybmap + 4 * hP[d].precision == key) ds.push_back(d);
    if (!ds.empty()) group(ds, rows(L.remap_cols), key, &L.remap[L.nremap++]);
  }
  group(all, rows(64), 0, &L.rows64);
  group(all, [](const KParams& P) { return (int64_t)P.kk * ((P.ii + 255) / 256); }, 0,
        &L.halo_j);
  group(all, [](const KParams& P) { return (int64_t)P.kk * ((P.jj + 2 + 255) / 256); }, 0,
        &L.halo_i);
  L.nuv = 0;
  for (int key = 0; key < kMaxGroups; key++) {
    std::vector<int> ds;
    for (int d = 0; d < nd; d++)
      if (uv_method_of(hP[d]) + 4 * hP[d].precision == key) ds.push_back(d);
    if (!ds.empty()) group(ds, uvblocks, key, &L.uv[L.nuv++]);
  }
  group(all, uvblocks, 0, &L.uvall);
}

void batch_bind(BatchLaunch& L, const int32_t* dI) {
  auto fix = [&](BatchGroup& G) {
    G.dom = dI + (uintptr_t)G.dom;
    G.prefix = dI + (uintptr_t)G.prefix;
  };
  fix(L.regrid);
  for (int q = 0; q < L.nremap; q++) fix(L.remap[q]);
  fix(L.rows64);
  fix(L.halo_j);
  fix(L.halo_i);
  for (int q = 0; q < L.nuv; q++) fix(L.uv[q]);
  fix(L.uvall);
}

}  // namespace hyb

namespace {

// device mirrors for the copy-mode call
struct Mirror {
  Shape shape;
  double *dp = nullptr, *dpo = nullptr, *dpu = nullptr, *dpv = nullptr;
  double *temp = nullptr, *saln = nullptr, *th3d = nullptr, *u = nullptr, *v = nullptr;
  double *tracer = nullptr, *p = nullptr, *pu = nullptr, *pv = nullptr, *theta = nullptr;
  double *depths = nullptr, *topiso = nullptr, *depthu = nullptr, *depthv = nullptr;
  double *ubavg = nullptr, *vbavg = nullptr;
  int32_t *ip = nullptr, *iu = nullptr, *iv = nullptr;
  hipStream_t stream = nullptr;
  int64_t up = 0, down = 0;  // bytes moved by the last call
  void release() {
    void* ptrs[] = {dp, dpo, dpu, dpv, temp, saln, th3d, u, v, tracer, p, pu, pv, theta,
                    depths, topiso, depthu, depthv, ubavg, vbavg, ip, iu, iv};
    for (void* x : ptrs)
      if (x) (void)hipFree(x);
    if (stream) (void)hipStreamDestroy(stream);
    Shape z;
    *this = Mirror();
    shape = z;
  }
};

std::mutex g_mirror_mu;
Mirror g_mirror;

// Copy mode's transfers of pageable host arrays go through two pinned
// staging buffers: the CPU copies a piece into one buffer (threads split
// large pieces) while the DMA engine moves the other, so uploads and
// downloads run near the PCIe rate.  The runtime's own path for pageable
// memory ran at about 27 GB/s up and 17 GB/s down on MI355X, the staged one
// at about 54 GB/s both ways.  Pieces are queued in order on one stream; a
// buffer is written again only after the event recorded behind its last
// DMA, and the downloads queued into it are copied out to the host then.
struct Stager {
  static constexpr size_t kCap = size_t(32) << 20;
  char* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool armed[2] = {false, false};
  struct Out {
    void* host;
    size_t off, n;
  };
  std::vector<Out> pending[2];
  int q = 0;
  size_t fill = 0;
  hipStream_t st = nullptr;

  static void pcopy(void* d, const void* s_, size_t n) {
    if (n < (size_t(4) << 20)) {
      memcpy(d, s_, n);
      return;
    }
    int nt = omp_get_max_threads();
    nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int t = 0; t < nt; t++) {
      const size_t a = n * (size_t)t / (size_t)nt, b = n * (size_t)(t + 1) / (size_t)nt;
      memcpy((char*)d + a, (const char*)s_ + a, b - a);
    }
  }
  hipError_t begin(hipStream_t s_) {
    st = s_;
    q = 0;
    fill = 0;
    for (int b = 0; b < 2; b++) {
      if (!buf[b]) {
        hipError_t e = hipHostMalloc((void**)&buf[b], kCap, hipHostMallocDefault);
        if (e != hipSuccess) return e;
      }
      if (!ev[b]) {
        hipError_t e = hipEventCreateWithFlags(&ev[b], hipEventDisableTiming);
        if (e != hipSuccess) return e;
      }
      armed[b] = false;
      pending[b].clear();
    }
    return hipSuccess;
  }
  // buffer b is free again: its DMA done, its downloads copied out
  hipError_t reuse(int b) {
    if (armed[b]) {
      hipError_t e = hipEventSynchronize(ev[b]);
      if (e != hipSuccess) return e;
      armed[b] = false;
    }
    for (const Out& o : pending[b]) pcopy(o.host, buf[b] + o.off, o.n);
    pending[b].clear();
    return hipSuccess;
  }
  hipError_t next() {
    hipError_t e = hipEventRecord(ev[q], st);
    if (e != hipSuccess) return e;
    armed[q] = true;
    q ^= 1;
    fill = 0;
    return reuse(q);
  }
  hipError_t up(void* dev, const void* host, size_t n) {
    while (n > 0) {
      if (fill == kCap) {
        hipError_t e = next();
        if (e != hipSuccess) return e;
      }
      const size_t m = n < kCap - fill ? n : kCap - fill;
      pcopy(buf[q] + fill, host, m);
      hipError_t e = hipMemcpyAsync(dev, buf[q] + fill, m, hipMemcpyHostToDevice, st);
      if (e != hipSuccess) return e;
      fill += m;
      dev = (char*)dev + m;
      host = (const char*)host + m;
      n -= m;
    }
    return hipSuccess;
  }
  hipError_t down(void* host, const void* dev, size_t n) {
    while (n > 0) {
      if (fill == kCap) {
        hipError_t e = next();
        if (e != hipSuccess) return e;
      }
      const size_t m = n < kCap - fill ? n : kCap - fill;
      hipError_t e = hipMemcpyAsync(buf[q] + fill, dev, m, hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) return e;
      pending[q].push_back({host, fill, m});
      fill += m;
      dev = (const char*)dev + m;
      host = (char*)host + m;
      n -= m;
    }
    return hipSuccess;
  }
  // everything queued has landed and every download is in host memory
  hipError_t finish() {
    hipError_t e = hipEventRecord(ev[q], st);
    if (e != hipSuccess) return e;
    armed[q] = true;
    for (int b = 0; b < 2; b++) {
      const int x = (q + 1 + b) & 1;  // the older buffer first
      if ((e = reuse(x)) != hipSuccess) return e;
    }
    fill = 0;
    return hipSuccess;
  }
  // after a failed call, once its stream has drained: nothing is in flight
  // and the downloads queued so far are not copied out
  void drop() {
    for (int b = 0; b < 2; b++) {
      armed[b] = false;
      pending[b].clear();
    }
    fill = 0;
  }
};

Stager g_stager;  // under g_mirror_mu; its pinned buffers live until the process ends

// measurement: per-phase HIP events of the calls since the last read
std::mutex g_timing_mu;
bool g_timing_on = false;
std::vector<PhaseTimer*> g_timers_pending, g_timers_free;

PhaseTimer* timer_take() {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  if (!g_timing_on) return nullptr;
  PhaseTimer* t;
  if (!g_timers_free.empty()) {
    t = g_timers_free.back();
    g_timers_free.pop_back();
  } else {
    t = new PhaseTimer;
    for (int q = 0; q < kPhases; q++) {
      if (hipEventCreate(&t->start[q]) != hipSuccess || hipEventCreate(&t->stop[q]) != hipSuccess) {
        delete t;
        return nullptr;
      }
    }
  }
  g_timers_pending.push_back(t);
  return t;
}

// Workspace of the streaming kernels (pres, kp, lcm, uvf, and ras for the
// Robert-Asselin branch, allocated with the first hybgen_raflag call), one
// per domain, keyed by the domain's p array; allocated on first use.
struct Workspace {
  const void* key = nullptr;
  int64_t nij = 0;
  int K = 0;
  double* pres = nullptr;
  int32_t* kp = nullptr;
  uint64_t* lcm = nullptr;
  int32_t* uvf = nullptr;
  double* ras = nullptr;
};
std::mutex g_ws_mu;
std::vector<Workspace> g_ws;
int g_kernel_path = 0;  // 0 auto, 1 per-lane arrays, 3 streaming (2 was register-resident)
StreamTuning g_tune;

int workspace_for(const void* key, int64_t nij, int K, Scratch* W, bool ra = false) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  for (auto& w : g_ws) {
    if (w.key == key && w.nij == nij && w.K == K) {
      if (ra && !w.ras) HIPCHK(hipMalloc(&w.ras, sizeof(double) * 3 * nij * K));
      *W = Scratch{w.pres, w.kp, w.lcm, w.uvf, w.ras};
      return 0;
    }
  }
  for (auto it = g_ws.begin(); it != g_ws.end(); ++it) {
    if (it->key == key) {  // same domain, new shape
      (void)hipFree(it->pres);
      (void)hipFree(it->kp);
      (void)hipFree(it->lcm);
      (void)hipFree(it->uvf);
      if (it->ras) (void)hipFree(it->ras);
      g_ws.erase(it);
      break;
    }
  }
  Workspace w;
  w.key = key;
  w.nij = nij;
  w.K = K;
  HIPCHK(hipMalloc(&w.pres, sizeof(double) * nij * (K + 1)));
  HIPCHK(hipMalloc(&w.kp, sizeof(int32_t) * nij));
  HIPCHK(hipMalloc(&w.lcm, sizeof(uint64_t) * nij));
  HIPCHK(hipMalloc(&w.uvf, sizeof(int32_t) * 2 * nij));
  if (ra) HIPCHK(hipMalloc(&w.ras, sizeof(double) * 3 * nij * K));
  g_ws.push_back(w);
  *W = Scratch{w.pres, w.kp, w.lcm, w.uvf, w.ras};
  return 0;
}

// the optional NaN/Inf scan: one int32 per scan block, grown as needed
std::mutex g_scan_mu;
int32_t* g_scan = nullptr;
size_t g_scan_n = 0;

const char* scan_field_name(int f) {
  static const char* names[] = {"dp", "temp", "saln", "th3d", "theta", "depths", "topiso",
                                "u", "dpu", "depthu", "ubavg", "v", "dpv", "depthv", "vbavg"};
  return f >= 0 && f < 15 ? names[f] : "tracer";
}

// HYB_ENONFINITE with the first bad point (row order) if an input read by
// the call is NaN or Inf; waits for the stream
int scan_finite(const KParams& K, const View& V, int n, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_scan_mu);
  const size_t nb = (size_t)scan_blocks(K);
  if (nb > g_scan_n) {
    if (g_scan) (void)hipFree(g_scan);
    g_scan = nullptr;
    g_scan_n = 0;
    HIPCHK(hipMalloc(&g_scan, sizeof(int32_t) * nb));
    g_scan_n = nb;
  }
  HIPCHK(launch_scan_finite(K, V, n, g_scan, s));
  std::vector<int32_t> h(nb);
  HIPCHK(hipMemcpyAsync(h.data(), g_scan, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int nbx = (int)(nb / (size_t)K.jj);
  for (size_t b = 0; b < nb; b++) {
    if (h[b] == INT32_MAX) continue;
    const int j = (int)(b / (size_t)nbx) + 1, i = h[b] / 128, f = h[b] % 128;
    if (f >= 16)
      return fail(HYB_ENONFINITE, "non-finite %s %d at i=%d j=%d", scan_field_name(f), f - 16, i,
                  j);
    return fail(HYB_ENONFINITE, "non-finite %s at i=%d j=%d", scan_field_name(f), i, j);
  }
  return 0;
}

thread_local int32_t* g_rec_fixlay = nullptr;
thread_local int32_t* g_rec_kp = nullptr;

}  // namespace

extern "C" {

int hycom_hybgen_abi_version(void) { return HYB_ABI_VERSION; }

const char* hycom_hybgen_strerror(int code) {
  switch (code) {
    case HYB_OK: return "ok";
    case HYB_EINVAL: return "invalid parameter";
    case HYB_EUNSUPPORTED: return "unsupported configuration";
    case HYB_EDEVICE: return "HIP runtime error";
    case HYB_ENOMEM: return "device out of memory";
    case HYB_ENODEVICE: return "no usable gfx950 device";
    case HYB_ENULL: return "required field pointer is NULL";
    case HYB_ENONFINITE: return "non-finite input";
    case HYB_EBUSY: return "no free slot or memory";
    default: return "unknown error";
  }
}

const char* hycom_hybgen_last_error(void) { return g_err; }

int hycom_hybgen_check_params(const hyb_params* P) {
  if (!P) return fail(HYB_ENULL, "params pointer is NULL");
  if (P->kdm < 4 || P->kdm > HYB_KMAX)
    return fail(HYB_EUNSUPPORTED, "kdm=%d outside 4..%d", P->kdm, HYB_KMAX);
  if (P->nbdy < 1) return fail(HYB_EINVAL, "nbdy=%d must be >= 1", P->nbdy);
  if (P->ii < 1 || P->jj < 1 || P->ii > P->idm || P->jj > P->jdm)
    return fail(HYB_EINVAL, "active extent %dx%d outside array %dx%d", P->ii, P->jj,
                P->idm, P->jdm);
  if (P->ntracr < 0 || P->ntracr > HYB_TRMAX)
    return fail(HYB_EINVAL, "ntracr=%d outside 0..%d", P->ntracr, HYB_TRMAX);
  if (P->ntracr > 64) return fail(HYB_EUNSUPPORTED, "ntracr=%d > 64", P->ntracr);
  // blkdat.F90:303-311 (nhybrd<=kdm); hybgen is only called when hybrid
  if (P->nhybrd < 1 || P->nhybrd > P->kdm)
    return fail(HYB_EINVAL, "nhybrd=%d outside 1..kdm", P->nhybrd);
  if (!(P->qhybrlx > 0.0 && P->qhybrlx <= 1.0))  // blkdat.F90:946-954 hybrlx>=1
    return fail(HYB_EINVAL, "qhybrlx=%g: hybrlx must be at least 1.0", P->qhybrlx);
  if (!(P->hybthn >= 1.0))  // blkdat.F90:956-964
    return fail(HYB_EINVAL, "hybthn=%g must be at least 1.0", P->hybthn);
  if (!(P->hybthk >= 0.0 && P->hybthk <= 0.9))  // blkdat.F90:966-974
    return fail(HYB_EINVAL, "hybthk=%g must be between 0.0 and 0.9", P->hybthk);
  if (P->hybmap < 0 || P->hybmap > 3)  // blkdat.F90:987-996
    return fail(HYB_EINVAL, "hybmap=%d must be 0..3", P->hybmap);
  if (P->hybflg < 0 || P->hybflg > 2)  // blkdat.F90:997-1005
    return fail(HYB_EINVAL, "hybflg=%d must be 0..2", P->hybflg);
  if (P->hybflg != 0)  // blkdat.F90:1006-1012 (17-term EOS)
    return fail(HYB_EINVAL, "hybflg must be 0 (T&S) for 17-term eqn. of state");
  if ((P->isopcm != 0) != (P->hybiso > 0.0))  // blkdat.F90:986
    return fail(HYB_EINVAL, "isopcm must equal hybiso>0");
  for (int k = P->nhybrd; k < P->kdm; k++)  // blkdat.F90:271-281
    if (P->dp0k[k] != 0.0) return fail(HYB_EINVAL, "dp0k must be zero for k>nhybrd");
  if (P->mxlmy) return fail(HYB_EUNSUPPORTED, "mxlmy (MY2.5 q2/q2l) is not supported");
  if (!P->isopcm && P->nhybrd < P->kdm)
    return fail(HYB_EUNSUPPORTED,
                "isopcm off with nhybrd<kdm carries PCM flags across columns of a row");
  int it[3] = {P->itracr801, P->itracr802, P->itracr803};
  for (int a = 0; a < 3; a++) {
    if (it[a] < 0 || it[a] > HYB_TRMAX)  // mxtrcr, mod_dimensions.F90:19-21
      return fail(HYB_EINVAL, "itracr(%d)=%d outside 0..%d", 801 + a, it[a], HYB_TRMAX);
    for (int b = 0; b < a; b++)
      if (it[a] && it[a] == it[b])
        return fail(HYB_EUNSUPPORTED, "itracr(%d) and itracr(%d) share a tracer", 801 + b,
                    801 + a);
    if (it[a] && !(P->baclin > 0.0))
      return fail(HYB_EINVAL, "baclin must be > 0 with itracr(%d)", 801 + a);
  }
  if (P->precision != HYB_PREC_FP64 && P->precision != HYB_PREC_MIXED)
    return fail(HYB_EINVAL, "precision=%d", P->precision);
  if (P->precision == HYB_PREC_MIXED && P->itracr801 + P->itracr802 + P->itracr803 > 0)
    return fail(HYB_EUNSUPPORTED, "diagnostic tracers need the FP64 path");
  return 0;
}

int hycom_hybgen_derive_vgrid(hyb_params* P, double dp00, double dp00f, double dp00x,
                              double ds00, double ds00f, double ds00x, double dp00i_m,
                              double dx00, double dx00f, double dx00x, double hybrlx,
                              double isotop_m, double* topiso_pa) {
  if (!P) return fail(HYB_ENULL, "params pointer is NULL");
  const int kk = P->kdm, nhybrd = P->nhybrd;
  if (kk < 4 || kk > HYB_KMAX) return fail(HYB_EUNSUPPORTED, "kdm=%d", kk);
  if (nhybrd < 1 || nhybrd > kk) return fail(HYB_EINVAL, "nhybrd=%d", nhybrd);
  if (hybrlx < 1.0) return fail(HYB_EINVAL, "hybrlx must be at least 1.0");
  if (isotop_m == 0.0) return fail(HYB_EINVAL, "isotop cannot be 0.0");  // blkdat.F90:403
  if (dp00 < 0.0) return fail(HYB_EUNSUPPORTED, "dp0k input (dp00<0) not derived here");
  // blkdat.F90:295-301
  int nsigma = P->nsigma <= 1 ? 1 : P->nsigma;
  if (nsigma > nhybrd) return fail(HYB_EINVAL, "maximum nsigma is nhybrd");
  P->nsigma = nsigma;
  // geopar.F90:449-482 maximum thickness
  if (dx00 > 0.0) {
    double x0 = kOnem * dx00, xx = kOnem * dx00x, f = 1.0;
    P->dx0k[0] = x0;
    for (int k = 2; k <= kk; k++) {
      f = f * dx00f;
      P->dx0k[k - 1] = dmin(x0 * f, xx);
    }
  } else {
    for (int k = 1; k <= kk; k++) P->dx0k[k - 1] = kOnem * P->dx0k[k - 1];
  }
  // geopar.F90:520-600 deep and shallow z-level spacing
  double d0 = kOnem * dp00, dx = kOnem * dp00x;
  P->dp00i = kOnem * dp00i_m;
  P->dp0k[0] = d0;
  double f = 1.0;
  for (int k = 2; k <= kk; k++) {
    f = f * dp00f;
    if (k <= nhybrd) {
      P->dp0k[k - 1] = dp00f >= 1.0 ? dmin(d0 * f, dx) : dmax(d0 * f, dx);
    } else {
      P->dp0k[k - 1] = 0.0;
    }
  }
  double s0 = kOnem * ds00, sx = kOnem * ds00x;
  P->ds0k[0] = s0;
  f = 1.0;
  for (int k = 2; k <= nsigma; k++) {
    f = f * ds00f;
    P->ds0k[k - 1] = ds00f >= 1.0 ? dmin(s0 * f, sx) : dmax(s0 * f, sx);
  }
  // geopar.F90:609-630 terrain-following start and stop depths (nsigma>=1)
  double dpns = 0.0, dsns = 0.0;
  for (int k = 1; k <= nsigma; k++) {
    dpns = dpns + P->dp0k[k - 1];
    dsns = dsns + P->ds0k[k - 1];
  }
  for (int k = nsigma + 1; k <= kk; k++) P->ds0k[k - 1] = 0.0;
  P->dpns = dpns * kQonem;
  P->dsns = dsns * kQonem;
  // blkdat.F90:954,986
  P->qhybrlx = 1.0 / hybrlx;
  P->isopcm = P->hybiso > 0.0;
  if (topiso_pa) *topiso_pa = kOnem * isotop_m;  // mod_hycom.F90:1447
  return 0;
}

int hycom_hybgen_record_decisions(int32_t* fixlay_dev, int32_t* kp_dev) {
  g_rec_fixlay = fixlay_dev;
  g_rec_kp = kp_dev;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Batched ragged domains
// ---------------------------------------------------------------------
namespace {

struct BatchPlan {
  std::vector<unsigned char> key;  // the KParams and View of every domain
  int ndom = 0;
  KParams* dP = nullptr;
  View* dV = nullptr;
  Scratch* dW = nullptr;
  int32_t* dI = nullptr;
  size_t nI = 0;
  BatchDevice B{};
  BatchLaunch L{};
  hipStream_t cap = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // per time level n
  // online autotuning: first the remap workgroups (columns, staged
  // interface levels) with 256-column regrid workgroups, then the regrid
  // widths with the best remap candidate
  struct Cand {
    int rc, ns, gc;  // remap columns, staged levels, regrid columns
  };
  std::vector<Cand> cands;
  std::vector<float> ms;  // best time of each candidate
  int cand = 0;           // candidate measured by the pending events
  int rep = 0;            // measurements of it so far
  bool second = false;    // the regrid widths are being searched
  bool pending = false, tuned = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t done = nullptr;  // after the last launch of this plan's graphs
  bool launched = false;
  std::vector<KParams> hP;
  std::vector<View> hV;
};
std::mutex g_batch_mu;
BatchPlan g_plan;
bool g_batch_autotune = true;

void plan_drop_graphs(BatchPlan& b) {
  for (int q = 0; q < 2; q++) {
    if (b.exec[q]) (void)hipGraphExecDestroy(b.exec[q]);
    if (b.graph[q]) (void)hipGraphDestroy(b.graph[q]);
    b.exec[q] = nullptr;
    b.graph[q] = nullptr;
  }
}

// waits for the last launch of the plan's graphs before its arrays change
int plan_quiesce(BatchPlan& b) {
  if (b.launched) HIPCHK(hipEventSynchronize(b.done));
  b.launched = false;
  return 0;
}

void plan_free(BatchPlan& b) {
  plan_drop_graphs(b);
  if (b.dP) (void)hipFree(b.dP);
  if (b.dV) (void)hipFree(b.dV);
  if (b.dW) (void)hipFree(b.dW);
  if (b.dI) (void)hipFree(b.dI);
  b.dP = nullptr;
  b.dV = nullptr;
  b.dW = nullptr;
  b.dI = nullptr;
  b.nI = 0;
  b.key.clear();
}

// block layouts of every launch for the current widths
int plan_layout(BatchPlan& b) {
  if (int rc = plan_quiesce(b)) return rc;
  std::vector<int32_t> I;
  batch_layout(b.hP.data(), b.ndom, b.L, I);
  if (I.size() > b.nI) {
    if (b.dI) (void)hipFree(b.dI);
    b.dI = nullptr;
    HIPCHK(hipMalloc(&b.dI, sizeof(int32_t) * I.size()));
    b.nI = I.size();
  }
  HIPCHK(hipMemcpy(b.dI, I.data(), sizeof(int32_t) * I.size(), hipMemcpyHostToDevice));
  batch_bind(b.L, b.dI);
  plan_drop_graphs(b);
  return 0;
}

int plan_graph(BatchPlan& b, int n) {
  const int q = n - 1;
  if (b.exec[q]) return 0;
  if (!b.cap) HIPCHK(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking));
  HIPCHK(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_batch(b.B, b.L, n, b.cap);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b.cap, &g);
  HIPCHK(e);
  HIPCHK(e2);
  b.graph[q] = g;
  HIPCHK(hipGraphInstantiate(&b.exec[q], g, nullptr, nullptr, 0));
  return 0;
}

int plan_build(BatchPlan& b, const hyb_domain* D, int ndom, std::vector<unsigned char>& key) {
  if (int rc = plan_quiesce(b)) return rc;
  plan_free(b);
  b.ndom = ndom;
  b.hP.resize(ndom);
  b.hV.resize(ndom);
  std::vector<Scratch> hW(ndom);
  for (int d = 0; d < ndom; d++) {
    b.hP[d] = kparams_of(D[d].P);
    b.hV[d] = view_of(D[d].P, D[d].F);
    int rc = workspace_for(D[d].F->p, b.hV[d].NIJ, D[d].P->kdm, &hW[d]);
    if (rc) return rc;
  }
  HIPCHK(hipMalloc(&b.dP, sizeof(KParams) * ndom));
  HIPCHK(hipMalloc(&b.dV, sizeof(View) * ndom));
  HIPCHK(hipMalloc(&b.dW, sizeof(Scratch) * ndom));
  HIPCHK(hipMemcpy(b.dP, b.hP.data(), sizeof(KParams) * ndom, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.dV, b.hV.data(), sizeof(View) * ndom, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.dW, hW.data(), sizeof(Scratch) * ndom, hipMemcpyHostToDevice));
  b.B = BatchDevice{b.dP, b.dV, b.dW};
  int kmax = 0;
  for (int d = 0; d < ndom; d++) kmax = b.hP[d].kk > kmax ? b.hP[d].kk : kmax;
  // remap candidates: a workgroup stages ns interface levels of each of
  // its columns in LDS (ns x cols doubles, at most 64 KB): all kmax-1, or
  // 24 or 12 with the deeper levels read from p as the walk reaches them
  b.cands.clear();
  const int lvl = kmax > 1 ? kmax - 1 : 1;
  for (int rc : {64, 128, 256})
    for (int ns : {lvl, 24, 12}) {
      if (ns != lvl && ns >= lvl) continue;  // a split stages fewer levels than all
      if (sizeof(double) * ns * rc <= 65536) b.cands.push_back({rc, ns, 256});
    }
  b.ms.assign(b.cands.size(), -1.0f);
  b.cand = 0;
  b.rep = 0;
  b.second = false;
  b.pending = false;
  b.tuned = !g_batch_autotune;
  // first candidate (and the untuned widths): 64-column remaps staging every
  // level, 256-column regrid
  b.L.remap_cols = b.cands[0].rc;
  b.L.remap_stage = b.cands[0].ns;
  b.L.regrid_cols = b.cands[0].gc;
  if (!b.ev0) HIPCHK(hipEventCreate(&b.ev0));
  if (!b.ev1) HIPCHK(hipEventCreate(&b.ev1));
  if (!b.done) HIPCHK(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
  b.key = key;
  return plan_layout(b);
}

}  // namespace

extern "C" {

int hycom_hybgen_batch(const hyb_domain* D, int ndom, int m, int n, void* stream) {
  if (!D || ndom < 1) return fail(HYB_EINVAL, "empty batch");
  if (m < 1 || m > 2 || n < 1 || n > 2) return fail(HYB_EINVAL, "time levels m=%d n=%d", m, n);
  int rc;
  for (int d = 0; d < ndom; d++) {
    if ((rc = batch_domain_ok(D[d].P, D[d].F, d))) return rc;
    for (int e = 0; e < d; e++)
      if (D[e].F->p == D[d].F->p) return fail(HYB_EINVAL, "domains %d and %d share p", e, d);
  }
  if ((rc = device_ready())) return rc;
  std::vector<unsigned char> key;
  key.reserve(ndom * (sizeof(KParams) + sizeof(View)));
  for (int d = 0; d < ndom; d++) {
    KParams k = kparams_of(D[d].P);
    View v = view_of(D[d].P, D[d].F);
    const unsigned char* a = (const unsigned char*)&k;
    key.insert(key.end(), a, a + sizeof k);
    a = (const unsigned char*)&v;
    key.insert(key.end(), a, a + sizeof v);
  }
  for (int d = 0; d < ndom; d++)
    if (D[d].P->check_finite) {
      const KParams K = kparams_of(D[d].P);
      const View V = view_of(D[d].P, D[d].F);
      if ((rc = scan_finite(K, V, n, (hipStream_t)stream))) return rc;
    }
  std::lock_guard<std::mutex> lk(g_batch_mu);
  BatchPlan& b = g_plan;
  if (b.key != key || b.ndom != ndom) {
    if ((rc = plan_build(b, D, ndom, key))) return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  // autotuning: read the pending measurement, move to the next candidate
  if (b.pending) {
    HIPCHK(hipEventSynchronize(b.ev1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, b.ev0, b.ev1));
    b.ms[b.cand] = (b.rep == 0 || t < b.ms[b.cand]) ? t : b.ms[b.cand];
    b.pending = false;
    // two launches per candidate: the first one also loads the graph
    if (++b.rep == 2) {
      b.rep = 0;
      int next = ++b.cand;
      if (next >= (int)b.cands.size()) {
        int best = 0;
        for (int q = 1; q < (int)b.cands.size(); q++)
          if (b.ms[q] < b.ms[best]) best = q;
        next = best;
        if (!b.second) {
          // the regrid widths, with the best remap workgroup
          b.second = true;
          const BatchPlan::Cand w = b.cands[best];
          for (int gc : {128, 64}) {
            b.cands.push_back({w.rc, w.ns, gc});
            b.ms.push_back(-1.0f);
          }
          next = b.cand;
        } else {
          b.tuned = true;
        }
      }
      b.L.remap_cols = b.cands[next].rc;
      b.L.remap_stage = b.cands[next].ns;
      b.L.regrid_cols = b.cands[next].gc;
      if ((rc = plan_layout(b))) return rc;
    }
  }
  if ((rc = plan_graph(b, n))) return rc;
  const bool measure = !b.tuned;
  if (measure) HIPCHK(hipEventRecord(b.ev0, s));
  HIPCHK(hipGraphLaunch(b.exec[n - 1], s));
  if (measure) {
    HIPCHK(hipEventRecord(b.ev1, s));
    b.pending = true;
  }
  HIPCHK(hipEventRecord(b.done, s));
  b.launched = true;
  (void)m;
  return 0;
}

int hycom_hybgen_batch_autotune(int enable) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  g_batch_autotune = enable != 0;
  return 0;
}

int hycom_hybgen_batch_config(int32_t* remap_cols, int32_t* regrid_cols, int32_t* tuned) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  if (remap_cols) *remap_cols = g_plan.L.remap_cols;
  if (regrid_cols) *regrid_cols = g_plan.L.regrid_cols;
  if (tuned) *tuned = g_plan.tuned ? 1 : 0;
  return 0;
}

int hycom_hybgen_batch_tuning(int32_t* remap_cols, int32_t* staged_levels,
                              int32_t* regrid_cols, float* ms, int cap) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  if (cap < 0) return fail(HYB_EINVAL, "cap %d", cap);
  const int nc = (int)g_plan.cands.size();
  for (int q = 0; q < nc && q < cap; q++) {
    if (remap_cols) remap_cols[q] = g_plan.cands[q].rc;
    if (staged_levels) staged_levels[q] = g_plan.cands[q].ns;
    if (regrid_cols) regrid_cols[q] = g_plan.cands[q].gc;
    if (ms) ms[q] = g_plan.ms[q];
  }
  return nc;
}

int hycom_hybgen_batch_staged_levels(int32_t* levels, int32_t* kmax) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  if (levels) *levels = g_plan.L.kmax > 0 ? batch_stage_levels(g_plan.L) : 0;
  if (kmax) *kmax = g_plan.L.kmax;
  return 0;
}

}  // extern "C"

extern "C" {

int hycom_hybgen_device(const hyb_params* P, const hyb_fields* F, int m, int n, int raflag,
                        void* stream) {
  int rc = hycom_hybgen_check_params(P);
  if (rc) return rc;
  if ((rc = check_fields(P, F, raflag))) return rc;
  if (m < 1 || m > 2 || n < 1 || n > 2 || (raflag && m == n))
    return fail(HYB_EINVAL, "time levels m=%d n=%d", m, n);
  if (P->precision == HYB_PREC_MIXED && !stream_path_ok(kparams_of(P)))
    return fail(HYB_EUNSUPPORTED, "the mixed form needs nhybrd==kdm");
  if (P->precision == HYB_PREC_MIXED && (g_kernel_path == 1 || g_kernel_path == 2))
    return fail(HYB_EUNSUPPORTED, "the mixed form runs on the streaming kernels only");
  if ((rc = device_ready())) return rc;
  KParams K = kparams_of(P);
  View V = view_of(P, F);
  if (P->check_finite && (rc = scan_finite(K, V, n, (hipStream_t)stream))) return rc;
  V.rec_fixlay = g_rec_fixlay;
  V.rec_kp = g_rec_kp;
  const int path = g_kernel_path;
  if (path == 3 && !stream_path_ok(K))
    return fail(HYB_EUNSUPPORTED, "streaming kernels need nhybrd==kdm");
  const bool use_stream = (path == 0 || path == 3) && stream_path_ok(K);
  if (use_stream) {
    Scratch W;
    if ((rc = workspace_for(F->p, V.NIJ, P->kdm, &W, raflag != 0))) return rc;
    HIPCHK(launch_hybgen_stream(K, V, W, m, n, raflag, g_tune, (hipStream_t)stream,
                                timer_take()));
  } else {
    HIPCHK(launch_hybgen_fp64(K, V, m, n, raflag, (hipStream_t)stream, timer_take()));
  }
  return 0;
}

}  // extern "C"

namespace {

// Workspace of hycom_hybgen_device_noalloc: one fixed device array per GPU,
// held from the library's load until the process ends.  The array's name
// is only a host-side handle; the device address of each GPU's copy comes
// from hipGetSymbolAddress with that GPU current, and only that address is
// ever handed to kernels.
constexpr size_t kNoallocBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[64] = {};

size_t noalloc_round(size_t b) { return (b + 255) & ~size_t(255); }

// bytes of pres, kp, lcm and uvf for NIJ points and K layers (and ras
// with hybgen_raflag), each piece 256-byte aligned
size_t noalloc_need(int64_t nij, int K, bool ra = false) {
  return noalloc_round(sizeof(double) * (size_t)nij * (size_t)(K + 1)) +
         noalloc_round(sizeof(int32_t) * (size_t)nij) +
         noalloc_round(sizeof(uint64_t) * (size_t)nij) +
         noalloc_round(sizeof(int32_t) * 2 * (size_t)nij) +
         (ra ? noalloc_round(sizeof(double) * 3 * (size_t)nij * (size_t)K) : 0);
}

// the current device's copy of scratch, looked up once per device
int noalloc_base(unsigned char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(HYB_EDEVICE, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(HYB_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_noalloc_ptr[dev] = p;
  }
  *out = (unsigned char*)g_noalloc_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

size_t hycom_hybgen_noalloc_capacity(void) { return kNoallocBytes; }

size_t hycom_hybgen_noalloc_need(const hyb_params* P, int raflag) {
  if (!P) return 0;
  const Shape s = shape_of(P);
  return noalloc_need(s.NIJ, s.K, raflag != 0);
}

int hycom_hybgen_device_noalloc(const hyb_params* P, const hyb_fields* F, int m, int n,
                                int raflag, void* stream) {
  int rc = hycom_hybgen_check_params(P);
  if (rc) return rc;
  if ((rc = check_fields(P, F, raflag))) return rc;
  if (m < 1 || m > 2 || n < 1 || n > 2 || (raflag && m == n))
    return fail(HYB_EINVAL, "time levels m=%d n=%d", m, n);
  if (P->check_finite)
    return fail(HYB_EUNSUPPORTED,
                "check_finite needs a result buffer and a host wait; use hycom_hybgen_device");
  if (P->precision == HYB_PREC_MIXED && !stream_path_ok(kparams_of(P)))
    return fail(HYB_EUNSUPPORTED, "the mixed form needs nhybrd==kdm");
  if (P->precision == HYB_PREC_MIXED && g_kernel_path == 1)
    return fail(HYB_EUNSUPPORTED, "the mixed form runs on the streaming kernels only");
  KParams K = kparams_of(P);
  const int path = g_kernel_path;
  if (path == 3 && !stream_path_ok(K))
    return fail(HYB_EUNSUPPORTED, "streaming kernels need nhybrd==kdm");
  const bool use_stream = (path == 0 || path == 3) && stream_path_ok(K);
  const int64_t nij = shape_of(P).NIJ;
  const size_t need = noalloc_need(nij, P->kdm, raflag != 0);
  if (use_stream && need > kNoallocBytes)
    return fail(HYB_EINVAL,
                "domain of %lld points x %d layers needs %zu bytes of workspace; the fixed "
                "array holds %zu (hycom_hybgen_noalloc_capacity)",
                (long long)nij, P->kdm, need, kNoallocBytes);
  if ((rc = device_ready())) return rc;
  View V = view_of(P, F);
  V.rec_fixlay = g_rec_fixlay;
  V.rec_kp = g_rec_kp;
  if (!use_stream) {  // the per-lane-array kernels use no workspace
    HIPCHK(launch_hybgen_fp64(K, V, m, n, raflag, (hipStream_t)stream, nullptr));
    return 0;
  }
  unsigned char* base = nullptr;
  if ((rc = noalloc_base(&base))) return rc;
  Scratch W;
  W.pres = (double*)base;
  base += noalloc_round(sizeof(double) * (size_t)V.NIJ * (size_t)(P->kdm + 1));
  W.kp = (int32_t*)base;
  base += noalloc_round(sizeof(int32_t) * (size_t)V.NIJ);
  W.lcm = (uint64_t*)base;
  base += noalloc_round(sizeof(uint64_t) * (size_t)V.NIJ);
  W.uvf = (int32_t*)base;
  base += noalloc_round(sizeof(int32_t) * 2 * (size_t)V.NIJ);
  W.ras = raflag ? (double*)base : nullptr;
  // the u/v phase stays on the caller's stream: a second stream would be
  // created on first use
  StreamTuning tune = g_tune;
  tune.concurrent_uv = 0;
  HIPCHK(launch_hybgen_stream(K, V, W, m, n, raflag, tune, (hipStream_t)stream, nullptr));
  return 0;
}

int hycom_hybgen_selftest_quotients(int which, const double* a, const double* b, double* fast,
                                    double* ieee, int64_t n, void* stream) {
  if (which < 0 || which > 4) return fail(HYB_EINVAL, "selftest quotient helper %d", which);
  if (n < 0 || n > ((int64_t)1 << 31)) return fail(HYB_EINVAL, "selftest length %lld", (long long)n);
  if (n > 0 && (!a || !b || !fast || !ieee)) return fail(HYB_ENULL, "selftest: NULL array");
  int rc;
  if ((rc = device_ready())) return rc;
  HIPCHK(launch_selftest_quot(which, a, b, fast, ieee, n, (hipStream_t)stream));
  return 0;
}

int hycom_hybgen_set_kernel_path(int path) {
  if (path < 0 || path > 3) return fail(HYB_EINVAL, "kernel path %d", path);
  if (path == 2) return fail(HYB_EUNSUPPORTED, "kernel path 2 (register-resident) was removed");
  g_kernel_path = path;
  return 0;
}

int hycom_hybgen_tune(int knob, int value) {
  if (knob == 0) {
    if (value != 3) return fail(HYB_EINVAL, "remap variant %d: only 3 is built", value);
    g_tune.remap_variant = value;
    return 0;
  }
  if (knob == 1) {
    if (value != 0 && value != 2 && value != 4)
      return fail(HYB_EINVAL, "u/v prefetch %d: 0, 2 or 4", value);
    g_tune.uv_prefetch = value;
    return 0;
  }
  if (knob == 2) {
    if (value != 0 && value != 1) return fail(HYB_EINVAL, "u/v recompute %d: 0 or 1", value);
    g_tune.uv_recompute = value;
    return 0;
  }
  if (knob == 3) {
    if (value != 0 && value != 1) return fail(HYB_EINVAL, "regrid read-ahead %d: 0 or 1", value);
    g_tune.regrid_ahead = value;
    return 0;
  }
  if (knob == 4) {
    if (value != 0 && value != 1) return fail(HYB_EINVAL, "concurrent u/v %d: 0 or 1", value);
    g_tune.concurrent_uv = value;
    return 0;
  }
  return fail(HYB_EINVAL, "tuning knob %d", knob);
}

int hycom_hybgen_timing(int enable) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing_on = enable != 0;
  return 0;
}

int hycom_hybgen_timing_read(double* ms_sum, int32_t* count, int nphase) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  for (int q = 0; q < nphase; q++) { ms_sum[q] = 0.0; count[q] = 0; }
  for (PhaseTimer* t : g_timers_pending) {
    for (int q = 0; q < kPhases && q < nphase; q++) {
      if (!t->used[q]) continue;
      HIPCHK(hipEventSynchronize(t->stop[q]));
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, t->start[q], t->stop[q]));
      ms_sum[q] += ms;
      count[q] += 1;
    }
    g_timers_free.push_back(t);
  }
  g_timers_pending.clear();
  return 0;
}

// Copy mode moves what the call reads and writes (hybgen.F90:14-306 with
// the fields of hybgenaj/bj/cj): time level n of dp, dpu, dpv, temp, saln,
// th3d, u, v and the tracers both ways, level n of ubavg/vbavg up, p, pu,
// pv both ways, and theta, depths, topiso, depthu, depthv and the masks
// up; with hybgen_raflag also level m of those fields and dpo.  Nothing is
// taken as unchanged from one call to the next: the model changes some of
// these in place between calls (dpthuv sets depthu/depthv after inicon's
// two hybgen calls, mod_hycom.F90:1495,1684; its halo updates rewrite the
// halo of p), and p, pu, pv come back whole, so their land and halo
// values must be the ones the caller holds.
int hycom_hybgen(const hyb_params* P, const hyb_fields* F, int m, int n, int raflag) {
  int rc = hycom_hybgen_check_params(P);
  if (rc) return rc;
  if ((rc = check_fields(P, F, raflag))) return rc;
  if (m < 1 || m > 2 || n < 1 || n > 2 || (raflag && m == n))
    return fail(HYB_EINVAL, "time levels m=%d n=%d", m, n);
  if ((rc = device_ready())) return rc;
  std::lock_guard<std::mutex> lk(g_mirror_mu);
  Mirror& M = g_mirror;
  Shape s = shape_of(P);
  const size_t n2 = (size_t)s.NIJ, n3 = n2 * s.K, n3p = n2 * (s.K + 1), n4 = 2 * n3;
  const int ntr = s.ntr > 0 ? s.ntr : 1;
  const size_t n5 = n4 * (size_t)ntr;
  if (!(M.shape == s) || !M.dp) {
    M.release();
    HIPCHK(hipStreamCreate(&M.stream));
    double** f4[] = {&M.dp, &M.dpo, &M.dpu, &M.dpv, &M.temp, &M.saln, &M.th3d, &M.u, &M.v};
    for (double** x : f4) HIPCHK(hipMalloc(x, n4 * sizeof(double)));
    HIPCHK(hipMalloc(&M.tracer, n5 * sizeof(double)));
    HIPCHK(hipMalloc(&M.p, n3p * sizeof(double)));
    HIPCHK(hipMalloc(&M.pu, n3p * sizeof(double)));
    HIPCHK(hipMalloc(&M.pv, n3p * sizeof(double)));
    HIPCHK(hipMalloc(&M.theta, n3 * sizeof(double)));
    double** f2[] = {&M.depths, &M.topiso, &M.depthu, &M.depthv};
    for (double** x : f2) HIPCHK(hipMalloc(x, n2 * sizeof(double)));
    HIPCHK(hipMalloc(&M.ubavg, 3 * n2 * sizeof(double)));
    HIPCHK(hipMalloc(&M.vbavg, 3 * n2 * sizeof(double)));
    HIPCHK(hipMalloc(&M.ip, n2 * sizeof(int32_t)));
    HIPCHK(hipMalloc(&M.iu, n2 * sizeof(int32_t)));
    HIPCHK(hipMalloc(&M.iv, n2 * sizeof(int32_t)));
    M.shape = s;
  }
  hipStream_t st = M.stream;
  Stager& X = g_stager;
  int64_t up = 0, down = 0;
  auto call = [&]() -> int {
    HIPCHK(X.begin(st));
    auto h2d = [&](void* d, const void* h, size_t bytes) {
      up += (int64_t)bytes;
      return X.up(d, h, bytes);
    };
    auto d2h = [&](void* h, const void* d, size_t bytes) {
      down += (int64_t)bytes;
      return X.down(h, d, bytes);
    };
    // the fields of time level l (1-based) that the call reads and writes
    const size_t b3 = n3 * sizeof(double);
    auto level = [&](int l, bool to_device) -> hipError_t {
      const size_t o = n3 * (size_t)(l - 1);
      double* dev[] = {M.dp, M.dpu, M.dpv, M.temp, M.saln, M.th3d, M.u, M.v};
      double* host[] = {F->dp, F->dpu, F->dpv, F->temp, F->saln, F->th3d, F->u, F->v};
      hipError_t e = hipSuccess;
      for (int q = 0; q < 8 && e == hipSuccess; q++)
        e = to_device ? h2d(dev[q] + o, host[q] + o, b3) : d2h(host[q] + o, dev[q] + o, b3);
      for (int t = 0; t < s.ntr && e == hipSuccess; t++) {
        const size_t ot = n4 * (size_t)t + o;
        e = to_device ? h2d(M.tracer + ot, F->tracer + ot, b3)
                      : d2h(F->tracer + ot, M.tracer + ot, b3);
      }
      if (to_device && e == hipSuccess) {
        const size_t ob = n2 * (size_t)(l - 1);
        e = h2d(M.ubavg + ob, F->ubavg + ob, n2 * sizeof(double));
        if (e == hipSuccess) e = h2d(M.vbavg + ob, F->vbavg + ob, n2 * sizeof(double));
      }
      return e;
    };
    HIPCHK(h2d(M.theta, F->theta, n3 * sizeof(double)));
    HIPCHK(h2d(M.depths, F->depths, n2 * sizeof(double)));
    HIPCHK(h2d(M.topiso, F->topiso, n2 * sizeof(double)));
    HIPCHK(h2d(M.depthu, F->depthu, n2 * sizeof(double)));
    HIPCHK(h2d(M.depthv, F->depthv, n2 * sizeof(double)));
    HIPCHK(h2d(M.ip, F->ip, n2 * sizeof(int32_t)));
    HIPCHK(h2d(M.iu, F->iu, n2 * sizeof(int32_t)));
    HIPCHK(h2d(M.iv, F->iv, n2 * sizeof(int32_t)));
    HIPCHK(h2d(M.p, F->p, n3p * sizeof(double)));
    HIPCHK(h2d(M.pu, F->pu, n3p * sizeof(double)));
    HIPCHK(h2d(M.pv, F->pv, n3p * sizeof(double)));
    HIPCHK(level(n, true));
    if (raflag) {
      HIPCHK(level(m, true));
      HIPCHK(h2d(M.dpo, F->dpo, n4 * sizeof(double)));
    }
    hyb_fields D;
    D.dp = M.dp; D.dpo = raflag ? M.dpo : nullptr; D.dpu = M.dpu; D.dpv = M.dpv;
    D.temp = M.temp; D.saln = M.saln; D.th3d = M.th3d; D.u = M.u; D.v = M.v;
    D.tracer = s.ntr > 0 ? M.tracer : nullptr;
    D.p = M.p; D.pu = M.pu; D.pv = M.pv; D.theta = M.theta;
    D.depths = M.depths; D.topiso = M.topiso; D.depthu = M.depthu; D.depthv = M.depthv;
    D.ubavg = M.ubavg; D.vbavg = M.vbavg; D.ip = M.ip; D.iu = M.iu; D.iv = M.iv;
    const int drc = hycom_hybgen_device(P, &D, m, n, raflag, st);
    if (drc) return drc;
    HIPCHK(level(n, false));
    if (raflag) HIPCHK(level(m, false));
    HIPCHK(d2h(F->p, M.p, n3p * sizeof(double)));
    HIPCHK(d2h(F->pu, M.pu, n3p * sizeof(double)));
    HIPCHK(d2h(F->pv, M.pv, n3p * sizeof(double)));
    HIPCHK(X.finish());
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  };
  rc = call();
  if (rc) {
    // copies queued before the failure may still read or fill the staging
    // buffers: let the stream drain before the next call reuses them
    (void)hipStreamSynchronize(st);
    X.drop();
    return rc;
  }
  M.up = up;
  M.down = down;
  return 0;
}

int hycom_hybgen_copy_invalidate(void) { return 0; }  // nothing is cached

int hycom_hybgen_copy_stats(int64_t* bytes_up, int64_t* bytes_down) {
  std::lock_guard<std::mutex> lk(g_mirror_mu);
  if (bytes_up) *bytes_up = g_mirror.up;
  if (bytes_down) *bytes_down = g_mirror.down;
  return 0;
}

// ---------------------------------------------------------------------
// synthetic inputs
// ---------------------------------------------------------------------
static double inv_normal_cdf(double pr) {
  double lo = -10.0, hi = 10.0;
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    if (0.5 * erfc(-mid / sqrt(2.0)) < pr) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

static int synth_domain(const hyb_params* P, const hyb_synth* S, SynthDomain* D) {
  if (!P || !S) return fail(HYB_ENULL, "NULL argument");
  if (P->kdm < 4 || P->kdm > HYB_KMAX) return fail(HYB_EUNSUPPORTED, "kdm=%d", P->kdm);
  if (P->ntracr > 16) return fail(HYB_EUNSUPPORTED, "generator supports ntracr<=16");
  if (S->itdm < 1 || S->jtdm < 1) return fail(HYB_EINVAL, "global extents");
  memset(D, 0, sizeof *D);
  D->seed = (uint64_t)S->seed;
  D->itdm = S->itdm;
  D->jtdm = S->jtdm;
  D->i0 = S->i0;
  D->j0 = S->j0;
  D->periodic_i = !(P->nreg == 0 || P->nreg == 4);
  D->periodic_j = P->nreg > 2;
  D->K = P->kdm;
  D->ntracr = P->ntracr;
  D->extreme = S->extreme != 0;
  D->topiso_pa = S->topiso_pa;
  D->dmin = S->depth_min;
  D->dmax = S->depth_max;
  D->shelf = S->shelf_frac > 0.0 ? S->shelf_frac : 0.0;
  D->sig_top = S->sigma_top;
  D->sig_bot = S->sigma_bot;
  D->uv_std = S->uv_std;
  D->ts_noise = S->ts_noise;
  D->thbase = P->thbase;
  for (int t = 0; t < P->ntracr; t++) D->trcflg[t] = P->trcflg[t];
  double ssq = 0.0;
  for (int q = 0; q < kSynthModes; q++) {
    uint64_t h = hash4((uint64_t)S->seed + 1, 1000 + q, 0, 0);
    int kx = (int)(mix64(h) % 9) - 4;
    int ky = (int)(mix64(h + 1) % 7) - 3;
    if (kx == 0 && ky == 0) kx = 1;
    double a = (0.5 + u01(mix64(h + 2))) / pow(1.0 + kx * kx + ky * ky, 0.75);
    D->kx[q] = kx;
    D->ky[q] = ky;
    D->amp[q] = a;
    D->ph[q] = kTwoPi * u01(mix64(h + 3));
    ssq += a * a;
  }
  double sd = sqrt(0.5 * ssq);
  for (int q = 0; q < kSynthModes; q++) D->amp[q] = D->amp[q] / sd;
  D->thr = inv_normal_cdf(S->land_frac);
  D->hmax = 2.5 - D->thr;
  return 0;
}

int hycom_hybgen_synth_host(const hyb_params* P, const hyb_synth* S, const hyb_fields* F,
                            int m, int n) {
  SynthDomain D;
  int rc = synth_domain(P, S, &D);
  if (rc) return rc;
  if ((rc = check_fields(P, F, 0))) return rc;
  View V = view_of(P, F);
  const int ii = P->ii, jj = P->jj;
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j <= jj + 1; j++)
    for (int i = 0; i <= ii + 1; i++) synth_point_pass1(D, V, ii, jj, i, j, m, n);
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 1; j <= jj; j++)
    for (int i = 1; i <= ii; i++) synth_point_pass2(D, V, ii, jj, i, j, m, n);
  return 0;
}

int hycom_hybgen_synth_device(const hyb_params* P, const hyb_synth* S, const hyb_fields* F,
                              int m, int n, void* stream) {
  SynthDomain D;
  int rc = synth_domain(P, S, &D);
  if (rc) return rc;
  if ((rc = check_fields(P, F, 0))) return rc;
  if ((rc = device_ready())) return rc;
  View V = view_of(P, F);
  HIPCHK(launch_synth(D, V, P->ii, P->jj, m, n, (hipStream_t)stream));
  return 0;
}

}  // extern "C"

// hyb_za.cpp -- HYCOM .a/.b files (hycom_za_read*, hycom_restart_read;
// README.src.mod_za:46-60, mod_za_sm.h:678-862, machi_c.c:50-67,
// mod_restart.F90:14-651).
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "hyb_host.h"
#include "hyb_math.h"

namespace hyb {
hipError_t launch_za_widen(const float* src, double* out, int idm, int jdm, int nbdy,
                           hipStream_t s);
}

using namespace hyb;

namespace {

constexpr float kVoid = 1.2676506002282294e30f;  // 2^100, the data-void marker

int64_t words_per_record(int idm, int jdm) {
  return (((int64_t)idm * jdm + 4095) / 4096) * 4096;
}

uint32_t swap4(uint32_t x) {
  return ((x >> 24) & 0xffu) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24);
}

// an open .a file read record by record into native 32-bit words
struct AFile {
  int fd = -1;
  int idm, jdm;
  int64_t nw;
  std::vector<uint32_t> buf;
  ~AFile() {
    if (fd >= 0) close(fd);
  }
  int open_(const char* path, int i, int j) {
    idm = i;
    jdm = j;
    nw = words_per_record(idm, jdm);
    fd = open(path, O_RDONLY);
    if (fd < 0) return fail(HYB_EINVAL, "cannot open %s", path);
    buf.resize((size_t)idm * jdm);
    return 0;
  }
  // record r (1-based) into buf, byte order swapped to native
  int read(int r, const char* path) {
    const size_t want = buf.size() * 4;
    const off_t off = (off_t)(r - 1) * nw * 4;
    size_t got = 0;
    while (got < want) {
      const ssize_t n = pread(fd, (char*)buf.data() + got, want - got, off + (off_t)got);
      if (n <= 0) return fail(HYB_EINVAL, "%s: record %d is short or missing", path, r);
      got += (size_t)n;
    }
    for (uint32_t& x : buf) x = swap4(x);
    return 0;
  }
  const float* f() const { return (const float*)buf.data(); }
};

// zaiord's range: over the points read, voids left out (with a mask, a void
// at a masked point is the reference's error at mod_za_sm.h:842-851)
int record_range(const float* w, int idm, int jdm, int nbdy, const int32_t* mask, double* mn,
                 double* mx, int r, const char* path) {
  const int64_t NI = (int64_t)idm + 2 * nbdy;
  float lo = kVoid, hi = -kVoid;
  for (int j = 1; j <= jdm; j++)
    for (int i = 1; i <= idm; i++) {
      const float x = w[(i - 1) + (int64_t)(j - 1) * idm];
      if (mask) {
        if (mask[(i - 1 + nbdy) + NI * (j - 1 + nbdy)] == 0) continue;
      } else if (x == kVoid) {
        continue;
      }
      lo = x < lo ? x : lo;
      hi = x > hi ? x : hi;
    }
  if (mask && hi == kVoid)
    return fail(HYB_EINVAL, "%s: record %d has 2^100 inside the mask", path, r);
  if (mn) *mn = (double)lo;
  if (mx) *mx = (double)hi;
  return 0;
}

int za_check(const char* path, int idm, int jdm, int nbdy, int rec0, int nrec, void* out) {
  if (!path || !out) return fail(HYB_EINVAL, "null path or output");
  if (idm < 1 || jdm < 1 || nbdy < 0 || rec0 < 1 || nrec < 1)
    return fail(HYB_EINVAL, "idm=%d jdm=%d nbdy=%d rec0=%d nrec=%d", idm, jdm, nbdy, rec0, nrec);
  return 0;
}

// one .b line of a restart: the name (columns 1-8) and the range after '='
struct BLine {
  std::string name;
  double hmin = 0.0, hmax = 0.0;
};

// A restart pair: the .b header and one line per .a record (restart_in,
// mod_restart.F90:45-77, and restart_in3d, 581-617).
struct RestartPair {
  const char* apath = nullptr;
  const char* bpath = nullptr;
  const hyb_params* P = nullptr;
  bool old_format = false;
  int32_t nstep = 0;
  double dtime = 0.0;
  int64_t NI = 0, NIJ = 0;
  std::vector<BLine> rec;

  int open_(const char* a, const char* b, const hyb_params* p) {
    apath = a;
    bpath = b;
    P = p;
    FILE* fb = fopen(bpath, "r");
    if (!fb) return fail(HYB_EINVAL, "cannot open %s", bpath);
    std::vector<std::string> lines;
    char line[512];
    while (fgets(line, sizeof line, fb)) lines.emplace_back(line);
    fclose(fb);
    if (lines.size() < 2) return fail(HYB_EINVAL, "%s: no restart header", bpath);
    // header: the type, then 'nstep,dtime = ...'
    if (lines[0].compare(0, 9, "RESTART: ") == 0) old_format = true;
    else if (lines[0].compare(0, 9, "RESTART2:") == 0) old_format = false;
    else return fail(HYB_EINVAL, "%s: unknown restart type", bpath);
    const size_t q = lines[1].find('=');
    long ns = 0;
    if (q == std::string::npos || sscanf(lines[1].c_str() + q + 1, "%ld %lf", &ns, &dtime) != 2)
      return fail(HYB_EINVAL, "%s: bad nstep,dtime line", bpath);
    nstep = (int32_t)ns;
    for (size_t k = 2; k < lines.size(); k++) {
      BLine bl;
      std::string& L = lines[k];
      while (!L.empty() && (L.back() == '\n' || L.back() == '\r')) L.pop_back();
      bl.name = L.substr(0, 8);
      bl.name.resize(8, ' ');
      const size_t e = L.find('=');
      int layer = 0, level = 0;
      if (e != std::string::npos)
        sscanf(L.c_str() + e + 1, "%d %d %lf %lf", &layer, &level, &bl.hmin, &bl.hmax);
      rec.push_back(bl);
    }
    NI = (int64_t)P->idm + 2 * P->nbdy;
    NIJ = NI * ((int64_t)P->jdm + 2 * P->nbdy);
    return 0;
  }

  // the first n records named `name`, which must be consecutive, into dst
  int field(const char* name, int n, double* dst) const {
    if (!dst) return 0;
    std::string nm(name);
    nm.resize(8, ' ');
    size_t first = rec.size();
    for (size_t r = 0; r < rec.size(); r++)
      if (rec[r].name == nm) {
        first = r;
        break;
      }
    if (first + n > rec.size()) return fail(HYB_EINVAL, "%s: no %d records '%s'", bpath, n, name);
    for (int k = 0; k < n; k++)
      if (rec[first + k].name != nm)
        return fail(HYB_EINVAL, "%s: record %d is '%s', expected '%s'", bpath, (int)(first + k + 1),
                    rec[first + k].name.c_str(), name);
    std::vector<double> mn(n, 0.0), mx(n, 0.0);
    int rc = hycom_za_read(apath, P->idm, P->jdm, P->nbdy, (int)first + 1, n, nullptr, dst,
                           mn.data(), mx.data());
    if (rc) return rc;
    for (int k = 0; k < n; k++) {  // restart_in3d's consistency check (604-616)
      const BLine& b = rec[first + k];
      if (fabs(mn[k] - b.hmin) > fabs(b.hmin) * 1.e-4 || fabs(mx[k] - b.hmax) > fabs(b.hmax) * 1.e-4)
        return fail(HYB_EINVAL, "%s and %s not consistent at record %d ('%s')", apath, bpath,
                    (int)(first + k + 1), name);
    }
    return 0;
  }
};

}  // namespace

extern "C" {

int hycom_za_read(const char* apath, int32_t idm, int32_t jdm, int32_t nbdy, int32_t rec0,
                  int32_t nrec, const int32_t* mask, double* out, double* hmin, double* hmax) {
  int rc = za_check(apath, idm, jdm, nbdy, rec0, nrec, out);
  if (rc) return rc;
  AFile A;
  if ((rc = A.open_(apath, idm, jdm))) return rc;
  const int64_t NI = (int64_t)idm + 2 * nbdy, NIJ = NI * ((int64_t)jdm + 2 * nbdy);
  for (int r = 0; r < nrec; r++) {
    if ((rc = A.read(rec0 + r, apath))) return rc;
    const float* w = A.f();
    if ((rc = record_range(w, idm, jdm, nbdy, mask, hmin ? hmin + r : nullptr,
                           hmax ? hmax + r : nullptr, rec0 + r, apath)))
      return rc;
    double* o = out + NIJ * r;
    for (int j = 1; j <= jdm; j++)
      for (int i = 1; i <= idm; i++) {
        const int64_t a = (i - 1 + nbdy) + NI * (j - 1 + nbdy);
        if (mask && mask[a] == 0) continue;
        o[a] = (double)w[(i - 1) + (int64_t)(j - 1) * idm];
      }
  }
  return 0;
}

int hycom_za_read_device(const char* apath, int32_t idm, int32_t jdm, int32_t nbdy, int32_t rec0,
                         int32_t nrec, const int32_t* mask, double* out, double* hmin,
                         double* hmax, void* stream) {
  int rc = za_check(apath, idm, jdm, nbdy, rec0, nrec, out);
  if (rc) return rc;
  if (mask) return fail(HYB_EUNSUPPORTED, "hycom_za_read_device takes no mask");
  if ((rc = device_ready())) return rc;
  AFile A;
  if ((rc = A.open_(apath, idm, jdm))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t NI = (int64_t)idm + 2 * nbdy, NIJ = NI * ((int64_t)jdm + 2 * nbdy);
  const size_t bytes = (size_t)idm * jdm * 4;
  float* stage = nullptr;  // pinned, two records: one in flight while the next is read
  float* dev = nullptr;
  HIPCHK(hipHostMalloc((void**)&stage, 2 * bytes, hipHostMallocDefault));
  hipError_t e = hipMalloc((void**)&dev, 2 * bytes);
  hipEvent_t done[2] = {nullptr, nullptr};
  if (e == hipSuccess) e = hipEventCreateWithFlags(&done[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&done[1], hipEventDisableTiming);
  for (int r = 0; r < nrec && e == hipSuccess && rc == 0; r++) {
    const int b = r & 1;
    if (r >= 2) e = hipEventSynchronize(done[b]);  // staging slot b free again
    if (e != hipSuccess) break;
    if ((rc = A.read(rec0 + r, apath))) break;
    if ((rc = record_range(A.f(), idm, jdm, nbdy, nullptr, hmin ? hmin + r : nullptr,
                           hmax ? hmax + r : nullptr, rec0 + r, apath)))
      break;
    float* hs = stage + (size_t)b * idm * jdm;
    float* ds = dev + (size_t)b * idm * jdm;
    memcpy(hs, A.f(), bytes);
    e = hipMemcpyAsync(ds, hs, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_za_widen(ds, out + NIJ * r, idm, jdm, nbdy, s);
    if (e == hipSuccess) e = hipEventRecord(done[b], s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (hipEvent_t ev : done)
    if (ev) (void)hipEventDestroy(ev);
  if (dev) (void)hipFree(dev);
  (void)hipHostFree(stage);
  if (rc) return rc;
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? HYB_ENOMEM : HYB_EDEVICE, "za read: %s",
                hipGetErrorString(e));
  return 0;
}

int hycom_restart_read(const char* apath, const char* bpath, const hyb_params* P,
                       const hyb_fields* F, int32_t* nstep, double* dtime) {
  if (!apath || !bpath || !P || !F) return fail(HYB_EINVAL, "null argument");
  const int kdm = P->kdm, ntr = P->ntracr;
  if (kdm < 1 || kdm > HYB_KMAX) return fail(HYB_EINVAL, "kdm=%d", kdm);
  RestartPair R;
  int rc0 = R.open_(apath, bpath, P);
  if (rc0) return rc0;
  if (nstep) *nstep = R.nstep;
  if (dtime) *dtime = R.dtime;
  const bool old_format = R.old_format;
  const int64_t NI = R.NI, NIJ = R.NIJ;
  auto field = [&](const char* name, int n, double* dst) -> int { return R.field(name, n, dst); };
  int rc;
  if ((rc = field("u", 2 * kdm, F->u))) return rc;
  if ((rc = field("v", 2 * kdm, F->v))) return rc;
  if ((rc = field("dp", 2 * kdm, F->dp))) return rc;
  if ((rc = field("temp", 2 * kdm, F->temp))) return rc;
  if ((rc = field("saln", 2 * kdm, F->saln))) return rc;
  if (old_format) {
    if ((rc = field("th3d", 2 * kdm, F->th3d))) return rc;
  } else if (F->th3d) {  // restart_in 83-97
    if (!F->ip || !F->temp || !F->saln)
      return fail(HYB_EINVAL, "th3d from temp and saln needs ip, temp and saln");
    for (int l = 0; l < 2; l++)
      for (int k = 0; k < kdm; k++)
        for (int j = 1; j <= P->jj; j++)
          for (int i = 1; i <= P->ii; i++) {
            const int64_t a = (i - 1 + P->nbdy) + NI * (j - 1 + P->nbdy);
            const int64_t o = a + NIJ * ((int64_t)k + (int64_t)kdm * l);
            F->th3d[o] = (F->ip[a] == 1) ? sig(F->temp[o], F->saln[o]) - P->thbase : 0.0;
          }
  }
  if ((rc = field("ubavg", 3, (double*)F->ubavg))) return rc;
  if ((rc = field("vbavg", 3, (double*)F->vbavg))) return rc;
  if (ntr > 0 && F->tracer)
    if ((rc = field("tracer", 2 * kdm * ntr, F->tracer))) return rc;
  return 0;
}

int hycom_restart_read_field(const char* apath, const char* bpath, const hyb_params* P,
                             const char* name, int32_t nrec, double* out) {
  if (!apath || !bpath || !P || !name || !out) return fail(HYB_EINVAL, "null argument");
  if (nrec < 1) return fail(HYB_EINVAL, "nrec=%d", nrec);
  RestartPair R;
  int rc = R.open_(apath, bpath, P);
  if (rc) return rc;
  return R.field(name, nrec, out);
}

}  // extern "C"

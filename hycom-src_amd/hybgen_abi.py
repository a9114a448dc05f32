"""ctypes mirror of include/hycom_hybgen.h plus field-array helpers.

Host fields are numpy arrays in HYCOM's mod_cb_arrays layout (Fortran order,
halo included); device fields are flat torch tensors of the same sizes
(torch is used for device memory and streams only).  The compute path is the
HIP library; loading fails loudly when it is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhycom_hybgen.so")

KMAX = 64
TRMAX = 99
PREC_FP64 = 0
PREC_MIXED = 1

HYB_OK = 0
ERRORS = {-1: "EINVAL", -2: "EUNSUPPORTED", -3: "EDEVICE", -4: "ENOMEM",
          -5: "ENODEVICE", -6: "ENULL", -7: "ENONFINITE", -8: "EBUSY"}
for _code, _name in ERRORS.items():
    globals()["HYB_" + _name] = _code

_INT_FIELDS = ["idm", "jdm", "ii", "jj", "nbdy", "kdm", "ntracr", "nhybrd",
               "nsigma", "hybmap", "hybflg", "nreg", "isopcm", "mxlmy",
               "ldenfix", "itracr801", "itracr802", "itracr803", "precision",
               "check_finite"]
_DBL_FIELDS = ["thbase", "qhybrlx", "hybiso", "hybthn", "hybthk", "thkbot",
               "dp00i", "dpns", "dsns", "baclin", "ra2fac", "vland"]


class HybParams(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int32) for n in _INT_FIELDS]
                + [("trcflg", ctypes.c_int32 * TRMAX)]
                + [(n, ctypes.c_double) for n in _DBL_FIELDS]
                + [("dp0k", ctypes.c_double * KMAX),
                   ("ds0k", ctypes.c_double * KMAX),
                   ("dx0k", ctypes.c_double * KMAX)])


FIELD_NAMES = ["dp", "dpo", "dpu", "dpv", "temp", "saln", "th3d", "u", "v",
               "tracer", "p", "pu", "pv", "theta", "depths", "topiso",
               "depthu", "depthv", "ubavg", "vbavg", "ip", "iu", "iv"]


class HybFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in FIELD_NAMES]


class HybSynth(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64),
                ("itdm", ctypes.c_int32), ("jtdm", ctypes.c_int32),
                ("i0", ctypes.c_int32), ("j0", ctypes.c_int32),
                ("depth_min", ctypes.c_double), ("depth_max", ctypes.c_double),
                ("land_frac", ctypes.c_double), ("shelf_frac", ctypes.c_double),
                ("sigma_top", ctypes.c_double), ("sigma_bot", ctypes.c_double),
                ("uv_std", ctypes.c_double), ("ts_noise", ctypes.c_double),
                ("topiso_pa", ctypes.c_double),
                ("extreme", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class HybDomain(ctypes.Structure):
    _fields_ = [("P", ctypes.c_void_p), ("F", ctypes.c_void_p)]


class HybServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("max_active", ctypes.c_int32),
                ("remap_cols", ctypes.c_int32), ("regrid_cols", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class HybCompletion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int64), ("domain", ctypes.c_int32), ("status", ctypes.c_int32),
                ("nsteps", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_done", ctypes.c_double), ("column_layers", ctypes.c_int64)]


class HybServerStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("column_layers", ctypes.c_int64), ("resident_bytes", ctypes.c_int64),
                ("budget_bytes", ctypes.c_int64), ("opens", ctypes.c_int64),
                ("evictions", ctypes.c_int64), ("resident_domains", ctypes.c_int32),
                ("max_active_seen", ctypes.c_int32), ("active_sum", ctypes.c_double)]


class HybStokesParams(ctypes.Structure):
    _fields_ = [("idm", ctypes.c_int32), ("jdm", ctypes.c_int32), ("ii", ctypes.c_int32),
                ("jj", ctypes.c_int32), ("nbdy", ctypes.c_int32), ("kdm", ctypes.c_int32),
                ("nsdzi", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("ws0", ctypes.c_double), ("ws1", ctypes.c_double)]


class HybStokesFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("usdz", "vsdz", "sdzi", "dp", "p", "ip", "usdp", "vsdp")]


class HybTrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("idm", "jdm", "ii", "jj", "nbdy", "ki", "ko", "ks", "method", "reserved")]


class HybTrcFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("si", "pi", "po", "mask", "so")]


class HybStartFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("pbot", "dp", "ip", "iu", "iv", "depthu", "depthv", "p", "dpu", "dpv")]


# every symbol include/hycom_hybgen.h declares, with its ctypes signature
_P = ctypes.POINTER
SIGNATURES = {
    "hycom_hybgen_abi_version": (ctypes.c_int, []),
    "hycom_hybgen_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "hycom_hybgen_last_error": (ctypes.c_char_p, []),
    "hycom_hybgen_check_params": (ctypes.c_int, [_P(HybParams)]),
    "hycom_hybgen_derive_vgrid": (ctypes.c_int, [_P(HybParams)] + [ctypes.c_double] * 12
                                  + [_P(ctypes.c_double)]),
    "hycom_hybgen": (ctypes.c_int, [_P(HybParams), _P(HybFields), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]),
    "hycom_hybgen_device": (ctypes.c_int, [_P(HybParams), _P(HybFields), ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "hycom_hybgen_copy_invalidate": (ctypes.c_int, []),
    "hycom_hybgen_copy_stats": (ctypes.c_int, [_P(ctypes.c_int64), _P(ctypes.c_int64)]),
    "hycom_hybgen_device_noalloc": (ctypes.c_int, [_P(HybParams), _P(HybFields), ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "hycom_hybgen_noalloc_capacity": (ctypes.c_size_t, []),
    "hycom_hybgen_noalloc_need": (ctypes.c_size_t, [_P(HybParams), ctypes.c_int]),
    "hycom_hybgen_record_decisions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "hycom_hybgen_selftest_quotients": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 4
                                        + [ctypes.c_int64, ctypes.c_void_p]),
    "hycom_hybgen_set_kernel_path": (ctypes.c_int, [ctypes.c_int]),
    "hycom_hybgen_tune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "hycom_hybgen_timing": (ctypes.c_int, [ctypes.c_int]),
    "hycom_hybgen_timing_read": (ctypes.c_int, [_P(ctypes.c_double), _P(ctypes.c_int32),
                                                ctypes.c_int]),
    "hycom_hybgen_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_void_p]),
    "hycom_hybgen_batch_autotune": (ctypes.c_int, [ctypes.c_int]),
    "hycom_hybgen_batch_config": (ctypes.c_int, [_P(ctypes.c_int32), _P(ctypes.c_int32),
                                                 _P(ctypes.c_int32)]),
    "hycom_hybgen_batch_staged_levels": (ctypes.c_int, [_P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "hycom_hybgen_batch_tuning": (ctypes.c_int, [_P(ctypes.c_int32), _P(ctypes.c_int32),
                                                 _P(ctypes.c_int32), _P(ctypes.c_float),
                                                 ctypes.c_int]),
    "hycom_hybgen_server_create": (ctypes.c_int, [_P(HybServerConfig), _P(ctypes.c_void_p)]),
    "hycom_hybgen_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "hycom_hybgen_server_add_domain": (ctypes.c_int, [ctypes.c_void_p, _P(HybParams),
                                                      _P(HybSynth), _P(HybFields),
                                                      _P(ctypes.c_int32)]),
    "hycom_hybgen_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                  ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.c_int64]),
    "hycom_hybgen_server_poll": (ctypes.c_int, [ctypes.c_void_p, _P(HybCompletion),
                                                ctypes.c_int32, ctypes.c_int32,
                                                _P(ctypes.c_int32)]),
    "hycom_hybgen_server_read_domain": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                       _P(HybFields)]),
    "hycom_hybgen_server_stats": (ctypes.c_int, [ctypes.c_void_p, _P(HybServerStats)]),
    "hycom_hybgen_synth_host": (ctypes.c_int, [_P(HybParams), _P(HybSynth), _P(HybFields),
                                               ctypes.c_int, ctypes.c_int]),
    "hycom_hybgen_synth_device": (ctypes.c_int, [_P(HybParams), _P(HybSynth), _P(HybFields),
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "hycom_stokes_vertical": (ctypes.c_int, [_P(HybStokesParams), _P(HybStokesFields),
                                             ctypes.c_int]),
    "hycom_stokes_vertical_device": (ctypes.c_int, [_P(HybStokesParams), _P(HybStokesFields),
                                                    ctypes.c_int, ctypes.c_void_p]),
    "hycom_trc_remap": (ctypes.c_int, [_P(HybTrcParams), _P(HybTrcFields)]),
    "hycom_za_read": (ctypes.c_int, [ctypes.c_char_p] + [ctypes.c_int32] * 5
                      + [ctypes.c_void_p] * 4),
    "hycom_za_read_device": (ctypes.c_int, [ctypes.c_char_p] + [ctypes.c_int32] * 5
                             + [ctypes.c_void_p] * 5),
    "hycom_restart_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _P(HybParams),
                                          _P(HybFields), ctypes.c_void_p, ctypes.c_void_p]),
    "hycom_restart_read_field": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _P(HybParams),
                                                ctypes.c_char_p, ctypes.c_int32,
                                                ctypes.c_void_p]),
    "hycom_grid_masks": (ctypes.c_int, [ctypes.c_int32] * 3 + [ctypes.c_void_p] * 5),
    "hycom_state_start": (ctypes.c_int, [_P(HybParams), _P(HybStartFields)]),
    "hycom_trc_remap_device": (ctypes.c_int, [_P(HybTrcParams), _P(HybTrcFields),
                                              ctypes.c_void_p]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with "
                           "`python hycom-src_amd/build_lib.py` (no fallback path exists)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HybError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hybgen error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def check(rc):
    if rc != HYB_OK:
        raise HybError(rc, load().hycom_hybgen_last_error().decode())
    return rc


# ---------------------------------------------------------------------------
# parameter presets for the documented workloads
# ---------------------------------------------------------------------------
WORKLOADS = {
    # 150x100x32, 2 tracers, sigma-2 28.10 -> 37.00, PPM
    "cfg1": dict(idm=150, jdm=100, kdm=32, ntracr=2, trcflg=[2, 0], hybmap=2,
                 sigma=(28.10, 37.00), depth=(50.0, 5500.0), land=0.30, shelf=0.10,
                 dp00=3.0, dp00f=1.125, dp00x=12.0, ds00=1.0, ds00f=1.125, ds00x=12.0,
                 nsigma=8, nreg=0),
    # GLBb 1/12 deg shape: 4500x3298x41, 4 tracers, sigma-2 17.00 -> 37.48
    "glbb": dict(idm=4500, jdm=3298, kdm=41, ntracr=4, trcflg=[0, 0, 0, 0], hybmap=2,
                 sigma=(17.00, 37.48), depth=(5.0, 6500.0), land=0.35, shelf=0.10,
                 dp00=1.0, dp00f=1.18, dp00x=450.0, ds00=1.0, ds00f=1.18, ds00x=450.0,
                 nsigma=14, nreg=1),
}


def make_params(kdm=32, idm=150, jdm=100, ntracr=2, trcflg=(2, 0), hybmap=2,
                nhybrd=None, nsigma=8, nreg=0, dp00=3.0, dp00f=1.125, dp00x=12.0,
                ds00=1.0, ds00f=1.125, ds00x=12.0, dp00i=1.0, dx00=9999.0,
                dx00f=1.0, dx00x=9999.0, hybrlx=8.0, hybiso=0.01, hybthn=1.0,
                hybthk=0.0, thkbot=10.0, isotop=6.0, thbase=34.0, ldenfix=0,
                nbdy=6, ii=None, jj=None, precision=PREC_FP64, baclin=300.0,
                ra2fac=0.125, **_unused):
    """hyb_params from blkdat.input-style values (geopar derivation in C)."""
    P = HybParams()
    P.idm, P.jdm = idm, jdm
    P.ii = idm if ii is None else ii
    P.jj = jdm if jj is None else jj
    P.nbdy = nbdy
    P.kdm = kdm
    P.ntracr = ntracr
    for t, f in enumerate(trcflg[:ntracr]):
        P.trcflg[t] = f
    P.nhybrd = kdm if nhybrd is None else nhybrd
    P.nsigma = nsigma
    P.hybmap = hybmap
    P.hybflg = 0
    P.nreg = nreg
    P.ldenfix = ldenfix
    P.precision = precision
    P.thbase = thbase
    P.hybiso = hybiso
    P.hybthn = hybthn
    P.hybthk = hybthk
    P.thkbot = thkbot
    P.baclin = baclin
    P.ra2fac = ra2fac
    P.vland = 0.0
    topiso = ctypes.c_double(0.0)
    check(load().hycom_hybgen_derive_vgrid(
        ctypes.byref(P), dp00, dp00f, dp00x, ds00, ds00f, ds00x, dp00i,
        dx00, dx00f, dx00x, hybrlx, isotop, ctypes.byref(topiso)))
    P._topiso_pa = topiso.value
    return P


def workload_params(name, **over):
    w = dict(WORKLOADS[name])
    w.update(over)
    P = make_params(**w)
    return P, w


def make_synth(P, w, seed=100, i0=0, j0=0, itdm=None, jtdm=None, extreme=0):
    S = HybSynth()
    S.seed = seed
    S.itdm = P.idm if itdm is None else itdm
    S.jtdm = P.jdm if jtdm is None else jtdm
    S.i0, S.j0 = i0, j0
    S.depth_min, S.depth_max = w["depth"]
    S.land_frac = w["land"]
    S.shelf_frac = w.get("shelf", 0.1)
    S.sigma_top, S.sigma_bot = w["sigma"]
    S.uv_std = 0.1
    S.ts_noise = 0.05
    S.topiso_pa = P._topiso_pa
    S.extreme = extreme
    return S


# ---------------------------------------------------------------------------
# field storage
# ---------------------------------------------------------------------------
def tracer_extent(P):
    """Tracers the tracer array holds: the ntracr remapped ones and the
    diagnostic tracers 801-803 wherever their indices put them (the model
    keeps them after the ntracr tracers)."""
    return max(P.ntracr, P.itracr801, P.itracr802, P.itracr803)


def field_shapes(P):
    NI = P.idm + 2 * P.nbdy
    NJ = P.jdm + 2 * P.nbdy
    K = P.kdm
    ntr = max(tracer_extent(P), 1)
    s4 = (NI, NJ, K, 2)
    return {
        "dp": s4, "dpo": s4, "dpu": s4, "dpv": s4, "temp": s4, "saln": s4,
        "th3d": s4, "u": s4, "v": s4, "tracer": (NI, NJ, K, 2, ntr),
        "p": (NI, NJ, K + 1), "pu": (NI, NJ, K + 1), "pv": (NI, NJ, K + 1),
        "theta": (NI, NJ, K), "depths": (NI, NJ), "topiso": (NI, NJ),
        "depthu": (NI, NJ), "depthv": (NI, NJ), "ubavg": (NI, NJ, 3),
        "vbavg": (NI, NJ, 3), "ip": (NI, NJ), "iu": (NI, NJ), "iv": (NI, NJ),
    }


INT_ARRAYS = ("ip", "iu", "iv")
OUTPUTS = ("dp", "dpu", "dpv", "temp", "saln", "th3d", "u", "v", "tracer",
           "p", "pu", "pv")


class HostDomain:
    """All hybgen fields of one domain as Fortran-ordered numpy arrays."""

    def __init__(self, P, arrays=None):
        self.P = P
        shapes = field_shapes(P)
        if arrays is None:
            arrays = {}
            for n, s in shapes.items():
                dt = np.int32 if n in INT_ARRAYS else np.float64
                arrays[n] = np.zeros(s, dtype=dt, order="F")
        self.a = arrays

    def copy(self):
        return HostDomain(self.P, {n: v.copy(order="F") for n, v in self.a.items()})

    def fields(self):
        F = HybFields()
        for n in FIELD_NAMES:
            setattr(F, n, self.a[n].ctypes.data)
        if tracer_extent(self.P) == 0:
            F.tracer = None
        return F

    def synth(self, S, m=1, n=2):
        F = self.fields()
        check(load().hycom_hybgen_synth_host(ctypes.byref(self.P), ctypes.byref(S),
                                             ctypes.byref(F), m, n))
        return self

    def hybgen(self, m=1, n=2, raflag=0):
        """Copy-mode drop-in call through the C ABI (needs a GPU)."""
        F = self.fields()
        check(load().hycom_hybgen(ctypes.byref(self.P), ctypes.byref(F), m, n, raflag))
        return self


class DeviceDomain:
    """All hybgen fields of one domain as flat device tensors."""

    def __init__(self, P, device="cuda", skip=()):
        import torch
        self.P = P
        self.t = {}
        for n, s in field_shapes(P).items():
            if n in skip:
                continue
            dt = torch.int32 if n in INT_ARRAYS else torch.float64
            self.t[n] = torch.zeros(int(np.prod(s)), dtype=dt, device=device)

    def fields(self):
        F = HybFields()
        for n in FIELD_NAMES:
            setattr(F, n, self.t[n].data_ptr() if n in self.t else None)
        if tracer_extent(self.P) == 0:
            F.tracer = None
        return F

    def upload(self, host):
        import torch
        for n, v in host.a.items():
            if n not in self.t:
                continue
            self.t[n].copy_(torch.from_numpy(np.ascontiguousarray(v.reshape(-1, order="F"))))
        return self

    def download(self, names=None):
        shapes = field_shapes(self.P)
        out = {}
        for n in (names or [k for k in shapes if k in self.t]):
            out[n] = self.t[n].cpu().numpy().reshape(shapes[n], order="F")
        return HostDomain(self.P, out) if names is None else out

    def synth(self, S, m=1, n=2, stream=None):
        F = self.fields()
        check(load().hycom_hybgen_synth_device(ctypes.byref(self.P), ctypes.byref(S),
                                               ctypes.byref(F), m, n, stream))
        return self

    def hybgen(self, m=1, n=2, raflag=0, stream=None, noalloc=False):
        F = self.fields()
        fn = load().hycom_hybgen_device_noalloc if noalloc else load().hycom_hybgen_device
        check(fn(ctypes.byref(self.P), ctypes.byref(F), m, n, raflag, stream))
        return self


class Batch:
    """A set of DeviceDomains stepped together by hycom_hybgen_batch."""

    def __init__(self, domains):
        self.domains = list(domains)
        self._F = [d.fields() for d in self.domains]
        self._D = (HybDomain * len(self.domains))()
        for q, (d, F) in enumerate(zip(self.domains, self._F)):
            self._D[q].P = ctypes.addressof(d.P)
            self._D[q].F = ctypes.addressof(F)

    def hybgen(self, m=1, n=2, stream=None):
        check(load().hycom_hybgen_batch(ctypes.addressof(self._D), len(self.domains), m, n,
                                        stream))
        return self

    @staticmethod
    def config():
        """(remap columns, regrid columns, tuned) of the last batch."""
        a, b, t = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(load().hycom_hybgen_batch_config(ctypes.byref(a), ctypes.byref(b), ctypes.byref(t)))
        return a.value, b.value, bool(t.value)

    @staticmethod
    def tuning():
        """The autotuner's candidates of the last batch: a list of
        (remap columns, staged levels, regrid columns, ms)."""
        cap = 64
        a, b, c = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
        t = (ctypes.c_float * cap)()
        n = load().hycom_hybgen_batch_tuning(a, b, c, t, cap)
        check(min(n, 0))
        return [(a[q], b[q], c[q], round(t[q], 3)) for q in range(min(n, cap))]

    @staticmethod
    def staged_levels():
        """(interface levels staged in LDS per remap column, kmax) of the last batch."""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        check(load().hycom_hybgen_batch_staged_levels(ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Server:
    """hycom_hybgen_server_*: resident domains stepped on request, with
    continuous batching of the active requests."""

    def __init__(self, hbm_budget=0, max_active=0, remap_cols=0, regrid_cols=0):
        C = HybServerConfig(hbm_budget, max_active, remap_cols, regrid_cols, 0)
        h = ctypes.c_void_p()
        check(load().hycom_hybgen_server_create(ctypes.byref(C), ctypes.byref(h)))
        self.h = h
        self.params = {}

    def add_domain(self, P, S=None, host=None):
        """Registers a domain from a synthetic description or a HostDomain."""
        did = ctypes.c_int32()
        F = host.fields() if host is not None else None
        check(load().hycom_hybgen_server_add_domain(
            self.h, ctypes.byref(P), ctypes.byref(S) if S is not None else None,
            ctypes.byref(F) if F is not None else None, ctypes.byref(did)))
        self.params[did.value] = P
        return did.value

    def submit(self, dom, nsteps, precision=PREC_FP64, tag=0):
        check(load().hycom_hybgen_server_submit(self.h, dom, nsteps, precision, tag))

    def poll(self, max_n=256, wait_ms=0):
        buf = (HybCompletion * max_n)()
        n = ctypes.c_int32()
        check(load().hycom_hybgen_server_poll(self.h, buf, max_n, wait_ms, ctypes.byref(n)))
        return [buf[q] for q in range(n.value)]

    def read_domain(self, dom):
        d = HostDomain(self.params[dom])
        F = d.fields()
        F.dpo = None
        check(load().hycom_hybgen_server_read_domain(self.h, dom, ctypes.byref(F)))
        return d

    def stats(self):
        st = HybServerStats()
        check(load().hycom_hybgen_server_stats(self.h, ctypes.byref(st)))
        return st

    def close(self):
        if self.h:
            check(load().hycom_hybgen_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# Stokes drift onto the model layers (hycom_stokes_vertical*)
# ---------------------------------------------------------------------------
ONEM = 9.806 / (1.0 / 1000.0)        # mod_cb_arrays.F90:843-846
DPTHIN = 0.001 * (ONEM * 0.001)      # 0.001*onemm (mod_stokes.F90:670)


class StokesDomain:
    """Inputs and outputs of the Stokes drift remap (mod_stokes.F90:650-705)
    over a hybgen domain's dp (time level n), p and ip, as Fortran-ordered
    numpy arrays.  The two snapshots are seeded profiles that decay with
    depth over fixed interfaces 0, 1, 2, 3.5, ... m; the first interface is
    moved to dpthin, as stokes_forfun does (mod_stokes.F90:372-376)."""

    def __init__(self, dom, nsdzi, seed=0, ws0=0.3, n=2):
        P = dom.P
        self.dom, self.n = dom, n
        self.P = HybStokesParams(idm=P.idm, jdm=P.jdm, ii=P.ii, jj=P.jj, nbdy=P.nbdy,
                                 kdm=P.kdm, nsdzi=nsdzi, reserved=0, ws0=ws0, ws1=1.0 - ws0)
        NI, NJ = P.idm + 2 * P.nbdy, P.jdm + 2 * P.nbdy
        rng = np.random.default_rng(seed)
        z = np.concatenate([[0.0], np.cumsum(np.geomspace(1.0, 60.0, nsdzi - 1))])
        self.sdzi = np.ascontiguousarray(z * ONEM)
        self.sdzi[0] = DPTHIN
        zmid = 0.5 * (z[:-1] + z[1:])
        zmid = np.concatenate([zmid, [z[-1]]])[:nsdzi]
        amp = 0.1 * rng.standard_normal((NI, NJ, 2))
        scale = rng.uniform(3.0, 15.0, (NI, NJ, 1))
        prof = np.exp(-zmid[None, None, :] / scale)
        self.usdz = np.asfortranarray(amp[:, :, :, None] * prof[:, :, None, :]
                                      + 0.002 * rng.standard_normal((NI, NJ, 2, nsdzi)))
        amp = 0.1 * rng.standard_normal((NI, NJ, 2))
        self.vsdz = np.asfortranarray(amp[:, :, :, None] * prof[:, :, None, :]
                                      + 0.002 * rng.standard_normal((NI, NJ, 2, nsdzi)))
        self.usdp = np.asfortranarray(np.full((NI, NJ, P.kdm), -7.0))
        self.vsdp = np.asfortranarray(np.full((NI, NJ, P.kdm), 7.0))

    def copy(self):
        c = StokesDomain.__new__(StokesDomain)
        c.dom, c.n, c.P = self.dom, self.n, self.P
        for k in ("sdzi", "usdz", "vsdz", "usdp", "vsdp"):
            setattr(c, k, getattr(self, k).copy(order="K"))
        return c

    def fields(self):
        a = self.dom.a
        F = HybStokesFields()
        F.usdz, F.vsdz = self.usdz.ctypes.data, self.vsdz.ctypes.data
        F.sdzi = self.sdzi.ctypes.data
        F.dp, F.p, F.ip = a["dp"].ctypes.data, a["p"].ctypes.data, a["ip"].ctypes.data
        F.usdp, F.vsdp = self.usdp.ctypes.data, self.vsdp.ctypes.data
        return F

    def run(self):
        """Copy-mode call through the C ABI (needs a GPU)."""
        F = self.fields()
        check(load().hycom_stokes_vertical(ctypes.byref(self.P), ctypes.byref(F), self.n))
        return self

#!/usr/bin/env python3
"""hybgen throughput on one MI355X.

One step = one  hybgen(m,n,.false.)  call (column regrid+remap of T, S and
the tracers, p halo, dpudpv + u/v remap) over the whole GLBb-shaped domain,
fields resident in HBM.  Each call regrids the state the previous call
left, as successive model steps do.

Prints one JSON line.  `roofline` is for the dominant kernel (the column
kernel, hybgenaj): algorithmic bytes per launch / its mean launch duration
from HIP events recorded on its own stream.  `cpu_baseline` is the reference
hybgen (flang build of HYCOM-src, oracle/_ref) timed on this host on a
bounded window of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hycom-src_amd"))
import hybgen_abi as abi  # noqa: E402

METRIC = "column-layers remapped/s, FP64 vs mixed, GLBb tile; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
PHASES = ["column", "halo", "uv", "ra", "halo2", "uv2", "regrid", "remap"]
# kernel of each p-column remap variant (hycom_hybgen_tune knob 0)
REMAP_KERNELS = {3: "k_remap_s<{m},3,2>"}


def bytes_per_column_phase_a(K, T):
    """Algorithmic HBM bytes of the column kernel per sea column: read and
    write dp, T, S, th3d and T tracers (K levels each), read theta (K),
    write p (K+1), read depths and topiso, read ip (int32)."""
    return 8 * (2 * (4 + T) * K + K + (K + 1)) + 8 * 2 + 4


def bytes_per_column_remap(K, T):
    """Algorithmic HBM bytes of the p-column remap kernel per sea column:
    read the new and the old interfaces (K+1 each), read and write T, S and
    T tracers (K levels each), write th3d (K), read kp and the PCM mask."""
    return 8 * (2 * (K + 1) + 2 * (2 + T) * K + K) + 4 + 8


def bytes_per_column_call(K, T):
    """Algorithmic bytes of the whole call per sea column (raflag off):
    8*[(20+2T)K + 10] + 12 (scalar fields, p, u/v points, masks)."""
    return 8 * ((20 + 2 * T) * K + 10) + 12


def sea_columns(dd):
    import torch
    P = dd.P
    NI, NJ = P.idm + 2 * P.nbdy, P.jdm + 2 * P.nbdy
    ip = dd.t["ip"].view(NJ, NI)[P.nbdy:P.nbdy + P.jj, P.nbdy:P.nbdy + P.ii]
    return int(torch.count_nonzero(ip).item())


def physical_cores():
    """Physical cores among the CPUs this process may run on: distinct
    (core, socket) pairs of `lscpu -p=cpu,core,socket` whose cpu is in the
    affinity set (logical CPUs of one core count once)."""
    import subprocess
    allowed = os.sched_getaffinity(0)
    try:
        out = subprocess.run(["lscpu", "-p=cpu,core,socket"], capture_output=True, text=True,
                             check=True).stdout
    except (OSError, subprocess.CalledProcessError):
        return len(allowed), len(allowed)
    cores = set()
    for line in out.splitlines():
        if not line or line.startswith("#"):
            continue
        cpu, core, sock = (int(x) if x else -1 for x in line.split(",")[:3])
        if cpu in allowed:
            cores.add((core, sock))
    return (len(cores) or len(allowed)), len(allowed)


def cpu_baseline(workload, hybmap, rows, seed):
    """Reference hybgen (flang build of the HYCOM-src sources) on a window of
    `rows` full-width j-rows of the same global field, on this host, with
    one OpenMP thread per physical core of the affinity set."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    if not ol.ref_available():
        return None
    cores, logical = physical_cores()
    lib = ol.ref(False)
    has_omp = hasattr(lib, "omp_set_num_threads") and hasattr(lib, "omp_get_max_threads")
    nmax = lib.omp_get_max_threads() if has_omp else 0
    if has_omp:
        lib.omp_set_num_threads(cores)
    w0 = abi.WORKLOADS[workload]
    itdm, jtdm = w0["idm"], w0["jdm"]
    j0 = max(0, jtdm // 2 - rows // 2)
    P, w = abi.workload_params(workload, hybmap=hybmap, jdm=rows)
    S = abi.make_synth(P, w, seed=seed, itdm=itdm, jtdm=jtdm, i0=0, j0=j0)
    d = abi.HostDomain(P).synth(S)
    sea = int((d.a["ip"][P.nbdy:P.nbdy + P.ii, P.nbdy:P.nbdy + P.jj] > 0).sum())
    times = []
    for _ in range(3):
        times.append(ol.run_ref(d.copy()))
    times.sort()
    t = times[len(times) // 2]
    threads = cores if has_omp else logical
    # the same reference on one thread, on the first rows // 8 rows of the band
    one = None
    if has_omp:
        P1, w1 = abi.workload_params(workload, hybmap=hybmap, jdm=max(rows // 8, 1))
        d1 = abi.HostDomain(P1).synth(abi.make_synth(P1, w1, seed=seed, itdm=itdm, jtdm=jtdm,
                                                     i0=0, j0=j0))
        sea1 = int((d1.a["ip"][P1.nbdy:P1.nbdy + P1.ii, P1.nbdy:P1.nbdy + P1.jj] > 0).sum())
        lib.omp_set_num_threads(1)
        try:
            t1 = sorted(ol.run_ref(d1.copy()) for _ in range(3))[1]
        finally:
            lib.omp_set_num_threads(nmax)
        one = {"value": sea1 * P1.kdm / t1, "rows": P1.jdm, "sea_columns": sea1,
               "seconds_per_call": t1}
    # the same reference on the whole config-1 domain (150x100x32, PPM), at
    # the same thread count and on one thread
    cfg1 = None
    try:
        Pc, wc = abi.workload_params("cfg1", hybmap=hybmap)
        dc = abi.HostDomain(Pc).synth(abi.make_synth(Pc, wc, seed=seed))
        seac = int((dc.a["ip"][Pc.nbdy:Pc.nbdy + Pc.ii, Pc.nbdy:Pc.nbdy + Pc.jj] > 0).sum())
        if has_omp:
            lib.omp_set_num_threads(cores)
        tc = sorted(ol.run_ref(dc.copy()) for _ in range(5))[2]
        tc1 = None
        if has_omp:
            lib.omp_set_num_threads(1)
            try:
                tc1 = sorted(ol.run_ref(dc.copy()) for _ in range(5))[2]
            finally:
                lib.omp_set_num_threads(nmax)
        cfg1 = {"workload": f"cfg1-{Pc.idm}x{Pc.jdm}x{Pc.kdm}", "sea_columns": seac,
                "value": seac * Pc.kdm / tc, "threads": threads, "seconds_per_call": tc,
                "one_thread_value": (seac * Pc.kdm / tc1) if tc1 else None,
                "about": "median of 5 calls of the whole domain"}
    except Exception as e:  # the band figure above stands on its own
        cfg1 = {"error": str(e)}
    return {
        "value": sea * P.kdm / t,
        "unit": "column-layers/s",
        "cores": threads,
        "logical_cpus": logical,
        "kind": "reference",
        "sample": (f"HYCOM-src hybgen (flang -O2, OpenMP over j) on rows {j0 + 1}-{j0 + rows} "
                   f"of the {itdm}x{jtdm}x{P.kdm} field ({sea} sea columns), median of 3 "
                   f"calls of {t:.3f} s, {threads} OpenMP threads = physical cores of the "
                   f"{logical} CPUs this process may use (lscpu)"),
        "one_thread": one,
        "config1": cfg1,
    }


def mixed_leg(lib, dd, S, sea, args, stream):
    """The same workload in HYB_PREC_MIXED (float reconstruction and piece
    integrals, FP64 interfaces and sums): throughput on the full domain, and
    the error against the FP64 path on a 600x400 window of the same field."""
    import torch
    P = dd.P
    P.precision = abi.PREC_MIXED
    dd.synth(S, stream=stream)
    for _ in range(args.warmup):
        dd.hybgen(m=1, n=2, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dd.hybgen(m=1, n=2, stream=stream)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    P.precision = abi.PREC_FP64
    print(f"[bench] mixed {1e3 * t:.2f} ms/step", file=sys.stderr, flush=True)
    # error on a window: FP64 and mixed from the same input
    w0 = abi.WORKLOADS[args.workload]
    Pw, ww = abi.workload_params(args.workload, hybmap=args.hybmap, idm=600, jdm=400)
    Sw = abi.make_synth(Pw, ww, seed=args.seed, itdm=w0["idm"], jtdm=w0["jdm"],
                        i0=w0["idm"] // 2 - 300, j0=w0["jdm"] // 2 - 200)
    a = abi.DeviceDomain(Pw).synth(Sw)
    Pm, _ = abi.workload_params(args.workload, hybmap=args.hybmap, idm=600, jdm=400)
    Pm.precision = abi.PREC_MIXED
    b = abi.DeviceDomain(Pm).synth(Sw)
    torch.cuda.synchronize()
    h0 = a.download()
    # both windows in one batched call (its kernels are the batch ones, so
    # the full-domain kernels' traces hold the full-domain calls only)
    abi.Batch([a, b]).hybgen(m=1, n=2)
    torch.cuda.synchronize()
    ha, hb = a.download(), b.download()
    import numpy as np
    nb = Pw.nbdy
    sea_w = ha.a["ip"][nb:nb + Pw.ii, nb:nb + Pw.jj] > 0
    err, rms = {}, {}
    for f in ("temp", "saln", "th3d", "tracer", "u", "v"):
        x, y = ha.a[f], hb.a[f]
        scale = max(float(abs(x).max()), 1e-300)
        err[f] = float(abs(x - y).max() / scale)
        rms[f] = float(np.sqrt(np.mean((x - y) ** 2)) / scale)
    # column integrals sum_k dp*x at level n over sea columns: the change
    # across the call (the remap's conservation residual, FP64 and mixed)
    # and the mixed integral against the FP64 one
    cons = {}

    def colint(h, f, tr=None):
        x = h.a[f][..., 1] if tr is None else h.a[f][..., 1, tr]
        return (h.a["dp"][..., 1] * x).sum(axis=2)[nb:nb + Pw.ii, nb:nb + Pw.jj][sea_w]

    fields = [("temp", None), ("saln", None)] + [("tracer", q) for q in range(Pw.ntracr)]
    for f, tr in fields:
        i0, i64, imx = colint(h0, f, tr), colint(ha, f, tr), colint(hb, f, tr)
        den = np.maximum(abs(i0), 1e-300)
        key = f if tr is None else f"tracer{tr + 1}"
        cons[key] = {"fp64_vs_input": float((abs(i64 - i0) / den).max()),
                     "mixed_vs_input": float((abs(imx - i0) / den).max()),
                     "mixed_vs_fp64": float((abs(imx - i64) / den).max())}
    same = all((ha.a[f] == hb.a[f]).all() for f in ("p", "dp", "dpu", "dpv", "pu", "pv"))
    return {"value": sea * P.kdm / t, "unit": "column-layers/s", "ms_per_step": 1e3 * t,
            "dtype": "f32 reconstruction, f64 interfaces and sums",
            "speedup_vs_fp64": None, "max_rel_error_vs_fp64": err,
            "rms_rel_error_vs_fp64": rms, "column_integral_residual": cons,
            "interfaces_bitwise_fp64": bool(same),
            "error_window": "600x400 centre window of the same field"}


def bytes_per_column_stokes(K, NZ):
    """Algorithmic HBM bytes of the Stokes drift remap per sea column: read
    both snapshots of usdz and vsdz (4*nsdzi), dp at the time level (K) and
    the bottom pressure; write usdp and vsdp (2K); read ip (int32)."""
    return 8 * (4 * NZ + K + 1 + 2 * K) + 4


def stokes_leg(lib, dd, P, sea, args, stream):
    """hycom_stokes_vertical_device on the same field's dp, p and ip:
    nsdzi fixed Stokes layers onto the kdm model layers, seeded snapshots
    decaying with depth, HIP events on the call's stream."""
    import numpy as np
    import torch
    nz = args.stokes_nsdzi
    NI, NJ = P.idm + 2 * P.nbdy, P.jdm + 2 * P.nbdy
    nij = NI * NJ
    z = np.concatenate([[0.0], np.cumsum(np.geomspace(1.0, 60.0, nz - 1))])
    sdzi = np.ascontiguousarray(z * abi.ONEM)
    sdzi[0] = abi.DPTHIN
    g = torch.Generator(device="cuda").manual_seed(args.seed)
    zmid = torch.tensor(np.concatenate([0.5 * (z[:-1] + z[1:]), [z[-1]]])[:nz], device="cuda")
    scale = 3.0 + 12.0 * torch.rand(nij, 1, 1, device="cuda", generator=g)
    prof = torch.exp(-zmid.view(1, 1, nz) / scale)                     # (nij, 1, nz)
    def snap():
        amp = 0.1 * torch.randn(nij, 2, 1, device="cuda", generator=g, dtype=torch.float64)
        x = amp * prof + 0.002 * torch.randn(nij, 2, nz, device="cuda", generator=g,
                                              dtype=torch.float64)
        return x.permute(2, 1, 0).contiguous().view(-1)                # (nz, 2, nij) = F order
    usdz, vsdz = snap(), snap()
    del prof, scale
    usdp = torch.zeros(nij * P.kdm, dtype=torch.float64, device="cuda")
    vsdp = torch.zeros_like(usdp)
    SP = abi.HybStokesParams(idm=P.idm, jdm=P.jdm, ii=P.ii, jj=P.jj, nbdy=P.nbdy, kdm=P.kdm,
                             nsdzi=nz, reserved=0, ws0=0.3, ws1=1.0 - 0.3)
    F = abi.HybStokesFields()
    F.usdz, F.vsdz, F.sdzi = usdz.data_ptr(), vsdz.data_ptr(), sdzi.ctypes.data
    F.dp, F.p, F.ip = dd.t["dp"].data_ptr(), dd.t["p"].data_ptr(), dd.t["ip"].data_ptr()
    F.usdp, F.vsdp = usdp.data_ptr(), vsdp.data_ptr()
    call = lambda: abi.check(lib.hycom_stokes_vertical_device(ctypes.byref(SP), ctypes.byref(F),
                                                              2, stream))
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    bpc = bytes_per_column_stokes(P.kdm, nz)
    out = {"value": sea * P.kdm / (ms * 1e-3), "unit": "column-layers/s", "ms_per_call": ms,
           "nsdzi": nz, "kdm": P.kdm, "bytes_per_sea_column": bpc,
           "achieved_GBs": sea * bpc / (ms * 1e-3) / 1e9,
           "kernels": "k_stokes_walk (walk_core PLM, 2 fields, snapshots blended as read)"}
    print(f"[bench] stokes {ms:.2f} ms/call", file=sys.stderr, flush=True)
    del usdz, vsdz, usdp, vsdp
    torch.cuda.empty_cache()
    return out


def bytes_per_column_trc(KI, KO, KS):
    """Algorithmic HBM bytes of the tracer column remap per masked column:
    read si (ki*ks), pi (ki+1) and po (ko+1), write so (ko*ks), read the
    mask (int32)."""
    return 8 * (KI * KS + KI + 1 + KO + 1 + KO * KS) + 4


def trc_leg(lib, dd, P, sea, args, stream):
    """hycom_trc_remap_device with plmtrc (method 1), as thermf's tracer
    relaxation calls it (thermf.F90:770-806): the field's tracers as seeded
    climatology on interfaces half a layer off the model's, onto the
    model's interfaces p, at every sea point."""
    import torch
    NI, NJ = P.idm + 2 * P.nbdy, P.jdm + 2 * P.nbdy
    nij, K, T = NI * NJ, P.kdm, max(P.ntracr, 1)
    p = dd.t["p"].view(K + 1, nij)
    pi = torch.empty_like(p)
    pi[0] = 0.0
    pi[1:K] = 0.5 * (p[1:K] + p[2:K + 1])
    pi[K] = p[K]
    g = torch.Generator(device="cuda").manual_seed(args.seed + 7)
    si = torch.cumsum(torch.randn(T, K, nij, device="cuda", generator=g, dtype=torch.float64),
                      dim=1) + 5.0
    so = torch.zeros(T, K, nij, device="cuda", dtype=torch.float64)
    TP = abi.HybTrcParams(idm=P.idm, jdm=P.jdm, ii=P.ii, jj=P.jj, nbdy=P.nbdy, ki=K, ko=K, ks=T,
                          method=1, reserved=0)
    F = abi.HybTrcFields()
    F.si, F.pi, F.po = si.data_ptr(), pi.data_ptr(), dd.t["p"].data_ptr()
    F.mask, F.so = dd.t["ip"].data_ptr(), so.data_ptr()
    call = lambda: abi.check(lib.hycom_trc_remap_device(ctypes.byref(TP), ctypes.byref(F), stream))
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    bpc = bytes_per_column_trc(K, K, T)
    out = {"value": sea * K / (ms * 1e-3), "unit": "column-layers/s", "ms_per_call": ms,
           "method": "plmtrc", "ki": K, "ko": K, "ks": T, "bytes_per_sea_column": bpc,
           "achieved_GBs": sea * bpc / (ms * 1e-3) / 1e9,
           "kernels": "k_trc<1,4,3> (plmtrc, 4 fields per scan, slopes formed as the scan goes)"}
    print(f"[bench] trc {ms:.2f} ms/call", file=sys.stderr, flush=True)
    del si, so, pi
    torch.cuda.empty_cache()
    return out


def copy_leg(args):
    """Copy mode (hycom_hybgen, host arrays in and out, the path the stock
    model's call takes) on a window of the GLBb field: ms per call with the
    PCIe transfers, the bytes each way, and the same window's device-only
    call for comparison.  The first call allocates the library's device
    mirror and pinned staging buffers and is not timed."""
    import torch
    lib = abi.load()
    P, w = abi.workload_params(args.workload, idm=args.copy_idm, jdm=args.copy_jdm,
                               hybmap=args.hybmap)
    g, _ = abi.workload_params(args.workload)
    i0, j0 = (g.idm - P.idm) // 2, (g.jdm - P.jdm) // 2
    S = abi.make_synth(P, w, seed=args.seed, i0=i0, j0=j0, itdm=g.idm, jtdm=g.jdm)
    print(f"[bench] copy mode: host window {P.idm}x{P.jdm}x{P.kdm}", file=sys.stderr, flush=True)
    import numpy as np
    h = abi.HostDomain(P).synth(S)
    sea = int(np.count_nonzero(h.a["ip"][P.nbdy:P.nbdy + P.ii, P.nbdy:P.nbdy + P.jj]))
    dd = abi.DeviceDomain(P, skip=("dpo",)).upload(h)
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        dd.hybgen(m=1, n=2, stream=stream)
    torch.cuda.synchronize()
    td = []
    for _ in range(5):
        ts = time.perf_counter()
        dd.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        td.append(time.perf_counter() - ts)
    del dd
    torch.cuda.empty_cache()
    h.hybgen(m=1, n=2)  # allocates the device mirror and staging buffers
    tc = []
    for q in range(3):
        ts = time.perf_counter()
        h.hybgen(m=1 + q % 2, n=2 - q % 2)
        tc.append(time.perf_counter() - ts)
    up, down = ctypes.c_int64(), ctypes.c_int64()
    abi.check(lib.hycom_hybgen_copy_stats(ctypes.byref(up), ctypes.byref(down)))
    tc.sort()
    td.sort()
    ms = 1e3 * tc[1]
    return {"window": f"{P.idm}x{P.jdm}x{P.kdm} centre window of the GLBb field, "
                      f"{P.ntracr} tracers, host arrays (pageable)",
            "sea_columns": sea, "ms_per_call": ms, "device_only_ms_per_call": 1e3 * td[2],
            "column_layers_per_s": sea * P.kdm / (ms * 1e-3),
            "bytes_up": up.value, "bytes_down": down.value,
            "pcie_GBs": (up.value + down.value) / (ms * 1e-3) / 1e9,
            "about": "median of 3 calls after one that allocates the device mirror and the "
                     "pinned staging buffers; levels n of the fields the call touches go both "
                     "ways, p/pu/pv both ways, the read-only fields up, on every call"}


def batch_leg(args):
    """64 regional domains of random shape (idm 100-1200, jdm 100-900, kdm
    22/32/41, shelf fraction 0-60%), FP64 PPM, stepped together by
    hycom_hybgen_batch (one graph launch) after the autotuner has settled;
    beside it the same domains stepped one hycom_hybgen_device call each."""
    import numpy as np
    import torch
    lib = abi.load()
    rng = np.random.default_rng(args.seed + 3)
    doms, sea = [], []
    for d in range(args.batch_domains):
        idm, jdm = int(rng.integers(100, 1201)), int(rng.integers(100, 901))
        kdm = int(rng.choice([22, 32, 41]))
        shelf = float(rng.uniform(0.0, 0.6))
        P, w = abi.workload_params("cfg1", idm=idm, jdm=jdm, kdm=kdm, shelf=shelf)
        dd = abi.DeviceDomain(P, skip=("dpo",))
        dd.synth(abi.make_synth(P, w, seed=args.seed + 17 * d))
        doms.append(dd)
        sea.append(sea_columns(dd) * kdm)
    torch.cuda.synchronize()
    layers = sum(sea)
    print(f"[bench] batch of {len(doms)} domains, {layers} sea column-layers",
          file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream().cuda_stream
    t_single = None
    if args.batch_baseline:
        # one hycom_hybgen_device call per domain (off by default: these
        # calls would mix other shapes into the full-domain kernels' traces)
        for dd in doms:
            dd.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for dd in doms:
                dd.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        t_single = (time.perf_counter() - t0) / args.steps
    # the batch: autotune, then time
    abi.check(lib.hycom_hybgen_batch_autotune(1))
    batch = abi.Batch(doms)
    calls = 0
    while not batch.config()[2] and calls < 64:
        batch.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        calls += 1
    batch.hybgen(m=1, n=2, stream=stream)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        batch.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    t_batch = (time.perf_counter() - t0) / args.steps
    rc, gc, tuned = batch.config()
    ns, kmax = batch.staged_levels()
    cands = batch.tuning()
    print(f"[bench] batch {1e3 * t_batch:.2f} ms/call; widths remap {rc} regrid {gc}, "
          f"{ns} of {kmax - 1} interface levels staged", file=sys.stderr, flush=True)
    lat.sort()
    shapes = [f"{dd.P.idm}x{dd.P.jdm}x{dd.P.kdm}" for dd in doms]
    del batch, doms
    torch.cuda.empty_cache()
    return {"value": layers / t_batch, "unit": "column-layers/s", "ms_per_call": 1e3 * t_batch,
            "domains": len(shapes), "sea_column_layers": layers,
            "one_call_per_domain": ({"value": layers / t_single, "ms_per_step": 1e3 * t_single,
                                     "speedup_of_batch": t_single / t_batch}
                                    if t_single else None),
            "autotuned": {"remap_cols": rc, "regrid_cols": gc, "remap_staged_levels": ns,
                          "kmax": kmax, "tuned": tuned, "tuning_calls": calls,
                          "candidates_ms": [{"remap_cols": a, "staged_levels": b,
                                             "regrid_cols": c, "ms": t}
                                            for a, b, c, t in cands]},
            "latency_ms": {"p50": 1e3 * lat[len(lat) // 2], "p99": 1e3 * lat[-1]},
            "shapes": shapes}


def serve_leg(args):
    """Request loop: Poisson arrivals (mean args.rate per s) for
    args.serve_seconds, each request nsteps ~ U{1..10} hybgen steps of one
    of args.serve_domains resident windows of the GLBb field (idm 150-600,
    jdm 100-400, kdm 41, 4 tracers, PPM), FP64 or mixed with equal odds.
    Latency is completion minus submission on the server's clock."""
    import numpy as np
    rng = np.random.default_rng(args.seed + 5)
    w0 = abi.WORKLOADS["glbb"]
    srv = abi.Server(max_active=args.serve_slots)
    ids, shapes = [], []
    for d in range(args.serve_domains):
        idm, jdm = int(rng.integers(150, 601)), int(rng.integers(100, 401))
        i0 = int(rng.integers(0, w0["idm"] - idm))
        j0 = int(rng.integers(0, w0["jdm"] - jdm))
        P, w = abi.workload_params("glbb", idm=idm, jdm=jdm, hybmap=args.hybmap)
        S = abi.make_synth(P, w, seed=args.seed, itdm=w0["idm"], jtdm=w0["jdm"], i0=i0, j0=j0)
        ids.append(srv.add_domain(P, S))
        shapes.append(f"{idm}x{jdm}@{i0},{j0}")
    # every domain made resident by a one-step request before the timed run
    for d in ids:
        srv.submit(d, 1, abi.PREC_FP64, tag=-1)
    warm = []
    while len(warm) < len(ids):
        warm += srv.poll(wait_ms=100)
    assert all(c.status == 0 for c in warm), "warm-up request failed"
    st0 = srv.stats()
    # arrivals
    n_req = int(rng.poisson(args.rate * args.serve_seconds))
    gaps = rng.exponential(1.0 / args.rate, size=n_req)
    t_arr = np.cumsum(gaps)
    doms = rng.integers(0, len(ids), size=n_req)
    steps = rng.integers(1, 11, size=n_req)
    prec = rng.integers(0, 2, size=n_req)
    print(f"[bench] serving {n_req} requests over {t_arr[-1]:.1f} s "
          f"({len(ids)} resident domains)", file=sys.stderr, flush=True)
    done = []
    t0 = time.perf_counter()
    q = 0
    last_print = t0
    while q < n_req or len(done) < n_req:
        now = time.perf_counter() - t0
        while q < n_req and t_arr[q] <= now:
            srv.submit(ids[int(doms[q])], int(steps[q]), int(prec[q]), tag=q)
            q += 1
        wait = 1 if q >= n_req else max(0, int(1e3 * (t_arr[q] - now)))
        done += srv.poll(wait_ms=min(wait, 2))
        if time.perf_counter() - last_print > 30:
            last_print = time.perf_counter()
            print(f"[bench] serving: {q} submitted, {len(done)} done", file=sys.stderr, flush=True)
        if now > 10 * args.serve_seconds + 60:
            raise RuntimeError(f"serving run stalled: {len(done)} of {n_req} done")
    st = srv.stats()
    srv.close()
    bad = [c for c in done if c.status != 0]
    assert not bad, f"{len(bad)} requests failed"
    lat = np.sort(np.array([c.t_done - c.t_submit for c in done]))
    qwait = np.array([c.t_start - c.t_submit for c in done])
    span = max(c.t_done for c in done) - min(c.t_submit for c in done)
    layers = sum(c.column_layers for c in done)
    iters = st.iterations - st0.iterations
    return {"value": layers / span, "unit": "column-layers/s",
            "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                           "p99": 1e3 * float(np.percentile(lat, 99)),
                           "max": 1e3 * float(lat[-1])},
            "queue_wait_ms_p50": 1e3 * float(np.median(qwait)),
            "requests": n_req, "rate_per_s": args.rate, "span_s": span,
            "mixed_fraction": float(prec.mean()),
            "iterations": iters,
            "mean_active_per_iteration": (st.active_sum - st0.active_sum) / max(iters, 1),
            "max_active": st.max_active_seen, "slots": args.serve_slots,
            "resident_domains": st.resident_domains,
            "resident_GB": st.resident_bytes / 1e9,
            "workload": (f"{len(ids)} windows of the GLBb 4500x3298x41 field, idm 150-600, "
                         f"jdm 100-400, 4 tracers, PPM; nsteps U{{1..10}}; FP64 or mixed 50/50")}


def model_timing(args):
    """The stand-alone model, stock and with the drop-in, on one case: a
    300x240x41 closed basin at 0.1 degree (tests/model_case.py), 450 s
    steps for an eighth of a day (inicon's two hybgen calls and 24 steps).
    Four builds: stock and with the drop-in (whose hybgen runs on the GPU in
    copy mode), each serial (config/generic-gnu-relo_one) and with OpenMP
    over j on 8 threads (relo_omp; mod_dimensions.F90 takes up to 8 with
    RELO).  Reports the model's own timers (xctmr: hybgen is timer 49
    around the call, mod_hycom.F90:2648-2653) and whether the runs'
    archives are the same bytes."""
    import re
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import model_case
    mdl = os.path.join(ROOT, "oracle", "_ref", "model")
    case = dict(idm=300, jdm=240, kdm=41, hybmap=2, days=0.125, res=0.1, baclin=450.0,
                batrop=15.0)
    out = {"case": "300x240x41 closed basin, 0.1 degree, z-level start, PPM, "
                   "inicon's 2 hybgen calls + 24 steps of 450 s",
           "runs": {}}
    dirs = {}
    # the serial builds keep the environment's thread count: the drop-in's
    # library splits its staging copies over those threads
    for name, exe, nthr in (("stock", "hycom", None), ("dropin", "hycom_hip", None),
                            ("stock_omp8", "omp/hycom", 8), ("dropin_omp8", "omp/hycom_hip", 8)):
        path = os.path.join(mdl, exe)
        if not os.path.exists(path):
            raise SystemExit(f"bench.py --model-timing: {path} is not built "
                             "(make -C oracle/ref model)")
        d = tempfile.mkdtemp(prefix=f"model_{name}_")
        model_case.write_case(d, **case)
        print(f"[bench] model timing: {name} ({exe})", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        env = dict(os.environ, OMP_NUM_THREADS=str(nthr)) if nthr else None
        with open(os.path.join(d, "run.log"), "w") as log:
            rc = subprocess.run([path], cwd=d, stdout=log, stderr=subprocess.STDOUT, env=env,
                                timeout=900).returncode
        wall = time.perf_counter() - t0
        text = open(os.path.join(d, "run.log")).read()
        if rc != 0 or "(normal)" not in text:
            raise SystemExit(f"bench.py --model-timing: {name} run failed (rc {rc})\n"
                             + text[-2000:])
        block = text[text.rindex("timer statistics"):]
        timers = {m.group(1): {"calls": int(m.group(2)), "s": float(m.group(3))}
                  for m in re.finditer(r"^\s+(\S+)\s+calls =\s+(\d+)\s+time =\s+([\d.]+)",
                                       block, re.M)}
        out["runs"][name] = {"threads": nthr, "wall_s": wall, "timers": timers}
        dirs[name] = d
        print(f"[bench] model timing: {name} hybgen {timers.get('hybgen')}, "
              f"total {timers.get('total')}", file=sys.stderr, flush=True)
    arch = sorted(f for f in os.listdir(dirs["stock"]) if f.startswith("archv.") and f.endswith(".a"))
    out["archives_identical"] = bool(arch) and all(
        open(os.path.join(dirs["stock"], f), "rb").read()
        == open(os.path.join(dirs[n], f), "rb").read()
        for f in arch for n in dirs if n != "stock")
    tm = {n: out["runs"][n]["timers"] for n in dirs}
    out["hybgen_ms_per_call"] = {n: 1e3 * t["hybgen"]["s"] / t["hybgen"]["calls"]
                                 for n, t in tm.items()}
    out["hybgen_speedup"] = {"serial": tm["stock"]["hybgen"]["s"] / tm["dropin"]["hybgen"]["s"],
                             "omp8": tm["stock_omp8"]["hybgen"]["s"]
                             / tm["dropin_omp8"]["hybgen"]["s"]}
    out["model_speedup"] = {"serial": tm["stock"]["total"]["s"] / tm["dropin"]["total"]["s"],
                            "omp8": tm["stock_omp8"]["total"]["s"]
                            / tm["dropin_omp8"]["total"]["s"]}
    print(json.dumps({"model_timing": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="glbb", choices=sorted(abi.WORKLOADS))
    ap.add_argument("--hybmap", type=int, default=2)
    ap.add_argument("--seed", type=int, default=200)
    ap.add_argument("--cpu-rows", type=int, default=160)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision (HYB_PREC_MIXED) leg")
    ap.add_argument("--no-weno", action="store_true", help="skip the WENO-remapper leg")
    ap.add_argument("--no-stokes", action="store_true",
                    help="skip the Stokes drift remap leg (hycom_stokes_vertical_device)")
    ap.add_argument("--stokes-nsdzi", type=int, default=24)
    ap.add_argument("--no-trc", action="store_true",
                    help="skip the tracer column remap leg (hycom_trc_remap_device, plmtrc)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched ragged domains leg")
    ap.add_argument("--batch-domains", type=int, default=64)
    ap.add_argument("--batch-baseline", action="store_true",
                    help="also time the batch's domains one hycom_hybgen_device call each")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-loop leg")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the HBM-counter calibration kernels (1 GiB fill, sum, clone)")
    ap.add_argument("--no-copy", action="store_true", help="skip the copy-mode leg")
    ap.add_argument("--model-timing", action="store_true",
                    help="only time the stand-alone model, stock and with the drop-in "
                         "(oracle/_ref/model, tests/model_case.py)")
    ap.add_argument("--no-denfix", action="store_true", help="skip the HYBGEN_DENFIX leg")
    ap.add_argument("--no-raflag", action="store_true", help="skip the Robert-Asselin leg")
    ap.add_argument("--copy-idm", type=int, default=1500)
    ap.add_argument("--copy-jdm", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=200.0, help="serving: requests per second")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-domains", type=int, default=64)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--kernel-path", type=int, default=0,
                    help="0 auto, 1 lane arrays, 3 streaming")
    ap.add_argument("--remap-variant", type=int, default=3,
                    help="streaming remap variant (hycom_hybgen_tune knob 0; 3, the "
                         "library default, is the only one built)")
    ap.add_argument("--uv-prefetch", type=int, default=None,
                    help="streaming u/v prefetch depth 0, 2 or 4 (knob 1)")
    ap.add_argument("--regrid-ahead", type=int, default=None,
                    help="regrid window loaded one iteration ahead (1) or not (0; knob 3)")
    ap.add_argument("--concurrent-uv", type=int, default=None,
                    help="u/v phase beside the p-column remap on a second stream, 0 or 1 "
                         "(knob 4)")
    ap.add_argument("--uv-recompute", type=int, default=None,
                    help="u/v new thicknesses from LDS (0) or formed again from p (1; knob 2)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the hybgen path is single-GPU here; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)

    if args.model_timing:
        model_timing(args)
        return

    import torch
    if args.calibrate:
        # known traffic for the FETCH_SIZE / WRITE_SIZE factors: a write-only
        # fill, a read-only sum and a copy, 1 GiB each, 3 times
        n = 1 << 27
        x = torch.empty(n, dtype=torch.float64, device="cuda")
        for _ in range(3):
            x.fill_(1.0)
            torch.cuda.synchronize()
            s_ = x.sum()
            torch.cuda.synchronize()
            y = x.clone()
            torch.cuda.synchronize()
            del y
        print(json.dumps({"calibration": "1 GiB float64: fill (writes 1 GiB), sum (reads 1 GiB), "
                                         "clone (reads and writes 1 GiB), 3 times",
                          "sum": float(s_.item())}), flush=True)
        return
    lib = abi.load()
    abi.check(lib.hycom_hybgen_set_kernel_path(args.kernel_path))
    abi.check(lib.hycom_hybgen_tune(0, args.remap_variant))
    if args.uv_prefetch is not None:
        abi.check(lib.hycom_hybgen_tune(1, args.uv_prefetch))
    if args.uv_recompute is not None:
        abi.check(lib.hycom_hybgen_tune(2, args.uv_recompute))
    if args.regrid_ahead is not None:
        abi.check(lib.hycom_hybgen_tune(3, args.regrid_ahead))
    if args.concurrent_uv is not None:
        abi.check(lib.hycom_hybgen_tune(4, args.concurrent_uv))
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    P, w = abi.workload_params(args.workload, hybmap=args.hybmap)
    print(f"[bench] allocating {args.workload} {P.idm}x{P.jdm}x{P.kdm} ntracr={P.ntracr}",
          file=sys.stderr, flush=True)
    dd = abi.DeviceDomain(P, skip=("dpo",))
    S = abi.make_synth(P, w, seed=args.seed)
    stream = torch.cuda.current_stream().cuda_stream
    dd.synth(S, stream=stream)
    torch.cuda.synchronize()
    sea = sea_columns(dd)
    print(f"[bench] sea columns {sea}; warmup {args.warmup}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        dd.hybgen(m=1, n=2, stream=stream)
    torch.cuda.synchronize()

    np_ = len(PHASES)
    ms = (ctypes.c_double * np_)()
    cnt = (ctypes.c_int32 * np_)()
    abi.check(lib.hycom_hybgen_timing(1))
    step_s = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        dd.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    abi.check(lib.hycom_hybgen_timing_read(ms, cnt, np_))
    abi.check(lib.hycom_hybgen_timing(0))
    print(f"[bench] timed {args.steps} steps in {t1 - t0:.3f} s", file=sys.stderr, flush=True)

    total = t1 - t0
    ms_step = 1e3 * total / args.steps
    value = sea * P.kdm * args.steps / total

    # The timed calls above regrid the state earlier calls left, as
    # successive model steps do.  The first call on freshly generated input
    # (perturbed isopycnal interfaces, far from the target grid) moves more
    # interfaces and sends more columns down the branches of the regrid:
    # timed alone here, three times, the generator outside the HIP events.
    first = []
    for q in range(3):
        dd.synth(S, stream=stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dd.hybgen(m=1, n=2, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        first.append(e0.elapsed_time(e1))
    first.sort()
    print(f"[bench] first call on fresh input {first[1]:.2f} ms", file=sys.stderr, flush=True)
    if cnt[7]:
        # the streaming kernels: the p-column remap kernel on its own
        kname = f"{REMAP_KERNELS[args.remap_variant].format(m=args.hybmap)} (p-column remap, hybgenaj)"
        ms_a = ms[7] / cnt[7]
        bpc_a = bytes_per_column_remap(P.kdm, P.ntracr)
    else:
        kname = "k_phase_a (column regrid + remap, hybgenaj)"
        ms_a = ms[0] / max(cnt[0], 1)
        bpc_a = bytes_per_column_phase_a(P.kdm, P.ntracr)
    achieved = sea * bpc_a / (ms_a * 1e-3) / 1e9
    srt = sorted(step_s)
    p50 = srt[len(srt) // 2] * 1e3
    p99 = srt[min(len(srt) - 1, int(round(0.99 * (len(srt) - 1))))] * 1e3

    weno = None
    if not args.no_weno and args.hybmap != 3:
        # the same field and steps with the WENO remapper (hybmap 3)
        P.hybmap = 3
        for _ in range(args.warmup):
            dd.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        tw = []
        for _ in range(args.steps):
            ts = time.perf_counter()
            dd.hybgen(m=1, n=2, stream=stream)
            torch.cuda.synchronize()
            tw.append(time.perf_counter() - ts)
        P.hybmap = args.hybmap
        tw.sort()
        weno = {"value": sea * P.kdm * len(tw) / sum(tw), "unit": "column-layers/s",
                "ms_per_step": 1e3 * sum(tw) / len(tw),
                "ms_p50": 1e3 * tw[len(tw) // 2],
                "ms_p99": 1e3 * tw[min(len(tw) - 1, int(round(0.99 * (len(tw) - 1))))]}
        print(f"[bench] weno {weno['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    denfix = None
    if not args.no_denfix:
        # the same field and steps with HYBGEN_DENFIX (the regional build's
        # fixed-layer search, hybgen.F90:496-549)
        P.ldenfix = 1
        for _ in range(args.warmup):
            dd.hybgen(m=1, n=2, stream=stream)
        torch.cuda.synchronize()
        tw = []
        for _ in range(args.steps):
            ts = time.perf_counter()
            dd.hybgen(m=1, n=2, stream=stream)
            torch.cuda.synchronize()
            tw.append(time.perf_counter() - ts)
        P.ldenfix = 0
        tw.sort()
        denfix = {"value": sea * P.kdm * len(tw) / sum(tw), "unit": "column-layers/s",
                  "ms_per_step": 1e3 * sum(tw) / len(tw),
                  "ms_p50": 1e3 * tw[len(tw) // 2]}
        print(f"[bench] denfix {denfix['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    raflag = None
    if not args.no_raflag:
        # hybgen(m,n,.true.): the Robert-Asselin branch (hybgencj) and the
        # second u/v pass on level m, on freshly generated input (both time
        # levels and dpo from the generator), generator outside the events
        import torch as _t
        dd.t["dpo"] = _t.zeros_like(dd.t["dp"])
        tr = []
        rms = (ctypes.c_double * np_)()
        rcnt = (ctypes.c_int32 * np_)()
        for q in range(4):
            if q == 1:
                abi.check(lib.hycom_hybgen_timing(1))
            dd.synth(S, stream=stream)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            dd.hybgen(m=1, n=2, raflag=1, stream=stream)
            e1.record()
            torch.cuda.synchronize()
            if q:
                tr.append(e0.elapsed_time(e1))
        abi.check(lib.hycom_hybgen_timing_read(rms, rcnt, np_))
        abi.check(lib.hycom_hybgen_timing(0))
        del dd.t["dpo"]
        torch.cuda.empty_cache()
        tr.sort()
        raflag = {"ms_per_call": tr[1], "all_ms": tr,
                  "phase_ms": {PHASES[q]: rms[q] / rcnt[q] for q in range(np_) if rcnt[q]},
                  "value": sea * P.kdm / (tr[1] * 1e-3), "unit": "column-layers/s",
                  "about": "hybgen(1,2,.true.) on freshly generated input, median of 3 "
                           "after one warm-up call"}
        print(f"[bench] raflag {tr[1]:.2f} ms/call", file=sys.stderr, flush=True)
    mixed = None
    if not args.no_mixed:
        mixed = mixed_leg(lib, dd, S, sea, args, stream)
    stokes = trc = None
    if not args.no_stokes:
        stokes = stokes_leg(lib, dd, P, sea, args, stream)
    if not args.no_trc:
        trc = trc_leg(lib, dd, P, sea, args, stream)
    batch = serve = copy = None
    if not (args.no_batch and args.no_serve and args.no_copy):
        del dd  # the GLBb fields make room for the other legs
        torch.cuda.empty_cache()
    if not args.no_copy:
        copy = copy_leg(args)
    if not args.no_batch:
        batch = batch_leg(args)
    if not args.no_serve:
        serve = serve_leg(args)

    traffic = valu = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_hybmap{args.hybmap}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        if cnt[7]:
            traffic = tj.get("bytes_per_launch_by_kernel", {}).get(kname.split(" ")[0])
        else:
            traffic = tj.get("bytes_per_launch")
        valu = tj.get("valu_instructions_per_launch", {}).get(kname.split(" ")[0])
        # HBM bytes per launch of the adjacent remaps' kernels (same profile)
        for leg, kn in ((stokes, "k_stokes_walk"), (trc, "k_trc<1,4,3>")):
            if leg is not None and kn in tj.get("bytes_per_launch_by_kernel", {}):
                leg["traffic_bytes_per_launch"] = tj["bytes_per_launch_by_kernel"][kn]

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.workload, args.hybmap, args.cpu_rows, args.seed)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "column-layers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, GLBb-shaped bathymetry and stratification)",
        "config": {
            "workload": f"{args.workload}-{P.idm}x{P.jdm}x{P.kdm}-"
                        f"{['pcm', 'plm', 'ppm', 'weno'][args.hybmap]}",
            "idm": P.idm, "jdm": P.jdm, "kdm": P.kdm, "ntracr": P.ntracr,
            "hybmap": args.hybmap, "precision": "fp64", "sea_columns": sea,
            "hybgen_raflag": False,
        },
        "latency_ms": {"p50": p50, "p99": p99},
        "first_call_ms": {"median": first[1], "all": first,
                          "about": "one call on freshly generated input (the configured "
                                   "perturbed-isopycnal state), median of 3; the timed steps "
                                   "regrid the state earlier calls left"},
        "phase_ms": {PHASES[q]: ms[q] / cnt[q] for q in range(np_) if cnt[q]},
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "kernel_ms": ms_a,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": sea * bpc_a,
            "bytes_per_sea_column": bpc_a,
            # VALU issue time of the kernel's SQ_INSTS_VALU (profiled run of the
            # same build) over its measured time: 4 cycles per wave64
            # instruction on 1024 SIMDs at 2.4 GHz
            "valu_issue_fraction": ((valu * 4 / (1024 * 2.4e9) * 1e3) / ms_a
                                    if valu and ms_a else None),
        },
        "call_bytes_per_sea_column": bytes_per_column_call(P.kdm, P.ntracr),
        "call_effective_GBs": sea * bytes_per_column_call(P.kdm, P.ntracr) / (ms_step * 1e-3) / 1e9,
        "cpu_baseline": cpu,
        "weno": weno,
        "denfix": denfix,
        "raflag": raflag,
        "mixed": mixed,
        "stokes": stokes,
        "trc": trc,
        "batch": batch,
        "serve": serve,
        "copy_mode": copy,
    }
    if mixed:
        mixed["speedup_vs_fp64"] = mixed["value"] / value
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""The stand-alone HYCOM model, stock against the drop-in.

oracle/ref/Makefile `model` builds the model from the HYCOM-src sources
twice: `hycom` as it lies, and `hycom_hip`, the same objects with the
Fortran shim's `hybgen` (hycom-src_amd/fortran/hybgen_hip.F90 over
libhycom_hybgen.so) in place of hybgen.F90's.  The two run side by side on
one case (model_case.py) with the model's own comparator: PIPE_MASTER in
one run directory and PIPE_SLAVE in the other name the same named pipe, and
after every phase of every step (mod_hycom.F90:2514-2651, the
`pipe_comparall` calls) the master compares every field of the slave with
its own, bit for bit, and stops at the first difference (mod_pipe.F90:
402-487).  The runs start from a z-level climatology, so hybgen regrids
every column hard in the first steps, and through inicon's two calls
(inicon.F90:544,559) before them.

CPU tests: the case runs and two stock runs agree; a PLM run against a PPM
run differs first after hybgen (the comparison sees hybgen's result).
The shim over the oracle (hycom_oracle) against the stock model, on the CPU.
GPU tests: stock against the drop-in, for each remap, the RA branch and
tracers.  Every case prints its name before it starts."""
import os
import re
import subprocess

import pytest

import model_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "oracle", "_ref", "model")
STOCK = os.path.join(MODEL, "hycom")
DROPIN = os.path.join(MODEL, "hycom_hip")
DENFIX = os.path.join(MODEL, "denfix")  # the three builds with HYBGEN_DENFIX
STEPS_PER_DAY = 96  # baclin 900 s, model_case's default

needs_stock = pytest.mark.skipif(not os.path.exists(STOCK),
                                 reason="oracle/_ref/model/hycom not built (make -C oracle/ref model)")


def run_pair(tmp, master, slave, master_case, slave_case=None, timeout=90):
    """Runs `master` and `slave` on their cases with a named pipe between
    them; returns the master's comparison log and both run logs."""
    m, s = os.path.join(tmp, "master"), os.path.join(tmp, "slave")
    model_case.write_case(m, **master_case)
    model_case.write_case(s, **(slave_case or master_case))
    fifo = os.path.join(tmp, "pipe")
    os.mkfifo(fifo)
    for d, name in ((m, "PIPE_MASTER"), (s, "PIPE_SLAVE")):
        with open(os.path.join(d, name), "w") as f:
            f.write(fifo + "\n")
    logs = {}
    procs = []
    for d, exe in ((m, master), (s, slave)):
        logs[d] = open(os.path.join(d, "run.log"), "w")
        procs.append(subprocess.Popen([exe], cwd=d, stdout=logs[d], stderr=subprocess.STDOUT))
    rcs = []
    try:
        for p in procs:
            rcs.append(p.wait(timeout=timeout))
    finally:
        for p in procs:
            if p.poll() is None:
                p.kill()
                p.wait()
        for f in logs.values():
            f.close()
    read = lambda p: open(p).read() if os.path.exists(p) else ""  # noqa: E731
    return (rcs, read(os.path.join(m, "PIPE_base.out")),
            read(os.path.join(m, "run.log")), read(os.path.join(s, "run.log")))


def ok_steps(base, phase):
    return [int(x) for x in re.findall(rf"^{phase}, step\s+(\d+) -- OK", base, re.M)]


def first_difference(base):
    """The phase and field of the master's first reported difference, and
    the lines reporting it."""
    bad = re.search(r"^i,j=.*(error|slave):|ERROR|out of sync", base, re.M)
    if not bad:
        return None
    head = base[:bad.start()]
    phase = re.findall(r"^(\w+, step\s+\d+)$", head, re.M)
    lines = base[bad.start():].splitlines()[:40]
    return f"first difference after [{phase[-1] if phase else '?'}]:\n" + "\n".join(lines)


def assert_identical(rcs, base, mlog, slog, days, baclin=900.0):
    diff = first_difference(base)
    assert diff is None, diff
    assert rcs == [0, 0], f"exit codes {rcs}\n{mlog[-1500:]}\n{slog[-1500:]}"
    assert "(normal)" in mlog and "(normal)" in slog
    # inicon's two calls (inicon.F90:544-560), then one per time step
    assert ok_steps(base, "inicn1") == [0] and ok_steps(base, "inicn2") == [0]
    steps = ok_steps(base, "hybgen")
    n = int(round(days * 86400.0 / baclin))
    assert len(steps) == n and steps[-1] - steps[0] == n - 1, steps[:5]
    # the comparisons read every field: u, v, dp, T, S, th3d of each layer
    assert len(re.findall(r"^reading for comparison: ", base, re.M)) > 100 * n


@needs_stock
def test_stock_against_stock(tmp_path):
    print("\n[case] stock model master and slave, PPM, 1 day", flush=True)
    rcs, base, mlog, slog = run_pair(str(tmp_path), STOCK, STOCK, dict(hybmap=2, days=1.0))
    assert_identical(rcs, base, mlog, slog, 1.0)
    # the start regrids: the z-level climatology is far from the targets
    assert "rdrlax" in mlog and "hybgen" in mlog


@needs_stock
def test_comparison_sees_a_hybgen_difference(tmp_path):
    """PPM master, PLM slave: every phase before the first hybgen call
    agrees, and the master stops at the comparison right after it (inicon's
    first call, inicon.F90:544-545)."""
    print("\n[case] stock PPM master against stock PLM slave", flush=True)
    rcs, base, mlog, slog = run_pair(str(tmp_path), STOCK, STOCK, dict(hybmap=2, days=0.25),
                                     dict(hybmap=1, days=0.25), timeout=60)
    assert "(pipe_compare_all)" in mlog, "the master did not stop at a difference"
    bad = re.search(r"^i,j=.*(error|slave):", base, re.M)  # mod_pipe.F90:458-466
    assert bad
    first = bad.start()
    phases = re.findall(r"^(\w+), step\s+(\d+)$", base[:first], re.M)
    assert phases[-1] == ("inicn1", "0"), phases
    assert all(re.search(rf"^{p}, step\s+{s} -- OK", base, re.M) for p, s in phases[:-1])


ORACLE_DROPIN = os.path.join(MODEL, "hycom_oracle")

CASES = [
    ("pcm", dict(hybmap=0)),
    ("plm", dict(hybmap=1)),
    ("ppm", dict(hybmap=2)),
    ("weno", dict(hybmap=3)),
    ("ppm_ra", dict(hybmap=2, hybraf=True)),
    ("ppm_tracers", dict(hybmap=2, ntracr=2, trcflg=23)),
    ("weno_k41", dict(hybmap=3, kdm=41, idm=64, jdm=48)),
    ("ppm_periodic", dict(hybmap=2, periodic=True)),
    ("ppm_sigma_thk", dict(hybmap=2, nsigma=4, hybthk=0.5, shelf=8.0)),
    ("plm_3days", dict(hybmap=1, days=3.0)),
    ("ppm_denfix", dict(hybmap=2, denfix=True)),
    ("ppm_ra_diag", dict(hybmap=2, hybraf=True, diag=(801, 802, 803))),
    ("plm_tracer_diag", dict(hybmap=1, ntracr=1, trcflg=3, diag=(803, 801))),
    ("ppm_dx0k_noiso_norelax", dict(hybmap=2, dx00=(30.0, 300.0, 1.1), hybiso=0.0,
                                    hybrlx=1.0)),
    # combinations
    ("weno_periodic_ra_tracers_denfix", dict(hybmap=3, periodic=True, hybraf=True, ntracr=2,
                                             trcflg=23, denfix=True)),
    ("plm_sigma_dx0k_ra_diag", dict(hybmap=1, nsigma=4, hybthk=0.9, shelf=8.0,
                                    dx00=(30.0, 300.0, 1.1), hybraf=True, ntracr=1, trcflg=2,
                                    diag=(802,))),
]


def builds(kw, dropin_name):
    """The stock build and the named drop-in build for a case: the
    HYBGEN_DENFIX builds where the case asks for them."""
    d = DENFIX if kw.get("denfix") else MODEL
    return os.path.join(d, "hycom"), os.path.join(d, dropin_name)


def case_kw(kw):
    return {k: v for k, v in kw.items() if k != "denfix"}


# on the GPU only (the oracle would take minutes): a domain several waves
# wide, 41 layers, two tracers, 450 s steps at 0.1 degree
GPU_CASES = CASES + [
    ("weno_160x120x41_tracers", dict(hybmap=3, kdm=41, idm=160, jdm=120, res=0.1, baclin=450.0,
                                     batrop=15.0, ntracr=2, trcflg=23, days=0.25)),
]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="oracle/_ref/model/hycom_hip not built (make -C oracle/ref model)")
@pytest.mark.parametrize("name,kw", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_stock_against_dropin(tmp_path, name, kw):
    days, baclin = kw.get("days", 1.0), kw.get("baclin", 900.0)
    print(f"\n[case] stock master (CPU) against the drop-in slave (GPU hybgen), {name}, "
          f"{int(round(days * 86400.0 / baclin))} steps", flush=True)
    stock, dropin = builds(kw, "hycom_hip")
    rcs, base, mlog, slog = run_pair(str(tmp_path), stock, dropin, dict(case_kw(kw), days=days),
                                     timeout=250)
    assert_identical(rcs, base, mlog, slog, days, baclin)
    check_tracers_and_archives(str(tmp_path), kw)


def check_tracers_and_archives(tmp, kw):
    """The pipe compares u, v, dp, dpu, dpv, T, S and th3d, not the
    tracers (mod_pipe.F90:1052-1700): the archives the two runs write
    (every field, tracers included, as float32 .a records, mod_archiv.F90)
    must be the same bytes."""
    m, s = os.path.join(tmp, "master"), os.path.join(tmp, "slave")
    arch = sorted(f for f in os.listdir(m) if f.startswith("archv.") and f.endswith(".a"))
    assert len(arch) >= 2 and arch == sorted(
        f for f in os.listdir(s) if f.startswith("archv.") and f.endswith(".a"))
    for f in arch:
        with open(os.path.join(m, f), "rb") as a, open(os.path.join(s, f), "rb") as b:
            assert a.read() == b.read(), f"archive {f} differs"


@pytest.mark.skipif(not os.path.exists(ORACLE_DROPIN),
                    reason="oracle/_ref/model/hycom_oracle not built (make -C oracle/ref model)")
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_stock_against_shim_over_oracle(tmp_path, name, kw):
    """The shim over the oracle (oracle/ref/oracle_as_dropin.c) in place of
    the GPU library, on the CPU: the oracle, and the shim's packing of the
    model's state, inside the model's own time loop."""
    days = kw.get("days", 1.0)
    print(f"\n[case] stock master against the shim over the oracle, {name}", flush=True)
    stock, dropin = builds(kw, "hycom_oracle")
    rcs, base, mlog, slog = run_pair(str(tmp_path), stock, dropin, dict(case_kw(kw), days=days))
    assert_identical(rcs, base, mlog, slog, days)
    check_tracers_and_archives(str(tmp_path), kw)


def archive_records(path, idm, jdm):
    """An archive's records (float32, voids as NaN) with the .b line that
    names each (mod_archiv.F90 writes one .b line per .a record after the
    header lines)."""
    import numpy as np
    n = idm * jdm
    nw = ((n + 4095) // 4096) * 4096
    a = np.fromfile(path, dtype=">f4").reshape(-1, nw)[:, :n].astype(np.float64)
    a[a > 1e29] = np.nan
    lines = open(path[:-2] + ".b").read().splitlines()
    first = next(q for q, ln in enumerate(lines) if ln.startswith("field")) + 1
    names = [ln.split()[0] + " k=" + ln.split("=")[1].split()[2] for ln in lines[first:] if ln.strip()]
    assert len(names) == a.shape[0], (len(names), a.shape)
    return a, names


def record_spread(d0, d1, idm, jdm):
    """Per archived record of the last archive: the largest difference
    between two runs over the record's largest magnitude, sorted."""
    import numpy as np
    last = sorted(f for f in os.listdir(d0) if f.startswith("archv.") and f.endswith(".a"))[-1]
    a, names = archive_records(os.path.join(d0, last), idm, jdm)
    b, _ = archive_records(os.path.join(d1, last), idm, jdm)
    assert a.shape == b.shape and np.array_equal(np.isnan(a), np.isnan(b))
    out = []
    for r in range(a.shape[0]):
        ok = ~np.isnan(a[r])
        if ok.any():
            scale = max(float(np.max(np.abs(a[r][ok]))), 1e-30)
            out.append((float(np.max(np.abs(a[r][ok] - b[r][ok]))) / scale, names[r]))
    return sorted(out, reverse=True)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="oracle/_ref/model/hycom_hip not built (make -C oracle/ref model)")
def test_mixed_dropin_stays_within_the_models_own_sensitivity(tmp_path):
    """HYCOM_HYBGEN_PRECISION=mixed runs the mixed form in the model.  The
    regrid's decisions are discontinuous in T and S, so after a day any
    tiny difference has grown wherever a decision flipped.  The yardstick
    is the stock model started from a T climatology one float32 ulp away
    (model_case tnudge): the mixed run may differ from the stock run by no
    more than that run does, per archived record, within a factor 2."""
    import numpy as np
    print("\n[case] stock, stock from T one ulp away, drop-in in mixed precision; PPM, 1 day",
          flush=True)
    kw = dict(hybmap=2, ntracr=1, trcflg=3, days=1.0)
    runs = {}
    for name, exe, env, extra in (("stock", STOCK, {}, {}),
                                  ("nudged", STOCK, {}, {"tnudge": 1}),
                                  ("mixed", DROPIN, {"HYCOM_HYBGEN_PRECISION": "mixed"}, {})):
        d = os.path.join(str(tmp_path), name)
        model_case.write_case(d, **kw, **extra)
        with open(os.path.join(d, "run.log"), "w") as log:
            rc = subprocess.run([exe], cwd=d, stdout=log, stderr=subprocess.STDOUT,
                                env=dict(os.environ, **env), timeout=90).returncode
        text = open(os.path.join(d, "run.log")).read()
        assert rc == 0 and "(normal)" in text, text[-1500:]
        runs[name] = d
    mixed = record_spread(runs["stock"], runs["mixed"], 40, 30)
    nudged = dict((n, w) for w, n in record_spread(runs["stock"], runs["nudged"], 40, 30))
    print(f"[mixed] largest differences over each record's magnitude: {mixed[:5]}", flush=True)
    print(f"[nudge] the same for the stock run from T one ulp away: "
          f"{sorted(((w, n) for n, w in nudged.items()), reverse=True)[:5]}", flush=True)
    assert mixed[0][0] > 0.0  # the mixed form did run
    top = max(nudged.values())
    assert mixed[0][0] <= 2.0 * top, (mixed[:5], top)
    p90m = float(np.percentile([w for w, _ in mixed], 90))
    p90n = float(np.percentile(list(nudged.values()), 90))
    assert p90m <= 2.0 * p90n + 1e-9, (p90m, p90n)

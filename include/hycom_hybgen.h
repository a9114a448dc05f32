/*
 * hycom_hybgen.h -- C ABI of the MI355X (gfx950) hybrid-coordinate generator.
 *
 * This library replaces HYCOM's
 *     subroutine hybgen(m,n,hybgen_raflag)            (hybgen.F90:14-306)
 * and everything below it on the call path (hybgenaj hybgen.F90:308-1756,
 * hybgenbj :1758-1907, hybgencj :1909-2274, the PCM/PLM/PPM/WENO remappers
 * :2276-3107, dpudpv dpudpv.F90:14-107 and the p halo of xctilr,
 * mod_xc_sm.h:1337-1429).  The reference has no array arguments: every field
 * is a module array of mod_cb_arrays (mod_cb_arrays.F90:13-33, 68-81,
 * 279-290).  Here the same arrays are passed as plain pointers, each one in
 * the unchanged Fortran layout
 *
 *     x(1-nbdy:idm+nbdy, 1-nbdy:jdm+nbdy, 1:kdm, 1:2 [, 1:ntracr])
 *
 * column-major, i fastest; element (i,j,k,l) sits at
 *     (i-1+nbdy) + NI*((j-1+nbdy) + NJ*((k-1) + kdm*(l-1)))
 * with NI = idm+2*nbdy, NJ = jdm+2*nbdy.  p, pu, pv have kdm+1 levels and no
 * time level; theta has kdm levels and no time level; ubavg/vbavg have 3
 * time levels (mod_cb_arrays.F90:1159-1160).
 *
 * Error convention: 0 on success, a negative HYB_E* code otherwise (the
 * reference calls xcstop instead, blkdat.F90:946-1012, dpudpv.F90:47-55).
 * hycom_hybgen_strerror() names the code; hycom_hybgen_last_error() gives a
 * sentence about the last failure on the calling thread.
 */
#ifndef HYCOM_HYBGEN_H
#define HYCOM_HYBGEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYB_ABI_VERSION 1

/* Largest supported layer count and tracer count.  kdm <= 64 lets one
 * 64-lane wavefront hold a whole column level-by-level; mxtrcr is the
 * reference's own tracer limit (mod_dimensions.F90:26). */
#define HYB_KMAX   64
#define HYB_TRMAX  99

/* return codes */
#define HYB_OK                0
#define HYB_EINVAL           -1  /* parameter outside what blkdat.F90 accepts   */
#define HYB_EUNSUPPORTED     -2  /* valid for the reference, not built here     */
#define HYB_EDEVICE          -3  /* HIP runtime error                           */
#define HYB_ENOMEM           -4  /* device allocation failed                    */
#define HYB_ENODEVICE        -5  /* no usable gfx950 device                     */
#define HYB_ENULL            -6  /* a required field pointer is NULL            */
#define HYB_ENONFINITE       -7  /* NaN/Inf found by the optional input scan    */
#define HYB_EBUSY            -8  /* serving: no slot or memory free right now   */

/* arithmetic modes */
#define HYB_PREC_FP64   0  /* FP64 throughout, bit-faithful to the reference   */
#define HYB_PREC_MIXED  1  /* FP32 reconstruction (edge values, slopes,
                              limiters, grid metrics, hardware reciprocals);
                              FP64 interfaces, FP64 piece overlap integrals
                              and TwoSum-compensated per-output sums; the
                              regrid stays bitwise FP64                         */

/*
 * Run-time parameters of one hybgen call.  Every member is the value of the
 * reference variable of the same name after blkdat/geopar set-up, in the
 * reference's units (pressure in Pa, i.e. metres times onem), so that a
 * Fortran shim copies them one-for-one:
 *   mod_dimensions.F90:45,49,68 (idm, jdm, kdm, ii, jj), :33 (nbdy)
 *   mod_cb_arrays.F90:279-290 (dx0k, dp0k, ds0k, dpns, dsns),
 *   :318-366 (isopcm, mxlmy), :807-822 (thbase, qhybrlx, hybiso, hybthn,
 *   hybthk, thkbot, dp00i, ra2fac, baclin), :824-833 (nhybrd, nsigma,
 *   hybmap, hybflg, ntracr, trcflg, itracr)
 *   mod_xc.F90 (nreg, vland)
 *   hybgen.F90:319-325 (the HYBGEN_DENFIX cpp macro becomes ldenfix)
 */
typedef struct hyb_params {
  int32_t idm, jdm;      /* array extents without halo                       */
  int32_t ii, jj;        /* active extents, 1..ii x 1..jj (ii<=idm, jj<=jdm) */
  int32_t nbdy;          /* halo width (6 in HYCOM)                          */
  int32_t kdm;           /* number of layers, kk == kdm                      */
  int32_t ntracr;        /* tracers in the trailing tracer index             */
  int32_t nhybrd;        /* number of hybrid layers                          */
  int32_t nsigma;        /* number of sigma levels (informational here)      */
  int32_t hybmap;        /* 0 PCM, 1 PLM, 2 PPM, 3 WENO-like                 */
  int32_t hybflg;        /* 0 T&S (the only value allowed with EOS_17T)      */
  int32_t nreg;          /* region type: 0,4 closed in i; else periodic in i;
                            <=2 closed in j, else periodic in j              */
  int32_t isopcm;        /* logical: hybiso > 0                              */
  int32_t mxlmy;         /* logical: MY2.5 q2/q2l fields (must be 0 here)     */
  int32_t ldenfix;       /* logical: the HYBGEN_DENFIX variant of fixlay     */
  int32_t itracr801;     /* tracer index or 0: dp change diagnostic          */
  int32_t itracr802;     /* tracer index or 0: temp change diagnostic        */
  int32_t itracr803;     /* tracer index or 0: saln change diagnostic        */
                         /* (1..99; above ntracr a diagnostic tracer is only
                            written, not remapped, as the model keeps its
                            mtracr diagnostics after the ntracr tracers,
                            blkdat.F90:1307-1330; the tracer array then
                            holds max(ntracr, itracr801..803) tracers)       */
  int32_t precision;     /* HYB_PREC_FP64 or HYB_PREC_MIXED                  */
  int32_t check_finite;  /* logical: scan the inputs the call reads at sea p-,
                            u- and v-points for NaN/Inf first; the first bad
                            point in row order is reported with
                            HYB_ENONFINITE and nothing is computed (the scan
                            waits for the stream).  Not used by the request
                            loop. */
  int32_t trcflg[HYB_TRMAX];
  double  thbase, qhybrlx, hybiso, hybthn, hybthk, thkbot, dp00i;
  double  dpns, dsns;    /* metres (geopar.F90:629-630)                      */
  double  baclin;        /* baroclinic time step, s (itracr 801-803 only)    */
  double  ra2fac;        /* Robert-Asselin weight (hybgen_raflag only)       */
  double  vland;         /* halo fill value for closed boundaries            */
  double  dp0k[HYB_KMAX], ds0k[HYB_KMAX], dx0k[HYB_KMAX];  /* Pa             */
} hyb_params;

/*
 * Field pointers, each in the layout described at the top of this file.
 * In hycom_hybgen() they are host pointers; in hycom_hybgen_device() and
 * the batched/serving calls they are device pointers.  Fields the call does
 * not touch may be NULL (dpo unless raflag, tracer when ntracr and the
 * itracr are all 0).
 */
typedef struct hyb_fields {
  double *dp, *dpo, *dpu, *dpv;        /* (NI,NJ,kdm,2)                     */
  double *temp, *saln, *th3d;          /* (NI,NJ,kdm,2)                     */
  double *u, *v;                       /* (NI,NJ,kdm,2)                     */
  double *tracer;                      /* (NI,NJ,kdm,2,max(ntracr,itracr))  */
  double *p, *pu, *pv;                 /* (NI,NJ,kdm+1)                     */
  const double *theta;                 /* (NI,NJ,kdm)                       */
  const double *depths, *topiso;       /* (NI,NJ); depths in m, topiso Pa   */
  const double *depthu, *depthv;       /* (NI,NJ), Pa                       */
  const double *ubavg, *vbavg;         /* (NI,NJ,3)                         */
  const int32_t *ip, *iu, *iv;         /* (NI,NJ) land/sea masks            */
} hyb_fields;

/* ---------------------------------------------------------------------- */
/* Library information                                                    */
/* ---------------------------------------------------------------------- */
int         hycom_hybgen_abi_version(void);
const char* hycom_hybgen_strerror(int code);
const char* hycom_hybgen_last_error(void);

/* ---------------------------------------------------------------------- */
/* Parameter set-up (host only, no device needed)                         */
/* ---------------------------------------------------------------------- */

/* Checks the parameter pack the way blkdat.F90:946-1012 checks the
 * blkdat.input values (hybrlx>=1 i.e. 0<qhybrlx<=1, hybthn>=1,
 * 0<=hybthk<=0.9, 0<=hybmap<=3, hybflg==0 for the 17-term EOS) plus the
 * limits of this build (kdm, ntracr, mxlmy). */
int hycom_hybgen_check_params(const hyb_params* P);

/* Derives dp0k, ds0k, dx0k, dpns, dsns exactly as geopar.F90:449-630 does
 * from the blkdat.input values (dp00* / ds00* / dx00* in metres, or
 * dx00 <= 0 for "dx0k given", in which case dx0k[] is taken as metres);
 * also sets qhybrlx = 1/hybrlx and isopcm = hybiso > 0
 * (blkdat.F90:954, 987).  isopyc (MICOM mode) is not supported. */
int hycom_hybgen_derive_vgrid(hyb_params* P,
                              double dp00, double dp00f, double dp00x,
                              double ds00, double ds00f, double ds00x,
                              double dp00i_m, double dx00, double dx00f,
                              double dx00x, double hybrlx, double isotop_m,
                              double* topiso_pa);

/* ---------------------------------------------------------------------- */
/* The hybgen call                                                        */
/* ---------------------------------------------------------------------- */

/* Drop-in for  call hybgen(m,n,hybgen_raflag)  (mod_hycom.F90:2650,
 * inicon.F90:544,559): host arrays in, host arrays out.  Copies the fields
 * the call reads to device mirrors the library keeps between calls, runs
 * the device path, copies the fields it writes back, and returns when the
 * host arrays hold the result.  m, n are the 1-based time levels. */
int hycom_hybgen(const hyb_params* P, const hyb_fields* F,
                 int m, int n, int raflag);

/* Copy mode moves only what the call reads and writes: time level n of
 * dp, dpu, dpv, temp, saln, th3d, u, v and the tracers both ways (and
 * level m plus dpo with hybgen_raflag), level n of ubavg/vbavg up, p, pu,
 * pv both ways, and theta, depths, topiso, depthu, depthv and the masks
 * up.  Every call moves all of it: the model changes some of these arrays
 * in place between calls (depthu/depthv after inicon's calls, the halo of
 * p).  The copies go through two pinned staging buffers of the library's
 * (2 x 32 MiB, allocated on the first copy-mode call and kept until the
 * process ends).  hycom_hybgen_copy_invalidate() has nothing to drop and returns 0
 * (kept for callers of the earlier cached form).  hycom_hybgen_copy_stats()
 * gives the bytes the last copy-mode call moved each way.  No reference
 * counterpart (the reference has no device copies). */
int hycom_hybgen_copy_invalidate(void);
int hycom_hybgen_copy_stats(int64_t* bytes_up, int64_t* bytes_down);

/* Same call on device-resident fields (device pointers in F), ordered on
 * the given HIP stream (hipStream_t passed as void*, NULL = default
 * stream).  Returns after enqueueing; the caller synchronises. */
int hycom_hybgen_device(const hyb_params* P, const hyb_fields* F,
                        int m, int n, int raflag, void* stream);

/* hycom_hybgen_device without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The streaming kernels' workspace (old interfaces,
 * kp, PCM mask and u/v fallback flags, plus the Robert-Asselin outputs
 * with hybgen_raflag: hycom_hybgen_noalloc_need() bytes, NIJ*(8*(kdm+1) +
 * 20), and 24*NIJ*kdm more with raflag, plus alignment) comes from one fixed device array
 * of hycom_hybgen_noalloc_capacity() bytes (256 MiB) per GPU, held from the
 * library's load until the process ends; larger domains fail with
 * HYB_EINVAL.  Nothing of it needs to start at zero: every word a kernel
 * reads is written earlier in the same call.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (two streams,
 * two threads) race and give undefined results.  check_finite is refused
 * (HYB_EUNSUPPORTED), the u/v phase always runs on the given stream, and
 * the call records no timing events.  Results are bitwise those of
 * hycom_hybgen_device. */
int hycom_hybgen_device_noalloc(const hyb_params* P, const hyb_fields* F,
                                int m, int n, int raflag, void* stream);
size_t hycom_hybgen_noalloc_capacity(void);
size_t hycom_hybgen_noalloc_need(const hyb_params* P, int raflag);

/* Per-column decisions of the last hycom_hybgen_device() call on this
 * thread's library context, for decision-parity checks: fixlay and kp of
 * every p-point (idm*jdm int32 each, i fastest, 0 on land).  Device
 * pointers; NULL disables recording (the default). */
int hycom_hybgen_record_decisions(int32_t* fixlay_dev, int32_t* kp_dev);

/* Testing hook (no reference counterpart): the FP64 quotient helpers the
 * hot kernels use in place of the compiler's division, beside that division,
 * on n device operands.  which: 0 a/b (qdiv), 1 a/b for a > 0 (qdiv_pos),
 * 2 1/b (qrcp), 3 a/b for any a (qdiv_any), 4 a/3 (div3).  fast[] gets the
 * helper's result and ieee[] the compiler's a/b (1/b, a/3).  Device
 * pointers; enqueued on the stream. */
int hycom_hybgen_selftest_quotients(int which, const double* a, const double* b,
                                    double* fast, double* ieee, int64_t n, void* stream);

/* Kernel selection (testing and measurement; no reference counterpart).
 * 0: automatic (streaming kernels when nhybrd==kdm, per-lane-array kernels
 * otherwise), 1: per-lane-array kernels only, 2: removed (HYB_EUNSUPPORTED),
 * 3: streaming kernels only (HYB_EUNSUPPORTED where they do not apply).
 * Process-wide. */
int hycom_hybgen_set_kernel_path(int path);

/* Variants of the streaming kernels (measurement and tuning; no reference
 * counterpart; results are bitwise the same for every value).  knob 0:
 * p-column remap variant, only 3 is built (3 fields per walk, 2 levels
 * loaded ahead, new interfaces staged in LDS); knob 1: levels loaded ahead
 * by the u/v walk, 0, 2 or 4 (the default); knob 2: the u/v walk's new
 * thicknesses kept in LDS (0) or formed again from p as the walk needs them
 * (1, the default, no LDS); knob 3: the regrid's window loaded one
 * iteration ahead (1, the default) or not (0); knob 4: the u/v phase on a
 * second stream of the library's, beside the p-column remap (1), or after
 * it on the caller's stream (0, the default).  Process-wide. */
int hycom_hybgen_tune(int knob, int value);

/* Measurement hooks (no reference counterpart; the reference's only timer
 * is xctmr 49 around the whole call, mod_hycom.F90:2648-2653).  While
 * enabled, every hycom_hybgen_device() call records HIP events around its
 * phases on its stream: 0 column kernel (hybgenaj), 1 p halo, 2 u/v kernel
 * (dpudpv + hybgenbj), 3 Robert-Asselin kernel (hybgencj), 4 second halo,
 * 5 second u/v kernel; within phase 0 of the streaming kernels, 6 the
 * regrid kernel alone and 7 the p-column remap kernel alone (without its
 * fallback columns).  _read waits for the recorded calls, returns the
 * summed milliseconds and call counts per phase, and forgets them. */
int hycom_hybgen_timing(int enable);
int hycom_hybgen_timing_read(double* ms_sum, int32_t* count, int nphase);

/* ---------------------------------------------------------------------- */
/* Synthetic inputs (workload generator; no reference counterpart)         */
/* ---------------------------------------------------------------------- */

/* Description of one synthetic domain.  The generator is a pure function
 * of (seed, global i, global j), evaluated with the same IEEE operations on
 * the host and on the device, so a window generated on the host is
 * bit-identical to the same window generated on the device. */
typedef struct hyb_synth {
  int64_t seed;
  int32_t itdm, jtdm;      /* global extents the window is part of         */
  int32_t i0, j0;          /* window offset in the global grid             */
  double  depth_min, depth_max;   /* m                                     */
  double  land_frac;       /* target land fraction                          */
  double  shelf_frac;      /* fraction of sea made shallow shelf (0..1)     */
  double  sigma_top, sigma_bot;   /* sigma-2 targets, evenly spaced         */
  double  uv_std;          /* m/s                                           */
  double  ts_noise;        /* psu, salinity noise std                       */
  double  topiso_pa;       /* value written to topiso (onem*isotop), Pa     */
  int32_t extreme;         /* 1: 40 C / 40 psu fronts and 1e-3 m layers     */
  int32_t reserved;
} hyb_synth;

/* ---------------------------------------------------------------------- */
/* Batched ragged domains (no reference counterpart: the reference runs    */
/* one domain per process; this serves ensembles and regional sets)       */
/* ---------------------------------------------------------------------- */

/* One domain of a batch: its parameters and its device field pointers. */
typedef struct hyb_domain {
  const hyb_params* P;
  const hyb_fields* F;
} hyb_domain;

/* One hybgen(m,n,.false.) step of each of ndom independent domains, of any
 * shapes and kdm (<= 64), FP64 or mixed per domain, enqueued on `stream`.
 * Every phase is one launch over all domains (one per remapper and
 * precision for the remaps), captured once into a HIP graph per batch and
 * replayed; the same domain set on the next call replays the same graph.
 * Each domain's result is bitwise the one of hycom_hybgen_device on that
 * domain alone.  Requires nhybrd == kdm and no itracr diagnostics
 * (HYB_EUNSUPPORTED otherwise); HYBGEN_DENFIX, FP64 and mixed domains may
 * share a batch. */
int hycom_hybgen_batch(const hyb_domain* D, int ndom, int m, int n, void* stream);

/* Online autotuning of the batch kernels' launch shape: while enabled (the
 * default), the first calls of a new domain set each try one candidate,
 * timed with HIP events, and the fastest is kept for that set.  First the
 * remap workgroups: columns per workgroup (64, 128, 256) times the LDS
 * split, the interface levels each column stages in LDS (all kmax-1, or
 * 24 or 12 with the deeper levels read from p as the walk reaches them;
 * at most 64 KB per workgroup); then the regrid's columns per workgroup
 * (256, 128, 64) with the best remap shape.  Results do not depend on the
 * shape. */
int hycom_hybgen_batch_autotune(int enable);

/* Widths in use for the last batch: columns per workgroup of the remap
 * and of the regrid; *tuned is 1 once the search has finished. */
int hycom_hybgen_batch_config(int32_t* remap_cols, int32_t* regrid_cols, int32_t* tuned);

/* The LDS split in use for the last batch: interface levels per column
 * staged in LDS by the remap (*levels) and the batch's largest kdm (*kmax;
 * kmax-1 levels = all staged). */
int hycom_hybgen_batch_staged_levels(int32_t* levels, int32_t* kmax);

/* The autotuner's candidates for the last batch, in search order: remap
 * columns, staged interface levels, regrid columns and the best of the two
 * timed launches in ms (-1 while not measured), up to cap of them.
 * Returns the number of candidates (or a negative HYB_E* code). */
int hycom_hybgen_batch_tuning(int32_t* remap_cols, int32_t* staged_levels,
                              int32_t* regrid_cols, float* ms, int cap);

/* ---------------------------------------------------------------------- */
/* Request loop over resident domains (no reference counterpart: serving) */
/* ---------------------------------------------------------------------- */

/* A server keeps many domains' state resident on the current device and
 * steps them on request.  A request asks for nsteps hybgen(1,2,.false.)
 * calls of one domain, FP64 or mixed.  A scheduler thread runs
 * iterations: each iteration steps every active request once, all
 * domains together in the batched kernels; requests that arrive while
 * others run are admitted at the next iteration as slots and device
 * memory free up (continuous batching), and a request leaves the active
 * set after its last step.  Requests of one domain run in submission
 * order.  A domain's state is created on the device (synthetic generator
 * or a copy of host arrays) when a request for it is first admitted;
 * when the memory budget is short, idle domains are evicted to host
 * memory, least recently used first, and come back on their next request.
 * Each result is bitwise the one of hycom_hybgen_device run alone. */
typedef struct hyb_server hyb_server;

typedef struct hyb_server_config {
  int64_t hbm_budget;   /* device bytes for resident domains; 0: 90% of free */
  int32_t max_active;   /* requests stepped per iteration, 1..256; 0: 64    */
  int32_t remap_cols;   /* columns per remap workgroup 64|128|256; 0: 64    */
  int32_t regrid_cols;  /* columns per regrid workgroup 64|128|256; 0: 256  */
  int32_t reserved;
} hyb_server_config;

typedef struct hyb_completion {
  int64_t tag;            /* the caller's tag from submit                    */
  int32_t domain;
  int32_t status;         /* 0, or the HYB_E* code the request failed with  */
  int32_t nsteps, precision;
  double t_submit;        /* seconds since the server was created           */
  double t_start;         /* admitted into an iteration                      */
  double t_done;          /* last iteration seen complete on the device      */
  int64_t column_layers;  /* nsteps x sea columns x kdm                      */
} hyb_completion;

typedef struct hyb_server_stats {
  int64_t iterations, completed, column_layers;
  int64_t resident_bytes, budget_bytes;
  int64_t opens, evictions;
  int32_t resident_domains, max_active_seen;
  double active_sum;      /* sum over iterations of the active requests     */
} hyb_server_stats;

int hycom_hybgen_server_create(const hyb_server_config* C, hyb_server** out);
/* Cancels queued requests (status HYB_EBUSY), waits for the device, frees. */
int hycom_hybgen_server_destroy(hyb_server* srv);
/* Registers a domain.  Its initial state is the synthetic generator S (both
 * time levels, m=1 n=2) or a copy of the host arrays F taken now; the
 * device copy is made at first admission.  P->precision is ignored (each
 * request names its own).  Same limits as hycom_hybgen_batch. */
int hycom_hybgen_server_add_domain(hyb_server* srv, const hyb_params* P, const hyb_synth* S,
                                   const hyb_fields* F, int32_t* domain_id);
int hycom_hybgen_server_submit(hyb_server* srv, int32_t domain_id, int32_t nsteps,
                               int32_t precision, int64_t tag);
/* Moves up to max completions into out (*n of them), waiting up to wait_ms
 * for the first one.  Returns the server's error code once it has failed. */
int hycom_hybgen_server_poll(hyb_server* srv, hyb_completion* out, int32_t max,
                             int32_t wait_ms, int32_t* n);
/* Copies a domain's current state into host arrays (NULL fields skipped);
 * HYB_EBUSY while requests of the domain are queued or running. */
int hycom_hybgen_server_read_domain(hyb_server* srv, int32_t domain_id, const hyb_fields* F);
int hycom_hybgen_server_stats(hyb_server* srv, hyb_server_stats* st);

/* Fills every field of F (host pointers) and theta/masks/depths for the
 * window described by S and P, in the layout above, time level n, and
 * copies level n to level m (m != n).  Parameters in P (dp0k, thbase...)
 * must already be set. */
int hycom_hybgen_synth_host(const hyb_params* P, const hyb_synth* S,
                            const hyb_fields* F, int m, int n);

/* Same on device pointers, ordered on `stream`. */
int hycom_hybgen_synth_device(const hyb_params* P, const hyb_synth* S,
                              const hyb_fields* F, int m, int n,
                              void* stream);

/* ---------------------------------------------------------------------
 * Stokes drift onto the model layers: stokes_vertical_j(n,j)
 * (mod_stokes.F90:650-705) for j = 1..jj, as stokes_forfun calls it
 * (mod_stokes.F90:541-546).  At every sea p-point the two snapshots of the
 * Stokes drift on nsdzi fixed layers are blended,
 *   s(k) = ws0*usdz(i,j,1,k) + ws1*usdz(i,j,2,k)   (and vsdz),
 * and remapped by PLM (hybgen_plm_coefs on max(dprs,dpthin), then
 * hybgen_plm_remap) from the interfaces min(sdzi, p(kk+1)) onto the model
 * interfaces prsf(k+1) = min(prsf(k) + dp(i,j,k,n), p(kk+1)).  Results go
 * to usdp, vsdp at sea points; land points are not written.
 * ------------------------------------------------------------------- */
typedef struct hyb_stokes_params {
  int32_t idm, jdm;      /* array extents without halo                      */
  int32_t ii, jj;        /* active extents                                  */
  int32_t nbdy;          /* halo width                                      */
  int32_t kdm;           /* model layers (1..HYB_KMAX)                      */
  int32_t nsdzi;         /* fixed Stokes interfaces (2..HYB_KMAX)           */
  int32_t reserved;      /* 0                                               */
  double  ws0, ws1;      /* snapshot weights (mod_stokes.F90:513-514)       */
} hyb_stokes_params;

typedef struct hyb_stokes_fields {
  const double *usdz, *vsdz;  /* (NI,NJ,2,nsdzi) the two snapshots          */
  const double *sdzi;         /* (nsdzi) interface depths, Pa; host memory  */
  const double *dp;           /* (NI,NJ,kdm,2)                              */
  const double *p;            /* (NI,NJ,kdm+1)                              */
  const int32_t *ip;          /* (NI,NJ)                                    */
  double *usdp, *vsdp;        /* (NI,NJ,kdm) results                        */
} hyb_stokes_fields;

/* Drop-in for the loop  do j=1,jj; call stokes_vertical_j(n,j)
 * (mod_stokes.F90:541-546): host arrays in and out, n the 1-based time
 * level of dp.  Returns when usdp and vsdp hold the result. */
int hycom_stokes_vertical(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n);

/* Same on device pointers (sdzi stays host memory), ordered on `stream`;
 * returns after enqueueing.  Allocates nothing. */
int hycom_stokes_vertical_device(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n,
                                 void* stream);

/* ---------------------------------------------------------------------
 * Tracer column remaps of trcupd.F90: pcmtrc (808-909, method 0) and
 * plmtrc with its slopes plmtrcx/plmtrcs (911-1147, method 1), applied at
 * every point where mask != 0, as initrc (trcupd.F90:167-199) and the
 * tracer relaxation of thermf (thermf.F90:770-806) call them per sea
 * column: si(ki,ks) on the interfaces pi(ki+1) onto so(ko,ks) on po(ko+1).
 * Interfaces are non-decreasing, pi(1) = 0 = po(1).  Where the reference
 * would stop ('bad lf', trcupd.F90:866-869: po out of order) that output
 * is NaN (thin outputs below it copy it) and the call returns HYB_EINVAL
 * (host call) after writing the rest.  A thin first output (po(2)-po(1) < 1e-6) takes si(1);
 * the reference reads outside its array there.
 * ------------------------------------------------------------------- */
typedef struct hyb_trc_params {
  int32_t idm, jdm;      /* array extents without halo                      */
  int32_t ii, jj;        /* active extents                                  */
  int32_t nbdy;          /* halo width                                      */
  int32_t ki, ko;        /* input and output layers (2..HYB_KMAX, 1..HYB_KMAX) */
  int32_t ks;            /* fields (1..HYB_TRMAX)                           */
  int32_t method;        /* 0 pcmtrc, 1 plmtrc                              */
  int32_t reserved;      /* 0                                               */
} hyb_trc_params;

typedef struct hyb_trc_fields {
  const double *si;      /* (NI,NJ,ki,ks)                                   */
  const double *pi;      /* (NI,NJ,ki+1)                                    */
  const double *po;      /* (NI,NJ,ko+1)                                    */
  const int32_t *mask;   /* (NI,NJ): columns remapped where != 0            */
  double *so;            /* (NI,NJ,ko,ks) results at masked points          */
} hyb_trc_fields;

/* Host arrays in and out; returns when so holds the result. */
int hycom_trc_remap(const hyb_trc_params* P, const hyb_trc_fields* F);

/* ---------------------------------------------------------------------
 * HYCOM .a/.b files (README.src.mod_za:46-60).  An .a file holds 2-D
 * records of idm*jdm 32-bit IEEE reals, big-endian, in Fortran order, each
 * padded with 2^100 to a multiple of 4096 words; 2^100 marks a data void.
 * ------------------------------------------------------------------- */

/* Reads records rec0..rec0+nrec-1 (1-based) of an .a file into
 * out(1-nbdy:idm+nbdy, 1-nbdy:jdm+nbdy, nrec), as zaiord3 does
 * (mod_za_sm.h:678-862, with zaio_endian machi_c.c:50-67): with mask NULL
 * every interior point is written; with a mask (NI,NJ) only points where
 * mask != 0, and a void there is an error (HYB_EINVAL).  Halo points are
 * not written.  hmin[r], hmax[r] (may be NULL) get each record's range
 * over the points read, voids left out, as zaiord returns it. */
int hycom_za_read(const char* apath, int32_t idm, int32_t jdm, int32_t nbdy, int32_t rec0,
                  int32_t nrec, const int32_t* mask, double* out, double* hmin, double* hmax);

/* Same into device memory (out a device pointer; mask must be NULL): the
 * records are staged as 32-bit words and widened on the device, ordered on
 * `stream`; returns when the records are in place. */
int hycom_za_read_device(const char* apath, int32_t idm, int32_t jdm, int32_t nbdy,
                         int32_t rec0, int32_t nrec, const int32_t* mask, double* out,
                         double* hmin, double* hmax, void* stream);

/* The state hybgen reads, from a model restart pair (restart_in,
 * mod_restart.F90:14-651): u, v, dp, temp, saln (2 time levels each),
 * th3d (read from an old "RESTART:" file, else sig(temp,saln)-thbase at
 * ip == 1 and 0 elsewhere, as restart_in forms it), ubavg, vbavg (3 levels)
 * and, when P->ntracr > 0, the tracers, into host fields of F (NULL fields
 * skipped).  Every record is found by its .b line (name in columns 1-8)
 * and its range checked against that line to 1e-4, as restart_in3d does
 * (HYB_EINVAL on a mismatch or a missing field).  P gives the extents,
 * kdm, ntracr and thbase; F->ip is read for th3d.  nstep and dtime (may be
 * NULL) get the header's values. */
int hycom_restart_read(const char* apath, const char* bpath, const hyb_params* P,
                       const hyb_fields* F, int32_t* nstep, double* dtime);

/* The nrec consecutive records of one named field of a restart pair (the
 * name as restart_out writes it in columns 1-8 of the .b line, e.g. "pbot",
 * "pbavg"), range-checked as above, into out(NI,NJ,nrec); halo points are
 * not written.  restart_in3d (mod_restart.F90:581-617). */
int hycom_restart_read_field(const char* apath, const char* bpath, const hyb_params* P,
                             const char* name, int32_t nrec, double* out);

/* ---------------------------------------------------------------------
 * Model start on one tile holding the whole domain (i0 = j0 = 0, ii = idm,
 * jj = jdm), host arrays (NI,NJ) with the halo:
 *   1. hycom_grid_masks: the bathymetry as geopar.F90:229-256 reads it
 *      (values above 2^99 are land) and bigrid.F90:1-330 turns it into the
 *      p, u and v masks;
 *   2. hycom_restart_read (th3d uses ip) and hycom_restart_read_field for
 *      "pbot";
 *   3. hycom_state_start: the fields mod_hycom.F90:1684-1722 derives before
 *      the first step (depthu, depthv, the dp halo, p, dpu, dpv).
 * The result is the state the model's first hybgen call sees.
 * ------------------------------------------------------------------- */

/* bigrid on the bathymetry depths (m): interior voids become 0
 * (geopar.F90:244-252), the halo is filled as bigrid fills it (land beyond
 * closed boundaries); ip, iu, iv are written at every point, halo included.
 * *nreg gets the region type bigrid finds in the bathymetry: 1 if the last
 * column has sea (periodic in i), 2 if the last row has sea too (global,
 * arctic fold: the tripole halo of the model's ARCTIC build), else 0.
 * HYB_EINVAL where bigrid stops: a sea point with 4 land neighbours, or
 * sea in the last row of a domain that is not periodic.  The f-plane
 * (mapflg 4) is not handled. */
int hycom_grid_masks(int32_t idm, int32_t jdm, int32_t nbdy, double* depths, int32_t* ip,
                     int32_t* iu, int32_t* iv, int32_t* nreg);

typedef struct hyb_start_fields {
  double *pbot;                  /* (NI,NJ) Pa, interior in; halo filled     */
  double *dp;                    /* (NI,NJ,kdm,2) interior in; halo filled   */
  const int32_t *ip, *iu, *iv;   /* (NI,NJ) from hycom_grid_masks            */
  double *depthu, *depthv;       /* (NI,NJ) out, Pa                          */
  double *p;                     /* (NI,NJ,kdm+1) out: level 2 at sea points,
                                    p(:,:,1) = 0 everywhere; other land values
                                    are left as they are                     */
  double *dpu, *dpv;             /* (NI,NJ,kdm,2) out                        */
} hyb_start_fields;

/* mod_hycom.F90:1684-1722 with no open boundaries (lbflag 0): dpthuv's
 * depthu = min(pbot, pbot(i-1)) and depthv at u and v points (dpthuv.F90:
 * 48-123), the dp halo, then for time levels 1 and 2 p from dp at sea
 * points and dpudpv (dpudpv.F90:85-100) out to nbdy-1.  Halos are those of
 * xctilr with P->vland at closed boundaries.  P gives idm = ii, jdm = jj,
 * nbdy, kdm, nreg (0, 1 or 2, from hycom_grid_masks) and vland (1.0 after
 * geopar). */
int hycom_state_start(const hyb_params* P, const hyb_start_fields* F);

/* Same on device pointers, ordered on `stream`; returns after enqueueing
 * (the output where the reference would stop is left NaN, not reported).
 * Calls on one device share the first-bad-column word: do not overlap them
 * on different streams. */
int hycom_trc_remap_device(const hyb_trc_params* P, const hyb_trc_fields* F, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HYCOM_HYBGEN_H */

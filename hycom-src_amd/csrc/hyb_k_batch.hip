// hyb_k_batch.hip -- batched ragged domains: every phase of hybgen over
// all domains of a batch in one launch per kernel (one per remapper and
// precision group for the remaps).  A launch's blocks are laid out domain
// after domain; block b finds its domain by binary search in the group's
// block prefix, then its row j and column slice from the domain's shape.
// The domain's parameters, field view and workspace are read from device
// arrays (uniform per block, so they come in through scalar loads).
#include "hyb_launch.h"
#include "hyb_stream.h"

namespace hyb {

// A pointer read from memory is a generic (flat) address to the compiler,
// so every access through it becomes a flat instruction, which counts
// against both the vector-memory and the LDS/scalar wait counters.  A
// pointer loaded from the constant address space is taken to be global
// (only the host fills that memory, and it can only name global memory),
// so the domain's field and workspace pointers are read through it, one
// member at a time: the accesses become global loads and stores, as in the
// single-domain kernels, whose pointers arrive as arguments.
template <class T>
using cview = __attribute__((address_space(4))) const T;
__device__ __forceinline__ View gview(const View* a, int d) {
  const cview<View>* q = (const cview<View>*)(a + d);
  View v;
  v.dp = q->dp;
  v.dpo = q->dpo;
  v.dpu = q->dpu;
  v.dpv = q->dpv;
  v.temp = q->temp;
  v.saln = q->saln;
  v.th3d = q->th3d;
  v.u = q->u;
  v.v = q->v;
  v.tracer = q->tracer;
  v.p = q->p;
  v.pu = q->pu;
  v.pv = q->pv;
  v.theta = q->theta;
  v.depths = q->depths;
  v.topiso = q->topiso;
  v.depthu = q->depthu;
  v.depthv = q->depthv;
  v.ubavg = q->ubavg;
  v.vbavg = q->vbavg;
  v.ip = q->ip;
  v.iu = q->iu;
  v.iv = q->iv;
  v.NI = q->NI;
  v.NIJ = q->NIJ;
  v.nb = q->nb;
  v.K = q->K;
  v.rec_fixlay = q->rec_fixlay;
  v.rec_kp = q->rec_kp;
  v.rec_ld = q->rec_ld;
  return v;
}
__device__ __forceinline__ Scratch gscratch(const Scratch* a, int d) {
  const cview<Scratch>* q = (const cview<Scratch>*)(a + d);
  return Scratch{q->pres, q->kp, q->lcm, q->uvf};
}

// largest g with prefix[g] <= b (prefix[0] = 0, prefix[ng] = total)
__device__ __forceinline__ int group_of(const int32_t* prefix, int ng, int b) {
  int lo = 0, hi = ng;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// block -> (domain, j, first i) for a row kernel with `bw` threads per
// block, rows mapped as i_aligned0 maps them
struct RowPos {
  int d, j, i0;
};
__device__ __forceinline__ RowPos row_pos(const BatchDevice& B, const BatchGroup& G, int bw,
                                          int b) {
  const int g = group_of(G.prefix, G.ndom, b);
  const int d = G.dom[g];
  const int local = b - G.prefix[g];
  const int nbi = (B.P[d].ii + kAlignPad + bw - 1) / bw;
  const int j = local / nbi + 1;
  return RowPos{d, j, i_aligned0(gview(B.V, d), j) + (local % nbi) * bw};
}

template <int BW>
__global__ void __launch_bounds__(BW) __attribute__((amdgpu_waves_per_eu(3)))
kb_regrid(BatchDevice B, BatchGroup G, int n) {
  const RowPos r = row_pos(B, G, BW, blockIdx.x);
  const KParams& P = B.P[r.d];
  const View V = gview(B.V, r.d);
  const int i = r.i0 + threadIdx.x;
  if (i < 1 || i > P.ii) return;
  if (V.ip[ij_of(V, i, r.j)] == 0) return;
  regrid_stream<1>(P, V, gscratch(B.W, r.d), i, r.j, n);
}

// CW columns per workgroup; each lane's LDS column has stride CW doubles
// and holds ns levels of the new interfaces: all of them (SPLIT false), or
// the upper ns with the rest read from p (the split costs the walk a few
// registers and branches, so the whole column keeps its own instance)
template <int METHOD, bool MIXED, int CW, bool SPLIT>
__global__ void __launch_bounds__(CW) kb_remap(BatchDevice B, BatchGroup G, int n, int ns) {
  extern __shared__ double lds[];
  const RowPos r = row_pos(B, G, CW, blockIdx.x);
  const KParams& P = B.P[r.d];
  const View V = gview(B.V, r.d);
  const Scratch W = gscratch(B.W, r.d);
  const int i = r.i0 + threadIdx.x;
  if (i < 1 || i > P.ii) return;
  const int64_t c0 = ij_of(V, i, r.j);
  const int32_t ipw = V.ip[c0], kpw = W.kp[c0];  // kp of a land column: read, not used
  if ((ipw == 0) | ((kpw & kUnsafe) != 0)) return;
  if (SPLIT) {
    SplitInterfaces po = stage_split(lds + threadIdx.x, CW, V.p + c0, V.NIJ, P.kk, ns);
    if (MIXED) remap_mixed<METHOD, 3, 2>(P, V, W, i, r.j, n, po);
    else remap_stream<METHOD, 3, 2>(P, V, W, i, r.j, n, po);
  } else {
    StagedInterfaces po = stage_interfaces(lds + threadIdx.x, CW, V.p + c0, V.NIJ, P.kk);
    if (MIXED) remap_mixed<METHOD, 3, 2>(P, V, W, i, r.j, n, po);
    else remap_stream<METHOD, 3, 2>(P, V, W, i, r.j, n, po);
  }
}

__global__ void __launch_bounds__(kWave) kb_remap_fb(BatchDevice B, BatchGroup G, int n) {
  const RowPos r = row_pos(B, G, kWave, blockIdx.x);
  const KParams& P = B.P[r.d];
  const View V = gview(B.V, r.d);
  const Scratch W = gscratch(B.W, r.d);
  const int i = r.i0 + threadIdx.x;
  if (i < 1 || i > P.ii) return;
  const int64_t c0 = ij_of(V, i, r.j);
  if (V.ip[c0] == 0) return;
  if (!(W.kp[c0] & kUnsafe)) return;
  remap_fallback(P, V, W, i, r.j, n);
}

// p halo of every domain, levels 2..kk+1: the j rows, then the i columns
// (mod_xc_sm.h:1374-1424).  Blocks: kk per domain and 256-wide slice.
__global__ void __launch_bounds__(kBlock) kb_halo_j(BatchDevice B, BatchGroup G) {
  const int g = group_of(G.prefix, G.ndom, blockIdx.x);
  const int d = G.dom[g];
  const KParams& P = B.P[d];
  const int local = blockIdx.x - G.prefix[g];
  const int nbi = (P.ii + kBlock - 1) / kBlock;
  const int k = local / nbi + 2;
  const int i = (local % nbi) * kBlock + threadIdx.x + 1;
  if (i > P.ii) return;
  p_halo_jrow(P, gview(B.V, d), i, k);
}
__global__ void __launch_bounds__(kBlock) kb_halo_i(BatchDevice B, BatchGroup G) {
  const int g = group_of(G.prefix, G.ndom, blockIdx.x);
  const int d = G.dom[g];
  const KParams& P = B.P[d];
  const int local = blockIdx.x - G.prefix[g];
  const int nbj = (P.jj + 2 + kBlock - 1) / kBlock;
  const int k = local / nbj + 2;
  const int j = (local % nbj) * kBlock + threadIdx.x;
  if (j > P.jj + 1) return;
  p_halo_icol(P, gview(B.V, d), j, k);
}

// u- and v-points: the first half of a domain's blocks are u, then v
template <int METHOD, bool MIXED>
__global__ void __launch_bounds__(kWave) kb_uv(BatchDevice B, BatchGroup G, int nl) {
  extern __shared__ double lds[];
  const int g = group_of(G.prefix, G.ndom, blockIdx.x);
  const int d = G.dom[g];
  const KParams& P = B.P[d];
  const View V = gview(B.V, d);
  const int nbi = (P.ii + kAlignPad + kWave - 1) / kWave;
  int local = blockIdx.x - G.prefix[g];
  const bool isv = local >= P.jj * nbi;
  if (isv) local -= P.jj * nbi;
  const int j = local / nbi + 1;
  const int i = i_aligned0(V, j) + (local % nbi) * kWave + threadIdx.x;
  if (i < 1 || i > P.ii) return;
  const int64_t c0 = ij_of(V, i, j);
  const int32_t* msk = isv ? V.iv : V.iu;
  if (msk[c0] == 0) return;
  LaneRow dcol{lds + threadIdx.x, kWave};  // unused: new thicknesses formed again from p
  const bool ok = uv_stream<METHOD, 4, MIXED, true>(P, V, i, j, nl, isv, dcol);
  gscratch(B.W, d).uvf[c0 + (isv ? V.NIJ : 0)] = ok ? 0 : 1;
}

__global__ void __launch_bounds__(kWave) kb_uv_fb(BatchDevice B, BatchGroup G, int nl) {
  const int g = group_of(G.prefix, G.ndom, blockIdx.x);
  const int d = G.dom[g];
  const KParams& P = B.P[d];
  const View V = gview(B.V, d);
  const int nbi = (P.ii + kAlignPad + kWave - 1) / kWave;
  int local = blockIdx.x - G.prefix[g];
  const bool isv = local >= P.jj * nbi;
  if (isv) local -= P.jj * nbi;
  const int j = local / nbi + 1;
  const int i = i_aligned0(V, j) + (local % nbi) * kWave + threadIdx.x;
  if (i < 1 || i > P.ii) return;
  const int64_t c0 = ij_of(V, i, j);
  const int32_t* msk = isv ? V.iv : V.iu;
  if (msk[c0] == 0) return;
  if (gscratch(B.W, d).uvf[c0 + (isv ? V.NIJ : 0)] == 0) return;
  point_c<64>(P, V, i, j, nl, isv);
}

template <int CW, bool SPLIT>
static void remap_group_cw(const BatchDevice& B, const BatchGroup& G, int n, int ns,
                           hipStream_t s) {
  const dim3 grid(G.nblocks), blk(CW);
  const size_t lds = sizeof(double) * ns * CW;
  switch (G.key) {  // hybmap + 4 * precision
    case 0: hipLaunchKernelGGL((kb_remap<0, false, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
    case 1: hipLaunchKernelGGL((kb_remap<1, false, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
    case 2: hipLaunchKernelGGL((kb_remap<2, false, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
    case 3: hipLaunchKernelGGL((kb_remap<3, false, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
    case 4: hipLaunchKernelGGL((kb_remap<0, true, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
    case 5: hipLaunchKernelGGL((kb_remap<1, true, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
    case 6: hipLaunchKernelGGL((kb_remap<2, true, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
    default: hipLaunchKernelGGL((kb_remap<3, true, CW, SPLIT>), grid, blk, lds, s, B, G, n, ns); break;
  }
}

// the interface levels a remap workgroup stages per column: remap_stage
// (0: all, kmax-1), at most kmax-1
int batch_stage_levels(const BatchLaunch& L) {
  const int all = lds_levels(L.kmax);
  return (L.remap_stage > 0 && L.remap_stage < all) ? L.remap_stage : all;
}

static void remap_group(const BatchDevice& B, const BatchGroup& G, int n, int ns, bool split,
                        int cw, hipStream_t s) {
  if (split) {
    if (cw == 256) remap_group_cw<256, true>(B, G, n, ns, s);
    else if (cw == 128) remap_group_cw<128, true>(B, G, n, ns, s);
    else remap_group_cw<64, true>(B, G, n, ns, s);
  } else {  // every level staged: at most 64 KB with 64 or 128 columns
    if (cw == 256) remap_group_cw<256, false>(B, G, n, ns, s);
    else if (cw == 128) remap_group_cw<128, false>(B, G, n, ns, s);
    else remap_group_cw<64, false>(B, G, n, ns, s);
  }
}

static void uv_group(const BatchDevice& B, const BatchGroup& G, int nl, size_t lds,
                     hipStream_t s) {
  const dim3 grid(G.nblocks), blk(kWave);
  switch (G.key) {  // uv_method + 4 * precision
    case 0: hipLaunchKernelGGL((kb_uv<0, false>), grid, blk, lds, s, B, G, nl); break;
    case 1: hipLaunchKernelGGL((kb_uv<1, false>), grid, blk, lds, s, B, G, nl); break;
    case 3: hipLaunchKernelGGL((kb_uv<3, false>), grid, blk, lds, s, B, G, nl); break;
    case 4: hipLaunchKernelGGL((kb_uv<0, true>), grid, blk, lds, s, B, G, nl); break;
    case 5: hipLaunchKernelGGL((kb_uv<1, true>), grid, blk, lds, s, B, G, nl); break;
    default: hipLaunchKernelGGL((kb_uv<3, true>), grid, blk, lds, s, B, G, nl); break;
  }
}

hipError_t launch_batch(const BatchDevice& B, const BatchLaunch& L, int n, hipStream_t s) {
  if (L.regrid.nblocks > 0) {
    const dim3 grid(L.regrid.nblocks);
    if (L.regrid_cols == 256)
      hipLaunchKernelGGL(kb_regrid<256>, grid, dim3(256), 0, s, B, L.regrid, n);
    else if (L.regrid_cols == 128)
      hipLaunchKernelGGL(kb_regrid<128>, grid, dim3(128), 0, s, B, L.regrid, n);
    else
      hipLaunchKernelGGL(kb_regrid<64>, grid, dim3(64), 0, s, B, L.regrid, n);
  }
  const int ns = batch_stage_levels(L);
  const bool split = ns < lds_levels(L.kmax);
  for (int q = 0; q < L.nremap; q++) remap_group(B, L.remap[q], n, ns, split, L.remap_cols, s);
  if (L.rows64.nblocks > 0)
    hipLaunchKernelGGL(kb_remap_fb, dim3(L.rows64.nblocks), dim3(kWave), 0, s, B, L.rows64, n);
  if (L.halo_j.nblocks > 0)
    hipLaunchKernelGGL(kb_halo_j, dim3(L.halo_j.nblocks), dim3(kBlock), 0, s, B, L.halo_j);
  if (L.halo_i.nblocks > 0)
    hipLaunchKernelGGL(kb_halo_i, dim3(L.halo_i.nblocks), dim3(kBlock), 0, s, B, L.halo_i);
  for (int q = 0; q < L.nuv; q++) uv_group(B, L.uv[q], n, 0, s);
  if (L.uvall.nblocks > 0)
    hipLaunchKernelGGL(kb_uv_fb, dim3(L.uvall.nblocks), dim3(kWave), 0, s, B, L.uvall, n);
  return hipGetLastError();
}

}  // namespace hyb

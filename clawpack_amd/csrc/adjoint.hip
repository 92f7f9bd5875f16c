// Adjoint of the fused 2D unsplit shallow-water step (FP64, dt fixed) for gfx950, and its C ABI.
//
// lam_out = (d q^{n+1} / d q^n)^T lam_in, with q^{n+1} the step of claw2d_step from q^n.  The
// step is a sum of per-interface contributions (see adjoint.hpp), so its transpose is a sum of
// per-interface reverses, scattered onto the four cells each interface reads.  Kernels:
//   adj_sweeps    both sweep directions in one launch (they are independent).  x part: one
//                 wavefront per (row line, segment of 58 columns); lane = column = the x-interface
//                 on its left.  Each lane computes its interface's Roe split once; the limiter
//                 partners' data and, in reverse, their adjoints travel one lane left or right by
//                 DPP wavefront rotations; the cell adjoint of column c is the sum of the shares of interfaces
//                 c and c+1.  It writes the x-sweep share of every extended cell (interior + 2
//                 ghost layers) to the workspace's first three (my+4) x (mx+4) planes.  y part: the
//                 same along columns (lane = row position, strided loads; the waves of a
//                 workgroup take adjacent columns), into the next three planes.
//   adj_fold      lam_out = lam_in + both shares (interior) + the ghost-cell shares folded back to
//                 the interior cells they were copied from (extrap / periodic / wall, with the
//                 wall's sign), in a fixed order (deterministic, no atomics).
// Branches (entropy fix, limiter cases, upwind side, transverse speed signs) are taken as the
// forward took them at q^n.  No counterpart upstream: PyClaw has no adjoint.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/clawpack_amd.h"
#include "adjoint.hpp"

namespace claw {
int set_error(int code, const char* msg);
int set_hip_error(hipError_t e, const char* where);

namespace adj {

constexpr int kOwnA = 58;  // owned positions per wavefront (3 halo lanes left, 3 right)
constexpr int kWPBA = 4;   // wavefronts per workgroup

// One-lane exchanges as DPP wavefront rotations (VALU moves, no LDS round trip): rotl gives lane i
// the value of lane i-1 and lane 0 that of lane 63; rotr gives lane i the value of lane i+1 and
// lane 63 that of lane 0.  At the sweep kernel's 2 wavefronts per SIMD the crossbar's latency sat
// on the dependency chain: a reverse step ran 3% faster with these than with ds_bpermute.  Every
// call site is in wave-uniform control flow (all 64 lanes active).
template <int CTRL> __device__ __forceinline__ double dpp_rot(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  const int rlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const int rhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)rhi << 32) | (unsigned int)rlo);
}
constexpr int kDppWaveRor1 = 0x13C, kDppWaveRol1 = 0x134;
__device__ __forceinline__ double rotl(double v) { return dpp_rot<kDppWaveRor1>(v); }
__device__ __forceinline__ double rotr(double v) { return dpp_rot<kDppWaveRol1>(v); }
__device__ __forceinline__ LimD rotl(const LimD& d) {
  LimD o;
  o.a1 = rotl(d.a1); o.a2 = rotl(d.a2); o.a3 = rotl(d.a3);
  o.s1 = rotl(d.s1); o.s3 = rotl(d.s3); o.v = rotl(d.v);
  return o;
}
__device__ __forceinline__ LimD rotr(const LimD& d) {
  LimD o;
  o.a1 = rotr(d.a1); o.a2 = rotr(d.a2); o.a3 = rotr(d.a3);
  o.s1 = rotr(d.s1); o.s3 = rotr(d.s3); o.v = rotr(d.v);
  return o;
}

struct AdjArgs {
  const double* q;
  const double* lam;
  double* ext;
  double* lam_out;
  int mx, my;
  double dtdx, dtdy;
  Method M;
  int bxl, bxh, byl, byh, rfx, rfy;
  int nseg_x, nseg_y;
};

// state at (x, y) of the extended grid, ghost cells by index mapping, as frame (h, hn, ht)
template <int DIR>
__device__ __forceinline__ void load_q(const AdjArgs& a, int x, int y, double o[3]) {
  bool fx, fy;
  const int ix = bc_src(x, a.mx, a.bxl, a.bxh, fx);
  const int iy = bc_src(y, a.my, a.byl, a.byh, fy);
  const size_t plane = (size_t)a.mx * (size_t)a.my;
  const size_t off = (size_t)iy * (size_t)a.mx + (size_t)ix;
  double c[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double v = a.q[(size_t)m * plane + off];
    c[m] = ((fx && m == a.rfx) != (fy && m == a.rfy)) ? -v : v;
  }
  o[0] = c[0];
  o[1] = DIR == 0 ? c[1] : c[2];
  o[2] = DIR == 0 ? c[2] : c[1];
}

// lambda at (x, y) as frame; zero outside the interior (ghost cells receive no update)
template <int DIR>
__device__ __forceinline__ void load_lam(const AdjArgs& a, int x, int y, double o[3]) {
  if (x < 0 || x >= a.mx || y < 0 || y >= a.my) {
    o[0] = o[1] = o[2] = 0.0;
    return;
  }
  const size_t plane = (size_t)a.mx * (size_t)a.my;
  const size_t off = (size_t)y * (size_t)a.mx + (size_t)x;
  const double c0 = a.lam[off], c1 = a.lam[plane + off], c2 = a.lam[2 * plane + off];
  o[0] = c0;
  o[1] = DIR == 0 ? c1 : c2;
  o[2] = DIR == 0 ? c2 : c1;
}

// One sweep direction's share of the reverse step; block `blk` of the `nblk` this direction has.
// Direction DIR writes its own workspace planes (ext + 3 DIR eplane); adj_fold adds the two.
template <int DIR>
__device__ __forceinline__ void sweep_body(const AdjArgs& a, const int blk, const int nblk) {
  const int lane = threadIdx.x & 63;
  const int n_al = DIR == 0 ? a.mx : a.my;  // cells along the sweep line
  const int n_ln = DIR == 0 ? a.my : a.mx;  // sweep lines
  const int nseg = DIR == 0 ? a.nseg_x : a.nseg_y;
  // DIR 0: the waves of a line's segments are consecutive.  DIR 1 (lanes walk a column, so every
  // access is strided by the row pitch): the waves of a workgroup take adjacent columns of one
  // segment, and the workgroups that land on one XCD (dealt to the 8 XCDs in turn) take
  // consecutive column groups, so the cache lines of neighbouring columns are shared in that
  // XCD's L2 instead of being fetched by every XCD.  `bb` is a bijection of [0, nblk).
  // (The same renumbering for DIR 0 measured 9% slower: its accesses are already coalesced.)
  int bb = blk;
  if (DIR == 1) {
    const int nb = nblk, q8 = nb >> 3, r8 = nb & 7, x = blk & 7, i = blk >> 3;
    bb = x < r8 ? x * (q8 + 1) + i : r8 * (q8 + 1) + (x - r8) * q8 + i;
  }
  const int wave = __builtin_amdgcn_readfirstlane(bb * kWPBA + (threadIdx.x >> 6));
  // The column sweep shares lambda between the waves of a workgroup through LDS (a barrier), so
  // every one of its waves runs to the end; a wave past the last line or outside the swept lines
  // computes on clamped, in-bounds data and only stores zeros or nothing.
  const bool active = wave < (n_ln + 4) * nseg;
  if (DIR == 0 && !active) return;
  const int li = DIR == 0 ? wave / nseg : wave % (n_ln + 4);
  const int seg = DIR == 0 ? wave - li * nseg : wave / (n_ln + 4);
  const int line = li - 2;                  // -2 .. n_ln + 1
  const int p = seg * kOwnA - 5 + lane;     // position along: lanes 3..60 own
  const bool own = active && lane >= 3 && lane < 3 + kOwnA && p <= n_al + 1;
  const int ew = a.mx + 4;
  const size_t eplane = (size_t)ew * (size_t)(a.my + 4);
  const size_t eoff = DIR == 0 ? (size_t)(line + 2) * ew + (size_t)(p + 2) : (size_t)(p + 2) * ew + (size_t)(line + 2);
  double* const ext = a.ext + (size_t)(3 * DIR) * eplane;
  // the forward sweeps lines -1 .. n_ln only
  const bool swept = line >= -1 && line <= n_ln;
  if (DIR == 0 && !swept) {
    if (own) {
#pragma unroll
      for (int m = 0; m < 3; ++m) ext[(size_t)m * eplane + eoff] = 0.0;
    }
    return;
  }
  const double r = DIR == 0 ? a.dtdx : a.dtdy, rt = DIR == 0 ? a.dtdy : a.dtdx;

  // ---- forward: cells p-1 | p, the Roe split and entropy fix of interface p
  double qo[3];
  if (DIR == 0) load_q<0>(a, p, line, qo); else load_q<1>(a, line, p, qo);
  double ql[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) ql[m] = rotl(qo[m]);
  const CellF cR = cell_fwd(qo[0], qo[1], qo[2]);
  const CellF cL = cell_fwd(ql[0], ql[1], ql[2]);
  const WaveF w = wave_fwd(cL, cR, a.M.g);
  const FixF fx = fix_fwd(cL, cR, w, a.M.g, a.M.sqg);
  const LimD me = lim_of(w);
  const LimD nl = rotl(me), nr = rotr(me);

  // ---- adjoints of the six output cells: (p-1 | p) x (line-1, line, line+1)
  double r0[3], rm[3], rp[3];
  if (DIR == 0) {
    load_lam<0>(a, p, line, r0); load_lam<0>(a, p, line - 1, rm); load_lam<0>(a, p, line + 1, rp);
  } else {
    // column line's lambda (strided loads) goes to LDS; the waves of this workgroup hold adjacent
    // columns at the same positions, so a neighbour column inside the workgroup is read from
    // there and only the workgroup's two edge columns are loaded again
    __shared__ double lam_cols[kWPBA][3][64];
    const int wv = threadIdx.x >> 6;
    load_lam<1>(a, line, p, r0);
#pragma unroll
    for (int m = 0; m < 3; ++m) lam_cols[wv][m][lane] = r0[m];
    __syncthreads();
    // the previous (next) wave of the workgroup holds column line - 1 (line + 1) when it is in
    // the same segment, i.e. when this wave's column is not the first (last) of the segment
    const bool prev_here = wv > 0 && li > 0;
    const bool next_here = wv < kWPBA - 1 && li < n_ln + 3 && wave + 1 < (n_ln + 4) * nseg;
    if (prev_here) {
#pragma unroll
      for (int m = 0; m < 3; ++m) rm[m] = lam_cols[wv - 1][m][lane];
    } else {
      load_lam<1>(a, line - 1, p, rm);
    }
    if (next_here) {
#pragma unroll
      for (int m = 0; m < 3; ++m) rp[m] = lam_cols[wv + 1][m][lane];
    } else {
      load_lam<1>(a, line + 1, p, rp);
    }
  }
  double bLX[3], bLG1[3], bLG2[3], bRX[3], bRG1[3], bRG2[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double l0 = rotl(r0[m]), lm = rotl(rm[m]), lp = rotl(rp[m]);
    bLX[m] = l0;
    bLG1[m] = l0 - lm;
    bLG2[m] = lp - l0;
    bRX[m] = r0[m];
    bRG1[m] = r0[m] - rm[m];
    bRG2[m] = rp[m] - r0[m];
  }

  // ---- reverse
  WaveB Wb;
  zero(Wb);
  LimD NLb, NRb;
  zero(NLb);
  zero(NRb);
  CellB Lb, Rb;
  zero(Lb);
  zero(Rb);
  iface_core(cL, cR, w, fx, nl, nr, bLX, bLG1, bLG2, bRX, bRG1, bRG2, r, rt, a.M, Wb, NLb, NRb, Lb, Rb);
  // this interface is the right partner of interface p-1 and the left partner of p+1
  add_lim(Wb, rotl(NRb));
  add_lim(Wb, rotr(NLb));
  wave_rev(cL, cR, w, a.M.g, Wb, Lb, Rb);
  // cell p: right cell of interface p, left cell of interface p+1
  CellB tot = Rb;
  tot.h += rotr(Lb.h);
  tot.hn += rotr(Lb.hn);
  tot.ht += rotr(Lb.ht);
  tot.s += rotr(Lb.s);
  tot.r += rotr(Lb.r);
  tot.P += rotr(Lb.P);
  tot.Q += rotr(Lb.Q);
  double o[3];
  cell_rev(cR, tot, o);
  if (own) {
    const double c[3] = {o[0], DIR == 0 ? o[1] : o[2], DIR == 0 ? o[2] : o[1]};
#pragma unroll
    for (int m = 0; m < 3; ++m) ext[(size_t)m * eplane + eoff] = swept ? c[m] : 0.0;
  }
}

// Both sweep directions in one launch: blocks [0, nbx) take the x-sweep lines, the rest the
// y-sweep columns.  The two are independent (each writes its own planes), so they overlap on the
// device instead of running as two dependent launches.
// The reverse steps of a backward pass form one dependent chain, which sets the pass's time; the
// forward recompute of the next checkpoint segment runs beside it on another stream and has slack.
// Both reverse kernels therefore issue at the highest wavefront priority, so that where their
// wavefronts share a SIMD with the recompute's, theirs are picked first (backward -4% on a
// 512 x 512 grid; neither result changes).
__global__ void __launch_bounds__(64 * kWPBA) adj_sweeps(AdjArgs a, int nbx) {
  __builtin_amdgcn_s_setprio(3);
  const int b = (int)blockIdx.x;
  if (b < nbx) sweep_body<0>(a, b, nbx);
  else sweep_body<1>(a, b - nbx, (int)gridDim.x - nbx);
}

// extended index k (0..4 = i, -2, -1, n, n+1) along one axis: does its value come from cell i, and
// with the wall's sign flip?  (k = 0 is cell i itself.)
__device__ __forceinline__ bool source_k(int k, int i, int n, int lo, int hi, int& e, bool& fl) {
  e = k == 0 ? i : (k == 1 ? -2 : (k == 2 ? -1 : (k == 3 ? n : n + 1)));
  fl = false;
  if (k == 0) return true;
  return bc_src(e, n, lo, hi, fl) == i;
}

__global__ void __launch_bounds__(256) adj_fold(AdjArgs a) {
  __builtin_amdgcn_s_setprio(3);  // on the backward's critical chain, like adj_sweeps
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long cells = (long long)a.mx * a.my;
  if (t >= cells) return;
  const int j = (int)(t / a.mx), i = (int)(t - (long long)j * a.mx);
  const int ew = a.mx + 4;
  const size_t eplane = (size_t)ew * (size_t)(a.my + 4);
  const size_t plane = (size_t)cells;
  double acc[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) acc[m] = a.lam[(size_t)m * plane + t];
  const double* const ey_ = a.ext + 3 * eplane;  // the y-sweep's planes
  if (i >= 2 && i < a.mx - 2 && j >= 2 && j < a.my - 2) {
    const size_t eo = (size_t)(j + 2) * ew + (size_t)(i + 2);
#pragma unroll
    for (int m = 0; m < 3; ++m) acc[m] += a.ext[(size_t)m * eplane + eo] + ey_[(size_t)m * eplane + eo];
  } else {
    // fixed order: y candidate outer, x candidate inner
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      int ey;
      bool fy;
      if (!source_k(b, j, a.my, a.byl, a.byh, ey, fy)) continue;
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        int ex;
        bool fx;
        if (!source_k(c, i, a.mx, a.bxl, a.bxh, ex, fx)) continue;
        const size_t eo = (size_t)(ey + 2) * ew + (size_t)(ex + 2);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const double v = a.ext[(size_t)m * eplane + eo] + ey_[(size_t)m * eplane + eo];
          acc[m] += ((fx && m == a.rfx) != (fy && m == a.rfy)) ? -v : v;
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) a.lam_out[(size_t)m * plane + t] = acc[m];
}

}  // namespace adj
}  // namespace claw

using namespace claw;
using namespace claw::adj;

static int check_adjoint_desc(const claw2d_desc* d) {
  const int rc = claw2d_validate(d);
  if (rc != 0) return rc;
  if (d->solver != CLAW_SOLVER_SHALLOW_ROE_EFIX)
    return set_error(-20, "the adjoint step is implemented for shallow water (shallow_roe_with_efix) only");
  if (d->dtype != CLAW_FP64) return set_error(-21, "the adjoint step runs in FP64 only");
  if (d->num_aux != 0) return set_error(-23, "the adjoint step takes no aux arrays (no capacity function)");
  return 0;
}

extern "C" {

size_t claw2d_adjoint_workspace_bytes(const claw2d_desc* d) {
  if (check_adjoint_desc(d) != 0) return 0;
  return (size_t)6 * (size_t)(d->mx + 4) * (size_t)(d->my + 4) * sizeof(double);
}

int claw2d_adjoint_step(const claw2d_desc* d, const void* q, const void* lam_in, void* lam_out, void* workspace,
                        double dt, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  const int rc = check_adjoint_desc(d);
  if (rc != 0) return rc;
  if (!q || !lam_in || !lam_out || !workspace) return set_error(-1, "null device pointer");
  if (!(dt > 0.0)) return set_error(-22, "dt must be positive");
  AdjArgs a;
  a.q = (const double*)q;
  a.lam = (const double*)lam_in;
  a.ext = (double*)workspace;
  a.lam_out = (double*)lam_out;
  a.mx = d->mx;
  a.my = d->my;
  a.dtdx = dt / d->dx;
  a.dtdy = dt / d->dy;
  a.M.g = d->grav;
  a.M.sqg = sqrt(d->grav);
  a.M.order = d->order;
  a.M.tw = d->transverse_waves;
  for (int p = 0; p < 3; ++p) a.M.mthlim[p] = d->mthlim[p];
  a.bxl = d->bc_lower[0];
  a.bxh = d->bc_upper[0];
  a.byl = d->bc_lower[1];
  a.byh = d->bc_upper[1];
  a.rfx = d->reflect_index[0];
  a.rfy = d->reflect_index[1];
  a.nseg_x = (d->mx + 4 + kOwnA - 1) / kOwnA;
  a.nseg_y = (d->my + 4 + kOwnA - 1) / kOwnA;
  const long long wx = (long long)(d->my + 4) * a.nseg_x, wy = (long long)(d->mx + 4) * a.nseg_y;
  const long long cells = (long long)d->mx * d->my;
  if (wx > (1LL << 30) || wy > (1LL << 30)) return set_error(-23, "grid too large for the adjoint launch");
  const int nbx = (int)((wx + kWPBA - 1) / kWPBA), nby = (int)((wy + kWPBA - 1) / kWPBA);
  hipLaunchKernelGGL(adj_sweeps, dim3((unsigned)(nbx + nby)), dim3(64 * kWPBA), 0, stream, a, nbx);
  hipLaunchKernelGGL(adj_fold, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, stream, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_hip_error(e, "adjoint launch");
  return 0;
}

}  // extern "C"

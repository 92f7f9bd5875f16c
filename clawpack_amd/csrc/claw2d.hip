// Fused 2D unsplit wave-propagation step for gfx950 (MI355X) and its C ABI.
//
// Replaces, in one kernel per step, PyClaw's ClawSolver2D.step_hyperbolic -> classic step2 ->
// flux2 -> rpn2 / limiter / rpt2 chain (upstream pyclaw/src/pyclaw/classic/{solver.py,step2.f90,
// flux2.f90,limiter.f90}; submodule empty in the reference tree, .gitmodules:1-3), together with
// the ghost-cell fill (solver.py _apply_bcs, done here by index mapping) and the CFL maximum.
//
// Kernel organisation ("marching strip"): one wavefront owns a strip of 60 columns (64 lanes:
// 2 halo lanes each side, one column per lane) and a segment of seg_h rows, and marches up the
// rows.  Per row it
//   * loads the next row (coalesced, one SoA plane per component; prefetched one row ahead),
//   * solves the y-interface below the new row in-lane (the column's history stays in VGPRs),
//   * solves the row's x-interfaces with the left neighbour's cell data taken over the LDS
//     crossbar (ds_bpermute), limits them with the right neighbour's wave data, splits the
//     fluctuations transversely, and hands the left-cell share back (ds_bpermute),
//   * finishes the y-interface one row lower (limiter needs the next y-wave) and passes the
//     transverse y-fluxes to the neighbouring columns,
//   * writes the cell two rows below, whose 21-cell domain of dependence is now complete.
// Every interface is solved once per owner (halo redundancy 64/60 in x, (seg_h+4)/seg_h in y), q
// is read once and written once per step, and no intermediate array goes to HBM.  The bound is
// HBM bandwidth (2*num_eqn*sizeof(T) bytes per cell) or FP64 VALU throughput, whichever binds.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/clawpack_amd.h"
#include "physics.hpp"

namespace claw {

constexpr int kLanes = 64;
constexpr int kOwn = 60;  // owned columns per strip
constexpr int kWPB = 4;   // independent wavefronts per workgroup
// Shallow-water steady rows per segment from which issue priority follows progress (shorter
// segments ran 0.5-2% slower with it on MI355X: 2048^2 and below)
constexpr int kPrioRows = 64;
// the same for the Euler kernels (about twice the work per row): marched rows per segment
constexpr int kPrioRowsEuler = 16;

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
static int hip_fail(hipError_t e, const char* where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return (int)e;
}
#ifndef CLAW_KERNEL_UNIT  // a kernel unit (see below) holds templates instantiated there only
// error reporting for the library's other translation units (claw2d_last_error reads g_err)
int set_error(int code, const char* msg) { return fail(code, msg); }
int set_hip_error(hipError_t e, const char* where) { return hip_fail(e, where); }
#endif

// ------------------------------------------------------------------------------------------
// cross-lane helpers: ds_bpermute (LDS crossbar) and DPP wavefront shifts (VALU)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double shfl(double v, int addr) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float shfl(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// Neighbour exchange within a wavefront: lane i receives lane i-1's value (from_left) or lane
// i+1's (from_right); lane 0 (from_left) and lane 63 (from_right) keep their own value.  DPP:
// wavefront shifts (wave_shr:1 / wave_shl:1; a source lane out of range keeps the old value), VALU
// moves with no LDS round trip; otherwise ds_bpermute.  Every call site is in wave-uniform control
// flow (all 64 lanes active), where both forms give the same values.
// ZERO: the edge lane receives 0 instead of keeping its own value, which saves the copy that
// pre-loads the destination.  Only halo lanes 0 and 63 see it, and no owned lane (2..61) reads a
// value derived from it: owned cells take interfaces 2..62, whose inputs come from lanes 1..63.
template <int CTRL, bool ZERO = false> __device__ __forceinline__ int dpp_shift(int v) {
  if constexpr (ZERO) return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
  else return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL, bool ZERO = false> __device__ __forceinline__ double dpp_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_shift<CTRL, ZERO>((int)(b & 0xffffffffLL));
  const int hi = dpp_shift<CTRL, ZERO>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, bool ZERO = false> __device__ __forceinline__ float dpp_shift(float v) {
  return __int_as_float(dpp_shift<CTRL, ZERO>(__float_as_int(v)));
}
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;
template <bool DPP, class T, bool ZERO = false> __device__ __forceinline__ T from_left(T v) {
  if constexpr (DPP) {
    return dpp_shift<kDppWaveShr1, ZERO>(v);
  } else {
    const int lane = threadIdx.x & 63;
    return shfl(v, (lane > 0 ? lane - 1 : 0) * 4);
  }
}
template <bool DPP, class T, bool ZERO = false> __device__ __forceinline__ T from_right(T v) {
  if constexpr (DPP) {
    return dpp_shift<kDppWaveShl1, ZERO>(v);
  } else {
    const int lane = threadIdx.x & 63;
    return shfl(v, (lane < 63 ? lane + 1 : 63) * 4);
  }
}
// Which form a step kernel uses (measured on MI355X): DPP for FP64, where at 2 wavefronts per SIMD
// the ds_bpermute latency is on the critical path (shallow water -4.7%, Euler -9%); ds_bpermute
// for FP32.  A DPP move issues like an FP64 instruction, one wave64 per 4 cycles, while FP32
// arithmetic issues in 2, so a 32-bit shift costs an FP32 kernel two arithmetic slots, and at 3-4
// wavefronts per SIMD the crossbar's latency is hidden (FP32 shallow water -8%, FP32 Euler -5.9%).
template <class T, class P> constexpr bool kDppShift = sizeof(T) == 8;

// max that lets a NaN through (so a NaN speed poisons the CFL number instead of vanishing)
template <class T> __device__ __forceinline__ T nmax(T acc, T s) { return (s > acc || s != s) ? s : acc; }
// FP32: the IEEE 754-2019 maximum, one v_maximum3_f32 on gfx950 (a NaN operand gives NaN; the speeds
// are never -0, so it equals the form above).  The FP32 step runs 2.3% faster with it.
template <> __device__ __forceinline__ float nmax(float acc, float s) { return __builtin_elementwise_maximum(acc, s); }

template <class T> __device__ __forceinline__ T wave_max(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = nmax(v, shfl(v, (lane ^ off) * 4));
  return v;
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Ghost cells by index mapping (pyclaw _bc_lower/_bc_upper: extrap copies the edge cell,
// periodic wraps, wall mirrors and negates the normal momentum).  x is filled before y in pyclaw,
// which for these three rules is the same as composing the two separable maps.
__device__ __forceinline__ void map_index(int i, int n, int lo, int hi, int& src, bool& flip) {
  flip = false;
  src = i;
  if (i < 0) {
    if (lo == CLAW_BC_PERIODIC) src = i + n;
    else if (lo == CLAW_BC_WALL) { src = -1 - i; flip = true; }
    else src = 0;
  } else if (i >= n) {
    if (hi == CLAW_BC_PERIODIC) src = i - n;
    else if (hi == CLAW_BC_WALL) { src = 2 * n - 1 - i; flip = true; }
    else src = n - 1;
  }
  src = src < 0 ? 0 : (src > n - 1 ? n - 1 : src);  // lanes far outside the grid carry unused data
}

// Workgroups are dealt round-robin to the 8 XCDs; renumber so that each XCD gets a contiguous
// range of strips/segments (neighbouring strips share halo lines in that XCD's L2).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int full = nb & ~7;
  if (b >= full) return b;
  const int per = full >> 3;
  return (b & 7) * per + (b >> 3);
}

template <class T, class P> struct StepArgs {
  T* q0;
  T* q1;
  claw2d_ctrl* ctrl;
  double dt_fixed;
  double dx, dy;
  int use_ctrl;
  int mx, my, nstrips, nsegs, seg_h;
  int bcx_lo, bcx_hi, bcy_lo, bcy_hi, rfx, rfy;
  int order, tw;
  int mthlim[4];
  // 0: no controller; 1: the last wavefront writes the CFL number of the fixed-dt step;
  // 2: the last wavefront runs the evolve_to_time controller on ctrl
  int fuse;
  int total_waves;
  double cfl_max, cfl_desired, dt_max;
  // aux plane 0 (the bathymetry of an f-wave policy) and the aux boundary conditions
  const T* aux;
  int abx_lo, abx_hi, aby_lo, aby_hi;
  // the capacity plane kappa (aux plane mcapa-1) of a capacity-form step, else null
  const T* capa;
  // launch shape: wavefronts per workgroup (1, 2 or 4) and XCD-contiguous order (1) or not (0)
  int wpb, xcd;
  typename P::Par par;
};

// Per-cell data of a policy, with its aux field when it reads one.
template <class P, class T>
__device__ __forceinline__ typename P::Cell prep_cell(const T q[P::NEQ], T b, const typename P::Par& pr) {
  if constexpr (P::NAUX > 0) return P::prep(q, b, pr);
  else return P::prep(q, pr);
}

// Limits one interface and forms its contributions (interface frame), following flux2.f90:
//   to the left/lower cell : LX = -r*(amdq + cqxx/2),  LG1/LG2 = rt * (-r/2) * B^-/B^+ (amdq + cqxx)
//   to the right/upper cell: RX = -r*(apdq - cqxx/2),  RG1/RG2 = rt * (-r/2) * B^-/B^+ (apdq - cqxx)
// (qadd, fadd += cqxx/2, cqxx folded into amdq/apdq for transverse_waves = 2, gadd), where r is
// dt/dx along the interface normal and rt the transverse dt/dy that step2.f90 applies to gadd.
// For an f-wave policy amdq = sum sf_p alpha_p e_p, apdq = sum (1 - sf_p) alpha_p e_p and the
// correction carries sign(s_p) where the wave form has |s_p| (flux2.f90 with use_fwave).
// upL[p] is the limiter's upwind term from the interface below/left along the sweep, alpha_p of
// that interface times e_p(neighbour).e_p(this), formed where the product is known; alR/mR are the
// wave strengths of the interface above/right and their eigenvector products.  (Given as one
// product, the left term reaches the select already multiplied: passed as alpha and m, the
// compiler rewrote select(s > 0, aL*mL, aR*mR) as select(aL, aR)*select(mL, mR), two selects more
// per wave.)
// TWC >= 0: the method is known at compile time (order 2, transverse_waves = TWC), so the
// order/transverse selects compile away; TWC < 0: order and transverse_waves from a.
// CAPA (capacity form, step2.f90/flux2.f90 with mcapa > 0): rkL/rkR are 1/kappa of the left and
// right cells.  Every contribution formed here is later scaled by 1/kappa of the cell that
// receives it (the capacity-form update divides a cell's flux differences by its kappa and qadd
// carries dt/(dx kappa) of the cell it updates); in addition the correction uses the average
// (dt/dx)(1/kappa_L + 1/kappa_R)/2 in place of dt/dx, and the transverse terms of amdq (apdq)
// carry dt/(dx kappa) of the left (right) cell.
template <class T, class P, int LIM, int TWC, bool CAPA = false>
__device__ __forceinline__ void finish_iface(const typename P::Iface& f, const T upL[P::NW],
                                             const T alR[P::NW], const T mR[P::NW], T r, T rt,
                                             const StepArgs<T, P>& a, T LX[P::NEQ], T LG1[P::NEQ],
                                             T LG2[P::NEQ], T RX[P::NEQ], T RG1[P::NEQ],
                                             T RG2[P::NEQ], T rkL = T(1), T rkR = T(1)) {
  constexpr int NEQ = P::NEQ, NW = P::NW;
  const bool second = TWC >= 0 ? true : a.order != 1;
  const int tw = TWC >= 0 ? TWC : a.tw;
  const bool fold = second && tw == 2;
  T in[NW];
  if (second) P::inv_norm(f, in);
  const T rc = CAPA ? T(0.5) * (r * rkL + r * rkR) : r;  // dt/dx of the correction
  T XL[NW], XR[NW], A[NW], B[NW];
#pragma unroll
  for (int p = 0; p < NW; ++p) {
    const T s = P::speed(f, p), al = f.al[p], sf = f.sf[p];
    const T e = sf * al, q = P::FWAVE ? al - e : (s - sf) * al;
    T k = T(0);
    if (second) {
      const T aL = upL[p], aR = alR[p], mr = mR[p];  // rvalues: no lvalue select
      const T t = (s > T(0) ? aL : aR * mr) * in[p];
      const int meth = LIM > 0 ? LIM : a.mthlim[p];
      const T sa = vabs(s);  // |s| as a source modifier (a select form costs two instructions)
      const T sw = P::FWAVE ? copysign(T(1), s) : sa;
      k = sw * (T(1) - sa * rc) * limited(al, t, meth);
    }
    XL[p] = -r * fma(T(0.5), k, e);
    XR[p] = -r * fma(T(-0.5), k, q);
    A[p] = fold ? e + k : e;
    B[p] = fold ? q - k : q;
  }
  P::expand_n(f, XL, LX);
  P::expand_n(f, XR, RX);
#pragma unroll
  for (int m = 0; m < NEQ; ++m) LG1[m] = LG2[m] = RG1[m] = RG2[m] = T(0);
  if (tw > 0) {
    T bA[NW], bB[NW];
    P::tbeta(f, A, bA);
    P::tbeta(f, B, bB);
    const T w = T(-0.5) * r * rt;
    const T wA = CAPA ? T(-0.5) * (r * rkL) * rt : w, wB = CAPA ? T(-0.5) * (r * rkR) * rt : w;
    // transversely subsonic on every lane (v - a < 0 < v + a): the down/up split has a fixed
    // pattern, so the zero entries are skipped; the result is the same bit for bit, because both
    // forms take their interface constants from the one tc formed here
    const typename P::TConst tc = P::tconst(f);
    const bool sub = tc.vm < T(0) && tc.vp > T(0);
    if (__ballot(!sub) == 0ull) {
      P::expand_t_split(f, tc, wA, bA, LG1, LG2);
      P::expand_t_split(f, tc, wB, bB, RG1, RG2);
    } else {
      T sg[NW], c1[NW], c2[NW], c3[NW], c4[NW];
      P::tspeeds(f, tc, sg);
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const T wm = wA * vmin(sg[j], T(0)), wp = wA * vmax(sg[j], T(0));
        c1[j] = wm * bA[j];
        c2[j] = wp * bA[j];
        if constexpr (CAPA) {
          c3[j] = (wB * vmin(sg[j], T(0))) * bB[j];
          c4[j] = (wB * vmax(sg[j], T(0))) * bB[j];
        } else {
          c3[j] = wm * bB[j];
          c4[j] = wp * bB[j];
        }
      }
      P::expand_t(f, tc, c1, LG1);
      P::expand_t(f, tc, c2, LG2);
      P::expand_t(f, tc, c3, RG1);
      P::expand_t(f, tc, c4, RG2);
    }
  }
}

// frame (h, normal, transverse, ...) of a y-interface -> (h, hu, hv, ...)
template <int NEQ, class T> __device__ __forceinline__ void yswap(T v[NEQ]) {
  const T t = v[1];
  v[1] = v[2];
  v[2] = t;
}

// One work item: the strip of columns [x0, x0+60) and rows [y0, y1) of one grid, advanced by dt
// from qin to qout.  Shared by the single-grid and the batched kernel.  The CFL speeds and the
// bad-state flag of the item are reduced over the wavefront and folded into *ctrl atomically.
// LIM > 0: every wave uses limiter LIM (compile-time); LIM == 0: per-wave ids from a.mthlim.
// PAD: qin is a ghost-padded state (num_eqn planes of my+4 rows of mx+4 cells, ghost cells filled
// by the caller, as classic step2 receives qbc); q's boundary conditions are then not applied here.
// SWEEP (split, padded): 0 both sweeps in one pass; 1 the x-sweep only, q* written into the
// interior of the padded qout (its ghost cells are left to the caller); 2 the y-sweep only, from a
// padded q* into the interior planes qout.
// CAPA: capacity form with kappa = a.capa (see finish_iface).  The accumulators hold each cell's
// update before its 1/kappa, applied when the cell is written.  For the split step the x-sweep of
// a ghost row takes kappa from the row its state comes from, so its q* equals the ghost fill of
// the x-swept interior (the y-sweep sees the aux ghost fill of kappa, as the reference's slice).
// The speed maxima of one work item over its x- and y-interfaces (NaN if a speed was NaN) and its
// bad-state flag, the same on every lane of the wavefront.
struct ItemMax {
  double sx, sy;
  bool bad;
};

// A wavefront-uniform double in scalar registers: the compiler then issues an atomic on it from
// one lane, instead of scanning the 64 lanes for a value it cannot prove uniform.
__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Folds maxima into a control block (every lane issues the atomics with the same operand, so the
// call may sit in wave-uniform control flow inside a work loop).
__device__ __forceinline__ void publish_max(claw2d_ctrl* c, const ItemMax& m) {
  atomic_max_nonneg(&c->smax_bits[0], m.sx);
  atomic_max_nonneg(&c->smax_bits[1], m.sy);
  if (m.bad) atomicOr(&c->status, (int)CLAW_STATUS_BAD_STATE);
}

template <class T, class P, int LIM, bool SPLIT = false, int TWC = -1, bool PAD = false, int SWEEP = 0,
          bool CAPA = false>
__device__ __forceinline__ ItemMax march_item(const StepArgs<T, P>& a, const T* __restrict__ qin,
                                           T* __restrict__ qout, const int mx, const int my,
                                           const int x0, const int y0, const int y1, const double dt,
                                           const double dx, const double dy) {
  static_assert(SWEEP == 0 || (SPLIT && PAD), "single sweeps run on the padded split path");
  static_assert(!(CAPA && PAD), "the capacity form runs on the index-mapped (unpadded) path");
  constexpr bool DPP = kDppShift<T, P>;
  // Zero-filled wavefront-shift edges (no copy pre-loads the destination): lanes 0 and 63 then
  // receive 0 from outside the wavefront.  Only halo data sees it: owned cells (lanes 2..61) take
  // the interfaces at lanes 2..62, whose inputs come from lanes 1..63 (measured on MI355X: FP64
  // shallow water -3.7% together with the one-way limiter exchange below, Euler FP32 -4%).
  constexpr bool DPPZ = true;
  constexpr int NEQ = P::NEQ, NW = P::NW, NC = P::NC, NL = P::NL;
  // 4 equations: add each y-interface's upper-cell shares at once (two more exchanges per row)
  // instead of carrying them to the next row, so the kernel fits 2 wavefronts per SIMD
  constexpr bool kImmY = NEQ == 4;
  using Cell = typename P::Cell;
  using Iface = typename P::Iface;
  const int lane = threadIdx.x & 63;
  const size_t plane = (size_t)mx * (size_t)my;
  const int c = x0 - 2 + lane;
  int cx;
  bool fx = false;
  if constexpr (PAD) cx = c < -2 ? -2 : (c > mx + 1 ? mx + 1 : c);  // lanes past the halo: unused data
  else map_index(c, mx, a.bcx_lo, a.bcx_hi, cx, fx);
  int cxa = cx;  // aux column (aux BCs may differ from the state's)
  if constexpr (P::NAUX > 0 || CAPA) {
    bool fa;
    map_index(c, mx, a.abx_lo, a.abx_hi, cxa, fa);
  }
  const bool own_col = lane >= 2 && lane < 2 + kOwn && c < mx;
  const bool xcfl_col = lane >= 1 && c >= 0 && c <= mx;   // x-interfaces 1..mx+1 (1-based)
  const bool ycfl_col = c >= -1 && c <= mx;               // y-sweeps on columns 0..mx+1
  const T dtdx = T(dt / dx), dtdy = T(dt / dy);

  // Loads row R of this lane's column.  The wall sign flips are returned, not applied, so the
  // consumer (one row later) and not the load waits for the data.
  const bool walls = !PAD && (fx || a.bcy_lo == CLAW_BC_WALL || a.bcy_hi == CLAW_BC_WALL);
  const ptrdiff_t ldq = PAD ? (ptrdiff_t)mx + 4 : (ptrdiff_t)mx;  // row stride of qin
  const size_t plane_in = PAD ? (size_t)(mx + 4) * (size_t)(my + 4) : plane;
  const T* const qin0 = PAD ? qin + 2 * ldq + 2 : qin;             // cell (0, 0)
  auto load_row = [&](int R, T q[NEQ], bool& fy, T& b, T& kap, T& kapx) {
    int ry = R;
    fy = false;
    if constexpr (!PAD) {
      if (R < 0 || R >= my) map_index(R, my, a.bcy_lo, a.bcy_hi, ry, fy);
    }
    ry = __builtin_amdgcn_readfirstlane(ry);
    const T* src = qin0 + (ptrdiff_t)ry * ldq + (ptrdiff_t)cx;
#pragma unroll
    for (int m = 0; m < NEQ; ++m) q[m] = src[(size_t)m * plane_in];
    if constexpr (P::NAUX > 0) {
      int ra = R;
      bool fa;
      if (R < 0 || R >= my) map_index(R, my, a.aby_lo, a.aby_hi, ra, fa);
      ra = __builtin_amdgcn_readfirstlane(ra);
      b = a.aux[(size_t)ra * (size_t)mx + (size_t)cxa];
    }
    if constexpr (CAPA) {
      int ra = R;
      bool fa;
      if (R < 0 || R >= my) map_index(R, my, a.aby_lo, a.aby_hi, ra, fa);
      ra = __builtin_amdgcn_readfirstlane(ra);
      kap = a.capa[(size_t)ra * (size_t)mx + (size_t)cxa];
      if constexpr (SPLIT) kapx = a.capa[(size_t)ry * (size_t)mx + (size_t)cxa];
      else kapx = kap;
    }
  };
  auto apply_flips = [&](T q[NEQ], bool fy) {
    if (!walls) return;
#pragma unroll
    for (int m = 0; m < NEQ; ++m) {
      const bool neg = (fx && m == a.rfx) != (fy && m == a.rfy);
      q[m] = neg ? -q[m] : q[m];
    }
  };

  // CFL speed of an interface: max |s| / kappa over its waves, a right-going wave against the
  // right cell's kappa and a left-going one against the left cell's (flux2.f90's cfl1d with
  // dtdx1d); every policy's extreme speeds are u - a and u + a
  auto cfl_speed = [&](const Iface& f, T rkl, T rkr) -> T {
    if constexpr (CAPA) return vmax((f.u + f.a) * rkr, (f.a - f.u) * rkl);
    else return P::smax(f);
  };
  // The split step's y-interfaces see q*, which a rejected attempt (too large a dt) can leave with
  // a negative pressure where the velocity is still finite: the sound speed a is then NaN and the
  // extreme speeds u -+ a with it.  The reference's maximum runs over every wave and skips NaN
  // speeds, so the middle waves (speed u) still count there; this adds them (they never exceed
  // the extremes when a is a number, so the result is otherwise the same bits).
  auto cfl_speed_q_star = [&](const Iface& f, T rkl, T rkr) -> T {
    const T mid = CAPA ? vmax(f.u * rkr, -f.u * rkl) : vabs(f.u);
    return vmax(cfl_speed(f, rkl, rkr), mid);
  };
  T rkprev = T(1), rkm2 = T(1);  // 1/kappa of rows R-1 and R-2 (capacity form)

  T smx = T(0), smy = T(0);
  // FP64: plain max plus a NaN flag folded in once at the end (the same result as a
  // NaN-propagating max at every update, in two instructions instead of five)
  constexpr bool kNanFlag = sizeof(T) == 8;
  bool nanx = false, nany = false;
  bool bad = false;
  Cell cprev = {};
  Iface fyprev;
  T yup[NW];  // y: alpha_p of interface R-2 times e_p(R-2).e_p(R-1), the upwind term of R-1
  T qm2[NEQ], qnext[NEQ];  // q of row R-2 (the row being finished), prefetched row R+1
  T acc0[NEQ], acc1[NEQ], acc2[NEQ], acc3[NEQ];
  T yX[NEQ], yG1[NEQ], yG2[NEQ];
#pragma unroll
  for (int m = 0; m < NEQ; ++m) {
    qm2[m] = T(0);
    acc0[m] = acc1[m] = acc2[m] = acc3[m] = T(0);
    yX[m] = yG1[m] = yG2[m] = T(0);
  }
#pragma unroll
  for (int p = 0; p < NW; ++p) yup[p] = T(0);

  // the x-sweep alone needs only the owned rows; every other form marches two halo rows each side
  const int rlo = SWEEP == 1 ? y0 : y0 - 2, rhi = SWEEP == 1 ? y1 - 1 : y1 + 1;
  bool fnext;
  T bnext = T(0), knext = T(1), kxnext = T(1);
  load_row(rlo, qnext, fnext, bnext, knext, kxnext);
  // One marched row.  ST (steady): a row y0+2..y1-1 of the full step, for which every row test
  // below holds, so that copy of the body is one straight block the scheduler can interleave
  // (shallow water: 2.4% faster in FP64 on MI355X; the Euler kernel would spill, so it keeps one
  // copy of the body).
  auto row = [&](const int R, auto steady_tag) {
    constexpr bool ST = decltype(steady_tag)::value;
    T qR[NEQ];
#pragma unroll
    for (int m = 0; m < NEQ; ++m) qR[m] = qnext[m];
    const T bR = bnext;
    // 1/kappa of row R as the y-sweeps see it (rkR) and as its x-sweep sees it (rkX)
    T rkR = T(1), rkX = T(1);
    if constexpr (CAPA) {
      rkR = frcp(knext);
      rkX = SPLIT ? frcp(kxnext) : rkR;
    }
    apply_flips(qR, fnext);
    if (ST || R < rhi) load_row(R + 1, qnext, fnext, bnext, knext, kxnext);
    const Cell cR = prep_cell<P>(qR, bR, a.par);

    // ---- request the left neighbour's cell data for the x-interfaces of row R now, so that the
    // y work below hides the ds_bpermute latency.  Shallow water sends it over ds_bpermute in
    // FP64 too: the y work hides the crossbar's latency there, and the 14 DPP moves it replaces
    // issue at the FP64 rate (4096^2 -2.5%).  The Euler and bathymetry kernels keep DPP here:
    // their y work does not hide it (Euler +0.6%, bathymetry +0.5%, split +1%).
    T pl[NC];
    if constexpr (SWEEP != 2) {
      constexpr bool kCellDpp = DPP && (P::NEQ == 4 || P::FWAVE);
      T pk[NC];
      P::pack(cR, pk);
#pragma unroll
      for (int k = 0; k < NC; ++k) pl[k] = from_left<kCellDpp, T, DPPZ>(pk[k]);
    }
    T rkLx = T(1);  // 1/kappa of the left cell of the row's x-interface
    if constexpr (CAPA && SWEEP != 2) rkLx = from_left<DPP, T, DPPZ>(rkX);

    // ---- y-sweep part of row R: the y-interface between rows R-1 and R (in-lane), then the
    // y-interface one row lower, whose limiter needed it.  cY is the cell data the y-sweeps see:
    // q^n for the unsplit step, the x-swept q* for the dimensionally split one.
    Iface fynew;
    auto y_part = [&](const Cell& cY) {
      // ---- y-interface R (between rows R-1 and R), in-lane
      if (ST || R >= y0 - 1) {
        P::template rpn2<1>(cprev, cY, a.par, fynew);
        // split step: the y-speeds come from q*, which a rejected attempt (too large a dt) can
        // leave with negative depths; as in the reference's max, NaN speeds are skipped (a NaN
        // in an accepted state is still caught by the stored-cell check)
        if ((SPLIT ? own_col : ycfl_col) && (ST || (R >= 0 && R <= my)))
        {
          const T sp = SPLIT ? cfl_speed_q_star(fynew, rkprev, rkR) : cfl_speed(fynew, rkprev, rkR);
          if constexpr (SPLIT) smy = vmax(smy, sp);
          else if constexpr (kNanFlag) { smy = vmax(smy, sp); nany |= sp != sp; }
          else smy = nmax(smy, sp);
        }
      }

      // ---- finish y-interface R-1 (between rows R-2 and R-1)
      if (ST || R >= y0) {
        T lp[NL], ln[NL], mR[NW];
        P::lim(fyprev, lp);
        P::lim(fynew, ln);
        P::mdot(lp, ln, mR);
        if (ST || R >= y0 + 1) {
          T LX[NEQ], LG1[NEQ], LG2[NEQ], RX[NEQ], RG1[NEQ], RG2[NEQ];
          finish_iface<T, P, LIM, TWC, CAPA>(fyprev, yup, fynew.al, mR, dtdy, dtdx, a, LX, LG1,
                                             LG2, RX, RG1, RG2, rkm2, rkprev);
          yswap<NEQ>(LX); yswap<NEQ>(LG1); yswap<NEQ>(LG2);
          yswap<NEQ>(RX); yswap<NEQ>(RG1); yswap<NEQ>(RG2);
          if constexpr (kImmY) {
            // both cells' shares go into their accumulators now (rows R-2 and R-1)
  #pragma unroll
            for (int m = 0; m < NEQ; ++m) {
              // (TWC 0: zero transverse shares, no exchange, as below)
              const T l2 = TWC == 0 ? T(0) : from_left<DPP, T, DPPZ>(LG2[m]);
              const T l1 = TWC == 0 ? T(0) : from_right<DPP, T, DPPZ>(LG1[m]);
              const T r2 = TWC == 0 ? T(0) : from_left<DPP, T, DPPZ>(RG2[m]);
              const T r1 = TWC == 0 ? T(0) : from_right<DPP, T, DPPZ>(RG1[m]);
              acc0[m] += ((LX[m] + LG1[m]) - LG2[m]) + (l2 - l1);
              acc1[m] += ((RX[m] + RG1[m]) - RG2[m]) + (r2 - r1);
            }
          }
          if (ST || R >= y0 + 2) {
            // cell (c, R-2) has all its y-contributions
            if constexpr (!kImmY) {
  #pragma unroll
              for (int m = 0; m < NEQ; ++m) {
                const T Y = yX[m] + LX[m];
                const T G1 = yG1[m] + LG1[m];
                const T G2 = yG2[m] + LG2[m];
                // without transverse propagation (TWC 0: the split step) the transverse shares
                // are zero and nothing crosses lanes; the compiler does not fold a DPP move of a
                // constant zero, so the exchange is left out here
                const T g2l = TWC == 0 ? T(0) : from_left<DPP, T, DPPZ>(G2);
                const T g1r = TWC == 0 ? T(0) : from_right<DPP, T, DPPZ>(G1);
                acc0[m] += ((Y + G1) - G2) + (g2l - g1r);
              }
            }
            if (own_col) {
              T qn[NEQ];
              const size_t off = (size_t)(R - 2) * (size_t)mx + (size_t)c;
  #pragma unroll
              for (int m = 0; m < NEQ; ++m) {
                qn[m] = CAPA ? qm2[m] + acc0[m] * rkm2 : qm2[m] + acc0[m];
                qout[(size_t)m * plane + off] = qn[m];
              }
              bad |= P::bad(qn, a.par);
            }
          }
          if constexpr (!kImmY) {
  #pragma unroll
            for (int m = 0; m < NEQ; ++m) { yX[m] = RX[m]; yG1[m] = RG1[m]; yG2[m] = RG2[m]; }
          }
        }
  #pragma unroll
        for (int p = 0; p < NW; ++p) {
          yup[p] = P::unit_m(p) ? fyprev.al[p] : fyprev.al[p] * mR[p];
        }
      }


    };

    if constexpr (!SPLIT) y_part(cR);

    // ---- x-interfaces of row R (the left cells were requested before the y work).  Split step:
    // every marched row (the y-sweeps need q* of the halo rows too); ghost rows are x-swept from
    // their mapped source rows, which equals refilling the ghosts of q* for extrap, periodic and
    // wall conditions.
    Cell cY = cR;
    if (SWEEP != 2 && (SPLIT || ST || (R >= y0 - 1 && R <= y1))) {
      const Cell cL = P::unpack(pl);
      Iface fxi;
      P::template rpn2<0>(cL, cR, a.par, fxi);
      if (xcfl_col && (ST || (SPLIT ? (R >= 0 && R < my) : (R >= -1 && R <= my))))
      {
        const T sp = cfl_speed(fxi, rkLx, rkX);
        if constexpr (kNanFlag) { smx = vmax(smx, sp); nanx |= sp != sp; }
        else smx = nmax(smx, sp);
      }
      // limiter partners.  The right neighbour's wave data come over one exchange; this lane
      // forms the eigenvector products m_p of its pair (this, right) and hands the right
      // neighbour the product it needs when its wave p is right-going, alpha_p(this) * m_p, which
      // is its left partner's upwind term (NW values instead of the left partner's NL).
      T lself[NL], lnb[NL], mR[NW], alR[NW], upL[NW];
      P::lim(fxi, lself);
#pragma unroll
      for (int k = 0; k < NL; ++k) lnb[k] = from_right<DPP, T, DPPZ>(lself[k]);
      P::mdot(lself, lnb, mR);
#pragma unroll
      for (int p = 0; p < NW; ++p) {
        const T up = P::unit_m(p) ? lself[p] : lself[p] * mR[p];
        upL[p] = from_left<DPP, T, DPPZ>(up);
        alR[p] = lnb[p];
      }
      T LX[NEQ], LG1[NEQ], LG2[NEQ], RX[NEQ], RG1[NEQ], RG2[NEQ];
      finish_iface<T, P, LIM, TWC, CAPA>(fxi, upL, alR, mR, dtdx, dtdy, a, LX, LG1, LG2, RX, RG1, RG2,
                                         rkLx, rkX);
      if constexpr (SPLIT) {
        // q* of row R: the x-sweep only (no transverse terms); valid on the owned columns
        T qs[NEQ];
#pragma unroll
        for (int m = 0; m < NEQ; ++m) {
          const T dq = RX[m] + from_right<DPP, T, DPPZ>(LX[m]);
          qs[m] = CAPA ? qR[m] + dq * rkX : qR[m] + dq;
        }
        if constexpr (SWEEP == 1) {
          if (own_col) {
            const size_t off = (size_t)(R + 2) * (size_t)(mx + 4) + (size_t)(c + 2);
#pragma unroll
            for (int m = 0; m < NEQ; ++m) qout[(size_t)m * plane_in + off] = qs[m];
          }
        } else {
          cY = prep_cell<P>(qs, bR, a.par);
        }
      } else {
#pragma unroll
        for (int m = 0; m < NEQ; ++m) {
          const T X = RX[m] + from_right<DPP, T, DPPZ>(LX[m]);
          const T G1 = RG1[m] + from_right<DPP, T, DPPZ>(LG1[m]);
          const T G2 = RG2[m] + from_right<DPP, T, DPPZ>(LG2[m]);
          acc1[m] -= G1;
          acc2[m] += (X + G1) - G2;
          acc3[m] = G2;  // row R+1 has no other contribution yet
        }
      }
    }
    if constexpr (SPLIT && SWEEP != 1) y_part(cY);

    // ---- roll the column window (the x-sweep alone keeps no column state)
    if constexpr (SWEEP != 1) {
#pragma unroll
      for (int m = 0; m < NEQ; ++m) {
        qm2[m] = cprev.q[m];  // row R-1 (cprev still holds it)
        acc0[m] = acc1[m];
        acc1[m] = acc2[m];
        acc2[m] = acc3[m];
        acc3[m] = T(0);
      }
      cprev = cY;
      fyprev = fynew;
      rkm2 = rkprev;
      rkprev = rkR;
    }
  };
  if constexpr (SWEEP == 0 && NEQ == 3) {
    int R = rlo;
    for (; R < y0 + 2 && R <= rhi; ++R) row(R, std::false_type{});
    // Issue priority from progress (segments of at least kPrioRows steady rows): 3 over the first
    // quarter of the steady rows, then 2, 1 and 0.  The wavefronts that share a SIMD start
    // together with equal work, and the arbiter otherwise favours the older one: it finishes
    // early and the others run their last stretch with fewer wavefronts per SIMD.  With the
    // priority falling as a wavefront advances, the one behind is issued first, and they finish
    // closer together.  Issue order only: no result bit changes.  (For FP32 it measured slower
    // while every wavefront still sent its own device-scope atomics at its end: finishing together
    // queued them; with one set per workgroup it is 7% faster at 4096^2, 10% at 8192^2.)
    const bool prio = y1 - R >= kPrioRows;
    const int pq = max((y1 - R) >> 2, 1);
    int Rq = prio ? R + pq : INT_MAX, pl = 3;  // next row at which the priority steps down
    if (prio) __builtin_amdgcn_s_setprio(3);
    auto prio_at = [&](int r) {
      if (r >= Rq) {
        Rq += pq;
        --pl;
        if (pl == 2) __builtin_amdgcn_s_setprio(2);
        else if (pl == 1) __builtin_amdgcn_s_setprio(1);
        else __builtin_amdgcn_s_setprio(0);
      }
    };
    if constexpr (SPLIT && sizeof(T) == 8) {
      // two steady rows per iteration: the rolling window's registers are renamed, not copied
      // (FP64 split: -2% on MI355X, at 158 of its 168 VGPRs; the unsplit kernels would spill)
      for (; R + 1 <= y1 - 1; R += 2) {
        prio_at(R);
        row(R, std::true_type{});
        row(R + 1, std::true_type{});
      }
    }
    for (; R <= y1 - 1; ++R) {
      prio_at(R);
      row(R, std::true_type{});
    }
    if (prio) __builtin_amdgcn_s_setprio(0);
    for (; R <= rhi; ++R) row(R, std::false_type{});
  } else {
    // Issue priority from progress, as above, over every marched row of an Euler segment (2048^2
    // on MI355X: FP64 -5%, FP32 -10%)
    const bool prio = NEQ == 4 && rhi - rlo >= kPrioRowsEuler;
    const int pq = max((rhi + 1 - rlo) >> 2, 1);
    int Rq = prio ? rlo + pq : INT_MAX, pl = 3;
    if (prio) __builtin_amdgcn_s_setprio(3);
    for (int R = rlo; R <= rhi; ++R) {
      if constexpr (NEQ == 4) {
        if (R >= Rq) {
          Rq += pq;
          --pl;
          if (pl == 2) __builtin_amdgcn_s_setprio(2);
          else if (pl == 1) __builtin_amdgcn_s_setprio(1);
          else __builtin_amdgcn_s_setprio(0);
        }
      }
      row(R, std::false_type{});
    }
    if (prio) __builtin_amdgcn_s_setprio(0);
  }

  // Wave-uniform epilogue: no lane-divergent branch (a lane-0-only branch here lets the compiler
  // peel lanes 1-63 off into their own loop when this function runs inside the batched kernel's
  // work loop).  The caller folds the maxima into the control block.
  if constexpr (kNanFlag) {
    if (nanx) smx = T(NAN);
    if (nany) smy = T(NAN);
  }
  const T wx = wave_max(smx), wy = wave_max(smy);
  return ItemMax{uniform((double)wx), uniform((double)wy), __ballot(bad) != 0ull};
}

// The speed maxima and the status bits were written by atomics of the step's wavefronts; read
// them with device-scope atomic loads (a plain load may hit a stale line in this CU's L1 when
// the controller runs inside the step kernel).
__device__ __forceinline__ double smax_read(claw2d_ctrl* c, int k) {
  return __longlong_as_double(
      (long long)__hip_atomic_load(&c->smax_bits[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ int status_read(claw2d_ctrl* c) {
  return __hip_atomic_load(&c->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// CFL of a fixed-dt step (classic step2 returns cfl = max over both sweeps of dt/dx*|s|).
__device__ __forceinline__ void finalize_one(claw2d_ctrl* c, double dt, double dx, double dy) {
  const double sx = smax_read(c, 0), sy = smax_read(c, 1);
  c->smax_last[0] = sx;
  c->smax_last[1] = sy;
  c->smax_bits[0] = 0ull;
  c->smax_bits[1] = 0ull;
  c->cfl = fmax(fmax(0.0, (dt / dx) * sx), (dt / dy) * sy);
}

__device__ __forceinline__ void control_one(claw2d_ctrl* c, double dx, double dy, double cfl_max,
                                            double cfl_desired, double dt_max) {
  if (c->done) return;
  const double sx = smax_read(c, 0), sy = smax_read(c, 1);
  c->smax_last[0] = sx;
  c->smax_last[1] = sy;
  c->smax_bits[0] = 0ull;
  c->smax_bits[1] = 0ull;
  const double dt = c->dt;
  const double cfl = fmax(fmax(0.0, (dt / dx) * sx), (dt / dy) * sy);
  const bool dtvar = (c->flags & CLAW_FLAG_DT_VARIABLE) != 0;
  const bool to_tend = (c->flags & CLAW_FLAG_RUN_TO_TEND) != 0;
  c->attempts += 1;
  c->cfl = cfl;
  // a rejected attempt is discarded whatever it produced; a non-positive depth/density or a NaN
  // speed in an attempt that would be accepted ends the run with CLAW_STATUS_BAD_STATE
  const int status = status_read(c);
  const bool bad = (status & CLAW_STATUS_BAD_STATE) || cfl != cfl;
  c->status = status & ~CLAW_STATUS_BAD_STATE;
  if (bad && !(cfl > cfl_max)) {
    c->status |= CLAW_STATUS_BAD_STATE;
    c->done = 1;
    return;
  }
  if (!(cfl > cfl_max)) {
    c->cflmax = fmax(cfl, c->cflmax);
    c->numsteps += 1;
    if (dtvar) c->t += dt;
    else c->t = c->tstart + (double)c->numsteps * dt;
    c->cur ^= 1;
  } else {
    c->rejected += 1;
    if (!dtvar) {
      c->cflmax = fmax(cfl, c->cflmax);
      c->status |= CLAW_STATUS_CFL_FIXED_DT;
      c->done = 1;
      return;
    }
  }
  if (to_tend ? (c->t >= c->tend) : (c->numsteps >= c->target_steps)) {
    c->done = 1;
    return;
  }
  if (c->attempts >= c->max_steps) {
    // pyclaw's evolve_to_time raises only with a variable dt and every attempt accepted
    if (dtvar && c->numsteps == c->max_steps) c->status |= CLAW_STATUS_MAX_STEPS;
    c->done = 1;
    return;
  }
  double ndt = dt;
  if (dtvar) {
    if (cfl > 0.0) {
      ndt = fmin(dt_max, dt * cfl_desired / cfl);
      c->dtmin = fmin(ndt, c->dtmin);
      c->dtmax = fmax(ndt, c->dtmax);
    } else {
      ndt = dt_max;
    }
  }
  if (to_tend) {
    if (c->t + ndt > c->tend && c->tstart < c->tend) ndt = c->tend - c->t;
    if (c->tend - c->t - ndt < 1.0e-14 * c->t) ndt = c->tend - c->t;
  }
  c->dt = ndt;
}

// Minimum resident wavefronts per SIMD the register allocator must allow (0 = its own choice).
#ifndef CLAW_MIN_WAVES_F64
#define CLAW_MIN_WAVES_F64 2
#endif
#ifndef CLAW_MIN_WAVES_F32
#define CLAW_MIN_WAVES_F32 0
#endif
template <class T> struct MinWaves {
  static constexpr int value = sizeof(T) == 8 ? CLAW_MIN_WAVES_F64 : CLAW_MIN_WAVES_F32;
};

template <class T, class P, int LIM, bool SPLIT, int TWC, bool PAD = false, int SWEEP = 0, bool CAPA = false>
__global__ void __launch_bounds__(kLanes* kWPB) __attribute__((amdgpu_waves_per_eu(MinWaves<T>::value)))
step2_fused(StepArgs<T, P> a) {
  // The wavefronts of a workgroup combine their speed maxima and bad-state flags in LDS; the last
  // of them to finish folds the workgroup's into the control block and counts the workgroup in
  // the step's arrival counter.  Device-scope atomics on the control block from every wavefront
  // had queued at the end of the step: removing them made FP32 shallow water at 4096^2 17%
  // faster on MI355X, and a workgroup of 4 wavefronts sends a quarter of them.
  __shared__ unsigned long long wg_smax[2];
  __shared__ int wg_bad, wg_arrived;
  if (threadIdx.x == 0) {
    wg_smax[0] = wg_smax[1] = 0ull;
    wg_bad = 0;
    wg_arrived = 0;
  }
  __syncthreads();  // before any wavefront of the workgroup leaves
  double dt = a.dt_fixed;
  const T* qin = a.q0;
  T* qout = a.q1;
  if (a.use_ctrl) {
    if (a.ctrl->done) return;
    dt = a.ctrl->dt;
    if (a.ctrl->cur) { qin = a.q1; qout = a.q0; }
  }
  // wave-uniform quantities live in SGPRs (readfirstlane), so the row loop branches on SCC
  const int blk = a.xcd ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(blk * a.wpb + (threadIdx.x >> 6));
  if (wave >= a.nstrips * a.nsegs) return;
  const int strip = wave % a.nstrips, seg = wave / a.nstrips;
  const int x0 = strip * kOwn, y0 = seg * a.seg_h;
  const int y1 = min(y0 + a.seg_h, a.my);
  const ItemMax m = march_item<T, P, LIM, SPLIT, TWC, PAD, SWEEP, CAPA>(a, qin, qout, a.mx, a.my, x0, y0, y1,
                                                                       dt, a.dx, a.dy);
  const int lane = threadIdx.x & 63;
  atomicMax(&wg_smax[0], (unsigned long long)__double_as_longlong(m.sx));
  atomicMax(&wg_smax[1], (unsigned long long)__double_as_longlong(m.sy));
  if (m.bad) atomicOr(&wg_bad, 1);
  // the workgroup's wavefronts that march (the last workgroup may hold fewer than wpb)
  const int nw = min(a.wpb, a.total_waves - blk * a.wpb);
  const int k = __builtin_amdgcn_readfirstlane(
      __hip_atomic_fetch_add(&wg_arrived, lane == 0 ? 1 : 0, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP));
  if (k != nw - 1) return;
  const ItemMax wg{
      uniform(__longlong_as_double(
          (long long)__hip_atomic_load(&wg_smax[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))),
      uniform(__longlong_as_double(
          (long long)__hip_atomic_load(&wg_smax[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))),
      __builtin_amdgcn_readfirstlane(__hip_atomic_load(&wg_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) != 0};
  publish_max(a.ctrl, wg);
  if (a.fuse) {
    // The last workgroup to finish closes the step: the speed maxima are then complete.  The wait
    // for this wavefront's outstanding memory operations (its max/or atomics are acknowledged once
    // performed device-wide) comes before its arrival; the controller reads the maxima with
    // device-scope atomic loads.  No cache writeback is needed: the controller reads nothing else
    // that this launch wrote.
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const int ticket = __builtin_amdgcn_readfirstlane(atomicAdd(&a.ctrl->arrive, lane == 0 ? 1 : 0));
    if (ticket == (int)gridDim.x - 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (lane == 0) {
        if (a.fuse == 1) finalize_one(a.ctrl, dt, a.dx, a.dy);
        else control_one(a.ctrl, a.dx, a.dy, a.cfl_max, a.cfl_desired, a.dt_max);
        a.ctrl->arrive = 0;
      }
    }
  }
}

// Batched grids: every grid's planes live in one ping-pong pair at its own offset, each grid has
// its own control block (its own CFL-chosen dt), and a persistent set of wavefronts takes
// (grid, strip, segment) work items from an atomic counter.  The counter is reset by the batch
// control kernel, so a captured sequence of launches contains no memset.
struct BatchDev {
  const claw2d_batch_grid* grids;
  claw2d_ctrl* ctrls;
  int* counter;
  int ngrids;
  int total_items;
};

template <class T, class P, int LIM, int TWC>
__global__ void __launch_bounds__(kLanes* kWPB) __attribute__((amdgpu_waves_per_eu(MinWaves<T>::value)))
batch_step(StepArgs<T, P> a, BatchDev b) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    // every lane takes part (lane 0 adds 1, the others 0) and lane 0's ticket is broadcast, so
    // no lane-divergent region sits inside this loop
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(b.counter, lane == 0 ? 1 : 0));
    if (item >= b.total_items) break;
    int lo = 0, hi = b.ngrids - 1;  // last grid whose item0 <= item
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (b.grids[mid].item0 <= item) lo = mid;
      else hi = mid - 1;
    }
    const int gi = __builtin_amdgcn_readfirstlane(lo);
    const claw2d_batch_grid& G = b.grids[gi];
    claw2d_ctrl* c = b.ctrls + gi;
    if (c->done) continue;
    const double dt = c->dt;
    const T* qin = (c->cur ? a.q1 : a.q0) + G.offset;
    T* qout = (c->cur ? a.q0 : a.q1) + G.offset;
    const int local = item - G.item0;
    const int strip = local % G.nstrips, seg = local / G.nstrips;
    const int x0 = strip * kOwn, y0 = seg * G.seg_h;
    const int y1 = min(y0 + G.seg_h, G.my);
    publish_max(c, march_item<T, P, LIM, false, TWC>(a, qin, qout, G.mx, G.my, x0, y0, y1, dt, G.dx, G.dy));
  }
}


// One evolve_to_time iteration after the step (upstream pyclaw solver.py): accept/reject,
// advance t, choose the next dt, clip it to tend.

#ifndef CLAW_KERNEL_UNIT
// One thread per grid; thread 0 also rearms the work counter of the batched step.
__global__ void batch_control(const claw2d_batch_grid* grids, claw2d_ctrl* ctrls, int* counter,
                              int ngrids, double cfl_max, double cfl_desired, double dt_max) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) *counter = 0;
  if (g >= ngrids) return;
  control_one(ctrls + g, grids[g].dx, grids[g].dy, cfl_max, cfl_desired, dt_max);
}
#endif

// ------------------------------------------------------------------------------------------
// Compensated FP64 sum of one component (Neumaier), fixed reduction order.
// ------------------------------------------------------------------------------------------
constexpr int kSumBlocks = 1024, kSumThreads = 256;

__device__ __forceinline__ void two_sum_acc(double& s, double& comp, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) comp += (s - t) + x;
  else comp += (x - t) + s;
  s = t;
}

template <class T>
__global__ void __launch_bounds__(kSumThreads) mass_partial(const T* q, size_t n, double* part) {
  __shared__ double ss[kSumThreads], sc[kSumThreads];
  double s = 0.0, comp = 0.0;
  const size_t stride = (size_t)kSumBlocks * kSumThreads;
  for (size_t i = (size_t)blockIdx.x * kSumThreads + threadIdx.x; i < n; i += stride)
    two_sum_acc(s, comp, (double)q[i]);
  ss[threadIdx.x] = s;
  sc[threadIdx.x] = comp;
  __syncthreads();
  for (int w = kSumThreads / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      double a = ss[threadIdx.x], ca = sc[threadIdx.x] + sc[threadIdx.x + w];
      two_sum_acc(a, ca, ss[threadIdx.x + w]);
      ss[threadIdx.x] = a;
      sc[threadIdx.x] = ca;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ss[0];
    part[2 * blockIdx.x + 1] = sc[0];
  }
}

#ifndef CLAW_KERNEL_UNIT
// offset: a constant added once to the sum (the reference depth times the cell count of a state
// that stores h - depth_ref), as one more compensated term
__global__ void __launch_bounds__(64) mass_final(const double* part, double* out, double offset) {
  if (threadIdx.x != 0) return;
  double s = 0.0, comp = 0.0;
  for (int b = 0; b < kSumBlocks; ++b) {
    comp += part[2 * b + 1];
    two_sum_acc(s, comp, part[2 * b]);
  }
  if (offset != 0.0) two_sum_acc(s, comp, offset);
  out[0] = s + comp;
}
#endif

// ------------------------------------------------------------------------------------------
// Step kernels instantiated in translation units of their own (claw2d_swe32.hip,
// claw2d_euler64.hip, claw2d_euler32.hip), which are built with another LLVM machine scheduler
// (clawpack_amd/build.py: FP32 shallow water -2.7% and FP32 Euler -1.1% with max-ilp, FP64 Euler
// -2.6% and no spill with iterative-minreg on MI355X; the same schedulers made other kernels up to
// 6% slower).  The host code below launches them through their kernel handles, which those units
// register with their own device code.  Scheduling does not change the arithmetic: contraction is
// decided per source expression.
extern template __global__ void step2_fused<float, SweP<float>, 4, false, 2, false, 0, false>(
    StepArgs<float, SweP<float>>);
extern template __global__ void step2_fused<double, EulerP<double>, 4, false, 2, false, 0, false>(
    StepArgs<double, EulerP<double>>);
extern template __global__ void step2_fused<float, EulerP<float>, 4, false, 2, false, 0, false>(
    StepArgs<float, EulerP<float>>);

#ifndef CLAW_KERNEL_UNIT
// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
// Resident wavefronts of a kernel on the current device (workgroups per CU x CUs x waves per
// workgroup), from the occupancy its register use allows.
// Wavefronts the work is cut for: the kernel's occupancy, at most 4 per SIMD.  A kernel that fits
// more (FP32 split: 6) gets shorter segments from the full count, and at 4096^2 their extra halo
// rows cost more than the occupancy returns (FP32 split: 0.241 ms at 6 per SIMD, 0.220 at 4).
constexpr long kMaxWavesPerCU = 16;
static long resident_waves(const void* kern, int wpb = kWPB) {
  int dev = 0, cus = 0, bpc = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kern, kLanes * wpb, 0) != hipSuccess) return 0;
  return (long)cus * std::min<long>((long)bpc * wpb, kMaxWavesPerCU);
}

// wavefronts per workgroup of a single-grid launch (desc->wg_waves; 0 = kWPB)
static int launch_wpb(const claw2d_desc* d) { return d->wg_waves > 0 ? d->wg_waves : kWPB; }
// occupancy cache index of a wavefronts-per-workgroup count 1, 2 or 4
static int wpb_index(int wpb) { return wpb == 1 ? 0 : (wpb == 2 ? 1 : 2); }

// SIMDs of the current device (4 per CU), 0 if the query fails.
static long device_simds() {
  static thread_local long simds = -1;
  if (simds < 0) {
    int dev = 0, cus = 0;
    simds = (hipGetDevice(&dev) == hipSuccess &&
             hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
                ? 4L * cus : 0;
  }
  return simds;
}

// Rows per marching segment.  A wavefront's run time is proportional to its rows, so the grid
// is cut into as many segments as fit in one resident round (every wavefront starts at once
// and ends at once), of at least 16 rows: each segment re-solves 4 halo rows.  A small grid whose
// 16-row segments would leave more than half the SIMDs idle runs latency-bound, one wavefront per
// SIMD, where time follows the rows per wavefront; it is cut for half the SIMDs instead, down to
// 4 rows (FP64 shallow water, 256^2: 43.3 -> 27.4 us per step; 512^2: 45.2 -> 40.1 us).  An
// explicit desc->seg_h overrides this.
static int choose_seg_h(const claw2d_desc* d, long resident) {
  if (d->seg_h > 0) return d->seg_h;
  const long nstrips = (d->mx + kOwn - 1) / kOwn;
  long nsegs = resident > 0 ? resident / nstrips : 32;
  long min_rows = 16;
  const long simds = device_simds(), half = simds / 2;
  if (nstrips * (((long)d->my + min_rows - 1) / min_rows) < half) {
    nsegs = half / nstrips;
    min_rows = 4;
  }
  nsegs = std::max(1L, std::min(nsegs, ((long)d->my + min_rows - 1) / min_rows));
  // More than 2 wavefronts per SIMD in segments under 32 rows re-solve more halo rows than the
  // extra occupancy returns: cut for 2 per SIMD instead.  At 2048^2 (18-row segments at 4 per
  // SIMD -> 36 rows at 2): shallow water FP32 135.9 -> 114.3 us, split 131.3 -> 107.1 us;
  // bathymetry FP32 138.1 -> 108.9 us, split 132.4 -> 101.7 us; FP64 split (24 rows at 3 per SIMD)
  // 128.4 -> 122.5 us.  The Euler kernels keep the occupancy sizing: FP32 135.2 -> 139.2 us there.
  long seg0 = ((long)d->my + nsegs - 1) / nsegs;
  long waves = nstrips * (((long)d->my + seg0 - 1) / seg0);
  if (simds > 0 && d->solver != CLAW_SOLVER_EULER_4WAVE && waves > 2 * simds && seg0 < 32) {
    nsegs = std::max(1L, 2 * simds / nstrips);
    seg0 = ((long)d->my + nsegs - 1) / nsegs;
    waves = nstrips * (((long)d->my + seg0 - 1) / seg0);
  }
  // A launch just over a whole number m of wavefronts per SIMD (by up to a quarter of the SIMDs)
  // leaves those few SIMDs a round more than the rest, and they set the time; it is cut to m
  // per SIMD instead (FP64 shallow water 1024^2: 1152 wavefronts 81.2 us, 936 wavefronts 66.9 us).
  if (simds > 0 && waves > simds) {
    const long m = waves / simds;
    if (waves - m * simds <= simds / 4 && m * simds / nstrips >= 1) nsegs = m * simds / nstrips;
  }
  return (int)((d->my + nsegs - 1) / nsegs);
}

static int check_desc(const claw2d_desc* d) {
  if (!d) return fail(-1, "null descriptor");
  if (d->solver != CLAW_SOLVER_SHALLOW_ROE_EFIX && d->solver != CLAW_SOLVER_EULER_4WAVE &&
      d->solver != CLAW_SOLVER_SHALLOW_BATHY_FWAVE)
    return fail(-2, "unsupported solver id");
  const int neq = d->solver == CLAW_SOLVER_EULER_4WAVE ? 4 : 3;
  if (d->num_eqn != neq || d->num_waves != neq) return fail(-3, "num_eqn/num_waves do not match the solver");
  if (d->dtype != CLAW_FP64 && d->dtype != CLAW_FP32) return fail(-4, "dtype must be FP64 or FP32");
  if (d->mx < 2 || d->my < 2) return fail(-5, "mx and my must be >= 2");
  if (d->num_ghost != 2) return fail(-6, "num_ghost must be 2");
  if (d->order != 1 && d->order != 2) return fail(-7, "order must be 1 or 2");
  if (d->transverse_waves < 0 || d->transverse_waves > 2) return fail(-8, "transverse_waves must be 0, 1 or 2");
  for (int p = 0; p < neq; ++p)
    if (d->mthlim[p] < 0 || d->mthlim[p] > 5) return fail(-9, "limiter id must be 0..5");
  for (int k = 0; k < 2; ++k) {
    const int lo = d->bc_lower[k], hi = d->bc_upper[k];
    if (lo < 1 || lo > 3 || hi < 1 || hi > 3) return fail(-10, "bc must be extrap(1), periodic(2) or wall(3)");
    if ((lo == CLAW_BC_PERIODIC) != (hi == CLAW_BC_PERIODIC)) return fail(-11, "periodic bc must be set on both sides");
    if (d->reflect_index[k] < 0 || d->reflect_index[k] >= neq) return fail(-12, "reflect_index out of range");
  }
  if (!(d->dx > 0) || !(d->dy > 0)) return fail(-13, "dx, dy must be positive");
  if (d->seg_h < 0) return fail(-14, "seg_h must be >= 0");
  if ((long long)d->mx * d->my > (1LL << 31)) return fail(-15, "grid too large for 32-bit row index math");
  if (d->mcapa < 0 || d->mcapa > d->num_aux)
    return fail(-29, "mcapa must be 0 (no capacity function) or a 1-based aux plane <= num_aux");
  if (d->solver == CLAW_SOLVER_SHALLOW_BATHY_FWAVE) {
    if (d->num_aux < 1) return fail(-16, "shallow_bathymetry_fwave needs num_aux >= 1 (aux plane 0 = bathymetry)");
    if (d->fwave != 1) return fail(-17, "shallow_bathymetry_fwave returns f-waves: fwave must be 1");
    if (!(d->dry_tolerance >= 0)) return fail(-18, "dry_tolerance must be >= 0");
  } else {
    if (d->num_aux != 0 && d->mcapa == 0)
      return fail(-16, "this Riemann solver reads no aux arrays: num_aux > 0 needs a capacity function (mcapa)");
    if (d->fwave != 0) return fail(-17, "fwave = 1 needs an f-wave Riemann solver (shallow_bathymetry_fwave)");
  }
  if (d->wg_waves != 0 && d->wg_waves != 1 && d->wg_waves != 2 && d->wg_waves != 4)
    return fail(-36, "wg_waves must be 0 (default), 1, 2 or 4");
  if (d->launch_order != 0 && d->launch_order != 1) return fail(-37, "launch_order must be 0 or 1");
  if (!std::isfinite(d->depth_ref)) return fail(-35, "depth_ref must be finite");
  if (d->depth_ref != 0.0 && (d->solver != CLAW_SOLVER_SHALLOW_ROE_EFIX || d->dtype != CLAW_FP32))
    return fail(-35, "depth_ref (stored depth h - depth_ref) is for FP32 shallow_roe_with_efix states only");
  if (d->num_aux > 0) {
    for (int k = 0; k < 2; ++k) {
      const int lo = d->aux_bc_lower[k], hi = d->aux_bc_upper[k];
      if (lo < 1 || lo > 3 || hi < 1 || hi > 3)
        return fail(-19, "aux bc must be extrap(1), periodic(2) or wall(3)");
      if ((lo == CLAW_BC_PERIODIC) != (hi == CLAW_BC_PERIODIC))
        return fail(-11, "periodic aux bc must be set on both sides");
    }
  }
  return 0;
}

template <class T, class P>
static void fill_args(StepArgs<T, P>& a, const claw2d_desc* d, long resident) {
  a.wpb = launch_wpb(d);
  a.xcd = d->launch_order == 0 ? 1 : 0;
  a.dx = d->dx;
  a.dy = d->dy;
  a.mx = d->mx;
  a.my = d->my;
  a.seg_h = choose_seg_h(d, resident);
  a.nstrips = (d->mx + kOwn - 1) / kOwn;
  a.nsegs = (d->my + a.seg_h - 1) / a.seg_h;
  a.bcx_lo = d->bc_lower[0];
  a.bcx_hi = d->bc_upper[0];
  a.bcy_lo = d->bc_lower[1];
  a.bcy_hi = d->bc_upper[1];
  a.rfx = d->reflect_index[0];
  a.rfy = d->reflect_index[1];
  a.order = d->order;
  a.tw = d->transverse_waves;
  for (int p = 0; p < 4; ++p) a.mthlim[p] = d->mthlim[p];
  a.aux = nullptr;
  a.capa = nullptr;
  a.abx_lo = d->aux_bc_lower[0];
  a.abx_hi = d->aux_bc_upper[0];
  a.aby_lo = d->aux_bc_lower[1];
  a.aby_hi = d->aux_bc_upper[1];
}

template <class T> static typename SweP<T>::Par swe_par(const claw2d_desc* d) {
  typename SweP<T>::Par p;
  p.g = T(d->grav);
  p.sqg = T(std::sqrt(d->grav));
  p.h0 = T(d->depth_ref);
  return p;
}
template <class T> static typename SweBathyP<T>::Par bathy_par(const claw2d_desc* d) {
  typename SweBathyP<T>::Par p;
  p.g = T(d->grav);
  p.hg = T(0.5 * d->grav);
  p.tol = T(d->dry_tolerance);
  return p;
}
template <class T> static typename EulerP<T>::Par euler_par(const claw2d_desc* d) {
  typename EulerP<T>::Par p;
  p.gamma = T(d->gamma);
  p.gamma1 = T(d->gamma - 1.0);
  return p;
}

// Kernel variant for a method: a uniform MC or minmod limiter at order 2 with the transverse
// mode fixed at compile time (split: none; unsplit: 1 or 2), else the general kernel (per-wave
// limiter ids, order and transverse_waves read at run time).  slot indexes the occupancy cache.
template <class T, class P, bool SPLIT>
static const void* step_kernel(const claw2d_desc* d, int* slot) {
  bool same = true;
  for (int p = 1; p < P::NW; ++p) same = same && d->mthlim[p] == d->mthlim[0];
  const int lim = same ? d->mthlim[0] : -1;
  if ((lim == 4 || lim == 1) && d->order == 2) {
    const bool mc = lim == 4;
    if constexpr (SPLIT) {
      *slot = mc ? 0 : 1;
      return mc ? (const void*)step2_fused<T, P, 4, true, 0> : (const void*)step2_fused<T, P, 1, true, 0>;
    } else {
      if (d->transverse_waves == 2) {
        *slot = mc ? 0 : 1;
        return mc ? (const void*)step2_fused<T, P, 4, false, 2> : (const void*)step2_fused<T, P, 1, false, 2>;
      }
      if (d->transverse_waves == 1) {
        *slot = mc ? 2 : 3;
        return mc ? (const void*)step2_fused<T, P, 4, false, 1> : (const void*)step2_fused<T, P, 1, false, 1>;
      }
    }
  }
  *slot = 4;
  return (const void*)step2_fused<T, P, 0, SPLIT, -1>;
}

template <class T, class P, bool SPLIT>
static int launch_step(const claw2d_desc* d, const typename P::Par& par, void* q0, void* q1,
                       const void* aux, claw2d_ctrl* ctrl, int use_ctrl, double dt, hipStream_t s,
                       bool pad = false) {
  int slot = 5;
  const void* kern = (const void*)step2_fused<T, P, 0, false, -1, true>;
  if (d->mcapa > 0) {
    slot = 6;  // the capacity form runs the general kernel
    kern = (const void*)step2_fused<T, P, 0, SPLIT, -1, false, 0, true>;
  } else if (!pad || SPLIT) {
    kern = step_kernel<T, P, SPLIT>(d, &slot);
  }
  const int wpb = launch_wpb(d);
  static thread_local long cached[7][3] = {{-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1},
                                           {-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1}};
  long& res = cached[slot][wpb_index(wpb)];
  if (res < 0) res = resident_waves(kern, wpb);
  StepArgs<T, P> a;
  fill_args(a, d, res);
  a.q0 = (T*)q0;
  a.q1 = (T*)q1;
  a.ctrl = ctrl;
  a.use_ctrl = use_ctrl;
  a.dt_fixed = dt;
  a.par = par;
  a.fuse = use_ctrl ? 2 : 1;
  a.cfl_max = d->cfl_max;
  a.cfl_desired = d->cfl_desired;
  a.dt_max = d->dt_max;
  if (SPLIT) a.tw = 0;  // dimensional splitting propagates no transverse waves (method(3) < 0)
  a.aux = (const T*)aux;
  if (d->mcapa > 0) a.capa = (const T*)aux + (size_t)(d->mcapa - 1) * (size_t)d->mx * (size_t)d->my;
  const int waves = a.nstrips * a.nsegs;
  a.total_waves = waves;
  const int blocks = (waves + wpb - 1) / wpb;
  void* args[] = {&a};
  const hipError_t e = hipLaunchKernel(kern, dim3(blocks), dim3(kLanes * wpb), args, 0, s);
  return e == hipSuccess ? 0 : hip_fail(e, "step2_fused launch");
}

// One sweep of the split step from a caller-filled padded state (the general-limiter kernel).
// Sweep 1 folds its x-speed maximum into the control block and leaves the step open; sweep 2 adds
// the y-speed maximum and closes it, so ctrl->cfl is the maximum over both sweeps.
template <class T, class P>
static int launch_sweep(const claw2d_desc* d, const typename P::Par& par, const void* qbc, void* qout,
                        const void* aux, claw2d_ctrl* ctrl, double dt, int ids, hipStream_t s) {
  const void* kern = ids == 1 ? (const void*)step2_fused<T, P, 0, true, -1, true, 1>
                              : (const void*)step2_fused<T, P, 0, true, -1, true, 2>;
  const int wpb = launch_wpb(d);
  static thread_local long cached[2][3] = {{-1, -1, -1}, {-1, -1, -1}};
  long& res = cached[ids - 1][wpb_index(wpb)];
  if (res < 0) res = resident_waves(kern, wpb);
  StepArgs<T, P> a;
  fill_args(a, d, res);
  a.q0 = (T*)qbc;
  a.q1 = (T*)qout;
  a.ctrl = ctrl;
  a.use_ctrl = 0;
  a.dt_fixed = dt;
  a.par = par;
  a.fuse = ids == 1 ? 0 : 1;
  a.cfl_max = d->cfl_max;
  a.cfl_desired = d->cfl_desired;
  a.dt_max = d->dt_max;
  a.tw = 0;
  a.aux = (const T*)aux;
  const int waves = a.nstrips * a.nsegs;
  a.total_waves = waves;
  const int blocks = (waves + wpb - 1) / wpb;
  void* args[] = {&a};
  const hipError_t e = hipLaunchKernel(kern, dim3(blocks), dim3(kLanes * wpb), args, 0, s);
  return e == hipSuccess ? 0 : hip_fail(e, "step2_fused sweep launch");
}

template <class T>
static int dispatch_sweep(const claw2d_desc* d, const void* qbc, void* qout, const void* aux,
                          claw2d_ctrl* ctrl, double dt, int ids, hipStream_t s) {
  switch (d->solver) {
    case CLAW_SOLVER_SHALLOW_ROE_EFIX:
      return launch_sweep<T, SweP<T>>(d, swe_par<T>(d), qbc, qout, aux, ctrl, dt, ids, s);
    case CLAW_SOLVER_SHALLOW_BATHY_FWAVE:
      return launch_sweep<T, SweBathyP<T>>(d, bathy_par<T>(d), qbc, qout, aux, ctrl, dt, ids, s);
    default:
      return launch_sweep<T, EulerP<T>>(d, euler_par<T>(d), qbc, qout, aux, ctrl, dt, ids, s);
  }
}

template <class T, bool SPLIT>
static int dispatch_typed(const claw2d_desc* d, void* q0, void* q1, const void* aux, claw2d_ctrl* ctrl,
                          int use_ctrl, double dt, hipStream_t s, bool pad) {
  switch (d->solver) {
    case CLAW_SOLVER_SHALLOW_ROE_EFIX:
      return launch_step<T, SweP<T>, SPLIT>(d, swe_par<T>(d), q0, q1, aux, ctrl, use_ctrl, dt, s, pad);
    case CLAW_SOLVER_SHALLOW_BATHY_FWAVE:
      return launch_step<T, SweBathyP<T>, SPLIT>(d, bathy_par<T>(d), q0, q1, aux, ctrl, use_ctrl, dt, s, pad);
    default:
      return launch_step<T, EulerP<T>, SPLIT>(d, euler_par<T>(d), q0, q1, aux, ctrl, use_ctrl, dt, s, pad);
  }
}

template <bool SPLIT>
static int dispatch_step(const claw2d_desc* d, void* q0, void* q1, const void* aux, claw2d_ctrl* ctrl,
                         int use_ctrl, double dt, hipStream_t s, bool pad = false) {
  return d->dtype == CLAW_FP64 ? dispatch_typed<double, SPLIT>(d, q0, q1, aux, ctrl, use_ctrl, dt, s, pad)
                               : dispatch_typed<float, SPLIT>(d, q0, q1, aux, ctrl, use_ctrl, dt, s, pad);
}

// The batched kernel for a method, chosen as step_kernel chooses the single-grid one, so that a
// grid advanced in a batch and alone runs the same arithmetic.
template <class T, class P>
static const void* batch_kernel(const claw2d_desc* d, int* slot) {
  bool same = true;
  for (int p = 1; p < P::NW; ++p) same = same && d->mthlim[p] == d->mthlim[0];
  const int lim = same ? d->mthlim[0] : -1;
  if ((lim == 4 || lim == 1) && d->order == 2 && (d->transverse_waves == 1 || d->transverse_waves == 2)) {
    const bool mc = lim == 4, t2 = d->transverse_waves == 2;
    *slot = (mc ? 0 : 1) + (t2 ? 0 : 2);
    if (mc) return t2 ? (const void*)batch_step<T, P, 4, 2> : (const void*)batch_step<T, P, 4, 1>;
    return t2 ? (const void*)batch_step<T, P, 1, 2> : (const void*)batch_step<T, P, 1, 1>;
  }
  *slot = 4;
  return (const void*)batch_step<T, P, 0, -1>;
}

template <class T, class P>
static int launch_batch(const claw2d_desc* d, const typename P::Par& par, const claw2d_batch_grid* grids,
                        int ngrids, int total_items, void* q0, void* q1, claw2d_ctrl* ctrls,
                        int32_t* counter, hipStream_t s) {
  int slot = 0;
  const void* kern = batch_kernel<T, P>(d, &slot);
  static thread_local long cached[5] = {-1, -1, -1, -1, -1};
  if (cached[slot] < 0) cached[slot] = resident_waves(kern);
  StepArgs<T, P> a;
  fill_args(a, d, cached[slot]);
  a.q0 = (T*)q0;
  a.q1 = (T*)q1;
  a.ctrl = nullptr;
  a.use_ctrl = 0;
  a.dt_fixed = 0.0;
  a.fuse = 0;
  a.total_waves = 0;
  a.cfl_max = a.cfl_desired = a.dt_max = 0.0;
  a.par = par;
  BatchDev b;
  b.grids = grids;
  b.ctrls = ctrls;
  b.counter = counter;
  b.ngrids = ngrids;
  b.total_items = total_items;
  const long res = cached[slot] > 0 ? cached[slot] : 4096;
  const long waves = std::min<long>(total_items, res);
  const int blocks = (int)((waves + kWPB - 1) / kWPB);
  void* args[] = {&a, &b};
  const hipError_t e = hipLaunchKernel(kern, dim3(blocks), dim3(kLanes * kWPB), args, 0, s);
  return e == hipSuccess ? 0 : hip_fail(e, "batch_step launch");
}

static int batch_attempt(const claw2d_desc* d, const claw2d_batch_grid* grids, int ngrids,
                         int total_items, void* q0, void* q1, claw2d_ctrl* ctrls, int32_t* counter,
                         hipStream_t s) {
  const bool swe = d->solver == CLAW_SOLVER_SHALLOW_ROE_EFIX;
  int rc;
  if (d->dtype == CLAW_FP64)
    rc = swe ? launch_batch<double, SweP<double>>(d, swe_par<double>(d), grids, ngrids, total_items, q0, q1, ctrls, counter, s)
             : launch_batch<double, EulerP<double>>(d, euler_par<double>(d), grids, ngrids, total_items, q0, q1, ctrls, counter, s);
  else
    rc = swe ? launch_batch<float, SweP<float>>(d, swe_par<float>(d), grids, ngrids, total_items, q0, q1, ctrls, counter, s)
             : launch_batch<float, EulerP<float>>(d, euler_par<float>(d), grids, ngrids, total_items, q0, q1, ctrls, counter, s);
  if (rc) return rc;
  const int threads = 256, blocks = (ngrids + threads - 1) / threads;
  hipLaunchKernelGGL(batch_control, dim3(blocks), dim3(threads), 0, s, grids, ctrls, (int*)counter,
                     ngrids, d->cfl_max, d->cfl_desired, d->dt_max);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "batch_control launch");
}

static int check_batch(const claw2d_desc* d, const void* grids, int32_t ngrids, int64_t total_items,
                       const void* q0, const void* q1, const void* ctrls, const void* counter) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (d->num_aux > 0) return fail(-25, "batched grids take no aux arrays (no capacity function)");
  if (!grids || !q0 || !q1 || !ctrls || !counter) return fail(-22, "null buffer");
  if (q0 == q1) return fail(-23, "q0 and q1 must differ");
  if (ngrids < 1) return fail(-30, "ngrids must be >= 1");
  if (total_items < 1 || total_items > (1LL << 30)) return fail(-31, "total_items out of range");
  return 0;
}

}  // namespace claw

using namespace claw;

extern "C" {

int claw2d_validate(const claw2d_desc* d) { return check_desc(d); }

size_t claw2d_state_bytes(const claw2d_desc* d) {
  if (check_desc(d) != 0) return 0;
  const size_t es = d->dtype == CLAW_FP64 ? sizeof(double) : sizeof(float);
  return (size_t)d->num_eqn * (size_t)d->mx * (size_t)d->my * es;
}

int claw2d_ctrl_init(const claw2d_desc* d, claw2d_ctrl* c, double t0, double dt0, double tend,
                     int32_t run_to_tend, long long target_steps, long long max_steps) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (!c) return fail(-20, "null control block");
  if (!(dt0 > 0)) return fail(-21, "dt must be positive");
  std::memset(c, 0, sizeof(*c));
  c->t = t0;
  c->tstart = t0;
  c->tend = tend;
  c->flags = (run_to_tend ? CLAW_FLAG_RUN_TO_TEND : 0) | (d->dt_variable ? CLAW_FLAG_DT_VARIABLE : 0);
  c->target_steps = target_steps;
  c->max_steps = max_steps > 0 ? max_steps : (1LL << 62);
  double dt = dt0;
  if (run_to_tend) {
    if (t0 + dt > tend && t0 < tend) dt = tend - t0;
    if (tend - t0 - dt < 1.0e-14 * t0) dt = tend - t0;
  }
  c->dt = dt;
  c->dtmin = dt;
  c->dtmax = dt;
  if (run_to_tend ? (t0 >= tend) : (target_steps <= 0)) c->done = 1;
  return 0;
}

}  // extern "C"

namespace claw {
// Shared bodies of the step / advance entry points (with_aux: the _aux variants).
template <bool SPLIT>
static int step_entry(const claw2d_desc* d, const void* q_in, void* q_out, const void* aux, double dt,
                      claw2d_ctrl* ctrl, void* stream, bool with_aux, const char* name) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (!q_in || !q_out || !ctrl) return fail(-22, "null buffer");
  if (q_in == q_out) return fail(-23, "q_in and q_out must differ");
  if (!(dt > 0)) return fail(-21, "dt must be positive");
  if (!with_aux && d->num_aux > 0) return fail(-26, std::string("num_aux > 0: use ") + name + "_aux");
  if (with_aux && d->num_aux > 0 && !aux) return fail(-22, "null aux buffer");
  return dispatch_step<SPLIT>(d, const_cast<void*>(q_in), q_out, aux, ctrl, 0, dt, (hipStream_t)stream);
}

template <bool SPLIT>
static int advance_entry(const claw2d_desc* d, void* q0, void* q1, const void* aux, claw2d_ctrl* ctrl,
                         long long attempts, void* stream, bool with_aux, const char* name) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (!q0 || !q1 || !ctrl) return fail(-22, "null buffer");
  if (q0 == q1) return fail(-23, "q0 and q1 must differ");
  if (!with_aux && d->num_aux > 0) return fail(-26, std::string("num_aux > 0: use ") + name + "_aux");
  if (with_aux && d->num_aux > 0 && !aux) return fail(-22, "null aux buffer");
  for (long long k = 0; k < attempts; ++k) {
    rc = dispatch_step<SPLIT>(d, q0, q1, aux, ctrl, 1, 0.0, (hipStream_t)stream);
    if (rc) return rc;
  }
  return 0;
}
}  // namespace claw

extern "C" {

int claw2d_step(const claw2d_desc* d, const void* q_in, void* q_out, double dt, claw2d_ctrl* ctrl,
                void* stream) {
  return step_entry<false>(d, q_in, q_out, nullptr, dt, ctrl, stream, false, "claw2d_step");
}

int claw2d_advance(const claw2d_desc* d, void* q0, void* q1, claw2d_ctrl* ctrl, long long attempts,
                   void* stream) {
  return advance_entry<false>(d, q0, q1, nullptr, ctrl, attempts, stream, false, "claw2d_advance");
}

int claw2d_step_ds(const claw2d_desc* d, const void* q_in, void* q_out, double dt, claw2d_ctrl* ctrl,
                   void* stream) {
  return step_entry<true>(d, q_in, q_out, nullptr, dt, ctrl, stream, false, "claw2d_step_ds");
}

int claw2d_advance_ds(const claw2d_desc* d, void* q0, void* q1, claw2d_ctrl* ctrl, long long attempts,
                      void* stream) {
  return advance_entry<true>(d, q0, q1, nullptr, ctrl, attempts, stream, false, "claw2d_advance_ds");
}

}  // extern "C"

namespace claw {
// Checks shared by the padded entry points.  The state's boundary conditions are in the caller's
// ghost cells, so any code, custom included, is accepted for them; dd gets extrap in their place
// (the aux BCs are still applied by index mapping).
static int padded_desc(const claw2d_desc* d, claw2d_desc* dd) {
  if (!d) return fail(-1, "null descriptor");
  *dd = *d;
  for (int k = 0; k < 2; ++k) {
    const int lo = dd->bc_lower[k], hi = dd->bc_upper[k];
    if (lo < 0 || lo > 3 || hi < 0 || hi > 3) return fail(-10, "bc must be custom(0), extrap(1), periodic(2) or wall(3)");
    dd->bc_lower[k] = dd->bc_upper[k] = CLAW_BC_EXTRAP;
  }
  if (dd->mcapa != 0) return fail(-28, "the padded (custom-BC) steps take no capacity function (mcapa must be 0)");
  return check_desc(dd);
}
}  // namespace claw

extern "C" {

int claw2d_sweep_padded(const claw2d_desc* d, const void* qbc_in, void* q_out, const void* aux,
                        double dt, int32_t ids, claw2d_ctrl* ctrl, void* stream) {
  claw2d_desc dd;
  int rc = padded_desc(d, &dd);
  if (rc) return rc;
  if (ids != 1 && ids != 2) return fail(-27, "ids must be 1 (x-sweeps) or 2 (y-sweeps)");
  if (!qbc_in || !q_out || !ctrl) return fail(-22, "null buffer");
  if (qbc_in == q_out) return fail(-23, "qbc_in and q_out must differ");
  if (!(dt > 0)) return fail(-21, "dt must be positive");
  if (dd.num_aux > 0 && !aux) return fail(-22, "null aux buffer");
  return dd.dtype == CLAW_FP64 ? dispatch_sweep<double>(&dd, qbc_in, q_out, aux, ctrl, dt, ids, (hipStream_t)stream)
                               : dispatch_sweep<float>(&dd, qbc_in, q_out, aux, ctrl, dt, ids, (hipStream_t)stream);
}

int claw2d_step_padded(const claw2d_desc* d, const void* qbc_in, void* q_out, const void* aux,
                       double dt, claw2d_ctrl* ctrl, void* stream) {
  claw2d_desc dd;
  int rc = padded_desc(d, &dd);
  if (rc) return rc;
  if (!qbc_in || !q_out || !ctrl) return fail(-22, "null buffer");
  if (qbc_in == q_out) return fail(-23, "qbc_in and q_out must differ");
  if (!(dt > 0)) return fail(-21, "dt must be positive");
  if (dd.num_aux > 0 && !aux) return fail(-22, "null aux buffer");
  return dispatch_step<false>(&dd, const_cast<void*>(qbc_in), q_out, aux, ctrl, 0, dt, (hipStream_t)stream,
                              true);
}

int claw2d_step_aux(const claw2d_desc* d, const void* q_in, void* q_out, const void* aux, double dt,
                    claw2d_ctrl* ctrl, void* stream) {
  return step_entry<false>(d, q_in, q_out, aux, dt, ctrl, stream, true, "claw2d_step");
}

int claw2d_advance_aux(const claw2d_desc* d, void* q0, void* q1, const void* aux, claw2d_ctrl* ctrl,
                       long long attempts, void* stream) {
  return advance_entry<false>(d, q0, q1, aux, ctrl, attempts, stream, true, "claw2d_advance");
}

int claw2d_step_ds_aux(const claw2d_desc* d, const void* q_in, void* q_out, const void* aux, double dt,
                       claw2d_ctrl* ctrl, void* stream) {
  return step_entry<true>(d, q_in, q_out, aux, dt, ctrl, stream, true, "claw2d_step_ds");
}

int claw2d_advance_ds_aux(const claw2d_desc* d, void* q0, void* q1, const void* aux, claw2d_ctrl* ctrl,
                          long long attempts, void* stream) {
  return advance_entry<true>(d, q0, q1, aux, ctrl, attempts, stream, true, "claw2d_advance_ds");
}

size_t claw2d_mass_sum_workspace_bytes(const claw2d_desc*) { return 2 * kSumBlocks * sizeof(double); }

int claw2d_mass_sum(const claw2d_desc* d, const void* q, int32_t m, double* out, void* ws, void* stream) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (!q || !out || !ws) return fail(-22, "null buffer");
  if (m < 0 || m >= d->num_eqn) return fail(-24, "component out of range");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)d->mx * d->my;
  if (d->dtype == CLAW_FP64)
    hipLaunchKernelGGL(mass_partial<double>, dim3(kSumBlocks), dim3(kSumThreads), 0, s,
                       (const double*)q + (size_t)m * n, n, (double*)ws);
  else
    hipLaunchKernelGGL(mass_partial<float>, dim3(kSumBlocks), dim3(kSumThreads), 0, s,
                       (const float*)q + (size_t)m * n, n, (double*)ws);
  hipLaunchKernelGGL(mass_final, dim3(1), dim3(64), 0, s, (const double*)ws, out,
                     m == 0 ? d->depth_ref * (double)n : 0.0);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "mass_sum launch");
}

// ---- claw2d_mass_sum_noalloc: the mass sum with its partials in a file-scope device array.
// Each GPU the library runs on holds one copy of `scratch` (kSumBlocks (sum, compensation) pairs,
// 16 KiB) until the process ends.  Host code never passes the array's name to a kernel (in host
// code it is a stand-in, not a device address): the device address is looked up once per device
// with hipGetSymbolAddress and kept in g_scratch_ptr.  Every call on one GPU shares its array, so
// calls that overlap on one GPU (different streams, or threads) race; mass_partial writes every
// pair before mass_final reads them, so nothing has to start at zero.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[2 * kSumBlocks];
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_for_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "mass_sum_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(-33, "mass_sum_noalloc: device index beyond the library's table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "mass_sum_noalloc: hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

size_t claw2d_mass_sum_noalloc_max_cells(void) { return (size_t)1 << 31; }

int claw2d_mass_sum_noalloc(const claw2d_desc* d, const void* q, int32_t m, double* out, void* stream) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (!q || !out) return fail(-22, "null buffer");
  if (m < 0 || m >= d->num_eqn) return fail(-24, "component out of range");
  // the partials need kSumBlocks pairs whatever the grid; the cell count is bounded by the
  // 32-bit row arithmetic of the kernels
  static_assert(sizeof(scratch) == 2 * kSumBlocks * sizeof(double), "scratch holds every partial pair");
  const size_t n = (size_t)d->mx * d->my;
  if (n > claw2d_mass_sum_noalloc_max_cells())
    return fail(-34, "mass_sum_noalloc: grid larger than 2^31 cells");
  double* ws = nullptr;
  rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (d->dtype == CLAW_FP64)
    hipLaunchKernelGGL(mass_partial<double>, dim3(kSumBlocks), dim3(kSumThreads), 0, s,
                       (const double*)q + (size_t)m * n, n, ws);
  else
    hipLaunchKernelGGL(mass_partial<float>, dim3(kSumBlocks), dim3(kSumThreads), 0, s,
                       (const float*)q + (size_t)m * n, n, ws);
  hipLaunchKernelGGL(mass_final, dim3(1), dim3(64), 0, s, (const double*)ws, out,
                     m == 0 ? d->depth_ref * (double)n : 0.0);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "mass_sum_noalloc launch");
}

size_t claw2d_abi_struct_size(int32_t which) {
  return which == 0 ? sizeof(claw2d_desc)
                    : (which == 1 ? sizeof(claw2d_ctrl) : (which == 2 ? sizeof(claw2d_batch_grid) : 0));
}

int claw2d_batch_prepare(const claw2d_desc* d, claw2d_batch_grid* grids, int32_t ngrids, int32_t seg_h,
                         int64_t* total_items) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (!grids || !total_items) return fail(-22, "null buffer");
  if (ngrids < 1) return fail(-30, "ngrids must be >= 1");
  if (seg_h < 0) return fail(-14, "seg_h must be >= 0");
  long long items = 0;
  for (int g = 0; g < ngrids; ++g) {
    claw2d_batch_grid& G = grids[g];
    if (G.mx < 2 || G.my < 2) return fail(-5, "every grid needs mx, my >= 2");
    if (!(G.dx > 0) || !(G.dy > 0)) return fail(-13, "dx, dy must be positive");
    if (G.offset < 0) return fail(-32, "negative grid offset");
    const int h = seg_h > 0 ? seg_h : std::max(16, std::min(64, (int)G.my));
    G.seg_h = h;
    G.nstrips = (G.mx + kOwn - 1) / kOwn;
    G.nsegs = (G.my + h - 1) / h;
    G.item0 = (int32_t)items;
    items += (long long)G.nstrips * G.nsegs;
    if (items > (1LL << 30)) return fail(-31, "too many work items");
  }
  *total_items = items;
  return 0;
}

int claw2d_batch_advance(const claw2d_desc* d, const claw2d_batch_grid* grids, int32_t ngrids,
                         int64_t total_items, void* q0, void* q1, claw2d_ctrl* ctrls, int32_t* counter,
                         long long attempts, void* stream) {
  int rc = check_batch(d, grids, ngrids, total_items, q0, q1, ctrls, counter);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  for (long long k = 0; k < attempts; ++k) {
    rc = batch_attempt(d, grids, ngrids, (int)total_items, q0, q1, ctrls, counter, s);
    if (rc) return rc;
  }
  return 0;
}

int claw2d_batch_graph_create(const claw2d_desc* d, const claw2d_batch_grid* grids, int32_t ngrids,
                              int64_t total_items, void* q0, void* q1, claw2d_ctrl* ctrls,
                              int32_t* counter, long long attempts, void** graph_exec_out) {
  int rc = check_batch(d, grids, ngrids, total_items, q0, q1, ctrls, counter);
  if (rc) return rc;
  if (!graph_exec_out || attempts < 1) return fail(-33, "bad graph arguments");
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  hipGraph_t graph = nullptr;
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(s);
    return hip_fail(e, "hipStreamBeginCapture");
  }
  for (long long k = 0; k < attempts && rc == 0; ++k)
    rc = batch_attempt(d, grids, ngrids, (int)total_items, q0, q1, ctrls, counter, s);
  e = hipStreamEndCapture(s, &graph);
  if (rc == 0 && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  if (rc == 0) {
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
  }
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(s);
  if (rc == 0) *graph_exec_out = (void*)exec;
  return rc;
}

int claw2d_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return fail(-22, "null graph");
  const hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "hipGraphLaunch");
}

int claw2d_graph_destroy(void* graph_exec) {
  if (!graph_exec) return 0;
  const hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph_exec);
  return e == hipSuccess ? 0 : hip_fail(e, "hipGraphExecDestroy");
}

const char* claw2d_last_error(void) { return g_err.c_str(); }

}  // extern "C"
#else
}  // namespace claw
#endif  // CLAW_KERNEL_UNIT

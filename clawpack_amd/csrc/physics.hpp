// Riemann-solver policies for the fused 2D step (gfx950).
//
// Each policy restates one Clawpack Riemann plugin in "wave-scalar" form: at an interface the
// jump is split into waves alpha_p * e_p along the Roe eigenvectors e_p, and everything the
// step needs afterwards is a per-wave scalar times e_p:
//
//   amdq = sum_p sf_p alpha_p e_p          (sf_p: left-going speed, entropy fixed)
//   apdq = sum_p (s_p - sf_p) alpha_p e_p
//   cqxx = sum_p |s_p|(1 - |s_p| dt/dx) phi_p alpha_p e_p
//
// and the transverse split of sum_p A_p e_p (rpt2 applied to amdq+cqxx or apdq-cqxx) has the
// closed form beta = M A with M = L_t E, the transverse left eigenvectors applied to the normal
// right eigenvectors (for both systems M has entries 1/2, +-a, +-1/(2a) and an identity block).
// The results equal upstream's component-wise evaluation up to rounding.
//
//   SweP      <- riemann/src/rpn2_shallow_roe_with_efix.f90, rpt2_shallow_roe_with_efix.f90
//   EulerP    <- riemann/src/rpn2_euler_4wave.f90, rpt2_euler.f90
//   SweBathyP <- riemann/src/rpn2_shallow_bathymetry_fwave.f90, rpt2_shallow_bathymetry_fwave.f90
//                (f-waves: alpha_p are f-wave strengths, sf_p the left-going fraction 1, 1/2 or 0,
//                 amdq = sum sf_p alpha_p e_p, apdq = sum (1 - sf_p) alpha_p e_p, and the
//                 correction uses sign(s_p) in place of |s_p|)
// (upstream riemann submodule; empty in the reference tree, .gitmodules:10-12).
//
// Limiter (upstream limiter.f90/philim): theta_p = W_p(upwind).W_p / W_p.W_p.  With
// W = alpha e, theta = t/alpha where t = alpha_upwind (e_upwind.e)/(e.e), and for minmod, MC and
// superbee phi(theta)*alpha is a positively homogeneous function of (alpha, t) that needs no
// division by alpha; e.e >= 1 so its reciprocal is always well conditioned.
//
// Vectors are in the interface frame (h, normal, transverse[, E]); DIR = 0 is an x-interface
// (normal = component 1), DIR = 1 a y-interface (normal = component 2).
//
// Entropy fix (Harten-Hyman, as upstream): the transonic tests "u - c > 0 right of wave 1" and
// "u + c < 0 left of the last wave" are evaluated in a division- and sqrt-free squared form; the
// exact speeds are only formed inside the (rare) transonic branch.  The squared test equals the
// upstream test except within rounding of the threshold, where both branches give the same
// fluctuation to within that rounding.
#pragma once
#include <hip/hip_runtime.h>

namespace claw {

// Reciprocal and reciprocal square root of a positive normal number: the hardware estimate
// (v_rcp_f64 / v_rsq_f64, about 23 correct bits) refined by two Newton-Raphson steps, which
// leaves about 1 ulp of error instead of the correctly rounded IEEE sequences.  FP32 uses the
// hardware instructions directly (about 1 ulp).
__device__ __forceinline__ double frcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-(x * y), y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-(x * y), y, 1.0);
  return fma(0.5 * y, e, y);
}
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ __forceinline__ double vmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float vmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double vabs(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float vabs(float a) { return __builtin_fabsf(a); }

// phi(t/alpha)*alpha for limiter `meth` (ids of pyclaw/limiters/tvd.py: minmod 1, superbee 2,
// van Leer 3, MC 4, beam-warming 5, 0 = unlimited).  alpha == 0 gives 0 (upstream skips a wave
// of zero norm, which is the zero wave).  minmod, superbee and MC are odd in (alpha, t), so they
// are evaluated on |alpha| and sign(alpha) t and the sign is put back: the same values, bit for bit,
// as the two-branch form (negation and the factors 2 and 1/2 are exact), with no selects.
template <class T> __device__ __forceinline__ T limited(T al, T t, int meth) {
  const T sg = copysign(T(1), al), a = vabs(al), ts = sg * t;
  switch (meth) {
    case 1:  // minmod: max(0, min(1, theta))
      return sg * vmax(T(0), vmin(a, ts));
    case 2: {  // superbee: max(0, min(1, 2 theta), min(2, theta))
      const T lo = vmax(vmin(a, T(2) * ts), vmin(T(2) * a, ts));
      return sg * vmax(T(0), lo);
    }
    case 3: {  // van Leer: (theta + |theta|)/(1 + |theta|)
      if (al == T(0)) return T(0);
      const T th = t / al, ath = th < T(0) ? -th : th;
      return al * ((th + ath) / (T(1) + ath));
    }
    case 4: {  // MC: max(0, min((1 + theta)/2, 2, 2 theta))
      const T c = T(0.5) * (a + ts);
      return sg * vmax(T(0), vmin(c, T(2) * vmin(a, ts)));
    }
    case 5: return al == T(0) ? T(0) : t;  // beam-warming: theta
    default: return al;
  }
}

// ------------------------------------------------------------------------------------------
// Shallow water, Roe + entropy fix.  Waves e1 = (1, u-a, v), e2 = (0, 0, 1), e3 = (1, u+a, v);
// transverse eigenvectors f1 = (1, u, v-a), f2 = (0, 1, 0), f3 = (1, u, v+a).
// ------------------------------------------------------------------------------------------
//
// Stored depth (FP32 only).  The FP32 state stores the depth as its deviation eta = h - h0 from a
// reference depth h0 (Par::h0, the descriptor's depth_ref; 0 stores h itself).  Rounding the
// stored depth to FP32 is the dominant error of the FP32 path, and it is relative to the stored
// value: storing eta, which is small where the water is near h0, makes it that much smaller,
// while the jumps across interfaces (the wave strengths) are the same differences of stored
// values.  Wherever the depth itself enters (square roots, the Roe average, the transonic tests)
// it is formed as h0 + eta.  The FP64 state always stores h (h0 is not read).
template <class T> struct SweP {
  static constexpr int NEQ = 3, NW = 3, NC = 7, NL = 6;
  static constexpr int NAUX = 0;        // aux fields read per cell
  static constexpr bool FWAVE = false;  // waves (not f-waves)
  static constexpr bool kRef = sizeof(T) == 4;  // stored depth is h - h0
  // e_p(A).e_p(B) is identically 1 for the shear wave
  __device__ __forceinline__ static constexpr bool unit_m(int p) { return p == 1; }
  struct Par { T g, sqg, h0; };
  // per-cell: q as stored, sqrt(h), 1/sqrt(h), hu/sqrt(h), hv/sqrt(h)
  struct Cell { T q[3]; T s, rs, P, Q; };
  struct Iface { T al[3]; T sf[3]; T u, v, a, ha; };

  // depth from a stored depth value
  __device__ __forceinline__ static T depth(T stored, const Par& pr) { return kRef ? stored + pr.h0 : stored; }

  __device__ __forceinline__ static bool bad(const T q[3], const Par& pr) { return !(depth(q[0], pr) > T(0)); }

  __device__ __forceinline__ static Cell prep(const T q[3], const Par& pr) {
    Cell c;
    c.q[0] = q[0]; c.q[1] = q[1]; c.q[2] = q[2];
    const T h = depth(q[0], pr);
    c.rs = frsq(h);
    c.s = h * c.rs;
    c.P = q[1] * c.rs;
    c.Q = q[2] * c.rs;
    return c;
  }
  __device__ __forceinline__ static void pack(const Cell& c, T o[NC]) {
    o[0] = c.q[0]; o[1] = c.q[1]; o[2] = c.q[2]; o[3] = c.s; o[4] = c.rs; o[5] = c.P; o[6] = c.Q;
  }
  __device__ __forceinline__ static Cell unpack(const T o[NC]) {
    Cell c;
    c.q[0] = o[0]; c.q[1] = o[1]; c.q[2] = o[2]; c.s = o[3]; c.rs = o[4]; c.P = o[5]; c.Q = o[6];
    return c;
  }

  template <int DIR>
  __device__ __forceinline__ static void rpn2(const Cell& L, const Cell& R, const Par& pr, Iface& f) {
    constexpr int mu = DIR == 0 ? 1 : 2, mv = DIR == 0 ? 2 : 1;
    const T pnL = DIR == 0 ? L.P : L.Q, ptL = DIR == 0 ? L.Q : L.P;
    const T pnR = DIR == 0 ? R.P : R.Q, ptR = DIR == 0 ? R.Q : R.P;
    const T inv = frcp(L.s + R.s);
    const T u = (pnL + pnR) * inv;
    const T v = (ptL + ptR) * inv;
    const T gh = kRef ? pr.g * (pr.h0 + (L.q[0] + R.q[0]) * T(0.5)) : pr.g * ((L.q[0] + R.q[0]) * T(0.5));
    const T ra = frsq(gh);
    const T a = gh * ra;
    const T ha = T(0.5) * ra;
    const T d0 = R.q[0] - L.q[0], dn = R.q[mu] - L.q[mu], dt = R.q[mv] - L.q[mv];
    const T s1 = u - a, s3 = u + a;
    const T a1 = (s3 * d0 - dn) * ha;
    const T a2 = dt - v * d0;
    const T a3 = (dn - s1 * d0) * ha;
    T sf1 = vmin(s1, T(0)), sf2 = T(0), sf3 = T(0);
    const T s0 = pnL * L.rs - pr.sqg * L.s;  // u - c in the left cell
    if (s0 < T(0)) {
      const T h1 = depth(L.q[0] + a1, pr), hu1 = L.q[mu] + a1 * s1;
      if (h1 > T(0) && hu1 > T(0) && hu1 * hu1 > pr.g * (h1 * h1) * h1) {
        const T s1r = hu1 / h1 - sqrt(pr.g * h1);  // u - c right of the 1-wave
        sf1 = s0 * (s1r - s1) / (s1r - s0);
      }
    }
    if (u < T(0)) {
      sf2 = u;
      sf3 = vmin(s3, T(0));
      const T h3 = depth(R.q[0] - a3, pr), hu3 = R.q[mu] - a3 * s3;
      if (h3 > T(0) && hu3 < T(0) && hu3 * hu3 > pr.g * (h3 * h3) * h3) {
        const T s3r = pnR * R.rs + pr.sqg * R.s;  // u + c in the right cell
        if (s3r > T(0)) {
          const T s3l = hu3 / h3 + sqrt(pr.g * h3);  // u + c left of the 3-wave
          sf3 = s3l * (s3r - s3) / (s3r - s3l);
        }
      }
    }
    f.al[0] = a1; f.al[1] = a2; f.al[2] = a3;
    f.sf[0] = sf1; f.sf[1] = sf2; f.sf[2] = sf3;
    f.u = u; f.v = v; f.a = a; f.ha = ha;
  }

  __device__ __forceinline__ static T smax(const Iface& f) { return vabs(f.u) + f.a; }
  __device__ __forceinline__ static T speed(const Iface& f, int p) {
    return p == 0 ? f.u - f.a : (p == 1 ? f.u : f.u + f.a);
  }
  // limiter data sent to neighbours: alpha1..3 and the eigenvector parameters s1, s3, v
  __device__ __forceinline__ static void lim(const Iface& f, T o[NL]) {
    o[0] = f.al[0]; o[1] = f.al[1]; o[2] = f.al[2]; o[3] = f.u - f.a; o[4] = f.u + f.a; o[5] = f.v;
  }
  // m_p = e_p(A).e_p(B)
  __device__ __forceinline__ static void mdot(const T A[NL], const T B[NL], T m[NW]) {
    const T vv1 = fma(A[5], B[5], T(1));
    m[0] = fma(A[3], B[3], vv1);
    m[1] = T(1);
    m[2] = fma(A[4], B[4], vv1);
  }
  __device__ __forceinline__ static void inv_norm(const Iface& f, T in[NW]) {
    const T vv1 = fma(f.v, f.v, T(1));
    const T s1 = f.u - f.a, s3 = f.u + f.a;
    const T n1 = fma(s1, s1, vv1), n3 = fma(s3, s3, vv1);
    const T r = frcp(n1 * n3);
    in[0] = n3 * r;
    in[1] = T(1);
    in[2] = n1 * r;
  }
  // sum_p c_p e_p
  __device__ __forceinline__ static void expand_n(const Iface& f, const T c[NW], T o[NEQ]) {
    o[0] = c[0] + c[2];
    o[1] = c[0] * (f.u - f.a) + c[2] * (f.u + f.a);
    o[2] = fma(o[0], f.v, c[1]);
  }
  // transverse strengths of sum_p A_p e_p
  __device__ __forceinline__ static void tbeta(const Iface& f, const T A[NW], T b[NW]) {
    const T m = T(0.5) * (A[0] + A[2]), d = A[1] * f.ha;
    b[0] = m - d;
    b[1] = f.a * (A[2] - A[0]);
    b[2] = m + d;
  }
  // Interface constants of the transverse eigenvectors, formed once per interface and shared by
  // both forms of the transverse split, so the two forms see the same bits.
  struct TConst { T vm, vp; };  // v - a, v + a
  __device__ __forceinline__ static TConst tconst(const Iface& f) { return TConst{f.v - f.a, f.v + f.a}; }
  __device__ __forceinline__ static void tspeeds(const Iface& f, const TConst& tc, T sg[NW]) {
    sg[0] = tc.vm; sg[1] = f.v; sg[2] = tc.vp;
  }
  // sum_j c_j f_j
  __device__ __forceinline__ static void expand_t(const Iface& f, const TConst& tc, const T c[NW], T o[NEQ]) {
    o[0] = c[0] + c[2];
    o[1] = fma(o[0], f.u, c[1]);
    o[2] = fma(c[2], tc.vp, c[0] * tc.vm);
  }
  // expand_t of the down-going (wm) and up-going (wp) halves of w*beta when v - a < 0 < v + a:
  // wave 1 is down-going, wave 3 up-going, the shear wave goes with the sign of v.  Bitwise equal
  // to expand_t of the general split, whose other entries are (signed) zeros.
  __device__ __forceinline__ static void expand_t_split(const Iface& f, const TConst& tc, T w, const T b[NW],
                                                        T Gm[NEQ], T Gp[NEQ]) {
    const T c0 = (w * tc.vm) * b[0];
    const T c1m = (w * vmin(f.v, T(0))) * b[1], c1p = (w * vmax(f.v, T(0))) * b[1];
    const T c2 = (w * tc.vp) * b[2];
    Gm[0] = c0;
    Gm[1] = fma(c0, f.u, c1m);
    Gm[2] = c0 * tc.vm;
    Gp[0] = c2;
    Gp[1] = fma(c2, f.u, c1p);
    Gp[2] = c2 * tc.vp;
  }
};

// ------------------------------------------------------------------------------------------
// Euler, 4 waves (u-c, entropy u, shear u, u+c), Roe + entropy fix on waves 1 and 4.
// e1 = (1, u-a, v, H-ua), e2 = (1, u, v, q2/2), e3 = (0, 0, 1, v), e4 = (1, u+a, v, H+ua);
// f1 = (1, u, v-a, H-va), f2 = (1, u, v, q2/2), f3 = (0, 1, 0, u), f4 = (1, u, v+a, H+va).
// ------------------------------------------------------------------------------------------
template <class T> struct EulerP {
  static constexpr int NEQ = 4, NW = 4, NC = 9, NL = 8;
  static constexpr int NAUX = 0;
  static constexpr bool FWAVE = false;
  __device__ __forceinline__ static constexpr bool unit_m(int) { return false; }
  struct Par { T gamma, gamma1; };
  // per-cell: q, 1/sqrt(rho), rho u/sqrt(rho), rho v/sqrt(rho), (E+p)/sqrt(rho), c
  // (sqrt(rho) = rho/sqrt(rho) is formed where it is used: one register pair less per cell)
  struct Cell { T q[4]; T rs, P, Q, Hs, c; };
  struct Iface { T al[4]; T sf[4]; T u, v, a, H, i2a; };

  __device__ __forceinline__ static bool bad(const T q[4], const Par&) { return !(q[0] > T(0)); }

  __device__ __forceinline__ static Cell prep(const T q[4], const Par& pr) {
    Cell c;
    c.q[0] = q[0]; c.q[1] = q[1]; c.q[2] = q[2]; c.q[3] = q[3];
    c.rs = frsq(q[0]);
    c.P = q[1] * c.rs;
    c.Q = q[2] * c.rs;
    const T p = pr.gamma1 * (q[3] - T(0.5) * (c.P * c.P + c.Q * c.Q));
    c.Hs = (q[3] + p) * c.rs;
    const T c2 = pr.gamma * p * (c.rs * c.rs);
    c.c = c2 * frsq(c2);
    return c;
  }
  __device__ __forceinline__ static void pack(const Cell& c, T o[NC]) {
    o[0] = c.q[0]; o[1] = c.q[1]; o[2] = c.q[2]; o[3] = c.q[3];
    o[4] = c.rs; o[5] = c.P; o[6] = c.Q; o[7] = c.Hs; o[8] = c.c;
  }
  __device__ __forceinline__ static Cell unpack(const T o[NC]) {
    Cell c;
    c.q[0] = o[0]; c.q[1] = o[1]; c.q[2] = o[2]; c.q[3] = o[3];
    c.rs = o[4]; c.P = o[5]; c.Q = o[6]; c.Hs = o[7]; c.c = o[8];
    return c;
  }

  // Squared-form test of "n/rho - c > 0" (right = true) or "n/rho + c < 0" (right = false) for
  // the state (rho, n, t, E) with c = sqrt(gamma p / rho), as upstream evaluates it.  A middle state
  // of the Roe split can have rho < 0 together with p < 0; p/rho > 0 there and c is real, so
  // upstream's test can pass, and the fix is applied.  With Prho = p rho = gamma1 (E rho -
  // (n^2 + t^2)/2), c = sqrt(gamma Prho)/|rho| is real iff Prho >= 0, and for either sign of rho
  //   n/rho - c > 0  <=>  n/rho > 0 and n^2 > gamma Prho,
  //   n/rho + c < 0  <=>  n/rho < 0 and n^2 > gamma Prho.
  __device__ __forceinline__ static bool sonic_cross(T rho, T n, T t, T E, const Par& pr, bool right) {
    const T Prho = pr.gamma1 * (E * rho - T(0.5) * (n * n + t * t));
    const bool pos = (n > T(0) && rho > T(0)) || (n < T(0) && rho < T(0));  // n/rho > 0
    const bool neg = (n < T(0) && rho > T(0)) || (n > T(0) && rho < T(0));  // n/rho < 0
    return (right ? pos : neg) && Prho >= T(0) && n * n > pr.gamma * Prho;
  }
  __device__ __forceinline__ static T sonic_speed(T rho, T n, T t, T E, const Par& pr, T sign) {
    const T p = pr.gamma1 * (E - T(0.5) * (n * n + t * t) / rho);
    return n / rho + sign * sqrt(pr.gamma * p / rho);
  }

  template <int DIR>
  __device__ __forceinline__ static void rpn2(const Cell& L, const Cell& R, const Par& pr, Iface& f) {
    constexpr int mu = DIR == 0 ? 1 : 2, mv = DIR == 0 ? 2 : 1;
    const T pnL = DIR == 0 ? L.P : L.Q, ptL = DIR == 0 ? L.Q : L.P;
    const T pnR = DIR == 0 ? R.P : R.Q, ptR = DIR == 0 ? R.Q : R.P;
    const T inv = frcp(L.q[0] * L.rs + R.q[0] * R.rs);
    const T u = (pnL + pnR) * inv;
    const T v = (ptL + ptR) * inv;
    const T H = (L.Hs + R.Hs) * inv;
    const T u2v2 = u * u + v * v;
    const T a2 = pr.gamma1 * (H - T(0.5) * u2v2);
    const T ra = frsq(a2);
    const T a = a2 * ra;
    const T g1a2 = pr.gamma1 * (ra * ra);
    const T euv = H - u2v2;
    const T i2a = T(0.5) * ra;
    const T d0 = R.q[0] - L.q[0], dn = R.q[mu] - L.q[mu], dt = R.q[mv] - L.q[mv],
            dE = R.q[3] - L.q[3];
    const T ae = g1a2 * (euv * d0 + u * dn + v * dt - dE);
    const T as = dt - v * d0;
    const T a4 = (dn + (a - u) * d0 - a * ae) * i2a;
    const T a1 = d0 - ae - a4;
    const T s1 = u - a, s4 = u + a;
    T sf1 = vmin(s1, T(0)), sfu = T(0), sf4 = T(0);
    const T s0 = pnL * L.rs - L.c;
    if (s0 < T(0)) {
      const T r1 = L.q[0] + a1, n1 = L.q[mu] + a1 * s1, t1 = L.q[mv] + a1 * v,
              E1 = L.q[3] + a1 * (H - u * a);
      if (sonic_cross(r1, n1, t1, E1, pr, true)) {
        const T s1r = sonic_speed(r1, n1, t1, E1, pr, T(-1));
        sf1 = s0 * (s1r - s1) / (s1r - s0);
      }
    }
    if (u < T(0)) {
      sfu = u;
      sf4 = vmin(s4, T(0));
      const T r4 = R.q[0] - a4, n4 = R.q[mu] - a4 * s4, t4 = R.q[mv] - a4 * v,
              E4 = R.q[3] - a4 * (H + u * a);
      if (sonic_cross(r4, n4, t4, E4, pr, false)) {
        const T s4r = pnR * R.rs + R.c;
        if (s4r > T(0)) {
          const T s4l = sonic_speed(r4, n4, t4, E4, pr, T(1));
          sf4 = s4l * (s4r - s4) / (s4r - s4l);
        }
      }
    }
    f.al[0] = a1; f.al[1] = ae; f.al[2] = as; f.al[3] = a4;
    f.sf[0] = sf1; f.sf[1] = sfu; f.sf[2] = sfu; f.sf[3] = sf4;
    f.u = u; f.v = v; f.a = a; f.H = H; f.i2a = i2a;
  }

  __device__ __forceinline__ static T smax(const Iface& f) { return vabs(f.u) + f.a; }
  __device__ __forceinline__ static T speed(const Iface& f, int p) {
    return p == 0 ? f.u - f.a : (p == 3 ? f.u + f.a : f.u);
  }
  // limiter data: alpha1..4, u, v, a, H
  __device__ __forceinline__ static void lim(const Iface& f, T o[NL]) {
    o[0] = f.al[0]; o[1] = f.al[1]; o[2] = f.al[2]; o[3] = f.al[3];
    o[4] = f.u; o[5] = f.v; o[6] = f.a; o[7] = f.H;
  }
  __device__ __forceinline__ static void mdot(const T A[NL], const T B[NL], T m[NW]) {
    const T uA = A[4], vA = A[5], aA = A[6], HA = A[7];
    const T uB = B[4], vB = B[5], aB = B[6], HB = B[7];
    const T vv1 = fma(vA, vB, T(1));
    m[0] = fma(HA - uA * aA, HB - uB * aB, fma(uA - aA, uB - aB, vv1));
    m[1] = fma(T(0.25) * (uA * uA + vA * vA), uB * uB + vB * vB, fma(uA, uB, vv1));
    m[2] = vv1;
    m[3] = fma(HA + uA * aA, HB + uB * aB, fma(uA + aA, uB + aB, vv1));
  }
  __device__ __forceinline__ static void inv_norm(const Iface& f, T in[NW]) {
    const T vv1 = fma(f.v, f.v, T(1));
    const T s1 = f.u - f.a, s4 = f.u + f.a, e1 = f.H - f.u * f.a, e4 = f.H + f.u * f.a;
    const T k = T(0.5) * (f.u * f.u + f.v * f.v);
    const T n1 = fma(e1, e1, fma(s1, s1, vv1));
    const T n2 = fma(k, k, fma(f.u, f.u, vv1));
    const T n4 = fma(e4, e4, fma(s4, s4, vv1));
    const T n14 = n1 * n4, n23 = n2 * vv1;
    const T r = frcp(n14 * n23);
    const T r14 = n23 * r, r23 = n14 * r;
    in[0] = n4 * r14;
    in[1] = vv1 * r23;
    in[2] = n2 * r23;
    in[3] = n1 * r14;
  }
  __device__ __forceinline__ static void expand_n(const Iface& f, const T c[NW], T o[NEQ]) {
    const T sum = c[0] + c[1] + c[3];
    o[0] = sum;
    o[1] = c[0] * (f.u - f.a) + c[1] * f.u + c[3] * (f.u + f.a);
    o[2] = fma(sum, f.v, c[2]);
    o[3] = c[0] * (f.H - f.u * f.a) + c[1] * (T(0.5) * (f.u * f.u + f.v * f.v)) + c[2] * f.v +
           c[3] * (f.H + f.u * f.a);
  }
  __device__ __forceinline__ static void tbeta(const Iface& f, const T A[NW], T b[NW]) {
    const T m = T(0.5) * (A[0] + A[3]), d = A[2] * f.i2a;
    b[0] = m - d;
    b[1] = A[1];
    b[2] = f.a * (A[3] - A[0]);
    b[3] = m + d;
  }
  // Interface constants of the transverse eigenvectors, formed once per interface and shared by
  // both forms of the transverse split.  Each form used to evaluate H -+ v a and (u^2 + v^2)/2
  // itself, and the compiler may contract such an expression differently in the two copies; the
  // form is chosen per wavefront, so the rounding then depended on a column's position in its strip.
  struct TConst { T vm, vp, k, Hm, Hp; };  // v - a, v + a, (u^2 + v^2)/2, H - v a, H + v a
  __device__ __forceinline__ static TConst tconst(const Iface& f) {
    TConst t;
    t.vm = f.v - f.a;
    t.vp = f.v + f.a;
    t.k = T(0.5) * (f.u * f.u + f.v * f.v);
    t.Hm = f.H - f.v * f.a;
    t.Hp = f.H + f.v * f.a;
    return t;
  }
  __device__ __forceinline__ static void tspeeds(const Iface& f, const TConst& tc, T sg[NW]) {
    sg[0] = tc.vm; sg[1] = f.v; sg[2] = f.v; sg[3] = tc.vp;
  }
  // Both forms below are written with explicit fma() and contraction off, so that they are the same
  // IEEE operations on every lane (a contracted "a*b + c" could be fused differently in the two).
  __device__ __forceinline__ static void expand_t(const Iface& f, const TConst& tc, const T c[NW], T o[NEQ]) {
    {
#pragma clang fp contract(off)
      const T sum = (c[0] + c[1]) + c[3];
      o[0] = sum;
      o[1] = fma(sum, f.u, c[2]);
      o[2] = fma(c[3], tc.vp, fma(c[1], f.v, c[0] * tc.vm));
      o[3] = fma(c[3], tc.Hp, fma(c[2], f.u, fma(c[1], tc.k, c[0] * tc.Hm)));
    }
  }
  // expand_t of the down-going and up-going halves of w*beta when v - a < 0 < v + a (wave 1
  // down, wave 4 up, entropy and shear waves with the sign of v); bitwise equal to expand_t of the
  // general split, whose other entries are (signed) zeros.
  __device__ __forceinline__ static void expand_t_split(const Iface& f, const TConst& tc, T w, const T b[NW],
                                                        T Gm[NEQ], T Gp[NEQ]) {
    {
#pragma clang fp contract(off)
      const T wvm = w * vmin(f.v, T(0)), wvp = w * vmax(f.v, T(0));
      const T c0 = (w * tc.vm) * b[0], c3 = (w * tc.vp) * b[3];
      const T c1m = wvm * b[1], c2m = wvm * b[2], c1p = wvp * b[1], c2p = wvp * b[2];
      Gm[0] = c0 + c1m;
      Gm[1] = fma(Gm[0], f.u, c2m);
      Gm[2] = fma(c1m, f.v, c0 * tc.vm);
      Gm[3] = fma(c2m, f.u, fma(c1m, tc.k, c0 * tc.Hm));
      Gp[0] = c1p + c3;
      Gp[1] = fma(Gp[0], f.u, c2p);
      Gp[2] = fma(c3, tc.vp, c1p * f.v);
      Gp[3] = fma(c3, tc.Hp, fma(c2p, f.u, c1p * tc.k));
    }
  }
};

// ------------------------------------------------------------------------------------------
// Shallow water over bathymetry b = aux(1), f-wave form (Bale, LeVeque, Mitran, Rossmanith 2002).
// The jump delta = f(q_r) - f(q_l) + (0, g hbar (b_r - b_l), 0) is split along SweP's Roe
// eigenvectors e1 = (1, u-a, v), e2 = (0, 0, 1), e3 = (1, u+a, v) into f-wave strengths
//   beta1 = (s3 d1 - d2)/(2a),  beta2 = d3 - v d1,  beta3 = (d2 - s1 d1)/(2a),
// with Roe speeds; the limiter, the transverse split and the eigenvector algebra are SweP's.
// A cell with h <= dry_tolerance has zero velocity (P = Q = 0); an interface between two dry
// cells has no waves (u = v = a = 0 there, which also keeps the transverse algebra finite).
// Flux terms from the per-cell P = hu/sqrt(h), Q = hv/sqrt(h), s = sqrt(h):
//   hu = s P,  hu^2 = P^2,  huv = P Q.
// ------------------------------------------------------------------------------------------
template <class T> struct SweBathyP : SweP<T> {
  static constexpr int NAUX = 1;
  static constexpr bool FWAVE = true;
  static constexpr int NC = 7;
  struct Par { T g, hg, tol; };  // g, g/2, dry_tolerance
  // per-cell: q, sqrt(h), hu/sqrt(h), hv/sqrt(h) (zero when dry), bathymetry
  struct Cell { T q[3]; T s, P, Q, b; };
  using Iface = typename SweP<T>::Iface;

  // depth 0 is a dry cell, a negative depth (or NaN) is a bad state
  __device__ __forceinline__ static bool bad(const T q[3], const Par&) { return !(q[0] >= T(0)); }

  __device__ __forceinline__ static Cell prep(const T q[3], T b, const Par& pr) {
    Cell c;
    c.q[0] = q[0]; c.q[1] = q[1]; c.q[2] = q[2];
    c.b = b;
    // clamped away from 0 so that a dry cell (h = 0) gets s = 0 without a 0 * inf
    const T rs = frsq(vmax(q[0], T(1e-30)));
    c.s = q[0] * rs;
    const bool wet = q[0] > pr.tol;
    c.P = wet ? q[1] * rs : T(0);
    c.Q = wet ? q[2] * rs : T(0);
    return c;
  }
  __device__ __forceinline__ static void pack(const Cell& c, T o[NC]) {
    o[0] = c.q[0]; o[1] = c.q[1]; o[2] = c.q[2]; o[3] = c.s; o[4] = c.P; o[5] = c.Q; o[6] = c.b;
  }
  __device__ __forceinline__ static Cell unpack(const T o[NC]) {
    Cell c;
    c.q[0] = o[0]; c.q[1] = o[1]; c.q[2] = o[2]; c.s = o[3]; c.P = o[4]; c.Q = o[5]; c.b = o[6];
    return c;
  }

  // left-going fraction of an f-wave of speed s
  __device__ __forceinline__ static T left_frac(T s) {
    return s < T(0) ? T(1) : (s > T(0) ? T(0) : T(0.5));
  }

  template <int DIR>
  __device__ __forceinline__ static void rpn2(const Cell& L, const Cell& R, const Par& pr, Iface& f) {
    const T pnL = DIR == 0 ? L.P : L.Q, ptL = DIR == 0 ? L.Q : L.P;
    const T pnR = DIR == 0 ? R.P : R.Q, ptR = DIR == 0 ? R.Q : R.P;
    const bool wet = L.q[0] > pr.tol || R.q[0] > pr.tol;
    const T inv = wet ? frcp(L.s + R.s) : T(0);
    const T u = (pnL + pnR) * inv;
    const T v = (ptL + ptR) * inv;
    const T hb = T(0.5) * (L.q[0] + R.q[0]);
    const T gh = wet ? pr.g * hb : T(1);
    const T ra = frsq(gh);
    const T a = wet ? gh * ra : T(0);
    const T ha = wet ? T(0.5) * ra : T(0);
    // Equal states must give exactly zero f-waves (an extrap ghost interface): a contracted
    // fma(a, b, -(c*d)) with a*b == c*d leaves the rounding error of c*d, and beam-warming turns
    // any nonzero wave, however small, into a full upwind wave.  So both products are rounded.
    T d1, d3;
    {
#pragma clang fp contract(off)
      d1 = R.s * pnR - L.s * pnL;
      d3 = pnR * ptR - pnL * ptL;
    }
    const T d2 = (fma(pnR, pnR, pr.hg * (R.q[0] * R.q[0])) - fma(pnL, pnL, pr.hg * (L.q[0] * L.q[0]))) +
                 (pr.g * hb) * (R.b - L.b);
    const T s1 = u - a, s3 = u + a;
    f.al[0] = (s3 * d1 - d2) * ha;
    f.al[1] = d3 - v * d1;
    f.al[2] = (d2 - s1 * d1) * ha;
    f.sf[0] = left_frac(s1);
    f.sf[1] = left_frac(u);
    f.sf[2] = left_frac(s3);
    f.u = u; f.v = v; f.a = a; f.ha = ha;
  }
};

}  // namespace claw

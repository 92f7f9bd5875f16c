// Riemann-solver policies for the fused 2D step (gfx950).
//
// Each policy restates one Clawpack Riemann plugin in a compact form suited to a register-resident
// stencil: a wave is kept as its strength alpha_p plus the few Roe averages that define its
// eigenvector, so neighbouring interfaces exchange 6 (shallow water) or 8 (Euler) numbers for the
// limiter instead of num_eqn*num_waves wave components.
//
//   SweP   <- riemann/src/rpn2_shallow_roe_with_efix.f90, rpt2_shallow_roe_with_efix.f90
//   EulerP <- riemann/src/rpn2_euler_4wave.f90, rpt2_euler.f90
// (upstream riemann submodule; empty in the reference tree, .gitmodules:10-12).
//
// All vectors inside a policy are in the interface frame (h, normal, transverse[, E]); DIR = 0 is
// an x-interface (normal = component 1), DIR = 1 a y-interface (normal = component 2).
//
// Entropy fix (Harten-Hyman, as upstream): the transonic tests "u - c > 0 right of wave 1" and
// "u + c < 0 left of the last wave" are evaluated in a division- and sqrt-free squared form; the
// exact speeds are only formed inside the (rare) transonic branch.  The squared test equals the
// upstream test except within rounding of the threshold, where both branches give the same
// fluctuation to within that rounding.
#pragma once
#include <hip/hip_runtime.h>

namespace claw {

template <class T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }

// philim of limiter.f90 with r = dot/wnorm2; ids as pyclaw/limiters/tvd.py (minmod 1,
// superbee 2, van Leer 3, MC 4, beam-warming 5).  wnorm2 == 0 leaves the wave unlimited.
template <class T> __device__ __forceinline__ T philim(T wn, T dot, int meth) {
  if (!(wn != T(0))) return T(1);
  const T r = dot / wn;
  switch (meth) {
    case 1: return tmax(T(0), tmin(T(1), r));
    case 2: return tmax(T(0), tmax(tmin(T(1), T(2) * r), tmin(T(2), r)));
    case 3: { const T ar = r < T(0) ? -r : r; return (r + ar) / (T(1) + ar); }
    case 4: return tmax(T(0), tmin((T(1) + r) * T(0.5), tmin(T(2), T(2) * r)));
    case 5: return r;
    default: return T(1);
  }
}

// ------------------------------------------------------------------------------------------
// Shallow water, Roe + entropy fix.
// ------------------------------------------------------------------------------------------
template <class T> struct SweP {
  static constexpr int NEQ = 3, NW = 3, NC = 7, NL = 6;
  struct Par { T g, sqg; };
  // per-cell: q, sqrt(h), 1/sqrt(h), hu/sqrt(h), hv/sqrt(h)
  struct Cell { T q[3]; T s, rs, P, Q; };
  struct Iface { T al[3]; T u, v, a, ha; T am[3], ap[3]; };

  __device__ __forceinline__ static bool bad(const T q[3]) { return !(q[0] > T(0)); }

  __device__ __forceinline__ static Cell prep(const T q[3], const Par&) {
    Cell c;
    c.q[0] = q[0]; c.q[1] = q[1]; c.q[2] = q[2];
    c.s = sqrt(q[0]);
    c.rs = T(1) / c.s;
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
    const T inv = T(1) / (L.s + R.s);
    const T u = (pnL + pnR) * inv;
    const T v = (ptL + ptR) * inv;
    const T a = sqrt(pr.g * ((L.q[0] + R.q[0]) * T(0.5)));
    const T ha = T(0.5) / a;
    const T d0 = R.q[0] - L.q[0], dn = R.q[mu] - L.q[mu], dt = R.q[mv] - L.q[mv];
    const T s1 = u - a, s3 = u + a;
    const T a1 = (s3 * d0 - dn) * ha;
    const T a2 = dt - v * d0;
    const T a3 = (dn - s1 * d0) * ha;
    T sf1 = T(0), sf2 = T(0), sf3 = T(0);
    const T s0 = pnL * L.rs - pr.sqg * L.s;  // u - c in the left cell
    if (!(s0 >= T(0) && s1 > T(0))) {
      sf1 = s1 < T(0) ? s1 : T(0);
      if (s0 < T(0)) {
        const T h1 = L.q[0] + a1, hu1 = L.q[mu] + a1 * s1;
        if (h1 > T(0) && hu1 > T(0) && hu1 * hu1 > pr.g * (h1 * h1) * h1) {
          const T s1r = hu1 / h1 - sqrt(pr.g * h1);  // u - c right of the 1-wave
          sf1 = s0 * (s1r - s1) / (s1r - s0);
        }
      }
      if (u < T(0)) {
        sf2 = u;
        sf3 = s3 < T(0) ? s3 : T(0);
        const T s3r = pnR * R.rs + pr.sqg * R.s;  // u + c in the right cell
        const T h3 = R.q[0] - a3, hu3 = R.q[mu] - a3 * s3;
        if (s3r > T(0) && h3 > T(0) && hu3 < T(0) && hu3 * hu3 > pr.g * (h3 * h3) * h3) {
          const T s3l = hu3 / h3 + sqrt(pr.g * h3);  // u + c left of the 3-wave
          sf3 = s3l * (s3r - s3) / (s3r - s3l);
        }
      }
    }
    const T e1 = sf1 * a1, e2 = sf2 * a2, e3 = sf3 * a3;
    const T p1 = (s1 - sf1) * a1, p2 = (u - sf2) * a2, p3 = (s3 - sf3) * a3;
    f.am[0] = e1 + e3; f.am[1] = e1 * s1 + e3 * s3; f.am[2] = (e1 + e3) * v + e2;
    f.ap[0] = p1 + p3; f.ap[1] = p1 * s1 + p3 * s3; f.ap[2] = (p1 + p3) * v + p2;
    f.al[0] = a1; f.al[1] = a2; f.al[2] = a3;
    f.u = u; f.v = v; f.a = a; f.ha = ha;
  }

  __device__ __forceinline__ static T smax(const Iface& f) { return (f.u < T(0) ? -f.u : f.u) + f.a; }
  __device__ __forceinline__ static T speed(const Iface& f, int p) {
    return p == 0 ? f.u - f.a : (p == 1 ? f.u : f.u + f.a);
  }
  // limiter data: alpha1..3, s1, s3, v
  __device__ __forceinline__ static void lim(const Iface& f, T o[NL]) {
    o[0] = f.al[0]; o[1] = f.al[1]; o[2] = f.al[2]; o[3] = f.u - f.a; o[4] = f.u + f.a; o[5] = f.v;
  }
  __device__ __forceinline__ static void dots(const T A[NL], const T B[NL], T d[NW]) {
    const T vv = A[5] * B[5];
    d[0] = (A[0] * B[0]) * ((T(1) + A[3] * B[3]) + vv);
    d[1] = A[1] * B[1];
    d[2] = (A[2] * B[2]) * ((T(1) + A[4] * B[4]) + vv);
  }
  // sum_p coef_p * alpha_p * e_p  (frame vector)
  __device__ __forceinline__ static void combine(const Iface& f, const T coef[NW], T o[NEQ]) {
    const T k1 = coef[0] * f.al[0], k2 = coef[1] * f.al[1], k3 = coef[2] * f.al[2];
    o[0] = k1 + k3;
    o[1] = k1 * (f.u - f.a) + k3 * (f.u + f.a);
    o[2] = (k1 + k3) * f.v + k2;
  }
  __device__ __forceinline__ static void rpt2(const Iface& f, const T as[NEQ], T bm[NEQ], T bp[NEQ]) {
    const T sg1 = f.v - f.a, sg3 = f.v + f.a;
    const T b1 = (sg3 * as[0] - as[2]) * f.ha;
    const T b2 = as[1] - f.u * as[0];
    const T b3 = (as[2] - sg1 * as[0]) * f.ha;
    const T m1 = tmin(sg1, T(0)) * b1, m2 = tmin(f.v, T(0)) * b2, m3 = tmin(sg3, T(0)) * b3;
    const T q1 = tmax(sg1, T(0)) * b1, q2 = tmax(f.v, T(0)) * b2, q3 = tmax(sg3, T(0)) * b3;
    bm[0] = m1 + m3; bm[1] = (m1 + m3) * f.u + m2; bm[2] = m1 * sg1 + m3 * sg3;
    bp[0] = q1 + q3; bp[1] = (q1 + q3) * f.u + q2; bp[2] = q1 * sg1 + q3 * sg3;
  }
};

// ------------------------------------------------------------------------------------------
// Euler, 4 waves (u-c, entropy u, shear u, u+c), Roe + entropy fix on waves 1 and 4.
// ------------------------------------------------------------------------------------------
template <class T> struct EulerP {
  static constexpr int NEQ = 4, NW = 4, NC = 10, NL = 8;
  struct Par { T gamma, gamma1; };
  // per-cell: q, sqrt(rho), 1/sqrt(rho), rho u/sqrt(rho), rho v/sqrt(rho), (E+p)/sqrt(rho), c
  struct Cell { T q[4]; T s, rs, P, Q, Hs, c; };
  struct Iface { T al[4]; T u, v, a, H, u2v2, g1a2, euv, i2a; T am[4], ap[4]; };

  __device__ __forceinline__ static bool bad(const T q[4]) { return !(q[0] > T(0)); }

  __device__ __forceinline__ static Cell prep(const T q[4], const Par& pr) {
    Cell c;
    c.q[0] = q[0]; c.q[1] = q[1]; c.q[2] = q[2]; c.q[3] = q[3];
    c.s = sqrt(q[0]);
    c.rs = T(1) / c.s;
    c.P = q[1] * c.rs;
    c.Q = q[2] * c.rs;
    const T p = pr.gamma1 * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    c.Hs = (q[3] + p) * c.rs;
    c.c = sqrt(pr.gamma * p / q[0]);
    return c;
  }
  __device__ __forceinline__ static void pack(const Cell& c, T o[NC]) {
    o[0] = c.q[0]; o[1] = c.q[1]; o[2] = c.q[2]; o[3] = c.q[3];
    o[4] = c.s; o[5] = c.rs; o[6] = c.P; o[7] = c.Q; o[8] = c.Hs; o[9] = c.c;
  }
  __device__ __forceinline__ static Cell unpack(const T o[NC]) {
    Cell c;
    c.q[0] = o[0]; c.q[1] = o[1]; c.q[2] = o[2]; c.q[3] = o[3];
    c.s = o[4]; c.rs = o[5]; c.P = o[6]; c.Q = o[7]; c.Hs = o[8]; c.c = o[9];
    return c;
  }

  // squared-form test of "n/rho - c > 0" (sign = +1) or "n/rho + c < 0" (sign = -1) for the
  // state (rho, n, t, E); P = p*rho.
  __device__ __forceinline__ static bool sonic_cross(T rho, T n, T t, T E, const Par& pr, bool right) {
    const T Prho = pr.gamma1 * (E * rho - T(0.5) * (n * n + t * t));
    const bool dir = right ? (n > T(0)) : (n < T(0));
    return rho > T(0) && dir && Prho > T(0) && n * n > pr.gamma * Prho;
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
    const T inv = T(1) / (L.s + R.s);
    const T u = (pnL + pnR) * inv;
    const T v = (ptL + ptR) * inv;
    const T H = (L.Hs + R.Hs) * inv;
    const T u2v2 = u * u + v * v;
    const T a2 = pr.gamma1 * (H - T(0.5) * u2v2);
    const T a = sqrt(a2);
    const T g1a2 = pr.gamma1 / a2;
    const T euv = H - u2v2;
    const T i2a = T(1) / (T(2) * a);
    const T d0 = R.q[0] - L.q[0], dn = R.q[mu] - L.q[mu], dt = R.q[mv] - L.q[mv],
            dE = R.q[3] - L.q[3];
    const T ae = g1a2 * (euv * d0 + u * dn + v * dt - dE);
    const T as = dt - v * d0;
    const T a4 = (dn + (a - u) * d0 - a * ae) * i2a;
    const T a1 = d0 - ae - a4;
    const T s1 = u - a, s4 = u + a;
    T sf1 = T(0), sfu = T(0), sf4 = T(0);
    const T s0 = pnL * L.rs - L.c;
    if (!(s0 >= T(0) && s1 > T(0))) {
      sf1 = s1 < T(0) ? s1 : T(0);
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
        sf4 = s4 < T(0) ? s4 : T(0);
        const T s4r = pnR * R.rs + R.c;
        const T r4 = R.q[0] - a4, n4 = R.q[mu] - a4 * s4, t4 = R.q[mv] - a4 * v,
                E4 = R.q[3] - a4 * (H + u * a);
        if (s4r > T(0) && sonic_cross(r4, n4, t4, E4, pr, false)) {
          const T s4l = sonic_speed(r4, n4, t4, E4, pr, T(1));
          sf4 = s4l * (s4r - s4) / (s4r - s4l);
        }
      }
    }
    f.al[0] = a1; f.al[1] = ae; f.al[2] = as; f.al[3] = a4;
    f.u = u; f.v = v; f.a = a; f.H = H; f.u2v2 = u2v2; f.g1a2 = g1a2; f.euv = euv; f.i2a = i2a;
    T cm[4] = {sf1, sfu, sfu, sf4};
    T cp[4] = {s1 - sf1, u - sfu, u - sfu, s4 - sf4};
    combine(f, cm, f.am);
    combine(f, cp, f.ap);
  }

  __device__ __forceinline__ static T smax(const Iface& f) { return (f.u < T(0) ? -f.u : f.u) + f.a; }
  __device__ __forceinline__ static T speed(const Iface& f, int p) {
    return p == 0 ? f.u - f.a : (p == 3 ? f.u + f.a : f.u);
  }
  // limiter data: alpha1..4, u, v, a, H
  __device__ __forceinline__ static void lim(const Iface& f, T o[NL]) {
    o[0] = f.al[0]; o[1] = f.al[1]; o[2] = f.al[2]; o[3] = f.al[3];
    o[4] = f.u; o[5] = f.v; o[6] = f.a; o[7] = f.H;
  }
  __device__ __forceinline__ static void dots(const T A[NL], const T B[NL], T d[NW]) {
    const T uA = A[4], vA = A[5], aA = A[6], HA = A[7];
    const T uB = B[4], vB = B[5], aB = B[6], HB = B[7];
    const T vv = vA * vB;
    d[0] = (A[0] * B[0]) * (((T(1) + (uA - aA) * (uB - aB)) + vv) + (HA - uA * aA) * (HB - uB * aB));
    d[1] = (A[1] * B[1]) * (((T(1) + uA * uB) + vv) +
                            (T(0.5) * (uA * uA + vA * vA)) * (T(0.5) * (uB * uB + vB * vB)));
    d[2] = (A[2] * B[2]) * (T(1) + vv);
    d[3] = (A[3] * B[3]) * (((T(1) + (uA + aA) * (uB + aB)) + vv) + (HA + uA * aA) * (HB + uB * aB));
  }
  __device__ __forceinline__ static void combine(const Iface& f, const T coef[NW], T o[NEQ]) {
    const T k1 = coef[0] * f.al[0], ke = coef[1] * f.al[1], ks = coef[2] * f.al[2],
            k4 = coef[3] * f.al[3];
    o[0] = k1 + ke + k4;
    o[1] = k1 * (f.u - f.a) + ke * f.u + k4 * (f.u + f.a);
    o[2] = (k1 + ke + k4) * f.v + ks;
    o[3] = k1 * (f.H - f.u * f.a) + ke * (T(0.5) * f.u2v2) + ks * f.v + k4 * (f.H + f.u * f.a);
  }
  __device__ __forceinline__ static void rpt2(const Iface& f, const T as[NEQ], T bm[NEQ], T bp[NEQ]) {
    const T a3 = f.g1a2 * (f.euv * as[0] + f.u * as[1] + f.v * as[2] - as[3]);
    const T a2 = as[1] - f.u * as[0];
    const T a4 = (as[2] + (f.a - f.v) * as[0] - f.a * a3) * f.i2a;
    const T a1 = as[0] - a3 - a4;
    const T sg1 = f.v - f.a, sg4 = f.v + f.a;
    const T m1 = tmin(sg1, T(0)) * a1, me = tmin(f.v, T(0)) * a3, ms = tmin(f.v, T(0)) * a2,
            m4 = tmin(sg4, T(0)) * a4;
    const T q1 = tmax(sg1, T(0)) * a1, qe = tmax(f.v, T(0)) * a3, qs = tmax(f.v, T(0)) * a2,
            q4 = tmax(sg4, T(0)) * a4;
    const T hk = T(0.5) * f.u2v2;
    bm[0] = m1 + me + m4;
    bm[1] = (m1 + me + m4) * f.u + ms;
    bm[2] = m1 * sg1 + me * f.v + m4 * sg4;
    bm[3] = m1 * (f.H - f.v * f.a) + me * hk + ms * f.u + m4 * (f.H + f.v * f.a);
    bp[0] = q1 + qe + q4;
    bp[1] = (q1 + qe + q4) * f.u + qs;
    bp[2] = q1 * sg1 + qe * f.v + q4 * sg4;
    bp[3] = q1 * (f.H - f.v * f.a) + qe * hk + qs * f.u + q4 * (f.H + f.v * f.a);
  }
};

}  // namespace claw

// Reverse mode (adjoint) of the 2D unsplit shallow-water step, FP64, dt fixed.
//
// The forward step (restating upstream step2.f90, flux2.f90, limiter.f90 and
// rpn2/rpt2_shallow_roe_with_efix.f90) updates every cell by a sum of interface contributions.
// An interface k of a sweep line (between cells k-1 = "L" and k = "R") with the neighbouring
// interfaces k-1 and k+1 (the limiter's upwind partners) sends, in the interface frame
// (h, normal momentum, transverse momentum),
//   cell L:  LX + LG1 - LG2,   cell L one line down: -LG1,   cell L one line up: +LG2,
//   cell R:  RX + RG1 - RG2,   cell R one line down: -RG1,   cell R one line up: +RG2.
// Its reverse takes the adjoints (lambda) at those six cells and returns adjoints of the
// interface's own Roe split, of the two neighbour interfaces' limiter data, and of its cells.
//
// The pieces below are split so that a sweep can compute each interface's Roe split once and
// trade neighbour data and their adjoints between interfaces (lanes of a wavefront on the GPU):
//   cell_fwd / cell_rev      per-cell sqrt(h), 1/sqrt(h), hn/sqrt(h), ht/sqrt(h)
//   wave_fwd / wave_rev      Roe averages, speeds and wave strengths of one interface
//   fix_fwd  / fix_rev       Harten-Hyman entropy fix (left-going speed fractions)
//   iface_core               the interface's outputs, reversed; branches (entropy fix,
//                            limiter case, upwind side, transverse speed signs) are taken as the
//                            forward took them at this state
// iface_vjp composes them for one interface with its four input cells (A, L, R, B).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>

namespace claw {
namespace adj {

// Reciprocal: IEEE on the host, hardware estimate + two Newton steps on the device (the adjoint
// only needs accuracy far below its 1e-6 gradient tolerance).
__host__ __device__ inline double arcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
#else
  return 1.0 / x;
#endif
}
__host__ __device__ inline double asqrt(double x) { return sqrt(x); }

struct CellF {
  double h, hn, ht, s, r, P, Q;  // s = sqrt(h), r = 1/s, P = hn*r, Q = ht*r
};
struct CellB {
  double h, hn, ht, s, r, P, Q;  // adjoints of the same quantities
};

__host__ __device__ inline CellF cell_fwd(double h, double hn, double ht) {
  CellF c;
  c.h = h; c.hn = hn; c.ht = ht;
  c.s = asqrt(h);
  c.r = arcp(c.s);
  c.P = hn * c.r;
  c.Q = ht * c.r;
  return c;
}
__host__ __device__ inline void zero(CellB& b) { b.h = b.hn = b.ht = b.s = b.r = b.P = b.Q = 0.0; }
__host__ __device__ inline void add(CellB& a, const CellB& b) {
  a.h += b.h; a.hn += b.hn; a.ht += b.ht; a.s += b.s; a.r += b.r; a.P += b.P; a.Q += b.Q;
}
// (s, r, P, Q) adjoints pushed down to (h, hn, ht); the frame adjoint goes to o[3]
__host__ __device__ inline void cell_rev(const CellF& c, const CellB& b, double o[3]) {
  const double rb = b.r + b.P * c.hn + b.Q * c.ht;
  const double sb = b.s - rb * c.r * c.r;
  o[0] = b.h + sb * 0.5 * c.r;
  o[1] = b.hn + b.P * c.r;
  o[2] = b.ht + b.Q * c.r;
}

struct WaveF {
  double S, inv, u, v, a, ha, d0, dn, dt, s1, s3, a1, a2, a3;
};
struct WaveB {
  double a1, a2, a3, u, v, a, ha, s1, s3;
};
__host__ __device__ inline void zero(WaveB& b) { b.a1 = b.a2 = b.a3 = b.u = b.v = b.a = b.ha = b.s1 = b.s3 = 0.0; }

__host__ __device__ inline WaveF wave_fwd(const CellF& L, const CellF& R, double g) {
  WaveF w;
  w.S = L.s + R.s;
  w.inv = arcp(w.S);
  w.u = (L.P + R.P) * w.inv;
  w.v = (L.Q + R.Q) * w.inv;
  w.a = asqrt(g * ((L.h + R.h) * 0.5));
  w.ha = 0.5 * arcp(w.a);
  w.d0 = R.h - L.h;
  w.dn = R.hn - L.hn;
  w.dt = R.ht - L.ht;
  w.s1 = w.u - w.a;
  w.s3 = w.u + w.a;
  w.a1 = (w.s3 * w.d0 - w.dn) * w.ha;
  w.a2 = w.dt - w.v * w.d0;
  w.a3 = (w.dn - w.s1 * w.d0) * w.ha;
  return w;
}

// Reverse of wave_fwd: adds the cell adjoints into Lb, Rb.
__host__ __device__ inline void wave_rev(const CellF& L, const CellF& R, const WaveF& w, double g, WaveB b,
                                         CellB& Lb, CellB& Rb) {
  double d0b = 0.0, dnb = 0.0, dtb = 0.0;
  // a3 = (dn - s1*d0)*ha
  b.ha += b.a3 * (w.dn - w.s1 * w.d0);
  dnb += b.a3 * w.ha;
  b.s1 -= b.a3 * w.ha * w.d0;
  d0b -= b.a3 * w.ha * w.s1;
  // a2 = dt - v*d0
  dtb += b.a2;
  b.v -= b.a2 * w.d0;
  d0b -= b.a2 * w.v;
  // a1 = (s3*d0 - dn)*ha
  b.ha += b.a1 * (w.s3 * w.d0 - w.dn);
  b.s3 += b.a1 * w.ha * w.d0;
  d0b += b.a1 * w.ha * w.s3;
  dnb -= b.a1 * w.ha;
  // s1 = u - a, s3 = u + a
  b.u += b.s1 + b.s3;
  b.a += b.s3 - b.s1;
  // ha = 0.5/a, a = sqrt(g*(hL + hR)/2): d a / d(gh) = 0.5/a = ha
  b.a -= b.ha * w.ha * arcp(w.a);
  const double ghb = b.a * w.ha;
  Lb.h += ghb * 0.5 * g;
  Rb.h += ghb * 0.5 * g;
  Rb.h += d0b; Lb.h -= d0b;
  Rb.hn += dnb; Lb.hn -= dnb;
  Rb.ht += dtb; Lb.ht -= dtb;
  // u = (PL + PR)*inv, v = (QL + QR)*inv, inv = 1/(sL + sR)
  Lb.P += b.u * w.inv; Rb.P += b.u * w.inv;
  Lb.Q += b.v * w.inv; Rb.Q += b.v * w.inv;
  const double Sb = -(b.u * (L.P + R.P) + b.v * (L.Q + R.Q)) * w.inv * w.inv;
  Lb.s += Sb;
  Rb.s += Sb;
}

// Entropy fix of one interface: left-going speed fractions sf1, sf2, sf3.  The transonic tests
// are the forward's squared forms (a non-positive middle depth fails them).
struct FixF {
  double sf1, sf2, sf3;
  bool t1, t3, neg1, neg3, uneg;
  double s0, h1, hu1, c1, s1r, D1;
  double s3r, h3, hu3, c3, s3l, D3;
};

__host__ __device__ inline FixF fix_fwd(const CellF& L, const CellF& R, const WaveF& w, double g, double sqg) {
  FixF f;
  f.t1 = f.t3 = false;
  f.neg1 = w.s1 < 0.0;
  f.sf1 = f.neg1 ? w.s1 : 0.0;
  f.sf2 = f.sf3 = 0.0;
  f.h1 = f.hu1 = f.c1 = f.s1r = f.D1 = 0.0;
  f.s3r = f.h3 = f.hu3 = f.c3 = f.s3l = f.D3 = 0.0;
  f.s0 = L.P * L.r - sqg * L.s;
  if (f.s0 < 0.0) {
    f.h1 = L.h + w.a1;
    f.hu1 = L.hn + w.a1 * w.s1;
    if (f.h1 > 0.0 && f.hu1 > 0.0 && f.hu1 * f.hu1 > g * (f.h1 * f.h1) * f.h1) {
      f.t1 = true;
      f.c1 = asqrt(g * f.h1);
      f.s1r = f.hu1 * arcp(f.h1) - f.c1;
      f.D1 = f.s1r - f.s0;
      f.sf1 = f.s0 * (f.s1r - w.s1) * arcp(f.D1);
    }
  }
  f.uneg = w.u < 0.0;
  f.neg3 = false;
  if (f.uneg) {
    f.sf2 = w.u;
    f.neg3 = w.s3 < 0.0;
    f.sf3 = f.neg3 ? w.s3 : 0.0;
    f.h3 = R.h - w.a3;
    f.hu3 = R.hn - w.a3 * w.s3;
    if (f.h3 > 0.0 && f.hu3 < 0.0 && f.hu3 * f.hu3 > g * (f.h3 * f.h3) * f.h3) {
      f.s3r = R.P * R.r + sqg * R.s;
      if (f.s3r > 0.0) {
        f.t3 = true;
        f.c3 = asqrt(g * f.h3);
        f.s3l = f.hu3 * arcp(f.h3) + f.c3;
        f.D3 = f.s3r - f.s3l;
        f.sf3 = f.s3l * (f.s3r - w.s3) * arcp(f.D3);
      }
    }
  }
  return f;
}

__host__ __device__ inline void fix_rev(const CellF& L, const CellF& R, const WaveF& w, const FixF& f, double g,
                                        double sqg, double sf1b, double sf2b, double sf3b, WaveB& wb,
                                        CellB& Lb, CellB& Rb) {
  if (f.uneg) {
    if (f.t3) {
      const double iD = arcp(f.D3), iD2 = iD * iD;
      const double s3lb = sf3b * (f.s3r - w.s3) * f.s3r * iD2;
      const double s3rb = sf3b * f.s3l * (w.s3 - f.s3l) * iD2;
      wb.s3 -= sf3b * f.s3l * iD;
      const double ih3 = arcp(f.h3);
      const double hu3b = s3lb * ih3;
      const double h3b = s3lb * (-f.hu3 * ih3 * ih3 + 0.5 * g * arcp(f.c3));
      Rb.h += h3b;
      wb.a3 -= h3b + hu3b * w.s3;
      Rb.hn += hu3b;
      wb.s3 -= hu3b * w.a3;
      Rb.P += s3rb * R.r;
      Rb.r += s3rb * R.P;
      Rb.s += s3rb * sqg;
    } else if (f.neg3) {
      wb.s3 += sf3b;
    }
    wb.u += sf2b;
  }
  if (f.t1) {
    const double iD = arcp(f.D1), iD2 = iD * iD;
    const double s0b = sf1b * (f.s1r - w.s1) * f.s1r * iD2;
    const double s1rb = sf1b * f.s0 * (w.s1 - f.s0) * iD2;
    wb.s1 -= sf1b * f.s0 * iD;
    const double ih1 = arcp(f.h1);
    const double hu1b = s1rb * ih1;
    const double h1b = s1rb * (-f.hu1 * ih1 * ih1 - 0.5 * g * arcp(f.c1));
    Lb.h += h1b;
    wb.a1 += h1b + hu1b * w.s1;
    Lb.hn += hu1b;
    wb.s1 += hu1b * w.a1;
    Lb.P += s0b * L.r;
    Lb.r += s0b * L.P;
    Lb.s -= s0b * sqg;
  } else if (f.neg1) {
    wb.s1 += sf1b;
  }
}

// phi(t/alpha)*alpha for the limiter `meth`, with its partials on the branch the value takes.
__host__ __device__ inline double limited_d(double al, double t, int meth, double& dal, double& dt) {
  dal = dt = 0.0;
  switch (meth) {
    case 1: {  // minmod
      if (al > 0.0) {
        if ((al < t ? al : t) <= 0.0) return 0.0;
        if (al < t) { dal = 1.0; return al; }
        dt = 1.0; return t;
      }
      if ((al > t ? al : t) >= 0.0) return 0.0;
      if (al > t) { dal = 1.0; return al; }
      dt = 1.0; return t;
    }
    case 2: {  // superbee: sign(a)*max(0, min(|a|, 2u), min(2|a|, u)), u = sign(a)*t
      if (al == 0.0) return 0.0;
      const double sgn = al > 0.0 ? 1.0 : -1.0;
      const double A = sgn * al, U = sgn * t;
      double x1, x1a, x1t, x2, x2a, x2t;
      if (A < 2.0 * U) { x1 = A; x1a = 1.0; x1t = 0.0; } else { x1 = 2.0 * U; x1a = 0.0; x1t = 2.0; }
      if (2.0 * A < U) { x2 = 2.0 * A; x2a = 2.0; x2t = 0.0; } else { x2 = U; x2a = 0.0; x2t = 1.0; }
      double m, ma, mt;
      if (x1 > x2) { m = x1; ma = x1a; mt = x1t; } else { m = x2; ma = x2a; mt = x2t; }
      if (m <= 0.0) return 0.0;
      dal = ma; dt = mt;
      return sgn * m;
    }
    case 3: {  // van Leer: 2 t al / (al + t) where t/al > 0, else 0
      if (al == 0.0) return 0.0;
      if (t / al <= 0.0) return 0.0;
      const double iden = arcp(al + t), id2 = iden * iden;
      dal = 2.0 * t * t * id2;
      dt = 2.0 * al * al * id2;
      return 2.0 * t * al * iden;
    }
    case 4: {  // MC: max(0, min((a+t)/2, 2t, 2a)) for a > 0 (mirrored for a <= 0)
      const double c = 0.5 * (al + t), t2 = 2.0 * t, a2 = 2.0 * al;
      double v = c;
      int k = 0;
      if (al > 0.0) {
        if (t2 < v) { v = t2; k = 1; }
        if (a2 < v) { v = a2; k = 2; }
        if (!(v > 0.0)) return 0.0;
      } else {
        if (t2 > v) { v = t2; k = 1; }
        if (a2 > v) { v = a2; k = 2; }
        if (!(v < 0.0)) return 0.0;
      }
      if (k == 0) { dal = 0.5; dt = 0.5; } else if (k == 1) { dt = 2.0; } else { dal = 2.0; }
      return v;
    }
    case 5:  // beam-warming: theta
      if (al == 0.0) return 0.0;
      dt = 1.0;
      return t;
    default:  // unlimited
      dal = 1.0;
      return al;
  }
}

struct Method {
  double g, sqg;
  int order, tw;
  int mthlim[3];
};

// Limiter data an interface shares with its neighbours (also used for its adjoint).
struct LimD {
  double a1, a2, a3, s1, s3, v;
};
__host__ __device__ inline LimD lim_of(const WaveF& w) {
  LimD d;
  d.a1 = w.a1; d.a2 = w.a2; d.a3 = w.a3; d.s1 = w.s1; d.s3 = w.s3; d.v = w.v;
  return d;
}
__host__ __device__ inline void zero(LimD& d) { d.a1 = d.a2 = d.a3 = d.s1 = d.s3 = d.v = 0.0; }
__host__ __device__ inline void add_lim(WaveB& b, const LimD& d) {
  b.a1 += d.a1; b.a2 += d.a2; b.a3 += d.a3; b.s1 += d.s1; b.s3 += d.s3; b.v += d.v;
}

// Reverse of one interface's outputs.  Inputs: its cells L, R, Roe split w, entropy fix fx, the
// neighbours' limiter data nl (interface k-1) and nr (k+1), and the adjoints of the six output
// vectors (frame).  Outputs (accumulated): Wb (own split), NLb / NRb (neighbour limiter data),
// Lb / Rb (cell quantities read by the entropy fix).
__host__ __device__ inline void iface_core(const CellF& L, const CellF& R, const WaveF& w, const FixF& fx,
                                           const LimD& nl, const LimD& nr, const double bLX[3],
                                           const double bLG1[3], const double bLG2[3], const double bRX[3],
                                           const double bRG1[3], const double bRG2[3], double r, double rt,
                                           const Method& M, WaveB& Wb, LimD& NLb, LimD& NRb, CellB& Lb,
                                           CellB& Rb) {
  const bool second = M.order != 1;
  const bool fold = second && M.tw == 2;
  const double al[3] = {w.a1, w.a2, w.a3};
  const double sp[3] = {w.s1, w.u, w.s3};
  const double sf[3] = {fx.sf1, fx.sf2, fx.sf3};
  const double aL[3] = {nl.a1, nl.a2, nl.a3}, aR[3] = {nr.a1, nr.a2, nr.a3};

  // ---- forward values the reverse needs: t_p = alpha_up m_p / n_p, k_p = |s|(1 - |s| r) phi
  double in[3] = {1.0, 1.0, 1.0}, ml[3] = {1.0, 1.0, 1.0}, mr[3] = {1.0, 1.0, 1.0};
  double U[3] = {0, 0, 0}, phi[3] = {0, 0, 0}, dpa[3] = {0, 0, 0}, dpt[3] = {0, 0, 0};
  double coef[3] = {0, 0, 0}, k[3] = {0, 0, 0};
  bool upl[3] = {false, false, false};
  if (second) {
    const double vv1 = 1.0 + w.v * w.v;
    in[0] = arcp(w.s1 * w.s1 + vv1);
    in[2] = arcp(w.s3 * w.s3 + vv1);
    ml[0] = 1.0 + nl.s1 * w.s1 + nl.v * w.v;
    ml[2] = 1.0 + nl.s3 * w.s3 + nl.v * w.v;
    mr[0] = 1.0 + w.s1 * nr.s1 + w.v * nr.v;
    mr[2] = 1.0 + w.s3 * nr.s3 + w.v * nr.v;
    for (int p = 0; p < 3; ++p) {
      upl[p] = sp[p] > 0.0;
      U[p] = upl[p] ? aL[p] * ml[p] : aR[p] * mr[p];
      phi[p] = limited_d(al[p], U[p] * in[p], M.mthlim[p], dpa[p], dpt[p]);
      const double sa = sp[p] < 0.0 ? -sp[p] : sp[p];
      coef[p] = sa * (1.0 - sa * r);
      k[p] = coef[p] * phi[p];
    }
  }
  double e[3], qq[3], Av[3], Bv[3];
  for (int p = 0; p < 3; ++p) {
    e[p] = sf[p] * al[p];
    qq[p] = (sp[p] - sf[p]) * al[p];
    Av[p] = fold ? e[p] + k[p] : e[p];
    Bv[p] = fold ? qq[p] - k[p] : qq[p];
  }

  // ---- LX = expand_n(XL), RX = expand_n(XR) with expand_n(c) = (c0 + c2, c0 s1 + c2 s3,
  //      (c0 + c2) v + c1), XL = -r (e + k/2), XR = -r (qq - k/2)
  double XLb[3], XRb[3];
  {
    const double XL0 = -r * (e[0] + 0.5 * k[0]), XL2 = -r * (e[2] + 0.5 * k[2]);
    const double XR0 = -r * (qq[0] - 0.5 * k[0]), XR2 = -r * (qq[2] - 0.5 * k[2]);
    double o0b = bLX[0] + bLX[2] * w.v;
    Wb.v += bLX[2] * (XL0 + XL2);
    XLb[1] = bLX[2];
    XLb[0] = o0b + bLX[1] * w.s1;
    XLb[2] = o0b + bLX[1] * w.s3;
    Wb.s1 += bLX[1] * XL0;
    Wb.s3 += bLX[1] * XL2;
    o0b = bRX[0] + bRX[2] * w.v;
    Wb.v += bRX[2] * (XR0 + XR2);
    XRb[1] = bRX[2];
    XRb[0] = o0b + bRX[1] * w.s1;
    XRb[2] = o0b + bRX[1] * w.s3;
    Wb.s1 += bRX[1] * XR0;
    Wb.s3 += bRX[1] * XR2;
  }
  // ---- transverse outputs G = expand_t(wgt * tbeta(V)) with expand_t(c) = (c0 + c2,
  //      (c0 + c2) u + c1, c0 sg0 + c2 sg2), sg = (v - a, v, v + a), wgt = wt min/max(sg, 0),
  //      tbeta(V) = (m - d, a (V2 - V0), m + d), m = (V0 + V2)/2, d = V1 ha
  double Avb[3] = {0, 0, 0}, Bvb[3] = {0, 0, 0};
  if (M.tw > 0) {
    const double sg[3] = {w.v - w.a, w.v, w.v + w.a};
    const double wt = -0.5 * r * rt;
    double wm[3], wp[3];
    for (int j = 0; j < 3; ++j) {
      wm[j] = sg[j] < 0.0 ? wt * sg[j] : 0.0;
      wp[j] = sg[j] > 0.0 ? wt * sg[j] : 0.0;
    }
    double bA[3], bB[3];
    {
      const double m = 0.5 * (Av[0] + Av[2]), d = Av[1] * w.ha;
      bA[0] = m - d; bA[1] = w.a * (Av[2] - Av[0]); bA[2] = m + d;
    }
    {
      const double m = 0.5 * (Bv[0] + Bv[2]), d = Bv[1] * w.ha;
      bB[0] = m - d; bB[1] = w.a * (Bv[2] - Bv[0]); bB[2] = m + d;
    }
    double wmb[3] = {0, 0, 0}, wpb[3] = {0, 0, 0}, bAb[3] = {0, 0, 0}, bBb[3] = {0, 0, 0};
    double sgb[3] = {0, 0, 0};
    for (int s = 0; s < 4; ++s) {
      const double* ob = s == 0 ? bLG1 : (s == 1 ? bLG2 : (s == 2 ? bRG1 : bRG2));
      const bool minus = (s == 0 || s == 2);
      const double* wgt = minus ? wm : wp;
      const double* bet = s < 2 ? bA : bB;
      const double c0 = wgt[0] * bet[0], c2 = wgt[2] * bet[2];
      const double o0b = ob[0] + ob[1] * w.u;
      Wb.u += ob[1] * (c0 + c2);
      const double cb[3] = {o0b + ob[2] * sg[0], ob[1], o0b + ob[2] * sg[2]};
      sgb[0] += ob[2] * c0;
      sgb[2] += ob[2] * c2;
      for (int j = 0; j < 3; ++j) {
        if (minus) wmb[j] += cb[j] * bet[j]; else wpb[j] += cb[j] * bet[j];
        if (s < 2) bAb[j] += cb[j] * wgt[j]; else bBb[j] += cb[j] * wgt[j];
      }
    }
    for (int j = 0; j < 3; ++j) {
      if (sg[j] < 0.0) sgb[j] += wmb[j] * wt;
      if (sg[j] > 0.0) sgb[j] += wpb[j] * wt;
    }
    Wb.v += sgb[0] + sgb[1] + sgb[2];
    Wb.a += sgb[2] - sgb[0];
    for (int s = 0; s < 2; ++s) {
      const double* bb = s == 0 ? bAb : bBb;
      const double* V = s == 0 ? Av : Bv;
      double* Vb = s == 0 ? Avb : Bvb;
      const double mb = bb[0] + bb[2], db = bb[2] - bb[0];
      Wb.a += bb[1] * (V[2] - V[0]);
      Vb[2] += bb[1] * w.a + 0.5 * mb;
      Vb[0] += -bb[1] * w.a + 0.5 * mb;
      Vb[1] += db * w.ha;
      Wb.ha += db * V[1];
    }
  }
  // ---- e, qq, k
  double eb[3], qqb[3], kb[3];
  for (int p = 0; p < 3; ++p) {
    eb[p] = -r * XLb[p] + Avb[p];
    qqb[p] = -r * XRb[p] + Bvb[p];
    kb[p] = 0.5 * r * (XRb[p] - XLb[p]) + (fold ? Avb[p] - Bvb[p] : 0.0);
  }
  // ---- limiter
  double alb[3] = {0, 0, 0}, spb[3] = {0, 0, 0};
  if (second) {
    double inb[3] = {0, 0, 0}, mlb[3] = {0, 0, 0}, mrb[3] = {0, 0, 0}, aLb[3] = {0, 0, 0}, aRb[3] = {0, 0, 0};
    for (int p = 0; p < 3; ++p) {
      const double phib = kb[p] * coef[p];
      const double sa = sp[p] < 0.0 ? -sp[p] : sp[p];
      const double sab = kb[p] * phi[p] * (1.0 - 2.0 * r * sa);
      spb[p] += sp[p] < 0.0 ? -sab : sab;
      alb[p] += phib * dpa[p];
      const double tb = phib * dpt[p];
      const double Ub = tb * in[p];
      inb[p] += tb * U[p];
      if (upl[p]) {
        mlb[p] += Ub * aL[p];
        aLb[p] += Ub * ml[p];
      } else {
        mrb[p] += Ub * aR[p];
        aRb[p] += Ub * mr[p];
      }
    }
    NLb.a1 += aLb[0]; NLb.a2 += aLb[1]; NLb.a3 += aLb[2];
    NRb.a1 += aRb[0]; NRb.a2 += aRb[1]; NRb.a3 += aRb[2];
    // in = 1/(s^2 + 1 + v^2) for waves 1, 3
    const double n1b = -inb[0] * in[0] * in[0], n3b = -inb[2] * in[2] * in[2];
    Wb.s1 += 2.0 * w.s1 * n1b;
    Wb.s3 += 2.0 * w.s3 * n3b;
    Wb.v += 2.0 * w.v * (n1b + n3b);
    // ml_p = 1 + sl_p s_p + vl v ; mr_p = 1 + s_p sr_p + v vr  (p = 1, 3)
    NLb.s1 += mlb[0] * w.s1; Wb.s1 += mlb[0] * nl.s1;
    NLb.s3 += mlb[2] * w.s3; Wb.s3 += mlb[2] * nl.s3;
    NLb.v += (mlb[0] + mlb[2]) * w.v; Wb.v += (mlb[0] + mlb[2]) * nl.v;
    NRb.s1 += mrb[0] * w.s1; Wb.s1 += mrb[0] * nr.s1;
    NRb.s3 += mrb[2] * w.s3; Wb.s3 += mrb[2] * nr.s3;
    NRb.v += (mrb[0] + mrb[2]) * w.v; Wb.v += (mrb[0] + mrb[2]) * nr.v;
  }
  // ---- e = sf al, qq = (s - sf) al
  double sfb[3];
  for (int p = 0; p < 3; ++p) {
    sfb[p] = (eb[p] - qqb[p]) * al[p];
    alb[p] += eb[p] * sf[p] + qqb[p] * (sp[p] - sf[p]);
    spb[p] += qqb[p] * al[p];
  }
  fix_rev(L, R, w, fx, M.g, M.sqg, sfb[0], sfb[1], sfb[2], Wb, Lb, Rb);
  Wb.s1 += spb[0];
  Wb.u += spb[1];
  Wb.s3 += spb[2];
  Wb.a1 += alb[0];
  Wb.a2 += alb[1];
  Wb.a3 += alb[2];
}

// One interface with its four input cells A = k-2, L = k-1, R = k, B = k+1 (frame vectors):
// accumulates the adjoints of the four cells (frame) into gA..gB.
__host__ __device__ inline void iface_vjp(const double qA[3], const double qL[3], const double qR[3],
                                          const double qB[3], const double bLX[3], const double bLG1[3],
                                          const double bLG2[3], const double bRX[3], const double bRG1[3],
                                          const double bRG2[3], double r, double rt, const Method& M,
                                          double gA[3], double gL[3], double gR[3], double gB[3]) {
  const CellF A = cell_fwd(qA[0], qA[1], qA[2]);
  const CellF L = cell_fwd(qL[0], qL[1], qL[2]);
  const CellF R = cell_fwd(qR[0], qR[1], qR[2]);
  const CellF B = cell_fwd(qB[0], qB[1], qB[2]);
  const WaveF w = wave_fwd(L, R, M.g);
  const WaveF wl = wave_fwd(A, L, M.g);
  const WaveF wr = wave_fwd(R, B, M.g);
  const FixF fx = fix_fwd(L, R, w, M.g, M.sqg);
  WaveB Wb, WLb, WRb;
  zero(Wb); zero(WLb); zero(WRb);
  LimD NLb, NRb;
  zero(NLb); zero(NRb);
  CellB Ab, Lb, Rb, Bb;
  zero(Ab); zero(Lb); zero(Rb); zero(Bb);
  iface_core(L, R, w, fx, lim_of(wl), lim_of(wr), bLX, bLG1, bLG2, bRX, bRG1, bRG2, r, rt, M, Wb, NLb, NRb,
             Lb, Rb);
  add_lim(WLb, NLb);
  add_lim(WRb, NRb);
  wave_rev(L, R, w, M.g, Wb, Lb, Rb);
  wave_rev(A, L, wl, M.g, WLb, Ab, Lb);
  wave_rev(R, B, wr, M.g, WRb, Rb, Bb);
  double o[3];
  cell_rev(A, Ab, o); gA[0] += o[0]; gA[1] += o[1]; gA[2] += o[2];
  cell_rev(L, Lb, o); gL[0] += o[0]; gL[1] += o[1]; gL[2] += o[2];
  cell_rev(R, Rb, o); gR[0] += o[0]; gR[1] += o[1]; gR[2] += o[2];
  cell_rev(B, Bb, o); gB[0] += o[0]; gB[1] += o[1]; gB[2] += o[2];
}

// Ghost-cell source index along one axis (n cells; BC codes 1 extrap, 2 periodic, 3 wall) with
// the wall's sign flip of the normal momentum.  Clamped so that far-out indices stay in bounds.
__host__ __device__ inline int bc_src(int i, int n, int lo, int hi, bool& flip) {
  flip = false;
  int s = i;
  if (i < 0) {
    if (lo == 2) s = i + n;
    else if (lo == 3) { s = -1 - i; flip = true; }
    else s = 0;
  } else if (i >= n) {
    if (hi == 2) s = i - n;
    else if (hi == 3) { s = 2 * n - 1 - i; flip = true; }
    else s = n - 1;
  }
  return s < 0 ? 0 : (s > n - 1 ? n - 1 : s);
}

}  // namespace adj
}  // namespace claw

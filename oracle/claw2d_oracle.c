/*
 * CPU FP64 oracle for the 2D unsplit wave-propagation step (test infrastructure).
 *
 * This file is the checker, never the product: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  It restates, slice by slice and in the same loop
 * order, the classic Clawpack 2D step:
 *
 *   step2  -> flux2 -> rpn2 / limiter / rpt2      (pyclaw/src/pyclaw/classic/{step2,flux2,limiter}.f90)
 *   rpn2/rpt2 shallow_roe_with_efix              (riemann/src/rpn2_shallow_roe_with_efix.f90,
 *                                                 riemann/src/rpt2_shallow_roe_with_efix.f90)
 *   rpn2 euler_4wave / rpt2 euler                (riemann/src/rpn2_euler_4wave.f90, rpt2_euler.f90)
 *   rpn2/rpt2 shallow_bathymetry_fwave           (riemann/src/rpn2_shallow_bathymetry_fwave.f90,
 *                                                 rpt2_shallow_bathymetry_fwave.f90), with the
 *                                                 f-wave form of flux2 (use_fwave) and aux arrays
 *   ghost fill + dt control                      (pyclaw/src/pyclaw/solver.py: _apply_bcs,
 *                                                 _bc_lower/_bc_upper, evolve_to_time)
 *
 * Those upstream files live in the pyclaw and riemann submodules, which are empty in the
 * reference tree (.gitmodules:1-3 pyclaw, :10-12 riemann; version 5.14.0 at
 * clawpack/__init__.py:27), so no line numbers can be cited; the algorithm is the published
 * one the reference cites (CITATION.md:17-42: LeVeque 1997 JCP 131, FVMHP 2002).  FP64 only,
 * as forced by .f2py_f2cmap:1.  Parity of this restatement with the upstream Fortran is pinned
 * only through invariants and closed forms (tests/test_oracle_*.py), not through reference
 * golden vectors, which do not exist in the reference tree.
 *
 * Array conventions (Fortran order, 1-based cell indices with ghosts, as in step2.f90):
 *   qbc(m, i, j)  m = 1..meqn fastest, i = 1-g..mx+g, j = 1-g..my+g
 *   q1d(m, i)     one slice, i = 1-g..n+g
 * Interface i of a slice lies between cells i-1 and i.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXEQN 4
#define ORC_MAXWAVES 4

enum { ORC_SWE = 0, ORC_EULER = 1, ORC_SWE_BATHY = 2 };
enum { BC_CUSTOM = 0, BC_EXTRAP = 1, BC_PERIODIC = 2, BC_WALL = 3 };

typedef struct {
    int solver;           /* ORC_SWE (shallow_roe_with_efix_2D) or ORC_EULER (euler_4wave_2D) */
    int num_eqn;
    int num_waves;
    int num_ghost;
    int mx, my;
    int order;            /* method(2) */
    int transverse_waves; /* method(3) */
    int mthlim[ORC_MAXWAVES];
    double dx, dy;
    double grav;          /* problem_data['grav']  -> cparam */
    double gamma;         /* problem_data['gamma'] -> cparam */
    int num_aux;          /* method(7); aux(1) is the bathymetry b for ORC_SWE_BATHY */
    int fwave;            /* solver.fwave: the Riemann solver returns f-waves Z (sum Z = flux jump) */
    double dry_tolerance; /* problem_data['dry_tolerance'] (ORC_SWE_BATHY) */
    int mcapa;            /* method(6): 1-based aux index of the capacity kappa, 0 = none
                             (pyclaw: state.index_capa + 1) */
} orc_desc;

typedef struct {
    double dt_initial, dt_max, cfl_max, cfl_desired;
    int dt_variable;
    int max_steps;
    int bc_lower[2], bc_upper[2];
    int reflect_index[2]; /* 0-based q component negated at a wall, per dimension */
    int dimensional_split; /* 1: Godunov splitting (step2ds), 0: unsplit (step2) */
    int aux_bc_lower[2], aux_bc_upper[2]; /* solver.aux_bc_lower/upper (wall mirrors, no sign flip) */
} orc_ctrl;

typedef struct {
    double t, dt, cflmax, dtmin, dtmax, last_cfl;
    long numsteps, rejected, attempts;
    int error; /* 0 ok, 1 max_steps reached, 2 cfl too large with fixed dt */
} orc_status;

/* ------------------------------------------------------------------------------------------
 * Slice work space (flux2 arrays), sized for one slice of length n with g ghosts.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    int cap;
    double *wave, *s, *amdq, *apdq, *cqxx, *bmasdq, *bpasdq, *qadd, *fadd, *gadd, *q1d;
    double *aux1d;  /* aux(maux, slice) */
    double *dtdx1d; /* dt/dx per slice cell: dtdx / kappa with a capacity array, else dtdx */
} orc_work;

static int work_alloc(orc_work *w, int nslice, int meqn, int mwaves, int maux) {
    w->cap = nslice;
    w->aux1d = calloc((size_t)nslice * (maux > 0 ? maux : 1), sizeof(double));
    w->wave = calloc((size_t)nslice * meqn * mwaves, sizeof(double));
    w->s = calloc((size_t)nslice * mwaves, sizeof(double));
    w->amdq = calloc((size_t)nslice * meqn, sizeof(double));
    w->apdq = calloc((size_t)nslice * meqn, sizeof(double));
    w->cqxx = calloc((size_t)nslice * meqn, sizeof(double));
    w->bmasdq = calloc((size_t)nslice * meqn, sizeof(double));
    w->bpasdq = calloc((size_t)nslice * meqn, sizeof(double));
    w->qadd = calloc((size_t)nslice * meqn, sizeof(double));
    w->fadd = calloc((size_t)nslice * meqn, sizeof(double));
    w->gadd = calloc((size_t)nslice * meqn * 2, sizeof(double));
    w->q1d = calloc((size_t)nslice * meqn, sizeof(double));
    w->dtdx1d = calloc((size_t)nslice, sizeof(double));
    return (w->wave && w->s && w->amdq && w->apdq && w->cqxx && w->bmasdq && w->bpasdq &&
            w->qadd && w->fadd && w->gadd && w->q1d && w->aux1d && w->dtdx1d) ? 0 : -1;
}

static void work_free(orc_work *w) {
    free(w->wave); free(w->s); free(w->amdq); free(w->apdq); free(w->cqxx);
    free(w->bmasdq); free(w->bpasdq); free(w->qadd); free(w->fadd); free(w->gadd); free(w->q1d);
    free(w->aux1d); free(w->dtdx1d);
}

/* 1-based slice index -> 0-based storage offset */
#define SL(i) ((i) + g - 1)
#define Q1(m, i) q1d[(m) + meqn * SL(i)]
#define WV(m, mw, i) wave[(m) + meqn * ((mw) + mwaves * SL(i))]
#define SP(mw, i) s[(mw) + mwaves * SL(i)]
#define A1(arr, m, i) arr[(m) + meqn * SL(i)]
#define GADD(m, k, i) gadd[(m) + meqn * ((k) + 2 * SL(i))]
#define DTDX(i) dtdx1d[SL(i)]

/* ------------------------------------------------------------------------------------------
 * rpn2, shallow water, Roe linearisation + Harten-Hyman entropy fix.
 * Follows riemann/src/rpn2_shallow_roe_with_efix.f90 (upstream, absent): Roe averages from
 * the two neighbouring cells, three waves (u-c, shear, u+c), amdq built from the left-going
 * part with the transonic-rarefaction split, apdq = sum(s*wave) - amdq.
 * Interfaces i = 2-g .. n+g.  mu/mv select normal/transverse momentum (ixy = 1: x).
 * ---------------------------------------------------------------------------------------- */
static void rpn2_swe(int ixy, int n, int g, int meqn, int mwaves, const double *q1d,
                     double grav, double *wave, double *s, double *amdq, double *apdq) {
    const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1; /* 0-based components */
    for (int i = 2 - g; i <= n + g; ++i) {
        const double hl = Q1(0, i - 1), hr = Q1(0, i);
        const double hbar = (hl + hr) * 0.5;
        const double hsqrtl = sqrt(hl), hsqrtr = sqrt(hr);
        const double hsq2 = hsqrtl + hsqrtr;
        const double u = (Q1(mu, i - 1) / hsqrtl + Q1(mu, i) / hsqrtr) / hsq2;
        const double v = (Q1(mv, i - 1) / hsqrtl + Q1(mv, i) / hsqrtr) / hsq2;
        const double a = sqrt(grav * hbar);
        const double d0 = Q1(0, i) - Q1(0, i - 1);
        const double d1 = Q1(mu, i) - Q1(mu, i - 1);
        const double d2 = Q1(mv, i) - Q1(mv, i - 1);
        const double a1 = ((u + a) * d0 - d1) * (0.5 / a);
        const double a2 = -v * d0 + d2;
        const double a3 = (-(u - a) * d0 + d1) * (0.5 / a);
        WV(0, 0, i) = a1;   WV(mu, 0, i) = a1 * (u - a); WV(mv, 0, i) = a1 * v;  SP(0, i) = u - a;
        WV(0, 1, i) = 0.0;  WV(mu, 1, i) = 0.0;          WV(mv, 1, i) = a2;      SP(1, i) = u;
        WV(0, 2, i) = a3;   WV(mu, 2, i) = a3 * (u + a); WV(mv, 2, i) = a3 * v;  SP(2, i) = u + a;
    }
    /* entropy fix */
    for (int i = 2 - g; i <= n + g; ++i) {
        double sfract;
        /* 1-wave: u-c in the left cell */
        const double s0 = Q1(mu, i - 1) / Q1(0, i - 1) - sqrt(grav * Q1(0, i - 1));
        if (s0 >= 0.0 && SP(0, i) > 0.0) {
            for (int m = 0; m < meqn; ++m) A1(amdq, m, i) = 0.0;
            goto done;
        }
        {
            const double h1 = Q1(0, i - 1) + WV(0, 0, i);
            const double hu1 = Q1(mu, i - 1) + WV(mu, 0, i);
            const double s1 = hu1 / h1 - sqrt(grav * h1); /* u-c right of the 1-wave */
            if (s0 < 0.0 && s1 > 0.0)
                sfract = s0 * (s1 - SP(0, i)) / (s1 - s0);
            else if (SP(0, i) < 0.0)
                sfract = SP(0, i);
            else
                sfract = 0.0;
            for (int m = 0; m < meqn; ++m) A1(amdq, m, i) = sfract * WV(m, 0, i);
        }
        /* 2-wave */
        if (SP(1, i) >= 0.0) goto done;
        for (int m = 0; m < meqn; ++m) A1(amdq, m, i) += SP(1, i) * WV(m, 1, i);
        {
            /* 3-wave: u+c in the right cell and left of the 3-wave */
            const double s3 = Q1(mu, i) / Q1(0, i) + sqrt(grav * Q1(0, i));
            const double h2 = Q1(0, i) - WV(0, 2, i);
            const double hu2 = Q1(mu, i) - WV(mu, 2, i);
            const double s2 = hu2 / h2 + sqrt(grav * h2);
            if (s2 < 0.0 && s3 > 0.0)
                sfract = s2 * (s3 - SP(2, i)) / (s3 - s2);
            else if (SP(2, i) < 0.0)
                sfract = SP(2, i);
            else
                goto done;
            for (int m = 0; m < meqn; ++m) A1(amdq, m, i) += sfract * WV(m, 2, i);
        }
    done:;
    }
    for (int m = 0; m < meqn; ++m)
        for (int i = 2 - g; i <= n + g; ++i) {
            double df = 0.0;
            for (int mw = 0; mw < mwaves; ++mw) df += SP(mw, i) * WV(m, mw, i);
            A1(apdq, m, i) = df - A1(amdq, m, i);
        }
}

/* ------------------------------------------------------------------------------------------
 * rpt2, shallow water.  Follows riemann/src/rpt2_shallow_roe_with_efix.f90 (upstream, absent):
 * the interface Roe averages of rpn2 are recomputed from cells i-1, i; asdq is split along the
 * transverse eigenvectors (speeds v-c, v, v+c) into bmasdq (down-going) and bpasdq (up-going).
 * ---------------------------------------------------------------------------------------- */
static void rpt2_swe(int ixy, int n, int g, int meqn, const double *q1d, double grav,
                     const double *asdq, double *bmasdq, double *bpasdq) {
    const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
    for (int i = 2 - g; i <= n + g; ++i) {
        const double hl = Q1(0, i - 1), hr = Q1(0, i);
        const double hbar = (hl + hr) * 0.5;
        const double hsqrtl = sqrt(hl), hsqrtr = sqrt(hr);
        const double hsq2 = hsqrtl + hsqrtr;
        const double u = (Q1(mu, i - 1) / hsqrtl + Q1(mu, i) / hsqrtr) / hsq2;
        const double v = (Q1(mv, i - 1) / hsqrtl + Q1(mv, i) / hsqrtr) / hsq2;
        const double a = sqrt(grav * hbar);
        const double as0 = A1(asdq, 0, i), asu = A1(asdq, mu, i), asv = A1(asdq, mv, i);
        const double b1 = (0.5 / a) * ((v + a) * as0 - asv);
        const double b2 = asu - u * as0;
        const double b3 = (0.5 / a) * (-(v - a) * as0 + asv);
        double wb[3][3], sb[3];
        wb[0][0] = b1;  wb[0][mu] = b1 * u;  wb[0][mv] = b1 * (v - a);  sb[0] = v - a;
        wb[1][0] = 0.0; wb[1][mu] = b2;      wb[1][mv] = 0.0;           sb[1] = v;
        wb[2][0] = b3;  wb[2][mu] = b3 * u;  wb[2][mv] = b3 * (v + a);  sb[2] = v + a;
        for (int m = 0; m < meqn; ++m) {
            double bm = 0.0, bp = 0.0;
            for (int mw = 0; mw < 3; ++mw) {
                bm += fmin(sb[mw], 0.0) * wb[mw][m];
                bp += fmax(sb[mw], 0.0) * wb[mw][m];
            }
            A1(bmasdq, m, i) = bm;
            A1(bpasdq, m, i) = bp;
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Shallow water over bathymetry, f-wave form.  Restates riemann/src/rpn2_shallow_bathymetry_
 * fwave.f90 and rpt2_shallow_bathymetry_fwave.f90 (upstream, absent) from the published method
 * (Bale, LeVeque, Mitran, Rossmanith, SIAM J. Sci. Comput. 24 (2002) 955-978, cited at
 * CITATION.md:74-88): the flux jump minus the bathymetry source,
 *   delta = f(q_r) - f(q_l) + (0, g hbar (b_r - b_l), 0),
 * is split along the Roe eigenvectors (1, u-c, v), (0, 0, 1), (1, u+c, v) into f-waves Z_p with
 * Roe speeds u-c, u, u+c; amdq/apdq take the f-waves of negative/positive speed (half of a
 * zero-speed one each).  A cell with h <= dry_tolerance has zero velocity; an interface with
 * both cells dry has no waves.  b = aux(1).  For a flat bottom and no entropy-fix branch the
 * f-waves are s_p times the Roe waves of rpn2_swe (the Roe property), a lake at rest
 * (u = v = 0, h + b constant) gives delta = 0.
 * ---------------------------------------------------------------------------------------- */
/* Roe averages of a bathymetry interface; 0 if both cells are dry. */
static int roe_bathy(double hl, double hul, double hvl, double hr, double hur, double hvr,
                     double grav, double tol, double *ul, double *vl, double *ur, double *vr,
                     double *uhat, double *vhat, double *chat) {
    if (!(hl > tol) && !(hr > tol)) return 0;
    *ul = hl > tol ? hul / hl : 0.0;
    *vl = hl > tol ? hvl / hl : 0.0;
    *ur = hr > tol ? hur / hr : 0.0;
    *vr = hr > tol ? hvr / hr : 0.0;
    const double sql = sqrt(hl), sqr = sqrt(hr);
    *uhat = (sql * *ul + sqr * *ur) / (sql + sqr);
    *vhat = (sql * *vl + sqr * *vr) / (sql + sqr);
    *chat = sqrt(grav * (0.5 * (hl + hr)));
    return 1;
}

static void rpn2_bathy(int ixy, int n, int g, int meqn, int mwaves, const double *q1d,
                       const double *aux1d, int maux, double grav, double tol, double *wave,
                       double *s, double *amdq, double *apdq) {
    const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
    for (int i = 2 - g; i <= n + g; ++i) {
        for (int m = 0; m < meqn; ++m) {
            A1(amdq, m, i) = 0.0;
            A1(apdq, m, i) = 0.0;
            for (int mw = 0; mw < mwaves; ++mw) WV(m, mw, i) = 0.0;
        }
        for (int mw = 0; mw < mwaves; ++mw) SP(mw, i) = 0.0;
        const double hl = Q1(0, i - 1), hr = Q1(0, i);
        double ul, vl, ur, vr, u, v, c;
        if (!roe_bathy(hl, Q1(mu, i - 1), Q1(mv, i - 1), hr, Q1(mu, i), Q1(mv, i), grav, tol,
                       &ul, &vl, &ur, &vr, &u, &v, &c))
            continue;
        const double bl = aux1d[maux * SL(i - 1)], br = aux1d[maux * SL(i)];
        const double d1 = hr * ur - hl * ul;
        const double d2 = (hr * ur * ur + 0.5 * grav * hr * hr) - (hl * ul * ul + 0.5 * grav * hl * hl)
                          + grav * (0.5 * (hl + hr)) * (br - bl);
        const double d3 = hr * ur * vr - hl * ul * vl;
        const double s1 = u - c, s3 = u + c;
        const double b1 = (s3 * d1 - d2) / (s3 - s1);
        const double b3 = (d2 - s1 * d1) / (s3 - s1);
        const double b2 = d3 - v * d1;
        WV(0, 0, i) = b1;  WV(mu, 0, i) = b1 * s1; WV(mv, 0, i) = b1 * v;  SP(0, i) = s1;
        WV(0, 1, i) = 0.0; WV(mu, 1, i) = 0.0;     WV(mv, 1, i) = b2;      SP(1, i) = u;
        WV(0, 2, i) = b3;  WV(mu, 2, i) = b3 * s3; WV(mv, 2, i) = b3 * v;  SP(2, i) = s3;
        for (int mw = 0; mw < mwaves; ++mw)
            for (int m = 0; m < meqn; ++m) {
                if (SP(mw, i) < 0.0) {
                    A1(amdq, m, i) += WV(m, mw, i);
                } else if (SP(mw, i) > 0.0) {
                    A1(apdq, m, i) += WV(m, mw, i);
                } else {
                    A1(amdq, m, i) += 0.5 * WV(m, mw, i);
                    A1(apdq, m, i) += 0.5 * WV(m, mw, i);
                }
            }
    }
}

/* Transverse split of asdq at a bathymetry interface: the rpt2_swe eigen-split with the Roe
 * averages of roe_bathy (zero velocity in dry cells); nothing moves across two dry cells. */
static void rpt2_bathy(int ixy, int n, int g, int meqn, const double *q1d, double grav, double tol,
                       const double *asdq, double *bmasdq, double *bpasdq) {
    const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
    for (int i = 2 - g; i <= n + g; ++i) {
        double ul, vl, ur, vr, u, v, a;
        if (!roe_bathy(Q1(0, i - 1), Q1(mu, i - 1), Q1(mv, i - 1), Q1(0, i), Q1(mu, i), Q1(mv, i),
                       grav, tol, &ul, &vl, &ur, &vr, &u, &v, &a)) {
            for (int m = 0; m < meqn; ++m) A1(bmasdq, m, i) = A1(bpasdq, m, i) = 0.0;
            continue;
        }
        const double as0 = A1(asdq, 0, i), asu = A1(asdq, mu, i), asv = A1(asdq, mv, i);
        const double b1 = (0.5 / a) * ((v + a) * as0 - asv);
        const double b2 = asu - u * as0;
        const double b3 = (0.5 / a) * (-(v - a) * as0 + asv);
        double wb[3][3], sb[3];
        wb[0][0] = b1;  wb[0][mu] = b1 * u;  wb[0][mv] = b1 * (v - a);  sb[0] = v - a;
        wb[1][0] = 0.0; wb[1][mu] = b2;      wb[1][mv] = 0.0;           sb[1] = v;
        wb[2][0] = b3;  wb[2][mu] = b3 * u;  wb[2][mv] = b3 * (v + a);  sb[2] = v + a;
        for (int m = 0; m < meqn; ++m) {
            double bm = 0.0, bp = 0.0;
            for (int mw = 0; mw < 3; ++mw) {
                bm += fmin(sb[mw], 0.0) * wb[mw][m];
                bp += fmax(sb[mw], 0.0) * wb[mw][m];
            }
            A1(bmasdq, m, i) = bm;
            A1(bpasdq, m, i) = bp;
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * rpn2, Euler, 4 waves (u-c, entropy u, shear u, u+c) with the Harten-Hyman fix on waves 1 and
 * 4.  Follows riemann/src/rpn2_euler_4wave.f90 (upstream, absent).  q = (rho, rho u, rho v, E).
 * ---------------------------------------------------------------------------------------- */
typedef struct { double u, v, enth, a, g1a2, euv, u2v2; } roe_euler;

static roe_euler roe_avg_euler(const double *ql, const double *qr, int mu, int mv, double gamma1) {
    roe_euler r;
    const double rhsqrtl = sqrt(ql[0]);
    const double rhsqrtr = sqrt(qr[0]);
    const double pl = gamma1 * (ql[3] - 0.5 * (ql[1] * ql[1] + ql[2] * ql[2]) / ql[0]);
    const double pr = gamma1 * (qr[3] - 0.5 * (qr[1] * qr[1] + qr[2] * qr[2]) / qr[0]);
    const double rhsq2 = rhsqrtl + rhsqrtr;
    r.u = (ql[mu] / rhsqrtl + qr[mu] / rhsqrtr) / rhsq2;
    r.v = (ql[mv] / rhsqrtl + qr[mv] / rhsqrtr) / rhsq2;
    r.enth = ((ql[3] + pl) / rhsqrtl + (qr[3] + pr) / rhsqrtr) / rhsq2;
    r.u2v2 = r.u * r.u + r.v * r.v;
    const double a2 = gamma1 * (r.enth - 0.5 * r.u2v2);
    r.a = sqrt(a2);
    r.g1a2 = gamma1 / a2;
    r.euv = r.enth - r.u2v2;
    return r;
}

static void rpn2_euler(int ixy, int n, int g, int meqn, int mwaves, const double *q1d,
                       double gamma, double *wave, double *s, double *amdq, double *apdq) {
    const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
    const double gamma1 = gamma - 1.0;
    for (int i = 2 - g; i <= n + g; ++i) {
        const double *ql = &Q1(0, i - 1), *qr = &Q1(0, i);
        const roe_euler r = roe_avg_euler(ql, qr, mu, mv, gamma1);
        const double d0 = qr[0] - ql[0], d1 = qr[mu] - ql[mu], d2 = qr[mv] - ql[mv],
                     d3 = qr[3] - ql[3];
        const double a3 = r.g1a2 * (r.euv * d0 + r.u * d1 + r.v * d2 - d3);
        const double a2 = d2 - r.v * d0;
        const double a4 = (d1 + (r.a - r.u) * d0 - r.a * a3) / (2.0 * r.a);
        const double a1 = d0 - a3 - a4;
        WV(0, 0, i) = a1; WV(mu, 0, i) = a1 * (r.u - r.a); WV(mv, 0, i) = a1 * r.v;
        WV(3, 0, i) = a1 * (r.enth - r.u * r.a);           SP(0, i) = r.u - r.a;
        WV(0, 1, i) = a3; WV(mu, 1, i) = a3 * r.u; WV(mv, 1, i) = a3 * r.v;
        WV(3, 1, i) = a3 * 0.5 * r.u2v2;                   SP(1, i) = r.u;
        WV(0, 2, i) = 0.0; WV(mu, 2, i) = 0.0; WV(mv, 2, i) = a2;
        WV(3, 2, i) = a2 * r.v;                            SP(2, i) = r.u;
        WV(0, 3, i) = a4; WV(mu, 3, i) = a4 * (r.u + r.a); WV(mv, 3, i) = a4 * r.v;
        WV(3, 3, i) = a4 * (r.enth + r.u * r.a);           SP(3, i) = r.u + r.a;
    }
    for (int i = 2 - g; i <= n + g; ++i) {
        double sfract;
        const double *ql = &Q1(0, i - 1), *qr = &Q1(0, i);
        /* 1-wave */
        const double rhoim1 = ql[0];
        const double pim1 = gamma1 * (ql[3] - 0.5 * (ql[mu] * ql[mu] + ql[mv] * ql[mv]) / rhoim1);
        const double cim1 = sqrt(gamma * pim1 / rhoim1);
        const double s0 = ql[mu] / rhoim1 - cim1;
        if (s0 >= 0.0 && SP(0, i) > 0.0) {
            for (int m = 0; m < meqn; ++m) A1(amdq, m, i) = 0.0;
            goto done;
        }
        {
            const double rho1 = ql[0] + WV(0, 0, i);
            const double rhou1 = ql[mu] + WV(mu, 0, i);
            const double rhov1 = ql[mv] + WV(mv, 0, i);
            const double en1 = ql[3] + WV(3, 0, i);
            const double p1 = gamma1 * (en1 - 0.5 * (rhou1 * rhou1 + rhov1 * rhov1) / rho1);
            const double c1 = sqrt(gamma * p1 / rho1);
            const double s1 = rhou1 / rho1 - c1;
            if (s0 < 0.0 && s1 > 0.0)
                sfract = s0 * (s1 - SP(0, i)) / (s1 - s0);
            else if (SP(0, i) < 0.0)
                sfract = SP(0, i);
            else
                sfract = 0.0;
            for (int m = 0; m < meqn; ++m) A1(amdq, m, i) = sfract * WV(m, 0, i);
        }
        /* contact and shear waves (both move at u) */
        if (SP(1, i) >= 0.0) goto done;
        for (int m = 0; m < meqn; ++m) {
            A1(amdq, m, i) += SP(1, i) * WV(m, 1, i);
            A1(amdq, m, i) += SP(2, i) * WV(m, 2, i);
        }
        {
            /* 4-wave */
            const double rhoi = qr[0];
            const double pi = gamma1 * (qr[3] - 0.5 * (qr[mu] * qr[mu] + qr[mv] * qr[mv]) / rhoi);
            const double ci = sqrt(gamma * pi / rhoi);
            const double s3 = qr[mu] / rhoi + ci;
            const double rho2 = qr[0] - WV(0, 3, i);
            const double rhou2 = qr[mu] - WV(mu, 3, i);
            const double rhov2 = qr[mv] - WV(mv, 3, i);
            const double en2 = qr[3] - WV(3, 3, i);
            const double p2 = gamma1 * (en2 - 0.5 * (rhou2 * rhou2 + rhov2 * rhov2) / rho2);
            const double c2 = sqrt(gamma * p2 / rho2);
            const double s2 = rhou2 / rho2 + c2;
            if (s2 < 0.0 && s3 > 0.0)
                sfract = s2 * (s3 - SP(3, i)) / (s3 - s2);
            else if (SP(3, i) < 0.0)
                sfract = SP(3, i);
            else
                goto done;
            for (int m = 0; m < meqn; ++m) A1(amdq, m, i) += sfract * WV(m, 3, i);
        }
    done:;
    }
    for (int m = 0; m < meqn; ++m)
        for (int i = 2 - g; i <= n + g; ++i) {
            double df = 0.0;
            for (int mw = 0; mw < mwaves; ++mw) df += SP(mw, i) * WV(m, mw, i);
            A1(apdq, m, i) = df - A1(amdq, m, i);
        }
}

/* rpt2, Euler.  Follows riemann/src/rpt2_euler.f90 (upstream, absent). */
static void rpt2_euler(int ixy, int n, int g, int meqn, const double *q1d, double gamma,
                       const double *asdq, double *bmasdq, double *bpasdq) {
    const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
    const double gamma1 = gamma - 1.0;
    for (int i = 2 - g; i <= n + g; ++i) {
        const double *ql = &Q1(0, i - 1), *qr = &Q1(0, i);
        const roe_euler r = roe_avg_euler(ql, qr, mu, mv, gamma1);
        const double *as = &A1(asdq, 0, i);
        const double a3 = r.g1a2 * (r.euv * as[0] + r.u * as[mu] + r.v * as[mv] - as[3]);
        const double a2 = as[mu] - r.u * as[0];
        const double a4 = (as[mv] + (r.a - r.v) * as[0] - r.a * a3) / (2.0 * r.a);
        const double a1 = as[0] - a3 - a4;
        double wb[4][4], sb[4];
        wb[0][0] = a1; wb[0][mu] = a1 * r.u; wb[0][mv] = a1 * (r.v - r.a);
        wb[0][3] = a1 * (r.enth - r.v * r.a);  sb[0] = r.v - r.a;
        wb[1][0] = a3; wb[1][mu] = a3 * r.u; wb[1][mv] = a3 * r.v;
        wb[1][3] = a3 * 0.5 * r.u2v2;          sb[1] = r.v;
        wb[2][0] = 0.0; wb[2][mu] = a2; wb[2][mv] = 0.0;
        wb[2][3] = a2 * r.u;                   sb[2] = r.v;
        wb[3][0] = a4; wb[3][mu] = a4 * r.u; wb[3][mv] = a4 * (r.v + r.a);
        wb[3][3] = a4 * (r.enth + r.v * r.a);  sb[3] = r.v + r.a;
        for (int m = 0; m < meqn; ++m) {
            double bm = 0.0, bp = 0.0;
            for (int mw = 0; mw < 4; ++mw) {
                bm += fmin(sb[mw], 0.0) * wb[mw][m];
                bp += fmax(sb[mw], 0.0) * wb[mw][m];
            }
            A1(bmasdq, m, i) = bm;
            A1(bpasdq, m, i) = bp;
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * philim / limiter.  Follows pyclaw/src/pyclaw/classic/limiter.f90 (upstream, absent) and the
 * limiter ids of pyclaw/limiters/tvd.py (minmod=1, superbee=2, vanleer=3, MC=4, beam-warming=5).
 * theta uses the unlimited neighbour wave upwind of the interface (i-1 if s > 0, else i+1);
 * a wave with zero norm is left alone.
 * ---------------------------------------------------------------------------------------- */
double orc_philim(double a, double b, int meth) {
    const double r = b / a;
    switch (meth) {
    case 1: return fmax(0.0, fmin(1.0, r));
    case 2: return fmax(0.0, fmax(fmin(1.0, 2.0 * r), fmin(2.0, r)));
    case 3: return (r + fabs(r)) / (1.0 + fabs(r));
    case 4: {
        const double c = (1.0 + r) / 2.0;
        return fmax(0.0, fmin(c, fmin(2.0, 2.0 * r)));
    }
    case 5: return r;
    default: return 1.0;
    }
}

static void limiter(int n, int g, int meqn, int mwaves, double *wave, const double *s,
                    const int *mthlim) {
    for (int mw = 0; mw < mwaves; ++mw) {
        if (mthlim[mw] == 0) continue;
        double dotr = 0.0;
        for (int i = 0; i <= n + 1; ++i) {
            double wnorm2 = 0.0, dotl = dotr;
            dotr = 0.0;
            for (int m = 0; m < meqn; ++m) {
                wnorm2 += WV(m, mw, i) * WV(m, mw, i);
                dotr += WV(m, mw, i) * WV(m, mw, i + 1);
            }
            if (i == 0) continue;
            if (wnorm2 == 0.0) continue;
            const double wlim = (SP(mw, i) > 0.0) ? orc_philim(wnorm2, dotl, mthlim[mw])
                                                  : orc_philim(wnorm2, dotr, mthlim[mw]);
            for (int m = 0; m < meqn; ++m) WV(m, mw, i) *= wlim;
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * flux2: one slice.  Follows pyclaw/src/pyclaw/classic/flux2.f90 (upstream, absent):
 * Godunov qadd, CFL, limited second-order correction cqxx (fadd += cqxx/2), cqxx folded into
 * amdq/apdq when transverse_waves == 2, then two transverse splittings into gadd with
 * 0.5*dt/dx weights (imp = 1 from amdq into cell i-1, imp = 2 from apdq into cell i).
 * ---------------------------------------------------------------------------------------- */
static void flux2(const orc_desc *d, int ixy, int n, double dtdx, orc_work *w, double *cfl1d) {
    const int g = d->num_ghost, meqn = d->num_eqn, mwaves = d->num_waves;
    double *q1d = w->q1d, *wave = w->wave, *s = w->s, *amdq = w->amdq, *apdq = w->apdq;
    double *cqxx = w->cqxx, *bmasdq = w->bmasdq, *bpasdq = w->bpasdq;
    double *qadd = w->qadd, *fadd = w->fadd, *gadd = w->gadd, *dtdx1d = w->dtdx1d;
    const int len = n + 2 * g;
    /* step2.f90: dtdx1d = dtdx / aux(mcapa, slice) with a capacity array, else dtdx */
    for (int k = 0; k < len; ++k)
        dtdx1d[k] = d->mcapa > 0 ? dtdx / w->aux1d[(d->mcapa - 1) + d->num_aux * k] : dtdx;
    memset(qadd, 0, sizeof(double) * meqn * len);
    memset(fadd, 0, sizeof(double) * meqn * len);
    memset(gadd, 0, sizeof(double) * meqn * 2 * len);
    memset(cqxx, 0, sizeof(double) * meqn * len);

    if (d->solver == ORC_SWE)
        rpn2_swe(ixy, n, g, meqn, mwaves, q1d, d->grav, wave, s, amdq, apdq);
    else if (d->solver == ORC_SWE_BATHY)
        rpn2_bathy(ixy, n, g, meqn, mwaves, q1d, w->aux1d, d->num_aux, d->grav, d->dry_tolerance,
                   wave, s, amdq, apdq);
    else
        rpn2_euler(ixy, n, g, meqn, mwaves, q1d, d->gamma, wave, s, amdq, apdq);

    for (int i = 1; i <= n + 1; ++i)
        for (int m = 0; m < meqn; ++m) {
            A1(qadd, m, i) -= DTDX(i) * A1(apdq, m, i);
            A1(qadd, m, i - 1) -= DTDX(i - 1) * A1(amdq, m, i);
        }

    /* a right-going wave is measured against the right cell's capacity, a left-going one
     * against the left cell's */
    double c = 0.0;
    for (int i = 1; i <= n + 1; ++i)
        for (int mw = 0; mw < mwaves; ++mw) {
            c = fmax(c, DTDX(i) * SP(mw, i));
            c = fmax(c, -DTDX(i - 1) * SP(mw, i));
        }
    *cfl1d = c;

    if (d->order != 1) {
        int limit = 0;
        for (int mw = 0; mw < mwaves; ++mw)
            if (d->mthlim[mw] > 0) limit = 1;
        if (limit) limiter(n, g, meqn, mwaves, wave, s, d->mthlim);
        for (int i = 1; i <= n + 1; ++i) {
            /* the average of dtdx1d over the two cells, in the correction only (not in the
             * transverse terms) */
            const double dtdxave = 0.5 * (DTDX(i - 1) + DTDX(i));
            for (int m = 0; m < meqn; ++m) {
                double cq = 0.0;
                for (int mw = 0; mw < mwaves; ++mw) {
                    /* f-waves already carry the speed: sign(s) in place of |s| (dsign(1, s)) */
                    const double abs_sign = d->fwave ? copysign(1.0, SP(mw, i)) : fabs(SP(mw, i));
                    cq += abs_sign * (1.0 - fabs(SP(mw, i)) * dtdxave) * WV(m, mw, i);
                }
                A1(cqxx, m, i) = cq;
                A1(fadd, m, i) += 0.5 * cq;
            }
        }
    }
    if (d->transverse_waves == 0) return;
    if (d->order != 1 && d->transverse_waves == 2)
        for (int i = 1; i <= n + 1; ++i)
            for (int m = 0; m < meqn; ++m) {
                A1(amdq, m, i) += A1(cqxx, m, i);
                A1(apdq, m, i) -= A1(cqxx, m, i);
            }

    if (d->solver == ORC_SWE)
        rpt2_swe(ixy, n, g, meqn, q1d, d->grav, amdq, bmasdq, bpasdq);
    else if (d->solver == ORC_SWE_BATHY)
        rpt2_bathy(ixy, n, g, meqn, q1d, d->grav, d->dry_tolerance, amdq, bmasdq, bpasdq);
    else
        rpt2_euler(ixy, n, g, meqn, q1d, d->gamma, amdq, bmasdq, bpasdq);
    for (int i = 1; i <= n + 1; ++i)
        for (int m = 0; m < meqn; ++m) {
            GADD(m, 0, i - 1) -= 0.5 * DTDX(i - 1) * A1(bmasdq, m, i);
            GADD(m, 1, i - 1) -= 0.5 * DTDX(i - 1) * A1(bpasdq, m, i);
        }
    if (d->solver == ORC_SWE)
        rpt2_swe(ixy, n, g, meqn, q1d, d->grav, apdq, bmasdq, bpasdq);
    else if (d->solver == ORC_SWE_BATHY)
        rpt2_bathy(ixy, n, g, meqn, q1d, d->grav, d->dry_tolerance, apdq, bmasdq, bpasdq);
    else
        rpt2_euler(ixy, n, g, meqn, q1d, d->gamma, apdq, bmasdq, bpasdq);
    for (int i = 1; i <= n + 1; ++i)
        for (int m = 0; m < meqn; ++m) {
            GADD(m, 0, i) -= 0.5 * DTDX(i) * A1(bmasdq, m, i);
            GADD(m, 1, i) -= 0.5 * DTDX(i) * A1(bpasdq, m, i);
        }
}

/* qbc element, 1-based cell indices with ghosts */
#define QB(arr, m, i, j) arr[(m) + meqn * (((i) + g - 1) + (size_t)(mx + 2 * g) * ((j) + g - 1))]
/* auxbc element (maux components), same indexing */
#define AB(arr, m, i, j) arr[(m) + maux * (((i) + g - 1) + (size_t)(mx + 2 * g) * ((j) + g - 1))]

/* ------------------------------------------------------------------------------------------
 * step2 (unsplit).  Follows pyclaw/src/pyclaw/classic/step2.f90 (upstream, absent): x-sweeps
 * over rows j = 0..my+1 then y-sweeps over columns i = 0..mx+1, each slice update applied to
 * qnew immediately (own row plus the transverse gadd into the rows either side).
 * qold must hold filled ghosts; qnew receives qold + update (ghost values are scratch).
 * Returns cfl = max over both sweeps.
 * ---------------------------------------------------------------------------------------- */
/* One x-sweep slice (row j) of step2, accumulated into qnew. */
static void xsweep_row(const orc_desc *d, int j, const double *qold, const double *auxbc,
                       double *qnew, orc_work *wk, double dtdx, double dtdy, double *cfl) {
    const int g = d->num_ghost, meqn = d->num_eqn, mx = d->mx, maux = d->num_aux;
    orc_work w = *wk;
    double cfl1d;
    for (int i = 1 - g; i <= mx + g; ++i) {
        for (int m = 0; m < meqn; ++m) w.q1d[m + meqn * (i + g - 1)] = QB(qold, m, i, j);
        if (auxbc)
            for (int m = 0; m < maux; ++m) w.aux1d[m + maux * (i + g - 1)] = AB(auxbc, m, i, j);
    }
    flux2(d, 1, mx, dtdx, &w, &cfl1d);
    *cfl = fmax(*cfl, cfl1d);
    const double *qadd = w.qadd, *fadd = w.fadd, *gadd = w.gadd;
    if (d->mcapa > 0) {
        /* capacity form: the flux differences of a cell are divided by its kappa */
        const int kc = d->mcapa - 1;
        for (int i = 1; i <= mx; ++i)
            for (int m = 0; m < meqn; ++m) {
                QB(qnew, m, i, j) = QB(qnew, m, i, j) + A1(qadd, m, i)
                                    - (dtdx * (A1(fadd, m, i + 1) - A1(fadd, m, i))
                                       + dtdy * (GADD(m, 1, i) - GADD(m, 0, i))) / AB(auxbc, kc, i, j);
                QB(qnew, m, i, j - 1) = QB(qnew, m, i, j - 1) - dtdy * GADD(m, 0, i) / AB(auxbc, kc, i, j - 1);
                QB(qnew, m, i, j + 1) = QB(qnew, m, i, j + 1) + dtdy * GADD(m, 1, i) / AB(auxbc, kc, i, j + 1);
            }
        return;
    }
    for (int i = 1; i <= mx; ++i)
        for (int m = 0; m < meqn; ++m) {
            QB(qnew, m, i, j) = QB(qnew, m, i, j) + A1(qadd, m, i)
                                - dtdx * (A1(fadd, m, i + 1) - A1(fadd, m, i))
                                - dtdy * (GADD(m, 1, i) - GADD(m, 0, i));
            QB(qnew, m, i, j - 1) = QB(qnew, m, i, j - 1) - dtdy * GADD(m, 0, i);
            QB(qnew, m, i, j + 1) = QB(qnew, m, i, j + 1) + dtdy * GADD(m, 1, i);
        }
}

/* One y-sweep slice (column i) of step2, accumulated into qnew. */
static void ysweep_col(const orc_desc *d, int i, const double *qold, const double *auxbc,
                       double *qnew, orc_work *wk, double dtdx, double dtdy, double *cfl) {
    const int g = d->num_ghost, meqn = d->num_eqn, mx = d->mx, my = d->my, maux = d->num_aux;
    orc_work w = *wk;
    double cfl1d;
    for (int j = 1 - g; j <= my + g; ++j) {
        for (int m = 0; m < meqn; ++m) w.q1d[m + meqn * (j + g - 1)] = QB(qold, m, i, j);
        if (auxbc)
            for (int m = 0; m < maux; ++m) w.aux1d[m + maux * (j + g - 1)] = AB(auxbc, m, i, j);
    }
    flux2(d, 2, my, dtdy, &w, &cfl1d);
    *cfl = fmax(*cfl, cfl1d);
    const double *qadd = w.qadd, *fadd = w.fadd, *gadd = w.gadd;
    if (d->mcapa > 0) {
        const int kc = d->mcapa - 1;
        for (int j = 1; j <= my; ++j)
            for (int m = 0; m < meqn; ++m) {
                QB(qnew, m, i, j) = QB(qnew, m, i, j) + A1(qadd, m, j)
                                    - (dtdy * (A1(fadd, m, j + 1) - A1(fadd, m, j))
                                       + dtdx * (GADD(m, 1, j) - GADD(m, 0, j))) / AB(auxbc, kc, i, j);
                QB(qnew, m, i - 1, j) = QB(qnew, m, i - 1, j) - dtdx * GADD(m, 0, j) / AB(auxbc, kc, i - 1, j);
                QB(qnew, m, i + 1, j) = QB(qnew, m, i + 1, j) + dtdx * GADD(m, 1, j) / AB(auxbc, kc, i + 1, j);
            }
        return;
    }
    for (int j = 1; j <= my; ++j)
        for (int m = 0; m < meqn; ++m) {
            QB(qnew, m, i, j) = QB(qnew, m, i, j) + A1(qadd, m, j)
                                - dtdy * (A1(fadd, m, j + 1) - A1(fadd, m, j))
                                - dtdx * (GADD(m, 1, j) - GADD(m, 0, j));
            QB(qnew, m, i - 1, j) = QB(qnew, m, i - 1, j) - dtdx * GADD(m, 0, j);
            QB(qnew, m, i + 1, j) = QB(qnew, m, i + 1, j) + dtdx * GADD(m, 1, j);
        }
}

/* Threads for orc_step2 (1 = the sequential slice order of step2.f90, the default).  With more
 * threads the slices of one sweep run in three colour classes (slice index mod 3), so no two
 * concurrent slices update the same cells; sums then associate in a different order than the
 * sequential loop (round-off level).  Used only to time a multi-core CPU baseline. */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

/* step2 with aux arrays (auxbc: ghost-filled aux(num_aux, 1-g:mx+g, 1-g:my+g), or NULL when
 * num_aux = 0). */
int orc_step2_aux(const orc_desc *d, const double *qold, const double *auxbc, double *qnew,
                  double dt, double *cfl_out) {
    const int g = d->num_ghost, meqn = d->num_eqn, mx = d->mx, my = d->my;
    const int maxm = mx > my ? mx : my;
    const size_t total = (size_t)meqn * (mx + 2 * g) * (my + 2 * g);
    memcpy(qnew, qold, total * sizeof(double));
    const double dtdx = dt / d->dx, dtdy = dt / d->dy;
    double cfl = 0.0;
    int err = 0;
    if (g_threads <= 1) {
        orc_work w;
        if (work_alloc(&w, maxm + 2 * g, meqn, d->num_waves, d->num_aux) != 0) { work_free(&w); return -1; }
        for (int j = 0; j <= my + 1; ++j) xsweep_row(d, j, qold, auxbc, qnew, &w, dtdx, dtdy, &cfl);
        for (int i = 0; i <= mx + 1; ++i) ysweep_col(d, i, qold, auxbc, qnew, &w, dtdx, dtdy, &cfl);
        work_free(&w);
    } else {
        for (int dir = 0; dir < 2; ++dir)
            for (int colour = 0; colour < 3; ++colour) {
                const int last = dir == 0 ? my + 1 : mx + 1;
#pragma omp parallel num_threads(g_threads) reduction(max : cfl) reduction(| : err)
                {
                    orc_work w;
                    if (work_alloc(&w, maxm + 2 * g, meqn, d->num_waves, d->num_aux) != 0) err = 1;
                    else {
#pragma omp for schedule(static)
                        for (int k = colour; k <= last; k += 3) {
                            if (dir == 0) xsweep_row(d, k, qold, auxbc, qnew, &w, dtdx, dtdy, &cfl);
                            else ysweep_col(d, k, qold, auxbc, qnew, &w, dtdx, dtdy, &cfl);
                        }
                    }
                    work_free(&w);
                }
            }
    }
    *cfl_out = cfl;
    return err ? -1 : 0;
}

int orc_step2(const orc_desc *d, const double *qold, double *qnew, double dt, double *cfl_out) {
    return orc_step2_aux(d, qold, NULL, qnew, dt, cfl_out);
}

/* ------------------------------------------------------------------------------------------
 * Ghost fill.  Follows pyclaw/src/pyclaw/solver.py _apply_bcs/_bc_lower/_bc_upper (upstream,
 * absent): interior copied in, then dimension x over all rows (ghost rows included), then
 * dimension y over all columns (ghost columns included), so corners get the y rule applied to
 * x-filled ghosts.  extrap: copy the nearest interior cell; periodic: wrap; wall: mirror with
 * the normal momentum (reflect_index) negated.
 * ---------------------------------------------------------------------------------------- */
/* Ghost fill of an array with `meqn` components under the given BC codes; rf[k] >= 0 is the
 * component a wall negates in dimension k (q), rf[k] < 0 a plain mirror (aux arrays). */
static void fill_ghosts_impl(const orc_desc *d, int meqn, const int *bclo, const int *bchi,
                             const int *rf, const double *q, double *qbc) {
    const int g = d->num_ghost, mx = d->mx, my = d->my;
    for (int j = 1; j <= my; ++j)
        for (int i = 1; i <= mx; ++i)
            for (int m = 0; m < meqn; ++m)
                QB(qbc, m, i, j) = q[m + meqn * ((i - 1) + (size_t)mx * (j - 1))];
    /* x dimension */
    for (int j = 1 - g; j <= my + g; ++j) {
        for (int k = 0; k < g; ++k) {
            const int il = 1 - g + k, iu = mx + g - k; /* lower ghost, upper ghost */
            for (int m = 0; m < meqn; ++m) {
                switch (bclo[0]) {
                case BC_EXTRAP: QB(qbc, m, il, j) = QB(qbc, m, 1, j); break;
                case BC_PERIODIC: QB(qbc, m, il, j) = QB(qbc, m, il + mx, j); break;
                case BC_WALL: QB(qbc, m, il, j) = QB(qbc, m, 1 - il, j); break;
                default: break;
                }
                switch (bchi[0]) {
                case BC_EXTRAP: QB(qbc, m, iu, j) = QB(qbc, m, mx, j); break;
                case BC_PERIODIC: QB(qbc, m, iu, j) = QB(qbc, m, iu - mx, j); break;
                case BC_WALL: QB(qbc, m, iu, j) = QB(qbc, m, 2 * mx + 1 - iu, j); break;
                default: break;
                }
            }
            if (bclo[0] == BC_WALL && rf[0] >= 0) {
                const int m = rf[0];
                QB(qbc, m, il, j) = -QB(qbc, m, 1 - il, j);
            }
            if (bchi[0] == BC_WALL && rf[0] >= 0) {
                const int m = rf[0];
                QB(qbc, m, iu, j) = -QB(qbc, m, 2 * mx + 1 - iu, j);
            }
        }
    }
    /* y dimension */
    for (int i = 1 - g; i <= mx + g; ++i) {
        for (int k = 0; k < g; ++k) {
            const int jl = 1 - g + k, ju = my + g - k;
            for (int m = 0; m < meqn; ++m) {
                switch (bclo[1]) {
                case BC_EXTRAP: QB(qbc, m, i, jl) = QB(qbc, m, i, 1); break;
                case BC_PERIODIC: QB(qbc, m, i, jl) = QB(qbc, m, i, jl + my); break;
                case BC_WALL: QB(qbc, m, i, jl) = QB(qbc, m, i, 1 - jl); break;
                default: break;
                }
                switch (bchi[1]) {
                case BC_EXTRAP: QB(qbc, m, i, ju) = QB(qbc, m, i, my); break;
                case BC_PERIODIC: QB(qbc, m, i, ju) = QB(qbc, m, i, ju - my); break;
                case BC_WALL: QB(qbc, m, i, ju) = QB(qbc, m, i, 2 * my + 1 - ju); break;
                default: break;
                }
            }
            if (bclo[1] == BC_WALL && rf[1] >= 0) {
                const int m = rf[1];
                QB(qbc, m, i, jl) = -QB(qbc, m, i, 1 - jl);
            }
            if (bchi[1] == BC_WALL && rf[1] >= 0) {
                const int m = rf[1];
                QB(qbc, m, i, ju) = -QB(qbc, m, i, 2 * my + 1 - ju);
            }
        }
    }
}

void orc_fill_ghosts(const orc_desc *d, const orc_ctrl *c, const double *q, double *qbc) {
    fill_ghosts_impl(d, d->num_eqn, c->bc_lower, c->bc_upper, c->reflect_index, q, qbc);
}

/* aux ghosts (pyclaw _apply_bcs on aux with aux_bc_lower/upper: a wall mirrors every
 * component without negating any) */
void orc_fill_aux_ghosts(const orc_desc *d, const orc_ctrl *c, const double *aux, double *auxbc) {
    const int none[2] = {-1, -1};
    fill_ghosts_impl(d, d->num_aux, c->aux_bc_lower, c->aux_bc_upper, none, aux, auxbc);
}

/* One hyperbolic step on interior q (num_eqn, mx, my Fortran order), in place.
 * Follows ClawSolver2D.step_hyperbolic (pyclaw/src/pyclaw/classic/solver.py, upstream,
 * absent): fill ghosts, qold = qbc copy, step2, copy the interior back. */
/* ------------------------------------------------------------------------------------------
 * step2ds: one sweep direction of the dimensionally split step.  Follows
 * pyclaw/src/pyclaw/classic/step2ds.f90 (upstream, absent): flux2 on every interior slice of
 * direction ids (1: rows, x-sweeps; 2: columns, y-sweeps) with transverse propagation off
 * (method(3) < 0), qnew += qadd - dt/dx (fadd(i+1) - fadd(i)).  Returns the sweep's CFL.
 * ---------------------------------------------------------------------------------------- */
int orc_step2ds_aux(const orc_desc *d, const double *qold, const double *auxbc, double *qnew,
                    double dt, int ids, double *cfl_out) {
    const int g = d->num_ghost, meqn = d->num_eqn, mx = d->mx, my = d->my, maux = d->num_aux;
    const int maxm = mx > my ? mx : my;
    orc_desc dd = *d;
    dd.transverse_waves = 0;
    orc_work w;
    if (work_alloc(&w, maxm + 2 * g, meqn, d->num_waves, maux) != 0) { work_free(&w); return -1; }
    const size_t total = (size_t)meqn * (mx + 2 * g) * (my + 2 * g);
    memcpy(qnew, qold, total * sizeof(double));
    double cfl = 0.0, cfl1d;
    if (ids == 1) {
        const double dtdx = dt / d->dx;
        for (int j = 1; j <= my; ++j) {
            for (int i = 1 - g; i <= mx + g; ++i) {
                for (int m = 0; m < meqn; ++m) w.q1d[m + meqn * (i + g - 1)] = QB(qold, m, i, j);
                if (auxbc)
                    for (int m = 0; m < maux; ++m) w.aux1d[m + maux * (i + g - 1)] = AB(auxbc, m, i, j);
            }
            flux2(&dd, 1, mx, dtdx, &w, &cfl1d);
            cfl = fmax(cfl, cfl1d);
            const double *qadd = w.qadd, *fadd = w.fadd;
            for (int i = 1; i <= mx; ++i)
                for (int m = 0; m < meqn; ++m)
                    QB(qnew, m, i, j) = QB(qnew, m, i, j) + A1(qadd, m, i)
                                        - (d->mcapa > 0 ? dtdx * (A1(fadd, m, i + 1) - A1(fadd, m, i))
                                                              / AB(auxbc, d->mcapa - 1, i, j)
                                                        : dtdx * (A1(fadd, m, i + 1) - A1(fadd, m, i)));
        }
    } else {
        const double dtdy = dt / d->dy;
        for (int i = 1; i <= mx; ++i) {
            for (int j = 1 - g; j <= my + g; ++j) {
                for (int m = 0; m < meqn; ++m) w.q1d[m + meqn * (j + g - 1)] = QB(qold, m, i, j);
                if (auxbc)
                    for (int m = 0; m < maux; ++m) w.aux1d[m + maux * (j + g - 1)] = AB(auxbc, m, i, j);
            }
            flux2(&dd, 2, my, dtdy, &w, &cfl1d);
            cfl = fmax(cfl, cfl1d);
            const double *qadd = w.qadd, *fadd = w.fadd;
            for (int j = 1; j <= my; ++j)
                for (int m = 0; m < meqn; ++m)
                    QB(qnew, m, i, j) = QB(qnew, m, i, j) + A1(qadd, m, j)
                                        - (d->mcapa > 0 ? dtdy * (A1(fadd, m, j + 1) - A1(fadd, m, j))
                                                              / AB(auxbc, d->mcapa - 1, i, j)
                                                        : dtdy * (A1(fadd, m, j + 1) - A1(fadd, m, j)));
        }
    }
    *cfl_out = cfl;
    work_free(&w);
    return 0;
}

int orc_step2ds(const orc_desc *d, const double *qold, double *qnew, double dt, int ids,
                double *cfl_out) {
    return orc_step2ds_aux(d, qold, NULL, qnew, dt, ids, cfl_out);
}

static void copy_interior(const orc_desc *d, const double *qbc, double *q) {
    const int g = d->num_ghost, meqn = d->num_eqn, mx = d->mx, my = d->my;
    for (int j = 1; j <= my; ++j)
        for (int i = 1; i <= mx; ++i)
            for (int m = 0; m < meqn; ++m)
                q[m + meqn * ((i - 1) + (size_t)mx * (j - 1))] = QB(qbc, m, i, j);
}

/* ClawSolver2D.step_hyperbolic (classic/solver.py, upstream): ghost fill, then step2, or with
 * dimensional splitting the x-sweeps of step2ds, a fresh ghost fill of the intermediate state and
 * the y-sweeps (Godunov splitting, x first); cfl = max of the two sweeps. */
int orc_step_hyperbolic_aux(const orc_desc *d, const orc_ctrl *c, double *q, const double *aux,
                            double dt, double *cfl) {
    const int g = d->num_ghost, meqn = d->num_eqn, mx = d->mx, my = d->my;
    const size_t nbc = (size_t)meqn * (mx + 2 * g) * (my + 2 * g);
    const size_t nabc = (size_t)(d->num_aux > 0 ? d->num_aux : 0) * (mx + 2 * g) * (my + 2 * g);
    double *qold = malloc(nbc * sizeof(double));
    double *qnew = malloc(nbc * sizeof(double));
    double *auxbc = (aux && nabc) ? malloc(nabc * sizeof(double)) : NULL;
    if (!qold || !qnew || (aux && nabc && !auxbc)) { free(qold); free(qnew); free(auxbc); return -1; }
    if (auxbc) orc_fill_aux_ghosts(d, c, aux, auxbc);
    orc_fill_ghosts(d, c, q, qold);
    int rc;
    if (c->dimensional_split) {
        double cx = 0.0, cy = 0.0;
        rc = orc_step2ds_aux(d, qold, auxbc, qnew, dt, 1, &cx);
        if (rc == 0) {
            copy_interior(d, qnew, q);
            orc_fill_ghosts(d, c, q, qold);
            rc = orc_step2ds_aux(d, qold, auxbc, qnew, dt, 2, &cy);
        }
        *cfl = fmax(cx, cy);
    } else {
        rc = orc_step2_aux(d, qold, auxbc, qnew, dt, cfl);
    }
    if (rc == 0) copy_interior(d, qnew, q);
    free(qold);
    free(qnew);
    free(auxbc);
    return rc;
}

int orc_step_hyperbolic(const orc_desc *d, const orc_ctrl *c, double *q, double dt, double *cfl) {
    return orc_step_hyperbolic_aux(d, c, q, NULL, dt, cfl);
}

/* ------------------------------------------------------------------------------------------
 * Time loop.  Follows Solver.evolve_to_time (pyclaw/src/pyclaw/solver.py, upstream, absent):
 * clip dt to tend, step, accept if cfl <= cfl_max (else restore q and retry when dt_variable),
 * then dt <- min(dt_max, dt*cfl_desired/cfl) (dt_max if cfl == 0).
 * nsteps_target > 0 stops after that many accepted steps (tend = +inf semantics otherwise
 * unchanged); nsteps_target <= 0 runs to tend.  st->dt carries dt in and out (set it to
 * dt_initial before the first call).
 * ---------------------------------------------------------------------------------------- */
int orc_evolve_aux(const orc_desc *d, const orc_ctrl *c, double *q, const double *aux, double tend,
                   long nsteps_target, orc_status *st) {
    const size_t n = (size_t)d->num_eqn * d->mx * d->my;
    double *qbackup = malloc(n * sizeof(double));
    if (!qbackup) return -1;
    const double tstart = st->t;
    const int take_one_step = 0;
    long num_steps = 0;
    st->error = 0;
    st->dtmin = st->dt;
    st->dtmax = st->dt;
    for (long it = 0; it < c->max_steps; ++it) {
        if (nsteps_target <= 0 && !take_one_step) {
            if (st->t + st->dt > tend && tstart < tend) st->dt = tend - st->t;
            if (tend - st->t - st->dt < 1.0e-14 * st->t) st->dt = tend - st->t;
        }
        const double told = st->t;
        if (c->dt_variable) memcpy(qbackup, q, n * sizeof(double));
        double cfl = 0.0;
        if (orc_step_hyperbolic_aux(d, c, q, aux, st->dt, &cfl) != 0) { free(qbackup); return -1; }
        st->attempts += 1;
        st->last_cfl = cfl;
        if (!(cfl > c->cfl_max)) {
            st->cflmax = fmax(cfl, st->cflmax);
            if (c->dt_variable)
                st->t += st->dt;
            else
                st->t = tstart + (double)(num_steps + 1) * st->dt;
            num_steps += 1;
            st->numsteps += 1;
        } else {
            st->rejected += 1;
            if (c->dt_variable) {
                memcpy(q, qbackup, n * sizeof(double));
                st->t = told;
            } else {
                st->cflmax = fmax(cfl, st->cflmax);
                st->error = 2;
                free(qbackup);
                return 0;
            }
        }
        if (nsteps_target > 0) {
            if (num_steps >= nsteps_target) break;
        } else if (st->t >= tend) {
            break;
        }
        if (c->dt_variable) {
            if (cfl > 0.0) {
                st->dt = fmin(c->dt_max, st->dt * c->cfl_desired / cfl);
                st->dtmin = fmin(st->dt, st->dtmin);
                st->dtmax = fmax(st->dt, st->dtmax);
            } else {
                st->dt = c->dt_max;
            }
        }
    }
    /* Solver.evolve_to_time raises "Maximum number of timesteps" only with a variable dt, short of
     * tend, after max_steps accepted steps: rejected attempts end the loop silently */
    if (nsteps_target > 0 ? num_steps < nsteps_target
                          : (c->dt_variable && st->t < tend && num_steps == c->max_steps))
        st->error = 1;
    free(qbackup);
    return 0;
}

int orc_evolve(const orc_desc *d, const orc_ctrl *c, double *q, double tend, long nsteps_target,
               orc_status *st) {
    return orc_evolve_aux(d, c, q, NULL, tend, nsteps_target, st);
}

/* ------------------------------------------------------------------------------------------
 * Unit entry points for tests: one slice of rpn2 / rpt2 / flux2 on a q1d(meqn, 1-g:n+g).
 * ---------------------------------------------------------------------------------------- */
int orc_rpn2_aux(const orc_desc *d, int ixy, int n, const double *q1d_in, const double *aux1d_in,
                 double *wave_out, double *s_out, double *amdq_out, double *apdq_out) {
    const int g = d->num_ghost, meqn = d->num_eqn, mwaves = d->num_waves;
    if (d->solver == ORC_SWE)
        rpn2_swe(ixy, n, g, meqn, mwaves, q1d_in, d->grav, wave_out, s_out, amdq_out, apdq_out);
    else if (d->solver == ORC_SWE_BATHY)
        rpn2_bathy(ixy, n, g, meqn, mwaves, q1d_in, aux1d_in, d->num_aux, d->grav, d->dry_tolerance,
                   wave_out, s_out, amdq_out, apdq_out);
    else
        rpn2_euler(ixy, n, g, meqn, mwaves, q1d_in, d->gamma, wave_out, s_out, amdq_out, apdq_out);
    return 0;
}

int orc_rpn2(const orc_desc *d, int ixy, int n, const double *q1d_in, double *wave_out,
             double *s_out, double *amdq_out, double *apdq_out) {
    if (d->num_aux > 0) return -1;
    return orc_rpn2_aux(d, ixy, n, q1d_in, NULL, wave_out, s_out, amdq_out, apdq_out);
}

int orc_rpt2(const orc_desc *d, int ixy, int n, const double *q1d_in, const double *asdq,
             double *bmasdq, double *bpasdq) {
    const int g = d->num_ghost, meqn = d->num_eqn;
    if (d->solver == ORC_SWE)
        rpt2_swe(ixy, n, g, meqn, q1d_in, d->grav, asdq, bmasdq, bpasdq);
    else if (d->solver == ORC_SWE_BATHY)
        rpt2_bathy(ixy, n, g, meqn, q1d_in, d->grav, d->dry_tolerance, asdq, bmasdq, bpasdq);
    else
        rpt2_euler(ixy, n, g, meqn, q1d_in, d->gamma, asdq, bmasdq, bpasdq);
    return 0;
}

int orc_limiter(const orc_desc *d, int n, double *wave, const double *s) {
    const int g = d->num_ghost, meqn = d->num_eqn, mwaves = d->num_waves;
    limiter(n, g, meqn, mwaves, wave, s, d->mthlim);
    return 0;
}

/* flux2 on one slice: outputs qadd, fadd (meqn x len) and gadd (meqn x 2 x len), cfl1d;
 * aux1d_in: aux(num_aux, len) of the slice, or NULL when num_aux = 0 */
int orc_flux2_aux(const orc_desc *d, int ixy, int n, double dtdx, const double *q1d_in,
                  const double *aux1d_in, double *qadd_out, double *fadd_out, double *gadd_out,
                  double *cfl1d) {
    const int g = d->num_ghost, meqn = d->num_eqn;
    if (d->num_aux > 0 && !aux1d_in) return -1;
    orc_work w;
    if (work_alloc(&w, n + 2 * g, meqn, d->num_waves, d->num_aux) != 0) { work_free(&w); return -1; }
    memcpy(w.q1d, q1d_in, sizeof(double) * meqn * (n + 2 * g));
    if (d->num_aux > 0) memcpy(w.aux1d, aux1d_in, sizeof(double) * d->num_aux * (n + 2 * g));
    flux2(d, ixy, n, dtdx, &w, cfl1d);
    memcpy(qadd_out, w.qadd, sizeof(double) * meqn * (n + 2 * g));
    memcpy(fadd_out, w.fadd, sizeof(double) * meqn * (n + 2 * g));
    memcpy(gadd_out, w.gadd, sizeof(double) * meqn * 2 * (n + 2 * g));
    work_free(&w);
    return 0;
}

int orc_flux2(const orc_desc *d, int ixy, int n, double dtdx, const double *q1d_in,
              double *qadd_out, double *fadd_out, double *gadd_out, double *cfl1d) {
    return orc_flux2_aux(d, ixy, n, dtdx, q1d_in, NULL, qadd_out, fadd_out, gadd_out, cfl1d);
}

/* Compensated (Neumaier) sum of one component over the interior, for conservation checks. */
double orc_component_sum(const orc_desc *d, const double *q, int m) {
    const size_t cells = (size_t)d->mx * d->my;
    double s = 0.0, comp = 0.0;
    for (size_t k = 0; k < cells; ++k) {
        const double x = q[m + d->num_eqn * k];
        const double t = s + x;
        if (fabs(s) >= fabs(x)) comp += (s - t) + x; else comp += (x - t) + s;
        s = t;
    }
    return s + comp;
}

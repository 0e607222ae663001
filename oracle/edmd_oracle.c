/*
 * CPU restatement of the reference's event-driven hard-disk engine.
 * TEST INFRASTRUCTURE ONLY: see edmd_oracle.h.  Build strict IEEE
 * (-O2 -fno-fast-math -ffp-contract=off, oracle/Makefile).
 *
 * Reference line numbers are those of src/EDMD.c unless a file is named.
 */
#include "edmd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NEVER 100000000000000000000000000.0 /* EDMD.h:8 */
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define NIL (-1)

typedef struct {
    int lft, rgt, top; /* BST links, or bucket-list links (lft = prev, rgt = next) */
    int i, j, q;       /* owner, partner / direction, bucket index */
    double t;
    int type;
    uint64_t coll_actual;
} ev_node;

struct orc_ctx {
    int n;
    double Lx, Ly, halfLx, halfLy;
    int ncx, ncy;
    double csx, csy, cfx, cfy;
    orc_params prm;
    double t;
    int status;
    /* particles (structure of arrays) */
    double *x, *y, *vx, *vy, *tl, *rad, *m;
    uint64_t* coll;
    int *cx, *cy, *crossX, *crossY;
    /* cell lists: head per cell, doubly linked through particles (EDMD.c:2053-2095) */
    int *head, *nxt, *prv;
    /* event node pool: [0,N) crossings, [N,2N) collisions, 2N root,
     * 2N+1 screenshot, 2N+2 noise, 2N+3 thermo, 2N+5 growth stop
     * (EDMD.c:1935-1948, 4421-4440, 4802-4808, 4814-4822) */
    ev_node* ev;
    int root, tree_min;
    double paul_time, dt_paul;
    int cur_list, n_lists;
    int* lists; /* n_lists + 1 heads, last = overflow */
    /* counters and accumulators */
    uint64_t ncol, ncross, nevents;
    double vir_xx, vir_yy, vir_xy, vir_yx;
    double E, px, py;
    double last_thermo;
    orc_thermo_row* rows;
    int nrows, caprows;
    orc_tape_rec* tape;
    int64_t ntape, captape;
    double wsum;
    /* growth start: per-disk radius growth rate, set while growing */
    double* vr;
    int growing;
    double dp; /* momentum given to the walls since the last thermo row */
    int* type; /* disk types (the mid wall stops type 1) */
    double dp_left, dp_mid, dp_right; /* dpLeft, dpMid, dpRight (EDMD.c:146-148) */
    uint32_t mt[624]; /* MT19937 stream of the thermal walls */
    int mti;
};

/* ------------------------------------------------------------ MT19937 */

/* init_genrand and genrand_int32 of mersenne.c (Matsumoto and Nishimura,
 * ACM TOMACS 8(1), 1998, with the 2002 initialisation), restated in order:
 * the state from a 32-bit seed, the twist of all 624 words when they are
 * used up, the tempering of the word drawn. */
static void mt_seed(uint32_t* mt, int* mti, uint32_t seed) {
    mt[0] = seed;
    for (int k = 1; k < 624; k++) mt[k] = 1812433253u * (mt[k - 1] ^ (mt[k - 1] >> 30)) + (uint32_t)k;
    *mti = 624;
}

static uint32_t mt_next(uint32_t* mt, int* mti) {
    if (*mti >= 624) {
        for (int k = 0; k < 624; k++) {
            uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
            mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        *mti = 0;
    }
    uint32_t y = mt[(*mti)++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

void orc_mt_draws(uint32_t seed, int n, uint32_t* out) {
    uint32_t mt[624];
    int mti;
    mt_seed(mt, &mti, seed);
    for (int k = 0; k < n; k++) out[k] = mt_next(mt, &mti);
}

/* drand(0, hi) (EDMD.c:5699-5701) and sign (EDMD.c:5703-5705). */
static double drand0(orc_ctx* s, double hi) {
    double g = mt_next(s->mt, &s->mti) * (1.0 / 4294967296.0);
    return (g * (hi - 0.0)) + 0.0;
}

static double ref_sign(double x) { return (x < 0) ? -1.0 : (double)(x > 0); }

/* ------------------------------------------------------------ kinematics */

/* PBC, EDMD.c:5896-5913: minimum image with >= L/2 and < -L/2, none
 * across walls. */
static void min_image(const orc_ctx* s, double* dx, double* dy) {
    if (!s->prm.wall_x && !s->prm.wall_circle) {
        if (*dx >= s->halfLx) *dx -= s->Lx;
        else if (*dx < -s->halfLx) *dx += s->Lx;
    }
    if (!s->prm.wall_y && !s->prm.wall_circle) {
        if (*dy >= s->halfLy) *dy -= s->Ly;
        else if (*dy < -s->halfLy) *dy += s->Ly;
    }
}

/* PBCinsideCellX/Y, EDMD.c:5922-5936. */
static double wrap_inside(double d, double half, double L) {
    if (d >= half) return d - L;
    if (d < -half) return d + L;
    return d;
}

/* PBCpostX/Y, EDMD.c:5948-5959. */
static double wrap_post(double v, double L) {
    if (v < 0) return v + L;
    if (v >= L) return v - L;
    return v;
}

/* e^x, from its Taylor series: x = k ln2 + r, |r| <= ln2/2, with ln2 in
 * two parts so that k ln2 is exact (Cody and Waite's reduction); e^r by
 * the nested form 1 + r(1 + r/2(1 + r/3(... (1 + r/17)))), whose
 * truncation error is below r^18/18! < 2e-24 relative; times 2^k.  Only
 * +, *, rint and ldexp, so any IEEE double machine computes the same
 * bits (the GPU library evaluates the same steps). */
static const double ORC_LN2_HI = 0x1.62e42fefa3800p-1; /* 43 significant bits */
static const double ORC_LN2_LO = 0x1.ef35793c76730p-45;
static const double ORC_INV[24] = {0, 1.0, 1.0 / 2, 1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 6, 1.0 / 7,
                                   1.0 / 8, 1.0 / 9, 1.0 / 10, 1.0 / 11, 1.0 / 12, 1.0 / 13,
                                   1.0 / 14, 1.0 / 15, 1.0 / 16, 1.0 / 17, 1.0 / 18, 1.0 / 19,
                                   1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23};

double orc_exp(double x) {
    if (x != x) return x;
    if (x > 709.5) return HUGE_VAL;
    if (x < -745.5) return 0.0;
    const double k = rint(x * 1.4426950408889634); /* x / ln2, to nearest */
    const double r = (x - k * ORC_LN2_HI) - k * ORC_LN2_LO;
    double p = 1.0;
    for (int n = 17; n >= 1; --n) p = 1.0 + (r * p) * ORC_INV[n];
    return ldexp(p, (int)k);
}

/* log x, from the series log m = 2 atanh s = 2 (s + s^3/3 + s^5/5 + ...),
 * s = (m - 1)/(m + 1), for x = m 2^e with m in [sqrt(1/2), sqrt(2)) so
 * |s| <= 0.1716 and the terms to s^23/23 leave less than 1e-19 relative;
 * plus e ln2 in two parts.  Same steps on the GPU. */
double orc_log(double x) {
    if (x != x || x < 0) return NAN;
    if (x == 0) return -HUGE_VAL;
    if (x == HUGE_VAL) return x;
    int e;
    double m = frexp(x, &e); /* [0.5, 1) */
    if (m < 0.70710678118654752) {
        m *= 2;
        e -= 1;
    }
    const double s = (m - 1) / (m + 1), s2 = s * s;
    double q = ORC_INV[23];
    for (int n = 21; n >= 1; n -= 2) q = ORC_INV[n] + s2 * q;
    const double lm = 2 * s * q;
    return e * ORC_LN2_HI + (e * ORC_LN2_LO + lm);
}

/* sin and cos of an angle in [-pi, pi] (atan2's range): theta = k pi/2 + r,
 * |r| <= pi/4, with pi/2 in three parts (33, 33 and 53 bits, Cody and Waite)
 * so that k times the first two is exact; sin r and cos r from their Taylor
 * series to r^19/19! and r^20/20! (truncation below 1e-19 relative) in the
 * nested forms r (1 - r^2/(2 3) (1 - r^2/(4 5) (...))) and
 * 1 - r^2/(1 2) (1 - r^2/(3 4) (...)); the quadrant k mod 4 picks and signs
 * them.  Constants: oracle/trig_constants.py.  Same steps on the GPU. */
static const double ORC_PIO2_1 = 0x1.921fb54400000p+0;
static const double ORC_PIO2_2 = 0x1.0b4611a600000p-34;
static const double ORC_PIO2_3 = 0x1.3198a2e037073p-69;
static const double ORC_PIO2_HI = 0x1.921fb54442d18p+0, ORC_PIO2_LO = 0x1.1a62633145c07p-54;
static const double ORC_PI_HI = 0x1.921fb54442d18p+1, ORC_PI_LO = 0x1.1a62633145c07p-53;
static const double ORC_2_OVER_PI = 0x1.45f306dc9c883p-1;
static const double ORC_ATAN_04 = 0x1.85a376b677dc0p-2; /* atan(0.4) */
static const double ORC_ATAN_08 = 0x1.5977a5103ea93p-1; /* atan(0.8) */

void orc_sincos(double th, double* sn, double* cs) {
    const double k = rint(th * ORC_2_OVER_PI);
    const double r = ((th - k * ORC_PIO2_1) - k * ORC_PIO2_2) - k * ORC_PIO2_3;
    const double r2 = r * r;
    double ps = 1.0, pc = 1.0;
    for (int n = 19; n >= 3; n -= 2) ps = 1.0 - (r2 * (ORC_INV[n] * ORC_INV[n - 1])) * ps;
    for (int n = 20; n >= 2; n -= 2) pc = 1.0 - (r2 * (ORC_INV[n] * ORC_INV[n - 1])) * pc;
    const double sr = r * ps, cr = pc;
    switch (((int)k) & 3) {
        case 0: *sn = sr; *cs = cr; break;
        case 1: *sn = cr; *cs = -sr; break;
        case 2: *sn = -sr; *cs = -cr; break;
        default: *sn = -cr; *cs = sr; break;
    }
}

/* atan z for z in [0, 1]: about 0.4 or 0.8 (atan z = atan c + atan w,
 * w = (z - c)/(1 + z c)) so that |w| <= 0.2, then the series
 * w (1 - w^2 (1/3 - w^2 (1/5 - ... w^2/23))), truncation below 1e-18
 * relative. */
static double orc_atan01(double z) {
    double base = 0, w = z;
    if (z > 0.6) {
        w = (z - 0.8) / (1 + z * 0.8);
        base = ORC_ATAN_08;
    } else if (z > 0.2) {
        w = (z - 0.4) / (1 + z * 0.4);
        base = ORC_ATAN_04;
    }
    const double w2 = w * w;
    double q = ORC_INV[23];
    for (int n = 21; n >= 3; n -= 2) q = ORC_INV[n] - w2 * q;
    return base + w * (1 - w2 * q);
}

/* atan2(y, x) in [-pi, pi]: the angle of the smaller over the larger
 * coordinate, reflected about pi/4 and pi/2 and signed by quadrant (pi/2
 * and pi as head + tail).  Same steps on the GPU. */
double orc_atan2(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    double a;
    if (ay > ax) a = (ORC_PIO2_HI - orc_atan01(ax / ay)) + ORC_PIO2_LO;
    else if (ax > 0) a = orc_atan01(ay / ax);
    else a = 0;
    if (signbit(x)) a = (ORC_PI_HI - a) + ORC_PI_LO;
    return signbit(y) ? -a : a;
}

/* The quartic of the circular wall under the field (quartic.c:27-158, after
 * github.com/sasamil/Quartic): Ferrari's method through the cubic
 * resolvent, step for step as there, with this oracle's own acos (atan2 of
 * sqrt((1 - t)(1 + t)) and t), cos and cube root (exp(log(x) / 3), the pow
 * with exponent 1./3) in place of the C library's.  Same steps on the GPU. */
#define ORC_QEPS 1e-12
static double orc_acos(double t) { return orc_atan2(sqrt((1 - t) * (1 + t)), t); }
static double orc_cosine(double th) {
    double sn, cs;
    orc_sincos(th, &sn, &cs);
    return cs;
}
static double orc_cbrt_pos(double x) { return x == 0 ? 0.0 : orc_exp(orc_log(x) * (1. / 3)); }

static unsigned orc_solve_p3(double* x, double a, double b, double c) {
    double a2 = a * a;
    double q = (a2 - 3 * b) / 9;
    double r = (a * (2 * a2 - 9 * b) + 27 * c) / 54;
    double r2 = r * r;
    double q3 = q * q * q;
    double A, B;
    if (r2 < q3) {
        double t = r / sqrt(q3);
        if (t < -1) t = -1;
        if (t > 1) t = 1;
        t = orc_acos(t);
        a /= 3;
        q = -2 * sqrt(q);
        x[0] = q * orc_cosine(t / 3) - a;
        x[1] = q * orc_cosine((t + (2.0 * M_PI)) / 3) - a;
        x[2] = q * orc_cosine((t - (2.0 * M_PI)) / 3) - a;
        return 3;
    }
    A = -orc_cbrt_pos(fabs(r) + sqrt(r2 - q3));
    if (r < 0) A = -A;
    B = (0 == A ? 0 : q / A);
    a /= 3;
    x[0] = (A + B) - a;
    x[1] = -0.5 * (A + B) - a;
    x[2] = 0.5 * sqrt(3.) * (A - B);
    if (fabs(x[2]) < ORC_QEPS) {
        x[2] = x[1];
        return 2;
    }
    return 1;
}

/* The smallest real root above 1e-9 of a t^4 + b t^3 + c t^2 + d t + e,
 * 1e10 when there is none (smallestRoot, quartic.c:143-157). */
double orc_smallest_root(double a, double b, double c, double d, double e) {
    const double qa = b / a, qb = c / a, qc = d / a, qd = e / a;
    double a3 = -qb;
    double b3 = qa * qc - 4. * qd;
    double c3 = -qa * qa * qd - qc * qc + 4. * qb * qd;
    double x3[3];
    unsigned zeroes = orc_solve_p3(x3, a3, b3, c3);
    double q1, q2, p1, p2, D, sqD, y = x3[0];
    if (zeroes != 1) {
        if (fabs(x3[1]) > fabs(y)) y = x3[1];
        if (fabs(x3[2]) > fabs(y)) y = x3[2];
    }
    D = y * y - 4 * qd;
    if (fabs(D) < ORC_QEPS) {
        q1 = q2 = y * 0.5;
        D = qa * qa - 4 * (qb - y);
        if (fabs(D) < ORC_QEPS) {
            p1 = p2 = qa * 0.5;
        } else {
            sqD = sqrt(D);
            p1 = (qa + sqD) * 0.5;
            p2 = (qa - sqD) * 0.5;
        }
    } else {
        sqD = sqrt(D);
        q1 = (y + sqD) * 0.5;
        q2 = (y - sqD) * 0.5;
        p1 = (qa * q1 - qc) / (q1 - q2);
        p2 = (qc - qa * q2) / (q1 - q2);
    }
    double re[4], im[4];
    D = p1 * p1 - 4 * q1;
    if (D < 0.0) {
        re[0] = -p1 * 0.5;
        im[0] = sqrt(-D) * 0.5;
        re[1] = re[0];
        im[1] = -im[0];
    } else {
        sqD = sqrt(D);
        re[0] = (-p1 + sqD) * 0.5;
        re[1] = (-p1 - sqD) * 0.5;
        im[0] = im[1] = 0.0;
    }
    D = p2 * p2 - 4 * q2;
    if (D < 0.0) {
        re[2] = -p2 * 0.5;
        im[2] = sqrt(-D) * 0.5;
        re[3] = re[2];
        im[3] = -im[2];
    } else {
        sqD = sqrt(D);
        re[2] = (-p2 + sqD) * 0.5;
        re[3] = (-p2 - sqD) * 0.5;
        im[2] = im[3] = 0.0;
    }
    double best = 10000000000;
    for (int i = 0; i < 4; i++)
        if (im[i] == 0 && re[i] > 0.000000001 && re[i] < best) best = re[i];
    return best;
}

/* logTime, EDMD.c:5709-5721: the time at which the damped motion has gone
 * as far as the undamped one in t_lin. */
static double log_time(const orc_ctx* s, double t_lin) {
    if (!s->prm.damping) return t_lin;
    double var = (1 - t_lin * s->prm.gamma);
    if (var < 0) return NEVER;
    return -orc_log(var) / s->prm.gamma;
}

/* freeFlyNormal, EDMD.c:4954-4990: straight, parabolic in y under the
 * field, exponentially slowed under damping; with both on, the field's
 * flight (its branch comes first, EDMD.c:4957). */
static void free_fly(orc_ctx* s, int p) {
    double dt = s->t - s->tl[p];
    s->tl[p] = s->t;
    if (s->prm.add_field) {
        s->x[p] += dt * s->vx[p];
        s->y[p] += dt * s->vy[p] + 0.5 * s->prm.field * dt * dt;
        s->vy[p] += dt * s->prm.field;
    } else if (s->prm.damping) {
        double g = s->prm.gamma;
        double ex = orc_exp(-g * dt);
        s->x[p] += s->vx[p] / g * (1 - ex);
        s->y[p] += s->vy[p] / g * (1 - ex);
        s->vx[p] *= ex;
        s->vy[p] *= ex;
    } else {
        s->x[p] += dt * s->vx[p];
        s->y[p] += dt * s->vy[p];
    }
    s->x[p] = wrap_post(s->x[p], s->Lx);
    s->y[p] = wrap_post(s->y[p], s->Ly);
}

/* freeFlyGrow, EDMD.c:4992-5007: the radius grows with the flight. */
static void free_fly_grow(orc_ctx* s, int p) {
    double dt = s->t - s->tl[p];
    s->tl[p] = s->t;
    s->x[p] += dt * s->vx[p];
    s->y[p] += dt * s->vy[p];
    s->rad[p] += dt * s->vr[p];
    s->x[p] = wrap_post(s->x[p], s->Lx);
    s->y[p] = wrap_post(s->y[p], s->Ly);
}

/* The freeFly dispatch pointer (EDMD.c:1979, 4757). */
static void fly(orc_ctx* s, int p) {
    if (s->growing) free_fly_grow(s, p);
    else free_fly(s, p);
}

/* DIST2, EDMD.c:35-43 (non-additive geometric mean contact). */
static double contact2(double r1, double r2) { return 4 * r1 * r2; }

/* ------------------------------------------------------------ calendar */

static void tree_insert(orc_ctx* s, int e) {
    /* addEventToTree, EDMD.c:2175-2203: equal keys go right. */
    ev_node* ev = s->ev;
    int w = s->root;
    for (;;) {
        if (ev[e].t < ev[w].t) {
            if (ev[w].lft == NIL) { ev[w].lft = e; break; }
            w = ev[w].lft;
        } else {
            if (ev[w].rgt == NIL) { ev[w].rgt = e; break; }
            w = ev[w].rgt;
        }
    }
    ev[e].lft = NIL;
    ev[e].rgt = NIL;
    ev[e].top = w;
    if (s->tree_min == NIL || ev[e].t < ev[s->tree_min].t) s->tree_min = e;
}

static void queue_add(orc_ctx* s, int e) {
    /* addEventToQueue, EDMD.c:2144-2170. */
    ev_node* ev = s->ev;
    double dt = ev[e].t - s->paul_time;
    if (dt < s->dt_paul) {
        ev[e].q = s->cur_list;
        tree_insert(s, e);
        return;
    }
    int idx;
    if (!(dt < s->dt_paul * s->n_lists)) { /* also NaN: parked in the overflow list */
        idx = s->n_lists;
    } else {
        idx = s->cur_list + (int)(dt / s->dt_paul);
        if (idx >= s->n_lists) idx -= s->n_lists;
    }
    ev[e].q = idx;
    ev[e].lft = NIL;
    ev[e].rgt = s->lists[idx];
    if (ev[e].rgt != NIL) ev[ev[e].rgt].lft = e;
    s->lists[idx] = e;
}

static void tree_remove(orc_ctx* s, int e) {
    /* removeEventFromTree, EDMD.c:2228-2280 (successor splice). */
    ev_node* ev = s->ev;
    int top = ev[e].top;
    int n;
    if (e == s->tree_min) {
        if (ev[e].rgt != NIL) {
            int m = ev[e].rgt;
            while (ev[m].lft != NIL) m = ev[m].lft;
            s->tree_min = m;
        } else {
            s->tree_min = (top == s->root) ? NIL : top;
        }
    }
    if (ev[e].lft == NIL && ev[e].rgt == NIL) {
        n = NIL;
    } else if (ev[e].lft == NIL) {
        n = ev[e].rgt;
        ev[n].top = top;
    } else if (ev[e].rgt == NIL) {
        n = ev[e].lft;
        ev[n].top = top;
    } else {
        n = ev[e].rgt;
        while (ev[n].lft != NIL) n = ev[n].lft;
        if (ev[n].top != e) {
            ev[ev[n].top].lft = ev[n].rgt;
            if (ev[n].rgt != NIL) ev[ev[n].rgt].top = ev[n].top;
            ev[ev[e].rgt].top = n;
            ev[n].rgt = ev[e].rgt;
        }
        ev[ev[e].lft].top = n;
        ev[n].lft = ev[e].lft;
        ev[n].top = top;
    }
    if (ev[top].lft == e) ev[top].lft = n;
    else ev[top].rgt = n;
}

static void queue_remove(orc_ctx* s, int e) {
    /* removeEventFromQueue, EDMD.c:2209-2221. */
    ev_node* ev = s->ev;
    if (ev[e].q != s->cur_list) {
        if (ev[e].rgt != NIL) ev[ev[e].rgt].lft = ev[e].lft;
        if (ev[e].lft != NIL) ev[ev[e].lft].rgt = ev[e].rgt;
        else s->lists[ev[e].q] = ev[e].rgt;
    } else {
        tree_remove(s, e);
    }
}

static void next_bucket(orc_ctx* s) {
    /* addNextPaulEvent, EDMD.c:2287-2307. */
    s->cur_list++;
    s->paul_time += s->dt_paul;
    if (s->cur_list == s->n_lists) {
        s->cur_list = 0;
        int e = s->lists[s->n_lists];
        s->lists[s->n_lists] = NIL;
        while (e != NIL) {
            int nx = s->ev[e].rgt;
            queue_add(s, e);
            e = nx;
        }
    }
    int e = s->lists[s->cur_list];
    while (e != NIL) {
        int nx = s->ev[e].rgt;
        tree_insert(s, e);
        e = nx;
    }
    s->lists[s->cur_list] = NIL;
}

static int queue_pop_min(orc_ctx* s) {
    /* findNextEvent, EDMD.c:2313-2327. */
    while (s->tree_min == NIL) next_bucket(s);
    return s->tree_min;
}

/* ------------------------------------------------------------ cell list */

static int cell_of(const orc_ctx* s, double a, double fac, int nc) {
    /* coordToCell, EDMD.c:2098-2107; clamped to the grid (a coordinate one
     * ulp below L can round to the grid size). */
    int c = (int)(a * fac);
    if (c >= nc) c = nc - 1;
    if (c < 0) c = 0;
    (void)s;
    return c;
}

static void cell_push_head(orc_ctx* s, int p) {
    /* head insertion, EDMD.c:2071-2077 and 2576-2585 */
    int c = s->cy[p] * s->ncx + s->cx[p];
    s->prv[p] = NIL;
    s->nxt[p] = s->head[c];
    s->head[c] = p;
    if (s->nxt[p] != NIL) s->prv[s->nxt[p]] = p;
}

static void cell_unlink(orc_ctx* s, int p) {
    /* removeFromCell, EDMD.c:2083-2095 */
    if (s->prv[p] == NIL) s->head[s->cy[p] * s->ncx + s->cx[p]] = s->nxt[p];
    else s->nxt[s->prv[p]] = s->nxt[p];
    if (s->nxt[p] != NIL) s->prv[s->nxt[p]] = s->prv[p];
}

static int wrap_cell(int a, int nc) {
    /* PBCcellX/Y, EDMD.c:2110-2124 */
    if (a < 0) return a + nc;
    if (a >= nc) return a - nc;
    return a;
}

/* ------------------------------------------------------------ predictions */

static double coll_time(orc_ctx* s, int i, int j) {
    /* collisionTimeNormal, EDMD.c:2661-2723 (no damping or shear); under the
     * field the neighbour is flown to the current time first, a state
     * update of the neighbour (EDMD.c:2664-2666). */
    if (s->prm.add_field || s->prm.damping) free_fly(s, j);
    double lat2 = s->t - s->tl[j];
    double dvx = s->vx[j] - s->vx[i];
    double dvy = s->vy[j] - s->vy[i];
    double dx = (s->x[j] + lat2 * s->vx[j]) - s->x[i];
    double dy = (s->y[j] + lat2 * s->vy[j]) - s->y[i];
    min_image(s, &dx, &dy);
    double b = dx * dvx + dy * dvy;
    if (b > 0) return NEVER;
    double v2 = dvx * dvx + dvy * dvy;
    double d = dx * dx + dy * dy - contact2(s->rad[i], s->rad[j]);
    double det = b * b - v2 * d;
    if (d < -0.01) {
        /* overlap abort, EDMD.c:2708-2717 */
        if (s->status == ORC_OK) s->status = ORC_ERR_OVERLAP;
    }
    if (det < 0) return NEVER;
    return log_time(s, (-b - sqrt(det)) / v2);
}

static double coll_time_grow(orc_ctx* s, int i, int j) {
    /* collisionTimeGrow, EDMD.c:2598-2659 (2D): contact distance
     * DIST(r_i, r_j(t)) growing at vr_i + vr_j; the earlier positive root,
     * the later one under the reference's tolerance rule, overlap abort. */
    double lat2 = s->t - s->tl[j];
    double dvx = s->vx[j] - s->vx[i];
    double dvy = s->vy[j] - s->vy[i];
    double dvr = s->vr[i] + s->vr[j];
    double dx = (s->x[j] + lat2 * s->vx[j]) - s->x[i];
    double dy = (s->y[j] + lat2 * s->vy[j]) - s->y[i];
    double dr = sqrt(4 * s->rad[i] * (s->rad[j] + lat2 * s->vr[j]));
    min_image(s, &dx, &dy);
    double b = dx * dvx + dy * dvy - dvr * dr;
    double v2 = dvx * dvx + dvy * dvy;
    double dist2 = dx * dx + dy * dy;
    double det = b * b - (v2 - dvr * dvr) * (dist2 - dr * dr);
    if (det < 0) return NEVER;
    double plus = (-b + sqrt(det)) / (v2 - dvr * dvr);
    double minus = (-b - sqrt(det)) / (v2 - dvr * dvr);
    if (((minus > 0) && (plus > 0) && (minus < plus)) || ((minus > 0) && (plus < 0)))
        return minus;
    else if (((minus > 0) && (plus > 0.000000001) && (plus < minus)) ||
             ((minus < 0) && (plus > 0.000000001)))
        return plus;
    if (dist2 - dr * dr < -0.01) {
        /* growth overlap abort, EDMD.c:2653-2656 */
        if (s->status == ORC_OK) s->status = ORC_ERR_OVERLAP;
    }
    return NEVER;
}

static void predict_collision_grow(orc_ctx* s, int i, double* t_abs, int* partner) {
    /* collisionEventGrow, EDMD.c:3104-3230 (hard walls, no circular or mid
     * wall): the minimum starts at 1e7, not never, and is scanned as in the
     * Normal version. */
    int best = 0;
    double dt = 10000000;
    int X = s->cx[i], Y = s->cy[i];
    /* walls first (EDMD.c:3118-3153): the surface of a disk in a boundary
     * cell closes on the wall at v -/+ vr; no logTime.  The x test sets
     * xy = 0 whenever it applies, as in the Normal version. */
    int wall = 0, xy = 2;
    const double vr = s->vr[i];
    if (s->prm.wall_y) {
        double vyp = s->vy[i] + vr, vym = s->vy[i] - vr;
        if (Y == 0 && vym < 0) {
            dt = -(s->y[i] - s->rad[i]) / vym;
            wall = 1;
            xy = 1;
        } else if (Y == s->ncy - 1 && vyp > 0) {
            dt = (s->Ly - s->y[i] - s->rad[i]) / vyp;
            wall = 1;
            xy = 1;
        }
    }
    if (s->prm.wall_x) {
        double vxp = s->vx[i] + vr, vxm = s->vx[i] - vr;
        if (X == 0 && vxm < 0) {
            double dtw = -(s->x[i] - s->rad[i]) / vxm;
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        } else if (X == s->ncx - 1 && vxp > 0) {
            double dtw = (s->Lx - s->x[i] - s->rad[i]) / vxp;
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        }
    }
    if (s->prm.mid_wall && s->type[i] == 1) {
        /* EDMD.c:3175-3195 */
        double xi = s->x[i], h = s->Lx / 2;
        double vxp = s->vx[i] + vr, vxm = s->vx[i] - vr;
        if (xi > h - 4 && xi < h && vxp > 0) {
            double dtw = (h - xi - s->rad[i]) / (vxp);
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        } else if (xi < h + 4 && xi > h && vxm < 0) {
            double dtw = ((xi - s->rad[i]) - h) / (-vxm);
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        }
    }
    if (s->prm.wall_circle) {
        /* EDMD.c:3158-3172: the surface closes on the circle at |v| - vr */
        double y0 = s->y[i] - s->halfLx, x0 = s->x[i] - s->halfLx;
        double v0x = s->vx[i], v0y = s->vy[i], rad = s->rad[i];
        double A = v0x * v0x + v0y * v0y - vr * vr;
        double B = 2 * (v0x * x0 + v0y * y0 + (s->halfLx - rad) * vr);
        double Cc = (x0 * x0 + y0 * y0) - (s->halfLx - rad) * (s->halfLx - rad);
        dt = (-B + sqrt(B * B - 4 * A * Cc)) / (2 * A);
        wall = 1;
        xy = 2;
    }
    for (int j = -1; j <= 1; j++) {
        for (int k = -1; k <= 1; k++) {
            int c = wrap_cell(Y + j, s->ncy) * s->ncx + wrap_cell(X + k, s->ncx);
            for (int p = s->head[c]; p != NIL; p = s->nxt[p]) {
                if (p == i) continue;
                double dtt = coll_time_grow(s, i, p);
                if (dt > dtt) {
                    best = p;
                    dt = dtt;
                    wall = 0;
                }
            }
        }
    }
    *t_abs = s->t + dt;
    *partner = wall ? ORC_WALL_PARTNER(xy) : best;
}

static void predict_collision(orc_ctx* s, int i, double* t_abs, int* partner) {
    if (s->growing) {
        predict_collision_grow(s, i, t_abs, partner);
        return;
    }
    /* collisionEventNormal, EDMD.c:2829-3003, 3091-3093: strict-less minimum
     * over the 3x3 cells, y offset outer, x offset inner, list order. */
    int best = 0;
    double dt = NEVER;
    int X = s->cx[i], Y = s->cy[i];
    /* walls first (EDMD.c:2845-2895, no field; logTime is the identity
     * without damping): a disk in a boundary cell moving towards the wall
     * gets the time to contact.  The x test sets xy = 0 whenever it applies,
     * also when the y wall stays the earlier, as the reference does. */
    int wall = 0, xy = 2;
    if (s->prm.wall_y && s->prm.add_field) {
        /* EDMD.c:2846-2860: the bottom wall from the bottom row whatever the
         * velocity, the top wall for any disk moving up (NaN when it cannot
         * reach it) */
        double f = s->prm.field, vy = s->vy[i];
        if (Y == 0) {
            dt = (-vy - sqrt(vy * vy - 2 * f * (s->y[i] - s->rad[i]))) / f;
            wall = 1;
            xy = 1;
        }
        if (vy > 0) {
            double dt2 = (-vy + sqrt(vy * vy - 2 * f * (s->y[i] - (s->Ly - s->rad[i])))) / f;
            if (!isnan(dt2) && dt2 < dt) {
                dt = dt2;
                wall = 1;
                xy = 1;
            }
        }
    } else if (s->prm.wall_y) {
        if (Y == 0 && s->vy[i] < 0) {
            dt = log_time(s, -(s->y[i] - s->rad[i]) / s->vy[i]);
            wall = 1;
            xy = 1;
        } else if (Y == s->ncy - 1 && s->vy[i] > 0) {
            dt = log_time(s, (s->Ly - s->y[i] - s->rad[i]) / s->vy[i]);
            wall = 1;
            xy = 1;
        }
    }
    if (s->prm.wall_x) {
        if (X == 0 && s->vx[i] < 0) {
            double dtw = log_time(s, -(s->x[i] - s->rad[i]) / s->vx[i]);
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        } else if (X == s->ncx - 1 && s->vx[i] > 0) {
            double dtw = log_time(s, (s->Lx - s->x[i] - s->rad[i]) / s->vx[i]);
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        }
    }
    if (s->prm.mid_wall && s->type[i] == 1) {
        /* EDMD.c:2897-2915: from within 4 of x = Lx/2, moving towards it */
        double xi = s->x[i], vx = s->vx[i], h = s->Lx / 2;
        if (xi > h - 4 && xi < h && vx > 0) {
            double dtw = log_time(s, (h - xi - s->rad[i]) / vx);
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        } else if (xi < h + 4 && xi > h && vx < 0) {
            double dtw = log_time(s, ((xi - s->rad[i]) - h) / (-vx));
            if (dt > dtw) {
                wall = 1;
                dt = dtw;
            }
            xy = 0;
        }
    }
    if (s->prm.wall_circle && s->prm.add_field) {
        /* EDMD.c:2918-2935: the parabolic flight meets the circle, the
         * smallest positive root of the quartic (y0 about halfLy here) */
        double y0 = s->y[i] - s->halfLy, x0 = s->x[i] - s->halfLx, rad = s->rad[i];
        double e = x0 * x0 + y0 * y0 - (s->halfLx - rad) * (s->halfLx - rad);
        double v0x = s->vx[i], v0y = s->vy[i], f = s->prm.field;
        double a = f * f / 4;
        double b = f * v0y;
        double c = v0x * v0x + v0y * v0y + f * y0;
        double d = 2 * v0x * x0 + 2 * v0y * y0;
        dt = orc_smallest_root(a, b, c, d, e);
        wall = 1;
        xy = 2;
    } else if (s->prm.wall_circle) {
        /* EDMD.c:2938-2954 (no field): the centre reaches the circle of
         * radius halfLx - rad; the time replaces any earlier wall's */
        double y0 = s->y[i] - s->halfLx, x0 = s->x[i] - s->halfLx, rad = s->rad[i];
        double Cc = (x0 * x0 + y0 * y0) - (s->halfLx - rad) * (s->halfLx - rad);
        double v0x = s->vx[i], v0y = s->vy[i];
        double A = (v0x * v0x + v0y * v0y);
        double B = 2 * (v0x * x0 + v0y * y0);
        dt = (-B + sqrt(B * B - 4 * A * Cc)) / (2 * A);
        wall = 1;
        xy = 2;
    }
    for (int j = -1; j <= 1; j++) {
        for (int k = -1; k <= 1; k++) {
            int c = wrap_cell(Y + j, s->ncy) * s->ncx + wrap_cell(X + k, s->ncx);
            for (int p = s->head[c]; p != NIL; p = s->nxt[p]) {
                if (p == i) continue;
                double dtt = coll_time(s, i, p);
                if (dt > dtt) {
                    best = p;
                    dt = dtt;
                    wall = 0;
                }
            }
        }
    }
    *t_abs = s->t + dt;
    *partner = wall ? ORC_WALL_PARTNER(xy) : best;
}

/* logTime where the crossing uses it: crossingEventGrow (EDMD.c:2343-2400)
 * takes the straight-line times as they are. */
static double cross_log_time(const orc_ctx* s, double t_lin) {
    return s->growing ? t_lin : log_time(s, t_lin);
}

static void predict_crossing(const orc_ctx* s, int i, double* t_abs, int* dir) {
    /* crossingEventNormal, EDMD.c:2405-2482, or crossingEventGrow while
     * growing (no field, no damping): ties go to y. */
    double tx, ty;
    int xx, yy;
    if (s->vx[i] < 0) {
        tx = cross_log_time(s, wrap_inside(s->cx[i] * s->csx - s->x[i], s->halfLx, s->Lx) /
                                   s->vx[i]);
        xx = 1;
    } else {
        tx = cross_log_time(s, wrap_inside((1 + s->cx[i]) * s->csx - s->x[i], s->halfLx, s->Lx) /
                                   s->vx[i]);
        xx = 2;
    }
    if (s->prm.add_field && !s->growing) {
        /* EDMD.c:2428-2442: up through the top edge when fast enough,
         * else down through the bottom edge */
        double f = s->prm.field, vy = s->vy[i];
        double dy = wrap_inside(s->y[i] - (1 + s->cy[i]) * s->csy, s->halfLy, s->Ly);
        if (dy < 0 && vy > sqrt(2 * f * dy)) {
            ty = (-vy + sqrt(vy * vy - 2 * f * dy)) / f;
            yy = 4;
        } else {
            dy = wrap_inside(s->y[i] - s->cy[i] * s->csy, s->halfLy, s->Ly);
            ty = (-vy - sqrt(vy * vy - 2 * f * dy)) / f;
            yy = 3;
        }
    } else if (s->vy[i] < 0) {
        ty = cross_log_time(s, wrap_inside(s->cy[i] * s->csy - s->y[i], s->halfLy, s->Ly) /
                                   s->vy[i]);
        yy = 3;
    } else {
        ty = cross_log_time(s, wrap_inside((1 + s->cy[i]) * s->csy - s->y[i], s->halfLy, s->Ly) /
                                   s->vy[i]);
        yy = 4;
    }
    if (tx < ty) {
        *t_abs = s->t + tx;
        *dir = xx;
    } else {
        *t_abs = s->t + ty;
        *dir = yy;
    }
}

static void schedule_crossing(orc_ctx* s, int i) {
    /* crossingEventNormal + addCrossingEvent, EDMD.c:2405-2496 */
    double tc;
    int dir;
    predict_crossing(s, i, &tc, &dir);
    ev_node* e = &s->ev[i];
    e->j = dir;
    e->type = ORC_CELLCROSS;
    e->t = tc;
    queue_add(s, i);
}

static void schedule_collision(orc_ctx* s, int i) {
    /* collisionEventNormal + addCollisionEvent, EDMD.c:3091-3093, 3286-3297 */
    double tc;
    int partner;
    predict_collision(s, i, &tc, &partner);
    ev_node* e = &s->ev[s->n + i];
    if (partner < 0) {
        /* addWallEvent, EDMD.c:3321-3330 */
        e->j = -5 - partner;
        e->type = ORC_WALL;
        e->coll_actual = 0;
    } else {
        e->j = partner;
        e->type = ORC_COLLISION;
        e->coll_actual = s->coll[partner];
    }
    e->t = tc;
    queue_add(s, s->n + i);
}

static void schedule_special(orc_ctx* s, int slot, int type, double t) {
    /* addEventScreenshot / addEventNoise / addEventThermo, EDMD.c:4421-4440, 4814-4822 */
    ev_node* e = &s->ev[2 * s->n + slot];
    e->type = type;
    e->t = t;
    e->j = 0;
    queue_add(s, 2 * s->n + slot);
}

/* ------------------------------------------------------------ handlers */

static void physical_q(orc_ctx* s) {
    /* physicalQ, EDMD.c:5968-6013 (2D, no shear) */
    s->px = 0;
    s->py = 0;
    s->E = 0;
    for (int i = 0; i < s->n; i++) {
        s->px += s->vx[i] * s->m[i];
        s->py += s->vy[i] * s->m[i];
        double vx = s->vx[i];
        s->E += 0.5 * s->m[i] * (vx * vx + s->vy[i] * s->vy[i]);
    }
}

static void do_collision_grow(orc_ctx* s, const ev_node* ev) {
    /* doTheCollisionGrow, EDMD.c:3465-3540 (2D): unit contact normal, the
     * growth speed enters the impulse; res and delta are the compile-time
     * 1 and 0 of EDMD.c:3467-3468; no virial. */
    const double delta = 0.1 * 0;
    const double res = 0.75 * 0 + 1;
    int i = ev->i, j = ev->j;
    free_fly_grow(s, i);
    if (ev->coll_actual != s->coll[j]) {
        schedule_collision(s, i);
        return;
    }
    s->ncol++;
    free_fly_grow(s, j);
    s->coll[i]++;
    s->coll[j]++;
    double dx = s->x[j] - s->x[i];
    double dy = s->y[j] - s->y[i];
    double dvx = s->vx[j] - s->vx[i];
    double dvy = s->vy[j] - s->vy[i];
    double dvr = s->vr[i] + s->vr[j];
    min_image(s, &dx, &dy);
    double dist = sqrt(dx * dx + dy * dy);
    double dxr = dx / dist;
    double dyr = dy / dist;
    double inv_mass = 1 / (s->m[i] + s->m[j]);
    double kernel = inv_mass * (1 + res) * (dxr * dvx + dyr * dvy - dvr);
    s->vx[i] += kernel * s->m[j] * dxr - 2 * s->m[j] * inv_mass * delta * dxr;
    s->vy[i] += kernel * s->m[j] * dyr - 2 * s->m[j] * inv_mass * delta * dyr;
    s->vx[j] -= kernel * s->m[i] * dxr - 2 * s->m[i] * inv_mass * delta * dxr;
    s->vy[j] -= kernel * s->m[i] * dyr - 2 * s->m[i] * inv_mass * delta * dyr;
    queue_remove(s, i);
    queue_remove(s, j);
    schedule_crossing(s, i);
    schedule_crossing(s, j);
    queue_remove(s, s->n + j);
    schedule_collision(s, i);
    schedule_collision(s, j);
}

static void do_collision(orc_ctx* s, const ev_node* ev) {
    /* doTheCollisionNormal, EDMD.c:3542-3562, 3583-3606, 3802-3827, 3869-3879 */
    if (s->growing) {
        do_collision_grow(s, ev);
        return;
    }
    int i = ev->i, j = ev->j;
    free_fly(s, i);
    if (ev->coll_actual != s->coll[j]) {
        schedule_collision(s, i);
        return;
    }
    s->ncol++;
    free_fly(s, j);
    s->coll[i]++;
    s->coll[j]++;
    double dx = s->x[j] - s->x[i];
    double dy = s->y[j] - s->y[i];
    double dvx = s->vx[j] - s->vx[i];
    double dvy = s->vy[j] - s->vy[i];
    double inv_mass = 1 / (s->m[i] + s->m[j]);
    min_image(s, &dx, &dy);
    double kernel = inv_mass * (1 + s->prm.res) * (dx * dvx + dy * dvy);
    double ff = kernel / contact2(s->rad[i], s->rad[j]);
    double w = s->m[i] * s->m[j] * ff;
    s->vir_xx += w * dx * dx;
    s->vir_yy += w * dy * dy;
    s->vir_xy += w * dx * dy;
    s->vir_yx += w * dy * dx;
    if (s->tape && s->ntape < s->captape) {
        orc_tape_rec* r = &s->tape[s->ntape];
        r->i = i;
        r->j = j;
        r->t = s->t;
        r->ax = dx;
        r->ay = dy;
        r->cx = dvx;
        r->cy = dvy;
        r->ff = ff;
        r->g = 0;
    }
    if (s->tape) {
        s->ntape++;
        s->wsum += w * (dx * dx + dy * dy);
    }
    s->vx[i] += ff * s->m[j] * dx;
    s->vy[i] += ff * s->m[j] * dy;
    s->vx[j] -= ff * s->m[i] * dx;
    s->vy[j] -= ff * s->m[i] * dy;
    queue_remove(s, i);
    queue_remove(s, j);
    schedule_crossing(s, i);
    schedule_crossing(s, j);
    queue_remove(s, s->n + j);
    schedule_collision(s, i);
    schedule_collision(s, j);
}

static void do_crossing(orc_ctx* s, const ev_node* ev) {
    /* doTheCrossing, EDMD.c:2501-2590 (no shear; freeFly is the dispatch
     * pointer, so the radius grows here while growing) */
    int i = ev->i;
    fly(s, i);
    cell_unlink(s, i);
    switch (ev->j) {
        case 1:
            if (s->cx[i] - 1 == -1) { s->cx[i] = s->ncx - 1; s->crossX[i] -= 1; }
            else s->cx[i] -= 1;
            break;
        case 2:
            if (s->cx[i] + 1 == s->ncx) { s->cx[i] = 0; s->crossX[i] += 1; }
            else s->cx[i] += 1;
            break;
        case 3:
            if (s->cy[i] - 1 == -1) { s->cy[i] = s->ncy - 1; s->crossY[i] -= 1; }
            else s->cy[i] -= 1;
            break;
        case 4:
            if (s->cy[i] + 1 == s->ncy) { s->cy[i] = 0; s->crossY[i] += 1; }
            else s->cy[i] += 1;
            break;
    }
    cell_push_head(s, i);
    schedule_crossing(s, i);
    queue_remove(s, s->n + i);
    schedule_collision(s, i);
}

static double sign_of(double x) { return (x < 0) ? -1 : (x > 0); }

static void do_wall_grow(orc_ctx* s, const ev_node* ev) {
    /* doTheWallGrow, EDMD.c:3333-3382 (hard walls): counted as a collision,
     * no momentum to the wall; the normal velocity turned away from the
     * wall with the growth rate added twice, which way for y decided by
     * the disk's height (y > Ly - 3: the top wall). */
    int i = ev->i, xy = ev->j;
    s->ncol++;
    free_fly_grow(s, i);
    const double vr = s->vr[i];
    if (xy == 0) {
        s->vx[i] = -sign_of(s->vx[i]) * (fabs(s->vx[i]) + 2 * vr);
    } else if (xy == 1) {
        if (s->y[i] > s->Ly - 3) s->vy[i] = -s->vy[i] - 2 * vr;
        else s->vy[i] = -s->vy[i] + 2 * vr;
    } else {
        /* the circle: reflected about the normal, 2 vr added along it */
        double nx, ny;
        orc_sincos(orc_atan2(s->y[i] - s->halfLx, s->x[i] - s->halfLx), &ny, &nx);
        double mix = -(1 + s->prm.res_wall) * (s->vx[i] * nx + s->vy[i] * ny);
        s->vx[i] += (mix - 2 * vr) * nx;
        s->vy[i] += (mix - 2 * vr) * ny;
    }
    s->coll[i]++;
    queue_remove(s, i);
    schedule_crossing(s, i);
    schedule_collision(s, i);
}

static void do_wall(orc_ctx* s, const ev_node* ev) {
    if (s->growing) {
        do_wall_grow(s, ev);
        return;
    }
    /* doTheWallNormal, EDMD.c:3384-3430, 3454-3460 (no mid wall, thermal
     * wall or shear wall): counted as a collision, momentum to the wall,
     * normal velocity reversed times resW, tangential scaled by resT, then
     * new crossing and collision slots.  One deliberate difference: the
     * scaled tangential velocity is made +0.0 where the product gives -0.0
     * (resT = 0, the reference's default, times a negative velocity).  The
     * reference's crossingEventNormal (EDMD.c:2416-2425) treats -0.0 as
     * non-negative and divides by it, so its crossing time is -inf and the
     * disk is flown to t = -inf (positions NaN); +0.0 gives +inf, no
     * crossing in that direction, as for any disk at rest along it. */
    int i = ev->i, xy = ev->j;
    s->ncol++;
    free_fly(s, i);
    double rw = s->prm.res_wall, rt = s->prm.res_tan;
    if (s->tape && xy < 2) {
        /* a wall record: j = -1 - xy, the normal and tangential velocities
         * before the hit, resW and resT, and the contact coordinate */
        if (s->ntape < s->captape) {
            orc_tape_rec* r = &s->tape[s->ntape];
            r->i = i;
            r->j = -1 - xy;
            r->t = s->t;
            r->ax = xy == 0 ? s->vx[i] : s->vy[i];
            r->ay = xy == 0 ? s->vy[i] : s->vx[i];
            r->cx = rw;
            r->cy = rt;
            r->ff = xy == 0 ? s->x[i] : s->y[i];
            r->g = s->prm.add_field ? s->prm.field : 0.0;
        }
        s->ntape++;
    }
    if (xy == 0) {
        s->dp += fabs((rw + 1) * s->m[i] * s->vx[i]);
        if (s->prm.mid_wall) {
            /* EDMD.c:3393-3409: the side walls within 3 of the box's edges */
            double dpi = fabs((rw + 1) * s->m[i] * s->vx[i]);
            if (s->vx[i] < 0) {
                if (s->x[i] - 3 < 0) s->dp_left += dpi;
                else s->dp_mid += dpi;
            } else {
                if (s->x[i] + 3 > s->Lx) s->dp_right += dpi;
                else s->dp_mid += dpi;
            }
        }
        if (s->prm.thermal_wall) {
            s->vx[i] = -ref_sign(s->vx[i]) * drand0(s, s->prm.temperature / s->m[i]);
        } else {
            s->vx[i] = -rw * s->vx[i];
            s->vy[i] = rt * s->vy[i] + 0.0;
        }
    } else if (xy == 1) {
        s->dp += fabs((rw + 1) * s->m[i] * s->vy[i]);
        if (s->prm.thermal_wall) {
            s->vy[i] = -ref_sign(s->vy[i]) * drand0(s, s->prm.temperature / s->m[i]);
        } else {
            s->vy[i] = -rw * s->vy[i];
            s->vx[i] = rt * s->vx[i] + 0.0;
        }
        if (s->prm.shear_wall) s->vx[i] = (s->y[i] + 2 * s->rad[i] > s->Ly) ? 0.1 : -0.1;
    } else {
        /* the circle (EDMD.c:3441-3454): the normal velocity reversed
         * times resW, the momentum |m mix| to the wall */
        double nx, ny;
        orc_sincos(orc_atan2(s->y[i] - s->halfLx, s->x[i] - s->halfLx), &ny, &nx);
        double mix = -(1 + rw) * (s->vx[i] * nx + s->vy[i] * ny);
        if (s->prm.thermal_wall) {
            s->vx[i] = -nx * drand0(s, s->prm.temperature);
            s->vy[i] = -ny * drand0(s, s->prm.temperature);
        } else {
            s->vx[i] += mix * nx;
            s->vy[i] += mix * ny;
        }
        s->dp += fabs(s->m[i] * mix);
    }
    s->coll[i]++;
    queue_remove(s, i);
    schedule_crossing(s, i);
    schedule_collision(s, i);
}

static void do_screenshot(orc_ctx* s) {
    /* takeAScreenshot, EDMD.c:4657-4677: free-fly everyone (a state update
     * point), reschedule; the dump write has no effect on the state. */
    for (int i = 0; i < s->n; i++) fly(s, i);
    schedule_special(s, 1, ORC_SCREENSHOT, s->t + s->prm.dtime);
}

static void do_thermo(orc_ctx* s) {
    /* takeAThermo + saveThermo, EDMD.c:4679-4713, 5232-5333 */
    physical_q(s);
    double area = s->Lx * s->Ly;
    if (s->prm.wall_circle) area = M_PI * s->halfLx * s->halfLx; /* EDMD.c:5254-5256 */
    double dT = s->t - s->last_thermo;
    s->last_thermo = s->t;
    double ex = 0, ey = 0, exy = 0, eyx = 0;
    for (int i = 0; i < s->n; i++) {
        double vx = s->vx[i], vy = s->vy[i];
        ex += s->m[i] * vx * vx;
        ey += s->m[i] * vy * vy;
        exy += s->m[i] * vx * vy;
        eyx += s->m[i] * vy * vx;
    }
    orc_thermo_row row;
    row.t = s->t;
    row.ncol = s->ncol;
    row.e_per_n = s->E / s->n;
    row.pxx = -1 * s->vir_xx / (area * dT) + ex / area;
    row.pyy = -1 * s->vir_yy / (area * dT) + ey / area;
    row.pxy = -1 * s->vir_xy / (area * dT) + exy / area;
    row.pyx = -1 * s->vir_yx / (area * dT) + eyx / area;
    row.pressure = (-1 / (dT)) * (s->vir_xx + s->vir_yy) / (2 * area) + s->E / area;
    s->vir_xx = s->vir_yy = s->vir_xy = s->vir_yx = 0;
    row.pwall = 0;
    row.pmid = row.dp_lr = row.pack_left = row.pack_right = 0;
    row.temp_left = row.temp_right = row.temp_solute = 0;
    if (s->prm.mid_wall) {
        int nl = 0, nr = 0, ns = 0;
        double tsol = 0, tl = 0, tr = 0;
        for (int i = 0; i < s->n; i++) {
            if (s->type[i] == 0) {
                nl += (s->x[i] < s->halfLx) ? 1 : 0;
                nr += (s->x[i] > s->halfLx) ? 1 : 0;
                double temp = 0.5 * s->m[i] * (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i]);
                tl += (s->x[i] < s->halfLx) ? temp : 0;
                tr += (s->x[i] > s->halfLx) ? temp : 0;
            } else {
                ns++;
                tsol += 0.5 * s->m[i] * (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i]);
            }
        }
        double ratio = M_PI / (s->halfLx * s->Ly);
        double sr = s->prm.sizeratio;
        row.pmid = s->dp_mid / (dT * s->Ly);
        row.dp_lr = -(s->dp_right - s->dp_left) / (dT * s->Ly);
        row.pack_left = ratio * nl * sr * sr;
        row.pack_right = ratio * nr * sr * sr;
        row.temp_left = tl / nl;
        row.temp_right = tr / nr;
        row.temp_solute = tsol / ns;
        s->dp_mid = s->dp_right = s->dp_left = 0;
    }
    if (s->prm.wall_x || s->prm.wall_y || s->prm.wall_circle) {
        double perimeter = 0;
        if (s->prm.wall_circle) perimeter = M_PI * s->Lx;
        if (s->prm.wall_x) perimeter += 2 * s->Ly;
        if (s->prm.wall_y) perimeter += 2 * s->Lx;
        row.pwall = s->dp / (dT * perimeter);
        s->dp = 0;
    }
    if (s->nrows == s->caprows) {
        int cap = s->caprows ? 2 * s->caprows : 16;
        orc_thermo_row* r = realloc(s->rows, (size_t)cap * sizeof *r);
        if (r) { s->rows = r; s->caprows = cap; }
    }
    if (s->nrows < s->caprows) s->rows[s->nrows++] = row;
    schedule_special(s, 3, ORC_THERMO, s->t + s->prm.dtime_thermo);
}

static void do_rescale(orc_ctx* s) {
    /* addNoise with the velocity-rescale branch (noise == 2),
     * EDMD.c:4828-4832, 4880-4915, 4922. */
    physical_q(s);
    for (int i = 0; i < s->n; i++) {
        s->coll[i]++;
        fly(s, i);
        s->vx[i] /= sqrt(s->E / s->n / s->prm.temperature);
        s->vy[i] /= sqrt(s->E / s->n / s->prm.temperature);
    }
    for (int j = 0; j < s->n; j++) {
        queue_remove(s, j);
        schedule_crossing(s, j);
        queue_remove(s, s->n + j);
        schedule_collision(s, j);
    }
    schedule_special(s, 2, ORC_ADDINGNOISE, s->t + s->prm.dtnoise);
}

static void normalize_physical_q(orc_ctx* s) {
    /* normalizePhysicalQ, EDMD.c:5723-5764: remove the mean momentum (in a
     * circle, the angular momentum about the centre instead), then scale
     * to E/N = Einit. */
    if (!s->prm.wall_circle) {
        physical_q(s);
        for (int i = 0; i < s->n; i++) {
            s->vx[i] -= s->px / (s->n * s->m[i]);
            s->vy[i] -= s->py / (s->n * s->m[i]);
        }
    } else {
        const double h = s->halfLx;
        double dvT = 0, dvM = 0;
        for (int i = 0; i < s->n; i++) {
            double r = sqrt((s->x[i] - h) * (s->x[i] - h) + (s->y[i] - h) * (s->y[i] - h));
            dvT -= s->m[i] * (-s->vx[i] * (s->y[i] - h) + s->vy[i] * (s->x[i] - h));
            dvM += s->m[i] * r;
        }
        double dv = dvT / dvM;
        for (int i = 0; i < s->n; i++) {
            double r = sqrt((s->x[i] - h) * (s->x[i] - h) + (s->y[i] - h) * (s->y[i] - h));
            s->vx[i] += -dv * (s->y[i] - h) / r;
            s->vy[i] += dv * (s->x[i] - h) / r;
        }
    }
    physical_q(s);
    for (int i = 0; i < s->n; i++) {
        s->vx[i] /= sqrt(s->E / s->n / s->prm.einit);
        s->vy[i] /= sqrt(s->E / s->n / s->prm.einit);
    }
}

static void do_stop_grow(orc_ctx* s) {
    /* stopGrow, EDMD.c:4740-4785 (2D, no wells, no intruder): final free
     * flight with growth, counters and image counters reset, Normal
     * dispatch, velocity normalisation, then per disk: counter++, new
     * collision slot, new crossing slot (this interleaving decides which
     * new predictions start stale). */
    for (int i = 0; i < s->n; i++) {
        free_fly_grow(s, i);
        s->coll[i] = 0;
        s->ncol = 0;
        s->ncross = 0;
        s->crossX[i] = 0;
        s->crossY[i] = 0;
    }
    s->growing = 0;
    normalize_physical_q(s);
    for (int i = 0; i < s->n; i++) {
        s->coll[i]++;
        queue_remove(s, s->n + i);
        schedule_collision(s, i);
        queue_remove(s, i);
        schedule_crossing(s, i);
    }
}

static int step(orc_ctx* s, orc_trace* rec) {
    /* One pass of the dispatch loop, EDMD.c:534-610. */
    int e = queue_pop_min(s);
    ev_node* ev = &s->ev[e];
    if (rec) {
        rec->type = ev->type;
        rec->i = ev->i;
        rec->j = ev->j;
        rec->t = ev->t;
        rec->valid = 1;
        if (ev->type == ORC_COLLISION) rec->valid = (ev->coll_actual == s->coll[ev->j]);
    }
    s->t = ev->t;
    queue_remove(s, e);
    switch (ev->type) {
        case ORC_COLLISION: do_collision(s, ev); break;
        case ORC_CELLCROSS: s->ncross++; do_crossing(s, ev); break;
        case ORC_WALL: do_wall(s, ev); break;
        case ORC_SCREENSHOT: do_screenshot(s); break;
        case ORC_THERMO: do_thermo(s); break;
        case ORC_ADDINGNOISE: do_rescale(s); break;
        case ORC_GROWSTOP: do_stop_grow(s); break;
        default: break;
    }
    s->nevents++;
    return s->status;
}

/* ------------------------------------------------------------ public API */

void orc_default_params(orc_params* p) {
    /* EDMD.c:244-248, 372, 379, 386 */
    p->res = 1;
    p->first_screen = 1;
    p->dtime = 100;
    p->first_thermo = 1000;
    p->dtime_thermo = 100;
    p->noise = 0;
    p->dtnoise = 0.5;
    p->temperature = 1;
    p->bucket_count = 0;
    p->vr = 0;
    p->einit = 1;
    p->wall_x = 0;
    p->wall_y = 0;
    p->res_wall = 1; /* resW, EDMD.c:372 */
    p->res_tan = 0;  /* resT, EDMD.c:373 */
    p->add_field = 0;
    p->field = -0.02; /* EDMD.c:366 */
    p->damping = 0;
    p->gamma = 0.02; /* gamm, EDMD.c:382 */
    p->wall_circle = 0;
    p->mid_wall = 0;
    p->type = NULL;
    p->sizeratio = 0.4;
    p->thermal_wall = 0;
    p->shear_wall = 0;
    p->rng_seed = 5489;
}

orc_ctx* orc_create(int n, double Lx, double Ly, const double* x, const double* y,
                    const double* vx, const double* vy, const double* rad,
                    const double* m, const orc_params* prm) {
    if (n < 2 || !(Lx > 0) || !(Ly > 0)) return NULL;
    orc_ctx* s = calloc(1, sizeof *s);
    if (!s) return NULL;
    s->n = n;
    if (prm) s->prm = *prm;
    else orc_default_params(&s->prm);
    if (s->prm.damping && !(s->prm.gamma > 0)) {
        free(s);
        return NULL;
    }
    /* boxConstantHelper, EDMD.c:679-715 */
    s->Lx = Lx;
    s->Ly = Ly;
    s->halfLx = Lx / 2;
    s->halfLy = Ly / 2;
    s->ncx = (int)(s->halfLx);
    s->ncy = (int)(s->halfLy);
    if (s->ncx < 1 || s->ncy < 1) { free(s); return NULL; }
    s->csx = Lx / s->ncx;
    s->csy = Ly / s->ncy;
    s->cfx = 1 / s->csx;
    s->cfy = 1 / s->csy;
    s->n_lists = s->prm.bucket_count > 0 ? s->prm.bucket_count : n;
    s->dt_paul = 5.0 / (double)n;

    size_t nd = (size_t)n;
    s->x = malloc(nd * sizeof(double));
    s->y = malloc(nd * sizeof(double));
    s->vx = malloc(nd * sizeof(double));
    s->vy = malloc(nd * sizeof(double));
    s->tl = calloc(nd, sizeof(double));
    s->rad = malloc(nd * sizeof(double));
    s->m = malloc(nd * sizeof(double));
    s->coll = calloc(nd, sizeof(uint64_t));
    s->cx = calloc(nd, sizeof(int));
    s->cy = calloc(nd, sizeof(int));
    s->crossX = calloc(nd, sizeof(int));
    s->crossY = calloc(nd, sizeof(int));
    s->head = malloc((size_t)s->ncx * s->ncy * sizeof(int));
    s->nxt = malloc(nd * sizeof(int));
    s->prv = malloc(nd * sizeof(int));
    s->ev = calloc(2 * nd + 10, sizeof(ev_node));
    s->lists = malloc(((size_t)s->n_lists + 1) * sizeof(int));
    s->vr = calloc(nd, sizeof(double));
    s->type = calloc(nd, sizeof(int));
    if (!s->x || !s->y || !s->vx || !s->vy || !s->tl || !s->rad || !s->m || !s->coll ||
        !s->cx || !s->cy || !s->crossX || !s->crossY || !s->head || !s->nxt || !s->prv ||
        !s->ev || !s->lists || !s->vr || !s->type) {
        orc_destroy(s);
        return NULL;
    }
    /* load path, EDMD.c:1424-1442: wrap positions into the box, t = 0 */
    for (int i = 0; i < n; i++) {
        s->x[i] = wrap_post(x[i], Lx);
        s->y[i] = wrap_post(y[i], Ly);
        s->vx[i] = vx[i];
        s->vy[i] = vy[i];
        s->rad[i] = rad[i];
        s->m[i] = m[i];
        s->type[i] = s->prm.type ? s->prm.type[i] : 0;
    }
    s->prm.type = NULL; /* copied */
    mt_seed(s->mt, &s->mti, s->prm.rng_seed);
    if (s->prm.vr > 0) {
        /* growth start (EDMD.c:1770-1773): radius 0, rate vr * target */
        s->growing = 1;
        for (int i = 0; i < n; i++) {
            s->vr[i] = s->prm.vr * s->rad[i];
            s->rad[i] = 0;
        }
    }
    /* cellListInit, EDMD.c:1906-1920 */
    for (int c = 0; c < s->ncx * s->ncy; c++) s->head[c] = NIL;
    for (int i = 0; i < n; i++) {
        s->cx[i] = cell_of(s, s->x[i], s->cfx, s->ncx);
        s->cy[i] = cell_of(s, s->y[i], s->cfy, s->ncy);
        cell_push_head(s, i);
    }
    for (int k = 0; k <= s->n_lists; k++) s->lists[k] = NIL;
    return s;
}

void orc_destroy(orc_ctx* s) {
    if (!s) return;
    free(s->x); free(s->y); free(s->vx); free(s->vy); free(s->tl); free(s->rad);
    free(s->m); free(s->coll); free(s->cx); free(s->cy); free(s->crossX);
    free(s->crossY); free(s->head); free(s->nxt); free(s->prv); free(s->ev);
    free(s->lists); free(s->rows); free(s->tape); free(s->vr); free(s->type);
    free(s);
}

int orc_init_events(orc_ctx* s) {
    /* eventListInit, EDMD.c:1935-2012 (load path: Normal functions) */
    int n = s->n;
    for (int e = 0; e < 2 * n + 10; e++) {
        s->ev[e].lft = s->ev[e].rgt = s->ev[e].top = NIL;
    }
    s->root = 2 * n;
    s->ev[s->root].t = NEVER + 1;
    s->tree_min = NIL;
    double first_screen = s->prm.first_screen, first_thermo = s->prm.first_thermo;
    if (s->growing) {
        /* EDMD.c:1957-1964: output starts after the growth */
        if (first_screen < 1 / s->prm.vr) first_screen = 1 / s->prm.vr + first_screen;
        if (first_thermo < 1 / s->prm.vr) first_thermo = 1 / s->prm.vr + first_thermo;
    }
    schedule_special(s, 3, ORC_THERMO, first_thermo + 0.000000001);
    if (s->growing) schedule_special(s, 5, ORC_GROWSTOP, 1 / s->prm.vr);
    schedule_special(s, 1, ORC_SCREENSHOT, first_screen);
    if (s->prm.noise) schedule_special(s, 2, ORC_ADDINGNOISE, s->prm.dtnoise);
    for (int i = 0; i < n; i++) {
        s->ev[i].i = i;
        s->ev[n + i].i = i;
        schedule_crossing(s, i);
        schedule_collision(s, i);
    }
    return s->status;
}

/* The slot of the reference's eventList a kind of event uses for particle
 * i: i (crossing), N + i (collision, wall), 2N + 1 screenshot, 2N + 2
 * noise, 2N + 3 thermo (EDMD.c:1935-2019, 2487-2496, 3286-3297, 3321-3330,
 * 4421-4440, 4814-4822); -1 for anything else. */
static int event_slot(const orc_ctx* s, int kind, int i) {
    switch (kind) {
        case ORC_CELLCROSS: return (i >= 0 && i < s->n) ? i : -1;
        case ORC_COLLISION:
        case ORC_WALL: return (i >= 0 && i < s->n) ? s->n + i : -1;
        case ORC_SCREENSHOT: return 2 * s->n + 1;
        case ORC_ADDINGNOISE: return 2 * s->n + 2;
        case ORC_THERMO: return 2 * s->n + 3;
        default: return -1;
    }
}

int orc_add_event(orc_ctx* s, int kind, int i, int j, double t, uint64_t coll_actual) {
    /* addEventToQueue (EDMD.c:2144-2170) of the slot's node after the
     * caller has filled it, as addCollisionEvent / addCrossingEvent /
     * addWallEvent / addEvent* fill it (EDMD.c:2487-2496, 3286-3297,
     * 3321-3330, 4421-4440, 4814-4822). */
    int e = event_slot(s, kind, i);
    if (e < 0) return ORC_ERR_ARG;
    ev_node* nd = &s->ev[e];
    nd->type = kind;
    nd->t = t;
    nd->j = j;
    if (kind == ORC_COLLISION) nd->coll_actual = coll_actual;
    if (kind == ORC_WALL) nd->coll_actual = 0;
    if (e < 2 * s->n) nd->i = i;
    queue_add(s, e);
    return ORC_OK;
}

int orc_remove_event(orc_ctx* s, int kind, int i) {
    /* removeEventFromQueue, EDMD.c:2209-2221 */
    int e = event_slot(s, kind, i);
    if (e < 0) return ORC_ERR_ARG;
    queue_remove(s, e);
    return ORC_OK;
}

uint64_t orc_coll_count(const orc_ctx* s, int i) { return s->coll[i]; }

/* crossingEvent(i) / collisionEvent(i) of the Normal model (EDMD.c:
 * 2405-2496, 2829-3093): predict and file the slot; the caller has taken
 * the old one out. */
void orc_crossing_event(orc_ctx* s, int i) { schedule_crossing(s, i); }
void orc_collision_event(orc_ctx* s, int i) { schedule_collision(s, i); }
/* freeFly(p_i) at the current time, EDMD.c:4954-4990 */
void orc_free_fly(orc_ctx* s, int i) { fly(s, i); }

int orc_run(orc_ctx* s, int64_t nevents, orc_trace* trace, int64_t* done) {
    int64_t k = 0;
    int st = s->status;
    while (k < nevents && st == ORC_OK) {
        st = step(s, trace ? trace + k : NULL);
        k++;
    }
    if (done) *done = k;
    return st;
}

int orc_run_until(orc_ctx* s, double tmax, int64_t max_events, int64_t* done) {
    int64_t k = 0;
    int st = s->status;
    while (s->t <= tmax && k < max_events && st == ORC_OK) {
        st = step(s, NULL);
        k++;
    }
    if (done) *done = k;
    return st;
}

int orc_run_collisions(orc_ctx* s, uint64_t target, int64_t max_events, int64_t* done) {
    int64_t k = 0;
    int st = s->status;
    while (s->ncol < target && k < max_events && st == ORC_OK) {
        st = step(s, NULL);
        k++;
    }
    if (done) *done = k;
    return st;
}

void orc_get_radii(const orc_ctx* s, double* rad) {
    memcpy(rad, s->rad, (size_t)s->n * sizeof(double));
}

int orc_n(const orc_ctx* s) { return s->n; }
double orc_time(const orc_ctx* s) { return s->t; }
int orc_status(const orc_ctx* s) { return s->status; }

void orc_grid(const orc_ctx* s, int* ncx, int* ncy, double* csx, double* csy) {
    *ncx = s->ncx;
    *ncy = s->ncy;
    *csx = s->csx;
    *csy = s->csy;
}

void orc_get_state(const orc_ctx* s, double* x, double* y, double* vx, double* vy,
                   double* tl, uint64_t* coll, int32_t* cx, int32_t* cy) {
    for (int i = 0; i < s->n; i++) {
        if (x) x[i] = s->x[i];
        if (y) y[i] = s->y[i];
        if (vx) vx[i] = s->vx[i];
        if (vy) vy[i] = s->vy[i];
        if (tl) tl[i] = s->tl[i];
        if (coll) coll[i] = s->coll[i];
        if (cx) cx[i] = s->cx[i];
        if (cy) cy[i] = s->cy[i];
    }
}

void orc_get_events(const orc_ctx* s, double* t_cross, int32_t* dir, double* t_coll,
                    int32_t* partner, uint64_t* coll_actual) {
    int n = s->n;
    for (int i = 0; i < n; i++) {
        if (t_cross) t_cross[i] = s->ev[i].t;
        if (dir) dir[i] = s->ev[i].j;
        if (t_coll) t_coll[i] = s->ev[n + i].t;
        if (partner)
            partner[i] = s->ev[n + i].type == ORC_WALL ? ORC_WALL_PARTNER(s->ev[n + i].j)
                                                       : s->ev[n + i].j;
        if (coll_actual) coll_actual[i] = s->ev[n + i].coll_actual;
    }
}

void orc_observables(orc_ctx* s, uint64_t* counters, double* sums) {
    physical_q(s);
    if (counters) {
        counters[0] = s->ncol;
        counters[1] = s->ncross;
        counters[2] = s->nevents;
    }
    if (sums) {
        sums[0] = s->t;
        sums[1] = s->vir_xx;
        sums[2] = s->vir_yy;
        sums[3] = s->vir_xy;
        sums[4] = s->vir_yx;
        sums[5] = s->E;
        sums[6] = s->px;
        sums[7] = s->py;
    }
}

int orc_thermo_rows(const orc_ctx* s, orc_thermo_row* out, int cap) {
    int k = s->nrows < cap ? s->nrows : cap;
    if (out) memcpy(out, s->rows, (size_t)k * sizeof *out);
    return s->nrows;
}

double orc_collision_time(orc_ctx* s, int i, int j) { return coll_time(s, i, j); }

void orc_predict_collision(orc_ctx* s, int i, double* t_abs, int32_t* partner) {
    int p;
    predict_collision(s, i, t_abs, &p);
    *partner = p;
}

void orc_predict_crossing(orc_ctx* s, int i, double* t_abs, int32_t* dir) {
    int d;
    predict_crossing(s, i, t_abs, &d);
    *dir = d;
}

void orc_sync_all(orc_ctx* s, double t) {
    s->t = t;
    for (int i = 0; i < s->n; i++) fly(s, i);
}

int orc_pcf(int n, const double* x, const double* y, double Lx, double Ly,
            double dr, double max_r, double* r, double* g) {
    /* calculate_pcf, pcf.c:16-75: O(N^2) histogram, each pair counted twice,
     * normalised by the 2D shell area 2 pi r dr times density times N. */
    int nb = (int)(max_r / dr);
    double rmax = dr * nb;
    double halfLx = Lx / 2, halfLy = Ly / 2;
    for (int b = 0; b < nb; b++) {
        r[b] = (b + 0.5) * dr;
        g[b] = 0;
    }
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            double dx = x[j] - x[i];
            double dy = y[j] - y[i];
            if (dx >= halfLx) dx -= Lx;
            else if (dx < -halfLx) dx += Lx;
            if (dy >= halfLy) dy -= Ly;
            else if (dy < -halfLy) dy += Ly;
            double rr = sqrt(dx * dx + dy * dy);
            if (rr < rmax) {
                int b = (int)(rr / dr);
                if (b < nb) g[b] += 2.0;
            }
        }
    }
    double density = n / (Lx * Ly);
    for (int b = 0; b < nb; b++) {
        double shell = 2 * M_PI * r[b] * dr;
        double norm = shell * density * n;
        if (norm > 0) g[b] /= norm;
        else g[b] = 0.0;
    }
    return nb;
}

/* ------------------------------------------------------------ derivatives */

int orc_tape_enable(orc_ctx* s, int64_t capacity) {
    free(s->tape);
    s->tape = malloc((size_t)(capacity > 0 ? capacity : 1) * sizeof(orc_tape_rec));
    s->captape = capacity;
    s->ntape = 0;
    s->wsum = 0;
    return s->tape ? ORC_OK : ORC_ERR_NOMEM;
}

int64_t orc_tape_count(const orc_ctx* s) { return s->ntape; }

void orc_get_tape(const orc_ctx* s, orc_tape_rec* out) {
    int64_t k = s->ntape < s->captape ? s->ntape : s->captape;
    memcpy(out, s->tape, (size_t)k * sizeof *out);
}

void orc_tape_clear(orc_ctx* s) {
    s->ntape = 0;
    s->wsum = 0;
}

double orc_tape_wsum(const orc_ctx* s) { return s->wsum; }

/* The tangent's free flight over h: dr += h dv, or under damping (rate g)
 * dr += dv (1 - e)/g, dv *= e with e = exp(-g h) (the own exp). */
static void tan_fly(double* v, double h, double g) {
    if (g == 0) {
        v[0] += h * v[2];
        v[1] += h * v[3];
        return;
    }
    const double e = orc_exp(-g * h);
    const double ph = (1 - e) / g;
    v[0] += ph * v[2];
    v[1] += ph * v[3];
    v[2] *= e;
    v[3] *= e;
}

/* Its transpose, the adjoint's flight back over h. */
static void adj_fly(double* l, double h, double g) {
    if (g == 0) {
        l[2] += h * l[0];
        l[3] += h * l[1];
        return;
    }
    const double e = orc_exp(-g * h);
    const double ph = (1 - e) / g;
    l[2] = e * l[2] + ph * l[0];
    l[3] = e * l[3] + ph * l[1];
}

double orc_tangent(int n, const double* m, const double* tl0, const orc_tape_rec* tape,
                   int64_t ntape, double t_end, double* tan, double* dt_last, double gamma) {
    const double g = gamma;
    double* tl = malloc((size_t)n * sizeof(double));
    memcpy(tl, tl0, (size_t)n * sizeof(double));
    double dsum = 0, dtl = 0;
    for (int64_t k = 0; k < ntape; k++) {
        const orc_tape_rec* r = &tape[k];
        const int i = r->i, j = r->j;
        double* ti = tan + 4 * i;
        tan_fly(ti, r->t - tl[i], g);
        tl[i] = r->t;
        if (j < 0) {
            /* wall hit: normal axis a, tangential axis b */
            const int a = -1 - j, b = 1 - a;
            const double vn = r->ax, vt = r->ay, rw = r->cx, rt = r->cy;
            const double dt = -ti[a] / vn;
            ti[a] = -rw * ti[a];
            ti[b] += (vt - rt * vt) * dt;
            ti[2 + a] = -rw * ti[2 + a];
            ti[2 + b] = rt * ti[2 + b];
            if (r->g != 0) {
                if (a == 0) ti[3] += (rt - 1) * r->g * dt;
                else ti[3] -= (1 + rw) * r->g * dt;
            }
            continue;
        }
        double* tj = tan + 4 * j;
        tan_fly(tj, r->t - tl[j], g);
        tl[j] = r->t;
        const double ax = r->ax, ay = r->ay, cx = r->cx, cy = r->cy, ff = r->ff;
        const double mi = m[i], mj = m[j];
        const double b = ax * cx + ay * cy;
        const double px = tj[0] - ti[0], py = tj[1] - ti[1];
        const double qx = tj[2] - ti[2], qy = tj[3] - ti[3];
        const double dt = -(ax * px + ay * py) / b;
        const double rcx = px + cx * dt, rcy = py + cy * dt;
        const double kap = ff / b;
        double dff = kap * ((rcx * cx + rcy * cy) + (ax * qx + ay * qy));
        /* damping: the relative velocity decays over the shift dt */
        if (g != 0) dff -= g * ff * dt;
        const double ux = dff * ax + ff * rcx, uy = dff * ay + ff * rcy;
        ti[2] += mj * ux;
        ti[3] += mj * uy;
        tj[2] -= mi * ux;
        tj[3] -= mi * uy;
        const double sh = ff * dt;
        ti[0] -= mj * sh * ax;
        ti[1] -= mj * sh * ay;
        tj[0] += mi * sh * ax;
        tj[1] += mi * sh * ay;
        if (g != 0) {
            /* damping: the jump, applied dt early, decays less */
            ti[2] += g * mj * sh * ax;
            ti[3] += g * mj * sh * ay;
            tj[2] -= g * mi * sh * ax;
            tj[3] -= g * mi * sh * ay;
        }
        dsum += mi * mj * (ax * ax + ay * ay) * dff;
        dtl = dt;
    }
    for (int i = 0; i < n; i++) tan_fly(tan + 4 * i, t_end - tl[i], g);
    free(tl);
    if (dt_last) *dt_last = dtl;
    return dsum;
}

void orc_adjoint(int n, const double* m, const double* tl0, const orc_tape_rec* tape,
                 int64_t ntape, double t_end, double lw, double lt, const double* lam_end,
                 double* grad, double gamma) {
    const double g = gamma;
    double* tl = malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) tl[i] = t_end;
    if (lam_end) memcpy(grad, lam_end, (size_t)n * 4 * sizeof(double));
    else memset(grad, 0, (size_t)n * 4 * sizeof(double));
    int pending = 1; /* lt goes to the last pair collision */
    for (int64_t k = ntape - 1; k >= 0; k--) {
        const orc_tape_rec* r = &tape[k];
        const int i = r->i, j = r->j;
        double* li = grad + 4 * i;
        adj_fly(li, tl[i] - r->t, g);
        tl[i] = r->t;
        if (j < 0) {
            const int a = -1 - j, b = 1 - a;
            const double vn = r->ax, vt = r->ay, rw = r->cx, rt = r->cy;
            double ldt = (vt - rt * vt) * li[b];
            if (r->g != 0) {
                if (a == 0) ldt += (rt - 1) * r->g * li[3];
                else ldt -= (1 + rw) * r->g * li[3];
            }
            li[a] = -rw * li[a] - ldt / vn;
            li[2 + a] = -rw * li[2 + a];
            li[2 + b] = rt * li[2 + b];
            continue;
        }
        double* lj = grad + 4 * j;
        adj_fly(lj, tl[j] - r->t, g);
        tl[j] = r->t;
        const double ax = r->ax, ay = r->ay, cx = r->cx, cy = r->cy, ff = r->ff;
        const double mi = m[i], mj = m[j];
        const double b = ax * cx + ay * cy;
        const double kap = ff / b;
        const double lux = mj * li[2] - mi * lj[2], luy = mj * li[3] - mi * lj[3];
        double lsh = -mj * (ax * li[0] + ay * li[1]) + mi * (ax * lj[0] + ay * lj[1]);
        if (g != 0) lsh += g * (ax * lux + ay * luy);
        const double ldff = (ax * lux + ay * luy) + lw * mi * mj * (ax * ax + ay * ay);
        const double lrcx = ff * lux + kap * ldff * cx, lrcy = ff * luy + kap * ldff * cy;
        const double lqx = kap * ldff * ax, lqy = kap * ldff * ay;
        double ldt = ff * lsh + (cx * lrcx + cy * lrcy);
        if (g != 0) ldt -= g * ff * ldff;
        if (pending) {
            ldt += lt;
            pending = 0;
        }
        const double lpx = lrcx - (ldt / b) * ax, lpy = lrcy - (ldt / b) * ay;
        li[0] -= lpx;
        li[1] -= lpy;
        lj[0] += lpx;
        lj[1] += lpy;
        li[2] -= lqx;
        li[3] -= lqy;
        lj[2] += lqx;
        lj[3] += lqy;
    }
    for (int i = 0; i < n; i++) adj_fly(grad + 4 * i, tl[i] - tl0[i], g);
    free(tl);
}

/* Free flight of disk i to time T under the field f along y, or damped at
 * rate g (velocities times exp(-g h), displacements v (1 - exp(-g h))/g). */
static void replay_fly(int i, double T, double f, double g, double* x, double* y, double* vx,
                       double* vy, double* tl) {
    const double h = T - tl[i];
    if (g != 0) {
        const double e = exp(-g * h);
        x[i] += vx[i] * (1 - e) / g;
        y[i] += vy[i] * (1 - e) / g;
        vx[i] *= e;
        vy[i] *= e;
    } else {
        x[i] += h * vx[i];
        y[i] += h * vy[i] + 0.5 * f * h * h;
        vy[i] += f * h;
    }
    tl[i] = T;
}

/* Under damping a displacement that straight flight covers in tau takes
 * -log(1 - g tau)/g (NaN when it is never covered). */
static double replay_damped_time(double tau, double g) {
    if (g == 0) return tau;
    return -log(1 - g * tau) / g;
}

/* The time s from now at which a disk moving along the wall normal at v
 * (with acceleration acc) covers gap: the root near 0 of acc/2 s^2 + v s =
 * gap, in the form without cancellation. */
static double replay_wall_shift(double gap, double v, double acc) {
    if (acc == 0) return gap / v;
    const double disc = v * v + 2 * acc * gap;
    return 2 * gap / (v + (v < 0 ? -sqrt(disc) : sqrt(disc)));
}

double orc_replay(int n, double Lx, double Ly, const double* rad, const double* m,
                  double res, const orc_tape_rec* tape, int64_t ntape, double t_end,
                  double* x, double* y, double* vx, double* vy, double* tl,
                  double* t_last, double field, double gamma) {
    double S = 0, tk = 0;
    const double f = field, g = gamma;
    for (int64_t k = 0; k < ntape; k++) {
        const orc_tape_rec* r = &tape[k];
        const int i = r->i, j = r->j;
        const double tr = r->t;
        replay_fly(i, tr, f, g, x, y, vx, vy, tl);
        if (j < 0) {
            /* wall hit: the disk flies on to the taped contact coordinate
             * and reflects */
            const int a = -1 - j;
            const double sh = replay_damped_time(
                a == 0 ? replay_wall_shift(r->ff - x[i], vx[i], 0.0)
                       : replay_wall_shift(r->ff - y[i], vy[i], f), g);
            replay_fly(i, tr + sh, f, g, x, y, vx, vy, tl);
            if (a == 0) {
                vx[i] = -r->cx * vx[i];
                vy[i] = r->cy * vy[i];
            } else {
                vy[i] = -r->cx * vy[i];
                vx[i] = r->cy * vx[i];
            }
            continue;
        }
        replay_fly(j, tr, f, g, x, y, vx, vy, tl);
        double ax = x[j] - x[i], ay = y[j] - y[i];
        ax -= Lx * nearbyint(ax / Lx);
        ay -= Ly * nearbyint(ay / Ly);
        const double cx = vx[j] - vx[i], cy = vy[j] - vy[i];
        const double sig2 = 4 * rad[i] * rad[j];
        const double b = ax * cx + ay * cy, c2 = cx * cx + cy * cy;
        const double d = ax * ax + ay * ay - sig2;
        const double det = b * b - c2 * d;
        if (!(det >= 0) || !(b < 0)) return NAN;
        const double tau = d / (-b + sqrt(det)); /* smaller root, cancellation free */
        tk = tr + replay_damped_time(tau, g);
        replay_fly(i, tk, f, g, x, y, vx, vy, tl);
        replay_fly(j, tk, f, g, x, y, vx, vy, tl);
        /* the relative motion is straight in the displacement: the contact
         * vector moves with c over tau; the relative velocity is the flown
         * disks' */
        ax += tau * cx;
        ay += tau * cy;
        const double ux = vx[j] - vx[i], uy = vy[j] - vy[i];
        const double ff = (1 + res) / (m[i] + m[j]) * (ax * ux + ay * uy) / sig2;
        S += m[i] * m[j] * ff * (ax * ax + ay * ay);
        vx[i] += ff * m[j] * ax; vy[i] += ff * m[j] * ay;
        vx[j] -= ff * m[i] * ax; vy[j] -= ff * m[i] * ay;
    }
    for (int i = 0; i < n; i++) replay_fly(i, t_end, f, g, x, y, vx, vy, tl);
    if (t_last) *t_last = tk;
    return S;
}

typedef long double ld;

static void replay_fly_ld(int i, ld T, ld f, ld g, ld* x, ld* y, ld* vx, ld* vy, ld* tl) {
    const ld h = T - tl[i];
    if (g != 0) {
        const ld e = expl(-g * h);
        x[i] += vx[i] * (1 - e) / g;
        y[i] += vy[i] * (1 - e) / g;
        vx[i] *= e;
        vy[i] *= e;
    } else {
        x[i] += h * vx[i];
        y[i] += h * vy[i] + 0.5L * f * h * h;
        vy[i] += f * h;
    }
    tl[i] = T;
}

static ld replay_damped_time_ld(ld tau, ld g) {
    if (g == 0) return tau;
    return -logl(1 - g * tau) / g;
}

static ld replay_wall_shift_ld(ld gap, ld v, ld acc) {
    if (acc == 0) return gap / v;
    const ld disc = v * v + 2 * acc * gap;
    return 2 * gap / (v + (v < 0 ? -sqrtl(disc) : sqrtl(disc)));
}

static ld replay_objective_ld(int n, ld Lx, ld Ly, const double* rad, const double* m, ld res,
                              const orc_tape_rec* tape, int64_t ntape, ld t_end, ld* x, ld* y,
                              ld* vx, ld* vy, ld* tl, ld lw, ld lt, const double* lam, ld f,
                              ld g) {
    ld S = 0, tk = 0;
    for (int64_t k = 0; k < ntape; k++) {
        const orc_tape_rec* r = &tape[k];
        const int i = r->i, j = r->j;
        const ld tr = r->t;
        replay_fly_ld(i, tr, f, g, x, y, vx, vy, tl);
        if (j < 0) {
            const int a = -1 - j;
            const ld sh = replay_damped_time_ld(
                a == 0 ? replay_wall_shift_ld((ld)r->ff - x[i], vx[i], 0)
                       : replay_wall_shift_ld((ld)r->ff - y[i], vy[i], f), g);
            replay_fly_ld(i, tr + sh, f, g, x, y, vx, vy, tl);
            if (a == 0) {
                vx[i] = -(ld)r->cx * vx[i];
                vy[i] = (ld)r->cy * vy[i];
            } else {
                vy[i] = -(ld)r->cx * vy[i];
                vx[i] = (ld)r->cy * vx[i];
            }
            continue;
        }
        replay_fly_ld(j, tr, f, g, x, y, vx, vy, tl);
        ld ax = x[j] - x[i], ay = y[j] - y[i];
        ax -= Lx * nearbyintl(ax / Lx);
        ay -= Ly * nearbyintl(ay / Ly);
        const ld cx = vx[j] - vx[i], cy = vy[j] - vy[i];
        const ld sig2 = 4 * (ld)rad[i] * (ld)rad[j];
        const ld b = ax * cx + ay * cy, c2 = cx * cx + cy * cy;
        const ld d = ax * ax + ay * ay - sig2;
        const ld det = b * b - c2 * d;
        if (!(det >= 0) || !(b < 0)) return NAN;
        const ld tau = d / (-b + sqrtl(det));
        tk = tr + replay_damped_time_ld(tau, g);
        replay_fly_ld(i, tk, f, g, x, y, vx, vy, tl);
        replay_fly_ld(j, tk, f, g, x, y, vx, vy, tl);
        ax += tau * cx;
        ay += tau * cy;
        const ld ux = vx[j] - vx[i], uy = vy[j] - vy[i];
        const ld mi = m[i], mj = m[j];
        const ld ff = (1 + res) / (mi + mj) * (ax * ux + ay * uy) / sig2;
        S += mi * mj * ff * (ax * ax + ay * ay);
        vx[i] += ff * mj * ax; vy[i] += ff * mj * ay;
        vx[j] -= ff * mi * ax; vy[j] -= ff * mi * ay;
    }
    ld J = lw * S + lt * tk;
    for (int i = 0; i < n; i++) {
        replay_fly_ld(i, t_end, f, g, x, y, vx, vy, tl);
        if (lam)
            J += lam[4 * i] * x[i] + lam[4 * i + 1] * y[i] + lam[4 * i + 2] * vx[i] +
                 lam[4 * i + 3] * vy[i];
    }
    return J;
}

double orc_replay_fd(int n, double Lx, double Ly, const double* rad, const double* m,
                     double res, const orc_tape_rec* tape, int64_t ntape, double t_end,
                     const double* x, const double* y, const double* vx, const double* vy,
                     const double* tl, const double* e, double h, double lw, double lt,
                     const double* lam, double field, double gamma) {
    ld* buf = malloc((size_t)n * 5 * sizeof(ld));
    if (!buf) return NAN;
    ld J[2];
    for (int side = 0; side < 2; side++) {
        const ld hh = side == 0 ? (ld)h : -(ld)h;
        ld *X = buf, *Y = buf + n, *VX = buf + 2 * n, *VY = buf + 3 * n, *TL = buf + 4 * n;
        for (int i = 0; i < n; i++) {
            X[i] = (ld)x[i] + hh * e[4 * i];
            Y[i] = (ld)y[i] + hh * e[4 * i + 1];
            VX[i] = (ld)vx[i] + hh * e[4 * i + 2];
            VY[i] = (ld)vy[i] + hh * e[4 * i + 3];
            TL[i] = tl[i];
        }
        J[side] = replay_objective_ld(n, Lx, Ly, rad, m, res, tape, ntape, t_end, X, Y, VX, VY,
                                      TL, lw, lt, lam, (ld)field, (ld)gamma);
    }
    free(buf);
    return (double)((J[0] - J[1]) / (2 * (ld)h));
}

/*
 * HIP kernels of libedmd_hip.so for gfx950 (CDNA4, 64-lane wavefronts).
 *
 *   k_sweep        prediction sweep: one lane per particle computes the
 *                  next cell crossing and the earliest collision partner over
 *                  the 3x3 cell neighbourhood (EDMD.c:2405-2482, 2661-2723,
 *                  2829-3003; invoked for all particles at EDMD.c:2007-2012
 *                  and 4909-4915)
 *   k_tree_level   builds one level of the 64-ary minimum calendar
 *   k_run          persistent event loop: one wavefront per system pops the
 *                  earliest event, applies the reference's handlers
 *                  (EDMD.c:534-610, 2501-2590, 3542-3915, 4657-4713,
 *                  4828-4923) and re-predicts with all 64 lanes
 *   k_sync_all     free flight of every particle to the system time
 *   k_fp64_check   IEEE division / square root probe for the tests
 *
 * Arithmetic: FP64 throughout, compiled with -ffp-contract=off so that every
 * product and sum is rounded where the reference's C rounds it; division and
 * square root use the compiler's correctly rounded FP64 expansions.
 * Reference line numbers are those of src/EDMD.c.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "edmd_internal.h"
#include "edmd_types.h"

#define INF_D __builtin_huge_val()

/* Compiler barrier for memory shared between lanes of one wavefront.  The
 * lanes of a wavefront need no hardware action to see each other's global
 * or LDS stores in program order; this only keeps the compiler from moving
 * memory operations across the point. */
#define WAVE_SYNC() __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront")

/* ------------------------------------------------------------ kinematics */

/* PBC, EDMD.c:5896-5913 */
__device__ __forceinline__ void min_image(const EdmdConst& C, double& dx, double& dy) {
    if (dx >= C.halfLx) dx -= C.Lx;
    else if (dx < -C.halfLx) dx += C.Lx;
    if (dy >= C.halfLy) dy -= C.Ly;
    else if (dy < -C.halfLy) dy += C.Ly;
}

/* PBCinsideCellX/Y, EDMD.c:5922-5936 */
__device__ __forceinline__ double wrap_inside(double d, double half, double L) {
    if (d >= half) return d - L;
    if (d < -half) return d + L;
    return d;
}

/* PBCpostX/Y, EDMD.c:5948-5959 */
__device__ __forceinline__ double wrap_post(double v, double L) {
    if (v < 0) return v + L;
    if (v >= L) return v - L;
    return v;
}

/* PBCcellX/Y, EDMD.c:2110-2124 */
__device__ __forceinline__ int wrap_cell(int a, int nc) {
    if (a < 0) return a + nc;
    if (a >= nc) return a - nc;
    return a;
}

/* freeFlyNormal, EDMD.c:4954-4990, on values held in registers */
__device__ __forceinline__ void fly(const EdmdConst& C, double t, double& x, double& y,
                                    double vx, double vy, double& tl) {
    double dt = t - tl;
    tl = t;
    x += dt * vx;
    y += dt * vy;
    x = wrap_post(x, C.Lx);
    y = wrap_post(y, C.Ly);
}

/* collisionTimeNormal, EDMD.c:2661-2723: relative time until contact of the
 * synchronised owner (x1..) with candidate j last updated at tlj. */
__device__ __forceinline__ double coll_time(const EdmdConst& C, double t, double x1, double y1,
                                            double vx1, double vy1, double r1, double x2,
                                            double y2, double vx2, double vy2, double tl2,
                                            double r2, bool& overlap) {
    double lat2 = t - tl2;
    double dvx = vx2 - vx1;
    double dvy = vy2 - vy1;
    double dx = (x2 + lat2 * vx2) - x1;
    double dy = (y2 + lat2 * vy2) - y1;
    min_image(C, dx, dy);
    double b = dx * dvx + dy * dvy;
    if (b > 0) return EDMD_NEVER;
    double v2 = dvx * dvx + dvy * dvy;
    double d = dx * dx + dy * dy - 4 * r1 * r2; /* DIST2, EDMD.c:35-43 */
    double det = b * b - v2 * d;
    if (d < -0.01) overlap = true; /* EDMD.c:2708-2717 */
    if (det < 0) return EDMD_NEVER;
    return (-b - sqrt(det)) / v2;
}

/* crossingEventNormal, EDMD.c:2405-2482: absolute crossing time and
 * direction code (1 -x, 2 +x, 3 -y, 4 +y); ties go to y. */
__device__ __forceinline__ void cross_time(const EdmdConst& C, double t, double x, double y,
                                           double vx, double vy, int cx, int cy, double& tout,
                                           int& dir) {
    double tx, ty;
    int xx, yy;
    if (vx < 0) {
        tx = wrap_inside(cx * C.csx - x, C.halfLx, C.Lx) / vx;
        xx = 1;
    } else {
        tx = wrap_inside((1 + cx) * C.csx - x, C.halfLx, C.Lx) / vx;
        xx = 2;
    }
    if (vy < 0) {
        ty = wrap_inside(cy * C.csy - y, C.halfLy, C.Ly) / vy;
        yy = 3;
    } else {
        ty = wrap_inside((1 + cy) * C.csy - y, C.halfLy, C.Ly) / vy;
        yy = 4;
    }
    if (tx < ty) {
        tout = t + tx;
        dir = xx;
    } else {
        tout = t + ty;
        dir = yy;
    }
}

/* Leaf key of a particle: its earlier slot; the crossing wins ties and a NaN
 * key is parked at +inf so that it is never popped. */
__device__ __forceinline__ double leaf_key(double tcoll, double tcross) {
    double k = (tcoll < tcross) ? tcoll : tcross;
    return (k == k) ? k : INF_D;
}

/* ------------------------------------------------------------ predictions */

/* Sequential neighbour scan for one particle (one lane): strict-less minimum
 * in the reference's order, y offset outer, x offset inner, newest list
 * entry first (EDMD.c:2959-3003).  Returns the relative time; NEVER and
 * partner 0 when no candidate qualifies (EDMD.c:2830-2831). */
__device__ __forceinline__ double predict_seq(const EdmdConst& C, const EdmdArrays& A, double t,
                                              int p, double x, double y, double vx, double vy,
                                              double r, int cx, int cy, int& partner,
                                              bool& overlap) {
    double dt = EDMD_NEVER;
    int best = 0;
    for (int j = -1; j <= 1; ++j) {
        int row = wrap_cell(cy + j, C.ncy) * C.ncx;
        for (int k = -1; k <= 1; ++k) {
            int c = row + wrap_cell(cx + k, C.ncx);
            int cnt = A.ccnt[c];
            const int32_t* sl = A.cslot + (size_t)c * C.cap;
            for (int s = cnt - 1; s >= 0; --s) {
                int q = sl[s];
                if (q == p) continue;
                double tt = coll_time(C, t, x, y, vx, vy, r, A.x[q], A.y[q], A.vx[q], A.vy[q],
                                      A.tl[q], A.rad[q], overlap);
                if (dt > tt) {
                    dt = tt;
                    best = q;
                }
            }
        }
    }
    partner = best;
    return dt;
}

/* Wave-cooperative neighbour scan for one owner: the candidates of the nine
 * cells are laid out in traversal order and spread over the 64 lanes; the
 * lexicographic minimum of (time, traversal rank) over candidates with time
 * < NEVER is the reference's strict-less first minimum.  All lanes return
 * the same relative time, partner and the partner's collision counter. */
__device__ __forceinline__ double predict_wave(const EdmdConst& C, const EdmdArrays& A, int lane,
                                               double t, int p, double x, double y, double vx,
                                               double vy, double r, int cx, int cy,
                                               int& partner, uint32_t& pcoll, bool& overlap) {
    int cell = 0, cnt = 0;
    if (lane < 9) {
        int dyo = lane / 3 - 1, dxo = lane % 3 - 1;
        cell = wrap_cell(cy + dyo, C.ncy) * C.ncx + wrap_cell(cx + dxo, C.ncx);
        cnt = A.ccnt[cell];
    }
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
        int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    int excl = incl - cnt;
    int total = __shfl(incl, 8, 64);

    double bt = INF_D;
    int br = 0x7fffffff;
    int bq = 0;
    uint32_t bc = 0;
    for (int base = 0; base < total; base += 64) {
        int f = base + lane;
        int seg = 0;
#pragma unroll
        for (int s = 1; s < 9; ++s) seg += (f >= __builtin_amdgcn_readlane(excl, s)) ? 1 : 0;
        int scell = __shfl(cell, seg, 64);
        int scnt = __shfl(cnt, seg, 64);
        int sexcl = __shfl(excl, seg, 64);
        if (f < total) {
            int local = f - sexcl;
            int q = A.cslot[(size_t)scell * C.cap + (scnt - 1 - local)];
            if (q != p) {
                double tt = coll_time(C, t, x, y, vx, vy, r, A.x[q], A.y[q], A.vx[q], A.vy[q],
                                      A.tl[q], A.rad[q], overlap);
                if (tt < EDMD_NEVER && tt < bt) {
                    bt = tt;
                    br = seg * C.cap + local;
                    bq = q;
                    bc = A.coll[q];
                }
            }
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        double ot = __shfl_xor(bt, m, 64);
        int orr = __shfl_xor(br, m, 64);
        int oq = __shfl_xor(bq, m, 64);
        uint32_t oc = (uint32_t)__shfl_xor((int)bc, m, 64);
        if (ot < bt || (ot == bt && orr < br)) {
            bt = ot;
            br = orr;
            bq = oq;
            bc = oc;
        }
    }
    if (bt < EDMD_NEVER) {
        partner = bq;
        pcoll = bc;
        return bt;
    }
    partner = 0;
    pcoll = A.coll[0];
    return EDMD_NEVER;
}

/* ------------------------------------------------------------ calendar */

__device__ __forceinline__ void wave_argmin(double& k, int& i) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        double ok = __shfl_xor(k, m, 64);
        int oi = __shfl_xor(i, m, 64);
        if (ok < k || (ok == k && oi < i)) {
            k = ok;
            i = oi;
        }
    }
}

/* Earliest pending particle event of the system: minimum of the top level. */
__device__ __forceinline__ void top_min(const EdmdConst& C, const EdmdArrays& A, int lane,
                                        double& key, int& leaf) {
    const int top = C.nlev;
    double k = INF_D;
    int i = 0x7fffffff;
    if (lane < C.lsize[top]) {
        if (top == 0) {
            k = A.key[lane];
            i = lane;
        } else {
            k = A.lkey[top - 1][lane];
            i = A.lidx[top - 1][lane];
        }
    }
    wave_argmin(k, i);
    key = k;
    leaf = i;
}

/* Stores the new key of leaf p and recomputes the minima on its path to the
 * top.  The freshly computed child value is taken from registers, so the
 * path needs no read-after-write through memory. */
__device__ __forceinline__ void ascend(const EdmdConst& C, const EdmdArrays& A, int lane, int p,
                                       double kp) {
    if (lane == 0) A.key[p] = kp;
    double v = kp;
    int vi = p;
    int child = p;
    for (int l = 1; l <= C.nlev; ++l) {
        int g = child >> 6;
        int c = (g << 6) + lane;
        double k = INF_D;
        int i = 0x7fffffff;
        if (c < C.lsize[l - 1]) {
            if (c == child) {
                k = v;
                i = vi;
            } else if (l == 1) {
                k = A.key[c];
                i = c;
            } else {
                k = A.lkey[l - 2][c];
                i = A.lidx[l - 2][c];
            }
        }
        wave_argmin(k, i);
        v = k;
        vi = i;
        if (lane == 0) {
            A.lkey[l - 1][g] = v;
            A.lidx[l - 1][g] = vi;
        }
        child = g;
    }
    WAVE_SYNC();
}

/* Sequential minimum over the children of node g of level l (l >= 1):
 * strict less in child order, i.e. ties go to the smaller leaf index. */
__device__ __forceinline__ void node_min(const EdmdConst& C, const EdmdArrays& A, int l, int g) {
    int lo = g << 6;
    int hi = lo + 64;
    if (hi > C.lsize[l - 1]) hi = C.lsize[l - 1];
    double k = INF_D;
    int idx = 0x7fffffff;
    for (int c = lo; c < hi; ++c) {
        double kc;
        int ic;
        if (l == 1) {
            kc = A.key[c];
            ic = c;
        } else {
            kc = A.lkey[l - 2][c];
            ic = A.lidx[l - 2][c];
        }
        if (kc < k || idx == 0x7fffffff) {
            k = kc;
            idx = ic;
        }
    }
    A.lkey[l - 1][g] = k;
    A.lidx[l - 1][g] = idx;
}

/* ------------------------------------------------------------ sweep */

/* One particle's full prediction (crossing + collision) written to its two
 * slots and its leaf key; the particle is assumed synchronised at t. */
__device__ __forceinline__ bool sweep_one(const EdmdConst& C, const EdmdArrays& A, double t, int p) {
    double x = A.x[p], y = A.y[p], vx = A.vx[p], vy = A.vy[p], r = A.rad[p];
    int cx = A.cx[p], cy = A.cy[p];
    double tc;
    int dir;
    cross_time(C, t, x, y, vx, vy, cx, cy, tc, dir);
    bool ov = false;
    int partner;
    double dt = predict_seq(C, A, t, p, x, y, vx, vy, r, cx, cy, partner, ov);
    double tk = t + dt;
    A.tcross[p] = tc;
    A.dir[p] = dir;
    A.tcoll[p] = tk;
    A.partner[p] = partner;
    A.cact[p] = A.coll[partner];
    A.key[p] = leaf_key(tk, tc);
    return ov;
}

__global__ __launch_bounds__(256) void k_sweep(const EdmdSys* __restrict__ systems) {
    const EdmdSys* S = systems + blockIdx.y;
    const int n = S->c.n;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double t = S->a.st->t;
    bool ov = sweep_one(S->c, S->a, t, p);
    if (ov) atomicExch(&S->a.st->status, EDMD_DEV_OVERLAP);
}

__global__ __launch_bounds__(256) void k_tree_level(const EdmdSys* __restrict__ systems, int l) {
    const EdmdSys* S = systems + blockIdx.y;
    if (l > S->c.nlev) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= S->c.lsize[l]) return;
    node_min(S->c, S->a, l, g);
}

__global__ __launch_bounds__(256) void k_sync_all(const EdmdSys* __restrict__ systems) {
    const EdmdSys* S = systems + blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S->c.n) return;
    const double t = S->a.st->t;
    double x = S->a.x[p], y = S->a.y[p], tl = S->a.tl[p];
    fly(S->c, t, x, y, S->a.vx[p], S->a.vy[p], tl);
    S->a.x[p] = x;
    S->a.y[p] = y;
    S->a.tl[p] = tl;
}

/* ------------------------------------------------------------ event loop */

struct LoopRegs {
    double t;
    double next_screen, next_thermo, next_noise, last_thermo;
    double vir[4];
    uint64_t ncol, ncross, nevents;
    int64_t ntrace;
    int32_t nthermo;
    int32_t status, err_i, err_j;
};

__device__ __forceinline__ void trace_push(const EdmdConst& C, const EdmdArrays& A, LoopRegs& R,
                                           int lane, int type, int i, int j, int valid) {
    if (R.ntrace < C.trace_cap) {
        if (lane == 0) {
            EdmdTraceRec rec;
            rec.type = type;
            rec.i = i;
            rec.j = j;
            rec.valid = valid;
            rec.t = R.t;
            A.trace[R.ntrace] = rec;
        }
        R.ntrace++;
    }
}

/* Ordered sum over all particles of f(q), accumulated in particle order as
 * the reference's serial loops do (physicalQ, EDMD.c:5968-6013). */
template <class F>
__device__ __forceinline__ double ordered_sum(int n, int lane, F f) {
    double acc = 0;
    for (int base = 0; base < n; base += 64) {
        int q = base + lane;
        double v = (q < n) ? f(q) : 0.0;
        int lim = n - base < 64 ? n - base : 64;
        for (int l = 0; l < lim; ++l) {
            double vl = __shfl(v, l, 64);
            acc += vl;
        }
    }
    return acc;
}

__device__ void ev_screenshot(const EdmdConst& C, const EdmdArrays& A, int lane, LoopRegs& R) {
    /* takeAScreenshot, EDMD.c:4657-4677: free-fly every particle */
    for (int q = lane; q < C.n; q += 64) {
        double x = A.x[q], y = A.y[q], tl = A.tl[q];
        fly(C, R.t, x, y, A.vx[q], A.vy[q], tl);
        A.x[q] = x;
        A.y[q] = y;
        A.tl[q] = tl;
    }
    WAVE_SYNC();
    R.next_screen = R.t + C.dtime;
}

__device__ void ev_thermo(const EdmdConst& C, const EdmdArrays& A, int lane, LoopRegs& R) {
    /* takeAThermo + saveThermo, EDMD.c:4679-4713, 5232-5333 */
    const double* vx = A.vx;
    const double* vy = A.vy;
    const double* m = A.m;
    double E = ordered_sum(C.n, lane, [&](int q) {
        double a = vx[q];
        return 0.5 * m[q] * (a * a + vy[q] * vy[q]);
    });
    double ex = ordered_sum(C.n, lane, [&](int q) { return m[q] * vx[q] * vx[q]; });
    double ey = ordered_sum(C.n, lane, [&](int q) { return m[q] * vy[q] * vy[q]; });
    double exy = ordered_sum(C.n, lane, [&](int q) { return m[q] * vx[q] * vy[q]; });
    double eyx = ordered_sum(C.n, lane, [&](int q) { return m[q] * vy[q] * vx[q]; });
    double area = C.Lx * C.Ly;
    double dT = R.t - R.last_thermo;
    R.last_thermo = R.t;
    if (R.nthermo < C.thermo_cap) {
        if (lane == 0) {
            EdmdThermoRow row;
            row.t = R.t;
            row.ncol = R.ncol;
            row.e_per_n = E / C.n;
            row.pxx = -1 * R.vir[0] / (area * dT) + ex / area;
            row.pyy = -1 * R.vir[1] / (area * dT) + ey / area;
            row.pxy = -1 * R.vir[2] / (area * dT) + exy / area;
            row.pyx = -1 * R.vir[3] / (area * dT) + eyx / area;
            row.pressure = (-1 / (dT)) * (R.vir[0] + R.vir[1]) / (2 * area) + E / area;
            A.thermo[R.nthermo] = row;
        }
        R.nthermo++;
    }
    R.vir[0] = R.vir[1] = R.vir[2] = R.vir[3] = 0;
    R.next_thermo = R.t + C.dtime_thermo;
}

/* Rebuilds every level of the calendar after a full re-prediction. */
__device__ void rebuild_tree(const EdmdConst& C, const EdmdArrays& A, int lane) {
    for (int l = 1; l <= C.nlev; ++l) {
        for (int g = lane; g < C.lsize[l]; g += 64) node_min(C, A, l, g);
        WAVE_SYNC();
    }
}

__device__ void ev_rescale(const EdmdConst& C, const EdmdArrays& A, int lane, LoopRegs& R) {
    /* addNoise, velocity-rescale branch (noise == 2): EDMD.c:4828-4832,
     * 4880-4915, 4922 */
    const double* vx = A.vx;
    const double* vy = A.vy;
    const double* m = A.m;
    double E = ordered_sum(C.n, lane, [&](int q) {
        double a = vx[q];
        return 0.5 * m[q] * (a * a + vy[q] * vy[q]);
    });
    for (int q = lane; q < C.n; q += 64) {
        A.coll[q] = A.coll[q] + 1;
        double x = A.x[q], y = A.y[q], tl = A.tl[q];
        double a = A.vx[q], b = A.vy[q];
        fly(C, R.t, x, y, a, b, tl);
        a /= sqrt(E / C.n / C.temperature);
        b /= sqrt(E / C.n / C.temperature);
        A.x[q] = x;
        A.y[q] = y;
        A.tl[q] = tl;
        A.vx[q] = a;
        A.vy[q] = b;
    }
    WAVE_SYNC();
    bool ov = false;
    for (int q = lane; q < C.n; q += 64) ov |= sweep_one(C, A, R.t, q);
    if (__ballot(ov) != 0ull) R.status = EDMD_DEV_OVERLAP;
    WAVE_SYNC();
    rebuild_tree(C, A, lane);
    R.next_noise = R.t + C.dtnoise;
}

/* Collision slot of owner p re-predicted by the whole wave and written. */
__device__ __forceinline__ double repredict_coll(const EdmdConst& C, const EdmdArrays& A,
                                                 int lane, LoopRegs& R, int p, double x, double y,
                                                 double vx, double vy, double r, int cx, int cy) {
    bool ov = false;
    int partner;
    uint32_t pc;
    double dt = predict_wave(C, A, lane, R.t, p, x, y, vx, vy, r, cx, cy, partner, pc, ov);
    if (__ballot(ov) != 0ull && R.status == EDMD_DEV_OK) {
        R.status = EDMD_DEV_OVERLAP;
        R.err_i = p;
    }
    double tk = R.t + dt;
    if (lane == 0) {
        A.tcoll[p] = tk;
        A.partner[p] = partner;
        A.cact[p] = pc;
    }
    return tk;
}

__device__ void ev_collision(const EdmdConst& C, const EdmdArrays& A, int lane, LoopRegs& R, int i) {
    /* doTheCollisionNormal, EDMD.c:3542-3562, 3583-3606, 3802-3827, 3869-3879 */
    const int j = A.partner[i];
    const uint32_t ca = A.cact[i];
    double xi = A.x[i], yi = A.y[i], vxi = A.vx[i], vyi = A.vy[i], tli = A.tl[i];
    double ri = A.rad[i], mi = A.m[i];
    int cxi = A.cx[i], cyi = A.cy[i];
    uint32_t colli = A.coll[i];
    double tci = A.tcross[i];
    double xj = A.x[j], yj = A.y[j], vxj = A.vx[j], vyj = A.vy[j], tlj = A.tl[j];
    double rj = A.rad[j], mj = A.m[j];
    int cxj = A.cx[j], cyj = A.cy[j];
    uint32_t collj = A.coll[j];

    fly(C, R.t, xi, yi, vxi, vyi, tli);
    const bool valid = (ca == collj);
    trace_push(C, A, R, lane, EDMD_EV_COLLISION, i, j, valid ? 1 : 0);
    if (!valid) {
        /* stale prediction: free flight of i and a new collision slot */
        if (lane == 0) {
            A.x[i] = xi;
            A.y[i] = yi;
            A.tl[i] = tli;
        }
        WAVE_SYNC();
        double tk = repredict_coll(C, A, lane, R, i, xi, yi, vxi, vyi, ri, cxi, cyi);
        ascend(C, A, lane, i, leaf_key(tk, tci));
        return;
    }
    R.ncol++;
    fly(C, R.t, xj, yj, vxj, vyj, tlj);
    colli++;
    collj++;
    double dx = xj - xi;
    double dy = yj - yi;
    double dvx = vxj - vxi;
    double dvy = vyj - vyi;
    double inv_mass = 1 / (mi + mj);
    min_image(C, dx, dy);
    double kernel = inv_mass * (1 + C.res) * (dx * dvx + dy * dvy);
    double ff = kernel / (4 * ri * rj);
    double w = mi * mj * ff;
    R.vir[0] += w * dx * dx;
    R.vir[1] += w * dy * dy;
    R.vir[2] += w * dx * dy;
    R.vir[3] += w * dy * dx;
    vxi += ff * mj * dx;
    vyi += ff * mj * dy;
    vxj -= ff * mi * dx;
    vyj -= ff * mi * dy;
    if (lane == 0) {
        A.x[i] = xi; A.y[i] = yi; A.vx[i] = vxi; A.vy[i] = vyi; A.tl[i] = tli; A.coll[i] = colli;
        A.x[j] = xj; A.y[j] = yj; A.vx[j] = vxj; A.vy[j] = vyj; A.tl[j] = tlj; A.coll[j] = collj;
    }
    WAVE_SYNC();
    double tci2, tcj2;
    int di, dj;
    cross_time(C, R.t, xi, yi, vxi, vyi, cxi, cyi, tci2, di);
    cross_time(C, R.t, xj, yj, vxj, vyj, cxj, cyj, tcj2, dj);
    if (lane == 0) {
        A.tcross[i] = tci2;
        A.dir[i] = di;
        A.tcross[j] = tcj2;
        A.dir[j] = dj;
    }
    double tki = repredict_coll(C, A, lane, R, i, xi, yi, vxi, vyi, ri, cxi, cyi);
    double tkj = repredict_coll(C, A, lane, R, j, xj, yj, vxj, vyj, rj, cxj, cyj);
    WAVE_SYNC();
    ascend(C, A, lane, i, leaf_key(tki, tci2));
    ascend(C, A, lane, j, leaf_key(tkj, tcj2));
}

__device__ void ev_crossing(const EdmdConst& C, const EdmdArrays& A, int lane, LoopRegs& R, int p) {
    /* doTheCrossing, EDMD.c:2501-2590 */
    const int d = A.dir[p];
    double x = A.x[p], y = A.y[p], vx = A.vx[p], vy = A.vy[p], tl = A.tl[p], r = A.rad[p];
    int cx = A.cx[p], cy = A.cy[p];
    int crx = A.crx[p], cry = A.cry[p];
    trace_push(C, A, R, lane, EDMD_EV_CELLCROSS, p, d, 1);
    R.ncross++;
    fly(C, R.t, x, y, vx, vy, tl);

    /* removeFromCell: drop p, keeping the order of the others */
    const int c0 = cy * C.ncx + cx;
    const int n0 = A.ccnt[c0];
    int v = (lane < n0) ? A.cslot[(size_t)c0 * C.cap + lane] : -1;
    uint64_t hit = __ballot(v == p);
    int pos = hit ? (int)__builtin_ctzll(hit) : n0;
    if (lane > pos && lane < n0) A.cslot[(size_t)c0 * C.cap + lane - 1] = v;
    if (lane == 0) A.ccnt[c0] = n0 - 1;

    switch (d) {
        case 1:
            if (cx - 1 == -1) { cx = C.ncx - 1; crx -= 1; } else cx -= 1;
            break;
        case 2:
            if (cx + 1 == C.ncx) { cx = 0; crx += 1; } else cx += 1;
            break;
        case 3:
            if (cy - 1 == -1) { cy = C.ncy - 1; cry -= 1; } else cy -= 1;
            break;
        default:
            if (cy + 1 == C.ncy) { cy = 0; cry += 1; } else cy += 1;
            break;
    }
    WAVE_SYNC();
    const int c1 = cy * C.ncx + cx;
    const int n1 = A.ccnt[c1];
    if (n1 >= C.cap) {
        if (R.status == EDMD_DEV_OK) {
            R.status = EDMD_DEV_CELL_OVERFLOW;
            R.err_i = p;
            R.err_j = c1;
        }
        return;
    }
    if (lane == 0) {
        A.cslot[(size_t)c1 * C.cap + n1] = p; /* newest entry: visited first */
        A.ccnt[c1] = n1 + 1;
        A.x[p] = x; A.y[p] = y; A.tl[p] = tl;
        A.cx[p] = cx; A.cy[p] = cy; A.crx[p] = crx; A.cry[p] = cry;
    }
    WAVE_SYNC();
    double tc;
    int dir;
    cross_time(C, R.t, x, y, vx, vy, cx, cy, tc, dir);
    if (lane == 0) {
        A.tcross[p] = tc;
        A.dir[p] = dir;
    }
    double tk = repredict_coll(C, A, lane, R, p, x, y, vx, vy, r, cx, cy);
    WAVE_SYNC();
    ascend(C, A, lane, p, leaf_key(tk, tc));
}

/* Persistent event loop, one wavefront per system (blockIdx.x). */
__global__ __launch_bounds__(64) void k_run(const EdmdSys* __restrict__ systems, EdmdRunArgs ra) {
    const EdmdConst& C = systems[blockIdx.x].c;
    const EdmdArrays& A = systems[blockIdx.x].a;
    const int lane = threadIdx.x;
    EdmdState* st = A.st;
    LoopRegs R;
    R.t = st->t;
    R.next_screen = st->next_screen;
    R.next_thermo = st->next_thermo;
    R.next_noise = st->next_noise;
    R.last_thermo = st->last_thermo;
    for (int k = 0; k < 4; ++k) R.vir[k] = st->vir[k];
    R.ncol = st->ncol;
    R.ncross = st->ncross;
    R.nevents = st->nevents;
    R.ntrace = st->ntrace;
    R.nthermo = st->nthermo;
    R.status = st->status;
    R.err_i = st->err_i;
    R.err_j = st->err_j;
    int reason = EDMD_STOP_NONE;
    int last_type = st->last_type;

    for (;;) {
        if (R.status != EDMD_DEV_OK) { reason = EDMD_STOP_ERROR; break; }
        if ((int64_t)R.nevents >= ra.max_events) { reason = EDMD_STOP_EVENTS; break; }
        if (R.ncol >= ra.max_coll) { reason = EDMD_STOP_COLLISIONS; break; }
        if (!(R.t <= ra.tmax)) { reason = EDMD_STOP_TIME; break; }
        double key;
        int p;
        top_min(C, A, lane, key, p);
        double tsp = R.next_screen;
        int sp = EDMD_EV_SCREENSHOT;
        if (R.next_thermo < tsp) { tsp = R.next_thermo; sp = EDMD_EV_THERMO; }
        if (R.next_noise < tsp) { tsp = R.next_noise; sp = EDMD_EV_ADDINGNOISE; }
        if (!(key < tsp)) {
            if (!(tsp < INF_D)) { reason = EDMD_STOP_EMPTY; break; }
            R.t = tsp;
            trace_push(C, A, R, lane, sp, 0, 0, 1);
            if (sp == EDMD_EV_SCREENSHOT) ev_screenshot(C, A, lane, R);
            else if (sp == EDMD_EV_THERMO) ev_thermo(C, A, lane, R);
            else ev_rescale(C, A, lane, R);
            R.nevents++;
            last_type = sp;
            if (ra.stop_mask & (1 << sp)) { reason = EDMD_STOP_SPECIAL; break; }
            continue;
        }
        R.t = key;
        const double tk = A.tcoll[p];
        const double tc = A.tcross[p];
        if (tk < tc) {
            ev_collision(C, A, lane, R, p);
            last_type = EDMD_EV_COLLISION;
        } else {
            ev_crossing(C, A, lane, R, p);
            last_type = EDMD_EV_CELLCROSS;
        }
        R.nevents++;
        if (ra.stop_mask & (1 << last_type)) { reason = EDMD_STOP_SPECIAL; break; }
    }

    if (lane == 0) {
        st->t = R.t;
        st->next_screen = R.next_screen;
        st->next_thermo = R.next_thermo;
        st->next_noise = R.next_noise;
        st->last_thermo = R.last_thermo;
        for (int k = 0; k < 4; ++k) st->vir[k] = R.vir[k];
        st->ncol = R.ncol;
        st->ncross = R.ncross;
        st->nevents = R.nevents;
        st->ntrace = R.ntrace;
        st->nthermo = R.nthermo;
        st->status = R.status;
        st->err_i = R.err_i;
        st->err_j = R.err_j;
        st->stop_reason = reason;
        st->last_type = last_type;
    }
}

/* ------------------------------------------------------------ single steps */

/* findNextEvent, EDMD.c:2313-2327: the earliest pending event, particle or
 * scheduled, without handling it. */
__global__ __launch_bounds__(64) void k_peek(const EdmdSys* __restrict__ systems, EdmdPeek* out) {
    const EdmdConst& C = systems[blockIdx.x].c;
    const EdmdArrays& A = systems[blockIdx.x].a;
    const int lane = threadIdx.x;
    const EdmdState* st = A.st;
    double key;
    int p;
    top_min(C, A, lane, key, p);
    double tsp = st->next_screen;
    int sp = EDMD_EV_SCREENSHOT;
    if (st->next_thermo < tsp) { tsp = st->next_thermo; sp = EDMD_EV_THERMO; }
    if (st->next_noise < tsp) { tsp = st->next_noise; sp = EDMD_EV_ADDINGNOISE; }
    if (lane != 0) return;
    EdmdPeek r;
    if (!(key < tsp)) {
        r.t = tsp; r.type = sp; r.i = 0; r.j = 0; r.coll_actual = 0;
    } else if (A.tcoll[p] < A.tcross[p]) {
        r.t = key; r.type = EDMD_EV_COLLISION; r.i = p; r.j = A.partner[p]; r.coll_actual = A.cact[p];
    } else {
        r.t = key; r.type = EDMD_EV_CELLCROSS; r.i = p; r.j = A.dir[p]; r.coll_actual = 0;
    }
    out[blockIdx.x] = r;
}

/* collisionTimeNormal(p1, p2) at the system time, EDMD.c:2661-2723; p1 is
 * taken at its stored position as the reference does. */
__global__ void k_pair_time(const EdmdSys* __restrict__ systems, int i, int j, double* out) {
    const EdmdConst& C = systems[0].c;
    const EdmdArrays& A = systems[0].a;
    bool ov = false;
    double t = A.st->t;
    out[0] = coll_time(C, t, A.x[i], A.y[i], A.vx[i], A.vy[i], A.rad[i], A.x[j], A.y[j], A.vx[j],
                       A.vy[j], A.tl[j], A.rad[j], ov);
    out[1] = ov ? 1.0 : 0.0;
}

/* collisionEvent(i) / crossingEvent(i) for one particle (which: bit 0 =
 * crossing, bit 1 = collision) at the system time, then the calendar path
 * of i is updated. */
__global__ __launch_bounds__(64) void k_repredict(const EdmdSys* __restrict__ systems, int i, int which) {
    const EdmdConst& C = systems[0].c;
    const EdmdArrays& A = systems[0].a;
    const int lane = threadIdx.x;
    LoopRegs R;
    R.t = A.st->t;
    R.status = EDMD_DEV_OK;
    double x = A.x[i], y = A.y[i], vx = A.vx[i], vy = A.vy[i], r = A.rad[i];
    int cx = A.cx[i], cy = A.cy[i];
    double tc = A.tcross[i], tk = A.tcoll[i];
    if (which & 1) {
        int dir;
        cross_time(C, R.t, x, y, vx, vy, cx, cy, tc, dir);
        if (lane == 0) {
            A.tcross[i] = tc;
            A.dir[i] = dir;
        }
    }
    if (which & 2) tk = repredict_coll(C, A, lane, R, i, x, y, vx, vy, r, cx, cy);
    WAVE_SYNC();
    ascend(C, A, lane, i, leaf_key(tk, tc));
    if (lane == 0 && R.status != EDMD_DEV_OK) {
        A.st->status = R.status;
        A.st->err_i = i;
    }
}

/* freeFlyNormal(p) at the system time, EDMD.c:4954-4990. */
__global__ void k_fly_one(const EdmdSys* __restrict__ systems, int i) {
    const EdmdConst& C = systems[0].c;
    const EdmdArrays& A = systems[0].a;
    double x = A.x[i], y = A.y[i], tl = A.tl[i];
    fly(C, A.st->t, x, y, A.vx[i], A.vy[i], tl);
    A.x[i] = x;
    A.y[i] = y;
    A.tl[i] = tl;
}

__global__ void k_fp64_check(int n, const double* __restrict__ a, const double* __restrict__ b,
                             double* __restrict__ q, double* __restrict__ s) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    q[k] = a[k] / b[k];
    s[k] = sqrt(a[k]);
}

/* ------------------------------------------------------------ launchers */

hipError_t edmd_launch_sweep(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s) {
    if (nsys <= 0 || max_n <= 0) return hipSuccess;
    dim3 grid((max_n + 255) / 256, nsys);
    hipLaunchKernelGGL(k_sweep, grid, dim3(256), 0, s, d_sys);
    return hipGetLastError();
}

hipError_t edmd_launch_tree(const EdmdSys* d_sys, int nsys, const int* max_lsize, int max_nlev,
                            hipStream_t s) {
    for (int l = 1; l <= max_nlev; ++l) {
        dim3 grid((max_lsize[l] + 255) / 256, nsys);
        hipLaunchKernelGGL(k_tree_level, grid, dim3(256), 0, s, d_sys, l);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t edmd_launch_sync_all(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s) {
    dim3 grid((max_n + 255) / 256, nsys);
    hipLaunchKernelGGL(k_sync_all, grid, dim3(256), 0, s, d_sys);
    return hipGetLastError();
}

hipError_t edmd_launch_run(const EdmdSys* d_sys, int nsys, EdmdRunArgs ra, hipStream_t s) {
    hipLaunchKernelGGL(k_run, dim3(nsys), dim3(64), 0, s, d_sys, ra);
    return hipGetLastError();
}

hipError_t edmd_launch_peek(const EdmdSys* d_sys, int nsys, EdmdPeek* out, hipStream_t s) {
    hipLaunchKernelGGL(k_peek, dim3(nsys), dim3(64), 0, s, d_sys, out);
    return hipGetLastError();
}

hipError_t edmd_launch_pair_time(const EdmdSys* d_sys, int i, int j, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_pair_time, dim3(1), dim3(1), 0, s, d_sys, i, j, out);
    return hipGetLastError();
}

hipError_t edmd_launch_repredict(const EdmdSys* d_sys, int i, int which, hipStream_t s) {
    hipLaunchKernelGGL(k_repredict, dim3(1), dim3(64), 0, s, d_sys, i, which);
    return hipGetLastError();
}

hipError_t edmd_launch_fly_one(const EdmdSys* d_sys, int i, hipStream_t s) {
    hipLaunchKernelGGL(k_fly_one, dim3(1), dim3(1), 0, s, d_sys, i);
    return hipGetLastError();
}

hipError_t edmd_launch_fp64_check(int n, const double* a, const double* b, double* q, double* sq,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_fp64_check, dim3((n + 255) / 256), dim3(256), 0, s, n, a, b, q, sq);
    return hipGetLastError();
}

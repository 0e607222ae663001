/*
 * Derivatives of the event map for a fixed collision sequence (gfx950).
 * The reference computes no derivatives; these kernels differentiate the
 * collision rule it applies (EDMD.c:3583-3606) and free flight
 * (EDMD.c:2381-2391) along the sequence of valid collisions that k_run
 * records on the tape (one EdmdTapeRec per collision).
 *
 * For a taped collision of i and j at time t with contact vector a = r_j - r_i,
 * relative velocity c = v_j - v_i, b = a.c and impulse factor ff, and tangent
 * (dr, dv) of both disks flown to t (p = dr_j - dr_i, q = dv_j - dv_i):
 *   dt  = -(a.p) / b                      shift of the collision time
 *   rc  = p + c dt                        change of the contact vector
 *   dff = (ff / b) (rc.c + a.q)
 *   u   = dff a + ff rc
 *   dv_i += m_j u,  dv_j -= m_i u
 *   dr_i -= m_j ff dt a,  dr_j += m_i ff dt a   (velocity jump applied dt early)
 *   d(m_i m_j ff |a|^2) = m_i m_j |a|^2 dff     (|a| stays at contact)
 * Free flight moves the tangent as dr += (t' - t) dv.  A flat wall's hit
 * (record j = -1 - xy, normal / tangential velocity vn, vt before it,
 * resW, resT) maps the disk's tangent flown to the hit, with n / s its
 * normal / tangential components, as
 *   dt = -dr_n / vn, dr_n' = -resW dr_n, dr_s' = dr_s + (vt - resT vt) dt,
 *   dv_n' = -resW dv_n, dv_s' = resT dv_s
 * (the reflected disk taken back to the unperturbed hit time); under the
 * field g along y the velocity along y also gains the field's pull over dt,
 * (resT - 1) g dt for a wall normal to x and -(1 + resW) g dt for one
 * normal to y.  Under damping at rate gamma the tangent's flight over h is
 * dr += dv (1 - e)/gamma, dv *= e with e = exp(-gamma h) (edmd_exp), a
 * pair's dff loses gamma ff dt and its velocity jump, applied dt early,
 * decays less (dv_i += gamma m_j ff dt a, dv_j -= gamma m_i ff dt a); wall
 * maps are unchanged.  k_adjoint applies the transpose of the same steps in
 * reverse order.
 *
 *   k_tangent      forward propagation of one tangent vector per system
 *   k_adjoint      reverse propagation of one cotangent vector per system
 *   k_deriv_init   reference times and accumulators before a pass
 *   k_deriv_flush  final free flight (tangent to the end time, cotangent back
 *                  to the particles' local times when taping began)
 *   k_seg_begin    per-system event bound and tape reset for one checkpoint
 *                  segment
 *
 * One wavefront per system.  The collision records are read 64 at a time
 * (one per lane, coalesced) and applied in tape order by the whole wavefront;
 * the vector, its reference times and the masses of the system sit in LDS
 * when they fit, in HBM otherwise.  Every operation is the one the oracle
 * performs, in the same order, with no contraction, so results are bitwise
 * reproducible.
 */
#define EDMD_DEVICE_ARRAYS 1
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "edmd_internal.h"
#include "edmd_math.h"
#include "edmd_types.h"

#define WAVE_SYNC() __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront")

__device__ __forceinline__ double lane_bcast(double v, int l) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)b, l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int lane_bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

/* Where the per-particle vector of one system lives during a pass. */
struct DerivView {
    double* v;  /* 4 per particle */
    double* tl; /* reference time of v */
    double* m;  /* masses (LDS copy) or nullptr: read from the particle records */
    const EdmdPart* p;
    __device__ __forceinline__ double mass(int q) const { return m ? m[q] : p[q].m; }
};

/* Copies the system's vector into LDS when `lds` is set. */
__device__ __forceinline__ DerivView view_open(const EdmdConst& C, const EdmdArrays& A, int lane,
                                               double* lds) {
    DerivView V;
    V.p = A.p;
    if (!lds) {
        V.v = A.dv;
        V.tl = A.dtl;
        V.m = nullptr;
        return V;
    }
    const int n = C.n;
    V.v = lds;
    V.tl = lds + 4 * n;
    V.m = lds + 5 * n;
    for (int q = lane; q < 4 * n; q += 64) V.v[q] = A.dv[q];
    for (int q = lane; q < n; q += 64) {
        V.tl[q] = A.dtl[q];
        V.m[q] = A.p[q].m;
    }
    __syncthreads();
    return V;
}

__device__ __forceinline__ void view_close(const EdmdConst& C, const EdmdArrays& A, int lane,
                                           const DerivView& V, bool lds) {
    if (!lds) return;
    __syncthreads();
    const int n = C.n;
    for (int q = lane; q < 4 * n; q += 64) A.dv[q] = V.v[q];
    for (int q = lane; q < n; q += 64) A.dtl[q] = V.tl[q];
}

/* The tangent's free flight over h (damped at rate g when g != 0). */
__device__ __forceinline__ void tan_fly(double& v0, double& v1, double& v2, double& v3, double h,
                                        double g) {
    if (g == 0) {
        v0 += h * v2;
        v1 += h * v3;
        return;
    }
    const double e = edmd_exp(-g * h);
    const double ph = (1 - e) / g;
    v0 += ph * v2;
    v1 += ph * v3;
    v2 *= e;
    v3 *= e;
}

/* Its transpose, the adjoint's flight back over h. */
__device__ __forceinline__ void adj_fly(double& l0, double& l1, double& l2, double& l3, double h,
                                        double g) {
    if (g == 0) {
        l2 += h * l0;
        l3 += h * l1;
        return;
    }
    const double e = edmd_exp(-g * h);
    const double ph = (1 - e) / g;
    l2 = e * l2 + ph * l0;
    l3 = e * l3 + ph * l1;
}

struct TapeLane {
    double t, ax, ay, cx, cy, ff, g;
    int i, j;
};

__device__ __forceinline__ TapeLane tape_load(const EdmdTapeRec* tape, int64_t k, bool ok) {
    TapeLane r{0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (ok) {
        const EdmdTapeRec q = tape[k];
        r.t = q.t; r.ax = q.ax; r.ay = q.ay; r.cx = q.cx; r.cy = q.cy; r.ff = q.ff;
        r.g = q.g;
        r.i = q.i; r.j = q.j;
    }
    return r;
}

__device__ __forceinline__ TapeLane tape_bcast(const TapeLane& r, int l) {
    TapeLane o;
    o.t = lane_bcast(r.t, l);
    o.ax = lane_bcast(r.ax, l);
    o.ay = lane_bcast(r.ay, l);
    o.cx = lane_bcast(r.cx, l);
    o.cy = lane_bcast(r.cy, l);
    o.ff = lane_bcast(r.ff, l);
    o.g = lane_bcast(r.g, l);
    o.i = lane_bcast(r.i, l);
    o.j = lane_bcast(r.j, l);
    return o;
}

template <bool LDS>
__global__ __launch_bounds__(64) void k_tangent(const EdmdSys* __restrict__ systems) {
    extern __shared__ double lds[];
    const EdmdConst& C = systems[blockIdx.x].c;
    const EdmdArrays& A = systems[blockIdx.x].a;
    const int lane = threadIdx.x;
    int64_t nt = A.st->ntape;
    if (nt > C.tape_cap) nt = C.tape_cap;
    const DerivView V = view_open(C, A, lane, LDS ? lds : nullptr);
    double dsum = A.dr->dsum, dtlast = A.dr->dtlast;
    const double g = C.damping ? C.gamma : 0.0;
    for (int64_t base = 0; base < nt; base += 64) {
        const int cnt = nt - base < 64 ? (int)(nt - base) : 64;
        const TapeLane mine = tape_load(A.tape, base + lane, lane < cnt);
        for (int l = 0; l < cnt; ++l) {
            const TapeLane r = tape_bcast(mine, l);
            const int i = r.i, j = r.j;
            double* ti = V.v + 4 * i;
            double ti0 = ti[0], ti1 = ti[1], ti2 = ti[2], ti3 = ti[3];
            tan_fly(ti0, ti1, ti2, ti3, r.t - V.tl[i], g);
            if (j < 0) { /* uniform: the record is the wave's */
                const double vn = r.ax, vt = r.ay, rw = r.cx, rt = r.cy;
                if (j == -1) {
                    const double dt = -ti0 / vn;
                    ti0 = -rw * ti0;
                    ti1 += (vt - rt * vt) * dt;
                    ti2 = -rw * ti2;
                    ti3 = rt * ti3;
                    if (r.g != 0) ti3 += (rt - 1) * r.g * dt;
                } else {
                    const double dt = -ti1 / vn;
                    ti1 = -rw * ti1;
                    ti0 += (vt - rt * vt) * dt;
                    ti3 = -rw * ti3;
                    ti2 = rt * ti2;
                    if (r.g != 0) ti3 -= (1 + rw) * r.g * dt;
                }
                if (lane == 0) {
                    ti[0] = ti0; ti[1] = ti1; ti[2] = ti2; ti[3] = ti3;
                    V.tl[i] = r.t;
                }
                WAVE_SYNC();
                continue;
            }
            double* tj = V.v + 4 * j;
            double tj0 = tj[0], tj1 = tj[1], tj2 = tj[2], tj3 = tj[3];
            tan_fly(tj0, tj1, tj2, tj3, r.t - V.tl[j], g);
            const double mi = V.mass(i), mj = V.mass(j);
            const double ax = r.ax, ay = r.ay, cx = r.cx, cy = r.cy, ff = r.ff;
            const double b = ax * cx + ay * cy;
            const double px = tj0 - ti0, py = tj1 - ti1;
            const double qx = tj2 - ti2, qy = tj3 - ti3;
            const double dt = -(ax * px + ay * py) / b;
            const double rcx = px + cx * dt, rcy = py + cy * dt;
            const double kap = ff / b;
            double dff = kap * ((rcx * cx + rcy * cy) + (ax * qx + ay * qy));
            if (g != 0) dff -= g * ff * dt;
            const double ux = dff * ax + ff * rcx, uy = dff * ay + ff * rcy;
            ti2 += mj * ux;
            ti3 += mj * uy;
            tj2 -= mi * ux;
            tj3 -= mi * uy;
            const double sh = ff * dt;
            ti0 -= mj * sh * ax;
            ti1 -= mj * sh * ay;
            tj0 += mi * sh * ax;
            tj1 += mi * sh * ay;
            if (g != 0) {
                ti2 += g * mj * sh * ax;
                ti3 += g * mj * sh * ay;
                tj2 -= g * mi * sh * ax;
                tj3 -= g * mi * sh * ay;
            }
            dsum += mi * mj * (ax * ax + ay * ay) * dff;
            dtlast = dt;
            if (lane == 0) {
                ti[0] = ti0; ti[1] = ti1; ti[2] = ti2; ti[3] = ti3;
                tj[0] = tj0; tj[1] = tj1; tj[2] = tj2; tj[3] = tj3;
                V.tl[i] = r.t;
                V.tl[j] = r.t;
            }
            WAVE_SYNC();
        }
    }
    view_close(C, A, lane, V, LDS);
    if (lane == 0) {
        A.dr->dsum = dsum;
        A.dr->dtlast = dtlast;
    }
}

template <bool LDS>
__global__ __launch_bounds__(64) void k_adjoint(const EdmdSys* __restrict__ systems) {
    extern __shared__ double lds[];
    const EdmdConst& C = systems[blockIdx.x].c;
    const EdmdArrays& A = systems[blockIdx.x].a;
    const int lane = threadIdx.x;
    int64_t nt = A.st->ntape;
    if (nt > C.tape_cap) nt = C.tape_cap;
    const DerivView V = view_open(C, A, lane, LDS ? lds : nullptr);
    const double lw = A.dr->lw, lt = A.dr->lt;
    const double g = C.damping ? C.gamma : 0.0;
    int pending = A.dr->lt_pending;
    for (int64_t top = nt; top > 0; top -= 64) {
        const int cnt = top < 64 ? (int)top : 64;
        const int64_t base = top - cnt;
        const TapeLane mine = tape_load(A.tape, base + lane, lane < cnt);
        for (int l = cnt - 1; l >= 0; --l) {
            const TapeLane r = tape_bcast(mine, l);
            const int i = r.i, j = r.j;
            double* li = V.v + 4 * i;
            double li0 = li[0], li1 = li[1], li2 = li[2], li3 = li[3];
            adj_fly(li0, li1, li2, li3, V.tl[i] - r.t, g);
            if (j < 0) {
                const double vn = r.ax, vt = r.ay, rw = r.cx, rt = r.cy;
                if (j == -1) {
                    double ldt = (vt - rt * vt) * li1;
                    if (r.g != 0) ldt += (rt - 1) * r.g * li3;
                    li0 = -rw * li0 - ldt / vn;
                    li2 = -rw * li2;
                    li3 = rt * li3;
                } else {
                    double ldt = (vt - rt * vt) * li0;
                    if (r.g != 0) ldt -= (1 + rw) * r.g * li3;
                    li1 = -rw * li1 - ldt / vn;
                    li3 = -rw * li3;
                    li2 = rt * li2;
                }
                if (lane == 0) {
                    li[0] = li0; li[1] = li1; li[2] = li2; li[3] = li3;
                    V.tl[i] = r.t;
                }
                WAVE_SYNC();
                continue;
            }
            double* lj = V.v + 4 * j;
            double lj0 = lj[0], lj1 = lj[1], lj2 = lj[2], lj3 = lj[3];
            adj_fly(lj0, lj1, lj2, lj3, V.tl[j] - r.t, g);
            const double mi = V.mass(i), mj = V.mass(j);
            const double ax = r.ax, ay = r.ay, cx = r.cx, cy = r.cy, ff = r.ff;
            const double b = ax * cx + ay * cy;
            const double kap = ff / b;
            const double lux = mj * li2 - mi * lj2, luy = mj * li3 - mi * lj3;
            double lsh = -mj * (ax * li0 + ay * li1) + mi * (ax * lj0 + ay * lj1);
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
            li0 -= lpx;
            li1 -= lpy;
            lj0 += lpx;
            lj1 += lpy;
            li2 -= lqx;
            li3 -= lqy;
            lj2 += lqx;
            lj3 += lqy;
            if (lane == 0) {
                li[0] = li0; li[1] = li1; li[2] = li2; li[3] = li3;
                lj[0] = lj0; lj[1] = lj1; lj[2] = lj2; lj[3] = lj3;
                V.tl[i] = r.t;
                V.tl[j] = r.t;
            }
            WAVE_SYNC();
        }
    }
    view_close(C, A, lane, V, LDS);
    if (lane == 0) A.dr->lt_pending = pending;
}

/* mode 0: tangent (reference times = local times when taping began),
 * mode 1: adjoint (reference times = the system's current time). */
__global__ __launch_bounds__(256) void k_deriv_init(const EdmdSys* __restrict__ systems, int mode) {
    const EdmdConst& C = systems[blockIdx.y].c;
    const EdmdArrays& A = systems[blockIdx.y].a;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < C.n) A.dtl[q] = mode == 0 ? A.otl[q] : A.st->t;
}

/* mode 0: tangent, free flight of every tangent to the system's time;
 * mode 1: adjoint, cotangent of the state at the taping origin. */
__global__ __launch_bounds__(256) void k_deriv_flush(const EdmdSys* __restrict__ systems, int mode) {
    const EdmdConst& C = systems[blockIdx.y].c;
    const EdmdArrays& A = systems[blockIdx.y].a;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C.n) return;
    double* v = A.dv + 4 * q;
    const double g = C.damping ? C.gamma : 0.0;
    double v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
    if (mode == 0) {
        tan_fly(v0, v1, v2, v3, A.st->t - A.dtl[q], g);
        A.dtl[q] = A.st->t;
    } else {
        adj_fly(v0, v1, v2, v3, A.dtl[q] - A.otl[q], g);
        A.dtl[q] = A.otl[q];
    }
    v[0] = v0;
    v[1] = v1;
    v[2] = v2;
    v[3] = v3;
}

/* Records the taping origin: local time of every particle, cleared tape and
 * sums. */
__global__ __launch_bounds__(256) void k_deriv_origin(const EdmdSys* __restrict__ systems) {
    const EdmdConst& C = systems[blockIdx.y].c;
    const EdmdArrays& A = systems[blockIdx.y].a;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < C.n) A.otl[q] = A.p[q].tl;
    if (q == 0) {
        A.st->ntape = 0;
        A.st->wsum = 0;
        A.st->tlast = A.st->t;
    }
}

/* One thread per system: bound the next run to `seg` more events (0: no
 * bound) and start an empty tape when `clear` is set. */
__global__ void k_seg_begin(const EdmdSys* __restrict__ systems, int nsys, int64_t seg, int clear) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    EdmdState* st = systems[s].a.st;
    st->ev_limit = seg > 0 ? (int64_t)st->nevents + seg : INT64_MAX;
    if (clear) st->ntape = 0;
}

/* ------------------------------------------------------------ launchers */

static const size_t kLdsLimit = 64 * 1024;

static size_t deriv_lds_bytes(int max_n) { return (size_t)max_n * 6 * sizeof(double); }

hipError_t edmd_launch_tangent(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s) {
    const size_t sh = deriv_lds_bytes(max_n);
    if (sh <= kLdsLimit)
        hipLaunchKernelGGL(k_tangent<true>, dim3(nsys), dim3(64), sh, s, d_sys);
    else
        hipLaunchKernelGGL(k_tangent<false>, dim3(nsys), dim3(64), 0, s, d_sys);
    return hipGetLastError();
}

hipError_t edmd_launch_adjoint(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s) {
    const size_t sh = deriv_lds_bytes(max_n);
    if (sh <= kLdsLimit)
        hipLaunchKernelGGL(k_adjoint<true>, dim3(nsys), dim3(64), sh, s, d_sys);
    else
        hipLaunchKernelGGL(k_adjoint<false>, dim3(nsys), dim3(64), 0, s, d_sys);
    return hipGetLastError();
}

hipError_t edmd_launch_deriv_init(const EdmdSys* d_sys, int nsys, int max_n, int mode,
                                  hipStream_t s) {
    dim3 grid((max_n + 255) / 256, nsys);
    hipLaunchKernelGGL(k_deriv_init, grid, dim3(256), 0, s, d_sys, mode);
    return hipGetLastError();
}

hipError_t edmd_launch_deriv_flush(const EdmdSys* d_sys, int nsys, int max_n, int mode,
                                   hipStream_t s) {
    dim3 grid((max_n + 255) / 256, nsys);
    hipLaunchKernelGGL(k_deriv_flush, grid, dim3(256), 0, s, d_sys, mode);
    return hipGetLastError();
}

hipError_t edmd_launch_deriv_origin(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s) {
    dim3 grid((max_n + 255) / 256, nsys);
    hipLaunchKernelGGL(k_deriv_origin, grid, dim3(256), 0, s, d_sys);
    return hipGetLastError();
}

hipError_t edmd_launch_seg_begin(const EdmdSys* d_sys, int nsys, int64_t seg, int clear,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_seg_begin, dim3((nsys + 255) / 256), dim3(256), 0, s, d_sys, nsys, seg,
                       clear);
    return hipGetLastError();
}

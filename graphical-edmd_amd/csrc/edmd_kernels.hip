/*
 * HIP kernels of libedmd_hip.so for gfx950 (CDNA4, 64-lane wavefronts).
 *
 *   k_sweep_tiled  prediction sweep: one workgroup per tile of cells (eight
 *                  tile shapes) stages the tile and its halo in LDS; one
 *                  lane per particle
 *                  computes the next cell crossing and the earliest
 *                  collision partner over the 3x3 cell neighbourhood
 *                  (EDMD.c:2405-2482, 2661-2723, 2829-3003; invoked for all
 *                  particles at EDMD.c:2007-2012 and 4909-4915)
 *   k_sweep_grow   the same for systems whose radii still grow
 *                  (collisionEventGrow, EDMD.c:3104-3230)
 *   k_sweep_count  counts a full sweep's insertions (the rank counter)
 *   k_tree_level   builds one level of the 64-ary minimum calendar
 *   k_run_small, k_run
 *                  persistent event loop: one wavefront per system pops the
 *                  earliest event, applies the reference's handlers
 *                  (EDMD.c:534-610, 2501-2590, 3542-3915, 4657-4713,
 *                  4740-4800, 4828-4923) and re-predicts with all 64 lanes;
 *                  instantiated for the calendar depth and for growing
 *                  systems (the Grow dispatch of EDMD.c:1974-1990)
 *   k_run_lds      the same loop with each system's arrays copied into its
 *                  workgroup's LDS for the duration of the launch
 *   k_epoch_begin, k_epoch_end
 *                  per-system bounds and counters of graph-captured epochs
 *   k_peek, k_pair_time, k_repredict, k_fly_one
 *                  single steps for the reference-compatible entry points
 *   k_sync_all     free flight of every particle to the system time
 *   k_pcf          g(r) (pcf.c:16-75)
 *   k_fp64_check   IEEE division / square root probe for the tests
 *
 * Arithmetic: FP64 throughout, compiled with -ffp-contract=off so that every
 * product and sum is rounded where the reference's C rounds it; division and
 * square root use the compiler's correctly rounded FP64 expansions.
 * Reference line numbers are those of src/EDMD.c.
 */
#define EDMD_DEVICE_ARRAYS 1
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "edmd_internal.h"
#include "edmd_math.h"
#include "edmd_types.h"

#define INF_D __builtin_huge_val()

/* Compiler barrier for memory shared between lanes of one wavefront.  The
 * lanes of a wavefront need no hardware action to see each other's global
 * or LDS stores in program order; this only keeps the compiler from moving
 * memory operations across the point. */
#define WAVE_SYNC() __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront")

/* Marks a value that is equal on every lane.  It stays in vector
 * registers: the event loop has few scalar registers to spare, and pinning
 * these values to them (readfirstlane) made the compiler spill about 130
 * scalar registers into vector-register lanes, whose reloads cost more
 * vector instructions than they saved (8 % of the batched loop's rate). */
__device__ __forceinline__ int uni(int v) { return v; }
__device__ __forceinline__ double uni(double v) { return v; }

/* The wavefront's index within its workgroup, in a scalar register: the
 * system a wavefront runs is addressed through it, so its record is read
 * with scalar loads. */
__device__ __forceinline__ int wave_in_group() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

__device__ __forceinline__ int64_t uni64(int64_t v) {
    return (int64_t)((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32 |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)v));
}

__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int64_t lane_value64(int64_t v, int l) {
    return (int64_t)((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), l) << 32 |
                     (uint32_t)__builtin_amdgcn_readlane((int)v, l));
}

/* A double moved across lanes by a DPP pattern (both halves). */
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

/* Minimum over the 64 lanes, returned on every lane.  Within each row of
 * 16 lanes: quad swaps (lane ^ 1, lane ^ 2), then the half-row mirror
 * (lane l <-> 7 - l within 8) and the row mirror (l <-> 15 - l), each
 * leaving the minimum of the lanes combined so far on every lane; the four
 * row minima are then read into scalar registers.  No LDS traffic, unlike
 * a shuffle butterfly.  The operands must not be NaN. */
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }

__device__ __forceinline__ double wave_min(double v) {
    v = dmin(v, dpp_f64<0xb1>(v));  /* quad_perm [1,0,3,2] */
    v = dmin(v, dpp_f64<0x4e>(v));  /* quad_perm [2,3,0,1] */
    v = dmin(v, dpp_f64<0x141>(v)); /* row_half_mirror */
    v = dmin(v, dpp_f64<0x140>(v)); /* row_mirror */
    const double r0 = lane_value(v, 0), r1 = lane_value(v, 16);
    const double r2 = lane_value(v, 32), r3 = lane_value(v, 48);
    return dmin(dmin(r0, r1), dmin(r2, r3));
}

__device__ __forceinline__ int imin(int a, int b) { return b < a ? b : a; }

/* Minimum of an int over the wave, on every lane: the same DPP pattern as
 * wave_min, each step a single v_min_i32 with a DPP source. */
__device__ __forceinline__ int wave_min_i32(int v) {
    v = imin(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0xb1, 0xf, 0xf, false));
    v = imin(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0x4e, 0xf, 0xf, false));
    v = imin(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0x141, 0xf, 0xf, false));
    v = imin(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0x140, 0xf, 0xf, false));
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return imin(imin(r0, r1), imin(r2, r3));
}

/* 32-bit key of v - base that never decreases when v increases: the float
 * nearest to v - base (both roundings are monotone), its sign-magnitude
 * bits turned into a signed order (-0 and +0 alike).  v is never NaN. */
__device__ __forceinline__ int order_key(double v, double base) {
    const int b = __float_as_int((float)(v - base));
    const int m = b & 0x7fffffff;
    return b < 0 ? -m : m;
}

/* Lane holding the minimum of v over the wave, the lowest lane among equal
 * values, and that minimum.  The search runs on order_key(v, base) with
 * single-instruction DPP steps; the lane holding the exact minimum always
 * has the least key, and only when several lanes share it are their
 * doubles compared.  base should lie near the values for resolution (any
 * finite base is correct). */
__device__ __forceinline__ int wave_argmin_lane(double v, double base, double& vmin) {
    const int k = order_key(v, base);
    const int mk = wave_min_i32(k);
    const uint64_t tied = __ballot(k == mk);
    if (__popcll(tied) == 1) {
        const int w = (int)__builtin_ctzll(tied);
        vmin = lane_value(v, w);
        return w;
    }
    const double vv = (k == mk) ? v : INF_D;
    vmin = wave_min(vv);
    return (int)__builtin_ctzll(__ballot(vv == vmin));
}

/* Two independent int minima over the wave at once: the same DPP steps as
 * wave_min_i32 with the two chains interleaved, so each DPP step's latency
 * is covered by the other chain's step. */
__device__ __forceinline__ void wave_min_i32x2(int& a, int& b) {
    a = imin(a, __builtin_amdgcn_update_dpp(0x7fffffff, a, 0xb1, 0xf, 0xf, false));
    b = imin(b, __builtin_amdgcn_update_dpp(0x7fffffff, b, 0xb1, 0xf, 0xf, false));
    a = imin(a, __builtin_amdgcn_update_dpp(0x7fffffff, a, 0x4e, 0xf, 0xf, false));
    b = imin(b, __builtin_amdgcn_update_dpp(0x7fffffff, b, 0x4e, 0xf, 0xf, false));
    a = imin(a, __builtin_amdgcn_update_dpp(0x7fffffff, a, 0x141, 0xf, 0xf, false));
    b = imin(b, __builtin_amdgcn_update_dpp(0x7fffffff, b, 0x141, 0xf, 0xf, false));
    a = imin(a, __builtin_amdgcn_update_dpp(0x7fffffff, a, 0x140, 0xf, 0xf, false));
    b = imin(b, __builtin_amdgcn_update_dpp(0x7fffffff, b, 0x140, 0xf, 0xf, false));
    const int a0 = __builtin_amdgcn_readlane(a, 0), a1 = __builtin_amdgcn_readlane(a, 16);
    const int a2 = __builtin_amdgcn_readlane(a, 32), a3 = __builtin_amdgcn_readlane(a, 48);
    const int b0 = __builtin_amdgcn_readlane(b, 0), b1 = __builtin_amdgcn_readlane(b, 16);
    const int b2 = __builtin_amdgcn_readlane(b, 32), b3 = __builtin_amdgcn_readlane(b, 48);
    a = imin(imin(a0, a1), imin(a2, a3));
    b = imin(imin(b0, b1), imin(b2, b3));
}

/* wave_argmin_lane of two values at once (the minima of two owners over the
 * same candidate pass): the key searches run interleaved; each result is
 * the one wave_argmin_lane gives for that value alone. */
__device__ __forceinline__ void wave_argmin_lane2(double va, double vb, double base, double& ma,
                                                  int& wa, double& mb, int& wb) {
    const int ka = order_key(va, base), kb = order_key(vb, base);
    int mka = ka, mkb = kb;
    wave_min_i32x2(mka, mkb);
    const uint64_t ta = __ballot(ka == mka), tb = __ballot(kb == mkb);
    if (__popcll(ta) == 1) {
        wa = (int)__builtin_ctzll(ta);
        ma = lane_value(va, wa);
    } else {
        const double vv = (ka == mka) ? va : INF_D;
        ma = wave_min(vv);
        wa = (int)__builtin_ctzll(__ballot(vv == ma));
    }
    if (__popcll(tb) == 1) {
        wb = (int)__builtin_ctzll(tb);
        mb = lane_value(vb, wb);
    } else {
        const double vv = (kb == mkb) ? vb : INF_D;
        mb = wave_min(vv);
        wb = (int)__builtin_ctzll(__ballot(vv == mb));
    }
}

/* ------------------------------------------------------------ kinematics */

/* PBC, EDMD.c:5896-5913 */
__device__ __forceinline__ void min_image(const EdmdConst& C, double& dx, double& dy) {
    /* mihx / mihy are halfLx / halfLy, or +inf across walls (no image) */
    if (dx >= C.mihx) dx -= C.Lx;
    else if (dx < -C.mihx) dx += C.Lx;
    if (dy >= C.mihy) dy -= C.Ly;
    else if (dy < -C.mihy) dy += C.Ly;
}

/* edmd_exp (e^x, the library's own) is in edmd_math.h. */

/* log x from log m = 2 atanh s = 2(s + s^3/3 + ... + s^23/23),
 * s = (m - 1)/(m + 1), x = m 2^e with m in [sqrt(1/2), sqrt(2)), plus
 * e ln2 in two parts; within 4 ulp, the oracle's steps. */
__device__ __forceinline__ double edmd_log(double x) {
    if (x != x || x < 0) return __builtin_nan("");
    if (x == 0) return -__builtin_huge_val();
    if (x == __builtin_huge_val()) return x;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752) {
        m *= 2;
        e -= 1;
    }
    const double s = (m - 1) / (m + 1), s2 = s * s;
    double q = 1.0 / 23;
#pragma unroll
    for (int n = 21; n >= 1; n -= 2) q = (1.0 / n) + s2 * q;
    const double lm = 2 * s * q;
    return e * 0x1.62e42fefa3800p-1 + (e * 0x1.ef35793c76730p-45 + lm);
}

/* sin and cos of an angle in [-pi, pi] (atan2's range): theta = k pi/2 + r,
 * |r| <= pi/4, pi/2 in three parts (33, 33 and 53 bits, Cody and Waite) so
 * that k times the first two is exact; sin r and cos r from their Taylor
 * series to r^19/19! and r^20/20! in nested form; the quadrant k mod 4
 * picks and signs them.  The oracle runs the same steps; both are within
 * 2 ulp of the C library. */
__device__ __forceinline__ void edmd_sincos(double th, double& sn, double& cs) {
    const double k = rint(th * 0x1.45f306dc9c883p-1); /* th * 2/pi, to nearest */
    const double r = ((th - k * 0x1.921fb54400000p+0) - k * 0x1.0b4611a600000p-34) -
                     k * 0x1.3198a2e037073p-69;
    const double r2 = r * r;
    double ps = 1.0, pc = 1.0;
#pragma unroll
    for (int n = 19; n >= 3; n -= 2) ps = 1.0 - (r2 * ((1.0 / n) * (1.0 / (n - 1)))) * ps;
#pragma unroll
    for (int n = 20; n >= 2; n -= 2) pc = 1.0 - (r2 * ((1.0 / n) * (1.0 / (n - 1)))) * pc;
    const double sr = r * ps, cr = pc;
    switch (((int)k) & 3) {
        case 0: sn = sr; cs = cr; break;
        case 1: sn = cr; cs = -sr; break;
        case 2: sn = -sr; cs = -cr; break;
        default: sn = -cr; cs = sr; break;
    }
}

/* atan z for z in [0, 1]: about 0.4 or 0.8 (atan z = atan c + atan w,
 * w = (z - c)/(1 + z c), |w| <= 0.2), then w (1 - w^2 (1/3 - w^2 (1/5 -
 * ... w^2/23))). */
__device__ __forceinline__ double edmd_atan01(double z) {
    double base = 0, w = z;
    if (z > 0.6) {
        w = (z - 0.8) / (1 + z * 0.8);
        base = 0x1.5977a5103ea93p-1; /* atan(0.8) */
    } else if (z > 0.2) {
        w = (z - 0.4) / (1 + z * 0.4);
        base = 0x1.85a376b677dc0p-2; /* atan(0.4) */
    }
    const double w2 = w * w;
    double q = 1.0 / 23;
#pragma unroll
    for (int n = 21; n >= 3; n -= 2) q = (1.0 / n) - w2 * q;
    return base + w * (1 - w2 * q);
}

/* atan2(y, x) in [-pi, pi], pi/2 and pi as head + tail. */
__device__ __forceinline__ double edmd_atan2(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    double a;
    if (ay > ax) a = (0x1.921fb54442d18p+0 - edmd_atan01(ax / ay)) + 0x1.1a62633145c07p-54;
    else if (ax > 0) a = edmd_atan01(ay / ax);
    else a = 0;
    if (__builtin_signbit(x)) a = (0x1.921fb54442d18p+1 - a) + 0x1.1a62633145c07p-53;
    return __builtin_signbit(y) ? -a : a;
}

/* The quartic of the circular wall under the field (quartic.c:27-158):
 * Ferrari's method through the cubic resolvent, step for step as there,
 * with this library's own acos (atan2 of sqrt((1 - t)(1 + t)) and t), cos
 * and cube root (exp(log(x) / 3), the pow with exponent 1./3).  The
 * oracle runs the same steps. */
__device__ __forceinline__ double edmd_cosine(double th) {
    double sn, cs;
    edmd_sincos(th, sn, cs);
    return cs;
}

__device__ __forceinline__ unsigned edmd_solve_p3(double* x, double a, double b, double c) {
    const double a2 = a * a;
    double q = (a2 - 3 * b) / 9;
    const double r = (a * (2 * a2 - 9 * b) + 27 * c) / 54;
    const double r2 = r * r;
    const double q3 = q * q * q;
    if (r2 < q3) {
        double t = r / sqrt(q3);
        if (t < -1) t = -1;
        if (t > 1) t = 1;
        t = edmd_atan2(sqrt((1 - t) * (1 + t)), t);
        a /= 3;
        q = -2 * sqrt(q);
        x[0] = q * edmd_cosine(t / 3) - a;
        x[1] = q * edmd_cosine((t + (2.0 * 3.14159265358979323846)) / 3) - a;
        x[2] = q * edmd_cosine((t - (2.0 * 3.14159265358979323846)) / 3) - a;
        return 3;
    }
    const double m = fabs(r) + sqrt(r2 - q3);
    double A = -(m == 0 ? 0.0 : edmd_exp(edmd_log(m) * (1. / 3)));
    if (r < 0) A = -A;
    const double B = (0 == A ? 0 : q / A);
    a /= 3;
    x[0] = (A + B) - a;
    x[1] = -0.5 * (A + B) - a;
    x[2] = 0.5 * sqrt(3.) * (A - B);
    if (fabs(x[2]) < 1e-12) {
        x[2] = x[1];
        return 2;
    }
    return 1;
}

/* The smallest real root above 1e-9 of a t^4 + b t^3 + c t^2 + d t + e,
 * 1e10 when there is none (smallestRoot, quartic.c:143-157). */
__device__ __forceinline__ double edmd_smallest_root(double a, double b, double c, double d,
                                                     double e) {
    const double qa = b / a, qb = c / a, qc = d / a, qd = e / a;
    const double a3 = -qb;
    const double b3 = qa * qc - 4. * qd;
    const double c3 = -qa * qa * qd - qc * qc + 4. * qb * qd;
    double x3[3];
    const unsigned zeroes = edmd_solve_p3(x3, a3, b3, c3);
    double q1, q2, p1, p2, D, sqD, y = x3[0];
    if (zeroes != 1) {
        if (fabs(x3[1]) > fabs(y)) y = x3[1];
        if (fabs(x3[2]) > fabs(y)) y = x3[2];
    }
    D = y * y - 4 * qd;
    if (fabs(D) < 1e-12) {
        q1 = q2 = y * 0.5;
        D = qa * qa - 4 * (qb - y);
        if (fabs(D) < 1e-12) {
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
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (im[i] == 0 && re[i] > 0.000000001 && re[i] < best) best = re[i];
    return best;
}

/* MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998), the generator
 * behind the reference's drand (mersenne.c genrand_int32; drand(a, b) =
 * a + (b - a) k / 2^32, EDMD.c:5699-5701): one stream per system, its 624
 * state words and the index of the next one (word 624) in HBM.  Word k of
 * the twist is mt[(k + 397) % 624] ^ (y >> 1) ^ (y odd ? 0x9908b0df : 0)
 * with y the upper bit of mt[k] and the lower 31 of mt[(k + 1) % 624]; in
 * the sequential order of the algorithm the words [0, 227) read only old
 * words, [227, 454) the new [0, 227), [454, 623) the new [227, 396) and
 * word 623 the new words 0 and 396, so each range runs wave-parallel once
 * the one before is stored. */
__device__ __forceinline__ void mt_twist_range(uint32_t EDMD_GLOBAL* mt, int lane, int a, int b) {
    uint32_t nw[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = a + lane + 64 * r;
        nw[r] = 0;
        if (k < b) {
            const uint32_t y = (mt[k] & 0x80000000u) | (mt[k + 1 == 624 ? 0 : k + 1] & 0x7fffffffu);
            nw[r] = mt[k + 397 < 624 ? k + 397 : k - 227] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = a + lane + 64 * r;
        if (k < b) mt[k] = nw[r];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

/* The system's next 32-bit output (uniform). */
template <class AR>
__device__ __forceinline__ uint32_t mt_next(const AR& A, int lane) {
    uint32_t EDMD_GLOBAL* mt = A.mt;
    int idx = uni((int)mt[624]);
    if (idx >= 624) {
        mt_twist_range(mt, lane, 0, 227);
        mt_twist_range(mt, lane, 227, 454);
        mt_twist_range(mt, lane, 454, 623);
        mt_twist_range(mt, lane, 623, 624);
        idx = 0;
    }
    uint32_t y = (uint32_t)uni((int)mt[idx]);
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    if (lane == 0) mt[624] = (uint32_t)(idx + 1);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    return y;
}

/* drand(0, hi) of the reference (EDMD.c:5699-5701) on the system's stream. */
template <class AR>
__device__ __forceinline__ double mt_drand0(const AR& A, int lane, double hi) {
    const double g = mt_next(A, lane) * (1.0 / 4294967296.0);
    return (g * (hi - 0.0)) + 0.0;
}

/* sign of EDMD.c:5703-5705: -1, 0 or 1. */
__device__ __forceinline__ double ref_sign(double x) { return (x < 0) ? -1.0 : (double)(x > 0); }

/* logTime, EDMD.c:5709-5721: under damping, the time at which the damped
 * motion has covered what the undamped one covers in t; never when it
 * cannot.  F = false compiles damping out. */
template <bool F>
__device__ __forceinline__ double log_time(const EdmdConst& C, double t) {
    if (!(F && C.damping)) return t;
    const double var = (1 - t * C.gamma);
    if (var < 0) return EDMD_NEVER;
    return -edmd_log(var) / C.gamma;
}

/* The wall part of collisionEventNormal (EDMD.c:2845-2895, no field, mid or
 * circular wall; logTime is the identity without damping): a disk in a
 * boundary cell moving towards the wall gets the time to contact, the
 * starting value of the strict-less minimum over the neighbours.  The x
 * test sets xy = 0 whenever it applies, also when the y wall stays the
 * earlier, as the reference does.  Returns whether a wall time was set. */
__device__ __forceinline__ bool wall_time(const EdmdConst& C, double x, double y, double vx,
                                          double vy, double r, int cx, int cy, double& dt,
                                          int& xy, int typ = 0) {
    bool wall = false;
    dt = EDMD_NEVER;
    xy = 2;
    if ((C.walls & 2) && C.field_on) {
        /* EDMD.c:2846-2860: the bottom wall from the bottom row whatever the
         * velocity; the top wall for any disk moving up (NaN when it cannot
         * reach it) */
        const double f = C.field;
        if (cy == 0) {
            dt = (-vy - sqrt(vy * vy - 2 * f * (y - r))) / f;
            wall = true;
            xy = 1;
        }
        if (vy > 0) {
            const double dt2 = (-vy + sqrt(vy * vy - 2 * f * (y - (C.Ly - r)))) / f;
            if (!__builtin_isnan(dt2) && dt2 < dt) {
                dt = dt2;
                wall = true;
                xy = 1;
            }
        }
    } else if (C.walls & 2) {
        if (cy == 0 && vy < 0) {
            dt = log_time<true>(C, -(y - r) / vy);
            wall = true;
            xy = 1;
        } else if (cy == C.ncy - 1 && vy > 0) {
            dt = log_time<true>(C, (C.Ly - y - r) / vy);
            wall = true;
            xy = 1;
        }
    }
    if (C.walls & 1) {
        if (cx == 0 && vx < 0) {
            const double dw = log_time<true>(C, -(x - r) / vx);
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        } else if (cx == C.ncx - 1 && vx > 0) {
            const double dw = log_time<true>(C, (C.Lx - x - r) / vx);
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        }
    }
    if ((C.walls & 8) && typ == 1) {
        /* the mid wall for the solute (EDMD.c:2897-2915): from within 4 of
         * x = Lx/2, moving towards it */
        const double h = C.Lx / 2;
        if (x > h - 4 && x < h && vx > 0) {
            const double dw = log_time<true>(C, (h - x - r) / vx);
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        } else if (x < h + 4 && x > h && vx < 0) {
            const double dw = log_time<true>(C, ((x - r) - h) / (-vx));
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        }
    }
    if ((C.walls & 4) && C.field_on) {
        /* the circle under the field (EDMD.c:2918-2935): the parabolic
         * flight meets it at the quartic's smallest positive root (y0 about
         * halfLy here) */
        const double y0 = y - C.halfLy, x0 = x - C.halfLx;
        const double e = x0 * x0 + y0 * y0 - (C.halfLx - r) * (C.halfLx - r);
        const double f = C.field;
        const double a = f * f / 4;
        const double b = f * vy;
        const double c = vx * vx + vy * vy + f * y0;
        const double d = 2 * vx * x0 + 2 * vy * y0;
        dt = edmd_smallest_root(a, b, c, d, e);
        wall = true;
        xy = 2;
    } else if (C.walls & 4) {
        /* the circle (EDMD.c:2938-2954, no field): the centre reaches the
         * circle of radius halfLx - r, for every disk; the time replaces
         * the other walls' */
        const double y0 = y - C.halfLx, x0 = x - C.halfLx;
        const double Cc = (x0 * x0 + y0 * y0) - (C.halfLx - r) * (C.halfLx - r);
        const double A = (vx * vx + vy * vy);
        const double B = 2 * (vx * x0 + vy * y0);
        dt = (-B + sqrt(B * B - 4 * A * Cc)) / (2 * A);
        wall = true;
        xy = 2;
    }
    return wall;
}

/* The wall slot of collisionEventGrow (EDMD.c:3118-3153): the surface of a
 * growing disk in a boundary cell closes on the wall at v -/+ vr; no
 * logTime.  dt starts at the growth version's 1e7; as in wall_time, the x
 * test sets xy = 0 whenever it applies. */
__device__ __forceinline__ bool wall_time_grow(const EdmdConst& C, double x, double y, double vx,
                                               double vy, double r, double vr, int cx, int cy,
                                               double& dt, int& xy,
                                               int typ = 0) {
    bool wall = false;
    dt = EDMD_GROW_NONE;
    xy = 2;
    if (C.walls & 2) {
        const double vyp = vy + vr, vym = vy - vr;
        if (cy == 0 && vym < 0) {
            dt = -(y - r) / vym;
            wall = true;
            xy = 1;
        } else if (cy == C.ncy - 1 && vyp > 0) {
            dt = (C.Ly - y - r) / vyp;
            wall = true;
            xy = 1;
        }
    }
    if (C.walls & 1) {
        const double vxp = vx + vr, vxm = vx - vr;
        if (cx == 0 && vxm < 0) {
            const double dw = -(x - r) / vxm;
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        } else if (cx == C.ncx - 1 && vxp > 0) {
            const double dw = (C.Lx - x - r) / vxp;
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        }
    }
    if ((C.walls & 8) && typ == 1) {
        /* EDMD.c:3175-3195 */
        const double h = C.Lx / 2;
        const double vxp = vx + vr, vxm = vx - vr;
        if (x > h - 4 && x < h && vxp > 0) {
            const double dw = (h - x - r) / (vxp);
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        } else if (x < h + 4 && x > h && vxm < 0) {
            const double dw = ((x - r) - h) / (-vxm);
            if (dt > dw) {
                wall = true;
                dt = dw;
            }
            xy = 0;
        }
    }
    if (C.walls & 4) {
        /* EDMD.c:3158-3172: the surface closes on the circle at |v| - vr */
        const double y0 = y - C.halfLx, x0 = x - C.halfLx;
        const double A = vx * vx + vy * vy - vr * vr;
        const double B = 2 * (vx * x0 + vy * y0 + (C.halfLx - r) * vr);
        const double Cc = (x0 * x0 + y0 * y0) - (C.halfLx - r) * (C.halfLx - r);
        dt = (-B + sqrt(B * B - 4 * A * Cc)) / (2 * A);
        wall = true;
        xy = 2;
    }
    return wall;
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

/* freeFlyNormal for a system with the field (EDMD.c:4957-4963: parabolic
 * in y, vy advanced) or damping (EDMD.c:4965-4975: velocities times
 * exp(-gamma dt)); F = false compiles both out. */
template <bool F>
__device__ __forceinline__ void fly_v(const EdmdConst& C, double t, double& x, double& y,
                                      double& vx, double& vy, double& tl) {
    /* the field first: with both on the flight is the field's alone
     * (EDMD.c:4957-4975), while logTime still applies the damping */
    if (F && C.field_on) {
        const double dt = t - tl;
        tl = t;
        x += dt * vx;
        y += dt * vy + 0.5 * C.field * dt * dt;
        vy += dt * C.field;
        x = wrap_post(x, C.Lx);
        y = wrap_post(y, C.Ly);
    } else if (F && C.damping) {
        const double dt = t - tl;
        tl = t;
        const double g = C.gamma;
        const double ex = edmd_exp(-g * dt);
        x += vx / g * (1 - ex);
        y += vy / g * (1 - ex);
        vx *= ex;
        vy *= ex;
        x = wrap_post(x, C.Lx);
        y = wrap_post(y, C.Ly);
    } else {
        fly(C, t, x, y, vx, vy, tl);
    }
}

/* freeFlyGrow, EDMD.c:4992-5007: the radius grows with the flight. */
__device__ __forceinline__ void fly_grow(const EdmdConst& C, double t, double& x, double& y,
                                         double vx, double vy, double& tl, double& rad,
                                         double vr) {
    double dt = t - tl;
    tl = t;
    x += dt * vx;
    y += dt * vy;
    rad += dt * vr;
    x = wrap_post(x, C.Lx);
    y = wrap_post(y, C.Ly);
}

/* collisionTimeNormal, EDMD.c:2661-2723: relative time until contact of the
 * synchronised owner (x1..) with candidate 2 last updated at tl2. */
/* The same with candidate 2's position already flown to t:
 * xt2 = x2 + (t - tl2) * vx2, yt2 likewise (EDMD.c:2667-2671). */
template <bool WRAP = true>
__device__ __forceinline__ double coll_time_at(const EdmdConst& C, double x1, double y1,
                                               double vx1, double vy1, double r1, double xt2,
                                               double yt2, double vx2, double vy2, double r2,
                                               bool& overlap) {
    double dvx = vx2 - vx1;
    double dvy = vy2 - vy1;
    double dx = xt2 - x1;
    double dy = yt2 - y1;
    if (WRAP) min_image(C, dx, dy);
    double b = dx * dvx + dy * dvy;
    if (b > 0) return EDMD_NEVER;
    double v2 = dvx * dvx + dvy * dvy;
    double d = dx * dx + dy * dy - 4 * r1 * r2; /* DIST2, EDMD.c:35-43 */
    double det = b * b - v2 * d;
    if (d < -0.01) overlap = true; /* EDMD.c:2708-2717 */
    if (det < 0) return EDMD_NEVER;
    return (-b - sqrt(det)) / v2;
}

/* collisionTimeNormal as coll_time computes it, without branches: every
 * term is evaluated and the result selected, so a wavefront whose lanes hold
 * different candidates runs one instruction stream.  Bitwise the same
 * result; the overlap test applies where the reference reaches it (b <= 0). */
__device__ __forceinline__ double coll_time_nb(const EdmdConst& C, double t, double x1, double y1,
                                               double vx1, double vy1, double r1, double x2,
                                               double y2, double vx2, double vy2, double tl2,
                                               double r2, bool& overlap) {
    const double lat2 = t - tl2;
    const double dvx = vx2 - vx1;
    const double dvy = vy2 - vy1;
    double dx = (x2 + lat2 * vx2) - x1;
    double dy = (y2 + lat2 * vy2) - y1;
    min_image(C, dx, dy);
    const double b = dx * dvx + dy * dvy;
    const double v2 = dvx * dvx + dvy * dvy;
    const double d = dx * dx + dy * dy - 4 * r1 * r2; /* DIST2, EDMD.c:35-43 */
    const double det = b * b - v2 * d;
    const bool approach = !(b > 0);
    overlap = overlap || (approach && d < -0.01); /* EDMD.c:2708-2717 */
    const double root = (-b - sqrt(det)) / v2;
    return (approach && !(det < 0)) ? root : EDMD_NEVER;
}

__device__ __forceinline__ double coll_time(const EdmdConst& C, double t, double x1, double y1,
                                            double vx1, double vy1, double r1, double x2,
                                            double y2, double vx2, double vy2, double tl2,
                                            double r2, bool& overlap) {
    double lat2 = t - tl2;
    return coll_time_at(C, x1, y1, vx1, vy1, r1, x2 + lat2 * vx2, y2 + lat2 * vy2, vx2, vy2, r2,
                        overlap);
}

/* collisionTimeGrow, EDMD.c:2598-2659 (2D): contact distance
 * DIST(r1, r2(t)) widening at vr1 + vr2; the earlier positive root, or the
 * later one under the reference's 1e-9 rule; an overlap is flagged only
 * when neither root is taken. */
__device__ __forceinline__ double coll_time_grow(const EdmdConst& C, double t, double x1,
                                                 double y1, double vx1, double vy1, double r1,
                                                 double vr1, double x2, double y2, double vx2,
                                                 double vy2, double tl2, double r2, double vr2,
                                                 bool& overlap) {
    double lat2 = t - tl2;
    double dvx = vx2 - vx1;
    double dvy = vy2 - vy1;
    double dvr = vr1 + vr2;
    double dx = (x2 + lat2 * vx2) - x1;
    double dy = (y2 + lat2 * vy2) - y1;
    double dr = sqrt(4 * r1 * (r2 + lat2 * vr2));
    min_image(C, dx, dy);
    double b = dx * dvx + dy * dvy - dvr * dr;
    double v2 = dvx * dvx + dvy * dvy;
    double dist2 = dx * dx + dy * dy;
    double det = b * b - (v2 - dvr * dvr) * (dist2 - dr * dr);
    if (det < 0) return EDMD_NEVER;
    double plus = (-b + sqrt(det)) / (v2 - dvr * dvr);
    double minus = (-b - sqrt(det)) / (v2 - dvr * dvr);
    if (((minus > 0) && (plus > 0) && (minus < plus)) || ((minus > 0) && (plus < 0)))
        return minus;
    else if (((minus > 0) && (plus > 0.000000001) && (plus < minus)) ||
             ((minus < 0) && (plus > 0.000000001)))
        return plus;
    if (dist2 - dr * dr < -0.01) overlap = true; /* EDMD.c:2653-2656 */
    return EDMD_NEVER;
}

/* crossingEventNormal, EDMD.c:2405-2482: absolute crossing time and
 * direction code (1 -x, 2 +x, 3 -y, 4 +y); ties go to y. */
template <bool F = true>
__device__ __forceinline__ void cross_time(const EdmdConst& C, double t, double x, double y,
                                           double vx, double vy, int cx, int cy, double& tout,
                                           int& dir) {
    double tx, ty;
    int xx, yy;
    if (vx < 0) {
        tx = log_time<F>(C, wrap_inside(cx * C.csx - x, C.halfLx, C.Lx) / vx);
        xx = 1;
    } else {
        tx = log_time<F>(C, wrap_inside((1 + cx) * C.csx - x, C.halfLx, C.Lx) / vx);
        xx = 2;
    }
    if (F && C.field_on) {
        /* EDMD.c:2428-2442: up through the top edge when fast enough, else
         * down through the bottom edge */
        const double f = C.field;
        double dy = wrap_inside(y - (1 + cy) * C.csy, C.halfLy, C.Ly);
        if (dy < 0 && vy > sqrt(2 * f * dy)) {
            ty = (-vy + sqrt(vy * vy - 2 * f * dy)) / f;
            yy = 4;
        } else {
            dy = wrap_inside(y - cy * C.csy, C.halfLy, C.Ly);
            ty = (-vy - sqrt(vy * vy - 2 * f * dy)) / f;
            yy = 3;
        }
    } else if (vy < 0) {
        ty = log_time<F>(C, wrap_inside(cy * C.csy - y, C.halfLy, C.Ly) / vy);
        yy = 3;
    } else {
        ty = log_time<F>(C, wrap_inside((1 + cy) * C.csy - y, C.halfLy, C.Ly) / vy);
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

/* Rank of an event scheduled at t_ev with the queue at paulTime = paul after
 * ctr earlier insertions (addEventToQueue, EDMD.c:2144-2170): straight into
 * the tree when t_ev - paulTime < dtPaul, into the overflow list when it is
 * dtPaul * nlists or more ahead, otherwise into a LIFO bucket. */
__device__ __forceinline__ int64_t rank_of(double t_ev, double paul, double dtp, double horizon,
                                           int64_t ctr) {
    const double dt = t_ev - paul;
    /* without branches (the loop files 4-6 events per collision):
     * bucket D - 1 - ctr, plus 2 ctr + 1 for the tree and the overflow
     * list, plus 2^61 for the overflow list */
    const int64_t up = (int64_t)(dt < dtp) | (int64_t)(dt >= horizon);
    const int64_t ovf = (int64_t)(dt >= horizon);
    return EDMD_RANK_DIRECT - 1 - ctr + (-up & (2 * ctr + 1)) + (ovf << 61);
}

/* Calendar leaf of a particle: its earlier slot by (time, rank) (a NaN key
 * is parked at +inf so that it is never popped), that slot's rank and its
 * event code: the partner for a collision, minus the direction for a
 * crossing. */
struct Leaf {
    double key;
    int64_t rank;
    int aux;
};

__device__ __forceinline__ Leaf make_leaf(double tcoll, int64_t rcoll, double tcross,
                                          int64_t rcross, int partner, int dir) {
    /* non-short-circuit: the compiler keeps a select, not a branch */
    const bool coll = (tcoll < tcross) | ((tcoll == tcross) & (rcoll < rcross));
    Leaf L;
    const double k = coll ? tcoll : tcross;
    L.key = (k == k) ? k : INF_D;
    L.rank = coll ? rcoll : rcross;
    L.aux = coll ? partner : -dir;
    return L;
}

__device__ __forceinline__ bool before(double ka, int64_t ra, double kb, int64_t rb) {
    return (ka < kb) | ((ka == kb) & (ra < rb));
}

/* ------------------------------------------------------------ predictions */

/* Sequential neighbour scan for one particle (one lane): strict-less minimum
 * in the reference's order, y offset outer, x offset inner, newest list
 * entry first (EDMD.c:2959-3003).  Returns the relative time; NEVER and
 * partner 0 when no candidate qualifies (EDMD.c:2830-2831). */
template <bool G, bool W = true, class AR>
__device__ __forceinline__ double predict_seq(const EdmdConst& C, const AR& A, double t,
                                              int p, const EdmdPart& o, int& partner,
                                              bool& overlap) {
    /* collisionEventGrow (EDMD.c:3104-3230, no walls) starts from 1e7 */
    double dt = G ? EDMD_GROW_NONE : EDMD_NEVER;
    int best = 0;
    const double ovr = G ? A.vr[p] : 0.0;
    int xy = 2;
    bool wall = false;
    const int typ = (W && (C.walls & 8)) ? A.ptype[p] : 0;
    if (W && !G && C.walls)
        wall = wall_time(C, o.x, o.y, o.vx, o.vy, o.rad, o.cx, o.cy, dt, xy, typ);
    if (W && G && C.walls)
        wall = wall_time_grow(C, o.x, o.y, o.vx, o.vy, o.rad, ovr, o.cx, o.cy, dt, xy, typ);
    for (int j = -1; j <= 1; ++j) {
        int row = wrap_cell(o.cy + j, C.ncy) * C.ncx;
        for (int k = -1; k <= 1; ++k) {
            int c = row + wrap_cell(o.cx + k, C.ncx);
            int cnt = A.ccnt[c];
            const auto sl = A.cslot + (size_t)c * C.cap;
            for (int s = cnt - 1; s >= 0; --s) {
                int q = sl[s];
                if (q == p) continue;
                EdmdPart c2 = A.p[q];
                if (W && !G && (C.field_on || C.damping)) {
                    /* collisionTimeNormal flies the neighbour first under
                     * the field or damping, a state update (EDMD.c:2664-2666) */
                    fly_v<true>(C, t, c2.x, c2.y, c2.vx, c2.vy, c2.tl);
                    const auto w = A.p + q;
                    w->x = c2.x;
                    w->y = c2.y;
                    w->vx = c2.vx;
                    w->vy = c2.vy;
                    w->tl = c2.tl;
                }
                double tt;
                if (G)
                    tt = coll_time_grow(C, t, o.x, o.y, o.vx, o.vy, o.rad, ovr, c2.x, c2.y, c2.vx,
                                        c2.vy, c2.tl, c2.rad, A.vr[q], overlap);
                else
                    tt = log_time<W && !G>(
                        C, coll_time(C, t, o.x, o.y, o.vx, o.vy, o.rad, c2.x, c2.y, c2.vx, c2.vy,
                                     c2.tl, c2.rad, overlap));
                if (dt > tt) {
                    dt = tt;
                    best = q;
                    wall = false;
                }
            }
        }
    }
    partner = wall ? EDMD_WALL_CODE(xy) : best;
    return dt;
}

struct Owner {
    int p;
    double x, y, vx, vy, r;
    int cx, cy;
    double vr; /* radius growth rate (growth starts only) */
};

/* Wave-cooperative neighbour scan for one or two owners at once.
 *
 * Lanes 0-8 read the nine cell counts of owner a, lanes 16-24 those of
 * owner b; a segmented scan lays each owner's candidates out in traversal
 * order (y offset outer, x offset inner, newest list entry first, EDMD.c:
 * 2959-3003).  Lane f evaluates candidate f of a and candidate f of b, so
 * both owners share the slot and record round trips.  The minimum over
 * (time, traversal rank) among candidates with time < NEVER is the
 * reference's strict-less first minimum; no candidate gives NEVER and
 * partner 0 (EDMD.c:2830-2831, 3091-3093).  Results are uniform. */
template <bool G, bool W = true, class AR>
__device__ __forceinline__ void predict_owners_gather(const EdmdConst& C, const AR& A,
                                                      int lane, double t, int nown,
                                                      const Owner& a, const Owner& b, double& dta,
                                                      int& pa, uint32_t& ca, double& dtb, int& pb,
                                                      uint32_t& cb, bool& overlap) {
    const int sub = lane & 15;
    const int grp = lane >> 4;
    int cell = 0, cnt = 0;
    if (sub < 9 && (grp == 0 || (grp == 1 && nown == 2))) {
        const int ocx = grp ? b.cx : a.cx;
        const int ocy = grp ? b.cy : a.cy;
        cell = wrap_cell(ocy + sub / 3 - 1, C.ncy) * C.ncx + wrap_cell(ocx + sub % 3 - 1, C.ncx);
        cnt = A.ccnt[cell];
    }
    /* inclusive scan within each row of 16 lanes: DPP row shifts right by
     * 1, 2, 4, 8, zeros shifted in at the row start */
    int incl = cnt;
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, true);
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, true);
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, true);
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, true);
    const int excl = incl - cnt;
    /* slot index of candidate f of the cell this lane read is rowbase - f
     * (newest entry first) */
    const int64_t rowbase = (int64_t)cell * C.cap + cnt - 1 + excl;
    const int totA = __builtin_amdgcn_readlane(incl, 8);
    const int totB = (nown == 2) ? __builtin_amdgcn_readlane(incl, 24) : 0;
    const int tot = totA > totB ? totA : totB;

    /* running minimum, uniform: candidates come in traversal order, so a
     * later pass only wins with a strictly earlier time, and within a pass
     * the lowest lane among equal times is the earliest in traversal */
    const double lim = G ? EDMD_GROW_NONE : EDMD_NEVER;
    double btA = INF_D, btB = INF_D;
    int bqA = 0, bqB = 0;
    uint32_t bcA = 0, bcB = 0;
    for (int base = 0; base < tot; base += 64) {
        const int f = base + lane;
        /* the cell holding candidate f: the last of the nine whose
         * exclusive prefix is <= f (scalar reads and selects, no LDS) */
        int64_t rbA = lane_value64(rowbase, 0), rbB = lane_value64(rowbase, 16);
#pragma unroll
        for (int s = 1; s < 9; ++s) {
            const int64_t ra = lane_value64(rowbase, s), rb = lane_value64(rowbase, 16 + s);
            if (f >= __builtin_amdgcn_readlane(excl, s)) rbA = ra;
            if (f >= __builtin_amdgcn_readlane(excl, 16 + s)) rbB = rb;
        }
        int qA = -1, qB = -1;
        if (f < totA) qA = A.cslot[rbA - f];
        if (f < totB) qB = A.cslot[rbB - f];
        const bool vA = qA >= 0 && qA != a.p;
        const bool vB = qB >= 0 && qB != b.p;
        EdmdPart rA, rB;
        double gA = 0, gB = 0;
        uint32_t cA = 0, cB = 0;
        if (vA) { rA = A.p[qA]; cA = rA.coll; }
        if (vB) { rB = A.p[qB]; cB = rB.coll; }
        if (W && !G && (C.field_on || C.damping)) {
            /* the neighbours flown to t and stored (EDMD.c:2664-2666) */
            if (vA) {
                fly_v<true>(C, t, rA.x, rA.y, rA.vx, rA.vy, rA.tl);
                const auto w = A.p + qA;
                w->x = rA.x; w->y = rA.y; w->vx = rA.vx; w->vy = rA.vy; w->tl = rA.tl;
            }
            if (vB) {
                fly_v<true>(C, t, rB.x, rB.y, rB.vx, rB.vy, rB.tl);
                const auto w = A.p + qB;
                w->x = rB.x; w->y = rB.y; w->vx = rB.vx; w->vy = rB.vy; w->tl = rB.tl;
            }
        }
        if (G && vA) gA = A.vr[qA];
        if (G && vB) gB = A.vr[qB];
        double tA = INF_D, tB = INF_D;
        if (vA) {
            double tt;
            if (G)
                tt = coll_time_grow(C, t, a.x, a.y, a.vx, a.vy, a.r, a.vr, rA.x, rA.y, rA.vx,
                                    rA.vy, rA.tl, rA.rad, gA, overlap);
            else
                tt = log_time<W>(C, coll_time_nb(C, t, a.x, a.y, a.vx, a.vy, a.r, rA.x, rA.y,
                                                 rA.vx, rA.vy, rA.tl, rA.rad, overlap));
            if (tt < lim) tA = tt;
        }
        if (vB) {
            double tt;
            if (G)
                tt = coll_time_grow(C, t, b.x, b.y, b.vx, b.vy, b.r, b.vr, rB.x, rB.y, rB.vx,
                                    rB.vy, rB.tl, rB.rad, gB, overlap);
            else
                tt = log_time<W>(C, coll_time_nb(C, t, b.x, b.y, b.vx, b.vy, b.r, rB.x, rB.y,
                                                 rB.vx, rB.vy, rB.tl, rB.rad, overlap));
            if (tt < lim) tB = tt;
        }
        double mA;
        const int wA = wave_argmin_lane(tA, 0.0, mA);
        if (mA < btA) {
            const int w = wA;
            btA = mA;
            bqA = __builtin_amdgcn_readlane(qA, w);
            bcA = (uint32_t)__builtin_amdgcn_readlane((int)cA, w);
        }
        if (nown == 2) {
            double mB;
            const int wB = wave_argmin_lane(tB, 0.0, mB);
            if (mB < btB) {
                const int w = wB;
                btB = mB;
                bqB = __builtin_amdgcn_readlane(qB, w);
                bcB = (uint32_t)__builtin_amdgcn_readlane((int)cB, w);
            }
        }
    }
    if (btA < lim) {
        dta = btA; pa = bqA; ca = bcA;
    } else {
        dta = lim; pa = 0; ca = A.p[0].coll;
    }
    if (btB < lim) {
        dtb = btB; pb = bqB; cb = bcB;
    } else {
        dtb = lim; pb = 0; cb = A.p[0].coll;
    }
}

/* Largest cell capacity the staged scan handles (six 16-byte row chunks
 * per cell), and its LDS stage: every slot of 18 cells. */
#define EDMD_STAGE_CAP 24
#define EDMD_STAGE_MAX (18 * EDMD_STAGE_CAP)

/* Inclusive prefix sum over the wave of values held by lanes 0-31 (rows 0
 * and 1): DPP row shifts within each row of 16, then row 0's total added
 * to row 1 (row_bcast:15). */
__device__ __forceinline__ int scan32(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    return v;
}

/* Neighbour scan for one or two owners, candidates staged in LDS.
 *
 * Lane e < 9 reads cell e of owner a's 3x3 block, lane 9 + e cell e of
 * owner b's (y offset outer, x offset inner): its count and its whole slot
 * row in the same round trip.  A scan over the 18 counts places every
 * candidate at its traversal position -- owner a's first, then owner b's,
 * each cell's newest entry first (EDMD.c:2959-3003) -- and each lane writes
 * its cell's entries there in LDS.  Lane f then evaluates candidate f
 * against its owner, so the two owners share one pass when their
 * candidates fit in 64 lanes.  Per owner, the minimum time over eligible
 * candidates (time < NEVER), the lowest lane among equal times, and across
 * passes only a strictly earlier time: the reference's strict-less first
 * minimum.  Requires cap <= EDMD_STAGE_CAP, cap % 4 == 0 and cell rows
 * aligned to 16 bytes. */
template <bool G, bool W = true, bool PF = false, class AR>
__device__ __forceinline__ void predict_owners_staged(const EdmdConst& C, const AR& A,
                                                      int lane, double t, int nown,
                                                      const Owner& a, const Owner& b, double& dta,
                                                      int& pa, uint32_t& ca, double& dtb, int& pb,
                                                      uint32_t& cb, bool& overlap,
                                                      int32_t* stage) {
    const int nent = 9 * nown;
    const bool ent = lane < nent;
    const bool inB = lane >= 9;
    const int sub = inB ? lane - 9 : lane;
    int cnt = 0;
    int4 row[EDMD_STAGE_CAP / 4];
    if (ent) {
        const int ocx = inB ? b.cx : a.cx;
        const int ocy = inB ? b.cy : a.cy;
        const int cell =
            wrap_cell(ocy + sub / 3 - 1, C.ncy) * C.ncx + wrap_cell(ocx + sub % 3 - 1, C.ncx);
        const int4* src = (const int4*)(A.cslot + (int64_t)cell * C.cap);
        cnt = A.ccnt[cell];
#pragma unroll
        for (int c = 0; c < EDMD_STAGE_CAP / 4; ++c)
            if (4 * c < C.cap) row[c] = src[c];
    }
    const int incl = scan32(cnt);
    const int excl = incl - cnt;
    const int totA = __builtin_amdgcn_readlane(incl, 8);
    const int tot = __builtin_amdgcn_readlane(incl, nent - 1);
    if (ent) {
        /* slot k (k-th oldest) goes to excl + cnt - 1 - k */
        const int top = excl + cnt - 1;
#pragma unroll
        for (int c = 0; c < EDMD_STAGE_CAP / 4; ++c) {
            if (4 * c >= cnt) break;
            stage[top - 4 * c] = row[c].x;
            if (4 * c + 1 < cnt) stage[top - 4 * c - 1] = row[c].y;
            if (4 * c + 2 < cnt) stage[top - 4 * c - 2] = row[c].z;
            if (4 * c + 3 < cnt) stage[top - 4 * c - 3] = row[c].w;
        }
    }
    WAVE_SYNC();

    const double lim = G ? EDMD_GROW_NONE : EDMD_NEVER;
    double btA = INF_D, btB = INF_D;
    int bqA = 0, bqB = 0;
    uint32_t bcA = 0, bcB = 0;
    /* PF: the next pass's candidate records are read while this pass is
     * evaluated (one fewer dependent round trip per extra pass; the records
     * cost registers, so only the loop of large systems does it) */
    EdmdPart rnext;
    int qnext = -1;
    if (PF) {
        qnext = lane < tot ? stage[lane] : -1;
        const int own0 = lane >= totA ? b.p : a.p;
        if (qnext >= 0 && qnext != own0) rnext = A.p[qnext];
    }
    for (int base = 0; base < tot; base += 64) {
        const int f = base + lane;
        const bool isB = f >= totA;
        int q;
        EdmdPart rcur;
        if (PF) {
            q = qnext;
            rcur = rnext;
            const int fn = f + 64;
            qnext = fn < tot ? stage[fn] : -1;
            const int ownn = fn >= totA ? b.p : a.p;
            if (qnext >= 0 && qnext != ownn) rnext = A.p[qnext];
        } else {
            q = f < tot ? stage[f] : -1;
        }
        const int own = isB ? b.p : a.p;
        const bool v = q >= 0 && q != own;
        double tq = INF_D;
        uint32_t cq = 0;
        if (v) {
            EdmdPart r = PF ? rcur : A.p[q];
            if (W && !G && (C.field_on || C.damping)) {
                /* the neighbour flown to t and stored (EDMD.c:2664-2666); a
                 * disk listed for both owners is written twice alike */
                fly_v<true>(C, t, r.x, r.y, r.vx, r.vy, r.tl);
                const auto w = A.p + q;
                w->x = r.x;
                w->y = r.y;
                w->vx = r.vx;
                w->vy = r.vy;
                w->tl = r.tl;
            }
            cq = r.coll;
            const double ox = isB ? b.x : a.x, oy = isB ? b.y : a.y;
            const double ovx = isB ? b.vx : a.vx, ovy = isB ? b.vy : a.vy;
            const double orad = isB ? b.r : a.r;
            double tt;
            if (G) {
                const double ovr = isB ? b.vr : a.vr;
                tt = coll_time_grow(C, t, ox, oy, ovx, ovy, orad, ovr, r.x, r.y, r.vx, r.vy, r.tl,
                                    r.rad, A.vr[q], overlap);
            } else {
                tt = log_time<W>(C, coll_time_nb(C, t, ox, oy, ovx, ovy, orad, r.x, r.y, r.vx,
                                                 r.vy, r.tl, r.rad, overlap));
            }
            if (tt < lim) tq = tt;
        }
        const bool hasA = base < totA, hasB = base + 64 > totA && nown == 2;
        const double tA = isB ? INF_D : tq;
        const double tB = isB ? tq : INF_D;
        if (hasA && hasB) {
            double mA, mB;
            int wA, wB;
            wave_argmin_lane2(tA, tB, 0.0, mA, wA, mB, wB);
            if (mA < btA) {
                btA = mA;
                bqA = __builtin_amdgcn_readlane(q, wA);
                bcA = (uint32_t)__builtin_amdgcn_readlane((int)cq, wA);
            }
            if (mB < btB) {
                btB = mB;
                bqB = __builtin_amdgcn_readlane(q, wB);
                bcB = (uint32_t)__builtin_amdgcn_readlane((int)cq, wB);
            }
        } else if (hasA) {
            double mA;
            const int w = wave_argmin_lane(tA, 0.0, mA);
            if (mA < btA) {
                btA = mA;
                bqA = __builtin_amdgcn_readlane(q, w);
                bcA = (uint32_t)__builtin_amdgcn_readlane((int)cq, w);
            }
        } else if (hasB) {
            double mB;
            const int w = wave_argmin_lane(tB, 0.0, mB);
            if (mB < btB) {
                btB = mB;
                bqB = __builtin_amdgcn_readlane(q, w);
                bcB = (uint32_t)__builtin_amdgcn_readlane((int)cq, w);
            }
        }
    }
    WAVE_SYNC();
    if (btA < lim) {
        dta = btA; pa = bqA; ca = bcA;
    } else {
        dta = lim; pa = 0; ca = A.p[0].coll;
    }
    if (btB < lim) {
        dtb = btB; pb = bqB; cb = bcB;
    } else {
        dtb = lim; pb = 0; cb = A.p[0].coll;
    }
}

/* The staged scan where the cell capacity allows it, else the gather scan
 * from HBM; both give the reference's result bit for bit. */
template <bool G, bool W = true, bool PF = false, class AR>
__device__ __forceinline__ void predict_owners(const EdmdConst& C, const AR& A, int lane,
                                               double t, int nown, const Owner& a, const Owner& b,
                                               double& dta, int& pa, uint32_t& ca, double& dtb,
                                               int& pb, uint32_t& cb, bool& overlap,
                                               int32_t* stage) {
    if (C.cap <= EDMD_STAGE_CAP && (C.cap & 3) == 0)
        predict_owners_staged<G, W, PF>(C, A, lane, t, nown, a, b, dta, pa, ca, dtb, pb, cb, overlap,
                                    stage);
    else
        predict_owners_gather<G, W>(C, A, lane, t, nown, a, b, dta, pa, ca, dtb, pb, cb, overlap);
    if (W && C.walls) {
        /* the wall time starts the reference's minimum: a neighbour
         * replaces it only when strictly earlier */
        auto wt = [&](const Owner& o, double& dw, int& xy) {
            const int typ = (C.walls & 8) ? A.ptype[o.p] : 0;
            return G ? wall_time_grow(C, o.x, o.y, o.vx, o.vy, o.r, o.vr, o.cx, o.cy, dw, xy, typ)
                     : wall_time(C, o.x, o.y, o.vx, o.vy, o.r, o.cx, o.cy, dw, xy, typ);
        };
        double dw;
        int xy;
        if (wt(a, dw, xy) && !(dta < dw)) {
            dta = dw;
            pa = EDMD_WALL_CODE(xy);
            ca = 0;
        }
        if (nown == 2 && wt(b, dw, xy) && !(dtb < dw)) {
            dtb = dw;
            pb = EDMD_WALL_CODE(xy);
            cb = 0;
        }
    }
}

/* ------------------------------------------------------------ calendar */

/* Minimum over the wave by (key, rank), carrying the leaf and the event
 * code.  A node's rank is that of its leaf (lrank[leaf]); ranks are read
 * only when several lanes hold the minimum key, which is rare.  Keys are
 * never NaN; among +inf keys (padding, parked events) the lowest lane wins.
 * base: a time near the keys (the current time), for the 32-bit search. */
/* The part of wave_argmin after the key search: kk is this lane's order
 * key, mkk the minimum key over the wave. */
template <class AR>
__device__ __forceinline__ void wave_argmin_finish(const AR& A, int kk, int mkk, double& k,
                                                   int& i, int& x) {
    const uint64_t near = __ballot(kk == mkk);
    double mk;
    uint64_t tied;
    if (__popcll(near) == 1) {
        tied = near;
        mk = lane_value(k, (int)__builtin_ctzll(near));
    } else {
        const double kv = (kk == mkk) ? k : INF_D;
        mk = wave_min(kv);
        tied = __ballot(kv == mk);
    }
    int w = (int)__builtin_ctzll(tied);
    if (__popcll(tied) > 1 && mk < INF_D) {
        /* equal keys: the earlier in the reference's queue (lower rank) */
        /* loaded by every lane (padding lanes read leaf 0) and selected,
         * not loaded under a divergent branch */
        const bool tie = ((tied >> (threadIdx.x & 63)) & 1) != 0;
        const int64_t rl = A.lrank[tie ? i : 0];
        int64_t r = tie ? rl : INT64_MAX;
        int l = tie ? (int)(threadIdx.x & 63) : 64;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const long long orr = __shfl_xor((long long)r, m, 64);
            const int ol = __shfl_xor(l, m, 64);
            if (orr < r || (orr == r && ol < l)) {
                r = orr;
                l = ol;
            }
        }
        w = uni(l);
    }
    k = mk;
    i = __builtin_amdgcn_readlane(i, w);
    x = __builtin_amdgcn_readlane(x, w);
}

template <class AR>
__device__ __forceinline__ void wave_argmin(const AR& A, double base, double& k, int& i,
                                            int& x) {
    const int kk = order_key(k, base);
    wave_argmin_finish(A, kk, wave_min_i32(kk), k, i, x);
}

/* wave_argmin of two independent (key, leaf, code) sets, the key searches
 * interleaved; each result is the one wave_argmin gives for its set. */
template <class AR>
__device__ __forceinline__ void wave_argmin2(const AR& A, double base, double& ka, int& ia,
                                             int& xa, double& kb, int& ib, int& xb) {
    const int kka = order_key(ka, base), kkb = order_key(kb, base);
    int ma = kka, mb = kkb;
    wave_min_i32x2(ma, mb);
    wave_argmin_finish(A, kka, ma, ka, ia, xa);
    wave_argmin_finish(A, kkb, mb, kb, ib, xb);
}

/* Level l of the calendar: the leaves for l = 0, else nodes of level l. */
template <class AR>
__device__ __forceinline__ auto level(const AR& A, int l) -> decltype(A.leaf) {
    return l == 0 ? A.leaf : A.lev[l - 1];
}

/* Earliest pending particle event: minimum of the top level, read from
 * `top` (the LDS copy the event loop keeps, or the level in HBM). */
template <class AR>
__device__ __forceinline__ void top_min(const EdmdConst& C, const AR& A, int lane,
                                        const EdmdNode* top, double base, double& key, int& leaf,
                                        int& aux) {
    double k = INF_D;
    int i = 0x7fffffff, x = 0;
    if (lane < C.lsize[C.nlev]) {
        const EdmdNode nd = top[lane];
        k = nd.key;
        i = nd.idx;
        x = nd.aux;
    }
    wave_argmin(A, base, k, i, x);
    key = k;
    leaf = uni(i);
    aux = uni(x);
}

/* The 64 children of every node on the calendar paths of up to two leaves,
 * read in one round trip before the event changes anything.  ML bounds the
 * number of levels above the leaves. */
template <int ML>
struct PathCache {
    double k[2][ML];
    int i[2][ML];
    int x[2][ML];
};

template <int ML, class AR>
__device__ __forceinline__ void path_prefetch(const EdmdConst& C, const AR& A, int lane,
                                              int pA, int pB, PathCache<ML>& P) {
#pragma unroll
    for (int l = 0; l < ML; ++l) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int p = h ? pB : pA;
            double k = INF_D;
            int i = 0x7fffffff, x = 0;
            if (l < C.nlev && p >= 0) {
                const int c = ((p >> (6 * (l + 1))) << 6) + lane;
                if (c < C.lsize[l]) {
                    const EdmdNode nd = level(A, l)[c];
                    k = nd.key;
                    i = nd.idx;
                    x = nd.aux;
                }
            }
            P.k[h][l] = k;
            P.i[h][l] = i;
            P.x[h][l] = x;
        }
    }
}

/* Stores the new leaves of pA (and pB >= 0) with their ranks and recomputes
 * every node on their paths from the prefetched children, substituting the
 * freshly computed path nodes; where the paths meet one node takes both.
 * With `top` (the event loop's LDS copy of the top level) the top-level
 * nodes are written there instead of to HBM. */
template <int ML, class AR>
__device__ __forceinline__ void path_update(const EdmdConst& C, const AR& A, int lane,
                                            const PathCache<ML>& P, int pA, Leaf LA, int pB,
                                            Leaf LB, EdmdNode* top, double base) {
    double kA = LA.key, kB = LB.key;
    int xA = LA.aux, xB = LB.aux;
    if (lane == 0) {
        EdmdNode n;
        n.key = kA;
        n.idx = pA;
        n.aux = xA;
        A.leaf[pA] = n;
        A.lrank[pA] = LA.rank;
        if (top && C.nlev == 0) top[pA] = n;
        if (pB >= 0) {
            n.key = kB;
            n.idx = pB;
            n.aux = xB;
            A.leaf[pB] = n;
            A.lrank[pB] = LB.rank;
            if (top && C.nlev == 0) top[pB] = n;
        }
    }
    WAVE_SYNC();
    int iA = pA, iB = pB;
    int cA = pA, cB = pB;
#pragma unroll
    for (int l = 0; l < ML; ++l) {
        if (l >= C.nlev) break;
        const int gA = cA >> 6;
        const int gB = pB >= 0 ? (cB >> 6) : -1;
        int c = (gA << 6) + lane;
        double k = P.k[0][l];
        int i = P.i[0][l], x = P.x[0][l];
        if (c == cA) { k = kA; i = iA; x = xA; }
        else if (pB >= 0 && c == cB) { k = kB; i = iB; x = xB; }
        double nkA, nkB;
        int niA, nxA, niB, nxB;
        if (pB >= 0 && gB != gA) {
            /* two different nodes: both minima at once */
            const int c2 = (gB << 6) + lane;
            double k2 = P.k[1][l];
            int i2 = P.i[1][l], x2 = P.x[1][l];
            if (c2 == cB) { k2 = kB; i2 = iB; x2 = xB; }
            wave_argmin2(A, base, k, i, x, k2, i2, x2);
            nkB = k2;
            niB = uni(i2);
            nxB = uni(x2);
            nkA = k;
            niA = uni(i);
            nxA = uni(x);
        } else {
            wave_argmin(A, base, k, i, x);
            nkA = k;
            niA = uni(i);
            nxA = uni(x);
            nkB = nkA;
            niB = niA;
            nxB = nxA;
        }
        if (lane == 0) {
            EdmdNode n;
            n.key = nkA;
            n.idx = niA;
            n.aux = nxA;
            EdmdNode nb;
            nb.key = nkB;
            nb.idx = niB;
            nb.aux = nxB;
            const bool two = pB >= 0 && gB != gA;
            if (top && l + 1 == C.nlev) {
                top[gA] = n;
                if (two) top[gB] = nb;
            } else {
                A.lev[l][gA] = n;
                if (two) A.lev[l][gB] = nb;
            }
        }
        kA = nkA; iA = niA; xA = nxA;
        kB = nkB; iB = niB; xB = nxB;
        cA = gA;
        cB = gB;
    }
    WAVE_SYNC();
}

/* Sequential minimum over the children of node g of level l (l >= 1) by
 * (key, rank).  Every node holds a real leaf index (0 <= idx < n), so each
 * child's rank is loaded unconditionally and the winner chosen by selects:
 * a rank load under a divergent branch here has been compiled into code
 * that kept the previous winner's event code on a tie (gfx950, ROCm 7.2). */
template <class AR>
__device__ __forceinline__ void node_min(const EdmdConst& C, const AR& A, int l, int g) {
    int lo = g << 6;
    int hi = lo + 64;
    if (hi > C.lsize[l - 1]) hi = C.lsize[l - 1];
    const auto ch = level(A, l - 1);
    const EdmdNode b0 = ch[lo];
    double bk = b0.key;
    int bi = b0.idx, bx = b0.aux;
    int64_t br = A.lrank[bi];
    for (int c = lo + 1; c < hi; ++c) {
        const EdmdNode nd = ch[c];
        const int64_t r = A.lrank[nd.idx];
        const bool take = nd.key < bk || (nd.key == bk && bk < INF_D && r < br);
        bk = take ? nd.key : bk;
        bi = take ? nd.idx : bi;
        bx = take ? nd.aux : bx;
        br = take ? r : br;
    }
    EdmdNode best;
    best.key = bk;
    best.idx = bi;
    best.aux = bx;
    A.lev[l - 1][g] = best;
}

/* ------------------------------------------------------------ sweep */

/* One particle's full prediction (crossing + collision) written to its two
 * slots and its leaf; the particle is assumed synchronised at t.  G selects
 * the growth functions.  The slots are scheduled as insertions number
 * ctr_cross and ctr_coll with the queue at paulTime = paul.  stop_grow: the
 * partner counter recorded is the one stopGrow's interleaved loop sees
 * (EDMD.c:4770-4776: every counter goes 0 -> 1 in particle order, each disk
 * predicting right after its own increment), i.e. 1 for a partner of lower
 * or equal index, else 0. */
template <bool G, bool W = true, class AR>
__device__ __forceinline__ bool sweep_one(const EdmdConst& C, const AR& A, double t, int p,
                                          double paul, int64_t ctr_cross, int64_t ctr_coll,
                                          bool stop_grow = false) {
    const EdmdPart o = A.p[p];
    double tc;
    int dir;
    /* crossingEventGrow while growing: no field, no logTime */
    cross_time<W && !G>(C, t, o.x, o.y, o.vx, o.vy, o.cx, o.cy, tc, dir);
    bool ov = false;
    int partner;
    double dt = predict_seq<G, W>(C, A, t, p, o, partner, ov);
    double tk = t + dt;
    EdmdEv e;
    e.tcoll = tk;
    e.tcross = tc;
    e.partner = partner;
    e.cact = stop_grow ? (partner <= p ? 1u : 0u) : (partner >= 0 ? A.p[partner].coll : 0u);
    e.dir = dir;
    e.pad = 0;
    A.e[p] = e;
    const int64_t rk = rank_of(tk, paul, C.dtp, C.horizon, ctr_coll);
    const int64_t rx = rank_of(tc, paul, C.dtp, C.horizon, ctr_cross);
    A.erank[2 * p] = rk;
    A.erank[2 * p + 1] = rx;
    const Leaf L = make_leaf(tk, rk, tc, rx, partner, dir);
    EdmdNode n;
    n.key = L.key;
    n.idx = p;
    n.aux = L.aux;
    A.leaf[p] = n;
    A.lrank[p] = L.rank;
    return ov;
}

/* Cell-tiled sweep: one workgroup per SW_TX x SW_TY cells.  The tile and a
 * one-cell halo (wrapped periodically) are staged into LDS cell by cell in
 * row-major order, newest list entry first within a cell, so the three
 * cells of one row of a disk's 3x3 neighbourhood are one contiguous run
 * that lists the candidates in the reference's scan order (y offset outer,
 * x offset inner, newest entry first, EDMD.c:2959-3003).  Each thread
 * predicts one of the tile's disks over its three runs as one flat
 * candidate loop whose trip count is uniform over the wave; the loop keeps
 * only the cheap terms (approach b, contact d, discriminant) and marks the
 * candidates that can collide, and the roots of the marked ones are solved
 * afterwards in the same order with the strict-less rule, so the result is
 * that of predict_seq bit for bit.  Every disk record is read from HBM by
 * the one tile that owns it and the (at most eight) tiles whose halo holds
 * it, instead of once per neighbour scan that touches it.  Workgroups are
 * dealt to the eight XCDs round robin; tiles are renumbered so that each
 * XCD sweeps a contiguous band of tile rows and halo re-reads hit its own
 * L2.  A tile whose staged disks exceed SW_STAGE falls back to predict_seq
 * from global memory. */
/* Tile shapes of the sweep: TX x TY cells per workgroup of THREADS
 * threads; STAGE staged disks at most (tile plus halo). */
template <int TX_, int TY_, int THREADS_>
struct SweepShape {
    static constexpr int TX = TX_, TY = TY_, THREADS = THREADS_;
    static constexpr int HW = TX + 2, HH = TY + 2, NH = HW * HH;
    static constexpr int STAGE = 512;
    static_assert(NH <= 255, "staged cells carry a u8 index");
    static_assert(THREADS >= 64 && THREADS % 64 == 0, "whole waves");
};

/* The shapes a batch chooses from (edmd_batch_autotune times them; the
 * default is picked from the disks per cell).  At phi = 0.8 with unit
 * diameter a cell holds ~4.1 disks, at phi = 0.7 ~3.6, with the reference
 * CLI's radius 1 about one. */
using SweepV0 = SweepShape<8, 8, 320>;
using SweepV1 = SweepShape<6, 10, 256>;
using SweepV2 = SweepShape<10, 6, 256>;
using SweepV3 = SweepShape<8, 7, 256>;
using SweepV4 = SweepShape<8, 6, 192>;
using SweepV5 = SweepShape<8, 8, 256>;
using SweepV6 = SweepShape<8, 8, 64>;
using SweepV7 = SweepShape<16, 8, 128>;

/* The workgroup's LDS.  POLY = false is the layout of launches whose
 * systems all have equal radii: no radius per staged disk (4 KB less, one
 * more workgroup per CU).  The cell counts and the own-cell counts of the
 * two scans live in the velocity stage, which is written only after them. */
template <class S, bool POLY_>
struct SweepLds {
    static constexpr bool POLY = POLY_;
    /* staged disks: position flown to the sweep time, velocity, radius,
     * index, staged cell */
    double2 pos[S::STAGE], vel[S::STAGE];
    double rad[POLY ? S::STAGE : 1];
    int32_t id[S::STAGE];
    uint32_t coll[S::STAGE]; /* collision counter (the partner counter a prediction records) */
    uint8_t hcell[S::STAGE];
    int16_t own[S::STAGE];     /* staged index of each of the tile's own disks */
    int32_t start[S::NH + 1];  /* first staged index of each cell; [S::NH] = total */
    int32_t ostart[S::NH];     /* first own-list index of each tile cell */
    int32_t gcell[S::NH];
    int32_t ninner, total;
    static_assert(2 * S::NH * sizeof(int32_t) <= sizeof(vel), "scan counts fit the velocity stage");
    __device__ int32_t* cnt() { return reinterpret_cast<int32_t*>(&vel[0]); }
    __device__ int32_t* owncnt() { return reinterpret_cast<int32_t*>(&vel[0]) + S::NH; }
};

__device__ __forceinline__ int mod_cell(int a, int nc) {
    int r = a % nc;
    return r < 0 ? r + nc : r;
}

/* Exclusive scan of v over the NH staged cells by one wave (PER
 * consecutive cells per lane); returns the total.  Cells are ordered
 * h = (ly + 1) * HW + lx + 1. */
template <int NH>
__device__ __forceinline__ int sweep_scan(int lane, const int* v, int* out, int base) {
    constexpr int PER = (NH + 63) / 64;
    int val[PER];
    int sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int h = PER * lane + i;
        val[i] = h < NH ? v[h] : 0;
        sum += val[i];
    }
    int incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int u = __shfl_up(incl, d, 64);
        if (lane >= d) incl += u;
    }
    int run = base + incl - sum;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int h = PER * lane + i;
        if (h < NH) out[h] = run;
        run += val[i];
    }
    return __shfl(incl, 63, 64);
}

/* The terms of collisionTime (EDMD.c:2661-2725) that decide whether a pair
 * can collide: relative position (minimum image when WRAP) and velocity,
 * approach b, |dv|^2, contact term d and discriminant, in the reference's
 * operation order. */
template <bool WRAP>
__device__ __forceinline__ void pair_terms(const EdmdConst& C, double ox, double oy, double ovx,
                                           double ovy, double orad, const double2 pv,
                                           const double2 vv, double r2, double& b, double& v2,
                                           double& d, double& det) {
    const double dvx = vv.x - ovx;
    const double dvy = vv.y - ovy;
    double dx = pv.x - ox;
    double dy = pv.y - oy;
    if (WRAP) min_image(C, dx, dy);
    b = dx * dvx + dy * dvy;
    v2 = dvx * dvx + dvy * dvy;
    d = dx * dx + dy * dy - 4 * orad * r2; /* DIST2, EDMD.c:35-43 */
    det = b * b - v2 * d;
}

/* The bits of the 64-bit mark word starting at candidate base that stay
 * set: candidates k < tot other than kself. */
__device__ __forceinline__ uint64_t keep_bits(int tot, int kself, int base) {
    const int r = tot - base;
    uint64_t m = r >= 64 ? ~(uint64_t)0 : (r <= 0 ? (uint64_t)0 : ((uint64_t)1 << r) - 1);
    const int sb = kself - base;
    if (sb >= 0 && sb < 64) m &= ~((uint64_t)1 << sb);
    return m;
}

/* The predictions of a staged tile's own disks, one thread per disk.  The
 * candidates of a disk are its three row runs, k = 0 .. tot-1 in scan
 * order.  Pass 1 runs k over the wave's largest tot (uniform) and marks in
 * a 256-bit mask the candidates with b <= 0 and det >= 0 (every other one
 * gives the reference's "never", which the strict-less minimum never
 * takes), flagging overlaps as collisionTime does; pass 2 solves the marked
 * roots in increasing k.  A disk with more than 256 candidates solves its
 * roots in pass 1 instead. */
template <class S, bool WRAP, bool X, bool MONO, class AR, class LD>
__device__ __forceinline__ void sweep_tile_predict(const EdmdConst& C, const AR& A, const LD& L,
                                                   int tid, int ninner, double t, int tx0, int ty0,
                                                   double paul, int64_t ctr0, double rad0,
                                                   bool& ov) {
    static_assert(MONO || LD::POLY, "unequal radii need the staged radii");
    const int lane = tid & 63;
    for (int u0 = tid - lane; u0 < ninner; u0 += S::THREADS) { /* uniform over the wave */
        const int u = u0 + lane;
        const bool act = u < ninner;
        const int me = L.own[act ? u : 0]; /* idle lanes shadow the first disk */
        const int h = L.hcell[me];
        const int lx = h % S::HW - 1, ly = h / S::HW - 1;
        /* the owner is synchronised at t (its staged position is its own) */
        const double2 opos = L.pos[me], ovel = L.vel[me];
        const double ox = opos.x, oy = opos.y, ovx = ovel.x, ovy = ovel.y;
        const double orad = LD::POLY ? L.rad[me] : rad0;
        const int p = L.id[me];
        /* row runs: cells (lx - 1 .. lx + 1, ly + j) are staged contiguously */
        const int h0 = ly * S::HW + lx;
        const int a0 = L.start[h0], n0 = L.start[h0 + 3] - a0;
        const int a1 = L.start[h0 + S::HW], n1 = L.start[h0 + S::HW + 3] - a1;
        const int a2 = L.start[h0 + 2 * S::HW], n2 = L.start[h0 + 2 * S::HW + 3] - a2;
        const int c1 = n0, c2 = n0 + n1;
        const int o0 = a0, o1 = a1 - c1, o2 = a2 - c2;
        const int tot = act ? c2 + n2 : 0;
        const bool direct = tot > 256;
        const int tot1 = direct ? 0 : tot;
        const int ktot = -wave_min_i32(-tot1);
        uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0, cur = 0;
        double dt = EDMD_NEVER;
        int best = -1;
        int wxy = 2;
        bool wall = false;
        if (X && C.walls)
            wall = wall_time(C, ox, oy, ovx, ovy, orad, tx0 + lx, ty0 + ly, dt, wxy,
                             (C.walls & 8) ? A.ptype[p] : 0);
        /* pass 1 without branches: every lane evaluates its k-th candidate;
         * the marks of consecutive k collect in cur and go to their word
         * when k leaves it (k is uniform over the wave) */
        if constexpr (!WRAP) {
            /* The staged index advances by one per k and moves to the next
             * row run where that run starts; a lane past its own count reads
             * a staged disk clamped into the stage.  Only the discriminant
             * terms stay in the loop: the owner itself (k = kself) and the
             * k past the count are cleared from the marks afterwards, and
             * the overlap test moves to pass 2, since a pair with d < 0 and
             * b <= 0 always has det >= 0 and is marked. */
            int v = a0 - 1;
            uint32_t c32 = 0;
            for (int k = 0; k < ktot; ++k) {
                v = v + 1;
                v = k == c1 ? a1 : v;
                v = k == c2 ? a2 : v;
                const int vc = min(v, S::STAGE - 1);
                double b, v2, d, det;
                pair_terms<false>(C, ox, oy, ovx, ovy, orad, L.pos[vc], L.vel[vc],
                                  MONO ? orad : L.rad[vc], b, v2, d, det);
                if (!(b > 0) && !(det < 0)) c32 |= 1u << (k & 31);
                if ((k & 31) == 31 || k == ktot - 1) {
                    const uint64_t c = (k & 32) ? (uint64_t)c32 << 32 : (uint64_t)c32;
                    switch (k >> 6) {
                        case 0: m0 |= c; break;
                        case 1: m1 |= c; break;
                        case 2: m2 |= c; break;
                        default: m3 |= c; break;
                    }
                    c32 = 0;
                }
            }
            const int kself = c1 + (me - a1); /* the owner lies in its middle row run */
            m0 &= keep_bits(tot1, kself, 0);
            m1 &= keep_bits(tot1, kself, 64);
            m2 &= keep_bits(tot1, kself, 128);
            m3 &= keep_bits(tot1, kself, 192);
        } else
        for (int k = 0; k < ktot; ++k) {
            const bool in = k < tot1;
            const int vk = k + (k < c1 ? o0 : (k < c2 ? o1 : o2));
            const int v = in ? vk : me;
            double b, v2, d, det;
            pair_terms<WRAP>(C, ox, oy, ovx, ovy, orad, L.pos[v], L.vel[v],
                             MONO ? orad : L.rad[v], b, v2, d, det);
            /* the owner itself; a halo that wraps (boxes under three tiles
             * wide) may stage the owner's cell more than once */
            const bool self = WRAP ? L.id[v] == p : v == me;
            const bool ok = in && !self && !(b > 0);
            ov = ov || (ok && d < -0.01); /* EDMD.c:2708-2717 */
            const bool mark = ok && !(det < 0);
            cur |= mark ? ((uint64_t)1 << (k & 63)) : (uint64_t)0;
            if ((k & 63) == 63 || k == ktot - 1) {
                switch (k >> 6) {
                    case 0: m0 = cur; break;
                    case 1: m1 = cur; break;
                    case 2: m2 = cur; break;
                    default: m3 = cur; break;
                }
                cur = 0;
            }
        }
        /* disks with more than 256 candidates solve their roots in order as
         * they go */
        if (__ballot(direct)) {
            if (direct) {
                for (int k = 0; k < tot; ++k) {
                    const int v = k + (k < c1 ? o0 : (k < c2 ? o1 : o2));
                    if (WRAP ? L.id[v] == p : v == me) continue;
                    double b, v2, d, det;
                    pair_terms<WRAP>(C, ox, oy, ovx, ovy, orad, L.pos[v], L.vel[v],
                                     MONO ? orad : L.rad[v], b, v2, d, det);
                    if (b > 0) continue;
                    if (d < -0.01) ov = true;
                    if (det < 0) continue;
                    const double tt = log_time<X>(C, (-b - sqrt(det)) / v2);
                    if (dt > tt) {
                        dt = tt;
                        best = v;
                    }
                }
            }
        }
        /* the screen (uniform; not under damping, whose logTime could
         * merge distinct times): with q = -b > 0, the computed root
         * (q - fl(sqrt(det)))/v2 of a pair with d > 0 is at least
         * d/(2q) - 2^-25 q/v2 (its exact value lies in [d/(2q), d/q] and
         * the cancellation in det costs at most sqrt(3 2^-53) q/v2), so a
         * mark with d v2 - 2^-19 q^2 > 2 q v2 dt (1 + 2^-40) has a root
         * above the lane's best so far dt and is skipped: the strict-less
         * minimum never takes it (a skip needs d > 0, so d <= 0 pairs are
         * always solved) */
        const bool scr = C.screen != 0 && !(X && C.damping);
        uint32_t nmark = 0, nsolve = 0, nmiss = 0;
        while (m0 | m1 | m2 | m3) {
            int k;
            if (m0) { k = __builtin_ctzll(m0); m0 &= m0 - 1; }
            else if (m1) { k = 64 + __builtin_ctzll(m1); m1 &= m1 - 1; }
            else if (m2) { k = 128 + __builtin_ctzll(m2); m2 &= m2 - 1; }
            else { k = 192 + __builtin_ctzll(m3); m3 &= m3 - 1; }
            const int v = k + (k < c1 ? o0 : (k < c2 ? o1 : o2));
            double b, v2, d, det;
            pair_terms<WRAP>(C, ox, oy, ovx, ovy, orad, L.pos[v], L.vel[v],
                             MONO ? orad : L.rad[v], b, v2, d, det);
            if (!WRAP) ov = ov || d < -0.01; /* EDMD.c:2708-2717 */
            if (scr) {
                ++nmark;
                const double q = -b;
                const bool skip = dt > 0 && dt < EDMD_NEVER &&
                                  d * v2 - 0x1p-19 * (q * q) > (2 * q * v2 * dt) * (1 + 0x1p-40);
                if (skip) {
                    if (C.screen == 2 && !(log_time<X>(C, (-b - sqrt(det)) / v2) > dt)) ++nmiss;
                    continue;
                }
                ++nsolve;
            }
            const double tt = log_time<X>(C, (-b - sqrt(det)) / v2);
            if (dt > tt) {
                dt = tt;
                best = v;
            }
        }
        if (scr && C.screen == 2) { /* counted in the check mode only: the
                                       atomics contend on one address */
            uint32_t s0 = act ? nmark : 0, s1 = act ? nsolve : 0, s2 = act ? nmiss : 0;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                s0 += __shfl_xor(s0, off, 64);
                s1 += __shfl_xor(s1, off, 64);
                s2 += __shfl_xor(s2, off, 64);
            }
            if (lane == 0) {
                atomicAdd((unsigned long long*)&A.st->scr_marks, (unsigned long long)s0);
                atomicAdd((unsigned long long*)&A.st->scr_solved, (unsigned long long)s1);
                if (s2) atomicAdd((unsigned long long*)&A.st->scr_miss, (unsigned long long)s2);
            }
        }
        if (!act) continue;
        double tc;
        int dir;
        cross_time<X>(C, t, ox, oy, ovx, ovy, tx0 + lx, ty0 + ly, tc, dir);
        const double tk = t + dt;
        const int partner = best >= 0 ? L.id[best] : (wall ? EDMD_WALL_CODE(wxy) : 0);
        EdmdEv e;
        e.tcoll = tk;
        e.tcross = tc;
        e.partner = partner;
        /* the partner's counter from the stage; no partner: disk 0's
         * (EDMD.c:3091-3093 files partner 0) */
        e.cact = best >= 0 ? L.coll[best] : (partner >= 0 ? A.p[partner].coll : 0u);
        e.dir = dir;
        e.pad = 0;
        A.e[p] = e;
        const int64_t rk = rank_of(tk, paul, C.dtp, C.horizon, ctr0 + 2 * (int64_t)p + 1);
        const int64_t rx = rank_of(tc, paul, C.dtp, C.horizon, ctr0 + 2 * (int64_t)p);
        A.erank[2 * p] = rk;
        A.erank[2 * p + 1] = rx;
        const Leaf Lf = make_leaf(tk, rk, tc, rx, partner, dir);
        EdmdNode n;
        n.key = Lf.key;
        n.idx = p;
        n.aux = Lf.aux;
        A.leaf[p] = n;
        A.lrank[p] = Lf.rank;
    }
}

template <class S, bool X, bool POLY>
__global__ __launch_bounds__(S::THREADS) void k_sweep_tiled(const EdmdSys* __restrict__ systems) {
    __shared__ SweepLds<S, POLY> L;
    int32_t* const cnt = L.cnt();
    int32_t* const vtmp = L.owncnt();
    const EdmdSys* Sy = systems + blockIdx.y;
    const EdmdConst& C = Sy->c;
    const EdmdArrays& A = Sy->a;
    const int tid = threadIdx.x;
    if (A.st->growing) return; /* k_sweep_grow predicts growing systems */
    const int ncx = C.ncx, ncy = C.ncy;
    const int ntx = (ncx + S::TX - 1) / S::TX, nty = (ncy + S::TY - 1) / S::TY;
    const int per = gridDim.x / 8; /* gridDim.x is a multiple of 8 */
    const int tile = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (tile >= ntx * nty) return; /* uniform over the workgroup */
    const int tx0 = (tile % ntx) * S::TX, ty0 = (tile / ntx) * S::TY;
    const int vw = min(S::TX, ncx - tx0), vh = min(S::TY, ncy - ty0);
    const double t = A.st->t;
    /* eventListInit order (EDMD.c:2007-2012): crossing then collision of
     * each particle in index order, after the insertions already counted */
    const double paul = A.st->paul;
    const int64_t ctr0 = A.st->ctr;

    for (int h = tid; h < S::NH; h += S::THREADS) {
        const int lx = h % S::HW - 1, ly = h / S::HW - 1;
        const int g = mod_cell(ty0 + ly, ncy) * ncx + mod_cell(tx0 + lx, ncx);
        L.gcell[h] = g;
        cnt[h] = A.ccnt[g];
    }
    __syncthreads();
    /* staged order: cells row-major (h), each cell newest entry first;
     * the tile's own cells also get consecutive places in the owner list */
    if (tid < 64) {
        constexpr int PER = (S::NH + 63) / 64;
        for (int h = PER * tid; h < PER * tid + PER && h < S::NH; ++h) {
            const int lx = h % S::HW - 1, ly = h / S::HW - 1;
            const bool own = lx >= 0 && lx < vw && ly >= 0 && ly < vh;
            vtmp[h] = own ? cnt[h] : 0;
        }
        const int nin = sweep_scan<S::NH>(tid, vtmp, L.ostart, 0);
        const int tot = sweep_scan<S::NH>(tid, cnt, L.start, 0);
        if (tid == 0) {
            L.start[S::NH] = tot;
            L.ninner = nin;
            L.total = tot;
        }
    }
    __syncthreads();
    const int total = L.total, ninner = L.ninner;
    bool ov = false;
    if (total <= S::STAGE) {
        const int cap = C.cap;
        for (int h = tid; h < S::NH; h += S::THREADS) {
            const int st0 = L.start[h], c = cnt[h];
            const int lx = h % S::HW - 1, ly = h / S::HW - 1;
            const bool own = lx >= 0 && lx < vw && ly >= 0 && ly < vh;
            for (int sl = 0; sl < c; ++sl) L.hcell[st0 + sl] = (uint8_t)h;
            if (own)
                for (int sl = 0; sl < c; ++sl) L.own[L.ostart[h] + sl] = (int16_t)(st0 + sl);
        }
        __syncthreads();
        /* two staged disks per thread and pass: both slot reads, then both
         * record reads, are in flight together; staged place st0 + i holds
         * list slot cnt - 1 - i */
        for (int u0 = tid; u0 < total; u0 += 2 * S::THREADS) {
            const int u1 = u0 + S::THREADS;
            const int h0 = L.hcell[u0];
            const int q0 = A.cslot[(size_t)L.gcell[h0] * cap + (L.start[h0 + 1] - 1 - u0)];
            int q1 = -1;
            if (u1 < total) {
                const int h1 = L.hcell[u1];
                q1 = A.cslot[(size_t)L.gcell[h1] * cap + (L.start[h1 + 1] - 1 - u1)];
            }
            const EdmdPart r0 = A.p[q0];
            EdmdPart r1;
            if (q1 >= 0) r1 = A.p[q1];
            const double lat0 = t - r0.tl;
            L.pos[u0] = make_double2(r0.x + lat0 * r0.vx, r0.y + lat0 * r0.vy);
            L.vel[u0] = make_double2(r0.vx, r0.vy);
            if (POLY) L.rad[u0] = r0.rad;
            L.id[u0] = q0;
            L.coll[u0] = r0.coll;
            if (q1 >= 0) {
                const double lat1 = t - r1.tl;
                L.pos[u1] = make_double2(r1.x + lat1 * r1.vx, r1.y + lat1 * r1.vy);
                L.vel[u1] = make_double2(r1.vx, r1.vy);
                if (POLY) L.rad[u1] = r1.rad;
                L.id[u1] = q1;
                L.coll[u1] = r1.coll;
            }
        }
        __syncthreads();
        /* a tile whose halo stays off the seam, in a box at least five
         * cells wide: every candidate lies within two cells of its owner,
         * so the minimum image (EDMD.c:5896-5913) changes nothing and is
         * skipped (uniform over the workgroup).  The last cell column and
         * row are seam cells: a leftward (downward) crossing out of cell 0
         * can leave x (y) at +0 there (PBCpostX wraps only x < 0,
         * EDMD.c:5948-5953), and a disk with t_local = t keeps it when
         * synchronised, so a halo holding them needs the minimum image.
         * Column and row 0 are only ever halo of tiles that wrap. */
        const bool nowrap = tx0 >= 1 && ty0 >= 1 && tx0 + vw <= ncx - 2 && ty0 + vh <= ncy - 2 &&
                            ncx >= 5 && ncy >= 5;
        /* equal radii: the contact term 4 r r of every pair is the owner's
         * own (bitwise the same product), no radius read per candidate;
         * without staged radii it is disk 0's (every radius is bitwise the
         * same) */
        const double rad0 = POLY ? 0.0 : A.p[0].rad;
        if constexpr (!POLY) {
            if (nowrap)
                sweep_tile_predict<S, false, X, true>(C, A, L, tid, ninner, t, tx0, ty0, paul, ctr0,
                                                      rad0, ov);
            else
                sweep_tile_predict<S, true, X, true>(C, A, L, tid, ninner, t, tx0, ty0, paul, ctr0,
                                                     rad0, ov);
        } else {
            if (nowrap && C.mono)
                sweep_tile_predict<S, false, X, true>(C, A, L, tid, ninner, t, tx0, ty0, paul, ctr0,
                                                      rad0, ov);
            else if (nowrap)
                sweep_tile_predict<S, false, X, false>(C, A, L, tid, ninner, t, tx0, ty0, paul,
                                                       ctr0, rad0, ov);
            else
                sweep_tile_predict<S, true, X, false>(C, A, L, tid, ninner, t, tx0, ty0, paul, ctr0,
                                                      rad0, ov);
        }
    } else {
        const int cap = C.cap;
        for (int w = tid; w < S::TX * S::TY * cap; w += S::THREADS) {
            const int c = w / cap, sl = w - c * cap;
            const int lx = c % S::TX, ly = c / S::TX;
            if (lx >= vw || ly >= vh) continue;
            const int g = (ty0 + ly) * ncx + tx0 + lx;
            if (sl < A.ccnt[g]) {
                const int q = A.cslot[(size_t)g * cap + sl];
                ov |= sweep_one<false, X>(C, A, t, q, paul, ctr0 + 2 * (int64_t)q,
                                       ctr0 + 2 * (int64_t)q + 1);
            }
        }
    }
    if (ov) atomicExch(&A.st->status, EDMD_DEV_OVERLAP);
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
    EdmdPart* r = S->a.p + p;
    double x = r->x, y = r->y, tl = r->tl;
    if (S->a.st->growing) {
        double rad = r->rad;
        fly_grow(S->c, t, x, y, r->vx, r->vy, tl, rad, S->a.vr[p]);
        r->rad = rad;
    } else {
        double vx = r->vx, vy = r->vy;
        fly_v<true>(S->c, t, x, y, vx, vy, tl);
        r->vx = vx;
        r->vy = vy;
    }
    r->x = x;
    r->y = y;
    r->tl = tl;
}

/* Prediction sweep of the systems that are growing (collisionEventGrow,
 * EDMD.c:3104-3230, crossingEventGrow, EDMD.c:2343-2400 = the Normal
 * crossing), one lane per particle. */
__global__ __launch_bounds__(256) void k_sweep_grow(const EdmdSys* __restrict__ systems) {
    const EdmdSys* S = systems + blockIdx.y;
    if (!S->a.st->growing) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S->c.n) return;
    const int64_t ctr0 = S->a.st->ctr;
    bool ov = sweep_one<true>(S->c, S->a, S->a.st->t, p, S->a.st->paul, ctr0 + 2 * (int64_t)p,
                              ctr0 + 2 * (int64_t)p + 1);
    if (ov) atomicExch(&S->a.st->status, EDMD_DEV_OVERLAP);
}

/* After a full sweep: one insertion per slot of every particle counted. */
__global__ void k_sweep_count(const EdmdSys* __restrict__ systems, int nsys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) systems[s].a.st->ctr += 2 * (int64_t)systems[s].c.n;
}

/* ------------------------------------------------------------ event loop */

/* Loop state kept in registers; the rest of EdmdState stays in memory and
 * is touched only by the events that use it. */
/* The emulated state of the reference's queue, kept in LDS (one wavefront
 * per workgroup): it changes rarely and would otherwise hold scalar
 * registers through the whole loop. */
struct QueueState {
    int32_t cand[EDMD_STAGE_MAX]; /* candidate stage of the neighbour scan */
    /* the top level of the calendar (at most 64 nodes): read by every
     * event, written only here while the loop runs and copied to HBM when
     * it ends */
    EdmdNode top[64];
    /* collision virial sums since the last thermo row (collTermX, collTermY,
     * collTermXY, collTermYX), accumulated by lane 0 in the reference's
     * order; here rather than in scalar registers, which the loop has too
     * few of */
    double vir[4];
    /* the calendar's minimum after a wrap moved particle events of the
     * overflow list (moved != 0) */
    double mkey;
    int32_t mleaf, maux, moved;
};

/* Top level of the calendar between HBM and the LDS copy. */
template <class AR>
__device__ __forceinline__ void top_load(const EdmdConst& C, const AR& A, int lane,
                                         EdmdNode* top) {
    if (lane < C.lsize[C.nlev]) top[lane] = level(A, C.nlev)[lane];
    WAVE_SYNC();
}

template <class AR>
__device__ __forceinline__ void top_store(const EdmdConst& C, const AR& A, int lane,
                                          const EdmdNode* top) {
    if (lane < C.lsize[C.nlev]) level(A, C.nlev)[lane] = top[lane];
    WAVE_SYNC();
}

struct LoopRegs {
    double t;
    double next_special; /* earliest of the screenshot / thermo / noise / growth-stop times */
    int special;         /* its kind */
    uint64_t ncol, nevents;
    int32_t status, err_i, err_j;
    /* the queue's position (uniform, kept in scalar registers) */
    double paul;   /* the reference's paulTime */
    int64_t ctr;   /* insertions so far */
    int64_t win;   /* buckets passed */
    int64_t srank; /* rank of the next special event */
    int64_t swin;  /* bucket of the next special event (-1: overflow list) */
    int cur;       /* win modulo the bucket count */
    QueueState* q; /* LDS: candidate stage and the calendar's top level */
};

__device__ __forceinline__ void next_special(const EdmdState* st, LoopRegs& R) {
    double tsp = st->next_screen;
    int64_t rsp = st->rank_screen, wsp = st->win_screen;
    int sp = EDMD_EV_SCREENSHOT;
    if (before(st->next_thermo, st->rank_thermo, tsp, rsp)) {
        tsp = st->next_thermo; rsp = st->rank_thermo; wsp = st->win_thermo; sp = EDMD_EV_THERMO;
    }
    if (before(st->next_noise, st->rank_noise, tsp, rsp)) {
        tsp = st->next_noise; rsp = st->rank_noise; wsp = st->win_noise; sp = EDMD_EV_ADDINGNOISE;
    }
    if (before(st->next_grow, st->rank_grow, tsp, rsp)) {
        tsp = st->next_grow; rsp = st->rank_grow; wsp = st->win_grow; sp = EDMD_EV_GROWSTOP;
    }
    R.next_special = uni(tsp);
    R.srank = uni64(rsp);
    R.swin = uni64(wsp);
    R.special = uni(sp);
}

/* Bucket the reference files an event at t_ev into (addEventToQueue,
 * EDMD.c:2144-2170): the current one when t_ev - paulTime < dtPaul, -1 for
 * the overflow list, else win + (int)((t_ev - paulTime) / dtPaul). */
__device__ __host__ __forceinline__ int64_t window_of(double t_ev, double paul, double dtp,
                                                      int nlists, int64_t win) {
    const double dt = t_ev - paul;
    if (dt < dtp) return win;
    if (dt >= dtp * nlists) return -1;
    return win + (int64_t)(int)(dt / dtp);
}

/* The overflow list's special events filed again when the ring of buckets
 * wraps (EDMD.c:2292-2302): they get their bucket and rank anew. */
__device__ __forceinline__ void refile_specials(const EdmdConst& C, EdmdState* st, double paul,
                                             int64_t win, int64_t& ctr) {
    double* ts[4] = {&st->next_screen, &st->next_thermo, &st->next_noise, &st->next_grow};
    int64_t* ws[4] = {&st->win_screen, &st->win_thermo, &st->win_noise, &st->win_grow};
    int64_t* rs[4] = {&st->rank_screen, &st->rank_thermo, &st->rank_noise, &st->rank_grow};
    for (int k = 0; k < 4; ++k) {
        const double tk = *ts[k];
        if (*ws[k] != -1 || !(tk < 1e25)) continue;
        const int64_t w = window_of(tk, paul, C.dtp, C.nlists, win);
        /* filed straight into the tree during the wrap, before the new
         * bucket is drained: ahead of both drained and later direct events */
        const int64_t r = (w == win) ? ctr : rank_of(tk, paul, C.dtp, C.horizon, ctr);
        ctr++;
        WAVE_SYNC();
        if ((threadIdx.x & 63) == 0) {
            *ws[k] = w;
            *rs[k] = r;
        }
        WAVE_SYNC();
    }
}

/* Rank of a particle event of the overflow list filed again at a wrap with
 * counter c (addEventToQueue from addNextPaulEvent, EDMD.c:2144-2170,
 * 2292-2302): straight into the tree, which is empty during a wrap, so
 * ahead of the drained bucket and of every later insertion (rank c, as a
 * special filed there); the overflow list again; or a bucket. */
__device__ __forceinline__ int64_t refiled_rank(const EdmdConst& C, double t_ev, double paul,
                                                int64_t c) {
    const double dt = t_ev - paul;
    if (dt < C.dtp) return c;
    if (dt >= C.horizon) return EDMD_RANK_OVF + c;
    return EDMD_RANK_DIRECT - 1 - c;
}

/* The particle events of the overflow list filed again at a wrap
 * (EDMD.c:2292-2302), after its special events, in the order of the list's
 * walk from its head (the newest first): a slot ranked EDMD_RANK_OVF + c
 * gets the counter 2 now - 1 - c and the rank of where it goes now; the
 * counter moves past them all (the list holds counters from the previous
 * wrap's on).  The leaves of the disks concerned are made again and the
 * calendar is rebuilt.  Once per wrap of the bucket ring (5 time units);
 * run from the loop's body, like the thermostat, not from the bucket walk,
 * which stays small. */
template <class AR>
__device__ void rebuild_tree(const EdmdConst& C, const AR& A, int lane);

/* The overflow list filed again at a wrap (EDMD.c:2292-2302): its special
 * events, then its particle events in the order of the list's walk from
 * its head (the newest first): a slot ranked EDMD_RANK_OVF + c gets the
 * counter 2 now - 1 - c and the rank of where it goes now, and the counter
 * moves past them all (the list holds counters from the previous wrap's
 * on).  When any slot moved, the leaves of the disks concerned are made
 * again, the calendar is rebuilt, its top level reloaded and its minimum
 * left in q (q->moved = 1).  Once per wrap of the bucket ring (5 time
 * units); kept out of line, with the loop's registers out of its way. */
template <class AR>
__device__ __noinline__ void refile_wrap(const EdmdConst& C, const AR& A, QueueState* q,
                                         double paul, int64_t win, double t, int64_t& ctr) {
    const int lane = threadIdx.x & 63;
    refile_specials(C, A.st, paul, win, ctr);
    const int64_t now = ctr;
    bool any = false;
    for (int k = lane; k < C.n; k += 64) {
        int64_t rk = A.erank[2 * k], rx = A.erank[2 * k + 1];
        const bool fk = EDMD_RANK_IS_OVF(rk), fx = EDMD_RANK_IS_OVF(rx);
        if (!fk && !fx) continue;
        const EdmdEv e = A.e[k];
        if (fk) rk = refiled_rank(C, e.tcoll, paul, 2 * now - 1 - (rk - EDMD_RANK_OVF));
        if (fx) rx = refiled_rank(C, e.tcross, paul, 2 * now - 1 - (rx - EDMD_RANK_OVF));
        A.erank[2 * k] = rk;
        A.erank[2 * k + 1] = rx;
        const Leaf Lf = make_leaf(e.tcoll, rk, e.tcross, rx, e.partner, e.dir);
        EdmdNode nd;
        nd.key = Lf.key;
        nd.idx = k;
        nd.aux = Lf.aux;
        A.leaf[k] = nd;
        A.lrank[k] = Lf.rank;
        any = true;
    }
    const bool moved = __ballot(any) != 0ull;
    const int64_t lo = A.st->ovf_lo;
    WAVE_SYNC();
    if (lane == 0) A.st->ovf_lo = now;
    ctr = 2 * now - lo;
    if (moved) {
        WAVE_SYNC();
        rebuild_tree(C, A, lane);
        top_load(C, A, lane, q->top);
        double key;
        int leaf, aux;
        top_min(C, A, lane, q->top, t, key, leaf, aux);
        if (lane == 0) {
            q->mkey = key;
            q->mleaf = leaf;
            q->maux = aux;
        }
    }
    if (lane == 0) q->moved = moved ? 1 : 0;
    WAVE_SYNC();
}

/* One bucket further (addNextPaulEvent, EDMD.c:2287-2307); true when the
 * ring wrapped (the overflow list filed again, refile_wrap). */
template <class AR>
__device__ __forceinline__ bool step_window(const EdmdConst& C, const AR& A, LoopRegs& R) {
    R.paul = uni(R.paul + C.dtp);
    R.win++;
    if (++R.cur < C.nlists) return false;
    R.cur = 0;
    int64_t ctr = R.ctr;
    refile_wrap(C, A, R.q, R.paul, R.win, R.t, ctr);
    R.ctr = uni64(ctr);
    return true;
}

/* findNextEvent (EDMD.c:2313-2327): while no pending event is in the
 * reference's tree -- a particle event within dtPaul of paulTime, or a
 * special event whose bucket has come up -- move on one bucket (refiling
 * the overflow list when the ring wraps).  Then the earlier in the tree by
 * (time, rank) is next.  Returns true for the special event, false for
 * the particle event (key, leaf, aux: the calendar's minimum, taken anew
 * when a wrap moved particle events of the overflow list).  A queue whose next event lies
 * more than 2^24 buckets ahead is not walked bucket by bucket (the
 * reference would); paulTime then jumps to that event. */
template <class AR>
__device__ __forceinline__ bool pick_next(const EdmdConst& C, const AR& A, EdmdState* st,
                                          LoopRegs& R, double& key, int& leaf, int& aux) {
    /* the particle event's rank matters only against a special at its time */
    auto krank = [&]() { return key == R.next_special ? uni64(A.lrank[leaf]) : (int64_t)0; };
    for (int guard = 0;; ++guard) {
        const bool part_in = key - R.paul < C.dtp;
        const int64_t ws = R.swin;
        const bool spec_in = ws >= 0 && ws <= R.win;
        if (part_in || spec_in)
            return spec_in && (!part_in || before(R.next_special, R.srank, key, krank()));
        if (!(key < 1e25) && !(R.next_special < 1e25)) return R.next_special <= key;
        if (guard >= (1 << 24)) {
            const double tn = key < R.next_special ? key : R.next_special;
            while (!(tn - R.paul < C.dtp)) {
                R.paul = uni(R.paul + C.dtp);
                R.win++;
            }
            return before(R.next_special, R.srank, key, krank());
        }
        if (step_window(C, A, R)) {
            next_special(st, R);
            if (R.q->moved) {
                key = R.q->mkey;
                leaf = R.q->mleaf;
                aux = R.q->maux;
            }
        }
    }
}

/* Rank of the next insertion; counts it. */
__device__ __forceinline__ int64_t take_rank(const EdmdConst& C, LoopRegs& R, double t_ev) {
    const int64_t r = rank_of(t_ev, R.paul, C.dtp, C.horizon, R.ctr);
    R.ctr++;
    return r;
}

template <class AR>
__device__ __forceinline__ void trace_push(const EdmdConst& C, const AR& A, int lane,
                                           double t, int type, int i, int j, int valid) {
    if (C.trace_cap <= 0) return;
    if (lane == 0) {
        const int64_t k = A.st->ntrace;
        if (k < C.trace_cap) {
            EdmdTraceRec rec;
            rec.type = type;
            rec.i = i;
            rec.j = j;
            rec.valid = valid;
            rec.t = t;
            A.trace[k] = rec;
        }
        A.st->ntrace = k + 1;
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

template <bool G, bool W, class AR>
__device__ __forceinline__ void ev_screenshot(const EdmdConst& C, const AR& A, int lane,
                                              LoopRegs& R) {
    /* takeAScreenshot, EDMD.c:4657-4677: free-fly every particle */
    const double t = R.t;
    for (int q = lane; q < C.n; q += 64) {
        const auto r = A.p + q;
        double x = r->x, y = r->y, tl = r->tl;
        if (G) {
            double rad = r->rad;
            fly_grow(C, t, x, y, r->vx, r->vy, tl, rad, A.vr[q]);
            r->rad = rad;
        } else {
            double vx = r->vx, vy = r->vy;
            fly_v<W>(C, t, x, y, vx, vy, tl);
            if (W) {
                r->vx = vx;
                r->vy = vy;
            }
        }
        r->x = x;
        r->y = y;
        r->tl = tl;
    }
    WAVE_SYNC();
    const int64_t wk = window_of(t + C.dtime, R.paul, C.dtp, C.nlists, R.win);
    const int64_t rk = take_rank(C, R, t + C.dtime);
    if (lane == 0) {
        A.st->next_screen = t + C.dtime;
        A.st->rank_screen = rk;
        A.st->win_screen = wk;
    }
    WAVE_SYNC();
}

template <class AR>
__device__ __forceinline__ void ev_thermo(const EdmdConst& C, const AR& A, int lane, LoopRegs& R) {
    /* takeAThermo + saveThermo, EDMD.c:4679-4713, 5232-5333 */
    const double t = R.t;
    const uint64_t ncol = R.ncol;
    const int64_t wk = window_of(t + C.dtime_thermo, R.paul, C.dtp, C.nlists, R.win);
    const int64_t rk = take_rank(C, R, t + C.dtime_thermo);
    const auto P = A.p;
    double E = ordered_sum(C.n, lane, [&](int q) {
        double a = P[q].vx;
        return 0.5 * P[q].m * (a * a + P[q].vy * P[q].vy);
    });
    double ex = ordered_sum(C.n, lane, [&](int q) { return P[q].m * P[q].vx * P[q].vx; });
    double ey = ordered_sum(C.n, lane, [&](int q) { return P[q].m * P[q].vy * P[q].vy; });
    double exy = ordered_sum(C.n, lane, [&](int q) { return P[q].m * P[q].vx * P[q].vy; });
    double eyx = ordered_sum(C.n, lane, [&](int q) { return P[q].m * P[q].vy * P[q].vx; });
    /* the mid wall's columns (EDMD.c:5453-5477), from the stored disks in
     * particle order; type-0 disks on the other side add +0 */
    double nl = 0, nr = 0, ns = 0, tl = 0, tr = 0, tsol = 0;
    if (C.walls & 8) {
        const auto T = A.ptype;
        const double h = C.halfLx;
        auto temp = [&](int q) {
            return 0.5 * P[q].m * (P[q].vx * P[q].vx + P[q].vy * P[q].vy);
        };
        nl = ordered_sum(C.n, lane, [&](int q) { return (T[q] == 0 && P[q].x < h) ? 1.0 : 0.0; });
        nr = ordered_sum(C.n, lane, [&](int q) { return (T[q] == 0 && P[q].x > h) ? 1.0 : 0.0; });
        ns = ordered_sum(C.n, lane, [&](int q) { return T[q] != 0 ? 1.0 : 0.0; });
        tl = ordered_sum(C.n, lane, [&](int q) { return (T[q] == 0 && P[q].x < h) ? temp(q) : 0.0; });
        tr = ordered_sum(C.n, lane, [&](int q) { return (T[q] == 0 && P[q].x > h) ? temp(q) : 0.0; });
        tsol = ordered_sum(C.n, lane, [&](int q) { return T[q] != 0 ? temp(q) : 0.0; });
    }
    EdmdState* st = A.st;
    /* EDMD.c:5253-5256: the disc's area inside a circular wall */
    const double area = (C.walls & 4) ? 3.14159265358979323846 * C.halfLx * C.halfLx : C.Lx * C.Ly;
    const double dT = t - st->last_thermo;
    const double v0 = R.q->vir[0], v1 = R.q->vir[1], v2 = R.q->vir[2], v3 = R.q->vir[3];
    const int k = st->nthermo;
    if (lane == 0) {
        if (k < C.thermo_cap) {
            EdmdThermoRow row;
            row.t = t;
            row.ncol = ncol;
            row.e_per_n = E / C.n;
            row.pxx = -1 * v0 / (area * dT) + ex / area;
            row.pyy = -1 * v1 / (area * dT) + ey / area;
            row.pxy = -1 * v2 / (area * dT) + exy / area;
            row.pyx = -1 * v3 / (area * dT) + eyx / area;
            row.pressure = (-1 / (dT)) * (v0 + v1) / (2 * area) + E / area;
            row.pwall = 0;
            row.pmid = row.dp_lr = row.pack_left = row.pack_right = 0;
            row.temp_left = row.temp_right = row.temp_solute = 0;
            if (C.walls & 8) {
                const double ratio = 3.14159265358979323846 / (C.halfLx * C.Ly);
                const double sr = C.sizeratio;
                row.pmid = st->dpm / (dT * C.Ly);
                row.dp_lr = -(st->dpr - st->dpl) / (dT * C.Ly);
                row.pack_left = ratio * (double)(int)nl * sr * sr;
                row.pack_right = ratio * (double)(int)nr * sr * sr;
                row.temp_left = tl / (int)nl;
                row.temp_right = tr / (int)nr;
                row.temp_solute = tsol / (int)ns;
            }
            if (C.walls & 7) {
                /* EDMD.c:5428-5451: dp / (deltaTime * perimeter) */
                double perimeter = 0;
                if (C.walls & 4) perimeter = 3.14159265358979323846 * C.Lx;
                if (C.walls & 1) perimeter += 2 * C.Ly;
                if (C.walls & 2) perimeter += 2 * C.Lx;
                row.pwall = st->dpw / (dT * perimeter);
            }
            A.thermo[k] = row;
        }
        if (C.walls & 7) st->dpw = 0;
        if (C.walls & 8) st->dpl = st->dpm = st->dpr = 0;
        st->nthermo = k + 1;
        st->last_thermo = t;
        st->next_thermo = t + C.dtime_thermo;
        st->rank_thermo = rk;
        st->win_thermo = wk;
    }
    if (lane == 0) R.q->vir[0] = R.q->vir[1] = R.q->vir[2] = R.q->vir[3] = 0;
    WAVE_SYNC();
}

/* Rebuilds every level of the calendar after a full re-prediction. */
template <class AR>
__device__ void rebuild_tree(const EdmdConst& C, const AR& A, int lane) {
    for (int l = 1; l <= C.nlev; ++l) {
        for (int g = lane; g < C.lsize[l]; g += 64) node_min(C, A, l, g);
        WAVE_SYNC();
    }
}

template <bool G, bool W, class AR>
__device__ __forceinline__ bool ev_rescale(const EdmdConst& C, const AR& A, int lane,
                                           LoopRegs& R) {
    const double t = R.t;
    /* addNoise, velocity-rescale branch (noise == 2): EDMD.c:4828-4832,
     * 4880-4915, 4922 */
    const auto P = A.p;
    double E = ordered_sum(C.n, lane, [&](int q) {
        double a = P[q].vx;
        return 0.5 * P[q].m * (a * a + P[q].vy * P[q].vy);
    });
    for (int q = lane; q < C.n; q += 64) {
        EdmdPart r = A.p[q];
        r.coll = r.coll + 1;
        if (G) fly_grow(C, t, r.x, r.y, r.vx, r.vy, r.tl, r.rad, A.vr[q]);
        else fly_v<W>(C, t, r.x, r.y, r.vx, r.vy, r.tl);
        r.vx /= sqrt(E / C.n / C.temperature);
        r.vy /= sqrt(E / C.n / C.temperature);
        A.p[q] = r;
    }
    WAVE_SYNC();
    bool ov = false;
    /* per particle: crossing, then collision (EDMD.c:4909-4915) */
    for (int q = lane; q < C.n; q += 64)
        ov |= sweep_one<G, W>(C, A, t, q, R.paul, R.ctr + 2 * (int64_t)q, R.ctr + 2 * (int64_t)q + 1);
    R.ctr += 2 * (int64_t)C.n;
    const bool any_ov = __ballot(ov) != 0ull;
    WAVE_SYNC();
    rebuild_tree(C, A, lane);
    top_load(C, A, lane, R.q->top);
    const int64_t wk = window_of(t + C.dtnoise, R.paul, C.dtp, C.nlists, R.win);
    const int64_t rk = take_rank(C, R, t + C.dtnoise);
    if (lane == 0) {
        A.st->next_noise = t + C.dtnoise;
        A.st->rank_noise = rk;
        A.st->win_noise = wk;
    }
    WAVE_SYNC();
    return any_ov;
}

__device__ __forceinline__ void flag_overlap(LoopRegs& R, bool ov, int p) {
    if (__ballot(ov) != 0ull && R.status == EDMD_DEV_OK) {
        R.status = EDMD_DEV_OVERLAP;
        R.err_i = p;
    }
}

template <int ML, bool W, class AR>
__device__ __forceinline__ void ev_collision(const EdmdConst& C, const AR& A, int lane, LoopRegs& R, int i,
                             int j) {
    /* doTheCollisionNormal, EDMD.c:3542-3562, 3583-3606, 3802-3827, 3869-3879 */
    const EdmdEv ei = A.e[i];
    const int64_t rxi_old = A.erank[2 * i + 1]; /* crossing slot kept by a stale pop */
    const EdmdPart pi = A.p[i];
    const EdmdPart pj = A.p[j];
    PathCache<ML> P;
    path_prefetch<ML>(C, A, lane, i, j, P);
    const double t = R.t;
    double xi = uni(pi.x), yi = uni(pi.y), vxi = uni(pi.vx), vyi = uni(pi.vy), tli = uni(pi.tl);
    const double ri = uni(pi.rad), mi = uni(pi.m);
    const int cxi = uni((int)pi.cx), cyi = uni((int)pi.cy);
    const uint32_t colli = (uint32_t)uni((int)pi.coll);
    const uint32_t collj = (uint32_t)uni((int)pj.coll);
    const uint32_t ca = (uint32_t)uni((int)ei.cact);

    fly_v<W>(C, t, xi, yi, vxi, vyi, tli);
    xi = uni(xi);
    yi = uni(yi);
    vxi = uni(vxi);
    vyi = uni(vyi);
    const bool valid = (ca == collj);
    trace_push(C, A, lane, t, EDMD_EV_COLLISION, i, j, valid ? 1 : 0);
    if (!valid) {
        /* stale prediction: free flight of i and a new collision slot */
        if (lane == 0) {
            A.p[i].x = xi;
            A.p[i].y = yi;
            A.p[i].vx = vxi;
            A.p[i].vy = vyi;
            A.p[i].tl = t;
        }
        WAVE_SYNC();
        const Owner oi{i, xi, yi, vxi, vyi, ri, cxi, cyi};
        bool ov = false;
        double dt, dtb;
        int pa, pb;
        uint32_t cpa, cpb;
        predict_owners<false, W, (ML > 2 && !W)>(C, A, lane, t, 1, oi, oi, dt, pa, cpa, dtb, pb, cpb,
                                                 ov, R.q->cand);
        flag_overlap(R, ov, i);
        const double tk = t + dt;
        const double tci = uni(ei.tcross);
        const int64_t rci = uni64(rxi_old);
        const int diri = uni(ei.dir);
        const int64_t rk = take_rank(C, R, tk);
        if (lane == 0) {
            A.e[i].tcoll = tk;
            A.e[i].partner = pa;
            A.e[i].cact = cpa;
            A.erank[2 * i] = rk;
        }
        path_update<ML>(C, A, lane, P, i, make_leaf(tk, rk, tci, rci, pa, diri), -1, Leaf{}, R.q->top, R.t);
        return;
    }
    R.ncol++;
    double xj = uni(pj.x), yj = uni(pj.y), vxj = uni(pj.vx), vyj = uni(pj.vy), tlj = uni(pj.tl);
    const double rj = uni(pj.rad), mj = uni(pj.m);
    const int cxj = uni((int)pj.cx), cyj = uni((int)pj.cy);
    fly_v<W>(C, t, xj, yj, vxj, vyj, tlj);
    vxj = uni(vxj);
    vyj = uni(vyj);
    double dx = xj - xi;
    double dy = yj - yi;
    double dvx = vxj - vxi;
    double dvy = vyj - vyi;
    double inv_mass = 1 / (mi + mj);
    min_image(C, dx, dy);
    double kernel = inv_mass * (1 + C.res) * (dx * dvx + dy * dvy);
    double ff = kernel / (4 * ri * rj);
    double w = mi * mj * ff;
    vxi = uni(vxi + ff * mj * dx);
    vyi = uni(vyi + ff * mj * dy);
    vxj = uni(vxj - ff * mi * dx);
    vyj = uni(vyj - ff * mi * dy);
    xj = uni(xj);
    yj = uni(yj);
    if (lane == 0) {
        double* vir = R.q->vir;
        vir[0] += w * dx * dx;
        vir[1] += w * dy * dy;
        vir[2] += w * dx * dy;
        vir[3] += w * dy * dx;
        EdmdState* st = A.st;
        auto q = A.p + i;
        q->x = xi; q->y = yi; q->vx = vxi; q->vy = vyi; q->tl = t; q->coll = colli + 1;
        q = A.p + j;
        q->x = xj; q->y = yj; q->vx = vxj; q->vy = vyj; q->tl = t; q->coll = collj + 1;
        if (C.tape_cap > 0) {
            const int64_t k = st->ntape;
            if (k < C.tape_cap) {
                EdmdTapeRec rec;
                rec.t = t; rec.ax = dx; rec.ay = dy; rec.cx = dvx; rec.cy = dvy; rec.ff = ff;
                rec.i = i; rec.j = j; rec.g = 0;
                A.tape[k] = rec;
            }
            st->ntape = k + 1;
            st->wsum += w * (dx * dx + dy * dy);
            st->tlast = t;
        }
    }
    WAVE_SYNC();
    const Owner oi{i, xi, yi, vxi, vyi, ri, cxi, cyi};
    const Owner oj{j, xj, yj, vxj, vyj, rj, cxj, cyj};
    bool ov = false;
    double dti, dtj;
    int pi_, pj_;
    uint32_t cpi, cpj;
    predict_owners<false, W, (ML > 2 && !W)>(C, A, lane, t, 2, oi, oj, dti, pi_, cpi, dtj, pj_, cpj,
                                             ov, R.q->cand);
    flag_overlap(R, ov, i);
    double tci2, tcj2;
    int di, dj;
    cross_time<W>(C, t, xi, yi, vxi, vyi, cxi, cyi, tci2, di);
    cross_time<W>(C, t, xj, yj, vxj, vyj, cxj, cyj, tcj2, dj);
    const double tki = t + dti, tkj = t + dtj;
    /* insertion order, EDMD.c:3869-3879: crossings of i and j, then the
     * collisions of i and j */
    const int64_t rxi = take_rank(C, R, tci2);
    const int64_t rxj = take_rank(C, R, tcj2);
    const int64_t rki = take_rank(C, R, tki);
    const int64_t rkj = take_rank(C, R, tkj);
    if (lane == 0) {
        EdmdEv e;
        e.tcoll = tki; e.tcross = tci2; e.partner = pi_; e.cact = cpi; e.dir = di; e.pad = 0;
        A.e[i] = e;
        e.tcoll = tkj; e.tcross = tcj2; e.partner = pj_; e.cact = cpj; e.dir = dj;
        A.e[j] = e;
        A.erank[2 * i] = rki;
        A.erank[2 * i + 1] = rxi;
        A.erank[2 * j] = rkj;
        A.erank[2 * j + 1] = rxj;
    }
    path_update<ML>(C, A, lane, P, i, make_leaf(tki, rki, tci2, rxi, pi_, di), j,
                    make_leaf(tkj, rkj, tcj2, rxj, pj_, dj), R.q->top, R.t);
}

template <int ML, class AR>
__device__ __forceinline__ void ev_collision_grow(const EdmdConst& C, const AR& A, int lane, LoopRegs& R,
                                  int i, int j) {
    /* doTheCollisionGrow, EDMD.c:3465-3540 (2D): both disks fly with their
     * radii growing; the unit contact normal and the summed growth speed
     * enter the impulse; res and delta are the compile-time 1 and 0 of
     * EDMD.c:3467-3468; no virial. */
    const double delta = 0.1 * 0;
    const double res = 0.75 * 0 + 1;
    const EdmdEv ei = A.e[i];
    const int64_t rxi_old = A.erank[2 * i + 1]; /* crossing slot kept by a stale pop */
    const EdmdPart pi = A.p[i];
    const EdmdPart pj = A.p[j];
    const double vri = uni(A.vr[i]), vrj = uni(A.vr[j]);
    PathCache<ML> P;
    path_prefetch<ML>(C, A, lane, i, j, P);
    const double t = R.t;
    double xi = uni(pi.x), yi = uni(pi.y), vxi = uni(pi.vx), vyi = uni(pi.vy), tli = uni(pi.tl);
    double ri = uni(pi.rad);
    const double mi = uni(pi.m);
    const int cxi = uni((int)pi.cx), cyi = uni((int)pi.cy);
    const uint32_t colli = (uint32_t)uni((int)pi.coll);
    const uint32_t collj = (uint32_t)uni((int)pj.coll);
    const uint32_t ca = (uint32_t)uni((int)ei.cact);

    fly_grow(C, t, xi, yi, vxi, vyi, tli, ri, vri);
    xi = uni(xi);
    yi = uni(yi);
    ri = uni(ri);
    const bool valid = (ca == collj);
    trace_push(C, A, lane, t, EDMD_EV_COLLISION, i, j, valid ? 1 : 0);
    if (!valid) {
        if (lane == 0) {
            A.p[i].x = xi;
            A.p[i].y = yi;
            A.p[i].tl = t;
            A.p[i].rad = ri;
        }
        WAVE_SYNC();
        const Owner oi{i, xi, yi, vxi, vyi, ri, cxi, cyi, vri};
        bool ov = false;
        double dt, dtb;
        int pa, pb;
        uint32_t cpa, cpb;
        predict_owners<true>(C, A, lane, t, 1, oi, oi, dt, pa, cpa, dtb, pb, cpb, ov, R.q->cand);
        flag_overlap(R, ov, i);
        const double tk = t + dt;
        const double tci = uni(ei.tcross);
        const int64_t rci = uni64(rxi_old);
        const int diri = uni(ei.dir);
        const int64_t rk = take_rank(C, R, tk);
        if (lane == 0) {
            A.e[i].tcoll = tk;
            A.e[i].partner = pa;
            A.e[i].cact = cpa;
            A.erank[2 * i] = rk;
        }
        path_update<ML>(C, A, lane, P, i, make_leaf(tk, rk, tci, rci, pa, diri), -1, Leaf{}, R.q->top, R.t);
        return;
    }
    R.ncol++;
    double xj = uni(pj.x), yj = uni(pj.y), vxj = uni(pj.vx), vyj = uni(pj.vy), tlj = uni(pj.tl);
    double rj = uni(pj.rad);
    const double mj = uni(pj.m);
    const int cxj = uni((int)pj.cx), cyj = uni((int)pj.cy);
    fly_grow(C, t, xj, yj, vxj, vyj, tlj, rj, vrj);
    double dx = xj - xi;
    double dy = yj - yi;
    double dvx = vxj - vxi;
    double dvy = vyj - vyi;
    double dvr = vri + vrj;
    min_image(C, dx, dy);
    double dist = sqrt(dx * dx + dy * dy);
    double dxr = dx / dist;
    double dyr = dy / dist;
    double inv_mass = 1 / (mi + mj);
    double kernel = inv_mass * (1 + res) * (dxr * dvx + dyr * dvy - dvr);
    vxi = uni(vxi + (kernel * mj * dxr - 2 * mj * inv_mass * delta * dxr));
    vyi = uni(vyi + (kernel * mj * dyr - 2 * mj * inv_mass * delta * dyr));
    vxj = uni(vxj - (kernel * mi * dxr - 2 * mi * inv_mass * delta * dxr));
    vyj = uni(vyj - (kernel * mi * dyr - 2 * mi * inv_mass * delta * dyr));
    xj = uni(xj);
    yj = uni(yj);
    rj = uni(rj);
    if (lane == 0) {
        auto q = A.p + i;
        q->x = xi; q->y = yi; q->vx = vxi; q->vy = vyi; q->tl = t; q->rad = ri; q->coll = colli + 1;
        q = A.p + j;
        q->x = xj; q->y = yj; q->vx = vxj; q->vy = vyj; q->tl = t; q->rad = rj; q->coll = collj + 1;
    }
    WAVE_SYNC();
    const Owner oi{i, xi, yi, vxi, vyi, ri, cxi, cyi, vri};
    const Owner oj{j, xj, yj, vxj, vyj, rj, cxj, cyj, vrj};
    bool ov = false;
    double dti, dtj;
    int pi_, pj_;
    uint32_t cpi, cpj;
    predict_owners<true>(C, A, lane, t, 2, oi, oj, dti, pi_, cpi, dtj, pj_, cpj, ov, R.q->cand);
    flag_overlap(R, ov, i);
    double tci2, tcj2;
    int di, dj;
    cross_time<false>(C, t, xi, yi, vxi, vyi, cxi, cyi, tci2, di);
    cross_time<false>(C, t, xj, yj, vxj, vyj, cxj, cyj, tcj2, dj);
    const double tki = t + dti, tkj = t + dtj;
    /* insertion order, EDMD.c:3869-3879: crossings of i and j, then the
     * collisions of i and j */
    const int64_t rxi = take_rank(C, R, tci2);
    const int64_t rxj = take_rank(C, R, tcj2);
    const int64_t rki = take_rank(C, R, tki);
    const int64_t rkj = take_rank(C, R, tkj);
    if (lane == 0) {
        EdmdEv e;
        e.tcoll = tki; e.tcross = tci2; e.partner = pi_; e.cact = cpi; e.dir = di; e.pad = 0;
        A.e[i] = e;
        e.tcoll = tkj; e.tcross = tcj2; e.partner = pj_; e.cact = cpj; e.dir = dj;
        A.e[j] = e;
        A.erank[2 * i] = rki;
        A.erank[2 * i + 1] = rxi;
        A.erank[2 * j] = rkj;
        A.erank[2 * j + 1] = rxj;
    }
    path_update<ML>(C, A, lane, P, i, make_leaf(tki, rki, tci2, rxi, pi_, di), j,
                    make_leaf(tkj, rkj, tcj2, rxj, pj_, dj), R.q->top, R.t);
}

/* stopGrow, EDMD.c:4740-4785 (2D, no wells): final flight with growth, all
 * counters and image counters reset, normalizePhysicalQ (EDMD.c:5723-5764:
 * remove the mean momentum, scale to E/N = Einit), then the Normal
 * prediction of every disk with stopGrow's partner counters, and a new
 * calendar.  The caller continues in the Normal loop. */
template <bool W, class AR>
__device__ __forceinline__ bool ev_stop_grow(const EdmdConst& C, const AR& A, int lane,
                                             LoopRegs& R) {
    const double t = R.t;
    for (int q = lane; q < C.n; q += 64) {
        EdmdPart r = A.p[q];
        fly_grow(C, t, r.x, r.y, r.vx, r.vy, r.tl, r.rad, A.vr[q]);
        r.coll = 0;
        A.p[q] = r;
        A.img[2 * q] = 0;
        A.img[2 * q + 1] = 0;
    }
    WAVE_SYNC();
    const auto P = A.p;
    const int n = C.n;
    if (!(W && (C.walls & 4))) {
        const double px = ordered_sum(n, lane, [&](int q) { return P[q].vx * P[q].m; });
        const double py = ordered_sum(n, lane, [&](int q) { return P[q].vy * P[q].m; });
        for (int q = lane; q < n; q += 64) {
            const double m = P[q].m;
            P[q].vx -= px / (n * m);
            P[q].vy -= py / (n * m);
        }
    } else {
        /* in a circle, the angular momentum about the centre
         * (EDMD.c:5730-5749): dvT accumulates -m (x0 vy - y0 vx), dvM m r */
        const double h = C.halfLx;
        const double dvT = -ordered_sum(n, lane, [&](int q) {
            return P[q].m * (-P[q].vx * (P[q].y - h) + P[q].vy * (P[q].x - h));
        });
        const double dvM = ordered_sum(n, lane, [&](int q) {
            return P[q].m * sqrt((P[q].x - h) * (P[q].x - h) + (P[q].y - h) * (P[q].y - h));
        });
        const double dv = dvT / dvM;
        for (int q = lane; q < n; q += 64) {
            const double x0 = P[q].x - h, y0 = P[q].y - h;
            const double r = sqrt(x0 * x0 + y0 * y0);
            P[q].vx += -dv * y0 / r;
            P[q].vy += dv * x0 / r;
        }
    }
    WAVE_SYNC();
    const double E = ordered_sum(n, lane, [&](int q) {
        double a = P[q].vx;
        return 0.5 * P[q].m * (a * a + P[q].vy * P[q].vy);
    });
    for (int q = lane; q < n; q += 64) {
        P[q].vx /= sqrt(E / n / C.einit);
        P[q].vy /= sqrt(E / n / C.einit);
        P[q].coll = 1;
    }
    WAVE_SYNC();
    bool ov = false;
    /* per particle: collision, then crossing (EDMD.c:4770-4776) */
    for (int q = lane; q < n; q += 64)
        ov |= sweep_one<false, W>(C, A, t, q, R.paul, R.ctr + 2 * (int64_t)q + 1,
                               R.ctr + 2 * (int64_t)q, true);
    R.ctr += 2 * (int64_t)n;
    WAVE_SYNC();
    const bool any_ov = __ballot(ov) != 0ull;
    rebuild_tree(C, A, lane);
    top_load(C, A, lane, R.q->top);
    if (lane == 0) {
        A.st->growing = 0;
        A.st->next_grow = INF_D;
    }
    WAVE_SYNC();
    return any_ov;
}

template <int ML, bool G, bool W, class AR>
__device__ __forceinline__ void ev_single(const EdmdConst& C, const AR& A, int lane, LoopRegs& R, int p,
                          int d) {
    /* The events that change one disk.  d = 1-4: a cell crossing in that
     * direction (doTheCrossing, EDMD.c:2501-2590; its freeFly is the
     * dispatch pointer, freeFlyGrow while growing).  d = 5 + xy: a wall hit
     * (doTheWallNormal, EDMD.c:3384-3460, without thermal or shear walls;
     * doTheWallGrow while growing): counted as a collision, the momentum
     * given to the wall accumulated for the thermo row, the normal velocity
     * reversed times resW and the tangential one scaled by resT, a flat
     * wall's hit written to the tape when taping.  Both then predict the
     * crossing slot and the collision slot again, in that order. */
    const bool wall = W && d >= 5;
    const EdmdPart pp = A.p[p];
    const double gvr = G ? uni(A.vr[p]) : 0.0;
    PathCache<ML> P;
    path_prefetch<ML>(C, A, lane, p, -1, P);
    const double t = R.t;
    double x = uni(pp.x), y = uni(pp.y), tl = uni(pp.tl);
    double vx = uni(pp.vx), vy = uni(pp.vy);
    double r = uni(pp.rad);
    int cx = uni((int)pp.cx), cy = uni((int)pp.cy);
    trace_push(C, A, lane, t, wall ? EDMD_EV_WALL : EDMD_EV_CELLCROSS, p, wall ? d - 5 : d, 1);
    if (G) {
        fly_grow(C, t, x, y, vx, vy, tl, r, gvr);
        r = uni(r);
    } else {
        fly_v<W>(C, t, x, y, vx, vy, tl);
    }
    x = uni(x);
    y = uni(y);
    vx = uni(vx);
    vy = uni(vy);

    if (wall && G) {
        /* doTheWallGrow, EDMD.c:3333-3382: a collision, no momentum to the
         * wall; the normal velocity turned away with 2 vr added (for y,
         * which way by the disk's height) */
        R.ncol++;
        if (d == 5) {
            const double sg = vx < 0 ? -1.0 : (vx > 0 ? 1.0 : 0.0);
            vx = -sg * (fabs(vx) + 2 * gvr);
        } else if (d == 6) {
            if (y > C.Ly - 3) vy = -vy - 2 * gvr;
            else vy = -vy + 2 * gvr;
        } else {
            /* the circle: reflected about the normal, 2 vr added along it */
            double nx, ny;
            edmd_sincos(edmd_atan2(y - C.halfLx, x - C.halfLx), ny, nx);
            const double mix = -(1 + C.res_wall) * (vx * nx + vy * ny);
            vx += (mix - 2 * gvr) * nx;
            vy += (mix - 2 * gvr) * ny;
        }
        vx = uni(vx);
        vy = uni(vy);
        if (lane == 0) {
            const auto q = A.p + p;
            q->x = x;
            q->y = y;
            q->tl = tl;
            q->vx = vx;
            q->vy = vy;
            q->rad = r;
            q->coll = pp.coll + 1;
        }
        WAVE_SYNC();
    } else if (wall) {
        R.ncol++;
        const double m = uni(pp.m);
        double dp;
        /* the tangential product is made +0.0 where it is -0.0 (resT = 0
         * times a negative velocity): the reference's crossing prediction
         * divides by a -0.0 velocity and flies the disk to t = -inf */
        double dpl = 0, dpm = 0, dpr = 0;
        const double vx0 = vx, vy0 = vy; /* before the hit, for the tape */
        if (d == 5) {
            dp = fabs((C.res_wall + 1) * m * vx);
            if (C.walls & 8) {
                /* EDMD.c:3393-3409: the side walls within 3 of the box's edges */
                if (vx < 0) {
                    if (x - 3 < 0) dpl = dp;
                    else dpm = dp;
                } else {
                    if (x + 3 > C.Lx) dpr = dp;
                    else dpm = dp;
                }
            }
            if (C.thermal_wall) {
                /* thermoWall (EDMD.c:3411-3413): the normal speed drawn */
                vx = -ref_sign(vx) * mt_drand0(A, lane, C.temperature / m);
            } else {
                vx = -C.res_wall * vx;
                vy = C.res_tan * vy + 0.0;
            }
        } else if (d == 6) {
            dp = fabs((C.res_wall + 1) * m * vy);
            if (C.thermal_wall) {
                vy = -ref_sign(vy) * mt_drand0(A, lane, C.temperature / m);
            } else {
                vy = -C.res_wall * vy;
                vx = C.res_tan * vx + 0.0;
            }
            /* addShearWall (EDMD.c:3429-3439, without the delta model's
             * kick): the top wall drives +0.1 along x, the bottom -0.1 */
            if (C.shear_wall) vx = (y + 2 * r > C.Ly) ? 0.1 : -0.1;
        } else {
            /* the circle (EDMD.c:3441-3454): the normal velocity reversed
             * times resW (thermal: both components drawn), |m mix| to the
             * wall */
            double nx, ny;
            edmd_sincos(edmd_atan2(y - C.halfLx, x - C.halfLx), ny, nx);
            const double mix = -(1 + C.res_wall) * (vx * nx + vy * ny);
            if (C.thermal_wall) {
                vx = -nx * mt_drand0(A, lane, C.temperature);
                vy = -ny * mt_drand0(A, lane, C.temperature);
            } else {
                vx += mix * nx;
                vy += mix * ny;
            }
            dp = fabs(m * mix);
        }
        vx = uni(vx);
        vy = uni(vy);
        if (lane == 0) {
            A.st->dpw += dp;
            if (C.walls & 8) {
                A.st->dpl += dpl;
                A.st->dpm += dpm;
                A.st->dpr += dpr;
            }
            if (C.tape_cap > 0 && d != 7) {
                /* a flat wall's record: j = -1 - xy, the normal and
                 * tangential velocities before the hit, resW, resT and the
                 * contact coordinate */
                EdmdState* st = A.st;
                const int64_t k = st->ntape;
                if (k < C.tape_cap) {
                    EdmdTapeRec rec;
                    rec.t = t;
                    rec.ax = d == 5 ? vx0 : vy0;
                    rec.ay = d == 5 ? vy0 : vx0;
                    rec.cx = C.res_wall;
                    rec.cy = C.res_tan;
                    rec.ff = d == 5 ? x : y;
                    rec.i = p;
                    rec.j = d == 5 ? -1 : -2;
                    rec.g = C.field_on ? C.field : 0.0;
                    A.tape[k] = rec;
                }
                st->ntape = k + 1;
            }
            const auto q = A.p + p;
            q->x = x;
            q->y = y;
            q->tl = tl;
            q->vx = vx;
            q->vy = vy;
            q->coll = pp.coll + 1;
        }
        WAVE_SYNC();
    } else {
        /* removeFromCell: drop p, keeping the order of the others */
        const int c0 = cy * C.ncx + cx;
        const int n0 = A.ccnt[c0];
        int v = (lane < n0) ? A.cslot[(size_t)c0 * C.cap + lane] : -1;
        uint64_t hit = __ballot(v == p);
        int pos = hit ? (int)__builtin_ctzll(hit) : n0;
        if (lane > pos && lane < n0) A.cslot[(size_t)c0 * C.cap + lane - 1] = v;
        if (lane == 0) A.ccnt[c0] = n0 - 1;

        int wx = 0, wy = 0; /* periodic image change */
        switch (d) {
            case 1:
                if (cx - 1 == -1) { cx = C.ncx - 1; wx = -1; } else cx -= 1;
                break;
            case 2:
                if (cx + 1 == C.ncx) { cx = 0; wx = 1; } else cx += 1;
                break;
            case 3:
                if (cy - 1 == -1) { cy = C.ncy - 1; wy = -1; } else cy -= 1;
                break;
            default:
                if (cy + 1 == C.ncy) { cy = 0; wy = 1; } else cy += 1;
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
            const auto q = A.p + p;
            q->x = x;
            q->y = y;
            q->tl = tl;
            if (W) { /* changed by the flight under the field or damping */
                q->vx = vx;
                q->vy = vy;
            }
            if (G) q->rad = r;
            q->cx = (int16_t)cx;
            q->cy = (int16_t)cy;
            if (wx) A.img[2 * p] += wx;
            if (wy) A.img[2 * p + 1] += wy;
        }
        WAVE_SYNC();
    }
    const Owner o{p, x, y, vx, vy, r, cx, cy, gvr};
    bool ov = false;
    double dt, dtb;
    int pa, pb;
    uint32_t cpa, cpb;
    predict_owners<G, W, (ML > 2 && !W)>(C, A, lane, t, 1, o, o, dt, pa, cpa, dtb, pb, cpb, ov,
                                         R.q->cand);
    flag_overlap(R, ov, p);
    double tc;
    int dir;
    cross_time<W && !G>(C, t, x, y, vx, vy, cx, cy, tc, dir);
    const double tk = t + dt;
    /* crossingEvent(i), then collisionEvent(i), EDMD.c:2587-2589, 3458-3459 */
    const int64_t rx = take_rank(C, R, tc);
    const int64_t rk = take_rank(C, R, tk);
    if (lane == 0) {
        EdmdEv e;
        e.tcoll = tk; e.tcross = tc; e.partner = pa; e.cact = cpa; e.dir = dir; e.pad = 0;
        A.e[p] = e;
        A.erank[2 * p] = rk;
        A.erank[2 * p + 1] = rx;
    }
    path_update<ML>(C, A, lane, P, p, make_leaf(tk, rk, tc, rx, pa, dir), -1, Leaf{}, R.q->top, R.t);
}

/* Persistent event loop of one system on one wavefront; ML bounds the
 * calendar depth of every system in the launch.  A holds the system's arrays
 * in HBM (EdmdArrays) or in LDS (EdmdArraysL); the caller has checked that
 * the system's growth phase matches G. */
template <int ML, bool G, bool W, int WPB = 1, class AR>
__device__ __forceinline__ void run_loop(const EdmdConst& C, const AR& A, EdmdRunArgs ra) {
    /* WPB wavefronts per workgroup, each looping over its own system with
     * its own LDS queue state */
    const int lane = threadIdx.x & 63;
    EdmdState* st = A.st;
    __shared__ QueueState Qs[WPB];
    LoopRegs R;
    R.q = &Qs[WPB > 1 ? wave_in_group() : 0];
    R.t = uni(st->t);
    R.ncol = st->ncol;
    R.nevents = st->nevents;
    R.status = st->status;
    R.err_i = st->err_i;
    R.err_j = st->err_j;
    R.paul = uni(st->paul);
    R.ctr = uni64(st->ctr);
    R.win = uni64(st->win);
    R.cur = uni((int)(R.win % C.nlists));
    if (lane == 0)
        for (int k = 0; k < 4; ++k) R.q->vir[k] = st->vir[k];
    next_special(st, R);
    top_load(C, A, lane, R.q->top);
    uint64_t ncross = st->ncross;
    int reason = EDMD_STOP_NONE;
    int last_type = st->last_type;

    const int64_t ev_limit = st->ev_limit < ra.max_events ? st->ev_limit : ra.max_events;
    const uint64_t coll_limit = st->coll_limit < ra.max_coll ? st->coll_limit : ra.max_coll;

    for (;;) {
        if (R.status != EDMD_DEV_OK) { reason = EDMD_STOP_ERROR; break; }
        if ((int64_t)R.nevents >= ev_limit) { reason = EDMD_STOP_EVENTS; break; }
        if (R.ncol >= coll_limit) { reason = EDMD_STOP_COLLISIONS; break; }
        if (!(R.t <= ra.tmax)) { reason = EDMD_STOP_TIME; break; }
        double key;
        int p, aux;
        top_min(C, A, lane, R.q->top, R.t, key, p, aux);
        if (!(key < INF_D) && !(R.next_special < INF_D)) { reason = EDMD_STOP_EMPTY; break; }
        if (pick_next(C, A, st, R, key, p, aux)) {
            const int sp = R.special;
            R.t = R.next_special;
            trace_push(C, A, lane, R.t, sp, 0, 0, 1);
            if (sp == EDMD_EV_SCREENSHOT) {
                ev_screenshot<G, W>(C, A, lane, R);
            } else if (sp == EDMD_EV_THERMO) {
                ev_thermo(C, A, lane, R);
            } else if (sp == EDMD_EV_GROWSTOP) {
                if (G) {
                    flag_overlap(R, ev_stop_grow<W>(C, A, lane, R), 0);
                    R.ncol = 0; /* stopGrow resets ncol and ncross, EDMD.c:4745-4746 */
                    ncross = 0;
                }
            } else {
                if (ev_rescale<G, W>(C, A, lane, R) && R.status == EDMD_DEV_OK)
                    R.status = EDMD_DEV_OVERLAP;
            }
            next_special(st, R);
            R.nevents++;
            last_type = sp;
            if (G && sp == EDMD_EV_GROWSTOP) { reason = EDMD_STOP_GROWN; break; }
            if (ra.stop_mask & (1 << sp)) { reason = EDMD_STOP_SPECIAL; break; }
            continue;
        }
        R.t = key;
        if (aux >= 0) {
            if (G) ev_collision_grow<ML>(C, A, lane, R, p, aux);
            else ev_collision<ML, W>(C, A, lane, R, p, aux);
            last_type = EDMD_EV_COLLISION;
        } else {
            /* a crossing (code -1..-4) or a wall hit (-5 - xy) */
            const bool wall = W && aux <= EDMD_WALL_CODE(0);
            ev_single<ML, G, W>(C, A, lane, R, p, wall ? 5 + EDMD_WALL_CODE(0) - aux : -aux);
            if (wall) {
                last_type = EDMD_EV_WALL;
            } else {
                ncross++;
                last_type = EDMD_EV_CELLCROSS;
            }
        }
        R.nevents++;
        if (ra.stop_mask & (1 << last_type)) { reason = EDMD_STOP_SPECIAL; break; }
    }

    if (lane == 0) {
        st->t = R.t;
        st->ncol = R.ncol;
        st->ncross = ncross;
        st->nevents = R.nevents;
        st->status = R.status;
        st->err_i = R.err_i;
        st->err_j = R.err_j;
        st->stop_reason = reason;
        st->last_type = last_type;
        st->paul = R.paul;
        st->ctr = R.ctr;
        st->win = R.win;
        for (int k = 0; k < 4; ++k) st->vir[k] = R.q->vir[k];
    }
    top_store(C, A, lane, R.q->top);
}

/* Systems of up to 64^2 disks (calendar depth <= 2): bounded to 128 VGPRs so
 * that four waves share each SIMD and a batch of 4096 replicas is resident
 * at once on 256 CUs. */
template <int ML, bool G, bool W, int WPB = 1>
__global__ __launch_bounds__(64 * WPB, 4) void k_run_small(const EdmdSys* __restrict__ systems,
                                                           EdmdRunArgs ra) {
    /* WPB systems per workgroup, one per wavefront (the launch's systems
     * per workgroup are tuned; the last workgroup may hold fewer) */
    const int sys = WPB > 1 ? (int)blockIdx.x * WPB + wave_in_group() : (int)blockIdx.x;
    if (WPB > 1 && sys >= ra.nsys) return;
    const EdmdSys& S = systems[sys];
    /* G launches run the systems that are growing, the others the rest */
    if ((uni(S.a.st->growing) != 0) != G) return;
    run_loop<ML, G, W, WPB>(S.c, S.a, ra);
}

/* Deeper calendars (large systems, typically one per launch): no register
 * bound, so the path caches stay in registers. */
template <int ML, bool G, bool W>
__global__ __launch_bounds__(64) void k_run(const EdmdSys* __restrict__ systems, EdmdRunArgs ra) {
    const EdmdSys& S = systems[blockIdx.x];
    if ((uni(S.a.st->growing) != 0) != G) return;
    run_loop<ML, G, W>(S.c, S.a, ra);
}

/* ----------------------------------------------------- LDS-resident systems */

#if defined(__HIP_DEVICE_COMPILE__)
#define EDMD_LDS __attribute__((address_space(3)))
#else
#define EDMD_LDS
#endif

/* The arrays of a system whose event loop runs out of LDS.  Everything the
 * loop reads and writes on each event (particles, event slots, calendar
 * nodes, ranks, image counters, cell lists, growth rates) sits in the
 * workgroup's LDS for the whole launch; the per-system record and the
 * output buffers (trace, thermo rows, tape, derivative state) stay in HBM.
 * Fields mean what the fields of EdmdArrays of the same name mean. */
struct EdmdArraysL {
    EdmdPart EDMD_LDS* p;
    EdmdEv EDMD_LDS* e;
    EdmdNode EDMD_LDS* leaf;
    EdmdNode EDMD_LDS* lev[EDMD_MAX_LEVELS];
    int32_t EDMD_LDS* img;
    int32_t EDMD_LDS* ccnt;
    int32_t EDMD_LDS* cslot;
    EdmdState EDMD_GLOBAL* st;
    EdmdTraceRec EDMD_GLOBAL* trace;
    EdmdThermoRow EDMD_GLOBAL* thermo;
    double EDMD_LDS* vr;
    int64_t EDMD_LDS* erank;
    int64_t EDMD_LDS* lrank;
    EdmdTapeRec EDMD_GLOBAL* tape;
    double EDMD_GLOBAL* dv;
    double EDMD_GLOBAL* dtl;
    double EDMD_GLOBAL* otl;
    EdmdDeriv EDMD_GLOBAL* dr;
    int32_t EDMD_GLOBAL* ptype;
    uint32_t EDMD_GLOBAL* mt;
};

/* Copies one array between HBM and LDS with the whole wavefront, in 16-,
 * 8- or 4-byte words as the HBM address and the length allow (LDS offsets
 * are 16-byte aligned; every length is a multiple of 4). */
template <bool IN>
__device__ __forceinline__ void lds_xfer(unsigned char EDMD_LDS* l, const void EDMD_GLOBAL* gv,
                                         uint32_t bytes, int lane) {
    unsigned char EDMD_GLOBAL* g = (unsigned char EDMD_GLOBAL*)gv;
    const uint32_t al = uni((int)((uint32_t)(uintptr_t)g | bytes));
    if ((al & 15) == 0) {
        int4 EDMD_LDS* L = (int4 EDMD_LDS*)l;
        int4 EDMD_GLOBAL* H = (int4 EDMD_GLOBAL*)g;
        for (uint32_t i = lane; i < bytes / 16; i += 64) {
            if (IN) L[i] = H[i]; else H[i] = L[i];
        }
    } else if ((al & 7) == 0) {
        int2 EDMD_LDS* L = (int2 EDMD_LDS*)l;
        int2 EDMD_GLOBAL* H = (int2 EDMD_GLOBAL*)g;
        for (uint32_t i = lane; i < bytes / 8; i += 64) {
            if (IN) L[i] = H[i]; else H[i] = L[i];
        }
    } else {
        int EDMD_LDS* L = (int EDMD_LDS*)l;
        int EDMD_GLOBAL* H = (int EDMD_GLOBAL*)g;
        for (uint32_t i = lane; i < bytes / 4; i += 64) {
            if (IN) L[i] = H[i]; else H[i] = L[i];
        }
    }
}

/* Moves a system's loop arrays into LDS (IN) or back to HBM (!IN); the
 * growth rates are read-only and only come in. */
template <bool IN, bool G>
__device__ __forceinline__ void lds_system(const EdmdConst& C, const EdmdArrays& H,
                                           const EdmdArraysL& A, int lane) {
    const uint32_t n = (uint32_t)C.n, ncell = (uint32_t)(C.ncx * C.ncy);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.p, H.p, n * 64, lane);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.e, H.e, n * 32, lane);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.leaf, H.leaf, n * 16, lane);
    for (int l = 0; l < C.nlev; l++)
        lds_xfer<IN>((unsigned char EDMD_LDS*)A.lev[l], H.lev[l], (uint32_t)C.lsize[l + 1] * 16, lane);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.erank, H.erank, n * 16, lane);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.lrank, H.lrank, n * 8, lane);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.img, H.img, n * 8, lane);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.ccnt, H.ccnt, ncell * 4, lane);
    lds_xfer<IN>((unsigned char EDMD_LDS*)A.cslot, H.cslot, ncell * (uint32_t)C.cap * 4, lane);
    if (IN && G) lds_xfer<true>((unsigned char EDMD_LDS*)A.vr, H.vr, n * 8, lane);
}

/* One workgroup (one wavefront) per system with the system's loop arrays
 * resident in LDS: they come in from HBM when the launch starts, the event
 * loop runs on them, and they go back when it stops.  The dynamic LDS of
 * the launch holds the largest system's arrays (edmd_lds_layout). */
template <int ML, bool G, bool W>
__global__ __launch_bounds__(64) void k_run_lds(const EdmdSys* __restrict__ systems,
                                                EdmdRunArgs ra) {
    extern __shared__ __align__(64) unsigned char lds_pool[];
    const EdmdSys& S = systems[blockIdx.x];
    const EdmdConst& C = S.c;
    const EdmdArrays& H = S.a;
    const int lane = threadIdx.x;
    if ((uni(H.st->growing) != 0) != G) return;
    const EdmdLdsLayout L = edmd_lds_layout(&C, G ? 1 : 0);
    unsigned char EDMD_LDS* base = (unsigned char EDMD_LDS*)lds_pool;
    /* the pointer table itself lives in LDS: the loop indexes lev[] with
     * run-time levels, which would put a register copy in scratch */
    __shared__ EdmdArraysL A;
    A.p = (EdmdPart EDMD_LDS*)(base + L.p);
    A.e = (EdmdEv EDMD_LDS*)(base + L.e);
    A.leaf = (EdmdNode EDMD_LDS*)(base + L.leaf);
    for (int l = 0; l < EDMD_MAX_LEVELS; l++) A.lev[l] = (EdmdNode EDMD_LDS*)(base + L.lev[l]);
    A.img = (int32_t EDMD_LDS*)(base + L.img);
    A.ccnt = (int32_t EDMD_LDS*)(base + L.ccnt);
    A.cslot = (int32_t EDMD_LDS*)(base + L.cslot);
    A.vr = (double EDMD_LDS*)(base + L.vr);
    A.erank = (int64_t EDMD_LDS*)(base + L.erank);
    A.lrank = (int64_t EDMD_LDS*)(base + L.lrank);
    A.st = H.st;
    A.trace = H.trace;
    A.thermo = H.thermo;
    A.tape = H.tape;
    A.dv = H.dv;
    A.dtl = H.dtl;
    A.otl = H.otl;
    A.dr = H.dr;
    A.ptype = H.ptype;
    A.mt = H.mt;
    WAVE_SYNC();
    lds_system<true, G>(C, H, A, lane);
    WAVE_SYNC();
    run_loop<ML, G, W>(C, A, ra);
    WAVE_SYNC();
    lds_system<false, G>(C, H, A, lane);
}

/* ------------------------------------------------------------ single steps */

/* findNextEvent, EDMD.c:2313-2327: the earliest pending event, particle or
 * scheduled, without handling it. */
__global__ __launch_bounds__(64) void k_peek(const EdmdSys* __restrict__ systems, EdmdPeek* out) {
    const EdmdConst& C = systems[blockIdx.x].c;
    const EdmdArrays& A = systems[blockIdx.x].a;
    const int lane = threadIdx.x;
    double key;
    int p, aux;
    top_min(C, A, lane, level(A, C.nlev), A.st->t, key, p, aux);
    __shared__ QueueState Qs;
    LoopRegs R;
    R.q = &Qs;
    next_special(A.st, R);
    const int64_t krank = key == R.next_special ? uni64(A.lrank[p]) : (int64_t)0;
    if (lane != 0) return;
    EdmdPeek r;
    if (!before(key, krank, R.next_special, R.srank)) {
        r.t = R.next_special; r.type = R.special; r.i = 0; r.j = 0; r.coll_actual = 0;
    } else if (aux >= 0) {
        r.t = key; r.type = EDMD_EV_COLLISION; r.i = p; r.j = aux; r.coll_actual = A.e[p].cact;
    } else if (aux <= EDMD_WALL_CODE(0)) {
        r.t = key; r.type = EDMD_EV_WALL; r.i = p; r.j = EDMD_WALL_CODE(0) - aux; r.coll_actual = 0;
    } else {
        r.t = key; r.type = EDMD_EV_CELLCROSS; r.i = p; r.j = -aux; r.coll_actual = 0;
    }
    out[blockIdx.x] = r;
}

/* collisionTimeNormal(p1, p2) at the system time, EDMD.c:2661-2723; p1 is
 * taken at its stored position as the reference does. */
__global__ void k_pair_time(const EdmdSys* __restrict__ systems, int i, int j, double* out) {
    const EdmdConst& C = systems[0].c;
    const EdmdArrays& A = systems[0].a;
    bool ov = false;
    const double t = A.st->t;
    const EdmdPart a = A.p[i], b = A.p[j];
    if (A.st->growing)
        out[0] = coll_time_grow(C, t, a.x, a.y, a.vx, a.vy, a.rad, A.vr[i], b.x, b.y, b.vx, b.vy,
                                b.tl, b.rad, A.vr[j], ov);
    else
        out[0] = coll_time(C, t, a.x, a.y, a.vx, a.vy, a.rad, b.x, b.y, b.vx, b.vy, b.tl, b.rad, ov);
    out[1] = ov ? 1.0 : 0.0;
}

/* collisionEvent(i) / crossingEvent(i) for one particle (which: bit 0 =
 * crossing, bit 1 = collision) at the system time, then the calendar path
 * of i is updated. */
template <int ML>
__global__ __launch_bounds__(64) void k_repredict(const EdmdSys* __restrict__ systems, int i,
                                                 int which) {
    const EdmdConst& C = systems[0].c;
    const EdmdArrays& A = systems[0].a;
    const int lane = threadIdx.x;
    __shared__ int32_t stage[EDMD_STAGE_MAX];
    const double t = A.st->t;
    const EdmdPart pi = A.p[i];
    EdmdEv e = A.e[i];
    PathCache<ML> P;
    path_prefetch<ML>(C, A, lane, i, -1, P);
    const double paul = A.st->paul;
    int64_t ctr = A.st->ctr;
    int64_t rcoll = A.erank[2 * i], rcross = A.erank[2 * i + 1];
    if (which & 1) {
        cross_time(C, t, pi.x, pi.y, pi.vx, pi.vy, pi.cx, pi.cy, e.tcross, e.dir);
        rcross = rank_of(e.tcross, paul, C.dtp, C.horizon, ctr++);
    }
    bool ov = false;
    if (which & 2) {
        const bool grow = uni(A.st->growing) != 0;
        const Owner o{i, pi.x, pi.y, pi.vx, pi.vy, pi.rad, pi.cx, pi.cy, grow ? A.vr[i] : 0.0};
        double dt, dtb;
        int pb;
        uint32_t cpb;
        if (grow)
            predict_owners<true>(C, A, lane, t, 1, o, o, dt, e.partner, e.cact, dtb, pb, cpb, ov, stage);
        else
            predict_owners<false>(C, A, lane, t, 1, o, o, dt, e.partner, e.cact, dtb, pb, cpb, ov, stage);
        e.tcoll = t + dt;
        rcoll = rank_of(e.tcoll, paul, C.dtp, C.horizon, ctr++);
    }
    if (lane == 0) {
        A.e[i] = e;
        A.erank[2 * i] = rcoll;
        A.erank[2 * i + 1] = rcross;
    }
    path_update<ML>(C, A, lane, P, i, make_leaf(e.tcoll, rcoll, e.tcross, rcross, e.partner, e.dir),
                    -1, Leaf{}, nullptr, t);
    const bool any_ov = __ballot(ov) != 0ull;
    if (lane == 0) {
        A.st->ctr = ctr;
        if (any_ov) {
            A.st->status = EDMD_DEV_OVERLAP;
            A.st->err_i = i;
        }
    }
}

/* addEventToQueue / removeEventFromQueue (EDMD.c:2144-2221) on one slot of
 * the reference's eventList, as a caller fills it (addCollisionEvent,
 * addCrossingEvent, addWallEvent, addEvent*: EDMD.c:2487-2496, 3286-3297,
 * 3321-3330, 4421-4440, 4814-4822).  mode 0: file the slot at time t with
 * partner / direction / wall side j and partner counter cact; mode 1: the
 * same with the partner's current counter (addCollisionEvent); mode 2:
 * take the slot out (it never pops until filed again).  A filed slot takes
 * the next insertion rank, as the reference's queue orders it. */
template <int ML>
__global__ __launch_bounds__(64) void k_queue_edit(const EdmdSys* __restrict__ systems, int kind,
                                                  int i, int j, double t, uint32_t cact,
                                                  int mode) {
    const EdmdConst& C = systems[0].c;
    const EdmdArrays& A = systems[0].a;
    const int lane = threadIdx.x;
    EdmdState* st = A.st;
    const double paul = st->paul;
    int64_t ctr = st->ctr;
    const bool remove = mode == 2;
    if (kind == EDMD_EV_SCREENSHOT || kind == EDMD_EV_THERMO || kind == EDMD_EV_ADDINGNOISE) {
        const double tn = remove ? INF_D : t;
        const int64_t rk = remove ? INT64_MAX : rank_of(t, paul, C.dtp, C.horizon, ctr);
        const int64_t wk = remove ? INT64_MAX : window_of(t, paul, C.dtp, C.nlists, st->win);
        if (!remove) ctr++;
        if (lane == 0) {
            if (kind == EDMD_EV_SCREENSHOT) {
                st->next_screen = tn; st->rank_screen = rk; st->win_screen = wk;
            } else if (kind == EDMD_EV_THERMO) {
                st->next_thermo = tn; st->rank_thermo = rk; st->win_thermo = wk;
            } else {
                st->next_noise = tn; st->rank_noise = rk; st->win_noise = wk;
            }
            st->ctr = ctr;
        }
        return;
    }
    EdmdEv e = A.e[i];
    int64_t rcoll = A.erank[2 * i], rcross = A.erank[2 * i + 1];
    PathCache<ML> P;
    path_prefetch<ML>(C, A, lane, i, -1, P);
    if (kind == EDMD_EV_CELLCROSS) {
        if (remove) {
            e.tcross = INF_D;
        } else {
            e.tcross = t;
            e.dir = j;
            rcross = rank_of(t, paul, C.dtp, C.horizon, ctr++);
        }
    } else {
        if (remove) {
            e.tcoll = INF_D;
        } else {
            e.tcoll = t;
            e.partner = kind == EDMD_EV_WALL ? EDMD_WALL_CODE(j) : j;
            e.cact = kind == EDMD_EV_WALL ? 0u : (mode == 1 ? A.p[j].coll : cact);
            rcoll = rank_of(t, paul, C.dtp, C.horizon, ctr++);
        }
    }
    if (lane == 0) {
        A.e[i] = e;
        A.erank[2 * i] = rcoll;
        A.erank[2 * i + 1] = rcross;
    }
    path_update<ML>(C, A, lane, P, i, make_leaf(e.tcoll, rcoll, e.tcross, rcross, e.partner, e.dir),
                    -1, Leaf{}, nullptr, st->t);
    if (lane == 0) st->ctr = ctr;
}

/* freeFlyNormal(p) at the system time, EDMD.c:4954-4990. */
__global__ void k_fly_one(const EdmdSys* __restrict__ systems, int i) {
    const EdmdConst& C = systems[0].c;
    const EdmdArrays& A = systems[0].a;
    EdmdPart* r = A.p + i;
    double x = r->x, y = r->y, tl = r->tl;
    if (A.st->growing) {
        double rad = r->rad;
        fly_grow(C, A.st->t, x, y, r->vx, r->vy, tl, rad, A.vr[i]);
        r->rad = rad;
    } else {
        double vx = r->vx, vy = r->vy;
        fly_v<true>(C, A.st->t, x, y, vx, vy, tl);
        r->vx = vx;
        r->vy = vy;
    }
    r->x = x;
    r->y = y;
    r->tl = tl;
}

/* calculate_pcf, pcf.c:16-75: histogram of minimum-image pair distances
 * below rmax = dr * nb, one count per unordered pair (the reference adds
 * 2.0 per pair; the host doubles the counts).  Blocks take 256 x 256 tiles
 * of the upper pair triangle and count into an LDS histogram.  extrap != 0
 * uses positions free-flown to time t without storing them. */
__global__ __launch_bounds__(256) void k_pcf(const double* __restrict__ px,
                                            const double* __restrict__ py,
                                            const EdmdPart* __restrict__ part, int n, double t,
                                            int extrap, double Lx, double Ly, double dr, int nb,
                                            unsigned long long* __restrict__ hist) {
    extern __shared__ unsigned int h[];
    __shared__ double sx[256], sy[256];
    const int ti = blockIdx.x, tj = blockIdx.y;
    if (tj < ti) return;
    const double halfLx = Lx / 2, halfLy = Ly / 2;
    const double rmax = dr * nb;
    for (int b = threadIdx.x; b < nb; b += 256) h[b] = 0;
    auto pos = [&](int q, double& xx, double& yy) {
        if (part) {
            const EdmdPart r = part[q];
            xx = r.x;
            yy = r.y;
            if (extrap) {
                double dt = t - r.tl;
                xx += dt * r.vx;
                yy += dt * r.vy;
                xx = wrap_post(xx, Lx);
                yy = wrap_post(yy, Ly);
            }
        } else {
            xx = px[q];
            yy = py[q];
        }
    };
    const int j0 = tj * 256;
    if (j0 + (int)threadIdx.x < n) pos(j0 + threadIdx.x, sx[threadIdx.x], sy[threadIdx.x]);
    __syncthreads();
    const int i = ti * 256 + threadIdx.x;
    if (i < n) {
        double xi, yi;
        pos(i, xi, yi);
        const int jn = min(256, n - j0);
        for (int k = 0; k < jn; ++k) {
            const int j = j0 + k;
            if (j <= i) continue;
            double dx = sx[k] - xi;
            double dy = sy[k] - yi;
            if (dx >= halfLx) dx -= Lx;
            else if (dx < -halfLx) dx += Lx;
            if (dy >= halfLy) dy -= Ly;
            else if (dy < -halfLy) dy += Ly;
            double r = sqrt(dx * dx + dy * dy);
            if (r < rmax) {
                int bin = (int)(r / dr);
                if (bin < nb) atomicAdd(&h[bin], 1u);
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256)
        if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

__global__ void k_fp64_check(int n, const double* __restrict__ a, const double* __restrict__ b,
                             double* __restrict__ q, double* __restrict__ s) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    q[k] = a[k] / b[k];
    s[k] = sqrt(a[k]);
}

/* ------------------------------------------------------------ launchers */

template <class S>
static void launch_sweep(const EdmdSys* d_sys, int nsys, int max_tiles, int ext, hipStream_t s) {
    dim3 grid((max_tiles + 7) / 8 * 8, nsys);
    if (ext & EDMD_SWEEP_EXT)
        hipLaunchKernelGGL((k_sweep_tiled<S, true, true>), grid, dim3(S::THREADS), 0, s, d_sys);
    else if (ext & EDMD_SWEEP_POLY)
        hipLaunchKernelGGL((k_sweep_tiled<S, false, true>), grid, dim3(S::THREADS), 0, s, d_sys);
    else
        hipLaunchKernelGGL((k_sweep_tiled<S, false, false>), grid, dim3(S::THREADS), 0, s, d_sys);
}

hipError_t edmd_launch_sweep_tiled(const EdmdSys* d_sys, int nsys, int max_tiles, int variant,
                                   int ext, hipStream_t s) {
    if (nsys <= 0 || max_tiles <= 0) return hipSuccess;
    switch (variant) {
        case 0: launch_sweep<SweepV0>(d_sys, nsys, max_tiles, ext, s); break;
        case 1: launch_sweep<SweepV1>(d_sys, nsys, max_tiles, ext, s); break;
        case 2: launch_sweep<SweepV2>(d_sys, nsys, max_tiles, ext, s); break;
        case 3: launch_sweep<SweepV3>(d_sys, nsys, max_tiles, ext, s); break;
        case 4: launch_sweep<SweepV4>(d_sys, nsys, max_tiles, ext, s); break;
        case 5: launch_sweep<SweepV5>(d_sys, nsys, max_tiles, ext, s); break;
        case 6: launch_sweep<SweepV6>(d_sys, nsys, max_tiles, ext, s); break;
        case 7: launch_sweep<SweepV7>(d_sys, nsys, max_tiles, ext, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class S>
static EdmdSweepShape shape_of() {
    return EdmdSweepShape{S::TX, S::TY, S::THREADS, S::STAGE};
}

EdmdSweepShape edmd_sweep_shape(int variant) {
    switch (variant) {
        case 0: return shape_of<SweepV0>();
        case 1: return shape_of<SweepV1>();
        case 2: return shape_of<SweepV2>();
        case 3: return shape_of<SweepV3>();
        case 4: return shape_of<SweepV4>();
        case 5: return shape_of<SweepV5>();
        case 6: return shape_of<SweepV6>();
        case 7: return shape_of<SweepV7>();
        default: return EdmdSweepShape{0, 0, 0, 0};
    }
}

int edmd_sweep_tiles(int ncx, int ncy, int variant) {
    const EdmdSweepShape sh = edmd_sweep_shape(variant);
    if (sh.tx <= 0) return 0;
    return ((ncx + sh.tx - 1) / sh.tx) * ((ncy + sh.ty - 1) / sh.ty);
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

/* W: some system of the launch has walls, the field or damping. */
template <bool G, bool W>
static void launch_run(const EdmdSys* d_sys, int nsys, int max_nlev, EdmdRunArgs ra, hipStream_t s) {
    ra.nsys = nsys;
    /* the _small loop is bounded to 128 vector registers so that four
     * wavefronts share each SIMD; a batch of at most one system per SIMD
     * gains nothing from that and runs the unbounded loop (the depth-2
     * loop spills to scratch under the bound) */
    const bool few = nsys <= 1024;
    if (max_nlev <= 1 && few) hipLaunchKernelGGL((k_run<1, G, W>), dim3(nsys), dim3(64), 0, s, d_sys, ra);
    else if (max_nlev <= 1) hipLaunchKernelGGL((k_run_small<1, G, W>), dim3(nsys), dim3(64), 0, s, d_sys, ra);
    else if (max_nlev == 2 && few) hipLaunchKernelGGL((k_run<2, G, W>), dim3(nsys), dim3(64), 0, s, d_sys, ra);
    else if (max_nlev == 2) hipLaunchKernelGGL((k_run_small<2, G, W>), dim3(nsys), dim3(64), 0, s, d_sys, ra);
    else if (max_nlev == 3) hipLaunchKernelGGL((k_run<3, G, W>), dim3(nsys), dim3(64), 0, s, d_sys, ra);
    else if (max_nlev <= 4) hipLaunchKernelGGL((k_run<4, G, W>), dim3(nsys), dim3(64), 0, s, d_sys, ra);
    else hipLaunchKernelGGL((k_run<EDMD_MAX_LEVELS, G, W>), dim3(nsys), dim3(64), 0, s, d_sys, ra);
}

/* Epochs of a graph-captured run: the bounds of the coming epoch relative
 * to each system's counters (one thread per system), and each system's
 * counters when the epoch has run. */
__global__ __launch_bounds__(256) void k_epoch_begin(const EdmdSys* __restrict__ systems, int nsys,
                                                     int64_t ev_inc, uint64_t coll_inc) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    EdmdState EDMD_GLOBAL* st = systems[s].a.st;
    const uint64_t ne = st->nevents, nc = st->ncol;
    st->ev_limit = ev_inc > INT64_MAX - (int64_t)ne ? INT64_MAX : (int64_t)ne + ev_inc;
    st->coll_limit = coll_inc > ~(uint64_t)0 - nc ? ~(uint64_t)0 : nc + coll_inc;
}

__global__ __launch_bounds__(256) void k_epoch_end(const EdmdSys* __restrict__ systems, int nsys,
                                                   EdmdEpochRow* __restrict__ rows) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const EdmdState EDMD_GLOBAL* st = systems[s].a.st;
    EdmdEpochRow r;
    r.t = st->t;
    r.ncol = st->ncol;
    r.ncross = st->ncross;
    r.nevents = st->nevents;
    r.status = st->status;
    r.stop_reason = st->stop_reason;
    rows[s] = r;
}

hipError_t edmd_launch_epoch_begin(const EdmdSys* d_sys, int nsys, int64_t ev_inc,
                                   uint64_t coll_inc, hipStream_t s) {
    hipLaunchKernelGGL(k_epoch_begin, dim3((nsys + 255) / 256), dim3(256), 0, s, d_sys, nsys,
                       ev_inc, coll_inc);
    return hipGetLastError();
}

hipError_t edmd_launch_epoch_end(const EdmdSys* d_sys, int nsys, EdmdEpochRow* rows,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_epoch_end, dim3((nsys + 255) / 256), dim3(256), 0, s, d_sys, nsys, rows);
    return hipGetLastError();
}

/* Systems of calendar depth <= 2 without growth, walls, field or damping,
 * wpb (2 or 4) systems per workgroup. */
template <int WPB>
static void launch_run_packed(const EdmdSys* d_sys, int nsys, int max_nlev, EdmdRunArgs ra,
                              hipStream_t s) {
    ra.nsys = nsys;
    const dim3 grid((nsys + WPB - 1) / WPB), block(64 * WPB);
    if (max_nlev <= 1) hipLaunchKernelGGL((k_run_small<1, false, false, WPB>), grid, block, 0, s, d_sys, ra);
    else hipLaunchKernelGGL((k_run_small<2, false, false, WPB>), grid, block, 0, s, d_sys, ra);
}

hipError_t edmd_launch_run(const EdmdSys* d_sys, int nsys, int max_nlev, int grow, int walls,
                           int wpb, EdmdRunArgs ra, hipStream_t s) {
    if (wpb != 1 && (grow || walls || max_nlev > 2 || (wpb != 2 && wpb != 4)))
        return hipErrorInvalidValue;
    if (wpb == 2) launch_run_packed<2>(d_sys, nsys, max_nlev, ra, s);
    else if (wpb == 4) launch_run_packed<4>(d_sys, nsys, max_nlev, ra, s);
    else if (grow && walls) launch_run<true, true>(d_sys, nsys, max_nlev, ra, s);
    else if (grow) launch_run<true, false>(d_sys, nsys, max_nlev, ra, s);
    else if (walls) launch_run<false, true>(d_sys, nsys, max_nlev, ra, s);
    else launch_run<false, false>(d_sys, nsys, max_nlev, ra, s);
    return hipGetLastError();
}

/* The LDS-resident loop exists for calendars of depth <= 2 (up to 4096
 * disks, far more than one workgroup's LDS holds). */
static const void* lds_kernel(int max_nlev, int grow, int walls) {
    if (max_nlev <= 1)
        return grow ? (walls ? (const void*)k_run_lds<1, true, true> : (const void*)k_run_lds<1, true, false>)
                    : walls ? (const void*)k_run_lds<1, false, true> : (const void*)k_run_lds<1, false, false>;
    if (max_nlev == 2)
        return grow ? (walls ? (const void*)k_run_lds<2, true, true> : (const void*)k_run_lds<2, true, false>)
                    : walls ? (const void*)k_run_lds<2, false, true> : (const void*)k_run_lds<2, false, false>;
    return nullptr;
}

hipError_t edmd_lds_occupancy(int max_nlev, int grow, int walls, size_t dyn_bytes,
                              int* blocks_per_cu, size_t* static_bytes) {
    const void* f = lds_kernel(max_nlev, grow, walls);
    *blocks_per_cu = 0;
    *static_bytes = 0;
    if (!f) return hipErrorInvalidValue;
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, f);
    if (e != hipSuccess) return e;
    *static_bytes = fa.sharedSizeBytes;
    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_bytes);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, 64, dyn_bytes);
}

template <bool G, bool W>
static void launch_lds(const EdmdSys* d_sys, int nsys, int max_nlev, size_t dyn_bytes,
                       EdmdRunArgs ra, hipStream_t s) {
    if (max_nlev <= 1) hipLaunchKernelGGL((k_run_lds<1, G, W>), dim3(nsys), dim3(64), dyn_bytes, s, d_sys, ra);
    else hipLaunchKernelGGL((k_run_lds<2, G, W>), dim3(nsys), dim3(64), dyn_bytes, s, d_sys, ra);
}

hipError_t edmd_launch_run_lds(const EdmdSys* d_sys, int nsys, int max_nlev, int grow, int walls,
                               size_t dyn_bytes, EdmdRunArgs ra, hipStream_t s) {
    if (max_nlev > 2) return hipErrorInvalidValue;
    if (grow && walls) launch_lds<true, true>(d_sys, nsys, max_nlev, dyn_bytes, ra, s);
    else if (grow) launch_lds<true, false>(d_sys, nsys, max_nlev, dyn_bytes, ra, s);
    else if (walls) launch_lds<false, true>(d_sys, nsys, max_nlev, dyn_bytes, ra, s);
    else launch_lds<false, false>(d_sys, nsys, max_nlev, dyn_bytes, ra, s);
    return hipGetLastError();
}

hipError_t edmd_launch_sweep_grow(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s) {
    if (nsys <= 0 || max_n <= 0) return hipSuccess;
    dim3 grid((max_n + 255) / 256, nsys);
    hipLaunchKernelGGL(k_sweep_grow, grid, dim3(256), 0, s, d_sys);
    return hipGetLastError();
}

hipError_t edmd_launch_sweep_count(const EdmdSys* d_sys, int nsys, hipStream_t s) {
    hipLaunchKernelGGL(k_sweep_count, dim3((nsys + 255) / 256), dim3(256), 0, s, d_sys, nsys);
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

hipError_t edmd_launch_repredict(const EdmdSys* d_sys, int max_nlev, int i, int which,
                                 hipStream_t s) {
    if (max_nlev <= 1) hipLaunchKernelGGL(k_repredict<1>, dim3(1), dim3(64), 0, s, d_sys, i, which);
    else if (max_nlev <= 4)
        hipLaunchKernelGGL(k_repredict<4>, dim3(1), dim3(64), 0, s, d_sys, i, which);
    else hipLaunchKernelGGL(k_repredict<EDMD_MAX_LEVELS>, dim3(1), dim3(64), 0, s, d_sys, i, which);
    return hipGetLastError();
}

hipError_t edmd_launch_queue_edit(const EdmdSys* d_sys, int max_nlev, int kind, int i, int j,
                                  double t, uint32_t cact, int mode, hipStream_t s) {
    if (max_nlev <= 1)
        hipLaunchKernelGGL(k_queue_edit<1>, dim3(1), dim3(64), 0, s, d_sys, kind, i, j, t, cact, mode);
    else if (max_nlev <= 4)
        hipLaunchKernelGGL(k_queue_edit<4>, dim3(1), dim3(64), 0, s, d_sys, kind, i, j, t, cact, mode);
    else
        hipLaunchKernelGGL(k_queue_edit<EDMD_MAX_LEVELS>, dim3(1), dim3(64), 0, s, d_sys, kind, i, j,
                           t, cact, mode);
    return hipGetLastError();
}

hipError_t edmd_launch_fly_one(const EdmdSys* d_sys, int i, hipStream_t s) {
    hipLaunchKernelGGL(k_fly_one, dim3(1), dim3(1), 0, s, d_sys, i);
    return hipGetLastError();
}

hipError_t edmd_launch_pcf(const double* x, const double* y, const EdmdPart* part, int n, double t,
                           int extrap, double Lx, double Ly, double dr, int nb,
                           unsigned long long* hist, hipStream_t s) {
    const int tiles = (n + 255) / 256;
    hipLaunchKernelGGL(k_pcf, dim3(tiles, tiles), dim3(256), (size_t)nb * sizeof(unsigned int), s,
                       x, y, part, n, t, extrap, Lx, Ly, dr, nb, hist);
    return hipGetLastError();
}

hipError_t edmd_launch_fp64_check(int n, const double* a, const double* b, double* q, double* sq,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_fp64_check, dim3((n + 255) / 256), dim3(256), 0, s, n, a, b, q, sq);
    return hipGetLastError();
}

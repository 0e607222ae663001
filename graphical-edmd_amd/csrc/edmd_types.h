/*
 * Device-resident layout of one hard-disk system, shared by host and device
 * code of libedmd_hip.so.
 *
 * A system is stored in HBM as three record arrays indexed by particle:
 *   particle     64 B: x y vx vy t_local rad m (f64), collision counter
 *                (u32), cell x / y (i16) -- one cache-line half per disk,
 *                gathered whole when the disk is a collision candidate
 *   event        32 B: collision slot (time, partner, partner's collision
 *                counter at prediction) and crossing slot (time, direction);
 *                their ranks among events of equal time sit in a side array
 *                (erank) that is read only when two keys are equal,
 *                the two slots the reference keeps per particle in
 *                eventList[N+i] and eventList[i] (EDMD.c:1935-1948)
 *   calendar     16 B nodes {key, leaf, event code}: one leaf per particle
 *                (key = earliest of its two slots by time, then rank) and a
 *                64-ary tournament tree above the leaves; every node keeps the
 *                minimum (key, rank) of its 64 children (a node's rank is that
 *                of its leaf, lrank[leaf]), the leaf holding it and that leaf's event
 *                code (partner index for a collision, minus the direction
 *                for a crossing), so one read of the top level names both
 *                disks of the next event
 * plus cell lists: per cell a count and `cap` slots holding particle indices
 * in insertion order; the newest entry is visited first, which is the order
 * of the reference's head-inserted linked lists (EDMD.c:2053-2095,
 * 2576-2585), and periodic image counters per particle.
 */
#ifndef EDMD_TYPES_H
#define EDMD_TYPES_H

#include <stdint.h>

#define EDMD_MAX_LEVELS 6
#define EDMD_NEVER 100000000000000000000000000.0 /* EDMD.h:8 */

/* Event kinds as numbered by the reference's enum event (EDMD.h:21-34). */
#define EDMD_EV_CELLCROSS 0
#define EDMD_EV_COLLISION 1
#define EDMD_EV_SCREENSHOT 2
#define EDMD_EV_THERMO 3
#define EDMD_EV_ADDINGNOISE 4
#define EDMD_EV_WALL 5
#define EDMD_EV_GROWSTOP 9

/* A collision slot holding a wall event: partner -5 - xy (xy = 0: wall
 * normal to x, 1: normal to y); the leaf's event code is the same value,
 * below the crossing codes -1..-4. */
#define EDMD_WALL_CODE(xy) (-5 - (xy))

/* collisionEventGrow's starting minimum (EDMD.c:3106): no partner found
 * within 1e7 time units gives partner 0 at t + 1e7 */
#define EDMD_GROW_NONE 10000000.0

/* Why a run launch returned. */
#define EDMD_STOP_NONE 0
#define EDMD_STOP_EVENTS 1
#define EDMD_STOP_COLLISIONS 2
#define EDMD_STOP_TIME 3
#define EDMD_STOP_ERROR 4
#define EDMD_STOP_EMPTY 5
#define EDMD_STOP_SPECIAL 6
#define EDMD_STOP_GROWN 7 /* growth ended; the run continues in the Normal loop */

/* Device status codes (mirrored in include/edmd.h). */
#define EDMD_DEV_OK 0
#define EDMD_DEV_OVERLAP 3
#define EDMD_DEV_CELL_OVERFLOW 7

typedef struct EdmdConst {
    int32_t n, ncx, ncy, cap;
    int32_t nlev;                          /* tree levels above the leaves */
    int32_t lsize[EDMD_MAX_LEVELS + 1];    /* lsize[0] = n, lsize[nlev] <= 64 */
    int32_t noise;                         /* 0 or 2 (velocity rescale) */
    int32_t trace_cap, thermo_cap;
    int32_t tape_cap;                      /* collision tape records kept, 0 = no tape */
    double Lx, Ly, halfLx, halfLy, csx, csy;
    double res, dtime, dtime_thermo, dtnoise, temperature;
    double einit; /* Einit: E/N after the growth stops (EDMD.c:369, 5759-5763) */
    double dtp;   /* dtPaul = 5/N, the reference's bucket width (EDMD.c:712-714) */
    int32_t nlists; /* bucket count of the reference's queue (N) */
    int32_t walls;  /* bit 0: walls normal to x, bit 1: normal to y, bit 2: the circle,
                       bit 3: the mid wall */
    double sizeratio; /* the mid wall's packing factor */
    double res_wall, res_tan; /* resW, resT (EDMD.c:372-373) */
    /* minimum-image half widths: halfLx / halfLy, +inf across walls */
    double mihx, mihy;
    int32_t field_on; /* uniform field along y (addField) */
    int32_t damping;  /* velocities decay as exp(-gamma t) in flight */
    double field;     /* its value (< 0) */
    double gamma;     /* damping rate (gamm) */
    int32_t mono;     /* every disk has the same radius (bitwise) */
    int32_t thermal_wall; /* wall hits draw the new normal speed (thermoWall) */
    int32_t shear_wall;   /* y-wall hits set vx to +-0.1 (addShearWall) */
    int32_t screen;       /* sweep root screen: 0 off, 1 on, 2 on and misses counted */
    double horizon;       /* dtPaul * paulListN: events this far ahead of paulTime or
                             more go to the overflow list (EDMD.c:2154) */
} EdmdConst;

/* Order among events of equal time.  The reference pops equal times in the
 * order its queue holds them (EDMD.c:2144-2327): events scheduled within
 * dtPaul of the current bucket start (paulTime) go straight into the tree,
 * where equal keys go right, so the earlier insertion pops first; later
 * events wait in LIFO bucket lists that are drained into the emptied tree
 * when their bucket comes up, so of those the later insertion pops first,
 * and all of them before anything inserted directly afterwards.  Every
 * scheduled event gets a rank from a per-system insertion counter:
 *   direct:  EDMD_RANK_DIRECT + counter
 *   bucket:  EDMD_RANK_DIRECT - 1 - counter
 * and the calendar orders by (time, rank).  Events of the overflow list
 * (scheduled 5 or more time units ahead) are filed again when the bucket
 * ring wraps, before the new bucket is drained; a special event filed
 * straight into the tree then gets rank = counter, ahead of both classes.
 * A particle event filed into the overflow list gets
 *   overflow: EDMD_RANK_OVF + counter
 * and keeps it until the next wrap.  The list is LIFO and is walked from
 * its head at the wrap, so its events are filed again newest first: the
 * wrap gives each the counter 2 now - 1 - c (now = the counter at the wrap,
 * c = its own), which orders them as the walk does without a sort, and
 * moves the counter past them (the reference refiles them one by one);
 * each is then ranked as the class it is filed into (straight into the
 * tree: rank = that counter, like a special; a bucket; the overflow list
 * again). */
#define EDMD_RANK_DIRECT ((int64_t)1 << 62)
#define EDMD_RANK_OVF ((int64_t)3 << 61)
/* rank r belongs to a particle event waiting in the overflow list */
#define EDMD_RANK_IS_OVF(r) ((r) >= EDMD_RANK_OVF && (r) < EDMD_RANK_OVF + ((int64_t)1 << 60))

typedef struct EdmdState {
    double t;
    double next_screen, next_thermo, next_noise, last_thermo;
    double vir[4]; /* xx yy xy yx, reset at every thermo event */
    uint64_t ncol, ncross, nevents;
    int32_t status, err_i, err_j, stop_reason;
    int64_t ntrace;
    int32_t nthermo, last_type; /* last_type: kind of the last event handled */
    int64_t ev_limit;           /* per-system bound on nevents (checkpoint segments) */
    int64_t ntape;              /* collisions taped since the tape was cleared */
    double wsum;                /* sum of m_i m_j ff |a|^2 over taped collisions */
    double tlast;               /* time of the last taped collision */
    double next_grow;           /* GROWSTOP time (1/vr) while growing, else +inf */
    int32_t growing;            /* radii grow (freeFlyGrow, collisionEventGrow) */
    int32_t pad_grow;
    double paul;                /* the reference's paulTime (bucket start) */
    int64_t ctr;                /* events scheduled so far (rank counter) */
    int64_t rank_screen, rank_thermo, rank_noise, rank_grow; /* ranks of the special events */
    int64_t win;                /* buckets passed (paulTime = win additions of dtPaul) */
    int64_t win_screen, win_thermo, win_noise, win_grow; /* their buckets, -1 = overflow list */
    uint64_t coll_limit;        /* per-system bound on ncol (epochs) */
    double dpw;                 /* momentum given to the walls since the last thermo row */
    double dpl, dpm, dpr;       /* dpLeft, dpMid, dpRight of the mid wall (EDMD.c:146-148) */
    /* the sweep's root screen: marked candidates, roots solved, and (check
     * mode) screened-out roots that the minimum would have taken */
    uint64_t scr_marks, scr_solved, scr_miss;
    /* counter at the last wrap of the bucket ring: the overflow list's
     * particle events hold counters from here on */
    int64_t ovf_lo;
} EdmdState;

/* LDS-resident runs: byte offsets of a system's arrays in its workgroup's
 * LDS (16-byte aligned), shared by the host (to size the launch) and the
 * device (to carve the pool).  Particle, event, calendar, rank, image and
 * cell arrays are resident; the state, trace, thermo rows and tape stay in
 * HBM. */
#if defined(__HIPCC__)
#define EDMD_HD __host__ __device__
#else
#define EDMD_HD
#endif
typedef struct EdmdLdsLayout {
    uint32_t p, e, leaf, lev[EDMD_MAX_LEVELS], erank, lrank, img, ccnt, cslot, vr, total;
} EdmdLdsLayout;

static inline EDMD_HD uint64_t edmd_lds_align(uint64_t b) { return (b + 15) & ~(uint64_t)15; }

static inline EDMD_HD EdmdLdsLayout edmd_lds_layout(const EdmdConst* c, int grow) {
    EdmdLdsLayout L;
    uint64_t o = 0;
    const uint64_t n = (uint64_t)c->n, ncell = (uint64_t)c->ncx * (uint64_t)c->ncy;
    L.p = (uint32_t)o;     o = edmd_lds_align(o + n * 64);
    L.e = (uint32_t)o;     o = edmd_lds_align(o + n * 32);
    L.leaf = (uint32_t)o;  o = edmd_lds_align(o + n * 16);
    for (int l = 0; l < EDMD_MAX_LEVELS; l++) {
        L.lev[l] = (uint32_t)o;
        if (l < c->nlev) o = edmd_lds_align(o + (uint64_t)c->lsize[l + 1] * 16);
    }
    L.erank = (uint32_t)o; o = edmd_lds_align(o + n * 16);
    L.lrank = (uint32_t)o; o = edmd_lds_align(o + n * 8);
    L.img = (uint32_t)o;   o = edmd_lds_align(o + n * 8);
    L.ccnt = (uint32_t)o;  o = edmd_lds_align(o + ncell * 4);
    L.cslot = (uint32_t)o; o = edmd_lds_align(o + ncell * (uint64_t)c->cap * 4);
    L.vr = (uint32_t)o;    o = edmd_lds_align(o + (grow ? n * 8 : 0));
    L.total = o > 0xffffffffu ? 0xffffffffu : (uint32_t)o;
    return L;
}

/* Per-system counters at the end of one epoch of a graph-captured run. */
typedef struct EdmdEpochRow {
    double t;
    uint64_t ncol, ncross, nevents;
    int32_t status, stop_reason;
} EdmdEpochRow;

/* One valid collision as the derivative kernels need it: time, contact
 * vector a = r_j - r_i (minimum image), relative velocity c = v_j - v_i
 * before the collision and the impulse factor ff of the collision rule. */
typedef struct __attribute__((aligned(64))) EdmdTapeRec {
    double t, ax, ay, cx, cy, ff;
    int32_t i, j;
    double g; /* wall records: the field along y (0 without it); pairs: 0 */
} EdmdTapeRec;

/* Per-system accumulators of the tangent / adjoint passes. */
typedef struct EdmdDeriv {
    double dsum;   /* tangent: derivative of wsum */
    double dtlast; /* tangent: derivative of tlast */
    double lw, lt; /* adjoint seeds on wsum and tlast */
    int32_t lt_pending; /* adjoint: the tlast seed is still to be applied */
    int32_t pad;
} EdmdDeriv;

typedef struct EdmdTraceRec {
    int32_t type, i, j, valid;
    double t;
} EdmdTraceRec;

typedef struct EdmdThermoRow {
    double t;
    uint64_t ncol;
    double e_per_n, pressure, pxx, pyy, pxy, pyx;
    double pwall;
    double pmid, dp_lr, pack_left, pack_right, temp_left, temp_right, temp_solute;
} EdmdThermoRow;

typedef struct __attribute__((aligned(64))) EdmdPart {
    double x, y, vx, vy, tl, rad, m;
    uint32_t coll;
    int16_t cx, cy;
} EdmdPart;

typedef struct __attribute__((aligned(32))) EdmdEv {
    double tcoll, tcross;
    int32_t partner;
    uint32_t cact;
    int32_t dir, pad;
} EdmdEv;

typedef struct __attribute__((aligned(16))) EdmdNode {
    double key;  /* sanitised: NaN -> +inf */
    int32_t idx; /* leaf holding the key; the event's rank is lrank[idx] */
    int32_t aux; /* that leaf's event: partner (>= 0) or -direction */
} EdmdNode;

/* Kernel sources (which define EDMD_DEVICE_ARRAYS) see every array as
 * global memory (address space 1) in their device pass, so that the
 * compiler emits global rather than flat memory instructions: flat accesses
 * are counted against the LDS counter too, and waiting for one then waits
 * for every LDS operation in flight.  Host code sees plain pointers; the
 * layout is the same. */
#if defined(__HIP_DEVICE_COMPILE__) && defined(EDMD_DEVICE_ARRAYS)
#define EDMD_GLOBAL __attribute__((address_space(1)))
#else
#define EDMD_GLOBAL
#endif

typedef struct EdmdArrays {
    EdmdPart EDMD_GLOBAL* p;
    EdmdEv EDMD_GLOBAL* e;
    EdmdNode EDMD_GLOBAL* leaf;
    EdmdNode EDMD_GLOBAL* lev[EDMD_MAX_LEVELS]; /* lev[l] = level l+1 nodes */
    int32_t EDMD_GLOBAL* img;                   /* periodic image counters, 2 per particle */
    int32_t EDMD_GLOBAL* ccnt;
    int32_t EDMD_GLOBAL* cslot;
    EdmdState EDMD_GLOBAL* st;
    EdmdTraceRec EDMD_GLOBAL* trace;
    EdmdThermoRow EDMD_GLOBAL* thermo;
    double EDMD_GLOBAL* vr;          /* radius growth rate per particle (growth starts) */
    int64_t EDMD_GLOBAL* erank;      /* ranks of the slots: [2p] collision, [2p + 1] crossing */
    int64_t EDMD_GLOBAL* lrank;      /* rank of each leaf's key (the earlier slot) */
    EdmdTapeRec EDMD_GLOBAL* tape;   /* tape_cap records */
    double EDMD_GLOBAL* dv;          /* tangent or adjoint vector, 4 per particle */
    double EDMD_GLOBAL* dtl;         /* the time each particle's dv refers to */
    double EDMD_GLOBAL* otl;         /* particle local times when taping began */
    EdmdDeriv EDMD_GLOBAL* dr;
    int32_t EDMD_GLOBAL* ptype;      /* disk types (mid wall) */
    uint32_t EDMD_GLOBAL* mt;        /* MT19937 state (624 words) and index (thermal walls) */
} EdmdArrays;

typedef struct EdmdSys {
    EdmdConst c;
    EdmdArrays a;
} EdmdSys;

typedef struct EdmdRunArgs {
    int64_t max_events; /* absolute bound on the system's event counter */
    uint64_t max_coll;  /* absolute bound on the valid-collision counter */
    double tmax;        /* the reference's loop bound: run while t <= tmax */
    int32_t stop_mask;  /* stop after handling an event of kind k when bit k is set */
    int32_t nsys;       /* systems in the launch (workgroups holding several) */
} EdmdRunArgs;

/* Next pending event as the reference's findNextEvent would return it. */
typedef struct EdmdPeek {
    double t;
    int32_t type, i, j;
    uint32_t coll_actual;
} EdmdPeek;

#endif

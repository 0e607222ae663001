/*
 * libedmd_hip.so — C ABI of the MI355X event-driven hard-disk engine.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (EDMD_OK = 0) and never exits the process; edmd_last_error() describes the
 * last failure of the calling thread.  Host arrays passed in are copied;
 * particle and event state lives on the GPU.  One host thread per handle;
 * handles are independent; each handle owns one HIP stream.
 *
 * The reference keeps its engine in global state behind void functions
 * (EDMD.h:78-193, EDMD.c:80-495).  This header replaces that path twice:
 *   - a handle API (edmd_ctx for one system, edmd_batch for many
 *     independent replicas), and
 *   - same-named compatibility entry points on a default context
 *     (include/edmd_compat.h).
 */
#ifndef EDMD_H
#define EDMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define EDMD_OK 0
#define EDMD_ERR_OVERLAP 3        /* reference: saveTXT + exit(3), EDMD.c:2708-2717 */
#define EDMD_ERR_IO 5             /* reference parser: exit(1), parser.c:32-35 */
#define EDMD_ERR_CELL_OVERFLOW 7  /* a cell list exceeded its capacity */
#define EDMD_ERR_TAPE_FULL 8      /* more collisions than the tape holds */
#define EDMD_ERR_NOMEM 12
#define EDMD_ERR_NODEVICE 19
#define EDMD_ERR_ARG 22
#define EDMD_ERR_HIP 30
#define EDMD_ERR_STATE 31

/* ---- event kinds: the reference's enum event, EDMD.h:21-34 -------------- */
#define EDMD_EV_CELLCROSS 0
#define EDMD_EV_COLLISION 1
#define EDMD_EV_SCREENSHOT 2
#define EDMD_EV_THERMO 3
#define EDMD_EV_ADDINGNOISE 4
#define EDMD_EV_WALL 5 /* trace j / peek j: 0 = wall normal to x, 1 = normal to y */
#define EDMD_EV_GROWSTOP 9

/* ---- why a run returned ------------------------------------------------- */
#define EDMD_STOP_NONE 0
#define EDMD_STOP_EVENTS 1
#define EDMD_STOP_COLLISIONS 2
#define EDMD_STOP_TIME 3
#define EDMD_STOP_ERROR 4
#define EDMD_STOP_EMPTY 5
#define EDMD_STOP_SPECIAL 6 /* stopped after an event kind listed in a stop mask */
#define EDMD_STOP_GROWN 7   /* growth ended (internal: the run continues) */

/* One system: N disks in a periodic Lx x Ly box.  Replaces the state that
 * constantInit/particlesInit build (EDMD.c:745-1219, 1356-1451).  rad is a
 * radius (contact distance 2 sqrt(r_i r_j), EDMD.c:35-43). */
typedef struct edmd_system_desc {
    int n;
    double Lx, Ly;
    const double *x, *y, *vx, *vy, *rad, *m; /* host arrays of n values */
    double res;             /* restitution coefficient, EDMD.c:372 (-r) */
    double first_screen;    /* first screenshot time, EDMD.c:244 */
    double dtime;           /* screenshot period, EDMD.c:245 (-D) */
    double first_thermo;    /* first thermo time (an extra 1e-9 is added), EDMD.c:248, 1970 */
    double dtime_thermo;    /* thermo period, EDMD.c:247 (-o) */
    int noise;              /* 0 none, 2 velocity-rescale thermostat, EDMD.c:4880-4915 */
    double dtnoise;         /* thermostat period, EDMD.c:386 */
    double temperature;     /* T: thermostat target E/N and thermal walls' scale, EDMD.c:383
                               (the reference's default is 0.05, this one's 1) */
    int cell_capacity;      /* max disks per cell, 0 = derived from the radii */
    int64_t trace_capacity; /* device event-trace records kept, 0 = none */
    int thermo_capacity;    /* thermo rows kept, 0 = 4096 */
    /* Growth start (the reference's default start, EDMD.c:1691-1810,
     * 1957-1981): vr > 0 starts every disk at radius 0 growing at
     * vr * rad[i] (rad[] holds the target radii) until GROWSTOP at
     * t = 1/vr, when stopGrow (EDMD.c:4740-4785) sets E/N = einit and the
     * Normal dynamics begin; screenshot and thermo times below 1/vr are
     * shifted by 1/vr as EDMD.c:1959-1964 does. */
    double vr;
    double einit;           /* Einit, EDMD.c:369 (-E); default 1 */
    /* Hard walls (addWallx / addWally of the reference's graphics build,
     * EDMD.c:267-268; its command-line build fixes them at 0): walls at
     * x = 0, Lx and/or y = 0, Ly, no minimum image across them
     * (EDMD.c:5896-5913).  A hit (EDMD.c:3384-3430) reverses the normal
     * velocity times res_wall (resW, default 1) and scales the tangential
     * one by res_tan (resT, default 0), and counts as a collision.  During
     * a growth start the wall slot closes the disk's surface at v -/+ vr on
     * the wall (EDMD.c:3118-3153) and a hit turns the normal velocity away
     * with 2 vr added, no momentum counted (doTheWallGrow, EDMD.c:
     * 3333-3382); after GROWSTOP the walls above apply.  Taped hits enter
     * the derivatives (edmd_tape_rec). */
    int wall_x, wall_y;
    double res_wall, res_tan;
    /* Uniform field along y (addField / field of the graphics build,
     * EDMD.c:265, 366, -f): free flight parabolic in y (EDMD.c:4957-4963),
     * every neighbour flown to the current time when its collision time is
     * taken (EDMD.c:2664-2666), crossings and y-wall hits solving the
     * quadratic (EDMD.c:2428-2442, 2846-2860).  field < 0 (default -0.02).
     * During a growth start the Grow dispatch ignores it (straight flights
     * and crossings, EDMD.c:2343-2400, 4992-5007) until GROWSTOP.  Taped
     * runs differentiate through it (edmd_tape_rec's g). */
    int add_field;
    double field;
    /* Damping (damping / gamm of the graphics build, EDMD.c:294, 382):
     * velocities decay as exp(-gamma t) in flight (EDMD.c:4965-4975), a
     * predicted time t of the undamped motion becomes -log(1 - gamma t) /
     * gamma, never when gamma t >= 1 (logTime, EDMD.c:5709-5721), and
     * neighbours are flown when their collision time is taken.  gamma > 0
     * (default 0.02).  exp and log are this library's own (a Taylor series
     * and an atanh series with Cody-Waite reduction), not the C library's.
     * With the field as well, the flight is the field's alone (its branch
     * of freeFlyNormal comes first, EDMD.c:4957-4975) while logTime still
     * converts the straight-line times of x crossings, x walls and
     * collisions; that pairing of damped times with undamped flights can
     * end in overlaps and NaN states, as in the reference.  During a growth
     * start damping waits for GROWSTOP, as the field does.  Taped runs
     * differentiate through it, not together with the field. */
    int damping;
    double gamma;
    /* Circular wall (addCircularWall of the graphics build, EDMD.c:270): a
     * circle of radius Lx/2 about (Lx/2, Lx/2) (the reference takes Lx/2
     * for both coordinates; boxes are square); a disk's centre meets the
     * circle of radius Lx/2 - rad (EDMD.c:2938-2954; with growth, its
     * surface at |v| - vr, EDMD.c:3158-3172), for every disk, replacing the
     * other walls' times; a hit reflects the velocity about the normal at
     * the angle atan2 of the disk's position (EDMD.c:3441-3454, growth
     * EDMD.c:3358-3376), this library's own atan2 / sin / cos (within 2 ulp
     * of the C library's, the same steps in the oracle).  No minimum
     * image; thermo rows take the area pi (Lx/2)^2 and the wall perimeter
     * pi Lx; at GROWSTOP the angular momentum about the centre is removed
     * instead of the momentum (EDMD.c:5723-5764).  Under the field the time
     * is the smallest positive root of the quartic of the parabolic flight
     * (EDMD.c:2918-2935, quartic.c's method with the library's own acos,
     * cos and cube root).  The disks must start inside the circle.  Not
     * with taping. */
    int wall_circle;
    /* Mid wall (addMidWall of the graphics build, EDMD.c:269): a wall at
     * x = Lx/2 for the disks of type 1 only (the solute; type 0, the
     * solvent, passes through), met from within 4 of it (EDMD.c:2897-2915;
     * with growth, at v +/- vr, EDMD.c:3175-3195); a hit is an x-wall hit
     * whose momentum also goes to dpLeft, dpMid or dpRight by the disk's
     * position (EDMD.c:3393-3409); thermo rows carry the mid-wall columns
     * of EDMD.c:5453-5477.  type: one int per disk (NULL: all 0), copied;
     * sizeratio: the solvent's packing factor in those columns (EDMD.c:83,
     * default 0.4).  Taped hits enter the derivatives like the side
     * walls'. */
    int mid_wall;
    const int32_t* type;
    double sizeratio;
    /* Thermal walls (thermoWall of the graphics build, EDMD.c:280): a hit
     * of a flat wall sets the normal velocity to -sign(v_n) drand(0, T/m)
     * and keeps the tangential one (EDMD.c:3411-3426); a hit of the
     * circular wall sets v = -n drand(0, T), one draw per component
     * (EDMD.c:3446-3452); T = temperature.  drand(a, b) = a + (b - a)
     * k / 2^32 for the next 32-bit output k of the system's own MT19937
     * stream (mersenne.c), seeded by init_genrand(rng_seed).  Needs walls;
     * not with taping. */
    int thermal_wall;
    /* Shear walls (addShearWall, EDMD.c:283): after a hit of a wall normal
     * to y the disk's x velocity is +0.1 at the top (y + 2 rad > Ly) and
     * -0.1 at the bottom (EDMD.c:3429-3435; the delta model's kick of
     * EDMD.c:3436-3438 is not built).  Needs wall_y; not with taping. */
    int shear_wall;
    uint32_t rng_seed; /* default 5489, mersenne.c's seed when none is given */
} edmd_system_desc;

/* A collision slot holding a wall event reports partner -5 - xy
 * (edmd_batch_get_events / edmd_get_events). */
#define EDMD_WALL_PARTNER(xy) (-5 - (xy))

typedef struct edmd_trace_rec {
    int32_t type, i, j, valid; /* crossing: j = direction 1..4; collision: j = partner */
    double t;
} edmd_trace_rec;

typedef struct edmd_thermo_row {
    double t;
    uint64_t ncol;
    double e_per_n, pressure, pxx, pyy, pxy, pyx; /* saveThermo, EDMD.c:5232-5333 */
    double pwall; /* wall pressure dp / (dt * perimeter), EDMD.c:5428-5451; 0 without walls */
    /* mid wall (EDMD.c:5453-5477), 0 without it: dpMid / (dt Ly),
     * -(dpRight - dpLeft) / (dt Ly), the solvent's packing pi n sizeratio^2
     * / (Lx/2 Ly) left and right of x = Lx/2, its mean kinetic energy left
     * and right, the solute's */
    double pmid, dp_lr, pack_left, pack_right, temp_left, temp_right, temp_solute;
} edmd_thermo_row;

typedef struct edmd_observables {
    double t;
    uint64_t ncol, ncross, nevents;
    double vir_xx, vir_yy, vir_xy, vir_yx; /* collTermX/Y/XY/YX since the last thermo */
    double E, px, py;                      /* physicalQ, EDMD.c:5968-6013 */
    int32_t status, stop_reason, err_i, err_j;
} edmd_observables;

void edmd_default_desc(edmd_system_desc* d);
const char* edmd_last_error(void);
int edmd_device_count(int* n);

/* ---- single system (plain path) ----------------------------------------- */
typedef struct edmd_ctx edmd_ctx;

/* Builds the cell grid and cell lists (boxConstantHelper + cellListInit,
 * EDMD.c:679-715, 1906-1920), uploads the state and runs the initial
 * prediction sweep and calendar build (eventListInit, EDMD.c:1935-2012). */
int edmd_ctx_create(const edmd_system_desc* desc, edmd_ctx** out);
int edmd_ctx_destroy(edmd_ctx* ctx);
/* The load path (EDMD.c:987-995, 1387-1451; parser.c): the last frame of a
 * LAMMPS dump as the initial state, the other fields from params (NULL =
 * edmd_default_desc). */
int edmd_ctx_create_from_dump(const char* path, const edmd_system_desc* params, edmd_ctx** out);
/* edmd_ctx_create_from_dump with the default parameters. */
int edmd_load_dump(const char* path, edmd_ctx** out);
/* Full prediction sweep at the current time (kernel k_sweep_tiled): free-flies
 * every particle to the current time first, then re-predicts every
 * crossing and collision slot and rebuilds the calendar (the pattern of
 * EDMD.c:2007-2012 and 4909-4915). */
int edmd_sweep_predict(edmd_ctx* ctx);
/* Event loop (main, EDMD.c:523-610): handles events while t <= tmax, until
 * the system's event counter reaches max_events or its valid-collision
 * counter reaches max_collisions (both absolute counts). */
int edmd_run(edmd_ctx* ctx, int64_t max_events, uint64_t max_collisions, double tmax);
/* edmd_run handing every handled event to cb in order, in chunks of at most
 * desc.trace_capacity records (which must be > 0); a nonzero return from cb
 * stops the run after the current chunk.  The reference prints or records
 * events in its loop (EDMD.c:534-610); this is that hook. */
typedef int (*edmd_trace_cb)(const edmd_trace_rec* recs, int64_t n, void* user);
int edmd_run_traced(edmd_ctx* ctx, int64_t max_events, uint64_t max_collisions, double tmax,
                    edmd_trace_cb cb, void* user);
int edmd_get_state(edmd_ctx* ctx, double* x, double* y, double* vx, double* vy,
                   double* t_local, uint32_t* coll, int32_t* cx, int32_t* cy);
int edmd_get_events(edmd_ctx* ctx, double* t_cross, int32_t* dir, double* t_coll,
                    int32_t* partner, uint32_t* coll_actual);
int edmd_get_observables(edmd_ctx* ctx, edmd_observables* out);
/* Copies up to cap trace records kept since the last call and clears them. */
int edmd_get_trace(edmd_ctx* ctx, edmd_trace_rec* out, int64_t cap, int64_t* n);
int edmd_get_thermo(edmd_ctx* ctx, edmd_thermo_row* out, int cap, int* n);
int edmd_get_grid(edmd_ctx* ctx, int* ncx, int* ncy, int* cap, double* csx, double* csy);
int edmd_get_box(edmd_ctx* ctx, int* n, double* Lx, double* Ly);
void* edmd_stream(edmd_ctx* ctx);

/* Step-level entry points of the single-system path, for drivers that
 * interleave their own work between events (the reference's GUI loop,
 * graphics.c:980-1016) and for tests. */
/* As edmd_run, but also stops right after handling an event whose kind k
 * has bit (1 << k) set in stop_mask; reports why it stopped and the kind of
 * the last event handled. */
int edmd_run_until_kind(edmd_ctx* ctx, int64_t max_events, uint64_t max_collisions, double tmax,
                        int stop_mask, int* stop_reason, int* last_kind);
/* findNextEvent (EDMD.c:2313-2327): the pending event that would be handled
 * next, without handling it.  For a crossing j is the direction code. */
int edmd_peek_event(edmd_ctx* ctx, double* t, int* kind, int* i, int* j, uint32_t* coll_actual);
/* (edmd_peek_event and edmd_collision_time use a result buffer allocated
 * with the context, no allocation per call.) */
/* collisionTimeNormal(p_i, p_j) (EDMD.c:2661-2723) at the current time:
 * relative time, EDMD_NEVER-valued 1e26 when the pair does not meet;
 * EDMD_ERR_OVERLAP when the pair overlaps by more than the tolerance. */
int edmd_collision_time(edmd_ctx* ctx, int i, int j, double* dt);
/* crossingEvent(i) (which & 1) and/or collisionEvent(i) (which & 2) at the
 * current time, EDMD.c:2405-2496, 2829-3093; i must be synchronised. */
int edmd_repredict(edmd_ctx* ctx, int i, int which);
/* freeFlyNormal(p_i) to the current time, EDMD.c:4954-4990. */
int edmd_free_fly(edmd_ctx* ctx, int i);
/* addEventToQueue(node*) (EDMD.c:2144-2170) of the reference's eventList
 * slot of (kind, i), filled as the reference's callers fill it:
 * EDMD_EV_COLLISION: particle i's collision slot, partner j, partner counter
 * coll_actual (addCollisionEvent, EDMD.c:3286-3297, sets it to the
 * partner's counter); EDMD_EV_WALL: that slot as a wall hit on side j
 * (addWallEvent, EDMD.c:3321-3330); EDMD_EV_CELLCROSS: i's crossing slot,
 * direction j = 1..4 (addCrossingEvent, EDMD.c:2487-2496);
 * EDMD_EV_SCREENSHOT / _THERMO / _ADDINGNOISE: that scheduled event (i, j
 * unused).  The event is filed at absolute time t and takes the next
 * insertion rank, so equal times pop in the reference's order.  A slot
 * filed twice keeps the second filing. */
int edmd_queue_add(edmd_ctx* ctx, int kind, int i, int j, double t, uint32_t coll_actual);
/* addCollisionEvent(i, j, t) (EDMD.c:3286-3297): edmd_queue_add of i's
 * collision slot with j's current collision counter. */
int edmd_add_collision_event(edmd_ctx* ctx, int i, int j, double t);
/* removeEventFromQueue(node*) (EDMD.c:2209-2221) of the slot of (kind, i):
 * its event never pops until the slot is filed again. */
int edmd_queue_remove(edmd_ctx* ctx, int kind, int i);

/* ---- batched replicas ---------------------------------------------------- */
typedef struct edmd_batch edmd_batch;

/* nsys independent systems of any sizes, one wavefront each in one
 * persistent launch. */
int edmd_batch_create(const edmd_system_desc* descs, int nsys, edmd_batch** out);
int edmd_batch_destroy(edmd_batch* b);
/* Every system runs until its own event counter reaches max_events or its
 * collision counter reaches max_collisions, or t > tmax. */
int edmd_batch_run(edmd_batch* b, int64_t max_events, uint64_t max_collisions, double tmax);
int edmd_batch_sweep_predict(edmd_batch* b);
/* Epochs (the reference has none: its loop is one process per system).
 * nepochs epochs in one hipGraph launch: each epoch lets every system handle
 * up to events_per_epoch more events and collisions_per_epoch more valid
 * collisions (while t <= tmax), then records its counters in
 * rows[epoch * nsys + sys] (may be NULL).  A GROWSTOP inside an epoch
 * resets the system's collision counter, as stopGrow does. */
typedef struct edmd_epoch_row {
    double t;
    uint64_t ncol, ncross, nevents;
    int32_t status, stop_reason;
} edmd_epoch_row;
int edmd_batch_run_epochs(edmd_batch* b, int nepochs, int64_t events_per_epoch,
                          uint64_t collisions_per_epoch, double tmax, edmd_epoch_row* rows);
int edmd_batch_get_observables(edmd_batch* b, int sys, edmd_observables* out);
int edmd_batch_get_state(edmd_batch* b, int sys, double* x, double* y, double* vx, double* vy,
                         double* t_local, uint32_t* coll, int32_t* cx, int32_t* cy);
int edmd_batch_get_events(edmd_batch* b, int sys, double* t_coll, int32_t* partner,
                          double* t_cross, int32_t* dir);
int edmd_batch_get_trace(edmd_batch* b, int sys, edmd_trace_rec* out, int64_t cap, int64_t* n);
/* Diagnostic, no counterpart in the reference (whose queue is a pointer
 * tree): the event calendar of one system, for consistency checks.  Nodes
 * level by level, the leaves (one per disk) first; each node holds the
 * earliest key below it, the leaf it came from and that leaf's event
 * (partner >= 0 or -direction).  slot_rank (2 per disk: collision,
 * crossing) and leaf_rank (1 per disk) are the insertion ranks that order
 * equal times as the reference's queue does.  Any output may be NULL;
 * *nnodes receives the node count, cap bounds key/idx/aux. */
int edmd_batch_get_calendar(edmd_batch* b, int sys, double* key, int32_t* idx, int32_t* aux,
                            int64_t cap, int64_t* nnodes, int64_t* slot_rank,
                            int64_t* leaf_rank);
int edmd_batch_get_thermo(edmd_batch* b, int sys, edmd_thermo_row* out, int cap, int* n);
int edmd_batch_size(edmd_batch* b, int* nsys);
void* edmd_batch_stream(edmd_batch* b);
/* Uses the caller's HIP stream (hipStream_t) for all later work. */
int edmd_batch_set_stream(edmd_batch* b, void* stream);

/* Where the event loop keeps each system's arrays while it runs.  No
 * counterpart in the reference (one system in host memory): the batched
 * form of its step loop, mainLoop (EDMD.c:5595-5630).  EDMD_RES_HBM: in
 * device memory, one wavefront per system.  EDMD_RES_LDS: each system's
 * particles, event slots, calendar, ranks and cell lists are copied into
 * its workgroup's LDS when a run launches and back when it stops (calendar
 * depth <= 2 and the largest system's arrays within one workgroup's LDS,
 * EDMD_ERR_ARG otherwise).  EDMD_RES_AUTO (the default): LDS when every
 * system fits and the whole batch is resident on the device at once, HBM
 * otherwise.  Results are identical in every mode. */
#define EDMD_RES_AUTO 0
#define EDMD_RES_HBM 1
#define EDMD_RES_LDS 2
int edmd_batch_set_residency(edmd_batch* b, int mode);
/* The mode in use (*in_lds 1 = LDS), the dynamic LDS per workgroup the LDS
 * mode takes, and the workgroups the device holds at once in that mode
 * (0 when the systems do not fit).  Any output may be NULL. */
int edmd_batch_get_residency(edmd_batch* b, int* in_lds, int64_t* lds_bytes,
                             int64_t* lds_slots);

/* Tile shape of the prediction sweep (collisionEventNormal and
 * crossingEventNormal for every disk, as eventListInit runs them,
 * EDMD.c:2007-2012): shape s tiles the cell grid into tile_x x tile_y
 * cells per workgroup of `threads` threads.  Every shape gives identical
 * predictions; the default is chosen from the disks per cell. */
#define EDMD_TUNE_SWEEP_SHAPES 8
int edmd_batch_set_sweep_shape(edmd_batch* b, int shape);
int edmd_batch_get_sweep_shape(edmd_batch* b, int* shape, int* tile_x, int* tile_y, int* threads);
/* The full sweep's conservative root screen (no counterpart in the
 * reference; the FP32-or-cheaper far-pair screen of the N = 2^20
 * configuration): mode 1 skips the square root and division of a marked
 * candidate whose root provably exceeds the best found so far for its
 * disk (with q = -b: d v2 - 2^-19 q^2 > 2 q v2 dt (1 + 2^-40), a bound that
 * covers the rounding of the reference's formula), so the predictions are
 * unchanged; mode 2 also counts, per system, the marked candidates seen,
 * the roots solved and the skipped roots the minimum would have taken
 * (misses, 0 by construction; the skipped roots are solved for that), which
 * edmd_batch_screen_stats sums over the batch; 0 (default) off.  Not under
 * damping.  Setting a mode clears the counters. */
int edmd_batch_set_screen(edmd_batch* b, int mode);
int edmd_batch_screen_stats(edmd_batch* b, uint64_t* marks, uint64_t* solved, uint64_t* misses);

/* Systems per workgroup of the event loop with arrays in HBM, no
 * counterpart in the reference: 1 (default), 2 or 4, one wavefront per
 * system either way.  More than one only for calendars of depth <= 2 (up to
 * 4096 disks) without walls, field or damping; the growing systems' launch
 * always takes one.  Results do not depend on it. */
int edmd_batch_set_loop_waves(edmd_batch* b, int systems_per_group);
int edmd_batch_get_loop_waves(edmd_batch* b, int* systems_per_group);

/* Autotuner, no counterpart in the reference: times sweep_reps sweeps in
 * every tile shape and a trial of trial_events events per system in each
 * event-loop configuration -- arrays in HBM with 1, 2 and 4 systems per
 * workgroup (2 and 4 where edmd_batch_set_loop_waves allows them) and, when
 * every system fits LDS, arrays in LDS -- then keeps the fastest sweep
 * shape (the one in use stays unless another is more than 3% faster) and
 * the fastest loop configuration.  The batch's state is saved before and
 * restored after every trial, so results are unchanged.  sweep_reps = 0 or
 * trial_events = 0 skips that part.  Not while taping. */
typedef struct edmd_tune_report {
    double sweep_ms[EDMD_TUNE_SWEEP_SHAPES]; /* per shape, < 0 when not timed */
    int sweep_shape;                         /* the shape kept */
    int in_lds;                              /* the residency kept (1 = LDS) */
    double loop_ms_hbm, loop_ms_lds;         /* trial times (HBM: one system per
                                                workgroup), < 0 when not run */
    double loop_ms_packed[2];                /* HBM, 2 and 4 systems per workgroup */
    int loop_systems_per_group;              /* the systems per workgroup kept */
} edmd_tune_report;
int edmd_batch_autotune(edmd_batch* b, int sweep_reps, int64_t trial_events,
                        edmd_tune_report* out);

/* ---- initial states and files -------------------------------------------- */
/* Seeded lattice start: lattice 0 = square (first n sites of a
 * ceil(sqrt n)^2 grid), 1 = triangular (the reference's Hex layout,
 * EDMD.c:1024-1049, 1582-1637; n must factor as nx*ny with ny even),
 * 2 = triangular stacking fitted to a square box for any n (ny even rows of
 * ceil(n/ny) sites, ny chosen for the largest nearest-neighbour distance);
 * a lattice whose spacing is below the diameter is refused;
 * Gaussian velocities with zero total momentum scaled to E/N = temperature
 * (normalizePhysicalQ, EDMD.c:5723-5764).  Writes n values to each array and
 * the box lengths. */
int edmd_init_lattice(int n, double phi, double rad, int lattice, uint64_t seed,
                      double temperature, double* x, double* y, double* vx, double* vy,
                      double* Lx, double* Ly);
/* Seeded random start for a growth run (the reference's default start,
 * EDMD.c:1690-1810, with this library's generator): centres uniform in a
 * square box sized so that the target radii fill the fraction phi; target
 * radii rad for the first n - (int)(n * fsmall) disks and rad * sizeratio
 * for the rest, masses 1 (EDMD.c:1776-1789); or, with poly > 0, radii
 * rad * d with d ~ N(1, poly) truncated to [1 - 3 poly, 1 + 3 poly] and
 * masses d^3; velocity components Gaussian with variance e_per / m (the
 * reference's EinitGrow = 0.05).  Pair it with desc.vr > 0 (the disks start
 * at radius 0).  Writes n values to each array and the box lengths. */
int edmd_init_random(int n, double phi, double rad, double poly, double sizeratio, double fsmall,
                     uint64_t seed, double e_per, double* x, double* y, double* vx, double* vy,
                     double* r, double* m, double* Lx, double* Ly);
/* optimizeGrowConstant (EDMD.c:5766-5800) for a periodic box: the growth
 * rate the reference picks for packing fraction phi. */
double edmd_growth_rate(double phi);
/* LAMMPS dump reader (the load path, EDMD.c:987-995, 1387-1451; parser.c):
 * last frame, box from row 0 column 1 of BOX, columns by name prefix. */
int edmd_dump_info(const char* path, int* n, double* Lx, double* Ly);
int edmd_dump_read(const char* path, int n, double* x, double* y, double* vx, double* vy,
                   double* rad, double* m);
/* Dump writer (saveTXT layout, EDMD.c:5012-5153); full_precision != 0
 * prints every value with 17 significant digits. */
int edmd_dump_write(const char* path, int append, double t, int n, double Lx, double Ly,
                    const double* x, const double* y, const double* vx, const double* vy,
                    const double* rad, const double* m, int full_precision);

/* Current radius of every disk of system sys (they change while growing). */
int edmd_batch_get_radii(edmd_batch* b, int sys, double* rad);
int edmd_get_radii(edmd_ctx* ctx, double* rad);

/* ---- observables ----------------------------------------------------------- */
/* calculate_pcf (pcf.c:16-75) on the GPU: g(r) of n positions with the
 * minimum image, bins of width dr up to (int)(max_r/dr)*dr, normalised by
 * 2 pi r dr * (n / (Lx Ly)) * n.  On entry *nbins is the capacity of r and
 * g; on return the bin count.  Pair counts are integers, so the result is
 * the reference's bit for bit. */
int edmd_pcf(int n, const double* x, const double* y, double Lx, double Ly, double dr,
             double max_r, double* r, double* g, int* nbins);
/* The same for system `sys` of a batch: its stored positions (as the
 * reference's thermo-time g(r) uses them) or, with extrapolate != 0, all
 * disks free-flown to the current time without changing the state. */
int edmd_batch_pcf(edmd_batch* b, int sys, int extrapolate, double dr, double max_r, double* r,
                   double* g, int* nbins);
/* edmd_pcf without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  x, y (n <= 2^22
 * positions) and r, g are DEVICE pointers the caller owns; the pair counts
 * live in a fixed per-GPU array of the library (at most
 * EDMD_PCF_NOALLOC_MAX_BINS bins, 256 KiB per GPU, held until the process
 * ends; more bins are an EDMD_ERR_ARG error).  Runs on the null stream of
 * the current device and waits for it.  Every call on one GPU shares that
 * array: calls that overlap on one GPU race.  The results equal edmd_pcf's
 * bit for bit. */
#define EDMD_PCF_NOALLOC_MAX_BINS 32768
int edmd_pcf_noalloc(int n, const double* x, const double* y, double Lx, double Ly, double dr,
                     double max_r, double* r, double* g, int* nbins);

/* ---- derivatives of the event map ------------------------------------------ */
/* The reference computes no derivatives; these entry points are new.  For
 * the sequence of valid collisions a run produced, they differentiate the
 * map from the state when taping began (x, y, vx, vy of every disk at its
 * local time) to
 *   W      = sum over taped collisions of m_i m_j ff |r_j - r_i|^2 (the
 *            collision virial XX + YY that the reference accumulates,
 *            EDMD.c:3804-3813, so the time-averaged virial is
 *            W / (t_last - t_origin)),
 *   t_last = time of the last taped collision,
 *   and the state of every disk free-flown to the final time,
 * holding the event sequence fixed (the map is smooth while it is).
 * Systems must not use the velocity-rescale thermostat, the circular wall,
 * thermal or shear walls, or the field together with damping; flat walls
 * (wall_x, wall_y, mid_wall), the field and damping are taped.
 *
 * One valid collision: time, contact vector a = r_j - r_i (minimum image),
 * relative velocity c = v_j - v_i before the collision, impulse factor ff
 * (v_i += ff m_j a, v_j -= ff m_i a).  One flat wall's hit: j = -1 - xy
 * (-1: a wall normal to x), ax / ay = the normal / tangential velocity
 * before the hit, cx / cy = res_wall / res_tan, ff = the contact
 * coordinate (x or y of the disk's centre at the hit); no virial term. */
typedef struct edmd_tape_rec {
    double t, ax, ay, cx, cy, ff;
    int32_t i, j;
    double g; /* wall records: the field along y at the hit (0 without it) */
} edmd_tape_rec;

/* Starts taping at the current state.  checkpoint_every = K > 0: the run
 * keeps a full checkpoint of every system every K events and the tape of
 * one K-event segment; the tangent and adjoint passes recompute each
 * segment's tape from its checkpoint.  K = 0: the whole tape is kept,
 * tape_capacity records (collisions and wall hits) per system
 * (EDMD_ERR_TAPE_FULL beyond). */
int edmd_batch_deriv_begin(edmd_batch* b, int64_t checkpoint_every, int64_t tape_capacity);
/* edmd_batch_run with taping (and checkpoints); may be called repeatedly. */
int edmd_batch_deriv_run(edmd_batch* b, int64_t max_events, uint64_t max_collisions, double tmax);
/* Tangent (forward) pass.  tan_in / tan_out hold 4 values (dx dy dvx dvy)
 * per disk, systems in batch order; tan_out is the tangent of the final
 * state at the systems' current time.  dw[s], dtlast[s] (may be NULL):
 * derivatives of W and t_last of system s. */
int edmd_batch_tangent(edmd_batch* b, const double* tan_in, double* tan_out, double* dw,
                       double* dtlast);
/* Adjoint (reverse) pass: gradient with respect to the taping-origin state
 * (4 per disk) of  sum_s lw[s] W_s + lt[s] t_last,s + <lam_end, final state>.
 * lw, lt: one per system (NULL = 0); lam_end: 4 per disk or NULL. */
int edmd_batch_adjoint(edmd_batch* b, const double* lw, const double* lt, const double* lam_end,
                       double* grad);
/* W, t_last and the number of tape records currently held for system sys. */
int edmd_batch_deriv_values(edmd_batch* b, int sys, double* w, double* tlast, int64_t* records);
/* Copies up to cap records of system sys's tape (the whole tape, or the
 * last segment with checkpoints); *n = records held. */
int edmd_batch_get_tape(edmd_batch* b, int sys, edmd_tape_rec* out, int64_t cap, int64_t* n);
/* Stops taping and frees tapes and checkpoints. */
int edmd_batch_deriv_end(edmd_batch* b);

/* The same for the single system of a context (no counterpart in the
 * reference, which has no derivatives): begin / run / tangent / gradient /
 * values / end, with the arguments of the batch calls for one system
 * (lw and lt are scalars here). */
int edmd_adjoint_begin(edmd_ctx* ctx, int64_t checkpoint_every, int64_t tape_capacity);
int edmd_adjoint_run(edmd_ctx* ctx, int64_t max_events, uint64_t max_collisions, double tmax);
int edmd_adjoint_tangent(edmd_ctx* ctx, const double* tan_in, double* tan_out, double* dw,
                         double* dtlast);
int edmd_adjoint_gradient(edmd_ctx* ctx, double lw, double lt, const double* lam_end,
                          double* grad);
int edmd_adjoint_values(edmd_ctx* ctx, double* w, double* tlast, int64_t* records);
int edmd_adjoint_end(edmd_ctx* ctx);

/* ---- diagnostics ----------------------------------------------------------- */
/* q[k] = a[k] / b[k] and s[k] = sqrt(a[k]) evaluated on the GPU, for the
 * IEEE-rounding tests of the FP64 operations the engine relies on. */
int edmd_fp64_check(int n, const double* a, const double* b, double* q, double* s);
/* Average duration (ms) of `reps` prediction sweeps of every system of the
 * batch, timed with HIP events on the batch's stream.  The sweeps re-predict
 * every event slot; the batch's state is saved before and restored after
 * them (one device allocation of the state's size for the call), so a run
 * that follows continues as if the call had not been made. */
int edmd_batch_time_sweep(edmd_batch* b, int reps, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif

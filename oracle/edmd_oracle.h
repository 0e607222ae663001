/*
 * CPU restatement of the reference's event-driven hard-disk engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is what the GPU library is
 * compared against; it is never linked into, called by, or shipped with the
 * product.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.
 *
 * It restates, in its own structure (one context struct, integer node
 * indices, no globals), the Normal-mode path of the reference (src/EDMD.c,
 * built with -DG=0): cell grid and cell lists, crossing and collision
 * prediction, the BST + bucket calendar, the collision and crossing handlers,
 * free flight, the screenshot / thermo / velocity-rescale events and the
 * observables.  Built strict IEEE (-O2 -fno-fast-math -ffp-contract=off), it
 * reproduces the reference's strict build bit for bit; every function cites
 * the reference lines it follows.
 */
#ifndef EDMD_ORACLE_H
#define EDMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Event kinds, numbered as the reference's enum event (EDMD.h:21-34). */
enum { ORC_CELLCROSS = 0, ORC_COLLISION = 1, ORC_SCREENSHOT = 2, ORC_THERMO = 3,
       ORC_ADDINGNOISE = 4, ORC_WALL = 5, ORC_GROWSTOP = 9 };

/* A collision slot holding a wall event reports partner -5 - xy (xy = 0:
 * a wall normal to x, 1: normal to y), apart from the disks (>= 0) and the
 * crossing directions (1-4). */
#define ORC_WALL_PARTNER(xy) (-5 - (xy))

/* Status codes. */
enum { ORC_OK = 0, ORC_ERR_OVERLAP = 3, ORC_ERR_ARG = 22, ORC_ERR_NOMEM = 12 };

typedef struct {
    double res;          /* restitution coefficient (EDMD.c:372, -r) */
    double first_screen; /* firstScreen (EDMD.c:244) */
    double dtime;        /* screenshot period (EDMD.c:245, -D) */
    double first_thermo; /* firstThermo (EDMD.c:248) */
    double dtime_thermo; /* dtimeThermo (EDMD.c:247, -o) */
    int noise;           /* 0: none, 2: velocity rescale thermostat (EDMD.c:4880-4901) */
    double dtnoise;      /* thermostat period (EDMD.c:386) */
    double temperature;  /* thermostat target T (EDMD.c:379) */
    int bucket_count;    /* calendar bucket count, 0 = N (EDMD.c:712-714) */
    /* Growth start (EDMD.c:1691-1810, 1957-1981): vr > 0 grows every disk
     * from radius 0 at rate vr * rad[i] (rad[] holds the target radii) until
     * the GROWSTOP event at t = 1/vr, then stopGrow (EDMD.c:4740-4800)
     * rescales velocities to E/N = einit. */
    double vr;
    double einit;        /* Einit (EDMD.c:369, -E) */
    /* Hard walls (the reference's addWallx / addWally, EDMD.c:267-268, which
     * its command-line build fixes at 0, EDMD.c:290-291): walls at x = 0, Lx
     * and/or y = 0, Ly; the minimum image is off across them
     * (EDMD.c:5896-5913).  A wall hit reverses the normal velocity times
     * res_wall and scales the tangential one by res_tan (resW, resT,
     * EDMD.c:372-373, 3384-3430). */
    int wall_x, wall_y;
    double res_wall, res_tan;
    /* Uniform field along y (addField / field, EDMD.c:265, 366, -f; the
     * command-line build fixes addField at 0): free flight is parabolic in
     * y (EDMD.c:4957-4963), every neighbour is flown to the current time
     * when its collision time is taken (EDMD.c:2664-2666), crossings and
     * y-wall hits solve the quadratic (EDMD.c:2428-2442, 2846-2860). */
    int add_field;
    double field;        /* must be < 0, default -0.02 (EDMD.c:365-366) */
    /* Damping (damping / gamm, EDMD.c:294, 382; fixed at 0 in the
     * command-line build): velocities decay as exp(-gamma t) in flight
     * (EDMD.c:4965-4975), every predicted time t_lin of the undamped motion
     * becomes logTime(t_lin) = -log(1 - gamma t_lin) / gamma, never when
     * gamma t_lin >= 1 (EDMD.c:5709-5721), and neighbours are flown before
     * their collision time is taken (EDMD.c:2664-2666).  exp and log are
     * this oracle's own (orc_exp, orc_log), the same algorithm the GPU
     * library uses, not the C library's. */
    int damping;
    double gamma;        /* gamm, default 0.02 */
    /* Circular wall (addCircularWall, EDMD.c:270): a circle of radius
     * halfLx about (halfLx, halfLx) (the reference uses halfLx for both
     * coordinates); a disk's centre meets the circle of radius
     * halfLx - rad (EDMD.c:2938-2954, growth EDMD.c:3158-3172); a hit
     * reflects the velocity about the normal at atan2 of the disk's
     * position (EDMD.c:3441-3454, growth 3358-3376), with this oracle's
     * own atan2 / sin / cos; no minimum image (EDMD.c:5896-5913); the area
     * of the thermo row is pi halfLx^2 and the wall perimeter pi Lx.  Under
     * the field the time is the quartic's smallest root (EDMD.c:2918-2935,
     * quartic.c) with this oracle's own acos, cos and cube root. */
    int wall_circle;
    /* Mid wall (addMidWall, EDMD.c:269): a wall at x = Lx/2 for the disks
     * of type 1 only (the solute; type 0, the solvent, passes), met from
     * within 4 of it (EDMD.c:2897-2915, growth EDMD.c:3175-3195); a hit is
     * an x-wall hit whose momentum also goes to dpLeft / dpMid / dpRight by
     * the disk's position (EDMD.c:3393-3409); thermo rows add pmid,
     * -(dpRight - dpLeft), the packing of the solvent on each side (with
     * sizeratio) and the solvent's and solute's kinetic energies
     * (EDMD.c:5453-5477).  type: one int per disk (NULL: all 0). */
    int mid_wall;
    const int32_t* type;
    double sizeratio; /* EDMD.c:83, default 0.4 */
    /* Thermal walls (thermoWall, EDMD.c:3411-3452): a flat wall's hit sets
     * the normal velocity to -sign(v_n) drand(0, T/m), the circle's sets
     * v = -n drand(0, T) (one draw per component); drand(a, b) = a + (b - a)
     * k / 2^32 for the next output k of an MT19937 stream seeded by
     * init_genrand(rng_seed).  Shear walls (addShearWall, EDMD.c:3429-3435):
     * a y-wall hit then sets vx to +0.1 (y + 2 rad > Ly) or -0.1. */
    int thermal_wall;
    int shear_wall;
    uint32_t rng_seed; /* default 5489 */
} orc_params;

typedef struct {
    int32_t type;
    int32_t i;
    int32_t j;
    int32_t valid;
    double t;
} orc_trace;

typedef struct {
    double t;
    uint64_t ncol;
    double e_per_n;
    double pressure;
    double pxx, pyy, pxy, pyx;
    double pwall; /* wall pressure dp / (dt * perimeter), 0 without walls (EDMD.c:5428-5451) */
    /* mid wall columns (EDMD.c:5453-5477), 0 without it: pmid, dp, packLeft,
     * packRight, tempSolventLeft, tempSolventRight, tempSolute */
    double pmid, dp_lr, pack_left, pack_right, temp_left, temp_right, temp_solute;
} orc_thermo_row;

typedef struct orc_ctx orc_ctx;

void orc_default_params(orc_params* p);
orc_ctx* orc_create(int n, double Lx, double Ly, const double* x, const double* y,
                    const double* vx, const double* vy, const double* rad,
                    const double* m, const orc_params* prm);
void orc_destroy(orc_ctx* s);
/* addEventToQueue of the eventList slot of (kind, i) filled with time t,
 * partner or direction j and, for a collision, the partner counter
 * coll_actual; kinds ORC_CELLCROSS, ORC_COLLISION, ORC_WALL (j = xy),
 * ORC_SCREENSHOT, ORC_THERMO, ORC_ADDINGNOISE.  removeEventFromQueue of that
 * slot; the slot must be queued. */
int orc_add_event(orc_ctx* s, int kind, int i, int j, double t, uint64_t coll_actual);
int orc_remove_event(orc_ctx* s, int kind, int i);
/* particles[i].coll */
uint64_t orc_coll_count(const orc_ctx* s, int i);
/* crossingEvent(i), collisionEvent(i) (predict and file; the old slot must
 * have been removed) and freeFly(p_i) at the current time */
void orc_crossing_event(orc_ctx* s, int i);
void orc_collision_event(orc_ctx* s, int i);
void orc_free_fly(orc_ctx* s, int i);

/* eventListInit (EDMD.c:1935-2019): thermo, screenshot, optional noise
 * event, then crossing and collision prediction for every particle. */
int orc_init_events(orc_ctx* s);
/* Pops and handles up to nevents events; writes one trace record per event
 * when trace != NULL.  Returns a status code, *done = events handled. */
int orc_run(orc_ctx* s, int64_t nevents, orc_trace* trace, int64_t* done);
/* The reference's main loop bound: while (t <= tmax) pop and handle. */
int orc_run_until(orc_ctx* s, double tmax, int64_t max_events, int64_t* done);
/* Runs until ncol >= target collisions (or max_events). */
int orc_run_collisions(orc_ctx* s, uint64_t target, int64_t max_events, int64_t* done);

int orc_n(const orc_ctx* s);
double orc_time(const orc_ctx* s);
void orc_grid(const orc_ctx* s, int* ncx, int* ncy, double* csx, double* csy);
void orc_get_state(const orc_ctx* s, double* x, double* y, double* vx, double* vy,
                   double* tl, uint64_t* coll, int32_t* cx, int32_t* cy);
/* Current radius of every disk (they change while growing). */
void orc_get_radii(const orc_ctx* s, double* rad);
/* Current calendar slots: crossing (time, direction) and collision (time,
 * partner, partner's counter at prediction) for every particle. */
void orc_get_events(const orc_ctx* s, double* t_cross, int32_t* dir, double* t_coll,
                    int32_t* partner, uint64_t* coll_actual);
/* counters[0..2] = ncol, ncross, nevents; sums[0..7] = t, virial XX YY XY YX,
 * E, px, py (E, px, py recomputed by physicalQ, EDMD.c:5968-6013). */
void orc_observables(orc_ctx* s, uint64_t* counters, double* sums);
int orc_thermo_rows(const orc_ctx* s, orc_thermo_row* out, int cap);
int orc_status(const orc_ctx* s);

/* Stand-alone pieces for unit tests.  All act at the context's current time. */
double orc_collision_time(orc_ctx* s, int i, int j);
void orc_predict_collision(orc_ctx* s, int i, double* t_abs, int32_t* partner);
void orc_predict_crossing(orc_ctx* s, int i, double* t_abs, int32_t* dir);
/* Moves the clock (no events handled) and free-flies every particle to it,
 * as the screenshot event does (EDMD.c:4659-4661). */
void orc_sync_all(orc_ctx* s, double t);

/* ---- derivatives of the event map (no reference counterpart) ----------
 * A tape records every valid collision: the pair, the time, the contact
 * vector a = r_j - r_i (minimum image), the relative velocity c = v_j - v_i
 * before the collision and the impulse factor ff of the collision rule
 * (EDMD.c:3583-3606).  For a fixed event sequence the tangent map of one
 * collision (dt = -(a.p)/b, b = a.c, p = dr_j - dr_i, q = dv_j - dv_i) is
 *   rc = p + c dt, dff = (ff/b)(rc.c + a.q), u = dff a + ff rc,
 *   dv_i += m_j u, dv_j -= m_i u, dr_i -= m_j ff dt a, dr_j += m_i ff dt a,
 *   dw = m_i m_j |a|^2 dff        (w = m_i m_j ff |a|^2, the virial term)
 * and free flight is dr += (t - t_last) dv.  A hit of a flat wall (x or y
 * walls, the mid wall) is a record with j = -1 - xy (xy = 0 for a wall
 * normal to x), ax / ay = the normal / tangential velocity vn, vt before the
 * hit, cx / cy = resW / resT and ff = the contact coordinate; with n / s
 * the normal / tangential components of the tangent flown to the hit,
 *   dt = -dr_n / vn, dr_n' = -resW dr_n, dr_s' = dr_s + (vt - resT vt) dt,
 *   dv_n' = -resW dv_n, dv_s' = resT dv_s
 * (the reflected disk evaluated at the unperturbed hit time); under the
 * field g along y (record field g) the velocity along y gains the field's
 * pull over dt as well: dv_y' += (resT - 1) g dt for a wall normal to x,
 * dv_y' -= (1 + resW) g dt for one normal to y (pair collisions and the
 * tangent's free flight are unchanged by a uniform field).  Under damping
 * at rate gamma (the gamma argument; 0 without) the tangent's flight over h
 * is dr += dv (1 - e)/gamma, dv *= e with e = exp(-gamma h) (orc_exp), a
 * pair collision's dff loses gamma ff dt and its velocity jump, applied dt
 * early, decays less: dv_i += gamma m_j ff dt a, dv_j -= gamma m_i ff dt a;
 * wall maps are unchanged.  t_last is the time of the last pair
 * collision. */
typedef struct {
    int32_t i, j;
    double t, ax, ay, cx, cy, ff;
    double g; /* wall records: the field along y at the hit (0 without it) */
} orc_tape_rec;

int orc_tape_enable(orc_ctx* s, int64_t capacity);
int64_t orc_tape_count(const orc_ctx* s);
void orc_get_tape(const orc_ctx* s, orc_tape_rec* out);
void orc_tape_clear(orc_ctx* s);
/* Tangent propagation of tan (4 per particle: dx dy dvx dvy at time tl0[i])
 * through ntape collisions, then free flight of every particle to t_end.
 * Returns the derivative of sum_k w_k; *dt_last = derivative of the time of
 * the last taped collision. */
double orc_tangent(int n, const double* m, const double* tl0, const orc_tape_rec* tape,
                   int64_t ntape, double t_end, double* tan, double* dt_last, double gamma);
/* Adjoint of the same map: the objective is lw * sum_k w_k + lt * t_last
 * + <lam_end, state at t_end> (lam_end may be NULL).  Writes its gradient
 * with respect to (x, y, vx, vy) stored at times tl0 into grad (4 per
 * particle). */
void orc_adjoint(int n, const double* m, const double* tl0, const orc_tape_rec* tape,
                 int64_t ntape, double t_end, double lw, double lt, const double* lam_end,
                 double* grad, double gamma);
/* The exponential and logarithm used under damping (see orc_params). */
double orc_exp(double x);
double orc_log(double x);
/* The library's own sin/cos (angles in [-pi, pi]) and atan2, the circular
 * wall's normal. */
void orc_sincos(double th, double* sn, double* cs);
double orc_atan2(double y, double x);
/* Smallest real root above 1e-9 of a t^4 + b t^3 + c t^2 + d t + e (1e10
 * if none), the circular wall's time under the field. */
double orc_smallest_root(double a, double b, double c, double d, double e);

/* The first n outputs of MT19937 seeded by init_genrand(seed) (the
 * generator of the thermal walls). */
void orc_mt_draws(uint32_t seed, int n, uint32_t* out);

/* Sum of w_k = m_i m_j ff |a|^2 over the taped collisions, in tape order. */
double orc_tape_wsum(const orc_ctx* s);

/* Fixed-sequence replay: the nonlinear map whose derivative the tangent and
 * adjoint compute.  Starting from (x, y, vx, vy) at times tl (updated in
 * place, 4 arrays + tl), every taped collision (i, j) is redone in tape
 * order: both disks fly to the taped time, the contact time is re-solved
 * from the perturbed state (stable root of |a + c s|^2 = 4 r_i r_j), both
 * fly to it and the collision rule is applied; a wall record's disk flies
 * on to the taped contact coordinate and reflects; finally every disk flies
 * to t_end.  Flights follow the field along y (field; 0 without it) or
 * are damped at rate gamma (0 without), a displacement tau of the straight
 * motion then taking -log(1 - gamma tau)/gamma.
 * Returns sum of w_k; *t_last = time of the last replayed collision.
 * Returns NaN when a pair no longer meets. */
double orc_replay(int n, double Lx, double Ly, const double* rad, const double* m,
                  double res, const orc_tape_rec* tape, int64_t ntape, double t_end,
                  double* x, double* y, double* vx, double* vy, double* tl,
                  double* t_last, double field, double gamma);

/* Central difference (J(s0 + h e) - J(s0 - h e)) / 2h of
 *   J = lw * sum_k w_k + lt * t_last + <lam, state at t_end>   (lam may be NULL)
 * along direction e (4 per particle) over the fixed-sequence replay above,
 * evaluated in extended precision (long double) so that rounding noise
 * stays far below the derivative. */
double orc_replay_fd(int n, double Lx, double Ly, const double* rad, const double* m,
                     double res, const orc_tape_rec* tape, int64_t ntape, double t_end,
                     const double* x, const double* y, const double* vx, const double* vy,
                     const double* tl, const double* e, double h, double lw, double lt,
                     const double* lam, double field, double gamma);

/* g(r) histogram of calculate_pcf (pcf.c:16-75).  Returns the bin count;
 * r and g must hold at least (int)(max_r/dr) values. */
int orc_pcf(int n, const double* x, const double* y, double Lx, double Ly,
            double dr, double max_r, double* r, double* g);

#ifdef __cplusplus
}
#endif
#endif

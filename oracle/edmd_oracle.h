/*
 * CPU restatement of the reference's event-driven hard-disk engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU library is
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
       ORC_ADDINGNOISE = 4 };

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
} orc_thermo_row;

typedef struct orc_ctx orc_ctx;

void orc_default_params(orc_params* p);
orc_ctx* orc_create(int n, double Lx, double Ly, const double* x, const double* y,
                    const double* vx, const double* vy, const double* rad,
                    const double* m, const orc_params* prm);
void orc_destroy(orc_ctx* s);
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

/* g(r) histogram of calculate_pcf (pcf.c:16-75).  Returns the bin count;
 * r and g must hold at least (int)(max_r/dr) values. */
int orc_pcf(int n, const double* x, const double* y, double Lx, double Ly,
            double dr, double max_r, double* r, double* g);

#ifdef __cplusplus
}
#endif
#endif

/*
 * Reference-compatible entry points of libedmd_hip.so.
 *
 * Same names, argument meaning and call order as the reference's engine API
 * (EDMD.h:78-193) and the globals its callers read (EDMD.c:80-495), acting
 * on one default single-system context whose state lives on the GPU.
 * Differences, all forced by the state being device-resident:
 *   - particle* arguments are particle indices;
 *   - findNextEvent returns a snapshot node whose tree links are NULL;
 *   - instead of exit(3) on an overlap (EDMD.c:2715-2716) the call prints
 *     the same diagnosis to stderr and records a status readable with
 *     edmd_compat_status(); the `edmd` command exits with that code.
 *
 * Call order, as main() (EDMD.c:508-522):
 *   constantInit(argc, argv); particlesInit(); cellListInit();
 *   eventListInit(); physicalQ();
 *   while (t <= tmax) { nextEvent = findNextEvent(); t = nextEvent->t;
 *     removeEventFromQueue(nextEvent); switch (nextEvent->type) {...} }
 */
#ifndef EDMD_COMPAT_H
#define EDMD_COMPAT_H

#include "edmd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's event node (EDMD.h:66-74); `type` holds an EDMD_EV_* kind. */
typedef struct node node;
struct node {
    node *lft, *rgt, *top;
    int i, j, q;
    double t;
    int type;
    unsigned long int collActual;
};

/* Globals of the reference read by its callers (EDMD.c:80-495); refreshed
 * after every call below. */
extern int N;
extern double phi, Lx, Ly, t, tmax, E, px, py, pressure, res;
extern unsigned long int ncol, ncross;
extern int Nxcells, Nycells;
extern node* nextEvent;
/* The reference's eventList (EDMD.c:488, 1935-1945), 2N + 10 nodes after
 * eventListInit: slot i is particle i's crossing, N + i its collision,
 * 2N + 1 the screenshot, 2N + 2 the thermostat, 2N + 3 the thermo row.
 * The nodes name the device slots they stand for: removeEventFromQueue and
 * addEventToQueue act on that slot (the latter with the node's t, j and
 * collActual, as the reference's callers fill them); addCollisionEvent and
 * addCrossingEvent also record t and j in the node.  Their fields are not
 * refreshed by the events the GPU handles. */
extern node** eventList;
/* The reference's dispatch pointers (EDMD.c:317-321) for the Normal model. */
extern void (*collisionEvent)(int);
extern void (*crossingEvent)(int);
extern void (*doTheCollision)(void);
extern void (*freeFly)(int);

/* CLI options of constantInit (EDMD.c:750-977), same letters and long names.
 * Options of model variants this engine does not implement are accepted and
 * reported on stderr.  Extra long options (no reference counterpart):
 *   --lattice square|tri   lattice start instead of the growth start
 *   --noise-period P       velocity-rescale thermostat every P (noise = 2)
 *   --first-screen T, --first-thermo T
 *   --max-collisions C     also stop after C valid collisions
 *   --full-precision       dump values with 17 significant digits
 *   --no-output            write no dump or thermo files
 *   --trace-capacity K     keep the first K handled events (edmd_get_trace
 *                          on edmd_compat_context())
 * Without -l the start is the reference's own (particlesInit, EDMD.c:
 * 1690-1810: the box from -N/-p/-a or -L/-X, random centres, Box-Muller
 * velocities, growth at optimizeGrowConstant's rate) drawing from the
 * reference's MT19937 stream, seeded by -v as the reference seeds it, so
 * that a given command line starts from the reference's disks bit for
 * bit; -l with a file not ending in 'p' is the reference's text format
 * ("N Lx [Ly]", then "type x y rad m" per disk). */
void constantInit(int argc, char* argv[]);
void particlesInit(void);
void cellListInit(void);
void eventListInit(void);
void freeArrays(void);
void physicalQ(void);
node* findNextEvent(void);
/* The node findNextEvent returned is taken out by the handler that runs its
 * event (main removes it right before dispatching it, EDMD.c:532-534). */
void removeEventFromQueue(node* toRemove);
void addEventToQueue(node* toAdd);
void addCollisionEvent(int i, int j, double t);
void addCrossingEvent(int i, int j, double t);
void doTheCollisionNormal(void);
void doTheCrossing(void);
void takeAScreenshot(void);
void takeAThermo(void);
void addNoise(void);
double collisionTimeNormal(int i, int j);
void collisionEventNormal(int i);
void crossingEventNormal(int i);
void freeFlyNormal(int i);
void saveTXT(void);

/* Status of the most recent compat call (EDMD_OK when none failed). */
int edmd_compat_status(void);
/* The default context (NULL before eventListInit). */
edmd_ctx* edmd_compat_context(void);
/* Fast path of main's loop: handles events on the GPU until t > tmax, the
 * collision cap, or (when stop_mask != 0) right after an event of a kind in
 * stop_mask; returns the kind of the last event handled or -1 on error. */
int edmd_compat_run(int stop_mask);
/* The disks particlesInit made, on the host, before eventListInit: x, y,
 * velocities, radius (a growth start's targets: 1 or sizeratio), mass and
 * growth rate (vr times the target; 0 without growth).  Any pointer may be
 * NULL.  EDMD_ERR_STATE before particlesInit. */
int edmd_compat_particles(double* x, double* y, double* vx, double* vy, double* rad, double* m,
                          double* vr);
/* Output settings parsed by constantInit, for the command-line host. */
int edmd_compat_outputs(const char** dump_path, const char** thermo_path, int* full_precision);

#ifdef __cplusplus
}
#endif
#endif

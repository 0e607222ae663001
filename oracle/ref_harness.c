/*
 * Trace harness around the reference engine (test infrastructure only).
 *
 * This file is NOT part of the product.  It includes the reference's own
 * translation unit (src/EDMD.c, compiled where it lies under the reference
 * tree, see oracle/Makefile) and drives the same start-up sequence and event
 * dispatch as the reference's main() (EDMD.c:498-618, G == 0 branch), with
 * every file the reference would write sent to /dev/null.  It is used to
 *   - write golden vectors (initial prediction sweep, event traces, final
 *     states, virial sums, g(r) histograms) for the tests, and
 *   - time the reference's own CPU path for the CPU baseline.
 *
 * Modes (all start from a LAMMPS dump loaded with "-l", EDMD.c:799-807,
 * 987-995, 1387-1451):
 *   sweep <dump> <out>                 initial sweep after eventListInit
 *   trace <dump> <out> <nevents>       event trace + final state
 *   pcf   <dump> <out> <dr> <max_r>    calculate_pcf on the loaded positions
 *   run   <dump> <out> <tmax> <thermo> main's loop to tmax, final state,
 *                                      the reference's thermo rows
 *   bench <dump> <ncoll> [maxsec]      valid collisions / wall second
 *   grow  <dump> <out> <nevents> <vr>  growth start: the dump's radii are the
 *                                      targets; every disk starts at radius
 *                                      0 growing at vr * target (EDMD.c:
 *                                      1770-1773) and eventListInit takes its
 *                                      growth branch (EDMD.c:1957-1981);
 *                                      event trace, final state and radii
 *   cli   <out> <nevents> <args...>   the reference's own start from its
 *                                      command line (no -l dump: its random
 *                                      growth start or its text load path,
 *                                      particlesInit, EDMD.c:1354-1838); the
 *                                      state right after particlesInit (with
 *                                      every disk's growth rate), the event
 *                                      trace, final state and radii
 *   noise <dump> <out> <nevents> <T> <dtnoise> <vr>
 *                                      velocity-rescale thermostat: after the
 *                                      start (a growth start when vr > 0) the
 *                                      harness sets T and dtnoise (plain
 *                                      globals, EDMD.c:383, 386) and
 *                                      schedules the first ADDINGNOISE event
 *                                      at dtnoise (addEventNoise, EDMD.c:
 *                                      4814-4822); event trace, final state,
 *                                      radii and E after every rescale
 * Extra arguments after the mode arguments are passed to constantInit.
 */
#define main edmd_reference_main
#include "EDMD.c"
#undef main

#include <stdint.h>
#include <time.h>

typedef struct {
    int32_t type;
    int32_t i;
    int32_t j;
    int32_t valid;
    double t;
} trace_rec;

static void fail(const char* msg) {
    fprintf(stderr, "ref_harness: %s\n", msg);
    exit(2);
}

static void open_sinks(void) {
    fichier = fopen("/dev/null", "w");
    thermo = fopen("/dev/null", "w");
    if (!fichier || !thermo) fail("cannot open /dev/null");
}

/* CS1: constantInit -> particlesInit -> cellListInit -> eventListInit ->
 * physicalQ, as EDMD.c:508-518 (customName and initThermo only name and open
 * output files and are replaced by open_sinks). */
static void start(const char* dump, int nextra, char** extra) {
    char* args[64];
    int n = 0;
    args[n++] = "edmd";
    args[n++] = "-l";
    args[n++] = (char*)dump;
    args[n++] = "-t";
    args[n++] = "1e300";
    for (int k = 0; k < nextra && n < 63; k++) args[n++] = extra[k];
    args[n] = NULL;
    init_genrand(1);
    constantInit(n, args);
    configureTransverseVelocityCorrelationSampling(&dtimeThermo, &logSpacing);
    /* the reference fixes firstThermo = 1000 (EDMD.c:248, no option);
     * EDMD_HARNESS_FIRST_THERMO moves it for short statistical runs */
    const char* ft = getenv("EDMD_HARNESS_FIRST_THERMO");
    if (ft) firstThermo = atof(ft);
    particlesInit();
    cellListInit();
    eventListInit();
    physicalQ();
    open_sinks();
}

/* start() with the growth branch of particlesInit / eventListInit applied to
 * the loaded disks: rad -> 0, vr_i = vr * target radius. */
static void start_grow(const char* dump, double rate, int nextra, char** extra) {
    char* args[64];
    int n = 0;
    args[n++] = "edmd";
    args[n++] = "-l";
    args[n++] = (char*)dump;
    args[n++] = "-t";
    args[n++] = "1e300";
    for (int k = 0; k < nextra && n < 63; k++) args[n++] = extra[k];
    args[n] = NULL;
    init_genrand(1);
    constantInit(n, args);
    configureTransverseVelocityCorrelationSampling(&dtimeThermo, &logSpacing);
    particlesInit();
    vr = rate;
    for (int i = 0; i < N; i++) {
        particles[i].vr = vr * particles[i].rad;
        particles[i].rad = 0;
    }
    cellListInit();
    int saved = load;
    load = 0;
    eventListInit();
    load = saved;
    physicalQ();
    open_sinks();
}

/* The command line as given (the reference's main, EDMD.c:498-518, with
 * init_genrand(1) standing in for its time(NULL) seed, which -v replaces);
 * the particles right after particlesInit are written to f. */
static void start_cli(int nargs, char** cli, FILE* f) {
    char* args[64];
    int n = 0;
    args[n++] = "edmd";
    for (int k = 0; k < nargs && n < 63; k++) args[n++] = cli[k];
    args[n] = NULL;
    init_genrand(1);
    constantInit(n, args);
    configureTransverseVelocityCorrelationSampling(&dtimeThermo, &logSpacing);
    particlesInit();
    int32_t n32 = N;
    double box[3] = {Lx, Ly, vr};
    fwrite(&n32, 4, 1, f);
    fwrite(box, 8, 3, f);
    for (int i = 0; i < N; i++) {
        particle* p = particles + i;
        double d[7] = {p->x, p->y, p->vx, p->vy, p->rad, p->m, p->vr};
        fwrite(d, 8, 7, f);
    }
    cellListInit();
    eventListInit();
    physicalQ();
    open_sinks();
}

/* E right after each rescale (noise mode), for the invariant E/N = T. */
static double* rescale_e;
static long n_rescale, cap_rescale;

/* One pass of the dispatch switch of EDMD.c:534-610. */
static int step(trace_rec* rec) {
    nextEvent = findNextEvent();
    rec->type = (int32_t)nextEvent->type;
    rec->i = nextEvent->i;
    rec->j = nextEvent->j;
    rec->t = nextEvent->t;
    rec->valid = 1;
    if (nextEvent->type == COLLISION)
        rec->valid = (nextEvent->collActual == particles[nextEvent->j].coll);
    t = nextEvent->t;
    removeEventFromQueue(nextEvent);
    switch (nextEvent->type) {
        case COLLISION: doTheCollision(); break;
        case WALL: doTheWall(); break;
        case CELLCROSS: ncross++; doTheCrossing(); break;
        case ADDINGNOISE:
            /* This build fixes noise at 0 (EDMD.c:289), so addNoise() takes
             * the velocity-rescale branch (v /= sqrt(E/N/T), EDMD.c:4880-
             * 4901) without the physicalQ() it makes first for noise == 2
             * (EDMD.c:4830-4832): make that call here. */
            physicalQ();
            addNoise();
            if (n_rescale < cap_rescale) {
                physicalQ();
                rescale_e[n_rescale++] = E;
            }
            break;
        case SCREENSHOT: takeAScreenshot(); break;
        case THERMO: takeAThermo(); break;
        case GROWSTOP: stopGrow(); break;
        default: fail("unexpected event type");
    }
    return 0;
}

static void write_all(FILE* f, const void* p, size_t sz, size_t n) {
    if (fwrite(p, sz, n, f) != n) fail("short write");
}

/* Final-state block: N, then per particle x y vx vy t (double), coll (u64),
 * cell[0] cell[1] (int32), then ncol ncross (u64), t, virial sums, E px py. */
static void write_state(FILE* f) {
    int32_t n32 = N;
    write_all(f, &n32, 4, 1);
    for (int i = 0; i < N; i++) {
        particle* p = particles + i;
        double d[5] = {p->x, p->y, p->vx, p->vy, p->t};
        uint64_t c = p->coll;
        int32_t cell[2] = {p->cell[0], p->cell[1]};
        write_all(f, d, 8, 5);
        write_all(f, &c, 8, 1);
        write_all(f, cell, 4, 2);
    }
    uint64_t counters[2] = {ncol, ncross};
    write_all(f, counters, 8, 2);
    physicalQ();
    double g[8] = {t, collTermX, collTermY, collTermXY, collTermYX, E, px, py};
    write_all(f, g, 8, 8);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char** argv) {
    if (argc < 3) fail("usage: mode dump ...");
    /* The reference prints a progress bar at every screenshot/thermo event. */
    if (!freopen("/dev/null", "w", stdout)) fail("cannot silence stdout");
    const char* mode = argv[1];
    const char* dump = argv[2];

    if (!strcmp(mode, "sweep")) {
        if (argc < 4) fail("sweep dump out");
        start(dump, argc - 4, argv + 4);
        FILE* f = fopen(argv[3], "wb");
        if (!f) fail("cannot open output");
        int32_t n32 = N;
        double box[2] = {Lx, Ly};
        int32_t cells[2] = {Nxcells, Nycells};
        write_all(f, &n32, 4, 1);
        write_all(f, box, 8, 2);
        write_all(f, cells, 4, 2);
        for (int i = 0; i < N; i++) {
            node* cr = eventList[i];
            node* co = eventList[N + i];
            double tt[2] = {cr->t, co->t};
            int32_t jj[2] = {cr->j, co->j};
            uint64_t ca = co->collActual;
            int32_t cell[2] = {particles[i].cell[0], particles[i].cell[1]};
            write_all(f, tt, 8, 2);
            write_all(f, jj, 4, 2);
            write_all(f, &ca, 8, 1);
            write_all(f, cell, 4, 2);
        }
        fclose(f);
        return 0;
    }
    if (!strcmp(mode, "trace")) {
        if (argc < 5) fail("trace dump out nevents");
        long nev = atol(argv[4]);
        start(dump, argc - 5, argv + 5);
        FILE* f = fopen(argv[3], "wb");
        if (!f) fail("cannot open output");
        int64_t n64 = nev;
        write_all(f, &n64, 8, 1);
        for (long k = 0; k < nev; k++) {
            trace_rec rec;
            step(&rec);
            write_all(f, &rec, sizeof rec, 1);
        }
        write_state(f);
        fclose(f);
        return 0;
    }
    if (!strcmp(mode, "grow")) {
        if (argc < 6) fail("grow dump out nevents vr");
        long nev = atol(argv[4]);
        start_grow(dump, atof(argv[5]), argc - 6, argv + 6);
        FILE* f = fopen(argv[3], "wb");
        if (!f) fail("cannot open output");
        int64_t n64 = nev;
        write_all(f, &n64, 8, 1);
        for (long k = 0; k < nev; k++) {
            trace_rec rec;
            step(&rec);
            write_all(f, &rec, sizeof rec, 1);
        }
        write_state(f);
        for (int i = 0; i < N; i++) write_all(f, &particles[i].rad, 8, 1);
        fclose(f);
        return 0;
    }
    if (!strcmp(mode, "cli")) {
        /* cli <out> <nevents> <args...>: argv[2] is the output here */
        if (argc < 5) fail("cli out nevents args...");
        long nev = atol(argv[3]);
        FILE* f = fopen(argv[2], "wb");
        if (!f) fail("cannot open output");
        start_cli(argc - 4, argv + 4, f);
        int64_t n64 = nev;
        write_all(f, &n64, 8, 1);
        for (long k = 0; k < nev; k++) {
            trace_rec rec;
            step(&rec);
            write_all(f, &rec, sizeof rec, 1);
        }
        write_state(f);
        for (int i = 0; i < N; i++) write_all(f, &particles[i].rad, 8, 1);
        fclose(f);
        return 0;
    }
    if (!strcmp(mode, "noise")) {
        if (argc < 8) fail("noise dump out nevents T dtnoise vr");
        long nev = atol(argv[4]);
        double temp = atof(argv[5]);
        double period = atof(argv[6]);
        double rate = atof(argv[7]);
        if (rate > 0)
            start_grow(dump, rate, argc - 8, argv + 8);
        else
            start(dump, argc - 8, argv + 8);
        T = temp;
        dtnoise = period;
        addEventNoise(dtnoise);
        cap_rescale = nev;
        rescale_e = calloc((size_t)nev + 1, sizeof(double));
        if (!rescale_e) fail("out of memory");
        FILE* f = fopen(argv[3], "wb");
        if (!f) fail("cannot open output");
        int64_t n64 = nev;
        write_all(f, &n64, 8, 1);
        for (long k = 0; k < nev; k++) {
            trace_rec rec;
            step(&rec);
            write_all(f, &rec, sizeof rec, 1);
        }
        write_state(f);
        for (int i = 0; i < N; i++) write_all(f, &particles[i].rad, 8, 1);
        int64_t nr = n_rescale;
        write_all(f, &nr, 8, 1);
        write_all(f, rescale_e, 8, (size_t)n_rescale);
        fclose(f);
        return 0;
    }
    if (!strcmp(mode, "pcf")) {
        if (argc < 6) fail("pcf dump out dr max_r");
        double dr = atof(argv[4]);
        double max_r = atof(argv[5]);
        start(dump, argc - 6, argv + 6);
        pcf_data* d = calculate_pcf(particles, N, dr, max_r, Lx, Ly);
        if (!d) fail("calculate_pcf failed");
        FILE* f = fopen(argv[3], "wb");
        if (!f) fail("cannot open output");
        int32_t nb = d->num_bins;
        write_all(f, &nb, 4, 1);
        write_all(f, d->r, 8, (size_t)nb);
        write_all(f, d->g_r, 8, (size_t)nb);
        fclose(f);
        return 0;
    }
    if (!strcmp(mode, "run")) {
        /* run <dump> <out> <tmax> <thermo_out>: main's loop bound (t <= tmax);
         * the reference's own thermo rows are written to thermo_out. */
        if (argc < 6) fail("run dump out tmax thermo_out");
        double tm = atof(argv[4]);
        start(dump, argc - 6, argv + 6);
        fclose(thermo);
        thermo = fopen(argv[5], "w");
        if (!thermo) fail("cannot open thermo output");
        trace_rec rec;
        long nev = 0;
        while (t <= tm) {
            step(&rec);
            nev++;
        }
        fclose(thermo);
        thermo = fopen("/dev/null", "w");
        FILE* f = fopen(argv[3], "wb");
        if (!f) fail("cannot open output");
        int64_t n64 = 0;
        write_all(f, &n64, 8, 1);
        write_state(f);
        fclose(f);
        fprintf(stderr, "{\"events\": %ld}\n", nev);
        return 0;
    }
    if (!strcmp(mode, "bench")) {
        if (argc < 4) fail("bench dump ncoll [maxsec]");
        unsigned long target = strtoul(argv[3], NULL, 10);
        double maxsec = argc > 4 ? atof(argv[4]) : 1e30;
        start(dump, argc - 5 > 0 ? argc - 5 : 0, argv + 5);
        unsigned long events = 0;
        double t0 = now_s();
        double el = 0;
        trace_rec rec;
        while (ncol < target) {
            step(&rec);
            events++;
            if ((events & 0xffff) == 0) {
                el = now_s() - t0;
                if (el > maxsec) break;
            }
        }
        el = now_s() - t0;
        fprintf(stderr,
                "{\"ncol\": %lu, \"events\": %lu, \"seconds\": %.6f, "
                "\"coll_per_s\": %.6e, \"events_per_s\": %.6e, \"t\": %.17g}\n",
                ncol, events, el, ncol / el, events / el, t);
        return 0;
    }
    fail("unknown mode");
    return 1;
}

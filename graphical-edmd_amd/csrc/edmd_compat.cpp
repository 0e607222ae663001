/*
 * Reference-compatible entry points (include/edmd_compat.h) on a default
 * single-system context.  Everything here is host bookkeeping: the particle
 * state, the calendar and every event handler run on the GPU through the
 * handle API of edmd_api.cpp.
 */
#include <getopt.h>
#include <math.h>
#include <time.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "../../include/edmd.h"
#include "../../include/edmd_compat.h"

/* ---- globals of the reference (EDMD.c:80-495) ------------------------- */
extern "C" {
int N = 500;
double phi = 0.38;
double Lx = 0, Ly = 0;
double t = 0;
double tmax = 4000;
double E = 0, px = 0, py = 0, pressure = 0;
double res = 1;
unsigned long int ncol = 0, ncross = 0;
int Nxcells = 0, Nycells = 0;
node* nextEvent = nullptr;
node** eventList = nullptr;
void (*collisionEvent)(int) = collisionEventNormal;
void (*crossingEvent)(int) = crossingEventNormal;
void (*doTheCollision)(void) = doTheCollisionNormal;
void (*freeFly)(int) = freeFlyNormal;
}

namespace {

struct Config {
    int load = 0;
    std::string file = "data.txt";
    double sizeratio = 0.4;      /* EDMD.c:83 */
    double fraction_small = 0.3; /* EDMD.c:84 */
    double aspect = 1.0;
    int control_length = 0;
    double dtime = 100, dtime_thermo = 100, first_screen = 1, first_thermo = 1000;
    double einit = 1;
    double temperature = 0.05; /* EDMD.c:379 */
    int version = 1;
    int lattice = -1; /* -1: the reference's growth start; 0 square, 1 triangular */
    double vr = 0;    /* growth rate of the growth start */
    double noise_period = 0;
    uint64_t max_coll = ~(uint64_t)0;
    int full_precision = 0;
    int no_output = 0;
    int64_t trace_capacity = 0;
};

Config cfg;

/* MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998; the 2002
 * initialisation by a 32-bit seed), the random stream the reference's
 * starts draw from (init_genrand / genrand of mersenne.c), with its draw
 * genrand() = k / 2^32 for the 32-bit output k and drand(a, b) =
 * genrand() (b - a) + a (EDMD.c:5699-5701).  One stream for the process, as
 * in the reference. */
struct Mt19937 {
    static constexpr int NN = 624, MM = 397;
    uint32_t mt[NN];
    int mti = NN + 1;
    void seed(uint32_t s) {
        mt[0] = s;
        for (mti = 1; mti < NN; mti++)
            mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (uint32_t)mti;
    }
    uint32_t next() {
        static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
        uint32_t y;
        if (mti >= NN) {
            if (mti == NN + 1) seed(5489u);
            int kk;
            for (kk = 0; kk < NN - MM; kk++) {
                y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
                mt[kk] = mt[kk + MM] ^ (y >> 1) ^ mag01[y & 1u];
            }
            for (; kk < NN - 1; kk++) {
                y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
                mt[kk] = mt[kk + (MM - NN)] ^ (y >> 1) ^ mag01[y & 1u];
            }
            y = (mt[NN - 1] & 0x80000000u) | (mt[0] & 0x7fffffffu);
            mt[NN - 1] = mt[MM - 1] ^ (y >> 1) ^ mag01[y & 1u];
            mti = 0;
        }
        y = mt[mti++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= (y >> 18);
        return y;
    }
    double genrand() { return next() * (1.0 / 4294967296.0); }
    double drand(double a, double b) { return (genrand() * (b - a)) + a; }
};
Mt19937 g_mt;

edmd_ctx* g_ctx = nullptr;
int g_status = EDMD_OK;
node g_node;
/* the reference's eventList (EDMD.c:1935-1945): 2N + 10 nodes, slot i the
 * crossing of particle i, N + i its collision, 2N + 1 the screenshot,
 * 2N + 2 the thermostat, 2N + 3 the thermo row, 2N + 5 the growth stop */
std::vector<node> g_nodes;
std::vector<node*> g_list;
std::vector<double> hx, hy, hvx, hvy, hrad, hm;
std::string dump_path, thermo_path;

void report(int rc, const char* where) {
    if (rc == EDMD_OK) return;
    g_status = rc;
    if (rc == EDMD_ERR_OVERLAP)
        fprintf(stderr, "\nERROR: Overlaps detected during SIMULATION (%s): %s\n", where,
                edmd_last_error());
    else
        fprintf(stderr, "\nERROR: %s: %s\n", where, edmd_last_error());
}

void refresh() {
    if (!g_ctx) return;
    edmd_observables o;
    int rc = edmd_get_observables(g_ctx, &o);
    if (rc) { report(rc, "observables"); return; }
    t = o.t;
    ncol = o.ncol;
    ncross = o.ncross;
    E = o.E;
    px = o.px;
    py = o.py;
    edmd_thermo_row rows[1];
    int nrow = 0;
    if (edmd_get_thermo(g_ctx, nullptr, 0, &nrow) == EDMD_OK && nrow > 0) {
        std::vector<edmd_thermo_row> all(nrow);
        if (edmd_get_thermo(g_ctx, all.data(), nrow, &nrow) == EDMD_OK) {
            rows[0] = all[nrow - 1];
            pressure = rows[0].pressure;
        }
    }
}

/* one event of the expected kind, handled on the GPU */
void step_one(int kind, const char* where) {
    if (!g_ctx) { report(EDMD_ERR_STATE, where); return; }
    double tt;
    int k, i, j;
    uint32_t ca;
    int rc = edmd_peek_event(g_ctx, &tt, &k, &i, &j, &ca);
    if (rc) { report(rc, where); return; }
    if (k != kind) {
        fprintf(stderr, "%s: the next event is of kind %d, not %d\n", where, k, kind);
        g_status = EDMD_ERR_STATE;
        return;
    }
    edmd_observables o;
    rc = edmd_get_observables(g_ctx, &o);
    if (!rc) rc = edmd_run(g_ctx, (int64_t)o.nevents + 1, ~(uint64_t)0, 1e300);
    report(rc, where);
    refresh();
}

std::string versioned(const std::string& base, const char* ext, int* version) {
    for (;;) {
        std::string f = base + std::to_string(*version) + ext;
        if (access(f.c_str(), F_OK) != 0) return f;
        (*version)++;
    }
}

}  // namespace

extern "C" void saveTXT(void);

extern "C" int edmd_compat_status(void) { return g_status; }
extern "C" edmd_ctx* edmd_compat_context(void) { return g_ctx; }

extern "C" void constantInit(int argc, char* argv[]) {
    cfg = Config();
    /* the reference's main seeds the stream with the clock right before
     * constantInit (EDMD.c:500); -v seeds it again with the version */
    g_mt.seed((uint32_t)time(NULL));
    /* option table of EDMD.c:750-797 plus this engine's extra options */
    static struct option longopt[] = {
        {"load", required_argument, NULL, 'l'},
        {"number", required_argument, NULL, 'N'},
        {"phi", required_argument, NULL, 'p'},
        {"res", required_argument, NULL, 'r'},
        {"delta", required_argument, NULL, 'd'},
        {"gamma", required_argument, NULL, 'g'},
        {"time", required_argument, NULL, 't'},
        {"dt", required_argument, NULL, 'D'},
        {"initial-energy", required_argument, NULL, 'E'},
        {"potential", required_argument, NULL, 'U'},
        {"well-radius", required_argument, NULL, 'R'},
        {"field", required_argument, NULL, 'f'},
        {"ts", required_argument, NULL, 's'},
        {"xs", required_argument, NULL, 'x'},
        {"sizeratio", required_argument, NULL, 'q'},
        {"temperature", required_argument, NULL, 'T'},
        {"aspect", required_argument, NULL, 'a'},
        {"Ly", required_argument, NULL, 'L'},
        {"Lx", required_argument, NULL, 'X'},
        {"deltaE", required_argument, NULL, 'e'},
        {"additionalEnergy", required_argument, NULL, 'A'},
        {"dtimeThermo", required_argument, NULL, 'o'},
        {"dr_pressure", required_argument, NULL, 1001},
        {"phil", required_argument, NULL, 'P'},
        {"phig", required_argument, NULL, 'G'},
        {"version", required_argument, NULL, 'v'},
        {"strainRate", required_argument, NULL, 'S'},
        {"cluster-mode", required_argument, NULL, 'C'},
        {"cluster-neighbor-threshold", required_argument, NULL, 'H'},
        {"muller-plathe", no_argument, NULL, 1002},
        {"mp-interval", required_argument, NULL, 1003},
        {"mp-slabs", required_argument, NULL, 1004},
        {"mp-hot", required_argument, NULL, 1005},
        {"mp-cold", required_argument, NULL, 1006},
        {"temp-profile", no_argument, NULL, 1007},
        {"temp-slabs", required_argument, NULL, 1008},
        {"intruderMass", required_argument, NULL, 'M'},
        {"intruderDelta", required_argument, NULL, 'z'},
        {"intruderRes", required_argument, NULL, 'Y'},
        {"intruderKMBurn", required_argument, NULL, 'B'},
        {"lattice", required_argument, NULL, 3001},
        {"noise-period", required_argument, NULL, 3002},
        {"first-screen", required_argument, NULL, 3003},
        {"first-thermo", required_argument, NULL, 3004},
        {"max-collisions", required_argument, NULL, 3005},
        {"full-precision", no_argument, NULL, 3006},
        {"no-output", no_argument, NULL, 3007},
        {"trace-capacity", required_argument, NULL, 3008},
        {NULL, 0, NULL, 0}};
    optind = 1;
    int c;
    while ((c = getopt_long(argc, argv, "l:N:p:r:d:g:t:D:E:U:R:f:s:x:q:T:a:L:X:e:A:o:P:G:v:S:C:H:M:B:z:Y:n:",
                            longopt, NULL)) != -1) {
        switch (c) {
            case 'l': {
                cfg.file = optarg;
                size_t len = cfg.file.size();
                cfg.load = (len > 0 && cfg.file[len - 1] == 'p') ? 2 : 1;
                break;
            }
            case 'N': cfg.load = 0; sscanf(optarg, "%d", &N); break;
            case 'p': cfg.load = 0; sscanf(optarg, "%lf", &phi); break;
            case 'x': cfg.load = 0; sscanf(optarg, "%lf", &cfg.fraction_small); break;
            case 'a': cfg.load = 0; sscanf(optarg, "%lf", &cfg.aspect); break;
            case 'q': cfg.load = 0; sscanf(optarg, "%lf", &cfg.sizeratio); break;
            case 'r': sscanf(optarg, "%lf", &res); break;
            case 't': sscanf(optarg, "%lf", &tmax); break;
            case 'D': sscanf(optarg, "%lf", &cfg.dtime); break;
            case 'E': sscanf(optarg, "%lf", &cfg.einit); break;
            case 'T': sscanf(optarg, "%lf", &cfg.temperature); break;
            case 'L': cfg.control_length = 1; cfg.load = 0; sscanf(optarg, "%lf", &Ly); break;
            case 'X': cfg.control_length = 2; cfg.load = 0; sscanf(optarg, "%lf", &Lx); break;
            case 'o': sscanf(optarg, "%lf", &cfg.dtime_thermo); break;
            case 'v':
                sscanf(optarg, "%d", &cfg.version);
                g_mt.seed((uint32_t)cfg.version); /* init_genrand(version), EDMD.c:903-906 */
                break;
            case 3001: cfg.lattice = (strcmp(optarg, "tri") == 0) ? 1 : 0; break;
            case 3002:
                sscanf(optarg, "%lf", &cfg.noise_period);
                break;
            case 3003: sscanf(optarg, "%lf", &cfg.first_screen); break;
            case 3004: sscanf(optarg, "%lf", &cfg.first_thermo); break;
            case 3005: cfg.max_coll = strtoull(optarg, NULL, 10); break;
            case 3006: cfg.full_precision = 1; break;
            case 3007: cfg.no_output = 1; break;
            case 3008: cfg.trace_capacity = strtoll(optarg, NULL, 10); break;
            case '?': break;
            default:
                fprintf(stderr, "WARNING: option -%c (long code %d) belongs to a model variant this "
                                "engine does not implement; ignored\n", c < 256 ? c : '-', c);
                break;
        }
    }
    g_status = EDMD_OK;
}

namespace {

/* cos and sin of one angle as the reference's build evaluates them: its
 * compiler turns the pair cos(a), sin(a) of the Box-Muller draws
 * (EDMD.c:1375-1376, 1800-1801) into one call of the C library's sincos,
 * whose results can differ from separate cos and sin calls in the last
 * bit. */
void ref_sincos(double a, double* s, double* c) { sincos(a, s, c); }

/* physicalQ and normalizePhysicalQ (EDMD.c:5968-5990, 5723-5764; periodic
 * box, no shear) on the host arrays, in the reference's order. */
void host_physical_q(double* E_out, double* px_out, double* py_out) {
    double ppx = 0, ppy = 0, ee = 0;
    for (int i = 0; i < N; i++) {
        ppx += hvx[i] * hm[i];
        ppy += hvy[i] * hm[i];
        const double vx = hvx[i];
        ee += 0.5 * hm[i] * (vx * vx + hvy[i] * hvy[i]);
    }
    *E_out = ee;
    *px_out = ppx;
    *py_out = ppy;
}

void host_normalize(double einit) {
    double ee, ppx, ppy;
    host_physical_q(&ee, &ppx, &ppy);
    for (int i = 0; i < N; i++) {
        hvx[i] -= ppx / (N * hm[i]);
        hvy[i] -= ppy / (N * hm[i]);
    }
    host_physical_q(&ee, &ppx, &ppy);
    for (int i = 0; i < N; i++) {
        hvx[i] /= sqrt(ee / N / einit);
        hvy[i] /= sqrt(ee / N / einit);
    }
}

/* The text load path (load == 1, EDMD.c:980-986, 1366-1386): a line
 * "N Lx [Ly]", then one line "type x y rad m" per disk; velocities drawn
 * from the stream (Box-Muller, E/N = Einit per component pair), then
 * normalizePhysicalQ.  Returns nonzero on error. */
int text_start() {
    FILE* f = fopen(cfg.file.c_str(), "r");
    if (!f) {
        fprintf(stderr, "ERROR: cannot open %s\n", cfg.file.c_str());
        g_status = EDMD_ERR_IO;
        return 1;
    }
    char line[4096];
    int n = 0;
    double lx = 0, ly = 0;
    int nv = fgets(line, sizeof line, f) ? sscanf(line, "%d %lf %lf", &n, &lx, &ly) : 0;
    if (nv == 2) ly = lx;
    if (nv < 2 || n < 2 || !(lx > 0) || !(ly > 0)) {
        fclose(f);
        fprintf(stderr, "ERROR: %s: first line is not \"N Lx [Ly]\"\n", cfg.file.c_str());
        g_status = EDMD_ERR_IO;
        return 1;
    }
    N = n;
    Lx = lx;
    Ly = ly;
    hx.resize(n); hy.resize(n); hvx.resize(n); hvy.resize(n); hrad.resize(n); hm.resize(n);
    int nbig = 0;
    for (int i = 0; i < n; i++) {
        int type = 0;
        if (!fgets(line, sizeof line, f) ||
            sscanf(line, "%d %lf %lf %lf %lf", &type, &hx[i], &hy[i], &hrad[i], &hm[i]) != 5) {
            fclose(f);
            fprintf(stderr, "ERROR: %s: disk line %d is not \"type x y rad m\"\n",
                    cfg.file.c_str(), i + 1);
            g_status = EDMD_ERR_IO;
            return 1;
        }
        nbig += type;
        const double u1 = g_mt.drand(0, 1);
        const double u2 = g_mt.drand(0, 1);
        double sn, cs;
        ref_sincos(2 * M_PI * u2, &sn, &cs);
        hvx[i] = sqrt(-2 * log(u1) / hm[i] * cfg.einit) * cs;
        hvy[i] = sqrt(-2 * log(u1) / hm[i] * cfg.einit) * sn;
    }
    fclose(f);
    host_normalize(cfg.einit);
    const int nsmall = N - nbig;
    cfg.sizeratio = hrad[N - 1];
    phi = M_PI * (cfg.sizeratio * cfg.sizeratio * nsmall + nbig) / (Lx * Ly);
    return 0;
}

/* The reference's default start (EDMD.c:1119-1184 box, 1690-1810
 * particlesInit, load == 0, periodic box, no polydispersity): the box from
 * N, phi and the aspect ratio (or Ly / Lx given by -L / -X), the growth
 * rate of optimizeGrowConstant, and per disk, from the stream: x, y
 * uniform in the box, radius 0 growing towards 1 (the first N - Nsmall
 * disks) or sizeratio, mass 1, Box-Muller velocities of energy EinitGrow =
 * 0.05 per disk, and one more draw (the charge draw, unused here). */
int random_start() {
    const double r2 = (1 + cfg.fraction_small * (cfg.sizeratio * cfg.sizeratio - 1));
    if (cfg.control_length == 0) {
        Lx = sqrt(M_PI * N / (phi * cfg.aspect) * r2);
        Ly = Lx * cfg.aspect;
    } else if (cfg.control_length == 1) {
        Lx = Ly / cfg.aspect;
    } else {
        Ly = Lx * cfg.aspect;
        N = (int)(phi * Lx * Ly / (M_PI * r2));
        phi = M_PI * N * r2 / (Lx * Ly);
    }
    if (N < 2 || !(Lx > 0) || !(Ly > 0)) {
        fprintf(stderr, "ERROR: no box for N = %d, Lx = %g, Ly = %g\n", N, Lx, Ly);
        g_status = EDMD_ERR_ARG;
        return 1;
    }
    cfg.vr = edmd_growth_rate(phi); /* optimizeGrowConstant, EDMD.c:5766-5799 */
    const double einit_grow = 0.05;
    const int nsmall = (int)(N * cfg.fraction_small);
    const int nbig = N - nsmall;
    hx.resize(N); hy.resize(N); hvx.resize(N); hvy.resize(N); hrad.resize(N); hm.resize(N);
    for (int i = 0; i < N; i++) {
        hx[i] = g_mt.drand(0, Lx);
        hy[i] = g_mt.drand(0, Ly);
        /* the target radius: the growth rate of disk i is vr * target */
        hrad[i] = i < nbig ? 1.0 : cfg.sizeratio;
        hm[i] = 1;
        const double u1 = g_mt.drand(0, 1);
        const double u2 = g_mt.drand(0, 1);
        double sn, cs;
        ref_sincos(2 * M_PI * u2, &sn, &cs);
        hvx[i] = sqrt(-2 * log(u1) / hm[i] * einit_grow) * cs;
        hvy[i] = sqrt(-2 * log(u1) / hm[i] * einit_grow) * sn;
        (void)g_mt.drand(0, 1); /* randomTemp, the charge draw, EDMD.c:1806 */
    }
    return 0;
}

}  // namespace

extern "C" int edmd_compat_particles(double* x, double* y, double* vx, double* vy, double* rad,
                                     double* m, double* vr) {
    if (hx.empty() || (int)hx.size() != N) return EDMD_ERR_STATE;
    for (int i = 0; i < N; i++) {
        if (x) x[i] = hx[i];
        if (y) y[i] = hy[i];
        if (vx) vx[i] = hvx[i];
        if (vy) vy[i] = hvy[i];
        if (rad) rad[i] = hrad[i];
        if (m) m[i] = hm[i];
        if (vr) vr[i] = cfg.load == 0 && cfg.lattice < 0 ? cfg.vr * hrad[i] : 0.0;
    }
    return EDMD_OK;
}

extern "C" void particlesInit(void) {
    /* particlesInit, EDMD.c:1356-1451 (load path) and the lattice starts */
    if (cfg.load == 2) {
        double lx, ly;
        int n;
        int rc = edmd_dump_info(cfg.file.c_str(), &n, &lx, &ly);
        if (rc) { report(rc, "particlesInit"); return; }
        N = n;
        Lx = lx;
        Ly = ly;
        hx.resize(n); hy.resize(n); hvx.resize(n); hvy.resize(n); hrad.resize(n); hm.resize(n);
        rc = edmd_dump_read(cfg.file.c_str(), n, hx.data(), hy.data(), hvx.data(), hvy.data(),
                            hrad.data(), hm.data());
        if (rc) { report(rc, "particlesInit"); return; }
        double area = 0;
        for (int i = 0; i < n; i++) area += M_PI * hrad[i] * hrad[i];
        phi = area / (Lx * Ly);
    } else if (cfg.load == 1) {
        if (text_start()) return;
    } else if (cfg.lattice >= 0) {
        /* monodisperse lattice start (--lattice sq|tri) */
        if (cfg.aspect != 1 || cfg.control_length != 0) {
            fprintf(stderr, "ERROR: the lattice starts are square (no -a/-L/-X)\n");
            g_status = EDMD_ERR_ARG;
            return;
        }
        hx.resize(N); hy.resize(N); hvx.resize(N); hvy.resize(N);
        hrad.assign(N, 1.0); /* the reference's large disks have radius 1 */
        hm.assign(N, 1.0);
        int rc = edmd_init_lattice(N, phi, 1.0, cfg.lattice, (uint64_t)cfg.version, cfg.einit,
                                   hx.data(), hy.data(), hvx.data(), hvy.data(), &Lx, &Ly);
        if (rc) { report(rc, "particlesInit"); return; }
    } else {
        if (random_start()) return;
    }
    /* output names in the spirit of customName, EDMD.c:6015-6065 */
    char base[400];
    snprintf(base, sizeof base, "dump/N_%dres_%.3lfphi_%.6lfLx_%.3lfLy_%.3lfv_", N, res, phi, Lx, Ly);
    int version = cfg.version;
    dump_path = versioned(base, ".dump", &version);
    thermo_path = std::string(base) + std::to_string(version) + ".thermo";
}

extern "C" void cellListInit(void) {
    /* the cell grid and lists are built on the GPU side by eventListInit */
}

extern "C" void eventListInit(void) {
    if (g_status) return;
    if (g_ctx) { edmd_ctx_destroy(g_ctx); g_ctx = nullptr; }
    edmd_system_desc d;
    edmd_default_desc(&d);
    d.n = N;
    d.Lx = Lx;
    d.Ly = Ly;
    d.x = hx.data(); d.y = hy.data(); d.vx = hvx.data(); d.vy = hvy.data();
    d.rad = hrad.data(); d.m = hm.data();
    d.res = res;
    d.first_screen = cfg.first_screen;
    d.dtime = cfg.dtime;
    d.first_thermo = cfg.first_thermo;
    d.dtime_thermo = cfg.dtime_thermo;
    d.vr = cfg.load == 0 && cfg.lattice < 0 ? cfg.vr : 0.0;
    d.trace_capacity = cfg.trace_capacity;
    d.einit = cfg.einit;
    if (cfg.noise_period > 0) {
        d.noise = 2;
        d.dtnoise = cfg.noise_period;
        d.temperature = cfg.temperature;
    }
    int rc = edmd_ctx_create(&d, &g_ctx);
    if (rc) { report(rc, "eventListInit"); g_ctx = nullptr; return; }
    int ncx, ncy;
    edmd_get_grid(g_ctx, &ncx, &ncy, nullptr, nullptr, nullptr);
    Nxcells = ncx;
    Nycells = ncy;
    g_nodes.assign(2 * (size_t)N + 10, node{});
    g_list.resize(g_nodes.size());
    for (size_t k = 0; k < g_nodes.size(); k++) g_list[k] = &g_nodes[k];
    for (int i = 0; i < N; i++) {
        g_nodes[i].i = i;
        g_nodes[i].type = EDMD_EV_CELLCROSS;
        g_nodes[N + i].i = i;
        g_nodes[N + i].type = EDMD_EV_COLLISION;
    }
    g_nodes[2 * N + 1].type = EDMD_EV_SCREENSHOT;
    g_nodes[2 * N + 2].type = EDMD_EV_ADDINGNOISE;
    g_nodes[2 * N + 3].type = EDMD_EV_THERMO;
    g_nodes[2 * N + 5].type = EDMD_EV_GROWSTOP;
    eventList = g_list.data();
    refresh();
}

extern "C" void freeArrays(void) {
    if (g_ctx) edmd_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    eventList = nullptr;
    g_list.clear();
    g_nodes.clear();
}

extern "C" void physicalQ(void) { refresh(); }

extern "C" node* findNextEvent(void) {
    memset(&g_node, 0, sizeof g_node);
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return &g_node; }
    double tt;
    int k, i, j;
    uint32_t ca;
    int rc = edmd_peek_event(g_ctx, &tt, &k, &i, &j, &ca);
    if (rc) { report(rc, "findNextEvent"); return &g_node; }
    g_node.t = tt;
    g_node.type = k;
    g_node.i = i;
    g_node.j = j;
    g_node.collActual = ca;
    nextEvent = &g_node;
    return &g_node;
}

namespace {
/* The (kind, particle) a node of eventList stands for, from its place in
 * the list; other nodes (the findNextEvent snapshot, a caller's own) from
 * their fields.  False for nodes that name no queue slot. */
bool node_slot(const node* n, int* kind, int* i) {
    if (!n) return false;
    if (!g_nodes.empty() && n >= g_nodes.data() && n < g_nodes.data() + g_nodes.size()) {
        const long k = n - g_nodes.data();
        if (k < N) { *kind = EDMD_EV_CELLCROSS; *i = (int)k; return true; }
        if (k < 2 * (long)N) {
            *kind = n->type == EDMD_EV_WALL ? EDMD_EV_WALL : EDMD_EV_COLLISION;
            *i = (int)(k - N);
            return true;
        }
        *i = 0;
        switch (k - 2 * (long)N) {
            case 1: *kind = EDMD_EV_SCREENSHOT; return true;
            case 2: *kind = EDMD_EV_ADDINGNOISE; return true;
            case 3: *kind = EDMD_EV_THERMO; return true;
            default: return false;
        }
    }
    *kind = n->type;
    *i = n->i;
    return true;
}
}  // namespace

extern "C" void removeEventFromQueue(node* toRemove) {
    /* The event findNextEvent returned is taken out by the handler that
     * runs it on the GPU (main's loop removes it right before dispatch,
     * EDMD.c:532-534); any other slot is taken out now. */
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return; }
    if (toRemove == &g_node) return;
    int kind, i;
    if (!node_slot(toRemove, &kind, &i)) {
        fprintf(stderr, "removeEventFromQueue: not a node of eventList\n");
        g_status = EDMD_ERR_ARG;
        return;
    }
    report(edmd_queue_remove(g_ctx, kind, i), "removeEventFromQueue");
}

extern "C" void addEventToQueue(node* toAdd) {
    /* files the node as its fields say (addEventToQueue, EDMD.c:2144-2170):
     * time t, partner or direction j, partner counter collActual */
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return; }
    int kind, i;
    if (!node_slot(toAdd, &kind, &i)) {
        fprintf(stderr, "addEventToQueue: not a node of eventList\n");
        g_status = EDMD_ERR_ARG;
        return;
    }
    report(edmd_queue_add(g_ctx, kind, i, toAdd->j, toAdd->t, (uint32_t)toAdd->collActual),
           "addEventToQueue");
}

extern "C" void addCollisionEvent(int i, int j, double tc) {
    /* addCollisionEvent, EDMD.c:3286-3297 */
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return; }
    if (i >= 0 && i < N && !g_nodes.empty()) {
        node* nd = &g_nodes[N + i];
        nd->t = tc;
        nd->j = j;
        nd->type = EDMD_EV_COLLISION;
    }
    report(edmd_add_collision_event(g_ctx, i, j, tc), "addCollisionEvent");
}

extern "C" void addCrossingEvent(int i, int j, double tc) {
    /* addCrossingEvent, EDMD.c:2487-2496: j is the direction code 1..4 */
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return; }
    if (i >= 0 && i < N && !g_nodes.empty()) {
        node* nd = &g_nodes[i];
        nd->t = tc;
        nd->j = j;
        nd->type = EDMD_EV_CELLCROSS;
    }
    report(edmd_queue_add(g_ctx, EDMD_EV_CELLCROSS, i, j, tc, 0), "addCrossingEvent");
}

extern "C" void doTheCollisionNormal(void) { step_one(EDMD_EV_COLLISION, "doTheCollisionNormal"); }
extern "C" void doTheCrossing(void) { step_one(EDMD_EV_CELLCROSS, "doTheCrossing"); }
extern "C" void takeAScreenshot(void) {
    /* free flight of every disk, next screenshot, then the dump frame
     * (takeAScreenshot, EDMD.c:4657-4677) */
    step_one(EDMD_EV_SCREENSHOT, "takeAScreenshot");
    if (g_status == EDMD_OK) saveTXT();
}
extern "C" void takeAThermo(void) { step_one(EDMD_EV_THERMO, "takeAThermo"); }
extern "C" void addNoise(void) { step_one(EDMD_EV_ADDINGNOISE, "addNoise"); }

extern "C" double collisionTimeNormal(int i, int j) {
    double dt = 1e26;
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return dt; }
    report(edmd_collision_time(g_ctx, i, j, &dt), "collisionTimeNormal");
    return dt;
}

extern "C" void collisionEventNormal(int i) {
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return; }
    report(edmd_repredict(g_ctx, i, 2), "collisionEventNormal");
}

extern "C" void crossingEventNormal(int i) {
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return; }
    report(edmd_repredict(g_ctx, i, 1), "crossingEventNormal");
}

extern "C" void freeFlyNormal(int i) {
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return; }
    report(edmd_free_fly(g_ctx, i), "freeFlyNormal");
}

extern "C" void saveTXT(void) {
    /* saveTXT, EDMD.c:5012-5153: appends one frame of the current state */
    if (!g_ctx || cfg.no_output) return;
    std::vector<double> x(N), y(N), vx(N), vy(N);
    int rc = edmd_get_state(g_ctx, x.data(), y.data(), vx.data(), vy.data(), nullptr, nullptr,
                            nullptr, nullptr);
    if (!rc) rc = edmd_dump_write(dump_path.c_str(), 1, t, N, Lx, Ly, x.data(), y.data(), vx.data(),
                                  vy.data(), hrad.data(), hm.data(), cfg.full_precision);
    report(rc, "saveTXT");
}

extern "C" int edmd_compat_run(int stop_mask) {
    if (!g_ctx) { g_status = EDMD_ERR_STATE; return -1; }
    int reason = 0, kind = -1;
    int rc = edmd_run_until_kind(g_ctx, INT64_MAX, cfg.max_coll, tmax, stop_mask, &reason, &kind);
    report(rc, "run");
    refresh();
    if (rc) return -1;
    return reason == EDMD_STOP_SPECIAL ? kind : -2;
}

extern "C" int edmd_compat_outputs(const char** dp, const char** tp, int* full) {
    if (dp) *dp = cfg.no_output ? nullptr : dump_path.c_str();
    if (tp) *tp = cfg.no_output ? nullptr : thermo_path.c_str();
    if (full) *full = cfg.full_precision;
    return EDMD_OK;
}

/*
 * Host side of libedmd_hip.so: handle management, device layout, launches,
 * lattice initial states and LAMMPS dump I/O.  See include/edmd.h.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/edmd.h"
#include "edmd_internal.h"
#include "edmd_types.h"

/* ------------------------------------------------------------ errors */

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(EDMD_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

extern "C" const char* edmd_last_error(void) { return g_err; }

extern "C" int edmd_device_count(int* n) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    if (n) *n = c;
    return c > 0 ? EDMD_OK : fail(EDMD_ERR_NODEVICE, "no HIP device visible");
}

extern "C" void edmd_default_desc(edmd_system_desc* d) {
    /* reference defaults, EDMD.c:244-248, 372, 379, 386 */
    memset(d, 0, sizeof *d);
    d->res = 1;
    d->first_screen = 1;
    d->dtime = 100;
    d->first_thermo = 1000;
    d->dtime_thermo = 100;
    d->noise = 0;
    d->dtnoise = 0.5;
    d->temperature = 1;
    d->vr = 0;
    d->einit = 1;
    d->wall_x = 0;
    d->wall_y = 0;
    d->res_wall = 1; /* resW, EDMD.c:372 */
    d->res_tan = 0;  /* resT, EDMD.c:373 */
    d->add_field = 0;
    d->field = -0.02; /* EDMD.c:366 */
    d->damping = 0;
    d->gamma = 0.02; /* gamm, EDMD.c:382 */
    d->wall_circle = 0;
    d->mid_wall = 0;
    d->type = nullptr;
    d->sizeratio = 0.4; /* EDMD.c:83 */
    d->thermal_wall = 0;
    d->shear_wall = 0;
    d->rng_seed = 5489; /* mersenne.c's seed when none is given */
}

/* ------------------------------------------------------------ handles */

struct HostSys {
    EdmdConst c;
    size_t p0;     /* first particle in the pooled particle arrays */
    size_t cell0;  /* first cell */
    size_t slot0;  /* first cell slot */
    size_t node0[EDMD_MAX_LEVELS];
    size_t trace0; /* first trace record */
    size_t thermo0;
};

struct edmd_batch {
    int nsys = 0;
    std::vector<HostSys> hs;
    std::vector<EdmdSys> sys;
    EdmdSys* d_sys = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int max_n = 0, max_nlev = 0, max_tiles = 0;
    int max_lsize[EDMD_MAX_LEVELS + 1] = {0};
    size_t np = 0, ncell = 0, nslot = 0, nnode = 0, ntrace = 0, nthermo = 0;
    std::vector<void*> allocs;
    /* pooled device arrays */
    EdmdPart* part = nullptr;
    EdmdEv* ev = nullptr;
    EdmdNode* leaf = nullptr;
    EdmdNode* nodes = nullptr;
    int32_t* img = nullptr;
    int32_t *ccnt = nullptr, *cslot = nullptr;
    EdmdState* st = nullptr;
    EdmdTraceRec* trace = nullptr;
    EdmdThermoRow* thermo = nullptr;
    int64_t *erank = nullptr, *lrank = nullptr; /* tie ranks: 2 slots and 1 leaf per particle */
    /* growth starts: per-particle growth rates; any_grow while a system grows */
    double* vr = nullptr;
    int32_t* ptype = nullptr; /* disk types, when a system has the mid wall */
    uint32_t* mt = nullptr;   /* MT19937 streams, 625 words per system (thermal walls) */
    bool any_grow = false;
    /* results of the step-level calls (peek, pair time), allocated once */
    EdmdPeek* step_peek = nullptr;
    double* step_pair = nullptr;
    /* derivatives: tape pools, checkpoints (one snapshot of every pool per
     * segment, with the run bounds the segment was run with) */
    bool deriv_on = false;
    int64_t seg = 0, tape_cap = 0;
    std::vector<void*> dallocs;
    EdmdTapeRec* tape = nullptr;
    double *dv = nullptr, *dtl = nullptr, *otl = nullptr;
    EdmdDeriv* dr = nullptr;
    struct Ckpt {
        void* mem;
        EdmdRunArgs ra;
    };
    std::vector<Ckpt> ckpts;
    void* ckpt_end = nullptr;
    /* residency of the event loop's arrays: requested mode (EDMD_RES_*),
     * whether every system fits one workgroup's LDS, the dynamic LDS per
     * workgroup of the Normal [0] and growing [1] launches, the workgroups
     * the device holds at once in LDS mode, and the mode in use */
    /* tile shape of the prediction sweep (edmd_sweep_shape) */
    int sweep_variant = 0;
    int residency = EDMD_RES_AUTO;
    bool lds_fit = false;
    size_t lds_dyn[2] = {0, 0};
    int64_t lds_slots = 0;
    bool use_lds = false;
    /* some system has walls, the field or damping: the kernels' variants for them */
    bool any_walls = false;
    /* some system's radii differ (or grow): the sweep stages each disk's radius */
    bool any_poly = false;
    /* systems per workgroup of the HBM-resident loop (1, 2 or 4; more than
     * one only for calendars of depth <= 2 without walls, field or damping,
     * and not for the growing systems' launch) */
    int loop_wpb = 1;
};

struct edmd_ctx {
    edmd_batch* b;
};

static int choose_residency(edmd_batch* b);
static void set_sweep_variant(edmd_batch* b, int v);
static int default_sweep_variant(const edmd_batch* b);

template <class T>
static int dalloc(edmd_batch* b, T** p, size_t n) {
    if (n == 0) n = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(EDMD_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    b->allocs.push_back(q);
    *p = (T*)q;
    return EDMD_OK;
}

static void deriv_release(edmd_batch* b) {
    for (auto& c : b->ckpts) (void)hipFree(c.mem);
    b->ckpts.clear();
    if (b->ckpt_end) (void)hipFree(b->ckpt_end);
    b->ckpt_end = nullptr;
    for (void* p : b->dallocs) (void)hipFree(p);
    b->dallocs.clear();
    b->tape = nullptr;
    b->dv = b->dtl = b->otl = nullptr;
    b->dr = nullptr;
    b->deriv_on = false;
    b->seg = b->tape_cap = 0;
}

static void batch_free(edmd_batch* b) {
    if (!b) return;
    deriv_release(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->d_sys) (void)hipFree(b->d_sys);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

static double wrap_post(double v, double L) {
    /* PBCpostX/Y, EDMD.c:5948-5959 */
    if (v < 0) return v + L;
    if (v >= L) return v - L;
    return v;
}

static int coord_to_cell(double a, double fac, int nc) {
    /* coordToCell, EDMD.c:2098-2107, clamped to the grid */
    int c = (int)(a * fac);
    if (c >= nc) c = nc - 1;
    if (c < 0) c = 0;
    return c;
}

/* Largest number of disk centres a cell can hold: centres are at least the
 * smallest contact distance apart (less 1% for the overlap tolerance of
 * EDMD.c:2708), so at most (s/d + 1)^2 per square cell of side s, with 20%
 * slack for hexagonal packing. */
static int derived_capacity(double csx, double csy, double rmin) {
    double d = 2 * rmin * 0.99;
    double b = ceil((csx / d + 1) * (csy / d + 1) * 1.2);
    int cap = (int)b;
    if (cap < 4) cap = 4;
    cap = (cap + 3) & ~3;
    return cap;
}

static int setup_system(const edmd_system_desc* d, HostSys* h, int idx) {
    if (!d) return fail(EDMD_ERR_ARG, "system %d: null descriptor", idx);
    if (d->n < 2) return fail(EDMD_ERR_ARG, "system %d: need at least 2 disks", idx);
    if (!(d->Lx > 0) || !(d->Ly > 0)) return fail(EDMD_ERR_ARG, "system %d: box must be positive", idx);
    if (!d->x || !d->y || !d->vx || !d->vy || !d->rad || !d->m)
        return fail(EDMD_ERR_ARG, "system %d: missing particle arrays", idx);
    EdmdConst& c = h->c;
    memset(&c, 0, sizeof c);
    c.n = d->n;
    /* boxConstantHelper, EDMD.c:679-715 */
    c.Lx = d->Lx;
    c.Ly = d->Ly;
    c.halfLx = d->Lx / 2;
    c.halfLy = d->Ly / 2;
    c.ncx = (int)(c.halfLx);
    c.ncy = (int)(c.halfLy);
    if (c.ncx < 1 || c.ncy < 1) return fail(EDMD_ERR_ARG, "system %d: box side below 2 (no cell grid)", idx);
    if (c.ncx > 32767 || c.ncy > 32767 || (double)c.ncx * c.ncy > 2.0e9)
        return fail(EDMD_ERR_ARG, "system %d: cell grid too large", idx);
    c.csx = d->Lx / c.ncx;
    c.csy = d->Ly / c.ncy;
    if (!(d->vr >= 0) || !(d->einit > 0))
        return fail(EDMD_ERR_ARG, "system %d: growth rate must be >= 0 and einit > 0", idx);
    c.einit = d->einit;
    if ((d->wall_x != 0 && d->wall_x != 1) || (d->wall_y != 0 && d->wall_y != 1))
        return fail(EDMD_ERR_ARG, "system %d: wall flags must be 0 or 1", idx);
    if (!(d->res_wall >= 0) || !(d->res_tan == d->res_tan))
        return fail(EDMD_ERR_ARG, "system %d: bad wall restitution", idx);
    if (d->add_field != 0 && d->add_field != 1)
        return fail(EDMD_ERR_ARG, "system %d: add_field must be 0 or 1", idx);
    if (d->add_field && !(d->field < 0))
        return fail(EDMD_ERR_ARG, "system %d: the field must be negative", idx);
    if (d->damping != 0 && d->damping != 1)
        return fail(EDMD_ERR_ARG, "system %d: damping must be 0 or 1", idx);
    if (d->damping && !(d->gamma > 0))
        return fail(EDMD_ERR_ARG, "system %d: the damping rate must be positive", idx);
    c.field_on = d->add_field;
    c.field = d->field;
    c.damping = d->damping;
    c.gamma = d->gamma;
    if (d->wall_circle != 0 && d->wall_circle != 1)
        return fail(EDMD_ERR_ARG, "system %d: wall_circle must be 0 or 1", idx);
    if (d->mid_wall != 0 && d->mid_wall != 1)
        return fail(EDMD_ERR_ARG, "system %d: mid_wall must be 0 or 1", idx);
    c.walls = (d->wall_x ? 1 : 0) | (d->wall_y ? 2 : 0) | (d->wall_circle ? 4 : 0) |
              (d->mid_wall ? 8 : 0);
    c.sizeratio = d->sizeratio;
    if ((d->thermal_wall != 0 && d->thermal_wall != 1) || (d->shear_wall != 0 && d->shear_wall != 1))
        return fail(EDMD_ERR_ARG, "system %d: thermal_wall and shear_wall must be 0 or 1", idx);
    if (d->thermal_wall && !c.walls)
        return fail(EDMD_ERR_ARG, "system %d: thermal walls need walls", idx);
    if (d->shear_wall && !d->wall_y)
        return fail(EDMD_ERR_ARG, "system %d: shear walls need the walls normal to y", idx);
    c.thermal_wall = d->thermal_wall;
    c.shear_wall = d->shear_wall;
    c.res_wall = d->res_wall;
    c.res_tan = d->res_tan;
    c.mihx = (d->wall_x || d->wall_circle) ? __builtin_huge_val() : c.halfLx;
    c.mihy = (d->wall_y || d->wall_circle) ? __builtin_huge_val() : c.halfLy;
    c.dtp = 5.0 / (double)d->n; /* dtPaul, EDMD.c:712-714 */
    c.nlists = d->n;
    c.horizon = c.dtp * c.nlists;
    c.res = d->res;
    c.dtime = d->dtime;
    c.dtime_thermo = d->dtime_thermo;
    c.noise = d->noise;
    c.dtnoise = d->dtnoise;
    c.temperature = d->temperature;
    if (c.noise != 0 && c.noise != 2) return fail(EDMD_ERR_ARG, "system %d: noise must be 0 or 2", idx);
    if (d->trace_capacity < 0 || d->trace_capacity > (int64_t)1 << 31)
        return fail(EDMD_ERR_ARG, "system %d: bad trace capacity", idx);
    c.trace_cap = (int32_t)d->trace_capacity;
    c.thermo_cap = d->thermo_capacity > 0 ? d->thermo_capacity : 4096;
    double rmin = d->rad[0];
    /* equal radii bit for bit; a growth start's radii reach their targets
     * through different products (rad + vr*dt per disk), so they are not */
    c.mono = d->vr > 0 ? 0 : 1;
    for (int i = 0; i < d->n; i++) {
        if (memcmp(&d->rad[i], &d->rad[0], sizeof(double)) != 0) c.mono = 0;
        if (!(d->rad[i] > 0) || !(d->m[i] > 0))
            return fail(EDMD_ERR_ARG, "system %d: particle %d has non-positive radius or mass", idx, i);
        rmin = std::min(rmin, d->rad[i]);
    }
    int cap = d->cell_capacity > 0 ? d->cell_capacity
                                   : derived_capacity(c.csx, c.csy, rmin);
    if (cap > 64) cap = 64;
    c.cap = cap;
    /* calendar levels */
    c.lsize[0] = c.n;
    int l = 0;
    while (c.lsize[l] > 64) {
        if (l + 1 > EDMD_MAX_LEVELS) return fail(EDMD_ERR_ARG, "system %d: too many disks", idx);
        c.lsize[l + 1] = (c.lsize[l] + 63) / 64;
        l++;
    }
    c.nlev = l;
    return EDMD_OK;
}

static int upload_system(edmd_batch* b, int s, const edmd_system_desc* d) {
    HostSys& h = b->hs[s];
    const EdmdConst& c = h.c;
    const int n = c.n;
    const double cfx = 1 / c.csx, cfy = 1 / c.csy;
    std::vector<double> x(n), y(n);
    std::vector<int32_t> cx(n), cy(n);
    const size_t ncell = (size_t)c.ncx * c.ncy;
    std::vector<int32_t> cnt(ncell, 0), slot(ncell * c.cap, -1);
    for (int i = 0; i < n; i++) {
        /* load path wraps positions into the box, EDMD.c:1429-1430 */
        x[i] = wrap_post(d->x[i], c.Lx);
        y[i] = wrap_post(d->y[i], c.Ly);
        cx[i] = coord_to_cell(x[i], cfx, c.ncx);
        cy[i] = coord_to_cell(y[i], cfy, c.ncy);
        size_t cell = (size_t)cy[i] * c.ncx + cx[i];
        if (cnt[cell] >= c.cap)
            return fail(EDMD_ERR_CELL_OVERFLOW,
                        "system %d: cell (%d,%d) holds more than %d disks; raise cell_capacity",
                        s, cx[i], cy[i], c.cap);
        /* cellListInit adds particles in index order, EDMD.c:1916-1919 */
        slot[cell * c.cap + cnt[cell]++] = i;
    }
    const size_t p0 = h.p0;
    std::vector<EdmdPart> part(n);
    for (int i = 0; i < n; i++) {
        EdmdPart& r = part[i];
        memset(&r, 0, sizeof r);
        r.x = x[i];
        r.y = y[i];
        r.vx = d->vx[i];
        r.vy = d->vy[i];
        r.tl = 0;
        r.rad = d->vr > 0 ? 0.0 : d->rad[i]; /* growth starts from radius 0, EDMD.c:1765 */
        r.m = d->m[i];
        r.coll = 0;
        r.cx = (int16_t)cx[i];
        r.cy = (int16_t)cy[i];
    }
    HIPCHK(hipMemcpy(b->part + p0, part.data(), (size_t)n * sizeof(EdmdPart), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b->img + 2 * p0, 0, (size_t)n * 2 * sizeof(int32_t)));
    HIPCHK(hipMemcpy(b->ccnt + h.cell0, cnt.data(), ncell * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->cslot + h.slot0, slot.data(), ncell * c.cap * sizeof(int32_t),
                     hipMemcpyHostToDevice));
    if (b->mt) {
        /* init_genrand(rng_seed) of mersenne.c; the index 624 makes the
         * first draw twist */
        uint32_t w[625];
        w[0] = d->rng_seed;
        for (int k = 1; k < 624; k++) w[k] = 1812433253u * (w[k - 1] ^ (w[k - 1] >> 30)) + (uint32_t)k;
        w[624] = 624;
        HIPCHK(hipMemcpy(b->mt + (size_t)s * 625, w, sizeof w, hipMemcpyHostToDevice));
    }
    if (b->ptype) {
        std::vector<int32_t> ty(n, 0);
        if (d->type)
            for (int i = 0; i < n; i++) ty[i] = d->type[i];
        HIPCHK(hipMemcpy(b->ptype + p0, ty.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    if (b->vr) {
        std::vector<double> g(n, 0.0);
        if (d->vr > 0)
            for (int i = 0; i < n; i++) g[i] = d->vr * d->rad[i]; /* EDMD.c:1772 */
        HIPCHK(hipMemcpy(b->vr + p0, g.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    }
    double first_screen = d->first_screen, first_thermo = d->first_thermo;
    if (d->vr > 0) {
        /* output starts after the growth, EDMD.c:1959-1964 */
        if (first_screen < 1 / d->vr) first_screen = 1 / d->vr + first_screen;
        if (first_thermo < 1 / d->vr) first_thermo = 1 / d->vr + first_thermo;
    }
    EdmdState st;
    memset(&st, 0, sizeof st);
    st.t = 0;
    st.next_screen = first_screen;
    st.next_thermo = first_thermo + 0.000000001; /* EDMD.c:1970 */
    st.growing = d->vr > 0 ? 1 : 0;
    st.next_grow = d->vr > 0 ? 1 / d->vr : __builtin_huge_val(); /* addEventGrow(1/vr), EDMD.c:1976 */
    /* eventListInit's insertion order (EDMD.c:1970-1995): thermo, growth
     * stop, screenshot, noise; the particles' slots follow (sweep). */
    st.paul = 0;
    st.ctr = 0;
    st.win = 0;
    const double dtp = c.dtp;
    auto rank0 = [&](double tev) {
        int64_t r = (tev - st.paul < dtp) ? EDMD_RANK_DIRECT + st.ctr : EDMD_RANK_DIRECT - 1 - st.ctr;
        st.ctr++;
        return r;
    };
    /* the bucket of a special event, as addEventToQueue files it */
    auto win0 = [&](double tev) {
        const double dt = tev - st.paul;
        if (dt < dtp) return (int64_t)0;
        if (dt >= dtp * c.nlists) return (int64_t)-1;
        return (int64_t)(int)(dt / dtp);
    };
    st.win_thermo = win0(st.next_thermo);
    st.rank_thermo = rank0(st.next_thermo);
    st.win_grow = d->vr > 0 ? win0(st.next_grow) : INT64_MAX;
    st.rank_grow = d->vr > 0 ? rank0(st.next_grow) : INT64_MAX;
    st.win_screen = win0(st.next_screen);
    st.rank_screen = rank0(st.next_screen);
    st.next_noise = d->noise ? d->dtnoise : __builtin_huge_val();
    st.win_noise = d->noise ? win0(st.next_noise) : INT64_MAX;
    st.rank_noise = d->noise ? rank0(st.next_noise) : INT64_MAX;
    st.last_thermo = 0;
    st.ev_limit = INT64_MAX;
    st.coll_limit = ~(uint64_t)0;
    HIPCHK(hipMemcpy(b->st + s, &st, sizeof st, hipMemcpyHostToDevice));
    return EDMD_OK;
}

/* The sweep kernel's variant for the batch's systems. */
static int sweep_ext(const edmd_batch* b) {
    return (b->any_walls ? EDMD_SWEEP_EXT : 0) | (b->any_poly ? EDMD_SWEEP_POLY : 0);
}

static int batch_predict_all(edmd_batch* b) {
    HIPCHK(edmd_launch_sweep_tiled(b->d_sys, b->nsys, b->max_tiles, b->sweep_variant,
                                   sweep_ext(b), b->stream));
    if (b->any_grow) HIPCHK(edmd_launch_sweep_grow(b->d_sys, b->nsys, b->max_n, b->stream));
    HIPCHK(edmd_launch_sweep_count(b->d_sys, b->nsys, b->stream));
    HIPCHK(edmd_launch_tree(b->d_sys, b->nsys, b->max_lsize, b->max_nlev, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return EDMD_OK;
}

static int check_status(edmd_batch* b) {
    std::vector<EdmdState> st(b->nsys);
    HIPCHK(hipMemcpy(st.data(), b->st, b->nsys * sizeof(EdmdState), hipMemcpyDeviceToHost));
    for (int s = 0; s < b->nsys; s++) {
        if (st[s].status == EDMD_DEV_OVERLAP)
            return fail(EDMD_ERR_OVERLAP, "system %d: overlapping disks near particle %d at t=%.17g",
                        s, st[s].err_i, st[s].t);
        if (st[s].status == EDMD_DEV_CELL_OVERFLOW)
            return fail(EDMD_ERR_CELL_OVERFLOW, "system %d: cell %d overflow (particle %d) at t=%.17g",
                        s, st[s].err_j, st[s].err_i, st[s].t);
        if (st[s].status != 0) return fail(EDMD_ERR_STATE, "system %d: status %d", s, st[s].status);
    }
    return EDMD_OK;
}

extern "C" int edmd_batch_create(const edmd_system_desc* descs, int nsys, edmd_batch** out) {
    if (!out) return fail(EDMD_ERR_ARG, "null output handle");
    *out = nullptr;
    if (!descs || nsys <= 0) return fail(EDMD_ERR_ARG, "need at least one system");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(EDMD_ERR_NODEVICE, "no HIP device visible");
    edmd_batch* b = new edmd_batch();
    b->nsys = nsys;
    b->hs.resize(nsys);
    for (int s = 0; s < nsys; s++) {
        int rc = setup_system(descs + s, &b->hs[s], s);
        if (rc) { batch_free(b); return rc; }
        HostSys& h = b->hs[s];
        h.p0 = b->np;
        b->np += h.c.n;
        h.cell0 = b->ncell;
        b->ncell += (size_t)h.c.ncx * h.c.ncy;
        h.slot0 = b->nslot; /* a multiple of 4: cell rows are read in 16-byte chunks */
        b->nslot += ((size_t)h.c.ncx * h.c.ncy * h.c.cap + 3) & ~(size_t)3;
        for (int l = 1; l <= h.c.nlev; l++) {
            h.node0[l - 1] = b->nnode;
            b->nnode += h.c.lsize[l];
        }
        h.trace0 = b->ntrace;
        b->ntrace += h.c.trace_cap;
        h.thermo0 = b->nthermo;
        b->nthermo += h.c.thermo_cap;
        b->max_n = std::max(b->max_n, h.c.n);
        b->max_nlev = std::max(b->max_nlev, h.c.nlev);
        for (int l = 0; l <= h.c.nlev; l++) b->max_lsize[l] = std::max(b->max_lsize[l], h.c.lsize[l]);
    }
    set_sweep_variant(b, default_sweep_variant(b));
    int rc = EDMD_OK;
    const size_t np = b->np;
    if (!rc) rc = dalloc(b, &b->part, np);
    if (!rc) rc = dalloc(b, &b->ev, np);
    if (!rc) rc = dalloc(b, &b->leaf, np);
    if (!rc) rc = dalloc(b, &b->nodes, b->nnode);
    if (!rc) rc = dalloc(b, &b->img, 2 * np);
    if (!rc) rc = dalloc(b, &b->ccnt, b->ncell);
    if (!rc) rc = dalloc(b, &b->cslot, b->nslot);
    if (!rc) rc = dalloc(b, &b->st, (size_t)nsys);
    if (!rc) rc = dalloc(b, &b->step_peek, 1);
    if (!rc) rc = dalloc(b, &b->step_pair, 2);
    if (!rc) rc = dalloc(b, &b->trace, b->ntrace);
    if (!rc) rc = dalloc(b, &b->thermo, b->nthermo);
    if (!rc) rc = dalloc(b, &b->erank, 2 * np);
    if (!rc) rc = dalloc(b, &b->lrank, np);
    for (int s = 0; s < nsys; s++) b->any_grow |= descs[s].vr > 0;
    for (int s = 0; s < nsys; s++)
        b->any_walls |= b->hs[s].c.walls != 0 || b->hs[s].c.field_on || b->hs[s].c.damping;
    for (int s = 0; s < nsys; s++) b->any_poly |= !b->hs[s].c.mono;
    if (!rc && b->any_grow) rc = dalloc(b, &b->vr, np);
    bool any_mid = false;
    for (int s = 0; s < nsys; s++) any_mid |= descs[s].mid_wall != 0;
    if (!rc && any_mid) rc = dalloc(b, &b->ptype, np);
    bool any_thermal = false;
    for (int s = 0; s < nsys; s++) any_thermal |= descs[s].thermal_wall != 0;
    if (!rc && any_thermal) rc = dalloc(b, &b->mt, (size_t)nsys * 625);
    if (rc) { batch_free(b); return rc; }
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        batch_free(b);
        return fail(EDMD_ERR_HIP, "hipStreamCreate failed");
    }
    b->own_stream = true;
    b->sys.resize(nsys);
    for (int s = 0; s < nsys; s++) {
        HostSys& h = b->hs[s];
        EdmdSys& S = b->sys[s];
        memset(&S, 0, sizeof S);
        S.c = h.c;
        EdmdArrays& A = S.a;
        size_t p0 = h.p0;
        A.p = b->part + p0;
        A.e = b->ev + p0;
        A.leaf = b->leaf + p0;
        A.img = b->img + 2 * p0;
        for (int l = 1; l <= h.c.nlev; l++) A.lev[l - 1] = b->nodes + h.node0[l - 1];
        A.ccnt = b->ccnt + h.cell0;
        A.cslot = b->cslot + h.slot0;
        A.st = b->st + s;
        A.trace = b->trace + h.trace0;
        A.thermo = b->thermo + h.thermo0;
        A.vr = b->vr ? b->vr + p0 : nullptr;
        A.ptype = b->ptype ? b->ptype + p0 : nullptr;
        A.mt = b->mt ? b->mt + (size_t)s * 625 : nullptr;
        A.erank = b->erank + 2 * p0;
        A.lrank = b->lrank + p0;
        rc = upload_system(b, s, descs + s);
        if (rc) { batch_free(b); return rc; }
    }
    if (hipMalloc(&b->d_sys, nsys * sizeof(EdmdSys)) != hipSuccess) {
        b->d_sys = nullptr;
        batch_free(b);
        return fail(EDMD_ERR_NOMEM, "hipMalloc of system table failed");
    }
    if (hipMemcpy(b->d_sys, b->sys.data(), nsys * sizeof(EdmdSys), hipMemcpyHostToDevice) != hipSuccess) {
        batch_free(b);
        return fail(EDMD_ERR_HIP, "upload of system table failed");
    }
    rc = batch_predict_all(b);
    if (!rc) rc = check_status(b);
    if (!rc) rc = choose_residency(b);
    if (rc) { batch_free(b); return rc; }
    *out = b;
    return EDMD_OK;
}

extern "C" int edmd_batch_destroy(edmd_batch* b) {
    batch_free(b);
    return EDMD_OK;
}

extern "C" int edmd_batch_size(edmd_batch* b, int* nsys) {
    if (!b || !nsys) return fail(EDMD_ERR_ARG, "null argument");
    *nsys = b->nsys;
    return EDMD_OK;
}

extern "C" void* edmd_batch_stream(edmd_batch* b) { return b ? (void*)b->stream : nullptr; }

extern "C" int edmd_batch_set_stream(edmd_batch* b, void* stream) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    b->stream = (hipStream_t)stream;
    b->own_stream = false;
    return EDMD_OK;
}

/* Sweep tile shape: the grid covers the largest system's tiles. */
static void set_sweep_variant(edmd_batch* b, int v) {
    b->sweep_variant = v;
    b->max_tiles = 0;
    for (int s = 0; s < b->nsys; s++)
        b->max_tiles = std::max(b->max_tiles, edmd_sweep_tiles(b->hs[s].c.ncx, b->hs[s].c.ncy, v));
}

/* Default sweep shape from the disks per cell rho of the batch: a tile's
 * disks take ceil(rho TX TY / 64) wave passes of the candidate loop, so
 * the cost per disk is that many passes of 64 over rho TX TY, times a
 * staging term for the halo; shapes whose staged disks would come near
 * STAGE are skipped, and idle waves break ties. */
static int default_sweep_variant(const edmd_batch* b) {
    double nd = 0, nc = 0;
    for (int s = 0; s < b->nsys; s++) {
        nd += b->hs[s].c.n;
        nc += (double)b->hs[s].c.ncx * b->hs[s].c.ncy;
    }
    const double rho = nd / nc;
    int best = 0;
    double bc = 1e300;
    for (int v = 0; v < EDMD_SWEEP_VARIANTS; v++) {
        const EdmdSweepShape sh = edmd_sweep_shape(v);
        const double cells = (double)sh.tx * sh.ty, halo = (double)(sh.tx + 2) * (sh.ty + 2);
        const double owners = rho * cells;
        if (rho * halo * 1.15 > sh.stage || owners <= 0) continue;
        const double waves = ceil(owners / 64);
        const double idle = std::max(0.0, sh.threads / 64.0 - waves);
        const double cost = waves * 64 / owners * (1 + 0.1 * halo / cells) + 0.01 * idle;
        if (cost < bc) {
            bc = cost;
            best = v;
        }
    }
    return best;
}

/* Decides where the event loop keeps each system's arrays.  LDS residency
 * needs a calendar of depth <= 2 and every system's arrays (edmd_lds_layout)
 * plus the loop's static LDS within one workgroup's LDS; in auto mode it is
 * used when, in addition, the whole batch is resident on the device at once
 * (one wavefront per CU-slot, no second round of workgroups). */
static int choose_residency(edmd_batch* b) {
    b->lds_fit = false;
    b->lds_slots = 0;
    b->use_lds = false;
    b->lds_dyn[0] = b->lds_dyn[1] = 0;
    for (int s = 0; s < b->nsys; s++)
        for (int g = 0; g < 2; g++)
            b->lds_dyn[g] = std::max(b->lds_dyn[g], (size_t)edmd_lds_layout(&b->hs[s].c, g).total);
    if (b->max_nlev <= 2) {
        int dev = 0, maxblk = 0, ncu = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&maxblk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        bool fit = true;
        int64_t slots = INT64_MAX;
        for (int g = 0; g < (b->any_grow ? 2 : 1) && fit; g++) {
            int per_cu = 0;
            size_t stat = 0;
            hipError_t e = edmd_lds_occupancy(b->max_nlev, g, b->any_walls ? 1 : 0, b->lds_dyn[g],
                                              &per_cu, &stat);
            if (e != hipSuccess || per_cu < 1 || stat + b->lds_dyn[g] > (size_t)maxblk) {
                (void)hipGetLastError();
                fit = false;
                break;
            }
            slots = std::min(slots, (int64_t)per_cu * ncu);
        }
        if (fit) {
            b->lds_fit = true;
            b->lds_slots = slots;
        }
    }
    if (b->residency == EDMD_RES_LDS && !b->lds_fit)
        return fail(EDMD_ERR_ARG,
                    "LDS residency needs calendars of depth <= 2 and %zu bytes of LDS per workgroup",
                    b->lds_dyn[1] > b->lds_dyn[0] ? b->lds_dyn[1] : b->lds_dyn[0]);
    b->use_lds = b->residency == EDMD_RES_LDS ||
                 (b->residency == EDMD_RES_AUTO && b->lds_fit && b->nsys <= b->lds_slots);
    return EDMD_OK;
}

/* Whether the HBM-resident loop may hold several systems per workgroup. */
static bool packable(const edmd_batch* b) { return !b->any_walls && b->max_nlev <= 2; }

/* One launch of the event loop over every system (grow: the growing ones). */
static hipError_t launch_loop(edmd_batch* b, int grow, const EdmdRunArgs& ra) {
    const int walls = b->any_walls ? 1 : 0;
    if (b->use_lds)
        return edmd_launch_run_lds(b->d_sys, b->nsys, b->max_nlev, grow, walls,
                                   b->lds_dyn[grow ? 1 : 0], ra, b->stream);
    const int wpb = (!grow && packable(b)) ? b->loop_wpb : 1;
    return edmd_launch_run(b->d_sys, b->nsys, b->max_nlev, grow, walls, wpb, ra, b->stream);
}

extern "C" int edmd_batch_set_loop_waves(edmd_batch* b, int systems_per_group) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (systems_per_group != 1 && systems_per_group != 2 && systems_per_group != 4)
        return fail(EDMD_ERR_ARG, "systems per workgroup must be 1, 2 or 4 (got %d)",
                    systems_per_group);
    if (systems_per_group > 1 && !packable(b))
        return fail(EDMD_ERR_ARG,
                    "several systems per workgroup need calendars of depth <= 2 and no walls, "
                    "field or damping");
    b->loop_wpb = systems_per_group;
    return EDMD_OK;
}

extern "C" int edmd_batch_get_loop_waves(edmd_batch* b, int* systems_per_group) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (systems_per_group) *systems_per_group = b->loop_wpb;
    return EDMD_OK;
}

extern "C" int edmd_batch_set_residency(edmd_batch* b, int mode) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (mode != EDMD_RES_AUTO && mode != EDMD_RES_HBM && mode != EDMD_RES_LDS)
        return fail(EDMD_ERR_ARG, "unknown residency mode %d", mode);
    const int old = b->residency;
    b->residency = mode;
    int rc = choose_residency(b);
    if (rc) {
        b->residency = old;
        (void)choose_residency(b);
    }
    return rc;
}

extern "C" int edmd_batch_get_residency(edmd_batch* b, int* in_lds, int64_t* lds_bytes,
                                        int64_t* lds_slots) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (in_lds) *in_lds = b->use_lds ? 1 : 0;
    if (lds_bytes) *lds_bytes = (int64_t)(b->any_grow ? b->lds_dyn[1] : b->lds_dyn[0]);
    if (lds_slots) *lds_slots = b->lds_fit ? b->lds_slots : 0;
    return EDMD_OK;
}

extern "C" int edmd_batch_set_sweep_shape(edmd_batch* b, int shape) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (shape < 0 || shape >= EDMD_SWEEP_VARIANTS)
        return fail(EDMD_ERR_ARG, "sweep shape %d out of range [0, %d)", shape, EDMD_SWEEP_VARIANTS);
    set_sweep_variant(b, shape);
    return EDMD_OK;
}

extern "C" int edmd_batch_get_sweep_shape(edmd_batch* b, int* shape, int* tile_x, int* tile_y,
                                          int* threads) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    const EdmdSweepShape sh = edmd_sweep_shape(b->sweep_variant);
    if (shape) *shape = b->sweep_variant;
    if (tile_x) *tile_x = sh.tx;
    if (tile_y) *tile_y = sh.ty;
    if (threads) *threads = sh.threads;
    return EDMD_OK;
}

static int upload_table(edmd_batch* b);

/* The sweep's conservative root screen for every system of the batch; the
 * counters start again from zero. */
extern "C" int edmd_batch_set_screen(edmd_batch* b, int mode) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (mode < 0 || mode > 2) return fail(EDMD_ERR_ARG, "screen mode %d not in {0, 1, 2}", mode);
    std::vector<EdmdState> st(b->nsys);
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(st.data(), b->st, b->nsys * sizeof(EdmdState), hipMemcpyDeviceToHost));
    for (int s = 0; s < b->nsys; s++) {
        st[s].scr_marks = st[s].scr_solved = st[s].scr_miss = 0;
        b->hs[s].c.screen = mode;
        b->sys[s].c.screen = mode;
    }
    HIPCHK(hipMemcpy(b->st, st.data(), b->nsys * sizeof(EdmdState), hipMemcpyHostToDevice));
    return upload_table(b);
}

extern "C" int edmd_batch_screen_stats(edmd_batch* b, uint64_t* marks, uint64_t* solved,
                                       uint64_t* misses) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    std::vector<EdmdState> st(b->nsys);
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(st.data(), b->st, b->nsys * sizeof(EdmdState), hipMemcpyDeviceToHost));
    uint64_t m = 0, so = 0, mi = 0;
    for (int s = 0; s < b->nsys; s++) {
        m += st[s].scr_marks;
        so += st[s].scr_solved;
        mi += st[s].scr_miss;
    }
    if (marks) *marks = m;
    if (solved) *solved = so;
    if (misses) *misses = mi;
    return EDMD_OK;
}

static int check_status(edmd_batch* b);
static size_t ckpt_bytes(edmd_batch* b);
static int ckpt_copy(edmd_batch* b, void* mem, bool save);

/* Times every sweep shape and, when the systems fit LDS, a trial run in
 * each residency, on the batch itself: its state is saved first and put
 * back after every trial, so the batch continues exactly as if untuned. */
extern "C" int edmd_batch_autotune(edmd_batch* b, int sweep_reps, int64_t trial_events,
                                   edmd_tune_report* out) {
    static_assert(EDMD_TUNE_SWEEP_SHAPES == EDMD_SWEEP_VARIANTS, "sweep shape count");
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (sweep_reps < 0 || trial_events < 0) return fail(EDMD_ERR_ARG, "negative trial sizes");
    if (b->deriv_on) return fail(EDMD_ERR_STATE, "autotuning is not available while taping");
    edmd_tune_report r;
    memset(&r, 0, sizeof r);
    for (int v = 0; v < EDMD_TUNE_SWEEP_SHAPES; v++) r.sweep_ms[v] = -1;
    r.loop_ms_hbm = r.loop_ms_lds = r.loop_ms_packed[0] = r.loop_ms_packed[1] = -1;
    HIPCHK(hipStreamSynchronize(b->stream));
    void* snap = nullptr;
    if (hipMalloc(&snap, ckpt_bytes(b)) != hipSuccess)
        return fail(EDMD_ERR_NOMEM, "hipMalloc of %zu bytes for the tuning snapshot failed",
                    ckpt_bytes(b));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int rc = e == hipSuccess ? ckpt_copy(b, snap, true) : fail(EDMD_ERR_HIP, "hipEventCreate failed");
    float ms = 0;
    if (!rc && sweep_reps > 0) {
        const int cur = b->sweep_variant;
        int best = cur;
        double bt = 1e300;
        for (int v = 0; v < EDMD_SWEEP_VARIANTS && e == hipSuccess; v++) {
            set_sweep_variant(b, v);
            e = edmd_launch_sweep_tiled(b->d_sys, b->nsys, b->max_tiles, v, sweep_ext(b),
                                        b->stream);
            if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
            for (int k = 0; k < sweep_reps && e == hipSuccess; k++)
                e = edmd_launch_sweep_tiled(b->d_sys, b->nsys, b->max_tiles, v, sweep_ext(b),
                                            b->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess) {
                r.sweep_ms[v] = ms / sweep_reps;
                if (r.sweep_ms[v] < bt) {
                    bt = r.sweep_ms[v];
                    best = v;
                }
            }
        }
        /* the shape in use stays unless another is more than 3% faster:
         * shapes within timing noise of each other do not flip */
        if (e == hipSuccess && r.sweep_ms[cur] > 0 && bt > 0.97 * r.sweep_ms[cur]) best = cur;
        set_sweep_variant(b, best);
        if (e != hipSuccess) rc = fail(EDMD_ERR_HIP, "sweep trial: %s", hipGetErrorString(e));
    }
    if (!rc) rc = ckpt_copy(b, snap, false);
    if (!rc && trial_events > 0 && (b->lds_fit || packable(b))) {
        EdmdRunArgs ra;
        memset(&ra, 0, sizeof ra);
        ra.max_events = INT64_MAX; /* the per-system trial bound applies */
        ra.max_coll = ~(uint64_t)0;
        ra.tmax = 1e300;
        /* trials: HBM with 1, 2 and 4 systems per workgroup, then LDS */
        const int wpbs[4] = {1, 2, 4, 1};
        double t[4] = {-1, -1, -1, -1};
        for (int m = 0; m < 4 && !rc; m++) {
            if (m == 3 ? !b->lds_fit : (m > 0 && !packable(b))) continue;
            b->use_lds = m == 3;
            b->loop_wpb = wpbs[m];
            e = edmd_launch_epoch_begin(b->d_sys, b->nsys, trial_events, ~(uint64_t)0, b->stream);
            if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
            if (e == hipSuccess && b->any_grow) e = launch_loop(b, 1, ra);
            if (e == hipSuccess) e = launch_loop(b, 0, ra);
            if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) rc = fail(EDMD_ERR_HIP, "loop trial: %s", hipGetErrorString(e));
            t[m] = ms;
            if (!rc) rc = check_status(b);
            int rc2 = ckpt_copy(b, snap, false);
            if (!rc) rc = rc2;
        }
        int best = 0;
        for (int m = 1; m < 4; m++)
            if (t[m] > 0 && t[m] < t[best]) best = m;
        /* one system per workgroup from HBM stays unless another
         * configuration is more than 3% faster: the trial runs the first
         * events of the batch, and differences within that margin have not
         * held over longer runs */
        if (best != 0 && t[0] > 0 && t[best] > 0.97 * t[0]) best = 0;
        b->loop_wpb = wpbs[best];
        if (!rc) {
            r.loop_ms_hbm = t[0];
            r.loop_ms_packed[0] = t[1];
            r.loop_ms_packed[1] = t[2];
            r.loop_ms_lds = t[3];
            b->residency = best == 3 ? EDMD_RES_LDS : EDMD_RES_HBM;
        }
        int rc3 = choose_residency(b);
        if (!rc) rc = rc3;
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess && !rc)
        rc = fail(EDMD_ERR_HIP, "tuning: stream synchronisation failed");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(snap);
    r.sweep_shape = b->sweep_variant;
    r.in_lds = b->use_lds ? 1 : 0;
    r.loop_systems_per_group = b->loop_wpb;
    if (out) *out = r;
    return rc;
}

static int batch_run_mask(edmd_batch* b, int64_t max_events, uint64_t max_collisions, double tmax,
                          int stop_mask) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    EdmdRunArgs ra;
    memset(&ra, 0, sizeof ra);
    ra.max_events = max_events;
    ra.max_coll = max_collisions;
    ra.tmax = tmax;
    ra.stop_mask = stop_mask & ~(1 << EDMD_EV_GROWSTOP);
    if (b->any_grow) {
        /* growing systems first: they run until GROWSTOP (then continue in
         * the Normal launch below) or until the bounds */
        HIPCHK(launch_loop(b, 1, ra));
        HIPCHK(hipStreamSynchronize(b->stream));
        int rc = check_status(b);
        if (rc) return rc;
        std::vector<EdmdState> st(b->nsys);
        HIPCHK(hipMemcpy(st.data(), b->st, b->nsys * sizeof(EdmdState), hipMemcpyDeviceToHost));
        bool still = false;
        for (int s = 0; s < b->nsys; s++) still |= st[s].growing != 0;
        b->any_grow = still;
    }
    HIPCHK(launch_loop(b, 0, ra));
    HIPCHK(hipStreamSynchronize(b->stream));
    return check_status(b);
}

extern "C" int edmd_batch_run(edmd_batch* b, int64_t max_events, uint64_t max_collisions,
                              double tmax) {
    return batch_run_mask(b, max_events, max_collisions, tmax, 0);
}

extern "C" int edmd_batch_run_epochs(edmd_batch* b, int nepochs, int64_t events_per_epoch,
                                     uint64_t collisions_per_epoch, double tmax,
                                     edmd_epoch_row* rows) {
    static_assert(sizeof(edmd_epoch_row) == sizeof(EdmdEpochRow), "epoch row layout");
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (nepochs <= 0 || events_per_epoch <= 0 || collisions_per_epoch == 0)
        return fail(EDMD_ERR_ARG, "epochs need nepochs > 0 and positive bounds");
    if (b->deriv_on) return fail(EDMD_ERR_STATE, "epochs are not available while taping");
    const int nsys = b->nsys;
    EdmdEpochRow* d_rows = nullptr;
    if (hipMalloc(&d_rows, (size_t)nepochs * nsys * sizeof(EdmdEpochRow)) != hipSuccess)
        return fail(EDMD_ERR_NOMEM, "hipMalloc of %d epoch rows failed", nepochs * nsys);
    EdmdRunArgs ra;
    memset(&ra, 0, sizeof ra);
    ra.max_events = INT64_MAX; /* the per-system epoch bounds apply */
    ra.max_coll = ~(uint64_t)0;
    ra.tmax = tmax;
    /* one graph holds every epoch: bounds, the growing systems' launch (a
     * no-op for systems that are not growing), the Normal launch and the
     * counters; captured once, launched once */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    for (int k = 0; e == hipSuccess && k < nepochs; k++) {
        e = edmd_launch_epoch_begin(b->d_sys, nsys, events_per_epoch, collisions_per_epoch,
                                    b->stream);
        if (e == hipSuccess && b->any_grow)
            e = launch_loop(b, 1, ra);
        if (e == hipSuccess) e = launch_loop(b, 0, ra);
        if (e == hipSuccess)
            e = edmd_launch_epoch_end(b->d_sys, nsys, d_rows + (size_t)k * nsys, b->stream);
    }
    hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphLaunch(exec, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e == hipSuccess && rows)
        e = hipMemcpy(rows, d_rows, (size_t)nepochs * nsys * sizeof(EdmdEpochRow),
                      hipMemcpyDeviceToHost);
    (void)hipFree(d_rows);
    /* lift the epoch bounds again */
    if (e == hipSuccess)
        e = edmd_launch_epoch_begin(b->d_sys, nsys, INT64_MAX, ~(uint64_t)0, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail(EDMD_ERR_HIP, "epoch graph: %s", hipGetErrorString(e));
    if (b->any_grow) {
        std::vector<EdmdState> st(nsys);
        HIPCHK(hipMemcpy(st.data(), b->st, nsys * sizeof(EdmdState), hipMemcpyDeviceToHost));
        bool still = false;
        for (int s = 0; s < nsys; s++) still |= st[s].growing != 0;
        b->any_grow = still;
    }
    return check_status(b);
}

extern "C" int edmd_batch_sweep_predict(edmd_batch* b) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    HIPCHK(edmd_launch_sync_all(b->d_sys, b->nsys, b->max_n, b->stream));
    int rc = batch_predict_all(b);
    if (rc) return rc;
    return check_status(b);
}

/* Times the prediction sweep on the batch's current state.  The sweep
 * rewrites every event slot, leaf and rank without rebuilding the calendar
 * above the leaves, so the state is saved first and put back afterwards:
 * the batch continues exactly as if the call had not been made. */
extern "C" int edmd_batch_time_sweep(edmd_batch* b, int reps, double* avg_ms) {
    if (!b || reps <= 0 || !avg_ms) return fail(EDMD_ERR_ARG, "bad arguments");
    if (b->deriv_on) return fail(EDMD_ERR_STATE, "sweep timing is not available while taping");
    HIPCHK(hipStreamSynchronize(b->stream));
    void* snap = nullptr;
    if (hipMalloc(&snap, ckpt_bytes(b)) != hipSuccess)
        return fail(EDMD_ERR_NOMEM, "hipMalloc of %zu bytes for the timing snapshot failed",
                    ckpt_bytes(b));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int rc = e == hipSuccess ? ckpt_copy(b, snap, true) : fail(EDMD_ERR_HIP, "hipEventCreate failed");
    float ms = 0;
    if (!rc) {
        e = hipEventRecord(e0, b->stream);
        for (int r = 0; r < reps && e == hipSuccess; r++)
            e = edmd_launch_sweep_tiled(b->d_sys, b->nsys, b->max_tiles, b->sweep_variant,
                                        sweep_ext(b), b->stream);
        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) rc = fail(EDMD_ERR_HIP, "sweep timing: %s", hipGetErrorString(e));
    }
    if (!rc) rc = check_status(b);
    int rc2 = ckpt_copy(b, snap, false);
    if (!rc) rc = rc2;
    if (hipStreamSynchronize(b->stream) != hipSuccess && !rc)
        rc = fail(EDMD_ERR_HIP, "sweep timing: stream synchronisation failed");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(snap);
    if (!rc) *avg_ms = ms / reps;
    return rc;
}

static int sys_check(edmd_batch* b, int s) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (s < 0 || s >= b->nsys) return fail(EDMD_ERR_ARG, "system index %d out of range", s);
    return EDMD_OK;
}

template <class T>
static int d2h(T* dst, const T* src, size_t n) {
    if (!dst) return EDMD_OK;
    HIPCHK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToHost));
    return EDMD_OK;
}

extern "C" int edmd_batch_get_state(edmd_batch* b, int s, double* x, double* y, double* vx,
                                    double* vy, double* t_local, uint32_t* coll, int32_t* cx,
                                    int32_t* cy) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    size_t n = h.c.n;
    std::vector<EdmdPart> part(n);
    if ((rc = d2h(part.data(), b->part + h.p0, n))) return rc;
    for (size_t i = 0; i < n; i++) {
        const EdmdPart& r = part[i];
        if (x) x[i] = r.x;
        if (y) y[i] = r.y;
        if (vx) vx[i] = r.vx;
        if (vy) vy[i] = r.vy;
        if (t_local) t_local[i] = r.tl;
        if (coll) coll[i] = r.coll;
        if (cx) cx[i] = r.cx;
        if (cy) cy[i] = r.cy;
    }
    return EDMD_OK;
}

extern "C" int edmd_batch_get_radii(edmd_batch* b, int s, double* rad) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    if (!rad) return fail(EDMD_ERR_ARG, "null output");
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    size_t n = h.c.n;
    std::vector<EdmdPart> part(n);
    if ((rc = d2h(part.data(), b->part + h.p0, n))) return rc;
    for (size_t i = 0; i < n; i++) rad[i] = part[i].rad;
    return EDMD_OK;
}

extern "C" int edmd_batch_get_events(edmd_batch* b, int s, double* t_coll, int32_t* partner,
                                     double* t_cross, int32_t* dir) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    size_t n = h.c.n;
    std::vector<EdmdEv> ev(n);
    if ((rc = d2h(ev.data(), b->ev + h.p0, n))) return rc;
    for (size_t i = 0; i < n; i++) {
        if (t_coll) t_coll[i] = ev[i].tcoll;
        if (partner) partner[i] = ev[i].partner;
        if (t_cross) t_cross[i] = ev[i].tcross;
        if (dir) dir[i] = ev[i].dir;
    }
    return EDMD_OK;
}

extern "C" int edmd_batch_get_calendar(edmd_batch* b, int s, double* key, int32_t* idx,
                                       int32_t* aux, int64_t cap, int64_t* nnodes,
                                       int64_t* slot_rank, int64_t* leaf_rank) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    const size_t n = h.c.n;
    int64_t total = 0;
    for (int l = 0; l <= h.c.nlev; l++) total += h.c.lsize[l];
    if (nnodes) *nnodes = total;
    if (key || idx || aux) {
        if (cap < total) return fail(EDMD_ERR_ARG, "calendar has %lld nodes, cap %lld",
                                     (long long)total, (long long)cap);
        std::vector<EdmdNode> nd(total);
        size_t off = 0;
        for (int l = 0; l <= h.c.nlev; l++) {
            const EdmdNode* src = l == 0 ? b->leaf + h.p0 : b->nodes + h.node0[l - 1];
            if ((rc = d2h(nd.data() + off, src, (size_t)h.c.lsize[l]))) return rc;
            off += h.c.lsize[l];
        }
        for (int64_t k = 0; k < total; k++) {
            if (key) key[k] = nd[k].key;
            if (idx) idx[k] = nd[k].idx;
            if (aux) aux[k] = nd[k].aux;
        }
    }
    if (slot_rank && (rc = d2h(slot_rank, b->erank + 2 * h.p0, 2 * n))) return rc;
    if (leaf_rank && (rc = d2h(leaf_rank, b->lrank + h.p0, n))) return rc;
    return EDMD_OK;
}

extern "C" int edmd_batch_get_observables(edmd_batch* b, int s, edmd_observables* o) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    if (!o) return fail(EDMD_ERR_ARG, "null output");
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    EdmdState st;
    HIPCHK(hipMemcpy(&st, b->st + s, sizeof st, hipMemcpyDeviceToHost));
    size_t n = h.c.n;
    std::vector<EdmdPart> part(n);
    if ((rc = d2h(part.data(), b->part + h.p0, n))) return rc;
    /* physicalQ, EDMD.c:5968-6013, summed in particle order */
    double E = 0, px = 0, py = 0;
    for (size_t i = 0; i < n; i++) {
        const EdmdPart& r = part[i];
        px += r.vx * r.m;
        py += r.vy * r.m;
        double a = r.vx;
        E += 0.5 * r.m * (a * a + r.vy * r.vy);
    }
    o->t = st.t;
    o->ncol = st.ncol;
    o->ncross = st.ncross;
    o->nevents = st.nevents;
    o->vir_xx = st.vir[0];
    o->vir_yy = st.vir[1];
    o->vir_xy = st.vir[2];
    o->vir_yx = st.vir[3];
    o->E = E;
    o->px = px;
    o->py = py;
    o->status = st.status;
    o->stop_reason = st.stop_reason;
    o->err_i = st.err_i;
    o->err_j = st.err_j;
    return EDMD_OK;
}

extern "C" int edmd_batch_get_trace(edmd_batch* b, int s, edmd_trace_rec* out, int64_t cap,
                                    int64_t* n) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    EdmdState st;
    HIPCHK(hipMemcpy(&st, b->st + s, sizeof st, hipMemcpyDeviceToHost));
    int64_t have = std::min<int64_t>(st.ntrace, h.c.trace_cap);
    int64_t k = std::min<int64_t>(have, cap);
    static_assert(sizeof(edmd_trace_rec) == sizeof(EdmdTraceRec), "trace layout");
    if (k > 0 && out)
        HIPCHK(hipMemcpy(out, b->trace + h.trace0, (size_t)k * sizeof(EdmdTraceRec),
                         hipMemcpyDeviceToHost));
    if (n) *n = k;
    int64_t zero = 0;
    HIPCHK(hipMemcpy((char*)(b->st + s) + offsetof(EdmdState, ntrace), &zero, sizeof zero,
                     hipMemcpyHostToDevice));
    return EDMD_OK;
}

extern "C" int edmd_batch_get_thermo(edmd_batch* b, int s, edmd_thermo_row* out, int cap, int* n) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    EdmdState st;
    HIPCHK(hipMemcpy(&st, b->st + s, sizeof st, hipMemcpyDeviceToHost));
    int have = std::min<int>(st.nthermo, h.c.thermo_cap);
    int k = std::min(have, cap);
    static_assert(sizeof(edmd_thermo_row) == sizeof(EdmdThermoRow), "thermo layout");
    if (k > 0 && out)
        HIPCHK(hipMemcpy(out, b->thermo + h.thermo0, (size_t)k * sizeof(EdmdThermoRow),
                         hipMemcpyDeviceToHost));
    if (n) *n = have;
    return EDMD_OK;
}

/* ------------------------------------------------------------ single system */

extern "C" int edmd_ctx_create(const edmd_system_desc* desc, edmd_ctx** out) {
    if (!out) return fail(EDMD_ERR_ARG, "null output handle");
    *out = nullptr;
    edmd_batch* b = nullptr;
    int rc = edmd_batch_create(desc, 1, &b);
    if (rc) return rc;
    edmd_ctx* c = new edmd_ctx;
    c->b = b;
    *out = c;
    return EDMD_OK;
}

extern "C" int edmd_ctx_create_from_dump(const char* path, const edmd_system_desc* params,
                                         edmd_ctx** out) {
    if (!out) return fail(EDMD_ERR_ARG, "null output handle");
    *out = nullptr;
    int n = 0;
    double Lx = 0, Ly = 0;
    int rc = edmd_dump_info(path, &n, &Lx, &Ly);
    if (rc) return rc;
    std::vector<double> x(n), y(n), vx(n), vy(n), rad(n), m(n);
    rc = edmd_dump_read(path, n, x.data(), y.data(), vx.data(), vy.data(), rad.data(), m.data());
    if (rc) return rc;
    edmd_system_desc d;
    if (params) d = *params;
    else edmd_default_desc(&d);
    d.n = n;
    d.Lx = Lx;
    d.Ly = Ly;
    d.x = x.data(); d.y = y.data(); d.vx = vx.data(); d.vy = vy.data();
    d.rad = rad.data(); d.m = m.data();
    return edmd_ctx_create(&d, out);
}

extern "C" int edmd_run_traced(edmd_ctx* c, int64_t max_events, uint64_t max_collisions,
                               double tmax, edmd_trace_cb cb, void* user) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    if (!cb) return fail(EDMD_ERR_ARG, "null trace callback");
    const int64_t cap = c->b->hs[0].c.trace_cap;
    if (cap <= 0) return fail(EDMD_ERR_ARG, "edmd_run_traced needs desc.trace_capacity > 0");
    std::vector<edmd_trace_rec> buf((size_t)cap);
    int64_t n = 0;
    int rc = edmd_get_trace(c, buf.data(), cap, &n); /* drop what earlier runs left */
    if (rc) return rc;
    for (;;) {
        edmd_observables o;
        if ((rc = edmd_get_observables(c, &o))) return rc;
        const int64_t done = (int64_t)o.nevents;
        if (done >= max_events || o.ncol >= max_collisions || !(o.t <= tmax)) return EDMD_OK;
        /* at most `cap` events per chunk, so the trace buffer never drops a record */
        const int64_t bound = max_events - done > cap ? done + cap : max_events;
        rc = edmd_run(c, bound, max_collisions, tmax);
        int64_t got = 0;
        int rc2 = edmd_get_trace(c, buf.data(), cap, &got);
        if (got > 0 && cb(buf.data(), got, user) != 0) return rc ? rc : rc2;
        if (rc) return rc;
        if (rc2) return rc2;
        if ((rc = edmd_get_observables(c, &o))) return rc;
        if (o.stop_reason != EDMD_STOP_EVENTS) return EDMD_OK; /* time, collisions or empty */
    }
}

extern "C" int edmd_ctx_destroy(edmd_ctx* c) {
    if (!c) return EDMD_OK;
    batch_free(c->b);
    delete c;
    return EDMD_OK;
}

extern "C" int edmd_sweep_predict(edmd_ctx* c) {
    return c ? edmd_batch_sweep_predict(c->b) : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_load_dump(const char* path, edmd_ctx** out) {
    return edmd_ctx_create_from_dump(path, nullptr, out);
}

/* Derivatives of a context's single system: the batch calls on its batch. */
extern "C" int edmd_adjoint_begin(edmd_ctx* c, int64_t checkpoint_every, int64_t tape_capacity) {
    return c ? edmd_batch_deriv_begin(c->b, checkpoint_every, tape_capacity)
             : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_adjoint_run(edmd_ctx* c, int64_t max_events, uint64_t max_collisions,
                                double tmax) {
    return c ? edmd_batch_deriv_run(c->b, max_events, max_collisions, tmax)
             : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_adjoint_tangent(edmd_ctx* c, const double* tan_in, double* tan_out,
                                    double* dw, double* dtlast) {
    return c ? edmd_batch_tangent(c->b, tan_in, tan_out, dw, dtlast)
             : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_adjoint_gradient(edmd_ctx* c, double lw, double lt, const double* lam_end,
                                     double* grad) {
    return c ? edmd_batch_adjoint(c->b, &lw, &lt, lam_end, grad)
             : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_adjoint_values(edmd_ctx* c, double* w, double* tlast, int64_t* records) {
    return c ? edmd_batch_deriv_values(c->b, 0, w, tlast, records)
             : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_adjoint_end(edmd_ctx* c) {
    return c ? edmd_batch_deriv_end(c->b) : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_run(edmd_ctx* c, int64_t max_events, uint64_t max_collisions, double tmax) {
    return c ? edmd_batch_run(c->b, max_events, max_collisions, tmax)
             : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_get_state(edmd_ctx* c, double* x, double* y, double* vx, double* vy,
                              double* t_local, uint32_t* coll, int32_t* cx, int32_t* cy) {
    return c ? edmd_batch_get_state(c->b, 0, x, y, vx, vy, t_local, coll, cx, cy)
             : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_get_radii(edmd_ctx* c, double* rad) {
    return c ? edmd_batch_get_radii(c->b, 0, rad) : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_get_events(edmd_ctx* c, double* t_cross, int32_t* dir, double* t_coll,
                               int32_t* partner, uint32_t* coll_actual) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    int rc = edmd_batch_get_events(c->b, 0, t_coll, partner, t_cross, dir);
    if (rc || !coll_actual) return rc;
    size_t n = c->b->hs[0].c.n;
    std::vector<EdmdEv> ev(n);
    if ((rc = d2h(ev.data(), c->b->ev + c->b->hs[0].p0, n))) return rc;
    for (size_t i = 0; i < n; i++) coll_actual[i] = ev[i].cact;
    return EDMD_OK;
}

extern "C" int edmd_get_observables(edmd_ctx* c, edmd_observables* o) {
    return c ? edmd_batch_get_observables(c->b, 0, o) : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_get_trace(edmd_ctx* c, edmd_trace_rec* out, int64_t cap, int64_t* n) {
    return c ? edmd_batch_get_trace(c->b, 0, out, cap, n) : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_get_thermo(edmd_ctx* c, edmd_thermo_row* out, int cap, int* n) {
    return c ? edmd_batch_get_thermo(c->b, 0, out, cap, n) : fail(EDMD_ERR_ARG, "null context");
}

extern "C" int edmd_get_grid(edmd_ctx* c, int* ncx, int* ncy, int* cap, double* csx, double* csy) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    const EdmdConst& k = c->b->hs[0].c;
    if (ncx) *ncx = k.ncx;
    if (ncy) *ncy = k.ncy;
    if (cap) *cap = k.cap;
    if (csx) *csx = k.csx;
    if (csy) *csy = k.csy;
    return EDMD_OK;
}

extern "C" void* edmd_stream(edmd_ctx* c) { return c ? (void*)c->b->stream : nullptr; }

extern "C" int edmd_run_until_kind(edmd_ctx* c, int64_t max_events, uint64_t max_collisions,
                                   double tmax, int stop_mask, int* stop_reason, int* last_kind) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    int rc = batch_run_mask(c->b, max_events, max_collisions, tmax, stop_mask);
    EdmdState st;
    HIPCHK(hipMemcpy(&st, c->b->st, sizeof st, hipMemcpyDeviceToHost));
    if (stop_reason) *stop_reason = st.stop_reason;
    if (last_kind) *last_kind = st.last_type;
    return rc;
}

extern "C" int edmd_peek_event(edmd_ctx* c, double* t, int* kind, int* i, int* j,
                               uint32_t* coll_actual) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    EdmdPeek* d = c->b->step_peek;
    EdmdPeek h;
    hipError_t e = edmd_launch_peek(c->b->d_sys, 1, d, c->b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c->b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->b->stream);
    if (e != hipSuccess) return fail(EDMD_ERR_HIP, "peek failed: %s", hipGetErrorString(e));
    if (t) *t = h.t;
    if (kind) *kind = h.type;
    if (i) *i = h.i;
    if (j) *j = h.j;
    if (coll_actual) *coll_actual = h.coll_actual;
    return EDMD_OK;
}

static int particle_check(edmd_ctx* c, int i) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    if (i < 0 || i >= c->b->hs[0].c.n) return fail(EDMD_ERR_ARG, "particle %d out of range", i);
    return EDMD_OK;
}

extern "C" int edmd_collision_time(edmd_ctx* c, int i, int j, double* dt) {
    int rc = particle_check(c, i);
    if (!rc) rc = particle_check(c, j);
    if (rc) return rc;
    double* d = c->b->step_pair;
    double h[2] = {0, 0};
    hipError_t e = edmd_launch_pair_time(c->b->d_sys, i, j, d, c->b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, c->b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->b->stream);
    if (e != hipSuccess) return fail(EDMD_ERR_HIP, "pair time failed: %s", hipGetErrorString(e));
    if (dt) *dt = h[0];
    if (h[1] != 0) return fail(EDMD_ERR_OVERLAP, "particles %d and %d overlap", i, j);
    return EDMD_OK;
}

extern "C" int edmd_repredict(edmd_ctx* c, int i, int which) {
    int rc = particle_check(c, i);
    if (rc) return rc;
    if (which < 1 || which > 3) return fail(EDMD_ERR_ARG, "which must be 1, 2 or 3");
    HIPCHK(edmd_launch_repredict(c->b->d_sys, c->b->max_nlev, i, which, c->b->stream));
    HIPCHK(hipStreamSynchronize(c->b->stream));
    return check_status(c->b);
}

/* addEventToQueue / removeEventFromQueue on one eventList slot. */
static int queue_edit(edmd_ctx* c, int kind, int i, int j, double t, uint32_t cact, int mode) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    const int n = c->b->hs[0].c.n;
    switch (kind) {
        case EDMD_EV_COLLISION:
            if (mode != 2 && (j < 0 || j >= n))
                return fail(EDMD_ERR_ARG, "collision partner %d out of range", j);
            break;
        case EDMD_EV_CELLCROSS:
            if (mode != 2 && (j < 1 || j > 4))
                return fail(EDMD_ERR_ARG, "crossing direction %d is not 1..4", j);
            break;
        case EDMD_EV_WALL:
            if (mode != 2 && (j < 0 || j > 1)) return fail(EDMD_ERR_ARG, "wall side %d is not 0 or 1", j);
            break;
        case EDMD_EV_SCREENSHOT:
        case EDMD_EV_THERMO:
        case EDMD_EV_ADDINGNOISE:
            i = 0;
            break;
        default:
            return fail(EDMD_ERR_ARG, "event kind %d cannot be queued by hand", kind);
    }
    if (i < 0 || i >= n) return fail(EDMD_ERR_ARG, "particle %d out of range", i);
    if (mode != 2 && t != t) return fail(EDMD_ERR_ARG, "event time is NaN");
    HIPCHK(edmd_launch_queue_edit(c->b->d_sys, c->b->max_nlev, kind, i, j, t, cact, mode,
                                  c->b->stream));
    HIPCHK(hipStreamSynchronize(c->b->stream));
    return EDMD_OK;
}

extern "C" int edmd_queue_add(edmd_ctx* c, int kind, int i, int j, double t,
                              uint32_t coll_actual) {
    return queue_edit(c, kind, i, j, t, coll_actual, 0);
}

extern "C" int edmd_add_collision_event(edmd_ctx* c, int i, int j, double t) {
    return queue_edit(c, EDMD_EV_COLLISION, i, j, t, 0, 1);
}

extern "C" int edmd_queue_remove(edmd_ctx* c, int kind, int i) {
    return queue_edit(c, kind, i, 0, 0.0, 0, 2);
}

extern "C" int edmd_free_fly(edmd_ctx* c, int i) {
    int rc = particle_check(c, i);
    if (rc) return rc;
    HIPCHK(edmd_launch_fly_one(c->b->d_sys, i, c->b->stream));
    HIPCHK(hipStreamSynchronize(c->b->stream));
    return EDMD_OK;
}

extern "C" int edmd_get_box(edmd_ctx* c, int* n, double* Lx, double* Ly) {
    if (!c) return fail(EDMD_ERR_ARG, "null context");
    const EdmdConst& k = c->b->hs[0].c;
    if (n) *n = k.n;
    if (Lx) *Lx = k.Lx;
    if (Ly) *Ly = k.Ly;
    return EDMD_OK;
}

/* ------------------------------------------------------------ diagnostics */

extern "C" int edmd_fp64_check(int n, const double* a, const double* b, double* q, double* s) {
    if (n <= 0 || !a || !b || !q || !s) return fail(EDMD_ERR_ARG, "bad arguments");
    double *da = nullptr, *db = nullptr, *dq = nullptr, *ds = nullptr;
    size_t nb = (size_t)n * sizeof(double);
    HIPCHK(hipMalloc(&da, nb));
    HIPCHK(hipMalloc(&db, nb));
    HIPCHK(hipMalloc(&dq, nb));
    HIPCHK(hipMalloc(&ds, nb));
    HIPCHK(hipMemcpy(da, a, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db, b, nb, hipMemcpyHostToDevice));
    HIPCHK(edmd_launch_fp64_check(n, da, db, dq, ds, nullptr));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(q, dq, nb, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s, ds, nb, hipMemcpyDeviceToHost));
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dq);
    (void)hipFree(ds);
    return EDMD_OK;
}

/* ------------------------------------------------------------ initial states */

static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform in (0, 1] with 53 random bits */
static double unit_open0(uint64_t* s) { return ((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0); }

extern "C" int edmd_init_lattice(int n, double phi, double rad, int lattice, uint64_t seed,
                                 double temperature, double* x, double* y, double* vx, double* vy,
                                 double* Lx, double* Ly) {
    if (n < 2 || !(phi > 0) || !(phi < 1) || !(rad > 0) || !x || !y || !vx || !vy || !Lx || !Ly)
        return fail(EDMD_ERR_ARG, "bad lattice arguments");
    const double pi = 3.14159265358979323846;
    if (lattice == 0) {
        double L = sqrt(pi * n * rad * rad / phi);
        int side = (int)ceil(sqrt((double)n));
        double a = L / side;
        if (a < 2 * rad) return fail(EDMD_ERR_ARG, "square lattice overlaps at phi=%g", phi);
        for (int k = 0; k < n; k++) {
            x[k] = (k % side + 0.5) * a;
            y[k] = (k / side + 0.5) * a;
        }
        *Lx = *Ly = L;
    } else if (lattice == 1) {
        int ny = (int)sqrt((double)n);
        if (ny % 2 == 1) ny += 1;
        int nx = n / ny;
        if (nx * ny != n) return fail(EDMD_ERR_ARG, "triangular lattice needs n = nx*ny with ny even");
        double ly = sqrt(sqrt(3.0) / 2 * ny / nx * pi * n * rad * rad / phi);
        double lx = 2 / sqrt(3.0) * nx / ny * ly;
        double dx = lx / nx, dy = ly / ny;
        if (dx < 2 * rad) return fail(EDMD_ERR_ARG, "triangular lattice overlaps at phi=%g", phi);
        int k = 0;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++, k++) {
                x[k] = i * dx + (j % 2) * (dx / 2) + dx / 4;
                y[k] = j * dy + dy / 2;
            }
        *Lx = lx;
        *Ly = ly;
    } else if (lattice == 2) {
        /* triangular stacking fitted to a square box for any n: ny (even)
         * rows of nx = ceil(n/ny) sites, the row count chosen to maximise
         * the nearest-neighbour distance; sites filled row by row */
        double L = sqrt(pi * n * rad * rad / phi);
        int best_ny = 0;
        double best_d = 0;
        int lim = 2 * (int)ceil(sqrt((double)n)) + 4;
        for (int ny = 2; ny <= lim; ny += 2) {
            int nx = (n + ny - 1) / ny;
            double dx = L / nx, dy = L / ny;
            double d = std::min(dx, sqrt(dx * dx / 4 + dy * dy));
            d = std::min(d, 2 * dy);
            if (d > best_d) { best_d = d; best_ny = ny; }
        }
        if (best_d < 2 * rad) return fail(EDMD_ERR_ARG, "fitted lattice overlaps at phi=%g", phi);
        int ny = best_ny, nx = (n + ny - 1) / ny;
        double dx = L / nx, dy = L / ny;
        for (int k = 0; k < n; k++) {
            int i = k % nx, j = k / nx;
            x[k] = (i + 0.25 + 0.5 * (j % 2)) * dx;
            y[k] = (j + 0.5) * dy;
        }
        *Lx = *Ly = L;
    } else {
        return fail(EDMD_ERR_ARG, "unknown lattice %d", lattice);
    }
    uint64_t s = seed;
    double mx = 0, my = 0;
    for (int k = 0; k < n; k++) {
        double u1 = unit_open0(&s), u2 = unit_open0(&s);
        double r = sqrt(-2 * log(u1));
        vx[k] = r * cos(2 * pi * u2);
        vy[k] = r * sin(2 * pi * u2);
        mx += vx[k];
        my += vy[k];
    }
    mx /= n;
    my /= n;
    double e = 0;
    for (int k = 0; k < n; k++) {
        vx[k] -= mx;
        vy[k] -= my;
        e += 0.5 * (vx[k] * vx[k] + vy[k] * vy[k]);
    }
    double f = sqrt(temperature / (e / n));
    for (int k = 0; k < n; k++) {
        vx[k] *= f;
        vy[k] *= f;
    }
    return EDMD_OK;
}

extern "C" int edmd_init_random(int n, double phi, double rad, double poly, double sizeratio,
                                double fsmall, uint64_t seed, double e_per, double* x, double* y,
                                double* vx, double* vy, double* r, double* m, double* Lx,
                                double* Ly) {
    if (n < 2 || !(phi > 0) || !(phi < 1) || !(rad > 0) || !(poly >= 0) || !(poly < 1.0 / 3) ||
        !(fsmall >= 0) || !(fsmall < 1) || !(sizeratio > 0) || !(e_per > 0) || !x || !y || !vx ||
        !vy || !r || !m || !Lx || !Ly)
        return fail(EDMD_ERR_ARG, "bad random-start arguments");
    const double pi = 3.14159265358979323846;
    uint64_t s = seed;
    /* target radii and masses */
    const int nsmall = (poly > 0) ? 0 : (int)(n * fsmall); /* Nsmall, EDMD.c:1703 */
    const int nbig = n - nsmall;
    for (int k = 0; k < n; k++) {
        if (poly > 0) {
            /* diameter ratio d ~ N(1, poly) truncated to [1 - 3 poly, 1 + 3 poly] */
            double d;
            do {
                const double u1 = unit_open0(&s), u2 = unit_open0(&s);
                d = 1 + poly * sqrt(-2 * log(u1)) * cos(2 * pi * u2);
            } while (d < 1 - 3 * poly || d > 1 + 3 * poly);
            r[k] = rad * d;
            m[k] = d * d * d;
        } else {
            r[k] = k < nbig ? rad : rad * sizeratio; /* EDMD.c:1776-1789 */
            m[k] = 1;
        }
    }
    double area = 0;
    for (int k = 0; k < n; k++) area += pi * r[k] * r[k];
    const double L = sqrt(area / phi);
    /* uniform centres, then Box-Muller velocities of variance e_per / m per
     * component (EDMD.c:1719-1790 with EinitGrow = e_per) */
    for (int k = 0; k < n; k++) {
        x[k] = L * (unit_open0(&s) - 0.5 / 9007199254740992.0);
        y[k] = L * (unit_open0(&s) - 0.5 / 9007199254740992.0);
        if (x[k] >= L) x[k] = 0;
        if (y[k] >= L) y[k] = 0;
    }
    for (int k = 0; k < n; k++) {
        const double u1 = unit_open0(&s), u2 = unit_open0(&s);
        const double a = sqrt(-2 * log(u1) / m[k] * e_per);
        vx[k] = a * cos(2 * pi * u2);
        vy[k] = a * sin(2 * pi * u2);
    }
    *Lx = *Ly = L;
    return EDMD_OK;
}

extern "C" double edmd_growth_rate(double phi) {
    /* optimizeGrowConstant, EDMD.c:5766-5800, for a periodic box without
     * walls, coexistence or intruder: 0.1 below phi = 0.8, 0.1 (0.8/phi)^30
     * above, divided by 8 above phi = 0.7 */
    double vr = phi < 0.8 ? 0.1 : 0.1 * pow(0.8 / phi, 30);
    if (phi > 0.7) vr /= 8;
    return vr;
}

/* error hook for the other translation units of the library */
int edmd_internal_fail(int code, const char* msg) { return fail(code, "%s", msg); }

/* ------------------------------------------------------------ g(r) */

/* Histogram of one system or of host arrays, normalised as calculate_pcf
 * (pcf.c:16-75): g = 2 count / (2 pi r dr * (N / (Lx Ly)) * N). */
static int pcf_common(const double* hx, const double* hy, const EdmdPart* dpart, int n, double t,
                      int extrap, double Lx, double Ly, double dr, double max_r, double* r,
                      double* g, int* nbins, hipStream_t stream) {
    if (n < 1 || !(dr > 0) || !(max_r > 0) || !r || !g || !nbins)
        return fail(EDMD_ERR_ARG, "bad g(r) arguments");
    const int nb = (int)(max_r / dr);
    if (nb < 1 || nb > 32768) return fail(EDMD_ERR_ARG, "g(r) needs 1..32768 bins, got %d", nb);
    if (*nbins < nb) {
        *nbins = nb;
        return fail(EDMD_ERR_ARG, "g(r) output arrays need %d entries", nb);
    }
    double *dx = nullptr, *dy = nullptr;
    unsigned long long* dh = nullptr;
    HIPCHK(hipMalloc(&dh, (size_t)nb * sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(dh, 0, (size_t)nb * sizeof(unsigned long long), stream);
    if (e == hipSuccess && hx) {
        e = hipMalloc(&dx, (size_t)n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&dy, (size_t)n * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(dx, hx, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dy, hy, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = edmd_launch_pcf(dx, dy, hx ? nullptr : dpart, n, t, extrap, Lx, Ly, dr, nb, dh, stream);
    std::vector<unsigned long long> cnt(nb);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess)
        e = hipMemcpy(cnt.data(), dh, (size_t)nb * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    (void)hipFree(dh);
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return fail(EDMD_ERR_HIP, "g(r) failed: %s", hipGetErrorString(e));
    const double pi = 3.14159265358979323846;
    const double volume = Lx * Ly;
    const double density = n / volume;
    for (int b = 0; b < nb; b++) {
        r[b] = (b + 0.5) * dr;
        const double shell = 2 * pi * r[b] * dr;
        const double norm = shell * density * n;
        const double c = 2.0 * (double)cnt[b];
        g[b] = norm > 0 ? c / norm : 0.0;
    }
    *nbins = nb;
    return EDMD_OK;
}

extern "C" int edmd_pcf(int n, const double* x, const double* y, double Lx, double Ly, double dr,
                        double max_r, double* r, double* g, int* nbins) {
    if (!x || !y) return fail(EDMD_ERR_ARG, "null positions");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(EDMD_ERR_NODEVICE, "no HIP device visible");
    return pcf_common(x, y, nullptr, n, 0, 0, Lx, Ly, dr, max_r, r, g, nbins, nullptr);
}

extern "C" int edmd_batch_pcf(edmd_batch* b, int s, int extrapolate, double dr, double max_r,
                              double* r, double* g, int* nbins) {
    int rc = sys_check(b, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    const HostSys& h = b->hs[s];
    EdmdState st;
    HIPCHK(hipMemcpy(&st, b->st + s, sizeof st, hipMemcpyDeviceToHost));
    return pcf_common(nullptr, nullptr, b->part + h.p0, h.c.n, st.t, extrapolate, h.c.Lx, h.c.Ly,
                      dr, max_r, r, g, nbins, b->stream);
}

/* ------------------------------------------------------------ derivatives */

static_assert(sizeof(edmd_tape_rec) == sizeof(EdmdTapeRec), "tape record layout");

/* Byte size of one checkpoint: every pool that k_run reads or writes,
 * except the trace and thermo rows, which a recomputed segment rewrites
 * with the same values. */
struct PoolSpan {
    void* p;
    size_t bytes;
};

static std::vector<PoolSpan> state_pools(edmd_batch* b) {
    return {{b->part, b->np * sizeof(EdmdPart)},
            {b->ev, b->np * sizeof(EdmdEv)},
            {b->leaf, b->np * sizeof(EdmdNode)},
            {b->nodes, b->nnode * sizeof(EdmdNode)},
            {b->img, 2 * b->np * sizeof(int32_t)},
            {b->ccnt, b->ncell * sizeof(int32_t)},
            {b->cslot, b->nslot * sizeof(int32_t)},
            {b->st, (size_t)b->nsys * sizeof(EdmdState)},
            {b->erank, 2 * b->np * sizeof(int64_t)},
            {b->lrank, b->np * sizeof(int64_t)},
            {b->mt, b->mt ? (size_t)b->nsys * 625 * sizeof(uint32_t) : 0}};
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t ckpt_bytes(edmd_batch* b) {
    size_t t = 0;
    for (auto& q : state_pools(b)) t += align256(q.bytes);
    return t;
}

static int ckpt_copy(edmd_batch* b, void* mem, bool save) {
    char* c = (char*)mem;
    for (auto& q : state_pools(b)) {
        if (q.bytes) {
            if (save) HIPCHK(hipMemcpyAsync(c, q.p, q.bytes, hipMemcpyDeviceToDevice, b->stream));
            else HIPCHK(hipMemcpyAsync(q.p, c, q.bytes, hipMemcpyDeviceToDevice, b->stream));
        }
        c += align256(q.bytes);
    }
    return EDMD_OK;
}

static int upload_table(edmd_batch* b) {
    HIPCHK(hipMemcpy(b->d_sys, b->sys.data(), b->nsys * sizeof(EdmdSys), hipMemcpyHostToDevice));
    return EDMD_OK;
}

template <class T>
static int deriv_alloc(edmd_batch* b, T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess)
        return fail(EDMD_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    b->dallocs.push_back(q);
    *p = (T*)q;
    return EDMD_OK;
}

extern "C" int edmd_batch_deriv_end(edmd_batch* b) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    deriv_release(b);
    for (int s = 0; s < b->nsys; s++) {
        EdmdSys& S = b->sys[s];
        S.c.tape_cap = 0;
        S.a.tape = nullptr;
        S.a.dv = S.a.dtl = S.a.otl = nullptr;
        S.a.dr = nullptr;
    }
    int rc = upload_table(b);
    if (rc) return rc;
    HIPCHK(edmd_launch_seg_begin(b->d_sys, b->nsys, 0, 1, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return EDMD_OK;
}

extern "C" int edmd_batch_deriv_begin(edmd_batch* b, int64_t checkpoint_every, int64_t tape_capacity) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (checkpoint_every < 0) return fail(EDMD_ERR_ARG, "checkpoint_every must be >= 0");
    const int64_t cap = checkpoint_every > 0 ? checkpoint_every : tape_capacity;
    if (cap < 1 || cap > INT32_MAX)
        return fail(EDMD_ERR_ARG, "tape capacity %lld out of range [1, 2^31)", (long long)cap);
    for (int s = 0; s < b->nsys; s++)
        if (b->hs[s].c.noise)
            return fail(EDMD_ERR_ARG, "system %d: derivatives need a system without the "
                                      "velocity-rescale thermostat", s);
    if (b->any_grow)
        return fail(EDMD_ERR_ARG, "derivatives need systems whose growth has ended");
    for (int s = 0; s < b->nsys; s++)
        if ((b->hs[s].c.walls & 4) || (b->hs[s].c.damping && b->hs[s].c.field_on) ||
            b->hs[s].c.thermal_wall || b->hs[s].c.shear_wall)
            return fail(EDMD_ERR_ARG,
                        "system %d: derivatives need a system without the circular wall, "
                        "thermal or shear walls, or the field together with damping (flat "
                        "walls, the field and damping are taped)", s);
    int rc = edmd_batch_deriv_end(b);
    if (rc) return rc;
    if ((double)b->nsys * (double)cap * sizeof(EdmdTapeRec) > 1.0e12)
        return fail(EDMD_ERR_NOMEM, "tape of %lld records for %d systems is too large",
                    (long long)cap, b->nsys);
    if (!rc) rc = deriv_alloc(b, &b->tape, (size_t)b->nsys * (size_t)cap);
    if (!rc) rc = deriv_alloc(b, &b->dv, 4 * b->np);
    if (!rc) rc = deriv_alloc(b, &b->dtl, b->np);
    if (!rc) rc = deriv_alloc(b, &b->otl, b->np);
    if (!rc) rc = deriv_alloc(b, &b->dr, (size_t)b->nsys);
    if (rc) { deriv_release(b); return rc; }
    for (int s = 0; s < b->nsys; s++) {
        EdmdSys& S = b->sys[s];
        const size_t p0 = b->hs[s].p0;
        S.c.tape_cap = (int32_t)cap;
        S.a.tape = b->tape + (size_t)s * (size_t)cap;
        S.a.dv = b->dv + 4 * p0;
        S.a.dtl = b->dtl + p0;
        S.a.otl = b->otl + p0;
        S.a.dr = b->dr + s;
    }
    rc = upload_table(b);
    if (rc) { deriv_release(b); return rc; }
    b->seg = checkpoint_every;
    b->tape_cap = cap;
    b->deriv_on = true;
    HIPCHK(hipMemsetAsync(b->dv, 0, 4 * b->np * sizeof(double), b->stream));
    HIPCHK(hipMemsetAsync(b->dr, 0, b->nsys * sizeof(EdmdDeriv), b->stream));
    HIPCHK(edmd_launch_deriv_origin(b->d_sys, b->nsys, b->max_n, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return EDMD_OK;
}

static int tape_overflow_check(edmd_batch* b) {
    std::vector<EdmdState> st(b->nsys);
    HIPCHK(hipMemcpy(st.data(), b->st, b->nsys * sizeof(EdmdState), hipMemcpyDeviceToHost));
    for (int s = 0; s < b->nsys; s++)
        if (st[s].ntape > b->tape_cap)
            return fail(EDMD_ERR_TAPE_FULL, "system %d: %lld collisions exceed the tape capacity %lld",
                        s, (long long)st[s].ntape, (long long)b->tape_cap);
    return EDMD_OK;
}

/* Runs one checkpoint segment: at most `seg` events per system, new tape. */
static int run_segment(edmd_batch* b, const EdmdRunArgs& ra) {
    HIPCHK(edmd_launch_seg_begin(b->d_sys, b->nsys, b->seg, 1, b->stream));
    HIPCHK(launch_loop(b, 0, ra));
    HIPCHK(hipStreamSynchronize(b->stream));
    return check_status(b);
}

extern "C" int edmd_batch_deriv_run(edmd_batch* b, int64_t max_events, uint64_t max_collisions,
                                    double tmax) {
    if (!b) return fail(EDMD_ERR_ARG, "null batch");
    if (!b->deriv_on) return fail(EDMD_ERR_STATE, "taping not started (edmd_batch_deriv_begin)");
    if (b->seg == 0) {
        int rc = batch_run_mask(b, max_events, max_collisions, tmax, 0);
        if (rc) return rc;
        return tape_overflow_check(b);
    }
    EdmdRunArgs ra;
    memset(&ra, 0, sizeof ra);
    ra.max_events = max_events;
    ra.max_coll = max_collisions;
    ra.tmax = tmax;
    const size_t bytes = ckpt_bytes(b);
    std::vector<EdmdState> st(b->nsys);
    for (;;) {
        void* mem = nullptr;
        hipError_t e = hipMalloc(&mem, bytes);
        if (e != hipSuccess)
            return fail(EDMD_ERR_NOMEM, "checkpoint %zu (%zu bytes): %s", b->ckpts.size(), bytes,
                        hipGetErrorString(e));
        b->ckpts.push_back({mem, ra});
        int rc = ckpt_copy(b, mem, true);
        if (!rc) rc = run_segment(b, ra);
        if (rc) return rc;
        HIPCHK(hipMemcpy(st.data(), b->st, b->nsys * sizeof(EdmdState), hipMemcpyDeviceToHost));
        bool more = false;
        for (int s = 0; s < b->nsys; s++)
            if (st[s].stop_reason == EDMD_STOP_EVENTS && (int64_t)st[s].nevents < max_events)
                more = true;
        if (!more) break;
    }
    HIPCHK(edmd_launch_seg_begin(b->d_sys, b->nsys, 0, 0, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return EDMD_OK;
}

/* Recomputes segment k's tape from its checkpoint. */
static int replay_segment(edmd_batch* b, size_t k) {
    int rc = ckpt_copy(b, b->ckpts[k].mem, false);
    if (!rc) rc = run_segment(b, b->ckpts[k].ra);
    return rc;
}

static int save_end(edmd_batch* b) {
    if (!b->ckpt_end) {
        hipError_t e = hipMalloc(&b->ckpt_end, ckpt_bytes(b));
        if (e != hipSuccess) {
            b->ckpt_end = nullptr;
            return fail(EDMD_ERR_NOMEM, "end-state snapshot: %s", hipGetErrorString(e));
        }
    }
    return ckpt_copy(b, b->ckpt_end, true);
}

static int restore_end(edmd_batch* b) {
    int rc = ckpt_copy(b, b->ckpt_end, false);
    if (rc) return rc;
    HIPCHK(edmd_launch_seg_begin(b->d_sys, b->nsys, 0, 0, b->stream));
    return EDMD_OK;
}

extern "C" int edmd_batch_tangent(edmd_batch* b, const double* tan_in, double* tan_out, double* dw,
                                  double* dtlast) {
    if (!b || !tan_in || !tan_out) return fail(EDMD_ERR_ARG, "null argument");
    if (!b->deriv_on) return fail(EDMD_ERR_STATE, "taping not started (edmd_batch_deriv_begin)");
    HIPCHK(hipMemcpy(b->dv, tan_in, 4 * b->np * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(b->dr, 0, b->nsys * sizeof(EdmdDeriv), b->stream));
    HIPCHK(edmd_launch_deriv_init(b->d_sys, b->nsys, b->max_n, 0, b->stream));
    int rc = EDMD_OK;
    if (b->seg == 0) {
        HIPCHK(edmd_launch_tangent(b->d_sys, b->nsys, b->max_n, b->stream));
    } else {
        rc = save_end(b);
        for (size_t k = 0; !rc && k < b->ckpts.size(); k++) {
            rc = replay_segment(b, k);
            if (!rc) HIPCHK(edmd_launch_tangent(b->d_sys, b->nsys, b->max_n, b->stream));
        }
        int rc2 = restore_end(b);
        if (!rc) rc = rc2;
        if (rc) return rc;
    }
    HIPCHK(edmd_launch_deriv_flush(b->d_sys, b->nsys, b->max_n, 0, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(tan_out, b->dv, 4 * b->np * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<EdmdDeriv> d(b->nsys);
    HIPCHK(hipMemcpy(d.data(), b->dr, b->nsys * sizeof(EdmdDeriv), hipMemcpyDeviceToHost));
    for (int s = 0; s < b->nsys; s++) {
        if (dw) dw[s] = d[s].dsum;
        if (dtlast) dtlast[s] = d[s].dtlast;
    }
    return EDMD_OK;
}

extern "C" int edmd_batch_adjoint(edmd_batch* b, const double* lw, const double* lt,
                                  const double* lam_end, double* grad) {
    if (!b || !grad) return fail(EDMD_ERR_ARG, "null argument");
    if (!b->deriv_on) return fail(EDMD_ERR_STATE, "taping not started (edmd_batch_deriv_begin)");
    std::vector<EdmdDeriv> d(b->nsys);
    for (int s = 0; s < b->nsys; s++) {
        memset(&d[s], 0, sizeof d[s]);
        d[s].lw = lw ? lw[s] : 0.0;
        d[s].lt = lt ? lt[s] : 0.0;
        d[s].lt_pending = 1;
    }
    HIPCHK(hipMemcpy(b->dr, d.data(), b->nsys * sizeof(EdmdDeriv), hipMemcpyHostToDevice));
    if (lam_end) HIPCHK(hipMemcpy(b->dv, lam_end, 4 * b->np * sizeof(double), hipMemcpyHostToDevice));
    else HIPCHK(hipMemsetAsync(b->dv, 0, 4 * b->np * sizeof(double), b->stream));
    HIPCHK(edmd_launch_deriv_init(b->d_sys, b->nsys, b->max_n, 1, b->stream));
    int rc = EDMD_OK;
    if (b->seg == 0) {
        HIPCHK(edmd_launch_adjoint(b->d_sys, b->nsys, b->max_n, b->stream));
    } else {
        rc = save_end(b);
        for (size_t k = b->ckpts.size(); !rc && k-- > 0;) {
            rc = replay_segment(b, k);
            if (!rc) HIPCHK(edmd_launch_adjoint(b->d_sys, b->nsys, b->max_n, b->stream));
        }
        int rc2 = restore_end(b);
        if (!rc) rc = rc2;
        if (rc) return rc;
    }
    HIPCHK(edmd_launch_deriv_flush(b->d_sys, b->nsys, b->max_n, 1, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(grad, b->dv, 4 * b->np * sizeof(double), hipMemcpyDeviceToHost));
    return EDMD_OK;
}

extern "C" int edmd_batch_deriv_values(edmd_batch* b, int sys, double* w, double* tlast,
                                       int64_t* records) {
    int rc = sys_check(b, sys);
    if (rc) return rc;
    if (!b->deriv_on) return fail(EDMD_ERR_STATE, "taping not started (edmd_batch_deriv_begin)");
    EdmdState st;
    HIPCHK(hipMemcpy(&st, b->st + sys, sizeof st, hipMemcpyDeviceToHost));
    if (w) *w = st.wsum;
    if (tlast) *tlast = st.tlast;
    if (records) *records = st.ntape < b->tape_cap ? st.ntape : b->tape_cap;
    return EDMD_OK;
}

extern "C" int edmd_batch_get_tape(edmd_batch* b, int sys, edmd_tape_rec* out, int64_t cap,
                                   int64_t* n) {
    int64_t k = 0;
    int rc = edmd_batch_deriv_values(b, sys, nullptr, nullptr, &k);
    if (rc) return rc;
    if (n) *n = k;
    if (out && cap > 0) {
        const int64_t c = std::min(k, cap);
        if (c > 0)
            HIPCHK(hipMemcpy(out, b->tape + (size_t)sys * (size_t)b->tape_cap, c * sizeof(EdmdTapeRec),
                             hipMemcpyDeviceToHost));
    }
    return EDMD_OK;
}

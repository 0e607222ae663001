/*
 * `edmd` — command-line host mirroring the reference's main()
 * (EDMD.c:498-661, G == 0): same options (constantInit), same start-up
 * order, same event loop bound (while t <= tmax), a LAMMPS dump frame at
 * every screenshot event and a thermo row at every thermo event.  The
 * events themselves are handled on the GPU by libedmd_hip.so; the host only
 * wakes up after screenshot and thermo events to write the files.
 *
 * Exit status: 0 on success, otherwise the library status (3 for an
 * overlap, as the reference's exit(3), EDMD.c:2715-2716).
 */
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "edmd_compat.h"

static void progress(FILE* out) {
    fprintf(out, "t = %.6e / %.2e | n coll = %.3e | E/N = %.8f | p = %.4f\n", t, tmax,
            (double)ncol, N > 0 ? E / N : 0.0, pressure);
    fflush(out);
}

int main(int argc, char* argv[]) {
    constantInit(argc, argv);
    particlesInit();
    if (edmd_compat_status()) return edmd_compat_status();
    cellListInit();
    eventListInit();
    if (edmd_compat_status()) return edmd_compat_status();
    physicalQ();

    const char *dump_path = NULL, *thermo_path = NULL;
    int full = 0;
    edmd_compat_outputs(&dump_path, &thermo_path, &full);
    FILE* thermo = NULL;
    if (dump_path) {
        mkdir("dump/", 0777);
        thermo = fopen(thermo_path, "w");
        if (!thermo) {
            fprintf(stderr, "cannot open %s\n", thermo_path);
            return 5;
        }
        fprintf(thermo, "t Ncol E p px py pxy pyx\n");
    }
    printf("N = %d  phi = %.6f  L = %.6f x %.6f  cells = %d x %d\n", N, phi, Lx, Ly, Nxcells,
           Nycells);

    int rows_done = 0;
    const int mask = (1 << EDMD_EV_SCREENSHOT) | (1 << EDMD_EV_THERMO);
    while (t <= tmax) {
        int kind = edmd_compat_run(mask);
        if (kind == EDMD_EV_SCREENSHOT) {
            saveTXT();
            progress(stdout);
        } else if (kind == EDMD_EV_THERMO) {
            int nrow = 0;
            edmd_get_thermo(edmd_compat_context(), NULL, 0, &nrow);
            if (thermo && nrow > rows_done) {
                edmd_thermo_row* rows = malloc(sizeof(edmd_thermo_row) * (size_t)nrow);
                if (rows && edmd_get_thermo(edmd_compat_context(), rows, nrow, &nrow) == EDMD_OK) {
                    for (int k = rows_done; k < nrow; k++)
                        fprintf(thermo, "%lf %lu %lf %lf %lf %lf %.10lf %.10lf\n", rows[k].t,
                                (unsigned long)rows[k].ncol, rows[k].e_per_n, rows[k].pressure,
                                rows[k].pxx, rows[k].pyy, rows[k].pxy, rows[k].pyx);
                    fflush(thermo);
                }
                free(rows);
                rows_done = nrow;
            }
            progress(stdout);
        } else {
            break; /* tmax, collision cap, or an error */
        }
    }
    progress(stdout);
    if (thermo) fclose(thermo);
    int status = edmd_compat_status();
    freeArrays();
    return status;
}

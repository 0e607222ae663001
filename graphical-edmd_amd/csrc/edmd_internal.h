/* Internal launch interface between the host code and the kernels of
 * libedmd_hip.so.  Not part of the public C ABI (include/edmd.h). */
#ifndef EDMD_INTERNAL_H
#define EDMD_INTERNAL_H

#include <hip/hip_runtime.h>

#include "edmd_types.h"

/* Tile shape of sweep variant v (tx = 0 for an unknown variant). */
typedef struct EdmdSweepShape {
    int tx, ty, threads, stage;
} EdmdSweepShape;
#define EDMD_SWEEP_VARIANTS 8
EdmdSweepShape edmd_sweep_shape(int variant);
/* ext: EDMD_SWEEP_EXT when some system has walls, the field or damping
 * (the variant of the kernel that handles them), EDMD_SWEEP_POLY when some
 * system's radii differ (the layout that stages each disk's radius) */
enum { EDMD_SWEEP_EXT = 1, EDMD_SWEEP_POLY = 2 };
hipError_t edmd_launch_sweep_tiled(const EdmdSys* d_sys, int nsys, int max_tiles, int variant,
                                   int ext, hipStream_t s);
int edmd_sweep_tiles(int ncx, int ncy, int variant);
hipError_t edmd_launch_tree(const EdmdSys* d_sys, int nsys, const int* max_lsize, int max_nlev,
                            hipStream_t s);
hipError_t edmd_launch_sync_all(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s);
hipError_t edmd_launch_epoch_begin(const EdmdSys* d_sys, int nsys, int64_t ev_inc,
                                   uint64_t coll_inc, hipStream_t s);
hipError_t edmd_launch_epoch_end(const EdmdSys* d_sys, int nsys, EdmdEpochRow* rows,
                                 hipStream_t s);
hipError_t edmd_launch_run(const EdmdSys* d_sys, int nsys, int max_nlev, int grow, int walls,
                           int wpb, EdmdRunArgs ra,
                           hipStream_t s);
/* Event loop with each system's arrays resident in LDS (calendar depth <= 2):
 * dyn_bytes of dynamic LDS per workgroup, at least edmd_lds_layout's total
 * for every system of the launch. */
hipError_t edmd_launch_run_lds(const EdmdSys* d_sys, int nsys, int max_nlev, int grow, int walls,
                               size_t dyn_bytes, EdmdRunArgs ra, hipStream_t s);
/* Workgroups of the LDS-resident loop one CU holds with dyn_bytes of dynamic
 * LDS each, and the loop's static LDS. */
hipError_t edmd_lds_occupancy(int max_nlev, int grow, int walls, size_t dyn_bytes,
                              int* blocks_per_cu, size_t* static_bytes);
hipError_t edmd_launch_sweep_grow(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s);
/* counts the 2N insertions of a full sweep (after the sweep kernels) */
hipError_t edmd_launch_sweep_count(const EdmdSys* d_sys, int nsys, hipStream_t s);
hipError_t edmd_launch_peek(const EdmdSys* d_sys, int nsys, EdmdPeek* out, hipStream_t s);
hipError_t edmd_launch_pair_time(const EdmdSys* d_sys, int i, int j, double* out, hipStream_t s);
hipError_t edmd_launch_repredict(const EdmdSys* d_sys, int max_nlev, int i, int which,
                                 hipStream_t s);
hipError_t edmd_launch_fly_one(const EdmdSys* d_sys, int i, hipStream_t s);
/* addEventToQueue / removeEventFromQueue of one eventList slot (mode 0:
 * file with cact, 1: file with the partner's counter, 2: remove) */
hipError_t edmd_launch_queue_edit(const EdmdSys* d_sys, int max_nlev, int kind, int i, int j,
                                  double t, uint32_t cact, int mode, hipStream_t s);
hipError_t edmd_launch_pcf(const double* x, const double* y, const EdmdPart* part, int n, double t,
                           int extrap, double Lx, double Ly, double dr, int nb,
                           unsigned long long* hist, hipStream_t s);
hipError_t edmd_launch_fp64_check(int n, const double* a, const double* b, double* q, double* sq,
                                  hipStream_t s);

/* derivatives of the event map (edmd_deriv.hip) */
hipError_t edmd_launch_tangent(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s);
hipError_t edmd_launch_adjoint(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s);
hipError_t edmd_launch_deriv_init(const EdmdSys* d_sys, int nsys, int max_n, int mode,
                                  hipStream_t s);
hipError_t edmd_launch_deriv_flush(const EdmdSys* d_sys, int nsys, int max_n, int mode,
                                   hipStream_t s);
hipError_t edmd_launch_deriv_origin(const EdmdSys* d_sys, int nsys, int max_n, hipStream_t s);
hipError_t edmd_launch_seg_begin(const EdmdSys* d_sys, int nsys, int64_t seg, int clear,
                                 hipStream_t s);

#endif

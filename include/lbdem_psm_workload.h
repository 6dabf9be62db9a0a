/*
 * Seeded synthetic particle workloads for the lbdem_psm drivers (host only).
 *
 * The reference takes particles from LIGGGHTS insertion
 * (examples/sedimentingSphere/in2.lbdem, examples/benchmark/in2.lbdem); the standalone drivers and benchmarks here need
 * the same kind of input without a DEM code: non-overlapping spheres at
 * uniform random positions (rejection sampling on a cell list), radii from
 * a constant, uniform or clipped-lognormal law, per-component normal
 * velocities and angular velocities.  Deterministic for a given seed
 * (splitmix64-seeded xoshiro256**, Box-Muller normals).
 */
#ifndef LBDEM_PSM_WORKLOAD_H
#define LBDEM_PSM_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBDEM_PSM_RADIUS_CONST 0     /* r = a */
#define LBDEM_PSM_RADIUS_UNIFORM 1   /* r ~ U[a, b] */
#define LBDEM_PSM_RADIUS_LOGNORMAL 2 /* r = a * exp(b * N(0,1)), clipped to [lo, hi] */

typedef struct lbdem_psm_workload_desc {
    int32_t n[3];            /* lattice size */
    int32_t periodic[3];     /* 1 periodic axis, 0 walls (centres kept r + wall_gap inside) */
    int64_t count;           /* number of spheres */
    int32_t radius_law;
    int32_t reserved;
    double a, b, lo, hi;     /* radius law parameters */
    double gap;              /* minimum surface-to-surface gap */
    double wall_gap;         /* minimum distance of a surface to a wall */
    double sigma_v;          /* std of each velocity component */
    double sigma_w;          /* std of each angular-velocity component */
    uint64_t seed;
    int64_t max_attempts;    /* per sphere, 0 = 100000 */
} lbdem_psm_workload_desc;

/* Fills x3, v3, w3 (3*count), r (count) and id (count, 1-based tags).
 * Positions are in the lattice frame, inside [-1/2, n - 1/2) per axis.
 * Returns 0, or -1 on bad arguments / when a sphere cannot be placed. */
int lbdem_psm_generate_spheres(const lbdem_psm_workload_desc* d, double* x3, double* v3, double* w3,
                               double* r, int64_t* id);

#ifdef __cplusplus
}
#endif
#endif

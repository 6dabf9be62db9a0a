/*
 * lbdem_psm -- C ABI of the MI355X (gfx950, HIP) partially-saturated-cells
 * (PSM) hot path of LBDEMcoupling: D3Q19 Noble-Torczynski collision fused
 * with streaming, sub-cell sphere voxelization into per-cell coupling
 * records, and per-particle force/torque summation.
 *
 * Conventions
 *  - Every function returns an int: LBDEM_PSM_OK (0) or a negative error
 *    code; lbdem_psm_last_error() gives the message.  No exception crosses
 *    the ABI and nothing aborts.
 *  - Caller-owned host arrays are read or written during the call only and
 *    never retained (the borrow-for-the-call lifetime of the reference's
 *    SetSingleSphere3D / SumForceTorque3D pointer arguments).
 *  - Device memory and one HIP stream are owned by the handle.  One host
 *    thread per handle at a time.
 *  - Lattice frame: cell (i,j,k) is centred at (i,j,k); cell index
 *    c = (i*ny + j)*nz + k (z fastest).  Host population arrays are
 *    pop-major, f[q*Nc + c], in the deviation form f_q - t_q, 19 D3Q19
 *    populations in the Palabos order (c_{q+9} = -c_q).  Per-cell vector
 *    fields are component-major, v[d*Nc + c].  Particle arrays are
 *    particle-major, x[3*p + d].
 */
#ifndef LBDEM_PSM_H
#define LBDEM_PSM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBDEM_PSM_OK 0
#define LBDEM_PSM_ERR_INVALID_ARG -1
#define LBDEM_PSM_ERR_HIP -2
#define LBDEM_PSM_ERR_OUT_OF_MEMORY -3
#define LBDEM_PSM_ERR_CAPACITY -4
#define LBDEM_PSM_ERR_NO_DEVICE -5
#define LBDEM_PSM_ERR_STATE -6
#define LBDEM_PSM_ERR_NUMERIC -7

/* Population storage precision. */
#define LBDEM_PSM_FP64 0        /* f - t in FP64, FP64 arithmetic */
#define LBDEM_PSM_FP32_DEV 1    /* f - t in FP32, FP32 arithmetic */
#define LBDEM_PSM_FP16_STORE 2  /* f - t stored in FP16, FP32 arithmetic */

/* Voxelization rule (lbdem_psm_desc.voxel_mode) */
#define LBDEM_PSM_VOXEL_SUBCELL 0  /* calcSolidFraction sub-cell count (the reference) */
#define LBDEM_PSM_VOXEL_OVERLAP 1  /* analytic sphere-cell overlap, differentiable in x */

/* lbdem_psm_set_spheres flags */
#define LBDEM_PSM_INIT_VEL 1    /* initVelFlag of setSpheresOnLattice */
#define LBDEM_PSM_AUTO_IMAGES 2 /* library appends periodic images (no ghosts given) */

/* lbdem_psm_step flags */
#define LBDEM_PSM_STEP_ADVANCE 1 /* prescribed kinematics: x += v (periodic wrap) first */
#define LBDEM_PSM_STEP_FROZEN 2  /* no voxelization: the records of the last one stay (particles at rest
                                  * in the lattice frame); excludes STEP_ADVANCE */

typedef struct lbdem_psm lbdem_psm;

typedef struct lbdem_psm_desc {
    int32_t nx, ny, nz;          /* lattice size (cells) */
    int32_t periodic[3];         /* 1 periodic, 0 half-way bounce-back walls */
    double tau;                  /* BGK relaxation time, omega = 1/tau, > 0.5 */
    int32_t precision;           /* LBDEM_PSM_FP64 / _FP32_DEV / _FP16_STORE */
    int32_t slices_per_dim;      /* sub-cell samples per axis (reference: 5) */
    double body_accel[3];        /* Guo body-force acceleration, lattice units */
    int32_t device;              /* HIP device ordinal */
    int32_t voxel_mode;          /* LBDEM_PSM_VOXEL_SUBCELL / _OVERLAP */
    int64_t max_solid_cells;     /* optional limit on live coupling records, 0 = none */
} lbdem_psm_desc;

/* Defaults: 1x1x1, periodic, tau 1, FP64, 5 slices, no force, device 0,
 * sub-cell count voxelization. */
int lbdem_psm_desc_init(lbdem_psm_desc* desc);

/* Creates the lattice (populations at rest, rho = 1) -- the Palabos
 * MultiBlockLattice3D with IBcompositeDynamics(new BGKdynamics(omega)) on
 * every cell, sedimentingSphere.cpp:39,181; src/ibCompositeDynamics3D.h:37-38. */
int lbdem_psm_create(const lbdem_psm_desc* desc, lbdem_psm** out);
int lbdem_psm_destroy(lbdem_psm* h);
/* Message of the last error on h, or of the last failed create when h is NULL. */
const char* lbdem_psm_last_error(const lbdem_psm* h);
const char* lbdem_psm_version(void);

/* Equilibrium populations from rho (Nc) and u (3*Nc); NULL means rho = 1 /
 * u = 0.  Replaces lattice.initialize() and the initial-field functionals. */
int lbdem_psm_init_equilibrium(lbdem_psm* h, const double* rho, const double* u);
/* Raw population access (checkpoint/restore, parity): 19*Nc doubles. */
int lbdem_psm_set_populations(lbdem_psm* h, const double* f);
int lbdem_psm_get_populations(lbdem_psm* h, double* f);

/* setSpheresOnLattice + AttributeFunctional
 * (src/ibDataExchangeWrappers3D.hh:33-146): one SetSingleSphere3D per
 * particle in array order (src/ibProcessors3D.h:37-45,
 * src/ibProcessors3D.hh:52-184), run as one ordered per-cell fold on the GPU.
 * x3, v3, w3: positions / velocities / angular velocities in lattice units;
 * r: radii; id: LIGGGHTS tags (>= 1); type may be NULL; exclude lists the
 * excluded types (excludeType).  flags: LBDEM_PSM_INIT_VEL,
 * LBDEM_PSM_AUTO_IMAGES.  The particles stay resident on the device. */
int lbdem_psm_set_spheres(lbdem_psm* h, int64_t n, const double* x3, const double* v3,
                          const double* w3, const double* r, const int64_t* id,
                          const int32_t* type, const int32_t* exclude, int32_t n_exclude,
                          int32_t flags);

/* lattice.collideAndStream() with IBcompositeDynamics::collide
 * (src/ibCompositeDynamics3D.hh:130-203) on every cell, nsteps times. */
int lbdem_psm_collide_stream(lbdem_psm* h, int32_t nsteps);

/* SumForceTorque3D (src/ibProcessors3D.hh:202-275) + the map(tag) fold of
 * getForcesFromLattice (src/ibDataExchangeWrappers3D.hh:160-243): force and
 * torque on each of the n particles of the last set_spheres, in lattice
 * units, torque arms about x3[map(id)] with the minimum image.  x3 may be
 * NULL (use the resident positions).  F3, T3: 3*n outputs. */
int lbdem_psm_sum_force_torque(lbdem_psm* h, int64_t n, const double* x3, double* F3,
                               double* T3);

/* One full coupled step on resident data, nsteps times: [x += v] ->
 * voxelize -> collide+stream -> force/torque sum (kept on the device; read
 * with lbdem_psm_get_particle_forces).  The loop of
 * sedimentingSphere.cpp:212-225 without the DEM substeps.  With
 * LBDEM_PSM_STEP_FROZEN the voxelization is skipped. */
int lbdem_psm_step(lbdem_psm* h, int32_t nsteps, int32_t flags);
int lbdem_psm_get_particle_forces(lbdem_psm* h, int64_t n, double* F3, double* T3);
int lbdem_psm_get_particles(lbdem_psm* h, int64_t n, double* x3);

/* Dense per-cell coupling record (IBdynamicsParticleData,
 * src/ibDynamicsParticleData.h:26-37): sf (Nc), id (Nc), us (3*Nc),
 * F (3*Nc).  Any pointer may be NULL. */
int lbdem_psm_get_cell_records(lbdem_psm* h, double* sf, int64_t* id, double* us,
                               double* F);
/* Restores records (checkpoint, or a prepared ownership history). */
int lbdem_psm_set_cell_records(lbdem_psm* h, const double* sf, const int64_t* id,
                               const double* us);

/* calcSolidFraction (src/ibProcessors3D.hh:116-158) evaluated on the GPU for
 * n (offset, radius) queries: d3 = cell centre - particle centre.  Outputs
 * sf, the sub-cell count and the branch (0 outside, 1 inside, 2 counted). */
int lbdem_psm_subcell_counts(lbdem_psm* h, int64_t n, const double* d3, const double* r,
                             double* sf, int32_t* count, int32_t* branch);

/* Allocation-free form of lbdem_psm_subcell_counts.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results; it calls no hipMalloc / hipFree
 * and stages the queries in a fixed device array of
 * LBDEM_PSM_STATIC_MAX_QUERIES x 48 bytes (3 MiB) that each GPU holds until
 * the process ends.  n above that limit returns LBDEM_PSM_ERR_CAPACITY.
 * All calls on one GPU share the array: calls that overlap on one GPU race. */
#define LBDEM_PSM_STATIC_MAX_QUERIES 65536
int lbdem_psm_subcell_counts_static(lbdem_psm* h, int64_t n, const double* d3, const double* r,
                                    double* sf, int32_t* count, int32_t* branch);

/* Density and velocity per cell, computed on the device (Palabos
 * computeDensity / computeVelocity of the BGK cell; with Guo forcing
 * u = j/rho + a/2): rho (Nc), u3 (3*Nc, component-major); either may be
 * NULL. */
int lbdem_psm_get_macroscopic(lbdem_psm* h, double* rho, double* u3);
/* The descriptor the handle was created with (tau as last set). */
int lbdem_psm_get_desc(const lbdem_psm* h, lbdem_psm_desc* out);

/* ---- On-disk formats ----
 * VTK XML ImageData (.vti), Float32 point data, x fastest, binary base64
 * with a UInt32 byte-count header -- the format of Palabos' VtkImageOutput3D
 * used by the reference's writeVTK (examples/benchmark/benchmark.cpp:74-99).
 * data[k]: ncomp[k]*Nc doubles in the host layout (component-major, z
 * fastest); scale may be NULL.  Host only, no device needed. */
int lbdem_psm_vti_write(const char* path, int32_t nx, int32_t ny, int32_t nz, double spacing, int32_t nfields,
                        const char* const* names, const int32_t* ncomp, const double* const* data,
                        const double* scale);
/* The reference's writeVTK fields of a lattice: velocity * vel_scale,
 * density * rho_scale, pressure (rho - 1) * p_scale, solid fraction. */
int lbdem_psm_write_vti(lbdem_psm* h, const char* path, double spacing, double vel_scale, double rho_scale,
                        double p_scale);
/* Checkpoint file: header, then per cell (z fastest) the 19 populations
 * f - t in FP64 followed by the coupling record in the serialize order of
 * IBcompositeDynamics (src/ibCompositeDynamics3D.hh:77-89): uPart[3],
 * hydrodynamicForce[3], solidFraction, partId (int64).  The _arrays forms
 * are host only (host layouts as get_populations / get_cell_records). */
int lbdem_psm_checkpoint_write_arrays(const char* path, const lbdem_psm_desc* desc, const double* f,
                                      const double* sf, const int64_t* id, const double* us, const double* F);
int lbdem_psm_checkpoint_read_desc(const char* path, lbdem_psm_desc* desc);
int lbdem_psm_checkpoint_read_arrays(const char* path, double* f, double* sf, int64_t* id, double* us,
                                     double* F);
int lbdem_psm_save_checkpoint(lbdem_psm* h, const char* path);
/* Restores populations and records (the records' forces are recomputed by
 * the next collision; all records are summed on the stale-record path
 * until the next set_spheres). */
int lbdem_psm_load_checkpoint(lbdem_psm* h, const char* path);

/* FP64 on-device global sums: mass deviation sum(f - t) and momentum. */
int lbdem_psm_conservation(lbdem_psm* h, double* mass, double* mom3);
/* Health check: non-finite populations and max |u|; returns
 * LBDEM_PSM_ERR_NUMERIC when a value is non-finite or max |u| >= 0.3. */
int lbdem_psm_check_health(lbdem_psm* h, double* max_u, int64_t* n_nonfinite);

/* Counters of the last voxelization: [0] coupling records, [1] candidate
 * entries, [2] voxelized particle images, [3] orphan records. */
int lbdem_psm_stats(lbdem_psm* h, int64_t* out4);
int lbdem_psm_synchronize(lbdem_psm* h);

/* Per-kernel timing with HIP events on the handle's stream.  Kernel
 * classes: 0 collide_stream, 1 voxelize (bin + fold), 2 force_sum,
 * 3 advance.  read: total ms and launch count per class; reset clears. */
int lbdem_psm_profile(lbdem_psm* h, int32_t enable);
int lbdem_psm_profile_read(lbdem_psm* h, double* ms4, int64_t* count4, int32_t reset);

/* Measurement utility (no counterpart in the reference): the device copy
 * ceiling of the handle's population store.  Copies the current population
 * buffer into the second one (which holds no state between steps) reps
 * times with a plain 16-byte-per-lane copy kernel (streaming loads and
 * stores, six wavefronts per SIMD) on the handle's stream, after one
 * untimed copy; ms_per_copy is the HIP-event time per copy and
 * bytes_per_copy the bytes read plus written by one copy.  Populations are
 * unchanged. */
int lbdem_psm_copy_bandwidth(lbdem_psm* h, int32_t reps, double* ms_per_copy, double* bytes_per_copy);

/* Zhang-Kwok periodic pressure difference across the periodic axis
 * `dimension` (ZhangPeriodicPressureFunctional3D,
 * src/periodicPressureFunctionals3D.hh:118-147, applied as in
 * examples/showcaseRectChannel/showcaseRectChannel.cpp:258-273) around every
 * subsequent collide+stream: the mean densities of the inlet plane (index 0)
 * and the outlet plane (index n-1) over the cross-section box are taken
 * before the step; after it the populations entering the inlet (c_d = +1)
 * are rescaled (in the f + t form) by rho_in / <rho>_outlet and those
 * entering the outlet (c_d = -1) by rho_out / <rho>_inlet.
 * range4 = {a0, a1, b0, b1}: inclusive box on the other two axes in
 * increasing axis order; NULL = 1..n-2 on both (the showcase's inlet and
 * outlet).  enable = 0 switches it off. */
int lbdem_psm_set_periodic_pressure(lbdem_psm* h, int32_t enable, int32_t dimension, double rho_in,
                                    double rho_out, const int32_t* range4);

/* Equilibrium populations of PoiseuilleProfileAndPressureGradient
 * (src/periodicPressureFunctionals3D.hh:259-298): density
 * rho_hi - (rho_hi - rho_lo) i / (n_d - 1) along `dimension` and velocity
 * u_max p(j, n_j) p(k, n_k) along it, p(i, n) = 1 - ((i - n/2)/(n/2))^2.
 * n3: the profile sizes (NULL = the lattice's; the showcase passes
 * nx - 1 along the axis). */
int lbdem_psm_init_poiseuille(lbdem_psm* h, int32_t dimension, double rho_hi, double rho_lo, double u_max,
                              const int32_t* n3);

/* ---- Differentiable window (no counterpart in the reference) ----
 *
 * A window of nsteps collide+stream+force-sum steps over frozen geometry:
 * the coupling records of the last set_spheres stay fixed while the
 * populations evolve (prescribed particle velocities act as moving-wall
 * data).  Objective J = sum_t sum_p g_ft[t][p] . (F_p(t), T_p(t))
 * + <g_fT, f(T)>.  The backward pass is the hand-written adjoint of the
 * fused PSM collide-stream and of the force/torque sum; populations are
 * checkpointed every checkpoint_every steps and recomputed per segment.
 * Gradients: particle velocities v, angular velocities w, positions x
 * (through the analytic overlap when voxel_mode = OVERLAP, plus the
 * rigid-body lever arms and the torque arms in either mode), tau, and the
 * per-cell solid fraction sf and weight B.  FP64 and FP32_DEV storage. */

/* Changes tau (omega = 1/tau) of an existing lattice. */
int lbdem_psm_set_tau(lbdem_psm* h, double tau);
/* Removes every coupling record (fresh ownership history). */
int lbdem_psm_clear_records(lbdem_psm* h);
/* sf += dsf[c] on every cell with a live record (dsf: Nc, dense).  The
 * result must stay in (0.001, 1]; dsf must be 0 on cells without a record. */
int lbdem_psm_perturb_solid_fraction(lbdem_psm* h, const double* dsf);
/* Runs the window from the current populations; ft (may be NULL):
 * nsteps*n*6 doubles, (F, T) per particle per step, folded onto map(id)
 * like sum_force_torque.  Leaves the lattice at step nsteps.
 * checkpoint_every > 0: a population checkpoint every that many steps, the
 * backward pass holding ceil(nsteps/K) + K - 1 states.  checkpoint_every =
 * -M < 0: a budget of M states in all (segments of M, M-1, ... steps, each
 * finished segment's checkpoint reused for recompute; needs M(M+1)/2 >=
 * nsteps), for windows too large for the first form; such a window allows
 * one diff_backward.  Every step is recomputed once in both forms, and the
 * gradients are bit-identical. */
int lbdem_psm_diff_forward(lbdem_psm* h, int32_t nsteps, int32_t checkpoint_every, double* ft);
/* Adjoint of the last diff_forward.  g_ft: nsteps*n*6 (may be NULL = 0);
 * g_fT: 19*Nc host populations layout (may be NULL = 0).  Outputs (each may
 * be NULL): g_x3, g_v3, g_w3 (3*n, on map(id), ghost slots 0), g_tau,
 * g_sf and g_B (Nc, dense, 0 where no partial record), g_f0 (19*Nc). */
int lbdem_psm_diff_backward(lbdem_psm* h, const double* g_ft, const double* g_fT, double* g_x3,
                            double* g_v3, double* g_w3, double* g_tau, double* g_sf, double* g_B,
                            double* g_f0);
/* Frees the checkpoints and adjoint buffers. */
int lbdem_psm_diff_release(lbdem_psm* h);

/* Host-side argument check of set_spheres, no device needed. */
int lbdem_psm_check_spheres(const lbdem_psm_desc* desc, int64_t n, const double* x3,
                            const double* v3, const double* w3, const double* r,
                            const int64_t* id);
/* Host-side check of a descriptor, no device needed. */
int lbdem_psm_check_desc(const lbdem_psm_desc* desc);

#ifdef __cplusplus
}
#endif
#endif

/*
 * CPU oracle of the LBDEMcoupling partially-saturated-cells (PSM) step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  Nothing in the product links or
 * calls it.
 *
 * It restates, in plain FP64 C++ (compiled with -ffp-contract=off), the
 * reference algorithm over a dense per-cell record, the way the reference
 * holds it (one IBdynamicsParticleData per cell):
 *   - PSM collision          src/ibCompositeDynamics3D.hh:109-203
 *   - equilibrium start      src/ibCompositeDynamics3D.hh:118-127 (defineVelocity)
 *   - voxelization           src/ibProcessors3D.hh:52-184, src/ibProcessors3D.h:43-45
 *   - force/torque sum       src/ibProcessors3D.hh:202-275
 *   - exchange loop          src/ibDataExchangeWrappers3D.hh:33-246
 *   - SOLFRAC limits         src/ibDef.h:24-27
 * The D3Q19 BGK / equilibrium / streaming arithmetic (Palabos, not part of
 * the reference tree) is restated from the textbook D3Q19 model with the
 * population convention f_i - t_i used by the reference
 * (src/periodicPressureFunctionals3D.hh:151-153).
 *
 * Population layout at this API: pop-major, f[i*Nc + c] with
 * c = (x*ny + y)*nz + z (z fastest, the iteration order of
 * src/ibProcessors3D.hh:57-59).  Internally the oracle keeps AoS cells.
 *
 * Parity pins (tests/test_oracle.py): closed forms and invariants
 * (equilibrium moments, mass/momentum conservation, B=0 reduces to BGK,
 * solid-velocity equilibrium gives zero force, Poiseuille profile,
 * brute-force and symmetry checks of sub-cell counts, hand-built
 * ownership cases, brute-force force/torque sums).
 */
#ifndef LBDEM_PSM_ORACLE_H
#define LBDEM_PSM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_lattice orc_lattice;

/* periodic[d] = 1: periodic axis; 0: half-way bounce-back walls.
 * body_accel: Guo forcing acceleration (lattice units), may be NULL. */
orc_lattice* orc_create(int nx, int ny, int nz, const int* periodic, double tau,
                        int slices_per_dim, const double* body_accel);
void orc_destroy(orc_lattice* L);

/* rho (Nc) and u (3*Nc, pop-major [d*Nc + c]) may be NULL (rho=1, u=0). */
void orc_init_equilibrium(orc_lattice* L, const double* rho, const double* u);
void orc_set_populations(orc_lattice* L, const double* f);
void orc_get_populations(const orc_lattice* L, double* f);

/* setSpheresOnLattice + AttributeFunctional.  Positions in the lattice
 * frame (cell (i,j,k) centred at (i,j,k)).  auto_images: append periodic
 * images in canonical order.  Returns 0 or a negative error. */
int orc_set_spheres(orc_lattice* L, int64_t n, const double* x3, const double* v3,
                    const double* w3, const double* r, const int64_t* id,
                    const int32_t* type, const int32_t* exclude, int n_exclude,
                    int init_vel, int auto_images);

/* Threads of orc_set_spheres and orc_sum_force_torque (default 1): each
 * thread owns an x-slab, as each MPI rank of the reference owns a block.
 * set_spheres gives the same records for any count; sum_force_torque adds
 * the slabs' partial sums in slab order (a different grouping than the
 * sequential scan). */
void orc_set_threads(orc_lattice* L, int nthreads);

/* nthreads <= 0: all OpenMP threads. */
void orc_collide_stream(orc_lattice* L, int nsteps, int nthreads);

/* SumForceTorque3D over the whole lattice, sequential in (x,y,z) order.
 * x3: positions of the n particles given to the last orc_set_spheres
 * (index = map(id)).  F3, T3: n*3 outputs (zeroed first). */
int orc_sum_force_torque(const orc_lattice* L, int64_t n, const double* x3, double* F3,
                         double* T3);

/* Dense per-cell coupling record: sf (Nc), id (Nc), us (3*Nc), F (3*Nc),
 * pop-major like the populations.  Any pointer may be NULL. */
void orc_get_cell_records(const orc_lattice* L, double* sf, int64_t* id, double* us,
                          double* F);
void orc_set_cell_records(orc_lattice* L, const double* sf, const int64_t* id,
                          const double* us);

/* calcSolidFraction for one (cell, particle) pair.  n_out: sub-cell count
 * (0 / N^3 for the two shortcuts), branch_out: 0 outside shortcut,
 * 1 inside shortcut, 2 counted. */
double orc_calc_solid_fraction(double dx, double dy, double dz, double r, int N,
                               int32_t* n_out, int32_t* branch_out);

/* Analytic overlap (the differentiable voxelization): the fraction of a
 * unit-volume cell covered by a sphere of radius r whose centre is at
 * distance d from the cell centre, with the cell represented by the sphere
 * of equal volume (radius (3/(4 pi))^(1/3)) at its centre -- the lens volume
 * of two spheres.  dsf_dd (may be NULL): derivative with respect to d.
 * branch_out (may be NULL): 0 outside, 1 cell inside the sphere, 3 lens,
 * 4 sphere inside the cell sphere. */
double orc_overlap_fraction(double d, double r, double* dsf_dd, int32_t* branch_out);

/* Voxelization rule of orc_set_spheres: 0 = sub-cell count of the reference
 * (default), 1 = analytic overlap (orc_overlap_fraction of the distance
 * sqrt((dx^2 + dy^2) + dz^2)).  Ownership, thresholds and setValues stay
 * the reference's. */
void orc_set_voxel_mode(orc_lattice* L, int mode);

/* Zhang-Kwok periodic pressure across the periodic axis dim around every
 * later orc_collide_stream step (src/periodicPressureFunctionals3D.hh:118-147,
 * examples/showcaseRectChannel/showcaseRectChannel.cpp:258-273).  range4:
 * inclusive box {a0,a1,b0,b1} on the other two axes (NULL: 1..n-2). */
void orc_set_periodic_pressure(orc_lattice* L, int enable, int dim, double rho_in, double rho_out,
                               const int32_t* range4);

/* PoiseuilleProfileAndPressureGradient at equilibrium
 * (src/periodicPressureFunctionals3D.hh:259-298); n3: profile sizes
 * (NULL: the lattice's). */
void orc_init_poiseuille(orc_lattice* L, int dim, double rho_hi, double rho_lo, double u_max,
                         const int32_t* n3);

/* FP64 global mass deviation sum_i sum_c fbar and momentum sum c_i fbar. */
void orc_conservation(const orc_lattice* L, double* mass, double* mom3);

/* Equilibrium of one cell (f - t form), for tests. */
void orc_equilibrium(double rho_bar, const double* j3, double* eq19);

/* Collide one cell in place (no streaming): f19 pre-collision populations,
 * sf / us the record, F3 the produced force.  For unit tests. */
void orc_collide_cell(double* f19, double tau, double sf, const double* us3, double* F3);

/* Number of image entries produced by the last orc_set_spheres (originals +
 * images that were voxelized), for tests. */
int64_t orc_last_n_voxelized(const orc_lattice* L);

#ifdef __cplusplus
}
#endif
#endif

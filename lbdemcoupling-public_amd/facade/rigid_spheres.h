/*
 * Minimal rigid-sphere integrator standing in for LIGGGHTS' fix nve/sphere
 * + fix gravity + fix couple/lb/onetoone for drivers without a DEM code.
 *
 * Reference behaviour kept (examples/sedimentingSphere/in.lbdem:36-42,
 * fix_lb_coupling_onetoone.cpp:130-163, sedimentingSphere.cpp:168,225):
 * the hydrodynamic force and torque from the last lattice step are held
 * constant over the DEM substeps and added to every owned particle
 * (torque only with use_torque), gravity is the buoyancy-reduced
 * g (rho_s - rho_f) / rho_s, and each substep is a velocity-Verlet step
 * of a free sphere (moment of inertia 2/5 m r^2).  Contacts and walls are
 * not modelled.
 */
#ifndef LBDEM_PSM_RIGID_SPHERES_H
#define LBDEM_PSM_RIGID_SPHERES_H

#include "plb_ib.h"

namespace plb {

struct RigidSphereIntegrator {
    double dt = 1.0;             /* DEM substep */
    double gravity[3] = {0.0, 0.0, 0.0};   /* buoyancy-reduced acceleration */

    /* n substeps on the owned atoms of `w` (rmass must be set) */
    void run(LiggghtsCouplingWrapper& w, int n) {
        AtomArrays* a = w.lmp->atom;
        FixLbCouplingOnetoone* fix = w.lmp->fix;
        double** fh = fix->get_force_ptr();
        double** th = fix->get_torque_ptr();
        for (int s = 0; s < n; ++s) {
            for (int i = 0; i < a->nlocal; ++i) {
                const double m = a->rmass[i];
                const double I = 0.4 * m * a->radius[i] * a->radius[i];
                for (int d = 0; d < 3; ++d) {
                    const double F = fh[i][d] + m * gravity[d];
                    const double Tq = fix->use_torque ? th[i][d] : 0.0;
                    a->v[i][d] += 0.5 * dt * F / m;          /* half kick */
                    a->x[i][d] += dt * a->v[i][d];           /* drift */
                    a->v[i][d] += 0.5 * dt * F / m;          /* half kick, force unchanged */
                    a->omega[i][d] += dt * Tq / I;
                    a->f[i][d] = F;
                    a->torque[i][d] = Tq;
                }
            }
        }
    }
};

}  // namespace plb

#endif

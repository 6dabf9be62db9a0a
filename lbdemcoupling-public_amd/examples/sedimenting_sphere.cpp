/*
 * Single sedimenting sphere through the plb_ib.h facade, shaped like
 * examples/sedimentingSphere/sedimentingSphere.cpp:160-240 of the reference:
 * 40 x 40 x 120 lattice, walls in x and y, periodic z, tau 0.8, one sphere
 * of radius 6 starting at rest at (20, 20, 100) (lattice frame), density
 * ratio 1.5, gravity 1e-6 on the particle only (buoyancy-reduced), 8^3
 * sub-cell sampling, 10 DEM substeps per lattice step.
 * Units: l_p = 1 cell, u_p = 1, N = 1, uMax = 1, rho_p = 1, so physical
 * positions are lattice positions + 1/2 and forces are lattice forces.
 *
 * usage: sedimenting_sphere [steps=2000] [log_every=100]
 */
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "../facade/plb_ib.h"
#include "../facade/rigid_spheres.h"

using namespace plb;
typedef double T;
#define DESCRIPTOR descriptors::D3Q19Descriptor

int main(int argc, char** argv) {
    const plint maxSteps = argc > 1 ? std::atol(argv[1]) : 2000;
    const plint logSteps = argc > 2 ? std::atol(argv[2]) : 100;
    const plint nx = 40, ny = 40, nz = 120;
    const T tau = 0.8, omega = 1.0 / tau, r = 6.0, rho_s = 1.5, g = 1e-6;
    const plint demSubsteps = 10;
    try {
        PhysUnits3D<T> units(1.0, 1.0, 0.1, (T)nx, (T)ny, (T)nz, 1, 1.0, 1.0);
        MultiBlockLattice3D<T, DESCRIPTOR> lattice(nx, ny, nz,
            new IBcompositeDynamics<T, DESCRIPTOR>(new BGKdynamics<T, DESCRIPTOR>(omega)));
        defineDynamics(lattice, lattice.getBoundingBox(),
                       new IBcompositeDynamics<T, DESCRIPTOR>(new BGKdynamics<T, DESCRIPTOR>(omega)));
        lattice.periodicity().toggleAll(false);
        lattice.periodicity().toggle(2, true);
        lattice.setSlicesPerDim(8);
        lattice.initialize();

        LiggghtsCouplingWrapper wrapper;
        wrapper.resize(1, 0);
        AtomArrays* atom = wrapper.lmp->atom;
        atom->x[0][0] = units.getPhysPosition(20.0);
        atom->x[0][1] = units.getPhysPosition(20.0);
        atom->x[0][2] = units.getPhysPosition(100.0);
        atom->radius[0] = units.getPhysLength(r);
        atom->rmass[0] = rho_s * 4.0 / 3.0 * M_PI * r * r * r;
        atom->tag[0] = 1;
        wrapper.rebuildMap();

        RigidSphereIntegrator dem;
        dem.dt = 1.0 / (T)demSubsteps;
        dem.gravity[2] = -g * (rho_s - 1.0) / rho_s;

        auto t0 = std::chrono::steady_clock::now();
        for (plint iT = 0; iT < maxSteps; ++iT) {
            setSpheresOnLattice(lattice, wrapper, units, false);
            lattice.collideAndStream();
            getForcesFromLattice(lattice, wrapper, units);
            dem.run(wrapper, (int)demSubsteps);
            /* periodic z: keep the particle inside [0, nz) */
            if (atom->x[0][2] < 0.0) atom->x[0][2] += nz;
            if (atom->x[0][2] >= nz) atom->x[0][2] -= nz;
            if ((iT + 1) % logSteps == 0) {
                const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                std::printf("step %ld z %.6f vz %.6e fz %.6e  %.1f MLU/s\n", (long)(iT + 1),
                            units.getLbPosition(atom->x[0][2]), atom->v[0][2],
                            wrapper.lmp->fix->get_force_ptr()[0][2],
                            (double)(nx * ny * nz) * (iT + 1) / s / 1e6);
            }
        }
    } catch (const PsmError& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}

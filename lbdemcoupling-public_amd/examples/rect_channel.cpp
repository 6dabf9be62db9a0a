/*
 * Pressure-driven rectangular channel with fixed spheres through the
 * plb_ib.h facade, shaped like examples/showcaseRectChannel/showcaseRectChannel.cpp:196-273
 * of the reference: periodic along x with the Zhang-Kwok pressure
 * difference, walls in y and z, a PoiseuilleProfileAndPressureGradient
 * initial field, spheres set on the lattice every step, VTK output and a
 * checkpoint at the end.
 *
 * usage: rect_channel [steps=300] [out_prefix=rect]
 * prints: "step N flux F fx FX" lines and "checkpoint ok" after a
 * save / load / continue check.
 */
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../facade/plb_ib.h"

using namespace plb;
typedef double T;
#define DESCRIPTOR descriptors::D3Q19Descriptor

int main(int argc, char** argv) {
    const plint maxSteps = argc > 1 ? std::atol(argv[1]) : 300;
    const std::string prefix = argc > 2 ? argv[2] : "rect";
    const plint nxL = 48, ny = 24, nz = 24;          /* lattice; the profile uses nx = nxL - 1 */
    const T tau = 0.6, omega = 1.0 / tau, uMax = 0.01, deltaRho = 2e-4;
    try {
        PhysUnits3D<T> units(1.0, 1.0, 0.1, (T)nxL, (T)ny, (T)nz, 1, 1.0, 1.0);
        MultiBlockLattice3D<T, DESCRIPTOR> lattice(nxL, ny, nz,
            new IBcompositeDynamics<T, DESCRIPTOR>(new BGKdynamics<T, DESCRIPTOR>(omega)));
        defineDynamics(lattice, lattice.getBoundingBox(),
                       new IBcompositeDynamics<T, DESCRIPTOR>(new BGKdynamics<T, DESCRIPTOR>(omega)));
        lattice.periodicity().toggleAll(false);
        lattice.periodicity().toggle(0, true);
        const plint nx = nxL - 1;
        Box3D inlet(0, 0, 1, ny - 2, 1, nz - 2);
        const T rhoHi = 1., rhoLo = 1. - deltaRho;
        initializeAtEquilibrium(lattice, lattice.getBoundingBox(),
                                PoiseuilleProfileAndPressureGradient<T>(rhoHi, rhoLo, uMax, nx, ny, nz, 0));
        lattice.initialize();
        setPeriodicPressure(lattice, 0, rhoHi, rhoLo, inlet);

        LiggghtsCouplingWrapper wrapper;
        wrapper.resize(3, 0);
        AtomArrays* atom = wrapper.lmp->atom;
        const T pos[3][3] = {{12.0, 11.5, 12.0}, {26.0, 8.0, 15.0}, {38.0, 15.0, 9.0}};
        for (int i = 0; i < 3; ++i) {
            for (int d = 0; d < 3; ++d) atom->x[i][d] = units.getPhysPosition(pos[i][d]);
            atom->radius[i] = units.getPhysLength(3.5);
            atom->rmass[i] = 1.0;
            atom->tag[i] = i + 1;
        }
        wrapper.rebuildMap();

        lbdem_psm* h = lattice.handle();
        const int64_t nc = (int64_t)nxL * ny * nz;
        std::vector<double> rho(nc), u(3 * nc);
        for (plint iT = 1; iT <= maxSteps; ++iT) {
            setSpheresOnLattice(lattice, wrapper, units, iT == 1);
            lattice.collideAndStream();
            getForcesFromLattice(lattice, wrapper, units);
            if (iT % 100 == 0 || iT == maxSteps) {
                psmCheck(lbdem_psm_get_macroscopic(h, rho.data(), u.data()), h);
                double flux = 0.0;   /* sum of u_x over the mid plane */
                const plint xm = nxL / 2;
                for (plint y = 0; y < ny; ++y)
                    for (plint z = 0; z < nz; ++z) flux += u[(xm * ny + y) * nz + z];
                double fx = 0.0;
                for (int i = 0; i < 3; ++i) fx += wrapper.lmp->fix->get_force_ptr()[i][0];
                std::printf("step %ld flux %.6e fx %.6e\n", (long)iT, flux, fx);
            }
        }
        writeVTK(lattice, units, prefix + ".vti");
        saveBinaryBlock(lattice, prefix + ".ckpt");
        /* a second lattice continues from the checkpoint exactly as the first */
        MultiBlockLattice3D<T, DESCRIPTOR> copy(nxL, ny, nz,
            new IBcompositeDynamics<T, DESCRIPTOR>(new BGKdynamics<T, DESCRIPTOR>(omega)));
        copy.periodicity().toggleAll(false);
        copy.periodicity().toggle(0, true);
        copy.initialize();
        loadBinaryBlock(copy, prefix + ".ckpt");
        setPeriodicPressure(copy, 0, rhoHi, rhoLo, inlet);
        for (MultiBlockLattice3D<T, DESCRIPTOR>* L : {&lattice, &copy}) {
            setSpheresOnLattice(*L, wrapper, units, false);
            L->collideAndStream();
        }
        std::vector<double> fa(19 * nc), fb(19 * nc);
        psmCheck(lbdem_psm_get_populations(lattice.handle(), fa.data()), lattice.handle());
        psmCheck(lbdem_psm_get_populations(copy.handle(), fb.data()), copy.handle());
        if (fa != fb) {
            std::fprintf(stderr, "checkpoint continuation differs\n");
            return 2;
        }
        std::printf("checkpoint ok\n");
    } catch (const PsmError& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}

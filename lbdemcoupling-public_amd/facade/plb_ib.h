/*
 * plb_ib.h -- C++ facade of the LBDEMcoupling PSM surface over the lbdem_psm
 * C ABI (include/lbdem_psm.h).
 *
 * A driver written against the reference (src/plb_ib.h and the examples)
 * keeps its shape: it builds a MultiBlockLattice3D with
 * IBcompositeDynamics(new BGKdynamics(omega)), calls setSpheresOnLattice,
 * collideAndStream and getForcesFromLattice each step, and reads the forces
 * from the coupling fix of a LiggghtsCouplingWrapper.  Behind these names
 * the lattice lives on one MI355X and every hot loop runs in the HIP
 * library; the host only copies particle arrays in and forces out.
 *
 *   reference                                        here
 *   IBcompositeDynamics (ibCompositeDynamics3D.h:32-66)  marker over BGKdynamics: omega
 *   SetSingleSphere3D (ibProcessors3D.h:34-58)           queued into the lattice's batch
 *   AttributeFunctional (ibProcessors3D.h:65-74)         flushes the batch: one GPU fold
 *   SumForceTorque3D (ibProcessors3D.h:76-94)            one GPU reduction
 *   setSpheresOnLattice / getForcesFromLattice           same loops, same unit conversion
 *     (ibDataExchangeWrappers3D.hh:33-246)
 *   PhysUnits3D (physunits.h:29-92)                      same conversions
 *   LiggghtsCouplingWrapper (liggghtsCouplingWrapper.h)  particle store shaped like the
 *                                                        LIGGGHTS atom arrays + coupling fix
 *
 * Errors from the library are thrown as plb::PsmError (the reference aborts
 * through PLB_PRECONDITION / error->all instead).
 */
#ifndef LBDEM_PSM_PLB_IB_H
#define LBDEM_PSM_PLB_IB_H

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lbdem_psm.h"

namespace plb {

typedef long plint;
typedef unsigned long pluint;

class PsmError : public std::runtime_error {
public:
    PsmError(int code, const std::string& msg)
        : std::runtime_error("lbdem_psm error " + std::to_string(code) + ": " + msg), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

inline void psmCheck(int rc, const lbdem_psm* h) {
    if (rc != LBDEM_PSM_OK) throw PsmError(rc, lbdem_psm_last_error(h));
}

template <typename T, int N>
struct Array {
    T v[N];
    Array() { for (int k = 0; k < N; ++k) v[k] = T(); }
    Array(T a, T b, T c) { static_assert(N == 3, "3-vector"); v[0] = a; v[1] = b; v[2] = c; }
    T& operator[](plint k) { return v[k]; }
    const T& operator[](plint k) const { return v[k]; }
    void from_cArray(const T* a) { for (int k = 0; k < N; ++k) v[k] = a[k]; }
    void resetToZero() { for (int k = 0; k < N; ++k) v[k] = T(); }
};

struct Dot3D { plint x, y, z; };

struct Box3D {
    plint x0, x1, y0, y1, z0, z1;
    Box3D() : x0(0), x1(0), y0(0), y1(0), z0(0), z1(0) {}
    Box3D(plint a0, plint a1, plint b0, plint b1, plint c0, plint c1)
        : x0(a0), x1(a1), y0(b0), y1(b1), z0(c0), z1(c1) {}
};

inline bool intersect(const Box3D& a, const Box3D& b, Box3D& out) {
    out = Box3D(std::max(a.x0, b.x0), std::min(a.x1, b.x1), std::max(a.y0, b.y0),
                std::min(a.y1, b.y1), std::max(a.z0, b.z0), std::min(a.z1, b.z1));
    return out.x0 <= out.x1 && out.y0 <= out.y1 && out.z0 <= out.z1;
}

namespace descriptors {
template <typename T> struct D3Q19Descriptor { enum { d = 3, q = 19 }; };
}

/* ---- dynamics markers (the collision itself runs on the GPU) ---- */
template <typename T, template <typename U> class Descriptor>
struct Dynamics {
    virtual ~Dynamics() {}
    virtual T getOmega() const = 0;
    virtual Dynamics* clone() const = 0;
};

template <typename T, template <typename U> class Descriptor>
struct BGKdynamics : public Dynamics<T, Descriptor> {
    explicit BGKdynamics(T omega_) : omega(omega_) {}
    T getOmega() const override { return omega; }
    BGKdynamics* clone() const override { return new BGKdynamics(*this); }
    T omega;
};

/* IBcompositeDynamics(Dynamics* base, bool automaticPrepareCollision = true),
 * src/ibCompositeDynamics3D.h:37-38: owns its base dynamics. */
template <typename T, template <typename U> class Descriptor>
struct IBcompositeDynamics : public Dynamics<T, Descriptor> {
    explicit IBcompositeDynamics(Dynamics<T, Descriptor>* base_, bool automaticPrepareCollision = true)
        : base(base_) { (void)automaticPrepareCollision; }
    IBcompositeDynamics(const IBcompositeDynamics& o) : Dynamics<T, Descriptor>(), base(o.base->clone()) {}
    T getOmega() const override { return base->getOmega(); }
    IBcompositeDynamics* clone() const override { return new IBcompositeDynamics(*this); }
    Dynamics<T, Descriptor>& getBaseDynamics() { return *base; }
    std::unique_ptr<Dynamics<T, Descriptor>> base;
};

/* ---- flow parameters and units ---- */
/* Palabos' IncomprFlowParam: lattice velocity, Reynolds number, resolution N
 * (cells per reference length) and the domain extent in reference lengths. */
template <typename T>
class IncomprFlowParam {
public:
    IncomprFlowParam(T latticeU_, T Re_, plint resolution_, T lx_, T ly_, T lz_)
        : latticeU(latticeU_), Re(Re_), resolution(resolution_), lx(lx_), ly(ly_), lz(lz_) {}
    T getLatticeU() const { return latticeU; }
    T getRe() const { return Re; }
    plint getResolution() const { return resolution; }
    T getDeltaX() const { return (T)1 / (T)resolution; }
    T getDeltaT() const { return getDeltaX() * latticeU; }
    T getLatticeNu() const { return latticeU * (T)resolution / Re; }
    T getTau() const { return (T)3 * getLatticeNu() + (T)0.5; }
    T getOmega() const { return (T)1 / getTau(); }
    plint getNx() const { return (plint)std::lround(lx * resolution) + 1; }
    plint getNy() const { return (plint)std::lround(ly * resolution) + 1; }
    plint getNz() const { return (plint)std::lround(lz * resolution) + 1; }
private:
    T latticeU, Re;
    plint resolution;
    T lx, ly, lz;
};

/* Physical <-> lattice conversions used by the exchange (physunits.h:29-92):
 * lattice position = x / l_p / dx - 1/2 (cell i covers [i, i+1) dx), forces
 * scale by dx^4/dt^2 rho l_p^4/t_p^2 and torques by dx^5/dt^2 rho l_p^5/t_p^2. */
template <typename T>
class PhysUnits3D {
public:
    PhysUnits3D(T l_p_, T u_p_, T nu_p_, T lx_, T ly_, T lz_, plint N, T uMax, T rho_p_)
        : l_p(l_p_), u_p(u_p_), t_p(l_p_ / u_p_), nu_p(nu_p_), rho_p(rho_p_),
          param(uMax, l_p_ * u_p_ / nu_p_, N, lx_ / l_p_, ly_ / l_p_, lz_ / l_p_) {
        const T dx = param.getDeltaX(), dt = param.getDeltaT();
        forceFactor = std::pow(dx, 4) / (dt * dt) * rho_p * std::pow(l_p, 4) / (t_p * t_p);
        torqueFactor = std::pow(dx, 5) / (dt * dt) * rho_p * std::pow(l_p, 5) / (t_p * t_p);
    }
    T getRe() const { return param.getRe(); }
    /* offsets of getLbPositionX/Y/Z only; the exchange's getLbPosition keeps 1/2 */
    void setLbOffset(T ox, T oy, T oz) { off[0] = ox; off[1] = oy; off[2] = oz; }
    T getLbLength(T x) const { return x / l_p / param.getDeltaX(); }
    T getLbPosition(T x) const { return x / l_p / param.getDeltaX() - 0.5; }
    T getLbPositionX(T x) const { return x / l_p / param.getDeltaX() - off[0]; }
    T getLbPositionY(T x) const { return x / l_p / param.getDeltaX() - off[1]; }
    T getLbPositionZ(T x) const { return x / l_p / param.getDeltaX() - off[2]; }
    T getLbDensity(T rho) const { return rho / rho_p; }
    /* lattice density deviation rho - 1 of a physical pressure (p = (rho - 1)
     * c_s^2 in the force / area unit); as in the reference it is not the
     * inverse of getPhysPress, which takes rho itself */
    T getLbRho(T p, T pOffset = 0) const { return (p - pOffset) * 3. / getPhysForce(1.) * std::pow(getPhysLength(1.), 2); }
    T getLbVel(T u) const { return u / u_p * param.getLatticeU(); }
    T getLbFreq(T f) const { return f * t_p * param.getDeltaT(); }
    T getLbAccel(T a) const { return a * (t_p * t_p / l_p) * (param.getDeltaT() * param.getDeltaT() / param.getDeltaX()); }
    T getLbTime(T t) const { return t / t_p / param.getDeltaT(); }
    plint getLbSteps(T t) const { return (plint)(t / param.getDeltaT() / t_p + 0.5); }
    T getPhysLength(T l) const { return l * param.getDeltaX() * l_p; }
    T getPhysPosition(T l) const { return (l + 0.5) * param.getDeltaX() * l_p; }
    T getPhysPositionX(T l) const { return (l + off[0]) * param.getDeltaX() * l_p; }
    T getPhysPositionY(T l) const { return (l + off[1]) * param.getDeltaX() * l_p; }
    T getPhysPositionZ(T l) const { return (l + off[2]) * param.getDeltaX() * l_p; }
    T getPhysDensity(T rho) const { return rho * rho_p; }
    T getPhysAccel(T a) const { return a / getLbAccel(1.); }
    T getPhysPress(T rho, T pOffset = 0) const { return (rho - 1.) / 3. * getPhysForce(1.) / std::pow(getPhysLength(1.), 2) + pOffset; }
    T getPhysVel(T u) const { return u / param.getLatticeU() * u_p; }
    T getPhysTime(T t) const { return t * param.getDeltaT() * t_p; }
    T getPhysForce(T f) const { return f * forceFactor; }
    T getPhysTorque(T t) const { return t * torqueFactor; }
    const IncomprFlowParam<T>& getLbParam() const { return param; }
private:
    T l_p, u_p, t_p, nu_p, rho_p;
    IncomprFlowParam<T> param;
    T forceFactor = 0, torqueFactor = 0;
    T off[3] = {0.5, 0.5, 0.5};
};

/* ---- the lattice ---- */
struct Periodicity {
    bool p[3] = {true, true, true};
    void toggleAll(bool v) { p[0] = p[1] = p[2] = v; }
    void toggle(plint d, bool v) { p[d] = v; }
    bool get(plint d) const { return p[d]; }
};

template <typename T, template <typename U> class Descriptor>
struct SetSingleSphere3D;

/* MultiBlockLattice3D(nx, ny, nz, backgroundDynamics): one lbdem_psm handle,
 * created on initialize() (after periodicity and dynamics are defined). */
template <typename T, template <typename U> class Descriptor>
class MultiBlockLattice3D {
public:
    MultiBlockLattice3D(plint nx_, plint ny_, plint nz_, Dynamics<T, Descriptor>* background)
        : nx(nx_), ny(ny_), nz(nz_) {
        omega = background->getOmega();
        delete background;
        lbdem_psm_desc_init(&desc);
    }
    ~MultiBlockLattice3D() { if (h) lbdem_psm_destroy(h); }
    MultiBlockLattice3D(const MultiBlockLattice3D&) = delete;
    MultiBlockLattice3D& operator=(const MultiBlockLattice3D&) = delete;

    plint getNx() const { return nx; }
    plint getNy() const { return ny; }
    plint getNz() const { return nz; }
    Box3D getBoundingBox() const { return Box3D(0, nx - 1, 0, ny - 1, 0, nz - 1); }
    Periodicity& periodicity() { return per; }

    /* build options beyond the reference surface */
    void setPrecision(int p) { desc.precision = p; }
    void setSlicesPerDim(int n) { desc.slices_per_dim = n; }
    void setBodyAccel(T ax, T ay, T az) { desc.body_accel[0] = ax; desc.body_accel[1] = ay; desc.body_accel[2] = az; }
    void setDevice(int d) { desc.device = d; }
    void setOmega(T om) { omega = om; }
    T getOmega() const { return omega; }

    /* Creates the device lattice (populations at rest, rho = 1) on first
     * use; later calls keep the populations (an initial field set by
     * initializeAtEquilibrium survives, as in Palabos). */
    void initialize() {
        if (!h) {
            desc.nx = (int32_t)nx; desc.ny = (int32_t)ny; desc.nz = (int32_t)nz;
            for (int d = 0; d < 3; ++d) desc.periodic[d] = per.get(d) ? 1 : 0;
            desc.tau = 1.0 / (double)omega;
            psmCheck(lbdem_psm_create(&desc, &h), nullptr);
        }
    }
    void collideAndStream() { psmCheck(lbdem_psm_collide_stream(handle(), 1), h); }

    lbdem_psm* handle() {
        if (!h) initialize();
        return h;
    }

    /* coupling batch: SetSingleSphere3D functionals queue here in call order */
    void queueSphere(const T* x, const T* v, const T* w, T r, int id, bool initVel) {
        for (int d = 0; d < 3; ++d) { qx.push_back(x[d]); qv.push_back(v[d]); qw.push_back(w[d]); }
        qr.push_back(r);
        qid.push_back(id);
        qInitVel = qInitVel || initVel;
    }
    void flushSpheres() {
        const int64_t n = (int64_t)qr.size();
        const int32_t flags = qInitVel ? LBDEM_PSM_INIT_VEL : 0;
        psmCheck(lbdem_psm_set_spheres(handle(), n, qx.data(), qv.data(), qw.data(), qr.data(),
                                       qid.data(), nullptr, nullptr, 0, flags), h);
        nSet = n;
        setIds = qid;
        qx.clear(); qv.clear(); qw.clear(); qr.clear(); qid.clear();
        qInitVel = false;
    }
    int64_t spheresOnLattice() const { return nSet; }
    const std::vector<int64_t>& sphereIds() const { return setIds; }

private:
    plint nx, ny, nz;
    T omega;
    Periodicity per;
    lbdem_psm_desc desc;
    lbdem_psm* h = nullptr;
    std::vector<double> qx, qv, qw, qr;
    std::vector<int64_t> qid, setIds;
    bool qInitVel = false;
    int64_t nSet = 0;
};

/* PoiseuilleProfileAndPressureGradient (src/periodicPressureFunctionals3D.h,
 * .hh:259-298) and initializeAtEquilibrium over the whole lattice: the field
 * is written on the device (lbdem_psm_init_poiseuille). */
template <typename T>
struct PoiseuilleProfileAndPressureGradient {
    PoiseuilleProfileAndPressureGradient(T pHi_, T pLo_, T uMax_, plint nx_, plint ny_, plint nz_, plint dim_)
        : pHi(pHi_), pLo(pLo_), uMax(uMax_), dimension(dim_) { n[0] = (int32_t)nx_; n[1] = (int32_t)ny_; n[2] = (int32_t)nz_; }
    T pHi, pLo, uMax;
    int32_t n[3];
    plint dimension;
};

template <typename T, template <typename U> class Descriptor>
void initializeAtEquilibrium(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D box,
                             const PoiseuilleProfileAndPressureGradient<T>& f) {
    (void)box;
    psmCheck(lbdem_psm_init_poiseuille(lattice.handle(), (int32_t)f.dimension, f.pHi, f.pLo, f.uMax, f.n),
             lattice.handle());
}

/* The reference's rectangular-channel driver computes the inlet/outlet mean
 * densities before collideAndStream and applies two
 * ZhangPeriodicPressureFunctional3D after it
 * (examples/showcaseRectChannel/showcaseRectChannel.cpp:258-273).  Here the
 * library does both around every collideAndStream once configured:
 * the inlet box is the cross-section of the plane at index 0 along
 * `dimension` (e.g. Box3D(0,0,1,ny-2,1,nz-2) for dimension 0). */
template <typename T, template <typename U> class Descriptor>
void setPeriodicPressure(MultiBlockLattice3D<T, Descriptor>& lattice, plint dimension, T rhoIn, T rhoOut,
                         Box3D inlet) {
    int32_t r[4];
    const plint lo[3] = {inlet.x0, inlet.y0, inlet.z0}, hi[3] = {inlet.x1, inlet.y1, inlet.z1};
    int k = 0;
    for (int d = 0; d < 3; ++d) {
        if (d == dimension) continue;
        r[k++] = (int32_t)lo[d];
        r[k++] = (int32_t)hi[d];
    }
    psmCheck(lbdem_psm_set_periodic_pressure(lattice.handle(), 1, (int32_t)dimension, rhoIn, rhoOut, r),
             lattice.handle());
}

/* Palabos' saveBinaryBlock / loadBinaryBlock checkpoint of a lattice: the
 * populations and the per-cell coupling record in the serialize order of
 * IBcompositeDynamics (src/ibCompositeDynamics3D.hh:77-107). */
template <typename T, template <typename U> class Descriptor>
void saveBinaryBlock(MultiBlockLattice3D<T, Descriptor>& lattice, const std::string& fname) {
    psmCheck(lbdem_psm_save_checkpoint(lattice.handle(), fname.c_str()), lattice.handle());
}

template <typename T, template <typename U> class Descriptor>
void loadBinaryBlock(MultiBlockLattice3D<T, Descriptor>& lattice, const std::string& fname) {
    psmCheck(lbdem_psm_load_checkpoint(lattice.handle(), fname.c_str()), lattice.handle());
}

/* writeVTK of the reference drivers (examples/benchmark/benchmark.cpp:74-99):
 * velocity, density, pressure (rho - 1) * pFact and solid fraction as VTK
 * ImageData in physical units. */
template <typename T, template <typename U> class Descriptor>
void writeVTK(MultiBlockLattice3D<T, Descriptor>& lattice, PhysUnits3D<T> const& units, const std::string& fname) {
    const T pFact = units.getPhysForce(1) / (units.getPhysLength(1) * units.getPhysLength(1)) / 3.;
    psmCheck(lbdem_psm_write_vti(lattice.handle(), fname.c_str(), units.getPhysLength(1), units.getPhysVel(1), 1.0,
                                 pFact),
             lattice.handle());
}

/* defineDynamics(lattice, box, new IBcompositeDynamics(new BGKdynamics(omega))):
 * every cell carries the same dynamics; only omega is kept. */
template <typename T, template <typename U> class Descriptor>
void defineDynamics(MultiBlockLattice3D<T, Descriptor>& lattice, Box3D box, Dynamics<T, Descriptor>* dyn) {
    (void)box;
    lattice.setOmega(dyn->getOmega());
    delete dyn;
}

/* ---- processing functionals ---- */
template <typename T, template <typename U> class Descriptor>
struct SetSingleSphere3D {
    SetSingleSphere3D(T* x_, T* v_, T* omega_, T* com_, T r_, int id_, bool initVelFlag_ = false)
        : r(r_), id(id_), initVelFlag(initVelFlag_) {
        for (int d = 0; d < 3; ++d) { x[d] = x_[d]; v[d] = v_[d]; omega[d] = omega_[d]; com[d] = com_[d]; }
    }
    /* Box3D(x-r-2, x+r+2) with truncation toward zero (ibProcessors3D.h:43-45) */
    Box3D getBoundingBox() const {
        return Box3D((plint)(x[0] - r - 2), (plint)(x[0] + r + 2), (plint)(x[1] - r - 2),
                     (plint)(x[1] + r + 2), (plint)(x[2] - r - 2), (plint)(x[2] + r + 2));
    }
    T x[3], v[3], omega[3], com[3], r;
    int id;
    bool initVelFlag;
};

template <typename T, template <typename U> class Descriptor>
struct AttributeFunctional {};

template <typename T>
struct ParticleData {
    typedef std::vector<Array<T, 3>> ParticleDataArrayVector;
    typedef std::vector<T> ParticleDataScalarVector;
};

class LiggghtsCouplingWrapper;

template <typename T, template <typename U> class Descriptor>
struct SumForceTorque3D {
    SumForceTorque3D(typename ParticleData<T>::ParticleDataArrayVector& x_, T* force_, T* torque_,
                     LiggghtsCouplingWrapper& wrapper_)
        : x(x_), force(force_), torque(torque_), wrapper(wrapper_) {}
    typename ParticleData<T>::ParticleDataArrayVector& x;
    T *force, *torque;
    LiggghtsCouplingWrapper& wrapper;
};

/* applyProcessingFunctional takes ownership of the functional, like Palabos. */
template <typename T, template <typename U> class Descriptor>
void applyProcessingFunctional(SetSingleSphere3D<T, Descriptor>* f, Box3D domain,
                               MultiBlockLattice3D<T, Descriptor>& lattice) {
    (void)domain;   /* the GPU fold intersects the same box with the lattice */
    lattice.queueSphere(f->x, f->v, f->omega, f->r, f->id, f->initVelFlag);
    delete f;
}

template <typename T, template <typename U> class Descriptor>
void applyProcessingFunctional(AttributeFunctional<T, Descriptor>* f, Box3D domain,
                               MultiBlockLattice3D<T, Descriptor>& lattice) {
    (void)domain;
    lattice.flushSpheres();
    delete f;
}

/* SumForceTorque3D over the lattice: force[3*map(id)+c], torque[...] for
 * every particle of the last sphere batch (ibProcessors3D.hh:202-275). */
template <typename T, template <typename U> class Descriptor>
void applyProcessingFunctional(SumForceTorque3D<T, Descriptor>* f, Box3D domain,
                               MultiBlockLattice3D<T, Descriptor>& lattice) {
    (void)domain;
    const int64_t n = lattice.spheresOnLattice();
    std::vector<double> xs(3 * n), F(3 * n), Tq(3 * n);
    for (int64_t i = 0; i < n && i < (int64_t)f->x.size(); ++i)
        for (int d = 0; d < 3; ++d) xs[3 * i + d] = f->x[i][d];
    psmCheck(lbdem_psm_sum_force_torque(lattice.handle(), n, xs.data(), F.data(), Tq.data()),
             lattice.handle());
    for (int64_t i = 0; i < 3 * n; ++i) { f->force[i] = F[i]; f->torque[i] = Tq[i]; }
    delete f;
}

/* ---- LIGGGHTS-shaped particle store ---- */
/* The arrays setSpheresOnLattice / getForcesFromLattice read
 * (ibDataExchangeWrappers3D.hh:48,77-84,171-190,237-241): x, v, omega,
 * radius, tag, type, nlocal, nghost and map(tag). */
struct AtomArrays {
    int nlocal = 0, nghost = 0;
    double** x = nullptr;
    double** v = nullptr;
    double** omega = nullptr;
    double** f = nullptr;
    double** torque = nullptr;
    double* radius = nullptr;
    double* rmass = nullptr;
    int* tag = nullptr;
    int* type = nullptr;
    std::map<int, int> tagmap;
    int map(int t) const {
        auto it = tagmap.find(t);
        return it == tagmap.end() ? -1 : it->second;
    }
};

/* fix couple/lb/onetoone sink (fix_lb_coupling_onetoone.h:53-55): per-atom
 * dragforce / hdtorque arrays; comm_force_torque sums ghost slots into
 * their owners (the reverse communication of a one-rank run). */
struct FixLbCouplingOnetoone {
    double** get_force_ptr() { return f_ptr.data(); }
    double** get_torque_ptr() { return t_ptr.data(); }
    void comm_force_torque();
    bool use_torque = true;
    std::vector<double> f_store, t_store;
    std::vector<double*> f_ptr, t_ptr;
    AtomArrays* atom = nullptr;
};

class LiggghtsCouplingWrapper {
public:
    struct Lmp {
        AtomArrays* atom;
        FixLbCouplingOnetoone* fix;
    };
    LiggghtsCouplingWrapper() : lmp(&lmp_) { lmp_.atom = &atom; lmp_.fix = &fix; fix.atom = &atom; }
    LiggghtsCouplingWrapper(const LiggghtsCouplingWrapper&) = delete;

    /* (re)size to nlocal owned + nghost ghost particles */
    void resize(int nlocal, int nghost) {
        const int n = nlocal + nghost;
        atom.nlocal = nlocal;
        atom.nghost = nghost;
        sx.assign(3 * n, 0.0); sv.assign(3 * n, 0.0); sw.assign(3 * n, 0.0);
        sf.assign(3 * n, 0.0); st.assign(3 * n, 0.0);
        fix.f_store.assign(3 * n, 0.0); fix.t_store.assign(3 * n, 0.0);
        rad.assign(n, 0.0); mass.assign(n, 0.0); tags.assign(n, 0); types.assign(n, 1);
        px.resize(n); pv.resize(n); pw.resize(n); pf.resize(n); pt.resize(n);
        fix.f_ptr.resize(n); fix.t_ptr.resize(n);
        for (int i = 0; i < n; ++i) {
            px[i] = &sx[3 * i]; pv[i] = &sv[3 * i]; pw[i] = &sw[3 * i];
            pf[i] = &sf[3 * i]; pt[i] = &st[3 * i];
            fix.f_ptr[i] = &fix.f_store[3 * i]; fix.t_ptr[i] = &fix.t_store[3 * i];
        }
        atom.x = px.data(); atom.v = pv.data(); atom.omega = pw.data();
        atom.f = pf.data(); atom.torque = pt.data();
        atom.radius = rad.data(); atom.rmass = mass.data();
        atom.tag = tags.data(); atom.type = types.data();
    }
    /* map(tag) -> owned index (owned atoms first, ghosts after) */
    void rebuildMap() {
        atom.tagmap.clear();
        for (int i = atom.nlocal + atom.nghost - 1; i >= 0; --i) atom.tagmap[atom.tag[i]] = i;
    }
    long getNumParticles() const { return atom.nlocal; }
    FixLbCouplingOnetoone* couplingFix() { return &fix; }

    Lmp* lmp;
private:
    Lmp lmp_;
    AtomArrays atom;
    FixLbCouplingOnetoone fix;
    std::vector<double> sx, sv, sw, sf, st, rad, mass;
    std::vector<int> tags, types;
    std::vector<double*> px, pv, pw, pf, pt;
};

inline void FixLbCouplingOnetoone::comm_force_torque() {
    const int n = atom->nlocal + atom->nghost;
    for (int i = atom->nlocal; i < n; ++i) {
        const int o = atom->map(atom->tag[i]);
        if (o < 0 || o == i) continue;
        for (int d = 0; d < 3; ++d) {
            f_ptr[o][d] += f_ptr[i][d];
            t_ptr[o][d] += t_ptr[i][d];
            f_ptr[i][d] = 0.0;
            t_ptr[i][d] = 0.0;
        }
    }
}

/* ---- exchange wrappers (ibDataExchangeWrappers3D.h:37-53) ---- */
template <typename T, template <typename U> class Descriptor>
void setSpheresOnLattice(MultiBlockLattice3D<T, Descriptor>& lattice, LiggghtsCouplingWrapper& wrapper,
                         PhysUnits3D<T> const& units, std::vector<plint>& excludeType, bool initVelFlag) {
    AtomArrays* atom = wrapper.lmp->atom;
    const plint nPart = atom->nlocal + atom->nghost;
    const Box3D localBB = lattice.getBoundingBox();
    for (plint iS = 0; iS < nPart; ++iS) {
        bool exclude = false;
        for (size_t k = 0; k < excludeType.size(); ++k)
            if ((plint)atom->type[iS] == excludeType[k]) { exclude = true; break; }
        if (exclude) continue;
        T x[3], v[3], omega[3];
        for (int d = 0; d < 3; ++d) {
            x[d] = units.getLbPosition(atom->x[iS][d]);
            v[d] = units.getLbVel(atom->v[iS][d]);
            omega[d] = units.getLbFreq(atom->omega[iS][d]);
        }
        const T r = units.getLbLength(atom->radius[iS]);
        auto* sss = new SetSingleSphere3D<T, Descriptor>(x, v, omega, x, r, atom->tag[iS], initVelFlag);
        Box3D inter;
        if (intersect(sss->getBoundingBox(), localBB, inter)) applyProcessingFunctional(sss, inter, lattice);
        else delete sss;
    }
    applyProcessingFunctional(new AttributeFunctional<T, Descriptor>(), lattice.getBoundingBox(), lattice);
}

template <typename T, template <typename U> class Descriptor>
void setSpheresOnLattice(MultiBlockLattice3D<T, Descriptor>& lattice, LiggghtsCouplingWrapper& wrapper,
                         PhysUnits3D<T> const& units, bool initVelFlag) {
    std::vector<plint> none;
    setSpheresOnLattice(lattice, wrapper, units, none, initVelFlag);
}

/* getForcesFromLattice: forces in lattice units from the GPU, converted to
 * physical units into the coupling fix at map(tag), then the reverse sum. */
template <typename T, template <typename U> class Descriptor>
void getForcesFromLattice(MultiBlockLattice3D<T, Descriptor>& lattice, LiggghtsCouplingWrapper& wrapper,
                          PhysUnits3D<T> const& units) {
    AtomArrays* atom = wrapper.lmp->atom;
    FixLbCouplingOnetoone* fix = wrapper.lmp->fix;
    const plint nPart = atom->nlocal + atom->nghost;
    if (nPart == 0) return;
    /* the batch holds the particles that met the lattice, in call order */
    const std::vector<int64_t>& ids = lattice.sphereIds();
    const plint nSet = (plint)ids.size();
    typename ParticleData<T>::ParticleDataArrayVector x_lb(nSet);
    std::map<int64_t, plint> first;
    for (plint i = 0; i < nPart; ++i)
        if (!first.count(atom->tag[i])) first[atom->tag[i]] = i;
    for (plint k = 0; k < nSet; ++k) {
        const plint i = first[ids[k]];
        for (int d = 0; d < 3; ++d) x_lb[k][d] = units.getLbPosition(atom->x[i][d]);
    }
    std::vector<T> force(3 * nSet, 0.0), torque(3 * nSet, 0.0);
    if (nSet > 0)
        applyProcessingFunctional(new SumForceTorque3D<T, Descriptor>(x_lb, force.data(), torque.data(), wrapper),
                                  lattice.getBoundingBox(), lattice);
    double** f = fix->get_force_ptr();
    double** t = fix->get_torque_ptr();
    for (plint i = 0; i < nPart; ++i)
        for (int d = 0; d < 3; ++d) { f[i][d] = 0.0; t[i][d] = 0.0; }
    for (plint k = 0; k < nSet; ++k) {
        const int ind = atom->map((int)ids[k]);
        if (ind < 0) continue;
        for (int d = 0; d < 3; ++d) {
            f[ind][d] += units.getPhysForce(force[3 * k + d]);
            t[ind][d] += units.getPhysTorque(torque[3 * k + d]);
        }
    }
    fix->comm_force_torque();
}

}  // namespace plb

#endif

/*
 * CPU oracle of the LBDEMcoupling PSM step -- TEST INFRASTRUCTURE ONLY.
 * See psm_oracle.h for what it restates and who may use it.
 *
 * Compiled with -ffp-contract=off: every floating-point expression below is
 * evaluated exactly as written (left to right, no fused multiply-add), so
 * that the operation order is a fixed, documented contract.
 */
#include "psm_oracle.h"

#include <cmath>
#include <cstring>
#include <map>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const int Q = 19;
/* D3Q19 in the Palabos order: c_{i+9} = -c_i. */
const int C[Q][3] = {
    {0, 0, 0},   {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
    {-1, 0, 1},  {0, -1, -1}, {0, -1, 1}, {1, 0, 0},  {0, 1, 0},   {0, 0, 1},  {1, 1, 0},
    {1, -1, 0},  {1, 0, 1},  {1, 0, -1}, {0, 1, 1},  {0, 1, -1}};
const double T0 = 1.0 / 3.0, T1 = 1.0 / 18.0, T2 = 1.0 / 36.0;
const double W[Q] = {T0, T1, T1, T1, T2, T2, T2, T2, T2, T2,
                     T1, T1, T1, T2, T2, T2, T2, T2, T2};

/* src/ibDef.h:26-27 */
const double SOLFRAC_MIN = 0.001;
const double SOLFRAC_MAX = 0.999;

/* rhoBar = sum_i f_i (i ascending); j_d = signed sum over the non-zero c_id
 * (i ascending), starting from 0.0. */
inline void moments(const double* f, double& rhoBar, double* j) {
    rhoBar = 0.0;
    for (int i = 0; i < Q; ++i) rhoBar = rhoBar + f[i];
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int i = 0; i < Q; ++i) {
            if (C[i][d] == 1) s = s + f[i];
            else if (C[i][d] == -1) s = s - f[i];
        }
        j[d] = s;
    }
}

inline double cdot(int i, const double* v) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        if (C[i][d] == 1) s = s + v[d];
        else if (C[i][d] == -1) s = s - v[d];
    }
    return s;
}

/* Second-order equilibrium in the f - t form:
 * eq_i = t_i [ rhoBar + 3 c.j + 1.5 invRho (3 (c.j)^2 - |j|^2) ]. */
inline void equilibria(double rhoBar, double invRho, const double* j, double jSqr,
                       double* eq) {
    const double p = 1.5 * invRho;
    for (int i = 0; i < Q; ++i) {
        const double cj = cdot(i, j);
        const double a = 3.0 * cj;
        const double b = a * cj;
        const double q = b - jSqr;
        eq[i] = W[i] * ((rhoBar + a) + p * q);
    }
}

inline double norm_sqr(const double* v) { return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]; }

/* Base dynamics: BGK, with Guo forcing when a != NULL. */
inline void bgk(double* f, double omega, const double* a) {
    double rhoBar, j[3];
    moments(f, rhoBar, j);
    const double rho = rhoBar + 1.0;
    const double invRho = 1.0 / rho;
    double jEq[3] = {j[0], j[1], j[2]};
    double u[3];
    if (a) {
        for (int d = 0; d < 3; ++d) {
            u[d] = j[d] * invRho + 0.5 * a[d];
            jEq[d] = rho * u[d];
        }
    }
    const double jSqr = norm_sqr(jEq);
    double eq[Q];
    equilibria(rhoBar, invRho, jEq, jSqr, eq);
    const double omm = 1.0 - omega;
    for (int i = 0; i < Q; ++i) f[i] = f[i] * omm + omega * eq[i];
    if (a) {
        const double pref = (1.0 - 0.5 * omega) * rho;
        const double ua = (u[0] * a[0] + u[1] * a[1]) + u[2] * a[2];
        for (int i = 0; i < Q; ++i) {
            const double ca = cdot(i, a);
            const double cu = cdot(i, u);
            const double g = 3.0 * (ca - ua) + 9.0 * (cu * ca);
            f[i] = f[i] + (pref * W[i]) * g;
        }
    }
}

/* IBcompositeDynamics::collide, src/ibCompositeDynamics3D.hh:130-203.
 * f: the cell (pre-collision in, post-collision out). */
inline void psm_collide(double* f, double omega, double ooo, double sf, const double* us,
                        const double* a, double* F) {
    F[0] = F[1] = F[2] = 0.0;                      /* :137 */
    double fPre[Q];
    for (int i = 0; i < Q; ++i) fPre[i] = f[i];   /* :109-116 */
    if (sf <= SOLFRAC_MAX) bgk(f, omega, a);       /* :139-140 */
    if (sf < SOLFRAC_MIN) return;                  /* :143-144 */

    double rhoBar, j[3];
    moments(f, rhoBar, j);                         /* :146-148, post-BGK cell */
    const double invRho = 1.0 / (rhoBar + 1.0);
    const double jSqr = norm_sqr(j);
    double fEq[Q], fEqSolid[Q];
    equilibria(rhoBar, invRho, j, jSqr, fEq);      /* :150-152 */
    const double onePlus = 1.0 + rhoBar;
    double uP[3];
    for (int d = 0; d < 3; ++d) uP[d] = us[d] * onePlus;   /* :154 */
    const double uPSqr = norm_sqr(uP);
    equilibria(rhoBar, invRho, uP, uPSqr, fEqSolid);       /* :156-157 */

    if (sf > SOLFRAC_MAX) {                        /* :160-175 */
        const double bb0 = fEq[0] - fEqSolid[0];
        f[0] = fPre[0] + bb0;
        for (int i = 1; i <= 9; ++i) {
            const int o = i + 9;
            const double bb = (fPre[o] - fEq[o]) - (fPre[i] - fEqSolid[i]);
            const double bbo = (fPre[i] - fEq[i]) - (fPre[o] - fEqSolid[o]);
            f[i] = fPre[i] + bb;
            f[o] = fPre[o] + bbo;
            const double diff = bb - bbo;
            for (int d = 0; d < 3; ++d) F[d] = F[d] - (double)C[i][d] * diff;
        }
    } else {                                       /* :176-201 */
        const double B = (sf * ooo) / ((1.0 - sf) + ooo);
        const double oneMinB = 1.0 - B;
        const double bb0 = fEq[0] - fEqSolid[0];
        f[0] = (fPre[0] + oneMinB * (f[0] - fPre[0])) + B * bb0;
        for (int i = 1; i <= 9; ++i) {
            const int o = i + 9;
            const double bb = B * ((fPre[o] - fEq[o]) - (fPre[i] - fEqSolid[i]));
            const double bbo = B * ((fPre[i] - fEq[i]) - (fPre[o] - fEqSolid[o]));
            f[i] = (fPre[i] + oneMinB * (f[i] - fPre[i])) + bb;
            f[o] = (fPre[o] + oneMinB * (f[o] - fPre[o])) + bbo;
            const double diff = bb - bbo;
            for (int d = 0; d < 3; ++d) F[d] = F[d] - (double)C[i][d] * diff;
        }
    }
}

struct Record {
    double sf = 0.0;
    int64_t id = 0;
    double us[3] = {0.0, 0.0, 0.0};
    double F[3] = {0.0, 0.0, 0.0};
};

}  // namespace

struct orc_lattice {
    int n[3];
    int periodic[3];
    double tau, omega, ooo;
    int slices;
    bool has_force;
    double accel[3];
    int64_t nc;
    std::vector<double> f, ftmp;    /* AoS: f[c*Q + i] */
    std::vector<Record> rec;
    std::vector<int64_t> map_ids;   /* id -> first index of the last set_spheres */
    std::map<int64_t, int64_t> idmap;
    int64_t last_n_vox = 0;
    int voxel_mode = 0;
    int threads = 1;                /* x-slabs of set_spheres / sum_force_torque */
    /* periodic pressure */
    bool pp = false;
    int pp_dim = 0;
    double pp_in = 1.0, pp_out = 1.0;
    int pp_range[4] = {0, 0, 0, 0};
};

namespace {
/* cell (i along dim, a, b along the other two axes in increasing order) */
int64_t plane_cell(const orc_lattice* L, int dim, int i, int a, int b) {
    int x, y, z;
    if (dim == 0) { x = i; y = a; z = b; }
    else if (dim == 1) { x = a; y = i; z = b; }
    else { x = a; y = b; z = i; }
    return ((int64_t)x * L->n[1] + y) * L->n[2] + z;
}

/* computeAverageDensity over the plane's box: 1 + mean rhoBar */
double plane_avg_density(const orc_lattice* L, int plane) {
    double sum = 0.0;
    int64_t cnt = 0;
    for (int a = L->pp_range[0]; a <= L->pp_range[1]; ++a)
        for (int b = L->pp_range[2]; b <= L->pp_range[3]; ++b) {
            const int64_t c = plane_cell(L, L->pp_dim, plane, a, b);
            double rb = 0.0;
            for (int i = 0; i < Q; ++i) rb = rb + L->f[c * Q + i];
            sum += rb;
            ++cnt;
        }
    return 1.0 + sum / (double)cnt;
}

/* ZhangPeriodicPressureFunctional3D::process on the plane's box */
void zhang_rescale(orc_lattice* L, int plane, int dir, double rho_target, double rho_avg) {
    const double rescale = rho_target / rho_avg;
    for (int a = L->pp_range[0]; a <= L->pp_range[1]; ++a)
        for (int b = L->pp_range[2]; b <= L->pp_range[3]; ++b) {
            const int64_t c = plane_cell(L, L->pp_dim, plane, a, b);
            for (int i = 0; i < Q; ++i) {
                if (C[i][L->pp_dim] != dir) continue;
                double fTmp = L->f[c * Q + i] + W[i];
                fTmp *= rescale;
                L->f[c * Q + i] = fTmp - W[i];
            }
        }
}

double poiseuille_velocity(double i, double n) {
    const double nHalf = n / 2.;
    const double x = (i - nHalf) / (nHalf);
    return (1. - x * x);
}
}  // namespace

extern "C" {

orc_lattice* orc_create(int nx, int ny, int nz, const int* periodic, double tau,
                        int slices_per_dim, const double* body_accel) {
    if (nx < 1 || ny < 1 || nz < 1 || !(tau > 0.5) || slices_per_dim < 1 || slices_per_dim > 16)
        return nullptr;
    orc_lattice* L = new orc_lattice();
    L->n[0] = nx; L->n[1] = ny; L->n[2] = nz;
    for (int d = 0; d < 3; ++d) L->periodic[d] = periodic ? (periodic[d] != 0) : 1;
    L->tau = tau;
    L->omega = 1.0 / tau;
    L->ooo = 1.0 / L->omega - 0.5;   /* src/ibCompositeDynamics3D.hh:178 */
    L->slices = slices_per_dim;
    L->has_force = false;
    for (int d = 0; d < 3; ++d) {
        L->accel[d] = body_accel ? body_accel[d] : 0.0;
        if (L->accel[d] != 0.0) L->has_force = true;
    }
    L->nc = (int64_t)nx * ny * nz;
    L->f.assign(L->nc * Q, 0.0);
    L->ftmp.assign(L->nc * Q, 0.0);
    L->rec.assign(L->nc, Record());
    return L;
}

void orc_destroy(orc_lattice* L) { delete L; }

void orc_equilibrium(double rho_bar, const double* j3, double* eq19) {
    const double invRho = 1.0 / (rho_bar + 1.0);
    equilibria(rho_bar, invRho, j3, norm_sqr(j3), eq19);
}

void orc_init_equilibrium(orc_lattice* L, const double* rho, const double* u) {
    for (int64_t c = 0; c < L->nc; ++c) {
        const double r = rho ? rho[c] : 1.0;
        const double rhoBar = r - 1.0;
        double j[3];
        for (int d = 0; d < 3; ++d) j[d] = u ? r * u[d * L->nc + c] : 0.0;
        const double invRho = 1.0 / (rhoBar + 1.0);
        equilibria(rhoBar, invRho, j, norm_sqr(j), &L->f[c * Q]);
    }
}

void orc_set_populations(orc_lattice* L, const double* f) {
    for (int64_t c = 0; c < L->nc; ++c)
        for (int i = 0; i < Q; ++i) L->f[c * Q + i] = f[(int64_t)i * L->nc + c];
}

void orc_get_populations(const orc_lattice* L, double* f) {
    for (int64_t c = 0; c < L->nc; ++c)
        for (int i = 0; i < Q; ++i) f[(int64_t)i * L->nc + c] = L->f[c * Q + i];
}

/* SetSingleSphere3D::calcSolidFraction, src/ibProcessors3D.hh:116-158. */
double orc_calc_solid_fraction(double dx, double dy, double dz, double r, int N,
                               int32_t* n_out, int32_t* branch_out) {
    const double sliceWidth = 1.0 / (double)N;
    const double fraction = 1.0 / (double)(N * N * N);
    const double sqrt3half = std::sqrt(3.1) / 2.0;
    const double dist = dx * dx + dy * dy + dz * dz;
    const double r_p = r + sqrt3half;
    if (dist > r_p * r_p) {
        if (n_out) *n_out = 0;
        if (branch_out) *branch_out = 0;
        return 0.0;
    }
    const double r_m = r - sqrt3half;
    if (dist < r_m * r_m) {
        if (n_out) *n_out = N * N * N;
        if (branch_out) *branch_out = 1;
        return 1.0;
    }
    const double r_sq = r * r;
    double xs[16], ys[16], zs[16];
    for (int i = 0; i < N; ++i) {
        const double delta = -0.5 + ((double)i + 0.5) * sliceWidth;
        const double ax = dx + delta; xs[i] = ax * ax;
        const double ay = dy + delta; ys[i] = ay * ay;
        const double az = dz + delta; zs[i] = az * az;
    }
    uint32_t n = 0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) n += (xs[i] + ys[j] + zs[k] < r_sq);
    if (n_out) *n_out = (int32_t)n;
    if (branch_out) *branch_out = 2;
    return fraction * (double)n;
}

double orc_overlap_fraction(double d, double r, double* dsf_dd, int32_t* branch_out) {
    const double pi = 3.14159265358979323846;
    const double rc = std::cbrt(3.0 / (4.0 * pi));
    const double vc = ((4.0 / 3.0) * pi) * ((rc * rc) * rc);
    double sf, der = 0.0;
    int br;
    if (d >= r + rc) {
        sf = 0.0; br = 0;
    } else if (d <= std::fabs(r - rc)) {
        if (r >= rc) { sf = 1.0; br = 1; }
        else { sf = (((4.0 / 3.0) * pi) * ((r * r) * r)) / vc; br = 4; }
    } else {
        /* lens: V = pi t^2 (d^2 + 2 d (r + rc) - 3 (r - rc)^2) / (12 d),
         * t = r + rc - d; dV/dd = -pi a^2 with a the radius of the
         * intersection circle */
        const double t = (r + rc) - d;
        const double rm = r - rc;
        const double poly = ((d * d) + (2.0 * d) * (r + rc)) - (3.0 * rm) * rm;
        sf = (((pi * t) * t) * poly) / (12.0 * d) / vc;
        const double s = ((d * d) - (rc * rc)) + (r * r);
        double a2 = (((4.0 * d) * d) * (r * r) - s * s) / ((4.0 * d) * d);
        if (a2 < 0.0) a2 = 0.0;
        der = -(pi * a2) / vc;
        br = 3;
    }
    if (dsf_dd) *dsf_dd = der;
    if (branch_out) *branch_out = br;
    return sf;
}

void orc_set_voxel_mode(orc_lattice* L, int mode) { L->voxel_mode = mode ? 1 : 0; }

int orc_set_spheres(orc_lattice* L, int64_t n, const double* x3, const double* v3,
                    const double* w3, const double* r, const int64_t* id,
                    const int32_t* type, const int32_t* exclude, int n_exclude, int init_vel,
                    int auto_images) {
    if (n < 0) return -2;
    /* map(id): first index holding the id (owned atoms come first). */
    L->idmap.clear();
    for (int64_t q = 0; q < n; ++q)
        if (L->idmap.find(id[q]) == L->idmap.end()) L->idmap[id[q]] = q;

    struct Img { int64_t q; int s[3]; };
    std::vector<Img> list;
    list.reserve(n);
    for (int64_t q = 0; q < n; ++q) list.push_back({q, {0, 0, 0}});

    /* Periodic images in generation order: per periodic axis d, a lower-side
     * pass (image at x + L) then an upper-side pass (image at x - L), each
     * over every entry present when axis d started.  An image is kept when
     * its bounding box along d meets [0, n_d - 1]. */
    auto bbox_lo = [](double x, double rr) { return (int64_t)(x - rr - 2); };
    auto bbox_hi = [](double x, double rr) { return (int64_t)(x + rr + 2); };
    if (auto_images) {
        for (int d = 0; d < 3; ++d) {
            if (!L->periodic[d]) continue;
            const size_t m = list.size();
            for (int side = 0; side < 2; ++side) {
                const int sh = side == 0 ? 1 : -1;
                for (size_t e = 0; e < m; ++e) {
                    const Img src = list[e];
                    const double xd = x3[3 * src.q + d] + (double)sh * (double)L->n[d];
                    const int64_t lo = bbox_lo(xd, r[src.q]);
                    const int64_t hi = bbox_hi(xd, r[src.q]);
                    if (hi < 0 || lo > L->n[d] - 1) continue;
                    Img im = src;
                    im.s[d] = sh;
                    list.push_back(im);
                }
            }
        }
    }
    L->last_n_vox = 0;

    const int N = L->slices;
    /* src/ibDataExchangeWrappers3D.hh:62-140: the voxelized images in order,
     * each with its box intersected with the lattice */
    struct Vox { int64_t q; double x[3]; int64_t lo[3], hi[3]; };
    std::vector<Vox> vox;
    vox.reserve(list.size());
    for (const Img& im : list) {
        const int64_t q = im.q;
        bool excl = false;
        for (int t = 0; t < n_exclude; ++t)
            if (type && type[q] == exclude[t]) { excl = true; break; }
        if (excl) continue;
        Vox vx;
        vx.q = q;
        for (int d = 0; d < 3; ++d) vx.x[d] = x3[3 * q + d] + (double)im.s[d] * (double)L->n[d];
        const double rr = r[q];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {   /* getBoundingBox + intersect */
            vx.lo[d] = std::max<int64_t>(bbox_lo(vx.x[d], rr), 0);
            vx.hi[d] = std::min<int64_t>(bbox_hi(vx.x[d], rr), L->n[d] - 1);
            if (vx.lo[d] > vx.hi[d]) ok = false;
        }
        if (!ok) continue;
        vox.push_back(vx);
    }
    L->last_n_vox = (int64_t)vox.size();
    /* With threads > 1 each thread owns an x-slab of the lattice and runs the
     * whole particle loop on its part of every box, as each MPI rank of the
     * reference does on its own block; a cell is visited by one thread in
     * the same particle order, so the records are the same. */
    const int nthr = std::max(1, std::min(L->threads, L->n[0]));
#pragma omp parallel for num_threads(nthr) schedule(static, 1)
    for (int slab = 0; slab < nthr; ++slab) {
      const int64_t sx0 = (int64_t)L->n[0] * slab / nthr, sx1 = (int64_t)L->n[0] * (slab + 1) / nthr - 1;
      for (const Vox& vx : vox) {
        const int64_t q = vx.q;
        const double* x = vx.x;
        const double* v = &v3[3 * q];
        const double* w = &w3[3 * q];
        const double rr = r[q];
        const int pid = (int)id[q];
        const int64_t* lo = vx.lo;
        const int64_t* hi = vx.hi;
        const int64_t xa = std::max(lo[0], sx0), xb = std::min(hi[0], sx1);
        /* SetSingleSphere3D::process, src/ibProcessors3D.hh:52-114 */
        for (int64_t iX = xa; iX <= xb; ++iX)
            for (int64_t iY = lo[1]; iY <= hi[1]; ++iY)
                for (int64_t iZ = lo[2]; iZ <= hi[2]; ++iZ) {
                    const int64_t c = (iX * L->n[1] + iY) * L->n[2] + iZ;
                    Record& p = L->rec[c];
                    const double xg = (double)iX, yg = (double)iY, zg = (double)iZ;
                    const double dx = xg - x[0], dy = yg - x[1], dz = zg - x[2];
                    const double sf = L->voxel_mode
                        ? orc_overlap_fraction(std::sqrt((dx * dx + dy * dy) + dz * dz), rr, nullptr, nullptr)
                        : orc_calc_solid_fraction(dx, dy, dz, rr, N, nullptr, nullptr);
                    const double sf_old = p.sf;
                    const int id_old = (int)p.id;
                    const int decFlag = (sf > SOLFRAC_MIN) + 2 * (sf_old > SOLFRAC_MIN);
                    bool set = false, zero = false;
                    switch (decFlag) {
                        case 0: zero = true; break;
                        case 1: set = true; break;
                        case 2: if (id_old == pid) zero = true; break;
                        case 3: if (sf > sf_old || id_old == pid) set = true; break;
                    }
                    if (zero) {   /* setToZero, :176-184 */
                        p.us[0] = p.us[1] = p.us[2] = 0.0;
                        p.sf = 0.0;
                        p.id = 0;
                    }
                    if (set) {    /* setValues, :162-174 (com = x) */
                        p.us[0] = v[0]; p.us[1] = v[1]; p.us[2] = v[2];
                        p.us[0] = p.us[0] + (w[1] * dz - w[2] * dy);
                        p.us[1] = p.us[1] + (-w[0] * dz + w[2] * dx);
                        p.us[2] = p.us[2] + (w[0] * dy - w[1] * dx);
                        p.sf = sf;
                        p.id = pid;
                    }
                    if (init_vel && sf > SOLFRAC_MAX) {   /* :106-108 -> defineVelocity */
                        double* cell = &L->f[c * Q];
                        const double uSqr = norm_sqr(p.us);
                        equilibria(0.0, 1.0 / (0.0 + 1.0), p.us, uSqr, cell);
                    }
                }
      }
    }
    return 0;
}

void orc_set_threads(orc_lattice* L, int nthreads) { L->threads = nthreads > 0 ? nthreads : 1; }

void orc_collide_stream(orc_lattice* L, int nsteps, int nthreads) {
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    const int nt = 1;
#endif
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const double* a = L->has_force ? L->accel : nullptr;
    for (int step = 0; step < nsteps; ++step) {
        double avg_in = 0.0, avg_out = 0.0;
        if (L->pp) {
            avg_in = plane_avg_density(L, 0);
            avg_out = plane_avg_density(L, L->n[L->pp_dim] - 1);
        }
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int iX = 0; iX < nx; ++iX) {
            double cell[Q];
            for (int iY = 0; iY < ny; ++iY)
                for (int iZ = 0; iZ < nz; ++iZ) {
                    const int64_t c = ((int64_t)iX * ny + iY) * nz + iZ;
                    for (int i = 0; i < Q; ++i) cell[i] = L->f[c * Q + i];
                    Record& p = L->rec[c];
                    psm_collide(cell, L->omega, L->ooo, p.sf, p.us, a, p.F);
                    /* push streaming; half-way bounce-back where x + c_i leaves a
                     * non-periodic axis */
                    for (int i = 0; i < Q; ++i) {
                        int y[3] = {iX + C[i][0], iY + C[i][1], iZ + C[i][2]};
                        bool wall = false;
                        for (int d = 0; d < 3; ++d) {
                            if (y[d] < 0 || y[d] >= L->n[d]) {
                                if (L->periodic[d]) y[d] = (y[d] + L->n[d]) % L->n[d];
                                else wall = true;
                            }
                        }
                        if (wall) {
                            const int o = i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9);
                            L->ftmp[c * Q + o] = cell[i];
                        } else {
                            const int64_t cn = ((int64_t)y[0] * ny + y[1]) * nz + y[2];
                            L->ftmp[cn * Q + i] = cell[i];
                        }
                    }
                }
        }
        L->f.swap(L->ftmp);
        if (L->pp) {
            zhang_rescale(L, 0, 1, L->pp_in, avg_out);
            zhang_rescale(L, L->n[L->pp_dim] - 1, -1, L->pp_out, avg_in);
        }
    }
}

void orc_set_periodic_pressure(orc_lattice* L, int enable, int dim, double rho_in, double rho_out,
                               const int32_t* range4) {
    L->pp = enable != 0;
    L->pp_dim = dim;
    L->pp_in = rho_in;
    L->pp_out = rho_out;
    const int ax = dim == 0 ? 1 : 0, bx = dim == 2 ? 1 : 2;
    if (range4) {
        for (int k = 0; k < 4; ++k) L->pp_range[k] = range4[k];
    } else {
        L->pp_range[0] = 1; L->pp_range[1] = L->n[ax] - 2;
        L->pp_range[2] = 1; L->pp_range[3] = L->n[bx] - 2;
    }
}

void orc_init_poiseuille(orc_lattice* L, int dim, double pHi, double pLo, double uMax, const int32_t* n3) {
    int pn[3];
    for (int d = 0; d < 3; ++d) pn[d] = n3 ? n3[d] : L->n[d];
    for (int x = 0; x < L->n[0]; ++x)
        for (int y = 0; y < L->n[1]; ++y)
            for (int z = 0; z < L->n[2]; ++z) {
                double density, u[3] = {0.0, 0.0, 0.0};
                if (dim == 0) {
                    density = pHi - (pHi - pLo) * (double)x / (double)(pn[0] - 1);
                    u[0] = uMax * poiseuille_velocity((double)y, (double)pn[1]) * poiseuille_velocity((double)z, (double)pn[2]);
                } else if (dim == 1) {
                    density = pHi - (pHi - pLo) * (double)y / (double)(pn[1] - 1);
                    u[1] = uMax * poiseuille_velocity((double)x, (double)pn[0]) * poiseuille_velocity((double)z, (double)pn[2]);
                } else {
                    density = pHi - (pHi - pLo) * (double)z / (double)(pn[2] - 1);
                    u[2] = uMax * poiseuille_velocity((double)x, (double)pn[0]) * poiseuille_velocity((double)y, (double)pn[1]);
                }
                const double rhoBar = density - 1.0;
                double j[3];
                for (int d = 0; d < 3; ++d) j[d] = density * u[d];
                const int64_t c = ((int64_t)x * L->n[1] + y) * L->n[2] + z;
                equilibria(rhoBar, 1.0 / (rhoBar + 1.0), j, norm_sqr(j), &L->f[c * Q]);
            }
}

/* SumForceTorque3D::process, src/ibProcessors3D.hh:202-263, on one block
 * spanning the whole lattice. */
int orc_sum_force_torque(const orc_lattice* L, int64_t n, const double* x3, double* F3,
                         double* T3) {
    for (int64_t k = 0; k < 3 * n; ++k) { F3[k] = 0.0; T3[k] = 0.0; }
    const int64_t nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const int nthr = std::max(1, std::min<int>(L->threads, (int)nx));
    if (nthr > 1) {
        /* one partial sum per x-slab (an MPI rank's SumForceTorque3D over its
         * block), then the slabs in order: the sum's grouping differs from
         * the sequential scan, so this form serves timing, not parity */
        std::vector<std::vector<double>> part(nthr, std::vector<double>(6 * (size_t)n, 0.0));
#pragma omp parallel for num_threads(nthr) schedule(static, 1)
        for (int slab = 0; slab < nthr; ++slab) {
            const int64_t x0 = nx * slab / nthr, x1 = nx * (slab + 1) / nthr;
            double* Fp = part[slab].data();
            double* Tp = Fp + 3 * n;
            for (int64_t iX = x0; iX < x1; ++iX)
                for (int64_t iY = 0; iY < ny; ++iY)
                    for (int64_t iZ = 0; iZ < nz; ++iZ) {
                        const Record& p = L->rec[(iX * ny + iY) * nz + iZ];
                        if (p.id < 1) continue;
                        auto it = L->idmap.find(p.id);
                        if (it == L->idmap.end() || it->second >= n) continue;
                        const int64_t ind = it->second;
                        double d[3] = {(double)iX - x3[3 * ind], (double)iY - x3[3 * ind + 1],
                                       (double)iZ - x3[3 * ind + 2]};
                        const int64_t nn[3] = {nx, ny, nz};
                        for (int k = 0; k < 3; ++k) {
                            if (d[k] > (double)(nn[k] / 2)) d[k] -= (double)nn[k];
                            else if (d[k] < (double)(-nn[k] / 2)) d[k] += (double)nn[k];
                        }
                        Fp[3 * ind] += p.F[0]; Fp[3 * ind + 1] += p.F[1]; Fp[3 * ind + 2] += p.F[2];
                        Tp[3 * ind] += d[1] * p.F[2] - d[2] * p.F[1];
                        Tp[3 * ind + 1] += -d[0] * p.F[2] + d[2] * p.F[0];
                        Tp[3 * ind + 2] += d[0] * p.F[1] - d[1] * p.F[0];
                    }
        }
        for (int slab = 0; slab < nthr; ++slab)
            for (int64_t k = 0; k < 3 * n; ++k) {
                F3[k] += part[slab][k];
                T3[k] += part[slab][3 * n + k];
            }
        return 0;
    }
    for (int64_t iX = 0; iX < nx; ++iX)
        for (int64_t iY = 0; iY < ny; ++iY)
            for (int64_t iZ = 0; iZ < nz; ++iZ) {
                const Record& p = L->rec[(iX * ny + iY) * nz + iZ];
                const int64_t id = p.id;
                if (id < 1) continue;
                auto it = L->idmap.find(id);
                if (it == L->idmap.end() || it->second >= n) continue;
                const int64_t ind = it->second;
                double dx = (double)iX - x3[3 * ind + 0];
                double dy = (double)iY - x3[3 * ind + 1];
                double dz = (double)iZ - x3[3 * ind + 2];
                /* minimum image with the block size, integer halves (:238-243) */
                if (dx > (double)(nx / 2)) dx -= (double)nx;
                else if (dx < (double)(-nx / 2)) dx += (double)nx;
                if (dy > (double)(ny / 2)) dy -= (double)ny;
                else if (dy < (double)(-ny / 2)) dy += (double)ny;
                if (dz > (double)(nz / 2)) dz -= (double)nz;
                else if (dz < (double)(-nz / 2)) dz += (double)nz;
                const double fx = p.F[0], fy = p.F[1], fz = p.F[2];
                const double tx = dy * fz - dz * fy;
                const double ty = -dx * fz + dz * fx;
                const double tz = dx * fy - dy * fx;
                F3[3 * ind + 0] += fx; F3[3 * ind + 1] += fy; F3[3 * ind + 2] += fz;
                T3[3 * ind + 0] += tx; T3[3 * ind + 1] += ty; T3[3 * ind + 2] += tz;
            }
    return 0;
}

void orc_get_cell_records(const orc_lattice* L, double* sf, int64_t* id, double* us,
                          double* F) {
    for (int64_t c = 0; c < L->nc; ++c) {
        const Record& p = L->rec[c];
        if (sf) sf[c] = p.sf;
        if (id) id[c] = p.id;
        for (int d = 0; d < 3; ++d) {
            if (us) us[d * L->nc + c] = p.us[d];
            if (F) F[d * L->nc + c] = p.F[d];
        }
    }
}

void orc_set_cell_records(orc_lattice* L, const double* sf, const int64_t* id,
                          const double* us) {
    for (int64_t c = 0; c < L->nc; ++c) {
        Record& p = L->rec[c];
        p.sf = sf ? sf[c] : 0.0;
        p.id = id ? id[c] : 0;
        for (int d = 0; d < 3; ++d) p.us[d] = us ? us[d * L->nc + c] : 0.0;
    }
}

void orc_conservation(const orc_lattice* L, double* mass, double* mom3) {
    double m = 0.0, j[3] = {0.0, 0.0, 0.0};
    for (int64_t c = 0; c < L->nc; ++c) {
        double rb, jj[3];
        moments(&L->f[c * Q], rb, jj);
        m += rb;
        for (int d = 0; d < 3; ++d) j[d] += jj[d];
    }
    *mass = m;
    for (int d = 0; d < 3; ++d) mom3[d] = j[d];
}

void orc_collide_cell(double* f19, double tau, double sf, const double* us3, double* F3) {
    const double omega = 1.0 / tau;
    const double ooo = 1.0 / omega - 0.5;
    psm_collide(f19, omega, ooo, sf, us3, nullptr, F3);
}

int64_t orc_last_n_voxelized(const orc_lattice* L) { return L->last_n_vox; }

}  // extern "C"

/*
 * Shared definitions of the gfx950 PSM library: D3Q19 tables, the per-cell
 * PSM arithmetic as device functions, the device-side views the kernels
 * take, and the handle.
 *
 * Floating-point contract: the FP64 path is compiled with
 * -ffp-contract=off and evaluates every expression in the written order,
 * so that it reproduces the FP64 CPU restatement bit for bit.
 */
#ifndef LBDEM_PSM_COMMON_H
#define LBDEM_PSM_COMMON_H

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/lbdem_psm.h"

namespace psm {

constexpr int Q = 19;
constexpr int BIN = 8;              /* voxelization bin edge (cells) */
constexpr int BIN_CELLS = BIN * BIN * BIN;
constexpr int QBITS = 27;           /* candidate key: rank << 27 | particle */
constexpr uint32_t QMASK = (1u << QBITS) - 1u;
constexpr double SOLFRAC_MIN = 0.001;   /* src/ibDef.h:26 */
constexpr double SOLFRAC_MAX = 0.999;   /* src/ibDef.h:27 */

/* D3Q19, Palabos order: c_{i+9} = -c_i. */
__host__ __device__ constexpr int cx(int i) {
    constexpr int v[Q] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
    return v[i];
}
__host__ __device__ constexpr int cy(int i) {
    constexpr int v[Q] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
    return v[i];
}
__host__ __device__ constexpr int cz(int i) {
    constexpr int v[Q] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
    return v[i];
}
__host__ __device__ constexpr int cc(int i, int d) { return d == 0 ? cx(i) : (d == 1 ? cy(i) : cz(i)); }
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }

template <typename T>
__host__ __device__ inline T wgt(int i) {
    return i == 0 ? T(1.0 / 3.0) : ((i <= 3 || (i >= 10 && i <= 12)) ? T(1.0 / 18.0) : T(1.0 / 36.0));
}

/* ---- per-cell arithmetic (same operation order as the CPU restatement) ---- */

template <typename T>
__device__ __forceinline__ void moments(const T* f, T& rhoBar, T* j) {
    rhoBar = T(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) rhoBar = rhoBar + f[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T s = T(0);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (cc(i, d) == 1) s = s + f[i];
            else if (cc(i, d) == -1) s = s - f[i];
        }
        j[d] = s;
    }
}

template <typename T>
__device__ __forceinline__ T cdot(int i, const T* v) {
    if (i == 0) return T(0);
    T s = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (cc(i, d) == 1) s = s + v[d];
        else if (cc(i, d) == -1) s = s - v[d];
    }
    return s;
}

template <typename T>
__device__ __forceinline__ T norm_sqr(const T* v) { return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]; }

template <typename T>
__device__ __forceinline__ T equilibrium(int i, T rhoBar, T p, const T* j, T jSqr) {
    const T cj = cdot<T>(i, j);
    const T a = T(3) * cj;
    const T b = a * cj;
    const T q = b - jSqr;
    return wgt<T>(i) * ((rhoBar + a) + p * q);
}

/* Base dynamics: BGK, with Guo forcing when FORCE. */
template <typename T, bool FORCE>
__device__ __forceinline__ void bgk(T* f, T omega, const T* acc) {
    T rhoBar, j[3];
    moments<T>(f, rhoBar, j);
    const T rho = rhoBar + T(1);
    const T invRho = T(1) / rho;
    T jEq[3] = {j[0], j[1], j[2]};
    T u[3] = {T(0), T(0), T(0)};
    if (FORCE) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            u[d] = j[d] * invRho + T(0.5) * acc[d];
            jEq[d] = rho * u[d];
        }
    }
    const T jSqr = norm_sqr<T>(jEq);
    const T p = T(1.5) * invRho;
    const T omm = T(1) - omega;
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = f[i] * omm + omega * equilibrium<T>(i, rhoBar, p, jEq, jSqr);
    if (FORCE) {
        const T pref = (T(1) - T(0.5) * omega) * rho;
        const T ua = (u[0] * acc[0] + u[1] * acc[1]) + u[2] * acc[2];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const T ca = cdot<T>(i, acc);
            const T cu = cdot<T>(i, u);
            const T g = T(3) * (ca - ua) + T(9) * (cu * ca);
            f[i] = f[i] + (pref * wgt<T>(i)) * g;
        }
    }
}

/* IBcompositeDynamics::collide (src/ibCompositeDynamics3D.hh:130-203) for a
 * cell with a coupling record (sf > SOLFRAC_MIN).  fPre: pre-collision
 * populations; f: output.  full = (sf > SOLFRAC_MAX) and the weight
 * B = sf (tau - 1/2) / ((1 - sf) + (tau - 1/2)) (:177-179) are decided by
 * the caller in FP64. */
template <typename T, bool FORCE>
__device__ __forceinline__ void psm_solid(const T* fPre, T* f, T omega, bool full, T B,
                                          const T* us, const T* acc, T* F) {
    F[0] = T(0); F[1] = T(0); F[2] = T(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = fPre[i];
    if (!full) bgk<T, FORCE>(f, omega, acc);
    T rhoBar, j[3];
    moments<T>(f, rhoBar, j);
    const T invRho = T(1) / (rhoBar + T(1));
    const T jSqr = norm_sqr<T>(j);
    const T p = T(1.5) * invRho;
    const T onePlus = T(1) + rhoBar;
    T uP[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) uP[d] = us[d] * onePlus;
    const T uPSqr = norm_sqr<T>(uP);
    if (full) {
        const T bb0 = equilibrium<T>(0, rhoBar, p, j, jSqr) - equilibrium<T>(0, rhoBar, p, uP, uPSqr);
        f[0] = fPre[0] + bb0;
#pragma unroll
        for (int i = 1; i <= 9; ++i) {
            const int o = i + 9;
            const T Ei = equilibrium<T>(i, rhoBar, p, j, jSqr);
            const T Eo = equilibrium<T>(o, rhoBar, p, j, jSqr);
            const T Si = equilibrium<T>(i, rhoBar, p, uP, uPSqr);
            const T So = equilibrium<T>(o, rhoBar, p, uP, uPSqr);
            const T bb = (fPre[o] - Eo) - (fPre[i] - Si);
            const T bbo = (fPre[i] - Ei) - (fPre[o] - So);
            f[i] = fPre[i] + bb;
            f[o] = fPre[o] + bbo;
            const T diff = bb - bbo;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (cc(i, d) == 1) F[d] = F[d] - diff;
                else if (cc(i, d) == -1) F[d] = F[d] + diff;
            }
        }
    } else {
        const T oneMinB = T(1) - B;
        const T bb0 = equilibrium<T>(0, rhoBar, p, j, jSqr) - equilibrium<T>(0, rhoBar, p, uP, uPSqr);
        f[0] = (fPre[0] + oneMinB * (f[0] - fPre[0])) + B * bb0;
#pragma unroll
        for (int i = 1; i <= 9; ++i) {
            const int o = i + 9;
            const T Ei = equilibrium<T>(i, rhoBar, p, j, jSqr);
            const T Eo = equilibrium<T>(o, rhoBar, p, j, jSqr);
            const T Si = equilibrium<T>(i, rhoBar, p, uP, uPSqr);
            const T So = equilibrium<T>(o, rhoBar, p, uP, uPSqr);
            const T bb = B * ((fPre[o] - Eo) - (fPre[i] - Si));
            const T bbo = B * ((fPre[i] - Ei) - (fPre[o] - So));
            f[i] = (fPre[i] + oneMinB * (f[i] - fPre[i])) + bb;
            f[o] = (fPre[o] + oneMinB * (f[o] - fPre[o])) + bbo;
            const T diff = bb - bbo;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (cc(i, d) == 1) F[d] = F[d] - diff;
                else if (cc(i, d) == -1) F[d] = F[d] + diff;
            }
        }
    }
}

/* Minimum image of a torque arm component over the lattice size n
 * (src/ibProcessors3D.hh:238-243, integer halves of the block size). */
__device__ __forceinline__ double min_image(double d, int n) {
    /* selects, not branches: the same two candidate values either way */
    const double h = (double)(n / 2), l = (double)(-n / 2);
    const double up = d - (double)n, dn = d + (double)n;
    return d > h ? up : (d < l ? dn : d);
}

/* Analytic sphere-cell overlap, the differentiable voxelization: the
 * fraction of a unit-volume cell covered by a sphere of radius r at centre
 * distance d, the cell taken as the sphere of equal volume (radius rc =
 * (3/(4 pi))^(1/3), volume vc) at its centre.  Lens volume of two spheres,
 * V = pi t^2 (d^2 + 2 d (r + rc) - 3 (r - rc)^2) / (12 d), t = r + rc - d;
 * its derivative is dV/dd = -pi a^2, a the radius of the intersection
 * circle.  br: 0 outside, 1 cell inside the sphere, 3 lens, 4 sphere inside
 * the cell sphere. */
__host__ __device__ inline double overlap_fraction(double d, double r, double rc, double vc,
                                                   double* dsf_dd, int* br) {
    const double pi = 3.14159265358979323846;
    double sf, der = 0.0;
    int b;
    if (d >= r + rc) {
        sf = 0.0; b = 0;
    } else if (d <= fabs(r - rc)) {
        if (r >= rc) { sf = 1.0; b = 1; }
        else { sf = (((4.0 / 3.0) * pi) * ((r * r) * r)) / vc; b = 4; }
    } else {
        const double t = (r + rc) - d;
        const double rm = r - rc;
        const double poly = ((d * d) + (2.0 * d) * (r + rc)) - (3.0 * rm) * rm;
        sf = (((pi * t) * t) * poly) / (12.0 * d) / vc;
        const double s = ((d * d) - (rc * rc)) + (r * r);
        double a2 = (((4.0 * d) * d) * (r * r) - s * s) / ((4.0 * d) * d);
        if (a2 < 0.0) a2 = 0.0;
        der = -(pi * a2) / vc;
        b = 3;
    }
    if (dsf_dd) *dsf_dd = der;
    if (br) *br = b;
    return sf;
}

/* ---- storage conversions ---- */
template <int PREC> struct Store;
template <> struct Store<LBDEM_PSM_FP64> {
    typedef double S; typedef double T;
    static __device__ __forceinline__ T load(const S* p) { return *p; }
    static __device__ __forceinline__ void store(S* p, T v) { *p = v; }
};
template <> struct Store<LBDEM_PSM_FP32_DEV> {
    typedef float S; typedef float T;
    static __device__ __forceinline__ T load(const S* p) { return *p; }
    static __device__ __forceinline__ void store(S* p, T v) { *p = v; }
};
template <> struct Store<LBDEM_PSM_FP16_STORE> {
    typedef _Float16 S; typedef float T;
    static __device__ __forceinline__ T load(const S* p) { return (float)(*p); }
    static __device__ __forceinline__ void store(S* p, T v) { *p = (_Float16)v; }
};

/* Coupling records (IBdynamicsParticleData), one slot per cell (slot = cell
 * index, rec_index).  A record is live when the cell's solid flag is set.
 * The values the collide reads (sf, u_s) and the force it writes sit
 * together per slot, 32 bytes (16 in float): a wavefront whose solid cells
 * are a short run of its lanes touches the few lines that hold them, where
 * one array per component spread the same run over a line of every array. */
struct Records {
    double* sus;     /* 4*cap: per slot (sf, us_x, us_y, us_z): the solid fraction and the
                      * particle velocity at the cell */
    int32_t* id;     /* particle tag */
    double* F;       /* 4*cap per slot (F_x, F_y, F_z, 0): the hydrodynamic force of the last
                      * collision, double with FP64 storage; with FP32/FP16 storage the first
                      * 4*cap floats of the buffer hold it as float */
    int32_t* cand;   /* candidate entry of the setter this step, -1 = none */
    float* m4;       /* FP32/FP16 storage only, 4*cap: per cell (s, us) in float for the
                      * collide, s = 2 for sf > SOLFRAC_MAX, else (float)sf; valid where the
                      * cell holds a live record (null with FP64 storage) */
};

/* The float view of a live record that the FP32/FP16 collide reads. */
__host__ __device__ inline void record_mirror(double sf, const double* us, float* out4) {
    out4[0] = sf > SOLFRAC_MAX ? 2.0f : (float)sf;
    out4[1] = (float)us[0];
    out4[2] = (float)us[1];
    out4[3] = (float)us[2];
}

/* A candidate particle image of one bin, gathered in key order (entry e of
 * the candidate array).  The first 80 bytes are what a fold wavefront reads
 * for every candidate it tests, one contiguous block of scalar loads; the
 * velocities are read once per cell the candidate sets. */
struct Cand {
    double x[3];     /* image centre */
    double rp2, rm2, rsq;
    double ri2, ro2;    /* squared radii of the exact no-count band (sub-cell rule) */
    uint64_t yzmask;    /* clipped bounding box within the entry's bin: bit 8 y + z */
    uint32_t xmask;     /* ... and bit x (bin-local coordinates) */
    int id;
    double r;
    double v[3];
    double w[3];
};

/* Record slot of the cell (x, y, z): the cell index itself, so that a
 * z-run of cells reads its records as contiguous stretches in the order the
 * collide sweep visits them. */
__host__ __device__ inline int64_t rec_index(int x, int y, int z, int ny, int nz) {
    return ((int64_t)x * ny + y) * nz + z;
}

__host__ __device__ inline void rec_cell(int64_t r, int ny, int nz, int& x, int& y, int& z) {
    z = (int)(r % nz);
    const int64_t xy = r / nz;
    y = (int)(xy % ny);
    x = (int)(xy / ny);
}

/* Population store layout, z-chunked row blocks: the z-line of row
 * r = x*ny + y is cut into nzc chunks of zc cells, and each (row, chunk)
 * block holds its 19 populations back to back,
 *   f[((r*nzc + z/zc)*19 + q)*zc + z%zc].
 * A workgroup owns one (row, chunk) block: its 19 loads read one contiguous
 * 19*zc-element stretch, and its pushes land in the blocks of the 8
 * neighbouring rows (and, across a chunk boundary, of the next chunk).
 * zc is a multiple of 64; with more than one chunk it is a power of two
 * (zsh = log2 zc). */
/* Workgroups are handed to the 8 XCDs round-robin (workgroup i to XCD
 * i % 8), each XCD with its own L2.  Renumbering them so that XCD k sweeps
 * one contiguous range of the grid keeps neighbouring blocks (the z chunks
 * of a row, adjacent rows) in one L2: a line that two neighbouring blocks
 * both write is then assembled in that L2 instead of leaving two partial
 * lines in two of them.  A bijection of [0, G). */
__device__ __forceinline__ int xcd_block(int i, int G) {
    const int xcd = i & 7, local = i >> 3;
    const int per = G >> 3, rem = G & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

/* Division by a run-time constant d > 0 for dividends n < 2^31, as a
 * multiply-high, an add and a shift (Granlund and Montgomery, "Division by
 * invariant integers using multiplication", PLDI 1994): with
 * l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
 * n / d = (mulhi(m, n) + n) >> l.  mulhi(m, n) < n, so the sum fits 32 bits
 * for n < 2^31.  The kernels' block and bin index splits used a general
 * division here (a software sequence of ~30 scalar instructions each). */
struct FastDiv {
    uint32_t d, m, l;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < (uint64_t)d) ++l;
    f.l = l;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    return f;
}

__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    const uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);   /* a multiply-high on the device */
    return (t + n) >> f.l;
}

/* Broadcast of lane L's double (L wave-uniform) into scalar registers. */
__device__ __forceinline__ double readlane_d(double v, int L) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, L);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), L);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* Cross-lane moves of a double that stay in the VALU (no LDS round trip).
 * swap32(a, b) exchanges lanes 32-63 of a with lanes 0-31 of b
 * (v_permlane32_swap); afterwards a + b is, on lanes 0-31, a[l] + a[l + 32]
 * and, on lanes 32-63, b[l - 32] + b[l].  swap16 does the same between
 * odd and even rows of 16 lanes (v_permlane16_swap). */
__device__ __forceinline__ void swap32(double& a, double& b) {
    const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
    const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ void swap16(double& a, double& b) {
    const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
    const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
/* DPP move of a double inside a row of 16 lanes: lanes whose bank (four
 * lanes) is in bank_mask take src from the lane ctrl names, the others keep
 * old. */
template <int CTRL, int BANKS>
__device__ __forceinline__ double dpp_d(double old, double src) {
    const unsigned long long uo = (unsigned long long)__double_as_longlong(old);
    const unsigned long long us = (unsigned long long)__double_as_longlong(src);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)uo, (int)(unsigned)us, CTRL, 0xf, BANKS, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(uo >> 32), (int)(unsigned)(us >> 32), CTRL,
                                                              0xf, BANKS, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
constexpr int DPP_QUAD_XOR1 = 0xB1;       /* quad_perm [1, 0, 3, 2] */
constexpr int DPP_QUAD_XOR2 = 0x4E;       /* quad_perm [2, 3, 0, 1] */
constexpr int DPP_ROW_ROR8 = 0x128;       /* row_ror:8, lane l <- lane l ^ 8 of its row */
constexpr int DPP_ROW_HALF_MIRROR = 0x141;   /* lane l <- lane 7 - l of its eight */

/* Sum of v[0..7] over the 64 lanes of a wavefront, reduce-scatter style:
 * three halving exchanges (lane bits 5, 4, 3) leave each lane one value,
 * index m = 4 b5 + 2 b4 + b3 of the lane, summed over the eight lanes that
 * share those bits; three more steps inside each group of eight (quad
 * exchanges, then the two quads) finish it.  The exchanges are permlane
 * swaps and DPP moves, so the whole tree runs in the VALU without waiting
 * on the LDS.  Every addition has a fixed pair of operands (a + b and b + a
 * round alike), so the sum is deterministic.  All 64 lanes must be
 * active. */
__device__ __forceinline__ double wave_sum8_scatter(const double (&v)[8], int lane) {
    double w4[4], w2[2];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        /* lanes 0-31 sum value m, lanes 32-63 value 4 + m */
        double a = v[m], b = v[4 + m];
        swap32(a, b);
        w4[m] = a + b;
    }
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        /* even rows value m (of the half), odd rows 2 + m */
        double a = w4[m], b = w4[2 + m];
        swap16(a, b);
        w2[m] = a + b;
    }
    /* lanes 0-7 of a row keep w2[0] and add it from lanes 8-15; lanes 8-15
     * keep w2[1] and add it from lanes 0-7 */
    const bool u3 = (lane & 8) != 0;
    double part = dpp_d<DPP_ROW_ROR8, 0x3>(0.0, w2[0]);
    part = dpp_d<DPP_ROW_ROR8, 0xC>(part, w2[1]);
    double w = (u3 ? w2[1] : w2[0]) + part;
    w += dpp_d<DPP_QUAD_XOR1, 0xF>(0.0, w);
    w += dpp_d<DPP_QUAD_XOR2, 0xF>(0.0, w);
    w += dpp_d<DPP_ROW_HALF_MIRROR, 0xF>(0.0, w);
    return w;
}

struct PopLayout {
    int zc;
    int zsh;
    int nzc;
    __host__ __device__ int64_t rowlen() const { return (int64_t)nzc * Q * zc; }   /* elements per row */
    /* offset of z inside a row, population 0 */
    __host__ __device__ uint32_t zoff(int z) const {
        return nzc > 1 ? (uint32_t)(z >> zsh) * (uint32_t)(Q * zc) + (uint32_t)(z & (zc - 1)) : (uint32_t)z;
    }
};

__host__ __device__ inline size_t pop_at(int64_t row, int q, int z, const PopLayout& L) {
    return (size_t)row * (size_t)L.rowlen() + (size_t)q * (size_t)L.zc + (size_t)L.zoff(z);
}

/* Chunk length for a z extent: one chunk (nz rounded up to 64) up to 256
 * cells, else the largest power of two in {256, 128, 64} that pads by at
 * most nz/8. */
inline PopLayout make_layout(int nz) {
    PopLayout L;
    if (nz <= 256) {
        L.zc = (nz + 63) / 64 * 64;
        L.zsh = 0;
        L.nzc = 1;
        return L;
    }
    const int cand[3] = {256, 128, 64};
    const int sh[3] = {8, 7, 6};
    for (int k = 0; k < 3; ++k) {
        const int nzc = (nz + cand[k] - 1) / cand[k];
        if ((int64_t)(nzc * cand[k] - nz) * 8 <= nz || k == 2) {
            L.zc = cand[k];
            L.zsh = sh[k];
            L.nzc = nzc;
            return L;
        }
    }
    return L;
}

struct Geom {
    int n[3];
    int periodic[3];
    int nb[3];       /* bins per axis */
    int64_t nc;
    PopLayout L;     /* population store layout */
};

}  // namespace psm

/* The handle. */
struct lbdem_psm {
    lbdem_psm_desc desc;
    psm::Geom g;
    double omega, ooo, sqrt3half;
    double cell_rc, cell_vc;   /* radius and volume of the equal-volume cell sphere */
    bool has_force;
    int precision;
    size_t pop_bytes;          /* bytes of one population buffer (row-blocked layout) */
    hipStream_t stream = nullptr;
    void* f[2] = {nullptr, nullptr};
    int cur_f = 0;
    /* dynamic LDS of the collide launches [body force][periodic z] that caps
     * their occupancy (collide_wave_cap), set at the first launch */
    unsigned cs_lds[2][2] = {{0u, 0u}, {0u, 0u}};
    bool cs_lds_ready = false;
    uint8_t* sflag = nullptr;        /* per cell: 1 when the cell holds a coupling record */
    psm::Records rec;                /* dense, bin-permuted, updated in place */
    int64_t rec_cap = 0;             /* one slot per cell */
    int64_t nbins = 0;
    uint32_t* bin_any[2] = {nullptr, nullptr};   /* per bin: live records (this / next step) */
    int cur_any = 0;
    int64_t max_records = 0;         /* optional limit on live records, 0 = none */

    /* particles (device) */
    int64_t np = 0, p_cap = 0;
    double *px = nullptr, *pv = nullptr, *pw = nullptr, *pr = nullptr;
    int32_t *pid = nullptr, *pexcl = nullptr, *powner = nullptr;
    int32_t* id_map = nullptr;       /* id -> first particle index */
    int64_t id_map_size = 0;
    std::vector<int32_t> h_owner;    /* host copy of powner */
    bool owner_identity = true;      /* h_owner[q] == q for every q (no user ghosts) */
    std::vector<int64_t> h_ids;
    /* the mapping inputs the device's pid / pexcl / powner / id_map were built
     * from (reused while a set_spheres call passes the same ones) */
    std::vector<int32_t> h_types, h_exclude;
    bool h_has_types = false;
    bool ids_valid = false;
    /* pinned host staging of the exchange calls: two buffers used in turn,
     * each free again once the event recorded after its copies completes */
    struct Stage { void* buf = nullptr; size_t bytes = 0; hipEvent_t ev = nullptr; bool inflight = false; } stage[2];
    int stage_next = 0;
    double* h_io = nullptr;          /* pinned: sum_force_torque positions in, forces out */
    size_t h_io_bytes = 0;
    bool pending_counters = false;   /* h_counters holds an unchecked copy (set_spheres) */
    int32_t flags = 0;
    bool have_particles = false;
    bool voxelized = false;
    bool records_from_host = false;   /* records last set by set_cell_records */

    /* binning */
    int32_t *bin_cand_count = nullptr, *bin_cand_start = nullptr, *bin_fill = nullptr;
    uint32_t* cand_key = nullptr;
    int32_t* cand_bin = nullptr;
    psm::Cand* cand_data = nullptr;  /* candidates in key order, gathered per step */
    int64_t ent_cap = 0;
    int32_t *p_nent = nullptr, *p_ent_start = nullptr, *owner_ent = nullptr;
    int32_t* orphan = nullptr;       /* record indices without a candidate this step */
    int64_t orphan_cap = 0;
    /* stale records grouped by owner in cell order (order_orphans): per
     * particle count / fill cursor / start, and the placed and sorted cells */
    int32_t *orph_cnt = nullptr, *orph_fill = nullptr, *orph_start = nullptr;
    int32_t *orph_slot = nullptr, *orph_sorted = nullptr;
    double* partial = nullptr;       /* 6 per candidate entry */
    double* pft = nullptr;           /* 6 per particle: F[3], T[3] */
    double* psum_x = nullptr;        /* positions for torque arms */
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    int64_t* counters = nullptr;     /* device: [0] live records, [1] overflow flags, [2] orphans,
                                        [3] unused, [4] images, [5] defineVelocity cells */
    int64_t* h_counters = nullptr;   /* pinned host mirror */
    double* red_tmp = nullptr;       /* reduction scratch */
    int32_t* dv_cell = nullptr;      /* defineVelocity list (initVel), lazily allocated */
    double* dv_us = nullptr;
    int64_t red_blocks = 0;

    /* Zhang-Kwok periodic pressure across one periodic axis */
    struct PeriodicPressure {
        bool enabled = false;
        int dim = 0;
        double rho_in = 1.0, rho_out = 1.0;
        int range[4] = {0, 0, 0, 0};   /* inclusive cross-section box on the other two axes */
        double* buf = nullptr;         /* [0] inlet rhoBar sum, [1] outlet sum, partials from [8] */
        double* cells = nullptr;       /* 2 x box cells: each plane cell's rhoBar, written by the collide */
        int64_t cells_cap = 0;
    } pp;

    /* differentiable window */
    int64_t geom_gen = 0;            /* bumped by every change of records or particles */
    struct Diff {
        bool ready = false;
        int64_t gen = -1;
        int32_t T = 0, K = 0;
        int64_t np = 0;
        std::vector<void*> ckpt;     /* population state at each segment start */
        std::vector<int32_t> seg_t0; /* segment start steps of the window */
        std::vector<void*> seg;      /* recomputed states of one segment */
        void* lam[2] = {nullptr, nullptr};
        double* d_ft = nullptr;      /* T*np*6, per-particle (F, T) of every step */
        double* gacc = nullptr;      /* 7 per record (list position): gB, gus[3], F x gT[3] */
        double* Fscr = nullptr;      /* 3 per record slot: forces of recomputed steps */
        double* opart = nullptr;     /* omega-gradient partial per workgroup */
        int64_t opart_n = 0;
        double* gpart = nullptr;     /* 9 per particle */
        double* cell_out = nullptr;  /* 2*Nc dense gsf, gB */
        double* tau_part = nullptr;
        int64_t tau_part_n = 0;
        void* pb = nullptr;          /* 19 per record (list position): stage-1 adjoint of the record cells */
        void* pc = nullptr;          /* 5 per record (list position): coefficients of the BGK adjoint */
        int32_t* rec_list = nullptr; /* record slots of the frozen geometry */
        int32_t* rpos = nullptr;     /* per cell: position in rec_list, -1 without a record */
        int32_t* rown = nullptr;     /* per list position: the record's particle index, -1 for none */
        int64_t pcap = 0;            /* entries of gacc, pb, pc (the record count) */
        size_t ptsz = 0;             /* arithmetic-type bytes of pb, pc */
        bool pp = false;             /* the window ran with the periodic-pressure rescale */
        void* pp_last = nullptr;     /* periodic pressure: inlet/outlet planes of f(T) */
        void* pp_next = nullptr;     /* ... of the state after the current segment */
        void* pp_cur = nullptr;      /* ... of f(t+1) inside a segment */
        int64_t* bin_off = nullptr;  /* per bin: record count, then the bin's first list position */
        int64_t nrec = 0;
    } diff;

    /* profiling */
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    struct Span { int cls; hipEvent_t a, b; };
    std::vector<Span> spans;
    double prof_ms[4] = {0, 0, 0, 0};
    int64_t prof_n[4] = {0, 0, 0, 0};

    std::string err;
};

/* internal entry points across translation units */
namespace psm {
int launch_collide_stream(lbdem_psm* h);
int launch_init_equilibrium(lbdem_psm* h, const double* d_rho, const double* d_u);
int launch_pop_convert(lbdem_psm* h, double* d_f64, bool to_device_store);
int launch_pop_convert_buf(lbdem_psm* h, void* buf, double* d_f64, bool to_device_store);
int launch_collide_stream_buf(lbdem_psm* h, const void* fin, void* fout, void* rF);
/* Periodic pressure with the plane sums taken inside the collide: the
 * buffer and box the collide writes to (plane_rho null when the pressure is
 * off), then the fixed-order sums of what it wrote. */
int periodic_pressure_cells(lbdem_psm* h, double** plane_rho, int* dim, int* a0, int* a1, int* b0, int* b1, int* nb,
                            int64_t* ncell);
int periodic_pressure_sum_cells(lbdem_psm* h);
int diff_free(lbdem_psm* h);
int launch_macroscopic(lbdem_psm* h, double* d_rho, double* d_u);
int periodic_pressure_post(lbdem_psm* h);
/* the same on caller-chosen population buffers (the differentiable window) */
int periodic_pressure_post_buf(lbdem_psm* h, void* f);
/* adjoint of the rescale, before the collide+stream adjoint: scales lam =
 * dJ/df(t+1) in place (gplanes = compact planes of f(t+1)) and leaves the
 * plane-mean terms that the collide+stream adjoint adds to dJ/df(t) in
 * pp.buf[4], pp.buf[5] */
int periodic_pressure_extract(lbdem_psm* h, const void* f, void* out);
int64_t periodic_pressure_plane_elems(const lbdem_psm* h);
int periodic_pressure_adjoint_pre(lbdem_psm* h, const void* fin, const void* gplanes, void* lam);
int launch_conservation(lbdem_psm* h, double* mass, double* mom3, double* max_u,
                        int64_t* n_nonfinite);
int run_voxelize(lbdem_psm* h);
int mirror_records(lbdem_psm* h);   /* float view of the live records (FP32/FP16 storage) */
int launch_subcell(lbdem_psm* h, int64_t n, const double* d_d3, const double* d_r, double* d_sf,
                   int32_t* d_cnt, int32_t* d_br);
/* xs: positions for the torque arms; out: 6 per particle (F, T), the
 * handle's own buffer when null */
int run_force_sum(lbdem_psm* h, const double* xs, double* out = nullptr);
/* Groups the stale-record list by owner particle, each particle's cells in
 * increasing cell order: orph_sorted[orph_start[q] .. + orph_cnt[q]) */
int order_orphans(lbdem_psm* h);
int count_records(lbdem_psm* h);
int launch_advance(lbdem_psm* h);
int launch_define_velocity(lbdem_psm* h);
int set_err(lbdem_psm* h, int code, const std::string& msg);
int hip_check(lbdem_psm* h, hipError_t e, const char* what);
void prof_begin(lbdem_psm* h, int cls, hipEvent_t* a);
void prof_end(lbdem_psm* h, int cls, hipEvent_t a);
}  // namespace psm

#define PSM_HIP(h, call)                                              \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return psm::hip_check((h), e_, #call);  \
    } while (0)

#endif

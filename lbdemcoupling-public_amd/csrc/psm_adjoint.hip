/*
 * Adjoint of the fused PSM collide+stream and of the force/torque sum on
 * gfx950: the backward pass of a differentiable window over frozen
 * geometry (the coupling records of the last voxelization stay fixed while
 * the populations evolve).
 *
 * Forward step t (per cell, src/ibCompositeDynamics3D.hh:130-203 plus push
 * streaming):  f(t+1)[x + c_i, i] = C_i(f(t)[x]; omega, B, us)
 * and the per-cell force F(x) that SumForceTorque3D
 * (src/ibProcessors3D.hh:202-275) adds into (F_p, T_p = sum d x F).
 * Backward step t: lambda*(x, i) = lambda(t+1)[x + c_i, i] (the transpose of
 * the push is a pull; at a wall the bounce-back is undone the same way),
 * then lambda(t)[x] = (dC/df)^T lambda* + (dF/df)^T gF with
 * gF = dJ/dF_p + dJ/dT_p x d, and the parameter adjoints are accumulated per
 * record slot (B, us, F x dJ/dT_p) and per workgroup (omega).  The
 * reverse-mode formulas are written out by hand below; nothing is taped.
 *
 * Memory: the window's populations are checkpointed every K steps in the
 * forward pass; the backward pass recomputes one segment (K - 1 states)
 * from its checkpoint and runs the adjoint steps of the segment backwards.
 */
#include "psm_collide_kernel.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace psm {

/* ---- reverse-mode pieces of the cell update ---- */

/* Adjoint accumulators of one set of equilibria
 * eq_i = W_i ((rhoBar + 3 c.j) + 1.5 invRho (3 (c.j)^2 - |j|^2)):
 * with e_i the adjoint of eq_i and w_i = W_i e_i,
 *   d/d rhoBar = sum w_i,  d/d invRho = 1.5 sum w_i (3 (c.j)^2 - |j|^2),
 *   d/d j = 3 sum w_i c_i + 9 invRho sum w_i (c.j) c_i - 3 invRho j sum w_i. */
template <typename T>
struct EqBar {
    T sW, sQ, sC[3], sCJ[3];
    __device__ __forceinline__ void zero() {
        sW = T(0); sQ = T(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) { sC[d] = T(0); sCJ[d] = T(0); }
    }
    __device__ __forceinline__ void add(int i, T e, T cj, T jSqr) {
        const T we = wgt<T>(i) * e;
        sW += we;
        sQ += we * (T(3) * cj * cj - jSqr);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (cc(i, d) == 1) { sC[d] += we; sCJ[d] += we * cj; }
            else if (cc(i, d) == -1) { sC[d] -= we; sCJ[d] -= we * cj; }
        }
    }
    __device__ __forceinline__ void j_bar(T invRho, const T* j, T* out) const {
#pragma unroll
        for (int d = 0; d < 3; ++d) out[d] = T(3) * sC[d] + T(9) * invRho * sCJ[d] - T(3) * invRho * j[d] * sW;
    }
};

/* Scalars of the base dynamics (BGK, Guo forcing when FORCE) of one cell,
 * computed once; fb(i) then gives the post-BGK population i with the
 * operation order of the forward kernel (bit-identical in FP64). */
template <typename T, bool FORCE>
struct BgkCell {
    T rhoBar, rho, invRho, j[3], u[3], jE[3], jSqr, p, omm, pref, ua;
    __device__ __forceinline__ void init(const T* f, T omega, const T* acc) {
        moments<T>(f, rhoBar, j);
        rho = rhoBar + T(1);
        invRho = T(1) / rho;
#pragma unroll
        for (int d = 0; d < 3; ++d) { jE[d] = j[d]; u[d] = T(0); }
        if (FORCE) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                u[d] = j[d] * invRho + T(0.5) * acc[d];
                jE[d] = rho * u[d];
            }
        }
        jSqr = norm_sqr<T>(jE);
        p = T(1.5) * invRho;
        omm = T(1) - omega;
        pref = (T(1) - T(0.5) * omega) * rho;
        ua = (u[0] * acc[0] + u[1] * acc[1]) + u[2] * acc[2];
    }
    __device__ __forceinline__ T eq(int i) const { return equilibrium<T>(i, rhoBar, p, jE, jSqr); }
    __device__ __forceinline__ T guo(int i, const T* acc) const {
        const T ca = cdot<T>(i, acc);
        const T cu = cdot<T>(i, u);
        return T(3) * (ca - ua) + T(9) * (cu * ca);
    }
    __device__ __forceinline__ T fb(int i, T fi, T omega, const T* acc) const {
        T v = fi * omm + omega * eq(i);
        if (FORCE) v = v + (pref * wgt<T>(i)) * guo(i, acc);
        return v;
    }
};

/* The adjoint of one cell's update splits in two stages.
 *
 * Forward (src/ibCompositeDynamics3D.hh:130-203): fb = BGK(fP); a cell
 * without a record streams fb.  A cell with a record (fb = fP, B = 1 when
 * full) streams
 *   out_0 = fP_0 + (1 - B)(fb_0 - fP_0) + B (E_0 - S_0)
 *   out_i = fP_i + (1 - B)(fb_i - fP_i) + B D_i,   i = 1..18,
 *   D_i = (fP_o - E_o) - (fP_i - S_i), o = opposite of i,
 * and produces F = -sum_{i=1..9} c_i B (D_i - D_{i+9}), with
 * E = eq(rho(fb), j(fb)) and S = eq(rho(fb), us (1 + rhoBar(fb))).
 *
 * Stage 1 (record cells only, blend_vjp): given lam = adjoint of out and
 * gF = adjoint of F, the direct part pb of the adjoint of fP, B_bar, us_bar,
 * and the adjoint of fb in the form gb_i = coef lam_i + gr + c_i . gj
 * (coef = 1 - B; a full cell has no BGK, its moment adjoint goes into pb
 * and coef = gr = gj = 0).
 * Stage 2 (every cell, bgk_vjp_store): the BGK adjoint of gb (a cell
 * without a record has gb = lam) plus pb, and omega_bar. */
template <typename T, bool FORCE>
__device__ __forceinline__ void blend_vjp(const T* fP, const T* lam, T* pb, T omega, const T* acc, bool full,
                                          T B, const T* us, const T* gF, T& B_bar, T* us_bar, T* F, T& coef,
                                          T& gr, T* gj) {
    BgkCell<T, FORCE> s;
    s.init(fP, omega, acc);
    T rhoB, jB[3];
    if (full) {
        rhoB = s.rhoBar;
#pragma unroll
        for (int d = 0; d < 3; ++d) jB[d] = s.j[d];
    } else {
        rhoB = T(0);
        jB[0] = jB[1] = jB[2] = T(0);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const T v = s.fb(i, fP[i], omega, acc);
            rhoB = rhoB + v;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (cc(i, d) == 1) jB[d] = jB[d] + v;
                else if (cc(i, d) == -1) jB[d] = jB[d] - v;
            }
        }
    }
    const T invB = T(1) / (rhoB + T(1));
    const T jSqr = norm_sqr<T>(jB);
    const T p = T(1.5) * invB;
    const T onePlus = T(1) + rhoB;
    T uP[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) uP[d] = us[d] * onePlus;
    const T uPSqr = norm_sqr<T>(uP);
    EqBar<T> e1, e2;
    e1.zero();
    e2.zero();
    T Bb = T(0);
    F[0] = F[1] = F[2] = T(0);
    {
        const T fb0 = full ? fP[0] : s.fb(0, fP[0], omega, acc);
        const T D0 = eq_rest<T>(rhoB, p, jSqr) - eq_rest<T>(rhoB, p, uPSqr);
        pb[0] = B * lam[0];
        Bb += lam[0] * (fP[0] - fb0) + lam[0] * D0;
        const T c0 = B * lam[0];
        e1.add(0, c0, T(0), jSqr);
        e2.add(0, -c0, T(0), uPSqr);
    }
#pragma unroll
    for (int i = 1; i <= 9; ++i) {
        const int o = i + 9;
        const T fbi = full ? fP[i] : s.fb(i, fP[i], omega, acc);
        const T fbo = full ? fP[o] : s.fb(o, fP[o], omega, acc);
        T Ei, Eo, Si, So;
        eq_pair<T>(i, rhoB, p, jB, jSqr, Ei, Eo);
        eq_pair<T>(i, rhoB, p, uP, uPSqr, Si, So);
        const T Di = (fP[o] - Eo) - (fP[i] - Si);
        const T Do = (fP[i] - Ei) - (fP[o] - So);
        const T diff = B * Di - B * Do;
        T sg = T(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (cc(i, d) == 1) { F[d] -= diff; sg += gF[d]; }
            else if (cc(i, d) == -1) { F[d] += diff; sg -= gF[d]; }
        }
        /* adjoints of B*D_i and B*D_o (F = -sum c_i (B D_i - B D_o)) */
        const T Dbi = lam[i] - sg, Dbo = lam[o] + sg;
        Bb += (lam[i] * (fP[i] - fbi) + lam[o] * (fP[o] - fbo)) + (Dbi * Di + Dbo * Do);
        const T ci = B * Dbi, co = B * Dbo;
        pb[i] = B * lam[i] + (co - ci);
        pb[o] = B * lam[o] + (ci - co);
        const T cjB = cdot<T>(i, jB), cjS = cdot<T>(i, uP);
        e1.add(i, -co, cjB, jSqr);
        e1.add(o, -ci, -cjB, jSqr);
        e2.add(i, ci, cjS, uPSqr);
        e2.add(o, co, -cjS, uPSqr);
    }
    T rb_bar = e1.sW + e2.sW;
    const T ir_bar = T(1.5) * (e1.sQ + e2.sQ);
    T jB_bar[3], uP_bar[3];
    e1.j_bar(invB, jB, jB_bar);
    e2.j_bar(invB, uP, uP_bar);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        us_bar[d] += onePlus * uP_bar[d];
        rb_bar += us[d] * uP_bar[d];
    }
    rb_bar += -(invB * invB) * ir_bar;
    if (full) {
#pragma unroll
        for (int i = 0; i < Q; ++i) pb[i] += rb_bar + cdot<T>(i, jB_bar);
        coef = T(0);
        gr = T(0);
        gj[0] = gj[1] = gj[2] = T(0);
    } else {
        B_bar += Bb;
        coef = T(1) - B;
        gr = rb_bar;
#pragma unroll
        for (int d = 0; d < 3; ++d) gj[d] = jB_bar[d];
    }
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

/* Fixed-order sum over a block of at most 256 threads (a multiple of 64);
 * the result is valid in thread 0. */
__device__ __forceinline__ double block_sum_256(double v, double* sh) {
    v = wave_sum_d(v);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wave] = v;
    __syncthreads();
    const int nw = (int)(blockDim.x >> 6);
    return (sh[0] + (nw > 1 ? sh[1] : 0.0)) + ((nw > 2 ? sh[2] : 0.0) + (nw > 3 ? sh[3] : 0.0));
}

struct AdjArgs {
    int nx, ny, nz;
    int px, py, pz;
    PopLayout L;
    double omega, ooo;
    double acc[3];
    const uint8_t* sflag;
    const double* rsus;         /* 4 per record slot: sf, us[3] */
    const int32_t* rid;
    int64_t rcap;
    const double* gft;          /* 6 per particle for this step, or null */
    const int32_t* id_map;
    int64_t id_map_size;
    const double* xs;           /* positions of the torque arms */
    /* gacc, pb and pc hold one entry per record at its position in the
     * window's record list (stride pcap), so the record stage writes them
     * contiguously; rpos maps a record cell to that position */
    double* gacc;               /* 7 per record (SoA): gB, gus[3], F x gT[3] */
    double* opart;              /* one omega-gradient partial per workgroup */
    const uint32_t* bin_live;   /* live records per bin */
    void* pb;                   /* 19 per record (SoA, arithmetic type): stage-1 adjoint of fP */
    void* pc;                   /* 5 per record (SoA): coef, gr, gj[3] */
    const int32_t* rpos;        /* per cell: position in the record list */
    const int32_t* rown;        /* per list position: the record's particle index, or -1 */
    int64_t pcap;
    /* periodic pressure: the plane-mean terms of its adjoint, added to every
     * population of the inlet (pp_add[0]) and outlet (pp_add[1]) box cells of
     * the output; none when pp_add is null */
    const double* pp_add;
    int pdim, plast, pa0, pa1, pb0, pb1;
};

/* Adjoint of the pushed populations at (x, y, z): pull from x + c_i, or
 * from the own opposite slot where the push bounced at a wall. */
template <int PREC>
__device__ __forceinline__ void pull_adjoint(const typename Store<PREC>::S* __restrict__ lin, const AdjArgs& a,
                                             int x, int y, int z, typename Store<PREC>::T* lam) {
    const int64_t row = (int64_t)x * a.ny + y;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int xd = x + cx(i), yd = y + cy(i), zd = z + cz(i);
        bool wall = false;
        if (xd < 0 || xd >= a.nx) { if (a.px) xd = xd < 0 ? xd + a.nx : xd - a.nx; else wall = true; }
        if (yd < 0 || yd >= a.ny) { if (a.py) yd = yd < 0 ? yd + a.ny : yd - a.ny; else wall = true; }
        if (zd < 0 || zd >= a.nz) { if (a.pz) zd = zd < 0 ? zd + a.nz : zd - a.nz; else wall = true; }
        if (wall) lam[i] = Store<PREC>::load(lin + pop_at(row, opp(i), z, a.L));
        else lam[i] = Store<PREC>::load(lin + pop_at((int64_t)xd * a.ny + yd, i, zd, a.L));
    }
}

/* pull_adjoint for a wavefront-uniform (x, y): the row-block bases are
 * uniform (scalar registers), only the 32-bit z offsets are per lane. */
template <int PREC, bool PZ>
__device__ __forceinline__ void pull_adjoint_row(const typename Store<PREC>::S* __restrict__ lin, const AdjArgs& a,
                                                 int x, int y, int z, typename Store<PREC>::T* lam) {
    typedef typename Store<PREC>::S S;
    const int64_t row = (int64_t)x * a.ny + y;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int xd = x + cx(i), yd = y + cy(i);
        bool wall = false;
        if (xd < 0 || xd >= a.nx) { if (a.px) xd = xd < 0 ? xd + a.nx : xd - a.nx; else wall = true; }
        if (yd < 0 || yd >= a.ny) { if (a.py) yd = yd < 0 ? yd + a.ny : yd - a.ny; else wall = true; }
        int zd = z + cz(i);
        bool zwall = false;
        if (cz(i) != 0) {
            if (PZ) zd = zd < 0 ? zd + a.nz : (zd >= a.nz ? zd - a.nz : zd);
            else zwall = zd < 0 || zd >= a.nz;
        }
        if (wall || zwall) {
            const char* own = (const char*)(lin + pop_at(row, opp(i), 0, a.L));
            lam[i] = Store<PREC>::load((const S*)(own + a.L.zoff(z) * (uint32_t)sizeof(S)));
        } else {
            const char* src = (const char*)(lin + pop_at((int64_t)xd * a.ny + yd, i, 0, a.L));
            lam[i] = Store<PREC>::load((const S*)(src + a.L.zoff(zd) * (uint32_t)sizeof(S)));
        }
    }
}

/* Stage 1 for the record slot r (= cell index) if it holds a record. */
template <int PREC, bool FORCE>
__device__ __forceinline__ void adjoint_record_cell(const typename Store<PREC>::S* __restrict__ fin,
                                                    const typename Store<PREC>::S* __restrict__ lin,
                                                    const AdjArgs& a, int64_t r, int64_t p) {
    typedef typename Store<PREC>::T T;
    /* everything that depends only on the list position is loaded first,
     * in flight with the gathers: the particle index and the accumulators */
    const int o = a.rown[p];
    const int64_t P = a.pcap;
    double* g = a.gacc;
    double gacc[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) gacc[m] = g[m * P + p];
    int x, y, z;
    rec_cell(r, a.ny, a.nz, x, y, z);
    const int64_t row = (int64_t)x * a.ny + y;
    const T omega = T(a.omega);
    const T acc[3] = {T(a.acc[0]), T(a.acc[1]), T(a.acc[2])};
    T lam[Q], fP[Q], pb[Q];
    pull_adjoint<PREC>(lin, a, x, y, z, lam);
#pragma unroll
    for (int i = 0; i < Q; ++i) fP[i] = Store<PREC>::load(fin + pop_at(row, i, z, a.L));
    const double sf = a.rsus[4 * r];
    const bool full = sf > SOLFRAC_MAX;
    const T B = full ? T(1) : T((sf * a.ooo) / ((1.0 - sf) + a.ooo));
    const T us[3] = {T(a.rsus[4 * r + 1]), T(a.rsus[4 * r + 2]), T(a.rsus[4 * r + 3])};
    T gF[3] = {T(0), T(0), T(0)};
    double gT[3] = {0.0, 0.0, 0.0};
    bool obj = false;
    if (a.gft) {
        if (o >= 0) {
            obj = true;
            const double d[3] = {min_image((double)x - a.xs[3 * o], a.nx),
                                 min_image((double)y - a.xs[3 * o + 1], a.ny),
                                 min_image((double)z - a.xs[3 * o + 2], a.nz)};
            const double* g = a.gft + 6 * (int64_t)o;
            gT[0] = g[3]; gT[1] = g[4]; gT[2] = g[5];
            /* d(gT . (d x F))/dF = gT x d */
            gF[0] = T(g[0] + (gT[1] * d[2] - gT[2] * d[1]));
            gF[1] = T(g[1] + (gT[2] * d[0] - gT[0] * d[2]));
            gF[2] = T(g[2] + (gT[0] * d[1] - gT[1] * d[0]));
        }
    }
    T Bb = T(0), usb[3] = {T(0), T(0), T(0)}, F[3], coef, gr, gj[3];
    blend_vjp<T, FORCE>(fP, lam, pb, omega, acc, full, B, us, gF, Bb, usb, F, coef, gr, gj);
    T* pbo = (T*)a.pb;
    T* pco = (T*)a.pc;
#pragma unroll
    for (int i = 0; i < Q; ++i) pbo[i * P + p] = pb[i];
    pco[p] = coef;
    pco[P + p] = gr;
    pco[2 * P + p] = gj[0];
    pco[3 * P + p] = gj[1];
    pco[4 * P + p] = gj[2];
    if (!full) g[p] = gacc[0] + (double)Bb;
    g[P + p] = gacc[1] + (double)usb[0];
    g[2 * P + p] = gacc[2] + (double)usb[1];
    g[3 * P + p] = gacc[3] + (double)usb[2];
    if (obj) {
        const double Fx = F[0], Fy = F[1], Fz = F[2];
        g[4 * P + p] = gacc[4] + (Fy * gT[2] - Fz * gT[1]);
        g[5 * P + p] = gacc[5] + (Fz * gT[0] - Fx * gT[2]);
        g[6 * P + p] = gacc[6] + (Fx * gT[1] - Fy * gT[0]);
    }
}

/* Stage 1 over the compact list of record slots (one lane per record). */
template <int PREC, bool FORCE>
__global__ __launch_bounds__(256) void k_adjoint_records(const typename Store<PREC>::S* __restrict__ fin,
                                                         const typename Store<PREC>::S* __restrict__ lin,
                                                         AdjArgs a, const int32_t* __restrict__ list,
                                                         int64_t n) {
    /* consecutive list positions are neighbouring cells: keep a range of
     * them on one XCD's L2, where their gathers share lines */
    const int64_t k = (int64_t)xcd_block((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    adjoint_record_cell<PREC, FORCE>(fin, lin, a, (int64_t)list[k], k);
}

/* Record-list groups: zb consecutive z-bins of one (bx, by) bin column, so
 * that a group spans 8 zb cells along z.  Group g -> (bx, by, gz). */
struct RecGroup {
    int bx, by, z0, zlen;   /* cells z0 .. z0 + zlen - 1 (8 zb; those beyond the lattice are skipped) */
    bool live;
};

__device__ __forceinline__ RecGroup rec_group(int64_t g, const uint32_t* __restrict__ bin_live, int nb1, int nb2,
                                             int zb) {
    RecGroup r;
    const int ngz = (nb2 + zb - 1) / zb;
    const int gz = (int)(g % ngz);
    const int64_t gxy = g / ngz;
    r.by = (int)(gxy % nb1);
    r.bx = (int)(gxy / nb1);
    r.z0 = gz * zb * BIN;
    r.zlen = zb * BIN;
    r.live = false;
    const int64_t b0 = ((int64_t)r.bx * nb1 + r.by) * nb2 + (int64_t)gz * zb;
    for (int k = 0; k < zb && gz * zb + k < nb2; ++k) r.live = r.live || bin_live[b0 + k] != 0u;
    return r;
}

/* Live record cells of each group (one workgroup per group). */
__global__ __launch_bounds__(256) void k_count_records(const uint8_t* __restrict__ sflag,
                                                       const uint32_t* __restrict__ bin_live, int n0, int n1,
                                                       int n2, int nb1, int nb2, int zb, int64_t* __restrict__ cnt) {
    const RecGroup gr = rec_group(blockIdx.x, bin_live, nb1, nb2, zb);
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    if (gr.live) {
        const int ncell = BIN * BIN * gr.zlen;
        for (int l = threadIdx.x; l < ncell; l += blockDim.x) {
            const int X = gr.bx * BIN + l / (BIN * gr.zlen), Y = gr.by * BIN + (l / gr.zlen) % BIN,
                      Z = gr.z0 + l % gr.zlen;
            const bool live = X < n0 && Y < n1 && Z < n2 && sflag[((int64_t)X * n1 + Y) * n2 + Z] != 0;
            const unsigned long long m = __ballot(live);
            if (__lane_id() == 0 && m != 0ull) atomicAdd(&s_n, (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = (int64_t)s_n;
}

/* The record slots of the frozen geometry by group (groups z fastest), and
 * within a group in cell order (x slice, then y, then z over the group's
 * 8 zb cells), from the group's first position off[g].  Consecutive list
 * positions are then neighbouring cells along z, so the record stage's
 * gathers of neighbouring lanes share the lines of the population stretches
 * (32 cells per 128-byte line in FP32) instead of using 8 cells of a line
 * per bin. */
__global__ __launch_bounds__(256) void k_list_records(const uint8_t* __restrict__ sflag,
                                                      const uint32_t* __restrict__ bin_live, int n0, int n1,
                                                      int n2, int nb1, int nb2, int zb,
                                                      const int64_t* __restrict__ off,
                                                      const int32_t* __restrict__ rid,
                                                      const int32_t* __restrict__ id_map, int64_t id_map_size,
                                                      int32_t* __restrict__ list, int32_t* __restrict__ rpos,
                                                      int32_t* __restrict__ rown) {
    const RecGroup gr = rec_group(blockIdx.x, bin_live, nb1, nb2, zb);
    if (!gr.live) return;
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    __shared__ uint32_t s_wn[16];
    int64_t base = off[blockIdx.x];
    const int ncell = BIN * BIN * gr.zlen;
    for (int l0 = 0; l0 < ncell; l0 += blockDim.x) {
        const int l = l0 + (int)threadIdx.x;
        const int X = gr.bx * BIN + l / (BIN * gr.zlen), Y = gr.by * BIN + (l / gr.zlen) % BIN,
                  Z = gr.z0 + l % gr.zlen;
        const bool live = l < ncell && X < n0 && Y < n1 && Z < n2 &&
                          sflag[((int64_t)X * n1 + Y) * n2 + Z] != 0;
        const unsigned long long m = __ballot(live);
        if (lane == 0) s_wn[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        int64_t wb = base;
        for (int w = 0; w < wave; ++w) wb += s_wn[w];
        int64_t tot = 0;
        for (int w = 0; w < nw; ++w) tot += s_wn[w];
        if (live) {
            const int64_t pos = wb + __popcll(m & ((1ull << lane) - 1ull));
            const int64_t c = rec_index(X, Y, Z, n1, n2);
            list[pos] = (int32_t)c;
            rpos[c] = (int32_t)pos;
            const int id = rid[c];
            rown[pos] = (id >= 1 && id < id_map_size) ? id_map[id] : -1;
        }
        base += tot;
        __syncthreads();
    }
}

/* Stage 2, every cell: lout = adjoint of f(t) from lin = adjoint of f(t+1).
 * One workgroup per (row, z chunk) block of the store, the chunk index
 * fastest, as the forward kernel.
 * The adjoint of fb is gb_i = coef lam_i + gr + c_i . gj (lam itself for a
 * cell without a record); the BGK adjoint is
 *   fPbar_i = pb_i + (1 - omega) gb_i + rb + c_i . jb,
 * with (rb, jb) the adjoint of the pre-collision moments, two passes over
 * gb (sums, then stores). */
template <int PREC, bool FORCE, bool PZ>
__global__ __launch_bounds__(256) void k_adjoint(const typename Store<PREC>::S* __restrict__ fin,
                                                 const typename Store<PREC>::S* __restrict__ lin,
                                                 typename Store<PREC>::S* __restrict__ lout, AdjArgs a) {
    typedef typename Store<PREC>::T T;
    __shared__ double sh[4];
    const int nzc = a.L.nzc;
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int row = blk / nzc;
    const int chunk = blk - row * nzc;
    const int z = chunk * a.L.zc + threadIdx.x;
    const int x = __builtin_amdgcn_readfirstlane(row / a.ny);
    const int y = __builtin_amdgcn_readfirstlane(row - x * a.ny);
    double wbar = 0.0;
    if (z < a.nz) {
        const T omega = T(a.omega);
        const T acc[3] = {T(a.acc[0]), T(a.acc[1]), T(a.acc[2])};
        typedef typename Store<PREC>::S S;
        T lam[Q], fP[Q];
        pull_adjoint_row<PREC, PZ>(lin, a, x, y, z, lam);
        const char* frow = (const char*)(fin + pop_at(row, 0, chunk * a.L.zc, a.L));
        const uint32_t qstride = (uint32_t)a.L.zc * (uint32_t)sizeof(S);
        const uint32_t zoff = (uint32_t)threadIdx.x * (uint32_t)sizeof(S);
#pragma unroll
        for (int i = 0; i < Q; ++i) fP[i] = Store<PREC>::load((const S*)(frow + i * qstride + zoff));
        /* a record cell of the window's list (its position r) */
        int64_t r = -1;
        if (a.sflag[(uint32_t)row * (uint32_t)a.nz + (uint32_t)z] != 0) r = a.rpos[rec_index(x, y, z, a.ny, a.nz)];
        const bool rec = r >= 0;
        T coef = T(1), gr = T(0), gj[3] = {T(0), T(0), T(0)};
        if (rec) {
            const T* pc = (const T*)a.pc;
            coef = pc[r];
            gr = pc[a.pcap + r];
            gj[0] = pc[2 * a.pcap + r];
            gj[1] = pc[3 * a.pcap + r];
            gj[2] = pc[4 * a.pcap + r];
        }
        BgkCell<T, FORCE> s;
        s.init(fP, omega, acc);
        EqBar<T> eb;
        eb.zero();
        T om = T(0), sgG = T(0), sgw = T(0), sga[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const T gb = coef * lam[i] + gr + cdot<T>(i, gj);
            lam[i] = gb;
            om += gb * (s.eq(i) - fP[i]);
            eb.add(i, omega * gb, cdot<T>(i, s.jE), s.jSqr);
            if (FORCE) {
                const T wg = wgt<T>(i) * gb;
                const T ca = cdot<T>(i, acc);
                sgG += wg * s.guo(i, acc);
                sgw += wg;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (cc(i, d) == 1) sga[d] += wg * ca;
                    else if (cc(i, d) == -1) sga[d] -= wg * ca;
                }
            }
        }
        T rb2 = eb.sW;
        T ir2 = T(1.5) * eb.sQ;
        T jE_bar[3], j2[3];
        eb.j_bar(s.invRho, s.jE, jE_bar);
        if (FORCE) {
            /* fb_i += pref W_i G_i: pref = (1 - omega/2) rho, G_i depends on
             * u; jE = rho u, u = j invRho + a/2 */
            om += T(-0.5) * s.rho * sgG;
            rb2 += (T(1) - T(0.5) * omega) * sgG;
            T u_bar[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                u_bar[d] = s.pref * (T(-3) * acc[d] * sgw + T(9) * sga[d]);
                rb2 += jE_bar[d] * s.u[d];
                u_bar[d] += s.rho * jE_bar[d];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                j2[d] = s.invRho * u_bar[d];
                ir2 += u_bar[d] * s.j[d];
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) j2[d] = jE_bar[d];
        }
        rb2 += -(s.invRho * s.invRho) * ir2;
#pragma unroll
        for (int i = 0; i < Q; ++i) lam[i] = s.omm * lam[i] + (rb2 + cdot<T>(i, j2));
        if (rec) {
            /* the record cell's direct part, all loads in flight before the adds */
            const T* pb = (const T*)a.pb;
#pragma unroll
            for (int i = 0; i < Q; ++i) fP[i] = pb[i * a.pcap + r];
#pragma unroll
            for (int i = 0; i < Q; ++i) lam[i] = fP[i] + lam[i];
        }
        char* orow = (char*)(lout + pop_at(row, 0, chunk * a.L.zc, a.L));
        bool pp_cell = false;
        double ppc = 0.0;
        if (a.pp_add != nullptr) {
            /* after the collide+stream adjoint: each population of an inlet
             * (outlet) box cell gains the outlet's (inlet's) plane-mean term,
             * added in double to the stored value and rounded once; the plane
             * mapping as selects */
            const int ci = a.pdim == 0 ? x : (a.pdim == 1 ? y : z);
            const int pl = ci == 0 ? 0 : (ci == a.plast ? 1 : -1);
            const int ca = a.pdim == 0 ? y : x;
            const int cb = a.pdim == 2 ? y : z;
            pp_cell = pl >= 0 && ca >= a.pa0 && ca <= a.pa1 && cb >= a.pb0 && cb <= a.pb1;
            if (pp_cell) ppc = a.pp_add[pl];
        }
        if (pp_cell) {
#pragma unroll
            for (int i = 0; i < Q; ++i)
                Store<PREC>::store((S*)(orow + i * qstride + zoff), (T)((double)(S)lam[i] + ppc));
        } else {
#pragma unroll
            for (int i = 0; i < Q; ++i) Store<PREC>::store((S*)(orow + i * qstride + zoff), lam[i]);
        }
        wbar = (double)om;
    }
    const double s = block_sum_256(wbar, sh);
    if (threadIdx.x == 0) a.opart[blockIdx.x] += s;
}

/* ---- reductions of the accumulated adjoints ---- */

__global__ __launch_bounds__(256) void k_sum_vec(const double* __restrict__ in, int64_t n,
                                                 double* __restrict__ out) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) v += in[k];
    const double s = block_sum_256(v, sh);
    if (threadIdx.x == 0) out[0] = s;
}

/* Per record slot: dense dJ/dB and dJ/dsf (partial cells; B = sf o / ((1 -
 * sf) + o), o = tau - 1/2) and the tau partial sum_r dJ/dB dB/do. */
__global__ __launch_bounds__(256) void k_grad_cells(
                                                    int64_t rcap, const uint8_t* __restrict__ sflag,
                                                    const double* __restrict__ rsus,
                                                    const double* __restrict__ gacc,
                                                    const int32_t* __restrict__ rpos, double ooo,
                                                    double* __restrict__ cell_out, int64_t nc,
                                                    double* __restrict__ tau_part) {
    __shared__ double sh[4];
    double tp = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rcap;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = r;   /* record slot = cell index */
        double gB = 0.0, gsf = 0.0;
        if (sflag[c] && rpos[c] >= 0) {
            const double sf = rsus[4 * r];
            if (!(sf > SOLFRAC_MAX)) {
                gB = gacc[rpos[c]];
                const double den = (1.0 - sf) + ooo;
                gsf = gB * (ooo * (1.0 + ooo)) / (den * den);
                tp += gB * (sf * (1.0 - sf)) / (den * den);
            }
        }
        cell_out[c] = gsf;
        cell_out[nc + c] = gB;
    }
    const double s = block_sum_256(tp, sh);
    if (threadIdx.x == 0) tau_part[blockIdx.x] = s;
}

struct GradArgs {
    int n[3];
    int nb[3];
    const uint8_t* sflag;
    const double* rsus;         /* 4 per record slot: sf, us[3] */
    const int32_t* rcand;
    const int32_t* rid;
    int64_t rcap;
    const double* gacc;         /* at the record's list position (stride pcap) */
    const int32_t* rpos;
    int64_t pcap;
    const double* xs;
    const double* pw;
    const double* pr;
    const int32_t* powner;
    const int32_t* p_ent_start;
    const int32_t* p_nent;
    const int32_t* owner_ent;
    const int32_t* cand_bin;
    int64_t ent_cap;
    double ooo;
    double rc, vc;
    int overlap;
    double* gpart;              /* 9 per particle: g_v[3], g_w[3], g_x[3] */
};

/* Contribution of record r of the cell (X, Y, Z) owned by particle o. */
__device__ __forceinline__ void record_grad(const GradArgs& a, int64_t r, int X, int Y, int Z, int o,
                                            double* acc) {
    const double d[3] = {min_image((double)X - a.xs[3 * o], a.n[0]),
                         min_image((double)Y - a.xs[3 * o + 1], a.n[1]),
                         min_image((double)Z - a.xs[3 * o + 2], a.n[2])};
    const int64_t p = a.rpos[r], P = a.pcap;
    if (p < 0) return;   /* not in the window's record list */
    const double g[3] = {a.gacc[P + p], a.gacc[2 * P + p], a.gacc[3 * P + p]};
    const double fxt[3] = {a.gacc[4 * P + p], a.gacc[5 * P + p], a.gacc[6 * P + p]};
    const double w[3] = {a.pw[3 * o], a.pw[3 * o + 1], a.pw[3 * o + 2]};
    /* us = v + w x d, d = cell - x */
    acc[0] += g[0];
    acc[1] += g[1];
    acc[2] += g[2];
    acc[3] += d[1] * g[2] - d[2] * g[1];
    acc[4] += d[2] * g[0] - d[0] * g[2];
    acc[5] += d[0] * g[1] - d[1] * g[0];
    /* d/dx: w x g (lever arm) - F x gT (torque arm) */
    double gx[3] = {(w[1] * g[2] - w[2] * g[1]) - fxt[0], (w[2] * g[0] - w[0] * g[2]) - fxt[1],
                    (w[0] * g[1] - w[1] * g[0]) - fxt[2]};
    if (a.overlap) {
        const double sf = a.rsus[4 * r];
        if (!(sf > SOLFRAC_MAX)) {
            const double den = (1.0 - sf) + a.ooo;
            const double gsf = a.gacc[p] * (a.ooo * (1.0 + a.ooo)) / (den * den);
            const double dist = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            double der = 0.0;
            int br;
            (void)overlap_fraction(dist, a.pr[o], a.rc, a.vc, &der, &br);
            if (dist > 0.0 && der != 0.0) {
                /* d dist / d x = -d / dist */
                const double k = -gsf * der / dist;
                gx[0] += k * d[0];
                gx[1] += k * d[1];
                gx[2] += k * d[2];
            }
        }
    }
    acc[6] += gx[0];
    acc[7] += gx[1];
    acc[8] += gx[2];
}

/* One workgroup per particle: the records its candidate entries set, in
 * entry order, then a fixed-order block sum. */
__global__ __launch_bounds__(256) void k_grad_particle(GradArgs a) {
    __shared__ double sh[4];
    const int64_t q = blockIdx.x;
    const int o = a.powner[q];
    const int64_t s = a.p_ent_start[q];
    const int ne = a.p_nent[q];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < ne; ++k) {
        if (s + k >= a.ent_cap) break;
        const int e = a.owner_ent[s + k];
        if (e < 0) continue;
        const int64_t b = a.cand_bin[e];
        const int bz = (int)(b % a.nb[2]);
        const int by = (int)((b / a.nb[2]) % a.nb[1]);
        const int bx = (int)(b / ((int64_t)a.nb[2] * a.nb[1]));
        for (int l = threadIdx.x; l < BIN_CELLS; l += blockDim.x) {
            const int X = bx * BIN + (l >> 6), Y = by * BIN + ((l >> 3) & 7), Z = bz * BIN + (l & 7);
            if (X >= a.n[0] || Y >= a.n[1] || Z >= a.n[2]) continue;
            const int64_t r = rec_index(X, Y, Z, a.n[1], a.n[2]);
            if (!a.sflag[r]) continue;
            if (a.rcand[r] != e) continue;
            record_grad(a, r, X, Y, Z, o, acc);
        }
    }
    for (int m = 0; m < 9; ++m) {
        const double v = block_sum_256(acc[m], sh);
        if (threadIdx.x == 0) a.gpart[9 * q + m] = v;
    }
}

/* Records no candidate set in the last voxelization (kept from earlier
 * steps), grouped by owner in cell order (order_orphans): one wavefront per
 * particle adds its records' contributions lane-strided in that order, then
 * a fixed-order tree, once onto the particle's sum of its entries.  Clears
 * the particle's count and fill cursor for the next grouping. */
__global__ __launch_bounds__(256) void k_grad_orphans(GradArgs a, int64_t np, int32_t* __restrict__ cnt,
                                                      int32_t* __restrict__ fill,
                                                      const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ sorted) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= np) return;
    const int m = cnt[q];
    if (m == 0) return;
    const int lane = (int)__lane_id();
    const int32_t* seg = sorted + start[q];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = lane; i < m; i += 64) {
        const int64_t r = seg[i];
        if (!a.sflag[r]) continue;
        int X, Y, Z;
        rec_cell(r, a.n[1], a.n[2], X, Y, Z);
        record_grad(a, r, X, Y, Z, (int)q, acc);
    }
    double first8[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) first8[k] = acc[k];
    const double s8 = wave_sum8_scatter(first8, lane);
    double s9 = acc[8];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s9 += __shfl_xor(s9, off, 64);
    const int mm = lane >> 3;
    if ((lane & 7) == 0) a.gpart[9 * q + mm] += s8;
    if (lane == 0) {
        a.gpart[9 * q + 8] += s9;
        cnt[q] = 0;
        fill[q] = 0;
    }
}

/* sf += dsf on live records; flags counters[1] bit 16 when a result leaves
 * (SOLFRAC_MIN, 1] or dsf is non-zero on a cell without a record. */
__global__ __launch_bounds__(256) void k_perturb_sf(int n0, int n1, int n2,
                                                    const uint8_t* __restrict__ sflag,
                                                    double* __restrict__ rsus,
                                                    const double* __restrict__ dsf,
                                                    int64_t* __restrict__ counters) {
    const int64_t nc = (int64_t)n0 * n1 * n2;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        const double d = dsf[c];
        if (d == 0.0) continue;
        if (!sflag[c]) {
            atomicOr((unsigned long long*)&counters[1], 16ull);
            continue;
        }
        const double v = rsus[4 * c] + d;   /* record slot = cell index */
        if (!(v > SOLFRAC_MIN && v <= 1.0)) atomicOr((unsigned long long*)&counters[1], 16ull);
        else rsus[4 * c] = v;
    }
}

/* ---- host side ---- */

template <typename P>
static hipError_t dalloc_d(P** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    return hipMalloc((void**)p, bytes);
}

static void dfree_d(void* p) {
    if (p) (void)hipFree(p);
}

static void free_states(std::vector<void*>& v) {
    for (void* p : v) dfree_d(p);
    v.clear();
}

static size_t state_elem_bytes(const lbdem_psm* h) {
    return h->precision == LBDEM_PSM_FP64 ? sizeof(double)
                                          : (h->precision == LBDEM_PSM_FP32_DEV ? sizeof(float) : sizeof(_Float16));
}

int diff_free(lbdem_psm* h) {
    auto& D = h->diff;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_states(D.ckpt);
    free_states(D.seg);
    dfree_d(D.lam[0]); dfree_d(D.lam[1]);
    D.lam[0] = D.lam[1] = nullptr;
    dfree_d(D.d_ft); dfree_d(D.gacc); dfree_d(D.Fscr); dfree_d(D.opart); dfree_d(D.gpart);
    dfree_d(D.cell_out); dfree_d(D.tau_part); dfree_d(D.pb); dfree_d(D.pc);
    dfree_d(D.rec_list); dfree_d(D.bin_off); dfree_d(D.rpos); dfree_d(D.rown);
    D.rpos = nullptr;
    D.pcap = 0;
    dfree_d(D.pp_last); dfree_d(D.pp_next); dfree_d(D.pp_cur);
    D.pp_last = D.pp_next = D.pp_cur = nullptr;
    D.pb = D.pc = nullptr;
    D.rec_list = nullptr;
    D.bin_off = nullptr;
    D.rown = nullptr;
    D.nrec = 0;
    D.d_ft = D.gacc = D.Fscr = D.opart = D.gpart = D.cell_out = D.tau_part = nullptr;
    D.opart_n = D.tau_part_n = 0;
    D.ready = false;
    return LBDEM_PSM_OK;
}

static int ensure_states(lbdem_psm* h, std::vector<void*>& v, size_t n) {
    while (v.size() > n) { dfree_d(v.back()); v.pop_back(); }
    while (v.size() < n) {
        void* p = nullptr;
        hipError_t e = dalloc_d(&p, h->pop_bytes);
        if (e != hipSuccess) return hip_check(h, e, "hipMalloc (window states)");
        v.push_back(p);
    }
    return LBDEM_PSM_OK;
}

static int adjoint_step(lbdem_psm* h, const void* fin, const void* lin, void* lout, const double* gft) {
    AdjArgs a;
    a.nx = h->g.n[0]; a.ny = h->g.n[1]; a.nz = h->g.n[2];
    a.px = h->g.periodic[0]; a.py = h->g.periodic[1]; a.pz = h->g.periodic[2];
    a.L = h->g.L;
    a.omega = h->omega; a.ooo = h->ooo;
    for (int d = 0; d < 3; ++d) a.acc[d] = h->desc.body_accel[d];
    a.sflag = h->sflag;
    a.rsus = h->rec.sus; a.rid = h->rec.id; a.rcap = h->rec_cap;
    a.gft = gft;
    a.id_map = h->id_map; a.id_map_size = h->id_map_size;
    a.xs = h->psum_x;
    a.gacc = h->diff.gacc;
    a.opart = h->diff.opart;
    a.bin_live = h->bin_any[h->cur_any];
    a.pb = h->diff.pb;
    a.pc = h->diff.pc;
    a.rpos = h->diff.rpos;
    a.rown = h->diff.rown;
    a.pcap = h->diff.pcap;
    a.pp_add = nullptr;
    a.pdim = a.plast = a.pa0 = a.pa1 = a.pb0 = a.pb1 = 0;
    if (h->pp.enabled) {
        /* the terms periodic_pressure_adjoint_pre left in its buffer */
        a.pp_add = h->pp.buf + 4;
        a.pdim = h->pp.dim;
        a.plast = h->g.n[a.pdim] - 1;
        a.pa0 = h->pp.range[0]; a.pa1 = h->pp.range[1];
        a.pb0 = h->pp.range[2]; a.pb1 = h->pp.range[3];
    }
    const int64_t nrec = h->diff.nrec;
    dim3 gridr((unsigned)std::max<int64_t>(1, (nrec + 255) / 256)), blockr(256);
    const int32_t* list = h->diff.rec_list;
    dim3 grid((unsigned)((int64_t)a.nx * a.ny * a.L.nzc)), block((unsigned)a.L.zc);
    if (h->precision == LBDEM_PSM_FP64) {
        typedef double S;
        if (h->has_force) {
            hipLaunchKernelGGL((k_adjoint_records<LBDEM_PSM_FP64, true>), gridr, blockr, 0, h->stream, (const S*)fin, (const S*)lin, a, list, nrec);
            if (a.pz) hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP64, true, true>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
            else hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP64, true, false>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
        } else {
            hipLaunchKernelGGL((k_adjoint_records<LBDEM_PSM_FP64, false>), gridr, blockr, 0, h->stream, (const S*)fin, (const S*)lin, a, list, nrec);
            if (a.pz) hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP64, false, true>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
            else hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP64, false, false>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
        }
    } else {
        typedef float S;
        if (h->has_force) {
            hipLaunchKernelGGL((k_adjoint_records<LBDEM_PSM_FP32_DEV, true>), gridr, blockr, 0, h->stream, (const S*)fin, (const S*)lin, a, list, nrec);
            if (a.pz) hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP32_DEV, true, true>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
            else hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP32_DEV, true, false>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
        } else {
            hipLaunchKernelGGL((k_adjoint_records<LBDEM_PSM_FP32_DEV, false>), gridr, blockr, 0, h->stream, (const S*)fin, (const S*)lin, a, list, nrec);
            if (a.pz) hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP32_DEV, false, true>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
            else hipLaunchKernelGGL((k_adjoint<LBDEM_PSM_FP32_DEV, false, false>), grid, block, 0, h->stream, (const S*)fin, (const S*)lin, (S*)lout, a);
        }
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

}  // namespace psm

using namespace psm;

extern "C" {

int lbdem_psm_set_tau(lbdem_psm* h, double tau) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    if (!(tau > 0.5) || !std::isfinite(tau))
        return set_err(h, LBDEM_PSM_ERR_NUMERIC, "tau must be finite and > 0.5 (omega < 2)");
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    h->desc.tau = tau;
    h->omega = 1.0 / tau;
    h->ooo = 1.0 / h->omega - 0.5;
    h->geom_gen += 1;
    return LBDEM_PSM_OK;
}

int lbdem_psm_clear_records(lbdem_psm* h) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    PSM_HIP(h, hipMemsetAsync(h->sflag, 0, h->g.nc * sizeof(uint8_t), h->stream));
    for (int k = 0; k < 2; ++k) PSM_HIP(h, hipMemsetAsync(h->bin_any[k], 0, h->nbins * sizeof(uint32_t), h->stream));
    PSM_HIP(h, hipMemsetAsync(h->bin_cand_count, 0, h->nbins * sizeof(int32_t), h->stream));
    if (h->p_cap > 0) PSM_HIP(h, hipMemsetAsync(h->p_nent, 0, h->p_cap * sizeof(int32_t), h->stream));
    PSM_HIP(h, hipMemsetAsync(h->counters, 0, 8 * sizeof(int64_t), h->stream));
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    h->geom_gen += 1;
    return LBDEM_PSM_OK;
}

int lbdem_psm_perturb_solid_fraction(lbdem_psm* h, const double* dsf) {
    if (!h || !dsf) return LBDEM_PSM_ERR_INVALID_ARG;
    const int64_t nc = h->g.nc;
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    double* d = nullptr;
    PSM_HIP(h, dalloc_d(&d, nc * sizeof(double)));
    hipError_t e = hipMemcpy(d, dsf, nc * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(h->counters + 1, 0, sizeof(int64_t), h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_perturb_sf, dim3((unsigned)std::min<int64_t>(8192, (nc + 255) / 256)), dim3(256), 0,
                           h->stream, h->g.n[0], h->g.n[1], h->g.n[2], h->sflag, h->rec.sus,
                           d, h->counters);
        e = hipGetLastError();
    }
    int64_t flags = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&flags, h->counters + 1, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree_d(d);
    PSM_HIP(h, e);
    h->geom_gen += 1;
    {
        const int mr = mirror_records(h);   /* the valid entries were applied */
        if (mr) return mr;
    }
    if (flags & 16)
        return set_err(h, LBDEM_PSM_ERR_INVALID_ARG,
                       "solid-fraction perturbation leaves (0.001, 1] or touches a cell without a record "
                       "(the valid entries were applied)");
    return LBDEM_PSM_OK;
}

/* Segment start steps of a window of T steps.  K > 0: a checkpoint every K
 * steps.  K < 0: a budget of M = -K population states for checkpoints and
 * recomputed states together, with segments of M, M-1, M-2, ... steps (the
 * last one cut to what is left).  The backward pass handles segment j
 * (0-based, from the last) while the checkpoints of segments 0..j are alive
 * and recomputes len_j - 1 = M - j - 1 states, so at most (j + 1) + (M - j - 1)
 * = M states are held: the checkpoint of each finished segment becomes
 * recompute storage.  Every step is recomputed once, as with K > 0.  Needs
 * M (M + 1) / 2 >= T. */
static bool window_segments(int T, int K, std::vector<int32_t>& t0) {
    t0.clear();
    if (K > 0) {
        for (int t = 0; t < T; t += K) t0.push_back(t);
        return true;
    }
    const int64_t M = -(int64_t)K;
    if (M < 1 || M * (M + 1) / 2 < T) return false;
    int t = 0;
    for (int len = (int)M; t < T; --len) {
        t0.push_back(t);
        t += len;
    }
    return true;
}

int lbdem_psm_diff_forward(lbdem_psm* h, int32_t nsteps, int32_t K, double* ft) {
    if (!h || nsteps < 1 || K == 0) return LBDEM_PSM_ERR_INVALID_ARG;
    if (h->precision == LBDEM_PSM_FP16_STORE)
        return set_err(h, LBDEM_PSM_ERR_STATE, "the differentiable window supports FP64 and FP32_DEV storage");
    if (!h->have_particles) return set_err(h, LBDEM_PSM_ERR_STATE, "diff_forward before set_spheres");
    auto& D = h->diff;
    D.ready = false;
    std::vector<int32_t> t0;
    if (!window_segments(nsteps, K, t0))
        return set_err(h, LBDEM_PSM_ERR_INVALID_ARG,
                       "a state budget M = -checkpoint_every needs M (M + 1) / 2 >= nsteps");
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    const int64_t np = h->np;
    const int ncp = (int)t0.size();
    /* checkpoints consumed by a budgeted backward pass were moved to the
     * recompute pool; drop the empty entries */
    D.ckpt.erase(std::remove(D.ckpt.begin(), D.ckpt.end(), nullptr), D.ckpt.end());
    int rc = LBDEM_PSM_OK;
    if (K < 0) {
        /* budgeted: all state buffers (checkpoints and recompute pool) are
         * one pool of at most M; the first ncp become checkpoints, the rest
         * stay in the recompute pool, so a repeated window allocates nothing */
        const size_t M = (size_t)(-(int64_t)K);
        std::vector<void*> pool(D.ckpt);
        pool.insert(pool.end(), D.seg.begin(), D.seg.end());
        D.ckpt.clear();
        D.seg.clear();
        while (pool.size() > M) { dfree_d(pool.back()); pool.pop_back(); }
        for (size_t k = 0; k < pool.size(); ++k) ((k < (size_t)ncp) ? D.ckpt : D.seg).push_back(pool[k]);
    }
    rc = ensure_states(h, D.ckpt, (size_t)ncp);
    if (rc) return rc;
    D.seg_t0 = t0;
    dfree_d(D.d_ft);
    D.d_ft = nullptr;
    PSM_HIP(h, dalloc_d(&D.d_ft, (size_t)nsteps * (size_t)std::max<int64_t>(np, 1) * 6 * sizeof(double)));
    if (np > 0) PSM_HIP(h, hipMemcpyAsync(h->psum_x, h->px, 3 * np * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    PSM_HIP(h, hipMemsetAsync(h->counters + 1, 0, sizeof(int64_t), h->stream));
    /* record slots of the frozen geometry, for the adjoint's first stage */
    if (!D.rec_list) PSM_HIP(h, dalloc_d(&D.rec_list, h->rec_cap * sizeof(int32_t)));
    if (!D.rpos) PSM_HIP(h, dalloc_d(&D.rpos, h->g.nc * sizeof(int32_t)));
    if (!D.bin_off) PSM_HIP(h, dalloc_d(&D.bin_off, h->nbins * sizeof(int64_t)));
    if (!D.rown) PSM_HIP(h, dalloc_d(&D.rown, h->rec_cap * sizeof(int32_t)));
    PSM_HIP(h, hipMemsetAsync(D.rpos, 0xff, h->g.nc * sizeof(int32_t), h->stream));   /* -1: no record */
    {
        /* groups of 8 z-bins, 64 cells along z: the backward pass of a 20-step
         * window at 512x512x768 (FP32) took 0.521 s with single bins, 0.497 /
         * 0.494 / 0.491 / 0.488 s with groups of 2 / 4 / 8 / 16 */
        const int zb = 8;
        const int64_t ngroups = (int64_t)h->g.nb[0] * h->g.nb[1] * ((h->g.nb[2] + zb - 1) / zb);
        hipLaunchKernelGGL(k_count_records, dim3((unsigned)ngroups), dim3(256), 0, h->stream, h->sflag,
                           h->bin_any[h->cur_any], h->g.n[0], h->g.n[1], h->g.n[2], h->g.nb[1], h->g.nb[2], zb,
                           D.bin_off);
        PSM_HIP(h, hipGetLastError());
        std::vector<int64_t> off((size_t)ngroups);
        PSM_HIP(h, hipMemcpyAsync(off.data(), D.bin_off, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                                  h->stream));
        PSM_HIP(h, hipStreamSynchronize(h->stream));
        int64_t tot = 0;
        for (auto& o : off) { const int64_t c = o; o = tot; tot += c; }
        if (tot > h->rec_cap) return LBDEM_PSM_ERR_CAPACITY;
        D.nrec = tot;
        PSM_HIP(h, hipMemcpyAsync(D.bin_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                                  h->stream));
        hipLaunchKernelGGL(k_list_records, dim3((unsigned)ngroups), dim3(256), 0, h->stream, h->sflag,
                           h->bin_any[h->cur_any], h->g.n[0], h->g.n[1], h->g.n[2], h->g.nb[1], h->g.nb[2], zb,
                           (const int64_t*)D.bin_off, h->rec.id, h->id_map, h->id_map_size, D.rec_list,
                           D.rpos, D.rown);
        PSM_HIP(h, hipGetLastError());
        /* the host copy of the offsets is released only after the upload */
        PSM_HIP(h, hipStreamSynchronize(h->stream));
    }
    for (int t = 0, nextc = 0; t < nsteps; ++t) {
        const bool segment_start = nextc < ncp && t == t0[nextc];
        rc = launch_collide_stream(h);
        if (rc) return rc;
        if (np > 0) {
            /* straight into the window's force history */
            rc = run_force_sum(h, h->psum_x, D.d_ft + (size_t)t * np * 6);
            if (rc) return rc;
        }
        if (segment_start) {
            /* the step only read its input f(t), which is this segment's
             * checkpoint: that buffer becomes the checkpoint, and the
             * checkpoint's old buffer the next step's output (no copy of
             * the state; all of them are pop_bytes device buffers) */
            std::swap(D.ckpt[nextc], h->f[1 - h->cur_f]);
            ++nextc;
        }
    }
    /* the handle's force buffer holds the last step's sums, as after a
     * coupled step */
    if (np > 0 && nsteps > 0)
        PSM_HIP(h, hipMemcpyAsync(h->pft, D.d_ft + (size_t)(nsteps - 1) * np * 6, np * 6 * sizeof(double),
                                  hipMemcpyDeviceToDevice, h->stream));
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    if (ft && np > 0 && h->owner_identity) {
        /* every particle is its own map(tag): the history is the fold */
        PSM_HIP(h, hipMemcpy(ft, D.d_ft, (size_t)nsteps * np * 6 * sizeof(double), hipMemcpyDeviceToHost));
    } else if (ft && np > 0) {
        std::vector<double> tmp((size_t)nsteps * np * 6);
        PSM_HIP(h, hipMemcpy(tmp.data(), D.d_ft, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int t = 0; t < nsteps; ++t) {
            double* out = ft + (size_t)t * np * 6;
            const double* in = tmp.data() + (size_t)t * np * 6;
            for (int64_t k = 0; k < 6 * np; ++k) out[k] = 0.0;
            /* getForcesFromLattice fold onto map(tag), as sum_force_torque */
            for (int64_t q = 0; q < np; ++q) {
                const int64_t o = h->h_owner[q];
                for (int m = 0; m < 6; ++m) out[6 * o + m] += in[6 * q + m];
            }
        }
    }
    if (h->pp.enabled) {
        /* the inlet/outlet planes of f(T), for the adjoint of the last step */
        const size_t pbytes = (size_t)periodic_pressure_plane_elems(h) * state_elem_bytes(h);
        dfree_d(D.pp_last);
        D.pp_last = nullptr;
        PSM_HIP(h, dalloc_d(&D.pp_last, pbytes));
        rc = periodic_pressure_extract(h, h->f[h->cur_f], D.pp_last);
        if (rc) return rc;
        PSM_HIP(h, hipStreamSynchronize(h->stream));
    }
    D.T = nsteps;
    D.K = K;
    D.np = np;
    D.pp = h->pp.enabled;
    D.gen = h->geom_gen;
    D.ready = true;
    return LBDEM_PSM_OK;
}

int lbdem_psm_diff_backward(lbdem_psm* h, const double* g_ft, const double* g_fT, double* g_x3, double* g_v3,
                            double* g_w3, double* g_tau, double* g_sf, double* g_B, double* g_f0) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    auto& D = h->diff;
    if (!D.ready) return set_err(h, LBDEM_PSM_ERR_STATE, "diff_backward without a diff_forward");
    if (D.gen != h->geom_gen || D.np != h->np || D.pp != h->pp.enabled)
        return set_err(h, LBDEM_PSM_ERR_STATE,
                       "records, particles, tau or the periodic-pressure setting changed since diff_forward");
    const int T = D.T, K = D.K;
    const int64_t np = D.np, nc = h->g.nc, rcap = h->rec_cap;
    const int nseg = (int)D.seg_t0.size();
    auto seg_len = [&](int s) { return (s + 1 < nseg ? D.seg_t0[s + 1] : T) - D.seg_t0[s]; };
    int maxlen = 0;
    for (int s = 0; s < nseg; ++s) maxlen = std::max(maxlen, seg_len(s));
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    /* a budgeted schedule hands its checkpoints to the recompute pool as it
     * goes: one backward pass per forward window */
    if (K < 0) D.ready = false;
    int rc = LBDEM_PSM_OK;
    if (K > 0) {
        rc = ensure_states(h, D.seg, (size_t)std::max(maxlen - 1, 0));
        if (rc) return rc;
    }
    for (int k = 0; k < 2; ++k)
        if (!D.lam[k]) PSM_HIP(h, dalloc_d(&D.lam[k], h->pop_bytes));
    if (!D.Fscr) PSM_HIP(h, dalloc_d(&D.Fscr, 4 * rcap * sizeof(double)));
    /* per-record adjoint arrays at the records' list positions: sized by the
     * window's record count */
    const size_t tsz = h->precision == LBDEM_PSM_FP64 ? sizeof(double) : sizeof(float);
    const int64_t P = std::max<int64_t>(D.nrec, 1);
    if (D.pcap != P || D.ptsz != tsz) {
        dfree_d(D.gacc); dfree_d(D.pb); dfree_d(D.pc);
        D.gacc = nullptr; D.pb = D.pc = nullptr;
        PSM_HIP(h, dalloc_d(&D.gacc, 7 * (size_t)P * sizeof(double)));
        PSM_HIP(h, dalloc_d(&D.pb, (size_t)Q * P * tsz));
        PSM_HIP(h, dalloc_d(&D.pc, (size_t)5 * P * tsz));
        D.pcap = P;
        D.ptsz = tsz;
    }
    const int64_t nblk = (int64_t)h->g.n[0] * h->g.n[1] * h->g.L.nzc;
    if (D.opart_n < nblk + 8) {
        dfree_d(D.opart);
        PSM_HIP(h, dalloc_d(&D.opart, (nblk + 8) * sizeof(double)));
        D.opart_n = nblk + 8;
    }
    dfree_d(D.gpart);
    D.gpart = nullptr;
    PSM_HIP(h, dalloc_d(&D.gpart, 9 * std::max<int64_t>(np, 1) * sizeof(double)));
    if (!D.cell_out) PSM_HIP(h, dalloc_d(&D.cell_out, 2 * nc * sizeof(double)));
    const int gblocks = (int)std::min<int64_t>(4096, std::max<int64_t>(1, (rcap + 255) / 256));
    if (D.tau_part_n < gblocks + 8) {
        dfree_d(D.tau_part);
        PSM_HIP(h, dalloc_d(&D.tau_part, (gblocks + 8) * sizeof(double)));
        D.tau_part_n = gblocks + 8;
    }

    /* objective weights (the forward outputs were already returned) */
    const bool have_gft = g_ft && np > 0;
    if (have_gft) PSM_HIP(h, hipMemcpy(D.d_ft, g_ft, (size_t)T * np * 6 * sizeof(double), hipMemcpyHostToDevice));
    int cur = 0;
    PSM_HIP(h, hipMemsetAsync(D.lam[0], 0, h->pop_bytes, h->stream));
    PSM_HIP(h, hipMemsetAsync(D.lam[1], 0, h->pop_bytes, h->stream));
    double* tmp64 = nullptr;
    if (g_fT || g_f0) PSM_HIP(h, dalloc_d(&tmp64, (size_t)Q * nc * sizeof(double)));
    if (g_fT) {
        hipError_t e = hipMemcpy(tmp64, g_fT, (size_t)Q * nc * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) { dfree_d(tmp64); return hip_check(h, e, "hipMemcpy g_fT"); }
        rc = launch_pop_convert_buf(h, D.lam[cur], tmp64, true);
        if (rc) { dfree_d(tmp64); return rc; }
    }
    PSM_HIP(h, hipMemsetAsync(D.gacc, 0, 7 * (size_t)D.pcap * sizeof(double), h->stream));
    PSM_HIP(h, hipMemsetAsync(D.opart, 0, D.opart_n * sizeof(double), h->stream));

    const bool pp = D.pp;
    if (pp) {
        const size_t pbytes = (size_t)periodic_pressure_plane_elems(h) * state_elem_bytes(h);
        dfree_d(D.pp_next);
        dfree_d(D.pp_cur);
        D.pp_next = D.pp_cur = nullptr;
        PSM_HIP(h, dalloc_d(&D.pp_next, pbytes));
        PSM_HIP(h, dalloc_d(&D.pp_cur, pbytes));
        PSM_HIP(h, hipMemcpyAsync(D.pp_next, D.pp_last, pbytes, hipMemcpyDeviceToDevice, h->stream));
    }
    std::vector<const void*> st((size_t)std::max(maxlen, 1));
    for (int s = nseg - 1; s >= 0; --s) {
        const int t0 = D.seg_t0[s];
        const int len = seg_len(s);
        if (K < 0) {
            /* budgeted: grow the recompute pool only as far as this
             * segment needs (freed checkpoints were added to it) */
            while ((int)D.seg.size() < len - 1) {
                void* p = nullptr;
                hipError_t e = dalloc_d(&p, h->pop_bytes);
                if (e != hipSuccess) { dfree_d(tmp64); return hip_check(h, e, "hipMalloc (window states)"); }
                D.seg.push_back(p);
            }
        }
        st[0] = D.ckpt[s];
        for (int j = 1; j < len; ++j) {
            rc = launch_collide_stream_buf(h, st[j - 1], D.seg[j - 1], D.Fscr);
            if (rc) { dfree_d(tmp64); return rc; }
            st[j] = D.seg[j - 1];
        }
        for (int j = len - 1; j >= 0; --j) {
            const int t = t0 + j;
            if (pp) {
                /* f(t+1): the next segment's first state, or recomputed */
                const void* gpl = D.pp_next;
                if (j < len - 1) {
                    rc = periodic_pressure_extract(h, st[j + 1], D.pp_cur);
                    if (rc) { dfree_d(tmp64); return rc; }
                    gpl = D.pp_cur;
                }
                rc = periodic_pressure_adjoint_pre(h, st[j], gpl, D.lam[cur]);
                if (rc) { dfree_d(tmp64); return rc; }
            }
            rc = adjoint_step(h, st[j], D.lam[cur], D.lam[1 - cur],
                              have_gft ? D.d_ft + (size_t)t * np * 6 : nullptr);
            if (rc) { dfree_d(tmp64); return rc; }
            cur = 1 - cur;
        }
        if (pp && s > 0) {
            /* f(t0) is the state after the last step of the segment before */
            rc = periodic_pressure_extract(h, D.ckpt[s], D.pp_next);
            if (rc) { dfree_d(tmp64); return rc; }
        }
        if (K < 0) {
            /* this segment's checkpoint is no longer needed: it becomes
             * recompute storage for the (longer) segment before it; the
             * stream order keeps the adjoint steps above ahead of its reuse */
            D.seg.push_back(D.ckpt[s]);
            D.ckpt[s] = nullptr;
        }
    }

    /* reductions */
    double* red = h->red_tmp;   /* [0] omega adjoint, [1] tau adjoint through B */
    hipLaunchKernelGGL(k_sum_vec, dim3(1), dim3(256), 0, h->stream, D.opart, nblk, red);
    hipLaunchKernelGGL(k_grad_cells, dim3((unsigned)gblocks), dim3(256), 0, h->stream, rcap, h->sflag, h->rec.sus,
                       D.gacc, D.rpos, h->ooo, D.cell_out, nc, D.tau_part);
    hipLaunchKernelGGL(k_sum_vec, dim3(1), dim3(256), 0, h->stream, D.tau_part, (int64_t)gblocks, red + 1);
    if (np > 0) {
        GradArgs ga;
        for (int d = 0; d < 3; ++d) { ga.n[d] = h->g.n[d]; ga.nb[d] = h->g.nb[d]; }
        ga.sflag = h->sflag;
        ga.rsus = h->rec.sus;
        ga.rcand = h->rec.cand;
        ga.rid = h->rec.id;
        ga.rcap = rcap;
        ga.gacc = D.gacc;
        ga.rpos = D.rpos;
        ga.pcap = D.pcap;
        ga.xs = h->psum_x;
        ga.pw = h->pw;
        ga.pr = h->pr;
        ga.powner = h->powner;
        ga.p_ent_start = h->p_ent_start;
        ga.p_nent = h->p_nent;
        ga.owner_ent = h->owner_ent;
        ga.cand_bin = h->cand_bin;
        ga.ent_cap = h->ent_cap;
        ga.ooo = h->ooo;
        ga.rc = h->cell_rc;
        ga.vc = h->cell_vc;
        ga.overlap = h->desc.voxel_mode == LBDEM_PSM_VOXEL_OVERLAP ? 1 : 0;
        ga.gpart = D.gpart;
        hipLaunchKernelGGL(k_grad_particle, dim3((unsigned)np), dim3(256), 0, h->stream, ga);
        rc = order_orphans(h);
        if (rc) { dfree_d(tmp64); return rc; }
        hipLaunchKernelGGL(k_grad_orphans, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, h->stream, ga, (int64_t)np,
                           h->orph_cnt, h->orph_fill, h->orph_start, h->orph_sorted);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { dfree_d(tmp64); return hip_check(h, e, "adjoint reductions"); }
    if (g_f0) {
        rc = launch_pop_convert_buf(h, D.lam[cur], tmp64, false);
        if (rc) { dfree_d(tmp64); return rc; }
    }
    double hred[2] = {0.0, 0.0};
    e = hipMemcpyAsync(hred, red, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && g_f0) e = hipMemcpy(g_f0, tmp64, (size_t)Q * nc * sizeof(double), hipMemcpyDeviceToHost);
    dfree_d(tmp64);
    PSM_HIP(h, e);
    if (g_tau) {
        /* omega = 1/tau, B depends on tau - 1/2 */
        *g_tau = -hred[0] * (h->omega * h->omega) + hred[1];
    }
    /* the per-cell gradients go straight into the caller's arrays */
    if (g_sf) PSM_HIP(h, hipMemcpy(g_sf, D.cell_out, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost));
    if (g_B) PSM_HIP(h, hipMemcpy(g_B, D.cell_out + nc, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost));
    if (g_x3 || g_v3 || g_w3) {
        std::vector<double> gp(9 * (size_t)np);
        if (np > 0) PSM_HIP(h, hipMemcpy(gp.data(), D.gpart, gp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < 3 * np; ++k) {
            if (g_v3) g_v3[k] = 0.0;
            if (g_w3) g_w3[k] = 0.0;
            if (g_x3) g_x3[k] = 0.0;
        }
        for (int64_t q = 0; q < np; ++q) {
            const int64_t o = h->h_owner[q];
            for (int d = 0; d < 3; ++d) {
                if (g_v3) g_v3[3 * o + d] += gp[9 * q + d];
                if (g_w3) g_w3[3 * o + d] += gp[9 * q + 3 + d];
                if (g_x3) g_x3[3 * o + d] += gp[9 * q + 6 + d];
            }
        }
    }
    return LBDEM_PSM_OK;
}

int lbdem_psm_diff_release(lbdem_psm* h) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    return diff_free(h);
}

}  // extern "C"

/*
 * Boundary conditions and initial fields around the PSM step on gfx950.
 *
 * - Zhang-Kwok periodic pressure difference across a periodic axis
 *   (ZhangPeriodicPressureFunctional3D, src/periodicPressureFunctionals3D.hh:118-147,
 *   as driven by examples/showcaseRectChannel/showcaseRectChannel.cpp:258-273):
 *   the mean densities of the inlet plane (index 0 along the axis) and of the
 *   outlet plane (index n-1) over a cross-section box are taken before the
 *   collide+stream; after it, the populations that entered the inlet
 *   (c_d = +1) are rescaled in the f + t form by rho_in / <rho>_outlet and
 *   those that entered the outlet (c_d = -1) by rho_out / <rho>_inlet.
 *   Everything stays on the device: the plane sums are reduced in a fixed
 *   order into device scalars that the rescale kernels read.
 * - The Poiseuille profile with a linear pressure ramp at equilibrium
 *   (PoiseuilleProfileAndPressureGradient, :259-298), written directly by a
 *   kernel (no host field).
 */
#include "psm_common.h"

#include <algorithm>
#include <cmath>

namespace psm {

/* Cell (i along the axis dim, a, b along the other two axes in increasing
 * order) -> (x, y, z).  Written as selects: the if / else-if form of the
 * same mapping was compiled for dim = 2 into code that took y = i (the
 * generated blocks for dim = 1 and dim = 2 shared the assignment of y),
 * which the z-axis tests caught. */
__device__ __forceinline__ void plane_cell(int dim, int i, int a, int b, int& x, int& y, int& z) {
    x = dim == 0 ? i : a;
    y = dim == 0 ? a : (dim == 1 ? i : b);
    z = dim == 2 ? i : b;
}

struct PlaneArgs {
    int n[3];
    int dim;
    int a0, a1, b0, b1;     /* inclusive cross-section box */
    PopLayout L;
};

/* Partial sums of rhoBar = sum_i (f_i - t_i) over a plane's box, one per
 * workgroup (fixed tree). */
template <int PREC>
__global__ __launch_bounds__(256) void k_plane_rhobar(const typename Store<PREC>::S* __restrict__ f, PlaneArgs p,
                                                      int plane, double* __restrict__ part) {
    __shared__ double sh[256];
    const int na = p.a1 - p.a0 + 1, nb = p.b1 - p.b0 + 1;
    const int64_t ncell = (int64_t)na * nb;
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ncell;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int a = p.a0 + (int)(k / nb), b = p.b0 + (int)(k % nb);
        int x, y, z;
        plane_cell(p.dim, plane, a, b, x, y, z);
        const int64_t row = (int64_t)x * p.n[1] + y;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) s = s + (double)Store<PREC>::load(f + pop_at(row, i, z, p.L));
        acc += s;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

/* The same partial sums from the per-cell rhoBar the collide wrote for one
 * plane (cells[k], k = (a - a0) nb + (b - b0)): the same grid, the same k per
 * thread and the same tree as k_plane_rhobar, so the sums agree to the bit. */
__global__ __launch_bounds__(256) void k_plane_cells_partial(const double* __restrict__ cells, int64_t ncell,
                                                             double* __restrict__ part) {
    /* plane blockIdx.y: its ncell values from cells + y * ncell, its
     * partials to part + y * gridDim.x */
    cells += (int64_t)blockIdx.y * ncell;
    part += (int64_t)blockIdx.y * gridDim.x;
    __shared__ double sh[256];
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ncell;
         k += (int64_t)gridDim.x * blockDim.x)
        acc += cells[k];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

/* Fixed-order sum of n partials into out[0]; with several workgroups,
 * workgroup g sums part + g * n into out[g]. */
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int n,
                                                      double* __restrict__ out) {
    part += (int64_t)blockIdx.x * n;
    out += blockIdx.x;
    __shared__ double sh[256];
    double v = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) v += part[k];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

/* f_i <- (f_i + t_i) * rho_target / <rho> - t_i for the populations with
 * c_i[dim] == dir on the plane's box; <rho> = 1 + sum / ncell (Palabos'
 * computeAverageDensity: the mean rhoBar, then the full density).  Both
 * planes in one launch (blockIdx.y): the inlet (index 0) rescales the
 * populations arriving with c_d = +1 against the outlet's mean (sums[1]),
 * the outlet (index last) those arriving with c_d = -1 against the inlet's
 * mean (sums[0]).  The two sets of populations are disjoint, so the order
 * between them does not matter. */
template <int PREC>
__global__ __launch_bounds__(256) void k_zhang_rescale(typename Store<PREC>::S* __restrict__ f, PlaneArgs p, int last,
                                                       double rho_in, double rho_out,
                                                       const double* __restrict__ sums) {
    typedef typename Store<PREC>::T T;
    const bool outlet = blockIdx.y != 0;
    const int plane = outlet ? last : 0, dir = outlet ? -1 : 1;
    const double rho_target = outlet ? rho_out : rho_in;
    const int na = p.a1 - p.a0 + 1, nb = p.b1 - p.b0 + 1;
    const int64_t ncell = (int64_t)na * nb;
    const double avg = 1.0 + sums[outlet ? 0 : 1] / (double)ncell;
    const double factor = rho_target / avg;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ncell;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int a = p.a0 + (int)(k / nb), b = p.b0 + (int)(k % nb);
        int x, y, z;
        plane_cell(p.dim, plane, a, b, x, y, z);
        const int64_t row = (int64_t)x * p.n[1] + y;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (cc(i, p.dim) != dir) continue;
            typename Store<PREC>::S* q = f + pop_at(row, i, z, p.L);
            double v = (double)Store<PREC>::load(q) + wgt<double>(i);
            v = v * factor;
            Store<PREC>::store(q, (T)(v - wgt<double>(i)));
        }
    }
}

static PlaneArgs plane_args(const lbdem_psm* h) {
    PlaneArgs p;
    for (int d = 0; d < 3; ++d) p.n[d] = h->g.n[d];
    p.dim = h->pp.dim;
    p.a0 = h->pp.range[0]; p.a1 = h->pp.range[1];
    p.b0 = h->pp.range[2]; p.b1 = h->pp.range[3];
    p.L = h->g.L;
    return p;
}

static unsigned plane_blocks(const PlaneArgs& p) {
    const int64_t n = (int64_t)(p.a1 - p.a0 + 1) * (p.b1 - p.b0 + 1);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256));
}

template <int PREC>
static void pp_pre_t(lbdem_psm* h, const PlaneArgs& p, const void* fv) {
    typedef typename Store<PREC>::S S;
    const S* f = (const S*)fv;
    const unsigned nb = plane_blocks(p);
    const int last = h->g.n[p.dim] - 1;
    /* [0] inlet sum, [1] outlet sum, partials from [8] on */
    hipLaunchKernelGGL(k_plane_rhobar<PREC>, dim3(nb), dim3(256), 0, h->stream, f, p, 0, h->pp.buf + 8);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, h->stream, h->pp.buf + 8, (int)nb, h->pp.buf);
    hipLaunchKernelGGL(k_plane_rhobar<PREC>, dim3(nb), dim3(256), 0, h->stream, f, p, last, h->pp.buf + 8);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, h->stream, h->pp.buf + 8, (int)nb, h->pp.buf + 1);
}

template <int PREC>
static void pp_post_t(lbdem_psm* h, const PlaneArgs& p, void* fv) {
    typedef typename Store<PREC>::S S;
    S* f = (S*)fv;
    const unsigned nb = plane_blocks(p);
    const int last = h->g.n[p.dim] - 1;
    /* [0] inlet sum, [1] outlet sum */
    hipLaunchKernelGGL(k_zhang_rescale<PREC>, dim3(nb, 2), dim3(256), 0, h->stream, f, p, last, h->pp.rho_in,
                       h->pp.rho_out, (const double*)h->pp.buf);
}

int periodic_pressure_post_buf(lbdem_psm* h, void* f) {
    if (!h->pp.enabled) return LBDEM_PSM_OK;
    const PlaneArgs p = plane_args(h);
    switch (h->precision) {
        case LBDEM_PSM_FP64: pp_post_t<LBDEM_PSM_FP64>(h, p, f); break;
        case LBDEM_PSM_FP32_DEV: pp_post_t<LBDEM_PSM_FP32_DEV>(h, p, f); break;
        default: pp_post_t<LBDEM_PSM_FP16_STORE>(h, p, f); break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int periodic_pressure_cells(lbdem_psm* h, double** plane_rho, int* dim, int* a0, int* a1, int* b0, int* b1, int* nb,
                            int64_t* ncell) {
    *plane_rho = nullptr;
    if (!h->pp.enabled) return LBDEM_PSM_OK;
    const PlaneArgs p = plane_args(h);
    const int64_t n = (int64_t)(p.a1 - p.a0 + 1) * (p.b1 - p.b0 + 1);
    if (h->pp.cells_cap < 2 * n) {
        if (h->pp.cells) PSM_HIP(h, hipFree(h->pp.cells));
        h->pp.cells = nullptr;
        h->pp.cells_cap = 0;
        PSM_HIP(h, hipMalloc((void**)&h->pp.cells, 2 * n * sizeof(double)));
        h->pp.cells_cap = 2 * n;
    }
    *plane_rho = h->pp.cells;
    *dim = p.dim;
    *a0 = p.a0; *a1 = p.a1; *b0 = p.b0; *b1 = p.b1;
    *nb = p.b1 - p.b0 + 1;
    *ncell = n;
    return LBDEM_PSM_OK;
}

int periodic_pressure_sum_cells(lbdem_psm* h) {
    if (!h->pp.enabled) return LBDEM_PSM_OK;
    const PlaneArgs p = plane_args(h);
    const int64_t n = (int64_t)(p.a1 - p.a0 + 1) * (p.b1 - p.b0 + 1);
    const unsigned nb = plane_blocks(p);
    /* [0] inlet sum, [1] outlet sum, each plane's partials from [8] on */
    hipLaunchKernelGGL(k_plane_cells_partial, dim3(nb, 2), dim3(256), 0, h->stream, h->pp.cells, n, h->pp.buf + 8);
    hipLaunchKernelGGL(k_sum_partials, dim3(2), dim3(256), 0, h->stream, h->pp.buf + 8, (int)nb, h->pp.buf);
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int periodic_pressure_post(lbdem_psm* h) { return periodic_pressure_post_buf(h, h->f[h->cur_f]); }

/* ---- adjoint of the periodic-pressure step ---------------------------
 *
 * Forward step t: S_in, S_out = plane sums of rhoBar over the inlet and
 * outlet boxes of f(t); h = collide+stream(f(t)); then g = f(t+1) equals h
 * except g_i + t_i = a_in (h_i + t_i) for c_i[d] = +1 on the inlet box,
 * a_in = rho_in / m_out, and g_i + t_i = a_out (h_i + t_i) for c_i[d] = -1 on
 * the outlet box, a_out = rho_out / m_in, with m = 1 + S / A (A cells per
 * box).  Given lam' = dJ/dg:
 *   dJ/dh_i   = a lam'_i on the rescaled populations, lam'_i elsewhere;
 *   dJ/dm_out = -(sum_inlet lam'_i (g_i + t_i)) / m_out  (= dJ/da_in a_in / -m_out),
 *   dJ/dm_in  = -(sum_outlet lam'_i (g_i + t_i)) / m_in;
 * and m depends on every population of its box with weight 1/A, so
 * dJ/df(t) gains -dot_in / (m_out A) on every population of the outlet box
 * and -dot_out / (m_in A) on every population of the inlet box, next to the
 * collide+stream adjoint of dJ/dh. */

/* The inlet (0) and outlet (n-1) boxes of f, all 19 populations, compact:
 * out[(plane * 19 + i) * A + k]. */
template <int PREC>
__global__ __launch_bounds__(256) void k_pp_extract(const typename Store<PREC>::S* __restrict__ f, PlaneArgs p,
                                                    typename Store<PREC>::S* __restrict__ out) {
    const int na = p.a1 - p.a0 + 1, nb = p.b1 - p.b0 + 1;
    const int64_t A = (int64_t)na * nb;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < 2 * A;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int pl = (int)(k / A);
        const int64_t kk = k - pl * A;
        const int a = p.a0 + (int)(kk / nb), b = p.b0 + (int)(kk % nb);
        int x, y, z;
        plane_cell(p.dim, pl == 0 ? 0 : p.n[p.dim] - 1, a, b, x, y, z);
        const int64_t row = (int64_t)x * p.n[1] + y;
#pragma unroll
        for (int i = 0; i < Q; ++i) out[((int64_t)pl * Q + i) * A + kk] = f[pop_at(row, i, z, p.L)];
    }
}

/* Partial sums of lam_i (g_i + t_i) over the populations with c_i[dim] ==
 * dir on one plane box; g is the compact plane copy of f(t+1). */
template <int PREC>
__global__ __launch_bounds__(256) void k_pp_adj_dot(const typename Store<PREC>::S* __restrict__ lam,
                                                    const typename Store<PREC>::S* __restrict__ g, PlaneArgs p,
                                                    int pl, int dir, double* __restrict__ part) {
    __shared__ double sh[256];
    const int na = p.a1 - p.a0 + 1, nb = p.b1 - p.b0 + 1;
    const int64_t A = (int64_t)na * nb;
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < A;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int a = p.a0 + (int)(k / nb), b = p.b0 + (int)(k % nb);
        int x, y, z;
        plane_cell(p.dim, pl == 0 ? 0 : p.n[p.dim] - 1, a, b, x, y, z);
        const int64_t row = (int64_t)x * p.n[1] + y;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (cc(i, p.dim) != dir) continue;
            const double gi = (double)Store<PREC>::load(g + ((int64_t)pl * Q + i) * A + k) + wgt<double>(i);
            s += (double)Store<PREC>::load(lam + pop_at(row, i, z, p.L)) * gi;
        }
        acc += s;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

/* v[0] = plane-sum inlet, v[1] outlet (of f(t)); v[2] = dot inlet, v[3]
 * dot outlet -> v[4] added on the inlet box, v[5] on the outlet box. */
__global__ void k_pp_adj_coef(double* v, double A) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double m_in = 1.0 + v[0] / A, m_out = 1.0 + v[1] / A;
    v[4] = -v[3] / (m_in * A);
    v[5] = -v[2] / (m_out * A);
}

/* lam_i *= rho_target / <rho> on the populations with c_i[dim] == dir of one
 * plane box (the rescale factor of the forward step). */
template <int PREC>
__global__ __launch_bounds__(256) void k_pp_adj_scale(typename Store<PREC>::S* __restrict__ lam, PlaneArgs p,
                                                      int plane, int dir, double rho_target,
                                                      const double* __restrict__ rhobar_sum) {
    typedef typename Store<PREC>::T T;
    const int na = p.a1 - p.a0 + 1, nb = p.b1 - p.b0 + 1;
    const int64_t A = (int64_t)na * nb;
    const double factor = rho_target / (1.0 + rhobar_sum[0] / (double)A);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < A;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int a = p.a0 + (int)(k / nb), b = p.b0 + (int)(k % nb);
        int x, y, z;
        plane_cell(p.dim, plane, a, b, x, y, z);
        const int64_t row = (int64_t)x * p.n[1] + y;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (cc(i, p.dim) != dir) continue;
            typename Store<PREC>::S* q = lam + pop_at(row, i, z, p.L);
            Store<PREC>::store(q, (T)((double)Store<PREC>::load(q) * factor));
        }
    }
}

int64_t periodic_pressure_plane_elems(const lbdem_psm* h) {
    const PlaneArgs p = plane_args(h);
    return 2 * (int64_t)Q * (p.a1 - p.a0 + 1) * (p.b1 - p.b0 + 1);
}

template <int PREC>
static void pp_extract_t(lbdem_psm* h, const PlaneArgs& p, const void* f, void* out) {
    typedef typename Store<PREC>::S S;
    hipLaunchKernelGGL(k_pp_extract<PREC>, dim3(2 * plane_blocks(p)), dim3(256), 0, h->stream, (const S*)f, p,
                       (S*)out);
}

int periodic_pressure_extract(lbdem_psm* h, const void* f, void* out) {
    if (!h->pp.enabled) return LBDEM_PSM_OK;
    const PlaneArgs p = plane_args(h);
    switch (h->precision) {
        case LBDEM_PSM_FP64: pp_extract_t<LBDEM_PSM_FP64>(h, p, f, out); break;
        case LBDEM_PSM_FP32_DEV: pp_extract_t<LBDEM_PSM_FP32_DEV>(h, p, f, out); break;
        default: pp_extract_t<LBDEM_PSM_FP16_STORE>(h, p, f, out); break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

template <int PREC>
static void pp_adj_pre_t(lbdem_psm* h, const PlaneArgs& p, const void* fin, const void* gplanes, void* lam) {
    typedef typename Store<PREC>::S S;
    const unsigned nb = plane_blocks(p);
    const int last = h->g.n[p.dim] - 1;
    double* v = h->pp.buf;
    const double A = (double)(p.a1 - p.a0 + 1) * (double)(p.b1 - p.b0 + 1);
    /* plane sums of f(t) into v[0], v[1] */
    pp_pre_t<PREC>(h, p, fin);
    /* dots with the unscaled lam' into v[2], v[3] */
    hipLaunchKernelGGL(k_pp_adj_dot<PREC>, dim3(nb), dim3(256), 0, h->stream, (const S*)lam, (const S*)gplanes, p, 0,
                       1, v + 8);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, h->stream, v + 8, (int)nb, v + 2);
    hipLaunchKernelGGL(k_pp_adj_dot<PREC>, dim3(nb), dim3(256), 0, h->stream, (const S*)lam, (const S*)gplanes, p, 1,
                       -1, v + 8);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, h->stream, v + 8, (int)nb, v + 3);
    hipLaunchKernelGGL(k_pp_adj_coef, dim3(1), dim3(64), 0, h->stream, v, A);
    /* lam' -> dJ/dh on the rescaled populations */
    hipLaunchKernelGGL(k_pp_adj_scale<PREC>, dim3(nb), dim3(256), 0, h->stream, (S*)lam, p, 0, 1, h->pp.rho_in,
                       (const double*)(v + 1));
    hipLaunchKernelGGL(k_pp_adj_scale<PREC>, dim3(nb), dim3(256), 0, h->stream, (S*)lam, p, last, -1, h->pp.rho_out,
                       (const double*)v);
}

int periodic_pressure_adjoint_pre(lbdem_psm* h, const void* fin, const void* gplanes, void* lam) {
    if (!h->pp.enabled) return LBDEM_PSM_OK;
    const PlaneArgs p = plane_args(h);
    switch (h->precision) {
        case LBDEM_PSM_FP64: pp_adj_pre_t<LBDEM_PSM_FP64>(h, p, fin, gplanes, lam); break;
        default: pp_adj_pre_t<LBDEM_PSM_FP32_DEV>(h, p, fin, gplanes, lam); break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}


/* PoiseuilleProfileAndPressureGradient::operator() and poiseuilleVelocity
 * (src/periodicPressureFunctionals3D.hh:272-298), then the equilibrium at
 * rhoBar = rho - 1, j = rho u as initializeAtEquilibrium does. */
__device__ __forceinline__ double poiseuille_velocity(double i, double n) {
    const double nHalf = n / 2.;
    const double x = (i - nHalf) / (nHalf);
    return (1. - x * x);
}

template <int PREC>
__global__ __launch_bounds__(256) void k_init_poiseuille(typename Store<PREC>::S* __restrict__ f, int nx, int ny,
                                                         int nz, PopLayout L, int dim, double pHi, double pLo,
                                                         double uMax, int px, int py, int pz) {
    const int64_t nc = (int64_t)nx * ny * nz;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = c / nz;
        const int z = (int)(c - row * nz);
        const int y = (int)(row % ny);
        const int x = (int)(row / ny);
        double density, u[3] = {0.0, 0.0, 0.0};
        if (dim == 0) {
            density = pHi - (pHi - pLo) * (double)x / (double)(px - 1);
            u[0] = uMax * poiseuille_velocity((double)y, (double)py) * poiseuille_velocity((double)z, (double)pz);
        } else if (dim == 1) {
            density = pHi - (pHi - pLo) * (double)y / (double)(py - 1);
            u[1] = uMax * poiseuille_velocity((double)x, (double)px) * poiseuille_velocity((double)z, (double)pz);
        } else {
            density = pHi - (pHi - pLo) * (double)z / (double)(pz - 1);
            u[2] = uMax * poiseuille_velocity((double)x, (double)px) * poiseuille_velocity((double)y, (double)py);
        }
        const double rhoBar = density - 1.0;
        double j[3];
        for (int d = 0; d < 3; ++d) j[d] = density * u[d];
        const double invRho = 1.0 / (rhoBar + 1.0);
        const double jSqr = norm_sqr<double>(j);
        const double p = 1.5 * invRho;
#pragma unroll
        for (int i = 0; i < Q; ++i)
            Store<PREC>::store(f + pop_at(row, i, z, L), (typename Store<PREC>::T)equilibrium<double>(i, rhoBar, p, j, jSqr));
    }
}

}  // namespace psm

using namespace psm;

extern "C" {

int lbdem_psm_set_periodic_pressure(lbdem_psm* h, int32_t enable, int32_t dimension, double rho_in, double rho_out,
                                    const int32_t* range4) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    if (!enable) {
        if (h->pp.enabled) h->geom_gen += 1;   /* a differentiable window ran with it */
        h->pp.enabled = false;
        return LBDEM_PSM_OK;
    }
    if (dimension < 0 || dimension > 2) return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "dimension must be 0, 1 or 2");
    if (!h->g.periodic[dimension])
        return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "the periodic-pressure axis must be periodic");
    if (!(rho_in > 0.0) || !(rho_out > 0.0) || !std::isfinite(rho_in) || !std::isfinite(rho_out))
        return set_err(h, LBDEM_PSM_ERR_NUMERIC, "target densities must be finite and > 0");
    if (h->g.n[dimension] < 2) return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "axis too short for inlet and outlet");
    const int ax = dimension == 0 ? 1 : 0, bx = dimension == 2 ? 1 : 2;
    int r[4];
    if (range4) {
        for (int k = 0; k < 4; ++k) r[k] = range4[k];
    } else {
        r[0] = 1; r[1] = h->g.n[ax] - 2; r[2] = 1; r[3] = h->g.n[bx] - 2;
    }
    if (r[0] < 0 || r[1] >= h->g.n[ax] || r[0] > r[1] || r[2] < 0 || r[3] >= h->g.n[bx] || r[2] > r[3])
        return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "cross-section box outside the lattice or empty");
    if (!h->pp.buf) PSM_HIP(h, hipMalloc((void**)&h->pp.buf, (8 + 2 * 1024) * sizeof(double)));
    h->pp.enabled = true;
    h->geom_gen += 1;   /* the step map changed: a window from before no longer applies */
    h->pp.dim = dimension;
    h->pp.rho_in = rho_in;
    h->pp.rho_out = rho_out;
    for (int k = 0; k < 4; ++k) h->pp.range[k] = r[k];
    return LBDEM_PSM_OK;
}

int lbdem_psm_init_poiseuille(lbdem_psm* h, int32_t dimension, double rho_hi, double rho_lo, double u_max,
                              const int32_t* n3) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    if (dimension < 0 || dimension > 2) return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "dimension must be 0, 1 or 2");
    if (!std::isfinite(rho_hi) || !std::isfinite(rho_lo) || !std::isfinite(u_max) || !(rho_hi > 0.0) ||
        !(rho_lo > 0.0) || std::fabs(u_max) >= 0.3)
        return set_err(h, LBDEM_PSM_ERR_NUMERIC, "densities must be > 0 and |u_max| < 0.3");
    int pn[3];
    for (int d = 0; d < 3; ++d) pn[d] = n3 ? n3[d] : h->g.n[d];
    if (pn[dimension] < 2 || pn[0] < 1 || pn[1] < 1 || pn[2] < 1)
        return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "profile sizes must be >= 1 (>= 2 along the axis)");
    const int64_t nc = h->g.nc;
    dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(65536, (nc + 255) / 256))), block(256);
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_init_poiseuille<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (double*)h->f[h->cur_f],
                               h->g.n[0], h->g.n[1], h->g.n[2], h->g.L, dimension, rho_hi, rho_lo, u_max, pn[0],
                               pn[1], pn[2]);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_init_poiseuille<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (float*)h->f[h->cur_f],
                               h->g.n[0], h->g.n[1], h->g.n[2], h->g.L, dimension, rho_hi, rho_lo, u_max, pn[0],
                               pn[1], pn[2]);
            break;
        default:
            hipLaunchKernelGGL(k_init_poiseuille<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream,
                               (_Float16*)h->f[h->cur_f], h->g.n[0], h->g.n[1], h->g.n[2], h->g.L, dimension, rho_hi,
                               rho_lo, u_max, pn[0], pn[1], pn[2]);
            break;
    }
    PSM_HIP(h, hipGetLastError());
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    return LBDEM_PSM_OK;
}

}  // extern "C"

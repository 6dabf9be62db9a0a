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
 * order) -> (x, y, z). */
__device__ __forceinline__ void plane_cell(int dim, int i, int a, int b, int& x, int& y, int& z) {
    if (dim == 0) { x = i; y = a; z = b; }
    else if (dim == 1) { x = a; y = i; z = b; }
    else { x = a; y = b; z = i; }
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

/* Fixed-order sum of n partials into out[0]. */
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int n,
                                                      double* __restrict__ out) {
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
 * computeAverageDensity: the mean rhoBar, then the full density). */
template <int PREC>
__global__ __launch_bounds__(256) void k_zhang_rescale(typename Store<PREC>::S* __restrict__ f, PlaneArgs p, int plane,
                                                       int dir, double rho_target,
                                                       const double* __restrict__ rhobar_sum) {
    typedef typename Store<PREC>::T T;
    const int na = p.a1 - p.a0 + 1, nb = p.b1 - p.b0 + 1;
    const int64_t ncell = (int64_t)na * nb;
    const double avg = 1.0 + rhobar_sum[0] / (double)ncell;
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
static void pp_pre_t(lbdem_psm* h, const PlaneArgs& p) {
    typedef typename Store<PREC>::S S;
    const S* f = (const S*)h->f[h->cur_f];
    const unsigned nb = plane_blocks(p);
    const int last = h->g.n[p.dim] - 1;
    /* [0] inlet sum, [1] outlet sum, partials from [8] on */
    hipLaunchKernelGGL(k_plane_rhobar<PREC>, dim3(nb), dim3(256), 0, h->stream, f, p, 0, h->pp.buf + 8);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, h->stream, h->pp.buf + 8, (int)nb, h->pp.buf);
    hipLaunchKernelGGL(k_plane_rhobar<PREC>, dim3(nb), dim3(256), 0, h->stream, f, p, last, h->pp.buf + 8);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, h->stream, h->pp.buf + 8, (int)nb, h->pp.buf + 1);
}

template <int PREC>
static void pp_post_t(lbdem_psm* h, const PlaneArgs& p) {
    typedef typename Store<PREC>::S S;
    S* f = (S*)h->f[h->cur_f];
    const unsigned nb = plane_blocks(p);
    const int last = h->g.n[p.dim] - 1;
    /* inlet: populations arriving with c_d = +1, against the outlet's mean */
    hipLaunchKernelGGL(k_zhang_rescale<PREC>, dim3(nb), dim3(256), 0, h->stream, f, p, 0, 1, h->pp.rho_in,
                       (const double*)(h->pp.buf + 1));
    /* outlet: populations arriving with c_d = -1, against the inlet's mean */
    hipLaunchKernelGGL(k_zhang_rescale<PREC>, dim3(nb), dim3(256), 0, h->stream, f, p, last, -1, h->pp.rho_out,
                       (const double*)(h->pp.buf));
}

int periodic_pressure_pre(lbdem_psm* h) {
    if (!h->pp.enabled) return LBDEM_PSM_OK;
    const PlaneArgs p = plane_args(h);
    switch (h->precision) {
        case LBDEM_PSM_FP64: pp_pre_t<LBDEM_PSM_FP64>(h, p); break;
        case LBDEM_PSM_FP32_DEV: pp_pre_t<LBDEM_PSM_FP32_DEV>(h, p); break;
        default: pp_pre_t<LBDEM_PSM_FP16_STORE>(h, p); break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int periodic_pressure_post(lbdem_psm* h) {
    if (!h->pp.enabled) return LBDEM_PSM_OK;
    const PlaneArgs p = plane_args(h);
    switch (h->precision) {
        case LBDEM_PSM_FP64: pp_post_t<LBDEM_PSM_FP64>(h, p); break;
        case LBDEM_PSM_FP32_DEV: pp_post_t<LBDEM_PSM_FP32_DEV>(h, p); break;
        default: pp_post_t<LBDEM_PSM_FP16_STORE>(h, p); break;
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
    if (!h->pp.buf) PSM_HIP(h, hipMalloc((void**)&h->pp.buf, (8 + 1024) * sizeof(double)));
    h->pp.enabled = true;
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

/*
 * Fused PSM collide + stream on gfx950.
 *
 * Replaces lattice.collideAndStream() over cells carrying
 * IBcompositeDynamics (src/ibCompositeDynamics3D.hh:130-203;
 * examples/sedimentingSphere/sedimentingSphere.cpp:221).
 *
 * Layout: populations are stored as f - t in FP64, FP32 or FP16, in
 * z-chunked row blocks (PopLayout): the 19 populations of a (row, z chunk)
 * are consecutive stretches.  One workgroup owns one block: every lane reads
 * its cell's 19 populations with unit stride (fully coalesced, one
 * contiguous stretch per workgroup), collides, and pushes each population to
 * its neighbour x + c_q, i.e. into one of 9 row blocks.
 * Cells with a coupling record are marked by a 1-byte flag; only for those
 * is the record (solid fraction, particle velocity) read and the per-cell
 * hydrodynamic force written.
 * Non-periodic axes use half-way bounce-back.
 */
#include "psm_collide_kernel.h"

namespace psm {

template <int PREC>
__global__ __launch_bounds__(256) void k_init_eq(typename Store<PREC>::S* __restrict__ f,
                                                 const double* __restrict__ rho,
                                                 const double* __restrict__ u, int64_t nc, int nz,
                                                 PopLayout L) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = c / nz;
        const int z = (int)(c - row * nz);
        const double r = rho ? rho[c] : 1.0;
        const double rhoBar = r - 1.0;
        double j[3];
        for (int d = 0; d < 3; ++d) j[d] = u ? r * u[d * nc + c] : 0.0;
        const double invRho = 1.0 / (rhoBar + 1.0);
        const double jSqr = norm_sqr<double>(j);
        const double p = 1.5 * invRho;
#pragma unroll
        for (int i = 0; i < Q; ++i)
            Store<PREC>::store(f + pop_at(row, i, z, L), (typename Store<PREC>::T)equilibrium<double>(i, rhoBar, p, j, jSqr));
    }
}

/* Host layout f64[q*Nc + c] <-> row-blocked store. */
template <int PREC>
__global__ __launch_bounds__(256) void k_pop_convert(typename Store<PREC>::S* __restrict__ f,
                                                     double* __restrict__ f64, int64_t nc, int nz,
                                                     PopLayout L, int to_store) {
    const int64_t total = (int64_t)Q * nc;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(k / nc);
        const int64_t c = k - (int64_t)q * nc;
        const int64_t row = c / nz;
        const int z = (int)(c - row * nz);
        const size_t at = pop_at(row, q, z, L);
        if (to_store) Store<PREC>::store(f + at, (typename Store<PREC>::T)f64[k]);
        else f64[k] = (double)Store<PREC>::load(f + at);
    }
}

/* FP64 global mass / momentum / health.  Stage 1: one FP64 partial per
 * block (fixed tree); stage 2: one block sums the partials in index order. */
template <int PREC>
__global__ __launch_bounds__(256) void k_cons_partial(const typename Store<PREC>::S* __restrict__ f,
                                                      int64_t nc, int nz, PopLayout L,
                                                      double* __restrict__ part) {
    __shared__ double sh[5][256];
    double m = 0.0, jx = 0.0, jy = 0.0, jz = 0.0, umax = 0.0, nonfin = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        double fc[Q];
        bool bad = false;
        const int64_t row = c / nz;
        const int z = (int)(c - row * nz);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            fc[i] = (double)Store<PREC>::load(f + pop_at(row, i, z, L));
            bad |= !isfinite(fc[i]);
        }
        double rb, j[3];
        moments<double>(fc, rb, j);
        m += rb; jx += j[0]; jy += j[1]; jz += j[2];
        if (bad) nonfin += 1.0;
        else {
            const double ir = 1.0 / (rb + 1.0);
            const double u2 = ((j[0] * j[0] + j[1] * j[1]) + j[2] * j[2]) * ir * ir;
            umax = fmax(umax, u2);
        }
    }
    sh[0][threadIdx.x] = m; sh[1][threadIdx.x] = jx; sh[2][threadIdx.x] = jy;
    sh[3][threadIdx.x] = jz; sh[4][threadIdx.x] = umax;
    __shared__ double shn[256];
    shn[threadIdx.x] = nonfin;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
            sh[4][threadIdx.x] = fmax(sh[4][threadIdx.x], sh[4][threadIdx.x + w]);
            shn[threadIdx.x] += shn[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 5; ++k) part[blockIdx.x * 6 + k] = sh[k][0];
        part[blockIdx.x * 6 + 5] = shn[0];
    }
}

__global__ __launch_bounds__(256) void k_cons_final(const double* __restrict__ part, int nblocks,
                                                    double* __restrict__ out) {
    __shared__ double sh[6][256];
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        for (int k = 0; k < 4; ++k) v[k] += part[b * 6 + k];
        v[4] = fmax(v[4], part[b * 6 + 4]);
        v[5] += part[b * 6 + 5];
    }
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
            sh[4][threadIdx.x] = fmax(sh[4][threadIdx.x], sh[4][threadIdx.x + w]);
            sh[5][threadIdx.x] += sh[5][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) out[k] = sh[k][0];
}

template <int PREC>
__global__ __launch_bounds__(256) void k_define_velocity(typename Store<PREC>::S* __restrict__ f,
                                                         const int32_t* __restrict__ dv_cell,
                                                         const double* __restrict__ dv_us,
                                                         const int64_t* __restrict__ count,
                                                         int64_t cap, int nz, PopLayout L) {
    const int64_t n = count[0] < cap ? count[0] : cap;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = dv_cell[k];
        const int64_t row = c / nz;
        const int z = (int)(c - row * nz);
        const double u[3] = {dv_us[3 * k], dv_us[3 * k + 1], dv_us[3 * k + 2]};
        /* IBcompositeDynamics::defineVelocity, src/ibCompositeDynamics3D.hh:118-127:
         * equilibrium at rhoBar = 0 with j = u. */
        const double uSqr = norm_sqr<double>(u);
        const double p = 1.5 * (1.0 / (0.0 + 1.0));
#pragma unroll
        for (int i = 0; i < Q; ++i)
            Store<PREC>::store(f + pop_at(row, i, z, L), (typename Store<PREC>::T)equilibrium<double>(i, 0.0, p, u, uSqr));
    }
}

static CollideArgs make_args(lbdem_psm* h) {
    CollideArgs a;
    a.nx = h->g.n[0]; a.ny = h->g.n[1]; a.nz = h->g.n[2];
    a.px = h->g.periodic[0]; a.py = h->g.periodic[1]; a.pz = h->g.periodic[2];
    a.nc = h->g.nc;
    a.L = h->g.L;
    a.omega = h->omega; a.ooo = h->ooo;
    for (int d = 0; d < 3; ++d) a.acc[d] = h->desc.body_accel[d];
    a.sflag = h->sflag;
    a.rsus = h->rec.sus; a.rF = h->rec.F;
    a.div_nzc = make_fastdiv((uint32_t)h->g.L.nzc);
    a.sub_shift = 0;
    while ((collide_wg(h->g.L.zc) << a.sub_shift) < h->g.L.zc) ++a.sub_shift;
    a.div_ny = make_fastdiv((uint32_t)h->g.n[1]);
    a.rm4 = h->rec.m4;
    a.plane_rho = nullptr;
    a.pdim = a.plast = a.pa0 = a.pa1 = a.pb0 = a.pb1 = a.pnb = 0;
    a.pncell = 0;
    return a;
}

template <>
void launch_cs_prec<LBDEM_PSM_FP64>(lbdem_psm* h, const CollideArgs& a, const void* vin, void* vout) {
    typedef typename Store<LBDEM_PSM_FP64>::S S;
    const S* fin = (const S*)vin;
    S* fout = (S*)vout;
    const int wgz = collide_wg(a.L.zc);
    dim3 grid((unsigned)((int64_t)a.nx * a.ny * a.L.nzc * (a.L.zc / wgz))), block((unsigned)wgz);
    if (!h->cs_lds_ready) {
        const int dev = h->desc.device, cap = collide_wave_cap(LBDEM_PSM_FP64);
        h->cs_lds[1][1] = collide_lds_for(k_collide_stream<LBDEM_PSM_FP64, true, true>, dev, wgz, cap);
        h->cs_lds[1][0] = collide_lds_for(k_collide_stream<LBDEM_PSM_FP64, true, false>, dev, wgz, cap);
        h->cs_lds[0][1] = collide_lds_for(k_collide_stream<LBDEM_PSM_FP64, false, true>, dev, wgz, cap);
        h->cs_lds[0][0] = collide_lds_for(k_collide_stream<LBDEM_PSM_FP64, false, false>, dev, wgz, cap);
        h->cs_lds_ready = true;
    }
    const unsigned lds = h->cs_lds[h->has_force ? 1 : 0][a.pz ? 1 : 0];
    if (h->has_force) {
        if (a.pz) hipLaunchKernelGGL((k_collide_stream<LBDEM_PSM_FP64, true, true>), grid, block, lds, h->stream, fin, fout, a);
        else hipLaunchKernelGGL((k_collide_stream<LBDEM_PSM_FP64, true, false>), grid, block, lds, h->stream, fin, fout, a);
    } else {
        if (a.pz) hipLaunchKernelGGL((k_collide_stream<LBDEM_PSM_FP64, false, true>), grid, block, lds, h->stream, fin, fout, a);
        else hipLaunchKernelGGL((k_collide_stream<LBDEM_PSM_FP64, false, false>), grid, block, lds, h->stream, fin, fout, a);
    }
}

static void launch_any(lbdem_psm* h, const CollideArgs& a, const void* fin, void* fout) {
    switch (h->precision) {
        case LBDEM_PSM_FP64: launch_cs_prec<LBDEM_PSM_FP64>(h, a, fin, fout); break;
        case LBDEM_PSM_FP32_DEV: launch_cs_prec<LBDEM_PSM_FP32_DEV>(h, a, fin, fout); break;
        default: launch_cs_prec<LBDEM_PSM_FP16_STORE>(h, a, fin, fout); break;
    }
}

int launch_collide_stream(lbdem_psm* h) {
    CollideArgs a = make_args(h);
    /* with the periodic pressure on, the collide writes the plane cells'
     * rhoBar of f(t) as it loads them; they are summed before the rescale */
    int rc = periodic_pressure_cells(h, &a.plane_rho, &a.pdim, &a.pa0, &a.pa1, &a.pb0, &a.pb1, &a.pnb, &a.pncell);
    if (rc) return rc;
    a.plast = h->g.n[a.pdim] - 1;
    hipEvent_t ev = nullptr;
    prof_begin(h, 0, &ev);
    launch_any(h, a, h->f[h->cur_f], h->f[1 - h->cur_f]);
    PSM_HIP(h, hipGetLastError());
    prof_end(h, 0, ev);
    h->cur_f = 1 - h->cur_f;
    rc = periodic_pressure_sum_cells(h);
    if (rc) return rc;
    return periodic_pressure_post(h);
}

/* One collide+stream between caller-chosen buffers (with the periodic-
 * pressure rescale when enabled); the per-cell forces go to rF (4 per
 * record slot) instead of the records. */
int launch_collide_stream_buf(lbdem_psm* h, const void* fin, void* fout, void* rF) {
    CollideArgs a = make_args(h);
    a.rF = rF;
    /* the plane sums of fin from the collide's own loads, as in
     * launch_collide_stream (the same sums, bit for bit) */
    int rc = periodic_pressure_cells(h, &a.plane_rho, &a.pdim, &a.pa0, &a.pa1, &a.pb0, &a.pb1, &a.pnb, &a.pncell);
    if (rc) return rc;
    a.plast = h->g.n[a.pdim] - 1;
    launch_any(h, a, fin, fout);
    PSM_HIP(h, hipGetLastError());
    rc = periodic_pressure_sum_cells(h);
    if (rc) return rc;
    return periodic_pressure_post_buf(h, fout);
}

static unsigned grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

int launch_init_equilibrium(lbdem_psm* h, const double* d_rho, const double* d_u) {
    const int64_t nc = h->g.nc;
    dim3 grid(grid_for(nc)), block(256);
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_init_eq<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (double*)h->f[h->cur_f], d_rho, d_u, nc, h->g.n[2], h->g.L);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_init_eq<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (float*)h->f[h->cur_f], d_rho, d_u, nc, h->g.n[2], h->g.L);
            break;
        default:
            hipLaunchKernelGGL(k_init_eq<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream, (_Float16*)h->f[h->cur_f], d_rho, d_u, nc, h->g.n[2], h->g.L);
            break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int launch_pop_convert(lbdem_psm* h, double* d_f64, bool to_store) {
    return launch_pop_convert_buf(h, h->f[h->cur_f], d_f64, to_store);
}

int launch_pop_convert_buf(lbdem_psm* h, void* buf, double* d_f64, bool to_store) {
    const int64_t nc = h->g.nc;
    const int64_t total = (int64_t)Q * nc;
    const int nz = h->g.n[2];
    const PopLayout L = h->g.L;
    dim3 grid(grid_for(total)), block(256);
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_pop_convert<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (double*)buf, d_f64, nc, nz, L, (int)to_store);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_pop_convert<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (float*)buf, d_f64, nc, nz, L, (int)to_store);
            break;
        default:
            hipLaunchKernelGGL(k_pop_convert<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream, (_Float16*)buf, d_f64, nc, nz, L, (int)to_store);
            break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int launch_conservation(lbdem_psm* h, double* mass, double* mom3, double* max_u,
                        int64_t* n_nonfinite) {
    const int64_t nc = h->g.nc;
    const int nblocks = (int)std::min<int64_t>((nc + 255) / 256, 4096);
    double* part = h->red_tmp;
    double* out = h->red_tmp + 6 * 4096;
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_cons_partial<LBDEM_PSM_FP64>, dim3(nblocks), dim3(256), 0, h->stream, (const double*)h->f[h->cur_f], nc, h->g.n[2], h->g.L, part);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_cons_partial<LBDEM_PSM_FP32_DEV>, dim3(nblocks), dim3(256), 0, h->stream, (const float*)h->f[h->cur_f], nc, h->g.n[2], h->g.L, part);
            break;
        default:
            hipLaunchKernelGGL(k_cons_partial<LBDEM_PSM_FP16_STORE>, dim3(nblocks), dim3(256), 0, h->stream, (const _Float16*)h->f[h->cur_f], nc, h->g.n[2], h->g.L, part);
            break;
    }
    hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(256), 0, h->stream, part, nblocks, out);
    PSM_HIP(h, hipGetLastError());
    double hv[6];
    PSM_HIP(h, hipMemcpyAsync(hv, out, sizeof(hv), hipMemcpyDeviceToHost, h->stream));
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    if (mass) *mass = hv[0];
    if (mom3) { mom3[0] = hv[1]; mom3[1] = hv[2]; mom3[2] = hv[3]; }
    if (max_u) *max_u = sqrt(hv[4]);
    if (n_nonfinite) *n_nonfinite = (int64_t)hv[5];
    return LBDEM_PSM_OK;
}

/* Density and velocity per cell (Palabos computeDensity / computeVelocity
 * of the BGK cell; with Guo forcing u = j/rho + a/2). Host order: c = (x*ny
 * + y)*nz + z, velocity component-major. */
template <int PREC>
__global__ __launch_bounds__(256) void k_macroscopic(const typename Store<PREC>::S* __restrict__ f, int64_t nc, int nz,
                                                     PopLayout L, double a0, double a1, double a2,
                                                     double* __restrict__ rho, double* __restrict__ u) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = c / nz;
        const int z = (int)(c - row * nz);
        double fc[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) fc[i] = (double)Store<PREC>::load(f + pop_at(row, i, z, L));
        double rb, j[3];
        moments<double>(fc, rb, j);
        const double r = rb + 1.0;
        if (rho) rho[c] = r;
        if (u) {
            u[c] = j[0] / r + 0.5 * a0;
            u[nc + c] = j[1] / r + 0.5 * a1;
            u[2 * nc + c] = j[2] / r + 0.5 * a2;
        }
    }
}

int launch_macroscopic(lbdem_psm* h, double* d_rho, double* d_u) {
    const int64_t nc = h->g.nc;
    dim3 grid(grid_for(nc)), block(256);
    const double* a = h->desc.body_accel;
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_macroscopic<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (const double*)h->f[h->cur_f], nc, h->g.n[2], h->g.L, a[0], a[1], a[2], d_rho, d_u);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_macroscopic<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (const float*)h->f[h->cur_f], nc, h->g.n[2], h->g.L, a[0], a[1], a[2], d_rho, d_u);
            break;
        default:
            hipLaunchKernelGGL(k_macroscopic<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream, (const _Float16*)h->f[h->cur_f], nc, h->g.n[2], h->g.L, a[0], a[1], a[2], d_rho, d_u);
            break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

/* Plain device copy of n16 16-byte words, one word per lane and one lane
 * per word (a grid-stride loop over the same store moved 4.7-5.2 TB/s, a
 * one-pass grid 6.5 TB/s), streaming loads and stores, each XCD on one
 * contiguous stretch as in the collide: the bandwidth ceiling the collide
 * is measured against (every byte of a population store read once and
 * written once, like the collide's algorithmic traffic), and a transfer of
 * known size for counter calibration. */
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src, u32x4* __restrict__ dst, int64_t n16) {
    const int64_t k = (int64_t)xcd_block((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
    if (k < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(src + k), dst + k);
}

/* Wavefronts per SIMD the copy is held to: six moved 0.7-3 % more than its
 * register limit of eight (512x512x768 stores, FP64 and FP32, three lattices
 * each); five or fewer moved less (four: 5-8 % less). */
constexpr int COPY_WAVES = 6;

/* The copy as one-pass launches of at most 2^31 words each (the work-item
 * count of a launch is a 32-bit number). */
static hipError_t copy16_launch(hipStream_t st, const u32x4* src, u32x4* dst, int64_t n16, unsigned lds) {
    const int64_t seg = (int64_t)1 << 31;
    for (int64_t o = 0; o < n16; o += seg) {
        const int64_t m = std::min<int64_t>(seg, n16 - o);
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)((m + 255) / 256)), dim3(256), lds, st, src + o, dst + o, m);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int launch_define_velocity(lbdem_psm* h) {
    if (!h->dv_cell) return LBDEM_PSM_OK;
    dim3 grid(1024), block(256);
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_define_velocity<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (double*)h->f[h->cur_f], h->dv_cell, h->dv_us, h->counters + 5, h->orphan_cap, h->g.n[2], h->g.L);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_define_velocity<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (float*)h->f[h->cur_f], h->dv_cell, h->dv_us, h->counters + 5, h->orphan_cap, h->g.n[2], h->g.L);
            break;
        default:
            hipLaunchKernelGGL(k_define_velocity<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream, (_Float16*)h->f[h->cur_f], h->dv_cell, h->dv_us, h->counters + 5, h->orphan_cap, h->g.n[2], h->g.L);
            break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

}  // namespace psm

using namespace psm;

extern "C" {

int lbdem_psm_copy_bandwidth(lbdem_psm* h, int32_t reps, double* ms_per_copy, double* bytes_per_copy) {
    if (!h || reps < 1) return LBDEM_PSM_ERR_INVALID_ARG;
    /* the current population store into the other one (which holds no state
     * between steps); the store size is a multiple of 16 bytes */
    const int64_t n16 = (int64_t)(h->pop_bytes / 16);
    const u32x4* src = (const u32x4*)h->f[h->cur_f];
    u32x4* dst = (u32x4*)h->f[1 - h->cur_f];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PSM_HIP(h, hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) { (void)hipEventDestroy(e0); return hip_check(h, e, "hipEventCreate"); }
    const unsigned lds = collide_lds_for(k_copy16, h->desc.device, 256, COPY_WAVES);
    e = copy16_launch(h->stream, src, dst, n16, lds);   /* warm-up */
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = copy16_launch(h->stream, src, dst, n16, lds);
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_check(h, e, "copy_bandwidth");
    if (ms_per_copy) *ms_per_copy = (double)ms / reps;
    if (bytes_per_copy) *bytes_per_copy = 2.0 * (double)(n16 * 16);
    return LBDEM_PSM_OK;
}

}  // extern "C"

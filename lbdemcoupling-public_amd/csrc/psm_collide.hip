/*
 * Fused PSM collide + stream on gfx950.
 *
 * Replaces lattice.collideAndStream() over cells carrying
 * IBcompositeDynamics (src/ibCompositeDynamics3D.hh:130-203;
 * examples/sedimentingSphere/sedimentingSphere.cpp:221).
 *
 * Layout: populations are structure-of-arrays, f[q][c] with
 * c = (x*ny + y)*nz + z, stored as f - t in FP64, FP32 or FP16.  One
 * workgroup owns whole z-rows: every lane reads its cell's 19 populations
 * with unit stride (fully coalesced), collides, and pushes each population
 * to its neighbour x + c_q.  A z-row is written only by its own workgroup,
 * so every cache line of the destination is filled by one workgroup.
 * Cells with a coupling record are found through slot[c] (int32, -1 for
 * plain fluid); the record (solid fraction, particle velocity) is read and
 * the per-cell hydrodynamic force written only for those cells.
 * Non-periodic axes use half-way bounce-back.
 */
#include "psm_common.h"

namespace psm {

struct CollideArgs {
    int nx, ny, nz;
    int px, py, pz;
    int64_t nc;
    double omega, ooo;
    double acc[3];
    const int32_t* slot;
    const double* rsf;
    const double* rus;
    double* rF;
    int64_t rcap;
};

/* BGK output of population i from the pre-collision value and the
 * pre-collision moments (the bgk() expression, evaluated for one i). */
template <typename T, bool FORCE>
struct BgkPre {
    T rhoBar, rho, p, omm, omega, jSqr;
    T jEq[3], u[3], acc[3], pref, ua;
    __device__ __forceinline__ void init(const T* f, T omega_, const T* acc_) {
        T j[3];
        moments<T>(f, rhoBar, j);
        rho = rhoBar + T(1);
        const T invRho = T(1) / rho;
        jEq[0] = j[0]; jEq[1] = j[1]; jEq[2] = j[2];
        u[0] = u[1] = u[2] = T(0);
        omega = omega_;
        if (FORCE) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                acc[d] = acc_[d];
                u[d] = j[d] * invRho + T(0.5) * acc[d];
                jEq[d] = rho * u[d];
            }
        }
        jSqr = norm_sqr<T>(jEq);
        p = T(1.5) * invRho;
        omm = T(1) - omega;
        if (FORCE) {
            pref = (T(1) - T(0.5) * omega) * rho;
            ua = (u[0] * acc[0] + u[1] * acc[1]) + u[2] * acc[2];
        }
    }
    __device__ __forceinline__ T out(int i, T fi) const {
        T r = fi * omm + omega * equilibrium<T>(i, rhoBar, p, jEq, jSqr);
        if (FORCE) {
            const T ca = cdot<T>(i, acc);
            const T cu = cdot<T>(i, u);
            const T g = T(3) * (ca - ua) + T(9) * (cu * ca);
            r = r + (pref * wgt<T>(i)) * g;
        }
        return r;
    }
};

/* One workgroup per z-row (x, y uniform): per direction the destination
 * row and the x/y wall test are uniform, only z varies per lane. */
template <typename S, bool PZ>
struct Pusher {
    S* fout;
    int64_t nc;
    int x, y, nx, ny, nz, px, py;
    __device__ __forceinline__ void put(int i, int z, uint32_t c, S v) const {
        int xd = x + cx(i), yd = y + cy(i);
        bool wall = false;
        if (xd < 0 || xd >= nx) { if (px) xd = xd < 0 ? xd + nx : xd - nx; else wall = true; }
        if (yd < 0 || yd >= ny) { if (py) yd = yd < 0 ? yd + ny : yd - ny; else wall = true; }
        int zd = z + cz(i);
        bool zwall = false;
        if (cz(i) != 0) {
            if (PZ) zd = zd < 0 ? zd + nz : (zd >= nz ? zd - nz : zd);
            else zwall = zd < 0 || zd >= nz;
        }
        if (wall) {
            fout[(size_t)opp(i) * nc + c] = v;
        } else {
            S* base = fout + (size_t)i * nc;
            const uint32_t rowd = (uint32_t)(xd * ny + yd) * (uint32_t)nz;
            if (zwall) fout[(size_t)opp(i) * nc + c] = v;
            else base[rowd + (uint32_t)zd] = v;
        }
    }
};

template <int PREC, bool FORCE, bool PZ>
__global__ __launch_bounds__(256) void k_collide_stream(const typename Store<PREC>::S* __restrict__ fin,
                                                        typename Store<PREC>::S* __restrict__ fout,
                                                        CollideArgs a) {
    typedef typename Store<PREC>::T T;
    typedef typename Store<PREC>::S S;
    const int row = blockIdx.x;            /* row = x*ny + y */
    const int x = row / a.ny;
    const int y = row - x * a.ny;
    const int64_t nc = a.nc;
    const T omega = T(a.omega);
    const T acc[3] = {T(a.acc[0]), T(a.acc[1]), T(a.acc[2])};
    Pusher<S, PZ> push{fout, nc, x, y, a.nx, a.ny, a.nz, a.px, a.py};

    for (int z = threadIdx.x; z < a.nz; z += blockDim.x) {
        const uint32_t c = (uint32_t)row * (uint32_t)a.nz + (uint32_t)z;
        const int s = a.slot[c];
        T fp[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) fp[i] = Store<PREC>::load(fin + (size_t)i * nc + c);
        BgkPre<T, FORCE> pre;
        pre.init(fp, omega, acc);
        if (s < 0) {
#pragma unroll
            for (int i = 0; i < Q; ++i) push.put(i, z, c, (S)pre.out(i, fp[i]));
            continue;
        }
        /* cell with a coupling record: IBcompositeDynamics::collide */
        const double sf = a.rsf[s];
        const bool full = sf > SOLFRAC_MAX;
        const T B = T((sf * a.ooo) / ((1.0 - sf) + a.ooo));
        const T us[3] = {T(a.rus[s]), T(a.rus[a.rcap + s]), T(a.rus[2 * a.rcap + s])};
        /* moments of the post-BGK cell (the BGK outputs are recomputed below,
         * bit-identically, instead of being kept) */
        T rhoB, jB[3];
        if (full) {
            rhoB = pre.rhoBar;
            moments<T>(fp, rhoB, jB);
        } else {
            rhoB = T(0); jB[0] = jB[1] = jB[2] = T(0);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const T fb = pre.out(i, fp[i]);
                rhoB = rhoB + fb;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (cc(i, d) == 1) jB[d] = jB[d] + fb;
                    else if (cc(i, d) == -1) jB[d] = jB[d] - fb;
                }
            }
        }
        const T invRho = T(1) / (rhoB + T(1));
        const T jSqr = norm_sqr<T>(jB);
        const T p = T(1.5) * invRho;
        const T onePlus = T(1) + rhoB;
        T uP[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) uP[d] = us[d] * onePlus;
        const T uPSqr = norm_sqr<T>(uP);
        const T oneMinB = T(1) - B;
        T F[3] = {T(0), T(0), T(0)};
        {
            const T bb0 = equilibrium<T>(0, rhoB, p, jB, jSqr) - equilibrium<T>(0, rhoB, p, uP, uPSqr);
            T o0;
            if (full) o0 = fp[0] + bb0;
            else o0 = (fp[0] + oneMinB * (pre.out(0, fp[0]) - fp[0])) + B * bb0;
            push.put(0, z, c, (S)o0);
        }
#pragma unroll
        for (int i = 1; i <= 9; ++i) {
            const int o = i + 9;
            const T Ei = equilibrium<T>(i, rhoB, p, jB, jSqr);
            const T Eo = equilibrium<T>(o, rhoB, p, jB, jSqr);
            const T Si = equilibrium<T>(i, rhoB, p, uP, uPSqr);
            const T So = equilibrium<T>(o, rhoB, p, uP, uPSqr);
            T bb, bbo, oi, oo;
            if (full) {
                bb = (fp[o] - Eo) - (fp[i] - Si);
                bbo = (fp[i] - Ei) - (fp[o] - So);
                oi = fp[i] + bb;
                oo = fp[o] + bbo;
            } else {
                bb = B * ((fp[o] - Eo) - (fp[i] - Si));
                bbo = B * ((fp[i] - Ei) - (fp[o] - So));
                oi = (fp[i] + oneMinB * (pre.out(i, fp[i]) - fp[i])) + bb;
                oo = (fp[o] + oneMinB * (pre.out(o, fp[o]) - fp[o])) + bbo;
            }
            push.put(i, z, c, (S)oi);
            push.put(o, z, c, (S)oo);
            const T diff = bb - bbo;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (cc(i, d) == 1) F[d] = F[d] - diff;
                else if (cc(i, d) == -1) F[d] = F[d] + diff;
            }
        }
        a.rF[s] = double(F[0]);
        a.rF[a.rcap + s] = double(F[1]);
        a.rF[2 * a.rcap + s] = double(F[2]);
    }
}

template <int PREC>
__global__ __launch_bounds__(256) void k_init_eq(typename Store<PREC>::S* __restrict__ f,
                                                 const double* __restrict__ rho,
                                                 const double* __restrict__ u, int64_t nc) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        const double r = rho ? rho[c] : 1.0;
        const double rhoBar = r - 1.0;
        double j[3];
        for (int d = 0; d < 3; ++d) j[d] = u ? r * u[d * nc + c] : 0.0;
        const double invRho = 1.0 / (rhoBar + 1.0);
        const double jSqr = norm_sqr<double>(j);
        const double p = 1.5 * invRho;
#pragma unroll
        for (int i = 0; i < Q; ++i)
            Store<PREC>::store(f + (int64_t)i * nc + c, (typename Store<PREC>::T)equilibrium<double>(i, rhoBar, p, j, jSqr));
    }
}

template <int PREC>
__global__ __launch_bounds__(256) void k_pop_convert(typename Store<PREC>::S* __restrict__ f,
                                                     double* __restrict__ f64, int64_t total,
                                                     int to_store) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        if (to_store) Store<PREC>::store(f + k, (typename Store<PREC>::T)f64[k]);
        else f64[k] = (double)Store<PREC>::load(f + k);
    }
}

/* FP64 global mass / momentum / health.  Stage 1: one FP64 partial per
 * block (fixed tree); stage 2: one block sums the partials in index order. */
template <int PREC>
__global__ __launch_bounds__(256) void k_cons_partial(const typename Store<PREC>::S* __restrict__ f,
                                                      int64_t nc, double* __restrict__ part) {
    __shared__ double sh[5][256];
    double m = 0.0, jx = 0.0, jy = 0.0, jz = 0.0, umax = 0.0, nonfin = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        double fc[Q];
        bool bad = false;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            fc[i] = (double)Store<PREC>::load(f + (int64_t)i * nc + c);
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
                                                         int64_t cap, int64_t nc) {
    const int64_t n = count[0] < cap ? count[0] : cap;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = dv_cell[k];
        const double u[3] = {dv_us[3 * k], dv_us[3 * k + 1], dv_us[3 * k + 2]};
        /* IBcompositeDynamics::defineVelocity, src/ibCompositeDynamics3D.hh:118-127:
         * equilibrium at rhoBar = 0 with j = u. */
        const double uSqr = norm_sqr<double>(u);
        const double p = 1.5 * (1.0 / (0.0 + 1.0));
#pragma unroll
        for (int i = 0; i < Q; ++i)
            Store<PREC>::store(f + (int64_t)i * nc + c, (typename Store<PREC>::T)equilibrium<double>(i, 0.0, p, u, uSqr));
    }
}

static CollideArgs make_args(lbdem_psm* h) {
    CollideArgs a;
    a.nx = h->g.n[0]; a.ny = h->g.n[1]; a.nz = h->g.n[2];
    a.px = h->g.periodic[0]; a.py = h->g.periodic[1]; a.pz = h->g.periodic[2];
    a.nc = h->g.nc;
    a.omega = h->omega; a.ooo = h->ooo;
    for (int d = 0; d < 3; ++d) a.acc[d] = h->desc.body_accel[d];
    const Records& r = h->rec[h->cur_rec];
    a.slot = h->slot; a.rsf = r.sf; a.rus = r.us; a.rF = r.F; a.rcap = h->rec_cap;
    return a;
}

template <int PREC>
static void launch_cs(lbdem_psm* h, const CollideArgs& a) {
    typedef typename Store<PREC>::S S;
    const S* fin = (const S*)h->f[h->cur_f];
    S* fout = (S*)h->f[1 - h->cur_f];
    dim3 grid((unsigned)(a.nx * a.ny)), block(256);
    if (h->has_force) {
        if (a.pz) hipLaunchKernelGGL((k_collide_stream<PREC, true, true>), grid, block, 0, h->stream, fin, fout, a);
        else hipLaunchKernelGGL((k_collide_stream<PREC, true, false>), grid, block, 0, h->stream, fin, fout, a);
    } else {
        if (a.pz) hipLaunchKernelGGL((k_collide_stream<PREC, false, true>), grid, block, 0, h->stream, fin, fout, a);
        else hipLaunchKernelGGL((k_collide_stream<PREC, false, false>), grid, block, 0, h->stream, fin, fout, a);
    }
}

int launch_collide_stream(lbdem_psm* h) {
    CollideArgs a = make_args(h);
    hipEvent_t ev = nullptr;
    prof_begin(h, 0, &ev);
    switch (h->precision) {
        case LBDEM_PSM_FP64: launch_cs<LBDEM_PSM_FP64>(h, a); break;
        case LBDEM_PSM_FP32_DEV: launch_cs<LBDEM_PSM_FP32_DEV>(h, a); break;
        default: launch_cs<LBDEM_PSM_FP16_STORE>(h, a); break;
    }
    PSM_HIP(h, hipGetLastError());
    prof_end(h, 0, ev);
    h->cur_f = 1 - h->cur_f;
    return LBDEM_PSM_OK;
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
            hipLaunchKernelGGL(k_init_eq<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (double*)h->f[h->cur_f], d_rho, d_u, nc);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_init_eq<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (float*)h->f[h->cur_f], d_rho, d_u, nc);
            break;
        default:
            hipLaunchKernelGGL(k_init_eq<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream, (_Float16*)h->f[h->cur_f], d_rho, d_u, nc);
            break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int launch_pop_convert(lbdem_psm* h, double* d_f64, bool to_store) {
    const int64_t total = (int64_t)Q * h->g.nc;
    dim3 grid(grid_for(total)), block(256);
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_pop_convert<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (double*)h->f[h->cur_f], d_f64, total, (int)to_store);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_pop_convert<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (float*)h->f[h->cur_f], d_f64, total, (int)to_store);
            break;
        default:
            hipLaunchKernelGGL(k_pop_convert<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream, (_Float16*)h->f[h->cur_f], d_f64, total, (int)to_store);
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
            hipLaunchKernelGGL(k_cons_partial<LBDEM_PSM_FP64>, dim3(nblocks), dim3(256), 0, h->stream, (const double*)h->f[h->cur_f], nc, part);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_cons_partial<LBDEM_PSM_FP32_DEV>, dim3(nblocks), dim3(256), 0, h->stream, (const float*)h->f[h->cur_f], nc, part);
            break;
        default:
            hipLaunchKernelGGL(k_cons_partial<LBDEM_PSM_FP16_STORE>, dim3(nblocks), dim3(256), 0, h->stream, (const _Float16*)h->f[h->cur_f], nc, part);
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

int launch_define_velocity(lbdem_psm* h) {
    if (!h->dv_cell) return LBDEM_PSM_OK;
    const int64_t nc = h->g.nc;
    dim3 grid(1024), block(256);
    switch (h->precision) {
        case LBDEM_PSM_FP64:
            hipLaunchKernelGGL(k_define_velocity<LBDEM_PSM_FP64>, grid, block, 0, h->stream, (double*)h->f[h->cur_f], h->dv_cell, h->dv_us, h->counters + 5, h->rec_cap, nc);
            break;
        case LBDEM_PSM_FP32_DEV:
            hipLaunchKernelGGL(k_define_velocity<LBDEM_PSM_FP32_DEV>, grid, block, 0, h->stream, (float*)h->f[h->cur_f], h->dv_cell, h->dv_us, h->counters + 5, h->rec_cap, nc);
            break;
        default:
            hipLaunchKernelGGL(k_define_velocity<LBDEM_PSM_FP16_STORE>, grid, block, 0, h->stream, (_Float16*)h->f[h->cur_f], h->dv_cell, h->dv_us, h->counters + 5, h->rec_cap, nc);
            break;
    }
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

}  // namespace psm

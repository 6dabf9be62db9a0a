/*
 * Per-particle hydrodynamic force and torque on gfx950.
 *
 * Reference: SumForceTorque3D::process (src/ibProcessors3D.hh:202-275),
 * which scans every cell and adds the per-cell force F and d x F
 * (d = cell - x[map(id)] with the minimum image over the block size,
 * :235-251) into the particle's slot.
 *
 * Here the sum runs over the bins that hold candidates, in two
 * deterministic stages:
 *  1. per bin (one wavefront): each lane holds eight cells of the bin in
 *     registers; for each of the bin's candidate entries the wavefront sums
 *     the records that candidate set this step, with a fixed reduce-scatter
 *     tree in FP64.  One 6-vector partial per candidate entry.
 *  2. per particle: its candidate entries are added in the fixed
 *     (image rank, bin) order.
 * Records whose owner set no candidate of their bin this step (stale
 * records the reference also keeps) are added afterwards with FP64 atomics.
 */
#include "psm_common.h"

namespace psm {

struct ForceArgs {
    int n[3];
    int nb[3];
    const int32_t* bin_cand_start;
    const int32_t* bin_cand_count;
    const uint32_t* key;
    const uint8_t* sflag;
    Records rec;
    int64_t rcap;
    int f32;               /* per-cell forces stored as float (FP32/FP16 storage) */
    const int32_t* powner;
    const double* xs;      /* positions for the torque arms */
    double* partial;
    int64_t ent_cap;
    int64_t nbins;
};

/* Sum of v[0..7] over the 64 lanes of a wavefront, reduce-scatter style:
 * three halving exchanges (lane bits 5, 4, 3) leave each lane one value,
 * index m = 4 b5 + 2 b4 + b3 of the lane, summed over the eight lanes that
 * share those bits; three butterflies over bits 2, 1, 0 finish it.  Ten
 * shuffles of a double in place of 36 for six separate trees, and the order
 * of every addition is fixed (a + b and b + a round alike), so the sum is
 * deterministic. */
__device__ __forceinline__ double wave_sum8_scatter(const double (&v)[8], int lane) {
    double w4[4], w2[2];
    const bool u5 = (lane & 32) != 0, u4 = (lane & 16) != 0, u3 = (lane & 8) != 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double keep = u5 ? v[4 + m] : v[m];
        const double send = u5 ? v[m] : v[4 + m];
        w4[m] = keep + __shfl_xor(send, 32, 64);
    }
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const double keep = u4 ? w4[2 + m] : w4[m];
        const double send = u4 ? w4[m] : w4[2 + m];
        w2[m] = keep + __shfl_xor(send, 16, 64);
    }
    const double keep = u3 ? w2[1] : w2[0];
    const double send = u3 ? w2[0] : w2[1];
    double w = keep + __shfl_xor(send, 8, 64);
    w += __shfl_xor(w, 4, 64);
    w += __shfl_xor(w, 2, 64);
    w += __shfl_xor(w, 1, 64);
    return w;
}

/* One wavefront per bin, four bins per workgroup.  Lane (y, z) of the bin
 * holds its eight cells (x, y, z), x = 0..7, in registers: the liveness
 * flag, the candidate entry that set the record and the force.  The
 * particle positions of the bin's candidate entries are fetched one entry
 * per lane, overlapping the cell loads.  Then, entry by entry, each lane
 * adds F and d x F of its cells set by that entry (in x order) and the
 * wavefront reduces the six sums with wave_sum8_scatter. */
template <typename FT>
__global__ __launch_bounds__(256) void k_force_bin(ForceArgs a) {
    const int lane = (int)(threadIdx.x & 63);
    /* neighbouring bins (z fastest) share the lines of the cell-ordered
     * arrays: keep them in one workgroup and on one XCD's L2 */
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t b = (int64_t)xcd_block((int)blockIdx.x, (int)gridDim.x) * 4 + wv;
    if (b >= a.nbins) return;
    const int ncand = a.bin_cand_count[b];
    if (ncand == 0) return;
    const int cstart = a.bin_cand_start[b];
    if ((int64_t)cstart + ncand > a.ent_cap) return;
    const int bi = (int)b;
    const int bz = bi % a.nb[2];
    const int by = (bi / a.nb[2]) % a.nb[1];
    const int bx = bi / (a.nb[2] * a.nb[1]);
    const int Y = by * BIN + (lane >> 3), Z = bz * BIN + (lane & 7);
    const bool yz = Y < a.n[1] && Z < a.n[2];
    /* Three rounds of loads, each issued at once: the eight liveness
     * flags and the bin's candidate keys (one entry per lane); the records
     * of the live cells and the entries' owners; the owners' positions.
     * The empty asm statements pin the loaded values in registers after
     * the last round, so no conversion or test waits inside a round. */
    const int64_t plane = (int64_t)a.n[1] * a.n[2];
    const int64_t lo = yz ? (int64_t)Y * a.n[2] + Z : 0;
    const int tl = lane < ncand ? lane : 0;
    const uint32_t kk = a.key[cstart + tl];
    int fl[BIN];
#pragma unroll
    for (int x = 0; x < BIN; ++x) {
        const int X = bx * BIN + x;
        const int64_t xb = (int64_t)(X < a.n[0] ? X : a.n[0] - 1) * plane;
        fl[x] = a.sflag[xb + lo];
        if (!yz || X >= a.n[0]) fl[x] = 0;
    }
    const int po = a.powner[(int)(kk & QMASK)];
    const FT* __restrict__ F = (const FT*)a.rec.F;
    int ce[BIN];
    FT fx[BIN], fy[BIN], fz[BIN];
#pragma unroll
    for (int x = 0; x < BIN; ++x) {
        ce[x] = -1;
        fx[x] = fy[x] = fz[x] = (FT)0;
        if (fl[x]) {
            const int64_t r = (int64_t)(bx * BIN + x) * plane + lo;
            ce[x] = a.rec.cand[r];
            fx[x] = F[r];
            fy[x] = F[a.rcap + r];
            fz[x] = F[2 * a.rcap + r];
        }
    }
    double px = a.xs[3 * po], py = a.xs[3 * po + 1], pz = a.xs[3 * po + 2];
#pragma unroll
    for (int x = 0; x < BIN; ++x) asm volatile("" : "+v"(ce[x]), "+v"(fx[x]), "+v"(fy[x]), "+v"(fz[x]));
    asm volatile("" : "+v"(px), "+v"(py), "+v"(pz));
    const double dy0 = (double)Y, dz0 = (double)Z;
    for (int t = 0; t < ncand; ++t) {
        const int e = cstart + t;
        /* an entry whose candidate set no cell of this bin (its box only
         * grazes the bin) contributes zeros, without the reduction */
        bool any = false;
#pragma unroll
        for (int x = 0; x < BIN; ++x) any |= ce[x] == e;
        if (__ballot(any) == 0ull) {
            if (lane < 6) a.partial[6 * (int64_t)e + lane] = 0.0;
            continue;
        }
        double xo, yo, zo;
        if (t < 64) {
            xo = readlane_d(px, t); yo = readlane_d(py, t); zo = readlane_d(pz, t);
        } else {
            const int o = a.powner[(int)(a.key[e] & QMASK)];
            xo = a.xs[3 * o]; yo = a.xs[3 * o + 1]; zo = a.xs[3 * o + 2];
        }
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const double dy = min_image(dy0 - yo, a.n[1]);
        const double dz = min_image(dz0 - zo, a.n[2]);
#pragma unroll
        for (int x = 0; x < BIN; ++x) {
            if (ce[x] == e) {
                /* converted here, not hoisted out of the entry loop, so the
                 * cells stay in registers at their stored width */
                FT h0 = fx[x], h1 = fy[x], h2 = fz[x];
                int X = bx * BIN + x;
                asm volatile("" : "+v"(h0), "+v"(h1), "+v"(h2), "+s"(X));
                const double dx = min_image((double)X - xo, a.n[0]);
                const double gx = (double)h0, gy = (double)h1, gz = (double)h2;
                acc[0] += gx;
                acc[1] += gy;
                acc[2] += gz;
                acc[3] += dy * gz - dz * gy;
                acc[4] += -dx * gz + dz * gx;
                acc[5] += dx * gy - dy * gx;
            }
        }
        const double s = wave_sum8_scatter(acc, lane);
        const int m = lane >> 3;
        if ((lane & 7) == 0 && m < 6) a.partial[6 * (int64_t)e + m] = s;
    }
}

/* One wavefront per particle: lane-strided partial sums of its candidate
 * entries, then the fixed reduce-scatter tree of wave_sum8_scatter. */
__global__ __launch_bounds__(256) void k_force_particle(int64_t np, const int32_t* __restrict__ p_ent_start,
                                                        const int32_t* __restrict__ p_nent,
                                                        const int32_t* __restrict__ owner_ent,
                                                        const double* __restrict__ partial,
                                                        double* __restrict__ pft, int64_t ent_cap) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = (int)__lane_id();
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (q < np) {
        const int64_t s = p_ent_start[q];
        const int n = p_nent[q];
        for (int k = lane; k < n; k += 64) {
            if (s + k >= ent_cap) break;
            const int e = owner_ent[s + k];
            if (e < 0) continue;
#pragma unroll
            for (int m = 0; m < 6; ++m) acc[m] += partial[6 * (int64_t)e + m];
        }
    }
    const double v = wave_sum8_scatter(acc, lane);
    const int m = lane >> 3;
    if (q < np && (lane & 7) == 0 && m < 6) pft[6 * q + m] = v;
}

__global__ __launch_bounds__(256) void k_force_orphans(ForceArgs a, const int32_t* __restrict__ orphan,
                                                       const int64_t* __restrict__ counters, int64_t ocap,
                                                       const int32_t* __restrict__ id_map,
                                                       int64_t id_map_size, double* __restrict__ pft) {
    const int64_t n = counters[2] < ocap ? counters[2] : ocap;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = orphan[k];
        const int id = a.rec.id[r];
        if (id < 1 || id >= id_map_size) continue;
        const int o = id_map[id];
        if (o < 0) continue;
        int X, Y, Z;
        rec_cell(r, a.n[1], a.n[2], X, Y, Z);
        const double dx = min_image((double)X - a.xs[3 * o], a.n[0]);
        const double dy = min_image((double)Y - a.xs[3 * o + 1], a.n[1]);
        const double dz = min_image((double)Z - a.xs[3 * o + 2], a.n[2]);
        double fx, fy, fz;
        if (a.f32) {
            const float* F32 = (const float*)a.rec.F;
            fx = F32[r]; fy = F32[a.rcap + r]; fz = F32[2 * a.rcap + r];
        } else {
            fx = a.rec.F[r]; fy = a.rec.F[a.rcap + r]; fz = a.rec.F[2 * a.rcap + r];
        }
        atomicAdd(&pft[6 * o + 0], fx);
        atomicAdd(&pft[6 * o + 1], fy);
        atomicAdd(&pft[6 * o + 2], fz);
        atomicAdd(&pft[6 * o + 3], dy * fz - dz * fy);
        atomicAdd(&pft[6 * o + 4], -dx * fz + dz * fx);
        atomicAdd(&pft[6 * o + 5], dx * fy - dy * fx);
    }
}

int run_force_sum(lbdem_psm* h, const double* xs, double* out) {
    double* pft = out ? out : h->pft;
    hipEvent_t ev = nullptr;
    prof_begin(h, 2, &ev);
    ForceArgs a;
    for (int d = 0; d < 3; ++d) { a.n[d] = h->g.n[d]; a.nb[d] = h->g.nb[d]; }
    a.bin_cand_start = h->bin_cand_start;
    a.bin_cand_count = h->bin_cand_count;
    a.key = h->cand_key;
    a.sflag = h->sflag;
    a.rec = h->rec;
    a.rcap = h->rec_cap;
    a.f32 = h->precision == LBDEM_PSM_FP64 ? 0 : 1;
    a.powner = h->powner;
    a.xs = xs;
    a.partial = h->partial;
    a.ent_cap = h->ent_cap;
    a.nbins = h->nbins;
    const dim3 grid((unsigned)((h->nbins + 3) / 4));
    if (a.f32) hipLaunchKernelGGL(k_force_bin<float>, grid, dim3(256), 0, h->stream, a);
    else hipLaunchKernelGGL(k_force_bin<double>, grid, dim3(256), 0, h->stream, a);
    if (h->np > 0) {
        hipLaunchKernelGGL(k_force_particle, dim3((unsigned)((h->np + 3) / 4)), dim3(256), 0, h->stream,
                           h->np, h->p_ent_start, h->p_nent, h->owner_ent, h->partial, pft, h->ent_cap);
        hipLaunchKernelGGL(k_force_orphans, dim3(64), dim3(256), 0, h->stream, a, h->orphan, h->counters,
                           h->orphan_cap, h->id_map, h->id_map_size, pft);
    }
    PSM_HIP(h, hipGetLastError());
    prof_end(h, 2, ev);
    return LBDEM_PSM_OK;
}

}  // namespace psm

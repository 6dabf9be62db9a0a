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
 *  1. per bin (one workgroup): the bin's records are staged in LDS; each
 *     wavefront takes some of the bin's candidate entries and sums, for
 *     each, the records that candidate set this step, with a fixed 64-lane
 *     shuffle tree in FP64.  One 6-vector partial per candidate entry.
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
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_force_bin(ForceArgs a) {
    __shared__ double sF[BIN_CELLS][3];
    __shared__ int sE[BIN_CELLS];
    /* neighbouring bins (z fastest) share the lines of the cell-ordered
     * arrays: keep them on one XCD's L2 */
    const int64_t b = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int ncand = a.bin_cand_count[b];
    if (ncand == 0) return;
    const int cstart = a.bin_cand_start[b];
    if ((int64_t)cstart + ncand > a.ent_cap) return;
    const int bz = (int)(b % a.nb[2]);
    const int by = (int)((b / a.nb[2]) % a.nb[1]);
    const int bx = (int)(b / ((int64_t)a.nb[2] * a.nb[1]));
    for (int l = threadIdx.x; l < BIN_CELLS; l += blockDim.x) {
        const int X = bx * BIN + (l >> 6), Y = by * BIN + ((l >> 3) & 7), Z = bz * BIN + (l & 7);
        const bool valid = X < a.n[0] && Y < a.n[1] && Z < a.n[2];
        const int64_t r = valid ? rec_index(X, Y, Z, a.n[1], a.n[2]) : 0;
        int e = -1;
        if (valid && a.sflag[r]) {
            e = a.rec.cand[r];
            if (a.f32) {
                const float* F32 = (const float*)a.rec.F;
                sF[l][0] = (double)F32[r];
                sF[l][1] = (double)F32[a.rcap + r];
                sF[l][2] = (double)F32[2 * a.rcap + r];
            } else {
                sF[l][0] = a.rec.F[r];
                sF[l][1] = a.rec.F[a.rcap + r];
                sF[l][2] = a.rec.F[2 * a.rcap + r];
            }
        }
        sE[l] = e;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = wave; j < ncand; j += 4) {
        const int e = cstart + j;
        const int q = (int)(a.key[e] & QMASK);
        const int o = a.powner[q];
        const double xo = a.xs[3 * o], yo = a.xs[3 * o + 1], zo = a.xs[3 * o + 2];
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int l = lane; l < BIN_CELLS; l += 64) {
            /* an x-slice of the bin with no cell of this entry is skipped by
             * the whole wavefront */
            const bool mine = sE[l] == e;
            if (__ballot(mine) == 0ull) continue;
            if (!mine) continue;
            const double dx = min_image((double)(bx * BIN + (l >> 6)) - xo, a.n[0]);
            const double dy = min_image((double)(by * BIN + ((l >> 3) & 7)) - yo, a.n[1]);
            const double dz = min_image((double)(bz * BIN + (l & 7)) - zo, a.n[2]);
            const double fx = sF[l][0], fy = sF[l][1], fz = sF[l][2];
            acc[0] += fx;
            acc[1] += fy;
            acc[2] += fz;
            acc[3] += dy * fz - dz * fy;
            acc[4] += -dx * fz + dz * fx;
            acc[5] += dx * fy - dy * fx;
        }
#pragma unroll
        for (int m = 0; m < 6; ++m) acc[m] = wave_sum(acc[m]);
        if (lane == 0) {
#pragma unroll
            for (int m = 0; m < 6; ++m) a.partial[6 * (int64_t)e + m] = acc[m];
        }
    }
}

/* One wavefront per particle: lane-strided partial sums of its candidate
 * entries, then the fixed 64-lane shuffle tree. */
__global__ __launch_bounds__(256) void k_force_particle(int64_t np, const int32_t* __restrict__ p_ent_start,
                                                        const int32_t* __restrict__ p_nent,
                                                        const int32_t* __restrict__ owner_ent,
                                                        const double* __restrict__ partial,
                                                        double* __restrict__ pft, int64_t ent_cap) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = (int)__lane_id();
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
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
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[m] = wave_sum(acc[m]);
    if (q < np && lane == 0) {
#pragma unroll
        for (int m = 0; m < 6; ++m) pft[6 * q + m] = acc[m];
    }
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

int run_force_sum(lbdem_psm* h) {
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
    a.xs = h->psum_x;
    a.partial = h->partial;
    a.ent_cap = h->ent_cap;
    hipLaunchKernelGGL(k_force_bin, dim3((unsigned)h->nbins), dim3(256), 0, h->stream, a);
    if (h->np > 0) {
        hipLaunchKernelGGL(k_force_particle, dim3((unsigned)((h->np + 3) / 4)), dim3(256), 0, h->stream,
                           h->np, h->p_ent_start, h->p_nent, h->owner_ent, h->partial, h->pft, h->ent_cap);
        hipLaunchKernelGGL(k_force_orphans, dim3(64), dim3(256), 0, h->stream, a, h->orphan, h->counters,
                           h->orphan_cap, h->id_map, h->id_map_size, h->pft);
    }
    PSM_HIP(h, hipGetLastError());
    prof_end(h, 2, ev);
    return LBDEM_PSM_OK;
}

}  // namespace psm

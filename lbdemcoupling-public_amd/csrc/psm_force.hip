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
 * records the reference also keeps) are grouped by owner and added
 * afterwards, each particle's in cell order (k_force_orphans).
 */
#include "psm_common.h"

#include <hipcub/hipcub.hpp>

namespace psm {

struct ForceArgs {
    int n[3];
    int nb[3];
    FastDiv div_nb2, div_nb12;   /* bin -> (bx, by, bz) */
    const int32_t* bin_cand_start;
    const int32_t* bin_cand_count;
    const uint32_t* key;
    const uint8_t* sflag;
    Records rec;
    int f32;               /* per-cell forces stored as float (FP32/FP16 storage) */
    const int32_t* powner;
    const double* xs;      /* positions for the torque arms */
    double* partial;
    int64_t ent_cap;
    int64_t nbins;
};

/* One wavefront per bin (FORCE_WAVES bins per workgroup).  Lane (y, z) of the bin
 * holds its eight cells (x, y, z), x = 0..7, in registers: the liveness
 * flag, the candidate entry that set the record and the force.  The
 * particle positions of the bin's candidate entries are fetched one entry
 * per lane, overlapping the cell loads.  Then, entry by entry, each lane
 * adds F and d x F of its cells set by that entry (in x order) and the
 * wavefront reduces the six sums with wave_sum8_scatter. */
/* bins (one wavefront each) per workgroup: one, so that a bin's wavefront
 * is scheduled on its own (measured faster than 2 or 4 per workgroup) */
#ifndef FORCE_WAVES
#define FORCE_WAVES 1
#endif

/* The three force components of one record slot (its fourth element is
 * padding): one 12-byte load in float, a 16-byte and an 8-byte one in
 * double. */
__device__ __forceinline__ void load_force3(const float* p, float& x, float& y, float& z) {
    typedef float f3 __attribute__((ext_vector_type(3)));
    const f3 v = *(const f3*)p;
    x = v.x; y = v.y; z = v.z;
}
__device__ __forceinline__ void load_force3(const double* p, double& x, double& y, double& z) {
    const double2 v = *(const double2*)p;
    x = v.x; y = v.y; z = p[2];
}

template <typename FT>
__global__ __launch_bounds__(64 * FORCE_WAVES) void k_force_bin(ForceArgs a) {
    const int lane = (int)(threadIdx.x & 63);
    /* neighbouring bins (z fastest) share the lines of the cell-ordered
     * arrays: keep them in one workgroup and on one XCD's L2 */
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t b = (int64_t)xcd_block((int)blockIdx.x, (int)gridDim.x) * FORCE_WAVES + wv;
    if (b >= a.nbins) return;
    /* both ends of the candidate range in one round of loads (the empty asm
     * keeps the second load from sinking below the test of the first) */
    const int ncand = a.bin_cand_count[b];
    const int cstart = a.bin_cand_start[b];
    asm volatile("" ::"s"(ncand), "s"(cstart));
    if (ncand == 0) return;
    if ((int64_t)cstart + ncand > a.ent_cap) return;
    const int bi = (int)b;
    const int bxy = (int)fdiv((uint32_t)bi, a.div_nb2);
    const int bz = bi - bxy * a.nb[2];
    const int bx = (int)fdiv((uint32_t)bi, a.div_nb12);
    const int by = bxy - bx * a.nb[1];
    const int Y = by * BIN + (lane >> 3), Z = bz * BIN + (lane & 7);
    const bool yz = Y < a.n[1] && Z < a.n[2];
    /* Three rounds of loads, each issued at once: the eight liveness
     * flags and the bin's candidate keys (one entry per lane); the records
     * of the live cells and the entries' owners; the owners' positions.
     * The scheduling barriers keep each round's address arithmetic behind
     * the previous round's loads, and the empty asm statements pin the
     * loaded values in registers after the last round, so no conversion or
     * test waits inside a round. */
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
    __builtin_amdgcn_sched_barrier(0);
    const int po = a.powner[(int)(kk & QMASK)];
    const FT* __restrict__ F = (const FT*)a.rec.F;
    int ce[BIN];
    FT fx[BIN], fy[BIN], fz[BIN];
    /* Every lane loads, without a branch: a cell without a live record reads
     * slot 0 (one line shared by the wavefront) and drops it below.  Loads
     * under a per-cell branch made the compiler wait for the first cell's
     * record before it issued the others (the branch's merge copies it). */
#pragma unroll
    for (int x = 0; x < BIN; ++x) {
        const int64_t r = fl[x] ? (int64_t)(bx * BIN + x) * plane + lo : 0;
        ce[x] = a.rec.cand[r];
        load_force3(F + 4 * r, fx[x], fy[x], fz[x]);
    }
    double px = a.xs[3 * po], py = a.xs[3 * po + 1], pz = a.xs[3 * po + 2];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int x = 0; x < BIN; ++x) {
        if (!fl[x]) {
            ce[x] = -1;
            fx[x] = fy[x] = fz[x] = (FT)0;
        }
    }
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
        /* six scalars, packed for the reduction only after the cell loop: an
         * eight-element array here was re-zeroed through a chain of ~26 register
         * moves per entry */
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0;
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
                s0 += gx;
                s1 += gy;
                s2 += gz;
                s3 += dy * gz - dz * gy;
                s4 += -dx * gz + dz * gx;
                s5 += dx * gy - dy * gx;
            }
        }
        const double acc[8] = {s0, s1, s2, s3, s4, s5, 0.0, 0.0};
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

/* ---- stale records, summed per owner in cell order ----
 * Records whose owner set no candidate of their bin this step (kept from
 * earlier steps, or all records after set_cell_records / load_checkpoint)
 * come as an unordered list of cells.  They are grouped by owner particle
 * (count, scan, place) and each particle's cells are put in increasing cell
 * order (each cell's rank among its particle's cells), the order of the
 * reference's lexicographic scan (src/ibProcessors3D.hh:210-260); the sum
 * over them then has a fixed order whatever order the list was built in. */

__device__ __forceinline__ int orphan_owner(int64_t r, const int32_t* __restrict__ rid,
                                            const int32_t* __restrict__ id_map, int64_t id_map_size) {
    const int id = rid[r];
    if (id < 1 || id >= id_map_size) return -1;
    return id_map[id];
}

__global__ __launch_bounds__(256) void k_orph_count(const int32_t* __restrict__ orphan,
                                                    const int64_t* __restrict__ counters, int64_t ocap,
                                                    const int32_t* __restrict__ rid,
                                                    const int32_t* __restrict__ id_map, int64_t id_map_size,
                                                    int32_t* __restrict__ cnt) {
    const int64_t n = counters[2] < ocap ? counters[2] : ocap;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int o = orphan_owner(orphan[k], rid, id_map, id_map_size);
        if (o >= 0) atomicAdd(&cnt[o], 1);
    }
}

__global__ __launch_bounds__(256) void k_orph_place(const int32_t* __restrict__ orphan,
                                                    const int64_t* __restrict__ counters, int64_t ocap,
                                                    const int32_t* __restrict__ rid,
                                                    const int32_t* __restrict__ id_map, int64_t id_map_size,
                                                    const int32_t* __restrict__ start, int32_t* __restrict__ fill,
                                                    int32_t* __restrict__ slot) {
    const int64_t n = counters[2] < ocap ? counters[2] : ocap;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = orphan[k];
        const int o = orphan_owner(r, rid, id_map, id_map_size);
        if (o >= 0) slot[start[o] + atomicAdd(&fill[o], 1)] = r;
    }
}

/* One wavefront per particle: each of its cells goes to its rank among
 * them (the cells are distinct), so orph_sorted holds them in cell order. */
__global__ __launch_bounds__(256) void k_orph_rank(int64_t np, const int32_t* __restrict__ cnt,
                                                   const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ slot,
                                                   int32_t* __restrict__ sorted) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= np) return;
    const int m = cnt[q];
    if (m == 0) return;
    const int lane = (int)__lane_id();
    const int32_t* seg = slot + start[q];
    for (int i0 = 0; i0 < m; i0 += 64) {
        const bool mine = i0 + lane < m;
        const int32_t ka = mine ? seg[i0 + lane] : 0;
        int rank = 0;
        for (int c0 = 0; c0 < m; c0 += 64) {
            const int32_t kb = c0 + lane < m ? seg[c0 + lane] : INT32_MAX;
            const int top = m - c0 < 64 ? m - c0 : 64;
            for (int s2 = 0; s2 < top; ++s2) rank += __shfl(kb, s2, 64) < ka ? 1 : 0;
        }
        if (mine) sorted[start[q] + rank] = ka;
    }
}

int order_orphans(lbdem_psm* h) {
    if (h->np == 0) return LBDEM_PSM_OK;
    const unsigned og = 256;
    hipLaunchKernelGGL(k_orph_count, dim3(og), dim3(256), 0, h->stream, h->orphan, h->counters, h->orphan_cap,
                       h->rec.id, h->id_map, h->id_map_size, h->orph_cnt);
    size_t tb = h->scan_tmp_bytes;
    PSM_HIP(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp, tb, h->orph_cnt, h->orph_start, (int)h->np, h->stream));
    hipLaunchKernelGGL(k_orph_place, dim3(og), dim3(256), 0, h->stream, h->orphan, h->counters, h->orphan_cap,
                       h->rec.id, h->id_map, h->id_map_size, h->orph_start, h->orph_fill, h->orph_slot);
    hipLaunchKernelGGL(k_orph_rank, dim3((unsigned)((h->np + 3) / 4)), dim3(256), 0, h->stream, h->np, h->orph_cnt,
                       h->orph_start, h->orph_slot, h->orph_sorted);
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

/* One wavefront per particle: its stale records in cell order, lane-strided,
 * then the fixed reduce-scatter tree; added once to the particle's sum of
 * its candidate entries.  Clears the particle's count and fill cursor for
 * the next grouping. */
__global__ __launch_bounds__(256) void k_force_orphans(ForceArgs a, int64_t np, int32_t* __restrict__ cnt,
                                                       int32_t* __restrict__ fill,
                                                       const int32_t* __restrict__ start,
                                                       const int32_t* __restrict__ sorted,
                                                       double* __restrict__ pft) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= np) return;
    const int m = cnt[q];
    if (m == 0) return;
    const int lane = (int)__lane_id();
    const int32_t* seg = sorted + start[q];
    const double xo = a.xs[3 * q], yo = a.xs[3 * q + 1], zo = a.xs[3 * q + 2];
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = lane; i < m; i += 64) {
        const int64_t r = seg[i];
        int X, Y, Z;
        rec_cell(r, a.n[1], a.n[2], X, Y, Z);
        const double dx = min_image((double)X - xo, a.n[0]);
        const double dy = min_image((double)Y - yo, a.n[1]);
        const double dz = min_image((double)Z - zo, a.n[2]);
        double fx, fy, fz;
        if (a.f32) {
            const float* F32 = (const float*)a.rec.F;
            fx = F32[4 * r]; fy = F32[4 * r + 1]; fz = F32[4 * r + 2];
        } else {
            fx = a.rec.F[4 * r]; fy = a.rec.F[4 * r + 1]; fz = a.rec.F[4 * r + 2];
        }
        acc[0] += fx;
        acc[1] += fy;
        acc[2] += fz;
        acc[3] += dy * fz - dz * fy;
        acc[4] += -dx * fz + dz * fx;
        acc[5] += dx * fy - dy * fx;
    }
    const double s = wave_sum8_scatter(acc, lane);
    const int mm = lane >> 3;
    if ((lane & 7) == 0 && mm < 6) pft[6 * q + mm] += s;
    if (lane == 0) {
        cnt[q] = 0;
        fill[q] = 0;
    }
}

int run_force_sum(lbdem_psm* h, const double* xs, double* out) {
    double* pft = out ? out : h->pft;
    hipEvent_t ev = nullptr;
    prof_begin(h, 2, &ev);
    ForceArgs a;
    for (int d = 0; d < 3; ++d) { a.n[d] = h->g.n[d]; a.nb[d] = h->g.nb[d]; }
    a.div_nb2 = make_fastdiv((uint32_t)h->g.nb[2]);
    a.div_nb12 = make_fastdiv((uint32_t)(h->g.nb[2] * h->g.nb[1]));
    a.bin_cand_start = h->bin_cand_start;
    a.bin_cand_count = h->bin_cand_count;
    a.key = h->cand_key;
    a.sflag = h->sflag;
    a.rec = h->rec;
    a.f32 = h->precision == LBDEM_PSM_FP64 ? 0 : 1;
    a.powner = h->powner;
    a.xs = xs;
    a.partial = h->partial;
    a.ent_cap = h->ent_cap;
    a.nbins = h->nbins;
    const dim3 grid((unsigned)((h->nbins + FORCE_WAVES - 1) / FORCE_WAVES));
    if (a.f32) hipLaunchKernelGGL(k_force_bin<float>, grid, dim3(64 * FORCE_WAVES), 0, h->stream, a);
    else hipLaunchKernelGGL(k_force_bin<double>, grid, dim3(64 * FORCE_WAVES), 0, h->stream, a);
    if (h->np > 0) {
        hipLaunchKernelGGL(k_force_particle, dim3((unsigned)((h->np + 3) / 4)), dim3(256), 0, h->stream,
                           h->np, h->p_ent_start, h->p_nent, h->owner_ent, h->partial, pft, h->ent_cap);
        int rc = order_orphans(h);
        if (rc) return rc;
        hipLaunchKernelGGL(k_force_orphans, dim3((unsigned)((h->np + 3) / 4)), dim3(256), 0, h->stream, a, h->np,
                           h->orph_cnt, h->orph_fill, h->orph_start, h->orph_sorted, pft);
    }
    PSM_HIP(h, hipGetLastError());
    prof_end(h, 2, ev);
    return LBDEM_PSM_OK;
}

}  // namespace psm

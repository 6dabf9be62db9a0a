/*
 * Sphere voxelization on gfx950: SetSingleSphere3D for every particle, as
 * an ordered per-cell fold.
 *
 * Reference: src/ibProcessors3D.hh:52-184 (process, calcSolidFraction,
 * setValues, setToZero), src/ibProcessors3D.h:43-45 (bounding box),
 * src/ibDataExchangeWrappers3D.hh:62-140 (particle loop, excludeType,
 * box intersection).
 *
 * The reference visits particles one after another and updates the
 * per-cell record in place; the outcome of a cell depends on the order of
 * the particles whose bounding box holds it and on the record left by the
 * previous step.  Here every particle image is binned into 8^3-cell bins
 * (key = image rank << 27 | particle index, which is the canonical order:
 * the given particles by index, then periodic images by shift rank), each
 * bin's candidate list is sorted, and one wavefront per x-slice of a bin
 * (one lane per cell) folds them, per cell and in key order,
 * through the reference's four-case rule; the candidates' data are
 * wave-uniform and come in through scalar loads.  Each cell's record lives
 * at the cell's own index and is updated in place (the fold of a cell
 * reads and writes only that cell's record).
 *
 * The sub-cell count is evaluated in FP64 with the reference's expression
 * order and without contraction (this file is compiled with
 * -ffp-contract=off), so counts are bit-exact.
 */
#include "psm_common.h"

#include <hipcub/hipcub.hpp>

namespace psm {

#ifndef FOLD_WAVES
#define FOLD_WAVES 1
#endif
/* wavefronts (x-slices of a bin) per fold workgroup: a bin is 8 / FOLD_WAVES
 * workgroups.  One: the slices of a bin carry very different amounts of
 * counting work, and a workgroup holds its slots until its slowest wavefront
 * ends; single-wavefront workgroups took the fold from 5.0 to 4.2 ms at
 * 512x512x768 (8 per workgroup before) */
constexpr int FOLD_SPLIT = 8 / FOLD_WAVES;

struct VoxArgs {
    int n[3];
    int per[3];
    int nb[3];
    int auto_images;
    int64_t np;
    const double* px;
    const double* pr;
    const int32_t* pexcl;
};

/* Truncated, clipped bounding box of image `rank` of a particle; false when
 * the image does not exist or misses the lattice.  Shift digits: 0 none,
 * 1 -> +n_d (image of the lower-side pass), 2 -> -n_d. */
__device__ __forceinline__ bool image_box(const double* x, double r, int rank, const VoxArgs& a,
                                          int* lo, int* hi, double* xi) {
    int dig[3] = {rank % 3, (rank / 3) % 3, rank / 9};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int sh = dig[d] == 0 ? 0 : (dig[d] == 1 ? 1 : -1);
        if (sh != 0 && !a.per[d]) return false;
        xi[d] = sh == 0 ? x[d] : x[d] + (double)sh * (double)a.n[d];
        int l = (int)((xi[d] - r) - 2.0);
        int h = (int)((xi[d] + r) + 2.0);
        l = l < 0 ? 0 : l;
        h = h > a.n[d] - 1 ? a.n[d] - 1 : h;
        if (l > h) return false;
        lo[d] = l;
        hi[d] = h;
    }
    return true;
}

/* One wavefront per particle: lane r < 27 evaluates image rank r; the
 * images are then visited in rank order and their bins spread over the
 * lanes, bins of an image in (bx, by, bz) order (bz fastest). */
struct ImageRanges {
    unsigned long long mask;   /* valid images */
    int b0, c0, d0, nyr, nzr, vol;   /* of the lane's own image */
};

__device__ __forceinline__ ImageRanges lane_image(const VoxArgs& a, int64_t q, bool live) {
    ImageRanges ir;
    const int lane = (int)__lane_id();
    const int nrank = a.auto_images ? 27 : 1;
    bool ok = false;
    int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
    if (live && lane < nrank && !a.pexcl[q]) {
        const double x[3] = {a.px[3 * q], a.px[3 * q + 1], a.px[3 * q + 2]};
        double xi[3];
        ok = image_box(x, a.pr[q], lane, a, lo, hi, xi);
    }
    ir.mask = __builtin_amdgcn_ballot_w64(ok);
    ir.b0 = lo[0] / BIN; ir.c0 = lo[1] / BIN; ir.d0 = lo[2] / BIN;
    ir.nyr = ok ? hi[1] / BIN - ir.c0 + 1 : 0;
    ir.nzr = ok ? hi[2] / BIN - ir.d0 + 1 : 0;
    ir.vol = ok ? (hi[0] / BIN - ir.b0 + 1) * ir.nyr * ir.nzr : 0;
    return ir;
}

__device__ __forceinline__ int64_t image_bin(const VoxArgs& a, int b0, int c0, int d0, int nyr, int nzr, int t) {
    const int bx = b0 + t / (nyr * nzr);
    const int rem = t - (bx - b0) * nyr * nzr;
    const int by = c0 + rem / nzr;
    const int bz = d0 + rem - (by - c0) * nzr;
    return ((int64_t)bx * a.nb[1] + by) * a.nb[2] + bz;
}

__global__ __launch_bounds__(256) void k_bin_count(VoxArgs a, int32_t* __restrict__ bin_count,
                                                   int32_t* __restrict__ p_nent) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool live = q < a.np;
    const ImageRanges ir = lane_image(a, q, live);
    const int lane = (int)__lane_id();
    unsigned long long m = ir.mask;
    int total = 0;
    while (m) {
        const int r = __ffsll(m) - 1;
        m &= m - 1;
        const int b0 = __shfl(ir.b0, r), c0 = __shfl(ir.c0, r), d0 = __shfl(ir.d0, r);
        const int nyr = __shfl(ir.nyr, r), nzr = __shfl(ir.nzr, r), vol = __shfl(ir.vol, r);
        for (int t = lane; t < vol; t += 64) atomicAdd(&bin_count[image_bin(a, b0, c0, d0, nyr, nzr, t)], 1);
        total += vol;
    }
    if (live && lane == 0) p_nent[q] = total;
}

__global__ __launch_bounds__(256) void k_bin_fill(VoxArgs a, const int32_t* __restrict__ bin_start,
                                                  int32_t* __restrict__ bin_fill,
                                                  uint32_t* __restrict__ key, int32_t* __restrict__ kbin,
                                                  int64_t cap, int64_t* __restrict__ counters) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool live = q < a.np;
    const ImageRanges ir = lane_image(a, q, live);
    const int lane = (int)__lane_id();
    if (lane == 0 && ir.mask > 1ull)
        atomicAdd((unsigned long long*)&counters[4], (unsigned long long)(__popcll(ir.mask) - (ir.mask & 1ull)));
    unsigned long long m = ir.mask;
    while (m) {
        const int r = __ffsll(m) - 1;
        m &= m - 1;
        const int b0 = __shfl(ir.b0, r), c0 = __shfl(ir.c0, r), d0 = __shfl(ir.d0, r);
        const int nyr = __shfl(ir.nyr, r), nzr = __shfl(ir.nzr, r), vol = __shfl(ir.vol, r);
        for (int t = lane; t < vol; t += 64) {
            const int64_t b = image_bin(a, b0, c0, d0, nyr, nzr, t);
            const int64_t pos = (int64_t)bin_start[b] + atomicAdd(&bin_fill[b], 1);
            if (pos < cap) {
                key[pos] = ((uint32_t)r << QBITS) | (uint32_t)q;
                kbin[pos] = (int32_t)b;
            } else {
                atomicOr((unsigned long long*)&counters[1], 2ull);
            }
        }
    }
}

/* Sort each bin's candidate keys (short lists: insertion sort per bin). */
__global__ __launch_bounds__(256) void k_bin_sort(const int32_t* __restrict__ bin_start,
                                                  const int32_t* __restrict__ bin_count,
                                                  uint32_t* __restrict__ key, int64_t nbins, int64_t cap) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    const int64_t s = bin_start[b];
    const int n = bin_count[b];
    if (s + n > cap) return;
    uint32_t* k = key + s;
    for (int i = 1; i < n; ++i) {
        const uint32_t v = k[i];
        int j = i - 1;
        while (j >= 0 && k[j] > v) { k[j + 1] = k[j]; --j; }
        k[j + 1] = v;
    }
}

/* For each particle, the entry indices of all its (image, bin) candidate
 * entries in (rank, bx, by, bz) order: the fixed summation order of the
 * per-particle force reduction. */
__global__ __launch_bounds__(256) void k_owner_entries(VoxArgs a, const int32_t* __restrict__ bin_start,
                                                       const int32_t* __restrict__ bin_count,
                                                       const uint32_t* __restrict__ key,
                                                       const int32_t* __restrict__ p_ent_start,
                                                       int32_t* __restrict__ owner_ent, int64_t cap) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool live = q < a.np;
    const ImageRanges ir = lane_image(a, q, live);
    const int lane = (int)__lane_id();
    unsigned long long m = ir.mask;
    int64_t out = live ? p_ent_start[q] : 0;
    while (m) {
        const int r = __ffsll(m) - 1;
        m &= m - 1;
        const int b0 = __shfl(ir.b0, r), c0 = __shfl(ir.c0, r), d0 = __shfl(ir.d0, r);
        const int nyr = __shfl(ir.nyr, r), nzr = __shfl(ir.nzr, r), vol = __shfl(ir.vol, r);
        const uint32_t want = ((uint32_t)r << QBITS) | (uint32_t)q;
        for (int t = lane; t < vol; t += 64) {
            const int64_t b = image_bin(a, b0, c0, d0, nyr, nzr, t);
            int lo_i = bin_start[b], hi_i = bin_start[b] + bin_count[b] - 1;
            int found = -1;
            if ((int64_t)hi_i < cap) {
                while (lo_i <= hi_i) {
                    const int mid = (lo_i + hi_i) >> 1;
                    const uint32_t v = key[mid];
                    if (v == want) { found = mid; break; }
                    if (v < want) lo_i = mid + 1;
                    else hi_i = mid - 1;
                }
            }
            if (out + t < cap) owner_ent[out + t] = found;
        }
        out += vol;
    }
}

/* ---- the fold ---- */

/* Candidate data in key order, one entry per thread: the image box and the
 * squared radii used by the sub-cell test. */
/* Squared radii of the exact no-count band of coop_solid_fraction: every
 * sub-cell sample lies within e of the cell centre, so a centre within
 * r - e - tol (beyond r + e + tol) of the sphere centre has all samples
 * inside (outside); tol is far above the sample test's rounding. */
__device__ __forceinline__ void exact_band(double r, double e, double& ri2, double& ro2) {
    const double tol = 1e-9 * (1.0 + r);
    const double ri = r - e - tol, ro = r + e + tol;
    ri2 = ri > 0.0 ? ri * ri : -1.0;
    ro2 = ro * ro;
}

__global__ __launch_bounds__(256) void k_cand_gather(VoxArgs a, const uint32_t* __restrict__ key,
                                                     const double* __restrict__ pv,
                                                     const double* __restrict__ pw,
                                                     const int32_t* __restrict__ pid, double h, double esamp,
                                                     const int32_t* __restrict__ bin_start,
                                                     const int32_t* __restrict__ bin_count,
                                                     const int32_t* __restrict__ cand_bin,
                                                     int64_t nbins, Cand* __restrict__ out, int64_t cap) {
    /* Each workgroup forms 256 consecutive records in LDS and writes them
     * out as contiguous 16-byte words: a record per lane stored straight
     * from registers is 17 eight-byte stores at a 136-byte stride, one
     * partial cache line each, and those line writes bounded the kernel. */
    __shared__ __align__(16) Cand tile[256];
    const int64_t nent = nbins > 0 ? (int64_t)bin_start[nbins - 1] + bin_count[nbins - 1] : 0;
    const int64_t nlim = nent < cap ? nent : cap;
    const int t = (int)threadIdx.x;
    for (int64_t e0 = (int64_t)blockIdx.x * 256; e0 < nlim; e0 += (int64_t)gridDim.x * 256) {
        const int64_t e = e0 + t;
        if (e < nlim) {
            const uint32_t kk = key[e];
            const int q = (int)(kk & QMASK);
            const int rank = (int)(kk >> QBITS);
            const double x[3] = {a.px[3 * q], a.px[3 * q + 1], a.px[3 * q + 2]};
            const double r = a.pr[q];
            double xi[3];
            int lo[3], hi[3];
            Cand cd;
            if (!image_box(x, r, rank, a, lo, hi, xi)) {
                lo[0] = lo[1] = lo[2] = 1;
                hi[0] = hi[1] = hi[2] = 0;
                xi[0] = xi[1] = xi[2] = 0.0;
            }
            for (int d = 0; d < 3; ++d) {
                cd.x[d] = xi[d];
                cd.v[d] = pv[3 * q + d]; cd.w[d] = pw[3 * q + d];
            }
            /* the box clipped to the entry's bin, as masks over the bin's cells */
            {
                const int b = cand_bin[e];
                const int bz = b % a.nb[2], by = (b / a.nb[2]) % a.nb[1], bx = b / (a.nb[2] * a.nb[1]);
                const int o[3] = {bx * BIN, by * BIN, bz * BIN};
                uint32_t m[3];
                for (int d = 0; d < 3; ++d) {
                    const int s0 = lo[d] - o[d] > 0 ? lo[d] - o[d] : 0;
                    const int s1 = hi[d] - o[d] < BIN - 1 ? hi[d] - o[d] : BIN - 1;
                    m[d] = s0 <= s1 ? ((2u << s1) - 1u) & ~((1u << s0) - 1u) : 0u;
                }
                uint64_t yz = 0;
                for (int yl = 0; yl < BIN; ++yl)
                    if ((m[1] >> yl) & 1u) yz |= (uint64_t)m[2] << (BIN * yl);
                cd.xmask = m[0];
                cd.yzmask = yz;
            }
            cd.r = r;
            const double r_p = r + h;
            const double r_m = r - h;
            cd.rp2 = r_p * r_p;
            cd.rm2 = r_m * r_m;
            cd.rsq = r * r;
            exact_band(r, esamp, cd.ri2, cd.ro2);
            cd.id = pid[q];
            tile[t] = cd;
        }
        __syncthreads();
        /* the workgroup's records, contiguous in out from e0 (16-byte aligned:
         * 256 records are 34 816 bytes), as 16-byte words and an 8-byte tail */
        const int64_t n = nlim - e0 < 256 ? nlim - e0 : 256;
        const int64_t bytes = n * (int64_t)sizeof(Cand);
        const uint4* src = (const uint4*)tile;
        uint4* dst = (uint4*)(out + e0);
        for (int64_t w = t; w < bytes / 16; w += 256) dst[w] = src[w];
        if ((bytes & 15) != 0 && t == 0)
            ((uint2*)(out + e0))[bytes / 8 - 1] = ((const uint2*)tile)[bytes / 8 - 1];
        __syncthreads();
    }
}

struct FoldArgs {
    int n[3];
    int nb[3];
    FastDiv div_nb2, div_nb12;   /* bin -> (bx, by, bz) */
    int N;
    double sw, frac;
    double rc, vc;             /* cell sphere of the analytic overlap */
    const int32_t* bin_cand_start;
    const int32_t* bin_cand_count;
    const Cand* cand;
    int64_t ent_cap;
    Records rec;
    uint8_t* sflag;
    const uint32_t* cnt_old;   /* per bin: live records of the previous step */
    uint32_t* cnt_new;         /* per bin: live records after this fold */
    int64_t* counters;
    int32_t* orphan;
    int64_t orphan_cap;
    int32_t* dv_cell;
    double* dv_us;
};

/* Sub-cell count of one (offset, radius) query by the whole wavefront:
 * lane L takes the sample column (i, j) = (L / N, L % N) and tests the N
 * samples along z; a ballot + popcount per k adds the column counts.  Every
 * sample test is the reference's FP64 expression
 * (xs_i + ys_j) + zs_k < r^2 (src/ibProcessors3D.hh:141-155); only the
 * integer summation order differs, so the count is exact.  The arguments
 * must be wave-uniform. */
template <int NT>
__device__ __forceinline__ unsigned wave_count(double dx, double dy, double dz, double rsq, int Nrt,
                                               double sw) {
    const int N = NT > 0 ? NT : Nrt;
    const int lane = (int)__lane_id();
    unsigned total = 0;
    for (int p0 = 0; p0 < N * N; p0 += 64) {
        const int pidx = p0 + lane;
        const bool act = pidx < N * N;
        const int i = act ? pidx / N : 0;
        const int j = act ? pidx - i * N : 0;
        const double ax = dx + (-0.5 + ((double)i + 0.5) * sw);
        const double ay = dy + (-0.5 + ((double)j + 0.5) * sw);
        const double sxy = ax * ax + ay * ay;
        if (NT > 0 && NT * NT == 64) {
            /* the 64 sample columns fill the wavefront: every lane is active,
             * and the compare's lane mask is the ballot */
#pragma unroll
            for (int k = 0; k < (NT > 0 ? NT : 1); ++k) {
                const double az = dz + (-0.5 + ((double)k + 0.5) * sw);
                total += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(sxy + az * az < rsq));
            }
        } else if (NT > 0) {
#pragma unroll
            for (int k = 0; k < (NT > 0 ? NT : 1); ++k) {
                const double az = dz + (-0.5 + ((double)k + 0.5) * sw);
                total += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(act && (sxy + az * az < rsq)));
            }
        } else {
            for (int k = 0; k < N; ++k) {
                const double az = dz + (-0.5 + ((double)k + 0.5) * sw);
                total += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(act && (sxy + az * az < rsq)));
            }
        }
    }
    return total;
}

/* Sub-cell counts of the pending lanes of one fold wavefront (N = 8).  The
 * wavefront is one x-slice of a bin with lane = 8 y + z, so the offset dx
 * and the radius are the same on every lane, lanes of one z-row share dz
 * bit for bit and lanes of one y-row share dy.  Two per-wavefront LDS
 * tables hold the squared sample offsets: zsq[8 r + k] = (dz_r + d_k)^2 for
 * z-row r, ysq[8 y + j] = (dy_y + d_j)^2 for y-row y (lane 8 y + j forms
 * its own entry: its dy and its d_j).  The pending lanes are taken z-row by
 * z-row: the row's eight zsq[k] are read once (wave-uniform); each query
 * then has lane 8 i + j read ysq of the query's y-row, form
 * dx_sq[i] + dy_sq[j] and test the eight k.  Every sample test is the same
 * FP64 expression as wave_count's, (dx_sq[i] + dy_sq[j]) + dz_sq[k] < r^2,
 * so the counts are the same.  A query's count goes into the querying
 * lane by a select (no branch); the fractions follow after. */
__device__ __forceinline__ void slice_counts8(unsigned long long pend, double dx, double dy, double dz,
                                              double rsq, double sw, double frac, double di, double dj, int& n,
                                              double& sf) {
    __shared__ double ztab[FOLD_WAVES][64];
    __shared__ double ytab[FOLD_WAVES][64];
    const int lane = (int)__lane_id();
    double* zt = ztab[threadIdx.x >> 6];
    double* yt = ytab[threadIdx.x >> 6];
    {
        const double dzr = __shfl(dz, lane >> 3, 64);   /* dz of z-row lane / 8 */
        const double az = dzr + dj;
        zt[lane] = az * az;
        const double ay = dy + dj;                      /* lane 8 y + j: its own dy and d_j */
        yt[lane] = ay * ay;
    }
    const double ax = readlane_d(dx, __ffsll(pend) - 1) + di;
    const double xsq = ax * ax;
    const unsigned long long counted = pend;
    int cnt_lane = n;
    while (pend) {
        const int L = __ffsll(pend) - 1;
        const unsigned long long row = pend & (0x0101010101010101ull << (L & 7));
        pend &= ~row;
        const double* zr = zt + (L & 7) * 8;
        double zsq[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) zsq[k] = zr[k];
        unsigned long long q = row;
        while (q) {
            const int M = __ffsll(q) - 1;
            q ^= 1ull << M;
            const double sxy = xsq + yt[(M & 0x38) + (lane & 7)];
            unsigned cnt = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) cnt += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(sxy + zsq[k] < rsq));
            cnt_lane = lane == M ? (int)cnt : cnt_lane;   /* a select, not a branch */
        }
    }
    if ((counted >> lane) & 1ull) {
        n = cnt_lane;
        sf = frac * (double)cnt_lane;
    }
}

/* calcSolidFraction (src/ibProcessors3D.hh:116-158) for one query per lane,
 * called by all 64 lanes together: the two distance shortcuts per lane, then
 * every lane that needs a count gets it from wave_count in turn (SLICE: from
 * slice_counts8, whose lane layout the caller guarantees).
 *
 * Between the reference's shortcut radii r -+ sqrt(3.1)/2, a cell whose
 * sub-cell samples all lie strictly inside (resp. outside) the sphere has
 * count N^3 (resp. 0) and solid fraction fraction * count, known without
 * counting.  Two bounds find such cells, each with a margin tol far above
 * the rounding error of the sample test, so the result is the same as the
 * full count's, bit for bit:
 *  - the samples lie within e = sqrt(3) (1/2 - 1/(2N)) of the cell centre:
 *    a centre within ri = r - e - tol (beyond ro = r + e + tol) of the
 *    sphere centre settles the cell;
 *  - tighter, the samples lie in the box of half-width s = 1/2 - 1/(2N)
 *    about the cell centre: the box's farthest point from the sphere centre
 *    inside r (nearest point outside r) settles it.  With the box, exactly
 *    the cells the sphere surface cuts are left to count. */
template <int NT, bool SLICE>
__device__ __forceinline__ double coop_solid_fraction(bool active, double dx, double dy, double dz,
                                                      double rp2, double rm2, double rsq, double ri2,
                                                      double ro2, int Nrt, double sw, double frac, int& n,
                                                      int& br, double di = 0.0, double dj = 0.0) {
    const int N = NT > 0 ? NT : Nrt;
    const int N3 = N * N * N;
    /* branch-free: every lane evaluates every test */
    const double dist = dx * dx + dy * dy + dz * dz;
    const bool far = dist > rp2;
    const bool core = !far && dist < rm2;
    const bool band = active && !far && !core;
    br = !active || far ? 0 : (core ? 1 : 2);
    const bool allin = band && dist < ri2;
    bool need = band && !allin && !(dist > ro2);
    n = (active && core) || allin ? N3 : 0;
    double sf = active && core ? 1.0 : (allin ? frac * (double)N3 : 0.0);
    if (__builtin_amdgcn_ballot_w64(need) != 0ull) {
        const double s = -0.5 + ((double)(N - 1) + 0.5) * sw;   /* outermost sample offset */
        const double tol = 1e-9 * (1.0 + rsq);
        const double ex = fabs(dx), ey = fabs(dy), ez = fabs(dz);
        const double fx = ex + s, fy = ey + s, fz = ez + s;
        const double nx = fmax(ex - s, 0.0), ny = fmax(ey - s, 0.0), nz = fmax(ez - s, 0.0);
        const bool hin = need & ((fx * fx + fy * fy) + fz * fz < rsq - tol);
        const bool hout = (nx * nx + ny * ny) + nz * nz > rsq + tol;
        n = hin ? N3 : n;
        sf = hin ? frac * (double)N3 : sf;
        need = need && !hin && !hout;
    }
    unsigned long long pend = __builtin_amdgcn_ballot_w64(need);
    if constexpr (SLICE) {
        static_assert(NT == 8, "one sample column per lane");
        if (pend) slice_counts8(pend, dx, dy, dz, rsq, sw, frac, di, dj, n, sf);
        return sf;
    }
    const int lane = (int)__lane_id();
    while (pend) {
        const int L = __ffsll(pend) - 1;
        pend &= pend - 1;
        const double bx = readlane_d(dx, L), by = readlane_d(dy, L), bz = readlane_d(dz, L);
        const double b2 = readlane_d(rsq, L);
        const unsigned cnt = wave_count<NT>(bx, by, bz, b2, N, sw);
        if (lane == L) { n = (int)cnt; sf = frac * (double)cnt; }
    }
    return sf;
}

/* The voxelization rule: NT >= 0 the reference's sub-cell count (NT = N
 * when specialized, 0 = runtime N), NT < 0 the analytic overlap of the
 * differentiable mode (per lane, no wavefront cooperation).  Called by all
 * 64 lanes together. */
template <int NT, bool SLICE = false>
__device__ __forceinline__ double solid_fraction(bool active, double dx, double dy, double dz,
                                                 double rp2, double rm2, double rsq, double r, double ri2,
                                                 double ro2, int Nrt, double sw, double frac, double rc,
                                                 double vc, int& n, int& br, double di = 0.0,
                                                 double dj = 0.0) {
    if constexpr (NT < 0) {
        n = 0;
        br = 0;
        if (!active) return 0.0;
        return overlap_fraction(sqrt((dx * dx + dy * dy) + dz * dz), r, rc, vc, nullptr, &br);
    } else {
        return coop_solid_fraction<NT, SLICE>(active, dx, dy, dz, rp2, rm2, rsq, ri2, ro2, Nrt, sw, frac, n, br, di,
                                              dj);
    }
}

enum { ST_NONE = -2, ST_KEPT = -1, ST_SET_FROM = 0 };

/* One wavefront per x-slice of an 8^3 bin (8 x 8 cells, one per lane;
 * FOLD_WAVES slices per workgroup).  The cell's record of the previous step is
 * folded with the bin's candidates in key order through the four-case rule
 * of SetSingleSphere3D::process (src/ibProcessors3D.hh:85-105) and written
 * back in place. */
template <int NT, bool INITVEL>
__global__ __launch_bounds__(64 * FOLD_WAVES) void k_fold(FoldArgs a) {
    /* neighbouring bins (z fastest) share the lines of the cell-ordered
     * arrays: keep them on one XCD's L2 */
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int b = blk / FOLD_SPLIT;
    const int ncand = a.bin_cand_count[b];
    if (ncand == 0 && a.cnt_old[b] == 0u) return;   /* nothing here, nothing was here */
    const int cstart = a.bin_cand_start[b];
    if ((int64_t)cstart + ncand > a.ent_cap) return;   /* overflow already flagged */
    const int bxy = (int)fdiv((uint32_t)b, a.div_nb2);
    const int bz = b - bxy * a.nb[2];
    const int bx = (int)fdiv((uint32_t)b, a.div_nb12);
    const int by = bxy - bx * a.nb[1];
    const int l = (int)threadIdx.x + (blk % FOLD_SPLIT) * (64 * FOLD_WAVES);   /* cell of the bin */
    const int X = bx * BIN + (l >> 6), Y = by * BIN + ((l >> 3) & 7), Z = bz * BIN + (l & 7);
    const bool valid = X < a.n[0] && Y < a.n[1] && Z < a.n[2];
    const int64_t cidx = valid ? ((int64_t)X * a.n[1] + Y) * a.n[2] + Z : 0;
    const int64_t ridx = cidx;   /* record slot = cell index */
    /* a wavefront is the x-slice X of the bin: 8 y by 8 z cells */
    const int wl = __builtin_amdgcn_readfirstlane(l >> 6);

    const bool had = valid && a.sflag[cidx] != 0;
    /* The record's state through the candidate loop: ST_SET_FROM + j when
     * candidate j set it last, ST_KEPT when it is the previous step's record
     * untouched, ST_NONE when there is none (never was or was reset). */
    int sj = had ? ST_KEPT : ST_NONE, sid = 0;
    double ssf = 0.0, sus[3] = {0.0, 0.0, 0.0};
    bool dv = false;
    double dvus[3] = {0.0, 0.0, 0.0};
    /* sample offsets of the lane's column (i, j) = (lane / 8, lane % 8) in
     * the N = 8 count, the reference's delta expression */
    const double sdi = -0.5 + ((double)((l & 63) >> 3) + 0.5) * a.sw;
    const double sdj = -0.5 + ((double)(l & 7) + 0.5) * a.sw;
    if (had) {
        ssf = a.rec.sus[4 * ridx];
        sid = a.rec.id[ridx];
    }
    const int oid = sid;
    const double fX = (double)X, fY = (double)Y, fZ = (double)Z;

    /* candidates in key order; their data are wave-uniform */
    const Cand* __restrict__ cands = a.cand + cstart;
    const int lane = (int)__lane_id();
    for (int c0 = 0; c0 < ncand; c0 += 64) {
        /* whole-wavefront rejection against the candidates' boxes (their
         * x-slices within the bin), 64 candidates at a time: lane c reads
         * candidate c0 + c's x mask, and the ballot lists the candidates
         * this x-slice has to take, in key order */
        const int cl = c0 + lane < ncand ? c0 + lane : ncand - 1;
        unsigned long long todo = __builtin_amdgcn_ballot_w64(c0 + lane < ncand && ((cands[cl].xmask >> wl) & 1u) != 0u);
        while (todo) {
            const int j = c0 + (__ffsll(todo) - 1);
            todo &= todo - 1;
            const Cand& cd = cands[j];
            /* the lane's cell against the candidate's (y, z) cells within
             * the bin */
            const bool inb = valid & (((cd.yzmask >> lane) & 1ull) != 0ull);
            const double dx = fX - cd.x[0];
            const double dy = fY - cd.x[1];
            const double dz = fZ - cd.x[2];
            /* Nothing changes for a lane outside the reach r + h of the
             * sphere (its solid fraction is 0) unless it holds a record of
             * this particle (the particle has left the cell, case 2 of the
             * rule); a lane with no record stays empty.  Skip the candidate
             * when that holds for every lane of the wavefront. */
            {
                /* non-short-circuit: the candidate's reach and id load
                 * with its centre, not behind a branch each */
                const bool near = inb & (dx * dx + dy * dy + dz * dz <= cd.rp2);
                const bool mine = inb & (ssf > SOLFRAC_MIN) & (sid == cd.id);
                if (__builtin_amdgcn_ballot_w64(near | mine) == 0ull) continue;
            }
            int nn, br;
            const double sf = solid_fraction<NT, NT == 8>(inb, dx, dy, dz, cd.rp2, cd.rm2, cd.rsq, cd.r, cd.ri2, cd.ro2,
                                                 a.N, a.sw, a.frac, a.rc, a.vc, nn, br, sdi, sdj);
            /* the four-case rule, branch-free: flag = new + 2 old;
             * 0 -> zero, 1 -> set, 2 -> zero if own record, 3 -> set if
             * larger or own record */
            const bool pnew = sf > SOLFRAC_MIN, pold = ssf > SOLFRAC_MIN, own = sid == cd.id;
            const bool zero = inb & !pnew & (!pold | own);
            const bool set = inb & pnew & (!pold | (sf > ssf) | own);
            sj = set ? ST_SET_FROM + j : (zero ? ST_NONE : sj);
            ssf = set ? sf : (zero ? 0.0 : ssf);
            sid = set ? cd.id : (zero ? 0 : sid);
            if constexpr (INITVEL) {
                /* setValues, src/ibProcessors3D.hh:162-174, here because
                 * initVel reads the velocity of the record in force */
                const double u0 = cd.v[0] + (cd.w[1] * dz - cd.w[2] * dy);
                const double u1 = cd.v[1] + (-cd.w[0] * dz + cd.w[2] * dx);
                const double u2 = cd.v[2] + (cd.w[0] * dy - cd.w[1] * dx);
                sus[0] = set ? u0 : (zero ? 0.0 : sus[0]);
                sus[1] = set ? u1 : (zero ? 0.0 : sus[1]);
                sus[2] = set ? u2 : (zero ? 0.0 : sus[2]);
            }
            if (INITVEL && inb && sf > SOLFRAC_MAX) {
                dv = true;
                if (sj == ST_KEPT) {
                    dvus[0] = a.rec.sus[4 * ridx + 1];
                    dvus[1] = a.rec.sus[4 * ridx + 2];
                    dvus[2] = a.rec.sus[4 * ridx + 3];
                } else {
                    dvus[0] = sus[0]; dvus[1] = sus[1]; dvus[2] = sus[2];
                }
            }
        }
    }

    const bool keep = valid && sj != ST_NONE;
    const unsigned long long m = __builtin_amdgcn_ballot_w64(keep);
    /* one atomic per wavefront on the bin's own counter (distinct addresses
     * across bins; a single global counter would serialize every wave) */
    if (lane == 0 && m) atomicAdd(&a.cnt_new[b], (uint32_t)__popcll(m));
    if (!valid) return;
    if (sj >= ST_SET_FROM) {
        const int js = sj - ST_SET_FROM;
        if constexpr (!INITVEL) {
            /* setValues (src/ibProcessors3D.hh:162-174) of the candidate
             * that set the record last, once per cell: the same FP64
             * expressions as in the candidate loop */
            const Cand& cw = cands[js];
            const double dx = fX - cw.x[0];
            const double dy = fY - cw.x[1];
            const double dz = fZ - cw.x[2];
            sus[0] = cw.v[0] + (cw.w[1] * dz - cw.w[2] * dy);
            sus[1] = cw.v[1] + (-cw.w[0] * dz + cw.w[2] * dx);
            sus[2] = cw.v[2] + (cw.w[0] * dy - cw.w[1] * dx);
        }
        /* the whole 32-byte slot, the fraction with the velocity (whole
         * lines where the wavefront's cells are live); a live record whose
         * owner did not change keeps its id */
        *(double2*)(a.rec.sus + 4 * ridx) = make_double2(ssf, sus[0]);
        *(double2*)(a.rec.sus + 4 * ridx + 2) = make_double2(sus[1], sus[2]);
        if (!had || sid != oid) a.rec.id[ridx] = sid;
        a.rec.cand[ridx] = cstart + js;   /* the setter's entry index */
        if (a.rec.m4) {
            float m[4];
            record_mirror(ssf, sus, m);
            *(float4*)(a.rec.m4 + 4 * ridx) = make_float4(m[0], m[1], m[2], m[3]);
        }
    } else if (sj == ST_KEPT) {
        /* untouched this step: values stay, summed on the orphan path */
        a.rec.cand[ridx] = -1;
        const unsigned long long p = atomicAdd((unsigned long long*)&a.counters[2], 1ull);
        if ((int64_t)p < a.orphan_cap) a.orphan[p] = (int32_t)ridx;
        else atomicOr((unsigned long long*)&a.counters[1], 4ull);
    }
    if (keep != had) a.sflag[cidx] = keep ? 1 : 0;
    if (INITVEL && dv && a.dv_cell) {
        const unsigned long long p = atomicAdd((unsigned long long*)&a.counters[5], 1ull);
        if ((int64_t)p < a.orphan_cap) {
            a.dv_cell[p] = (int32_t)cidx;
            a.dv_us[3 * p] = dvus[0];
            a.dv_us[3 * p + 1] = dvus[1];
            a.dv_us[3 * p + 2] = dvus[2];
        } else {
            atomicOr((unsigned long long*)&a.counters[1], 8ull);
        }
    }
}

template <int NT>
__global__ __launch_bounds__(256) void k_subcell(int64_t n, const double* __restrict__ d3,
                                                 const double* __restrict__ r, int N, double sw,
                                                 double frac, double h, double rc, double vc,
                                                 double esamp, double* __restrict__ sf,
                                                 int32_t* __restrict__ cnt, int32_t* __restrict__ br) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = k < n;
    const double rr = act ? r[k] : 1.0;
    const double r_p = rr + h, r_m = rr - h;
    const double dx = act ? d3[3 * k] : 0.0, dy = act ? d3[3 * k + 1] : 0.0, dz = act ? d3[3 * k + 2] : 0.0;
    int nn, bb;
    double ri2, ro2;
    exact_band(rr, esamp, ri2, ro2);
    const double v = solid_fraction<NT>(act, dx, dy, dz, r_p * r_p, r_m * r_m, rr * rr, rr, ri2, ro2, N, sw, frac,
                                        rc, vc, nn, bb);
    if (act) {
        sf[k] = v;
        cnt[k] = nn;
        br[k] = bb;
    }
}

/* Prescribed kinematics: x += v, wrapped into [-1/2, n - 1/2) on periodic axes. */
__global__ __launch_bounds__(256) void k_advance(int64_t np, double* __restrict__ px,
                                                 const double* __restrict__ pv, int n0, int n1, int n2,
                                                 int p0, int p1, int p2) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int n[3] = {n0, n1, n2};
    const int per[3] = {p0, p1, p2};
    for (int d = 0; d < 3; ++d) {
        double x = px[3 * q + d] + pv[3 * q + d];
        if (per[d]) {
            const double L = (double)n[d];
            if (x >= L - 0.5) x = x - L;
            else if (x < -0.5) x = x + L;
        }
        px[3 * q + d] = x;
    }
}

/* The per-step clears of the voxelization in one launch: counters 0 and
 * 2..7 (1 holds the overflow flags), the per-bin candidate counts and
 * fill cursors, and the live-record counts the fold is about to write. */
__global__ __launch_bounds__(256) void k_clear_step(int64_t* __restrict__ counters, int32_t* __restrict__ bcount,
                                                    int32_t* __restrict__ bfill, uint32_t* __restrict__ cnt_new,
                                                    int64_t nbins) {
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 < 8 && i0 != 1) counters[i0] = 0;
    for (int64_t i = i0; i < nbins; i += (int64_t)gridDim.x * blockDim.x) {
        bcount[i] = 0;
        bfill[i] = 0;
        cnt_new[i] = 0u;
    }
}

static VoxArgs vox_args(lbdem_psm* h) {
    VoxArgs a;
    for (int d = 0; d < 3; ++d) { a.n[d] = h->g.n[d]; a.per[d] = h->g.periodic[d]; a.nb[d] = h->g.nb[d]; }
    a.auto_images = (h->flags & LBDEM_PSM_AUTO_IMAGES) ? 1 : 0;
    a.np = h->np;
    a.px = h->px; a.pr = h->pr; a.pexcl = h->pexcl;
    return a;
}

int run_voxelize(lbdem_psm* h) {
    hipEvent_t ev = nullptr;
    prof_begin(h, 1, &ev);
    const int64_t nbins = h->nbins;
    VoxArgs va = vox_args(h);
    /* counters[1] (overflow flags) is cleared by the public entry points only */
    uint32_t* cnt_new = h->bin_any[1 - h->cur_any];
    hipLaunchKernelGGL(k_clear_step, dim3((unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, (nbins + 255) / 256))),
                       dim3(256), 0, h->stream, h->counters, h->bin_cand_count, h->bin_fill, cnt_new, nbins);
    const unsigned pg = (unsigned)std::max<int64_t>(1, (h->np + 3) / 4);   /* a wave per particle */
    if (h->np > 0)
        hipLaunchKernelGGL(k_bin_count, dim3(pg), dim3(256), 0, h->stream, va, h->bin_cand_count, h->p_nent);
    size_t tb = h->scan_tmp_bytes;
    PSM_HIP(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp, tb, h->bin_cand_count, h->bin_cand_start, (int)nbins, h->stream));
    if (h->np > 0) {
        tb = h->scan_tmp_bytes;
        PSM_HIP(h, hipcub::DeviceScan::ExclusiveSum(h->scan_tmp, tb, h->p_nent, h->p_ent_start, (int)h->np, h->stream));
        hipLaunchKernelGGL(k_bin_fill, dim3(pg), dim3(256), 0, h->stream, va, h->bin_cand_start, h->bin_fill, h->cand_key, h->cand_bin, h->ent_cap, h->counters);
    }
    hipLaunchKernelGGL(k_bin_sort, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, h->stream, h->bin_cand_start, h->bin_cand_count, h->cand_key, nbins, h->ent_cap);
    if (h->np > 0)
        hipLaunchKernelGGL(k_owner_entries, dim3(pg), dim3(256), 0, h->stream, va, h->bin_cand_start, h->bin_cand_count, h->cand_key, h->p_ent_start, h->owner_ent, h->ent_cap);

    /* farthest sub-cell sample from the cell centre, sqrt(3) (1/2 - 1/(2N)) */
    const double esamp = std::sqrt(3.0) * (0.5 - 0.5 / (double)h->desc.slices_per_dim) + 1e-12;
    hipLaunchKernelGGL(k_cand_gather, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (h->ent_cap + 255) / 256))),
                       dim3(256), 0, h->stream, va, h->cand_key, h->pv, h->pw, h->pid, h->sqrt3half, esamp,
                       h->bin_cand_start, h->bin_cand_count, h->cand_bin, nbins, h->cand_data, h->ent_cap);

    FoldArgs fa;
    for (int d = 0; d < 3; ++d) { fa.n[d] = h->g.n[d]; fa.nb[d] = h->g.nb[d]; }
    fa.div_nb2 = make_fastdiv((uint32_t)h->g.nb[2]);
    fa.div_nb12 = make_fastdiv((uint32_t)(h->g.nb[2] * h->g.nb[1]));
    fa.N = h->desc.slices_per_dim;
    fa.sw = 1.0 / (double)fa.N;
    fa.frac = 1.0 / (double)(fa.N * fa.N * fa.N);
    fa.rc = h->cell_rc;
    fa.vc = h->cell_vc;
    const bool overlap = h->desc.voxel_mode == LBDEM_PSM_VOXEL_OVERLAP;
    const bool init_vel = (h->flags & LBDEM_PSM_INIT_VEL) != 0;
    fa.bin_cand_start = h->bin_cand_start;
    fa.bin_cand_count = h->bin_cand_count;
    fa.cand = h->cand_data;
    fa.ent_cap = h->ent_cap;
    fa.rec = h->rec;
    fa.sflag = h->sflag;
    fa.cnt_old = h->bin_any[h->cur_any];
    fa.cnt_new = cnt_new;
    fa.counters = h->counters;
    fa.orphan = h->orphan;
    fa.orphan_cap = h->orphan_cap;
    fa.dv_cell = init_vel ? h->dv_cell : nullptr;
    fa.dv_us = h->dv_us;
    dim3 grid((unsigned)(nbins * FOLD_SPLIT)), block(64 * FOLD_WAVES);
    if (overlap) {
        if (init_vel) hipLaunchKernelGGL((k_fold<-1, true>), grid, block, 0, h->stream, fa);
        else hipLaunchKernelGGL((k_fold<-1, false>), grid, block, 0, h->stream, fa);
    } else if (init_vel) {
        if (fa.N == 5) hipLaunchKernelGGL((k_fold<5, true>), grid, block, 0, h->stream, fa);
        else if (fa.N == 8) hipLaunchKernelGGL((k_fold<8, true>), grid, block, 0, h->stream, fa);
        else hipLaunchKernelGGL((k_fold<0, true>), grid, block, 0, h->stream, fa);
    } else {
        if (fa.N == 5) hipLaunchKernelGGL((k_fold<5, false>), grid, block, 0, h->stream, fa);
        else if (fa.N == 8) hipLaunchKernelGGL((k_fold<8, false>), grid, block, 0, h->stream, fa);
        else hipLaunchKernelGGL((k_fold<0, false>), grid, block, 0, h->stream, fa);
    }
    PSM_HIP(h, hipGetLastError());
    h->cur_any = 1 - h->cur_any;
    h->records_from_host = false;
    prof_end(h, 1, ev);
    if (init_vel) {
        int rc = launch_define_velocity(h);
        if (rc) return rc;
    }
    return LBDEM_PSM_OK;
}

/* The float view of every live record, after a writer other than the fold
 * (host records, a checkpoint, the sf perturbation of the adjoint checks). */
__global__ __launch_bounds__(256) void k_mirror_records(int64_t nc, const uint8_t* __restrict__ sflag,
                                                        const double* __restrict__ sus,
                                                        float* __restrict__ m4) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x) {
        if (!sflag[c]) continue;
        const double u[3] = {sus[4 * c + 1], sus[4 * c + 2], sus[4 * c + 3]};
        float m[4];
        record_mirror(sus[4 * c], u, m);
        *(float4*)(m4 + 4 * c) = make_float4(m[0], m[1], m[2], m[3]);
    }
}

int mirror_records(lbdem_psm* h) {
    if (!h->rec.m4) return LBDEM_PSM_OK;
    const int64_t nc = h->g.nc;
    hipLaunchKernelGGL(k_mirror_records, dim3((unsigned)std::min<int64_t>(8192, (nc + 255) / 256)), dim3(256), 0,
                       h->stream, nc, h->sflag, h->rec.sus, h->rec.m4);
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

/* Total live records: sum of the per-bin counts into counters[0]. */
__global__ __launch_bounds__(256) void k_sum_bins(const uint32_t* __restrict__ cnt, int64_t nbins,
                                                  int64_t* __restrict__ out) {
    __shared__ unsigned long long sh[256];
    unsigned long long v = 0;
    for (int64_t b = threadIdx.x; b < nbins; b += blockDim.x) v += cnt[b];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (int64_t)sh[0];
}

int count_records(lbdem_psm* h) {
    hipLaunchKernelGGL(k_sum_bins, dim3(1), dim3(256), 0, h->stream, h->bin_any[h->cur_any], h->nbins, h->counters);
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int launch_subcell(lbdem_psm* h, int64_t n, const double* d_d3, const double* d_r, double* d_sf,
                   int32_t* d_cnt, int32_t* d_br) {
    const int N = h->desc.slices_per_dim;
    const double sw = 1.0 / (double)N, frac = 1.0 / (double)(N * N * N);
    dim3 grid((unsigned)std::max<int64_t>(1, (n + 255) / 256)), block(256);
    const double rc = h->cell_rc, vc = h->cell_vc;
    const double es = std::sqrt(3.0) * (0.5 - 0.5 / (double)N) + 1e-12;
    if (h->desc.voxel_mode == LBDEM_PSM_VOXEL_OVERLAP)
        hipLaunchKernelGGL(k_subcell<-1>, grid, block, 0, h->stream, n, d_d3, d_r, N, sw, frac, h->sqrt3half, rc, vc, es, d_sf, d_cnt, d_br);
    else if (N == 5) hipLaunchKernelGGL(k_subcell<5>, grid, block, 0, h->stream, n, d_d3, d_r, N, sw, frac, h->sqrt3half, rc, vc, es, d_sf, d_cnt, d_br);
    else if (N == 8) hipLaunchKernelGGL(k_subcell<8>, grid, block, 0, h->stream, n, d_d3, d_r, N, sw, frac, h->sqrt3half, rc, vc, es, d_sf, d_cnt, d_br);
    else hipLaunchKernelGGL(k_subcell<0>, grid, block, 0, h->stream, n, d_d3, d_r, N, sw, frac, h->sqrt3half, rc, vc, es, d_sf, d_cnt, d_br);
    PSM_HIP(h, hipGetLastError());
    return LBDEM_PSM_OK;
}

int launch_advance(lbdem_psm* h) {
    if (h->np == 0) return LBDEM_PSM_OK;
    hipEvent_t ev = nullptr;
    prof_begin(h, 3, &ev);
    hipLaunchKernelGGL(k_advance, dim3((unsigned)((h->np + 255) / 256)), dim3(256), 0, h->stream, h->np, h->px, h->pv,
                       h->g.n[0], h->g.n[1], h->g.n[2], h->g.periodic[0], h->g.periodic[1], h->g.periodic[2]);
    PSM_HIP(h, hipGetLastError());
    prof_end(h, 3, ev);
    return LBDEM_PSM_OK;
}

size_t scan_temp_bytes(int64_t n) {
    size_t tb = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (int32_t*)nullptr, (int32_t*)nullptr, (int)n, (hipStream_t)0);
    return tb;
}

}  // namespace psm

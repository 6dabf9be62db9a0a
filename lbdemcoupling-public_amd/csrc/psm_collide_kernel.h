/*
 * Fused PSM collide + stream kernel template, shared by the FP64 translation
 * unit (psm_collide.hip, compiled without contraction: bit-exact with the
 * FP64 restatement) and the FP32 / FP16-storage unit (psm_collide_lp.hip,
 * compiled with FMA contraction and the hardware reciprocal).
 */
#ifndef LBDEM_PSM_COLLIDE_KERNEL_H
#define LBDEM_PSM_COLLIDE_KERNEL_H

#include "psm_common.h"

namespace psm {

/* 1/x: exact division in FP64, the hardware reciprocal (1 ulp) in FP32. */
template <typename T> __device__ __forceinline__ T recip(T x);
template <> __device__ __forceinline__ double recip<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }

struct CollideArgs {
    int nx, ny, nz;
    int px, py, pz;
    int64_t nc;
    PopLayout L;
    double omega, ooo;
    double acc[3];
    const uint8_t* sflag;
    const double* rsus;   /* 4 per record slot: sf, us[3] */
    void* rF;             /* 4 per record slot: the cell's force, double for FP64 storage,
                           * float otherwise, and a padding zero */
    const float* rm4;  /* float view of the records (FP32/FP16 storage) */
    FastDiv div_nzc, div_ny;   /* block -> (row, chunk), row -> (x, y) */
    int sub_shift;             /* workgroups per (row, chunk) block: 1 << sub_shift */
    /* periodic pressure: each inlet / outlet plane cell of the cross-section
     * box writes its rhoBar (FP64 sum of its 19 pre-collision populations)
     * to plane_rho[plane * pncell + (a - pa0) * pnb + (b - pb0)]; none when
     * plane_rho is null */
    double* plane_rho;
    int pdim, plast, pa0, pa1, pb0, pb1, pnb;   /* plast = n[pdim] - 1 */
    int64_t pncell;
};

/* The plane sums of the periodic pressure from the collide's own loads:
 * the lanes of the inlet (index 0 along pdim) and outlet (index n-1) plane
 * cells inside the box write the cell's rhoBar, summed in FP64 in
 * population order (the order of the separate plane-sum kernel, so the sums
 * are the same to the bit).  Wave-uniform tests first: only the wavefronts
 * that hold such a cell go on. */
template <typename T>
__device__ __forceinline__ void plane_rhobar(const CollideArgs& a, const T* fp, int x, int y, int z) {
    if (a.plane_rho == nullptr) return;
    const int lane = (int)__lane_id();
    const int zw0 = __builtin_amdgcn_readfirstlane(z - lane);   /* z of lane 0 of this wavefront */
    bool near;
    if (a.pdim == 2)
        near = (zw0 == 0 || zw0 + 63 >= a.nz - 1) && x >= a.pa0 && x <= a.pa1 && y >= a.pb0 && y <= a.pb1;
    else if (a.pdim == 0)
        near = (x == 0 || x == a.nx - 1) && y >= a.pa0 && y <= a.pa1;
    else
        near = (y == 0 || y == a.ny - 1) && x >= a.pa0 && x <= a.pa1;
    if (!near) return;
    const int ci = a.pdim == 0 ? x : (a.pdim == 1 ? y : z);
    const int pl = ci == 0 ? 0 : (ci == a.plast ? 1 : -1);
    const int ca = a.pdim == 0 ? y : x;
    const int cb = a.pdim == 2 ? y : z;
    if (pl >= 0 && ca >= a.pa0 && ca <= a.pa1 && cb >= a.pb0 && cb <= a.pb1) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) s = s + (double)fp[i];
        a.plane_rho[pl * a.pncell + (int64_t)(ca - a.pa0) * a.pnb + (cb - a.pb0)] = s;
    }
}

#ifndef COLLIDE_WG
#define COLLIDE_WG 256
#endif
/* Lanes per collide workgroup: a (row, chunk) block of zc cells is
 * zc / COLLIDE_WG workgroups (one when zc is smaller). */
inline int collide_wg(int zc) { return zc < COLLIDE_WG ? zc : COLLIDE_WG; }

/* Resident wavefronts per SIMD the collide is held to (0: as many as its
 * registers allow).  Fewer wavefronts in flight than the registers allow
 * moved the store faster: on MI355X, stepping one lattice alternately capped
 * and uncapped (512x512x768 channel, 512^3 and 256^3 boxes, the channel's
 * 64x512x768 slab), three per SIMD made the FP64 collide 1-3 % faster than
 * its register limit of four and the FP32 collide 4-8 % faster than its
 * eight; FP16 storage ran fastest uncapped (seven; four cost 13 %). */
inline int collide_wave_cap(int precision) { return precision == LBDEM_PSM_FP16_STORE ? 0 : 3; }

/* Dynamic LDS per workgroup that holds a kernel with wgz-lane workgroups at
 * `waves` wavefronts per SIMD on a CU of lds_cu bytes of LDS and four SIMDs.
 * The collide uses no LDS: the reservation only limits how many of its
 * workgroups share a CU.  The size is taken from the middle of the range
 * that fits exactly that many workgroups, clear of both ends, because the
 * hardware rounds each allocation up: on MI355X (160 KiB per CU) 54 272
 * bytes let only two workgroups in, not three, while 41 984 to 53 248 let
 * three.  0 when the cap cannot be expressed. */
inline unsigned collide_cap_lds(int lds_cu, int wgz, int waves) {
    const int wg_waves = (wgz + 63) / 64;
    if (waves <= 0 || lds_cu <= 0 || (4 * waves) % wg_waves != 0) return 0;
    const unsigned k = (unsigned)(4 * waves / wg_waves), cu = (unsigned)lds_cu;
    const unsigned bytes = ((cu / (k + 1) + cu / k) / 2) & ~511u;
    const unsigned slack = 2048;   /* the allocation granularity or above */
    return (k + 1) * bytes > cu && k * (bytes + slack) <= cu ? bytes : 0;
}

/* The cap's LDS for one kernel (a collide instance, the device copy), used
 * only when the runtime's occupancy for the kernel confirms it: exactly the
 * capped number of workgroups per CU, fewer than its registers allow. */
template <typename K>
unsigned collide_lds_for(K kernel, int device, int wgz, int waves) {
    int lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess)
        return 0;
    const unsigned bytes = collide_cap_lds(lds_cu, wgz, waves);
    if (!bytes) return 0;
    int uncapped = 0, capped = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&uncapped, reinterpret_cast<const void*>(kernel), wgz, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&capped, reinterpret_cast<const void*>(kernel), wgz, bytes) !=
            hipSuccess)
        return 0;
    const int per_cu = 4 * waves / ((wgz + 63) / 64);
    return capped == per_cu && capped < uncapped ? bytes : 0;
}

/* Equilibria of a population pair (i, i+9): c_{i+9} = -c_i, so the pair
 * shares c.j, (3 c.j)(c.j) - |j|^2 and its product with 1.5/rho; each value
 * is bit-identical to equilibrium(i, ...) / equilibrium(i+9, ...). */
template <typename T>
__device__ __forceinline__ void eq_pair(int i, T rhoBar, T p, const T* j, T jSqr, T& ei, T& eo) {
    const T cj = cdot<T>(i, j);
    const T a = T(3) * cj;
    const T b = a * cj;
    const T q = b - jSqr;
    const T pq = p * q;
    ei = wgt<T>(i) * ((rhoBar + a) + pq);
    eo = wgt<T>(i) * ((rhoBar - a) + pq);
}

template <typename T>
__device__ __forceinline__ T eq_rest(T rhoBar, T p, T jSqr) {
    return equilibrium<T>(0, rhoBar, p, nullptr, jSqr);
}

/* BGK (with Guo forcing when FORCE) of all 19 populations of one cell, in
 * the operation order of the base dynamics. */
template <typename T, bool FORCE>
__device__ __forceinline__ void bgk_all(const T* f, T* fb, T omega, const T* acc, T& rhoBar, T* j) {
    moments<T>(f, rhoBar, j);
    const T rho = rhoBar + T(1);
    const T invRho = recip<T>(rho);
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
    fb[0] = f[0] * omm + omega * eq_rest<T>(rhoBar, p, jSqr);
#pragma unroll
    for (int i = 1; i <= 9; ++i) {
        T ei, eo;
        eq_pair<T>(i, rhoBar, p, jEq, jSqr, ei, eo);
        fb[i] = f[i] * omm + omega * ei;
        fb[i + 9] = f[i + 9] * omm + omega * eo;
    }
    if (FORCE) {
        const T pref = (T(1) - T(0.5) * omega) * rho;
        const T ua = (u[0] * acc[0] + u[1] * acc[1]) + u[2] * acc[2];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const T ca = cdot<T>(i, acc);
            const T cu = cdot<T>(i, u);
            const T g = T(3) * (ca - ua) + T(9) * (cu * ca);
            fb[i] = fb[i] + (pref * wgt<T>(i)) * g;
        }
    }
}

/* Push of one cell's post-collision populations.  x, y are wave-uniform, so
 * the nine destination rows (one per (cx, cy)) with their periodic wrap or
 * wall test are set up once, in scalar registers; per direction the push
 * adds the population's stretch offset and one of the lane's three z
 * offsets.  A population leaving through a wall bounces back into its own
 * cell as the opposite population. */
template <typename S, bool PZ>
struct Pusher {
    char* out;         /* fout as bytes */
    int64_t row[9];    /* byte offset of the destination row of class (cx+1)*3 + (cy+1) */
    bool wall[9];      /* that row lies beyond a non periodic x or y boundary */
    int64_t qb;        /* bytes of one population stretch (zc cells) */
    uint32_t zo[3];    /* lane's byte offset of z-1, z, z+1 inside a row */
    bool zin[3];       /* lane: z-1, z+1 inside the domain (always, for periodic z) */

    __device__ __forceinline__ Pusher(S* fout, const PopLayout& L, int x, int y, int z, int nx, int ny,
                                      int nz, int px, int py) {
        out = (char*)fout;
        const int64_t rowbytes = L.rowlen() * (int64_t)sizeof(S);
        int64_t xo[3], yo[3];
        bool xw[3], yw[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int xd = x + d - 1, yd = y + d - 1;
            xw[d] = yw[d] = false;
            if (xd < 0 || xd >= nx) {
                if (px) xd = xd < 0 ? xd + nx : xd - nx;
                else { xw[d] = true; xd = x; }
            }
            if (yd < 0 || yd >= ny) {
                if (py) yd = yd < 0 ? yd + ny : yd - ny;
                else { yw[d] = true; yd = y; }
            }
            xo[d] = (int64_t)xd * ny * rowbytes;
            yo[d] = (int64_t)yd * rowbytes;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            row[k] = xo[k / 3] + yo[k % 3];
            wall[k] = xw[k / 3] || yw[k % 3];
        }
        qb = (int64_t)L.zc * (int64_t)sizeof(S);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int zd = z + d - 1;
            zin[d] = true;
            if (zd < 0 || zd >= nz) {
                if (PZ) zd = zd < 0 ? zd + nz : zd - nz;
                else { zin[d] = false; zd = z; }
            }
            zo[d] = L.zoff(zd) * (uint32_t)sizeof(S);
        }
    }
    /* any wall test of this wave can succeed (x/y boundary row, or a z end
     * chunk of a non periodic z) */
    __device__ __forceinline__ bool near_wall(int chunk, int nzc) const {
        bool w = false;
#pragma unroll
        for (int k = 0; k < 9; ++k) w = w || wall[k];
        if (!PZ) w = w || chunk == 0 || chunk == nzc - 1;
        return w;
    }
    __device__ __forceinline__ void put(int i, S v) const {
        const int k = (cx(i) + 1) * 3 + (cy(i) + 1), m = cz(i) + 1;
        const int64_t back = row[4] + opp(i) * qb + zo[1];
        const int64_t dst = wall[k] ? back : row[k] + i * qb + zo[m];
        if (PZ || m == 1 || zin[m]) *(S*)(out + dst) = v;
        else *(S*)(out + back) = v;
    }
    /* no wall test can succeed: the destination row and the lane's z offset */
    __device__ __forceinline__ void put_inner(int i, S v) const {
        const int k = (cx(i) + 1) * 3 + (cy(i) + 1), m = cz(i) + 1;
        *(S*)(out + (row[k] + i * qb) + zo[m]) = v;
    }
};

/* FP32 arithmetic of the same collision, restructured for FMAs: with
 * K1 = 3 W, K2 = 3 p W and base = W (rhoBar - p |j|^2) per weight class,
 * eq_i = base + cj (K1 + K2 cj) and eq_{i+9} = base + cj (K2 cj - K1). */
struct EqCoef {
    float base1, k1a, k1b, base2, k2a, k2b, base0;
    __device__ __forceinline__ void init(float scale, float rhoBar, float p, float jSqr) {
        const float w0 = 1.0f / 3.0f, w1 = 1.0f / 18.0f, w2 = 1.0f / 36.0f;
        const float r0 = rhoBar - p * jSqr;
        base0 = scale * w0 * r0;
        base1 = scale * w1 * r0;
        base2 = scale * w2 * r0;
        k1a = scale * 3.0f * w1;
        k2a = scale * 3.0f * w2;
        k1b = k1a * p;
        k2b = k2a * p;
    }
    __device__ __forceinline__ void pair(int i, float cj, float& ei, float& eo) const {
        const bool c1 = (i <= 3);
        const float base = c1 ? base1 : base2, ka = c1 ? k1a : k2a, kb = c1 ? k1b : k2b;
        ei = fmaf(cj, fmaf(kb, cj, ka), base);
        eo = fmaf(cj, fmaf(kb, cj, -ka), base);
    }
};

template <bool FORCE>
__device__ __forceinline__ void bgk_all_fast(const float* f, float* fb, float omega, const float* acc,
                                             float& rhoBar, float* j) {
    moments<float>(f, rhoBar, j);
    const float rho = rhoBar + 1.0f;
    const float invRho = recip<float>(rho);
    float jEq[3] = {j[0], j[1], j[2]};
    float u[3] = {0.0f, 0.0f, 0.0f};
    if (FORCE) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            u[d] = j[d] * invRho + 0.5f * acc[d];
            jEq[d] = rho * u[d];
        }
    }
    const float jSqr = norm_sqr<float>(jEq);
    const float omm = 1.0f - omega;
    EqCoef ec;
    ec.init(omega, rhoBar, 1.5f * invRho, jSqr);
    fb[0] = fmaf(f[0], omm, ec.base0);
#pragma unroll
    for (int i = 1; i <= 9; ++i) {
        float ei, eo;
        ec.pair(i, cdot<float>(i, jEq), ei, eo);
        fb[i] = fmaf(f[i], omm, ei);
        fb[i + 9] = fmaf(f[i + 9], omm, eo);
    }
    if (FORCE) {
        const float pref = (1.0f - 0.5f * omega) * rho;
        const float ua = (u[0] * acc[0] + u[1] * acc[1]) + u[2] * acc[2];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const float ca = cdot<float>(i, acc);
            const float cu = cdot<float>(i, u);
            const float g = 3.0f * (ca - ua) + 9.0f * (cu * ca);
            fb[i] = fb[i] + (pref * wgt<float>(i)) * g;
        }
    }
}

/* The solid blend in FP32 (same structure as the FP64 path). */
__device__ __forceinline__ void psm_blend_fast(const float* fp, float* fo, bool full, float B,
                                               const float* us, float rhoP, const float* jP, float* F) {
    const float oneMinB = 1.0f - B;
    /* moments of the cell the blend sees: the pre-collision cell (already
     * summed by the collision) when full, the post-collision cell otherwise.
     * With B = 1 the (1 - B) terms below vanish, so fo serves both cases. */
    float rhoB, jB[3];
    moments<float>(fo, rhoB, jB);
    if (full) {
        rhoB = rhoP;
        jB[0] = jP[0];
        jB[1] = jP[1];
        jB[2] = jP[2];
    }
    const float invRho = recip<float>(rhoB + 1.0f);
    const float jSqr = norm_sqr<float>(jB);
    const float p = 1.5f * invRho;
    const float onePlus = 1.0f + rhoB;
    const float uP[3] = {us[0] * onePlus, us[1] * onePlus, us[2] * onePlus};
    const float uPSqr = norm_sqr<float>(uP);
    /* Pair (i, o = i+9) with cj = c_i.j, cu = c_i.uP, per weight class W:
     * E_i/o = base_j +- K1 cj + K2 cj^2, S_i/o = base_u +- K1 cu + K2 cu^2
     * (K1 = 3W, K2 = 3pW, base = W (rhoB - p |.|^2)).  The solid terms
     *   bb  = B ((f_o - E_o) - (f_i - S_i)) = B (sym + asym),
     *   bbo = B ((f_i - E_i) - (f_o - S_o)) = B (sym - asym),
     * with asym = (f_o - f_i) + K1 (cj + cu) and
     * sym = (base_u - base_j) + K2 (cu - cj)(cu + cj); bb - bbo = 2 B asym. */
    const float w1 = 1.0f / 18.0f, w2 = 1.0f / 36.0f;
    const float dq = p * (jSqr - uPSqr);          /* (base_u - base_j) / W */
    const float twoB = B + B;
    F[0] = F[1] = F[2] = 0.0f;
    /* rest population: B (E_0 - S_0) = -B W0 dq */
    fo[0] = fmaf(oneMinB, fo[0] - fp[0], fp[0]) - B * ((1.0f / 3.0f) * dq);
#pragma unroll
    for (int i = 1; i <= 9; ++i) {
        const int o = i + 9;
        const float W = i <= 3 ? w1 : w2;
        const float K1 = 3.0f * W, K2 = K1 * p, dbase = W * dq;
        const float cj = cdot<float>(i, jB), cu = cdot<float>(i, uP);
        const float sum = cj + cu;
        const float sym = fmaf(K2, (cu - cj) * sum, dbase);
        const float asym = fmaf(K1, sum, fp[o] - fp[i]);
        /* f + (1 - B)(fB - f): exactly f for a full cell (B = 1) */
        fo[i] = fmaf(oneMinB, fo[i] - fp[i], fp[i]) + B * (sym + asym);
        fo[o] = fmaf(oneMinB, fo[o] - fp[o], fp[o]) + B * (sym - asym);
        const float diff = twoB * asym;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (cc(i, d) == 1) F[d] = F[d] - diff;
            else if (cc(i, d) == -1) F[d] = F[d] + diff;
        }
    }
}

/* One lane per cell, one workgroup per (row, z chunk) block of the store
 * (blockDim.x = chunk length), the chunk index fastest: the grid sweeps the
 * store once, in address order. */
template <int PREC, bool FORCE, bool PZ>
__global__ __launch_bounds__(256) void k_collide_stream(const typename Store<PREC>::S* __restrict__ fin,
                                                        typename Store<PREC>::S* __restrict__ fout,
                                                        CollideArgs a) {
    typedef typename Store<PREC>::T T;
    typedef typename Store<PREC>::S S;
    const int nzc = a.L.nzc;
    const int wg = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int blk = wg >> a.sub_shift;
    const int row = (int)fdiv((uint32_t)blk, a.div_nzc);
    const int chunk = blk - row * nzc;
    const int zl = (wg - (blk << a.sub_shift)) * (int)blockDim.x + (int)threadIdx.x;
    const int z = chunk * a.L.zc + zl;
    if (z >= a.nz) return;
    /* x, y are wave-uniform: keep them (and the row-block addresses derived
     * from them) in scalar registers */
    const int x = __builtin_amdgcn_readfirstlane((int)fdiv((uint32_t)row, a.div_ny));
    const int y = __builtin_amdgcn_readfirstlane(row - x * a.ny);
    const T omega = T(a.omega);
    const T acc[3] = {T(a.acc[0]), T(a.acc[1]), T(a.acc[2])};
    Pusher<S, PZ> push(fout, a.L, x, y, z, a.nx, a.ny, a.nz, a.px, a.py);
    /* this workgroup's (row, chunk) block: 19 contiguous stretches of zc */
    const char* frow = (const char*)(fin + pop_at(row, 0, chunk * a.L.zc, a.L));
    const size_t qstride = (size_t)a.L.zc * sizeof(S);

    const int64_t c = rec_index(x, y, z, a.ny, a.nz);
    const uint32_t zoff = (uint32_t)zl * (uint32_t)sizeof(S);
    /* Issue order: the flag byte, the 19 populations, then the cell's record.
     * The flag is back before the populations (loads return in order), so
     * the record loads go out without a stall of their own and are in flight
     * during the BGK collision; waiting on the flag before the populations
     * would add a full memory round trip to every wave. */
    const uint8_t flag = a.sflag[c];
    S raw[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        /* plain loads: at three wavefronts per SIMD streaming loads were
         * 0-2 % slower for FP64 and 1-2 % slower for FP32 / FP16 */
        raw[i] = *(const S*)(frow + i * qstride + zoff);
    }
    __builtin_amdgcn_sched_barrier(0);
    const bool solid = flag != 0;
    /* fluid lanes read record 0 (one shared line); their values are unused */
    const int64_t s = solid ? c : 0;
    /* FP64: the record itself; FP32/FP16: its 16-byte float view (half the
     * bytes of the four doubles, one load) */
    double rsf = 0.0, rus[3] = {0.0, 0.0, 0.0};
    float4 rm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (sizeof(T) == 4) {
        rm = *(const float4*)(a.rm4 + 4 * s);
    } else {
        /* the 32-byte slot in two loads */
        const double2 r0 = *(const double2*)(a.rsus + 4 * s);
        const double2 r1 = *(const double2*)(a.rsus + 4 * s + 2);
        rsf = r0.x;
        rus[0] = r0.y;
        rus[1] = r1.x;
        rus[2] = r1.y;
    }
    __builtin_amdgcn_sched_barrier(0);
    T fp[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fp[i] = (T)raw[i];
    plane_rhobar<T>(a, fp, x, y, z);
    T fo[Q];
    double Fc[3] = {0.0, 0.0, 0.0};
    if constexpr (sizeof(T) == 4) {
        float rhoP, jP[3];
        bgk_all_fast<FORCE>(fp, fo, omega, acc, rhoP, jP);
        if (solid) {
            const float sf = rm.x;
            const bool full = rm.x > 1.5f;   /* the view's mark of sf > SOLFRAC_MAX */
            const float ooo = (float)a.ooo;
            const float B = full ? 1.0f : (sf * ooo) / ((1.0f - sf) + ooo);
            const float us[3] = {rm.y, rm.z, rm.w};
            float F[3];
            psm_blend_fast(fp, fo, full, B, us, rhoP, jP, F);
            Fc[0] = double(F[0]);
            Fc[1] = double(F[1]);
            Fc[2] = double(F[2]);
        }
    } else {
        T rhoP, jP[3];
        bgk_all<T, FORCE>(fp, fo, omega, acc, rhoP, jP);
        if (solid) {
            /* IBcompositeDynamics::collide (src/ibCompositeDynamics3D.hh:130-203).
             * sf > 0.999 runs as B = 1 on the pre-collision cell: with B = 1
             * and fB = fP the blend (fP + (1-B)(fB - fP)) + B*Omega reduces
             * exactly to fP + Omega, the reference's full-solid branch. */
            const T sf = T(rsf);
            const bool full = rsf > SOLFRAC_MAX;
            const T ooo = T(a.ooo);
            const T B = full ? T(1) : (sf * ooo) / ((T(1) - sf) + ooo);
            const T oneMinB = T(1) - B;
            const T us[3] = {T(rus[0]), T(rus[1]), T(rus[2])};
            /* the reference sums the moments of the cell it blends: the
             * pre-collision cell when full (the collision summed those
             * already, in the same order), the post-collision cell otherwise.
             * With B = 1 the (1 - B) terms below vanish, so fo serves both. */
            T rhoB, jB[3];
            moments<T>(fo, rhoB, jB);
            if (full) {
                rhoB = rhoP;
                jB[0] = jP[0];
                jB[1] = jP[1];
                jB[2] = jP[2];
            }
            const T invRho = recip<T>(rhoB + T(1));
            const T jSqr = norm_sqr<T>(jB);
            const T p = T(1.5) * invRho;
            const T onePlus = T(1) + rhoB;
            T uP[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) uP[d] = us[d] * onePlus;
            const T uPSqr = norm_sqr<T>(uP);
            T F[3] = {T(0), T(0), T(0)};
            {
                const T bb0 = eq_rest<T>(rhoB, p, jSqr) - eq_rest<T>(rhoB, p, uPSqr);
                fo[0] = (fp[0] + oneMinB * (fo[0] - fp[0])) + B * bb0;
            }
#pragma unroll
            for (int i = 1; i <= 9; ++i) {
                const int o = i + 9;
                T Ei, Eo, Si, So;
                eq_pair<T>(i, rhoB, p, jB, jSqr, Ei, Eo);
                eq_pair<T>(i, rhoB, p, uP, uPSqr, Si, So);
                const T bb = B * ((fp[o] - Eo) - (fp[i] - Si));
                const T bbo = B * ((fp[i] - Ei) - (fp[o] - So));
                fo[i] = (fp[i] + oneMinB * (fo[i] - fp[i])) + bb;
                fo[o] = (fp[o] + oneMinB * (fo[o] - fp[o])) + bbo;
                const T diff = bb - bbo;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (cc(i, d) == 1) F[d] = F[d] - diff;
                    else if (cc(i, d) == -1) F[d] = F[d] + diff;
                }
            }
            Fc[0] = double(F[0]);
            Fc[1] = double(F[1]);
            Fc[2] = double(F[2]);
        }
    }
    /* The lanes of every 128-byte line of the force slots that holds a solid
     * cell write the force of all their cells, zero for the fluid ones (the
     * reference resets the force of every cell in each collision,
     * src/ibCompositeDynamics3D.hh:137): whole lines, not the partial-line
     * writes of the solid cells alone, which cost a read-modify-write each,
     * and not the whole wavefront either (the lines without a solid cell are
     * left alone).  A slot is 32 bytes in double, 16 in float, so a line is
     * four or eight cells: a short run of solid cells writes little more
     * than its own slots.  Readers look at solid cells only. */
    const unsigned long long solid_bal = __builtin_amdgcn_ballot_w64(solid);
    constexpr int FG = sizeof(T) == 4 ? 8 : 4;   /* lanes per 128-byte line */
    const int lane_id = (int)__lane_id();
    if ((solid_bal >> (lane_id & ~(FG - 1))) & ((1ull << FG) - 1)) {
        if constexpr (sizeof(T) == 4) {
            /* the FP32 arithmetic's force, stored exactly in float: half the
             * bytes here and in the force sum */
            *((float4*)a.rF + c) = make_float4((float)Fc[0], (float)Fc[1], (float)Fc[2], 0.0f);
        } else {
            double* q = (double*)a.rF + 4 * c;
            *(double2*)q = make_double2(Fc[0], Fc[1]);
            *(double2*)(q + 2) = make_double2(Fc[2], 0.0);
        }
    }
    if (push.near_wall(chunk, nzc)) {
#pragma unroll
        for (int i = 0; i < Q; ++i) push.put(i, (S)fo[i]);
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) push.put_inner(i, (S)fo[i]);
    }
}

/* Launches one collide+stream from fin into fout (row-blocked stores of the
 * handle's precision). */
template <int PREC>
void launch_cs_prec(lbdem_psm* h, const CollideArgs& a, const void* fin, void* fout);

}  // namespace psm

#endif

/*
 * C ABI of the gfx950 PSM library: handle life cycle, device memory,
 * host-side argument checks and the host halves of the exchange calls
 * (setSpheresOnLattice / getForcesFromLattice,
 * src/ibDataExchangeWrappers3D.hh:33-246).
 */
#include "psm_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace psm {
size_t scan_temp_bytes(int64_t n);

static std::string g_create_err;

int set_err(lbdem_psm* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_err = msg;
    return code;
}

int hip_check(lbdem_psm* h, hipError_t e, const char* what) {
    std::string m = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return set_err(h, LBDEM_PSM_ERR_OUT_OF_MEMORY, m);
    return set_err(h, LBDEM_PSM_ERR_HIP, m);
}

void prof_begin(lbdem_psm* h, int cls, hipEvent_t* a) {
    (void)cls;
    *a = nullptr;
    if (!h->profiling) return;
    if (h->ev_pool.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        h->ev_pool.push_back(e);
    }
    *a = h->ev_pool.back();
    h->ev_pool.pop_back();
    (void)hipEventRecord(*a, h->stream);
}

void prof_end(lbdem_psm* h, int cls, hipEvent_t a) {
    if (!h->profiling || !a) return;
    hipEvent_t b;
    if (h->ev_pool.empty()) {
        if (hipEventCreate(&b) != hipSuccess) return;
    } else {
        b = h->ev_pool.back();
        h->ev_pool.pop_back();
    }
    (void)hipEventRecord(b, h->stream);
    h->spans.push_back({cls, a, b});
}

template <typename P>
static hipError_t dalloc(P** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    return hipMalloc((void**)p, bytes);
}

static void dfree(void* p) {
    if (p) (void)hipFree(p);
}

static int alloc_records(lbdem_psm* h, Records& r) {
    const int64_t cap = h->rec_cap;
    PSM_HIP(h, dalloc(&r.sus, 4 * cap * sizeof(double)));
    PSM_HIP(h, dalloc(&r.id, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&r.F, 4 * cap * (h->precision == LBDEM_PSM_FP64 ? sizeof(double) : sizeof(float))));
    PSM_HIP(h, dalloc(&r.cand, cap * sizeof(int32_t)));
    r.m4 = nullptr;
    if (h->precision != LBDEM_PSM_FP64) PSM_HIP(h, dalloc(&r.m4, 4 * cap * sizeof(float)));
    return LBDEM_PSM_OK;
}

static void free_records(Records& r) {
    dfree(r.sus); dfree(r.id); dfree(r.F); dfree(r.cand); dfree(r.m4);
    memset(&r, 0, sizeof(r));
}

/* Upper bound of the (image, bin) candidate entries of particles with these
 * radii, valid for any positions. */
static int64_t entry_bound(const lbdem_psm* h, int64_t n, const double* r, bool auto_images) {
    int64_t total = 0;
    for (int64_t q = 0; q < n; ++q) {
        const int64_t ext = (int64_t)std::floor(2.0 * r[q]) + 6;
        int64_t prod = 1;
        for (int d = 0; d < 3; ++d) {
            const int64_t span = (ext + BIN - 1) / BIN + 1;
            int64_t m;
            if (auto_images && h->g.periodic[d]) m = std::min<int64_t>(span + 2, 3 * (int64_t)h->g.nb[d]);
            else m = std::min<int64_t>(span, h->g.nb[d]);
            prod *= m;
        }
        total += prod;
    }
    return total + 1024;
}

static int ensure_particles(lbdem_psm* h, int64_t n) {
    if (n <= h->p_cap && h->px) return LBDEM_PSM_OK;
    const int64_t cap = std::max<int64_t>(n, std::max<int64_t>(2 * h->p_cap, 256));
    /* queued kernels may still use the buffers being replaced */
    if (h->stream) PSM_HIP(h, hipStreamSynchronize(h->stream));
    h->ids_valid = false;
    dfree(h->px); dfree(h->pv); dfree(h->pw); dfree(h->pr); dfree(h->pid); dfree(h->pexcl);
    dfree(h->powner); dfree(h->p_nent); dfree(h->p_ent_start); dfree(h->pft); dfree(h->psum_x);
    dfree(h->orph_cnt); dfree(h->orph_fill); dfree(h->orph_start);
    h->orph_cnt = h->orph_fill = h->orph_start = nullptr;
    PSM_HIP(h, dalloc(&h->orph_cnt, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->orph_fill, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->orph_start, cap * sizeof(int32_t)));
    /* the grouping kernels keep these zero between uses */
    PSM_HIP(h, hipMemset(h->orph_cnt, 0, cap * sizeof(int32_t)));
    PSM_HIP(h, hipMemset(h->orph_fill, 0, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->px, 3 * cap * sizeof(double)));
    PSM_HIP(h, dalloc(&h->pv, 3 * cap * sizeof(double)));
    PSM_HIP(h, dalloc(&h->pw, 3 * cap * sizeof(double)));
    PSM_HIP(h, dalloc(&h->pr, cap * sizeof(double)));
    PSM_HIP(h, dalloc(&h->pid, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->pexcl, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->powner, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->p_nent, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->p_ent_start, cap * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->pft, 6 * cap * sizeof(double)));
    PSM_HIP(h, dalloc(&h->psum_x, 3 * cap * sizeof(double)));
    h->p_cap = cap;
    const size_t need = std::max(scan_temp_bytes(h->nbins), scan_temp_bytes(cap));
    if (need > h->scan_tmp_bytes) {
        dfree(h->scan_tmp);
        PSM_HIP(h, dalloc(&h->scan_tmp, need));
        h->scan_tmp_bytes = need;
    }
    return LBDEM_PSM_OK;
}

static int ensure_entries(lbdem_psm* h, int64_t need) {
    if (need > (int64_t)INT32_MAX)
        return set_err(h, LBDEM_PSM_ERR_CAPACITY, "candidate entries exceed 2^31");
    if (need <= h->ent_cap && h->cand_key) return LBDEM_PSM_OK;
    if (h->stream) PSM_HIP(h, hipStreamSynchronize(h->stream));
    dfree(h->cand_key); dfree(h->cand_bin); dfree(h->owner_ent); dfree(h->partial); dfree(h->cand_data);
    PSM_HIP(h, dalloc(&h->cand_data, need * sizeof(Cand)));
    PSM_HIP(h, dalloc(&h->cand_key, need * sizeof(uint32_t)));
    PSM_HIP(h, dalloc(&h->cand_bin, need * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->owner_ent, need * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&h->partial, 6 * need * sizeof(double)));
    h->ent_cap = need;
    return LBDEM_PSM_OK;
}

static int read_counters(lbdem_psm* h) {
    PSM_HIP(h, hipMemcpyAsync(h->h_counters, h->counters, 8 * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    return LBDEM_PSM_OK;
}

/* The error flags of a fresh copy of the device counters in h_counters. */
static int counters_status(lbdem_psm* h) {
    h->pending_counters = false;
    if (h->h_counters[1] & 2)
        return set_err(h, LBDEM_PSM_ERR_CAPACITY, "candidate-entry capacity exceeded; lattice state is invalid");
    if (h->h_counters[1] & 4)
        return set_err(h, LBDEM_PSM_ERR_CAPACITY, "too many stale records; forces incomplete");
    if (h->h_counters[1] & 8)
        return set_err(h, LBDEM_PSM_ERR_CAPACITY, "too many initVel cells");
    if (h->max_records > 0 && h->h_counters[0] > h->max_records)
        return set_err(h, LBDEM_PSM_ERR_CAPACITY, "coupling records exceed max_solid_cells");
    return LBDEM_PSM_OK;
}

static int check_overflow(lbdem_psm* h) {
    if (h->max_records > 0) {
        int rc0 = count_records(h);
        if (rc0) return rc0;
    }
    int rc = read_counters(h);
    if (rc) return rc;
    return counters_status(h);
}

/* After a stream synchronization: the flags of the counters a set_spheres
 * call copied without waiting. */
static int pending_status(lbdem_psm* h) {
    return h->pending_counters ? counters_status(h) : LBDEM_PSM_OK;
}

/* A pinned staging buffer of at least `bytes`, the older of the two; waits
 * for the copies still reading it. */
static int stage_acquire(lbdem_psm* h, size_t bytes, int* which) {
    const int k = h->stage_next;
    h->stage_next ^= 1;
    lbdem_psm::Stage& s = h->stage[k];
    if (s.inflight) {
        PSM_HIP(h, hipEventSynchronize(s.ev));
        s.inflight = false;
    }
    if (!s.ev) PSM_HIP(h, hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    if (bytes > s.bytes) {
        if (s.buf) (void)hipHostFree(s.buf);
        s.buf = nullptr;
        s.bytes = 0;
        PSM_HIP(h, hipHostMalloc(&s.buf, bytes, hipHostMallocDefault));
        s.bytes = bytes;
    }
    *which = k;
    return LBDEM_PSM_OK;
}

static int stage_release(lbdem_psm* h, int which) {
    PSM_HIP(h, hipEventRecord(h->stage[which].ev, h->stream));
    h->stage[which].inflight = true;
    return LBDEM_PSM_OK;
}

static int io_buffer(lbdem_psm* h, size_t bytes) {
    if (bytes <= h->h_io_bytes) return LBDEM_PSM_OK;
    if (h->h_io) (void)hipHostFree(h->h_io);
    h->h_io = nullptr;
    h->h_io_bytes = 0;
    PSM_HIP(h, hipHostMalloc((void**)&h->h_io, bytes, hipHostMallocDefault));
    h->h_io_bytes = bytes;
    return LBDEM_PSM_OK;
}

static int fold_forces_to_host(lbdem_psm* h, int64_t n, double* F3, double* T3) {
    int rc = io_buffer(h, 6 * (size_t)std::max<int64_t>(n, 1) * sizeof(double));
    if (rc) return rc;
    const double* pft = h->h_io;
    if (n > 0) PSM_HIP(h, hipMemcpyAsync(h->h_io, h->pft, 6 * n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    rc = pending_status(h);
    if (rc) return rc;
    /* getForcesFromLattice: f[map(tag[i])] += force[i]; the ghost slots end 0. */
    for (int64_t q = 0; q < n; ++q)
        for (int d = 0; d < 3; ++d) {
            if (F3) F3[3 * q + d] = 0.0;
            if (T3) T3[3 * q + d] = 0.0;
        }
    for (int64_t q = 0; q < n; ++q) {
        const int64_t o = h->h_owner[q];
        for (int d = 0; d < 3; ++d) {
            if (F3) F3[3 * o + d] += pft[6 * q + d];
            if (T3) T3[3 * o + d] += pft[6 * q + 3 + d];
        }
    }
    return LBDEM_PSM_OK;
}

}  // namespace psm

using namespace psm;

extern "C" {

const char* lbdem_psm_version(void) { return "lbdem_psm 0.1.0 (gfx950 HIP)"; }

/* Error message setter for the host-only units of the library. */
__attribute__((visibility("hidden"))) int lbdem_psm_internal_set_error(lbdem_psm* h, int code, const char* msg) {
    return set_err(h, code, msg ? std::string(msg) : std::string());
}

int lbdem_psm_desc_init(lbdem_psm_desc* d) {
    if (!d) return LBDEM_PSM_ERR_INVALID_ARG;
    memset(d, 0, sizeof(*d));
    d->nx = d->ny = d->nz = 1;
    d->periodic[0] = d->periodic[1] = d->periodic[2] = 1;
    d->tau = 1.0;
    d->precision = LBDEM_PSM_FP64;
    d->slices_per_dim = 5;
    return LBDEM_PSM_OK;
}

int lbdem_psm_check_desc(const lbdem_psm_desc* d) {
    if (!d) return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "desc is NULL");
    if (d->nx < 1 || d->ny < 1 || d->nz < 1)
        return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "lattice sizes must be >= 1");
    const int64_t nc = (int64_t)d->nx * d->ny * d->nz;
    if (nc > (int64_t)INT32_MAX)
        return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "more than 2^31-1 cells");
    if (!(d->tau > 0.5) || !std::isfinite(d->tau))
        return set_err(nullptr, LBDEM_PSM_ERR_NUMERIC, "tau must be finite and > 0.5 (omega < 2)");
    if (d->precision < LBDEM_PSM_FP64 || d->precision > LBDEM_PSM_FP16_STORE)
        return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "unknown precision");
    if (d->slices_per_dim < 1 || d->slices_per_dim > 16)
        return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "slices_per_dim must be in [1,16]");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(d->body_accel[k]))
            return set_err(nullptr, LBDEM_PSM_ERR_NUMERIC, "body_accel must be finite");
    if (d->voxel_mode != LBDEM_PSM_VOXEL_SUBCELL && d->voxel_mode != LBDEM_PSM_VOXEL_OVERLAP)
        return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "unknown voxel_mode");
    if (d->max_solid_cells < 0 || d->max_solid_cells > nc)
        return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "max_solid_cells must be in [0, cells]");
    return LBDEM_PSM_OK;
}

int lbdem_psm_check_spheres(const lbdem_psm_desc* d, int64_t n, const double* x3, const double* v3,
                            const double* w3, const double* r, const int64_t* id) {
    if (!d) return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "desc is NULL");
    if (n < 0) return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "negative particle count");
    if (n >= (int64_t)QMASK) return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "at most 2^27-2 particles");
    if (n > 0 && (!x3 || !v3 || !w3 || !r || !id))
        return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "NULL particle array");
    const int nn[3] = {d->nx, d->ny, d->nz};
    for (int64_t q = 0; q < n; ++q) {
        if (!(r[q] > 0.0) || !std::isfinite(r[q]))
            return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "particle radius must be finite and > 0");
        if (id[q] < 1 || id[q] >= (int64_t)INT32_MAX)
            return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "particle ids must be in [1, 2^31-1)");
        for (int k = 0; k < 3; ++k) {
            const double x = x3[3 * q + k];
            if (!std::isfinite(x) || !std::isfinite(v3[3 * q + k]) || !std::isfinite(w3[3 * q + k]))
                return set_err(nullptr, LBDEM_PSM_ERR_NUMERIC, "non-finite particle state");
            if (x < -(double)nn[k] || x >= 2.0 * nn[k])
                return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "particle outside [-n, 2n) of the lattice");
            if (std::fabs(v3[3 * q + k]) >= 0.3)
                return set_err(nullptr, LBDEM_PSM_ERR_NUMERIC, "particle velocity >= 0.3 lattice units");
        }
        if (r[q] + 2.0 >= 1e6) return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "radius too large");
    }
    return LBDEM_PSM_OK;
}

int lbdem_psm_create(const lbdem_psm_desc* desc, lbdem_psm** out) {
    if (!out) return set_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int rc = lbdem_psm_check_desc(desc);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(nullptr, LBDEM_PSM_ERR_NO_DEVICE, "no HIP device available");
    if (desc->device < 0 || desc->device >= ndev)
        return set_err(nullptr, LBDEM_PSM_ERR_NO_DEVICE, "device ordinal out of range");
    if (hipSetDevice(desc->device) != hipSuccess)
        return set_err(nullptr, LBDEM_PSM_ERR_NO_DEVICE, "hipSetDevice failed");

    lbdem_psm* h = new lbdem_psm();
    h->desc = *desc;
    const int nn[3] = {desc->nx, desc->ny, desc->nz};
    for (int d = 0; d < 3; ++d) {
        h->g.n[d] = nn[d];
        h->g.periodic[d] = desc->periodic[d] ? 1 : 0;
        h->g.nb[d] = (nn[d] + BIN - 1) / BIN;
    }
    h->g.nc = (int64_t)nn[0] * nn[1] * nn[2];
    h->g.L = make_layout(nn[2]);
    h->nbins = (int64_t)h->g.nb[0] * h->g.nb[1] * h->g.nb[2];
    h->omega = 1.0 / desc->tau;
    h->ooo = 1.0 / h->omega - 0.5;          /* src/ibCompositeDynamics3D.hh:178 */
    h->sqrt3half = std::sqrt(3.1) / 2.0;    /* src/ibProcessors3D.hh:127 */
    {
        const double pi = 3.14159265358979323846;
        h->cell_rc = std::cbrt(3.0 / (4.0 * pi));
        h->cell_vc = ((4.0 / 3.0) * pi) * ((h->cell_rc * h->cell_rc) * h->cell_rc);
    }
    h->has_force = desc->body_accel[0] != 0.0 || desc->body_accel[1] != 0.0 || desc->body_accel[2] != 0.0;
    h->precision = desc->precision;
    const size_t S = h->precision == LBDEM_PSM_FP64 ? 8 : (h->precision == LBDEM_PSM_FP32_DEV ? 4 : 2);
    h->pop_bytes = (size_t)nn[0] * (size_t)nn[1] * (size_t)h->g.L.rowlen() * S;
    h->rec_cap = h->g.nc;                        /* one record slot per cell */
    h->max_records = desc->max_solid_cells;
    h->orphan_cap = std::min<int64_t>(h->rec_cap, (int64_t)1 << 24);
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { rc = hip_check(nullptr, e, "hipStreamCreate"); delete h; return rc; }
    std::string msg;
#define CREATE_CHECK(call)                                                   \
    do {                                                                     \
        hipError_t ee = (call);                                              \
        if (ee != hipSuccess) {                                              \
            rc = hip_check(nullptr, ee, #call);                              \
            msg = g_create_err;                                              \
            lbdem_psm_destroy(h);                                            \
            g_create_err = msg;                                              \
            return rc;                                                       \
        }                                                                    \
    } while (0)
    CREATE_CHECK(dalloc(&h->f[0], h->pop_bytes));
    CREATE_CHECK(dalloc(&h->f[1], h->pop_bytes));
    CREATE_CHECK(dalloc(&h->sflag, h->g.nc * sizeof(uint8_t)));
    CREATE_CHECK(hipMemset(h->sflag, 0, h->g.nc * sizeof(uint8_t)));
    for (int k = 0; k < 2; ++k) {
        CREATE_CHECK(dalloc(&h->bin_any[k], h->nbins * sizeof(uint32_t)));
        CREATE_CHECK(hipMemset(h->bin_any[k], 0, h->nbins * sizeof(uint32_t)));
    }
    CREATE_CHECK(dalloc(&h->orphan, h->orphan_cap * sizeof(int32_t)));
    CREATE_CHECK(dalloc(&h->orph_slot, h->orphan_cap * sizeof(int32_t)));
    CREATE_CHECK(dalloc(&h->orph_sorted, h->orphan_cap * sizeof(int32_t)));
    CREATE_CHECK(dalloc(&h->bin_cand_count, h->nbins * sizeof(int32_t)));
    CREATE_CHECK(dalloc(&h->bin_cand_start, h->nbins * sizeof(int32_t)));
    CREATE_CHECK(dalloc(&h->bin_fill, h->nbins * sizeof(int32_t)));
    CREATE_CHECK(hipMemset(h->bin_cand_count, 0, h->nbins * sizeof(int32_t)));
    CREATE_CHECK(hipMemset(h->bin_cand_start, 0, h->nbins * sizeof(int32_t)));
    CREATE_CHECK(dalloc(&h->counters, 8 * sizeof(int64_t)));
    CREATE_CHECK(hipMemset(h->counters, 0, 8 * sizeof(int64_t)));
    CREATE_CHECK(hipHostMalloc((void**)&h->h_counters, 8 * sizeof(int64_t), hipHostMallocDefault));
    CREATE_CHECK(dalloc(&h->red_tmp, (6 * 4096 + 8) * sizeof(double)));
    rc = alloc_records(h, h->rec);
    if (rc) { msg = h->err; lbdem_psm_destroy(h); g_create_err = msg; return rc; }
    rc = ensure_particles(h, 256);
    if (rc) { msg = h->err; lbdem_psm_destroy(h); g_create_err = msg; return rc; }
    rc = ensure_entries(h, 1024);
    if (rc) { msg = h->err; lbdem_psm_destroy(h); g_create_err = msg; return rc; }
    rc = launch_init_equilibrium(h, nullptr, nullptr);
    if (rc) { msg = h->err; lbdem_psm_destroy(h); g_create_err = msg; return rc; }
    CREATE_CHECK(hipStreamSynchronize(h->stream));
#undef CREATE_CHECK
    *out = h;
    return LBDEM_PSM_OK;
}

int lbdem_psm_destroy(lbdem_psm* h) {
    if (!h) return LBDEM_PSM_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    diff_free(h);
    dfree(h->f[0]); dfree(h->f[1]); dfree(h->sflag);
    dfree(h->bin_any[0]); dfree(h->bin_any[1]); dfree(h->cand_data);
    free_records(h->rec);
    dfree(h->px); dfree(h->pv); dfree(h->pw); dfree(h->pr); dfree(h->pid); dfree(h->pexcl);
    dfree(h->powner); dfree(h->id_map); dfree(h->bin_cand_count); dfree(h->bin_cand_start);
    dfree(h->bin_fill); dfree(h->cand_key); dfree(h->cand_bin); dfree(h->p_nent);
    dfree(h->p_ent_start); dfree(h->owner_ent); dfree(h->partial); dfree(h->pft); dfree(h->psum_x);
    dfree(h->scan_tmp); dfree(h->counters); dfree(h->red_tmp); dfree(h->orphan);
    dfree(h->orph_cnt); dfree(h->orph_fill); dfree(h->orph_start); dfree(h->orph_slot); dfree(h->orph_sorted);
    dfree(h->dv_cell); dfree(h->dv_us); dfree(h->pp.buf); dfree(h->pp.cells);
    if (h->h_counters) (void)hipHostFree(h->h_counters);
    for (auto& st : h->stage) {
        if (st.buf) (void)hipHostFree(st.buf);
        if (st.ev) (void)hipEventDestroy(st.ev);
    }
    if (h->h_io) (void)hipHostFree(h->h_io);
    for (auto& s : h->spans) { (void)hipEventDestroy(s.a); (void)hipEventDestroy(s.b); }
    for (auto e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return LBDEM_PSM_OK;
}

const char* lbdem_psm_last_error(const lbdem_psm* h) {
    return h ? h->err.c_str() : g_create_err.c_str();
}

int lbdem_psm_init_equilibrium(lbdem_psm* h, const double* rho, const double* u) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    const int64_t nc = h->g.nc;
    double *d_rho = nullptr, *d_u = nullptr;
    if (rho) {
        PSM_HIP(h, dalloc(&d_rho, nc * sizeof(double)));
        PSM_HIP(h, hipMemcpy(d_rho, rho, nc * sizeof(double), hipMemcpyHostToDevice));
    }
    if (u) {
        PSM_HIP(h, dalloc(&d_u, 3 * nc * sizeof(double)));
        PSM_HIP(h, hipMemcpy(d_u, u, 3 * nc * sizeof(double), hipMemcpyHostToDevice));
    }
    int rc = launch_init_equilibrium(h, d_rho, d_u);
    hipError_t e = hipStreamSynchronize(h->stream);
    dfree(d_rho);
    dfree(d_u);
    if (rc) return rc;
    PSM_HIP(h, e);
    return LBDEM_PSM_OK;
}

int lbdem_psm_set_populations(lbdem_psm* h, const double* f) {
    if (!h || !f) return LBDEM_PSM_ERR_INVALID_ARG;
    const size_t bytes = (size_t)Q * h->g.nc * sizeof(double);
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    double* tmp = nullptr;
    PSM_HIP(h, dalloc(&tmp, bytes));
    PSM_HIP(h, hipMemcpy(tmp, f, bytes, hipMemcpyHostToDevice));
    int rc = launch_pop_convert(h, tmp, true);
    hipError_t e = hipStreamSynchronize(h->stream);
    dfree(tmp);
    if (rc) return rc;
    PSM_HIP(h, e);
    return LBDEM_PSM_OK;
}

int lbdem_psm_get_populations(lbdem_psm* h, double* f) {
    if (!h || !f) return LBDEM_PSM_ERR_INVALID_ARG;
    const size_t bytes = (size_t)Q * h->g.nc * sizeof(double);
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    double* tmp = nullptr;
    PSM_HIP(h, dalloc(&tmp, bytes));
    int rc = launch_pop_convert(h, tmp, false);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (!rc && e == hipSuccess) e = hipMemcpy(f, tmp, bytes, hipMemcpyDeviceToHost);
    dfree(tmp);
    if (rc) return rc;
    PSM_HIP(h, e);
    return LBDEM_PSM_OK;
}

int lbdem_psm_set_spheres(lbdem_psm* h, int64_t n, const double* x3, const double* v3,
                          const double* w3, const double* r, const int64_t* id, const int32_t* type,
                          const int32_t* exclude, int32_t n_exclude, int32_t flags) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    int rc = lbdem_psm_check_spheres(&h->desc, n, x3, v3, w3, r, id);
    if (rc) return set_err(h, rc, g_create_err);
    if (n_exclude < 0 || (n_exclude > 0 && !exclude))
        return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "bad exclude list");
    /* Stream-ordered: the particle arrays go up through pinned staging with
     * asynchronous copies on the handle's stream, the voxelization is queued
     * behind them, and nothing waits for the device (unless a buffer must
     * grow, or max_solid_cells asks for the record count now). */
    rc = ensure_particles(h, n);
    if (rc) return rc;
    const bool auto_images = (flags & LBDEM_PSM_AUTO_IMAGES) != 0;
    rc = ensure_entries(h, entry_bound(h, n, r, auto_images));
    if (rc) return rc;

    /* the map(tag) inputs: rebuilt and uploaded only when they changed */
    const bool same_ids = h->ids_valid && (int64_t)h->h_ids.size() == n &&
                          (n == 0 || std::memcmp(h->h_ids.data(), id, (size_t)n * sizeof(int64_t)) == 0) &&
                          h->h_has_types == (type != nullptr) &&
                          (!type || std::memcmp(h->h_types.data(), type, (size_t)n * sizeof(int32_t)) == 0) &&
                          (int64_t)h->h_exclude.size() == n_exclude &&
                          (n_exclude == 0 || std::memcmp(h->h_exclude.data(), exclude, (size_t)n_exclude * sizeof(int32_t)) == 0);
    int64_t max_id = 0;
    if (!same_ids)
        for (int64_t q = 0; q < n; ++q) max_id = std::max<int64_t>(max_id, id[q]);
    const size_t nd = 10 * (size_t)n;                      /* x, v, w (3 n each), r (n) */
    const size_t ni = same_ids ? 0 : 3 * (size_t)n + (size_t)(max_id + 1);   /* pid, excl, owner, id map */
    int sk = 0;
    rc = stage_acquire(h, std::max<size_t>(8, nd * sizeof(double) + ni * sizeof(int32_t)), &sk);
    if (rc) return rc;
    double* sd = (double*)h->stage[sk].buf;
    std::memcpy(sd, x3, 3 * (size_t)n * sizeof(double));
    std::memcpy(sd + 3 * n, v3, 3 * (size_t)n * sizeof(double));
    std::memcpy(sd + 6 * n, w3, 3 * (size_t)n * sizeof(double));
    std::memcpy(sd + 9 * n, r, (size_t)n * sizeof(double));
    if (n > 0) {
        PSM_HIP(h, hipMemcpyAsync(h->px, sd, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        PSM_HIP(h, hipMemcpyAsync(h->pv, sd + 3 * n, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        PSM_HIP(h, hipMemcpyAsync(h->pw, sd + 6 * n, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        PSM_HIP(h, hipMemcpyAsync(h->pr, sd + 9 * n, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
    if (!same_ids) {
        if (max_id + 1 > h->id_map_size) {
            PSM_HIP(h, hipStreamSynchronize(h->stream));
            dfree(h->id_map);
            h->id_map = nullptr;
            h->id_map_size = 0;
            PSM_HIP(h, dalloc(&h->id_map, (max_id + 1) * sizeof(int32_t)));
            h->id_map_size = max_id + 1;
        }
        int32_t* pid = (int32_t*)(sd + nd);
        int32_t* excl = pid + n;
        int32_t* owner = excl + n;
        int32_t* idmap = owner + n;
        std::fill(idmap, idmap + (max_id + 1), -1);
        /* map(tag): the first index holding each id (owned atoms precede ghosts) */
        h->h_owner.assign(n, 0);
        h->owner_identity = true;
        for (int64_t q = 0; q < n; ++q) {
            if (idmap[id[q]] < 0) idmap[id[q]] = (int32_t)q;
            owner[q] = idmap[id[q]];
            h->h_owner[q] = owner[q];
            h->owner_identity = h->owner_identity && owner[q] == q;
            pid[q] = (int32_t)id[q];
            excl[q] = 0;
            if (type)
                for (int t = 0; t < n_exclude; ++t)
                    if (type[q] == exclude[t]) { excl[q] = 1; break; }
        }
        if (n > 0) {
            PSM_HIP(h, hipMemcpyAsync(h->pid, pid, n * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
            PSM_HIP(h, hipMemcpyAsync(h->pexcl, excl, n * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
            PSM_HIP(h, hipMemcpyAsync(h->powner, owner, n * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
        }
        PSM_HIP(h, hipMemsetAsync(h->id_map, 0xff, h->id_map_size * sizeof(int32_t), h->stream));
        PSM_HIP(h, hipMemcpyAsync(h->id_map, idmap, (max_id + 1) * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
        h->h_ids.assign(id, id + n);
        h->h_has_types = type != nullptr;
        if (type) h->h_types.assign(type, type + n);
        else h->h_types.clear();
        h->h_exclude.assign(exclude, exclude + n_exclude);
        h->ids_valid = true;
    }
    rc = stage_release(h, sk);
    if (rc) return rc;
    h->np = n;
    h->flags = flags;
    h->have_particles = true;
    h->geom_gen += 1;
    if ((flags & LBDEM_PSM_INIT_VEL) && !h->dv_cell) {
        PSM_HIP(h, dalloc(&h->dv_cell, h->orphan_cap * sizeof(int32_t)));
        PSM_HIP(h, dalloc(&h->dv_us, 3 * h->orphan_cap * sizeof(double)));
    }
    PSM_HIP(h, hipMemsetAsync(h->counters + 1, 0, sizeof(int64_t), h->stream));
    rc = run_voxelize(h);
    if (rc) return rc;
    h->voxelized = true;
    if (h->max_records > 0) return check_overflow(h);
    /* the capacity flags are checked by the next call that waits for the
     * device (sum_force_torque, get_particle_forces, step, synchronize) */
    PSM_HIP(h, hipMemcpyAsync(h->h_counters, h->counters, 8 * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    h->pending_counters = true;
    return LBDEM_PSM_OK;
}

int lbdem_psm_collide_stream(lbdem_psm* h, int32_t nsteps) {
    if (!h || nsteps < 0) return LBDEM_PSM_ERR_INVALID_ARG;
    for (int s = 0; s < nsteps; ++s) {
        int rc = launch_collide_stream(h);
        if (rc) return rc;
    }
    return LBDEM_PSM_OK;
}

int lbdem_psm_sum_force_torque(lbdem_psm* h, int64_t n, const double* x3, double* F3, double* T3) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    if (!h->have_particles) return set_err(h, LBDEM_PSM_ERR_STATE, "sum_force_torque before set_spheres");
    if (n != h->np) return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "particle count differs from set_spheres");
    if (n > 0) {
        if (x3) {
            /* through pinned staging, queued behind the collision */
            int sk = 0;
            int rc0 = stage_acquire(h, 3 * (size_t)n * sizeof(double), &sk);
            if (rc0) return rc0;
            std::memcpy(h->stage[sk].buf, x3, 3 * (size_t)n * sizeof(double));
            PSM_HIP(h, hipMemcpyAsync(h->psum_x, h->stage[sk].buf, 3 * n * sizeof(double), hipMemcpyHostToDevice,
                                      h->stream));
            rc0 = stage_release(h, sk);
            if (rc0) return rc0;
        } else {
            PSM_HIP(h, hipMemcpyAsync(h->psum_x, h->px, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        }
    }
    int rc = run_force_sum(h, h->psum_x);
    if (rc) return rc;
    return fold_forces_to_host(h, n, F3, T3);
}

int lbdem_psm_step(lbdem_psm* h, int32_t nsteps, int32_t flags) {
    if (!h || nsteps < 0) return LBDEM_PSM_ERR_INVALID_ARG;
    if (!h->have_particles) return set_err(h, LBDEM_PSM_ERR_STATE, "step before set_spheres");
    const bool frozen = (flags & LBDEM_PSM_STEP_FROZEN) != 0;
    if (frozen && (flags & LBDEM_PSM_STEP_ADVANCE))
        return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "STEP_FROZEN excludes STEP_ADVANCE");
    if (!frozen) h->geom_gen += 1;
    PSM_HIP(h, hipMemsetAsync(h->counters + 1, 0, sizeof(int64_t), h->stream));
    for (int s = 0; s < nsteps; ++s) {
        int rc;
        if (flags & LBDEM_PSM_STEP_ADVANCE) {
            rc = launch_advance(h);
            if (rc) return rc;
        }
        if (!frozen) {
            rc = run_voxelize(h);
            if (rc) return rc;
        }
        rc = launch_collide_stream(h);
        if (rc) return rc;
        rc = run_force_sum(h, h->px);
        if (rc) return rc;
    }
    h->pending_counters = false;
    return check_overflow(h);
}

int lbdem_psm_get_particle_forces(lbdem_psm* h, int64_t n, double* F3, double* T3) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    if (n != h->np) return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "particle count differs from set_spheres");
    return fold_forces_to_host(h, n, F3, T3);
}

int lbdem_psm_get_particles(lbdem_psm* h, int64_t n, double* x3) {
    if (!h || !x3) return LBDEM_PSM_ERR_INVALID_ARG;
    if (n != h->np) return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "particle count differs from set_spheres");
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    if (n > 0) PSM_HIP(h, hipMemcpy(x3, h->px, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    return LBDEM_PSM_OK;
}

int lbdem_psm_get_cell_records(lbdem_psm* h, double* sf, int64_t* id, double* us, double* F) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    const int64_t nc = h->g.nc, cap = h->rec_cap;
    const int n0 = h->g.n[0], n1 = h->g.n[1], n2 = h->g.n[2];
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    std::vector<uint8_t> flag(nc);
    PSM_HIP(h, hipMemcpy(flag.data(), h->sflag, nc, hipMemcpyDeviceToHost));
    std::vector<double> rsus(4 * cap), rF(4 * cap);
    std::vector<int32_t> rid(cap);
    PSM_HIP(h, hipMemcpy(rsus.data(), h->rec.sus, 4 * cap * sizeof(double), hipMemcpyDeviceToHost));
    PSM_HIP(h, hipMemcpy(rid.data(), h->rec.id, cap * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (h->precision == LBDEM_PSM_FP64) {
        PSM_HIP(h, hipMemcpy(rF.data(), h->rec.F, 4 * cap * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        /* FP32/FP16 storage keeps the per-cell force as float */
        std::vector<float> f32(4 * cap);
        PSM_HIP(h, hipMemcpy(f32.data(), h->rec.F, 4 * cap * sizeof(float), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < 4 * cap; ++k) rF[k] = (double)f32[k];
    }
    for (int X = 0; X < n0; ++X)
        for (int Y = 0; Y < n1; ++Y)
            for (int Z = 0; Z < n2; ++Z) {
                const int64_t c = ((int64_t)X * n1 + Y) * n2 + Z;
                const bool live = flag[c] != 0;
                const int64_t r = rec_index(X, Y, Z, n1, n2);
                if (sf) sf[c] = live ? rsus[4 * r] : 0.0;
                if (id) id[c] = live ? rid[r] : 0;
                for (int d = 0; d < 3; ++d) {
                    if (us) us[d * nc + c] = live ? rsus[4 * r + 1 + d] : 0.0;
                    if (F) F[d * nc + c] = live ? rF[4 * r + d] : 0.0;
                }
            }
    return LBDEM_PSM_OK;
}

int lbdem_psm_set_cell_records(lbdem_psm* h, const double* sf, const int64_t* id, const double* us) {
    if (!h || !sf) return LBDEM_PSM_ERR_INVALID_ARG;
    const int64_t nc = h->g.nc, cap = h->rec_cap, nbins = h->nbins;
    const int n0 = h->g.n[0], n1 = h->g.n[1], n2 = h->g.n[2];
    std::vector<uint8_t> flag(nc, 0);
    std::vector<double> rsus(4 * cap, 0.0);
    std::vector<int32_t> rid(cap, 0), cand(cap, -1), orph;
    std::vector<uint32_t> any(nbins, 0u);   /* live records per bin */
    for (int X = 0; X < n0; ++X)
        for (int Y = 0; Y < n1; ++Y)
            for (int Z = 0; Z < n2; ++Z) {
                const int64_t c = ((int64_t)X * n1 + Y) * n2 + Z;
                /* only records with sf > SOLFRAC_MIN take part in collision and ownership */
                if (!(sf[c] > SOLFRAC_MIN)) continue;
                const int64_t pid = id ? id[c] : 0;
                if (pid < 0 || pid >= (int64_t)INT32_MAX) return set_err(h, LBDEM_PSM_ERR_INVALID_ARG, "bad record id");
                const int64_t r = rec_index(X, Y, Z, n1, n2);
                flag[c] = 1;
                rsus[4 * r] = sf[c];
                rid[r] = (int32_t)pid;
                for (int d = 0; d < 3; ++d) rsus[4 * r + 1 + d] = us ? us[d * nc + c] : 0.0;
                any[((int64_t)(X >> 3) * h->g.nb[1] + (Y >> 3)) * h->g.nb[2] + (Z >> 3)] += 1u;
                orph.push_back((int32_t)r);
            }
    const int64_t m = (int64_t)orph.size();
    if (m > h->orphan_cap) return set_err(h, LBDEM_PSM_ERR_CAPACITY, "too many records");
    if (h->max_records > 0 && m > h->max_records)
        return set_err(h, LBDEM_PSM_ERR_CAPACITY, "records exceed max_solid_cells");
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    PSM_HIP(h, hipMemcpy(h->sflag, flag.data(), nc, hipMemcpyHostToDevice));
    PSM_HIP(h, hipMemcpy(h->rec.sus, rsus.data(), 4 * cap * sizeof(double), hipMemcpyHostToDevice));
    PSM_HIP(h, hipMemcpy(h->rec.id, rid.data(), cap * sizeof(int32_t), hipMemcpyHostToDevice));
    PSM_HIP(h, hipMemcpy(h->rec.cand, cand.data(), cap * sizeof(int32_t), hipMemcpyHostToDevice));
    PSM_HIP(h, hipMemcpy(h->bin_any[h->cur_any], any.data(), nbins * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (m > 0) PSM_HIP(h, hipMemcpy(h->orphan, orph.data(), m * sizeof(int32_t), hipMemcpyHostToDevice));
    /* Until the next voxelization every record is summed on the orphan
     * path: no candidate lists describe these records. */
    PSM_HIP(h, hipMemset(h->bin_cand_count, 0, nbins * sizeof(int32_t)));
    if (h->p_cap > 0) PSM_HIP(h, hipMemset(h->p_nent, 0, h->p_cap * sizeof(int32_t)));
    int64_t hc[8] = {m, 0, m, 0, 0, 0, 0, 0};
    PSM_HIP(h, hipMemcpy(h->counters, hc, sizeof(hc), hipMemcpyHostToDevice));
    h->records_from_host = true;
    h->geom_gen += 1;
    return mirror_records(h);
}

int lbdem_psm_subcell_counts(lbdem_psm* h, int64_t n, const double* d3, const double* r, double* sf,
                             int32_t* count, int32_t* branch) {
    if (!h || n < 0 || (n > 0 && (!d3 || !r || !sf || !count || !branch))) return LBDEM_PSM_ERR_INVALID_ARG;
    if (n == 0) return LBDEM_PSM_OK;
    double *dd = nullptr, *dr = nullptr, *dsf = nullptr;
    int32_t *dc = nullptr, *db = nullptr;
    PSM_HIP(h, dalloc(&dd, 3 * n * sizeof(double)));
    PSM_HIP(h, dalloc(&dr, n * sizeof(double)));
    PSM_HIP(h, dalloc(&dsf, n * sizeof(double)));
    PSM_HIP(h, dalloc(&dc, n * sizeof(int32_t)));
    PSM_HIP(h, dalloc(&db, n * sizeof(int32_t)));
    PSM_HIP(h, hipMemcpy(dd, d3, 3 * n * sizeof(double), hipMemcpyHostToDevice));
    PSM_HIP(h, hipMemcpy(dr, r, n * sizeof(double), hipMemcpyHostToDevice));
    int rc = launch_subcell(h, n, dd, dr, dsf, dc, db);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (!rc && e == hipSuccess) e = hipMemcpy(sf, dsf, n * sizeof(double), hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess) e = hipMemcpy(count, dc, n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess) e = hipMemcpy(branch, db, n * sizeof(int32_t), hipMemcpyDeviceToHost);
    dfree(dd); dfree(dr); dfree(dsf); dfree(dc); dfree(db);
    if (rc) return rc;
    PSM_HIP(h, e);
    return LBDEM_PSM_OK;
}

int lbdem_psm_get_macroscopic(lbdem_psm* h, double* rho, double* u3) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    const int64_t nc = h->g.nc;
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    double *dr = nullptr, *du = nullptr;
    if (rho) PSM_HIP(h, dalloc(&dr, nc * sizeof(double)));
    if (u3) {
        hipError_t e = dalloc(&du, 3 * nc * sizeof(double));
        if (e != hipSuccess) { dfree(dr); return hip_check(h, e, "hipMalloc"); }
    }
    int rc = (rho || u3) ? launch_macroscopic(h, dr, du) : LBDEM_PSM_OK;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (!rc && e == hipSuccess && rho) e = hipMemcpy(rho, dr, nc * sizeof(double), hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess && u3) e = hipMemcpy(u3, du, 3 * nc * sizeof(double), hipMemcpyDeviceToHost);
    dfree(dr);
    dfree(du);
    if (rc) return rc;
    PSM_HIP(h, e);
    return LBDEM_PSM_OK;
}

int lbdem_psm_get_desc(const lbdem_psm* h, lbdem_psm_desc* out) {
    if (!h || !out) return LBDEM_PSM_ERR_INVALID_ARG;
    *out = h->desc;
    return LBDEM_PSM_OK;
}

int lbdem_psm_conservation(lbdem_psm* h, double* mass, double* mom3) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    return launch_conservation(h, mass, mom3, nullptr, nullptr);
}

int lbdem_psm_check_health(lbdem_psm* h, double* max_u, int64_t* n_nonfinite) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    double mu = 0.0;
    int64_t nf = 0;
    int rc = launch_conservation(h, nullptr, nullptr, &mu, &nf);
    if (rc) return rc;
    if (max_u) *max_u = mu;
    if (n_nonfinite) *n_nonfinite = nf;
    if (nf > 0) return set_err(h, LBDEM_PSM_ERR_NUMERIC, "non-finite populations");
    if (!(mu < 0.3)) return set_err(h, LBDEM_PSM_ERR_NUMERIC, "max |u| >= 0.3 lattice units");
    return LBDEM_PSM_OK;
}

int lbdem_psm_stats(lbdem_psm* h, int64_t* out4) {
    if (!h || !out4) return LBDEM_PSM_ERR_INVALID_ARG;
    int rc = count_records(h);
    if (rc) return rc;
    rc = read_counters(h);
    if (rc) return rc;
    int32_t last_start = 0, last_count = 0;
    if (h->nbins > 0) {
        PSM_HIP(h, hipMemcpy(&last_start, h->bin_cand_start + h->nbins - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
        PSM_HIP(h, hipMemcpy(&last_count, h->bin_cand_count + h->nbins - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    out4[0] = h->h_counters[0];
    out4[1] = (int64_t)last_start + last_count;
    out4[2] = h->h_counters[4];
    out4[3] = h->h_counters[2];
    return LBDEM_PSM_OK;
}

int lbdem_psm_synchronize(lbdem_psm* h) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    return pending_status(h);
}

int lbdem_psm_profile(lbdem_psm* h, int32_t enable) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    h->profiling = enable != 0;
    return LBDEM_PSM_OK;
}

int lbdem_psm_profile_read(lbdem_psm* h, double* ms4, int64_t* count4, int32_t reset) {
    if (!h) return LBDEM_PSM_ERR_INVALID_ARG;
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    for (auto& s : h->spans) {
        float ms = 0.0f;
        PSM_HIP(h, hipEventElapsedTime(&ms, s.a, s.b));
        h->prof_ms[s.cls] += ms;
        h->prof_n[s.cls] += 1;
        h->ev_pool.push_back(s.a);
        h->ev_pool.push_back(s.b);
    }
    h->spans.clear();
    for (int k = 0; k < 4; ++k) {
        if (ms4) ms4[k] = h->prof_ms[k];
        if (count4) count4[k] = h->prof_n[k];
    }
    if (reset)
        for (int k = 0; k < 4; ++k) { h->prof_ms[k] = 0.0; h->prof_n[k] = 0; }
    return LBDEM_PSM_OK;
}

}  // extern "C"

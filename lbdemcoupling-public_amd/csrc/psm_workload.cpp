/*
 * Seeded synthetic sphere packings (host only); see lbdem_psm_workload.h.
 */
#include "../../include/lbdem_psm_workload.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int k = 0; k < 4; ++k) {   /* splitmix64 */
            z += 0x9e3779b97f4a7c15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
            s[k] = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {   /* xoshiro256** */
        const uint64_t res = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return res;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }   /* [0,1) */
    double normal() {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

extern "C" int lbdem_psm_generate_spheres(const lbdem_psm_workload_desc* d, double* x3, double* v3,
                                          double* w3, double* r, int64_t* id) {
    if (!d || d->count < 0) return -1;
    const int64_t n = d->count;
    if (n > 0 && (!x3 || !v3 || !w3 || !r || !id)) return -1;
    for (int k = 0; k < 3; ++k)
        if (d->n[k] < 1) return -1;
    Rng rng(d->seed);
    const int64_t max_att = d->max_attempts > 0 ? d->max_attempts : 100000;

    /* radii first, so that the largest radius sizes the cell list */
    double rmax = 0.0;
    for (int64_t q = 0; q < n; ++q) {
        double rr;
        if (d->radius_law == LBDEM_PSM_RADIUS_CONST) rr = d->a;
        else if (d->radius_law == LBDEM_PSM_RADIUS_UNIFORM) rr = d->a + (d->b - d->a) * rng.uniform();
        else if (d->radius_law == LBDEM_PSM_RADIUS_LOGNORMAL) {
            rr = d->a * std::exp(d->b * rng.normal());
            rr = std::min(std::max(rr, d->lo), d->hi);
        } else return -1;
        if (!(rr > 0.0)) return -1;
        r[q] = rr;
        rmax = std::max(rmax, rr);
    }
    const double cs = 2.0 * rmax + d->gap;
    int nc[3];
    double L[3];
    for (int k = 0; k < 3; ++k) {
        L[k] = (double)d->n[k];
        nc[k] = std::max(1, (int)std::floor(L[k] / cs));
    }
    std::vector<std::vector<int64_t>> cells((size_t)nc[0] * nc[1] * nc[2]);
    auto cell_of = [&](const double* x, int* c) {
        for (int k = 0; k < 3; ++k) {
            int v = (int)std::floor((x[k] + 0.5) / L[k] * nc[k]);
            c[k] = std::min(std::max(v, 0), nc[k] - 1);
        }
    };
    for (int64_t q = 0; q < n; ++q) {
        bool placed = false;
        for (int64_t att = 0; att < max_att && !placed; ++att) {
            double x[3];
            bool ok = true;
            for (int k = 0; k < 3; ++k) {
                if (d->periodic[k]) x[k] = -0.5 + L[k] * rng.uniform();
                else {
                    const double m = r[q] + d->wall_gap;
                    const double span = L[k] - 2.0 * m;
                    if (span <= 0.0) return -1;
                    x[k] = -0.5 + m + span * rng.uniform();
                }
            }
            int c[3];
            cell_of(x, c);
            for (int ox = -1; ox <= 1 && ok; ++ox)
                for (int oy = -1; oy <= 1 && ok; ++oy)
                    for (int oz = -1; oz <= 1 && ok; ++oz) {
                        int cc[3] = {c[0] + ox, c[1] + oy, c[2] + oz};
                        bool skip = false;
                        for (int k = 0; k < 3; ++k) {
                            if (cc[k] < 0 || cc[k] >= nc[k]) {
                                if (d->periodic[k]) cc[k] = (cc[k] + nc[k]) % nc[k];
                                else skip = true;
                            }
                        }
                        if (skip) continue;
                        for (int64_t p : cells[((size_t)cc[0] * nc[1] + cc[1]) * nc[2] + cc[2]]) {
                            double d2 = 0.0;
                            for (int k = 0; k < 3; ++k) {
                                double dd = x[k] - x3[3 * p + k];
                                if (d->periodic[k]) {
                                    if (dd > 0.5 * L[k]) dd -= L[k];
                                    else if (dd < -0.5 * L[k]) dd += L[k];
                                }
                                d2 += dd * dd;
                            }
                            const double md = r[q] + r[p] + d->gap;
                            if (d2 < md * md) { ok = false; break; }
                        }
                    }
            if (!ok) continue;
            for (int k = 0; k < 3; ++k) x3[3 * q + k] = x[k];
            cells[((size_t)c[0] * nc[1] + c[1]) * nc[2] + c[2]].push_back(q);
            placed = true;
        }
        if (!placed) return -1;
    }
    for (int64_t q = 0; q < n; ++q)
        for (int k = 0; k < 3; ++k) v3[3 * q + k] = d->sigma_v * rng.normal();
    for (int64_t q = 0; q < n; ++q)
        for (int k = 0; k < 3; ++k) w3[3 * q + k] = d->sigma_w * rng.normal();
    for (int64_t q = 0; q < n; ++q) id[q] = q + 1;
    return 0;
}

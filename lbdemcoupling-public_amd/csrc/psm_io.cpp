/*
 * On-disk formats around the PSM path (host code, no device code):
 *
 * - VTK XML ImageData (.vti) output of per-cell fields, the format of
 *   Palabos' VtkImageOutput3D as used by the reference's writeVTK
 *   (examples/benchmark/benchmark.cpp:74-99; field extraction
 *   src/ibDataWritingFunctionals3D.hh:32-114): Float32 point data, x fastest,
 *   binary base64 with a UInt32 byte-count header.
 * - The checkpoint record: a header, then per cell (x, y, z fastest) the 19
 *   populations f - t in FP64 followed by the coupling record in the order
 *   IBcompositeDynamics::serialize writes it
 *   (src/ibCompositeDynamics3D.hh:77-89): uPart[3], hydrodynamicForce[3],
 *   solidFraction, partId.
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/lbdem_psm.h"

extern "C" __attribute__((visibility("hidden"))) int lbdem_psm_internal_set_error(lbdem_psm* h, int code,
                                                                                   const char* msg);

namespace {

const int Q = 19;
const char kMagic[8] = {'L', 'B', 'D', 'E', 'M', 'P', 'S', 'M'};
const uint32_t kVersion = 1;

struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t n[3];
    int32_t periodic[3];
    int32_t precision;
    int32_t slices;
    double tau;
    double body_accel[3];
    int64_t ncell;
    uint32_t cell_bytes;     /* 27 * 8 */
    uint32_t reserved;
};

int io_err(lbdem_psm* h, int code, const std::string& msg) {
    return lbdem_psm_internal_set_error(h, code, msg.c_str());
}

/* ---- base64 ---- */
const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

void base64_append(const unsigned char* p, size_t n, std::string& out) {
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        const uint32_t v = ((uint32_t)p[i] << 16) | ((uint32_t)p[i + 1] << 8) | p[i + 2];
        out.push_back(kB64[(v >> 18) & 63]);
        out.push_back(kB64[(v >> 12) & 63]);
        out.push_back(kB64[(v >> 6) & 63]);
        out.push_back(kB64[v & 63]);
    }
    if (i < n) {
        uint32_t v = (uint32_t)p[i] << 16;
        if (i + 1 < n) v |= (uint32_t)p[i + 1] << 8;
        out.push_back(kB64[(v >> 18) & 63]);
        out.push_back(kB64[(v >> 12) & 63]);
        out.push_back(i + 1 < n ? kB64[(v >> 6) & 63] : '=');
        out.push_back('=');
    }
}

}  // namespace

extern "C" {

int lbdem_psm_vti_write(const char* path, int32_t nx, int32_t ny, int32_t nz, double spacing, int32_t nfields,
                        const char* const* names, const int32_t* ncomp, const double* const* data,
                        const double* scale) {
    if (!path || nx < 1 || ny < 1 || nz < 1 || nfields < 0 || (nfields > 0 && (!names || !ncomp || !data)))
        return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "vti_write: bad arguments");
    const int64_t nc = (int64_t)nx * ny * nz;
    for (int k = 0; k < nfields; ++k) {
        if (!names[k] || !data[k] || ncomp[k] < 1 || ncomp[k] > 9)
            return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "vti_write: bad field");
        if ((uint64_t)nc * (uint64_t)ncomp[k] * 4u > 0xffffffffull)
            return io_err(nullptr, LBDEM_PSM_ERR_CAPACITY, "vti_write: field over 4 GiB (UInt32 header)");
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, std::string("vti_write: cannot open ") + path);
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n");
    fprintf(fp, "<ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" Spacing=\"%.17g %.17g %.17g\">\n",
            nx - 1, ny - 1, nz - 1, spacing, spacing, spacing);
    fprintf(fp, "<Piece Extent=\"0 %d 0 %d 0 %d\">\n<PointData>\n", nx - 1, ny - 1, nz - 1);
    std::vector<float> buf;
    std::string enc;
    for (int k = 0; k < nfields; ++k) {
        const int m = ncomp[k];
        const double s = scale ? scale[k] : 1.0;
        buf.assign((size_t)nc * m + 1, 0.0f);
        const uint32_t nbytes = (uint32_t)((uint64_t)nc * m * 4u);
        memcpy(&buf[0], &nbytes, 4);
        /* VTK point order: x fastest; the arrays here are z fastest, component-major */
        size_t o = 1;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const int64_t c = ((int64_t)x * ny + y) * nz + z;
                    for (int d = 0; d < m; ++d) buf[o++] = (float)(data[k][(int64_t)d * nc + c] * s);
                }
        enc.clear();
        enc.reserve(((size_t)nc * m * 4 + 4) / 3 * 4 + 8);
        base64_append((const unsigned char*)buf.data(), (size_t)nc * m * 4 + 4, enc);
        fprintf(fp, "<DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"binary\">\n",
                names[k], m);
        fwrite(enc.data(), 1, enc.size(), fp);
        fprintf(fp, "\n</DataArray>\n");
    }
    fprintf(fp, "</PointData>\n</Piece>\n</ImageData>\n</VTKFile>\n");
    const bool ok = ferror(fp) == 0;
    fclose(fp);
    if (!ok) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "vti_write: write error");
    return LBDEM_PSM_OK;
}

int lbdem_psm_write_vti(lbdem_psm* h, const char* path, double spacing, double vel_scale, double rho_scale,
                        double p_scale) {
    if (!h || !path) return LBDEM_PSM_ERR_INVALID_ARG;
    lbdem_psm_desc d;
    int rc = lbdem_psm_get_desc(h, &d);
    if (rc) return rc;
    const int64_t nc = (int64_t)d.nx * d.ny * d.nz;
    std::vector<double> rho(nc), u(3 * nc), p(nc), sf(nc);
    rc = lbdem_psm_get_macroscopic(h, rho.data(), u.data());
    if (rc) return rc;
    rc = lbdem_psm_get_cell_records(h, sf.data(), nullptr, nullptr, nullptr);
    if (rc) return rc;
    /* writeVTK of the reference: velocity, density, pressure (rho - 1) * p_fact, solid fraction */
    for (int64_t c = 0; c < nc; ++c) p[c] = rho[c] - 1.0;
    const char* names[4] = {"velocity", "density", "pressure", "solidfraction"};
    const int32_t ncomp[4] = {3, 1, 1, 1};
    const double* data[4] = {u.data(), rho.data(), p.data(), sf.data()};
    const double scale[4] = {vel_scale, rho_scale, p_scale, 1.0};
    rc = lbdem_psm_vti_write(path, d.nx, d.ny, d.nz, spacing, 4, names, ncomp, data, scale);
    if (rc) return io_err(h, rc, lbdem_psm_last_error(nullptr));
    return LBDEM_PSM_OK;
}

int lbdem_psm_checkpoint_write_arrays(const char* path, const lbdem_psm_desc* d, const double* f, const double* sf,
                                      const int64_t* id, const double* us, const double* F) {
    if (!path || !d || !f || !sf || !id || !us || !F)
        return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: NULL argument");
    const int64_t nc = (int64_t)d->nx * d->ny * d->nz;
    CkptHeader hd;
    memset(&hd, 0, sizeof(hd));
    memcpy(hd.magic, kMagic, 8);
    hd.version = kVersion;
    hd.header_bytes = (uint32_t)sizeof(CkptHeader);
    hd.n[0] = d->nx; hd.n[1] = d->ny; hd.n[2] = d->nz;
    for (int k = 0; k < 3; ++k) { hd.periodic[k] = d->periodic[k]; hd.body_accel[k] = d->body_accel[k]; }
    hd.precision = d->precision;
    hd.slices = d->slices_per_dim;
    hd.tau = d->tau;
    hd.ncell = nc;
    hd.cell_bytes = (Q + 8) * 8;
    FILE* fp = fopen(path, "wb");
    if (!fp) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, std::string("checkpoint: cannot open ") + path);
    fwrite(&hd, sizeof(hd), 1, fp);
    const int64_t chunk = 1 << 16;
    std::vector<double> buf((size_t)chunk * (Q + 8));
    for (int64_t c0 = 0; c0 < nc; c0 += chunk) {
        const int64_t m = nc - c0 < chunk ? nc - c0 : chunk;
        for (int64_t k = 0; k < m; ++k) {
            const int64_t c = c0 + k;
            double* o = &buf[(size_t)k * (Q + 8)];
            for (int i = 0; i < Q; ++i) o[i] = f[(int64_t)i * nc + c];
            /* IBcompositeDynamics::serialize: uPart, hydrodynamicForce, solidFraction, partId */
            for (int dd = 0; dd < 3; ++dd) o[Q + dd] = us[(int64_t)dd * nc + c];
            for (int dd = 0; dd < 3; ++dd) o[Q + 3 + dd] = F[(int64_t)dd * nc + c];
            o[Q + 6] = sf[c];
            memcpy(&o[Q + 7], &id[c], 8);
        }
        fwrite(buf.data(), sizeof(double) * (Q + 8), (size_t)m, fp);
    }
    const bool ok = ferror(fp) == 0;
    fclose(fp);
    if (!ok) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: write error");
    return LBDEM_PSM_OK;
}

static int read_header(FILE* fp, CkptHeader* hd) {
    if (fread(hd, sizeof(*hd), 1, fp) != 1) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: short header");
    if (memcmp(hd->magic, kMagic, 8) != 0 || hd->version != kVersion || hd->header_bytes != sizeof(CkptHeader) ||
        hd->cell_bytes != (Q + 8) * 8)
        return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: not an lbdem_psm checkpoint (version 1)");
    if (hd->n[0] < 1 || hd->n[1] < 1 || hd->n[2] < 1 || hd->ncell != (int64_t)hd->n[0] * hd->n[1] * hd->n[2])
        return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: inconsistent sizes");
    return LBDEM_PSM_OK;
}

int lbdem_psm_checkpoint_read_desc(const char* path, lbdem_psm_desc* d) {
    if (!path || !d) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: NULL argument");
    FILE* fp = fopen(path, "rb");
    if (!fp) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, std::string("checkpoint: cannot open ") + path);
    CkptHeader hd;
    const int rc = read_header(fp, &hd);
    fclose(fp);
    if (rc) return rc;
    lbdem_psm_desc_init(d);
    d->nx = hd.n[0]; d->ny = hd.n[1]; d->nz = hd.n[2];
    for (int k = 0; k < 3; ++k) { d->periodic[k] = hd.periodic[k]; d->body_accel[k] = hd.body_accel[k]; }
    d->tau = hd.tau;
    d->precision = hd.precision;
    d->slices_per_dim = hd.slices;
    return LBDEM_PSM_OK;
}

int lbdem_psm_checkpoint_read_arrays(const char* path, double* f, double* sf, int64_t* id, double* us, double* F) {
    if (!path) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: NULL path");
    FILE* fp = fopen(path, "rb");
    if (!fp) return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, std::string("checkpoint: cannot open ") + path);
    CkptHeader hd;
    int rc = read_header(fp, &hd);
    if (rc) { fclose(fp); return rc; }
    const int64_t nc = hd.ncell;
    const int64_t chunk = 1 << 16;
    std::vector<double> buf((size_t)chunk * (Q + 8));
    for (int64_t c0 = 0; c0 < nc; c0 += chunk) {
        const int64_t m = nc - c0 < chunk ? nc - c0 : chunk;
        if (fread(buf.data(), sizeof(double) * (Q + 8), (size_t)m, fp) != (size_t)m) {
            fclose(fp);
            return io_err(nullptr, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: truncated file");
        }
        for (int64_t k = 0; k < m; ++k) {
            const int64_t c = c0 + k;
            const double* o = &buf[(size_t)k * (Q + 8)];
            if (f)
                for (int i = 0; i < Q; ++i) f[(int64_t)i * nc + c] = o[i];
            if (us)
                for (int dd = 0; dd < 3; ++dd) us[(int64_t)dd * nc + c] = o[Q + dd];
            if (F)
                for (int dd = 0; dd < 3; ++dd) F[(int64_t)dd * nc + c] = o[Q + 3 + dd];
            if (sf) sf[c] = o[Q + 6];
            if (id) memcpy(&id[c], &o[Q + 7], 8);
        }
    }
    fclose(fp);
    return LBDEM_PSM_OK;
}

int lbdem_psm_save_checkpoint(lbdem_psm* h, const char* path) {
    if (!h || !path) return LBDEM_PSM_ERR_INVALID_ARG;
    lbdem_psm_desc d;
    int rc = lbdem_psm_get_desc(h, &d);
    if (rc) return rc;
    const int64_t nc = (int64_t)d.nx * d.ny * d.nz;
    std::vector<double> f((size_t)Q * nc), sf(nc), us(3 * nc), F(3 * nc);
    std::vector<int64_t> id(nc);
    rc = lbdem_psm_get_populations(h, f.data());
    if (rc) return rc;
    rc = lbdem_psm_get_cell_records(h, sf.data(), id.data(), us.data(), F.data());
    if (rc) return rc;
    rc = lbdem_psm_checkpoint_write_arrays(path, &d, f.data(), sf.data(), id.data(), us.data(), F.data());
    if (rc) return io_err(h, rc, lbdem_psm_last_error(nullptr));
    return LBDEM_PSM_OK;
}

int lbdem_psm_load_checkpoint(lbdem_psm* h, const char* path) {
    if (!h || !path) return LBDEM_PSM_ERR_INVALID_ARG;
    lbdem_psm_desc d, fd;
    int rc = lbdem_psm_get_desc(h, &d);
    if (rc) return rc;
    rc = lbdem_psm_checkpoint_read_desc(path, &fd);
    if (rc) return io_err(h, rc, lbdem_psm_last_error(nullptr));
    if (fd.nx != d.nx || fd.ny != d.ny || fd.nz != d.nz)
        return io_err(h, LBDEM_PSM_ERR_INVALID_ARG, "checkpoint: lattice sizes differ");
    const int64_t nc = (int64_t)d.nx * d.ny * d.nz;
    std::vector<double> f((size_t)Q * nc), sf(nc), us(3 * nc);
    std::vector<int64_t> id(nc);
    rc = lbdem_psm_checkpoint_read_arrays(path, f.data(), sf.data(), id.data(), us.data(), nullptr);
    if (rc) return io_err(h, rc, lbdem_psm_last_error(nullptr));
    rc = lbdem_psm_set_populations(h, f.data());
    if (rc) return rc;
    return lbdem_psm_set_cell_records(h, sf.data(), id.data(), us.data());
}

}  // extern "C"

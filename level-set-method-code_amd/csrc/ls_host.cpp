// ls_host.cpp -- host-only parts of the C ABI: defaults, CFL time step and the
// reference G output format.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ls_api.h"
#include "ls_internal.hpp"

extern "C" {

// initDefaultParams (Local_code/level-set_MPI_2D/include/config.h:166-213,
// Local_code/level-set_MPI_3D/include/config.h:180-234) with runtime n.
int ls_default_params(ls_params_t* p, int32_t dim, int32_t n) {
    if (!p) return lsi::set_error(LS_ERR_INVALID, "params is NULL");
    if ((dim != 2 && dim != 3) || n < 3) return lsi::set_error(LS_ERR_INVALID, "bad dim or n");
    std::memset(p, 0, sizeof(*p));
    p->dim = dim;
    p->nx = p->ny = n;
    p->nz = dim == 3 ? n : 1;
    p->nghost = 3;
    p->bc[0] = LS_BC_PERIODIC;
    p->bc[1] = dim == 3 ? LS_BC_PERIODIC : LS_BC_ZERO_GRAD;
    p->bc[2] = LS_BC_PERIODIC;
    p->mode = LS_MODE_EXACT;
    p->velocity = LS_VEL_ARRAYS;
    p->dx = p->dy = (1.0 - 0.0) / (n - 1);
    p->dz = dim == 3 ? (1.0 - 0.0) / (n - 1) : 1.0;
    p->x0 = p->y0 = p->z0 = 0.0;
    p->s_l = 0.0;
    p->eps = 1.0e-10;
    p->reinit_interval = 10;
    p->reinit_iters = 2;
    p->reinit_dtau = 0.25 * (dim == 3 ? std::fmin(p->dx, std::fmin(p->dy, p->dz)) : std::fmin(p->dx, p->dy));
    p->reinit_beta = 0.5;
    p->vel_period = 1.5;
    return LS_OK;
}

// calculateTimeStep: 2D config.h:251-263 (speed floor eps), 3D
// config.h:272-293 (speed floor 2.0 for the deformation test).
double ls_cfl_dt(const ls_params_t* p, double cfl, double u, double v, double w) {
    if (!p) return -1.0;
    if (p->dim == 3) {
        double sp = std::fabs(u) + std::fabs(v) + std::fabs(w) + p->s_l;
        if (sp < p->eps) sp = 2.0;
        double dmin = std::fmin(p->dx, std::fmin(p->dy, p->dz));
        return cfl * dmin / sp;
    }
    double sp = std::fabs(u) + std::fabs(v) + p->s_l;
    if (sp < p->eps) sp = p->eps;
    return cfl * std::fmin(p->dx, p->dy) / sp;
}

// saveFieldBinary (2D io.h:34-114) / saveFieldBinary3D (3D io.h:34-136) for
// one rank: header of int32, then the padded field, y-ghost rows replaced by
// the adjacent interior row.
int ls_write_G(const ls_params_t* p, const double* host_G, const char* path) {
    int rc = lsi::validate(p);
    if (rc) return rc;
    if (!host_G || !path) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    const int ng = p->nghost;
    const long long nxt = p->nx + 2 * ng, nyt = p->ny + 2 * ng;
    const long long nzt = p->dim == 3 ? p->nz + 2 * ng : 1;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return lsi::set_error(LS_ERR_IO, "cannot open output file");
    int32_t hdr[4];
    size_t nh;
    if (p->dim == 3) {
        hdr[0] = p->nx; hdr[1] = p->ny; hdr[2] = p->nz; hdr[3] = ng; nh = 4;
    } else {
        hdr[0] = p->nx; hdr[1] = p->ny; hdr[2] = ng; nh = 3;
    }
    bool ok = std::fwrite(hdr, sizeof(int32_t), nh, fp) == nh;
    for (long long k = 0; ok && k < nzt; k++) {
        for (long long j = 0; ok && j < nyt; j++) {
            long long js = j < ng ? ng : (j >= p->ny + ng ? p->ny + ng - 1 : j);
            const double* row = host_G + (k * nyt + js) * nxt;
            ok = std::fwrite(row, sizeof(double), (size_t)nxt, fp) == (size_t)nxt;
        }
    }
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return lsi::set_error(LS_ERR_IO, "short write");
    return LS_OK;
}

// saveFieldVTK / saveFieldVTK3D (2D io.h:119-180, 3D io.h:145-215): the
// header lines and number formats of the reference, interior values only.
int ls_write_G_vtk(const ls_params_t* p, const double* host_G, const char* path) {
    int rc = lsi::validate(p);
    if (rc) return rc;
    if (!host_G || !path) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return lsi::set_error(LS_ERR_IO, "cannot open output file");
    const int ng = p->nghost;
    const bool d3 = p->dim == 3;
    const long long nxt = p->nx + 2 * ng, nyt = p->ny + 2 * ng;
    const long long nz = d3 ? p->nz : 1;
    bool ok = std::fprintf(fp, "# vtk DataFile Version 3.0\n%s\nASCII\nDATASET STRUCTURED_POINTS\n",
                           d3 ? "G-equation Level-Set Field 3D" : "G-equation Level-Set Field") > 0;
    if (d3) {
        ok = ok && std::fprintf(fp, "DIMENSIONS %d %d %d\n", p->nx, p->ny, p->nz) > 0 &&
             std::fprintf(fp, "ORIGIN %f %f %f\n", p->x0, p->y0, p->z0) > 0 &&
             std::fprintf(fp, "SPACING %f %f %f\n", p->dx, p->dy, p->dz) > 0;
    } else {
        ok = ok && std::fprintf(fp, "DIMENSIONS %d %d 1\n", p->nx, p->ny) > 0 &&
             std::fprintf(fp, "ORIGIN %f %f 0\n", p->x0, p->y0) > 0 &&
             std::fprintf(fp, "SPACING %f %f 1\n", p->dx, p->dy) > 0;
    }
    ok = ok && std::fprintf(fp, "POINT_DATA %lld\nSCALARS G double 1\nLOOKUP_TABLE default\n",
                            (long long)p->nx * p->ny * nz) > 0;
    for (long long k = 0; ok && k < nz; k++)
        for (long long j = 0; ok && j < p->ny; j++) {
            const double* row = host_G + ((d3 ? k + ng : 0) * nyt + j + ng) * nxt + ng;
            for (long long i = 0; ok && i < p->nx; i++) ok = std::fprintf(fp, "%e\n", row[i]) > 0;
        }
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return lsi::set_error(LS_ERR_IO, "short write");
    return LS_OK;
}

// save_G_to_bin (sample/main.cpp:34-51): the interior rows only, no header.
int ls_sample_write_G(int32_t nx, int32_t ny, const double* host_G, const char* path) {
    if (nx < 6 || ny < 6) return lsi::set_error(LS_ERR_INVALID, "the sample scheme needs nx, ny >= 6");
    if (!host_G || !path) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return lsi::set_error(LS_ERR_IO, "cannot open output file");
    const long long nxt = nx + 6;
    bool ok = true;
    for (long long j = 3; ok && j <= ny + 2; j++)
        ok = std::fwrite(host_G + j * nxt + 3, sizeof(double), (size_t)nx, fp) == (size_t)nx;
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return lsi::set_error(LS_ERR_IO, "short write");
    return LS_OK;
}

}  // extern "C"

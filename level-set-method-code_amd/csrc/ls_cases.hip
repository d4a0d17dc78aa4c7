// ls_cases.hip -- synthetic inputs generated on the device.
//
// Shapes and velocity fields of the reference test cases
// (Local_code/level-set_MPI_2D/include/initial_conditions.h:67-117,209-281;
//  Local_code/level-set_MPI_3D/include/initial_conditions.h:52-58,232-343)
// and the seeded workloads of the benchmark.  Seeded draws use
// SplitMix64 on the host; the kernels receive the drawn mode tables by value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "ls_api.h"
#include "ls_internal.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

constexpr int kMaxModes = 16;

// Mode tables: G-surface modes (amplitude, wavenumbers, phase) and solenoidal
// velocity modes (unit amplitude direction a_hat perpendicular to k).
struct Modes {
    int ng_modes, nv_modes;
    double ga[kMaxModes], gky[kMaxModes], gkz[kMaxModes], gph[kMaxModes];
    double vk[kMaxModes][3], va[kMaxModes][3], vph[kMaxModes];
    double vamp;
    double base[3];
    // flame2d shape
    int circle;
    double cx, cy, r, x_off;
};

struct Coord {
    double x0, y0, z0, dx, dy, dz;
    int ng, nxt, nyt, nzt;
};

__device__ __forceinline__ void coord(const Coord& c, int i, int j, int k, double& x, double& y, double& z) {
    x = c.x0 + (i - c.ng) * c.dx;
    y = c.y0 + (j - c.ng) * c.dy;
    z = c.z0 + (k - c.ng) * c.dz;
}

__device__ __forceinline__ void deformation(double x, double y, double z, double t, double T,
                                            double& u, double& v, double& w) {
    // initial_conditions.h:232-247 (3D)
    const double tf = cos(kPi * t / T);
    const double spx = sin(kPi * x), s2px = sin(2.0 * kPi * x);
    const double spy = sin(kPi * y), s2py = sin(2.0 * kPi * y);
    const double spz = sin(kPi * z), s2pz = sin(2.0 * kPi * z);
    u = 2.0 * spx * spx * s2py * s2pz * tf;
    v = -s2px * spy * spy * s2pz * tf;
    w = -s2px * s2py * spz * spz * tf;
}

__device__ __forceinline__ double zalesak2d(double x, double y, double cx, double cy, double r,
                                            double sw, double sl) {
    // initial_conditions.h:91-117 (2D): disk minus a slot open at the bottom.
    const double ddx = x - cx, ddy = y - cy;
    const double disk = sqrt(ddx * ddx + ddy * ddy) - r;
    const double hw = sw / 2.0;
    const double bot = cy - r, top = bot + sl;
    const double sx = fabs(x - cx) - hw;
    const double sy = fmax(bot - y, y - top);
    double slot;
    if (sx < 0.0 && sy < 0.0) slot = fmax(sx, sy);
    else if (sx < 0.0) slot = sy;
    else if (sy < 0.0) slot = sx;
    else slot = sqrt(sx * sx + sy * sy);
    return fmax(disk, -slot);
}

__global__ __launch_bounds__(256) void k_case(Coord c, int case_id, Modes m, double period,
                                              double* __restrict__ G, double* __restrict__ u,
                                              double* __restrict__ v, double* __restrict__ w,
                                              int64_t ntot) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntot) return;
    const int i = (int)(t % c.nxt);
    const int j = (int)((t / c.nxt) % c.nyt);
    const int k = (int)(t / ((int64_t)c.nxt * c.nyt));
    double x, y, z;
    coord(c, i, j, k, x, y, z);
    double g = 0.0, uu = 0.0, vv = 0.0, ww = 0.0;
    switch (case_id) {
        case LS_CASE_CIRCLE2D: {
            const double a = x - 0.5, b = y - 0.5;
            g = sqrt(a * a + b * b) - 0.25;
            uu = 0.5;
            break;
        }
        case LS_CASE_SPHERE_SWIRL3D: {
            const double a = x - 0.5, b = y - 0.5, d = z - 0.5;
            g = sqrt(a * a + b * b + d * d) - 0.2;
            uu = sin(kPi * x) * cos(kPi * y);
            vv = -cos(kPi * x) * sin(kPi * y);
            ww = 0.2 * sin(2.0 * kPi * z);
            break;
        }
        case LS_CASE_PLANE_MODES3D: {
            double eta = 0.0;
            for (int q = 0; q < m.ng_modes; q++)
                eta += m.ga[q] * sin(2.0 * kPi * (m.gky[q] * y + m.gkz[q] * z) + m.gph[q]);
            g = x - 0.3 - eta;
            uu = m.base[0];
            vv = m.base[1];
            ww = m.base[2];
            for (int q = 0; q < m.nv_modes; q++) {
                const double s = m.vamp * sin(2.0 * kPi * (m.vk[q][0] * x + m.vk[q][1] * y + m.vk[q][2] * z) + m.vph[q]);
                uu += s * m.va[q][0];
                vv += s * m.va[q][1];
                ww += s * m.va[q][2];
            }
            break;
        }
        case LS_CASE_ZALESAK2D: {
            g = zalesak2d(x, y, 0.5, 0.75, 0.15, 0.05, 0.25);
            const double om = 2.0 * kPi / period;
            uu = -om * (y - 0.5);
            vv = om * (x - 0.5);
            break;
        }
        case LS_CASE_SPHERE_DEFORM3D: {
            const double a = x - 0.35, b = y - 0.35, d = z - 0.35;
            g = sqrt(a * a + b * b + d * d) - 0.15;
            deformation(x, y, z, 0.0, period, uu, vv, ww);
            break;
        }
        case LS_CASE_FLAME2D: {
            if (m.circle) {
                const double a = x - m.cx, b = y - m.cy;
                g = sqrt(a * a + b * b) - m.r;
            } else {
                g = x - m.x_off;
            }
            uu = m.base[0];
            vv = m.base[1];
            for (int q = 0; q < m.nv_modes; q++) {
                const double s = m.vamp * sin(2.0 * kPi * (m.vk[q][0] * x + m.vk[q][1] * y) + m.vph[q]);
                uu += s * m.va[q][0];
                vv += s * m.va[q][1];
            }
            break;
        }
        default:
            break;
    }
    G[t] = g;
    u[t] = uu;
    v[t] = vv;
    if (w) w[t] = ww;
}

__global__ __launch_bounds__(256) void k_deform(Coord c, double tt, double T, double* __restrict__ u,
                                                double* __restrict__ v, double* __restrict__ w,
                                                int64_t ntot) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntot) return;
    const int i = (int)(t % c.nxt);
    const int j = (int)((t / c.nxt) % c.nyt);
    const int k = (int)(t / ((int64_t)c.nxt * c.nyt));
    double x, y, z;
    coord(c, i, j, k, x, y, z);
    deformation(x, y, z, tt, T, u[t], v[t], w[t]);
}

Coord make_coord(const ls_params_t& p) {
    Coord c;
    c.x0 = p.x0;
    c.y0 = p.y0;
    c.z0 = p.z0;
    c.dx = p.dx;
    c.dy = p.dy;
    c.dz = p.dim == 3 ? p.dz : 0.0;
    c.ng = p.nghost;
    c.nxt = p.nx + 2 * p.nghost;
    c.nyt = p.ny + 2 * p.nghost;
    c.nzt = p.dim == 3 ? p.nz + 2 * p.nghost : 1;
    return c;
}

// Integer wavevector with 1 <= |k| <= 4 (components in [-4,4]; 2D: kz = 0),
// then a unit direction perpendicular to it (Gram-Schmidt on a random vector).
void draw_solenoidal(SplitMix64& rng, int dim, double k[3], double a[3]) {
    for (;;) {
        k[0] = std::floor(9.0 * rng.uniform()) - 4.0;
        k[1] = std::floor(9.0 * rng.uniform()) - 4.0;
        k[2] = dim == 3 ? std::floor(9.0 * rng.uniform()) - 4.0 : 0.0;
        const double n2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        if (n2 >= 1.0 && n2 <= 16.0) break;
    }
    const double kn = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    const double kh[3] = {k[0] / kn, k[1] / kn, k[2] / kn};
    for (;;) {
        double r[3] = {rng.uniform() - 0.5, rng.uniform() - 0.5, dim == 3 ? rng.uniform() - 0.5 : 0.0};
        const double d = r[0] * kh[0] + r[1] * kh[1] + r[2] * kh[2];
        for (int q = 0; q < 3; q++) r[q] -= d * kh[q];
        const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (rn > 1e-3) {
            for (int q = 0; q < 3; q++) a[q] = r[q] / rn;
            return;
        }
    }
}


// ---------------------------------------------------------------------------
// Flame-inversion parameter maps.  The flame speed is the trilinear
// interpolant of nc^3 coarse node values spread uniformly over the interior
// box [x0, x0 + (nx-1) dx] x ...; the velocity is the swirl field of the
// sphere test plus sum_q amp_q * a_q sin(2 pi k_q . x + ph_q) over seeded
// solenoidal unit modes.  Both maps are linear in the parameters, so their
// adjoints are the transposed sums below.
struct FlameModes {
    int n;
    double k[kMaxModes][3], a[kMaxModes][3], ph[kMaxModes];
};

struct Coarse {
    int nc;
    double h[3];  // coarse node spacing per axis
};

__device__ __forceinline__ void trilin(double s, int nc, int& i0, double& f) {
    s = fmin(fmax(s, 0.0), (double)(nc - 1));
    int i = (int)floor(s);
    if (i > nc - 2) i = nc - 2;
    i0 = i;
    f = s - i;
}

__device__ __forceinline__ void swirl(double x, double y, double z, double& u, double& v, double& w) {
    u = sin(kPi * x) * cos(kPi * y);
    v = -cos(kPi * x) * sin(kPi * y);
    w = 0.2 * sin(2.0 * kPi * z);
}

__global__ __launch_bounds__(256) void k_flame_fields(Coord c, FlameModes m, Coarse cg,
                                                      const double* __restrict__ slc,
                                                      const double* __restrict__ amps,
                                                      double* __restrict__ sl, double* __restrict__ u,
                                                      double* __restrict__ v, double* __restrict__ w,
                                                      int64_t ntot) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntot) return;
    const int i = (int)(t % c.nxt);
    const int j = (int)((t / c.nxt) % c.nyt);
    const int k = (int)(t / ((int64_t)c.nxt * c.nyt));
    double x, y, z;
    coord(c, i, j, k, x, y, z);
    if (sl) {
        int ix, iy, iz;
        double fx, fy, fz;
        trilin((x - c.x0) / cg.h[0], cg.nc, ix, fx);
        trilin((y - c.y0) / cg.h[1], cg.nc, iy, fy);
        trilin((z - c.z0) / cg.h[2], cg.nc, iz, fz);
        const int nc = cg.nc;
        double acc = 0.0;
        for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    const double wgt = (dx ? fx : 1.0 - fx) * (dy ? fy : 1.0 - fy) * (dz ? fz : 1.0 - fz);
                    acc += wgt * slc[((iz + dz) * nc + (iy + dy)) * nc + (ix + dx)];
                }
        sl[t] = acc;
    }
    if (u) {
        double uu, vv, ww;
        swirl(x, y, z, uu, vv, ww);
        for (int q = 0; q < m.n; q++) {
            const double s = amps[q] * sin(2.0 * kPi * (m.k[q][0] * x + m.k[q][1] * y + m.k[q][2] * z) + m.ph[q]);
            uu += s * m.a[q][0];
            vv += s * m.a[q][1];
            ww += s * m.a[q][2];
        }
        u[t] = uu;
        v[t] = vv;
        w[t] = ww;
    }
}

// g_coarse[node] = sum over interior cells of weight(node, cell) * gsl[cell];
// one block per coarse node walks the cells whose interpolation stencil can
// hold the node, in a fixed order.
__global__ __launch_bounds__(256) void k_flame_project_sl(Coord c, int nx, int ny, int nz, Coarse cg,
                                                          const double* __restrict__ gsl,
                                                          double* __restrict__ gc) {
    __shared__ double red[256];
    const int nc = cg.nc;
    const int node = blockIdx.x;
    const int I = node % nc, J = (node / nc) % nc, K = node / (nc * nc);
    int lo[3], hi[3];
    const int nn[3] = {nx, ny, nz};
    const int ID[3] = {I, J, K};
    const double d[3] = {c.dx, c.dy, c.dz};
    for (int a = 0; a < 3; a++) {
        lo[a] = (int)floor((ID[a] - 1) * cg.h[a] / d[a]) - 1;
        hi[a] = (int)ceil((ID[a] + 1) * cg.h[a] / d[a]) + 1;
        lo[a] = lo[a] < 0 ? 0 : lo[a];
        hi[a] = hi[a] > nn[a] - 1 ? nn[a] - 1 : hi[a];
    }
    const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
    const int64_t cnt = (int64_t)ex * ey * ez;
    double acc = 0.0;
    for (int64_t q = threadIdx.x; q < cnt; q += blockDim.x) {
        const int i = lo[0] + (int)(q % ex), j = lo[1] + (int)((q / ex) % ey), k = lo[2] + (int)(q / ((int64_t)ex * ey));
        double x, y, z;
        coord(c, i + c.ng, j + c.ng, k + c.ng, x, y, z);
        int ix, iy, iz;
        double fx, fy, fz;
        trilin((x - c.x0) / cg.h[0], nc, ix, fx);
        trilin((y - c.y0) / cg.h[1], nc, iy, fy);
        trilin((z - c.z0) / cg.h[2], nc, iz, fz);
        const int dx = I - ix, dy = J - iy, dz = K - iz;
        if (dx < 0 || dx > 1 || dy < 0 || dy > 1 || dz < 0 || dz > 1) continue;
        const double wgt = (dx ? fx : 1.0 - fx) * (dy ? fy : 1.0 - fy) * (dz ? fz : 1.0 - fz);
        acc += wgt * gsl[((int64_t)(k + c.ng) * c.nyt + (j + c.ng)) * c.nxt + (i + c.ng)];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) gc[node] = red[0];
}

constexpr int kProjBlocks = 512;

// Per-block partial sums over interior cells of mode_q(x) . (gu, gv, gw).
__global__ __launch_bounds__(256) void k_flame_project_modes(Coord c, int nx, int ny, int nz, FlameModes m,
                                                             const double* __restrict__ gu,
                                                             const double* __restrict__ gv,
                                                             const double* __restrict__ gw,
                                                             double* __restrict__ part) {
    __shared__ double red[kMaxModes][256];
    double acc[kMaxModes];
    for (int q = 0; q < kMaxModes; q++) acc[q] = 0.0;
    const int64_t ncell = (int64_t)nx * ny * nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
        const int64_t p = ((int64_t)(k + c.ng) * c.nyt + (j + c.ng)) * c.nxt + (i + c.ng);
        double x, y, z;
        coord(c, i + c.ng, j + c.ng, k + c.ng, x, y, z);
        const double a = gu[p], b = gv[p], e = gw[p];
        for (int q = 0; q < m.n; q++) {
            const double s = sin(2.0 * kPi * (m.k[q][0] * x + m.k[q][1] * y + m.k[q][2] * z) + m.ph[q]);
            acc[q] += s * (m.a[q][0] * a + m.a[q][1] * b + m.a[q][2] * e);
        }
    }
    for (int q = 0; q < m.n; q++) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < m.n; q++) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if ((int)threadIdx.x < m.n) part[blockIdx.x * kMaxModes + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_flame_project_final(int n, const double* __restrict__ part, double* __restrict__ ga) {
    const int q = threadIdx.x;
    if (q >= n) return;
    double s = 0.0;
    for (int b = 0; b < kProjBlocks; b++) s += part[b * kMaxModes + q];
    ga[q] = s;
}

// Per-block mode partials of ls_flame_project (kProjBlocks x kMaxModes
// doubles, 64 KB): one buffer per device, allocated by the first call on that
// device and kept until the process ends, so repeated projections (one per
// gradient) neither allocate nor free.  Calls on one device share it; the
// entry point synchronises its stream before returning, so calls from one
// host thread never overlap.
constexpr int kMaxProjDevices = 64;
std::mutex g_proj_mu;
double* g_proj_part[kMaxProjDevices] = {};

int project_scratch(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxProjDevices)
        return lsi::set_error(LS_ERR_DEVICE, "no usable current device");
    std::lock_guard<std::mutex> lock(g_proj_mu);
    if (!g_proj_part[dev] && hipMalloc((void**)&g_proj_part[dev], kProjBlocks * kMaxModes * sizeof(double)) != hipSuccess) {
        g_proj_part[dev] = nullptr;
        return lsi::set_error(LS_ERR_NOMEM, "projection scratch allocation failed");
    }
    *out = g_proj_part[dev];
    return LS_OK;
}

FlameModes draw_flame_modes(uint64_t seed, int n) {
    FlameModes m;
    std::memset(&m, 0, sizeof(m));
    m.n = n;
    SplitMix64 rng{seed};
    for (int q = 0; q < n; q++) {
        draw_solenoidal(rng, 3, m.k[q], m.a[q]);
        m.ph[q] = 2.0 * kPi * rng.uniform();
    }
    return m;
}

int flame_check(const ls_params_t* p, int32_t nc, int32_t nmodes) {
    int rc = lsi::validate(p);
    if (rc) return rc;
    if (p->dim != 3) return lsi::set_error(LS_ERR_INVALID, "the flame maps are 3D");
    if (nc < 2 || nc > 64) return lsi::set_error(LS_ERR_INVALID, "coarse grid size must be in [2, 64]");
    if (nmodes < 0 || nmodes > kMaxModes) return lsi::set_error(LS_ERR_INVALID, "at most 16 velocity modes");
    return LS_OK;
}

Coarse make_coarse(const ls_params_t& p, int nc) {
    Coarse cg;
    cg.nc = nc;
    cg.h[0] = (p.nx - 1) * p.dx / (nc - 1);
    cg.h[1] = (p.ny - 1) * p.dy / (nc - 1);
    cg.h[2] = (p.nz - 1) * p.dz / (nc - 1);
    return cg;
}

}  // namespace

namespace lsi {

int launch_deformation_velocity(const ls_params_t& p, double t, double period, double* u, double* v,
                                double* w, hipStream_t st) {
    const Coord c = make_coord(p);
    const int64_t n = (int64_t)c.nxt * c.nyt * c.nzt;
    hipLaunchKernelGGL(k_deform, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c, t, period, u, v, w, n);
    if (hipGetLastError() != hipSuccess) return set_error(LS_ERR_DEVICE, "deformation velocity launch failed");
    return LS_OK;
}

}  // namespace lsi

extern "C" int ls_init_case(const ls_params_t* p, int32_t case_id, uint64_t seed, double period,
                            double* G, double* u, double* v, double* w, void* stream) {
    int rc = lsi::validate(p);
    if (rc) return rc;
    if (!G || !u || !v) return lsi::set_error(LS_ERR_INVALID, "NULL field pointer");
    const bool is3 = case_id == LS_CASE_SPHERE_SWIRL3D || case_id == LS_CASE_PLANE_MODES3D ||
                     case_id == LS_CASE_SPHERE_DEFORM3D;
    const bool is2 = case_id == LS_CASE_CIRCLE2D || case_id == LS_CASE_ZALESAK2D || case_id == LS_CASE_FLAME2D;
    if (!is3 && !is2) return lsi::set_error(LS_ERR_INVALID, "unknown case id");
    if ((is3 && p->dim != 3) || (is2 && p->dim != 2))
        return lsi::set_error(LS_ERR_INVALID, "case dimension does not match params");
    if (is3 && !w) return lsi::set_error(LS_ERR_INVALID, "w is NULL for a 3D case");
    if ((case_id == LS_CASE_ZALESAK2D || case_id == LS_CASE_SPHERE_DEFORM3D) && !(period > 0.0))
        return lsi::set_error(LS_ERR_INVALID, "period must be > 0");

    Modes m;
    std::memset(&m, 0, sizeof(m));
    SplitMix64 rng{seed};
    if (case_id == LS_CASE_PLANE_MODES3D) {
        m.ng_modes = 8;
        for (int q = 0; q < 8; q++) {
            m.ga[q] = 0.02 * rng.uniform();
            m.gky[q] = 1.0 + std::floor(8.0 * rng.uniform());
            m.gkz[q] = 1.0 + std::floor(8.0 * rng.uniform());
            m.gph[q] = 2.0 * kPi * rng.uniform();
        }
        m.nv_modes = 16;
        m.vamp = 0.1;
        m.base[0] = 1.0;
        for (int q = 0; q < 16; q++) {
            draw_solenoidal(rng, 3, m.vk[q], m.va[q]);
            m.vph[q] = 2.0 * kPi * rng.uniform();
        }
    } else if (case_id == LS_CASE_FLAME2D) {
        m.circle = rng.uniform() < 0.5 ? 1 : 0;
        m.cx = 0.3 + 0.4 * rng.uniform();
        m.cy = 0.3 + 0.4 * rng.uniform();
        m.r = 0.1 + 0.1 * rng.uniform();
        m.x_off = 0.2 + 0.2 * rng.uniform();
        m.nv_modes = 8;
        m.vamp = 0.1;
        m.base[0] = 0.5;
        for (int q = 0; q < 8; q++) {
            draw_solenoidal(rng, 2, m.vk[q], m.va[q]);
            m.vph[q] = 2.0 * kPi * rng.uniform();
        }
    }
    const Coord c = make_coord(*p);
    const int64_t n = (int64_t)c.nxt * c.nyt * c.nzt;
    hipLaunchKernelGGL(k_case, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), c, (int)case_id, m, period, G, u, v,
                       p->dim == 3 ? w : nullptr, n);
    if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "case launch failed");
    return LS_OK;
}

extern "C" int ls_deformation_velocity(const ls_params_t* p, double t, double period, double* u,
                                       double* v, double* w, void* stream) {
    int rc = lsi::validate(p);
    if (rc) return rc;
    if (p->dim != 3 || !u || !v || !w) return lsi::set_error(LS_ERR_INVALID, "3D fields required");
    if (!(period > 0.0)) return lsi::set_error(LS_ERR_INVALID, "period must be > 0");
    return lsi::launch_deformation_velocity(*p, t, period, u, v, w, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ls_flame_fields(const ls_params_t* p, uint64_t mode_seed, int32_t nc, const double* sl_coarse,
                               int32_t nmodes, const double* amps, double* sl_field, double* u, double* v,
                               double* w, void* stream) {
    int rc = flame_check(p, nc, nmodes);
    if (rc) return rc;
    if (sl_field && !sl_coarse) return lsi::set_error(LS_ERR_INVALID, "sl_coarse is NULL");
    if (u && (!v || !w || (nmodes > 0 && !amps))) return lsi::set_error(LS_ERR_INVALID, "NULL velocity argument");
    if (!sl_field && !u) return LS_OK;
    const Coord c = make_coord(*p);
    const int64_t n = (int64_t)c.nxt * c.nyt * c.nzt;
    hipLaunchKernelGGL(k_flame_fields, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), c, draw_flame_modes(mode_seed, nmodes),
                       make_coarse(*p, nc), sl_coarse, amps, sl_field, u, v, w, n);
    if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "flame field launch failed");
    return LS_OK;
}

extern "C" int ls_flame_project(const ls_params_t* p, uint64_t mode_seed, int32_t nc, int32_t nmodes,
                                const double* gsl, const double* gu, const double* gv, const double* gw,
                                double* g_coarse, double* g_amps, void* stream) {
    int rc = flame_check(p, nc, nmodes);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const Coord c = make_coord(*p);
    if (g_coarse) {
        if (!gsl) return lsi::set_error(LS_ERR_INVALID, "gsl is NULL");
        hipLaunchKernelGGL(k_flame_project_sl, dim3((unsigned)(nc * nc * nc)), dim3(256), 0, st, c, p->nx, p->ny,
                           p->nz, make_coarse(*p, nc), gsl, g_coarse);
        if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "projection launch failed");
    }
    if (g_amps && nmodes > 0) {
        if (!gu || !gv || !gw) return lsi::set_error(LS_ERR_INVALID, "NULL velocity adjoint");
        double* part = nullptr;
        rc = project_scratch(&part);
        if (rc) return rc;
        hipLaunchKernelGGL(k_flame_project_modes, dim3(kProjBlocks), dim3(256), 0, st, c, p->nx, p->ny, p->nz,
                           draw_flame_modes(mode_seed, nmodes), gu, gv, gw, part);
        hipLaunchKernelGGL(k_flame_project_final, dim3(1), dim3(64), 0, st, (int)nmodes, part, g_amps);
        const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
        if (!ok) return lsi::set_error(LS_ERR_DEVICE, "projection launch failed");
    }
    return LS_OK;
}

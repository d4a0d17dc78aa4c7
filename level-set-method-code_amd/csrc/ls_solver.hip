// ls_solver.hip -- HIP kernels and the C ABI of the gfx950 level-set solver.
//
// Path per RK3 stage: one fused kernel computes the WENO-5 / u_eff RHS of every
// interior cell and applies the Shu-Osher combine in registers (the RHS field
// of the reference, rk3.h:54-150, is never stored), then one ghost-fill kernel
// refreshes all 3 ghost layers from the interior (boundary.h).  HCR-2 runs one
// kernel per pseudo-time iteration that recomputes the crossing flags and r~
// from phi_0 on the fly (reinitialization.h:46-193), followed by a ghost fill.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>

#include "ls_api.h"
#include "ls_exact.hpp"
#include "ls_fast.hpp"
#include "ls_internal.hpp"

using lsx::Geom;

namespace {

thread_local std::string g_err = "no error";

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define LS_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(LS_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kTX = 64;  // x extent of a block tile: one wavefront along x
constexpr int kTY = 4;   // y extent (4 wavefronts per block)
constexpr int kRedBlocks = 1024;

// ---------------------------------------------------------------------------
// Fused RHS + stage combine, exact mode.  One thread per interior cell; block
// (64,4) covers a 64x4 x-y tile (x-fastest, 512-byte coalesced rows); grid z
// walks the z planes.
template <int DIM, int STAGE>
__global__ __launch_bounds__(256) void k_stage_exact(Geom g, const double* __restrict__ Gs,
                                                     const double* __restrict__ Gn,
                                                     double* __restrict__ out,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ v,
                                                     const double* __restrict__ w,
                                                     const double* __restrict__ sl, double dt) {
    const int i = blockIdx.x * kTX + threadIdx.x;
    const int j = blockIdx.y * kTY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int k = DIM == 3 ? (int)blockIdx.z + g.ng : 0;
    const int64_t c = (int64_t)k * g.sz + (int64_t)(j + g.ng) * g.sy + (i + g.ng);
    const double s_l = sl ? sl[c] : g.s_l;
    const double r = lsx::rhs_cell<DIM>(g, Gs + c, u[c], v[c], DIM == 3 ? w[c] : 0.0, s_l);
    out[c] = lsx::stage_combine<STAGE>(Gn[c], Gs[c], r, dt);
}

// ---------------------------------------------------------------------------
// Ghost fill.  Each ghost cell takes G at the per-axis image of its index
// (periodic wrap or zero-gradient clamp), read from the interior only, so the
// three face sets can be written in one launch in any order: corner cells get
// the same value from every face set.  Linear extrapolation (y axis) computes
// from the two edge rows of the x/z-mapped column.
__device__ __forceinline__ int img(int q, int n, int ng, int bc) {
    if (q < ng) return bc == LS_BC_PERIODIC ? q + n : ng;
    if (q >= n + ng) return bc == LS_BC_PERIODIC ? q - n : n + ng - 1;
    return q;
}

__global__ __launch_bounds__(256) void k_fill_ghosts(Geom g, double* __restrict__ G,
                                                     int64_t nx_face, int64_t ny_face,
                                                     int64_t ntot) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntot) return;
    const int ng = g.ng;
    const int gw = 2 * ng;
    int i, j, k;
    if (t < nx_face) {  // x faces: (k, j, q)
        int64_t r = t;
        int q = (int)(r % gw); r /= gw;
        j = (int)(r % g.nyt); k = (int)(r / g.nyt);
        i = q < ng ? q : g.nx + q;
    } else if (t < nx_face + ny_face) {  // y faces: (k, q, i)
        int64_t r = t - nx_face;
        i = (int)(r % g.nxt); r /= g.nxt;
        int q = (int)(r % gw); k = (int)(r / gw);
        j = q < ng ? q : g.ny + q;
    } else {  // z faces: (q, j, i)
        int64_t r = t - nx_face - ny_face;
        i = (int)(r % g.nxt); r /= g.nxt;
        j = (int)(r % g.nyt); int q = (int)(r / g.nyt);
        k = q < ng ? q : g.nz + q;
    }
    const int si = img(i, g.nx, ng, g.bc[0]);
    const int sk = g.dim == 3 ? img(k, g.nz, ng, g.bc[2]) : 0;
    const int64_t col = (int64_t)sk * g.sz + si;
    const int64_t dst = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    double val;
    if (g.bc[1] == LS_BC_LIN_EXTRAP && (j < ng || j >= g.ny + ng)) {
        if (j < ng) {
            const double b0 = G[col + (int64_t)ng * g.sy];
            const double b1 = G[col + (int64_t)(ng + 1) * g.sy];
            const double s = b0 - b1;
            val = b0 + (ng - j) * s;  // ghost ng-1-q gets b0 + (q+1) s
        } else {
            const double t0 = G[col + (int64_t)(g.ny + ng - 1) * g.sy];
            const double t1 = G[col + (int64_t)(g.ny + ng - 2) * g.sy];
            const double s = t0 - t1;
            val = t0 + (j - g.ny - ng + 1) * s;
        }
    } else {
        const int sj = img(j, g.ny, ng, g.bc[1]);
        val = G[col + (int64_t)sj * g.sy];
    }
    G[dst] = val;
}

// ---------------------------------------------------------------------------
// HCR-2 pseudo-time step, exact mode.
template <int DIM>
__global__ __launch_bounds__(256) void k_hcr2_exact(Geom g, const double* __restrict__ pn,
                                                    const double* __restrict__ p0,
                                                    double* __restrict__ pnew) {
    const int i = blockIdx.x * kTX + threadIdx.x;
    const int j = blockIdx.y * kTY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int k = DIM == 3 ? (int)blockIdx.z + g.ng : 0;
    const int64_t c = (int64_t)k * g.sz + (int64_t)(j + g.ng) * g.sy + (i + g.ng);
    pnew[c] = lsx::hcr2_cell<DIM>(g, pn + c, p0 + c);
}

// HCR-2 set-up arrays (computeInterfaceCrossings), for callers that read them.
template <int DIM>
__global__ __launch_bounds__(256) void k_crossings(Geom g, const double* __restrict__ p0,
                                                   double* __restrict__ rt, int32_t* __restrict__ flag) {
    const int i = blockIdx.x * kTX + threadIdx.x;
    const int j = blockIdx.y * kTY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int k = DIM == 3 ? (int)blockIdx.z + g.ng : 0;
    const int64_t c = (int64_t)k * g.sz + (int64_t)(j + g.ng) * g.sy + (i + g.ng);
    const int64_t off[6] = {-1, 1, -g.sy, g.sy, -g.sz, g.sz};
    constexpr int NB = DIM == 3 ? 6 : 4;
    const double pc = p0[c];
    double sum = 0.0;
    int cnt = 0, f = 0;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const double pq = p0[c + off[q]];
        if (pc * pq < 0.0) {
            sum += pq;
            cnt++;
            f |= 1 << q;
        }
    }
    flag[c] = f;
    rt[c] = (cnt > 0 && fabs(sum) > 1e-15) ? pc / sum : 0.0;
}

// ---------------------------------------------------------------------------
// Diagnostics: per-block partial sums in a fixed order, then one block sums the
// partials in index order, so the result is deterministic run to run.
struct DiagPart {
    double sq;
    long long neg;
    long long bad;
};

__global__ __launch_bounds__(256) void k_diag_partial(Geom g, const double* __restrict__ G,
                                                      const double* __restrict__ G0,
                                                      DiagPart* __restrict__ part, int64_t ncell) {
    __shared__ double s_sq[256];
    __shared__ long long s_neg[256];
    __shared__ long long s_bad[256];
    double sq = 0.0;
    long long neg = 0, bad = 0;
    const int64_t nxy = (int64_t)g.nx * g.ny;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t / nxy);
        const int64_t r = t - (int64_t)k * nxy;
        const int j = (int)(r / g.nx), i = (int)(r - (int64_t)j * g.nx);
        const int64_t c = (int64_t)(g.dim == 3 ? k + g.ng : 0) * g.sz + (int64_t)(j + g.ng) * g.sy + i + g.ng;
        const double x = G[c];
        if (G0) {
            const double d = x - G0[c];
            sq += d * d;
        }
        neg += x < 0.0 ? 1 : 0;
        bad += isfinite(x) ? 0 : 1;
    }
    s_sq[threadIdx.x] = sq;
    s_neg[threadIdx.x] = neg;
    s_bad[threadIdx.x] = bad;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_sq[threadIdx.x] += s_sq[threadIdx.x + h];
            s_neg[threadIdx.x] += s_neg[threadIdx.x + h];
            s_bad[threadIdx.x] += s_bad[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = DiagPart{s_sq[0], s_neg[0], s_bad[0]};
}

__global__ __launch_bounds__(256) void k_diag_final(DiagPart* __restrict__ part, int nparts) {
    __shared__ double s_sq[256];
    __shared__ long long s_neg[256];
    __shared__ long long s_bad[256];
    double sq = 0.0;
    long long neg = 0, bad = 0;
    for (int t = threadIdx.x; t < nparts; t += blockDim.x) {
        sq += part[t].sq;
        neg += part[t].neg;
        bad += part[t].bad;
    }
    s_sq[threadIdx.x] = sq;
    s_neg[threadIdx.x] = neg;
    s_bad[threadIdx.x] = bad;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_sq[threadIdx.x] += s_sq[threadIdx.x + h];
            s_neg[threadIdx.x] += s_neg[threadIdx.x + h];
            s_bad[threadIdx.x] += s_bad[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[nparts] = DiagPart{s_sq[0], s_neg[0], s_bad[0]};
}

// The final partial as the three numbers ls_diag reports, on the device:
// the same correctly rounded quotient, square root and product as the host
// side of ls_diag, so the values are the same bits.
__global__ __launch_bounds__(64) void k_diag_emit(const DiagPart* __restrict__ fin, double* __restrict__ out,
                                                  double ncell, double cell, int have_g0) {
    if (threadIdx.x == 0) {
        const DiagPart r = *fin;
        out[0] = have_g0 ? sqrt(r.sq / ncell) : 0.0;
        out[1] = (double)r.neg * cell;
        out[2] = (double)r.bad;
    }
}

// Partial sums of ls_diag_noalloc: kRedBlocks block partials and the final
// one (24,600 bytes), one copy per GPU, held until the process ends.  Every
// call overwrites all of it before reading it, so nothing needs clearing;
// calls that overlap on one GPU (two streams, or two host threads) race on it.
// Host code never uses the symbol itself, only the address that
// hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) DiagPart scratch[kRedBlocks + 1];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
DiagPart* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call for that device and kept.
int scratch_on_device(DiagPart** out) {
    int dev = 0;
    LS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(LS_ERR_INVALID, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(LS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<DiagPart*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return LS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Solver object.
struct ls_solver {
    ls_params_t p;
    Geom g;
    int64_t n;          // doubles per padded field
    double* buf[2];     // workspace fields
    DiagPart* diag;     // kRedBlocks + 1 partials (device)
    DiagPart* diag_host;  // pinned host copy of the final partial
    int device;
    // ls_advance's plain RK3 loop as a captured graph of kGraphSteps steps
    // (made on first use, kept for the arguments in gkey/gdt)
    hipStream_t cap = nullptr;        // private stream the graph is captured on
    hipGraphExec_t gexec = nullptr;
    const void* gkey[6] = {};         // cur, a, u, v, w, sl at capture
    double gdt = 0.0;
};

// Steps per graph replay.  Even, so that the buffer rotation of the loop is
// back where it started after each replay.
constexpr int kGraphSteps = 16;

namespace lsi {

Geom make_geom(const ls_params_t& p) {
    Geom g;
    std::memset(&g, 0, sizeof(g));
    g.dim = p.dim;
    g.nx = p.nx;
    g.ny = p.ny;
    g.nz = p.dim == 3 ? p.nz : 1;
    g.ng = p.nghost;
    g.nxt = p.nx + 2 * p.nghost;
    g.nyt = p.ny + 2 * p.nghost;
    g.nzt = p.dim == 3 ? p.nz + 2 * p.nghost : 1;
    for (int a = 0; a < 3; a++) g.bc[a] = p.bc[a];
    g.sy = g.nxt;
    g.sz = (int64_t)g.nxt * g.nyt;
    g.dx = p.dx;
    g.dy = p.dy;
    g.dz = p.dim == 3 ? p.dz : 1.0;
    g.rdx = 1.0 / g.dx;
    g.rdy = 1.0 / g.dy;
    g.rdz = 1.0 / g.dz;
    g.s_l = p.s_l;
    g.eps = p.eps;
    g.dtau = p.reinit_dtau;
    g.beta = p.reinit_beta;
    return g;
}

int validate(const ls_params_t* p) {
    if (!p) return fail(LS_ERR_INVALID, "params is NULL");
    if (p->dim != 2 && p->dim != 3) return fail(LS_ERR_INVALID, "dim must be 2 or 3");
    if (p->nghost != 3) return fail(LS_ERR_INVALID, "nghost must be 3 (WENO-5 stencil)");
    const int nd = p->dim;
    const int n[3] = {p->nx, p->ny, nd == 3 ? p->nz : 3};
    for (int a = 0; a < 3; a++)
        if (p->bc[a] < 0 || p->bc[a] > 2) return fail(LS_ERR_INVALID, "unknown boundary kind");
    for (int a = 0; a < nd; a++) {
        if (n[a] < 3) return fail(LS_ERR_INVALID, "every interior extent must be >= 3");
        if (p->bc[a] < 0 || p->bc[a] > 2) return fail(LS_ERR_INVALID, "unknown boundary kind");
        if (p->bc[a] == LS_BC_LIN_EXTRAP && a != 1)
            return fail(LS_ERR_INVALID, "linear extrapolation is defined on the y axis only");
    }
    if (!(p->dx > 0.0) || !(p->dy > 0.0) || (nd == 3 && !(p->dz > 0.0)))
        return fail(LS_ERR_INVALID, "grid spacing must be positive");
    if (p->mode != LS_MODE_EXACT && p->mode != LS_MODE_FAST)
        return fail(LS_ERR_INVALID, "unknown arithmetic mode");
    if (p->velocity != LS_VEL_ARRAYS && p->velocity != LS_VEL_DEFORM3D)
        return fail(LS_ERR_INVALID, "unknown velocity kind");
    if (p->velocity == LS_VEL_DEFORM3D && (nd != 3 || !(p->vel_period > 0.0)))
        return fail(LS_ERR_INVALID, "deformation velocity needs dim 3 and vel_period > 0");
    if (p->reinit_interval < 0 || p->reinit_iters < 0)
        return fail(LS_ERR_INVALID, "reinit interval and iterations must be >= 0");
    const double cells = (double)(p->nx + 6) * (p->ny + 6) * (nd == 3 ? p->nz + 6 : 1);
    if (cells > 9.0e18) return fail(LS_ERR_INVALID, "grid too large");
    return LS_OK;
}

}  // namespace lsi

using lsi::make_geom;

static hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// The exact mode's division sequences on the device, one quotient per thread:
// kind 0 rdiv_u(a, b); kind 1 qdiv_u(a, b, y) with y = rdiv_u(1, b), as the
// weights use it; kind 2 div6_u(a).
__global__ void k_div_check(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out,
                            int64_t n, int kind) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = a[i];
    double q;
    if (kind == 0) q = lsx::rdiv_u(x, b[i]);
    else if (kind == 1) q = lsx::qdiv_u(x, b[i], lsx::rdiv_u(1.0, b[i]));
    else if (kind == 2) q = lsx::div6_u(x);
    else if (kind == 3) q = lsx::sqrt_u(x);
    else q = lsx::sqrt_u(fmax(x, 0x1p-767)) - 1.0;
    out[i] = q;
}

static dim3 cell_grid(const Geom& g) {
    return dim3((g.nx + kTX - 1) / kTX, (g.ny + kTY - 1) / kTY, g.dim == 3 ? g.nz : 1);
}

static int launch_fill(const ls_solver* s, double* G, hipStream_t st) {
    const Geom& g = s->g;
    const int64_t gw = 2 * g.ng;
    const int64_t nxf = gw * g.nyt * g.nzt;
    const int64_t nyf = gw * g.nxt * g.nzt;
    const int64_t nzf = g.dim == 3 ? gw * g.nxt * g.nyt : 0;
    const int64_t tot = nxf + nyf + nzf;
    const int64_t blocks = (tot + 255) / 256;
    hipLaunchKernelGGL(k_fill_ghosts, dim3((unsigned)blocks), dim3(256), 0, st, g, G, nxf, nyf, tot);
    LS_HIP(hipGetLastError());
    return LS_OK;
}

template <int DIM>
static int launch_stage(const ls_solver* s, int stage, const double* Gs, const double* Gn,
                        double* out, const double* u, const double* v, const double* w,
                        const double* sl, double dt, hipStream_t st) {
    const Geom& g = s->g;
    if (s->p.mode == LS_MODE_FAST) return lsf::launch_stage_fast<DIM>(g, stage, Gs, Gn, out, u, v, w, sl, dt, st);
    if constexpr (DIM == 3) {
        return lsf::launch_stage_exact3d(g, stage, Gs, Gn, out, u, v, w, sl, dt, st);
    } else {
        if (lsf::ymarch_fits(g)) return lsf::launch_stage_ymarch(g, stage, Gs, Gn, out, u, v, sl, dt, st, true);
        dim3 grid = cell_grid(g), block(kTX, kTY, 1);
        if (stage == 1)
            hipLaunchKernelGGL((k_stage_exact<DIM, 1>), grid, block, 0, st, g, Gs, Gn, out, u, v, w, sl, dt);
        else if (stage == 2)
            hipLaunchKernelGGL((k_stage_exact<DIM, 2>), grid, block, 0, st, g, Gs, Gn, out, u, v, w, sl, dt);
        else
            hipLaunchKernelGGL((k_stage_exact<DIM, 3>), grid, block, 0, st, g, Gs, Gn, out, u, v, w, sl, dt);
        LS_HIP(hipGetLastError());
        return LS_OK;
    }
}

static int stage(const ls_solver* s, int k, const double* Gs, const double* Gn, double* out,
                 const double* u, const double* v, const double* w, const double* sl, double dt,
                 hipStream_t st) {
    int rc = s->g.dim == 3 ? launch_stage<3>(s, k, Gs, Gn, out, u, v, w, sl, dt, st)
                           : launch_stage<2>(s, k, Gs, Gn, out, u, v, w, sl, dt, st);
    if (rc < 0) return rc;
    if (rc == 1) return LS_OK;  // the stage kernel wrote the ghost cells itself
    return launch_fill(s, out, st);
}

// One RK3 step Gn -> Gnew with temporaries t1, t2 (t2 != Gnew; t1 may equal Gnew).
static int rk3(const ls_solver* s, const double* Gn, double* Gnew, double* t1, double* t2,
               const double* u, const double* v, const double* w, const double* sl, double dt,
               hipStream_t st) {
    int rc;
    if ((rc = stage(s, 1, Gn, Gn, t1, u, v, w, sl, dt, st))) return rc;
    if ((rc = stage(s, 2, t1, Gn, t2, u, v, w, sl, dt, st))) return rc;
    return stage(s, 3, t2, Gn, Gnew, u, v, w, sl, dt, st);
}

static int hcr2_iter(const ls_solver* s, const double* pn, const double* p0, double* pnew,
                     hipStream_t st) {
    const Geom& g = s->g;
    if (s->p.mode == LS_MODE_FAST) {
        int rc = lsf::launch_hcr2_fast(g, pn, p0, pnew, st);
        if (rc < 0) return rc;
        if (rc == 1) return LS_OK;  // the kernel wrote the ghost cells itself
    } else {
        dim3 grid = cell_grid(g), block(kTX, kTY, 1);
        if (g.dim == 3)
            hipLaunchKernelGGL((k_hcr2_exact<3>), grid, block, 0, st, g, pn, p0, pnew);
        else
            hipLaunchKernelGGL((k_hcr2_exact<2>), grid, block, 0, st, g, pn, p0, pnew);
        LS_HIP(hipGetLastError());
    }
    return launch_fill(s, pnew, st);
}

// HCR-2 with phi_0 = *cur (already ghost-filled).  Iterates into the two
// spare buffers a/b; on return *cur points at the buffer holding the result
// and *a, *b at the two free buffers.
static int hcr2_rotate(const ls_solver* s, double** cur, double** a, double** b, hipStream_t st) {
    const int iters = s->p.reinit_iters;
    if (iters <= 0) return LS_OK;
    double* p0 = *cur;
    double* src = p0;
    double* dst = *a;
    int it = 0;
    if (iters >= 2) {
        // the first two iterations in one pass where the grid allows it
        int rc = lsf::launch_hcr2_pair(s->g, p0, *b, st, s->p.mode != LS_MODE_FAST);
        if (rc < 0) return rc;
        if (rc == 1) {
            src = *b;
            dst = *a;
            it = 2;
        }
    }
    for (; it < iters; it++) {
        int rc = hcr2_iter(s, src, p0, dst, st);
        if (rc) return rc;
        src = dst;
        dst = (src == *a) ? *b : *a;
    }
    // src holds the result; the two other buffers of {p0, a, b} are free.
    double* bufs[3] = {p0, *a, *b};
    double* fr[2];
    int nf = 0;
    for (int q = 0; q < 3; q++)
        if (bufs[q] != src) fr[nf++] = bufs[q];
    *cur = src;
    *a = fr[0];
    *b = fr[1];
    return LS_OK;
}

extern "C" {

int64_t ls_field_size(const ls_params_t* p) {
    if (lsi::validate(p)) return -1;
    return (int64_t)(p->nx + 2 * p->nghost) * (p->ny + 2 * p->nghost) *
           (p->dim == 3 ? (p->nz + 2 * p->nghost) : 1);
}

int ls_create(const ls_params_t* p, ls_solver_t** out) {
    if (!out) return fail(LS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = lsi::validate(p);
    if (rc) return rc;
    int device = 0;
    LS_HIP(hipGetDevice(&device));
    ls_solver* s = new ls_solver();
    s->p = *p;
    s->g = make_geom(*p);
    s->n = ls_field_size(p);
    s->device = device;
    for (int q = 0; q < 2; q++) {
        if (hipMalloc((void**)&s->buf[q], (size_t)s->n * sizeof(double)) != hipSuccess) {
            for (int r = 0; r < q; r++) (void)hipFree(s->buf[r]);
            delete s;
            return fail(LS_ERR_NOMEM, "hipMalloc of solver workspace failed");
        }
    }
    if (hipMalloc((void**)&s->diag, (kRedBlocks + 1) * sizeof(DiagPart)) != hipSuccess ||
        hipHostMalloc((void**)&s->diag_host, sizeof(DiagPart), hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(s->buf[0]);
        (void)hipFree(s->buf[1]);
        delete s;
        return fail(LS_ERR_NOMEM, "allocation of reduction scratch failed");
    }
    // Zero the workspace so that never-written cells are defined.
    for (int q = 0; q < 2; q++) {
        if (hipMemset(s->buf[q], 0, (size_t)s->n * sizeof(double)) != hipSuccess) {
            ls_destroy(s);
            return fail(LS_ERR_DEVICE, "hipMemset of solver workspace failed");
        }
    }
    *out = s;
    return LS_OK;
}

int ls_destroy(ls_solver_t* s) {
    if (!s) return LS_OK;
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->cap) (void)hipStreamDestroy(s->cap);
    (void)hipFree(s->buf[0]);
    (void)hipFree(s->buf[1]);
    (void)hipFree(s->diag);
    (void)hipHostFree(s->diag_host);
    delete s;
    return LS_OK;
}

int ls_fill_ghosts(ls_solver_t* s, double* G, void* stream) {
    if (!s || !G) return fail(LS_ERR_INVALID, "NULL argument");
    return launch_fill(s, G, S(stream));
}

int ls_rk3_step(ls_solver_t* s, const double* G_n, double* G_new, const double* u,
                const double* v, const double* w, const double* sl_field, double dt,
                void* stream) {
    if (!s || !G_n || !G_new || !u || !v || (s->g.dim == 3 && !w))
        return fail(LS_ERR_INVALID, "NULL field pointer");
    if (G_new == G_n) return fail(LS_ERR_INVALID, "G_new must not alias G_n");
    if (!std::isfinite(dt)) return fail(LS_ERR_INVALID, "dt must be finite");
    return rk3(s, G_n, G_new, s->buf[0], s->buf[1], u, v, w, sl_field, dt, S(stream));
}

int ls_rk3_stage(ls_solver_t* s, int32_t k, const double* G_s, const double* G_n, double* out,
                 const double* u, const double* v, const double* w, const double* sl_field,
                 double dt, void* stream) {
    if (!s || !G_s || !G_n || !out || !u || !v || (s->g.dim == 3 && !w))
        return fail(LS_ERR_INVALID, "NULL field pointer");
    if (k < 1 || k > 3) return fail(LS_ERR_INVALID, "stage must be 1, 2 or 3");
    if (out == G_s || out == G_n) return fail(LS_ERR_INVALID, "out must not alias an input");
    if (!std::isfinite(dt)) return fail(LS_ERR_INVALID, "dt must be finite");
    return stage(s, k, G_s, G_n, out, u, v, w, sl_field, dt, S(stream));
}

int ls_hcr2_step(ls_solver_t* s, const double* phi_n, const double* phi_0, double* phi_new,
                 void* stream) {
    if (!s || !phi_n || !phi_0 || !phi_new) return fail(LS_ERR_INVALID, "NULL argument");
    if (phi_new == phi_n || phi_new == phi_0) return fail(LS_ERR_INVALID, "phi_new must not alias an input");
    return hcr2_iter(s, phi_n, phi_0, phi_new, S(stream));
}

int ls_hcr2_crossings(ls_solver_t* s, const double* phi_0, double* r_tilde, int32_t* flag,
                      void* stream) {
    if (!s || !phi_0 || !r_tilde || !flag) return fail(LS_ERR_INVALID, "NULL argument");
    const Geom& g = s->g;
    dim3 grid = cell_grid(g), block(kTX, kTY, 1);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_crossings<3>), grid, block, 0, S(stream), g, phi_0, r_tilde, flag);
    else
        hipLaunchKernelGGL((k_crossings<2>), grid, block, 0, S(stream), g, phi_0, r_tilde, flag);
    LS_HIP(hipGetLastError());
    return LS_OK;
}

int ls_reinit_hcr2(ls_solver_t* s, double* G, void* stream) {
    if (!s || !G) return fail(LS_ERR_INVALID, "NULL argument");
    hipStream_t st = S(stream);
    // phi_0 is G after a ghost fill (reinitialization.h:333-336); filling a
    // field whose ghosts are already current changes nothing.
    int rc = launch_fill(s, G, st);
    if (rc) return rc;
    double* cur = G;
    double* a = s->buf[0];
    double* b = s->buf[1];
    rc = hcr2_rotate(s, &cur, &a, &b, st);
    if (rc) return rc;
    if (cur != G) LS_HIP(hipMemcpyAsync(G, cur, (size_t)s->n * sizeof(double), hipMemcpyDeviceToDevice, st));
    return LS_OK;
}

// Captures kGraphSteps RK3 steps of the plain loop (state in cur, spare
// buffers a and b) into s->gexec, on the solver's private stream.  Nothing runs.
static int capture_loop(ls_solver* s, double* cur, double* a, double* b, const double* u,
                        const double* v, const double* w, const double* sl, double dt) {
    if (s->gexec) {
        (void)hipGraphExecDestroy(s->gexec);
        s->gexec = nullptr;
    }
    if (!s->cap) LS_HIP(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
    LS_HIP(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
    int rc = LS_OK;
    for (int q = 0; q < kGraphSteps && rc == LS_OK; q++) {
        rc = rk3(s, cur, a, a, b, u, v, w, sl, dt, s->cap);
        std::swap(cur, a);
    }
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s->cap, &graph);
    if (rc != LS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(LS_ERR_DEVICE, "graph capture of the RK3 loop failed");
    const hipError_t ei = hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        s->gexec = nullptr;
        return fail(LS_ERR_DEVICE, "graph instantiation of the RK3 loop failed");
    }
    return LS_OK;
}

int ls_advance(ls_solver_t* s, double* G, double* u, double* v, double* w,
               const double* sl_field, double dt, double t_final, int64_t step0,
               int64_t max_steps, double* t_io, int64_t* steps_done, double* dt_last,
               void* stream) {
    if (!s || !G || !u || !v || (s->g.dim == 3 && !w) || !t_io)
        return fail(LS_ERR_INVALID, "NULL argument");
    if (!(dt > 0.0) || !std::isfinite(t_final)) return fail(LS_ERR_INVALID, "dt must be > 0");
    hipStream_t st = S(stream);
    double t = *t_io;
    int64_t step = step0;
    const int64_t stop = step0 + max_steps;
    const int ri = s->p.reinit_interval;
    // Buffer rotation: the state lives in `cur`; `a` and `b` are free.  An RK3
    // step writes stage 1 into a, stage 2 into b and stage 3 back into a.
    double* cur = G;
    double* a = s->buf[0];
    double* b = s->buf[1];
    int rc;
    // Without reinitialisation and with fixed velocity arrays, runs of
    // kGraphSteps full steps replay a captured graph: the same launches with
    // the same arguments, so the same bits, at a fraction of the launch cost
    // (the small grids are launch-bound).  The host still advances t step by
    // step, so the loop's time and sliver step are unchanged.
    const bool plain = ri == 0 && s->p.velocity != LS_VEL_DEFORM3D;
    while (plain && step + kGraphSteps <= stop) {
        double tt = t;
        bool full = true;
        for (int q = 0; q < kGraphSteps && full; q++) {
            if (!(tt < t_final) || tt + dt > t_final) full = false;
            tt += dt;
        }
        if (!full) break;
        const void* key[6] = {cur, a, u, v, w, sl_field};
        if (!s->gexec || s->gdt != dt || std::memcmp(key, s->gkey, sizeof(key)) != 0) {
            if ((rc = capture_loop(s, cur, a, b, u, v, w, sl_field, dt))) return rc;
            std::memcpy(s->gkey, key, sizeof(key));
            s->gdt = dt;
        }
        LS_HIP(hipGraphLaunch(s->gexec, st));
        t = tt;
        step += kGraphSteps;
    }
    while (t < t_final && step < stop) {
        if (t + dt > t_final) dt = t_final - t;
        if (s->p.velocity == LS_VEL_DEFORM3D) {
            if ((rc = lsi::launch_deformation_velocity(s->p, t, s->p.vel_period, u, v, w, st))) return rc;
        }
        if ((rc = rk3(s, cur, a, a, b, u, v, w, sl_field, dt, st))) return rc;
        double* old = cur;
        cur = a;
        a = old;
        if (ri > 0 && step > 0 && step % ri == 0) {
            if ((rc = hcr2_rotate(s, &cur, &a, &b, st))) return rc;
        }
        t += dt;
        step++;
    }
    if (cur != G) LS_HIP(hipMemcpyAsync(G, cur, (size_t)s->n * sizeof(double), hipMemcpyDeviceToDevice, st));
    LS_HIP(hipStreamSynchronize(st));
    *t_io = t;
    if (steps_done) *steps_done = step - step0;
    if (dt_last) *dt_last = dt;
    return LS_OK;
}

int ls_diag(ls_solver_t* s, const double* G, const double* G0, double* l2, double* area,
            void* stream) {
    if (!s || !G) return fail(LS_ERR_INVALID, "NULL argument");
    hipStream_t st = S(stream);
    const Geom& g = s->g;
    const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_diag_partial, dim3(kRedBlocks), dim3(256), 0, st, g, G, G0, s->diag, ncell);
    LS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, s->diag, kRedBlocks);
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpyAsync(s->diag_host, s->diag + kRedBlocks, sizeof(DiagPart), hipMemcpyDeviceToHost, st));
    LS_HIP(hipStreamSynchronize(st));
    const DiagPart r = *s->diag_host;
    if (l2 && G0) *l2 = std::sqrt(r.sq / (double)ncell);
    if (area) {
        const double cell = g.dim == 3 ? g.dx * g.dy * g.dz : g.dx * g.dy;
        *area = (double)r.neg * cell;
    }
    return LS_OK;
}

int ls_check_finite(ls_solver_t* s, const double* G, void* stream) {
    if (!s || !G) return fail(LS_ERR_INVALID, "NULL argument");
    hipStream_t st = S(stream);
    const Geom& g = s->g;
    const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_diag_partial, dim3(kRedBlocks), dim3(256), 0, st, g, G, (const double*)nullptr, s->diag, ncell);
    LS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, s->diag, kRedBlocks);
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpyAsync(s->diag_host, s->diag + kRedBlocks, sizeof(DiagPart), hipMemcpyDeviceToHost, st));
    LS_HIP(hipStreamSynchronize(st));
    if (s->diag_host->bad != 0) return fail(LS_ERR_NONFINITE, "non-finite value in G");
    return LS_OK;
}

int ls_diag_noalloc(const ls_params_t* p, const double* G, const double* G0, double* out, void* stream) {
    int rc = lsi::validate(p);
    if (rc) return rc;
    if (!G || !out) return fail(LS_ERR_INVALID, "NULL argument");
    DiagPart* part = nullptr;
    rc = scratch_on_device(&part);
    if (rc) return rc;
    hipStream_t st = S(stream);
    const Geom g = make_geom(*p);
    const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
    const double cell = g.dim == 3 ? g.dx * g.dy * g.dz : g.dx * g.dy;
    hipLaunchKernelGGL(k_diag_partial, dim3(kRedBlocks), dim3(256), 0, st, g, G, G0, part, ncell);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, part, kRedBlocks);
    hipLaunchKernelGGL(k_diag_emit, dim3(1), dim3(64), 0, st, (const DiagPart*)(part + kRedBlocks), out,
                       (double)ncell, cell, G0 ? 1 : 0);
    LS_HIP(hipGetLastError());
    return LS_OK;
}

int ls_div_check(const double* a, const double* b, double* out, int64_t n, int32_t kind, void* stream) {
    if (!a || !out || (kind < 2 && !b)) return fail(LS_ERR_INVALID, "NULL argument");
    if (n < 0 || kind < 0 || kind > 4) return fail(LS_ERR_INVALID, "bad count or kind");
    if (n == 0) return LS_OK;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffLL) return fail(LS_ERR_INVALID, "count too large for one launch");
    hipLaunchKernelGGL(k_div_check, dim3((unsigned)blocks), dim3(256), 0, S(stream), a, b, out, n, kind);
    LS_HIP(hipGetLastError());
    return LS_OK;
}

const char* ls_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace lsi {
int set_error(int code, const char* msg) { return fail(code, msg); }
const ls_params_t& solver_params(const ls_solver_t* s) { return s->p; }
int solver_device(const ls_solver_t* s) { return s->device; }
void* solver_scratch(ls_solver_t* s, size_t* bytes) {
    *bytes = (kRedBlocks + 1) * sizeof(DiagPart);
    return s->diag;
}
}  // namespace lsi

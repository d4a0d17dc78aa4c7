// ls_fast.hip -- fast-mode kernels: the same scheme as the exact mode with the
// arithmetic reorganised for CDNA4 FP64 throughput.  Results agree with the
// exact mode / CPU oracle to rounding (tested at 1e-12 relative in G and in
// zero-level-set position), not bit for bit.
//
// Per axis and cell the reference performs 2 WENO-5 reconstructions with 9
// correctly rounded divisions each (Local_code/level-set_MPI_3D/include/weno5.h:39-81),
// which the exact mode reproduces bit for bit.  Here:
//  * the right-biased branch is the left-biased formula on the reversed
//    stencil with the linear weights (0.1, 0.6, 0.3) swapped to (0.3, 0.6, 0.1)
//    (this is the reference's weno5_right, weno5.h:60-81), so the upwind
//    choice is a select of six stencil values, not a divergent branch;
//  * nonlinear weights use the product form
//        w_k = d_k * prod_{j != k} t_j,   t_j = (eps + beta_j)^2,
//    which equals d_k/t_k / sum_j d_j/t_j without any division (the kernels
//    hold 16 t_j, a power-of-two factor that cancels exactly: see tsq);
//  * the two face values enter the derivative as one fraction
//        (N_b D_a - N_a D_b) / (6 h D_a D_b),
//    and all axes of a cell go over one denominator, so a cell costs a single
//    reciprocal (v_rcp_f64 + Newton steps);
//  * the stage kernels reconstruct each face once, in the orientations its
//    two cells need, and share it between them (k_stage_zface in 3D, the
//    y-march of ls_ymarch.hpp in 2D);
//  * constant divisions (by 6, 2h, h) become multiplies, FMA is used freely.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ls_fast.hpp"
#include "ls_fast_math.hpp"
#include "ls_internal.hpp"
#include "ls_ymarch.hpp"

namespace lsf {

using lsx::Geom;

struct FastGeom {
    Geom g;
    double inv2dx, inv2dy, inv2dz;
    double inv6dx, inv6dy, inv6dz;
    double invdx, invdy, invdz;
    int tiles_x, tiles_y;      // block tiles per plane
    int64_t nblocks;
};

constexpr int kFX = 64;
constexpr int kFY = 4;

// Block id -> tile with an XCD-aware order: workgroups are dealt round-robin
// to the 8 XCDs, so block b runs on XCD b%8.  Give XCD e the contiguous range
// [e*per, (e+1)*per) of tiles (x fastest, then y, then z) so that the y and z
// neighbours of a tile are read through the same L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
    const int64_t per = (nb + 7) / 8;
    const int64_t e = b & 7, r = b >> 3;
    return e * per + r;
}

// 2D stage: one wavefront per (64-column strip, chunk of rows), y-marching
// (ls_ymarch.hpp).  Units are dealt XCD-contiguously, strips fastest.  The
// exact variant is held to four waves per SIMD (127 VGPRs).
template <int STAGE, bool EXACT>
__global__ __launch_bounds__(kYW) __attribute__((amdgpu_waves_per_eu(EXACT ? 4 : 1)))
void k_stage_ymarch(YGrid m, int strips, int ch, int64_t units, int64_t nblocks, const double* __restrict__ Gs,
                    const double* __restrict__ Gn, double* __restrict__ out, const double* __restrict__ u,
                    const double* __restrict__ v, const double* __restrict__ sl) {
    __shared__ YLds<EXACT> L;
    const int64_t t = xcd_tile(blockIdx.x, nblocks);
    if (t >= units) return;
    const int sx = (int)(t % strips), cy = (int)(t / strips);
    const int y0 = cy * ch;
    ymarch_strip<STAGE, EXACT>(m, sx * kYW, y0, min(y0 + ch, m.ny), Gs, Gn, out, u, v, sl, L);
}

// 2D stage for small grids: one thread per cell, 64x4 tiles, the y-march's
// arithmetic (ycell).
template <int STAGE>
__global__ __launch_bounds__(256) void k_stage_ycell(YGrid m, int tiles_x, int64_t tiles, int64_t nblocks,
                                                     const double* __restrict__ Gs, const double* __restrict__ Gn,
                                                     double* __restrict__ out, const double* __restrict__ u,
                                                     const double* __restrict__ v, const double* __restrict__ sl) {
    const int64_t t = xcd_tile(blockIdx.x, nblocks);
    if (t >= tiles) return;
    const int i = (int)(t % tiles_x) * kFX + (int)threadIdx.x;
    const int j = (int)(t / tiles_x) * kFY + (int)threadIdx.y;
    if (i >= m.nx || j >= m.ny) return;
    ycell<STAGE>(m, i, j, Gs, Gn, out, u, v, sl);
}

// ---------------------------------------------------------------------------
// 3D kernels march along z: a block owns a column of cells of the x-y plane
// and walks a chunk of z planes, keeping the column's z stencil in registers
// and the current plane (tile + halo) in LDS.  Chunks are assigned so that
// each XCD walks one z-slab: the y-halo rows a block reads were loaded as its
// neighbours' planes moments before and are still in that XCD's L2.  Ghost
// cells of the output are written by the interior cell they copy (periodic /
// zero-gradient images), which replaces the separate ghost-fill launch.
constexpr int ZX = 64, ZY = 4, ZH = 3;
#ifndef LS_ZMARCH_KZ
#define LS_ZMARCH_KZ 32  // z planes per block at large nz
#endif
// Planes per block for a z-march over tiles of tile_cols x tile_rows cells:
// LS_ZMARCH_KZ where the grid is large, shorter chunks (down to 8 planes)
// where that keeps at least ~2048 blocks in flight (two rounds of four
// blocks on each of 256 CUs), since a chunk's 6-plane start-up is cheaper
// than idle CUs.
static int zmarch_chunk(const Geom& g, int tile_cols, int tile_rows) {
    const int64_t tiles = (int64_t)((g.nx + tile_cols - 1) / tile_cols) * ((g.ny + tile_rows - 1) / tile_rows);
    const int64_t want = 2048;
    int kz = LS_ZMARCH_KZ;
    while (kz > 8 && tiles * ((g.nz + kz - 1) / kz) < want) kz /= 2;
    return kz;
}

// Kernel arguments of the z-marching kernels, kept small so that they stay in
// scalar registers.
struct ZGeom {
    int32_t nx, ny, nz;
    int32_t sy, sz;               // strides in doubles (sz*nzt < 2^31 checked at launch)
    int32_t bc[3];
    int32_t tiles_x, tiles_y, kz, fuse_ghosts;
    int64_t ntiles, nblocks;
    double s_l, eps;
    double inv2dx, inv2dy, inv2dz, inv6dx, inv6dy, inv6dz;
    double dx, dy, dz;            // exact-mode stage kernel: the reference divides by these
    double rdx, rdy, rdz;         // RN(1/dx), RN(1/dy), RN(1/dz) for lsx::qdiv_u
    int32_t crdiv;                // spacings, eps and a scalar S_L admit the FMA division sequences
};


// ---------------------------------------------------------------------------
// 3D stage kernel with shared faces.  The upwind derivative of a cell along
// an axis is (F(c+1/2) - F(c-1/2)) / h, where both faces are reconstructed
// in the cell's upwind orientation: left-biased (weno5_left on G(c-2..c+2)
// for the face c+1/2) when u_eff >= 0, mirrored (weno5_right) otherwise
// (Local_code/level-set_MPI_3D/include/weno5.h:90-196).  A face is shared by
// the two cells beside it, so it is reconstructed once, in each orientation
// one of those cells needs -- usually both cells agree and one reconstruction
// serves two cells:
//  * x and y faces: after the cells of a plane have their u_eff, each cell
//    builds the face on its high side in the orientations of itself and its
//    high neighbour and leaves it in LDS; the faces on the low edge of the
//    tile are built by spare lanes of two waves;
//  * z faces: a column builds the face k+1/2 in the orientation of cell k
//    and carries it to plane k+1, which rebuilds it only where its own
//    orientation differs.
// A face is kept as numerator n and denominator d, F = n / (6 d), with the
// product-form weights (see face_core).  The three axis derivatives enter one
// fraction, so a cell costs a single refined reciprocal:
//    rhs = -(sum_a (u_a / 6h_a) N_a prod_{b != a} D_b) / prod_a D_a,
//    N_a = n_hi d_lo - n_lo d_hi,  D_a = d_hi d_lo.
// Smoothness terms T(c) = 16 (eps + beta)^2 of the sub-stencil centred at c
// do not depend on the orientation and are computed once per point.
// Ghost images of a freshly computed interior value (periodic / zero-gradient).
__device__ __forceinline__ void store_images(const ZGeom& z, double* __restrict__ out, int i, int j, int k,
                                             double o) {
    const int64_t sy = z.sy, sz = z.sz;
    const int nxi = n_images(i, z.nx, z.bc[0]);
    const int nyi = n_images(j, z.ny, z.bc[1]);
    const int nzi = n_images(k, z.nz, z.bc[2]);
    for (int a = 0; a < nzi; a++) {
        const int64_t pk = (int64_t)image(k, z.nz, z.bc[2], a) * sz;
        for (int b = 0; b < nyi; b++) {
            const int64_t pj = pk + (int64_t)image(j, z.ny, z.bc[1], b) * sy;
            for (int e = 0; e < nxi; e++)
                if (a | b | e) out[pj + image(i, z.nx, z.bc[0], e)] = o;
        }
    }
}

// Tile of the shared-face kernel: 64 x rows cells per plane, one cell per
// thread, one wave per row.  Besides its row, a wave may take a share of the
// tile's edge work: waves 0-3 the y-edge smoothness terms (one row each),
// wave kWaveTx the x-edge terms, waves kWaveFx and kWaveFy the x and y faces
// on the low tile edges.  The edge faces cost two wave passes per plane
// whatever the height, so the exact mode, whose faces are several times
// dearer, uses 8 rows (3.62 -> 3.58 ms at 512^3; 16 rows: 4.02 ms, from
// barrier skew over 16 waves); the fast mode keeps 4 (8 rows: 1.78 -> 1.80 ms).
#ifndef LS_ZFACE_ROWS
#define LS_ZFACE_ROWS 4  // fast mode
#endif
#ifndef LS_ZFACE_ROWS_EXACT
#define LS_ZFACE_ROWS_EXACT 8
#endif
constexpr int FX = 64, FGW = FX + 2 * ZH;  // LDS plane width: tile + 3-cell halo
template <bool EXACT>
struct ZTile {
    static constexpr int FY = EXACT ? LS_ZFACE_ROWS_EXACT : LS_ZFACE_ROWS;
    static_assert(FY == 4 || FY == 8 || FY == 16, "tile rows");
    static constexpr int kWaveTx = FY == 4 ? 2 : 5, kWaveFx = FY == 4 ? 0 : FY - 2, kWaveFy = FY == 4 ? 1 : FY - 1;
    static constexpr int FGH = FY + 2 * ZH;
    static constexpr int kFaceThreads = FX * FY;
    static constexpr int kHaloItems = 6 * FX + 6 * FY;   // G halo values per plane
    static constexpr int NH = (kHaloItems + kFaceThreads - 1) / kFaceThreads;  // per thread
};

// A stored face as the value (exact mode) or the (n, d) pair (fast mode);
// each overload is called only in its own mode.
__device__ __forceinline__ double val(double f) { return f; }
__device__ __forceinline__ double val(const double2& f) { return f.x; }
__device__ __forceinline__ double2 as2(const double2& f) { return f; }
__device__ __forceinline__ double2 as2(double f) { return make_double2(f, 1.0); }

// Exact-mode faces: the reference's weno5_left on G(c-2..c+2) and
// weno5_right on G(c-1..c+3) for the face c+1/2 (weno5.h:39-81), in its own
// operation order.  A face value depends only on those five values, so the
// two cells beside a face get the same bits whether it is built once or
// twice.  g points at G(c) with stride GS; D is the division flavour
// (ls_crdiv.hpp).
template <int GS, int D>
__device__ __forceinline__ double xface_lds(const double* g, bool mirror) {
    return mirror ? lsx::weno_r<D>(g[-GS], g[0], g[GS], g[2 * GS], g[3 * GS])
                  : lsx::weno_l<D>(g[-2 * GS], g[-GS], g[0], g[GS], g[2 * GS]);
}

using DivIEEE = std::integral_constant<int, lsx::kDivIEEE>;
using DivFma = std::integral_constant<int, lsx::kDivFma>;

// EXACT: the exact mode's arithmetic (bitwise the CPU oracle) on the same
// shared-face z-march; faces are then single values and the smoothness terms
// are computed inside each reconstruction, as the reference does.  The
// divisions run as the FMA sequences of ls_crdiv.hpp while every G and S_L
// value the block has loaded passed crdiv_safe (each thread checks what it
// loads; a sticky block flag in LDS collects the failures) and the host found
// the spacings in range (z.crdiv); otherwise as divisions.  Both give the
// same bits, so waves may differ in which they use.
// Launch variants of k_stage_zface (VAR bits): a per-cell S_L field; G_s is
// G_n (stage 1 of a step: no separate G_n load); every tile full (no bounds
// tests).  Each drops registers or lane masks the general kernel carries.
constexpr int kVarSLF = 1, kVarSame = 2, kVarFull = 4;

template <int STAGE, bool EXACT, int VAR>
__global__ __launch_bounds__(ZTile<EXACT>::kFaceThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_stage_zface(
    ZGeom z, const double* __restrict__ Gs, const double* __restrict__ Gn, double* __restrict__ out,
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
    const double* __restrict__ sl, double dt) {
    using TT = ZTile<EXACT>;
    constexpr int FY = TT::FY, kWaveTx = TT::kWaveTx, kWaveFx = TT::kWaveFx, kWaveFy = TT::kWaveFy;
    constexpr int FGH = TT::FGH, kFaceThreads = TT::kFaceThreads, kHaloItems = TT::kHaloItems;
    constexpr int NH = TT::NH;
    using FV = typename std::conditional<EXACT, double, double2>::type;  // a stored face
    constexpr bool SLF = VAR & kVarSLF, SAME = VAR & kVarSame, FULL = VAR & kVarFull;
    __shared__ double sG[FGH][FGW];                      // plane k: tile + 3-cell halo
    __shared__ double sTx[EXACT ? 1 : FY][FX + 4];       // T along x at local x = -2..FX+1
    __shared__ double sTy[EXACT ? 1 : FY + 4][FX];       // T along y at local y = -2..FY+1
    // exact mode: sub-stencil data (t, rcp_u(t)) per centre, [left-biased, mirrored]
    __shared__ double2 sSx[EXACT ? 2 : 1][EXACT ? FY : 1][EXACT ? FX + 4 : 1];  // local x = -2..FX+1
    __shared__ double2 sSy[EXACT ? 2 : 1][EXACT ? FY + 4 : 1][EXACT ? FX : 1];  // local y = -2..FY+1
    __shared__ FV sFx[2][FY][FX + 1];                    // [left-biased, mirrored] x face between local x s-1 and s
    __shared__ FV sFy[2][FY + 1][FX];                    // [left-biased, mirrored] y face between local y s-1 and s
    __shared__ unsigned char sOri[FY][FX];     // bit 0: u_eff >= 0, bit 1: v_eff >= 0
    __shared__ int sBad;                       // exact mode: some loaded value failed crdiv_safe
    const int64_t t = xcd_tile(blockIdx.x, z.nblocks);
    if (t >= z.ntiles) return;  // block-uniform
    const int tx = (int)(t % z.tiles_x);
    const int64_t r = t / z.tiles_x;
    const int ty = (int)(r % z.tiles_y);
    const int ch = (int)(r / z.tiles_y);
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * FX + lx;
    const int x0 = tx * FX, y0 = ty * FY;
    const int i = x0 + lx, j = y0 + ly;
    const bool active = FULL || (i < z.nx && j < z.ny);
    const int kb = ch * z.kz;
    const int ke = min(kb + z.kz, z.nz);
    const int64_t sz = z.sz;
    const double c2 = 0.25 * dt, c3 = (2.0 / 3.0) * dt;
    bool bad = false;  // exact mode: a value this thread loaded failed crdiv_safe
    if (EXACT) {
        if (tid == 0) sBad = 0;
        __syncthreads();  // before any thread can raise the flag
    }

    // Loads are unconditional: a position past the padded array reads the
    // nearest padded cell instead.  Such positions feed only the stencils of
    // cells outside the interior.  Offsets within a plane are int32 (the
    // launcher checks the padded size).
    auto pad = [&](int gx, int gy) { return (min(gy, z.ny + 2) + 3) * z.sy + min(gx, z.nx + 2) + 3; };
    const int32_t col = pad(i, j);
    // Halo items tid + a * kFaceThreads: below 6 FX the rows at local y
    // -3..-1 and FY..FY+2, then 6 FY items the columns at local x -3..-1 and
    // FX..FX+2.
    int hrow[NH], hcol[NH];
    int32_t ohalo[NH];
#pragma unroll
    for (int a = 0; a < NH; a++) {
        const int h = min(tid + a * kFaceThreads, kHaloItems - 1);
        if (h < 6 * FX) {
            const int q = h / FX;
            hrow[a] = q < 3 ? q : q + FY;
            hcol[a] = h - q * FX + 3;
        } else {
            const int e = h - 6 * FX, c6 = e % 6;
            hrow[a] = e / 6 + 3;
            hcol[a] = c6 < 3 ? c6 : c6 + FX;
        }
        ohalo[a] = pad(x0 - 3 + hcol[a], y0 - 3 + hrow[a]);
    }

    struct Next {
        double lead, u, v, w, gn, sl, halo[NH];
    };
    // The plane loads are buffer loads: a descriptor based at the plane
    // (block-uniform, built in scalar registers) and a 32-bit byte offset per
    // thread that stays fixed for the whole march, so no load needs
    // per-thread 64-bit address arithmetic.  Plane offsets fit: sz < 2^29
    // follows from the padded-size check at launch.
    const uint32_t colb = (uint32_t)col * 8u;
    uint32_t ohb[NH];
#pragma unroll
    for (int a = 0; a < NH; a++) ohb[a] = (uint32_t)ohalo[a] * 8u;
    auto ld = [](const double* plane, uint32_t off) {
        const __amdgpu_buffer_rsrc_t r =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(plane), (short)0, -1, 0x00020000);
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    };
    auto st = [](double* plane, uint32_t off, double val) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(plane, (short)0, -1, 0x00020000);
        typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, val), r, (int)off, 0, 0);
    };
    // does any lane of this wave own an x/y edge cell (whose ghost images it writes)?
    const bool wave_edge =
        __builtin_amdgcn_ballot_w64(active && (i < 3 || i >= z.nx - 3 || j < 3 || j >= z.ny - 3)) != 0;
    auto fetch = [&](int kk, Next& n) {  // plane kk's point data and halo values, and G(kk+3) of the column
        const int64_t pl = (int64_t)(kk + 3) * sz;
        n.lead = ld(Gs + pl + 3 * sz, colb);
        n.u = ld(u + pl, colb);
        n.v = ld(v + pl, colb);
        n.w = ld(w + pl, colb);
        if constexpr (!SAME) n.gn = ld(Gn + pl, colb);
        if constexpr (SLF) n.sl = ld(sl + pl, colb);
#pragma unroll
        for (int a = 0; a < NH; a++) n.halo[a] = ld(Gs + pl, ohb[a]);
    };
    auto put_plane = [&](double own, const Next& n) {
        sG[ly + 3][lx + 3] = own;
#pragma unroll
        for (int a = 0; a < NH; a++)
            if (tid + a * kFaceThreads < kHaloItems) {
                sG[hrow[a]][hcol[a]] = n.halo[a];
                if (EXACT) bad |= !lsx::crdiv_safe(n.halo[a]);
            }
    };

    // Column window along z: zG[q] = G(k-3+q), zt[q] = T(k-2+q).
    double zG[7], zt[5];
#pragma unroll
    for (int q = 0; q < 6; q++) zG[q] = Gs[(int64_t)(kb + q) * sz + col];
    zG[6] = 0.0;
    if (EXACT)
#pragma unroll
        for (int q = 0; q < 6; q++) bad |= !lsx::crdiv_safe(zG[q]);
#pragma unroll
    for (int q = 0; q < 4; q++) zt[q] = tsq(zG[q], zG[q + 1], zG[q + 2]);
    zt[4] = 0.0;
    // The fast mode's constants as vector-register copies: the VALU reads them
    // there as well as from scalar registers, and the scalar registers they
    // free hold loop-invariant lane masks that would otherwise spill.  (The
    // empty asm only pins each value to a vector register.)
    double vi2x = z.inv2dx, vi2y = z.inv2dy, vi2z = z.inv2dz, vi6x = z.inv6dx, vi6y = z.inv6dy, vi6z = z.inv6dz;
    double vsl = z.s_l, veps = z.eps, vc = STAGE == 1 ? dt : STAGE == 2 ? c2 : c3;
    if constexpr (!EXACT)
        asm volatile("" : "+v"(vi2x), "+v"(vi2y), "+v"(vi2z), "+v"(vi6x), "+v"(vi6y), "+v"(vi6z), "+v"(vsl),
                     "+v"(veps), "+v"(vc));
    Next nx;
    fetch(kb, nx);
    put_plane(zG[3], nx);
    Face zlo;          // z face k-1/2, built in orientation zlo_L
    zlo.n = zlo.d = 0.0;
    int zlo_L = -1;    // none yet

    for (int k = kb; k < ke; k++) {
        const int64_t plane = (int64_t)(k + 3) * sz;
        zG[6] = nx.lead;
        zt[4] = tsq(zG[4], zG[5], zG[6]);
        const double cu = nx.u, cv = nx.v, cw = nx.w, cgn = SAME ? zG[3] : nx.gn,
                     csl = SLF ? nx.sl : (EXACT ? z.s_l : vsl);
        if (EXACT) {
            bad |= !lsx::crdiv_safe(zG[6]) || (SLF && !lsx::crdiv_safe(csl));
            if (bad) sBad = 1;  // sticky; read after the barrier
        }
        if (k + 1 < ke) fetch(k + 1, nx);
        __syncthreads();  // plane k in sG
        // every value of plane k's stencils was checked before this barrier
        const bool fd = EXACT && z.crdiv && __builtin_amdgcn_readfirstlane(sBad) == 0;
        // --- exact mode with the FMA sequences: each sub-stencil's (t, 1/t)
        // once per centre, for the three x (y) faces of an orientation that use it
        auto subx = [&](int row, int q, bool l, bool r) {  // centre sG[row + 3][q + 1], local x q - 2
            if constexpr (EXACT) {
                const double gm = sG[row + 3][q], g0 = sG[row + 3][q + 1], gp = sG[row + 3][q + 2];
                if (l) sSx[0][row][q] = lsx::sub_l(gm, g0, gp);
                if (r) sSx[1][row][q] = lsx::sub_r(gm, g0, gp);
            }
        };
        if constexpr (EXACT) {
            if (fd) {
                auto suby = [&](int r, int c) {  // centre sG[r + 1][c + 3], local y r - 2
                    const double gm = sG[r][c + 3], g0 = sG[r + 1][c + 3], gp = sG[r + 2][c + 3];
                    sSy[0][r][c] = lsx::sub_l(gm, g0, gp);
                    sSy[1][r][c] = lsx::sub_r(gm, g0, gp);
                };
                suby(ly + 2, lx);
                if (ly < 4) suby(ly < 2 ? ly : ly + FY, lx);  // local rows -2, -1, FY, FY+1
                if (ly == kWaveTx && lx < 4 * FY) {           // local x -2, -1, FX, FX+1 of each row
                    const int c4 = lx & 3;
                    subx(lx >> 2, c4 < 2 ? c4 : c4 + FX, true, true);
                }
            }
        }
        // --- smoothness terms along x and y, effective velocity, orientation
        if (!EXACT) {
        sTx[ly][lx + 2] = tsq(sG[ly + 3][lx + 2], sG[ly + 3][lx + 3], sG[ly + 3][lx + 4]);
        sTy[ly + 2][lx] = tsq(sG[ly + 2][lx + 3], sG[ly + 3][lx + 3], sG[ly + 4][lx + 3]);
        if (ly < 4) {
            const int ry = ly < 2 ? ly : ly + FY;  // local rows -2, -1, FY, FY+1
            sTy[ry][lx] = tsq(sG[ry][lx + 3], sG[ry + 1][lx + 3], sG[ry + 2][lx + 3]);
        }
        if (ly == kWaveTx && lx < 4 * FY) {  // local x -2, -1, FX, FX+1 of each row
            const int row = lx >> 2, c4 = lx & 3;
            const int q = c4 < 2 ? c4 : c4 + FX;
            sTx[row][q] = tsq(sG[row + 3][q], sG[row + 3][q + 1], sG[row + 3][q + 2]);
        }
        }
        double ue = cu, ve = cv, we = cw;
        if (EXACT) {  // computeRHS3D (rk3.h:31-81), C++ evaluation order
            auto ueff = [&](auto dtag) {
                constexpr int D = decltype(dtag)::value;
                double gx = 0.0, gy = 0.0, gz = 0.0, gm = 1.0;
                if (csl > z.eps) {  // 2h is exact, and RN(1/(2h)) = RN(1/h) / 2
                    gx = lsx::dq<D>(sG[ly + 3][lx + 4] - sG[ly + 3][lx + 2], 2.0 * z.dx, 0.5 * z.rdx);
                    gy = lsx::dq<D>(sG[ly + 4][lx + 3] - sG[ly + 2][lx + 3], 2.0 * z.dy, 0.5 * z.rdy);
                    gz = lsx::dq<D>(zG[4] - zG[2], 2.0 * z.dz, 0.5 * z.rdz);
                    gm = sqrt(gx * gx + gy * gy + gz * gz + z.eps);
                }
                const double ygm = D == lsx::kDivIEEE ? 0.0 : lsx::dr<D>(1.0, gm);
                ue = cu - lsx::dq<D>(csl * gx, gm, ygm);
                ve = cv - lsx::dq<D>(csl * gy, gm, ygm);
                we = cw - lsx::dq<D>(csl * gz, gm, ygm);
            };
            if (fd) ueff(DivFma());
            else ueff(DivIEEE());
        } else if (csl > veps) {
            const double gx = (sG[ly + 3][lx + 4] - sG[ly + 3][lx + 2]) * vi2x;
            const double gy = (sG[ly + 4][lx + 3] - sG[ly + 2][lx + 3]) * vi2y;
            const double gz = (zG[4] - zG[2]) * vi2z;
            const double s2 = fma(gz, gz, fma(gx, gx, gy * gy));
            const double kk = csl * rsqrt_nr(s2 + veps);
            ue = fma(-kk, gx, ue);
            ve = fma(-kk, gy, ve);
            we = fma(-kk, gz, we);
        }
        const bool Lx = ue >= 0.0, Ly = ve >= 0.0, Lz = we >= 0.0;
        sOri[ly][lx] = (unsigned char)((Lx ? 1 : 0) | (Ly ? 2 : 0));
        // the x sub-stencils of this row serve only the row's x faces, whose
        // orientations are its cells' (the tile-edge ones are computed above
        // in both), so a wavefront computes the sets its row needs
        if (EXACT && fd)
            subx(ly, lx + 2, __builtin_amdgcn_ballot_w64(Lx) != 0, __builtin_amdgcn_ballot_w64(!Lx) != 0);
        __syncthreads();  // T, orientations
        // --- faces
        // face c+1/2 of an axis into slot[0] (left-biased) and slot[OS] (mirrored)
        // t: the smoothness data at the face's low cell c (fast: T; exact: (t, 1/t)
        // per orientation), TS its stride along the axis
        auto put = [&](auto gs, auto ts, auto os, const double* g, auto t, bool nL, bool nR, FV* slot) {
            constexpr int GS = decltype(gs)::value, TS = decltype(ts)::value, OS = decltype(os)::value;
            if constexpr (EXACT) {
                if (fd) {
                    // weno_l's beta_0..2 sit at centres c-1, c, c+1; weno_r's at c+2, c+1, c
                    constexpr int SO = GS == 1 ? FY * (FX + 4) : (FY + 4) * FX;  // to the mirrored set
                    if (nL)
                        slot[0] = lsx::weno_t<false>(g[-2 * GS], g[-GS], g[0], g[GS], g[2 * GS], t[-TS], t[0], t[TS]);
                    if (nR)
                        slot[OS] = lsx::weno_t<true>(g[-GS], g[0], g[GS], g[2 * GS], g[3 * GS], t[SO + 2 * TS],
                                                     t[SO + TS], t[SO]);
                } else {
                    (void)t;
                    if (nL) slot[0] = xface_lds<GS, lsx::kDivIEEE>(g, false);
                    if (nR) slot[OS] = xface_lds<GS, lsx::kDivIEEE>(g, true);
                }
            } else {
                face_put<GS, TS, OS>(g, t, nL, nR, slot);
            }
        };
        auto tbx = [&](int row, int q) {  // smoothness data along x at local x q - 2 of a row
            if constexpr (EXACT) return &sSx[0][row][q];
            else return &sTx[row][q];
        };
        auto tby = [&](int r, int c) {  // smoothness data along y at local y r - 2 of a column
            if constexpr (EXACT) return &sSy[0][r][c];
            else return &sTy[r][c];
        };
        using C1 = std::integral_constant<int, 1>;
        using CGW = std::integral_constant<int, FGW>;
        using CFX = std::integral_constant<int, FX>;
        using COX = std::integral_constant<int, FY * (FX + 1)>;
        using COY = std::integral_constant<int, (FY + 1) * FX>;
        Face zhi;
        zhi.n = zhi.d = 0.0;
        if (active) {
            const double* g = &sG[ly + 3][lx + 3];
            {  // x face i+1/2, for this cell and its high neighbour
                bool nL = Lx, nR = !Lx;
                if (lx + 1 < FX && (FULL || i + 1 < z.nx)) {
                    const bool o = sOri[ly][lx + 1] & 1;
                    nL = nL || o;
                    nR = nR || !o;
                }
                put(C1(), C1(), COX(), g, tbx(ly, lx + 2), nL, nR, &sFx[0][ly][lx + 1]);
            }
            {  // y face j+1/2
                bool nL = Ly, nR = !Ly;
                if (ly + 1 < FY && (FULL || j + 1 < z.ny)) {
                    const bool o = (sOri[ly + 1][lx] >> 1) & 1;
                    nL = nL || o;
                    nR = nR || !o;
                }
                put(CGW(), CFX(), COY(), g, tby(ly + 2, lx), nL, nR, &sFy[0][ly + 1][lx]);
            }
            if constexpr (EXACT) {
                // z faces as single values in zlo.n
                // weno_l(zG[q..q+4]) or weno_r(zG[q+1..q+5]), q = 0, 1, with the
                // orientation a per-lane choice of the stencil order (lsx::weno_u):
                // no branch when Lz differs within the wavefront.  (The x and y
                // faces keep their branches: selecting their LDS stencils too
                // raised the register count into spills.)
                auto zfaces = [&](auto dtag) {
                    constexpr int D = decltype(dtag)::value;
                    const bool m = !Lz;
                    if (zlo_L != (int)Lz)
                        zlo.n = lsx::weno_u<D>(m ? zG[5] : zG[0], m ? zG[4] : zG[1], m ? zG[3] : zG[2],
                                               m ? zG[2] : zG[3], m ? zG[1] : zG[4], m);
                    zhi.n = lsx::weno_u<D>(m ? zG[6] : zG[1], m ? zG[5] : zG[2], m ? zG[4] : zG[3],
                                           m ? zG[3] : zG[4], m ? zG[2] : zG[5], m);
                };
                if (fd) zfaces(DivFma());
                else zfaces(DivIEEE());
            } else {
                // z face k-1/2 in this cell's orientation (carried unless it differs)
                if (zlo_L != (int)Lz) {
                    if (Lz) zlo = face_core<false>(zG[0], zG[1], zG[2], zG[3], zG[4], zt[0], zt[1], zt[2]);
                    else zlo = face_core<true>(zG[5], zG[4], zG[3], zG[2], zG[1], zt[3], zt[2], zt[1]);
                }
                // z face k+1/2
                if (Lz) zhi = face_core<false>(zG[1], zG[2], zG[3], zG[4], zG[5], zt[1], zt[2], zt[3]);
                else zhi = face_core<true>(zG[6], zG[5], zG[4], zG[3], zG[2], zt[4], zt[3], zt[2]);
            }
        }
        if (ly == kWaveFx && lx < FY) {  // x face on the low edge of row lx, in its tile cell's orientation
            if (FULL || y0 + lx < z.ny) {
                const bool o = sOri[lx][0] & 1;
                put(C1(), C1(), COX(), &sG[lx + 3][2], tbx(lx, 1), o, !o, &sFx[0][lx][0]);
            }
        } else if (ly == kWaveFy) {  // y face on the low edge of column lx
            if (FULL || x0 + lx < z.nx) {
                const bool o = (sOri[0][lx] >> 1) & 1;
                put(CGW(), CFX(), COY(), &sG[2][lx + 3], tby(1, lx), o, !o, &sFy[0][0][lx]);
            }
        }
        __syncthreads();  // faces
        // Plane k+1 goes into sG (free after the face barrier) before this
        // plane's stores are issued: vmcnt counts stores too, so the wait for
        // the prefetched halo would otherwise also wait for them.
        if (k + 1 < ke) put_plane(zG[4], nx);
        // --- derivatives, RHS, stage combine
        double o = 0.0;
        if (active && EXACT) {
            // dweno_*_3d: (face(c+1/2) - face(c-1/2)) / h; rhs and stage in the
            // reference's order (rk3.h:31-125)
            const int ox = Lx ? 0 : 1, oy = Ly ? 0 : 1;
            const double ax = val(sFx[ox][ly][lx + 1]) - val(sFx[ox][ly][lx]);
            const double ay = val(sFy[oy][ly + 1][lx]) - val(sFy[oy][ly][lx]);
            const double az = zhi.n - zlo.n;
            double dgx, dgy, dgz;
            if (fd) {
                dgx = lsx::qdiv_u(ax, z.dx, z.rdx);
                dgy = lsx::qdiv_u(ay, z.dy, z.rdy);
                dgz = lsx::qdiv_u(az, z.dz, z.rdz);
            } else {
                dgx = ax / z.dx;
                dgy = ay / z.dy;
                dgz = az / z.dz;
            }
            const double rr = -(ue * dgx + ve * dgy + we * dgz);
            o = lsx::stage_combine<STAGE>(cgn, zG[3], rr, dt);
        } else if (active) {
            const int ox = Lx ? 0 : 1, oy = Ly ? 0 : 1;
            const double2 xl = as2(sFx[ox][ly][lx]), xh = as2(sFx[ox][ly][lx + 1]);
            const double2 yl = as2(sFy[oy][ly][lx]), yh = as2(sFy[oy][ly + 1][lx]);
            const double Nx = fma(xh.x, xl.y, -(xl.x * xh.y)), Dx = xh.y * xl.y;
            const double Ny = fma(yh.x, yl.y, -(yl.x * yh.y)), Dy = yh.y * yl.y;
            const double Nz = fma(zhi.n, zlo.d, -(zlo.n * zhi.d)), Dz = zhi.d * zlo.d;
            const double Dxy = Dx * Dy;
            const double num = fma((ue * vi6x) * Nx, Dy * Dz,
                                   fma((ve * vi6y) * Ny, Dx * Dz, ((we * vi6z) * Nz) * Dxy));
            const double rr = -(num * rcp_nr(Dxy * Dz));
            const double gc = zG[3];
            if (STAGE == 1) o = fma(vc, rr, cgn);
            else if (STAGE == 2) o = fma(vc, rr, fma(0.25, gc, 0.75 * cgn));
            else o = fma(vc, rr, fma(2.0 / 3.0, gc, (1.0 / 3.0) * cgn));
        }
        if (active) st(out + plane, colb, o);
        // ghost images: a scalar test skips them for the waves and planes
        // that own none
        if (z.fuse_ghosts && (wave_edge || k < 3 || k >= z.nz - 3)) {
            if (active && (i < 3 || i >= z.nx - 3 || j < 3 || j >= z.ny - 3 || k < 3 || k >= z.nz - 3))
                store_images(z, out, i, j, k, o);
        }
        zlo = zhi;
        zlo_L = (int)Lz;
#pragma unroll
        for (int q = 0; q < 6; q++) zG[q] = zG[q + 1];
#pragma unroll
        for (int q = 0; q < 4; q++) zt[q] = zt[q + 1];
    }
}

// ---------------------------------------------------------------------------
// HCR-2 pseudo-time iteration, z-marching (3D).  Same structure as the stage
// kernel with a 1-cell halo: the current planes of phi_n and phi_0 go through
// LDS for the x/y neighbours, the z neighbours stay in registers, the next
// plane is prefetched, and the ghost cells of phi_new are written by the
// interior cells they copy.
constexpr int HX = 64, HY = 4;

struct HGeom {
    ZGeom z;
    double dx2, invdx, invdy, invdz, dtau, beta;
};

// One HCR-2 pseudo-time update of a cell (reinitialization.h:136-193, 3D),
// hcr2_point, from phi_n (P) and phi_0 (Q) at the cell and its neighbours in
// flag order x-, x+, y-, y+, z-, z+ (reinitialization.h:46-89; entries 5 and 6
// unused in 2D).  Crossing flags and r~ come from phi_0 with the arithmetic of
// computeInterfaceCrossings.  The update splits into hcr2_q, the part that
// depends on phi_0 alone (crossing flags, the smoothed sign S, r~), which both
// iterations of an invocation share, and hcr2_p, the rest.  H is HGeom or
// HPar (the members dx2, invdx, invdy, invdz, dtau, beta).
struct HQ {
    double S, rt;
    int flag;
};
template <int DIM = 3, class H>
__device__ __forceinline__ HQ hcr2_q(const H& hg, const double (&Q)[7]) {
    constexpr int NB = DIM == 3 ? 6 : 4;
    const double pc = Q[0];
    int flag = 0;
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (pc * Q[q + 1] < 0.0) flag |= 1 << q;
    HQ h;
    h.S = pc * rsqrt_nr(fma(pc, pc, hg.dx2));
    h.flag = flag;
    h.rt = 0.0;
    // r~ = phi_0 / sum of crossing neighbours, 0 when that sum vanishes.  Only
    // cells beside the interface have a flag, so most wavefronts skip this.
    if (flag != 0) {
        double sum0 = 0.0;
#pragma unroll
        for (int q = 0; q < NB; q++)
            if ((flag >> q) & 1) sum0 += Q[q + 1];
        if (fabs(sum0) > 1e-15) h.rt = div_fast(pc, sum0);
    }
    return h;
}

// The rest of the update from phi_n (P), phi_0 (Q) and their shared part.
template <int DIM = 3, class H>
__device__ __forceinline__ double hcr2_p(const H& hg, const double (&P)[7], const double (&Q)[7], const HQ& h) {
    constexpr int NB = DIM == 3 ? 6 : 4;
    const double pc = Q[0];
    const double phi = P[0];
    const double S = h.S;
    // Godunov |grad phi| (weno5.h:202-248): for S > 0 an axis contributes
    // max(max(D-, 0)^2, min(D+, 0)^2), for S <= 0 the mirrored min/max.  With
    // s = +-1 the sign of S, both read max(A, B)^2 over the non-negative
    // A = max(s (phi_c - phi_lo) / h, 0), B = max(s (phi_c - phi_hi) / h, 0),
    // the same value bit for bit (A^2 >= B^2 exactly when A >= B).  Since
    // negation is exact and rounding is monotone, max(A, B) is
    // max(RN(RN(s phi_c - min(s phi_lo, s phi_hi)) / h), 0) with the products
    // by 1/h: one subtraction and one product per axis.
    const int sgn = S > 0.0 ? 0 : (int)0x80000000u;  // the sign bit of -s
    double R[7];
#pragma unroll
    for (int q = 0; q < 7; q++) R[q] = __hiloint2double(__double2hiint(P[q]) ^ sgn, __double2loint(P[q]));
    const double gx = fmax((R[0] - fmin(R[1], R[2])) * hg.invdx, 0.0);
    const double gy = fmax((R[0] - fmin(R[3], R[4])) * hg.invdy, 0.0);
    double s2 = gx * gx + gy * gy;
    if (DIM == 3) {
        const double gz = fmax((R[0] - fmin(R[5], R[6])) * hg.invdz, 0.0);
        s2 = s2 + gz * gz;
    }
    // sqrt(s2) - 1 through sqrt_u on max(s2, 2^-767): below that bound both
    // square roots are under 2^-54, so the difference rounds to -1 either way
    const double term = -S * (lsx::sqrt_u(fmax(s2, 0x1p-767)) - 1.0);
    double forcing = 0.0;
    if (h.flag != 0) {
        bool stable = (P[0] > 0.0) == (pc > 0.0);
#pragma unroll
        for (int q = 0; q < NB; q++)
            if ((h.flag >> q) & 1) stable = stable && ((P[q + 1] > 0.0) == (Q[q + 1] > 0.0));
        if (stable) {
            double sum = 0.0;
#pragma unroll
            for (int q = 0; q < NB; q++)
                if ((h.flag >> q) & 1) sum += P[q + 1];
            forcing = hg.beta * fma(h.rt, sum, -phi) * hg.invdx;
        }
    }
    return fma(hg.dtau, term + forcing, phi);
}

template <int DIM = 3, class H>
__device__ __forceinline__ double hcr2_point(const H& hg, const double (&P)[7], const double (&Q)[7]) {
    return hcr2_p<DIM>(hg, P, Q, hcr2_q<DIM>(hg, Q));
}

// The parameters a point update reads (the members of HGeom it uses).
struct HPar {
    double dx2, invdx, invdy, invdz, dtau, beta;
};

// 2D HCR-2 pseudo-time step, one thread per cell, with the 3D kernels' point
// update (x-, x+, y-, y+ neighbours; entries 5 and 6 unused).
__global__ __launch_bounds__(256) void k_hcr2_fast2(FastGeom f, HPar hp, const double* __restrict__ pn,
                                                    const double* __restrict__ p0, double* __restrict__ pnew) {
    const int64_t t = xcd_tile(blockIdx.x, f.nblocks);
    if (t >= (int64_t)f.tiles_x * f.tiles_y) return;
    const int i = (int)(t % f.tiles_x) * kFX + (int)threadIdx.x;
    const int j = (int)(t / f.tiles_x) * kFY + (int)threadIdx.y;
    if (i >= f.g.nx || j >= f.g.ny) return;
    const Geom& g = f.g;
    const int64_t c = (int64_t)(j + g.ng) * g.sy + (i + g.ng);
    const int64_t off[4] = {-1, 1, -g.sy, g.sy};
    double P[7], Q[7];
    P[0] = pn[c];
    Q[0] = p0[c];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        P[q + 1] = pn[c + off[q]];
        Q[q + 1] = p0[c + off[q]];
    }
    P[5] = P[6] = Q[5] = Q[6] = 0.0;
    pnew[c] = hcr2_point<2>(hp, P, Q);
}

template <int W, int H>
__device__ __forceinline__ void h_load_tile(const HGeom& hg, int64_t plane, int x0, int y0, int tid,
                                            const double* __restrict__ F, double (&s)[H + 2][W + 2],
                                            double own, bool active, int lx, int ly) {
    const ZGeom& z = hg.z;
    s[ly + 1][lx + 1] = own;
    // halo ring: rows -1 and H (W cells each), columns -1 and W (H cells each)
    for (int h = tid; h < 2 * W + 2 * H; h += W * H) {
        int gx, gy, sx, syy;
        if (h < 2 * W) {
            sx = h % W + 1;
            syy = h < W ? 0 : H + 1;
        } else {
            const int q = h - 2 * W;
            syy = q % H + 1;
            sx = q < H ? 0 : W + 1;
        }
        gx = x0 - 1 + sx;
        gy = y0 - 1 + syy;
        s[syy][sx] = (gx < z.nx + 3 && gy < z.ny + 3 && gx >= -3 && gy >= -3)
                         ? F[plane + (int64_t)(gy + 3) * z.sy + (gx + 3)] : 0.0;
    }
    (void)active;
}

__global__ __launch_bounds__(256) void k_hcr2_zmarch(HGeom hg, const double* __restrict__ pn,
                                                     const double* __restrict__ p0, double* __restrict__ pnew) {
    __shared__ double sP[HY + 2][HX + 2];
    __shared__ double sQ[HY + 2][HX + 2];
    const ZGeom& z = hg.z;
    const int64_t t = xcd_tile(blockIdx.x, z.nblocks);
    if (t >= z.ntiles) return;  // block-uniform
    const int tx = (int)(t % z.tiles_x);
    const int64_t r = t / z.tiles_x;
    const int ty = (int)(r % z.tiles_y);
    const int ch = (int)(r / z.tiles_y);
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * HX + lx;
    const int x0 = tx * HX, y0 = ty * HY;
    const int i = x0 + lx, j = y0 + ly;
    const bool active = i < z.nx && j < z.ny;
    const bool inpad = i < z.nx + 3 && j < z.ny + 3;
    const int kb = ch * z.kz;
    const int ke = min(kb + z.kz, z.nz);
    const int64_t sy = z.sy, sz = z.sz;
    const int64_t col = (int64_t)(j + 3) * sy + (i + 3);
    // column windows at k-1, k, k+1 (padded k index = k + 3)
    double a0 = inpad ? pn[(int64_t)(kb + 2) * sz + col] : 0.0;
    double a1 = inpad ? pn[(int64_t)(kb + 3) * sz + col] : 0.0;
    double b0 = inpad ? p0[(int64_t)(kb + 2) * sz + col] : 0.0;
    double b1 = inpad ? p0[(int64_t)(kb + 3) * sz + col] : 0.0;
    double na = inpad ? pn[(int64_t)(kb + 4) * sz + col] : 0.0;
    double nb = inpad ? p0[(int64_t)(kb + 4) * sz + col] : 0.0;
    for (int k = kb; k < ke; k++) {
        const int64_t plane = (int64_t)(k + 3) * sz;
        const double a2 = na, b2 = nb;
        if (k + 1 < ke && inpad) {
            na = pn[plane + 2 * sz + col];
            nb = p0[plane + 2 * sz + col];
        }
        __syncthreads();
        h_load_tile<HX, HY>(hg, plane, x0, y0, tid, pn, sP, a1, active, lx, ly);
        h_load_tile<HX, HY>(hg, plane, x0, y0, tid, p0, sQ, b1, active, lx, ly);
        __syncthreads();
        if (active) {
            const double P[7] = {a1, sP[ly + 1][lx], sP[ly + 1][lx + 2], sP[ly][lx + 1], sP[ly + 2][lx + 1], a0, a2};
            const double Q[7] = {b1, sQ[ly + 1][lx], sQ[ly + 1][lx + 2], sQ[ly][lx + 1], sQ[ly + 2][lx + 1], b0, b2};
            const double o = hcr2_point(hg, P, Q);
            pnew[plane + col] = o;
            if (z.fuse_ghosts && (i < 3 || i >= z.nx - 3 || j < 3 || j >= z.ny - 3 || k < 3 || k >= z.nz - 3))
                store_images(z, pnew, i, j, k, o);
        }
        a0 = a1;
        a1 = a2;
        b0 = b1;
        b1 = b2;
    }
}

// ---------------------------------------------------------------------------
// The first two HCR-2 iterations in one z-marching pass.  The first iteration
// has phi_n = phi_0, so phi_1 = step(phi_0, phi_0) needs phi_0 only, and
// phi_2 = step(phi_1, phi_0) needs phi_1 at the cell and its six neighbours.
// Plane k computes phi_1 on plane k+1 over the tile plus a one-cell ring
// (into LDS), then phi_2 on plane k.  The phi_0 planes k-2..k+2 (tile plus a
// two-cell ring) sit in a five-plane LDS ring, so a plane needs two barriers.
// phi_1 at a ghost position takes the value the ghost fill between the two
// iterations would give it (reinitialization.h:349-355):
//  * a periodic axis computes it in place, because phi_0 is periodic over the
//    padded array, so the stencil there equals the stencil of the image;
//  * a zero-gradient axis takes it from the clamped cell.
// Both iterations use the single iteration's point update (hcr2_point in the
// fast mode, the reference's arithmetic lsx::hcr2_core in the exact mode), so
// the result is bitwise that of two single iterations with a ghost fill
// between them.  Needs the ghost images of the output (no
// linear-extrapolation axis).
// Tiles are 64x8 (8 waves): the first iteration then covers 66x10 cells per
// 512 outputs, 1.29 per cell, against 66x6 per 256 (1.55) on 64x4 tiles
// (1.53 against 1.66 ms per pair at 512^3; 64x16 tiles were slower).
// The fast mode keeps, per thread, the phi_0-only part of its own cell's
// update (HQ: the sign flags, S, the constraint divisor) from the first
// iteration to the second, one plane later, and is held to 76 VGPRs so six
// waves fit per SIMD (three blocks per CU, as the LDS allows): 1.88 against
// 1.93 ms for ls_reinit_hcr2 at 512^3 (at 88 VGPRs and four waves, 1.97 ms).
#ifndef LS_PAIR_ROWS
#define LS_PAIR_ROWS 8
#endif
constexpr int PY = LS_PAIR_ROWS;                               // tile rows of the pair kernel
constexpr int PQX = HX + 4, PQY = PY + 4;                      // phi_0 plane: tile + 2-cell ring
constexpr int P1X = HX + 2, P1Y = PY + 2;                      // phi_1 plane: tile + 1-cell ring
constexpr int kQItems = PQX * PQY, kQPer = (kQItems + HX * PY - 1) / (HX * PY);

template <bool EXACT>
__global__ __launch_bounds__(HX * PY) __attribute__((amdgpu_waves_per_eu(EXACT ? 1 : 6))) void k_hcr2_pair0(HGeom hg, const double* __restrict__ p0,
                                                       double* __restrict__ pnew) {
    __shared__ double sQ[5][PQY][PQX];   // phi_0 plane kk in slot (kk + 10) % 5 (kk >= -2)
    __shared__ double sP1[2][P1Y][P1X];  // phi_1 plane kk in slot kk & 1
    const ZGeom& z = hg.z;
    const int64_t t = xcd_tile(blockIdx.x, z.nblocks);
    if (t >= z.ntiles) return;  // block-uniform
    const int tx = (int)(t % z.tiles_x);
    const int64_t r = t / z.tiles_x;
    const int ty = (int)(r % z.tiles_y);
    const int ch = (int)(r / z.tiles_y);
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * HX + lx;
    const int x0 = tx * HX, y0 = ty * PY;
    const int i = x0 + lx, j = y0 + ly;
    const bool active = i < z.nx && j < z.ny;
    const int kb = ch * z.kz;
    const int ke = min(kb + z.kz, z.nz);
    const int64_t sz = z.sz;
    const bool zgx = z.bc[0] == LS_BC_ZERO_GRAD, zgy = z.bc[1] == LS_BC_ZERO_GRAD, zgz = z.bc[2] == LS_BC_ZERO_GRAD;
    auto point = [&](const double (&P)[7], const double (&Q)[7]) {
        if constexpr (EXACT) return lsx::hcr2_core<3>(P, Q, z.dx, z.dy, z.dz, hg.dtau, hg.beta);
        else return hcr2_point(hg, P, Q);
    };

    // phi_0 items: local (x, y) in [-2, HX+1] x [-2, PY+1]; positions past the
    // padded array read the nearest padded cell (they feed no interior cell)
    int qx[kQPer], qy[kQPer];
    int32_t qoff[kQPer];
#pragma unroll
    for (int a = 0; a < kQPer; a++) {
        const int e = min(tid + a * HX * PY, kQItems - 1);
        qy[a] = e / PQX;
        qx[a] = e - qy[a] * PQX;
        const int gx = min(x0 - 2 + qx[a], z.nx + 2), gy = min(y0 - 2 + qy[a], z.ny + 2);
        qoff[a] = (gy + 3) * z.sy + gx + 3;
    }
    auto slot = [](int kk) { return (kk + 10) % 5; };
    auto load_q = [&](int kk, double (&val)[kQPer]) {
        const int64_t pl = (int64_t)(kk + 3) * sz;
#pragma unroll
        for (int a = 0; a < kQPer; a++) val[a] = p0[pl + qoff[a]];
    };
    auto put_q = [&](int kk, const double (&val)[kQPer]) {
        const int sl = slot(kk);
#pragma unroll
        for (int a = 0; a < kQPer; a++)
            if (tid + a * HX * PY < kQItems) sQ[sl][qy[a]][qx[a]] = val[a];
    };
    // phi_1 items: pass 0 the thread's own cell (local x = lx, y = ly), pass 1
    // the one-cell ring around the tile (kRingItems items, threads below that
    // count); each evaluated at (ex, ey) (sQ indices; the clamped cell on a
    // zero-gradient axis).  In the fast mode the part of the update that
    // depends on phi_0 alone (HQ) is kept from pass 0 for the thread's own
    // cell and reused by the second iteration there one plane later.
    constexpr int kRingItems = 2 * P1X + 2 * PY;
    static_assert(kRingItems <= HX * PY, "ring items per thread");
    int px[2], py[2], ex[2], ey[2];
    px[0] = lx + 1;
    py[0] = ly + 1;
    {
        const int r = min(tid, kRingItems - 1);
        if (r < P1X) {
            px[1] = r;
            py[1] = 0;
        } else if (r < 2 * P1X) {
            px[1] = r - P1X;
            py[1] = PY + 1;
        } else if (r < 2 * P1X + PY) {
            px[1] = 0;
            py[1] = r - 2 * P1X + 1;
        } else {
            px[1] = HX + 1;
            py[1] = r - 2 * P1X - PY + 1;
        }
    }
    const bool ring = tid < kRingItems;
#pragma unroll
    for (int a = 0; a < 2; a++) {
        int gx = x0 - 1 + px[a], gy = y0 - 1 + py[a];
        if (zgx) gx = min(max(gx, 0), z.nx - 1);
        if (zgy) gy = min(max(gy, 0), z.ny - 1);
        ex[a] = gx - x0 + 2;
        ey[a] = gy - y0 + 2;
    }
    auto zev = [&](int kk) { return zgz ? min(max(kk, 0), z.nz - 1) : kk; };
    // phi_0 at sQ position (X, Y) of plane kk and its six neighbours
    auto qget = [&](int kk, int X, int Y, double (&Q)[7]) {
        const int sc = slot(kk), sm = slot(kk - 1), sp = slot(kk + 1);
        Q[0] = sQ[sc][Y][X];
        Q[1] = sQ[sc][Y][X - 1];
        Q[2] = sQ[sc][Y][X + 1];
        Q[3] = sQ[sc][Y - 1][X];
        Q[4] = sQ[sc][Y + 1][X];
        Q[5] = sQ[sm][Y][X];
        Q[6] = sQ[sp][Y][X];
    };
    // phi_1 on plane kk at sQ position (X, Y): the first iteration, phi_n = phi_0
    auto phi1 = [&](int kk, int X, int Y) {
        double Q[7];
        qget(kk, X, Y, Q);
        return point(Q, Q);
    };
    // phi_1 of the thread's own cell on plane kk, keeping the phi_0 part
    HQ hq_cur, hq_next;
    hq_cur.S = hq_cur.rt = hq_next.S = hq_next.rt = 0.0;
    hq_cur.flag = hq_next.flag = 0;
    auto phi1_own = [&](int kk, HQ& hq) {
        double Q[7];
        qget(kk, ex[0], ey[0], Q);
        if constexpr (EXACT) {
            (void)hq;
            return point(Q, Q);
        } else {
            hq = hcr2_q(hg, Q);
            return hcr2_p(hg, Q, Q, hq);
        }
    };

    double val[kQPer];
    for (int kk = kb - 2; kk <= kb + 2; kk++) {
        load_q(kk, val);
        put_q(kk, val);
    }
    __syncthreads();
    sP1[kb & 1][py[0]][px[0]] = phi1_own(zev(kb), hq_cur);
    if (ring) sP1[kb & 1][py[1]][px[1]] = phi1(zev(kb), ex[1], ey[1]);
    double fm = phi1(zev(kb - 1), lx + 2, ly + 2);  // phi_1 of this column one plane below

    for (int k = kb; k < ke; k++) {
        if (k + 1 < ke) load_q(k + 3, val);
        __syncthreads();  // phi_0 plane k+2 in place; phi_1 plane k-1 no longer read
        sP1[(k + 1) & 1][py[0]][px[0]] = phi1_own(zev(k + 1), hq_next);
        if (ring) sP1[(k + 1) & 1][py[1]][px[1]] = phi1(zev(k + 1), ex[1], ey[1]);
        __syncthreads();  // phi_1 plane k+1 in place
        // phi_0 plane k+3 into the slot of plane k-2, last read in the previous
        // plane, before this plane's stores: vmcnt counts stores too, so a wait
        // for the loads placed after them would also wait for the stores.
        if (k + 1 < ke) put_q(k + 3, val);
        if (active) {
            const int s0 = k & 1, X = lx + 1, Y = ly + 1;
            const double fc = sP1[s0][Y][X];
            const double P[7] = {fc,
                                 sP1[s0][Y][X - 1],
                                 sP1[s0][Y][X + 1],
                                 sP1[s0][Y - 1][X],
                                 sP1[s0][Y + 1][X],
                                 fm,
                                 sP1[(k + 1) & 1][Y][X]};
            double Q[7];
            qget(k, lx + 2, ly + 2, Q);
            // an active cell is interior, so its pass-0 phi_0 position was
            // (lx + 2, ly + 2) on plane k: hq_cur is this cell's phi_0 part
            double o;
            if constexpr (EXACT) o = point(P, Q);
            else o = hcr2_p(hg, P, Q, hq_cur);
            const int64_t c = (int64_t)(k + 3) * sz + (int64_t)(j + 3) * z.sy + (i + 3);
            pnew[c] = o;
            if (i < 3 || i >= z.nx - 3 || j < 3 || j >= z.ny - 3 || k < 3 || k >= z.nz - 3)
                store_images(z, pnew, i, j, k, o);
            fm = fc;
        }
        hq_cur = hq_next;
    }
}

static FastGeom make_fast(const Geom& g) {
    FastGeom f;
    f.g = g;
    f.inv2dx = 1.0 / (2.0 * g.dx);
    f.inv2dy = 1.0 / (2.0 * g.dy);
    f.inv2dz = 1.0 / (2.0 * g.dz);
    f.inv6dx = 1.0 / (6.0 * g.dx);
    f.inv6dy = 1.0 / (6.0 * g.dy);
    f.inv6dz = 1.0 / (6.0 * g.dz);
    f.invdx = 1.0 / g.dx;
    f.invdy = 1.0 / g.dy;
    f.invdz = 1.0 / g.dz;
    f.tiles_x = (g.nx + kFX - 1) / kFX;
    f.tiles_y = (g.ny + kFY - 1) / kFY;
    const int64_t tiles = (int64_t)f.tiles_x * f.tiles_y * (g.dim == 3 ? g.nz : 1);
    f.nblocks = ((tiles + 7) / 8) * 8;
    return f;
}

static ZGeom make_z(const Geom& g, int kz, int tile_cols = ZX, int tile_rows = ZY) {
    const FastGeom f = make_fast(g);
    ZGeom z;
    z.nx = g.nx;
    z.ny = g.ny;
    z.nz = g.nz;
    z.sy = (int32_t)g.sy;
    z.sz = (int32_t)g.sz;
    for (int a = 0; a < 3; a++) z.bc[a] = g.bc[a];
    z.tiles_x = (g.nx + tile_cols - 1) / tile_cols;
    z.tiles_y = (g.ny + tile_rows - 1) / tile_rows;
    z.kz = kz;
    const int nchunks = (g.nz + kz - 1) / kz;
    z.ntiles = (int64_t)z.tiles_x * z.tiles_y * nchunks;
    z.nblocks = ((z.ntiles + 7) / 8) * 8;
    z.fuse_ghosts = g.bc[0] != LS_BC_LIN_EXTRAP && g.bc[1] != LS_BC_LIN_EXTRAP && g.bc[2] != LS_BC_LIN_EXTRAP;
    z.s_l = g.s_l;
    z.eps = g.eps;
    z.inv2dx = f.inv2dx;
    z.inv2dy = f.inv2dy;
    z.inv2dz = f.inv2dz;
    z.inv6dx = f.inv6dx;
    z.inv6dy = f.inv6dy;
    z.inv6dz = f.inv6dz;
    z.dx = g.dx;
    z.dy = g.dy;
    z.dz = g.dz;
    z.rdx = g.rdx;
    z.rdy = g.rdy;
    z.rdz = g.rdz;
    z.crdiv = lsx::crdiv_geom_ok(g.dx, g.dy, g.dz, g.eps) && lsx::crdiv_safe(g.s_l);
    return z;
}

// Launches stage `stage` of k_stage_zface with variant bits V (kVarSame is
// added for stage 1 when G_s is G_n; stages 2 and 3 always load G_n, which is
// also correct when the caller passes the same array).
template <bool E, int V>
static void zface_stage(int stage, bool same, dim3 grid, dim3 block, hipStream_t st, const ZGeom& z,
                        const double* Gs, const double* Gn, double* out, const double* u, const double* v,
                        const double* w, const double* sl, double dt) {
    if (stage == 1) {
        if (same)
            hipLaunchKernelGGL((k_stage_zface<1, E, V | kVarSame>), grid, block, 0, st, z, Gs, Gn, out, u, v, w, sl, dt);
        else
            hipLaunchKernelGGL((k_stage_zface<1, E, V>), grid, block, 0, st, z, Gs, Gn, out, u, v, w, sl, dt);
    } else if (stage == 2) {
        hipLaunchKernelGGL((k_stage_zface<2, E, V>), grid, block, 0, st, z, Gs, Gn, out, u, v, w, sl, dt);
    } else {
        hipLaunchKernelGGL((k_stage_zface<3, E, V>), grid, block, 0, st, z, Gs, Gn, out, u, v, w, sl, dt);
    }
}

// Returns 1 when the launched kernel also filled the ghost cells of out.
int launch_stage_zmarch(const Geom& g, int stage, const double* Gs, const double* Gn, double* out,
                        const double* u, const double* v, const double* w, const double* sl, double dt,
                        hipStream_t st, bool exact) {
    const int FY = exact ? ZTile<true>::FY : ZTile<false>::FY;
    const int kz = zmarch_chunk(g, FX, FY);
    if (g.sz * (int64_t)g.nzt >= 0x7fffffffLL)
        return lsi::set_error(LS_ERR_INVALID, "padded grid exceeds 2^31 cells");
    const ZGeom z = make_z(g, kz, FX, FY);
    if (z.nblocks > 0x7fffffffLL) return lsi::set_error(LS_ERR_INVALID, "grid too large for one launch");
    const dim3 grid((unsigned)z.nblocks), block(FX, FY, 1);
    const bool same = stage == 1 && Gs == Gn;
    const bool full = g.nx % FX == 0 && g.ny % FY == 0;
    if (exact) {
        if (full) {
            if (sl) zface_stage<true, kVarSLF | kVarFull>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
            else zface_stage<true, kVarFull>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
        } else {
            if (sl) zface_stage<true, kVarSLF>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
            else zface_stage<true, 0>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
        }
    } else {
        if (full) {
            if (sl) zface_stage<false, kVarSLF | kVarFull>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
            else zface_stage<false, kVarFull>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
        } else {
            if (sl) zface_stage<false, kVarSLF>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
            else zface_stage<false, 0>(stage, same, grid, block, st, z, Gs, Gn, out, u, v, w, sl, dt);
        }
    }
    if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "z-march stage launch failed");
    return z.fuse_ghosts ? 1 : 0;
}

int launch_stage_exact3d(const Geom& g, int stage, const double* Gs, const double* Gn, double* out,
                         const double* u, const double* v, const double* w, const double* sl, double dt,
                         hipStream_t st) {
    return launch_stage_zmarch(g, stage, Gs, Gn, out, u, v, w, sl, dt, st, true);
}

// Rows per y-march chunk.  Short chunks win on this part: 8-row chunks give
// the most wavefronts to hide load latency behind, and the 6 halo rows each
// chunk re-reads come mostly from L2, because the chunks of a strip run on one
// XCD (measured over 96 uneven 2D grids: 93 us per stage with 8 rows, 102 us
// with 4, 107 us with 16, 110 us with 32).  Returns 0 when the grid gives
// fewer than 2048 units: it then runs one thread per cell.
int ymarch_chunk(int nx, int ny) {
    const int64_t strips = (nx + kYW - 1) / kYW;
    const int ch = 8;
    return strips * ((ny + ch - 1) / ch) < 2048 ? 0 : ch;
}

static YGrid make_y(const Geom& g, double dt) {
    const FastGeom f = make_fast(g);
    YGrid m;
    m.nx = g.nx;
    m.ny = g.ny;
    m.bc0 = g.bc[0];
    m.bc1 = g.bc[1];
    m.fuse_ghosts = g.bc[0] != LS_BC_LIN_EXTRAP && g.bc[1] != LS_BC_LIN_EXTRAP;
    m.pad = 0;
    m.sy = g.sy;
    m.off = 0;
    m.s_l = g.s_l;
    m.eps = g.eps;
    m.dt = dt;
    m.inv2dx = f.inv2dx;
    m.inv2dy = f.inv2dy;
    m.inv6dx = f.inv6dx;
    m.inv6dy = f.inv6dy;
    m.dx = g.dx;
    m.dy = g.dy;
    m.rdx = g.rdx;
    m.rdy = g.rdy;
    m.crdiv = lsx::crdiv_geom_ok(g.dx, g.dy, 1.0, g.eps) && lsx::crdiv_safe(g.s_l);
    return m;
}

bool ymarch_fits(const Geom& g) { return g.dim == 2 && ymarch_chunk(g.nx, g.ny) != 0; }

// The 2D stage as a y-march (the grid must fit it).  Returns 1 when the
// kernel also filled the ghost cells of out.
int launch_stage_ymarch(const Geom& g, int stage, const double* Gs, const double* Gn, double* out, const double* u,
                        const double* v, const double* sl, double dt, hipStream_t st, bool exact) {
    const YGrid m = make_y(g, dt);
    const int ch = ymarch_chunk(g.nx, g.ny);
    const int strips = (g.nx + kYW - 1) / kYW;
    const int64_t units = (int64_t)strips * ((g.ny + ch - 1) / ch);
    const int64_t nblocks = ((units + 7) / 8) * 8;
    if (ch == 0 || nblocks > 0x7fffffffLL) return lsi::set_error(LS_ERR_INVALID, "grid does not fit the y-march");
    const dim3 grid((unsigned)nblocks), block(kYW, 1, 1);
#define LS_YMARCH_LAUNCH(E)                                                                                      \
    if (stage == 1)                                                                                              \
        hipLaunchKernelGGL((k_stage_ymarch<1, E>), grid, block, 0, st, m, strips, ch, units, nblocks, Gs, Gn, out, u, v, sl); \
    else if (stage == 2)                                                                                         \
        hipLaunchKernelGGL((k_stage_ymarch<2, E>), grid, block, 0, st, m, strips, ch, units, nblocks, Gs, Gn, out, u, v, sl); \
    else                                                                                                         \
        hipLaunchKernelGGL((k_stage_ymarch<3, E>), grid, block, 0, st, m, strips, ch, units, nblocks, Gs, Gn, out, u, v, sl);
    if (exact) {
        LS_YMARCH_LAUNCH(true)
    } else {
        LS_YMARCH_LAUNCH(false)
    }
#undef LS_YMARCH_LAUNCH
    if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "y-march stage launch failed");
    return m.fuse_ghosts ? 1 : 0;
}

template <int DIM>
int launch_stage_fast(const Geom& g, int stage, const double* Gs, const double* Gn, double* out,
                      const double* u, const double* v, const double* w, const double* sl, double dt,
                      hipStream_t st) {
    if (DIM == 3) return launch_stage_zmarch(g, stage, Gs, Gn, out, u, v, w, sl, dt, st, false);
    if (ymarch_fits(g)) return launch_stage_ymarch(g, stage, Gs, Gn, out, u, v, sl, dt, st, false);
    const FastGeom f = make_fast(g);
    const YGrid m = make_y(g, dt);
    const int64_t tiles = (int64_t)f.tiles_x * f.tiles_y;
    if (f.nblocks > 0x7fffffffLL) return lsi::set_error(LS_ERR_INVALID, "grid too large for one launch");
    const dim3 grid((unsigned)f.nblocks), block(kFX, kFY, 1);
    if (stage == 1)
        hipLaunchKernelGGL((k_stage_ycell<1>), grid, block, 0, st, m, f.tiles_x, tiles, f.nblocks, Gs, Gn, out, u, v, sl);
    else if (stage == 2)
        hipLaunchKernelGGL((k_stage_ycell<2>), grid, block, 0, st, m, f.tiles_x, tiles, f.nblocks, Gs, Gn, out, u, v, sl);
    else
        hipLaunchKernelGGL((k_stage_ycell<3>), grid, block, 0, st, m, f.tiles_x, tiles, f.nblocks, Gs, Gn, out, u, v, sl);
    if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "fast stage launch failed");
    return m.fuse_ghosts ? 1 : 0;
}

template int launch_stage_fast<2>(const Geom&, int, const double*, const double*, double*, const double*,
                                  const double*, const double*, const double*, double, hipStream_t);
template int launch_stage_fast<3>(const Geom&, int, const double*, const double*, double*, const double*,
                                  const double*, const double*, const double*, double, hipStream_t);

// The first two HCR-2 iterations from phi_0 in one launch (3D), in the fast
// or the exact arithmetic.  Returns 1 when it ran (the output's ghost cells
// included), 0 when this grid needs the single-iteration path (a
// linear-extrapolation axis).
int launch_hcr2_pair(const Geom& g, const double* p0, double* pnew, hipStream_t st, bool exact) {
    if (g.dim != 3) return 0;
    if (g.sz * (int64_t)g.nzt >= 0x7fffffffLL)
        return lsi::set_error(LS_ERR_INVALID, "padded grid exceeds 2^31 cells");
    HGeom hg;
    hg.z = make_z(g, zmarch_chunk(g, HX, PY), HX, PY);
    if (!hg.z.fuse_ghosts) return 0;
    hg.dx2 = g.dx * g.dx;
    hg.invdx = 1.0 / g.dx;
    hg.invdy = 1.0 / g.dy;
    hg.invdz = 1.0 / g.dz;
    hg.dtau = g.dtau;
    hg.beta = g.beta;
    if (hg.z.nblocks > 0x7fffffffLL) return lsi::set_error(LS_ERR_INVALID, "grid too large for one launch");
    if (exact)
        hipLaunchKernelGGL(k_hcr2_pair0<true>, dim3((unsigned)hg.z.nblocks), dim3(HX, PY, 1), 0, st, hg, p0, pnew);
    else
        hipLaunchKernelGGL(k_hcr2_pair0<false>, dim3((unsigned)hg.z.nblocks), dim3(HX, PY, 1), 0, st, hg, p0, pnew);
    if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "HCR-2 pair launch failed");
    return 1;
}

// Returns 1 when the kernel also filled the ghost cells of pnew.
int launch_hcr2_fast(const Geom& g, const double* pn, const double* p0, double* pnew, hipStream_t st) {
    if (g.dim == 3) {
        if (g.sz * (int64_t)g.nzt >= 0x7fffffffLL)
            return lsi::set_error(LS_ERR_INVALID, "padded grid exceeds 2^31 cells");
        HGeom hg;
        hg.z = make_z(g, zmarch_chunk(g, HX, HY));
        hg.dx2 = g.dx * g.dx;
        hg.invdx = 1.0 / g.dx;
        hg.invdy = 1.0 / g.dy;
        hg.invdz = 1.0 / g.dz;
        hg.dtau = g.dtau;
        hg.beta = g.beta;
        if (hg.z.nblocks > 0x7fffffffLL) return lsi::set_error(LS_ERR_INVALID, "grid too large for one launch");
        hipLaunchKernelGGL(k_hcr2_zmarch, dim3((unsigned)hg.z.nblocks), dim3(HX, HY, 1), 0, st, hg, pn, p0, pnew);
        if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "fast HCR-2 launch failed");
        return hg.z.fuse_ghosts ? 1 : 0;
    }
    const FastGeom f = make_fast(g);
    if (f.nblocks > 0x7fffffffLL) return lsi::set_error(LS_ERR_INVALID, "grid too large for one launch");
    HPar hp;
    hp.dx2 = g.dx * g.dx;
    hp.invdx = f.invdx;
    hp.invdy = f.invdy;
    hp.invdz = 0.0;
    hp.dtau = g.dtau;
    hp.beta = g.beta;
    const dim3 grid((unsigned)f.nblocks), block(kFX, kFY, 1);
    hipLaunchKernelGGL(k_hcr2_fast2, grid, block, 0, st, f, hp, pn, p0, pnew);
    if (hipGetLastError() != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "fast HCR-2 launch failed");
    return LS_OK;
}

}  // namespace lsf

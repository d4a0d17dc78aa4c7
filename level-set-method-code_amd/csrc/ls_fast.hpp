// ls_fast.hpp -- launchers of the fast-mode kernels (ls_fast.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ls_exact.hpp"

namespace lsf {
template <int DIM>
int launch_stage_fast(const lsx::Geom& g, int stage, const double* Gs, const double* Gn,
                      double* out, const double* u, const double* v, const double* w,
                      const double* sl, double dt, hipStream_t st);
// Exact-mode 3D stage on the shared-face z-march (bitwise the CPU oracle).
// Returns 1 when the kernel also wrote the ghost cells of out.
int launch_stage_exact3d(const lsx::Geom& g, int stage, const double* Gs, const double* Gn, double* out,
                         const double* u, const double* v, const double* w, const double* sl, double dt,
                         hipStream_t st);
// 2D stage as a wavefront y-march (ls_ymarch.hpp), fast or exact arithmetic;
// ymarch_fits says whether the grid gives it enough wavefronts.  Returns 1
// when the kernel also wrote the ghost cells of out.
bool ymarch_fits(const lsx::Geom& g);
int launch_stage_ymarch(const lsx::Geom& g, int stage, const double* Gs, const double* Gn, double* out,
                        const double* u, const double* v, const double* sl, double dt, hipStream_t st, bool exact);
int launch_hcr2_fast(const lsx::Geom& g, const double* pn, const double* p0, double* pnew,
                     hipStream_t st);
// The first two HCR-2 iterations in one z-march (3D), fast or exact arithmetic;
// returns 1 when it ran, 0 when the grid needs single iterations.
int launch_hcr2_pair(const lsx::Geom& g, const double* p0, double* pnew, hipStream_t st, bool exact);
}  // namespace lsf

// ls_internal.hpp -- declarations shared by the translation units of liblsamd.
#pragma once
#include <hip/hip_runtime.h>

#include "ls_api.h"
#include "ls_exact.hpp"

namespace lsi {
// Records msg as the thread's last error and returns code.
int set_error(int code, const char* msg);
int validate(const ls_params_t* p);
lsx::Geom make_geom(const ls_params_t& p);
// The parameters a solver was created with.
const ls_params_t& solver_params(const ls_solver_t* s);
// The solver's device reduction scratch (allocated by ls_create) and its size
// in bytes.  Entry points that take the handle may use it between their own
// launches; the one-handle-per-thread rule keeps uses apart.
void* solver_scratch(ls_solver_t* s, size_t* bytes);
// The HIP device a solver was created on.
int solver_device(const ls_solver_t* s);
// Deformation velocity of the 3D sphere test at time t (ls_cases.hip).
int launch_deformation_velocity(const ls_params_t& p, double t, double period, double* u,
                                double* v, double* w, hipStream_t st);
}  // namespace lsi

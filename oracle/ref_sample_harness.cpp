// ref_sample_harness.cpp -- C entry points over the reference's own sample/
// solver, compiled from its sources where they lie (oracle/Makefile target
// `ref`).  TEST INFRASTRUCTURE ONLY: used to pin oracle/ls_sample.c.
//
// The reference keeps a field as std::vector<std::vector<double>> of
// (ny+6) x (nx+6) (sample/field.hpp:5-24); these wrappers copy flat padded
// arrays in and out around the reference calls.
#include <cstdint>
#include <cstring>
#include <vector>

#include "field.hpp"
#include "first_order_upwind.hpp"
#include "weno5_rk3.hpp"

namespace {

void load(Field2D& f, const double* a) {
    const int nxt = f.nx + 6;
    for (int j = 0; j < f.ny + 6; ++j) std::memcpy(f.data[j].data(), a + (int64_t)j * nxt, nxt * sizeof(double));
}

void store(const Field2D& f, double* a) {
    const int nxt = f.nx + 6;
    for (int j = 0; j < f.ny + 6; ++j) std::memcpy(a + (int64_t)j * nxt, f.data[j].data(), nxt * sizeof(double));
}

}  // namespace

extern "C" {

void refs_rhs(int nx, int ny, const double* G, const double* u, const double* v, double* rhs, double dx,
              double dy, double S_L) {
    Field2D g(nx, ny), uu(nx, ny), vv(nx, ny), r(nx, ny);
    load(g, G);
    load(uu, u);
    load(vv, v);
    compute_geqn_rhs(g.data, uu.data, vv.data, r.data, dx, dy, S_L);
    store(r, rhs);
}

void refs_rk3_step(int nx, int ny, double* G, const double* u, const double* v, double dx, double dy, double dt,
                   double S_L, const double* G_initial) {
    Field2D g(nx, ny), uu(nx, ny), vv(nx, ny), g0(nx, ny);
    load(g, G);
    load(uu, u);
    load(vv, v);
    load(g0, G_initial);
    RK3_step(g, uu, vv, dx, dy, dt, S_L, g0);
    store(g, G);
}

void refs_apply_periodic(int nx, int ny, double* a) {
    Field2D f(nx, ny);
    load(f, a);
    f.apply_periodic();
    store(f, a);
}

void refs_apply_anchored(int nx, int ny, double* a, const double* init) {
    Field2D f(nx, ny), g0(nx, ny);
    load(f, a);
    load(g0, init);
    f.apply_anchored_bc(g0);
    store(f, a);
}

void refs_reinitialize(int nx, int ny, double* a, const double* init, double dx, double dy, int iters) {
    Field2D f(nx, ny), g0(nx, ny);
    load(f, a);
    load(g0, init);
    f.reinitialize(g0, dx, dy, iters);
    store(f, a);
}

void refs_first_order_upwind(int nx, int ny, double* G, const double* u, const double* v, double dx, double dy,
                             double dt, double S_L) {
    Field2D g(nx, ny), uu(nx, ny), vv(nx, ny);
    load(g, G);
    load(uu, u);
    load(vv, v);
    compute_first_order_upwind(g.data, uu.data, vv.data, dx, dy, dt, S_L);
    store(g, G);
}

}  // extern "C"

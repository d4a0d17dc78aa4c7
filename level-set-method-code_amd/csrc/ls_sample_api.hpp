// ls_sample_api.hpp -- C++ drop-in for the reference's sample/ interface
// (sample/field.hpp, sample/weno5_rk3.hpp, save_G_to_bin of sample/main.cpp)
// on top of the C ABI.  Same class and function names, argument order and
// meaning; the solver runs on the GPU.
//
// Field2D keeps the reference's host array `data` ([ny+6][nx+6]) as a mirror
// and a device copy `d` that every solver call works on.  Differences from
// the reference a port must observe:
//   * after writing `data` on the host, call to_device(); before reading it,
//     call to_host() (the reference's main writes the initial field and the
//     velocity on the host: see ls_sample_main.cpp for the port that keeps
//     the velocity on the device instead);
//   * errors print a message and exit, as the reference's file helpers do.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "ls_api.h"

namespace ls_sample_detail {

inline void check(int rc, const char* what) {
    if (rc != LS_OK) {
        std::fprintf(stderr, "%s: %s\n", what, ls_last_error());
        std::exit(EXIT_FAILURE);
    }
}

// One solver context per grid shape and spacing, created on first use.
inline ls_sample_t* context(int nx, int ny, double dx, double dy) {
    static std::map<std::tuple<int, int, double, double>, ls_sample_t*> ctx;
    auto key = std::make_tuple(nx, ny, dx, dy);
    auto it = ctx.find(key);
    if (it != ctx.end()) return it->second;
    ls_sample_t* s = nullptr;
    check(ls_sample_create(nx, ny, dx, dy, &s), "ls_sample_create");
    ctx[key] = s;
    return s;
}

}  // namespace ls_sample_detail

class Field2D {
public:
    int nx, ny;
    std::vector<std::vector<double>> data;  // host mirror [ny+6][nx+6]
    double* d = nullptr;                    // device copy, (ny+6)*(nx+6) doubles

    Field2D(int nx_, int ny_) : nx(nx_), ny(ny_) {
        data.resize(ny + 6, std::vector<double>(nx + 6, 0.0));
        if (hipMalloc((void**)&d, bytes()) != hipSuccess || hipMemset(d, 0, bytes()) != hipSuccess) {
            std::fprintf(stderr, "Field2D: device allocation failed\n");
            std::exit(EXIT_FAILURE);
        }
    }
    Field2D(const Field2D&) = delete;
    Field2D& operator=(const Field2D&) = delete;
    ~Field2D() { (void)hipFree(d); }

    size_t bytes() const { return (size_t)(nx + 6) * (ny + 6) * sizeof(double); }

    void to_device() {
        for (int j = 0; j < ny + 6; ++j)
            (void)hipMemcpy(d + (size_t)j * (nx + 6), data[j].data(), (nx + 6) * sizeof(double), hipMemcpyHostToDevice);
    }
    void to_host() {
        (void)hipDeviceSynchronize();
        for (int j = 0; j < ny + 6; ++j)
            (void)hipMemcpy(data[j].data(), d + (size_t)j * (nx + 6), (nx + 6) * sizeof(double), hipMemcpyDeviceToHost);
    }

    // field.hpp:12 -- host mirror and device copy
    void fill(double value = 0.0) {
        for (auto& row : data) std::fill(row.begin(), row.end(), value);
        to_device();
    }
    // field.cpp:16-35
    void apply_periodic() {
        ls_sample_detail::check(ls_sample_apply_periodic(ctx(), d, nullptr), "apply_periodic");
    }
    // field.cpp:37-63
    void apply_anchored_bc(const Field2D& initial) {
        ls_sample_detail::check(ls_sample_apply_anchored_bc(ctx(), d, initial.d, nullptr), "apply_anchored_bc");
    }
    // field.cpp:66-130
    void reinitialize(const Field2D& initial, double dx, double dy, int num_iters) {
        ls_sample_detail::check(
            ls_sample_reinitialize(ls_sample_detail::context(nx, ny, dx, dy), d, initial.d, num_iters, nullptr),
            "reinitialize");
    }

private:
    // the boundary functions do not depend on the spacing
    ls_sample_t* ctx() const { return ls_sample_detail::context(nx, ny, 1.0, 1.0); }
};

// weno5_rk3.hpp:10-16 (device version: rhs is a Field2D).
inline void compute_geqn_rhs(const Field2D& G, const Field2D& u, const Field2D& v, Field2D& rhs, double dx,
                             double dy, double S_L) {
    ls_sample_detail::check(
        ls_sample_rhs(ls_sample_detail::context(G.nx, G.ny, dx, dy), G.d, u.d, v.d, S_L, rhs.d, nullptr),
        "compute_geqn_rhs");
}

// weno5_rk3.hpp:20-26
inline void RK3_step(Field2D& G_field, const Field2D& u_field, const Field2D& v_field, double dx, double dy,
                     double dt, double S_L, const Field2D& G_initial) {
    ls_sample_detail::check(ls_sample_rk3_step(ls_sample_detail::context(G_field.nx, G_field.ny, dx, dy), G_field.d,
                                               u_field.d, v_field.d, dt, S_L, G_initial.d, nullptr),
                            "RK3_step");
}

// first_order_upwind.hpp (device version on a Field2D)
inline void compute_first_order_upwind(Field2D& G, const Field2D& u, const Field2D& v, double dx, double dy,
                                       double dt, double S_L) {
    ls_sample_detail::check(ls_sample_first_order_upwind(ls_sample_detail::context(G.nx, G.ny, dx, dy), G.d, u.d,
                                                         v.d, dt, S_L, nullptr),
                            "compute_first_order_upwind");
}

// sample/main.cpp:34-51, from the device copy.
inline void save_G_to_bin(Field2D& G, const std::string& filename) {
    G.to_host();
    std::vector<double> flat((size_t)(G.nx + 6) * (G.ny + 6));
    for (int j = 0; j < G.ny + 6; ++j) std::copy(G.data[j].begin(), G.data[j].end(), flat.begin() + (size_t)j * (G.nx + 6));
    ls_sample_detail::check(ls_sample_write_G(G.nx, G.ny, flat.data(), filename.c_str()), "save_G_to_bin");
}

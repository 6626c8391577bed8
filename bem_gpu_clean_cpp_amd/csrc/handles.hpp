// The opaque handles of bem.h, shared by the C-ABI (capi.cpp) and the
// batch / request-server layer (serve.cpp).
#pragma once

#include <mutex>

#include "bem_internal.hpp"

struct bem_mesh {
    bem::Mesh m;
    bem::DeviceProblem dp;
    uint64_t dp_version = ~0ull;
};

struct bem_solver {
    bem::Mesh mesh;
    bem::DeviceProblem dp;
    float k = 0.f;
    int prec = 0;
    float2 *A32 = nullptr;
    double2 *A64 = nullptr;
    bem::RhsSpec rhs;
    double2 *B64 = nullptr;
    double2 *X64 = nullptr;
    float2 *X32 = nullptr;
    size_t a_cap = 0;  // elements allocated for A32 / A64
    size_t x_cap = 0;  // elements allocated for each of B64, X64, X32
    int xr = 0;        // columns of the last solution in X64
    mutable bem_solve_stats stats{};
    // events around the last assembly: its time is read once it has finished
    // (when the stats are read or a solve synchronises), so assembling does
    // not wait for the device
    hipEvent_t ae0 = nullptr, ae1 = nullptr;
    mutable bool asm_pending = false;
    hipEvent_t se0 = nullptr, se1 = nullptr;  // around the last solve
    bem::GmresCache *gcache = nullptr;         // GMRES workspace of the last solve's shape
    // exterior field as a matrix: the mesh with the field points as extra rows
    bem::Mesh fmesh;
    bem::DeviceProblem fdp;
    float2 *F = nullptr;  // P x N radiation operator, ld ldF
    double2 *fbterm = nullptr;
    long ldF = 0;
    size_t f_cap = 0;
    float fk = 0.f;
    int fP = 0;
    hipStream_t stream = nullptr;  // each solver runs on its own stream
    mutable std::mutex mu;         // one call at a time per solver
};

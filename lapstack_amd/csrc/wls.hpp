// wls.hpp -- the edge-aware WLS solver (depth.rs:653-1087) on the device.
#pragma once
#include <memory>

#include "common.hpp"

namespace lsh {

// Solve modes.  REFERENCE is the reference's solver (depth.rs:1006-1077): FP32
// vectors, f64 dot products, the CG recurrence's residual, a V-cycle ending in
// 32 damped-Jacobi sweeps.  REFINED keeps that FP32 MG-PCG as the inner solver
// of an iterative refinement: the iterate and the true residual b - A x are
// FP64, the inner solve finds the correction in FP32, and its V-cycle ends in an
// exact FP64 solve of the coarsest grid -- the only FP64 steps.  FP64 is the
// oracle: FP64 vectors, the FP64 operator (the f32 coefficients promoted), an
// FP64 V-cycle with the exact coarse solve.
enum WlsMode { WLS_REFERENCE = 0, WLS_REFINED = 1, WLS_FP64 = 2 };

struct WlsResult {
    uint32_t iters = 0;  // CG iterations (summed over the inner solves)
    uint32_t outer = 0;  // refinement steps
    double rel = 0.0;    // final relative residual: the recurrence's for REFERENCE, the FP64 true one otherwise
};

struct WlsSolver;
struct WlsDeleter {
    void operator()(WlsSolver *s) const;
};
using WlsPtr = std::unique_ptr<WlsSolver, WlsDeleter>;

// depth.rs:976-988: edges of the guide, the multigrid hierarchy
WlsPtr wls_create(const float *guide, int w, int h, float lambda, float sigma_c);
// The same solver with every buffer in a caller's device block of cap bytes
// (no allocation); wls_scratch_bytes is what a w x h grid needs there.
size_t wls_scratch_bytes(int w, int h);
WlsPtr wls_create_in(void *block, size_t cap, const float *guide, int w, int h, float lambda, float sigma_c);
// depth.rs:992-1002: data weights max(c, 0) + 1e-4 from conf, or, with u set, the
// Huber reweighting conf * min(tau / |d - u|, 1) first (depth.rs:1168)
void wls_set_weights(WlsSolver *s, const float *conf, const float *d, const float *u, float tau);
// solve (W + lambda L) u = W d from the FGS guess (depth.rs:1019-1034), or from
// `warm` (device, w*h) when given; u (device f32) receives the solution; u64
// (device f64, nullable) receives the FP64 iterate in the REFINED and FP64 modes.
WlsResult wls_solve(WlsSolver *s, const float *d, uint32_t max_iters, double tol, int mode, const float *warm,
                    float *u, double *u64);
// ||W d - A u|| / ||W d|| in FP64 for an FP32 u (device buffers)
double wls_true_rel(WlsSolver *s, const float *d, const float *u);
// max |u - u64| over the grid (device buffers)
double wls_max_abs_diff(WlsSolver *s, const float *u, const double *u64);

}  // namespace lsh

// Device-side convergence test and Chebyshev weights shared by the Jacobi
// solvers (Horn-Schunck and Liu-Shen).  Included by k_hs.hip and k_ls.hip.
#pragma once

#include "ofri_internal.h"

namespace ofri {
namespace {

// Test of one launch of `nsw` fused sweeps: `while err > tol and k < maxnum`
// (sweeps before s0 carry no partial sums: they are counted, not tested)
// with err = (||du||_2 + ||dv||_2) / (H*W) (the reference's Liu-Shen update
// norm, PhysicsBasedOpticalFlowLiuShen.py:79,141; the same residual is used
// for Horn-Schunck when a tolerance is given).  Run by the last workgroup of
// the launch to finish, with all its threads: the per-sweep squared-update
// sums of every workgroup are added in a fixed order, the float32 norm of
// each sweep is tested in turn (and recorded in `hist` when given), and the
// control word records the sweep count, the last norm, the buffer holding
// the result and, when the test held before the launch's last sweep (or at
// any sweep, with always_redo), how many sweeps the stopping launch must be
// re-run with.
__device__ void solve_check(const double* part, int nblk, int nsw, int dst, double tol, int maxnum,
                            size_t N, SolveCtrl* ctrl, float* hist, bool always_redo, int s0) {
    // one pass: wavefront q sums sweeps q, q + nwaves, ... over the
    // workgroups (lane-strided, then a fixed butterfly), thread 0 then tests
    // the sweeps in order
    __shared__ double tot[64][2];
    const int T = blockDim.x * blockDim.y;
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = t & 63, wq = t >> 6, nw = (T + 63) >> 6;
    for (int s = s0 + wq; s < nsw; s += nw) {
        double a = 0.0, b = 0.0;
        for (int i = lane; i < nblk; i += 64) {
            // other workgroups' sums: read at device scope, past this CU's cache
            a = a + __hip_atomic_load(part + ((size_t)s * nblk + i) * 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            b = b + __hip_atomic_load(part + ((size_t)s * nblk + i) * 2 + 1, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int off = 32; off > 0; off >>= 1) {
            a = a + __shfl_xor(a, off);
            b = b + __shfl_xor(b, off);
        }
        if (lane == 0) {
            tot[s][0] = a;
            tot[s][1] = b;
        }
    }
    __syncthreads();
    if (t != 0) return;
    ctrl->sweeps += s0;   // sweeps run but not reduced
    for (int s = s0; s < nsw; ++s) {
        float err = (float)((sqrt(tot[s][0]) + sqrt(tot[s][1])) / (double)N);
        if (hist) hist[ctrl->sweeps] = err;
        ctrl->sweeps += 1;
        ctrl->last_err = err;
        ctrl->final_buf = dst;
        if (!((double)err > tol)) {
            ctrl->done = 1;
            if (s + 1 < nsw || always_redo) {
                ctrl->redo = s + 1;
                ctrl->redo_in = dst ^ 1;
            }
            return;
        }
        if (ctrl->sweeps >= maxnum || (ctrl->limit > 0 && ctrl->sweeps >= ctrl->limit)) {
            ctrl->done = 1;
            return;
        }
    }
}

// Called by every workgroup of an image (nblk of them) at the end of a launch
// with per-sweep partial sums: the last one to arrive runs solve_check and
// re-arms the counter.
__device__ __forceinline__ void arrive_and_check(const double* part, int nblk, int nsw, int dst,
                                                 double tol, int maxnum, size_t N, SolveCtrl* ctrl,
                                                 float* hist, bool always_redo = false, int s0 = 0) {
    __shared__ int last;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        __threadfence();
        last = atomicAdd(&ctrl->arrived, 1) == nblk - 1;
    }
    __syncthreads();
    if (last) {
        __threadfence();
        solve_check(part, nblk, nsw, dst, tol, maxnum, N, ctrl, hist, always_redo, s0);
        if (threadIdx.x == 0 && threadIdx.y == 0) ctrl->arrived = 0;
    }
}

// Re-arms a finished solve for a further phase of at most `budget` sweeps
// on the same sweep count and history (the phase starts unless the overall
// budget maxnum is spent).
__global__ void ctrl_continue_k(SolveCtrl* c, int maxnum, int budget) {
    c->done = c->sweeps >= maxnum;
    c->limit = c->sweeps + budget;
    c->final_buf = 0;
    c->redo = 0;
    c->redo_in = 0;
    c->arrived = 0;
}

// Chebyshev weights omega_1..omega_n for a spectrum in [-sigma, sigma]
// (omega_1 = 1, omega_2 = 2 / (2 - sigma^2), omega_{k+1} = 1 / (1 - sigma^2
// omega_k / 4)), formed in double on the host.  The recurrence converges
// geometrically with ratio sigma^2 omega_inf^2 / 4 (at most 0.92 for the
// bounds accepted here), so past the table's 512 entries every weight equals
// the last one in float32.
constexpr int kOmegaTab = 512;
struct OmegaTab {
    float w[kOmegaTab];
};

OmegaTab cheb_omegas(double sigma) {
    OmegaTab t;
    const double s2 = sigma * sigma;
    double om = 1.0;
    for (int k = 0; k < kOmegaTab; ++k) {
        if (k == 1) om = 2.0 / (2.0 - s2);
        else if (k > 1) om = 1.0 / (1.0 - 0.25 * s2 * om);
        t.w[k] = (float)om;
    }
    return t;
}

__global__ void cheb_omega_k(OmegaTab t, float* w, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) w[k] = t.w[k < kOmegaTab ? k : kOmegaTab - 1];
}

__global__ void ctrl_init_k(SolveCtrl* c) {
    c->done = 0;
    c->sweeps = 0;
    c->final_buf = 0;
    c->redo = 0;
    c->redo_in = 0;
    c->last_err = 0.0f;
    c->arrived = 0;
    c->limit = 0;
}

}  // namespace
}  // namespace ofri

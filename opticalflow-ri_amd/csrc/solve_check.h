// Device-side convergence test and Chebyshev weights shared by the Jacobi
// solvers (Horn-Schunck and Liu-Shen).  Included by k_hs.hip and k_ls.hip.
#pragma once

#include "ofri_internal.h"

namespace ofri {
namespace {

// Lane-strided sums of n (x, y) pairs: a += p[2i], b += p[2i + 1] for
// i = lane, lane + 64, ... in that order.  Four strides are loaded before
// any is added, so their latencies overlap (a stride past n re-reads pair
// n - 1 and is not added).
__device__ __forceinline__ void lane_sums(const double* p, int n, int lane, double& a, double& b) {
    for (int i0 = lane; i0 < n; i0 += 256) {
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = min(i0 + 64 * u, n - 1);
            x[u] = p[(size_t)i * 2];
            y[u] = p[(size_t)i * 2 + 1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + 64 * u < n) {
                a = a + x[u];
                b = b + y[u];
            }
    }
}

// Stop test at the start of a launch.  Every workgroup of launch j reads the
// state launch j-1 published (slot `prev`) and, when that launch is pending,
// reduces its per-sweep partial sums (part_prev: [nsw][nblk][2] doubles, the
// squared updates of u and v per workgroup) itself, in one fixed order, so
// that all workgroups reach the same decision with no synchronisation
// between them.  The test is `while err > tol and k < maxnum` with
// err = (||du||_2 + ||dv||_2) / (H*W) (the reference's Liu-Shen update norm,
// PhysicsBasedOpticalFlowLiuShen.py:79,141; the same residual serves
// Horn-Schunck when a tolerance is given), evaluated sweep by sweep in float32
// (recorded in `hist` when given).  When it holds before the pending launch's
// last sweep (or at any sweep, with always_redo) the state asks for that
// launch to be re-run with the sweeps up to the stopping one.  Returns the
// updated state, the same in every thread; nsw <= 32.  rel: the tolerance
// is relative -- tol * (||u||_2 + ||v||_2) / (H*W), the norm of the tested
// launch's result, whose per-workgroup squared sums sit in row K of the
// partials (written once per launch, after its last sweep; the test then
// covers that sweep alone).
__device__ SolveState take_test(const SolveCtrl* ctrl, int prev, const double* part_prev, int nblk, double tol,
                                int maxnum, size_t N, float* hist, bool always_redo, int rel, int K) {
    __shared__ double tot[32][2];
    __shared__ double totn[2];
    __shared__ SolveState sh;
    SolveState s = ctrl->slot[prev];
    const bool ev = !s.done && s.pend_nsw > 0;
    if (!ev) {
        // nothing to test (a finished solve, or the first launch): every
        // thread already holds the state
        s.pend_nsw = 0;
        return s;
    }
    const int T = blockDim.x * blockDim.y;
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = t & 63, wq = t >> 6, nw = (T + 63) >> 6;
    {
        // wavefront q sums sweeps q, q + nw, ... over the workgroups
        // (lane-strided, then a fixed butterfly)
        for (int k = s.pend_s0 + wq; k < s.pend_nsw; k += nw) {
            double a = 0.0, b = 0.0;
            lane_sums(part_prev + (size_t)k * nblk * 2, nblk, lane, a, b);
            for (int off = 32; off > 0; off >>= 1) {
                a = a + __shfl_xor(a, off);
                b = b + __shfl_xor(b, off);
            }
            if (lane == 0) {
                tot[k][0] = a;
                tot[k][1] = b;
            }
        }
        if (rel && wq == nw - 1) {
            // the tested launch's solution norm partials (row K of the slot)
            const double* pn = part_prev + (size_t)K * nblk * 2;
            double a = 0.0, b = 0.0;
            lane_sums(pn, nblk, lane, a, b);
            for (int off = 32; off > 0; off >>= 1) {
                a = a + __shfl_xor(a, off);
                b = b + __shfl_xor(b, off);
            }
            if (lane == 0) {
                totn[0] = a;
                totn[1] = b;
            }
        }
    }
    __syncthreads();
    if (t == 0) {
        {
            const int nsw = s.pend_nsw, dst = s.pend_dst;
            s.sweeps += s.pend_s0;   // sweeps run but not reduced
            // relative mode: the tolerance scales with the launch's solution norm
            double teff = tol;
            if (rel) {
                s.last_nrm = (float)((sqrt(totn[0]) + sqrt(totn[1])) / (double)N);
                teff = tol * (double)s.last_nrm;
            }
            for (int k = s.pend_s0; k < nsw; ++k) {
                float err = (float)((sqrt(tot[k][0]) + sqrt(tot[k][1])) / (double)N);
                if (hist) hist[s.sweeps] = err;
                s.sweeps += 1;
                s.last_err = err;
                s.final_buf = dst;
                if (!((double)err > teff)) {
                    s.done = 1;
                    if (k + 1 < nsw || always_redo) {
                        s.redo = k + 1;
                        s.redo_in = dst ^ 1;
                    }
                    break;
                }
                if (s.sweeps >= maxnum || (s.limit > 0 && s.sweeps >= s.limit)) {
                    s.done = 1;
                    break;
                }
            }
        }
        s.pend_nsw = 0;
        sh = s;
    }
    __syncthreads();
    return sh;
}

// A launch's sweep count, input buffer and first global sweep in residual
// mode.  Pass provides ctrl, part (two slots of part_stride doubles), slot
// (launch parity), final_, next_budget, rel_src, nsweeps, src, sweep0,
// check_last, tol, maxnum, N and hist.  A regular launch runs unless the
// solve is done, and its workgroup 0 publishes the state with this launch
// pending; its buffers are relative to ctrl->base_buf when rel_src.  The
// phase's final launch (final_) runs only the re-run of a launch whose test
// held inside it, and publishes the result; with next_budget > 0 it also
// arms a following phase of at most that many sweeps (starting from the
// result's buffer, reading the state slot this launch writes: its launches
// take the opposite parity).  fin (optional) receives the buffer holding the
// solve's result as this launch found it.
template <class Pass>
__device__ __forceinline__ bool launch_args(const Pass& a, int blk, int nblk, bool always_redo, int& nsw, int& src,
                                            int& base, int* fin = nullptr) {
    const int prev = a.slot ^ 1;
    SolveState s = take_test(a.ctrl, prev, a.part + (size_t)prev * a.part_stride, nblk, a.tol, a.maxnum, a.N,
                             blk == 0 ? a.hist : nullptr, always_redo, a.rel, a.K);
    const bool lead = blk == 0 && threadIdx.x == 0 && threadIdx.y == 0;
    if (fin) *fin = s.final_buf;
    if (a.final_) {
        nsw = s.done ? s.redo : 0;
        src = s.redo_in;
        base = s.sweeps - nsw;
        if (lead) {
            SolveCtrl* c = a.ctrl;
            c->done = s.done;
            c->sweeps = s.sweeps;
            c->final_buf = s.final_buf;
            c->redo = s.redo;
            c->redo_in = s.redo_in;
            c->last_err = s.last_err;
            c->last_nrm = s.last_nrm;
            c->limit = s.limit;
            if (a.next_budget > 0) {
                SolveState n{};
                n.done = s.sweeps >= a.maxnum;
                n.sweeps = s.sweeps;
                n.final_buf = s.final_buf;
                n.last_err = s.last_err;
                n.last_nrm = s.last_nrm;
                n.limit = s.sweeps + a.next_budget;
                c->slot[a.slot] = n;
                c->base_buf = s.final_buf;
            }
        }
        return nsw > 0;
    }
    nsw = a.nsweeps;
    src = a.src ^ (a.rel_src ? a.ctrl->base_buf : 0);
    base = a.sweep0;
    const bool runs = !s.done;
    if (lead) {
        if (runs) {
            s.pend_nsw = nsw;
            s.pend_dst = src ^ 1;
            s.pend_s0 = a.check_last ? nsw - 1 : 0;
        }
        a.ctrl->slot[a.slot] = s;
    }
    return runs;
}

// Relative mode: the squared sums of a launch's result over the owned
// pixels of this workgroup's tile (u, v: the lane's R rows after the last
// sweep), stored at dst (row K of the launch's partials): wave sums in a
// fixed butterfly, then the waves in order.
template <int R, int WAVES>
__device__ __forceinline__ void store_norm(const float (&u)[R], const float (&v)[R], bool own_x, int w, int gy0,
                                           int K, int H, double* dst) {
    __shared__ double nred[WAVES][2];
    const int lane = threadIdx.x & 63;
    double su = 0.0, sv = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ry = w * R + r, gy = gy0 + r;
        if (own_x && ry >= K && ry < WAVES * R - K && gy < H) {
            su = su + (double)u[r] * (double)u[r];
            sv = sv + (double)v[r] * (double)v[r];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        su = su + __shfl_xor(su, off);
        sv = sv + __shfl_xor(sv, off);
    }
    if (lane == 0) {
        nred[w][0] = su;
        nred[w][1] = sv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < WAVES; ++q) {
            a = a + nred[q][0];
            b = b + nred[q][1];
        }
        dst[0] = a;
        dst[1] = b;
    }
}

// Chebyshev weights omega_1..omega_n for a spectrum in [-sigma, sigma]
// (omega_1 = 1, omega_2 = 2 / (2 - sigma^2), omega_{k+1} = 1 / (1 - sigma^2
// omega_k / 4)), formed in double on the host.  The recurrence converges
// geometrically with ratio sigma^2 omega_inf^2 / 4 (at most 0.92 for the
// bounds accepted here), so past the table's 512 entries every weight equals
// the last one in float32: sweep k reads entry min(k, kOmegaTab - 1) of the
// table, which lives in device memory (omega_table, one per sigma).
constexpr int kOmegaTab = 512;

__device__ __forceinline__ int omega_index(int k) { return k < kOmegaTab ? k : kOmegaTab - 1; }

}  // namespace

// Device table of kOmegaTab Chebyshev weights for sigma, owned by the
// context (built and uploaded on first use, outside any graph capture).
int omega_table(ofri_ctx* ctx, double sigma, const float** out);

namespace {

}  // namespace
}  // namespace ofri

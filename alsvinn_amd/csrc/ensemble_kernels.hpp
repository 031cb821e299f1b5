// Elementwise, ghost-fill and bookkeeping kernels of the hip platform.
#pragma once

#include "../../include/alsvinn_hip.h"
#include "march.hpp"

namespace alsvinn_hip {

// out = X*c + Y over every cell of sample D's padded volume (ghosts
// included), c = dt/dx_x of the sample: the integrator's
// `output *= dt/dx; output += U` (ForwardEuler.cpp:42-46, RungeKutta2.cpp:53-57).
// The calling threads take the cells first, first + stride, ... With check
// set, interior cells are tested against the equation's constraints
// (Simulator::checkConstraints); returns whether one of this thread's failed.
// axpy_sample forms c from the sample's wave speeds (one thread, shared
// through LDS) and records a failure in the sample's status.
template <class T, class Eq>
__device__ __forceinline__ bool axpy_cells(const SampleDesc& D, T c, const T* X, const T* Y, T* out,
                                           int check, double gamma, long long first,
                                           long long stride) {
    const long long n = D.vstride;
    const T* x = X + D.base;
    const T* y = Y + D.base;
    T* o = out + D.base;
    Eq eq;
    eq.gamma = T(gamma);
    bool bad = false;
    for (long long i = first; i < n; i += stride) {
        T r[Eq::nvar];
#pragma unroll
        for (int v = 0; v < Eq::nvar; ++v) {
            r[v] = x[i + v * n] * c + y[i + v * n];
            o[i + v * n] = r[v];
        }
        if (check) {
            const int px = int(i % D.tx);
            const int py = int((i / D.tx) % D.ty);
            const int pz = int(i / ((long long)D.tx * D.ty));
            const bool inside = px >= D.gx && px < D.gx + D.nx && py >= D.gy && py < D.gy + D.ny &&
                                pz >= D.gz && pz < D.gz + D.nz;
            if (inside && !eq.valid(r)) bad = true;
        }
    }
    return bad;
}

template <class T, class Eq>
__device__ __forceinline__ bool axpy_sample(const SampleDesc& D, SampleState* Sp,
                                            const unsigned long long* __restrict__ wss, const T* X,
                                            const T* Y, T* out, int check, double gamma,
                                            long long first, long long stride) {
    __shared__ double cs;
    if (threadIdx.x == 0) cs = sample_dt(D, *Sp, wss) / D.dx[0];
    __syncthreads();
    const bool bad = axpy_cells<T, Eq>(D, T(cs), X, Y, out, check, gamma, first, stride);
    if (bad) atomicOr(&Sp->status, 2);
    return bad;
}

// The kernel: blockIdx.y selects the sample, blockIdx.x of gridDim.x shares
// its cells; `end` ends the step after the work when this is the step's
// last kernel.
template <class T, class Eq>
__global__ __launch_bounds__(256) void axpy_kernel(const SampleDesc* __restrict__ desc,
                                                  SampleState* state,
                                                  const unsigned long long* __restrict__ ws,
                                                  const T* X, const T* Y, T* out, int check,
                                                  double gamma, StepEnd end) {
    const int s = blockIdx.y;
    const SampleDesc D = desc[s];
    SampleState* Sp = state + s;
    const bool bad = runnable(*Sp) &&
                     axpy_sample<T, Eq>(D, Sp, ws + 3 * s, X, Y, out, check, gamma,
                                        (long long)blockIdx.x * blockDim.x + threadIdx.x,
                                        (long long)gridDim.x * blockDim.x);
    end_step(end, bad);
}

// Small 1D solves, whose steps are too short to fill the GPU or to hide a
// launch: one workgroup per sample runs up to `nsteps` whole timesteps
// (Simulator::performStep, Simulator.cpp:77-85) in one launch, with a
// workgroup barrier where the multi-launch step has a kernel boundary. Its
// wavefronts take the sample's strips of W = 64 - 2G columns in turn
// through the same marching code, and the combinations, the dt and the step
// end are the same functions, so the results are the same bits. The
// sample's state stays in registers (every thread holds the same copy, so
// the step loop and its exit are uniform over the workgroup; thread 0 writes
// it back after each step), and the wave-speed maxima and the failure flag
// are kept in LDS: the step needs no global round trip besides the volumes.
template <class T>
struct Solve1DArgs {
    StageArgs<T> a;  // desc, state, ws, gamma
    T* U;
    T* A;
    T* B;
    int integrator;  // ALSVINN_HIP_INTEGRATOR_*: forward Euler, RK2, SSP-RK3
    int nsteps;
    double* dt_hist;
    long long dt_cap;
};

// One stage over the strips of sample s (single-sample waves).
template <class T, class Eq, class Recon, class Flux, int MODE>
__device__ __forceinline__ void march_sample_1d(const StageArgs<T>& a, const Eq& eq, int s, int nx,
                                                const LocalStep& ls, bool& flagged) {
    constexpr int G = Recon::ghosts;
    constexpr int W = 64 - 2 * G;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int ntasks = (nx + W - 1) / W;
    for (int t = wave; t < ntasks; t += nw) {
        const int x0 = t * W;
        const int ncols = nx - x0 < W ? nx - x0 : W;
        const bool in_strip = lane < ncols + 2 * G;
        march_12<T, Eq, Recon, Flux, 1, MODE, 0, true, true>(
            a, eq, t, s, x0, ncols, 0, in_strip, in_strip ? 0 : -1, 0, 1, flagged, ls);
    }
}

template <class T, class Eq, class Recon, class Flux>
__global__ __launch_bounds__(1024) void solve_kernel_1d(Solve1DArgs<T> p) {
    // maxima of even and odd steps: a step raises one entry while the other,
    // read by every thread in the step before, is cleared for the next
    __shared__ unsigned long long s_ws[2][3];
    __shared__ int s_status;
    const int s = blockIdx.x;
    StageArgs<T> a = p.a;
    Eq eq;
    eq.gamma = T(a.gamma);
    const SampleDesc D = a.desc[s];
    SampleState S = a.state[s];
    if (threadIdx.x < 6) s_ws[threadIdx.x / 3][threadIdx.x % 3] = 0;
    if (threadIdx.x == 0) s_status = 0;
    bool flagged = false;
    for (int k = 0; k < p.nsteps && runnable(S); ++k) {
        LocalStep ls{s_ws[k & 1], &s_status, 0.0};
        __syncthreads();  // the previous step's results and the cleared maxima
        // first stage: A = L(U) and the wave speeds
        a.in = p.U;
        a.in2 = p.U;
        a.out = p.A;
        march_sample_1d<T, Eq, Recon, Flux, MODE_RHS>(a, eq, s, D.nx, ls, flagged);
        __syncthreads();
        const unsigned long long w3[3] = {ls.ws[0], ls.ws[1], ls.ws[2]};
        ls.dt = sample_dt(D, S, w3);
        if (p.integrator == ALSVINN_HIP_INTEGRATOR_FORWARD_EULER) {
            if (axpy_cells<T, Eq>(D, T(ls.dt / D.dx[0]), p.A, p.U, p.U, 1, a.gamma, threadIdx.x,
                                  blockDim.x))
                s_status = 2;
        } else if (p.integrator == ALSVINN_HIP_INTEGRATOR_RK2) {
            axpy_cells<T, Eq>(D, T(ls.dt / D.dx[0]), p.A, p.U, p.A, 0, a.gamma, threadIdx.x,
                              blockDim.x);
            __syncthreads();
            a.in = p.A;
            a.in2 = p.U;
            a.out = p.U;
            march_sample_1d<T, Eq, Recon, Flux, MODE_RK2_1>(a, eq, s, D.nx, ls, flagged);
        } else {
            a.in = p.A;
            a.in2 = p.U;
            a.out = p.B;
            march_sample_1d<T, Eq, Recon, Flux, MODE_RK3_1>(a, eq, s, D.nx, ls, flagged);
            __syncthreads();
            a.in = p.B;
            a.in2 = p.U;
            a.out = p.U;
            march_sample_1d<T, Eq, Recon, Flux, MODE_RK3_2>(a, eq, s, D.nx, ls, flagged);
        }
        __syncthreads();
        // the step's end (finish_sample) in every thread's copy
        if (threadIdx.x < 3) s_ws[(k + 1) & 1][threadIdx.x] = 0;
        if (threadIdx.x == 0 && p.dt_hist && S.steps < p.dt_cap)
            p.dt_hist[(long long)s * p.dt_cap + S.steps] = ls.dt;
        S.status |= s_status;
        advance_sample(D, S, ls.dt);
        if (threadIdx.x == 0) a.state[s] = S;
    }
}

// Stage combinations of the classical RK4 integrator (RungeKutta4.cpp:50-96)
// over whole volumes (ghost images included), c = dt/dx per sample:
//   kind 0: out = ((X*c)*0.5) + Y                       (U1, U2; Y = U0)
//   kind 1: out = (((((c/6)*X + (1/3)*Z) + (2/3)*W) + (1/3)*V) + (-1/3)*Y)
//           with X = L(U3), Z = U1, W = U2, V = U3, Y = U0
//           (HostMemory::addLinearCombination, HostMemory.cpp:266-279),
//           followed by the constraint check when `check` is set.
template <class T, class Eq>
__device__ __forceinline__ bool rk4_sample(const SampleDesc& D, SampleState* Sp,
                                           const unsigned long long* __restrict__ wss, int kind,
                                           const T* X, const T* Y, const T* Z, const T* W,
                                           const T* V, T* out, int check, double gamma) {
    __shared__ double cs;
    if (threadIdx.x == 0) cs = sample_dt(D, *Sp, wss) / D.dx[0];
    __syncthreads();
    const T c = T(cs);
    const T a1 = T((1.0 / 6.0) * cs);
    const long long n = D.vstride;
    const long long b = D.base;
    Eq eq;
    eq.gamma = T(gamma);
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        T r[Eq::nvar];
#pragma unroll
        for (int v = 0; v < Eq::nvar; ++v) {
            const long long j = b + i + v * n;
            if (kind == 0) {
                r[v] = ((X[j] * c) * T(0.5)) + Y[j];
            } else {
                r[v] = ((((a1 * X[j] + T(1.0 / 3.0) * Z[j]) + T(2.0 / 3.0) * W[j]) +
                         T(1.0 / 3.0) * V[j]) +
                        T(-1.0 / 3.0) * Y[j]);
            }
            out[j] = r[v];
        }
        if (check) {
            const int px = int(i % D.tx);
            const int py = int((i / D.tx) % D.ty);
            const int pz = int(i / ((long long)D.tx * D.ty));
            const bool inside = px >= D.gx && px < D.gx + D.nx && py >= D.gy && py < D.gy + D.ny &&
                                pz >= D.gz && pz < D.gz + D.nz;
            if (inside && !eq.valid(r)) bad = true;
        }
    }
    if (bad) atomicOr(&Sp->status, 2);
    return bad;
}

// The kernel, laid out as axpy_kernel.
template <class T, class Eq>
__global__ __launch_bounds__(256) void rk4_kernel(const SampleDesc* __restrict__ desc,
                                                 SampleState* state,
                                                 const unsigned long long* __restrict__ ws,
                                                 int kind, const T* X, const T* Y, const T* Z,
                                                 const T* W, const T* V, T* out, int check,
                                                 double gamma, StepEnd end) {
    const int s = blockIdx.y;
    const SampleDesc D = desc[s];
    SampleState* Sp = state + s;
    const bool bad =
        runnable(*Sp) && rk4_sample<T, Eq>(D, Sp, ws + 3 * s, kind, X, Y, Z, W, V, out, check, gamma);
    end_step(end, bad);
}

// Source cell of padded position p on one axis (BoundaryCPU.cpp:53-104 with
// Neumann.hpp / Periodic.hpp; corners compose axis by axis).
__device__ __forceinline__ int ghost_source(int p, int N, int G, int bc) {
    if (p >= G && p < N + G) return p;
    if (p < G) return bc == 1 ? p + N : 2 * G - 1 - p;
    return bc == 1 ? p - N : 2 * N + 2 * G - 1 - p;
}

// Boundary::applyBoundaryConditions for every sample: fills each ghost cell
// from its interior source. blockIdx.y selects the sample (first + y).
template <class T>
__global__ __launch_bounds__(256) void ghost_fill_kernel(const SampleDesc* __restrict__ desc,
                                                        T* U, int nvar, int first) {
    const int s = first + blockIdx.y;
    const SampleDesc D = desc[s];
    const long long n = D.vstride;
    T* u = U + D.base;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int px = int(i % D.tx);
        const int py = int((i / D.tx) % D.ty);
        const int pz = int(i / ((long long)D.tx * D.ty));
        const int sx = ghost_source(px, D.nx, D.gx, D.bc);
        const int sy = ghost_source(py, D.ny, D.gy, D.bc);
        const int sz = ghost_source(pz, D.nz, D.gz, D.bc);
        if (sx == px && sy == py && sz == pz) continue;
        const long long j = ((long long)sz * D.ty + sy) * D.tx + sx;
        for (int v = 0; v < nvar; ++v) u[i + v * n] = u[j + v * n];
    }
}

// Boundary::applyBoundaryConditions on one reference Volume (per-variable
// memories; BoundaryCPU.cpp:30-121): the same ghost sources as
// ghost_fill_kernel. D describes the volume (base 0).
template <class T>
__global__ __launch_bounds__(256) void ghost_fill_volume_kernel(PerVarVol<T> u, int nvar,
                                                               SampleDesc D) {
    const long long n = D.vstride;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int px = int(i % D.tx);
        const int py = int((i / D.tx) % D.ty);
        const int pz = int(i / ((long long)D.tx * D.ty));
        const int sx = ghost_source(px, D.nx, D.gx, D.bc);
        const int sy = ghost_source(py, D.ny, D.gy, D.bc);
        const int sz = ghost_source(pz, D.nz, D.gz, D.bc);
        if (sx == px && sy == py && sz == pz) continue;
        const long long j = ((long long)sz * D.ty + sy) * D.tx + sx;
        for (int v = 0; v < nvar; ++v) u(v, i) = u(v, j);
    }
}

}  // namespace alsvinn_hip

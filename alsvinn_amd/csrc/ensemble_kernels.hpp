// Elementwise, ghost-fill and bookkeeping kernels of the hip platform.
#pragma once

#include "march.hpp"

namespace alsvinn_hip {

// out = X*c + Y over every cell of each sample's padded volume (ghosts
// included), c = dt/dx_x of the sample: the integrator's
// `output *= dt/dx; output += U` (ForwardEuler.cpp:42-46, RungeKutta2.cpp:53-57).
// blockIdx.y selects the sample. With check set, interior cells are tested
// against the equation's constraints (Simulator::checkConstraints).
template <class T, class Eq>
__device__ __forceinline__ bool axpy_sample(const SampleDesc& D, SampleState* Sp,
                                            const unsigned long long* __restrict__ wss, const T* X,
                                            const T* Y, T* out, int check, double gamma) {
    __shared__ double cs;
    if (threadIdx.x == 0) cs = sample_dt(D, *Sp, wss) / D.dx[0];
    __syncthreads();
    const T c = T(cs);
    const long long n = D.vstride;
    const T* x = X + D.base;
    const T* y = Y + D.base;
    T* o = out + D.base;
    Eq eq;
    eq.gamma = T(gamma);
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
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
            if (inside && !eq.obeys(r)) bad = true;
        }
    }
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
    const bool bad = runnable(*Sp) && axpy_sample<T, Eq>(D, Sp, ws + 3 * s, X, Y, out, check, gamma);
    end_step(end, bad);
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
            if (inside && !eq.obeys(r)) bad = true;
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

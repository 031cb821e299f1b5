// Fused stage kernels of the hip platform.
//
// One launch evaluates one SSP-RK stage for every sample of an ensemble:
// reconstruction + face flux + divergence (NumericalFluxCPU.cpp:32-118 with
// ConservedSystem.cpp:31-103), the integrator's stage combination
// (ForwardEuler.cpp:42-46, RungeKutta2.cpp:53-64, RungeKutta3.cpp:65-80), the
// ghost-cell fill of the result (BoundaryCPU.cpp:53-104), the per-sample
// max-wave-speed reduction that feeds the CFL condition (Integrator.cpp:20-46)
// and the constraint check (Simulator.cpp:180-189).
//
// Work decomposition. A wavefront owns a strip of columns (x) of one sample
// and marches upward in y (and z for 3D), keeping a sliding register window
// of the rows it needs. x-neighbours come from the adjacent lanes (ds_bpermute
// shuffles); the G = ghost-count lanes at each end of a strip are halo lanes
// that only feed their neighbours. y-neighbours come from the window, so each
// reconstruction and each face flux is evaluated once per cell. Nothing is
// staged through LDS and no barrier is needed: waves are independent.
#pragma once

#include "physics.hpp"

namespace alsvinn_hip {

enum StageMode : int {
    MODE_RHS = 0,    // out = L(U)                   (first stage; wave speeds kept)
    MODE_RK2_1 = 1,  // out = ((L*c + U1) + U0)*0.5
    MODE_RK3_1 = 2,  // out = (((L*c + U1)*(1/3)) + U0)*(3/4)
    MODE_RK3_2 = 3,  // out = (((L*c + U2)*2) + U0)*(1/3)
};

struct SampleDesc {
    long long base;     // element offset of variable 0 of this sample
    long long vstride;  // elements per variable (tx*ty*tz)
    int nx, ny, nz;     // interior cells
    int tx, ty, tz;     // padded extents
    int gx, gy, gz;     // ghost layers per axis (0 on inactive axes)
    int bc;             // 0 neumann, 1 periodic
    double dx[3];       // cell lengths (1 on inactive axes)
    double cfl;
    double end_time;
    int nadj;
    int pad0;
    double interval[2];
    long long max_steps;
};

struct SampleState {
    double t;
    double last_dt;
    long long steps;
    long long saved[2];
    int active;  // 1 while the sample still steps
    int status;  // 0 ok, 2 constraint violation
};

// A strip of one sample: interior columns [x0, x0+ncols), rows [y0, y1),
// planes [z0, z1).
struct WaveTask {
    int sample;
    int x0, ncols;
    int y0, y1;
    int z0, z1;
    int pad;
};

template <class T>
struct StageArgs {
    const SampleDesc* desc;
    SampleState* state;
    unsigned long long* ws;  // 3 per sample: bit patterns of non-negative doubles
    const WaveTask* tasks;
    int ntasks;
    int mode;
    int lazy;   // stage input = in*c + in2 (formed on the fly)
    int check;  // evaluate the constraint check on the result
    const T* in;
    const T* in2;  // U0
    T* out;
    double gamma;
};

// Integrator::computeTimestep + FixedInterval adjusters (Integrator.cpp:20-66,
// FixedIntervalWriter.cpp:45-53, FixedIntervalStatistics.cpp:33-41), in FP64.
__device__ __forceinline__ double sample_dt(const SampleDesc& D, const SampleState& S,
                                            const unsigned long long* ws) {
    double total = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (D.dx[d] == 0) continue;
        total += __longlong_as_double((long long)ws[d]) / D.dx[d];
    }
    double dt = D.cfl / total;
    for (int k = 0; k < D.nadj; ++k) {
        if (S.saved[k] > 0) {
            const double next = double(S.saved[k]) * D.interval[k] + 0.0;
            const double cand = next - S.t;
            dt = (cand < dt) ? cand : dt;
        }
    }
    return dt;
}

// Padded positions holding the value of interior cell c on one axis: the
// cell itself plus its ghost images (Neumann mirror / periodic wrap).
__device__ __forceinline__ int axis_images(int c, int N, int G, int bc, int* pos) {
    pos[0] = c + G;
    int n = 1;
    if (G == 0) return n;
    if (bc == 1) {
        if (c >= N - G) pos[n++] = c - N + G;
        if (c < G) pos[n++] = c + G + N;
    } else {
        if (c < G) pos[n++] = G - 1 - c;
        if (c >= N - G) pos[n++] = 2 * N + G - 1 - c;
    }
    return n;
}

template <class T>
__device__ __forceinline__ T shfl_up1(T v) {
    return __shfl_up(v, 1, 64);
}
template <class T>
__device__ __forceinline__ T shfl_down1(T v) {
    return __shfl_down(v, 1, 64);
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = std_max(v, __shfl_xor(v, off, 64));
    return v;
}

// Stage input at flat index i (variable stride vs): plain or lazy U1 = L0*c + U0.
template <class T, int NV>
__device__ __forceinline__ void load_cell(const T* __restrict__ in, const T* __restrict__ in2,
                                          bool lazy, T c, long long i, long long vs, T* q) {
    if (lazy) {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = in[i + v * vs] * c + in2[i + v * vs];
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = in[i + v * vs];
    }
}

template <class T, class Eq, class Flux, int d>
__device__ __forceinline__ T face_flux(const Eq& eq, const T* qa, const T* qb, T* F) {
    typename Eq::All A, B;
    eq.make(qa, A);
    eq.make(qb, B);
    return Flux::template compute<d>(eq, A, B, F);
}

template <class T, int NV>
__device__ __forceinline__ void combine(int mode, T c, const T* Lv, const T* uin, const T* u0,
                                        T* res) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        T r = Lv[v];
        if (mode == MODE_RK2_1) {
            r = r * c;
            r = r + uin[v];
            r = r + u0[v];
            r = r * T(0.5);
        } else if (mode == MODE_RK3_1) {
            r = r * c;
            r = r + uin[v];
            r = r * T(1. / 3.);
            r = r + u0[v];
            r = r * T(3. / 4.);
        } else if (mode == MODE_RK3_2) {
            r = r * c;
            r = r + uin[v];
            r = r * T(2);
            r = r + u0[v];
            r = r * T(1. / 3.);
        }
        res[v] = r;
    }
}

// 1D and 2D stage kernel (DIM = active grid dimensions).
template <class T, class Eq, class Recon, class Flux, int DIM>
__global__ __launch_bounds__(256) void stage_kernel_12(StageArgs<T> a) {
    constexpr int NV = Eq::nvar;
    constexpr int G = Recon::ghosts;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    Eq eq;
    eq.gamma = T(a.gamma);

    for (int w = wave; w < a.ntasks; w += nwaves) {
        const WaveTask tk = a.tasks[w];
        const SampleDesc D = a.desc[tk.sample];
        SampleState* Sp = a.state + tk.sample;
        if (!Sp->active) continue;  // uniform across the wave
        const bool need_dt = a.lazy || a.mode != MODE_RHS;
        T c = T(0);
        if (need_dt) c = T(sample_dt(D, *Sp, a.ws + 3 * tk.sample) / D.dx[0]);

        const int col = tk.x0 - G + lane;  // interior x coordinate of this lane
        const bool in_strip = lane < tk.ncols + 2 * G;
        const bool interior = lane >= G && lane < G + tk.ncols;
        const bool xface = lane >= G - 1 && lane < G + tk.ncols;  // face col+1/2
        const int px = in_strip ? col + D.gx : D.gx;               // padded column
        const long long vs = D.vstride;
        const T* in = a.in + D.base;
        const T* in2 = a.in2 + D.base;
        T* out = a.out + D.base;

        T wsx = T(0), wsy = T(0);

        // Rows are addressed in padded coordinates; loads outside the volume
        // are clamped (they only occur for G == 1, whose neighbours are unused).
        auto row_index = [&](int py) -> long long {
            py = py < 0 ? 0 : (py >= D.ty ? D.ty - 1 : py);
            return (long long)py * D.tx + px;
        };

        T yc[NV], yp[NV];          // rows r, r+1 of the stage input
        T Ry[NV];                  // right (r+1/2) reconstruction of row r in y
        T Gm[NV];                  // y face flux at r-1/2
        const int gy = D.gy;
        if constexpr (DIM >= 2) {
            const int pr = tk.y0 + gy;
            T qa[NV], ym[NV];
            load_cell<T, NV>(in, in2, a.lazy, c, row_index(pr - 2), vs, qa);
            load_cell<T, NV>(in, in2, a.lazy, c, row_index(pr - 1), vs, ym);
            load_cell<T, NV>(in, in2, a.lazy, c, row_index(pr), vs, yc);
            load_cell<T, NV>(in, in2, a.lazy, c, row_index(pr + 1), vs, yp);
            T Lt[NV], Rt[NV];
            Recon::template apply<T, NV>(qa, ym, yc, Lt, Rt);  // row r-1: keep R
            T L0[NV], R0[NV];
            Recon::template apply<T, NV>(ym, yc, yp, L0, R0);  // row r
            const T s = face_flux<T, Eq, Flux, 1>(eq, Rt, L0, Gm);
            if (interior) wsy = std_max(wsy, s);
#pragma unroll
            for (int v = 0; v < NV; ++v) Ry[v] = R0[v];
        } else {
            load_cell<T, NV>(in, in2, a.lazy, c, row_index(0), vs, yc);
        }

        for (int r = tk.y0; r < tk.y1; ++r) {
            const int pr = r + gy;
            T out_v[NV];
            // ---- x direction on row r (first: out = 0 + x divergence)
            {
                T xm[NV], xp[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    xm[v] = shfl_up1(yc[v]);
                    xp[v] = shfl_down1(yc[v]);
                }
                T Lx[NV], Rx[NV];
                Recon::template apply<T, NV>(xm, yc, xp, Lx, Rx);
                T Ln[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) Ln[v] = shfl_down1(Lx[v]);
                T Fp[NV];
                const T s = face_flux<T, Eq, Flux, 0>(eq, Rx, Ln, Fp);
                if (xface) wsx = std_max(wsx, s);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const T Fm = shfl_up1(Fp[v]);
                    const T t0 = (T(-1.0) * Fp[v]) + Fm;
                    out_v[v] = T(0) + t0;
                }
            }
            // ---- y direction: out += (-G(r+1/2)) + G(r-1/2)
            if constexpr (DIM >= 2) {
                T qn[NV];
                load_cell<T, NV>(in, in2, a.lazy, c, row_index(pr + 2), vs, qn);
                T Ln[NV], Rn[NV];
                Recon::template apply<T, NV>(yc, yp, qn, Ln, Rn);  // row r+1
                T Gp[NV];
                const T s = face_flux<T, Eq, Flux, 1>(eq, Ry, Ln, Gp);
                if (interior) wsy = std_max(wsy, s);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    out_v[v] = out_v[v] + ((T(-1.0) * Gp[v]) + Gm[v]);
                    Gm[v] = Gp[v];
                    Ry[v] = Rn[v];
                }
                // epilogue below uses yc (row r); shift afterwards
                if (interior) {
                    const long long i0 = (long long)pr * D.tx + px;
                    T u0[NV], res[NV];
                    if (a.mode != MODE_RHS) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) u0[v] = in2[i0 + v * vs];
                    }
                    combine<T, NV>(a.mode, c, out_v, yc, u0, res);
                    if (a.check && !eq.obeys(res)) atomicOr(&Sp->status, 2);
                    int xs[3], ys[3];
                    const int nxi = axis_images(col, D.nx, D.gx, D.bc, xs);
                    const int nyi = axis_images(r, D.ny, D.gy, D.bc, ys);
                    for (int iy = 0; iy < nyi; ++iy)
                        for (int ix = 0; ix < nxi; ++ix) {
                            const long long o = (long long)ys[iy] * D.tx + xs[ix];
#pragma unroll
                            for (int v = 0; v < NV; ++v) out[o + v * vs] = res[v];
                        }
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    yc[v] = yp[v];
                    yp[v] = qn[v];
                }
            } else {
                if (interior) {
                    const long long i0 = px;
                    T u0[NV], res[NV];
                    if (a.mode != MODE_RHS) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) u0[v] = in2[i0 + v * vs];
                    }
                    combine<T, NV>(a.mode, c, out_v, yc, u0, res);
                    if (a.check && !eq.obeys(res)) atomicOr(&Sp->status, 2);
                    int xs[3];
                    const int nxi = axis_images(col, D.nx, D.gx, D.bc, xs);
                    for (int ix = 0; ix < nxi; ++ix) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) out[xs[ix] + v * vs] = res[v];
                    }
                }
            }
        }

        if (a.mode == MODE_RHS && !a.lazy) {
            const T mx = wave_max(wsx);
            const T my = wave_max(wsy);
            if (lane == 0) {
                atomicMax(a.ws + 3 * tk.sample + 0,
                          (unsigned long long)__double_as_longlong(double(mx)));
                if (DIM >= 2)
                    atomicMax(a.ws + 3 * tk.sample + 1,
                              (unsigned long long)__double_as_longlong(double(my)));
            }
        }
    }
}

}  // namespace alsvinn_hip

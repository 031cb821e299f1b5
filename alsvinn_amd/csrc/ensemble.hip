// Host side of the hip platform's C ABI: the ensemble object (a batch of
// solves of one scheme), its buffers, the per-step launch sequence and the
// captured HIP graph of the timestep loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <climits>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/alsvinn_hip.h"
#include "ensemble_kernels.hpp"
#include "internal.hpp"

using namespace alsvinn_hip;

namespace alsvinn_hip {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

// Pool slot (re)assignment in stream order: descriptor, fresh state and
// cleared wave-speed slots, passed by value.
__global__ void slot_update_kernel(SampleDesc* desc, SampleState* state, unsigned long long* ws,
                                   int slot, SampleDesc D, SampleState S) {
    if (threadIdx.x != 0) return;
    desc[slot] = D;
    state[slot] = S;
    ws[3 * slot + 0] = 0;
    ws[3 * slot + 1] = 0;
    ws[3 * slot + 2] = 0;
}

// Each listed sample's maximum wave speeds from the partial maxima its
// tasks stored in the first stage (StageArgs::ws_part; bit patterns of
// non-negative doubles, whose unsigned maximum is the maximum of the
// values): workgroup i reduces the entries idx[off[i] .. off[i+1]) of sample
// samples[i] and stores its three slots (the same maximum the atomics
// would have formed, in any order).
__global__ __launch_bounds__(1024) void ws_reduce_kernel(const unsigned long long* __restrict__ part,
                                                         const int* __restrict__ off,
                                                         const int* __restrict__ idx,
                                                         const int* __restrict__ samples,
                                                         unsigned long long* ws) {
    __shared__ unsigned long long red[3][16];
    const int i = blockIdx.x;
    unsigned long long m[3] = {0, 0, 0};
    for (int j = off[i] + (int)threadIdx.x; j < off[i + 1]; j += (int)blockDim.x) {
        const long long e = idx[j];
#pragma unroll
        for (int d = 0; d < 3; ++d) m[d] = part[e * 3 + d] > m[d] ? part[e * 3 + d] : m[d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long x = __shfl_xor(m[d], o, 64);
            m[d] = x > m[d] ? x : m[d];
        }
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 3; ++d) red[d][wv] = m[d];
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned long long r = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r = red[threadIdx.x][k] > r ? red[threadIdx.x][k] : r;
        ws[3 * (long long)samples[i] + threadIdx.x] = r;
    }
}

__global__ void set_hold_kernel(SampleState* state, int nsamples, int k, long long limit) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsamples) state[s].hold[k] = limit;
}

__global__ void count_active_kernel(const SampleState* __restrict__ state, int nsamples,
                                    int* count) {
    __shared__ int acc;
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    int local = 0;
    for (int s = threadIdx.x; s < nsamples; s += blockDim.x) local += runnable(state[s]) ? 1 : 0;
    atomicAdd(&acc, local);
    __syncthreads();
    if (threadIdx.x == 0) *count = acc;
}


}  // namespace alsvinn_hip

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return fail(ALSVINN_HIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)


namespace alsvinn_hip {

int validate(const alsvinn_hip_scheme* s) {
    if (!s) return fail(ALSVINN_HIP_E_INVALID, "scheme is NULL");
    if (s->equation < 0 || s->equation > 3)
        return fail(ALSVINN_HIP_E_INVALID, "unknown equation " + std::to_string(s->equation));
    if (s->reconstruction < 0 || s->reconstruction > 8)
        return fail(ALSVINN_HIP_E_INVALID, "unknown reconstruction " + std::to_string(s->reconstruction));
    if (s->integrator < 0 || s->integrator > 3)
        return fail(ALSVINN_HIP_E_INVALID, "unknown integrator " + std::to_string(s->integrator));
    if (s->boundary < 0 || s->boundary > 1)
        return fail(ALSVINN_HIP_E_INVALID, "unknown boundary " + std::to_string(s->boundary));
    if (s->precision < 0 || s->precision > 1)
        return fail(ALSVINN_HIP_E_INVALID, "unknown precision " + std::to_string(s->precision));
    if (s->equation == ALSVINN_HIP_EQ_BURGERS && s->flux != ALSVINN_HIP_FLUX_GODUNOV &&
        s->flux != ALSVINN_HIP_FLUX_CENTRAL && s->flux != ALSVINN_HIP_FLUX_RUSANOV &&
        s->flux != ALSVINN_HIP_FLUX_ROE)
        return fail(ALSVINN_HIP_E_INVALID, "burgers supports the godunov, central, rusanov and roe fluxes");
    if (s->equation == ALSVINN_HIP_EQ_BURGERS && s->reconstruction == ALSVINN_HIP_RECON_WENOF2)
        return fail(ALSVINN_HIP_E_INVALID, "wenof2 needs the Euler equations (as in the reference)");
    if (s->equation != ALSVINN_HIP_EQ_BURGERS && s->flux != ALSVINN_HIP_FLUX_HLL &&
        s->flux != ALSVINN_HIP_FLUX_HLL3 && s->flux != ALSVINN_HIP_FLUX_CENTRAL)
        return fail(ALSVINN_HIP_E_INVALID, "euler supports the hll, hll3 and central fluxes");
    return 0;
}

int ghost_cells(const alsvinn_hip_scheme& s) {
    // NumericalFluxCPU.cpp:226-231: max(flux stencil - 1, reconstruction ghosts)
    if (s.reconstruction == ALSVINN_HIP_RECON_NONE) return 1;
    if (s.reconstruction == ALSVINN_HIP_RECON_ENO3) return 3;
    if (s.reconstruction == ALSVINN_HIP_RECON_ENO4) return 4;
    if (s.reconstruction == ALSVINN_HIP_RECON_WENO3) return 3;
    return 2;
}

}  // namespace alsvinn_hip

namespace {

template <class T>
void* buf_at(void* base, long long off) {
    return (void*)((T*)base + off);
}

SampleDesc make_desc(const alsvinn_hip_scheme& sc, int G, const alsvinn_hip_sample& p,
                     long long base) {
    SampleDesc D;
    std::memset(&D, 0, sizeof(D));
    D.nx = p.nx;
    D.ny = p.ny;
    D.nz = p.nz;
    D.gx = G;
    D.gy = p.ny > 1 ? G : 0;
    D.gz = p.nz > 1 ? G : 0;
    D.tx = p.nx + 2 * D.gx;
    D.ty = p.ny + 2 * D.gy;
    D.tz = p.nz + 2 * D.gz;
    D.vstride = (long long)D.tx * D.ty * D.tz;
    D.base = base;
    D.bc = sc.boundary;
    // Grid.cpp:24-41: (top - origin) / N, inactive axes 1
    const int n[3] = {p.nx, p.ny, p.nz};
    for (int k = 0; k < 3; ++k) D.dx[k] = (p.top[k] - p.origin[k]) / double(n[k]);
    if (p.ny == 1) D.dx[1] = 1;
    if (p.nz == 1) D.dx[2] = 1;
    D.cfl = sc.cfl;
    D.end_time = p.end_time;
    D.nadj = p.num_adjusters;
    D.interval[0] = p.interval[0];
    D.interval[1] = p.interval[1];
    D.max_steps = p.max_steps;
    return D;
}

SampleState fresh_state(const alsvinn_hip_sample& p) {
    SampleState S;
    std::memset(&S, 0, sizeof(S));
    S.saved[0] = p.saved0[0];
    S.saved[1] = p.saved0[1];
    S.active = (0 < p.end_time && p.max_steps != 0) ? 1 : 0;
    S.hold[0] = S.hold[1] = LLONG_MAX;
    return S;
}

// Band height (rows per wave task) for a sample of ny rows: whole columns
// unless the ensemble would then launch fewer than ~4 waves per SIMD.
int band_rows(long long tasks_per_band, int ny) {
    const long long target_waves = 4096;
    int ty = ny;
    while (ty > 8 && tasks_per_band * ((ny + ty - 1) / ty) < target_waves) ty = (ty + 1) / 2;
    return ty;
}

// Wave tasks: every sample is cut into bands of rows and strips of W = 64-2G
// interior columns. A full strip fills a wave; the narrower last strips of
// samples with the same rows are packed together, up to kMaxSeg per wave.
void build_tasks(alsvinn_hip_ensemble* e, std::vector<WaveTask>& tasks) {
    const int G = e->ghosts;
    const int W = 64 - 2 * G;
    long long waves = 0;
    for (int s = 0; s < e->nsamples; ++s)
        if (e->assigned[s])
            waves += (long long)((e->desc[s].nx + W - 1) / W) * (e->dim == 3 ? e->desc[s].ny : 1);
    // (ny, nz, y0, y1, z0, z1) -> index of the task still open for packing
    std::map<std::tuple<int, int, int, int, int, int>, size_t> open;
    auto add_strips = [&](int s, int y0, int y1, int z0, int z1) {
        const SampleDesc& D = e->desc[s];
        int x0 = 0;
        for (; x0 + W <= D.nx; x0 += W) {
            WaveTask t;
            std::memset(&t, 0, sizeof(t));
            t.nseg = 1;
            t.y0 = y0;
            t.y1 = y1;
            t.z0 = z0;
            t.z1 = z1;
            t.sample[0] = s;
            t.x0[0] = x0;
            t.ncols[0] = W;
            tasks.push_back(t);
        }
        if (x0 < D.nx) {
            const int ncols = D.nx - x0;
            const int lanes = ncols + 2 * G;
            const auto key = std::make_tuple(D.ny, D.nz, y0, y1, z0, z1);
            auto it = open.find(key);
            if (it != open.end()) {
                WaveTask& t = tasks[it->second];
                const int used = t.lane0[t.nseg - 1] + t.ncols[t.nseg - 1] + 2 * G;
                if (t.nseg < kMaxSeg && used + lanes <= 64) {
                    t.sample[t.nseg] = s;
                    t.x0[t.nseg] = x0;
                    t.ncols[t.nseg] = ncols;
                    t.lane0[t.nseg] = used;
                    t.nseg++;
                    return;
                }
            }
            WaveTask t;
            std::memset(&t, 0, sizeof(t));
            t.nseg = 1;
            t.y0 = y0;
            t.y1 = y1;
            t.z0 = z0;
            t.z1 = z1;
            t.sample[0] = s;
            t.x0[0] = x0;
            t.ncols[0] = ncols;
            open[key] = tasks.size();
            tasks.push_back(t);
        }
    };
    for (int s = 0; s < e->nsamples; ++s) {
        if (!e->assigned[s]) continue;
        const SampleDesc& D = e->desc[s];
        if (e->dim == 3) {
            // kGroupRows rows per task (one workgroup), bands of planes marched in z
            const int tz = e->band_rows > 0 ? std::min(e->band_rows, D.nz) : band_rows(waves, D.nz);
            for (int z0 = 0; z0 < D.nz; z0 += tz)
                for (int y = 0; y < D.ny; y += kGroupRows)
                    add_strips(s, y, std::min(D.ny, y + kGroupRows), z0, std::min(D.nz, z0 + tz));
        } else {
            const int ty = e->dim < 2 ? 1
                           : (e->band_rows > 0 ? std::min(e->band_rows, D.ny) : band_rows(waves, D.ny));
            for (int y0 = 0; y0 < D.ny; y0 += ty) add_strips(s, y0, std::min(D.ny, y0 + ty), 0, 1);
        }
    }
}

// Optional per-kernel timing: a pair of HIP events around every launch on
// the ensemble's stream (no host synchronisation between launches); the
// durations are summed once the stream has drained.
struct KernelTimer {
    std::vector<hipEvent_t> ev;  // created before any launch is enqueued
    std::vector<int> kind;
    size_t next = 0;
    hipStream_t s = nullptr;
    void reserve(size_t launches) {
        ev.resize(2 * launches);
        for (auto& x : ev) hipEventCreate(&x);
    }
    void begin() {
        if (next + 1 < ev.size()) hipEventRecord(ev[next], s);
    }
    void end(int k) {
        if (next + 1 < ev.size()) {
            hipEventRecord(ev[next + 1], s);
            kind.push_back(k);
            next += 2;
        }
    }
    void collect(double ms[3], long long count[3]) {
        hipStreamSynchronize(s);
        for (int k = 0; k < 3; ++k) {
            ms[k] = 0;
            count[k] = 0;
        }
        for (size_t i = 0; i < kind.size(); ++i) {
            float f = 0;
            hipEventElapsedTime(&f, ev[2 * i], ev[2 * i + 1]);
            ms[kind[i]] += f;
            count[kind[i]] += 1;
        }
        for (auto x : ev) hipEventDestroy(x);
        ev.clear();
        kind.clear();
    }
};

// Work list of a stage launch: an ensemble launches one wave per task; a
// pool launches its fixed grid, which reads the current number of tasks
// from the device (a.ntasks is then the 3D kernel's grid size).
template <class T>
int set_tasks(const alsvinn_hip_ensemble* e, StageArgs<T>& a) {
    a.tasks = e->d_tasks;
    a.ws_filter = e->ws_filter;
    if (e->pool) {
        a.ntasks_dev = (const int*)e->d_taskbuf;
        a.ntasks = e->pool_blocks;
        return e->pool_blocks;
    }
    a.ntasks_dev = nullptr;
    a.ntasks = e->ntasks;
    return (e->ntasks + 3) / 4;
}

void launch_ws_reduce(const alsvinn_hip_ensemble* e);

template <class T>
int enqueue_step(alsvinn_hip_ensemble* e, const KernelSet<T>& k, KernelTimer* tm = nullptr) {
    StageArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.desc = e->d_desc;
    a.state = e->d_state;
    a.ws = e->d_ws;
    a.gamma = e->scheme.gamma;
    const int blocks = set_tasks(e, a);
    T* U = (T*)e->U;
    T* A = (T*)e->A;
    T* B = (T*)e->B;
    T* C = (T*)e->C;
    T* Dv = (T*)e->D;
    const int bx = (int)std::max<long long>(1, std::min<long long>(64, (e->max_vstride + 1023) / 1024));
    const dim3 pgrid(bx, e->nsamples);
    // the step's last kernel ends the step (t += dt, counters, stop flags,
    // wave-speed reset) once all its workgroups are done
    StepEnd end{};
    end.counter = e->d_end_counter;
    end.desc = e->d_desc;
    end.state = e->d_state;
    end.ws = e->d_ws;
    end.dt_hist = e->d_dt_hist;
    end.dt_cap = e->dt_cap;
    end.nsamples = e->nsamples;
    const StepEnd none{};

    // stage 1: A = L(U), wave speeds
    a.mode = MODE_RHS;
    a.check = 0;
    a.in = U;
    a.in2 = U;
    a.out = A;
    a.ws_part = e->d_ws_part;
    auto stage = [&]() {
        if (tm) tm->begin();
        k.stage[e->variant[stage_kind(a.mode)]](a, blocks, e->stream);
        if (tm) tm->end(0);
    };
    auto axpy = [&](const T* X, const T* Y, T* O, int check, const StepEnd& en) {
        if (tm) tm->begin();
        k.axpy(pgrid, e->stream, e->d_desc, e->d_state, e->d_ws, X, Y, O, check, e->scheme.gamma,
               en);
        if (tm) tm->end(1);
    };
    stage();
    launch_ws_reduce(e);
    if (e->scheme.integrator == ALSVINN_HIP_INTEGRATOR_FORWARD_EULER) {
        axpy(A, U, U, 1, end);
    } else if (e->scheme.integrator == ALSVINN_HIP_INTEGRATOR_RK2) {
        axpy(A, U, A, 0, none);
        a.mode = MODE_RK2_1;
        a.in = A;
        a.in2 = U;
        a.out = U;
        a.check = 1;
        a.end = end;
        stage();
    } else if (e->scheme.integrator == ALSVINN_HIP_INTEGRATOR_RK4) {
        // RungeKutta4.cpp:50-96: U1 = ((L0*c)*0.5) + U0, U2 = ((L1*c)*0.5) + U0,
        // U3 = L2*c + U0, U' = (c/6)L3 + U1/3 + 2U2/3 + U3/3 - U0/3; only the
        // first evaluation feeds the CFL condition
        auto rk4 = [&](int kind, const T* X, const T* Z, const T* W, const T* V, T* O, int check,
                       const StepEnd& en) {
            if (tm) tm->begin();
            k.rk4(pgrid, e->stream, e->d_desc, e->d_state, e->d_ws, kind, X, U, Z, W, V, O, check,
                  e->scheme.gamma, en);
            if (tm) tm->end(1);
        };
        rk4(0, A, nullptr, nullptr, nullptr, B, 0, none);  // U1
        a.no_ws = 1;
        a.in = B;
        a.in2 = B;
        a.out = C;
        stage();                                      // L1
        rk4(0, C, nullptr, nullptr, nullptr, C, 0, none);  // U2
        a.in = C;
        a.in2 = C;
        a.out = Dv;
        stage();                                      // L2
        axpy(Dv, U, Dv, 0, none);                           // U3
        a.in = Dv;
        a.in2 = Dv;
        a.out = A;
        stage();                                      // L3
        rk4(1, A, B, C, Dv, U, 1, end);                    // U^{n+1}
    } else {
        a.mode = MODE_RK3_1;
        a.in = A;
        a.in2 = U;
        a.out = B;
        stage();
        a.mode = MODE_RK3_2;
        a.in = B;
        a.in2 = U;
        a.out = U;
        a.check = 1;
        a.end = end;
        stage();
    }
    return 0;
}

// Whether the ensemble's steps run as whole-step launches of
// solve_kernel_1d (sets solve1d_threads: 0, or 64 per wavefront of a
// sample's workgroup, one wavefront per strip up to 16).
void update_solve1d(alsvinn_hip_ensemble* e) {
    e->solve1d_threads = 0;
    const bool has = e->esize == 8 ? e->k64.solve1d != nullptr : e->k32.solve1d != nullptr;
    if (!has || e->dim != 1 || e->pool || e->solve1d_rounds <= 0 || e->nsamples > kSolve1DSamples ||
        e->scheme.integrator == ALSVINN_HIP_INTEGRATOR_RK4)
        return;
    const int W = 64 - 2 * e->ghosts;
    long long strips = 1;
    for (const SampleDesc& D : e->desc) strips = std::max<long long>(strips, (D.nx + W - 1) / W);
    if (strips > 16LL * e->solve1d_rounds) return;
    e->solve1d_threads = 64 * (int)std::min<long long>(16, strips);
}

template <class T>
void enqueue_solve1d(alsvinn_hip_ensemble* e, const KernelSet<T>& k, int nsteps) {
    Solve1DArgs<T> p;
    std::memset(&p, 0, sizeof(p));
    p.a.desc = e->d_desc;
    p.a.state = e->d_state;
    p.a.ws = e->d_ws;
    p.a.gamma = e->scheme.gamma;
    p.U = (T*)e->U;
    p.A = (T*)e->A;
    p.B = (T*)e->B;
    p.integrator = e->scheme.integrator;
    p.nsteps = nsteps;
    p.dt_hist = e->d_dt_hist;
    p.dt_cap = e->dt_cap;
    k.solve1d(p, e->nsamples, e->solve1d_threads, e->stream);
}

int enqueue_steps(alsvinn_hip_ensemble* e, int nsteps, KernelTimer* tm = nullptr) {
    if (e->solve1d_threads > 0 && !tm) {  // per-kernel timing keeps the multi-launch step
        if (nsteps > 0) {
            if (e->esize == 8)
                enqueue_solve1d<double>(e, e->k64, nsteps);
            else
                enqueue_solve1d<float>(e, e->k32, nsteps);
        }
        HIP_TRY(hipGetLastError());
        return 0;
    }
    for (int i = 0; i < nsteps; ++i) {
        if (e->esize == 8)
            enqueue_step<double>(e, e->k64, tm);
        else
            enqueue_step<float>(e, e->k32, tm);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int init_state(alsvinn_hip_ensemble* e) {
    std::vector<SampleState> st(e->nsamples);
    for (int s = 0; s < e->nsamples; ++s) {
        st[s] = fresh_state(e->spec[s]);
        if (!e->assigned[s]) st[s].active = 0;
    }
    HIP_TRY(hipMemcpy(e->d_state, st.data(), sizeof(SampleState) * st.size(), hipMemcpyHostToDevice));
    std::vector<unsigned long long> z(3 * (size_t)e->nsamples, 0ull);
    HIP_TRY(hipMemcpy(e->d_ws, z.data(), sizeof(unsigned long long) * z.size(), hipMemcpyHostToDevice));
    return 0;
}

void destroy_graph(alsvinn_hip_ensemble* e) {
    if (e->graph_exec) hipGraphExecDestroy(e->graph_exec);
    if (e->graph) hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_steps = 0;
}

// Device work list of a pool, capacity `cap` tasks: an int count (16 bytes
// with padding) followed by the tasks, and its pinned staging buffers.
int alloc_pool_tasks(alsvinn_hip_ensemble* e, long long cap) {
    for (int j = 0; j < alsvinn_hip_ensemble::kRing; ++j) {
        if (e->stage_ev[j]) HIP_TRY(hipEventSynchronize(e->stage_ev[j]));
        if (e->h_stage[j]) (void)hipHostFree(e->h_stage[j]);
        e->h_stage[j] = nullptr;
    }
    (void)hipFree(e->d_taskbuf);
    e->d_taskbuf = nullptr;
    e->d_tasks = nullptr;
    const size_t bytes = 16 + sizeof(WaveTask) * (size_t)std::max<long long>(1, cap);
    HIP_TRY(hipMalloc(&e->d_taskbuf, bytes));
    HIP_TRY(hipMemset(e->d_taskbuf, 0, 16));
    for (int j = 0; j < alsvinn_hip_ensemble::kRing; ++j)
        HIP_TRY(hipHostMalloc(&e->h_stage[j], bytes));
    e->d_tasks = (WaveTask*)(e->d_taskbuf + 16);
    e->task_cap = cap;
    e->pool_blocks = (int)std::max<long long>(
        1, std::min<long long>(1024, e->dim == 3 ? cap : (cap + 3) / 4));
    return 0;
}

void free_ws_part(alsvinn_hip_ensemble* e) {
    (void)hipFree(e->d_ws_part);
    (void)hipFree(e->d_ws_off);
    (void)hipFree(e->d_ws_idx);
    (void)hipFree(e->d_ws_samples);
    e->d_ws_part = nullptr;
    e->d_ws_off = e->d_ws_idx = e->d_ws_samples = nullptr;
    e->ws_nlist = 0;
}

// Ensembles with large samples: the partial-maxima buffer of the first
// stage and each sample's list of its entries (task, row in 3D, segment).
int build_ws_part(alsvinn_hip_ensemble* e, const std::vector<WaveTask>& tasks) {
    const int R = e->dim == 3 ? kGroupRows : 1;
    std::vector<std::vector<int>> lists(e->nsamples);
    for (size_t t = 0; t < tasks.size(); ++t)
        for (int r = 0; r < R; ++r)
            for (int k = 0; k < tasks[t].nseg; ++k)
                lists[tasks[t].sample[k]].push_back(int((t * R + r) * kMaxSeg + k));
    std::vector<int> off(1, 0), idx, smp;
    for (int s = 0; s < e->nsamples; ++s) {
        if (lists[s].empty()) continue;
        smp.push_back(s);
        idx.insert(idx.end(), lists[s].begin(), lists[s].end());
        off.push_back((int)idx.size());
    }
    if (smp.empty()) return 0;
    const size_t entries = tasks.size() * R * kMaxSeg * 3;
    HIP_TRY(hipMalloc(&e->d_ws_part, sizeof(unsigned long long) * entries));
    HIP_TRY(hipMemset(e->d_ws_part, 0, sizeof(unsigned long long) * entries));
    HIP_TRY(hipMalloc(&e->d_ws_off, sizeof(int) * off.size()));
    HIP_TRY(hipMalloc(&e->d_ws_idx, sizeof(int) * idx.size()));
    HIP_TRY(hipMalloc(&e->d_ws_samples, sizeof(int) * smp.size()));
    HIP_TRY(hipMemcpy(e->d_ws_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_ws_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_ws_samples, smp.data(), sizeof(int) * smp.size(), hipMemcpyHostToDevice));
    e->ws_nlist = (int)smp.size();
    return 0;
}

// After the first stage of a step with partial maxima: each sample's slots.
void launch_ws_reduce(const alsvinn_hip_ensemble* e) {
    if (!e->d_ws_part || e->ws_nlist == 0) return;
    hipLaunchKernelGGL(ws_reduce_kernel, dim3(e->ws_nlist), dim3(1024), 0, e->stream, e->d_ws_part,
                       e->d_ws_off, e->d_ws_idx, e->d_ws_samples, e->d_ws);
}

int upload_tasks(alsvinn_hip_ensemble* e) {
    std::vector<WaveTask> tasks;
    build_tasks(e, tasks);
    if (e->pool) {
        // asynchronous, in stream order: steps enqueued before see the old
        // list, steps enqueued after the new one
        if ((long long)tasks.size() > e->task_cap) {  // a schedule change outgrew the buffer
            HIP_TRY(hipStreamSynchronize(e->stream));
            destroy_graph(e);
            if (int rc = alloc_pool_tasks(e, 2 * (long long)tasks.size())) return rc;
        }
        const int j = e->stage_next;
        e->stage_next = (j + 1) % alsvinn_hip_ensemble::kRing;
        HIP_TRY(hipEventSynchronize(e->stage_ev[j]));  // its previous copy has run
        const int count = (int)tasks.size();
        std::memcpy(e->h_stage[j], &count, sizeof(int));
        if (count) std::memcpy(e->h_stage[j] + 16, tasks.data(), sizeof(WaveTask) * tasks.size());
        HIP_TRY(hipMemcpyAsync(e->d_taskbuf, e->h_stage[j], 16 + sizeof(WaveTask) * tasks.size(),
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipEventRecord(e->stage_ev[j], e->stream));
        e->ntasks = count;
        return 0;
    }
    hipFree(e->d_tasks);
    e->d_tasks = nullptr;
    e->ntasks = (int)tasks.size();
    // wavefronts per sample: many on one sample filter their wave-speed atomics
    std::vector<int> per(e->nsamples, 0);
    for (const auto& t : tasks)
        for (int k = 0; k < t.nseg; ++k) per[t.sample[k]] += 1;
    const int most = per.empty() ? 0 : *std::max_element(per.begin(), per.end());
    e->ws_filter = most * (e->dim == 3 ? kGroupRows : 1) > kWsFilterWaves ? 1 : 0;
    // large samples reduce their wave speeds through partial maxima and a
    // reduction kernel instead of thousands of same-address atomics
    free_ws_part(e);
    if (e->ws_filter)
        if (int rc = build_ws_part(e, tasks)) return rc;
    if (tasks.empty()) return 0;
    HIP_TRY(hipMalloc(&e->d_tasks, sizeof(WaveTask) * tasks.size()));
    HIP_TRY(hipMemcpy(e->d_tasks, tasks.data(), sizeof(WaveTask) * tasks.size(), hipMemcpyHostToDevice));
    return 0;
}

// Times `reps` first-stage launches (reconstruct + flux + divergence +
// wave speeds on the current state; writes only the scratch buffer and the
// wave-speed slots, which are cleared afterwards).
template <class T>
void time_stages(alsvinn_hip_ensemble* e, const KernelSet<T>& k, int reps, double ms[3]) {
    // A dry run of the step's stage kernels on the live state: every output
    // goes to the scratch volumes A and B, the constraint check is off and the
    // wave-speed slots are cleared by the caller afterwards, so U and the
    // samples' state are untouched. SSP-RK3 times each of its three stages
    // (the second and third read two volumes and use more registers than the
    // first); the other integrators time the first stage. ms[kind] receives
    // the average time of each stage kind (0 where not timed).
    StageArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.desc = e->d_desc;
    // the samples' states are copied for the dry run: the last stage's
    // constraint check may flag a sample on the scratch volumes' values
    (void)hipMemcpyAsync(e->d_state_dry, e->d_state, sizeof(SampleState) * e->nsamples,
                         hipMemcpyDeviceToDevice, e->stream);
    a.state = e->d_state_dry;
    a.ws = e->d_ws;
    a.gamma = e->scheme.gamma;
    const int blocks = set_tasks(e, a);
    T* U = (T*)e->U;
    T* A = (T*)e->A;
    T* B = (T*)e->B;
    const bool rk3 = e->scheme.integrator == ALSVINN_HIP_INTEGRATOR_RK3;
    hipEvent_t ev[4];
    for (auto& x : ev) hipEventCreate(&x);
    auto launch_all = [&](bool timed) {
        if (timed) hipEventRecord(ev[0], e->stream);
        a.mode = MODE_RHS;
        a.in = U;
        a.in2 = U;
        a.out = A;
        a.ws_part = e->d_ws_part;
        k.stage[e->variant[0]](a, blocks, e->stream);
        launch_ws_reduce(e);
        if (timed) hipEventRecord(ev[1], e->stream);
        if (!rk3) return;
        a.mode = MODE_RK3_1;
        a.in = A;
        a.in2 = U;
        a.out = B;
        k.stage[e->variant[1]](a, blocks, e->stream);
        if (timed) hipEventRecord(ev[2], e->stream);
        a.mode = MODE_RK3_2;
        a.in = B;
        a.in2 = U;
        a.out = A;
        k.stage[e->variant[2]](a, blocks, e->stream);
        if (timed) hipEventRecord(ev[3], e->stream);
    };
    launch_all(false);  // warm-up
    ms[0] = ms[1] = ms[2] = 0;
    for (int i = 0; i < reps; ++i) {
        launch_all(true);
        hipEventSynchronize(rk3 ? ev[3] : ev[1]);
        for (int s = 0; s < (rk3 ? 3 : 1); ++s) {
            float t = 0;
            hipEventElapsedTime(&t, ev[s], ev[s + 1]);
            ms[s] += t / reps;
        }
    }
    for (auto x : ev) hipEventDestroy(x);
}

}  // namespace

static int finish_create(alsvinn_hip_ensemble* e, const alsvinn_hip_scheme* scheme, void* stream,
                         alsvinn_hip_ensemble** out) {
    const int nsamples = e->nsamples;
    int rc = 0;
    if (e->esize == 8)
        rc = select_kernels<double>(*scheme, e->dim, e->k64);
    else
        rc = select_kernels<float>(*scheme, e->dim, e->k32);
    if (rc) {
        delete e;
        return rc;
    }
    auto cleanup = [&](int code) {
        alsvinn_hip_ensemble_destroy(e);
        return code;
    };
#define HIP_TRY_C(expr)                                                                       \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return cleanup(fail(ALSVINN_HIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e))); \
    } while (0)
    if (stream) {
        e->stream = (hipStream_t)stream;
    } else {
        HIP_TRY_C(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        e->own_stream = true;
    }
    const size_t bytes = (size_t)e->total * e->esize;
    if (e->own_U) HIP_TRY_C(hipMalloc(&e->U, bytes));
    HIP_TRY_C(hipMalloc(&e->A, bytes));
    const bool rk4 = scheme->integrator == ALSVINN_HIP_INTEGRATOR_RK4;
    if (scheme->integrator == ALSVINN_HIP_INTEGRATOR_RK3 || rk4) HIP_TRY_C(hipMalloc(&e->B, bytes));
    if (rk4) HIP_TRY_C(hipMalloc(&e->C, bytes));
    if (rk4) HIP_TRY_C(hipMalloc(&e->D, bytes));
    if (e->own_U) HIP_TRY_C(hipMemset(e->U, 0, bytes));
    HIP_TRY_C(hipMemset(e->A, 0, bytes));
    if (e->B) HIP_TRY_C(hipMemset(e->B, 0, bytes));
    if (e->C) HIP_TRY_C(hipMemset(e->C, 0, bytes));
    if (e->D) HIP_TRY_C(hipMemset(e->D, 0, bytes));
    HIP_TRY_C(hipMalloc(&e->d_desc, sizeof(SampleDesc) * nsamples));
    HIP_TRY_C(hipMalloc(&e->d_state, sizeof(SampleState) * nsamples));
    HIP_TRY_C(hipMalloc(&e->d_state_dry, sizeof(SampleState) * nsamples));
    HIP_TRY_C(hipMalloc(&e->d_ws, sizeof(unsigned long long) * 3 * nsamples));
    HIP_TRY_C(hipMalloc(&e->d_count, sizeof(int)));
    HIP_TRY_C(hipMalloc(&e->d_end_counter, sizeof(unsigned int)));
    HIP_TRY_C(hipMemset(e->d_end_counter, 0, sizeof(unsigned int)));
    HIP_TRY_C(hipHostMalloc(&e->h_count, sizeof(int)));
    HIP_TRY_C(hipMemcpy(e->d_desc, e->desc.data(), sizeof(SampleDesc) * nsamples, hipMemcpyHostToDevice));
    if (e->pool) {
        for (int j = 0; j < alsvinn_hip_ensemble::kRing; ++j) {
            HIP_TRY_C(hipEventCreateWithFlags(&e->stage_ev[j], hipEventDisableTiming));
            HIP_TRY_C(hipEventCreateWithFlags(&e->snap_ev[j], hipEventDisableTiming));
            HIP_TRY_C(hipHostMalloc(&e->h_snap[j], sizeof(SampleState) * nsamples));
        }
        if (int r = alloc_pool_tasks(e, e->task_cap)) return cleanup(r);
        e->ntasks = 0;
    } else {
        std::vector<WaveTask> tasks;
        build_tasks(e, tasks);
        e->ntasks = (int)tasks.size();
        if (!tasks.empty()) {
            HIP_TRY_C(hipMalloc(&e->d_tasks, sizeof(WaveTask) * tasks.size()));
            HIP_TRY_C(hipMemcpy(e->d_tasks, tasks.data(), sizeof(WaveTask) * tasks.size(),
                                hipMemcpyHostToDevice));
        }
    }
    HIP_TRY_C(hipEventCreate(&e->ev0));
    HIP_TRY_C(hipEventCreate(&e->ev1));
#undef HIP_TRY_C
    update_solve1d(e);
    if (int r = init_state(e)) return cleanup(r);
    *out = e;
    return 0;
}


extern "C" {

const char* alsvinn_hip_last_error(void) { return g_last_error.c_str(); }

int alsvinn_hip_ghost_cells(const alsvinn_hip_scheme* scheme) {
    if (validate(scheme)) return -1;
    return ghost_cells(*scheme);
}

// Descriptors of an ensemble's samples (layout of the concatenated volumes);
// the caller owns the returned object until finish_create takes it.
static int describe_samples(const alsvinn_hip_scheme* scheme, int nsamples,
                            const alsvinn_hip_sample* samples, alsvinn_hip_ensemble** out) {
    if (int rc = validate(scheme)) return rc;
    if (!samples || nsamples < 1) return fail(ALSVINN_HIP_E_INVALID, "bad ensemble arguments");
    auto* e = new alsvinn_hip_ensemble();
    e->scheme = *scheme;
    e->nsamples = nsamples;
    e->nvar = scheme->equation == ALSVINN_HIP_EQ_BURGERS ? 1 : scheme->equation + 2;
    e->ghosts = ghost_cells(*scheme);
    e->esize = scheme->precision == ALSVINN_HIP_FP64 ? 8 : 4;
    const int G = e->ghosts;
    long long off = 0;
    for (int s = 0; s < nsamples; ++s) {
        const alsvinn_hip_sample& p = samples[s];
        const int dim = 1 + (p.ny > 1) + (p.nz > 1);
        if (s == 0) e->dim = dim;
        if (dim != e->dim || p.nx < 1 || p.ny < 1 || p.nz < 1) {
            delete e;
            return fail(ALSVINN_HIP_E_INVALID, "all samples of an ensemble must have the same dimension");
        }
        if (p.nx < G || (p.ny > 1 && p.ny < G) || (p.nz > 1 && p.nz < G)) {
            delete e;
            return fail(ALSVINN_HIP_E_INVALID, "grid smaller than the number of ghost cells");
        }
        if (p.num_adjusters < 0 || p.num_adjusters > 2) {
            delete e;
            return fail(ALSVINN_HIP_E_INVALID, "num_adjusters must be 0..2");
        }
        const SampleDesc D = make_desc(*scheme, G, p, off);
        off += D.vstride * e->nvar;
        e->max_vstride = std::max(e->max_vstride, D.vstride);
        e->desc.push_back(D);
        e->spec.push_back(p);
        e->assigned.push_back(1);
    }
    e->total = off;
    *out = e;
    return 0;
}

int alsvinn_hip_state_elements(const alsvinn_hip_scheme* scheme, int nsamples,
                               const alsvinn_hip_sample* samples, long long* elements) {
    if (!elements) return fail(ALSVINN_HIP_E_INVALID, "elements is NULL");
    alsvinn_hip_ensemble* e = nullptr;
    if (int rc = describe_samples(scheme, nsamples, samples, &e)) return rc;
    *elements = e->total;
    delete e;
    return 0;
}

int alsvinn_hip_ensemble_create(const alsvinn_hip_scheme* scheme, int nsamples,
                                const alsvinn_hip_sample* samples, void* stream,
                                alsvinn_hip_ensemble** out) {
    return alsvinn_hip_ensemble_create_bound(scheme, nsamples, samples, stream, nullptr, out);
}

int alsvinn_hip_ensemble_create_bound(const alsvinn_hip_scheme* scheme, int nsamples,
                                      const alsvinn_hip_sample* samples, void* stream, void* state,
                                      alsvinn_hip_ensemble** out) {
    if (!out) return fail(ALSVINN_HIP_E_INVALID, "bad ensemble arguments");
    *out = nullptr;
    alsvinn_hip_ensemble* e = nullptr;
    if (int rc = describe_samples(scheme, nsamples, samples, &e)) return rc;
    if (state) {
        e->U = state;
        e->own_U = false;
    }
    return finish_create(e, scheme, stream, out);
}

int alsvinn_hip_ensemble_destroy(alsvinn_hip_ensemble* e) {
    if (!e) return 0;
    if (e->stream) hipStreamSynchronize(e->stream);
    destroy_graph(e);
    if (e->own_U) hipFree(e->U);
    hipFree(e->A);
    hipFree(e->B);
    hipFree(e->C);
    hipFree(e->D);
    hipFree(e->d_desc);
    hipFree(e->d_state);
    hipFree(e->d_state_dry);
    hipFree(e->d_ws);
    free_ws_part(e);
    if (e->pool) {
        (void)hipFree(e->d_taskbuf);
        for (int j = 0; j < alsvinn_hip_ensemble::kRing; ++j) {
            if (e->h_stage[j]) (void)hipHostFree(e->h_stage[j]);
            if (e->h_snap[j]) (void)hipHostFree(e->h_snap[j]);
            if (e->stage_ev[j]) (void)hipEventDestroy(e->stage_ev[j]);
            if (e->snap_ev[j]) (void)hipEventDestroy(e->snap_ev[j]);
        }
    } else {
        hipFree(e->d_tasks);
    }
    hipFree(e->d_count);
    hipFree(e->d_end_counter);
    hipFree(e->d_dt_hist);
    if (e->h_count) hipHostFree(e->h_count);
    if (e->ev0) hipEventDestroy(e->ev0);
    if (e->ev1) hipEventDestroy(e->ev1);
    if (e->own_stream && e->stream) hipStreamDestroy(e->stream);
    delete e;
    return 0;
}

int alsvinn_hip_pool_create(const alsvinn_hip_scheme* scheme, int nslots, const int max_n[3],
                            long long arena_elements, void* stream, alsvinn_hip_ensemble** out) {
    if (int rc = validate(scheme)) return rc;
    if (!out || nslots < 1 || !max_n || max_n[0] < 1 || max_n[1] < 1 || max_n[2] < 1 ||
        (max_n[2] > 1 && max_n[1] < 2) || arena_elements < 1)
        return fail(ALSVINN_HIP_E_INVALID, "bad pool arguments");
    *out = nullptr;
    auto* e = new alsvinn_hip_ensemble();
    e->scheme = *scheme;
    e->nsamples = nslots;
    e->nvar = scheme->equation == ALSVINN_HIP_EQ_BURGERS ? 1 : scheme->equation + 2;
    e->ghosts = ghost_cells(*scheme);
    e->esize = scheme->precision == ALSVINN_HIP_FP64 ? 8 : 4;
    e->dim = 1 + (max_n[1] > 1) + (max_n[2] > 1);
    e->pool = true;
    for (int a = 0; a < 3; ++a) e->max_n[a] = max_n[a];
    // fixed band heights (rows per 2D task, planes per 3D task), so that the
    // work list of every slot is bounded by that of the largest grid
    e->band_rows = e->dim == 1 ? 1 : 16;
    const int G = e->ghosts, W = 64 - 2 * G;
    long long per_slot = (max_n[0] + W - 1) / W;
    if (e->dim == 2) per_slot *= (max_n[1] + e->band_rows - 1) / e->band_rows;
    if (e->dim == 3)
        per_slot *= (long long)((max_n[1] + kGroupRows - 1) / kGroupRows) *
                    ((max_n[2] + e->band_rows - 1) / e->band_rows);
    e->task_cap = per_slot * nslots;
    e->ws_filter = per_slot * (e->dim == 3 ? kGroupRows : 1) > kWsFilterWaves ? 1 : 0;
    // the elementwise kernels' grid is sized once, for a sample the size of
    // the arena's share of one variable
    e->max_vstride = std::max<long long>(1, arena_elements / e->nvar);
    alsvinn_hip_sample blank;
    std::memset(&blank, 0, sizeof(blank));
    SampleDesc D;
    std::memset(&D, 0, sizeof(D));
    for (int s = 0; s < nslots; ++s) {
        e->desc.push_back(D);
        e->spec.push_back(blank);
        e->assigned.push_back(0);
    }
    e->total = arena_elements;
    return finish_create(e, scheme, stream, out);
}

int alsvinn_hip_pool_assign(alsvinn_hip_ensemble* e, int slot, const alsvinn_hip_sample* p,
                            long long offset) {
    if (!e || !e->pool || !p || slot < 0 || slot >= e->nsamples)
        return fail(ALSVINN_HIP_E_INVALID, "bad slot");
    if (e->assigned[slot]) return fail(ALSVINN_HIP_E_INVALID, "slot is in use");
    const int G = e->ghosts;
    const int dim = 1 + (p->ny > 1) + (p->nz > 1);
    if (dim != e->dim || p->nx < G || (p->ny > 1 && p->ny < G) || (p->nz > 1 && p->nz < G))
        return fail(ALSVINN_HIP_E_INVALID, "grid does not fit the pool's dimension or ghost cells");
    if (p->nx > e->max_n[0] || p->ny > e->max_n[1] || p->nz > e->max_n[2])
        return fail(ALSVINN_HIP_E_INVALID, "grid larger than the pool's largest grid");
    if (p->num_adjusters < 0 || p->num_adjusters > 2)
        return fail(ALSVINN_HIP_E_INVALID, "num_adjusters must be 0..2");
    const SampleDesc D = make_desc(e->scheme, G, *p, offset);
    if (offset < 0 || offset + D.vstride * e->nvar > e->total)
        return fail(ALSVINN_HIP_E_INVALID, "sample does not fit in the pool arena at this offset");
    for (int s = 0; s < e->nsamples; ++s) {
        if (s == slot || !e->assigned[s]) continue;
        const long long a0 = e->desc[s].base, a1 = a0 + e->desc[s].vstride * e->nvar;
        if (offset < a1 && offset + D.vstride * e->nvar > a0)
            return fail(ALSVINN_HIP_E_INVALID, "arena range overlaps slot " + std::to_string(s));
    }
    e->desc[slot] = D;
    e->spec[slot] = *p;
    e->assigned[slot] = 1;
    // in stream order: steps already enqueued still see the slot unassigned
    hipLaunchKernelGGL(slot_update_kernel, dim3(1), dim3(64), 0, e->stream, e->d_desc, e->d_state,
                       e->d_ws, slot, D, fresh_state(*p));
    HIP_TRY(hipGetLastError());
    return upload_tasks(e);
}

int alsvinn_hip_pool_release(alsvinn_hip_ensemble* e, int slot) {
    if (!e || !e->pool || slot < 0 || slot >= e->nsamples) return fail(ALSVINN_HIP_E_INVALID, "bad slot");
    e->assigned[slot] = 0;
    SampleState S;
    std::memset(&S, 0, sizeof(S));
    SampleDesc D;
    std::memset(&D, 0, sizeof(D));
    e->desc[slot] = D;
    hipLaunchKernelGGL(slot_update_kernel, dim3(1), dim3(64), 0, e->stream, e->d_desc, e->d_state,
                       e->d_ws, slot, D, S);
    HIP_TRY(hipGetLastError());
    return upload_tasks(e);
}

int alsvinn_hip_pool_upload(alsvinn_hip_ensemble* e, int slot, const double* host) {
    if (!e || !e->pool || !host || slot < 0 || slot >= e->nsamples || !e->assigned[slot])
        return fail(ALSVINN_HIP_E_INVALID, "bad upload arguments");
    const SampleDesc& D = e->desc[slot];
    const size_t n = (size_t)D.vstride * e->nvar;
    // a synchronous copy that does not wait for the pool's stream: no step
    // enqueued before the slot's assignment reads or writes its memory
    if (e->esize == 8) {
        HIP_TRY(hipMemcpy(buf_at<double>(e->U, D.base), host, n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
        HIP_TRY(hipMemcpy(buf_at<float>(e->U, D.base), tmp.data(), n * 4, hipMemcpyHostToDevice));
    }
    return 0;
}

int alsvinn_hip_pool_download(alsvinn_hip_ensemble* e, int slot, double* host) {
    if (!e || !e->pool || !host || slot < 0 || slot >= e->nsamples || !e->assigned[slot])
        return fail(ALSVINN_HIP_E_INVALID, "bad download arguments");
    const SampleDesc& D = e->desc[slot];
    const size_t n = (size_t)D.vstride * e->nvar;
    // does not wait for the pool's stream: the slot has finished (a snapshot
    // showed it inactive), and no step writes an inactive slot
    if (e->esize == 8) {
        HIP_TRY(hipMemcpy(host, buf_at<double>(e->U, D.base), n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        HIP_TRY(hipMemcpy(tmp.data(), buf_at<float>(e->U, D.base), n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) host[i] = tmp[i];
    }
    return 0;
}

int alsvinn_hip_pool_apply_boundary(alsvinn_hip_ensemble* e, int slot) {
    if (!e || !e->pool || slot < 0 || slot >= e->nsamples || !e->assigned[slot])
        return fail(ALSVINN_HIP_E_INVALID, "bad slot");
    const int bx = (int)std::max<long long>(1, std::min<long long>(64, (e->desc[slot].vstride + 1023) / 1024));
    if (e->esize == 8)
        hipLaunchKernelGGL(ghost_fill_kernel<double>, dim3(bx, 1), dim3(256), 0, e->stream, e->d_desc,
                           (double*)e->U, e->nvar, slot);
    else
        hipLaunchKernelGGL(ghost_fill_kernel<float>, dim3(bx, 1), dim3(256), 0, e->stream, e->d_desc,
                           (float*)e->U, e->nvar, slot);
    HIP_TRY(hipGetLastError());
    return 0;
}

int alsvinn_hip_pool_snapshot(alsvinn_hip_ensemble* e, long long* ticket) {
    if (!e || !e->pool || !ticket) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    const long long tk = e->next_ticket++;
    const int j = int(tk % alsvinn_hip_ensemble::kRing);
    HIP_TRY(hipEventSynchronize(e->snap_ev[j]));  // the copy of ticket tk - kRing has run
    HIP_TRY(hipMemcpyAsync(e->h_snap[j], e->d_state, sizeof(SampleState) * e->nsamples,
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipEventRecord(e->snap_ev[j], e->stream));
    e->snap_ticket[j] = tk;
    *ticket = tk;
    return 0;
}

int alsvinn_hip_pool_poll(alsvinn_hip_ensemble* e, long long ticket, int wait, double* t,
                          long long* steps, int* active, int* status, int* ready) {
    if (!e || !e->pool || !ready || ticket < 0) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    const int j = int(ticket % alsvinn_hip_ensemble::kRing);
    if (e->snap_ticket[j] != ticket)
        return fail(ALSVINN_HIP_E_INVALID, "snapshot ticket " + std::to_string(ticket) +
                                               " was overwritten (at most 3 are kept)");
    *ready = 0;
    if (wait) {
        HIP_TRY(hipEventSynchronize(e->snap_ev[j]));
    } else {
        const hipError_t q = hipEventQuery(e->snap_ev[j]);
        if (q == hipErrorNotReady) return 0;
        if (q != hipSuccess) return fail(ALSVINN_HIP_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
    }
    for (int s = 0; s < e->nsamples; ++s) {
        const SampleState& S = e->h_snap[j][s];
        if (t) t[s] = S.t;
        if (steps) steps[s] = S.steps;
        if (active) active[s] = S.active;
        if (status) status[s] = S.status;
    }
    *ready = 1;
    return 0;
}

int alsvinn_hip_device_memory(long long* free_bytes, long long* total_bytes) {
    size_t f = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&f, &tot));
    if (free_bytes) *free_bytes = (long long)f;
    if (total_bytes) *total_bytes = (long long)tot;
    return 0;
}

int alsvinn_hip_ensemble_states(alsvinn_hip_ensemble* e, double* t, long long* steps, int* active,
                                int* status) {
    if (!e) return fail(ALSVINN_HIP_E_INVALID, "ensemble is NULL");
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<SampleState> st(e->nsamples);
    HIP_TRY(hipMemcpy(st.data(), e->d_state, sizeof(SampleState) * st.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < e->nsamples; ++s) {
        if (t) t[s] = st[s].t;
        if (steps) steps[s] = st[s].steps;
        if (active) active[s] = st[s].active;
        if (status) status[s] = st[s].status;
    }
    return 0;
}

int alsvinn_hip_ensemble_layout(const alsvinn_hip_ensemble* e, int sample, int dims6[6],
                                long long* offset, long long* elements_per_variable) {
    if (!e || sample < 0 || sample >= e->nsamples) return fail(ALSVINN_HIP_E_INVALID, "bad sample index");
    const SampleDesc& D = e->desc[sample];
    if (dims6) {
        dims6[0] = D.tx;
        dims6[1] = D.ty;
        dims6[2] = D.tz;
        dims6[3] = D.gx;
        dims6[4] = D.gy;
        dims6[5] = D.gz;
    }
    if (offset) *offset = D.base;
    if (elements_per_variable) *elements_per_variable = D.vstride;
    return 0;
}

int alsvinn_hip_ensemble_upload(alsvinn_hip_ensemble* e, int sample, const double* host) {
    if (!e || !host || sample < 0 || sample >= e->nsamples) return fail(ALSVINN_HIP_E_INVALID, "bad upload arguments");
    const SampleDesc& D = e->desc[sample];
    const size_t n = (size_t)D.vstride * e->nvar;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->esize == 8) {
        HIP_TRY(hipMemcpy(buf_at<double>(e->U, D.base), host, n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
        HIP_TRY(hipMemcpy(buf_at<float>(e->U, D.base), tmp.data(), n * 4, hipMemcpyHostToDevice));
    }
    return 0;
}

int alsvinn_hip_ensemble_download(alsvinn_hip_ensemble* e, int sample, double* host) {
    if (!e || !host || sample < 0 || sample >= e->nsamples) return fail(ALSVINN_HIP_E_INVALID, "bad download arguments");
    const SampleDesc& D = e->desc[sample];
    const size_t n = (size_t)D.vstride * e->nvar;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->esize == 8) {
        HIP_TRY(hipMemcpy(host, buf_at<double>(e->U, D.base), n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        HIP_TRY(hipMemcpy(tmp.data(), buf_at<float>(e->U, D.base), n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) host[i] = tmp[i];
    }
    return 0;
}

int alsvinn_hip_ensemble_state_pointer(alsvinn_hip_ensemble* e, void** ptr) {
    if (!e || !ptr) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    *ptr = e->U;
    return 0;
}

int alsvinn_hip_ensemble_apply_boundary(alsvinn_hip_ensemble* e) {
    if (!e) return fail(ALSVINN_HIP_E_INVALID, "ensemble is NULL");
    const int bx = (int)std::max<long long>(1, std::min<long long>(64, (e->max_vstride + 1023) / 1024));
    if (e->esize == 8)
        hipLaunchKernelGGL(ghost_fill_kernel<double>, dim3(bx, e->nsamples), dim3(256), 0, e->stream,
                           e->d_desc, (double*)e->U, e->nvar, 0);
    else
        hipLaunchKernelGGL(ghost_fill_kernel<float>, dim3(bx, e->nsamples), dim3(256), 0, e->stream,
                           e->d_desc, (float*)e->U, e->nvar, 0);
    HIP_TRY(hipGetLastError());
    return 0;
}

int alsvinn_hip_ensemble_reset_time(alsvinn_hip_ensemble* e) {
    if (!e) return fail(ALSVINN_HIP_E_INVALID, "ensemble is NULL");
    HIP_TRY(hipStreamSynchronize(e->stream));
    return init_state(e);
}

int alsvinn_hip_ensemble_set_schedule(alsvinn_hip_ensemble* e, int rows, int variant) {
    if (!e || rows < 0 || variant < 0 || variant >= kStageVariants)
        return fail(ALSVINN_HIP_E_INVALID, "bad schedule");
    HIP_TRY(hipStreamSynchronize(e->stream));
    destroy_graph(e);
    e->band_rows = rows;
    e->variant[0] = e->variant[1] = e->variant[2] = variant;
    e->tuned = true;
    return upload_tasks(e);
}

int alsvinn_hip_ensemble_set_small_1d(alsvinn_hip_ensemble* e, int max_rounds) {
    if (!e || max_rounds < 0) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    HIP_TRY(hipStreamSynchronize(e->stream));
    destroy_graph(e);
    e->solve1d_rounds = max_rounds;
    update_solve1d(e);
    return 0;
}

int alsvinn_hip_ensemble_get_small_1d(const alsvinn_hip_ensemble* e, int* threads) {
    if (!e || !threads) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    *threads = e->solve1d_threads;
    return 0;
}

int alsvinn_hip_ensemble_get_stage_variants(const alsvinn_hip_ensemble* e, int* variants) {
    if (!e || !variants) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    for (int s = 0; s < 3; ++s) variants[s] = e->variant[s];
    return 0;
}

int alsvinn_hip_ensemble_get_stage_waves(const alsvinn_hip_ensemble* e, int* waves) {
    if (!e || !waves) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    const int* w = e->esize == 8 ? e->k64.waves : e->k32.waves;
    for (int s = 0; s < 3; ++s) waves[s] = w[e->variant[s]];
    return 0;
}

int alsvinn_hip_ensemble_get_schedule(const alsvinn_hip_ensemble* e, int* rows, int* variant,
                                      int* ntasks) {
    if (!e) return fail(ALSVINN_HIP_E_INVALID, "ensemble is NULL");
    if (rows) *rows = e->band_rows;
    if (variant) *variant = e->variant[0];
    if (ntasks) *ntasks = e->ntasks;
    return 0;
}

int alsvinn_hip_ensemble_autotune(alsvinn_hip_ensemble* e, int* rows, int* variant) {
    if (!e) return fail(ALSVINN_HIP_E_INVALID, "ensemble is NULL");
    HIP_TRY(hipStreamSynchronize(e->stream));
    destroy_graph(e);
    int maxny = 1;
    for (const auto& D : e->desc) maxny = std::max(maxny, e->dim == 3 ? D.nz : D.ny);
    std::vector<int> cand;
    if (e->dim >= 2) {
        // powers-of-two halvings of the tallest grid, plus the band heights
        // that fill the GPU's wave slots: for w = 1..3 waves per SIMD and
        // one or two dispatch rounds, the lowest band whose work list still
        // fits in w * SIMDs * rounds waves (an almost full last round, where
        // a power of two can leave the second round nearly empty); grids too
        // small to fill the GPU with 8-row bands also try 1..7 rows
        std::set<int> c;
        for (int r = maxny; r >= 8; r = (r + 1) / 2) {
            c.insert(r);
            if (r == 8) break;
        }
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const long long simds = 4LL * std::max(1, cus);
        const int saved = e->band_rows;
        auto waves_of = [&](int h) {
            std::vector<WaveTask> t;
            e->band_rows = h;
            build_tasks(e, t);
            return (long long)t.size() * (e->dim == 3 ? kGroupRows : 1);
        };
        for (int w = 1; w <= 3; ++w)
            for (int rounds = 1; rounds <= 2; ++rounds) {
                const long long slots = simds * w * rounds;
                if (waves_of(maxny) <= slots) {  // some band height fits
                    int lo = 1, hi = maxny;  // waves_of decreases with h: smallest h that fits
                    while (lo < hi) {
                        const int mid = (lo + hi) / 2;
                        if (waves_of(mid) <= slots)
                            hi = mid;
                        else
                            lo = mid + 1;
                    }
                    c.insert(lo);
                }
            }
        if (waves_of(std::min(8, maxny)) < simds)
            for (int h = 1; h < std::min(8, maxny); ++h) c.insert(h);
        e->band_rows = saved;
        cand.assign(c.rbegin(), c.rend());
    } else {
        cand.push_back(1);
    }
    // Each (band height, budget) pair is timed twice, the budgets in opposite
    // orders, so that the clock's drift under sustained load does not favour
    // the configurations timed first; every stage kind keeps its own fastest
    // budget, and the band height with the smallest total wins.
    const int nc = (int)cand.size();
    std::vector<double> t((size_t)nc * kStageVariants * 3, 0.0);
    for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < nc; ++c) {
            e->band_rows = cand[c];
            if (int rc = upload_tasks(e)) return rc;
            for (int i = 0; i < kStageVariants; ++i) {
                const int v = pass == 0 ? i : kStageVariants - 1 - i;
                e->variant[0] = e->variant[1] = e->variant[2] = v;
                double ms[3];
                if (e->esize == 8)
                    time_stages<double>(e, e->k64, 1, ms);
                else
                    time_stages<float>(e, e->k32, 1, ms);
                for (int s = 0; s < 3; ++s) t[((size_t)c * kStageVariants + v) * 3 + s] += ms[s];
            }
        }
    double best = 1e300;
    int best_rows = cand[0], best_var[3] = {0, 0, 0};
    for (int c = 0; c < nc; ++c) {
        double total = 0;
        int vs[3];
        for (int s = 0; s < 3; ++s) {
            vs[s] = 0;
            for (int v = 1; v < kStageVariants; ++v)
                if (t[((size_t)c * kStageVariants + v) * 3 + s] <
                    t[((size_t)c * kStageVariants + vs[s]) * 3 + s])
                    vs[s] = v;
            total += t[((size_t)c * kStageVariants + vs[s]) * 3 + s];
        }
        if (total < best) {
            best = total;
            best_rows = cand[c];
            for (int s = 0; s < 3; ++s) best_var[s] = vs[s];
        }
    }
    std::vector<unsigned long long> z(3 * (size_t)e->nsamples, 0ull);
    HIP_TRY(hipMemcpy(e->d_ws, z.data(), sizeof(unsigned long long) * z.size(), hipMemcpyHostToDevice));
    for (int s = 0; s < 3; ++s) e->variant[s] = best_var[s];
    e->band_rows = best_rows;
    e->tuned = true;
    if (int rc = upload_tasks(e)) return rc;
    HIP_TRY(hipGetLastError());
    if (rows) *rows = best_rows;
    if (variant) *variant = best_var[0];
    return 0;
}

int alsvinn_hip_ensemble_step(alsvinn_hip_ensemble* e, int nsteps) {
    if (!e || nsteps < 0) return fail(ALSVINN_HIP_E_INVALID, "bad step arguments");
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    if (int rc = enqueue_steps(e, nsteps)) return rc;
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    return 0;
}

static int ensure_graph(alsvinn_hip_ensemble* e, int steps) {
    if (e->graph_exec && e->graph_steps == steps) return 0;
    destroy_graph(e);
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_steps(e, steps);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) return fail(ALSVINN_HIP_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
    e->graph = g;
    HIP_TRY(hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
    e->graph_steps = steps;
    return 0;
}

int alsvinn_hip_ensemble_set_hold(alsvinn_hip_ensemble* e, int adjuster, long long saved_limit) {
    if (!e || adjuster < 0 || adjuster > 1) return fail(ALSVINN_HIP_E_INVALID, "adjuster must be 0 or 1");
    const long long limit = saved_limit < 0 ? LLONG_MAX : saved_limit;
    hipLaunchKernelGGL(set_hold_kernel, dim3((e->nsamples + 255) / 256), dim3(256), 0, e->stream,
                       e->d_state, e->nsamples, adjuster, limit);
    HIP_TRY(hipGetLastError());
    return 0;
}

int alsvinn_hip_ensemble_saves(alsvinn_hip_ensemble* e, int sample, long long saved[2]) {
    if (!e || sample < 0 || sample >= e->nsamples || !saved)
        return fail(ALSVINN_HIP_E_INVALID, "bad sample index");
    SampleState S;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(&S, e->d_state + sample, sizeof(S), hipMemcpyDeviceToHost));
    saved[0] = S.saved[0];
    saved[1] = S.saved[1];
    return 0;
}

int alsvinn_hip_ensemble_run_steps(alsvinn_hip_ensemble* e, int nsteps, int steps_per_graph) {
    if (!e || nsteps < 0 || steps_per_graph < 1 || nsteps % steps_per_graph != 0)
        return fail(ALSVINN_HIP_E_INVALID, "nsteps must be a multiple of steps_per_graph");
    if (int rc = ensure_graph(e, steps_per_graph)) return rc;
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    for (int i = 0; i < nsteps / steps_per_graph; ++i) HIP_TRY(hipGraphLaunch(e->graph_exec, e->stream));
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    return 0;
}

int alsvinn_hip_ensemble_step_timed(alsvinn_hip_ensemble* e, int nsteps, double kernel_ms[3],
                                    long long launches[3]) {
    if (!e || nsteps < 0) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    KernelTimer tm;
    tm.s = e->stream;
    tm.reserve((size_t)nsteps * 5);
    HIP_TRY(hipStreamSynchronize(e->stream));
    int rc = enqueue_steps(e, nsteps, &tm);
    double ms[3];
    long long n[3];
    tm.collect(ms, n);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    for (int k = 0; k < 3; ++k) {
        if (kernel_ms) kernel_ms[k] = ms[k];
        if (launches) launches[k] = n[k];
    }
    return 0;
}

int alsvinn_hip_ensemble_run(alsvinn_hip_ensemble* e, int steps_per_sync) {
    if (!e || steps_per_sync < 1) return fail(ALSVINN_HIP_E_INVALID, "bad run arguments");
    if (int rc = ensure_graph(e, steps_per_sync)) return rc;
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    for (;;) {
        HIP_TRY(hipGraphLaunch(e->graph_exec, e->stream));
        hipLaunchKernelGGL(count_active_kernel, dim3(1), dim3(256), 0, e->stream, e->d_state,
                           e->nsamples, e->d_count);
        HIP_TRY(hipMemcpyAsync(e->h_count, e->d_count, sizeof(int), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (*e->h_count == 0) break;
    }
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<SampleState> st(e->nsamples);
    HIP_TRY(hipMemcpy(st.data(), e->d_state, sizeof(SampleState) * st.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < e->nsamples; ++s)
        if (st[s].status != 0)
            return fail(ALSVINN_HIP_E_CONSTRAINT,
                        "Simulation state does not obey constraints! sample " + std::to_string(s) +
                            " at time " + std::to_string(st[s].t) + ", number of timesteps performed " +
                            std::to_string(st[s].steps));
    return 0;
}

int alsvinn_hip_ensemble_synchronize(alsvinn_hip_ensemble* e) {
    if (!e) return fail(ALSVINN_HIP_E_INVALID, "ensemble is NULL");
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int alsvinn_hip_ensemble_sample_info(alsvinn_hip_ensemble* e, int sample, double* t,
                                     long long* steps, double* last_dt, int* status, int* active) {
    if (!e || sample < 0 || sample >= e->nsamples) return fail(ALSVINN_HIP_E_INVALID, "bad sample index");
    HIP_TRY(hipStreamSynchronize(e->stream));
    SampleState S;
    HIP_TRY(hipMemcpy(&S, e->d_state + sample, sizeof(S), hipMemcpyDeviceToHost));
    if (t) *t = S.t;
    if (steps) *steps = S.steps;
    if (last_dt) *last_dt = S.last_dt;
    if (status) *status = S.status;
    if (active) *active = S.active;
    return 0;
}

int alsvinn_hip_ensemble_record_dt(alsvinn_hip_ensemble* e, long long capacity) {
    if (!e || capacity < 0) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    HIP_TRY(hipStreamSynchronize(e->stream));
    destroy_graph(e);
    hipFree(e->d_dt_hist);
    e->d_dt_hist = nullptr;
    e->dt_cap = 0;
    if (capacity > 0) {
        HIP_TRY(hipMalloc(&e->d_dt_hist, sizeof(double) * capacity * e->nsamples));
        HIP_TRY(hipMemset(e->d_dt_hist, 0, sizeof(double) * capacity * e->nsamples));
        e->dt_cap = capacity;
    }
    return 0;
}

int alsvinn_hip_ensemble_dt_history(alsvinn_hip_ensemble* e, int sample, double* host,
                                    long long capacity) {
    if (!e || sample < 0 || sample >= e->nsamples || !host) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    if (!e->d_dt_hist) return fail(ALSVINN_HIP_E_INVALID, "dt recording is off");
    HIP_TRY(hipStreamSynchronize(e->stream));
    const long long n = std::min(capacity, e->dt_cap);
    HIP_TRY(hipMemcpy(host, e->d_dt_hist + (long long)sample * e->dt_cap, sizeof(double) * n,
                      hipMemcpyDeviceToHost));
    return 0;
}

int alsvinn_hip_ensemble_rhs(alsvinn_hip_ensemble* e, int sample, double* host_out,
                             double wave_speeds[3]) {
    if (!e || sample < 0 || sample >= e->nsamples) return fail(ALSVINN_HIP_E_INVALID, "bad sample index");
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<unsigned long long> z(3 * (size_t)e->nsamples, 0ull);
    HIP_TRY(hipMemcpy(e->d_ws, z.data(), sizeof(unsigned long long) * z.size(), hipMemcpyHostToDevice));
    const SampleDesc& D = e->desc[sample];
    const size_t n = (size_t)D.vstride * e->nvar;
    if (e->esize == 8) {
        StageArgs<double> a;
        std::memset(&a, 0, sizeof(a));
        a.desc = e->d_desc;
        a.state = e->d_state;
        a.ws = e->d_ws;
        const int blocks = set_tasks(e, a);
        a.gamma = e->scheme.gamma;
        a.mode = MODE_RHS;
        a.in = (double*)e->U;
        a.in2 = (double*)e->U;
        a.out = (double*)e->A;
        e->k64.stage[e->variant[0]](a, blocks, e->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (host_out) HIP_TRY(hipMemcpy(host_out, buf_at<double>(e->A, D.base), n * 8, hipMemcpyDeviceToHost));
    } else {
        StageArgs<float> a;
        std::memset(&a, 0, sizeof(a));
        a.desc = e->d_desc;
        a.state = e->d_state;
        a.ws = e->d_ws;
        const int blocks = set_tasks(e, a);
        a.gamma = e->scheme.gamma;
        a.mode = MODE_RHS;
        a.in = (float*)e->U;
        a.in2 = (float*)e->U;
        a.out = (float*)e->A;
        e->k32.stage[e->variant[0]](a, blocks, e->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (host_out) {
            std::vector<float> tmp(n);
            HIP_TRY(hipMemcpy(tmp.data(), buf_at<float>(e->A, D.base), n * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) host_out[i] = tmp[i];
        }
    }
    unsigned long long w[3];
    HIP_TRY(hipMemcpy(w, e->d_ws + 3 * sample, sizeof(w), hipMemcpyDeviceToHost));
    if (wave_speeds)
        for (int k = 0; k < 3; ++k) std::memcpy(&wave_speeds[k], &w[k], 8);
    HIP_TRY(hipMemcpy(e->d_ws, z.data(), sizeof(unsigned long long) * z.size(), hipMemcpyHostToDevice));
    return 0;
}

int alsvinn_hip_ensemble_interval_ms(alsvinn_hip_ensemble* e, alsvinn_hip_ensemble* ref,
                                     double* start_ms, double* end_ms) {
    if (!e || !ref || !start_ms || !end_ms) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    HIP_TRY(hipEventSynchronize(e->ev1));
    HIP_TRY(hipEventSynchronize(ref->ev1));
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, ref->ev0, e->ev0));
    HIP_TRY(hipEventElapsedTime(&b, ref->ev0, e->ev1));
    *start_ms = a;
    *end_ms = b;
    return 0;
}

int alsvinn_hip_ensemble_last_elapsed_ms(alsvinn_hip_ensemble* e, double* ms) {
    if (!e || !ms) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    HIP_TRY(hipEventSynchronize(e->ev1));
    float f = 0;
    HIP_TRY(hipEventElapsedTime(&f, e->ev0, e->ev1));
    *ms = f;
    return 0;
}

}  // extern "C"

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
// of the rows it needs. x-neighbours come from the adjacent lanes (DPP lane
// moves); the G = ghost-count lanes at each end of a strip are halo lanes
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
    int active;  // 1 until the sample reaches its end (or fails)
    int status;  // 0 ok, 2 constraint violation
    long long hold[2];  // the sample pauses while saved[k] >= hold[k] (a host-side writer's turn)
};

// A sample advances while it is active and not held at a save.
__device__ __forceinline__ bool runnable(const SampleState& S) {
    return S.active && S.saved[0] < S.hold[0] && S.saved[1] < S.hold[1];
}

// A wavefront's work: up to four segments, each a strip of interior
// columns [x0, x0+ncols) of one sample occupying lanes
// [lane0, lane0 + ncols + 2G) (G halo lanes on each side), all marching
// over the same rows [y0, y1). Packing the narrow last strips of several
// samples into one wave keeps the lanes busy.
constexpr int kMaxSeg = 4;
struct WaveTask {
    int nseg;
    int y0, y1;
    int z0, z1;
    int pad[3];
    int sample[kMaxSeg];
    int x0[kMaxSeg];
    int ncols[kMaxSeg];
    int lane0[kMaxSeg];
};

// Where the step's last kernel ends the step (end_step).
struct StepEnd {
    unsigned int* counter;  // null: no step end in this launch
    const SampleDesc* desc;
    SampleState* state;
    unsigned long long* ws;
    double* dt_hist;
    long long dt_cap;
    int nsamples;
};

template <class T>
struct StageArgs {
    const SampleDesc* desc;
    SampleState* state;
    unsigned long long* ws;  // 3 per sample: bit patterns of non-negative doubles
    const WaveTask* tasks;
    int ntasks;
    // when set, the number of tasks is read here at run time (pools: the
    // work list changes as solves come and go, the launch and a captured
    // graph stay the same; ntasks then is the grid size of the 3D kernel)
    const int* ntasks_dev;
    int mode;   // StageMode; MODE_RK3_1 forms its stage input as in*c + in2
    int check;  // unused by the stage kernels: the step's last stage checks
    int no_ws;  // MODE_RHS without the wave-speed reduction (RK4 stages 2-4)
    // wave speeds through raise_wave_speed (read, then atomic only if
    // larger): set when a sample spans many wavefronts, whose atomics on the
    // sample's slots would otherwise serialise; with few wavefronts per
    // sample the plain atomic (no wait for the read) is cheaper
    int ws_filter;
    // when set (ensembles with large samples), the first stage stores each
    // task's segment maxima here instead: entry ((slot * kMaxSeg + seg) * 3 +
    // d), slot the task's index (3D: times kGroupRows, plus the row), and a
    // reduction kernel forms each sample's maximum after the launch
    // (thousands of same-address atomics at the end of a launch serialise)
    unsigned long long* ws_part;
    const T* in;
    const T* in2;  // U0
    T* out;
    double gamma;
    // volume operators (OP kernels): one device pointer per conserved
    // variable of the input and output volumes (a reference Volume's
    // per-variable memories), used instead of in / out
    const T* vin[5];
    T* vout[5];
    StepEnd end;  // the step's last stage: ends the step after its work
};

// Memory of one volume: variable v of padded cell i. The ensemble's buffers
// hold a sample's variables vstride elements apart; a reference Volume has
// one memory per variable (Volume.cpp:44-60).
template <class T>
struct StridedVol {
    T* p;
    long long vs;
    __device__ __forceinline__ T& operator()(int v, long long i) const { return p[i + v * vs]; }
};
template <class T>
struct PerVarVol {
    T* p[5];
    __device__ __forceinline__ T& operator()(int v, long long i) const { return p[v][i]; }
};

// A cell of a wavefront whose lanes all belong to one sample: its row's
// byte offset inside the variable's memory, the same for every lane (a
// scalar register), and the lane's column in bytes, fixed for the task.
struct UniIdx {
    unsigned row;   // wave-uniform
    unsigned lane;  // per lane
};
// The volumes of such a wave are read and written through buffer resources
// (scalar base, size and format): an access is the resource, the scalar
// byte offset of the variable and row, and the lane's column offset, which
// the hardware adds itself, so loads and stores need no vector address
// arithmetic. Every offset used stays below the resource's size.
constexpr int kBufferWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, unsigned lane, unsigned off) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, lane, off, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, lane, off, 0));
}
template <class T>
__device__ __forceinline__ void buf_store(T x, __amdgpu_buffer_rsrc_t r, unsigned lane,
                                          unsigned off) {
    if constexpr (sizeof(T) == 8) {
        using W = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0u, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(W, x), r, lane, off, 0);
    } else {
        using W = decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0u, 0u, 0));
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(W, x), r, lane, off, 0);
    }
}
template <class T>
struct BufRef {
    __amdgpu_buffer_rsrc_t r;
    unsigned lane, off;
    __device__ __forceinline__ operator T() const { return buf_load<T>(r, lane, off); }
    __device__ __forceinline__ void operator=(T x) const { buf_store<T>(x, r, lane, off); }
};
// Variable v of a volume: resource r[v] at byte offset vo[v] (one resource
// for a whole sample of the ensemble, variables vstride elements apart; one
// per memory for a reference Volume).
template <class T>
struct UniVol {
    __amdgpu_buffer_rsrc_t r[5];
    unsigned vo[5];
    __device__ __forceinline__ BufRef<std::remove_const_t<T>> operator()(int v, UniIdx i) const {
        return {r[v], i.lane, vo[v] + i.row};
    }
};
// A struct read from a wave-uniform address into scalar registers.
template <class S>
__device__ __forceinline__ S uniform_load(const S* p) {
    static_assert(sizeof(S) % 4 == 0, "whole words");
    S r;
    const int* src = (const int*)p;
    int* dst = (int*)&r;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(S) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    return r;
}

// Input, second input and output accessors of a stage kernel for sample D:
// OP = 0 the ensemble's buffers, OP = 1 the volume operators' pointers.
template <int OP, class T>
__device__ __forceinline__ auto stage_in(const StageArgs<T>& a, const SampleDesc& D, bool second) {
    if constexpr (OP == 0) {
        return StridedVol<const T>{(second ? a.in2 : a.in) + D.base, D.vstride};
    } else {
        PerVarVol<const T> r;
#pragma unroll
        for (int v = 0; v < 5; ++v) r.p[v] = a.vin[v];
        return r;
    }
}
template <int OP, class T>
__device__ __forceinline__ auto stage_out(const StageArgs<T>& a, const SampleDesc& D) {
    if constexpr (OP == 0) {
        return StridedVol<T>{a.out + D.base, D.vstride};
    } else {
        PerVarVol<T> r;
#pragma unroll
        for (int v = 0; v < 5; ++v) r.p[v] = a.vout[v];
        return r;
    }
}

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
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (k < D.nadj && S.saved[k] > 0) {
            const double next = double(S.saved[k]) * D.interval[k] + 0.0;
            const double cand = next - S.t;
            dt = (cand < dt) ? cand : dt;
        }
    }
    return dt;
}

// End of a timestep of sample s (Simulator::incrementSolution tail,
// Simulator::atEnd, FixedIntervalWriter::write): t += dt, save counters,
// stop flag, wave-speed reset; the dt is recorded when asked.
__device__ __forceinline__ void advance_sample(const SampleDesc& D, SampleState& S, double dt) {
    S.t = S.t + dt;
    S.steps += 1;
    S.last_dt = dt;
    // constant indices: the state stays in registers
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (k < D.nadj && S.t >= double(S.saved[k]) * D.interval[k] + 0.0) S.saved[k] += 1;
    if (S.status != 0 || S.t >= D.end_time || (D.max_steps >= 0 && S.steps >= D.max_steps))
        S.active = 0;
}

__device__ __forceinline__ void finish_sample(const SampleDesc* __restrict__ desc, SampleState* state,
                                              unsigned long long* ws, double* dt_hist,
                                              long long dt_cap, int s) {
    const SampleDesc D = desc[s];
    SampleState S = state[s];
    if (!runnable(S)) return;
    const double dt = sample_dt(D, S, ws + 3 * s);
    if (dt_hist && S.steps < dt_cap) dt_hist[(long long)s * dt_cap + S.steps] = dt;
    ws[3 * s + 0] = 0;
    ws[3 * s + 1] = 0;
    ws[3 * s + 2] = 0;
    advance_sample(D, S, dt);
    state[s] = S;
}

// A sample that one workgroup advances alone for whole steps
// (solve_kernel_1d): the stages read the step's dt from here and raise the
// wave-speed maxima and the failure flag in LDS, instead of the sample's
// global state and slots (its copy of the state lives in registers).
struct LocalStep {
    unsigned long long* ws;  // LDS: this step's maxima, bit patterns (3)
    int* status;             // LDS: 2 once a constraint check failed
    double dt;               // the step's dt (the stages after the first)
};

// The end of a step fused into the step's last kernel: every workgroup
// counts itself out once its work is done, and the last one to finish (its
// count is the grid size) ends the step of every sample and clears the
// counter for the next launch. This saves the step a launch; the counter is
// the only state, and it is zero again after every launch.
// The only values the last workgroup reads that this launch writes are the
// status flags of failed constraint checks (`flagged`: this thread set one).
// So only a workgroup that set one publishes its writes before counting
// itself out: the threads that set a flag wait for their atomics and write
// back the L2 (an agent-scope release fence) before the barrier, and the
// counting thread releases again after it. The others count out with a
// relaxed atomic, and the last workgroup acquires (its fence synchronises
// with every released workgroup through the counter's read-modify-write
// chain). A release per workgroup would write back the L2 once per
// workgroup, thousands of times per launch.
__device__ __forceinline__ void end_step(const StepEnd& E, bool flagged) {
    if (!E.counter) return;
    __shared__ int s_last;
    if (flagged) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const bool publish = __syncthreads_or(flagged);
    if (threadIdx.x == 0) {
        if (publish) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
        s_last = __hip_atomic_fetch_add(E.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                 nb - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the published status flags
    for (int s = threadIdx.x; s < E.nsamples; s += blockDim.x)
        finish_sample(E.desc, E.state, E.ws, E.dt_hist, E.dt_cap, s);
    if (threadIdx.x == 0) __hip_atomic_store(E.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane i receives the value of lane i-1 (lane_from_lower) or lane i+1
// (lane_from_upper); the lane without a source (lane 0, resp. 63) receives
// 0. DPP wave_shr:1 / wave_shl:1 moves, two 32-bit halves for a double. Lanes
// 0 and 63 are always halo lanes of their segment (every segment carries
// ghost-width halo lanes on both sides), so no face or cell that is written
// or reduced reads the value they receive; writing 0 there (bound_ctrl)
// instead of keeping the lane's own value lets the move write a fresh
// register without first copying its source into it.
__device__ __forceinline__ int dpp_shr1(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int dpp_shl1(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ double lane_from_lower(double v) {
    return __hiloint2double(dpp_shr1(__double2hiint(v)), dpp_shr1(__double2loint(v)));
}
__device__ __forceinline__ double lane_from_upper(double v) {
    return __hiloint2double(dpp_shl1(__double2hiint(v)), dpp_shl1(__double2loint(v)));
}
__device__ __forceinline__ float lane_from_lower(float v) {
    return __int_as_float(dpp_shr1(__float_as_int(v)));
}
__device__ __forceinline__ float lane_from_upper(float v) {
    return __int_as_float(dpp_shl1(__float_as_int(v)));
}

// Max over the lanes of each contiguous segment; the first lane of a
// segment ends up holding its segment's max.
template <class T>
__device__ __forceinline__ T segment_max(T v, int seg) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T o = __shfl_down(v, off, 64);
        const int so = __shfl_down(seg, off, 64);
        if (lane + off < 64 && so == seg) v = std_max(v, o);
    }
    return v;
}

// Stage input at flat index i (variable stride vs): plain or lazy
// U1 = L0*c + U0.
// Raises a sample's wave-speed slot to v (the bit pattern of a non-negative
// double, whose unsigned order is the order of the values). With `filter`
// the slot is read first and the atomic skipped when it already holds v or
// more: the result is the same maximum, and a single solve's thousands of
// wavefronts then issue a few dozen atomics on the slot instead of one
// each, which would serialise at the slot.
__device__ __forceinline__ void raise_wave_speed(unsigned long long* slot, double v, int filter) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    if (!filter || b > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(slot, b);
}

// A segment's maximum wave speed m in direction d, reported by its first
// lane (`leader`) for a sample that is (`run`) or is not stepping: into the
// task's partial-maxima entry (a.ws_part; 0 for a sample not stepping, so
// the entry is written every launch) or onto the sample's slot.
template <class T>
__device__ __forceinline__ void report_wave_speed(const StageArgs<T>& a, long long slot, int seg,
                                                  int sample, int d, double m, bool leader,
                                                  bool run) {
    if (!leader) return;
    if (a.ws_part)
        a.ws_part[(slot * kMaxSeg + seg) * 3 + d] =
            run ? (unsigned long long)__double_as_longlong(m) : 0ull;
    else if (run)
        raise_wave_speed(a.ws + 3 * sample + d, m, a.ws_filter);
}

template <class T, int NV>
__device__ __forceinline__ void load_strided(const T* __restrict__ in, const T* __restrict__ in2,
                                             bool lazy, T c, long long i, long long vs, T* q) {
    if (lazy) {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = in[i + v * vs] * c + in2[i + v * vs];
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = in[i + v * vs];
    }
}

template <class T, int NV, class Acc, class I>
__device__ __forceinline__ void load_cell(const Acc& in, const Acc& in2, bool lazy, T c, I i,
                                          T* q) {
    if constexpr (std::is_same_v<Acc, StridedVol<const T>>) {
        load_strided<T, NV>(in.p, in2.p, lazy, c, i, in.vs, q);
    } else if (lazy) {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = in(v, i) * c + in2(v, i);
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = in(v, i);
    }
}

// A prefetched cell of the stage input: issue() starts the loads and finish()
// forms the lazy combination U1 = L0*c + U0, so that the work in between
// hides the loads' latency (forming it at the load makes the wave wait for
// both loads at once: +12% for the second SSP-RK3 stage in 2D, +40-80% in 3D).
template <class T, int NV>
struct PendingCell {
    T a[NV], b[NV];
    __device__ __forceinline__ void issue_strided(const T* __restrict__ in,
                                                  const T* __restrict__ in2, bool lazy,
                                                  long long i, long long vs) {
#pragma unroll
        for (int v = 0; v < NV; ++v) a[v] = in[i + v * vs];
        if (lazy) {
#pragma unroll
            for (int v = 0; v < NV; ++v) b[v] = in2[i + v * vs];
        }
    }
    template <class Acc, class I>
    __device__ __forceinline__ void issue(const Acc& in, const Acc& in2, bool lazy, I i) {
        if constexpr (std::is_same_v<Acc, StridedVol<const T>>) {
            issue_strided(in.p, in2.p, lazy, i, in.vs);
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) a[v] = in(v, i);
            if (lazy) {
#pragma unroll
                for (int v = 0; v < NV; ++v) b[v] = in2(v, i);
            }
        }
    }
    __device__ __forceinline__ void finish(bool lazy, T c, T* q) const {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = lazy ? a[v] * c + b[v] : a[v];
    }
};

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

// Padded positions of interior index c on one axis: the cell itself (a)
// plus the ghost cells that mirror it (b: bottom side, c: top side; -1 when
// none).
__device__ __forceinline__ void axis_images(int c, int N, int G, int bc, int& a, int& b, int& d) {
    a = c + G;
    b = -1;
    d = -1;
    if (G == 0) return;
    if (bc == 1) {
        if (c >= N - G) b = c - N + G;
        if (c < G) d = c + G + N;
    } else {
        if (c < G) b = G - 1 - c;
        if (c >= N - G) d = 2 * N + G - 1 - c;
    }
}

// The padded column of a lane's cell and of the (at most two) ghost columns
// that mirror it: computed once per task, since a lane keeps its column
// while it marches.
struct ColumnImages {
    int px, xb, xd;  // xb / xd: -1 when there is none
};
__device__ __forceinline__ ColumnImages column_images(const SampleDesc& D, int col) {
    ColumnImages c;
    axis_images(col, D.nx, D.gx, D.bc, c.px, c.xb, c.xd);
    return c;
}

// Writes one result cell of row `row` (and plane `plane`) to its padded
// position and, with IMAGES, to every ghost cell that mirrors it. The row
// and plane are the same for every lane of a wave: their images are
// computed on the scalar unit and skipped by a uniform branch on the rows
// that have none (all but the G rows next to each boundary).
template <class T, int NV, int DIM, bool IMAGES, class Acc>
__device__ __forceinline__ void store_result(const Acc& out, const SampleDesc& D,
                                             const ColumnImages& C, int row, int plane,
                                             const T* res) {
    row = __builtin_amdgcn_readfirstlane(row);
    plane = __builtin_amdgcn_readfirstlane(plane);
    const long long tx = D.tx;
    auto put_row = [&](int py, int pz) {
        const long long base = ((long long)pz * D.ty + py) * tx;
#pragma unroll
        for (int v = 0; v < NV; ++v) out(v, base + C.px) = res[v];
        if constexpr (IMAGES) {
            if (C.xb >= 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) out(v, base + C.xb) = res[v];
            }
            if (C.xd >= 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) out(v, base + C.xd) = res[v];
            }
        }
    };
    int ya = DIM >= 2 ? row + D.gy : 0, yb = -1, yd = -1;
    int za = DIM >= 3 ? plane + D.gz : 0, zb = -1, zd = -1;
    if constexpr (IMAGES && DIM >= 2) axis_images(row, D.ny, D.gy, D.bc, ya, yb, yd);
    if constexpr (IMAGES && DIM >= 3) axis_images(plane, D.nz, D.gz, D.bc, za, zb, zd);
    auto put_plane = [&](int pz) {
        put_row(ya, pz);
        if constexpr (IMAGES && DIM >= 2) {
            if (yb >= 0) put_row(yb, pz);
            if (yd >= 0) put_row(yd, pz);
        }
    };
    put_plane(za);
    if constexpr (IMAGES && DIM >= 3) {
        if (zb >= 0) put_plane(zb);
        if (zd >= 0) put_plane(zd);
    }
}

// Workgroups are dispatched round-robin over the GPU's 8 XCDs (each with its
// own L2); this maps dispatch order to a logical block index so that each XCD
// runs a contiguous range of logical blocks, i.e. of neighbouring wave tasks,
// which share halo rows and columns in that XCD's L2.
__device__ __forceinline__ int xcd_block() {
    constexpr int kXcd = 8;
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb / kXcd, r = nb % kXcd;
    const int x = b % kXcd, k = b / kXcd;
    return x * q + (x < r ? x : r) + k;
}

// Reconstructions whose weights come from one quantity per face, shared by
// the face's two cells (ReconWENO2::face_weight).
template <class R, class = void>
struct recon_face_weights : std::false_type {};
template <class R>
struct recon_face_weights<R, std::void_t<decltype(R::face_weights)>> : std::true_type {};

// max(acc, s) for the running wave speed of a lane: s is a NaN (skipped, as
// std::max(acc, s) skips it) or non-negative, acc starts at +0 and stays
// finite, so the hardware maximum (IEEE maxNum: a quiet NaN operand yields
// the other) is std::max here. One instruction; the compiler's fmax would
// first canonicalise both operands.
__device__ __forceinline__ double ws_max(double acc, double s) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(acc), "v"(s));
    return r;
}
__device__ __forceinline__ float ws_max(float acc, float s) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(acc), "v"(s));
    return r;
}

// A struct read through the scalar data cache (one s_load per few words,
// the values land in scalar registers). Only for data that no wave of the
// running kernel writes (work lists and sample descriptors are written by
// the host or by earlier launches).
template <class S>
__device__ __forceinline__ S scalar_load(const S* p) {
    static_assert(sizeof(S) % 4 == 0, "whole words");
    using CW = const __attribute__((address_space(4))) int;
    CW* src = (CW*)p;
    S r;
    int* dst = (int*)&r;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(S) / 4); ++k) dst[k] = src[k];
    return r;
}

// Accessors of a wave whose lanes all belong to one sample (buffer
// resources: OP = 0 one per volume, covering the sample's NV variables;
// OP = 1 one per reference memory). The single-sample march runs only where
// these fit 32-bit offsets (uni_fits).
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const T* p, unsigned long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(unsigned)bytes, kBufferWord3);
}
template <int OP, int NV, class Vol, class T>
__device__ __forceinline__ Vol uni_vol(const T* base, const T* const* per_var, const SampleDesc& D) {
    Vol r;
    const unsigned long long vbytes = (unsigned long long)D.vstride * sizeof(T);
    if constexpr (OP == 0) {
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(base + D.base, NV * vbytes);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            r.r[v] = rs;
            r.vo[v] = unsigned(v * vbytes);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            r.r[v] = make_rsrc(per_var[v], vbytes);
            r.vo[v] = 0;
        }
    }
    return r;
}
template <int OP, int NV, class T>
__device__ __forceinline__ UniVol<const T> uni_in(const StageArgs<T>& a, const SampleDesc& D,
                                                  bool second) {
    return uni_vol<OP, NV, UniVol<const T>>(second ? a.in2 : a.in, a.vin, D);
}
template <int OP, int NV, class T>
__device__ __forceinline__ UniVol<T> uni_out(const StageArgs<T>& a, const SampleDesc& D) {
    return uni_vol<OP, NV, UniVol<T>>((const T*)a.out, a.vout, D);
}
// A sample's volume fits the buffer resources' 32-bit offsets.
template <class T, int NV>
__device__ __forceinline__ bool uni_fits(const StageArgs<T>& a, int sample) {
    const long long vs = scalar_load(&a.desc[sample].vstride);
    return (unsigned long long)vs * sizeof(T) * NV < (1ull << 32);
}

// store_result for a single-sample wave: the row's (and plane's) byte
// offset and their ghost images are scalar, the lane's column offset and its
// ghost columns' offsets (negative when none) are fixed for the task.
template <class T, int NV, int DIM, bool IMAGES>
__device__ __forceinline__ void store_result_uni(const UniVol<T>& out, const SampleDesc& D,
                                                 unsigned lo, int xbo, int xdo, int row, int plane,
                                                 const T* res) {
    auto put_row = [&](int py, int pz) {
        const unsigned rb =
            (unsigned(pz) * unsigned(D.ty) + unsigned(py)) * unsigned(D.tx) * unsigned(sizeof(T));
#pragma unroll
        for (int v = 0; v < NV; ++v) out(v, UniIdx{rb, lo}) = res[v];
        if constexpr (IMAGES) {
            if (xbo >= 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) out(v, UniIdx{rb, unsigned(xbo)}) = res[v];
            }
            if (xdo >= 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) out(v, UniIdx{rb, unsigned(xdo)}) = res[v];
            }
        }
    };
    int ya = DIM >= 2 ? row + D.gy : 0, yb = -1, yd = -1;
    int za = DIM >= 3 ? plane + D.gz : 0, zb = -1, zd = -1;
    if constexpr (IMAGES && DIM >= 2) axis_images(row, D.ny, D.gy, D.bc, ya, yb, yd);
    if constexpr (IMAGES && DIM >= 3) axis_images(plane, D.nz, D.gz, D.bc, za, zb, zd);
    auto put_plane = [&](int pz) {
        put_row(ya, pz);
        if constexpr (IMAGES && DIM >= 2) {
            if (yb >= 0) put_row(yb, pz);
            if (yd >= 0) put_row(yd, pz);
        }
    };
    put_plane(za);
    if constexpr (IMAGES && DIM >= 3) {
        if (zb >= 0) put_plane(zb);
        if (zd >= 0) put_plane(zd);
    }
}

// One wave task of the 1D/2D stage kernel: the strip [x0, x0 + ncols) of
// `sample` occupying lanes [lane0, lane0 + ncols + 2G), marched over rows
// [y0, y1). UNI: every lane of the wave belongs to this segment's sample,
// so sample, x0, ncols and lane0 are wave-uniform and the volumes are
// addressed from scalar bases; otherwise each lane passes its own segment.
template <class T, class Eq, class Recon, class Flux, int DIM, int MODE, int OP, bool UNI,
          bool LOCAL = false>
__device__ __forceinline__ void march_12(const StageArgs<T>& a, const Eq& eq, int task, int sample,
                                         int x0, int ncols, int lane0, bool in_strip, int seg,
                                         int y0, int y1, bool& flagged,
                                         const LocalStep ls = LocalStep{}) {
    constexpr bool LAZY = MODE == MODE_RK3_1;  // stage input U1 = L0*c + U0, formed on the fly
    constexpr int NV = Eq::nvar;
    constexpr int G = Recon::ghosts;
    // lazy stage: U^n of rows r, r+1 stays in registers from the lazy load
    // that formed the stage input there, instead of being loaded again for
    // the combination
    constexpr bool KEEP_U0 = LAZY && DIM >= 2;
    constexpr bool CHECK = MODE == MODE_RK2_1 || MODE == MODE_RK3_2;  // the step's last stage
    constexpr bool PRIM = recon_prim<Recon>::value;
    constexpr bool FW = recon_face_weights<Recon>::value && !PRIM;
    const int lane = threadIdx.x & 63;
    const SampleDesc D = [&] {
        if constexpr (UNI)
            return scalar_load(a.desc + sample);
        else
            return a.desc[sample];
    }();
    SampleState* Sp = a.state + sample;
    const bool active = in_strip && (LOCAL || runnable(*Sp));
    if (!__any(active)) return;  // wave-uniform
    const bool need_dt = LAZY || MODE != MODE_RHS;
    T c = T(0);
    if (need_dt && active)
        c = T((LOCAL ? ls.dt : sample_dt(D, *Sp, a.ws + 3 * sample)) / D.dx[0]);

    const int li = lane - lane0;
    const int col = x0 - G + li;  // interior x coordinate of this lane
    const bool interior = active && li >= G && li < G + ncols;
    const bool xface = active && li >= G - 1 && li < G + ncols;  // face col+1/2
    const int px = in_strip ? col + D.gx : D.gx;                 // padded column
    const ColumnImages cimg = column_images(D, col);
    // byte offsets of the lane's column and of its ghost images (UNI)
    const unsigned lo = unsigned(px) * unsigned(sizeof(T));
    // XLOAD: the x-neighbours of a row are loaded (UNI, stage input read
    // as stored) instead of moved across lanes: two loads per variable, no
    // VALU, issued one row ahead; the lanes at the ends of the padded row
    // read their own cell (they are halo lanes, whose neighbours are unused)
    constexpr bool XLOAD = UNI && !LAZY && !PRIM && G >= 2;
    const unsigned lom = unsigned(px > 0 ? px - 1 : px) * unsigned(sizeof(T));
    const unsigned lop = unsigned(px < D.tx - 1 ? px + 1 : px) * unsigned(sizeof(T));
    const int xbo = cimg.xb >= 0 ? cimg.xb * int(sizeof(T)) : -1;
    const int xdo = cimg.xd >= 0 ? cimg.xd * int(sizeof(T)) : -1;
    auto make_in = [&](bool second) {
        if constexpr (UNI)
            return uni_in<OP, NV>(a, D, second);
        else
            return stage_in<OP>(a, D, second);
    };
    const auto in = make_in(false);
    const auto in2 = make_in(true);
    const auto out = [&] {
        if constexpr (UNI)
            return uni_out<OP, NV>(a, D);
        else
            return stage_out<OP>(a, D);
    }();

    T wsx = T(0), wsy = T(0);

    // Rows are addressed in padded coordinates; loads outside the volume
    // are clamped (they only occur for G == 1, whose neighbours are unused).
    auto row_at = [&](int py) {
        if constexpr (UNI)
            return UniIdx{unsigned(py) * unsigned(D.tx) * unsigned(sizeof(T)), lo};
        else
            return (long long)py * D.tx + px;
    };
    auto row_index = [&](int py) {
        py = py < 0 ? 0 : (py >= D.ty ? D.ty - 1 : py);
        return row_at(py);
    };

    T yc[NV], yp[NV];  // rows r, r+1 of the stage input
    T u_yc[NV], u_yp[NV];  // KEEP_U0: U^n of rows r, r+1
    T xmn[NV], xpn[NV];  // XLOAD: x-neighbours of the next row to sweep
    auto load_xn = [&](int py) {
        if constexpr (XLOAD) {
            const UniIdx im{unsigned(py) * unsigned(D.tx) * unsigned(sizeof(T)), lom};
            const UniIdx ip{im.row, lop};
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                xmn[v] = in(v, im);
                xpn[v] = in(v, ip);
            }
        }
    };
    T Ry[NV];          // right (r+1/2) reconstruction of row r in y
    T Gm[NV];          // y face flux at r-1/2
    T wy = T(0);       // FW: the weight of face r+1/2 in y
    // reconstructions from primitive variables: those of rows r, r+1,
    // computed once per row and shared by the x and y stencils
    T Pc[NV], Pp[NV];
    const int gy = D.gy;
    if constexpr (DIM >= 2) {
        const int pr = y0 + gy;
        T qa[NV], ym[NV];
        load_cell<T, NV>(in, in2, LAZY, c, row_index(pr - 2), qa);
        load_cell<T, NV>(in, in2, LAZY, c, row_index(pr - 1), ym);
        if constexpr (KEEP_U0) {
            // rows pr, pr + 1 <= ny + gy lie inside the padded volume
            PendingCell<T, NV> p0, p1;
            p0.issue(in, in2, LAZY, row_at(pr));
            p1.issue(in, in2, LAZY, row_at(pr + 1));
            p0.finish(LAZY, c, yc);
            p1.finish(LAZY, c, yp);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                u_yc[v] = p0.b[v];
                u_yp[v] = p1.b[v];
            }
        } else {
            load_cell<T, NV>(in, in2, LAZY, c, row_index(pr), yc);
            load_cell<T, NV>(in, in2, LAZY, c, row_index(pr + 1), yp);
        }
        T Lt[NV], Rt[NV], L0[NV], R0[NV];
        if constexpr (PRIM) {
            T Pa[NV], Pm[NV];
            Recon::template prim<T, NV>(qa, Pa, eq.gamma);
            Recon::template prim<T, NV>(ym, Pm, eq.gamma);
            Recon::template prim<T, NV>(yc, Pc, eq.gamma);
            Recon::template prim<T, NV>(yp, Pp, eq.gamma);
            Recon::template apply_prim<T, NV>(Pa, Pm, Pc, Lt, Rt, eq.gamma);  // row r-1
            Recon::template apply_prim<T, NV>(Pm, Pc, Pp, L0, R0, eq.gamma);  // row r
        } else if constexpr (FW) {
            const T wa = Recon::face_weight(qa[0], ym[0]);  // face r-3/2
            const T wm = Recon::face_weight(ym[0], yc[0]);  // face r-1/2
            wy = Recon::face_weight(yc[0], yp[0]);          // face r+1/2
            Recon::template apply_w<T, NV>(qa, ym, yc, wm, wa, Lt, Rt);  // row r-1: keep R
            Recon::template apply_w<T, NV>(ym, yc, yp, wy, wm, L0, R0);  // row r
        } else {
            Recon::template apply<T, NV>(qa, ym, yc, Lt, Rt, eq.gamma);  // row r-1: keep R
            Recon::template apply<T, NV>(ym, yc, yp, L0, R0, eq.gamma);  // row r
        }
        wsy = ws_max(wsy, face_flux<T, Eq, Flux, 1>(eq, Rt, L0, Gm));
#pragma unroll
        for (int v = 0; v < NV; ++v) Ry[v] = R0[v];
    } else {
        load_cell<T, NV>(in, in2, LAZY, c, row_index(0), yc);
        if constexpr (PRIM) Recon::template prim<T, NV>(yc, Pc, eq.gamma);
    }

    load_xn(y0 + gy);
    for (int r = y0; r < y1; ++r) {
        const int pr = r + gy;
        T qn[NV], Pn[NV];
        // next row first: its latency overlaps the x sweep
        PendingCell<T, NV> pn;
        if constexpr (DIM >= 2) pn.issue(in, in2, LAZY, row_index(pr + 2));
        // U^n of this row for the stage combination, also issued before
        // the sweep (loaded after it, its latency would stall the wave);
        // every lane's column lies inside the padded row
        T u0[NV];
        if constexpr (KEEP_U0) {
            // loaded at row_index(pr) two rows ago: the same cell (an
            // interior row lies gy or more rows inside the padded volume,
            // where no clamping reaches)
#pragma unroll
            for (int v = 0; v < NV; ++v) u0[v] = u_yc[v];
        } else if constexpr (MODE != MODE_RHS) {
            const auto i0 = row_at(pr);
#pragma unroll
            for (int v = 0; v < NV; ++v) u0[v] = in2(v, i0);
        }
        T out_v[NV];
        // ---- x direction on row r (first: out = 0 + x divergence)
        {
            T xm[NV], xp[NV];
            const T* xc = PRIM ? Pc : yc;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                xm[v] = XLOAD ? xmn[v] : lane_from_lower(xc[v]);
                xp[v] = XLOAD ? xpn[v] : lane_from_upper(xc[v]);
            }
            T Lx[NV], Rx[NV];
            if constexpr (PRIM) {
                Recon::template apply_prim<T, NV>(xm, Pc, xp, Lx, Rx, eq.gamma);
            } else if constexpr (FW) {
                // face col+1/2 of this lane; face col-1/2 is the lower lane's
                const T wf = Recon::face_weight(yc[0], xp[0]);
                Recon::template apply_w<T, NV>(xm, yc, xp, wf, lane_from_lower(wf), Lx, Rx);
            } else {
                Recon::template apply<T, NV>(xm, yc, xp, Lx, Rx, eq.gamma);
            }
            // the next row's x-neighbours, in flight during the x flux and
            // the y sweep (after the last row this reads row y1 + gy <= ty - 1,
            // a ghost row). Issued only once this row's are consumed: the
            // loads then write the registers that held them, no copies.
            if constexpr (XLOAD && DIM >= 2) {
                __builtin_amdgcn_sched_barrier(0);
                load_xn(pr + 1);
            }
            T Ln[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) Ln[v] = lane_from_upper(Lx[v]);
            T Fp[NV];
            wsx = ws_max(wsx, face_flux<T, Eq, Flux, 0>(eq, Rx, Ln, Fp));
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const T Fm = lane_from_lower(Fp[v]);
                const T t0 = (T(-1.0) * Fp[v]) + Fm;
                out_v[v] = T(0) + t0;
            }
        }
        // ---- y direction: out += (-G(r+1/2)) + G(r-1/2)
        T wn = T(0);
        if constexpr (DIM >= 2) {
            pn.finish(LAZY, c, qn);
            T Ln[NV], Rn[NV];
            if constexpr (PRIM) {
                Recon::template prim<T, NV>(qn, Pn, eq.gamma);
                Recon::template apply_prim<T, NV>(Pc, Pp, Pn, Ln, Rn, eq.gamma);  // row r+1
            } else if constexpr (FW) {
                wn = Recon::face_weight(yp[0], qn[0]);  // face r+3/2
                Recon::template apply_w<T, NV>(yc, yp, qn, wn, wy, Ln, Rn);  // row r+1
            } else {
                Recon::template apply<T, NV>(yc, yp, qn, Ln, Rn, eq.gamma);  // row r+1
            }
            T Gp[NV];
            wsy = ws_max(wsy, face_flux<T, Eq, Flux, 1>(eq, Ry, Ln, Gp));
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                out_v[v] = out_v[v] + ((T(-1.0) * Gp[v]) + Gm[v]);
                Gm[v] = Gp[v];
                Ry[v] = Rn[v];
            }
        }
        if (interior) {
            T res[NV];
            combine<T, NV>(MODE, c, out_v, yc, u0, res);
            if (CHECK && !eq.valid(res)) {
                if constexpr (LOCAL)
                    *ls.status = 2;
                else
                    atomicOr(&Sp->status, 2);
                flagged = true;
            }
            if constexpr (UNI)
                store_result_uni<T, NV, DIM, OP == 0>(out, D, lo, xbo, xdo, r, 0, res);
            else
                store_result<T, NV, DIM, OP == 0>(out, D, cimg, r, 0, res);
        }
        if constexpr (DIM >= 2) {
            wy = wn;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                yc[v] = yp[v];
                yp[v] = qn[v];
                if constexpr (KEEP_U0) {
                    u_yc[v] = u_yp[v];
                    u_yp[v] = pn.b[v];
                }
                if constexpr (PRIM) {
                    Pc[v] = Pp[v];
                    Pp[v] = Pn[v];
                }
            }
        }
    }

    if (MODE == MODE_RHS && !LAZY && !a.no_ws) {
        // the lanes' running maxima cover every row of their faces: only
        // the faces of the strip count (x: col+1/2 for the lanes from G-1,
        // y: the interior columns)
        const T mx = segment_max(xface ? wsx : T(0), seg);
        const T my = segment_max(interior ? wsy : T(0), seg);
        const bool leader = in_strip && li == 0;
        if constexpr (LOCAL) {
            if (leader) {
                atomicMax(ls.ws + 0, (unsigned long long)__double_as_longlong(double(mx)));
                if (DIM >= 2)
                    atomicMax(ls.ws + 1, (unsigned long long)__double_as_longlong(double(my)));
            }
        } else {
            const bool run = leader && runnable(*Sp);
            report_wave_speed(a, task, seg, sample, 0, double(mx), leader, run);
            if (DIM >= 2) report_wave_speed(a, task, seg, sample, 1, double(my), leader, run);
        }
    }
}

// 1D and 2D stage kernel (DIM = active grid dimensions). WPS is the
// register budget expressed as waves per SIMD (2: no spills for FP64 Euler
// 2D; 3 and 4 trade spills for occupancy); the autotuner picks one. OP = 1
// is the volume operator (computeFlux on a reference Volume's per-variable
// memories over a region; the result cells only, no ghost images).
// A task whose wave holds one segment (every full-width strip) runs the
// single-sample march with scalar addressing; packed narrow strips of
// several samples run the per-lane one.
template <class T, class Eq, class Recon, class Flux, int DIM, int WPS, int MODE, int OP = 0>
__global__ __launch_bounds__(256, WPS) void stage_kernel_12(StageArgs<T> a) {
    constexpr int G = Recon::ghosts;
    constexpr bool CHECK = MODE == MODE_RK2_1 || MODE == MODE_RK3_2;  // the step's last stage
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((xcd_block() * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    Eq eq;
    eq.gamma = T(a.gamma);

    const int ntasks = a.ntasks_dev ? *a.ntasks_dev : a.ntasks;
    bool flagged = false;  // this thread set a sample's failure status
    for (int w = wave; w < ntasks; w += nwaves) {
        const WaveTask tk = scalar_load(a.tasks + w);
        if (tk.nseg == 1 && uni_fits<T, Eq::nvar>(a, tk.sample[0])) {
            const bool in_strip = lane >= tk.lane0[0] && lane < tk.lane0[0] + tk.ncols[0] + 2 * G;
            march_12<T, Eq, Recon, Flux, DIM, MODE, OP, true>(
                a, eq, w, tk.sample[0], tk.x0[0], tk.ncols[0], tk.lane0[0], in_strip,
                in_strip ? 0 : -1, tk.y0, tk.y1, flagged);
        } else {
            // this lane's segment
            int seg = -1, sample = tk.sample[0], x0 = tk.x0[0], ncols = tk.ncols[0],
                lane0 = tk.lane0[0];
#pragma unroll
            for (int k = 0; k < kMaxSeg; ++k) {
                if (k < tk.nseg && lane >= tk.lane0[k] &&
                    lane < tk.lane0[k] + tk.ncols[k] + 2 * G) {
                    seg = k;
                    sample = tk.sample[k];
                    x0 = tk.x0[k];
                    ncols = tk.ncols[k];
                    lane0 = tk.lane0[k];
                }
            }
            march_12<T, Eq, Recon, Flux, DIM, MODE, OP, false>(a, eq, w, sample, x0, ncols, lane0,
                                                               seg >= 0, seg, tk.y0, tk.y1,
                                                               flagged);
        }
    }
    if constexpr (CHECK) end_step(a.end, flagged);
}

// 3D stage kernel. A workgroup of kGroupRows wavefronts owns the x-strips of
// rows y0 .. y0+kGroupRows-1 (one row per wavefront) and marches in z with a
// register window (z neighbours); x neighbours come from adjacent lanes. The
// y direction is shared through LDS, plane by plane: every wavefront writes
// its row of the stage input (the first and last also the two halo rows on
// their side), reconstructs its own row in y and publishes the left face
// value, forms the flux of the face above its row from the published value of
// the row above and publishes it for that row. Each y face flux and each
// reconstruction is thus evaluated once (the halo rows' once more per group)
// instead of per row.
constexpr int kGroupRows = 4;

// One task of the 3D stage kernel, run by the whole workgroup: wave w
// marches row y0 + w of the strip [x0, x0 + ncols) of `sample` (lanes
// [lane0, lane0 + ncols + 2G)) over planes [z0, z1). UNI as in march_12:
// the workgroup's lanes all belong to one sample, addressed from scalar
// bases through buffer resources. s_rows, s_L, s_G: the group's LDS rows.
template <class T, class Eq, class Recon, class Flux, int MODE, int OP, bool UNI, int NV = Eq::nvar>
__device__ __forceinline__ void march_3(const StageArgs<T>& a, const Eq& eq, int task, int sample,
                                        int x0, int ncols, int lane0, bool in_strip, int seg, int y0,
                                        int y1, int z0, int z1, T (*s_rows)[NV][64],
                                        T (*s_L)[NV][64], T (*s_G)[NV][64], bool& flagged) {
    constexpr bool LAZY = MODE == MODE_RK3_1;
    constexpr bool CHECK = MODE == MODE_RK2_1 || MODE == MODE_RK3_2;  // the step's last stage
    // FP32 lazy stage: U^n of the two planes ahead of the combination stays
    // in registers from the lazy load that formed the stage input there,
    // instead of being loaded again for the combination (FP64 keeps the
    // reload: twice the registers)
    constexpr bool KEEP_U0 = LAZY && std::is_same_v<T, float>;
    constexpr int G = Recon::ghosts;
    constexpr int NW = kGroupRows;
    constexpr bool FW = recon_face_weights<Recon>::value;
    // the x-neighbours of the next plane loaded (single-sample groups, stage
    // input read as stored), one plane ahead, instead of moved across lanes
    constexpr bool XLOAD = UNI && !LAZY && G >= 2;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const SampleDesc D = [&] {
        if constexpr (UNI)
            return scalar_load(a.desc + sample);
        else
            return a.desc[sample];
    }();
    SampleState* Sp = a.state + sample;
    const bool active = in_strip && runnable(*Sp);
    // the samples of a task are the same for every wavefront of the group
    if (!__syncthreads_or(active)) return;
    const bool need_dt = LAZY || MODE != MODE_RHS;
    T c = T(0);
    if (need_dt && active) c = T(sample_dt(D, *Sp, a.ws + 3 * sample) / D.dx[0]);

    const int li = lane - lane0;
    const int row = y0 + w;            // this wavefront's row
    const bool row_in = row < y1;      // rows past the group end only feed neighbours
    const int col = x0 - G + li;
    const bool lanes_in = active && li >= G && li < G + ncols;
    const bool interior = lanes_in && row_in;
    const bool xface = active && row_in && li >= G - 1 && li < G + ncols;
    const int px = in_strip ? col + D.gx : D.gx;
    const int py = row + D.gy;
    const ColumnImages cimg = column_images(D, col);
    const unsigned lo = unsigned(px) * unsigned(sizeof(T));
    const unsigned lom = unsigned(px > 0 ? px - 1 : px) * unsigned(sizeof(T));
    const unsigned lop = unsigned(px < D.tx - 1 ? px + 1 : px) * unsigned(sizeof(T));
    const int xbo = cimg.xb >= 0 ? cimg.xb * int(sizeof(T)) : -1;
    const int xdo = cimg.xd >= 0 ? cimg.xd * int(sizeof(T)) : -1;
    const long long plane_stride = (long long)D.tx * D.ty;
    auto make_in = [&](bool second) {
        if constexpr (UNI)
            return uni_in<OP, NV>(a, D, second);
        else
            return stage_in<OP>(a, D, second);
    };
    const auto in = make_in(false);
    const auto in2 = make_in(true);
    const auto out = [&] {
        if constexpr (UNI)
            return uni_out<OP, NV>(a, D);
        else
            return stage_out<OP>(a, D);
    }();

    T wsx = T(0), wsy = T(0), wsz = T(0);
    // cell (padded row pyy, padded plane pz) of the lane's column, clamped
    // to the padded volume (rows past the group end and the planes of the
    // prologue's far halo)
    auto idx = [&](int pyy, int pz) {
        pz = pz < 0 ? 0 : (pz >= D.tz ? D.tz - 1 : pz);
        pyy = pyy < 0 ? 0 : (pyy >= D.ty ? D.ty - 1 : pyy);
        if constexpr (UNI)
            return UniIdx{(unsigned(pz) * unsigned(D.ty) + unsigned(pyy)) * unsigned(D.tx) *
                              unsigned(sizeof(T)),
                          lo};
        else
            return (long long)pz * plane_stride + (long long)pyy * D.tx + px;
    };
    // halo rows: the first wavefront loads y0-2, y0-1, the last y0+NW, y0+NW+1
    const bool lo_halo = w == 0;
    const bool hi_halo = w == NW - 1;
    const int h0 = lo_halo ? y0 - 2 + D.gy : y0 + NW + D.gy;

    T zc[NV], zp[NV], Rz[NV], Hm[NV];
    T u_zc[NV], u_zp[NV];  // KEEP_U0: U^n at planes zc, zp
    T xmn[NV], xpn[NV];    // XLOAD: x-neighbours of the next plane's row
    T wz = T(0);           // FW: the weight of the z face above plane zc
    auto load_xn = [&](int pz) {
        if constexpr (XLOAD) {
            const UniIdx c0 = idx(py, pz);
            const UniIdx im{c0.row, lom}, ip{c0.row, lop};
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                xmn[v] = in(v, im);
                xpn[v] = in(v, ip);
            }
        }
    };
    {
        const int pz = z0 + D.gz;
        T qa[NV], zm[NV];
        load_cell<T, NV>(in, in2, LAZY, c, idx(py, pz - 2), qa);
        load_cell<T, NV>(in, in2, LAZY, c, idx(py, pz - 1), zm);
        if constexpr (KEEP_U0) {
            PendingCell<T, NV> p0, p1;
            p0.issue(in, in2, LAZY, idx(py, pz));
            p1.issue(in, in2, LAZY, idx(py, pz + 1));
            p0.finish(LAZY, c, zc);
            p1.finish(LAZY, c, zp);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                u_zc[v] = p0.b[v];
                u_zp[v] = p1.b[v];
            }
        } else {
            load_cell<T, NV>(in, in2, LAZY, c, idx(py, pz), zc);
            load_cell<T, NV>(in, in2, LAZY, c, idx(py, pz + 1), zp);
        }
        load_xn(pz);
        T Lt[NV], Rt[NV], L0[NV], R0[NV];
        if constexpr (FW) {
            const T wa = Recon::face_weight(qa[0], zm[0]);  // face pz-3/2
            const T wm = Recon::face_weight(zm[0], zc[0]);  // face pz-1/2
            wz = Recon::face_weight(zc[0], zp[0]);          // face pz+1/2
            Recon::template apply_w<T, NV>(qa, zm, zc, wm, wa, Lt, Rt);
            Recon::template apply_w<T, NV>(zm, zc, zp, wz, wm, L0, R0);
        } else {
            Recon::template apply<T, NV>(qa, zm, zc, Lt, Rt, eq.gamma);
            Recon::template apply<T, NV>(zm, zc, zp, L0, R0, eq.gamma);
        }
        wsz = ws_max(wsz, face_flux<T, Eq, Flux, 2>(eq, Rt, L0, Hm));
#pragma unroll
        for (int v = 0; v < NV; ++v) Rz[v] = R0[v];
    }

    for (int zz = z0; zz < z1; ++zz) {
        const int pz = zz + D.gz;
        T qn[NV];
        PendingCell<T, NV> pq, ph0, ph1;
        pq.issue(in, in2, LAZY, idx(py, pz + 2));
        if (lo_halo || hi_halo) {
            ph0.issue(in, in2, LAZY, idx(h0, pz));
            ph1.issue(in, in2, LAZY, idx(h0 + 1, pz));
        }
        T u0[NV];
        if constexpr (KEEP_U0) {
            // loaded at idx(py, pz) two planes ago: the same cell (no
            // clamping reaches the interior planes, which lie G or more
            // planes inside the padded volume)
#pragma unroll
            for (int v = 0; v < NV; ++v) u0[v] = u_zc[v];
        } else if constexpr (MODE != MODE_RHS && UNI) {
            // every lane's cell (clamped rows past the group end) lies in
            // the volume: loaded unconditionally, before the sweep
            const auto i0 = idx(py, pz);
#pragma unroll
            for (int v = 0; v < NV; ++v) u0[v] = in2(v, i0);
        } else if constexpr (MODE != MODE_RHS) {
            if (interior) {
                const long long i0 = (long long)pz * plane_stride + (long long)py * D.tx + px;
#pragma unroll
                for (int v = 0; v < NV; ++v) u0[v] = in2(v, i0);
            }
        }
        T out_v[NV];
        {  // x
            T xm[NV], xp[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                xm[v] = XLOAD ? xmn[v] : lane_from_lower(zc[v]);
                xp[v] = XLOAD ? xpn[v] : lane_from_upper(zc[v]);
            }
            T Lx[NV], Rx[NV];
            if constexpr (FW) {
                const T wf = Recon::face_weight(zc[0], xp[0]);  // face col+1/2
                Recon::template apply_w<T, NV>(xm, zc, xp, wf, lane_from_lower(wf), Lx, Rx);
            } else {
                Recon::template apply<T, NV>(xm, zc, xp, Lx, Rx, eq.gamma);
            }
            // the next plane's x-neighbours, in flight during the rest of
            // the plane (after the last plane this reads plane z1 + gz <=
            // tz - 1, a ghost plane); issued once this plane's are consumed
            if constexpr (XLOAD) {
                __builtin_amdgcn_sched_barrier(0);
                load_xn(pz + 1);
            }
            T Ln[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) Ln[v] = lane_from_upper(Lx[v]);
            T Fp[NV];
            wsx = ws_max(wsx, face_flux<T, Eq, Flux, 0>(eq, Rx, Ln, Fp));
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const T Fm = lane_from_lower(Fp[v]);
                out_v[v] = T(0) + ((T(-1.0) * Fp[v]) + Fm);
            }
        }
        // ---- y, phase 1: publish the plane's rows
#pragma unroll
        for (int v = 0; v < NV; ++v) s_rows[w + 2][v][lane] = zc[v];
        if (lo_halo || hi_halo) {
            T q0[NV], q1[NV];
            ph0.finish(LAZY, c, q0);
            ph1.finish(LAZY, c, q1);
            const int k0 = lo_halo ? 0 : NW + 2;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                s_rows[k0][v][lane] = q0[v];
                s_rows[k0 + 1][v][lane] = q1[v];
            }
        }
        __syncthreads();
        // ---- y, phase 2: reconstruct the own row (and the halo row on the edge)
        T L0[NV], R0[NV], Rb[NV];
        {
            T ym1[NV], yp1[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                ym1[v] = s_rows[w + 1][v][lane];
                yp1[v] = s_rows[w + 3][v][lane];
            }
            T wy1 = T(0), wy2 = T(0);  // FW: faces below and above the own row
            if constexpr (FW) {
                wy1 = Recon::face_weight(ym1[0], zc[0]);
                wy2 = Recon::face_weight(zc[0], yp1[0]);
                Recon::template apply_w<T, NV>(ym1, zc, yp1, wy2, wy1, L0, R0);
            } else {
                Recon::template apply<T, NV>(ym1, zc, yp1, L0, R0, eq.gamma);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) s_L[w][v][lane] = L0[v];
            if (lo_halo) {  // right value of row y0-1
                T q0[NV], Lb[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) q0[v] = s_rows[0][v][lane];
                if constexpr (FW)
                    Recon::template apply_w<T, NV>(q0, ym1, zc, wy1,
                                                   Recon::face_weight(q0[0], ym1[0]), Lb, Rb);
                else
                    Recon::template apply<T, NV>(q0, ym1, zc, Lb, Rb, eq.gamma);
            }
            if (hi_halo) {  // left value of row y0+NW
                T q1[NV], La[NV], Ra[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) q1[v] = s_rows[NW + 3][v][lane];
                if constexpr (FW)
                    Recon::template apply_w<T, NV>(zc, yp1, q1, Recon::face_weight(yp1[0], q1[0]),
                                                   wy2, La, Ra);
                else
                    Recon::template apply<T, NV>(zc, yp1, q1, La, Ra, eq.gamma);
#pragma unroll
                for (int v = 0; v < NV; ++v) s_L[NW][v][lane] = La[v];
            }
        }
        // ---- z (own row only; placed here to overlap the exchange)
        T zpart[NV];
        T wn = T(0);
        {
            pq.finish(LAZY, c, qn);
            T Ln[NV], Rn[NV];
            if constexpr (FW) {
                wn = Recon::face_weight(zp[0], qn[0]);  // face pz+3/2
                Recon::template apply_w<T, NV>(zc, zp, qn, wn, wz, Ln, Rn);
            } else {
                Recon::template apply<T, NV>(zc, zp, qn, Ln, Rn, eq.gamma);
            }
            T Hp[NV];
            wsz = ws_max(wsz, face_flux<T, Eq, Flux, 2>(eq, Rz, Ln, Hp));
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                zpart[v] = (T(-1.0) * Hp[v]) + Hm[v];
                Hm[v] = Hp[v];
                Rz[v] = Rn[v];
            }
        }
        __syncthreads();
        // ---- y, phase 3: the flux of the face above the own row
        T Gp[NV], Gm[NV];
        {
            T La[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) La[v] = s_L[w + 1][v][lane];
            wsy = ws_max(wsy, face_flux<T, Eq, Flux, 1>(eq, R0, La, Gp));
#pragma unroll
            for (int v = 0; v < NV; ++v) s_G[w + 1][v][lane] = Gp[v];
            if (lo_halo)  // the face below row y0
                wsy = ws_max(wsy, face_flux<T, Eq, Flux, 1>(eq, Rb, L0, Gm));
        }
        __syncthreads();
        if (!lo_halo) {
#pragma unroll
            for (int v = 0; v < NV; ++v) Gm[v] = s_G[w][v][lane];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            out_v[v] = out_v[v] + ((T(-1.0) * Gp[v]) + Gm[v]);
            out_v[v] = out_v[v] + zpart[v];
        }
        if (interior) {
            T res[NV];
            combine<T, NV>(MODE, c, out_v, zc, u0, res);
            if (CHECK && !eq.valid(res)) {
                atomicOr(&Sp->status, 2);
                flagged = true;
            }
            if constexpr (UNI)
                store_result_uni<T, NV, 3, OP == 0>(out, D, lo, xbo, xdo, row, zz, res);
            else
                store_result<T, NV, 3, OP == 0>(out, D, cimg, row, zz, res);
        }
        wz = wn;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            zc[v] = zp[v];
            zp[v] = qn[v];
            if constexpr (KEEP_U0) {
                u_zc[v] = u_zp[v];
                u_zp[v] = pq.b[v];
            }
        }
    }

    if (MODE == MODE_RHS && !LAZY && !a.no_ws) {
        // the lanes' running maxima over the planes: only the strip's
        // faces count (x: col+1/2 from lane G-1; y, z: interior cells)
        const T mx = segment_max(xface ? wsx : T(0), seg);
        const T my = segment_max(interior ? wsy : T(0), seg);
        const T mz = segment_max(interior ? wsz : T(0), seg);
        const bool leader = in_strip && li == 0, run = leader && runnable(*Sp) && row_in;
        const long long slot = (long long)task * NW + w;
        report_wave_speed(a, slot, seg, sample, 0, double(mx), leader, run);
        report_wave_speed(a, slot, seg, sample, 1, double(my), leader, run);
        report_wave_speed(a, slot, seg, sample, 2, double(mz), leader, run);
    }
}

template <class T, class Eq, class Recon, class Flux, int WPS, int MODE, int OP = 0>
__global__ __launch_bounds__(64 * kGroupRows, WPS) void stage_kernel_3(StageArgs<T> a) {
    constexpr bool CHECK = MODE == MODE_RK2_1 || MODE == MODE_RK3_2;  // the step's last stage
    constexpr int NV = Eq::nvar;
    constexpr int G = Recon::ghosts;
    constexpr int NW = kGroupRows;
    __shared__ T s_rows[NW + 4][NV][64];  // rows y0-2 .. y0+NW+1 of the plane's stage input
    __shared__ T s_L[NW + 1][NV][64];     // y-left face values of rows y0 .. y0+NW
    __shared__ T s_G[NW + 1][NV][64];     // y fluxes of the faces above rows y0-1 .. y0+NW-1
    const int lane = threadIdx.x & 63;
    Eq eq;
    eq.gamma = T(a.gamma);

    const int ntasks = a.ntasks_dev ? *a.ntasks_dev : a.ntasks;
    bool flagged = false;  // this thread set a sample's failure status
    for (int t = xcd_block(); t < ntasks; t += gridDim.x) {  // uniform over the workgroup
        const WaveTask tk = scalar_load(a.tasks + t);
        if (tk.nseg == 1 && uni_fits<T, NV>(a, tk.sample[0])) {
            const bool in_strip = lane >= tk.lane0[0] && lane < tk.lane0[0] + tk.ncols[0] + 2 * G;
            march_3<T, Eq, Recon, Flux, MODE, OP, true>(
                a, eq, t, tk.sample[0], tk.x0[0], tk.ncols[0], tk.lane0[0], in_strip,
                in_strip ? 0 : -1, tk.y0, tk.y1, tk.z0, tk.z1, s_rows, s_L, s_G, flagged);
        } else {
            int seg = -1, sample = tk.sample[0], x0 = tk.x0[0], ncols = tk.ncols[0],
                lane0 = tk.lane0[0];
#pragma unroll
            for (int k = 0; k < kMaxSeg; ++k) {
                if (k < tk.nseg && lane >= tk.lane0[k] &&
                    lane < tk.lane0[k] + tk.ncols[k] + 2 * G) {
                    seg = k;
                    sample = tk.sample[k];
                    x0 = tk.x0[k];
                    ncols = tk.ncols[k];
                    lane0 = tk.lane0[k];
                }
            }
            march_3<T, Eq, Recon, Flux, MODE, OP, false>(a, eq, t, sample, x0, ncols, lane0, seg >= 0,
                                                         seg, tk.y0, tk.y1, tk.z0, tk.z1, s_rows,
                                                         s_L, s_G, flagged);
        }
        __syncthreads();  // the next task reuses the LDS rows
    }
    if constexpr (CHECK) end_step(a.end, flagged);
}

// Stage kernel for reconstructions wider than three cells (ENO3, ENO4,
// WENO3): the same tasks, arithmetic, outputs and wave-speed reduction as the
// marching kernels, with the stencils read from memory (L1/L2) instead of a
// register window. x: every lane of the strip (and the halo lane each side)
// reconstructs its cell and evaluates the flux of its right face; the left
// face's flux comes from the lane below (DPP), as in the marching kernels.
// The wavefront marches along y (2D) or z (3D), keeping the last face flux
// and right face value, so each face flux in that direction is evaluated
// once; in 3D the y faces of a row are evaluated by the row itself (twice
// per face). The divergence is summed in the marching kernels' order:
// ((0 + x part) + y part) + z part.
template <class T, class Eq, class Recon, class Flux, int DIM, int WPS, int MODE, int OP = 0>
__global__ __launch_bounds__(256, WPS) void stage_kernel_wide(StageArgs<T> a) {
    constexpr bool LAZY = MODE == MODE_RK3_1;
    constexpr bool CHECK = MODE == MODE_RK2_1 || MODE == MODE_RK3_2;  // the step's last stage
    constexpr int NV = Eq::nvar;
    constexpr int G = Recon::ghosts;
    constexpr int RAD = recon_radius<Recon>::value;
    constexpr int MD = DIM == 3 ? 2 : 1;  // marched direction (DIM >= 2)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((xcd_block() * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    Eq eq;
    eq.gamma = T(a.gamma);

    const int ntasks = a.ntasks_dev ? *a.ntasks_dev : a.ntasks;
    bool flagged = false;  // this thread set a sample's failure status
    for (int w = wave; w < ntasks; w += nwaves) {
        const WaveTask& tk = a.tasks[w];
        int seg = -1;
        for (int k = 0; k < tk.nseg; ++k)
            if (lane >= tk.lane0[k] && lane < tk.lane0[k] + tk.ncols[k] + 2 * G) seg = k;
        const bool in_strip = seg >= 0;
        const int sk = in_strip ? seg : 0;
        const int sample = tk.sample[sk];
        const int li = lane - tk.lane0[sk];
        const int ncols = tk.ncols[sk];
        const SampleDesc D = a.desc[sample];
        SampleState* Sp = a.state + sample;
        const bool active = in_strip && runnable(*Sp);
        if (!__any(active)) continue;  // wave-uniform
        const bool need_dt = LAZY || MODE != MODE_RHS;
        T c = T(0);
        if (need_dt && active) c = T(sample_dt(D, *Sp, a.ws + 3 * sample) / D.dx[0]);
        const int col = tk.x0[sk] - G + li;
        const bool interior = active && li >= G && li < G + ncols;
        const bool xlane = active && li >= G - 1 && li <= G + ncols;  // reconstructs in x
        const bool xface = active && li >= G - 1 && li < G + ncols;   // face col+1/2
        const ColumnImages cimg = column_images(D, col);
        const auto in = stage_in<OP>(a, D, false);
        const auto in2 = stage_in<OP>(a, D, true);
        const auto out = stage_out<OP>(a, D);
        const long long sxd[3] = {1, (long long)D.tx, (long long)D.tx * D.ty};
        T ws[3] = {T(0), T(0), T(0)};

        // stencil of 2*RAD+1 cells around flat index i in direction d; the
        // centre cell (the stage input there) is left in q[RAD]
        auto recon_at = [&](auto dc, long long i, T (*q)[NV], T* L, T* R) {
            constexpr int d = decltype(dc)::value;
#pragma unroll
            for (int k = 0; k <= 2 * RAD; ++k)
                load_cell<T, NV>(in, in2, LAZY, c, i + (k - RAD) * sxd[d], q[k]);
            recon_stencil<Recon, T, NV>(q, L, R, eq.gamma);
        };

        const int o0 = DIM == 3 ? tk.y0 : 0, o1 = DIM == 3 ? tk.y1 : 1;
        const int m0 = DIM == 3 ? tk.z0 : (DIM == 2 ? tk.y0 : 0);
        const int m1 = DIM == 3 ? tk.z1 : (DIM == 2 ? tk.y1 : 1);
        for (int o = o0; o < o1; ++o) {  // uniform over the wavefront
            auto flat = [&](int m) -> long long {
                const int y = DIM == 3 ? o : (DIM == 2 ? m : 0);
                const int z = DIM == 3 ? m : 0;
                return ((long long)(z + D.gz) * D.ty + (y + D.gy)) * D.tx + (col + D.gx);
            };
            T Rprev[NV], Gm[NV];  // marched direction: right value of cell m-1, flux at m-1/2
            if constexpr (DIM >= 2) {
                if (interior) {
                    T q[2 * RAD + 1][NV], Lt[NV], L0[NV];
                    recon_at(std::integral_constant<int, MD>(), flat(m0 - 1), q, Lt, Rprev);
                    T Rt[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) Rt[v] = Rprev[v];
                    recon_at(std::integral_constant<int, MD>(), flat(m0), q, L0, Rprev);
                    const T s = face_flux<T, Eq, Flux, MD>(eq, Rt, L0, Gm);
                    ws[MD] = std_max(ws[MD], s);
                }
            }
            // 1D/2D: the x stencil comes from the neighbouring lanes (DPP
            // moves of the centre row, which every lane of a segment holds)
            // and the marched stencil from a register window of the rows
            // m+1-RAD .. m+1+RAD, so each row of the stage input is loaded
            // once; 3D loads its stencils (cached).
            constexpr bool WINDOW = DIM <= 2;
            const int wcol = in_strip ? col + D.gx : D.gx;  // a column inside the volume
            auto row_flat = [&](int m) -> long long {
                int py = (DIM == 2 ? m : 0) + D.gy;
                py = py < 0 ? 0 : (py >= D.ty ? D.ty - 1 : py);
                return (long long)py * D.tx + wcol;
            };
            T Wn[2 * RAD + 1][NV];  // window rows m+1-RAD .. m+1+RAD (DIM == 2); row m at RAD-1
            if constexpr (WINDOW) {
                if constexpr (DIM == 2) {
#pragma unroll
                    for (int k = 0; k <= 2 * RAD; ++k)
                        load_cell<T, NV>(in, in2, LAZY, c, row_flat(m0 + 1 - RAD + k), Wn[k]);
                } else {
                    load_cell<T, NV>(in, in2, LAZY, c, row_flat(0), Wn[RAD - 1]);
                }
            }
            for (int m = m0; m < m1; ++m) {  // uniform over the wavefront
                const long long i0 = flat(m);
                T out_v[NV], uin[NV];
                PendingCell<T, NV> pn;
                if constexpr (DIM == 2) pn.issue(in, in2, LAZY, row_flat(m + 2 + RAD));
                {  // x: own right face, left face from the lane below
                    T q[2 * RAD + 1][NV], Lx[NV], Rx[NV];
                    if constexpr (WINDOW) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            q[RAD][v] = Wn[RAD - 1][v];
#pragma unroll
                            for (int k = RAD - 1; k >= 0; --k) q[k][v] = lane_from_lower(q[k + 1][v]);
#pragma unroll
                            for (int k = RAD + 1; k <= 2 * RAD; ++k)
                                q[k][v] = lane_from_upper(q[k - 1][v]);
                            uin[v] = q[RAD][v];
                        }
                        // lanes outside [G-1, G+ncols] reconstruct from values
                        // past their segment; no face that is used reads them
                        recon_stencil<Recon, T, NV>(q, Lx, Rx, eq.gamma);
                    } else if (xlane) {
                        recon_at(std::integral_constant<int, 0>(), i0, q, Lx, Rx);
#pragma unroll
                        for (int v = 0; v < NV; ++v) uin[v] = q[RAD][v];
                    } else {
#pragma unroll
                        for (int v = 0; v < NV; ++v) Lx[v] = Rx[v] = uin[v] = T(1);
                    }
                    T Ln[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) Ln[v] = lane_from_upper(Lx[v]);
                    T Fp[NV];
                    const T s = face_flux<T, Eq, Flux, 0>(eq, Rx, Ln, Fp);
                    if (xface) ws[0] = std_max(ws[0], s);
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const T Fm = lane_from_lower(Fp[v]);
                        out_v[v] = T(0) + ((T(-1.0) * Fp[v]) + Fm);
                    }
                }
                if constexpr (DIM == 3) {  // y: both faces of the row
                    T part[NV];
                    if (interior) {
                        T q[2 * RAD + 1][NV], Lm[NV], Rm[NV], L0[NV], R0[NV], Lp[NV], Rp[NV];
                        recon_at(std::integral_constant<int, 1>(), i0 - sxd[1], q, Lm, Rm);
                        recon_at(std::integral_constant<int, 1>(), i0, q, L0, R0);
                        recon_at(std::integral_constant<int, 1>(), i0 + sxd[1], q, Lp, Rp);
                        T Fm[NV], Fp[NV];
                        const T sm = face_flux<T, Eq, Flux, 1>(eq, Rm, L0, Fm);
                        const T sp = face_flux<T, Eq, Flux, 1>(eq, R0, Lp, Fp);
                        ws[1] = std_max(std_max(ws[1], sm), sp);
#pragma unroll
                        for (int v = 0; v < NV; ++v) part[v] = (T(-1.0) * Fp[v]) + Fm[v];
                    } else {
#pragma unroll
                        for (int v = 0; v < NV; ++v) part[v] = T(0);
                    }
#pragma unroll
                    for (int v = 0; v < NV; ++v) out_v[v] = out_v[v] + part[v];
                }
                if constexpr (DIM >= 2) {  // marched direction
                    T Gp[NV];
                    if (interior) {
                        T Ln[NV], Rn[NV];
                        if constexpr (WINDOW) {
                            recon_stencil<Recon, T, NV>(Wn, Ln, Rn, eq.gamma);
                        } else {
                            T q[2 * RAD + 1][NV];
                            recon_at(std::integral_constant<int, MD>(), flat(m + 1), q, Ln, Rn);
                        }
                        const T s = face_flux<T, Eq, Flux, MD>(eq, Rprev, Ln, Gp);
                        ws[MD] = std_max(ws[MD], s);
#pragma unroll
                        for (int v = 0; v < NV; ++v) Rprev[v] = Rn[v];
                    } else {
#pragma unroll
                        for (int v = 0; v < NV; ++v) Gp[v] = T(0);
                    }
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        out_v[v] = out_v[v] + ((T(-1.0) * Gp[v]) + Gm[v]);
                        Gm[v] = Gp[v];
                    }
                }
                if (interior) {
                    T u0[NV], res[NV];
                    if (MODE != MODE_RHS) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) u0[v] = in2(v, i0);
                    }
                    combine<T, NV>(MODE, c, out_v, uin, u0, res);
                    if (CHECK && !eq.valid(res)) {
                        atomicOr(&Sp->status, 2);
                        flagged = true;
                    }
                    store_result<T, NV, DIM, OP == 0>(out, D, cimg, DIM == 3 ? o : m,
                                                      DIM == 3 ? m : 0, res);
                }
                if constexpr (DIM == 2) {  // advance the window by one row
#pragma unroll
                    for (int k = 0; k < 2 * RAD; ++k)
#pragma unroll
                        for (int v = 0; v < NV; ++v) Wn[k][v] = Wn[k + 1][v];
                    pn.finish(LAZY, c, Wn[2 * RAD]);
                }
            }
        }

        if (MODE == MODE_RHS && !LAZY && !a.no_ws) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const T m = segment_max(ws[d], seg);
                const bool leader = in_strip && li == 0;
                report_wave_speed(a, w, seg, sample, d, double(m), leader, leader && runnable(*Sp));
            }
        }
    }
    if constexpr (CHECK) end_step(a.end, flagged);
}

// ---- Reconstruction::performReconstruction on a reference Volume
// (ReconstructionCPU.cpp:38-97, ENOCPU.cpp:41-176, WENOCPU.cpp:32-186,
// NoReconstruction.cpp:42-73): the left (i-1/2) and right (i+1/2) face
// values of every cell of the padded range [lo, hi) along `direction`, from
// the same reconstruction arithmetic as the stage kernels. One thread per
// cell; the host checks that every stencil stays inside the volume.
template <class T>
struct ReconArgs {
    const T* vin[5];
    T* vleft[5];
    T* vright[5];
    int tx, ty, tz;      // padded extents
    int lo[3], hi[3];    // padded cell range per axis
    int direction;
    double gamma;
};

template <class T, class Eq, class Recon>
__global__ __launch_bounds__(256) void recon_volume_kernel(ReconArgs<T> a) {
    constexpr int NV = Eq::nvar;
    constexpr int RAD = recon_radius<Recon>::value;
    const int ex = a.hi[0] - a.lo[0], ey = a.hi[1] - a.lo[1], ez = a.hi[2] - a.lo[2];
    const long long n = (long long)ex * ey * ez;
    const long long st = a.direction == 0 ? 1LL
                         : (a.direction == 1 ? (long long)a.tx : (long long)a.tx * a.ty);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const int x = a.lo[0] + int(c % ex);
        const int y = a.lo[1] + int((c / ex) % ey);
        const int z = a.lo[2] + int(c / ((long long)ex * ey));
        const long long i = ((long long)z * a.ty + y) * a.tx + x;
        T L[NV], R[NV];
        if constexpr (std::is_same_v<Recon, ReconNone>) {
#pragma unroll
            for (int v = 0; v < NV; ++v) L[v] = R[v] = a.vin[v][i];
        } else {
            T q[2 * RAD + 1][NV];
#pragma unroll
            for (int k = 0; k <= 2 * RAD; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v) q[k][v] = a.vin[v][i + (k - RAD) * st];
            recon_stencil<Recon, T, NV>(q, L, R, T(a.gamma));
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            a.vleft[v][i] = L[v];
            a.vright[v][i] = R[v];
        }
    }
}

}  // namespace alsvinn_hip

// Internal declarations shared by the hip platform's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>
#include <vector>

#include "../../include/alsvinn_hip.h"
#include "ensemble_kernels.hpp"

namespace alsvinn_hip {

int fail(int code, const std::string& msg);
// Checks a scheme's option values and their combination (0 = valid).
int validate(const alsvinn_hip_scheme* s);
// Ghost layers of a scheme: max(flux stencil - 1, reconstruction ghosts)
// (NumericalFluxCPU.cpp:226-231).
int ghost_cells(const alsvinn_hip_scheme& s);

// Launchers of the kernels one scheme (equation, reconstruction, flux, grid
// dimension, precision) needs.
constexpr int kStageVariants = 3;  // register budgets (KernelSet::waves gives each one's waves per SIMD)
constexpr int kDefaultVariant = 1;
// A sample spanning more wavefronts than this reads its wave-speed slots
// before raising them (StageArgs::ws_filter).
constexpr int kWsFilterWaves = 128;
// Small 1D solves run whole timesteps per launch (solve_kernel_1d) when
// each sample's strips take at most this many turns of a workgroup's 16
// wavefronts and the ensemble has at most kSolve1DSamples samples (one
// workgroup each; larger ensembles fill the GPU with the multi-launch step).
constexpr int kSolve1DRounds = 4;
constexpr int kSolve1DSamples = 256;
// Stage kind of a stage mode, for the per-kind register budget.
inline int stage_kind(int mode) { return mode == MODE_RK3_1 ? 1 : (mode == MODE_RK3_2 ? 2 : 0); }

template <class T>
struct KernelSet {
    void (*stage[kStageVariants])(const StageArgs<T>& a, int blocks, hipStream_t s) = {};
    int waves[kStageVariants] = {1, 2, 3};  // waves per SIMD of each variant's register budget
    void (*axpy)(dim3 grid, hipStream_t s, const SampleDesc* desc, SampleState* state,
                 const unsigned long long* ws, const T* X, const T* Y, T* out, int check,
                 double gamma, const StepEnd& end) = nullptr;
    void (*rk4)(dim3 grid, hipStream_t s, const SampleDesc* desc, SampleState* state,
                const unsigned long long* ws, int kind, const T* X, const T* Y, const T* Z,
                const T* W, const T* V, T* out, int check, double gamma,
                const StepEnd& end) = nullptr;
    // volume operators: NumericalFlux::computeFlux on a reference Volume
    // (the first-stage kernel on per-variable memories) and
    // Reconstruction::performReconstruction
    void (*op_flux)(const StageArgs<T>& a, int blocks, hipStream_t s) = nullptr;
    void (*op_recon)(const ReconArgs<T>& a, int blocks, hipStream_t s) = nullptr;
    int recon_radius = 1;
    // whole timesteps of small 1D samples, one workgroup per sample (1D
    // schemes on the marching kernel; null otherwise)
    void (*solve1d)(const Solve1DArgs<T>& p, int nsamples, int threads, hipStream_t s) = nullptr;
};

// Each stage kind (first stage, RK2 second stage, SSP-RK3 second and third
// stages) is its own instantiation: the combination, the lazy stage input of
// the second SSP-RK3 stage and the wave-speed reduction are resolved at
// compile time instead of by branches in the marching loop.
template <class T, class Eq, class Recon, class Flux, int DIM, int WPS, int MODE>
void launch_stage_mode(const StageArgs<T>& a, int blocks, hipStream_t s) {
    if constexpr (DIM == 3)  // one workgroup of kGroupRows wavefronts per task
        hipLaunchKernelGGL((stage_kernel_3<T, Eq, Recon, Flux, WPS, MODE>), dim3(a.ntasks),
                           dim3(64 * kGroupRows), 0, s, a);
    else
        hipLaunchKernelGGL((stage_kernel_12<T, Eq, Recon, Flux, DIM, WPS, MODE>), dim3(blocks), dim3(256),
                           0, s, a);
}

template <class T, class Eq, class Recon, class Flux, int DIM, int WPS>
void launch_stage(const StageArgs<T>& a, int blocks, hipStream_t s) {
    if (blocks <= 0) return;
    switch (a.mode) {
        case MODE_RHS: launch_stage_mode<T, Eq, Recon, Flux, DIM, WPS, MODE_RHS>(a, blocks, s); break;
        case MODE_RK2_1: launch_stage_mode<T, Eq, Recon, Flux, DIM, WPS, MODE_RK2_1>(a, blocks, s); break;
        case MODE_RK3_1: launch_stage_mode<T, Eq, Recon, Flux, DIM, WPS, MODE_RK3_1>(a, blocks, s); break;
        case MODE_RK3_2: launch_stage_mode<T, Eq, Recon, Flux, DIM, WPS, MODE_RK3_2>(a, blocks, s); break;
    }
}

// The schemes of the measured configurations get all three register
// budgets for the autotuner; the others one (2 waves/SIMD), which bounds the
// number of instantiations.
template <class Recon, class Flux>
constexpr bool kAllBudgets =
    (std::is_same_v<Recon, ReconNone> || std::is_same_v<Recon, ReconWENO2> ||
     std::is_same_v<Recon, ReconENO2>) &&
    (std::is_same_v<Flux, FluxHLL> || std::is_same_v<Flux, FluxHLL3> ||
     std::is_same_v<Flux, FluxGodunov>);

// Reconstructions wider than three cells run on the per-cell wide kernel.
template <class T, class Eq, class Recon, class Flux, int DIM, int WPS>
void launch_stage_wide(const StageArgs<T>& a, int blocks, hipStream_t s) {
    if (blocks <= 0) return;
    switch (a.mode) {
        case MODE_RHS:
            hipLaunchKernelGGL((stage_kernel_wide<T, Eq, Recon, Flux, DIM, WPS, MODE_RHS>),
                               dim3(blocks), dim3(256), 0, s, a);
            break;
        case MODE_RK2_1:
            hipLaunchKernelGGL((stage_kernel_wide<T, Eq, Recon, Flux, DIM, WPS, MODE_RK2_1>),
                               dim3(blocks), dim3(256), 0, s, a);
            break;
        case MODE_RK3_1:
            hipLaunchKernelGGL((stage_kernel_wide<T, Eq, Recon, Flux, DIM, WPS, MODE_RK3_1>),
                               dim3(blocks), dim3(256), 0, s, a);
            break;
        case MODE_RK3_2:
            hipLaunchKernelGGL((stage_kernel_wide<T, Eq, Recon, Flux, DIM, WPS, MODE_RK3_2>),
                               dim3(blocks), dim3(256), 0, s, a);
            break;
    }
}

// computeFlux on a reference Volume: the first-stage kernel with per-variable
// memories, writing only the cells of the region (2 waves/SIMD budget).
template <class T, class Eq, class Recon, class Flux, int DIM>
void launch_op_flux(const StageArgs<T>& a, int blocks, hipStream_t s) {
    if (blocks <= 0) return;
    if constexpr (recon_radius<Recon>::value > 1)
        hipLaunchKernelGGL((stage_kernel_wide<T, Eq, Recon, Flux, DIM, 2, MODE_RHS, 1>), dim3(blocks),
                           dim3(256), 0, s, a);
    else if constexpr (DIM == 3)
        hipLaunchKernelGGL((stage_kernel_3<T, Eq, Recon, Flux, 2, MODE_RHS, 1>), dim3(a.ntasks),
                           dim3(64 * kGroupRows), 0, s, a);
    else
        hipLaunchKernelGGL((stage_kernel_12<T, Eq, Recon, Flux, DIM, 2, MODE_RHS, 1>), dim3(blocks),
                           dim3(256), 0, s, a);
}

template <class T, class Eq, class Recon>
void launch_op_recon(const ReconArgs<T>& a, int blocks, hipStream_t s) {
    if (blocks <= 0) return;
    hipLaunchKernelGGL((recon_volume_kernel<T, Eq, Recon>), dim3(blocks), dim3(256), 0, s, a);
}

template <class T, class Eq, class Recon, class Flux>
void launch_solve1d(const Solve1DArgs<T>& p, int nsamples, int threads, hipStream_t s) {
    hipLaunchKernelGGL((solve_kernel_1d<T, Eq, Recon, Flux>), dim3(nsamples), dim3(threads), 0, s, p);
}

template <class T, class Eq, class Recon, class Flux, int DIM>
void set_stage(KernelSet<T>& k) {
    if constexpr (DIM == 1 && recon_radius<Recon>::value <= 1)
        k.solve1d = &launch_solve1d<T, Eq, Recon, Flux>;
    k.op_flux = &launch_op_flux<T, Eq, Recon, Flux, DIM>;
    k.op_recon = &launch_op_recon<T, Eq, Recon>;
    k.recon_radius = recon_radius<Recon>::value;
    if constexpr (recon_radius<Recon>::value > 1) {
        // wide kernels wait on their cached stencil loads: the autotuner
        // also tries the tighter register budgets for HLL3 (the Euler
        // configurations); the others get 2 waves/SIMD
        if constexpr (std::is_same_v<Flux, FluxHLL3>) {
            k.stage[0] = &launch_stage_wide<T, Eq, Recon, Flux, DIM, 2>;
            k.stage[1] = &launch_stage_wide<T, Eq, Recon, Flux, DIM, 3>;
            k.stage[2] = &launch_stage_wide<T, Eq, Recon, Flux, DIM, 4>;
            k.waves[0] = 2, k.waves[1] = 3, k.waves[2] = 4;
        } else {
            k.stage[0] = k.stage[1] = k.stage[2] = &launch_stage_wide<T, Eq, Recon, Flux, DIM, 2>;
            k.waves[0] = k.waves[1] = k.waves[2] = 2;
        }
    } else if constexpr (kAllBudgets<Recon, Flux>) {
        k.stage[0] = &launch_stage<T, Eq, Recon, Flux, DIM, 1>;
        k.stage[1] = &launch_stage<T, Eq, Recon, Flux, DIM, 2>;
        k.stage[2] = &launch_stage<T, Eq, Recon, Flux, DIM, 3>;
        k.waves[0] = 1, k.waves[1] = 2, k.waves[2] = 3;
    } else {
        k.stage[0] = k.stage[1] = k.stage[2] = &launch_stage<T, Eq, Recon, Flux, DIM, 2>;
        k.waves[0] = k.waves[1] = k.waves[2] = 2;
    }
}

template <class T, class Eq>
void launch_axpy(dim3 grid, hipStream_t s, const SampleDesc* desc, SampleState* state,
                 const unsigned long long* ws, const T* X, const T* Y, T* out, int check,
                 double gamma, const StepEnd& end) {
    hipLaunchKernelGGL((axpy_kernel<T, Eq>), grid, dim3(256), 0, s, desc, state, ws, X, Y, out,
                       check, gamma, end);
}

template <class T, class Eq>
void launch_rk4(dim3 grid, hipStream_t s, const SampleDesc* desc, SampleState* state,
                const unsigned long long* ws, int kind, const T* X, const T* Y, const T* Z,
                const T* W, const T* V, T* out, int check, double gamma, const StepEnd& end) {
    hipLaunchKernelGGL((rk4_kernel<T, Eq>), grid, dim3(256), 0, s, desc, state, ws, kind, X, Y,
                       Z, W, V, out, check, gamma, end);
}

template <class T, class Eq, class Flux, int DIM>
int select_recon(const alsvinn_hip_scheme& sc, KernelSet<T>& k) {
    k.axpy = &launch_axpy<T, Eq>;
    k.rk4 = &launch_rk4<T, Eq>;
    switch (sc.reconstruction) {
        case ALSVINN_HIP_RECON_NONE: set_stage<T, Eq, ReconNone, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_WENO2: set_stage<T, Eq, ReconWENO2, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_ENO2: set_stage<T, Eq, ReconENO2, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_MC: set_stage<T, Eq, ReconMC, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_MM: set_stage<T, Eq, ReconMM, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_ENO3: set_stage<T, Eq, ReconENO<3>, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_ENO4: set_stage<T, Eq, ReconENO<4>, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_WENO3: set_stage<T, Eq, ReconWENO3, Flux, DIM>(k); return 0;
        case ALSVINN_HIP_RECON_WENOF2:
            if constexpr (Eq::is_euler) {
                set_stage<T, Eq, ReconWENOF2, Flux, DIM>(k);
                return 0;
            }
            return fail(ALSVINN_HIP_E_INVALID, "wenof2 needs the Euler equations");
    }
    return fail(ALSVINN_HIP_E_INVALID, "unknown reconstruction");
}

// Defined per equation family in kernels_*.hip.
template <class T>
int select_burgers(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k);
template <class T>
int select_euler1(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k);
template <class T>
int select_euler2(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k);
template <class T>
int select_euler3(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k);

template <class T>
int select_kernels(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k) {
    switch (sc.equation) {
        case ALSVINN_HIP_EQ_BURGERS: return select_burgers<T>(sc, dim, k);
        case ALSVINN_HIP_EQ_EULER1: return select_euler1<T>(sc, dim, k);
        case ALSVINN_HIP_EQ_EULER2: return select_euler2<T>(sc, dim, k);
        case ALSVINN_HIP_EQ_EULER3: return select_euler3<T>(sc, dim, k);
    }
    return fail(ALSVINN_HIP_E_INVALID, "equation not available on the hip platform: " +
                                            std::to_string(sc.equation));
}

}  // namespace alsvinn_hip

// The ensemble object behind the C ABI's opaque handle.
struct alsvinn_hip_ensemble {
    alsvinn_hip_scheme scheme;
    int nsamples = 0;
    int nvar = 0;
    int dim = 0;
    int ghosts = 0;
    size_t esize = 8;
    std::vector<alsvinn_hip::SampleDesc> desc;
    std::vector<alsvinn_hip_sample> spec;
    std::vector<char> assigned;  // pool slots: 0 = free
    long long total = 0;  // elements per buffer
    void* U = nullptr;  // conserved state (the caller's buffer when own_U is false)
    bool own_U = true;
    void* A = nullptr;
    void* B = nullptr;  // SSP-RK3 U2; RK4 U1
    void* C = nullptr;  // RK4 U2
    void* D = nullptr;  // RK4 U3
    alsvinn_hip::SampleDesc* d_desc = nullptr;
    alsvinn_hip::SampleState* d_state = nullptr;
    alsvinn_hip::SampleState* d_state_dry = nullptr;  // the autotuner's dry runs
    unsigned long long* d_ws = nullptr;
    alsvinn_hip::WaveTask* d_tasks = nullptr;
    int ntasks = 0;
    int* d_count = nullptr;
    unsigned int* d_end_counter = nullptr;  // workgroups done in the step's last kernel (0 between launches)
    int ws_filter = 0;                      // StageArgs::ws_filter for this work list
    // ensembles with large samples: the first stage's partial maxima
    // (StageArgs::ws_part) and, per sample that has tasks, its entries
    // (ws_idx[ws_off[i] .. ws_off[i+1]) for sample ws_samples[i])
    unsigned long long* d_ws_part = nullptr;
    int* d_ws_off = nullptr;
    int* d_ws_idx = nullptr;
    int* d_ws_samples = nullptr;
    int ws_nlist = 0;
    int* h_count = nullptr;
    double* d_dt_hist = nullptr;
    long long dt_cap = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    long long max_vstride = 0;
    // stage kernel register budget per stage kind: [0] first stage (and the
    // RK2 / RK4 evaluations), [1] second and [2] third SSP-RK3 stage
    int variant[3] = {alsvinn_hip::kDefaultVariant, alsvinn_hip::kDefaultVariant,
                      alsvinn_hip::kDefaultVariant};
    int band_rows = 0;  // rows per wave task; 0 = heuristic
    bool tuned = false;
    // small 1D solves (solve_kernel_1d): eligible when every sample's strips
    // take at most solve1d_rounds turns of a 16-wavefront workgroup (0: never);
    // solve1d_threads is the workgroup size in use (0: the multi-launch step)
    int solve1d_rounds = alsvinn_hip::kSolve1DRounds;
    int solve1d_threads = 0;
    // pools: the work list lives in a preallocated device buffer (an int
    // count, then up to task_cap tasks) that asynchronous copies from pinned
    // staging buffers update, and the stage kernels read the count at run
    // time, so admitting or releasing a solve neither allocates, nor
    // synchronises, nor invalidates the captured step graph
    static constexpr int kRing = 3;
    bool pool = false;
    int max_n[3] = {1, 1, 1};        // the largest grid a slot may hold
    long long task_cap = 0;
    int pool_blocks = 0;             // fixed grid of the pool's stage launches
    unsigned char* d_taskbuf = nullptr;
    unsigned char* h_stage[kRing] = {};
    hipEvent_t stage_ev[kRing] = {};
    int stage_next = 0;
    // state snapshots for non-blocking polls (pinned, one per ticket slot)
    alsvinn_hip::SampleState* h_snap[kRing] = {};
    hipEvent_t snap_ev[kRing] = {};
    long long snap_ticket[kRing] = {-1, -1, -1};
    long long next_ticket = 0;
    alsvinn_hip::KernelSet<double> k64;
    alsvinn_hip::KernelSet<float> k32;
};


/*
 * alsvinn_hip.h — C ABI of the MI355X-native "hip" platform for the alsfvm
 * finite-volume timestep and the alsuq sample statistics.
 *
 * Conventions (the reference's DLLWriter precedent,
 * alsfvm/include/alsfvm/io/DLLWriter.hpp:16-55): plain C types only; every
 * function returns 0 on success and a non-zero status otherwise, with the
 * message available from alsvinn_hip_last_error() (thread-local). The C++
 * side of a caller turns a non-zero status into THROW, as the reference's
 * CUDA_SAFE_CALL does. Calls are asynchronous on the ensemble's HIP stream
 * unless stated otherwise; host copies synchronise.
 *
 * Volumes are structure-of-arrays with ghost layers on the active axes,
 * index var*(tx*ty*tz) + z*tx*ty + y*tx + x over the padded extents
 * (alsfvm/src/volume/Volume.cpp:44-60, alsfvm/include/alsfvm/memory/View.hpp:116-118).
 */
#ifndef ALSVINN_HIP_H
#define ALSVINN_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ALSVINN_HIP_EQ_BURGERS = 0,
    ALSVINN_HIP_EQ_EULER1 = 1,
    ALSVINN_HIP_EQ_EULER2 = 2,
    ALSVINN_HIP_EQ_EULER3 = 3
};
/* Fluxes: hll (numflux/euler/HLL.hpp), hll3 (HLL3.hpp), godunov
 * (numflux/burgers/Godunov.hpp), central (numflux/Central.hpp; Euler and
 * Burgers), rusanov (numflux/Rusanov.hpp; Burgers), roe (numflux/burgers/Roe.hpp). */
enum {
    ALSVINN_HIP_FLUX_HLL = 0,
    ALSVINN_HIP_FLUX_HLL3 = 1,
    ALSVINN_HIP_FLUX_GODUNOV = 2,
    ALSVINN_HIP_FLUX_CENTRAL = 3,
    ALSVINN_HIP_FLUX_RUSANOV = 4,
    ALSVINN_HIP_FLUX_ROE = 5
};
/* Reconstructions: none (NoReconstruction.cpp), weno2 (WENO2.hpp), eno2,
 * eno3, eno4 (ENOCPU.cpp, orders 2-4; ghost cells = order), mc (MC.hpp), mm
 * (MM.hpp), wenof2 (WENOF2.hpp, Euler only), weno3 (WENOCPU.cpp, order 3,
 * 3 ghost cells); ReconstructionFactory.cpp:59-160. */
enum {
    ALSVINN_HIP_RECON_NONE = 0,
    ALSVINN_HIP_RECON_WENO2 = 1,
    ALSVINN_HIP_RECON_ENO2 = 2,
    ALSVINN_HIP_RECON_MC = 3,
    ALSVINN_HIP_RECON_MM = 4,
    ALSVINN_HIP_RECON_WENOF2 = 5,
    ALSVINN_HIP_RECON_ENO3 = 6,
    ALSVINN_HIP_RECON_ENO4 = 7,
    ALSVINN_HIP_RECON_WENO3 = 8
};
/* Integrators: forward Euler (ForwardEuler.cpp), SSP-RK2 (RungeKutta2.cpp),
 * SSP-RK3 (RungeKutta3.cpp), classical RK4 (RungeKutta4.cpp:50-96);
 * IntegratorFactory.cpp:31-44. */
enum {
    ALSVINN_HIP_INTEGRATOR_FORWARD_EULER = 0,
    ALSVINN_HIP_INTEGRATOR_RK2 = 1,
    ALSVINN_HIP_INTEGRATOR_RK3 = 2,
    ALSVINN_HIP_INTEGRATOR_RK4 = 3
};
enum { ALSVINN_HIP_BC_NEUMANN = 0, ALSVINN_HIP_BC_PERIODIC = 1 };
enum { ALSVINN_HIP_FP64 = 0, ALSVINN_HIP_FP32 = 1 };

/* Status codes. */
enum {
    ALSVINN_HIP_OK = 0,
    ALSVINN_HIP_E_INVALID = 1,   /* bad argument / unsupported combination */
    ALSVINN_HIP_E_HIP = 2,       /* HIP runtime error */
    ALSVINN_HIP_E_CONSTRAINT = 3 /* a sample violated the state constraints */
};

/* The numerical scheme: equation, flux, reconstruction, integrator, boundary,
 * storage precision. Mirrors the <fvm> options the reference's factories are
 * keyed on (NumericalFluxFactory.cpp:43-200, ReconstructionFactory.cpp:41-355,
 * IntegratorFactory.cpp:31-44, BoundaryFactory.cpp:38-75,
 * SimulatorSetup.cpp:306-337 for cfl). */
typedef struct alsvinn_hip_scheme {
    int equation;
    int flux;
    int reconstruction;
    int integrator;
    int boundary;
    int precision;
    double gamma;
    double cfl;
} alsvinn_hip_scheme;

/* One solve of an ensemble: its grid (Grid.cpp:24-41), end time and the
 * FixedInterval time-step adjusters attached to it (FixedIntervalWriter.cpp:45-53,
 * FixedIntervalStatistics.cpp:33-41: next save = saved*interval). */
typedef struct alsvinn_hip_sample {
    int nx, ny, nz;
    int num_adjusters;   /* 0..2 */
    double origin[3];
    double top[3];
    double end_time;
    double interval[2];
    long long saved0[2]; /* initial save counters (1 when the initial state is not written) */
    long long max_steps; /* -1: unbounded */
} alsvinn_hip_sample;

typedef struct alsvinn_hip_ensemble alsvinn_hip_ensemble;

const char* alsvinn_hip_last_error(void);

/* Number of ghost layers the scheme needs: max(flux stencil - 1, reconstruction
 * ghosts) (NumericalFluxCPU.cpp:226-231). */
int alsvinn_hip_ghost_cells(const alsvinn_hip_scheme* scheme);

/* Ensemble: many solves of one scheme advanced together. Replaces the
 * reference's sequential sample loop (alsuq/src/run/Runner.cpp:42-78) and the
 * Simulator objects it creates per sample (Simulator.cpp:26-64). stream may be
 * NULL (the library creates one). The ensemble owns its state buffers. */
int alsvinn_hip_ensemble_create(const alsvinn_hip_scheme* scheme, int nsamples,
                                const alsvinn_hip_sample* samples, void* stream,
                                alsvinn_hip_ensemble** out);
int alsvinn_hip_ensemble_destroy(alsvinn_hip_ensemble* e);

/* Caller-owned state (the reference's Simulator owns its conserved volumes,
 * Simulator.cpp:59-63, and hands device pointers to its platform kernels):
 * alsvinn_hip_state_elements gives the number of reals of the concatenated
 * sample volumes (each in the reference's SoA layout with ghost layers, in
 * sample order; offsets from alsvinn_hip_ensemble_layout), and
 * alsvinn_hip_ensemble_create_bound advances the samples in that device buffer
 * in place. The buffer must outlive the ensemble; the library never frees it
 * and does not clear it. Scratch volumes stay library-owned. */
int alsvinn_hip_state_elements(const alsvinn_hip_scheme* scheme, int nsamples,
                               const alsvinn_hip_sample* samples, long long* elements);
int alsvinn_hip_ensemble_create_bound(const alsvinn_hip_scheme* scheme, int nsamples,
                                      const alsvinn_hip_sample* samples, void* stream, void* state,
                                      alsvinn_hip_ensemble** out);

/* Pools for continuous batching (no counterpart in the reference, whose
 * Runner creates one Simulator per sample): nslots sample slots over an
 * arena of arena_elements state elements (per buffer); max_n is the largest
 * grid a slot may hold (its dimension is the pool's). A slot is assigned a
 * solve at an arena offset (the caller places solves, e.g. first-fit by
 * footprint nvar*(nx+2g)*(ny+2g)*(nz+2g)) and released when its result was
 * read; other slots keep stepping. Unassigned slots cost nothing.
 *
 * Assignment and release are asynchronous, in the pool's stream order: they
 * allocate nothing (the work list lives in a device buffer sized at creation
 * for nslots grids of max_n cells, updated by copies from pinned staging
 * buffers), do not synchronise the stream, and keep the captured step graph
 * valid, so a request can be admitted while the pool is stepping. */
int alsvinn_hip_pool_create(const alsvinn_hip_scheme* scheme, int nslots, const int max_n[3],
                            long long arena_elements, void* stream, alsvinn_hip_ensemble** out);
int alsvinn_hip_pool_assign(alsvinn_hip_ensemble* e, int slot, const alsvinn_hip_sample* sample,
                            long long offset);
int alsvinn_hip_pool_release(alsvinn_hip_ensemble* e, int slot);
/* Initial state of a just-assigned slot (full padded volume, FP64 on the
 * host): a synchronous copy that does not wait for the pool's stream (no
 * step enqueued before the assignment touches the slot's memory). */
int alsvinn_hip_pool_upload(alsvinn_hip_ensemble* e, int slot, const double* host);
/* Boundary::applyBoundaryConditions on one slot, in stream order. */
int alsvinn_hip_pool_apply_boundary(alsvinn_hip_ensemble* e, int slot);
/* Result of a finished slot (a snapshot showed it inactive; no step writes
 * an inactive slot), without waiting for the pool's stream. Synchronous. */
int alsvinn_hip_pool_download(alsvinn_hip_ensemble* e, int slot, double* host);
/* Non-blocking progress checks: pool_snapshot enqueues a copy of every
 * slot's t, steps, active flag and status into pinned host memory and
 * returns its ticket; pool_poll reports it (*ready = 1) once the copy has
 * run, or returns at once with *ready = 0 (wait = 0), or waits (wait = 1).
 * The last 3 tickets are kept. */
int alsvinn_hip_pool_snapshot(alsvinn_hip_ensemble* e, long long* ticket);
int alsvinn_hip_pool_poll(alsvinn_hip_ensemble* e, long long ticket, int wait, double* t,
                          long long* steps, int* active, int* status, int* ready);

/* Free and total device memory (hipMemGetInfo), for sizing pools against
 * the device's memory budget. */
int alsvinn_hip_device_memory(long long* free_bytes, long long* total_bytes);

/* Per-sample t, steps, active flag and status of every sample/slot in one
 * copy. Synchronous. */
int alsvinn_hip_ensemble_states(alsvinn_hip_ensemble* e, double* t, long long* steps, int* active,
                                int* status);

/* Padded extents (tx,ty,tz), ghost layers (gx,gy,gz) and the element offset of
 * a sample's volume in the state buffer. */
int alsvinn_hip_ensemble_layout(const alsvinn_hip_ensemble* e, int sample, int dims6[6],
                                long long* offset, long long* elements_per_variable);

/* Host copies of a sample's full conserved volume (nvar x tz x ty x tx, FP64 on
 * the host whatever the storage precision). Synchronous. */
int alsvinn_hip_ensemble_upload(alsvinn_hip_ensemble* e, int sample, const double* host);
int alsvinn_hip_ensemble_download(alsvinn_hip_ensemble* e, int sample, double* host);

/* Device pointer of the current conserved state buffer (all samples). */
int alsvinn_hip_ensemble_state_pointer(alsvinn_hip_ensemble* e, void** ptr);

/* Boundary::applyBoundaryConditions on every sample's current state
 * (Simulator::setInitialValue, Simulator.cpp:160). */
int alsvinn_hip_ensemble_apply_boundary(alsvinn_hip_ensemble* e);

/* Resets every sample to t = 0, steps = 0, initial save counters, active. */
int alsvinn_hip_ensemble_reset_time(alsvinn_hip_ensemble* e);

/* Work schedule of the stage kernels: rows per wavefront task (0 = built-in
 * heuristic) and the register-budget variant (0, 1, 2; the marching kernels'
 * budgets allow 1, 2, 3 waves per SIMD, the wide-stencil kernel's with HLL3
 * 2, 3, 4, as alsvinn_hip_ensemble_get_stage_waves reports). Results do not
 * depend on the schedule, only the speed does. */
int alsvinn_hip_ensemble_set_schedule(alsvinn_hip_ensemble* e, int rows, int variant);
int alsvinn_hip_ensemble_get_schedule(const alsvinn_hip_ensemble* e, int* rows, int* variant,
                                      int* ntasks);

/* Autotuner: times a dry run of the step's stage kernels (each of the three
 * SSP-RK3 stages, the first stage otherwise; outputs into scratch volumes) on
 * the current state for every variant and a ladder of band heights, twice in
 * opposite orders; keeps the band height with the smallest total and, per
 * stage kind, the fastest variant (*variant: the first stage's).
 * Synchronous; leaves the conserved state untouched. set_schedule sets one
 * variant for every stage kind. */
int alsvinn_hip_ensemble_autotune(alsvinn_hip_ensemble* e, int* rows, int* variant);
/* The register-budget variant of each stage kind: [0] first stage (and the
 * RK2/RK4 evaluations), [1] second and [2] third SSP-RK3 stage. */
int alsvinn_hip_ensemble_get_stage_variants(const alsvinn_hip_ensemble* e, int* variants);
/* Waves per SIMD that the chosen register budget of each stage kind allows,
 * in the same order (no counterpart in the reference: a report of the
 * autotuned schedule). */
int alsvinn_hip_ensemble_get_stage_waves(const alsvinn_hip_ensemble* e, int* waves);

/* Small 1D solves (no counterpart in the reference, whose Simulator runs
 * one sample's step as a sequence of host calls): when the ensemble is 1D,
 * not a pool, integrated by forward Euler, RK2 or SSP-RK3, has at most 256
 * samples and every sample's strips of 64 - 2*ghosts cells take at most
 * max_rounds turns of 16 wavefronts, one workgroup per sample runs whole
 * timesteps in one launch (workgroup barriers between the stages) instead of
 * two or three launches per step. The results are the same bits either way.
 * max_rounds = 0 keeps the multi-launch step; the default is 4 (up to 3840
 * cells with two ghost layers). get_small_1d reports the workgroup size in
 * use (0: the multi-launch step). */
int alsvinn_hip_ensemble_set_small_1d(alsvinn_hip_ensemble* e, int max_rounds);
int alsvinn_hip_ensemble_get_small_1d(const alsvinn_hip_ensemble* e, int* threads);

/* Enqueues nsteps timesteps (Simulator::performStep, Simulator.cpp:81-85) for
 * every sample that has not reached its end time. Asynchronous. */
int alsvinn_hip_ensemble_step(alsvinn_hip_ensemble* e, int nsteps);

/* Steps until every sample reached its end time (or failed), replaying a
 * captured HIP graph of steps_per_sync steps between host checks. Synchronous.
 * Returns ALSVINN_HIP_E_CONSTRAINT if a sample violated the constraints. */
int alsvinn_hip_ensemble_run(alsvinn_hip_ensemble* e, int steps_per_sync);

/* Pause at saves (FixedIntervalWriter::write and FixedIntervalStatistics::write,
 * FixedIntervalWriter.cpp:31-43, FixedIntervalStatistics.cpp:44-62, run after
 * the step that reaches a save time): every sample holds (stays unchanged and
 * is not counted as running) once its save counter of adjuster k (0 or 1)
 * reaches saved_limit, so a host writer can take the snapshot that the
 * reference writes after that step. saved_limit < 0 removes the hold. The
 * dt sequence is unaffected. */
int alsvinn_hip_ensemble_set_hold(alsvinn_hip_ensemble* e, int adjuster, long long saved_limit);
/* Save counters of a sample's two adjusters (initial value saved0 + 1 per
 * save time reached). Synchronises the ensemble's stream. */
int alsvinn_hip_ensemble_saves(alsvinn_hip_ensemble* e, int sample, long long saved[2]);

/* Replays the captured graph of steps_per_graph timesteps nsteps/steps_per_graph
 * times without host checks (samples past their end time skip on the device).
 * Asynchronous. */
int alsvinn_hip_ensemble_run_steps(alsvinn_hip_ensemble* e, int nsteps, int steps_per_graph);

/* Measurement mode: enqueues nsteps timesteps with HIP events around every
 * kernel (synchronising after each) and returns the summed kernel times in
 * milliseconds and launch counts for [stage kernels, elementwise kernels,
 * separate step-end kernels]; the last is 0: the step's last kernel ends the
 * step (time, counters, stop flags). Synchronous. */
int alsvinn_hip_ensemble_step_timed(alsvinn_hip_ensemble* e, int nsteps, double kernel_ms[3],
                                    long long launches[3]);

int alsvinn_hip_ensemble_synchronize(alsvinn_hip_ensemble* e);

/* Per-sample time, steps performed, last dt, status (0 ok, 2 constraint
 * violation) and active flag. Synchronous. */
int alsvinn_hip_ensemble_sample_info(alsvinn_hip_ensemble* e, int sample, double* t,
                                     long long* steps, double* last_dt, int* status,
                                     int* active);

/* Records the dt of the first `capacity` steps of every sample; read back with
 * alsvinn_hip_ensemble_dt_history. */
int alsvinn_hip_ensemble_record_dt(alsvinn_hip_ensemble* e, long long capacity);
int alsvinn_hip_ensemble_dt_history(alsvinn_hip_ensemble* e, int sample, double* host,
                                    long long capacity);

/* One evaluation of the spatial operator on every sample's current state:
 * ConservedSystem::operator() = NumericalFlux::computeFlux over all regions
 * (ConservedSystem.cpp:31-103, NumericalFluxCPU.cpp:170-221). The result
 * (interior cells) and the per-direction max wave speeds of one sample are
 * copied to the host. Synchronous. */
int alsvinn_hip_ensemble_rhs(alsvinn_hip_ensemble* e, int sample, double* host_out,
                             double wave_speeds[3]);

/* Time of the kernels enqueued by the last alsvinn_hip_ensemble_step/run
 * call, from HIP events on the ensemble's stream (milliseconds). */
int alsvinn_hip_ensemble_last_elapsed_ms(alsvinn_hip_ensemble* e, double* ms);
/* Start and end of e's last step/run call on the device, in milliseconds
 * after the start of ref's last call (HIP events on the two streams; no
 * counterpart in the reference): shows whether two ensembles or pools ran at
 * the same time. Synchronous. */
int alsvinn_hip_ensemble_interval_ms(alsvinn_hip_ensemble* e, alsvinn_hip_ensemble* ref,
                                     double* start_ms, double* end_ms);

/* ---- alsuq statistics on the device (accumulated in FP64) ----
 * Each call takes a group of sample indices whose grids are identical and
 * reads their current states (the states at the save time when called after
 * alsvinn_hip_ensemble_run with one save at the end time). Synchronous. */

/* meanvar (alsuq/src/stats/MeanVariance.cpp:30-67, StatisticsHelper.cpp:35-79):
 * full padded volumes (nvar x tz x ty x tx). welford = 0 reproduces the
 * reference (sum u and pow(u,2) in sample order; var = S2/M - mean^2);
 * welford = 1 accumulates (mean, M2) in sample order and returns M2/M. */
int alsvinn_hip_stats_meanvar(alsvinn_hip_ensemble* e, const int* samples, int m, int welford,
                              double* host_mean, double* host_var);

/* OnePointMoment (alsuq/src/stats/OnePointMoment.cpp:30-45 with
 * HostMemory::addPower, HostMemory.cpp:283-290, and the combine's division by
 * M, StatisticsHelper.cpp:35-79): per cell (ghosts included), the sum over the
 * listed samples, in order, of std::pow(u, p), divided by M. Output in the
 * volume layout of the samples. Bit-identical to the reference for p = 1, 2. */
int alsvinn_hip_stats_onepoint_moment(alsvinn_hip_ensemble* e, const int* samples, int m,
                                      double p, double* host_out);
/* BoundedVariation (BoundedVariation.cpp:38-53 with
 * HostMemory::getTotalVariation, HostMemory.cpp:325-362): per variable, the
 * sum over the listed samples of sum_cells pow(sqrt((u - u_below)^2 +
 * (u - u_left)^2), p) over the reference's index ranges, divided by M; 1D and
 * 2D (the reference rejects 3D). out: nvar reals. */
int alsvinn_hip_stats_bounded_variation(alsvinn_hip_ensemble* e, const int* samples, int m,
                                        double p, double* host_out);

/* structure_cube (alsuq/src/stats/StructureCube.cpp:37-76,
 * alsfvm/include/alsfvm/functional/structure_common.hpp:37-150): out[var*H + h]
 * for h = 1..H-1 (out[var*H] = 0), combined over the group (divided by m). */
int alsvinn_hip_stats_structure_cube(alsvinn_hip_ensemble* e, const int* samples, int m,
                                     int number_of_h, double p, double* host_out);

/* structure_basic (alsuq/src/stats/StructureBasic.cpp:67-120): one direction,
 * h = 0..H-1, out[var*H + h], divided by m. */
int alsvinn_hip_stats_structure_basic(alsvinn_hip_ensemble* e, const int* samples, int m,
                                      int number_of_h, double p, int direction,
                                      double* host_out);

/* FP64 sum of every conserved variable over each sample's interior cells:
 * host[sample*nvar + var]. */
int alsvinn_hip_ensemble_totals(alsvinn_hip_ensemble* e, double* host);

/* ---- Operators on a reference Volume (device memory, one area per variable)
 *
 * The reference's platform operators take a volume::Volume whose conserved
 * variables are separate memory areas of the padded extents
 * (alsfvm/src/volume/Volume.cpp:44-60, View.hpp:116-118). These entry points
 * take such a volume as device pointers, one per variable, and never copy it
 * to the host. Calls are asynchronous on the operator's stream except where
 * a result goes to the host. */
typedef struct alsvinn_hip_volume {
    int nvar;        /* conserved variables: 1 (Burgers), 3/4/5 (Euler 1D/2D/3D) */
    int precision;   /* ALSVINN_HIP_FP64 or ALSVINN_HIP_FP32: element type of var[] */
    int nx, ny, nz;  /* interior cells per axis (1 on inactive axes) */
    int gx, gy, gz;  /* ghost layers per axis (0 on inactive axes) */
    void* var[5];    /* device memory of variable v: (nx+2gx)(ny+2gy)(nz+2gz) elements, x fastest */
} alsvinn_hip_volume;

typedef struct alsvinn_hip_operator alsvinn_hip_operator;

/* A numerical flux with its reconstruction for grids of nx x ny x nz interior
 * cells: the object NumericalFluxFactory::createNumericalFlux(grid) returns
 * (NumericalFluxFactory.cpp:159-200; ReconstructionFactory.cpp:41-355 for the
 * reconstruction it owns). Only the scheme's equation, flux, reconstruction,
 * precision and gamma are used. stream may be NULL (the library creates one). */
int alsvinn_hip_operator_create(const alsvinn_hip_scheme* scheme, int nx, int ny, int nz,
                                void* stream, alsvinn_hip_operator** out);
int alsvinn_hip_operator_destroy(alsvinn_hip_operator* op);

/* NumericalFlux::computeFlux(conservedVariables, waveSpeed, computeWaveSpeed,
 * output, start, end) (NumericalFlux.hpp:44-52, NumericalFluxCPU.cpp:170-221):
 * for every cell of the region [g+start, N+g+end) on each active axis (start
 * >= 0, end <= 0, in cells), output = sum over directions of
 * (-F(i+1/2)) + F(i-1/2), the unscaled divergence of the face fluxes; other
 * cells of output, ghosts included, keep their contents. wave_speeds[d] gets
 * the max wave speed over the faces of direction d of the region (every
 * active direction; the reference computes it whatever compute_wave_speed
 * says); the entry synchronises the operator's stream to return it. The
 * ConservedSystem of the reference calls this once for the inner region and
 * twice per direction for the sides (ConservedSystem.cpp:31-103). start and
 * end may be NULL (whole interior). */
int alsvinn_hip_numflux_compute(alsvinn_hip_operator* op, const alsvinn_hip_volume* conserved,
                                double wave_speeds[3], int compute_wave_speed,
                                const alsvinn_hip_volume* output, const int start[3],
                                const int end[3]);

/* Reconstruction::performReconstruction(input, direction, indicatorVariable,
 * left, right, start, end) (Reconstruction.hpp:46-58) with the reference's
 * index ranges: ReconstructionCPU.cpp:71-77 for weno2, wenof2, mc and mm
 * (ghost 2), ENOCPU.cpp:92-98 for eno2-4 (ghost = order), WENOCPU.cpp:101-107
 * for weno3 (start and end unused, every row of the other axes),
 * NoReconstruction.cpp:42-64 for none (every cell). left[i] is the value at
 * the cell's lower face, right[i] at its upper face. indicator_variable must
 * be 0 (the reference's numerical fluxes pass 0). Asynchronous. */
int alsvinn_hip_reconstruction_compute(alsvinn_hip_operator* op, const alsvinn_hip_volume* input,
                                       int direction, int indicator_variable,
                                       const alsvinn_hip_volume* left,
                                       const alsvinn_hip_volume* right, const int start[3],
                                       const int end[3]);

/* Boundary::applyBoundaryConditions(volume, grid) (Boundary.hpp:30-31,
 * BoundaryCPU.cpp:30-121 with Neumann.hpp:24-35 / Periodic.hpp:24-42): fills
 * every ghost layer of the volume (corners included) from the interior.
 * boundary: ALSVINN_HIP_BC_NEUMANN or ALSVINN_HIP_BC_PERIODIC. stream may be
 * NULL (the default stream). Asynchronous. */
int alsvinn_hip_boundary_apply(int boundary, const alsvinn_hip_volume* volume, void* stream);

/* ---- Statistics fed one sample at a time (alsuq::stats::Statistics,
 * Statistics.hpp:25-62), as the reference's Runner feeds them
 * (Runner.cpp:42-87): computeStatistics for every sample, then
 * combineStatistics (StatisticsHelper.cpp:35-79: / number of samples) and
 * finalizeStatistics. Accumulators are FP64 device memory owned by the
 * object, whatever the sample precision. One object is one snapshot (one
 * save time); REGISTER_STATISTICS(hip, ...) classes keep one per time. */
enum {
    ALSVINN_HIP_STAT_MEANVAR = 0,          /* MeanVariance.cpp:30-67: outputs mean, variance */
    ALSVINN_HIP_STAT_ONEPOINTMOMENT = 1,   /* OnePointMoment.cpp:30-45: output m_p */
    ALSVINN_HIP_STAT_STRUCTURE_CUBE = 2,   /* StructureCube.cpp:37-76 */
    ALSVINN_HIP_STAT_STRUCTURE_BASIC = 3,  /* StructureBasic.cpp:67-120 */
    ALSVINN_HIP_STAT_BV = 4                /* BoundedVariation.cpp:38-53 (1D, 2D) */
};

typedef struct alsvinn_hip_statistics alsvinn_hip_statistics;

/* kind: ALSVINN_HIP_STAT_*; p: the exponent (onepointmoment, structure
 * functions, bv); number_of_h and direction: structure functions (direction
 * for structure_basic only); boundary: ALSVINN_HIP_BC_* of the grid (the
 * structure functions' ValueAtBoundary, ValueAtBoundary.hpp:20-98). shape
 * gives nvar, precision, extents and ghosts of the samples (its pointers are
 * not used). Outputs: meanvar 2 volumes (mean, variance) of nvar x padded
 * cells; onepointmoment 1 such volume; structure functions nvar x number_of_h
 * values (out[var*H + h]); bv nvar values. bv refuses a shape more than
 * one cell deep in z, ghost layers included, as HostMemory::
 * getTotalVariation throws for such memory (HostMemory.cpp:333-335). */
int alsvinn_hip_statistics_create(int kind, double p, int number_of_h, int direction, int boundary,
                                  const alsvinn_hip_volume* shape, void* stream,
                                  alsvinn_hip_statistics** out);
int alsvinn_hip_statistics_destroy(alsvinn_hip_statistics* st);
/* computeStatistics: adds one sample's contribution. Asynchronous. */
int alsvinn_hip_statistics_compute(alsvinn_hip_statistics* st, const alsvinn_hip_volume* sample);
/* combineStatistics on one rank: every output divided by samples (the
 * structure functions' sums first by the number of interior cells). */
int alsvinn_hip_statistics_combine(alsvinn_hip_statistics* st, long long samples);
/* finalizeStatistics: meanvar variance -= mean^2; nothing for the others. */
int alsvinn_hip_statistics_finalize(alsvinn_hip_statistics* st);
/* Zeroes the accumulators (a new snapshot). */
int alsvinn_hip_statistics_reset(alsvinn_hip_statistics* st);
/* Number of outputs and FP64 elements per output. */
int alsvinn_hip_statistics_outputs(const alsvinn_hip_statistics* st, int* count,
                                   long long* elements);
/* Device pointer of output `index` (FP64), valid until the object is destroyed. */
int alsvinn_hip_statistics_result(alsvinn_hip_statistics* st, int index, double** device);
/* Host copy of output `index`. Synchronous. */
int alsvinn_hip_statistics_download(alsvinn_hip_statistics* st, int index, double* host);

/* ---- Allocation-free computeFlux.
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same computation as
 * alsvinn_hip_numflux_compute for a scheme and one region, without an
 * operator object: the volume descriptor, the wave-speed slots and the
 * region's work list live in a library-owned 1 MiB device array per GPU
 * (held until the process ends), written by a setup kernel on `stream`; the
 * call makes no allocation, no host copy and no synchronisation. A region
 * that needs more than about 10900 work items (for example a 1D grid of more
 * than 650k cells, or 2D/3D grids whose x-strips times row bands exceed it)
 * is rejected with ALSVINN_HIP_E_INVALID. device_wave_speeds (may be NULL):
 * device memory of 3 doubles that receives the max wave speed of each active
 * direction. Calls on one GPU share its array: calls that overlap in time on
 * that GPU (different streams, or different host threads) race. */
int alsvinn_hip_numflux_compute_noalloc(const alsvinn_hip_scheme* scheme,
                                        const alsvinn_hip_volume* conserved,
                                        const alsvinn_hip_volume* output, const int start[3],
                                        const int end[3], double* device_wave_speeds,
                                        void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ALSVINN_HIP_H */

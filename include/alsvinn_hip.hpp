// C++ host layer of the hip platform, over the C ABI of alsvinn_hip.h.
//
// The classes mirror the reference interfaces the platform stands behind, with
// the same names, argument meaning and error behaviour (a failing call throws
// std::runtime_error, as the reference's THROW / CUDA_SAFE_CALL do):
//
//   alsfvm::hip::Simulator        alsfvm/include/alsfvm/simulator/Simulator.hpp
//                                 (setInitialValue, performStep, atEnd,
//                                 getCurrentTime, getEndTime; Simulator.cpp:77-214)
//   alsfvm::hip::NumericalFlux    numflux::NumericalFlux::computeFlux on a device
//                                 Volume over a region (NumericalFlux.hpp:44-52)
//   alsfvm::hip::Reconstruction   reconstruction::Reconstruction::performReconstruction
//                                 (Reconstruction.hpp:46-58)
//   alsfvm::hip::Boundary         boundary::Boundary::applyBoundaryConditions (Boundary.hpp:30-31)
//   alsfvm::hip::ConservedSystem  simulator::ConservedSystem::operator() from computeFlux
//                                 regions (ConservedSystem.cpp:31-103)
//   alsuq::hip::Runner            alsuq::run::Runner::run (Runner.cpp:39-90): all samples
//                                 of the rank advance together
//   alsuq::hip::MeanVariance, StructureCube, StructureBasic, OnePointMoment,
//   BoundedVariation              alsuq::stats (MeanVariance.cpp:30-67, StructureCube.cpp:37-76,
//                                 StructureBasic.cpp:67-120, OnePointMoment.cpp:30-45,
//                                 BoundedVariation.cpp:38-53) over a Runner's samples
//   alsuq::hip::Statistics        alsuq::stats::Statistics fed one device Volume at a
//                                 time (Statistics.hpp:25-62)
//
// The operators take device Volumes (alsvinn_hip_volume: one device pointer
// per conserved variable, the reference Volume's layout, Volume.cpp:44-60)
// and never copy them to the host. The Simulator and Runner take initial
// data as host vectors in the same layout (variable-major, z*tx*ty + y*tx +
// x over padded extents), as the reference evaluates initial data on the host.
#pragma once

#include <algorithm>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "alsvinn_hip.h"

namespace alsvinn_hip_cxx {

inline void check(int rc, const char* what) {
    if (rc != ALSVINN_HIP_OK)
        throw std::runtime_error(std::string(what) + ": " + alsvinn_hip_last_error());
}

struct EnsembleHandle {
    alsvinn_hip_ensemble* e = nullptr;
    EnsembleHandle() = default;
    EnsembleHandle(const EnsembleHandle&) = delete;
    EnsembleHandle& operator=(const EnsembleHandle&) = delete;
    ~EnsembleHandle() {
        if (e) alsvinn_hip_ensemble_destroy(e);
    }
};

// Grid of the reference (Grid.cpp:24-41): origin, top, cells per axis.
struct Grid {
    double origin[3] = {0, 0, 0};
    double top[3] = {1, 1, 1};
    int n[3] = {1, 1, 1};
};

inline alsvinn_hip_sample make_sample(const Grid& g, double end_time, int number_of_saves,
                                      long long max_steps = -1) {
    alsvinn_hip_sample p;
    std::memset(&p, 0, sizeof(p));
    p.nx = g.n[0];
    p.ny = g.n[1];
    p.nz = g.n[2];
    for (int k = 0; k < 3; ++k) {
        p.origin[k] = g.origin[k];
        p.top[k] = g.top[k];
    }
    p.end_time = end_time;
    // FixedIntervalWriter / FixedIntervalStatistics without the initial
    // timestep: interval T / numberOfSaves, first save at one interval.
    p.num_adjusters = number_of_saves > 0 ? 1 : 0;
    p.interval[0] = number_of_saves > 0 ? end_time / number_of_saves : 0.0;
    p.saved0[0] = 1;
    p.max_steps = max_steps;
    return p;
}

inline int number_of_variables(const alsvinn_hip_scheme& s) {
    return s.equation == ALSVINN_HIP_EQ_BURGERS ? 1 : s.equation + 2;
}

}  // namespace alsvinn_hip_cxx

namespace alsfvm {
namespace hip {

using alsvinn_hip_cxx::check;
using alsvinn_hip_cxx::Grid;

// One solve on the device: the hip counterpart of simulator::Simulator.
class Simulator {
public:
    Simulator(const alsvinn_hip_scheme& scheme, const Grid& grid, double endTime,
              int numberOfSaves = 1)
        : scheme(scheme), grid(grid), endTime(endTime) {
        auto p = alsvinn_hip_cxx::make_sample(grid, endTime, numberOfSaves);
        check(alsvinn_hip_ensemble_create(&scheme, 1, &p, nullptr, &h.e), "Simulator");
    }

    // Simulator::setInitialValue: conserved volume (padded, reference layout);
    // applies the boundary conditions (Simulator.cpp:160).
    void setInitialValue(const std::vector<double>& conserved) {
        if (conserved.size() != volumeSize())
            throw std::runtime_error("setInitialValue: wrong volume size");
        check(alsvinn_hip_ensemble_upload(h.e, 0, conserved.data()), "setInitialValue");
        check(alsvinn_hip_ensemble_apply_boundary(h.e), "setInitialValue");
    }

    // Simulator::performStep: one timestep; THROWs on a constraint violation.
    void performStep() {
        check(alsvinn_hip_ensemble_step(h.e, 1), "performStep");
        int status = 0;
        check(alsvinn_hip_ensemble_sample_info(h.e, 0, &t, &steps, &lastDt, &status, &active),
              "performStep");
        if (status != 0)
            throw std::runtime_error("Simulation state does not obey constraints! At time " +
                                     std::to_string(t) + ", number of timesteps performed " +
                                     std::to_string(steps));
    }

    // Runs to the end time inside the device loop (HIP graph replays).
    void runToEnd(int stepsPerSync = 32) {
        check(alsvinn_hip_ensemble_run(h.e, stepsPerSync), "runToEnd");
        int status = 0;
        check(alsvinn_hip_ensemble_sample_info(h.e, 0, &t, &steps, &lastDt, &status, &active),
              "runToEnd");
    }

    bool atEnd() const { return t >= endTime; }
    double getCurrentTime() const { return t; }
    double getEndTime() const { return endTime; }
    long long getTimestepsPerformed() const { return steps; }

    std::vector<double> getConservedVolume() {
        std::vector<double> out(volumeSize());
        check(alsvinn_hip_ensemble_download(h.e, 0, out.data()), "getConservedVolume");
        return out;
    }

    size_t volumeSize() const {
        int d[6];
        long long off = 0, n = 0;
        check(alsvinn_hip_ensemble_layout(h.e, 0, d, &off, &n), "layout");
        return (size_t)n * alsvinn_hip_cxx::number_of_variables(scheme);
    }

    alsvinn_hip_ensemble* handle() { return h.e; }

private:
    alsvinn_hip_scheme scheme;
    Grid grid;
    double endTime;
    double t = 0, lastDt = 0;
    long long steps = 0;
    int active = 1;
    alsvinn_hip_cxx::EnsembleHandle h;
};

// numflux::NumericalFlux of the hip platform with its reconstruction, for one
// grid: computeFlux on a device Volume over a region [start, end) of
// interior cells (start >= 0, end <= 0), as ConservedSystem calls it.
class NumericalFlux {
public:
    NumericalFlux(const alsvinn_hip_scheme& scheme, const Grid& grid, void* stream = nullptr)
        : scheme(scheme) {
        check(alsvinn_hip_operator_create(&scheme, grid.n[0], grid.n[1], grid.n[2], stream, &op),
              "NumericalFlux");
    }
    ~NumericalFlux() { alsvinn_hip_operator_destroy(op); }
    NumericalFlux(const NumericalFlux&) = delete;
    NumericalFlux& operator=(const NumericalFlux&) = delete;

    void computeFlux(const alsvinn_hip_volume& conservedVariables, double waveSpeed[3],
                     bool computeWaveSpeed, const alsvinn_hip_volume& output,
                     const int start[3] = nullptr, const int end[3] = nullptr) {
        check(alsvinn_hip_numflux_compute(op, &conservedVariables, waveSpeed, computeWaveSpeed ? 1 : 0,
                                          &output, start, end),
              "computeFlux");
    }

    size_t getNumberOfGhostCells() const { return (size_t)alsvinn_hip_ghost_cells(&scheme); }
    alsvinn_hip_operator* handle() { return op; }

private:
    alsvinn_hip_scheme scheme;
    alsvinn_hip_operator* op = nullptr;
};

// reconstruction::Reconstruction of the hip platform (the scheme's
// reconstruction), with the reference's index ranges per reconstruction.
class Reconstruction {
public:
    Reconstruction(const alsvinn_hip_scheme& scheme, const Grid& grid, void* stream = nullptr)
        : flux(scheme, grid, stream), scheme(scheme) {}

    void performReconstruction(const alsvinn_hip_volume& inputVariables, size_t direction,
                               size_t indicatorVariable, const alsvinn_hip_volume& leftOut,
                               const alsvinn_hip_volume& rightOut, const int start[3] = nullptr,
                               const int end[3] = nullptr) {
        check(alsvinn_hip_reconstruction_compute(flux.handle(), &inputVariables, (int)direction,
                                                 (int)indicatorVariable, &leftOut, &rightOut, start,
                                                 end),
              "performReconstruction");
    }

    size_t getNumberOfGhostCells() const {
        switch (scheme.reconstruction) {
            case ALSVINN_HIP_RECON_NONE: return 1;
            case ALSVINN_HIP_RECON_ENO3:
            case ALSVINN_HIP_RECON_WENO3: return 3;
            case ALSVINN_HIP_RECON_ENO4: return 4;
            default: return 2;
        }
    }

private:
    NumericalFlux flux;
    alsvinn_hip_scheme scheme;
};

// boundary::Boundary of the hip platform: fills the ghost layers of a device
// Volume in place (BoundaryFactory.cpp:38-75 names: "neumann", "periodic").
class Boundary {
public:
    explicit Boundary(int boundary, void* stream = nullptr) : bc(boundary), stream(stream) {}
    void applyBoundaryConditions(const alsvinn_hip_volume& volume) {
        check(alsvinn_hip_boundary_apply(bc, &volume, stream), "applyBoundaryConditions");
    }

private:
    int bc;
    void* stream;
};

// simulator::ConservedSystem::operator() (ConservedSystem.cpp:31-103) on one
// rank: computeFlux on the inner region, then on both sides of every active
// direction; the wave speeds are the max over the regions.
class ConservedSystem {
public:
    explicit ConservedSystem(NumericalFlux& flux) : flux(flux) {}
    void operator()(const alsvinn_hip_volume& conservedVariables, double waveSpeed[3],
                    bool computeWaveSpeed, const alsvinn_hip_volume& output) {
        const int g[3] = {output.gx, output.gy, output.gz};
        const int size[3] = {output.nx + 2 * output.gx, output.ny + 2 * output.gy,
                             output.nz + 2 * output.gz};
        const int dims = 1 + (output.ny > 1) + (output.nz > 1);
        const int negg[3] = {-g[0], -g[1], -g[2]};
        flux.computeFlux(conservedVariables, waveSpeed, computeWaveSpeed, output, g, negg);
        double sides[3] = {0, 0, 0};
        for (int d = 0; d < dims; ++d) {
            int sb[3] = {0, 0, 0}, eb[3] = {0, 0, 0}, st[3] = {0, 0, 0}, et[3] = {0, 0, 0};
            eb[d] = -size[d] + 3 * g[d];
            st[d] = size[d] - 3 * g[d];
            double wb[3] = {0, 0, 0}, wt[3] = {0, 0, 0};
            flux.computeFlux(conservedVariables, wb, computeWaveSpeed, output, sb, eb);
            flux.computeFlux(conservedVariables, wt, computeWaveSpeed, output, st, et);
            for (int k = 0; k < 3; ++k) sides[k] = std::max(sides[k], std::max(wb[k], wt[k]));
        }
        for (int k = 0; k < 3; ++k) waveSpeed[k] = std::max(waveSpeed[k], sides[k]);
    }

private:
    NumericalFlux& flux;
};

}  // namespace hip
}  // namespace alsfvm

namespace alsuq {
namespace hip {

using alsvinn_hip_cxx::check;
using alsvinn_hip_cxx::Grid;

// generator::STLMersenne + distribution::Uniform (STLMersenne.cpp:25-85,
// Uniform.cpp:27-34): one engine shared by the process, draws addressed by
// (sample, component) and fast-forwarded with discard.
class STLMersenneUniform {
public:
    STLMersenneUniform(long dimension, double lower = 0, double upper = 1)
        : dimension(dimension), lower(lower), upper(upper) {}
    double generate(long component, long sample) {
        if (component >= dimension) throw std::runtime_error("Component given higher than dimension");
        if (count / dimension < sample) {
            const long add = sample * dimension - count;
            engine.discard(add);
            count += add;
        }
        if (count % dimension < component) {
            const long add = component - count % dimension;
            engine.discard(add);
            count += add;
        }
        count++;
        return dist(engine) * (upper - lower) + lower;
    }

private:
    std::mt19937_64 engine;
    long count = 0;
    std::uniform_real_distribution<double> dist{0.0, 1.0};
    long dimension;
    double lower, upper;
};

// run::Runner: every sample of [sampleStart, sampleStart + samples) is set up
// from its parameter vector and all advance together to the end time.
class Runner {
public:
    using InitialData = std::function<std::vector<double>(const std::vector<double>& parameters,
                                                          long sample)>;

    Runner(const alsvinn_hip_scheme& scheme, const Grid& grid, double endTime, long samples,
           long parameterDimension, InitialData init, long sampleStart = 0)
        : scheme(scheme), grid(grid), endTime(endTime), samples(samples) {
        std::vector<alsvinn_hip_sample> specs(samples, alsvinn_hip_cxx::make_sample(grid, endTime, 1));
        check(alsvinn_hip_ensemble_create(&scheme, (int)samples, specs.data(), nullptr, &h.e),
              "Runner");
        STLMersenneUniform gen(parameterDimension);
        for (long i = 0; i < samples; ++i) {
            std::vector<double> prm(parameterDimension);
            for (long c = 0; c < parameterDimension; ++c) prm[c] = gen.generate(c, sampleStart + i);
            const auto u = init(prm, sampleStart + i);
            check(alsvinn_hip_ensemble_upload(h.e, (int)i, u.data()), "Runner");
        }
        check(alsvinn_hip_ensemble_apply_boundary(h.e), "Runner");
    }

    void run(int stepsPerSync = 32) { check(alsvinn_hip_ensemble_run(h.e, stepsPerSync), "run"); }

    alsvinn_hip_ensemble* handle() { return h.e; }
    long getNumberOfSamples() const { return samples; }

    std::vector<double> getSample(int i) {
        int d[6];
        long long off = 0, n = 0;
        check(alsvinn_hip_ensemble_layout(h.e, i, d, &off, &n), "layout");
        std::vector<double> out((size_t)n * alsvinn_hip_cxx::number_of_variables(scheme));
        check(alsvinn_hip_ensemble_download(h.e, i, out.data()), "getSample");
        return out;
    }

private:
    alsvinn_hip_scheme scheme;
    Grid grid;
    double endTime;
    long samples;
    alsvinn_hip_cxx::EnsembleHandle h;
};

// stats::MeanVariance over all samples of a Runner (combine + finalize:
// mean = S1/M, variance = S2/M - mean^2).
class MeanVariance {
public:
    explicit MeanVariance(Runner& r) : runner(r) {}
    void computeStatistics(std::vector<double>& mean, std::vector<double>& variance,
                           bool welford = false) {
        std::vector<int> idx(runner.getNumberOfSamples());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        const size_t n = runner.getSample(0).size();
        mean.resize(n);
        variance.resize(n);
        check(alsvinn_hip_stats_meanvar(runner.handle(), idx.data(), (int)idx.size(), welford ? 1 : 0,
                                        mean.data(), variance.data()),
              "MeanVariance");
    }

private:
    Runner& runner;
};

// stats::StructureCube (p, numberOfH) over all samples of a Runner.
class StructureCube {
public:
    StructureCube(Runner& r, double p, int numberOfH) : runner(r), p(p), numberOfH(numberOfH) {}
    std::vector<double> computeStatistics(int nvar) {
        std::vector<int> idx(runner.getNumberOfSamples());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        std::vector<double> out((size_t)nvar * numberOfH);
        check(alsvinn_hip_stats_structure_cube(runner.handle(), idx.data(), (int)idx.size(), numberOfH,
                                               p, out.data()),
              "StructureCube");
        return out;
    }

private:
    Runner& runner;
    double p;
    int numberOfH;
};

// stats::StructureBasic (p, numberOfH, direction) over all samples of a Runner.
class StructureBasic {
public:
    StructureBasic(Runner& r, double p, int numberOfH, int direction)
        : runner(r), p(p), numberOfH(numberOfH), direction(direction) {}
    std::vector<double> computeStatistics(int nvar) {
        std::vector<int> idx = all_samples();
        std::vector<double> out((size_t)nvar * numberOfH);
        check(alsvinn_hip_stats_structure_basic(runner.handle(), idx.data(), (int)idx.size(),
                                                numberOfH, p, direction, out.data()),
              "StructureBasic");
        return out;
    }

private:
    std::vector<int> all_samples() const {
        std::vector<int> idx(runner.getNumberOfSamples());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        return idx;
    }
    Runner& runner;
    double p;
    int numberOfH, direction;
};

// stats::OnePointMoment (integer p, OnePointMoment.cpp:21-45): per cell, the
// mean over all samples of pow(u, p), in the volume layout.
class OnePointMoment {
public:
    OnePointMoment(Runner& r, int p) : runner(r), p(p) {}
    std::vector<double> computeStatistics() {
        std::vector<int> idx(runner.getNumberOfSamples());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        std::vector<double> out(runner.getSample(0).size());
        check(alsvinn_hip_stats_onepoint_moment(runner.handle(), idx.data(), (int)idx.size(),
                                                double(p), out.data()),
              "OnePointMoment");
        return out;
    }

private:
    Runner& runner;
    int p;
};

// stats::BoundedVariation (integer p, BoundedVariation.cpp:23-53): per
// variable, the mean over all samples of the total variation; 1D and 2D.
class BoundedVariation {
public:
    BoundedVariation(Runner& r, int p) : runner(r), p(p) {}
    std::vector<double> computeStatistics(int nvar) {
        std::vector<int> idx(runner.getNumberOfSamples());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        std::vector<double> out(nvar);
        check(alsvinn_hip_stats_bounded_variation(runner.handle(), idx.data(), (int)idx.size(),
                                                  double(p), out.data()),
              "BoundedVariation");
        return out;
    }

private:
    Runner& runner;
    int p;
};

// alsuq::stats::Statistics of the hip platform fed one device Volume at a
// time, as the reference's Runner feeds its statistics (Runner.cpp:42-87):
// computeStatistics per sample, then combineStatistics (one rank: divide by
// the number of samples) and finalizeStatistics. kind: ALSVINN_HIP_STAT_*.
class Statistics {
public:
    Statistics(int kind, const alsvinn_hip_volume& shape, double p = 2.0, int numberOfH = 1,
               int direction = 0, int boundary = ALSVINN_HIP_BC_PERIODIC, void* stream = nullptr) {
        check(alsvinn_hip_statistics_create(kind, p, numberOfH, direction, boundary, &shape, stream,
                                            &st),
              "Statistics");
        check(alsvinn_hip_statistics_outputs(st, &outputs, &elements), "Statistics");
    }
    ~Statistics() { alsvinn_hip_statistics_destroy(st); }
    Statistics(const Statistics&) = delete;
    Statistics& operator=(const Statistics&) = delete;

    void computeStatistics(const alsvinn_hip_volume& conservedVariables) {
        check(alsvinn_hip_statistics_compute(st, &conservedVariables), "computeStatistics");
    }
    void combineStatistics(long long samples) {
        check(alsvinn_hip_statistics_combine(st, samples), "combineStatistics");
    }
    void finalizeStatistics() { check(alsvinn_hip_statistics_finalize(st), "finalizeStatistics"); }
    // Host copy of output i (meanvar: 0 mean, 1 variance).
    std::vector<double> getStatistics(int i = 0) {
        std::vector<double> out((size_t)elements);
        check(alsvinn_hip_statistics_download(st, i, out.data()), "getStatistics");
        return out;
    }
    // Device pointer of output i (FP64), for a writer that reads device memory.
    double* getDevicePointer(int i = 0) {
        double* p = nullptr;
        check(alsvinn_hip_statistics_result(st, i, &p), "getDevicePointer");
        return p;
    }
    int getNumberOfOutputs() const { return outputs; }

private:
    alsvinn_hip_statistics* st = nullptr;
    int outputs = 0;
    long long elements = 0;
};

}  // namespace hip
}  // namespace alsuq

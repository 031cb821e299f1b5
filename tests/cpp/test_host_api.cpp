// The C++ host layer (include/alsvinn_hip.hpp) against the oracle, on the GPU:
// the reference-style Simulator loop, ConservedSystem from NumericalFlux::computeFlux
// regions on device Volumes (hipMalloc'd per variable), Boundary, the batched
// Runner with the STLMersenne sampler, MeanVariance, OnePointMoment,
// BoundedVariation, StructureBasic and the per-sample Statistics. Prints one
// line per case and exits non-zero on the first mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/alsvinn_hip.hpp"
#include "../../oracle/oracle.h"

static int failures = 0;
#define EXPECT(cond, msg)                                  \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("FAIL %s\n", msg);                 \
            failures++;                                    \
        } else {                                           \
            std::printf("ok   %s\n", msg);                 \
        }                                                  \
        std::fflush(stdout);                               \
    } while (0)

static bool same_bits(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * 8) == 0;
}

// A device Volume: one hipMalloc per variable (as a Memory<real> per
// variable of the reference's Volume).
struct DeviceVolume {
    alsvinn_hip_volume v;
    size_t per_var = 0;
    DeviceVolume(int nvar, int nx, int ny, int nz, int g) {
        std::memset(&v, 0, sizeof(v));
        v.nvar = nvar;
        v.precision = ALSVINN_HIP_FP64;
        v.nx = nx;
        v.ny = ny;
        v.nz = nz;
        v.gx = g;
        v.gy = ny > 1 ? g : 0;
        v.gz = nz > 1 ? g : 0;
        per_var = (size_t)(nx + 2 * v.gx) * (ny + 2 * v.gy) * (nz + 2 * v.gz);
        for (int k = 0; k < nvar; ++k)
            if (hipMalloc(&v.var[k], per_var * 8) != hipSuccess) throw std::runtime_error("hipMalloc");
    }
    ~DeviceVolume() {
        for (int k = 0; k < v.nvar; ++k) (void)hipFree(v.var[k]);
    }
    void upload(const std::vector<double>& host) {
        for (int k = 0; k < v.nvar; ++k)
            if (hipMemcpy(v.var[k], host.data() + k * per_var, per_var * 8, hipMemcpyHostToDevice) !=
                hipSuccess)
                throw std::runtime_error("hipMemcpy");
    }
    std::vector<double> download() {
        std::vector<double> host(per_var * v.nvar);
        if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize");
        for (int k = 0; k < v.nvar; ++k)
            if (hipMemcpy(host.data() + k * per_var, v.var[k], per_var * 8, hipMemcpyDeviceToHost) !=
                hipSuccess)
                throw std::runtime_error("hipMemcpy");
        return host;
    }
};

static oracle_setup setup_of(const alsvinn_hip_scheme& s, const alsvinn_hip_cxx::Grid& g, double T) {
    oracle_setup o;
    std::memset(&o, 0, sizeof(o));
    o.equation = s.equation;
    o.flux = s.flux;
    o.recon = s.reconstruction;
    o.integrator = s.integrator;
    o.boundary = s.boundary;
    o.nx = g.n[0];
    o.ny = g.n[1];
    o.nz = g.n[2];
    for (int k = 0; k < 3; ++k) {
        o.origin[k] = g.origin[k];
        o.top[k] = g.top[k];
    }
    o.gamma = s.gamma;
    o.cfl = s.cfl;
    o.end_time = T;
    o.num_adjusters = 1;
    o.adj_interval[0] = T;
    o.adj_saved0[0] = 1;
    o.max_steps = -1;
    return o;
}

// Kelvin-Helmholtz initial data from a 40-vector (band 0.25+s0 < y < 0.75+s1).
static std::vector<double> kh(int n, const std::vector<double>& X) {
    const int g = 2, t = n + 2 * g;
    std::vector<double> u(4 * (size_t)t * t, 0.0);
    const double pi = M_PI;
    auto pert = [&](double x, int a0, int b0) {
        double norm = 0;
        for (int k = 0; k < 10; ++k) norm += X[a0 + k];
        double S = 0;
        for (int k = 0; k < 10; ++k)
            S += X[a0 + k] * std::cos((X[b0 + k] * (pi - (-pi)) + (-pi)) + (2 * pi * (k + 1)) * x);
        return (0.01 * S) / norm;
    };
    const double dx = 1.0 / n;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const double x = 0.0 + dx * i + dx / 2, y = 0.0 + dx * j + dx / 2;
            const bool mid = y > 0.25 + pert(x, 0, 20) && y < 0.75 + pert(x, 10, 30);
            const double rho = mid ? 2.0 : 1.0, ux = mid ? -0.5 : 0.5, p = 2.5;
            const size_t c = (size_t)(j + g) * t + (i + g), v = (size_t)t * t;
            u[c] = rho;
            u[v + c] = rho * ux;
            u[2 * v + c] = rho * 0.0;
            u[3 * v + c] = p / (1.4 - 1) + (0.5 * rho) * (ux * ux + 0.0 * 0.0);
        }
    return u;
}

int main() {
    // 1. Sod through the reference-style step loop
    {
        alsvinn_hip_scheme s = {ALSVINN_HIP_EQ_EULER1, ALSVINN_HIP_FLUX_HLL, ALSVINN_HIP_RECON_NONE,
                                ALSVINN_HIP_INTEGRATOR_FORWARD_EULER, ALSVINN_HIP_BC_NEUMANN,
                                ALSVINN_HIP_FP64, 1.4, 0.9};
        alsvinn_hip_cxx::Grid g;
        g.origin[0] = -1;
        g.top[0] = 1;
        g.n[0] = 1024;
        alsfvm::hip::Simulator sim(s, g, 0.2);
        std::vector<double> u(3 * 1026, 0.0);
        for (int i = 0; i < 1024; ++i) {
            const double x = -1.0 + (2.0 / 1024) * i + (2.0 / 1024) / 2;
            const double rho = x < 0 ? 1.0 : 0.125, p = x < 0 ? 1.0 : 0.1;
            u[1 + i] = rho;
            u[1026 + 1 + i] = 0.0;
            u[2 * 1026 + 1 + i] = p / (1.4 - 1) + (0.5 * rho) * 0.0;
        }
        sim.setInitialValue(u);
        while (!sim.atEnd()) sim.performStep();
        const auto got = sim.getConservedVolume();
        oracle_setup o = setup_of(s, g, 0.2);
        std::vector<double> ref(u.size());
        double t;
        long steps;
        oracle_run(&o, u.data(), ref.data(), &t, &steps, nullptr, 0);
        EXPECT(steps == sim.getTimestepsPerformed() && t == sim.getCurrentTime(), "sod steps and time");
        EXPECT(same_bits(ref, got), "sod conserved variables bit-identical");
    }
    // 2. Runner + MeanVariance on KH 32^2 x 3 samples
    alsvinn_hip_scheme kh_s = {ALSVINN_HIP_EQ_EULER2, ALSVINN_HIP_FLUX_HLL3, ALSVINN_HIP_RECON_WENO2,
                               ALSVINN_HIP_INTEGRATOR_RK3, ALSVINN_HIP_BC_PERIODIC, ALSVINN_HIP_FP64,
                               1.4, 0.475};
    alsvinn_hip_cxx::Grid g2;
    g2.top[2] = 0;
    g2.n[0] = g2.n[1] = 32;
    {
        const double T = 0.1;
        alsuq::hip::Runner runner(kh_s, g2, T, 3, 40,
                                  [](const std::vector<double>& X, long) { return kh(32, X); });
        runner.run();
        std::vector<double> X(3 * 40);
        oracle_stl_mersenne_uniform(0, 3, 40, 0.0, 1.0, X.data());
        oracle_setup o = setup_of(kh_s, g2, T);
        std::vector<double> all;
        bool ok = true;
        for (int i = 0; i < 3; ++i) {
            const auto u0 = kh(32, std::vector<double>(X.begin() + 40 * i, X.begin() + 40 * (i + 1)));
            std::vector<double> ref(u0.size());
            double t;
            long steps;
            oracle_run(&o, u0.data(), ref.data(), &t, &steps, nullptr, 0);
            ok = ok && same_bits(ref, runner.getSample(i));
            all.insert(all.end(), ref.begin(), ref.end());
        }
        EXPECT(ok, "runner samples bit-identical to oracle runs");
        std::vector<double> mean, var, mref(all.size() / 3), vref(all.size() / 3);
        alsuq::hip::MeanVariance mv(runner);
        mv.computeStatistics(mean, var);
        oracle_meanvar(3, (long)mref.size(), all.data(), mref.data(), vref.data());
        EXPECT(same_bits(mean, mref) && same_bits(var, vref), "meanvar bit-identical");
        // OnePointMoment p = 2 and BoundedVariation p = 1: sequential sums, bit for bit
        std::vector<double> m2ref(mref.size());
        oracle_onepoint_moment(3, (long)m2ref.size(), all.data(), 2.0, m2ref.data());
        EXPECT(same_bits(alsuq::hip::OnePointMoment(runner, 2).computeStatistics(), m2ref),
               "onepointmoment bit-identical");
        const size_t vol = (size_t)36 * 36;
        std::vector<double> bvref(4, 0.0);
        for (int v = 0; v < 4; ++v) {
            double acc = 0;
            for (int k = 0; k < 3; ++k)
                acc += oracle_total_variation(36, 36, 1, 2, 2, 0, all.data() + (size_t)k * 4 * vol + v * vol,
                                              1.0);
            bvref[v] = acc / 3;
        }
        EXPECT(same_bits(alsuq::hip::BoundedVariation(runner, 1).computeStatistics(4), bvref),
               "bounded variation bit-identical");
        // StructureBasic in y, p = 2: parallel summation, 1e-12 relative
        std::vector<double> sbref(4 * 5, 0.0);
        for (int k = 0; k < 3; ++k)
            oracle_structure_basic(4, 32, 32, 1, 2, all.data() + (size_t)k * 4 * vol, 5, 2.0, 1, 1,
                                   sbref.data());
        const auto sb = alsuq::hip::StructureBasic(runner, 2.0, 5, 1).computeStatistics(4);
        bool sok = sb.size() == sbref.size();
        for (size_t i = 0; sok && i < sb.size(); ++i)
            sok = std::fabs(sb[i] - sbref[i] / 3) <= 1e-12 * std::fabs(sbref[i] / 3);
        EXPECT(sok, "structure_basic within 1e-12");
    }
    // 3. ConservedSystem from NumericalFlux::computeFlux regions on device
    //    Volumes, and Boundary, vs the oracle
    {
        std::vector<double> X(40);
        oracle_stl_mersenne_uniform(5, 1, 40, 0.0, 1.0, X.data());
        auto u = kh(32, X);
        for (size_t i = 0; i < u.size() / 4; ++i) u[i] *= 1.0 + 0.01 * std::sin(0.37 * i);
        oracle_setup o = setup_of(kh_s, g2, 1.0);
        auto ub = u;
        oracle_apply_boundary(&o, ub.data());
        DeviceVolume U(4, 32, 32, 1, 2), O(4, 32, 32, 1, 2);
        U.upload(u);
        alsfvm::hip::Boundary(ALSVINN_HIP_BC_PERIODIC).applyBoundaryConditions(U.v);
        EXPECT(same_bits(U.download(), ub), "Boundary::applyBoundaryConditions bit-identical");
        std::vector<double> ref(u.size(), 0.0);
        O.upload(ref);
        double wsr[3], ws[3] = {0, 0, 0};
        oracle_system_rhs(&o, ub.data(), ref.data(), wsr);
        alsfvm::hip::NumericalFlux flux(kh_s, g2);
        alsfvm::hip::ConservedSystem system(flux);
        system(U.v, ws, true, O.v);
        EXPECT(flux.getNumberOfGhostCells() == 2, "getNumberOfGhostCells");
        EXPECT(ws[0] == wsr[0] && ws[1] == wsr[1], "ConservedSystem wave speeds bit-identical");
        EXPECT(same_bits(O.download(), ref), "ConservedSystem divergence bit-identical (every cell)");
        // per-sample Statistics over two device samples vs the oracle's meanvar
        DeviceVolume V(4, 32, 32, 1, 2);
        auto u2 = ub;
        for (auto& x : u2) x *= 1.25;
        V.upload(u2);
        alsuq::hip::Statistics mv(ALSVINN_HIP_STAT_MEANVAR, U.v);
        mv.computeStatistics(U.v);
        mv.computeStatistics(V.v);
        mv.combineStatistics(2);
        mv.finalizeStatistics();
        std::vector<double> both(ub);
        both.insert(both.end(), u2.begin(), u2.end());
        std::vector<double> mref(ub.size()), vref(ub.size());
        oracle_meanvar(2, (long)ub.size(), both.data(), mref.data(), vref.data());
        EXPECT(same_bits(mv.getStatistics(0), mref) && same_bits(mv.getStatistics(1), vref),
               "per-sample Statistics meanvar bit-identical");
    }
    // 4. error behaviour: an unknown flux throws
    {
        alsvinn_hip_scheme bad = kh_s;
        bad.flux = ALSVINN_HIP_FLUX_GODUNOV;
        bool threw = false;
        try {
            alsfvm::hip::Simulator sim(bad, g2, 1.0);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        EXPECT(threw, "invalid scheme throws std::runtime_error");
    }
    std::printf(failures ? "FAILED\n" : "PASSED\n");
    return failures ? 1 : 0;
}

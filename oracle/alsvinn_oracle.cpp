// CPU oracle for the alsfvm finite-volume timestep and the alsuq sample statistics.
//
// TEST INFRASTRUCTURE ONLY. This file is a clean-room restatement of the
// reference's `cpu` platform, written so that the HIP product can be checked
// cell-for-cell against it. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load the library built from it; the product
// never links or calls it.
//
// What it restates (reference file:line, paths relative to the alsvinn tree):
//   Euler helpers          alsfvm/include/alsfvm/equation/euler/Euler.hpp:172-288
//   vector dot / division   alsutils/include/alsutils/vec{1,2,3}.hpp (component-wise)
//   HLL flux               alsfvm/include/alsfvm/numflux/euler/HLL.hpp:51-115
//   HLL3 flux              alsfvm/include/alsfvm/numflux/euler/HLL3.hpp:52-148
//   Burgers Godunov        alsfvm/include/alsfvm/numflux/burgers/Godunov.hpp:40-60
//   face flux stencil      alsfvm/include/alsfvm/numflux/numflux_util.hpp:62-92
//   net flux / makeZero    alsfvm/src/numflux/NumericalFluxCPU.cpp:32-146,170-231
//   WENO2                  alsfvm/include/alsfvm/reconstruction/WENO2.hpp:31-95 (eps 1e-8,
//                          WENOCoefficients.hpp:21); ranges ReconstructionCPU.cpp:38-97
//   ENO2, ENO3, ENO4       alsfvm/src/reconstruction/ENOCPU.cpp:41-267, ENOCoefficients.cpp:27-50
//   WENO3                  alsfvm/src/reconstruction/WENOCPU.cpp:32-186,
//                          WENOCoefficients.hpp:51-63, WENOCoefficients.cpp:26-38
//   WENOF2                 alsfvm/include/alsfvm/reconstruction/WENOF2.hpp:31-115,
//                          euler/PrimitiveVariables.hpp, Euler.hpp:236-245,290-297
//   no reconstruction      alsfvm/src/reconstruction/NoReconstruction.cpp:42-73
//   region split           alsfvm/src/simulator/ConservedSystem.cpp:31-103
//   CFL + adjusters        alsfvm/src/integrator/Integrator.cpp:20-66,
//                          alsfvm/src/io/FixedIntervalWriter.cpp:45-53,
//                          alsuq/src/stats/FixedIntervalStatistics.cpp:33-41
//   integrators            ForwardEuler.cpp:32-51, RungeKutta2.cpp:37-67, RungeKutta3.cpp:49-86,
//                          RungeKutta4.cpp:50-96
//   boundary               alsfvm/src/boundary/BoundaryCPU.cpp:30-121, Neumann.hpp:24-35,
//                          Periodic.hpp:24-42
//   simulator loop         alsfvm/src/simulator/Simulator.cpp:77-85,180-214
//   constraints            alsfvm/src/equation/CPUCellComputer.cpp:79-125
//   grid                   alsfvm/src/grid/Grid.cpp:24-41,148-171
//   sampler                alsuq/src/generator/STLMersenne.cpp:25-85, Uniform.cpp:27-34
//   meanvar                alsuq/src/stats/MeanVariance.cpp:30-67, StatisticsHelper.cpp:35-79
//   onepointmoment         alsuq/src/stats/OnePointMoment.cpp:30-45, HostMemory.cpp:283-290
//   bv                     alsuq/src/stats/BoundedVariation.cpp:38-53, HostMemory.cpp:325-362
//   structure functions    alsuq/src/stats/StructureBasic.cpp:67-120,
//                          alsfvm/include/alsfvm/functional/structure_common.hpp:37-180,
//                          alsfvm/include/alsfvm/boundary/ValueAtBoundary.hpp:20-98
//
// Every expression keeps the reference's association order; the file must be
// compiled with -ffp-contract=off and without fast-math.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace orc {

typedef double real;

struct ivec3 {
    int x = 0, y = 0, z = 0;
    int& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    int operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

static ivec3 iv(int x, int y, int z) {
    ivec3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}

// Structure-of-arrays volume with ghost layers only on active axes
// (Volume.cpp:44-60), index z*tx*ty + y*tx + x (View.hpp:116-118).
struct Vol {
    int nvar, nx, ny, nz, ngx, ngy, ngz, tx, ty, tz;
    std::vector<real> buf;
    Vol(int nvar_, int nx_, int ny_, int nz_, int ng)
        : nvar(nvar_), nx(nx_), ny(ny_), nz(nz_), ngx(ng), ngy(ny_ > 1 ? ng : 0),
          ngz(nz_ > 1 ? ng : 0) {
        tx = nx + 2 * ngx;
        ty = ny + 2 * ngy;
        tz = nz + 2 * ngz;
        // HostMemory starts filled with 42 (HostMemory.cpp:36-40).
        buf.assign((size_t)nvar * cells(), real(42));
    }
    size_t cells() const { return (size_t)tx * ty * tz; }
    real* v(int var) { return buf.data() + (size_t)var * cells(); }
    const real* v(int var) const { return buf.data() + (size_t)var * cells(); }
    size_t idx(int x, int y, int z) const {
        return (size_t)z * tx * ty + (size_t)y * tx + (size_t)x;
    }
    int dims() const { return 1 + (ny > 1) + (nz > 1); }
    ivec3 total() const { return iv(tx, ty, tz); }
    ivec3 ghosts() const { return iv(ngx, ngy, ngz); }
};

// ---------------------------------------------------------------- equations

template <int NSD>
struct Euler {
    static constexpr int nvar = NSD + 2;
    static constexpr int nsd = NSD;
    real gamma;
    struct All {
        real q[nvar];
        real u[NSD];
        real p;
    };
    // Euler.hpp:204-211 computeExtra: ie = E - 0.5*(m.m)/rho, u = m/rho, p = (gamma-1)*ie
    void extra(const real* q, real* u, real& p) const {
        real dot = q[1] * q[1];
        if (NSD > 1) dot = dot + q[2] * q[2];
        if (NSD > 2) dot = dot + q[3] * q[3];
        const real ie = q[nvar - 1] - 0.5 * dot / q[0];
        for (int k = 0; k < NSD; ++k) u[k] = q[1 + k] / q[0];
        p = (gamma - 1) * ie;
    }
    All make(const real* q) const {
        All a;
        for (int i = 0; i < nvar; ++i) a.q[i] = q[i];
        extra(q, a.u, a.p);
        return a;
    }
    // Euler.hpp:172-181
    template <int d>
    void pointFlux(const All& a, real* F) const {
        F[0] = a.q[1 + d];
        for (int k = 0; k < NSD; ++k) F[1 + k] = a.q[1 + k] * a.u[d];
        F[1 + d] += a.p;
        F[nvar - 1] = (a.q[nvar - 1] + a.p) * a.u[d];
    }
    static real weight(const real* q) { return q[0]; }
    // Euler.hpp:251-258 computeWaveSpeed: |u_d| + sqrt(gamma * p / rho)
    template <int d>
    real waveSpeed(const All& a) const {
        return std::fabs(a.u[d]) + std::sqrt(gamma * a.p / a.q[0]);
    }
    // CPUCellComputer.cpp:79-125 + Euler.hpp:269-277
    bool obeys(const real* q) const {
        real u[NSD], p;
        extra(q, u, p);
        bool ok = true;
        for (int i = 0; i < nvar; ++i)
            if (std::isnan(q[i]) || std::isinf(q[i])) ok = false;
        for (int k = 0; k < NSD; ++k)
            if (std::isnan(u[k]) || std::isinf(u[k])) ok = false;
        if (std::isnan(p) || std::isinf(p)) ok = false;
        if (!(std::isfinite(q[0]) && (q[0] == q[0]) && (q[0] > 0) && (p > 0))) ok = false;
        return ok;
    }
};

struct Burgers {
    static constexpr int nvar = 1;
    static constexpr int nsd = 3;
    real gamma;
    struct All {
        real q[1];
    };
    All make(const real* q) const {
        All a;
        a.q[0] = q[0];
        return a;
    }
    // Burgers.hpp:158-164: F = u*u/2 in every direction
    template <int d>
    void pointFlux(const All& a, real* F) const {
        F[0] = a.q[0] * a.q[0] / 2;
    }
    static real weight(const real* q) { return q[0]; }
    // Burgers.hpp:203-209 computeWaveSpeed = |u|; 298-301 computeEigenValues = u
    template <int d>
    real waveSpeed(const All& a) const {
        return std::fabs(a.q[0]);
    }
    bool obeys(const real* q) const {
        if (std::isnan(q[0]) || std::isinf(q[0])) return false;
        return q[0] < INFINITY && (q[0] == q[0]);
    }
};

// ---------------------------------------------------------------- fluxes

struct HLL {
    // HLL.hpp:51-115
    template <int d, int NSD>
    static real compute(const Euler<NSD>& eq, const typename Euler<NSD>::All& L,
                        const typename Euler<NSD>::All& R, real* F) {
        constexpr int nv = NSD + 2;
        const real g = eq.gamma;
        const real wL = std::sqrt(L.q[0]);
        const real wR = std::sqrt(R.q[0]);
        const real rho = (L.q[0] + R.q[0]) / 2;
        real u[NSD];
        for (int k = 0; k < NSD; ++k) u[k] = (L.u[k] * wL + R.u[k] * wR) / (wL + wR);
        const real p = (L.p * wL + R.p * wR) / (wL + wR);
        const real cs = std::sqrt(g * p / rho);
        const real sL = std::fmin(L.u[d] - std::sqrt(g * L.p / L.q[0]), u[d] - cs);
        const real sR = std::fmax(R.u[d] + std::sqrt(g * R.p / R.q[0]), u[d] + cs);
        if (sL == 0) {
            for (int i = 0; i < nv; ++i) F[i] = 0;
        } else if (sL > 0) {
            eq.template pointFlux<d>(L, F);
        } else if (sR < 0) {
            eq.template pointFlux<d>(R, F);
        } else {
            real fL[nv], fR[nv];
            eq.template pointFlux<d>(L, fL);
            eq.template pointFlux<d>(R, fR);
            for (int i = 0; i < nv; ++i)
                F[i] = (fL[i] * sR - fR[i] * sL + (R.q[i] - L.q[i]) * (sR * sL)) / (sR - sL);
        }
        return std::fmax(std::fabs(sL), std::fabs(sR));
    }
};

struct HLL3 {
    // HLL3.hpp:52-148
    template <int d, int NSD>
    static real compute(const Euler<NSD>& eq, const typename Euler<NSD>::All& L,
                        const typename Euler<NSD>::All& R, real* F) {
        constexpr int nv = NSD + 2;
        const real g = eq.gamma;
        const real wL = std::sqrt(L.q[0]);
        const real wR = std::sqrt(R.q[0]);
        const real rho = (L.q[0] + R.q[0]) / 2;
        real u[NSD];
        for (int k = 0; k < NSD; ++k) u[k] = (L.u[k] * wL + R.u[k] * wR) / (wL + wR);
        const real p = (wL * L.p + wR * R.p) / (wL + wR);
        const real cfL = std::sqrt(g * L.p / L.q[0]);
        const real cfR = std::sqrt(g * R.p / R.q[0]);
        const real corr = 0.5 * std::fmax(real(0), L.u[d] - R.u[d]);
        const real cs = std::sqrt(g * p / rho);
        const real sL = std::fmin(L.u[d] + corr - cfL, u[d] - cs);
        const real sR = std::fmax(R.u[d] - corr + cfR, u[d] + cs);

        real fL[nv], fR[nv];
        eq.template pointFlux<d>(L, fL);
        eq.template pointFlux<d>(R, fR);
        const real pL = L.p, pR = R.p;
        const real udl = L.u[d] - sL;
        const real udr = R.u[d] - sR;
        const real aa = udr * R.q[0] - udl * L.q[0];
        const real sm = (R.q[1 + d] * udr - L.q[1 + d] * udl + pR - pL) / aa;
        real us[NSD];
        for (int k = 0; k < NSD; ++k)
            us[k] = (fR[1 + k] - fL[1 + k] - R.q[1 + k] * sR + L.q[1 + k] * sL) / aa;
        us[d] = sm;

        if (sL == 0) {
            for (int i = 0; i < nv; ++i) F[i] = 0;
        } else if (sL > 0) {
            for (int i = 0; i < nv; ++i) F[i] = fL[i];
        } else if (sR < 0) {
            for (int i = 0; i < nv; ++i) F[i] = fR[i];
        } else if (sm >= 0) {
            real mid[nv];
            mid[0] = L.q[0] * udl / (sm - sL);
            for (int k = 0; k < NSD; ++k) mid[1 + k] = us[k] * mid[0];
            const real ps = pL + L.q[0] * (L.u[d] - sm) * udl;
            mid[nv - 1] = (udl * L.q[nv - 1] + pL * L.u[d] - ps * sm) / (sm - sL);
            for (int i = 0; i < nv; ++i) F[i] = fL[i] + (mid[i] - L.q[i]) * sL;
        } else {
            real mid[nv];
            mid[0] = R.q[0] * udr / (sm - sR);
            for (int k = 0; k < NSD; ++k) mid[1 + k] = us[k] * mid[0];
            const real ps = pR + R.q[0] * (R.u[d] - sm) * udr;
            mid[nv - 1] = (udr * R.q[nv - 1] + pR * R.u[d] - ps * sm) / (sm - sR);
            for (int i = 0; i < nv; ++i) F[i] = fR[i] + (mid[i] - R.q[i]) * sR;
        }
        return std::fmax(std::fabs(sL), std::fabs(sR));
    }
};

struct Godunov {
    // burgers/Godunov.hpp:40-60
    template <int d>
    static real compute(const Burgers& eq, const Burgers::All& L, const Burgers::All& R,
                        real* F) {
        Burgers::All a, b;
        a.q[0] = std::fmax(L.q[0], real(0.0));
        b.q[0] = std::fmin(R.q[0], real(0.0));
        real fl[1], fr[1];
        eq.pointFlux<d>(a, fl);
        eq.pointFlux<d>(b, fr);
        F[0] = std::fmax(fl[0], fr[0]);
        return std::fmax(std::fabs(L.q[0]), std::fabs(R.q[0]));
    }
};

struct Central {
    // numflux/Central.hpp:38-56: F = 0.5*(f(L) + f(R)), speed max of the two
    template <int d, class Eq>
    static real compute(const Eq& eq, const typename Eq::All& L, const typename Eq::All& R,
                        real* F) {
        constexpr int nv = Eq::nvar;
        real fl[nv], fr[nv];
        eq.template pointFlux<d>(L, fl);
        eq.template pointFlux<d>(R, fr);
        for (int i = 0; i < nv; ++i) F[i] = 0.5 * (fl[i] + fr[i]);
        return std::fmax(eq.template waveSpeed<d>(L), eq.template waveSpeed<d>(R));
    }
};

struct Rusanov {
    // numflux/Rusanov.hpp:38-60 for Burgers (eigenvalue u):
    // F = 0.5*(f(L) + f(R)) - 0.5*max(|uL|, |uR|)*(uR - uL)
    template <int d>
    static real compute(const Burgers& eq, const Burgers::All& L, const Burgers::All& R,
                        real* F) {
        real fl[1], fr[1];
        eq.pointFlux<d>(L, fl);
        eq.pointFlux<d>(R, fr);
        const real eigenLeft = std::fabs(L.q[0]);
        const real eigenRight = std::fabs(R.q[0]);
        const real eigenMax = std::fmax(eigenLeft, eigenRight);
        F[0] = 0.5 * (fl[0] + fr[0]) - 0.5 * eigenMax * (R.q[0] - L.q[0]);
        return std::fmax(eq.waveSpeed<d>(L), eq.waveSpeed<d>(R));
    }
};

struct Roe {
    // numflux/burgers/Roe.hpp:38-62: upwind by the sign of uL + uR
    template <int d>
    static real compute(const Burgers& eq, const Burgers::All& L, const Burgers::All& R,
                        real* F) {
        Burgers::All a;
        if (L.q[0] >= -R.q[0])
            a.q[0] = std::fmax(L.q[0], real(0.0));
        else
            a.q[0] = std::fmin(R.q[0], real(0.0));
        eq.pointFlux<d>(a, F);
        return std::fmax(eq.waveSpeed<d>(L), eq.waveSpeed<d>(R));
    }
};

// ---------------------------------------------------------------- reconstruction

enum {
    RECON_NONE = 0,
    RECON_WENO2 = 1,
    RECON_ENO2 = 2,
    RECON_MC = 3,
    RECON_MM = 4,
    RECON_WENOF2 = 5,
    RECON_ENO3 = 6,
    RECON_ENO4 = 7,
    RECON_WENO3 = 8
};

static int reconGhosts(int recon) {
    if (recon == RECON_NONE) return 1;
    if (recon == RECON_ENO3) return 3;
    if (recon == RECON_ENO4) return 4;
    if (recon == RECON_WENO3) return 3;
    return 2;
}

// reconstruction/sign.hpp and minmod.hpp
static real refSign(real a) {
    if (a > 0) return 1;
    if (a == 0) return 0;
    return -1;
}
static real minmod(real a, real b) {
    if (refSign(a) == refSign(b)) return refSign(a) * std::fmin(std::fabs(a), std::fabs(b));
    return 0;
}
static real minmod(real a, real b, real c) {
    if (refSign(a) == refSign(b) && refSign(b) == refSign(c))
        return refSign(a) * std::fmin(std::fabs(a), std::fmin(std::fabs(b), std::fabs(c)));
    return 0;
}

// ReconstructionCPU.cpp:38-97 with MC.hpp:29-55 / MM.hpp:29-54: per variable,
// slope sigma from the limiter, left = middle - sigma/2, right = middle + sigma/2.
static void reconLimiter(const Vol& in, int d, Vol& Lo, Vol& Ro, ivec3 start, ivec3 end, bool mc) {
    const int nx = in.tx, ny = in.ty, nz = in.tz;
    const int dim = in.dims();
    const int g = 2;
    const int ngx = g, ngy = (dim > 1) * g, ngz = (dim > 2) * g;
    const int dx = d == 0, dy = d == 1, dz = d == 2;
    const int sx = (ngx - dx) + start.x, sy = (ngy - dy) + start.y, sz = (ngz - dz) + start.z;
    const int ex = nx - (ngx - dx) + end.x, ey = ny - (ngy - dy) + end.y,
              ez = nz - (ngz - dz) + end.z;
    for (int z = sz; z < ez; z++) {
#pragma omp parallel for
        for (int y = sy; y < ey; y++) {
            for (int x = sx; x < ex; x++) {
                const size_t iO = in.idx(x, y, z);
                const size_t iR = in.idx(x + dx, y + dy, z + dz);
                const size_t iL = in.idx(x - dx, y - dy, z - dz);
                for (int var = 0; var < in.nvar; ++var) {
                    const real left = in.v(var)[iL];
                    const real middle = in.v(var)[iO];
                    const real right = in.v(var)[iR];
                    const real sigma = mc ? minmod(2 * (right - middle), (right - left) / 2,
                                                   2 * (middle - left))
                                          : minmod((right - middle), (middle - left));
                    Lo.v(var)[iO] = middle - sigma / 2;
                    Ro.v(var)[iO] = middle + sigma / 2;
                }
            }
        }
    }
}

// NoReconstruction.cpp:42-73: every cell, left = right = input.
static void reconNone(const Vol& in, Vol& L, Vol& R) {
    for (int var = 0; var < in.nvar; ++var) {
        const real* pin = in.v(var);
        real* pl = L.v(var);
        real* pr = R.v(var);
        for (size_t i = 0; i < in.cells(); ++i) {
            pl[i] = pin[i];
            pr[i] = pin[i];
        }
    }
}

// ReconstructionCPU.cpp:38-97 with WENO2.hpp:37-71.
static void reconWeno2(const Vol& in, int d, Vol& Lo, Vol& Ro, ivec3 start, ivec3 end) {
    const int nx = in.tx, ny = in.ty, nz = in.tz;
    const int dim = in.dims();
    const int g = 2;
    const int ngx = g, ngy = (dim > 1) * g, ngz = (dim > 2) * g;
    const int dx = d == 0, dy = d == 1, dz = d == 2;
    const int sx = (ngx - dx) + start.x, sy = (ngy - dy) + start.y, sz = (ngz - dz) + start.z;
    const int ex = nx - (ngx - dx) + end.x, ey = ny - (ngy - dy) + end.y,
              ez = nz - (ngz - dz) + end.z;
    const real eps = 1e-8;
    const real* w = in.v(0);
    for (int z = sz; z < ez; z++) {
#pragma omp parallel for
        for (int y = sy; y < ey; y++) {
            for (int x = sx; x < ex; x++) {
                const size_t iO = in.idx(x, y, z);
                const size_t iR = in.idx(x + dx, y + dy, z + dz);
                const size_t iL = in.idx(x - dx, y - dy, z - dz);
                const real i0 = w[iO];
                const real b0 = (w[iR] - i0) * (w[iR] - i0);
                const real b1 = (i0 - w[iL]) * (i0 - w[iL]);
                const real a0L = 1 / (3 * (eps + b0) * (eps + b0));
                const real a1L = 2 / (3 * (eps + b1) * (eps + b1));
                const real w0L = a0L / (a0L + a1L);
                const real w1L = a1L / (a0L + a1L);
                const real a0R = 2 / (3 * (eps + b0) * (eps + b0));
                const real a1R = 1 / (3 * (eps + b1) * (eps + b1));
                const real w0R = a0R / (a0R + a1R);
                const real w1R = a1R / (a0R + a1R);
                for (int var = 0; var < in.nvar; ++var) {
                    const real* q = in.v(var);
                    Lo.v(var)[iO] = 0.5 * (w1L * q[iL] + (3 * w0L + w1L) * q[iO] - w0L * q[iR]);
                    Ro.v(var)[iO] = 0.5 * (w0R * q[iR] + (3 * w1R + w0R) * q[iO] - w1R * q[iL]);
                }
            }
        }
    }
}

// Primitive state of the Euler equations (euler/PrimitiveVariables.hpp):
// rho, velocity (nsd components), p, with component-wise arithmetic.
struct Prim {
    real rho, u[3], p;
    int nsd;
};
static Prim primOf(const Vol& in, size_t i, int nsd, real gamma) {
    // Euler::computePrimitiveVariables (Euler.hpp:290-297)
    Prim P;
    P.nsd = nsd;
    const real rho = in.v(0)[i];
    real m[3] = {0, 0, 0};
    for (int k = 0; k < nsd; ++k) m[k] = in.v(1 + k)[i];
    const real E = in.v(nsd + 1)[i];
    real mm = m[0] * m[0];
    for (int k = 1; k < nsd; ++k) mm = mm + m[k] * m[k];
    P.rho = rho;
    for (int k = 0; k < 3; ++k) P.u[k] = k < nsd ? m[k] / rho : 0;
    P.p = (gamma - 1) * (E - 0.5 * mm / rho);
    return P;
}
static real udot(const Prim& a, const Prim& b) {
    real d = a.u[0] * b.u[0];
    for (int k = 1; k < a.nsd; ++k) d = d + a.u[k] * b.u[k];
    return d;
}
static Prim lin(real wa, const Prim& a, real wb, const Prim& b) {  // wa*a + wb*b
    Prim r = a;
    r.rho = wa * a.rho + wb * b.rho;
    for (int k = 0; k < 3; ++k) r.u[k] = wa * a.u[k] + wb * b.u[k];
    r.p = wa * a.p + wb * b.p;
    return r;
}
static Prim diff(const Prim& a, const Prim& b) {  // a - b
    Prim r = a;
    r.rho = a.rho - b.rho;
    for (int k = 0; k < 3; ++k) r.u[k] = a.u[k] - b.u[k];
    r.p = a.p - b.p;
    return r;
}

// WENOF2.hpp:31-115 (Euler only) inside ReconstructionCPU.cpp:38-97: WENO2
// weights from the density applied to the primitive slopes; rho and p slopes
// clamped to +-1.9 times the middle value; each side scaled by
// 0.5*sqrt(R/max(R, LL)), R = p/(gamma-1); converted back with
// Euler::computeConserved (Euler.hpp:236-245).
static void reconWenoF2(const Vol& in, int d, Vol& Lo, Vol& Ro, ivec3 start, ivec3 end,
                        real gamma) {
    const int nx = in.tx, ny = in.ty, nz = in.tz;
    const int dim = in.dims();
    const int nsd = in.nvar - 2;
    const int g = 2;
    const int ngx = g, ngy = (dim > 1) * g, ngz = (dim > 2) * g;
    const int dx = d == 0, dy = d == 1, dz = d == 2;
    const int sx = (ngx - dx) + start.x, sy = (ngy - dy) + start.y, sz = (ngz - dz) + start.z;
    const int ex = nx - (ngx - dx) + end.x, ey = ny - (ngy - dy) + end.y,
              ez = nz - (ngz - dz) + end.z;
    const real eps = 1e-8;
    const real BOUND = 1.9;
    const real* w = in.v(0);
    auto store = [&](Vol& out, size_t i, const Prim& P) {
        out.v(0)[i] = P.rho;
        for (int k = 0; k < nsd; ++k) out.v(1 + k)[i] = P.rho * P.u[k];
        out.v(nsd + 1)[i] = P.p / (gamma - 1) + 0.5 * P.rho * udot(P, P);
    };
    for (int z = sz; z < ez; z++)
        for (int y = sy; y < ey; y++)
            for (int x = sx; x < ex; x++) {
                const size_t iO = in.idx(x, y, z);
                const size_t iR = in.idx(x + dx, y + dy, z + dz);
                const size_t iL = in.idx(x - dx, y - dy, z - dz);
                const real i0 = w[iO];
                const real b0 = (w[iR] - i0) * (w[iR] - i0);
                const real b1 = (i0 - w[iL]) * (i0 - w[iL]);
                const real a0L = 1 / (3 * (eps + b0) * (eps + b0));
                const real a1L = 2 / (3 * (eps + b1) * (eps + b1));
                const real w0L = a0L / (a0L + a1L);
                const real w1L = a1L / (a0L + a1L);
                const real a0R = 2 / (3 * (eps + b0) * (eps + b0));
                const real a1R = 1 / (3 * (eps + b1) * (eps + b1));
                const real w0R = a0R / (a0R + a1R);
                const real w1R = a1R / (a0R + a1R);
                const Prim pL = primOf(in, iL, nsd, gamma);
                const Prim pM = primOf(in, iO, nsd, gamma);
                const Prim pR = primOf(in, iR, nsd, gamma);
                const real Rr = pM.p / (gamma - 1);
                for (int side = 0; side < 2; ++side) {
                    const real wl = side == 0 ? w1L : w1R, wr = side == 0 ? w0L : w0R;
                    Prim dP = lin(wl, diff(pM, pL), wr, diff(pR, pM));
                    dP.rho = std::fmax(-BOUND * pM.rho, std::fmin(BOUND * pM.rho, dP.rho));
                    dP.p = std::fmax(-BOUND * pM.p, std::fmin(BOUND * pM.p, dP.p));
                    const real LL = 0.125 * pM.rho * udot(dP, dP) -
                                    0.5 * std::fmin(real(0.0), dP.rho * udot(pM, dP)) +
                                    0.5 * dP.rho * dP.rho * udot(dP, dP) / pM.rho;
                    const real a = 0.5 * std::sqrt(Rr / std::fmax(Rr, LL));
                    // left: middle - a*dP, right: middle + a*dP
                    Prim P = pM;
                    P.rho = side == 0 ? pM.rho - a * dP.rho : pM.rho + a * dP.rho;
                    for (int k = 0; k < 3; ++k)
                        P.u[k] = side == 0 ? pM.u[k] - a * dP.u[k] : pM.u[k] + a * dP.u[k];
                    P.p = side == 0 ? pM.p - a * dP.p : pM.p + a * dP.p;
                    store(side == 0 ? Lo : Ro, iO, P);
                }
            }
}

// ENOCoefficients.cpp:27-50: rows r = 0..order of the reconstruction
// coefficients over cells i-r .. i-r+order-1 (row r left face, r+1 right).
static const real kEno2[3][2] = {{real(3.0 / 2.0), real(-1.0 / 2.0)},
                                 {real(1.0 / 2.0), real(1.0 / 2.0)},
                                 {real(-1.0 / 2.0), real(3.0 / 2.0)}};
static const real kEno3[4][3] = {{real(11.0 / 6.0), real(-7.0 / 6.0), real(1.0 / 3.0)},
                                 {real(1.0 / 3.0), real(5.0 / 6.0), real(-1.0 / 6.0)},
                                 {real(-1.0 / 6.0), real(5.0 / 6.0), real(1.0 / 3.0)},
                                 {real(1.0 / 3.0), real(-7.0 / 6.0), real(11.0 / 6.0)}};
static const real kEno4[5][4] = {
    {real(25.0 / 12.0), real(-23.0 / 12.0), real(13.0 / 12.0), real(-1.0 / 4.0)},
    {real(1.0 / 4.0), real(13.0 / 12.0), real(-5.0 / 12.0), real(1.0 / 12.0)},
    {real(-1.0 / 12.0), real(7.0 / 12.0), real(7.0 / 12.0), real(-1.0 / 12.0)},
    {real(1.0 / 12.0), real(-5.0 / 12.0), real(13.0 / 12.0), real(1.0 / 4.0)},
    {real(-1.0 / 4.0), real(13.0 / 12.0), real(-23.0 / 12.0), real(25.0 / 12.0)}};
static real enoCoefficient(int order, int row, int j) {
    if (order == 2) return kEno2[row][j];
    if (order == 3) return kEno3[row][j];
    return kEno4[row][j];
}

// ENOCPU.cpp:41-185 (orders 2-4) with the divided differences of 205-264:
// (order-1) difference arrays of variable 0, each level from the one below
// over the reference's index ranges (kept between calls, as the reference
// keeps its arrays); at every level the shift grows when the difference at
// i-1 is smaller in magnitude than the one at i.
static void reconEno(int order, const Vol& in, int d, Vol& Lo, Vol& Ro, std::vector<real>& dd,
                     ivec3 start, ivec3 end) {
    const int nx = in.tx, ny = in.ty, nz = in.tz;
    const int dxv = d == 0, dyv = d == 1, dzv = d == 2;
    const size_t cells = in.cells();
    if (dd.size() != (order - 1) * cells) dd.assign((order - 1) * cells, 0.0);
    for (int level = 1; level < order; ++level) {
        const real* pin = level == 1 ? in.v(0) : dd.data() + (level - 2) * cells;
        real* pout = dd.data() + (level - 1) * cells;
        const int sx = (dxv == 0) * order + dxv * level + start.x;
        const int sy = (ny > 1) * (dyv == 0) * order + dyv * level + start.y;
        const int sz = (nz > 1) * (dzv == 0) * order + dzv * level + start.z;
        const int ex = nx - ((dxv == 0) * (order - 1) + dxv * (level - 1)) + end.x;
        const int ey = ny - ((ny > 1) * (dyv == 0) * (order - 1) + dyv * (level - 1)) + end.y;
        const int ez = nz - ((nz > 1) * (dzv == 0) * (order - 1) + dzv * (level - 1)) + end.z;
        for (int z = sz; z < ez; z++)
            for (int y = sy; y < ey; y++)
                for (int x = sx; x < ex; x++) {
                    const size_t iR = in.idx(x, y, z);
                    const size_t iL = in.idx(x - dxv, y - dyv, z - dzv);
                    pout[iL] = pin[iR] - pin[iL];
                }
    }
    const int sx = (order - dxv) + start.x;
    const int sy = (ny > 1) * (order - dyv) + start.y;
    const int sz = (nz > 1) * (order - dzv) + start.z;
    const int ex = nx - (order - dxv) + end.x;
    const int ey = ny - (ny > 1) * (order - dyv) + end.y;
    const int ez = nz - (nz > 1) * (order - dzv) + end.z;
    for (int var = 0; var < in.nvar; ++var) {
        const real* pin = in.v(var);
        real* pl = Lo.v(var);
        real* pr = Ro.v(var);
        for (int z = sz; z < ez; z++) {
#pragma omp parallel for
            for (int y = sy; y < ey; y++) {
                for (int x = sx; x < ex; x++) {
                    const size_t iR = in.idx(x, y, z);
                    const size_t iL = in.idx(x - dxv, y - dyv, z - dzv);
                    int shift = 0;
                    for (int level = 0; level < order - 1; ++level) {
                        const real* ddl = dd.data() + level * cells;
                        if (std::fabs(ddl[iL]) < std::fabs(ddl[iR])) shift++;
                    }
                    real lv = 0.0, rv = 0.0;
                    for (int j = 0; j < order; j++) {
                        const size_t index = in.idx(x - (shift - j) * dxv, y - (shift - j) * dyv,
                                                    z - (shift - j) * dzv);
                        const real value = pin[index];
                        lv += enoCoefficient(order, shift, j) * value;
                        rv += enoCoefficient(order, shift + 1, j) * value;
                    }
                    pl[iR] = lv;
                    pr[iR] = rv;
                }
            }
        }
    }
}

// WENOCPU<3> (WENOCPU.cpp:32-186) with WENOCoefficients<3> (beta:
// WENOCoefficients.hpp:51-63; epsilon 1e-8 and linear weights 3/10, 3/5,
// 1/10: WENOCoefficients.cpp:26-38): every cell with two neighbours each side
// in direction d (all rows of the other axes), indicator variable 0.
static void reconWeno3(const Vol& in, int d, Vol& Lo, Vol& Ro) {
    const int order = 3;
    const int nx = in.tx, ny = in.ty, nz = in.tz;
    const int dxv = d == 0, dyv = d == 1, dzv = d == 2;
    const real epsilon = 1e-8;
    const real lin[3] = {3.0 / 10.0, 3.0 / 5.0, 1.0 / 10.0};
    const real* w = in.v(0);
    auto sq = [](real x) { return std::pow(x, 2); };
    for (int z = dzv * (order - 1); z < nz - dzv * (order - 1); z++)
        for (int y = dyv * (order - 1); y < ny - dyv * (order - 1); y++)
            for (int x = dxv * (order - 1); x < nx - dxv * (order - 1); x++) {
                const size_t out = in.idx(x, y, z);
                real V[5];
                for (int i = -order + 1; i < order; i++)
                    V[i + order - 1] = w[in.idx(x + i * dxv, y + i * dyv, z + i * dzv)];
                real beta[3];
                beta[0] = 13.0 / 12.0 * sq(V[2] - 2 * V[3] + V[4]) +
                          1 / 4.0 * sq(3 * V[2] - 4 * V[3] + V[4]);
                beta[1] = 13.0 / 12.0 * sq(V[1] - 2 * V[2] + V[3]) + 1 / 4.0 * sq(V[1] - V[3]);
                beta[2] = 13.0 / 12.0 * sq(V[0] - 2 * V[1] + V[2]) +
                          1 / 4.0 * sq(V[0] - 4 * V[1] + 3 * V[2]);
                real aR[3], aL[3], sumR = 0.0, sumL = 0.0;
                for (int i = order - 1; i >= 0; --i) {  // computeAlpha recursion: i = 2, 1, 0
                    aR[i] = lin[i] / sq(beta[i] + epsilon);
                    sumR += aR[i];
                    aL[i] = lin[order - 1 - i] / sq(beta[i] + epsilon);
                    sumL += aL[i];
                }
                for (int var = 0; var < in.nvar; ++var) {
                    const real* pin = in.v(var);
                    real lw = 0.0, rw = 0.0;
                    for (int shift = 0; shift < order; shift++) {
                        real lv = 0.0, rv = 0.0;
                        for (int j = 0; j < order; j++) {
                            const real value = pin[in.idx(x - (shift - j) * dxv,
                                                          y - (shift - j) * dyv,
                                                          z - (shift - j) * dzv)];
                            lv += enoCoefficient(order, shift, j) * value;
                            rv += enoCoefficient(order, shift + 1, j) * value;
                        }
                        lw += lv * (aL[shift] / sumL);
                        rw += rv * (aR[shift] / sumR);
                    }
                    Lo.v(var)[out] = lw;
                    Ro.v(var)[out] = rw;
                }
            }
}

// ---------------------------------------------------------------- numerical flux

// NumericalFluxCPU.cpp:121-146
static void makeZero(Vol& out, ivec3 start, ivec3 end) {
    for (int z = out.ngz + start.z; z < out.tz - out.ngz + end.z; ++z)
#pragma omp parallel for
        for (int y = out.ngy + start.y; y < out.ty - out.ngy + end.y; ++y)
            for (int x = out.ngx + start.x; x < out.tx - out.ngx + end.x; ++x)
                for (int var = 0; var < out.nvar; ++var) out.v(var)[out.idx(x, y, z)] = 0;
}

// NumericalFluxCPU.cpp:32-118
template <class Eq, class Flux, int d>
static void netFlux(const Eq& eq, const Vol& left, const Vol& right, Vol& out, Vol& temp,
                    real& waveSpeed, ivec3 start, ivec3 end) {
    constexpr int nv = Eq::nvar;
    const int xD = d == 0, yD = d == 1, zD = d == 2;
    const int nx = out.tx, ny = out.ty, nz = out.tz;
    const int ngx = out.ngx, ngy = out.ngy, ngz = out.ngz;
    real wsc = 0;
    for (int z = ngz - zD + start.z; z < nz - ngz + end.z; ++z) {
#pragma omp parallel for reduction(max : wsc)
        for (int y = ngy - yD + start.y; y < ny - ngy + end.y; ++y) {
            for (int x = ngx - xD + start.x; x < nx - ngx + end.x; ++x) {
                const size_t i0 = out.idx(x, y, z);
                const size_t i1 = out.idx(x + xD, y + yD, z + zD);
                real qa[nv], qb[nv];
                for (int v = 0; v < nv; ++v) {
                    qa[v] = right.v(v)[i0];
                    qb[v] = left.v(v)[i1];
                }
                const typename Eq::All A = eq.make(qa);
                const typename Eq::All B = eq.make(qb);
                real F[nv];
                const real wl = Flux::template compute<d>(eq, A, B, F);
                for (int v = 0; v < nv; ++v) temp.v(v)[i0] = (-1.0) * F[v];
                wsc = std::max(wsc, wl);
            }
        }
    }
    waveSpeed = wsc;
    for (int z = ngz - zD + start.z; z < nz - ngz + end.z - zD; ++z) {
#pragma omp parallel for
        for (int y = ngy - yD + start.y; y < ny - ngy + end.y - yD; ++y) {
            for (int x = ngx - xD + start.x; x < nx - ngx + end.x - xD; ++x) {
                const size_t iR = out.idx(x + xD, y + yD, z + zD);
                const size_t iM = out.idx(x, y, z);
                for (int v = 0; v < nv; ++v) {
                    const real a = temp.v(v)[iR];
                    const real b = (-1.0) * temp.v(v)[iM];
                    out.v(v)[iR] += a + b;
                }
            }
        }
    }
}

struct Setup {
    int equation;    // 0 burgers, 1..3 euler<nsd>
    int flux;        // 0 hll, 1 hll3, 2 godunov, 3 central, 4 rusanov, 5 roe
    int recon;       // RECON_*
    int integrator;  // 0 forward euler, 1 ssp-rk2, 2 ssp-rk3, 3 rk4
    int boundary;    // 0 neumann, 1 periodic
    int nx, ny, nz;
    double origin[3], top[3];
    double gamma, cfl, end_time;
    int num_adjusters;
    double adj_interval[2];
    long adj_saved0[2];
    long max_steps;
};

// Abstract evaluation of one ConservedSystem::operator() (RHS + wave speeds).
struct SystemBase {
    virtual ~SystemBase() {}
    virtual void rhs(const Vol& U, real ws[3], Vol& out) = 0;
    virtual bool obeys(const Vol& U) = 0;
    // NumericalFlux::computeFlux over one region
    virtual void fluxRegion(const Vol& U, real ws[3], Vol& out, ivec3 s, ivec3 e) = 0;
    // Reconstruction::performReconstruction into L, R (prior contents kept
    // outside the reconstruction's range)
    virtual void reconstructInto(const Vol& U, int d, ivec3 s, ivec3 e, Vol& L, Vol& R) = 0;
};

template <class Eq, class Flux>
struct System : SystemBase {
    Eq eq;
    int recon;
    Vol left, right, temp;
    std::vector<real> dd;
    System(const Eq& e, int recon_, const Vol& shape)
        : eq(e), recon(recon_), left(shape.nvar, shape.nx, shape.ny, shape.nz, shape.ngx),
          right(shape.nvar, shape.nx, shape.ny, shape.nz, shape.ngx),
          temp(shape.nvar, shape.nx, shape.ny, shape.nz, shape.ngx) {
        std::fill(left.buf.begin(), left.buf.end(), 0.0);
        std::fill(right.buf.begin(), right.buf.end(), 0.0);
    }

    void reconstruct(const Vol& U, int d, ivec3 s, ivec3 e) {
        if (recon == RECON_NONE)
            reconNone(U, left, right);
        else if (recon == RECON_WENO2)
            reconWeno2(U, d, left, right, s, e);
        else if (recon == RECON_MC || recon == RECON_MM)
            reconLimiter(U, d, left, right, s, e, recon == RECON_MC);
        else if (recon == RECON_WENOF2)
            reconWenoF2(U, d, left, right, s, e, eq.gamma);
        else if (recon == RECON_WENO3)
            reconWeno3(U, d, left, right);
        else
            reconEno(recon == RECON_ENO2 ? 2 : (recon == RECON_ENO3 ? 3 : 4), U, d, left, right,
                     dd, s, e);
    }

    // NumericalFluxCPU::computeFlux (NumericalFluxCPU.cpp:170-221)
    void computeFlux(const Vol& U, real ws[3], Vol& out, ivec3 s, ivec3 e) {
        makeZero(out, s, e);
        const int dim = U.dims();
        reconstruct(U, 0, s, e);
        netFlux<Eq, Flux, 0>(eq, left, right, out, temp, ws[0], s, e);
        if constexpr (Eq::nsd > 1) {
            if (dim > 1) {
                reconstruct(U, 1, s, e);
                netFlux<Eq, Flux, 1>(eq, left, right, out, temp, ws[1], s, e);
            }
        }
        if constexpr (Eq::nsd > 2) {
            if (dim > 2) {
                reconstruct(U, 2, s, e);
                netFlux<Eq, Flux, 2>(eq, left, right, out, temp, ws[2], s, e);
            }
        }
    }

    void fluxRegion(const Vol& U, real ws[3], Vol& out, ivec3 s, ivec3 e) override {
        computeFlux(U, ws, out, s, e);
    }

    void reconstructInto(const Vol& U, int d, ivec3 s, ivec3 e, Vol& L, Vol& R) override {
        left.buf = L.buf;
        right.buf = R.buf;
        reconstruct(U, d, s, e);
        L.buf = left.buf;
        R.buf = right.buf;
    }

    // ConservedSystem::operator() (ConservedSystem.cpp:31-103): inner region
    // first, then the 2*d side regions; wave speeds are the max over regions.
    void rhs(const Vol& U, real ws[3], Vol& out) override {
        const ivec3 g = out.ghosts();
        const ivec3 size = out.total();
        const int dim = out.dims();
        ivec3 negg = iv(-g.x, -g.y, -g.z);
        computeFlux(U, ws, out, g, negg);
        real side[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            ivec3 sb, eb;
            eb[d] = -size[d] + 3 * g[d];
            real wb[3] = {0, 0, 0};
            computeFlux(U, wb, out, sb, eb);
            for (int k = 0; k < 3; ++k) side[k] = std::max(side[k], wb[k]);
            ivec3 st, et;
            st[d] = size[d] - 3 * g[d];
            real wt[3] = {0, 0, 0};
            computeFlux(U, wt, out, st, et);
            for (int k = 0; k < 3; ++k) side[k] = std::max(side[k], wt[k]);
        }
        for (int k = 0; k < 3; ++k) ws[k] = std::max(ws[k], side[k]);
    }

    bool obeys(const Vol& U) override {
        bool ok = true;
        const size_t n = U.cells();
        for (size_t i = 0; i < n; ++i) {
            real q[Eq::nvar];
            for (int v = 0; v < Eq::nvar; ++v) q[v] = U.v(v)[i];
            if (!eq.obeys(q)) ok = false;
        }
        return ok;
    }
};

// BoundaryCPU.cpp:30-104 with Neumann.hpp / Periodic.hpp.
static void applyBoundary(Vol& vol, int periodic, int ng) {
    const int dims = vol.dims();
    const int nx = vol.tx, ny = vol.ty, nz = vol.tz;
    for (int var = 0; var < vol.nvar; var++) {
        real* q = vol.v(var);
        for (int d = 0; d < dims; d++) {
            for (int i = 0; i < 2; i++) {
                const bool zD = d == 2, yD = d == 1, xD = d == 0;
                const int zs = zD ? (i == 0 ? ng : nz - ng - 1) : 0;
                const int ze = zD ? zs + 1 : nz;
                const int ys = yD ? (i == 0 ? ng : ny - ng - 1) : 0;
                const int ye = yD ? ys + 1 : ny;
                const int xs = xD ? (i == 0 ? ng : nx - ng - 1) : 0;
                const int xe = xD ? xs + 1 : nx;
                const int sign = i == 1 ? -1 : 1;
                for (int z = zs; z < ze; z++)
                    for (int y = ys; y < ye; y++)
                        for (int x = xs; x < xe; x++)
                            for (int gc = 1; gc <= ng; gc++) {
                                const size_t dst =
                                    vol.idx(x - sign * gc * xD, y - sign * gc * yD, z - sign * gc * zD);
                                size_t src;
                                if (!periodic) {
                                    src = vol.idx(x + sign * (gc - 1) * xD, y + sign * (gc - 1) * yD,
                                                  z + sign * (gc - 1) * zD);
                                } else {
                                    const int Nx = nx - 2 * ng, Ny = ny - 2 * ng, Nz = nz - 2 * ng;
                                    src = vol.idx(x + sign * (-gc + Nx) * xD, y + sign * (-gc + Ny) * yD,
                                                  z + sign * (-gc + Nz) * zD);
                                }
                                q[dst] = q[src];
                            }
            }
        }
    }
}

static int systemGhosts(const Setup& s) {
    // NumericalFluxCPU.cpp:226-231: max(stencil - 1, reconstruction ghosts)
    return std::max(1, reconGhosts(s.recon));
}

static int nvarOf(int equation) { return equation == 0 ? 1 : equation + 2; }

static SystemBase* makeSystem(const Setup& s, const Vol& shape) {
    if (s.equation == 0) {
        Burgers b;
        b.gamma = s.gamma;
        if (s.flux == 2) return new System<Burgers, Godunov>(b, s.recon, shape);
        if (s.flux == 3) return new System<Burgers, Central>(b, s.recon, shape);
        if (s.flux == 4) return new System<Burgers, Rusanov>(b, s.recon, shape);
        if (s.flux == 5) return new System<Burgers, Roe>(b, s.recon, shape);
        return nullptr;
    }
    if (s.flux != 0 && s.flux != 1 && s.flux != 3) return nullptr;
#define ORC_EULER(N)                                                  \
    if (s.equation == N) {                                            \
        Euler<N> e;                                                   \
        e.gamma = s.gamma;                                            \
        if (s.flux == 0) return new System<Euler<N>, HLL>(e, s.recon, shape);  \
        if (s.flux == 3) return new System<Euler<N>, Central>(e, s.recon, shape);  \
        return new System<Euler<N>, HLL3>(e, s.recon, shape);         \
    }
    ORC_EULER(1)
    ORC_EULER(2)
    ORC_EULER(3)
#undef ORC_EULER
    return nullptr;
}

static void cellLengths(const Setup& s, real L[3]) {
    // Grid.cpp:24-41
    const int n[3] = {s.nx, s.ny, s.nz};
    for (int k = 0; k < 3; ++k) L[k] = (s.top[k] - s.origin[k]) / real(n[k]);
    if (s.ny == 1) L[1] = 1;
    if (s.nz == 1) L[2] = 1;
}

struct Adjuster {
    real interval;
    long saved;
};

// Integrator.cpp:20-66 (no MPI wave-speed adjuster on one rank).
static real computeTimestep(const real ws[3], const real dx[3], real cfl, real t,
                            const std::vector<Adjuster>& adj) {
    real total = 0;
    for (int d = 0; d < 3; ++d) {
        if (dx[d] == 0) continue;
        total += ws[d] / dx[d];
    }
    real dt = cfl / total;
    for (const auto& a : adj) {
        // FixedIntervalWriter.cpp:45-53 / FixedIntervalStatistics.cpp:33-41
        if (a.saved > 0) {
            const real next = a.saved * a.interval + 0.0;
            dt = std::min(dt, next - t);
        }
    }
    return dt;
}

// Elementwise Memory ops over whole volumes incl. ghosts (HostMemory.cpp:96-201).
static void scaleVol(Vol& a, real s) {
    for (auto& x : a.buf) x *= s;
}
static void addVol(Vol& a, const Vol& b) {
    for (size_t i = 0; i < a.buf.size(); ++i) a.buf[i] += b.buf[i];
}

struct Sim {
    Setup s;
    int ng, nvar;
    real dx[3];
    std::vector<Vol> U;
    SystemBase* sys;
    std::vector<Adjuster> adj;
    real t = 0;
    long steps = 0;
    int substeps;

    explicit Sim(const Setup& setup) : s(setup) {
        ng = systemGhosts(s);
        nvar = nvarOf(s.equation);
        cellLengths(s, dx);
        substeps = s.integrator + 1;
        for (int i = 0; i <= substeps; ++i) U.emplace_back(nvar, s.nx, s.ny, s.nz, ng);
        sys = makeSystem(s, U[0]);
        for (int k = 0; k < s.num_adjusters; ++k) adj.push_back({s.adj_interval[k], s.adj_saved0[k]});
    }
    ~Sim() { delete sys; }

    // ForwardEuler.cpp:32-51, RungeKutta2.cpp:37-67, RungeKutta3.cpp:49-86,
    // RungeKutta4.cpp:50-96
    real substep(int k, real dt) {
        real ws[3] = {0, 0, 0};
        Vol& out = U[k + 1];
        sys->rhs(U[k], ws, out);
        if (k == 0) dt = computeTimestep(ws, dx, s.cfl, t, adj);
        const real c = dt / dx[0];
        if (s.integrator == 3) {
            if (k < 3) scaleVol(out, c);
            if (k < 2) scaleVol(out, 0.5);
            if (k < 3) addVol(out, U[0]);
            if (k == 3) {
                // Volume::addLinearCombination -> HostMemory.cpp:266-279
                const real a1 = (1.0 / 6.0) * c, a2 = 1.0 / 3.0, a3 = 2.0 / 3.0, a4 = 1.0 / 3.0,
                           a5 = -1.0 / 3.0;
                for (size_t i = 0; i < out.buf.size(); ++i)
                    out.buf[i] = a1 * out.buf[i] + a2 * U[1].buf[i] + a3 * U[2].buf[i] +
                                 a4 * U[3].buf[i] + a5 * U[0].buf[i];
            }
            return dt;
        }
        scaleVol(out, c);
        addVol(out, U[k]);
        if (s.integrator == 1 && k == 1) {
            addVol(out, U[0]);
            scaleVol(out, 0.5);
        } else if (s.integrator == 2 && k == 1) {
            scaleVol(out, 1. / 3.);
            addVol(out, U[0]);
            scaleVol(out, 3. / 4.);
        } else if (s.integrator == 2 && k == 2) {
            scaleVol(out, 2);
            addVol(out, U[0]);
            scaleVol(out, 1. / 3.);
        }
        return dt;
    }

    // Simulator::incrementSolution + checkConstraints + callWriters
    // (Simulator.cpp:81-85,180-214; FixedIntervalWriter.cpp:33-43)
    int step(real* dt_out) {
        real dt = 0;
        for (int k = 0; k < substeps; ++k) {
            dt = substep(k, dt);
            applyBoundary(U[k + 1], s.boundary, ng);
        }
        std::swap(U[0], U[substeps]);
        t += dt;
        steps++;
        if (dt_out) *dt_out = dt;
        if (!sys->obeys(U[0])) return 2;
        for (auto& a : adj)
            if (t >= a.saved * a.interval + 0.0) a.saved++;
        return 0;
    }
};

static bool validSetup(const Setup& s) {
    if (s.equation < 0 || s.equation > 3) return false;
    if (s.recon < 0 || s.recon > 8) return false;
    if (s.recon == RECON_WENOF2 && s.equation == 0) return false;  // Euler only
    if (s.integrator < 0 || s.integrator > 3) return false;
    if (s.num_adjusters < 0 || s.num_adjusters > 2) return false;
    if (s.nx < 1 || s.ny < 1 || s.nz < 1) return false;
    return true;
}

}  // namespace orc

// ---------------------------------------------------------------- C ABI (ctypes)

extern "C" {

typedef orc::Setup oracle_setup;

int oracle_ghost_cells(const oracle_setup* s) { return orc::systemGhosts(*s); }
int oracle_nvar(int equation) { return orc::nvarOf(equation); }

// Runs Simulator steps from u0 (a full volume incl. ghosts, variable-major)
// until t >= end_time or max_steps steps. Boundary conditions are applied to
// u0 first (Simulator::setInitialValue, Simulator.cpp:160).
// Returns 0 ok, 1 bad setup, 2 constraint violation (the reference THROWs).
int oracle_run(const oracle_setup* setup, const double* u0, double* uout, double* t_out,
               long* steps_out, double* dt_hist, long dt_cap) {
    if (!orc::validSetup(*setup)) return 1;
    orc::Sim sim(*setup);
    if (!sim.sys) return 1;
    std::memcpy(sim.U[0].buf.data(), u0, sim.U[0].buf.size() * sizeof(double));
    orc::applyBoundary(sim.U[0], setup->boundary, sim.ng);
    int rc = 0;
    while (sim.t < setup->end_time && (setup->max_steps < 0 || sim.steps < setup->max_steps)) {
        double dt = 0;
        rc = sim.step(&dt);
        if (dt_hist && sim.steps - 1 < dt_cap) dt_hist[sim.steps - 1] = dt;
        if (rc) break;
    }
    std::memcpy(uout, sim.U[0].buf.data(), sim.U[0].buf.size() * sizeof(double));
    if (t_out) *t_out = sim.t;
    if (steps_out) *steps_out = sim.steps;
    return rc;
}

// One ConservedSystem::operator() evaluation. `out` must hold the prior
// contents of the output volume (only interior cells are written).
int oracle_system_rhs(const oracle_setup* setup, const double* u, double* out, double* ws3) {
    if (!orc::validSetup(*setup)) return 1;
    const int ng = orc::systemGhosts(*setup);
    const int nv = orc::nvarOf(setup->equation);
    orc::Vol U(nv, setup->nx, setup->ny, setup->nz, ng);
    orc::Vol O(nv, setup->nx, setup->ny, setup->nz, ng);
    std::memcpy(U.buf.data(), u, U.buf.size() * sizeof(double));
    std::memcpy(O.buf.data(), out, O.buf.size() * sizeof(double));
    orc::SystemBase* sys = orc::makeSystem(*setup, U);
    if (!sys) return 1;
    double ws[3] = {0, 0, 0};
    sys->rhs(U, ws, O);
    delete sys;
    std::memcpy(out, O.buf.data(), O.buf.size() * sizeof(double));
    for (int k = 0; k < 3; ++k) ws3[k] = ws[k];
    return 0;
}

// One NumericalFlux::computeFlux(U, ws, computeWaveSpeed, out, start, end)
// call (NumericalFluxCPU.cpp:170-221). `out` holds the prior contents of the
// output volume; ws3 gets the wave speeds of the active directions (0 for
// the others).
int oracle_compute_flux(const oracle_setup* setup, const double* u, const int* start,
                        const int* end, double* out, double* ws3) {
    if (!orc::validSetup(*setup)) return 1;
    const int ng = orc::systemGhosts(*setup);
    const int nv = orc::nvarOf(setup->equation);
    orc::Vol U(nv, setup->nx, setup->ny, setup->nz, ng);
    orc::Vol O(nv, setup->nx, setup->ny, setup->nz, ng);
    std::memcpy(U.buf.data(), u, U.buf.size() * sizeof(double));
    std::memcpy(O.buf.data(), out, O.buf.size() * sizeof(double));
    orc::SystemBase* sys = orc::makeSystem(*setup, U);
    if (!sys) return 1;
    double ws[3] = {0, 0, 0};
    sys->fluxRegion(U, ws, O, orc::iv(start[0], start[1], start[2]), orc::iv(end[0], end[1], end[2]));
    delete sys;
    std::memcpy(out, O.buf.data(), O.buf.size() * sizeof(double));
    for (int k = 0; k < 3; ++k) ws3[k] = ws[k];
    return 0;
}

// One Reconstruction::performReconstruction(U, direction, 0, left, right,
// start, end) call of the setup's reconstruction on a fresh reconstruction
// object; left/right hold their prior contents.
int oracle_reconstruct(const oracle_setup* setup, const double* u, int direction, const int* start,
                       const int* end, double* left, double* right) {
    if (!orc::validSetup(*setup)) return 1;
    const int ng = orc::systemGhosts(*setup);
    const int nv = orc::nvarOf(setup->equation);
    orc::Vol U(nv, setup->nx, setup->ny, setup->nz, ng);
    orc::Vol L(nv, setup->nx, setup->ny, setup->nz, ng);
    orc::Vol R(nv, setup->nx, setup->ny, setup->nz, ng);
    std::memcpy(U.buf.data(), u, U.buf.size() * sizeof(double));
    std::memcpy(L.buf.data(), left, L.buf.size() * sizeof(double));
    std::memcpy(R.buf.data(), right, R.buf.size() * sizeof(double));
    orc::SystemBase* sys = orc::makeSystem(*setup, U);
    if (!sys) return 1;
    sys->reconstructInto(U, direction, orc::iv(start[0], start[1], start[2]),
                         orc::iv(end[0], end[1], end[2]), L, R);
    delete sys;
    std::memcpy(left, L.buf.data(), L.buf.size() * sizeof(double));
    std::memcpy(right, R.buf.data(), R.buf.size() * sizeof(double));
    return 0;
}

int oracle_apply_boundary(const oracle_setup* setup, double* u) {
    if (!orc::validSetup(*setup)) return 1;
    const int ng = orc::systemGhosts(*setup);
    orc::Vol U(orc::nvarOf(setup->equation), setup->nx, setup->ny, setup->nz, ng);
    std::memcpy(U.buf.data(), u, U.buf.size() * sizeof(double));
    orc::applyBoundary(U, setup->boundary, ng);
    std::memcpy(u, U.buf.data(), U.buf.size() * sizeof(double));
    return 0;
}

int oracle_obeys_constraints(const oracle_setup* setup, const double* u) {
    if (!orc::validSetup(*setup)) return -1;
    const int ng = orc::systemGhosts(*setup);
    orc::Vol U(orc::nvarOf(setup->equation), setup->nx, setup->ny, setup->nz, ng);
    std::memcpy(U.buf.data(), u, U.buf.size() * sizeof(double));
    orc::SystemBase* sys = orc::makeSystem(*setup, U);
    if (!sys) return -1;
    const int ok = sys->obeys(U) ? 1 : 0;
    delete sys;
    return ok;
}

// Single-face flux for known-answer tests: state vectors are conserved
// (rho, m..., E) for Euler or (u) for Burgers; returns the wave speed.
double oracle_face_flux(int equation, int flux, int direction, double gamma, const double* qL,
                        const double* qR, double* F) {
#define ORC_FACE(EQ, FL, D) \
    return FL::compute<D>(EQ, EQ.make(qL), EQ.make(qR), F);
    if (equation == 0) {
        orc::Burgers b;
        b.gamma = gamma;
        if (flux == 3) ORC_FACE(b, orc::Central, 0)
        if (flux == 4) ORC_FACE(b, orc::Rusanov, 0)
        if (flux == 5) ORC_FACE(b, orc::Roe, 0)
        if (direction == 0) ORC_FACE(b, orc::Godunov, 0)
        if (direction == 1) ORC_FACE(b, orc::Godunov, 1)
        ORC_FACE(b, orc::Godunov, 2)
    }
#define ORC_EQ(N)                                                              \
    if (equation == N) {                                                       \
        orc::Euler<N> e;                                                       \
        e.gamma = gamma;                                                       \
        if (flux == 3) {                                                       \
            if (direction == 0) ORC_FACE(e, orc::Central, 0)                    \
            if (N > 1 && direction == 1) ORC_FACE(e, orc::Central, (N > 1 ? 1 : 0)) \
            if (N > 2 && direction == 2) ORC_FACE(e, orc::Central, (N > 2 ? 2 : 0)) \
        }                                                                      \
        if (flux == 0) {                                                       \
            if (direction == 0) ORC_FACE(e, orc::HLL, 0)                        \
            if (N > 1 && direction == 1) ORC_FACE(e, orc::HLL, (N > 1 ? 1 : 0)) \
            if (N > 2 && direction == 2) ORC_FACE(e, orc::HLL, (N > 2 ? 2 : 0)) \
        } else {                                                               \
            if (direction == 0) ORC_FACE(e, orc::HLL3, 0)                       \
            if (N > 1 && direction == 1) ORC_FACE(e, orc::HLL3, (N > 1 ? 1 : 0)) \
            if (N > 2 && direction == 2) ORC_FACE(e, orc::HLL3, (N > 2 ? 2 : 0)) \
        }                                                                      \
    }
    ORC_EQ(1)
    ORC_EQ(2)
    ORC_EQ(3)
#undef ORC_EQ
#undef ORC_FACE
    return -1.0;
}

// Euler point flux in direction d for a conserved state.
void oracle_point_flux(int nsd, int direction, double gamma, const double* q, double* F) {
    if (nsd == 1) {
        orc::Euler<1> e;
        e.gamma = gamma;
        e.pointFlux<0>(e.make(q), F);
    } else if (nsd == 2) {
        orc::Euler<2> e;
        e.gamma = gamma;
        if (direction == 0) e.pointFlux<0>(e.make(q), F);
        else e.pointFlux<1>(e.make(q), F);
    } else {
        orc::Euler<3> e;
        e.gamma = gamma;
        if (direction == 0) e.pointFlux<0>(e.make(q), F);
        else if (direction == 1) e.pointFlux<1>(e.make(q), F);
        else e.pointFlux<2>(e.make(q), F);
    }
}

// Reconstruction of a 1D line (nvar variables, n cells incl. ghosts):
// left/right face values at every cell where the reference computes them.
int oracle_reconstruct_1d(int recon, int nvar, int n, const double* q, double* left,
                          double* right, double gamma) {
    const int ng = orc::reconGhosts(recon);
    if (n <= 2 * ng) return 1;
    orc::Vol U(nvar, n - 2 * ng, 1, 1, ng), L(nvar, n - 2 * ng, 1, 1, ng),
        R(nvar, n - 2 * ng, 1, 1, ng);
    std::memcpy(U.buf.data(), q, U.buf.size() * sizeof(double));
    std::memcpy(L.buf.data(), left, L.buf.size() * sizeof(double));
    std::memcpy(R.buf.data(), right, R.buf.size() * sizeof(double));
    std::vector<double> dd;
    if (recon == orc::RECON_NONE) orc::reconNone(U, L, R);
    else if (recon == orc::RECON_WENO2) orc::reconWeno2(U, 0, L, R, orc::ivec3(), orc::ivec3());
    else if (recon == orc::RECON_MC || recon == orc::RECON_MM)
        orc::reconLimiter(U, 0, L, R, orc::ivec3(), orc::ivec3(), recon == orc::RECON_MC);
    else if (recon == orc::RECON_WENOF2) {
        if (nvar < 3) return 1;
        orc::reconWenoF2(U, 0, L, R, orc::ivec3(), orc::ivec3(), gamma);
    } else if (recon == orc::RECON_WENO3) {
        orc::reconWeno3(U, 0, L, R);
    } else
        orc::reconEno(recon == orc::RECON_ENO2 ? 2 : (recon == orc::RECON_ENO3 ? 3 : 4), U, 0, L, R,
                      dd, orc::ivec3(), orc::ivec3());
    std::memcpy(left, L.buf.data(), L.buf.size() * sizeof(double));
    std::memcpy(right, R.buf.data(), R.buf.size() * sizeof(double));
    return 0;
}

// Integrator::computeTimestep with FixedInterval adjusters.
double oracle_compute_timestep(const double* ws3, const double* dx3, double cfl, double t,
                               int nadj, const double* interval, const long* saved) {
    std::vector<orc::Adjuster> adj;
    for (int k = 0; k < nadj; ++k) adj.push_back({interval[k], saved[k]});
    return orc::computeTimestep(ws3, dx3, cfl, t, adj);
}

// ------------------------------------------------ sampler (STLMersenne + Uniform)

// Draws nsamples x dim values exactly as the reference Runner asks for them:
// sample-major, component-minor, from one default-constructed mt19937_64
// shared by the process (STLMersenne.cpp:25-85), scaled x*(upper-lower)+lower
// (Uniform.cpp:27-34). sample_start follows Setup.cpp:338-347.
void oracle_stl_mersenne_uniform(long sample_start, long nsamples, long dim, double lower,
                                 double upper, double* out) {
    std::mt19937_64 engine;
    long count = 0;
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    for (long i = 0; i < nsamples; ++i) {
        const long sample = sample_start + i;
        for (long c = 0; c < dim; ++c) {
            if (count / dim < sample) {
                const long add = sample * dim - count;
                engine.discard(add);
                count += add;
            }
            if (count % dim < c) {
                const long add = c - count % dim;
                engine.discard(add);
                count += add;
            }
            count++;
            const double x = dist(engine);
            out[i * dim + c] = x * (upper - lower) + lower;
        }
    }
}

unsigned long long oracle_mt19937_64_nth(long n) {
    std::mt19937_64 e;
    e.discard(n - 1);
    return e();
}

// ------------------------------------------------ statistics

// meanvar over M full volumes (variable-major, incl. ghosts) stored
// contiguously: samples[s * len + i]. MeanVariance.cpp:30-67 and
// StatisticsHelper.cpp:65-74 (sum in sample order, /= M, var = E[u^2]-E[u]^2).
// The reference's addPower(u, 2) and subtractPower(mean, 2.) call libm's pow
// with a run-time exponent; here pow(., 2.0) has a constant exponent, which
// the compiler evaluates as the correctly rounded square (libm's pow(x, 2.0)
// is within 1 ulp of it; the tests bound the difference).
void oracle_meanvar(long nsamples, long len, const double* samples, double* mean, double* var) {
    for (long i = 0; i < len; ++i) {
        mean[i] = 0;
        var[i] = 0;
    }
    for (long s = 0; s < nsamples; ++s) {
        const double* u = samples + s * len;
        for (long i = 0; i < len; ++i) mean[i] += u[i];
        for (long i = 0; i < len; ++i) var[i] += std::pow(u[i], 2.0);
    }
    const double M = double(nsamples);
    for (long i = 0; i < len; ++i) {
        mean[i] /= M;
        var[i] /= M;
    }
    for (long i = 0; i < len; ++i) var[i] -= std::pow(mean[i], 2.);
}

static inline double powP(double x, double p) {
    // dispatchComputeStructureCubeCPU: FastPower<k> for integer p <= 5
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * (x * x);
    if (p == 4.0) return x * (x * (x * x));
    if (p == 5.0) return x * (x * (x * (x * x)));
    return std::pow(x, p);
}

static inline double valueAt(const orc::Vol& V, const double* q, int x, int y, int z, int periodic) {
    const int N[3] = {V.nx, V.ny, V.nz};
    int P[3] = {x, y, z};
    for (int k = 0; k < 3; ++k) {
        if (periodic) {
            while (P[k] < 0) P[k] += N[k];
            P[k] = P[k] % N[k];
        } else {
            P[k] = std::max(0, std::min(P[k], N[k] - 1));
        }
    }
    return q[V.idx(P[0] + V.ngx, P[1] + V.ngy, P[2] + V.ngz)];
}

// std::pow(x, p) with a run-time exponent (HostMemory::addPower,
// HostMemory.cpp:283-290; HostMemory::getTotalVariation, HostMemory.cpp:
// 325-362). p = 1 gives x: libm's pow(x, 1) returns x, the exact result being
// representable. p = 2 gives the correctly rounded x*x: glibc's pow(x, 2.0)
// is within 1 ulp of it (glibc 2.35 differs in the last bit for about 0.08 %
// of inputs, and x86-64 glibc picks an FMA or a non-FMA pow at run time, so
// the reference's own last bit there depends on the machine). Other exponents
// call libm's pow, as the reference does.
static inline double statPow(double x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    return std::pow(x, p);
}

// OnePointMoment over M full volumes stored contiguously (samples[s * len +
// i]): out[i] = (sum over s in order of pow(u_s[i], p)) / M.
// OnePointMoment.cpp:30-45 (addPower per sample) and StatisticsHelper.cpp:
// 65-74 (division by the number of samples).
void oracle_onepoint_moment(long nsamples, long len, const double* samples, double p,
                            double* out) {
    for (long i = 0; i < len; ++i) out[i] = 0;
    for (long s = 0; s < nsamples; ++s) {
        const double* u = samples + s * len;
        for (long i = 0; i < len; ++i) out[i] += statPow(u[i], p);
    }
    for (long i = 0; i < len; ++i) out[i] /= double(nsamples);
}

// Total variation of one variable of one volume of total extents (tx, ty,
// tz) with ghost layers (gx, gy, gz): HostMemory::getTotalVariation
// (HostMemory.cpp:325-362) with start = ghosts, end = size - ghosts as
// BoundedVariation.cpp:38-53 passes them. Returns -1 for 3D (the reference
// throws).
double oracle_total_variation(int tx, int ty, int tz, int gx, int gy, int gz, const double* u,
                              double p) {
    if (tz > 1) return -1;
    const int startX = gx + 1;
    const int startY = gy + (ty > 1 ? 1 : 0);
    const int endX = tx - gx, endY = ty - gy, endZ = tz - gz;
    double bv = 0;
    for (int z = 0; z < endZ; z++)
        for (int y = startY; y < endY; y++)
            for (int x = startX; x < endX; x++) {
                const long idx = (long)z * tx * ty + (long)y * tx + x;
                const long left = (long)z * tx * ty + (long)y * tx + (x - 1);
                const int yBottom = ty > 1 ? y - 1 : 0;
                const long below = (long)z * tx * ty + (long)yBottom * tx + x;
                const double dy = u[idx] - u[below], dx = u[idx] - u[left];
                bv += statPow(std::sqrt(dy * dy + dx * dx), p);
            }
    return bv;
}

// Adds one sample's structure_cube contribution to out[var * H + h]
// (structure_common.hpp:37-150, h = 1..H-1; ValueAtBoundary.hpp:20-98).
void oracle_structure_cube(int nvar, int nx, int ny, int nz, int ng, const double* u, int H,
                           double p, int periodic, double* out) {
    orc::Vol V(nvar, nx, ny, nz, ng);
    std::memcpy(V.buf.data(), u, V.buf.size() * sizeof(double));
    const int dims = V.dims();
    for (int var = 0; var < nvar; ++var) {
        const double* q = V.v(var);
        double* o = out + (size_t)var * H;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (int h = 1; h < H; ++h) {
                        const double uc = q[V.idx(i + V.ngx, j + V.ngy, k + V.ngz)];
                        for (int d = 0; d < dims; d++)
                            for (int side = 0; side < 2; side++) {
                                const bool zD = d == 2, yD = d == 1, xD = d == 0;
                                const int zs = zD ? (side == 0 ? k - h : k + h) : (dims > 2 ? k - h : 0);
                                const int ze = zD ? zs + 1 : (dims > 2 ? k + h + 1 : 1);
                                int ys = yD ? (side == 0 ? j - h : j + h)
                                            : (dims == 2 ? j - h + 1 : (dims == 3 ? j - h + 1 : 0));
                                int ye = yD ? ys + 1 : (dims == 2 ? j + h : (dims == 3 ? j + h : 1));
                                int xs = xD ? (side == 0 ? i - h : i + h) : i - h;
                                int xe = xD ? xs + 1 : i + h + 1;
                                if (zD) {
                                    xs += 1;
                                    xe -= 1;
                                }
                                for (int z = zs; z < ze; z++)
                                    for (int y = ys; y < ye; y++)
                                        for (int x = xs; x < xe; x++) {
                                            const double uh = valueAt(V, q, x, y, z, periodic);
                                            o[h] += powP(std::fabs(uc - uh), p) / (nx * ny * nz);
                                        }
                            }
                    }
    }
}

// Adds one sample's structure_basic contribution (StructureBasic.cpp:67-120,
// h = 0..H-1, one direction).
void oracle_structure_basic(int nvar, int nx, int ny, int nz, int ng, const double* u, int H,
                            double p, int direction, int periodic, double* out) {
    orc::Vol V(nvar, nx, ny, nz, ng);
    std::memcpy(V.buf.data(), u, V.buf.size() * sizeof(double));
    const int ex = direction == 0, ey = direction == 1, ez = direction == 2;
    for (int var = 0; var < nvar; ++var) {
        const double* q = V.v(var);
        double* o = out + (size_t)var * H;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (int h = 0; h < H; ++h) {
                        const double uc = q[V.idx(i + V.ngx, j + V.ngy, k + V.ngz)];
                        const double uh = valueAt(V, q, i + h * ex, j + h * ey, k + h * ez, periodic);
                        o[h] += std::pow(std::fabs(uc - uh), p) / (nx * ny * nz);
                    }
    }
}

}  // extern "C"

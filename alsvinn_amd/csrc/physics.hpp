// Device-side physics for the hip platform: Euler<nsd> / Burgers equations,
// their face fluxes (HLL, HLL3, Godunov, central, Rusanov, Roe) and
// reconstructions (none, WENO2, WENOF2, ENO2-4, MC, MM).
//
// Every expression keeps the association order of the reference's CPU path so
// that FP64 results are bit-identical (the library is built with
// -ffp-contract=off; '/' and sqrt are IEEE correctly rounded on gfx950):
//   Euler::computeExtra / computePointFlux   alsfvm/include/alsfvm/equation/euler/Euler.hpp:172-211
//   HLL::computeFlux / computeHLLSpeeds      alsfvm/include/alsfvm/numflux/euler/HLL.hpp:51-115
//   HLL3::computeFlux / computeHLLSpeeds     alsfvm/include/alsfvm/numflux/euler/HLL3.hpp:52-148
//   burgers::Godunov::computeFlux            alsfvm/include/alsfvm/numflux/burgers/Godunov.hpp:40-60
//   WENO2::reconstruct                       alsfvm/include/alsfvm/reconstruction/WENO2.hpp:37-71
//   ENOCPU<2..4> shift + coefficients        alsfvm/src/reconstruction/ENOCPU.cpp:143-176,
//                                            alsfvm/src/reconstruction/ENOCoefficients.cpp:27-50
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

namespace alsvinn_hip {

// glibc's fmin/fmax on x86-64 (minsd/maxsd with NaN fall-through): for
// ordered operands min returns x only if x < y, max only if x > y.
template <class T>
__device__ __forceinline__ T ref_fmin(T x, T y) {
    if (x != x) return y;
    if (y != y) return x;
    return (x < y) ? x : y;
}
template <class T>
__device__ __forceinline__ T ref_fmax(T x, T y) {
    if (x != x) return y;
    if (y != y) return x;
    return (x > y) ? x : y;
}
// fmin/fmax by the hardware (v_min_f64 / v_max_f64: IEEE minNum/maxNum, the
// other operand when one is a quiet NaN). They agree with glibc's fmin/fmax
// except possibly for the pair (+0, -0), where glibc returns its second
// operand; used only where such a pair cannot occur or its sign cannot
// reach the result (see the call sites).
template <class T>
__device__ __forceinline__ T hw_fmin(T x, T y) {
    return __builtin_fmin(x, y);
}
template <class T>
__device__ __forceinline__ T hw_fmax(T x, T y) {
    return __builtin_fmax(x, y);
}
// std::max(a, b) = (a < b) ? b : a (keeps a when b is NaN).
template <class T>
__device__ __forceinline__ T std_max(T a, T b) {
    return (a < b) ? b : a;
}

// Division and square root. FP64 (the parity path): the correctly rounded
// IEEE results, i.e. exactly the reference's operations, computed by gfx950's
// own IEEE sequences without their range-scaling and special-case steps.
// The compiler's FP64 division is v_div_scale (x2), v_rcp_f64, two Newton
// steps on the reciprocal, the quotient, its residual, v_div_fmas and
// v_div_fixup. For operands that need no scaling (y and x/y normal, |x| >=
// 2^-969, exponents less than 768 apart) the scale steps return their
// inputs, v_div_fmas is a plain FMA and v_div_fixup returns the quotient, so
// the sequence below gives the same bits. A zero numerator gives the signed
// zero x*r, the IEEE -0/+0 result: dv() gives its result the sign of x*r
// (the sign of every quotient, zero or not); dvp(), for
// denominators that are positive in every valid state (densities, pressures,
// sums of positive weights), gets it from the residual step itself (the
// residual is formed as y*q - x and subtracted, which keeps the zero's sign
// when y > 0). The refined reciprocal depends on y only, so quotients with a
// common denominator share it (the compiler merges the identical chains).
// The square root is the compiler's sequence (v_rsq_f64 and Newton
// corrections) without the rescaling of inputs below 2^-767, exact for zero. Every operand of a valid state (positive density and pressure,
// finite values) is in these ranges; zero, infinite or NaN denominators and
// infinite square-root inputs, which arise only in states the constraint
// check rejects, give NaN here instead of the IEEE special values.
// FP32-state variant: the hardware reciprocal and square root (about 1 ulp
// each), traded for speed; the variant's error against the FP64 path is
// measured, not assumed.
__device__ __forceinline__ double rcp_refined(double y) {
    double r = __builtin_amdgcn_rcp(y);
    double t = __builtin_fma(-y, r, 1.0);
    r = __builtin_fma(r, t, r);
    t = __builtin_fma(-y, r, 1.0);
    r = __builtin_fma(r, t, r);
    return r;
}
__device__ __forceinline__ double dvp(double a, double b) {
    const double r = rcp_refined(b);
    const double q = a * r;
    const double e = __builtin_fma(b, q, -a);
    return __builtin_fma(-e, r, q);
}
__device__ __forceinline__ double dv(double a, double b) {
    const double r = rcp_refined(b);
    const double q = a * r;
    const double e = __builtin_fma(b, q, -a);
    const double x = __builtin_fma(-e, r, q);
    // sign(q) = sign(a) xor sign(b), the quotient's sign also when it is
    // zero; for a nonzero quotient x has it already
    return __builtin_copysign(x, q);
}
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float dvp(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double dsqrt(double x) {
    // x = +-0: the estimate 1/sqrt(x) is +-inf; 2^1023 in its place (the
    // same low word, 0) keeps every step below exact, and the sequence then
    // returns x itself, +-0, the IEEE square root of +-0
    const double y1 = __builtin_amdgcn_rsq(x);
    const double y0 = __hiloint2double(x == 0.0 ? 0x7fe00000 : __double2hiint(y1), __double2loint(y1));
    double g = x * y0;
    double h = y0 * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
// The same sequence without the zero pass-through, for the square roots of
// face densities: a zero density is also the divisor of the face's velocity,
// so its flux is not finite either way (zero denominators give NaN here, as
// above).
__device__ __forceinline__ double dsqrt_rho(double x) {
    const double y0 = __builtin_amdgcn_rsq(x);
    double g = x * y0;
    double h = y0 * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
__device__ __forceinline__ float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float dsqrt_rho(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }

// ------------------------------------------------------------------ equations

template <class T, int NSD>
struct EulerEq {
    static constexpr int nvar = NSD + 2;
    static constexpr int nsd = NSD;
    static constexpr bool is_euler = true;
    T gamma;
    struct All {
        T q[nvar];
        T u[NSD];
        T p;
    };
    __device__ __forceinline__ void make(const T* q, All& a) const {
#pragma unroll
        for (int i = 0; i < nvar; ++i) a.q[i] = q[i];
        T dot = q[1] * q[1];
        if (NSD > 1) dot = dot + q[2] * q[2];
        if (NSD > 2) dot = dot + q[3] * q[3];
        const T ie = q[nvar - 1] - dvp(T(0.5) * dot, q[0]);
#pragma unroll
        for (int k = 0; k < NSD; ++k) a.u[k] = dvp(q[1 + k], q[0]);
        a.p = (gamma - T(1)) * ie;
    }
    template <int d>
    __device__ __forceinline__ void pointFlux(const All& a, T* F) const {
        F[0] = a.q[1 + d];
#pragma unroll
        for (int k = 0; k < NSD; ++k) F[1 + k] = a.q[1 + k] * a.u[d];
        F[1 + d] += a.p;
        F[nvar - 1] = (a.q[nvar - 1] + a.p) * a.u[d];
    }
    // Euler::computeWaveSpeed (Euler.hpp:251-258): |u_d| + sqrt(gamma * p / rho)
    template <int d>
    __device__ __forceinline__ T waveSpeed(const All& a) const {
        return dabs(a.u[d]) + dsqrt(dvp(gamma * a.p, a.q[0]));
    }
    // CPUCellComputer::obeysConstraints + Euler::obeysConstraints:
    // every conserved and extra value finite, rho > 0, p > 0.
    __device__ __forceinline__ bool obeys(const T* q) const {
        All a;
        make(q, a);
        bool ok = true;
#pragma unroll
        for (int i = 0; i < nvar; ++i) ok = ok && __builtin_isfinite(q[i]);
#pragma unroll
        for (int k = 0; k < NSD; ++k) ok = ok && __builtin_isfinite(a.u[k]);
        ok = ok && __builtin_isfinite(a.p);
        return ok && (q[0] > T(0)) && (a.p > T(0));
    }
    // A sufficient condition for obeys(q) without its divisions (FP64):
    // rho > 2^-100, E > 2^-100, rho*E < 2^200 and rho*E > K*(1 + 2^-20) with
    // K = 0.5*m.m formed as make() forms it, for 1 + 2^-40 < gamma < 2^100.
    // Then every q is finite, |m_k| < 2^101 and |u_k| < 2^201; E exceeds
    // K/rho by a relative margin above 2^-21, more than the roundings of the
    // products, the quotient K/rho and the subtraction E - K/rho can take
    // away, so the internal energy, and p = (gamma - 1)*ie, are positive
    // normal numbers below 2^400. A false result decides nothing: the
    // caller then evaluates obeys(q) itself.
    __device__ __forceinline__ bool surely_obeys(const T* q) const {
        if constexpr (!std::is_same_v<T, double>) {
            return false;
        } else {
            T dot = q[1] * q[1];
            if (NSD > 1) dot = dot + q[2] * q[2];
            if (NSD > 2) dot = dot + q[3] * q[3];
            const T K = T(0.5) * dot;
            const T rE = q[0] * q[nvar - 1];
            return gamma > 1.0 + 0x1p-40 && gamma < 0x1p100 && q[0] > 0x1p-100 &&
                   q[nvar - 1] > 0x1p-100 && rE < 0x1p200 && rE > K * (1.0 + 0x1p-20);
        }
    }
    // The constraint check, with the division-free test first: obeys(q)
    // runs only on lanes the test does not clear.
    __device__ __forceinline__ bool valid(const T* q) const {
        if (surely_obeys(q)) return true;
        return obeys(q);
    }
};

template <class T>
struct BurgersEq {
    static constexpr int nvar = 1;
    static constexpr int nsd = 3;
    static constexpr bool is_euler = false;
    T gamma;
    struct All {
        T q[1];
    };
    __device__ __forceinline__ void make(const T* q, All& a) const { a.q[0] = q[0]; }
    template <int d>
    __device__ __forceinline__ void pointFlux(const All& a, T* F) const {
        F[0] = a.q[0] * a.q[0] / T(2);
    }
    __device__ __forceinline__ bool obeys(const T* q) const { return __builtin_isfinite(q[0]); }
    __device__ __forceinline__ bool valid(const T* q) const { return obeys(q); }
    // Burgers::computeWaveSpeed (Burgers.hpp:203-209): |u|
    template <int d>
    __device__ __forceinline__ T waveSpeed(const All& a) const {
        return dabs(a.q[0]);
    }
};

// ------------------------------------------------------------------ fluxes

struct FluxHLL {
    template <int d, class T, int NSD>
    __device__ __forceinline__ static T compute(const EulerEq<T, NSD>& eq,
                                                const typename EulerEq<T, NSD>::All& L,
                                                const typename EulerEq<T, NSD>::All& R, T* F) {
        constexpr int nv = NSD + 2;
        const T g = eq.gamma;
        const T wL = dsqrt_rho(L.q[0]);
        const T wR = dsqrt_rho(R.q[0]);
        const T rho = (L.q[0] + R.q[0]) / T(2);
        const T wsum = wL + wR;
        const T ud = dvp(L.u[d] * wL + R.u[d] * wR, wsum);
        const T p = dvp(L.p * wL + R.p * wR, wsum);
        const T cs = dsqrt(dvp(g * p, rho));
        const T sL = ref_fmin(L.u[d] - dsqrt(dvp(g * L.p, L.q[0])), ud - cs);
        const T sR = ref_fmax(R.u[d] + dsqrt(dvp(g * R.p, R.q[0])), ud + cs);
        if (sL == T(0)) {
#pragma unroll
            for (int i = 0; i < nv; ++i) F[i] = T(0);
        } else if (sL > T(0)) {
            eq.template pointFlux<d>(L, F);
        } else if (sR < T(0)) {
            eq.template pointFlux<d>(R, F);
        } else {
            T fL[nv], fR[nv];
            eq.template pointFlux<d>(L, fL);
            eq.template pointFlux<d>(R, fR);
            const T srl = sR * sL;
            const T den = sR - sL;
#pragma unroll
            for (int i = 0; i < nv; ++i)
                F[i] = dv(fL[i] * sR - fR[i] * sL + (R.q[i] - L.q[i]) * srl, den);
        }
        return ref_fmax(dabs(sL), dabs(sR));
    }
};

struct FluxHLL3 {
    template <int d, class T, int NSD>
    __device__ __forceinline__ static T compute(const EulerEq<T, NSD>& eq,
                                                const typename EulerEq<T, NSD>::All& L,
                                                const typename EulerEq<T, NSD>::All& R, T* F) {
        constexpr int nv = NSD + 2;
        const T g = eq.gamma;
        const T wL = dsqrt_rho(L.q[0]);
        const T wR = dsqrt_rho(R.q[0]);
        const T rho = (L.q[0] + R.q[0]) / T(2);
        const T wsum = wL + wR;
        const T ud = dvp(L.u[d] * wL + R.u[d] * wR, wsum);
        const T p = dvp(wL * L.p + wR * R.p, wsum);
        const T cfL = dsqrt(dvp(g * L.p, L.q[0]));
        const T cfR = dsqrt(dvp(g * R.p, R.q[0]));
        // hw_fmin/hw_fmax: corr is a zero of either sign only when uL == uR
        // or for uL = -0, uR = +0, and uL + corr, uR - corr then lose the
        // zero's sign against cfL, cfR > 0; sL, sR compare two differences,
        // each +0 when it vanishes (round to nearest), never a (+0, -0) pair
        const T corr = T(0.5) * hw_fmax(T(0), L.u[d] - R.u[d]);
        const T cs = dsqrt(dvp(g * p, rho));
        const T sL = hw_fmin(L.u[d] + corr - cfL, ud - cs);
        const T sR = hw_fmax(R.u[d] - corr + cfR, ud + cs);
        if (sL == T(0)) {
#pragma unroll
            for (int i = 0; i < nv; ++i) F[i] = T(0);
        } else if (sL > T(0)) {
            eq.template pointFlux<d>(L, F);
        } else if (sR < T(0)) {
            eq.template pointFlux<d>(R, F);
        } else {
            // The star state: only the branch taken is evaluated; the values
            // are the ones the reference computes before branching.
            T fL[nv], fR[nv];
            eq.template pointFlux<d>(L, fL);
            eq.template pointFlux<d>(R, fR);
            const T udl = L.u[d] - sL;
            const T udr = R.u[d] - sR;
            const T aa = udr * R.q[0] - udl * L.q[0];
            const T sm = dv(R.q[1 + d] * udr - L.q[1 + d] * udl + R.p - L.p, aa);
            T us[NSD];
#pragma unroll
            for (int k = 0; k < NSD; ++k) {
                if (k == d)
                    us[k] = sm;
                else
                    us[k] = dv(fR[1 + k] - fL[1 + k] - R.q[1 + k] * sR + L.q[1 + k] * sL, aa);
            }
            // The reference evaluates the two branches with the same formula
            // on mirrored operands; select the operands, then evaluate once.
            // The side's point flux is evaluated on the selected state (the
            // same function of the same values as selecting fL or fR; only
            // the transverse momentum fluxes of both sides feed us[] above).
            const bool left = sm >= T(0);
            typename EulerEq<T, NSD>::All X;
#pragma unroll
            for (int i = 0; i < nv; ++i) X.q[i] = left ? L.q[i] : R.q[i];
#pragma unroll
            for (int k = 0; k < NSD; ++k) X.u[k] = left ? L.u[k] : R.u[k];
            X.p = left ? L.p : R.p;
            T fX[nv];
            eq.template pointFlux<d>(X, fX);
            const T sX = left ? sL : sR;
            const T udx = X.u[d] - sX;  // udl or udr: the same difference
            const T rhoX = X.q[0];
            const T pX = X.p;
            const T uX = X.u[d];
            const T EX = X.q[nv - 1];
            T mid[nv];
            const T smx = sm - sX;
            mid[0] = dv(rhoX * udx, smx);
#pragma unroll
            for (int k = 0; k < NSD; ++k) mid[1 + k] = us[k] * mid[0];
            const T ps = pX + rhoX * (uX - sm) * udx;
            mid[nv - 1] = dv(udx * EX + pX * uX - ps * sm, smx);
#pragma unroll
            for (int i = 0; i < nv; ++i) F[i] = fX[i] + (mid[i] - X.q[i]) * sX;
        }
        return hw_fmax(dabs(sL), dabs(sR));  // non-negative operands
    }
};

struct FluxGodunov {
    template <int d, class T>
    __device__ __forceinline__ static T compute(const BurgersEq<T>& eq,
                                                const typename BurgersEq<T>::All& L,
                                                const typename BurgersEq<T>::All& R, T* F) {
        typename BurgersEq<T>::All a, b;
        a.q[0] = ref_fmax(L.q[0], T(0));
        b.q[0] = ref_fmin(R.q[0], T(0));
        T fl[1], fr[1];
        eq.template pointFlux<d>(a, fl);
        eq.template pointFlux<d>(b, fr);
        F[0] = ref_fmax(fl[0], fr[0]);
        return ref_fmax(dabs(L.q[0]), dabs(R.q[0]));
    }
};

// Central.hpp:38-56 (Euler and Burgers): F = 0.5*(f(L) + f(R)); wave speed
// max of the two cells' computeWaveSpeed.
struct FluxCentral {
    template <int d, class T, class Eq>
    __device__ __forceinline__ static T compute(const Eq& eq, const typename Eq::All& L,
                                                const typename Eq::All& R, T* F) {
        constexpr int nv = Eq::nvar;
        T fL[nv], fR[nv];
        eq.template pointFlux<d>(L, fL);
        eq.template pointFlux<d>(R, fR);
#pragma unroll
        for (int i = 0; i < nv; ++i) F[i] = T(0.5) * (fL[i] + fR[i]);
        return ref_fmax(eq.template waveSpeed<d>(L), eq.template waveSpeed<d>(R));
    }
};

// Rusanov.hpp:38-60 for Burgers (eigenvalue u):
// F = 0.5*(f(L) + f(R)) - 0.5*max(|uL|, |uR|)*(uR - uL).
struct FluxRusanov {
    template <int d, class T>
    __device__ __forceinline__ static T compute(const BurgersEq<T>& eq,
                                                const typename BurgersEq<T>::All& L,
                                                const typename BurgersEq<T>::All& R, T* F) {
        T fL[1], fR[1];
        eq.template pointFlux<d>(L, fL);
        eq.template pointFlux<d>(R, fR);
        const T eigenMax = ref_fmax(dabs(L.q[0]), dabs(R.q[0]));
        F[0] = T(0.5) * (fL[0] + fR[0]) - T(0.5) * eigenMax * (R.q[0] - L.q[0]);
        return ref_fmax(eq.template waveSpeed<d>(L), eq.template waveSpeed<d>(R));
    }
};

// burgers/Roe.hpp:38-62: f(max(uL, 0)) when uL >= -uR, else f(min(uR, 0)).
struct FluxRoe {
    template <int d, class T>
    __device__ __forceinline__ static T compute(const BurgersEq<T>& eq,
                                                const typename BurgersEq<T>::All& L,
                                                const typename BurgersEq<T>::All& R, T* F) {
        typename BurgersEq<T>::All a;
        a.q[0] = (L.q[0] >= -R.q[0]) ? ref_fmax(L.q[0], T(0)) : ref_fmin(R.q[0], T(0));
        eq.template pointFlux<d>(a, F);
        return ref_fmax(eq.template waveSpeed<d>(L), eq.template waveSpeed<d>(R));
    }
};

// ------------------------------------------------------------------ reconstruction
// Each reconstruction maps the three cells (i-1, i, i+1) along one direction
// to the left (i-1/2) and right (i+1/2) face values of cell i.

struct ReconNone {
    static constexpr int ghosts = 1;
    template <class T, int NV>
    __device__ __forceinline__ static void apply(const T* qm, const T* q0, const T* qp, T* L,
                                                 T* R, T /*gamma*/) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            L[v] = q0[v];
            R[v] = q0[v];
        }
    }
};

struct ReconWENO2 {
    static constexpr int ghosts = 2;
    static constexpr bool face_weights = true;
    // 1/D of the face between two neighbouring cells whose first variables
    // are lo (lower cell) and hi (upper cell): D = 3(eps + (hi - lo)^2)^2,
    // eps = 1e-8 (WENOCoefficients.hpp:21). It is a0L of the lower cell
    // (its d0 = q[i+1] - q[i]) and a1R of the upper cell (its
    // d1 = q[i+1] - q[i]): the same expression of the same two values, so a
    // marching kernel computes it once per face and hands it to both cells.
    template <class T>
    __device__ __forceinline__ static T face_weight(T lo, T hi) {
        const T eps = T(1e-8);
        const T d = hi - lo;
        const T b = d * d;
        const T e = eps + b;
        const T D = T(3) * e * e;
        return dvp(T(1), D);
    }
    // The reconstruction from the two face quantities of the cell:
    // a0L = face_weight(q0, qp) (face i+1/2), a1R = face_weight(qm, q0)
    // (face i-1/2).
    template <class T, int NV>
    __device__ __forceinline__ static void apply_w(const T* qm, const T* q0, const T* qp, T a0L,
                                                   T a1R, T* L, T* R) {
        // 2/D == 2*(1/D) exactly while 1/D is a normal number (D < 2^1021,
        // i.e. |w_{i+1}-w_i| < 2^510), so two of the four divisions are shared.
        const T a1L = T(2) * a1R;
        const T a0R = T(2) * a0L;
        const T w0L = dvp(a0L, a0L + a1L);
        const T w1L = dvp(a1L, a0L + a1L);
        const T w0R = dvp(a0R, a0R + a1R);
        const T w1R = dvp(a1R, a0R + a1R);
        const T cL = T(3) * w0L + w1L;
        const T cR = T(3) * w1R + w0R;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            L[v] = T(0.5) * (w1L * qm[v] + cL * q0[v] - w0L * qp[v]);
            R[v] = T(0.5) * (w0R * qp[v] + cR * q0[v] - w1R * qm[v]);
        }
    }
    template <class T, int NV>
    __device__ __forceinline__ static void apply(const T* qm, const T* q0, const T* qp, T* L,
                                                 T* R, T /*gamma*/) {
        apply_w<T, NV>(qm, q0, qp, face_weight(q0[0], qp[0]), face_weight(qm[0], q0[0]), L, R);
    }
};

// WENOF2 (WENOF2.hpp:31-115), Euler only: WENO2's weights from the density
// applied to the slopes of the primitive variables (rho, u, p); the density
// and pressure slopes are clamped to +-1.9 times the middle value, and each
// slope is scaled by 0.5 * sqrt(R / max(R, LL)) with R = p / (gamma - 1) and
// LL the kinetic-energy change the slope would cause, which keeps the
// reconstructed pressure positive. Expressions in the reference's order.
struct ReconWENOF2 {
    static constexpr int ghosts = 2;
    template <class T, int NV>
    __device__ __forceinline__ static void prim(const T* q, T* P, T gamma) {
        // Euler::computePrimitiveVariables (Euler.hpp:290-297): rho, m/rho,
        // (gamma-1)(E - 0.5 m.m / rho)
        constexpr int NSD = NV - 2;
        T dot = q[1] * q[1];
        if (NSD > 1) dot = dot + q[2] * q[2];
        if (NSD > 2) dot = dot + q[3] * q[3];
        P[0] = q[0];
#pragma unroll
        for (int k = 0; k < NSD; ++k) P[1 + k] = dvp(q[1 + k], q[0]);
        P[NV - 1] = (gamma - T(1)) * (q[NV - 1] - dvp(T(0.5) * dot, q[0]));
    }
    template <class T, int NV>
    __device__ __forceinline__ static T dotu(const T* a, const T* b) {
        constexpr int NSD = NV - 2;
        T d = a[1] * b[1];
        if (NSD > 1) d = d + a[2] * b[2];
        if (NSD > 2) d = d + a[3] * b[3];
        return d;
    }
    // the primitive slope dP (in place) clamped, its scale a, and the
    // conserved state of P0 + sgn * a * dP (Euler::computeConserved,
    // Euler.hpp:236-245)
    template <class T, int NV>
    __device__ __forceinline__ static void side(const T* P0, T* dP, T Rr, T sgn, T gamma, T* out) {
        const T BOUND = T(1.9);
        dP[0] = ref_fmax(-BOUND * P0[0], ref_fmin(BOUND * P0[0], dP[0]));
        dP[NV - 1] = ref_fmax(-BOUND * P0[NV - 1], ref_fmin(BOUND * P0[NV - 1], dP[NV - 1]));
        const T duu = dotu<T, NV>(dP, dP);
        const T LL = (T(0.125) * P0[0] * duu - T(0.5) * ref_fmin(T(0), dP[0] * dotu<T, NV>(P0, dP))) +
                     dvp(T(0.5) * dP[0] * dP[0] * duu, P0[0]);
        const T a = T(0.5) * dsqrt(dvp(Rr, ref_fmax(Rr, LL)));
        T P[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) P[v] = sgn > T(0) ? P0[v] + a * dP[v] : P0[v] - a * dP[v];
        out[0] = P[0];
#pragma unroll
        for (int k = 1; k < NV - 1; ++k) out[k] = P[k] * P[0];
        out[NV - 1] = dvp(P[NV - 1], gamma - T(1)) + T(0.5) * P[0] * dotu<T, NV>(P, P);
    }
    // The reconstruction from the primitive variables of the three cells
    // (P[0] is the density, the weights' input): apply() computes them;
    // a caller that already holds them (each cell's primitives serve three
    // stencils) calls this directly.
    template <class T, int NV>
    __device__ __forceinline__ static void apply_prim(const T* Pm, const T* P0, const T* Pp, T* L,
                                                      T* R, T gamma) {
        const T eps = T(1e-8);  // WENOCoefficients.hpp:21
        const T i0 = P0[0];
        const T d0 = Pp[0] - i0;
        const T d1 = i0 - Pm[0];
        const T e0 = eps + d0 * d0;
        const T e1 = eps + d1 * d1;
        const T D0 = T(3) * e0 * e0;
        const T D1 = T(3) * e1 * e1;
        // 2/D == 2*(1/D) exactly while 1/D is a normal number (as in WENO2)
        const T a0L = dvp(T(1), D0);
        const T a1R = dvp(T(1), D1);
        const T a1L = T(2) * a1R;
        const T a0R = T(2) * a0L;
        const T w0L = dvp(a0L, a0L + a1L);
        const T w1L = dvp(a1L, a0L + a1L);
        const T w0R = dvp(a0R, a0R + a1R);
        const T w1R = dvp(a1R, a0R + a1R);
        T dL[NV], dR[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            dL[v] = w1L * (P0[v] - Pm[v]) + w0L * (Pp[v] - P0[v]);
            dR[v] = w1R * (P0[v] - Pm[v]) + w0R * (Pp[v] - P0[v]);
        }
        const T Rr = dvp(P0[NV - 1], gamma - T(1));
        side<T, NV>(P0, dL, Rr, T(-1), gamma, L);
        side<T, NV>(P0, dR, Rr, T(1), gamma, R);
    }
    template <class T, int NV>
    __device__ __forceinline__ static void apply(const T* qm, const T* q0, const T* qp, T* L,
                                                 T* R, T gamma) {
        T Pm[NV], P0[NV], Pp[NV];
        prim<T, NV>(qm, Pm, gamma);
        prim<T, NV>(q0, P0, gamma);
        prim<T, NV>(qp, Pp, gamma);
        apply_prim<T, NV>(Pm, P0, Pp, L, R, gamma);
    }
    static constexpr bool uses_primitives = true;
};

struct ReconENO2 {
    static constexpr int ghosts = 2;
    template <class T, int NV>
    __device__ __forceinline__ static void apply(const T* qm, const T* q0, const T* qp, T* L,
                                                 T* R, T /*gamma*/) {
        // Divided differences of variable 0: dd_{i-1} = q_i - q_{i-1},
        // dd_i = q_{i+1} - q_i; shift left when |dd_{i-1}| < |dd_i|.
        const T ddl = q0[0] - qm[0];
        const T ddr = qp[0] - q0[0];
        const bool shift = dabs(ddl) < dabs(ddr);
        // Coefficient rows {3/2,-1/2}, {1/2,1/2}, {-1/2,3/2}: with shift the
        // stencil is (i-1, i) with rows 1 (left) and 2 (right); else (i, i+1)
        // with rows 0 and 1.
        const T cl0 = shift ? T(0.5) : T(1.5);
        const T cl1 = shift ? T(0.5) : T(-0.5);
        const T cr0 = shift ? T(-0.5) : T(0.5);
        const T cr1 = shift ? T(1.5) : T(0.5);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const T a = shift ? qm[v] : q0[v];
            const T b = shift ? q0[v] : qp[v];
            T l = T(0);
            T r = T(0);
            l += cl0 * a;
            r += cr0 * a;
            l += cl1 * b;
            r += cr1 * b;
            L[v] = l;
            R[v] = r;
        }
    }
};

// ENO of order K = 3, 4 (ENOCPU.cpp:41-185, differences 205-264,
// coefficients ENOCoefficients.cpp:34-50) on a stencil of 2K-1 cells
// q[0 .. 2K-2] around cell c = K-1. Differences of variable 0 up to level
// K-2; at every level the shift grows by one when the difference at position
// c-1 is smaller in magnitude than the one at c (the reference compares the
// same two positions at every level). Left/right face values: coefficient
// rows shift and shift+1 over cells c-shift .. c-shift+K-1, summed from 0 in
// stencil order.
template <int K>
struct EnoRows;
template <>
struct EnoRows<3> {
    static constexpr double c[4][3] = {{11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0},
                                       {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0},
                                       {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0},
                                       {1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0}};
};
template <>
struct EnoRows<4> {
    static constexpr double c[5][4] = {{25.0 / 12.0, -23.0 / 12.0, 13.0 / 12.0, -1.0 / 4.0},
                                       {1.0 / 4.0, 13.0 / 12.0, -5.0 / 12.0, 1.0 / 12.0},
                                       {-1.0 / 12.0, 7.0 / 12.0, 7.0 / 12.0, -1.0 / 12.0},
                                       {1.0 / 12.0, -5.0 / 12.0, 13.0 / 12.0, 1.0 / 4.0},
                                       {-1.0 / 4.0, 13.0 / 12.0, -23.0 / 12.0, 25.0 / 12.0}};
};

template <int K>
struct ReconENO {
    static constexpr int ghosts = K;
    static constexpr int radius = K - 1;
    template <class T, int NV>
    __device__ __forceinline__ static void apply_wide(const T (*q)[NV], T* L, T* R, T /*gamma*/) {
        constexpr int c = K - 1;
        T dd[2 * K - 2];
#pragma unroll
        for (int j = 0; j < 2 * K - 2; ++j) dd[j] = q[j + 1][0] - q[j][0];
        int shift = 0;
#pragma unroll
        for (int l = 0; l < K - 1; ++l) {
            if (dabs(dd[c - 1]) < dabs(dd[c])) ++shift;
#pragma unroll
            for (int j = 0; j < 2 * K - 3 - l; ++j) dd[j] = dd[j + 1] - dd[j];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T l = T(0);
            T r = T(0);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                T val = q[c + j][v];
                T cl = T(EnoRows<K>::c[0][j]);
                T cr = T(EnoRows<K>::c[1][j]);
#pragma unroll
                for (int s = 1; s < K; ++s)
                    if (shift == s) {
                        val = q[c - s + j][v];
                        cl = T(EnoRows<K>::c[s][j]);
                        cr = T(EnoRows<K>::c[s + 1][j]);
                    }
                l += cl * val;
                r += cr * val;
            }
            L[v] = l;
            R[v] = r;
        }
    }
};

// WENO of order 3 (WENOCPU.cpp:32-186, WENOCoefficients.hpp:51-63,
// WENOCoefficients.cpp:26-38) on the five cells q[0..4] around cell 2:
// smoothness indicators of variable 0 for the three candidate stencils,
// alpha = c / (beta + 1e-8)^2 (linear weights 3/10, 3/5, 1/10, mirrored for
// the left face), summed from the third stencil down; each candidate is the
// ENO3 row combination, weighted by alpha / sum in stencil order.
struct ReconWENO3 {
    static constexpr int ghosts = 3;
    static constexpr int radius = 2;
    template <class T, int NV>
    __device__ __forceinline__ static void apply_wide(const T (*q)[NV], T* L, T* R, T /*gamma*/) {
        const T V0 = q[0][0], V1 = q[1][0], V2 = q[2][0], V3 = q[3][0], V4 = q[4][0];
        const T c1312 = T(13.0 / 12.0), c14 = T(1 / 4.0);
        T beta[3];
        {
            const T a = V2 - T(2) * V3 + V4, b = T(3) * V2 - T(4) * V3 + V4;
            beta[0] = c1312 * (a * a) + c14 * (b * b);
        }
        {
            const T a = V1 - T(2) * V2 + V3, b = V1 - V3;
            beta[1] = c1312 * (a * a) + c14 * (b * b);
        }
        {
            const T a = V0 - T(2) * V1 + V2, b = V0 - T(4) * V1 + T(3) * V2;
            beta[2] = c1312 * (a * a) + c14 * (b * b);
        }
        const T lin[3] = {T(3.0 / 10.0), T(3.0 / 5.0), T(1.0 / 10.0)};
        const T eps = T(1e-8);
        T aR[3], aL[3];
        T sR = T(0), sL = T(0);
#pragma unroll
        for (int i = 2; i >= 0; --i) {
            const T e = beta[i] + eps;
            const T e2 = e * e;
            aR[i] = dvp(lin[i], e2);
            sR += aR[i];
            aL[i] = dvp(lin[2 - i], e2);
            sL += aL[i];
        }
        T wL[3], wR[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            wL[i] = dvp(aL[i], sL);
            wR[i] = dvp(aR[i], sR);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T lw = T(0), rw = T(0);
#pragma unroll
            for (int sh = 0; sh < 3; ++sh) {
                T l = T(0), r = T(0);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const T val = q[2 - sh + j][v];
                    l += T(EnoRows<3>::c[sh][j]) * val;
                    r += T(EnoRows<3>::c[sh + 1][j]) * val;
                }
                lw += l * wL[sh];
                rw += r * wR[sh];
            }
            L[v] = lw;
            R[v] = rw;
        }
    }
};

// Stencil radius of a reconstruction: cells each side of the reconstructed
// cell (1 unless the reconstruction declares more).
template <class R, class = void>
struct recon_radius : std::integral_constant<int, 1> {};
template <class R>
struct recon_radius<R, std::void_t<decltype(R::radius)>> : std::integral_constant<int, R::radius> {};
// Reconstructions computed from per-cell primitive variables (apply_prim).
template <class R, class = void>
struct recon_prim : std::false_type {};
template <class R>
struct recon_prim<R, std::void_t<decltype(R::uses_primitives)>> : std::true_type {};

// Reconstruction of cell q[radius] of a stencil of 2*radius+1 cells.
template <class Recon, class T, int NV>
__device__ __forceinline__ void recon_stencil(const T (*q)[NV], T* L, T* R, T gamma) {
    if constexpr (recon_radius<Recon>::value == 1)
        Recon::template apply<T, NV>(q[0], q[1], q[2], L, R, gamma);
    else
        Recon::template apply_wide<T, NV>(q, L, R, gamma);
}

// sign.hpp and minmod.hpp of the reference.
template <class T>
__device__ __forceinline__ T ref_sign(T a) {
    return a > T(0) ? T(1) : (a == T(0) ? T(0) : T(-1));
}
template <class T>
__device__ __forceinline__ T minmod2(T a, T b) {
    return ref_sign(a) == ref_sign(b) ? ref_sign(a) * ref_fmin(dabs(a), dabs(b)) : T(0);
}
template <class T>
__device__ __forceinline__ T minmod3(T a, T b, T c) {
    return (ref_sign(a) == ref_sign(b) && ref_sign(b) == ref_sign(c))
               ? ref_sign(a) * ref_fmin(dabs(a), ref_fmin(dabs(b), dabs(c)))
               : T(0);
}

// MC.hpp:29-55 (monotonized central) and MM.hpp:29-54 (minmod): per
// variable, slope sigma from the limiter; left = m - sigma/2, right = m + sigma/2.
struct ReconMC {
    static constexpr int ghosts = 2;
    template <class T, int NV>
    __device__ __forceinline__ static void apply(const T* qm, const T* q0, const T* qp, T* L,
                                                 T* R, T /*gamma*/) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const T sigma = minmod3(T(2) * (qp[v] - q0[v]), (qp[v] - qm[v]) / T(2),
                                    T(2) * (q0[v] - qm[v]));
            L[v] = q0[v] - sigma / T(2);
            R[v] = q0[v] + sigma / T(2);
        }
    }
};

struct ReconMM {
    static constexpr int ghosts = 2;
    template <class T, int NV>
    __device__ __forceinline__ static void apply(const T* qm, const T* q0, const T* qp, T* L,
                                                 T* R, T /*gamma*/) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const T sigma = minmod2(qp[v] - q0[v], q0[v] - qm[v]);
            L[v] = q0[v] - sigma / T(2);
            R[v] = q0[v] + sigma / T(2);
        }
    }
};

}  // namespace alsvinn_hip

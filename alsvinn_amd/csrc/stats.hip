// Sample statistics of the hip platform (alsuq::stats on the device) and the
// FP64 conserved totals.
//
//   meanvar          MeanVariance.cpp:30-67 + StatisticsHelper.cpp:35-79: per
//                    cell (ghosts included) S1 = sum u, S2 = sum pow(u,2) in
//                    sample order; mean = S1/M, var = S2/M - pow(mean,2).
//                    The Welford variant accumulates (mean, M2) in sample
//                    order and finalises var = M2/M.
//   structure_cube   structure_common.hpp:37-150 with ValueAtBoundary.hpp:20-98:
//                    S_p(h) = sum over cells and the cube shell of radius h of
//                    |u - u_h|^p / (nx*ny*nz), h = 1..H-1, summed over samples,
//                    divided by M.
//   structure_basic  StructureBasic.cpp:67-120: one direction, h = 0..H-1.
//   onepointmoment   OnePointMoment.cpp:30-45 + HostMemory::addPower
//                    (HostMemory.cpp:283-290): per cell sum over samples (in
//                    order) of std::pow(u, p), divided by M.
//   bv               BoundedVariation.cpp:38-53 + HostMemory::getTotalVariation
//                    (HostMemory.cpp:325-362): per variable, sum over samples
//                    of sum_cells pow(sqrt((u - u_below)^2 + (u - u_left)^2), p)
//                    over the reference's index ranges, divided by M (1D/2D);
//                    each sample's sum is sequential, in the reference's order
//                    (one lane per sample and variable adds LDS tiles of
//                    terms that other waves form).
// Accumulation is FP64 whatever the storage precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace alsvinn_hip;

// One statistic fed one sample at a time (include/alsvinn_hip.h,
// alsvinn_hip_statistics_*): FP64 accumulators of one snapshot plus the
// scratch of the structure-function partial sums and the bv terms.
struct alsvinn_hip_statistics {
    int kind = 0;
    double p = 1;
    int H = 1, direction = 0, periodic = 0;
    int nvar = 0, precision = 0, dims = 1;
    int n[3] = {1, 1, 1}, g[3] = {0, 0, 0};
    long long vstride = 0;   // padded cells per variable
    long long elements = 0;  // FP64 values per output
    int noutputs = 1;
    double* out[2] = {nullptr, nullptr};
    double* part = nullptr;  // structure functions: nvar x nblocks x H partial sums
    double* tv = nullptr;    // bv: nvar total variations of the last sample
    int nblocks = 0, tiles_x = 0;
    int tile = 0;  // structure_cube: cells per tile side of the LDS-tiled kernels, 0 = per-cell kernel
    size_t lds = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
};

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return fail(ALSVINN_HIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace {

template <class T>
__global__ __launch_bounds__(256) void meanvar_kernel(const T* __restrict__ U,
                                                     const long long* __restrict__ bases, int m,
                                                     long long n, int welford,
                                                     double* __restrict__ mean,
                                                     double* __restrict__ var) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double a = 0, b = 0;
        if (!welford) {
            // samples in order; batches of loads issued ahead of the sums
            constexpr int kBatch = 16;
            int k = 0;
            for (; k + kBatch <= m; k += kBatch) {
                double u[kBatch];
#pragma unroll
                for (int j = 0; j < kBatch; ++j) u[j] = double(U[bases[k + j] + i]);
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    a += u[j];
                    b += u[j] * u[j];
                }
            }
            for (; k < m; ++k) {
                const double u = double(U[bases[k] + i]);
                a += u;
                b += u * u;
            }
            const double M = double(m);
            const double mu = a / M;
            mean[i] = mu;
            var[i] = b / M - mu * mu;
        } else {
            for (int k = 0; k < m; ++k) {
                const double u = double(U[bases[k] + i]);
                const double d = u - a;
                a += d / double(k + 1);
                b += d * (u - a);
            }
            mean[i] = a;
            var[i] = b / double(m);
        }
    }
}

// ValueAtBoundary<PERIODIC> (wrap) / <NEUMANN> (clamp) on one axis.
__device__ __forceinline__ int wrap_index(int p, int N, int periodic) {
    if (periodic) {
        if (p < 0) p += N;
        if (p >= N) p -= N;
        if (p < 0 || p >= N) {
            p %= N;
            if (p < 0) p += N;
        }
        return p;
    }
    return p < 0 ? 0 : (p > N - 1 ? N - 1 : p);
}

__device__ __forceinline__ double pow_p(double x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * (x * x);
    if (p == 4.0) return x * (x * (x * x));
    if (p == 5.0) return x * (x * (x * (x * x)));
    return pow(x, p);
}

// x^n, 3 <= n <= 16, correctly rounded for results in the normal range: the
// running power is kept as an unevaluated sum hi + lo (Dekker's product,
// with the rounding error of hi*x recovered exactly by one FMA), so before
// the final addition hi + lo carries about 104 bits and the sum rounds once.
// Infinite or NaN hi is returned as it is (an overflowed power stays
// infinite, a NaN stays NaN); lo == 0 returns hi, so (-0)^3 is -0.
__device__ __forceinline__ double pow_int_rounded(double x, int n) {
    double hi = x, lo = 0.0;
    for (int k = 1; k < n; ++k) {
        const double h = hi * x;
        lo = fma(lo, x, fma(hi, x, -h));
        hi = h;
    }
    return (lo == 0.0 || !isfinite(hi)) ? hi : hi + lo;
}

// std::pow(x, p) with a run-time exponent, as OnePointMoment and
// BoundedVariation call it: p = 1 returns x (libm's pow(x, 1) is x, the
// exact result being representable), p = 2 the correctly rounded x*x, an
// integer p from 3 to 16 the correctly rounded power (pow_int_rounded).
// glibc's pow, which the reference calls, is within 1 ulp of these (it
// differs from the correctly rounded power in the last bit for about 0.09 %
// of inputs). Other exponents call the device pow.
__device__ __forceinline__ double ref_pow(double x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p >= 3.0 && p <= 16.0 && p == double(int(p))) return pow_int_rounded(x, int(p));
    return pow(x, p);
}

template <class T>
__global__ __launch_bounds__(256) void onepoint_kernel(const T* __restrict__ U,
                                                      const long long* __restrict__ bases, int m,
                                                      long long n, double p,
                                                      double* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double acc = 0;
        constexpr int kBatch = 16;  // loads issued ahead of the ordered sum
        int k = 0;
        for (; k + kBatch <= m; k += kBatch) {
            double u[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) u[j] = double(U[bases[k + j] + i]);
#pragma unroll
            for (int j = 0; j < kBatch; ++j) acc += ref_pow(u[j], p);
        }
        for (; k < m; ++k) acc += ref_pow(double(U[bases[k] + i]), p);
        out[i] = acc / double(m);
    }
}

// Device scratch of one statistics call, freed on every return path.
template <class T>
struct DeviceArray {
    T* p = nullptr;
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() {
        if (p) hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * (n ? n : 1)); }
};

struct StructGeom {
    int nx, ny, nz, gx, gy, gz, tx, ty, dims, periodic;
    long long vstride;
};

// Where the statistics kernels find variable `var` of sample k: the
// ensemble's buffer (a sample's volume at bases[k], variables vstride apart)
// or one reference Volume's per-variable memories (k = 0).
template <class T>
struct EnsembleSrc {
    const T* U;
    const long long* bases;
    long long vstride;
    __device__ __forceinline__ const T* at(int k, int var) const {
        return U + bases[k] + (long long)var * vstride;
    }
};
template <class T>
struct VolumeSrc {
    const T* p[5];
    __device__ __forceinline__ const T* at(int, int var) const { return p[var]; }
};

template <class T>
__device__ __forceinline__ double value_at(const T* u, const StructGeom& g, int x, int y, int z) {
    const int X = wrap_index(x, g.nx, g.periodic) + g.gx;
    const int Y = wrap_index(y, g.ny, g.periodic) + g.gy;
    const int Z = wrap_index(z, g.nz, g.periodic) + g.gz;
    return double(u[((long long)Z * g.ty + Y) * g.tx + X]);
}

// The sum of v over the wave's 64 lanes, left in lane 63: DPP row shifts
// add within each row of 16 lanes, then the row broadcasts add the rows
// (vector moves only; a shuffle tree would use the LDS pipe, which the
// structure kernels' reads keep busy). update_dpp with old = 0 and
// bound_ctrl gives 0 to lanes without a source and to rows masked off.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add_source(double v) {
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_to_last(double v) {
    v += dpp_add_source<0x111, 0xf>(v);  // row_shr:1
    v += dpp_add_source<0x112, 0xf>(v);  // row_shr:2
    v += dpp_add_source<0x114, 0xf>(v);  // row_shr:4
    v += dpp_add_source<0x118, 0xf>(v);  // row_shr:8: lane 15 of a row holds its sum
    v += dpp_add_source<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_add_source<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Block-wide sum (256 threads); thread 0 gets the result.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// One block = 256 interior cells of one sample (blockIdx.z) and one variable
// (blockIdx.y); writes partial[((k*nvar + var)*nblocks + block)*H + h].
template <class T, class Src>
__global__ __launch_bounds__(256) void structure_cube_kernel(Src src, StructGeom g, int H, double p,
                                                            int nvar, double* partial) {
    __shared__ double wsum[4][256];  // wave sums of up to 256 values of h
    const int k = blockIdx.z, var = blockIdx.y;
    const long long ncell = (long long)g.nx * g.ny * g.nz;
    const long long cell = (long long)blockIdx.x * 256 + threadIdx.x;
    const bool valid = cell < ncell;
    const int i = valid ? int(cell % g.nx) : 0;
    const int j = valid ? int((cell / g.nx) % g.ny) : 0;
    const int l = valid ? int(cell / ((long long)g.nx * g.ny)) : 0;
    const T* u = src.at(k, var);
    const double uc = value_at(u, g, i, j, l);
    const int dims = g.dims;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* out = partial + (((long long)k * nvar + var) * gridDim.x + blockIdx.x) * H;
    for (int h0 = 1; h0 < H; h0 += 256) {
        const int h1 = H < h0 + 256 ? H : h0 + 256;
        for (int h = h0; h < h1; ++h) {
            double acc = 0;
            if (valid) {
                for (int d = 0; d < dims; d++) {
                    for (int side = 0; side < 2; side++) {
                        const bool zD = d == 2, yD = d == 1, xD = d == 0;
                        const int zs = zD ? (side == 0 ? l - h : l + h) : (dims > 2 ? l - h : 0);
                        const int ze = zD ? zs + 1 : (dims > 2 ? l + h + 1 : 1);
                        const int ys = yD ? (side == 0 ? j - h : j + h) : (dims >= 2 ? j - h + 1 : 0);
                        const int ye = yD ? ys + 1 : (dims >= 2 ? j + h : 1);
                        int xs = xD ? (side == 0 ? i - h : i + h) : i - h;
                        int xe = xD ? xs + 1 : i + h + 1;
                        if (zD) {
                            xs += 1;
                            xe -= 1;
                        }
                        // value_at(u, g, x, y, z) along each row: the row's
                        // base once, then x wrapped by stepping (periodic) or
                        // clamped (Neumann)
                        for (int z = zs; z < ze; z++) {
                            const int Z = wrap_index(z, g.nz, g.periodic) + g.gz;
                            for (int y = ys; y < ye; y++) {
                                const int Y = wrap_index(y, g.ny, g.periodic) + g.gy;
                                const T* row = u + ((long long)Z * g.ty + Y) * g.tx + g.gx;
                                if (g.periodic) {
                                    int X = wrap_index(xs, g.nx, 1);
                                    for (int x = xs; x < xe; x++) {
                                        acc += pow_p(fabs(uc - double(row[X])), p);
                                        X = X + 1 == g.nx ? 0 : X + 1;
                                    }
                                } else {
                                    for (int x = xs; x < xe; x++) {
                                        const int X = x < 0 ? 0 : (x > g.nx - 1 ? g.nx - 1 : x);
                                        acc += pow_p(fabs(uc - double(row[X])), p);
                                    }
                                }
                            }
                        }
                    }
                }
            }
            acc = wave_sum_to_last(acc);
            if (lane == 63) wsum[wave][h - h0] = acc;
        }
        __syncthreads();
        for (int h = h0 + (int)threadIdx.x; h < h1; h += 256) {
            const int t = h - h0;
            out[h] = ((wsum[0][t] + wsum[1][t]) + wsum[2][t]) + wsum[3][t];
        }
        __syncthreads();
    }
}

// 2D structure_cube on LDS tiles: a block of 256 threads owns a 16 x 16 tile
// of cells of one sample (blockIdx.z) and variable (blockIdx.y) and first
// stages the tile with a halo of R = H-1 cells on every side (the boundary's
// wrap or clamp applied as it is loaded), so each of a cell's 8h shell
// neighbours is an LDS read (periodic grids run the quad kernel below).
// Every thread keeps one running sum per shell side (the same terms as
// structure_cube_kernel, in four independent chains for the FP64 adder);
// the block total of each h (wave sums by DPP, then the four waves in
// order) goes to
// partial[((k*nvar + var)*ntiles + tile)*H + h], tiles numbered row-major.
// The total is a parallel sum either way, within rounding of the
// reference's sequential loop.
constexpr int kTile = 16;

__host__ __device__ constexpr int tile_width(int H) { return kTile + 2 * (H - 1); }

// |x|^PK for a positive integer PK, as pow_p evaluates it.
template <int PK>
__device__ __forceinline__ double pow_int(double x) {
    if constexpr (PK == 1) return x;
    if constexpr (PK == 2) return x * x;
    if constexpr (PK == 3) return x * (x * x);
    if constexpr (PK == 4) return x * (x * (x * x));
    return x * (x * (x * (x * x)));
}

// PK: the exponent p when it is 1..5 (evaluated as pow_p does), 0 for any
// other p (device pow).
template <class T, int PK, class Src>
__global__ __launch_bounds__(256) void structure_cube_tiled_kernel(Src src, StructGeom g, int H,
                                                                  double p, int nvar, int tiles_x,
                                                                  double* partial) {
    extern __shared__ double lds[];
    const int W = tile_width(H);
    const int R = H - 1;
    double* tile = lds;               // W x W staged values
    double* wsum = lds + W * W;       // [4][H] wave sums
    const int k = blockIdx.z, var = blockIdx.y;
    const int x0 = (blockIdx.x % tiles_x) * kTile;
    const int y0 = (blockIdx.x / tiles_x) * kTile;
    const T* u = src.at(k, var);
    for (int idx = threadIdx.x; idx < W * W; idx += 256) {
        const int lx = idx % W, ly = idx / W;
        tile[idx] = value_at(u, g, x0 - R + lx, y0 - R + ly, 0);
    }
    __syncthreads();
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const bool valid = x0 + tx < g.nx && y0 + ty < g.ny;
    const double* c = tile + (ty + R) * W + (tx + R);  // this cell
    const double uc = *c;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto term = [&](double v) { return PK > 0 ? pow_int<PK>(fabs(uc - v)) : pow(fabs(uc - v), p); };
    // a running sum plus one term; for p = 2 the square and the addition are
    // one FMA (the sum is a parallel sum within rounding of the reference's
    // sequential loop either way; the FMA drops the product's rounding)
    auto add = [&](double acc, double v) {
        if constexpr (PK == 2) {
            const double d = uc - v;
            return __builtin_fma(d, d, acc);
        } else {
            return acc + term(v);
        }
    };
    for (int h = 1; h < H; ++h) {
        double acc = 0;
        if (valid) {
            // one running sum per shell side (independent chains for the
            // FP64 adder), added together at the end of the shell
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            // x sides: columns i-h and i+h, rows j-h+1 .. j+h-1
            const double* cl = c - h + (1 - h) * W;
            const double* cr = c + h + (1 - h) * W;
            for (int y = 0; y < 2 * h - 1; ++y) {
                a0 = add(a0, cl[y * W]);
                a1 = add(a1, cr[y * W]);
            }
            // y sides: rows j-h and j+h, columns i-h .. i+h
            const double* rb = c - h * W - h;
            const double* rt = c + h * W - h;
            for (int x = 0; x < 2 * h + 1; ++x) {
                a2 = add(a2, rb[x]);
                a3 = add(a3, rt[x]);
            }
            acc = (a0 + a1) + (a2 + a3);
        }
        acc = wave_sum_to_last(acc);
        if (lane == 63) wsum[wave * H + h] = acc;
    }
    __syncthreads();
    for (int h = 1 + (int)threadIdx.x; h < H; h += 256) {
        const double tot = ((wsum[h] + wsum[H + h]) + wsum[2 * H + h]) + wsum[3 * H + h];
        partial[(((long long)k * nvar + var) * gridDim.x + blockIdx.x) * H + h] = tot;
    }
}

// Periodic 2D structure_cube with 2 x 2 cells per thread: a block of 256
// threads owns a 32 x 32 tile of one sample and variable, staged in LDS with
// an (H-1)-cell halo on every side (the periodic wrap applied at the load).
// Each cell sums half of its shell, doubled: cell -> cell + d and
// cell -> cell - d are both bijections of the sample's cells on a periodic
// grid and the term is even in the difference, so the shell's opposite
// halves give the same total over the sample; the top row and the right
// column are the half kept (4h terms). The four cells of
// a thread share the LDS reads: the top rows j+h and j+1+h over columns
// i-h .. i+1+h and the right columns i+h and i+1+h over rows j-h+1 .. j+h,
// 8h + 4 reads per h for the four cells instead of 16h. Four running sums
// (one per cell, independent chains), DPP wave sums, block partials as in
// the tiled kernel (the tiles numbered row-major, 32 x 32 cells each).
constexpr int kQuadTile = 32;

__host__ __device__ constexpr int quad_tile_width(int H) { return kQuadTile + 2 * (H - 1); }

template <class T, int PK, class Src>
__global__ __launch_bounds__(256) void structure_cube_quad_kernel(Src src, StructGeom g, int H,
                                                                 double p, int nvar, int tiles_x,
                                                                 double* partial) {
    extern __shared__ double lds[];
    const int W = quad_tile_width(H);
    const int R = H - 1;
    double* tile = lds;          // W x W staged values
    double* wsum = lds + W * W;  // [4][H] wave sums
    const int k = blockIdx.z, var = blockIdx.y;
    const int x0 = (blockIdx.x % tiles_x) * kQuadTile;
    const int y0 = (blockIdx.x / tiles_x) * kQuadTile;
    const T* u = src.at(k, var);
    for (int idx = threadIdx.x; idx < W * W; idx += 256) {
        const int lx = idx % W, ly = idx / W;
        tile[idx] = value_at(u, g, x0 - R + lx, y0 - R + ly, 0);
    }
    __syncthreads();
    // this thread's cells (i + a, j + b), a, b in {0, 1}, tile coordinates
    const int i = 2 * ((int)threadIdx.x % 16), j = 2 * ((int)threadIdx.x / 16);
    const double* c = tile + (j + R) * W + (i + R);
    const double u00 = c[0], u10 = c[1], u01 = c[W], u11 = c[W + 1];
    const bool in_x0 = x0 + i < g.nx, in_x1 = x0 + i + 1 < g.nx;
    const bool in_y0 = y0 + j < g.ny, in_y1 = y0 + j + 1 < g.ny;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // a running sum plus the term of cell value uc and shell value v (p = 2:
    // one FMA, as in the tiled kernel)
    auto add = [&](double acc, double uc, double v) {
        if constexpr (PK == 2) {
            const double d = uc - v;
            return __builtin_fma(d, d, acc);
        } else {
            const double d = fabs(uc - v);
            return acc + (PK > 0 ? pow_int<PK>(d) : pow(d, p));
        }
    };
    for (int h = 1; h < H; ++h) {
        double a00 = 0, a10 = 0, a01 = 0, a11 = 0;
        // top rows: r0 (row j+h) for the cells of row j, r1 (row j+1+h) for
        // those of row j+1; column offsets 0 .. 2h+1 are columns i-h .. i+1+h
        const double* r0 = c + h * W - h;
        const double* r1 = r0 + W;
        a00 = add(a00, u00, r0[0]);
        a01 = add(a01, u01, r1[0]);
        for (int x = 1; x <= 2 * h; ++x) {
            const double v = r0[x], w = r1[x];
            a00 = add(a00, u00, v);
            a10 = add(a10, u10, v);
            a01 = add(a01, u01, w);
            a11 = add(a11, u11, w);
        }
        a10 = add(a10, u10, r0[2 * h + 1]);
        a11 = add(a11, u11, r1[2 * h + 1]);
        // right columns: q0 (column i+h) for the cells of column i, q1
        // (column i+1+h) for those of column i+1; row offsets 0 .. 2h-1 are
        // rows j-h+1 .. j+h
        const double* q0 = c + h + (1 - h) * W;
        const double* q1 = q0 + 1;
        a00 = add(a00, u00, q0[0]);
        a10 = add(a10, u10, q1[0]);
        for (int y = 1; y <= 2 * h - 2; ++y) {
            const double v = q0[y * W], w = q1[y * W];
            a00 = add(a00, u00, v);
            a01 = add(a01, u01, v);
            a10 = add(a10, u10, w);
            a11 = add(a11, u11, w);
        }
        a01 = add(a01, u01, q0[(2 * h - 1) * W]);
        a11 = add(a11, u11, q1[(2 * h - 1) * W]);
        const double acc = ((in_x0 && in_y0 ? a00 : 0.0) + (in_x1 && in_y0 ? a10 : 0.0)) +
                           ((in_x0 && in_y1 ? a01 : 0.0) + (in_x1 && in_y1 ? a11 : 0.0));
        const double tot = wave_sum_to_last(2.0 * acc);
        if (lane == 63) wsum[wave * H + h] = tot;
    }
    __syncthreads();
    for (int h = 1 + (int)threadIdx.x; h < H; h += 256) {
        const double tot = ((wsum[h] + wsum[H + h]) + wsum[2 * H + h]) + wsum[3 * H + h];
        partial[(((long long)k * nvar + var) * gridDim.x + blockIdx.x) * H + h] = tot;
    }
}

// PK: the exponent p when it is 1..5 (evaluated as pow_p does), 0 for any
// other p (device pow). A block takes kBasicCells x 256 cells, each thread
// kBasicCells of them 256 apart, so that one wave sum per h covers
// kBasicCells cells per lane.
constexpr int kBasicCells = 4;

template <class T, int PK, class Src>
__global__ __launch_bounds__(256) void structure_basic_kernel(Src src, StructGeom g, int H, double p,
                                                             int direction, int nvar,
                                                             double* partial) {
    // block totals per h: wave sums, then the four waves in order, for up to
    // 256 values of h between two barriers
    __shared__ double wsum[4][256];
    const int k = blockIdx.z, var = blockIdx.y;
    const long long ncell = (long long)g.nx * g.ny * g.nz;
    const T* u = src.at(k, var);
    // the neighbour at h along `direction` sits at pos = (c_d + h) mod N
    // (periodic) or min(c_d + h, N - 1) (Neumann), the wrap_index of c_d + h:
    // each cell keeps pos and its offset and steps them once per h
    const int N = direction == 0 ? g.nx : (direction == 1 ? g.ny : g.nz);
    const long long stride = direction == 0 ? 1 : (direction == 1 ? (long long)g.tx
                                                                  : (long long)g.tx * g.ty);
    const long long wrap_back = (long long)(N - 1) * stride;
    bool valid[kBasicCells];
    int pos[kBasicCells];
    long long off[kBasicCells];
    double uc[kBasicCells];
#pragma unroll
    for (int c = 0; c < kBasicCells; ++c) {
        const long long cell = ((long long)blockIdx.x * kBasicCells + c) * 256 + threadIdx.x;
        valid[c] = cell < ncell;
        const int ci = valid[c] ? int(cell % g.nx) : 0;
        const int cj = valid[c] ? int((cell / g.nx) % g.ny) : 0;
        const int cl = valid[c] ? int(cell / ((long long)g.nx * g.ny)) : 0;
        pos[c] = direction == 0 ? ci : (direction == 1 ? cj : cl);
        off[c] = ((long long)(cl + g.gz) * g.ty + (cj + g.gy)) * g.tx + (ci + g.gx);
        uc[c] = double(u[off[c]]);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* out = partial + (((long long)k * nvar + var) * gridDim.x + blockIdx.x) * H;
    for (int h0 = 0; h0 < H; h0 += 256) {
        const int h1 = H < h0 + 256 ? H : h0 + 256;
        for (int h = h0; h < h1; ++h) {
            double acc = 0;
#pragma unroll
            for (int c = 0; c < kBasicCells; ++c) {
                if (valid[c]) {
                    const double uh = double(u[off[c]]);
                    acc += PK > 0 ? pow_int<PK>(fabs(uc[c] - uh)) : pow(fabs(uc[c] - uh), p);
                }
                if (pos[c] + 1 < N) {
                    ++pos[c];
                    off[c] += stride;
                } else if (g.periodic) {
                    pos[c] = 0;
                    off[c] -= wrap_back;
                }
            }
            acc = wave_sum_to_last(acc);
            if (lane == 63) wsum[wave][h - h0] = acc;
        }
        __syncthreads();
        for (int h = h0 + (int)threadIdx.x; h < h1; h += 256) {
            const int t = h - h0;
            out[h] = ((wsum[0][t] + wsum[1][t]) + wsum[2][t]) + wsum[3][t];
        }
        __syncthreads();
    }
}

// Total variation of one variable of one sample (a chain) over the
// reference's ranges: x in [gx+1, tx-gx), y in [gy+(ty>1), ty-gy) (1D and 2D
// volumes: one z plane); the cell below is y-1 (y itself in 1D), the cell to
// the left x-1. Each chain's terms are added in the reference's loop order (x
// fastest) into one FP64 accumulator, so the sum is the reference's
// sequential sum, bit for bit. A block takes kBvChains chains: producer
// waves form kBvTile terms of every chain into one of two LDS tiles (term t
// of chain c at t * kBvPad + c) while the last wave adds the other tile's
// terms, one lane per chain; each producer keeps the loads of its next tile
// in flight across the barrier. Cells past the range give the term +0,
// which leaves the (non-negative) sum as it is. out[chain], chain = k * nvar
// + var.
constexpr int kBvChains = 8;
constexpr int kBvProducers = 4;
constexpr int kBvTile = 256;
constexpr int kBvPad = kBvChains + 1;  // tile row stride: producer stores spread over the banks
constexpr int kBvThreads = 64 * (kBvProducers + 1);
constexpr int kBvBatches = kBvTile / 64;
constexpr int kBvPerWave = kBvChains / kBvProducers;

template <class T, class Src>
__global__ __launch_bounds__(kBvThreads) void bv_kernel(Src src, StructGeom g, double p, int nvar,
                                                       int nchains, double* out) {
    __shared__ double tile_buf[2][kBvTile * kBvPad];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int chain0 = blockIdx.x * kBvChains;
    const int x0 = g.gx + 1, x1 = g.tx - g.gx;
    const int y0 = g.gy + (g.ty > 1 ? 1 : 0), y1 = g.ty - g.gy;
    const long long wx = x1 > x0 ? x1 - x0 : 0, wy = y1 > y0 ? y1 - y0 : 0;
    const long long ncell = wx * wy;
    const long long ntiles = (ncell + kBvTile - 1) / kBvTile;
    if (wave == kBvProducers) {
        double acc = 0;
        for (long long i = 0; i < ntiles; ++i) {
            __syncthreads();
            if (lane < kBvChains) {
                const double* b = tile_buf[i & 1] + lane;
#pragma unroll 64
                for (int t = 0; t < kBvTile; ++t) acc += b[t * kBvPad];
            }
        }
        __syncthreads();
        if (lane < kBvChains && chain0 + lane < nchains) out[chain0 + lane] = acc;
        return;
    }
    // producers: chains wave + kBvProducers * i; cell c = tile * kBvTile + j * 64 + lane
    // sits at (x[j], y[j]), advanced by kBvTile cells per tile
    const T* u[kBvPerWave];
#pragma unroll
    for (int i = 0; i < kBvPerWave; ++i) {
        const int chain = chain0 + wave + kBvProducers * i;
        u[i] = chain < nchains ? src.at(chain / nvar, chain % nvar) : nullptr;
    }
    const int dy_tile = wx > 0 ? int(kBvTile / wx) : 0, dx_tile = wx > 0 ? int(kBvTile % wx) : 0;
    int x[kBvBatches], y[kBvBatches];
#pragma unroll
    for (int j = 0; j < kBvBatches; ++j) {
        const int c = j * 64 + lane;
        x[j] = wx > 0 ? x0 + int(c % wx) : x0;
        y[j] = wx > 0 ? y0 + int(c / wx) : y0;
    }
    double uc[kBvPerWave][kBvBatches], ub[kBvPerWave][kBvBatches], ul[kBvPerWave][kBvBatches];
    auto load_tile = [&](long long tile) {
#pragma unroll
        for (int j = 0; j < kBvBatches; ++j) {
            const bool in = tile * kBvTile + j * 64 + lane < ncell;
            const long long row = (long long)y[j] * g.tx + x[j];
            const long long below = g.ty > 1 ? row - g.tx : x[j];
#pragma unroll
            for (int i = 0; i < kBvPerWave; ++i) {
                uc[i][j] = ub[i][j] = ul[i][j] = 0;
                if (in && u[i]) {
                    uc[i][j] = double(u[i][row]);
                    ub[i][j] = double(u[i][below]);
                    ul[i][j] = double(u[i][row - 1]);
                }
            }
            x[j] += dx_tile;
            y[j] += dy_tile;
            if (x[j] >= x1) {
                x[j] -= wx;
                ++y[j];
            }
        }
    };
    auto store_tile = [&](long long tile, double* b) {
#pragma unroll
        for (int j = 0; j < kBvBatches; ++j) {
            const bool in = tile * kBvTile + j * 64 + lane < ncell;
#pragma unroll
            for (int i = 0; i < kBvPerWave; ++i) {
                const double dyv = uc[i][j] - ub[i][j], dxv = uc[i][j] - ul[i][j];
                const double t = in ? ref_pow(sqrt(dyv * dyv + dxv * dxv), p) : 0.0;
                b[(j * 64 + lane) * kBvPad + wave + kBvProducers * i] = t;
            }
        }
    };
    if (ntiles > 0) {
        load_tile(0);
        store_tile(0, tile_buf[0]);
        if (ntiles > 1) load_tile(1);
    }
    for (long long i = 0; i < ntiles; ++i) {
        __syncthreads();
        if (i + 1 < ntiles) {
            store_tile(i + 1, tile_buf[(i + 1) & 1]);
            if (i + 2 < ntiles) load_tile(i + 2);
        }
    }
    __syncthreads();
}

// FP64 sum of every variable over the interior cells of each sample:
// partial[(s*nvar + var)*gridDim.x + block].
template <class T>
__global__ __launch_bounds__(256) void totals_kernel(const T* __restrict__ U,
                                                    const SampleDesc* __restrict__ desc, int nvar,
                                                    double* partial) {
    __shared__ double sh[256];
    const int s = blockIdx.z, var = blockIdx.y;
    const SampleDesc D = desc[s];
    const long long ncell = (long long)D.nx * D.ny * D.nz;
    double acc = 0;
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * 256) {
        const int i = int(c % D.nx), j = int((c / D.nx) % D.ny), l = int(c / ((long long)D.nx * D.ny));
        acc += double(U[D.base + var * D.vstride +
                        ((long long)(l + D.gz) * D.ty + (j + D.gy)) * D.tx + (i + D.gx)]);
    }
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[((long long)s * nvar + var) * gridDim.x + blockIdx.x] = tot;
}

int group_bases(alsvinn_hip_ensemble* e, const int* samples, int m, std::vector<long long>& bases) {
    if (!e || !samples || m < 1) return fail(ALSVINN_HIP_E_INVALID, "bad sample group");
    const SampleDesc& D0 = e->desc[samples[0] >= 0 && samples[0] < e->nsamples ? samples[0] : 0];
    for (int k = 0; k < m; ++k) {
        const int s = samples[k];
        if (s < 0 || s >= e->nsamples) return fail(ALSVINN_HIP_E_INVALID, "sample index out of range");
        const SampleDesc& D = e->desc[s];
        if (D.nx != D0.nx || D.ny != D0.ny || D.nz != D0.nz)
            return fail(ALSVINN_HIP_E_INVALID, "statistics need samples on identical grids");
        bases.push_back(D.base);
    }
    return 0;
}

StructGeom geom_of(const alsvinn_hip_ensemble* e, const SampleDesc& D) {
    StructGeom g;
    g.nx = D.nx;
    g.ny = D.ny;
    g.nz = D.nz;
    g.gx = D.gx;
    g.gy = D.gy;
    g.gz = D.gz;
    g.tx = D.tx;
    g.ty = D.ty;
    g.dims = e->dim;
    g.periodic = D.bc == 1;
    g.vstride = D.vstride;
    return g;
}

// Per-sample totals of the structure-function partials: thread (k, var, h)
// adds partial[((k*nvar + var)*nblocks + b)*H + h] over the blocks b in index
// order into sums[(k*nvar + var)*H + h]; the host then adds the samples in
// list order (each sample's structure function, then their combination, as
// the reference computes it per sample). The order is fixed, so repeated
// calls give the same bits.
template <bool ACCUMULATE>
__global__ __launch_bounds__(256) void sample_partials_kernel(const double* __restrict__ part,
                                                             long long n, int nblocks, int H,
                                                             double* __restrict__ sums) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long kv = t / H;
    const int h = int(t % H);
    const double* src = part + kv * nblocks * H + h;
    double acc = 0;
    constexpr int kBatch = 16;  // loads issued ahead of the ordered additions
    int b = 0;
    for (; b + kBatch <= nblocks; b += kBatch) {
        double v[kBatch];
#pragma unroll
        for (int i = 0; i < kBatch; ++i) v[i] = src[(long long)(b + i) * H];
#pragma unroll
        for (int i = 0; i < kBatch; ++i) acc += v[i];
    }
    for (; b < nblocks; ++b) acc += src[(long long)b * H];
    if (ACCUMULATE)
        sums[t] += acc;  // the per-sample statistics: the samples added in call order
    else
        sums[t] = acc;
}

template <class T, class Src>
void launch_basic(int pk, dim3 grid, hipStream_t st, Src src, StructGeom g, int H, double p, int dir,
                  int nvar, double* part) {
    switch (pk) {
        case 1: hipLaunchKernelGGL((structure_basic_kernel<T, 1, Src>), grid, dim3(256), 0, st, src, g, H, p, dir, nvar, part); break;
        case 2: hipLaunchKernelGGL((structure_basic_kernel<T, 2, Src>), grid, dim3(256), 0, st, src, g, H, p, dir, nvar, part); break;
        case 3: hipLaunchKernelGGL((structure_basic_kernel<T, 3, Src>), grid, dim3(256), 0, st, src, g, H, p, dir, nvar, part); break;
        case 4: hipLaunchKernelGGL((structure_basic_kernel<T, 4, Src>), grid, dim3(256), 0, st, src, g, H, p, dir, nvar, part); break;
        case 5: hipLaunchKernelGGL((structure_basic_kernel<T, 5, Src>), grid, dim3(256), 0, st, src, g, H, p, dir, nvar, part); break;
        default: hipLaunchKernelGGL((structure_basic_kernel<T, 0, Src>), grid, dim3(256), 0, st, src, g, H, p, dir, nvar, part); break;
    }
}

template <class T, class Src>
void launch_tiled(int pk, dim3 grid, size_t lds, hipStream_t st, Src src, StructGeom g, int H,
                  double p, int nvar, int tiles_x, double* part) {
    switch (pk) {
        case 1: hipLaunchKernelGGL((structure_cube_tiled_kernel<T, 1, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 2: hipLaunchKernelGGL((structure_cube_tiled_kernel<T, 2, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 3: hipLaunchKernelGGL((structure_cube_tiled_kernel<T, 3, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 4: hipLaunchKernelGGL((structure_cube_tiled_kernel<T, 4, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 5: hipLaunchKernelGGL((structure_cube_tiled_kernel<T, 5, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        default: hipLaunchKernelGGL((structure_cube_tiled_kernel<T, 0, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
    }
}

// How a 2D structure_cube runs: periodic grids on 32 x 32 tiles with 2 x 2
// cells per thread (the staged window and the wave sums within the 160 KB
// of LDS a block may use on gfx950), other grids on 16 x 16 tiles (within
// 64 KB); 0: the per-cell kernel reading global memory (3D, H < 2, or a
// halo too wide). *lds: the block's dynamic LDS bytes.
int cube_tile(int dims, int H, bool periodic, size_t* lds) {
    if (dims != 2 || H < 2) return 0;
    const size_t wsum = 4 * size_t(H) * sizeof(double);
    const size_t quad = size_t(quad_tile_width(H)) * quad_tile_width(H) * sizeof(double) + wsum;
    const size_t single = size_t(tile_width(H)) * tile_width(H) * sizeof(double) + wsum;
    if (periodic && quad <= 160 * 1024) {
        *lds = quad;
        return kQuadTile;
    }
    if (single <= 65536) {
        *lds = single;
        return kTile;
    }
    return 0;
}

template <class T, class Src>
void launch_cube(int tile, int pk, dim3 grid, size_t lds, hipStream_t st, Src src, StructGeom g,
                 int H, double p, int nvar, int tiles_x, double* part) {
    if (tile != kQuadTile) {
        launch_tiled<T>(pk, grid, lds, st, src, g, H, p, nvar, tiles_x, part);
        return;
    }
    switch (pk) {
        case 1: hipLaunchKernelGGL((structure_cube_quad_kernel<T, 1, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 2: hipLaunchKernelGGL((structure_cube_quad_kernel<T, 2, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 3: hipLaunchKernelGGL((structure_cube_quad_kernel<T, 3, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 4: hipLaunchKernelGGL((structure_cube_quad_kernel<T, 4, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        case 5: hipLaunchKernelGGL((structure_cube_quad_kernel<T, 5, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
        default: hipLaunchKernelGGL((structure_cube_quad_kernel<T, 0, Src>), grid, dim3(256), lds, st, src, g, H, p, nvar, tiles_x, part); break;
    }
}

// nblocks_override > 0: the kernel's own block count per (sample, variable)
// (the tiled kernels' tiles, structure_basic's blocks of kBasicCells x 256
// cells); otherwise one block per 256 cells.
template <class Launch>
int run_structure(alsvinn_hip_ensemble* e, const int* samples, int m, int H, double* host_out,
                  Launch launch, int nblocks_override = 0) {
    std::vector<long long> bases;
    if (int rc = group_bases(e, samples, m, bases)) return rc;
    const SampleDesc& D = e->desc[samples[0]];
    const long long ncell = (long long)D.nx * D.ny * D.nz;
    const int nblocks = nblocks_override > 0 ? nblocks_override : int((ncell + 255) / 256);
    DeviceArray<long long> d_bases;
    DeviceArray<double> d_part;
    const size_t npart = (size_t)m * e->nvar * nblocks * H;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(d_bases.alloc(m));
    HIP_TRY(d_part.alloc(npart));
    HIP_TRY(hipMemcpy(d_bases.p, bases.data(), sizeof(long long) * m, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d_part.p, 0, sizeof(double) * npart));
    DeviceArray<double> d_sum;
    const long long nsum = (long long)m * e->nvar * H;
    HIP_TRY(d_sum.alloc((size_t)nsum));
    launch(dim3(nblocks, e->nvar, m), d_bases.p, geom_of(e, D), d_part.p);
    hipError_t le = hipGetLastError();
    // fixed-order reduction: each sample's blocks in index order on the
    // device, then the samples in list order on the host
    if (le == hipSuccess) {
        hipLaunchKernelGGL(sample_partials_kernel<false>, dim3(unsigned((nsum + 255) / 256)), dim3(256), 0,
                           e->stream, d_part.p, nsum, nblocks, H, d_sum.p);
        le = hipGetLastError();
    }
    hipError_t se = hipStreamSynchronize(e->stream);
    if (le == hipSuccess) le = se;
    std::vector<double> sums((size_t)nsum);
    hipError_t ce = hipMemcpy(sums.data(), d_sum.p, sizeof(double) * sums.size(), hipMemcpyDeviceToHost);
    if (le != hipSuccess) return fail(ALSVINN_HIP_E_HIP, hipGetErrorString(le));
    if (ce != hipSuccess) return fail(ALSVINN_HIP_E_HIP, hipGetErrorString(ce));
    std::vector<double> acc((size_t)e->nvar * H, 0.0);
    for (int k = 0; k < m; ++k)
        for (size_t i = 0; i < acc.size(); ++i) acc[i] += sums[(size_t)k * acc.size() + i];
    const double N = double(ncell);
    for (int var = 0; var < e->nvar; ++var)
        for (int h = 0; h < H; ++h)
            host_out[(size_t)var * H + h] = (acc[(size_t)var * H + h] / N) / double(m);
    return 0;
}

// ---- statistics fed one sample at a time (a reference Volume per call)

// meanvar: a1 += u, a2 += u*u; onepointmoment: a1 += pow(u, p); every
// padded cell of every variable (index var*n + i), as the ensemble kernels
// add the samples in order.
template <class T>
__global__ __launch_bounds__(256) void sample_moments_kernel(VolumeSrc<T> src, long long n, int nvar,
                                                            int meanvar, double p,
                                                            double* __restrict__ a1,
                                                            double* __restrict__ a2) {
    const long long total = n * nvar;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int var = int(idx / n);
        const double u = double(src.p[var][idx - (long long)var * n]);
        if (meanvar) {
            a1[idx] += u;
            a2[idx] += u * u;
        } else {
            a1[idx] += ref_pow(u, p);
        }
    }
}

__global__ void add_vec_kernel(double* __restrict__ acc, const double* __restrict__ x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] += x[i];
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double* x, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = 0;
}

// combineStatistics: x / M, or (x / N) / M for the structure functions' sums
// over N interior cells (the ensemble entry points' order of operations).
__global__ __launch_bounds__(256) void combine_kernel(double* x, long long n, double N, double M,
                                                     int per_cell) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = per_cell ? (x[i] / N) / M : x[i] / M;
}

// MeanVariance::finalizeStatistics: variance -= pow(mean, 2).
__global__ __launch_bounds__(256) void variance_kernel(const double* __restrict__ mean,
                                                      double* __restrict__ var, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        var[i] = var[i] - mean[i] * mean[i];
}

int grid_blocks(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 65536)); }

}  // namespace

extern "C" {

int alsvinn_hip_stats_meanvar(alsvinn_hip_ensemble* e, const int* samples, int m, int welford,
                              double* host_mean, double* host_var) {
    std::vector<long long> bases;
    if (int rc = group_bases(e, samples, m, bases)) return rc;
    if (!host_mean || !host_var) return fail(ALSVINN_HIP_E_INVALID, "output is NULL");
    const SampleDesc& D = e->desc[samples[0]];
    const long long n = D.vstride * e->nvar;
    DeviceArray<long long> d_bases;
    DeviceArray<double> d_mean, d_var;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(d_bases.alloc(m));
    HIP_TRY(d_mean.alloc(n));
    HIP_TRY(d_var.alloc(n));
    HIP_TRY(hipMemcpy(d_bases.p, bases.data(), sizeof(long long) * m, hipMemcpyHostToDevice));
    const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
    if (e->esize == 8)
        hipLaunchKernelGGL(meanvar_kernel<double>, dim3(blocks), dim3(256), 0, e->stream,
                           (const double*)e->U, d_bases.p, m, n, welford, d_mean.p, d_var.p);
    else
        hipLaunchKernelGGL(meanvar_kernel<float>, dim3(blocks), dim3(256), 0, e->stream,
                           (const float*)e->U, d_bases.p, m, n, welford, d_mean.p, d_var.p);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(e->stream);
    hipError_t c1 = hipMemcpy(host_mean, d_mean.p, sizeof(double) * n, hipMemcpyDeviceToHost);
    hipError_t c2 = hipMemcpy(host_var, d_var.p, sizeof(double) * n, hipMemcpyDeviceToHost);
    for (hipError_t x : {le, se, c1, c2})
        if (x != hipSuccess) return fail(ALSVINN_HIP_E_HIP, hipGetErrorString(x));
    return 0;
}

int alsvinn_hip_stats_onepoint_moment(alsvinn_hip_ensemble* e, const int* samples, int m,
                                      double p, double* host_out) {
    std::vector<long long> bases;
    if (int rc = group_bases(e, samples, m, bases)) return rc;
    if (!host_out) return fail(ALSVINN_HIP_E_INVALID, "output is NULL");
    const SampleDesc& D = e->desc[samples[0]];
    const long long n = D.vstride * e->nvar;
    DeviceArray<long long> d_bases;
    DeviceArray<double> d_out;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(d_bases.alloc(m));
    HIP_TRY(d_out.alloc(n));
    HIP_TRY(hipMemcpy(d_bases.p, bases.data(), sizeof(long long) * m, hipMemcpyHostToDevice));
    const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
    if (e->esize == 8)
        hipLaunchKernelGGL(onepoint_kernel<double>, dim3(blocks), dim3(256), 0, e->stream,
                           (const double*)e->U, d_bases.p, m, n, p, d_out.p);
    else
        hipLaunchKernelGGL(onepoint_kernel<float>, dim3(blocks), dim3(256), 0, e->stream,
                           (const float*)e->U, d_bases.p, m, n, p, d_out.p);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(e->stream);
    hipError_t c1 = hipMemcpy(host_out, d_out.p, sizeof(double) * n, hipMemcpyDeviceToHost);
    for (hipError_t x : {le, se, c1})
        if (x != hipSuccess) return fail(ALSVINN_HIP_E_HIP, hipGetErrorString(x));
    return 0;
}

int alsvinn_hip_stats_bounded_variation(alsvinn_hip_ensemble* e, const int* samples, int m,
                                        double p, double* host_out) {
    if (!e || !host_out) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    if (e->dim > 2) return fail(ALSVINN_HIP_E_INVALID, "bv: not supported for 3d (as in the reference)");
    std::vector<long long> bases;
    if (int rc = group_bases(e, samples, m, bases)) return rc;
    const SampleDesc& D = e->desc[samples[0]];
    if (D.tz != 1) return fail(ALSVINN_HIP_E_INVALID, "bv: not supported for 3d (as in the reference)");
    const size_t np = (size_t)m * e->nvar;
    DeviceArray<long long> d_bases;
    DeviceArray<double> d_tv;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(d_bases.alloc(m));
    HIP_TRY(d_tv.alloc(np));
    HIP_TRY(hipMemcpy(d_bases.p, bases.data(), sizeof(long long) * m, hipMemcpyHostToDevice));
    const int nchains = int(np);
    const dim3 grid((nchains + kBvChains - 1) / kBvChains);
    const StructGeom gd = geom_of(e, D);
    if (e->esize == 8)
        hipLaunchKernelGGL((bv_kernel<double, EnsembleSrc<double>>), grid, dim3(kBvThreads), 0,
                           e->stream, EnsembleSrc<double>{(const double*)e->U, d_bases.p, gd.vstride},
                           gd, p, e->nvar, nchains, d_tv.p);
    else
        hipLaunchKernelGGL((bv_kernel<float, EnsembleSrc<float>>), grid, dim3(kBvThreads), 0,
                           e->stream, EnsembleSrc<float>{(const float*)e->U, d_bases.p, gd.vstride},
                           gd, p, e->nvar, nchains, d_tv.p);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(e->stream);
    std::vector<double> tv(np);
    hipError_t ce = hipMemcpy(tv.data(), d_tv.p, sizeof(double) * np, hipMemcpyDeviceToHost);
    for (hipError_t x : {le, se, ce})
        if (x != hipSuccess) return fail(ALSVINN_HIP_E_HIP, hipGetErrorString(x));
    // the statistic's snapshot adds the samples' variations in order, then
    // the combine divides by M
    for (int var = 0; var < e->nvar; ++var) {
        double acc = 0;
        for (int k = 0; k < m; ++k) acc += tv[(size_t)k * e->nvar + var];
        host_out[var] = acc / double(m);
    }
    return 0;
}

int alsvinn_hip_stats_structure_cube(alsvinn_hip_ensemble* e, const int* samples, int m,
                                     int number_of_h, double p, double* host_out) {
    if (!e || number_of_h < 1 || !host_out) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    const int H = number_of_h;
    // 2D with a halo that fits in LDS: the tiled kernels
    if (m > 0 && samples) {
        if (samples[0] < 0 || samples[0] >= e->nsamples)
            return fail(ALSVINN_HIP_E_INVALID, "sample index out of range");
        const SampleDesc& D = e->desc[samples[0]];
        size_t lds = 0;
        const int tile = cube_tile(e->dim, H, D.bc == 1, &lds);
        if (tile) {
            const int tiles_x = (D.nx + tile - 1) / tile;
            const int tiles = tiles_x * ((D.ny + tile - 1) / tile);
            return run_structure(
                e, samples, m, H, host_out,
                [&](dim3 grid, const long long* bases, StructGeom g, double* part) {
                    const int pk =
                        (p == 1.0 || p == 2.0 || p == 3.0 || p == 4.0 || p == 5.0) ? int(p) : 0;
                    if (e->esize == 8)
                        launch_cube<double>(tile, pk, grid, lds, e->stream,
                                            EnsembleSrc<double>{(const double*)e->U, bases, g.vstride},
                                            g, H, p, e->nvar, tiles_x, part);
                    else
                        launch_cube<float>(tile, pk, grid, lds, e->stream,
                                           EnsembleSrc<float>{(const float*)e->U, bases, g.vstride},
                                           g, H, p, e->nvar, tiles_x, part);
                },
                tiles);
        }
    }
    return run_structure(e, samples, m, number_of_h, host_out,
                         [&](dim3 grid, const long long* bases, StructGeom g, double* part) {
                             if (e->esize == 8)
                                 hipLaunchKernelGGL((structure_cube_kernel<double, EnsembleSrc<double>>),
                                                    grid, dim3(256), 0, e->stream,
                                                    EnsembleSrc<double>{(const double*)e->U, bases, g.vstride},
                                                    g, number_of_h, p, e->nvar, part);
                             else
                                 hipLaunchKernelGGL((structure_cube_kernel<float, EnsembleSrc<float>>),
                                                    grid, dim3(256), 0, e->stream,
                                                    EnsembleSrc<float>{(const float*)e->U, bases, g.vstride},
                                                    g, number_of_h, p, e->nvar, part);
                         });
}

int alsvinn_hip_stats_structure_basic(alsvinn_hip_ensemble* e, const int* samples, int m,
                                      int number_of_h, double p, int direction,
                                      double* host_out) {
    if (!e || number_of_h < 1 || !host_out || direction < 0 || direction >= e->dim)
        return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    int nblocks = 0;  // kBasicCells x 256 cells per block
    if (m > 0 && samples && samples[0] >= 0 && samples[0] < e->nsamples) {
        const SampleDesc& D = e->desc[samples[0]];
        const long long ncell = (long long)D.nx * D.ny * D.nz;
        nblocks = int((ncell + 256 * kBasicCells - 1) / (256 * kBasicCells));
    }
    return run_structure(e, samples, m, number_of_h, host_out,
                         [&](dim3 grid, const long long* bases, StructGeom g, double* part) {
                             const int pk = (p == 1.0 || p == 2.0 || p == 3.0 || p == 4.0 || p == 5.0)
                                                ? int(p) : 0;
                             if (e->esize == 8)
                                 launch_basic<double>(pk, grid, e->stream,
                                                      EnsembleSrc<double>{(const double*)e->U, bases, g.vstride},
                                                      g, number_of_h, p, direction, e->nvar, part);
                             else
                                 launch_basic<float>(pk, grid, e->stream,
                                                     EnsembleSrc<float>{(const float*)e->U, bases, g.vstride},
                                                     g, number_of_h, p, direction, e->nvar, part);
                         },
                         nblocks);
}

int alsvinn_hip_ensemble_totals(alsvinn_hip_ensemble* e, double* host) {
    if (!e || !host) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    const int blocks = 64;
    const size_t np = (size_t)e->nsamples * e->nvar * blocks;
    DeviceArray<double> d_part;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(d_part.alloc(np));
    if (e->esize == 8)
        hipLaunchKernelGGL(totals_kernel<double>, dim3(blocks, e->nvar, e->nsamples), dim3(256), 0,
                           e->stream, (const double*)e->U, e->d_desc, e->nvar, d_part.p);
    else
        hipLaunchKernelGGL(totals_kernel<float>, dim3(blocks, e->nvar, e->nsamples), dim3(256), 0,
                           e->stream, (const float*)e->U, e->d_desc, e->nvar, d_part.p);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(e->stream);
    if (le == hipSuccess) le = se;
    std::vector<double> part(np);
    hipError_t ce = hipMemcpy(part.data(), d_part.p, sizeof(double) * np, hipMemcpyDeviceToHost);
    if (le != hipSuccess) return fail(ALSVINN_HIP_E_HIP, hipGetErrorString(le));
    if (ce != hipSuccess) return fail(ALSVINN_HIP_E_HIP, hipGetErrorString(ce));
    for (int s = 0; s < e->nsamples; ++s)
        for (int v = 0; v < e->nvar; ++v) {
            double acc = 0;
            for (int b = 0; b < blocks; ++b) acc += part[((size_t)s * e->nvar + v) * blocks + b];
            host[(size_t)s * e->nvar + v] = acc;
        }
    return 0;
}

int alsvinn_hip_statistics_create(int kind, double p, int number_of_h, int direction, int boundary,
                                  const alsvinn_hip_volume* shape, void* stream,
                                  alsvinn_hip_statistics** out) {
    if (!out) return fail(ALSVINN_HIP_E_INVALID, "out is NULL");
    *out = nullptr;
    if (kind < ALSVINN_HIP_STAT_MEANVAR || kind > ALSVINN_HIP_STAT_BV)
        return fail(ALSVINN_HIP_E_INVALID, "unknown statistics kind " + std::to_string(kind));
    if (!shape || shape->nvar < 1 || shape->nvar > 5 || shape->nx < 1 || shape->ny < 1 ||
        shape->nz < 1 || shape->gx < 0 || shape->gy < 0 || shape->gz < 0 ||
        (shape->precision != ALSVINN_HIP_FP64 && shape->precision != ALSVINN_HIP_FP32))
        return fail(ALSVINN_HIP_E_INVALID, "bad sample shape");
    if (boundary != ALSVINN_HIP_BC_NEUMANN && boundary != ALSVINN_HIP_BC_PERIODIC)
        return fail(ALSVINN_HIP_E_INVALID, "unknown boundary " + std::to_string(boundary));
    const bool structure = kind == ALSVINN_HIP_STAT_STRUCTURE_CUBE || kind == ALSVINN_HIP_STAT_STRUCTURE_BASIC;
    const int dims = 1 + (shape->ny > 1) + (shape->nz > 1);
    if (structure && number_of_h < 1) return fail(ALSVINN_HIP_E_INVALID, "numberOfH must be >= 1");
    if (kind == ALSVINN_HIP_STAT_STRUCTURE_BASIC && (direction < 0 || direction >= dims))
        return fail(ALSVINN_HIP_E_INVALID, "structure_basic direction must be an active axis");
    // HostMemory::getTotalVariation throws for any memory more than one cell
    // deep in z, ghost layers included
    if (kind == ALSVINN_HIP_STAT_BV && (dims > 2 || shape->nz + 2 * shape->gz > 1))
        return fail(ALSVINN_HIP_E_INVALID, "bv: not supported for 3d (as in the reference)");
    auto* st = new alsvinn_hip_statistics();
    st->kind = kind;
    st->p = p;
    st->H = structure ? number_of_h : 1;
    st->direction = direction;
    st->periodic = boundary == ALSVINN_HIP_BC_PERIODIC;
    st->nvar = shape->nvar;
    st->precision = shape->precision;
    st->dims = dims;
    st->n[0] = shape->nx;
    st->n[1] = shape->ny;
    st->n[2] = shape->nz;
    st->g[0] = shape->gx;
    st->g[1] = shape->gy;
    st->g[2] = shape->gz;
    st->vstride = (long long)(shape->nx + 2 * shape->gx) * (shape->ny + 2 * shape->gy) *
                  (shape->nz + 2 * shape->gz);
    const long long ncell = (long long)shape->nx * shape->ny * shape->nz;
    if (kind == ALSVINN_HIP_STAT_MEANVAR || kind == ALSVINN_HIP_STAT_ONEPOINTMOMENT) {
        st->elements = st->vstride * st->nvar;
        st->noutputs = kind == ALSVINN_HIP_STAT_MEANVAR ? 2 : 1;
    } else if (structure) {
        st->elements = (long long)st->nvar * st->H;
        const int H = st->H;
        st->tile = kind == ALSVINN_HIP_STAT_STRUCTURE_CUBE ? cube_tile(dims, H, st->periodic, &st->lds) : 0;
        if (st->tile) {
            st->tiles_x = (shape->nx + st->tile - 1) / st->tile;
            st->nblocks = st->tiles_x * ((shape->ny + st->tile - 1) / st->tile);
        } else if (kind == ALSVINN_HIP_STAT_STRUCTURE_BASIC) {
            st->nblocks = int((ncell + 256 * kBasicCells - 1) / (256 * kBasicCells));
        } else {
            st->nblocks = int((ncell + 255) / 256);
        }
    } else {
        st->elements = st->nvar;
    }
    auto cleanup = [&](hipError_t e, const char* what) {
        alsvinn_hip_statistics_destroy(st);
        return fail(ALSVINN_HIP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    if (stream) {
        st->stream = (hipStream_t)stream;
    } else {
        if ((e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking)) != hipSuccess)
            return cleanup(e, "hipStreamCreateWithFlags");
        st->own_stream = true;
    }
    for (int k = 0; k < st->noutputs; ++k)
        if ((e = hipMalloc(&st->out[k], sizeof(double) * st->elements)) != hipSuccess)
            return cleanup(e, "hipMalloc");
    if (structure) {
        const size_t np = (size_t)st->nvar * st->nblocks * st->H;
        if ((e = hipMalloc(&st->part, sizeof(double) * np)) != hipSuccess) return cleanup(e, "hipMalloc");
        // entries a kernel never writes (structure_cube's h = 0) stay zero
        if ((e = hipMemset(st->part, 0, sizeof(double) * np)) != hipSuccess) return cleanup(e, "hipMemset");
    }
    if (kind == ALSVINN_HIP_STAT_BV)
        if ((e = hipMalloc(&st->tv, sizeof(double) * st->nvar)) != hipSuccess) return cleanup(e, "hipMalloc");
    if (int rc = alsvinn_hip_statistics_reset(st)) {
        alsvinn_hip_statistics_destroy(st);
        return rc;
    }
    if ((e = hipStreamSynchronize(st->stream)) != hipSuccess) return cleanup(e, "hipStreamSynchronize");
    *out = st;
    return 0;
}

int alsvinn_hip_statistics_destroy(alsvinn_hip_statistics* st) {
    if (!st) return 0;
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    for (int k = 0; k < 2; ++k) (void)hipFree(st->out[k]);
    (void)hipFree(st->part);
    (void)hipFree(st->tv);
    if (st->own_stream && st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
    return 0;
}

int alsvinn_hip_statistics_reset(alsvinn_hip_statistics* st) {
    if (!st) return fail(ALSVINN_HIP_E_INVALID, "statistics is NULL");
    for (int k = 0; k < st->noutputs; ++k)
        hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_blocks(st->elements)), dim3(256), 0, st->stream,
                           st->out[k], st->elements);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"

template <class T>
static void statistics_compute(alsvinn_hip_statistics* st, const alsvinn_hip_volume& v) {
    VolumeSrc<T> src;
    for (int k = 0; k < 5; ++k) src.p[k] = k < st->nvar ? (const T*)v.var[k] : nullptr;
    StructGeom g;
    g.nx = st->n[0];
    g.ny = st->n[1];
    g.nz = st->n[2];
    g.gx = st->g[0];
    g.gy = st->g[1];
    g.gz = st->g[2];
    g.tx = st->n[0] + 2 * st->g[0];
    g.ty = st->n[1] + 2 * st->g[1];
    g.dims = st->dims;
    g.periodic = st->periodic;
    g.vstride = st->vstride;
    const int pk = (st->p == 1.0 || st->p == 2.0 || st->p == 3.0 || st->p == 4.0 || st->p == 5.0)
                       ? int(st->p) : 0;
    switch (st->kind) {
        case ALSVINN_HIP_STAT_MEANVAR:
        case ALSVINN_HIP_STAT_ONEPOINTMOMENT:
            hipLaunchKernelGGL(sample_moments_kernel<T>, dim3(grid_blocks(st->elements)), dim3(256), 0,
                               st->stream, src, st->vstride, st->nvar,
                               st->kind == ALSVINN_HIP_STAT_MEANVAR ? 1 : 0, st->p, st->out[0],
                               st->out[1]);
            break;
        case ALSVINN_HIP_STAT_STRUCTURE_CUBE:
        case ALSVINN_HIP_STAT_STRUCTURE_BASIC: {
            const dim3 grid(st->nblocks, st->nvar, 1);
            if (st->kind == ALSVINN_HIP_STAT_STRUCTURE_BASIC)
                launch_basic<T>(pk, grid, st->stream, src, g, st->H, st->p, st->direction, st->nvar,
                                st->part);
            else if (st->tile)
                launch_cube<T>(st->tile, pk, grid, st->lds, st->stream, src, g, st->H, st->p,
                               st->nvar, st->tiles_x, st->part);
            else
                hipLaunchKernelGGL((structure_cube_kernel<T, VolumeSrc<T>>), grid, dim3(256), 0,
                                   st->stream, src, g, st->H, st->p, st->nvar, st->part);
            const long long nsum = st->elements;
            hipLaunchKernelGGL(sample_partials_kernel<true>, dim3(unsigned((nsum + 255) / 256)), dim3(256),
                               0, st->stream, st->part, nsum, st->nblocks, st->H, st->out[0]);
            break;
        }
        case ALSVINN_HIP_STAT_BV:
            hipLaunchKernelGGL((bv_kernel<T, VolumeSrc<T>>), dim3((st->nvar + kBvChains - 1) / kBvChains),
                               dim3(kBvThreads), 0, st->stream, src, g, st->p, st->nvar, st->nvar,
                               st->tv);
            hipLaunchKernelGGL(add_vec_kernel, dim3(1), dim3(64), 0, st->stream, st->out[0], st->tv,
                               st->nvar);
            break;
    }
}

extern "C" {

int alsvinn_hip_statistics_compute(alsvinn_hip_statistics* st, const alsvinn_hip_volume* sample) {
    if (!st || !sample) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    if (sample->nvar != st->nvar || sample->precision != st->precision || sample->nx != st->n[0] ||
        sample->ny != st->n[1] || sample->nz != st->n[2] || sample->gx != st->g[0] ||
        sample->gy != st->g[1] || sample->gz != st->g[2])
        return fail(ALSVINN_HIP_E_INVALID, "sample volume does not match the statistics' shape");
    for (int k = 0; k < st->nvar; ++k)
        if (!sample->var[k]) return fail(ALSVINN_HIP_E_INVALID, "sample variable pointer is NULL");
    if (st->precision == ALSVINN_HIP_FP64)
        statistics_compute<double>(st, *sample);
    else
        statistics_compute<float>(st, *sample);
    HIP_TRY(hipGetLastError());
    return 0;
}

int alsvinn_hip_statistics_combine(alsvinn_hip_statistics* st, long long samples) {
    if (!st || samples < 1) return fail(ALSVINN_HIP_E_INVALID, "bad arguments");
    const bool per_cell = st->kind == ALSVINN_HIP_STAT_STRUCTURE_CUBE ||
                          st->kind == ALSVINN_HIP_STAT_STRUCTURE_BASIC;
    const double N = double((long long)st->n[0] * st->n[1] * st->n[2]);
    for (int k = 0; k < st->noutputs; ++k)
        hipLaunchKernelGGL(combine_kernel, dim3(grid_blocks(st->elements)), dim3(256), 0, st->stream,
                           st->out[k], st->elements, N, double(samples), per_cell ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return 0;
}

int alsvinn_hip_statistics_finalize(alsvinn_hip_statistics* st) {
    if (!st) return fail(ALSVINN_HIP_E_INVALID, "statistics is NULL");
    if (st->kind == ALSVINN_HIP_STAT_MEANVAR) {
        hipLaunchKernelGGL(variance_kernel, dim3(grid_blocks(st->elements)), dim3(256), 0, st->stream,
                           st->out[0], st->out[1], st->elements);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

int alsvinn_hip_statistics_outputs(const alsvinn_hip_statistics* st, int* count,
                                   long long* elements) {
    if (!st) return fail(ALSVINN_HIP_E_INVALID, "statistics is NULL");
    if (count) *count = st->noutputs;
    if (elements) *elements = st->elements;
    return 0;
}

int alsvinn_hip_statistics_result(alsvinn_hip_statistics* st, int index, double** device) {
    if (!st || !device || index < 0 || index >= st->noutputs)
        return fail(ALSVINN_HIP_E_INVALID, "bad output index");
    *device = st->out[index];
    return 0;
}

int alsvinn_hip_statistics_download(alsvinn_hip_statistics* st, int index, double* host) {
    if (!st || !host || index < 0 || index >= st->noutputs)
        return fail(ALSVINN_HIP_E_INVALID, "bad output index");
    HIP_TRY(hipMemcpyAsync(host, st->out[index], sizeof(double) * st->elements, hipMemcpyDeviceToHost,
                           st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return 0;
}

}  // extern "C"

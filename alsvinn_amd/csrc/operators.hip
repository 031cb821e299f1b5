// Host side of the operators on a reference Volume (include/alsvinn_hip.h):
// NumericalFlux::computeFlux, Reconstruction::performReconstruction and
// Boundary::applyBoundaryConditions on device memories with one area per
// conserved variable (alsfvm/src/volume/Volume.cpp:44-60). computeFlux runs
// the ensemble's fused first-stage kernels (reconstruction + face flux +
// divergence + wave speeds in one launch) over the work list of the region;
// nothing is copied to the host except the three wave speeds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/alsvinn_hip.h"
#include "ensemble_kernels.hpp"
#include "internal.hpp"

using namespace alsvinn_hip;

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return fail(ALSVINN_HIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct alsvinn_hip_operator {
    alsvinn_hip_scheme scheme;
    int nvar = 0, dim = 0, ghosts = 0;
    int n[3] = {1, 1, 1};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    SampleDesc* d_desc = nullptr;
    SampleState* d_state = nullptr;
    unsigned long long* d_ws = nullptr;
    unsigned long long* h_ws = nullptr;  // pinned
    SampleDesc desc{};                   // the last volume's geometry on the device
    bool desc_valid = false;
    // work lists of the regions seen so far (ConservedSystem uses 1 + 2d)
    std::map<std::array<int, 6>, std::pair<WaveTask*, int>> tasks;
    KernelSet<double> k64;
    KernelSet<float> k32;
};

namespace {

// Clears the wave-speed slots (a kernel rather than a memset, so that a
// caller's stream capture records no memset node).
__global__ void clear_u64_kernel(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

bool axis_active(const alsvinn_hip_volume& v, int a) {
    return a == 0 || (a == 1 ? v.ny > 1 : v.nz > 1);
}

// Shape checks of a volume against the operator's grid and scheme.
int check_volume(const alsvinn_hip_volume* v, int nvar, int precision, const int n[3], int G,
                 const char* what) {
    const std::string w(what);
    if (!v) return fail(ALSVINN_HIP_E_INVALID, w + " volume is NULL");
    if (v->nvar != nvar)
        return fail(ALSVINN_HIP_E_INVALID, w + " volume has " + std::to_string(v->nvar) +
                                               " variables, the equation " + std::to_string(nvar));
    if (v->precision != precision) return fail(ALSVINN_HIP_E_INVALID, w + " volume precision differs");
    const int vn[3] = {v->nx, v->ny, v->nz};
    const int vg[3] = {v->gx, v->gy, v->gz};
    for (int a = 0; a < 3; ++a) {
        if (n && vn[a] != n[a]) return fail(ALSVINN_HIP_E_INVALID, w + " volume does not match the grid");
        if (vn[a] < 1 || vg[a] < 0) return fail(ALSVINN_HIP_E_INVALID, w + " volume extents");
        if (axis_active(*v, a) ? vg[a] < G : vg[a] != 0)
            return fail(ALSVINN_HIP_E_INVALID, w + " volume needs " + std::to_string(G) +
                                                   " ghost layers on its active axes and none on the others");
    }
    for (int k = 0; k < nvar; ++k)
        if (!v->var[k]) return fail(ALSVINN_HIP_E_INVALID, w + " volume variable pointer is NULL");
    return 0;
}

SampleDesc volume_desc(const alsvinn_hip_volume& v, int bc) {
    SampleDesc D;
    std::memset(&D, 0, sizeof(D));
    D.nx = v.nx;
    D.ny = v.ny;
    D.nz = v.nz;
    D.gx = v.gx;
    D.gy = v.gy;
    D.gz = v.gz;
    D.tx = v.nx + 2 * v.gx;
    D.ty = v.ny + 2 * v.gy;
    D.tz = v.nz + 2 * v.gz;
    D.vstride = (long long)D.tx * D.ty * D.tz;
    D.bc = bc;
    D.dx[0] = D.dx[1] = D.dx[2] = 1;
    D.end_time = 1;
    D.max_steps = -1;
    return D;
}

// Work list of the interior region [lo, hi) of one volume: x-strips of
// W = 64 - 2G columns, bands of rows (2D) or groups of kGroupRows rows and
// bands of planes (3D), about 4096 waves for the whole grid.
std::vector<WaveTask> region_tasks(int G, int dim, const int lo[3], const int hi[3]) {
    std::vector<WaveTask> out;
    const int W = 64 - 2 * G;
    const int width = hi[0] - lo[0], rows = hi[1] - lo[1], planes = hi[2] - lo[2];
    if (width <= 0 || rows <= 0 || planes <= 0) return out;
    const long long strips = (width + W - 1) / W;
    const long long target = 4096;
    auto band = [&](int extent, long long per_band) {
        int t = extent;
        while (t > 8 && per_band * ((extent + t - 1) / t) < target) t = (t + 1) / 2;
        return t;
    };
    auto add = [&](int y0, int y1, int z0, int z1) {
        for (int x0 = lo[0]; x0 < hi[0]; x0 += W) {
            WaveTask t;
            std::memset(&t, 0, sizeof(t));
            t.nseg = 1;
            t.y0 = y0;
            t.y1 = y1;
            t.z0 = z0;
            t.z1 = z1;
            t.sample[0] = 0;
            t.x0[0] = x0;
            t.ncols[0] = std::min(W, hi[0] - x0);
            out.push_back(t);
        }
    };
    if (dim == 3) {
        const long long groups = (rows + kGroupRows - 1) / kGroupRows;
        const int tz = band(planes, strips * groups * kGroupRows);
        for (int z0 = lo[2]; z0 < hi[2]; z0 += tz)
            for (int y = lo[1]; y < hi[1]; y += kGroupRows)
                add(y, std::min(hi[1], y + kGroupRows), z0, std::min(hi[2], z0 + tz));
    } else if (dim == 2) {
        const int ty = band(rows, strips);
        for (int y0 = lo[1]; y0 < hi[1]; y0 += ty) add(y0, std::min(hi[1], y0 + ty), 0, 1);
    } else {
        add(0, 1, 0, 1);
    }
    return out;
}

template <class T>
int numflux_compute(alsvinn_hip_operator* op, const KernelSet<T>& k, const alsvinn_hip_volume& U,
                    const alsvinn_hip_volume& O, const int lo[3], const int hi[3], double ws[3]) {
    const SampleDesc D = volume_desc(U, op->scheme.boundary);
    if (!op->desc_valid || std::memcmp(&D, &op->desc, sizeof(D)) != 0) {
        HIP_TRY(hipMemcpyAsync(op->d_desc, &D, sizeof(D), hipMemcpyHostToDevice, op->stream));
        HIP_TRY(hipStreamSynchronize(op->stream));  // D lives on the host stack
        op->desc = D;
        op->desc_valid = true;
    }
    const std::array<int, 6> key = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
    auto it = op->tasks.find(key);
    if (it == op->tasks.end()) {
        const std::vector<WaveTask> t = region_tasks(op->ghosts, op->dim, lo, hi);
        WaveTask* d = nullptr;
        if (!t.empty()) {
            HIP_TRY(hipMalloc(&d, sizeof(WaveTask) * t.size()));
            HIP_TRY(hipMemcpy(d, t.data(), sizeof(WaveTask) * t.size(), hipMemcpyHostToDevice));
        }
        it = op->tasks.emplace(key, std::make_pair(d, (int)t.size())).first;
    }
    hipLaunchKernelGGL(clear_u64_kernel, dim3(1), dim3(64), 0, op->stream, op->d_ws, 3);
    StageArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.desc = op->d_desc;
    a.state = op->d_state;
    a.ws = op->d_ws;
    a.tasks = it->second.first;
    a.ntasks = it->second.second;
    a.mode = MODE_RHS;
    a.ws_filter = 1;  // one sample, any number of wavefronts
    a.gamma = op->scheme.gamma;
    for (int v = 0; v < op->nvar; ++v) {
        a.vin[v] = (const T*)U.var[v];
        a.vout[v] = (T*)O.var[v];
    }
    a.in = a.in2 = a.vin[0];
    a.out = a.vout[0];
    k.op_flux(a, (a.ntasks + 3) / 4, op->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(op->h_ws, op->d_ws, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           op->stream));
    HIP_TRY(hipStreamSynchronize(op->stream));
    for (int d = 0; d < 3; ++d) std::memcpy(&ws[d], &op->h_ws[d], sizeof(double));
    return 0;
}

template <class T>
void recon_launch(const alsvinn_hip_operator* op, const KernelSet<T>& k,
                  const alsvinn_hip_volume& in, const alsvinn_hip_volume& L,
                  const alsvinn_hip_volume& R, int direction, const int lo[3], const int hi[3]) {
    ReconArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    for (int v = 0; v < op->nvar; ++v) {
        a.vin[v] = (const T*)in.var[v];
        a.vleft[v] = (T*)L.var[v];
        a.vright[v] = (T*)R.var[v];
    }
    a.tx = in.nx + 2 * in.gx;
    a.ty = in.ny + 2 * in.gy;
    a.tz = in.nz + 2 * in.gz;
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = lo[d];
        a.hi[d] = hi[d];
    }
    a.direction = direction;
    a.gamma = op->scheme.gamma;
    const long long n = (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
    k.op_recon(a, blocks, op->stream);
}

// ---- allocation-free computeFlux: every piece of device memory the call
// needs beyond its input and output volumes (the volume's descriptor and
// run state, three wave-speed slots and the region's work list) lives in
// this file-scope array, one copy per GPU, held until the process ends.
// Host code never uses the array's name as an address: it takes the
// device's pointer from hipGetSymbolAddress (noalloc_scratch_pointer).
constexpr size_t kScratchBytes = size_t(1) << 20;
constexpr size_t kScratchDesc = 0, kScratchState = 256, kScratchWs = 512, kScratchTasks = 1024;
constexpr int kScratchMaxTasks = int((kScratchBytes - kScratchTasks) / sizeof(WaveTask));
static_assert(sizeof(SampleDesc) <= kScratchState - kScratchDesc, "descriptor slot");
static_assert(sizeof(SampleState) <= kScratchWs - kScratchState, "state slot");
static_assert(kScratchTasks % alignof(WaveTask) == 0, "work list alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of the scratch array, looked up once per device.
int noalloc_scratch_pointer(unsigned char** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ALSVINN_HIP_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Band heights of region_tasks (rows per 2D task, planes per 3D task).
int region_band(int G, int dim, const int lo[3], const int hi[3]) {
    const int W = 64 - 2 * G;
    const long long strips = (hi[0] - lo[0] + W - 1) / W;
    const int rows = hi[1] - lo[1], planes = hi[2] - lo[2];
    auto band = [&](int extent, long long per_band) {
        int t = extent;
        while (t > 8 && per_band * ((extent + t - 1) / t) < 4096) t = (t + 1) / 2;
        return t;
    };
    if (dim == 3) return band(planes, strips * ((rows + kGroupRows - 1) / kGroupRows) * kGroupRows);
    if (dim == 2) return band(rows, strips);
    return 1;
}

struct RegionArgs {
    int lo[3], hi[3];
    int G, dim, band, ntasks;
};

// Writes the descriptor, the run state and cleared wave-speed slots, and the
// region's work list (the tasks of region_tasks, in its order: planes or rows
// outer, x-strips inner) into the scratch memory at `base`.
__global__ __launch_bounds__(256) void noalloc_setup_kernel(unsigned char* base, SampleDesc D,
                                                           RegionArgs r) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *(SampleDesc*)(base + kScratchDesc) = D;
        SampleState S;
        S.t = 0;
        S.last_dt = 0;
        S.steps = 0;
        S.saved[0] = S.saved[1] = 0;
        S.active = 1;
        S.status = 0;
        S.hold[0] = S.hold[1] = LLONG_MAX;
        *(SampleState*)(base + kScratchState) = S;
        unsigned long long* ws = (unsigned long long*)(base + kScratchWs);
        ws[0] = ws[1] = ws[2] = 0;
    }
    if (t >= r.ntasks) return;
    const int W = 64 - 2 * r.G;
    const int strips = (r.hi[0] - r.lo[0] + W - 1) / W;
    const int xs = t % strips, outer = t / strips;
    WaveTask k;
    k.nseg = 1;
    k.pad[0] = k.pad[1] = k.pad[2] = 0;
    for (int j = 0; j < kMaxSeg; ++j) k.sample[j] = k.x0[j] = k.ncols[j] = k.lane0[j] = 0;
    k.x0[0] = r.lo[0] + xs * W;
    k.ncols[0] = min(W, r.hi[0] - k.x0[0]);
    if (r.dim == 3) {
        const int groups = (r.hi[1] - r.lo[1] + kGroupRows - 1) / kGroupRows;
        const int yg = outer % groups, zb = outer / groups;
        k.y0 = r.lo[1] + yg * kGroupRows;
        k.y1 = min(r.hi[1], k.y0 + kGroupRows);
        k.z0 = r.lo[2] + zb * r.band;
        k.z1 = min(r.hi[2], k.z0 + r.band);
    } else if (r.dim == 2) {
        k.y0 = r.lo[1] + outer * r.band;
        k.y1 = min(r.hi[1], k.y0 + r.band);
        k.z0 = 0;
        k.z1 = 1;
    } else {
        k.y0 = 0;
        k.y1 = 1;
        k.z0 = 0;
        k.z1 = 1;
    }
    ((WaveTask*)(base + kScratchTasks))[t] = k;
}

// Wave speeds of the active directions from the scratch slots (bit patterns
// of non-negative doubles) into the caller's device array.
__global__ void noalloc_ws_kernel(const unsigned char* base, double* out, int dim) {
    const int d = threadIdx.x;
    if (d < dim) out[d] = __longlong_as_double((long long)((const unsigned long long*)(base + kScratchWs))[d]);
}

}  // namespace

extern "C" {

int alsvinn_hip_operator_create(const alsvinn_hip_scheme* scheme, int nx, int ny, int nz,
                                void* stream, alsvinn_hip_operator** out) {
    if (!out) return fail(ALSVINN_HIP_E_INVALID, "out is NULL");
    *out = nullptr;
    if (int rc = validate(scheme)) return rc;
    if (nx < 1 || ny < 1 || nz < 1 || (nz > 1 && ny < 2))
        return fail(ALSVINN_HIP_E_INVALID, "bad grid extents");
    auto* op = new alsvinn_hip_operator();
    op->scheme = *scheme;
    op->n[0] = nx;
    op->n[1] = ny;
    op->n[2] = nz;
    op->dim = 1 + (ny > 1) + (nz > 1);
    op->nvar = scheme->equation == ALSVINN_HIP_EQ_BURGERS ? 1 : scheme->equation + 2;
    op->ghosts = ghost_cells(*scheme);
    const int G = op->ghosts;
    if (nx < G || (ny > 1 && ny < G) || (nz > 1 && nz < G)) {
        delete op;
        return fail(ALSVINN_HIP_E_INVALID, "grid smaller than the number of ghost cells");
    }
    int rc = scheme->precision == ALSVINN_HIP_FP64 ? select_kernels<double>(*scheme, op->dim, op->k64)
                                                   : select_kernels<float>(*scheme, op->dim, op->k32);
    if (rc) {
        delete op;
        return rc;
    }
    auto cleanup = [&](hipError_t e, const char* what) {
        alsvinn_hip_operator_destroy(op);
        return fail(ALSVINN_HIP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    if (stream) {
        op->stream = (hipStream_t)stream;
    } else {
        if ((e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking)) != hipSuccess)
            return cleanup(e, "hipStreamCreateWithFlags");
        op->own_stream = true;
    }
    if ((e = hipMalloc(&op->d_desc, sizeof(SampleDesc))) != hipSuccess) return cleanup(e, "hipMalloc");
    if ((e = hipMalloc(&op->d_state, sizeof(SampleState))) != hipSuccess) return cleanup(e, "hipMalloc");
    if ((e = hipMalloc(&op->d_ws, 3 * sizeof(unsigned long long))) != hipSuccess)
        return cleanup(e, "hipMalloc");
    if ((e = hipHostMalloc(&op->h_ws, 3 * sizeof(unsigned long long))) != hipSuccess)
        return cleanup(e, "hipHostMalloc");
    SampleState S;
    std::memset(&S, 0, sizeof(S));
    S.active = 1;
    S.hold[0] = S.hold[1] = LLONG_MAX;
    if ((e = hipMemcpy(op->d_state, &S, sizeof(S), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(e, "hipMemcpy");
    *out = op;
    return 0;
}

int alsvinn_hip_operator_destroy(alsvinn_hip_operator* op) {
    if (!op) return 0;
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    for (auto& kv : op->tasks) (void)hipFree(kv.second.first);
    (void)hipFree(op->d_desc);
    (void)hipFree(op->d_state);
    (void)hipFree(op->d_ws);
    if (op->h_ws) (void)hipHostFree(op->h_ws);
    if (op->own_stream && op->stream) (void)hipStreamDestroy(op->stream);
    delete op;
    return 0;
}

int alsvinn_hip_numflux_compute(alsvinn_hip_operator* op, const alsvinn_hip_volume* conserved,
                                double wave_speeds[3], int compute_wave_speed,
                                const alsvinn_hip_volume* output, const int start[3],
                                const int end[3]) {
    (void)compute_wave_speed;  // the wave speeds are always computed, as in the reference
    if (!op) return fail(ALSVINN_HIP_E_INVALID, "operator is NULL");
    if (int rc = check_volume(conserved, op->nvar, op->scheme.precision, op->n, op->ghosts, "input"))
        return rc;
    if (int rc = check_volume(output, op->nvar, op->scheme.precision, op->n, op->ghosts, "output"))
        return rc;
    if (output->gx != conserved->gx || output->gy != conserved->gy || output->gz != conserved->gz)
        return fail(ALSVINN_HIP_E_INVALID, "input and output volumes need the same ghost layers");
    for (int v = 0; v < op->nvar; ++v)
        for (int w = 0; w < op->nvar; ++w)
            if (output->var[v] == conserved->var[w])
                return fail(ALSVINN_HIP_E_INVALID, "output must not alias the input");
    int lo[3] = {0, 0, 0}, hi[3] = {op->n[0], op->n[1], op->n[2]};
    for (int a = 0; a < 3; ++a) {
        const int s = start ? start[a] : 0, t = end ? end[a] : 0;
        if (a >= op->dim) {
            if (s != 0 || t != 0) return fail(ALSVINN_HIP_E_INVALID, "start/end on an inactive axis");
            continue;
        }
        if (s < 0 || t > 0) return fail(ALSVINN_HIP_E_INVALID, "start must be >= 0 and end <= 0");
        lo[a] = s;
        hi[a] = op->n[a] + t;
    }
    double ws[3] = {0, 0, 0};
    int rc = op->scheme.precision == ALSVINN_HIP_FP64
                 ? numflux_compute<double>(op, op->k64, *conserved, *output, lo, hi, ws)
                 : numflux_compute<float>(op, op->k32, *conserved, *output, lo, hi, ws);
    if (rc) return rc;
    if (wave_speeds)
        for (int d = 0; d < op->dim; ++d) wave_speeds[d] = ws[d];
    return 0;
}

int alsvinn_hip_reconstruction_compute(alsvinn_hip_operator* op, const alsvinn_hip_volume* input,
                                       int direction, int indicator_variable,
                                       const alsvinn_hip_volume* left,
                                       const alsvinn_hip_volume* right, const int start[3],
                                       const int end[3]) {
    if (!op) return fail(ALSVINN_HIP_E_INVALID, "operator is NULL");
    if (direction < 0 || direction >= op->dim)
        return fail(ALSVINN_HIP_E_INVALID, "Direction can only be 0, 1 or 2 (an active axis), was given: " +
                                               std::to_string(direction));
    if (indicator_variable != 0)
        return fail(ALSVINN_HIP_E_INVALID, "the hip reconstructions select their stencils from variable 0");
    const int recon = op->scheme.reconstruction;
    int G = 2;  // the reconstruction's own ghost count (ReconstructionFactory.cpp:59-160)
    if (recon == ALSVINN_HIP_RECON_NONE) G = 1;
    if (recon == ALSVINN_HIP_RECON_ENO3 || recon == ALSVINN_HIP_RECON_WENO3) G = 3;
    if (recon == ALSVINN_HIP_RECON_ENO4) G = 4;
    for (auto* v : {input, left, right})
        if (int rc = check_volume(v, op->nvar, op->scheme.precision, op->n, G,
                                  v == input ? "input" : (v == left ? "left" : "right")))
            return rc;
    if (left->gx != input->gx || left->gy != input->gy || left->gz != input->gz ||
        right->gx != input->gx || right->gy != input->gy || right->gz != input->gz)
        return fail(ALSVINN_HIP_E_INVALID, "input, left and right volumes need the same ghost layers");
    const int tot[3] = {input->nx + 2 * input->gx, input->ny + 2 * input->gy,
                        input->nz + 2 * input->gz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const int s = start ? start[a] : 0, t = end ? end[a] : 0;
        const int dv = a == direction;
        if (recon == ALSVINN_HIP_RECON_NONE) {  // NoReconstruction: every cell
            lo[a] = 0;
            hi[a] = tot[a];
        } else if (recon == ALSVINN_HIP_RECON_WENO3) {  // WENOCPU: start/end unused
            lo[a] = dv * 2;
            hi[a] = tot[a] - dv * 2;
        } else {  // ReconstructionCPU (ghost 2), ENOCPU (ghost = order)
            const int act = axis_active(*input, a) ? 1 : 0;
            lo[a] = act * (G - dv) + s;
            hi[a] = tot[a] - act * (G - dv) + t;
        }
    }
    const int RAD = op->scheme.precision == ALSVINN_HIP_FP64 ? op->k64.recon_radius : op->k32.recon_radius;
    for (int a = 0; a < 3; ++a) {
        if (lo[a] >= hi[a]) return 0;  // empty range: nothing to do
        const int r = (a == direction && recon != ALSVINN_HIP_RECON_NONE) ? RAD : 0;
        if (lo[a] - r < 0 || hi[a] + r > tot[a])
            return fail(ALSVINN_HIP_E_INVALID, "the reconstruction's range leaves the volume (start/end)");
    }
    if (op->scheme.precision == ALSVINN_HIP_FP64)
        recon_launch<double>(op, op->k64, *input, *left, *right, direction, lo, hi);
    else
        recon_launch<float>(op, op->k32, *input, *left, *right, direction, lo, hi);
    HIP_TRY(hipGetLastError());
    return 0;
}

int alsvinn_hip_boundary_apply(int boundary, const alsvinn_hip_volume* volume, void* stream) {
    if (boundary != ALSVINN_HIP_BC_NEUMANN && boundary != ALSVINN_HIP_BC_PERIODIC)
        return fail(ALSVINN_HIP_E_INVALID, "unknown boundary " + std::to_string(boundary));
    if (!volume || volume->nvar < 1 || volume->nvar > 5)
        return fail(ALSVINN_HIP_E_INVALID, "bad volume");
    if (volume->precision != ALSVINN_HIP_FP64 && volume->precision != ALSVINN_HIP_FP32)
        return fail(ALSVINN_HIP_E_INVALID, "unknown precision");
    if (int rc = check_volume(volume, volume->nvar, volume->precision, nullptr, 0, "boundary"))
        return rc;
    const int n[3] = {volume->nx, volume->ny, volume->nz};
    const int g[3] = {volume->gx, volume->gy, volume->gz};
    for (int a = 0; a < 3; ++a)
        if (g[a] > n[a])
            return fail(ALSVINN_HIP_E_INVALID, "more ghost layers than interior cells on an axis");
    const SampleDesc D = volume_desc(*volume, boundary);
    const int blocks = (int)std::max<long long>(1, std::min<long long>((D.vstride + 255) / 256, 4096));
    hipStream_t s = (hipStream_t)stream;
    if (volume->precision == ALSVINN_HIP_FP64) {
        PerVarVol<double> u;
        for (int v = 0; v < 5; ++v) u.p[v] = v < volume->nvar ? (double*)volume->var[v] : nullptr;
        hipLaunchKernelGGL(ghost_fill_volume_kernel<double>, dim3(blocks), dim3(256), 0, s, u,
                           volume->nvar, D);
    } else {
        PerVarVol<float> u;
        for (int v = 0; v < 5; ++v) u.p[v] = v < volume->nvar ? (float*)volume->var[v] : nullptr;
        hipLaunchKernelGGL(ghost_fill_volume_kernel<float>, dim3(blocks), dim3(256), 0, s, u,
                           volume->nvar, D);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int alsvinn_hip_numflux_compute_noalloc(const alsvinn_hip_scheme* scheme,
                                        const alsvinn_hip_volume* conserved,
                                        const alsvinn_hip_volume* output, const int start[3],
                                        const int end[3], double* device_wave_speeds,
                                        void* stream) {
    if (int rc = validate(scheme)) return rc;
    if (!conserved) return fail(ALSVINN_HIP_E_INVALID, "input volume is NULL");
    const int n[3] = {conserved->nx, conserved->ny, conserved->nz};
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || (n[2] > 1 && n[1] < 2))
        return fail(ALSVINN_HIP_E_INVALID, "bad grid extents");
    const int dim = 1 + (n[1] > 1) + (n[2] > 1);
    const int nvar = scheme->equation == ALSVINN_HIP_EQ_BURGERS ? 1 : scheme->equation + 2;
    const int G = ghost_cells(*scheme);
    if (n[0] < G || (n[1] > 1 && n[1] < G) || (n[2] > 1 && n[2] < G))
        return fail(ALSVINN_HIP_E_INVALID, "grid smaller than the number of ghost cells");
    if (int rc = check_volume(conserved, nvar, scheme->precision, n, G, "input")) return rc;
    if (int rc = check_volume(output, nvar, scheme->precision, n, G, "output")) return rc;
    if (output->gx != conserved->gx || output->gy != conserved->gy || output->gz != conserved->gz)
        return fail(ALSVINN_HIP_E_INVALID, "input and output volumes need the same ghost layers");
    for (int v = 0; v < nvar; ++v)
        for (int w = 0; w < nvar; ++w)
            if (output->var[v] == conserved->var[w])
                return fail(ALSVINN_HIP_E_INVALID, "output must not alias the input");
    RegionArgs r;
    std::memset(&r, 0, sizeof(r));
    for (int a = 0; a < 3; ++a) {
        const int s = start ? start[a] : 0, t = end ? end[a] : 0;
        r.lo[a] = 0;
        r.hi[a] = n[a];
        if (a >= dim) {
            if (s != 0 || t != 0) return fail(ALSVINN_HIP_E_INVALID, "start/end on an inactive axis");
            continue;
        }
        if (s < 0 || t > 0) return fail(ALSVINN_HIP_E_INVALID, "start must be >= 0 and end <= 0");
        r.lo[a] = s;
        r.hi[a] = n[a] + t;
    }
    r.G = G;
    r.dim = dim;
    const bool empty = r.hi[0] <= r.lo[0] || r.hi[1] <= r.lo[1] || r.hi[2] <= r.lo[2];
    if (!empty) {
        r.band = region_band(G, dim, r.lo, r.hi);
        const long long W = 64 - 2 * G;
        const long long strips = (r.hi[0] - r.lo[0] + W - 1) / W;
        long long outer = 1;
        if (dim == 3)
            outer = (long long)((r.hi[1] - r.lo[1] + kGroupRows - 1) / kGroupRows) *
                    ((r.hi[2] - r.lo[2] + r.band - 1) / r.band);
        else if (dim == 2)
            outer = (r.hi[1] - r.lo[1] + r.band - 1) / r.band;
        const long long ntasks = strips * outer;
        if (ntasks > kScratchMaxTasks)
            return fail(ALSVINN_HIP_E_INVALID,
                        "the region needs " + std::to_string(ntasks) +
                            " work items; the allocation-free computeFlux holds at most " +
                            std::to_string(kScratchMaxTasks) +
                            " (1 MiB of device scratch): use alsvinn_hip_numflux_compute");
        r.ntasks = (int)ntasks;
    }
    KernelSet<double> k64;
    KernelSet<float> k32;
    const int rc = scheme->precision == ALSVINN_HIP_FP64 ? select_kernels<double>(*scheme, dim, k64)
                                                         : select_kernels<float>(*scheme, dim, k32);
    if (rc) return rc;
    unsigned char* base = nullptr;
    if (int e = noalloc_scratch_pointer(&base)) return e;
    hipStream_t s = (hipStream_t)stream;
    const SampleDesc D = volume_desc(*conserved, scheme->boundary);
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(std::max(1, (r.ntasks + 255) / 256)), dim3(256), 0, s,
                       base, D, r);
    HIP_TRY(hipGetLastError());
    if (!empty) {
        if (scheme->precision == ALSVINN_HIP_FP64) {
            StageArgs<double> a;
            std::memset(&a, 0, sizeof(a));
            a.desc = (const SampleDesc*)(base + kScratchDesc);
            a.state = (SampleState*)(base + kScratchState);
            a.ws = (unsigned long long*)(base + kScratchWs);
            a.tasks = (const WaveTask*)(base + kScratchTasks);
            a.ntasks = r.ntasks;
            a.mode = MODE_RHS;
            a.ws_filter = 1;  // one sample, any number of wavefronts
            a.gamma = scheme->gamma;
            for (int v = 0; v < nvar; ++v) {
                a.vin[v] = (const double*)conserved->var[v];
                a.vout[v] = (double*)output->var[v];
            }
            a.in = a.in2 = a.vin[0];
            a.out = a.vout[0];
            k64.op_flux(a, (a.ntasks + 3) / 4, s);
        } else {
            StageArgs<float> a;
            std::memset(&a, 0, sizeof(a));
            a.desc = (const SampleDesc*)(base + kScratchDesc);
            a.state = (SampleState*)(base + kScratchState);
            a.ws = (unsigned long long*)(base + kScratchWs);
            a.tasks = (const WaveTask*)(base + kScratchTasks);
            a.ntasks = r.ntasks;
            a.mode = MODE_RHS;
            a.ws_filter = 1;  // one sample, any number of wavefronts
            a.gamma = scheme->gamma;
            for (int v = 0; v < nvar; ++v) {
                a.vin[v] = (const float*)conserved->var[v];
                a.vout[v] = (float*)output->var[v];
            }
            a.in = a.in2 = a.vin[0];
            a.out = a.vout[0];
            k32.op_flux(a, (a.ntasks + 3) / 4, s);
        }
        HIP_TRY(hipGetLastError());
    }
    if (device_wave_speeds) {
        hipLaunchKernelGGL(noalloc_ws_kernel, dim3(1), dim3(64), 0, s, base, device_wave_speeds, dim);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

}  // extern "C"

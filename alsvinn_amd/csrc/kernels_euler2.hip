// Kernel instantiations for the 2D Euler equations (HLL, HLL3).
#include "internal.hpp"

// Built twice (alsvinn_amd/build.py): once with ALSVINN_HIP_FP64_ONLY for the
// FP64-state kernels and once with ALSVINN_HIP_FP32_ONLY for the FP32-state ones.

namespace alsvinn_hip {

template <class T>
int select_euler2(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k) {
    if (dim != 2) return fail(ALSVINN_HIP_E_INVALID, "euler2 needs a 2D grid");
    if (sc.flux == ALSVINN_HIP_FLUX_HLL) return select_recon<T, EulerEq<T, 2>, FluxHLL, 2>(sc, k);
    if (sc.flux == ALSVINN_HIP_FLUX_CENTRAL) return select_recon<T, EulerEq<T, 2>, FluxCentral, 2>(sc, k);
    return select_recon<T, EulerEq<T, 2>, FluxHLL3, 2>(sc, k);
}

#ifndef ALSVINN_HIP_FP32_ONLY
template int select_euler2<double>(const alsvinn_hip_scheme&, int, KernelSet<double>&);
#endif
#ifndef ALSVINN_HIP_FP64_ONLY
template int select_euler2<float>(const alsvinn_hip_scheme&, int, KernelSet<float>&);
#endif

}  // namespace alsvinn_hip

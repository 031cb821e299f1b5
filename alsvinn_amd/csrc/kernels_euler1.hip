// Kernel instantiations for the 1D Euler equations (HLL, HLL3).
#include "internal.hpp"

namespace alsvinn_hip {

template <class T>
int select_euler1(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k) {
    if (dim != 1) return fail(ALSVINN_HIP_E_INVALID, "euler1 needs a 1D grid");
    if (sc.flux == ALSVINN_HIP_FLUX_HLL) return select_recon<T, EulerEq<T, 1>, FluxHLL, 1>(sc, k);
    if (sc.flux == ALSVINN_HIP_FLUX_CENTRAL) return select_recon<T, EulerEq<T, 1>, FluxCentral, 1>(sc, k);
    return select_recon<T, EulerEq<T, 1>, FluxHLL3, 1>(sc, k);
}

template int select_euler1<double>(const alsvinn_hip_scheme&, int, KernelSet<double>&);
template int select_euler1<float>(const alsvinn_hip_scheme&, int, KernelSet<float>&);

}  // namespace alsvinn_hip

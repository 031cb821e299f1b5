// Kernel instantiations for the Burgers equation (Godunov flux), 1D and 2D.
#include "internal.hpp"

// Built twice (alsvinn_amd/build.py): once with ALSVINN_HIP_FP64_ONLY for the
// FP64-state kernels and once with ALSVINN_HIP_FP32_ONLY for the FP32-state ones.

namespace alsvinn_hip {

template <class T>
int select_burgers(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k) {
    if (dim != 1 && dim != 2)
        return fail(ALSVINN_HIP_E_INVALID, "burgers on the hip platform supports 1D and 2D grids");
    switch (sc.flux) {
        case ALSVINN_HIP_FLUX_GODUNOV:
            return dim == 1 ? select_recon<T, BurgersEq<T>, FluxGodunov, 1>(sc, k)
                            : select_recon<T, BurgersEq<T>, FluxGodunov, 2>(sc, k);
        case ALSVINN_HIP_FLUX_CENTRAL:
            return dim == 1 ? select_recon<T, BurgersEq<T>, FluxCentral, 1>(sc, k)
                            : select_recon<T, BurgersEq<T>, FluxCentral, 2>(sc, k);
        case ALSVINN_HIP_FLUX_RUSANOV:
            return dim == 1 ? select_recon<T, BurgersEq<T>, FluxRusanov, 1>(sc, k)
                            : select_recon<T, BurgersEq<T>, FluxRusanov, 2>(sc, k);
        case ALSVINN_HIP_FLUX_ROE:
            return dim == 1 ? select_recon<T, BurgersEq<T>, FluxRoe, 1>(sc, k)
                            : select_recon<T, BurgersEq<T>, FluxRoe, 2>(sc, k);
    }
    return fail(ALSVINN_HIP_E_INVALID, "flux not available for burgers");
}

#ifndef ALSVINN_HIP_FP32_ONLY
template int select_burgers<double>(const alsvinn_hip_scheme&, int, KernelSet<double>&);
#endif
#ifndef ALSVINN_HIP_FP64_ONLY
template int select_burgers<float>(const alsvinn_hip_scheme&, int, KernelSet<float>&);
#endif

}  // namespace alsvinn_hip

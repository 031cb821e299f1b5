// Kernel instantiations for the Burgers equation (Godunov flux), 1D and 2D.
#include "internal.hpp"

namespace alsvinn_hip {

template <class T>
int select_burgers(const alsvinn_hip_scheme& sc, int dim, KernelSet<T>& k) {
    if (dim == 1) return select_recon<T, BurgersEq<T>, FluxGodunov, 1>(sc, k);
    if (dim == 2) return select_recon<T, BurgersEq<T>, FluxGodunov, 2>(sc, k);
    return fail(ALSVINN_HIP_E_INVALID, "burgers on the hip platform supports 1D and 2D grids");
}

template int select_burgers<double>(const alsvinn_hip_scheme&, int, KernelSet<double>&);
template int select_burgers<float>(const alsvinn_hip_scheme&, int, KernelSet<float>&);

}  // namespace alsvinn_hip

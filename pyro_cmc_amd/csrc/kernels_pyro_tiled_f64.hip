// Instantiation unit: pyrochlore-specialised tiled sweep kernels, double.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_tiled<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, int, int, int, unsigned long long, hipStream_t, bool);
template hipError_t launch_pyro_energy_tiled<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, double*, int*, hipStream_t);
}  // namespace pyro

// Instantiation unit: pyrochlore-specialised tiled sweep kernels, float.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_tiled<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, int, int, int, unsigned long long, hipStream_t, bool);
template hipError_t launch_pyro_energy_tiled<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, double*, int*, hipStream_t);
}  // namespace pyro

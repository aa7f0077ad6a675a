// Instantiation unit: pyrochlore-specialised colour sweep kernels, float.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_colour<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, int, unsigned long long, hipStream_t, double*, bool, int);
template hipError_t launch_pyro_energy<float>(const DeviceState<float>&, const PyroCouplings<float>&, double*, hipStream_t);
template hipError_t launch_pyro_fields<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, float*, hipStream_t);
}  // namespace pyro

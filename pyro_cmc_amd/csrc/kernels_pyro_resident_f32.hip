// Instantiation unit: pyrochlore-specialised resident and persistent sweep kernels, float.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_resident<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, int, int, unsigned long long, hipStream_t);
template hipError_t launch_pyro_persist<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, int, int, unsigned long long, unsigned int*, unsigned int*, int, hipStream_t);
}  // namespace pyro

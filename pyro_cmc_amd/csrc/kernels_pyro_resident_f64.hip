// Instantiation unit: pyrochlore-specialised resident and persistent sweep kernels, double.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_resident<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, int, int, unsigned long long, hipStream_t);
template hipError_t launch_pyro_persist<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, int, int, unsigned long long, unsigned int*, unsigned int*, int, hipStream_t);
}  // namespace pyro

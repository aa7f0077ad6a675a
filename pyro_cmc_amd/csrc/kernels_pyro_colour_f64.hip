// Instantiation unit: pyrochlore-specialised colour sweep kernels, double.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_colour<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, int, unsigned long long, hipStream_t, double*, bool, int);
template hipError_t launch_pyro_energy<double>(const DeviceState<double>&, const PyroCouplings<double>&, double*, hipStream_t);
template hipError_t launch_pyro_fields<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, double*, hipStream_t);
}  // namespace pyro

// Instantiation unit: pyrochlore-specialised resident sweep kernels, double.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_resident<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, int, int, unsigned long long, hipStream_t, double*, bool);
}  // namespace pyro

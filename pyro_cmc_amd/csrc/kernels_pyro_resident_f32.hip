// Instantiation unit: pyrochlore-specialised resident sweep kernels, float.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_resident<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, int, int, unsigned long long, hipStream_t, double*, bool);
}  // namespace pyro

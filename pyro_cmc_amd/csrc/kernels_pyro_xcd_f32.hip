// Instantiation unit: XCD-resident persistent sweep kernels, float.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_xcd<float>(const DeviceState<float>&, const PyroCouplings<float>&, int, int, int, unsigned long long, hipStream_t, double*, bool, unsigned int*, int*, unsigned long long);
}  // namespace pyro

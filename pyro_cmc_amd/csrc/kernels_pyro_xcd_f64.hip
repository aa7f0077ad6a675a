// Instantiation unit: XCD-resident persistent sweep kernels, double.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_xcd<double>(const DeviceState<double>&, const PyroCouplings<double>&, int, int, int, unsigned long long, hipStream_t, double*, bool, unsigned int*, int*, unsigned long long);
}  // namespace pyro

// Instantiation unit: persistent batch sweep kernels, double.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_batch<double>(const BatchMember<double>*, const int4*, int, int, int, bool, int, int, int, bool, bool, unsigned int*, hipStream_t);
template int pyro_batch_capacity<double>(int, bool, int, bool, bool);
}  // namespace pyro

// Instantiation unit: persistent batch sweep kernels, float.
#include "kernels_pyro_impl.hpp"

namespace pyro {
template hipError_t launch_pyro_batch<float>(const BatchMember<float>*, const int4*, int, int, int, bool, int, int, int, bool, bool, unsigned int*, hipStream_t);
template int pyro_batch_capacity<float>(int, bool, int, bool, bool);
}  // namespace pyro

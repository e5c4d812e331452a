// Normal variates on the device from the counter-based generator of philox.hpp:
// element i of the output is normal number offset + i of stream `seed`, the
// same bits the host function of the C-ABI produces.
#include "philox.hpp"
#include "kernels.hpp"

namespace aphi_dev {

__global__ void __launch_bounds__(256) k_random_normal(uint64_t seed, uint64_t offset, int64_t count,
                                                       double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t k = offset + static_cast<uint64_t>(i);
    double z0, z1;
    aphi_rng::normal_pair(seed, k >> 1, z0, z1);
    out[i] = (k & 1) ? z1 : z0;
}

}  // namespace aphi_dev

// Persistent signing engine: mode dispatch and workgroup tiling defaults.
#include "sign_persistent.hpp"

namespace dil {

// clears the queue head, status and attempt count, and the phase sums
__global__ void zero_ctl_kernel(uint32_t *p, unsigned long long *phase_sums) {
    p[0] = 0;
    p[1] = 0;
    p[2] = 0;
    if (phase_sums)
        for (int w = 0; w < PERSIST_PHASES; ++w) phase_sums[w] = 0;
}

int launch_zero_ctl(uint32_t *d_head, hipStream_t st) {
    hipLaunchKernelGGL(zero_ctl_kernel, dim3(1), dim3(1), 0, st, d_head, nullptr);
    return (int)hipGetLastError();
}

int launch_sign_persistent(int mode, const SignArgs &a, const SlotRef &sr, size_t nslots, int slots_per_block,
                           int occupancy, int blocks, int claim, uint32_t *d_head, unsigned long long *phase_sums,
                           hipStream_t st) {
    if (a.n == 0) return 0;
    if (occupancy == 0) occupancy = 2;
    switch (mode) {
        case 2: return launch_persistent_p<Params<2>>(a, sr, nslots, slots_per_block, occupancy, blocks, claim, d_head,
                                                      phase_sums, st);
        case 3: return launch_persistent_p<Params<3>>(a, sr, nslots, slots_per_block, occupancy, blocks, claim, d_head,
                                                      phase_sums, st);
        case 5: return launch_persistent_p<Params<5>>(a, sr, nslots, slots_per_block, occupancy, blocks, claim, d_head,
                                                      phase_sums, st);
        default: return -1;
    }
}

// Slots per workgroup used when the caller does not choose: the largest
// supported value that still leaves one workgroup per compute unit.
int default_slots_per_block(size_t nslots) {
    const size_t cus = (size_t)device_cu_count();
    for (int s : {32, 16})
        if (nslots / s >= cus) return s;
    return nslots >= 8 ? 8 : 1;
}

}  // namespace dil

// The reference functor formulas (shared with the generic engine) and the kernels' DPP row-neighbour helpers.
#pragma once

#include "../../include/gt355/functors.hpp"

namespace gt355 {

    // Row neighbours through DPP wavefront shifts.  With bound control off, a lane whose source lies outside the
    // wavefront keeps the destination's old value: passing the side column as `old` hands it to the edge lane.
    constexpr int DPP_WF_SL1 = 0x130; // lane L reads lane L+1; lane 63 keeps its old value
    constexpr int DPP_WF_SR1 = 0x138; // lane L reads lane L-1; lane 0 keeps its old value

    template <int CTRL>
    __device__ __forceinline__ float dpp_keep(float old, float src) {
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                         __builtin_bit_cast(int, src), CTRL, 0xF, 0xF, false));
    }
    template <int CTRL>
    __device__ __forceinline__ double dpp_keep(double old, double src) {
        const long long o = __builtin_bit_cast(long long, old), s = __builtin_bit_cast(long long, src);
        const int lo = __builtin_amdgcn_update_dpp(int(o), int(s), CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(int(o >> 32), int(s >> 32), CTRL, 0xF, 0xF, false);
        return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    }

} // namespace gt355

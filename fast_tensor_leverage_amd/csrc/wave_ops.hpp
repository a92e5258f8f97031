// wave_ops.hpp -- wavefront-level helpers (64-lane CDNA waves): LDS ordering
// within a wave, DPP permutations of doubles, and group sums / scans over
// power-of-two lane groups.
#pragma once

#include <hip/hip_runtime.h>

namespace ftl {

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP lane permutation of a double (two 32-bit halves).
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v[lane & ~16] + v[lane | 16] in every lane (rows 0+1 and 2+3 of a wave):
// v_permlane16_swap of a register with itself returns the even-row value in
// its first result and the odd-row value in its second.
__device__ inline double pair_sum16(double v) {
    const long long b = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((int)(b & 0xffffffff), (int)(b & 0xffffffff), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((int)(b >> 32), (int)(b >> 32), false, false);
    const double even = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
    const double odd = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
    return even + odd;
}

// v[lane & ~32] + v[lane | 32] in every lane (v_permlane32_swap likewise).
__device__ inline double pair_sum32(double v) {
    const long long b = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((int)(b & 0xffffffff), (int)(b & 0xffffffff), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((int)(b >> 32), (int)(b >> 32), false, false);
    const double low = __longlong_as_double(((long long)hi[0] << 32) | (unsigned int)lo[0]);
    const double high = __longlong_as_double(((long long)hi[1] << 32) | (unsigned int)lo[1]);
    return low + high;
}

// Sum over a group of SG consecutive lanes.  Each step adds a lane's value to
// that of a partner lane such that both end with the same operands, so every
// lane of the group ends with the bitwise identical sum (IEEE addition is
// commutative).  Within a 16-lane row the partners come from DPP (quad
// permutations, half-row and row mirrors); across rows from the gfx950 lane
// swaps (register to register, no LDS round trip).
template <int SG>
__device__ inline double group_sum(double v) {
    static_assert(SG >= 1 && SG <= 64 && (SG & (SG - 1)) == 0, "group width");
    if constexpr (SG == 1) return v;
    v += dpp_f64<0xB1>(v);                     // quad_perm [1,0,3,2]
    if constexpr (SG >= 4) v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    if constexpr (SG >= 8) v += dpp_f64<0x141>(v);  // row_half_mirror
    if constexpr (SG >= 16) v += dpp_f64<0x140>(v); // row_mirror
    if constexpr (SG >= 32) v = pair_sum16(v);
    if constexpr (SG >= 64) v = pair_sum32(v);
    return v;
}

template <int SG>
__device__ inline double group_incl_scan(double v, int g) {
#pragma unroll
    for (int o = 1; o < SG; o <<= 1) {
        double y = __shfl_up(v, o, SG);
        if (g >= o) v += y;
    }
    return v;
}

}  // namespace ftl

// grp_dev.hpp -- lane-level helpers shared by the flash kernels: lane id,
// wave-uniform work fetch, and the 2S-lane cell groups (S lanes per phase,
// one lane per component) with their butterfly reductions.
#pragma once
#include <hip/hip_runtime.h>

namespace osr {

__device__ __forceinline__ unsigned lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Contraction stays off for every operand of a butterfly: a product fused
// into the first add (fma(z, td, partner's rounded z*td)) rounds differently
// on the two lanes of a pair and the lanes of a group would then disagree.
template <int S>
__device__ __forceinline__ double half_sum(double v) {
#pragma clang fp contract(off)
#pragma unroll
    for (int o = 1; o < S; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One atomic per wave whose result is wave-uniform by construction (SGPR),
// so loops that exit on it have a uniform exit and the compiler keeps every
// lane of the wave on the same iteration (cross-lane reads inside the loop
// then always see their partners' current values).  All 64 lanes must be
// active at the call.
__device__ __forceinline__ unsigned long long wave_fetch_add(unsigned long long* p, unsigned long long v) {
    unsigned long long r = 0;
    if (lane_id() == 0) r = atomicAdd(p, v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(r & 0xffffffffull));
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(r >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Lanes of one group: h = phase (0 liquid, 1 vapour), j = component slot.
template <int S>
struct Grp {
    int lane, gl, gbase, h, j, jc;
    bool valid;
};

template <int NC, int S>
__device__ __forceinline__ Grp<S> grp_lane() {
    Grp<S> g;
    g.lane = (int)lane_id();
    g.gl = g.lane % (2 * S);
    g.gbase = g.lane - g.gl;
    g.h = g.gl / S;
    g.j = g.gl % S;
    g.valid = g.j < NC;
    g.jc = g.valid ? g.j : 0;
    return g;
}

// max over a half-group (fmax is commutative, so every lane gets the same)
template <int S>
__device__ __forceinline__ double half_max(double v) {
#pragma unroll
    for (int o = 1; o < S; o <<= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// (max value, smallest index among the maxima) over a half-group; v must not
// be NaN, so every lane of the half ends with the same pair
template <int S>
__device__ __forceinline__ void half_argmax(double& v, int& idx) {
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
}

}  // namespace osr

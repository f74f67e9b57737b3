// grp_dev.hpp -- lane-level helpers shared by the flash kernels: lane id,
// wave-uniform work fetch, and the 2S-lane cell groups (S lanes per phase,
// one lane per component) with their butterfly reductions.
#pragma once
#include <hip/hip_runtime.h>

namespace osr {

__device__ __forceinline__ unsigned lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Butterfly reductions inside a half-group of S lanes (S = 8: half a DPP
// row, S = 16: one DPP row), with DPP lane moves at VALU latency instead of
// LDS-path permutes: xor 1 and xor 2 by quad_perm, then row_half_mirror
// (lane i <-> 7 - i, joining the two quads of 8 lanes) and, for S = 16,
// row_mirror (i <-> 15 - i, joining the two halves of the row).  Every
// stage pairs lanes symmetrically, so the two lanes of a pair add the same
// operands (a + b = b + a exactly) and every lane of the half ends with
// bitwise the same value -- the invariant that keeps a group's control flow
// uniform.  Contraction stays off: a product fused into the first add would
// round differently on the two lanes of a pair.
constexpr int DPP_QUAD_XOR1 = 0xB1;   // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_XOR2 = 0x4E;   // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

template <int S>
__device__ __forceinline__ double half_sum(double v) {
#pragma clang fp contract(off)
    static_assert(S == 8 || S == 16, "half-groups are 8 or 16 lanes");
    v += dpp_f64<DPP_QUAD_XOR1>(v);
    v += dpp_f64<DPP_QUAD_XOR2>(v);
    v += dpp_f64<DPP_ROW_HALF_MIRROR>(v);
    if (S == 16) v += dpp_f64<DPP_ROW_MIRROR>(v);
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
    v = fmax(v, dpp_f64<DPP_QUAD_XOR1>(v));
    v = fmax(v, dpp_f64<DPP_QUAD_XOR2>(v));
    v = fmax(v, dpp_f64<DPP_ROW_HALF_MIRROR>(v));
    if (S == 16) v = fmax(v, dpp_f64<DPP_ROW_MIRROR>(v));
    return v;
}

// (max value, smallest index among the maxima) over a half-group; v must not
// be NaN, so every lane of the half ends with the same pair
template <int CTRL>
__device__ __forceinline__ void argmax_step(double& v, int& idx) {
    const double ov = dpp_f64<CTRL>(v);
    const int oi = dpp_i32<CTRL>(idx);
    if (ov > v || (ov == v && oi < idx)) {
        v = ov;
        idx = oi;
    }
}

template <int S>
__device__ __forceinline__ void half_argmax(double& v, int& idx) {
    argmax_step<DPP_QUAD_XOR1>(v, idx);
    argmax_step<DPP_QUAD_XOR2>(v, idx);
    argmax_step<DPP_ROW_HALF_MIRROR>(v, idx);
    if (S == 16) argmax_step<DPP_ROW_MIRROR>(v, idx);
}

}  // namespace osr

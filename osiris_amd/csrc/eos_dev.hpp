// eos_dev.hpp -- device-side cubic-EOS arithmetic for gfx950.
//
// Everything here is templated on the real type R (double or float) and on
// the compile-time component count NC, so per-cell vectors live in VGPRs and
// every component loop is fully unrolled.  Fluid constants sit in a
// read-only device struct (FluidDev) indexed with wave-uniform offsets, which
// the compiler turns into scalar (SMEM) loads.
//
// Reference arithmetic restated here:
//   EOS_Init         compsim/Init_Calcs.cpp:66-97
//   Solve_Z          compsim/Globals.cpp:212-264
//   mixing + ln phi  compsim/Init_Calcs.cpp:650-712 (canonical inline copy)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace osr {

constexpr int NCMAX = 16;
constexpr double RGAS = 8.3144621;                     // Functions.h:216
constexpr double SQRT2 = 1.4142135623730950488016887;  // Functions.h:219
constexpr double NEWTON_TOL = 1e-8;                    // Functions.h:236
constexpr double FLASH_TOL = 1e-14;                    // Functions.h:237
constexpr double DZERO = 1e-20;                        // Functions.h:239
constexpr int MAX_FLASH_N_ITR = 100;                   // Functions.h:126
constexpr int MAX_SS_ITER = 500;                       // Init_Calcs.cpp:714

// Per-precision copy of the temperature-independent fluid constants.
template <class R>
struct FluidConsts {
    R Tc[NCMAX];     // K
    R Pc[NCMAX];     // Pa
    R sac[NCMAX];    // sqrt(ac_i): a_i(T) = ac_i (1 + m_i (1 - sqrt(T/Tc_i)))^2
    R m[NCMAX];
    R b[NCMAX];      // co-volume b_i (T-independent)
    R wk[NCMAX];     // Wilson exponent factor 5.37 (1 + omega_i)
    R bic[NCMAX * NCMAX];  // 1 - k_ij
    R d1, d2;        // EOS delta constants: PR 1 +- sqrt2, SRK 1, 0
};

struct FluidDev {
    int32_t nc, eos;
    FluidConsts<double> d;
    FluidConsts<float> f;
};

template <class R> __device__ __forceinline__ const FluidConsts<R>& consts(const FluidDev* F);
template <> __device__ __forceinline__ const FluidConsts<double>& consts<double>(const FluidDev* F) { return F->d; }
template <> __device__ __forceinline__ const FluidConsts<float>& consts<float>(const FluidDev* F) { return F->f; }

// Math wrappers so templates pick the right-precision device library call.
__device__ __forceinline__ double m_sqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float m_sqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double m_log(double v) { return log(v); }
__device__ __forceinline__ float m_log(float v) { return logf(v); }
__device__ __forceinline__ double m_exp(double v) { return exp(v); }
__device__ __forceinline__ float m_exp(float v) { return expf(v); }
__device__ __forceinline__ double m_acos(double v) { return acos(v); }
__device__ __forceinline__ float m_acos(float v) { return acosf(v); }
__device__ __forceinline__ double m_cbrt(double v) { return cbrt(v); }
__device__ __forceinline__ float m_cbrt(float v) { return cbrtf(v); }
__device__ __forceinline__ void m_sincos(double v, double* s, double* c) { sincos(v, s, c); }
__device__ __forceinline__ void m_sincos(float v, float* s, float* c) { sincosf(v, s, c); }
__device__ __forceinline__ bool m_isnan(double v) { return v != v; }
__device__ __forceinline__ bool m_isnan(float v) { return v != v; }

// Root of Z^3 + a1 Z^2 + a2 Z + a3 chosen as Solve_Z does (Globals.cpp:212-264).
// The three trigonometric roots use one sincos: cos(t +- 2pi/3) =
// -cos(t)/2 -+ sin(t) sqrt(3)/2.  The root-selection comparisons are the
// reference's, written out literally (not an assumed ordering).
template <class R>
__device__ __forceinline__ R solve_z(R a1, R a2, R a3, bool gas) {
    // divisions by the constants 9, 54 and 3 as products with their
    // reciprocals (one rounding each instead of a division sequence)
    const R third = R(0.33333333333333333333333333333333);
    const R Q = (R(3) * a2 - a1 * a1) * R(0.11111111111111111111111111111111);
    const R J = (R(9) * a1 * a2 - R(27) * a3 - R(2) * a1 * a1 * a1) * R(0.018518518518518518518518518518519);
    const R D = Q * Q * Q + J * J;
    R Z;
    if (D < 0) {
        const R r = R(2) * m_sqrt(-Q);
        const R s = a1 * third;
        const R t = m_acos(J / m_sqrt(-Q * Q * Q)) * third;
        R sn, cs;
        m_sincos(t, &sn, &cs);
        const R h = R(0.86602540378443864676372317075294);
        const R x1 = r * cs - s;
        const R x2 = r * (R(-0.5) * cs - h * sn) - s;
        const R x3 = r * (R(-0.5) * cs + h * sn) - s;
        if (gas) {
            if (x1 >= x2 && x1 >= x3) Z = x1;
            else if (x2 >= x1 && x2 >= x3) Z = x2;
            else Z = x3;
        } else {
            if (x1 <= x2 && x1 <= x3 && x1 > 0) Z = x1;
            else if (x2 <= x1 && x2 <= x3 && x2 > 0) Z = x2;
            else Z = x3;
        }
    } else if (D == 0) {
        const R t = m_cbrt(J);
        const R x1 = R(2) * t - a1 * third;
        const R x2 = -t - a1 * third;
        if (gas) Z = (x2 >= x1) ? x2 : x1;
        else Z = (x2 <= x1 && x2 > 0) ? x2 : x1;
    } else {
        const R sd = m_sqrt(D);
        Z = m_cbrt(J + sd) + m_cbrt(J - sd) - a1 * third;
    }
    return Z;
}

// Z of one phase from its scaled A, B (U = 2B, W = B for PR; U = B, W = 0 SRK).
template <class R>
__device__ __forceinline__ R phase_z(bool srk, R A, R B, bool gas) {
    const R U = srk ? B : R(2) * B;
    const R W = srk ? R(0) : B;
    return solve_z<R>(-(R(1) + B - U), A - B * U - U - W * W, -(A * B - B * W * W - W * W), gas);
}

// Mixing of one phase: t_i = sqrt(a_i) sum_j bic_ij sqrt(a_j) x_j,
// Ca = sum_i x_i t_i, Cb = sum_i x_i b_i.  sa = sqrt(a_i(T)).
template <int NC, class R>
__device__ __forceinline__ void mix(const FluidConsts<R>& C, const R (&sa)[NC], const R (&x)[NC],
                                    R (&t)[NC], R& Ca, R& Cb) {
    R u[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) u[j] = sa[j] * x[j];
    Ca = 0;
    Cb = 0;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        R s = 0;
#pragma unroll
        for (int j = 0; j < NC; j++) s += C.bic[i * NC + j] * u[j];
        t[i] = sa[i] * s;
        Ca += u[i] * s;
        Cb += x[i] * C.b[i];
    }
}

// Phase invariants of ln phi_i = t1_i (Z-1) - ln(Z-B) + DD * EE_i * FF with
// t1_i = b_i/Cb, EE_i = 2 t_i/Ca - t1_i (Init_Calcs.cpp:694-711).
template <class R>
struct PhaseInv {
    R Z, A, B, invCa, invCb, lnZB, DDFF;
};

// pA = P/(R T)^2, pB = P/(R T): per-cell constants of the flash.
template <int NC, class R>
__device__ __forceinline__ PhaseInv<R> phase_eval(const FluidConsts<R>& C, bool srk, const R (&sa)[NC],
                                                  const R (&x)[NC], R pA, R pB, bool gas, R (&t)[NC]) {
    R Ca, Cb;
    mix<NC, R>(C, sa, x, t, Ca, Cb);
    PhaseInv<R> p;
    p.A = Ca * pA;
    p.B = Cb * pB;
    p.Z = phase_z<R>(srk, p.A, p.B, gas);
    p.invCa = R(1) / Ca;
    p.invCb = R(1) / Cb;
    p.lnZB = m_log(p.Z - p.B);
    const R DD = p.A / (p.B * (C.d1 - C.d2));
    p.DDFF = DD * m_log((p.Z + C.d2 * p.B) / (p.Z + C.d1 * p.B));
    return p;
}

template <class R>
__device__ __forceinline__ R lnphi(const PhaseInv<R>& p, R bi, R ti) {
    const R t1 = bi * p.invCb;
    const R EE = R(2) * ti * p.invCa - t1;
    return t1 * (p.Z - R(1)) - p.lnZB + p.DDFF * EE;
}

// sqrt(a_i(T)) for a cell temperature T.
template <int NC, class R>
__device__ __forceinline__ void sqrt_a(const FluidConsts<R>& C, R T, R (&sa)[NC]) {
#pragma unroll
    for (int i = 0; i < NC; i++) sa[i] = C.sac[i] * (R(1) + C.m[i] * (R(1) - m_sqrt(T / C.Tc[i])));
}

// Wilson K_i = Pc_i/P exp(5.37 (1+w_i)(1 - Tc_i/T)) (Init_Calcs.cpp:586-588).
template <int NC, class R>
__device__ __forceinline__ void wilson(const FluidConsts<R>& C, R T, R P, R (&K)[NC]) {
#pragma unroll
    for (int i = 0; i < NC; i++) K[i] = (C.Pc[i] / P) * m_exp(C.wk[i] * (R(1) - C.Tc[i] / T));
}

}  // namespace osr

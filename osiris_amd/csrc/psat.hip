// psat.hip -- saturation pressure (phase-appearance test) of independent
// cells on gfx950: Bubl_Succ / Dew_Succ (compsim/Minor_NR.cpp:491-871,
// 873-1249).
//
// Stage 1: successive substitution on K (<= 21 passes, |sum - 1|/k <= 1e-15),
// seeded -- as in both reference routines -- from comp[..][1]/K, with a
// triviality test.  Stage 2 (when stage 1 fails): Newton on the (nc+1)
// system [fixed-phase fugacity - incipient fugacity = 0 ; sum Xm = 1] in the
// incipient mole numbers and P, Wilson start, clamps, <= 21 iterations,
// mean relative step <= 1e-8.  The (nc+1)^2 Newton matrix of each lane lives
// in LDS, laid out [element][lane] so the 64 lanes of a wave hit 64 distinct
// banks; LU with partial pivoting (the dgesv arithmetic) runs in place.
#include <algorithm>

#include "osr_internal.hpp"
#include "phase_dev.hpp"

namespace osr {

constexpr double SXMTOL = 1e-15;   // Functions.h:251
constexpr double XMTOL = 1e-8;     // Functions.h:224
constexpr double TRIV_K = 1e-6;    // Functions.h:269
constexpr int MAX_MNR_ITER = 20;   // Functions.h:270
constexpr double PZERO = 1000000;  // Functions.h:247

struct PsatArgs {
    int64_t n;
    const double* T;
    double T0;
    const double* P;
    const int8_t* kind;
    const double* x;
    const double* y;
    double* psat;
    double* xm;
    const int32_t* list;                  // cells with a test to run (kind +-1)
    const unsigned long long* count;      // their number
};

// Phase state with the sensitivities the Newton stage needs.
template <int NC>
struct PhaseD {
    PhaseInv<double> p;
    double t[NC];
    double phi[NC];
    double Ca, Cb, G, den;   // dZ = -(dA (Z-B) + dB G) / den
};

template <int NC>
__device__ __forceinline__ void phase_full(const FluidConsts<double>& C, bool srk, const double (&sa)[NC],
                                           const double (&x)[NC], double P, double T, bool gas,
                                           PhaseD<NC>& s) {
    mix<NC, double>(C, sa, x, s.t, s.Ca, s.Cb);
    const double rt = RGAS * T;
    PhaseInv<double>& p = s.p;
    p.A = s.Ca * P / (rt * rt);
    p.B = s.Cb * P / rt;
    p.Z = phase_z<double>(srk, p.A, p.B, gas);
    p.invCa = 1.0 / s.Ca;
    p.invCb = 1.0 / s.Cb;
    p.lnZB = log(p.Z - p.B);
    const double DD = p.A / (p.B * (C.d1 - C.d2));
    p.DDFF = DD * log((p.Z + C.d2 * p.B) / (p.Z + C.d1 * p.B));
#pragma unroll
    for (int i = 0; i < NC; i++) s.phi[i] = exp(lnphi<double>(p, C.b[i], s.t[i]));
    const double Z = p.Z, A = p.A, B = p.B;
    const double U = srk ? B : 2 * B, W = srk ? 0.0 : B;
    const double gB = -Z * Z - U * Z - A + W * W;
    const double gU = Z * Z - B * Z - Z;
    const double gW = -2 * W * Z + 2 * B * W + 2 * W;
    s.G = srk ? (gB + gU) : (gB + 2 * gU + gW);   // dU = 2dB, dW = dB (PR); dU = dB (SRK)
    s.den = 3 * Z * Z - 2 * Z * (1 + B - U) + A - B * U - U - W * W;
}

// d ln(phi_i)/dP (Minor_NR.cpp:731-772).
template <int NC>
__device__ __forceinline__ void dlnphi_dp(const FluidConsts<double>& C, double T, const PhaseD<NC>& s,
                                          double (&out)[NC]) {
    const PhaseInv<double>& p = s.p;
    const double rt = RGAS * T;
    const double dA = s.Ca / (rt * rt), dB = s.Cb / rt;
    const double dZ = -(dA * (p.Z - p.B) + dB * s.G) / s.den;
    const double dd = C.d1 - C.d2;
    const double DD = p.A / (p.B * dd);
    const double FF = p.DDFF / DD;
    const double dDD = (p.B * dA - p.A * dB) / (p.B * p.B * dd);
    const double dFF = (dZ + C.d2 * dB) / (p.Z + C.d2 * p.B) - (dZ + C.d1 * dB) / (p.Z + C.d1 * p.B);
    const double dBB = (dB - dZ) / (p.Z - p.B);
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const double t1 = C.b[i] * p.invCb;
        const double EE = 2 * s.t[i] * p.invCa - t1;
        out[i] = t1 * dZ + dBB + EE * (dDD * FF + DD * dFF);
    }
}

// Cells without a test (kind 0) get psat = -1 and xm = 0 here; the others
// are appended to the work list (one atomic per wave, wave-aggregated), so
// the saturation kernel only schedules cells that run a test.
__global__ void k_psat_select(int64_t n, int nc, const int8_t* kind, double* psat, double* xm, int32_t* list,
                              unsigned long long* count) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool act = false;
    if (c < n) {
        const int kd = kind[c];
        act = kd == 1 || kd == -1;
        if (!act) {
            psat[c] = -1;
            if (xm)
                for (int i = 0; i < nc; i++) xm[(int64_t)i * n + c] = 0;
        }
    }
    const unsigned long long m = __ballot(act);
    if (m == 0) return;
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = (unsigned long long)__shfl((long long)base, leader, 64);
    if (act) list[base + __popcll(m & ((1ull << lane) - 1))] = (int32_t)c;
}

__global__ void k_psat_zero(unsigned long long* count) { *count = 0; }

template <int NC>
__device__ __forceinline__ void psat_cell(const FluidDev* __restrict__ F, const PsatArgs& a, int64_t c, int lane,
                                          double* sJ);

// One lane per listed cell (grid-stride over the work list).  The Newton
// matrix of each lane lives in LDS, laid out [element][lane].
template <int NC>
__global__ __launch_bounds__(64) void k_psat(const FluidDev* __restrict__ F, PsatArgs a) {
    extern __shared__ double sJ[];   // [(NC+1)^2][64]
    const int lane = threadIdx.x;
    const int64_t cnt = (int64_t)*a.count;
    for (int64_t i = (int64_t)blockIdx.x * 64 + lane; i < cnt; i += (int64_t)gridDim.x * 64)
        psat_cell<NC>(F, a, (int64_t)a.list[i], lane, sJ);
}

template <int NC>
__device__ __forceinline__ void psat_cell(const FluidDev* __restrict__ F, const PsatArgs& a, int64_t c, int lane,
                                          double* sJ) {
    constexpr int N1 = NC + 1;
    const FluidConsts<double>& C = F->d;
    const bool srk = F->eos == OSR_EOS_SRK;
    const int kind = a.kind[c];
    const bool bubble = kind == 1;
    const double T = a.T ? a.T[c] : a.T0;
    const double P0 = a.P[c];
    double fixed[NC], ystale[NC], sa[NC], Kw[NC], K[NC], X[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        ystale[i] = a.y[(int64_t)i * a.n + c];
        fixed[i] = bubble ? a.x[(int64_t)i * a.n + c] : ystale[i];
    }
    sqrt_a<NC, double>(C, T, sa);
    wilson<NC, double>(C, T, P0, Kw);

    // ---- stage 1: successive substitution
    double Pm = P0, sumK = 0, SXm = 0;
    int k = 0, nIter = 0;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < NC; i++) K[i] = Kw[i];
    PhaseD<NC> sf, si;
    do {
        double xs = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            double v = ystale[i] / K[i];
            v = v < 0 ? 0.0 : (v > 1 ? 1.0 : v);
            X[i] = v;
            xs += v;
        }
        if (Pm < 0) Pm = PZERO;
#pragma unroll
        for (int i = 0; i < NC; i++) X[i] /= xs;
        if (bubble) {
            phase_full<NC>(C, srk, sa, fixed, Pm, T, false, sf);
            phase_full<NC>(C, srk, sa, X, Pm, T, true, si);
        } else {
            phase_full<NC>(C, srk, sa, X, Pm, T, false, si);
            phase_full<NC>(C, srk, sa, fixed, Pm, T, true, sf);
        }
        SXm = 0;
        sumK = 0;
        k = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            if (fixed[i] != 0) {
                if (bubble) {
                    K[i] *= (fixed[i] * sf.phi[i]) / (X[i] * si.phi[i]);
                    SXm += K[i] * fixed[i];
                } else {
                    K[i] *= (X[i] * si.phi[i]) / (fixed[i] * sf.phi[i]);
                    SXm += fixed[i] / K[i];
                }
                sumK += fabs(1 - K[i]);
                k++;
            }
        }
        Pm *= SXm;
        nIter++;
        if (nIter > MAX_MNR_ITER) {
            bad = true;
            break;
        }
    } while (fabs(SXm - 1) / k > SXMTOL);
    if (sumK / k < TRIV_K) bad = true;
    double result;
    if (!bad && isfinite(Pm)) {   // the reference's (Pm - Pm) == 0, robust to FMA contraction
        result = Pm;
    } else {
        // ---- stage 2: Newton in (Xm, P)
#define JM(r, cc) sJ[((r) * N1 + (cc)) * 64 + lane]
        Pm = P0;
#pragma unroll
        for (int i = 0; i < NC; i++) X[i] = bubble ? Kw[i] * fixed[i] : fixed[i] / Kw[i];
        nIter = 0;
        double XmTol;
        result = 0;
        bool capped = false;
        do {
            double rhs[N1];
            double s = 0;
#pragma unroll
            for (int i = 0; i < NC; i++) {
                s += X[i];
                JM(NC, i) = 1;
            }
            rhs[NC] = 1 - s;
            JM(NC, NC) = 0;
            if (bubble) {
                phase_full<NC>(C, srk, sa, fixed, Pm, T, false, sf);
                phase_full<NC>(C, srk, sa, X, Pm, T, true, si);
            } else {
                phase_full<NC>(C, srk, sa, X, Pm, T, false, si);
                phase_full<NC>(C, srk, sa, fixed, Pm, T, true, sf);
            }
            double dpf[NC], dpi[NC];
            dlnphi_dp<NC>(C, T, sf, dpf);
            dlnphi_dp<NC>(C, T, si, dpi);
            // Incipient-phase composition sensitivities (Minor_NR.cpp:778-808).
            const PhaseInv<double>& p = si.p;
            const double rt = RGAS * T;
            const double dd = C.d1 - C.d2;
            const double DD = p.A / (p.B * dd);
            const double FF = p.DDFF / DD;
            const double dBk_s = Pm / rt;              // dB/dX_k = b_k * dBk_s
            const double dAk_s = 2 * Pm / (rt * rt);   // dA/dX_k = t_k * dAk_s
            const double sgn = bubble ? -1.0 : 1.0;
#pragma unroll
            for (int i = 0; i < NC; i++) {
                const double fixedterm = fixed[i] * sf.phi[i];
                const double incterm = X[i] * si.phi[i];
                rhs[i] = bubble ? (-fixedterm + incterm) : (-incterm + fixedterm);
                // bubble: J = x dphiL/dp - Xm dphiG/dp ; dew: Xm dphiL/dp - y dphiG/dp
                JM(i, NC) = bubble ? (fixedterm * dpf[i] - incterm * dpi[i])
                                   : (incterm * dpi[i] - fixedterm * dpf[i]);
            }
#pragma unroll
            for (int kk = 0; kk < NC; kk++) {
                const double dA = si.t[kk] * dAk_s;
                const double dB = C.b[kk] * dBk_s;
                const double dZ = -(dA * (p.Z - p.B) + dB * si.G) / si.den;
                const double dDD = (dA * p.B - dB * p.A) / (p.B * p.B * dd);
                const double dFF = (dZ + C.d2 * dB) / (p.Z + C.d2 * p.B) - (dZ + C.d1 * dB) / (p.Z + C.d1 * p.B);
                const double dBB = (dB - dZ) / (p.Z - p.B);
                const double bk = C.b[kk];
#pragma unroll
                for (int i = 0; i < NC; i++) {
                    const double bi = C.b[i];
                    const double t1 = bi * p.invCb;
                    const double EE = 2 * si.t[i] * p.invCa - t1;
                    const double dAA = -bk * bi * (p.Z - 1) * p.invCb * p.invCb + t1 * dZ;
                    const double aik = C.bic[i * NC + kk] * sa[i] * sa[kk];
                    const double dEE = 2 * (aik * p.invCa - 2 * si.t[kk] * si.t[i] * p.invCa * p.invCa) +
                                       bi * bk * p.invCb * p.invCb;
                    const double dln = dAA + dBB + dDD * EE * FF + DD * dEE * FF + DD * EE * dFF;
                    const double diag = (i == kk) ? si.phi[i] : 0.0;
                    JM(i, kk) = sgn * (diag + X[i] * si.phi[i] * dln);
                }
            }
            // LU with partial pivoting (first maximal |a|), then solve.
            int piv[N1];
#pragma unroll
            for (int col = 0; col < N1; col++) {
                int pr = col;
                double best = fabs(JM(col, col));
                for (int r = col + 1; r < N1; r++) {
                    const double v = fabs(JM(r, col));
                    if (v > best) { best = v; pr = r; }
                }
                piv[col] = pr;
                if (JM(pr, col) != 0) {
                    if (pr != col)
                        for (int cc = 0; cc < N1; cc++) {
                            const double t = JM(col, cc);
                            JM(col, cc) = JM(pr, cc);
                            JM(pr, cc) = t;
                        }
                    const double inv = 1.0 / JM(col, col);
                    for (int r = col + 1; r < N1; r++) JM(r, col) *= inv;
                }
                for (int r = col + 1; r < N1; r++) {
                    const double l = JM(r, col);
                    if (l != 0)
                        for (int cc = col + 1; cc < N1; cc++) JM(r, cc) -= l * JM(col, cc);
                }
            }
            double sol[N1];
#pragma unroll
            for (int i = 0; i < N1; i++) sol[i] = rhs[i];
#pragma unroll
            for (int col = 0; col < N1; col++) {
                const int pr = piv[col];
                if (pr != col) {
                    // swap sol[col] and sol[pr] without dynamic register indexing
                    double vp = 0;
#pragma unroll
                    for (int r = 0; r < N1; r++) if (r == pr) vp = sol[r];
#pragma unroll
                    for (int r = 0; r < N1; r++) if (r == pr) sol[r] = sol[col];
                    sol[col] = vp;
                }
            }
#pragma unroll
            for (int i = 0; i < N1; i++)
#pragma unroll
                for (int cc = 0; cc < i; cc++) sol[i] -= JM(i, cc) * sol[cc];
#pragma unroll
            for (int i = N1 - 1; i >= 0; i--) {
#pragma unroll
                for (int cc = i + 1; cc < N1; cc++) sol[i] -= JM(i, cc) * sol[cc];
                sol[i] /= JM(i, i);
            }
            // Update with clamps (Minor_NR.cpp:836-852).
            XmTol = 0;
            double xs = 0;
#pragma unroll
            for (int i = 0; i < NC; i++) {
                double v = X[i] + sol[i];
                v = v < 0 ? 0.0 : (v > 1 ? 1.0 : v);
                X[i] = v;
                xs += v;
                XmTol += (v != 0) ? fabs(sol[i] / v) : fabs(sol[i]);
            }
            Pm += sol[NC];
            if (Pm < 0) Pm = PZERO;
            XmTol += (Pm != 0) ? fabs(sol[NC] / Pm) : fabs(sol[NC]);
            XmTol /= N1;
#pragma unroll
            for (int i = 0; i < NC; i++) X[i] /= xs;
            nIter++;
            if (nIter > MAX_MNR_ITER) {
                capped = true;
                break;
            }
        } while (XmTol > XMTOL);
#undef JM
        if (capped) {
            result = -1;
        } else {
            double sk = 0;
            int kk = 0;
#pragma unroll
            for (int i = 0; i < NC; i++) {
                if (fixed[i] != 0) {
                    sk += bubble ? fabs(1 - X[i] / fixed[i]) : fabs(1 - fixed[i] / X[i]);
                    kk++;
                }
            }
            result = (sk / kk < TRIV_K) ? -1.0 : Pm;
        }
    }
    a.psat[c] = result;
    if (a.xm)
#pragma unroll
        for (int i = 0; i < NC; i++) a.xm[(int64_t)i * a.n + c] = X[i];
}

template <int NC>
static int launch_psat_nc(osr_ctx* ctx, const PsatArgs& a, hipStream_t stream) {
    const size_t lds = (size_t)(NC + 1) * (NC + 1) * 64 * sizeof(double);
    const int64_t per_cu = 4;   // more workgroups than fit at once at nc >= 8 (LDS-bound), for balance
    const int64_t grid = std::min<int64_t>((a.n + 63) / 64, (int64_t)(ctx->num_cu > 0 ? ctx->num_cu : 256) * per_cu);
    if (lds > 65536)
        OSR_HIP(hipFuncSetAttribute((const void*)k_psat<NC>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_psat<NC>, dim3((unsigned)grid), dim3(64), lds, stream, ctx->d_fluid, a);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

int launch_psat(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                const int8_t* kind, const double* x, const double* y, double* psat,
                double* xm, hipStream_t stream) {
    if (!ctx || n < 0) return OSR_E_ARG;
    if (n == 0) return OSR_OK;
    if (!P || !kind || !x || !y || !psat) return OSR_E_ARG;
    if (!T && !(T0 > 0)) return OSR_E_ARG;
    const size_t need = 256 + (size_t)n * sizeof(int32_t);
    if (need > ctx->plist_bytes) {
        if (ctx->ws_frozen) return OSR_E_SIZE;
        if (ctx->d_plist) OSR_HIP(hipFree(ctx->d_plist));
        ctx->d_plist = nullptr;
        ctx->plist_bytes = 0;
        OSR_HIP(hipMalloc(&ctx->d_plist, need));
        ctx->plist_bytes = need;
    }
    unsigned long long* count = (unsigned long long*)ctx->d_plist;
    int32_t* list = (int32_t*)((char*)ctx->d_plist + 256);
    hipLaunchKernelGGL(k_psat_zero, dim3(1), dim3(1), 0, stream, count);
    hipLaunchKernelGGL(k_psat_select, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, ctx->nc, kind,
                       psat, xm, list, count);
    PsatArgs a{n, T, T0, P, kind, x, y, psat, xm, list, count};
    switch (ctx->nc) {
#define CASE(NC) case NC: return launch_psat_nc<NC>(ctx, a, stream);
        OSR_NC_LIST(CASE)
#undef CASE
        default: return OSR_E_NC;
    }
}

// PhaseCtrl / fPhaseCtrl (Do_Cycle.cpp:1558-1722), split in three launches:
// k_phase_kind picks the saturation test per cell (dew at the gas pressure
// for gas cells, bubble at the oil pressure for liquid cells), k_psat runs
// it, k_phase_apply applies the label/saturation/composition rules.
__global__ void k_phase_kind(int64_t n, const int8_t* label, const double* Po, const double* Pg,
                             int8_t* kind, double* Pin) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int8_t L = label[c];
    kind[c] = (L == 1 || L == -1) ? L : 0;
    Pin[c] = L == -1 ? Pg[c] : Po[c];
}

template <int NC>
__global__ void k_phase_apply(int64_t n, int8_t* label, const double* Po, const double* Pg, double* x,
                              double* y, double* S, const double* psat, const double* xm) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    phase_apply_cell<NC>(n, c, label, Po[c], Pg[c], x, y, S, psat[c], xm);
}

int phase_ctrl(osr_ctx* ctx, int64_t n, const double* T, double T0, int8_t* label, const double* Po,
               const double* Pg, double* x, double* y, double* S, double* psat_out, hipStream_t st) {
    if (!ctx || n < 0) return OSR_E_ARG;
    if (n == 0) return OSR_OK;
    if (!label || !Po || !Pg || !x || !y || !S) return OSR_E_ARG;
    if (!T && !(T0 > 0)) return OSR_E_ARG;
    const size_t nc = ctx->nc;
    const size_t need = n * (sizeof(int8_t) + 2 * sizeof(double) + nc * sizeof(double)) + 1024;
    if (need > ctx->queue_bytes) {
        if (ctx->ws_frozen) return OSR_E_SIZE;
        if (ctx->d_queue) hipFree(ctx->d_queue);
        ctx->d_queue = nullptr;
        ctx->queue_bytes = 0;
        OSR_HIP(hipMalloc(&ctx->d_queue, need));
        ctx->queue_bytes = need;
    }
    char* base = (char*)ctx->d_queue;
    double* ps = (double*)base;
    double* Pin = ps + n;
    double* xm = Pin + n;
    int8_t* kind = (int8_t*)(xm + nc * n);
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_phase_kind, dim3(g), dim3(256), 0, st, n, label, Po, Pg, kind, Pin);
    int rc = launch_psat(ctx, n, T, T0, Pin, kind, x, y, ps, xm, st);
    if (rc) return rc;
    switch (ctx->nc) {
#define APPLY(NC) case NC: hipLaunchKernelGGL(k_phase_apply<NC>, dim3(g), dim3(256), 0, st, n, label, Po, Pg, x, y, S, ps, xm); break;
        OSR_NC_LIST(APPLY)
#undef APPLY
        default: return OSR_E_NC;
    }
    if (psat_out) OSR_HIP(hipMemcpyAsync(psat_out, ps, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

// The saturation pressures PhaseCtrl tests (kind and Bubl/Dew_Succ) without
// applying them: psat [n] and the incipient compositions xm [nc*n] into the
// caller's arrays, for a kernel that applies the rule itself.
int phase_psat(osr_ctx* ctx, int64_t n, const double* T, double T0, const int8_t* label, const double* Po,
               const double* Pg, const double* x, const double* y, double* psat, double* xm, hipStream_t st) {
    if (!ctx || n < 0) return OSR_E_ARG;
    if (n == 0) return OSR_OK;
    if (!label || !Po || !Pg || !x || !y || !psat || !xm) return OSR_E_ARG;
    if (!T && !(T0 > 0)) return OSR_E_ARG;
    const size_t need = n * (sizeof(int8_t) + sizeof(double)) + 1024;
    if (need > ctx->queue_bytes) {
        if (ctx->ws_frozen) return OSR_E_SIZE;
        if (ctx->d_queue) hipFree(ctx->d_queue);
        ctx->d_queue = nullptr;
        ctx->queue_bytes = 0;
        OSR_HIP(hipMalloc(&ctx->d_queue, need));
        ctx->queue_bytes = need;
    }
    double* Pin = (double*)ctx->d_queue;
    int8_t* kind = (int8_t*)(Pin + n);
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_phase_kind, dim3(g), dim3(256), 0, st, n, label, Po, Pg, kind, Pin);
    return launch_psat(ctx, n, T, T0, Pin, kind, x, y, psat, xm, st);
}

}  // namespace osr

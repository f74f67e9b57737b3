// flash_ad.hip -- tangent and adjoint of a converged two-phase flash.
//
// No counterpart in the reference (which finite-differences nothing here but
// also never differentiates its flash); these are the sensitivities of the
// phase split that SuccFalsh (compsim/Init_Calcs.cpp:571-738) computes.
//
// At a converged two-phase state the unknowns are u = ln K and
//   g_i(u; z, P) = ln K_i - ln phiL_i(x) + ln phiG_i(y) = 0,
// with x_i = z_i / D_i, y_i = K_i x_i, D_i = L + K_i (1 - L) and L the exact
// Rachford-Rice root sum_i z_i (1 - K_i)/D_i = 0.  By the implicit-function
// theorem, for a perturbation (dz, dP)
//   J du = -(B dz + p dP),
//   J = dg/du = I - PhiL (dX + u_x w^T) + PhiG (dY + u_y w^T),
//   B = dg/dz =   - PhiL (diag(1/D) + u_x v^T) + PhiG (diag(K/D) + u_y v^T),
//   p = dg/dP = -psiL + psiG,
// where Phi = d ln phi / dx (cubic EOS, composition at fixed P), psi =
// d ln phi / dP (fixed composition), dX = diag(d x_i/d u_i) at fixed L, u_x =
// d x / dL, and w = dL/du, v = dL/dz from the RR root.  Outputs follow from
//   dL = w.du + v.dz,  dx = dz/D + u_x dL + dX du,  dy = K (dx + du) x.
// The adjoint runs the same chain backwards with J^T lambda = ubar.
//
// Single-phase cells (label 1 or -1): the phase composition is z and L is
// constant, so dx = dz (L = 1) or dy = dz (L = 0); the absent phase's
// composition is reported with K held fixed (y = K z, x = z / K).
// Components with z_i = 0 keep K_i fixed (the reference does not update
// them, Init_Calcs.cpp:705).
//
// Form: one cell per group of 2S lanes, lane (h, j) = phase h, component j,
// as in the mixed Newton; row j of J, B lives in lane j of each half (the
// two halves hold identical rows), the factorisation pivots across a half.
#include "osr_internal.hpp"
#include "grp_dev.hpp"

namespace osr {

template <int NC>
struct LinRow {
    double J[NC];   // row j of dg/du
    double B[NC];   // row j of dg/dz
    double p;       // dg_j/dP
    double w, v;    // dL/du_j, dL/dz_j
    double D, x, y, K, z, L;
    double dxd, ux; // d x_j/d u_j at fixed L, d x_j / dL
    bool two;       // cell is two-phase with L in (0, 1)
    bool active;    // component j carries an equation
};

// Exact RR split and the linearisation of g at (K, z, P).  Group-uniform.
template <int NC, int S>
__device__ __forceinline__ void flash_linearise(const FluidConsts<double>& C, bool srk, const Grp<S>& g,
                                                double T, double P, double z, double K, int label,
                                                LinRow<NC>& R) {
#pragma clang fp contract(off)
    const bool valid = g.valid;
    const int h = g.h, j = g.j;
    const int hb = g.gbase + h * S;
    R.K = K;
    R.z = z;
    R.active = valid && z != 0;
    // exact Rachford-Rice root (Newton from the reference's start, run to
    // convergence)
    double L = 0.6, step;
    int it = 0;
    do {
        double td = 0;
        if (valid) td = (1.0 - K) / (L + (1.0 - L) * K);
        const double Fs = half_sum<S>(z * td);
        const double dF = -half_sum<S>(z * td * td);
        step = Fs / dF;
        L -= step;
        it++;
    } while (fabs(step) > 1e-15 && it < 200);
    R.two = label == 0 && L > 0 && L < 1;
    if (!R.two) L = label == 1 ? 1.0 : (label == -1 ? 0.0 : (L >= 1 ? 1.0 : 0.0));
    R.L = L;
    const double D = L + K * (1.0 - L);
    R.D = D;
    R.x = valid ? z / D : 0.0;
    R.y = K * R.x;
    const double iD2 = 1.0 / (D * D);
    R.dxd = valid ? -z * K * (1.0 - L) * iD2 : 0.0;
    R.ux = valid ? -z * (1.0 - K) * iD2 : 0.0;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        R.J[k] = (k == j) ? 1.0 : 0.0;
        R.B[k] = 0.0;
    }
    R.p = 0;
    R.w = R.v = 0;
    if (!R.two) return;   // group-uniform
    const double td = valid ? (1.0 - K) / D : 0.0;
    const double dF = -half_sum<S>(z * td * td);
    R.w = valid ? (z * K * iD2) / dF : 0.0;    // -(dR/du_j)/(dR/dL), dR/du_j = -z_j K_j / D_j^2
    R.v = valid ? -td / dF : 0.0;              // -(dR/dz_j)/(dR/dL)
    // phase h
    const double cph = h == 0 ? R.x : R.y;
    const double dd = h == 0 ? R.dxd : K * (R.dxd + R.x);   // d comp_j / d u_j at fixed L
    const double uu = h == 0 ? R.ux : K * R.ux;             // d comp_j / dL
    const double bb = valid ? (h == 0 ? 1.0 / D : K / D) : 0.0;   // d comp_j / d z_j at fixed L
    const double rt = RGAS * T;
    const double pB = P / rt, pA = pB / rt;
    const double sa = valid ? C.sac[g.jc] * (1.0 + C.m[g.jc] * (1.0 - sqrt(T / C.Tc[g.jc]))) : 0.0;
    const double bj = C.b[g.jc];
    const double u = sa * cph;
    double sacc = 0;
#pragma unroll
    for (int k = 0; k < NC; k++) sacc += C.bic[g.jc * NC + k] * __shfl(u, hb + k, 64);
    const double t = valid ? sa * sacc : 0.0;
    const double Ca = half_sum<S>(u * sacc);
    const double Cb = half_sum<S>(cph * (valid ? bj : 0.0));
    const double A = Ca * pA, Bv = Cb * pB;
    const double Z = phase_z<double>(srk, A, Bv, h == 1);
    const double iCa = 1.0 / Ca, iCb = 1.0 / Cb;
    const double q = A / (Bv * (C.d1 - C.d2));
    const double zd1 = Z + C.d1 * Bv, zd2 = Z + C.d2 * Bv;
    const double Lam = log(zd2 / zd1);
    const double uE = srk ? 1.0 : 2.0, wE = srk ? 0.0 : 1.0;
    const double c2p = -1.0 + (uE - 1.0) * Bv;
    const double c1p = A - uE * Bv * Bv - uE * Bv - wE * Bv * Bv;
    const double FZ = 3.0 * Z * Z + 2.0 * c2p * Z + c1p;
    const double FA = Z - Bv;
    const double FB = (uE - 1.0) * Z * Z - (2.0 * (uE + wE) * Bv + uE) * Z - A + 3.0 * wE * Bv * Bv + 2.0 * wE * Bv;
    const double qL = q * Lam;
    const double bi = bj * iCb;
    const double Ei = 2.0 * t * iCa - bi;
    // d ln phi_j / dP at fixed composition (A, B proportional to P, q not)
    const double ZP = -(FA * A + FB * Bv) / (FZ * P);
    const double LamP = (ZP + C.d2 * Bv / P) / zd2 - (ZP + C.d1 * Bv / P) / zd1;
    const double psi = valid ? bi * ZP - (ZP - Bv / P) / (Z - Bv) + q * Ei * LamP : 0.0;
    // column factors of Phi (own column k = j)
    double Uc = 0, Vc = 0, Wc = 0;
    if (valid) {
        const double Ak = 2.0 * t * pA, Bk = bj * pB;
        const double Zk = -(FA * Ak + FB * Bk) / FZ;
        const double qk = q * (Ak / A - Bk / Bv);
        const double Lk = (Zk + C.d2 * Bk) / zd2 - (Zk + C.d1 * Bk) / zd1;
        Uc = Zk - bj * (Z - 1.0) * iCb + qL * bj * iCb;
        Vc = (Zk - Bk) / (Z - Bv);
        Wc = Lam * qk + q * Lk;
    }
    const double c1 = 4.0 * qL * iCa * iCa, c2 = 2.0 * qL * iCa;
    const double c1i = c1 * t, c2i = c2 * sa;
    const double sgn = h == 0 ? -1.0 : 1.0;
    double rJ = 0;
    double partJ[NC], partB[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const int src = hb + k;
        const double Uk = __shfl(Uc, src, 64), Vk = __shfl(Vc, src, 64), Wk = __shfl(Wc, src, 64);
        const double tk = __shfl(t, src, 64), sk = __shfl(sa, src, 64);
        const double phi = bi * Uk - Vk + Ei * Wk - c1i * tk + c2i * C.bic[g.jc * NC + k] * sk;
        partJ[k] = sgn * phi * __shfl(dd, src, 64);
        partB[k] = sgn * phi * __shfl(bb, src, 64);
        rJ += sgn * phi * __shfl(uu, src, 64);
    }
    // sum the two phases' parts (same operand order in both halves)
    const double rO = __shfl_xor(rJ, S, 64);
    const double r = h == 0 ? rJ + rO : rO + rJ;
    const double psiO = __shfl_xor(psi, S, 64);
    R.p = R.active ? (h == 0 ? -psi + psiO : -psiO + psi) : 0.0;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const double wk = __shfl(R.w, g.gbase + k, 64);
        const double vk = __shfl(R.v, g.gbase + k, 64);
        const double oJ = __shfl_xor(partJ[k], S, 64), oB = __shfl_xor(partB[k], S, 64);
        const double sJ = h == 0 ? partJ[k] + oJ : oJ + partJ[k];
        const double sB = h == 0 ? partB[k] + oB : oB + partB[k];
        R.J[k] = R.active ? ((k == j ? 1.0 : 0.0) + sJ + r * wk) : (k == j ? 1.0 : 0.0);
        R.B[k] = R.active ? sB + r * vk : 0.0;
    }
}

// Solve M u = rhs for the rows of a half (lane j holds row j and rhs_j);
// returns u_j in lane j.  Partial pivoting across the lanes; group-uniform.
template <int NC, int S>
__device__ __forceinline__ double grp_solve(const Grp<S>& g, double (&row)[NC], double rhs, bool& sing) {
#pragma clang fp contract(off)
    const int j = g.j;
    const int hb = g.gbase + g.h * S;
    sing = false;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        double cand = (g.valid && j >= k) ? fabs(row[k]) : -1.0;
        if (cand != cand) cand = -1.0;
        int piv = j;
        half_argmax<S>(cand, piv);
        if (!(cand > 0.0)) sing = true;
        const int src = (j == k) ? piv : (j == piv ? k : j);
#pragma unroll
        for (int c = k; c < NC; c++) row[c] = __shfl(row[c], hb + src, 64);
        rhs = __shfl(rhs, hb + src, 64);
        const double pkk = __shfl(row[k], hb + k, 64);
        const double prh = __shfl(rhs, hb + k, 64);
        const double f = (g.valid && j > k) ? row[k] / pkk : 0.0;
#pragma unroll
        for (int c = k + 1; c < NC; c++) row[c] -= f * __shfl(row[c], hb + k, 64);
        rhs -= f * prh;
    }
    double acc = rhs, out = 0;
#pragma unroll
    for (int k = NC - 1; k >= 0; k--) {
        const double dk = __shfl(acc / row[k], hb + k, 64);
        if (j == k) out = dk;
        if (j < k) acc -= row[k] * dk;
    }
    return g.valid ? out : 0.0;
}

struct AdArgs {
    int64_t n;
    const double* T;
    double T0;
    const double* P;
    const double* z;
    const double* K;
    const int8_t* label;
    // tangent
    const double* dP;
    const double* dz;
    double* dx;
    double* dy;
    double* dL;
    // adjoint
    const double* xbar;
    const double* ybar;
    const double* Lbar;
    double* Pbar;
    double* zbar;
};

template <int NC, int S>
__global__ __launch_bounds__(256) void k_flash_tangent(const FluidDev* __restrict__ F, AdArgs a) {
#pragma clang fp contract(off)
    constexpr int G = 2 * S;
    const Grp<S> g = grp_lane<NC, S>();
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (c >= a.n) return;   // group-uniform
    const FluidConsts<double>& C = F->d;
    const bool srk = F->eos == OSR_EOS_SRK;
    const int64_t n = a.n;
    const int jc = g.jc;
    const double T = a.T ? a.T[c] : a.T0;
    const double P = a.P[c];
    const double z = g.valid ? a.z[(int64_t)jc * n + c] : 0.0;
    const double K = g.valid ? a.K[(int64_t)jc * n + c] : 1.0;
    const double dP = a.dP ? a.dP[c] : 0.0;
    const double dz = (g.valid && a.dz) ? a.dz[(int64_t)jc * n + c] : 0.0;
    LinRow<NC> R;
    flash_linearise<NC, S>(C, srk, g, T, P, z, K, a.label[c], R);
    double du = 0, dL = 0, dx, dy;
    if (R.two) {
        double rhs = -R.p * dP;
#pragma unroll
        for (int k = 0; k < NC; k++) rhs -= R.B[k] * __shfl(dz, g.gbase + k, 64);
        bool sing;
        du = grp_solve<NC, S>(g, R.J, rhs, sing);
        if (sing || !R.active) du = R.active ? NAN : 0.0;
        dL = half_sum<S>(g.valid ? R.w * du + R.v * dz : 0.0);
        dx = g.valid ? dz / R.D + R.ux * dL + R.dxd * du : 0.0;
        dy = K * dx + R.y * du;
    } else if (R.L >= 1.0) {
        dx = dz;
        dy = K * dz;
    } else {
        dy = dz;
        dx = g.valid ? dz / K : 0.0;
    }
    if (g.h == 0 && g.valid) {
        if (a.dx) a.dx[(int64_t)jc * n + c] = dx;
        if (a.dy) a.dy[(int64_t)jc * n + c] = dy;
    }
    if (g.gl == 0 && a.dL) a.dL[c] = R.two ? dL : 0.0;
}

template <int NC, int S>
__global__ __launch_bounds__(256) void k_flash_adjoint(const FluidDev* __restrict__ F, AdArgs a) {
#pragma clang fp contract(off)
    constexpr int G = 2 * S;
    const Grp<S> g = grp_lane<NC, S>();
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (c >= a.n) return;   // group-uniform
    const FluidConsts<double>& C = F->d;
    const bool srk = F->eos == OSR_EOS_SRK;
    const int64_t n = a.n;
    const int j = g.j, jc = g.jc;
    const int hb = g.gbase + g.h * S;
    const double T = a.T ? a.T[c] : a.T0;
    const double P = a.P[c];
    const double z = g.valid ? a.z[(int64_t)jc * n + c] : 0.0;
    const double K = g.valid ? a.K[(int64_t)jc * n + c] : 1.0;
    const double xb = (g.valid && a.xbar) ? a.xbar[(int64_t)jc * n + c] : 0.0;
    const double yb = (g.valid && a.ybar) ? a.ybar[(int64_t)jc * n + c] : 0.0;
    const double Lb = a.Lbar ? a.Lbar[c] : 0.0;
    LinRow<NC> R;
    flash_linearise<NC, S>(C, srk, g, T, P, z, K, a.label[c], R);
    double zb, Pb = 0;
    if (R.two) {
        // backwards through y = K x, x = z/D(L, u), L(u, z)
        const double xt = xb + K * yb;
        double ub = yb * R.y + R.dxd * xt;
        const double Lt = Lb + half_sum<S>(g.valid ? R.ux * xt : 0.0);
        zb = g.valid ? xt / R.D + R.v * Lt : 0.0;
        ub += R.w * Lt;
        if (!R.active) ub = 0.0;
        // J^T lambda = ubar: lane j takes column j of J
        double JT[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            double col = 0;
#pragma unroll
            for (int cc = 0; cc < NC; cc++) {
                const double v = __shfl(R.J[cc], hb + k, 64);   // J[k][cc]
                if (cc == j) col = v;
            }
            JT[k] = col;
        }
        bool sing;
        double lam = grp_solve<NC, S>(g, JT, ub, sing);
        if (sing) lam = NAN;
        // zbar -= B^T lambda, Pbar = -p.lambda
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const double s = half_sum<S>(g.valid ? lam * R.B[k] : 0.0);
            if (j == k) zb -= s;
        }
        Pb = -half_sum<S>(g.valid ? R.p * lam : 0.0);
    } else if (R.L >= 1.0) {
        zb = xb + K * yb;
    } else {
        zb = yb + (g.valid ? xb / K : 0.0);
    }
    if (g.h == 0 && g.valid && a.zbar) a.zbar[(int64_t)jc * n + c] = zb;
    if (g.gl == 0 && a.Pbar) a.Pbar[c] = Pb;
}

template <int NC>
static int launch_ad(osr_ctx* ctx, const AdArgs& a, bool adjoint, hipStream_t st) {
    constexpr int S = NC <= 8 ? 8 : 16;
    const int64_t threads = a.n * 2 * S;
    const int64_t grid = (threads + 255) / 256;
    if (adjoint) hipLaunchKernelGGL((k_flash_adjoint<NC, S>), dim3((unsigned)grid), dim3(256), 0, st, ctx->d_fluid, a);
    else hipLaunchKernelGGL((k_flash_tangent<NC, S>), dim3((unsigned)grid), dim3(256), 0, st, ctx->d_fluid, a);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

static int dispatch_ad(osr_ctx* ctx, const AdArgs& a, bool adjoint, hipStream_t st) {
    switch (ctx->nc) {
#define CASE(NC) case NC: return launch_ad<NC>(ctx, a, adjoint, st);
        OSR_NC_LIST(CASE)
#undef CASE
        default: return OSR_E_NC;
    }
}

}  // namespace osr

extern "C" {

int osr_flash_tangent_device(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P, const double* z,
                             const double* K, const int8_t* label, const double* dP, const double* dz, double* dx,
                             double* dy, double* dL, void* stream) {
    if (!ctx || n < 0) return OSR_E_ARG;
    if (n == 0) return OSR_OK;
    if (!P || !z || !K || !label || (!T && !(T0 > 0))) return OSR_E_ARG;
    osr::AdArgs a{};
    a.n = n; a.T = T; a.T0 = T0; a.P = P; a.z = z; a.K = K; a.label = label;
    a.dP = dP; a.dz = dz; a.dx = dx; a.dy = dy; a.dL = dL;
    return osr::dispatch_ad(ctx, a, false, (hipStream_t)stream);
}

int osr_flash_adjoint_device(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P, const double* z,
                             const double* K, const int8_t* label, const double* xbar, const double* ybar,
                             const double* Lbar, double* Pbar, double* zbar, void* stream) {
    if (!ctx || n < 0) return OSR_E_ARG;
    if (n == 0) return OSR_OK;
    if (!P || !z || !K || !label || (!T && !(T0 > 0))) return OSR_E_ARG;
    osr::AdArgs a{};
    a.n = n; a.T = T; a.T0 = T0; a.P = P; a.z = z; a.K = K; a.label = label;
    a.xbar = xbar; a.ybar = ybar; a.Lbar = Lbar; a.Pbar = Pbar; a.zbar = zbar;
    return osr::dispatch_ad(ctx, a, true, (hipStream_t)stream);
}

}  // extern "C"

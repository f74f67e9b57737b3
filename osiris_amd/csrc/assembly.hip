// assembly.hip -- residual and Jacobian assembly of the dual-porosity
// compositional model on gfx950, in the reference's variable-block AIJ
// layout (matrix blocks, then fracture blocks; block sizes nc+3 for single-
// phase and 2nc+4 for two-phase cells).
//
// Reference arithmetic restated (compsim/...):
//   property sweep  PetscAUX.cpp:500-547 order: CalcIFT Do_Cycle.cpp:2438-2456,
//                   CalcSatFs :99-259 / fCalcSatFs :261-437, Calc_Phase_P
//                   :1418-1446, Calc_Visco :438-616 / fCalc_Visco :619-799,
//                   CalcDiffusion Mass_Tran.cpp:21-62, CalcPhaseGrav
//                   Do_Cycle.cpp:2022-2057, Pre_Mul Do_Time_Step.cpp:54-84
//   face terms      Calc_Transes Do_Cycle.cpp:800-1108 / fCalc_Transes :1109-1417
//   residual rows   IteratorFunction Petsc_Functions.cpp:10-6896
//   layout          PJac Init_Calcs.cpp:1013-1116
//
// GPU structure: one lane per block (continuum, cell).  k_props writes the
// cell properties structure-of-arrays ([field][2n]) so the stencil reads of
// the residual kernel are coalesced along x.  The residual of a block is the
// sum of separable pieces -- accumulation, six face fluxes, the matrix-
// fracture transfer, the fugacity and constraint rows.  The Jacobian takes
// the property derivatives of every block once (k_dprops) and then
// differentiates only the pieces that touch each (row block, dependency
// block) pair, one lane per column, forward mode in registers (k_jacobian);
// entries go straight into the CSR arrays, so no atomics and the result is
// bitwise reproducible.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <type_traits>

#include "dual_dev.hpp"
#include "grp_dev.hpp"
#include "osr_internal.hpp"
#include "phase_dev.hpp"

namespace osr {

constexpr double SI_C = 9.869232667160128e-13;   // Functions.h:207
constexpr double WAT_M_RO = 55.5084;             // Functions.h:206
constexpr double G_ACC = 9.80665;                // Functions.h:198
constexpr double G_ACCG = 9.80665e-3;            // Functions.h:199
constexpr double IFTBASECASE = 50;               // Functions.h:279
constexpr double MISCIBLEIFT = 1;                // Functions.h:281
constexpr double IFTPOWER = 7;                   // Functions.h:280
constexpr double RELPERM0 = 0;                   // Functions.h:266
constexpr int MAXTAB = 32;

// Constants of one assembly context: EOS at the run temperature, the
// component data the property correlations need and the binary diffusion
// table (all T-only), read with wave-uniform loads.
struct AsmConsts {
    int nc, eos;
    double T, d1, d2;
    double a[NCMAX], b[NCMAX], sa[NCMAX];
    double bic[NCMAX * NCMAX];
    double mw[NCMAX], sqmw[NCMAX], vcm3[NCMAX], mu0[NCMAX], tc[NCMAX], pc[NCMAX], par[NCMAX];
    double v53[NCMAX], v23[NCMAX];
    double pdij[NCMAX * NCMAX];   // pDij (Mass_Tran.cpp:41-45), 0 on the diagonal
    double ipdij[NCMAX * NCMAX];  // 1 / pDij, 0 on the diagonal
    double pdlast[NCMAX];         // pDij of the last k != j (used when x_j == 1)
    double wk[NCMAX];             // Wilson exponent factor 5.37 (1 + omega_i)
    double swt[4 * MAXTAB], sgt[4 * MAXTAB];
    int nswt, nsgt;
    double refP, cpor, dcpor, dt, wat_ro, wat_mu, sigma, lf, cdiff;
    int const_diff;
};

struct AsmArgs {
    int nx, ny, nz, n;
    const AsmConsts* K;
    const double *dx, *dy, *dz, *h;
    const double *phi;       // [2n]
    const double *perm_m;    // [3n]
    const double *iftran;    // [2][3][n]
    const double *tor;       // [n]
    // state
    const int8_t* lab[2];
    const double* Po[2];
    const double* x[2];
    const double* y[2];
    const double* S[2];
    const double* rho_prev;  // [2n] or null
    const double* pre[2];    // [(nc+1) n] or null
    double pcoeff;
    // derived
    double* props;           // [NF][2n]
    const int64_t* boff;     // [2n+1]
    int xcd;                 // 1: XCD-contiguous workgroup order (wg_index)
    const double* perm_f;    // [3n] fracture permeabilities
    // wells (osr_asm_set_wells; nw = 0: none)
    int nw;
    const double* wsc;       // [2] simTime, Re_f
    const int* wi;           // [nw][5] type, Ix, Iy, first and last layer
    const double* wf;        // [nw][nc + 6]
    const int8_t* wcond;     // [nw] WellCondition
    const double* wre;       // [nw] Calc_Re of the well's block
    const double* winj;      // [nw][2nc + 6] FlashInjection of types 3 and 8
};

// Logical workgroup index.  The dispatcher deals workgroups round-robin to
// the chip's 8 XCDs (workgroup i to XCD i mod 8), each with its own L2, so
// consecutive workgroups -- neighbouring cells of the stencil -- land in
// different L2s.  With A.xcd set, XCD x walks one contiguous eighth of the
// grid instead: hardware workgroup i = 8 j + x is logical workgroup
// start_x + j, start_x = x q + min(x, r) for g = 8 q + r workgroups (a
// bijection on [0, g)).
__device__ __forceinline__ int64_t wg_index(const AsmArgs& A) {
    const unsigned b = blockIdx.x;
    if (!A.xcd) return b;
    const unsigned g = gridDim.x, q = g >> 3, r = g & 7, x = b & 7, j = b >> 3;
    return (int64_t)x * q + (x < r ? x : r) + j;
}

__device__ __forceinline__ int nfields(int nc) { return 16 + 2 * nc; }

// ---------------------------------------------------------------- EOS (T)
// Root of Z^3 + a1 Z^2 + a2 Z + a3 by the branches of Solve_Z
// (Globals.cpp:212-264): the largest root for the gas, the smallest positive
// one for the liquid.
__device__ __forceinline__ double zroot(double a1, double a2, double a3, bool gas) {
    using T = double;
    const T Q = (3.0 * a2 - a1 * a1) / 9.0;
    const T J = (9.0 * a1 * a2 - 27.0 * a3 - 2.0 * a1 * a1 * a1) / 54.0;
    const T D = Q * Q * Q + J * J;
    T Z;
    if (D < 0) {
        const T r = 2.0 * fsqrt(-Q);
        const T s = a1 / 3.0;
        const T t = facos(J / fsqrt(-Q * Q * Q)) / 3.0;
        const T x1 = r * fcos(t) - s;
        const T x2 = r * fcos(t + 2.0 * 3.1415926535897932384626433 / 3) - s;
        const T x3 = r * fcos(t + 4.0 * 3.1415926535897932384626433 / 3) - s;
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
        const T t = fcbrt(J);
        const T x1 = 2.0 * t - a1 / 3.0;
        const T x2 = -t - a1 / 3.0;
        if (gas) Z = (x2 >= x1) ? x2 : x1;
        else Z = (x2 <= x1 && x2 > 0) ? x2 : x1;
    } else {
        const T sd = fsqrt(D);
        Z = fcbrt(J + sd) + fcbrt(J - sd) - a1 / 3.0;
    }
    return Z;
}

// Forward mode: the value from the branches above, the derivative from the
// implicit function theorem on f(Z) = 0, dZ = -(Z^2 da1 + Z da2 + da3) / f'(Z),
// as the reference's hand Jacobian takes dZ (Do_Cycle.cpp:438-616).
// Differentiating the explicit formulas instead fails where they cancel: in
// the one-root branch with Q -> 0, J - sqrt(D) rounds to exactly 0 and its
// cube root has an infinite slope.
template <int W>
__device__ __forceinline__ DualV<W> zroot(const DualV<W>& a1, const DualV<W>& a2, const DualV<W>& a3, bool gas) {
    const double Z = zroot(a1.v, a2.v, a3.v, gas);
    const double inv = -1.0 / ((3.0 * Z + 2.0 * a1.v) * Z + a2.v);
    DualV<W> r = (Z * Z * a1 + Z * a2 + a3) * inv;
    r.v = Z;
    return r;
}

// Forward mode with the root's value already known (zk > 0, e.g. from the
// property table): only the implicit-function derivative is formed.
__device__ __forceinline__ double zroot_k(double a1, double a2, double a3, bool gas, double zk) {
    return zroot(a1, a2, a3, gas);
}
template <int W>
__device__ __forceinline__ DualV<W> zroot_k(const DualV<W>& a1, const DualV<W>& a2, const DualV<W>& a3, bool gas,
                                            double zk) {
    if (!(zk > 0)) return zroot(a1, a2, a3, gas);
    const double inv = -1.0 / ((3.0 * zk + 2.0 * a1.v) * zk + a2.v);
    DualV<W> r = (zk * zk * a1 + zk * a2 + a3) * inv;
    r.v = zk;
    return r;
}

template <int NC, class T>
__device__ __forceinline__ void mixT(const AsmConsts& K, const T* x, T& Ca, T& Cb, T* t) {
    T u[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) u[j] = x[j] * K.sa[j];
    Ca = 0.0;
    Cb = 0.0;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        T s = 0.0;
#pragma unroll
        for (int j = 0; j < NC; j++) s += u[j] * K.bic[i * NC + j];
        Ca += u[i] * s;
        Cb += x[i] * K.b[i];
        if (t) t[i] = s * K.sa[i];
    }
}

template <int NC, class T>
__device__ __forceinline__ T zphase(const AsmConsts& K, const T* x, const T& P, bool gas, double zk = 0.0) {
    T Ca, Cb;
    mixT<NC, T>(K, x, Ca, Cb, (T*)nullptr);
    const double rt = RGAS * K.T;
    const T A = Ca * P / (rt * rt);
    const T B = Cb * P / rt;
    const bool srk = K.eos == OSR_EOS_SRK;
    const T U = srk ? B : 2.0 * B;
    const T W = srk ? T(0.0) : B;
    return zroot_k(-(1.0 + B - U), A - B * U - U - W * W, -(A * B - B * W * W - W * W), gas, zk);
}

// ------------------------------------------------------- cell properties
template <int NC, class T>
struct PropsT {
    T porpor, Pw, Po, Pg, bift;
    T ro[2], mu[2], gam[2];
    T kr[3], pcw, pcg;
    T D[2][NC];
};

// slope (s - s1) + v1 with the product rounded before the sum, as the
// reference evaluates it: not fused into one FMA, so a saturation on the
// segment's left node gives that node's value exactly (a zero relative
// permeability stays zero, where the reference divides by it)
__device__ __forceinline__ double seg_rn(double slope, double s, double s1, double v1) {
#pragma clang fp contract(off)
    return slope * (s - s1) + v1;
}
template <int W>
__device__ __forceinline__ DualV<W> seg_rn(double slope, const DualV<W>& s, double s1, double v1) {
    DualV<W> r;
    r.v = seg_rn(slope, s.v, s1, v1);
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = slope * s.d[q];
    return r;
}

template <class T>
__device__ __forceinline__ void table(const double* tb, int nr, const T& s, T& kr, T& kro, T& pc) {
    int r = 0;
    while (r < nr && !(s < tb[4 * r])) r++;
    if (r == 0) {
        kr = T(RELPERM0);
        pc = T(tb[3]);
        kro = T(tb[2]);
    } else if (r == nr) {
        kr = T(tb[4 * (nr - 1) + 1]);
        pc = T(tb[4 * (nr - 1) + 3]);
        kro = T(tb[4 * (nr - 1) + 2]);
    } else {
        // slope as a quotient, as the reference writes it (Do_Cycle.cpp:183-185),
        // so a saturation on a table node gives the node value exactly
        const double s0 = tb[4 * (r - 1)], s1 = tb[4 * r];
        kr = seg_rn((tb[4 * r + 1] - tb[4 * (r - 1) + 1]) / (s1 - s0), s, s1, tb[4 * r + 1]);
        kro = seg_rn((tb[4 * r + 2] - tb[4 * (r - 1) + 2]) / (s1 - s0), s, s1, tb[4 * r + 2]);
        pc = seg_rn((tb[4 * r + 3] - tb[4 * (r - 1) + 3]) / (s1 - s0), s, s1, tb[4 * r + 3]);
    }
}

template <int NC, class T>
__device__ __forceinline__ void viscoT(const AsmConsts& K, const T* x, const T& P, bool gas, T& ro, T& mu,
                                       double zk = 0.0) {
    const T Z = zphase<NC, T>(K, x, P, gas, zk);
    ro = P / (Z * (RGAS * K.T));
    T s1 = 0.0, s2 = 0.0, rv = 0.0, tcm = 0.0, pcm = 0.0, mwm = 0.0;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const T tm = x[i] * K.sqmw[i];
        s1 += tm * K.mu0[i];
        s2 += tm;
        rv += x[i] * K.vcm3[i];
        tcm += x[i] * K.tc[i];
        pcm += x[i] * (K.pc[i] / 101325);
        mwm += x[i] * K.mw[i];
    }
    const T Er = ro * rv;
    const T etha = fpow(tcm, 1.0 / 6) / (fsqrt(mwm) * fpow(pcm, 2.0 / 3));
    const T r1 = 0.10230 + 0.023364 * Er + 0.058533 * Er * Er - 0.040758 * Er * Er * Er +
                 0.0093324 * Er * Er * Er * Er;
    const T r2 = r1 * r1;
    mu = s1 / s2 + (r2 * r2 - 1e-4) / etha;
}

template <int NC, class T>
__device__ __forceinline__ void diffT(const AsmConsts& K, const T* x, const T& ro, T* D) {
    T s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int j = 0; j < NC; j++) {
        s1 += x[j] * K.v53[j];
        s2 += x[j] * K.v23[j];
    }
    const T rr = ro * s1 / s2;
    const T roRD = (0.99589 + 0.096016 * rr - 0.22035 * rr * rr + 0.032874 * rr * rr * rr) / ro;
    // sum_k x_k / (pDij roRD) (Mass_Tran.cpp:40-49) as (sum_k x_k / pDij) / roRD,
    // with 1 / pDij tabulated: one division per component instead of nc
#pragma unroll
    for (int j = 0; j < NC; j++) {
        T s = 0.0;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            if (k == j) continue;
            s += x[k] * K.ipdij[j * NC + k];
        }
        const T dt = (1.0 - x[j]) * roRD / (s * 10000.0);
        T dv;
        if (x[j] == 1.0) dv = roRD * K.pdlast[j];
        else if (dt < 0 || s == 0.0) dv = T(0.0);
        else dv = dt;
        if (K.const_diff) dv = T(K.cdiff);
        D[j] = dv;
    }
}

// zl, zg: the phase roots when already known (forward mode only; 0 = solve)
template <int NC, class T>
__device__ __forceinline__ void propsT(const AsmArgs& A, int k, int c, int lab, const T& Po, const T* x,
                                       const T* y, const T* S, double rpo, double rpg, PropsT<NC, T>& p,
                                       double zl = 0.0, double zg = 0.0) {
    const AsmConsts& K = *A.K;
    const double phi = A.phi[k * A.n + c];
    const T dP = Po - K.refP;
    p.Po = Po;
    p.porpor = phi * (1.0 + K.cpor * dP + K.dcpor * dP * dP);
    if (k == 0) {
        if (lab == 0) {
            T ps = 0.0;
#pragma unroll
            for (int i = 0; i < NC; i++) ps += K.par[i] * (x[i] * rpo - y[i] * rpg);
            p.bift = fpow(ps * 1e-6, 3.88);
        } else {
            p.bift = T(IFTBASECASE);
        }
        T krw, krow, pcw, krg, krog, pcg;
        table<T>(K.swt, K.nswt, S[0], krw, krow, pcw);
        table<T>(K.sgt, K.nsgt, S[2], krg, krog, pcg);
        const double kro0 = K.swt[2];
        const T AA = krw + krow / kro0;
        const T BB = krg + krog / kro0;
        T kro = kro0 * (AA * BB - krw - krg);
        if (p.bift < MISCIBLEIFT) {
            const T F = fpow(p.bift / IFTBASECASE, 1.0 / IFTPOWER);
            kro = F * kro + (1.0 - F) * S[1] * (1 - K.swt[0]);
            krg = F * krg + (1.0 - F) * S[2] * (1 - K.swt[0]);
        }
        if (kro < 0) kro = T(RELPERM0);
        if (kro > 1) kro = T(1.0);
        p.kr[0] = krw;
        p.kr[1] = kro;
        p.kr[2] = krg;
        p.pcw = pcw;
        p.pcg = pcg;
        p.Pw = Po - pcw;
        if (p.bift < MISCIBLEIFT) p.Pg = Po + pcg * p.bift / IFTBASECASE;
        else p.Pg = Po + pcg;
    } else {
        p.bift = T(IFTBASECASE);
        p.kr[0] = S[0];
        p.kr[1] = S[1];
        p.kr[2] = S[2];
        p.pcw = T(0.0);
        p.pcg = T(0.0);
        p.Pw = Po;
        p.Pg = Po;
    }
    viscoT<NC, T>(K, x, p.Po, false, p.ro[0], p.mu[0], zl);
    viscoT<NC, T>(K, y, p.Pg, true, p.ro[1], p.mu[1], zg);
    diffT<NC, T>(K, x, p.ro[0], p.D[0]);
    diffT<NC, T>(K, y, p.ro[1], p.D[1]);
    T ml = 0.0, mg = 0.0;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        ml += x[i] * K.mw[i];
        mg += y[i] * K.mw[i];
    }
    p.gam[0] = G_ACCG * p.ro[0] * ml;
    p.gam[1] = G_ACCG * p.ro[1] * mg;
}

// the two phase roots of block blk from the property table (Z = P / (rho R T))
__device__ __forceinline__ void table_roots(const AsmArgs& A, int blk, double& zl, double& zg) {
    const int64_t s2 = 2 * (int64_t)A.n;
    const double* q = A.props + blk;
    const double rt = RGAS * A.K->T;
    zl = q[2 * s2] / (q[5 * s2] * rt);
    zg = q[3 * s2] / (q[6 * s2] * rt);
}

// state of one block as T
template <int NC, class T>
struct VarsT {
    int lab;
    T Po;
    T x[NC], y[NC], S[3];
};

template <int NC>
__device__ __forceinline__ void load_vars(const AsmArgs& A, int k, int c, VarsT<NC, double>& v) {
    const int n = A.n;
    v.lab = A.lab[k][c];
    v.Po = A.Po[k][c];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        v.x[i] = A.x[k][(int64_t)i * n + c];
        v.y[i] = A.y[k][(int64_t)i * n + c];
    }
#pragma unroll
    for (int j = 0; j < 3; j++) v.S[j] = A.S[k][(int64_t)j * n + c];
}

template <int NC>
__device__ __forceinline__ void load_props(const AsmArgs& A, int k, int c, PropsT<NC, double>& p) {
    const int64_t s = 2 * (int64_t)A.n;
    const double* q = A.props + (int64_t)k * A.n + c;
    p.porpor = q[0 * s]; p.Pw = q[1 * s]; p.Po = q[2 * s]; p.Pg = q[3 * s]; p.bift = q[4 * s];
    p.ro[0] = q[5 * s]; p.ro[1] = q[6 * s]; p.mu[0] = q[7 * s]; p.mu[1] = q[8 * s];
    p.gam[0] = q[9 * s]; p.gam[1] = q[10 * s];
    p.kr[0] = q[11 * s]; p.kr[1] = q[12 * s]; p.kr[2] = q[13 * s]; p.pcw = q[14 * s]; p.pcg = q[15 * s];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        p.D[0][i] = q[(16 + i) * s];
        p.D[1][i] = q[(16 + NC + i) * s];
    }
}

template <int NC>
__device__ __forceinline__ void rho_prev_of(const AsmArgs& A, int c, double& ro, double& rg) {
    if (A.rho_prev) {
        ro = A.rho_prev[c];
        rg = A.rho_prev[A.n + c];
    } else {
        ro = rg = 0;
    }
}

// k_props: one lane per block; pass 0 (without previous densities) writes the
// zero-density sweep whose densities CalcIFT reads in pass 1.
template <int NC>
__global__ __launch_bounds__(256) void k_props(AsmArgs A) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * A.n) return;
    const int k = blk / A.n, c = blk % A.n;
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    double ro = 0, rg = 0;
    rho_prev_of<NC>(A, c, ro, rg);
    PropsT<NC, double> p;
    propsT<NC, double>(A, k, c, v.lab, v.Po, v.x, v.y, v.S, ro, rg, p);
    const int64_t s = 2 * (int64_t)A.n;
    double* q = A.props + blk;
    q[0 * s] = p.porpor; q[1 * s] = p.Pw; q[2 * s] = p.Po; q[3 * s] = p.Pg; q[4 * s] = p.bift;
    q[5 * s] = p.ro[0]; q[6 * s] = p.ro[1]; q[7 * s] = p.mu[0]; q[8 * s] = p.mu[1];
    q[9 * s] = p.gam[0]; q[10 * s] = p.gam[1];
    q[11 * s] = p.kr[0]; q[12 * s] = p.kr[1]; q[13 * s] = p.kr[2]; q[14 * s] = p.pcw; q[15 * s] = p.pcg;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        q[(16 + i) * s] = p.D[0][i];
        q[(16 + NC + i) * s] = p.D[1][i];
    }
}

// Copy the first-pass matrix densities into the [2n] previous-density buffer.
__global__ void k_take_rho(const double* props, int n, double* rp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int64_t s = 2 * (int64_t)n;
    rp[c] = props[5 * s + c];
    rp[n + c] = props[6 * s + c];
}

// ------------------------------------------------------------ geometry
// Per-cell geometry in scalars only (no arrays indexed by a run-time face
// number, which would put the struct in scratch memory).
struct Geo {
    int c, ix, iy, iz;
    double ds0, ds1, ds2;   // own sizes
    double h, Vb;
};

__device__ __forceinline__ Geo geom(const AsmArgs& A, int c) {
    Geo g;
    const int nxy = A.nx * A.ny;
    g.c = c;
    g.iz = c / nxy;
    const int rem = c - g.iz * nxy;
    g.iy = rem / A.nx;
    g.ix = rem - g.iy * A.nx;
    g.ds0 = A.dx[g.ix];
    g.ds1 = A.dy[g.iy];
    g.ds2 = A.dz[g.iz];
    g.h = A.h[g.iz];
    g.Vb = g.ds0 * g.ds1 * g.ds2;
    return g;
}

// neighbour cell across face q (+x -x +y -y +z -z), -1 outside the grid
__device__ __forceinline__ int nb_of(const AsmArgs& A, const Geo& G, int q) {
    const int nxy = A.nx * A.ny;
    switch (q) {
        case 0: return G.ix < A.nx - 1 ? G.c + 1 : -1;
        case 1: return G.ix > 0 ? G.c - 1 : -1;
        case 2: return G.iy < A.ny - 1 ? G.c + A.nx : -1;
        case 3: return G.iy > 0 ? G.c - A.nx : -1;
        case 4: return G.iz < A.nz - 1 ? G.c + nxy : -1;
        default: return G.iz > 0 ? G.c - nxy : -1;
    }
}

struct FaceGeo {
    int o;          // neighbour cell (or -1)
    double dl, dn;  // own and neighbour size along the face normal
    double hn;      // neighbour block height (own for x/y faces)
    double area;
};

__device__ __forceinline__ FaceGeo face_geo(const AsmArgs& A, const Geo& G, int q) {
    FaceGeo f;
    f.o = nb_of(A, G, q);
    const bool in = f.o >= 0;
    switch (q >> 1) {
        case 0:
            f.dl = G.ds0;
            f.dn = in ? A.dx[G.ix + ((q & 1) ? -1 : 1)] : 0;
            f.hn = G.h;
            f.area = G.ds1 * G.ds2;
            break;
        case 1:
            f.dl = G.ds1;
            f.dn = in ? A.dy[G.iy + ((q & 1) ? -1 : 1)] : 0;
            f.hn = G.h;
            f.area = G.ds0 * G.ds2;
            break;
        default:
            f.dl = G.ds2;
            f.dn = in ? A.dz[G.iz + ((q & 1) ? -1 : 1)] : 0;
            f.hn = in ? A.h[G.iz + ((q & 1) ? -1 : 1)] : 0;
            f.area = G.ds0 * G.ds1;
            break;
    }
    return f;
}

// ------------------------------------------------------- residual pieces
// All pieces add into a RowAcc (the block's residual, F = LSide - RSide, so
// inflow terms are subtracted), indexed statically so it stays in registers:
//   bal[0..NC-2] component balances, bal[NC-1] total hydrocarbon, bal[NC]
//   water -- reference rows r0 + i (r0 = NC two-phase, 0 single-phase);
//   loc[] the block-local rows -- two-phase: fugacity i -> row i, then
//   sum x, sum y, sum S -> rows 2NC+1..2NC+3; single-phase: sum x|y, sum S
//   -> rows NC+1, NC+2 (IteratorFunction's row order).
template <int NC, class T>
struct RowAcc {
    T bal[NC + 1];
    T loc[NC + 3];
    __device__ __forceinline__ RowAcc() {
#pragma unroll
        for (int i = 0; i <= NC; i++) bal[i] = T(0.0);
#pragma unroll
        for (int i = 0; i < NC + 3; i++) loc[i] = T(0.0);
    }
};

// reference row of local row i
__device__ __forceinline__ int loc_row(int lab, int nc, int i) { return (lab == 0 && i < nc) ? i : nc + 1 + i; }
__device__ __forceinline__ int nloc(int lab, int nc) { return lab == 0 ? nc + 3 : 2; }

// Flux across face q between self (s) and neighbour (o): convective and
// gravity terms of every component, total and water row, and the Fickian
// flux of each component (Calc_Transes + IteratorFunction face terms).
//
// MIRROR also accumulates, into mb[0..NC], the same face's terms of the
// NEIGHBOUR's balance rows.  Every convective, gravity and water term is
// antisymmetric between the two rows: the potential, the upstream choice,
// the transmissibility and the mixed gravity are formed in (a, b) order
// whichever cell is "self", and the pressure and height differences change
// sign exactly.  The one exception is the Fickian upstream coefficient:
// each side takes the other's D_i only where the other's composition is
// strictly larger, so where x_s == x_o each row takes its own D_i; the
// neighbour's row therefore forms its own coefficient below.
template <int NC, class TS, class TO, bool MIRROR = false>
__device__ __forceinline__ void face_piece(const AsmArgs& A, int k, const Geo& G, const FaceGeo& FG, int q, int c,
                                           int o, const VarsT<NC, TS>& vs, const PropsT<NC, TS>& ps,
                                           const VarsT<NC, TO>& vo, const PropsT<NC, TO>& po,
                                           RowAcc<NC, decltype(TS() + TO())>& R, bool fick = true,
                                           decltype(TS() + TO())* mb = nullptr, bool conv = true) {
    using T = decltype(TS() + TO());
    const AsmConsts& K = *A.K;
    const int d = q >> 1;
    const bool self_is_a = (q & 1) == 0;
    const int ca = self_is_a ? c : o;
    const double Kf = A.iftran[((int64_t)k * 3 + d) * A.n + ca];
    const double da = self_is_a ? FG.dl : FG.dn;
    const double db = self_is_a ? FG.dn : FG.dl;
    const double ha = self_is_a ? G.h : FG.hn;
    const double hb = self_is_a ? FG.hn : G.h;
    const double dh = FG.hn - G.h;
    // water (upstream by water potential)
    if (conv) {
        const T pot = (self_is_a ? (ps.Pw - po.Pw) : (po.Pw - ps.Pw)) + G_ACC * K.wat_ro * (ha - hb);
        const bool up_a = pot > 0;
        const bool up_self = (up_a == self_is_a);
        const T tw = (Kf * (K.wat_ro * WAT_M_RO / K.wat_mu)) * (up_self ? T(ps.kr[0]) : T(po.kr[0]));
        const T fw = tw * (po.Pw - ps.Pw) + G_ACC * K.wat_ro * tw * dh;
        R.bal[NC] -= fw;
        if constexpr (MIRROR) mb[NC] += fw;
    }
    // oil and gas
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
        if (!conv) break;
        const int absent = ph == 0 ? -1 : 1;
        const double cs = vs.lab == absent ? 0.0 : 1.0;
        const double co = vo.lab == absent ? 0.0 : 1.0;
        const double c1 = self_is_a ? cs : co, c2 = self_is_a ? co : cs;
        const T gs = T(ps.gam[ph]);
        const T go = T(po.gam[ph]);
        const T gaa = self_is_a ? gs : go, gbb = self_is_a ? go : gs;
        T mg;
        if (c1 == 0 && c2 == 0) mg = T(0.0);
        else mg = (gaa * da * c1 + gbb * db * c2) / (c1 * da + db * c2);
        const TS Psel = ph == 0 ? ps.Po : ps.Pg;
        const TO Poth = ph == 0 ? po.Po : po.Pg;
        const T pot = (self_is_a ? (Psel - Poth) : (Poth - Psel)) + mg * (ha - hb);
        const bool up_a = pot > 0;
        const bool up_self = (up_a == self_is_a);
        const T ros = T(ps.ro[ph]), roo = T(po.ro[ph]);
        const T mus = T(ps.mu[ph]), muo = T(po.mu[ph]);
        const T kru = up_self ? T(ps.kr[1 + ph]): T(po.kr[1 + ph]);
        const T roa = self_is_a ? ros : roo, rob = self_is_a ? roo : ros;
        const T mua = self_is_a ? mus : muo, mub = self_is_a ? muo : mus;
        T t = Kf * (roa * da * c1 + rob * db * c2) / (mua * da * c1 + mub * db * c2) * kru;
        T tg = mg * t;
        if (c1 == 0 && c2 == 0) {
            t = T(0.0);
            tg = T(0.0);
        }
        const T dPp = Poth - Psel;
        const T base = t * dPp + tg * dh;
        // total hydrocarbon row
        R.bal[NC - 1] -= base;
        if constexpr (MIRROR) mb[NC - 1] += base;
#pragma unroll
        for (int i = 0; i < NC - 1; i++) {
            const T cu = up_self ? (ph == 0 ? T(vs.x[i]): T(vs.y[i]))
                                 : (ph == 0 ? T(vo.x[i]): T(vo.y[i]));
            if constexpr (MIRROR) {
                const T cb = cu * base;
                R.bal[i] -= cb;
                mb[i] += cb;
            } else {
                R.bal[i] -= cu * base;
            }
        }
    }
    if (!fick) return;
    // Fickian inter-cell flux of the components (tortuosity 1 in the fracture)
    const double dl = FG.dl, dn = FG.dn;
    const double NN = 2.0 * FG.area / (dn + dl);
    const double tn = k == 0 ? A.tor[o] : 1.0, ts = k == 0 ? A.tor[c] : 1.0;
    const bool zerophi = A.phi[k * A.n + c] < 1e-10 || A.phi[k * A.n + o] < 1e-10;
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
        const int absent = ph == 0 ? -1 : 1;
        const double Cf = vo.lab == absent ? 0.0 : 1.0;
        const double dcs = vs.lab == absent ? 0.0 : 1.0;
        // neither cell has the phase: every coefficient is 0 and the rows
        // (finite compositions) are unchanged
        if (Cf == 0 && dcs == 0) continue;
        const T AA = (po.ro[ph] * (dn * Cf) + ps.ro[ph] * (dl * dcs)) *
                     ((vo.S[1 + ph] * (dn * Cf) + vs.S[1 + ph] * (dl * dcs)) / (dn + dl));
        const T BB = (tn * dn * Cf) / po.porpor + (ts * dl * dcs) / ps.porpor;
        const T MN = AA / BB * NN;
        if (zerophi) {
            // a zero-porosity cell takes D = 0: the coefficient is MN * 0,
            // NaN where MN is (0/0 for a neighbour lacking the phase), as in
            // the reference
            const T Dt = MN * T(0.0);
#pragma unroll
            for (int i = 0; i < NC - 1; i++) {
                const TS xs = ph == 0 ? vs.x[i] : vs.y[i];
                const TO xn = ph == 0 ? vo.x[i] : vo.y[i];
                R.bal[i] -= Dt * (xn - xs);
                if constexpr (MIRROR) mb[i] += Dt * (xn - xs);
            }
            continue;
        }
#pragma unroll
        for (int i = 0; i < NC - 1; i++) {
            const TS xs = ph == 0 ? vs.x[i] : vs.y[i];
            const TO xn = ph == 0 ? vo.x[i] : vo.y[i];
            const T Dup = (xn > xs) ? T(po.D[ph][i]): T(ps.D[ph][i]);
            R.bal[i] -= (MN * Dup) * (xn - xs);
            if constexpr (MIRROR) {
                // the neighbour's row: upstream by its own comparison
                const T Dum = (xs > xn) ? T(ps.D[ph][i]): T(po.D[ph][i]);
                mb[i] += (MN * Dum) * (xn - xs);
            }
        }
    }
}

// Matrix-fracture transfer, as seen from continuum k (sign + for matrix
// rows, - for fracture rows), Petsc_Functions.cpp:5257-5315 / :731-806.
template <int NC, class TM, class TF>
__device__ __forceinline__ void transfer_piece(const AsmArgs& A, int k, const Geo& G, int c,
                                               const VarsT<NC, TM>& vm, const PropsT<NC, TM>& pm,
                                               const VarsT<NC, TF>& vf, const PropsT<NC, TF>& pf,
                                               RowAcc<NC, decltype(TM() + TF())>& R, bool diffuse = true) {
    using T = decltype(TM() + TF());
    const AsmConsts& K = *A.K;
    const int n = A.n;
    const double Vb = G.Vb;
    const double dx = G.ds0, dy = G.ds1, dz = G.ds2;
    const double sf2 = 4 * K.sigma * (A.perm_m[c] / (dx * dx) + A.perm_m[n + c] / (dy * dy) +
                                      A.perm_m[2 * n + c] / (dz * dz));
    const bool off = K.const_diff && K.cdiff == 0;
    const double sf1 = off ? 0 : 4 * K.sigma * (1 / (dx * dx) + 1 / (dy * dy) + 1 / (dz * dz));
    const double sf3 = off ? 0 : 4 * K.sigma * ((1 / dx) + (1 / dy) + (1 / dz)) / K.lf;
    const double sgn = k == 0 ? 1.0 : -1.0;
    const T kro = pm.Po >= pf.Po ? T(pm.kr[1]): T(pf.kr[1]);
    const T krg = pm.Pg >= pf.Pg ? T(pm.kr[2]): T(pf.kr[2]);
    const T krw = pm.Pw >= pf.Pw ? T(pm.kr[0]): T(pf.kr[0]);
    const T sp = pm.porpor + pf.porpor;
    const T muo = (pm.mu[0] * pm.porpor + pf.mu[0] * pf.porpor) / sp;
    const T mug = (pm.mu[1] * pm.porpor + pf.mu[1] * pf.porpor) / sp;
    const T convo = (SI_C * sf2 * Vb) * pm.ro[0] * kro / muo * (pf.Po - pm.Po);
    const T convg = (SI_C * sf2 * Vb) * pm.ro[1] * krg / mug * ((pf.Pg - pm.Pg) + (pf.pcg - pm.pcg));
    R.bal[NC - 1] -= sgn * (convo + convg);
    const T convw = (SI_C * K.wat_ro * WAT_M_RO * sf2 * Vb / K.wat_mu) * krw * ((pf.Pw - pm.Pw) + (pm.pcw - pf.pcw));
    R.bal[NC] -= sgn * convw;
    const int lm = vm.lab, lf = vf.lab;
    const double tor = A.tor[c];
    if (!diffuse) {
#pragma unroll
        for (int i = 0; i < NC - 1; i++) R.bal[i] -= (vm.x[i] * convo + vm.y[i] * convg) * sgn;
        return;
    }
    const T mo0 = (sf1 * Vb / tor) * pm.porpor * vm.S[1] * pm.ro[0];
    const T mg0 = (sf1 * Vb / tor) * pm.porpor * vm.S[2] * pm.ro[1];
    const T fo0 = (sf3 * Vb / tor) * pf.porpor * vf.S[1] * pf.ro[0];
    const T fg0 = (sf3 * Vb / tor) * pf.porpor * vf.S[2] * pf.ro[1];
#pragma unroll
    for (int i = 0; i < NC - 1; i++) {
        T rs = (vm.x[i] * convo + vm.y[i] * convg) * sgn;
        const T mo = mo0 * pm.D[0][i], mg = mg0 * pm.D[1][i];
        const T dxo = vf.x[i] - vm.x[i], dyo = vf.y[i] - vm.y[i];
        if (k == 0) {
            const T exo = T(-vm.x[i]), eyo = T(-vm.y[i]);
            if (lm == 1 && lf == 1) rs += mo * dxo;
            if (lm == -1 && lf == -1) rs += mg * dyo;
            if (lm == 0 && lf == 0) { rs += mo * dxo; rs += mg * dyo; }
            if (lm == 1 && lf == -1) rs += mo * exo;
            if (lm == -1 && lf == 1) rs += mg * eyo;
            if (lm == 0 && lf == 1) { rs += mg * eyo; rs += mo * dxo; }
            if (lm == 0 && lf == -1) { rs += mg * dyo; rs += mo * exo; }
            if (lm == 1 && lf == 0) { rs += mo * exo; rs += mo * dxo; }
            if (lm == -1 && lf == 0) { rs += mg * dyo; rs += mg * eyo; }
        } else {
            const T fo = fo0 * pf.D[0][i], fg = fg0 * pf.D[1][i];
            const T fxo = T(vf.x[i]), fyo = T(vf.y[i]);
            if (lm == 1 && lf == 1) rs -= mo * dxo;
            if (lm == -1 && lf == -1) rs -= mg * dyo;
            if (lm == 0 && lf == 0) { rs -= mo * dxo; rs -= mg * dyo; }
            if (lm == 1 && lf == -1) rs -= fg * fyo;
            if (lm == -1 && lf == 1) rs -= fo * fxo;
            if (lm == 0 && lf == 1) { rs -= fo * fxo; rs -= mo * dxo; }
            if (lm == 0 && lf == -1) { rs -= mg * dyo; rs -= fg * fyo; }
            if (lm == 1 && lf == 0) { rs -= fg * fyo; rs -= mo * dxo; }
            if (lm == -1 && lf == 0) { rs -= mg * dyo; rs -= fo * fxo; }
        }
        R.bal[i] -= rs;
    }
}

// Accumulation (LSide) of the balances, the fugacity rows of a two-phase
// block and the constraint rows: everything that depends on the block's own
// unknowns only.
template <int NC, class T>
__device__ __forceinline__ void loc_piece(const AsmArgs& A, const VarsT<NC, T>& v, const PropsT<NC, T>& p,
                                          RowAcc<NC, T>& R);

// the accumulation of the balance rows alone
template <int NC, class T>
__device__ __forceinline__ void acc_piece(const AsmArgs& A, int k, int c, const Geo& G, const VarsT<NC, T>& v,
                                          const PropsT<NC, T>& p, RowAcc<NC, T>& R) {
    const AsmConsts& K = *A.K;
    const int n = A.n;
    const double f = G.Vb / K.dt;
    const double* pre = A.pre[k];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        T acc;
        if (i < NC - 1) acc = p.porpor * (v.x[i] * p.ro[0] * v.S[1] + v.y[i] * p.ro[1] * v.S[2]);
        else acc = p.porpor * (p.ro[0] * v.S[1] + p.ro[1] * v.S[2]);
        const double pr = pre ? pre[(int64_t)i * n + c] : dval(acc);
        R.bal[i] += f * (acc - pr);
    }
    {
        const T acc = p.porpor * (K.wat_ro * WAT_M_RO) * v.S[0];
        const double pr = pre ? pre[(int64_t)NC * n + c] : dval(acc);
        R.bal[NC] += f * (acc - pr);
    }
}

template <int NC, class T>
__device__ __forceinline__ void own_piece(const AsmArgs& A, int k, int c, const Geo& G, const VarsT<NC, T>& v,
                                          const PropsT<NC, T>& p, RowAcc<NC, T>& R) {
    acc_piece<NC, T>(A, k, c, G, v, p, R);
    loc_piece<NC, T>(A, v, p, R);
}

// The block-local rows: fugacity equalities of a two-phase block and the
// constraint rows (they read Po, Pg and the two phase densities only).
template <int NC, class T>
__device__ __forceinline__ void loc_piece(const AsmArgs& A, const VarsT<NC, T>& v, const PropsT<NC, T>& p,
                                          RowAcc<NC, T>& R) {
    const AsmConsts& K = *A.K;
    if (v.lab == 0) {
        T Cal, Cbl, Cag, Cbg, tl[NC], tg[NC];
        mixT<NC, T>(K, v.x, Cal, Cbl, tl);
        mixT<NC, T>(K, v.y, Cag, Cbg, tg);
        const double rt = RGAS * K.T;
        const T Al = Cal * p.Po / (rt * rt), Bl = Cbl * p.Po / rt;
        const T Ag = Cag * p.Pg / (rt * rt), Bg = Cbg * p.Pg / rt;
        const T Zl = p.Po / (p.ro[0] * rt), Zg = p.Pg / (p.ro[1] * rt);
        const T DDl = Al / (Bl * (K.d1 - K.d2)), DDg = Ag / (Bg * (K.d1 - K.d2));
        const T FFl = flog((Zl + K.d2 * Bl) / (Zl + K.d1 * Bl));
        const T FFg = flog((Zg + K.d2 * Bg) / (Zg + K.d1 * Bg));
        const T lzl = flog(Zl - Bl), lzg = flog(Zg - Bg);
        // b_i / Cb and 2 t_i / Ca through the reciprocals (one division per phase)
        const T iCal = 1.0 / Cal, iCbl = 1.0 / Cbl, iCag = 1.0 / Cag, iCbg = 1.0 / Cbg;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const T t1l = K.b[i] * iCbl, t1g = K.b[i] * iCbg;
            const T EEl = 2.0 * tl[i] * iCal - t1l, EEg = 2.0 * tg[i] * iCag - t1g;
            const T phiL = fexp(t1l * (Zl - 1.0) - lzl + DDl * EEl * FFl);
            const T phiG = fexp(t1g * (Zg - 1.0) - lzg + DDg * EEg * FFg);
            R.loc[i] += v.x[i] * p.Po * phiL - v.y[i] * p.Pg * phiG;
        }
        T sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            sx += v.x[i];
            sy += v.y[i];
        }
        R.loc[NC] += sx - 1.0;
        R.loc[NC + 1] += sy - 1.0;
        R.loc[NC + 2] += v.S[0] + v.S[1] + v.S[2] - 1.0;
    } else {
        T s = 0.0;
#pragma unroll
        for (int i = 0; i < NC; i++) s += (v.lab == 1 ? v.x[i] : v.y[i]);
        R.loc[0] += s - 1.0;
        R.loc[1] += v.S[0] + v.S[1] + v.S[2] - 1.0;
    }
}

// ------------------------------------------------------ residual kernel
template <int NC>
__global__ __launch_bounds__(256) void k_residual(AsmArgs A, double* F) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * A.n) return;
    const int n = A.n;
    const int k = blk / n, c = blk % n;
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    const int bs = v.lab == 0 ? 2 * NC + 4 : NC + 3;
    double* out = F + A.boff[blk];
    if (A.phi[k * n + c] < 1e-10) {
        for (int r = 0; r < bs; r++) out[r] = 0;
        return;
    }
    const int r0 = v.lab == 0 ? NC : 0;
    RowAcc<NC, double> R;
    const Geo G = geom(A, c);
    PropsT<NC, double> p;
    load_props<NC>(A, k, c, p);
    own_piece<NC, double>(A, k, c, G, v, p, R);
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const FaceGeo FG = face_geo(A, G, q);
        const int o = FG.o;
        if (o < 0) continue;
        VarsT<NC, double> vo;
        PropsT<NC, double> po;
        load_vars<NC>(A, k, o, vo);
        load_props<NC>(A, k, o, po);
        face_piece<NC, double, double>(A, k, G, FG, q, c, o, v, p, vo, po, R);
    }
    {
        VarsT<NC, double> vx;
        PropsT<NC, double> px;
        load_vars<NC>(A, 1 - k, c, vx);
        load_props<NC>(A, 1 - k, c, px);
        if (k == 0) transfer_piece<NC, double, double>(A, 0, G, c, v, p, vx, px, R);
        else transfer_piece<NC, double, double>(A, 1, G, c, vx, px, v, p, R);
    }
#pragma unroll
    for (int i = 0; i <= NC; i++) out[r0 + i] = R.bal[i];
    const int nl = nloc(v.lab, NC);
#pragma unroll
    for (int i = 0; i < NC + 3; i++)
        if (i < nl) out[loc_row(v.lab, NC, i)] = R.loc[i];
}

// ------------------------------------------------------ Jacobian pattern
// Dependency slots of a row block: 0 own, 1..6 face neighbours (+x -x +y -y
// +z -z), 7 the other continuum's cell.  Columns of a row are emitted in
// increasing block offset; component-balance and total rows take every
// column of every existing dependency, the water row the P and Sw columns of
// non-own dependencies, fugacity and constraint rows their own block.
__device__ __forceinline__ int dep_block(const AsmArgs& A, int k, const Geo& G, int c, int slot) {
    if (slot == 0) return k * A.n + c;
    if (slot == 7) return (1 - k) * A.n + c;
    const int o = nb_of(A, G, slot - 1);
    return o < 0 ? -1 : k * A.n + o;
}

__device__ __forceinline__ int bsize_of(const AsmArgs& A, int blk, int nc) {
    const int k = blk / A.n, c = blk % A.n;
    return A.lab[k][c] == 0 ? 2 * nc + 4 : nc + 3;
}

// nnz of each block's rows (for the row pointer scan)
template <int NC>
__global__ void k_jac_rownnz(AsmArgs A, int64_t* rownnz) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * A.n) return;
    const int n = A.n, k = blk / n, c = blk % n;
    const Geo G = geom(A, c);
    const int lab = A.lab[k][c];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    const int r0 = lab == 0 ? NC : 0;
    int full = 0, nother = 0;
    for (int s = 0; s < 8; s++) {
        const int d = dep_block(A, k, G, c, s);
        if (d < 0) continue;
        full += bsize_of(A, d, NC);
        if (s != 0) nother++;
    }
    int64_t* out = rownnz + A.boff[blk];
    if (A.phi[k * n + c] < 1e-10) {
        for (int r = 0; r < bs; r++) out[r] = 1;   // identity row
        return;
    }
    for (int r = 0; r < bs; r++) {
        int64_t v;
        if (r >= r0 && r < r0 + NC) v = full;
        else if (r == r0 + NC) v = bs + 2 * nother;
        else v = bs;
        out[r] = v;
    }
}

// Start of dependency block `d` within row r's column list.
__device__ __forceinline__ int64_t seg_start(const AsmArgs& A, int k, const Geo& G, int c, int row_kind,
                                             int d, int nc) {
    // row_kind: 0 mass (all deps), 1 water, 2 own-only
    int64_t pos = 0;
    const int64_t od = A.boff[d];
    for (int s = 0; s < 8; s++) {
        const int e = dep_block(A, k, G, c, s);
        if (e < 0 || e == d) continue;
        if (A.boff[e] >= od) continue;
        if (row_kind == 0) pos += bsize_of(A, e, nc);
        else if (row_kind == 1) pos += (s == 0) ? bsize_of(A, e, nc) : 2;
    }
    return pos;
}

template <int NC, int W>
__device__ __forceinline__ void seed_vars(const VarsT<NC, double>& v, int chunk, double pcoeff, VarsT<NC, DualV<W>>& out) {
    out.lab = v.lab;
    out.Po = DualV<W>(v.Po);
#pragma unroll
    for (int i = 0; i < NC; i++) {
        out.x[i] = DualV<W>(v.x[i]);
        out.y[i] = DualV<W>(v.y[i]);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) out.S[j] = DualV<W>(v.S[j]);
    // unknown order (InitGuess2, PetscAUX.cpp:892-957)
    int col = 0;
    const int base = chunk * W;
    auto mark = [&](DualV<W>& t, double sc) {
        const int rel = col - base;
        if (rel >= 0 && rel < W) t.d[rel] = sc;
        col++;
    };
    if (v.lab != -1)
        for (int i = 0; i < NC; i++) mark(out.x[i], 1.0);
    if (v.lab != 1)
        for (int i = 0; i < NC; i++) mark(out.y[i], 1.0);
    mark(out.Po, pcoeff);
    mark(out.S[0], 1.0);
    if (v.lab != -1) mark(out.S[1], 1.0);
    if (v.lab != 1) mark(out.S[2], 1.0);
}

// Jacobian in two kernels.  Every residual piece of a row block depends on a
// dependency block d through d's unknowns directly and through d's cell
// properties, and the properties of d are shared by the eight row blocks
// that touch d.  So:
//   k_dprops    one lane per (block d, unknown j): the property sweep of d in
//               forward mode with unknown j seeded (one direction per lane,
//               DualV<1>, everything in registers); writes d props / d u_j to
//               a workspace laid out [block][field][column];
//   k_jacobian  one lane per (row block, dependency slot, column j): the
//               pieces that touch the dependency, in DualV<1>, with the
//               dependency's properties rebuilt from the property table
//               (value) and the workspace (derivative).  The lanes of one
//               (row block, dependency) write consecutive CSR columns.
template <int NC>
struct JacShape {
    static constexpr int NF = 16 + 2 * NC;
    static constexpr int MAXC = 2 * NC + 4;
    // direct reduction: per direction, the NF property derivatives and the
    // seeds of the block's variables in variable order (x, y, P, Sw, So, Sg)
    static constexpr int RF = NF + 2 * NC + 4;
};

template <int NC, class T>
__device__ __forceinline__ T prop_field(const PropsT<NC, T>& p, int f) {
    switch (f) {
        case 0: return p.porpor;
        case 1: return p.Pw;
        case 2: return p.Po;
        case 3: return p.Pg;
        case 4: return p.bift;
        case 5: return p.ro[0];
        case 6: return p.ro[1];
        case 7: return p.mu[0];
        case 8: return p.mu[1];
        case 9: return p.gam[0];
        case 10: return p.gam[1];
        case 11: return p.kr[0];
        case 12: return p.kr[1];
        case 13: return p.kr[2];
        case 14: return p.pcw;
        case 15: return p.pcg;
        default: return f < 16 + NC ? p.D[0][f - 16] : p.D[1][f - 16 - NC];
    }
}

template <int NC>
__global__ __launch_bounds__(256) void k_dprops(AsmArgs A, double* dp) {
    using D1 = DualV<1>;
    constexpr int NF = JacShape<NC>::NF, MAXC = JacShape<NC>::MAXC;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (int64_t)2 * A.n * MAXC) return;
    const int blk = (int)(item / MAXC);
    const int j = (int)(item % MAXC);
    const int n = A.n, k = blk / n, c = blk % n;
    VarsT<NC, double> vd;
    load_vars<NC>(A, k, c, vd);
    const int bs = vd.lab == 0 ? 2 * NC + 4 : NC + 3;
    if (j >= bs) return;
    VarsT<NC, D1> sv;
    seed_vars<NC, 1>(vd, j, A.pcoeff, sv);
    double rpo = 0, rpg = 0;
    if (k == 0 && A.rho_prev) {
        rpo = A.rho_prev[c];
        rpg = A.rho_prev[n + c];
    }
    PropsT<NC, D1> sp;
    propsT<NC, D1>(A, k, c, sv.lab, sv.Po, sv.x, sv.y, sv.S, rpo, rpg, sp);
    double* o = dp + (int64_t)blk * NF * MAXC + j;
#pragma unroll
    for (int f = 0; f < NF; f++) o[f * MAXC] = prop_field<NC, D1>(sp, f).d[0];
}

template <int NC>
__device__ __forceinline__ void load_props_dual(const AsmArgs& A, const double* dp, int blk, int j,
                                                PropsT<NC, DualV<1>>& p, int ncol = JacShape<NC>::MAXC,
                                                int nfield = JacShape<NC>::NF) {
    const int64_t s = 2 * (int64_t)A.n;
    const double* q = A.props + blk;
    const double* g = dp + (int64_t)blk * nfield * ncol + j;
    auto F = [&](int f) {
        DualV<1> r;
        r.v = q[f * s];
        r.d[0] = g[f * ncol];
        return r;
    };
    p.porpor = F(0); p.Pw = F(1); p.Po = F(2); p.Pg = F(3); p.bift = F(4);
    p.ro[0] = F(5); p.ro[1] = F(6); p.mu[0] = F(7); p.mu[1] = F(8);
    p.gam[0] = F(9); p.gam[1] = F(10);
    p.kr[0] = F(11); p.kr[1] = F(12); p.kr[2] = F(13); p.pcw = F(14); p.pcg = F(15);
#pragma unroll
    for (int i = 0; i < NC; i++) {
        p.D[0][i] = F(16 + i);
        p.D[1][i] = F(16 + NC + i);
    }
}

// Column j of dependency d, for the rows of block blk, into the CSR.  Mass
// rows list every dependency in block-offset order, the water row lists its
// own block whole and only P, Sw of the others, local rows their own block.
template <int NC>
__device__ __forceinline__ void scatter_col(const AsmArgs& A, const Geo& G, int k, int c, int blk, int lab,
                                            int slot, int d, int j, const RowAcc<NC, DualV<1>>& R,
                                            const int64_t* rowptr, int32_t* cols, double* vals) {
    const int n = A.n;
    const int r0 = lab == 0 ? NC : 0;
    const int64_t rb = A.boff[blk];
    const int32_t col = (int32_t)(A.boff[d] + j);
    const int64_t s_mass = seg_start(A, k, G, c, 0, d, NC);
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const int64_t pos = rowptr[rb + r0 + i] + s_mass + j;
        cols[pos] = col;
        vals[pos] = R.bal[i].d[0];
    }
    const int64_t s_water = seg_start(A, k, G, c, 1, d, NC);
    if (slot == 0) {
        const int64_t pos = rowptr[rb + r0 + NC] + s_water + j;
        cols[pos] = col;
        vals[pos] = R.bal[NC].d[0];
        const int nl = nloc(lab, NC);
#pragma unroll
        for (int i = 0; i < NC + 3; i++) {
            if (i >= nl) break;
            const int64_t p2 = rowptr[rb + loc_row(lab, NC, i)] + j;
            cols[p2] = col;
            vals[p2] = R.loc[i].d[0];
        }
    } else {
        const int dk = d / n, dc = d % n;
        const int pcol = (A.lab[dk][dc] == 0) ? 2 * NC : NC;
        if (j == pcol || j == pcol + 1) {
            const int64_t pos = rowptr[rb + r0 + NC] + s_water + (j - pcol);
            cols[pos] = col;
            vals[pos] = R.bal[NC].d[0];
        }
    }
}

// own block (slot 0): accumulation, fugacity and constraint rows, the six
// faces and the transfer, all with the block's own unknown j seeded
template <int NC>
__global__ __launch_bounds__(256) void k_jac_own(AsmArgs A, const double* dp, const int64_t* rowptr,
                                                 int32_t* cols, double* vals) {
    using D1 = DualV<1>;
    constexpr int MAXC = JacShape<NC>::MAXC;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (int64_t)2 * A.n * MAXC) return;
    const int j = (int)(item % MAXC);
    const int blk = (int)(item / MAXC);
    const int n = A.n, k = blk / n, c = blk % n;
    const int lab = A.lab[k][c];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    if (j >= bs) return;
    if (A.phi[k * n + c] < 1e-10) {
        // zero-porosity block: identity rows (Petsc_Functions.cpp:42051-42056)
        const int64_t row = A.boff[blk] + j;
        cols[rowptr[row]] = (int32_t)row;
        vals[rowptr[row]] = 1.0;
        return;
    }
    const Geo G = geom(A, c);
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    VarsT<NC, D1> sv;
    seed_vars<NC, 1>(v, j, A.pcoeff, sv);
    PropsT<NC, D1> sp;
    load_props_dual<NC>(A, dp, blk, j, sp);
    RowAcc<NC, D1> R;
    own_piece<NC, D1>(A, k, c, G, sv, sp, R);
    for (int q = 0; q < 6; q++) {
        const FaceGeo FG = face_geo(A, G, q);
        const int o = FG.o;
        if (o < 0) continue;
        VarsT<NC, double> vo;
        PropsT<NC, double> po;
        load_vars<NC>(A, k, o, vo);
        load_props<NC>(A, k, o, po);
        face_piece<NC, D1, double>(A, k, G, FG, q, c, o, sv, sp, vo, po, R);
    }
    {
        VarsT<NC, double> vx;
        PropsT<NC, double> px;
        load_vars<NC>(A, 1 - k, c, vx);
        load_props<NC>(A, 1 - k, c, px);
        if (k == 0) transfer_piece<NC, D1, double>(A, 0, G, c, sv, sp, vx, px, R);
        else transfer_piece<NC, double, D1>(A, 1, G, c, vx, px, sv, sp, R);
    }
    scatter_col<NC>(A, G, k, c, blk, lab, 0, blk, j, R, rowptr, cols, vals);
}

// dependency slots 1..7: one face (neighbour seeded) or the transfer (the
// other continuum seeded); only the balance and water rows depend on them
template <int NC>
__global__ __launch_bounds__(256) void k_jac_nb(AsmArgs A, const double* dp, const int64_t* rowptr,
                                                int32_t* cols, double* vals) {
    using D1 = DualV<1>;
    constexpr int MAXC = JacShape<NC>::MAXC;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (int64_t)2 * A.n * 7 * MAXC) return;
    // block-major: a block's seven slots write neighbouring CSR positions
    const int j = (int)(item % MAXC);
    const int slot = 1 + (int)((item / MAXC) % 7);
    const int blk = (int)(item / (7 * MAXC));
    const int n = A.n, k = blk / n, c = blk % n;
    if (A.phi[k * n + c] < 1e-10) return;
    const Geo G = geom(A, c);
    const int d = dep_block(A, k, G, c, slot);
    if (d < 0) return;
    const int dk = d / n, dc = d % n;
    if (j >= bsize_of(A, d, NC)) return;
    const int lab = A.lab[k][c];
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    PropsT<NC, double> p;
    load_props<NC>(A, k, c, p);
    VarsT<NC, double> vd;
    load_vars<NC>(A, dk, dc, vd);
    VarsT<NC, D1> sv;
    seed_vars<NC, 1>(vd, j, A.pcoeff, sv);
    PropsT<NC, D1> sp;
    load_props_dual<NC>(A, dp, d, j, sp);
    RowAcc<NC, D1> R;
    if (slot == 7) {
        if (k == 0) transfer_piece<NC, double, D1>(A, 0, G, c, v, p, sv, sp, R);
        else transfer_piece<NC, D1, double>(A, 1, G, c, sv, sp, v, p, R);
    } else {
        const FaceGeo FG = face_geo(A, G, slot - 1);
        face_piece<NC, double, D1>(A, k, G, FG, slot - 1, c, dc, v, p, sv, sp, R);
    }
    scatter_col<NC>(A, G, k, c, blk, lab, slot, d, j, R, rowptr, cols, vals);
}

// ------------------------------------------------ permeability adjoint
// d(Fbar . F)/d ln k.  Permeability enters F only through the face
// transmissibilities Kf = SI_C 2 A / (d_a/k_a + d_b/k_b) (IFK_Calc,
// Init_Calcs.cpp:21-42), which scale the convective face fluxes, and through
// the transfer shape factor sf2 = 4 sigma sum_d k_m,d / ds_d^2, which scales
// the convective transfer terms; both enter linearly, so d F / d Kf is the
// convective piece divided by Kf (same for sf2).
//   k_perm_face: one lane per (continuum, cell, direction) owning the face
//     to the + neighbour: s = sum over both blocks' balance rows of
//     Fbar . (convective piece)/Kf; writes s dKf/dln k of the lower cell and
//     of the upper cell to two face slots;
//   k_perm_gather: per cell and direction, the lower-face slot of the cell
//     plus the upper-face slot of its - neighbour (no atomics);
//   k_perm_transfer: per cell, sum over both continua's balance rows of
//     Fbar . (convective transfer)/sf2 times d sf2 / d ln k_m,d.
template <int NC>
__device__ __forceinline__ double fbar_dot(const AsmArgs& A, const double* Fbar, int blk, int lab,
                                           const RowAcc<NC, double>& R) {
    const int r0 = lab == 0 ? NC : 0;
    const double* fb = Fbar + A.boff[blk] + r0;
    double s = 0;
#pragma unroll
    for (int i = 0; i <= NC; i++) s += fb[i] * R.bal[i];
    return s;
}

template <int NC>
__global__ __launch_bounds__(256) void k_perm_face(AsmArgs A, const double* Fbar, const double* perm_m,
                                                   const double* perm_f, double* slot_lo, double* slot_hi) {
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (item >= (int64_t)2 * 3 * n) return;
    const int c = (int)(item % n);
    const int d = (int)((item / n) % 3);
    const int k = (int)(item / (3 * (int64_t)n));
    const int64_t out = ((int64_t)k * 3 + d) * n + c;
    slot_lo[out] = 0;
    slot_hi[out] = 0;
    const Geo G = geom(A, c);
    const FaceGeo FG = face_geo(A, G, 2 * d);
    const int o = FG.o;
    if (o < 0) return;
    const double Kf = A.iftran[((int64_t)k * 3 + d) * n + c];
    if (!(Kf > 0)) return;   // no flow through the face (zero porosity or zero area)
    VarsT<NC, double> vc, vo;
    PropsT<NC, double> pc, po;
    load_vars<NC>(A, k, c, vc);
    load_vars<NC>(A, k, o, vo);
    load_props<NC>(A, k, c, pc);
    load_props<NC>(A, k, o, po);
    double s = 0;
    if (A.phi[k * n + c] >= 1e-10) {
        RowAcc<NC, double> R;
        face_piece<NC, double, double>(A, k, G, FG, 2 * d, c, o, vc, pc, vo, po, R, false);
        s += fbar_dot<NC>(A, Fbar, k * n + c, vc.lab, R);
    }
    if (A.phi[k * n + o] >= 1e-10) {
        const Geo Go = geom(A, o);
        const FaceGeo FGo = face_geo(A, Go, 2 * d + 1);
        RowAcc<NC, double> R;
        face_piece<NC, double, double>(A, k, Go, FGo, 2 * d + 1, o, c, vo, po, vc, pc, R, false);
        s += fbar_dot<NC>(A, Fbar, k * n + o, vo.lab, R);
    }
    s /= Kf;
    // dKf/dln k_c = Kf (d_c/k_c) / (d_c/k_c + d_o/k_o), likewise for o
    const double* pk = (k == 0 ? perm_m : perm_f) + (int64_t)d * n;
    const double rc = FG.dl / pk[c], ro = FG.dn / pk[o];
    slot_lo[out] = s * Kf * rc / (rc + ro);
    slot_hi[out] = s * Kf * ro / (rc + ro);
}

__global__ void k_perm_gather(const AsmArgs A, const double* slot_lo, const double* slot_hi, double* g_m,
                              double* g_f) {
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (item >= (int64_t)2 * 3 * n) return;
    const int c = (int)(item % n);
    const int d = (int)((item / n) % 3);
    const int k = (int)(item / (3 * (int64_t)n));
    const Geo G = geom(A, c);
    const int lower = nb_of(A, G, 2 * d + 1);   // the - neighbour owns the face below
    double v = slot_lo[((int64_t)k * 3 + d) * n + c];
    if (lower >= 0) v += slot_hi[((int64_t)k * 3 + d) * n + lower];
    double* g = k == 0 ? g_m : g_f;
    if (g) g[(int64_t)d * n + c] = v;
}

template <int NC>
__global__ __launch_bounds__(256) void k_perm_transfer(AsmArgs A, const double* Fbar, double* g_m) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (c >= n || !g_m) return;
    const AsmConsts& K = *A.K;
    const Geo G = geom(A, c);
    const double ds[3] = {G.ds0, G.ds1, G.ds2};
    const double sf2 = 4 * K.sigma * (A.perm_m[c] / (ds[0] * ds[0]) + A.perm_m[n + c] / (ds[1] * ds[1]) +
                                      A.perm_m[2 * n + c] / (ds[2] * ds[2]));
    VarsT<NC, double> vm, vf;
    PropsT<NC, double> pm, pf;
    load_vars<NC>(A, 0, c, vm);
    load_vars<NC>(A, 1, c, vf);
    load_props<NC>(A, 0, c, pm);
    load_props<NC>(A, 1, c, pf);
    double s = 0;
    if (A.phi[c] >= 1e-10) {
        RowAcc<NC, double> R;
        transfer_piece<NC, double, double>(A, 0, G, c, vm, pm, vf, pf, R, false);
        s += fbar_dot<NC>(A, Fbar, c, vm.lab, R);
    }
    if (A.phi[n + c] >= 1e-10) {
        RowAcc<NC, double> R;
        transfer_piece<NC, double, double>(A, 1, G, c, vm, pm, vf, pf, R, false);
        s += fbar_dot<NC>(A, Fbar, n + c, vf.lab, R);
    }
    if (!(sf2 > 0)) return;
    s /= sf2;
#pragma unroll
    for (int d = 0; d < 3; d++)
        g_m[(int64_t)d * n + c] += s * 4 * K.sigma * A.perm_m[(int64_t)d * n + c] / (ds[d] * ds[d]);
}

// ------------------------------------------- reduced (nc+1)^2 block CSR
// The local rows of a block (fugacity and constraint rows; they involve the
// block's own unknowns only) are eliminated: with primary unknowns p (nc+1
// per block) and secondary unknowns s,
//   J_ls ds + J_lp dp = -F_l  =>  ds = -S dp - t,  S = J_ls^-1 J_lp, t = J_ls^-1 F_l,
// and the balance rows become
//   sum_d (J_mp,ad - J_ms,ad S_d) dp_d = -F_m,a + sum_d J_ms,ad t_d.
// Primary unknowns (in this order; the rest of the block is secondary):
//   two-phase    x_0..x_{nc-3}, P, Sw, So   (secondary x_{nc-2}, x_{nc-1}, y, Sg)
//   liquid       x_0..x_{nc-2}, P, Sw       (secondary x_{nc-1}, So)
//   vapour       y_0..y_{nc-2}, P, Sw       (secondary y_{nc-1}, Sg)
// sum S = 1 is the only local row with saturations, so exactly one
// saturation can be secondary; the compositions carry the rest.  Zero-
// porosity blocks (identity rows, F = 0) reduce to an identity block.
__device__ __forceinline__ int red_prim_col(int lab, int nc, int k) {
    if (lab == 0) return k < nc - 2 ? k : 2 * nc + (k - (nc - 2));   // x.., P, Sw, So
    return k < nc - 1 ? k : nc + (k - (nc - 1));                     // comp.., P, Sw
}
__device__ __forceinline__ int red_sec_col(int lab, int nc, int i) {
    if (lab == 0) return i < 2 + nc ? nc - 2 + i : 2 * nc + 3;       // x_{nc-2}, x_{nc-1}, y.., Sg
    return i == 0 ? nc - 1 : nc + 2;                                 // comp_{nc-1}, So|Sg
}

// role of unknown column j of a block: primary index (>= 0), or -1 - the
// secondary index (the inverse of red_prim_col / red_sec_col)
__device__ __forceinline__ int red_role(int lab, int nc, int j) {
    if (lab == 0) {
        if (j < nc - 2) return j;
        if (j < 2 * nc) return -1 - (j - (nc - 2));
        if (j <= 2 * nc + 2) return nc - 2 + (j - 2 * nc);
        return -1 - (nc + 2);
    }
    if (j < nc - 1) return j;
    if (j == nc - 1) return -1;
    if (j <= nc + 1) return nc - 1 + (j - nc);
    return -2;
}

// entry (row r of block blk, column col of dependency block d) of the CSR;
// sm, sw: the dependency's segment start in the balance rows and the water
// row (seg_start, kinds 0 and 1)
template <int NC>
__device__ __forceinline__ double csr_at(const AsmArgs& A, int blk, int lab, int d, int64_t sm, int64_t sw, int r,
                                         int col, const int64_t* rowptr, const double* vals) {
    const int r0 = lab == 0 ? NC : 0;
    const int64_t rb = A.boff[blk];
    if (r >= r0 && r < r0 + NC) return vals[rowptr[rb + r] + sm + col];
    if (r == r0 + NC) {
        if (d == blk) return vals[rowptr[rb + r] + sw + col];
        const int n = A.n;
        const int pcol = (A.lab[d / n][d % n] == 0) ? 2 * NC : NC;
        if (col != pcol && col != pcol + 1) return 0.0;
        return vals[rowptr[rb + r] + sw + (col - pcol)];
    }
    return d == blk ? vals[rowptr[rb + r] + col] : 0.0;   // local rows: own block only
}

// Gauss-Jordan elimination of one block's local system by a group of
// RED_GROUP = 32 lanes of a wave: lane j < block size holds column j of the
// local rows (the block's unknown j), lane 31 the residual column.  The
// pivots run over the secondary columns in order; each is chosen by partial
// pivoting (largest magnitude at or below the diagonal, the first on ties)
// by the lane holding that column and broadcast with a cross-lane read, so
// every lane updates its own column in registers.  Afterwards the primary
// lanes hold S = J_ls^-1 J_lp column by column and lane 31 holds
// t = J_ls^-1 F_l.  nl and lab are uniform over the group.
constexpr int RED_GROUP = 32;

// Calls f(std::integral_constant<int, B>) ... f(<E - 1>): straight-line code
// with every array index a constant from the start, so per-lane arrays stay
// in registers (a loop whose body selects an element by a run-time index,
// such as a pivot-row swap, is otherwise turned into an indexed access of a
// stack array).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// NL = the block's number of local rows, a compile-time shape: nc + 3 for a
// two-phase block (lab = 0), 2 for a single-phase one.
template <int NC, int NL>
__device__ __forceinline__ void group_eliminate(double (&col)[NC + 3], int lab, int gbase) {
    if (NL == NC + 3) lab = 0;
    static_for<0, NL>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int P = gbase + red_sec_col(lab, NC, q);
        int piv = q;
        double best = fabs(col[q]);
        static_for<q + 1, NL>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            const double v = fabs(col[r]);
            if (v > best) {
                best = v;
                piv = r;
            }
        });
        piv = __shfl(piv, P, 64);
        const double cq = col[q];
        double cp = cq;
        static_for<q + 1, NL>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            const bool m = r == piv;
            cp = m ? col[r] : cp;
            col[r] = m ? cq : col[r];
        });
        col[q] = cp;
        double f[NL];
        static_for<0, NL>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            f[r] = __shfl(col[r], P, 64);
        });
        const double cn = col[q] * (1.0 / f[q]);
        static_for<0, NL>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if (r != q) col[r] -= f[r] * cn;
        });
        col[q] = cn;
    });
}

// S, t of one block from the group's columns (zero beyond the block's nl rows)
template <int NC>
__device__ __forceinline__ void group_write_st(const double (&col)[NC + 3], int lab, int nl, int j, int bs,
                                               double* Sb, double* tb) {
    constexpr int NLM = NC + 3, NP = NC + 1;
    if (j == RED_GROUP - 1) {
#pragma unroll
        for (int r = 0; r < NLM; r++) tb[r] = r < nl ? col[r] : 0.0;
    } else if (j < bs) {
        const int role = red_role(lab, NC, j);
        if (role >= 0)
#pragma unroll
            for (int r = 0; r < NLM; r++) Sb[r * NP + role] = r < nl ? col[r] : 0.0;
    }
}

// block counts of the reduced rows (zero-porosity rows keep only their own block)
__global__ void k_reduce_bcount(AsmArgs A, int64_t* cnt) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (blk >= 2 * n) return;
    const int k = blk / n, c = blk % n;
    if (A.phi[k * n + c] < 1e-10) {
        cnt[blk] = 1;
        return;
    }
    const Geo G = geom(A, c);
    int m = 0;
    for (int sl = 0; sl < 8; sl++)
        if (dep_block(A, k, G, c, sl) >= 0) m++;
    cnt[blk] = m;
}

// one lane per (row block, dependency slot, reduced column q): column q of
// the reduced block, and (slot 0, q 0) also the reduced right-hand side
template <int NC>
__global__ __launch_bounds__(256) void k_reduce_blocks(AsmArgs A, const double* F, const int64_t* rowptr,
                                                       const double* vals, const double* S, const double* t,
                                                       const int64_t* browptr, int32_t* bcols, double* bvals,
                                                       double* brhs) {
    constexpr int NLM = NC + 3, NP = NC + 1;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (item >= (int64_t)2 * n * 8 * NP) return;
    const int q = (int)(item % NP);
    const int slot = (int)((item / NP) % 8);
    const int blk = (int)(item / (8 * NP));
    const int k = blk / n, c = blk % n;
    const Geo G = geom(A, c);
    const int lab = A.lab[k][c];
    const int64_t b0 = browptr[blk];
    if (A.phi[k * n + c] < 1e-10) {
        if (slot != 0) return;
        bcols[b0] = blk;
        for (int i = 0; i < NP; i++) bvals[b0 * NP * NP + i * NP + q] = (i == q) ? 1.0 : 0.0;
        if (q == 0)
            for (int i = 0; i < NP; i++) brhs[(int64_t)blk * NP + i] = 0.0;
        return;
    }
    const int d = dep_block(A, k, G, c, slot);
    if (d < 0) return;
    // rank of d among the existing dependencies in block order
    int rank = 0;
    for (int sl = 0; sl < 8; sl++) {
        const int e = dep_block(A, k, G, c, sl);
        if (e >= 0 && e < d) rank++;
    }
    const int64_t bz = b0 + rank;
    const int dk = d / n, dc = d % n;
    const int dlab = A.lab[dk][dc];
    const int dnl = A.phi[dk * n + dc] < 1e-10 ? 0 : nloc(dlab, NC);
    const double* Sd = S + (int64_t)d * NLM * NP;
    const int r0 = lab == 0 ? NC : 0;
    if (q == 0) bcols[bz] = d;
    const int64_t sm = seg_start(A, k, G, c, 0, d, NC), sw = seg_start(A, k, G, c, 1, d, NC);
    for (int i = 0; i < NP; i++) {
        const int r = r0 + i;
        double v = csr_at<NC>(A, blk, lab, d, sm, sw, r, red_prim_col(dlab, NC, q), rowptr, vals);
        for (int sidx = 0; sidx < dnl; sidx++)
            v -= csr_at<NC>(A, blk, lab, d, sm, sw, r, red_sec_col(dlab, NC, sidx), rowptr, vals) * Sd[sidx * NP + q];
        bvals[bz * NP * NP + i * NP + q] = v;
    }
    if (slot == 0 && q == 0) {
        const int64_t rb = A.boff[blk];
        for (int i = 0; i < NP; i++) {
            const int r = r0 + i;
            double v = -F[rb + r];
            for (int sl = 0; sl < 8; sl++) {
                const int e = dep_block(A, k, G, c, sl);
                if (e < 0) continue;
                const int ek = e / n, ec = e % n;
                if (A.phi[ek * n + ec] < 1e-10) continue;
                const int elab = A.lab[ek][ec];
                const int enl = nloc(elab, NC);
                const double* te = t + (int64_t)e * NLM;
                const int64_t em = seg_start(A, k, G, c, 0, e, NC), ew = seg_start(A, k, G, c, 1, e, NC);
                for (int sidx = 0; sidx < enl; sidx++)
                    v += csr_at<NC>(A, blk, lab, e, em, ew, r, red_sec_col(elab, NC, sidx), rowptr, vals) * te[sidx];
            }
            brhs[(int64_t)blk * NP + i] = v;
        }
    }
}

// full Newton update from the reduced solution: du_p = dp, du_s = -S dp - t
template <int NC>
__global__ void k_reduce_expand(AsmArgs A, const double* S, const double* t, const double* dp, double* du) {
    constexpr int NLM = NC + 3, NP = NC + 1;
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (blk >= 2 * n) return;
    const int k = blk / n, c = blk % n;
    const int lab = A.lab[k][c];
    const int64_t rb = A.boff[blk];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    if (A.phi[k * n + c] < 1e-10) {
        for (int j = 0; j < bs; j++) du[rb + j] = 0.0;
        return;
    }
    const double* p = dp + (int64_t)blk * NP;
    for (int q = 0; q < NP; q++) du[rb + red_prim_col(lab, NC, q)] = p[q];
    const double* Sb = S + (int64_t)blk * NLM * NP;
    const double* tb = t + (int64_t)blk * NLM;
    const int nl = nloc(lab, NC);
    for (int i = 0; i < nl; i++) {
        double v = -tb[i];
        for (int q = 0; q < NP; q++) v -= Sb[i * NP + q] * p[q];
        du[rb + red_sec_col(lab, NC, i)] = v;
    }
}

// ------------------------------- reduced block CSR straight from the state
// The same reduced system without the full CSR.  The Jacobian is linear in
// the seed direction, so the reduced column q of dependency d is the
// directional derivative of the balance rows along
//   v_q = e_{primary q} - sum_s S_d[s][q] e_{secondary s}   (q < NP)
// and the right-hand-side correction sum_d J_ms,d t_d the one along
//   w   = sum_s t_d[s] e_{secondary s}                        (q = NP).
//   k_red_loc    one lane per (block, unknown j): the local rows' column j
//                (own_piece in forward mode);
//   k_red_local  one lane group per block: the columns eliminated to S, t;
//   k_red_dprops one lane per (block, direction q <= NP): the property
//                derivatives along v_q / w, combined from the per-unknown
//                ones (the property sweep is linear in the seed), and the
//                direction's seeds of the block's variables;
//   k_red_own    one lane per (block, direction q): the balance rows along
//                direction q of the own block and of the blocks that depend
//                on it (six face neighbours from the same face evaluations,
//                the other continuum's block), written to the reduced
//                blocks' column q, or (q = NP) to per-slot partials;
//   k_red_rhs    one lane per block: -F_m plus the slot partials.

// unknown column of variable m of a block in variable order (x_0..x_{nc-1},
// y_0..y_{nc-1}, P, Sw, So, Sg), or -1 where the block has no such unknown
// (InitGuess2's order, PetscAUX.cpp:892-957)
__device__ __forceinline__ int var_col(int lab, int nc, int m) {
    if (lab == 0) return m;
    if (m < nc) return lab == 1 ? m : -1;
    if (m < 2 * nc) return lab == -1 ? m - nc : -1;
    const int sidx = m - 2 * nc;   // 0 P, 1 Sw, 2 So, 3 Sg
    if (sidx <= 1) return nc + sidx;
    if (sidx == 2) return lab == 1 ? nc + 2 : -1;
    return lab == -1 ? nc + 2 : -1;
}

// the block's variables seeded along reduced direction q from the seeds
// k_red_dprops tabulated (g: the block's column q, stride ncol)
template <int NC>
__device__ __forceinline__ void seed_tab(const VarsT<NC, double>& v, const double* g, int ncol, double pcoeff,
                                         VarsT<NC, DualV<1>>& out) {
    constexpr int NF = JacShape<NC>::NF;
    out.lab = v.lab;
    auto D = [&](double val, int m, double sc) {
        DualV<1> r;
        r.v = val;
        r.d[0] = g[(NF + m) * ncol] * sc;
        return r;
    };
#pragma unroll
    for (int i = 0; i < NC; i++) {
        out.x[i] = D(v.x[i], i, 1.0);
        out.y[i] = D(v.y[i], NC + i, 1.0);
    }
    out.Po = D(v.Po, 2 * NC, pcoeff);
#pragma unroll
    for (int j = 0; j < 3; j++) out.S[j] = D(v.S[j], 2 * NC + 1 + j, 1.0);
}

// The local rows of a block (fugacity and constraint rows) as the Po, Pg and
// density part of propsT and then loc_piece compute them, phase by phase:
// each phase's mixing sums serve
// both its density (the root's derivative) and its ln phi, and only one
// phase's temporaries are live at a time.  Values of Po, Pg and the
// densities are the property table's (tab, stride s2), derivatives from the
// forward-mode sweep.
template <int NC>
__device__ __forceinline__ void red_loc_rows(const AsmArgs& A, int k, int lab, const VarsT<NC, DualV<1>>& v,
                                             double rpo, double rpg, double zl, double zg, const double* tab,
                                             int64_t s2, RowAcc<NC, DualV<1>>& R) {
    using T = DualV<1>;
    const AsmConsts& K = *A.K;
    const double rt = RGAS * K.T;
    if (lab != 0) {
        T sm = 0.0;
#pragma unroll
        for (int i = 0; i < NC; i++) sm += (lab == 1 ? v.x[i] : v.y[i]);
        R.loc[0] += sm - 1.0;
        R.loc[1] += v.S[0] + v.S[1] + v.S[2] - 1.0;
        return;
    }
    // Pg as propsT forms it (IFT-scaled gas capillary pressure in the matrix)
    T Po = v.Po, Pg;
    if (k == 0) {
        T ps = 0.0;
#pragma unroll
        for (int i = 0; i < NC; i++) ps += K.par[i] * (v.x[i] * rpo - v.y[i] * rpg);
        const T bift = fpow(ps * 1e-6, 3.88);
        T krg, krog, pcg;
        table<T>(K.sgt, K.nsgt, v.S[2], krg, krog, pcg);
        if (bift < MISCIBLEIFT) Pg = Po + pcg * bift / IFTBASECASE;
        else Pg = Po + pcg;
    } else {
        Pg = Po;
    }
    const T Pg_sweep = Pg;   // zphase's pressure: the sweep's own value
    Po.v = tab[2 * s2];
    Pg.v = tab[3 * s2];
    T fL[NC];
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
        const T* xp = ph == 0 ? v.x : v.y;
        const T& P = ph == 0 ? Po : Pg;
        T Ca, Cb, t[NC];
        mixT<NC, T>(K, xp, Ca, Cb, t);
        // density: the root's derivative from the implicit function (the
        // value from the table), as zphase
        T ro;
        {
            // zphase forms A, B from the sweep's own pressure (its value the state's)
            const T Pz = ph == 0 ? v.Po : Pg_sweep;
            const T Az = Ca * Pz / (rt * rt);
            const T Bz = Cb * Pz / rt;
            const bool srk = K.eos == OSR_EOS_SRK;
            const T U = srk ? Bz : 2.0 * Bz;
            const T W = srk ? T(0.0) : Bz;
            const T Z = zroot_k(-(1.0 + Bz - U), Az - Bz * U - U - W * W, -(Az * Bz - Bz * W * W - W * W), ph == 1,
                                ph == 0 ? zl : zg);
            ro = Pz / (Z * rt);
            ro.v = tab[(5 + ph) * s2];
        }
        const T Ap = Ca * P / (rt * rt), Bp = Cb * P / rt;
        const T Zp = P / (ro * rt);
        const T DD = Ap / (Bp * (K.d1 - K.d2));
        const T FF = flog((Zp + K.d2 * Bp) / (Zp + K.d1 * Bp));
        const T lz = flog(Zp - Bp);
        const T iCa = 1.0 / Ca, iCb = 1.0 / Cb;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const T t1 = K.b[i] * iCb;
            const T EE = 2.0 * t[i] * iCa - t1;
            const T phi = fexp(t1 * (Zp - 1.0) - lz + DD * EE * FF);
            if (ph == 0) fL[i] = xp[i] * P * phi;
            else R.loc[i] += fL[i] - xp[i] * P * phi;
        }
    }
    T sx = 0.0, sy = 0.0;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        sx += v.x[i];
        sy += v.y[i];
    }
    R.loc[NC] += sx - 1.0;
    R.loc[NC + 1] += sy - 1.0;
    R.loc[NC + 2] += v.S[0] + v.S[1] + v.S[2] - 1.0;
}

// one lane per (block, unknown j): column j of the block's local rows
// (fugacity and constraint rows) by forward mode with unknown j seeded --
// they read only Po, Pg and the phase densities, so only that part of the
// property sweep runs (red_loc_rows), with the values of the property table
// and its phase roots -- into Wl laid out [block][local row][column]
template <int NC>
__global__ __launch_bounds__(64) void k_red_loc(AsmArgs A, double* Wl) {
    using D1 = DualV<1>;
    constexpr int MAXC = JacShape<NC>::MAXC;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (int64_t)2 * A.n * MAXC) return;
    const int j = (int)(item % MAXC);
    const int blk = (int)(item / MAXC);
    const int n = A.n, k = blk / n, c = blk % n;
    const int lab = A.lab[k][c];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    if (j >= bs || A.phi[k * n + c] < 1e-10) return;
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    VarsT<NC, D1> sv;
    seed_vars<NC, 1>(v, j, A.pcoeff, sv);
    double rpo = 0, rpg = 0;
    if (k == 0 && A.rho_prev) {
        rpo = A.rho_prev[c];
        rpg = A.rho_prev[n + c];
    }
    double zl, zg;
    table_roots(A, blk, zl, zg);
    const int64_t s2 = 2 * (int64_t)n;
    RowAcc<NC, D1> R;
    red_loc_rows<NC>(A, k, lab, sv, rpo, rpg, zl, zg, A.props + blk, s2, R);
    const int nl = nloc(lab, NC);
    double* w = Wl + (int64_t)blk * MAXC * (NC + 3) + j;
#pragma unroll
    for (int r = 0; r < NC + 3; r++)
        if (r < nl) w[r * MAXC] = R.loc[r].d[0];
}

// One group of RED_GROUP lanes per block: each lane loads its column of the
// local rows -- from the full CSR (CSR = true) or from the columns k_red_loc
// wrote (Wl) -- and the group eliminates (group_eliminate) to S, t.
// Zero-porosity blocks (identity rows) get S = 0, t = 0.
template <int NC, bool CSR>
__global__ __launch_bounds__(256) void k_red_local(AsmArgs A, const double* Wl, const double* F,
                                                   const int64_t* rowptr, const double* vals, double* S,
                                                   double* t) {
    constexpr int NLM = NC + 3, NP = NC + 1;
    static_assert(2 * NC + 4 < RED_GROUP, "a block's unknowns and its residual column fit one lane group");
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int blk = (int)(item / RED_GROUP);
    const int j = (int)(item % RED_GROUP);
    const int n = A.n;
    if (blk >= 2 * n) return;                    // whole groups leave together
    const int gbase = (int)(threadIdx.x & 63) & ~(RED_GROUP - 1);
    const int k = blk / n, c = blk % n;
    const int lab = A.lab[k][c];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    double* Sb = S + (int64_t)blk * NLM * NP;
    double* tb = t + (int64_t)blk * NLM;
    double col[NLM];
#pragma unroll
    for (int r = 0; r < NLM; r++) col[r] = 0.0;
    if (A.phi[k * n + c] < 1e-10) {
        group_write_st<NC>(col, lab, 0, j, bs, Sb, tb);
        return;
    }
    const int nl = nloc(lab, NC);
    const int64_t rb = A.boff[blk];
    if (j == RED_GROUP - 1) {
#pragma unroll
        for (int r = 0; r < NLM; r++)
            if (r < nl) col[r] = F[rb + loc_row(lab, NC, r)];
    } else if (j < bs) {
        if (CSR) {
#pragma unroll
            for (int r = 0; r < NLM; r++)
                if (r < nl) col[r] = vals[rowptr[rb + loc_row(lab, NC, r)] + j];
        } else {
            const double* w = Wl + (int64_t)blk * JacShape<NC>::MAXC * NLM + j;
#pragma unroll
            for (int r = 0; r < NLM; r++)
                if (r < nl) col[r] = w[r * JacShape<NC>::MAXC];
        }
    }
    // a group is one block, so each group runs one static shape
    if (nl == NLM) group_eliminate<NC, NC + 3>(col, lab, gbase);
    else group_eliminate<NC, 2>(col, lab, gbase);
    group_write_st<NC>(col, lab, nl, j, bs, Sb, tb);
}

// The local rows of a block and their elimination in one group of RED_GROUP
// lanes, the local-row Jacobian formed from the closed-form derivatives of
// the cubic EOS instead of forward mode.  Same values as k_red_loc +
// k_red_local<NC, false> up to rounding; no per-column workspace.
//
// Fugacity row i of a two-phase block is R_i = fL_i - fG_i with
// f_i = w_i P phi_i(w, P) of the phase's composition w (x or y) and pressure
// P (Po; Pg = Po + pcg(Sg) bift(x, y) / IFTBASE in the matrix, Pg = Po in
// the fracture), ln phi_i = t1_i (Z - 1) - ln(Z - B) + DD EE_i FF with
// t1_i = b_i / Cb, EE_i = 2 t_i / Ca - t1_i, DD = A / (B (d1 - d2)),
// FF = ln((Z + d2 B) / (Z + d1 B)), t_i = sum_m a_im w_m, Ca = sum w_i t_i,
// Cb = sum b_i w_i, A = Ca P / (RT)^2, B = Cb P / RT.  Along composition m of
// the same phase (dCa = sa_m sum_j u_j (bic_mj + bic_jm), dCb = b_m,
// dt_i = a_im, dZ = Z_A dA + Z_B dB from the cubic's implicit function):
//   d ln phi_i = b_i alpha_m + EE_i beta_m + t_i gamma_m + a_im kappa - dlz_m
//   alpha_m = (DD FF - (Z - 1)) b_m / Cb^2 + dZ_m / Cb,
//   beta_m  = dDD_m FF + DD dFF_m,  gamma_m = -2 DD FF dCa_m / Ca^2,
//   kappa   = 2 DD FF / Ca,  dlz_m = (dZ_m - dB_m) / (Z - B),
// and df_i = delta_im P phi_i + f_i d ln phi_i; along P (unit dP)
// g_i = df_i/dP = w_i phi_i + f_i (t1_i dZ_P - dlz_P + DD EE_i dFF_P).  The
// matrix gas pressure adds -g_i^G dPg to every column that moves Pg
// (compositions through the parachor IFT, Sg through the Pc table).
//
// Lanes [0, 16) hold the liquid phase and [16, 32) the vapour, lane i of a
// half component i: the mixing row, the per-component quantities f_i,
// f_i EE_i, f_i t_i, g_i.  Then lane j holds column j of the local rows
// (InitGuess2's unknown order: x, y, P, Sw, So, Sg), gathering the row
// quantities by cross-lane reads, lane 31 the residual column, and the group
// eliminates as k_red_local does.
template <int NC>
__global__ __launch_bounds__(256, 3) void k_red_locel(AsmArgs A, const double* F, double* S, double* t) {
    constexpr int NLM = NC + 3, NP = NC + 1;
    static_assert(NC <= 16, "one component per lane of a 16-lane half");
    static_assert(2 * NC + 4 < RED_GROUP, "a block's unknowns and its residual column fit one lane group");
    const int64_t item = wg_index(A) * blockDim.x + threadIdx.x;
    const int blk = (int)(item / RED_GROUP);
    const int j = (int)(item % RED_GROUP);
    const int n = A.n;
    if (blk >= 2 * n) return;                    // whole groups leave together
    const int gbase = (int)(threadIdx.x & 63) & ~(RED_GROUP - 1);
    const int k = blk / n, c = blk % n;
    const int lab = A.lab[k][c];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    double* Sb = S + (int64_t)blk * NLM * NP;
    double* tb = t + (int64_t)blk * NLM;
    double col[NLM];
#pragma unroll
    for (int r = 0; r < NLM; r++) col[r] = 0.0;
    if (A.phi[k * n + c] < 1e-10) {
        group_write_st<NC>(col, lab, 0, j, bs, Sb, tb);
        return;
    }
    const int nl = nloc(lab, NC);
    const int64_t rb = A.boff[blk];
    if (j == RED_GROUP - 1) {
#pragma unroll
        for (int r = 0; r < NLM; r++)
            if (r < nl) col[r] = F[rb + loc_row(lab, NC, r)];
    }
    if (lab != 0) {
        // sum of the present phase's compositions - 1, sum S - 1 (columns:
        // compositions, P, Sw, So|Sg)
        if (j < NC) col[0] = 1.0;
        else if (j == NC + 1 || j == NC + 2) col[1] = 1.0;
        group_eliminate<NC, 2>(col, lab, gbase);
        group_write_st<NC>(col, lab, nl, j, bs, Sb, tb);
        return;
    }
    const AsmConsts& K = *A.K;
    const double rt = RGAS * K.T;
    const double dd = K.d1 - K.d2;
    const bool srk = K.eos == OSR_EOS_SRK;
    const int64_t s2 = 2 * (int64_t)n;
    const double* tab = A.props + blk;
    // ---- per half: phase h, component i
    const int h = (j >> 4) & 1;
    const int i = j & 15;
    const bool ival = i < NC;
    const int ic = ival ? i : 0;
    const double* wsrc = (h == 0 ? A.x[k] : A.y[k]) + c;
    double u[NC];
    double Cb = 0.0, spar = 0.0;
#pragma unroll
    for (int m = 0; m < NC; m++) {
        const double w = wsrc[(int64_t)m * n];
        u[m] = w * K.sa[m];
        Cb += w * K.b[m];
        spar += w * K.par[m];
    }
    double srow = 0.0, scol = 0.0;
#pragma unroll
    for (int m = 0; m < NC; m++) {
        srow += u[m] * K.bic[ic * NC + m];
        scol += u[m] * K.bic[m * NC + ic];
    }
    const double ti = srow * K.sa[ic];
    const double dCai = (srow + scol) * K.sa[ic];   // dCa / dw_i
    const double Ca = half_sum<16>(ival ? u[ic] * srow : 0.0);
    const double wi = ival ? wsrc[(int64_t)ic * n] : 0.0;
    const double P = tab[(2 + h) * s2];
    const double Z = P / (tab[(5 + h) * s2] * rt);   // the sweep's root
    const double Bb = Cb * P / rt, Aa = Ca * P / (rt * rt);
    const double a1 = srk ? -1.0 : Bb - 1.0;
    const double a2 = srk ? Aa - Bb * Bb - Bb : Aa - 3.0 * Bb * Bb - 2.0 * Bb;
    const double ifz = -1.0 / ((3.0 * Z + 2.0 * a1) * Z + a2);
    const double ZA = (Z - Bb) * ifz;
    const double ZB = (srk ? -(2.0 * Bb + 1.0) * Z - Aa : (Z - 6.0 * Bb - 2.0) * Z - Aa + (3.0 * Bb + 2.0) * Bb) * ifz;
    const double DD = Aa / (Bb * dd);
    const double Zd2 = Z + K.d2 * Bb, Zd1 = Z + K.d1 * Bb;
    const double FF = log(Zd2 / Zd1);
    const double ZmB = Z - Bb;
    const double lz = log(ZmB);
    const double iCa = 1.0 / Ca, iCb = 1.0 / Cb;
    const double t1 = K.b[ic] * iCb;
    const double EE = 2.0 * ti * iCa - t1;
    const double phi = exp(t1 * (Z - 1.0) - lz + DD * EE * FF);
    const double fi = wi * P * phi;
    // along P: dA = Ca / RT^2, dB = Cb / RT
    const double dAP = Ca / (rt * rt), dBP = Cb / rt;
    const double dZP = ZA * dAP + ZB * dBP;
    const double dFFP = (dZP + K.d2 * dBP) / Zd2 - (dZP + K.d1 * dBP) / Zd1;
    const double dlzP = (dZP - dBP) / ZmB;
    const double gi = wi * phi + fi * (t1 * dZP - dlzP + DD * EE * dFFP);
    // ---- per column
    const int hc = j < NC ? 0 : 1;                 // phase of a composition column
    const int m = j < NC ? j : (j < 2 * NC ? j - NC : 0);
    const int own = gbase + 16 * hc + m;           // owner lane of the column's component
    const int same = gbase + 16 * hc + i;          // this lane's place in the column's half
    const double cZ = __shfl(Z, same, 64), cB = __shfl(Bb, same, 64), cCa = __shfl(Ca, same, 64);
    const double cCb = __shfl(Cb, same, 64), cP = __shfl(P, same, 64);
    const double cZA = __shfl(ZA, same, 64), cZB = __shfl(ZB, same, 64), cDD = __shfl(DD, same, 64);
    const double cFF = __shfl(FF, same, 64);
    const double cdCa = __shfl(dCai, own, 64), cPphi = __shfl(P * phi, own, 64);
    const double sparL = __shfl(spar, gbase, 64), sparG = __shfl(spar, gbase + 16, 64);
    // composition column m of phase hc
    const double bm = K.b[m];
    const double dA = cdCa * cP / (rt * rt), dB = bm * cP / rt;
    const double dZ = cZA * dA + cZB * dB;
    const double cZd2 = cZ + K.d2 * cB, cZd1 = cZ + K.d1 * cB;
    const double dFF = (dZ + K.d2 * dB) / cZd2 - (dZ + K.d1 * dB) / cZd1;
    const double dlz = (dZ - dB) / (cZ - cB);
    const double icCa = 1.0 / cCa, icCb = 1.0 / cCb;
    const double dDD = (cdCa - cCa * bm * icCb) * icCb / (rt * dd);
    const double alpha = (cDD * cFF - (cZ - 1.0)) * bm * icCb * icCb + dZ * icCb;
    const double beta = dDD * cFF + cDD * dFF;
    const double gamma = -2.0 * cDD * cFF * cdCa * icCa * icCa;
    const double kappa = 2.0 * cDD * cFF * icCa;
    // the matrix gas pressure's parachor-IFT and Pc-table dependence
    double dPgx = 0.0, dPgy = 0.0, dPgS = 0.0;
    if (k == 0) {
        const double bift = tab[4 * s2];
        int r = 0;
        const double sg = A.S[0][2 * (int64_t)n + c];
        while (r < K.nsgt && !(sg < K.sgt[4 * r])) r++;
        const double slope = (r == 0 || r == K.nsgt)
                                 ? 0.0
                                 : (K.sgt[4 * r + 3] - K.sgt[4 * (r - 1) + 3]) / (K.sgt[4 * r] - K.sgt[4 * (r - 1)]);
        if (bift < MISCIBLEIFT) {
            const double rpo = A.rho_prev ? A.rho_prev[c] : 0.0, rpg = A.rho_prev ? A.rho_prev[n + c] : 0.0;
            const double ps = sparL * rpo - sparG * rpg;
            const double cpl = tab[15 * s2] / IFTBASECASE * 3.88 * bift / ps;
            dPgx = cpl * K.par[m] * rpo;
            dPgy = -cpl * K.par[m] * rpg;
            dPgS = slope * bift / IFTBASECASE;
        } else {
            dPgS = slope;
        }
    }
    const bool xcol = j < NC, ycol = j >= NC && j < 2 * NC;
    const double dPgc = xcol ? dPgx : dPgy;
#pragma unroll
    for (int r = 0; r < NC; r++) {
        const int src = gbase + 16 * hc + r;
        const double f_r = __shfl(fi, src, 64);
        const double fe_r = __shfl(fi * EE, src, 64);
        const double ft_r = __shfl(fi * ti, src, 64);
        const double gL = __shfl(gi, gbase + r, 64);
        const double gG = __shfl(gi, gbase + 16 + r, 64);
        if (xcol || ycol) {
            const double a_rm = K.sa[r] * K.sa[m] * K.bic[r * NC + m];
            double dfr = f_r * (K.b[r] * alpha - dlz + kappa * a_rm) + fe_r * beta + ft_r * gamma;
            if (r == m) dfr += cPphi;
            col[r] = (xcol ? dfr : -dfr) - gG * dPgc;
        } else if (j == 2 * NC) {
            col[r] = A.pcoeff * (gL - gG);
        } else if (j == 2 * NC + 3) {
            col[r] = -gG * dPgS;
        }
    }
    if (xcol) col[NC] = 1.0;
    if (ycol) col[NC + 1] = 1.0;
    if (j >= 2 * NC + 1 && j <= 2 * NC + 3) col[NC + 2] = 1.0;
    group_eliminate<NC, NC + 3>(col, lab, gbase);
    group_write_st<NC>(col, lab, nl, j, bs, Sb, tb);
}

// one lane per (block, direction q <= NP): the property sweep in forward
// mode along the reduced direction q of the block (v_q, or w for q = NP;
// the phase roots from the property table), written with the direction's
// seeds of the block's variables.  Built for two waves per SIMD (registers
// capped at 256 per lane, the rest spilled): the latency of the sweep's
// chains hides better than at one wave with no spill.
template <int NC>
__global__ __launch_bounds__(256, 2) void k_red_dprops(AsmArgs A, const double* S, const double* t, double* dred) {
    using D1 = DualV<1>;
    constexpr int NF = JacShape<NC>::NF, NLM = NC + 3, NP = NC + 1;
    const int64_t item = wg_index(A) * blockDim.x + threadIdx.x;
    if (item >= (int64_t)2 * A.n * (NP + 1)) return;
    const int q = (int)(item % (NP + 1));
    const int blk = (int)(item / (NP + 1));
    const int n = A.n, k = blk / n, c = blk % n;
    const int lab = A.lab[k][c];
    const int nl = A.phi[k * n + c] < 1e-10 ? 0 : nloc(lab, NC);
    const double* Sb = S + (int64_t)blk * NLM * NP;
    const double* tb = t + (int64_t)blk * NLM;
    double* o = dred + (int64_t)blk * JacShape<NC>::RF * (NP + 1) + q;
    // the variables' seeds: 1 on primary q, -S[s][q] (or t[s] for q = NP) on
    // secondary s, 0 elsewhere and for variables the block does not have
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    VarsT<NC, D1> sv;
    sv.lab = v.lab;
    auto seed = [&](double val, int m, double sc, D1& out) -> double {
        const int jc = var_col(lab, NC, m);
        double dv = 0.0;
        if (jc >= 0) {
            const int role = red_role(lab, NC, jc);
            if (role >= 0) {
                dv = role == q ? 1.0 : 0.0;
            } else {
                const int si = -1 - role;
                if (si < nl) dv = q < NP ? -Sb[si * NP + q] : tb[si];
            }
        }
        out.v = val;
        out.d[0] = dv * sc;
        return dv;
    };
#pragma unroll
    for (int i = 0; i < NC; i++) {
        seed(v.x[i], i, 1.0, sv.x[i]);
        seed(v.y[i], NC + i, 1.0, sv.y[i]);
    }
    const double dvP = seed(v.Po, 2 * NC, A.pcoeff, sv.Po);
#pragma unroll
    for (int j = 0; j < 3; j++) seed(v.S[j], 2 * NC + 1 + j, 1.0, sv.S[j]);
    double rpo = 0, rpg = 0;
    if (k == 0 && A.rho_prev) {
        rpo = A.rho_prev[c];
        rpg = A.rho_prev[n + c];
    }
    double zl, zg;
    table_roots(A, blk, zl, zg);
    PropsT<NC, D1> sp;
    propsT<NC, D1>(A, k, c, sv.lab, sv.Po, sv.x, sv.y, sv.S, rpo, rpg, sp, zl, zg);
#pragma unroll
    for (int f = 0; f < NF; f++) o[f * (NP + 1)] = prop_field<NC, D1>(sp, f).d[0];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        o[(NF + i) * (NP + 1)] = sv.x[i].d[0];
        o[(NF + NC + i) * (NP + 1)] = sv.y[i].d[0];
    }
    o[(NF + 2 * NC) * (NP + 1)] = dvP;
#pragma unroll
    for (int j = 0; j < 3; j++) o[(NF + 2 * NC + 1 + j) * (NP + 1)] = sv.S[j].d[0];
}

// Position of a dependency in the reduced block row of row block b
// (continuum kb, cell coordinates ix, iy, iz): the row's existing
// dependencies in block order are the other continuum's block first when
// kb = 1, then b's own continuum by cell offset -nx*ny, -nx, -1, 0, +1, +nx,
// +nx*ny (faces 5, 3, 1, b itself, 0, 2, 4; a face exists inside the grid),
// then the other continuum's block when kb = 0.  slot: 0 b itself, 1 + g
// the neighbour across b's face g, 7 the other continuum.
__device__ __forceinline__ int red_rank(const AsmArgs& A, int kb, int ix, int iy, int iz, int slot) {
    const int e5 = iz > 0, e3 = iy > 0, e1 = ix > 0;
    const int e0 = ix < A.nx - 1, e2 = iy < A.ny - 1, e4 = iz < A.nz - 1;
    if (slot == 7) return kb == 1 ? 0 : 1 + e0 + e1 + e2 + e3 + e4 + e5;
    switch (slot) {
        case 6: return kb;                                     // face 5
        case 4: return kb + e5;                                // face 3
        case 2: return kb + e5 + e3;                           // face 1
        case 0: return kb + e5 + e3 + e1;                      // b itself
        case 1: return kb + e5 + e3 + e1 + 1;                  // face 0
        case 3: return kb + e5 + e3 + e1 + 1 + e0;             // face 2
        default: return kb + e5 + e3 + e1 + 1 + e0 + e2;       // face 4
    }
}

// One lane per (block a, direction q): every reduced column that a's
// direction q feeds.  The lane seeds a along q once (a's variables and
// property derivatives in registers) and forms
//   * a's own block: accumulation, its six faces and its transfer term
//     (its local rows are k_red_loc's: forming them here too would be dead
//     work that the compiler does not remove -- 2.0 k instructions and most
//     of the register spill at nc = 10);
//   * the blocks of its face neighbours b, (b, a): each face's terms of b's
//     balance rows, taken from the same face evaluation (face_piece MIRROR);
//   * the other continuum's block of the same cell: its transfer term.
// Each is written to column q of the reduced block, or (q = NP) to the
// per-slot partial of the row block's right-hand side (slot 0 own, 1 + f
// the row block's face f, 7 the transfer).  A zero-porosity block gets the
// identity (Petsc_Functions.cpp:42051-42056) and still feeds its
// neighbours; no row of a zero-porosity block is written here otherwise.
template <int NC>
__global__ __launch_bounds__(256) void k_red_own(AsmArgs A, const double* dred, const int64_t* browptr,
                                                 int32_t* bcols, double* bvals, double* rpart) {
    using D1 = DualV<1>;
    constexpr int NP = NC + 1;
    const int64_t item = wg_index(A) * blockDim.x + threadIdx.x;
    if (item >= (int64_t)2 * A.n * (NP + 1)) return;
    const int q = (int)(item % (NP + 1));
    const int blk = (int)(item / (NP + 1));
    const int n = A.n, k = blk / n, c = blk % n;
    const int64_t nb = 2 * (int64_t)n;
    const int64_t b0 = browptr[blk];
    const bool live = A.phi[k * n + c] >= 1e-10;
    if (!live && q < NP) {
        if (q == 0) bcols[b0] = blk;
        for (int i = 0; i < NP; i++) bvals[b0 * NP * NP + i * NP + q] = (i == q) ? 1.0 : 0.0;
    }
    // column q of reduced block (row block b, column block blk), or the
    // slot partial of b's right-hand side
    // (slot: b's dependency slot of a, bslot: the same for b's right-hand side;
    // ix, iy, iz: b's cell)
    auto emit = [&](int b, int slot, int kb, int ix, int iy, int iz, const double (&col)[NP]) {
        if (q == NP) {
#pragma unroll
            for (int i = 0; i < NP; i++) rpart[((int64_t)slot * NP + i) * nb + b] = col[i];
            return;
        }
        const int64_t bz = browptr[b] + red_rank(A, kb, ix, iy, iz, slot);
        if (q == 0) bcols[bz] = blk;
#pragma unroll
        for (int i = 0; i < NP; i++) bvals[bz * NP * NP + i * NP + q] = col[i];
    };
    const Geo G = geom(A, c);
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    VarsT<NC, D1> sv;
    seed_tab<NC>(v, dred + (int64_t)blk * JacShape<NC>::RF * (NP + 1) + q, NP + 1, A.pcoeff, sv);
    PropsT<NC, D1> sp;
    load_props_dual<NC>(A, dred, blk, q, sp, NP + 1, JacShape<NC>::RF);
    RowAcc<NC, D1> R;
    if (live) acc_piece<NC, D1>(A, k, c, G, sv, sp, R);
    for (int f = 0; f < 6; f++) {
        const FaceGeo FG = face_geo(A, G, f);
        const int o = FG.o;
        if (o < 0) continue;
        VarsT<NC, double> vo;
        PropsT<NC, double> po;
        load_vars<NC>(A, k, o, vo);
        {
            // the fields of the convective and gravity terms only
            const int64_t s2 = 2 * (int64_t)n;
            const double* qn = A.props + (int64_t)k * n + o;
            po.Pw = qn[1 * s2]; po.Po = qn[2 * s2]; po.Pg = qn[3 * s2];
            po.ro[0] = qn[5 * s2]; po.ro[1] = qn[6 * s2]; po.mu[0] = qn[7 * s2]; po.mu[1] = qn[8 * s2];
            po.gam[0] = qn[9 * s2]; po.gam[1] = qn[10 * s2];
            po.kr[0] = qn[11 * s2]; po.kr[1] = qn[12 * s2]; po.kr[2] = qn[13 * s2];
        }
        D1 mb[NP];
#pragma unroll
        for (int i = 0; i < NP; i++) mb[i] = D1(0.0);
        face_piece<NC, D1, double, true>(A, k, G, FG, f, c, o, sv, sp, vo, po, R, false, mb, true);
        // the Fickian terms' fields are loaded only now, so they are not
        // live across the convective part
        __builtin_amdgcn_sched_barrier(0);
        {
            const int64_t s2 = 2 * (int64_t)n;
            const double* qn = A.props + (int64_t)k * n + o;
            po.porpor = qn[0];
#pragma unroll
            for (int i = 0; i < NC; i++) {
                po.D[0][i] = qn[(16 + i) * s2];
                po.D[1][i] = qn[(16 + NC + i) * s2];
            }
        }
        face_piece<NC, D1, double, true>(A, k, G, FG, f, c, o, sv, sp, vo, po, R, true, mb, false);
        if (A.phi[k * n + o] >= 1e-10) {
            double col[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) col[i] = mb[i].d[0];
            // a sits across b's face f ^ 1
            const int sx = f == 0 ? 1 : (f == 1 ? -1 : 0), sy = f == 2 ? 1 : (f == 3 ? -1 : 0);
            const int sz = f == 4 ? 1 : (f == 5 ? -1 : 0);
            emit(k * n + o, (f ^ 1) + 1, k, G.ix + sx, G.iy + sy, G.iz + sz, col);
        }
    }
    {
        const int kx = 1 - k;
        VarsT<NC, double> vx;
        PropsT<NC, double> px;
        load_vars<NC>(A, kx, c, vx);
        load_props<NC>(A, kx, c, px);
        const bool xlive = A.phi[kx * n + c] >= 1e-10;
        // With equal labels in the two continua the transfer term of the
        // other continuum's rows is the negated term of this one's, operation
        // for operation (the convective part changes only its sign, the
        // diffusive cases (1,1), (-1,-1), (0,0) add and subtract the same
        // products), so it is formed once; the mixed-state cases differ.
        const bool same = vx.lab == sv.lab;
        RowAcc<NC, D1> Rt;
        if (live || (xlive && same)) {
            if (k == 0) transfer_piece<NC, D1, double>(A, 0, G, c, sv, sp, vx, px, Rt);
            else transfer_piece<NC, double, D1>(A, 1, G, c, vx, px, sv, sp, Rt);
        }
        if (live) {
#pragma unroll
            for (int i = 0; i < NP; i++) R.bal[i] += Rt.bal[i];
        }
        if (xlive) {
            double col[NP];
            if (same) {
#pragma unroll
                for (int i = 0; i < NP; i++) col[i] = -Rt.bal[i].d[0];
            } else {
                RowAcc<NC, D1> Rx;
                if (k == 0) transfer_piece<NC, D1, double>(A, 1, G, c, sv, sp, vx, px, Rx);
                else transfer_piece<NC, double, D1>(A, 0, G, c, vx, px, sv, sp, Rx);
#pragma unroll
                for (int i = 0; i < NP; i++) col[i] = Rx.bal[i].d[0];
            }
            emit(kx * n + c, 7, kx, G.ix, G.iy, G.iz, col);
        }
    }
    if (!live) return;
    double col[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) col[i] = R.bal[i].d[0];
    emit(blk, 0, k, G.ix, G.iy, G.iz, col);
}

template <int NC>
__global__ void k_red_rhs(AsmArgs A, const double* F, const double* rpart, double* brhs) {
    constexpr int NP = NC + 1;
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (blk >= 2 * n) return;
    const int k = blk / n, c = blk % n;
    double* o = brhs + (int64_t)blk * NP;
    if (A.phi[k * n + c] < 1e-10) {
        for (int i = 0; i < NP; i++) o[i] = 0.0;
        return;
    }
    const Geo G = geom(A, c);
    const int r0 = A.lab[k][c] == 0 ? NC : 0;
    const int64_t rb = A.boff[blk];
    const int64_t nb = 2 * (int64_t)n;
    bool has[8];
    for (int sl = 0; sl < 8; sl++) has[sl] = dep_block(A, k, G, c, sl) >= 0;
    for (int i = 0; i < NP; i++) {
        double v = -F[rb + r0 + i];
        for (int sl = 0; sl < 8; sl++)
            if (has[sl]) v += rpart[((int64_t)sl * NP + i) * nb + blk];
        o[i] = v;
    }
}

// block sizes of the layout
__global__ void k_block_sizes(const int8_t* lm, const int8_t* lf, int n, int nc, int64_t* bs) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * n) return;
    const int8_t L = blk < n ? lm[blk] : lf[blk - n];
    bs[blk] = L == 0 ? 2 * nc + 4 : nc + 3;
}

// UpdateAll(X, true)'s last two stages in one pass per block (PetscAUX.cpp
// :176-507): the PhaseCtrl / fPhaseCtrl rule of the block's cell (its
// saturation pressure and incipient composition already evaluated; Pg0 the
// gas pressure of the previous sweep, which the reference's PhaseCtrl reads
// before the sweep refreshes it), then UpdateDependencies' property sweep
// of the new state with the previous sweep's densities (k_props), the
// block size of the new label (PJac) and, for a matrix block, the densities
// the next sweep's CalcIFT reads.  The state arrays of A alias the writable
// ones passed here; each lane reads back only what it wrote.
template <int NC>
__global__ __launch_bounds__(256) void k_apply_sweep(AsmArgs A, int8_t* lab_m, int8_t* lab_f, double* x_m,
                                                     double* y_m, double* x_f, double* y_f, double* S_m,
                                                     double* S_f, const double* psat, const double* xm,
                                                     int64_t* bs, double* rho_next) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * A.n) return;
    const int n = A.n, k = blk / n, c = blk % n;
    const int64_t s = 2 * (int64_t)n;
    double* q = A.props + blk;
    phase_apply_cell<NC>(n, c, k == 0 ? lab_m : lab_f, A.Po[k][c], q[3 * s], k == 0 ? x_m : x_f,
                         k == 0 ? y_m : y_f, k == 0 ? S_m : S_f, psat[blk], xm + (int64_t)k * NC * n);
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    double ro = 0, rg = 0;
    rho_prev_of<NC>(A, c, ro, rg);
    PropsT<NC, double> p;
    propsT<NC, double>(A, k, c, v.lab, v.Po, v.x, v.y, v.S, ro, rg, p);
    q[0 * s] = p.porpor; q[1 * s] = p.Pw; q[2 * s] = p.Po; q[3 * s] = p.Pg; q[4 * s] = p.bift;
    q[5 * s] = p.ro[0]; q[6 * s] = p.ro[1]; q[7 * s] = p.mu[0]; q[8 * s] = p.mu[1];
    q[9 * s] = p.gam[0]; q[10 * s] = p.gam[1];
    q[11 * s] = p.kr[0]; q[12 * s] = p.kr[1]; q[13 * s] = p.kr[2]; q[14 * s] = p.pcw; q[15 * s] = p.pcg;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        q[(16 + i) * s] = p.D[0][i];
        q[(16 + NC + i) * s] = p.D[1][i];
    }
    bs[blk] = v.lab == 0 ? 2 * NC + 4 : NC + 3;
    if (k == 0 && rho_next) {
        rho_next[c] = p.ro[0];
        rho_next[n + c] = p.ro[1];
    }
}

__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }

}  // namespace osr

#include "wells_dev.hpp"
#include "refjac_dev.hpp"

// ============================================================== host side
struct osr_asm {
    osr_ctx* ctx = nullptr;
    int nx = 0, ny = 0, nz = 0, n = 0, nc = 0;
    osr::AsmConsts hK{};
    osr::AsmConsts* dK = nullptr;
    double *dx = nullptr, *dy = nullptr, *dz = nullptr, *h = nullptr, *phi = nullptr, *perm_m = nullptr,
           *iftran = nullptr, *tor = nullptr;
    double* props = nullptr;       // [NF][2n]
    double* rho_tmp = nullptr;     // [2n]
    int64_t* boff = nullptr;       // [2n+1]
    int64_t* bs = nullptr;         // [2n]
    double* dprops = nullptr;      // Jacobian workspace [2n][NF][2nc+4]
    double* perm_f = nullptr;      // [3n] fracture permeabilities (permeability adjoint)
    double* pslots = nullptr;      // permeability adjoint face slots [2][2][3][n]
    int64_t* bcnt = nullptr;       // [2n+1] block counts of the reduced rows
    double* wloc = nullptr;        // direct reduction: local-row columns [2n][2nc+4][nc+3]
    double* dred = nullptr;        // direct reduction: property derivatives along the reduced directions
    double* rpart = nullptr;       // direct reduction: per-slot right-hand-side partials [8][nc+1][2n]
    bool ws_external = false;      // dprops, wloc, dred, rpart live in a caller's workspace
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    int64_t nrows = -1;
    // wells (osr_asm_set_wells): host cell sizes for Calc_Re; device copies
    // of the well table, Calc_Re and FlashInjection per well, the completed
    // blocks, [simTime, Re_f]; kept while counts fit, so a captured graph
    // reads the values of the last call
    std::vector<double> hdx, hdy;
    int nw = 0, nwb = 0, cap_w = 0, cap_b = 0;
    int *d_wi = nullptr, *d_wblk = nullptr;
    double *d_wf = nullptr, *d_wre = nullptr, *d_winj = nullptr, *d_wsc = nullptr;
    int8_t* d_wcond = nullptr;
};

namespace {

template <class Tv>
hipError_t upload(Tv** dst, const Tv* src, size_t count) {
    hipError_t e = hipMalloc((void**)dst, count * sizeof(Tv));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src, count * sizeof(Tv), hipMemcpyHostToDevice);
}

int exclusive_scan(osr_asm* a, const int64_t* in, int64_t* out, int64_t count, hipStream_t s) {
    size_t need = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)count, s) != hipSuccess) return OSR_E_HIP;
    if (need > a->scan_bytes) {
        if (a->scan_tmp) hipFree(a->scan_tmp);
        a->scan_tmp = nullptr;
        OSR_HIP(hipMalloc(&a->scan_tmp, need));
        a->scan_bytes = need;
    }
    OSR_HIP(hipcub::DeviceScan::ExclusiveSum(a->scan_tmp, need, in, out, (int)count, s));
    return OSR_OK;
}

osr::AsmArgs make_args(osr_asm* a, const osr_state_dev* s) {
    osr::AsmArgs A{};
    A.nx = a->nx; A.ny = a->ny; A.nz = a->nz; A.n = a->n;
    A.K = a->dK;
    A.dx = a->dx; A.dy = a->dy; A.dz = a->dz; A.h = a->h;
    A.phi = a->phi; A.perm_m = a->perm_m; A.iftran = a->iftran; A.tor = a->tor;
    A.lab[0] = s->label_m; A.lab[1] = s->label_f;
    A.Po[0] = s->Po_m; A.Po[1] = s->Po_f;
    A.x[0] = s->x_m; A.x[1] = s->x_f;
    A.y[0] = s->y_m; A.y[1] = s->y_f;
    A.S[0] = s->S_m; A.S[1] = s->S_f;
    A.rho_prev = s->rho_prev_m;
    A.pre[0] = s->pre_m; A.pre[1] = s->pre_f;
    A.pcoeff = s->pcoeff;
    A.props = a->props;
    A.boff = a->boff;
    A.xcd = a->ctx->xcd_remap;
    A.perm_f = a->perm_f;
    A.nw = a->nw;
    A.wsc = a->d_wsc;
    A.wi = a->d_wi;
    A.wf = a->d_wf;
    A.wcond = a->d_wcond;
    A.wre = a->d_wre;
    A.winj = a->d_winj;
    return A;
}

// the well kernels after the kernels that write the completed blocks' rows
template <int NC>
int wells_res(osr_asm* a, const osr::AsmArgs& A, double* F, hipStream_t st) {
    if (a->nw == 0 || a->nwb == 0) return OSR_OK;
    hipLaunchKernelGGL(osr::k_well_res<NC>, dim3((a->nwb + 63) / 64), dim3(64), 0, st, A, (const int*)a->d_wblk,
                       a->nwb, F);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

template <int NC>
int run_props(osr_asm* a, osr::AsmArgs& A, hipStream_t st) {
    const int nb = 2 * a->n;
    const int grid = (nb + 255) / 256;
    if (!A.rho_prev) {
        // first sweep with zero previous densities, its matrix densities feed CalcIFT
        hipLaunchKernelGGL(osr::k_props<NC>, dim3(grid), dim3(256), 0, st, A);
        hipLaunchKernelGGL(osr::k_take_rho, dim3((a->n + 255) / 256), dim3(256), 0, st, a->props, a->n, a->rho_tmp);
        A.rho_prev = a->rho_tmp;
    }
    hipLaunchKernelGGL(osr::k_props<NC>, dim3(grid), dim3(256), 0, st, A);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

int layout(osr_asm* a, const osr_state_dev* s, hipStream_t st) {
    const int nb = 2 * a->n;
    hipLaunchKernelGGL(osr::k_block_sizes, dim3((nb + 255) / 256), dim3(256), 0, st, s->label_m, s->label_f, a->n,
                       a->nc, a->bs);
    // inclusive end: scan nb+1 entries with a trailing zero
    hipLaunchKernelGGL(osr::k_set_i64, dim3(1), dim3(1), 0, st, a->bs + nb, (int64_t)0);
    return exclusive_scan(a, a->bs, a->boff, nb + 1, st);
}

template <int NC>
int residual_nc(osr_asm* a, const osr_state_dev* s, double* F, hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    int rc = run_props<NC>(a, A, st);
    if (rc) return rc;
    const int nb = 2 * a->n;
    hipLaunchKernelGGL(osr::k_residual<NC>, dim3((nb + 255) / 256), dim3(256), 0, st, A, F);
    if (hipGetLastError() != hipSuccess) return OSR_E_HIP;
    return wells_res<NC>(a, A, F, st);
}

template <int NC>
int rowptr_nc(osr_asm* a, const osr_state_dev* s, int64_t* rowptr, int64_t nrows, hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    const int nb = 2 * a->n;
    // per-row nnz into rowptr[0..nrows), trailing 0, exclusive scan in place
    hipLaunchKernelGGL(osr::k_jac_rownnz<NC>, dim3((nb + 255) / 256), dim3(256), 0, st, A, rowptr);
    hipLaunchKernelGGL(osr::k_set_i64, dim3(1), dim3(1), 0, st, rowptr + nrows, (int64_t)0);
    return exclusive_scan(a, rowptr, rowptr, nrows + 1, st);
}

template <int NC>
int jacobian_nc(osr_asm* a, const osr_state_dev* s, const int64_t* rowptr, int32_t* cols, double* vals,
                hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    int rc = run_props<NC>(a, A, st);
    if (rc) return rc;
    constexpr int NF = osr::JacShape<NC>::NF, MAXC = osr::JacShape<NC>::MAXC;
    const int64_t nb = 2 * (int64_t)a->n;
    if (!a->dprops) OSR_HIP(hipMalloc((void**)&a->dprops, (size_t)nb * NF * MAXC * sizeof(double)));
    const int64_t ga = (nb * MAXC + 255) / 256;
    hipLaunchKernelGGL(osr::k_dprops<NC>, dim3((unsigned)ga), dim3(256), 0, st, A, a->dprops);
    const int64_t go = (nb * MAXC + 255) / 256;
    hipLaunchKernelGGL(osr::k_jac_own<NC>, dim3((unsigned)go), dim3(256), 0, st, A, (const double*)a->dprops,
                       rowptr, cols, vals);
    const int64_t gn = (nb * 7 * MAXC + 255) / 256;
    hipLaunchKernelGGL(osr::k_jac_nb<NC>, dim3((unsigned)gn), dim3(256), 0, st, A, (const double*)a->dprops,
                       rowptr, cols, vals);
    if (a->nw > 0 && a->nwb > 0)
        hipLaunchKernelGGL(osr::k_well_jac<NC>, dim3((unsigned)(((int64_t)a->nwb * MAXC + 63) / 64)), dim3(64), 0, st,
                           A, (const int*)a->d_wblk, a->nwb, (const double*)a->dprops, rowptr, vals);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

// The reference's hand-written Jacobian (IteratorJacobian): its pattern
// depends on the labels only; its values on the property sweep, the
// reference-rule derivative table (in the Jacobian workspace) and the faces.
template <int NC>
int rj_rowptr_nc(osr_asm* a, const osr_state_dev* s, int64_t* rowptr, int64_t nrows, hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    const int nb = 2 * a->n;
    hipLaunchKernelGGL(osr::k_rj_rownnz<NC>, dim3((nb + 255) / 256), dim3(256), 0, st, A, rowptr);
    hipLaunchKernelGGL(osr::k_set_i64, dim3(1), dim3(1), 0, st, rowptr + nrows, (int64_t)0);
    return exclusive_scan(a, rowptr, rowptr, nrows + 1, st);
}

template <int NC>
int rj_jacobian_nc(osr_asm* a, const osr_state_dev* s, const int64_t* rowptr, int32_t* cols, double* vals,
                   hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    int rc = run_props<NC>(a, A, st);
    if (rc) return rc;
    constexpr int NF = osr::JacShape<NC>::NF, MAXC = osr::JacShape<NC>::MAXC;
    static_assert(osr::RJ<NC>::NF <= NF * MAXC, "reference-rule table exceeds the Jacobian workspace");
    const int64_t nb = 2 * (int64_t)a->n;
    if (!a->dprops) OSR_HIP(hipMalloc((void**)&a->dprops, (size_t)nb * NF * MAXC * sizeof(double)));
    hipLaunchKernelGGL(osr::k_rj_cells<NC>, dim3((unsigned)((nb * (NC + 1) + 255) / 256)), dim3(256), 0, st, A,
                       a->dprops);
    hipLaunchKernelGGL(osr::k_rj_rows<NC>, dim3((unsigned)((nb * MAXC + 127) / 128)), dim3(128), 0, st, A,
                       (const double*)a->dprops, rowptr, cols, vals);
    if (a->nw > 0 && a->nwb > 0)
        hipLaunchKernelGGL(osr::k_well_refjac<NC>, dim3((a->nwb + 63) / 64), dim3(64), 0, st, A,
                           (const int*)a->d_wblk, a->nwb, rowptr, (const int32_t*)cols, vals);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

template <int NC>
int perm_vjp_nc(osr_asm* a, const osr_state_dev* s, const double* Fbar, double* g_m, double* g_f, hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    int rc = run_props<NC>(a, A, st);
    if (rc) return rc;
    const int64_t n = a->n;
    if (!a->pslots) OSR_HIP(hipMalloc((void**)&a->pslots, (size_t)2 * 2 * 3 * n * sizeof(double)));
    double* lo = a->pslots;
    double* hi = a->pslots + 2 * 3 * n;
    const int64_t g6 = (6 * n + 255) / 256;
    hipLaunchKernelGGL(osr::k_perm_face<NC>, dim3((unsigned)g6), dim3(256), 0, st, A, Fbar, (const double*)a->perm_m,
                       (const double*)a->perm_f, lo, hi);
    hipLaunchKernelGGL(osr::k_perm_gather, dim3((unsigned)g6), dim3(256), 0, st, A, (const double*)lo,
                       (const double*)hi, g_m, g_f);
    hipLaunchKernelGGL(osr::k_perm_transfer<NC>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A, Fbar, g_m);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

template <int NC>
int reduced_nc(osr_asm* a, const osr_state_dev* s, const double* F, const int64_t* rowptr, const double* vals,
               const int64_t* browptr, int32_t* bcols, double* bvals, double* brhs, double* S, double* t,
               hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    constexpr int NP = NC + 1;
    const int64_t nb = 2 * (int64_t)a->n;
    hipLaunchKernelGGL((osr::k_red_local<NC, true>), dim3((unsigned)((nb * osr::RED_GROUP + 255) / 256)), dim3(256),
                       0, st, A, (const double*)nullptr, F, rowptr, vals, S, t);
    const int64_t g = (nb * 8 * NP + 255) / 256;
    hipLaunchKernelGGL(osr::k_reduce_blocks<NC>, dim3((unsigned)g), dim3(256), 0, st, A, F, rowptr, vals,
                       (const double*)S, (const double*)t, browptr, bcols, bvals, brhs);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

template <int NC>
int reduced_direct_nc(osr_asm* a, const osr_state_dev* s, const double* F, const int64_t* browptr, int32_t* bcols,
                      double* bvals, double* brhs, double* S, double* t, hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    int rc = run_props<NC>(a, A, st);
    if (rc) return rc;
    constexpr int NF = osr::JacShape<NC>::NF, MAXC = osr::JacShape<NC>::MAXC;
    constexpr int NP = NC + 1;
    const int64_t nb = 2 * (int64_t)a->n;
    const bool fwd_local = a->ctx->red_local_form == 0;
    if (fwd_local && !a->wloc) OSR_HIP(hipMalloc((void**)&a->wloc, (size_t)nb * MAXC * (NC + 3) * sizeof(double)));
    if (!a->dred)
        OSR_HIP(hipMalloc((void**)&a->dred, (size_t)nb * osr::JacShape<NC>::RF * (NP + 1) * sizeof(double)));
    if (!a->rpart) OSR_HIP(hipMalloc((void**)&a->rpart, (size_t)nb * 8 * NP * sizeof(double)));
    const unsigned g1 = (unsigned)((nb + 255) / 256);
    // workgroup sizes of the direction sweep and the balance-row kernel
    // (launch knobs; the per-lane arithmetic does not depend on them)
    const int wq = a->ctx->red_dprops_wg, wo = a->ctx->red_own_wg;
    const unsigned gq = (unsigned)((nb * (NP + 1) + wq - 1) / wq);
    // the one-wave-per-SIMD kernels run in one-wave workgroups, so each SIMD
    // takes its next wave as soon as its own finishes
    if (fwd_local) {
        // the local rows in forward mode, one lane per (block, unknown), then
        // their elimination
        hipLaunchKernelGGL(osr::k_red_loc<NC>, dim3((unsigned)((nb * MAXC + 63) / 64)), dim3(64), 0, st, A, a->wloc);
        hipLaunchKernelGGL((osr::k_red_local<NC, false>), dim3((unsigned)((nb * osr::RED_GROUP + 255) / 256)),
                           dim3(256), 0, st, A, (const double*)a->wloc, F, (const int64_t*)nullptr,
                           (const double*)nullptr, S, t);
    } else {
        hipLaunchKernelGGL(osr::k_red_locel<NC>, dim3((unsigned)((nb * osr::RED_GROUP + 255) / 256)), dim3(256), 0,
                           st, A, F, S, t);
    }
    hipLaunchKernelGGL(osr::k_red_dprops<NC>, dim3(gq), dim3(wq), 0, st, A, (const double*)S, (const double*)t,
                       a->dred);
    hipLaunchKernelGGL(osr::k_red_own<NC>, dim3((unsigned)((nb * (NP + 1) + wo - 1) / wo)), dim3(wo), 0, st, A,
                       (const double*)a->dred, browptr, bcols,
                       bvals, a->rpart);
    if (a->nw > 0 && a->nwb > 0)
        hipLaunchKernelGGL(osr::k_well_red<NC>, dim3((unsigned)(((int64_t)a->nwb * (NP + 1) + 63) / 64)), dim3(64), 0,
                           st, A, (const int*)a->d_wblk, a->nwb, (const double*)a->dred, browptr, bvals, a->rpart);
    hipLaunchKernelGGL(osr::k_red_rhs<NC>, dim3(g1), dim3(256), 0, st, A, F, (const double*)a->rpart, brhs);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

template <int NC>
int expand_nc(osr_asm* a, const osr_state_dev* s, const double* S, const double* t, const double* dp, double* du,
              hipStream_t st) {
    osr::AsmArgs A = make_args(a, s);
    const int64_t nb = 2 * (int64_t)a->n;
    hipLaunchKernelGGL(osr::k_reduce_expand<NC>, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, A, S, t, dp,
                       du);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

// Calc_Re (Minor_NR.cpp:2700-3249) of a well at the centre of its block
// (X_posi = Y_posi = 0, where the reference places it): with face factors
// T_j = dx/dy, dy/dx, dx/dy, dy/dx and distances r_j = dy, dx, dy, dx to the
// face neighbours j = 1..4 (+y, +x, -y, -x) and images r'_aj of the well in
// the reservoir boundaries, at distances a,
//     Re = (e^{-2 pi} prod_j r_j^T_j prod_a (r'_aj / a)^T_j)^(1 / sum T),
// 0.2079 d on an interior square block (Peaceman's e^{-pi/2} d).  The
// diagonal neighbours carry T = 0 at the centre and drop out.  Each
// boundary case is one row of the table below, applied in the reference's
// order on shared factors: on a one-block-wide grid several rows apply in
// turn, and the later ones see what the earlier ones left, as there.
struct ReImage {
    int8_t a, j, mx, my;   // image a of face j at (mx dx, my dy)
};
struct ReRow {
    int8_t kind;           // 0 interior, 1 side (faces in set cleared), 2 corner (faces in set kept)
    int8_t set[2], nset;
    int8_t aj[3][2];       // distance to image a = 2, 3, 4 as (mx, my)
    ReImage img[6];
    int8_t nimg;
};
const ReRow RE_ROWS[9] = {
    {0, {0, 0}, 0, {{0, 0}, {0, 0}, {0, 0}}, {}, 0},
    {1, {1, 0}, 1, {{0, 1}, {0, 0}, {0, 0}}, {{2, 2, 1, 1}, {2, 3, 0, 2}, {2, 4, 1, 1}}, 3},   // Iy = ny - 1
    {1, {3, 0}, 1, {{0, 1}, {0, 0}, {0, 0}}, {{2, 2, 1, 1}, {2, 1, 0, 2}, {2, 4, 1, 1}}, 3},   // Iy = 0
    {1, {2, 0}, 1, {{1, 0}, {0, 0}, {0, 0}}, {{2, 1, 1, 1}, {2, 4, 2, 0}, {2, 3, 1, 1}}, 3},   // Ix = nx - 1
    {1, {4, 0}, 1, {{1, 0}, {0, 0}, {0, 0}}, {{2, 1, 1, 1}, {2, 2, 2, 0}, {2, 3, 1, 1}}, 3},   // Ix = 0
    {2, {2, 3}, 2, {{0, 1}, {1, 1}, {1, 0}},                                                  // (0, ny - 1)
     {{2, 2, 1, 1}, {2, 3, 0, 2}, {3, 2, 2, 1}, {3, 3, 1, 2}, {4, 2, 2, 0}, {4, 3, 1, 1}}, 6},
    {2, {1, 2}, 2, {{0, 1}, {1, 1}, {1, 0}},                                                  // (0, 0)
     {{2, 2, 1, 1}, {2, 1, 0, 2}, {3, 2, 2, 1}, {3, 1, 1, 2}, {4, 2, 2, 0}, {4, 1, 1, 1}}, 6},
    {2, {3, 4}, 2, {{1, 0}, {1, 1}, {0, 1}},                                                  // (nx - 1, ny - 1)
     {{2, 4, 2, 0}, {2, 3, 1, 1}, {3, 4, 2, 1}, {3, 3, 1, 2}, {4, 3, 0, 2}, {4, 4, 1, 1}}, 6},
    {2, {1, 4}, 2, {{1, 0}, {1, 1}, {0, 1}},                                                  // (nx - 1, 0)
     {{2, 1, 1, 1}, {2, 4, 2, 0}, {3, 4, 2, 1}, {3, 1, 1, 2}, {4, 1, 0, 2}, {4, 4, 1, 1}}, 6},
};

double calc_re(int nx, int ny, double dx, double dy, int ix, int iy) {
    auto len = [&](int mx, int my) {
        const double u = mx * dx, v = my * dy;
        return std::sqrt(u * u + v * v);
    };
    double T[5] = {0, dx / dy, dy / dx, dx / dy, dy / dx};
    double r[5] = {0, dy, dx, dy, dx};
    double rij[5][5] = {}, aj[5] = {};
    double bsum = 0, iz = 1, re = 0;
    const double e2pi = std::exp(-2 * 3.1415926535897932384626433);
    const bool inx = ix != 0 && ix != nx - 1, iny = iy != 0 && iy != ny - 1;
    const bool apply[9] = {0 < ix && ix < nx - 1 && 0 < iy && iy < ny - 1,
                           iy == ny - 1 && inx, iy == 0 && inx, ix == nx - 1 && iny, ix == 0 && iny,
                           ix == 0 && iy == ny - 1, ix == 0 && iy == 0, ix == nx - 1 && iy == ny - 1,
                           ix == nx - 1 && iy == 0};
    for (int b = 0; b < 9; b++) {
        if (!apply[b]) continue;
        const ReRow& R = RE_ROWS[b];
        for (int a = 0; a < 3; a++)
            if (R.aj[a][0] || R.aj[a][1]) aj[a + 2] = len(R.aj[a][0], R.aj[a][1]);
        for (int m = 0; m < R.nimg; m++) rij[R.img[m].a][R.img[m].j] = len(R.img[m].mx, R.img[m].my);
        for (int j = 1; j <= 4; j++) {
            bool in = false;
            for (int m = 0; m < R.nset; m++) in = in || R.set[m] == j;
            if ((R.kind == 1 && in) || (R.kind == 2 && !in)) {
                T[j] = 0;
                r[j] = 1;
                rij[2][j] = 1;
                if (R.kind == 2) rij[3][j] = rij[4][j] = 1;
            }
        }
        if (R.kind == 0)
            for (int j = 1; j <= 4; j++) iz *= std::pow(r[j], T[j]);
        for (int j = 1; j <= 4; j++) bsum += T[j];
        bsum = 1 / bsum;
        if (R.kind == 1)
            for (int j = 1; j <= 4; j++) iz *= std::pow(r[j], T[j]) * std::pow(rij[2][j] / aj[2], T[j]);
        if (R.kind == 2) {
            iz = 1;
            for (int j = 1; j <= 4; j++) {
                double jz = 1;
                for (int a = 2; a <= 4; a++) jz *= std::pow(rij[a][j] / aj[a], T[j]);
                iz *= std::pow(r[j], T[j]) * jz;
            }
        }
        re = std::pow(e2pi * iz, bsum);
    }
    return re;
}

template <class Tv>
int ensure(Tv** p, int need, int cap) {
    if (*p && need <= cap) return OSR_OK;
    if (*p) hipFree(*p);
    *p = nullptr;
    OSR_HIP(hipMalloc((void**)p, (size_t)(need > 0 ? need : 1) * sizeof(Tv)));
    return OSR_OK;
}

template <int NC>
int flash_injection_nc(osr_asm* a, hipStream_t st) {
    hipLaunchKernelGGL(osr::k_flash_injection<NC>, dim3((a->nw + 63) / 64), dim3(64), 0, st,
                       (const osr::AsmConsts*)a->dK, a->nw, (const int*)a->d_wi, (const double*)a->d_wf, a->d_winj);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

// switch over the compiled component counts; the call site defines CALL(NC)
#define OSR_NC_RET_(N) case N: return CALL(N);
#define OSR_NC_RC_(N) case N: rc = CALL(N); break;
#define OSR_NC_SWITCH(nc, CALL_)                   \
    switch (nc) {                                  \
        OSR_NC_LIST(OSR_NC_RET_)                   \
        default: return OSR_E_NC;                  \
    }
#define OSR_NC_SWITCH_RC(nc)                       \
    switch (nc) {                                  \
        OSR_NC_LIST(OSR_NC_RC_)                    \
        default: return OSR_E_NC;                  \
    }

}  // namespace

extern "C" {

int osr_asm_create(osr_ctx* ctx, const osr_fluid_ext* fe, const osr_grid* g, osr_asm** out) {
    if (!ctx || !fe || !g || !out) return OSR_E_ARG;
    if (!fe->MW || !fe->Vc || !fe->Zc || !fe->parachor || !g->dx || !g->dy || !g->dz || !g->phi_m ||
        !g->phi_f || !g->perm_m || !g->perm_f || !g->swt || !g->sgt)
        return OSR_E_ARG;
    if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->nswt < 1 || g->nsgt < 1 || g->nswt > osr::MAXTAB ||
        g->nsgt > osr::MAXTAB || !(g->T > 0) || !(g->dt > 0))
        return OSR_E_ARG;
    const int64_t n64 = (int64_t)g->nx * g->ny * g->nz;
    if (n64 * 2 * (2 * ctx->nc + 4) > INT32_MAX) return OSR_E_ARG;   // column indices are int32
    osr_asm* a = new (std::nothrow) osr_asm();
    if (!a) return OSR_E_ARG;
    a->ctx = ctx;
    a->nx = g->nx; a->ny = g->ny; a->nz = g->nz; a->n = (int)n64; a->nc = ctx->nc;
    const int n = a->n, nc = a->nc;
    a->hdx.assign(g->dx, g->dx + g->nx);
    a->hdy.assign(g->dy, g->dy + g->ny);
    osr::AsmConsts& K = a->hK;
    K.nc = nc;
    K.eos = ctx->eos;
    K.T = g->T;
    const osr::FluidConsts<double>& F = ctx->host_fluid.d;
    K.d1 = F.d1;
    K.d2 = F.d2;
    for (int i = 0; i < nc; i++) {
        const double sa = F.sac[i] * (1 + F.m[i] * (1 - std::sqrt(g->T / F.Tc[i])));
        K.sa[i] = sa;
        K.a[i] = sa * sa;
        K.b[i] = F.b[i];
        K.tc[i] = F.Tc[i];
        K.pc[i] = F.Pc[i];
        K.mw[i] = fe->MW[i];
        K.sqmw[i] = std::sqrt(fe->MW[i]);
        K.vcm3[i] = fe->Vc[i] / 1000000.0;
        K.v53[i] = std::pow(K.vcm3[i], 5.0 / 3.0);
        K.v23[i] = std::pow(K.vcm3[i], 2.0 / 3.0);
        K.par[i] = fe->parachor[i];
        K.wk[i] = F.wk[i];
        const double tr = g->T / F.Tc[i];
        const double gam = std::pow(F.Tc[i], 1.0 / 6) / (std::sqrt(fe->MW[i]) * std::pow(F.Pc[i] / 101325, 2.0 / 3));
        K.mu0[i] = (4.61 * std::pow(tr, 0.618) - 2.04 * std::exp(-0.449 * tr) + 1.94 * std::exp(-4.058 * tr) + 0.1) *
                   1e-4 / gam;
        for (int j = 0; j < nc; j++) K.bic[i * nc + j] = F.bic[i * nc + j];
    }
    // binary diffusion table (Mass_Tran.cpp:41-45), T-only
    double coldia[osr::NCMAX], epdia[osr::NCMAX];
    for (int i = 0; i < nc; i++) {
        coldia[i] = 0.1866 * std::pow(fe->Vc[i], 1.0 / 3.0) * std::pow(fe->Zc[i], -1.2);
        epdia[i] = 65.3 * F.Tc[i] * std::pow(fe->Zc[i], 3.6);
    }
    for (int j = 0; j < nc; j++) {
        K.pdlast[j] = 0;
        K.ipdij[j * nc + j] = 0;
        for (int k = 0; k < nc; k++) {
            if (j == k) { K.pdij[j * nc + k] = 0; continue; }
            const double Sij = (coldia[j] + coldia[k]) / 2;
            const double Tij = g->T / std::sqrt(epdia[j] * epdia[k]);
            const double Oij = 1.06306 * std::pow(Tij, -0.1561) + 0.193 * std::exp(-0.47635 * Tij) +
                               1.03587 * std::exp(-1.52996 * Tij) + 1.76474 * std::exp(-3.89411 * Tij);
            K.pdij[j * nc + k] = 188.2922475 * std::sqrt(g->T * (1.0 / fe->MW[j] + 1.0 / fe->MW[k])) /
                                 (osr::RGAS * Oij * Sij * Sij);
            K.pdlast[j] = K.pdij[j * nc + k];
            K.ipdij[j * nc + k] = 1.0 / K.pdij[j * nc + k];
        }
    }
    for (int r = 0; r < 4 * g->nswt; r++) K.swt[r] = g->swt[r];
    for (int r = 0; r < 4 * g->nsgt; r++) K.sgt[r] = g->sgt[r];
    K.nswt = g->nswt;
    K.nsgt = g->nsgt;
    K.refP = g->refP; K.cpor = g->cpor; K.dcpor = g->dcpor; K.dt = g->dt;
    K.wat_ro = g->wat_ro; K.wat_mu = g->wat_mu; K.sigma = g->sigma; K.lf = g->lf;
    K.cdiff = g->cdiff; K.const_diff = g->const_diff;
    // host geometry: block heights (CalcBlockHeight) and face kA/h (IFK_Calc)
    std::vector<double> hh(g->nz), phi(2 * (size_t)n), ift(6 * (size_t)n, 0.0), tor(n, 1.0), permm(3 * (size_t)n);
    double acc = 0;
    for (int k = g->nz - 1; k >= 0; k--) {
        if (k == g->nz - 1) acc = 0.5 * g->dz[g->nz - 1];
        else acc += 0.5 * (g->dz[k] + g->dz[k + 1]);
        hh[k] = acc;
    }
    for (int c = 0; c < n; c++) {
        phi[c] = g->phi_m[c];
        phi[n + c] = g->phi_f[c];
        if (g->tor) tor[c] = g->tor[c];
    }
    for (int i = 0; i < 3 * n; i++) permm[i] = g->perm_m[i];
    for (int iz = 0; iz < g->nz; iz++)
        for (int iy = 0; iy < g->ny; iy++)
            for (int ix = 0; ix < g->nx; ix++) {
                const int c = iz * g->nx * g->ny + iy * g->nx + ix;
                for (int k = 0; k < 2; k++) {
                    const double* ph = k == 0 ? g->phi_m : g->phi_f;
                    const double* pm = k == 0 ? g->perm_m : g->perm_f;
                    double* o = ift.data() + (size_t)k * 3 * n;
                    if (g->nx != 1 && ix != g->nx - 1 && ph[c] > 1e-10 && ph[c + 1] > 1e-10)
                        o[c] = osr::SI_C * 2 * g->dy[iy] * g->dz[iz] / (g->dx[ix] / pm[c] + g->dx[ix + 1] / pm[c + 1]);
                    const int cy = c + g->nx;
                    if (g->ny != 1 && iy != g->ny - 1 && ph[c] > 1e-10 && ph[cy] > 1e-10)
                        o[n + c] = osr::SI_C * 2 * g->dx[ix] * g->dz[iz] /
                                   (g->dy[iy] / pm[n + c] + g->dy[iy + 1] / pm[n + cy]);
                    const int cz = c + g->nx * g->ny;
                    if (g->nz != 1 && iz != g->nz - 1 && ph[c] > 1e-10 && ph[cz] > 1e-10)
                        o[2 * n + c] = osr::SI_C * 2 * g->dx[ix] * g->dy[iy] /
                                       (g->dz[iz] / pm[2 * n + c] + g->dz[iz + 1] / pm[2 * n + cz]);
                }
            }
    const int nf = 16 + 2 * nc;
    bool ok = upload(&a->dK, &a->hK, 1) == hipSuccess && upload(&a->dx, g->dx, g->nx) == hipSuccess &&
              upload(&a->dy, g->dy, g->ny) == hipSuccess && upload(&a->dz, g->dz, g->nz) == hipSuccess &&
              upload(&a->h, hh.data(), hh.size()) == hipSuccess && upload(&a->phi, phi.data(), phi.size()) == hipSuccess &&
              upload(&a->perm_m, permm.data(), permm.size()) == hipSuccess &&
              upload(&a->perm_f, g->perm_f, 3 * (size_t)n) == hipSuccess &&
              upload(&a->iftran, ift.data(), ift.size()) == hipSuccess &&
              upload(&a->tor, tor.data(), tor.size()) == hipSuccess &&
              hipMalloc((void**)&a->props, (size_t)nf * 2 * n * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&a->rho_tmp, 2 * (size_t)n * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&a->boff, (2 * (size_t)n + 1) * sizeof(int64_t)) == hipSuccess &&
              hipMalloc((void**)&a->bs, (2 * (size_t)n + 1) * sizeof(int64_t)) == hipSuccess;
    if (!ok) {
        osr_asm_destroy(a);
        return OSR_E_HIP;
    }
    *out = a;
    return OSR_OK;
}

int osr_asm_set_wells(osr_asm* a, const osr_wells* w, void* stream) {
    if (!a || !w || w->nw < 0) return OSR_E_ARG;
    const int nw = w->nw, nc = a->nc, n = a->n;
    if (nw == 0) {
        a->nw = 0;
        a->nwb = 0;
        return OSR_OK;
    }
    if (!w->welli || !w->wellf || !(w->re_f >= 0)) return OSR_E_ARG;
    std::vector<double> re(nw);
    std::vector<int> blocks;
    std::vector<char> seen(2 * (size_t)n, 0);
    for (int i = 0; i < nw; i++) {
        const int32_t* wi = w->welli + 5 * i;
        const double* wf = w->wellf + (size_t)(nc + 6) * i;
        if (wi[0] < 0 || wi[0] > 8 || wi[1] < 0 || wi[1] >= a->nx || wi[2] < 0 || wi[2] >= a->ny || wi[3] < 0 ||
            wi[4] < wi[3] || wi[4] >= a->nz || !(wf[3] > 0))
            return OSR_E_ARG;
        re[i] = calc_re(a->nx, a->ny, a->hdx[wi[1]], a->hdy[wi[2]], wi[1], wi[2]);
        for (int k = 0; k < 2; k++)
            for (int z = wi[3]; z <= wi[4]; z++) {
                const int blk = k * n + (z * a->ny + wi[2]) * a->nx + wi[1];
                if (!seen[blk]) {
                    seen[blk] = 1;
                    blocks.push_back(blk);
                }
            }
    }
    std::sort(blocks.begin(), blocks.end());
    const int nwb = (int)blocks.size();
    std::vector<int8_t> cond(nw, 0);
    if (w->condition)
        for (int i = 0; i < nw; i++) cond[i] = w->condition[i] ? 1 : 0;
    const double sc[2] = {w->sim_time, w->re_f};
    OSR_HIP_RET(ensure(&a->d_wi, 5 * nw, 5 * a->cap_w));
    OSR_HIP_RET(ensure(&a->d_wf, (nc + 6) * nw, (nc + 6) * a->cap_w));
    OSR_HIP_RET(ensure(&a->d_wcond, nw, a->cap_w));
    OSR_HIP_RET(ensure(&a->d_wre, nw, a->cap_w));
    OSR_HIP_RET(ensure(&a->d_winj, (2 * nc + 6) * nw, (2 * nc + 6) * a->cap_w));
    a->cap_w = std::max(a->cap_w, nw);
    OSR_HIP_RET(ensure(&a->d_wblk, nwb, a->cap_b));
    a->cap_b = std::max(a->cap_b, nwb);
    OSR_HIP_RET(ensure(&a->d_wsc, 2, a->d_wsc ? 2 : 0));
    // the previous call's kernels may still read the tables: copy in stream order
    const hipStream_t st = (hipStream_t)stream;
    OSR_HIP(hipStreamSynchronize(st));
    OSR_HIP(hipMemcpy(a->d_wi, w->welli, 5 * (size_t)nw * sizeof(int), hipMemcpyHostToDevice));
    OSR_HIP(hipMemcpy(a->d_wf, w->wellf, (size_t)(nc + 6) * nw * sizeof(double), hipMemcpyHostToDevice));
    OSR_HIP(hipMemcpy(a->d_wcond, cond.data(), (size_t)nw, hipMemcpyHostToDevice));
    OSR_HIP(hipMemcpy(a->d_wre, re.data(), (size_t)nw * sizeof(double), hipMemcpyHostToDevice));
    OSR_HIP(hipMemcpy(a->d_wblk, blocks.data(), (size_t)nwb * sizeof(int), hipMemcpyHostToDevice));
    OSR_HIP(hipMemcpy(a->d_wsc, sc, sizeof(sc), hipMemcpyHostToDevice));
    OSR_HIP(hipMemset(a->d_winj, 0, (size_t)(2 * nc + 6) * nw * sizeof(double)));
    a->nw = nw;
    a->nwb = nwb;
    int rc = OSR_OK;
#define CALL(NC) flash_injection_nc<NC>(a, st)
    OSR_NC_SWITCH_RC(nc)
#undef CALL
    if (rc) return rc;
    OSR_HIP(hipStreamSynchronize(st));
    return OSR_OK;
}

int osr_asm_well_data(osr_asm* a, double* re, double* inj) {
    if (!a) return OSR_E_ARG;
    if (a->nw == 0) return OSR_OK;
    if (re) OSR_HIP(hipMemcpy(re, a->d_wre, (size_t)a->nw * sizeof(double), hipMemcpyDeviceToHost));
    if (inj)
        OSR_HIP(hipMemcpy(inj, a->d_winj, (size_t)a->nw * (2 * a->nc + 6) * sizeof(double), hipMemcpyDeviceToHost));
    return OSR_OK;
}

int osr_asm_destroy(osr_asm* a) {
    if (!a) return OSR_E_ARG;
    void* ptrs[] = {a->dK, a->dx, a->dy, a->dz, a->h, a->phi, a->perm_m, a->iftran, a->tor,
                    a->props, a->rho_tmp, a->boff, a->bs, a->perm_f, a->pslots, a->bcnt, a->scan_tmp,
                    a->d_wi, a->d_wblk, a->d_wf, a->d_wre, a->d_winj, a->d_wsc, a->d_wcond};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (!a->ws_external) {
        void* ws[] = {a->dprops, a->wloc, a->dred, a->rpart};
        for (void* p : ws)
            if (p) hipFree(p);
    }
    delete a;
    return OSR_OK;
}

// Jacobian workspace layout (256-byte aligned): the CSR Jacobian uses dprops
// only, the direct reduced system wloc | dred | rpart only, so the two
// overlay each other from offset 0 (calls on one stream run in order)
static size_t ws_layout(const osr_asm* a, size_t off[4]) {
    const size_t nb = 2 * (size_t)a->n, nc = (size_t)a->nc;
    const size_t NF = 16 + 2 * nc, MAXC = 2 * nc + 4, NP = nc + 1;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t dprops = al(nb * NF * MAXC * 8);
    const size_t wloc = al(nb * MAXC * (nc + 3) * 8), dred = al(nb * (NF + MAXC) * (NP + 1) * 8),
                 rpart = al(nb * 8 * NP * 8);
    off[0] = 0;
    off[1] = 0;
    off[2] = wloc;
    off[3] = wloc + dred;
    return std::max(dprops, wloc + dred + rpart);
}

int osr_asm_props_table(osr_asm* a, const double** props) {
    if (!a || !props) return OSR_E_ARG;
    *props = a->props;
    return OSR_OK;
}

int osr_asm_workspace_size(osr_asm* a, size_t* bytes) {
    if (!a || !bytes) return OSR_E_ARG;
    size_t off[4];
    *bytes = ws_layout(a, off);
    return OSR_OK;
}

int osr_asm_set_workspace(osr_asm* a, void* ptr, size_t bytes) {
    if (!a) return OSR_E_ARG;
    if (a->ctx && a->ctx->ws_frozen) return OSR_E_ARG;   // a captured graph may address the current one
    size_t off[4];
    const size_t need = ws_layout(a, off);
    if (ptr && (bytes < need || ((uintptr_t)ptr & 255))) return OSR_E_ARG;
    if (!a->ws_external) {
        void* ws[] = {a->dprops, a->wloc, a->dred, a->rpart};
        for (void* p : ws)
            if (p) OSR_HIP(hipFree(p));
    }
    if (ptr) {
        char* b = (char*)ptr;
        a->dprops = (double*)(b + off[0]);
        a->wloc = (double*)(b + off[1]);
        a->dred = (double*)(b + off[2]);
        a->rpart = (double*)(b + off[3]);
        a->ws_external = true;
    } else {
        a->dprops = a->wloc = a->dred = a->rpart = nullptr;
        a->ws_external = false;
    }
    return OSR_OK;
}

int osr_asm_layout(osr_asm* a, const osr_state_dev* s, int64_t* offsets, int64_t* nrows, void* stream) {
    if (!a || !s || !s->label_m || !s->label_f) return OSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    int rc = layout(a, s, st);
    if (rc) return rc;
    int64_t total = 0;
    OSR_HIP(hipMemcpyAsync(&total, a->boff + 2 * a->n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    OSR_HIP(hipStreamSynchronize(st));
    a->nrows = total;
    if (nrows) *nrows = total;
    if (offsets)
        OSR_HIP(hipMemcpyAsync(offsets, a->boff, (2 * (size_t)a->n + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    return OSR_OK;
}

static int check_state(osr_asm* a, const osr_state_dev* s) {
    if (!a || !s || !s->label_m || !s->label_f || !s->Po_m || !s->Po_f || !s->x_m || !s->y_m || !s->x_f ||
        !s->y_f || !s->S_m || !s->S_f || !(s->pcoeff > 0))
        return OSR_E_ARG;
    if (a->nrows < 0) return OSR_E_ARG;   // osr_asm_layout first
    return OSR_OK;
}

int osr_residual_device(osr_asm* a, const osr_state_dev* s, double* F, void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!F) return OSR_E_ARG;
#define CALL(NC) residual_nc<NC>(a, s, F, (hipStream_t)stream)
    OSR_NC_SWITCH(a->nc, CALL)
#undef CALL
}

int osr_jacobian_rowptr_device(osr_asm* a, const osr_state_dev* s, int64_t* rowptr, int64_t* nnz, void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!rowptr) return OSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
#define CALL(NC) rowptr_nc<NC>(a, s, rowptr, a->nrows, st)
    OSR_NC_SWITCH_RC(a->nc)
#undef CALL
    if (rc) return rc;
    if (nnz) {
        OSR_HIP(hipMemcpyAsync(nnz, rowptr + a->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        OSR_HIP(hipStreamSynchronize(st));
    }
    return OSR_OK;
}

int osr_jacobian_device(osr_asm* a, const osr_state_dev* s, const int64_t* rowptr, int32_t* cols, double* vals,
                        void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!rowptr || !cols || !vals) return OSR_E_ARG;
#define CALL(NC) jacobian_nc<NC>(a, s, rowptr, cols, vals, (hipStream_t)stream)
    OSR_NC_SWITCH(a->nc, CALL)
#undef CALL
}

int osr_jacobian_ref_rowptr_device(osr_asm* a, const osr_state_dev* s, int64_t* rowptr, int64_t* nnz,
                                   void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!rowptr) return OSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
#define CALL(NC) rj_rowptr_nc<NC>(a, s, rowptr, a->nrows, st)
    OSR_NC_SWITCH_RC(a->nc)
#undef CALL
    if (rc) return rc;
    if (nnz) {
        OSR_HIP(hipMemcpyAsync(nnz, rowptr + a->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        OSR_HIP(hipStreamSynchronize(st));
    }
    return OSR_OK;
}

int osr_jacobian_ref_device(osr_asm* a, const osr_state_dev* s, const int64_t* rowptr, int32_t* cols, double* vals,
                            void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!rowptr || !cols || !vals) return OSR_E_ARG;
#define CALL(NC) rj_jacobian_nc<NC>(a, s, rowptr, cols, vals, (hipStream_t)stream)
    OSR_NC_SWITCH(a->nc, CALL)
#undef CALL
}

int osr_residual_perm_vjp_device(osr_asm* a, const osr_state_dev* s, const double* Fbar, double* g_perm_m,
                                 double* g_perm_f, void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!Fbar || (!g_perm_m && !g_perm_f)) return OSR_E_ARG;
#define CALL(NC) perm_vjp_nc<NC>(a, s, Fbar, g_perm_m, g_perm_f, (hipStream_t)stream)
    OSR_NC_SWITCH(a->nc, CALL)
#undef CALL
}

int osr_bsr_reduced_rowptr_device(osr_asm* a, const osr_state_dev* s, int64_t* browptr, int64_t* nnzb,
                                  void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!browptr) return OSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    osr::AsmArgs A = make_args(a, s);
    const int64_t nb = 2 * (int64_t)a->n;
    hipLaunchKernelGGL(osr::k_reduce_bcount, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, A, browptr);
    hipLaunchKernelGGL(osr::k_set_i64, dim3(1), dim3(1), 0, st, browptr + nb, (int64_t)0);
    rc = exclusive_scan(a, browptr, browptr, nb + 1, st);
    if (rc) return rc;
    if (nnzb) {
        OSR_HIP(hipMemcpyAsync(nnzb, browptr + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        OSR_HIP(hipStreamSynchronize(st));
    }
    return OSR_OK;
}

int osr_jacobian_bsr_reduced_device(osr_asm* a, const osr_state_dev* s, const double* F, const int64_t* rowptr,
                                    const double* vals, const int64_t* browptr, int32_t* bcols, double* bvals,
                                    double* brhs, double* S, double* t, void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!F || !rowptr || !vals || !browptr || !bcols || !bvals || !brhs || !S || !t) return OSR_E_ARG;
#define CALL(NC) reduced_nc<NC>(a, s, F, rowptr, vals, browptr, bcols, bvals, brhs, S, t, (hipStream_t)stream)
    OSR_NC_SWITCH(a->nc, CALL)
#undef CALL
}

int osr_jacobian_bsr_reduced_direct_device(osr_asm* a, const osr_state_dev* s, const double* F,
                                           const int64_t* browptr, int32_t* bcols, double* bvals, double* brhs,
                                           double* S, double* t, void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!F || !browptr || !bcols || !bvals || !brhs || !S || !t) return OSR_E_ARG;
#define CALL(NC) reduced_direct_nc<NC>(a, s, F, browptr, bcols, bvals, brhs, S, t, (hipStream_t)stream)
    OSR_NC_SWITCH(a->nc, CALL)
#undef CALL
}

int osr_bsr_reduced_expand_device(osr_asm* a, const osr_state_dev* s, const double* S, const double* t,
                                  const double* dp, double* du, void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!S || !t || !dp || !du) return OSR_E_ARG;
#define CALL(NC) expand_nc<NC>(a, s, S, t, dp, du, (hipStream_t)stream)
    OSR_NC_SWITCH(a->nc, CALL)
#undef CALL
}

int osr_props_device(osr_asm* a, const osr_state_dev* s, double* out, void* stream) {
    int rc = check_state(a, s);
    if (rc) return rc;
    if (!out) return OSR_E_ARG;
    hipStream_t st = (hipStream_t)stream;
    osr::AsmArgs A = make_args(a, s);
#define CALL(NC) run_props<NC>(a, A, st)
    OSR_NC_SWITCH_RC(a->nc)
#undef CALL
    if (rc) return rc;
    const size_t nf = 16 + 2 * (size_t)a->nc;
    OSR_HIP(hipMemcpyAsync(out, a->props, nf * 2 * a->n * sizeof(double), hipMemcpyDeviceToDevice, st));
    return OSR_OK;
}

}  // extern "C"

// ============================================================ UpdateAll
namespace osr {

constexpr double MAX_P_COEFF = 0.01;   // Functions.h:4
constexpr double MINPPP = 5000000;     // Functions.h:259
constexpr double EPS_COMP = 0;         // Functions.h:283

// UpdateAll (PetscAUX.cpp:176-498): one lane per block reads its slice of the
// unknown vector X (reference layout) and updates the state in place:
// mole fractions clamped to [EPS_COMP, 1] and renormalised (uniform if they
// sum to 0), the pressure limiter that adds +-MAX_P_COEFF Pa when the new
// value exceeds the old by more than 1 %, the MINPPP floor, and saturations
// clamped and renormalised.  Stale x (gas blocks) / y (liquid blocks) stay.
template <int NC>
__global__ void k_unpack(int n, const int64_t* boff, const int8_t* lm, const int8_t* lf, double* Po_m,
                         double* Po_f, double* x_m, double* y_m, double* x_f, double* y_f, double* S_m,
                         double* S_f, double pcoeff, const double* X) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * n) return;
    const int k = blk / n, c = blk % n;
    const int L = (k == 0 ? lm : lf)[c];
    double* Po = k == 0 ? Po_m : Po_f;
    double* x = k == 0 ? x_m : x_f;
    double* y = k == 0 ? y_m : y_f;
    double* S = k == 0 ? S_m : S_f;
    const double* u = X + boff[blk];
    int p = 0;
    auto comp = [&](double* arr) {
        double v[NC];
        double sum = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            double t = u[p++];
            if (t < EPS_COMP) t = 0;
            else if (t > 1) t = 1;
            v[i] = t;
            sum += t;
        }
#pragma unroll
        for (int i = 0; i < NC; i++) arr[(int64_t)i * n + c] = sum != 0 ? v[i] / sum : 1.0 / NC;
    };
    if (L != -1) comp(x);
    if (L != 1) comp(y);
    const double xp = u[p++];
    double P = Po[c];
    if (fabs(pcoeff * xp) > P * (1 + MAX_P_COEFF)) {
        const double dv = xp * pcoeff - P;
        P += MAX_P_COEFF * (dv > 0 ? 1.0 : (dv < 0 ? -1.0 : 0.0));
    } else {
        P = xp * pcoeff;
    }
    if (P < MINPPP) P = MINPPP;
    Po[c] = P;
    double Sw = u[p++];
    double So = S[(int64_t)n + c], Sg = S[2 * (int64_t)n + c];
    if (L != -1) So = u[p++];
    if (L != 1) Sg = u[p++];
    Sw = Sw < 0 ? 0 : (Sw > 1 ? 1 : Sw);
    if (L != -1) So = So > 1 ? 1 : (So < 0 ? 0 : So);
    if (L != 1) Sg = Sg > 1 ? 1 : (Sg < 0 ? 0 : Sg);
    const double sum = Sw + (L != -1 ? So : 0.0) + (L != 1 ? Sg : 0.0);
    if (sum != 0) {
        Sw /= sum;
        if (L != -1) So /= sum;
        if (L != 1) Sg /= sum;
    }
    S[c] = Sw;
    if (L != -1) S[(int64_t)n + c] = So;
    if (L != 1) S[2 * (int64_t)n + c] = Sg;
}

// InitGuess2 (PetscAUX.cpp:878-964): state -> unknown vector X.
template <int NC>
__global__ void k_pack(int n, const int64_t* boff, const int8_t* lm, const int8_t* lf, const double* Po_m,
                       const double* Po_f, const double* x_m, const double* y_m, const double* x_f,
                       const double* y_f, const double* S_m, const double* S_f, double pcoeff, double* X) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * n) return;
    const int k = blk / n, c = blk % n;
    const int L = (k == 0 ? lm : lf)[c];
    const double* x = k == 0 ? x_m : x_f;
    const double* y = k == 0 ? y_m : y_f;
    const double* S = k == 0 ? S_m : S_f;
    double* u = X + boff[blk];
    int p = 0;
    if (L != -1)
#pragma unroll
        for (int i = 0; i < NC; i++) u[p++] = x[(int64_t)i * n + c];
    if (L != 1)
#pragma unroll
        for (int i = 0; i < NC; i++) u[p++] = y[(int64_t)i * n + c];
    u[p++] = (k == 0 ? Po_m : Po_f)[c] / pcoeff;
    u[p++] = S[c];
    if (L != -1) u[p++] = S[(int64_t)n + c];
    if (L != 1) u[p++] = S[2 * (int64_t)n + c];
}

}  // namespace osr

extern "C" {

int osr_unpack_device(osr_asm* a, const osr_state_mut_dev* s, const double* X, void* stream) {
    if (!a || !s || !X || !s->label_m || !s->label_f || !s->Po_m || !s->Po_f || !s->x_m || !s->y_m ||
        !s->x_f || !s->y_f || !s->S_m || !s->S_f || !(s->pcoeff > 0))
        return OSR_E_ARG;
    if (a->nrows < 0) return OSR_E_ARG;
    const int nb = 2 * a->n;
    hipStream_t st = (hipStream_t)stream;
#define UNPACK(NC) case NC: hipLaunchKernelGGL(osr::k_unpack<NC>, dim3((nb + 255) / 256), dim3(256), 0, st, a->n, a->boff, \
        s->label_m, s->label_f, s->Po_m, s->Po_f, s->x_m, s->y_m, s->x_f, s->y_f, s->S_m, s->S_f, s->pcoeff, X); break;
    switch (a->nc) {
        OSR_NC_LIST(UNPACK)
        default: return OSR_E_NC;
    }
#undef UNPACK
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

int osr_pack_device(osr_asm* a, const osr_state_dev* s, double* X, void* stream) {
    if (!a || !s || !X || !s->label_m || !s->label_f || !s->Po_m || !s->Po_f || !s->x_m || !s->y_m ||
        !s->x_f || !s->y_f || !s->S_m || !s->S_f || !(s->pcoeff > 0))
        return OSR_E_ARG;
    if (a->nrows < 0) return OSR_E_ARG;
    const int nb = 2 * a->n;
    hipStream_t st = (hipStream_t)stream;
#define PACK(NC) case NC: hipLaunchKernelGGL(osr::k_pack<NC>, dim3((nb + 255) / 256), dim3(256), 0, st, a->n, a->boff, \
        s->label_m, s->label_f, s->Po_m, s->Po_f, s->x_m, s->y_m, s->x_f, s->y_f, s->S_m, s->S_f, s->pcoeff, X); break;
    switch (a->nc) {
        OSR_NC_LIST(PACK)
        default: return OSR_E_NC;
    }
#undef PACK
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

}  // extern "C"

// ------------------------------------------------ for the Newton step (newton.hip)
namespace osr {

void asm_info(const osr_asm* a, osr_ctx** ctx, int64_t* n, int* nc, double* T) {
    *ctx = a->ctx;
    *n = a->n;
    *nc = a->nc;
    *T = a->hK.T;
}

void asm_dims(const osr_asm* a, int* nx, int* ny, int* nz) {
    *nx = a->nx;
    *ny = a->ny;
    *nz = a->nz;
}

// The property sweep of a state into the assembler's table (UpdateDependencies'
// sweep without a residual); rho_out [2n] (may be null) receives the matrix
// oil/gas densities of it, the densities the next sweep's CalcIFT reads.
int asm_sweep(osr_asm* a, const osr_state_dev* s, double* rho_out, hipStream_t st) {
    const int rc0 = check_state(a, s);
    if (rc0) return rc0;
    osr::AsmArgs A = make_args(a, s);
    int rc = OSR_E_NC;
#define CALL(NC) run_props<NC>(a, A, st)
    OSR_NC_SWITCH_RC(a->nc)
#undef CALL
    if (rc) return rc;
    return rho_out ? asm_take_rho(a, rho_out, st) : OSR_OK;
}

// PhaseCtrl's rule, the property sweep and the layout of the new state in
// one kernel plus the block-offset scan (k_apply_sweep); *nrows (host)
// receives the new MatrixSize after one synchronisation of the stream.
int asm_apply_sweep(osr_asm* a, const osr_state_mut_dev* m, const double* rho_prev, const double* psat,
                    const double* xm, double* rho_next, int64_t* nrows, hipStream_t st) {
    osr_state_dev s{};
    s.label_m = m->label_m; s.label_f = m->label_f; s.Po_m = m->Po_m; s.Po_f = m->Po_f;
    s.x_m = m->x_m; s.y_m = m->y_m; s.x_f = m->x_f; s.y_f = m->y_f; s.S_m = m->S_m; s.S_f = m->S_f;
    s.rho_prev_m = rho_prev;
    s.pcoeff = m->pcoeff;
    if (!rho_prev || !psat || !xm || !nrows) return OSR_E_ARG;
    const int rc0 = check_state(a, &s);
    if (rc0) return rc0;
    osr::AsmArgs A = make_args(a, &s);
    const int nb = 2 * a->n;
    int8_t* lm = (int8_t*)m->label_m;
    int8_t* lf = (int8_t*)m->label_f;
    switch (a->nc) {
#define APPLY_(NC)                                                                                            \
    case NC:                                                                                                  \
        hipLaunchKernelGGL(osr::k_apply_sweep<NC>, dim3((nb + 255) / 256), dim3(256), 0, st, A, lm, lf, m->x_m, \
                           m->y_m, m->x_f, m->y_f, m->S_m, m->S_f, psat, xm, a->bs, rho_next);                 \
        break;
        OSR_NC_LIST(APPLY_)
#undef APPLY_
        default: return OSR_E_NC;
    }
    OSR_HIP(hipGetLastError());
    hipLaunchKernelGGL(osr::k_set_i64, dim3(1), dim3(1), 0, st, a->bs + nb, (int64_t)0);
    OSR_HIP_RET(exclusive_scan(a, a->bs, a->boff, nb + 1, st));
    int64_t total = 0;
    OSR_HIP(hipMemcpyAsync(&total, a->boff + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    OSR_HIP(hipStreamSynchronize(st));
    a->nrows = total;
    *nrows = total;
    return OSR_OK;
}

// The matrix oil/gas densities of the table as the last sweep left it.
int asm_take_rho(osr_asm* a, double* rho_out, hipStream_t st) {
    hipLaunchKernelGGL(osr::k_take_rho, dim3((a->n + 255) / 256), dim3(256), 0, st, a->props, a->n, rho_out);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

}  // namespace osr

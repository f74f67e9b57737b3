// osr_model.hpp -- model, cell state, property sweep, face terms and residual
// rows shared by the oracle's assembly restatements (osr_assembly.cpp: the
// residual and its forward-mode Jacobian; osr_refjac.cpp: the reference's
// hand-written Jacobian).
//
// TEST INFRASTRUCTURE ONLY (see osr_oracle.h): the HIP product never links,
// loads or calls this file.  Reference citations are in osr_assembly.cpp's
// header and beside each routine.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dual.hpp"
#include "osr_oracle.h"

namespace {

constexpr double RG = 8.3144621;                   // Functions.h:216
constexpr double SI_C = 9.869232667160128e-13;     // Functions.h:207
constexpr double WAT_M_RO = 55.5084;               // Functions.h:206
constexpr double G_ACC = 9.80665;                  // Functions.h:198
constexpr double G_ACCG = 9.80665e-3;              // Functions.h:199
constexpr double IFTBASECASE = 50;                 // Functions.h:279
constexpr double MISCIBLEIFT = 1;                  // Functions.h:281
constexpr double IFTPOWER = 7;                     // Functions.h:280
constexpr double RELPERM0 = 0;                     // Functions.h:266
constexpr double MAX_P_COEFF = 0.01;               // Functions.h:4
constexpr double MINPPP = 5000000;                 // Functions.h:259
constexpr double EPS_COMP = 0;                     // Functions.h:283
constexpr int NCM = 16;
constexpr int WMAX = 2 * NCM + 4;
using D = orad::Dual<WMAX>;

using orad::val;

// ------------------------------------------------------------------ model
struct Model {
    int nx, ny, nz, n, nc, eos;
    double T, dt, refP, cpor, dcpor, sigma, lf, cdiff, wat_ro, wat_mu;
    bool const_diff;
    int nswt, nsgt;
    std::vector<double> swt, sgt;     // [row][4]: SAT KR KRO PC
    std::vector<double> dx, dy, dz, h;  // h: block height per layer
    std::vector<double> a, b, bic, mw, vcm3, coldia, epdia, mu0, par, tc, pc;
    std::vector<double> phi[2], perm[2], tor;
    std::vector<double> iftran[2][3];  // kA/h of the face to the +dir neighbour
    double d1, d2;
    std::vector<double> om;             // acentric factors (FlashInjection's Wilson start)
    // wells (deck WELLS and the run's simTime / WellCondition)
    int nw = 0;
    std::vector<int> welli;             // [nw][5] type, Ix, Iy, first and last layer
    std::vector<double> wellf;          // [nw][nc + 6]
    std::vector<int8_t> wcond;          // [nw] WellCondition
    double sim_time = 0, re_f = 0;
    std::vector<double> well_re;        // [nw] Calc_Re of the well's block (matrix completions)
    std::vector<double> inj;            // [nw][2nc + 6] FlashInjection of types 3 and 8
};

void well_setup(Model& M, const or_wells* w);

inline int cid(const Model& M, int ix, int iy, int iz) { return iz * M.nx * M.ny + iy * M.nx + ix; }

void build_model(Model& M, const or_fluid* f, const or_fluid_ext* fe, const or_grid* g) {
    M.nx = g->nx; M.ny = g->ny; M.nz = g->nz; M.n = g->nx * g->ny * g->nz;
    M.nc = f->nc; M.eos = f->eos;
    M.T = g->T; M.dt = g->dt; M.refP = g->refP; M.cpor = g->cpor; M.dcpor = g->dcpor;
    M.sigma = g->sigma; M.lf = g->lf; M.const_diff = g->const_diff != 0; M.cdiff = g->cdiff;
    M.wat_ro = g->wat_ro; M.wat_mu = g->wat_mu;
    M.nswt = g->nswt; M.nsgt = g->nsgt;
    M.swt.assign(g->swt, g->swt + 4 * g->nswt);
    M.sgt.assign(g->sgt, g->sgt + 4 * g->nsgt);
    M.dx.assign(g->dx, g->dx + M.nx);
    M.dy.assign(g->dy, g->dy + M.ny);
    M.dz.assign(g->dz, g->dz + M.nz);
    // CalcBlockHeight (Init_Calcs.cpp:993-1011): height above the bottom layer.
    M.h.assign(M.nz, 0);
    double hh = 0;
    for (int k = M.nz - 1; k >= 0; k--) {
        if (k == M.nz - 1) hh = 0.5 * M.dz[M.nz - 1];
        else hh += 0.5 * (M.dz[k] + M.dz[k + 1]);
        M.h[k] = hh;
    }
    const int nc = M.nc;
    M.a.resize(nc); M.b.resize(nc);
    or_eos_ab(f, M.T, M.a.data(), M.b.data());
    M.bic.assign(f->bic, f->bic + nc * nc);
    M.tc.assign(f->Tc, f->Tc + nc);
    M.pc.assign(f->Pc, f->Pc + nc);
    M.mw.resize(nc); M.vcm3.resize(nc); M.coldia.resize(nc); M.epdia.resize(nc);
    M.mu0.resize(nc); M.par.resize(nc);
    for (int i = 0; i < nc; i++) {
        const double Vc = fe->Vc[i], Zc = fe->Zc[i], Tc = f->Tc[i], Pc = f->Pc[i], MW = fe->MW[i];
        // EOS_Init (Init_Calcs.cpp:71-76): diameters use Vc in cm^3/mol, then Vc /= 1e6.
        M.coldia[i] = 0.1866 * std::pow(Vc, 1.0 / 3.0) * std::pow(Zc, -1.2);
        M.epdia[i] = 65.3 * Tc * std::pow(Zc, 3.6);
        M.vcm3[i] = Vc / 1000000.0;
        M.mw[i] = MW;
        M.par[i] = fe->parachor[i];
        // Tr_Mu (Init_Calcs.cpp:48-64): low-pressure viscosity.
        const double tr = M.T / Tc;
        const double gam = std::pow(Tc, 1.0 / 6) / (std::sqrt(MW) * std::pow(Pc / 101325, 2.0 / 3));
        M.mu0[i] = (4.61 * std::pow(tr, 0.618) - 2.04 * std::exp(-0.449 * tr) +
                    1.94 * std::exp(-4.058 * tr) + 0.1) * 1e-4 / gam;
    }
    if (M.eos == OR_SRK) { M.d1 = 1; M.d2 = 0; }
    else { M.d1 = 1 + OR_SQRT2; M.d2 = 1 - OR_SQRT2; }
    M.om.assign(f->w, f->w + nc);
    const int n = M.n;
    M.phi[0].assign(g->phi_m, g->phi_m + n);
    M.phi[1].assign(g->phi_f, g->phi_f + n);
    M.perm[0].assign(g->perm_m, g->perm_m + 3 * n);
    M.perm[1].assign(g->perm_f, g->perm_f + 3 * n);
    if (g->tor) M.tor.assign(g->tor, g->tor + n);
    else M.tor.assign(n, 1.0);
    // IFK_Calc (Init_Calcs.cpp:21-42).
    for (int k = 0; k < 2; k++)
        for (int d = 0; d < 3; d++) M.iftran[k][d].assign(n, 0.0);
    for (int iz = 0; iz < M.nz; iz++)
        for (int iy = 0; iy < M.ny; iy++)
            for (int ix = 0; ix < M.nx; ix++) {
                const int c = cid(M, ix, iy, iz);
                for (int k = 0; k < 2; k++) {
                    const std::vector<double>& ph = M.phi[k];
                    const std::vector<double>& pm = M.perm[k];
                    if (M.nx != 1 && ix != M.nx - 1 && ph[c] > 1e-10 && ph[c + 1] > 1e-10)
                        M.iftran[k][0][c] = SI_C * 2 * M.dy[iy] * M.dz[iz] /
                                            (M.dx[ix] / pm[c] + M.dx[ix + 1] / pm[c + 1]);
                    const int cy = c + M.nx;
                    if (M.ny != 1 && iy != M.ny - 1 && ph[c] > 1e-10 && ph[cy] > 1e-10)
                        M.iftran[k][1][c] = SI_C * 2 * M.dx[ix] * M.dz[iz] /
                                            (M.dy[iy] / pm[n + c] + M.dy[iy + 1] / pm[n + cy]);
                    const int cz = c + M.nx * M.ny;
                    if (M.nz != 1 && iz != M.nz - 1 && ph[c] > 1e-10 && ph[cz] > 1e-10)
                        M.iftran[k][2][c] = SI_C * 2 * M.dx[ix] * M.dy[iy] /
                                            (M.dz[iz] / pm[2 * n + c] + M.dz[iz + 1] / pm[2 * n + cz]);
                }
            }
    if (g->wells && g->wells->nw > 0) well_setup(M, g->wells);
}

// ------------------------------------------------------------ cell state
template <class T>
struct Vars {
    int label;
    T Po;
    T x[NCM], y[NCM];
    T S[3];
};

template <class T>
struct Props {
    T porpor, Pw, Po, Pg, bift;
    T ro[2], mu[2], gam[2];
    T kr[3], pcw, pcg;
    T D[2][NCM];
};

template <class T>
T solve_z_val(T a1, T a2, T a3, bool gas) {
    using std::acos; using std::cos; using std::pow; using std::sqrt;
    using orad::acos; using orad::cos; using orad::pow; using orad::sqrt;
    const T Q = (3 * a2 - a1 * a1) / 9;
    const T J = (9 * a1 * a2 - 27 * a3 - 2 * a1 * a1 * a1) / 54;
    const T Dd = Q * Q * Q + J * J;
    T Z;
    if (Dd < 0) {
        const T r = 2 * sqrt(-Q);
        const T s = a1 / 3;
        const T t = acos(J / sqrt(-Q * Q * Q)) / 3;
        const T x1 = r * cos(t) - s;
        const T x2 = r * cos(t + 2 * OR_PI / 3) - s;
        const T x3 = r * cos(t + 4 * OR_PI / 3) - s;
        if (gas) {
            if (x1 >= x2 && x1 >= x3) Z = x1;
            else if (x2 >= x1 && x2 >= x3) Z = x2;
            else Z = x3;
        } else {
            if (x1 <= x2 && x1 <= x3 && x1 > 0) Z = x1;
            else if (x2 <= x1 && x2 <= x3 && x2 > 0) Z = x2;
            else Z = x3;
        }
    } else if (Dd == 0) {
        const T t = J > 0 ? pow(J, 1.0 / 3) : -pow(-J, 1.0 / 3);
        const T x1 = 2 * t - a1 / 3;
        const T x2 = -t - a1 / 3;
        if (gas) Z = (x2 >= x1) ? x2 : x1;
        else Z = (x2 <= x1 && x2 > 0) ? x2 : x1;
    } else {
        const T sd = sqrt(Dd);
        T t = J + sd;
        Z = t > 0 ? pow(t, 1.0 / 3) : -pow(-t, 1.0 / 3);
        t = J - sd;
        if (t > 0) Z = Z + pow(t, 1.0 / 3);
        else Z = Z - pow(-t, 1.0 / 3);
        Z = Z - a1 / 3;
    }
    return Z;
}

template <class T>
T solve_z_t(T a1, T a2, T a3, bool gas) {
    if constexpr (!std::is_same<T, double>::value) {
        // Forward mode: the root's value from the branches below (the
        // reference's arithmetic), its derivative from the implicit function
        // theorem on f(Z) = Z^3 + a1 Z^2 + a2 Z + a3 = 0, as the reference's
        // hand Jacobian takes dZ (Do_Cycle.cpp:438-616).  Differentiating the
        // explicit formulas instead fails where they cancel: in the one-root
        // branch with Q -> 0, J - sqrt(D) rounds to exactly 0 and its cube
        // root has an infinite slope.
        const double Z = solve_z_t<double>(a1.v, a2.v, a3.v, gas);
        const double inv = -1.0 / ((3 * Z + 2 * a1.v) * Z + a2.v);
        T r = (Z * Z * a1 + Z * a2 + a3) * inv;
        r.v = Z;
        return r;
    } else {
        return solve_z_val(a1, a2, a3, gas);
    }
}

// Mixing sums with unscaled a: Ca = sum x_i x_j sqrt(a_i a_j) bic_ij, Cb = sum x b.
template <class T>
void mix_t(const Model& M, const T* x, T& Ca, T& Cb, T* t) {
    const int nc = M.nc;
    Ca = 0;
    Cb = 0;
    for (int i = 0; i < nc; i++) {
        Cb += x[i] * M.b[i];
        T ti = 0;
        for (int j = 0; j < nc; j++) {
            const double aij = std::sqrt(M.a[i] * M.a[j]) * M.bic[i * nc + j];
            Ca += x[i] * x[j] * aij;
            ti += x[j] * aij;
        }
        if (t) t[i] = ti;
    }
}

template <class T>
T phase_z_t(const Model& M, const T* x, T P, bool gas) {
    T Ca, Cb;
    mix_t(M, x, Ca, Cb, (T*)nullptr);
    const T A = Ca * P / (RG * RG * M.T * M.T);
    const T B = Cb * P / (RG * M.T);
    const bool srk = M.eos == OR_SRK;
    const T U = srk ? B : 2 * B;
    const T W = srk ? T(0) : B;
    return solve_z_t<T>(-(1 + B - U), A - B * U - U - W * W, -(A * B - B * W * W - W * W), gas);
}

// Table interpolation used by CalcSatFs (Do_Cycle.cpp:163-219): returns the
// segment index rule of the reference (first row with s < SAT).
template <class T>
void table_eval(const std::vector<double>& tb, int nr, T s, T& kr, T& kro, T& pc) {
    int r;
    for (r = 0; r < nr; r++)
        if (s < tb[4 * r + 0]) break;
    if (r == 0) {
        kr = RELPERM0;
        pc = tb[3];
        kro = tb[2];
    } else if (r == nr) {
        kr = tb[4 * (nr - 1) + 1];
        pc = tb[4 * (nr - 1) + 3];
        kro = tb[4 * (nr - 1) + 2];
    } else {
        const double s0 = tb[4 * (r - 1)], s1 = tb[4 * r];
        kr = (tb[4 * r + 1] - tb[4 * (r - 1) + 1]) / (s1 - s0) * (s - s1) + tb[4 * r + 1];
        kro = (tb[4 * r + 2] - tb[4 * (r - 1) + 2]) / (s1 - s0) * (s - s1) + tb[4 * r + 2];
        pc = (tb[4 * r + 3] - tb[4 * (r - 1) + 3]) / (s1 - s0) * (s - s1) + tb[4 * r + 3];
    }
}

// Calc_Visco density and LBC viscosity of one phase (Do_Cycle.cpp:561-612).
template <class T>
void visco_phase(const Model& M, const T* x, T P, bool gas, T& ro, T& mu) {
    using std::pow; using std::sqrt; using orad::pow; using orad::sqrt;
    const T Z = phase_z_t<T>(M, x, P, gas);
    ro = P / (Z * RG * M.T);
    T s1 = 0, s2 = 0, rv = 0, tcm = 0, pcm = 0, mwm = 0;
    for (int n = 0; n < M.nc; n++) {
        const double sq = std::sqrt(M.mw[n]);
        const T tm = x[n] * sq;
        s1 += tm * M.mu0[n];
        s2 += tm;
        rv += x[n] * M.vcm3[n];
        tcm += x[n] * M.tc[n];
        pcm += x[n] * M.pc[n] / 101325;
        mwm += x[n] * M.mw[n];
    }
    const T mu0 = s1 / s2;
    const T Er = ro * rv;
    const T etha = pow(tcm, 1.0 / 6) / (sqrt(mwm) * pow(pcm, 2.0 / 3));
    const T r1 = 0.10230 + 0.023364 * Er + 0.058533 * Er * Er - 0.040758 * Er * Er * Er +
                 0.0093324 * Er * Er * Er * Er;
    mu = mu0 + (r1 * r1 * r1 * r1 - 1e-4) / etha;
}

// CalcDiffusion (Mass_Tran.cpp:21-62) for one phase.
template <class T>
void diffusion_phase(const Model& M, const T* x, T ro, T* Dout) {
    const int nc = M.nc;
    T s1 = 0, s2 = 0;
    for (int j = 0; j < nc; j++) {
        s1 += x[j] * std::pow(M.vcm3[j], 5.0 / 3.0);
        s2 += x[j] * std::pow(M.vcm3[j], 2.0 / 3.0);
    }
    const T rr = ro * s1 / s2;
    const T roRD = (0.99589 + 0.096016 * rr - 0.22035 * rr * rr + 0.032874 * rr * rr * rr) / ro;
    for (int j = 0; j < nc; j++) {
        T s = 0;
        double pDij = 0;
        for (int k = 0; k < nc; k++) {
            if (j == k) continue;
            const double Sij = (M.coldia[j] + M.coldia[k]) / 2;
            const double Tij = M.T / std::sqrt(M.epdia[j] * M.epdia[k]);
            const double Oij = 1.06306 * std::pow(Tij, -0.1561) + 0.193 * std::exp(-0.47635 * Tij) +
                               1.03587 * std::exp(-1.52996 * Tij) + 1.76474 * std::exp(-3.89411 * Tij);
            pDij = 188.2922475 * std::sqrt(M.T * (1.0 / M.mw[j] + 1.0 / M.mw[k])) / (RG * Oij * Sij * Sij);
            s += x[k] / (pDij * roRD);
        }
        const T dt = (1 - x[j]) / (s * 10000);
        T dv;
        if (x[j] == 1) dv = roRD * pDij;
        else if (dt < 0 || s == 0) dv = T(0);
        else dv = dt;   // DIFF_INTERNAL_COEFF = 1
        if (M.const_diff) dv = T(M.cdiff);
        Dout[j] = dv;
    }
}

template <class T>
T porpor_t(const Model& M, int k, int c, T P) {
    return M.phi[k][c] * (1 + M.cpor * (P - M.refP) + M.dcpor * (P - M.refP) * (P - M.refP));
}

// Property sweep of one cell (matrix k = 0, fracture k = 1).
template <class T>
void cell_props(const Model& M, int k, int c, const Vars<T>& v, const double* rho_prev, Props<T>& p) {
    using std::pow; using orad::pow;
    const int nc = M.nc;
    p.Po = v.Po;
    p.porpor = porpor_t<T>(M, k, c, v.Po);
    if (k == 0) {
        // CalcIFT (Do_Cycle.cpp:2438-2456), densities from the previous sweep.
        if (v.label == 0) {
            T ps = 0;
            for (int i = 0; i < nc; i++) ps += M.par[i] * (v.x[i] * rho_prev[0] - v.y[i] * rho_prev[1]);
            p.bift = pow(ps * 1e-6, 3.88);
        } else {
            p.bift = T(IFTBASECASE);
        }
        // CalcSatFs (Do_Cycle.cpp:163-257): tables, Stone II, IFT scaling.
        T krw, krow, pcw, krg, krog, pcg;
        table_eval<T>(M.swt, M.nswt, v.S[0], krw, krow, pcw);
        table_eval<T>(M.sgt, M.nsgt, v.S[2], krg, krog, pcg);
        const double kro0 = M.swt[2];
        const T AA = krw + krow / kro0;
        const T BB = krg + krog / kro0;
        T kro = kro0 * (AA * BB - krw - krg);
        if (p.bift < MISCIBLEIFT) {
            const T F = pow(p.bift / IFTBASECASE, 1.0 / IFTPOWER);
            kro = F * kro + (1 - F) * v.S[1] * (1 - M.swt[0]);
            krg = F * krg + (1 - F) * v.S[2] * (1 - M.swt[0]);
        }
        if (kro < 0) kro = T(RELPERM0);
        if (kro > 1) kro = T(1);
        p.kr[0] = krw;
        p.kr[1] = kro;
        p.kr[2] = krg;
        p.pcw = pcw;
        p.pcg = pcg;
        // Calc_Phase_P (Do_Cycle.cpp:1418-1428).
        p.Pw = v.Po - pcw;
        if (p.bift < MISCIBLEIFT) p.Pg = v.Po + pcg * p.bift / IFTBASECASE;
        else p.Pg = v.Po + pcg;
    } else {
        // fCalcSatFs (Do_Cycle.cpp:426-435): kr = S, no capillary pressure;
        // fCalc_Phase_P (:1430-1446): all phase pressures equal.
        p.bift = T(IFTBASECASE);
        p.kr[0] = v.S[0];
        p.kr[1] = v.S[1];
        p.kr[2] = v.S[2];
        p.pcw = T(0);
        p.pcg = T(0);
        p.Pw = v.Po;
        p.Pg = v.Po;
    }
    // Calc_Visco: oil at Po, gas at Pg.
    visco_phase<T>(M, v.x, p.Po, false, p.ro[0], p.mu[0]);
    visco_phase<T>(M, v.y, p.Pg, true, p.ro[1], p.mu[1]);
    // CalcDiffusion, CalcPhaseGrav.
    diffusion_phase<T>(M, v.x, p.ro[0], p.D[0]);
    diffusion_phase<T>(M, v.y, p.ro[1], p.D[1]);
    T mwl = 0, mwg = 0;
    for (int i = 0; i < nc; i++) {
        mwl += v.x[i] * M.mw[i];
        mwg += v.y[i] * M.mw[i];
    }
    p.gam[0] = G_ACCG * p.ro[0] * mwl;
    p.gam[1] = G_ACCG * p.ro[1] * mwg;
}

// ----------------------------------------------------------- face terms
// Calc_Transes / fCalc_Transes for the face between a (lower index) and b
// along direction d: oil/gas/water transmissibilities, gravity mobilities
// (trans[2], trans[3]) and upstream flags (0: a upstream, 1: b upstream).
template <class T>
struct Face {
    T to, tg, tw, tgo, tgg;
    int up_o, up_g, up_w;
};

template <class T>
Face<T> face_terms(const Model& M, int k, int d, int ca, int cb, double da, double db, double ha, double hb,
                   const Vars<T>& va, const Props<T>& pa, const Vars<T>& vb, const Props<T>& pb) {
    Face<T> f;
    const double K = M.iftran[k][d][ca];
    // water
    {
        const T pot = pa.Pw - pb.Pw + G_ACC * M.wat_ro * (ha - hb);
        f.up_w = pot > 0 ? 0 : 1;
        const T kr = f.up_w == 0 ? pa.kr[0] : pb.kr[0];
        f.tw = K * (M.wat_ro * WAT_M_RO / M.wat_mu) * kr;
    }
    for (int ph = 0; ph < 2; ph++) {
        const int absent = ph == 0 ? -1 : 1;
        const double c1 = va.label == absent ? 0 : 1;
        const double c2 = vb.label == absent ? 0 : 1;
        T mg;
        if (c1 == 0 && c2 == 0) mg = T(0);
        else mg = (pa.gam[ph] * da * c1 + pb.gam[ph] * db * c2) / (c1 * da + db * c2);
        const T Pa = ph == 0 ? pa.Po : pa.Pg;
        const T Pb = ph == 0 ? pb.Po : pb.Pg;
        const T pot = Pa - Pb + mg * (ha - hb);
        const int up = pot > 0 ? 0 : 1;
        const T kr = up == 0 ? pa.kr[1 + ph] : pb.kr[1 + ph];
        T t = K * (pa.ro[ph] * da * c1 + pb.ro[ph] * db * c2) / (pa.mu[ph] * da * c1 + pb.mu[ph] * db * c2) * kr;
        T tgv = mg * t;
        if (c1 == 0 && c2 == 0) {
            t = T(0);
            tgv = T(0);
        }
        if (ph == 0) { f.to = t; f.tgo = tgv; f.up_o = up; }
        else { f.tg = t; f.tgg = tgv; f.up_g = up; }
    }
    return f;
}

// ---------------------------------------------------------------- wells
// Wells of IteratorFunction (Petsc_Functions.cpp:164-697 fracture blocks,
// :3505-4040 matrix blocks; the liquid and gas blocks repeat the same code at
// :1329, :2421, :4683, :5810), Calc_Re (Minor_NR.cpp:2700-3249) and
// FlashInjection (Init_Calcs.cpp:1161-1377).
constexpr double PI_R = 3.1415926535897932384626433;  // Functions.h:218
constexpr double WELL_MIN_P = 7000000;                // Functions.h:6
constexpr double WELL_MAX_P = 14000000;               // Functions.h:5
constexpr double TSTEPIF = 1e-5;                      // Functions.h:265
constexpr int W_OF = 0, W_GF = 1, W_OD = 2, W_GD = 3, W_OV = 4, W_GV = 5;  // after the 2 nc compositions

// Calc_Re (Minor_NR.cpp:2700-3249): equivalent radius of the well in block
// (Ix, Iy) from the interface transmissibility factors and the image wells
// of the reservoir boundaries.  The well sits at X_posi = Y_posi = 0 of its
// block; the branches run in the reference's order on shared factors (on a
// one-block-wide grid more than one branch runs, each on what the previous
// left, as there).
double calc_re(const Model& M, int Ix, int Iy) {
    const double dxi = M.dx[Ix], dyi = M.dy[Iy];
    const int Nx = M.nx, Ny = M.ny;
    const double X = 0, Y = 0;
    double T[9], r1[9], rij[5][9], aj[5];
    for (int j = 0; j < 9; j++) { T[j] = 0; r1[j] = 0; for (int a = 0; a < 5; a++) rij[a][j] = 0; }
    for (int a = 0; a < 5; a++) aj[a] = 0;
    double bSum = 0, iZ = 1, jZ = 1, Re = 0, frac = 1;
    using std::sqrt; using std::pow; using std::exp;
    r1[1] = sqrt(X * X + (dyi + Y) * (dyi + Y));
    r1[2] = sqrt((dxi - X) * (dxi - X) + Y * Y);
    r1[3] = sqrt(X * X + (dyi - Y) * (dyi - Y));
    r1[4] = sqrt((dxi + X) * (dxi + X) + Y * Y);
    r1[5] = sqrt((dxi + X) * (dxi + X) + (dyi + Y) * (dyi + Y));
    r1[6] = sqrt((dxi - X) * (dxi - X) + (dyi + Y) * (dyi + Y));
    r1[7] = sqrt((dxi - X) * (dxi - X) + (dyi - Y) * (dyi - Y));
    r1[8] = sqrt((dxi + X) * (dxi + X) + (dyi - Y) * (dyi - Y));
    T[5] = T[6] = T[7] = T[8] = 0;
    T[1] = dxi / dyi - (T[5] + T[6]);
    T[2] = dyi / dxi - (T[6] + T[7]);
    T[3] = dxi / dyi - (T[7] + T[8]);
    T[4] = dyi / dxi - (T[5] + T[8]);
    auto sum_b = [&]() { for (int j = 1; j < 9; j++) bSum += T[j]; bSum = 1 / bSum; };
    auto edge = [&](const int* zero) {
        // one image well: rij[2][.] / aj[2]
        for (int q = 0; q < 3; q++) { rij[2][zero[q]] = 1; r1[zero[q]] = 1; T[zero[q]] = 0; }
        sum_b();
        for (int j = 1; j < 5; j++) iZ *= pow(r1[j], T[j]) * pow(rij[2][j] / aj[2], T[j]);
        Re = pow(exp(-2 * PI_R * frac) * iZ, bSum);
    };
    auto corner = [&](const int* keep) {
        // three image wells: rij[2..4][.] / aj[2..4]
        for (int j = 1; j < 9; j++)
            if (j != keep[0] && j != keep[1] && j != keep[2]) {
                T[j] = 0; r1[j] = 1; rij[2][j] = 1; rij[3][j] = 1; rij[4][j] = 1;
            }
        sum_b();
        iZ = 1;
        for (int t = 1; t < 5; t++) {
            jZ = 1;
            for (int j = 2; j < 5; j++) jZ *= pow(rij[j][t] / aj[j], T[t]);
            iZ *= pow(r1[t], T[t]) * jZ;
        }
        Re = pow(exp(-2 * PI_R * frac) * iZ, bSum);
    };
    if (0 < Ix && Ix < Nx - 1 && 0 < Iy && Iy < Ny - 1) {   // interior block
        frac = 1;
        for (int j = 1; j < 5; j++) iZ *= pow(r1[j], T[j]);
        sum_b();
        Re = pow(exp(-2 * PI_R * frac) * iZ, bSum);
    }
    if (Iy == Ny - 1 && Ix != 0 && Ix != Nx - 1) {          // boundary Iy = Ny - 1
        frac = 1;
        aj[2] = dyi + 2 * Y;
        rij[2][2] = sqrt((dxi - X) * (dxi - X) + (dyi + Y) * (dyi + Y));
        rij[2][7] = sqrt((dxi - X) * (dxi - X) + (2 * dyi + Y) * (2 * dyi + Y));
        rij[2][3] = sqrt(X * X + (2 * dyi + Y) * (2 * dyi + Y));
        rij[2][8] = sqrt((dxi + X) * (dxi + X) + (2 * dyi + Y) * (2 * dyi + Y));
        rij[2][4] = sqrt((dxi + X) * (dxi + X) + (dyi + Y) * (dyi + Y));
        const int z[3] = {1, 5, 6};
        edge(z);
    }
    if (Iy == 0 && Ix != 0 && Ix != Nx - 1) {               // boundary Iy = 0
        frac = 1;
        aj[2] = dyi - 2 * Y;
        rij[2][2] = sqrt((dxi - X) * (dxi - X) + (dyi - Y) * (dyi - Y));
        rij[2][6] = sqrt((dxi - X) * (dxi - X) + (2 * dyi - Y) * (2 * dyi - Y));
        rij[2][1] = sqrt(X * X + (2 * dyi - Y) * (2 * dyi - Y));
        rij[2][5] = sqrt((dxi + X) * (dxi + X) + (2 * dyi - Y) * (2 * dyi - Y));
        rij[2][4] = sqrt((dxi + X) * (dxi + X) + (dyi - Y) * (dyi - Y));
        const int z[3] = {3, 7, 8};
        edge(z);
    }
    if (Ix == Nx - 1 && Iy != 0 && Iy != Ny - 1) {          // boundary Ix = Nx - 1
        frac = 1;
        aj[2] = dxi - 2 * X;
        rij[2][1] = sqrt((dxi - X) * (dxi - X) + (dyi + Y) * (dyi + Y));
        rij[2][5] = sqrt((2 * dxi - X) * (2 * dxi - X) + (dyi + Y) * (dyi + Y));
        rij[2][4] = sqrt((2 * dxi - X) * (2 * dxi - X) + Y * Y);
        rij[2][8] = sqrt((2 * dxi - X) * (2 * dxi - X) + (dyi - Y) * (dyi - Y));
        rij[2][3] = sqrt((dxi - X) * (dxi - X) + (dyi - Y) * (dyi - Y));
        const int z[3] = {2, 7, 6};
        edge(z);
    }
    if (Ix == 0 && Iy != 0 && Iy != Ny - 1) {               // boundary Ix = 0
        frac = 1;
        aj[2] = dxi + 2 * X;
        rij[2][1] = sqrt((dxi + X) * (dxi + X) + (dyi + Y) * (dyi + Y));
        rij[2][7] = sqrt((2 * dxi + X) * (2 * dxi + X) + (dyi - Y) * (dyi - Y));
        rij[2][2] = sqrt((2 * dxi + X) * (2 * dxi + X) + Y * Y);
        rij[2][6] = sqrt((2 * dxi + X) * (2 * dxi + X) + (dyi + Y) * (dyi + Y));
        rij[2][3] = sqrt((dxi + X) * (dxi + X) + (dyi - Y) * (dyi - Y));
        const int z[3] = {8, 4, 5};
        edge(z);
    }
    if (Ix == 0 && Iy == Ny - 1) {                          // corner
        frac = 1;
        aj[2] = dyi + 2 * Y;
        aj[4] = dxi + 2 * X;
        aj[3] = sqrt((dyi + 2 * Y) * (dyi + 2 * Y) + (dxi + 2 * X) * (dxi + 2 * X));
        rij[2][2] = sqrt((dxi - X) * (dxi - X) + (dyi + Y) * (dyi + Y));
        rij[2][7] = sqrt((dxi - X) * (dxi - X) + (2 * dyi + Y) * (2 * dyi + Y));
        rij[2][3] = sqrt(X * X + (2 * dyi + Y) * (2 * dyi + Y));
        rij[3][2] = sqrt((2 * dxi + X) * (2 * dxi + X) + (dyi + Y) * (dyi + Y));
        rij[3][3] = sqrt((dxi + X) * (dxi + X) + (2 * dyi + Y) * (2 * dyi + Y));
        rij[3][7] = sqrt((2 * dxi + X) * (2 * dxi + X) + (2 * dyi + Y) * (2 * dyi + Y));
        rij[4][2] = sqrt((2 * dxi + X) * (2 * dxi + X) + Y * Y);
        rij[4][3] = sqrt((dxi + X) * (dxi + X) + (dyi - Y) * (dyi - Y));
        rij[4][7] = sqrt((2 * dxi + X) * (2 * dxi + X) + (dyi - Y) * (dyi - Y));
        const int k3[3] = {2, 3, 7};
        corner(k3);
    }
    if (Ix == 0 && Iy == 0) {                               // corner
        frac = 1;
        aj[2] = dyi - 2 * Y;
        aj[4] = dxi + 2 * X;
        aj[3] = sqrt((dyi - 2 * Y) * (dyi - 2 * Y) + (dxi + 2 * X) * (dxi + 2 * X));
        rij[2][2] = sqrt((dxi - X) * (dxi - X) + (dyi - Y) * (dyi - Y));
        rij[2][6] = sqrt((dxi - X) * (dxi - X) + (2 * dyi - Y) * (2 * dyi - Y));
        rij[2][1] = sqrt(X * X + (2 * dyi - Y) * (2 * dyi - Y));
        rij[3][2] = sqrt((2 * dxi + X) * (2 * dxi + X) + (dyi - Y) * (dyi - Y));
        rij[3][1] = sqrt((dxi + X) * (dxi + X) + (2 * dyi - Y) * (2 * dyi - Y));
        rij[3][6] = sqrt((2 * dxi + X) * (2 * dxi + X) + (2 * dyi - Y) * (2 * dyi - Y));
        rij[4][2] = sqrt((2 * dxi + X) * (2 * dxi + X) + Y * Y);
        rij[4][1] = sqrt((dxi + X) * (dxi + X) + (dyi + Y) * (dyi + Y));
        rij[4][6] = sqrt((2 * dxi + X) * (2 * dxi + X) + (dyi + Y) * (dyi + Y));
        const int k3[3] = {1, 2, 6};
        corner(k3);
    }
    if (Ix == Nx - 1 && Iy == Ny - 1) {                     // corner
        frac = 1;
        aj[2] = dxi - 2 * X;
        aj[4] = dyi + 2 * Y;
        aj[3] = sqrt((dxi - 2 * X) * (dxi - 2 * X) + (dyi + 2 * Y) * (dyi + 2 * Y));
        rij[2][4] = sqrt((2 * dxi - X) * (2 * dxi - X) + Y * Y);
        rij[2][8] = sqrt((2 * dxi - X) * (2 * dxi - X) + (dyi - Y) * (dyi - Y));
        rij[2][3] = sqrt((dxi - X) * (dxi - X) + (dyi - Y) * (dyi - Y));
        rij[3][4] = sqrt((2 * dxi - X) * (2 * dxi - X) + (dyi + Y) * (dyi + Y));
        rij[3][3] = sqrt((dxi - X) * (dxi - X) + (2 * dyi + Y) * (2 * dyi + Y));
        rij[3][8] = sqrt((2 * dxi - X) * (2 * dxi - X) + (2 * dyi + Y) * (2 * dyi + Y));
        rij[4][3] = sqrt((2 * dyi + Y) * (2 * dyi + Y) + X * X);
        rij[4][4] = sqrt((dxi + X) * (dxi + X) + (dyi + Y) * (dyi + Y));
        rij[4][8] = sqrt((dxi + X) * (dxi + X) + (2 * dyi + Y) * (2 * dyi + Y));
        const int k3[3] = {3, 4, 8};
        corner(k3);
    }
    if (Ix == Nx - 1 && Iy == 0) {                          // corner
        frac = 1;
        aj[2] = dxi - 2 * X;
        aj[4] = dyi - 2 * Y;
        aj[3] = sqrt((dxi - 2 * X) * (dxi - 2 * X) + (dyi - 2 * Y) * (dyi - 2 * Y));
        rij[2][1] = sqrt((dxi - X) * (dxi - X) + (dyi + Y) * (dyi + Y));
        rij[2][5] = sqrt((2 * dxi - X) * (2 * dxi - X) + (dyi + Y) * (dyi + Y));
        rij[2][4] = sqrt((2 * dxi - X) * (2 * dxi - X) + Y * Y);
        rij[3][4] = sqrt((2 * dxi - X) * (2 * dxi - X) + (dyi - Y) * (dyi - Y));
        rij[3][1] = sqrt((dxi - X) * (dxi - X) + (2 * dyi - Y) * (2 * dyi - Y));
        rij[3][5] = sqrt((2 * dxi - X) * (2 * dxi - X) + (2 * dyi - Y) * (2 * dyi - Y));
        rij[4][1] = sqrt((2 * dyi - Y) * (2 * dyi - Y) + X * X);
        rij[4][4] = sqrt((dxi + X) * (dxi + X) + (dyi - Y) * (dyi - Y));
        rij[4][5] = sqrt((dxi + X) * (dxi + X) + (2 * dyi - Y) * (2 * dyi - Y));
        const int k3[3] = {1, 4, 5};
        corner(k3);
    }
    return Re;
}

// FlashInjection (Init_Calcs.cpp:1161-1377): successive substitution on the
// injection stream z at pressure P from Wilson K-values, the Rachford-Rice
// Newton restarted at L = 0.6 every pass, until L moves by at most
// FLASH_TOL; then the phase split, densities and LBC viscosities.  A K-value
// turned negative makes the vapour composition y_i / (1 - L) of the previous
// pass's y_i (the reference divides the output array, :1219).
// out: [x (nc), y (nc), OilFraction, GasFraction, OilDensity, GasDensity,
//       OilViscosity, GasViscosity]
void flash_injection(const Model& M, const double* z, double P, double* out) {
    const int nc = M.nc;
    const double RT = RG * M.T;
    double K[NCM], x[NCM], y[NCM];
    for (int i = 0; i < nc; i++) {
        K[i] = (M.pc[i] / P) * std::exp(5.37 * (1 + M.om[i]) * (1 - M.tc[i] / M.T));
        x[i] = 0;
        y[i] = 0;
    }
    double L = 0.6, Lb = 10, ESum, Zl = 0, Zg = 0;
    int k = 0;
    const bool srk = M.eos == OR_SRK;
    do {
        L = 0.6;
        int j = 0;
        double tD;
        do {
            double F = 0, dF = 0;
            for (int i = 0; i < nc; i++) {
                if (K[i] >= 0) tD = (1 - K[i]) / (L + (1 - L) * K[i]);
                else tD = 1 / (L - 1);
                F += z[i] * tD;
                dF -= z[i] * tD * tD;
            }
            tD = F / dF;
            L -= tD;
            j++;
        } while ((tD * tD) > 1e-8 && j < 100);   // NEWTON_TOL, MAX_FLASH_N_ITR
        if (L > 1) L = 1;
        else if (L < 0) L = 0;
        ESum = std::fabs(Lb - L);
        Lb = L;
        for (int i = 0; i < nc; i++) {
            if (K[i] >= 0) {
                x[i] = z[i] / (L + K[i] * (1 - L));
                y[i] = K[i] * x[i];
            } else {
                x[i] = 0;
                y[i] = y[i] / (1 - L);
            }
        }
        double Al = 0, Bl = 0, Ag = 0, Bg = 0;
        for (int i = 0; i < nc; i++) {
            Bl += x[i] * M.b[i];
            Bg += y[i] * M.b[i];
            for (int jj = 0; jj < nc; jj++) {
                const double sq = std::sqrt(M.a[i] * M.a[jj]) * M.bic[i * nc + jj];
                Al += x[i] * x[jj] * sq;
                Ag += y[i] * y[jj] * sq;
            }
        }
        const double Cal = Al, Cbl = Bl, Cag = Ag, Cbg = Bg;
        Al *= P / (RG * RG * M.T * M.T);
        Bl *= P / RT;
        Ag *= P / (RG * RG * M.T * M.T);
        Bg *= P / RT;
        const double Ul = srk ? Bl : 2 * Bl, Wl = srk ? 0 : Bl;
        const double Ug = srk ? Bg : 2 * Bg, Wg = srk ? 0 : Bg;
        Zl = or_solve_z(-(1 + Bl - Ul), Al - Bl * Ul - Ul - Wl * Wl, -(Al * Bl - Bl * Wl * Wl - Wl * Wl), 0);
        Zg = or_solve_z(-(1 + Bg - Ug), Ag - Bg * Ug - Ug - Wg * Wg, -(Ag * Bg - Bg * Wg * Wg - Wg * Wg), 1);
        for (int i = 0; i < nc; i++) {
            double tl = 0, tg = 0;
            for (int jj = 0; jj < nc; jj++) {
                const double sq = std::sqrt(M.a[i] * M.a[jj]) * M.bic[i * nc + jj];
                tl += x[jj] * sq;
                tg += y[jj] * sq;
            }
            const double t1l = M.b[i] / Cbl, t1g = M.b[i] / Cbg;
            const double EEl = 2 * tl / Cal - t1l, EEg = 2 * tg / Cag - t1g;
            const double DDl = Al / (Bl * (M.d1 - M.d2)), DDg = Ag / (Bg * (M.d1 - M.d2));
            const double FFl = std::log((Zl + M.d2 * Bl) / (Zl + M.d1 * Bl));
            const double FFg = std::log((Zg + M.d2 * Bg) / (Zg + M.d1 * Bg));
            const double phiL = std::exp(t1l * (Zl - 1) - std::log(Zl - Bl) + DDl * EEl * FFl);
            const double phiG = std::exp(t1g * (Zg - 1) - std::log(Zg - Bg) + DDg * EEg * FFg);
            const double fL = x[i] * phiL, fG = y[i] * phiG;
            if (z[i] != 0) {
                if (fG != 0) K[i] *= fL / fG;
                else K[i] = -1;
                if (fL != fL) K[i] = -1;
                if (fG != fG) K[i] = 1;
            }
        }
        k++;
    } while (ESum > 1e-14 && k < 500);   // FLASH_TOL
    double OF, GF;
    if (std::fabs(L) < 1e-20) { OF = 0; GF = 1; }
    else if (std::fabs(1 - L) < 1e-20) { OF = 1; GF = 0; }
    else { OF = L; GF = 1 - L; }
    const double OD = P / (Zl * RT), GD = P / (Zg * RT);
    double mu[2];
    for (int ph = 0; ph < 2; ph++) {
        const double* w = ph == 0 ? x : y;
        const double ro = ph == 0 ? OD : GD;
        double s1 = 0, s2 = 0, rv = 0, tcm = 0, pcm = 0, mwm = 0;
        for (int i = 0; i < nc; i++) {
            const double tm = w[i] * std::sqrt(M.mw[i]);
            s1 += tm * M.mu0[i];
            s2 += tm;
            rv += w[i] * M.vcm3[i];
            tcm += w[i] * M.tc[i];
            pcm += w[i] * M.pc[i] / 101325;
            mwm += w[i] * M.mw[i];
        }
        const double Er = ro * rv;
        const double eta = std::pow(tcm, 1.0 / 6) / (std::sqrt(mwm) * std::pow(pcm, 2.0 / 3));
        const double r1 = 0.10230 + 0.023364 * Er + 0.058533 * Er * Er - 0.040758 * Er * Er * Er +
                          0.0093324 * Er * Er * Er * Er;
        mu[ph] = s1 / s2 + (r1 * r1 * r1 * r1 - 1e-4) / eta;
    }
    for (int i = 0; i < nc; i++) { out[i] = x[i]; out[nc + i] = y[i]; }
    double* t = out + 2 * nc;
    t[W_OF] = OF; t[W_GF] = GF; t[W_OD] = OD; t[W_GD] = GD; t[W_OV] = mu[0]; t[W_GV] = mu[1];
}

void well_setup(Model& M, const or_wells* w) {
    const int nc = M.nc, nw = w->nw;
    M.nw = nw;
    M.welli.assign(w->welli, w->welli + 5 * nw);
    M.wellf.assign(w->wellf, w->wellf + (nc + 6) * nw);
    M.wcond.assign(nw, 0);
    if (w->condition) M.wcond.assign(w->condition, w->condition + nw);
    M.sim_time = w->sim_time;
    M.re_f = w->re_f;
    M.well_re.assign(nw, 0);
    M.inj.assign((size_t)nw * (2 * nc + 6), 0);
    for (int i = 0; i < nw; i++) {
        M.well_re[i] = calc_re(M, M.welli[5 * i + 1], M.welli[5 * i + 2]);
        const int type = M.welli[5 * i];
        const double* wf = &M.wellf[(size_t)(nc + 6) * i];
        if (type == 3 || type == 8)
            flash_injection(M, wf + 4, type == 3 ? wf[0] : wf[1], &M.inj[(size_t)(2 * nc + 6) * i]);
    }
}

// Rates of the wells completed in block (k, c) (one evaluation per block, as
// the reference's loop over wells at :181-697 does it): the last active well
// that matches sets the rates; an injector (types 3, 8) marks the block so
// its component rows take the injection stream's compositions.  Rates are
// positive into the block.  The allocation sums over the well's completions
// (SSo1 ...) take the other blocks' properties as values: the derivative is
// the one with respect to the block's own unknowns.
template <class T>
struct WellQ {
    T Qw = T(0), Qo = T(0), Qg = T(0), PIw = T(0), PIo = T(0), PIg = T(0);
    int inj = -1;   // injector whose stream the component rows take, or -1
};

template <class T>
WellQ<T> well_terms(const Model& M, int k, int c, int ix, int iy, int iz, const Props<T>& p,
                    const std::vector<Props<T>>* Pp) {
    WellQ<T> q;
    const int nc = M.nc, n = M.n;
    const double* pk = M.perm[k].data();
    const double* px = M.perm[1 - k].data();
    const Props<T>& pxc = Pp[1 - k][c];
    auto nan_or = [](const T& v) { const double a = val(v); return (a - a) != 0; };
    for (int i = 0; i < M.nw; i++) {
        const int* wi = &M.welli[5 * i];
        if (!(ix == wi[1] && iy == wi[2] && iz >= wi[3] && iz <= wi[4])) continue;
        const double* wf = &M.wellf[(size_t)(nc + 6) * i];
        const int type = wi[0];
        const int ts = (type == 3 || type == 8) ? nc + 4 : 4;
        if (!((std::fabs(M.sim_time - wf[ts]) <= TSTEPIF) || (std::fabs(M.sim_time - wf[ts + 1]) <= TSTEPIF) ||
              ((M.sim_time >= wf[ts]) && (M.sim_time <= wf[ts + 1]))))
            continue;
        const double rw = wf[3], skin = wf[2];
        double Re = k == 1 ? M.re_f : M.well_re[i];
        if (k == 0 && (type == 0 || type == 3 || type == 8) && Re <= rw) Re += rw;
        const double Lf = std::log((rw + M.re_f) / rw) + skin;
        const double Lm = std::log(Re / rw) + skin;
        const double Lown = k == 1 ? Lf : Lm, Loth = k == 1 ? Lm : Lf;
        const double S1pi = 2 * SI_C * PI_R * M.dz[iz] * std::sqrt(pk[n + c] * pk[c]) / Lown;
        const double S1r = M.dz[iz] * std::sqrt(pk[c] * pk[n + c]) / Lown;
        const bool cond = M.wcond[i] != 0;
        const T krw = p.kr[0], kro = p.kr[1], krg = p.kr[2];
        const T roo = p.ro[0], rog = p.ro[1], muo = p.mu[0], mug = p.mu[1];
        const double watf = M.wat_ro * WAT_M_RO;
        auto set_pi = [&](double S1) {
            q.PIw = S1 * M.wat_ro * WAT_M_RO * krw / M.wat_mu;
            q.PIo = S1 * roo * kro / muo;
            q.PIg = S1 * rog * krg / mug;
        };
        auto pwf_rates = [&](double Pwf) {
            q.Qw = q.PIw * (Pwf - p.Pw);
            q.Qo = q.PIo * (Pwf - p.Po);
            q.Qg = q.PIg * (Pwf - p.Pg);
        };
        auto clamp_prod = [&]() {
            if (val(q.Qw) > 0 || nan_or(q.Qw)) { q.Qw = T(0); q.PIw = T(0); }
            if (val(q.Qo) > 0 || nan_or(q.Qo)) { q.Qo = T(0); q.PIo = T(0); }
            if (val(q.Qg) > 0 || nan_or(q.Qg)) { q.Qg = T(0); q.PIg = T(0); }
        };
        // completion sums over layers j: own continuum first, then the other;
        // per completion f(ro, kr, mu of that block) x S2, over mu_den (the
        // phase viscosity of this layer in the completion's continuum)
        auto own_at = [&](int j) -> Props<T> {
            if (j == iz) return p;
            Props<T> r;
            const Props<T>& s = Pp[k][cid(M, ix, iy, j)];
            for (int ph = 0; ph < 2; ph++) { r.ro[ph] = T(val(s.ro[ph])); r.mu[ph] = T(val(s.mu[ph])); }
            for (int a = 0; a < 3; a++) r.kr[a] = T(val(s.kr[a]));
            return r;
        };
        auto oth_at = [&](int j) -> Props<T> {
            Props<T> r;
            const Props<T>& s = Pp[1 - k][cid(M, ix, iy, j)];
            for (int ph = 0; ph < 2; ph++) { r.ro[ph] = T(val(s.ro[ph])); r.mu[ph] = T(val(s.mu[ph])); }
            for (int a = 0; a < 3; a++) r.kr[a] = T(val(s.kr[a]));
            return r;
        };
        auto S2own = [&](int j) { const int cj = cid(M, ix, iy, j); return M.dz[j] * std::sqrt(pk[cj] * pk[n + cj]) / Lown; };
        auto S2oth = [&](int j) { const int cj = cid(M, ix, iy, j); return M.dz[j] * std::sqrt(px[cj] * px[n + cj]) / Loth; };
        // SSo1 of a phase ph (ro S2 kr / mu of this layer); ph = 2: water (watRo S2 krw / watMu)
        auto sso1 = [&](int ph) {
            T s = T(0);
            const T muown = ph < 2 ? p.mu[ph] : T(0);
            const T muoth = ph < 2 ? T(val(pxc.mu[ph])) : T(0);
            for (int j = wi[3]; j <= wi[4]; j++) {
                const Props<T> a = own_at(j), b = oth_at(j);
                if (ph < 2) {
                    s += a.ro[ph] * S2own(j) * a.kr[1 + ph] / muown;
                    s += b.ro[ph] * S2oth(j) * b.kr[1 + ph] / muoth;
                } else {
                    s += M.wat_ro * S2own(j) * a.kr[0] / M.wat_mu;
                    s += M.wat_ro * S2oth(j) * b.kr[0] / M.wat_mu;
                }
            }
            return s;
        };
        auto sso1_geo = [&]() {
            double s = 0;
            for (int j = wi[3]; j <= wi[4]; j++) { s += S2own(j); s += S2oth(j); }
            return s;
        };
        const double* ij = &M.inj[(size_t)(2 * nc + 6) * i + 2 * nc];
        switch (type) {
            case 0:   // constant Pwf, oil production (:186-266)
                set_pi(S1pi);
                pwf_rates(wf[0]);
                if (val(p.Po) < WELL_MIN_P) {
                    q.Qw = q.Qo = q.Qg = T(0);
                    q.PIw = q.PIo = q.PIg = T(0);
                } else if (wf[1] != 0 && -val(q.Qo) > wf[1]) {
                    q.Qo = T(-wf[1]);
                    q.Qw = q.Qo * krw * M.wat_ro * WAT_M_RO * muo / (kro * roo * M.wat_mu);
                    q.Qg = q.Qo * krg * rog * muo / (kro * roo * mug);
                    q.PIw = q.PIo = q.PIg = T(0);
                }
                clamp_prod();
                break;
            case 1:   // constant Pwf, gas production (:268-318)
                set_pi(S1pi);
                pwf_rates(wf[0]);
                clamp_prod();
                if (cond) {
                    const T s = sso1(1);
                    if (val(s) != 0) q.Qg = -S1r * wf[1] * rog * krg / (mug * s);
                    q.Qw = q.Qg * krw * M.wat_ro * WAT_M_RO * mug / (krg * rog * M.wat_mu);
                    q.Qo = q.Qg * kro * roo * mug / (krg * rog * muo);
                }
                break;
            case 2:   // constant Pwf, water injection (:320-350)
                q.PIw = S1pi * M.wat_ro * WAT_M_RO * krw / M.wat_mu;
                q.Qw = q.PIw * (wf[0] - p.Pw);
                if (val(q.Qw) < 0 || nan_or(q.Qw)) { q.Qw = T(0); q.PIw = T(0); }
                if (cond) {
                    const T s = sso1(2);
                    if (val(s) != 0) q.Qw = S1r * wf[1] * M.wat_ro * WAT_M_RO * krw / (M.wat_mu * s);
                }
                break;
            case 3: {   // constant Pwf, gas injection (:352-400)
                q.PIg = S1pi * ij[W_GD] * M.sgt[4 * (M.nsgt - 1) + 1] * ij[W_GF] / ij[W_GV];
                q.Qg = q.PIg * (wf[0] - p.Pg);
                q.PIo = q.PIg * ij[W_OF] / ij[W_GF];
                q.Qo = q.PIo * (wf[0] - p.Pg);
                if (val(q.Qo) < 0 || nan_or(q.Qo)) { q.Qo = T(0); q.PIo = T(0); }
                if (val(q.Qg) < 0 || nan_or(q.Qg)) { q.Qg = T(0); q.PIg = T(0); }
                q.inj = i;
                if (cond) {
                    const double s = sso1_geo();
                    if (s != 0) q.Qg = T(S1r * wf[1] / s);
                    if (s != 0) q.Qo = q.Qg * ij[W_OF] / ij[W_GF];
                }
                break;
            }
            case 4: {   // constant oil rate (:402-461)
                const T s = sso1(0);
                if (val(s) != 0) q.Qo = -S1r * wf[0] * roo * kro / (muo * s);
                q.Qw = q.Qo * krw * M.wat_ro * WAT_M_RO * muo / (kro * roo * M.wat_mu);
                q.Qg = q.Qo * krg * rog * muo / (kro * roo * mug);
                if (cond) {
                    set_pi(S1pi);
                    pwf_rates(wf[0]);
                    clamp_prod();
                }
                if (k == 1 && val(p.Po) < WELL_MIN_P) {
                    q.Qw = q.Qo = q.Qg = T(0);
                    q.PIw = q.PIo = q.PIg = T(0);
                }
                break;
            }
            case 5: {   // constant gas rate (:463-515)
                const T s = sso1(1);
                if (val(s) != 0) q.Qg = -S1r * wf[0] * rog * krg / (mug * s);
                q.Qw = q.Qg * krw * M.wat_ro * WAT_M_RO * mug / (krg * rog * M.wat_mu);
                q.Qo = q.Qg * kro * roo * mug / (krg * rog * muo);
                if (cond) {
                    set_pi(S1pi);
                    pwf_rates(wf[0]);
                    clamp_prod();
                }
                break;
            }
            case 6: {   // constant total production rate (:517-576)
                T SSw = T(0), SSo = T(0), SSg = T(0), s = T(0);
                for (int j = wi[3]; j <= wi[4]; j++) {
                    for (int side = 0; side < 2; side++) {
                        const Props<T> a = side == 0 ? own_at(j) : oth_at(j);
                        const T mo_den = side == 0 ? p.mu[0] : T(val(pxc.mu[0]));
                        const T S1 = (a.kr[0] / M.wat_mu) + (a.kr[1] / a.mu[0]) + (a.kr[2] / a.mu[1]);
                        SSw += (a.kr[0] / M.wat_mu) / S1;
                        SSo += (a.kr[1] / a.mu[0]) / S1;
                        SSg += (a.kr[2] / a.mu[1]) / S1;
                        s += a.ro[0] * (side == 0 ? S2own(j) : S2oth(j)) * a.kr[1] / mo_den;
                    }
                }
                const T QoP = SSo * wf[0] / (SSw + SSo + SSg);
                if (val(s) != 0) q.Qo = -S1r * QoP * roo * kro / (muo * s);
                q.Qw = q.Qo * krw * M.wat_ro * WAT_M_RO * muo / (kro * roo * M.wat_mu);
                q.Qg = q.Qo * krg * rog * muo / (kro * roo * mug);
                if (cond) {
                    set_pi(S1pi);
                    pwf_rates(wf[0]);
                    clamp_prod();
                }
                break;
            }
            case 7:   // constant water injection rate (:578-604): its rate
                      // branch tests a sum that is never set, so only the
                      // Pwf-limited branch ever injects
                if (cond) {
                    q.PIw = S1pi * M.wat_ro * WAT_M_RO * krw / M.wat_mu;
                    q.Qw = q.PIw * (wf[0] - p.Pw);
                    if (val(q.Qw) < 0 || nan_or(q.Qw)) { q.Qw = T(0); q.PIw = T(0); }
                }
                break;
            case 8: {   // constant gas injection rate (:606-691)
                const double s = sso1_geo();
                if (s != 0) q.Qg = T(S1r * wf[0] / s);
                if (s != 0) q.Qo = q.Qg * ij[W_OF] / ij[W_GF];
                if (cond) {
                    q.PIg = S1pi * ij[W_GD] * M.sgt[4 * (M.nsgt - 1) + 1] * ij[W_GF] / ij[W_GV];
                    q.Qg = q.PIg * (wf[1] - p.Pg);
                    q.PIo = q.PIg * ij[W_OF] / ij[W_GF];
                    q.Qo = q.PIo * (wf[1] - p.Pg);
                    if (val(q.Qo) < 0 || nan_or(q.Qo)) { q.Qo = T(0); q.PIo = T(0); }
                    if (val(q.Qg) < 0 || nan_or(q.Qg)) { q.Qg = T(0); q.PIg = T(0); }
                }
                if (val(p.Po) > WELL_MAX_P) {
                    q.Qw = q.Qo = q.Qg = T(0);
                    q.PIw = q.PIo = q.PIg = T(0);
                }
                q.inj = i;
                break;
            }
            default:
                break;
        }
        (void)watf;
    }
    return q;
}

// ------------------------------------------------------------- residual
template <class T>
struct Neigh {
    int valid;
    int c;
    double dl;      // neighbour cell size along the direction
    double area;    // face area for the diffusion term (2 A / (dl + dself))
    double h;
};

// Row-scale mode: every term enters |term| so a row's value is the sum of
// the magnitudes it is built from (the yardstick of the parity tests).
thread_local bool g_absmode = false;
template <class T> inline void acc_add(T& a, const T& t) { a += g_absmode ? T(std::fabs(val(t))) : t; }
inline double absif(double v) { return g_absmode ? std::fabs(v) : v; }

// Residual rows of one block.  V/Pp are indexed [continuum][cell].
template <class T>
void block_rows(const Model& M, int k, int c, const std::vector<Vars<T>>* V, const std::vector<Props<T>>* Pp,
                const double* pre, T* out) {
    const int nc = M.nc;
    const int n = M.n;
    const int iz = c / (M.nx * M.ny), rem = c % (M.nx * M.ny), iy = rem / M.nx, ix = rem % M.nx;
    const Vars<T>& v = V[k][c];
    const Props<T>& p = Pp[k][c];
    const Vars<T>& vm = V[0][c];
    const Props<T>& pm = Pp[0][c];
    const Vars<T>& vf = V[1][c];
    const Props<T>& pf = Pp[1][c];
    const int lab = v.label;
    const int two = lab == 0;
    const int bs = two ? 2 * nc + 4 : nc + 3;
    for (int r = 0; r < bs; r++) out[r] = T(0);
    if (M.phi[k][c] < 1e-10) return;  // Petsc_Functions.cpp:6878-6883

    const double Vb = M.dx[ix] * M.dy[iy] * M.dz[iz];
    const double dsz[3] = {M.dx[ix], M.dy[iy], M.dz[iz]};
    const double hself = M.h[iz];

    // fugacity rows (two-phase): x_i Po phiL_i - y_i Pg phiG_i (:4559-4622)
    if (two) {
        T Cal, Cbl, Cag, Cbg, tl[NCM], tg[NCM];
        mix_t(M, v.x, Cal, Cbl, tl);
        mix_t(M, v.y, Cag, Cbg, tg);
        const T Al = Cal * p.Po / (RG * RG * M.T * M.T), Bl = Cbl * p.Po / (RG * M.T);
        const T Ag = Cag * p.Pg / (RG * RG * M.T * M.T), Bg = Cbg * p.Pg / (RG * M.T);
        const T Zl = p.Po / (p.ro[0] * RG * M.T);
        const T Zg = p.Pg / (p.ro[1] * RG * M.T);
        using std::exp; using std::log; using orad::exp; using orad::log;
        for (int i = 0; i < nc; i++) {
            const T t1l = M.b[i] / Cbl, t1g = M.b[i] / Cbg;
            const T EEl = 2 * tl[i] / Cal - t1l, EEg = 2 * tg[i] / Cag - t1g;
            const T DDl = Al / (Bl * (M.d1 - M.d2)), DDg = Ag / (Bg * (M.d1 - M.d2));
            const T FFl = log((Zl + M.d2 * Bl) / (Zl + M.d1 * Bl));
            const T FFg = log((Zg + M.d2 * Bg) / (Zg + M.d1 * Bg));
            const T phiL = exp(t1l * (Zl - 1) - log(Zl - Bl) + DDl * EEl * FFl);
            const T phiG = exp(t1g * (Zg - 1) - log(Zg - Bg) + DDg * EEg * FFg);
            out[i] = g_absmode ? T(std::fabs(val(v.x[i] * p.Po * phiL)) + std::fabs(val(v.y[i] * p.Pg * phiG)))
                               : v.x[i] * p.Po * phiL - v.y[i] * p.Pg * phiG;
        }
    }
    const int r0 = two ? nc : 0;   // first component-balance row

    // neighbours (+x, -x, +y, -y, +z, -z)
    int nb[6];
    nb[0] = ix < M.nx - 1 ? c + 1 : -1;
    nb[1] = ix > 0 ? c - 1 : -1;
    nb[2] = iy < M.ny - 1 ? c + M.nx : -1;
    nb[3] = iy > 0 ? c - M.nx : -1;
    nb[4] = iz < M.nz - 1 ? c + M.nx * M.ny : -1;
    nb[5] = iz > 0 ? c - M.nx * M.ny : -1;
    double nd[6], nh[6], narea[6];
    for (int q = 0; q < 6; q++) {
        const int d = q / 2;
        const int s = (q % 2 == 0) ? 1 : -1;
        int j;
        if (d == 0) { j = ix + s; nd[q] = (j >= 0 && j < M.nx) ? M.dx[j] : 0; nh[q] = hself; }
        else if (d == 1) { j = iy + s; nd[q] = (j >= 0 && j < M.ny) ? M.dy[j] : 0; nh[q] = hself; }
        else { j = iz + s; nd[q] = (j >= 0 && j < M.nz) ? M.dz[j] : 0; nh[q] = (j >= 0 && j < M.nz) ? M.h[j] : 0; }
        narea[q] = d == 0 ? M.dy[iy] * M.dz[iz] : (d == 1 ? M.dx[ix] * M.dz[iz] : M.dx[ix] * M.dy[iy]);
    }
    // face terms, oriented: for q even the face is (self=a, nb=b); odd (nb=a, self=b)
    Face<T> F6[6];
    for (int q = 0; q < 6; q++) {
        if (nb[q] < 0) continue;
        const int d = q / 2;
        const int o = nb[q];
        if (q % 2 == 0)
            F6[q] = face_terms<T>(M, k, d, c, o, dsz[d], nd[q], hself, nh[q], v, p, V[k][o], Pp[k][o]);
        else
            F6[q] = face_terms<T>(M, k, d, o, c, nd[q], dsz[d], nh[q], hself, V[k][o], Pp[k][o], v, p);
    }

    // matrix-fracture transfer pieces (:4626-4680 upstream kr; :5257-5258)
    const double sf2 = 4 * M.sigma * (M.perm[0][c] / (M.dx[ix] * M.dx[ix]) + M.perm[0][n + c] / (M.dy[iy] * M.dy[iy]) +
                                      M.perm[0][2 * n + c] / (M.dz[iz] * M.dz[iz]));
    const bool sfoff = M.const_diff && M.cdiff == 0;
    const double sf1 = sfoff ? 0 : 4 * M.sigma * (1 / (M.dx[ix] * M.dx[ix]) + 1 / (M.dy[iy] * M.dy[iy]) + 1 / (M.dz[iz] * M.dz[iz]));
    const double sf3 = sfoff ? 0 : 4 * M.sigma * ((1 / M.dx[ix]) + (1 / M.dy[iy]) + (1 / M.dz[iz])) / M.lf;
    const T kro_mf = pm.Po >= pf.Po ? pm.kr[1] : pf.kr[1];
    const T krg_mf = pm.Pg >= pf.Pg ? pm.kr[2] : pf.kr[2];
    const T krw_mf = pm.Pw >= pf.Pw ? pm.kr[0] : pf.kr[0];
    const T phm = pm.porpor, phf = pf.porpor;
    const T muo = (pm.mu[0] * phm + pf.mu[0] * phf) / (phm + phf);
    const T mug = (pm.mu[1] * phm + pf.mu[1] * phf) / (phm + phf);
    const T dPo = pf.Po - pm.Po;
    const T dPg = (pf.Pg - pm.Pg) + (pf.pcg - pm.pcg);
    const T convo = SI_C * pm.ro[0] * sf2 * Vb * kro_mf / muo * dPo;   // times x_m,i
    const T convg = SI_C * pm.ro[1] * sf2 * Vb * krg_mf / mug * dPg;   // times y_m,i
    const double sgn = k == 0 ? 1.0 : -1.0;   // fracture rows take the opposite sign

    const T dcl = T(lab == -1 ? 0.0 : 1.0), dcg = T(lab == 1 ? 0.0 : 1.0);
    // wells completed in this block (Petsc_Functions.cpp:164-697, :3505-4040)
    const WellQ<T> wq = M.nw ? well_terms<T>(M, k, c, ix, iy, iz, p, Pp) : WellQ<T>();
    const double* injx = wq.inj >= 0 ? &M.inj[(size_t)(2 * nc + 6) * wq.inj] : nullptr;

    // component balances i = 0..nc-2, then total hydrocarbon
    for (int i = 0; i < nc; i++) {
        const bool total = (i == nc - 1);
        T acc;
        if (!total) acc = p.porpor * (v.x[i] * p.ro[0] * v.S[1] + v.y[i] * p.ro[1] * v.S[2]);
        else acc = p.porpor * (p.ro[0] * v.S[1] + p.ro[1] * v.S[2]);
        const T prev = T(pre ? pre[i] : val(acc));   // Pre_Mul (Do_Time_Step.cpp:54-67)
        const T LS = (Vb / M.dt) * (acc - prev);
        T RS = T(0);
        for (int q = 0; q < 6; q++) {
            if (nb[q] < 0) continue;
            const int o = nb[q];
            const Face<T>& f = F6[q];
            const Vars<T>& vo = V[k][o];
            const Props<T>& po = Pp[k][o];
            // upstream compositions: flag 0 -> cell a, 1 -> cell b
            const Vars<T>& ua_o = (q % 2 == 0) ? v : vo;   // cell a
            const Vars<T>& ub_o = (q % 2 == 0) ? vo : v;   // cell b
            T co = T(1), cg = T(1);
            if (!total) {
                co = (f.up_o == 0 ? ua_o : ub_o).x[i];
                cg = (f.up_g == 0 ? ua_o : ub_o).y[i];
            }
            // flux into self: (P_nb - P_self) for both orientations
            acc_add(RS, co * f.to * (po.Po - p.Po));
            acc_add(RS, cg * f.tg * (po.Pg - p.Pg));
            acc_add(RS, co * f.tgo * (nh[q] - hself));
            acc_add(RS, cg * f.tgg * (nh[q] - hself));
        }
        // convective matrix-fracture transfer
        if (!total) { acc_add(RS, sgn * vm.x[i] * convo); acc_add(RS, sgn * vm.y[i] * convg); }
        else {
            acc_add(RS, sgn * convo);
            acc_add(RS, sgn * convg);
            if (M.nw) { acc_add(RS, wq.Qo); acc_add(RS, wq.Qg); }   // :1143
        }
        if (!total) {
            // diffusive matrix-fracture transfer (:5259-5313 matrix rows,
            // :740-806 fracture rows); interface compositions are zero
            // (EdgeComp is zeroed before every row, :95-98).
            const int lm = vm.label, lf = vf.label;
            const T mo = sf1 * Vb * pm.porpor / M.tor[c] * vm.S[1] * pm.ro[0] * pm.D[0][i];
            const T mg = sf1 * Vb * pm.porpor / M.tor[c] * vm.S[2] * pm.ro[1] * pm.D[1][i];
            if (k == 0) {
                const T dxo = vf.x[i] - vm.x[i], dyo = vf.y[i] - vm.y[i];
                const T exo = T(0) - vm.x[i], eyo = T(0) - vm.y[i];
                if (lm == 1 && lf == 1) acc_add(RS, mo * dxo);
                if (lm == -1 && lf == -1) acc_add(RS, mg * dyo);
                if (lm == 0 && lf == 0) { acc_add(RS, mo * dxo); acc_add(RS, mg * dyo); }
                if (lm == 1 && lf == -1) acc_add(RS, mo * exo);
                if (lm == -1 && lf == 1) acc_add(RS, mg * eyo);
                if (lm == 0 && lf == 1) { acc_add(RS, mg * eyo); acc_add(RS, mo * dxo); }
                if (lm == 0 && lf == -1) { acc_add(RS, mg * dyo); acc_add(RS, mo * exo); }
                if (lm == 1 && lf == 0) { acc_add(RS, mo * exo); acc_add(RS, mo * dxo); }
                if (lm == -1 && lf == 0) { acc_add(RS, mg * dyo); acc_add(RS, mg * eyo); }
            } else {
                const T fo = sf3 * Vb * pf.porpor / M.tor[c] * vf.S[1] * pf.ro[0] * pf.D[0][i];
                const T fg = sf3 * Vb * pf.porpor / M.tor[c] * vf.S[2] * pf.ro[1] * pf.D[1][i];
                const T dxo = vf.x[i] - vm.x[i], dyo = vf.y[i] - vm.y[i];
                const T fxo = vf.x[i] - T(0), fyo = vf.y[i] - T(0);
                if (lm == 1 && lf == 1) acc_add(RS, -(mo * dxo));
                if (lm == -1 && lf == -1) acc_add(RS, -(mg * dyo));
                if (lm == 0 && lf == 0) { acc_add(RS, -(mo * dxo)); acc_add(RS, -(mg * dyo)); }
                if (lm == 1 && lf == -1) acc_add(RS, -(fg * fyo));
                if (lm == -1 && lf == 1) acc_add(RS, -(fo * fxo));
                if (lm == 0 && lf == 1) { acc_add(RS, -(fo * fxo)); acc_add(RS, -(mo * dxo)); }
                if (lm == 0 && lf == -1) { acc_add(RS, -(mg * dyo)); acc_add(RS, -(fg * fyo)); }
                if (lm == 1 && lf == 0) { acc_add(RS, -(fg * fyo)); acc_add(RS, -(mo * dxo)); }
                if (lm == -1 && lf == 0) { acc_add(RS, -(mg * dyo)); acc_add(RS, -(fo * fxo)); }
            }
            // wells: the block's own compositions, or an injector's stream
            // (:787-808)
            if (M.nw) {
                const T oc = injx ? T(injx[i]) : v.x[i];
                const T gc = injx ? T(injx[nc + i]) : v.y[i];
                acc_add(RS, oc * wq.Qo);
                acc_add(RS, gc * wq.Qg);
            }
            // inter-cell Fickian flux (:5346-5656): thickness-weighted
            // density x saturation over tortuosity/porosity (tortuosity 1
            // in the fracture), upstream-by-composition D.
            for (int q = 0; q < 6; q++) {
                if (nb[q] < 0) continue;
                const int o = nb[q];
                const int d = q / 2;
                const Vars<T>& vo = V[k][o];
                const Props<T>& po = Pp[k][o];
                const double dl = dsz[d], dn = nd[q];
                const T pp2 = po.porpor;
                const double NN = 2 * narea[q] / (dn + dl);
                const double tn = k == 0 ? M.tor[o] : 1.0, ts = k == 0 ? M.tor[c] : 1.0;
                for (int ph = 0; ph < 2; ph++) {
                    const int absent = ph == 0 ? -1 : 1;
                    const double Cf = vo.label == absent ? 0 : 1;
                    const T dcs = ph == 0 ? dcl : dcg;
                    const T xs = ph == 0 ? v.x[i] : v.y[i];
                    const T xn = ph == 0 ? vo.x[i] : vo.y[i];
                    T Dup = (xn > xs) ? po.D[ph][i] : p.D[ph][i];
                    if (M.phi[k][c] < 1e-10 || M.phi[k][o] < 1e-10) Dup = T(0);
                    const T AA = (po.ro[ph] * dn * Cf + p.ro[ph] * dl * dcs) *
                                 ((vo.S[1 + ph] * dn * Cf + v.S[1 + ph] * dl * dcs) / (dn + dl));
                    const T BB = tn * dn * Cf / pp2 + ts * dl * dcs / p.porpor;
                    T Dt = AA / BB * NN * Dup;
                    if (Cf == 0 && val(dcs) == 0) Dt = T(0);
                    acc_add(RS, Dt * (xn - xs));
                }
            }
        }
        out[r0 + i] = g_absmode ? T(std::fabs(val(LS)) + val(RS)) : LS - RS;
    }
    // water balance
    {
        const T acc = p.porpor * M.wat_ro * WAT_M_RO * v.S[0];
        const T LS = (Vb / M.dt) * (acc - T(pre ? pre[nc] : val(acc)));
        T RS = T(0);
        for (int q = 0; q < 6; q++) {
            if (nb[q] < 0) continue;
            const int o = nb[q];
            const Props<T>& po = Pp[k][o];
            acc_add(RS, F6[q].tw * (po.Pw - p.Pw));
            acc_add(RS, G_ACC * M.wat_ro * F6[q].tw * (nh[q] - hself));
        }
        acc_add(RS, sgn * (SI_C * M.wat_ro * WAT_M_RO * sf2 * krw_mf / M.wat_mu * Vb *
                     ((pf.Pw - pm.Pw) + (pm.pcw - pf.pcw))));
        if (M.nw) acc_add(RS, wq.Qw);   // :1168
        out[r0 + nc] = g_absmode ? T(std::fabs(val(LS)) + val(RS)) : LS - RS;
    }
    // local constraints
    if (two) {
        T sx = 0, sy = 0;
        for (int i = 0; i < nc; i++) { sx += v.x[i]; sy += v.y[i]; }
        out[2 * nc + 1] = g_absmode ? T(val(sx) + 1) : sx - 1;
        out[2 * nc + 2] = g_absmode ? T(val(sy) + 1) : sy - 1;
        out[2 * nc + 3] = g_absmode ? T(val(v.S[0] + v.S[1] + v.S[2]) + 1) : v.S[0] + v.S[1] + v.S[2] - 1;
    } else {
        T s = 0;
        for (int i = 0; i < nc; i++) s += (lab == 1 ? v.x[i] : v.y[i]);
        out[nc + 1] = g_absmode ? T(val(s) + 1) : s - 1;
        out[nc + 2] = g_absmode ? T(val(v.S[0] + v.S[1] + v.S[2]) + 1) : v.S[0] + v.S[1] + v.S[2] - 1;
    }
}

// ------------------------------------------------------------ state I/O

template <class T>
void load_vars(const Model& M, const or_state* s, int k, int c, Vars<T>& v) {
    const int n = M.n, nc = M.nc;
    const int8_t* lab = k == 0 ? s->label_m : s->label_f;
    const double* Po = k == 0 ? s->Po_m : s->Po_f;
    const double* x = k == 0 ? s->x_m : s->x_f;
    const double* y = k == 0 ? s->y_m : s->y_f;
    const double* S = k == 0 ? s->S_m : s->S_f;
    v.label = lab[c];
    v.Po = T(Po[c]);
    for (int i = 0; i < nc; i++) {
        v.x[i] = T(x[(int64_t)i * n + c]);
        v.y[i] = T(y[(int64_t)i * n + c]);
    }
    for (int j = 0; j < 3; j++) v.S[j] = T(S[(int64_t)j * n + c]);
}

inline int block_size(int nc, int label) { return label == 0 ? 2 * nc + 4 : nc + 3; }

// Seed the unknowns of one block (InitGuess2 order, PetscAUX.cpp:892-957):
// liquid [x, P/PCoeff, Sw, So]; gas [y, P/PCoeff, Sw, Sg]; two-phase
// [x, y, P/PCoeff, Sw, So, Sg].
inline void seed(Vars<D>& v, int nc, double pcoeff) {
    const int bs = block_size(nc, v.label);
    int col = 0;
    auto set = [&](D& t, double scale) {
        t.n = bs;
        for (int q = 0; q < bs; q++) t.d[q] = 0;
        t.d[col++] = scale;
    };
    auto zero = [&](D& t) { t.n = bs; for (int q = 0; q < bs; q++) t.d[q] = 0; };
    for (int i = 0; i < nc; i++) { zero(v.x[i]); zero(v.y[i]); }
    zero(v.Po); zero(v.S[0]); zero(v.S[1]); zero(v.S[2]);
    if (v.label == 1) for (int i = 0; i < nc; i++) set(v.x[i], 1);
    else if (v.label == -1) for (int i = 0; i < nc; i++) set(v.y[i], 1);
    else {
        for (int i = 0; i < nc; i++) set(v.x[i], 1);
        for (int i = 0; i < nc; i++) set(v.y[i], 1);
    }
    set(v.Po, pcoeff);   // d/d(P/PCoeff) = PCoeff d/dP
    set(v.S[0], 1);
    if (v.label == 1) set(v.S[1], 1);
    else if (v.label == -1) set(v.S[2], 1);
    else { set(v.S[1], 1); set(v.S[2], 1); }
}

template <class A, class B>
void lift(const Vars<A>& a, Vars<B>& b, int nc) {
    b.label = a.label;
    b.Po = B(val(a.Po));
    for (int i = 0; i < nc; i++) { b.x[i] = B(val(a.x[i])); b.y[i] = B(val(a.y[i])); }
    for (int j = 0; j < 3; j++) b.S[j] = B(val(a.S[j]));
}

template <class B>
void liftp(const Props<double>& a, Props<B>& b, int nc) {
    b.porpor = B(a.porpor); b.Pw = B(a.Pw); b.Po = B(a.Po); b.Pg = B(a.Pg); b.bift = B(a.bift);
    for (int q = 0; q < 2; q++) { b.ro[q] = B(a.ro[q]); b.mu[q] = B(a.mu[q]); b.gam[q] = B(a.gam[q]); }
    for (int q = 0; q < 3; q++) b.kr[q] = B(a.kr[q]);
    b.pcw = B(a.pcw); b.pcg = B(a.pcg);
    for (int q = 0; q < 2; q++) for (int i = 0; i < nc; i++) b.D[q][i] = B(a.D[q][i]);
}

struct Sweep {
    std::vector<Vars<double>> V[2];
    std::vector<Props<double>> P[2];
    std::vector<double> rhoprev;   // [2n]: matrix oil/gas densities CalcIFT reads
    const or_state* st = nullptr;
    // preProp of (continuum, cell) gathered into buf, or NULL when not given
    const double* pre(int k, int n, int nc, int c, double* buf) const {
        const double* a = k == 0 ? st->pre_m : st->pre_f;
        if (!a) return nullptr;
        for (int i = 0; i <= nc; i++) buf[i] = a[(int64_t)i * n + c];
        return buf;
    }
    const double* rp(int n, int c, double* buf) const {
        buf[0] = rhoprev[c];
        buf[1] = rhoprev[n + c];
        return buf;
    }
};

// Property sweep in double for every cell.  CalcIFT reads the densities of
// the previous sweep; without them (rho_prev_m NULL) a first sweep with
// zero densities supplies them (so its IFT is 0 and its gas phase pressure
// equals the oil pressure), and the second sweep is the one used.
void sweep(const Model& M, const or_state* s, Sweep& S, int nthreads = 1) {
    const int n = M.n;
    S.st = s;
    for (int k = 0; k < 2; k++) {
        S.V[k].resize(n);
        S.P[k].resize(n);
        for (int c = 0; c < n; c++) load_vars<double>(M, s, k, c, S.V[k][c]);
    }
    S.rhoprev.assign(2 * n, 0.0);
    for (int c = 0; c < n; c++) {
        if (s->rho_prev_m) {
            S.rhoprev[c] = s->rho_prev_m[c];
            S.rhoprev[n + c] = s->rho_prev_m[n + c];
        } else {
            Props<double> tmp;
            const double zero[2] = {0, 0};
            cell_props<double>(M, 0, c, S.V[0][c], zero, tmp);
            S.rhoprev[c] = tmp.ro[0];
            S.rhoprev[n + c] = tmp.ro[1];
        }
    }
    for (int k = 0; k < 2; k++) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int c = 0; c < n; c++) {
            double rp[2];
            cell_props<double>(M, k, c, S.V[k][c], S.rp(n, c, rp), S.P[k][c]);
        }
    }
}

}  // namespace

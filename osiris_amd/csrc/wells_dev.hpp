// wells_dev.hpp -- well terms of the residual (included by assembly.hip after
// the residual pieces).
//
// IteratorFunction evaluates, for every block, the wells completed in it
// (Petsc_Functions.cpp:164-697 fracture blocks, :3505-4040 matrix blocks; the
// liquid and gas blocks repeat the same code at :1329, :2421, :4683, :5810):
// the last active matching well sets the water, oil and gas rates Qw, Qo, Qg
// (positive into the block), added to the right side of the component rows
// as x_i Qo + y_i Qg (an injector's stream compositions instead, :787-808),
// of the total row as Qo + Qg (:1143) and of the water row as Qw (:1168).
// The allocation sums over the well's completed layers read the other
// blocks' property-table values: the derivative (forward mode) is the one
// with respect to the block's own unknowns.
#pragma once

namespace osr {

constexpr double PI_W = 3.1415926535897932384626433;   // Functions.h:218
constexpr double WELL_MIN_P = 7000000;                 // Functions.h:6
constexpr double WELL_MAX_P = 14000000;                // Functions.h:5
constexpr double TSTEPIF = 1e-5;                       // Functions.h:265
// FlashInjection results after the 2 nc compositions
constexpr int W_OF = 0, W_GF = 1, W_OD = 2, W_GD = 3, W_OV = 4, W_GV = 5;

// The reference's test (Q - Q) != 0 on a stored rate: true for NaN and
// +-inf.  Written as !isfinite: with FMA contraction a compiler may fuse the
// subtraction into the product that formed Q, fma(a, b, -(a b)), whose
// value is the product's rounding error, not 0.
template <class T>
__device__ __forceinline__ bool nonfinite_or(const T& v) {
    return !isfinite(dval(v));
}

// Rates and productivity indices of the wells of one block, carried from
// well to well as the reference's loop carries its variables (a later well
// overwrites what its type sets and leaves the rest).
template <class T>
struct WellAcc {
    T Qw, Qo, Qg, PIw, PIo, PIg;
    int inj;   // injector whose stream the component rows take, or -1
};

// One well i of the block's loop (Petsc_Functions.cpp:181-697), if it is
// completed in block (k, c) and open at simTime.
template <int NC, class T>
__device__ __forceinline__ void well_one(const AsmArgs& A, int i, int k, int c, const Geo& G, const PropsT<NC, T>& p,
                                      WellAcc<T>& W) {
    const AsmConsts& K = *A.K;
    const int n = A.n;
    const int64_t s2 = 2 * (int64_t)n;
    const double* pk = k == 0 ? A.perm_m : A.perm_f;
    const double* px = k == 0 ? A.perm_f : A.perm_m;
    const double* tabx = A.props + (int64_t)(1 - k) * n;   // the other continuum's table
    const double* tabk = A.props + (int64_t)k * n;
    const int nxy = A.nx * A.ny;
    const int col0 = G.iy * A.nx + G.ix;                      // the block's column
    const double wtime = A.wsc[0], re_f = A.wsc[1];
    T& Qw = W.Qw;
    T& Qo = W.Qo;
    T& Qg = W.Qg;
    T& PIw = W.PIw;
    T& PIo = W.PIo;
    T& PIg = W.PIg;
    {
        const int* wi = A.wi + 5 * i;
        const double* wf = A.wf + (int64_t)(NC + 6) * i;
        const int type = wi[0];
        const int ts = (type == 3 || type == 8) ? NC + 4 : 4;
        if (!((fabs(wtime - wf[ts]) <= TSTEPIF) || (fabs(wtime - wf[ts + 1]) <= TSTEPIF) ||
              ((wtime >= wf[ts]) && (wtime <= wf[ts + 1]))))
            return;
        const double rw = wf[3], skin = wf[2];
        double Re = k == 1 ? re_f : A.wre[i];
        if (k == 0 && (type == 0 || type == 3 || type == 8) && Re <= rw) Re += rw;
        const double Lf = log((rw + re_f) / rw) + skin;
        const double Lm = log(Re / rw) + skin;
        const double Lown = k == 1 ? Lf : Lm, Loth = k == 1 ? Lm : Lf;
        const double S1pi = 2 * SI_C * PI_W * A.dz[G.iz] * sqrt(pk[n + c] * pk[c]) / Lown;
        const double S1r = A.dz[G.iz] * sqrt(pk[c] * pk[n + c]) / Lown;
        const bool cond = A.wcond[i] != 0;
        const T krw = p.kr[0], kro = p.kr[1], krg = p.kr[2];
        const T roo = p.ro[0], rog = p.ro[1], muo = p.mu[0], mug = p.mu[1];
        auto set_pi = [&](double S1) {
            PIw = S1 * K.wat_ro * WAT_M_RO * krw / K.wat_mu;
            PIo = S1 * roo * kro / muo;
            PIg = S1 * rog * krg / mug;
        };
        auto pwf_rates = [&](double Pwf) {
            Qw = PIw * (Pwf - p.Pw);
            Qo = PIo * (Pwf - p.Po);
            Qg = PIg * (Pwf - p.Pg);
        };
        auto clamp_prod = [&]() {
            if (dval(Qw) > 0 || nonfinite_or(Qw)) { Qw = T(0.0); PIw = T(0.0); }
            if (dval(Qo) > 0 || nonfinite_or(Qo)) { Qo = T(0.0); PIo = T(0.0); }
            if (dval(Qg) > 0 || nonfinite_or(Qg)) { Qg = T(0.0); PIg = T(0.0); }
        };
        // completion factor of layer j in this block's continuum and in the other
        auto S2own = [&](int j) { const int cj = j * nxy + col0; return A.dz[j] * sqrt(pk[cj] * pk[n + cj]) / Lown; };
        auto S2oth = [&](int j) { const int cj = j * nxy + col0; return A.dz[j] * sqrt(px[cj] * px[n + cj]) / Loth; };
        // property f of layer j in this block's continuum: the block's own
        // (T) value on its layer, the table's elsewhere
        auto ownf = [&](int j, int f, const T& mine) -> T {
            return j == G.iz ? mine : T(tabk[f * s2 + j * nxy + col0]);
        };
        auto othf = [&](int j, int f) -> double { return tabx[f * s2 + j * nxy + col0]; };
        // allocation sum of phase ph: ro S2 kr / mu of this layer (in each
        // completion's continuum); ph = 2: water, watRo S2 krw / watMu
        auto sso1 = [&](int ph) -> T {
            T s = T(0.0);
            for (int j = wi[3]; j <= wi[4]; j++) {
                if (ph < 2) {
                    const T ro = ownf(j, 5 + ph, p.ro[ph]), kr = ownf(j, 12 + ph, p.kr[1 + ph]);
                    s += ro * S2own(j) * kr / p.mu[ph];
                    s += othf(j, 5 + ph) * S2oth(j) * othf(j, 12 + ph) / tabx[(7 + ph) * s2 + c];
                } else {
                    const T kr = ownf(j, 11, p.kr[0]);
                    s += K.wat_ro * S2own(j) * kr / K.wat_mu;
                    s += K.wat_ro * S2oth(j) * othf(j, 11) / K.wat_mu;
                }
            }
            return s;
        };
        auto sso1_geo = [&]() {
            double s = 0;
            for (int j = wi[3]; j <= wi[4]; j++) {
                s += S2own(j);
                s += S2oth(j);
            }
            return s;
        };
        const double* ij = A.winj + (int64_t)(2 * NC + 6) * i + 2 * NC;
        const double sgt_kr = K.sgt[4 * (K.nsgt - 1) + 1];
        switch (type) {
            case 0:   // constant Pwf, oil production (:186-266)
                set_pi(S1pi);
                pwf_rates(wf[0]);
                if (dval(p.Po) < WELL_MIN_P) {
                    Qw = Qo = Qg = T(0.0);
                    PIw = PIo = PIg = T(0.0);
                } else if (wf[1] != 0 && -dval(Qo) > wf[1]) {
                    Qo = T(-wf[1]);
                    Qw = Qo * krw * K.wat_ro * WAT_M_RO * muo / (kro * roo * K.wat_mu);
                    Qg = Qo * krg * rog * muo / (kro * roo * mug);
                    PIw = PIo = PIg = T(0.0);
                }
                clamp_prod();
                break;
            case 1:   // constant Pwf, gas production (:268-318)
                set_pi(S1pi);
                pwf_rates(wf[0]);
                clamp_prod();
                if (cond) {
                    const T s = sso1(1);
                    if (dval(s) != 0) Qg = -S1r * wf[1] * rog * krg / (mug * s);
                    Qw = Qg * krw * K.wat_ro * WAT_M_RO * mug / (krg * rog * K.wat_mu);
                    Qo = Qg * kro * roo * mug / (krg * rog * muo);
                }
                break;
            case 2:   // constant Pwf, water injection (:320-350)
                PIw = S1pi * K.wat_ro * WAT_M_RO * krw / K.wat_mu;
                Qw = PIw * (wf[0] - p.Pw);
                if (dval(Qw) < 0 || nonfinite_or(Qw)) { Qw = T(0.0); PIw = T(0.0); }
                if (cond) {
                    const T s = sso1(2);
                    if (dval(s) != 0) Qw = S1r * wf[1] * K.wat_ro * WAT_M_RO * krw / (K.wat_mu * s);
                }
                break;
            case 3:   // constant Pwf, gas injection (:352-400)
                PIg = S1pi * ij[W_GD] * sgt_kr * ij[W_GF] / ij[W_GV];
                Qg = PIg * (wf[0] - p.Pg);
                PIo = PIg * ij[W_OF] / ij[W_GF];
                Qo = PIo * (wf[0] - p.Pg);
                if (dval(Qo) < 0 || nonfinite_or(Qo)) { Qo = T(0.0); PIo = T(0.0); }
                if (dval(Qg) < 0 || nonfinite_or(Qg)) { Qg = T(0.0); PIg = T(0.0); }
                W.inj = i;
                if (cond) {
                    const double s = sso1_geo();
                    if (s != 0) Qg = T(S1r * wf[1] / s);
                    if (s != 0) Qo = Qg * ij[W_OF] / ij[W_GF];
                }
                break;
            case 4: {   // constant oil rate (:402-461)
                const T s = sso1(0);
                if (dval(s) != 0) Qo = -S1r * wf[0] * roo * kro / (muo * s);
                Qw = Qo * krw * K.wat_ro * WAT_M_RO * muo / (kro * roo * K.wat_mu);
                Qg = Qo * krg * rog * muo / (kro * roo * mug);
                if (cond) {
                    set_pi(S1pi);
                    pwf_rates(wf[0]);
                    clamp_prod();
                }
                if (k == 1 && dval(p.Po) < WELL_MIN_P) {
                    Qw = Qo = Qg = T(0.0);
                    PIw = PIo = PIg = T(0.0);
                }
                break;
            }
            case 5: {   // constant gas rate (:463-515)
                const T s = sso1(1);
                if (dval(s) != 0) Qg = -S1r * wf[0] * rog * krg / (mug * s);
                Qw = Qg * krw * K.wat_ro * WAT_M_RO * mug / (krg * rog * K.wat_mu);
                Qo = Qg * kro * roo * mug / (krg * rog * muo);
                if (cond) {
                    set_pi(S1pi);
                    pwf_rates(wf[0]);
                    clamp_prod();
                }
                break;
            }
            case 6: {   // constant total production rate (:517-576)
                T SSw = T(0.0), SSo = T(0.0), SSg = T(0.0), s = T(0.0);
                for (int j = wi[3]; j <= wi[4]; j++) {
                    {
                        const T kw = ownf(j, 11, p.kr[0]), ko = ownf(j, 12, p.kr[1]), kg = ownf(j, 13, p.kr[2]);
                        const T mo = ownf(j, 7, p.mu[0]), mg = ownf(j, 8, p.mu[1]);
                        const T S1 = (kw / K.wat_mu) + (ko / mo) + (kg / mg);
                        SSw += (kw / K.wat_mu) / S1;
                        SSo += (ko / mo) / S1;
                        SSg += (kg / mg) / S1;
                        s += ownf(j, 5, p.ro[0]) * S2own(j) * ko / p.mu[0];
                    }
                    {
                        const double kw = othf(j, 11), ko = othf(j, 12), kg = othf(j, 13);
                        const double mo = othf(j, 7), mg = othf(j, 8);
                        const double S1 = (kw / K.wat_mu) + (ko / mo) + (kg / mg);
                        SSw += (kw / K.wat_mu) / S1;
                        SSo += (ko / mo) / S1;
                        SSg += (kg / mg) / S1;
                        s += othf(j, 5) * S2oth(j) * ko / tabx[7 * s2 + c];
                    }
                }
                const T QoP = SSo * wf[0] / (SSw + SSo + SSg);
                if (dval(s) != 0) Qo = -S1r * QoP * roo * kro / (muo * s);
                Qw = Qo * krw * K.wat_ro * WAT_M_RO * muo / (kro * roo * K.wat_mu);
                Qg = Qo * krg * rog * muo / (kro * roo * mug);
                if (cond) {
                    set_pi(S1pi);
                    pwf_rates(wf[0]);
                    clamp_prod();
                }
                break;
            }
            case 7:   // constant water injection rate (:578-604): its rate branch
                      // tests a sum that is never set, so only the Pwf branch injects
                if (cond) {
                    PIw = S1pi * K.wat_ro * WAT_M_RO * krw / K.wat_mu;
                    Qw = PIw * (wf[0] - p.Pw);
                    if (dval(Qw) < 0 || nonfinite_or(Qw)) { Qw = T(0.0); PIw = T(0.0); }
                }
                break;
            case 8: {   // constant gas injection rate (:606-691)
                const double s = sso1_geo();
                if (s != 0) Qg = T(S1r * wf[0] / s);
                if (s != 0) Qo = Qg * ij[W_OF] / ij[W_GF];
                if (cond) {
                    PIg = S1pi * ij[W_GD] * sgt_kr * ij[W_GF] / ij[W_GV];
                    Qg = PIg * (wf[1] - p.Pg);
                    PIo = PIg * ij[W_OF] / ij[W_GF];
                    Qo = PIo * (wf[1] - p.Pg);
                    if (dval(Qo) < 0 || nonfinite_or(Qo)) { Qo = T(0.0); PIo = T(0.0); }
                    if (dval(Qg) < 0 || nonfinite_or(Qg)) { Qg = T(0.0); PIg = T(0.0); }
                }
                if (dval(p.Po) > WELL_MAX_P) {
                    Qw = Qo = Qg = T(0.0);
                    PIw = PIo = PIg = T(0.0);
                }
                W.inj = i;
                break;
            }
            default:
                break;
        }
    }
}

// Rates and productivity indices of the wells completed in block (k, c).
template <int NC, class T>
__device__ __forceinline__ WellAcc<T> well_acc(const AsmArgs& A, int k, int c, const Geo& G, const PropsT<NC, T>& p) {
    WellAcc<T> W{T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), -1};
    for (int i = 0; i < A.nw; i++) {
        const int* wi = A.wi + 5 * i;
        if (G.ix == wi[1] && G.iy == wi[2] && G.iz >= wi[3] && G.iz <= wi[4]) well_one<NC, T>(A, i, k, c, G, p, W);
    }
    return W;
}

// Rates of the wells completed in block (k, c); returns the injector whose
// stream the component rows take, or -1.
template <int NC, class T>
__device__ __forceinline__ int well_rates(const AsmArgs& A, int k, int c, const Geo& G, const PropsT<NC, T>& p,
                                          T& Qw, T& Qo, T& Qg) {
    const WellAcc<T> W = well_acc<NC, T>(A, k, c, G, p);
    Qw = W.Qw;
    Qo = W.Qo;
    Qg = W.Qg;
    return W.inj;
}

// The well terms of block (k, c)'s balance rows.
template <int NC, class T>
__device__ __forceinline__ void well_piece(const AsmArgs& A, int k, int c, const Geo& G, const VarsT<NC, T>& v,
                                           const PropsT<NC, T>& p, RowAcc<NC, T>& R) {
    if (A.nw == 0) return;
    T Qw, Qo, Qg;
    const int inj = well_rates<NC, T>(A, k, c, G, p, Qw, Qo, Qg);
    const double* ic = inj >= 0 ? A.winj + (int64_t)(2 * NC + 6) * inj : nullptr;
#pragma unroll
    for (int i = 0; i < NC - 1; i++) {
        const T oc = ic ? T(ic[i]) : v.x[i];
        const T gc = ic ? T(ic[NC + i]) : v.y[i];
        R.bal[i] -= oc * Qo;
        R.bal[i] -= gc * Qg;
    }
    R.bal[NC - 1] -= Qo;
    R.bal[NC - 1] -= Qg;
    R.bal[NC] -= Qw;
}

// FlashInjection (Init_Calcs.cpp:1161-1377) of every injector's stream, one
// lane per well: successive substitution from Wilson K-values with the
// Rachford-Rice Newton restarted at L = 0.6 every pass, until L moves by at
// most FLASH_TOL; a K-value turned negative makes the vapour composition the
// previous pass's y_i / (1 - L) (the reference divides its output array,
// :1219).  out[w]: x (nc), y (nc), oil and gas fraction, density, viscosity.
template <int NC>
__global__ void k_flash_injection(const AsmConsts* Kp, int nw, const int* wi, const double* wf, double* out) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const int type = wi[5 * w];
    if (type != 3 && type != 8) return;
    const AsmConsts& K = *Kp;
    const double* f = wf + (int64_t)(NC + 6) * w;
    const double* z = f + 4;
    const double P = type == 3 ? f[0] : f[1];
    const double RT = RGAS * K.T;
    const bool srk = K.eos == OSR_EOS_SRK;
    double Kv[NC], x[NC], y[NC];
    for (int i = 0; i < NC; i++) {
        Kv[i] = (K.pc[i] / P) * exp(K.wk[i] * (1 - K.tc[i] / K.T));
        x[i] = 0;
        y[i] = 0;
    }
    double L = 0.6, Lb = 10, ESum, Zl = 0, Zg = 0;
    int it = 0;
    do {
        L = 0.6;
        int j = 0;
        double tD;
        do {
            double Fv = 0, dF = 0;
            for (int i = 0; i < NC; i++) {
                if (Kv[i] >= 0) tD = (1 - Kv[i]) / (L + (1 - L) * Kv[i]);
                else tD = 1 / (L - 1);
                Fv += z[i] * tD;
                dF -= z[i] * tD * tD;
            }
            tD = Fv / dF;
            L -= tD;
            j++;
        } while ((tD * tD) > 1e-8 && j < 100);   // NEWTON_TOL, MAX_FLASH_N_ITR
        if (L > 1) L = 1;
        else if (L < 0) L = 0;
        ESum = fabs(Lb - L);
        Lb = L;
        for (int i = 0; i < NC; i++) {
            if (Kv[i] >= 0) {
                x[i] = z[i] / (L + Kv[i] * (1 - L));
                y[i] = Kv[i] * x[i];
            } else {
                x[i] = 0;
                y[i] = y[i] / (1 - L);
            }
        }
        double Al = 0, Bl = 0, Ag = 0, Bg = 0;
        for (int i = 0; i < NC; i++) {
            Bl += x[i] * K.b[i];
            Bg += y[i] * K.b[i];
            for (int jj = 0; jj < NC; jj++) {
                const double sq = sqrt(K.a[i] * K.a[jj]) * K.bic[i * NC + jj];
                Al += x[i] * x[jj] * sq;
                Ag += y[i] * y[jj] * sq;
            }
        }
        const double Cal = Al, Cbl = Bl, Cag = Ag, Cbg = Bg;
        Al *= P / (RGAS * RGAS * K.T * K.T);
        Bl *= P / RT;
        Ag *= P / (RGAS * RGAS * K.T * K.T);
        Bg *= P / RT;
        const double Ul = srk ? Bl : 2 * Bl, Wl = srk ? 0 : Bl;
        const double Ug = srk ? Bg : 2 * Bg, Wg = srk ? 0 : Bg;
        Zl = zroot(-(1 + Bl - Ul), Al - Bl * Ul - Ul - Wl * Wl, -(Al * Bl - Bl * Wl * Wl - Wl * Wl), false);
        Zg = zroot(-(1 + Bg - Ug), Ag - Bg * Ug - Ug - Wg * Wg, -(Ag * Bg - Bg * Wg * Wg - Wg * Wg), true);
        for (int i = 0; i < NC; i++) {
            double tl = 0, tg = 0;
            for (int jj = 0; jj < NC; jj++) {
                const double sq = sqrt(K.a[i] * K.a[jj]) * K.bic[i * NC + jj];
                tl += x[jj] * sq;
                tg += y[jj] * sq;
            }
            const double t1l = K.b[i] / Cbl, t1g = K.b[i] / Cbg;
            const double EEl = 2 * tl / Cal - t1l, EEg = 2 * tg / Cag - t1g;
            const double DDl = Al / (Bl * (K.d1 - K.d2)), DDg = Ag / (Bg * (K.d1 - K.d2));
            const double FFl = log((Zl + K.d2 * Bl) / (Zl + K.d1 * Bl));
            const double FFg = log((Zg + K.d2 * Bg) / (Zg + K.d1 * Bg));
            const double phiL = exp(t1l * (Zl - 1) - log(Zl - Bl) + DDl * EEl * FFl);
            const double phiG = exp(t1g * (Zg - 1) - log(Zg - Bg) + DDg * EEg * FFg);
            const double fL = x[i] * phiL, fG = y[i] * phiG;
            if (z[i] != 0) {
                if (fG != 0) Kv[i] *= fL / fG;
                else Kv[i] = -1;
                if (fL != fL) Kv[i] = -1;
                if (fG != fG) Kv[i] = 1;
            }
        }
        it++;
    } while (ESum > 1e-14 && it < 500);   // FLASH_TOL
    double* o = out + (int64_t)(2 * NC + 6) * w;
    double OF, GF;
    if (fabs(L) < 1e-20) { OF = 0; GF = 1; }
    else if (fabs(1 - L) < 1e-20) { OF = 1; GF = 0; }
    else { OF = L; GF = 1 - L; }
    const double OD = P / (Zl * RT), GD = P / (Zg * RT);
    for (int ph = 0; ph < 2; ph++) {
        const double* wv = ph == 0 ? x : y;
        const double ro = ph == 0 ? OD : GD;
        double s1 = 0, s2 = 0, rv = 0, tcm = 0, pcm = 0, mwm = 0;
        for (int i = 0; i < NC; i++) {
            const double tm = wv[i] * K.sqmw[i];
            s1 += tm * K.mu0[i];
            s2 += tm;
            rv += wv[i] * K.vcm3[i];
            tcm += wv[i] * K.tc[i];
            pcm += wv[i] * (K.pc[i] / 101325);
            mwm += wv[i] * K.mw[i];
        }
        const double Er = ro * rv;
        const double eta = pow(tcm, 1.0 / 6) / (sqrt(mwm) * pow(pcm, 2.0 / 3));
        const double r1 = 0.10230 + 0.023364 * Er + 0.058533 * Er * Er - 0.040758 * Er * Er * Er +
                          0.0093324 * Er * Er * Er * Er;
        o[2 * NC + (ph == 0 ? W_OV : W_GV)] = s1 / s2 + (r1 * r1 * r1 * r1 - 1e-4) / eta;
    }
    for (int i = 0; i < NC; i++) {
        o[i] = x[i];
        o[NC + i] = y[i];
    }
    o[2 * NC + W_OF] = OF;
    o[2 * NC + W_GF] = GF;
    o[2 * NC + W_OD] = OD;
    o[2 * NC + W_GD] = GD;
}


// ------------------------------------------------------------ well kernels
// The well terms touch only the balance rows of a completed block and depend
// on that block's own unknowns, so they are added by small kernels over the
// list of completed blocks (wblk), launched after the kernels that write
// those rows: the stencil kernels carry no well code.

// residual: F rows += well terms
template <int NC>
__global__ void k_well_res(AsmArgs A, const int* wblk, int nwb, double* F) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwb) return;
    const int blk = wblk[t], n = A.n, k = blk / n, c = blk % n;
    if (A.phi[k * n + c] < 1e-10) return;
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    PropsT<NC, double> p;
    load_props<NC>(A, k, c, p);
    const Geo G = geom(A, c);
    RowAcc<NC, double> R;
    well_piece<NC, double>(A, k, c, G, v, p, R);
    double* out = F + A.boff[blk] + (v.lab == 0 ? NC : 0);
#pragma unroll
    for (int i = 0; i <= NC; i++) out[i] += R.bal[i];
}

// exact Jacobian (CSR): the own-block columns of the balance rows, one lane
// per (completed block, own unknown j)
template <int NC>
__global__ void k_well_jac(AsmArgs A, const int* wblk, int nwb, const double* dp, const int64_t* rowptr,
                           double* vals) {
    using D1 = DualV<1>;
    constexpr int MAXC = JacShape<NC>::MAXC;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (int64_t)nwb * MAXC) return;
    const int j = (int)(item % MAXC);
    const int blk = wblk[item / MAXC], n = A.n, k = blk / n, c = blk % n;
    const int lab = A.lab[k][c];
    if (j >= (lab == 0 ? 2 * NC + 4 : NC + 3)) return;
    if (A.phi[k * n + c] < 1e-10) return;
    const Geo G = geom(A, c);
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    VarsT<NC, D1> sv;
    seed_vars<NC, 1>(v, j, A.pcoeff, sv);
    PropsT<NC, D1> sp;
    load_props_dual<NC>(A, dp, blk, j, sp);
    RowAcc<NC, D1> R;
    well_piece<NC, D1>(A, k, c, G, sv, sp, R);
    const int64_t rb = A.boff[blk] + (lab == 0 ? NC : 0);
    const int64_t s_mass = seg_start(A, k, G, c, 0, blk, NC);
#pragma unroll
    for (int i = 0; i < NC; i++) vals[rowptr[rb + i] + s_mass + j] += R.bal[i].d[0];
    vals[rowptr[rb + NC] + seg_start(A, k, G, c, 1, blk, NC) + j] += R.bal[NC].d[0];
}

// reference-mode Jacobian: the reference's well entries (IteratorJacobian,
// matrix :32543, :32808-32813, :35425, :36029; fracture :15033, :17868,
// :18474) added into the block's own columns of its balance and water rows:
// PCoeff (x_i PIo + y_i PIg) in the P column of component row i (y_i the
// injection stream's in an injector's block; x_i the block's own in every
// case, as there), PCoeff (PIo + PIg) in the total row's, PCoeff PIw in the
// water row's, and for a producer -Qo, -Qg in the x_i, y_i columns of
// component row i.  One lane per completed block; the columns are found in
// the row's sorted column list.
__device__ __forceinline__ void csr_add(const int64_t* rowptr, const int32_t* cols, double* vals, int64_t row,
                                        int64_t col, double v) {
    int64_t lo = rowptr[row], hi = rowptr[row + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cols[mid] < col) lo = mid + 1;
        else hi = mid;
    }
    if (lo < rowptr[row + 1] && cols[lo] == col) vals[lo] += v;
}

template <int NC>
__global__ void k_well_refjac(AsmArgs A, const int* wblk, int nwb, const int64_t* rowptr, const int32_t* cols,
                              double* vals) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwb) return;
    const int blk = wblk[t], n = A.n, k = blk / n, c = blk % n;
    if (A.phi[k * n + c] < 1e-10) return;
    const int lab = A.lab[k][c];
    PropsT<NC, double> p;
    load_props<NC>(A, k, c, p);
    const Geo G = geom(A, c);
    const WellAcc<double> W = well_acc<NC, double>(A, k, c, G, p);
    const int64_t b0 = A.boff[blk];
    const int r0 = lab == 0 ? NC : 0, reCol = lab == 0 ? 0 : NC;
    const int64_t pcol = b0 + 2 * NC - reCol;
    const double* ic = W.inj >= 0 ? A.winj + (int64_t)(2 * NC + 6) * W.inj : nullptr;
    for (int i = 0; i < NC - 1; i++) {
        const double xi = A.x[k][(int64_t)i * n + c];
        const double gc = ic ? ic[NC + i] : A.y[k][(int64_t)i * n + c];
        const int64_t row = b0 + r0 + i;
        csr_add(rowptr, cols, vals, row, pcol, A.pcoeff * (xi * W.PIo + gc * W.PIg));
        if (!ic) {
            if (lab != -1) csr_add(rowptr, cols, vals, row, b0 + i, -W.Qo);
            if (lab != 1) csr_add(rowptr, cols, vals, row, b0 + NC + i - reCol, -W.Qg);
        }
    }
    csr_add(rowptr, cols, vals, b0 + r0 + NC - 1, pcol, A.pcoeff * (W.PIo + W.PIg));
    csr_add(rowptr, cols, vals, b0 + r0 + NC, pcol, A.pcoeff * W.PIw);
}

// reduced system (direct): the own column block of the completed block's
// reduced rows (directions q < nc + 1) and its right-hand-side partial
// (direction nc + 1), one lane per (completed block, direction)
template <int NC>
__global__ void k_well_red(AsmArgs A, const int* wblk, int nwb, const double* dred, const int64_t* browptr,
                           double* bvals, double* rpart) {
    using D1 = DualV<1>;
    constexpr int NP = NC + 1;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (int64_t)nwb * (NP + 1)) return;
    const int q = (int)(item % (NP + 1));
    const int blk = wblk[item / (NP + 1)], n = A.n, k = blk / n, c = blk % n;
    if (A.phi[k * n + c] < 1e-10) return;
    const Geo G = geom(A, c);
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    VarsT<NC, D1> sv;
    seed_tab<NC>(v, dred + (int64_t)blk * JacShape<NC>::RF * (NP + 1) + q, NP + 1, A.pcoeff, sv);
    PropsT<NC, D1> sp;
    load_props_dual<NC>(A, dred, blk, q, sp, NP + 1, JacShape<NC>::RF);
    RowAcc<NC, D1> R;
    well_piece<NC, D1>(A, k, c, G, sv, sp, R);
    if (q == NP) {
#pragma unroll
        for (int i = 0; i < NP; i++) rpart[(int64_t)i * 2 * n + blk] += R.bal[i].d[0];
        return;
    }
    const int64_t bz = browptr[blk] + red_rank(A, k, G.ix, G.iy, G.iz, 0);
#pragma unroll
    for (int i = 0; i < NP; i++) bvals[bz * NP * NP + i * NP + q] += R.bal[i].d[0];
}

}  // namespace osr

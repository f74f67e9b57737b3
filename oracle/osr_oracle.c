/*
 * osr_oracle.c -- CPU FP64 restatement of the Osiris thermodynamic kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see osr_oracle.h): the HIP product never links,
 * loads or calls this file.  It is the checker for parity tests and the
 * timed CPU baseline.
 *
 * Citations are file:line in the reference repository (compsim/...).
 */
#include "osr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NCMAX 32

/* ---------------------------------------------------------------- EOS_Init */
/* Init_Calcs.cpp:66-97.  PR uses the PR78 m(omega) cubic for every omega,
 * SRK the original quadratic.  a_i = ac_i (1 + m_i (1 - sqrt(T/Tc_i)))^2. */
void or_eos_ab(const or_fluid *f, double T, double *a, double *b)
{
    for (int i = 0; i < f->nc; i++) {
        const double Tc = f->Tc[i], Pc = f->Pc[i], w = f->w[i];
        const double tr = T / Tc;                        /* Init_Calcs.cpp:75 */
        double ac, m, bc;
        if (f->eos == OR_SRK) {                          /* :80-86 */
            ac = 0.42747 * OR_RGAS * OR_RGAS * Tc * Tc / Pc;
            m = 0.48508 + 1.55171 * w - 0.15613 * w * w;
            bc = 0.08664 * OR_RGAS * Tc / Pc;
        } else {                                         /* :88-95 */
            ac = 0.457235 * OR_RGAS * OR_RGAS * Tc * Tc / Pc;
            m = 0.3796 + 1.485 * w - 0.1644 * w * w + 0.01667 * w * w * w;
            bc = 0.077796 * OR_RGAS * Tc / Pc;
        }
        const double s = 1 + m * (1 - sqrt(tr));
        a[i] = ac * s * s;
        b[i] = bc;
    }
}

/* ----------------------------------------------------------------- Solve_Z */
/* Globals.cpp:212-264.  Three real roots (D<0): trigonometric form, vapour =
 * largest, liquid = smallest positive else the third root.  D==0: double
 * root form.  D>0: Cardano's single real root (both states). */
double or_solve_z(double a1, double a2, double a3, int gas)
{
    const double Q = (3 * a2 - a1 * a1) / 9;
    const double J = (9 * a1 * a2 - 27 * a3 - 2 * a1 * a1 * a1) / 54;
    const double D = Q * Q * Q + J * J;
    double Z;
    if (D < 0) {
        const double r = 2 * sqrt(-Q);
        const double s = a1 / 3;
        const double t = acos(J / sqrt(-Q * Q * Q)) / 3;
        const double x1 = r * cos(t) - s;
        const double x2 = r * cos(t + 2 * OR_PI / 3) - s;
        const double x3 = r * cos(t + 4 * OR_PI / 3) - s;
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
        const double t = J > 0 ? pow(J, 1.0 / 3) : -pow(-J, 1.0 / 3);
        const double x1 = 2 * t - a1 / 3;
        const double x2 = -t - a1 / 3;
        if (gas) Z = (x2 >= x1) ? x2 : x1;
        else Z = (x2 <= x1 && x2 > 0) ? x2 : x1;
    } else {
        const double sd = sqrt(D);
        double t = J + sd;
        Z = t > 0 ? pow(t, 1.0 / 3) : -pow(-t, 1.0 / 3);
        t = J - sd;
        if (t > 0) Z += pow(t, 1.0 / 3);
        else Z -= pow(-t, 1.0 / 3);
        Z -= a1 / 3;
    }
    return Z;
}

/* Cubic coefficients of the generic two-parameter EOS as the reference writes
 * them (e.g. Init_Calcs.cpp:675-692): U=2B, W=B for PR; U=B, W=0 for SRK. */
static double phase_z(int eos, double A, double B, int gas)
{
    const double U = (eos == OR_SRK) ? B : 2 * B;
    const double W = (eos == OR_SRK) ? 0 : B;
    return or_solve_z(-(1 + B - U), A - B * U - U - W * W, -(A * B - B * W * W - W * W), gas);
}

static void eos_d12(int eos, double *d1, double *d2)
{
    if (eos == OR_SRK) { *d1 = 1; *d2 = 0; }
    else { *d1 = 1 + OR_SQRT2; *d2 = 1 - OR_SQRT2; }
}

/* Mixing sums (Init_Calcs.cpp:650-660): Ca = sum x_i x_j sqrt(a_i a_j) bic_ij,
 * Cb = sum x_i b_i, and t_i = sum_j x_j sqrt(a_i a_j) bic_ij (:697-702). */
static void mix(const or_fluid *f, const double *a, const double *b,
                const double *x, double *Ca, double *Cb, double *t)
{
    const int nc = f->nc;
    double ca = 0, cb = 0;
    for (int i = 0; i < nc; i++) {
        cb += x[i] * b[i];
        double ti = 0;
        for (int j = 0; j < nc; j++) {
            const double aij = sqrt(a[i] * a[j]) * f->bic[i * nc + j];
            ca += x[i] * x[j] * aij;
            ti += x[j] * aij;
        }
        t[i] = ti;
    }
    *Ca = ca;
    *Cb = cb;
}

/* ln(phi_i) at given Z (Init_Calcs.cpp:694-712): the exponent of phiL/phiG. */
static double lnphi_i(double bi, double ti, double Ca, double Cb, double A,
                      double B, double Z, double d1, double d2)
{
    const double t1 = bi / Cb;
    const double EE = 2 * ti / Ca - t1;
    const double DD = A / (B * (d1 - d2));
    const double FF = log((Z + d2 * B) / (Z + d1 * B));
    return t1 * (Z - 1) - log(Z - B) + DD * EE * FF;
}

void or_phase(const or_fluid *f, const double *a, const double *b, double T,
              double P, const double *x, int gas, double *Z, double *lnphi)
{
    double t[NCMAX], Ca, Cb, d1, d2;
    mix(f, a, b, x, &Ca, &Cb, t);
    const double A = Ca * P / (OR_RGAS * OR_RGAS * T * T);
    const double B = Cb * P / (OR_RGAS * T);
    eos_d12(f->eos, &d1, &d2);
    const double z = phase_z(f->eos, A, B, gas);
    *Z = z;
    if (lnphi)
        for (int i = 0; i < f->nc; i++)
            lnphi[i] = lnphi_i(b[i], t[i], Ca, Cb, A, B, z, d1, d2);
}

/* Wilson K-value (Init_Calcs.cpp:586-588, Minor_NR.cpp:530-532). */
static double wilson(const or_fluid *f, int i, double T, double P)
{
    return (f->Pc[i] / P) * exp(5.37 * (1 + f->w[i]) * (1 - f->Tc[i] / T));
}

/* --------------------------------------------------------------- SuccFalsh */
void or_flash(const or_fluid *f, double T, double P, const double *z, double sw,
              double *x, double *y, double *K, or_flash_out *o)
{
    const int nc = f->nc;
    double a[NCMAX], b[NCMAX], tl[NCMAX], tg[NCMAX];
    double d1, d2, L = 0, Lb = 10, ESum, Zl = 0, Zg = 0;
    int k = 0, rr = 0;
    or_eos_ab(f, T, a, b);
    eos_d12(f->eos, &d1, &d2);
    for (int i = 0; i < nc; i++) K[i] = wilson(f, i, T, P);

    do {
        /* Rachford-Rice by Newton from L=0.6 every outer pass (:592-606). */
        L = 0.6;
        int j = 0;
        double step;
        do {
            double F = 0, dF = 0;
            for (int i = 0; i < nc; i++) {
                const double td = (K[i] >= 0) ? (1 - K[i]) / (L + (1 - L) * K[i]) : 1 / (L - 1);
                F += z[i] * td;
                dF -= z[i] * td * td;
            }
            step = F / dF;
            L -= step;
            j++;
        } while (step * step > OR_NEWTON_TOL && j < OR_MAX_FLASH_N_ITR);
        rr += j;

        if (L > 1) L = 1;                                /* :608-615 */
        else if (L < 0) L = 0;
        ESum = fabs(Lb - L);
        Lb = L;

        for (int i = 0; i < nc; i++) {                   /* :620-629 */
            if (K[i] >= 0) {
                x[i] = z[i] / (L + K[i] * (1 - L));
                y[i] = K[i] * x[i];
            } else {
                x[i] = 0;
                y[i] = z[i] / (1 - L);
            }
        }

        /* Both phases evaluated at the oil pressure (:660-664). */
        double Cal, Cbl, Cag, Cbg;
        mix(f, a, b, x, &Cal, &Cbl, tl);
        mix(f, a, b, y, &Cag, &Cbg, tg);
        const double Al = Cal * P / (OR_RGAS * OR_RGAS * T * T);
        const double Bl = Cbl * P / (OR_RGAS * T);
        const double Ag = Cag * P / (OR_RGAS * OR_RGAS * T * T);
        const double Bg = Cbg * P / (OR_RGAS * T);
        Zl = phase_z(f->eos, Al, Bl, 0);
        Zg = phase_z(f->eos, Ag, Bg, 1);

        for (int i = 0; i < nc; i++) {                   /* :694-711 */
            const double phiL = exp(lnphi_i(b[i], tl[i], Cal, Cbl, Al, Bl, Zl, d1, d2));
            const double phiG = exp(lnphi_i(b[i], tg[i], Cag, Cbg, Ag, Bg, Zg, d1, d2));
            const double fL = x[i] * phiL;
            const double fG = y[i] * phiG;
            if (z[i] != 0) {
                if (fG != 0) K[i] *= fL / fG;
                else K[i] = -1;
                if (fL != fL) K[i] = -1;
                if (fG != fG) K[i] = 1;
            }
        }
        k++;
    } while (ESum > OR_FLASH_TOL && k < OR_MAX_SS_ITER);

    /* Label from the clamped L (:717-727), saturation split and densities
     * at the oil pressure (:730-736). */
    if (fabs(L) < OR_DZERO) o->label = -1;
    else if (fabs(1 - L) < OR_DZERO) o->label = 1;
    else o->label = 0;
    const double newD = Zl * L / (Zl * L + Zg * (1 - L));
    o->so = newD * (1 - sw);
    o->sg = (1 - newD) * (1 - sw);
    o->rho_l = P / (Zl * OR_RGAS * T);
    o->rho_g = P / (Zg * OR_RGAS * T);
    o->L = L;
    o->Zl = Zl;
    o->Zg = Zg;
    o->iters = k;
    o->rr_iters = rr;
}

void or_flash_batch(const or_fluid *f, int64_t n, const double *T, double T0,
                    const double *P, const double *z, double sw,
                    double *x, double *y, double *K, double *L, int8_t *label,
                    double *Zl, double *Zg, int32_t *iters, int nthreads)
{
    or_flash_batch_rr(f, n, T, T0, P, z, sw, x, y, K, L, label, Zl, Zg, iters, NULL, nthreads);
}

void or_flash_batch_rr(const or_fluid *f, int64_t n, const double *T, double T0,
                       const double *P, const double *z, double sw,
                       double *x, double *y, double *K, double *L, int8_t *label,
                       double *Zl, double *Zg, int32_t *iters, int32_t *rr, int nthreads)
{
    const int nc = f->nc;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
#endif
    for (int64_t c = 0; c < n; c++) {
        double zc[NCMAX], xc[NCMAX], yc[NCMAX], kc[NCMAX];
        or_flash_out o;
        for (int i = 0; i < nc; i++) zc[i] = z[(int64_t)i * n + c];
        or_flash(f, T ? T[c] : T0, P[c], zc, sw, xc, yc, kc, &o);
        for (int i = 0; i < nc; i++) {
            if (x) x[(int64_t)i * n + c] = xc[i];
            if (y) y[(int64_t)i * n + c] = yc[i];
            if (K) K[(int64_t)i * n + c] = kc[i];
        }
        if (L) L[c] = o.L;
        if (label) label[c] = (int8_t)o.label;
        if (Zl) Zl[c] = o.Zl;
        if (Zg) Zg[c] = o.Zg;
        if (iters) iters[c] = o.iters;
        if (rr) rr[c] = o.rr_iters;
    }
    (void)nthreads;
}

/* ---------------------------------------------------------------- LU solve */
int or_lu_solve(int n, double *A, const double *rhs, double *sol)
{
    int piv[NCMAX + 2];
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int r = k + 1; r < n; r++)          /* first maximal |a| (idamax) */
            if (fabs(A[r * n + k]) > best) { best = fabs(A[r * n + k]); p = r; }
        piv[k] = p;
        if (A[p * n + k] != 0) {
            if (p != k)
                for (int c = 0; c < n; c++) {
                    double t = A[k * n + c]; A[k * n + c] = A[p * n + c]; A[p * n + c] = t;
                }
            const double inv = 1.0 / A[k * n + k];
            for (int r = k + 1; r < n; r++) A[r * n + k] *= inv;
        } else if (!info) {
            info = k + 1;
        }
        for (int r = k + 1; r < n; r++) {
            const double l = A[r * n + k];
            if (l != 0)
                for (int c = k + 1; c < n; c++) A[r * n + c] -= l * A[k * n + c];
        }
    }
    for (int i = 0; i < n; i++) sol[i] = rhs[i];
    for (int k = 0; k < n; k++)
        if (piv[k] != k) { double t = sol[k]; sol[k] = sol[piv[k]]; sol[piv[k]] = t; }
    for (int i = 0; i < n; i++)
        for (int c = 0; c < i; c++) sol[i] -= A[i * n + c] * sol[c];
    for (int i = n - 1; i >= 0; i--) {
        for (int c = i + 1; c < n; c++) sol[i] -= A[i * n + c] * sol[c];
        sol[i] /= A[i * n + i];
    }
    return info;
}

/* ------------------------------------------------- saturation-pressure core */
/* Derivatives of ln(phi_i) for one phase with respect to its own (unnormalised)
 * mole numbers x_k and to P, the exact partials the reference assembles by
 * hand in Bubl_Succ/Dew_Succ (Minor_NR.cpp:731-817, 1111-1197). */
typedef struct {
    double Ca, Cb, A, B, Z, U, W, dZdp;
    double t[NCMAX], phi[NCMAX];
} phase_state;

static void phase_eval(const or_fluid *f, const double *a, const double *b,
                       double T, double P, const double *x, int gas, phase_state *s)
{
    double d1, d2;
    eos_d12(f->eos, &d1, &d2);
    mix(f, a, b, x, &s->Ca, &s->Cb, s->t);
    s->A = s->Ca * P / (OR_RGAS * OR_RGAS * T * T);
    s->B = s->Cb * P / (OR_RGAS * T);
    s->U = (f->eos == OR_SRK) ? s->B : 2 * s->B;
    s->W = (f->eos == OR_SRK) ? 0 : s->B;
    s->Z = phase_z(f->eos, s->A, s->B, gas);
    for (int i = 0; i < f->nc; i++)
        s->phi[i] = exp(lnphi_i(b[i], s->t[i], s->Ca, s->Cb, s->A, s->B, s->Z, d1, d2));
}

/* dZ from the implicit cubic (Minor_NR.cpp:731-732): the partials of
 * Z^3-(1+B-U)Z^2+(A-BU-U-W^2)Z-(AB-BW^2-W^2) in A, B, U, W and Z. */
static double dz_implicit(const phase_state *s, double dA, double dB, double dU, double dW)
{
    const double Z = s->Z, A = s->A, B = s->B, U = s->U, W = s->W;
    const double num = dA * (Z - B) + dB * (-Z * Z - U * Z - A + W * W) +
                       dU * (Z * Z - B * Z - Z) + dW * (-2 * W * Z + 2 * B * W + 2 * W);
    const double den = 3 * Z * Z - 2 * Z * (1 + B - U) + A - B * U - U - W * W;
    return -num / den;
}

/* d(phi_i)/dP (Minor_NR.cpp:760-772). */
static double dphi_dp(const or_fluid *f, const double *b, double T,
                      const phase_state *s, int i)
{
    double d1, d2;
    eos_d12(f->eos, &d1, &d2);
    const double dA = s->Ca / (OR_RGAS * OR_RGAS * T * T);
    const double dB = s->Cb / (OR_RGAS * T);
    const double dU = (f->eos == OR_SRK) ? dB : 2 * dB;
    const double dW = (f->eos == OR_SRK) ? 0 : dB;
    const double dZ = dz_implicit(s, dA, dB, dU, dW);
    const double A = s->A, B = s->B, Z = s->Z;
    const double t1 = b[i] / s->Cb;
    const double EE = 2 * s->t[i] / s->Ca - t1;
    const double DD = A / (B * (d1 - d2));
    const double FF = log((Z + d2 * B) / (Z + d1 * B));
    const double dAA = b[i] * dZ / s->Cb;
    const double dBB = (dB - dZ) / (Z - B);
    const double dDD = (B * dA - A * dB) / (B * B * (d1 - d2));
    const double dFF = (dZ + d2 * dB) / (Z + d2 * B) - (dZ + d1 * dB) / (Z + d1 * B);
    return s->phi[i] * (dAA + dBB + EE * (dDD * FF + DD * dFF));
}

/* d(phi_i)/dx_k (Minor_NR.cpp:778-808). */
static double dphi_dx(const or_fluid *f, const double *a, const double *b,
                      double T, double P, const double *x, const phase_state *s,
                      int i, int k)
{
    const int nc = f->nc;
    double d1, d2;
    eos_d12(f->eos, &d1, &d2);
    double sk = 0;
    for (int n = 0; n < nc; n++) sk += x[n] * sqrt(a[n] * a[k]) * f->bic[n * nc + k];
    const double dA = sk * 2 * P / (OR_RGAS * OR_RGAS * T * T);
    const double dB = b[k] * P / (OR_RGAS * T);
    const double dU = (f->eos == OR_SRK) ? dB : 2 * dB;
    const double dW = (f->eos == OR_SRK) ? 0 : dB;
    const double dZ = dz_implicit(s, dA, dB, dU, dW);
    const double A = s->A, B = s->B, Z = s->Z, Ca = s->Ca, Cb = s->Cb;
    const double t1 = b[i] / Cb;
    const double EE = 2 * s->t[i] / Ca - t1;
    const double DD = A / (B * (d1 - d2));
    const double FF = log((Z + d2 * B) / (Z + d1 * B));
    const double dAA = -b[k] * b[i] * (Z - 1) / (Cb * Cb) + b[i] * dZ / Cb;
    const double dBB = (dB - dZ) / (Z - B);
    const double dDD = (dA * B - dB * A) / (B * B * (d1 - d2));
    const double dFF = (dZ + d2 * dB) / (Z + d2 * B) - (dZ + d1 * dB) / (Z + d1 * B);
    const double dEE = 2 * (f->bic[i * nc + k] * sqrt(a[i] * a[k]) / Ca -
                            dA * OR_RGAS * OR_RGAS * T * T * s->t[i] / (P * Ca * Ca)) +
                       b[i] * b[k] / (Cb * Cb);
    return s->phi[i] * (dAA + dBB + dDD * EE * FF + DD * dEE * FF + DD * EE * dFF);
}

/* Shared body of Bubl_Succ and Dew_Succ.  bubble=1: the fixed phase is the
 * liquid (fixed = x, incipient vapour), pressure P; bubble=0: fixed vapour y,
 * incipient liquid. */
static double psat_core(const or_fluid *f, double T, double P, const double *fixed,
                        const double *ystale, int bubble, double *xm)
{
    const int nc = f->nc;
    double a[NCMAX], b[NCMAX], Kw[NCMAX], K[NCMAX], X[NCMAX + 1], dX[NCMAX + 1];
    double J[(NCMAX + 1) * (NCMAX + 1)], r[NCMAX + 1];
    phase_state sf, si;
    or_eos_ab(f, T, a, b);
    for (int i = 0; i < nc; i++) Kw[i] = wilson(f, i, T, P);

    /* Stage 1: successive substitution (Minor_NR.cpp:529-637 / 908-1021).
     * Both routines seed the incipient phase from comp[..][1]/K. */
    int ret_bad = 0, nIter = 0, k = 0;
    double sumK = 0, SXm;
    X[nc] = P;
    for (int i = 0; i < nc; i++) K[i] = Kw[i];
    do {
        double xs = 0;
        for (int i = 0; i < nc; i++) {
            X[i] = ystale[i] / K[i];
            if (X[i] < 0) X[i] = 0;
            else if (X[i] > 1) X[i] = 1;
            xs += X[i];
        }
        if (X[nc] < 0) X[nc] = OR_PZERO;
        for (int i = 0; i < nc; i++) X[i] /= xs;

        if (bubble) {
            phase_eval(f, a, b, T, X[nc], fixed, 0, &sf);
            phase_eval(f, a, b, T, X[nc], X, 1, &si);
        } else {
            phase_eval(f, a, b, T, X[nc], X, 0, &si);
            phase_eval(f, a, b, T, X[nc], fixed, 1, &sf);
        }
        SXm = 0;
        k = 0;
        sumK = 0;
        for (int i = 0; i < nc; i++) {
            if (bubble) {
                const double fL = fixed[i] * sf.phi[i];
                const double fG = X[i] * si.phi[i];
                if (fixed[i] != 0) {
                    K[i] *= fL / fG;
                    SXm += K[i] * fixed[i];
                    k++;
                    sumK += fabs(1 - K[i]);
                }
            } else {
                const double fL = X[i] * si.phi[i];
                const double fG = fixed[i] * sf.phi[i];
                if (fixed[i] != 0) {
                    K[i] *= fL / fG;
                    SXm += fixed[i] / K[i];
                    sumK += fabs(1 - K[i]);
                    k++;
                }
            }
        }
        X[nc] *= SXm;
        nIter++;
        if (nIter > OR_MAX_MNR_ITER) { ret_bad = 1; break; }
    } while (fabs(SXm - 1) / k > OR_SXMTOL);

    if (sumK / k < OR_TRIV_K) ret_bad = 1;
    if (!ret_bad && (X[nc] - X[nc]) == 0) {
        for (int i = 0; i < nc; i++) xm[i] = X[i];
        return X[nc];
    }

    /* Stage 2: Newton on [x_i phiL_i - Xm_i phiG_i = 0 ; sum Xm = 1] in
     * (Xm, P), Wilson start (Minor_NR.cpp:648-870 / 1027-1249). */
    X[nc] = P;
    for (int i = 0; i < nc; i++) X[i] = bubble ? Kw[i] * fixed[i] : fixed[i] / Kw[i];
    nIter = 0;
    double XmTol;
    const int n1 = nc + 1;
    do {
        double s = 0;
        for (int i = 0; i < nc; i++) {
            s += X[i];
            J[nc * n1 + i] = 1;
        }
        r[nc] = -s + 1;
        J[nc * n1 + nc] = 0;
        if (bubble) {
            phase_eval(f, a, b, T, X[nc], fixed, 0, &sf);
            phase_eval(f, a, b, T, X[nc], X, 1, &si);
        } else {
            phase_eval(f, a, b, T, X[nc], X, 0, &si);
            phase_eval(f, a, b, T, X[nc], fixed, 1, &sf);
        }
        /* Residual g_i = (liquid fugacity) - (vapour fugacity), with the
         * incipient phase's mole numbers as unknowns. */
        for (int i = 0; i < nc; i++) {
            const double dpf = dphi_dp(f, b, T, &sf, i);
            const double dpi = dphi_dp(f, b, T, &si, i);
            if (bubble) {
                r[i] = -fixed[i] * sf.phi[i] + X[i] * si.phi[i];
                J[i * n1 + nc] = fixed[i] * dpf - X[i] * dpi;
            } else {
                r[i] = -X[i] * si.phi[i] + fixed[i] * sf.phi[i];
                J[i * n1 + nc] = X[i] * dpi - fixed[i] * dpf;
            }
            for (int kk = 0; kk < nc; kk++) {
                const double dx = dphi_dx(f, a, b, T, X[nc], X, &si, i, kk);
                const double diag = (i == kk) ? si.phi[i] : 0;
                J[i * n1 + kk] = bubble ? (-diag - X[i] * dx) : (diag + X[i] * dx);
            }
        }
        or_lu_solve(n1, J, r, dX);
        XmTol = 0;
        double xs = 0;
        for (int i = 0; i < n1; i++) {
            X[i] += dX[i];
            if (i < nc) {
                if (X[i] < 0) X[i] = 0;
                else if (X[i] > 1) X[i] = 1;
                xs += X[i];
            } else if (X[nc] < 0) {
                X[nc] = OR_PZERO;
            }
            if (X[i] != 0) XmTol += fabs(dX[i] / X[i]);
            else XmTol += fabs(dX[i]);
        }
        XmTol /= n1;
        for (int i = 0; i < nc; i++) X[i] /= xs;
        nIter++;
        if (nIter > OR_MAX_MNR_ITER) {
            for (int i = 0; i < nc; i++) xm[i] = X[i];
            return -1;
        }
    } while (XmTol > OR_XMTOL);

    for (int i = 0; i < nc; i++) xm[i] = X[i];
    sumK = 0;
    k = 0;
    for (int i = 0; i < nc; i++) {
        if (fixed[i] != 0) {
            sumK += bubble ? fabs(1 - X[i] / fixed[i]) : fabs(1 - fixed[i] / X[i]);
            k++;
        }
    }
    if (sumK / k < OR_TRIV_K) return -1;
    return X[nc];
}

double or_bubble(const or_fluid *f, double T, double P, const double *x,
                 const double *ystale, double *xm)
{
    return psat_core(f, T, P, x, ystale, 1, xm);
}

double or_dew(const or_fluid *f, double T, double P, const double *y, double *xm)
{
    return psat_core(f, T, P, y, y, 0, xm);
}

void or_psat_batch(const or_fluid *f, int64_t n, const double *T, double T0,
                   const double *P, const int8_t *kind, const double *x,
                   const double *y, double *psat, double *xm, int nthreads)
{
    const int nc = f->nc;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
#endif
    for (int64_t c = 0; c < n; c++) {
        double xc[NCMAX], yc[NCMAX], mc[NCMAX];
        for (int i = 0; i < nc; i++) {
            xc[i] = x[(int64_t)i * n + c];
            yc[i] = y[(int64_t)i * n + c];
            mc[i] = 0;
        }
        double ps = -1;
        const double Tc = T ? T[c] : T0;
        if (kind[c] == 1) ps = or_bubble(f, Tc, P[c], xc, yc, mc);
        else if (kind[c] == -1) ps = or_dew(f, Tc, P[c], yc, mc);
        psat[c] = ps;
        if (xm)
            for (int i = 0; i < nc; i++) xm[(int64_t)i * n + c] = mc[i];
    }
    (void)nthreads;
}

/* ------------------------------------------------------------- PhaseCtrl */
/* PhaseCtrl / fPhaseCtrl (Do_Cycle.cpp:1558-1643 / 1644-1722) for n cells of
 * one continuum: gas cells whose gas pressure is at or below 0.95 x the dew
 * pressure, and liquid cells whose oil pressure is at or below 0.95 x the
 * bubble pressure, turn two-phase with an incipient phase of saturation
 * EPS_SAT and the saturation-pressure routine's incipient composition; any
 * other cell whose oil (gas) saturation is <= 0 becomes gas (liquid); the
 * saturations are renormalised.  Arrays are SoA; label, x, y, S are updated
 * in place; psat (may be NULL) receives the saturation pressure evaluated
 * (-1 where none was). */
void or_phase_ctrl(const or_fluid *f, int64_t n, const double *T, double T0,
                   int8_t *label, const double *Po, const double *Pg,
                   double *x, double *y, double *S, double *psat, int nthreads)
{
    const int nc = f->nc;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
#endif
    for (int64_t c = 0; c < n; c++) {
        double xc[NCMAX], yc[NCMAX], xm[NCMAX];
        for (int i = 0; i < nc; i++) {
            xc[i] = x[(int64_t)i * n + c];
            yc[i] = y[(int64_t)i * n + c];
        }
        const double Tc = T ? T[c] : T0;
        double *Sw = S + c, *So = S + n + c, *Sg = S + 2 * n + c;
        double ps = -1;
        int lab = label[c];
        int done = 0;
        if (lab == -1) {
            ps = or_dew(f, Tc, Pg[c], yc, xm);
            if (Pg[c] <= OR_PHASE_ENV * ps) {
                *So = OR_EPS_SAT;
                *Sg -= OR_EPS_SAT;
                lab = 0;
                for (int i = 0; i < nc; i++) x[(int64_t)i * n + c] = xm[i];
                done = 1;
            }
        } else if (lab == 1) {
            ps = or_bubble(f, Tc, Po[c], xc, yc, xm);
            if (Po[c] <= OR_PHASE_ENV * ps) {
                *Sg = OR_EPS_SAT;
                *So -= OR_EPS_SAT;
                lab = 0;
                for (int i = 0; i < nc; i++) y[(int64_t)i * n + c] = xm[i];
                done = 1;
            }
        }
        if (!done) {
            if (*So <= OR_EPS_SAT * OR_EPS_SAT_FRAC) {
                *So = 0;
                *Sg = 1 - *Sw;
                lab = -1;
            }
            if (*Sg <= OR_EPS_SAT * OR_EPS_SAT_FRAC) {
                *Sg = 0;
                *So = 1 - *Sw;
                lab = 1;
            }
        }
        const double inv = 1.0 / (*Sw + *So + *Sg);
        *Sw *= inv;
        *So *= inv;
        *Sg *= inv;
        label[c] = (int8_t)lab;
        if (psat) psat[c] = ps;
    }
    (void)nthreads;
}

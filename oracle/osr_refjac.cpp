// osr_refjac.cpp -- CPU restatement of the reference's hand-written Jacobian,
// IteratorJacobian (compsim/Petsc_Functions.cpp:6898-42073), row by row, in
// the reference's row and column layout.
//
// TEST INFRASTRUCTURE ONLY (see osr_oracle.h): the HIP product never links,
// loads or calls this file.
//
// The reference assembles each row from hand-derived partial derivatives of
// the quantities its residual is built from (property sweeps of
// PetscAUX.cpp:500-547), not from the exact derivative of IteratorFunction.
// This file follows those rules, including where they differ from the exact
// derivative; the differences are listed where they occur ("rule:").  The
// six phase-state sections of IteratorJacobian (fracture -1/0/1, matrix
// -1/0/1) run the same row code with phase tests inside; the per-section and
// per-face differences are reproduced here as flags.
//
// Inputs per cell, all in the reference's meaning:
//   relPerm / dRelPerm     CalcSatFs Do_Cycle.cpp:99-257 (matrix),
//                          fCalcSatFs :261-437 (fracture: kr = S)
//   dE                     Calc_Visco Do_Cycle.cpp:438-616 (d rho / d x_k,
//                          d rho / d P of each phase)
//   DiffVisco / fDiffVisco Mass_Tran.cpp:307-399 (d mu / d x_k, d mu / d P)
//   DifDiffusion           Mass_Tran.cpp:107-305: d[(1 - x_i) / S] / d x_k,
//                          i.e. 1e4 times the derivative of CalcDiffusion's
//                          coefficient (Mass_Tran.cpp:21-62 divides by 1e4)
//   trans / transS         Calc_Transes Do_Cycle.cpp:800-1108,
//                          fCalc_Transes :1109-1417
// dE and DiffVisco are the exact partial derivatives of Calc_Visco's density
// and LBC viscosity (implicit dZ), so they are taken here by forward-mode
// differentiation of the same property functions the residual uses.
//
// Pins (tests/test_oracle_refjac.py): the column pattern against the row
// code's structure and PJac's count; with no capillary pressure, diffusion or
// gravity, equality with the exact Jacobian up to the reference's sign rules
// and listed copy-paste rules; the transfer terms' sign rule; fugacity and
// constraint rows exact.  Parity unpinned (the restatement alone stands for
// them): the Fickian and diffusive-transfer derivative terms, the z-face
// gravity terms and the terms that depend on capillary pressure or IFT.
#include <omp.h>

#include "osr_model.hpp"

namespace {

// ----------------------------------------------------- per-cell quantities
struct RefCells {
    int n = 0, nc = 0;
    std::vector<int> lab;
    std::vector<double> x, y;        // [c*nc + i]
    std::vector<double> P;           // [c*3 + j]: water, oil, gas pressure
    std::vector<double> S;           // [c*3 + j]
    std::vector<double> ro, mu, gam, pc;   // [c*2 + ph]
    std::vector<double> kr;          // [c*3 + j]
    std::vector<double> dkr;         // [c*5 + j]: dRelPerm[0..4]
    std::vector<double> dE, dMu;     // [(c*(nc+1) + k)*2 + ph]
    std::vector<double> Dif;         // [(c*nc + i)*2 + ph]
    std::vector<double> dDif;        // [((c*nc + i)*(nc+1) + k)*2 + ph]
    std::vector<double> porpor, dpor, phi;

    double X(int c, int i, int ph) const { return ph == 0 ? x[(size_t)c * nc + i] : y[(size_t)c * nc + i]; }
    double dEk(int c, int k, int ph) const { return dE[((size_t)c * (nc + 1) + k) * 2 + ph]; }
    double dMuk(int c, int k, int ph) const { return dMu[((size_t)c * (nc + 1) + k) * 2 + ph]; }
    double Di(int c, int i, int ph) const { return Dif[((size_t)c * nc + i) * 2 + ph]; }
    double dDi(int c, int i, int k, int ph) const { return dDif[(((size_t)c * nc + i) * (nc + 1) + k) * 2 + ph]; }
};

// dRelPerm of the matrix (CalcSatFs, Do_Cycle.cpp:163-256): table slopes,
// Stone II derivative with respect to Sw and Sg, IFT scaling, then the sign
// clamps of :254-255 (kro's derivatives are never positive).
void matrix_dkr(const Model& M, double sw, double sg, double bift, double* d) {
    auto seg = [](const std::vector<double>& tb, int nr, double s, double& dkr, double& dkro, double& kr, double& kro) {
        int r;
        for (r = 0; r < nr; r++)
            if (s < tb[4 * r + 0]) break;
        if (r == 0) { kr = RELPERM0; kro = tb[2]; dkr = 0; dkro = 0; }
        else if (r == nr) { kr = tb[4 * (nr - 1) + 1]; kro = tb[4 * (nr - 1) + 2]; dkr = 0; dkro = 0; }
        else {
            const double ds = tb[4 * r] - tb[4 * (r - 1)];
            kr = (tb[4 * r + 1] - tb[4 * (r - 1) + 1]) / ds * (s - tb[4 * r]) + tb[4 * r + 1];
            kro = (tb[4 * r + 2] - tb[4 * (r - 1) + 2]) / ds * (s - tb[4 * r]) + tb[4 * r + 2];
            dkr = (tb[4 * r + 1] - tb[4 * (r - 1) + 1]) / ds;
            dkro = (tb[4 * r + 2] - tb[4 * (r - 1) + 2]) / ds;
        }
    };
    double dkrw, sdKrow, krw, Krow, dkrg, sdKrog, krg, Krog;
    seg(M.swt, M.nswt, sw, dkrw, sdKrow, krw, Krow);
    seg(M.sgt, M.nsgt, sg, dkrg, sdKrog, krg, Krog);
    const double kro0 = M.swt[2];
    const double AA = krw + Krow / kro0, BB = krg + Krog / kro0;
    const double dAA = dkrw + sdKrow / kro0, dBB = dkrg + sdKrog / kro0;
    d[0] = dkrw;
    d[3] = dkrg;
    d[4] = 0;
    d[1] = kro0 * (dAA * BB - dkrw);
    d[2] = kro0 * (AA * dBB - dkrg);
    if (bift < MISCIBLEIFT) {
        const double F = std::pow(bift / IFTBASECASE, 1.0 / IFTPOWER);
        d[3] = F * d[3] + (1 - F) * (1 - M.swt[0]);
        d[1] *= F;
        d[2] *= F;
        d[4] = (1 - F) * (1 - M.swt[0]);
    }
    if (d[1] > 0) d[1] = 0;
    if (d[2] > 0) d[2] = 0;
}

void ref_cells(const Model& M, const Sweep& Sw, int k, RefCells& R, int nthreads) {
    const int n = M.n, nc = M.nc;
    R.n = n;
    R.nc = nc;
    R.lab.resize(n);
    R.x.resize((size_t)n * nc); R.y.resize((size_t)n * nc);
    R.P.resize(3 * (size_t)n); R.S.resize(3 * (size_t)n); R.kr.resize(3 * (size_t)n); R.dkr.resize(5 * (size_t)n);
    R.ro.resize(2 * (size_t)n); R.mu.resize(2 * (size_t)n); R.gam.resize(2 * (size_t)n); R.pc.resize(2 * (size_t)n);
    R.dE.resize((size_t)n * (nc + 1) * 2); R.dMu.resize((size_t)n * (nc + 1) * 2);
    R.Dif.resize((size_t)n * nc * 2); R.dDif.resize((size_t)n * nc * (nc + 1) * 2);
    R.porpor.resize(n); R.dpor.resize(n); R.phi.resize(n);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int c = 0; c < n; c++) {
        const Vars<double>& v = Sw.V[k][c];
        const Props<double>& p = Sw.P[k][c];
        R.lab[c] = v.label;
        for (int i = 0; i < nc; i++) { R.x[(size_t)c * nc + i] = v.x[i]; R.y[(size_t)c * nc + i] = v.y[i]; }
        R.P[3 * c + 0] = p.Pw; R.P[3 * c + 1] = p.Po; R.P[3 * c + 2] = p.Pg;
        for (int j = 0; j < 3; j++) { R.S[3 * c + j] = v.S[j]; R.kr[3 * c + j] = p.kr[j]; }
        for (int ph = 0; ph < 2; ph++) {
            R.ro[2 * c + ph] = p.ro[ph]; R.mu[2 * c + ph] = p.mu[ph]; R.gam[2 * c + ph] = p.gam[ph];
        }
        R.pc[2 * c + 0] = p.pcw;
        R.pc[2 * c + 1] = p.pcg;
        if (k == 0) matrix_dkr(M, v.S[0], v.S[2], p.bift, &R.dkr[5 * c]);
        else {
            // fCalcSatFs (Do_Cycle.cpp:426-435): kr = S, fdRelPerm = {1, 0, 0, 1, 1}.
            const double d[5] = {1, 0, 0, 1, 1};
            for (int j = 0; j < 5; j++) R.dkr[5 * c + j] = d[j];
        }
        R.phi[c] = M.phi[k][c];
        R.porpor[c] = p.porpor;
        R.dpor[c] = M.phi[k][c] * (M.cpor + 2 * M.dcpor * (v.Po - M.refP));
        // d rho, d mu, D and dD of each phase at the phase's pressure:
        // directions 0..nc-1 = x_k, nc = P.
        for (int ph = 0; ph < 2; ph++) {
            D xd[NCM];
            for (int i = 0; i < nc; i++) {
                xd[i] = D(ph == 0 ? v.x[i] : v.y[i]);
                xd[i].n = nc + 1;
                for (int q = 0; q <= nc; q++) xd[i].d[q] = 0;
                xd[i].d[i] = 1;
            }
            D Pd(ph == 0 ? p.Po : p.Pg);
            Pd.n = nc + 1;
            for (int q = 0; q <= nc; q++) Pd.d[q] = 0;
            Pd.d[nc] = 1;
            D rod, mud, Dd[NCM];
            visco_phase<D>(M, xd, Pd, ph == 1, rod, mud);
            diffusion_phase<D>(M, xd, rod, Dd);
            for (int q = 0; q <= nc; q++) {
                R.dE[((size_t)c * (nc + 1) + q) * 2 + ph] = rod.d[q];
                R.dMu[((size_t)c * (nc + 1) + q) * 2 + ph] = mud.d[q];
            }
            for (int i = 0; i < nc; i++) {
                R.Dif[((size_t)c * nc + i) * 2 + ph] = p.D[ph][i];
                // DifDiffusion has no 1e-4 factor; a pure phase (x_i = 1)
                // gives sSum2 = 0 there and a zero derivative.
                const bool pure = (ph == 0 ? v.x[i] : v.y[i]) == 1;
                for (int q = 0; q <= nc; q++)
                    R.dDif[(((size_t)c * nc + i) * (nc + 1) + q) * 2 + ph] = pure ? 0.0 : 10000.0 * Dd[i].d[q];
            }
        }
    }
}

// Calc_Transes / fCalc_Transes for the face between a and b = a + e_d.
struct RefFace {
    double t[12];   // trans[..][0..11]
    int up[3];      // transS: 0 a upstream, 1 b upstream (oil, gas, water)
    double W, dW;   // Wtran, dWtran
};

RefFace ref_face(const Model& M, const RefCells& R, int k, int d, int a, int b, double da, double db, double ha,
                 double hb) {
    RefFace f;
    const double K = M.iftran[k][d][a];
    const double wc = M.wat_ro * WAT_M_RO / M.wat_mu;
    {
        const double pot = R.P[3 * a] - R.P[3 * b] + G_ACC * M.wat_ro * (ha - hb);
        const int u = pot > 0 ? a : b;
        f.up[2] = pot > 0 ? 0 : 1;
        f.W = K * wc * R.kr[3 * u + 0];
        f.dW = k == 0 ? K * wc * R.dkr[5 * u + 0] : K * wc;
    }
    for (int ph = 0; ph < 2; ph++) {
        const int absent = ph == 0 ? -1 : 1;
        const double c1 = R.lab[a] == absent ? 0 : 1, c2 = R.lab[b] == absent ? 0 : 1;
        double mg = (R.gam[2 * a + ph] * da * c1 + R.gam[2 * b + ph] * db * c2) / (c1 * da + db * c2);
        if (c1 == 0 && c2 == 0) mg = 0;
        const double pot = R.P[3 * a + 1 + ph] - R.P[3 * b + 1 + ph] + mg * (ha - hb);
        const int u = pot > 0 ? a : b;
        f.up[ph] = pot > 0 ? 0 : 1;
        const double T = K * (R.ro[2 * a + ph] * da * c1 + R.ro[2 * b + ph] * db * c2) /
                         (R.mu[2 * a + ph] * da * c1 + R.mu[2 * b + ph] * db * c2);
        const double kru = R.kr[3 * u + 1 + ph];
        if (ph == 0) {
            f.t[0] = T * kru;
            f.t[2] = mg * T * kru;
            if (k == 0) {
                const double* dk = &R.dkr[5 * u];
                f.t[4] = T * dk[1]; f.t[6] = mg * T * dk[1];
                f.t[8] = T * dk[2]; f.t[9] = mg * T * dk[2];
                f.t[10] = T * dk[4]; f.t[11] = mg * T * dk[4];
            } else {
                // fCalc_Transes: d kro / d Sw = d kro / d Sg = -1, d kro / d So = 1
                f.t[4] = T * (-1); f.t[6] = mg * T * (-1);
                f.t[8] = T * (-1); f.t[9] = mg * T * (-1);
                f.t[10] = T; f.t[11] = mg * T;
            }
            if (c1 == 0 && c2 == 0) f.t[0] = f.t[2] = f.t[4] = f.t[6] = f.t[8] = f.t[9] = 0;
        } else {
            f.t[1] = T * kru;
            f.t[3] = mg * T * kru;
            if (k == 0) { f.t[5] = T * R.dkr[5 * u + 3]; f.t[7] = mg * T * R.dkr[5 * u + 3]; }
            else { f.t[5] = T; f.t[7] = mg * T; }
            if (c1 == 0 && c2 == 0) f.t[1] = f.t[3] = f.t[5] = f.t[7] = 0;
        }
    }
    return f;
}

// ------------------------------------------------------------ row emitter
struct RowOut {
    std::vector<int64_t> col;
    std::vector<double> val;
    void add(int64_t c, double v) { col.push_back(c); val.push_back(v); }
    void clear() { col.clear(); val.clear(); }
};

struct Ctx {
    const Model& M;
    const Sweep* S = nullptr;     // property sweep (the well terms read it)
    const RefCells* R[2];
    std::vector<RefFace> F[2];   // [c*3 + d]: face to the +d neighbour
    std::vector<int64_t> boff;   // preConRow
    double pcoeff;
    int n, nc;
    explicit Ctx(const Model& m) : M(m) {}
};

// Geometry of one face of cell c seen from c.
struct FaceGeo {
    int q, d, s;     // face index (x+, x-, y+, y-, z+, z-), direction, +1/-1
    int o;           // neighbour cell
    int a, b;        // lower and upper cell of the face
    double GS, GN;   // cell sizes along d: self, neighbour
    double hs, hn;   // block heights
    double area;     // 2 x the face's transverse cell dimensions (NNd numerator)
};

bool face_geo(const Model& M, int c, int q, FaceGeo& g) {
    const int iz = c / (M.nx * M.ny), rem = c % (M.nx * M.ny), iy = rem / M.nx, ix = rem % M.nx;
    g.q = q;
    g.d = q / 2;
    g.s = (q % 2 == 0) ? 1 : -1;
    const int stride[3] = {1, M.nx, M.nx * M.ny};
    const int idx[3] = {ix, iy, iz};
    const int dim[3] = {M.nx, M.ny, M.nz};
    const int j = idx[g.d] + g.s;
    if (j < 0 || j >= dim[g.d]) return false;
    g.o = c + g.s * stride[g.d];
    g.a = g.s > 0 ? c : g.o;
    g.b = g.s > 0 ? g.o : c;
    const std::vector<double>* sz[3] = {&M.dx, &M.dy, &M.dz};
    g.GS = (*sz[g.d])[idx[g.d]];
    g.GN = (*sz[g.d])[j];
    g.hs = M.h[iz];
    g.hn = g.d == 2 ? M.h[j] : M.h[iz];
    g.area = g.d == 0 ? 2 * M.dy[iy] * M.dz[iz] : (g.d == 1 ? 2 * M.dx[ix] * M.dz[iz] : 2 * M.dx[ix] * M.dy[iy]);
    return true;
}

// Which side of the face is upstream for phase p (0 oil, 1 gas, 2 water):
// true if the neighbour is.
inline bool nb_up(const RefFace& f, const FaceGeo& g, int p) { return g.s > 0 ? f.up[p] == 1 : f.up[p] == 0; }

}  // namespace

namespace {

// Per-face deviations of the unrolled face code from the x+/x- pattern
// (component rows of the matrix, Petsc_Functions.cpp:31012-34627 and its
// copies; total rows :34627-36182 and copies).
struct FaceQuirk {
    bool gas_gam_coef_swapped = false;  // comp, y+: the gas MoshtaqGama carries Coef, BmoshtaqGama not
    bool gas_overwrites_nbp = false;    // comp, x-: the gas part assigns dBsidedp instead of adding
    bool tot_oil_grav_overwrite = false;   // total, y-: oil gravity (neighbour upstream) assigns dBsidedp
    bool tot_gas_p_overwrite = false;      // total, y-: gas pressure term (neighbour upstream) assigns dBsidedp
};

FaceQuirk comp_quirk(int q) {
    FaceQuirk f;
    if (q == 2) f.gas_gam_coef_swapped = true;   // y+ (:31539-31541)
    if (q == 1) f.gas_overwrites_nbp = true;     // x- (:31466, :31473)
    return f;
}
FaceQuirk tot_quirk(int q) {
    FaceQuirk f;
    if (q == 3) { f.tot_oil_grav_overwrite = true; f.tot_gas_p_overwrite = true; }  // y-
    return f;
}

// Shape factors (Init_Calcs.cpp:970-983).
struct Shape { double sf1, sf2, sf3; };
Shape shape(const Model& M, int c) {
    const int n = M.n;
    const int iz = c / (M.nx * M.ny), rem = c % (M.nx * M.ny), iy = rem / M.nx, ix = rem % M.nx;
    Shape s;
    s.sf2 = 4 * M.sigma * (M.perm[0][c] / (M.dx[ix] * M.dx[ix]) + M.perm[0][n + c] / (M.dy[iy] * M.dy[iy]) +
                           M.perm[0][2 * n + c] / (M.dz[iz] * M.dz[iz]));
    const bool off = M.const_diff && M.cdiff == 0;
    s.sf1 = off ? 0 : 4 * M.sigma * (1 / (M.dx[ix] * M.dx[ix]) + 1 / (M.dy[iy] * M.dy[iy]) + 1 / (M.dz[iz] * M.dz[iz]));
    s.sf3 = off ? 0 : 4 * M.sigma * ((1 / M.dx[ix]) + (1 / M.dy[iy]) + (1 / M.dz[iz])) / M.lf;
    return s;
}

// f_m_relperm / f_m_drelperm ("sadeq" regions, e.g. Petsc_Functions.cpp:30409-30466):
// upstream kr between the two continua by phase pressure.
struct FM {
    double kr[3];        // water, oil, gas
    double d[3][10];     // f_m_drelperm[ph][..], zero where unset
};
FM fm_terms(const RefCells& m, const RefCells& f, int c) {
    FM r;
    for (int a = 0; a < 3; a++) for (int b = 0; b < 10; b++) r.d[a][b] = 0;
    if (m.P[3 * c + 1] >= f.P[3 * c + 1]) {
        r.kr[1] = m.kr[3 * c + 1];
        r.d[1][1] = m.dkr[5 * c + 1]; r.d[1][2] = m.dkr[5 * c + 2]; r.d[1][4] = m.dkr[5 * c + 4];
    } else {
        r.kr[1] = f.kr[3 * c + 1];
        r.d[1][6] = f.dkr[5 * c + 1]; r.d[1][7] = f.dkr[5 * c + 2]; r.d[1][9] = f.dkr[5 * c + 4];
    }
    if (m.P[3 * c + 2] >= f.P[3 * c + 2]) { r.kr[2] = m.kr[3 * c + 2]; r.d[2][3] = m.dkr[5 * c + 3]; }
    else { r.kr[2] = f.kr[3 * c + 2]; r.d[2][8] = f.dkr[5 * c + 3]; }
    if (m.P[3 * c + 0] >= f.P[3 * c + 0]) { r.kr[0] = m.kr[3 * c + 0]; r.d[1][0] = m.dkr[5 * c + 0]; }
    else { r.kr[0] = f.kr[3 * c + 0]; r.d[1][5] = f.dkr[5 * c + 0]; }
    return r;
}

// Diffusive-transfer label combinations of the matrix rows: for each
// (matrix label, fracture label) the phase terms the residual adds
// (Petsc_Functions.cpp:5259-5313), as (phase, uses fracture composition):
// 'F' = (x_f - x_m), 'E' = (EdgeComp - x_m) = (0 - x_m).
struct DTerm { int ph; char kind; };
int mdiff_terms(int lm, int lf, DTerm* t) {
    int n = 0;
    if (lm == 1 && lf == 1) t[n++] = {0, 'F'};
    if (lm == -1 && lf == -1) t[n++] = {1, 'F'};
    if (lm == 0 && lf == 0) { t[n++] = {0, 'F'}; t[n++] = {1, 'F'}; }
    if (lm == 1 && lf == -1) t[n++] = {0, 'E'};
    if (lm == -1 && lf == 1) t[n++] = {1, 'E'};
    if (lm == 0 && lf == 1) { t[n++] = {1, 'E'}; t[n++] = {0, 'F'}; }
    if (lm == 0 && lf == -1) { t[n++] = {1, 'F'}; t[n++] = {0, 'E'}; }
    if (lm == 1 && lf == 0) { t[n++] = {0, 'E'}; t[n++] = {0, 'F'}; }
    if (lm == -1 && lf == 0) { t[n++] = {1, 'F'}; t[n++] = {1, 'E'}; }
    return n;
}

}  // namespace

namespace {

// Diffusive-transfer terms of the fracture rows (Petsc_Functions.cpp:731-806):
// for each (matrix label, fracture label), the terms built from matrix
// properties with shape factor 1 (x_f - x_m) and from fracture properties
// with shape factor 3 (x_f - EdgeComp), as (phase).
int fdiff_sf1(int lm, int lf, int* ph) {
    int n = 0;
    if (lm == 1 && lf == 1) ph[n++] = 0;
    if (lm == -1 && lf == -1) ph[n++] = 1;
    if (lm == 0 && lf == 0) { ph[n++] = 0; ph[n++] = 1; }
    if (lm == 0 && lf == 1) ph[n++] = 0;
    if (lm == 0 && lf == -1) ph[n++] = 1;
    if (lm == 1 && lf == 0) ph[n++] = 0;
    if (lm == -1 && lf == 0) ph[n++] = 1;
    return n;
}
int fdiff_sf3(int lm, int lf, int* ph) {
    int n = 0;
    if (lm == 1 && lf == -1) ph[n++] = 1;
    if (lm == -1 && lf == 1) ph[n++] = 0;
    if (lm == 0 && lf == 1) ph[n++] = 0;
    if (lm == 1 && lf == 0) ph[n++] = 1;
    if (lm == 0 && lf == -1) ph[n++] = 1;
    if (lm == -1 && lf == 0) ph[n++] = 0;
    return n;
}

// Own-cell pieces of a balance row that are not face fluxes: accumulation
// derivatives, the matrix-fracture transfer and the diffusive transfer.
// "O" = the other continuum of the same cell.
struct OwnP {
    double dLp, dLSo, dLSg;          // dLsidedp, dLsidedSo, dLsidedSg
    double dRp, dROp;                // own P accumulator part, other-continuum P column
    double dRSo, dRSg, dRSw;         // own S accumulators
    double dROSo, dROSg, dROSw;      // other-continuum S columns
};
struct OwnX { double dR[2], dRO[2], dL[2]; };

// One component-balance row (total = false, component i) or the total
// hydrocarbon row (total = true) of block (k, c).  Matrix component rows:
// Petsc_Functions.cpp:31012-34627 (two-phase section; the liquid and gas
// sections :25028-28651, :36841-40456 run the same code); matrix total
// rows :34627-36182 (and copies); fracture rows :13505-18623 (and the
// liquid and gas copies), whose face code is the matrix's with tortuosity 1.
// Column layout (InitGuess2): two-phase [x, y, P/PCoeff, Sw, So, Sg];
// liquid [x, P, Sw, So]; gas [y, P, Sw, Sg].
//
// Sign conventions kept from the reference: the row value is
// dLSide - dRSide for x and S columns and dLSide + dRsidedp for the own P
// column, where dRsidedp collects -dRSide/dP of the face fluxes but
// +dRSide/dP of the matrix-fracture transfer terms; the other continuum's
// columns are emitted as +dRSide/d(unknown).  (rule: the transfer part of
// the P column and every other-continuum column carry the opposite sign of
// the residual's derivative.)
void balance_row(const Ctx& X, int kc, int c, int i, bool total, RowOut& out) {
    const Model& M = X.M;
    const RefCells& s = *X.R[kc];        // the row's continuum
    const RefCells& o = *X.R[1 - kc];    // the other continuum of the cell
    const RefCells& m = *X.R[0];
    const RefCells& f = *X.R[1];
    const int nc = X.nc, n = X.n;
    const int lab = s.lab[c], olab = o.lab[c];
    const int lm = m.lab[c], lf = f.lab[c];
    const int reCol = lab == 0 ? 0 : nc;
    const int reColF1 = olab == 0 ? 0 : nc;
    const int64_t jacCol = X.boff[(size_t)kc * n + c], OEq = X.boff[(size_t)(1 - kc) * n + c];
    const int iz = c / (M.nx * M.ny), rem = c % (M.nx * M.ny), iy = rem / M.nx, ix = rem % M.nx;
    const double Vb = M.dx[ix] * M.dy[iy] * M.dz[iz];
    const double Dt = M.dt;
    const Shape sh = shape(M, c);
    const FM fm = fm_terms(m, f, c);
    const double tor = M.tor[c];                     // diffusive transfer: matrix tortuosity in both
    auto ftor = [&](int cell) { return kc == 0 ? M.tor[cell] : 1.0; };   // inter-cell Fickian flux
    const double porpor = s.porpor[c], dPordp = s.dpor[c];
    const double mpor = m.porpor[c], fpor = f.porpor[c], mdpor = m.dpor[c], fdpor = f.dpor[c];
    const double MUP[2] = {m.mu[2 * c] * mpor + f.mu[2 * c] * fpor, m.mu[2 * c + 1] * mpor + f.mu[2 * c + 1] * fpor};
    const double PP = mpor + fpor;
    const double DPh[2] = {f.P[3 * c + 1] - m.P[3 * c + 1],
                           (f.P[3 * c + 2] - m.P[3 * c + 2]) + (f.pc[2 * c + 1] - m.pc[2 * c + 1])};
    const double cPh[2] = {sh.sf2 * Vb * fm.kr[1], sh.sf2 * Vb * fm.kr[2]};
    auto cx = [&](const RefCells& r, int cell, int ph) { return total ? 1.0 : r.X(cell, i, ph); };
    const double sx[2] = {cx(s, c, 0), cx(s, c, 1)};
    const double mx[2] = {cx(m, c, 0), cx(m, c, 1)};   // the transfer term is built on the matrix
    const double ro_s[2] = {s.ro[2 * c], s.ro[2 * c + 1]}, ro_m[2] = {m.ro[2 * c], m.ro[2 * c + 1]};
    const double S_s[2] = {s.S[3 * c + 1], s.S[3 * c + 2]};
    const int dfd[2] = {1, 7}, dfo[2] = {4, 9};   // f_m_drelperm slots (unused helper names)
    (void)dfd; (void)dfo;

    out.add(jacCol + (lab == 0 ? (total ? 2 * nc - 1 : nc + i) : (total ? nc - 1 : i)), 0.0);  // diagonal placeholder

    // ------------------------------------------------------- own-cell pieces
    OwnP w{};
    w.dLp = (Vb / Dt) * (sx[0] * S_s[0] * (ro_s[0] * dPordp + s.dEk(c, nc, 0) * porpor) +
                         sx[1] * S_s[1] * (ro_s[1] * dPordp + s.dEk(c, nc, 1) * porpor));
    w.dLSo = (Vb / Dt) * (sx[0] * porpor * ro_s[0]);
    w.dLSg = (Vb / Dt) * (sx[1] * porpor * ro_s[1]);
    if (kc == 0) {
        // matrix rows (:31080-31206, total :34629-34665)
        for (int ph = 0; ph < 2; ph++) {
            w.dRp += SI_C * m.dEk(c, nc, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
            w.dRp += -SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            w.dRp += -SI_C * (m.dMuk(c, nc, ph) * mpor + mdpor * m.mu[2 * c + ph]) * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            w.dRp += SI_C * mdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        if (!total) {
            DTerm dt[4];
            const int nt = mdiff_terms(lm, lf, dt);
            for (int t = 0; t < nt; t++) {
                const int ph = dt[t].ph;
                const double dl = dt[t].kind == 'F' ? f.X(c, i, ph) - m.X(c, i, ph) : 0.0 - m.X(c, i, ph);
                w.dRp += mdpor * sh.sf1 * S_s[ph] * Vb / tor * ro_s[ph] * m.Di(c, i, ph) * dl;
                w.dRp += m.dEk(c, nc, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * m.Di(c, i, ph) * dl;
                w.dRp += m.dDi(c, i, nc, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_s[ph] * dl;
                if (ph == 0) w.dRSo += sh.sf1 * Vb * mpor / tor * ro_s[0] * m.Di(c, i, 0) * dl;
                // rule: (matrix two-phase, fracture liquid) has no gas Sg term (:31172-31189)
                if (ph == 1 && !(lm == 0 && lf == 1)) w.dRSg += sh.sf1 * Vb * mpor / tor * ro_s[1] * m.Di(c, i, 1) * dl;
            }
        }
        for (int ph = 0; ph < 2; ph++) {
            w.dROp += SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            w.dROp += -SI_C * (f.dMuk(c, nc, ph) * fpor + fdpor * f.mu[2 * c + ph]) * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            w.dROp += SI_C * fdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        if (!total && lm == 1 && lf == 0) {
            // liquid-section component rows only (:25170-25174): the fracture
            // gas-phase diffusion term differentiated by the fracture pressure
            const double dl = f.X(c, i, 1) - 0.0;
            w.dROp += fdpor * sh.sf1 * f.S[3 * c + 2] * Vb / tor * f.ro[2 * c + 1] * f.Di(c, i, 1) * dl;
            w.dROp += f.dEk(c, nc, 1) * sh.sf1 * f.S[3 * c + 2] * fpor * Vb / tor * f.Di(c, i, 1) * dl;
            w.dROp += f.dDi(c, i, nc, 1) * sh.sf1 * f.S[3 * c + 2] * fpor * Vb / tor * f.ro[2 * c + 1] * dl;
        }
        w.dRSo += SI_C * fm.d[1][4] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dROSo += SI_C * fm.d[1][9] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dRSg += SI_C * fm.d[1][2] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dRSg += SI_C * fm.d[2][3] * mx[1] * ro_m[1] * (sh.sf2 * Vb) / (MUP[1] / PP) * DPh[1];
        w.dROSg += SI_C * fm.d[1][7] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dROSg += SI_C * fm.d[2][8] * mx[1] * ro_m[1] * (sh.sf2 * Vb) / (MUP[1] / PP) * DPh[1];
        w.dRSw += SI_C * fm.d[1][1] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dROSw += SI_C * fm.d[1][6] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
    } else {
        // fracture rows (:13568-13640, total :17078-17114): the transfer term
        // is the matrix row's, subtracted; shape factor 3 terms use the
        // fracture's properties.
        for (int ph = 0; ph < 2; ph++) {
            w.dROp -= SI_C * m.dEk(c, nc, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
            w.dROp -= -SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            w.dROp -= -SI_C * (m.dMuk(c, nc, ph) * mpor + mdpor * m.mu[2 * c + ph]) * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            w.dROp -= SI_C * mdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        int p1[4], p3[4];
        const int n1 = total ? 0 : fdiff_sf1(lm, lf, p1);
        const int n3 = total ? 0 : fdiff_sf3(lm, lf, p3);
        for (int t = 0; t < n1; t++) {
            const int ph = p1[t];
            const double dl = f.X(c, i, ph) - m.X(c, i, ph);
            const double Sm = m.S[3 * c + 1 + ph];
            w.dROp -= mdpor * sh.sf1 * Sm * Vb / tor * ro_m[ph] * m.Di(c, i, ph) * dl;
            w.dROp -= m.dEk(c, nc, ph) * sh.sf1 * Sm * mpor * Vb / tor * m.Di(c, i, ph) * dl;
            w.dROp -= m.dDi(c, i, nc, ph) * sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * dl;
            if (ph == 0) w.dROSo -= sh.sf1 * Vb * mpor / tor * ro_m[0] * m.Di(c, i, 0) * dl;
            else w.dROSg -= sh.sf1 * Vb * mpor / tor * ro_m[1] * m.Di(c, i, 1) * dl;
        }
        for (int ph = 0; ph < 2; ph++) {
            w.dRp -= SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            w.dRp -= -SI_C * (f.dMuk(c, nc, ph) * fpor + fdpor * f.mu[2 * c + ph]) * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            w.dRp -= SI_C * fdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        for (int t = 0; t < n3; t++) {
            const int ph = p3[t];
            const double dl = f.X(c, i, ph) - 0.0;
            w.dRp -= fdpor * sh.sf3 * S_s[ph] * Vb / tor * ro_s[ph] * f.Di(c, i, ph) * dl;
            w.dRp -= f.dEk(c, nc, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * f.Di(c, i, ph) * dl;
            w.dRp -= f.dDi(c, i, nc, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * ro_s[ph] * dl;
            if (ph == 0) w.dRSo -= sh.sf3 * Vb * fpor / tor * ro_s[0] * f.Di(c, i, 0) * dl;
            else w.dRSg -= sh.sf3 * Vb * fpor / tor * ro_s[1] * f.Di(c, i, 1) * dl;
        }
        w.dROSo -= SI_C * fm.d[1][4] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dRSo -= SI_C * fm.d[1][9] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dROSg -= SI_C * fm.d[1][2] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dROSg -= SI_C * fm.d[2][3] * mx[1] * ro_m[1] * (sh.sf2 * Vb) / (MUP[1] / PP) * DPh[1];
        w.dRSg -= SI_C * fm.d[1][7] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dRSg -= SI_C * fm.d[2][8] * mx[1] * ro_m[1] * (sh.sf2 * Vb) / (MUP[1] / PP) * DPh[1];
        w.dROSw -= SI_C * fm.d[1][1] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        w.dRSw -= SI_C * fm.d[1][6] * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
    }

    // advective fluxes: d(P_nb - P)/dP of every face, upstream composition
    FaceGeo G[6];
    bool has[6];
    for (int q = 0; q < 6; q++) has[q] = face_geo(M, c, q, G[q]);
    double dRsidedp = w.dRp;
    for (int ph = 0; ph < 2; ph++)
        for (int q = 0; q < 6; q++) {
            if (!has[q]) continue;
            const RefFace& F = X.F[kc][(size_t)G[q].a * 3 + G[q].d];
            const int up = F.up[ph] == 0 ? G[q].a : G[q].b;
            dRsidedp += cx(s, up, ph) * F.t[ph];
        }
    const double dCoef[2] = {lab == -1 ? 0.0 : 1.0, lab == 1 ? 0.0 : 1.0};
    const double dMu_p[2] = {s.dMuk(c, nc, 0), s.dMuk(c, nc, 1)};
    double dRsidedSw = w.dRSw, dRsidedSg = w.dRSg, dRsidedSo = w.dRSo;

    // ------------------------------------------------ faces: P and S columns
    for (int q = 0; q < 6; q++) {
        if (!has[q]) continue;
        const FaceGeo& g = G[q];
        const int on = g.o, sg = g.s;
        const RefFace& F = X.F[kc][(size_t)g.a * 3 + g.d];
        const FaceQuirk qk = total ? tot_quirk(q) : comp_quirk(q);
        const int nlab = s.lab[on];
        const int reCol2 = nlab ? nc : 0, reCol4 = nlab ? 1 : 0;
        const int64_t mBLNo = X.boff[(size_t)kc * n + on] - reCol2;
        const int upo = F.up[0] == 0 ? g.a : g.b, upg = F.up[1] == 0 ? g.a : g.b;
        const bool nbo = nb_up(F, g, 0), nbg = nb_up(F, g, 1);
        const double dPo = s.P[3 * g.b + 1] - s.P[3 * g.a + 1];
        const double dPg = s.P[3 * g.b + 2] - s.P[3 * g.a + 2];
        const double dh = (sg > 0 ? g.hn - g.hs : g.hs - g.hn);   // h_b - h_a
        const double dsw = sg * (cx(s, upo, 0) * F.t[4] * dPo + cx(s, upo, 0) * F.t[6] * dh);
        if (nbo) out.add(mBLNo + 2 * nc + 1, dsw);
        else { dRsidedSw += dsw; out.add(mBLNo + 2 * nc + 1, 0.0); }
        const double dso = sg * (cx(s, upo, 0) * F.t[10] * dPo + cx(s, upo, 0) * F.t[11] * dh);
        if (nbo) out.add(mBLNo + 2 * nc + 2, nlab == -1 ? 0.0 : dso);
        else { dRsidedSo += dso; out.add(mBLNo + 2 * nc + 2, 0.0); }
        const double dsgo = sg * (cx(s, upo, 0) * F.t[8] * dPo + cx(s, upo, 0) * F.t[9] * dh);
        const double dsgg = sg * (cx(s, upg, 1) * F.t[5] * dPg + cx(s, upg, 1) * F.t[7] * dh);
        double vg;
        if (nbo) vg = nlab == 1 ? 0.0 : dsgo;
        else { dRsidedSg += dsgo; vg = 0; }
        if (nbg) vg += nlab == 1 ? 0.0 : dsgg;
        else dRsidedSg += dsgg;
        out.add(mBLNo + 2 * nc + 3 - reCol4, vg);

        double dBsidedp = 0;
        const double IF = M.iftran[kc][g.d][g.a];
        for (int ph = 0; ph < 2; ph++) {
            const double Coef = ph == 0 ? (nlab == -1 ? 0.0 : 1.0) : (nlab == 1 ? 0.0 : 1.0);
            const double Surat = s.ro[2 * c + ph] * g.GS + s.ro[2 * on + ph] * g.GN * Coef;
            const double Makhraj = s.mu[2 * c + ph] * g.GS + s.mu[2 * on + ph] * g.GN * Coef;
            const double Mo = (s.dEk(c, nc, ph) * g.GS * Makhraj - dMu_p[ph] * g.GS * Surat) / (Makhraj * Makhraj);
            const double Bm = (s.dEk(on, nc, ph) * g.GN * Coef * Makhraj - s.dMuk(on, nc, ph) * g.GN * Coef * Surat) /
                              (Makhraj * Makhraj);
            double MG = g.GS * (s.gam[2 * c + ph] * s.dEk(c, nc, ph) / s.ro[2 * c + ph]) / (g.GS + g.GN * Coef);
            double BMG = Coef * g.GN * (s.gam[2 * on + ph] * s.dEk(on, nc, ph) / s.ro[2 * on + ph]) / (g.GS + g.GN * Coef);
            if (ph == 1 && qk.gas_gam_coef_swapped) {
                MG = Coef * g.GS * (s.gam[2 * c + ph] * s.dEk(c, nc, ph) / s.ro[2 * c + ph]) / (g.GS + g.GN * Coef);
                BMG = g.GN * (s.gam[2 * on + ph] * s.dEk(on, nc, ph) / s.ro[2 * on + ph]) / (g.GS + g.GN * Coef);
            }
            const double tempdx = (s.gam[2 * c + ph] * g.GS + s.gam[2 * on + ph] * g.GN * Coef) / (g.GS + g.GN * Coef);
            const bool nbu = ph == 0 ? nbo : nbg;
            const int u = nbu ? on : c;
            const double cu = cx(s, u, ph) * IF * s.kr[3 * u + 1 + ph];
            const double dP = ph == 0 ? dPo : dPg;
            // gravity: the minus faces' gas term takes the neighbour's
            // derivative for the own column and the own one for the
            // neighbour's (:31468-31479 and copies)
            const bool swap = ph == 1 && sg < 0;
            const double gOwn = swap ? (Bm * tempdx + BMG * Surat / Makhraj) : (Mo * tempdx + MG * Surat / Makhraj);
            const double gNb = swap ? (Mo * tempdx + MG * Surat / Makhraj) : (Bm * tempdx + BMG * Surat / Makhraj);
            dRsidedp -= sg * cu * Mo * dP;
            const double pNb = sg * cu * Bm * dP;
            if (ph == 0) dBsidedp = pNb;
            else if (qk.gas_overwrites_nbp || (qk.tot_gas_p_overwrite && nbu)) dBsidedp = pNb;
            else dBsidedp += pNb;
            dRsidedp -= sg * cu * gOwn * dh;
            if (ph == 0 && qk.tot_oil_grav_overwrite && nbu) dBsidedp = sg * cu * gNb * dh;
            else dBsidedp += sg * cu * gNb * dh;
            if (total) continue;
            // inter-cell Fickian flux, pressure derivative (:31247-31283)
            const double dC = dCoef[ph];
            const double porpor2 = s.porpor[on], dPordp2 = s.dpor[on];
            double Dup, dDdp, dDdp2;
            if (s.X(on, i, ph) > s.X(c, i, ph)) { Dup = s.Di(on, i, ph); dDdp = 0; dDdp2 = s.dDi(on, i, nc, ph); }
            else { Dup = s.Di(c, i, ph); dDdp = s.dDi(c, i, nc, ph); dDdp2 = 0; }
            if (s.phi[c] < 1e-10 || s.phi[on] < 1e-10) { Dup = 0; dDdp = 0; dDdp2 = 0; }
            const double Sv = s.S[3 * on + 1 + ph], Ss = s.S[3 * c + 1 + ph];
            double AAd = s.ro[2 * on + ph] * g.GN * Coef + s.ro[2 * c + ph] * g.GS * dC;
            AAd *= (Sv * g.GN * Coef + Ss * g.GS * dC) / (g.GN + g.GS);
            const double BBd = ftor(on) * g.GN * Coef / porpor2 + ftor(c) * g.GS * dC / porpor;
            const double MMd = AAd / BBd;
            const double NNd = g.area / (g.GN + g.GS);
            const double dX = s.X(g.b, i, ph) - s.X(g.a, i, ph);
            const double dAAddp = s.dEk(c, nc, ph) * g.GS * dC;
            const double dAAddp2 = s.dEk(on, nc, ph) * g.GN * Coef;
            const double dBBddp = -dPordp * ftor(c) * g.GS * dC / (porpor * porpor);
            const double dBBddp2 = -dPordp2 * ftor(on) * g.GN * Coef / (porpor2 * porpor2);
            const double dMMddp = (dAAddp * BBd - dBBddp * AAd) / (BBd * BBd);
            const double dMMddp2 = (dAAddp2 * BBd - dBBddp2 * AAd) / (BBd * BBd);
            if (Coef != 0 && dC != 0) {
                dRsidedp += -sg * NNd * dX * (MMd * dDdp + Dup * dMMddp);
                dBsidedp += sg * NNd * dX * (MMd * dDdp2 + Dup * dMMddp2);
            }
        }
        out.add(mBLNo + 2 * nc, X.pcoeff * (-cx(s, upo, 0) * F.t[0] - cx(s, upg, 1) * F.t[1] - dBsidedp));
    }

    // ------------------------------------------------ own and other P, S
    out.add(jacCol + 2 * nc - reCol, X.pcoeff * (w.dLp + dRsidedp));
    // wells (matrix :32543, total :35425; fracture :15033, :17868): the
    // productivity indices in the P column -- the oil part on the block's
    // own x_i, the gas part on the injection stream's y_i in an injector's
    // block -- and, for a producer, -Qo, -Qg in the x_i, y_i columns
    // (:32808-32813; zero for injectors, and in the total row)
    WellQ<double> wq;
    if (M.nw) wq = well_terms<double>(M, kc, c, ix, iy, iz, X.S->P[kc][c], X.S->P);
    if (M.nw) {
        if (total) out.add(jacCol + 2 * nc - reCol, X.pcoeff * (wq.PIo + wq.PIg));
        else {
            const double gc = wq.inj >= 0 ? M.inj[(size_t)(2 * nc + 6) * wq.inj + nc + i] : s.X(c, i, 1);
            out.add(jacCol + 2 * nc - reCol, X.pcoeff * (s.X(c, i, 0) * wq.PIo + gc * wq.PIg));
            if (wq.inj < 0) {
                if (lab != -1) out.add(jacCol + i, -wq.Qo);
                if (lab != 1) out.add(jacCol + nc + i - reCol, -wq.Qg);
            }
        }
    }
    int reCol3;
    if (lab != -1) {
        out.add(jacCol + 2 * nc + 1 - reCol, -dRsidedSw);
        out.add(jacCol + 2 * nc + 2 - reCol, w.dLSo - dRsidedSo);
        reCol3 = 0;
    } else {
        reCol3 = 1;   // rule: a gas block's own Sw column is not emitted here
    }
    if (lab != 1) out.add(jacCol + 2 * nc + 3 - reCol - reCol3, w.dLSg - dRsidedSg);
    if (olab == 1) out.add(OEq + 2 * nc + 2 - reColF1, w.dROSo);
    else if (olab == -1) out.add(OEq + 2 * nc + 2 - reColF1, w.dROSg);
    else {
        out.add(OEq + 2 * nc + 2 - reColF1, w.dROSo);
        out.add(OEq + 2 * nc + 3 - reColF1, w.dROSg);
    }
    out.add(OEq + 2 * nc - reColF1, X.pcoeff * w.dROp);
    out.add(OEq + 2 * nc + 1 - reColF1, w.dROSw);

    // -------------------------------------------------- composition columns
    for (int k = 0; k < nc; k++) {
        const bool same = !total && k == i;   // the k == i branch (:32238-32950)
        OwnX v{};
        if (kc == 0) {
            for (int ph = 0; ph < 2; ph++) {
                if (same) v.dR[ph] += SI_C * ro_m[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                v.dR[ph] += SI_C * m.dEk(c, k, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                v.dR[ph] += -SI_C * m.dMuk(c, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * mpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
                v.dRO[ph] += -SI_C * f.dMuk(c, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * fpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            }
            if (!total) {
                DTerm dt[4];
                const int nt = mdiff_terms(lm, lf, dt);
                for (int t = 0; t < nt; t++) {
                    const int ph = dt[t].ph;
                    const double dl = dt[t].kind == 'F' ? f.X(c, i, ph) - m.X(c, i, ph) : 0.0 - m.X(c, i, ph);
                    v.dR[ph] += m.dEk(c, k, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * m.Di(c, i, ph) * dl;
                    if (same) v.dR[ph] += -sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_m[ph] * m.Di(c, i, ph);
                    v.dR[ph] += m.dDi(c, i, k, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_m[ph] * dl;
                }
                if (same) {
                    // d(x_f,i - x_m,i)/d x_f,i of the combinations that use the
                    // fracture composition (:32100-32120, :32206-32225); the
                    // (gas, two-phase) one sits only in the liquid and
                    // two-phase sections, where the matrix is never gas.
                    for (int ph = 0; ph < 2; ph++) {
                        bool on = false;
                        if (ph == 0) on = (lm == 1 && lf == 1) || (lm == 0 && lf == 0) || (lm == 0 && lf == 1) || (lm == 1 && lf == 0);
                        else on = (lm == -1 && lf == -1) || (lm == 0 && lf == 0) || (lm == 0 && lf == -1);
                        if (on) v.dRO[ph] += sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_m[ph] * m.Di(c, i, ph);
                    }
                }
            }
            for (int ph = 0; ph < 2; ph++)
                v.dL[ph] = total ? m.dEk(c, k, ph) : (same ? ro_m[ph] : 0.0) + sx[ph] * m.dEk(c, k, ph);
        } else {
            for (int ph = 0; ph < 2; ph++) {
                if (same) v.dRO[ph] -= SI_C * ro_m[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                v.dRO[ph] -= SI_C * m.dEk(c, k, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                v.dRO[ph] -= -SI_C * m.dMuk(c, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * mpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
                v.dR[ph] -= -SI_C * f.dMuk(c, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * fpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            }
            if (!total) {
                int p1[4], p3[4];
                const int n1 = fdiff_sf1(lm, lf, p1), n3 = fdiff_sf3(lm, lf, p3);
                for (int t = 0; t < n1; t++) {
                    const int ph = p1[t];
                    const double dl = f.X(c, i, ph) - m.X(c, i, ph);
                    const double Sm = m.S[3 * c + 1 + ph];
                    v.dRO[ph] -= m.dEk(c, k, ph) * sh.sf1 * Sm * mpor * Vb / tor * m.Di(c, i, ph) * dl;
                    if (same) v.dRO[ph] -= -sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * m.Di(c, i, ph);
                    v.dRO[ph] -= m.dDi(c, i, k, ph) * sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * dl;
                    // d(x_f,i - x_m,i)/d x_f,i of the same terms
                    if (same) v.dR[ph] -= sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * m.Di(c, i, ph);
                }
                for (int t = 0; t < n3; t++) {
                    const int ph = p3[t];
                    const double dl = f.X(c, i, ph) - 0.0;
                    v.dR[ph] -= f.dEk(c, k, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * f.Di(c, i, ph) * dl;
                    if (same) v.dR[ph] -= sh.sf3 * S_s[ph] * fpor * Vb / tor * ro_s[ph] * f.Di(c, i, ph);
                    v.dR[ph] -= f.dDi(c, i, k, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * ro_s[ph] * dl;
                }
            }
            for (int ph = 0; ph < 2; ph++)
                v.dL[ph] = total ? f.dEk(c, k, ph) : (same ? ro_s[ph] : 0.0) + sx[ph] * f.dEk(c, k, ph);
        }
        const double dMuk[2] = {s.dMuk(c, k, 0), s.dMuk(c, k, 1)};

        for (int q = 0; q < 6; q++) {
            if (!has[q]) continue;
            const FaceGeo& g = G[q];
            const int on = g.o, sg = g.s;
            const RefFace& F = X.F[kc][(size_t)g.a * 3 + g.d];
            const int nlab = s.lab[on];
            const bool zf = g.d == 2;
            for (int ph = 0; ph < 2; ph++) {
                double Coef = ph == 0 ? (nlab == -1 ? 0.0 : 1.0) : (nlab == 1 ? 0.0 : 1.0);
                // rule: in the k != i branch the x and y faces drop the gas
                // term for any single-phase neighbour (:33015 and copies)
                if (!total && !same && ph == 1 && !zf) Coef = nlab ? 0.0 : 1.0;
                const double Surat = s.ro[2 * c + ph] * g.GS + s.ro[2 * on + ph] * g.GN * Coef;
                const double Makhraj = s.mu[2 * c + ph] * g.GS + s.mu[2 * on + ph] * g.GN * Coef;
                const double Mo = (s.dEk(c, k, ph) * g.GS * Makhraj - dMuk[ph] * g.GS * Surat) / (Makhraj * Makhraj);
                const double Bm = (s.dEk(on, k, ph) * g.GN * Coef * Makhraj - s.dMuk(on, k, ph) * g.GN * Coef * Surat) /
                                  (Makhraj * Makhraj);
                const double MG = g.GS * (G_ACCG * s.ro[2 * c + ph] * M.mw[k] + s.gam[2 * c + ph] * s.dEk(c, k, ph) / s.ro[2 * c + ph]) /
                                  (g.GS + g.GN * Coef);
                const double BMG = Coef * g.GN * (G_ACCG * s.ro[2 * on + ph] * M.mw[k] + s.gam[2 * on + ph] * s.dEk(on, k, ph) / s.ro[2 * on + ph]) /
                                   (g.GS + g.GN * Coef);
                const double tempdx = (s.gam[2 * c + ph] * g.GS + s.gam[2 * on + ph] * g.GN * Coef) / (g.GS + g.GN * Coef);
                const bool nbu = nb_up(F, g, ph);
                const int u = nbu ? on : c;
                const double IF = M.iftran[kc][g.d][g.a];
                const double cu = cx(s, u, ph) * IF * s.kr[3 * u + 1 + ph];
                const double dP = s.P[3 * g.b + 1 + ph] - s.P[3 * g.a + 1 + ph];
                const double dh = (sg > 0 ? g.hn - g.hs : g.hs - g.hn);
                const double trP = same ? F.t[ph] : 0.0;        // d(upstream x_i)/dx_i
                const double trG = same ? F.t[2 + ph] : 0.0;
                double dB;
                if (!nbu) {
                    v.dR[ph] += sg * (trP + cu * Mo) * dP;
                    dB = sg * cu * Bm * dP;
                } else {
                    v.dR[ph] += sg * cu * Mo * dP;
                    dB = sg * (trP + cu * Bm) * dP;
                }
                if (zf) {
                    // gravity only on the z faces (the x and y faces of the
                    // composition columns have none)
                    const bool swap = ph == 1 && sg < 0;
                    const double gOwn = swap ? (Bm * tempdx + BMG * Surat / Makhraj) : (Mo * tempdx + MG * Surat / Makhraj);
                    const double gNb = swap ? (Mo * tempdx + MG * Surat / Makhraj) : (Bm * tempdx + BMG * Surat / Makhraj);
                    double sgn = sg;
                    // k != i, z- face, neighbour upstream, oil: both gravity
                    // terms carry the opposite sign (:34020-34021)
                    if (!total && !same && sg < 0 && nbu && ph == 0) sgn = -sgn;
                    if (!nbu) {
                        v.dR[ph] += sgn * (trG + cu * gOwn) * dh;
                        dB += sgn * cu * gNb * dh;
                    } else {
                        // k == i, z+ face, oil: the own-column gravity term
                        // uses the y transmissibility of the cell (:33190)
                        double cuo = cu;
                        if (same && sg > 0 && ph == 0) cuo = cx(s, u, ph) * M.iftran[kc][1][c] * s.kr[3 * u + 1 + ph];
                        v.dR[ph] += sgn * cuo * gOwn * dh;
                        dB += sgn * (trG + cu * gNb) * dh;
                    }
                }
                if (!total) {
                    // Fickian flux, composition derivative (:32303-32318):
                    // DeltaX is taken neighbour - self on both sides, and
                    // the minus faces negate the whole expression.
                    const double dC = dCoef[ph];
                    double Dup, dDdx, dDdx2;
                    if (s.X(on, i, ph) > s.X(c, i, ph)) { Dup = s.Di(on, i, ph); dDdx = 0; dDdx2 = s.dDi(on, i, k, ph); }
                    else { Dup = s.Di(c, i, ph); dDdx = s.dDi(c, i, k, ph); dDdx2 = 0; }
                    const double Sv = s.S[3 * on + 1 + ph], Ss = s.S[3 * c + 1 + ph];
                    double AAd = s.ro[2 * on + ph] * g.GN * Coef + s.ro[2 * c + ph] * g.GS * dC;
                    AAd *= (Sv * g.GN * Coef + Ss * g.GS * dC) / (g.GN + g.GS);
                    const double BBd = ftor(on) * g.GN * Coef / s.porpor[on] + ftor(c) * g.GS * dC / porpor;
                    const double MMd = AAd / BBd;
                    const double NNd = g.area / (g.GN + g.GS);
                    const double dX = s.X(on, i, ph) - s.X(c, i, ph);
                    const double dMMddx = s.dEk(c, k, ph) * g.GS * dC / BBd;
                    const double dMMddx2 = s.dEk(on, k, ph) * g.GN * Coef / BBd;
                    if (Coef != 0 && dC != 0) {
                        const double dd = same ? MMd * Dup : 0.0;
                        v.dR[ph] += sg * NNd * (-dd + dDdx * MMd * dX + dMMddx * Dup * dX);
                        dB += sg * NNd * (dd + dDdx2 * MMd * dX + dMMddx2 * Dup * dX);
                    }
                }
                const int reCol2 = nlab ? nc : 0;
                if (ph == 0) out.add(X.boff[(size_t)kc * n + on] + k, -dB * Coef);
                else out.add(X.boff[(size_t)kc * n + on] - reCol2 + k + nc, -dB * Coef);
            }
        }
        if (olab == 1) out.add(OEq + k, v.dRO[0]);
        else if (olab == -1) out.add(OEq + k, v.dRO[1]);
        else { out.add(OEq + k, v.dRO[0]); out.add(OEq + nc + k, v.dRO[1]); }
        if (lab != -1) out.add(jacCol + k, (Vb / Dt) * porpor * S_s[0] * v.dL[0] - v.dR[0]);
        if (lab != 1) out.add(jacCol + nc + k - reCol, (Vb / Dt) * porpor * S_s[1] * v.dL[1] - v.dR[1]);
    }
}

}  // namespace

namespace {

// Water row of block (kc, c) (matrix :36076-36170, fracture :18523-18621,
// and the copies in the single-phase sections).
void water_row(const Ctx& X, int kc, int c, RowOut& out) {
    const Model& M = X.M;
    const RefCells& s = *X.R[kc];
    const RefCells& m = *X.R[0];
    const RefCells& f = *X.R[1];
    const int nc = X.nc, n = X.n;
    const int lab = s.lab[c], olab = X.R[1 - kc]->lab[c];
    const int reCol = lab == 0 ? 0 : nc, reColF1 = olab == 0 ? 0 : nc;
    const int64_t jacCol = X.boff[(size_t)kc * n + c], OEq = X.boff[(size_t)(1 - kc) * n + c];
    const int iz = c / (M.nx * M.ny), rem = c % (M.nx * M.ny), iy = rem / M.nx, ix = rem % M.nx;
    const double Vb = M.dx[ix] * M.dy[iy] * M.dz[iz];
    const Shape sh = shape(M, c);
    const FM fm = fm_terms(m, f, c);
    out.add(jacCol + (lab == 0 ? 2 * nc : nc), 0.0);   // diagonal placeholder
    const double dLsidedp = (Vb / M.dt) * M.wat_ro * WAT_M_RO * s.S[3 * c] * s.dpor[c];
    const double dLsidedSw = (Vb / M.dt) * (s.porpor[c] * M.wat_ro * WAT_M_RO);
    // rule: transfer pressure terms with the opposite sign of the residual's
    // (in both continua: -c own, +c other)
    const double cW = SI_C * M.wat_ro * WAT_M_RO * sh.sf2 * fm.kr[0] / M.wat_mu * Vb;
    double dRsidedp = -cW;
    const double dRsideOdp = cW;
    FaceGeo G[6];
    bool has[6];
    for (int q = 0; q < 6; q++) has[q] = face_geo(M, c, q, G[q]);
    for (int q = 0; q < 6; q++)
        if (has[q]) dRsidedp += X.F[kc][(size_t)G[q].a * 3 + G[q].d].W;
    out.add(jacCol + 2 * nc - reCol, X.pcoeff * (dLsidedp + dRsidedp));
    if (M.nw) {   // wells: PIw in the P column (matrix :36029, fracture :18474)
        const WellQ<double> wq = well_terms<double>(M, kc, c, ix, iy, iz, X.S->P[kc][c], X.S->P);
        out.add(jacCol + 2 * nc - reCol, X.pcoeff * wq.PIw);
    }
    const double dPw = (f.P[3 * c] - m.P[3 * c]) + (m.pc[2 * c] - f.pc[2 * c]);
    const double cS = SI_C * WAT_M_RO;
    double dRsidedSw, dRsideOdSw;
    if (kc == 0) {
        dRsidedSw = cS * fm.d[1][0] * M.wat_ro * sh.sf2 / M.wat_mu * Vb * dPw;
        dRsideOdSw = cS * fm.d[1][5] * M.wat_ro * sh.sf2 / M.wat_mu * Vb * dPw;
    } else {
        dRsidedSw = -(cS * fm.d[1][5] * M.wat_ro * sh.sf2 / M.wat_mu * Vb * dPw);
        dRsideOdSw = -(cS * fm.d[1][0] * M.wat_ro * sh.sf2 / M.wat_mu * Vb * dPw);
    }
    for (int q = 0; q < 6; q++) {
        if (!has[q]) continue;
        const FaceGeo& g = G[q];
        const RefFace& F = X.F[kc][(size_t)g.a * 3 + g.d];
        const int reCol2 = s.lab[g.o] ? nc : 0;
        const int64_t mBLNo = X.boff[(size_t)kc * n + g.o] - reCol2;
        out.add(mBLNo + 2 * nc, X.pcoeff * (-F.W));
        const double dP = s.P[3 * g.b] - s.P[3 * g.a];
        const double dh = g.s > 0 ? g.hn - g.hs : g.hs - g.hn;
        const double dsw = g.s * (F.dW * dP + G_ACC * M.wat_ro * F.dW * dh);
        if (nb_up(F, g, 2)) out.add(mBLNo + 2 * nc + 1, dsw);
        else { dRsidedSw += dsw; out.add(mBLNo + 2 * nc + 1, 0.0); }
    }
    out.add(jacCol + 2 * nc + 1 - reCol, dLsidedSw - dRsidedSw);
    out.add(OEq + 2 * nc - reColF1, X.pcoeff * dRsideOdp);
    out.add(OEq + 2 * nc + 1 - reColF1, dRsideOdSw);
}

// Fugacity row i of a two-phase block (:30240-30405; fracture :12760-12916):
// x_i Po phiL_i - y_i Pg phiG_i with Z recovered from the stored densities
// and its derivatives from dE.  rule: no saturation columns (the capillary
// pressure inside Pg is held fixed).
void fugacity_row(const Ctx& X, int k, int c, int i, RowOut& out) {
    const Model& M = X.M;
    const RefCells& r = *X.R[k];
    const int nc = X.nc;
    const int64_t jacCol = X.boff[(size_t)k * X.n + c];
    out.add(jacCol + i, 0.0);
    const double T = M.T, Po = r.P[3 * c + 1], Pg = r.P[3 * c + 2];
    double Al = 0, Bl = 0, Ag = 0, Bg = 0;
    for (int a = 0; a < nc; a++) {
        Bl += r.X(c, a, 0) * M.b[a];
        Bg += r.X(c, a, 1) * M.b[a];
        for (int b = 0; b < nc; b++) {
            const double t = std::sqrt(M.a[a] * M.a[b]) * M.bic[a * nc + b];
            Al += r.X(c, a, 0) * r.X(c, b, 0) * t;
            Ag += r.X(c, a, 1) * r.X(c, b, 1) * t;
        }
    }
    const double Cal = Al, Cbl = Bl, Cag = Ag, Cbg = Bg;
    Al *= Po / (RG * RG * T * T); Bl *= Po / (RG * T);
    Ag *= Pg / (RG * RG * T * T); Bg *= Pg / (RG * T);
    const double dAdpl = Cal / (RG * RG * T * T), dAdpg = Cag / (RG * RG * T * T);
    const double dBdpl = Cbl / (RG * T), dBdpg = Cbg / (RG * T);
    const double Zl = Po / (r.ro[2 * c] * RG * T), Zg = Pg / (r.ro[2 * c + 1] * RG * T);
    const double dZdpl = (1 - r.dEk(c, nc, 0) * RG * T * Zl) * (Zl / Po);
    const double dZdpg = (1 - r.dEk(c, nc, 1) * RG * T * Zg) * (Zg / Pg);
    const double d1 = M.d1, d2 = M.d2;
    double tl = 0, tg = 0;
    for (int j = 0; j < nc; j++) {
        const double t = std::sqrt(M.a[i] * M.a[j]) * M.bic[i * nc + j];
        tl += r.X(c, j, 0) * t;
        tg += r.X(c, j, 1) * t;
    }
    const double t1l = M.b[i] / Cbl, t1g = M.b[i] / Cbg;
    const double EEl = 2 * tl / Cal - t1l, EEg = 2 * tg / Cag - t1g;
    const double DDl = Al / (Bl * (d1 - d2)), DDg = Ag / (Bg * (d1 - d2));
    const double FFl = std::log((Zl + d2 * Bl) / (Zl + d1 * Bl)), FFg = std::log((Zg + d2 * Bg) / (Zg + d1 * Bg));
    const double phiL = std::exp(t1l * (Zl - 1) - std::log(Zl - Bl) + DDl * EEl * FFl);
    const double phiG = std::exp(t1g * (Zg - 1) - std::log(Zg - Bg) + DDg * EEg * FFg);
    const double dAAdpl = M.b[i] * dZdpl / Cbl, dAAdpg = M.b[i] * dZdpg / Cbg;
    const double dBBdpl = (dBdpl - dZdpl) / (Zl - Bl), dBBdpg = (dBdpg - dZdpg) / (Zg - Bg);
    const double dDDdpl = (Bl * dAdpl - Al * dBdpl) / (Bl * Bl * (d1 - d2));
    const double dDDdpg = (Bg * dAdpg - Ag * dBdpg) / (Bg * Bg * (d1 - d2));
    const double dFFdpl = (dZdpl + d2 * dBdpl) / (Zl + d2 * Bl) - (dZdpl + d1 * dBdpl) / (Zl + d1 * Bl);
    const double dFFdpg = (dZdpg + d2 * dBdpg) / (Zg + d2 * Bg) - (dZdpg + d1 * dBdpg) / (Zg + d1 * Bg);
    const double dPhidpl = phiL * (dAAdpl + dBBdpl + EEl * (dDDdpl * FFl + DDl * dFFdpl));
    const double dPhidpg = phiG * (dAAdpg + dBBdpg + EEg * (dDDdpg * FFg + DDg * dFFdpg));
    const double xi = r.X(c, i, 0), yi = r.X(c, i, 1);
    out.add(jacCol + 2 * nc, X.pcoeff * (xi * (phiL + Po * dPhidpl) - yi * (phiG + Pg * dPhidpg)));
    for (int kk = 0; kk < nc; kk++) {
        double dAdxl = 0, dAdxg = 0;
        for (int q = 0; q < nc; q++) {
            const double t = std::sqrt(M.a[kk] * M.a[q]) * M.bic[kk * nc + q];
            dAdxl += r.X(c, q, 0) * t;
            dAdxg += r.X(c, q, 1) * t;
        }
        dAdxl *= 2 * Po / (RG * RG * T * T);
        dAdxg *= 2 * Pg / (RG * RG * T * T);
        const double dBdxl = M.b[kk] * Po / (RG * T), dBdxg = M.b[kk] * Pg / (RG * T);
        const double dZdxl = -r.dEk(c, kk, 0) * RG * T * Zl * Zl / Po;
        const double dZdxg = -r.dEk(c, kk, 1) * RG * T * Zg * Zg / Pg;
        const double dAAdxl = -M.b[kk] * M.b[i] * (Zl - 1) / (Cbl * Cbl) + M.b[i] * dZdxl / Cbl;
        const double dAAdxg = -M.b[kk] * M.b[i] * (Zg - 1) / (Cbg * Cbg) + M.b[i] * dZdxg / Cbg;
        const double dBBdxl = (dBdxl - dZdxl) / (Zl - Bl), dBBdxg = (dBdxg - dZdxg) / (Zg - Bg);
        const double dDDdxl = (dAdxl * Bl - dBdxl * Al) / (Bl * Bl * (d1 - d2));
        const double dDDdxg = (dAdxg * Bg - dBdxg * Ag) / (Bg * Bg * (d1 - d2));
        const double dFFdxl = (dZdxl + d2 * dBdxl) / (Zl + d2 * Bl) - (dZdxl + d1 * dBdxl) / (Zl + d1 * Bl);
        const double dFFdxg = (dZdxg + d2 * dBdxg) / (Zg + d2 * Bg) - (dZdxg + d1 * dBdxg) / (Zg + d1 * Bg);
        const double tik = std::sqrt(M.a[i] * M.a[kk]) * M.bic[i * nc + kk];
        const double dEEdxl = 2 * (tik / Cal - dAdxl * RG * RG * T * T * tl / (Po * Cal * Cal)) + M.b[i] * M.b[kk] / (Cbl * Cbl);
        const double dEEdxg = 2 * (tik / Cag - dAdxg * RG * RG * T * T * tg / (Pg * Cag * Cag)) + M.b[i] * M.b[kk] / (Cbg * Cbg);
        const double dPhidxl = phiL * (dAAdxl + dBBdxl + dDDdxl * EEl * FFl + DDl * dEEdxl * FFl + DDl * EEl * dFFdxl);
        const double dPhidxg = phiG * (dAAdxg + dBBdxg + dDDdxg * EEg * FFg + DDg * dEEdxg * FFg + DDg * EEg * dFFdxg);
        if (i == kk) {
            out.add(jacCol + kk, Po * phiL + Po * xi * dPhidxl);
            out.add(jacCol + nc + kk, -(Pg * phiG + Pg * yi * dPhidxg));
        } else {
            out.add(jacCol + kk, Po * xi * dPhidxl);
            out.add(jacCol + nc + kk, -Pg * yi * dPhidxg);
        }
    }
}

// Constraint rows (:36182-36228, :30207-30235, :42005-42029 and the
// fracture copies): sum x = 1, sum y = 1, sum S = 1.
void constraint_row(const Ctx& X, int k, int c, int r, RowOut& out) {
    const int nc = X.nc;
    const int lab = X.R[k]->lab[c];
    const int64_t jacCol = X.boff[(size_t)k * X.n + c];
    out.add(jacCol + r, 0.0);
    if (lab == 0) {
        if (r == 2 * nc + 1) for (int i = 0; i < nc; i++) out.add(jacCol + i, 1.0);
        else if (r == 2 * nc + 2) for (int i = 0; i < nc; i++) out.add(jacCol + nc + i, 1.0);
        else { out.add(jacCol + 2 * nc + 1, 1.0); out.add(jacCol + 2 * nc + 2, 1.0); out.add(jacCol + 2 * nc + 3, 1.0); }
    } else {
        if (r == nc + 1) for (int i = 0; i < nc; i++) out.add(jacCol + i, 1.0);
        else { out.add(jacCol + nc + 1, 1.0); out.add(jacCol + nc + 2, 1.0); }
    }
}

}  // namespace

extern "C" {

// IteratorJacobian in the reference's CSR form: every row lists the columns
// its code emits (duplicates summed in emission order, as MatSetValues with
// ADD_VALUES does on a zeroed matrix), sorted.  Returns the nnz, -1 if cap
// is too small (rowptr still filled), -2 if a row emits a column outside
// the matrix (the reference would fail there).
int64_t or_jacobian_ref_mt(const or_fluid* fl, const or_fluid_ext* fe, const or_grid* g, const or_state* st,
                           int64_t* rowptr, int32_t* cols, double* vals, int64_t cap, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Model M;
    build_model(M, fl, fe, g);
    Sweep S;
    sweep(M, st, S, nthreads);
    const int n = M.n, nc = M.nc;
    RefCells R[2];
    ref_cells(M, S, 0, R[0], nthreads);
    ref_cells(M, S, 1, R[1], nthreads);
    Ctx X(M);
    X.S = &S;
    X.R[0] = &R[0];
    X.R[1] = &R[1];
    X.n = n;
    X.nc = nc;
    X.pcoeff = st->pcoeff;
    X.boff.resize(2 * (size_t)n + 1);
    {
        int64_t off = 0;
        for (int k = 0; k < 2; k++)
            for (int c = 0; c < n; c++) {
                X.boff[(size_t)k * n + c] = off;
                off += block_size(nc, R[k].lab[c]);
            }
        X.boff[2 * (size_t)n] = off;
    }
    const int64_t nrows = X.boff[2 * (size_t)n];
    for (int k = 0; k < 2; k++) {
        X.F[k].resize(3 * (size_t)n);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int c = 0; c < n; c++)
            for (int d = 0; d < 3; d++) {
                FaceGeo gg;
                if (!face_geo(M, c, 2 * d, gg)) continue;
                X.F[k][(size_t)c * 3 + d] = ref_face(M, R[k], k, d, gg.a, gg.b, gg.GS, gg.GN, gg.hs, gg.hn);
            }
    }
    // rows of each block in parallel (a block's rows into its own buffers),
    // then concatenated in row order
    const int64_t nb = 2 * (int64_t)n;
    std::vector<std::vector<int32_t>> bcol(nb);
    std::vector<std::vector<double>> bval(nb);
    std::vector<std::vector<int32_t>> blen(nb);
    bool bad = false;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16) reduction(|| : bad)
    for (int64_t b = 0; b < nb; b++) {
        const int k = (int)(b / n), c = (int)(b % n);
        const int lab = R[k].lab[c];
        const int bs = block_size(nc, lab);
        RowOut row;
        std::vector<int> ord;
        for (int r = 0; r < bs; r++) {
            const int64_t rowno = X.boff[(size_t)k * n + c] + r;
            row.clear();
            if (M.phi[k][c] < 1e-10) {
                row.add(rowno, 1.0);   // zero-porosity block (:42051-42056)
            } else if (lab == 0) {
                if (r < nc) fugacity_row(X, k, c, r, row);
                else if (r < 2 * nc - 1) balance_row(X, k, c, r - nc, false, row);
                else if (r == 2 * nc - 1) balance_row(X, k, c, nc - 1, true, row);
                else if (r == 2 * nc) water_row(X, k, c, row);
                else constraint_row(X, k, c, r, row);
            } else {
                if (r < nc - 1) balance_row(X, k, c, r, false, row);
                else if (r == nc - 1) balance_row(X, k, c, nc - 1, true, row);
                else if (r == nc) water_row(X, k, c, row);
                else constraint_row(X, k, c, r, row);
            }
            ord.resize(row.col.size());
            for (size_t q = 0; q < ord.size(); q++) ord[q] = (int)q;
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b2) { return row.col[a] < row.col[b2]; });
            int32_t len = 0;
            for (size_t q = 0; q < ord.size();) {
                const int64_t col = row.col[ord[q]];
                if (col < 0 || col >= nrows) bad = true;
                double v = 0.0;
                size_t e = q;
                while (e < ord.size() && row.col[ord[e]] == col) v += row.val[ord[e++]];
                bcol[b].push_back((int32_t)col);
                bval[b].push_back(v);
                len++;
                q = e;
            }
            blen[b].push_back(len);
        }
    }
    int64_t nnz = 0;
    bool overflow = false;
    for (int64_t b = 0; b < nb; b++) {
        int64_t rowno = X.boff[b], p = 0;
        for (int32_t len : blen[b]) {
            rowptr[rowno++] = nnz;
            for (int32_t t = 0; t < len; t++, p++, nnz++) {
                if (nnz < cap) { cols[nnz] = bcol[b][p]; vals[nnz] = bval[b][p]; }
                else overflow = true;
            }
        }
    }
    rowptr[nrows] = nnz;
    if (bad) return -2;
    return overflow ? -1 : nnz;
}

int64_t or_jacobian_ref(const or_fluid* fl, const or_fluid_ext* fe, const or_grid* g, const or_state* st,
                        int64_t* rowptr, int32_t* cols, double* vals, int64_t cap) {
    return or_jacobian_ref_mt(fl, fe, g, st, rowptr, cols, vals, cap, 1);
}

}  // extern "C"

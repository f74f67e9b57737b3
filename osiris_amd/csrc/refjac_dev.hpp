// refjac_dev.hpp -- the reference's hand-written Jacobian (IteratorJacobian,
// compsim/Petsc_Functions.cpp:6898-42073) on gfx950: its column pattern and
// its values, including the places where its hand rules differ from the
// exact derivative of IteratorFunction.  Included by assembly.hip inside
// namespace osr (uses AsmArgs, the property table and the sweep helpers).
//
// The reference builds each row from per-cell quantities its property sweep
// leaves in globals: relPerm/dRelPerm (CalcSatFs Do_Cycle.cpp:99-257, with
// the sign clamps of :254-255 and the IFT rules of :241-249), dE (Calc_Visco
// :438-616), DiffVisco (Mass_Tran.cpp:307-399), DifDiffusion (:107-305,
// i.e. 1e4 times the derivative of CalcDiffusion's coefficient), and the
// face coefficients of Calc_Transes / fCalc_Transes (Do_Cycle.cpp:800-1417).
//
// GPU structure:
//   k_rj_cells   one lane per (block, direction q = x_0..x_{nc-1}, P): the
//                density, viscosity and diffusion coefficients of both phases
//                in forward mode along q (DualV<1>), written to a
//                [field][2n] table next to the property table; the q = 0
//                lane also writes dRelPerm and d porosity / dP;
//   k_rj_rownnz  one lane per block: the length of each of its rows;
//   k_rj_rows    one lane per (block, row): the row's entries.  A row's
//                columns are known from the labels alone (own block, six
//                neighbours and the other continuum's block in column order,
//                each either whole, whole but the own Sw, or {P, Sw}), so the
//                lane writes the column indices first and then adds every
//                term straight into its entry; it owns the row, so no atomics
//                are needed and the sum order is the emission order.
#pragma once

namespace osr {

template <int NC>
struct RJ {
    static constexpr int NQ = NC + 1;
    static constexpr int DKR = 0;                  // dRelPerm[0..4]
    static constexpr int DPOR = 5;                 // d porpor / d P
    static constexpr int DE = 6;                   // [q][ph]
    static constexpr int DMU = DE + 2 * NQ;        // [q][ph]
    static constexpr int DDIF = DMU + 2 * NQ;      // [i][q][ph]
    static constexpr int NF = DDIF + 2 * NC * NQ;
};

// Read access to the per-block quantities (all [field][2n], block = k n + c).
template <int NC>
struct RjView {
    const AsmArgs& A;
    const double* t;
    int64_t s2;
    int n;
    __device__ RjView(const AsmArgs& a, const double* tab) : A(a), t(tab), s2(2 * (int64_t)a.n), n(a.n) {}
    __device__ double pr(int f, int b) const { return A.props[f * s2 + b]; }
    __device__ double porpor(int b) const { return pr(0, b); }
    __device__ double P(int b, int j) const { return pr(1 + j, b); }     // 0 water, 1 oil, 2 gas
    __device__ double ro(int b, int ph) const { return pr(5 + ph, b); }
    __device__ double mu(int b, int ph) const { return pr(7 + ph, b); }
    __device__ double gam(int b, int ph) const { return pr(9 + ph, b); }
    __device__ double kr(int b, int j) const { return pr(11 + j, b); }   // 0 water, 1 oil, 2 gas
    __device__ double pc(int b, int j) const { return pr(14 + j, b); }   // 0 pcw, 1 pcg
    __device__ double Di(int b, int i, int ph) const { return pr(16 + ph * NC + i, b); }
    __device__ double dkr(int b, int j) const { return t[(RJ<NC>::DKR + j) * s2 + b]; }
    __device__ double dpor(int b) const { return t[RJ<NC>::DPOR * s2 + b]; }
    __device__ double dE(int b, int q, int ph) const { return t[(RJ<NC>::DE + 2 * q + ph) * s2 + b]; }
    __device__ double dMu(int b, int q, int ph) const { return t[(RJ<NC>::DMU + 2 * q + ph) * s2 + b]; }
    __device__ double dDi(int b, int i, int q, int ph) const {
        return t[(RJ<NC>::DDIF + (i * (NC + 1) + q) * 2 + ph) * s2 + b];
    }
    __device__ int lab(int b) const { return b < n ? A.lab[0][b] : A.lab[1][b - n]; }
    __device__ double X(int b, int i, int ph) const {
        const int k = b < n ? 0 : 1, c = b - k * n;
        return ph == 0 ? A.x[k][(int64_t)i * n + c] : A.y[k][(int64_t)i * n + c];
    }
    __device__ double S(int b, int j) const {
        const int k = b < n ? 0 : 1, c = b - k * n;
        return A.S[k][(int64_t)j * n + c];
    }
    __device__ double phi(int b) const { return A.phi[b]; }
};

// dRelPerm of the matrix (CalcSatFs, Do_Cycle.cpp:163-256): table slopes, the
// Stone II derivatives with respect to Sw and Sg, the IFT rule (F frozen),
// then the sign clamps of :254-255.
__device__ inline void rj_seg(const double* tb, int nr, double s, double& dkr, double& dkro, double& kr,
                              double& kro) {
    int r = 0;
    while (r < nr && !(s < tb[4 * r])) r++;
    if (r == 0) {
        kr = RELPERM0; kro = tb[2]; dkr = 0; dkro = 0;
    } else if (r == nr) {
        kr = tb[4 * (nr - 1) + 1]; kro = tb[4 * (nr - 1) + 2]; dkr = 0; dkro = 0;
    } else {
        const double ds = tb[4 * r] - tb[4 * (r - 1)];
        kr = (tb[4 * r + 1] - tb[4 * (r - 1) + 1]) / ds * (s - tb[4 * r]) + tb[4 * r + 1];
        kro = (tb[4 * r + 2] - tb[4 * (r - 1) + 2]) / ds * (s - tb[4 * r]) + tb[4 * r + 2];
        dkr = (tb[4 * r + 1] - tb[4 * (r - 1) + 1]) / ds;
        dkro = (tb[4 * r + 2] - tb[4 * (r - 1) + 2]) / ds;
    }
}

__device__ inline void rj_matrix_dkr(const AsmConsts& K, double sw, double sg, double bift, double* d) {
    double dkrw, sdKrow, krw, Krow, dkrg, sdKrog, krg, Krog;
    rj_seg(K.swt, K.nswt, sw, dkrw, sdKrow, krw, Krow);
    rj_seg(K.sgt, K.nsgt, sg, dkrg, sdKrog, krg, Krog);
    const double kro0 = K.swt[2];
    const double AA = krw + Krow / kro0, BB = krg + Krog / kro0;
    const double dAA = dkrw + sdKrow / kro0, dBB = dkrg + sdKrog / kro0;
    d[0] = dkrw;
    d[3] = dkrg;
    d[4] = 0;
    d[1] = kro0 * (dAA * BB - dkrw);
    d[2] = kro0 * (AA * dBB - dkrg);
    if (bift < MISCIBLEIFT) {
        const double F = pow(bift / IFTBASECASE, 1.0 / IFTPOWER);
        d[3] = F * d[3] + (1 - F) * (1 - K.swt[0]);
        d[1] *= F;
        d[2] *= F;
        d[4] = (1 - F) * (1 - K.swt[0]);
    }
    if (d[1] > 0) d[1] = 0;
    if (d[2] > 0) d[2] = 0;
}

template <int NC>
__global__ __launch_bounds__(256) void k_rj_cells(AsmArgs A, double* t) {
    const int64_t nb = 2 * (int64_t)A.n;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb * (NC + 1)) return;
    const int q = (int)(idx / nb);
    const int blk = (int)(idx - q * nb);
    const int k = blk / A.n, c = blk - k * A.n;
    const AsmConsts& K = *A.K;
    VarsT<NC, double> v;
    load_vars<NC>(A, k, c, v);
    double zl, zg;
    table_roots(A, blk, zl, zg);
    const double* pr = A.props + blk;
    using D1 = DualV<1>;
    for (int ph = 0; ph < 2; ph++) {
        D1 xd[NC];
#pragma unroll
        for (int i = 0; i < NC; i++) {
            xd[i] = D1(ph == 0 ? v.x[i] : v.y[i]);
            xd[i].d[0] = (i == q) ? 1.0 : 0.0;
        }
        D1 Pd(pr[(ph == 0 ? 2 : 3) * nb]);
        Pd.d[0] = (q == NC) ? 1.0 : 0.0;
        D1 ro, mu, Dd[NC];
        viscoT<NC, D1>(K, xd, Pd, ph == 1, ro, mu, ph == 0 ? zl : zg);
        diffT<NC, D1>(K, xd, ro, Dd);
        t[(RJ<NC>::DE + 2 * q + ph) * nb + blk] = ro.d[0];
        t[(RJ<NC>::DMU + 2 * q + ph) * nb + blk] = mu.d[0];
#pragma unroll
        for (int i = 0; i < NC; i++) {
            // DifDiffusion has no 1e-4 factor; a pure phase gives 0
            const bool pure = (ph == 0 ? v.x[i] : v.y[i]) == 1.0;
            t[(RJ<NC>::DDIF + (i * (NC + 1) + q) * 2 + ph) * nb + blk] = pure ? 0.0 : 10000.0 * Dd[i].d[0];
        }
    }
    if (q == 0) {
        double d[5];
        if (k == 0) {
            rj_matrix_dkr(K, v.S[0], v.S[2], pr[4 * nb], d);
        } else {
            // fCalcSatFs (Do_Cycle.cpp:426-435): kr = S, fdRelPerm = {1, 0, 0, 1, 1}
            d[0] = 1; d[1] = 0; d[2] = 0; d[3] = 1; d[4] = 1;
        }
#pragma unroll
        for (int j = 0; j < 5; j++) t[(RJ<NC>::DKR + j) * nb + blk] = d[j];
        t[RJ<NC>::DPOR * nb + blk] = A.phi[blk] * (K.cpor + 2 * K.dcpor * (v.Po - K.refP));
    }
}

// Calc_Transes / fCalc_Transes for the face between cells a and b = a + e_d
// of continuum k.
struct RjFace {
    double t[12];   // trans[..][0..11]
    int up[3];      // transS: 0 a upstream, 1 b upstream (oil, gas, water)
    double W, dW;   // Wtran, dWtran
};

template <int NC>
__device__ inline RjFace rj_face(const RjView<NC>& V, int k, int d, int a, int b, double da, double db, double ha,
                                 double hb) {
    const AsmConsts& K = *V.A.K;
    const int n = V.n;
    const int ba = k * n + a, bb = k * n + b;
    RjFace f;
    const double Kf = V.A.iftran[((int64_t)k * 3 + d) * n + a];
    const double wc = K.wat_ro * WAT_M_RO / K.wat_mu;
    {
        const double pot = V.P(ba, 0) - V.P(bb, 0) + G_ACC * K.wat_ro * (ha - hb);
        const int u = pot > 0 ? ba : bb;
        f.up[2] = pot > 0 ? 0 : 1;
        f.W = Kf * wc * V.kr(u, 0);
        f.dW = k == 0 ? Kf * wc * V.dkr(u, 0) : Kf * wc;
    }
    for (int ph = 0; ph < 2; ph++) {
        const int absent = ph == 0 ? -1 : 1;
        const double c1 = V.lab(ba) == absent ? 0 : 1, c2 = V.lab(bb) == absent ? 0 : 1;
        double mg = (V.gam(ba, ph) * da * c1 + V.gam(bb, ph) * db * c2) / (c1 * da + db * c2);
        if (c1 == 0 && c2 == 0) mg = 0;
        const double pot = V.P(ba, 1 + ph) - V.P(bb, 1 + ph) + mg * (ha - hb);
        const int u = pot > 0 ? ba : bb;
        f.up[ph] = pot > 0 ? 0 : 1;
        const double T = Kf * (V.ro(ba, ph) * da * c1 + V.ro(bb, ph) * db * c2) /
                         (V.mu(ba, ph) * da * c1 + V.mu(bb, ph) * db * c2);
        const double kru = V.kr(u, 1 + ph);
        if (ph == 0) {
            f.t[0] = T * kru;
            f.t[2] = mg * T * kru;
            if (k == 0) {
                const double d1 = V.dkr(u, 1), d2 = V.dkr(u, 2), d4 = V.dkr(u, 4);
                f.t[4] = T * d1; f.t[6] = mg * T * d1;
                f.t[8] = T * d2; f.t[9] = mg * T * d2;
                f.t[10] = T * d4; f.t[11] = mg * T * d4;
            } else {
                // fCalc_Transes: d kro / d Sw = d kro / d Sg = -1, d kro / d So = 1
                f.t[4] = -T; f.t[6] = -mg * T;
                f.t[8] = -T; f.t[9] = -mg * T;
                f.t[10] = T; f.t[11] = mg * T;
            }
            if (c1 == 0 && c2 == 0) f.t[0] = f.t[2] = f.t[4] = f.t[6] = f.t[8] = f.t[9] = 0;
        } else {
            f.t[1] = T * kru;
            f.t[3] = mg * T * kru;
            if (k == 0) {
                const double d3 = V.dkr(u, 3);
                f.t[5] = T * d3; f.t[7] = mg * T * d3;
            } else {
                f.t[5] = T; f.t[7] = mg * T;
            }
            if (c1 == 0 && c2 == 0) f.t[1] = f.t[3] = f.t[5] = f.t[7] = 0;
        }
    }
    return f;
}

// One face of cell c seen from c: q = 0..5 is x+, x-, y+, y-, z+, z-.
struct RjGeo {
    int d, s;        // direction, +1 / -1
    int o;           // neighbour cell
    int a, b;        // lower and upper cell of the face
    double GS, GN;   // cell sizes along d: self, neighbour
    double hs, hn;   // block heights
    double area;     // 2 x the face's transverse cell dimensions
};

__device__ inline bool rj_geo(const AsmArgs& A, const Geo& G, int q, RjGeo& g) {
    g.d = q >> 1;
    g.s = (q & 1) ? -1 : 1;
    const int idx = g.d == 0 ? G.ix : (g.d == 1 ? G.iy : G.iz);
    const int dim = g.d == 0 ? A.nx : (g.d == 1 ? A.ny : A.nz);
    const int j = idx + g.s;
    if (j < 0 || j >= dim) return false;
    const int stride = g.d == 0 ? 1 : (g.d == 1 ? A.nx : A.nx * A.ny);
    g.o = G.c + g.s * stride;
    g.a = g.s > 0 ? G.c : g.o;
    g.b = g.s > 0 ? g.o : G.c;
    const double* sz = g.d == 0 ? A.dx : (g.d == 1 ? A.dy : A.dz);
    g.GS = sz[idx];
    g.GN = sz[j];
    g.hs = A.h[G.iz];
    g.hn = g.d == 2 ? A.h[j] : A.h[G.iz];
    g.area = g.d == 0 ? 2 * A.dy[G.iy] * A.dz[G.iz] : (g.d == 1 ? 2 * A.dx[G.ix] * A.dz[G.iz] : 2 * A.dx[G.ix] * A.dy[G.iy]);
    return true;
}

template <int NC>
__device__ inline RjFace rj_face_of(const RjView<NC>& V, int k, const RjGeo& g) {
    // the face is stored from its lower cell: sizes and heights of (a, b)
    const double da = g.s > 0 ? g.GS : g.GN, db = g.s > 0 ? g.GN : g.GS;
    const double ha = g.s > 0 ? g.hs : g.hn, hb = g.s > 0 ? g.hn : g.hs;
    return rj_face<NC>(V, k, g.d, g.a, g.b, da, db, ha, hb);
}

// true if the neighbour is upstream for phase p (0 oil, 1 gas, 2 water)
__device__ inline bool rj_nb_up(const RjFace& f, const RjGeo& g, int p) {
    return g.s > 0 ? f.up[p] == 1 : f.up[p] == 0;
}

// ------------------------------------------------------------- row layout
enum { RJ_FULL = 0, RJ_NOSW = 1, RJ_PSW = 2 };

// The column blocks of one row in column order; each is the whole block, the
// whole block without its Sw (a gas block's own columns in its balance
// rows) or {P, Sw} (water rows).
struct RjRow {
    int ns;
    int64_t base[8];
    int bsz[8], kind[8], pidx[8];
    int64_t start[8];
    int len;
};

__device__ inline int rj_count(int kind, int bsz) { return kind == RJ_FULL ? bsz : (kind == RJ_NOSW ? bsz - 1 : 2); }

// slots of a balance (water = false) or water row of block (k, c)
__device__ inline void rj_slots(const AsmArgs& A, const Geo& G, int k, bool water, RjRow& R) {
    const int n = A.n, nc = A.K->nc;
    const int c = G.c;
    int blks[8];
    int m = 0;
    const int own = k * n + c, other = (1 - k) * n + c;
    if (k == 1) blks[m++] = other;
    const int qs[7] = {5, 3, 1, -1, 0, 2, 4};   // z-, y-, x-, own, x+, y+, z+
    for (int j = 0; j < 7; j++) {
        if (qs[j] < 0) { blks[m++] = own; continue; }
        const int o = nb_of(A, G, qs[j]);
        if (o >= 0) blks[m++] = k * n + o;
    }
    if (k == 0) blks[m++] = other;
    R.ns = m;
    int64_t pos = 0;
    for (int j = 0; j < m; j++) {
        const int b = blks[j];
        const int lb = b < n ? A.lab[0][b] : A.lab[1][b - n];
        R.base[j] = A.boff[b];
        R.bsz[j] = lb == 0 ? 2 * nc + 4 : nc + 3;
        R.pidx[j] = lb == 0 ? 2 * nc : nc;
        R.kind[j] = water ? RJ_PSW : ((b == own && lb == -1) ? RJ_NOSW : RJ_FULL);
        R.start[j] = pos;
        pos += rj_count(R.kind[j], R.bsz[j]);
    }
    R.len = (int)pos;
}

// Emission of (column, value) into a row whose columns are laid out by RjRow.
struct RjEmit {
    const RjRow& R;
    double* v;   // the row's values
    __device__ RjEmit(const RjRow& r, double* vals) : R(r), v(vals) {}
    __device__ void add(int64_t col, double x) const {
        for (int j = 0; j < R.ns; j++) {
            if (col >= R.base[j] && col < R.base[j] + R.bsz[j]) {
                const int l = (int)(col - R.base[j]);
                int idx = l;
                if (R.kind[j] == RJ_NOSW) {
                    if (l == R.pidx[j] + 1) return;   // not in the pattern
                    idx = l - (l > R.pidx[j] + 1 ? 1 : 0);
                } else if (R.kind[j] == RJ_PSW) {
                    idx = l - R.pidx[j];
                    if (idx < 0 || idx > 1) return;   // not in the pattern
                }
                v[R.start[j] + idx] += x;
                return;
            }
        }
    }
};

__device__ inline void rj_fill_cols(const RjRow& R, int32_t* cols, double* vals) {
    for (int j = 0; j < R.ns; j++) {
        int64_t p = R.start[j];
        if (R.kind[j] == RJ_PSW) {
            cols[p] = (int32_t)(R.base[j] + R.pidx[j]);
            cols[p + 1] = (int32_t)(R.base[j] + R.pidx[j] + 1);
            vals[p] = vals[p + 1] = 0.0;
            continue;
        }
        for (int l = 0; l < R.bsz[j]; l++) {
            if (R.kind[j] == RJ_NOSW && l == R.pidx[j] + 1) continue;
            cols[p] = (int32_t)(R.base[j] + l);
            vals[p] = 0.0;
            p++;
        }
    }
}

// ------------------------------------------------------------ row kernels
// Shape factors (Init_Calcs.cpp:970-983).
struct RjShape {
    double sf1, sf2, sf3;
};
__device__ inline RjShape rj_shape(const AsmArgs& A, const Geo& G) {
    const AsmConsts& K = *A.K;
    const int n = A.n, c = G.c;
    RjShape s;
    s.sf2 = 4 * K.sigma * (A.perm_m[c] / (G.ds0 * G.ds0) + A.perm_m[n + c] / (G.ds1 * G.ds1) +
                           A.perm_m[2 * n + c] / (G.ds2 * G.ds2));
    const bool off = K.const_diff && K.cdiff == 0;
    s.sf1 = off ? 0 : 4 * K.sigma * (1 / (G.ds0 * G.ds0) + 1 / (G.ds1 * G.ds1) + 1 / (G.ds2 * G.ds2));
    s.sf3 = off ? 0 : 4 * K.sigma * ((1 / G.ds0) + (1 / G.ds1) + (1 / G.ds2)) / K.lf;
    return s;
}

// f_m_relperm / f_m_drelperm (e.g. Petsc_Functions.cpp:30409-30466): the
// upstream kr between the two continua by phase pressure.
struct RjFM {
    double kr[3];       // water, oil, gas
    double d[3][10];    // f_m_drelperm[ph][..]
};
template <int NC>
__device__ inline RjFM rj_fm(const RjView<NC>& V, int c) {
    const int bm = c, bf = V.n + c;
    RjFM r;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 10; b++) r.d[a][b] = 0;
    if (V.P(bm, 1) >= V.P(bf, 1)) {
        r.kr[1] = V.kr(bm, 1);
        r.d[1][1] = V.dkr(bm, 1); r.d[1][2] = V.dkr(bm, 2); r.d[1][4] = V.dkr(bm, 4);
    } else {
        r.kr[1] = V.kr(bf, 1);
        r.d[1][6] = V.dkr(bf, 1); r.d[1][7] = V.dkr(bf, 2); r.d[1][9] = V.dkr(bf, 4);
    }
    if (V.P(bm, 2) >= V.P(bf, 2)) { r.kr[2] = V.kr(bm, 2); r.d[2][3] = V.dkr(bm, 3); }
    else { r.kr[2] = V.kr(bf, 2); r.d[2][8] = V.dkr(bf, 3); }
    if (V.P(bm, 0) >= V.P(bf, 0)) { r.kr[0] = V.kr(bm, 0); r.d[1][0] = V.dkr(bm, 0); }
    else { r.kr[0] = V.kr(bf, 0); r.d[1][5] = V.dkr(bf, 0); }
    return r;
}

// Phase terms of the diffusive transfer by (matrix label, fracture label).
// Matrix rows (:5259-5313): (phase, uses the fracture composition 'F' or
// EdgeComp = 0 'E').
__device__ inline int rj_mdiff(int lm, int lf, int* ph, char* kind) {
    int n = 0;
    auto put = [&](int p, char k) { ph[n] = p; kind[n] = k; n++; };
    if (lm == 1 && lf == 1) put(0, 'F');
    if (lm == -1 && lf == -1) put(1, 'F');
    if (lm == 0 && lf == 0) { put(0, 'F'); put(1, 'F'); }
    if (lm == 1 && lf == -1) put(0, 'E');
    if (lm == -1 && lf == 1) put(1, 'E');
    if (lm == 0 && lf == 1) { put(1, 'E'); put(0, 'F'); }
    if (lm == 0 && lf == -1) { put(1, 'F'); put(0, 'E'); }
    if (lm == 1 && lf == 0) { put(0, 'E'); put(0, 'F'); }
    if (lm == -1 && lf == 0) { put(1, 'F'); put(1, 'E'); }
    return n;
}
// Fracture rows (:731-806): matrix-property terms with shape factor 1 ...
__device__ inline int rj_fsf1(int lm, int lf, int* ph) {
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
// ... and fracture-property terms with shape factor 3 (x_f - EdgeComp).
__device__ inline int rj_fsf3(int lm, int lf, int* ph) {
    int n = 0;
    if (lm == 1 && lf == -1) ph[n++] = 1;
    if (lm == -1 && lf == 1) ph[n++] = 0;
    if (lm == 0 && lf == 1) ph[n++] = 0;
    if (lm == 1 && lf == 0) ph[n++] = 1;
    if (lm == 0 && lf == -1) ph[n++] = 1;
    if (lm == -1 && lf == 0) ph[n++] = 0;
    return n;
}

// One component-balance row (total = false, component i) or the total
// hydrocarbon row (total = true) of block (kc, c): matrix component rows
// Petsc_Functions.cpp:31012-34627 (two-phase section; the liquid and gas
// sections :25028-28651, :36841-40456 run the same code), matrix total rows
// :34627-36182, fracture rows :13505-18623 (face code as the matrix's with
// tortuosity 1).  Value conventions kept from the reference: x and S columns
// carry dLSide - dRSide; the own P column dLSide + dRsidedp, where
// dRsidedp collects -dRSide/dP of the face fluxes and +dRSide/dP of the
// transfer; neighbour S columns and every column of the other continuum
// carry +dRSide/d(unknown).
template <int NC>
__device__ void rj_balance(const RjView<NC>& s, const Geo& G, int kc, int i, bool total, double pcoeff,
                           const RjEmit& out) {
    const AsmArgs& A = s.A;
    const AsmConsts& K = *A.K;
    constexpr int nc = NC;
    const int n = s.n, c = G.c;
    const int bs_ = kc * n + c, bo = (1 - kc) * n + c, bm = c, bf = n + c;
    const int lab = s.lab(bs_), olab = s.lab(bo);
    const int lm = s.lab(bm), lf = s.lab(bf);
    const int reCol = lab == 0 ? 0 : nc;
    const int reColF1 = olab == 0 ? 0 : nc;
    const int64_t jacCol = A.boff[bs_], OEq = A.boff[bo];
    const double Vb = G.Vb, Dt = K.dt;
    const RjShape sh = rj_shape(A, G);
    const RjFM fm = rj_fm<NC>(s, c);
    const double tor = A.tor[c];
    const double porpor = s.porpor(bs_), dPordp = s.dpor(bs_);
    const double mpor = s.porpor(bm), fpor = s.porpor(bf), mdpor = s.dpor(bm), fdpor = s.dpor(bf);
    const double MUP[2] = {s.mu(bm, 0) * mpor + s.mu(bf, 0) * fpor, s.mu(bm, 1) * mpor + s.mu(bf, 1) * fpor};
    const double PP = mpor + fpor;
    const double DPh[2] = {s.P(bf, 1) - s.P(bm, 1), (s.P(bf, 2) - s.P(bm, 2)) + (s.pc(bf, 1) - s.pc(bm, 1))};
    const double cPh[2] = {sh.sf2 * Vb * fm.kr[1], sh.sf2 * Vb * fm.kr[2]};
    auto cx = [&](int b, int ph) { return total ? 1.0 : s.X(b, i, ph); };
    auto ftor = [&](int cell) { return kc == 0 ? A.tor[cell] : 1.0; };
    const double sx[2] = {cx(bs_, 0), cx(bs_, 1)};
    const double mx[2] = {cx(bm, 0), cx(bm, 1)};
    const double ro_s[2] = {s.ro(bs_, 0), s.ro(bs_, 1)}, ro_m[2] = {s.ro(bm, 0), s.ro(bm, 1)};
    const double S_s[2] = {s.S(bs_, 1), s.S(bs_, 2)};

    out.add(jacCol + (lab == 0 ? (total ? 2 * nc - 1 : nc + i) : (total ? nc - 1 : i)), 0.0);   // placeholder

    // ------------------------------------------------------ own-cell pieces
    double dLp = (Vb / Dt) * (sx[0] * S_s[0] * (ro_s[0] * dPordp + s.dE(bs_, nc, 0) * porpor) +
                              sx[1] * S_s[1] * (ro_s[1] * dPordp + s.dE(bs_, nc, 1) * porpor));
    const double dLSo = (Vb / Dt) * (sx[0] * porpor * ro_s[0]);
    const double dLSg = (Vb / Dt) * (sx[1] * porpor * ro_s[1]);
    double dRp = 0, dROp = 0, dRSo = 0, dRSg = 0, dRSw = 0, dROSo = 0, dROSg = 0, dROSw = 0;
    if (kc == 0) {
        for (int ph = 0; ph < 2; ph++) {
            dRp += SI_C * s.dE(bm, nc, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
            dRp += -SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            dRp += -SI_C * (s.dMu(bm, nc, ph) * mpor + mdpor * s.mu(bm, ph)) * mx[ph] * ro_m[ph] * cPh[ph] /
                   (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            dRp += SI_C * mdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        if (!total) {
            int tph[4];
            char tk[4];
            const int nt = rj_mdiff(lm, lf, tph, tk);
            for (int t = 0; t < nt; t++) {
                const int ph = tph[t];
                const double dl = tk[t] == 'F' ? s.X(bf, i, ph) - s.X(bm, i, ph) : 0.0 - s.X(bm, i, ph);
                dRp += mdpor * sh.sf1 * S_s[ph] * Vb / tor * ro_s[ph] * s.Di(bm, i, ph) * dl;
                dRp += s.dE(bm, nc, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * s.Di(bm, i, ph) * dl;
                dRp += s.dDi(bm, i, nc, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_s[ph] * dl;
                if (ph == 0) dRSo += sh.sf1 * Vb * mpor / tor * ro_s[0] * s.Di(bm, i, 0) * dl;
                // the (matrix two-phase, fracture liquid) combination has no gas Sg term
                if (ph == 1 && !(lm == 0 && lf == 1)) dRSg += sh.sf1 * Vb * mpor / tor * ro_s[1] * s.Di(bm, i, 1) * dl;
            }
        }
        for (int ph = 0; ph < 2; ph++) {
            dROp += SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            dROp += -SI_C * (s.dMu(bf, nc, ph) * fpor + fdpor * s.mu(bf, ph)) * mx[ph] * ro_m[ph] * cPh[ph] /
                    (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            dROp += SI_C * fdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        if (!total && lm == 1 && lf == 0) {
            // liquid-section component rows only: the fracture gas-phase
            // diffusion term differentiated by the fracture pressure
            const double dl = s.X(bf, i, 1);
            dROp += fdpor * sh.sf1 * s.S(bf, 2) * Vb / tor * s.ro(bf, 1) * s.Di(bf, i, 1) * dl;
            dROp += s.dE(bf, nc, 1) * sh.sf1 * s.S(bf, 2) * fpor * Vb / tor * s.Di(bf, i, 1) * dl;
            dROp += s.dDi(bf, i, nc, 1) * sh.sf1 * s.S(bf, 2) * fpor * Vb / tor * s.ro(bf, 1) * dl;
        }
        const double c0 = SI_C * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        const double c1 = SI_C * mx[1] * ro_m[1] * (sh.sf2 * Vb) / (MUP[1] / PP) * DPh[1];
        dRSo += fm.d[1][4] * c0;
        dROSo += fm.d[1][9] * c0;
        dRSg += fm.d[1][2] * c0;
        dRSg += fm.d[2][3] * c1;
        dROSg += fm.d[1][7] * c0;
        dROSg += fm.d[2][8] * c1;
        dRSw += fm.d[1][1] * c0;
        dROSw += fm.d[1][6] * c0;
    } else {
        for (int ph = 0; ph < 2; ph++) {
            dROp -= SI_C * s.dE(bm, nc, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
            dROp -= -SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            dROp -= -SI_C * (s.dMu(bm, nc, ph) * mpor + mdpor * s.mu(bm, ph)) * mx[ph] * ro_m[ph] * cPh[ph] /
                    (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            dROp -= SI_C * mdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        int p1[4], p3[4];
        const int n1 = total ? 0 : rj_fsf1(lm, lf, p1);
        const int n3 = total ? 0 : rj_fsf3(lm, lf, p3);
        for (int t = 0; t < n1; t++) {
            const int ph = p1[t];
            const double dl = s.X(bf, i, ph) - s.X(bm, i, ph);
            const double Sm = s.S(bm, 1 + ph);
            dROp -= mdpor * sh.sf1 * Sm * Vb / tor * ro_m[ph] * s.Di(bm, i, ph) * dl;
            dROp -= s.dE(bm, nc, ph) * sh.sf1 * Sm * mpor * Vb / tor * s.Di(bm, i, ph) * dl;
            dROp -= s.dDi(bm, i, nc, ph) * sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * dl;
            if (ph == 0) dROSo -= sh.sf1 * Vb * mpor / tor * ro_m[0] * s.Di(bm, i, 0) * dl;
            else dROSg -= sh.sf1 * Vb * mpor / tor * ro_m[1] * s.Di(bm, i, 1) * dl;
        }
        for (int ph = 0; ph < 2; ph++) {
            dRp -= SI_C * mx[ph] * ro_m[ph] * cPh[ph] / (MUP[ph] / PP);
            dRp -= -SI_C * (s.dMu(bf, nc, ph) * fpor + fdpor * s.mu(bf, ph)) * mx[ph] * ro_m[ph] * cPh[ph] /
                   (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            dRp -= SI_C * fdpor * mx[ph] * ro_m[ph] * cPh[ph] / MUP[ph] * DPh[ph];
        }
        for (int t = 0; t < n3; t++) {
            const int ph = p3[t];
            const double dl = s.X(bf, i, ph);
            dRp -= fdpor * sh.sf3 * S_s[ph] * Vb / tor * ro_s[ph] * s.Di(bf, i, ph) * dl;
            dRp -= s.dE(bf, nc, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * s.Di(bf, i, ph) * dl;
            dRp -= s.dDi(bf, i, nc, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * ro_s[ph] * dl;
            if (ph == 0) dRSo -= sh.sf3 * Vb * fpor / tor * ro_s[0] * s.Di(bf, i, 0) * dl;
            else dRSg -= sh.sf3 * Vb * fpor / tor * ro_s[1] * s.Di(bf, i, 1) * dl;
        }
        const double c0 = SI_C * mx[0] * ro_m[0] * (sh.sf2 * Vb) / (MUP[0] / PP) * DPh[0];
        const double c1 = SI_C * mx[1] * ro_m[1] * (sh.sf2 * Vb) / (MUP[1] / PP) * DPh[1];
        dROSo -= fm.d[1][4] * c0;
        dRSo -= fm.d[1][9] * c0;
        dROSg -= fm.d[1][2] * c0;
        dROSg -= fm.d[2][3] * c1;
        dRSg -= fm.d[1][7] * c0;
        dRSg -= fm.d[2][8] * c1;
        dROSw -= fm.d[1][1] * c0;
        dRSw -= fm.d[1][6] * c0;
    }

    // advective fluxes: d(P_nb - P)/dP of every face, upstream composition
    RjGeo Gq[6];
    bool has[6];
    RjFace Fq[6];
    for (int q = 0; q < 6; q++) {
        has[q] = rj_geo(A, G, q, Gq[q]);
        if (has[q]) Fq[q] = rj_face_of<NC>(s, kc, Gq[q]);
    }
    double dRsidedp = dRp;
    for (int ph = 0; ph < 2; ph++)
        for (int q = 0; q < 6; q++) {
            if (!has[q]) continue;
            const int up = Fq[q].up[ph] == 0 ? Gq[q].a : Gq[q].b;
            dRsidedp += cx(kc * n + up, ph) * Fq[q].t[ph];
        }
    const double dCoef[2] = {lab == -1 ? 0.0 : 1.0, lab == 1 ? 0.0 : 1.0};
    double dRsidedSw = dRSw, dRsidedSg = dRSg, dRsidedSo = dRSo;

    // ------------------------------------------------- faces: P and S columns
    for (int q = 0; q < 6; q++) {
        if (!has[q]) continue;
        const RjGeo& g = Gq[q];
        const RjFace& F = Fq[q];
        const int on = kc * n + g.o, sg = g.s;
        const int ga = kc * n + g.a, gb = kc * n + g.b;
        // copy-paste rules of the unrolled face code: component rows, y+: the
        // gas MoshtaqGama carries Coef and BmoshtaqGama not (:31539-31541);
        // x-: the gas part assigns dBsidedp (:31466, :31473); total rows, y-:
        // the oil gravity and gas pressure terms with the neighbour upstream
        // assign dBsidedp
        const bool q_gas_swap = !total && q == 2;
        const bool q_gas_assign = !total && q == 1;
        const bool q_tot = total && q == 3;
        const int nlab = s.lab(on);
        const int reCol2 = nlab ? nc : 0, reCol4 = nlab ? 1 : 0;
        const int64_t mBLNo = A.boff[on] - reCol2;
        const int upo = F.up[0] == 0 ? ga : gb, upg = F.up[1] == 0 ? ga : gb;
        const bool nbo = rj_nb_up(F, g, 0), nbg = rj_nb_up(F, g, 1);
        const double dPo = s.P(gb, 1) - s.P(ga, 1);
        const double dPg = s.P(gb, 2) - s.P(ga, 2);
        const double dh = sg > 0 ? g.hn - g.hs : g.hs - g.hn;
        const double dsw = sg * (cx(upo, 0) * F.t[4] * dPo + cx(upo, 0) * F.t[6] * dh);
        if (nbo) out.add(mBLNo + 2 * nc + 1, dsw);
        else { dRsidedSw += dsw; out.add(mBLNo + 2 * nc + 1, 0.0); }
        const double dso = sg * (cx(upo, 0) * F.t[10] * dPo + cx(upo, 0) * F.t[11] * dh);
        if (nbo) out.add(mBLNo + 2 * nc + 2, nlab == -1 ? 0.0 : dso);
        else { dRsidedSo += dso; out.add(mBLNo + 2 * nc + 2, 0.0); }
        const double dsgo = sg * (cx(upo, 0) * F.t[8] * dPo + cx(upo, 0) * F.t[9] * dh);
        const double dsgg = sg * (cx(upg, 1) * F.t[5] * dPg + cx(upg, 1) * F.t[7] * dh);
        double vg;
        if (nbo) vg = nlab == 1 ? 0.0 : dsgo;
        else { dRsidedSg += dsgo; vg = 0; }
        if (nbg) vg += nlab == 1 ? 0.0 : dsgg;
        else dRsidedSg += dsgg;
        out.add(mBLNo + 2 * nc + 3 - reCol4, vg);

        double dBsidedp = 0;
        const double IF = A.iftran[((int64_t)kc * 3 + g.d) * n + g.a];
        for (int ph = 0; ph < 2; ph++) {
            const double Coef = ph == 0 ? (nlab == -1 ? 0.0 : 1.0) : (nlab == 1 ? 0.0 : 1.0);
            const double Surat = s.ro(bs_, ph) * g.GS + s.ro(on, ph) * g.GN * Coef;
            const double Makhraj = s.mu(bs_, ph) * g.GS + s.mu(on, ph) * g.GN * Coef;
            const double Mo = (s.dE(bs_, nc, ph) * g.GS * Makhraj - s.dMu(bs_, nc, ph) * g.GS * Surat) / (Makhraj * Makhraj);
            const double Bm = (s.dE(on, nc, ph) * g.GN * Coef * Makhraj - s.dMu(on, nc, ph) * g.GN * Coef * Surat) /
                              (Makhraj * Makhraj);
            double MG = g.GS * (s.gam(bs_, ph) * s.dE(bs_, nc, ph) / s.ro(bs_, ph)) / (g.GS + g.GN * Coef);
            double BMG = Coef * g.GN * (s.gam(on, ph) * s.dE(on, nc, ph) / s.ro(on, ph)) / (g.GS + g.GN * Coef);
            if (ph == 1 && q_gas_swap) {
                MG = Coef * g.GS * (s.gam(bs_, ph) * s.dE(bs_, nc, ph) / s.ro(bs_, ph)) / (g.GS + g.GN * Coef);
                BMG = g.GN * (s.gam(on, ph) * s.dE(on, nc, ph) / s.ro(on, ph)) / (g.GS + g.GN * Coef);
            }
            const double tempdx = (s.gam(bs_, ph) * g.GS + s.gam(on, ph) * g.GN * Coef) / (g.GS + g.GN * Coef);
            const bool nbu = ph == 0 ? nbo : nbg;
            const int u = nbu ? on : bs_;
            const double cu = cx(u, ph) * IF * s.kr(u, 1 + ph);
            const double dP = ph == 0 ? dPo : dPg;
            // the minus faces' gas gravity takes the neighbour's derivative
            // for the own column and the own one for the neighbour's
            const bool swap = ph == 1 && sg < 0;
            const double gOwn = swap ? (Bm * tempdx + BMG * Surat / Makhraj) : (Mo * tempdx + MG * Surat / Makhraj);
            const double gNb = swap ? (Mo * tempdx + MG * Surat / Makhraj) : (Bm * tempdx + BMG * Surat / Makhraj);
            dRsidedp -= sg * cu * Mo * dP;
            const double pNb = sg * cu * Bm * dP;
            if (ph == 0) dBsidedp = pNb;
            else if (q_gas_assign || (q_tot && nbu)) dBsidedp = pNb;
            else dBsidedp += pNb;
            dRsidedp -= sg * cu * gOwn * dh;
            if (ph == 0 && q_tot && nbu) dBsidedp = sg * cu * gNb * dh;
            else dBsidedp += sg * cu * gNb * dh;
            if (total) continue;
            // inter-cell Fickian flux, pressure derivative (:31247-31283)
            const double dC = dCoef[ph];
            const double porpor2 = s.porpor(on), dPordp2 = s.dpor(on);
            double Dup, dDdp, dDdp2;
            if (s.X(on, i, ph) > s.X(bs_, i, ph)) { Dup = s.Di(on, i, ph); dDdp = 0; dDdp2 = s.dDi(on, i, nc, ph); }
            else { Dup = s.Di(bs_, i, ph); dDdp = s.dDi(bs_, i, nc, ph); dDdp2 = 0; }
            if (s.phi(bs_) < 1e-10 || s.phi(on) < 1e-10) { Dup = 0; dDdp = 0; dDdp2 = 0; }
            const double Sv = s.S(on, 1 + ph), Ss = s.S(bs_, 1 + ph);
            double AAd = s.ro(on, ph) * g.GN * Coef + s.ro(bs_, ph) * g.GS * dC;
            AAd *= (Sv * g.GN * Coef + Ss * g.GS * dC) / (g.GN + g.GS);
            const double BBd = ftor(g.o) * g.GN * Coef / porpor2 + ftor(c) * g.GS * dC / porpor;
            const double MMd = AAd / BBd;
            const double NNd = g.area / (g.GN + g.GS);
            const double dX = s.X(gb, i, ph) - s.X(ga, i, ph);
            const double dAAddp = s.dE(bs_, nc, ph) * g.GS * dC;
            const double dAAddp2 = s.dE(on, nc, ph) * g.GN * Coef;
            const double dBBddp = -dPordp * ftor(c) * g.GS * dC / (porpor * porpor);
            const double dBBddp2 = -dPordp2 * ftor(g.o) * g.GN * Coef / (porpor2 * porpor2);
            const double dMMddp = (dAAddp * BBd - dBBddp * AAd) / (BBd * BBd);
            const double dMMddp2 = (dAAddp2 * BBd - dBBddp2 * AAd) / (BBd * BBd);
            if (Coef != 0 && dC != 0) {
                dRsidedp += -sg * NNd * dX * (MMd * dDdp + Dup * dMMddp);
                dBsidedp += sg * NNd * dX * (MMd * dDdp2 + Dup * dMMddp2);
            }
        }
        out.add(mBLNo + 2 * nc, pcoeff * (-cx(upo, 0) * F.t[0] - cx(upg, 1) * F.t[1] - dBsidedp));
    }

    // ------------------------------------------------- own and other P, S
    out.add(jacCol + 2 * nc - reCol, pcoeff * (dLp + dRsidedp));
    int reCol3;
    if (lab != -1) {
        out.add(jacCol + 2 * nc + 1 - reCol, -dRsidedSw);
        out.add(jacCol + 2 * nc + 2 - reCol, dLSo - dRsidedSo);
        reCol3 = 0;
    } else {
        reCol3 = 1;   // a gas block's own Sw column is not emitted here
    }
    if (lab != 1) out.add(jacCol + 2 * nc + 3 - reCol - reCol3, dLSg - dRsidedSg);
    if (olab == 1) out.add(OEq + 2 * nc + 2 - reColF1, dROSo);
    else if (olab == -1) out.add(OEq + 2 * nc + 2 - reColF1, dROSg);
    else {
        out.add(OEq + 2 * nc + 2 - reColF1, dROSo);
        out.add(OEq + 2 * nc + 3 - reColF1, dROSg);
    }
    out.add(OEq + 2 * nc - reColF1, pcoeff * dROp);
    out.add(OEq + 2 * nc + 1 - reColF1, dROSw);

    // -------------------------------------------------- composition columns
    for (int k = 0; k < nc; k++) {
        const bool same = !total && k == i;   // the k == i branch (:32238-32950)
        double vdR[2] = {0, 0}, vdRO[2] = {0, 0}, vdL[2] = {0, 0};
        if (kc == 0) {
            for (int ph = 0; ph < 2; ph++) {
                if (same) vdR[ph] += SI_C * ro_m[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                vdR[ph] += SI_C * s.dE(bm, k, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                vdR[ph] += -SI_C * s.dMu(bm, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * mpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
                vdRO[ph] += -SI_C * s.dMu(bf, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * fpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            }
            if (!total) {
                int tph[4];
                char tk[4];
                const int nt = rj_mdiff(lm, lf, tph, tk);
                for (int t = 0; t < nt; t++) {
                    const int ph = tph[t];
                    const double dl = tk[t] == 'F' ? s.X(bf, i, ph) - s.X(bm, i, ph) : 0.0 - s.X(bm, i, ph);
                    vdR[ph] += s.dE(bm, k, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * s.Di(bm, i, ph) * dl;
                    if (same) vdR[ph] += -sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_m[ph] * s.Di(bm, i, ph);
                    vdR[ph] += s.dDi(bm, i, k, ph) * sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_m[ph] * dl;
                }
                if (same) {
                    // d(x_f,i - x_m,i)/d x_f,i of the combinations that use
                    // the fracture composition (:32100-32120, :32206-32225)
                    for (int ph = 0; ph < 2; ph++) {
                        bool on = false;
                        if (ph == 0) on = (lm == 1 && lf == 1) || (lm == 0 && lf == 0) || (lm == 0 && lf == 1) || (lm == 1 && lf == 0);
                        else on = (lm == -1 && lf == -1) || (lm == 0 && lf == 0) || (lm == 0 && lf == -1);
                        if (on) vdRO[ph] += sh.sf1 * S_s[ph] * mpor * Vb / tor * ro_m[ph] * s.Di(bm, i, ph);
                    }
                }
            }
            for (int ph = 0; ph < 2; ph++)
                vdL[ph] = total ? s.dE(bm, k, ph) : (same ? ro_m[ph] : 0.0) + sx[ph] * s.dE(bm, k, ph);
        } else {
            for (int ph = 0; ph < 2; ph++) {
                if (same) vdRO[ph] -= SI_C * ro_m[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                vdRO[ph] -= SI_C * s.dE(bm, k, ph) * mx[ph] * cPh[ph] / (MUP[ph] / PP) * DPh[ph];
                vdRO[ph] -= -SI_C * s.dMu(bm, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * mpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
                vdR[ph] -= -SI_C * s.dMu(bf, k, ph) * mx[ph] * ro_m[ph] * cPh[ph] * fpor / (MUP[ph] * MUP[ph]) * DPh[ph] * PP;
            }
            if (!total) {
                int p1[4], p3[4];
                const int n1 = rj_fsf1(lm, lf, p1), n3 = rj_fsf3(lm, lf, p3);
                for (int t = 0; t < n1; t++) {
                    const int ph = p1[t];
                    const double dl = s.X(bf, i, ph) - s.X(bm, i, ph);
                    const double Sm = s.S(bm, 1 + ph);
                    vdRO[ph] -= s.dE(bm, k, ph) * sh.sf1 * Sm * mpor * Vb / tor * s.Di(bm, i, ph) * dl;
                    if (same) vdRO[ph] -= -sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * s.Di(bm, i, ph);
                    vdRO[ph] -= s.dDi(bm, i, k, ph) * sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * dl;
                    if (same) vdR[ph] -= sh.sf1 * Sm * mpor * Vb / tor * ro_m[ph] * s.Di(bm, i, ph);
                }
                for (int t = 0; t < n3; t++) {
                    const int ph = p3[t];
                    const double dl = s.X(bf, i, ph);
                    vdR[ph] -= s.dE(bf, k, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * s.Di(bf, i, ph) * dl;
                    if (same) vdR[ph] -= sh.sf3 * S_s[ph] * fpor * Vb / tor * ro_s[ph] * s.Di(bf, i, ph);
                    vdR[ph] -= s.dDi(bf, i, k, ph) * sh.sf3 * S_s[ph] * fpor * Vb / tor * ro_s[ph] * dl;
                }
            }
            for (int ph = 0; ph < 2; ph++)
                vdL[ph] = total ? s.dE(bf, k, ph) : (same ? ro_s[ph] : 0.0) + sx[ph] * s.dE(bf, k, ph);
        }

        for (int q = 0; q < 6; q++) {
            if (!has[q]) continue;
            const RjGeo& g = Gq[q];
            const RjFace& F = Fq[q];
            const int on = kc * n + g.o, sg = g.s;
            const int nlab = s.lab(on);
            const bool zf = g.d == 2;
            for (int ph = 0; ph < 2; ph++) {
                double Coef = ph == 0 ? (nlab == -1 ? 0.0 : 1.0) : (nlab == 1 ? 0.0 : 1.0);
                // in the k != i branch the x and y faces drop the gas term for
                // any single-phase neighbour (:33015 and copies)
                if (!total && !same && ph == 1 && !zf) Coef = nlab ? 0.0 : 1.0;
                const double Surat = s.ro(bs_, ph) * g.GS + s.ro(on, ph) * g.GN * Coef;
                const double Makhraj = s.mu(bs_, ph) * g.GS + s.mu(on, ph) * g.GN * Coef;
                const double Mo = (s.dE(bs_, k, ph) * g.GS * Makhraj - s.dMu(bs_, k, ph) * g.GS * Surat) / (Makhraj * Makhraj);
                const double Bm = (s.dE(on, k, ph) * g.GN * Coef * Makhraj - s.dMu(on, k, ph) * g.GN * Coef * Surat) /
                                  (Makhraj * Makhraj);
                const double MG = g.GS * (G_ACCG * s.ro(bs_, ph) * K.mw[k] + s.gam(bs_, ph) * s.dE(bs_, k, ph) / s.ro(bs_, ph)) /
                                  (g.GS + g.GN * Coef);
                const double BMG = Coef * g.GN * (G_ACCG * s.ro(on, ph) * K.mw[k] + s.gam(on, ph) * s.dE(on, k, ph) / s.ro(on, ph)) /
                                   (g.GS + g.GN * Coef);
                const double tempdx = (s.gam(bs_, ph) * g.GS + s.gam(on, ph) * g.GN * Coef) / (g.GS + g.GN * Coef);
                const bool nbu = rj_nb_up(F, g, ph);
                const int u = nbu ? on : bs_;
                const double IF = A.iftran[((int64_t)kc * 3 + g.d) * n + g.a];
                const double cu = cx(u, ph) * IF * s.kr(u, 1 + ph);
                const double dP = s.P(kc * n + g.b, 1 + ph) - s.P(kc * n + g.a, 1 + ph);
                const double dh = sg > 0 ? g.hn - g.hs : g.hs - g.hn;
                const double trP = same ? F.t[ph] : 0.0;   // d(upstream x_i)/dx_i
                const double trG = same ? F.t[2 + ph] : 0.0;
                double dB;
                if (!nbu) {
                    vdR[ph] += sg * (trP + cu * Mo) * dP;
                    dB = sg * cu * Bm * dP;
                } else {
                    vdR[ph] += sg * cu * Mo * dP;
                    dB = sg * (trP + cu * Bm) * dP;
                }
                if (zf) {
                    // gravity only on the z faces in the composition columns
                    const bool swap = ph == 1 && sg < 0;
                    const double gOwn = swap ? (Bm * tempdx + BMG * Surat / Makhraj) : (Mo * tempdx + MG * Surat / Makhraj);
                    const double gNb = swap ? (Mo * tempdx + MG * Surat / Makhraj) : (Bm * tempdx + BMG * Surat / Makhraj);
                    double sgn = sg;
                    // k != i, z- face, neighbour upstream, oil: both gravity
                    // terms carry the opposite sign (:34020-34021)
                    if (!total && !same && sg < 0 && nbu && ph == 0) sgn = -sgn;
                    if (!nbu) {
                        vdR[ph] += sgn * (trG + cu * gOwn) * dh;
                        dB += sgn * cu * gNb * dh;
                    } else {
                        // k == i, z+ face, oil: the own-column gravity term
                        // uses the y transmissibility of the cell (:33190)
                        double cuo = cu;
                        if (same && sg > 0 && ph == 0)
                            cuo = cx(u, ph) * A.iftran[((int64_t)kc * 3 + 1) * n + c] * s.kr(u, 1 + ph);
                        vdR[ph] += sgn * cuo * gOwn * dh;
                        dB += sgn * (trG + cu * gNb) * dh;
                    }
                }
                if (!total) {
                    // Fickian flux, composition derivative (:32303-32318):
                    // DeltaX is neighbour - self on both sides, and the
                    // minus faces negate the whole expression
                    const double dC = dCoef[ph];
                    double Dup, dDdx, dDdx2;
                    if (s.X(on, i, ph) > s.X(bs_, i, ph)) { Dup = s.Di(on, i, ph); dDdx = 0; dDdx2 = s.dDi(on, i, k, ph); }
                    else { Dup = s.Di(bs_, i, ph); dDdx = s.dDi(bs_, i, k, ph); dDdx2 = 0; }
                    const double Sv = s.S(on, 1 + ph), Ss = s.S(bs_, 1 + ph);
                    double AAd = s.ro(on, ph) * g.GN * Coef + s.ro(bs_, ph) * g.GS * dC;
                    AAd *= (Sv * g.GN * Coef + Ss * g.GS * dC) / (g.GN + g.GS);
                    const double BBd = ftor(g.o) * g.GN * Coef / s.porpor(on) + ftor(c) * g.GS * dC / porpor;
                    const double MMd = AAd / BBd;
                    const double NNd = g.area / (g.GN + g.GS);
                    const double dX = s.X(on, i, ph) - s.X(bs_, i, ph);
                    const double dMMddx = s.dE(bs_, k, ph) * g.GS * dC / BBd;
                    const double dMMddx2 = s.dE(on, k, ph) * g.GN * Coef / BBd;
                    if (Coef != 0 && dC != 0) {
                        const double dd = same ? MMd * Dup : 0.0;
                        vdR[ph] += sg * NNd * (-dd + dDdx * MMd * dX + dMMddx * Dup * dX);
                        dB += sg * NNd * (dd + dDdx2 * MMd * dX + dMMddx2 * Dup * dX);
                    }
                }
                const int reCol2 = nlab ? nc : 0;
                if (ph == 0) out.add(A.boff[on] + k, -dB * Coef);
                else out.add(A.boff[on] - reCol2 + k + nc, -dB * Coef);
            }
        }
        if (olab == 1) out.add(OEq + k, vdRO[0]);
        else if (olab == -1) out.add(OEq + k, vdRO[1]);
        else { out.add(OEq + k, vdRO[0]); out.add(OEq + nc + k, vdRO[1]); }
        if (lab != -1) out.add(jacCol + k, (Vb / Dt) * porpor * S_s[0] * vdL[0] - vdR[0]);
        if (lab != 1) out.add(jacCol + nc + k - reCol, (Vb / Dt) * porpor * S_s[1] * vdL[1] - vdR[1]);
    }
}

// Water row of block (kc, c) (matrix :36076-36170, fracture :18523-18621).
template <int NC>
__device__ void rj_water(const RjView<NC>& s, const Geo& G, int kc, double pcoeff, const RjEmit& out) {
    const AsmArgs& A = s.A;
    const AsmConsts& K = *A.K;
    constexpr int nc = NC;
    const int n = s.n, c = G.c;
    const int bs_ = kc * n + c, bo = (1 - kc) * n + c, bm = c, bf = n + c;
    const int lab = s.lab(bs_), olab = s.lab(bo);
    const int reCol = lab == 0 ? 0 : nc, reColF1 = olab == 0 ? 0 : nc;
    const int64_t jacCol = A.boff[bs_], OEq = A.boff[bo];
    const double Vb = G.Vb;
    const RjShape sh = rj_shape(A, G);
    const RjFM fm = rj_fm<NC>(s, c);
    out.add(jacCol + (lab == 0 ? 2 * nc : nc), 0.0);   // placeholder
    const double dLsidedp = (Vb / K.dt) * K.wat_ro * WAT_M_RO * s.S(bs_, 0) * s.dpor(bs_);
    const double dLsidedSw = (Vb / K.dt) * (s.porpor(bs_) * K.wat_ro * WAT_M_RO);
    // transfer pressure terms with the opposite sign of the residual's
    const double cW = SI_C * K.wat_ro * WAT_M_RO * sh.sf2 * fm.kr[0] / K.wat_mu * Vb;
    double dRsidedp = -cW;
    const double dRsideOdp = cW;
    RjGeo Gq[6];
    bool has[6];
    RjFace Fq[6];
    for (int q = 0; q < 6; q++) {
        has[q] = rj_geo(A, G, q, Gq[q]);
        if (has[q]) Fq[q] = rj_face_of<NC>(s, kc, Gq[q]);
    }
    for (int q = 0; q < 6; q++)
        if (has[q]) dRsidedp += Fq[q].W;
    out.add(jacCol + 2 * nc - reCol, pcoeff * (dLsidedp + dRsidedp));
    const double dPw = (s.P(bf, 0) - s.P(bm, 0)) + (s.pc(bm, 0) - s.pc(bf, 0));
    const double cS = SI_C * WAT_M_RO;
    double dRsidedSw, dRsideOdSw;
    if (kc == 0) {
        dRsidedSw = cS * fm.d[1][0] * K.wat_ro * sh.sf2 / K.wat_mu * Vb * dPw;
        dRsideOdSw = cS * fm.d[1][5] * K.wat_ro * sh.sf2 / K.wat_mu * Vb * dPw;
    } else {
        dRsidedSw = -(cS * fm.d[1][5] * K.wat_ro * sh.sf2 / K.wat_mu * Vb * dPw);
        dRsideOdSw = -(cS * fm.d[1][0] * K.wat_ro * sh.sf2 / K.wat_mu * Vb * dPw);
    }
    for (int q = 0; q < 6; q++) {
        if (!has[q]) continue;
        const RjGeo& g = Gq[q];
        const RjFace& F = Fq[q];
        const int on = kc * n + g.o;
        const int reCol2 = s.lab(on) ? nc : 0;
        const int64_t mBLNo = A.boff[on] - reCol2;
        out.add(mBLNo + 2 * nc, pcoeff * (-F.W));
        const double dP = s.P(kc * n + g.b, 0) - s.P(kc * n + g.a, 0);
        const double dh = g.s > 0 ? g.hn - g.hs : g.hs - g.hn;
        const double dsw = g.s * (F.dW * dP + G_ACC * K.wat_ro * F.dW * dh);
        if (rj_nb_up(F, g, 2)) out.add(mBLNo + 2 * nc + 1, dsw);
        else { dRsidedSw += dsw; out.add(mBLNo + 2 * nc + 1, 0.0); }
    }
    out.add(jacCol + 2 * nc + 1 - reCol, dLsidedSw - dRsidedSw);
    out.add(OEq + 2 * nc - reColF1, pcoeff * dRsideOdp);
    out.add(OEq + 2 * nc + 1 - reColF1, dRsideOdSw);
}

// Fugacity row i of a two-phase block (:30339-30405, fracture :12400-12466):
// d[x_i phiL_i(x, Po) Po - y_i phiG_i(y, Pg) Pg], Z from the stored density,
// dZ from dE.  Columns 0..2nc of the block, written directly.
template <int NC>
__device__ void rj_fugacity(const RjView<NC>& s, int blk, int i, double pcoeff, double* v) {
    const AsmConsts& K = *s.A.K;
    constexpr int nc = NC;
    const double T = K.T, RT = RGAS * T, RT2 = RT * RT;
    const double Po = s.P(blk, 1), Pg = s.P(blk, 2);
    double xl[NC], xg[NC];
#pragma unroll
    for (int a = 0; a < NC; a++) { xl[a] = s.X(blk, a, 0); xg[a] = s.X(blk, a, 1); }
    double Al = 0, Bl = 0, Ag = 0, Bg = 0;
    double tl = 0, tg = 0;
#pragma unroll
    for (int a = 0; a < NC; a++) {
        Bl += xl[a] * K.b[a];
        Bg += xg[a] * K.b[a];
        double sl = 0, sgs = 0;
#pragma unroll
        for (int b = 0; b < NC; b++) {
            const double t = K.sa[a] * K.sa[b] * K.bic[a * NC + b];
            sl += xl[b] * t;
            sgs += xg[b] * t;
        }
        Al += xl[a] * sl;
        Ag += xg[a] * sgs;
        if (a == i) { tl = sl; tg = sgs; }
    }
    const double Cal = Al, Cbl = Bl, Cag = Ag, Cbg = Bg;
    Al *= Po / RT2; Bl *= Po / RT;
    Ag *= Pg / RT2; Bg *= Pg / RT;
    const double dAdpl = Cal / RT2, dAdpg = Cag / RT2;
    const double dBdpl = Cbl / RT, dBdpg = Cbg / RT;
    const double Zl = Po / (s.ro(blk, 0) * RT), Zg = Pg / (s.ro(blk, 1) * RT);
    const double dZdpl = (1 - s.dE(blk, nc, 0) * RT * Zl) * (Zl / Po);
    const double dZdpg = (1 - s.dE(blk, nc, 1) * RT * Zg) * (Zg / Pg);
    const double d1 = K.d1, d2 = K.d2;
    const double t1l = K.b[i] / Cbl, t1g = K.b[i] / Cbg;
    const double EEl = 2 * tl / Cal - t1l, EEg = 2 * tg / Cag - t1g;
    const double DDl = Al / (Bl * (d1 - d2)), DDg = Ag / (Bg * (d1 - d2));
    const double FFl = log((Zl + d2 * Bl) / (Zl + d1 * Bl)), FFg = log((Zg + d2 * Bg) / (Zg + d1 * Bg));
    const double phiL = exp(t1l * (Zl - 1) - log(Zl - Bl) + DDl * EEl * FFl);
    const double phiG = exp(t1g * (Zg - 1) - log(Zg - Bg) + DDg * EEg * FFg);
    const double dAAdpl = K.b[i] * dZdpl / Cbl, dAAdpg = K.b[i] * dZdpg / Cbg;
    const double dBBdpl = (dBdpl - dZdpl) / (Zl - Bl), dBBdpg = (dBdpg - dZdpg) / (Zg - Bg);
    const double dDDdpl = (Bl * dAdpl - Al * dBdpl) / (Bl * Bl * (d1 - d2));
    const double dDDdpg = (Bg * dAdpg - Ag * dBdpg) / (Bg * Bg * (d1 - d2));
    const double dFFdpl = (dZdpl + d2 * dBdpl) / (Zl + d2 * Bl) - (dZdpl + d1 * dBdpl) / (Zl + d1 * Bl);
    const double dFFdpg = (dZdpg + d2 * dBdpg) / (Zg + d2 * Bg) - (dZdpg + d1 * dBdpg) / (Zg + d1 * Bg);
    const double dPhidpl = phiL * (dAAdpl + dBBdpl + EEl * (dDDdpl * FFl + DDl * dFFdpl));
    const double dPhidpg = phiG * (dAAdpg + dBBdpg + EEg * (dDDdpg * FFg + DDg * dFFdpg));
    const double xi = xl[i], yi = xg[i];
    v[2 * nc] = pcoeff * (xi * (phiL + Po * dPhidpl) - yi * (phiG + Pg * dPhidpg));
#pragma unroll
    for (int kk = 0; kk < NC; kk++) {
        double dAdxl = 0, dAdxg = 0;
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const double t = K.sa[kk] * K.sa[q] * K.bic[kk * NC + q];
            dAdxl += xl[q] * t;
            dAdxg += xg[q] * t;
        }
        dAdxl *= 2 * Po / RT2;
        dAdxg *= 2 * Pg / RT2;
        const double dBdxl = K.b[kk] * Po / RT, dBdxg = K.b[kk] * Pg / RT;
        const double dZdxl = -s.dE(blk, kk, 0) * RT * Zl * Zl / Po;
        const double dZdxg = -s.dE(blk, kk, 1) * RT * Zg * Zg / Pg;
        const double dAAdxl = -K.b[kk] * K.b[i] * (Zl - 1) / (Cbl * Cbl) + K.b[i] * dZdxl / Cbl;
        const double dAAdxg = -K.b[kk] * K.b[i] * (Zg - 1) / (Cbg * Cbg) + K.b[i] * dZdxg / Cbg;
        const double dBBdxl = (dBdxl - dZdxl) / (Zl - Bl), dBBdxg = (dBdxg - dZdxg) / (Zg - Bg);
        const double dDDdxl = (dAdxl * Bl - dBdxl * Al) / (Bl * Bl * (d1 - d2));
        const double dDDdxg = (dAdxg * Bg - dBdxg * Ag) / (Bg * Bg * (d1 - d2));
        const double dFFdxl = (dZdxl + d2 * dBdxl) / (Zl + d2 * Bl) - (dZdxl + d1 * dBdxl) / (Zl + d1 * Bl);
        const double dFFdxg = (dZdxg + d2 * dBdxg) / (Zg + d2 * Bg) - (dZdxg + d1 * dBdxg) / (Zg + d1 * Bg);
        const double tik = K.sa[i] * K.sa[kk] * K.bic[i * NC + kk];
        const double dEEdxl = 2 * (tik / Cal - dAdxl * RT2 * tl / (Po * Cal * Cal)) + K.b[i] * K.b[kk] / (Cbl * Cbl);
        const double dEEdxg = 2 * (tik / Cag - dAdxg * RT2 * tg / (Pg * Cag * Cag)) + K.b[i] * K.b[kk] / (Cbg * Cbg);
        const double dPhidxl = phiL * (dAAdxl + dBBdxl + dDDdxl * EEl * FFl + DDl * dEEdxl * FFl + DDl * EEl * dFFdxl);
        const double dPhidxg = phiG * (dAAdxg + dBBdxg + dDDdxg * EEg * FFg + DDg * dEEdxg * FFg + DDg * EEg * dFFdxg);
        if (i == kk) {
            v[kk] = Po * phiL + Po * xi * dPhidxl;
            v[nc + kk] = -(Pg * phiG + Pg * yi * dPhidxg);
        } else {
            v[kk] = Po * xi * dPhidxl;
            v[nc + kk] = -Pg * yi * dPhidxg;
        }
    }
}

// row length of row r of block blk
template <int NC>
__device__ inline int rj_row_len(const AsmArgs& A, const Geo& G, int k, int lab, int r) {
    constexpr int nc = NC;
    if (A.phi[k * A.n + G.c] < 1e-10) return 1;
    int kind;   // 0 balance, 1 water, 2 fugacity, 3 constraint
    if (lab == 0) kind = r < nc ? 2 : (r < 2 * nc ? 0 : (r == 2 * nc ? 1 : 3));
    else kind = r < nc ? 0 : (r == nc ? 1 : 3);
    if (kind == 2) return 2 * nc + 1;
    if (kind == 3) {
        if (lab == 0) return r == 2 * nc + 3 ? 3 : nc + 1;
        return r == nc + 1 ? nc + 1 : 2;
    }
    RjRow R;
    rj_slots(A, G, k, kind == 1, R);
    return R.len;
}

template <int NC>
__global__ __launch_bounds__(256) void k_rj_rownnz(AsmArgs A, int64_t* rowlen) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= 2 * A.n) return;
    const int k = blk / A.n, c = blk - k * A.n;
    const int lab = A.lab[k][c];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    const Geo G = geom(A, c);
    const int64_t r0 = A.boff[blk];
    for (int r = 0; r < bs; r++) rowlen[r0 + r] = rj_row_len<NC>(A, G, k, lab, r);
}

// one lane per (block, row); lanes past the block's size return
template <int NC>
__global__ __launch_bounds__(128) void k_rj_rows(AsmArgs A, const double* tab, const int64_t* rowptr,
                                                 int32_t* cols, double* vals) {
    constexpr int MAXR = 2 * NC + 4;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = 2 * (int64_t)A.n;
    if (idx >= nb * MAXR) return;
    const int blk = (int)(idx % nb);
    const int r = (int)(idx / nb);
    const int k = blk / A.n, c = blk - k * A.n;
    const int lab = A.lab[k][c];
    const int bs = lab == 0 ? 2 * NC + 4 : NC + 3;
    if (r >= bs) return;
    constexpr int nc = NC;
    const int64_t rowno = A.boff[blk] + r;
    const int64_t p0 = rowptr[rowno];
    int32_t* cr = cols + p0;
    double* vr = vals + p0;
    const int64_t jacCol = A.boff[blk];
    if (A.phi[blk] < 1e-10) {
        // zero-porosity block (:42051-42056): a unit diagonal
        cr[0] = (int32_t)rowno;
        vr[0] = 1.0;
        return;
    }
    const Geo G = geom(A, c);
    const RjView<NC> V(A, tab);
    int kind;
    if (lab == 0) kind = r < nc ? 2 : (r < 2 * nc ? 0 : (r == 2 * nc ? 1 : 3));
    else kind = r < nc ? 0 : (r == nc ? 1 : 3);
    if (kind == 2) {
        for (int l = 0; l <= 2 * nc; l++) cr[l] = (int32_t)(jacCol + l);
        rj_fugacity<NC>(V, blk, r, A.pcoeff, vr);
        return;
    }
    if (kind == 3) {
        // sum x = 1, sum y = 1, sum S = 1 with the diagonal placeholder
        // (:36182-36228 and copies); columns in order
        if (lab == 0) {
            if (r == 2 * nc + 3) {
                for (int l = 0; l < 3; l++) { cr[l] = (int32_t)(jacCol + 2 * nc + 1 + l); vr[l] = 1.0; }
            } else {
                const int b0 = r == 2 * nc + 1 ? 0 : nc;
                for (int l = 0; l < nc; l++) { cr[l] = (int32_t)(jacCol + b0 + l); vr[l] = 1.0; }
                cr[nc] = (int32_t)(jacCol + r);
                vr[nc] = 0.0;
            }
        } else {
            if (r == nc + 1) {
                for (int l = 0; l < nc; l++) { cr[l] = (int32_t)(jacCol + l); vr[l] = 1.0; }
                cr[nc] = (int32_t)(jacCol + r);
                vr[nc] = 0.0;
            } else {
                cr[0] = (int32_t)(jacCol + nc + 1); vr[0] = 1.0;
                cr[1] = (int32_t)(jacCol + nc + 2); vr[1] = 1.0;
            }
        }
        return;
    }
    RjRow R;
    rj_slots(A, G, k, kind == 1, R);
    rj_fill_cols(R, cr, vr);
    const RjEmit out(R, vr);
    if (kind == 1) {
        rj_water<NC>(V, G, k, A.pcoeff, out);
    } else {
        const int i = lab == 0 ? r - nc : r;   // component (nc - 1 = total)
        rj_balance<NC>(V, G, k, i == nc - 1 ? nc - 1 : i, i == nc - 1, A.pcoeff, out);
    }
}

}  // namespace osr

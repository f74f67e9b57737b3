// osr_assembly.cpp -- CPU FP64 restatement of the Osiris residual and
// Jacobian assembly for the matrix and fracture continua.
//
// TEST INFRASTRUCTURE ONLY (see osr_oracle.h): the HIP product never links,
// loads or calls this file.
//
// Restated reference code (compsim/...):
//   property sweep   PetscAUX.cpp:500-547 (UpdateDependencies order):
//                    CalcIFT Do_Cycle.cpp:2438-2456, CalcSatFs :99-259,
//                    fCalcSatFs :261-437, Calc_Phase_P :1418-1446,
//                    Calc_Visco :438-616 / fCalc_Visco :619-799,
//                    CalcDiffusion Mass_Tran.cpp:21-62, CalcPhaseGrav
//                    Do_Cycle.cpp:2022-2057, Pre_Mul Do_Time_Step.cpp:54-84,
//                    Tr_Mu Init_Calcs.cpp:48-64, EOS_Init :66-97
//   face terms       Calc_Transes Do_Cycle.cpp:800-1108, fCalc_Transes
//                    :1109-1417, IFK_Calc Init_Calcs.cpp:21-42,
//                    CalcBlockHeight Init_Calcs.cpp:993-1011
//   residual rows    IteratorFunction Petsc_Functions.cpp:10-6896
//                    (fracture rows :102-3443, matrix rows :3446-6870,
//                    zero-porosity rows :6878-6883)
//   layout           PJac Init_Calcs.cpp:1013-1116, InitGuess2
//                    PetscAUX.cpp:878-964, UpdateAll :176-498
//
// The Jacobian is the forward-mode derivative of this residual with respect
// to the reference's unknowns (x|y, P/PCoeff, Sw, So|Sg per block), holding
// the previous-step accumulation (preProp) and the previous-sweep densities
// that CalcIFT reads fixed; see dual.hpp for the branch convention.
#include "osr_model.hpp"


extern "C" {

int or_layout(const or_grid* g, const or_state* s, int nc, int32_t* bsize, int64_t* offset) {
    const int n = g->nx * g->ny * g->nz;
    int64_t off = 0;
    for (int k = 0; k < 2; k++) {
        const int8_t* lab = k == 0 ? s->label_m : s->label_f;
        for (int c = 0; c < n; c++) {
            const int bs = block_size(nc, lab[c]);
            bsize[k * n + c] = bs;
            offset[k * n + c] = off;
            off += bs;
        }
    }
    offset[2 * n] = off;
    return 0;
}

int or_residual_scale(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, const or_state* s, double* F);

int or_residual(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, const or_state* s, double* F) {
    Model M;
    build_model(M, f, fe, g);
    Sweep S;
    sweep(M, s, S);
    const int n = M.n;
    int64_t off = 0;
    double rows[2 * NCM + 4];
    for (int k = 0; k < 2; k++)
        for (int c = 0; c < n; c++) {
            double pb[NCM + 1];
            block_rows<double>(M, k, c, S.V, S.P, S.pre(k, n, M.nc, c, pb), rows);
            const int bs = block_size(M.nc, S.V[k][c].label);
            for (int r = 0; r < bs; r++) F[off + r] = rows[r];
            off += bs;
        }
    return 0;
}

// Per-well Calc_Re and FlashInjection outputs of the grid's wells:
// re[nw], inj[nw][2nc + 6] (zeros for wells that are not types 3 and 8).
int or_well_data(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, double* re, double* inj) {
    Model M;
    build_model(M, f, fe, g);
    for (int i = 0; i < M.nw; i++) re[i] = M.well_re[i];
    for (size_t i = 0; i < M.inj.size(); i++) inj[i] = M.inj[i];
    return M.nw;
}

// Per-row sum of the magnitudes of the terms the residual row adds up.
int or_residual_scale(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, const or_state* s, double* F) {
    g_absmode = true;
    const int rc = or_residual(f, fe, g, s, F);
    g_absmode = false;
    return rc;
}

// Cell properties (for property-level parity): per continuum and cell
// [porpor, Pw, Po, Pg, bift, ro_o, ro_g, mu_o, mu_g, gam_o, gam_g, krw, kro,
//  krg, pcw, pcg, D_o[nc], D_g[nc]] -> 16 + 2 nc values.
int or_props(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, const or_state* s, double* out) {
    Model M;
    build_model(M, f, fe, g);
    Sweep S;
    sweep(M, s, S);
    const int n = M.n, nc = M.nc, w = 16 + 2 * nc;
    for (int k = 0; k < 2; k++)
        for (int c = 0; c < n; c++) {
            const Props<double>& p = S.P[k][c];
            double* o = out + ((int64_t)k * n + c) * w;
            const double v[16] = {p.porpor, p.Pw, p.Po, p.Pg, p.bift, p.ro[0], p.ro[1], p.mu[0], p.mu[1],
                                  p.gam[0], p.gam[1], p.kr[0], p.kr[1], p.kr[2], p.pcw, p.pcg};
            for (int q = 0; q < 16; q++) o[q] = v[q];
            for (int i = 0; i < nc; i++) { o[16 + i] = p.D[0][i]; o[16 + nc + i] = p.D[1][i]; }
        }
    return 0;
}

// IFK_Calc's face coefficients kA/h (Init_Calcs.cpp:21-42) of the model:
// out [2][3][n] (continuum, direction, cell; the face to the +direction
// neighbour, 0 where there is none).
int or_iftran(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, double* out) {
    Model M;
    build_model(M, f, fe, g);
    const int n = M.n;
    for (int k = 0; k < 2; k++)
        for (int d = 0; d < 3; d++)
            for (int c = 0; c < n; c++) out[((int64_t)k * 3 + d) * n + c] = M.iftran[k][d][c];
    return 0;
}

// Jacobian in the reference layout as CSR.  Row pattern: every row of a
// block depends on its own block; component-balance and total rows on the
// blocks of the six face neighbours (same continuum) and of the other
// continuum's cell; the water row on the P and Sw columns of those blocks;
// fugacity and constraint rows on their own block only.  Columns sorted.
// Returns the nnz, or -1 if cap is too small (rowptr is still filled).
// The rows of the listed blocks (all blocks in order when blocks is NULL),
// consecutively in list order: rowptr has one entry per listed row plus one.
int64_t or_jacobian_blocks(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, const or_state* s,
                           const int32_t* blocks, int32_t nblocks, int64_t* rowptr, int32_t* cols, double* vals,
                           int64_t cap) {
    Model M;
    build_model(M, f, fe, g);
    Sweep S;
    sweep(M, s, S);
    const int n = M.n, nc = M.nc;
    std::vector<int64_t> boff(2 * n + 1);
    std::vector<int32_t> bsz(2 * n);
    {
        int64_t off = 0;
        for (int k = 0; k < 2; k++)
            for (int c = 0; c < n; c++) {
                bsz[k * n + c] = block_size(nc, S.V[k][c].label);
                boff[k * n + c] = off;
                off += bsz[k * n + c];
            }
        boff[2 * n] = off;
    }
    // dual copies of all cells (constants) reused per dependency block
    std::vector<Vars<D>> VD[2];
    std::vector<Props<D>> PD[2];
    for (int k = 0; k < 2; k++) {
        VD[k].resize(n);
        PD[k].resize(n);
        for (int c = 0; c < n; c++) {
            lift(S.V[k][c], VD[k][c], nc);
            liftp(S.P[k][c], PD[k][c], nc);
        }
    }
    int64_t nnz = 0, rowno = 0;
    bool overflow = false;
    D rows[2 * NCM + 4];
    std::vector<double> blockJ;
    const int32_t nlist = blocks ? nblocks : 2 * n;
    for (int32_t bi = 0; bi < nlist; bi++) {
        {
            const int b = blocks ? blocks[bi] : bi;
            if (b < 0 || b >= 2 * n) return -2;
            const int k = b / n, c = b % n;
            const int iz = c / (M.nx * M.ny), rem = c % (M.nx * M.ny), iy = rem / M.nx, ix = rem % M.nx;
            const int bs = bsz[k * n + c];
            const int lab = S.V[k][c].label;
            const int r0 = lab == 0 ? nc : 0;
            // dependency blocks: (continuum, cell), in column order
            std::vector<std::pair<int, int>> deps;
            deps.push_back({k, c});
            const int cand[6] = {ix < M.nx - 1 ? c + 1 : -1, ix > 0 ? c - 1 : -1,
                                 iy < M.ny - 1 ? c + M.nx : -1, iy > 0 ? c - M.nx : -1,
                                 iz < M.nz - 1 ? c + M.nx * M.ny : -1, iz > 0 ? c - M.nx * M.ny : -1};
            for (int q = 0; q < 6; q++)
                if (cand[q] >= 0) deps.push_back({k, cand[q]});
            deps.push_back({1 - k, c});
            std::sort(deps.begin(), deps.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                return boff[a.first * n + a.second] < boff[b.first * n + b.second];
            });
            // derivative block for each dependency: J[r][col]
            std::vector<std::vector<double>> Jd(deps.size());
            for (size_t q = 0; q < deps.size(); q++) {
                const int dk = deps[q].first, dc = deps[q].second;
                const int dbs = bsz[dk * n + dc];
                Vars<D> saved = VD[dk][dc];
                Props<D> savedp = PD[dk][dc];
                seed(VD[dk][dc], nc, s->pcoeff);
                double rp[2];
                cell_props<D>(M, dk, dc, VD[dk][dc], S.rp(n, dc, rp), PD[dk][dc]);
                double pb[NCM + 1];
                block_rows<D>(M, k, c, VD, PD, S.pre(k, n, nc, c, pb), rows);
                Jd[q].assign((size_t)bs * dbs, 0.0);
                for (int r = 0; r < bs; r++)
                    for (int j = 0; j < dbs && j < rows[r].n; j++) Jd[q][(size_t)r * dbs + j] = rows[r].d[j];
                VD[dk][dc] = saved;
                PD[dk][dc] = savedp;
            }
            if (M.phi[k][c] < 1e-10) {
                // zero-porosity block: identity rows (Petsc_Functions.cpp:42051-42056)
                for (int r = 0; r < bs; r++) {
                    const int64_t row = boff[k * n + c] + r;
                    rowptr[rowno++] = nnz;
                    if (nnz < cap) { cols[nnz] = (int32_t)row; vals[nnz] = 1.0; }
                    else overflow = true;
                    nnz++;
                }
                continue;
            }
            for (int r = 0; r < bs; r++) {
                rowptr[rowno++] = nnz;
                const bool mass = (r >= r0 && r < r0 + nc);
                const bool water = (r == r0 + nc);
                for (size_t q = 0; q < deps.size(); q++) {
                    const int dk = deps[q].first, dc = deps[q].second;
                    const bool own = (dk == k && dc == c);
                    if (!own && !mass && !water) continue;
                    const int dbs = bsz[dk * n + dc];
                    const int dlab = S.V[dk][dc].label;
                    const int pcol = dlab == 0 ? 2 * nc : nc;   // P/PCoeff column within the block
                    for (int j = 0; j < dbs; j++) {
                        if (water && !own && !(j == pcol || j == pcol + 1)) continue;
                        if (nnz < cap) {
                            cols[nnz] = (int32_t)(boff[dk * n + dc] + j);
                            vals[nnz] = Jd[q][(size_t)r * dbs + j];
                        } else {
                            overflow = true;
                        }
                        nnz++;
                    }
                }
            }
        }
    }
    rowptr[rowno] = nnz;
    return overflow ? -1 : nnz;
}

int64_t or_jacobian(const or_fluid* f, const or_fluid_ext* fe, const or_grid* g, const or_state* s,
                    int64_t* rowptr, int32_t* cols, double* vals, int64_t cap) {
    return or_jacobian_blocks(f, fe, g, s, nullptr, 0, rowptr, cols, vals, cap);
}

// UpdateAll (PetscAUX.cpp:176-498): unknown vector X (reference layout) ->
// state, with the clamps, renormalisations, pressure limiter and floor.
// The state arrays are updated in place (stale x or y of single-phase
// cells are kept, as in the reference).
int or_unpack(const or_grid* g, or_state_mut* s, int nc, const double* X) {
    const int n = g->nx * g->ny * g->nz;
    int64_t p = 0;
    for (int k = 0; k < 2; k++) {
        const int8_t* lab = k == 0 ? s->label_m : s->label_f;
        double* Po = k == 0 ? s->Po_m : s->Po_f;
        double* x = k == 0 ? s->x_m : s->x_f;
        double* y = k == 0 ? s->y_m : s->y_f;
        double* S = k == 0 ? s->S_m : s->S_f;
        for (int c = 0; c < n; c++) {
            auto comp_in = [&](double* arr) {
                double sum = 0;
                for (int i = 0; i < nc; i++) {
                    double v = X[p++];
                    if (v < EPS_COMP) v = 0;
                    else if (v > 1) v = 1;
                    arr[(int64_t)i * n + c] = v;
                    sum += v;
                }
                return sum;
            };
            auto comp_norm = [&](double* arr, double sum) {
                for (int i = 0; i < nc; i++) {
                    if (sum != 0) arr[(int64_t)i * n + c] /= sum;
                    else arr[(int64_t)i * n + c] = 1.0 / nc;
                }
            };
            const int L = lab[c];
            if (L == 1) comp_norm(x, comp_in(x));
            else if (L == -1) comp_norm(y, comp_in(y));
            else {
                const double sl = comp_in(x);
                const double sg = comp_in(y);
                comp_norm(x, sl);
                comp_norm(y, sg);
            }
            // pressure limiter: adds MAX_P_COEFF (in Pa) when the step is above 1 %
            const double xp = X[p++];
            double P = Po[c];
            if (std::fabs(s->pcoeff * xp) > P * (1 + MAX_P_COEFF)) {
                const double dv = xp * s->pcoeff - P;
                P += MAX_P_COEFF * (dv > 0 ? 1.0 : (dv < 0 ? -1.0 : 0.0));
            } else {
                P = xp * s->pcoeff;
            }
            if (P < MINPPP) P = MINPPP;
            Po[c] = P;
            double* Sw = S + c;
            double* So = S + n + c;
            double* Sg = S + 2 * n + c;
            *Sw = X[p++];
            if (L == 1) {
                *So = X[p++];
                if (*Sw < 0) *Sw = 0;
                if (*Sw > 1) *Sw = 1;
                if (*So > 1) *So = 1;
                if (*So < 0) *So = 0;
                const double sum = *Sw + *So;
                if (sum != 0) { *Sw /= sum; *So /= sum; }
            } else if (L == -1) {
                *Sg = X[p++];
                if (*Sw < 0) *Sw = 0;
                if (*Sw > 1) *Sw = 1;
                if (*Sg > 1) *Sg = 1;
                if (*Sg < 0) *Sg = 0;
                const double sum = *Sw + *Sg;
                if (sum != 0) { *Sw /= sum; *Sg /= sum; }
            } else {
                *So = X[p++];
                *Sg = X[p++];
                if (*Sw < 0) *Sw = 0;
                if (*Sw > 1) *Sw = 1;
                if (*So > 1) *So = 1;
                if (*Sg > 1) *Sg = 1;
                if (*So < 0) *So = 0;
                if (*Sg < 0) *Sg = 0;
                const double sum = *Sw + *So + *Sg;
                if (sum != 0) { *Sw /= sum; *So /= sum; *Sg /= sum; }
            }
        }
    }
    return 0;
}

// InitGuess2 (PetscAUX.cpp:878-964): state -> unknown vector X.
int or_pack(const or_grid* g, const or_state* s, int nc, double* X) {
    const int n = g->nx * g->ny * g->nz;
    int64_t p = 0;
    for (int k = 0; k < 2; k++) {
        const int8_t* lab = k == 0 ? s->label_m : s->label_f;
        const double* Po = k == 0 ? s->Po_m : s->Po_f;
        const double* x = k == 0 ? s->x_m : s->x_f;
        const double* y = k == 0 ? s->y_m : s->y_f;
        const double* S = k == 0 ? s->S_m : s->S_f;
        for (int c = 0; c < n; c++) {
            const int L = lab[c];
            if (L != -1) for (int i = 0; i < nc; i++) X[p++] = x[(int64_t)i * n + c];
            if (L != 1) for (int i = 0; i < nc; i++) X[p++] = y[(int64_t)i * n + c];
            X[p++] = Po[c] / s->pcoeff;
            X[p++] = S[c];
            if (L != -1) X[p++] = S[n + c];
            if (L != 1) X[p++] = S[2 * n + c];
        }
    }
    return 0;
}

}  // extern "C"

/*
 * osr_oracle.h -- CPU FP64 restatement of the Osiris per-cell Newton hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, never by the product path.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root, compsim/...).  Arithmetic is plain IEEE
 * double, written to follow the reference's operation order where that order
 * is observable (iteration stopping rules, clamps, special cases).
 */
#ifndef OSR_ORACLE_H
#define OSR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PR 0
#define OR_SRK 1

/* Constants the reference compiles in (compsim/Functions.h). */
#define OR_RGAS 8.3144621             /* Functions.h:216 */
#define OR_PI 3.1415926535897932384626433
#define OR_SQRT2 1.4142135623730950488016887
#define OR_NEWTON_TOL 1e-8            /* Functions.h:236 */
#define OR_FLASH_TOL 1e-14            /* Functions.h:237 */
#define OR_DZERO 1e-20                /* Functions.h:239 */
#define OR_MAX_FLASH_N_ITR 100        /* Functions.h:126 */
#define OR_MAX_SS_ITER 500            /* Init_Calcs.cpp:714 */
#define OR_SXMTOL 1e-15               /* Functions.h:251 */
#define OR_XMTOL 1e-8                 /* Functions.h:224 */
#define OR_TRIV_K 1e-6                /* Functions.h:269 */
#define OR_MAX_MNR_ITER 20            /* Functions.h:270 */
#define OR_PZERO 1000000.0            /* Functions.h:247 */
#define OR_EPS_SAT 1e-3               /* Functions.h:226 */
#define OR_PHASE_ENV 0.95             /* Functions.h:11 */
#define OR_EPS_SAT_FRAC 0.0           /* Functions.h:12 */

typedef struct {
    int nc;
    int eos;              /* OR_PR or OR_SRK */
    const double *Tc;     /* [nc] K   */
    const double *Pc;     /* [nc] Pa  */
    const double *w;      /* [nc] acentric factor */
    const double *bic;    /* [nc*nc] row-major, holds 1 - k_ij (Data_Input.cpp:977-990) */
} or_fluid;

/* EOS_Init (Init_Calcs.cpp:66-97): a_i(T), b_i for PR or SRK. */
void or_eos_ab(const or_fluid *f, double T, double *a, double *b);

/* Solve_Z (Globals.cpp:212-264): root of Z^3 + a1 Z^2 + a2 Z + a3, gas=1 picks
 * the vapour root, gas=0 the liquid root. */
double or_solve_z(double a1, double a2, double a3, int gas);

/* One phase: mixing, Z root and ln(phi_i) (canonical form Init_Calcs.cpp:633-712).
 * x need not be normalised (the reference mixes unnormalised compositions). */
void or_phase(const or_fluid *f, const double *a, const double *b, double T,
              double P, const double *x, int gas, double *Z, double *lnphi);

typedef struct {
    double L, Zl, Zg, so, sg, rho_l, rho_g;
    int label;   /* 1 liquid, 0 two-phase, -1 gas (Functions.h:327) */
    int iters;   /* outer successive-substitution iterations */
    int rr_iters;/* Rachford-Rice Newton iterations over all passes */
} or_flash_out;

/* SuccFalsh (Init_Calcs.cpp:571-738) for one cell. */
void or_flash(const or_fluid *f, double T, double P, const double *z, double sw,
              double *x, double *y, double *K, or_flash_out *o);

/* Batched SuccFalsh over n cells, structure-of-arrays ([comp][cell]).
 * T: per-cell temperatures or NULL (then T0 is used).  Outputs may be NULL. */
void or_flash_batch(const or_fluid *f, int64_t n, const double *T, double T0,
                    const double *P, const double *z, double sw,
                    double *x, double *y, double *K, double *L, int8_t *label,
                    double *Zl, double *Zg, int32_t *iters, int nthreads);
/* The same, with the Rachford-Rice iteration count of each cell (rr [n]). */
void or_flash_batch_rr(const or_fluid *f, int64_t n, const double *T, double T0,
                       const double *P, const double *z, double sw,
                       double *x, double *y, double *K, double *L, int8_t *label,
                       double *Zl, double *Zg, int32_t *iters, int32_t *rr, int nthreads);

/* Bubl_Succ (Minor_NR.cpp:491-871).  x: liquid composition (comp[..][0]),
 * ystale: the cell's stored vapour composition (comp[..][1]), which the
 * reference's successive-substitution stage reads.  Returns Psat or -1;
 * xm[0..nc-1] receives the incipient composition left in the global Xm. */
double or_bubble(const or_fluid *f, double T, double P, const double *x,
                 const double *ystale, double *xm);

/* Dew_Succ (Minor_NR.cpp:873-1249).  y: vapour composition. */
double or_dew(const or_fluid *f, double T, double P, const double *y, double *xm);

/* Batched saturation pressure: kind[c] = 1 bubble, -1 dew, 0 skip (psat = -1). */
void or_psat_batch(const or_fluid *f, int64_t n, const double *T, double T0,
                   const double *P, const int8_t *kind, const double *x,
                   const double *y, double *psat, double *xm, int nthreads);

/* PhaseCtrl / fPhaseCtrl (Do_Cycle.cpp:1558-1722) over n cells of one
 * continuum; label, x, y, S ([3*n]: Sw, So, Sg) updated in place. */
void or_phase_ctrl(const or_fluid *f, int64_t n, const double *T, double T0,
                   int8_t *label, const double *Po, const double *Pg,
                   double *x, double *y, double *S, double *psat, int nthreads);

/* Dense LU with partial pivoting + solve, the arithmetic LAPACK dgesv performs
 * (called through MKLS, MKL_Solver.Cpp:66-90).  A is n*n row-major and is
 * overwritten.  Returns 0, or k>0 if U(k-1,k-1) is exactly zero. */
int or_lu_solve(int n, double *A, const double *rhs, double *sol);


/* ------------------------------------------------------------ assembly */
/* Component data beyond the flash (Data_Input.cpp:657-951 library fields). */
typedef struct {
    const double *MW;        /* [nc] g/mol */
    const double *Vc;        /* [nc] cm^3/mol (the deck unit; EOS_Init divides by 1e6) */
    const double *Zc;        /* [nc] */
    const double *parachor;  /* [nc] */
} or_fluid_ext;

/* Wells (deck WELLS, Data_Input.cpp:1185-1330; run state simTime and
 * WellCondition).  welli rows as Data_Input stores them: type 0..8, then the
 * 0-based block Ix, Iy and the first and last completed layer; wellf rows of
 * nc + 6 values in the type's order (Pwf or rate, limit, skin, rw, for types
 * 3 and 8 the nc injection mole fractions, then start and end time, s). */
typedef struct {
    int32_t nw;
    const int32_t *welli;    /* [nw*5] */
    const double *wellf;     /* [nw*(nc+6)] */
    const int8_t *condition; /* [nw] WellCondition, or NULL = all false */
    double sim_time;         /* simTime, s */
    double re_f;             /* Re_f, half fracture length, m */
} or_wells;

/* Grid, rock and run constants (deck GRID, DI/DJ/DK, POR/FPOR, PERM*, TOR,
 * REFP/CPOR/DCPOR, SIGMA, Lf, CDIF, RESTEMP, DT, WATERPROPS, SWT/SGT). */
typedef struct {
    const double *dx, *dy, *dz;          /* [nx], [ny], [nz] m */
    const double *phi_m, *phi_f;         /* [n] porosity */
    const double *perm_m, *perm_f;       /* [3*n] (i, j, k) mD */
    const double *tor;                   /* [n] tortuosity, NULL = 1 */
    const double *swt, *sgt;             /* [rows*4]: SAT, KR, KRO, PC */
    double refP, cpor, dcpor, sigma, lf, cdiff, T, dt, wat_ro, wat_mu;
    int32_t nx, ny, nz, const_diff, nswt, nsgt;
    const or_wells *wells;               /* NULL = no wells */
} or_grid;

/* Per-cell state of both continua (n = nx*ny*nz, cell = iz*nx*ny+iy*nx+ix). */
typedef struct {
    const int8_t *label_m, *label_f;     /* [n] 1 L, 0 2ph, -1 G */
    const double *Po_m, *Po_f;           /* [n] oil pressure, Pa */
    const double *x_m, *y_m, *x_f, *y_f; /* [nc*n] */
    const double *S_m, *S_f;             /* [3*n] Sw, So, Sg */
    const double *rho_prev_m;            /* [2*n] matrix oil/gas density of the
                                            previous sweep (read by CalcIFT), or NULL */
    const double *pre_m, *pre_f;         /* [(nc+1)*n] preProp / fpreProp (Pre_Mul), or
                                            NULL = computed from this state (then the
                                            accumulation terms vanish, as at X0) */
    double pcoeff;                       /* PCoeff (CalcPCoeff) */
} or_state;

typedef struct {
    const int8_t *label_m, *label_f;
    double *Po_m, *Po_f, *x_m, *y_m, *x_f, *y_f, *S_m, *S_f;
    const double *rho_prev_m;
    const double *pre_m, *pre_f;
    double pcoeff;
} or_state_mut;

int or_layout(const or_grid *g, const or_state *s, int nc, int32_t *bsize, int64_t *offset);
int or_residual(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, const or_state *s, double *F);
/* Same rows with every term taken by magnitude: the scale of each row. */
int or_residual_scale(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, const or_state *s, double *F);
int or_props(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, const or_state *s, double *out);
/* The grid's wells: Calc_Re (Minor_NR.cpp:2700-3249) re[nw] and the
 * FlashInjection outputs (Init_Calcs.cpp:1161-1377) inj[nw][2nc+6]; returns nw. */
int or_well_data(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, double *re, double *inj);
/* IFK_Calc's kA/h (Init_Calcs.cpp:21-42): out [2][3][n]. */
int or_iftran(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, double *out);
int64_t or_jacobian(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, const or_state *s,
                    int64_t *rowptr, int32_t *cols, double *vals, int64_t cap);
/* The rows of the listed blocks only (block = k n + c), consecutively in
 * list order; rowptr [sum of the blocks' sizes + 1].  -2 for a bad block. */
int64_t or_jacobian_blocks(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, const or_state *s,
                           const int32_t *blocks, int32_t nblocks, int64_t *rowptr, int32_t *cols,
                           double *vals, int64_t cap);
/* IteratorJacobian (Petsc_Functions.cpp:6898-42073) restated: the reference's
 * hand-written Jacobian, its emitted columns (duplicates summed, sorted).
 * Returns nnz, -1 if cap is too small, -2 if a row emits a column outside
 * the matrix. */
int64_t or_jacobian_ref(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, const or_state *s,
                        int64_t *rowptr, int32_t *cols, double *vals, int64_t cap);
/* The same with the cells, faces and blocks spread over nthreads OpenMP
 * threads (<= 0: all); identical output. */
int64_t or_jacobian_ref_mt(const or_fluid *f, const or_fluid_ext *fe, const or_grid *g, const or_state *s,
                           int64_t *rowptr, int32_t *cols, double *vals, int64_t cap, int nthreads);
int or_unpack(const or_grid *g, or_state_mut *s, int nc, const double *X);
int or_pack(const or_grid *g, const or_state *s, int nc, double *X);

/* ------------------------------------------------------------ restart file */
/* WriteRST / ReadRST (Do_Time_Step.cpp:179-294) into / out of a memory image
 * of or_rst_bytes bytes.  P_m, P_f: [3*n] water, oil, gas pressures of the
 * interior cells (may be NULL when reading); the other arrays as or_state.
 * Both return the bytes written / read. */
int64_t or_rst_bytes(int nx, int ny, int nz, int nc);
int64_t or_rst_write(int nx, int ny, int nz, int nc, const double *P_m, const double *P_f, const double *S_m,
                     const double *x_m, const double *y_m, const int8_t *label_m, const double *S_f,
                     const double *x_f, const double *y_f, const int8_t *label_f, double sim_time, double dt,
                     int32_t tsm_ctrl, double smethq, uint8_t *img);
int64_t or_rst_read(int nx, int ny, int nz, int nc, const uint8_t *img, double *P_m, double *P_f, double *S_m,
                    double *x_m, double *y_m, int8_t *label_m, double *S_f, double *x_f, double *y_f,
                    int8_t *label_f, double *sim_time, double *dt, int32_t *tsm_ctrl, double *smethq);

#ifdef __cplusplus
}
#endif
#endif

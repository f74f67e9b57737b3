/*
 * osiris_amd.h -- C ABI of the MI355X (gfx950) per-cell Newton hot path of
 * the Osiris compositional dual-porosity simulator.
 *
 * Plain C types only: pointers, sizes, int status codes.  All array
 * arguments of the *_device entry points are device pointers (HBM) owned by
 * the caller; arrays indexed by component are structure-of-arrays,
 * element (i, c) at [i * n + c] for component i of cell c.  `stream` is a
 * hipStream_t passed as void* (NULL = the default stream); calls are
 * asynchronous with respect to the host unless stated otherwise.
 *
 * Status: 0 ok; > 0 count of cells that hit an iteration cap (results still
 * written); < 0 error (see OSR_E_*).  No entry point exits the process.
 * Reference interfaces each call replaces are cited as file:line of the
 * reference repository (compsim/...).
 */
#ifndef OSIRIS_AMD_H
#define OSIRIS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSR_OK 0
#define OSR_E_ARG (-1)      /* bad argument (null pointer, size, nc) */
#define OSR_E_HIP (-2)      /* a HIP runtime call failed */
#define OSR_E_NC (-3)       /* component count not compiled in */

#define OSR_EOS_PR 0
#define OSR_EOS_SRK 1

/* Fluid description (host memory).  Replaces the globals the reference fills
 * from the deck: fluidProp[Nc][COMP_PROPS] (Functions.h:96-111, filled at
 * Data_Input.cpp:572-951) and bic[Nc][Nc] = 1 - k_ij (Data_Input.cpp:977-990). */
typedef struct osr_fluid {
    int32_t nc;              /* 2..12 compiled (osr_supported_nc); 1..16 else OSR_E_NC */
    int32_t eos;             /* OSR_EOS_PR or OSR_EOS_SRK (deck MODEL keyword) */
    const double *Tc;        /* [nc] critical temperature, K */
    const double *Pc;        /* [nc] critical pressure, Pa */
    const double *omega;     /* [nc] acentric factor */
    const double *kij;       /* [nc*nc] binary interaction coefficients k_ij */
} osr_fluid;

typedef struct osr_ctx osr_ctx;

/* Create a context on the current HIP device: uploads the fluid constants
 * (the EOS_Init coefficients, Init_Calcs.cpp:66-97, kept T-independent so
 * cells may carry their own temperature). */
int osr_ctx_create(const osr_fluid *fluid, osr_ctx **out);
int osr_ctx_destroy(osr_ctx *ctx);

/* Pin the context's workspaces (no counterpart in the reference): with
 * frozen = 1, a call that would grow a workspace the context owns -- the
 * flash's park store, the saturation-pressure work list, the phase-control
 * queue -- returns OSR_E_SIZE (launching nothing) instead of freeing it and
 * allocating a larger one, and osr_asm_set_workspace on an assembler of this
 * context returns OSR_E_ARG; frozen = 0 lifts it.  For callers that capture
 * calls of this context in a HIP graph: the graph keeps the workspaces'
 * addresses, so they must not move while it can replay.  Eager calls made
 * while such a graph replays still share its queue counters and race. */
int osr_ctx_freeze_workspaces(osr_ctx *ctx, int frozen);

/* Flash algorithm selector for osr_flash_device. */
#define OSR_FLASH_PLAIN 0        /* one lane per cell, FP64, loop to convergence */
#define OSR_FLASH_PERSISTENT 1   /* FP64 persistent kernel, lanes refilled from a work queue */
#define OSR_FLASH_MIXED 2        /* FP32 successive substitution, FP64 finish (persistent) */

/* Output arrays of osr_flash_device (device pointers; any may be NULL). */
typedef struct osr_flash_out {
    double *x;        /* [nc*n] liquid composition (comp[..][i][0]) */
    double *y;        /* [nc*n] vapour composition (comp[..][i][1]) */
    double *K;        /* [nc*n] final K-values */
    double *L;        /* [n] liquid mole fraction, clamped to [0,1] */
    int8_t *label;    /* [n] 1 liquid, 0 two-phase, -1 gas (phaseStat) */
    double *Zl;       /* [n] liquid compressibility factor */
    double *Zg;       /* [n] vapour compressibility factor */
    double *so;       /* [n] oil saturation (1-sw) split by Z-weighted L */
    double *sg;       /* [n] gas saturation */
    double *rho_l;    /* [n] liquid molar density, mol/m^3 */
    double *rho_g;    /* [n] vapour molar density, mol/m^3 */
    int32_t *iters;   /* [n] outer successive-substitution iterations */
    int32_t *rr_iters;/* [n] total Rachford-Rice Newton iterations */
} osr_flash_out;

/* Two-phase flash of n independent cells: SuccFalsh (Init_Calcs.cpp:571-738)
 * as driven by AllFlash (Init_Calcs.cpp:949-962).  T: per-cell temperature
 * [n] or NULL to use T0 everywhere; P: [n] pressure, Pa; z: [nc*n] feed.
 * sw: water saturation used for the so/sg split.  `mode` is OSR_FLASH_*.
 * Returns the number of cells that hit the 500-iteration cap, or < 0. */
int osr_flash_device(osr_ctx *ctx, int64_t n, const double *T, double T0,
                     const double *P, const double *z, double sw,
                     const osr_flash_out *out, int mode, void *stream);

/* Allocation-free variant of osr_flash_device, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as osr_flash_device; the parked-cell store
 * of the persistent and mixed forms lives in a fixed 64 MiB workspace per
 * device (allocated when the library is loaded, kept until the process ends)
 * instead of a buffer grown on demand.  Calls no hipMalloc/hipFree.  Returns
 * OSR_E_SIZE, launching nothing, when n exceeds
 * osr_flash_noalloc_max_cells(nc, mode).  Every call on one device shares
 * that workspace and the context's queue counters: calls that overlap on one
 * device (several streams, or several contexts on the same device) race. */
#define OSR_E_SIZE (-4)     /* input larger than a fixed-size workspace allows */
int osr_flash_device_noalloc(osr_ctx *ctx, int64_t n, const double *T, double T0,
                             const double *P, const double *z, double sw,
                             const osr_flash_out *out, int mode, void *stream);

/* Largest n osr_flash_device_noalloc accepts for nc components and `mode`:
 * 8 Mi doubles / (nc + 5) persistent, / (2 nc + 9) mixed, unbounded (INT64_MAX)
 * plain.  Host-only; OSR_E_NC / OSR_E_ARG for an unsupported nc / mode. */
int64_t osr_flash_noalloc_max_cells(int32_t nc, int32_t mode);

/* Saturation pressure of n cells: kind[c] = 1 runs Bubl_Succ
 * (Minor_NR.cpp:491-871) on liquid x with stored vapour y, kind[c] = -1 runs
 * Dew_Succ (Minor_NR.cpp:873-1249) on vapour y, 0 skips (psat = -1).  P is
 * the cell pressure the iteration starts from.  psat[c] = -1 when the
 * reference routine returns -1 (trivial or non-convergent); xm [nc*n]
 * (may be NULL) receives the incipient-phase composition (the global Xm). */
int osr_psat_device(osr_ctx *ctx, int64_t n, const double *T, double T0,
                    const double *P, const int8_t *kind, const double *x,
                    const double *y, double *psat, double *xm, void *stream);

/* Phase-state control of n cells of one continuum after an accepted Newton
 * update: PhaseCtrl / fPhaseCtrl (Do_Cycle.cpp:1558-1643 / 1644-1722).  Gas
 * cells with Pg <= 0.95 x dew pressure and liquid cells with Po <= 0.95 x
 * bubble pressure become two-phase (incipient saturation 1e-3, incipient
 * composition from the saturation routine); other cells with So <= 0 (Sg <= 0)
 * become gas (liquid); saturations are renormalised.  label [n], x, y [nc*n]
 * and S [3*n] (Sw, So, Sg) are updated in place; Po, Pg [n] are the oil and
 * gas phase pressures; psat [n] (may be NULL) receives the saturation
 * pressure evaluated (-1 where none).  Uses a library-owned workspace. */
int osr_phase_ctrl_device(osr_ctx *ctx, int64_t n, const double *T, double T0, int8_t *label,
                          const double *Po, const double *Pg, double *x, double *y, double *S,
                          double *psat, void *stream);

/* ================================================================ assembly
 * Residual and Jacobian of the dual-porosity model in the reference's layout:
 * all matrix blocks then all fracture blocks, cell index iz*nx*ny+iy*nx+ix,
 * block size nc+3 (single-phase) or 2nc+4 (two-phase) (PJac,
 * Init_Calcs.cpp:1013-1116; MatSetVariableBlockSizes, PetscAUX.cpp:160-163).
 * Unknowns per block in InitGuess2 order (PetscAUX.cpp:892-957): liquid
 * [x(nc), P/PCoeff, Sw, So]; gas [y(nc), P/PCoeff, Sw, Sg]; two-phase
 * [x(nc), y(nc), P/PCoeff, Sw, So, Sg].  Rows in IteratorFunction order
 * (Petsc_Functions.cpp:3446-6870): two-phase 0..nc-1 fugacity, nc..2nc-2
 * component balances 0..nc-2, 2nc-1 total hydrocarbon, 2nc water, then
 * sum x - 1, sum y - 1, sum S - 1; single-phase 0..nc-2 components, nc-1
 * total, nc water, sum x|y - 1, sum S - 1. */

/* Component data beyond the flash (library fields of Data_Input.cpp:657-951). */
typedef struct osr_fluid_ext {
    const double *MW;        /* [nc] g/mol */
    const double *Vc;        /* [nc] cm^3/mol (deck unit; EOS_Init divides by 1e6) */
    const double *Zc;        /* [nc] */
    const double *parachor;  /* [nc] */
} osr_fluid_ext;

/* Grid, rock and run constants (host memory; deck GRID, DI/DJ/DK, POR/FPOR,
 * PERMI..K/FPERMI..K, TOR, REFP/CPOR/DCPOR, SIGMA, Lf, CDIF, RESTEMP, DT,
 * WATERPROPS, SWT/SGT -- Data_Input.cpp:13-1349). */
typedef struct osr_grid {
    const double *dx, *dy, *dz;          /* [nx], [ny], [nz] m */
    const double *phi_m, *phi_f;         /* [n] porosity (matrix, fracture) */
    const double *perm_m, *perm_f;       /* [3*n] (i, j, k) mD */
    const double *tor;                   /* [n] tortuosity, NULL = 1 */
    const double *swt, *sgt;             /* [rows*4] SAT, KR, KRO, PC(Pa); <= 32 rows */
    double refP, cpor, dcpor;            /* rock compressibility */
    double sigma;                        /* SIGMA (shape-factor multiplier) */
    double lf;                           /* Lf, half fracture opening, m */
    double cdiff;                        /* CDIF value */
    double T;                            /* reservoir temperature, K */
    double dt;                           /* time step, s */
    double wat_ro, wat_mu;               /* WATERPROPS */
    int32_t nx, ny, nz;
    int32_t const_diff;                  /* 1 when the deck has CDIF */
    int32_t nswt, nsgt;
} osr_grid;

/* Cell state of both continua (device pointers; n = nx*ny*nz). */
typedef struct osr_state_dev {
    const int8_t *label_m, *label_f;     /* [n] phaseStat / fphaseStat */
    const double *Po_m, *Po_f;           /* [n] oil pressure P[..][1], Pa */
    const double *x_m, *y_m, *x_f, *y_f; /* [nc*n] comp[..][i][0..1] */
    const double *S_m, *S_f;             /* [3*n] sat[..][0..2] */
    const double *rho_prev_m;            /* [2*n] matrix oil/gas molar density of the
                                            previous property sweep (CalcIFT reads them
                                            before Calc_Visco refreshes them), or NULL:
                                            then a first sweep with zero densities
                                            supplies them */
    const double *pre_m, *pre_f;         /* [(nc+1)*n] preProp/fpreProp (Pre_Mul) or NULL =
                                            this state's accumulation (the residual at X0) */
    double pcoeff;                       /* PCoeff (CalcPCoeff, Init_Calcs.cpp:1118-1130) */
} osr_state_dev;

/* The same state with writable arrays (for osr_unpack_device). */
typedef struct osr_state_mut_dev {
    const int8_t *label_m, *label_f;
    double *Po_m, *Po_f, *x_m, *y_m, *x_f, *y_f, *S_m, *S_f;
    const double *rho_prev_m, *pre_m, *pre_f;
    double pcoeff;
} osr_state_mut_dev;

typedef struct osr_asm osr_asm;

/* Assembly context for one grid: derives the face kA/h (IFK_Calc,
 * Init_Calcs.cpp:21-42), block heights (CalcBlockHeight, :993-1011), the
 * low-pressure viscosities (Tr_Mu, :48-64) and the binary diffusion table
 * (Mass_Tran.cpp:38-45), and allocates the property workspace. */
int osr_asm_create(osr_ctx *ctx, const osr_fluid_ext *fe, const osr_grid *grid, osr_asm **out);
int osr_asm_destroy(osr_asm *a);

/* Jacobian workspace (no counterpart in the reference, whose PETSc Mat is the
 * only buffer).  The Jacobian and reduced-system calls need scratch of
 * *bytes (per block, the larger of (16+2nc)(2nc+4) doubles for the CSR
 * Jacobian and (2nc+4)(nc+3) + (20+4nc)(nc+2) + 8(nc+1) for the direct
 * reduced system: 8.8 KB at nc = 10); by default each
 * osr_asm allocates its own on first use.  osr_asm_set_workspace hands it
 * a caller-owned device buffer instead (256-byte aligned, at least *bytes;
 * NULL returns to a library-owned one), so assemblers whose calls are
 * serialised on one stream -- e.g. many grids replayed in one graph -- can
 * share one buffer sized for the largest.  Calls on different streams must
 * not share it. */
int osr_asm_workspace_size(osr_asm *a, size_t *bytes);
int osr_asm_set_workspace(osr_asm *a, void *ptr, size_t bytes);

/* Device pointer to the assembler's cell-property table, [16+2nc][2n] in the
 * field order of osr_props_device, rewritten by every residual/Jacobian call
 * on the stream it runs on (e.g. Pg = table + 3*2n for PhaseCtrl).  Valid
 * until osr_asm_destroy.  The table replaces the per-cell property globals
 * the reference's sweeps fill (blockFProps and kin, Do_Cycle.cpp:99-437). */
int osr_asm_props_table(osr_asm *a, const double **props);

/* Wells (deck WELLS, Data_Input.cpp:1196-1330, with the run's simTime and
 * WellCondition).  welli rows as the reference stores them: type 0..8, then
 * the 0-based block Ix, Iy and the first and last completed layer; wellf rows
 * of nc + 6 values in the type's order (Pwf or rate, limit, skin, rw, for
 * types 3 and 8 the nc injection mole fractions, then start and end time,
 * s).  Host memory. */
typedef struct osr_wells {
    int32_t nw;
    const int32_t *welli;    /* [nw*5] */
    const double *wellf;     /* [nw*(nc+6)] */
    const int8_t *condition; /* [nw] WellCondition, NULL = all false */
    double sim_time;         /* simTime, s */
    double re_f;             /* Re_f, m */
} osr_wells;

/* Completes the wells in the assembler's grid: from then on the residual,
 * Jacobian and reduced-system calls add the well terms of IteratorFunction
 * (Petsc_Functions.cpp:164-697 fracture blocks, :3505-4040 matrix blocks;
 * the rates Qw, Qo, Qg of the last active well of a block enter its
 * component rows as x_i Qo + y_i Qg -- an injector's stream compositions
 * instead, :787-808 --, its total row as Qo + Qg and its water row as Qw).
 * Computes each well's equivalent radius (Calc_Re, Minor_NR.cpp:2700-3249;
 * the matrix blocks use it, the fracture blocks Re_f) on the host and the
 * injection stream's flash (FlashInjection, Init_Calcs.cpp:1161-1377) of
 * types 3 and 8 on the device, on `stream`; returns when both are done.
 * The allocation sums over a well's layers (SSo1 ...) read the other
 * blocks' property-table values, so the Jacobian takes the derivative with
 * respect to the block's own unknowns, as the reference's does.  Call again
 * whenever simTime or WellCondition change (a call with the same number of
 * wells and completed blocks keeps the device tables, so a graph captured
 * earlier reads the new values); nw = 0 removes the wells.  The reference-
 * mode Jacobian (osr_jacobian_ref_device) takes the reference's own well
 * entries instead: PCoeff x the productivity indices in the block's P
 * column and, for a producer, -Qo, -Qg in its x_i, y_i columns. */
int osr_asm_set_wells(osr_asm *a, const osr_wells *w, void *stream);

/* Per-well results of the last osr_asm_set_wells (host outputs, either may
 * be NULL): re[nw] the equivalent radius; inj[nw*(2nc+6)] the injection
 * stream's x (nc), y (nc), oil and gas fraction, densities (mol/m^3) and
 * viscosities -- the outputs of FlashInjection; zeros for the other types. */
int osr_asm_well_data(osr_asm *a, double *re, double *inj);

/* Block offsets of the layout for the state's labels (PJac).  offsets
 * (device, may be NULL) receives 2n+1 entries; *nrows (host) = MatrixSize.
 * Must be called after any label change, before residual/Jacobian. */
int osr_asm_layout(osr_asm *a, const osr_state_dev *s, int64_t *offsets, int64_t *nrows, void *stream);

/* Residual F[nrows] (device): the rows IteratorFunction writes for every
 * block (Petsc_Functions.cpp:10-6896), after the property sweep
 * UpdateDependencies runs (PetscAUX.cpp:500-547).  Replaces the
 * FormFunction row loop (PetscAUX.cpp:59-104). */
int osr_residual_device(osr_asm *a, const osr_state_dev *s, double *F, void *stream);

/* Jacobian row pointer (device, nrows+1) and, on the host, the nnz.
 * Row pattern: balance and total rows couple to every unknown of the own
 * block, the six face neighbours and the other continuum's cell; the water
 * row to the own block and the P, Sw columns of the others; fugacity and
 * constraint rows to the own block; zero-porosity rows hold the diagonal. */
int osr_jacobian_rowptr_device(osr_asm *a, const osr_state_dev *s, int64_t *rowptr, int64_t *nnz, void *stream);

/* Jacobian values and column indices (device CSR, columns ascending within a
 * row; duplicates never occur).  The values are the derivatives of the
 * residual above in the unknowns listed, with preProp and the previous-sweep
 * densities held fixed; zero-porosity rows are identity rows
 * (Petsc_Functions.cpp:42051-42056).  Replaces the FormJacobian row loop
 * (PetscAUX.cpp:106-173, IteratorJacobian Petsc_Functions.cpp:6898-42073). */
int osr_jacobian_device(osr_asm *a, const osr_state_dev *s, const int64_t *rowptr, int32_t *cols,
                        double *vals, void *stream);

/* The reference's own Jacobian: IteratorJacobian (Petsc_Functions.cpp:
 * 6898-42073) as the FormJacobian row loop (PetscAUX.cpp:106-173) assembles
 * it with MatSetValues(ADD_VALUES) into PJac's matrix (Init_Calcs.cpp:
 * 1013-1116): the columns each row's code emits (own block, six neighbours
 * and the other continuum's block whole for the balance rows, less a gas
 * block's own Sw; {P, Sw} of the same blocks for the water row; x, y, P of
 * the own block for fugacity rows; the constraint rows' columns with the
 * diagonal placeholder of :7004-7006), ascending, and the values of its
 * hand rules: dRelPerm with the sign clamps and frozen IFT (Do_Cycle.cpp:
 * 163-256), no capillary-pressure derivatives, dE-based dZ, DiffVisco and
 * 1e4-scaled DifDiffusion (Mass_Tran.cpp:107-399), its sign conventions for
 * neighbour saturation and transfer columns and its per-face copy-paste
 * assignments.  Same state and layout contract as osr_jacobian_device.
 * osr_jacobian_ref_rowptr_device: rowptr [nrows+1] (device), *nnz (host). */
int osr_jacobian_ref_rowptr_device(osr_asm *a, const osr_state_dev *s, int64_t *rowptr, int64_t *nnz,
                                   void *stream);
int osr_jacobian_ref_device(osr_asm *a, const osr_state_dev *s, const int64_t *rowptr, int32_t *cols,
                            double *vals, void *stream);

/* UpdateAll (PetscAUX.cpp:176-498): unknown vector X [nrows] (device) ->
 * state, in place: mole fractions clamped to [0, 1] and renormalised, the
 * pressure limiter (+-0.01 Pa when the new pressure exceeds the old one by
 * more than 1 %, PetscAUX.cpp:227), the 5 MPa floor, saturations clamped and
 * renormalised.  Layout from the last osr_asm_layout. */
int osr_unpack_device(osr_asm *a, const osr_state_mut_dev *s, const double *X, void *stream);

/* InitGuess2 (PetscAUX.cpp:878-964): state -> unknown vector X [nrows]. */
int osr_pack_device(osr_asm *a, const osr_state_dev *s, double *X, void *stream);

/* Reduced (nc+1)^2 block-CSR of the Newton system (no counterpart in the reference, whose PETSc
 * solve takes the full AIJ matrix with point-block Jacobi, PetscAUX.cpp:
 * 628-646).  Each block's local rows (fugacity and constraint rows, which
 * involve its own unknowns only) are eliminated:
 *   ds = -S dp - t,   S = J_ls^-1 J_lp,   t = J_ls^-1 F_l,
 *   sum_d (J_mp,ad - J_ms,ad S_d) dp_d = -F_m,a + sum_d J_ms,ad t_d.
 * Primary unknowns per block (nc+1, in this order): two-phase x_0..x_{nc-3},
 * P, Sw, So; liquid x_0..x_{nc-2}, P, Sw; vapour y_0..y_{nc-2}, P, Sw.  Rows:
 * the nc+1 balance rows (nc-1 components, total, water) in reference order.
 * Zero-porosity blocks reduce to an identity block with zero right-hand side.
 *
 * osr_bsr_reduced_rowptr_device: browptr [2n+1] (device), *nnzb (host).
 * osr_jacobian_bsr_reduced_device: from this state's residual F and CSR
 * (rowptr, vals of osr_jacobian_device) -> bcols [nnzb] (block index),
 * bvals [nnzb][(nc+1)^2] row-major, brhs [2n][nc+1], and the elimination
 * factors S [2n][nc+3][nc+1], t [2n][nc+3] (rows beyond a block's local-row
 * count are zero).
 * osr_bsr_reduced_expand_device: the full Newton update du [nrows] (reference
 * unknown layout) from the reduced solution dp [2n][nc+1]. */
int osr_bsr_reduced_rowptr_device(osr_asm *a, const osr_state_dev *s, int64_t *browptr, int64_t *nnzb,
                                  void *stream);
int osr_jacobian_bsr_reduced_device(osr_asm *a, const osr_state_dev *s, const double *F, const int64_t *rowptr,
                                    const double *vals, const int64_t *browptr, int32_t *bcols, double *bvals,
                                    double *brhs, double *S, double *t, void *stream);
int osr_bsr_reduced_expand_device(osr_asm *a, const osr_state_dev *s, const double *S, const double *t,
                                  const double *dp, double *du, void *stream);

/* The same reduced system (bcols, bvals, brhs, S, t as above, browptr from
 * osr_bsr_reduced_rowptr_device) assembled straight from the state without
 * the full CSR: the Jacobian is linear in the seed direction, so column q of
 * a reduced block is the forward-mode derivative of the balance rows along
 * e_{primary q} - S_d e_{secondaries}, and the right-hand-side correction
 * the derivative along t_d on the secondaries.  F is this state's residual
 * (osr_residual_device).  No counterpart in the reference (see above); it is
 * the form that fits one GPU at the largest grids, where the full AIJ matrix
 * of FormJacobian (PetscAUX.cpp:106-173) would not. */
int osr_jacobian_bsr_reduced_direct_device(osr_asm *a, const osr_state_dev *s, const double *F,
                                           const int64_t *browptr, int32_t *bcols, double *bvals, double *brhs,
                                           double *S, double *t, void *stream);

/* Adjoint of the residual with respect to log-permeability (no counterpart
 * in the reference): g_perm_m[d][c] = sum_r Fbar_r dF_r / d ln perm_m[d][c],
 * likewise g_perm_f, for a cotangent Fbar [nrows] in the row layout of the
 * last osr_asm_layout (e.g. F itself on the fracture balance rows and 0
 * elsewhere gives the gradient of 1/2 |F_frac,mass|^2).  Permeability enters
 * F through the face kA/h of IFK_Calc (Init_Calcs.cpp:21-42) and the
 * transfer shape factor (ShapeFactor_2, :970-983); state held fixed.
 * Either output may be NULL. */
int osr_residual_perm_vjp_device(osr_asm *a, const osr_state_dev *s, const double *Fbar, double *g_perm_m,
                                 double *g_perm_f, void *stream);

/* Cell properties after the sweep, [16+2nc][2n] (device): porpor, Pw, Po, Pg,
 * IFT, rho_o, rho_g, mu_o, mu_g, gamma_o, gamma_g, krw, kro, krg, Pcw, Pcg,
 * D_o[nc], D_g[nc]. */
int osr_props_device(osr_asm *a, const osr_state_dev *s, double *out, void *stream);

/* Launch-shape knobs (no counterpart in the reference; set by the
 * autotuner, osiris_amd/tune.py): "flash.refill_min" (1..64: idle lanes a
 * wave waits for before refilling from the work queue), "flash.park_after"
 * (0..500: passes after which the persistent flash hands a cell to the
 * lane-group kernel; 0 = never), "flash.blocks_per_cu" (0 = occupancy-
 * derived; else 256-thread blocks per CU of the persistent flash grid), and
 * for OSR_FLASH_MIXED "flash.mixed_warm_iters" (FP64 passes before the FP32
 * stage, 1..100), "flash.mixed_tol32" (FP32 stage stop on max |K'/K - 1|) and
 * "flash.mixed_max32" (FP32 stage pass cap); for the reduced system straight
 * from the state "reduced.local_rows" (1: the local rows from closed-form
 * cubic-EOS derivatives and eliminated in one lane group per block -- the
 * default; 0: forward mode per (block, unknown), then the elimination) and
 * "assembly.xcd_remap" (1: each of the chip's 8 XCDs takes a contiguous
 * eighth of the reduced-system kernels' workgroups; 0: dispatch order),
 * "reduced.dprops_wg" and "reduced.own_wg" (64, 128 or 256: workgroup
 * sizes of the reduced system's direction sweep and balance-row kernels). */
int osr_tune_set(osr_ctx *ctx, const char *key, double value);

/* Tangent of the flash at a converged state (no counterpart in the
 * reference; the sensitivities of the phase split SuccFalsh computes,
 * Init_Calcs.cpp:571-738).  By the implicit-function theorem on
 * g(ln K; z, P) = ln K - ln phiL(x) + ln phiG(y) = 0 with the exact
 * Rachford-Rice split.  Inputs: T (or T0), P, z [nc][n] and the state the
 * flash returned, K [nc][n] and label [n]; a perturbation dP [n], dz [nc][n]
 * (NULL = 0).  Outputs dx, dy [nc][n], dL [n] (NULL = not written).
 * Single-phase cells: the phase composition is z, L is constant, the absent
 * phase's composition moves with K held fixed. */
int osr_flash_tangent_device(osr_ctx *ctx, int64_t n, const double *T, double T0, const double *P,
                             const double *z, const double *K, const int8_t *label, const double *dP,
                             const double *dz, double *dx, double *dy, double *dL, void *stream);

/* Adjoint (vector-Jacobian product) of the same map: cotangents xbar, ybar
 * [nc][n], Lbar [n] (NULL = 0) -> Pbar [n], zbar [nc][n]. */
int osr_flash_adjoint_device(osr_ctx *ctx, int64_t n, const double *T, double T0, const double *P,
                             const double *z, const double *K, const int8_t *label, const double *xbar,
                             const double *ybar, const double *Lbar, double *Pbar, double *zbar,
                             void *stream);

/* Stage counts of the last persistent/mixed flash on this context (no
 * counterpart in the reference), out[0..nout): [0] cells handed on after the
 * warm FP64 passes, [1] cells that reached the FP64 Newton polish, [2] cells
 * finished by FP64 successive substitution from their parked state (the FP32
 * stage or the Newton declined them), [3] Newton cells accepted, then the
 * Newton's other exits: [4] L within 1e-3 of 0 or 1, [5] non-finite
 * residual, [6] singular factorisation, [7] step > 1 or non-finite,
 * [8] iteration cap, [9] rejected by the final reference pass.  Synchronises
 * the stream. */
int osr_flash_stats(osr_ctx *ctx, int64_t *out, int32_t nout, void *stream);

/* ====================================================== linear solve
 * BiCGStab with block-Jacobi preconditioning on a block-CSR system with
 * square blocks of size bs (the reduced (nc+1)^2 system: bs = nc + 1 in
 * {3, 4, 5, 6, 7, 9, 11, 13}).  Replaces the reference's linear solve inside
 * SNESSolve -- KSPSetType(SolverType = "bcgs"), PCSetType(PCVPBJACOBI),
 * KSPSetTolerances(KSPRelTol, KSPAbsTol, KSPDivTol, MaxIters)
 * (PetscAUX.cpp:628-646, 669; defaults 1e-18, 1e-50, 1e3, 5000 from
 * Data_Input.cpp:232-266) -- moved onto the reduced system: the
 * preconditioner inverts the (nc+1)^2 diagonal blocks of the reduced matrix
 * rather than the variable-size diagonal blocks of the full AIJ matrix.
 * Convergence is tested on the unpreconditioned residual norm:
 * |r| <= max(rtol |b|, atol) converged, |r| >= dtol |r0| diverged.
 *
 * osr_ksp_setup_device keeps the (device) matrix pointers and inverts the
 * diagonal blocks; it returns the number of singular diagonal blocks (>= 0,
 * the solve then yields non-finite values) or < 0.  osr_ksp_solve_device
 * solves A x = b (x: initial guess if x0_nonzero, else zeroed) and
 * synchronises the stream once per iteration to test convergence. */
typedef struct osr_ksp osr_ksp;
typedef struct osr_ksp_opts {
    double rtol, atol, dtol;
    int32_t maxit;
    int32_t x0_nonzero;
} osr_ksp_opts;
#define OSR_KSP_CONVERGED_RTOL 2
#define OSR_KSP_CONVERGED_ATOL 3
#define OSR_KSP_DIVERGED_ITS (-3)
#define OSR_KSP_DIVERGED_DTOL (-4)
#define OSR_KSP_DIVERGED_BREAKDOWN (-5)
#define OSR_KSP_DIVERGED_NANORINF (-9)
typedef struct osr_ksp_result {
    int32_t iterations;
    int32_t reason;       /* OSR_KSP_* (the KSPConvergedReason codes) */
    double rnorm, bnorm, r0norm;
} osr_ksp_result;
int osr_ksp_create(int32_t bs, int64_t nbrows, osr_ksp **out);
int osr_ksp_destroy(osr_ksp *k);
int osr_ksp_setup_device(osr_ksp *k, const int64_t *browptr, const int32_t *bcols, const double *bvals,
                         void *stream);
int osr_ksp_solve_device(osr_ksp *k, const double *b, double *x, const osr_ksp_opts *opts,
                         osr_ksp_result *res, void *stream);

/* ====================================================== Newton step
 * One time step's nonlinear solve as the reference runs it: SolveNonLinear
 * (PetscAUX.cpp:566-730) with SNESSetType(SNESKSPONLY) -- one Jacobian at X0,
 * one linear solve J Y = F and the step X1 = X0 - Y -- its acceptance test
 * (PetscAUX.cpp:703), and for an accepted step UpdateAll(X, true)
 * (PetscAUX.cpp:176-507): X -> state (osr_unpack_device's clamps and
 * limiter), PhaseCtrl and fPhaseCtrl with the gas pressures of the last
 * property sweep (that of X0), CalcPCoeff (Init_Calcs.cpp:1118-1130), PJac
 * and the property sweep of the new state.
 *
 * On the device: InitGuess2 packs X0, the residual and the reduced
 * (nc+1)^2 system are assembled straight from the state
 * (osr_jacobian_bsr_reduced_direct_device), BiCGStab with block Jacobi
 * solves it (osr_ksp_*, opts->ksp as KSPSetTolerances), and the full update
 * du = -Y is expanded from its solution.  Accepted, as PetscAUX.cpp:703,
 * when the SNES reason (SNESKSPONLY: SNES_CONVERGED_ITS = 5 when the linear
 * solve converged, SNES_DIVERGED_LINEAR_SOLVE = -3 otherwise) and the KSP
 * reason are positive and vnorm = |Y|_1 / MatrixSize is finite and at most
 * opts->tolmain (TOLMAIN, Functions.h:241, 1e-2).  A rejected step leaves
 * the state as it was (the caller's time-step control then shortens Dt, as
 * TimeStepCtrl does after SolveNonLinear returns 0).  An accepted step
 * updates s in place, s->pcoeff included, and writes to rho_next (if not
 * NULL) the matrix oil/gas densities [2n] of the new state's sweep: the
 * rho_prev_m of the next step (CalcIFT reads them); rho_next may alias
 * s->rho_prev_m.  pre_m/pre_f must be NULL (Pre_Mul recomputes preProp in
 * every sweep, so the accumulation of X0 is the reference's).
 *
 * osr_newton_create allocates the step's vectors for the assembler's grid
 * (the reduced blocks on the first step).  osr_newton_solution_update
 * copies the last step's update du [*nrows] into du (device, may be NULL
 * to query *nrows), asynchronously on the stream: X1 = X0 + du
 * (the reference's SNESGetSolutionUpdate, PetscAUX.cpp:681, is -du).  A
 * step synchronises the stream (layout counts, once per BiCGStab
 * iteration, the acceptance test). */
typedef struct osr_newton osr_newton;
typedef struct osr_newton_opts {
    osr_ksp_opts ksp;        /* KSPRelTol, KSPAbsTol, KSPDivTol, MaxIters */
    double tolmain;          /* TOLMAIN */
} osr_newton_opts;
#define OSR_SNES_CONVERGED_ITS 5
#define OSR_SNES_DIVERGED_LINEAR_SOLVE (-3)
typedef struct osr_newton_result {
    int32_t accepted;        /* SolveNonLinear's return value: 1 state updated, 0 not */
    int32_t snes_reason;     /* OSR_SNES_* (PETSc's SNESConvergedReason codes) */
    int32_t singular_blocks; /* singular diagonal blocks of the reduced system */
    int32_t reserved;
    osr_ksp_result ksp;
    double vnorm;            /* |du|_1 / nrows */
    int64_t nrows;           /* MatrixSize of the step (the old labels) */
    int64_t nrows_next;      /* MatrixSize of the new state (PJac); nrows if rejected */
} osr_newton_result;
/* Cell state updated in place by the step (labels included). */
typedef struct osr_state_rw_dev {
    int8_t *label_m, *label_f;
    double *Po_m, *Po_f, *x_m, *y_m, *x_f, *y_f, *S_m, *S_f;
    const double *rho_prev_m;
    double pcoeff;
} osr_state_rw_dev;
int osr_newton_create(osr_asm *a, osr_newton **out);
int osr_newton_destroy(osr_newton *w);
int osr_newton_step_device(osr_newton *w, osr_state_rw_dev *s, const osr_newton_opts *opts, double *rho_next,
                           osr_newton_result *res, void *stream);
int osr_newton_solution_update(osr_newton *w, double *du, int64_t *nrows, void *stream);

/* ====================================================== restart file
 * The reference's restart file (WriteRST / ReadRST, Do_Time_Step.cpp:179-294,
 * "test.rst") as a byte image in device memory, byte for byte the file:
 * P[Nx+2][Ny+2][Nz+2][3] and fP (water, oil, gas pressure; the ghost halo
 * holds 0), then per matrix cell and per fracture cell, Ix outer and Iz
 * inner, sat[3], (comp[i][0], comp[i][1]) for i < nc and phaseStat (int8),
 * then simTime, Dt (double), TSMCtrl (int32), SMethQ (double); little-endian,
 * unpadded.  osr_rst_bytes gives the image size (= the file size).
 * osr_rst_write_device takes the state and the property table of its sweep
 * for the water and gas pressures (osr_asm_props_table after a residual,
 * Jacobian or accepted step, or the output of osr_props_device).
 * osr_rst_read_device fills the state (labels, Po, x, y, S) and the scalars
 * (it synchronises the stream) and, when phase_p is not NULL, the water and
 * gas pressures [4][n] (Pw_m, Pg_m, Pw_f, Pg_f), which the reference reads
 * and then recomputes (UpdateDependencies after ReadRST, PrimeSim.cpp:84-85).
 * img must be 8-byte aligned; moving the image to and from the file is the
 * caller's (one contiguous transfer). */
typedef struct osr_rst_scalars {
    double sim_time, dt;     /* simTime, Dt */
    int32_t tsm_ctrl;        /* TSMCtrl */
    int32_t reserved;
    double smethq;           /* SMethQ */
} osr_rst_scalars;
int osr_rst_bytes(int32_t nx, int32_t ny, int32_t nz, int32_t nc, int64_t *bytes);
int osr_rst_write_device(osr_asm *a, const osr_state_dev *s, const double *props, const osr_rst_scalars *sc,
                         uint8_t *img, void *stream);
int osr_rst_read_device(osr_asm *a, const uint8_t *img, osr_state_rw_dev *s, double *phase_p, osr_rst_scalars *sc,
                        void *stream);

/* Library version string and the component counts compiled in: every nc
 * from 2 to 12 (osr_supported_nc returns 1 for those, 0 otherwise). */
const char *osr_version(void);
int osr_supported_nc(int32_t nc);

#ifdef __cplusplus
}
#endif
#endif

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
    int32_t nc;              /* 1..16 */
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

/* Saturation pressure of n cells: kind[c] = 1 runs Bubl_Succ
 * (Minor_NR.cpp:491-871) on liquid x with stored vapour y, kind[c] = -1 runs
 * Dew_Succ (Minor_NR.cpp:873-1249) on vapour y, 0 skips (psat = -1).  P is
 * the cell pressure the iteration starts from.  psat[c] = -1 when the
 * reference routine returns -1 (trivial or non-convergent); xm [nc*n]
 * (may be NULL) receives the incipient-phase composition (the global Xm). */
int osr_psat_device(osr_ctx *ctx, int64_t n, const double *T, double T0,
                    const double *P, const int8_t *kind, const double *x,
                    const double *y, double *psat, double *xm, void *stream);

/* Library version string and the component counts compiled in. */
const char *osr_version(void);
int osr_supported_nc(int32_t nc);

#ifdef __cplusplus
}
#endif
#endif

/*
 * osr_restart.c -- CPU restatement of the reference's restart file,
 * WriteRST / ReadRST (compsim/Do_Time_Step.cpp:179-235 / 237-294), into and
 * out of a memory image with the file's exact byte layout.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP restart kernels).
 *
 * The file is the raw bytes of, in this order (FType = double, phaseStat =
 * char, TSMCtrl = int, native little-endian, no padding):
 *   P  [Nx+2][Ny+2][Nz+2][3]   (water, oil, gas pressure; ghost halo included)
 *   fP [Nx+2][Ny+2][Nz+2][3]
 *   per matrix cell, Ix outer, Iz inner: sat[3], then (comp[i][0], comp[i][1])
 *       for i < Nc, then phaseStat
 *   the same per fracture cell (fsat, fcomp, fphaseStat)
 *   simTime, Dt (double), TSMCtrl (int), SMethQ (double).
 * Cells here are indexed iz*nx*ny + iy*nx + ix (SoA, component-major).
 * Halo entries: Edge_Trans / Edge_fTrans zero the face halos
 * (Init_Calcs.cpp:262-305, 473-516); edges and corners of the halo are never
 * written by the reference (P is malloc'd, Data_Input.cpp:1523-1540) nor read,
 * and are written as 0 here.
 */
#include <stdint.h>
#include <string.h>

#include "osr_oracle.h"

int64_t or_rst_bytes(int nx, int ny, int nz, int nc) {
    const int64_t npad = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
    const int64_t n = (int64_t)nx * ny * nz;
    return 2 * npad * 3 * 8 + 2 * n * (3 * 8 + 2 * 8 * (int64_t)nc + 1) + 8 + 8 + 4 + 8;
}

static uint8_t *put(uint8_t *o, const void *v, size_t b) {
    memcpy(o, v, b);
    return o + b;
}

static const uint8_t *get(const uint8_t *o, void *v, size_t b) {
    memcpy(v, o, b);
    return o + b;
}

/* P, fP sections: Do_Time_Step.cpp:185-196 (write), 243-254 (read) */
static uint8_t *put_pressures(uint8_t *o, int nx, int ny, int nz, const double *P) {
    const int64_t n = (int64_t)nx * ny * nz;
    for (int Ix = 0; Ix < nx + 2; Ix++)
        for (int Iy = 0; Iy < ny + 2; Iy++)
            for (int Iz = 0; Iz < nz + 2; Iz++)
                for (int i = 0; i < 3; i++) {
                    double v = 0.0;
                    if (Ix >= 1 && Ix <= nx && Iy >= 1 && Iy <= ny && Iz >= 1 && Iz <= nz)
                        v = P[i * n + (int64_t)(Iz - 1) * nx * ny + (int64_t)(Iy - 1) * nx + (Ix - 1)];
                    o = put(o, &v, 8);
                }
    return o;
}

static const uint8_t *get_pressures(const uint8_t *o, int nx, int ny, int nz, double *P) {
    const int64_t n = (int64_t)nx * ny * nz;
    for (int Ix = 0; Ix < nx + 2; Ix++)
        for (int Iy = 0; Iy < ny + 2; Iy++)
            for (int Iz = 0; Iz < nz + 2; Iz++)
                for (int i = 0; i < 3; i++) {
                    double v;
                    o = get(o, &v, 8);
                    if (P && Ix >= 1 && Ix <= nx && Iy >= 1 && Iy <= ny && Iz >= 1 && Iz <= nz)
                        P[i * n + (int64_t)(Iz - 1) * nx * ny + (int64_t)(Iy - 1) * nx + (Ix - 1)] = v;
                }
    return o;
}

/* cell records: Do_Time_Step.cpp:197-226 (write), 255-284 (read) */
static uint8_t *put_cells(uint8_t *o, int nx, int ny, int nz, int nc, const double *S, const double *x,
                          const double *y, const int8_t *label) {
    const int64_t n = (int64_t)nx * ny * nz;
    for (int Ix = 0; Ix < nx; Ix++)
        for (int Iy = 0; Iy < ny; Iy++)
            for (int Iz = 0; Iz < nz; Iz++) {
                const int64_t c = (int64_t)Iz * nx * ny + (int64_t)Iy * nx + Ix;
                for (int i = 0; i < 3; i++) o = put(o, &S[i * n + c], 8);
                for (int i = 0; i < nc; i++) {
                    o = put(o, &x[i * n + c], 8);
                    o = put(o, &y[i * n + c], 8);
                }
                o = put(o, &label[c], 1);
            }
    return o;
}

static const uint8_t *get_cells(const uint8_t *o, int nx, int ny, int nz, int nc, double *S, double *x, double *y,
                                int8_t *label) {
    const int64_t n = (int64_t)nx * ny * nz;
    for (int Ix = 0; Ix < nx; Ix++)
        for (int Iy = 0; Iy < ny; Iy++)
            for (int Iz = 0; Iz < nz; Iz++) {
                const int64_t c = (int64_t)Iz * nx * ny + (int64_t)Iy * nx + Ix;
                for (int i = 0; i < 3; i++) o = get(o, &S[i * n + c], 8);
                for (int i = 0; i < nc; i++) {
                    o = get(o, &x[i * n + c], 8);
                    o = get(o, &y[i * n + c], 8);
                }
                o = get(o, &label[c], 1);
            }
    return o;
}

int64_t or_rst_write(int nx, int ny, int nz, int nc, const double *P_m, const double *P_f, const double *S_m,
                     const double *x_m, const double *y_m, const int8_t *label_m, const double *S_f,
                     const double *x_f, const double *y_f, const int8_t *label_f, double sim_time, double dt,
                     int32_t tsm_ctrl, double smethq, uint8_t *img) {
    uint8_t *o = img;
    o = put_pressures(o, nx, ny, nz, P_m);
    o = put_pressures(o, nx, ny, nz, P_f);
    o = put_cells(o, nx, ny, nz, nc, S_m, x_m, y_m, label_m);
    o = put_cells(o, nx, ny, nz, nc, S_f, x_f, y_f, label_f);
    /* Do_Time_Step.cpp:228-232 */
    o = put(o, &sim_time, 8);
    o = put(o, &dt, 8);
    o = put(o, &tsm_ctrl, 4);
    o = put(o, &smethq, 8);
    return (int64_t)(o - img);
}

int64_t or_rst_read(int nx, int ny, int nz, int nc, const uint8_t *img, double *P_m, double *P_f, double *S_m,
                    double *x_m, double *y_m, int8_t *label_m, double *S_f, double *x_f, double *y_f,
                    int8_t *label_f, double *sim_time, double *dt, int32_t *tsm_ctrl, double *smethq) {
    const uint8_t *o = img;
    o = get_pressures(o, nx, ny, nz, P_m);
    o = get_pressures(o, nx, ny, nz, P_f);
    o = get_cells(o, nx, ny, nz, nc, S_m, x_m, y_m, label_m);
    o = get_cells(o, nx, ny, nz, nc, S_f, x_f, y_f, label_f);
    /* Do_Time_Step.cpp:286-290 */
    o = get(o, sim_time, 8);
    o = get(o, dt, 8);
    o = get(o, tsm_ctrl, 4);
    o = get(o, smethq, 8);
    return (int64_t)(o - img);
}

// restart.hip -- the reference's restart file (WriteRST / ReadRST,
// compsim/Do_Time_Step.cpp:179-294) as a byte image in HBM.
//
// Layout (FType = double, phaseStat = char, TSMCtrl = int, little-endian,
// no padding): P[Nx+2][Ny+2][Nz+2][3] and fP (water, oil, gas pressure with
// the ghost halo), then per matrix cell in Ix-outer, Iz-inner order sat[3],
// (comp[i][0], comp[i][1]) for i < Nc and phaseStat, the same per fracture
// cell, then simTime, Dt, TSMCtrl, SMethQ.  The halo holds 0 (the reference
// zeroes its faces, Init_Calcs.cpp:262-305, and never touches its edges).
//
// The cell records are 24 + 16 nc + 1 bytes, so their doubles sit at every
// byte alignment.  The writer therefore runs one lane per 4-byte word of the
// image: each lane assembles its word from the (transposed) state bytes it
// covers and issues one aligned store, so a wave writes 256 contiguous
// bytes.  The reader runs one lane per destination value and gathers its
// bytes.  Both are HBM-bound layout transforms.
#include <hip/hip_runtime.h>

#include <cstring>

#include "osr_internal.hpp"

namespace osr {

struct RstArgs {
    int nx, ny, nz, nc;
    int64_t n, npad, rec;        // cells, padded cells, record bytes
    int64_t b_cells, b_tail;     // byte offsets of the matrix records and of the scalars
    int64_t total;               // image bytes
    // write sources / read destinations (SoA, cell = iz*nx*ny + iy*nx + ix)
    const double* Pw[2];
    const double* Po[2];
    const double* Pg[2];
    const double* S[2];
    const double* x[2];
    const double* y[2];
    const int8_t* lab[2];
    double sc[2];                // simTime, Dt
    int32_t tsm;                 // TSMCtrl
    double smethq;               // SMethQ
};

__device__ __forceinline__ uint32_t byte_of(double v, int b) {
    return (uint32_t)((__double_as_longlong(v) >> (8 * b)) & 0xff);
}

// byte o (>= b_cells) of the image
__device__ uint32_t rec_byte(const RstArgs& A, int64_t o) {
    if (o >= A.b_tail) {
        const int t = (int)(o - A.b_tail);
        if (t < 16) return byte_of(A.sc[t >> 3], t & 7);
        if (t < 20) return ((uint32_t)A.tsm >> (8 * (t - 16))) & 0xff;
        return byte_of(A.smethq, t - 20);
    }
    int64_t rel = o - A.b_cells;
    const int k = rel >= A.n * A.rec ? 1 : 0;
    rel -= (int64_t)k * A.n * A.rec;
    const int64_t r = rel / A.rec;
    const int b = (int)(rel - r * A.rec);
    // file order Ix outer, Iz inner
    const int64_t nyz = (int64_t)A.ny * A.nz;
    const int ix = (int)(r / nyz), iy = (int)((r / A.nz) % A.ny), iz = (int)(r % A.nz);
    const int64_t c = (int64_t)iz * A.nx * A.ny + (int64_t)iy * A.nx + ix;
    if (b < 24) return byte_of(A.S[k][(int64_t)(b >> 3) * A.n + c], b & 7);
    const int cb = b - 24;
    if (cb < 16 * A.nc) {
        const int w = cb >> 3, i = w >> 1;
        const double* src = (w & 1) ? A.y[k] : A.x[k];
        return byte_of(src[(int64_t)i * A.n + c], cb & 7);
    }
    return (uint32_t)(uint8_t)A.lab[k][c];
}

// the pressure sections: one lane per double of P and fP (8-byte aligned)
__global__ void k_rst_write_p(RstArgs A, double* img) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 6 * A.npad) return;
    const int k = e >= 3 * A.npad ? 1 : 0;
    const int64_t q = e - (int64_t)k * 3 * A.npad;
    const int64_t p = q / 3;
    const int ph = (int)(q - p * 3);
    const int NY = A.ny + 2, NZ = A.nz + 2;
    const int ix = (int)(p / ((int64_t)NY * NZ)), iy = (int)((p / NZ) % NY), iz = (int)(p % NZ);
    double v = 0.0;
    if (ix >= 1 && ix <= A.nx && iy >= 1 && iy <= A.ny && iz >= 1 && iz <= A.nz) {
        const int64_t c = (int64_t)(iz - 1) * A.nx * A.ny + (int64_t)(iy - 1) * A.nx + (ix - 1);
        v = ph == 0 ? A.Pw[k][c] : (ph == 1 ? A.Po[k][c] : A.Pg[k][c]);
    }
    img[e] = v;
}

// the cell records and the scalars: one lane per 4-byte word from b_cells
// (a multiple of 8); the last, partial word is written bytewise
__global__ void k_rst_write_rec(RstArgs A, uint8_t* img) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t o = A.b_cells + 4 * w;
    if (o >= A.total) return;
    if (o + 4 <= A.total) {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) v |= rec_byte(A, o + i) << (8 * i);
        *reinterpret_cast<uint32_t*>(img + o) = v;
    } else {
        for (int64_t b = o; b < A.total; b++) img[b] = (uint8_t)rec_byte(A, b);
    }
}

__device__ __forceinline__ double load_unaligned(const uint8_t* p) {
    uint64_t v = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return __longlong_as_double((long long)v);
}

struct RstOut {
    double* Po[2];
    double* S[2];
    double* x[2];
    double* y[2];
    int8_t* lab[2];
    double* pwpg;                // [4][n]: Pw_m, Pg_m, Pw_f, Pg_f, or null
};

// one lane per (continuum, cell, value): 3 saturations, 2nc compositions, the
// label, and the three interior pressures (8-byte aligned in the image)
__global__ void k_rst_read(RstArgs A, RstOut O, const uint8_t* img) {
    const int per = 3 + 2 * A.nc + 1 + 3;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * A.n * per) return;
    const int k = e >= A.n * per ? 1 : 0;
    const int64_t q = e - (int64_t)k * A.n * per;
    const int f = (int)(q / A.n);
    const int64_t c = q - (int64_t)f * A.n;
    const int ix = (int)(c % A.nx), iy = (int)((c / A.nx) % A.ny), iz = (int)(c / ((int64_t)A.nx * A.ny));
    if (f >= per - 3) {
        const int ph = f - (per - 3);
        const int64_t p = ((int64_t)(ix + 1) * (A.ny + 2) + (iy + 1)) * (A.nz + 2) + (iz + 1);
        const double v = reinterpret_cast<const double*>(img)[(int64_t)k * 3 * A.npad + 3 * p + ph];
        if (ph == 1) O.Po[k][c] = v;
        else if (O.pwpg) O.pwpg[(int64_t)(2 * k + (ph >> 1)) * A.n + c] = v;
        return;
    }
    const int64_t r = ((int64_t)ix * A.ny + iy) * A.nz + iz;
    const uint8_t* rec = img + A.b_cells + (int64_t)k * A.n * A.rec + r * A.rec;
    if (f < 3) {
        O.S[k][(int64_t)f * A.n + c] = load_unaligned(rec + 8 * f);
    } else if (f < 3 + 2 * A.nc) {
        const int w = f - 3, i = w >> 1;
        double* dst = (w & 1) ? O.y[k] : O.x[k];
        dst[(int64_t)i * A.n + c] = load_unaligned(rec + 24 + 8 * w);
    } else {
        O.lab[k][c] = (int8_t)rec[24 + 16 * A.nc];
    }
}

int rst_shape(int nx, int ny, int nz, int nc, RstArgs& A) {
    if (nx < 1 || ny < 1 || nz < 1 || nc < 1 || nc > 64) return OSR_E_ARG;
    A.nx = nx;
    A.ny = ny;
    A.nz = nz;
    A.nc = nc;
    A.n = (int64_t)nx * ny * nz;
    A.npad = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
    A.rec = 24 + 16 * (int64_t)nc + 1;
    A.b_cells = 2 * A.npad * 24;
    A.b_tail = A.b_cells + 2 * A.n * A.rec;
    A.total = A.b_tail + 28;
    return OSR_OK;
}

}  // namespace osr

extern "C" {

int osr_rst_bytes(int32_t nx, int32_t ny, int32_t nz, int32_t nc, int64_t* bytes) {
    osr::RstArgs A{};
    if (!bytes) return OSR_E_ARG;
    const int rc = osr::rst_shape(nx, ny, nz, nc, A);
    if (rc) return rc;
    *bytes = A.total;
    return OSR_OK;
}

int osr_rst_write_device(osr_asm* a, const osr_state_dev* s, const double* props, const osr_rst_scalars* sc,
                         uint8_t* img, void* stream) {
    if (!a || !s || !props || !sc || !img || !s->label_m || !s->label_f || !s->Po_m || !s->Po_f || !s->x_m ||
        !s->y_m || !s->x_f || !s->y_f || !s->S_m || !s->S_f)
        return OSR_E_ARG;
    if (((uintptr_t)img & 7) != 0) return OSR_E_ARG;
    osr_ctx* ctx;
    int64_t n;
    int nc;
    double T;
    osr::asm_info(a, &ctx, &n, &nc, &T);
    int nx, ny, nz;
    osr::asm_dims(a, &nx, &ny, &nz);
    osr::RstArgs A{};
    OSR_HIP_RET(osr::rst_shape(nx, ny, nz, nc, A));
    const int64_t s2 = 2 * n;   // the property table is [16+2nc][2n], matrix cells first
    A.Pw[0] = props + 1 * s2;
    A.Pw[1] = props + 1 * s2 + n;
    A.Pg[0] = props + 3 * s2;
    A.Pg[1] = props + 3 * s2 + n;
    A.Po[0] = s->Po_m;
    A.Po[1] = s->Po_f;
    A.S[0] = s->S_m;
    A.S[1] = s->S_f;
    A.x[0] = s->x_m;
    A.x[1] = s->x_f;
    A.y[0] = s->y_m;
    A.y[1] = s->y_f;
    A.lab[0] = s->label_m;
    A.lab[1] = s->label_f;
    A.sc[0] = sc->sim_time;
    A.sc[1] = sc->dt;
    A.tsm = sc->tsm_ctrl;
    A.smethq = sc->smethq;
    hipStream_t st = (hipStream_t)stream;
    const int64_t np = 6 * A.npad;
    hipLaunchKernelGGL(osr::k_rst_write_p, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, A,
                       reinterpret_cast<double*>(img));
    const int64_t nw = (A.total - A.b_cells + 3) / 4;
    hipLaunchKernelGGL(osr::k_rst_write_rec, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, A, img);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

int osr_rst_read_device(osr_asm* a, const uint8_t* img, osr_state_rw_dev* s, double* phase_p,
                        osr_rst_scalars* sc, void* stream) {
    if (!a || !img || !s || !sc || !s->label_m || !s->label_f || !s->Po_m || !s->Po_f || !s->x_m || !s->y_m ||
        !s->x_f || !s->y_f || !s->S_m || !s->S_f)
        return OSR_E_ARG;
    if (((uintptr_t)img & 7) != 0) return OSR_E_ARG;
    osr_ctx* ctx;
    int64_t n;
    int nc;
    double T;
    osr::asm_info(a, &ctx, &n, &nc, &T);
    int nx, ny, nz;
    osr::asm_dims(a, &nx, &ny, &nz);
    osr::RstArgs A{};
    OSR_HIP_RET(osr::rst_shape(nx, ny, nz, nc, A));
    osr::RstOut O{};
    O.Po[0] = s->Po_m;
    O.Po[1] = s->Po_f;
    O.S[0] = s->S_m;
    O.S[1] = s->S_f;
    O.x[0] = s->x_m;
    O.x[1] = s->x_f;
    O.y[0] = s->y_m;
    O.y[1] = s->y_f;
    O.lab[0] = s->label_m;
    O.lab[1] = s->label_f;
    O.pwpg = phase_p;
    hipStream_t st = (hipStream_t)stream;
    const int64_t ne = 2 * n * (3 + 2 * nc + 1 + 3);
    hipLaunchKernelGGL(osr::k_rst_read, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, A, O, img);
    OSR_HIP(hipGetLastError());
    // the four scalars of the tail (simTime, Dt, TSMCtrl, SMethQ)
    unsigned char tail[28];
    OSR_HIP(hipMemcpyAsync(tail, img + A.b_tail, 28, hipMemcpyDeviceToHost, st));
    OSR_HIP(hipStreamSynchronize(st));
    memcpy(&sc->sim_time, tail, 8);
    memcpy(&sc->dt, tail + 8, 8);
    memcpy(&sc->tsm_ctrl, tail + 16, 4);
    memcpy(&sc->smethq, tail + 20, 8);
    return OSR_OK;
}

}  // extern "C"

// ksp.hip -- Krylov solve of the reduced (nc+1)^2 block-CSR Newton system on
// gfx950: BiCGStab (van der Vorst) with block-Jacobi preconditioning, the
// role of the reference's KSPBCGS + PCVPBJACOBI (PetscAUX.cpp:628-646,
// KSPSolve inside SNESSolve :669) moved onto the reduced system.
//
// Kernels: block-Jacobi inverse (one group of 16 lanes per block row,
// Gauss-Jordan with partial pivoting, one row per lane, the pivot row
// broadcast by cross-lane reads, the pivot chosen by a DPP argmax), BSR
// matrix-vector product and preconditioner apply (one lane per scalar row),
// fused vector updates that read their scalars from device memory, and
// dot products in two fixed-shape stages (a fixed grid of block partials,
// then one block) so every reduction has one summation order and the solve
// is bitwise reproducible.  The stopping test runs on the device after each
// iteration (k_ksp_check: residual norm, tolerances, divergence, breakdown,
// iteration cap) and sets a done flag that every iteration kernel reads
// first, so the host launches iterations in growing chunks (1, 2, 4, ...,
// 16) and synchronises once per chunk; the iterations after the stop are
// empty launches and leave every vector as the stopping iteration left it.
#include <algorithm>
#include <cmath>
#include <new>

#include "grp_dev.hpp"
#include "osr_internal.hpp"

namespace osr {

constexpr int KG = 16;          // lanes per block row in the inverse kernel
constexpr int DOT_BLOCKS = 1024;

// scalar slots on the device
enum {
    SC_RHO = 0, SC_ALPHA, SC_OMEGA, SC_BETA, SC_RN2, SC_FLAG, SC_NSING,
    // the device stopping test: state, then its parameters
    SC_DONE, SC_ITS, SC_RN, SC_REASON, SC_TOL, SC_ATOL, SC_DTOLR0, SC_MAXIT,
    SC_COUNT = 16
};
// true once the device stopping test has ended the solve (sc may be null
// outside the iteration loop)
__device__ __forceinline__ bool ksp_done(const double* sc) { return sc && sc[SC_DONE] != 0.0; }

template <int BS>
__global__ __launch_bounds__(256) void k_bj_inv(int64_t nbr, const int64_t* browptr, const int32_t* bcols,
                                                const double* bvals, double* minv, double* sc) {
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t br = item / KG;
    const int r = (int)(item % KG);
    if (br >= nbr) return;   // whole groups leave together
    const int gbase = (int)(threadIdx.x & 63) & ~(KG - 1);
    int64_t kd = -1;
    for (int64_t k = browptr[br]; k < browptr[br + 1]; k++)
        if (bcols[k] == br) {
            kd = k;
            break;
        }
    double a[BS], e[BS];
#pragma unroll
    for (int j = 0; j < BS; j++) {
        a[j] = (r < BS && kd >= 0) ? bvals[kd * BS * BS + r * BS + j] : 0.0;
        e[j] = (r == j) ? 1.0 : 0.0;
    }
    bool used = r >= BS;
    int myq = -1;
    bool sing = false;
#pragma unroll
    for (int q = 0; q < BS; q++) {
        double v = used ? -1.0 : fabs(a[q]);
        int idx = r;
        half_argmax<KG>(v, idx);   // the unused row with the largest |a[q]|, first on ties
        if (!(v > 0)) sing = true;
        const int P = gbase + idx;
        double pa[BS], pe[BS];
#pragma unroll
        for (int j = 0; j < BS; j++) {
            pa[j] = __shfl(a[j], P, 64);
            pe[j] = __shfl(e[j], P, 64);
        }
        const double ip = 1.0 / pa[q];
#pragma unroll
        for (int j = 0; j < BS; j++) {
            pa[j] *= ip;
            pe[j] *= ip;
        }
        if (r == idx) {
#pragma unroll
            for (int j = 0; j < BS; j++) {
                a[j] = pa[j];
                e[j] = pe[j];
            }
            used = true;
            myq = q;
        } else if (r < BS) {
            const double f = a[q];
#pragma unroll
            for (int j = 0; j < BS; j++) {
                a[j] -= f * pa[j];
                e[j] -= f * pe[j];
            }
        }
    }
    // the lane that pivoted at step q holds row q of the inverse
    if (myq >= 0)
#pragma unroll
        for (int j = 0; j < BS; j++) minv[br * BS * BS + myq * BS + j] = e[j];
    if (sing && r == 0) atomicAdd(&sc[SC_NSING], 1.0);
}

// y = A x (BSR), one lane per scalar row
template <int BS>
__global__ __launch_bounds__(256) void k_bsr_spmv(int64_t nbr, const int64_t* browptr, const int32_t* bcols,
                                                  const double* bvals, const double* x, double* y,
                                                  const double* sc) {
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nbr * BS || ksp_done(sc)) return;
    const int64_t br = item / BS;
    const int r = (int)(item % BS);
    double acc = 0.0;
    for (int64_t k = browptr[br]; k < browptr[br + 1]; k++) {
        const double* row = bvals + k * BS * BS + r * BS;
        const double* xv = x + (int64_t)bcols[k] * BS;
#pragma unroll
        for (int j = 0; j < BS; j++) acc += row[j] * xv[j];
    }
    y[item] = acc;
}

// z = M^-1 y with M the block diagonal
template <int BS>
__global__ __launch_bounds__(256) void k_bj_apply(int64_t nbr, const double* minv, const double* y, double* z,
                                                  const double* sc) {
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nbr * BS || ksp_done(sc)) return;
    const int64_t br = item / BS;
    const int r = (int)(item % BS);
    const double* row = minv + br * BS * BS + r * BS;
    const double* yv = y + br * BS;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < BS; j++) acc += row[j] * yv[j];
    z[item] = acc;
}

// partial sums of up to two dot products (a.b, c.d) over a fixed grid
__global__ __launch_bounds__(256) void k_dot2(int64_t n, const double* a, const double* b, const double* c,
                                              const double* d, double* part, const double* sc) {
    if (ksp_done(sc)) return;
    __shared__ double s0[256], s1[256];
    double v0 = 0.0, v1 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        v0 += a[i] * b[i];
        if (c) v1 += c[i] * d[i];
    }
    s0[threadIdx.x] = v0;
    s1[threadIdx.x] = v1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0[0];
        part[DOT_BLOCKS + blockIdx.x] = s1[0];
    }
}

enum { OP_RHO = 0, OP_ALPHA, OP_OMEGA, OP_RNORM, OP_INIT };

// final reduction of the partials and the scalar recurrences of BiCGStab
__global__ __launch_bounds__(256) void k_dot_final(const double* part, double* sc, int op) {
    if (sc[SC_DONE] != 0.0) return;
    __shared__ double s0[256], s1[256];
    double v0 = 0.0, v1 = 0.0;
    for (int i = threadIdx.x; i < DOT_BLOCKS; i += 256) {
        v0 += part[i];
        v1 += part[DOT_BLOCKS + i];
    }
    s0[threadIdx.x] = v0;
    s1[threadIdx.x] = v1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double d0 = s0[0], d1 = s1[0];
    switch (op) {
        case OP_INIT:   // d0 = |r0|^2: rho = alpha = omega = 1
            sc[SC_RN2] = d0;
            sc[SC_RHO] = 1.0;
            sc[SC_ALPHA] = 1.0;
            sc[SC_OMEGA] = 1.0;
            sc[SC_FLAG] = 0.0;
            break;
        case OP_RHO:    // d0 = (rhat, r)
            if (d0 == 0.0) sc[SC_FLAG] = 1.0;
            sc[SC_BETA] = (d0 / sc[SC_RHO]) * (sc[SC_ALPHA] / sc[SC_OMEGA]);
            sc[SC_RHO] = d0;
            break;
        case OP_ALPHA:  // d0 = (rhat, v)
            if (d0 == 0.0) sc[SC_FLAG] = 1.0;
            sc[SC_ALPHA] = sc[SC_RHO] / d0;
            break;
        case OP_OMEGA:  // d0 = (t, s), d1 = (t, t)
            if (d1 == 0.0) {
                sc[SC_FLAG] = 1.0;
                sc[SC_OMEGA] = 0.0;
            } else {
                sc[SC_OMEGA] = d0 / d1;
            }
            break;
        default:        // OP_RNORM: d0 = |r|^2
            sc[SC_RN2] = d0;
            break;
    }
}

__global__ void k_vec_init(int64_t n, const double* b, const double* ax, double* r, double* rhat, double* p,
                           double* v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ri = b[i] - ax[i];
    r[i] = ri;
    rhat[i] = ri;
    p[i] = 0.0;
    v[i] = 0.0;
}

// p = r + beta (p - omega v)
__global__ void k_vec_p(int64_t n, const double* sc, const double* r, const double* v, double* p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || ksp_done(sc)) return;
    p[i] = r[i] + sc[SC_BETA] * (p[i] - sc[SC_OMEGA] * v[i]);
}

// s = r - alpha v
__global__ void k_vec_s(int64_t n, const double* sc, const double* r, const double* v, double* s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || ksp_done(sc)) return;
    s[i] = r[i] - sc[SC_ALPHA] * v[i];
}

// x += alpha phat + omega shat;  r = s - omega t
__global__ void k_vec_xr(int64_t n, const double* sc, const double* ph, const double* sh, const double* s,
                         const double* t, double* x, double* r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || ksp_done(sc)) return;
    const double om = sc[SC_OMEGA];
    x[i] += sc[SC_ALPHA] * ph[i] + om * sh[i];
    r[i] = s[i] - om * t[i];
}

__global__ void k_vec_zero(int64_t n, double* x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

__global__ void k_sc_zero(double* sc) {
    if (threadIdx.x < SC_COUNT) sc[threadIdx.x] = 0.0;
}

// the stopping test's parameters; the solve is not done, no iteration run
__global__ void k_ksp_params(double* sc, double tol, double atol, double dtolr0, double maxit) {
    sc[SC_DONE] = 0.0;
    sc[SC_ITS] = 0.0;
    sc[SC_REASON] = 0.0;
    sc[SC_TOL] = tol;
    sc[SC_ATOL] = atol;
    sc[SC_DTOLR0] = dtolr0;
    sc[SC_MAXIT] = maxit;
}

// after an iteration's |r|^2: the host loop's tests in its order -- a
// non-finite norm, the tolerance (absolute if below atol), divergence
// against dtol |r0|, a breakdown flag of the iteration, the iteration cap
__global__ void k_ksp_check(double* sc) {
    if (sc[SC_DONE] != 0.0) return;
    const double its = sc[SC_ITS] + 1.0;
    sc[SC_ITS] = its;
    const double rn = sqrt(sc[SC_RN2]);
    sc[SC_RN] = rn;
    double reason = 0.0;
    if (!isfinite(rn)) reason = OSR_KSP_DIVERGED_NANORINF;
    else if (rn <= sc[SC_TOL]) reason = rn <= sc[SC_ATOL] ? OSR_KSP_CONVERGED_ATOL : OSR_KSP_CONVERGED_RTOL;
    else if (rn >= sc[SC_DTOLR0]) reason = OSR_KSP_DIVERGED_DTOL;
    else if (sc[SC_FLAG] != 0.0) reason = OSR_KSP_DIVERGED_BREAKDOWN;
    else if (its >= sc[SC_MAXIT]) reason = OSR_KSP_DIVERGED_ITS;
    if (reason != 0.0) {
        sc[SC_REASON] = reason;
        sc[SC_DONE] = 1.0;
    }
}

}  // namespace osr

struct osr_ksp {
    int bs = 0;
    int64_t nbr = 0, n = 0;
    const int64_t* browptr = nullptr;
    const int32_t* bcols = nullptr;
    const double* bvals = nullptr;
    double* minv = nullptr;    // [nbr][bs][bs]
    double* vec = nullptr;     // r, rhat, p, v, s, t, phat, shat: 8 x n
    double* part = nullptr;    // [2][DOT_BLOCKS]
    double* sc = nullptr;      // [SC_COUNT]
    double* host_sc = nullptr; // pinned [SC_COUNT]
};

namespace {

template <int BS>
int ksp_setup_bs(osr_ksp* k, hipStream_t st) {
    hipLaunchKernelGGL(osr::k_sc_zero, dim3(1), dim3(64), 0, st, k->sc);
    const unsigned g = (unsigned)((k->nbr * osr::KG + 255) / 256);
    hipLaunchKernelGGL(osr::k_bj_inv<BS>, dim3(g), dim3(256), 0, st, k->nbr, k->browptr, k->bcols, k->bvals, k->minv,
                       k->sc);
    OSR_HIP(hipGetLastError());
    OSR_HIP(hipMemcpyAsync(k->host_sc, k->sc, osr::SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, st));
    OSR_HIP(hipStreamSynchronize(st));
    return (int)k->host_sc[osr::SC_NSING];
}

// sc: the device stopping state the kernels test first, or null
template <int BS>
void spmv(osr_ksp* k, const double* x, double* y, const double* sc, hipStream_t st) {
    const unsigned g = (unsigned)((k->n + 255) / 256);
    hipLaunchKernelGGL(osr::k_bsr_spmv<BS>, dim3(g), dim3(256), 0, st, k->nbr, k->browptr, k->bcols, k->bvals, x, y,
                       sc);
}

template <int BS>
void apply(osr_ksp* k, const double* y, double* z, hipStream_t st) {
    const unsigned g = (unsigned)((k->n + 255) / 256);
    hipLaunchKernelGGL(osr::k_bj_apply<BS>, dim3(g), dim3(256), 0, st, k->nbr, (const double*)k->minv, y, z,
                       (const double*)k->sc);
}

void dot(osr_ksp* k, const double* a, const double* b, const double* c, const double* d, int op, hipStream_t st) {
    hipLaunchKernelGGL(osr::k_dot2, dim3(osr::DOT_BLOCKS), dim3(256), 0, st, k->n, a, b, c, d, k->part,
                       (const double*)k->sc);
    hipLaunchKernelGGL(osr::k_dot_final, dim3(1), dim3(256), 0, st, (const double*)k->part, k->sc, op);
}

int read_sc(osr_ksp* k, hipStream_t st) {
    OSR_HIP(hipMemcpyAsync(k->host_sc, k->sc, osr::SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, st));
    OSR_HIP(hipStreamSynchronize(st));
    return OSR_OK;
}

template <int BS>
int ksp_solve_bs(osr_ksp* k, const double* b, double* x, const osr_ksp_opts* o, osr_ksp_result* res,
                 hipStream_t st) {
    const int64_t n = k->n;
    double* r = k->vec;
    double* rhat = r + n;
    double* p = rhat + n;
    double* v = p + n;
    double* s = v + n;
    double* t = s + n;
    double* ph = t + n;
    double* sh = ph + n;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (!o->x0_nonzero) hipLaunchKernelGGL(osr::k_vec_zero, dim3(g), dim3(256), 0, st, n, x);
    // a fresh solve: the stopping state cleared before the first dot
    hipLaunchKernelGGL(osr::k_ksp_params, dim3(1), dim3(1), 0, st, k->sc, 0.0, 0.0, 0.0, 0.0);
    // |b| and r0 = b - A x0
    dot(k, b, b, nullptr, nullptr, osr::OP_RNORM, st);
    OSR_HIP_RET(read_sc(k, st));
    const double bnorm = std::sqrt(k->host_sc[osr::SC_RN2]);
    spmv<BS>(k, x, t, nullptr, st);
    hipLaunchKernelGGL(osr::k_vec_init, dim3(g), dim3(256), 0, st, n, b, (const double*)t, r, rhat, p, v);
    dot(k, r, r, nullptr, nullptr, osr::OP_INIT, st);
    OSR_HIP_RET(read_sc(k, st));
    const double r0 = std::sqrt(k->host_sc[osr::SC_RN2]);
    res->bnorm = bnorm;
    res->r0norm = r0;
    res->rnorm = r0;
    res->iterations = 0;
    const double tol = std::max(o->rtol * bnorm, o->atol);
    if (!std::isfinite(r0)) {
        res->reason = OSR_KSP_DIVERGED_NANORINF;
        return OSR_OK;
    }
    if (r0 <= tol) {
        res->reason = r0 <= o->atol ? OSR_KSP_CONVERGED_ATOL : OSR_KSP_CONVERGED_RTOL;
        return OSR_OK;
    }
    res->reason = OSR_KSP_DIVERGED_ITS;
    if (o->maxit < 1) return OSR_OK;
    hipLaunchKernelGGL(osr::k_ksp_params, dim3(1), dim3(1), 0, st, k->sc, tol, o->atol, o->dtol * r0,
                       (double)o->maxit);
    const double* scd = k->sc;
    int launched = 0, chunk = 1;
    while (launched < o->maxit) {
        const int m = std::min(chunk, o->maxit - launched);
        for (int it = 0; it < m; it++) {
            dot(k, rhat, r, nullptr, nullptr, osr::OP_RHO, st);
            hipLaunchKernelGGL(osr::k_vec_p, dim3(g), dim3(256), 0, st, n, scd, (const double*)r, (const double*)v,
                               p);
            apply<BS>(k, p, ph, st);
            spmv<BS>(k, ph, v, scd, st);
            dot(k, rhat, v, nullptr, nullptr, osr::OP_ALPHA, st);
            hipLaunchKernelGGL(osr::k_vec_s, dim3(g), dim3(256), 0, st, n, scd, (const double*)r, (const double*)v,
                               s);
            apply<BS>(k, s, sh, st);
            spmv<BS>(k, sh, t, scd, st);
            dot(k, t, s, t, t, osr::OP_OMEGA, st);
            hipLaunchKernelGGL(osr::k_vec_xr, dim3(g), dim3(256), 0, st, n, scd, (const double*)ph,
                               (const double*)sh, (const double*)s, (const double*)t, x, r);
            dot(k, r, r, nullptr, nullptr, osr::OP_RNORM, st);
            hipLaunchKernelGGL(osr::k_ksp_check, dim3(1), dim3(1), 0, st, k->sc);
        }
        OSR_HIP(hipGetLastError());
        launched += m;
        chunk = std::min(2 * chunk, 16);
        OSR_HIP_RET(read_sc(k, st));
        if (k->host_sc[osr::SC_DONE] != 0.0) break;
    }
    res->iterations = (int)k->host_sc[osr::SC_ITS];
    res->rnorm = k->host_sc[osr::SC_RN];
    res->reason = (int)k->host_sc[osr::SC_REASON];
    return OSR_OK;
}

// block sizes nc + 1 of the compiled component counts
#define OSR_BS_RET_(NC) case NC + 1: return CALL(NC + 1);
#define OSR_BS_SWITCH(bs, CALL_)          \
    switch (bs) {                         \
        OSR_NC_LIST(OSR_BS_RET_)          \
        default: return OSR_E_ARG;        \
    }

}  // namespace

extern "C" {

int osr_ksp_create(int32_t bs, int64_t nbrows, osr_ksp** out) {
    if (!out || nbrows < 1) return OSR_E_ARG;
    if (!osr::nc_supported(bs - 1)) return OSR_E_ARG;
    osr_ksp* k = new (std::nothrow) osr_ksp();
    if (!k) return OSR_E_ARG;
    k->bs = bs;
    k->nbr = nbrows;
    k->n = nbrows * bs;
    bool ok = hipMalloc((void**)&k->minv, (size_t)nbrows * bs * bs * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&k->vec, (size_t)8 * k->n * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&k->part, (size_t)2 * osr::DOT_BLOCKS * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&k->sc, osr::SC_COUNT * sizeof(double)) == hipSuccess &&
              hipHostMalloc((void**)&k->host_sc, osr::SC_COUNT * sizeof(double), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        osr_ksp_destroy(k);
        return OSR_E_HIP;
    }
    *out = k;
    return OSR_OK;
}

int osr_ksp_destroy(osr_ksp* k) {
    if (!k) return OSR_E_ARG;
    void* d[] = {k->minv, k->vec, k->part, k->sc};
    for (void* p : d)
        if (p) (void)hipFree(p);
    if (k->host_sc) (void)hipHostFree(k->host_sc);
    delete k;
    return OSR_OK;
}

int osr_ksp_setup_device(osr_ksp* k, const int64_t* browptr, const int32_t* bcols, const double* bvals,
                         void* stream) {
    if (!k || !browptr || !bcols || !bvals) return OSR_E_ARG;
    k->browptr = browptr;
    k->bcols = bcols;
    k->bvals = bvals;
#define CALL(BS) ksp_setup_bs<BS>(k, (hipStream_t)stream)
    OSR_BS_SWITCH(k->bs, CALL)
#undef CALL
}

int osr_ksp_solve_device(osr_ksp* k, const double* b, double* x, const osr_ksp_opts* opts, osr_ksp_result* res,
                         void* stream) {
    if (!k || !b || !x || !opts || !res || !k->bvals) return OSR_E_ARG;
    if (opts->maxit < 0 || !(opts->dtol > 0)) return OSR_E_ARG;
#define CALL(BS) ksp_solve_bs<BS>(k, b, x, opts, res, (hipStream_t)stream)
    OSR_BS_SWITCH(k->bs, CALL)
#undef CALL
}

}  // extern "C"

// newton.hip -- one time step's nonlinear solve on the device: SolveNonLinear
// (compsim/PetscAUX.cpp:566-730, SNESKSPONLY) and, for an accepted step,
// UpdateAll(X, true) with UpdateDependencies(true) (PetscAUX.cpp:176-507).
//
// The step composes the library's own entry points (pack, residual, reduced
// system, BiCGStab, expansion, unpack, PhaseCtrl, layout) on one stream and
// adds three reductions of its own: the 1-norm of the update (the acceptance
// test), the largest oil pressure (CalcPCoeff) and X0 + du.  Reductions run
// over a fixed grid and combine the partials in a fixed order, so a step is
// bitwise reproducible.
#include <hip/hip_runtime.h>

#include <cmath>

#include "osr_internal.hpp"

namespace osr {

constexpr int NRM_BLOCKS = 512;

// partials of sum |du| and of the count of non-finite entries
__global__ __launch_bounds__(256) void k_step_norm(int64_t n, const double* du, double* part) {
    __shared__ double s0[256], s1[256];
    double a = 0.0, bad = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double v = du[i];
        if (v - v != 0.0) bad += 1.0;   // NaN or infinite
        else a += fabs(v);
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = bad;
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
        part[NRM_BLOCKS + blockIdx.x] = s1[0];
    }
}

// partials of CalcPCoeff: the largest oil pressure of both continua, from 0
// (a cell enters only when its pressure is greater, so NaN never does)
__global__ __launch_bounds__(256) void k_step_pmax(int64_t n, const double* Pm, const double* Pf, double* part) {
    __shared__ double s0[256];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double a = Pm[i], b = Pf[i];
        if (a > m) m = a;
        if (b > m) m = b;
    }
    s0[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && s0[threadIdx.x + w] > s0[threadIdx.x]) s0[threadIdx.x] = s0[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[2 * NRM_BLOCKS + blockIdx.x] = s0[0];
}

// out[0] = sum |du|, out[1] = non-finite count, out[2] = PCoeff (if pmax)
__global__ __launch_bounds__(256) void k_step_final(const double* part, double* out, int pmax) {
    __shared__ double s0[256], s1[256], s2[256];
    double a = 0.0, b = 0.0, m = 0.0;
    for (int i = threadIdx.x; i < NRM_BLOCKS; i += 256) {
        a += part[i];
        b += part[NRM_BLOCKS + i];
        if (pmax && part[2 * NRM_BLOCKS + i] > m) m = part[2 * NRM_BLOCKS + i];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    s2[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
            if (s2[threadIdx.x + w] > s2[threadIdx.x]) s2[threadIdx.x] = s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (pmax) {
            out[2] = s2[0];
        } else {
            out[0] = s0[0];
            out[1] = s1[0];
        }
    }
}

__global__ void k_step_axpy(int64_t n, double* X, const double* du) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) X[i] += du[i];
}

}  // namespace osr

struct osr_newton {
    osr_asm* a = nullptr;
    osr_ctx* ctx = nullptr;
    int64_t n = 0;
    int nc = 0;
    double T = 0;
    osr_ksp* ksp = nullptr;
    int64_t rows_cap = 0;       // 2n (2nc+4)
    double *X = nullptr, *F = nullptr, *du = nullptr;
    double *brhs = nullptr, *dp = nullptr, *S = nullptr, *t = nullptr;
    int64_t* browptr = nullptr;
    int32_t* bcols = nullptr;
    double* bvals = nullptr;
    int64_t nnzb_cap = 0;
    double *rho0 = nullptr, *psat = nullptr, *xm = nullptr, *part = nullptr, *sc = nullptr;
    double* host_sc = nullptr;  // pinned [4]
    int64_t last_nrows = 0;
};

namespace {

void release(osr_newton* w) {
    if (w->ksp) osr_ksp_destroy(w->ksp);
    double* d[] = {w->X, w->F, w->du, w->brhs, w->dp, w->S, w->t, w->bvals, w->rho0, w->psat, w->xm, w->part,
                   w->sc};
    for (double* p : d)
        if (p) hipFree(p);
    if (w->browptr) hipFree(w->browptr);
    if (w->bcols) hipFree(w->bcols);
    if (w->host_sc) hipHostFree(w->host_sc);
    delete w;
}

osr_state_dev const_view(const osr_state_rw_dev* s, const double* rho_prev) {
    osr_state_dev d{};
    d.label_m = s->label_m;
    d.label_f = s->label_f;
    d.Po_m = s->Po_m;
    d.Po_f = s->Po_f;
    d.x_m = s->x_m;
    d.y_m = s->y_m;
    d.x_f = s->x_f;
    d.y_f = s->y_f;
    d.S_m = s->S_m;
    d.S_f = s->S_f;
    d.rho_prev_m = rho_prev;
    d.pre_m = nullptr;
    d.pre_f = nullptr;
    d.pcoeff = s->pcoeff;
    return d;
}

}  // namespace

extern "C" {

int osr_newton_create(osr_asm* a, osr_newton** out) {
    if (!a || !out) return OSR_E_ARG;
    *out = nullptr;
    osr_newton* w = new osr_newton();
    w->a = a;
    osr::asm_info(a, &w->ctx, &w->n, &w->nc, &w->T);
    const int64_t nb = 2 * w->n;
    const int nc = w->nc;
    w->rows_cap = nb * (2 * nc + 4);
    int rc = osr_ksp_create(nc + 1, nb, &w->ksp);
    if (rc) {
        release(w);
        return rc;
    }
    bool ok = hipMalloc((void**)&w->X, w->rows_cap * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->F, w->rows_cap * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->du, w->rows_cap * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->brhs, nb * (nc + 1) * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->dp, nb * (nc + 1) * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->S, nb * (nc + 3) * (nc + 1) * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->t, nb * (nc + 3) * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->browptr, (nb + 1) * sizeof(int64_t)) == hipSuccess &&
              hipMalloc((void**)&w->rho0, nb * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->psat, nb * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->xm, nb * nc * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->part, 3 * osr::NRM_BLOCKS * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->sc, 4 * sizeof(double)) == hipSuccess &&
              hipHostMalloc((void**)&w->host_sc, 4 * sizeof(double), hipHostMallocDefault) == hipSuccess;
    // dp is the initial guess of a solve with x0_nonzero: start from zero
    if (!ok || hipMemset(w->dp, 0, nb * (nc + 1) * sizeof(double)) != hipSuccess) {
        release(w);
        return OSR_E_HIP;
    }
    *out = w;
    return OSR_OK;
}

int osr_newton_destroy(osr_newton* w) {
    if (!w) return OSR_E_ARG;
    release(w);
    return OSR_OK;
}

int osr_newton_solution_update(osr_newton* w, double* du, int64_t* nrows, void* stream) {
    if (!w || !nrows) return OSR_E_ARG;
    *nrows = w->last_nrows;
    if (du && w->last_nrows > 0)
        OSR_HIP(hipMemcpyAsync(du, w->du, w->last_nrows * sizeof(double), hipMemcpyDeviceToDevice,
                               (hipStream_t)stream));
    return OSR_OK;
}

int osr_newton_step_device(osr_newton* w, osr_state_rw_dev* s, const osr_newton_opts* o, double* rho_next,
                           osr_newton_result* r, void* stream) {
    if (!w || !s || !o || !r || !s->label_m || !s->label_f) return OSR_E_ARG;
    *r = osr_newton_result{};
    hipStream_t st = (hipStream_t)stream;
    osr_asm* a = w->a;
    const int64_t n = w->n;
    const osr_state_dev sd = const_view(s, s->rho_prev_m);

    // FormFunction / FormJacobian at X0 (SNESKSPONLY evaluates both once)
    int64_t nrows = 0, nnzb = 0;
    OSR_HIP_RET(osr_asm_layout(a, &sd, nullptr, &nrows, stream));                       // PJac
    if (nrows > w->rows_cap) return OSR_E_ARG;
    w->last_nrows = nrows;
    r->nrows = r->nrows_next = nrows;
    OSR_HIP_RET(osr_pack_device(a, &sd, w->X, stream));                                 // InitGuess2
    OSR_HIP_RET(osr_residual_device(a, &sd, w->F, stream));
    OSR_HIP_RET(osr_bsr_reduced_rowptr_device(a, &sd, w->browptr, &nnzb, stream));
    if (nnzb > w->nnzb_cap) {
        OSR_HIP(hipStreamSynchronize(st));
        if (w->bcols) hipFree(w->bcols);
        if (w->bvals) hipFree(w->bvals);
        w->bcols = nullptr;
        w->bvals = nullptr;
        w->nnzb_cap = 0;
        const size_t bb = (size_t)(w->nc + 1) * (w->nc + 1);
        OSR_HIP(hipMalloc((void**)&w->bcols, nnzb * sizeof(int32_t)));
        OSR_HIP(hipMalloc((void**)&w->bvals, nnzb * bb * sizeof(double)));
        w->nnzb_cap = nnzb;
    }
    OSR_HIP_RET(osr_jacobian_bsr_reduced_direct_device(a, &sd, w->F, w->browptr, w->bcols, w->bvals, w->brhs,
                                                       w->S, w->t, stream));
    // densities of X0's sweep: the next sweep's CalcIFT reads them
    OSR_HIP_RET(osr::asm_take_rho(a, w->rho0, st));

    // KSPSolve, then the update du = -Y in the reference's unknown layout
    const int nsing = osr_ksp_setup_device(w->ksp, w->browptr, w->bcols, w->bvals, stream);
    if (nsing < 0) return nsing;
    r->singular_blocks = nsing;
    OSR_HIP_RET(osr_ksp_solve_device(w->ksp, w->brhs, w->dp, &o->ksp, &r->ksp, stream));
    OSR_HIP_RET(osr_bsr_reduced_expand_device(a, &sd, w->S, w->t, w->dp, w->du, stream));

    // vnorm = |Y|_1 / MatrixSize (PetscAUX.cpp:682-683) and the acceptance test (:703)
    hipLaunchKernelGGL(osr::k_step_norm, dim3(osr::NRM_BLOCKS), dim3(256), 0, st, nrows, w->du, w->part);
    hipLaunchKernelGGL(osr::k_step_final, dim3(1), dim3(256), 0, st, w->part, w->sc, 0);
    OSR_HIP(hipGetLastError());
    OSR_HIP(hipMemcpyAsync(w->host_sc, w->sc, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    OSR_HIP(hipStreamSynchronize(st));
    const double vnorm = w->host_sc[1] > 0 ? NAN : w->host_sc[0] / (double)nrows;
    r->vnorm = vnorm;
    r->snes_reason = r->ksp.reason > 0 ? OSR_SNES_CONVERGED_ITS : OSR_SNES_DIVERGED_LINEAR_SOLVE;
    if (r->snes_reason <= 0 || r->ksp.reason <= 0 || vnorm > o->tolmain || vnorm - vnorm != 0) return OSR_OK;

    // UpdateAll(X, true): X1 = X0 + du -> state (clamps, limiter, floor)
    hipLaunchKernelGGL(osr::k_step_axpy, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, st, nrows, w->X,
                       w->du);
    OSR_HIP(hipGetLastError());
    osr_state_mut_dev sm{};
    sm.label_m = s->label_m;
    sm.label_f = s->label_f;
    sm.Po_m = s->Po_m;
    sm.Po_f = s->Po_f;
    sm.x_m = s->x_m;
    sm.y_m = s->y_m;
    sm.x_f = s->x_f;
    sm.y_f = s->y_f;
    sm.S_m = s->S_m;
    sm.S_f = s->S_f;
    sm.pcoeff = s->pcoeff;
    OSR_HIP_RET(osr_unpack_device(a, &sm, w->X, stream));

    // UpdateDependencies(true): PhaseCtrl, fPhaseCtrl (gas pressures of X0's
    // sweep, still in the property table), CalcPCoeff, PJac, then the sweep.
    // The saturation pressures the phase test needs first (matrix, then
    // fracture), the largest pressure (PCoeff), then one kernel per block
    // applies the phase rule and runs the new state's property sweep and
    // block size, and the offsets are scanned; one synchronisation reads
    // PCoeff and MatrixSize.
    const double* props = nullptr;
    OSR_HIP_RET(osr_asm_props_table(a, &props));
    const double* Pg = props + 3 * 2 * n;
    OSR_HIP_RET(osr::phase_psat(w->ctx, n, nullptr, w->T, s->label_m, s->Po_m, Pg, s->x_m, s->y_m, w->psat,
                                w->xm, st));
    OSR_HIP_RET(osr::phase_psat(w->ctx, n, nullptr, w->T, s->label_f, s->Po_f, Pg + n, s->x_f, s->y_f,
                                w->psat + n, w->xm + (size_t)w->nc * n, st));
    hipLaunchKernelGGL(osr::k_step_pmax, dim3(osr::NRM_BLOCKS), dim3(256), 0, st, n, s->Po_m, s->Po_f, w->part);
    hipLaunchKernelGGL(osr::k_step_final, dim3(1), dim3(256), 0, st, w->part, w->sc, 1);
    OSR_HIP(hipGetLastError());
    OSR_HIP(hipMemcpyAsync(w->host_sc + 2, w->sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    int64_t nrows_next = 0;
    OSR_HIP_RET(osr::asm_apply_sweep(a, &sm, w->rho0, w->psat, w->xm, rho_next, &nrows_next, st));
    if (!(w->host_sc[2] > 0)) return OSR_E_ARG;   // no positive pressure left: no PCoeff to scale by
    s->pcoeff = w->host_sc[2];
    r->nrows_next = nrows_next;
    r->accepted = 1;
    return OSR_OK;
}

}  // extern "C"

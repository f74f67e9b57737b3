// flash.hip -- two-phase flash of independent cells on gfx950.
//
// Restates SuccFalsh (compsim/Init_Calcs.cpp:571-738): Wilson start, outer
// successive substitution (SS) on K with an inner Rachford-Rice (RR) Newton
// restarted from L = 0.6, L clamped to [0,1], stop at |dL| <= 1e-14 or 500
// passes; the phase label comes from the clamped L.
//
// Three launch forms share one per-cell iteration body (ss_pass):
//  * plain      -- one lane per cell, loop to convergence.  A wave runs as
//                  long as its slowest cell (the reference form on a GPU).
//  * persistent -- a grid sized to the resident-wave capacity; each lane owns
//                  a cell slot and, when its cell converges, the wave refills
//                  idle lanes from a global work-queue head with ONE atomic per
//                  wave (ballot + popcount + prefix rank).  Lanes stay busy
//                  however uneven the iteration counts are.
//  * mixed      -- stage A (FP64, persistent) runs the first passes of every
//                  cell exactly as the reference does and compacts cells that
//                  are still iterating into a queue; stage B runs FP32 SS on the
//                  queue until |dL| < tol32; stage C finishes in FP64 with the
//                  reference stopping rule, so outputs are FP64 iterates.
#include "osr_internal.hpp"

namespace osr {

struct FlashArgs {
    int64_t n;
    const double* T;
    double T0;
    const double* P;
    const double* z;
    double sw;
    osr_flash_out out;
    unsigned long long* counter;
    int refill_min;
};

template <int NC, class R>
struct SSCell {
    R K[NC];
    R z[NC];
    R sa[NC];
    R P, T, L, Lb, Zl, Zg;
    R pA, pB;   // P/(R T)^2, P/(R T)
    int it, rr;
};

// One outer SS pass (Init_Calcs.cpp:590-713).  Returns |L - Lb| (ESum).
// x, y receive the compositions computed from the incoming K (the values the
// reference stores in comp[..][0..1] during that pass).
template <int NC, class R>
__device__ __forceinline__ R ss_pass(const FluidConsts<R>& C, bool srk, SSCell<NC, R>& s,
                                     R (&x)[NC], R (&y)[NC]) {
    // Rachford-Rice Newton from L = 0.6 (:592-606).
    R L = R(0.6);
    int j = 0;
    R step;
    do {
        R F = 0, dF = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            const R Ki = s.K[i];
            const R td = (Ki >= 0) ? (R(1) - Ki) / (L + (R(1) - L) * Ki) : R(1) / (L - R(1));
            F += s.z[i] * td;
            dF -= s.z[i] * td * td;
        }
        step = F / dF;
        L -= step;
        j++;
    } while (step * step > R(NEWTON_TOL) && j < MAX_FLASH_N_ITR);
    s.rr += j;
    if (L > 1) L = 1;
    else if (L < 0) L = 0;
    const R esum = fabs(s.Lb - L);
    s.Lb = L;
    s.L = L;

#pragma unroll
    for (int i = 0; i < NC; i++) {
        const R Ki = s.K[i];
        if (Ki >= 0) {
            x[i] = s.z[i] / (L + Ki * (R(1) - L));
            y[i] = Ki * x[i];
        } else {
            x[i] = 0;
            y[i] = s.z[i] / (R(1) - L);
        }
    }

    // Both phases at the oil pressure (:660-664).
    R tl[NC], tg[NC];
    const PhaseInv<R> pl = phase_eval<NC, R>(C, srk, s.sa, x, s.pA, s.pB, false, tl);
    const PhaseInv<R> pg = phase_eval<NC, R>(C, srk, s.sa, y, s.pA, s.pB, true, tg);
    s.Zl = pl.Z;
    s.Zg = pg.Z;

    // K_i <- K_i * (x_i phiL_i) / (y_i phiG_i) for z_i != 0 (:705-710).  With
    // K_i >= 0 and x_i finite and nonzero, y_i = K_i x_i, so the update is
    // exp(lnphiL_i - lnphiG_i) with a single exponential.
#pragma unroll
    for (int i = 0; i < NC; i++) {
        if (s.z[i] != 0) {
            const R a = lnphi<R>(pl, C.b[i], tl[i]);
            const R g = lnphi<R>(pg, C.b[i], tg[i]);
            const R Ki = s.K[i];
            R r;
            if (Ki > 0 && x[i] != 0 && isfinite(x[i])) {
                r = m_exp(a - g);
            } else {
                const R fL = x[i] * m_exp(a);
                const R fG = y[i] * m_exp(g);
                r = (fG != 0) ? Ki * (fL / fG) : R(-1);
            }
            if (m_isnan(a)) r = -1;
            if (m_isnan(g)) r = 1;
            s.K[i] = r;
        }
    }
    s.it++;
    return esum;
}

template <int NC, class R>
__device__ __forceinline__ void cell_init(const FluidConsts<R>& C, const FlashArgs& a, int64_t c,
                                          SSCell<NC, R>& s) {
    s.T = R(a.T ? a.T[c] : a.T0);
    s.P = R(a.P[c]);
#pragma unroll
    for (int i = 0; i < NC; i++) s.z[i] = R(a.z[(int64_t)i * a.n + c]);
    sqrt_a<NC, R>(C, s.T, s.sa);
    wilson<NC, R>(C, s.T, s.P, s.K);
    const R rt = R(RGAS) * s.T;
    s.pB = s.P / rt;
    s.pA = s.pB / rt;
    s.L = 0;
    s.Lb = R(10);
    s.Zl = s.Zg = 0;
    s.it = 0;
    s.rr = 0;
}

// Outputs of a finished cell (Init_Calcs.cpp:717-736).
template <int NC>
__device__ __forceinline__ void cell_write(const FlashArgs& a, int64_t c, const SSCell<NC, double>& s,
                                           const double (&x)[NC], const double (&y)[NC]) {
    const osr_flash_out& o = a.out;
    const int64_t n = a.n;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        if (o.x) o.x[(int64_t)i * n + c] = x[i];
        if (o.y) o.y[(int64_t)i * n + c] = y[i];
        if (o.K) o.K[(int64_t)i * n + c] = s.K[i];
    }
    const double L = s.L;
    int8_t lab = 0;
    if (fabs(L) < DZERO) lab = -1;
    else if (fabs(1 - L) < DZERO) lab = 1;
    if (o.L) o.L[c] = L;
    if (o.label) o.label[c] = lab;
    if (o.Zl) o.Zl[c] = s.Zl;
    if (o.Zg) o.Zg[c] = s.Zg;
    const double newD = s.Zl * L / (s.Zl * L + s.Zg * (1 - L));
    if (o.so) o.so[c] = newD * (1 - a.sw);
    if (o.sg) o.sg[c] = (1 - newD) * (1 - a.sw);
    if (o.rho_l) o.rho_l[c] = s.P / (s.Zl * RGAS * s.T);
    if (o.rho_g) o.rho_g[c] = s.P / (s.Zg * RGAS * s.T);
    if (o.iters) o.iters[c] = s.it;
    if (o.rr_iters) o.rr_iters[c] = s.rr;
}

// ------------------------------------------------------------ plain form
template <int NC>
__global__ __launch_bounds__(256) void k_flash_plain(const FluidDev* __restrict__ F, FlashArgs a) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const FluidConsts<double>& C = F->d;
    const bool srk = F->eos == OSR_EOS_SRK;
    SSCell<NC, double> s;
    double x[NC], y[NC];
    cell_init<NC, double>(C, a, c, s);
    double esum;
    do {
        esum = ss_pass<NC, double>(C, srk, s, x, y);
    } while (esum > FLASH_TOL && s.it < MAX_SS_ITER);
    cell_write<NC>(a, c, s, x, y);
}

// ------------------------------------------------------- persistent form
__device__ __forceinline__ unsigned lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

template <int NC>
__global__ __launch_bounds__(256) void k_flash_persistent(const FluidDev* __restrict__ F, FlashArgs a) {
    const FluidConsts<double>& C = F->d;
    const bool srk = F->eos == OSR_EOS_SRK;
    const unsigned lane = lane_id();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    SSCell<NC, double> s;
    double x[NC], y[NC];
    int64_t c = -1;
    bool active = false;
    bool exhausted = false;
    const int refill_min = a.refill_min;
    while (true) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= refill_min) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader) base = atomicAdd(a.counter, (unsigned long long)nidle);
            base = __shfl(base, leader, 64);
            if (base + (unsigned long long)nidle >= (unsigned long long)a.n) exhausted = true;
            if (!active) {
                const int64_t mine = (int64_t)base + __popcll(idle & lt_mask);
                if (mine < a.n) {
                    c = mine;
                    cell_init<NC, double>(C, a, c, s);
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            const double esum = ss_pass<NC, double>(C, srk, s, x, y);
            if (!(esum > FLASH_TOL && s.it < MAX_SS_ITER)) {
                cell_write<NC>(a, c, s, x, y);
                active = false;
            }
        }
    }
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
    const int i = threadIdx.x;
    if (i < count) p[i] = 0ull;
}

template <int NC>
static int launch_nc(osr_ctx* ctx, const FlashArgs& a, int mode, hipStream_t stream) {
    const int block = 256;
    if (mode == OSR_FLASH_PLAIN) {
        const int64_t grid = (a.n + block - 1) / block;
        hipLaunchKernelGGL(k_flash_plain<NC>, dim3((unsigned)grid), dim3(block), 0, stream,
                           ctx->d_fluid, a);
        return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
    }
    int per_cu = 0;
    OSR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_flash_persistent<NC>, block, 0));
    if (per_cu < 1) per_cu = 1;
    if (ctx->flash_blocks_per_cu > 0) per_cu = ctx->flash_blocks_per_cu;  // 4 waves per block = 1 per SIMD
    int64_t grid = (int64_t)per_cu * ctx->num_cu;
    const int64_t need = (a.n + block - 1) / block;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, ctx->d_counters, 8);
    hipLaunchKernelGGL(k_flash_persistent<NC>, dim3((unsigned)grid), dim3(block), 0, stream,
                       ctx->d_fluid, a);
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

bool nc_supported(int nc) {
    switch (nc) {
        case 2: case 3: case 4: case 5: case 6: case 8: case 10: case 12: return true;
        default: return false;
    }
}

int launch_flash(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                 const double* z, double sw, const osr_flash_out* out, int mode,
                 hipStream_t stream) {
    if (!ctx || n < 0 || !out) return OSR_E_ARG;
    if (mode != OSR_FLASH_PLAIN && mode != OSR_FLASH_PERSISTENT && mode != OSR_FLASH_MIXED)
        return OSR_E_ARG;
    if (n == 0) return OSR_OK;
    if (!P || !z || (!T && !(T0 > 0))) return OSR_E_ARG;
    FlashArgs a;
    a.n = n;
    a.T = T;
    a.T0 = T0;
    a.P = P;
    a.z = z;
    a.sw = sw;
    a.out = *out;
    a.counter = ctx->d_counters;
    a.refill_min = ctx->refill_min < 1 ? 1 : (ctx->refill_min > 64 ? 64 : ctx->refill_min);
    if (mode == OSR_FLASH_MIXED) mode = OSR_FLASH_PERSISTENT;  // mixed stage kernels: see flash_mixed
    switch (ctx->nc) {
        case 2: return launch_nc<2>(ctx, a, mode, stream);
        case 3: return launch_nc<3>(ctx, a, mode, stream);
        case 4: return launch_nc<4>(ctx, a, mode, stream);
        case 5: return launch_nc<5>(ctx, a, mode, stream);
        case 6: return launch_nc<6>(ctx, a, mode, stream);
        case 8: return launch_nc<8>(ctx, a, mode, stream);
        case 10: return launch_nc<10>(ctx, a, mode, stream);
        case 12: return launch_nc<12>(ctx, a, mode, stream);
        default: return OSR_E_NC;
    }
}

}  // namespace osr

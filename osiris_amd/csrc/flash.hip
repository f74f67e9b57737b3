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
#include "grp_dev.hpp"

namespace osr {

struct FlashArgs {
    int64_t n;
    const double* T;
    double T0;
    const double* P;
    const double* z;
    double sw;
    osr_flash_out out;
    unsigned long long* counter;   // [0] work-queue head, [1] park-queue tail, [2] park-queue head
    int refill_min;
    int park_after;                // 0: never park; else park cells still iterating after this many passes
    double* park;                  // parked-cell store, [NC + 5][cap] (K, L, Lb, it, rr, cell)
    int64_t park_cap;
    // mixed precision (park slots q index every array below)
    double* kb;                    // [NC][cap] K after the FP32 stage
    double* itb;                   // [2][cap] passes and RR iterations after the FP32 stage
    int64_t* list2;                // slots for the FP64 Newton polish
    int64_t* list3;                // slots sent back to FP64 SS from the park state
    const int64_t* glist;          // group kernel: slots to finish (nullptr = every parked slot)
    int glen_slot;                 // counter holding the length of glist (or of the park queue)
    float tol32;                   // FP32 stage stop: max_i |K_i'/K_i - 1| below this
    int max32;                     // FP32 stage pass cap
};

// counter slots
constexpr int CNT_WORK = 0, CNT_PARK = 1, CNT_GROUP = 2, CNT_B = 3, CNT_NT_TAIL = 4, CNT_NT_HEAD = 5,
              CNT_FB_TAIL = 6;
// Newton-stage outcomes (telemetry, read by osr_flash_stats)
constexpr int CNT_NT_OK = 7, CNT_NT_BOUNDARY = 8, CNT_NT_NONFINITE = 9, CNT_NT_SINGULAR = 10,
              CNT_NT_DIVERGED = 11, CNT_NT_MAXIT = 12, CNT_NT_REJECT = 13;
constexpr int NCOUNTERS = 16;

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
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_flash_plain(const FluidDev* __restrict__ F, FlashArgs a) {
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
template <int NC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_flash_persistent(const FluidDev* __restrict__ F, FlashArgs a) {
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
            if ((int)lane == leader) base = atomicAdd(a.counter + CNT_WORK, (unsigned long long)nidle);
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
        bool park = false;
        if (active) {
            const double esum = ss_pass<NC, double>(C, srk, s, x, y);
            if (!(esum > FLASH_TOL && s.it < MAX_SS_ITER)) {
                cell_write<NC>(a, c, s, x, y);
                active = false;
            } else if (a.park_after > 0 && s.it >= a.park_after) {
                park = true;
            }
        }
        // Slow cells move to the park queue (one atomic per wave) and are
        // finished by the lane-group kernel, which runs a pass with far
        // shorter latency than one lane can.
        const unsigned long long pk = __ballot(park);
        if (pk) {
            const int leader = __ffsll((long long)pk) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader) base = atomicAdd(a.counter + CNT_PARK, (unsigned long long)__popcll(pk));
            base = __shfl(base, leader, 64);
            if (park) {
                const int64_t q = (int64_t)base + __popcll(pk & lt_mask);
                const int64_t cap = a.park_cap;
#pragma unroll
                for (int i = 0; i < NC; i++) a.park[(int64_t)i * cap + q] = s.K[i];
                a.park[(int64_t)(NC + 0) * cap + q] = s.L;
                a.park[(int64_t)(NC + 1) * cap + q] = s.Lb;
                a.park[(int64_t)(NC + 2) * cap + q] = (double)s.it;
                a.park[(int64_t)(NC + 3) * cap + q] = (double)s.rr;
                a.park[(int64_t)(NC + 4) * cap + q] = (double)c;
                active = false;
            }
        }
    }
}

// ------------------------------------------------------- lane-group form
// G = 2S lanes finish one parked cell: lanes [0, S) evaluate the liquid
// phase, [S, 2S) the vapour phase, slot j of each half owns component j.
// Rachford-Rice sums and mixing sums are xor-butterflies inside a half (each
// lane adds the same operands in the same pairing, so every lane of the group
// holds bitwise the same L and the group's control flow is uniform); the
// mixing matrix-vector product gathers the half's S compositions by shuffle.
// A pass then costs one cubic, two logs and one exp per lane instead of a
// whole cell's worth per lane, which is what bounds the slowest cells.
// One reference SS pass (Init_Calcs.cpp:590-713) in group form.  Per lane:
// z, K of component j; returns ESum = |Lb - L| (group-uniform) and leaves x,
// y of component j (from the incoming K), Zl, Zg, L, and the updated K.
template <int NC, int S>
__device__ __forceinline__ double grp_pass(const FluidConsts<double>& C, bool srk, const Grp<S>& g, double z,
                                           double sa, double bj, double pA, double pB, double& K, double& L,
                                           double& Lb, double& x, double& y, double& Zl, double& Zg, int& rr) {
#pragma clang fp contract(off)
    const bool valid = g.valid;
    // Rachford-Rice Newton from 0.6
    double Lr = 0.6, step;
    int jr = 0;
    do {
        double td = 0;
        if (valid) td = (K >= 0) ? (1.0 - K) / (Lr + (1.0 - Lr) * K) : 1.0 / (Lr - 1.0);
        const double Fs = half_sum<S>(z * td);
        const double dF = -half_sum<S>(z * td * td);
        step = Fs / dF;
        Lr -= step;
        jr++;
    } while (step * step > NEWTON_TOL && jr < MAX_FLASH_N_ITR);
    rr += jr;
    if (Lr > 1) Lr = 1;
    else if (Lr < 0) Lr = 0;
    const double esum = fabs(Lb - Lr);
    Lb = Lr;
    L = Lr;
    if (K >= 0) {
        x = z / (L + K * (1.0 - L));
        y = K * x;
    } else {
        x = 0;
        y = z / (1.0 - L);
    }
    if (!valid) x = y = 0;
    const double cph = g.h == 0 ? x : y;
    const double u = sa * cph;
    double sacc = 0;
#pragma unroll
    for (int k2 = 0; k2 < NC; k2++) sacc += C.bic[g.jc * NC + k2] * __shfl(u, g.gbase + g.h * S + k2, 64);
    const double t = valid ? sa * sacc : 0.0;
    const double Ca = half_sum<S>(u * sacc);
    const double Cb = half_sum<S>(cph * (valid ? bj : 0.0));
    PhaseInv<double> ph;
    ph.A = Ca * pA;
    ph.B = Cb * pB;
    ph.Z = phase_z<double>(srk, ph.A, ph.B, g.h == 1);
    ph.invCa = 1.0 / Ca;
    ph.invCb = 1.0 / Cb;
    ph.lnZB = log(ph.Z - ph.B);
    const double DD = ph.A / (ph.B * (C.d1 - C.d2));
    ph.DDFF = DD * log((ph.Z + C.d2 * ph.B) / (ph.Z + C.d1 * ph.B));
    const double lp = lnphi<double>(ph, bj, t);
    const double lo = __shfl_xor(lp, S, 64);
    const double Zo = __shfl_xor(ph.Z, S, 64);
    Zl = g.h == 0 ? ph.Z : Zo;
    Zg = g.h == 0 ? Zo : ph.Z;
    const double av = g.h == 0 ? lp : lo;   // ln phi liquid of component j
    const double gv = g.h == 0 ? lo : lp;   // ln phi vapour of component j
    if (valid && z != 0) {
        double r;
        if (K > 0 && x != 0 && isfinite(x)) {
            r = exp(av - gv);
        } else {
            const double fL = x * exp(av);
            const double fG = y * exp(gv);
            r = (fG != 0) ? K * (fL / fG) : -1.0;
        }
        if (av != av) r = -1;
        if (gv != gv) r = 1;
        K = r;
    }
    return esum;
}

// Outputs of a cell finished in group form (Init_Calcs.cpp:717-736).
template <int NC, int S>
__device__ __forceinline__ void grp_write(const FlashArgs& a, const Grp<S>& g, int64_t c, double P, double T,
                                          double x, double y, double K, double L, double Zl, double Zg, int it,
                                          int rr) {
    const osr_flash_out& o = a.out;
    const int64_t n = a.n;
    if (g.h == 0 && g.valid) {
        if (o.x) o.x[(int64_t)g.jc * n + c] = x;
        if (o.y) o.y[(int64_t)g.jc * n + c] = y;
        if (o.K) o.K[(int64_t)g.jc * n + c] = K;
    }
    if (g.gl == 0) {
        int8_t lab = 0;
        if (fabs(L) < DZERO) lab = -1;
        else if (fabs(1 - L) < DZERO) lab = 1;
        if (o.L) o.L[c] = L;
        if (o.label) o.label[c] = lab;
        if (o.Zl) o.Zl[c] = Zl;
        if (o.Zg) o.Zg[c] = Zg;
        const double newD = Zl * L / (Zl * L + Zg * (1 - L));
        if (o.so) o.so[c] = newD * (1 - a.sw);
        if (o.sg) o.sg[c] = (1 - newD) * (1 - a.sw);
        if (o.rho_l) o.rho_l[c] = P / (Zl * RGAS * T);
        if (o.rho_g) o.rho_g[c] = P / (Zg * RGAS * T);
        if (o.iters) o.iters[c] = it;
        if (o.rr_iters) o.rr_iters[c] = rr;
    }
}

template <int NC, int S>
__global__ __launch_bounds__(256) void k_flash_group(const FluidDev* __restrict__ F, FlashArgs a) {
    constexpr int G = 2 * S;
    constexpr int GPW = 64 / G;    // groups per wave
    const FluidConsts<double>& C = F->d;
    const bool srk = F->eos == OSR_EOS_SRK;
    const Grp<S> g = grp_lane<NC, S>();
    const unsigned long long qlen = a.counter[a.glen_slot];
    const int64_t cap = a.park_cap;
    const double bj = C.b[g.jc];
    while (true) {
        const unsigned long long base = wave_fetch_add(a.counter + CNT_GROUP, (unsigned long long)GPW);
        if (base >= qlen) break;   // wave-uniform
        const unsigned long long qi = base + (unsigned long long)(g.lane / G);
        if (qi >= qlen) continue;  // group-uniform: this group has no cell this round
        const int64_t q = a.glist ? a.glist[qi] : (int64_t)qi;
        const int64_t c = (int64_t)a.park[(int64_t)(NC + 4) * cap + q];
        const double T = a.T ? a.T[c] : a.T0;
        const double P = a.P[c];
        const double z = g.valid ? a.z[(int64_t)g.jc * a.n + c] : 0.0;
        double K = g.valid ? a.park[(int64_t)g.jc * cap + q] : 0.0;
        double L = a.park[(int64_t)(NC + 0) * cap + q];
        double Lb = a.park[(int64_t)(NC + 1) * cap + q];
        int it = (int)a.park[(int64_t)(NC + 2) * cap + q];
        int rr = (int)a.park[(int64_t)(NC + 3) * cap + q];
        const double sa = g.valid ? C.sac[g.jc] * (1.0 + C.m[g.jc] * (1.0 - sqrt(T / C.Tc[g.jc]))) : 0.0;
        const double rt = RGAS * T;
        const double pB = P / rt, pA = pB / rt;
        double x = 0, y = 0, Zl = 0, Zg = 0;
        double esum;
        do {
            esum = grp_pass<NC, S>(C, srk, g, z, sa, bj, pA, pB, K, L, Lb, x, y, Zl, Zg, rr);
            it++;
        } while (esum > FLASH_TOL && it < MAX_SS_ITER);
        grp_write<NC, S>(a, g, c, P, T, x, y, K, L, Zl, Zg, it, rr);
    }
}

// ------------------------------------------------------- mixed precision
// Cells still iterating after the warm FP64 passes (parked above) run:
//  stage B  FP32 successive substitution (one lane per cell, work queue over
//           the park slots) until max_i |K_i'/K_i - 1| < tol32.  Cells whose
//           L comes within 1e-3 of 0 or 1, that stop improving, or that turn
//           non-finite go back to FP64 SS from their park state (the
//           reference trajectory), finished by the lane-group kernel;
//  stage C  FP64 Newton on ln K for the fixed point of the reference's own
//           SS map: L(K) is the Rachford-Rice Newton from 0.6 with the
//           reference's stopping rule, differentiated forward through its
//           iterations (that RR stops at |step| <= 1e-4, so the reference
//           converges to the fixed point of this inexact map, not to exact
//           fugacity equality -- a polish towards the latter would differ by
//           up to ~1e-7).  Residual and update are FP64; the Jacobian is
//           assembled analytically (derivatives of the cubic-EOS ln phi in x)
//           and factorised in FP32 in LDS, which only affects the rate.  One
//           final FP64 reference pass then produces the outputs with
//           SuccFalsh's semantics (x, y from K, K updated, L, Z, labels).
__device__ __forceinline__ int64_t wave_push(unsigned long long* tail, bool want, unsigned lane,
                                             unsigned long long lt_mask) {
    const unsigned long long m = __ballot(want);
    if (!m) return -1;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if ((int)lane == leader) base = atomicAdd(tail, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    return want ? (int64_t)base + __popcll(m & lt_mask) : -1;
}

template <int NC>
__global__ __launch_bounds__(256) void k_mixed_ss32(const FluidDev* __restrict__ F, FlashArgs a) {
    const FluidConsts<float>& C = F->f;
    const bool srk = F->eos == OSR_EOS_SRK;
    const unsigned lane = lane_id();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t len = (int64_t)a.counter[CNT_PARK];
    const int64_t cap = a.park_cap;
    SSCell<NC, float> s;
    float x[NC], y[NC];
    int64_t q = -1;
    int npass = 0;
    bool active = false;
    bool exhausted = len == 0;
    while (true) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= 16) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader) base = atomicAdd(a.counter + CNT_B, (unsigned long long)nidle);
            base = __shfl(base, leader, 64);
            if (base + (unsigned long long)nidle >= (unsigned long long)len) exhausted = true;
            if (!active) {
                const int64_t mine = (int64_t)base + __popcll(idle & lt_mask);
                if (mine < len) {
                    q = mine;
                    const int64_t c = (int64_t)a.park[(int64_t)(NC + 4) * cap + q];
                    const double T = a.T ? a.T[c] : a.T0;
                    const double P = a.P[c];
                    s.T = (float)T;
                    s.P = (float)P;
#pragma unroll
                    for (int i = 0; i < NC; i++) {
                        s.z[i] = (float)a.z[(int64_t)i * a.n + c];
                        s.K[i] = (float)a.park[(int64_t)i * cap + q];
                    }
                    sqrt_a<NC, float>(C, s.T, s.sa);
                    const double rt = RGAS * T;
                    s.pB = (float)(P / rt);
                    s.pA = (float)(P / rt / rt);
                    s.L = (float)a.park[(int64_t)(NC + 0) * cap + q];
                    s.Lb = (float)a.park[(int64_t)(NC + 1) * cap + q];
                    s.it = (int)a.park[(int64_t)(NC + 2) * cap + q];
                    s.rr = (int)a.park[(int64_t)(NC + 3) * cap + q];
                    s.Zl = s.Zg = 0;
                    npass = 0;
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        bool to_newton = false, to_fallback = false;
        if (active) {
            float Kold[NC];
#pragma unroll
            for (int i = 0; i < NC; i++) Kold[i] = s.K[i];
            ss_pass<NC, float>(C, srk, s, x, y);
            npass++;
            float dk = 0.f;
            bool ok = true;
#pragma unroll
            for (int i = 0; i < NC; i++) {
                if (s.z[i] != 0.f) {
                    ok = ok && (s.K[i] > 0.f) && isfinite(s.K[i]);
                    dk = fmaxf(dk, fabsf(s.K[i] / Kold[i] - 1.f));
                }
            }
            const float L = s.L;
            if (!ok || !(L > 1e-3f && L < 1.f - 1e-3f)) to_fallback = true;
            else if (dk < a.tol32) to_newton = true;
            // a cell not contracting to FP32 resolution within the cap may
            // still be drifting (towards a phase boundary, say): only the
            // reference trajectory decides where it ends, so FP64 SS resumes
            // from its park state
            else if (npass >= a.max32) to_fallback = true;
        }
        const int64_t p2 = wave_push(a.counter + CNT_NT_TAIL, to_newton, lane, lt_mask);
        if (to_newton) {
            a.list2[p2] = q;
#pragma unroll
            for (int i = 0; i < NC; i++) a.kb[(int64_t)i * cap + q] = (double)s.K[i];
            a.itb[q] = (double)s.it;
            a.itb[cap + q] = (double)s.rr;
            active = false;
        }
        const int64_t p3 = wave_push(a.counter + CNT_FB_TAIL, to_fallback, lane, lt_mask);
        if (to_fallback) {
            a.list3[p3] = q;
            active = false;
        }
    }
}

// Stage C in group form (2S lanes per cell, lane (h, j) = phase h,
// component j).  The unknowns are ln K; the residual is
//   g_i = ln K_i - ln phiL_i(x) + ln phiG_i(y)
// with x, y from the reference's RR (Newton from 0.6, |step| <= 1e-4), whose
// L(K) is differentiated forward through its iterations (lane j carries
// dL/d lnK_j).  The Jacobian
//   J = I - PhiL (dX + u_x w^T) + PhiG (dY + u_y w^T)
// (Phi = d ln phi / d x of the cubic EOS, dX = d x / d lnK at fixed L,
// u_x = d x / dL, w = dL / d lnK) is assembled row-wise: lane (h, i) forms
// phase h's part of row i from per-column factors gathered by shuffle, the
// two halves add, and the rows are factorised with partial pivoting across
// the lanes of a half (both halves hold identical rows and take identical
// decisions, so control flow stays group-uniform).  All FP64.  Phi of one
// phase, with E_i = 2 t_i/Ca - b_i/Cb, q = A/(B(d1-d2)), Lam = ln((Z+d2 B)/
// (Z+d1 B)) and Z_k from the implicit cubic:
//   Phi_ik = (b_i/Cb) U_k - V_k + E_i W_k - c1 t_i t_k + c2 a_ik,
//   U_k = Z_k - b_k (Z-1)/Cb + q Lam b_k/Cb, V_k = (Z_k - b_k pB)/(Z - B),
//   W_k = Lam q_k + q Lam_k, c1 = 4 q Lam/Ca^2, c2 = 2 q Lam/Ca.
constexpr int MIXED_NEWTON_MAX = 8;
constexpr double MIXED_ACCEPT = 1e-10;   // max |ln K'/K| of the final reference pass

// Built for two waves per SIMD (registers capped at 256 per lane; a few
// spill at nc >= 10): the polish's cross-lane chains hide better with the
// second wave.
template <int NC, int S>
__global__ __launch_bounds__(256, 2) void k_mixed_newton(const FluidDev* __restrict__ F, FlashArgs a) {
#pragma clang fp contract(off)
    constexpr int G = 2 * S;
    constexpr int GPW = 64 / G;
    const FluidConsts<double>& C = F->d;
    const bool srk = F->eos == OSR_EOS_SRK;
    const Grp<S> g = grp_lane<NC, S>();
    const int h = g.h, j = g.j;
    const bool valid = g.valid;
    const int hb = g.gbase + h * S;    // first lane of this lane's half
    const unsigned long long qlen = a.counter[CNT_NT_TAIL];
    const int64_t cap = a.park_cap;
    const double bj = C.b[g.jc];
    const double uE = srk ? 1.0 : 2.0, wE = srk ? 0.0 : 1.0;   // U = uE B, W^2 = wE B^2
    while (true) {
        const unsigned long long base = wave_fetch_add(a.counter + CNT_NT_HEAD, (unsigned long long)GPW);
        if (base >= qlen) break;   // wave-uniform
        const unsigned long long qi = base + (unsigned long long)(g.lane / G);
        if (qi >= qlen) continue;  // group-uniform
        const int64_t q = a.list2[qi];
        const int64_t c = (int64_t)a.park[(int64_t)(NC + 4) * cap + q];
        const double T = a.T ? a.T[c] : a.T0;
        const double P = a.P[c];
        const double z = valid ? a.z[(int64_t)g.jc * a.n + c] : 0.0;
        double K = valid ? a.kb[(int64_t)g.jc * cap + q] : 1.0;
        const double sa = valid ? C.sac[g.jc] * (1.0 + C.m[g.jc] * (1.0 - sqrt(T / C.Tc[g.jc]))) : 0.0;
        const double rt = RGAS * T;
        const double pB = P / rt, pA = pB / rt;
        const bool active = valid && z != 0;     // rows of absent components stay identity
        int rr = (int)a.itb[cap + q];
        int why = CNT_NT_MAXIT;
        bool ok = false;
        int itn = 0;
        for (; itn < MIXED_NEWTON_MAX; itn++) {
            // --- RR exactly as the reference pass, dL/dlnK_j carried by lane j
            double L = 0.6, step, dL = 0;
            int jr = 0;
            do {
                double td = 0, gK = 0;
                if (valid) {
                    const double D = L + (1.0 - L) * K;
                    td = (1.0 - K) / D;
                    gK = -z * K / (D * D);             // z_j d td_j / d lnK_j
                }
                const double Fs = half_sum<S>(z * td);
                const double dF = -half_sum<S>(z * td * td);
                const double sB = -half_sum<S>(z * td * td * td);
                step = Fs / dF;
                const double dFm = dF * dL + gK;
                const double ddFm = -2.0 * (sB * dL + td * gK);
                dL -= (dFm * dF - Fs * ddFm) / (dF * dF);
                L -= step;
                jr++;
            } while (step * step > NEWTON_TOL && jr < MAX_FLASH_N_ITR);
            rr += jr;
            if (!(L > 1e-3 && L < 1.0 - 1e-3)) {   // near or past a boundary: FP64 SS decides
                why = CNT_NT_BOUNDARY;
                break;
            }
            // --- compositions and their derivatives (own component)
            const double D = L + K * (1.0 - L);
            const double x = valid ? z / D : 0.0;
            const double iD2 = 1.0 / (D * D);
            const double dxd = valid ? -z * K * (1.0 - L) * iD2 : 0.0;   // d x_j / d lnK_j (L fixed)
            const double ux = valid ? -z * (1.0 - K) * iD2 : 0.0;        // d x_j / dL
            const double cph = h == 0 ? x : K * x;
            const double dd = h == 0 ? dxd : K * (dxd + x);
            const double uu = h == 0 ? ux : K * ux;
            // --- phase h: mixing, Z, ln phi_j, column factors of d ln phi / dx
            const double u = sa * cph;
            double sacc = 0;
#pragma unroll
            for (int k2 = 0; k2 < NC; k2++) sacc += C.bic[g.jc * NC + k2] * __shfl(u, hb + k2, 64);
            const double t = valid ? sa * sacc : 0.0;
            const double Ca = half_sum<S>(u * sacc);
            const double Cb = half_sum<S>(cph * (valid ? bj : 0.0));
            const double A = Ca * pA, B = Cb * pB;
            const double Z = phase_z<double>(srk, A, B, h == 1);
            const double iCa = 1.0 / Ca, iCb = 1.0 / Cb;
            const double q = A / (B * (C.d1 - C.d2));
            const double zd1 = Z + C.d1 * B, zd2 = Z + C.d2 * B;
            const double Lam = log(zd2 / zd1);
            const double lp = (bj * iCb) * (Z - 1.0) - log(Z - B) + q * Lam * (2.0 * t * iCa - bj * iCb);
            const double c2p = -1.0 + (uE - 1.0) * B;
            const double c1p = A - uE * B * B - uE * B - wE * B * B;
            const double FZ = 3.0 * Z * Z + 2.0 * c2p * Z + c1p;
            const double FA = Z - B;
            const double FB = (uE - 1.0) * Z * Z - (2.0 * (uE + wE) * B + uE) * Z - A + 3.0 * wE * B * B + 2.0 * wE * B;
            const double qL = q * Lam;
            double Uc = 0, Vc = 0, Wc = 0;
            if (valid) {
                const double Ak = 2.0 * t * pA, Bk = bj * pB;
                const double Zk = -(FA * Ak + FB * Bk) / FZ;
                const double qk = q * (Ak / A - Bk / B);
                const double Lk = (Zk + C.d2 * Bk) / zd2 - (Zk + C.d1 * Bk) / zd1;
                Uc = Zk - bj * (Z - 1.0) * iCb + qL * bj * iCb;
                Vc = (Zk - Bk) / (Z - B);
                Wc = Lam * qk + q * Lk;
            }
            const double c1 = 4.0 * qL * iCa * iCa, c2 = 2.0 * qL * iCa;
            const double Uu = half_sum<S>(Uc * uu), Vu = half_sum<S>(Vc * uu);
            const double Wu = half_sum<S>(Wc * uu), Tu = half_sum<S>(t * uu);
            // --- phase h's part of row i = j
            const double sgn = h == 0 ? -1.0 : 1.0;
            const double bi = bj * iCb;
            const double Ei = 2.0 * t * iCa - bi;
            const double c1i = c1 * t, c2i = c2 * sa;
            double row[NC];
            double bicu = 0;
#pragma unroll
            for (int k2 = 0; k2 < NC; k2++) {
                const int src = hb + k2;
                const double Ud = __shfl(Uc * dd, src, 64), Vd = __shfl(Vc * dd, src, 64);
                const double Wd = __shfl(Wc * dd, src, 64), Td = __shfl(t * dd, src, 64);
                const double Sd = __shfl(sa * dd, src, 64), Su = __shfl(sa * uu, src, 64);
                const double bik = C.bic[g.jc * NC + k2];
                row[k2] = sgn * (bi * Ud - Vd + Ei * Wd - c1i * Td + c2i * bik * Sd);
                bicu += bik * Su;
            }
            double rpart = sgn * (bi * Uu - Vu + Ei * Wu - c1i * Tu + c2i * bicu);
            // --- combine the halves: J row i and the rank-1 coefficient
            const double lo = __shfl_xor(lp, S, 64);
            const double lnL = h == 0 ? lp : lo, lnG = h == 0 ? lo : lp;
            rpart += __shfl_xor(rpart, S, 64);
            const double wj = (double)dL;
#pragma unroll
            for (int k2 = 0; k2 < NC; k2++) {
                const double other = __shfl_xor(row[k2], S, 64);
                const double wk = __shfl(wj, g.gbase + k2, 64);
                const double v = (h == 0 ? row[k2] + other : other + row[k2]) + rpart * wk;
                row[k2] = (k2 == j ? 1.0 : 0.0) + v;
                if (!active) row[k2] = (k2 == j) ? 1.0 : 0.0;
            }
            const double gi = active ? log(K) - lnL + lnG : 0.0;
            double rhs = -gi;
            const double bad = half_max<S>(isfinite(gi) ? 0.0 : 1.0);
            if (bad != 0.0) {
                why = CNT_NT_NONFINITE;
                break;
            }
            // --- LU with partial pivoting across the lanes of the half
            bool sing = false;
#pragma unroll
            for (int k2 = 0; k2 < NC; k2++) {
                double cand = (valid && j >= k2) ? fabs(row[k2]) : -1.0;
                if (cand != cand) cand = -1.0;
                int piv = j;
                half_argmax<S>(cand, piv);
                if (!(cand > 0.0)) sing = true;
                // swap rows k2 and piv (lane k2 takes piv's row and vice versa)
                const int src = (j == k2) ? piv : (j == piv ? k2 : j);
#pragma unroll
                for (int cc = k2; cc < NC; cc++) row[cc] = __shfl(row[cc], hb + src, 64);
                rhs = __shfl(rhs, hb + src, 64);
                // eliminate below the pivot row
                const double pkk = __shfl(row[k2], hb + k2, 64);
                const double prh = __shfl(rhs, hb + k2, 64);
                const double f = (valid && j > k2) ? row[k2] / pkk : 0.0;
#pragma unroll
                for (int cc = k2 + 1; cc < NC; cc++) row[cc] -= f * __shfl(row[cc], hb + k2, 64);
                rhs -= f * prh;
            }
            if (sing) {
                why = CNT_NT_SINGULAR;
                break;
            }
            // back substitution: lane k2 owns unknown k2
            double acc = rhs, delta = 0;
#pragma unroll
            for (int k2 = NC - 1; k2 >= 0; k2--) {
                const double dk = __shfl(acc / row[k2], hb + k2, 64);
                if (j == k2) delta = dk;
                if (j < k2) acc -= row[k2] * dk;
            }
            if (!active) delta = 0;
            const double dmax = half_max<S>(fabs(delta));
            if (!isfinite(dmax) || dmax > 1.0) {
                why = CNT_NT_DIVERGED;
                break;
            }
            if (active) K *= exp(delta);
            if (dmax < 1e-9) {
                ok = true;
                itn++;
                break;
            }
        }
        if (ok) {
            // final reference pass from the polished K; accept only a fixed point
            const double K0 = K;
            double L = 0, Lb = 10, x = 0, y = 0, Zl = 0, Zg = 0;
            grp_pass<NC, S>(C, srk, g, z, sa, bj, pA, pB, K, L, Lb, x, y, Zl, Zg, rr);
            const double gres = half_max<S>(active ? fabs(log(K / K0)) : 0.0);
            if (gres < MIXED_ACCEPT) {
                grp_write<NC, S>(a, g, c, P, T, x, y, K, L, Zl, Zg, (int)a.itb[q] + itn, rr);
                if (g.gl == 0) atomicAdd(a.counter + CNT_NT_OK, 1ull);
                continue;
            }
            why = CNT_NT_REJECT;
        }
        if (g.gl == 0) {
            atomicAdd(a.counter + why, 1ull);
            const unsigned long long p3 = atomicAdd(a.counter + CNT_FB_TAIL, 1ull);
            a.list3[p3] = q;
        }
    }
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
    const int i = threadIdx.x;
    if (i < count) p[i] = 0ull;
}

// Workspace for parked cells: park store [NC + 5][n], then (mixed) kb
// [NC][n], itb [2][n], list2 [n], list3 [n] -- every cell parks at most once.
static size_t park_words(int nc, int64_t n, bool mixed) {
    return (size_t)(nc + 5) * (size_t)n + (mixed ? (size_t)(nc + 4) * (size_t)n : 0);
}

// Fixed park workspace of the allocation-free entry point
// (osr_flash_device_noalloc).  One copy per device, held until the process
// ends.  Its device address is taken once per device, with that device
// current, when a context is created (flash_scratch_init), and only that
// address is ever handed to kernels or copies: on the host, `scratch` names a
// host-side shadow, not device memory.
static __device__ __align__(256) double scratch[FLASH_SCRATCH_WORDS];

constexpr int SCRATCH_MAX_DEVICES = 64;
static double* g_scratch_dev[SCRATCH_MAX_DEVICES];
static std::mutex g_scratch_mu;

int flash_scratch_init(int device, double** out) {
    if (!out || device < 0 || device >= SCRATCH_MAX_DEVICES) return OSR_E_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return OSR_E_HIP;
        g_scratch_dev[device] = (double*)p;
    }
    *out = g_scratch_dev[device];
    return OSR_OK;
}

int64_t flash_noalloc_max_cells(int nc, int mode) {
    if (mode == OSR_FLASH_PLAIN) return INT64_MAX;
    const size_t per_cell = park_words(nc, 1, mode == OSR_FLASH_MIXED);
    return (int64_t)(FLASH_SCRATCH_WORDS / per_cell);
}

static int ensure_park(osr_ctx* ctx, int nc, int64_t n, bool mixed) {
    const size_t words = park_words(nc, n, mixed);
    const size_t need = words * sizeof(double);
    if (need > ctx->park_bytes) {
        if (ctx->ws_frozen) return OSR_E_SIZE;
        if (ctx->d_park) (void)hipFree(ctx->d_park);
        ctx->d_park = nullptr;
        ctx->park_bytes = 0;
        OSR_HIP(hipMalloc(&ctx->d_park, need));
        ctx->park_bytes = need;
    }
    return OSR_OK;
}

template <int NC>
static int launch_nc(osr_ctx* ctx, const FlashArgs& a, int mode, bool noalloc, hipStream_t stream) {
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
    const bool mixed = mode == OSR_FLASH_MIXED;
    FlashArgs b = a;
    const int park_after = mixed ? (ctx->mixed_warm_iters < 1 ? 1 : ctx->mixed_warm_iters) : ctx->park_after;
    if (park_after > 0) {
        double* base;
        if (noalloc) {
            // size checked against FLASH_SCRATCH_WORDS by launch_flash
            if (!ctx->d_scratch) return OSR_E_HIP;
            base = ctx->d_scratch;
        } else {
            OSR_HIP_RET(ensure_park(ctx, NC, a.n, mixed));
            base = (double*)ctx->d_park;
        }
        const int64_t n = a.n;
        b.park_after = park_after;
        b.park = base;
        b.park_cap = n;
        if (mixed) {
            b.kb = base + (int64_t)(NC + 5) * n;
            b.itb = b.kb + (int64_t)NC * n;
            b.list2 = (int64_t*)(b.itb + 2 * n);
            b.list3 = b.list2 + n;
            b.glist = b.list3;
            b.glen_slot = CNT_FB_TAIL;
        }
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, ctx->d_counters, NCOUNTERS);
    hipLaunchKernelGGL(k_flash_persistent<NC>, dim3((unsigned)grid), dim3(block), 0, stream,
                       ctx->d_fluid, b);
    if (mixed) {
        int pb = 0;
        OSR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pb, k_mixed_ss32<NC>, block, 0));
        if (pb < 1) pb = 1;
        int64_t gB = (int64_t)pb * ctx->num_cu;
        if (gB > need) gB = need;
        if (gB < 1) gB = 1;
        hipLaunchKernelGGL(k_mixed_ss32<NC>, dim3((unsigned)gB), dim3(block), 0, stream, ctx->d_fluid, b);
        constexpr int SN = NC <= 8 ? 8 : 16;
        const int64_t gpbN = 256 / (2 * SN);
        int64_t gN = (int64_t)ctx->num_cu * 8;
        const int64_t maxN = (a.n + gpbN - 1) / gpbN;
        if (gN > maxN) gN = maxN;
        if (gN < 1) gN = 1;
        hipLaunchKernelGGL((k_mixed_newton<NC, SN>), dim3((unsigned)gN), dim3(256), 0, stream, ctx->d_fluid, b);
    }
    if (park_after > 0) {
        constexpr int S = NC <= 8 ? 8 : 16;
        const int64_t groups_per_block = 256 / (2 * S);
        int64_t gb = (int64_t)ctx->num_cu * 4;
        const int64_t maxb = (a.n + groups_per_block - 1) / groups_per_block;
        if (gb > maxb) gb = maxb;
        if (gb < 1) gb = 1;
        hipLaunchKernelGGL((k_flash_group<NC, S>), dim3((unsigned)gb), dim3(256), 0, stream, ctx->d_fluid, b);
    }
    return hipGetLastError() == hipSuccess ? OSR_OK : OSR_E_HIP;
}

bool nc_supported(int nc) {
    switch (nc) {
#define CASE(NC) case NC:
        OSR_NC_LIST(CASE)
#undef CASE
        return true;
        default: return false;
    }
}

int launch_flash(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                 const double* z, double sw, const osr_flash_out* out, int mode,
                 bool noalloc, hipStream_t stream) {
    if (!ctx || n < 0 || !out) return OSR_E_ARG;
    if (mode != OSR_FLASH_PLAIN && mode != OSR_FLASH_PERSISTENT && mode != OSR_FLASH_MIXED)
        return OSR_E_ARG;
    if (noalloc && n > flash_noalloc_max_cells(ctx->nc, mode)) return OSR_E_SIZE;
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
    a.park_after = 0;
    a.park = nullptr;
    a.park_cap = 0;
    a.kb = nullptr;
    a.itb = nullptr;
    a.list2 = nullptr;
    a.list3 = nullptr;
    a.glist = nullptr;
    a.glen_slot = CNT_PARK;
    a.tol32 = ctx->mixed_tol32;
    a.max32 = ctx->mixed_max32;
    switch (ctx->nc) {
#define CASE(NC) case NC: return launch_nc<NC>(ctx, a, mode, noalloc, stream);
        OSR_NC_LIST(CASE)
#undef CASE
        default: return OSR_E_NC;
    }
}

}  // namespace osr

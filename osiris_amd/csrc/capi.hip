// capi.hip -- the extern "C" entry points of libosiris_amd.so.
#include <cmath>
#include <cstring>
#include <string>
#include <new>

#include "osr_internal.hpp"

namespace {

// Temperature-independent EOS constants (EOS_Init, compsim/Init_Calcs.cpp:66-97):
// ac_i and m_i are kept apart so each cell can form a_i(T) for its own T.
template <class R>
void fill_consts(osr::FluidConsts<R>& C, const osr_fluid* f) {
    const int nc = f->nc;
    const bool srk = f->eos == OSR_EOS_SRK;
    const double R2 = osr::RGAS * osr::RGAS;
    for (int i = 0; i < nc; i++) {
        const double Tc = f->Tc[i], Pc = f->Pc[i], w = f->omega[i];
        double ac, m, b;
        if (srk) {
            ac = 0.42747 * R2 * Tc * Tc / Pc;
            m = 0.48508 + 1.55171 * w - 0.15613 * w * w;
            b = 0.08664 * osr::RGAS * Tc / Pc;
        } else {
            ac = 0.457235 * R2 * Tc * Tc / Pc;
            m = 0.3796 + 1.485 * w - 0.1644 * w * w + 0.01667 * w * w * w;
            b = 0.077796 * osr::RGAS * Tc / Pc;
        }
        C.Tc[i] = R(Tc);
        C.Pc[i] = R(Pc);
        C.sac[i] = R(std::sqrt(ac));
        C.m[i] = R(m);
        C.b[i] = R(b);
        C.wk[i] = R(5.37 * (1 + w));
        for (int j = 0; j < nc; j++) C.bic[i * nc + j] = R(1.0 - f->kij[i * nc + j]);
    }
    C.d1 = R(srk ? 1.0 : 1.0 + osr::SQRT2);
    C.d2 = R(srk ? 0.0 : 1.0 - osr::SQRT2);
}

}  // namespace

extern "C" {

const char* osr_version(void) { return "osiris_amd 0.1 gfx950"; }

int osr_supported_nc(int32_t nc) { return osr::nc_supported(nc) ? 1 : 0; }

int osr_ctx_create(const osr_fluid* fluid, osr_ctx** out) {
    if (!fluid || !out || !fluid->Tc || !fluid->Pc || !fluid->omega || !fluid->kij) return OSR_E_ARG;
    if (fluid->nc < 1 || fluid->nc > osr::NCMAX) return OSR_E_ARG;
    if (fluid->eos != OSR_EOS_PR && fluid->eos != OSR_EOS_SRK) return OSR_E_ARG;
    if (!osr::nc_supported(fluid->nc)) return OSR_E_NC;
    for (int i = 0; i < fluid->nc; i++)
        if (!(fluid->Tc[i] > 0) || !(fluid->Pc[i] > 0)) return OSR_E_ARG;
    osr_ctx* c = new (std::nothrow) osr_ctx();
    if (!c) return OSR_E_ARG;
    c->nc = fluid->nc;
    c->eos = fluid->eos;
    c->host_fluid.nc = fluid->nc;
    c->host_fluid.eos = fluid->eos;
    fill_consts<double>(c->host_fluid.d, fluid);
    fill_consts<float>(c->host_fluid.f, fluid);
    hipDeviceProp_t prop;
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
        delete c;
        return OSR_E_HIP;
    }
    c->num_cu = prop.multiProcessorCount;
    if (hipMalloc(&c->d_fluid, sizeof(osr::FluidDev)) != hipSuccess ||
        hipMalloc(&c->d_counters, 16 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemcpy(c->d_fluid, &c->host_fluid, sizeof(osr::FluidDev), hipMemcpyHostToDevice) != hipSuccess) {
        osr_ctx_destroy(c);
        return OSR_E_HIP;
    }
    // The allocation-free flash's workspace address on this device; if it
    // cannot be had, d_scratch stays null and osr_flash_device_noalloc
    // returns OSR_E_HIP (the allocating entry points are unaffected).
    if (osr::flash_scratch_init(c->device, &c->d_scratch) != OSR_OK) c->d_scratch = nullptr;
    *out = c;
    return OSR_OK;
}

int osr_ctx_destroy(osr_ctx* ctx) {
    if (!ctx) return OSR_E_ARG;
    int dev = 0;
    hipGetDevice(&dev);
    hipSetDevice(ctx->device);
    if (ctx->d_fluid) hipFree(ctx->d_fluid);
    if (ctx->d_counters) hipFree(ctx->d_counters);
    if (ctx->d_queue) hipFree(ctx->d_queue);
    if (ctx->d_plist) hipFree(ctx->d_plist);
    if (ctx->d_park) hipFree(ctx->d_park);
    hipSetDevice(dev);
    delete ctx;
    return OSR_OK;
}

int osr_flash_device(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                     const double* z, double sw, const osr_flash_out* out, int mode,
                     void* stream) {
    return osr::launch_flash(ctx, n, T, T0, P, z, sw, out, mode, false, (hipStream_t)stream);
}

int osr_flash_device_noalloc(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                             const double* z, double sw, const osr_flash_out* out, int mode,
                             void* stream) {
    return osr::launch_flash(ctx, n, T, T0, P, z, sw, out, mode, true, (hipStream_t)stream);
}

int64_t osr_flash_noalloc_max_cells(int32_t nc, int32_t mode) {
    if (!osr::nc_supported(nc)) return OSR_E_NC;
    if (mode != OSR_FLASH_PLAIN && mode != OSR_FLASH_PERSISTENT && mode != OSR_FLASH_MIXED) return OSR_E_ARG;
    return osr::flash_noalloc_max_cells(nc, mode);
}

int osr_psat_device(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                    const int8_t* kind, const double* x, const double* y, double* psat,
                    double* xm, void* stream) {
    return osr::launch_psat(ctx, n, T, T0, P, kind, x, y, psat, xm, (hipStream_t)stream);
}

int osr_ctx_freeze_workspaces(osr_ctx* ctx, int frozen) {
    if (!ctx || (frozen != 0 && frozen != 1)) return OSR_E_ARG;
    ctx->ws_frozen = frozen;
    return OSR_OK;
}

int osr_tune_set(osr_ctx* ctx, const char* key, double value) {
    if (!ctx || !key) return OSR_E_ARG;
    const std::string k(key);
    if (k == "flash.refill_min") {
        if (value < 1 || value > 64) return OSR_E_ARG;
        ctx->refill_min = (int)value;
    } else if (k == "flash.mixed_warm_iters") {
        if (value < 1 || value > 100) return OSR_E_ARG;
        ctx->mixed_warm_iters = (int)value;
    } else if (k == "flash.mixed_tol32") {
        if (!(value > 0 && value < 1)) return OSR_E_ARG;
        ctx->mixed_tol32 = (float)value;
    } else if (k == "flash.mixed_max32") {
        if (value < 1 || value > 500) return OSR_E_ARG;
        ctx->mixed_max32 = (int)value;
    } else if (k == "flash.park_after") {
        if (value < 0 || value > 500) return OSR_E_ARG;
        ctx->park_after = (int)value;
    } else if (k == "reduced.dprops_wg" || k == "reduced.own_wg") {
        if (value != 64 && value != 128 && value != 256) return OSR_E_ARG;
        (k == "reduced.dprops_wg" ? ctx->red_dprops_wg : ctx->red_own_wg) = (int)value;
    } else if (k == "assembly.xcd_remap") {
        if (value != 0 && value != 1) return OSR_E_ARG;
        ctx->xcd_remap = (int)value;
    } else if (k == "reduced.local_rows") {
        if (value != 0 && value != 1) return OSR_E_ARG;
        ctx->red_local_form = (int)value;
    } else if (k == "flash.blocks_per_cu") {
        if (value < 0 || value > 64) return OSR_E_ARG;
        ctx->flash_blocks_per_cu = (int)value;
    } else {
        return OSR_E_ARG;
    }
    return OSR_OK;
}

int osr_flash_stats(osr_ctx* ctx, int64_t* out, int32_t nout, void* stream) {
    if (!ctx || (!out && nout > 0) || nout < 0) return OSR_E_ARG;
    unsigned long long h[16];
    hipStream_t st = (hipStream_t)stream;
    OSR_HIP(hipMemcpyAsync(h, ctx->d_counters, sizeof(h), hipMemcpyDeviceToHost, st));
    OSR_HIP(hipStreamSynchronize(st));
    // parked, newton, fallback, newton accepted, then the Newton exits:
    // boundary, non-finite, singular, diverged, iteration cap, final-pass reject
    const int slot[10] = {1, 4, 6, 7, 8, 9, 10, 11, 12, 13};
    for (int i = 0; i < nout && i < 10; i++) out[i] = (int64_t)h[slot[i]];
    for (int i = 10; i < nout; i++) out[i] = 0;
    return OSR_OK;
}

int osr_phase_ctrl_device(osr_ctx* ctx, int64_t n, const double* T, double T0, int8_t* label,
                          const double* Po, const double* Pg, double* x, double* y, double* S,
                          double* psat, void* stream) {
    return osr::phase_ctrl(ctx, n, T, T0, label, Po, Pg, x, y, S, psat, (hipStream_t)stream);
}

}  // extern "C"

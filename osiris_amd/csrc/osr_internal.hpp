// osr_internal.hpp -- host-side context and helpers shared by the HIP
// translation units of libosiris_amd.so.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/osiris_amd.h"
#include "eos_dev.hpp"

struct osr_ctx {
    int device = 0;
    int nc = 0;
    int eos = 0;
    int num_cu = 0;
    osr::FluidDev host_fluid{};
    osr::FluidDev* d_fluid = nullptr;
    // Work-queue heads, one per queue used inside a call (zeroed by a fill
    // kernel on the call's stream, so the call stays graph-capturable).
    unsigned long long* d_counters = nullptr;
    // Compaction queue of unconverged flash cells (grown on demand).
    void* d_queue = nullptr;
    size_t queue_bytes = 0;
    // Tuning knobs (set by osr_tune_set; defaults chosen from measurement).
    int flash_block = 256;
    int flash_blocks_per_cu = 0;   // 0 = from occupancy query
    int refill_min = 48;
    int park_after = 40;            // persistent flash: passes before a cell moves to the lane-group kernel
    void* d_park = nullptr;
    size_t park_bytes = 0;
    // Saturation-pressure work list: [count (8 bytes) | cell indices int32[n]]
    void* d_plist = nullptr;
    size_t plist_bytes = 0;
    // Device address of this device's fixed park workspace (flash.hip
    // `scratch`), taken at context creation; used by osr_flash_device_noalloc.
    double* d_scratch = nullptr;
    // while set, calls that would grow a workspace above (park store, psat
    // list, phase-control queue) refuse with OSR_E_SIZE instead of freeing
    // buffers a captured graph may still address (osr_ctx_freeze_workspaces)
    int ws_frozen = 0;
    // local rows of the reduced system straight from the state: 1 = closed-form
    // EOS derivatives and elimination in one lane group per block
    // (k_red_locel), 0 = forward mode per (block, unknown) then elimination
    int red_local_form = 1;
    // reduced-system kernels: 1 = each XCD takes a contiguous eighth of the
    // grid's workgroups (neighbouring cells share an L2), 0 = dispatch order
    int xcd_remap = 0;
    // workgroup sizes (64, 128 or 256 lanes) of the reduced system's
    // direction sweep and balance-row kernels; the one-wave-per-SIMD
    // balance-row kernel defaults to one-wave workgroups, so each SIMD takes
    // its next wave as soon as its own finishes
    int red_dprops_wg = 256;
    int red_own_wg = 64;
    int mixed_warm_iters = 3;
    float mixed_tol32 = 1e-4f;
    int mixed_max32 = 150;
};

// Component counts compiled in: every nc from 2 to 12.  Each kernel keeps a
// cell's per-component vectors in registers with its component loops
// unrolled, so nc is a template argument; OSR_NC_LIST(X) expands X(nc) for
// each compiled count and every dispatch switch is built from it.
#define OSR_NC_LIST(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

#define OSR_HIP(call)                                   \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return OSR_E_HIP;         \
    } while (0)

#define OSR_HIP_RET(call)                               \
    do {                                                \
        const int r_ = (call);                          \
        if (r_ != OSR_OK) return r_;                    \
    } while (0)

namespace osr {
// Size of flash.hip's fixed park workspace, in doubles (64 MiB).
constexpr size_t FLASH_SCRATCH_WORDS = (size_t)8 << 20;
int launch_flash(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                 const double* z, double sw, const osr_flash_out* out, int mode,
                 bool noalloc, hipStream_t stream);
int flash_scratch_init(int device, double** out);
int64_t flash_noalloc_max_cells(int nc, int mode);
int launch_psat(osr_ctx* ctx, int64_t n, const double* T, double T0, const double* P,
                const int8_t* kind, const double* x, const double* y, double* psat,
                double* xm, hipStream_t stream);
int phase_ctrl(osr_ctx* ctx, int64_t n, const double* T, double T0, int8_t* label, const double* Po,
               const double* Pg, double* x, double* y, double* S, double* psat_out, hipStream_t st);
bool nc_supported(int nc);
// assembly.hip, for the Newton step
void asm_info(const osr_asm* a, osr_ctx** ctx, int64_t* n, int* nc, double* T);
void asm_dims(const osr_asm* a, int* nx, int* ny, int* nz);
int asm_sweep(osr_asm* a, const osr_state_dev* s, double* rho_out, hipStream_t st);
int asm_take_rho(osr_asm* a, double* rho_out, hipStream_t st);
int asm_apply_sweep(osr_asm* a, const osr_state_mut_dev* m, const double* rho_prev, const double* psat,
                    const double* xm, double* rho_next, int64_t* nrows, hipStream_t st);
int phase_psat(osr_ctx* ctx, int64_t n, const double* T, double T0, const int8_t* label, const double* Po,
               const double* Pg, const double* x, const double* y, double* psat, double* xm, hipStream_t st);
}  // namespace osr

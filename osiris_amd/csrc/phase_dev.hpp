// phase_dev.hpp -- the per-cell rule of PhaseCtrl / fPhaseCtrl
// (Do_Cycle.cpp:1558-1643 / 1644-1722), shared by the phase-control kernel
// and the Newton step's fused update-and-sweep kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace osr {

constexpr double PHASE_ENV = 0.95;   // PAHSE_ENV_BOUNDARY, Functions.h:11
constexpr double EPS_SAT = 1e-3;     // Functions.h:226
constexpr double EPS_SAT_FRAC = 0;   // Functions.h:12

// Cell c of a continuum with n cells: a gas cell at Pg <= 0.95 x its dew
// pressure or a liquid cell at Po <= 0.95 x its bubble pressure turns
// two-phase with the incipient phase's composition xm and saturation 1e-3;
// otherwise a cell without oil (gas) saturation turns gas (liquid); the
// saturations are renormalised.  Writes label, x or y, S in place.
template <int NC>
__device__ __forceinline__ void phase_apply_cell(int64_t n, int64_t c, int8_t* label, double Po, double Pg,
                                                 double* x, double* y, double* S, double ps, const double* xm) {
    int L = label[c];
    double Sw = S[c], So = S[n + c], Sg = S[2 * n + c];
    bool done = false;
    if (L == -1 && Pg <= PHASE_ENV * ps) {
        So = EPS_SAT;
        Sg -= EPS_SAT;
        L = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) x[(int64_t)i * n + c] = xm[(int64_t)i * n + c];
        done = true;
    } else if (L == 1 && Po <= PHASE_ENV * ps) {
        Sg = EPS_SAT;
        So -= EPS_SAT;
        L = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) y[(int64_t)i * n + c] = xm[(int64_t)i * n + c];
        done = true;
    }
    if (!done) {
        if (So <= EPS_SAT * EPS_SAT_FRAC) {
            So = 0;
            Sg = 1 - Sw;
            L = -1;
        }
        if (Sg <= EPS_SAT * EPS_SAT_FRAC) {
            Sg = 0;
            So = 1 - Sw;
            L = 1;
        }
    }
    const double inv = 1.0 / (Sw + So + Sg);
    S[c] = Sw * inv;
    S[n + c] = So * inv;
    S[2 * n + c] = Sg * inv;
    label[c] = (int8_t)L;
}

}  // namespace osr

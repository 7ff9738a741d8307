// ao_plan.hpp -- host-side planning of the orbital loop for the AO kernels.
//
// The reference maps orbital o to radial iR[o] and harmonic iY[o]
// (src/orbitals/atomicorbitals.jl:56-75).  For the maps AtomicOrbitals builds, the
// orbitals of radial k are 2l+1 consecutive columns with harmonics l^2 .. l^2+2l, so
// the kernels walk "runs": (k, l, first column).  Any other map is still accepted:
// orbitals that do not form a full shell become single-orbital runs (l < 0 encodes
// harmonic -1-l).  The table travels by value in the kernel arguments.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/aok.h"

namespace aok {

constexpr int AO_MAXRUN = 192;
constexpr int AO_MAXRAD = 256;

// Runs packed in 32-bit words so the kernels read them with scalar loads (the table is
// indexed by a wave-uniform counter):  k | (l + 128) << 8 | o0 << 16, and r^poly of
// radial k in byte k % 4 of kpoly4[k / 4].
struct AoRunTable {
    int nrun;
    int norb;
    uint32_t run[AO_MAXRUN];
    uint32_t kpoly4[AO_MAXRAD / 4];
};

__host__ __device__ inline int run_k(uint32_t w) { return static_cast<int>(w & 0xffu); }
__host__ __device__ inline int run_l(uint32_t w) { return static_cast<int>((w >> 8) & 0xffu) - 128; }
__host__ __device__ inline int run_o0(uint32_t w) { return static_cast<int>(w >> 16); }
__host__ __device__ inline int run_poly(const AoRunTable &t, int k)
{
    return static_cast<int>((t.kpoly4[k >> 2] >> (8 * (k & 3))) & 0xffu);
}

struct AoArgs {
    const void *X;
    const void *sidx;
    int itype;
    int64_t nX;
    const void *zeta;
    const void *D;
    int nRad, K, NZ;
    int L;
    double scale;
    int decay;
    void *P;
    void *dP;
    const void *dPin;
};

// Build the run table; returns an empty string on success, else the error message.
inline std::string build_runs(const int64_t *poly, const int64_t *iR, const int64_t *iY, int64_t nOrb,
                              int64_t nRad, int L, AoRunTable &t)
{
    t.nrun = 0;
    t.norb = static_cast<int>(nOrb);
    if (nRad > AO_MAXRAD) return "nRad exceeds " + std::to_string(AO_MAXRAD);
    if (nOrb > 65535) return "nOrb too large";
    for (int q = 0; q < AO_MAXRAD / 4; ++q) t.kpoly4[q] = 0u;
    for (int64_t k = 0; k < nRad; ++k) {
        const int64_t p = poly ? poly[k] : 0;
        if (p < 0 || p > 64) return "poly out of range";
        t.kpoly4[k >> 2] |= static_cast<uint32_t>(p) << (8 * (k & 3));
    }
    const int nY = (L + 1) * (L + 1);
    int64_t o = 0;
    while (o < nOrb) {
        const int64_t k = iR[o] - 1, y = iY[o] - 1;
        if (k < 0 || k >= nRad) return "iR[" + std::to_string(o) + "] out of range";
        if (y < 0 || y >= nY) return "iY[" + std::to_string(o) + "] out of range for L";
        if (t.nrun >= AO_MAXRUN) return "too many orbital runs";
        // full shell?  y must be l^2 and the next 2l columns must follow
        int l = 0;
        while ((l + 1) * (l + 1) <= y) ++l;
        bool full = (y == l * l) && (o + 2 * l < nOrb);
        if (full)
            for (int j = 0; j <= 2 * l; ++j)
                if (iR[o + j] - 1 != k || iY[o + j] - 1 != y + j) {
                    full = false;
                    break;
                }
        // l >= 0: full shell of 2l + 1 columns; l < 0: one column with harmonic -1 - l
        const int lenc = full ? l : static_cast<int>(-1 - y);
        t.run[t.nrun++] = static_cast<uint32_t>(k) | (static_cast<uint32_t>(lenc + 128) << 8) |
                          (static_cast<uint32_t>(o) << 16);
        o += full ? 2 * l + 1 : 1;
    }
    return std::string();
}

template <typename T>
aok_status ao_eval_launch(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs);
template <typename T>
aok_status ao_pullback_launch(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs, void *dX, void *dzeta, void *dD);

}  // namespace aok

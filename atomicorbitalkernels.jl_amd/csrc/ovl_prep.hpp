// ovl_prep.hpp -- device helpers shared by the 2- and 3-centre planning kernels.
#pragma once

#include "aok_common.hpp"
#include "ovl_plan.hpp"

namespace aok {

// Bucket offsets from per-block histograms, inside one 256-thread block: for every key a
// block-wide exclusive scan over the hist blocks (thread t owns a contiguous run of
// blocks), then the scan over keys; blk_off[b][key] = first output slot of block b's
// elements with that key (stable order).
__device__ inline void bucket_offsets(int NK, int nblk_h, const int *__restrict__ blk_hist, int *__restrict__ blk_off,
                               int *__restrict__ bcnt, int *__restrict__ boff, int *__restrict__ toff)
{
    __shared__ int tsum[256];
    __shared__ int sboff[OVL_MAXNZ * OVL_MAXNZ > 512 ? OVL_MAXNZ * OVL_MAXNZ : 512];
    const int per = (nblk_h + 255) / 256;
    const int b0 = threadIdx.x * per, b1 = min(nblk_h, b0 + per);
    for (int key = 0; key < NK; ++key) {
        int s = 0;
        for (int b = b0; b < b1; ++b) s += blk_hist[static_cast<int64_t>(b) * NK + key];
        tsum[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int q = 0; q < 256; ++q) {
                const int v = tsum[q];
                tsum[q] = run;
                run += v;
            }
            bcnt[key] = run;
        }
        __syncthreads();
        int run = tsum[threadIdx.x];
        for (int b = b0; b < b1; ++b) {
            blk_off[static_cast<int64_t>(b) * NK + key] = run;
            run += blk_hist[static_cast<int64_t>(b) * NK + key];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int o = 0, t = 0;
        for (int key = 0; key < NK; ++key) {
            boff[key] = o;
            sboff[key] = o;
            toff[key] = t;
            o += bcnt[key];
            t += (bcnt[key] + OVL_TILE - 1) / OVL_TILE;
        }
        toff[NK] = t;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NK * nblk_h; q += blockDim.x) blk_off[q] += sboff[q % NK];
}

}  // namespace aok

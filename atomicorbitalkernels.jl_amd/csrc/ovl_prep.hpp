// ovl_prep.hpp -- device helpers shared by the 2- and 3-centre planning kernels.
#pragma once

#include "aok_common.hpp"
#include "ovl_plan.hpp"

namespace aok {

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

__device__ __forceinline__ int64_t wave_incl_scan64(int64_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Fixed-order sums of two doubles per thread over a 256-thread block, returned to every
// thread: a butterfly inside each wave (lanes i and i ^ d add the same two operands, so
// every lane holds the identical wave total), then the four wave totals in wave order.
// `red` holds 8 doubles.
__device__ __forceinline__ void block_sum256x2(double &c, double &a, double *red)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        c += __shfl_xor(c, d, 64);
        a += __shfl_xor(a, d, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave] = c;
        red[4 + wave] = a;
    }
    __syncthreads();
    c = (red[0] + red[1]) + (red[2] + red[3]);
    a = (red[4] + red[5]) + (red[6] + red[7]);
    __syncthreads();
}

// Exclusive prefix sum over a block of whole waves; `wtot` holds one int64 per wave.
// `total` receives the block total.
__device__ inline int64_t block_excl_scan64(int64_t v, int64_t *wtot, int64_t &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int64_t incl = wave_incl_scan64(v);
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    int64_t base = 0, tot = 0;
    for (int w = 0; w < nwave; ++w) {
        if (w < wave) base += wtot[w];
        tot += wtot[w];
    }
    total = tot;
    __syncthreads();
    return base + incl - v;
}

constexpr int OVL_MAXKEYS = 512;  // species pairs (16^2) or triples (8^3)

// Bucket offsets from per-block histograms, inside one block of whole waves: each wave
// sums its run of histogram blocks per key, the waves' exclusive prefix per key and the
// scan over keys give boff (first element of each key's bucket), toff (first tile of each
// key, 64 elements per tile) and blk_off[b][key] = first output slot of block b's elements
// with that key (stable order).
__device__ inline void bucket_offsets(int NK, int nblk_h, const int *__restrict__ blk_hist, int *__restrict__ blk_off,
                                      int *__restrict__ bcnt, int *__restrict__ boff, int *__restrict__ toff)
{
    // Wave w takes the histogram rows of blocks [w * per, (w + 1) * per), its lanes 64
    // consecutive keys, so every load and store of a row segment is coalesced (one lane per
    // block instead read one 4-byte word per cache line: C2's 1024 blocks x 36 keys took
    // 75-85 us on the one CU).  Rows are read PF at a time, all in flight together.
    constexpr int PF = 16;
    constexpr int MAXW = 16;
    __shared__ int s_cnt[OVL_MAXKEYS];
    __shared__ int s_base[OVL_MAXKEYS];
    __shared__ int wpre[MAXW * OVL_MAXKEYS];  // per wave and key: its blocks' sum, then their offset
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwave = min(static_cast<int>(blockDim.x >> 6), MAXW);
    const int per = (nblk_h + nwave - 1) / nwave;
    const int b0 = wave * per, b1 = min(nblk_h, b0 + per);
    if (wave < nwave) {
        for (int k0 = 0; k0 < NK; k0 += 64) {
            const int key = k0 + lane;
            int s = 0;
            if (key < NK)
                for (int b = b0; b < b1; b += PF) {
                    int v[PF];
#pragma unroll
                    for (int u = 0; u < PF; ++u) v[u] = b + u < b1 ? blk_hist[static_cast<int64_t>(b + u) * NK + key] : 0;
#pragma unroll
                    for (int u = 0; u < PF; ++u) s += v[u];
                }
            if (key < NK) wpre[wave * OVL_MAXKEYS + key] = s;
        }
    }
    __syncthreads();
    // per key: the waves' exclusive offsets and the key's total
    for (int key = threadIdx.x; key < NK; key += blockDim.x) {
        int c = 0;
        for (int w = 0; w < nwave; ++w) {
            const int x = wpre[w * OVL_MAXKEYS + key];
            wpre[w * OVL_MAXKEYS + key] = c;
            c += x;
        }
        s_cnt[key] = c;
    }
    __syncthreads();
    if (wave == 0) {
        int carry_o = 0, carry_t = 0;
        for (int k0 = 0; k0 < NK; k0 += 64) {
            const int key = k0 + lane;
            const int c = key < NK ? s_cnt[key] : 0;
            const int t = (c + OVL_TILE - 1) / OVL_TILE;
            const int ic = wave_incl_scan(c), it = wave_incl_scan(t);
            if (key < NK) {
                const int o = carry_o + ic - c;
                s_base[key] = o;
                bcnt[key] = c;
                boff[key] = o;
                toff[key] = carry_t + it - t;
            }
            carry_o += __shfl(ic, 63, 64);
            carry_t += __shfl(it, 63, 64);
        }
        if (lane == 0) toff[NK] = carry_t;
    }
    __syncthreads();
    if (wave < nwave) {
        for (int k0 = 0; k0 < NK; k0 += 64) {
            const int key = k0 + lane;
            if (key >= NK) continue;
            int run = s_base[key] + wpre[wave * OVL_MAXKEYS + key];
            for (int b = b0; b < b1; b += PF) {
                int v[PF];
#pragma unroll
                for (int u = 0; u < PF; ++u) v[u] = b + u < b1 ? blk_hist[static_cast<int64_t>(b + u) * NK + key] : 0;
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    if (b + u < b1) blk_off[static_cast<int64_t>(b + u) * NK + key] = run;
                    run += v[u];
                }
            }
        }
    }
}

// Stable in-block ranks and per-block key histogram (256 threads, one element each;
// key < 0 marks no element).  Lanes of one wave with equal keys are grouped by ballots
// (one round per distinct key in the wave); earlier waves' counts come from LDS.
// `whist` holds 4 * NK ints.  Returns the element's rank among equal keys of the block.
__device__ inline int block_rank(int key, int NK, int *whist, int *__restrict__ blk_hist_row)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = threadIdx.x; q < 4 * NK; q += blockDim.x) whist[q] = 0;
    __syncthreads();
    unsigned long long todo = __ballot(key >= 0);
    int r = 0;
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const int lk = __shfl(key, leader, 64);
        const unsigned long long m = __ballot(key == lk);
        if (key == lk) r = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == leader) whist[wave * NK + lk] = __popcll(m);
        todo &= ~m;
    }
    __syncthreads();
    if (key >= 0)
        for (int w = 0; w < wave; ++w) r += whist[w * NK + key];
    for (int q = threadIdx.x; q < NK; q += blockDim.x)
        blk_hist_row[q] = whist[q] + whist[NK + q] + whist[2 * NK + q] + whist[3 * NK + q];
    return r;
}

}  // namespace aok

// Working order of point clouds (3D Hilbert curve) and the bounding boxes of
// groups of consecutive points, shared by the EM-ICP and ICP kernels.
#ifndef EMSA_ORDER_H_
#define EMSA_ORDER_H_

#include <cfloat>

#include "emsa_device.h"

namespace emsa {

// Orders LDS writes of some lanes before reads by other lanes of the same
// wave (a wave's LDS instructions execute in order; this keeps the compiler
// from moving accesses across and waits for outstanding LDS operations).
__device__ inline void order_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kOrdThreads = 1024;
constexpr int kOrdWaves = kOrdThreads / kWave;
constexpr int kOrdRounds = 4;                        // keys per thread per tile
constexpr int kOrdKeys = kOrdThreads * kOrdRounds;   // keys per tile
constexpr int kOrdBits = 10;
// Index of cell (x, y, z) (kOrdBits bits each) along the 3D Hilbert curve:
// J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707 (2004):
// the axes are turned into the transposed index, whose bits are then
// interleaved from the most significant level down.
// Bits 0..9 of x moved to bits 0, 3, 6, ..., 27.
__host__ __device__ inline unsigned order_spread3(unsigned x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__host__ __device__ inline unsigned hilbert_key(unsigned x, unsigned y, unsigned z) {
    unsigned v[3] = {x, y, z};
    for (unsigned q = 1u << (kOrdBits - 1); q > 1; q >>= 1) {
        const unsigned p = q - 1;
        for (int i = 0; i < 3; ++i) {
            if (v[i] & q) {
                v[0] ^= p;
            } else {
                const unsigned t = (v[0] ^ v[i]) & p;
                v[0] ^= t;
                v[i] ^= t;
            }
        }
    }
    v[1] ^= v[0];
    v[2] ^= v[1];
    unsigned t = 0;
    for (unsigned q = 1u << (kOrdBits - 1); q > 1; q >>= 1)
        if (v[2] & q) t ^= q - 1;
    // bit b of axis i goes to key bit 3 b + 2 - i (axis 0 most significant)
    return (order_spread3(v[0] ^ t) << 2) | (order_spread3(v[1] ^ t) << 1) | order_spread3(v[2] ^ t);
}

constexpr int kOrdLdsTiles = 4;
// Clouds of at most this many points are sorted with every key in registers
// and each pass's scatter through LDS (8 bytes per point); larger ones keep
// their keys in the global scratch between passes.
constexpr int kOrdLdsMax = kOrdLdsTiles * kOrdKeys;

// Hilbert key of point i of P ([3][n]) on the grid over [lo, lo + 1023 /
// scale]; non-finite points get the largest key.
__device__ inline unsigned order_point_key(const float* __restrict__ P, int n, int i,
                                           const float lo[3], const float scale[3]) {
    unsigned q[3];
    bool fin = true;
    for (int a = 0; a < 3; ++a) {
        const float v = P[(size_t)a * n + i];
        fin = fin && fabsf(v) <= FLT_MAX;
        // NaN from an infinite range times a zero scale clamps to 0
        const float f = fminf(fmaxf((v - lo[a]) * scale[a], 0.f), (float)((1 << kOrdBits) - 1));
        q[a] = (unsigned)f;
    }
    return fin ? hilbert_key(q[0], q[1], q[2]) : 0xFFFFFFFFu;
}

// base[d] = number of keys with a digit below d, from the digit counts tot
// (waves 0..3 scan; barrier inside, base valid after the caller's next one).
__device__ inline void order_digit_scan(const unsigned* tot, unsigned* base, unsigned* wsum) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    unsigned h = 0, x = 0;
    if (tid < 256) {
        h = tot[tid];
        x = h;
        for (int o = 1; o < kWave; o <<= 1) {
            const unsigned y = __shfl_up(x, o, kWave);
            if (lane >= o) x += y;
        }
        if (lane == kWave - 1) wsum[wave] = x;
    }
    __syncthreads();
    if (tid < 256) {
        unsigned off = 0;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        base[tid] = off + x - h;
    }
}

// Rank of a key among this round's earlier lanes with the same digit plus the
// wave's running count of that digit in wcnt (all lanes read it before the
// digit's first lane adds the round's count: one wave's LDS instructions
// execute in order).
__device__ inline unsigned order_rank(unsigned dig, bool valid, unsigned* wcnt_wave) {
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long same = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
        const bool bit = ((dig >> b) & 1u) != 0;
        const unsigned long long bb = __ballot(bit);
        same &= bit ? bb : ~bb;
    }
    const unsigned r = (unsigned)__popcll(same & below);
    const unsigned pre = valid ? wcnt_wave[dig] : 0u;
    order_lds_sync();
    if (valid && r == 0) wcnt_wave[dig] = pre + (unsigned)__popcll(same);
    order_lds_sync();
    return pre + r;
}

// Per digit: wcnt[w][d] becomes the count of digit d in waves below w, tot[d]
// the tile's count (threads 0..255).
__device__ inline void order_wave_offsets(unsigned* wcnt, unsigned* tot) {
    const int tid = threadIdx.x;
    if (tid < 256) {
        unsigned run = 0;
        for (int w = 0; w < kOrdWaves; ++w) {
            const unsigned c = wcnt[w * 256 + tid];
            wcnt[w * 256 + tid] = run;
            run += c;
        }
        tot[tid] = run;
    }
}

// The sort of order_cloud for n <= kOrdLdsMax. Thread (wave w, lane l) holds
// the keys i = kOrdKeys tt + kWave (kOrdRounds w + j) + l (tile tt, round j)
// in registers through all four passes; a pass ranks them tile by tile as
// the global-memory sort does, scatters each tile's (key, index) into kv at
// the ranked positions, then every thread reads its slots back (the previous
// pass's reads of kv end before this pass's first barrier, so the scatter
// may overwrite kv in place).
__device__ inline void order_cloud_lds(const float* __restrict__ P, int n, int* __restrict__ perm,
                                       float* __restrict__ W, int ld, float pad, const float lo[3],
                                       const float scale[3], unsigned* base, unsigned* tot,
                                       unsigned* wsum, unsigned* wcnt, unsigned* hist, uint2* kv) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int ntiles = (n + kOrdKeys - 1) / kOrdKeys;  // workgroup-uniform
    unsigned key[kOrdLdsTiles][kOrdRounds], val[kOrdLdsTiles][kOrdRounds];
#pragma unroll
    for (int tt = 0; tt < kOrdLdsTiles; ++tt)
#pragma unroll
        for (int j = 0; j < kOrdRounds; ++j) {
            const int i = tt * kOrdKeys + kWave * (kOrdRounds * wave + j) + lane;
            key[tt][j] = (i < n) ? order_point_key(P, n, i, lo, scale) : 0u;
            val[tt][j] = (unsigned)i;
        }
    // the four passes' digit counts in one sweep (a count does not depend on
    // the keys' order); wcnt starts at zero and each tile leaves it so
    for (int q = tid; q < 4 * 256; q += kOrdThreads) hist[q] = 0;
    for (int q = tid; q < kOrdWaves * 256; q += kOrdThreads) wcnt[q] = 0;
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < kOrdLdsTiles; ++tt)
#pragma unroll
        for (int j = 0; j < kOrdRounds; ++j)
            if (val[tt][j] < (unsigned)n)
#pragma unroll
                for (int d = 0; d < 4; ++d) atomicAdd(&hist[256 * d + ((key[tt][j] >> (8 * d)) & 255u)], 1u);
    __syncthreads();
    for (int pass = 0; pass < 4; ++pass) {
        const int sh = 8 * pass;
        order_digit_scan(hist + 256 * pass, base, wsum);
#pragma unroll
        for (int tt = 0; tt < kOrdLdsTiles; ++tt) {
            if (tt < ntiles) {
                unsigned rank[kOrdRounds];
#pragma unroll
                for (int j = 0; j < kOrdRounds; ++j)
                    rank[j] = order_rank((key[tt][j] >> sh) & 255u, val[tt][j] < (unsigned)n,
                                         wcnt + wave * 256);
                __syncthreads();
                order_wave_offsets(wcnt, tot);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < kOrdRounds; ++j)
                    if (val[tt][j] < (unsigned)n) {
                        const unsigned dig = (key[tt][j] >> sh) & 255u;
                        kv[base[dig] + wcnt[wave * 256 + dig] + rank[j]] =
                            make_uint2(key[tt][j], val[tt][j]);
                    }
                __syncthreads();
                if (tid < 256) base[tid] += tot[tid];
                for (int q = tid; q < kOrdWaves * 256; q += kOrdThreads) wcnt[q] = 0;
                __syncthreads();
            }
        }
        // slot i now holds the i-th key of this pass's order; slots at or
        // past n keep their own index, which marks them unused
#pragma unroll
        for (int tt = 0; tt < kOrdLdsTiles; ++tt)
#pragma unroll
            for (int j = 0; j < kOrdRounds; ++j) {
                const int i = tt * kOrdKeys + kWave * (kOrdRounds * wave + j) + lane;
                if (i < n) {
                    const uint2 e = kv[i];
                    key[tt][j] = e.x;
                    val[tt][j] = e.y;
                }
            }
        __syncthreads();  // reads of kv before the next pass's scatter
    }
    // perm from the registers; the cloud in working order one axis at a time:
    // the axis staged in LDS (coalesced loads), then gathered from there (a
    // gather from global memory touches a cache line per point); `pad` past n
    float* stage = reinterpret_cast<float*>(kv);  // 2 kOrdLdsMax floats >= n
#pragma unroll
    for (int tt = 0; tt < kOrdLdsTiles; ++tt)
#pragma unroll
        for (int j = 0; j < kOrdRounds; ++j) {
            const int i = tt * kOrdKeys + kWave * (kOrdRounds * wave + j) + lane;
            if (i < n) perm[i] = (int)val[tt][j];
        }
    for (int a = 0; a < 3; ++a) {
        for (int i = tid; i < n; i += kOrdThreads) stage[i] = P[(size_t)a * n + i];
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < kOrdLdsTiles; ++tt)
#pragma unroll
            for (int j = 0; j < kOrdRounds; ++j) {
                const int i = tt * kOrdKeys + kWave * (kOrdRounds * wave + j) + lane;
                if (i < ld) W[(size_t)a * ld + i] = (i < n) ? stage[val[tt][j]] : pad;
            }
        __syncthreads();  // gathers of this axis before the next one's stage
    }
    for (int i = kOrdLdsMax + tid; i < ld; i += kOrdThreads)
        for (int a = 0; a < 3; ++a) W[(size_t)a * ld + i] = pad;
}

// One workgroup of kOrdThreads threads orders the n points of P ([3][n]):
// the Hilbert key of every point (over the bounding box of the finite ones;
// non-finite points get the largest key), then a stable least-significant-
// digit radix sort of (key, index), 8 bits per pass. A pass histograms the
// digits, then scatters tiles of kOrdKeys keys in order; within a tile, equal
// digits are ranked inside a wave with ballots (kOrdRounds rounds of 64
// consecutive keys) and across waves with per-digit prefix counts in LDS.
// Writes perm[i] = the index in P of the point at working position i, and W
// ([3][ld], ld >= n) = the points in that order, `pad` at positions n..ld-1.
// tmp: 4 n unsigned of scratch (used when n > kOrdLdsMax).
__device__ inline void order_cloud(const float* __restrict__ P, int n, int* __restrict__ perm,
                                   unsigned* __restrict__ tmp, float* __restrict__ W, int ld,
                                   float pad) {
    unsigned* kA = tmp;
    unsigned* vA = kA + n;
    unsigned* kB = vA + n;
    unsigned* vB = kB + n;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;

    __shared__ float boxs[kOrdWaves][6];
    __shared__ unsigned base[256], tot[256], wsum[4];
    __shared__ unsigned wcnt[kOrdWaves * 256];
    __shared__ unsigned hist[4 * 256];
    __shared__ uint2 kv[kOrdLdsMax];

    // bounding box of the finite points
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = tid; i < n; i += kOrdThreads)
        for (int a = 0; a < 3; ++a) {
            const float v = P[(size_t)a * n + i];
            if (fabsf(v) <= FLT_MAX) {
                lo[a] = fminf(lo[a], v);
                hi[a] = fmaxf(hi[a], v);
            }
        }
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            boxs[wave][a] = lo[a];
            boxs[wave][3 + a] = hi[a];
        }
    __syncthreads();
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        for (int w = 0; w < kOrdWaves; ++w) {
            lo[a] = fminf(lo[a], boxs[w][a]);
            hi[a] = fmaxf(hi[a], boxs[w][3 + a]);
        }
        scale[a] = (hi[a] > lo[a]) ? (float)((1 << kOrdBits) - 1) / (hi[a] - lo[a]) : 0.f;
    }
    if (n <= kOrdLdsMax) {  // workgroup-uniform
        order_cloud_lds(P, n, perm, W, ld, pad, lo, scale, base, tot, wsum, wcnt, hist, kv);
        return;
    }
    for (int i = tid; i < n; i += kOrdThreads) {
        kA[i] = order_point_key(P, n, i, lo, scale);
        vA[i] = (unsigned)i;
    }
    __syncthreads();

    for (int pass = 0; pass < 4; ++pass) {
        const int sh = 8 * pass;
        const unsigned* ki = (pass & 1) ? kB : kA;
        const unsigned* vi = (pass & 1) ? vB : vA;
        unsigned* ko = (pass & 1) ? kA : kB;
        unsigned* vo = (pass & 1) ? vA : vB;
        if (tid < 256) tot[tid] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += kOrdThreads) atomicAdd(&tot[(ki[i] >> sh) & 255u], 1u);
        __syncthreads();
        order_digit_scan(tot, base, wsum);
        // tiles of kOrdKeys keys; wave w ranks keys t0 + 256 w + 64 j + lane in
        // rounds j = 0..3 (index order), keeping its running count per digit
        // in wcnt (all lanes read it before the digit's first lane adds the
        // round's count: one wave's LDS instructions execute in order)
        for (int q = tid; q < kOrdWaves * 256; q += kOrdThreads) wcnt[q] = 0;
        __syncthreads();
        for (int t0 = 0; t0 < n; t0 += kOrdKeys) {
            unsigned key[kOrdRounds], val[kOrdRounds], rank[kOrdRounds];
#pragma unroll
            for (int j = 0; j < kOrdRounds; ++j) {
                const int i = t0 + kWave * (kOrdRounds * wave + j) + lane;
                const bool valid = i < n;
                key[j] = valid ? ki[i] : 0u;
                val[j] = valid ? vi[i] : 0u;
                rank[j] = order_rank((key[j] >> sh) & 255u, valid, wcnt + wave * 256);
            }
            __syncthreads();
            order_wave_offsets(wcnt, tot);
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kOrdRounds; ++j) {
                const int i = t0 + kWave * (kOrdRounds * wave + j) + lane;
                if (i < n) {
                    const unsigned dig = (key[j] >> sh) & 255u;
                    const unsigned pos = base[dig] + wcnt[wave * 256 + dig] + rank[j];
                    ko[pos] = key[j];
                    vo[pos] = val[j];
                }
            }
            __syncthreads();
            if (tid < 256) base[tid] += tot[tid];
            for (int q = tid; q < kOrdWaves * 256; q += kOrdThreads) wcnt[q] = 0;
            __syncthreads();
        }
        __syncthreads();  // this pass's output is the next pass's input
    }
    for (int i = tid; i < n; i += kOrdThreads) perm[i] = (int)vA[i];

    // the cloud in working order, `pad` past n; 4 points per thread in flight
    for (int i0 = 0; i0 < ld; i0 += 4 * kOrdThreads) {
        float v[4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kOrdThreads + tid;
            const int o = (i < n) ? (int)vA[i] : -1;
#pragma unroll
            for (int a = 0; a < 3; ++a) v[u][a] = (o >= 0) ? P[(size_t)a * n + o] : pad;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kOrdThreads + tid;
            if (i < ld)
#pragma unroll
                for (int a = 0; a < 3; ++a) W[(size_t)a * ld + i] = v[u][a];
        }
    }
}

// Bounding boxes of the groups of G consecutive points of W ([3][ld], n
// valid points): gbox[g] = {lo x, y, z, 0, hi x, y, z, 0}. Call after W is
// complete (a barrier between the writes of W and this).
template <int G>
__device__ inline void group_boxes(const float* __restrict__ W, int ld, int n,
                                   float* __restrict__ gbox) {
    static_assert(G % 4 == 0, "groups of whole float4s");
    for (int g = threadIdx.x; g < ld / G; g += blockDim.x) {
        // empty group (padding only): lo > hi, always skipped; a group with a
        // non-finite point: the whole space, never skipped
        float lo3[3] = {INFINITY, INFINITY, INFINITY}, hi3[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool fin = true;
        const int cn = min(G, n - g * G);
        for (int a = 0; a < 3; ++a) {
            const float4* q = reinterpret_cast<const float4*>(W + (size_t)a * ld + g * G);
#pragma unroll
            for (int h = 0; h < G / 4; ++h) {
                const float4 qh = q[h];
                const float c4[4] = {qh.x, qh.y, qh.z, qh.w};
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (4 * h + u < cn) {
                        fin = fin && fabsf(c4[u]) <= FLT_MAX;
                        lo3[a] = fminf(lo3[a], c4[u]);
                        hi3[a] = fmaxf(hi3[a], c4[u]);
                    }
            }
        }
        float* b = gbox + (size_t)g * 8;
        *reinterpret_cast<float4*>(b) =
            make_float4(fin ? lo3[0] : -INFINITY, fin ? lo3[1] : -INFINITY, fin ? lo3[2] : -INFINITY, 0.f);
        *reinterpret_cast<float4*>(b + 4) =
            make_float4(fin ? hi3[0] : INFINITY, fin ? hi3[1] : INFINITY, fin ? hi3[2] : INFINITY, 0.f);
    }
}

}  // namespace emsa

#endif

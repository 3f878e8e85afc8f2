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
    unsigned key = 0;
    for (int b = kOrdBits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) key = (key << 1) | (((v[i] ^ t) >> b) & 1u);
    return key;
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
// tmp: 4 n unsigned of scratch.
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
    for (int i = tid; i < n; i += kOrdThreads) {
        unsigned q[3];
        bool fin = true;
        for (int a = 0; a < 3; ++a) {
            const float v = P[(size_t)a * n + i];
            fin = fin && fabsf(v) <= FLT_MAX;
            // NaN from an infinite range times a zero scale clamps to 0
            const float f = fminf(fmaxf((v - lo[a]) * scale[a], 0.f), (float)((1 << kOrdBits) - 1));
            q[a] = (unsigned)f;
        }
        kA[i] = fin ? hilbert_key(q[0], q[1], q[2]) : 0xFFFFFFFFu;
        vA[i] = (unsigned)i;
    }
    __syncthreads();

    const unsigned long long below = (1ull << lane) - 1ull;
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
        {  // exclusive scan of the 256 digit counts by waves 0..3
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
                const unsigned dig = (key[j] >> sh) & 255u;
                // lanes of this round holding the same digit
                unsigned long long same = __ballot(valid);
                for (int b = 0; b < 8; ++b) {
                    const bool bit = ((dig >> b) & 1u) != 0;
                    const unsigned long long bb = __ballot(bit);
                    same &= bit ? bb : ~bb;
                }
                const unsigned r = (unsigned)__popcll(same & below);
                const unsigned pre = valid ? wcnt[wave * 256 + dig] : 0u;
                order_lds_sync();
                if (valid && r == 0) wcnt[wave * 256 + dig] = pre + (unsigned)__popcll(same);
                order_lds_sync();
                rank[j] = pre + r;
            }
            __syncthreads();
            if (tid < 256) {  // per digit: offsets of the waves, then the tile total
                unsigned run = 0;
                for (int w = 0; w < kOrdWaves; ++w) {
                    const unsigned c = wcnt[w * 256 + tid];
                    wcnt[w * 256 + tid] = run;
                    run += c;
                }
                tot[tid] = run;
            }
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

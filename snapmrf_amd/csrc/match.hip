// match.hip -- fused complex-GEMM + argmax matcher for gfx950.
//
// S[atom][voxel] = sum_t conj(a_t) s_t is computed as two real GEMMs over the
// interleaved (re, im) index k:
//     Re S = sum_k A[k] X[k],   Im S = sum_k A[k] Y[k],
//     X = (xr0, xi0, xr1, xi1, ...),  Y = (xi0, -xr0, xi1, -xr1, ...)
// on v_mfma_f32_16x16x4f32 with rows = atoms, columns = voxels.  Each lane
// reads 4 consecutive k of its row/column with one ds_read_b128 and uses
// component s in MFMA step s; Y comes from the same 4 floats in registers
// (x1, -x0, x3, -x2), so the signal is staged once.
//
// The score matrix is never written: after each 128-atom sub-tile every lane
// folds its |S|^2 values into a running (score, index) per voxel, first index
// winning ties as in generate_maps (src/functions.cu:1317-1349).  Work-groups
// that share a voxel tile but cover different atom chunks merge through a
// 64-bit atomicMax on (score bits << 32 | ~index), valid because scores >= 0.
//
// B1-map mode matches each voxel only against the slice its B1 map selects
// (merge_maps(b1map), src/functions.cu:1351-1384): voxels are bucketed by
// slice with a counting sort and tiled per slice.  Without a B1 map all NATOMS
// atoms form one slice, which equals per-slice argmax + merge_maps(PD)
// (src/functions.cu:1386-1416): max PD, ties to the lowest slice and index.
#include "kernels.h"
#include "snapmrf.h"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrf {
namespace {

constexpr int kBlock = 256;
constexpr int VT = 128;          // voxels per work-group tile
constexpr int AT = 128;          // atoms per sub-tile
constexpr int KC = 16;           // timepoints per K chunk
constexpr int KF = 2 * KC;       // floats per row per chunk
constexpr int LS = KF + 4;       // LDS row stride (floats), padded for ds_read_b128
constexpr int kAtomChunk = 8192; // atoms per work-group (grid.y splits the slice)

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct Tile {
    int slice;
    int pstart;   // first position in the permuted voxel list
    int count;    // voxels in this tile (<= VT)
    int pad;
};

struct MatchArgs {
    const float2 *atoms;
    const float2 *sig;
    const int32_t *perm;        // permuted voxel ids
    const Tile *tiles;
    const int *ntiles;
    unsigned long long *keys;   // [V]
    size_t natoms;              // atoms per slice
    int T;
};

__device__ __forceinline__ unsigned long long make_key(float score, uint32_t idx)
{
    return ((unsigned long long)__float_as_uint(score) << 32) | (unsigned long long)(~idx);
}

// Two consecutive complex values (t, t+1) of a row, zero outside [0, T) or
// for a null row.  Rows are 16-byte aligned only when T is even.
__device__ __forceinline__ float4 load_pair(const float2 *row, int t, int T, bool even_t)
{
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!row) return v;
    if (t + 1 < T) {
        if (even_t) return *reinterpret_cast<const float4 *>(row + t);
        const float2 x = row[t], y = row[t + 1];
        return make_float4(x.x, x.y, y.x, y.y);
    }
    if (t < T) {
        const float2 x = row[t];
        v = make_float4(x.x, x.y, 0.f, 0.f);
    }
    return v;
}

__global__ __launch_bounds__(kBlock) void match_fp32_kernel(MatchArgs a)
{
    const bool even_t = (a.T & 1) == 0;
    const int tile_id = blockIdx.x;
    if (tile_id >= *a.ntiles) return;
    const Tile tile = a.tiles[tile_id];
    const size_t a_begin = (size_t)blockIdx.y * kAtomChunk;
    if (a_begin >= a.natoms) return;
    const size_t a_end = (a_begin + kAtomChunk < a.natoms) ? a_begin + kAtomChunk : a.natoms;
    const size_t slice_base = (size_t)tile.slice * a.natoms;

    __shared__ __attribute__((aligned(16))) float lds_a[AT * LS];
    __shared__ __attribute__((aligned(16))) float lds_b[VT * LS];
    __shared__ float red_s[2][4][VT];
    __shared__ uint32_t red_i[2][4][VT];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wa = wave & 1;          // atom half of the sub-tile (64 atoms)
    const int wv = wave >> 1;         // voxel half of the tile (64 voxels)
    const int g = lane >> 4;          // k group
    const int r16 = lane & 15;

    // staging map: 8 float4 per row-chunk (KF = 32 floats); 256 threads cover
    // 32 rows per pass, 4 passes for 128 rows.
    const int st_row = tid >> 3;      // 0..31
    const int st_col = (tid & 7) * 4; // float offset within the chunk row

    // voxel rows of this tile (gathered through the permutation)
    const float2 *vrow[4];
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int vr = st_row + 32 * p;
        vrow[p] = (vr < tile.count) ? a.sig + (size_t)a.perm[tile.pstart + vr] * (size_t)a.T
                                    : nullptr;
    }

    float best[4];
    uint32_t bidx[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        best[u] = -1.0f;
        bidx[u] = 0xffffffffu;
    }

    const int nchunks = (a.T + KC - 1) / KC;
    for (size_t sub = a_begin; sub < a_end; sub += AT) {
        const float2 *arow[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const size_t ar = sub + st_row + 32 * p;
            arow[p] = (ar < a_end) ? a.atoms + (slice_base + ar) * (size_t)a.T : nullptr;
        }
        floatx4 acc_re[4][4], acc_im[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                acc_re[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc_im[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
            }

        // global -> registers (zero outside [0, T) and outside the tile); the
        // next chunk's loads are issued before the current chunk's MFMAs
        float4 ga[4], gb[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            ga[p] = load_pair(arow[p], st_col >> 1, a.T, even_t);
            gb[p] = load_pair(vrow[p], st_col >> 1, a.T, even_t);
        }
        for (int c = 0; c < nchunks; c++) {
            __syncthreads();
#pragma unroll
            for (int p = 0; p < 4; p++) {
                *reinterpret_cast<float4 *>(&lds_a[(st_row + 32 * p) * LS + st_col]) = ga[p];
                *reinterpret_cast<float4 *>(&lds_b[(st_row + 32 * p) * LS + st_col]) = gb[p];
            }
            __syncthreads();
            if (c + 1 < nchunks) {
                const int tc = (c + 1) * KC + (st_col >> 1);
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    ga[p] = load_pair(arow[p], tc, a.T, even_t);
                    gb[p] = load_pair(vrow[p], tc, a.T, even_t);
                }
            }
#pragma unroll
            for (int q = 0; q < KF / 16; q++) {
                floatx4 fa[4], fx[4], fy[4];
#pragma unroll
                for (int i = 0; i < 4; i++)
                    fa[i] = *reinterpret_cast<const floatx4 *>(
                        &lds_a[(wa * 64 + i * 16 + r16) * LS + q * 16 + g * 4]);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    fx[v] = *reinterpret_cast<const floatx4 *>(
                        &lds_b[(wv * 64 + v * 16 + r16) * LS + q * 16 + g * 4]);
                    fy[v] = floatx4{fx[v][1], -fx[v][0], fx[v][3], -fx[v][2]};
                }
#pragma unroll
                for (int s = 0; s < 4; s++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            acc_re[i][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                fa[i][s], fx[v][s], acc_re[i][v], 0, 0, 0);
                            acc_im[i][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                fa[i][s], fy[v][s], acc_im[i][v], 0, 0, 0);
                        }
            }
        }
        // epilogue: D[row = 4g + r][col = r16] of each 16x16 tile
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const size_t atom = sub + wa * 64 + i * 16 + 4 * g + r;
                if (atom < a_end) {
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const float re = acc_re[i][v][r], im = acc_im[i][v][r];
                        const float sc = re * re + im * im;
                        if (sc > best[v]) {
                            best[v] = sc;
                            bidx[v] = (uint32_t)(slice_base + atom);
                        }
                    }
                }
            }
    }

    // reduce the 8 candidates per voxel (4 k-groups x 2 atom halves)
#pragma unroll
    for (int v = 0; v < 4; v++) {
        red_s[wa][g][wv * 64 + v * 16 + r16] = best[v];
        red_i[wa][g][wv * 64 + v * 16 + r16] = bidx[v];
    }
    __syncthreads();
    if (tid < VT && tid < tile.count) {
        float bs = -1.0f;
        uint32_t bi = 0xffffffffu;
        for (int h = 0; h < 2; h++)
            for (int q = 0; q < 4; q++) {
                const float s2 = red_s[h][q][tid];
                const uint32_t i2 = red_i[h][q][tid];
                if (s2 > bs || (s2 == bs && i2 < bi)) {
                    bs = s2;
                    bi = i2;
                }
            }
        if (bs >= 0.0f) {
            const int vox = a.perm[tile.pstart + tid];
            atomicMax(&a.keys[vox], make_key(bs, bi));
        }
    }
}

// --- voxel bucketing by B1 slice ------------------------------------------

__global__ void b1_select_kernel(const float *b1map, const float *params, size_t natoms,
                                 int n_slices, size_t V, int32_t *vox_slice)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    // merge_maps(b1map): strict '<' from 10, first slice on ties, float math
    const float b = b1map[v];
    int m = 0;
    float bd = 10.0f;
    for (int k = 0; k < n_slices; k++) {
        const float d = fabsf(b - params[(size_t)k * natoms * 4 + 3]);
        if (d < bd) {
            bd = d;
            m = k;
        }
    }
    vox_slice[v] = m;
}

// Slice histogram: per-block LDS counts, one global atomic per (block, slice).
constexpr int kMaxLdsSlices = 1024;

__global__ void count_kernel(const int32_t *vox_slice, size_t V, int n_slices, int *counts)
{
    __shared__ int h[kMaxLdsSlices];
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool lds = n_slices <= kMaxLdsSlices;
    if (lds)
        for (int k = threadIdx.x; k < n_slices; k += blockDim.x) h[k] = 0;
    __syncthreads();
    if (v < V) {
        const int k = vox_slice ? vox_slice[v] : 0;
        if (lds) atomicAdd(&h[k], 1);
        else atomicAdd(&counts[k], 1);
    }
    __syncthreads();
    if (lds)
        for (int k = threadIdx.x; k < n_slices; k += blockDim.x)
            if (h[k]) atomicAdd(&counts[k], h[k]);
}

// single work-item: exclusive scan of the slice counts and the tile list
__global__ void tiles_kernel(const int *counts, int n_slices, int *offsets, Tile *tiles,
                             int *ntiles)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int off = 0, nt = 0;
    for (int k = 0; k < n_slices; k++) {
        offsets[k] = off;
        const int c = counts[k];
        for (int p = 0; p < c; p += VT) {
            Tile t;
            t.slice = k;
            t.pstart = off + p;
            t.count = (c - p < VT) ? c - p : VT;
            t.pad = 0;
            tiles[nt++] = t;
        }
        off += c;
    }
    *ntiles = nt;
}

// Voxel permutation grouped by slice: ranks within the block from LDS
// atomics, one global reservation per (block, slice).
__global__ void scatter_kernel(const int32_t *vox_slice, size_t V, int n_slices,
                               const int *offsets, int *fill, int32_t *perm)
{
    __shared__ int h[kMaxLdsSlices];
    __shared__ int base[kMaxLdsSlices];
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = (v < V && vox_slice) ? vox_slice[v] : 0;
    if (n_slices > kMaxLdsSlices) {
        if (v < V) perm[offsets[k] + atomicAdd(&fill[k], 1)] = (int32_t)v;
        return;
    }
    for (int q = threadIdx.x; q < n_slices; q += blockDim.x) h[q] = 0;
    __syncthreads();
    const int rank = (v < V) ? atomicAdd(&h[k], 1) : 0;
    __syncthreads();
    for (int q = threadIdx.x; q < n_slices; q += blockDim.x)
        if (h[q]) base[q] = offsets[q] + atomicAdd(&fill[q], h[q]);
    __syncthreads();
    if (v < V) perm[base[k] + rank] = (int32_t)v;
}

__global__ void maps_kernel(const unsigned long long *keys, const float *params, size_t V,
                            float *maps, int32_t *idx)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const unsigned long long k = keys[v];
    uint32_t i = ~(uint32_t)(k & 0xffffffffull);
    float sc = __uint_as_float((uint32_t)(k >> 32));
    if (k == 0ull) {   // no candidate (empty slice): reference leaves index 0
        i = 0;
        sc = 0.0f;
    }
    const float *p = params + (size_t)i * 4;
    maps[v * 5 + 0] = p[0];
    maps[v * 5 + 1] = p[1];
    maps[v * 5 + 2] = p[2];
    maps[v * 5 + 3] = p[3];
    maps[v * 5 + 4] = sqrtf(sc);
    if (idx) idx[v] = (int32_t)i;
}

__global__ void zero_u64_kernel(unsigned long long *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void zero_i32_kernel(int *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------------
// BF16 matcher: v_mfma_f32_16x16x32_bf16 over BF16 copies of the dictionary
// and the signals, FP32 accumulation, a per-voxel top-8 candidate list and an
// FP32 rescore of the candidates from the complex64 dictionary.  SPLIT adds
// the hi*lo and lo*hi products of the split-BF16 representation
// x = hi + lo (hi = bf16(x), lo = bf16(x - hi)).
// ------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kTopK = 8;
constexpr int kMaxChunks = 32;

struct LowpArgs {
    const uint32_t *ahi, *alo;   // [NATOMS][T] bf16x2 (re low half, im high half)
    const uint32_t *shi, *slo;   // [V][T] bf16x2, original voxel order
    const int32_t *perm;
    const Tile *tiles;
    const int *ntiles;
    float *part_s;               // [chunks][V][kTopK] candidate scores (-1 = empty)
    uint32_t *part_i;            // [chunks][V][kTopK] global atom indices
    size_t natoms;               // atoms per slice
    size_t chunk;                // atoms per work-group
    size_t V;
    int T;
};

// (re, im) -> (im, -re) on a bf16x2 word: the Y operand of Im S.
__device__ __forceinline__ uint32_t swap_neg(uint32_t w)
{
    return ((w >> 16) | (w << 16)) ^ 0x80000000u;
}

// Four consecutive bf16x2 words (timepoints t..t+3) of a row, zero outside
// [0, T) or for a null row; rows are 16-byte aligned when T % 4 == 0.
__device__ __forceinline__ u32x4 load4(const uint32_t *row, int t, int T, bool al4)
{
    u32x4 v = u32x4{0u, 0u, 0u, 0u};
    if (!row) return v;
    if (al4 && t + 3 < T) return *reinterpret_cast<const u32x4 *>(row + t);
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (t + u < T) v[u] = row[t + u];
    return v;
}

__device__ __forceinline__ floatx4 mfma_bf16(u32x4 a, u32x4 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

template <bool SPLIT, int KC>
__global__ __launch_bounds__(kBlock) void match_lowp_kernel(LowpArgs a)
{
    constexpr int LSB = KC + 4;             // dwords per LDS row (padded)
    constexpr int SEGS = KC / 4;            // 16-byte segments per row chunk
    constexpr int RPP = kBlock / SEGS;      // rows staged per pass
    constexpr int PASSES = AT / RPP;
    constexpr int KSTEPS = KC / 16;         // 16 complex = 32 bf16 per MFMA
    static_assert(AT == VT, "square work-group tile");

    const int tile_id = blockIdx.x;
    if (tile_id >= *a.ntiles) return;
    const Tile tile = a.tiles[tile_id];
    const size_t a_begin = (size_t)blockIdx.y * a.chunk;
    if (a_begin >= a.natoms) return;
    const size_t a_end = (a_begin + a.chunk < a.natoms) ? a_begin + a.chunk : a.natoms;
    const size_t slice_base = (size_t)tile.slice * a.natoms;
    const bool al4 = (a.T & 3) == 0;

    __shared__ __attribute__((aligned(16))) uint32_t lA[AT * LSB];
    __shared__ __attribute__((aligned(16))) uint32_t lX[VT * LSB];
    __shared__ __attribute__((aligned(16))) uint32_t lY[VT * LSB];
    __shared__ __attribute__((aligned(16))) uint32_t lAl[SPLIT ? AT * LSB : 4];
    __shared__ __attribute__((aligned(16))) uint32_t lXl[SPLIT ? VT * LSB : 4];
    __shared__ __attribute__((aligned(16))) uint32_t lYl[SPLIT ? VT * LSB : 4];
    __shared__ float cs[2 * 4 * VT * kTopK];
    __shared__ uint32_t ci[2 * 4 * VT * kTopK];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wa = wave & 1;
    const int wv = wave >> 1;
    const int g = lane >> 4;
    const int r16 = lane & 15;
    const int st_row = tid / SEGS;
    const int st_t = (tid % SEGS) * 4;      // timepoint offset within the chunk

    for (int q = tid; q < 2 * 4 * VT * kTopK; q += kBlock) {
        cs[q] = -1.0f;
        ci[q] = 0xffffffffu;
    }
    float th[4] = {-1.0f, -1.0f, -1.0f, -1.0f};   // 8th best of this lane's lists

    const uint32_t *vh[PASSES], *vl[PASSES];
#pragma unroll
    for (int p = 0; p < PASSES; p++) {
        const int vr = st_row + RPP * p;
        const size_t off = (vr < tile.count) ? (size_t)a.perm[tile.pstart + vr] * (size_t)a.T : 0;
        vh[p] = (vr < tile.count) ? a.shi + off : nullptr;
        vl[p] = (SPLIT && vr < tile.count) ? a.slo + off : nullptr;
    }
    const int nchunks = (a.T + KC - 1) / KC;

    for (size_t sub = a_begin; sub < a_end; sub += AT) {
        const uint32_t *ah[PASSES], *al[PASSES];
#pragma unroll
        for (int p = 0; p < PASSES; p++) {
            const size_t ar = sub + st_row + RPP * p;
            const size_t off = (slice_base + ar) * (size_t)a.T;
            ah[p] = (ar < a_end) ? a.ahi + off : nullptr;
            al[p] = (SPLIT && ar < a_end) ? a.alo + off : nullptr;
        }
        floatx4 acc_re[4][4], acc_im[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                acc_re[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc_im[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
            }
        u32x4 ga[PASSES], gx[PASSES], gal[PASSES], gxl[PASSES];
#pragma unroll
        for (int p = 0; p < PASSES; p++) {
            ga[p] = load4(ah[p], st_t, a.T, al4);
            gx[p] = load4(vh[p], st_t, a.T, al4);
            if (SPLIT) {
                gal[p] = load4(al[p], st_t, a.T, al4);
                gxl[p] = load4(vl[p], st_t, a.T, al4);
            }
        }
        for (int c = 0; c < nchunks; c++) {
            __syncthreads();
#pragma unroll
            for (int p = 0; p < PASSES; p++) {
                const int o = (st_row + RPP * p) * LSB + st_t;
                *reinterpret_cast<u32x4 *>(&lA[o]) = ga[p];
                *reinterpret_cast<u32x4 *>(&lX[o]) = gx[p];
                *reinterpret_cast<u32x4 *>(&lY[o]) =
                    u32x4{swap_neg(gx[p][0]), swap_neg(gx[p][1]), swap_neg(gx[p][2]),
                          swap_neg(gx[p][3])};
                if (SPLIT) {
                    *reinterpret_cast<u32x4 *>(&lAl[o]) = gal[p];
                    *reinterpret_cast<u32x4 *>(&lXl[o]) = gxl[p];
                    *reinterpret_cast<u32x4 *>(&lYl[o]) =
                        u32x4{swap_neg(gxl[p][0]), swap_neg(gxl[p][1]), swap_neg(gxl[p][2]),
                              swap_neg(gxl[p][3])};
                }
            }
            __syncthreads();
            if (c + 1 < nchunks) {
                const int tc = (c + 1) * KC + st_t;
#pragma unroll
                for (int p = 0; p < PASSES; p++) {
                    ga[p] = load4(ah[p], tc, a.T, al4);
                    gx[p] = load4(vh[p], tc, a.T, al4);
                    if (SPLIT) {
                        gal[p] = load4(al[p], tc, a.T, al4);
                        gxl[p] = load4(vl[p], tc, a.T, al4);
                    }
                }
            }
#pragma unroll
            for (int ks = 0; ks < KSTEPS; ks++) {
                const int ko = ks * 16 + g * 4;
                u32x4 fa[4], fx[4], fy[4];
#pragma unroll
                for (int i = 0; i < 4; i++)
                    fa[i] = *reinterpret_cast<const u32x4 *>(&lA[(wa * 64 + i * 16 + r16) * LSB + ko]);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int o = (wv * 64 + v * 16 + r16) * LSB + ko;
                    fx[v] = *reinterpret_cast<const u32x4 *>(&lX[o]);
                    fy[v] = *reinterpret_cast<const u32x4 *>(&lY[o]);
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        acc_re[i][v] = mfma_bf16(fa[i], fx[v], acc_re[i][v]);
                        acc_im[i][v] = mfma_bf16(fa[i], fy[v], acc_im[i][v]);
                    }
                if (SPLIT) {
                    u32x4 fal[4], fxl[4], fyl[4];
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        fal[i] = *reinterpret_cast<const u32x4 *>(
                            &lAl[(wa * 64 + i * 16 + r16) * LSB + ko]);
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int o = (wv * 64 + v * 16 + r16) * LSB + ko;
                        fxl[v] = *reinterpret_cast<const u32x4 *>(&lXl[o]);
                        fyl[v] = *reinterpret_cast<const u32x4 *>(&lYl[o]);
                    }
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            acc_re[i][v] = mfma_bf16(fa[i], fxl[v], acc_re[i][v]);
                            acc_im[i][v] = mfma_bf16(fa[i], fyl[v], acc_im[i][v]);
                            acc_re[i][v] = mfma_bf16(fal[i], fx[v], acc_re[i][v]);
                            acc_im[i][v] = mfma_bf16(fal[i], fy[v], acc_im[i][v]);
                        }
                }
            }
        }
        // epilogue: fold |S|^2 of this sub-tile into the lane's top-8 lists
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int voxel = wv * 64 + v * 16 + r16;
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const size_t atom = sub + wa * 64 + i * 16 + 4 * g + r;
                    const float re = acc_re[i][v][r], im = acc_im[i][v][r];
                    const float sc = re * re + im * im;
                    if (atom < a_end && sc > th[v]) {
                        float *ls = &cs[((wa * 4 + g) * VT + voxel) * kTopK];
                        uint32_t *li = &ci[((wa * 4 + g) * VT + voxel) * kTopK];
                        int pos = kTopK - 1;
                        while (pos > 0 && ls[pos - 1] < sc) {
                            ls[pos] = ls[pos - 1];
                            li[pos] = li[pos - 1];
                            pos--;
                        }
                        ls[pos] = sc;
                        li[pos] = (uint32_t)(slice_base + atom);
                        th[v] = ls[kTopK - 1];
                    }
                }
        }
    }
    __syncthreads();
    // merge the 8 sorted lists of each voxel into its top-8 for this chunk
    if (tid < tile.count) {
        int head[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const size_t vox = (size_t)a.perm[tile.pstart + tid];
        float *ps = a.part_s + ((size_t)blockIdx.y * a.V + vox) * kTopK;
        uint32_t *pi = a.part_i + ((size_t)blockIdx.y * a.V + vox) * kTopK;
        for (int e = 0; e < kTopK; e++) {
            int bh = -1;
            float bs = -1.0f;
            uint32_t bi = 0xffffffffu;
            for (int h = 0; h < 8; h++) {
                if (head[h] >= kTopK) continue;
                const int q = (h * VT + tid) * kTopK + head[h];
                const float s2 = cs[q];
                const uint32_t i2 = ci[q];
                if (s2 > bs || (s2 == bs && s2 >= 0.0f && i2 < bi)) {
                    bs = s2;
                    bi = i2;
                    bh = h;
                }
            }
            if (bh >= 0 && bs >= 0.0f) head[bh]++;
            else {
                bs = -1.0f;
                bi = 0xffffffffu;
            }
            ps[e] = bs;
            pi[e] = bi;
        }
    }
}

__device__ __forceinline__ bool cand_better(float s1, uint32_t i1, float s2, uint32_t i2)
{
    return s1 > s2 || (s1 == s2 && i1 < i2);
}

// Per voxel (one wave): merge the chunk candidate lists into the top-8,
// rescore them in FP32 from the complex64 dictionary and keep the best
// (first index on ties).
__global__ __launch_bounds__(kBlock) void rescore_kernel(const float *part_s,
                                                         const uint32_t *part_i, int nchunks,
                                                         const float2 *atoms, const float2 *sig,
                                                         const float *params, size_t V, int T,
                                                         float *maps, int32_t *idx)
{
    const size_t vox = (size_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (vox >= V) return;
    const int n = nchunks * kTopK;   // <= kMaxChunks * kTopK = 256
    float s[4];
    uint32_t id[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int e = lane + 64 * q;
        s[q] = -1.0f;
        id[q] = 0xffffffffu;
        if (e < n) {
            const int ch = e / kTopK, k = e % kTopK;
            const size_t o = ((size_t)ch * V + vox) * kTopK + k;
            s[q] = part_s[o];
            id[q] = part_i[o];
        }
    }
    uint32_t cand[kTopK];
    int nc = 0;
    for (int r = 0; r < kTopK; r++) {
        float bs = -1.0f;
        uint32_t bi = 0xffffffffu;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (s[q] >= 0.0f && cand_better(s[q], id[q], bs, bi)) {
                bs = s[q];
                bi = id[q];
            }
        for (int off = 32; off >= 1; off >>= 1) {
            const float os = __shfl_xor(bs, off);
            const uint32_t oi = (uint32_t)__shfl_xor((int)bi, off);
            if (cand_better(os, oi, bs, bi)) {
                bs = os;
                bi = oi;
            }
        }
        if (bs < 0.0f) break;
        cand[nc++] = bi;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (id[q] == bi) s[q] = -1.0f;
    }
    // FP32 rescore
    float re[kTopK], im[kTopK];
#pragma unroll
    for (int c = 0; c < kTopK; c++) re[c] = im[c] = 0.0f;
    const float2 *x = sig + vox * (size_t)T;
    for (int t = lane; t < T; t += 64) {
        const float2 xv = x[t];
#pragma unroll
        for (int c = 0; c < kTopK; c++)
            if (c < nc) {
                const float2 av = atoms[(size_t)cand[c] * T + t];
                re[c] += av.x * xv.x + av.y * xv.y;
                im[c] += av.x * xv.y - av.y * xv.x;
            }
    }
#pragma unroll
    for (int c = 0; c < kTopK; c++)
        for (int off = 32; off >= 1; off >>= 1) {
            re[c] += __shfl_xor(re[c], off);
            im[c] += __shfl_xor(im[c], off);
        }
    float bs = -1.0f;
    uint32_t bi = 0u;
#pragma unroll
    for (int c = 0; c < kTopK; c++)
        if (c < nc) {
            const float sc = re[c] * re[c] + im[c] * im[c];
            if (cand_better(sc, cand[c], bs, bi)) {
                bs = sc;
                bi = cand[c];
            }
        }
    if (nc == 0) {
        bs = 0.0f;
        bi = 0u;
    }
    if (lane == 0) {
        const float *p = params + (size_t)bi * 4;
        maps[vox * 5 + 0] = p[0];
        maps[vox * 5 + 1] = p[1];
        maps[vox * 5 + 2] = p[2];
        maps[vox * 5 + 3] = p[3];
        maps[vox * 5 + 4] = sqrtf(bs);
        if (idx) idx[vox] = (int32_t)bi;
    }
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Scratch {
    unsigned long long *keys;
    int32_t *perm;
    int *counts, *offsets, *fill, *ntiles;
    Tile *tiles;
    size_t max_tiles;
    // low-precision path
    uint32_t *shi, *slo;
    float *part_s;
    uint32_t *part_i;
    size_t chunk, nchunks;
    size_t bytes;
};

// Work-group atom range of the low-precision path: enough chunks to give
// about 2048 work-groups, at most kMaxChunks, a multiple of the sub-tile.
void lowp_chunks(const MatchLaunch &l, size_t nat, size_t max_tiles, size_t *chunk,
                 size_t *nchunks)
{
    size_t want = (2048 + max_tiles - 1) / max_tiles;
    if (want < 1) want = 1;
    if (want > (size_t)kMaxChunks) want = kMaxChunks;
    size_t c = (nat + want - 1) / want;
    c = (c + AT - 1) / AT * AT;
    *chunk = c;
    *nchunks = (nat + c - 1) / c;
}

Scratch carve(void *base, const MatchLaunch &l)
{
    Scratch s{};
    const int ns = l.vox_slice ? l.n_slices : 1;
    const size_t nat = l.vox_slice ? l.natoms : l.natoms * (size_t)l.n_slices;
    s.max_tiles = (l.V + VT - 1) / VT + (size_t)ns;
    uint8_t *p = (uint8_t *)base;
    s.keys = (unsigned long long *)p;
    p += align256(l.V * 8);
    s.perm = (int32_t *)p;
    p += align256(l.V * 4);
    s.counts = (int *)p;
    p += align256((size_t)ns * 4);
    s.offsets = (int *)p;
    p += align256((size_t)ns * 4);
    s.fill = (int *)p;
    p += align256((size_t)ns * 4);
    s.ntiles = (int *)p;
    p += 256;
    s.tiles = (Tile *)p;
    p += align256(s.max_tiles * sizeof(Tile));
    if (l.mode != MRF_MATCH_FP32) {
        lowp_chunks(l, nat, s.max_tiles, &s.chunk, &s.nchunks);
        s.shi = (uint32_t *)p;
        p += align256(l.V * (size_t)l.T * 4);
        if (l.mode == MRF_MATCH_SPLIT3) {
            s.slo = (uint32_t *)p;
            p += align256(l.V * (size_t)l.T * 4);
        }
        s.part_s = (float *)p;
        p += align256(s.nchunks * l.V * kTopK * 4);
        s.part_i = (uint32_t *)p;
        p += align256(s.nchunks * l.V * kTopK * 4);
    }
    s.bytes = (size_t)(p - (uint8_t *)base);
    return s;
}

}  // namespace

size_t match_scratch_bytes(const MatchLaunch &l)
{
    return carve(nullptr, l).bytes;
}

int launch_b1_select(const float *b1map, const float *params, size_t natoms, int n_slices,
                     size_t V, int32_t *vox_slice, hipStream_t s)
{
    if (V == 0) return 0;
    b1_select_kernel<<<dim3((unsigned)((V + 255) / 256)), 256, 0, s>>>(b1map, params, natoms,
                                                                        n_slices, V, vox_slice);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "b1 select: %s", hipGetErrorString(e));
}

int launch_match(const MatchLaunch &l, void *scratch, hipStream_t s)
{
    if (l.V == 0) return 0;
    if (l.mode != MRF_MATCH_FP32 && l.mode != MRF_MATCH_BF16_TOP8 && l.mode != MRF_MATCH_SPLIT3)
        return mrf_fail(MRF_E_UNSUPPORTED, "matching mode %d not available", l.mode);
    if (l.mode != MRF_MATCH_FP32 && (!l.lowp_hi || (l.mode == MRF_MATCH_SPLIT3 && !l.lowp_lo)))
        return mrf_fail(MRF_E_ARG, "low-precision matching needs the bf16 dictionary");
    const int ns = l.vox_slice ? l.n_slices : 1;
    const size_t nat = l.vox_slice ? l.natoms : l.natoms * (size_t)l.n_slices;
    if ((size_t)ns * nat > 0xfffffffeULL)
        return mrf_fail(MRF_E_ARG, "dictionary larger than 2^32-1 atoms");
    if (l.V > 0x7fffffffULL) return mrf_fail(MRF_E_ARG, "too many voxels");
    Scratch sc = carve(scratch, l);
    const unsigned vb = (unsigned)((l.V + 255) / 256);
    zero_u64_kernel<<<vb, 256, 0, s>>>(sc.keys, l.V);
    const unsigned nb = (unsigned)((ns + 255) / 256);
    zero_i32_kernel<<<nb, 256, 0, s>>>(sc.counts, (size_t)ns);
    zero_i32_kernel<<<nb, 256, 0, s>>>(sc.fill, (size_t)ns);
    count_kernel<<<vb, 256, 0, s>>>(l.vox_slice, l.V, ns, sc.counts);
    tiles_kernel<<<1, 1, 0, s>>>(sc.counts, ns, sc.offsets, sc.tiles, sc.ntiles);
    scatter_kernel<<<vb, 256, 0, s>>>(l.vox_slice, l.V, ns, sc.offsets, sc.fill, sc.perm);

    if (l.mode != MRF_MATCH_FP32) {
        if (int rc = launch_lowp_prepare(l.sig, l.V, l.T, sc.shi,
                                         l.mode == MRF_MATCH_SPLIT3 ? sc.slo : nullptr, s))
            return rc;
        LowpArgs b{};
        b.ahi = l.lowp_hi;
        b.alo = l.lowp_lo;
        b.shi = sc.shi;
        b.slo = sc.slo;
        b.perm = sc.perm;
        b.tiles = sc.tiles;
        b.ntiles = sc.ntiles;
        b.part_s = sc.part_s;
        b.part_i = sc.part_i;
        b.natoms = nat;
        b.chunk = sc.chunk;
        b.V = l.V;
        b.T = l.T;
        dim3 grid((unsigned)sc.max_tiles, (unsigned)sc.nchunks);
        if (l.mode == MRF_MATCH_SPLIT3) match_lowp_kernel<true, 16><<<grid, kBlock, 0, s>>>(b);
        else match_lowp_kernel<false, 32><<<grid, kBlock, 0, s>>>(b);
        rescore_kernel<<<dim3((unsigned)((l.V + 3) / 4)), kBlock, 0, s>>>(
            sc.part_s, sc.part_i, (int)sc.nchunks, l.atoms, l.sig, l.params, l.V, l.T, l.maps,
            l.idx);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "match launch: %s", hipGetErrorString(e));
    }

    MatchArgs a{};
    a.atoms = l.atoms;
    a.sig = l.sig;
    a.perm = sc.perm;
    a.tiles = sc.tiles;
    a.ntiles = sc.ntiles;
    a.keys = sc.keys;
    a.natoms = nat;
    a.T = l.T;
    const size_t chunks = (nat + kAtomChunk - 1) / kAtomChunk;
    dim3 grid((unsigned)sc.max_tiles, (unsigned)chunks);
    match_fp32_kernel<<<grid, kBlock, 0, s>>>(a);
    maps_kernel<<<vb, 256, 0, s>>>(sc.keys, l.params, l.V, l.maps, l.idx);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "match launch: %s", hipGetErrorString(e));
}

}  // namespace mrf

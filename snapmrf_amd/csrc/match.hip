// match.hip -- fused complex-GEMM + argmax matcher for gfx950.
//
// S[atom][voxel] = sum_t conj(a_t) s_t.  The FP32 kernel forms it from three
// real GEMMs over t (Gauss's three-multiplication complex product):
//     P1 = sum ar xr,   P2 = sum ai xi,   P3 = sum (ar - ai)(xr + xi),
//     Re S = P1 + P2,   Im S = P3 - P1 + P2
// on v_mfma_f32_16x16x4f32 with rows = atoms, columns = voxels: 6 MFMAs per
// 8 timepoints where the direct Re/Im split needs 8.  Each lane reads two
// (re, im) pairs of its row/column with one ds_read_b128; the sums and
// differences are formed in registers, so atoms and signals are staged once.
// The cancellation in Im S costs about one extra ulp of the summed
// magnitudes, far below the score gaps the argmax has to resolve.
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
#include <functional>
#include <type_traits>

namespace mrf {
namespace {

constexpr int kBlock = 256;
constexpr int VT = 128;          // voxels per work-group tile
constexpr int AT = 128;          // atoms per sub-tile
constexpr int KC = 16;           // timepoints per K chunk
constexpr int KF = 2 * KC;       // floats per row per chunk
constexpr int LS = KF + 4;       // LDS row stride (floats), padded for ds_read_b128
constexpr int kAtomChunk = 8192; // max atoms per work-group (grid.y splits the slice)

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
    size_t natoms;              // atoms per slice in this launch
    size_t chunk;               // atoms per work-group (multiple of AT)
    size_t row_stride;          // atom rows per slice in `atoms`
    size_t gidx_stride;         // global atom index step per slice
    size_t gidx_base;           // global index of atom 0 of slice 0
    size_t sig_bytes;           // bytes of `sig` (V T complex values)
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

// Eight waves: 4 atom groups of 32 x 2 voxel halves of 64 (two waves per
// SIMD; 2 % faster than four waves of 64 x 64 at one per SIMD).
constexpr int kFp32AtomGroups = 4;
constexpr int kFp32Threads = 128 * kFp32AtomGroups;
constexpr uint32_t kRsrcWord3 = 0x00020000u;   // raw 32-bit buffer, gfx9 family
// a buffer offset past every resource used here (resources stay at or below
// it, and offset + 16 does not wrap)
constexpr uint32_t kOob = 0xFFFFFF00u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, size_t bytes)
{
    const uint32_t n = bytes > 0xffffffffu ? 0xffffffffu : (uint32_t)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)n,
                                             (int)kRsrcWord3);
}

typedef __attribute__((address_space(3))) void lds_void;

// DMA = true (even T, signals below 4 GB): the chunks go from global memory
// straight into LDS (buffer_load ... lds), no registers or LDS stores on the
// way.  A wave's copy of 64 x 16 bytes lands as one 1 KB block, lane l at
// byte 16 l, so the LDS holds a chunk fragment-major: block (operand, row
// block rb of 16 rows, half q of the chunk) has lane 16 g + r holding row
// 16 rb + r, timepoints 8 q + 2 g and + 1 -- the 16 bytes that lane (g, r)
// of an MFMA fragment reads, so fragment reads are ds_read_b128 at 16 l.
// Rows outside the sub-tile or tile and timepoints past T get an offset past
// the resource, which the range check turns into zeros.
// DMA = false: the chunks are staged through registers into padded rows.
template <bool DMA>
__global__ __launch_bounds__(kFp32Threads) void match_fp32_kernel(MatchArgs a)
{
    constexpr int NWA = kFp32AtomGroups;
    constexpr int NT = kFp32Threads;    // NWA atom groups x 2 voxel halves of waves
    constexpr int NIB = AT / NWA / 16;  // 16-atom MFMA blocks per wave
    constexpr int RP = NT / 8;          // staged rows per pass
    constexpr int NP = AT / RP;         // passes (atoms and voxels: 128 rows each)
    constexpr int NRB = AT / 16;        // 16-row blocks per operand (AT == VT)
    constexpr int NBW = 2 * NRB * 2 / (NT / 64);   // DMA blocks per wave and chunk
    static_assert(AT == VT, "one block layout for both operands");
    const bool even_t = (a.T & 1) == 0;
    const int tile_id = blockIdx.x;
    if (tile_id >= *a.ntiles) return;
    const Tile tile = a.tiles[tile_id];
    const size_t a_begin = (size_t)blockIdx.y * a.chunk;
    if (a_begin >= a.natoms) return;
    const size_t a_end = (a_begin + a.chunk < a.natoms) ? a_begin + a.chunk : a.natoms;
    const size_t row_base = (size_t)tile.slice * a.row_stride;
    const size_t idx_base = a.gidx_base + (size_t)tile.slice * a.gidx_stride;

    // two LDS buffers: chunk c + 1 is stored while chunk c is multiplied
    __shared__ __attribute__((aligned(16))) float lds_a[DMA ? 1 : 2][DMA ? 4 : AT * LS];
    __shared__ __attribute__((aligned(16))) float lds_b[DMA ? 1 : 2][DMA ? 4 : VT * LS];
    // DMA: two buffers (separate arrays, so the compiler sees that copies
    // into one do not touch reads of the other), [operand][row block][half][lane]
    __shared__ floatx4 lds_d0[DMA ? 2 * NRB * 2 * 64 : 1];
    __shared__ floatx4 lds_d1[DMA ? 2 * NRB * 2 * 64 : 1];
    __shared__ float red_s[NWA][4][VT];
    __shared__ uint32_t red_i[NWA][4][VT];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: DMA operands
    const int wa = wave % NWA;        // atom group of the sub-tile (AT / NWA atoms)
    const int wv = wave / NWA;        // voxel half of the tile (64 voxels)
    const int g = lane >> 4;          // k group
    const int r16 = lane & 15;

    // staging map: 8 float4 per row-chunk (KF = 32 floats); NT threads cover
    // RP rows per pass, NP passes for 128 rows.
    const int st_row = tid >> 3;      // 0..RP-1
    const int st_col = (tid & 7) * 4; // float offset within the chunk row

    // voxel rows of this tile (gathered through the permutation)
    const float2 *vrow[NP];
    // DMA: the wave's blocks w NBW .. w NBW + NBW - 1 (block b: operand
    // b / (2 NRB), row block (b / 2) % NRB, half b % 2); the lane's row byte
    // offset in its operand's resource, kOob for a row outside
    uint32_t drow[NBW];
    const __amdgpu_buffer_rsrc_t rsig = make_rsrc(a.sig, DMA ? a.sig_bytes : 0);
    if constexpr (DMA) {
#pragma unroll
        for (int j = 0; j < NBW; j++) {
            const int b = wave * NBW + j;
            const int row = ((b >> 1) % NRB) * 16 + r16;
            if (b >= 2 * NRB) {   // voxels
                drow[j] = row < tile.count
                              ? (uint32_t)a.perm[tile.pstart + row] * (uint32_t)a.T * 8u
                              : kOob;
            }
        }
    } else {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int vr = st_row + RP * p;
            vrow[p] = (vr < tile.count) ? a.sig + (size_t)a.perm[tile.pstart + vr] * (size_t)a.T
                                        : nullptr;
        }
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
        const size_t nsub = (a_end - sub < (size_t)AT) ? a_end - sub : (size_t)AT;
        const float2 *arow[NP];
        __amdgpu_buffer_rsrc_t ratom;
        if constexpr (DMA) {
            ratom = make_rsrc(a.atoms + (row_base + sub) * (size_t)a.T, nsub * (size_t)a.T * 8);
#pragma unroll
            for (int j = 0; j < NBW; j++) {
                const int b = wave * NBW + j;
                const int row = ((b >> 1) % NRB) * 16 + r16;
                if (b < 2 * NRB) drow[j] = row < (int)nsub ? (uint32_t)row * (uint32_t)a.T * 8u : kOob;
            }
        } else {
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const size_t ar = sub + st_row + RP * p;
                arow[p] = (ar < a_end) ? a.atoms + (row_base + ar) * (size_t)a.T : nullptr;
            }
        }
        // DMA: issue chunk c's copies into buffer d
        auto dma_chunk = [&](int c, floatx4 *d) {
#pragma unroll
            for (int j = 0; j < NBW; j++) {
                const int b = wave * NBW + j;
                const int t = c * KC + (b & 1) * 8 + 2 * g;
                const uint32_t off = (drow[j] != kOob && t < a.T) ? drow[j] + (uint32_t)t * 8u : kOob;
                lds_void *dst = (lds_void *)&d[b * 64];
                __builtin_amdgcn_raw_ptr_buffer_load_lds(b < 2 * NRB ? ratom : rsig, dst, 16,
                                                         (int)off, 0, 0, 0);
            }
        };
        // three real products (Gauss): P1 = ar.xr, P2 = ai.xi,
        // P3 = (ar - ai).(xr + xi); Re S = P1 + P2, Im S = P3 - P1 + P2
        floatx4 acc_p1[NIB][4], acc_p2[NIB][4], acc_p3[NIB][4];
#pragma unroll
        for (int i = 0; i < NIB; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                acc_p1[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc_p2[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc_p3[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
            }

        // global -> registers (zero outside [0, T) and outside the tile); the
        // next chunk's loads are issued before the current chunk's MFMAs and
        // stored into the other LDS buffer after three quarters of them (one
        // barrier per chunk)
        float4 ga[NP], gb[NP];
        // (the previous sub-tile's reads ended at its last chunk's barrier)
        if constexpr (DMA) {
            dma_chunk(0, lds_d0);
        } else {
#pragma unroll
            for (int p = 0; p < NP; p++) {
                ga[p] = load_pair(arow[p], st_col >> 1, a.T, even_t);
                gb[p] = load_pair(vrow[p], st_col >> 1, a.T, even_t);
            }
#pragma unroll
            for (int p = 0; p < NP; p++) {
                *reinterpret_cast<float4 *>(&lds_a[0][(st_row + RP * p) * LS + st_col]) = ga[p];
                *reinterpret_cast<float4 *>(&lds_b[0][(st_row + RP * p) * LS + st_col]) = gb[p];
            }
        }
        __syncthreads();
        // LDS fragments are double-buffered in registers: the reads for the
        // next half-chunk go out before the current MFMAs.  Per chunk: MFMAs
        // of timepoints 0-7 while 8-15 are read; the first half of 8-15;
        // the next chunk's LDS store and the barrier; its first fragments
        // are read while the second half runs.  Every read of a buffer is
        // issued before the barrier that precedes the next store into it.
        floatx4 fa[2][NIB], fx[2][4];
        // buffer b (DMA: d)
        auto read_frags = [&](int b, const floatx4 *d, int q, floatx4 (&ra)[NIB],
                              floatx4 (&rx)[4]) {
            if constexpr (DMA) {
#pragma unroll
                for (int i = 0; i < NIB; i++)
                    ra[i] = d[((0 * NRB + wa * NIB + i) * 2 + q) * 64 + lane];
#pragma unroll
                for (int v = 0; v < 4; v++)
                    rx[v] = d[((1 * NRB + wv * 4 + v) * 2 + q) * 64 + lane];
            } else {
#pragma unroll
                for (int i = 0; i < NIB; i++)
                    ra[i] = *reinterpret_cast<const floatx4 *>(
                        &lds_a[b][(wa * (AT / NWA) + i * 16 + r16) * LS + q * 16 + g * 4]);
#pragma unroll
                for (int v = 0; v < 4; v++)
                    rx[v] = *reinterpret_cast<const floatx4 *>(
                        &lds_b[b][(wv * 64 + v * 16 + r16) * LS + q * 16 + g * 4]);
            }
        };
        // MFMA step u of a fragment set: timepoint u of every lane group
        auto mfma_step = [&](const floatx4 (&ra)[NIB], const floatx4 (&rx)[4], int u) {
            float fd[NIB], fs[4];
#pragma unroll
            for (int i = 0; i < NIB; i++) fd[i] = ra[i][2 * u] - ra[i][2 * u + 1];
#pragma unroll
            for (int v = 0; v < 4; v++) fs[v] = rx[v][2 * u] + rx[v][2 * u + 1];
#pragma unroll
            for (int i = 0; i < NIB; i++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    acc_p1[i][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        ra[i][2 * u], rx[v][2 * u], acc_p1[i][v], 0, 0, 0);
                    acc_p2[i][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        ra[i][2 * u + 1], rx[v][2 * u + 1], acc_p2[i][v], 0, 0, 0);
                    acc_p3[i][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        fd[i], fs[v], acc_p3[i][v], 0, 0, 0);
                }
        };
        static_assert(KF == 32, "two 8-timepoint halves per chunk");
        read_frags(0, lds_d0, 0, fa[0], fx[0]);
        // chunk c from buffer cb (DMA: cur), the next into the other (nxt)
        auto chunk = [&](int c, int cb, floatx4 *cur, floatx4 *nxt) {
            const bool more = c + 1 < nchunks;
            if (more) {
                if constexpr (DMA) {
                    // nxt was last read in chunk c - 1, before the barrier
                    // that ended it
                    dma_chunk(c + 1, nxt);
                } else {
                    const int tc = (c + 1) * KC + (st_col >> 1);
#pragma unroll
                    for (int p = 0; p < NP; p++) {
                        ga[p] = load_pair(arow[p], tc, a.T, even_t);
                        gb[p] = load_pair(vrow[p], tc, a.T, even_t);
                    }
                }
            }
            read_frags(cb, cur, 1, fa[1], fx[1]);
            __builtin_amdgcn_sched_barrier(0);
            mfma_step(fa[0], fx[0], 0);
            mfma_step(fa[0], fx[0], 1);
            mfma_step(fa[1], fx[1], 0);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (!DMA) {
                if (more) {
#pragma unroll
                    for (int p = 0; p < NP; p++) {
                        *reinterpret_cast<float4 *>(&lds_a[cb ^ 1][(st_row + RP * p) * LS + st_col]) = ga[p];
                        *reinterpret_cast<float4 *>(&lds_b[cb ^ 1][(st_row + RP * p) * LS + st_col]) = gb[p];
                    }
                }
            }
            __syncthreads();
            if (more) read_frags(cb ^ 1, nxt, 0, fa[0], fx[0]);
            __builtin_amdgcn_sched_barrier(0);
            mfma_step(fa[1], fx[1], 1);
        };
        // two chunks per iteration: the buffers are fixed in each copy
        int c = 0;
        for (; c + 1 < nchunks; c += 2) {
            chunk(c, 0, lds_d0, lds_d1);
            chunk(c + 1, 1, lds_d1, lds_d0);
        }
        if (c < nchunks) chunk(c, 0, lds_d0, lds_d1);
        // epilogue: D[row = 4g + r][col = r16] of each 16x16 tile
#pragma unroll
        for (int i = 0; i < NIB; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const size_t atom = sub + wa * (AT / NWA) + i * 16 + 4 * g + r;
                if (atom < a_end) {
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const float p1 = acc_p1[i][v][r], p2 = acc_p2[i][v][r];
                        const float re = p1 + p2, im = acc_p3[i][v][r] - p1 + p2;
                        const float sc = re * re + im * im;
                        if (sc > best[v]) {
                            best[v] = sc;
                            bidx[v] = (uint32_t)(idx_base + atom);
                        }
                    }
                }
            }
    }

    // reduce the 4 * NWA candidates per voxel (4 k-groups x NWA atom groups)
#pragma unroll
    for (int v = 0; v < 4; v++) {
        red_s[wa][g][wv * 64 + v * 16 + r16] = best[v];
        red_i[wa][g][wv * 64 + v * 16 + r16] = bidx[v];
    }
    __syncthreads();
    if (tid < VT && tid < tile.count) {
        float bs = -1.0f;
        uint32_t bi = 0xffffffffu;
        for (int h = 0; h < NWA; h++)
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
__global__ void tiles_kernel(const int *counts, int n_slices, int vt, int *offsets, Tile *tiles,
                             int *ntiles)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int off = 0, nt = 0;
    for (int k = 0; k < n_slices; k++) {
        offsets[k] = off;
        const int c = counts[k];
        for (int p = 0; p < c; p += vt) {
            Tile t;
            t.slice = k;
            t.pstart = off + p;
            t.count = (c - p < vt) ? c - p : vt;
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
// Low-precision matchers: v_mfma_f32_16x16x32_bf16 over BF16 copies of the
// dictionary and the signals with FP32 accumulation, candidate lists per
// voxel, and a rescore (FP64 accumulation) of the final candidates from the complex64
// dictionary.  NPROD selects the products of the split representation
// x = hi + lo (hi = bf16(x), lo = bf16(x - hi)):
//   1: a_hi*s_hi                       (BF16)
//   2: a_hi*s_hi + a_lo*s_hi           (dictionary split; the signal's own
//                                       rounding shifts every atom's score
//                                       together, so the ranking is kept)
//   3: a_hi*s_hi + a_lo*s_hi + a_hi*s_lo (split-BF16, 3 products)
//
// Operands are bf16x2 words (re low, im high), rows padded with zeros to a
// stride `ld` (a multiple of 32 words); the signals are converted in tile
// order, so a tile's voxels are contiguous rows.  Rows are fetched with
// buffer loads whose range check returns zeros for rows past the end of a
// sub-tile or tile, so the staging needs no per-lane guards.
//
// Work-group tile: 128 atoms x 256 voxels, 8 waves of 64 atoms x 64 voxels
// (two per SIMD, one per atom half and voxel quarter; acc 4 x 4 MFMA blocks,
// Re and Im); Y = (xi, -xr) is formed in registers from the X fragment.
// match_lowp_kernel (the split forms) stages K chunks of 16 or 32
// timepoints through registers into unpadded LDS rows with a per-row
// segment swizzle, double-buffered (one barrier per chunk);
// match_bf16_dma_kernel (the one-product forms) copies 32-timepoint chunks
// straight into LDS.  Work-groups are numbered so that the ones on one XCD
// (work-group id mod 8) take consecutive atom chunks of the same voxel
// tiles, which then stream from that XCD's L2.
//
// Candidates: MFMA row 4g + r of block i holds atom sub + g + 4r + 16i, so
// lane group g sees the atoms = g (mod 4) and neighbouring atoms (adjacent
// T1 grid points) fall to different lanes.  Each lane keeps a sorted top-LK
// list per voxel; the 8 lists of a voxel (4 lanes x 2 atom halves) form its
// candidates per atom chunk; the rescore kernel merges the chunks, keeps the
// top K and rescores them with
// FP64 sums (first index on ties, as generate_maps, src/functions.cu:1317-1349).
// ------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int LAT = 128;          // atoms per work-group sub-tile
constexpr int kAtomHalves = 2;    // waves per SIMD: 64-atom wave tiles
constexpr int kLowpThreads = 4 * 64 * kAtomHalves;
constexpr int LVT = 256;          // voxels per work-group tile (64 per wave)
// LK candidates per lane and voxel (4, or 6 for BF16_TOP32); a voxel's four
// lane lists make its 4*LK candidates per atom chunk.
constexpr int kMaxChunks = 32;    // atom chunks per voxel tile (rescore merges their lists)
constexpr int kMinChunkAtoms = 4 * LAT;

struct LowpArgs {
    const uint32_t *ahi, *alo;   // [NATOMS][ld] bf16x2
    const uint32_t *shi, *slo;   // [V][ld] bf16x2 in tile (perm) order
    const int32_t *perm;
    const Tile *tiles;
    const int *ntiles;
    float *part_s;               // [chunks][V][4 * LK] scores (-1 = empty)
    uint32_t *part_i;            // [chunks][V][4 * LK] global atom indices
    size_t natoms;               // atoms per slice in this launch
    size_t row_stride;           // dictionary rows per slice in ahi / alo
    size_t gidx_stride;          // global atom index step per slice
    size_t gidx_base;            // global index of atom 0 of slice 0
    size_t V;
    int ld;                      // row stride (words)
    int nk;                      // K chunks, ceil(T / KC); nk * KC <= ld
    int nch;                     // atom chunks per tile
    int nsl;                     // slices every tile meets (1: the tile's own slice)
    int phase;                   // first sub-tile's dictionary sub-tile index, mod nch
    int kstep;                   // dictionary sub-tiles per slice, mod nch
};

// (re, im) -> (im, -re) on a bf16x2 word: the Y operand of Im S.
__device__ __forceinline__ uint32_t swap_neg(uint32_t w)
{
    return ((w >> 16) | (w << 16)) ^ 0x80000000u;
}

__device__ __forceinline__ u32x4 swap_neg4(u32x4 x)
{
    return u32x4{swap_neg(x[0]), swap_neg(x[1]), swap_neg(x[2]), swap_neg(x[3])};
}

// 16 bytes at byte offset voff + soff; zeros when voff is past the resource
__device__ __forceinline__ u32x4 bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
}

// Insert (sc, gi) into the sorted LK-entry list at ls/li (score descending;
// equal scores keep the earlier, lower index first), given that sc beats the
// last entry.  Branch-free; returns the new last score.
template <int LK>
__device__ __forceinline__ float lane_insert(float *ls, uint32_t *li, float sc, uint32_t gi)
{
    float S[LK], So[LK];
    uint32_t I[LK], Io[LK];
    bool keep[LK];
#pragma unroll
    for (int j = 0; j < LK; j++) {
        S[j] = ls[j];
        I[j] = li[j];
        keep[j] = S[j] >= sc;
    }
    So[0] = keep[0] ? S[0] : sc;
    Io[0] = keep[0] ? I[0] : gi;
#pragma unroll
    for (int j = 1; j < LK; j++) {
        So[j] = keep[j] ? S[j] : (keep[j - 1] ? sc : S[j - 1]);
        Io[j] = keep[j] ? I[j] : (keep[j - 1] ? gi : I[j - 1]);
    }
#pragma unroll
    for (int j = 0; j < LK; j++) {
        ls[j] = So[j];
        li[j] = Io[j];
    }
    return So[LK - 1];
}

__device__ __forceinline__ floatx4 mfma_bf16(u32x4 a, u32x4 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// atom offset within the sub-tile held by LDS row s = 16i + 4g + r
__device__ __forceinline__ int row_atom(int s)
{
    return ((s >> 2) & 3) + 4 * (s & 3) + (s & ~15);
}

// Fold the |S|^2 of one sub-tile (a wave's NI x 4 blocks of Re and Im sums:
// 16 NI atoms x 64 voxels) into the lane's top-LK lists at cs / ci.  16
// scores at a time are tested against the list threshold th[v]; the (rare)
// insertions run in a short divergent loop in increasing atom order.  Lane
// (g, r16) holds rows 4 g + r of each block: atoms lane_atom0 + 4 r + 16 i
// (of the sub-tile's first nvalid) against voxel w 64 + 16 v + r16.
template <int NI, int LK, bool FULL>
__device__ __forceinline__ void fold_subtile_body(const floatx4 (&acc_re)[NI][4],
                                                  const floatx4 (&acc_im)[NI][4], float (&th)[4],
                                                  float *cs, uint32_t *ci, int ah, int w, int g,
                                                  int r16, uint32_t lane_atom0, int nvalid)
{
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int q = ((ah * 4 + g) * LVT + w * 64 + v * 16 + r16) * LK;
#pragma unroll
        for (int h = 0; h < NI / 4; h++) {
            float sc[16];
            uint32_t m = 0;
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = 4 * h + ii, j = 4 * ii + r;
                    const float re = acc_re[i][v][r], im = acc_im[i][v][r];
                    sc[j] = re * re + im * im;
                    const bool valid = FULL || g + 4 * r + 16 * (ah * NI + i) < nvalid;
                    if (valid && sc[j] > th[v]) m |= 1u << j;
                }
            while (m) {
                const int j = __builtin_ctz(m);
                m &= m - 1;
                float s1 = sc[0];
#pragma unroll
                for (int k = 1; k < 16; k++) s1 = (j == k) ? sc[k] : s1;
                if (s1 > th[v])
                    th[v] = lane_insert<LK>(&cs[q], &ci[q], s1,
                                            lane_atom0 + 4u * (j & 3) +
                                                16u * (ah * NI + 4 * h + (j >> 2)));
            }
        }
    }
}

// a whole sub-tile (every row a valid atom) skips the per-score bound test
// (the one-product kernel; the split kernels, at the edge of their
// registers, keep the single form)
template <int NI, int LK>
__device__ __forceinline__ void fold_subtile(const floatx4 (&acc_re)[NI][4],
                                             const floatx4 (&acc_im)[NI][4], float (&th)[4],
                                             float *cs, uint32_t *ci, int ah, int w, int g,
                                             int r16, uint32_t lane_atom0, int nvalid)
{
    if (nvalid >= LAT)
        fold_subtile_body<NI, LK, true>(acc_re, acc_im, th, cs, ci, ah, w, g, r16, lane_atom0,
                                        nvalid);
    else
        fold_subtile_body<NI, LK, false>(acc_re, acc_im, th, cs, ci, ah, w, g, r16, lane_atom0,
                                         nvalid);
}

// KC timepoints per K chunk (16 or 32), NB LDS staging buffers (2: one
// barrier per chunk; 1: a second barrier once the fragments are read, which
// leaves LDS for longer lane lists); LK candidates per
// lane and voxel.
template <int NPROD, int KC, int LK, int NB>
__global__ __launch_bounds__(kLowpThreads) void match_lowp_kernel(LowpArgs a)
{
    constexpr int AH = kAtomHalves;
    constexpr int NT = kLowpThreads;        // threads: AH atom halves x 4 voxel quarters
    constexpr int NI = 8 / AH;              // 16-atom MFMA blocks per wave
    constexpr bool ASPLIT = NPROD >= 2;
    constexpr bool SSPLIT = NPROD >= 3;
    constexpr int LSB = KC;                 // dwords per LDS row
    constexpr int SEGS = KC / 4;            // 16-byte segments per row chunk
    // segment s of LDS row r sits at segment s ^ rkey(r) of the row: with
    // these keys the fragment reads (16 rows, lane group g reading segment
    // 4 ks + g) have no LDS bank conflicts (rows padded to KC + 4 words cost
    // 4 conflict cycles per ds_read_b128, measured with SQ_LDS_BANK_CONFLICT)
    auto rkey = [](int r) { return KC == 32 ? (r & 7) : (((r >> 2) & 1) << 1); };
    constexpr int RPP = NT / SEGS;          // rows staged per pass
    constexpr int PA = LAT / RPP;
    constexpr int PX = LVT / RPP;
    constexpr int BA = LAT * LSB, BX = LVT * LSB;
    static_assert(KC == 16 || KC == 32, "K chunk");
    static_assert(NB == 1 || NB == 2, "staging buffers");
    static_assert(4 * (NB * (BA * (ASPLIT ? 2 : 1) + BX * (SSPLIT ? 2 : 1)) +
                       2 * AH * 4 * LVT * LK) <= 160 * 1024,
                  "LDS budget");

    // XCD-aware numbering: work-group L runs on XCD L % 8
    unsigned L = blockIdx.x;
    if ((gridDim.x & 7u) == 0) L = (L & 7u) * (gridDim.x >> 3) + (L >> 3);
    const int tile_id = (int)(L / (unsigned)a.nch);
    const int chunk_id = (int)(L % (unsigned)a.nch);
    if (tile_id >= *a.ntiles) return;
    const Tile tile = a.tiles[tile_id];
    // The dictionary's sub-tiles (128 atoms of one slice, numbered slice
    // by slice) go to the chunks of a voxel tile round-robin: chunk c takes
    // sub-tiles j = c (mod nch).  Neighbouring grid points (near-equal
    // scores) then fall in different chunks and compete in different lane
    // lists, and the work-groups of one tile still read every sub-tile once
    // between them.  A launch over a block of the dictionary (atoms a0 ..
    // a0 + natoms - 1 of each slice, a0 a multiple of 128) numbers its
    // sub-tiles as the whole dictionary does (phase, kstep), so each of its
    // lane lists sees a subset of one lane list of the whole-dictionary
    // launch.  A chunk with no sub-tile writes empty lists.
    const size_t a_end = a.natoms;
    const size_t a_step = (size_t)a.nch * LAT;
    const size_t ldb = (size_t)a.ld * 4;    // row stride in bytes

    __shared__ __attribute__((aligned(16))) uint32_t lA[NB * BA];
    __shared__ __attribute__((aligned(16))) uint32_t lAl[ASPLIT ? NB * BA : 4];
    __shared__ __attribute__((aligned(16))) uint32_t lX[NB * BX];
    __shared__ __attribute__((aligned(16))) uint32_t lXl[SSPLIT ? NB * BX : 4];
    __shared__ __attribute__((aligned(16))) float cs[AH * 4 * LVT * LK];
    __shared__ __attribute__((aligned(16))) uint32_t ci[AH * 4 * LVT * LK];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = (tid >> 6) & 3;           // 64-voxel quarter of the tile
    const int ah = tid >> 8;                // atom half
    const int g = lane >> 4;
    const int r16 = lane & 15;
    const int st_row = tid / SEGS;
    const uint32_t st_b = (uint32_t)(tid % SEGS) * 16u;

    for (int q = tid; q < AH * 4 * LVT * LK; q += NT) {
        cs[q] = -1.0f;
        ci[q] = 0xffffffffu;
    }
    float th[4] = {-1.0f, -1.0f, -1.0f, -1.0f};

    const size_t xrow0 = (size_t)tile.pstart * a.ld;
    const __amdgpu_buffer_rsrc_t rX = make_rsrc(a.shi + xrow0, (size_t)tile.count * ldb);
    __amdgpu_buffer_rsrc_t rXl = rX;
    if (SSPLIT) rXl = make_rsrc(a.slo + xrow0, (size_t)tile.count * ldb);
    uint32_t xoff[PX], aoff[PA];
#pragma unroll
    for (int p = 0; p < PX; p++) xoff[p] = (uint32_t)((st_row + RPP * p) * ldb) + st_b;
#pragma unroll
    for (int p = 0; p < PA; p++) aoff[p] = (uint32_t)(row_atom(st_row + RPP * p) * ldb) + st_b;

    int ks = 0;                             // slice of the launch
    int r0 = chunk_id - a.phase;            // sub-tile residue of this chunk in slice ks
    if (r0 < 0) r0 += a.nch;
    size_t sub = (size_t)r0 * LAT;
    while (true) {
        if (sub >= a_end) {
            if (++ks >= a.nsl) break;
            r0 -= a.kstep;
            if (r0 < 0) r0 += a.nch;
            sub = (size_t)r0 * LAT;
            continue;
        }
        const size_t row_base = (size_t)(tile.slice + ks) * a.row_stride;
        const size_t idx_base = a.gidx_base + (size_t)(tile.slice + ks) * a.gidx_stride;
        const size_t nsub = (a_end - sub < (size_t)LAT) ? a_end - sub : (size_t)LAT;
        const size_t arow0 = (row_base + sub) * a.ld;
        const __amdgpu_buffer_rsrc_t rA = make_rsrc(a.ahi + arow0, nsub * ldb);
        __amdgpu_buffer_rsrc_t rAl = rA;
        if (ASPLIT) rAl = make_rsrc(a.alo + arow0, nsub * ldb);

        floatx4 acc_re[NI][4], acc_im[NI][4];
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                acc_re[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc_im[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
            }
        u32x4 ga[PA], gal[PA], gx[PX], gxl[PX];
        auto stage_load = [&](int c) {
            const uint32_t so = (uint32_t)c * (KC * 4);
#pragma unroll
            for (int p = 0; p < PA; p++) {
                ga[p] = bload(rA, aoff[p], so);
                if (ASPLIT) gal[p] = bload(rAl, aoff[p], so);
            }
#pragma unroll
            for (int p = 0; p < PX; p++) {
                gx[p] = bload(rX, xoff[p], so);
                if (SSPLIT) gxl[p] = bload(rXl, xoff[p], so);
            }
        };
        auto stage_store = [&](int b) {
#pragma unroll
            for (int p = 0; p < PA; p++) {
                const int row = st_row + RPP * p;
                const int o = b * BA + row * LSB + ((int)(st_b >> 4) ^ rkey(row)) * 4;
                *reinterpret_cast<u32x4 *>(&lA[o]) = ga[p];
                if (ASPLIT) *reinterpret_cast<u32x4 *>(&lAl[o]) = gal[p];
            }
#pragma unroll
            for (int p = 0; p < PX; p++) {
                const int row = st_row + RPP * p;
                const int o = b * BX + row * LSB + ((int)(st_b >> 4) ^ rkey(row)) * 4;
                *reinterpret_cast<u32x4 *>(&lX[o]) = gx[p];
                if (SSPLIT) *reinterpret_cast<u32x4 *>(&lXl[o]) = gxl[p];
            }
        };
        // All fragments of a chunk are read from LDS before its MFMAs, so the
        // LDS latency is paid once per chunk rather than once per A block.
        constexpr int KS = KC / 16;
        u32x4 fx[KS][4], fxl[KS][4], fa[KS][NI], fal[KS][NI];
        auto read_frags = [&](int b) {
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const int ko = ((ks * 4 + g) ^ rkey(r16)) * 4;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int o = b * BX + (w * 64 + v * 16 + r16) * LSB + ko;
                    fx[ks][v] = *reinterpret_cast<const u32x4 *>(&lX[o]);
                    if (SSPLIT) fxl[ks][v] = *reinterpret_cast<const u32x4 *>(&lXl[o]);
                }
#pragma unroll
                for (int i = 0; i < NI; i++) {
                    const int o = b * BA + ((ah * NI + i) * 16 + r16) * LSB + ko;
                    fa[ks][i] = *reinterpret_cast<const u32x4 *>(&lA[o]);
                    if (ASPLIT) fal[ks][i] = *reinterpret_cast<const u32x4 *>(&lAl[o]);
                }
            }
        };
        auto mfmas = [&]() {
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                u32x4 fy[4], fyl[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    fy[v] = swap_neg4(fx[ks][v]);
                    if (SSPLIT) fyl[v] = swap_neg4(fxl[ks][v]);
                }
#pragma unroll
                for (int i = 0; i < NI; i++) {
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        acc_re[i][v] = mfma_bf16(fa[ks][i], fx[ks][v], acc_re[i][v]);
                        acc_im[i][v] = mfma_bf16(fa[ks][i], fy[v], acc_im[i][v]);
                    }
                    if (ASPLIT) {
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            acc_re[i][v] = mfma_bf16(fal[ks][i], fx[ks][v], acc_re[i][v]);
                            acc_im[i][v] = mfma_bf16(fal[ks][i], fy[v], acc_im[i][v]);
                        }
                    }
                    if (SSPLIT) {
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            acc_re[i][v] = mfma_bf16(fa[ks][i], fxl[ks][v], acc_re[i][v]);
                            acc_im[i][v] = mfma_bf16(fa[ks][i], fyl[v], acc_im[i][v]);
                        }
                    }
                }
            }
        };
        // One K chunk: its fragments are read, then its MFMAs run while the
        // next chunk is staged into the other LDS buffer.  The global loads
        // are spread over the first half of the MFMAs and the LDS stores
        // over the second half, so the staging never holds up MFMA issue.
        constexpr int NMF = KS * NI * 4 * 2 * NPROD;
        constexpr int NLD = PA * (ASPLIT ? 2 : 1) + PX * (SSPLIT ? 2 : 1);
        constexpr int MPL = NMF / 2 / NLD;
        // The last chunk restages itself (in bounds, unused) so that every
        // step has the same straight-line body.
        const int nk = a.nk;
        stage_load(0);
        stage_store(0);
        __syncthreads();
        for (int c = 0; c < nk; c++) {
            read_frags(NB == 2 ? (c & 1) : 0);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (NB == 1) {
                // one staging buffer: once every wave holds its fragments
                // the buffer takes the next chunk
                __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0)
                __syncthreads();
            }
            stage_load(c + 1 < nk ? c + 1 : c);
            mfmas();
            stage_store(NB == 2 ? ((c + 1) & 1) : 0);
#pragma unroll
            for (int j = 0; j < NLD; j++) {
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);   // VMEM read
                __builtin_amdgcn_sched_group_barrier(0x008, MPL, 0); // MFMA
            }
#pragma unroll
            for (int j = 0; j < NLD; j++) {
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);   // DS write
                __builtin_amdgcn_sched_group_barrier(0x008, MPL, 0); // MFMA
            }
            __syncthreads();
        }
        fold_subtile_body<NI, LK, false>(acc_re, acc_im, th, cs, ci, ah, w, g, r16,
                             (uint32_t)(idx_base + sub) + (uint32_t)g, (int)nsub);
        sub += a_step;
    }
    __syncthreads();
    // the 4 lane lists of each voxel are its 4 * LK candidates of this chunk
    if (tid < tile.count) {
        constexpr int CK = 4 * AH * LK;
        const size_t vox = (size_t)a.perm[tile.pstart + tid];
        float *ps = a.part_s + ((size_t)chunk_id * a.V + vox) * CK;
        uint32_t *pi = a.part_i + ((size_t)chunk_id * a.V + vox) * CK;
#pragma unroll
        for (int gg = 0; gg < 4 * AH; gg++) {
            const int q = (gg * LVT + tid) * LK;
#pragma unroll
            for (int j = 0; j < LK; j++) {
                ps[gg * LK + j] = cs[q + j];
                pi[gg * LK + j] = ci[q + j];
            }
        }
    }
}

// One-product BF16 matcher with direct-to-LDS chunk copies (the TOP8,
// TOP16 and TOP64 candidate kernel): the tile, waves, sub-tile order and
// lane lists of match_lowp_kernel<1, 32, 3, 2>, with the 32-timepoint chunks
// going from global memory straight into LDS (buffer_load ... lds) -- no
// staging registers and no LDS stores in the MFMA stream.  A chunk row is
// 128 bytes (32 bf16x2 words), eight 16-byte segments.  A copy (one
// instruction of a wave) moves 8 rows: lane l takes row l / 8 of them,
// segment (l mod 8) ^ (row mod 8), and lands at byte 16 l of a 1 KB block,
// so each copy reads whole cache lines and segment s of row r sits at
// 16-byte slot 8 r + (s ^ (r mod 8)).  MFMA fragment reads (16 rows, lane
// group g reading segment 4 ks + g) are then free of LDS bank conflicts
// (measured with SQ_LDS_BANK_CONFLICT: a key of (r + r / 8) mod 8 costs 4
// conflict cycles per ds_read_b128, rows padded to 36 words as many).  The
// MFMA's K order within a chunk is the segment order, the same for atoms
// and signals, so the sums are those of the K-ordered product up to FP32
// summation order.
//
// Per chunk c (two LDS buffers, one barrier): the second-half fragments of
// c are read, the copies of chunk c + 1 go out (after the reads, so that no
// read waits for them), the MFMAs of the first half and of half the second
// run, the wave waits for its copies and meets the others at the barrier,
// the first-half fragments of c + 1 are read, and the rest of c's MFMAs run
// while they arrive.
__device__ __forceinline__ int seg_key(int r)
{
    return r & 7;
}

template <int LK>
__global__ __launch_bounds__(kLowpThreads) void match_bf16_dma_kernel(LowpArgs a)
{
    constexpr int AH = kAtomHalves;
    constexpr int NT = kLowpThreads;
    constexpr int NI = 8 / AH;              // 16-atom MFMA blocks per wave
    constexpr int NW = NT / 64;             // waves
    constexpr int CPA = LAT / 8 / NW;       // atom copies per wave and chunk
    constexpr int CPX = LVT / 8 / NW;       // signal copies per wave and chunk
    constexpr int BUF = (LAT + LVT) * 8;    // 16-byte slots per buffer: atom rows, then signal rows
    static_assert(CPA * 8 * NW == LAT && CPX * 8 * NW == LVT, "copies cover the tile");
    static_assert(16 * 2 * BUF + 4 * 2 * AH * 4 * LVT * LK <= 160 * 1024, "LDS budget");

    // XCD-aware numbering as match_lowp_kernel: work-group L runs on XCD
    // L % 8, and an XCD takes consecutive L.  Its concurrent work-groups
    // (about 32, one per CU) take 4 voxel tiles x 8 atom chunks, so each atom
    // sub-tile they stream is shared by four of them in the XCD's L2 while
    // the four voxel tiles stay there; more than 8 chunks go 8 at a time.
    unsigned L = blockIdx.x;
    if ((gridDim.x & 7u) == 0) L = (L & 7u) * (gridDim.x >> 3) + (L >> 3);
    unsigned tile_u = L / (unsigned)a.nch, chunk_u = L % (unsigned)a.nch;
    if (a.nch > 8 && (a.nch & 7) == 0) {
        const unsigned ngrp = (unsigned)a.nch >> 3;   // chunk groups
        const unsigned rest = L >> 5;
        tile_u = (rest / ngrp) * 4 + ((L >> 3) & 3);
        chunk_u = (rest % ngrp) * 8 + (L & 7);
    }
    const int tile_id = (int)tile_u;
    const int chunk_id = (int)chunk_u;
    if (tile_id >= *a.ntiles) return;
    const Tile tile = a.tiles[tile_id];
    const size_t a_end = a.natoms;
    const size_t a_step = (size_t)a.nch * LAT;
    const uint32_t ldb = (uint32_t)a.ld * 4u;   // row stride in bytes

    // two chunk buffers
    __shared__ u32x4 lb[2 * BUF];
    __shared__ __attribute__((aligned(16))) float cs[AH * 4 * LVT * LK];
    __shared__ __attribute__((aligned(16))) uint32_t ci[AH * 4 * LVT * LK];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: copy targets
    const int w = wave & 3;                 // 64-voxel quarter of the tile
    const int ah = wave >> 2;               // atom half
    const int g = lane >> 4;
    const int r16 = lane & 15;

    for (int q = tid; q < AH * 4 * LVT * LK; q += NT) {
        cs[q] = -1.0f;
        ci[q] = 0xffffffffu;
    }
    float th[4] = {-1.0f, -1.0f, -1.0f, -1.0f};

    const size_t xrow0 = (size_t)tile.pstart * a.ld;
    const __amdgpu_buffer_rsrc_t rX = make_rsrc(a.shi + xrow0, (size_t)tile.count * ldb);
    // the lane's byte offsets of its copies (row offset + segment); rows past
    // the sub-tile or tile fall outside the resource and read as zeros
    uint32_t aoff[CPA], xoff[CPX];
#pragma unroll
    for (int j = 0; j < CPA; j++) {
        const int row = (wave * CPA + j) * 8 + (lane >> 3);
        aoff[j] = (uint32_t)row_atom(row) * ldb + (uint32_t)((lane & 7) ^ seg_key(row & 15)) * 16u;
    }
#pragma unroll
    for (int j = 0; j < CPX; j++) {
        const int xr = (wave * CPX + j) * 8 + (lane >> 3);
        xoff[j] = (uint32_t)xr * ldb + (uint32_t)((lane & 7) ^ seg_key(xr & 15)) * 16u;
    }
    // the lane's fragment slots within a 16-row block, per K half
    const int fslot0 = r16 * 8 + ((0 + g) ^ seg_key(r16));
    const int fslot1 = r16 * 8 + ((4 + g) ^ seg_key(r16));

    int ks = 0;                             // slice of the launch
    int r0 = chunk_id - a.phase;            // sub-tile residue of this chunk in slice ks
    if (r0 < 0) r0 += a.nch;
    size_t sub = (size_t)r0 * LAT;
    while (true) {
        if (sub >= a_end) {
            if (++ks >= a.nsl) break;
            r0 -= a.kstep;
            if (r0 < 0) r0 += a.nch;
            sub = (size_t)r0 * LAT;
            continue;
        }
        const size_t row_base = (size_t)(tile.slice + ks) * a.row_stride;
        const size_t idx_base = a.gidx_base + (size_t)(tile.slice + ks) * a.gidx_stride;
        const size_t nsub = (a_end - sub < (size_t)LAT) ? a_end - sub : (size_t)LAT;
        const __amdgpu_buffer_rsrc_t rA = make_rsrc(a.ahi + (row_base + sub) * a.ld, nsub * ldb);

        floatx4 acc_re[NI][4], acc_im[NI][4];
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                acc_re[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc_im[i][v] = floatx4{0.f, 0.f, 0.f, 0.f};
            }
        // chunk c's copies into buffer d (each wave: CPA atom and CPX signal
        // copies of 8 rows)
        auto copy_chunk = [&](int c, u32x4 *d) {
            const int so = c * 128;
#pragma unroll
            for (int j = 0; j < CPA; j++)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rA, (lds_void *)&d[(wave * CPA + j) * 64], 16, (int)aoff[j], so, 0, 0);
#pragma unroll
            for (int j = 0; j < CPX; j++)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rX, (lds_void *)&d[LAT * 8 + (wave * CPX + j) * 64], 16, (int)xoff[j], so, 0, 0);
        };
        auto read_frags = [&](const u32x4 *d, int fslot, u32x4 (&fa)[NI], u32x4 (&fx)[4]) {
#pragma unroll
            for (int i = 0; i < NI; i++) fa[i] = d[((ah * NI + i) * 16) * 8 + fslot];
#pragma unroll
            for (int v = 0; v < 4; v++) fx[v] = d[(LAT + w * 64 + v * 16) * 8 + fslot];
        };
        auto mfmas = [&](const u32x4 (&fa)[NI], const u32x4 (&fx)[4], int i0, int i1) {
            u32x4 fy[4];
#pragma unroll
            for (int v = 0; v < 4; v++) fy[v] = swap_neg4(fx[v]);
#pragma unroll
            for (int i = i0; i < i1; i++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    acc_re[i][v] = mfma_bf16(fa[i], fx[v], acc_re[i][v]);
                    acc_im[i][v] = mfma_bf16(fa[i], fy[v], acc_im[i][v]);
                }
        };
        u32x4 fa0[NI], fx0[4], fa1[NI], fx1[4];
        const int nk = a.nk;
        // (every read of both buffers ended before the previous sub-tile's
        // last barrier)
        copy_chunk(0, lb);
        __builtin_amdgcn_s_waitcnt(0x0F70);     // vmcnt(0): this wave's copies landed
        __syncthreads();
        read_frags(lb, fslot0, fa0, fx0);
        for (int c = 0; c < nk; c++) {
            const bool more = c + 1 < nk;
            u32x4 *cur = lb + (c & 1) * BUF;
            u32x4 *nxt = lb + ((c + 1) & 1) * BUF;
            read_frags(cur, fslot1, fa1, fx1);
            __builtin_amdgcn_sched_barrier(0);
            // the copies of chunk c + 1 go out after this chunk's reads (so no
            // read waits for them); their buffer was last read in chunk c - 1,
            // before the barrier that ended it
            if (more) copy_chunk(c + 1, nxt);
            mfmas(fa0, fx0, 0, NI);
            mfmas(fa1, fx1, 0, NI / 2);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): chunk c + 1 landed
            __syncthreads();
            if (more) read_frags(nxt, fslot0, fa0, fx0);
            __builtin_amdgcn_sched_barrier(0);
            mfmas(fa1, fx1, NI / 2, NI);
        }
        fold_subtile<NI, LK>(acc_re, acc_im, th, cs, ci, ah, w, g, r16,
                             (uint32_t)(idx_base + sub) + (uint32_t)g, (int)nsub);
        sub += a_step;
    }
    __syncthreads();
    // the lane lists of each voxel are its 4 AH LK candidates of this chunk
    if (tid < tile.count) {
        constexpr int CK = 4 * AH * LK;
        const size_t vox = (size_t)a.perm[tile.pstart + tid];
        float *ps = a.part_s + ((size_t)chunk_id * a.V + vox) * CK;
        uint32_t *pi = a.part_i + ((size_t)chunk_id * a.V + vox) * CK;
#pragma unroll
        for (int gg = 0; gg < 4 * AH; gg++) {
            const int q = (gg * LVT + tid) * LK;
#pragma unroll
            for (int j = 0; j < LK; j++) {
                ps[gg * LK + j] = cs[q + j];
                pi[gg * LK + j] = ci[q + j];
            }
        }
    }
}

__device__ __forceinline__ bool cand_better(float s1, uint32_t i1, float s2, uint32_t i2)
{
    return s1 > s2 || (s1 == s2 && i1 < i2);
}

// Per voxel (one wave): merge the chunk candidate lists, keep the top K by
// low-precision score, rescore them (FP64 accumulation) from the complex64
// dictionary and keep the best (first index on ties).
struct RescoreArgs {
    const float *part_s;
    const uint32_t *part_i;
    int nchunks;
    const float2 *atoms;        // complex64 rows (the whole dictionary or one block)
    const float2 *sig;          // [V][T]
    const float *params;        // [NATOMS][4], global index
    size_t V;
    int T;
    // block layout of `atoms` (nb > 0): global atom gi of slice k = gi / gnat
    // is row k * nb + (gi - k * gnat - a0); nb = 0: row = gi
    size_t gnat, a0, nb;
    // running best over blocks (or null): the best of this call is merged
    // with it, first index on equal FP64 scores, and the maps are written from it
    double *best_s;
    uint32_t *best_i;
    float *maps;
    int32_t *idx;
    const uint32_t *order;      // voxel of wave i (or null: voxel i)
};

__device__ __forceinline__ size_t atom_row(const RescoreArgs &r, uint32_t gi)
{
    if (r.nb == 0) return gi;
    const size_t k = gi / r.gnat;
    return k * r.nb + (gi - k * r.gnat - r.a0);
}

// Acc: double (default; products of two floats are exact, so the pick is the
// exact argmax among the candidates) or float (sums in FP32, the order below:
// per lane over t = lane + 64 i, then a butterfly over the 64 lanes).
// Per voxel (one wave): the key of the rescore order, the global index of
// its best low-precision candidate (first index on ties; ~0 for none), and
// the voxel itself.
template <int CK>
__global__ __launch_bounds__(kBlock) void top1_kernel(const float *part_s, const uint32_t *part_i,
                                                      int nchunks, size_t V, uint32_t *key,
                                                      uint32_t *vox_out)
{
    const size_t vox = (size_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (vox >= V) return;
    float bs = -1.0f;
    uint32_t bi = 0xffffffffu;
    for (int e = lane; e < nchunks * CK; e += 64) {
        const size_t o = ((size_t)(e / CK) * V + vox) * CK + e % CK;
        const float sc = part_s[o];
        const uint32_t id = part_i[o];
        if (sc >= 0.0f && cand_better(sc, id, bs, bi)) {
            bs = sc;
            bi = id;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const float os = __shfl_xor(bs, off);
        const uint32_t oi = (uint32_t)__shfl_xor((int)bi, off);
        if (os >= 0.0f && cand_better(os, oi, bs, bi)) {
            bs = os;
            bi = oi;
        }
    }
    if (lane == 0) {
        key[vox] = bi;
        vox_out[vox] = (uint32_t)vox;
    }
}

// lane l's value of x, on every lane (no LDS)
__device__ __forceinline__ float wave_read(float x, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}
__device__ __forceinline__ double wave_read(double x, int l)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <int K, int CK, typename Acc>
__global__ __launch_bounds__(kBlock, K > 32 ? 2 : 1) void rescore_kernel(RescoreArgs ra)
{
    const float *part_s = ra.part_s;
    const uint32_t *part_i = ra.part_i;
    const int nchunks = ra.nchunks;
    const float2 *atoms = ra.atoms;
    const float2 *sig = ra.sig;
    const float *params = ra.params;
    const size_t V = ra.V;
    const int T = ra.T;
    constexpr int PER = kMaxChunks * CK / 64;   // entries per lane
    static_assert(kMaxChunks * CK % 64 == 0, "candidate entries per lane");
    // XCD-aware numbering: work-group L runs on XCD L % 8 and takes a
    // contiguous eighth of the (ordered) voxels, so the voxels that share an
    // XCD's L2 at one time are neighbours in the order
    unsigned L = blockIdx.x;
    if ((gridDim.x & 7u) == 0) L = (L & 7u) * (gridDim.x >> 3) + (L >> 3);
    const size_t wi = (size_t)L * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wi >= V) return;
    const size_t vox = ra.order ? (size_t)ra.order[wi] : wi;
    const int n = nchunks * CK;
    float s[PER];
    uint32_t id[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int e = lane + 64 * q;
        s[q] = -1.0f;
        id[q] = 0xffffffffu;
        if (e < n) {
            const int ch = e / CK, k = e % CK;
            const size_t o = ((size_t)ch * V + vox) * CK + k;
            s[q] = part_s[o];
            id[q] = part_i[o];
        }
    }
    // the selected candidates are wave-uniform; above 32 of them they live
    // in LDS (written by lane 0, read back as broadcasts) rather than in K
    // registers per lane, which would spill
    constexpr bool kCandLds = K > 32;
    __shared__ uint32_t cand_lds[kBlock / 64][kCandLds ? K : 1];
    uint32_t cand[kCandLds ? 1 : K];
    int nc = 0;
#pragma unroll
    for (int r = 0; r < K; r++) {
        float bs = -1.0f;
        uint32_t bi = 0xffffffffu;
#pragma unroll
        for (int q = 0; q < PER; q++)
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
        if constexpr (kCandLds) {
            if (lane == 0) cand_lds[threadIdx.x >> 6][r] = bi;
        } else {
            cand[r] = bi;
        }
        if (bs >= 0.0f) nc = r + 1;
#pragma unroll
        for (int q = 0; q < PER; q++)
            if (id[q] == bi) s[q] = -1.0f;
    }
    if constexpr (kCandLds) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // Rescore with FP64 accumulation: products of two floats are exact in
    // double, so the candidates are ranked by their (nearly) exact scores
    // and equal scores resolve to the lowest index.  Candidates go in groups
    // of at most 16 (the signal row is re-read from cache per group), which
    // keeps the accumulators within 64 registers.  Every slot of a group
    // loads a valid row (slots past the nc candidates re-read candidate 0's
    // row and are ignored), so the loop body has no branches and a group's
    // row loads are all in flight together.
    // 64 candidates in groups of 8: groups of 16 spill FP64 sums at the two
    // waves per SIMD this kernel is bounded for, and take FP32 sums to 162
    // registers (27.3 ms against 23.2 ms for FP64 in groups of 8)
    constexpr int G = K < 16 ? K : (K > 32 ? 8 : 16);
    Acc bs = -1.0f;
    uint32_t bi = 0u;
    const float2 *x = sig + vox * (size_t)T;
    // one group of G candidates from c0 (returns false past the last)
    auto group = [&](int c0) {
        if (c0 >= nc) return false;
        Acc re[G], im[G];
        const float2 *row[G];
#pragma unroll
        for (int c = 0; c < G; c++) {
            re[c] = im[c] = 0.0;
            uint32_t ci;
            if constexpr (kCandLds) ci = cand_lds[threadIdx.x >> 6][c0 + c < nc ? c0 + c : 0];
            else ci = c0 + c < nc ? cand[c0 + c] : cand[0];
            row[c] = atoms + atom_row(ra, ci) * T;
        }
        if ((T & 1) == 0) {
            // even T: rows are 16-byte aligned, two timepoints per load
            const float4 *x4 = reinterpret_cast<const float4 *>(x);
            for (int t = lane; t < T / 2; t += 64) {
                const float4 xv = x4[t];
                float4 av[G];
#pragma unroll
                for (int c = 0; c < G; c++) av[c] = reinterpret_cast<const float4 *>(row[c])[t];
#pragma unroll
                for (int c = 0; c < G; c++) {
                    re[c] += (Acc)av[c].x * xv.x + (Acc)av[c].y * xv.y +
                             ((Acc)av[c].z * xv.z + (Acc)av[c].w * xv.w);
                    im[c] += (Acc)av[c].x * xv.y - (Acc)av[c].y * xv.x +
                             ((Acc)av[c].z * xv.w - (Acc)av[c].w * xv.z);
                }
            }
        } else {
            for (int t = lane; t < T; t += 64) {
                const float2 xv = x[t];
                float2 av[G];
#pragma unroll
                for (int c = 0; c < G; c++) av[c] = row[c][t];
#pragma unroll
                for (int c = 0; c < G; c++) {
                    re[c] += (Acc)av[c].x * xv.x + (Acc)av[c].y * xv.y;
                    im[c] += (Acc)av[c].x * xv.y - (Acc)av[c].y * xv.x;
                }
            }
        }
        // Sum the 2G lane partials over the wave as a reduce-scatter: each
        // step over lane bit b halves the values a lane holds (the half of
        // lanes with bit b set keeps the upper half) and adds the partner's
        // copy, so after log2(2G) steps lane l holds value l >> (6 - S) (Re of
        // candidate k for k < G, Im of candidate k - G) summed over the lanes
        // that share its low 6 - S bits; those are then summed all-round.
        // 2G + log2(2G) shuffles instead of 12 G.
        constexpr int NV = 2 * G;
        constexpr int S = NV == 32 ? 5 : NV == 16 ? 4 : 3;
        static_assert((1 << S) == NV, "2G is 8, 16 or 32");
        Acc v[NV];
#pragma unroll
        for (int c = 0; c < G; c++) {
            v[c] = re[c];
            v[G + c] = im[c];
        }
#pragma unroll
        for (int st = 0; st < S; st++) {
            const int b = 5 - st;
            const int n = NV >> (st + 1);
            const bool upper = (lane >> b) & 1;
#pragma unroll
            for (int k = 0; k < n; k++) {
                const Acc send = upper ? v[k] : v[k + n];
                const Acc keep = upper ? v[k + n] : v[k];
                v[k] = keep + __shfl_xor(send, 1 << b);
            }
        }
#pragma unroll
        for (int b = 5 - S; b >= 0; b--) v[0] += __shfl_xor(v[0], 1 << b);
        // lanes l < 32 hold Re of candidate l >> (6 - S), lane l + 32 its Im
        const Acc vo = __shfl_xor(v[0], 32);
        const Acc sc_lane = v[0] * v[0] + vo * vo;
#pragma unroll
        for (int c = 0; c < G; c++)
            if (c0 + c < nc) {
                const Acc sc = wave_read(sc_lane, c << (6 - S));
                uint32_t ci;
                if constexpr (kCandLds) ci = cand_lds[threadIdx.x >> 6][c0 + c];
                else ci = cand[c0 + c];
                if (sc > bs || (sc == bs && ci < bi)) {
                    bs = sc;
                    bi = ci;
                }
            }
        return true;
    };
    if constexpr (kCandLds) {
        // candidates in LDS: a rolled loop keeps one group's registers live
#pragma unroll 1
        for (int c0 = 0; c0 < K; c0 += G)
            if (!group(c0)) break;
    } else {
#pragma unroll
        for (int c0 = 0; c0 < K; c0 += G)
            if (!group(c0)) break;
    }
    double bsd = (double)bs;   // an FP32 score is exact in double
    if (nc == 0) {
        bsd = -1.0;
        bi = 0xffffffffu;
    }
    if (lane == 0) {
        if (ra.best_s) {
            const double ps = ra.best_s[vox];
            const uint32_t pi = ra.best_i[vox];
            if (!(bsd > ps || (bsd == ps && bi < pi))) {
                bsd = ps;
                bi = pi;
            }
            ra.best_s[vox] = bsd;
            ra.best_i[vox] = bi;
        }
        if (bsd < 0.0) {   // no candidate: the reference leaves index 0
            bsd = 0.0;
            bi = 0u;
        }
        const float *p = params + (size_t)bi * 4;
        ra.maps[vox * 5 + 0] = p[0];
        ra.maps[vox * 5 + 1] = p[1];
        ra.maps[vox * 5 + 2] = p[2];
        ra.maps[vox * 5 + 3] = p[3];
        ra.maps[vox * 5 + 4] = (float)sqrt(bsd);
        if (ra.idx) ra.idx[vox] = (int32_t)bi;
    }
}

__global__ void init_best_kernel(double *best_s, uint32_t *best_i, size_t V)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    best_s[v] = -1.0;
    best_i[v] = 0xffffffffu;
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
    size_t chunk, nchunks, part_chunks;
    // the rescore's voxel order: best-candidate keys and voxels, sorted
    uint32_t *rkey, *rvox, *rkey_s, *rvox_s;
    void *sort_tmp;
    size_t sort_tmp_bytes;
    size_t bytes;
};

// Atom chunks per voxel tile of the low-precision path: at least 8 (so the
// work-groups of one XCD share a voxel tile), enough for about 2048
// work-groups, at most kMaxChunks; chunks hold whole sub-tiles and at least
// kMinChunkAtoms atoms.
// candidates per lane and voxel of a low-precision mode
int lane_k(int mode)
{
    switch (mode) {
    case MRF_MATCH_BF16_TOP32: return 3;
    case MRF_MATCH_BF16_TOP64: return 3;
    case MRF_MATCH_SPLIT2: return 4;
    case MRF_MATCH_SPLIT3: return 2;
    default: return 3;
    }
}



// atom chunks per voxel tile: about 2048 work-groups, 8 to kMaxChunks.
// BF16_TOP64 always takes kMaxChunks: its lane lists are the short 3-entry
// ones of TOP8/TOP16, and lists that each cover 1/kMaxChunks of the atoms
// keep the FP32 argmax among its 64 rescored candidates.  BF16_TOP32 splits
// each of TOP16's chunks in two where that stays within kMaxChunks (else it
// takes TOP16's chunks): every TOP16 candidate is then a TOP32 candidate,
// and at most 30 TOP32 candidates outscore any of TOP16's best 16 (an extra
// candidate outscores it only where all 3 of its TOP16 list do, at most 5
// lists), so TOP32 rescores every candidate TOP16 rescores.
size_t lowp_want(size_t max_tiles, int mode)
{
    if (mode == MRF_MATCH_BF16_TOP64) return kMaxChunks;
    size_t want = (2048 + max_tiles - 1) / max_tiles;
    if (want < 8) want = 8;
    if (want > (size_t)kMaxChunks) want = kMaxChunks;
    if (mode == MRF_MATCH_BF16_TOP32 && 2 * want <= (size_t)kMaxChunks) want *= 2;
    return want;
}

// chunking of nat atoms; at most lowp_want chunks for any nat (*chunk: the
// atoms per chunk, for BF16_TOP32 those of the TOP16 chunk it splits)
void lowp_chunks(size_t nat, size_t max_tiles, int mode, size_t *chunk, size_t *nchunks)
{
    const bool top32 = mode == MRF_MATCH_BF16_TOP32;
    const size_t want = lowp_want(max_tiles, top32 ? MRF_MATCH_BF16_TOP16 : mode);
    size_t c = (nat + want - 1) / want;
    if (c < (size_t)kMinChunkAtoms) c = kMinChunkAtoms;
    c = (c + LAT - 1) / LAT * LAT;
    *chunk = c;
    *nchunks = (nat + c - 1) / c;
    if (*nchunks == 0) *nchunks = 1;
    if (top32 && 2 * *nchunks <= (size_t)kMaxChunks) *nchunks *= 2;
}

Scratch carve(void *base, const MatchLaunch &l)
{
    Scratch s{};
    const int ns = l.vox_slice ? l.n_slices : 1;
    const size_t nat = l.vox_slice ? l.natoms : l.natoms * (size_t)l.n_slices;
    const size_t vt = (l.mode == MRF_MATCH_FP32) ? VT : LVT;
    s.max_tiles = (l.V + vt - 1) / vt + (size_t)ns;
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
        lowp_chunks(nat, s.max_tiles, l.mode, &s.chunk, &s.nchunks);
        const size_t ld = lowp_ld(l.T);
        s.shi = (uint32_t *)p;
        p += align256(l.V * ld * 4);
        if (l.mode == MRF_MATCH_SPLIT3) {
            s.slo = (uint32_t *)p;
            p += align256(l.V * ld * 4);
        }
        // candidate lists for up to lowp_want chunks (any launch's nat)
        s.part_chunks = lowp_want(s.max_tiles, l.mode);
        s.part_s = (float *)p;
        const size_t ck = 4 * (size_t)kAtomHalves * lane_k(l.mode);
        p += align256(s.part_chunks * l.V * ck * 4);
        s.part_i = (uint32_t *)p;
        p += align256(s.part_chunks * l.V * ck * 4);
        uint32_t **ord[4] = {&s.rkey, &s.rvox, &s.rkey_s, &s.rvox_s};
        for (uint32_t **q : ord) {
            *q = (uint32_t *)p;
            p += align256(l.V * 4);
        }
        s.sort_tmp_bytes = voxel_sort_tmp_bytes(l.V);
        s.sort_tmp = p;
        p += align256(s.sort_tmp_bytes);
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

namespace {

// Argument checks shared by the matching launches.
int match_check(const MatchLaunch &l)
{
    const bool lowp = l.mode != MRF_MATCH_FP32;
    if (l.mode < MRF_MATCH_FP32 || l.mode > MRF_MATCH_BF16_TOP64)
        return mrf_fail(MRF_E_UNSUPPORTED, "matching mode %d not available", l.mode);
    const bool need_lo = l.mode == MRF_MATCH_SPLIT3 || l.mode == MRF_MATCH_SPLIT2;
    if (lowp && (!l.lowp_hi || (need_lo && !l.lowp_lo)))
        return mrf_fail(MRF_E_ARG, "low-precision matching needs the bf16 dictionary%s",
                        need_lo ? " and its lo part" : "");
    if (lowp && l.T > kLowpMaxT)
        return mrf_fail(MRF_E_ARG, "low-precision matching supports T <= %d", kLowpMaxT);
    const int ns = l.vox_slice ? l.n_slices : 1;
    const size_t nat = l.vox_slice ? l.natoms : l.natoms * (size_t)l.n_slices;
    if ((size_t)ns * nat > 0xfffffffeULL)
        return mrf_fail(MRF_E_ARG, "dictionary larger than 2^32-1 atoms");
    if (l.V > 0x7fffffffULL) return mrf_fail(MRF_E_ARG, "too many voxels");
    return 0;
}

// Slice bucketing of the voxels and zeroed keys.
Scratch match_setup(const MatchLaunch &l, void *scratch, hipStream_t s)
{
    const int ns = l.vox_slice ? l.n_slices : 1;
    Scratch sc = carve(scratch, l);
    const unsigned vb = (unsigned)((l.V + 255) / 256);
    zero_u64_kernel<<<vb, 256, 0, s>>>(sc.keys, l.V);
    const unsigned nb = (unsigned)((ns + 255) / 256);
    zero_i32_kernel<<<nb, 256, 0, s>>>(sc.counts, (size_t)ns);
    zero_i32_kernel<<<nb, 256, 0, s>>>(sc.fill, (size_t)ns);
    count_kernel<<<vb, 256, 0, s>>>(l.vox_slice, l.V, ns, sc.counts);
    tiles_kernel<<<1, 1, 0, s>>>(sc.counts, ns, l.mode == MRF_MATCH_FP32 ? VT : LVT, sc.offsets,
                                 sc.tiles, sc.ntiles);
    scatter_kernel<<<vb, 256, 0, s>>>(l.vox_slice, l.V, ns, sc.offsets, sc.fill, sc.perm);
    return sc;
}

// One FP32 launch over atoms [0, nat) of every slice of `atoms` (rows
// row_stride apart per slice); keys take global index
// gbase + slice * gstride + atom.
void launch_fp32(const MatchLaunch &l, const Scratch &sc, const float2 *atoms, size_t nat,
                 size_t row_stride, size_t gstride, size_t gbase, hipStream_t s)
{
    MatchArgs a{};
    a.atoms = atoms;
    a.sig = l.sig;
    a.perm = sc.perm;
    a.tiles = sc.tiles;
    a.ntiles = sc.ntiles;
    a.keys = sc.keys;
    a.natoms = nat;
    a.row_stride = row_stride;
    a.gidx_stride = gstride;
    a.gidx_base = gbase;
    a.T = l.T;
    // at most kAtomChunk atoms per work-group; small problems split the
    // atoms further so that there are about 2048 work-groups
    size_t want = (2048 + sc.max_tiles - 1) / sc.max_tiles;
    size_t chunk = (nat + want - 1) / want;
    chunk = (chunk + AT - 1) / AT * AT;
    if (chunk > (size_t)kAtomChunk) chunk = kAtomChunk;
    a.chunk = chunk;
    const size_t chunks = (nat + chunk - 1) / chunk;
    dim3 grid((unsigned)sc.max_tiles, (unsigned)chunks);
    // direct-to-LDS copies need 16-byte rows (even T) and 32-bit byte offsets
    // into the signals and a sub-tile's atoms, below kOob
    a.sig_bytes = l.V * (size_t)l.T * sizeof(float2);
    if ((l.T & 1) == 0 && a.sig_bytes <= kOob && (size_t)AT * l.T * sizeof(float2) <= kOob)
        match_fp32_kernel<true><<<grid, kFp32Threads, 0, s>>>(a);
    else
        match_fp32_kernel<false><<<grid, kFp32Threads, 0, s>>>(a);
}

// tile-ordered BF16 (and lo) copies of the signals
int lowp_prepare_signals(const MatchLaunch &l, const Scratch &sc, hipStream_t s)
{
    return launch_lowp_prepare(l.sig, sc.perm, l.V, l.T, (int)lowp_ld(l.T), sc.shi,
                               l.mode == MRF_MATCH_SPLIT3 ? sc.slo : nullptr, s);
}

RescoreArgs rescore_args(const MatchLaunch &l, const Scratch &sc)
{
    RescoreArgs r{};
    r.part_s = sc.part_s;
    r.part_i = sc.part_i;
    r.sig = l.sig;
    r.params = l.params;
    r.V = l.V;
    r.T = l.T;
    r.maps = l.maps;
    r.idx = l.idx;
    return r;
}

// One low-precision launch (candidates, then their rescore) over atoms
// [0, nat) of nsl slices of hi / lo (rows row_stride apart per slice; nsl = 1:
// each voxel tile's own slice), global index gbase + slice * gstride + atom.
// The launch's atom 0 is atom a0 = gbase mod gstride of its slice; its
// sub-tiles take the chunks of the whole dictionary's sub-tiles.
int launch_lowp(const MatchLaunch &l, const Scratch &sc, const uint32_t *hi, const uint32_t *lo,
                size_t nat, int nsl, size_t row_stride, size_t gstride, size_t gbase,
                RescoreArgs r, hipStream_t s)
{
    const size_t nchunks = sc.nchunks;
    if (nchunks > sc.part_chunks) return mrf_fail(MRF_E_ARG, "candidate storage too small");
    const size_t a0 = gbase % gstride;
    if (a0 % LAT) return mrf_fail(MRF_E_ARG, "block start not a multiple of %d atoms", LAT);
    const size_t nsub_w = (l.natoms + LAT - 1) / LAT;   // sub-tiles per dictionary slice
    LowpArgs b{};
    b.ahi = hi;
    b.alo = lo;
    b.shi = sc.shi;
    b.slo = sc.slo;
    b.perm = sc.perm;
    b.tiles = sc.tiles;
    b.ntiles = sc.ntiles;
    b.part_s = sc.part_s;
    b.part_i = sc.part_i;
    b.natoms = nat;
    b.row_stride = row_stride;
    b.gidx_stride = gstride;
    b.gidx_base = gbase;
    b.V = l.V;
    b.ld = (int)lowp_ld(l.T);
    b.nch = (int)nchunks;
    b.nsl = nsl;
    b.phase = (int)((a0 / LAT) % nchunks);
    b.kstep = (int)(nsub_w % nchunks);
    // whole groups of 4 tiles (the DMA kernel's numbering), whole groups of
    // eight work-groups (the XCD numbering)
    const size_t nwg = ((sc.max_tiles + 3) / 4 * 4 * nchunks + 7) / 8 * 8;
    if (nwg > 0x7fffffffULL) return mrf_fail(MRF_E_ARG, "too many work-groups");
    const dim3 grid((unsigned)nwg);
    r.nchunks = (int)nchunks;
    // Launch shapes measured on the 446k-atom x 460,800-voxel workload:
    // eight waves (two per SIMD) were 6-13 % faster than four waves of
    // 128-atom tiles in every mode; the LDS (chunk buffers + the lane lists)
    // sets the chunk length and list size per mode.
    // TOP8/TOP16/TOP32/TOP64: match_bf16_dma_kernel (32-timepoint chunks
    // copied straight to LDS, two buffers, 3-entry lists) over 8-32, 16 and
    // 32 strided atom chunks; on 500,000 atoms of the 871k-atom grid (B0
    // steps of 2 Hz) TOP32 keeps the FP32 pick on as many voxels as 6-entry
    // lists over 8 chunks (99.66 %) in 282 ms against 359 ms, TOP64 on
    // 99.92 % in 304 ms.  SPLIT2: match_lowp_kernel with 16-timepoint chunks
    // and 4-entry lists.  SPLIT3: 16, 2-entry (their fragments of
    // 32-timepoint chunks do not fit the registers of two waves per SIMD).
    // rescore grid: whole groups of eight work-groups (the XCD numbering)
    const dim3 rg((unsigned)(((l.V + 3) / 4 + 7) / 8 * 8));
    const dim3 blk(kLowpThreads);
    // the rescore order: voxels sorted by their best candidate's index
    int kbits = 1;
    while (kbits < 32 && (l.natoms * (size_t)l.n_slices >> kbits) != 0) kbits++;
    r.order = sc.rvox_s;
#define MRF_RESCORE(K, CK)                                                  \
    do {                                                                    \
        top1_kernel<CK><<<(unsigned)((l.V + 3) / 4), kBlock, 0, s>>>(       \
            sc.part_s, sc.part_i, (int)nchunks, l.V, sc.rkey, sc.rvox);     \
        if (int rc = sort_voxels(sc.sort_tmp, sc.sort_tmp_bytes, sc.rkey, sc.rkey_s, sc.rvox, \
                                 sc.rvox_s, l.V, kbits, s))                 \
            return rc;                                                      \
        if (l.rescore_fp32) rescore_kernel<K, CK, float><<<rg, kBlock, 0, s>>>(r);   \
        else rescore_kernel<K, CK, double><<<rg, kBlock, 0, s>>>(r);                 \
    } while (0)
    if (l.mode == MRF_MATCH_SPLIT3) {
        b.nk = (l.T + 15) / 16;
        match_lowp_kernel<3, 16, 2, 2><<<grid, blk, 0, s>>>(b);
        MRF_RESCORE(8, 16);
    } else if (l.mode == MRF_MATCH_SPLIT2) {
        b.nk = (l.T + 15) / 16;
        match_lowp_kernel<2, 16, 4, 2><<<grid, blk, 0, s>>>(b);
        MRF_RESCORE(8, 32);
    } else if (l.mode == MRF_MATCH_BF16_TOP64) {
        b.nk = (l.T + 31) / 32;
        match_bf16_dma_kernel<3><<<grid, blk, 0, s>>>(b);
        MRF_RESCORE(64, 24);
    } else if (l.mode == MRF_MATCH_BF16_TOP32) {
        b.nk = (l.T + 31) / 32;
        match_bf16_dma_kernel<3><<<grid, blk, 0, s>>>(b);
        MRF_RESCORE(32, 24);
    } else {
        b.nk = (l.T + 31) / 32;
        match_bf16_dma_kernel<3><<<grid, blk, 0, s>>>(b);
        if (l.mode == MRF_MATCH_BF16_TOP16) MRF_RESCORE(16, 24);
        else MRF_RESCORE(8, 24);
    }
#undef MRF_RESCORE
    return 0;
}

}  // namespace

int launch_match(const MatchLaunch &l, void *scratch, hipStream_t s)
{
    if (l.V == 0) return 0;
    if (int rc = match_check(l)) return rc;
    const size_t nat = l.vox_slice ? l.natoms : l.natoms * (size_t)l.n_slices;
    Scratch sc = match_setup(l, scratch, s);
    const unsigned vb = (unsigned)((l.V + 255) / 256);
    if (l.mode != MRF_MATCH_FP32) {
        if (int rc = lowp_prepare_signals(l, sc, s)) return rc;
        RescoreArgs r = rescore_args(l, sc);
        r.atoms = l.atoms;
        if (int rc = launch_lowp(l, sc, l.lowp_hi, l.lowp_lo, l.natoms,
                                 l.vox_slice ? 1 : l.n_slices, l.natoms, l.natoms, 0, r, s))
            return rc;
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "match launch: %s", hipGetErrorString(e));
    }

    if ((nat + kAtomChunk - 1) / kAtomChunk > 65535)
        return mrf_fail(MRF_E_ARG, "dictionary slice too large");
    launch_fp32(l, sc, l.atoms, nat, nat, nat, 0, s);
    maps_kernel<<<vb, 256, 0, s>>>(sc.keys, l.params, l.V, l.maps, l.idx);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "match launch: %s", hipGetErrorString(e));
}

int launch_match_blocks(const MatchLaunch &l, size_t block, float2 *buf,
                        const LowpBlockBufs *lb, void *scratch, hipStream_t s,
                        const std::function<int(size_t, size_t)> &gen)
{
    const bool lowp = l.mode != MRF_MATCH_FP32;
    if (block == 0) return mrf_fail(MRF_E_ARG, "block size must be positive");
    if (lowp && (!lb || !lb->hi || !lb->best_s || !lb->best_i))
        return mrf_fail(MRF_E_ARG, "block-streamed low-precision matching needs its buffers");
    const bool need_lo = l.mode == MRF_MATCH_SPLIT2 || l.mode == MRF_MATCH_SPLIT3;
    if (need_lo && !lb->lo) return mrf_fail(MRF_E_ARG, "split mode needs the lo block buffer");
    if (l.V == 0) return 0;
    MatchLaunch lc = l;
    if (lowp) lc.lowp_hi = lc.lowp_lo = reinterpret_cast<const uint32_t *>(lb->hi);
    if (int rc = match_check(lc)) return rc;
    if ((block + kAtomChunk - 1) / kAtomChunk > 65535)
        return mrf_fail(MRF_E_ARG, "block too large");
    Scratch sc = match_setup(l, scratch, s);
    const unsigned vb = (unsigned)((l.V + 255) / 256);
    const int ld = (int)lowp_ld(l.T);
    if (lowp) {
        if (int rc = lowp_prepare_signals(l, sc, s)) return rc;
        init_best_kernel<<<vb, 256, 0, s>>>(lb->best_s, lb->best_i, l.V);
    }
    for (size_t a0 = 0; a0 < l.natoms; a0 += block) {
        const size_t nb = (l.natoms - a0 < block) ? l.natoms - a0 : block;
        if (int rc = gen(a0, nb)) return rc;
        if (!lowp) {
            if (l.vox_slice) {
                launch_fp32(l, sc, buf, nb, nb, l.natoms, a0, s);
            } else {
                // every voxel meets every slice: one launch per slice
                for (int k = 0; k < l.n_slices; k++)
                    launch_fp32(l, sc, buf + (size_t)k * nb * l.T, nb, 0, 0,
                                (size_t)k * l.natoms + a0, s);
            }
            continue;
        }
        // BF16 copies of the block, candidates, and their rescore against
        // the block's complex64 rows, merged into the running best per voxel
        const size_t rows = (size_t)l.n_slices * nb;
        if (int rc = launch_lowp_prepare(buf, nullptr, rows, l.T, ld, lb->hi,
                                         need_lo ? lb->lo : nullptr, s))
            return rc;
        RescoreArgs r = rescore_args(l, sc);
        r.atoms = buf;
        r.gnat = l.natoms;
        r.a0 = a0;
        r.nb = nb;
        r.best_s = lb->best_s;
        r.best_i = lb->best_i;
        if (int rc = launch_lowp(l, sc, lb->hi, need_lo ? lb->lo : nullptr, nb,
                                 l.vox_slice ? 1 : l.n_slices, nb, l.natoms, a0, r, s))
            return rc;
    }
    if (!lowp) maps_kernel<<<vb, 256, 0, s>>>(sc.keys, l.params, l.V, l.maps, l.idx);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "match launch: %s", hipGetErrorString(e));
}

}  // namespace mrf

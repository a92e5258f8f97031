// sampler_kernels.hpp -- gfx950 kernels of the exact KRP leverage sampler.
//
//   k_leaf_gram     leaf Grams U_blk^T U_blk (partition_tree.hpp:136-154)
//   k_level_sum     parent = left + right, one tree level (partition_tree.hpp:156-168)
//   k_tree_to_f32   FP32 copy of the bottom levels of an FP64 tree
//   k_gram_pinv     computeM part 1: G = Hadamard of sampled root Grams, its
//                   eigendecomposition and pinv (efficient_krp_sampler.hpp:90-150)
//   k_eigen_modes   computeM part 2: per-mode eigen matrices E_k and eigen
//                   trees (efficient_krp_sampler.hpp:152-193)
//   k_krp_draw      the whole KRPDrawSamples after computeM, fused: for every
//                   sample, every sampled mode's eigen descent, Gram descent,
//                   leaf row scan and h update, then the weight
//                   (efficient_krp_sampler.hpp:209-268, partition_tree.hpp:232-405)
//
// Trees are stored level-major (heap order) with each node's R x R Gram as a
// packed upper triangle.  In FP32 mode the top `d64` levels stay FP64 and the
// levels below are FP32: a descent compares u against a running cutoff whose
// absolute error is set by the top levels, so those must be FP64, while the
// bottom levels (where the HBM traffic is) tolerate FP32.
#pragma once

#include "ftl_common.hpp"
#include "jacobi.hpp"
#include "wave_ops.hpp"

namespace ftl {

// ---------------------------------------------------------------------------
// Tree build kernels (FP64).  Leaf blocks are staged in LDS and each packed
// upper-triangle entry is a sequential sum over the leaf's rows (no
// contraction, so the result is the exactly rounded sequential sum the
// reference's dsyrk restatement produces); every internal node is the sum of
// its two children.
// ---------------------------------------------------------------------------
struct BuildArgs {
    const double* U;   // n x R
    double* tree;      // node_count x ts64
    Shape sh;
    uint32_t R;
    uint32_t ts64;
    uint64_t first_leaf;  // node_count - leaf_count
    // FP32 samplers (else null): the factor's FP32 copy, and the FP32 copy of
    // the nodes from n64 on (ts32 floats each, zero past ts64)
    float* U32;
    float* tree32;
    uint64_t n64;
    uint32_t ts32;
};

// FP32 copy of entry e (< ts32) of node c from its FP64 value (entries past
// ts64 are zero), when c is in the FP32 levels
__device__ inline void store_f32(float* tree32, uint64_t n64, uint32_t ts32, uint32_t ts64, uint64_t c, uint32_t e,
                                 double v) {
    if (tree32 && c >= n64) tree32[(c - n64) * ts32 + e] = e < ts64 ? (float)v : 0.0f;
}

// lpb leaves per block (BS / lpb threads each).  Dynamic LDS: the leaves'
// row blocks (lpb x F x R doubles), then the (row, column) of every packed
// entry (T uint16).  Entry e of a leaf is the sequential sum over its rows.
template <int BS>
__global__ void __launch_bounds__(BS) k_leaf_gram(BuildArgs a, uint32_t lpb) {
#pragma clang fp contract(off)
    extern __shared__ double lds_u[];
    const uint32_t R = a.R;
    const uint32_t F = (uint32_t)a.sh.F;
    const uint32_t T = (uint32_t)tri_size(R);
    uint16_t* pairs = reinterpret_cast<uint16_t*>(lds_u + (uint64_t)lpb * F * R);
    for (uint32_t r = threadIdx.x; r < R; r += BS) {
        const uint32_t base = r * R - r * (r - 1) / 2;
        for (uint32_t c = r; c < R; c++) pairs[base + c - r] = (uint16_t)(r | (c << 8));
    }
    const uint32_t tpl = BS / lpb;
    const uint32_t q = threadIdx.x / tpl, lt = threadIdx.x - q * tpl;
    const uint64_t i = (uint64_t)blockIdx.x * lpb + q;  // leaf position
    const bool have = q < lpb && i < a.sh.leaf_count;
    uint64_t node = 0, r0 = 0;
    uint32_t rows = 0;
    double* lu = lds_u + (uint64_t)q * F * R;
    if (have) {
        node = a.first_leaf + i;
        r0 = leaf_index(a.sh, node) * F;
        rows = (uint32_t)min<uint64_t>(F, a.sh.n - r0);
        const double* src = a.U + r0 * R;
        for (uint32_t e = lt; e < rows * R; e += tpl) {
            const double x = src[e];
            lu[e] = x;
            if (a.U32) a.U32[r0 * R + e] = (float)x;
        }
    }
    __syncthreads();
    if (!have) return;
    double* out = a.tree + node * (uint64_t)a.ts64;
    const uint32_t tsw = a.tree32 && a.ts32 > a.ts64 ? a.ts32 : a.ts64;
    for (uint32_t e = lt; e < tsw; e += tpl) {
        double acc = 0.0;
        if (e < T) {
            const uint32_t ab = pairs[e];
            const uint32_t ra = ab & 0xffu, cb = ab >> 8;
            for (uint32_t t = 0; t < rows; t++) acc = acc + lu[t * R + ra] * lu[t * R + cb];
        }
        if (e < a.ts64) out[e] = acc;
        store_f32(a.tree32, a.n64, a.ts32, a.ts64, node, e, acc);
    }
}

// Leaf Grams of FP32 samplers on the FP64 matrix cores (R <= 16 NB <= 64): a
// wave per leaf, four leaves per block.  v_mfma_f64_16x16x4f64 takes A
// (16 x 4) with A[i][k] in lane 16k + i and B (4 x 16) with B[k][j] in lane
// 16k + j and accumulates D (16 x 16) with D[4v + l/16][l%16] in item v of
// lane l; with lane l holding x_b = U[row 4s + l/16][16b + l%16] of the leaf,
// mfma(x_a, x_b) adds rows 4s..4s+3 to Gram block (a, b), a <= b.  The sums
// are those of the rows in groups of four (their order inside the matrix
// core), so the FP64 nodes equal k_leaf_gram's to rounding, not bitwise --
// used only where the bottom levels are stored FP32.  Writes the FP32 factor
// rows, the FP64 leaf and its FP32 copy like k_leaf_gram.
typedef double lg_f64x4 __attribute__((ext_vector_type(4)));

// One wave: the leaf Gram of `node` as NB (NB + 1) / 2 accumulator tiles
// (upper blocks (p, q), p <= q, of 16 x 16, in the order p, then q), and the
// leaf's FP32 factor rows when a.U32 is set.
template <int NB>
__device__ inline void leaf_mfma_tiles(const BuildArgs& a, uint64_t node, uint32_t lane,
                                       lg_f64x4 (&acc)[NB * (NB + 1) / 2]) {
    constexpr int NT = NB * (NB + 1) / 2;
    const uint32_t R = a.R;
    const uint32_t F = (uint32_t)a.sh.F;
    const uint64_t r0 = leaf_index(a.sh, node) * F;
    const uint32_t rows = (uint32_t)min<uint64_t>(F, a.sh.n - r0);
    const uint32_t kr = lane >> 4, ci = lane & 15;
    const double* __restrict__ U = a.U + r0 * R;
    float* __restrict__ U32 = a.U32 ? a.U32 + r0 * R : nullptr;
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = lg_f64x4{0.0, 0.0, 0.0, 0.0};
    // four 4-row steps per trip: their 4 NB loads are issued before the MFMAs
    for (uint32_t s16 = 0; s16 < rows; s16 += 16) {
        double x[4][NB];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t row = s16 + 4 * u + kr;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const uint32_t col = 16 * b + ci;
                x[u][b] = row < rows && col < R ? U[row * R + col] : 0.0;
            }
        }
        if (U32) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t row = s16 + 4 * u + kr;
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const uint32_t col = 16 * b + ci;
                    if (row < rows && col < R) U32[row * R + col] = (float)x[u][b];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int t = 0;
#pragma unroll
            for (int p = 0; p < NB; p++)
#pragma unroll
                for (int q = p; q < NB; q++) {
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[u][p], x[u][q], acc[t], 0, 0, 0);
                    t++;
                }
        }
    }
}

// The packed entries of the tiles held by this lane: put(e, value) for each
// upper-triangle entry e (< T) the lane holds.
template <int NB, class Put>
__device__ inline void leaf_mfma_put(const lg_f64x4 (&acc)[NB * (NB + 1) / 2], uint32_t lane, uint32_t R, Put put) {
    const uint32_t kr = lane >> 4, ci = lane & 15;
    int t = 0;
#pragma unroll
    for (int p = 0; p < NB; p++)
#pragma unroll
        for (int q = p; q < NB; q++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const uint32_t r = 16 * p + 4 * v + kr, c = 16 * q + ci;
                if (r < R && c < R && r <= c) put(r * R - r * (r - 1) / 2 + c - r, acc[t][v]);
            }
            t++;
        }
}

template <int NB>
__global__ void __launch_bounds__(256) k_leaf_gram_mfma(BuildArgs a) {
    constexpr int NT = NB * (NB + 1) / 2;
    const uint32_t T = (uint32_t)tri_size(a.R);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t i = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= a.sh.leaf_count) return;  // whole waves; no barrier follows
    const uint64_t node = a.first_leaf + i;
    lg_f64x4 acc[NT];
    leaf_mfma_tiles<NB>(a, node, lane, acc);
    double* out = a.tree + node * (uint64_t)a.ts64;
    leaf_mfma_put<NB>(acc, lane, a.R, [&](uint32_t e, double v) {
        out[e] = v;
        store_f32(a.tree32, a.n64, a.ts32, a.ts64, node, e, v);
    });
    const uint32_t tsw = a.tree32 && a.ts32 > a.ts64 ? a.ts32 : a.ts64;
    for (uint32_t e = T + lane; e < tsw; e += 64) {
        if (e < a.ts64) out[e] = 0.0;
        store_f32(a.tree32, a.n64, a.ts32, a.ts64, node, e, 0.0);
    }
}

// FP32 samplers whose FP32 levels are stored only in FP32: a block of WPB
// waves builds the whole subtree of one node c at depth s, the lowest FP64
// level.  The subtree's leaves are taken in order, WPB at a time (a wave
// each, as k_leaf_gram_mfma; thread 0 lists them in LDS), staged as FP64
// packed nodes in LDS and written FP32.  Each leaf is then folded in: a left
// child waits in the slot of its depth; a right child is added to it (parent
// = left + right in FP64, the additions of the level-wise build, so every
// node equals that build's) and the parent goes on up, written FP32, until
// c, which is written FP64.  Every thread touches the LDS entries
// e = threadIdx.x + 64 WPB k only, except the waves' leaf stores, which
// barriers separate from the rest.  Dynamic LDS: (WPB + depth - s) slots of
// ts64 doubles.
template <int NB, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_subtree_build(BuildArgs a, uint64_t c0, uint64_t count, uint32_t s) {
    constexpr int NT = NB * (NB + 1) / 2;
    constexpr uint32_t BS = 64 * WPB;
    extern __shared__ double slot[];
    __shared__ uint64_t lv[WPB];
    __shared__ uint64_t next_leaf;
    __shared__ int n_batch, more_s;
    const uint32_t ts = a.ts64, ts32 = a.ts32;
    const uint32_t T = (uint32_t)tri_size(a.R);
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t first_leaf = a.first_leaf;
    for (uint64_t b = blockIdx.x; b < count; b += gridDim.x) {
        const uint64_t c = c0 + b;
        if (c >= first_leaf) {
            // a node of depth s that is a leaf: an FP64 node of its own
            if (w == 0) {
                lg_f64x4 acc[NT];
                leaf_mfma_tiles<NB>(a, c, lane, acc);
                double* out = a.tree + c * (uint64_t)ts;
                leaf_mfma_put<NB>(acc, lane, a.R, [&](uint32_t e, double v) { out[e] = v; });
                for (uint32_t e = T + lane; e < ts; e += 64) out[e] = 0.0;
            }
            continue;
        }
        if (threadIdx.x == 0) {
            uint64_t x = c;  // the subtree's first leaf
            while (x < first_leaf) x = 2 * x + 1;
            next_leaf = x;
            more_s = 1;
        }
        int more = 1;
        while (more) {
            if (threadIdx.x == 0) {
                // up to WPB leaves in order from next_leaf
                uint64_t x = next_leaf;
                int nl = 0;
                int m = 1;
                while (m && nl < WPB) {
                    lv[nl++] = x;
                    uint64_t y = x;
                    while (y != c && (y & 1) == 0) y = (y - 1) >> 1;  // up while a right child
                    if (y == c) {
                        m = 0;
                    } else {
                        y = y + 1;  // the right sibling, then down to its leftmost leaf
                        while (y < first_leaf) y = 2 * y + 1;
                        x = y;
                    }
                }
                next_leaf = x;
                n_batch = nl;
                more_s = m;
            }
            __syncthreads();
            const int nl = n_batch;
            more = more_s;
            if ((int)w < nl) {
                lg_f64x4 acc[NT];
                leaf_mfma_tiles<NB>(a, lv[w], lane, acc);
                double* sl = slot + (uint64_t)w * ts;
                leaf_mfma_put<NB>(acc, lane, a.R, [&](uint32_t e, double v) { sl[e] = v; });
                for (uint32_t e = T + lane; e < ts; e += 64) sl[e] = 0.0;
            }
            __syncthreads();
            for (int i = 0; i < nl; i++) {
                double* cur = slot + (uint64_t)i * ts;
                uint64_t node = lv[i];
                float* o = a.tree32 + (node - a.n64) * (uint64_t)ts32;
                for (uint32_t e = threadIdx.x; e < ts32; e += BS) o[e] = e < ts ? (float)cur[e] : 0.0f;
                while (node != c) {
                    double* pend = slot + (uint64_t)(WPB + node_depth(node) - s - 1) * ts;
                    if (node & 1) {  // a left child waits for its sibling
                        for (uint32_t e = threadIdx.x; e < ts; e += BS) pend[e] = cur[e];
                        break;
                    }
                    node = (node - 1) >> 1;
                    if (node == c) {
                        double* out = a.tree + c * (uint64_t)ts;
                        for (uint32_t e = threadIdx.x; e < ts; e += BS) out[e] = pend[e] + cur[e];
                    } else {
                        float* op = a.tree32 + (node - a.n64) * (uint64_t)ts32;
                        for (uint32_t e = threadIdx.x; e < ts32; e += BS) {
                            const double v = e < ts ? pend[e] + cur[e] : 0.0;
                            if (e < ts) cur[e] = v;
                            op[e] = (float)v;
                        }
                    }
                }
            }
            __syncthreads();  // leaf slots and the leaf list are rewritten next
        }
    }
}

// Compact FP32 trees, regular subtrees: node c at depth s (the lowest FP64
// level) whose leaves all lie Lc levels below it (Lc = L or L - 1; the one
// subtree that straddles a partial bottom level goes to k_subtree_build).  A
// block of 2^(L-1) waves: wave w takes the sibling leaves 2w and 2w + 1, each
// on the FP64 matrix cores as k_leaf_gram_mfma, and adds them (the tile
// layouts match entry for entry); then at each level above, the right one of
// two waves hands its node to the left one through its LDS slot, which adds
// it (left + right, as the level-wise build).  Every node passes through its
// wave's slot, from which the wave writes it in packed order (coalesced),
// FP32 -- or FP64 for c.  Dynamic LDS: a node per wave in the accumulator
// layout (NT x 4 x 64 doubles), then the (row, column) of each packed entry.
//
// At NB = 4 (R = 49-64) a slot in that layout is 20 KB, and eight waves
// (four FP32 levels) would not fit the LDS: there the slot holds the packed
// upper triangle instead (T entries, 16.6 KB at R = 64).  A lane then finds
// the packed index of each tile entry it holds (row 16p + 4v + lane / 16,
// column 16q + lane % 16) and moves only the entries on or above the
// diagonal and inside R; the others (the lower half of diagonal tiles, the
// zero padding past R) are never written out.
template <int NB>
constexpr bool pairs_packed_slot() {
    return NB == 4;
}

// Doubles of one wave's slot (packed: T rounded up to even, so every slot
// starts 16-byte aligned).  Packed slots are written out entry by entry and
// need no (row, column) table; at R = 49 eight of them take 78.5 KB, so two
// blocks fit a CU's LDS.
template <int NB>
__host__ __device__ inline uint64_t pairs_slot_doubles(uint64_t R) {
    return pairs_packed_slot<NB>() ? (tri_size(R) + 1) / 2 * 2 : (uint64_t)(NB * (NB + 1) / 2) * 256;
}

// Packed index of the entry of tile t (the t-th of the upper tile triangle)
// that `lane` holds in accumulator v, or -1 when it is below the diagonal
// or past R.
template <int NB>
__device__ inline int packed_tile_index(int t, int v, uint32_t lane, uint32_t R) {
    int p = 0, rem = t;
    while (rem >= NB - p) {
        rem -= NB - p;
        p++;
    }
    const int q = p + rem;
    const int r = 16 * p + 4 * v + (int)(lane >> 4), c = 16 * q + (int)(lane & 15);
    if (r > c || c >= (int)R) return -1;
    return r * (int)R - r * (r - 1) / 2 + (c - r);
}

template <int NB>
__device__ inline void tiles_to_slot(const lg_f64x4 (&acc)[NB * (NB + 1) / 2], double* slot, uint32_t lane,
                                     uint32_t R) {
    if constexpr (pairs_packed_slot<NB>()) {
        // the lane index is made opaque so that the 40 packed indices are
        // formed at each use instead of being held in registers
        uint32_t ln = lane;
        asm volatile("" : "+v"(ln));
#pragma unroll
        for (int t = 0; t < NB * (NB + 1) / 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int e = packed_tile_index<NB>(t, v, ln, R);
                if (e >= 0) slot[e] = acc[t][v];
            }
    } else {
#pragma unroll
        for (int t = 0; t < NB * (NB + 1) / 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) slot[(t * 4 + v) * 64 + lane] = acc[t][v];
    }
}

// The slot holds the left sibling and acc the right one: afterwards acc holds
// their sum (left + right) and the slot the right sibling, entry by entry
// (each entry a lane holds is its own in either layout, so no other lane
// reads it in between).
template <int NB>
__device__ inline void swap_add_slot(lg_f64x4 (&acc)[NB * (NB + 1) / 2], double* slot, uint32_t lane, uint32_t R) {
    if constexpr (pairs_packed_slot<NB>()) {
        uint32_t ln = lane;
        asm volatile("" : "+v"(ln));
#pragma unroll
        for (int t = 0; t < NB * (NB + 1) / 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int e = packed_tile_index<NB>(t, v, ln, R);
                if (e >= 0) {
                    const double right = acc[t][v];
                    acc[t][v] = slot[e] + right;
                    slot[e] = right;
                }
            }
    } else {
#pragma unroll
        for (int t = 0; t < NB * (NB + 1) / 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                double& d = slot[(t * 4 + v) * 64 + lane];
                const double right = acc[t][v];
                acc[t][v] = d + right;
                d = right;
            }
    }
}

// acc[t][v] = s[t][v] + the slot's entry (left + right, in that order: `slot_left`
// says which side the slot holds); entries the slot does not hold keep acc.
template <int NB>
__device__ inline void add_slot(lg_f64x4 (&acc)[NB * (NB + 1) / 2], const double* slot, bool slot_left,
                                uint32_t lane, uint32_t R) {
    if constexpr (pairs_packed_slot<NB>()) {
        uint32_t ln = lane;
        asm volatile("" : "+v"(ln));
#pragma unroll
        for (int t = 0; t < NB * (NB + 1) / 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int e = packed_tile_index<NB>(t, v, ln, R);
                if (e >= 0) acc[t][v] = slot_left ? slot[e] + acc[t][v] : acc[t][v] + slot[e];
            }
    } else {
#pragma unroll
        for (int t = 0; t < NB * (NB + 1) / 2; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double d = slot[(t * 4 + v) * 64 + lane];
                acc[t][v] = slot_left ? d + acc[t][v] : acc[t][v] + d;
            }
    }
}

// Entries [0, ts) of a node from its wave's slot (zero from T on), FP32 or FP64.
template <int NB, typename OutT>
__device__ inline void slot_write(const double* slot, const uint16_t* pairs, OutT* out, uint32_t T, uint32_t ts,
                                  uint32_t lane) {
    if constexpr (pairs_packed_slot<NB>()) {
        for (uint32_t e = lane; e < ts; e += 64) out[e] = (OutT)(e < T ? slot[e] : 0.0);
        return;
    }
    for (uint32_t e = lane; e < ts; e += 64) {
        double v = 0.0;
        if (e < T) {
            const uint32_t rc = pairs[e], r = rc & 0xffu, c = rc >> 8;
            const uint32_t p = r >> 4, q = c >> 4;
            const uint32_t t = p * NB - p * (p - 1) / 2 + (q - p);
            v = slot[(t * 4 + ((r & 15) >> 2)) * 64 + (r & 3) * 16 + (c & 15)];
        }
        out[e] = (OutT)v;
    }
}

template <int NB>
__global__ void __launch_bounds__(NB <= 2 ? 1024 : 512)
    k_subtree_pairs(BuildArgs a, uint64_t c0, uint64_t count, uint32_t L, uint64_t skip) {
    constexpr int NT = NB * (NB + 1) / 2;
    extern __shared__ double xch[];
    const uint32_t R = a.R, T = (uint32_t)tri_size(R);
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t waves = blockDim.x >> 6;
    const uint64_t slot_n = pairs_slot_doubles<NB>(R);
    uint16_t* pairs_rc = reinterpret_cast<uint16_t*>(xch + (uint64_t)waves * slot_n);
    if constexpr (!pairs_packed_slot<NB>()) {
        for (uint32_t r = threadIdx.x; r < R; r += blockDim.x) {
            const uint32_t base = r * R - r * (r - 1) / 2;
            for (uint32_t c = r; c < R; c++) pairs_rc[base + c - r] = (uint16_t)(r | (c << 8));
        }
    }
    __syncthreads();
    const uint64_t first_leaf = a.first_leaf, nodes = a.sh.node_count;
    double* mine = xch + (uint64_t)w * slot_n;
    auto emit = [&](uint64_t node, bool f64) {
        wave_sync_lds();
        if (f64) slot_write<NB>(mine, pairs_rc, a.tree + node * (uint64_t)a.ts64, T, a.ts64, lane);
        else slot_write<NB>(mine, pairs_rc, a.tree32 + (node - a.n64) * (uint64_t)a.ts32, T, a.ts32, lane);
        wave_sync_lds();
    };
    for (uint64_t b = blockIdx.x; b < count; b += gridDim.x) {
        const uint64_t c = c0 + b;
        if (c == skip) continue;
        if (c >= first_leaf) {  // a node of depth s that is a leaf: FP64
            if (w == 0) {
                lg_f64x4 acc[NT];
                leaf_mfma_tiles<NB>(a, c, lane, acc);
                tiles_to_slot<NB>(acc, mine, lane, R);
                emit(c, true);
            }
            continue;
        }
        // leaves L levels below c when its first descendant there exists
        const uint32_t Lc = (((c + 1) << L) - 1 < nodes) ? L : L - 1;
        const uint32_t npairs = 1u << (Lc - 1);
        lg_f64x4 A[NT];
        uint64_t nd = 0;
        if (w < npairs) {
            const uint64_t la = ((c + 1) << Lc) - 1 + 2 * (uint64_t)w;
            leaf_mfma_tiles<NB>(a, la, lane, A);
            tiles_to_slot<NB>(A, mine, lane, R);
            emit(la, false);
            leaf_mfma_tiles<NB>(a, la + 1, lane, A);
            // the parent in A, the right leaf in the slot: each lane swaps
            // the entries it holds itself
            swap_add_slot<NB>(A, mine, lane, R);
            emit(la + 1, false);
            tiles_to_slot<NB>(A, mine, lane, R);
            nd = (la - 1) >> 1;
            emit(nd, nd == c);
        }
        for (uint32_t k = 1; k < Lc; k++) {
            const uint32_t half = 1u << (k - 1);
            __syncthreads();  // the right children's slots hold their nodes
            if (w < npairs && (w & (2 * half - 1)) == 0)  // a left child: add its sibling
                add_slot<NB>(A, xch + (uint64_t)(w + half) * slot_n, false, lane, R);
            __syncthreads();  // the sibling slots are read before the left children rewrite theirs
            if (w < npairs && (w & (2 * half - 1)) == 0) {
                tiles_to_slot<NB>(A, mine, lane, R);
                nd = (nd - 1) >> 1;
                emit(nd, nd == c);
            }
        }
        __syncthreads();  // the slots are rewritten by the next subtree
    }
}

// Nodes [start, start + count) of one level: parent = left + right, a block
// per node.
__global__ void k_level_sum(double* tree, uint64_t start, uint64_t count, uint32_t ts64, float* tree32,
                            uint64_t n64, uint32_t ts32) {
    const uint32_t tsw = tree32 && ts32 > ts64 ? ts32 : ts64;
    for (uint64_t b = blockIdx.x; b < count; b += gridDim.x) {
        const uint64_t c = start + b;
        double* p = tree + c * ts64;
        const double* l = tree + (2 * c + 1) * ts64;
        const double* r = tree + (2 * c + 2) * ts64;
        for (uint32_t e = threadIdx.x; e < tsw; e += blockDim.x) {
            const double v = e < ts64 ? l[e] + r[e] : 0.0;
            if (e < ts64) p[e] = v;
            store_f32(tree32, n64, ts32, ts64, c, e, v);
        }
    }
}

// Three complete levels at once: a block takes the subtree of root
// c = root0 + b (b < count), reads its 8 great-grandchildren and writes the 4
// grandchildren, 2 children and the root, each the sum of its two children
// (the same additions as three single-level passes).
__global__ void k_level_sum3(double* tree, uint64_t root0, uint64_t count, uint32_t ts64, float* tree32,
                             uint64_t n64, uint32_t ts32) {
    const uint32_t tsw = tree32 && ts32 > ts64 ? ts32 : ts64;
    for (uint64_t b = blockIdx.x; b < count; b += gridDim.x)
    for (uint32_t e = threadIdx.x; e < tsw; e += blockDim.x) {
        const uint64_t c = root0 + b;
        const bool in64 = e < ts64;
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = in64 ? tree[(8 * c + 7 + k) * ts64 + e] : 0.0;
        double g[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            g[k] = v[2 * k] + v[2 * k + 1];
            if (in64) tree[(4 * c + 3 + k) * ts64 + e] = g[k];
            store_f32(tree32, n64, ts32, ts64, 4 * c + 3 + k, e, g[k]);
        }
        const double c0 = g[0] + g[1], c1 = g[2] + g[3];
        if (in64) {
            tree[(2 * c + 1) * ts64 + e] = c0;
            tree[(2 * c + 2) * ts64 + e] = c1;
            tree[c * ts64 + e] = c0 + c1;
        }
        store_f32(tree32, n64, ts32, ts64, 2 * c + 1, e, c0);
        store_f32(tree32, n64, ts32, ts64, 2 * c + 2, e, c1);
        store_f32(tree32, n64, ts32, ts64, c, e, c0 + c1);
    }
}

// FP32 copy of nodes [n64, n64 + count): a block per node.
__global__ void k_tree_to_f32(const double* tree64, float* tree32, uint64_t n64, uint64_t count, uint32_t ts64,
                              uint32_t ts32) {
    for (uint64_t b = blockIdx.x; b < count; b += gridDim.x) {
        const double* src = tree64 + (n64 + b) * ts64;
        float* dst = tree32 + b * ts32;
        for (uint32_t e = threadIdx.x; e < ts32; e += blockDim.x) dst[e] = e < ts64 ? (float)src[e] : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// computeM kernels
// ---------------------------------------------------------------------------
struct ComputeMArgs {
    const double* roots[kMaxModes];  // packed FP64 root Gram of each mode (node 0)
    int modes[kMaxModes];            // sampled modes in increasing order
    int m;                           // number of sampled modes
    int R;
    double* pinv;    // R x R full symmetric
    double* W;       // R x R, columns = eigvecs / sqrt(lambda), descending lambda
    double* E;       // N x R x R
    double* etree;   // N x etree_nodes x ts64
    double* work;    // per block: 3 (R+1) R doubles (V and eigenvectors when they do not fit LDS)
    double* lam_out; // (m+1) x R eigenvalues (diagnostics)
    uint32_t ts64;
    uint32_t etree_nodes;
};

__device__ inline double packed_get(const double* P, int R, int a, int b) {
    return a <= b ? P[tri_offset(R, a, b)] : P[tri_offset(R, b, a)];
}

// Packed-entry -> (row, col) table of the upper triangle (entry e < T).
__device__ inline void fill_pair_table(uint8_t* pa, uint8_t* pb, int R) {
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        const int base = r * R - r * (r - 1) / 2;
        for (int c = r; c < R; c++) {
            pa[base + c - r] = (uint8_t)r;
            pb[base + c - r] = (uint8_t)c;
        }
    }
}

template <int BS>
__device__ __forceinline__ void gram_pinv_body(const ComputeMArgs& a, double* smem, JacobiShared& js,
                                               double* diag_tmp) {
#pragma clang fp contract(off)
    const int R = a.R, R2 = R * R;
    const int LD = R + 1;
    const JacobiStore st = jacobi_store(smem, a.work, R);
    double* A = st.A;
    double* lam = a.lam_out;
    // G = Hadamard product of the sampled root Grams, multiplied in the
    // reference's suffix order G_{m-1} (.) ... (.) G_0 (efficient_krp_sampler.hpp:95-105);
    // each pass issues its loads for all entries of the thread at once.
    for (int e = threadIdx.x; e < R2; e += BS) A[(e / R) * LD + e % R] = 1.0;
    for (int t = a.m - 1; t >= 0; t--) {
        const double* G = a.roots[a.modes[t]];
        for (int e = threadIdx.x; e < R2; e += BS) {
            const int r = e / R, c = e % R;
            A[r * LD + c] = packed_get(G, R, r, c) * A[r * LD + c];
        }
    }
    __syncthreads();
    jacobi_block<BS>(A, st.V, R, lam, st.vecs, js, diag_tmp);
    const double* vecs = st.vecs;
    // W[u][R-1-v] = V[u][v] / sqrt(lam[v]) when lam[v] > 1e-8 (126-138); W is
    // staged in the (now free) V buffer and copied out for the last mode's E.
    double* W = st.V;
    for (int e = threadIdx.x; e < R2; e += BS) {
        const int u = e / R, v = e % R;
        const double l = lam[v];
        const double w = l > 1e-8 ? vecs[u * R + v] / sqrt(l) : 0.0;
        W[u * R + (R - 1 - v)] = w;
        a.W[u * R + (R - 1 - v)] = w;
    }
    __syncthreads();
    // pinv = W W^T (140-150), symmetric storage
    for (int e = threadIdx.x; e < R2; e += BS) {
        const int r = e / R, c = e % R;
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        double acc = 0.0;
        for (int k = 0; k < R; k++) acc = acc + W[lo * R + k] * W[hi * R + k];
        a.pinv[e] = acc;
    }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_gram_pinv(ComputeMArgs a) {
    extern __shared__ double smem[];
    __shared__ JacobiShared js;
    __shared__ double diag_tmp[128];
    gram_pinv_body<BS>(a, smem, js, diag_tmp);
}

// computeM part 1 of several draws (the same R): block b takes args[b].
template <int BS>
__global__ void __launch_bounds__(BS) k_gram_pinv_batched(const ComputeMArgs* __restrict__ args) {
    extern __shared__ double smem[];
    __shared__ JacobiShared js;
    __shared__ double diag_tmp[128];
    gram_pinv_body<BS>(args[blockIdx.x], smem, js, diag_tmp);
}

template <int BS>
__device__ __forceinline__ void eigen_modes_body(const ComputeMArgs& a, int t, double* smem, JacobiShared& js,
                                                 double* diag_tmp, uint8_t* pa, uint8_t* pb) {
#pragma clang fp contract(off)
    const int k = a.modes[t];  // t: position in the sampled-mode list
    const int R = a.R, R2 = R * R;
    const int LD = R + 1;
    const JacobiStore st = jacobi_store(smem, a.work + (uint64_t)(t + 1) * 3 * LD * R, R);
    double* A = st.A;
    double* E = a.E + (uint64_t)k * R2;
    double* El = st.V;  // E staged next to the Jacobi buffers
    fill_pair_table(pa, pb, R);
    if (t + 1 < a.m) {
        double* lam = a.lam_out + (uint64_t)(t + 1) * R;
        // P_k = (G_{modes[t+1]} (.) (... (.) G_{modes[m-1]})) (.) pinv
        for (int e = threadIdx.x; e < R2; e += BS) A[(e / R) * LD + e % R] = 1.0;
        for (int s = a.m - 1; s > t; s--) {
            const double* G = a.roots[a.modes[s]];
            for (int e = threadIdx.x; e < R2; e += BS) {
                const int r = e / R, c = e % R;
                A[r * LD + c] = packed_get(G, R, r, c) * A[r * LD + c];
            }
        }
        for (int e = threadIdx.x; e < R2; e += BS) {
            const int r = e / R, c = e % R;
            A[r * LD + c] = A[r * LD + c] * a.pinv[e];
        }
        __syncthreads();
        jacobi_block<BS>(A, st.V, R, lam, st.vecs, js, diag_tmp);
        const double* vecs = st.vecs;
        for (int e = threadIdx.x; e < R2; e += BS) {
            const int r = e / R, u = e % R;
            // sqrt(mu) clamped at 0 (the reference takes sqrt of a negative
            // eigenvalue and produces NaN, efficient_krp_sampler.hpp:176-180)
            const double mu = lam[r];
            const double v = vecs[u * R + r] * sqrt(mu > 0.0 ? mu : 0.0);
            El[r * R + u] = v;
            E[r * R + u] = v;
        }
    } else {
        for (int e = threadIdx.x; e < R2; e += BS) {
            const int r = e / R, u = e % R;
            const double v = a.W[u * R + r];
            El[r * R + u] = v;
            E[r * R + u] = v;
        }
    }
    __syncthreads();
    // Eigen tree over the R rows of E with F = 1, then every node (.) G_k
    // (187-193).  Leaves: node L-1+i holds row leaf_index(node).
    const Shape es = make_shape(R, 1);
    const uint32_t T = (uint32_t)tri_size(R), ts = a.ts64;
    double* tree = a.etree + (uint64_t)k * a.etree_nodes * ts;
    const uint32_t L = (uint32_t)es.leaf_count, first_leaf = (uint32_t)(es.node_count - es.leaf_count);
    for (uint32_t w = threadIdx.x; w < L * ts; w += BS) {
        const uint32_t i = w / ts, e = w - i * ts;
        const uint32_t node = first_leaf + i;
        const uint32_t row = (uint32_t)leaf_index(es, node);
        tree[node * ts + e] = e < T ? El[row * R + pa[e]] * El[row * R + pb[e]] : 0.0;
    }
    __syncthreads();
    int start = (int)es.nodes_before_lfill, end = (int)first_leaf;
    for (int level = (int)es.lfill_level; level >= 0; level--) {
        for (uint32_t w = threadIdx.x; w < (uint32_t)(end - start) * ts; w += BS) {
            const uint32_t cc = w / ts, e = w - cc * ts;
            const uint32_t c = start + cc;
            tree[c * ts + e] = tree[(2 * c + 1) * ts + e] + tree[(2 * c + 2) * ts + e];
        }
        __syncthreads();
        end = start;
        start = ((start + 1) / 2) - 1;
    }
    const double* Gk = a.roots[k];
    for (uint32_t w = threadIdx.x; w < (uint32_t)es.node_count * ts; w += BS) {
        const uint32_t node = w / ts, e = w - node * ts;
        if (e < T) tree[w] = tree[w] * Gk[e];
    }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_eigen_modes(ComputeMArgs a) {
    extern __shared__ double smem[];
    __shared__ JacobiShared js;
    __shared__ double diag_tmp[128];
    __shared__ uint8_t pa[8256], pb[8256];
    eigen_modes_body<BS>(a, (int)blockIdx.x, smem, js, diag_tmp, pa, pb);
}

// computeM part 2 of several draws (the same R): block b takes sampled-mode
// position blk_pos[b] of args[blk_req[b]].
template <int BS>
__global__ void __launch_bounds__(BS) k_eigen_modes_batched(const ComputeMArgs* __restrict__ args,
                                                            const uint32_t* __restrict__ blk_req,
                                                            const uint32_t* __restrict__ blk_pos) {
    extern __shared__ double smem[];
    __shared__ JacobiShared js;
    __shared__ double diag_tmp[128];
    __shared__ uint8_t pa[8256], pb[8256];
    eigen_modes_body<BS>(args[blk_req[blockIdx.x]], (int)blk_pos[blockIdx.x], smem, js, diag_tmp, pa, pb);
}

// ---------------------------------------------------------------------------
// Fused draw kernel
// ---------------------------------------------------------------------------
struct ModeDev {
    const double* tree64;  // nodes [0, n64), stride ts64
    const float* tree32;   // nodes [n64, node_count) at (c - n64) * ts32
    const double* U64;     // I x R
    const float* U32;      // I x R, FP32 mode only
    const double* etree;   // eigen tree, stride ts64
    const double* E;       // R x R
    Shape sh;
    uint64_t n64;          // first FP32 node (= node_count in FP64 mode)
    uint32_t d64;          // levels [0, d64) are FP64
    uint32_t T;            // packed entries per node, R (R + 1) / 2
};

struct DrawArgs {
    ModeDev mode[kMaxModes];
    int order[kMaxModes];  // sampled modes, increasing
    int m;                 // number of sampled modes
    int N;
    uint64_t J;
    int R;
    uint32_t ts64, ts32;
    Shape esh;             // eigen tree shape (R, 1)
    const double* pinv;    // R x R
    const double* draws;   // 2*N*J or null
    uint64_t seed, call;
    uint64_t* samples;     // N x J
    double* h_out;         // J x R or null
    double* w_out;         // J or null
    // sampled-mode positions [mi_begin, mi_end) of this launch; h carried
    // between launches in h_state (J x R) when the draw is split by mode
    int mi_begin, mi_end;
    double* h_state;
    // FP32 samplers: leaf scans over the FP32 factor copy (1) or the FP64
    // factors (0)
    int leaf32;
    // subtree-bucketed descent of one mode (k_draw_prefix / k_draw_suffix):
    // per-sample x (J x R) and {low, high, m, node} (J x 4) after the first
    // split_level levels, the level-split_level subtree of each sample and
    // the samples grouped by subtree (offsets: kBuckets + 1 entries)
    uint32_t split_level;
    double* st_x;
    double* st_s;
    uint32_t* bucket;
    uint32_t* blist;
    uint32_t* boff;
};

// Subtrees of the bucketed descent: one per XCD (8 on MI355X).  Workgroups
// are dispatched to the XCDs round-robin, so block b runs on XCD b % 8 and
// the suffix kernel gives bucket b % 8 to block b: each XCD's L2 then holds
// the nodes of one subtree only, eight times fewer than the whole tree.
constexpr uint32_t kBuckets = 8;
constexpr uint32_t kBucketLevel = 3;

template <int RMAX, int SG>
struct DrawCfg {
    static constexpr int TMAX = RMAX * (RMAX + 1) / 2;
    // a lane reads VEC consecutive packed entries per load: entries
    // VEC (g + SG i) + v, v < VEC, for i < NLD.  Measured at N=3, I=2^23,
    // R=32 (FP32 draw, 3 launches): VEC=2 0.940 ms, VEC=4 0.967 ms; single
    // entries were 10% slower than pairs.  The texture-data return path is
    // busy ~88% of the kernel's cycles (TD_TD_BUSY)
    static constexpr int VEC = 2;
    static constexpr int NLD = (TMAX + VEC * SG - 1) / (VEC * SG);  // loads per lane per node
    static constexpr int EPL = VEC * NLD;                            // packed entries per lane
    static constexpr int HPL = (RMAX + SG - 1) / SG;   // vector entries per lane
    static constexpr int RPL = (RMAX + SG - 1) / SG;   // leaf rows per lane (scan)
    // leaf dot products: LPR lanes per row, 8 columns per lane
    static constexpr int LCOLS = 8;
    static constexpr int LPR = RMAX / LCOLS;
    static constexpr int RPP = SG / LPR;                // rows per pass
};

// (row, column) of every packed upper-triangle entry e < T, as a | b << 8
// (R <= 128), filled once per block: make_coefs then needs two LDS reads per
// entry instead of walking the triangle.
template <int BS>
__device__ inline void fill_pairs(uint16_t* pt, int R) {
    for (int r = threadIdx.x; r < R; r += BS) {
        const int base = r * R - r * (r - 1) / 2;
        for (int c = r; c < R; c++) pt[base + c - r] = (uint16_t)(r | (c << 8));
    }
    __syncthreads();
}

// Coefficients of x^T G x over the packed entries this lane owns:
// c = x_a x_b (diagonal) or 2 x_a x_b (off-diagonal); c[VEC i + v] belongs to
// entry VEC (g + SG i) + v, zero past the packed size T.
template <int RMAX, int SG>
__device__ inline void make_coefs(const double* xs, const uint16_t* pt, int T, int g,
                                  double (&c)[DrawCfg<RMAX, SG>::EPL]) {
    using Cfg = DrawCfg<RMAX, SG>;
    // the table entries are read at every call rather than hoisted: held in
    // registers they would stay live through every descent
    int gg = g;
    asm volatile("" : "+v"(gg));
#pragma unroll
    for (int i = 0; i < Cfg::NLD; i++) {
#pragma unroll
        for (int v = 0; v < Cfg::VEC; v++) {
            const int e = Cfg::VEC * (gg + SG * i) + v;
            double val = 0.0;
            if (e < T) {
                const uint32_t ab = pt[e];
                const uint32_t ra = ab & 0xffu, rb = ab >> 8;
                const double p = xs[ra] * xs[rb];
                val = ra == rb ? p : 2.0 * p;
            }
            c[Cfg::VEC * i + v] = val;
        }
    }
}

// x^T G x over one packed node.  The EPL/2 paired loads per lane are
// unconditional and use one base address with immediate offsets, so all of
// them are in flight together.  Entries past the packed size T (the node
// stride is rounded to 16/32 entries, which SG * EPL can exceed) belong to the
// next node or to the zeroed tail pad every tree allocation carries; their
// coefficients are zero and the values are finite, so they add nothing.
// Nodes of FP32 levels are widened to FP64 before the multiply-add, so the
// only FP32 error is the storage rounding of each entry.
template <typename TN, int EPL, int SG>
__device__ inline double node_qform(const TN* __restrict__ node, const double (&c)[EPL], int g, uint32_t T) {
    // pairs of entries: one 8-byte (FP32) or 16-byte (FP64) load each; node
    // starts are multiples of 16 (FP64) / 32 (FP32) entries
    typedef TN V2 __attribute__((ext_vector_type(2)));
    constexpr int NLD = EPL / 2;
    V2 v[NLD];
    const V2* base = reinterpret_cast<const V2*>(node) + g;
#pragma unroll
    for (int i = 0; i + 1 < NLD; i++) v[i] = base[SG * i];
    // the last pair of a lane past the node's T entries re-reads its first
    // pair (same cache lines, coefficient zero) instead of the next node
    v[NLD - 1] = *(2u * (uint32_t)(g + SG * (NLD - 1)) < T ? base + SG * (NLD - 1) : base);
    double acc0 = 0, acc1 = 0;
#pragma unroll
    for (int i = 0; i < NLD; i++) {
        acc0 += c[2 * i] * (double)v[i].x;
        acc1 += c[2 * i + 1] * (double)v[i].y;
    }
    return group_sum<SG>(acc0 + acc1);
}

// Quadratic form of the left child `child` at depth `depth`: FP64 node for
// depth < d64 (or every level in FP64 mode), FP32 node below.
template <int EPL, int SG, bool USE32>
__device__ inline double child_qform(const ModeDev& md, const double* tree64, uint64_t child, uint32_t depth,
                                     uint32_t ts64, uint32_t ts32, const double (&c)[EPL], int g) {
    if (!USE32 || depth < md.d64) return node_qform<double, EPL, SG>(tree64 + child * ts64, c, g, md.T);
    return node_qform<float, EPL, SG>(md.tree32 + (child - md.n64) * ts32, c, g, md.T);
}

// One inverse-CDF descent (partition_tree.hpp:251-332) for the group's sample.
template <int RMAX, int SG, bool USE32>
__device__ inline uint64_t descend(const ModeDev& md, const Shape& sh, uint32_t ts64, uint32_t ts32,
                                   const double (&c64)[DrawCfg<RMAX, SG>::EPL], const double* tree64, double u,
                                   int g, double& low, double& high) {
    constexpr int EPL = DrawCfg<RMAX, SG>::EPL;
    const double m = node_qform<double, EPL, SG>(tree64, c64, g, md.T);
    uint64_t c = 0;
    low = 0.0;
    high = 1.0;
    for (uint32_t d = 0; d < sh.lfill_level; d++) {
        const uint64_t child = 2 * c + 1;
        const double mL = child_qform<EPL, SG, USE32>(md, tree64, child, d + 1, ts64, ts32, c64, g);
        const double cutoff = low + mL / m;
        if (u <= cutoff) {
            c = child;
            high = cutoff;
        } else {
            c = child + 1;
            low = cutoff;
        }
    }
    if (sh.node_count > sh.nodes_before_lfill) {
        const bool leaf = 2 * c + 1 >= sh.node_count;
        if (!leaf) {
            const uint64_t child = 2 * c + 1;
            const double mL =
                child_qform<EPL, SG, USE32>(md, tree64, child, (uint32_t)sh.lfill_level + 1, ts64, ts32, c64, g);
            const double cutoff = low + mL / m;
            if (u <= cutoff) {
                c = child;
                high = cutoff;
            } else {
                c = child + 1;
                low = cutoff;
            }
        }
    }
    return c;
}

// Descent steps d_begin .. d_end-1 (and the guarded tail level when tail)
// from the state (c, low, high) with root mass m: the loop of descend, split
// so that a descent can stop after the top levels and resume elsewhere.
template <int RMAX, int SG, bool USE32>
__device__ inline void descend_steps(const ModeDev& md, const Shape& sh, uint32_t ts64, uint32_t ts32,
                                     const double (&c64)[DrawCfg<RMAX, SG>::EPL], const double* tree64, double u,
                                     int g, double m, uint32_t d_begin, uint32_t d_end, bool tail, uint64_t& c,
                                     double& low, double& high) {
    constexpr int EPL = DrawCfg<RMAX, SG>::EPL;
    for (uint32_t d = d_begin; d < d_end; d++) {
        const uint64_t child = 2 * c + 1;
        const double mL = child_qform<EPL, SG, USE32>(md, tree64, child, d + 1, ts64, ts32, c64, g);
        const double cutoff = low + mL / m;
        if (u <= cutoff) {
            c = child;
            high = cutoff;
        } else {
            c = child + 1;
            low = cutoff;
        }
    }
    if (tail && sh.node_count > sh.nodes_before_lfill) {
        const bool leaf = 2 * c + 1 >= sh.node_count;
        if (!leaf) {
            const uint64_t child = 2 * c + 1;
            const double mL =
                child_qform<EPL, SG, USE32>(md, tree64, child, (uint32_t)sh.lfill_level + 1, ts64, ts32, c64, g);
            const double cutoff = low + mL / m;
            if (u <= cutoff) {
                c = child;
                high = cutoff;
            } else {
                c = child + 1;
                low = cutoff;
            }
        }
    }
}

// Leaf row scan (partition_tree.hpp:336-391): q_t = (U_t . x)^2 for the rows
// of the leaf, inclusive prefix over the group, P_{t+1} = prefix_t / total,
// result = first t <= F-2 with frac < P_{t+1}, else F-1 (clamped to the last
// real row of a partial leaf).
// The dot products take LPR lanes per row, 8 consecutive columns per lane
// (16-byte loads when R == RMAX, else element loads with the column clamped,
// x being zero past R), summed over the LPR lanes with the DPP butterfly, and
// RPP = SG / LPR rows per pass; q goes to LDS (qs, RMAX entries per group)
// and lane g then owns rows g, g+SG, ... for the scan.  Few registers per
// lane: the kernel's occupancy, not its bandwidth, bounds the descents.
template <typename TU, int RMAX, int SG>
__device__ inline uint32_t leaf_scan(const TU* __restrict__ U, uint64_t r0, uint32_t rows, uint32_t F,
                                     int R, const double* xs, double* qs, double frac, int g, int lane) {
    using Cfg = DrawCfg<RMAX, SG>;
    constexpr int LC = Cfg::LCOLS, LPR = Cfg::LPR, RPP = Cfg::RPP, RPL = Cfg::RPL;
    constexpr int W = 16 / sizeof(TU);
    typedef TU V __attribute__((ext_vector_type(W)));
    const int sub = g % LPR, rsub = g / LPR;
    const int col0 = sub * LC;
    const uint32_t passes = (F + RPP - 1) / RPP;
    // the loads of up to PCH passes are in flight together
    constexpr int PMAX = (RMAX + RPP - 1) / RPP;
    constexpr int PCH = PMAX < 4 ? PMAX : 4;
    for (uint32_t p0 = 0; p0 < passes; p0 += PCH) {
        TU buf[PCH][LC];
#pragma unroll
        for (int pp = 0; pp < PCH; pp++) {
            const uint32_t t = (p0 + pp) * RPP + rsub;
            const uint32_t tc = t < rows ? t : rows - 1;
            const TU* row = U + (r0 + tc) * (uint64_t)R;
            if (R == RMAX) {
                const V* rv = reinterpret_cast<const V*>(row + col0);
#pragma unroll
                for (int q = 0; q < LC / W; q++) {
                    const V v = rv[q];
#pragma unroll
                    for (int w = 0; w < W; w++) buf[pp][q * W + w] = v[w];
                }
            } else {
                // formed per pass: hoisted, the eight clamped offsets would
                // hold registers through every descent of the kernel
                int cb = col0;
                asm volatile("" : "+v"(cb));
#pragma unroll
                for (int w = 0; w < LC; w++) {
                    const int b = cb + w;
                    buf[pp][w] = row[b < R ? b : R - 1];
                }
            }
        }
#pragma unroll
        for (int pp = 0; pp < PCH; pp++) {
            double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
            for (int w = 0; w < LC; w += 2) {
                acc0 += (double)buf[pp][w] * xs[col0 + w];
                acc1 += (double)buf[pp][w + 1] * xs[col0 + w + 1];
            }
            const double d = group_sum<LPR>(acc0 + acc1);
            const uint32_t t = (p0 + pp) * RPP + rsub;
            if (sub == 0 && t < F) qs[t] = t < rows ? d * d : 0.0;
        }
    }
    wave_sync_lds();
    double incl[RPL];
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        if (SG * k < (int)F) {
            const uint32_t t = g + SG * k;
            const double q2 = t < F ? qs[t] : 0.0;
            const double v = group_incl_scan<SG>(q2, g) + run;
            incl[k] = v;
            run = __shfl(v, SG - 1, SG);
        } else {
            incl[k] = run;
        }
    }
    const double total = run;
    uint32_t res = F - 1;
    bool found = false;
    const int gbase = lane - g;
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        const uint32_t t = g + SG * k;
        const bool pred = (t + 1 < F) && (frac < incl[k] / total);
        const uint64_t bal = __ballot(pred);
        const uint64_t bits = (bal >> gbase) & (SG == 64 ? ~0ull : ((1ull << SG) - 1));
        if (!found && bits) {
            res = SG * k + __builtin_ctzll(bits);
            found = true;
        }
    }
    // clamp for a partial last leaf (partition_tree.hpp:353, 373-391)
    if (res >= rows) res = rows - 1;
    return res;
}

// One wave item of the fused draw: the GPW = 64 / SG samples
// item * GPW + (lane / SG) of the draw described by `a` (every sampled mode
// in [a.mi_begin, a.mi_end), then the weight when the launch ends the draw).
// xs / qs: the group's x and leaf-scan buffers in LDS; pair_tab: the packed
// (row, column) table of a.R.
template <int RMAX, int SG, bool USE32>
__device__ __forceinline__ void krp_draw_item(const DrawArgs& a, uint64_t item, double* xs, double* qs,
                                              const uint16_t* pair_tab, int lane, int g) {
    using Cfg = DrawCfg<RMAX, SG>;
    constexpr int EPL = Cfg::EPL, HPL = Cfg::HPL;
    constexpr int GPW = 64 / SG;  // groups per wave
    const int R = a.R;
    const int T = (int)tri_size(R);
    constexpr bool fp32 = USE32;
    const uint64_t i = item * GPW + lane / SG;
    const bool active = i < a.J;
    double h[HPL];
#pragma unroll
    for (int k = 0; k < HPL; k++) h[k] = 1.0;
    if (a.mi_begin > 0 && active) {
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) h[q] = a.h_state[i * R + g + SG * q];
    }

    for (int mi = a.mi_begin; mi < a.mi_end; mi++) {
        const int k = a.order[mi];
        const ModeDev& md = a.mode[k];
        // ---- eigen stage: x = h -----------------------------------------
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) xs[g + SG * q] = h[q];
        wave_sync_lds();
        double c64[EPL];
        make_coefs<RMAX, SG>(xs, pair_tab, T, g, c64);
        double ue = active ? (a.draws ? a.draws[(uint64_t)k * a.J + i] : philox_uniform(a.seed, a.call, i, 2 * k))
                           : 0.5;
        double low, high;
        const uint64_t ce = descend<RMAX, SG, false>(md, a.esh, a.ts64, a.ts32, c64, md.etree, ue, g, low, high);
        const uint64_t r = leaf_index(a.esh, ce);
        wave_sync_lds();
        // scaled_h = h (.) E_k[r]
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) xs[g + SG * q] = h[q] * md.E[r * R + g + SG * q];
        wave_sync_lds();
        // ---- Gram stage: x = scaled_h -------------------------------------
        make_coefs<RMAX, SG>(xs, pair_tab, T, g, c64);
        const double ug = active ? (a.draws ? a.draws[(uint64_t)a.N * a.J + (uint64_t)k * a.J + i]
                                            : philox_uniform(a.seed, a.call, i, 2 * k + 1))
                                 : 0.5;
        const uint64_t c = descend<RMAX, SG, fp32>(md, md.sh, a.ts64, a.ts32, c64, md.tree64, ug, g, low, high);
        const uint64_t idx = leaf_index(md.sh, c);
        uint32_t res = 0;
        if (md.sh.F > 1) {
            const double frac = (ug - low) / (high - low);
            const uint64_t r0 = idx * md.sh.F;
            const uint32_t rows = (uint32_t)min<uint64_t>(md.sh.F, md.sh.n - r0);
            if (fp32 && a.leaf32)
                res = leaf_scan<float, RMAX, SG>(md.U32, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac, g, lane);
            else
                res = leaf_scan<double, RMAX, SG>(md.U64, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac, g, lane);
        }
        const uint64_t srow = idx * md.sh.F + res;
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) h[q] *= md.U64[srow * R + g + SG * q];
        if (active && g == 0) a.samples[(uint64_t)k * a.J + i] = srow;
        wave_sync_lds();
    }
    if (a.mi_end < a.m) {  // the next launch continues from h
        if (active) {
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (g + SG * q < R) a.h_state[i * R + g + SG * q] = h[q];
        }
        return;
    }
    // ---- weight: R / (J * h^T pinv h) (sampler.hpp:94-122) --------------
#pragma unroll
    for (int q = 0; q < HPL; q++)
        if (g + SG * q < R) xs[g + SG * q] = h[q];
    wave_sync_lds();
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < HPL; q++) {
        const int ra = g + SG * q;
        if (ra < R) {
            double y = 0.0;
            for (int b = 0; b < R; b++) y += a.pinv[ra * R + b] * xs[b];
            part += h[q] * y;
        }
    }
    const double lev = group_sum<SG>(part);
    if (active) {
        if (a.w_out && g == 0) a.w_out[i] = (double)R / (lev * (double)a.J);
        if (a.h_out) {
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (g + SG * q < R) a.h_out[i * R + g + SG * q] = h[q];
        }
    }
    wave_sync_lds();
}

template <int RMAX, int SG, int BS, bool USE32>
__global__ void __launch_bounds__(BS) k_krp_draw(DrawArgs a) {
    using Cfg = DrawCfg<RMAX, SG>;
    constexpr int GPW = 64 / SG;  // groups per wave
    __shared__ double lds_x[BS / SG][RMAX];
    __shared__ double lds_q[BS / SG][RMAX];
    __shared__ uint16_t pair_tab[Cfg::TMAX];
    fill_pairs<BS>(pair_tab, a.R);
    const int lane = threadIdx.x & 63;
    const int g = lane % SG;
    const int wave = threadIdx.x / 64;
    const int grp_in_block = threadIdx.x / SG;
    double* xs = lds_x[grp_in_block];
    double* qs = lds_q[grp_in_block];
    // entries R..RMAX-1 of x stay zero (row_dot's generic path reads them)
    for (int b = g; b < RMAX; b += SG) xs[b] = 0.0;
    wave_sync_lds();
    const uint64_t waves_total = (uint64_t)gridDim.x * (BS / 64);
    const uint64_t wave_id = (uint64_t)blockIdx.x * (BS / 64) + wave;
    const uint64_t n_items = (a.J + GPW - 1) / GPW;
    for (uint64_t item = wave_id; item < n_items; item += waves_total)
        krp_draw_item<RMAX, SG, USE32>(a, item, xs, qs, pair_tab, lane, g);
}

// Packed draw of several requests (the same R and precision) in one launch:
// request r owns the wave items [item_off[r], item_off[r + 1]) (one item per
// GPW of its samples); a wave takes items of the whole launch in turn and
// draws them with its request's arguments (fused: every sampled mode, then
// the weight).  Each request's samples are those of its own k_krp_draw.
template <int RMAX, int SG, int BS, bool USE32>
__global__ void __launch_bounds__(BS) k_krp_draw_packed(const DrawArgs* __restrict__ reqs,
                                                        const uint64_t* __restrict__ item_off, uint32_t n_req,
                                                        int R) {
    using Cfg = DrawCfg<RMAX, SG>;
    __shared__ double lds_x[BS / SG][RMAX];
    __shared__ double lds_q[BS / SG][RMAX];
    __shared__ uint16_t pair_tab[Cfg::TMAX];
    fill_pairs<BS>(pair_tab, R);
    const int lane = threadIdx.x & 63;
    const int g = lane % SG;
    const int wave = threadIdx.x / 64;
    const int grp_in_block = threadIdx.x / SG;
    double* xs = lds_x[grp_in_block];
    double* qs = lds_q[grp_in_block];
    for (int b = g; b < RMAX; b += SG) xs[b] = 0.0;
    wave_sync_lds();
    const uint64_t waves_total = (uint64_t)gridDim.x * (BS / 64);
    const uint64_t wave_id = (uint64_t)blockIdx.x * (BS / 64) + wave;
    const uint64_t n_items = item_off[n_req];
    for (uint64_t item = wave_id; item < n_items; item += waves_total) {
        // request of this item: the last r with item_off[r] <= item
        uint32_t lo = 0, hi = n_req;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (item_off[mid] <= item) lo = mid;
            else hi = mid;
        }
        const uint32_t r = (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
        krp_draw_item<RMAX, SG, USE32>(reqs[r], item - item_off[r], xs, qs, pair_tab, lane, g);
    }
}

// ---------------------------------------------------------------------------
// One-sample-per-wave draw kernel.  The whole wave (64 lanes) works on one
// sample, so every per-sample quantity (cutoffs, low/high, node index, the
// uniform) is wave-uniform and lives in scalar registers, the descent branch
// is a scalar branch, and each lane holds only ceil(T/64) coefficients and
// node values.  The lower register count admits more waves per SIMD, i.e.
// more descents in flight: the descents are chains of dependent node reads,
// bounded by how many are in flight rather than by bandwidth.  Same
// arithmetic and comparisons as k_krp_draw.
// ---------------------------------------------------------------------------
__device__ inline double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int RMAX, bool USE32>
__device__ inline uint64_t descend_wave(const ModeDev& md, const Shape& sh, uint32_t ts64, uint32_t ts32,
                                        const double (&c64)[DrawCfg<RMAX, 64>::EPL], const double* tree64, double u,
                                        int g, double& low, double& high) {
    constexpr int EPL = DrawCfg<RMAX, 64>::EPL;
    const double m = wave_uniform(node_qform<double, EPL, 64>(tree64, c64, g, md.T));
    uint64_t c = 0;
    double lo = 0.0, hi = 1.0;
    for (uint32_t d = 0; d < sh.lfill_level; d++) {
        const uint64_t child = 2 * c + 1;
        const double mL = wave_uniform(child_qform<EPL, 64, USE32>(md, tree64, child, d + 1, ts64, ts32, c64, g));
        const double cutoff = lo + mL / m;
        if (u <= cutoff) {
            c = child;
            hi = cutoff;
        } else {
            c = child + 1;
            lo = cutoff;
        }
    }
    if (sh.node_count > sh.nodes_before_lfill) {
        const bool leaf = 2 * c + 1 >= sh.node_count;
        if (!leaf) {
            const uint64_t child = 2 * c + 1;
            const double mL = wave_uniform(
                child_qform<EPL, 64, USE32>(md, tree64, child, (uint32_t)sh.lfill_level + 1, ts64, ts32, c64, g));
            const double cutoff = lo + mL / m;
            if (u <= cutoff) {
                c = child;
                hi = cutoff;
            } else {
                c = child + 1;
                lo = cutoff;
            }
        }
    }
    low = lo;
    high = hi;
    return c;
}

template <int RMAX, int BS, bool USE32>
__global__ void __launch_bounds__(BS) k_krp_draw_wave(DrawArgs a) {
    using Cfg = DrawCfg<RMAX, 64>;
    constexpr int EPL = Cfg::EPL, HPL = Cfg::HPL;
    constexpr int WPB = BS / 64;
    __shared__ double lds_x[WPB][RMAX];
    __shared__ double lds_q[WPB][RMAX];
    __shared__ uint16_t pair_tab[Cfg::TMAX];
    fill_pairs<BS>(pair_tab, a.R);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double* xs = lds_x[wave];
    double* qs = lds_q[wave];
    const int R = a.R;
    const int T = (int)tri_size(R);
    for (int b = lane; b < RMAX; b += 64) xs[b] = 0.0;
    wave_sync_lds();
    const uint64_t waves_total = (uint64_t)gridDim.x * WPB;
    for (uint64_t i = (uint64_t)blockIdx.x * WPB + wave; i < a.J; i += waves_total) {
        double h[HPL];
#pragma unroll
        for (int q = 0; q < HPL; q++) h[q] = 1.0;
        if (a.mi_begin > 0) {
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (lane + 64 * q < R) h[q] = a.h_state[i * R + lane + 64 * q];
        }
        for (int mi = a.mi_begin; mi < a.mi_end; mi++) {
            const int k = a.order[mi];
            const ModeDev& md = a.mode[k];
            // ---- eigen stage: x = h ---------------------------------------
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (lane + 64 * q < R) xs[lane + 64 * q] = h[q];
            wave_sync_lds();
            double c64[EPL];
            make_coefs<RMAX, 64>(xs, pair_tab, T, lane, c64);
            const double ue = wave_uniform(a.draws ? a.draws[(uint64_t)k * a.J + i]
                                                   : philox_uniform(a.seed, a.call, i, 2 * k));
            double low, high;
            const uint64_t ce = descend_wave<RMAX, false>(md, a.esh, a.ts64, a.ts32, c64, md.etree, ue, lane, low, high);
            const uint64_t r = leaf_index(a.esh, ce);
            wave_sync_lds();
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (lane + 64 * q < R) xs[lane + 64 * q] = h[q] * md.E[r * R + lane + 64 * q];
            wave_sync_lds();
            // ---- Gram stage: x = scaled_h -----------------------------------
            make_coefs<RMAX, 64>(xs, pair_tab, T, lane, c64);
            const double ug = wave_uniform(a.draws ? a.draws[(uint64_t)a.N * a.J + (uint64_t)k * a.J + i]
                                                   : philox_uniform(a.seed, a.call, i, 2 * k + 1));
            const uint64_t c = descend_wave<RMAX, USE32>(md, md.sh, a.ts64, a.ts32, c64, md.tree64, ug, lane, low, high);
            const uint64_t idx = leaf_index(md.sh, c);
            uint32_t res = 0;
            if (md.sh.F > 1) {
                const double frac = (ug - low) / (high - low);
                const uint64_t r0 = idx * md.sh.F;
                const uint32_t rows = (uint32_t)min<uint64_t>(md.sh.F, md.sh.n - r0);
                if constexpr (USE32)
                    res = leaf_scan<float, RMAX, 64>(md.U32, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac, lane, lane);
                else
                    res = leaf_scan<double, RMAX, 64>(md.U64, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac, lane,
                                                      lane);
            }
            const uint64_t srow = idx * md.sh.F + res;
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (lane + 64 * q < R) h[q] *= md.U64[srow * R + lane + 64 * q];
            if (lane == 0) a.samples[(uint64_t)k * a.J + i] = srow;
            wave_sync_lds();
        }
        if (a.mi_end < a.m) {
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (lane + 64 * q < R) a.h_state[i * R + lane + 64 * q] = h[q];
            continue;
        }
        // ---- weight: R / (J * h^T pinv h) (sampler.hpp:94-122) --------------
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (lane + 64 * q < R) xs[lane + 64 * q] = h[q];
        wave_sync_lds();
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < HPL; q++) {
            const int ra = lane + 64 * q;
            if (ra < R) {
                double y = 0.0;
                for (int b = 0; b < R; b++) y += a.pinv[ra * R + b] * xs[b];
                part += h[q] * y;
            }
        }
        const double lev = group_sum<64>(part);
        if (a.w_out && lane == 0) a.w_out[i] = (double)R / (lev * (double)a.J);
        if (a.h_out) {
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (lane + 64 * q < R) a.h_out[i * R + lane + 64 * q] = h[q];
        }
        wave_sync_lds();
    }
}

// ---------------------------------------------------------------------------
// One-sample-per-block draw kernel for large ranks (R = 64, 128).  A packed
// R = 128 node is 8,256 doubles: split over one group of 64 lanes it needs
// 129 coefficients and 129 node values per lane, more than the register file
// holds.  Here a block of BS threads (4-8 waves) works on one sample: each
// thread owns the packed entries tid, tid + BS, ... (ceil(T/BS) of them),
// a level sums the per-thread products with the DPP wave sum and one LDS
// exchange between the waves (one barrier per level; the exchange slots
// alternate between two sets so a level's writes never meet the previous
// level's reads), and every thread then holds the same node mass, so the
// descent branch is uniform over the block.  The leaf scan computes the F
// row dot products with the whole block and scans them in every wave.
// Same arithmetic and comparisons as k_krp_draw up to the order of the
// partial sums.
// ---------------------------------------------------------------------------
template <int BS>
__device__ inline double block_sum(double v, double* red, int& par) {
    constexpr int NW = BS / 64;
    v = group_sum<64>(v);
    if ((threadIdx.x & 63) == 0) red[par * NW + (threadIdx.x >> 6)] = v;
    __syncthreads();
    double t = red[par * NW];
#pragma unroll
    for (int w = 1; w < NW; w++) t += red[par * NW + w];
    par ^= 1;
    return t;
}

template <typename TN, int EPL, int BS>
__device__ inline double node_qform_block(const TN* __restrict__ node, const double (&c)[EPL], double* red, int& par,
                                          uint32_t T) {
    typedef TN V2 __attribute__((ext_vector_type(2)));
    constexpr int NLD = EPL / 2;
    V2 v[NLD];
    const V2* base = reinterpret_cast<const V2*>(node) + threadIdx.x;
#pragma unroll
    for (int i = 0; i + 1 < NLD; i++) v[i] = base[BS * i];
    v[NLD - 1] = *(2u * (uint32_t)(threadIdx.x + BS * (NLD - 1)) < T ? base + BS * (NLD - 1) : base);
    double acc0 = 0, acc1 = 0;
#pragma unroll
    for (int i = 0; i < NLD; i++) {
        acc0 += c[2 * i] * (double)v[i].x;
        acc1 += c[2 * i + 1] * (double)v[i].y;
    }
    return block_sum<BS>(acc0 + acc1, red, par);
}

template <int EPL, int BS, bool USE32>
__device__ inline double child_qform_block(const ModeDev& md, const double* tree64, uint64_t child, uint32_t depth,
                                           uint32_t ts64, uint32_t ts32, const double (&c)[EPL], double* red,
                                           int& par) {
    if (!USE32 || depth < md.d64) return node_qform_block<double, EPL, BS>(tree64 + child * ts64, c, red, par, md.T);
    return node_qform_block<float, EPL, BS>(md.tree32 + (child - md.n64) * ts32, c, red, par, md.T);
}

// The inverse-CDF descent of descend (partition_tree.hpp:251-332) with
// block-wide node masses.
template <int EPL, int BS, bool USE32>
__device__ inline uint64_t descend_block(const ModeDev& md, const Shape& sh, uint32_t ts64, uint32_t ts32,
                                         const double (&c64)[EPL], const double* tree64, double u, double* red,
                                         int& par, double& low, double& high) {
    const double m = node_qform_block<double, EPL, BS>(tree64, c64, red, par, md.T);
    uint64_t c = 0;
    low = 0.0;
    high = 1.0;
    const uint32_t levels = sh.lfill_level + ((sh.node_count > sh.nodes_before_lfill) ? 1u : 0u);
    for (uint32_t d = 0; d < levels; d++) {
        // the guarded tail level (partition_tree.hpp:304-332): a leaf on the
        // lfill level stops the descent
        if (d == sh.lfill_level && 2 * c + 1 >= sh.node_count) break;
        const uint64_t child = 2 * c + 1;
        const double mL = child_qform_block<EPL, BS, USE32>(md, tree64, child, d + 1, ts64, ts32, c64, red, par);
        const double cutoff = low + mL / m;
        if (u <= cutoff) {
            c = child;
            high = cutoff;
        } else {
            c = child + 1;
            low = cutoff;
        }
    }
    return c;
}

// Leaf scan of one sample with the whole block (partition_tree.hpp:336-391):
// LPR = RMAX/8 threads per row, 8 columns each, RPP = BS/LPR rows per pass;
// loads for PCH passes are in flight together.  q_t = (U_t . x)^2 goes to
// LDS, then every wave forms the same inclusive scan over the F values and
// finds the first t < F-1 with frac < P[t+1] (else F-1).
template <typename TU, int RMAX, int BS>
__device__ inline uint32_t leaf_scan_block(const TU* __restrict__ U, uint64_t r0, uint32_t rows, uint32_t F, int R,
                                           const double* xs, double* qs, double frac) {
    constexpr int LC = 8, LPR = RMAX / LC, RPP = BS / LPR;
    constexpr int PMAX = (RMAX + RPP - 1) / RPP;
    constexpr int PCH = PMAX < 4 ? PMAX : 4;
    constexpr int W = 16 / sizeof(TU);
    typedef TU V __attribute__((ext_vector_type(W)));
    const int tid = threadIdx.x;
    const int sub = tid % LPR, rsub = tid / LPR;
    const int col0 = sub * LC;
    const uint32_t passes = (F + RPP - 1) / RPP;
    for (uint32_t p0 = 0; p0 < passes; p0 += PCH) {
        TU buf[PCH][LC];
#pragma unroll
        for (int pp = 0; pp < PCH; pp++) {
            const uint32_t t = (p0 + pp) * RPP + rsub;
            const uint32_t tc = t < rows ? t : rows - 1;
            const TU* row = U + (r0 + tc) * (uint64_t)R;
            if (R == RMAX) {
                const V* rv = reinterpret_cast<const V*>(row + col0);
#pragma unroll
                for (int q = 0; q < LC / W; q++) {
                    const V v = rv[q];
#pragma unroll
                    for (int w = 0; w < W; w++) buf[pp][q * W + w] = v[w];
                }
            } else {
#pragma unroll
                for (int w = 0; w < LC; w++) {
                    const int b = col0 + w;
                    buf[pp][w] = row[b < R ? b : R - 1];
                }
            }
        }
#pragma unroll
        for (int pp = 0; pp < PCH; pp++) {
            double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
            for (int w = 0; w < LC; w += 2) {
                acc0 += (double)buf[pp][w] * xs[col0 + w];
                acc1 += (double)buf[pp][w + 1] * xs[col0 + w + 1];
            }
            const double d = group_sum<LPR>(acc0 + acc1);
            const uint32_t t = (p0 + pp) * RPP + rsub;
            if (sub == 0 && t < F) qs[t] = t < rows ? d * d : 0.0;
        }
    }
    __syncthreads();
    constexpr int RPL = (RMAX + 63) / 64;
    const int lane = tid & 63;
    double incl[RPL];
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        if (64 * k < (int)F) {
            const uint32_t t = lane + 64 * k;
            const double q2 = t < F ? qs[t] : 0.0;
            const double v = group_incl_scan<64>(q2, lane) + run;
            incl[k] = v;
            run = __shfl(v, 63, 64);
        } else {
            incl[k] = run;
        }
    }
    const double total = run;
    uint32_t res = F - 1;
    bool found = false;
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        const uint32_t t = lane + 64 * k;
        const bool pred = (t + 1 < F) && (frac < incl[k] / total);
        const uint64_t bal = __ballot(pred);
        if (!found && bal) {
            res = 64 * k + __builtin_ctzll(bal);
            found = true;
        }
    }
    if (res >= rows) res = rows - 1;
    return res;
}

template <int RMAX, int BS, bool USE32>
__global__ void __launch_bounds__(BS) k_krp_draw_block(DrawArgs a) {
    static_assert(BS >= RMAX && BS % 64 == 0, "one h entry per thread");
    using Cfg = DrawCfg<RMAX, BS>;
    constexpr int EPL = Cfg::EPL;
    constexpr int NW = BS / 64;
    __shared__ double xs[RMAX];
    __shared__ double qs[RMAX];
    __shared__ double red[2 * NW];
    __shared__ uint16_t pair_tab[Cfg::TMAX];
    fill_pairs<BS>(pair_tab, a.R);
    const int tid = threadIdx.x;
    const int R = a.R;
    const int T = (int)tri_size(R);
    int par = 0;
    // entries R..RMAX-1 of x stay zero (the leaf scan's generic path reads them)
    for (int b = tid; b < RMAX; b += BS) xs[b] = 0.0;
    __syncthreads();

    for (uint64_t i = blockIdx.x; i < a.J; i += gridDim.x) {
        double h = 1.0;  // thread tid < R owns h[tid]
        if (a.mi_begin > 0 && tid < R) h = a.h_state[i * R + tid];
        for (int mi = a.mi_begin; mi < a.mi_end; mi++) {
            const int k = a.order[mi];
            const ModeDev& md = a.mode[k];
            // ---- eigen stage: x = h -----------------------------------------
            if (tid < R) xs[tid] = h;
            __syncthreads();
            double c64[EPL];
            make_coefs<RMAX, BS>(xs, pair_tab, T, tid, c64);
            const double ue = a.draws ? a.draws[(uint64_t)k * a.J + i] : philox_uniform(a.seed, a.call, i, 2 * k);
            double low, high;
            const uint64_t ce = descend_block<EPL, BS, false>(md, a.esh, a.ts64, a.ts32, c64, md.etree, ue, red, par,
                                                              low, high);
            const uint64_t r = leaf_index(a.esh, ce);
            // every thread has passed the descent's barriers after its
            // make_coefs reads of xs
            if (tid < R) xs[tid] = h * md.E[r * R + tid];
            __syncthreads();
            // ---- Gram stage: x = scaled_h -------------------------------------
            make_coefs<RMAX, BS>(xs, pair_tab, T, tid, c64);
            const double ug = a.draws ? a.draws[(uint64_t)a.N * a.J + (uint64_t)k * a.J + i]
                                      : philox_uniform(a.seed, a.call, i, 2 * k + 1);
            const uint64_t c = descend_block<EPL, BS, USE32>(md, md.sh, a.ts64, a.ts32, c64, md.tree64, ug, red, par,
                                                             low, high);
            const uint64_t idx = leaf_index(md.sh, c);
            uint32_t res = 0;
            if (md.sh.F > 1) {
                const double frac = (ug - low) / (high - low);
                const uint64_t r0 = idx * md.sh.F;
                const uint32_t rows = (uint32_t)min<uint64_t>(md.sh.F, md.sh.n - r0);
                if constexpr (USE32)
                    res = leaf_scan_block<float, RMAX, BS>(md.U32, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac);
                else
                    res = leaf_scan_block<double, RMAX, BS>(md.U64, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac);
            }
            const uint64_t srow = idx * md.sh.F + res;
            if (tid < R) h *= md.U64[srow * R + tid];
            if (tid == 0) a.samples[(uint64_t)k * a.J + i] = srow;
            __syncthreads();  // xs / qs reads done before the next stage writes them
        }
        if (a.mi_end < a.m) {  // the next launch continues from h
            if (tid < R) a.h_state[i * R + tid] = h;
            continue;
        }
        // ---- weight: R / (J * h^T pinv h) (sampler.hpp:94-122), over the
        // packed triangle with the coefficients of x = h ----------------------
        if (tid < R) xs[tid] = h;
        __syncthreads();
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            const int e = tid + BS * q;
            if (e < T) {
                const uint32_t ab = pair_tab[e];
                const uint32_t ra = ab & 0xffu, rb = ab >> 8;
                const double p = xs[ra] * xs[rb] * a.pinv[ra * R + rb];
                part += ra == rb ? p : 2.0 * p;
            }
        }
        const double lev = block_sum<BS>(part, red, par);
        if (tid == 0 && a.w_out) a.w_out[i] = (double)R / (lev * (double)a.J);
        if (a.h_out && tid < R) a.h_out[i * R + tid] = h;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Subtree-bucketed draw of one sampled mode (positions mi_begin, one launch
// each): k_draw_prefix runs the eigen stage and the first split_level Gram
// levels and records the state, k_bucket_partition groups the samples by
// their subtree, k_draw_suffix finishes the descents bucket by bucket on
// XCD-affine blocks.  The arithmetic and comparisons are those of k_krp_draw,
// so the results are identical.
// ---------------------------------------------------------------------------
template <int RMAX, int SG, int BS>
__global__ void __launch_bounds__(BS) k_draw_prefix(DrawArgs a) {
    using Cfg = DrawCfg<RMAX, SG>;
    constexpr int EPL = Cfg::EPL, HPL = Cfg::HPL;
    constexpr int GPW = 64 / SG;
    __shared__ double lds_x[BS / SG][RMAX];
    __shared__ uint16_t pair_tab[Cfg::TMAX];
    fill_pairs<BS>(pair_tab, a.R);
    const int lane = threadIdx.x & 63;
    const int g = lane % SG;
    const int wave = threadIdx.x / 64;
    double* xs = lds_x[threadIdx.x / SG];
    const int R = a.R;
    const int T = (int)tri_size(R);
    for (int b = g; b < RMAX; b += SG) xs[b] = 0.0;
    wave_sync_lds();
    const int mi = a.mi_begin;
    const int k = a.order[mi];
    const ModeDev& md = a.mode[k];
    const uint64_t waves_total = (uint64_t)gridDim.x * (BS / 64);
    const uint64_t n_items = (a.J + GPW - 1) / GPW;
    for (uint64_t item = (uint64_t)blockIdx.x * (BS / 64) + wave; item < n_items; item += waves_total) {
        const uint64_t i = item * GPW + lane / SG;
        const bool active = i < a.J;
        double h[HPL];
#pragma unroll
        for (int q = 0; q < HPL; q++) h[q] = 1.0;
        if (mi > 0 && active) {
#pragma unroll
            for (int q = 0; q < HPL; q++)
                if (g + SG * q < R) h[q] = a.h_state[i * R + g + SG * q];
        }
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) xs[g + SG * q] = h[q];
        wave_sync_lds();
        double c64[EPL];
        make_coefs<RMAX, SG>(xs, pair_tab, T, g, c64);
        const double ue = active ? (a.draws ? a.draws[(uint64_t)k * a.J + i] : philox_uniform(a.seed, a.call, i, 2 * k))
                                 : 0.5;
        double low, high;
        const uint64_t ce = descend<RMAX, SG, false>(md, a.esh, a.ts64, a.ts32, c64, md.etree, ue, g, low, high);
        const uint64_t r = leaf_index(a.esh, ce);
        wave_sync_lds();
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) {
                const double x = h[q] * md.E[r * R + g + SG * q];
                xs[g + SG * q] = x;
                if (active) a.st_x[i * R + g + SG * q] = x;
            }
        wave_sync_lds();
        make_coefs<RMAX, SG>(xs, pair_tab, T, g, c64);
        const double ug = active ? (a.draws ? a.draws[(uint64_t)a.N * a.J + (uint64_t)k * a.J + i]
                                            : philox_uniform(a.seed, a.call, i, 2 * k + 1))
                                 : 0.5;
        const double m = node_qform<double, EPL, SG>(md.tree64, c64, g, md.T);
        uint64_t c = 0;
        low = 0.0;
        high = 1.0;
        descend_steps<RMAX, SG, false>(md, md.sh, a.ts64, a.ts32, c64, md.tree64, ug, g, m, 0, a.split_level, false, c,
                                       low, high);
        if (active && g == 0) {
            a.st_s[i * 4 + 0] = low;
            a.st_s[i * 4 + 1] = high;
            a.st_s[i * 4 + 2] = m;
            a.st_s[i * 4 + 3] = (double)c;
            a.bucket[i] = (uint32_t)(c - ((1ull << a.split_level) - 1)) % kBuckets;
        }
        wave_sync_lds();
    }
}

// Stable partition of the samples by bucket (one block): list holds the
// sample indices bucket by bucket, off[b] the start of bucket b.
template <int BS>
__global__ void __launch_bounds__(BS) k_bucket_partition(const uint32_t* __restrict__ bucket, uint64_t J,
                                                         uint32_t* __restrict__ list, uint32_t* __restrict__ off) {
    __shared__ uint32_t cnt[kBuckets][BS];
    __shared__ uint32_t base[kBuckets + 1];
    const uint32_t t = threadIdx.x;
    const uint64_t chunk = (J + BS - 1) / BS;
    const uint64_t s0 = min<uint64_t>(J, t * chunk), s1 = min<uint64_t>(J, s0 + chunk);
    uint32_t c[kBuckets];
#pragma unroll
    for (uint32_t b = 0; b < kBuckets; b++) c[b] = 0;
    for (uint64_t s = s0; s < s1; s++) c[bucket[s]]++;
#pragma unroll
    for (uint32_t b = 0; b < kBuckets; b++) cnt[b][t] = c[b];
    __syncthreads();
    if (t < kBuckets) {  // exclusive scan of bucket t over the threads
        uint32_t run = 0;
        for (uint32_t q = 0; q < BS; q++) {
            const uint32_t v = cnt[t][q];
            cnt[t][q] = run;
            run += v;
        }
        base[t + 1] = run;  // size of bucket t
    }
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (uint32_t b = 0; b < kBuckets; b++) {
            const uint32_t n = base[b + 1];
            base[b] = run;
            run += n;
            off[b] = base[b];
        }
        base[kBuckets] = run;
        off[kBuckets] = run;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t b = 0; b < kBuckets; b++) c[b] = base[b] + cnt[b][t];
    for (uint64_t s = s0; s < s1; s++) list[c[bucket[s]]++] = (uint32_t)s;
}

template <int RMAX, int SG, int BS, bool USE32>
__global__ void __launch_bounds__(BS) k_draw_suffix(DrawArgs a) {
    using Cfg = DrawCfg<RMAX, SG>;
    constexpr int EPL = Cfg::EPL, HPL = Cfg::HPL;
    constexpr int GPW = 64 / SG;
    __shared__ double lds_x[BS / SG][RMAX];
    __shared__ double lds_q[BS / SG][RMAX];
    __shared__ uint16_t pair_tab[Cfg::TMAX];
    fill_pairs<BS>(pair_tab, a.R);
    const int lane = threadIdx.x & 63;
    const int g = lane % SG;
    const int wave = threadIdx.x / 64;
    double* xs = lds_x[threadIdx.x / SG];
    double* qs = lds_q[threadIdx.x / SG];
    const int R = a.R;
    const int T = (int)tri_size(R);
    for (int b = g; b < RMAX; b += SG) xs[b] = 0.0;
    wave_sync_lds();
    const int mi = a.mi_begin;
    const int k = a.order[mi];
    const ModeDev& md = a.mode[k];
    const bool last = mi + 1 >= a.m;
    const uint32_t bk = blockIdx.x % kBuckets;
    const uint64_t nb = (gridDim.x + kBuckets - 1 - bk) / kBuckets;  // blocks serving bucket bk
    const uint64_t sub = blockIdx.x / kBuckets;
    const uint64_t b0 = a.boff[bk], n = a.boff[bk + 1] - b0;
    const uint64_t waves_total = nb * (BS / 64);
    const uint64_t n_items = (n + GPW - 1) / GPW;
    for (uint64_t item = sub * (BS / 64) + wave; item < n_items; item += waves_total) {
        const uint64_t t = item * GPW + lane / SG;
        const bool active = t < n;
        const uint64_t i = active ? a.blist[b0 + t] : 0;
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) xs[g + SG * q] = active ? a.st_x[i * R + g + SG * q] : 0.0;
        wave_sync_lds();
        double c64[EPL];
        make_coefs<RMAX, SG>(xs, pair_tab, T, g, c64);
        const double ug = active ? (a.draws ? a.draws[(uint64_t)a.N * a.J + (uint64_t)k * a.J + i]
                                            : philox_uniform(a.seed, a.call, i, 2 * k + 1))
                                 : 0.5;
        double low = active ? a.st_s[i * 4 + 0] : 0.0;
        double high = active ? a.st_s[i * 4 + 1] : 1.0;
        const double m = active ? a.st_s[i * 4 + 2] : 1.0;
        uint64_t c = active ? (uint64_t)a.st_s[i * 4 + 3] : ((1ull << a.split_level) - 1);
        descend_steps<RMAX, SG, USE32>(md, md.sh, a.ts64, a.ts32, c64, md.tree64, ug, g, m, a.split_level,
                                       (uint32_t)md.sh.lfill_level, true, c, low, high);
        const uint64_t idx = leaf_index(md.sh, c);
        uint32_t res = 0;
        if (md.sh.F > 1) {
            const double frac = (ug - low) / (high - low);
            const uint64_t r0 = idx * md.sh.F;
            const uint32_t rows = (uint32_t)min<uint64_t>(md.sh.F, md.sh.n - r0);
            if constexpr (USE32)
                res = leaf_scan<float, RMAX, SG>(md.U32, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac, g, lane);
            else
                res = leaf_scan<double, RMAX, SG>(md.U64, r0, rows, (uint32_t)md.sh.F, R, xs, qs, frac, g, lane);
        }
        const uint64_t srow = idx * md.sh.F + res;
        double h[HPL];
#pragma unroll
        for (int q = 0; q < HPL; q++) {
            h[q] = 1.0;
            if (mi > 0 && active && g + SG * q < R) h[q] = a.h_state[i * R + g + SG * q];
            if (g + SG * q < R) h[q] *= md.U64[srow * R + g + SG * q];
        }
        if (active && g == 0) a.samples[(uint64_t)k * a.J + i] = srow;
        wave_sync_lds();
        if (!last) {
            if (active) {
#pragma unroll
                for (int q = 0; q < HPL; q++)
                    if (g + SG * q < R) a.h_state[i * R + g + SG * q] = h[q];
            }
            continue;
        }
        // ---- weight: R / (J * h^T pinv h) (sampler.hpp:94-122) --------------
#pragma unroll
        for (int q = 0; q < HPL; q++)
            if (g + SG * q < R) xs[g + SG * q] = h[q];
        wave_sync_lds();
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < HPL; q++) {
            const int ra = g + SG * q;
            if (ra < R) {
                double y = 0.0;
                for (int b = 0; b < R; b++) y += a.pinv[ra * R + b] * xs[b];
                part += h[q] * y;
            }
        }
        const double lev = group_sum<SG>(part);
        if (active) {
            if (a.w_out && g == 0) a.w_out[i] = (double)R / (lev * (double)a.J);
            if (a.h_out) {
#pragma unroll
                for (int q = 0; q < HPL; q++)
                    if (g + SG * q < R) a.h_out[i * R + g + SG * q] = h[q];
            }
        }
        wave_sync_lds();
    }
}

}  // namespace ftl

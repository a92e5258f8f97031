// als_kernels.hpp -- gfx950 kernels of the sparse tensor and the sampled
// CP-ALS step.
//
//   k_nz_keys / k_nz_gather    per-mode sorted lookup of the nonzeros
//                              (SortIdxLookup ctor, sort_lookup.hpp:33-60)
//   k_sample_keys ...          sort/unique of the drawn samples with counts
//   k_dedup_rows               w' = w c, h' = sqrt(w') h, h'' = sqrt(w') h'
//                              (als.hpp:76-160)
//   k_gram_partial             sampled Gram h'^T h' and factor Grams U^T U
//   k_pinv_partials            eigendecomposition-based pinv with the 1e-11
//                              cut and target rank (common.h:228-323)
//   k_probe / k_expand         equal_range of every sample on the sorted
//                              lookup and the list of matched nonzeros
//                              (sort_lookup.hpp:69-132)
//   k_seg_* / k_item_partial   sampled MTTKRP as a segmented reduction over
//   k_seg_final                matches sorted by output row, fused with the
//                              product by pinv (als.hpp:189-203)
//   k_colsq_partial ...        renormalize_columns (low_rank_tensor.hpp:203-260)
//   k_residual_partial         sum over nonzeros of x^2 - 2 x m
//   k_chain_final              ||M||^2 = sum sigma sigma^T (*) prod_k U_k^T U_k
//                              (sort_lookup.hpp:134-197, common.h:152-226)
//
// Every reduction has a fixed association (block partials summed in block
// order), so results do not depend on scheduling; there are no floating-point
// atomics.  A sum over the matches of one output row is taken sequentially
// in (sample, nonzero) order when the row has at most kSegPiece matches --
// the order of the CPU restatement, so those rows agree bitwise -- and in
// pieces of kSegPiece otherwise.
#pragma once

#include "ftl_common.hpp"
#include "jacobi.hpp"
#include "wave_ops.hpp"

namespace ftl {

// Matches of one output row summed sequentially per piece.
constexpr uint32_t kSegPiece = 256;
// A row with more pieces than this has them summed by a whole block
// (k_seg_heavy_sum) before the row finalise.
constexpr uint32_t kHeavyPieces = 32;
// Blocks of the partial-sum kernels (fixed per launch shape, so the
// association of every sum is fixed).
constexpr uint32_t kGramBlocks = 128;      // sampled Gram (J' rows)
constexpr uint32_t kGramBlocksMax = 512;   // factor Grams (I_k rows)
constexpr uint32_t kColBlocks = 256;       // column norms over all rows
constexpr uint32_t kSegBlocksMax = 2048;   // row finalise (+ column norms of touched rows)
constexpr uint32_t kResidBlocks = 4096;    // fit

// Coordinates of every mode but `skip` packed into one 64-bit key, the
// lowest-numbered mode in the most significant field, so that key order is
// the lexicographic order over modes != skip of the reference comparators
// (sort_lookup.hpp:52-58, als.hpp:92-101).
struct KeyLayout {
    uint32_t N;
    uint32_t skip;
    uint32_t shift[kMaxModes];
    uint32_t bits[kMaxModes];
    uint32_t total_bits;
};

struct FactorPtrs {
    const double* U[kMaxModes];
};

// Counters the ALS pipeline keeps on the device: unique samples J', matched
// nonzeros T, touched rows D, reduction pieces W, heavy rows (segments of more
// than kHeavyPieces pieces).
enum {
    kCntUnique = 0,
    kCntMatches = 1,
    kCntRows = 2,
    kCntPieces = 3,
    kCntHeavy = 4,
    kCntMatchesHi = 5,  // bits 32..63 of the 64-bit match total (host samples)
    kCntSlots = 6
};

__device__ inline uint64_t pack_tuple(const KeyLayout& L, const uint32_t* t) {
    uint64_t key = 0;
    for (uint32_t k = 0; k < L.N; k++)
        if (k != L.skip) key |= (uint64_t)t[k] << L.shift[k];
    return key;
}

// 64-bit lane broadcast within a group of `width` lanes
__device__ inline uint64_t shfl_u64(uint64_t v, int src, int width) {
    const uint32_t lo = __shfl((uint32_t)v, src, width), hi = __shfl((uint32_t)(v >> 32), src, width);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint32_t key_field(const KeyLayout& L, uint64_t key, uint32_t k) {
    const uint32_t b = L.bits[k];
    return b == 0 ? 0u : (uint32_t)((key >> L.shift[k]) & (b >= 64 ? ~0ull : ((1ull << b) - 1)));
}

// ---------------------------------------------------------------------------
// sparse tensor construction
// ---------------------------------------------------------------------------
__global__ void k_coord_max(const uint32_t* __restrict__ idx, uint64_t nnz, uint32_t N, uint32_t* __restrict__ mx) {
    __shared__ uint32_t loc[kMaxModes];
    if (threadIdx.x < kMaxModes) loc[threadIdx.x] = 0;
    __syncthreads();
    uint32_t m[kMaxModes];
    for (uint32_t k = 0; k < N; k++) m[k] = 0;
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x)
        for (uint32_t k = 0; k < N; k++) m[k] = max(m[k], idx[z * N + k]);
    for (uint32_t k = 0; k < N; k++) atomicMax(&loc[k], m[k]);
    __syncthreads();
    if (threadIdx.x < N) atomicMax(&mx[threadIdx.x], loc[threadIdx.x]);
}

__global__ void k_nz_keys(const uint32_t* __restrict__ idx, uint64_t nnz, KeyLayout L, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ perm) {
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz;
         z += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t key = 0;
        for (uint32_t k = 0; k < L.N; k++)
            if (k != L.skip) key |= (uint64_t)idx[z * L.N + k] << L.shift[k];
        keys[z] = key;
        perm[z] = (uint32_t)z;
    }
}

__global__ void k_nz_gather(const uint32_t* __restrict__ idx, const double* __restrict__ val,
                            const uint32_t* __restrict__ perm, uint64_t nnz, uint32_t N, uint32_t j,
                            uint32_t* __restrict__ row, double* __restrict__ vout) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t z = perm[i];
        row[i] = idx[z * N + j];
        vout[i] = val[z];
    }
}

// Block partial sums of v^2 (fixed grid; summed in block order by k_sum_final).
template <int BS>
__global__ void __launch_bounds__(BS) k_sumsq_partial(const double* __restrict__ v, uint64_t n,
                                                      double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double red[BS];
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BS)
        acc += v[i] * v[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] = sum of partial[0..n): thread t sums t, t+256, ... in order, then
// a fixed tree over the 256 thread sums.  One block of 256 threads.
__global__ void __launch_bounds__(256) k_sum_final(const double* __restrict__ partial, uint32_t n,
                                                   double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double red[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// sample sort / unique (als.hpp:76-160)
// ---------------------------------------------------------------------------
// samples: N x J (row k = mode k, the sampler's layout); coordinates cast to
// uint32 as the reference's transpose does (als.hpp:79).
__global__ void k_sample_keys(const uint64_t* __restrict__ samples, uint64_t J, KeyLayout L,
                              uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < J; s += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t key = 0;
        for (uint32_t k = 0; k < L.N; k++)
            if (k != L.skip) key |= (uint64_t)(uint32_t)samples[k * J + s] << L.shift[k];
        keys[s] = key;
        vals[s] = (uint32_t)s;
    }
}

template <typename K>
__global__ void k_head_flags(const K* __restrict__ keys, uint64_t n, uint32_t* __restrict__ flags) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// At each run head of the sorted sample keys: the unique tuple (coordinates
// of its first sample, mode `skip` zero), that first sample and the run
// start.  The last thread closes the run list and stores J'.
__global__ void k_dedup_heads(const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ svals,
                              const uint32_t* __restrict__ uid, uint64_t J, KeyLayout L,
                              uint32_t* __restrict__ utup, uint32_t* __restrict__ ufirst,
                              uint32_t* __restrict__ ustart, uint32_t* __restrict__ counters) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < J; i += (uint64_t)gridDim.x * blockDim.x) {
        const bool head = i == 0 || skeys[i] != skeys[i - 1];
        if (head) {
            const uint32_t u = uid[i] - 1;
            ufirst[u] = svals[i];
            ustart[u] = (uint32_t)i;
            for (uint32_t k = 0; k < L.N; k++) utup[(uint64_t)u * L.N + k] = key_field(L, skeys[i], k);
        }
        if (i == J - 1) {
            ustart[uid[i]] = (uint32_t)J;
            counters[kCntUnique] = uid[i];
        }
    }
}

// w' = w[first] * count, h' = h[first] * sqrt(w'), h'' = h' * sqrt(w')
// (als.hpp:136-170: the same two roundings).
// hdd rows have the even stride Rp = R + (R & 1) (a zero pad column for odd
// R), so the MTTKRP reads each row in 16-byte column pairs.
__global__ void k_dedup_rows(const uint32_t* __restrict__ ufirst, const uint32_t* __restrict__ ustart,
                             const uint32_t* __restrict__ counters, uint64_t J, uint32_t R,
                             const double* __restrict__ h, const double* __restrict__ w,
                             double* __restrict__ hs, double* __restrict__ hdd) {
#pragma clang fp contract(off)
    const uint64_t Jp = counters[kCntUnique];
    const uint64_t n = Jp * R;
    const uint64_t Rp = R + (R & 1);
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t u = e / R, c = e - u * R;
        const uint64_t f = ufirst[u];
        const double wd = w[f] * (double)(ustart[u + 1] - ustart[u]);
        const double sq = sqrt(wd);
        const double a = h[f * R + c] * sq;
        hs[e] = a;
        hdd[u * Rp + c] = a * sq;
        if (Rp != R && c == R - 1) hdd[u * Rp + R] = 0.0;
    }
    (void)J;
}

// ---------------------------------------------------------------------------
// Gram of the rows of X (n x R, row-major): block b sums rows
// [b*chunk, (b+1)*chunk) sequentially into the packed upper triangle
// partial[b][0..T).  n comes from *n_dev when given.
// ---------------------------------------------------------------------------
template <int BS, int RMAX>
__global__ void __launch_bounds__(BS) k_gram_partial(const double* __restrict__ X, const uint32_t* n_dev,
                                                     uint64_t n_host, uint32_t R, double* __restrict__ partial) {
#pragma clang fp contract(off)
    constexpr int TMAX = RMAX * (RMAX + 1) / 2;
    constexpr int EPT = (TMAX + BS - 1) / BS;
    constexpr int TILE = 32;
    __shared__ double tile[TILE * RMAX];
    __shared__ uint8_t pa[TMAX], pb[TMAX];
    const uint64_t n = n_dev ? (uint64_t)*n_dev : n_host;
    const int T = (int)(R * (R + 1) / 2);
    for (int r = threadIdx.x; r < (int)R; r += BS) {
        const int base = r * (int)R - r * (r - 1) / 2;
        for (int c = r; c < (int)R; c++) {
            pa[base + c - r] = (uint8_t)r;
            pb[base + c - r] = (uint8_t)c;
        }
    }
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t r0 = min<uint64_t>(n, (uint64_t)blockIdx.x * chunk), r1 = min<uint64_t>(n, r0 + chunk);
    double acc[EPT];
#pragma unroll
    for (int q = 0; q < EPT; q++) acc[q] = 0.0;
    __syncthreads();
    for (uint64_t t0 = r0; t0 < r1; t0 += TILE) {
        const int rows = (int)min<uint64_t>(TILE, r1 - t0);
        for (int e = threadIdx.x; e < rows * (int)R; e += BS) tile[e] = X[t0 * R + e];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < EPT; q++) {
            const int e = threadIdx.x + q * BS;
            if (e < T) {
                const int a = pa[e], b = pb[e];
                double s = acc[q];
                for (int r = 0; r < rows; r++) s += tile[r * R + a] * tile[r * R + b];
                acc[q] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < EPT; q++) {
        const int e = threadIdx.x + q * BS;
        if (e < T) partial[(uint64_t)blockIdx.x * T + e] = acc[q];
    }
}

// compute_pinv over the block partials of the sampled Gram: M = sum of the
// partials (in block order), eigendecomposition, keep eigenvalue v when
// v >= R - target and lambda_v > 1e-11 with target = min(J', R), then
// pinv = M' M'^T with M'[:, v] = V[:, v] / sqrt(lambda_v) (common.h:228-323).
// One block; pinv written full symmetric (R x R).
// pinv of a symmetric R x R matrix A (LDS, leading dimension R + 1) as its
// inverse by Cholesky, when that equals the eigendecomposition-based pinv
// with the 1e-11 cut at full target rank (compute_pinv_square,
// common.h:228-269): lambda_min(A) >= 1 / ||A^-1||_F > 2e-11 certifies that
// no eigenvalue would be cut.  Also required: ||A||_F ||A^-1||_F < 1e6, so
// the result agrees with the eigendecomposition route to rounding.  C (R
// (R+1) doubles) and Li (R^2) are scratch.  Right-looking Cholesky (a column
// per step), the inverse of the factor a column per thread (forward
// substitution), pinv = Li^T Li; every sum in a fixed order.  Returns false
// (pinv untouched) when a pivot is not positive or a test fails; the caller
// then runs the eigensolver.
constexpr int kCholMaxR = 64;

template <int BS>
__device__ bool pinv_cholesky(const double* A, double* C, double* Li, int R, double* __restrict__ pinv) {
#pragma clang fp contract(off)
    __shared__ double red[BS];
    __shared__ int fail;
    const int tid = threadIdx.x, LD = R + 1;
    for (int e = tid; e < R * R; e += BS) {
        const int i = e / R, j = e - i * R;
        C[i * LD + j] = A[i * LD + j];
    }
    if (tid == 0) fail = 0;
    __syncthreads();
    for (int k = 0; k < R; k++) {
        const double piv = C[k * LD + k];
        __syncthreads();
        if (!(piv > 0.0)) return false;  // uniform: every thread read the same pivot
        const double lkk = sqrt(piv);
        for (int i = k + 1 + tid; i < R; i += BS) C[i * LD + k] = C[i * LD + k] / lkk;
        if (tid == 0) C[k * LD + k] = lkk;
        __syncthreads();
        const int m = R - k - 1;  // trailing block, lower triangle
        for (int e = tid; e < m * m; e += BS) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            if (j <= i) C[i * LD + j] = C[i * LD + j] - C[i * LD + k] * C[j * LD + k];
        }
        __syncthreads();
    }
    // column j of L^-1 (lower): x_j = 1 / l_jj, x_i = -(sum_{m=j}^{i-1} l_im x_m) / l_ii
    for (int j = tid; j < R; j += BS) {
        for (int i = 0; i < j; i++) Li[i * R + j] = 0.0;
        Li[j * R + j] = 1.0 / C[j * LD + j];
        for (int i = j + 1; i < R; i++) {
            double s = 0.0;
            for (int m = j; m < i; m++) s = s + C[i * LD + m] * Li[m * R + j];
            Li[i * R + j] = (0.0 - s) / C[i * LD + i];
        }
    }
    __syncthreads();
    // A^-1 = Li^T Li into pinv; Frobenius norms of A^-1 and A
    double f_inv = 0.0, f_a = 0.0;
    for (int e = tid; e < R * R; e += BS) {
        const int a = e / R, b = e - a * R;
        const int m0 = a > b ? a : b;
        double s = 0.0;
        for (int m = m0; m < R; m++) s = s + Li[m * R + a] * Li[m * R + b];
        pinv[e] = s;
        f_inv = f_inv + s * s;
        const double x = A[a * LD + b];
        f_a = f_a + x * x;
        if (!(s == s) || s - s != 0.0) fail = 1;  // NaN or infinity
    }
    red[tid] = f_inv;
    __syncthreads();
    for (int h = BS / 2; h > 0; h >>= 1) {
        if (tid < h) red[tid] = red[tid] + red[tid + h];
        __syncthreads();
    }
    const double fi = red[0];
    __syncthreads();
    red[tid] = f_a;
    __syncthreads();
    for (int h = BS / 2; h > 0; h >>= 1) {
        if (tid < h) red[tid] = red[tid] + red[tid + h];
        __syncthreads();
    }
    const double fa = red[0];
    const bool ok = fail == 0 && fi > 0.0 && 1.0 / sqrt(fi) > 2e-11 && sqrt(fa) * sqrt(fi) < 1e6;
    __syncthreads();
    return ok;  // when false the caller overwrites pinv
}

template <int BS>
__global__ void __launch_bounds__(BS) k_pinv_partials(const double* __restrict__ partial, uint32_t nparts,
                                                      const uint32_t* __restrict__ counters, uint32_t R,
                                                      double* __restrict__ pinv, double* __restrict__ work,
                                                      double* __restrict__ lam_out) {
#pragma clang fp contract(off)
    extern __shared__ double smem[];
    __shared__ JacobiShared js;
    __shared__ double lam[128], diag_tmp[128];
    const int tid = threadIdx.x;
    const int Ri = (int)R, LD = Ri + 1, T = Ri * (Ri + 1) / 2;
    JacobiStore st = jacobi_store(smem, work, Ri);
    for (int e = tid; e < T; e += BS) {
        // packed index e -> (a, b), a <= b
        int a = 0, base = 0;
        while (base + (Ri - a) <= e) {
            base += Ri - a;
            a++;
        }
        const int b = a + (e - base);
        // partials summed in block order; loads issued 16 at a time
        double s = 0.0;
        uint32_t p = 0;
        for (; p + 16 <= nparts; p += 16) {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = partial[(uint64_t)(p + u) * T + e];
#pragma unroll
            for (int u = 0; u < 16; u++) s += v[u];
        }
        for (; p < nparts; p++) s += partial[(uint64_t)p * T + e];
        st.A[a * LD + b] = s;
        st.A[b * LD + a] = s;
    }
    __syncthreads();
    const uint32_t Jp = counters[kCntUnique];
    const int target = (int)min<uint32_t>(Jp, R);
    // full target rank and a well-conditioned Gram: pinv = G^-1 by Cholesky
    if (target == Ri && Ri <= kCholMaxR && 3 * LD * Ri <= kJacobiLdsDoubles &&
        pinv_cholesky<BS>(st.A, st.V, st.vecs, Ri, pinv))
        return;
    jacobi_block<BS>(st.A, st.V, Ri, lam, st.vecs, js, diag_tmp);
    // M' in place of the eigenvectors
    for (int w = tid; w < Ri * Ri; w += BS) {
        const int u = w / Ri, v = w - u * Ri;
        const bool keep = v >= Ri - target && lam[v] > 1e-11;
        st.vecs[u * Ri + v] = keep ? st.vecs[u * Ri + v] / sqrt(lam[v]) : 0.0;
    }
    if (tid < Ri) lam_out[tid] = lam[tid];
    __syncthreads();
    for (int w = tid; w < Ri * Ri; w += BS) {
        const int a = w / Ri, b = w - a * Ri;
        if (a > b) continue;
        double s = 0.0;
        for (int c = 0; c < Ri; c++) s += st.vecs[a * Ri + c] * st.vecs[b * Ri + c];
        pinv[a * Ri + b] = s;
        pinv[b * Ri + a] = s;
    }
}

// ---------------------------------------------------------------------------
// probe: equal_range of each tuple's key on mode j's sorted keys
// (sort_lookup.hpp:94-103).  A coordinate outside the tensor's range cannot
// match (its field would not hold it) and gives an empty range.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Coordinate hash of mode j's lookup: an open-addressing table (linear
// probing, load factor <= 1/2, a power-of-two number of slots) holding, for
// every distinct key of the sorted nonzeros, the key and its run [lo,
// lo + cnt) in the sorted arrays -- the range a binary search would find.
// Empty slots hold key kHashEmpty, which no key of fewer than 64 bits takes.
// ---------------------------------------------------------------------------
struct HashSlot {
    uint64_t key;
    uint32_t lo, cnt;
};
constexpr uint64_t kHashEmpty = ~0ull;

// splitmix64's finalizer (Steele, Lea and Flood, OOPSLA 2014)
__device__ inline uint64_t hash_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Number of distinct keys (run heads) of the sorted keys, added to *count.
__global__ void k_count_heads(const uint64_t* __restrict__ keys, uint64_t nnz, unsigned long long* count) {
    uint64_t c = 0;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (uint64_t)gridDim.x * blockDim.x)
        c += (p == 0 || keys[p - 1] != keys[p]) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}

// Inserts every run head p of the sorted keys with its run length (the
// first later position holding a larger key, by binary search).
__global__ void k_hash_build(const uint64_t* __restrict__ keys, uint64_t nnz, HashSlot* __restrict__ table,
                             uint64_t mask) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[p];
        if (p > 0 && keys[p - 1] == key) continue;
        uint64_t l = p + 1, h = nnz;
        while (l < h) {  // first position with keys[q] > key
            const uint64_t m = (l + h) >> 1;
            if (keys[m] <= key) l = m + 1;
            else h = m;
        }
        uint64_t slot = hash_mix(key) & mask;
        for (;;) {
            const unsigned long long prev = atomicCAS((unsigned long long*)&table[slot].key,
                                                      (unsigned long long)kHashEmpty, (unsigned long long)key);
            if (prev == kHashEmpty) {
                table[slot].lo = (uint32_t)p;
                table[slot].cnt = (uint32_t)(l - p);
                break;
            }
            slot = (slot + 1) & mask;
        }
    }
}

struct ProbeArgs {
    const uint32_t* tup;       // n x N, row-major
    const uint32_t* n_dev;     // J' on the device, or null (n_host)
    uint64_t n_host;
    uint64_t n_cap;            // entries of lo / cnt written (zero past n)
    const uint64_t* keys;      // sorted lookup keys of mode j
    uint64_t nnz;
    KeyLayout L;               // the tensor's layout for mode j
    uint64_t dims[kMaxModes];  // coordinate bound per mode
    uint32_t* lo;
    uint32_t* cnt;
    const HashSlot* table;     // mode j's coordinate hash, or null (binary search)
    uint64_t mask;             // table slots - 1
};

__global__ void k_probe(ProbeArgs a) {
    const uint64_t n = a.n_dev ? (uint64_t)*a.n_dev : a.n_host;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.n_cap;
         s += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, c = 0;
        if (s < n) {
            const uint32_t* t = a.tup + s * a.L.N;
            bool ok = true;
            uint64_t key = 0;
            for (uint32_t k = 0; k < a.L.N; k++) {
                if (k == a.L.skip) continue;
                ok &= t[k] < a.dims[k];
                key |= (uint64_t)t[k] << a.L.shift[k];
            }
            if (ok && a.table) {
                uint64_t slot = hash_mix(key) & a.mask;
                for (;;) {
                    const HashSlot e = a.table[slot];
                    if (e.key == key) {
                        lo = e.lo;
                        c = e.cnt;
                        break;
                    }
                    if (e.key == kHashEmpty) break;
                    slot = (slot + 1) & a.mask;
                }
            } else if (ok) {
                uint64_t l = 0, h = a.nnz;
                while (l < h) {  // first position with keys[p] >= key
                    const uint64_t m = (l + h) >> 1;
                    if (a.keys[m] < key) l = m + 1;
                    else h = m;
                }
                uint64_t l2 = l, h2 = a.nnz;
                while (l2 < h2) {  // first position with keys[p] > key
                    const uint64_t m = (l2 + h2) >> 1;
                    if (a.keys[m] <= key) l2 = m + 1;
                    else h2 = m;
                }
                lo = (uint32_t)l;
                c = (uint32_t)(l2 - l);
            }
        }
        a.lo[s] = lo;
        a.cnt[s] = c;
    }
}

// counters[slot] = off[n-1] + cnt[n-1] (n >= 1): total of an exclusive scan.
__global__ void k_scan_total(const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint64_t n,
                             uint32_t* __restrict__ counters, int slot) {
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[slot] = off[n - 1] + cnt[n - 1];
}

// Bits 32..63 of sum(cnt[0..n)) into counters[slot]: the uint32 exclusive
// scan of cnt wraps once the match total reaches 2^32, which only a host
// sample list with repeated tuples can reach (a deduplicated one matches
// disjoint ranges, so at most nnz < 2^32).  One block; the sum is exact.
__global__ void __launch_bounds__(1024) k_total_hi(const uint32_t* __restrict__ cnt, uint64_t n,
                                                   uint32_t* __restrict__ counters, int slot) {
    __shared__ uint64_t part[1024];
    uint64_t acc = 0;
    for (uint64_t u = threadIdx.x; u < n; u += 1024) acc += cnt[u];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) counters[slot] = (uint32_t)(part[0] >> 32);
}

// mark[off[u]] = u for every sample u with matches, on a zeroed array of T
// entries: its inclusive max-scan gives every match the sample that owns it
// (the one u with off[u] <= m < off[u] + cnt[u]; samples without matches
// share their offset with the next sample that has some)
__global__ void k_mark_owners(const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint64_t n,
                              uint32_t* __restrict__ mark) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x)
        if (cnt[u] > 0) mark[off[u]] = (uint32_t)u;
}

// Match m (sample u = owner[m], the (m - off[u])-th nonzero of its range):
// output row and payload (u << 32 | sorted position).  Matches are numbered
// in (sample, position) order.
__global__ void k_expand(const uint32_t* __restrict__ owner, const uint32_t* __restrict__ off,
                         const uint32_t* __restrict__ lo, uint64_t T, const uint32_t* __restrict__ rowj,
                         uint32_t* __restrict__ mrow, uint64_t* __restrict__ mpay) {
    for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < T; m += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t l = owner[m];
        const uint64_t pos = (uint64_t)lo[l] + (m - off[l]);
        mrow[m] = rowj[pos];
        mpay[m] = (l << 32) | pos;
    }
}

// Row segments of the matches sorted by row: start of segment d, its row,
// and the number of reduction pieces (ceil(len / kSegPiece)).
__global__ void k_seg_heads(const uint32_t* __restrict__ srow, const uint32_t* __restrict__ segid, uint64_t T,
                            uint32_t* __restrict__ seg_start, uint32_t* __restrict__ seg_row,
                            uint32_t* __restrict__ counters) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == 0 || srow[i] != srow[i - 1]) {
            const uint32_t d = segid[i] - 1;
            seg_start[d] = (uint32_t)i;
            seg_row[d] = srow[i];
        }
        if (i == T - 1) {
            seg_start[segid[i]] = (uint32_t)T;
            counters[kCntRows] = segid[i];
        }
    }
}

__global__ void k_seg_pieces(const uint32_t* __restrict__ seg_start, const uint32_t* __restrict__ counters,
                             uint64_t cap, uint32_t* __restrict__ npieces) {
    const uint64_t D = counters[kCntRows];
    for (uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; d < cap; d += (uint64_t)gridDim.x * blockDim.x)
        npieces[d] = d < D ? (seg_start[d + 1] - seg_start[d] + kSegPiece - 1) / kSegPiece : 0u;
}

// Piece w of the segmented reduction: the sequential sum over its matches of
// lhs[u] * value (a group of G lanes, CPL columns per lane).  The payloads and
// values of G matches are fetched at once (one per lane) and broadcast, so the
// loop only waits for the lhs rows, which are independent across matches.
template <int G, int CPL>
__global__ void __launch_bounds__(256) k_item_partial(const uint32_t* __restrict__ seg_start,
                                                      const uint32_t* __restrict__ piece_off,
                                                      const uint32_t* __restrict__ piece_seg,
                                                      const uint32_t* __restrict__ counters,
                                                      const uint64_t* __restrict__ spay,
                                                      const double* __restrict__ valj,
                                                      const double* __restrict__ lhs, uint32_t R,
                                                      double* __restrict__ partial) {
#pragma clang fp contract(off)
    const uint32_t W = counters[kCntPieces];
    const uint32_t lane = threadIdx.x % G;
    const uint64_t groups = (uint64_t)gridDim.x * (256 / G);
    for (uint64_t w = (uint64_t)blockIdx.x * (256 / G) + threadIdx.x / G; w < W; w += groups) {
        const uint32_t l = piece_seg[w];  // segment of piece w
        const uint64_t i0 = (uint64_t)seg_start[l] + (uint64_t)(w - piece_off[l]) * kSegPiece;
        const uint64_t i1 = min<uint64_t>(i0 + kSegPiece, seg_start[l + 1]);
        double acc[CPL];
#pragma unroll
        for (int q = 0; q < CPL; q++) acc[q] = 0.0;
        for (uint64_t c0 = i0; c0 < i1; c0 += G) {
            const int cnt = (int)min<uint64_t>(G, i1 - c0);
            uint32_t my_u = 0;
            double my_v = 0.0;
            if ((int)lane < cnt) {
                const uint64_t pay = spay[c0 + lane];
                my_u = (uint32_t)(pay >> 32);
                my_v = valj[pay & 0xffffffffull];
            }
            // the lhs rows of 8 matches are fetched before they are added (in
            // match order, so the sum is the same)
            for (int t0 = 0; t0 < cnt; t0 += 8) {
                double x[8][CPL], vv[8];
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const int t = t0 + s < cnt ? t0 + s : cnt - 1;
                    const uint64_t u = (uint32_t)__shfl(my_u, t, G);
                    vv[s] = __shfl(my_v, t, G);
#pragma unroll
                    for (int q = 0; q < CPL; q++) {
                        const uint32_t c = lane + q * G;
                        x[s][q] = c < R ? lhs[u * R + c] : 0.0;
                    }
                }
#pragma unroll
                for (int s = 0; s < 8; s++)
                    if (t0 + s < cnt) {
#pragma unroll
                        for (int q = 0; q < CPL; q++) {
                            const uint32_t c = lane + q * G;
                            if (c < R) acc[q] = acc[q] + x[s][q] * vv[s];
                        }
                    }
            }
        }
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            const uint32_t c = lane + q * G;
            if (c < R) partial[(uint64_t)w * R + c] = acc[q];
        }
    }
}

// k_item_partial for lhs rows of even stride Rp (16-byte aligned): lane l of
// a G-lane group sums columns 2l and 2l + 1, reading them with one 16-byte
// load per match -- half the load instructions (and addresses) of one
// column per lane.  Each column's sum is the same sequential sum over the
// piece's matches in order, so the partials equal k_item_partial's.
template <int G>
__global__ void __launch_bounds__(256) k_item_partial_pairs(const uint32_t* __restrict__ seg_start,
                                                            const uint32_t* __restrict__ piece_off,
                                                            const uint32_t* __restrict__ piece_seg,
                                                            const uint32_t* __restrict__ counters,
                                                            const uint64_t* __restrict__ spay,
                                                            const double* __restrict__ valj,
                                                            const double* __restrict__ lhs, uint32_t R, uint32_t Rp,
                                                            double* __restrict__ partial) {
#pragma clang fp contract(off)
    const uint32_t W = counters[kCntPieces];
    const uint32_t lane = threadIdx.x % G;
    const uint32_t c0 = 2 * lane;
    const bool on = c0 < R;
    const uint64_t groups = (uint64_t)gridDim.x * (256 / G);
    for (uint64_t w = (uint64_t)blockIdx.x * (256 / G) + threadIdx.x / G; w < W; w += groups) {
        const uint32_t l = piece_seg[w];  // segment of piece w
        const uint64_t i0 = (uint64_t)seg_start[l] + (uint64_t)(w - piece_off[l]) * kSegPiece;
        const uint64_t i1 = min<uint64_t>(i0 + kSegPiece, seg_start[l + 1]);
        double a0 = 0.0, a1 = 0.0;
        for (uint64_t cb = i0; cb < i1; cb += G) {
            const int cnt = (int)min<uint64_t>(G, i1 - cb);
            uint32_t my_u = 0;
            double my_v = 0.0;
            if ((int)lane < cnt) {
                const uint64_t pay = spay[cb + lane];
                my_u = (uint32_t)(pay >> 32);
                my_v = valj[pay & 0xffffffffull];
            }
            // the lhs pairs of 8 matches are fetched before they are added (in
            // match order, so the sums are the same)
            for (int t0 = 0; t0 < cnt; t0 += 8) {
                double2 x[8];
                double vv[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int t = t0 + q < cnt ? t0 + q : cnt - 1;
                    const uint64_t u = (uint32_t)__shfl(my_u, t, G);
                    vv[q] = __shfl(my_v, t, G);
                    x[q] = on ? *reinterpret_cast<const double2*>(lhs + u * Rp + c0) : double2{0.0, 0.0};
                }
#pragma unroll
                for (int q = 0; q < 8; q++)
                    if (t0 + q < cnt) {
                        a0 = a0 + x[q].x * vv[q];
                        a1 = a1 + x[q].y * vv[q];
                    }
            }
        }
        if (on) {
            partial[(uint64_t)w * R + c0] = a0;
            if (c0 + 1 < R) partial[(uint64_t)w * R + c0 + 1] = a1;
        }
    }
}

// Piece w of the exact MTTKRP (sort_lookup.hpp:199-254): the nonzeros of
// one output row, in the order of the mode-j lookup (pos[] lists lookup
// positions sorted stably by the mode-j coordinate), each adding
// prod_{k != j} U_k[i_k] * value -- the product formed from 1 over the modes
// in order, then times the value, as the reference's partial_prod -- summed
// sequentially per piece.  Keys and values of G nonzeros are fetched at once
// (one per lane) and broadcast; lane c holds column c (CPL per lane).
template <int G, int CPL>
__global__ void __launch_bounds__(256) k_exact_item_partial(const uint32_t* __restrict__ seg_start,
                                                            const uint32_t* __restrict__ piece_off,
                                                            const uint32_t* __restrict__ piece_seg,
                                                            const uint32_t* __restrict__ counters,
                                                            const uint32_t* __restrict__ pos,
                                                            const uint64_t* __restrict__ keys,
                                                            const double* __restrict__ vals, KeyLayout L,
                                                            FactorPtrs F, uint32_t R, double* __restrict__ partial) {
#pragma clang fp contract(off)
    const uint32_t W = counters[kCntPieces];
    const uint32_t lane = threadIdx.x % G;
    const uint64_t groups = (uint64_t)gridDim.x * (256 / G);
    for (uint64_t w = (uint64_t)blockIdx.x * (256 / G) + threadIdx.x / G; w < W; w += groups) {
        const uint32_t l = piece_seg[w];  // segment of piece w
        const uint64_t i0 = (uint64_t)seg_start[l] + (uint64_t)(w - piece_off[l]) * kSegPiece;
        const uint64_t i1 = min<uint64_t>(i0 + kSegPiece, seg_start[l + 1]);
        double acc[CPL];
#pragma unroll
        for (int q = 0; q < CPL; q++) acc[q] = 0.0;
        for (uint64_t c0 = i0; c0 < i1; c0 += G) {
            const int cnt = (int)min<uint64_t>(G, i1 - c0);
            uint64_t my_key = 0;
            double my_v = 0.0;
            if ((int)lane < cnt) {
                const uint32_t p = pos[c0 + lane];
                my_key = keys[p];
                my_v = vals[p];
            }
            for (int t = 0; t < cnt; t++) {
                const uint64_t key = shfl_u64(my_key, t, G);
                const double v = __shfl(my_v, t, G);
#pragma unroll
                for (int q = 0; q < CPL; q++) {
                    const uint32_t c = lane + q * G;
                    if (c < R) {
                        double pr = 1.0;
                        for (uint32_t k = 0; k < L.N; k++)
                            if (k != L.skip) pr = pr * F.U[k][(uint64_t)key_field(L, key, k) * R + c];
                        acc[q] = acc[q] + pr * v;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            const uint32_t c = lane + q * G;
            if (c < R) partial[(uint64_t)w * R + c] = acc[q];
        }
    }
}

// Packed Gram of the exact update: (*)_{k != skip} U_k^T U_k from the
// per-mode packed Grams, starting from ones and multiplying in mode order
// (ATB_chain_prod with unit sigmas, common.h:152-213).
__global__ void k_gram_hadamard(const double* __restrict__ G, uint32_t N, uint32_t T, uint32_t skip,
                                double* __restrict__ out) {
#pragma clang fp contract(off)
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < T; e += gridDim.x * blockDim.x) {
        double p = 1.0;
        for (uint32_t k = 0; k < N; k++)
            if (k != skip) p = p * G[(uint64_t)k * T + e];
        out[e] = p;
    }
}

__global__ void k_iota(uint32_t* __restrict__ out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)i;
}

// piece_seg[piece_off[d]] = d for every segment (the entries between stay
// zero); an inclusive max-scan then gives every piece its segment.
__global__ void k_mark_segs(const uint32_t* __restrict__ piece_off, const uint32_t* __restrict__ counters,
                            uint32_t* __restrict__ piece_seg) {
    const uint64_t D = counters[kCntRows];
    for (uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; d < D; d += (uint64_t)gridDim.x * blockDim.x)
        piece_seg[piece_off[d]] = (uint32_t)d;
}

// Rows with more than kHeavyPieces pieces, listed in any order (each is
// reduced on its own, so the order does not change any result).
__global__ void k_heavy_list(const uint32_t* __restrict__ piece_off, uint32_t* __restrict__ counters,
                             uint32_t* __restrict__ list) {
    const uint32_t D = counters[kCntRows], W = counters[kCntPieces];
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < D; d += gridDim.x * blockDim.x) {
        const uint32_t p1 = d + 1 < D ? piece_off[d + 1] : W;
        if (p1 - piece_off[d] > kHeavyPieces) list[atomicAdd(&counters[kCntHeavy], 1u)] = d;
    }
}

// Sum of the pieces of every heavy row into its first piece, with a fixed
// association: 8 piece lanes per column sum every 8th piece in order, then
// the 8 lane sums are added in lane order.  One block per row at a time.
__global__ void __launch_bounds__(256) k_seg_heavy_sum(const uint32_t* __restrict__ piece_off,
                                                       const uint32_t* __restrict__ counters,
                                                       const uint32_t* __restrict__ list, uint32_t R,
                                                       double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double red[8][32];
    const uint32_t nh = counters[kCntHeavy], D = counters[kCntRows], W = counters[kCntPieces];
    const uint32_t l = threadIdx.x / 32, c = threadIdx.x % 32;
    for (uint32_t i = blockIdx.x; i < nh; i += gridDim.x) {
        const uint32_t d = list[i];
        const uint64_t p0 = piece_off[d], p1 = d + 1 < D ? piece_off[d + 1] : W;
        for (uint32_t c0 = 0; c0 < R; c0 += 32) {
            const uint32_t col = c0 + c;
            double s = 0.0;
            if (col < R) {
#pragma unroll 4
                for (uint64_t p = p0 + l; p < p1; p += 8) s = s + partial[p * R + col];
            }
            red[l][c] = s;
            __syncthreads();
            if (l == 0 && col < R) {
                double t = red[0][c];
                for (int q = 1; q < 8; q++) t = t + red[q][c];
                partial[p0 * R + col] = t;
            }
            __syncthreads();
        }
    }
}

// Segment d: sum of its pieces in order, then either the row itself (plain
// MTTKRP) or its product with pinv (the ALS solve U_j = MTTKRP pinv,
// als.hpp:189-203), written to row seg_row[d] of out.  With colsq given, the
// block also sums the squares of the rows it wrote per column (block order,
// then group order: a fixed association) into colsq[block][c], the input of
// renormalize_columns over the touched rows (all other rows are zero).
template <int G, int CPL, bool SOLVE>
__global__ void __launch_bounds__(256) k_seg_final(const uint32_t* __restrict__ seg_row,
                                                   const uint32_t* __restrict__ piece_off,
                                                   const uint32_t* __restrict__ counters,
                                                   const double* __restrict__ partial, uint32_t R,
                                                   const double* __restrict__ pinv, double* __restrict__ out,
                                                   double* __restrict__ colsq) {
#pragma clang fp contract(off)
    constexpr int GPB = 256 / G;
    constexpr int PINV_LDS = 64 * 64;
    // one column per lane (CPL == 1): the lane keeps its column of pinv in
    // registers, so a row's product reads only the row sums from LDS; the
    // column-loop form stages pinv in LDS (static LDS bounds the blocks a CU
    // holds, so the register form allocates none for it)
    constexpr bool kRegP = SOLVE && CPL == 1;
    __shared__ double mt_s[GPB][G * CPL];
    __shared__ double pinv_s[(SOLVE && !kRegP) ? PINV_LDS : 1];
    __shared__ double sq_s[GPB][G * CPL];
    const uint32_t D = counters[kCntRows], W = counters[kCntPieces];
    const uint32_t lane = threadIdx.x % G, gi = threadIdx.x / G;
    const bool pinv_in_lds = SOLVE && !kRegP && R * R <= (uint32_t)PINV_LDS;
    if constexpr (SOLVE) {
        if (pinv_in_lds)
            for (uint32_t e = threadIdx.x; e < R * R; e += 256) pinv_s[e] = pinv[e];
    }
    const double* P = pinv_in_lds ? pinv_s : pinv;
    double sq[CPL];
#pragma unroll
    for (int q = 0; q < CPL; q++) sq[q] = 0.0;
    __syncthreads();
    double pc[kRegP ? G : 1];  // pinv column `lane` (zero past R)
    if constexpr (kRegP) {
#pragma unroll
        for (int k = 0; k < G; k++) pc[k] = ((uint32_t)k < R && lane < R) ? P[k * R + lane] : 0.0;
    }
    const uint64_t groups = (uint64_t)gridDim.x * GPB;
    // groups run independently: a group stages its row in its own LDS slot,
    // and a group is at most one wave, whose LDS accesses are ordered.  A
    // group's next segment has its first piece's values in flight while the
    // current one is summed and multiplied, and the one after that its piece
    // range and row.
    uint64_t d = (uint64_t)blockIdx.x * GPB + gi;
    uint32_t nx0 = 0, nx1 = 0, nxrow = 0, n20 = 0, n21 = 0, n2row = 0;
    double nxv[CPL];
#pragma unroll
    for (int q = 0; q < CPL; q++) nxv[q] = 0.0;
    if (d < D) {
        nx0 = piece_off[d];
        nx1 = d + 1 < D ? piece_off[d + 1] : W;
        nxrow = seg_row[d];
#pragma unroll
        for (int q = 0; q < CPL; q++)
            if (lane + q * G < R) nxv[q] = partial[(uint64_t)nx0 * R + lane + q * G];
    }
    if (d + groups < D) {
        const uint64_t d2 = d + groups;
        n20 = piece_off[d2];
        n21 = d2 + 1 < D ? piece_off[d2 + 1] : W;
        n2row = seg_row[d2];
    }
    for (; d < D; d += groups) {
        double mt[CPL];
        const uint32_t p0 = nx0;
        uint32_t p1 = nx1;
        const uint64_t row = nxrow;
#pragma unroll
        for (int q = 0; q < CPL; q++) mt[q] = nxv[q];
        if (d + groups < D) {
            nx0 = n20;
            nx1 = n21;
            nxrow = n2row;
#pragma unroll
            for (int q = 0; q < CPL; q++)
                if (lane + q * G < R) nxv[q] = partial[(uint64_t)nx0 * R + lane + q * G];
        }
        if (d + 2 * groups < D) {
            const uint64_t d2 = d + 2 * groups;
            n20 = piece_off[d2];
            n21 = d2 + 1 < D ? piece_off[d2 + 1] : W;
            n2row = seg_row[d2];
        }
        if (p1 - p0 > kHeavyPieces) p1 = p0 + 1;  // k_seg_heavy_sum left the row's sum in piece p0
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            const uint32_t c = lane + q * G;
            double s = mt[q];  // piece p0 (0 past R)
            if (c < R)
                for (uint32_t p = p0 + 1; p < p1; p++) s = s + partial[(uint64_t)p * R + c];
            mt[q] = s;
        }
        if constexpr (!SOLVE) {
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                const uint32_t c = lane + q * G;
                if (c < R) out[row * R + c] = mt[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < CPL; q++) mt_s[gi][lane + q * G] = mt[q];
            wave_sync_lds();
            if constexpr (kRegP) {
                // the same sequential sum over k < R as below
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < G; k++)
                    if ((uint32_t)k < R) s = s + mt_s[gi][k] * pc[k];
                if (lane < R) {
                    out[row * R + lane] = s;
                    sq[0] = sq[0] + s * s;
                }
            } else {
#pragma unroll
                for (int q = 0; q < CPL; q++) {
                    const uint32_t c = lane + q * G;
                    if (c < R) {
                        double s = 0.0;
                        for (uint32_t k = 0; k < R; k++) s = s + mt_s[gi][k] * P[k * R + c];
                        out[row * R + c] = s;
                        sq[q] = sq[q] + s * s;
                    }
                }
            }
            wave_sync_lds();  // this row's reads of mt_s before the next row's writes
        }
    }
    if (SOLVE && colsq) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < CPL; q++) sq_s[gi][lane + q * G] = sq[q];
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < R; c += 256) {
            double s = 0.0;
            for (int g = 0; g < GPB; g++) s += sq_s[g][c];
            colsq[(uint64_t)blockIdx.x * R + c] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// column norms (renormalize_columns, low_rank_tensor.hpp:203-260).  Rows are
// rows[0..n) when rows is given (n from *n_dev or n_host), else 0..n.
// ---------------------------------------------------------------------------
template <int BS>
__global__ void __launch_bounds__(BS) k_colsq_partial(const double* __restrict__ U, const uint32_t* __restrict__ rows,
                                                      const uint32_t* n_dev, uint64_t n_host, uint32_t R,
                                                      double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double red[BS];
    const uint64_t n = n_dev ? (uint64_t)*n_dev : n_host;
    const uint32_t lanes = (BS / R);  // rows in flight per block
    const uint32_t rl = threadIdx.x / R, c = threadIdx.x - rl * R;
    double acc = 0.0;
    if (rl < lanes) {
        for (uint64_t i = (uint64_t)blockIdx.x * lanes + rl; i < n; i += (uint64_t)gridDim.x * lanes) {
            const uint64_t r = rows ? (uint64_t)rows[i] : i;
            const double x = U[r * R + c];
            acc += x * x;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < R) {
        double s = 0.0;
        for (uint32_t q = 0; q < lanes; q++) s += red[q * R + threadIdx.x];
        partial[(uint64_t)blockIdx.x * R + threadIdx.x] = s;
    }
}

// norms[c] = sqrt(sum of the partials), 1 when <= 1e-7; sigma either the
// norms or sigma * norms (multiply).  One block of 256 threads: wave w sums
// columns w, w+4, ...; lane l adds partials l, l+64, ... in order, then a
// fixed xor butterfly over the 64 lanes.
__global__ void __launch_bounds__(256) k_colnorm_final(const double* __restrict__ partial, uint32_t nparts,
                                                       uint32_t R, double* __restrict__ norms,
                                                       double* __restrict__ sigma, int multiply) {
#pragma clang fp contract(off)
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t c = wv; c < R; c += 4) {
        double s = 0.0;
        for (uint32_t p = lane; p < nparts; p += 64) s += partial[(uint64_t)p * R + c];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) {
            double nrm = sqrt(s);
            if (nrm <= 1e-7) nrm = 1.0;
            norms[c] = nrm;
            if (sigma) sigma[c] = multiply ? sigma[c] * nrm : nrm;
        }
    }
}

__global__ void k_scale_rows(double* __restrict__ U, const uint32_t* __restrict__ rows, const uint32_t* n_dev,
                             uint64_t n_host, uint32_t R, const double* __restrict__ norms) {
    const uint64_t n = n_dev ? (uint64_t)*n_dev : n_host;
    const uint32_t rpb = blockDim.x / R;  // rows a block takes per step
    if (rpb == 0) {  // R > blockDim: one entry per thread
        for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * R;
             e += (uint64_t)gridDim.x * blockDim.x) {
            const uint64_t i = e / R, c = e - i * R;
            const uint64_t r = rows ? (uint64_t)rows[i] : i;
            U[r * R + c] = U[r * R + c] / norms[c];
        }
        return;
    }
    // thread t: column t % R of row t / R of the block's batch (no index
    // division inside the loop)
    const uint32_t rl = threadIdx.x / R, c = threadIdx.x - rl * R;
    if (rl >= rpb) return;
    const double nc = norms[c];
    for (uint64_t i = (uint64_t)blockIdx.x * rpb + rl; i < n; i += (uint64_t)gridDim.x * rpb) {
        const uint64_t r = rows ? (uint64_t)rows[i] : i;
        U[r * R + c] = U[r * R + c] / nc;
    }
}

__global__ void k_fill(double* __restrict__ p, uint64_t n, double v) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x)
        p[e] = v;
}

// ---------------------------------------------------------------------------
// residual: sum over nonzeros of x^2 - 2 x m, m = sum_r sigma_r prod_k
// U_k[i_k][r] (sort_lookup.hpp:134-197), walking the nonzeros in the order of
// the lookup that leaves out mode s (L.skip): the partial product over the
// modes != s is reused while consecutive nonzeros share those coordinates,
// and each nonzero gathers one row of U_s (the caller picks s with the
// smallest factor, which stays cache resident).  A group of G lanes takes a
// run of G consecutive nonzeros: their keys, coordinates and values are
// fetched one per lane and broadcast; lane r holds column r (CPL per lane).
// ---------------------------------------------------------------------------
template <int G, int CPL, int BS>
__global__ void __launch_bounds__(BS) k_residual_partial(const uint64_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ rows,
                                                         const double* __restrict__ vals, uint64_t nnz,
                                                         KeyLayout L, FactorPtrs F, const double* __restrict__ sigma,
                                                         uint32_t R, double* __restrict__ partial) {
#pragma clang fp contract(off)
    constexpr int GPB = BS / G;
    __shared__ double red[GPB];
    const uint32_t lane = threadIdx.x % G, gi = threadIdx.x / G;
    const uint32_t s = L.skip;
    const uint64_t nruns = (nnz + G - 1) / G;
    const uint64_t groups = (uint64_t)gridDim.x * GPB;
    double acc = 0.0;  // identical in every lane of a group
    for (uint64_t run = (uint64_t)blockIdx.x * GPB + gi; run < nruns; run += groups) {
        const uint64_t base = run * G;
        const int cnt = (int)min<uint64_t>(G, nnz - base);
        uint64_t my_key = 0;
        uint32_t my_row = 0;
        double my_val = 0.0;
        if ((int)lane < cnt) {
            my_key = keys[base + lane];
            my_row = rows[base + lane];
            my_val = vals[base + lane];
        }
        double pp[CPL];
        uint64_t prev = 0;
        for (int t = 0; t < cnt; t++) {
            const uint64_t key = shfl_u64(my_key, t, G);
            const uint64_t row = __shfl(my_row, t, G);
            const double x = __shfl(my_val, t, G);
            if (t == 0 || key != prev) {
#pragma unroll
                for (int q = 0; q < CPL; q++) {
                    const uint32_t c = lane + q * G;
                    double p = 0.0;
                    if (c < R) {
                        p = sigma[c];
                        for (uint32_t k = 0; k < L.N; k++)
                            if (k != s) p = p * F.U[k][(uint64_t)key_field(L, key, k) * R + c];
                    }
                    pp[q] = p;
                }
                prev = key;
            }
            double part = 0.0;
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                const uint32_t c = lane + q * G;
                if (c < R) part = part + pp[q] * F.U[s][row * R + c];
            }
            const double m = group_sum<G>(part);
            acc = acc + (x * x - 2.0 * m * x);
        }
    }
    if (lane == 0) red[gi] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int g = 0; g < GPB; g++) t += red[g];
        partial[blockIdx.x] = t;
    }
}

// Sum of nparts block partials of packed Grams, per entry (grid over
// nmodes x T entries; each thread adds its entry's partials in order).
__global__ void k_gram_reduce(const double* __restrict__ parts, uint32_t nparts, uint32_t T, uint32_t nmodes,
                              double* __restrict__ out) {
#pragma clang fp contract(off)
    const uint64_t n = (uint64_t)nmodes * T;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = w / T, e = w - k * T;
        const double* P = parts + k * nparts * T;
        double s = 0.0;
        for (uint32_t p = 0; p < nparts; p++) s += P[(uint64_t)p * T + e];
        out[w] = s;
    }
}

// ||M||^2 = sum_{a,b} sigma_a sigma_b prod_k G_k[a][b] (common.h:152-226) from
// the reduced packed Grams G (nmodes x T); one block of 256 threads.
__global__ void __launch_bounds__(256) k_chain_final(const double* __restrict__ G, uint32_t N, uint32_t R,
                                                     const double* __restrict__ sigma, double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double red[256];
    const uint32_t T = R * (R + 1) / 2;
    double acc = 0.0;
    for (uint32_t e = threadIdx.x; e < T; e += blockDim.x) {
        uint32_t a = 0, base = 0;
        while (base + (R - a) <= e) {
            base += R - a;
            a++;
        }
        const uint32_t b = a + (e - base);
        double v = sigma[a] * sigma[b];
        for (uint32_t k = 0; k < N; k++) v = v * G[(uint64_t)k * T + e];
        acc += a == b ? v : 2.0 * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// Larsen-Kolda leverage sampler (larsen_kolda_sampler.hpp): per factor its
// Gram U^T U (k_gram_mfma), the pinv of the Gram (k_pinv_partials), the row
// leverage scores lev_i = u_i^T pinv u_i (compute_DAGAT, sampler.hpp:94-122;
// k_lev_rows) with their sums over tiles of 64 rows, the prefix sums of the
// scores (k_tile_offsets, k_cum_rows), and per sample and mode an independent
// inverse-CDF draw with the log-space weight of larsen_kolda_sampler.hpp:105-128
// (k_lk_draw).  Every sum has a fixed association, so the scores, prefix sums
// and draws do not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr uint32_t kLevTile = 64;  // rows of one leverage tile: one wave, one row per lane

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Gram of a tall factor (I x R, R <= 16 NB) on the FP64 matrix cores:
// v_mfma_f64_16x16x4f64 takes A (16 x 4) with A[i][k] in lane 16k + i and
// B (4 x 16) with B[k][j] in lane 16k + j, and accumulates D (16 x 16) with
// D[4v + l/16][l%16] in item v of lane l.  With lane l holding
// U[r0 + l/16][16a + l%16] as x_a, mfma(x_a, x_b) adds the four rows' share of
// the Gram block (a, b); only blocks a <= b are formed.  Block b0 sums rows
// [b0 chunk, (b0+1) chunk) (chunk a multiple of 16); wave w of it takes the
// 4-row groups r0 = start + 4w + 16s; the four waves' blocks are added in
// wave order.  partial[b0] holds the packed upper triangle (entry (r, c),
// r <= c, at r R - r(r-1)/2 + c - r), the layout k_pinv_partials reduces.
template <int NB>
__global__ void __launch_bounds__(256) k_gram_mfma(const double* __restrict__ U, uint64_t I, uint32_t R,
                                                   double* __restrict__ partial) {
    constexpr int NT = NB * (NB + 1) / 2;
    constexpr int UNR = NB <= 2 ? 8 : 4;  // 16-row steps per loop trip (loads in flight)
    __shared__ double red[4][4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t chunk = (((I + gridDim.x - 1) / gridDim.x) + 15) & ~15ull;
    const uint64_t rb = (uint64_t)blockIdx.x * chunk;
    const uint64_t re = min<uint64_t>(I, rb + chunk);
    const int kr = lane >> 4, ci = lane & 15;
    f64x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (uint64_t r0 = rb + 4 * w; r0 < re; r0 += 16 * UNR) {
        double x[UNR][NB];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint64_t row = r0 + 16 * u + kr;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const uint32_t col = 16 * b + ci;
                x[u][b] = (row < re && col < R) ? U[row * R + col] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            int t = 0;
#pragma unroll
            for (int a = 0; a < NB; a++)
#pragma unroll
                for (int b = a; b < NB; b++) {
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[u][a], x[u][b], acc[t], 0, 0, 0);
                    t++;
                }
        }
    }
    double* out = partial + (uint64_t)blockIdx.x * (R * (R + 1) / 2);
    int t = 0;
#pragma unroll
    for (int a = 0; a < NB; a++)
#pragma unroll
        for (int b = a; b < NB; b++) {
#pragma unroll
            for (int v = 0; v < 4; v++) red[w][v][lane] = acc[t][v];
            __syncthreads();
            if (w == 0) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const double sum = ((red[0][v][lane] + red[1][v][lane]) + red[2][v][lane]) + red[3][v][lane];
                    const uint32_t r = 16 * a + 4 * v + (lane >> 4), c = 16 * b + (lane & 15);
                    if (r < R && c < R && r <= c) out[r * R - r * (r - 1) / 2 + c - r] = sum;
                }
            }
            __syncthreads();
            t++;
        }
}

// P (R x R) into Pp (RM x RM), zero outside R x R.
__global__ void k_pad_square(const double* __restrict__ P, uint32_t R, uint32_t RM, double* __restrict__ Pp) {
    for (uint32_t e = threadIdx.x; e < RM * RM; e += blockDim.x) {
        const uint32_t r = e / RM, c = e - r * RM;
        Pp[e] = (r < R && c < R) ? P[r * R + c] : 0.0;
    }
}

// Row leverage scores of an I x R factor (R <= RMAX), one row per lane: wave
// w of block b takes tile t = 4b + w (64 rows; the grid covers all tiles, one
// per wave, so the scalar loads of Pp stay in straight-line code instead of
// being hoisted out of a tile loop into spilled registers), stages it through
// LDS with coalesced loads (16-byte column pairs when R is even; row stride
// RMAX + 2 doubles, so the row reads are free of bank conflicts), and lane l
// forms
// lev = sum_k u_k (P_kk u_k + 2 sum_{c>k} P_kc u_c) from its row in registers
// and the wave-uniform entries of Pp (RMAX x RMAX, zero outside R x R: padded
// columns contribute exact zeros).  tsum[t] = the inclusive 64-lane scan's
// last value over tile t (rows past I count 0), the same scan k_cum_rows
// takes, so the tile offsets and the prefix sums agree bitwise.
// Two blocks per CU up to RMAX = 32; the 64-column tile of RMAX = 64 needs
// the registers of one.
template <int RMAX>
__global__ void __launch_bounds__(256, (RMAX <= 32 ? 2 : 1)) k_lev_rows(const double* __restrict__ U, uint64_t I, uint32_t R,
                                                     const double* __restrict__ Pp, double* __restrict__ lev,
                                                     double* __restrict__ tsum) {
    constexpr int ST = RMAX + 2;
    constexpr int RPI = 64 / RMAX;  // rows per staging load
    __shared__ __attribute__((aligned(16))) double stage[4][kLevTile * ST];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double* tile = stage[w];
    const uint64_t ntiles = (I + kLevTile - 1) / kLevTile;
    const uint64_t t = (uint64_t)blockIdx.x * 4 + w;
    if (t >= ntiles) return;
    const uint64_t row0 = t * kLevTile;
    if ((R & 1u) == 0 && ((uintptr_t)U & 15u) == 0) {
        // column pairs, one 16-byte load each (R even and U 16-byte aligned
        // keep every pair aligned): RMAX / 2 loads per tile
        constexpr int PPR = RMAX / 2;   // pairs per row
        constexpr int RPL = 64 / PPR;   // rows per load
        const int lr = lane / PPR, lp = lane % PPR;
        double2 v[PPR];
#pragma unroll
        for (int q = 0; q < PPR; q++) {
            const uint64_t row = row0 + q * RPL + lr;
            const uint32_t c = 2u * (uint32_t)lp;
            v[q] = (row < I && c < R) ? *reinterpret_cast<const double2*>(U + row * R + c) : double2{0.0, 0.0};
        }
#pragma unroll
        for (int q = 0; q < PPR; q++) *reinterpret_cast<double2*>(tile + (q * RPL + lr) * ST + 2 * lp) = v[q];
    } else {
        const int lr = lane / RMAX, lc = lane % RMAX;
        double v[RMAX];
#pragma unroll
        for (int q = 0; q < RMAX; q++) {
            const uint64_t row = row0 + q * RPI + lr;
            v[q] = (row < I && (uint32_t)lc < R) ? U[row * R + lc] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < RMAX; q++) tile[(q * RPI + lr) * ST + lc] = v[q];
    }
    wave_sync_lds();
    double u[RMAX];
#pragma unroll
    for (int c = 0; c < RMAX; c += 2) {
        const double2 p = *reinterpret_cast<const double2*>(tile + lane * ST + c);
        u[c] = p.x;
        u[c + 1] = p.y;
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < RMAX; k++) {
        double sk = 0.0;
#pragma unroll
        for (int c = k + 1; c < RMAX; c++) sk = __builtin_fma(Pp[k * RMAX + c], u[c], sk);
        acc = __builtin_fma(u[k], __builtin_fma(Pp[k * RMAX + k], u[k], 2.0 * sk), acc);
    }
    const bool valid = row0 + lane < I;
    if (valid) lev[row0 + lane] = acc;
    const double sc = group_incl_scan<64>(valid ? acc : 0.0, lane);
    if (lane == 63) tsum[t] = sc;
}

// Leverage scores for R > 64: R-lane groups, one row per group (lanes take
// columns c = lane + q G; y_c = (u^T P)_c, P symmetric), group sum.
template <int G, int CPL>
__global__ void __launch_bounds__(256) k_row_leverage(const double* __restrict__ U, uint64_t I, uint32_t R,
                                                      const double* __restrict__ P, double* __restrict__ lev) {
    const uint32_t lane = threadIdx.x % G;
    const uint64_t groups = (uint64_t)gridDim.x * (256 / G);
    for (uint64_t i = (uint64_t)blockIdx.x * (256 / G) + threadIdx.x / G; i < I; i += groups) {
        const double* u = U + i * R;
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            const uint32_t c = lane + q * G;
            if (c < R) {
                double y = 0.0;
                for (uint32_t k = 0; k < R; k++) y = __builtin_fma(u[k], P[k * R + c], y);
                part = __builtin_fma(u[c], y, part);
            }
        }
        const double s = group_sum<G>(part);
        if (lane == 0) lev[i] = s;
    }
}

// tsum[t] for the scores of k_row_leverage: the 64-lane inclusive scan of
// tile t's scores (rows past I count 0), last value.
__global__ void __launch_bounds__(256) k_tile_sums(const double* __restrict__ lev, uint64_t I,
                                                   double* __restrict__ tsum) {
    const int lane = threadIdx.x & 63;
    const uint64_t ntiles = (I + kLevTile - 1) / kLevTile;
    for (uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += (uint64_t)gridDim.x * 4) {
        const uint64_t i = t * kLevTile + lane;
        const double sc = group_incl_scan<64>(i < I ? lev[i] : 0.0, lane);
        if (lane == 63) tsum[t] = sc;
    }
}

// Exclusive tile offsets off[t] = off[t-1] + tsum[t-1] (off[0] = 0), taken in
// order: thread h of one 1024-thread block owns tiles [h per, (h+1) per); the
// threads' sums are scanned in thread order by thread 0; every thread then
// runs its tiles from its offset.  *total = the running value after the last
// tile, which equals the last prefix sum k_cum_rows writes.
__global__ void __launch_bounds__(1024) k_tile_offsets(const double* __restrict__ tsum, uint64_t ntiles,
                                                       double* __restrict__ off, double* __restrict__ total) {
    __shared__ double part[1024];
    const uint64_t per = (ntiles + 1023) / 1024;
    const uint64_t tb = min<uint64_t>(ntiles, threadIdx.x * per), te = min<uint64_t>(ntiles, tb + per);
    // sums in tile order; loads issued 16 at a time
    double s = 0.0;
    uint64_t t0 = tb;
    for (; t0 + 16 <= te; t0 += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = tsum[t0 + u];
#pragma unroll
        for (int u = 0; u < 16; u++) s = s + v[u];
    }
    for (; t0 < te; t0++) s = s + tsum[t0];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int h = 0; h < 1024; h++) {
            const double x = part[h];
            part[h] = run;
            run = run + x;
        }
    }
    __syncthreads();
    double run = part[threadIdx.x];
    uint64_t t1 = tb;
    for (; t1 + 16 <= te; t1 += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = tsum[t1 + u];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            off[t1 + u] = run;
            run = run + v[u];
        }
    }
    for (; t1 < te; t1++) {
        off[t1] = run;
        run = run + tsum[t1];
    }
    if (tb < te && te == ntiles) *total = run;
}

// Tile offsets for many tiles, in segments of kOffSeg tiles (one block of
// 256 threads each; thread h owns 16 consecutive tiles):
//   k_seg_tile_sums   hsum[s 256 + h] = thread h's tiles summed in order,
//                     segsum[s] = the 256 thread sums in thread order;
//   k_tile_offsets    on segsum: segoff (exclusive, in segment order);
//   k_seg_tile_offsets off[t] = segoff[s] + thread sums before h (in order)
//                     + tiles before t in h's range (in order); *total = the
//                     running value after the last tile (the last prefix sum
//                     k_cum_rows writes).
constexpr uint32_t kOffSeg = 4096;

__global__ void __launch_bounds__(256) k_seg_tile_sums(const double* __restrict__ tsum, uint64_t ntiles,
                                                       double* __restrict__ hsum, double* __restrict__ segsum) {
    __shared__ double part[256];
    const uint64_t tb = min<uint64_t>(ntiles, (uint64_t)blockIdx.x * kOffSeg + 16 * threadIdx.x);
    const uint64_t te = min<uint64_t>(ntiles, tb + 16);
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = tb + u < te ? tsum[tb + u] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 16; u++)
        if (tb + u < te) s = s + v[u];
    hsum[(uint64_t)blockIdx.x * 256 + threadIdx.x] = s;
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int h = 0; h < 256; h++) run = run + part[h];
        segsum[blockIdx.x] = run;
    }
}

__global__ void __launch_bounds__(256) k_seg_tile_offsets(const double* __restrict__ tsum, uint64_t ntiles,
                                                          const double* __restrict__ hsum,
                                                          const double* __restrict__ segoff, double* __restrict__ off,
                                                          double* __restrict__ total) {
    __shared__ double base[256];
    const uint64_t tb = min<uint64_t>(ntiles, (uint64_t)blockIdx.x * kOffSeg + 16 * threadIdx.x);
    const uint64_t te = min<uint64_t>(ntiles, tb + 16);
    base[threadIdx.x] = hsum[(uint64_t)blockIdx.x * 256 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = segoff[blockIdx.x];
        for (int h = 0; h < 256; h++) {
            const double x = base[h];
            base[h] = run;
            run = run + x;
        }
    }
    __syncthreads();
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = tb + u < te ? tsum[tb + u] : 0.0;
    double run = base[threadIdx.x];
#pragma unroll
    for (int u = 0; u < 16; u++)
        if (tb + u < te) {
            off[tb + u] = run;
            run = run + v[u];
        }
    if (tb < te && te == ntiles) *total = run;
}

// Block partials of a packed Gram (nparts x T) reduced to 8 chunk sums
// (out[h T + e] = partials [h c, (h+1) c) of entry e in order, c =
// ceil(nparts / 8)), so the one-block pinv kernel reads 8 partials.
__global__ void __launch_bounds__(256) k_partials_to8(const double* __restrict__ partial, uint32_t nparts, uint32_t T,
                                                      double* __restrict__ out) {
    const uint32_t idx = blockIdx.x * 256 + threadIdx.x;
    const uint32_t e = idx >> 3, h = idx & 7;
    if (e >= T) return;
    const uint32_t c = (nparts + 7) / 8;
    const uint32_t pb = min(nparts, h * c), pe = min(nparts, pb + c);
    double s = 0.0;
    uint32_t p = pb;
    for (; p + 16 <= pe; p += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = partial[(uint64_t)(p + u) * T + e];
#pragma unroll
        for (int u = 0; u < 16; u++) s = s + v[u];
    }
    for (; p < pe; p++) s = s + partial[(uint64_t)p * T + e];
    out[(uint64_t)h * T + e] = s;
}

// cum[i] = off[t] + the inclusive 64-lane scan of tile t's scores up to i.
__global__ void __launch_bounds__(256) k_cum_rows(const double* __restrict__ lev, uint64_t I,
                                                  const double* __restrict__ off, double* __restrict__ cum) {
    const int lane = threadIdx.x & 63;
    const uint64_t ntiles = (I + kLevTile - 1) / kLevTile;
    for (uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += (uint64_t)gridDim.x * 4) {
        const uint64_t i = t * kLevTile + lane;
        const double sc = group_incl_scan<64>(i < I ? lev[i] : 0.0, lane);
        if (i < I) cum[i] = off[t] + sc;
    }
}

__global__ void k_set_u32(uint32_t* __restrict__ p, uint32_t v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

struct LKArgs {
    const double* cum[kMaxModes];
    const double* off[kMaxModes];  // tile offsets (cum at the end of the previous tile)
    const double* lev[kMaxModes];
    const double* total;  // N entries
    uint64_t I[kMaxModes];
    int N, j;
    uint64_t J;
    const double* draws;  // uniform of (sample i, mode k) at draws[k J + i], or null
    uint64_t seed, call;
    uint64_t* samples;    // N x J
    double* w;            // J or null
};

// First row of mode k whose prefix sum exceeds x (I_k - 1 if none): the tile
// from its offsets (off[t] equals the prefix sum at the end of tile t - 1, so
// the first row past x lies in the tile before the first offset past x),
// then the row within the tile's 64 prefix sums.
__device__ inline uint64_t lk_search(const double* __restrict__ off, const double* __restrict__ cum, uint64_t I,
                                     double x) {
    const uint64_t nt = (I + kLevTile - 1) / kLevTile;
    uint64_t lo = 1, hi = nt;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (off[mid] > x) hi = mid;
        else lo = mid + 1;
    }
    const uint64_t t = lo - 1;
    uint64_t a = t * kLevTile, b = min<uint64_t>(I, a + kLevTile);
    while (a < b) {
        const uint64_t mid = (a + b) >> 1;
        if (cum[mid] > x) b = mid;
        else a = mid + 1;
    }
    return a < I ? a : I - 1;
}

// Sample i: per mode k != j, u in [0,1), the first row whose inclusive prefix
// sum exceeds u * total (rows of zero leverage are never drawn), and the
// weight exp(-log J + sum_k (log total_k - log lev_k[row])).
__global__ void k_lk_draw(LKArgs a) {
#pragma clang fp contract(off)
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.J; i += (uint64_t)gridDim.x * blockDim.x) {
        double logw = 0.0 - log((double)a.J);
        for (int k = 0; k < a.N; k++) {
            if (k == a.j) continue;
            const double u = a.draws ? a.draws[(uint64_t)k * a.J + i] : philox_uniform(a.seed, a.call, i, 2 * k);
            const double tot = a.total[k];
            const uint64_t row = lk_search(a.off[k], a.cum[k], a.I[k], u * tot);
            a.samples[(uint64_t)k * a.J + i] = row;
            logw = logw + (log(tot) - log(a.lev[k][row]));
        }
        if (a.w) a.w[i] = exp(logw);
    }
}

// ---------------------------------------------------------------------------
// Larsen-Kolda hybrid sampler (larsen_kolda_hybrid.hpp): tau = 1/J; alpha_k =
// max_i lev_k[i] / total_k; the tuples whose product probability
// p = prod_{k != j} lev_k / total_k reaches tau are taken deterministically
// (weight 1), the other samples by rejection from the product distribution
// (weight (1 - p_det) / (J p), p_det the mass of the deterministic tuples).
// Candidates of mode k are the rows with lev_k >= (tau alpha_k / alpha*) total_k
// (alpha* = prod_{k != j} alpha_k), listed in ascending row order; tuple c
// of the candidate product has mixed-radix digits over k != j ascending (the
// lowest mode varies fastest).
// ---------------------------------------------------------------------------
constexpr uint32_t kHybChunk = 4096;        // candidates per enumeration block (16 rounds of 256)
constexpr uint32_t kHybMaxAttempts = 65536; // rejection attempts per sample

struct HybArgs {
    const uint64_t* cand[kMaxModes];
    const uint64_t* ncand;  // N
    const double* lev[kMaxModes];
    const double* total;    // N
    int N, j;
    uint64_t J;
    uint64_t C;             // candidates (product of ncand over k != j)
};

// per block: maxima of lev over a contiguous chunk
__global__ void __launch_bounds__(256) k_block_max(const double* __restrict__ v, uint64_t n,
                                                   double* __restrict__ part) {
    __shared__ double red[256];
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b = (uint64_t)blockIdx.x * chunk, e = min<uint64_t>(n, b + chunk);
    double m = -INFINITY;
    for (uint64_t i = b + threadIdx.x; i < e; i += 256) m = fmax(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// alpha[k] = max over the block maxima / total[k]
__global__ void k_alpha(const double* __restrict__ part, uint32_t nb, const double* __restrict__ total, uint32_t k,
                        double* __restrict__ alpha) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double m = -INFINITY;
    for (uint32_t b = 0; b < nb; b++) m = fmax(m, part[b]);
    alpha[k] = m / total[k];
}

// cut[k] = (tau alpha_k / alpha*) total_k for k != j (larsen_kolda_hybrid.hpp:150-165)
__global__ void k_hyb_cuts(const double* __restrict__ alpha, const double* __restrict__ total, int N, int j,
                           uint64_t J, double* __restrict__ cut) {
#pragma clang fp contract(off)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double tau = 1.0 / (double)J;
    double astar = 1.0;
    for (int k = 0; k < N; k++)
        if (k != j) astar = astar * alpha[k];
    for (int k = 0; k < N; k++) cut[k] = k == j ? 0.0 : tau * alpha[k] / astar * total[k];
}

// rows of tuple c; log p of the tuple (sum over k != j of log lev - log total)
__device__ inline double hyb_tuple(const HybArgs& a, uint64_t c, uint64_t* rows) {
#pragma clang fp contract(off)
    double wlog = 0.0;
    for (int k = 0; k < a.N; k++) {
        if (k == a.j) continue;
        const uint64_t nk = a.ncand[k];
        const uint64_t d = c % nk;
        c /= nk;
        const uint64_t row = a.cand[k][d];
        rows[k] = row;
        wlog = wlog + (log(a.lev[k][row]) - log(a.total[k]));
    }
    return wlog;
}

// Fixed-shape exclusive scan of 256 thread values in LDS (Hillis-Steele on
// a double buffer); returns the exclusive prefix, *tot the block total.
__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t* buf, uint32_t* tot) {
    const int t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t x = t >= o ? buf[t - o] : 0u;
        __syncthreads();
        buf[t] += x;
        __syncthreads();
    }
    const uint32_t incl = buf[t];
    *tot = buf[255];
    __syncthreads();
    return incl - v;
}

// Pass 1 over the candidates: block b takes [b kHybChunk, (b+1) kHybChunk)
// in 16 rounds of 256; cnt[b] = tuples kept (p >= tau); pp[b] = their p
// summed per round by a fixed tree, rounds in order.
__global__ void __launch_bounds__(256) k_hyb_count(HybArgs a, uint32_t* __restrict__ cnt, double* __restrict__ pp) {
#pragma clang fp contract(off)
    __shared__ double red[256];
    const double tau = 1.0 / (double)a.J;
    uint32_t kept = 0;
    double psum = 0.0;
    for (int r = 0; r < (int)(kHybChunk / 256); r++) {
        const uint64_t c = (uint64_t)blockIdx.x * kHybChunk + (uint64_t)r * 256 + threadIdx.x;
        double p = 0.0;
        if (c < a.C) {
            uint64_t rows[kMaxModes];
            const double e = exp(hyb_tuple(a, c, rows));
            if (e >= tau) {
                p = e;
                kept++;
            }
        }
        red[threadIdx.x] = p;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) psum = psum + red[0];
        __syncthreads();
    }
    __shared__ uint32_t kc[256];
    kc[threadIdx.x] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int t = 0; t < 256; t++) s += kc[t];
        cnt[blockIdx.x] = s;
        pp[blockIdx.x] = psum;
    }
}

// off[b] = exclusive sum of cnt (in block order); *ndet = the total
__global__ void k_hyb_offsets(const uint32_t* __restrict__ cnt, uint32_t nb, uint64_t* __restrict__ off,
                              uint64_t* __restrict__ ndet) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint64_t run = 0;
    for (uint32_t b = 0; b < nb; b++) {
        off[b] = run;
        run += cnt[b];
    }
    *ndet = run;
}

// Pass 2: the kept tuples in candidate order at positions off[b] + ...;
// rows k != j written (positions < J only), weight 1.
__global__ void __launch_bounds__(256) k_hyb_write(HybArgs a, const uint64_t* __restrict__ off,
                                                   uint64_t* __restrict__ samples, double* __restrict__ w) {
    __shared__ uint32_t buf[256];
    const double tau = 1.0 / (double)a.J;
    uint64_t base = off[blockIdx.x];
    for (int r = 0; r < (int)(kHybChunk / 256); r++) {
        const uint64_t c = (uint64_t)blockIdx.x * kHybChunk + (uint64_t)r * 256 + threadIdx.x;
        uint64_t rows[kMaxModes];
        bool keep = false;
        if (c < a.C) keep = exp(hyb_tuple(a, c, rows)) >= tau;
        uint32_t tot;
        const uint32_t ex = block_excl_scan_u32(keep ? 1u : 0u, buf, &tot);
        if (keep) {
            const uint64_t pos = base + ex;
            if (pos < a.J) {
                for (int k = 0; k < a.N; k++)
                    if (k != a.j) samples[(uint64_t)k * a.J + pos] = rows[k];
                if (w) w[pos] = 1.0;
            }
        }
        base += tot;
    }
}

__global__ void k_hyb_zero(uint64_t* __restrict__ ndet, double* __restrict__ pdet) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *ndet = 0;
        *pdet = 0.0;
    }
}

// Samples i >= ndet: tuples from the product distribution until p < tau
// (larsen_kolda_hybrid.hpp:213-241); weight (1 - p_det) exp(-log J - log p).
// Attempt 0 takes the uniform of (i, k) from draws (when given) or the
// built-in stream at stage 2k; attempt a >= 1 the built-in stream at stage
// 2k of call (call XOR a 2^40).  A sample still rejected after
// kHybMaxAttempts attempts keeps its last tuple with weight 0 and counts in
// *fail: the mass below tau it failed to find is then below what that many
// attempts reach (all of it is deterministic when p_det = 1, where the
// reference's loop never ends).
__global__ void k_hyb_reject(LKArgs a, const uint64_t* __restrict__ ndet, const double* __restrict__ pdet,
                             uint32_t* __restrict__ fail) {
#pragma clang fp contract(off)
    const double tau = 1.0 / (double)a.J;
    const uint64_t nd = *ndet;
    const double scale = 1.0 - *pdet;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.J; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < nd) continue;
        double wlog = 0.0;
        uint32_t att = 0;
        bool exhausted = false;
        for (;; att++) {
            wlog = 0.0;
            for (int k = 0; k < a.N; k++) {
                if (k == a.j) continue;
                double u;
                if (att == 0)
                    u = a.draws ? a.draws[(uint64_t)k * a.J + i] : philox_uniform(a.seed, a.call, i, 2 * k);
                else
                    u = philox_uniform(a.seed, a.call ^ ((uint64_t)att << 40), i, 2 * k);
                const double tot = a.total[k];
                const uint64_t row = lk_search(a.off[k], a.cum[k], a.I[k], u * tot);
                a.samples[(uint64_t)k * a.J + i] = row;
                wlog = wlog + (log(a.lev[k][row]) - log(tot));
            }
            if (exp(wlog) < tau) break;
            if (att + 1 >= kHybMaxAttempts) {
                atomicAdd(fail, 1u);
                exhausted = true;
                break;
            }
        }
        if (a.w) a.w[i] = exhausted ? 0.0 : scale * exp((0.0 - log((double)a.J)) - wlog);
    }
}

// h = product over modes k != j of the sampled rows (fill_h_by_samples,
// sampler.hpp:75-92)
__global__ void k_fill_h(const uint64_t* __restrict__ samples, uint64_t J, uint32_t R, int N, int j, FactorPtrs F,
                         double* __restrict__ h) {
#pragma clang fp contract(off)
    const uint64_t n = J * R;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / R, c = e - i * R;
        double v = 1.0;
        for (int k = 0; k < N; k++)
            if (k != j) v = v * F.U[k][samples[(uint64_t)k * J + i] * R + c];
        h[e] = v;
    }
}

}  // namespace ftl

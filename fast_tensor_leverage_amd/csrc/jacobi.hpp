// jacobi.hpp -- block-level cyclic Jacobi eigensolver for small symmetric
// matrices (R <= 128), used by computeM and by the sampled-Gram pseudo-inverse
// of the ALS step.
#pragma once

#include "ftl_common.hpp"

namespace ftl {

// ---------------------------------------------------------------------------
// Cyclic Jacobi eigensolver for one symmetric R x R matrix held by a block
// (A in LDS, V in LDS or global scratch).  Round-robin pair order; a round
// first computes its R/2 rotations from the current A (one thread each),
// then applies J^T A J in place: each thread owns the 2x2 blocks where a row
// pair meets a column pair and applies the row rotation, then the column
// rotation (the same operations, in the same order, as applying all row
// rotations and then all column rotations); V is rotated in place, stored
// transposed (buffer row v holds eigenvector column v, so the entries one
// rotation touches are contiguous).  Rotation
// annihilating A[p][q] (Golub & Van Loan Alg. 8.5.2), skipped when
// |A[p][q]| <= eps*sqrt(|A[p][p] A[q][q]|).  Products and sums are rounded
// individually (no contraction).  On return lam is ascending and Vout column
// v holds the eigenvector of lam[v].
// ---------------------------------------------------------------------------
struct JacobiShared {
    int p[64], q[64], act[64];
    double c[64], s[64];
};

// 1/x and 1/sqrt(x): v_rcp_f64 / v_rsq_f64 estimates plus one Newton step.
__device__ inline double fast_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ inline double fast_rsq(double x) {
    double r = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x * r;
    const double e = fma(-h, r, 0.5);
    return fma(r, e, r);
}

// Pair k of round `round` in the circle-method ordering of n2 (even) indices:
// position 0 fixed, positions 1..n2-1 rotating; returned with p < q.
__device__ inline void jacobi_pair(int k, int round, int n2, int& p, int& q) {
    p = k == 0 ? 0 : 1 + ((k - 1 + round) % (n2 - 1));
    const int kk = n2 - 1 - k;
    q = kk == 0 ? 0 : 1 + ((kk - 1 + round) % (n2 - 1));
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
}

// Rotation (c, s) annihilating A[p][q]; false (c = 1, s = 0) when the entry
// is negligible or q is the padding index of an odd size.
__device__ inline bool jacobi_rotation(const double* A, int LD, int R, int p, int q, double& c, double& s) {
#pragma clang fp contract(off)
    c = 1.0;
    s = 0.0;
    if (q >= R) return false;
    const double eps = 2.220446049250313e-16;
    const double apq = A[p * LD + q], app = A[p * LD + p], aqq = A[q * LD + q];
    // |apq| <= eps sqrt(|app aqq|), compared in squares (no square root on
    // the round's critical path)
    if (apq * apq <= (eps * eps) * fabs(app * aqq) || apq == 0.0) return false;
    // hardware reciprocal / reciprocal square root, each refined by one
    // Newton step (error ~1 ulp; the rotation stays orthogonal to working
    // precision)
    const double tau = (aqq - app) * fast_rcp(2.0 * apq);
    double t;
    if (fabs(tau) > 1e150) {
        t = 0.5 * fast_rcp(tau);
    } else {
        const double x = 1.0 + tau * tau;
        const double sq = x * fast_rsq(x);
        t = (tau >= 0 ? 1.0 : -1.0) * fast_rcp(fabs(tau) + sq);
    }
    c = fast_rsq(1.0 + t * t);
    s = t * c;
    return true;
}

// Returns false (in every thread) when the sweep cap ended the iteration
// before every pair met the skip test.
// Sweep cap of the cyclic Jacobi iteration (quadratic convergence takes
// 6-10 sweeps for R <= 128 in FP64).
constexpr int kJacobiMaxSweeps = 40;

// Eigensolves of this translation unit's kernels that stopped at the sweep
// cap (one device counter per source file and device; read on the host with
// jacobi_capped_read).
static __device__ uint32_t g_jacobi_capped = 0;

static inline uint64_t jacobi_capped_read() {
    uint32_t v = 0;
    FTL_HIP_CHECK(hipDeviceSynchronize());
    FTL_HIP_CHECK(hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_jacobi_capped), sizeof(v), 0, hipMemcpyDeviceToHost));
    return v;
}

template <int BS>
__device__ __forceinline__ bool jacobi_block(double* A, double* V, int R, double* lam, double* Vout, JacobiShared& js,
                             double* diag_tmp) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    const int LD = R + 1;  // odd leading dimension: rows start in different LDS banks
    const int n2 = R + (R & 1);
    const int npairs = n2 / 2;
    for (int e = tid; e < R * R; e += BS) V[(e / R) * LD + e % R] = (e / R == e % R) ? 1.0 : 0.0;
    __syncthreads();
    bool converged = false;
    const int nblk = npairs * npairs, nv = R * npairs;
    // When the block holds at most one 2x2-block item and two V items per
    // thread (R <= 64 at 1024 threads), their indices are the same in every
    // round and are computed once; the thread then issues the loads of all
    // its items before any arithmetic, one dependent LDS chain per round.
    const bool fused = nblk <= BS && nv <= 2 * BS;
    // (pair fk, pair fk2): the block item; (pair fk0, row fi0) and (pair fk1,
    // row fi1): the V items.  Clamped to 0 past the item counts so that every
    // js read stays inside its 64 entries.
    const int fk = tid < nblk ? tid / npairs : 0, fk2 = tid < nblk ? tid - (tid / npairs) * npairs : 0;
    // V entries: item it is row it % R of pair it / R (V is stored
    // transposed, so consecutive lanes touch consecutive entries of a column)
    const int fk0 = tid < nv ? tid / R : 0, fi0 = tid < nv ? tid - (tid / R) * R : 0;
    const int fk1 = tid + BS < nv ? (tid + BS) / R : 0, fi1 = tid + BS < nv ? tid + BS - ((tid + BS) / R) * R : 0;
    for (int sweep = 0; sweep < kJacobiMaxSweeps; sweep++) {
        int any = 0;
        // circle-method positions of pair tid (jacobi_pair), advanced by one
        // (mod n2 - 1) per round instead of recomputed by division
        int pa = tid - 1, pb = n2 - 2 - tid;
        for (int round = 0; round + 1 < n2; round++) {
            // rotations of this round, each computed once from the current A
            if (tid < npairs) {
                int p = tid == 0 ? 0 : 1 + pa, q = 1 + pb;  // tid < n2/2: q's position is never the fixed one
                if (p > q) {
                    const int t = p;
                    p = q;
                    q = t;
                }
                double c, s;
                const bool act = jacobi_rotation(A, LD, R, p, q, c, s);
                js.p[tid] = p;
                js.q[tid] = q;
                js.c[tid] = c;
                js.s[tid] = s;
                js.act[tid] = act;
                any |= act;
            }
            pa = pa + 1 == n2 - 1 ? 0 : pa + 1;
            pb = pb + 1 == n2 - 1 ? 0 : pb + 1;
            __syncthreads();
            if (fused) {
                // J^T A J on block (fk, fk2) and V J on entries (fi0, pair fk0)
                // and (fi1, pair fk1)
                const bool bon = tid < nblk, v0on = tid < nv, v1on = tid + BS < nv;
                const bool act1 = bon && js.act[fk], act2 = bon && js.act[fk2];
                const bool a0 = v0on && js.act[fk0], a1 = v1on && js.act[fk1];
                const int p = js.p[fk], q = js.q[fk], p2 = js.p[fk2], q2 = js.q[fk2];
                const int p0 = js.p[fk0], q0 = js.q[fk0], p3 = js.p[fk1], q3 = js.q[fk1];
                const bool blk = act1 || act2;
                const bool qv = q < R, q2v = q2 < R;
                double app2 = 0.0, apq2 = 0.0, aqp2 = 0.0, aqq2 = 0.0, vp0 = 0.0, vq0 = 0.0, vp1 = 0.0, vq1 = 0.0;
                if (blk) {
                    app2 = A[p * LD + p2];
                    if (q2v) apq2 = A[p * LD + q2];
                    if (qv) aqp2 = A[q * LD + p2];
                    if (qv && q2v) aqq2 = A[q * LD + q2];
                }
                if (a0) {
                    vp0 = V[p0 * LD + fi0];
                    vq0 = V[q0 * LD + fi0];
                }
                if (a1) {
                    vp1 = V[p3 * LD + fi1];
                    vq1 = V[q3 * LD + fi1];
                }
                if (blk) {
                    const double c = js.c[fk], s = js.s[fk], c2 = js.c[fk2], s2 = js.s[fk2];
                    double rpp2 = app2, rpq2 = apq2, rqp2 = aqp2, rqq2 = aqq2;
                    if (act1) {
                        rpp2 = c * app2 - s * aqp2;
                        rqp2 = s * app2 + c * aqp2;
                        rpq2 = c * apq2 - s * aqq2;
                        rqq2 = s * apq2 + c * aqq2;
                    }
                    double npp2 = rpp2, npq2 = rpq2, nqp2 = rqp2, nqq2 = rqq2;
                    if (act2) {
                        npp2 = c2 * rpp2 - s2 * rpq2;
                        npq2 = s2 * rpp2 + c2 * rpq2;
                        nqp2 = c2 * rqp2 - s2 * rqq2;
                        nqq2 = s2 * rqp2 + c2 * rqq2;
                    }
                    A[p * LD + p2] = npp2;
                    if (q2v) A[p * LD + q2] = npq2;
                    if (qv) A[q * LD + p2] = nqp2;
                    if (qv && q2v) A[q * LD + q2] = nqq2;
                }
                if (a0) {
                    const double c = js.c[fk0], s = js.s[fk0];
                    V[p0 * LD + fi0] = c * vp0 - s * vq0;
                    V[q0 * LD + fi0] = s * vp0 + c * vq0;
                }
                if (a1) {
                    const double c = js.c[fk1], s = js.s[fk1];
                    V[p3 * LD + fi1] = c * vp1 - s * vq1;
                    V[q3 * LD + fi1] = s * vp1 + c * vq1;
                }
                __syncthreads();
                continue;
            }
            // J^T A J in place: the 2x2 block where row pair k meets column
            // pair k2 belongs to one thread, which applies the row rotation
            // and then the column rotation (the operation order of applying
            // all row rotations, then all column rotations)
            for (int it = tid; it < npairs * npairs; it += BS) {
                const int k = it / npairs, k2 = it - (it / npairs) * npairs;
                const bool act1 = js.act[k], act2 = js.act[k2];
                if (!act1 && !act2) continue;
                const int p = js.p[k], q = js.q[k], p2 = js.p[k2], q2 = js.q[k2];
                const double c = js.c[k], s = js.s[k], c2 = js.c[k2], s2 = js.s[k2];
                // an index of value R (padding of an odd size) only occurs in
                // inactive pairs; its entries are treated as absent
                const bool qv = q < R, q2v = q2 < R;
                const double app2 = A[p * LD + p2];
                const double apq2 = q2v ? A[p * LD + q2] : 0.0;
                const double aqp2 = qv ? A[q * LD + p2] : 0.0;
                const double aqq2 = qv && q2v ? A[q * LD + q2] : 0.0;
                double rpp2 = app2, rpq2 = apq2, rqp2 = aqp2, rqq2 = aqq2;
                if (act1) {
                    rpp2 = c * app2 - s * aqp2;
                    rqp2 = s * app2 + c * aqp2;
                    rpq2 = c * apq2 - s * aqq2;
                    rqq2 = s * apq2 + c * aqq2;
                }
                double npp2 = rpp2, npq2 = rpq2, nqp2 = rqp2, nqq2 = rqq2;
                if (act2) {
                    npp2 = c2 * rpp2 - s2 * rpq2;
                    npq2 = s2 * rpp2 + c2 * rpq2;
                    nqp2 = c2 * rqp2 - s2 * rqq2;
                    nqq2 = s2 * rqp2 + c2 * rqq2;
                }
                A[p * LD + p2] = npp2;
                if (q2v) A[p * LD + q2] = npq2;
                if (qv) A[q * LD + p2] = nqp2;
                if (qv && q2v) A[q * LD + q2] = nqq2;
            }
            // V J, on the transposed store: entry (i, pair k2), i fastest
            for (int it = tid; it < R * npairs; it += BS) {
                const int k2 = it / R, i = it - (it / R) * R;
                if (!js.act[k2]) continue;
                const int p = js.p[k2], q = js.q[k2];
                const double c = js.c[k2], s = js.s[k2];
                const double vp = V[p * LD + i], vq = V[q * LD + i];
                V[p * LD + i] = c * vp - s * vq;
                V[q * LD + i] = s * vp + c * vq;
            }
            __syncthreads();
        }
        if (!__syncthreads_or(any)) {
            converged = true;
            break;
        }
        // converged when every pair meets the skip test: the next sweep would
        // rotate nothing, so stopping here gives the same result
        int open = 0;
        for (int w = tid; w < R * R; w += BS) {
            const int p = w / R, q = w - (w / R) * R;
            if (p < q) {
                const double apq = A[p * LD + q];
                const double thr = (2.220446049250313e-16 * 2.220446049250313e-16) * fabs(A[p * LD + p] * A[q * LD + q]);
                open |= !(apq * apq <= thr || apq == 0.0);
            }
        }
        if (!__syncthreads_or(open)) {
            converged = true;
            break;
        }
    }
    // stable ascending sort of the diagonal; Vout (R x R, dense) column v is
    // the eigenvector of lam[v]
    for (int v = tid; v < R; v += BS) diag_tmp[v] = A[v * LD + v];
    __syncthreads();
    for (int w = tid; w < R * R; w += BS) {
        const int v = w / R, i = w - v * R;
        const double d = diag_tmp[v];
        int rank = 0;
        for (int u = 0; u < R; u++) {
            const double du = diag_tmp[u];
            rank += (du < d) || (du == d && u < v);
        }
        if (i == 0) lam[rank] = d;
        Vout[i * R + rank] = V[v * LD + i];
    }
    if (!converged && tid == 0) atomicAdd(&g_jacobi_capped, 1u);
    __syncthreads();
    return converged;
}

// Jacobi storage of a block: A ((R+1) R doubles, rows padded to an odd
// length) always in LDS; V ((R+1) R) and the sorted eigenvectors (R^2) in LDS
// as well when all three fit in kJacobiLdsDoubles (R <= 64), otherwise in the
// block's global scratch (R = 128).
constexpr int kJacobiLdsDoubles = 3 * 65 * 64;
constexpr int kJacobiThreads = 1024;

// Block size of the one-block Jacobi kernels.  Measured on computeM at R=32
// (N=3, I=2^23): 1024 threads 80 + 72 us, 256 threads 98 + 127 us -- the
// element-wise phases around the rotations (P_k, V, E, eigen tree) gain
// more from the extra waves than the barriers cost.
inline uint32_t jacobi_threads_for(uint64_t R) { (void)R; return 1024u; }

struct JacobiStore {
    double *A, *V, *vecs;
};

__device__ inline JacobiStore jacobi_store(double* smem, double* work, int R) {
    const int LD = R + 1;
    JacobiStore st;
    st.A = smem;
    if (3 * LD * R <= kJacobiLdsDoubles) {
        st.V = smem + LD * R;
        st.vecs = smem + 2 * LD * R;
    } else {
        st.V = work;
        st.vecs = work + LD * R;
    }
    return st;
}

}  // namespace ftl

// ftl_oracle.cpp -- CPU FP64 restatement of the exact Khatri-Rao leverage-score
// sampler of fast_tensor_leverage.  TEST INFRASTRUCTURE ONLY: it is the checker
// for the HIP library and the timed CPU baseline.  Nothing in the product path
// links, loads or calls this file.
//
// Every function names the reference file:line whose behaviour it restates.
// The reference itself (cpp_ext/*.hpp) cannot be compiled in this image: it
// includes <cblas.h> and <lapacke.h>, which are not installed.  The BLAS and
// LAPACK calls it makes are restated here from their published definitions:
//   cblas_dsyrk / dsymv / dgemv / dsymm  (reference BLAS, netlib 3.x semantics),
//   LAPACKE_dsyev ('V','U'): eigenvalues in ascending order with orthonormal
//   eigenvectors, restated as a cyclic Jacobi eigensolver (Golub & Van Loan,
//   "Matrix Computations" 4th ed., Alg. 8.5.3, round-robin ordering of 8.5.4)
//   followed by an ascending sort.
//
// Parity status of each function (how it is pinned) is recorded in the header
// comment of oracle/oracle.py.

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <limits>
#include <numeric>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Tree shape: common.h:17-33 (divide_and_roundup, log2_round_down) and
// partition_tree.hpp:83-118 (ctor, is_leaf, leaf_idx).
// ---------------------------------------------------------------------------
struct Shape {
    uint64_t n, F, leaf_count, node_count;
    uint64_t lfill_level, lfill_count;
    uint64_t nodes_upto_lfill, nodes_before_lfill, complete_level_offset;
};

Shape make_shape(uint64_t n, uint64_t F) {
    Shape s;
    s.n = n;
    s.F = F;
    s.leaf_count = (n + F - 1) / F;
    s.node_count = 2 * s.leaf_count - 1;
    s.lfill_level = 0;
    s.lfill_count = 1;
    while (s.lfill_count * 2 <= s.leaf_count) {
        s.lfill_level++;
        s.lfill_count *= 2;
    }
    s.nodes_upto_lfill = 2 * s.lfill_count - 1;
    s.nodes_before_lfill = s.lfill_count - 1;
    uint64_t partial_div2 = (s.node_count - s.nodes_upto_lfill) / 2;
    s.complete_level_offset = s.nodes_before_lfill - partial_div2;
    return s;
}

inline bool is_leaf(const Shape& s, uint64_t c) { return 2 * c + 1 >= s.node_count; }

inline uint64_t leaf_idx(const Shape& s, uint64_t c) {
    return c >= s.nodes_upto_lfill ? c - s.nodes_upto_lfill : c - s.complete_level_offset;
}

// ---------------------------------------------------------------------------
// build_tree: partition_tree.hpp:120-170.  G is node_count x R*R, row-major
// per node, upper triangle filled, lower triangle zero (zero-fill at 131-134,
// dsyrk(Upper, Trans) at 143-153, child sums level by level at 156-168).
// ---------------------------------------------------------------------------
void build_tree(const Shape& s, const double* U, uint64_t R, double* G) {
    const uint64_t R2 = R * R;
    const uint64_t first_leaf = s.node_count - s.leaf_count;
    #pragma omp parallel
    {
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < s.node_count * R2; i++) G[i] = 0.0;

        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < s.leaf_count; i++) {
            uint64_t node = first_leaf + i;
            uint64_t idx = leaf_idx(s, node);
            uint64_t rows = std::min(s.F, s.n - idx * s.F);
            double* C = G + node * R2;
            const double* A = U + idx * s.F * R;
            for (uint64_t t = 0; t < rows; t++) {
                const double* row = A + t * R;
                for (uint64_t a = 0; a < R; a++) {
                    double ra = row[a];
                    for (uint64_t b = a; b < R; b++) C[a * R + b] += ra * row[b];
                }
            }
        }

        int64_t start = (int64_t)s.nodes_before_lfill;
        int64_t end = (int64_t)first_leaf;
        for (int64_t level = (int64_t)s.lfill_level; level >= 0; level--) {
            #pragma omp for schedule(static)
            for (int64_t c = start; c < end; c++) {
                for (uint64_t e = 0; e < R2; e++)
                    G[c * R2 + e] += G[(2 * c + 1) * R2 + e] + G[(2 * c + 2) * R2 + e];
            }
            end = start;
            start = ((start + 1) / 2) - 1;
        }
    }
}

// x^T Gsym x with Gsym the symmetric matrix whose upper triangle is stored in
// G: dsymv(Upper) followed by a dot product (partition_tree.hpp:66-81, 201-214).
inline double quad_form(const double* G, const double* x, uint64_t R) {
    double m = 0.0;
    for (uint64_t a = 0; a < R; a++) {
        double y = 0.0;
        for (uint64_t b = 0; b < R; b++) {
            double g = a <= b ? G[a * R + b] : G[b * R + a];
            y += g * x[b];
        }
        m += x[a] * y;
    }
    return m;
}

// ---------------------------------------------------------------------------
// Batched inverse-CDF descent: partition_tree.hpp:232-405.
//   x      : J x R vectors the quadratic forms are taken against ("scaled_h")
//   h      : J x R, updated in place h_i *= U[sample_i]
//   margin : optional, J; min over the sample's comparisons of the distance in
//            uniform units between u and the boundary it was compared with.
// ---------------------------------------------------------------------------
void pt_sample(const Shape& s, const double* G, uint64_t R, const double* U,
               const double* x, double* h, uint64_t* samples, const double* draws,
               uint64_t J, double* margin) {
    const uint64_t R2 = R * R;
    #pragma omp parallel
    {
        std::vector<double> q(s.F);
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < J; i++) {
            const double* xi = x + i * R;
            const double u = draws[i];
            double mg = std::numeric_limits<double>::infinity();
            double m = quad_form(G, xi, R);
            uint64_t c = 0;
            double low = 0.0, high = 1.0;
            for (uint64_t level = 0; level < s.lfill_level; level++) {
                double mL = quad_form(G + (2 * c + 1) * R2, xi, R);
                double cutoff = low + mL / m;
                mg = std::min(mg, std::fabs(u - cutoff));
                if (u <= cutoff) { c = 2 * c + 1; high = cutoff; }
                else             { c = 2 * c + 2; low = cutoff; }
            }
            // Tail level (partition_tree.hpp:304-332): one guarded step for
            // samples that stopped on an internal node of the lfill level.
            if (s.node_count > s.nodes_before_lfill && !is_leaf(s, c)) {
                double mL = quad_form(G + (2 * c + 1) * R2, xi, R);
                double cutoff = low + mL / m;
                mg = std::min(mg, std::fabs(u - cutoff));
                if (u <= cutoff) { c = 2 * c + 1; high = cutoff; }
                else             { c = 2 * c + 2; low = cutoff; }
            }
            const uint64_t idx = leaf_idx(s, c);
            uint64_t res = 0;
            if (s.F > 1) {
                // Leaf row scan (partition_tree.hpp:336-391): q_t = (U_t . x)^2,
                // exclusive prefix, divide by the total, first t < F-1 with
                // frac < P[t+1], else F-1.
                const double frac = (u - low) / (high - low);
                const uint64_t rows = std::min(s.F, s.n - idx * s.F);
                for (uint64_t t = 0; t < s.F; t++) {
                    double v = 0.0;
                    if (t < rows) {
                        const double* row = U + (idx * s.F + t) * R;
                        for (uint64_t b = 0; b < R; b++) v += row[b] * xi[b];
                    }
                    q[t] = v;
                }
                double run = 0.0;
                for (uint64_t t = 0; t < s.F; t++) {
                    double tmp = q[t] * q[t];
                    q[t] = run;
                    run += tmp;
                }
                for (uint64_t t = 0; t < s.F; t++) q[t] /= run;
                res = s.F - 1;
                for (uint64_t t = 0; t + 1 < s.F; t++) {
                    if (frac < q[t + 1]) { res = t; break; }
                }
                for (uint64_t t = 1; t < rows; t++)
                    mg = std::min(mg, std::fabs(frac - q[t]) * (high - low));
                // A partial last leaf with frac >= 1 would return a row past
                // the end of the factor (partition_tree.hpp:353, 373-391);
                // clamp to the last real row.
                if (res >= rows) res = rows - 1;
            }
            const uint64_t sample = idx * s.F + res;
            samples[i] = sample;
            const double* urow = U + sample * R;
            for (uint64_t b = 0; b < R; b++) h[i * R + b] *= urow[b];
            if (margin) margin[i] = std::min(margin[i], mg);
        }
    }
}

// ---------------------------------------------------------------------------
// Symmetric eigensolver standing in for LAPACKE_dsyev('V','U')
// (efficient_krp_sampler.hpp:118, 168; common.h:233): cyclic Jacobi with the
// round-robin pair ordering, rotations chosen to annihilate A[p][q]
// (Golub & Van Loan Alg. 8.5.2), skipped when |A[p][q]| is below
// eps*sqrt(|A[p][p] A[q][q]|) (Demmel & Veselic 1992 stopping rule).
// A: R x R, only the upper triangle is read.  On return lam ascending and
// V (row-major, column v = eigenvector of lam[v]).
// ---------------------------------------------------------------------------
void jacobi_eigh(uint64_t R, const double* Ain, double* lam, double* Vout) {
    const uint64_t n2 = R + (R & 1);
    std::vector<double> A(R * R), V(R * R, 0.0);
    for (uint64_t a = 0; a < R; a++)
        for (uint64_t b = 0; b < R; b++)
            A[a * R + b] = a <= b ? Ain[a * R + b] : Ain[b * R + a];
    for (uint64_t a = 0; a < R; a++) V[a * R + a] = 1.0;

    const double eps = std::numeric_limits<double>::epsilon();
    std::vector<uint64_t> P(n2 / 2), Q(n2 / 2);
    std::vector<double> C(n2 / 2), S(n2 / 2);
    std::vector<char> act(n2 / 2);
    for (int sweep = 0; sweep < 40; sweep++) {
        bool rotated = false;
        for (uint64_t round = 0; round + 1 < n2; round++) {
            // circle method: position 0 fixed, positions 1..n2-1 rotate
            auto pos = [&](uint64_t k) -> uint64_t {
                return k == 0 ? 0 : 1 + ((k - 1 + round) % (n2 - 1));
            };
            for (uint64_t k = 0; k < n2 / 2; k++) {
                uint64_t p = pos(k), q = pos(n2 - 1 - k);
                if (p > q) std::swap(p, q);
                P[k] = p; Q[k] = q; act[k] = 0;
                if (q >= R) continue;  // dummy index of an odd-sized matrix
                double apq = A[p * R + q];
                double app = A[p * R + p], aqq = A[q * R + q];
                if (std::fabs(apq) <= eps * std::sqrt(std::fabs(app * aqq)) || apq == 0.0) continue;
                double tau = (aqq - app) / (2.0 * apq);
                double t;
                if (std::fabs(tau) > 1e150) t = 0.5 / tau;
                else t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double c = 1.0 / std::sqrt(1.0 + t * t);
                C[k] = c; S[k] = t * c; act[k] = 1;
                rotated = true;
            }
            for (uint64_t k = 0; k < n2 / 2; k++) {
                if (!act[k]) continue;
                uint64_t p = P[k], q = Q[k];
                double c = C[k], s = S[k];
                for (uint64_t j = 0; j < R; j++) {
                    double ap = A[p * R + j], aq = A[q * R + j];
                    A[p * R + j] = c * ap - s * aq;
                    A[q * R + j] = s * ap + c * aq;
                }
            }
            for (uint64_t k = 0; k < n2 / 2; k++) {
                if (!act[k]) continue;
                uint64_t p = P[k], q = Q[k];
                double c = C[k], s = S[k];
                for (uint64_t i = 0; i < R; i++) {
                    double ap = A[i * R + p], aq = A[i * R + q];
                    A[i * R + p] = c * ap - s * aq;
                    A[i * R + q] = s * ap + c * aq;
                    double vp = V[i * R + p], vq = V[i * R + q];
                    V[i * R + p] = c * vp - s * vq;
                    V[i * R + q] = s * vp + c * vq;
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<uint64_t> order(R);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint64_t a, uint64_t b) { return A[a * R + a] < A[b * R + b]; });
    for (uint64_t v = 0; v < R; v++) {
        lam[v] = A[order[v] * R + order[v]];
        for (uint64_t i = 0; i < R; i++) Vout[i * R + v] = V[i * R + order[v]];
    }
}

// ---------------------------------------------------------------------------
// Sampler state: EfficientKRPSampler (efficient_krp_sampler.hpp:16-67).
// The factors are aliased, as the reference aliases the NumPy arrays.
// ---------------------------------------------------------------------------
struct Sampler {
    uint64_t N, J, R;
    std::vector<const double*> U;
    std::vector<uint64_t> I;
    std::vector<Shape> gshape;
    std::vector<std::vector<double>> gtree;
    // outputs of the last computeM
    std::vector<double> pinv;                  // R x R, upper valid
    std::vector<std::vector<double>> E;        // per mode, R x R rows
    std::vector<std::vector<double>> etree;    // per mode eigen tree
    std::vector<double> h, scaled_h, weights;
};

// Which modes are sampled for exclusion index j.  The reference treats
// j in [0, N) as "exclude mode j" and j >= N+1 as "all modes"; j == N reads
// an unrelated slot (efficient_krp_sampler.hpp:113, 166, 189, 219), so here
// every j >= N means all modes.
inline bool sampled(uint64_t k, uint64_t j, uint64_t N) { return j >= N || k != j; }

// computeM(j): efficient_krp_sampler.hpp:90-194, written over the sampled-mode
// list instead of the reference's slot arithmetic.  G = Hadamard product of
// the sampled root Grams; pinv(G) from its eigendecomposition with absolute
// cut 1e-8 (44, 126-150).  For each sampled mode k except the last,
// P_k = pinv (.) prod_{k' sampled, k' > k} G_k', eigendecomposed ascending,
// E_k[r] = sqrt(mu_r) v_r^T (152-185).  For the last sampled mode E rows are
// v_{R-1-r}^T / sqrt(lam_{R-1-r}) (zero when lam <= 1e-8).  Eigen trees are
// built over the R rows of E_k with F = 1 and multiplied by the root Gram of
// mode k (187-193).
void compute_M(Sampler& S, uint64_t j) {
    const uint64_t N = S.N, R = S.R, R2 = R * R;
    std::vector<uint64_t> modes;
    for (uint64_t k = 0; k < N; k++) if (sampled(k, j, N)) modes.push_back(k);
    const uint64_t m = modes.size();

    // suffix Hadamard products: suf[i] = prod_{t >= i} G_{modes[t]} (upper)
    std::vector<std::vector<double>> suf(m + 1, std::vector<double>(R2, 1.0));
    for (int64_t t = (int64_t)m - 1; t >= 0; t--) {
        const double* G0 = S.gtree[modes[t]].data();
        for (uint64_t e = 0; e < R2; e++) suf[t][e] = G0[e] * suf[t + 1][e];
    }
    std::vector<double> lam(R), V(R2), W(R2, 0.0);
    jacobi_eigh(R, suf[0].data(), lam.data(), V.data());
    const double tol = 1e-8;
    for (uint64_t v = 0; v < R; v++) {
        for (uint64_t u = 0; u < R; u++)
            W[u * R + (R - 1 - v)] = lam[v] > tol ? V[u * R + v] / std::sqrt(lam[v]) : 0.0;
    }
    S.pinv.assign(R2, 0.0);
    for (uint64_t a = 0; a < R; a++)
        for (uint64_t b = a; b < R; b++) {
            double acc = 0.0;
            for (uint64_t c = 0; c < R; c++) acc += W[a * R + c] * W[b * R + c];
            S.pinv[a * R + b] = acc;
        }
    S.E.assign(N, std::vector<double>());
    S.etree.assign(N, std::vector<double>());
    for (uint64_t t = 0; t < m; t++) {
        const uint64_t k = modes[t];
        std::vector<double> E(R2);
        if (t + 1 < m) {
            std::vector<double> Pk(R2);
            for (uint64_t e = 0; e < R2; e++) Pk[e] = suf[t + 1][e] * S.pinv[e];
            std::vector<double> mu(R), Vk(R2);
            jacobi_eigh(R, Pk.data(), mu.data(), Vk.data());
            for (uint64_t r = 0; r < R; r++) {
                // No clamp in the reference (176-180): sqrt of a negative
                // eigenvalue is NaN there.  Clamped to zero here.
                double sm = std::sqrt(std::max(mu[r], 0.0));
                for (uint64_t u = 0; u < R; u++) E[r * R + u] = Vk[u * R + r] * sm;
            }
        } else {
            for (uint64_t r = 0; r < R; r++)
                for (uint64_t u = 0; u < R; u++) E[r * R + u] = W[u * R + r];
        }
        S.E[k] = E;
        Shape es = make_shape(R, 1);
        std::vector<double> T(es.node_count * R2);
        build_tree(es, E.data(), R, T.data());
        const double* Gk = S.gtree[k].data();
        for (uint64_t nidx = 0; nidx < es.node_count; nidx++)
            for (uint64_t e = 0; e < R2; e++) T[nidx * R2 + e] *= Gk[e];
        S.etree[k] = std::move(T);
    }
}

}  // namespace

// ===========================================================================
// C ABI used by the Python wrapper (oracle/oracle.py).
// ===========================================================================
extern "C" {

int ftlo_tree_shape(uint64_t n, uint64_t F, uint64_t* out) {
    if (n == 0 || F == 0) return 1;
    Shape s = make_shape(n, F);
    out[0] = s.leaf_count; out[1] = s.node_count; out[2] = s.lfill_level;
    out[3] = s.lfill_count; out[4] = s.nodes_upto_lfill; out[5] = s.nodes_before_lfill;
    out[6] = s.complete_level_offset;
    return 0;
}

int64_t ftlo_leaf_idx(uint64_t n, uint64_t F, uint64_t c) {
    Shape s = make_shape(n, F);
    if (c >= s.node_count || !is_leaf(s, c)) return -1;
    return (int64_t)leaf_idx(s, c);
}

// PartitionTree::build_tree on a standalone tree (partition_tree.cpp:42-50).
void ftlo_build_tree(const double* U, uint64_t n, uint64_t R, uint64_t F, double* G) {
    Shape s = make_shape(n, F);
    build_tree(s, U, R, G);
}

// PartitionTree::PTSample (partition_tree.hpp:216-230): x = scaled_h.
void ftlo_pt_sample(const double* G, uint64_t n, uint64_t F, uint64_t R, const double* U,
                    const double* x, double* h, uint64_t* samples, const double* draws,
                    uint64_t J, double* margin) {
    Shape s = make_shape(n, F);
    pt_sample(s, G, R, U, x, h, samples, draws, J, margin);
}

void ftlo_jacobi_eigh(uint64_t R, const double* A, double* lam, double* V) {
    jacobi_eigh(R, A, lam, V);
}

void* ftlo_sampler_create(const double** U, const uint64_t* I, uint64_t N, uint64_t J, uint64_t R) {
    Sampler* S = new Sampler();
    S->N = N; S->J = J; S->R = R;
    for (uint64_t k = 0; k < N; k++) {
        S->U.push_back(U[k]);
        S->I.push_back(I[k]);
        uint64_t F = R < I[k] ? R : I[k];
        S->gshape.push_back(make_shape(I[k], F));
        S->gtree.emplace_back(S->gshape[k].node_count * R * R);
        build_tree(S->gshape[k], U[k], R, S->gtree[k].data());
    }
    S->h.assign(J * R, 1.0);
    S->scaled_h.assign(J * R, 1.0);
    S->weights.assign(J, 0.0);
    return S;
}

// update_sampler(j): efficient_krp_sampler.hpp:73-75.
void ftlo_sampler_update(void* handle, uint64_t j, const double* Uj) {
    Sampler* S = (Sampler*)handle;
    if (Uj) S->U[j] = Uj;
    build_tree(S->gshape[j], S->U[j], S->R, S->gtree[j].data());
}

void ftlo_sampler_destroy(void* handle) { delete (Sampler*)handle; }

// Root Gram of mode k's tree (PartitionTree::get_G0, partition_tree.hpp:172-177).
void ftlo_sampler_root(void* handle, uint64_t k, double* out) {
    Sampler* S = (Sampler*)handle;
    std::memcpy(out, S->gtree[k].data(), S->R * S->R * sizeof(double));
}

// Node Grams of mode k's tree (node_count x R*R).
void ftlo_sampler_tree(void* handle, uint64_t k, double* out) {
    Sampler* S = (Sampler*)handle;
    std::memcpy(out, S->gtree[k].data(), S->gtree[k].size() * sizeof(double));
}

// computeM(j) outputs: pinv (R x R, upper valid) and E (N x R x R, rows of the
// eigen matrix of each sampled mode; zero for the excluded mode).
void ftlo_sampler_compute_m(void* handle, uint64_t j, double* pinv, double* E) {
    Sampler* S = (Sampler*)handle;
    compute_M(*S, j);
    const uint64_t R2 = S->R * S->R;
    std::memcpy(pinv, S->pinv.data(), R2 * sizeof(double));
    for (uint64_t k = 0; k < S->N; k++) {
        if (S->E[k].empty()) std::memset(E + k * R2, 0, R2 * sizeof(double));
        else std::memcpy(E + k * R2, S->E[k].data(), R2 * sizeof(double));
    }
}

// KRPDrawSamples(j, samples, random_draws): efficient_krp_sampler.hpp:209-268.
//   draws   : 2*N*J uniforms; eigen draws of mode k at [k*J, (k+1)*J), gram
//             draws at [N*J + k*J, ...) (221-229)
//   samples : N x J row-major; rows of sampled modes are written
//   h, w    : optional J x R and J outputs (KRPDrawSamples_materialize,
//             als_module.cpp:70-81)
//   margin  : optional J, min distance of each sample's uniforms to the CDF
//             boundaries they were compared against
//   eig_idx : optional N x J, eigen-component index drawn per mode
int ftlo_sampler_draw(void* handle, uint64_t j, const double* draws, uint64_t* samples,
                      double* h_out, double* w_out, double* margin, uint64_t* eig_idx) {
    Sampler* S = (Sampler*)handle;
    const uint64_t N = S->N, J = S->J, R = S->R, R2 = R * R;
    compute_M(*S, j);
    std::fill(S->h.begin(), S->h.end(), 1.0);
    if (margin) for (uint64_t i = 0; i < J; i++) margin[i] = std::numeric_limits<double>::infinity();
    std::vector<uint64_t> tmp(J);
    Shape es = make_shape(R, 1);
    for (uint64_t k = 0; k < N; k++) {
        if (!sampled(k, j, N)) continue;
        std::copy(S->h.begin(), S->h.end(), S->scaled_h.begin());
        // eigen stage: x = h, scaled_h *= E_k[r]
        pt_sample(es, S->etree[k].data(), R, S->E[k].data(), S->h.data(), S->scaled_h.data(),
                  tmp.data(), draws + k * J, J, margin);
        if (eig_idx) std::memcpy(eig_idx + k * J, tmp.data(), J * sizeof(uint64_t));
        // gram stage: x = scaled_h, h *= U_k[s]
        pt_sample(S->gshape[k], S->gtree[k].data(), R, S->U[k], S->scaled_h.data(), S->h.data(),
                  samples + k * J, draws + N * J + k * J, J, margin);
    }
    // compute_DAGAT (sampler.hpp:94-122) and weights R / (lev * J)
    // (efficient_krp_sampler.hpp:257-267).
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < J; i++) {
        double lev = quad_form(S->pinv.data(), S->h.data() + i * R, R);
        S->weights[i] = (double)R / (lev * (double)J);
    }
    if (h_out) std::memcpy(h_out, S->h.data(), J * R * sizeof(double));
    if (w_out) std::memcpy(w_out, S->weights.data(), J * sizeof(double));
    (void)R2;
    return 0;
}

int ftlo_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

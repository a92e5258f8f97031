// ftl_oracle_als.cpp -- CPU FP64 restatement of the sampled CP-ALS step of
// fast_tensor_leverage (cpp_ext/als.hpp, sparse_tensor.hpp, sort_lookup.hpp,
// low_rank_tensor.hpp, common.h).  TEST INFRASTRUCTURE ONLY: the checker for
// the HIP library's ALS path and the timed CPU baseline; nothing in the
// product links or loads it.
//
// Where the reference is order-nondeterministic (parallel unstable sort,
// OpenMP atomics in compute_pinv and execute_spmm, thread-chunked residual),
// this restatement fixes one order: nonzeros sorted stably by (key, original
// position), samples processed in ascending key order, sums taken
// sequentially.  The results equal the reference's up to the rounding of
// those sums.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include <omp.h>

// ---- helpers from ftl_oracle.cpp (same translation unit set) ---------------
extern "C" void ftlo_jacobi_eigh(uint64_t R, const double* A, double* lam, double* V);
extern "C" int ftlo_sampler_draw(void* handle, uint64_t j, const double* draws, uint64_t* samples,
                                 double* h_out, double* w_out, double* margin, uint64_t* eig_idx);
extern "C" void ftlo_sampler_update(void* handle, uint64_t j, const double* Uj);

namespace {

// SortIdxLookup (sort_lookup.hpp:21-60): nonzero positions sorted
// lexicographically over every mode except mode_to_leave.
struct Lookup {
    uint64_t mode;
    std::vector<uint64_t> order;  // nonzero positions in sorted order
};

struct Sparse {
    uint64_t N = 0, nnz = 0;
    const uint32_t* idx = nullptr;  // nnz x N, aliased
    const double* val = nullptr;    // nnz, aliased
    std::vector<Lookup> lookups;
    double normsq = 0.0;

    // a < b over modes != mode (sort_lookup.hpp:52-58)
    bool less(uint64_t mode, uint64_t a, uint64_t b) const {
        for (uint64_t k = 0; k < N; k++) {
            if (k == mode) continue;
            const uint32_t x = idx[a * N + k], y = idx[b * N + k];
            if (x != y) return x < y;
        }
        return false;
    }
    // compare nonzero a with a coordinate tuple t (N entries; t[mode] ignored)
    int cmp_tuple(uint64_t mode, uint64_t a, const uint32_t* t) const {
        for (uint64_t k = 0; k < N; k++) {
            if (k == mode) continue;
            const uint32_t x = idx[a * N + k];
            if (x != t[k]) return x < t[k] ? -1 : 1;
        }
        return 0;
    }
};

// compute_pinv_square (common.h:228-269): eigendecomposition, keep
// eigenvalue v when v >= R - target_rank and lambda_v > 1e-11, M[:,v] =
// V[:,v]/sqrt(lambda_v), out = M M^T (upper triangle written, symmetric).
void pinv_square(const double* Min, uint64_t R, uint64_t target_rank, double* out) {
    std::vector<double> lam(R), V(R * R);
    ftlo_jacobi_eigh(R, Min, lam.data(), V.data());
    std::vector<double> M(R * R, 0.0);
    for (uint64_t v = 0; v < R; v++) {
        const bool keep = v + target_rank >= R && lam[v] > 1e-11;
        for (uint64_t u = 0; u < R; u++) M[u * R + v] = keep ? V[u * R + v] / std::sqrt(lam[v]) : 0.0;
    }
    for (uint64_t a = 0; a < R; a++)
        for (uint64_t b = a; b < R; b++) {
            double acc = 0.0;
            for (uint64_t c = 0; c < R; c++) acc += M[a * R + c] * M[b * R + c];
            out[a * R + b] = acc;
            out[b * R + a] = acc;
        }
}

// ATB_chain_prod_sum(U, U, sigma, sigma) (common.h:152-226): sum over all
// (a, b) of sigma_a sigma_b prod_k (U_k^T U_k)[a][b].
double chain_prod_sum(const double* const* U, const uint64_t* dims, uint64_t N, uint64_t R, const double* sigma) {
    std::vector<double> res(R * R);
    for (uint64_t a = 0; a < R; a++)
        for (uint64_t b = 0; b < R; b++) res[a * R + b] = sigma[a] * sigma[b];
    std::vector<double> g(R * R);
    for (uint64_t k = 0; k < N; k++) {
        std::fill(g.begin(), g.end(), 0.0);
        #pragma omp parallel
        {
            std::vector<double> loc(R * R, 0.0);
            #pragma omp for schedule(static)
            for (uint64_t i = 0; i < dims[k]; i++) {
                const double* row = U[k] + i * R;
                for (uint64_t a = 0; a < R; a++)
                    for (uint64_t b = 0; b < R; b++) loc[a * R + b] += row[a] * row[b];
            }
            #pragma omp critical
            for (uint64_t e = 0; e < R * R; e++) g[e] += loc[e];
        }
        for (uint64_t e = 0; e < R * R; e++) res[e] *= g[e];
    }
    double s = 0.0;
    for (uint64_t e = 0; e < R * R; e++) s += res[e];
    return s;
}

}  // namespace

// std::stable_sort's result (unique for a given order) computed in parallel:
// P contiguous chunks stable-sorted by separate threads, then merged pairwise
// in rounds; std::merge takes equal elements from the left (earlier) chunk
// first, so the merged order is the stable order.
template <class Cmp>
static void stable_sort_par(std::vector<uint64_t>& v, Cmp cmp) {
    const size_t n = v.size();
    const int T = omp_get_max_threads();
    if (n < (1u << 20) || T <= 1) {
        std::stable_sort(v.begin(), v.end(), cmp);
        return;
    }
    size_t P = 1;
    while (P < (size_t)T) P <<= 1;
    std::vector<size_t> b(P + 1);
    for (size_t i = 0; i <= P; i++) b[i] = n * i / P;
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)P; i++) std::stable_sort(v.begin() + b[i], v.begin() + b[i + 1], cmp);
    std::vector<uint64_t> tmp(n);
    uint64_t* src = v.data();
    uint64_t* dst = tmp.data();
    for (size_t w = 1; w < P; w <<= 1) {
#pragma omp parallel for schedule(dynamic, 1)
        for (long i = 0; i < (long)P; i += (long)(2 * w)) {
            const size_t lo = b[i], mid = b[std::min(i + w, P)], hi = b[std::min(i + 2 * w, P)];
            std::merge(src + lo, src + mid, src + mid, src + hi, dst + lo, cmp);
        }
        std::swap(src, dst);
    }
    if (src != v.data()) v.swap(tmp);
}

extern "C" {

// SparseTensor(idx, values, "sort") (sparse_tensor.hpp:46-77): one lookup
// per mode and the squared norm of the values.
void* ftlo_sparse_create(const uint32_t* idx, const double* val, uint64_t nnz, uint64_t N) {
    Sparse* S = new Sparse();
    S->N = N;
    S->nnz = nnz;
    S->idx = idx;
    S->val = val;
    S->lookups.resize(N);
    for (uint64_t j = 0; j < N; j++) {
        Lookup& L = S->lookups[j];
        L.mode = j;
        L.order.resize(nnz);
        std::iota(L.order.begin(), L.order.end(), 0);
        stable_sort_par(L.order, [&](uint64_t a, uint64_t b) { return S->less(j, a, b); });
        if (nnz >= (1u << 24)) {
            fprintf(stderr, "[oracle] lookup of mode %llu sorted (%llu nonzeros)\n", (unsigned long long)j,
                    (unsigned long long)nnz);
            fflush(stderr);
        }
    }
    double ns = 0.0;
    for (uint64_t i = 0; i < nnz; i++) ns += val[i] * val[i];
    S->normsq = ns;
    return S;
}

void ftlo_sparse_destroy(void* h) { delete (Sparse*)h; }

double ftlo_sparse_normsq(void* h) { return ((Sparse*)h)->normsq; }

// execute_downsampled_mttkrp (sparse_tensor.hpp:95-115) + execute_spmm
// (sort_lookup.hpp:69-132): samples (Jp x N, cast to uint32), lhs (Jp x R)
// -> result (Ij x R, zero-filled here).  For each sample the nonzeros whose
// coordinates equal the sample's on every mode but j are found by
// equal_range; result[z_j] += value_z * lhs[sample].  match_counts (Jp,
// optional) receives the number of matched nonzeros per sample; matched_ids
// (optional, capacity match_cap) the matched nonzero positions in order.
// Returns the total number of matches (found_count).
uint64_t ftlo_sparse_mttkrp(void* h, uint64_t j, const uint64_t* samples, uint64_t Jp, const double* lhs,
                            uint64_t R, uint64_t Ij, double* result, uint64_t* match_counts,
                            uint64_t* matched_ids, uint64_t match_cap) {
    Sparse* S = (Sparse*)h;
    const uint64_t N = S->N;
    const Lookup& L = S->lookups[j];
    std::fill(result, result + Ij * R, 0.0);
    std::vector<uint32_t> t(N);
    uint64_t found = 0;
    for (uint64_t s = 0; s < Jp; s++) {
        for (uint64_t k = 0; k < N; k++) t[k] = (uint32_t)samples[s * N + k];
        auto lo = std::lower_bound(L.order.begin(), L.order.end(), 0,
                                   [&](uint64_t a, int) { return S->cmp_tuple(j, a, t.data()) < 0; });
        auto hi = std::upper_bound(lo, L.order.end(), 0,
                                   [&](int, uint64_t a) { return S->cmp_tuple(j, a, t.data()) > 0; });
        uint64_t cnt = 0;
        for (auto it = lo; it != hi; ++it) {
            const uint64_t z = *it;
            const double v = S->val[z];
            const uint64_t row = S->idx[z * N + j];
            for (uint64_t c = 0; c < R; c++) result[row * R + c] += lhs[s * R + c] * v;
            if (matched_ids && found + cnt < match_cap) matched_ids[found + cnt] = z;
            cnt++;
        }
        if (match_counts) match_counts[s] = cnt;
        found += cnt;
    }
    return found;
}

// SortIdxLookup::compute_residual_normsq via lookups[0]
// (sparse_tensor.hpp:117-119, sort_lookup.hpp:134-197):
// sum_z (x_z^2 - 2 x_z m_z) + ||M||^2, m_z = sum_r sigma_r prod_k U_k[z_k][r].
double ftlo_sparse_residual_normsq(void* h, const double* sigma, const double* const* U, const uint64_t* dims,
                                   uint64_t R) {
    Sparse* S = (Sparse*)h;
    const uint64_t N = S->N;
    const Lookup& L = S->lookups[0];
    double acc = 0.0;
    #pragma omp parallel reduction(+ : acc)
    {
        std::vector<double> pp(R);
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < S->nnz; i++) {
            const uint64_t z = L.order[i];
            const uint32_t* id = S->idx + z * N;
            for (uint64_t u = 0; u < R; u++) pp[u] = sigma[u];
            for (uint64_t k = 1; k < N; k++)
                for (uint64_t u = 0; u < R; u++) pp[u] *= U[k][(uint64_t)id[k] * R + u];
            double m = 0.0;
            for (uint64_t u = 0; u < R; u++) m += pp[u] * U[0][(uint64_t)id[0] * R + u];
            const double x = S->val[z];
            acc += x * x - 2 * m * x;
        }
    }
    return acc + chain_prod_sum(U, dims, N, R, sigma);
}

// ||[[sigma; U]]||^2 = ATB_chain_prod_sum(U, U, sigma, sigma) (common.h:215-226).
double ftlo_lowrank_normsq(const double* const* U, const uint64_t* dims, uint64_t N, uint64_t R, const double* sigma) {
    return chain_prod_sum(U, dims, N, R, sigma);
}

// compute_pinv (common.h:271-323): pinv(in^T in) with target rank min(I, R).
void ftlo_compute_pinv(const double* in, uint64_t I, uint64_t R, double* out) {
    std::vector<double> M(R * R, 0.0);
    for (uint64_t i = 0; i < I; i++)
        for (uint64_t a = 0; a < R; a++)
            for (uint64_t b = a; b < R; b++) M[a * R + b] += in[i * R + a] * in[i * R + b];
    pinv_square(M.data(), R, std::min(I, R), out);
}

void ftlo_compute_pinv_square(const double* M, uint64_t R, uint64_t target_rank, double* out) {
    pinv_square(M, R, target_rank, out);
}

// LowRankTensor::renormalize_columns for one mode (low_rank_tensor.hpp:203-260):
// norms = column 2-norms (<= 1e-7 -> 1), U /= norms; norms returned.
void ftlo_renormalize_columns(double* U, uint64_t I, uint64_t R, double* norms) {
    std::vector<double> acc(R, 0.0);
    for (uint64_t i = 0; i < I; i++)
        for (uint64_t c = 0; c < R; c++) acc[c] += U[i * R + c] * U[i * R + c];
    for (uint64_t c = 0; c < R; c++) {
        double n = std::sqrt(acc[c]);
        if (n <= 1e-7) n = 1.0;
        norms[c] = n;
    }
    for (uint64_t i = 0; i < I; i++)
        for (uint64_t c = 0; c < R; c++) U[i * R + c] /= norms[c];
}

// Sample dedup of ALS::execute_ds_als_update (als.hpp:76-160): samples
// (N x J, row j ignored) -> unique tuples (ascending lexicographic order over
// modes != j), their first sample index and multiplicity.  Returns J'.
uint64_t ftlo_als_dedup(const uint64_t* samples, uint64_t N, uint64_t J, uint64_t j, uint64_t* uniq_first,
                        uint64_t* uniq_count, uint64_t* uniq_tuples, void*, void*, void*) {
    std::vector<uint64_t> ord(J);
    std::iota(ord.begin(), ord.end(), 0);
    auto key = [&](uint64_t s, uint64_t k) -> uint32_t { return (uint32_t)samples[k * J + s]; };
    auto lt = [&](uint64_t a, uint64_t b) {
        for (uint64_t k = 0; k < N; k++) {
            if (k == j) continue;
            if (key(a, k) != key(b, k)) return key(a, k) < key(b, k);
        }
        return false;
    };
    std::stable_sort(ord.begin(), ord.end(), lt);
    uint64_t u = 0;
    for (uint64_t i = 0; i < J;) {
        uint64_t e = i + 1;
        while (e < J && !lt(ord[i], ord[e])) e++;
        uniq_first[u] = ord[i];
        uniq_count[u] = e - i;
        for (uint64_t k = 0; k < N; k++) uniq_tuples[u * N + k] = k == j ? 0 : key(ord[i], k);
        u++;
        i = e;
    }
    return u;
}

// One ALS::execute_ds_als_update(j, renormalize, update_sampler) (als.hpp:57-220)
// with the sampler's uniforms injected.  U (N pointers, I_k x R each) are the
// arrays the oracle sampler aliases; U[j] is overwritten in place, sigma
// receives the new column norms when renormalize is set.  diag (optional,
// 4 entries): J', found_count, 0, 0.
int ftlo_als_update_drawn(void* sparse, double* const* U, const uint64_t* dims, uint64_t N, uint64_t J, uint64_t R,
                          uint64_t j, const uint64_t* samples_in, const double* h_in, const double* w_in,
                          int renormalize, double* sigma, double* diag);

int ftlo_als_update(void* sampler, void* sparse, double* const* U, const uint64_t* dims, uint64_t N, uint64_t J,
                    uint64_t R, uint64_t j, const double* draws, int renormalize, int update_sampler, double* sigma,
                    double* diag) {
    std::vector<uint64_t> samples(N * J, 0);
    std::vector<double> h(J * R), w(J);
    ftlo_sampler_draw(sampler, j, draws, samples.data(), h.data(), w.data(), nullptr, nullptr);
    const int rc = ftlo_als_update_drawn(sparse, U, dims, N, J, R, j, samples.data(), h.data(), w.data(), renormalize,
                                         sigma, diag);
    if (rc == 0 && update_sampler) ftlo_sampler_update(sampler, j, U[j]);
    return rc;
}

// The update of execute_ds_als_update after the draw (als.hpp:76-220): the
// samples (N x J), h (J x R) and weights (J) of any sampler -- the efficient
// sampler above, the Larsen-Kolda sampler -- go through dedup, pinv, sampled
// MTTKRP, solve and renormalize.  The caller updates its sampler.
int ftlo_als_update_drawn(void* sparse, double* const* U, const uint64_t* dims, uint64_t N, uint64_t J, uint64_t R,
                          uint64_t j, const uint64_t* samples_in, const double* h_in, const double* w_in,
                          int renormalize, double* sigma, double* diag) {
    std::vector<uint64_t> samples(samples_in, samples_in + N * J);
    std::vector<double> h(h_in, h_in + J * R), w(w_in, w_in + J);
    std::vector<uint64_t> first(J), count(J), tuples(J * N);
    const uint64_t Jp = ftlo_als_dedup(samples.data(), N, J, j, first.data(), count.data(), tuples.data(), 0, 0, 0);
    std::vector<double> hd(Jp * R), wd(Jp);
    for (uint64_t u = 0; u < Jp; u++) {
        wd[u] = w[first[u]] * (double)count[u];
        for (uint64_t c = 0; c < R; c++) hd[u * R + c] = h[first[u] * R + c];
    }
    for (uint64_t u = 0; u < Jp; u++) {
        const double s = std::sqrt(wd[u]);
        for (uint64_t c = 0; c < R; c++) hd[u * R + c] *= s;
    }
    std::vector<double> pinv(R * R);
    ftlo_compute_pinv(hd.data(), Jp, R, pinv.data());
    for (uint64_t u = 0; u < Jp; u++) {
        const double s = std::sqrt(wd[u]);
        for (uint64_t c = 0; c < R; c++) hd[u * R + c] *= s;
    }
    const uint64_t Ij = dims[j];
    std::vector<double> mt(Ij * R);
    const uint64_t found =
        ftlo_sparse_mttkrp(sparse, j, tuples.data(), Jp, hd.data(), R, Ij, mt.data(), nullptr, nullptr, 0);
    // U_j = mttkrp * pinv (cblas_dsymm Right Upper, als.hpp:189-203)
    double* Uj = U[j];
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < Ij; i++)
        for (uint64_t c = 0; c < R; c++) {
            double acc = 0.0;
            for (uint64_t k = 0; k < R; k++) acc += mt[i * R + k] * pinv[k * R + c];
            Uj[i * R + c] = acc;
        }
    if (renormalize) ftlo_renormalize_columns(Uj, Ij, R, sigma);
    if (diag) {
        diag[0] = (double)Jp;
        diag[1] = (double)found;
        diag[2] = 0.0;
        diag[3] = 0.0;
    }
    return 0;
}

// SortIdxLookup::execute_exact_mttkrp through lookups[j] (sparse_tensor.hpp:79-82,
// sort_lookup.hpp:199-254): result (Ij x R, zero-filled here) gets, for every
// nonzero z in lookup-j order, value_z * prod_{k != j} U_k[z_k] added to row
// z_j.  The product starts from 1 and takes the modes in order, then the
// value (the reference's partial_prod, reused along runs of equal
// coordinates off mode j, which gives the same numbers).  One sequential
// pass: the reference's OpenMP chunks and atomics fix no order.
void ftlo_sparse_exact_mttkrp(void* h, uint64_t j, const double* const* U, uint64_t R, uint64_t Ij, double* result) {
    Sparse* S = (Sparse*)h;
    const uint64_t N = S->N;
    const Lookup& L = S->lookups[j];
    std::fill(result, result + Ij * R, 0.0);
    std::vector<double> pp(R);
    for (uint64_t i = 0; i < S->nnz; i++) {
        const uint64_t z = L.order[i];
        const uint32_t* id = S->idx + z * N;
        for (uint64_t u = 0; u < R; u++) pp[u] = 1.0;
        for (uint64_t k = 0; k < N; k++) {
            if (k == j) continue;
            for (uint64_t u = 0; u < R; u++) pp[u] *= U[k][(uint64_t)id[k] * R + u];
        }
        const double x = S->val[z];
        double* out = result + (uint64_t)id[j] * R;
        for (uint64_t u = 0; u < R; u++) out[u] += pp[u] * x;
    }
}

// ALS::execute_exact_als_update(j, renormalize, update_sampler) (als.hpp:222-270):
// gram = (*)_{k != j} U_k^T U_k (ATB_chain_prod with unit sigmas, common.h:152-213:
// start from ones, multiply by each mode's Gram in mode order), pinv =
// compute_pinv_square(gram, R) (1e-11 cut, full target rank), U_j = exact
// MTTKRP * pinv (cblas_dsymm), then renormalize_columns(j) (sigma <- norms).
// U[j] is overwritten in place.
int ftlo_als_exact_update(void* sparse, double* const* U, const uint64_t* dims, uint64_t N, uint64_t R, uint64_t j,
                          int renormalize, double* sigma) {
    std::vector<double> gram(R * R, 1.0), G(R * R);
    for (uint64_t k = 0; k < N; k++) {
        if (k == j) continue;
        std::fill(G.begin(), G.end(), 0.0);
        for (uint64_t i = 0; i < dims[k]; i++)
            for (uint64_t a = 0; a < R; a++)
                for (uint64_t b = 0; b < R; b++) G[a * R + b] += U[k][i * R + a] * U[k][i * R + b];
        for (uint64_t e = 0; e < R * R; e++) gram[e] *= G[e];
    }
    std::vector<double> pinv(R * R);
    pinv_square(gram.data(), R, R, pinv.data());
    const uint64_t Ij = dims[j];
    std::vector<double> mt(Ij * R);
    ftlo_sparse_exact_mttkrp(sparse, j, U, R, Ij, mt.data());
    double* Uj = U[j];
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < Ij; i++)
        for (uint64_t c = 0; c < R; c++) {
            double acc = 0.0;
            for (uint64_t k = 0; k < R; k++) acc += mt[i * R + k] * pinv[k * R + c];
            Uj[i * R + c] = acc;
        }
    if (renormalize) ftlo_renormalize_columns(Uj, Ij, R, sigma);
    return 0;
}

}  // extern "C"


// ftl_oracle_lk.cpp -- CPU FP64 restatement of the Larsen-Kolda leverage
// sampler of fast_tensor_leverage (cpp_ext/larsen_kolda_sampler.hpp,
// compute_DAGAT of cpp_ext/sampler.hpp, fill_h_by_samples of the same file).
// TEST INFRASTRUCTURE ONLY: the checker for the HIP library's Larsen-Kolda
// path; nothing in the product links or loads it.
//
// The reference draws each row with std::discrete_distribution on mt19937;
// here the uniform is injected (draws[k J + i] for sample i, mode k) and a row
// is the first whose inclusive prefix sum of leverage exceeds u * total -- the
// inverse CDF that discrete_distribution evaluates on its normalised
// cumulative probabilities.  Sums are taken sequentially.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" void ftlo_compute_pinv(const double* in, uint64_t I, uint64_t R, double* out);

namespace {

struct LK {
    uint64_t N = 0, J = 0, R = 0;
    std::vector<uint64_t> I;
    std::vector<const double*> U;  // aliased, I_k x R row-major
    std::vector<std::vector<double>> lev, cum;
    std::vector<double> total;
};

// update_sampler(j) (larsen_kolda_sampler.hpp:45-100): P = compute_pinv(U_j)
// (target rank min(I_j, R)), lev_i = u_i . (u_i P) (compute_DAGAT,
// sampler.hpp:94-122: temp = U P by dsymm, res_i = sum_c U_ic temp_ic), total =
// sum of lev.
void lk_update(LK& s, uint64_t j) {
    const uint64_t I = s.I[j], R = s.R;
    const double* U = s.U[j];
    std::vector<double> P(R * R);
    ftlo_compute_pinv(U, I, R, P.data());
    auto& lev = s.lev[j];
    auto& cum = s.cum[j];
    lev.assign(I, 0.0);
    cum.assign(I, 0.0);
    std::vector<double> t(R);
    double run = 0.0;
    for (uint64_t i = 0; i < I; i++) {
        const double* u = U + i * R;
        for (uint64_t c = 0; c < R; c++) {
            double y = 0.0;
            for (uint64_t k = 0; k < R; k++) y += u[k] * P[k * R + c];
            t[c] = y;
        }
        double l = 0.0;
        for (uint64_t c = 0; c < R; c++) l += u[c] * t[c];
        lev[i] = l;
        run += l;
        cum[i] = run;
    }
    s.total[j] = run;
}

}  // namespace

extern "C" {

void* ftlo_lk_create(const double** U, const uint64_t* I, uint64_t N, uint64_t J, uint64_t R) {
    LK* s = new LK;
    s->N = N;
    s->J = J;
    s->R = R;
    s->I.assign(I, I + N);
    s->U.assign(U, U + N);
    s->lev.resize(N);
    s->cum.resize(N);
    s->total.assign(N, 0.0);
    for (uint64_t k = 0; k < N; k++) lk_update(*s, k);
    return s;
}

void ftlo_lk_destroy(void* h) { delete (LK*)h; }

// update_sampler(j) from the factor now at Uj (aliased from then on).
void ftlo_lk_update(void* h, uint64_t j, const double* Uj) {
    LK* s = (LK*)h;
    if (Uj) s->U[j] = Uj;
    lk_update(*s, j);
}

void ftlo_lk_leverage(void* h, uint64_t k, double* lev, double* total) {
    LK* s = (LK*)h;
    std::copy(s->lev[k].begin(), s->lev[k].end(), lev);
    if (total) *total = s->total[k];
}

// KRPDrawSamples(j) (larsen_kolda_sampler.hpp:102-133) with injected uniforms:
// samples (N x J, rows k != j), weights_i = exp(-log J + sum_k (log total_k -
// log lev_k[row])), h = prod_{k != j} U_k[row] in mode order from 1
// (fill_h_by_samples, sampler.hpp:74-89).  margin (optional, J): per sample
// the smallest distance, over its modes, of u * total to the two prefix sums
// bracketing it, divided by total -- how far the draw is from flipping rows.
int ftlo_lk_draw(void* h, uint64_t j, const double* draws, uint64_t* samples, double* h_out, double* w_out,
                 double* margin) {
    LK* s = (LK*)h;
    const uint64_t N = s->N, J = s->J, R = s->R;
    for (uint64_t i = 0; i < J; i++) {
        double logw = 0.0 - std::log((double)J);
        double mg = 1.0;
        if (h_out)
            for (uint64_t c = 0; c < R; c++) h_out[i * R + c] = 1.0;
        for (uint64_t k = 0; k < N; k++) {
            if (k == j) continue;
            const auto& cum = s->cum[k];
            const double tot = s->total[k];
            const double t = draws[k * J + i] * tot;
            uint64_t row = (uint64_t)(std::upper_bound(cum.begin(), cum.end(), t) - cum.begin());
            if (row >= s->I[k]) row = s->I[k] - 1;
            samples[k * J + i] = row;
            const double below = row == 0 ? t : t - cum[row - 1];
            const double above = cum[row] - t;
            mg = std::min(mg, std::min(below, above) / tot);
            logw += std::log(tot) - std::log(s->lev[k][row]);
            if (h_out)
                for (uint64_t c = 0; c < R; c++) h_out[i * R + c] *= s->U[k][row * R + c];
        }
        if (w_out) w_out[i] = std::exp(logw);
        if (margin) margin[i] = mg;
    }
    return 0;
}

}  // extern "C"

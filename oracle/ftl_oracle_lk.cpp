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
    // hybrid (larsen_kolda_hybrid.hpp): alpha_k = max lev / total, rows in
    // ascending leverage order (ties by row)
    bool hybrid = false;
    std::vector<double> alpha;
    std::vector<std::vector<uint64_t>> sorted;
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
    if (s.hybrid) {
        // alpha_j and the sort of update_sampler (larsen_kolda_hybrid.hpp:101-133)
        double m = 0.0;
        for (uint64_t i = 0; i < I; i++) m = std::max(m, lev[i]);
        s.alpha[j] = m / run;
        auto& o = s.sorted[j];
        o.resize(I);
        for (uint64_t i = 0; i < I; i++) o[i] = i;
        std::stable_sort(o.begin(), o.end(), [&](uint64_t a, uint64_t b) { return lev[a] < lev[b]; });
    }
}

}  // namespace

extern "C" {

void* ftlo_lk_create2(const double** U, const uint64_t* I, uint64_t N, uint64_t J, uint64_t R, int hybrid) {
    LK* s = new LK;
    s->hybrid = hybrid != 0;
    s->alpha.assign(N, 0.0);
    s->sorted.resize(N);
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

void* ftlo_lk_create(const double** U, const uint64_t* I, uint64_t N, uint64_t J, uint64_t R) {
    return ftlo_lk_create2(U, I, N, J, R, 0);
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

// KRPDrawSamples of LarsenKoldaHybrid (larsen_kolda_hybrid.hpp:139-246) with
// injected uniforms: attempt a of (sample i, mode k) takes
// utab[a N J + k J + i] (A attempts available).  Candidates of mode k: the
// rows from the lower bound of (tau alpha_k / alpha*) total_k in the
// leverage order, listed in ascending row order; the kept tuples (p >= tau)
// in mixed-radix order of the candidate lists (lowest mode fastest), weight 1,
// their mass p_det summed in that order; the other samples by rejection
// (p < tau), weight (1 - p_det) exp(-log J - log p).  info (8): deterministic
// count, p_det, candidates, 1 if skipped (2^62 or more candidates), the
// smallest relative distance of a leverage score to its mode's cut, of a
// candidate tuple's p to tau, of a drawn uniform to a row boundary, and the
// most attempts one sample took.  Returns 0; 2 if a sample needs more than A
// attempts; 3 if more than J tuples are kept; 4 if there are more than 2^32
// candidate tuples (and fewer than 2^62).
int ftlo_lkh_draw(void* h, uint64_t j, const double* utab, uint64_t A, uint64_t* samples, double* h_out,
                  double* w_out, double* info) {
    LK* s = (LK*)h;
    const uint64_t N = s->N, J = s->J, R = s->R;
    const double tau = 1.0 / (double)J;
    double alpha_star = 1.0;
    for (uint64_t k = 0; k < N; k++)
        if (k != j) alpha_star *= s->alpha[k];
    std::vector<std::vector<uint64_t>> cand(N);
    double log_c = 0.0, cut_margin = 1.0;
    uint64_t C = 1;
    bool zero = false;
    for (uint64_t k = 0; k < N; k++) {
        if (k == j) continue;
        const double threshold = tau * s->alpha[k] / alpha_star;
        const double cut = threshold * s->total[k];
        const auto& lev = s->lev[k];
        const auto& o = s->sorted[k];
        const auto lb = std::lower_bound(o.begin(), o.end(), cut, [&](uint64_t r, double t) { return lev[r] < t; });
        cand[k].assign(lb, o.end());
        std::sort(cand[k].begin(), cand[k].end());
        for (uint64_t i = 0; i < s->I[k]; i++) cut_margin = std::min(cut_margin, std::fabs(lev[i] - cut) / cut);
        if (cand[k].empty()) zero = true;
        log_c += std::log((double)cand[k].size());
        C *= cand[k].empty() ? 1 : cand[k].size();
    }
    const bool skip = !zero && log_c >= 62.0 * std::log(2.0);
    if (zero) C = 0;
    if (!zero && !skip && log_c > 32.0 * std::log(2.0) + 1e-9) return 4;
    uint64_t ndet = 0;
    double p_det = 0.0, keep_margin = 1.0;
    std::vector<uint64_t> rows(N, 0);
    if (!skip) {
        for (uint64_t c = 0; c < C; c++) {
            uint64_t x = c;
            double wlog = 0.0;
            for (uint64_t k = 0; k < N; k++) {
                if (k == j) continue;
                const uint64_t d = x % cand[k].size();
                x /= cand[k].size();
                rows[k] = cand[k][d];
                wlog += std::log(s->lev[k][rows[k]]) - std::log(s->total[k]);
            }
            const double p = std::exp(wlog);
            keep_margin = std::min(keep_margin, std::fabs(p - tau) / tau);
            if (p >= tau) {
                if (ndet >= J) return 3;
                for (uint64_t k = 0; k < N; k++)
                    if (k != j) samples[k * J + ndet] = rows[k];
                if (w_out) w_out[ndet] = 1.0;
                ndet++;
                p_det += p;
            }
        }
    }
    double draw_margin = 1.0;
    uint64_t max_att = 0;
    for (uint64_t i = ndet; i < J; i++) {
        double wlog = 0.0;
        uint64_t a = 0;
        for (;; a++) {
            if (a >= A) return 2;
            wlog = 0.0;
            for (uint64_t k = 0; k < N; k++) {
                if (k == j) continue;
                const auto& cum = s->cum[k];
                const double tot = s->total[k];
                const double t = utab[a * N * J + k * J + i] * tot;
                uint64_t row = (uint64_t)(std::upper_bound(cum.begin(), cum.end(), t) - cum.begin());
                if (row >= s->I[k]) row = s->I[k] - 1;
                samples[k * J + i] = row;
                const double below = row == 0 ? t : t - cum[row - 1];
                draw_margin = std::min(draw_margin, std::min(below, cum[row] - t) / tot);
                wlog += std::log(s->lev[k][row]) - std::log(tot);
            }
            if (std::exp(wlog) < tau) break;
        }
        max_att = std::max(max_att, a + 1);
        if (w_out) w_out[i] = (1.0 - p_det) * std::exp((0.0 - std::log((double)J)) - wlog);
    }
    if (h_out)
        for (uint64_t i = 0; i < J; i++)
            for (uint64_t c = 0; c < R; c++) {
                double v = 1.0;
                for (uint64_t k = 0; k < N; k++)
                    if (k != j) v *= s->U[k][samples[k * J + i] * R + c];
                h_out[i * R + c] = v;
            }
    if (info) {
        info[0] = (double)ndet;
        info[1] = p_det;
        info[2] = zero ? 0.0 : std::exp(log_c);
        info[3] = skip ? 1.0 : 0.0;
        info[4] = cut_margin;
        info[5] = keep_margin;
        info[6] = draw_margin;
        info[7] = (double)max_att;
    }
    return 0;
}

void ftlo_lk_alpha(void* h, double* out) {
    LK* s = (LK*)h;
    std::copy(s->alpha.begin(), s->alpha.end(), out);
}

}  // extern "C"

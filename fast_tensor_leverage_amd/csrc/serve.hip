// serve.hip -- the draw-request server: requests for KRP leverage samples over
// different factor sets (uneven N, I, R, J) are admitted by HBM footprint into
// a fixed device arena and served concurrently.  The reference has no
// server; its one-request surface is PySampler (als_module.cpp:30-82).
//
// Two kinds of request: one over factors generated on the device from the
// request's seed (built, drawn and released with the request), and one over a
// registered factor set whose Gram trees stay resident in the arena between
// requests and are rebuilt only for the modes a caller updates.  A request's
// tree builds run on a stream of its own; every request whose trees are
// complete is drawn in a packed launch -- the computeM kernels and the fused
// draw kernel each launched once for all such requests of the same rank and
// precision (k_gram_pinv_batched, k_eigen_modes_batched, k_krp_draw_packed) on
// the server's draw stream.  A request's samples depend only on its trees and
// its seed, so they are the same served alone or in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "ftl_common.hpp"
#include "sampler_internal.hpp"

namespace ftl {

namespace {

// First-fit allocator over one device allocation (offsets kept sorted; freed
// blocks merge with their neighbours).
class ArenaAllocator final : public DeviceAllocator {
public:
    static constexpr size_t kAlign = 256;

    explicit ArenaAllocator(size_t bytes) : size_(round_up(bytes)) {
        FTL_HIP_CHECK(hipMalloc((void**)&base_, size_));
        free_[0] = size_;
    }
    ~ArenaAllocator() override {
        if (base_) (void)hipFree(base_);
    }

    static size_t round_up(size_t b) { return (std::max<size_t>(b, 1) + kAlign - 1) / kAlign * kAlign; }

    void* alloc(size_t bytes) override {
        const size_t need = round_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > need) free_[off + need] = len - need;
            used_ += need;
            peak_ = std::max(peak_, used_);
            return base_ + off;
        }
        return nullptr;
    }

    void release(void* p, size_t bytes) override {
        if (!p) return;
        const size_t off = (size_t)((char*)p - base_), len = round_up(bytes);
        used_ -= len;
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && off + len == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t size() const { return size_; }

private:
    char* base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

// Factor entries N(0,1) in FP32 (stored FP64): Box-Muller in FP32 on two
// 24-bit uniforms of one Philox4x32-10 block keyed by (seed, mode, entry
// pair); each block gives two entries.
__global__ void k_normal_factor(double* __restrict__ U, uint64_t n, uint64_t seed, uint32_t mode) {
    const uint64_t pairs = (n + 1) / 2;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs;
         p += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c[4] = {(uint32_t)p, (uint32_t)(p >> 32), mode, 0x9E3779B9u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const float u1 = ((float)(c[0] >> 8) + 1.0f) * (1.0f / 16777217.0f);  // (0, 1)
        const float u2 = (float)(c[1] >> 8) * (1.0f / 16777216.0f);           // [0, 1)
        const float rad = sqrtf(-2.0f * logf(u1));
        float sn, cs;
        sincosf(6.2831853f * u2, &sn, &cs);
        U[2 * p] = (double)(rad * cs);
        if (2 * p + 1 < n) U[2 * p + 1] = (double)(rad * sn);
    }
}

// A registered factor set: factors and trees resident in the arena.
struct FactorSet {
    uint64_t id = 0;
    uint32_t N = 0;
    uint64_t I[kMaxModes] = {};
    uint64_t R = 0, Jmax = 0;
    int precision = FTL_FP64;
    std::vector<std::pair<void*, size_t>> blocks;
    double* U[kMaxModes] = {};
    Sampler* sampler = nullptr;
    hipStream_t stream = nullptr;  // factor copies and tree builds
    hipEvent_t ready = nullptr;    // after the last build queued
    hipEvent_t last_draw = nullptr;
    hipEvent_t copied = nullptr;   // after the last copy of caller factors
    bool drew = false;             // last_draw recorded at least once
    uint64_t in_flight = 0;        // requests admitted and not finished
    bool busy = false;             // a draw of this set is queued and not done
    bool removing = false;
};

enum class Phase { Waiting, Building, Drawing };

struct Request {
    uint64_t id = 0, seed = 0;
    uint32_t N = 0, j = 0;
    uint64_t I[kMaxModes] = {};
    uint64_t R = 0, J = 0;
    int precision = FTL_FP64;
    size_t footprint = 0;
    bool cancelled = false;
    bool failed = false;  // could not be placed in the arena
    int slot = -1;
    FactorSet* fs = nullptr;  // the registered set it draws from (else its own factors)
    Sampler* sampler = nullptr;
    std::vector<std::pair<void*, size_t>> blocks;
    double* U[kMaxModes] = {};
    uint64_t* samples = nullptr;
    Phase phase = Phase::Waiting;
    hipEvent_t built = nullptr;  // trees complete (own factors)
    hipEvent_t done = nullptr;   // samples complete
    Sampler* drawer() const { return fs ? fs->sampler : sampler; }
};

}  // namespace

struct Server {
    std::unique_ptr<ArenaAllocator> arena;
    uint32_t max_active = 0;
    std::vector<hipStream_t> streams;
    std::vector<int> free_slots;
    uint64_t next_id = 1, next_fs = 1;
    std::deque<Request*> waiting;
    std::vector<Request*> active;
    std::map<uint64_t, Request*> finished;
    std::map<uint64_t, FactorSet*> sets;
    std::vector<uint64_t> unreported;  // finished ids not yet returned by progress()
    uint64_t n_admitted = 0, n_completed = 0, n_cancelled = 0, n_deferred = 0, n_packs = 0, n_packed = 0;
    // packed draws: a ring of kStages draw lanes, each a stream with its own
    // device copy of a pack's arguments and pinned host staging buffer (reused
    // once its previous copy has run), so consecutive packs overlap
    static constexpr int kStages = 4;
    hipStream_t draw_stream[kStages] = {};
    void* pack_dev[kStages] = {};
    size_t pack_bytes = 0;
    void* stage[kStages] = {};
    hipEvent_t stage_ev[kStages] = {};
    bool stage_used[kStages] = {};
    int next_stage = 0;

    ~Server() {
        for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
        for (hipStream_t d : draw_stream)
            if (d) (void)hipStreamSynchronize(d);
        for (auto& kv : sets)
            if (kv.second->stream) (void)hipStreamSynchronize(kv.second->stream);
        for (Request* r : active) release(r);
        for (auto& kv : finished) release(kv.second);
        for (Request* r : waiting) delete r;
        for (auto& kv : sets) free_set(kv.second);
        for (int i = 0; i < kStages; i++) {
            if (stage_ev[i]) (void)hipEventDestroy(stage_ev[i]);
            if (stage[i]) (void)hipHostFree(stage[i]);
            if (pack_dev[i]) (void)hipFree(pack_dev[i]);
            if (draw_stream[i]) (void)hipStreamDestroy(draw_stream[i]);
        }
        for (hipStream_t s : streams) (void)hipStreamDestroy(s);
    }

    void init(uint64_t budget, uint32_t slots) {
        FTL_REQUIRE(budget >= (1ull << 20), "the HBM budget must be at least 1 MiB");
        FTL_REQUIRE(slots >= 1 && slots <= 64, "active slots must be in [1, 64]");
        arena.reset(new ArenaAllocator(budget));
        max_active = slots;
        streams.resize(slots);
        for (uint32_t i = 0; i < slots; i++) {
            FTL_HIP_CHECK(hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking));
            free_slots.push_back((int)(slots - 1 - i));
        }
        // the packed draws' arguments (a few KB per request) live outside the
        // arena, which is the budget of the requests' own buffers
        pack_bytes = packed_args_bytes(slots);
        for (int i = 0; i < kStages; i++) {
            FTL_HIP_CHECK(hipStreamCreateWithFlags(&draw_stream[i], hipStreamNonBlocking));
            FTL_HIP_CHECK(hipMalloc(&pack_dev[i], pack_bytes));
            FTL_HIP_CHECK(hipHostMalloc(&stage[i], pack_bytes, hipHostMallocDefault));
            FTL_HIP_CHECK(hipEventCreateWithFlags(&stage_ev[i], hipEventDisableTiming));
        }
    }

    static size_t request_footprint(const Request& r) {
        size_t b = ArenaAllocator::round_up(r.N * r.J * sizeof(uint64_t));
        if (r.fs) return b;
        b += sampler_footprint(r.N, r.I, r.J, r.R, r.precision, ArenaAllocator::kAlign);
        for (uint32_t k = 0; k < r.N; k++) b += ArenaAllocator::round_up(r.I[k] * r.R * sizeof(double));
        return b;
    }

    uint64_t enqueue(std::unique_ptr<Request> r) {
        r->footprint = request_footprint(*r);
        if (r->footprint > arena->size())
            throw Error{FTL_ENOMEM, "request needs " + std::to_string(r->footprint) +
                                        " bytes, more than the whole HBM budget (" + std::to_string(arena->size()) +
                                        ")"};
        if (r->fs) r->fs->in_flight++;
        const uint64_t id = r->id;
        waiting.push_back(r.release());
        admit();
        return id;
    }

    uint64_t submit(uint32_t N, const uint64_t* I, uint64_t R, uint64_t J, uint32_t j, uint64_t seed, int prec) {
        FTL_REQUIRE(N >= 1 && N <= (uint32_t)kMaxModes, "N must be in [1, 16]");
        FTL_REQUIRE(R >= 1 && R <= 128, "R must be in [1, 128]");
        FTL_REQUIRE(J >= 1 && J <= (1ull << 32), "J must be in [1, 2^32]");
        FTL_REQUIRE(I, "null row counts");
        FTL_REQUIRE(prec == FTL_FP64 || prec == FTL_FP32, "precision must be FTL_FP64 or FTL_FP32");
        FTL_REQUIRE(!(N == 1 && j == 0), "no mode left to sample (N == 1 and j == 0)");
        std::unique_ptr<Request> r(new Request());
        r->id = next_id++;
        r->seed = seed;
        r->N = N;
        r->j = j;
        r->R = R;
        r->J = J;
        r->precision = prec;
        for (uint32_t k = 0; k < N; k++) {
            FTL_REQUIRE(I[k] >= 1 && I[k] < (1ull << 40), "factor row counts must be in [1, 2^40)");
            r->I[k] = I[k];
        }
        return enqueue(std::move(r));
    }

    // ---- registered factor sets ------------------------------------------
    FactorSet* find_set(uint64_t fs) {
        auto it = sets.find(fs);
        FTL_REQUIRE(it != sets.end() && !it->second->removing, "unknown factor set " + std::to_string(fs));
        return it->second;
    }

    void free_set(FactorSet* f) {
        if (f->stream) (void)hipStreamSynchronize(f->stream);
        if (f->sampler) sampler_delete(f->sampler);
        for (auto& b : f->blocks) arena->release(b.first, b.second);
        if (f->ready) (void)hipEventDestroy(f->ready);
        if (f->last_draw) (void)hipEventDestroy(f->last_draw);
        if (f->copied) (void)hipEventDestroy(f->copied);
        if (f->stream) (void)hipStreamDestroy(f->stream);
        delete f;
    }

    uint64_t add_factors(uint32_t N, const uint64_t* I, uint64_t R, const double* const* U, bool on_device, int prec,
                         uint64_t Jmax) {
        FTL_REQUIRE(N >= 1 && N <= (uint32_t)kMaxModes, "N must be in [1, 16]");
        FTL_REQUIRE(R >= 1 && R <= 128, "R must be in [1, 128]");
        FTL_REQUIRE(Jmax >= 1 && Jmax <= (1ull << 32), "J_max must be in [1, 2^32]");
        FTL_REQUIRE(I && U, "null argument");
        FTL_REQUIRE(prec == FTL_FP64 || prec == FTL_FP32, "precision must be FTL_FP64 or FTL_FP32");
        std::unique_ptr<FactorSet> f(new FactorSet());
        f->N = N;
        f->R = R;
        f->Jmax = Jmax;
        f->precision = prec;
        for (uint32_t k = 0; k < N; k++) {
            FTL_REQUIRE(I[k] >= 1 && I[k] < (1ull << 40), "factor row counts must be in [1, 2^40)");
            FTL_REQUIRE(U[k], "null factor");
            f->I[k] = I[k];
        }
        size_t need = sampler_footprint(N, f->I, Jmax, R, prec, ArenaAllocator::kAlign);
        for (uint32_t k = 0; k < N; k++) need += ArenaAllocator::round_up(I[k] * R * sizeof(double));
        if (arena->used() + need > arena->size())
            throw Error{FTL_ENOMEM, "the factor set needs " + std::to_string(need) + " bytes; " +
                                        std::to_string(arena->size() - arena->used()) + " of the budget are free"};
        FTL_HIP_CHECK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        try {
            FTL_HIP_CHECK(hipEventCreateWithFlags(&f->copied, hipEventDisableTiming));
            for (uint32_t k = 0; k < N; k++) {
                const size_t b = I[k] * R * sizeof(double);
                void* p = arena->alloc(b);
                if (!p) throw Error{FTL_ENOMEM, "arena full"};
                f->blocks.push_back({p, b});
                f->U[k] = (double*)p;
                FTL_HIP_CHECK(hipMemcpyAsync(p, U[k], b, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                             f->stream));
            }
            // the caller's factors (host or device) are read before the call
            // returns, so they may be freed or overwritten at once; the tree
            // builds stay queued
            FTL_HIP_CHECK(hipEventRecord(f->copied, f->stream));
            FTL_HIP_CHECK(hipEventSynchronize(f->copied));
            f->sampler = sampler_new_async(N, f->I, Jmax, R, f->U, prec, arena.get(), f->stream);
            // the group draw kernel (the packed launch's; the same samples as
            // the one-sample-per-wave kernel a tuning row may pick)
            if (R <= 64) sampler_set_option(f->sampler, "draw_kernel", 0);
            FTL_HIP_CHECK(hipEventCreateWithFlags(&f->ready, hipEventDisableTiming));
            FTL_HIP_CHECK(hipEventCreateWithFlags(&f->last_draw, hipEventDisableTiming));
            FTL_HIP_CHECK(hipEventRecord(f->ready, f->stream));
        } catch (...) {
            free_set(f.release());
            throw;
        }
        f->id = next_fs++;
        const uint64_t id = f->id;
        sets[id] = f.release();
        return id;
    }

    // update_sampler(k) of a registered set: after the set's queued draws,
    // copy the new factor k and rebuild tree k only.
    void update_factor(uint64_t fs, uint32_t k, const double* Uk, bool on_device) {
        FactorSet* f = find_set(fs);
        FTL_REQUIRE(k < f->N, "mode index out of range");
        FTL_REQUIRE(Uk, "null factor");
        if (f->drew) FTL_HIP_CHECK(hipStreamWaitEvent(f->stream, f->last_draw, 0));
        FTL_HIP_CHECK(hipMemcpyAsync(f->U[k], Uk, f->I[k] * f->R * sizeof(double),
                                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, f->stream));
        // Uk is read before the call returns (after the set's queued draws,
        // which the copy waits for); the rebuild stays queued
        FTL_HIP_CHECK(hipEventRecord(f->copied, f->stream));
        sampler_rebuild(f->sampler, k, f->stream);
        FTL_HIP_CHECK(hipEventRecord(f->ready, f->stream));
        FTL_HIP_CHECK(hipEventSynchronize(f->copied));
    }

    void remove_factors(uint64_t fs) {
        FactorSet* f = find_set(fs);
        f->removing = true;
        if (f->in_flight == 0) {
            sets.erase(fs);
            free_set(f);
        }
    }

    uint64_t submit_on(uint64_t fs, uint64_t J, uint32_t j, uint64_t seed) {
        FactorSet* f = find_set(fs);
        FTL_REQUIRE(J >= 1 && J <= f->Jmax, "J must be in [1, J_max] of the factor set");
        FTL_REQUIRE(sampler_packable(f->sampler) || J == f->Jmax,
                    "draws on a factor set of R > 64 take J_max samples");
        FTL_REQUIRE(!(f->N == 1 && j == 0), "no mode left to sample (N == 1 and j == 0)");
        std::unique_ptr<Request> r(new Request());
        r->id = next_id++;
        r->seed = seed;
        r->N = f->N;
        r->j = j;
        r->R = f->R;
        r->J = J;
        r->precision = f->precision;
        for (uint32_t k = 0; k < f->N; k++) r->I[k] = f->I[k];
        r->fs = f;
        return enqueue(std::move(r));
    }

    void release(Request* r) {
        if (r->sampler) sampler_delete(r->sampler);
        for (auto& b : r->blocks) arena->release(b.first, b.second);
        if (r->built) (void)hipEventDestroy(r->built);
        if (r->done) (void)hipEventDestroy(r->done);
        if (r->slot >= 0) free_slots.push_back(r->slot);
        if (r->fs) {
            FactorSet* f = r->fs;
            f->in_flight--;
            if (f->removing && f->in_flight == 0) {
                sets.erase(f->id);
                free_set(f);
            }
        }
        delete r;
    }

    // Starts r on a free stream: its factors and tree builds (own factors),
    // or nothing to build (a registered set).  false (nothing held) if the
    // arena cannot place its blocks right now.
    bool start(Request* r) {
        const int slot = free_slots.back();
        hipStream_t st = streams[slot];
        auto undo = [&]() {
            (void)hipStreamSynchronize(st);
            if (r->sampler) sampler_delete(r->sampler);
            r->sampler = nullptr;
            for (auto& b : r->blocks) arena->release(b.first, b.second);
            r->blocks.clear();
        };
        auto take = [&](size_t bytes) -> void* {
            void* p = arena->alloc(bytes);
            if (!p) throw Error{FTL_ENOMEM, "arena full"};
            r->blocks.push_back({p, bytes});
            return p;
        };
        try {
            r->samples = (uint64_t*)take(r->N * r->J * sizeof(uint64_t));
            FTL_HIP_CHECK(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
            if (!r->fs) {
                for (uint32_t k = 0; k < r->N; k++) r->U[k] = (double*)take(r->I[k] * r->R * sizeof(double));
                for (uint32_t k = 0; k < r->N; k++) {
                    const uint64_t n = r->I[k] * r->R;
                    const unsigned g =
                        (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, (n / 2 + 255) / 256), 16384);
                    hipLaunchKernelGGL(k_normal_factor, dim3(g), dim3(256), 0, st, r->U[k], n, r->seed, k);
                    FTL_HIP_CHECK(hipGetLastError());
                }
                r->sampler = sampler_new_async(r->N, r->I, r->J, r->R, r->U, r->precision, arena.get(), st);
                if (r->R <= 64) sampler_set_option(r->sampler, "draw_kernel", 0);
                sampler_set_seed(r->sampler, r->seed, 0);
                if (!sampler_packable(r->sampler)) {
                    // one sample per block above R = 64: its own launch
                    FTL_HIP_CHECK(hipMemsetAsync(r->samples, 0, r->N * r->J * sizeof(uint64_t), st));
                    sampler_draw(r->sampler, r->j, r->samples, nullptr, nullptr, nullptr, st);
                    FTL_HIP_CHECK(hipEventRecord(r->done, st));
                    r->phase = Phase::Drawing;
                } else {
                    FTL_HIP_CHECK(hipEventCreateWithFlags(&r->built, hipEventDisableTiming));
                    FTL_HIP_CHECK(hipMemsetAsync(r->samples, 0, r->N * r->J * sizeof(uint64_t), st));
                    FTL_HIP_CHECK(hipEventRecord(r->built, st));
                    r->phase = Phase::Building;
                }
            } else {
                FTL_HIP_CHECK(hipMemsetAsync(r->samples, 0, r->N * r->J * sizeof(uint64_t), st));
                FTL_HIP_CHECK(hipEventCreateWithFlags(&r->built, hipEventDisableTiming));
                FTL_HIP_CHECK(hipEventRecord(r->built, st));
                r->phase = Phase::Building;
            }
        } catch (const Error& e) {
            undo();
            if (r->done) {
                (void)hipEventDestroy(r->done);
                r->done = nullptr;
            }
            if (r->built) {
                (void)hipEventDestroy(r->built);
                r->built = nullptr;
            }
            if (e.code == FTL_ENOMEM) return false;
            throw;
        }
        free_slots.pop_back();
        r->slot = slot;
        return true;
    }

    // FIFO admission while a slot is free and the head request fits the
    // budget next to the requests in flight.
    void admit() {
        while (!waiting.empty() && !free_slots.empty()) {
            Request* r = waiting.front();
            if (arena->used() + r->footprint > arena->size()) break;
            if (!start(r)) {
                n_deferred++;
                if (active.empty() && finished.empty()) {
                    // nothing will free memory: the request cannot be placed
                    waiting.pop_front();
                    r->failed = true;
                    finished[r->id] = r;
                    unreported.push_back(r->id);
                    continue;
                }
                break;
            }
            waiting.pop_front();
            active.push_back(r);
            n_admitted++;
        }
    }

    // The next draw lane (its staging buffer free for the host to fill).
    int take_stage() {
        const int i = next_stage;
        next_stage = (next_stage + 1) % kStages;
        if (stage_used[i]) FTL_HIP_CHECK(hipEventSynchronize(stage_ev[i]));
        stage_used[i] = true;
        return i;
    }

    // Every building request whose trees are complete (and, on a registered
    // set, whose set has no draw in flight) is drawn now: one packed launch
    // per group of the same rank, precision and launch options.
    void launch_ready() {
        std::map<uint64_t, std::vector<Request*>> groups;
        for (Request* r : active) {
            if (r->phase != Phase::Building || r->cancelled) continue;
            if (hipEventQuery(r->built) != hipSuccess) continue;
            if (r->fs) {
                if (r->fs->busy || hipEventQuery(r->fs->ready) != hipSuccess) continue;
                r->fs->busy = true;  // one draw of a set at a time (its computeM buffers)
                sampler_set_seed(r->fs->sampler, r->seed, 0);
            }
            if (!sampler_packable(r->drawer())) {
                // one sample per block (R > 64): its own launch
                hipStream_t ds = draw_stream[next_stage];
                next_stage = (next_stage + 1) % kStages;
                sampler_draw(r->drawer(), r->j, r->samples, nullptr, nullptr, nullptr, ds);
                FTL_HIP_CHECK(hipEventRecord(r->done, ds));
                if (r->fs) {
                    FTL_HIP_CHECK(hipEventRecord(r->fs->last_draw, ds));
                    r->fs->drew = true;
                }
                r->phase = Phase::Drawing;
                continue;
            }
            groups[sampler_pack_key(r->drawer())].push_back(r);
        }
        for (auto& kv : groups) {
            // smallest draws first; a pack closes once it holds about two
            // device-wide rounds of wave items, so that a small request is not
            // held back by a large one packed with it (a larger request is
            // a pack of its own)
            std::vector<Request*>& rs = kv.second;
            std::stable_sort(rs.begin(), rs.end(), [](const Request* a, const Request* b) { return a->J < b->J; });
            size_t b = 0;
            while (b < rs.size()) {
                size_t e = b + 1;
                uint64_t items_in = rs[b]->J;
                while (e < rs.size() && items_in + rs[e]->J <= kPackSamples) items_in += rs[e++]->J;
                launch_pack(std::vector<Request*>(rs.begin() + (long)b, rs.begin() + (long)e));
                b = e;
            }
        }
    }

    // Samples per packed launch (beyond a single request): about two rounds
    // of the resident waves of the group draw kernel on 256 CUs.
    static constexpr uint64_t kPackSamples = 12288;

    void launch_pack(const std::vector<Request*>& rs) {
        {
            std::vector<PackItem> items;
            for (Request* r : rs) items.push_back({r->drawer(), r->j, r->samples, r->J});
            const int si = take_stage();
            hipStream_t ds = draw_stream[si];
            sampler_draw_packed(items.data(), (uint32_t)items.size(), pack_dev[si], stage[si], ds);
            FTL_HIP_CHECK(hipEventRecord(stage_ev[si], ds));
            for (Request* r : rs) {
                FTL_HIP_CHECK(hipEventRecord(r->done, ds));
                if (r->fs) {
                    FTL_HIP_CHECK(hipEventRecord(r->fs->last_draw, ds));
                    r->fs->drew = true;
                }
                r->phase = Phase::Drawing;
            }
            n_packs++;
            n_packed += rs.size();
        }
    }

    // Launches ready draws, retires finished requests (their ids queued for
    // progress()), then admits.
    void step() {
        for (size_t i = 0; i < active.size();) {
            Request* r = active[i];
            if (r->phase != Phase::Drawing) {
                if (r->cancelled && r->phase == Phase::Building && hipEventQuery(r->built) == hipSuccess) {
                    active.erase(active.begin() + (long)i);
                    n_cancelled++;
                    release(r);
                    continue;
                }
                i++;
                continue;
            }
            const hipError_t q = hipEventQuery(r->done);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            FTL_HIP_CHECK(q);
            active.erase(active.begin() + (long)i);
            if (r->fs) r->fs->busy = false;
            if (r->cancelled) {
                n_cancelled++;
                release(r);
                continue;
            }
            n_completed++;
            // the stream is free for the next request; the result stays in the arena
            free_slots.push_back(r->slot);
            r->slot = -1;
            finished[r->id] = r;
            unreported.push_back(r->id);
        }
        admit();
        launch_ready();
    }

    // Up to `capacity` finished ids not reported before (the rest stay
    // queued for the next call).
    void progress(std::vector<uint64_t>& done, uint64_t capacity) {
        step();
        const size_t n = (size_t)std::min<uint64_t>(capacity, unreported.size());
        done.insert(done.end(), unreported.begin(), unreported.begin() + (long)n);
        unreported.erase(unreported.begin(), unreported.begin() + (long)n);
    }

    // Drives the server until request id (0: any request) has finished or
    // timeout_s passed, sleeping between polls of the device events.
    bool wait(uint64_t id, double timeout_s) {
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            step();
            if (id ? finished.count(id) != 0 : !unreported.empty()) return true;
            if (id && finished.count(id) == 0) {
                bool known = false;
                for (Request* r : active) known |= r->id == id;
                for (Request* r : waiting) known |= r->id == id;
                FTL_REQUIRE(known, "unknown request " + std::to_string(id));
            }
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el >= timeout_s) return false;
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }

    Request* finished_request(uint64_t id) {
        auto it = finished.find(id);
        FTL_REQUIRE(it != finished.end(), "request " + std::to_string(id) + " has not finished (or was released)");
        return it->second;
    }

    void result(uint64_t id, uint64_t* samples_host) {
        Request* r = finished_request(id);
        if (r->failed) {
            finished.erase(id);
            release(r);
            throw Error{FTL_ENOMEM, "request " + std::to_string(id) + " could not be placed in the HBM arena"};
        }
        if (samples_host)
            FTL_HIP_CHECK(hipMemcpy(samples_host, r->samples, r->N * r->J * sizeof(uint64_t), hipMemcpyDeviceToHost));
        finished.erase(id);
        release(r);
        admit();
    }

    void factor(uint64_t id, uint32_t k, double* out) {
        Request* r = finished_request(id);
        FTL_REQUIRE(k < r->N, "mode index out of range");
        FTL_REQUIRE(!r->fs, "request " + std::to_string(id) + " drew from a registered factor set");
        FTL_HIP_CHECK(hipMemcpy(out, r->U[k], r->I[k] * r->R * sizeof(double), hipMemcpyDeviceToHost));
    }

    void cancel(uint64_t id) {
        for (auto it = waiting.begin(); it != waiting.end(); ++it)
            if ((*it)->id == id) {
                Request* r = *it;
                waiting.erase(it);
                release(r);
                n_cancelled++;
                admit();
                return;
            }
        for (Request* r : active)
            if (r->id == id) {
                r->cancelled = true;  // released when its queued work ends
                return;
            }
        auto it = finished.find(id);
        FTL_REQUIRE(it != finished.end(), "unknown request " + std::to_string(id));
        release(it->second);
        finished.erase(it);
        n_cancelled++;
        admit();
    }

    void stats(double* out) const {
        out[0] = (double)active.size();
        out[1] = (double)waiting.size();
        out[2] = (double)finished.size();
        out[3] = (double)arena->used();
        out[4] = (double)arena->peak();
        out[5] = (double)n_admitted;
        out[6] = (double)n_completed;
        out[7] = (double)n_cancelled;
    }
};

}  // namespace ftl

struct ftl_server {
    ftl::Server impl;
};

using ftl::guarded;

extern "C" {

int ftl_server_create(uint64_t hbm_budget_bytes, uint32_t max_active, ftl_server** out) {
    return guarded([&] {
        FTL_REQUIRE(out, "null output handle");
        *out = nullptr;
        std::unique_ptr<ftl_server> s(new ftl_server());
        s->impl.init(hbm_budget_bytes, max_active);
        *out = s.release();
    });
}

int ftl_server_destroy(ftl_server* s) {
    return guarded([&] { delete s; });
}

int ftl_server_submit(ftl_server* s, uint32_t N, const uint64_t* I, uint64_t R, uint64_t J, uint32_t j, uint64_t seed,
                      int precision, uint64_t* id_out) {
    return guarded([&] {
        FTL_REQUIRE(s && id_out, "null argument");
        *id_out = s->impl.submit(N, I, R, J, j, seed, precision);
    });
}

int ftl_server_add_factors(ftl_server* s, uint32_t N, const uint64_t* I, uint64_t R, const double* const* U,
                           int on_device, int precision, uint64_t J_max, uint64_t* fs_out) {
    return guarded([&] {
        FTL_REQUIRE(s && fs_out, "null argument");
        *fs_out = s->impl.add_factors(N, I, R, U, on_device != 0, precision, J_max);
    });
}

int ftl_server_add_sampler(ftl_server* s, ftl_sampler* smp, uint64_t* fs_out) {
    return guarded([&] {
        FTL_REQUIRE(s && smp && fs_out, "null argument");
        ftl::SamplerView v = ftl::sampler_view(smp);
        *fs_out = s->impl.add_factors(v.N, v.I, v.R, v.U, true, v.precision, v.J);
    });
}

int ftl_server_update_factor(ftl_server* s, uint64_t fs, uint32_t k, const double* Uk, int on_device) {
    return guarded([&] {
        FTL_REQUIRE(s, "null server");
        s->impl.update_factor(fs, k, Uk, on_device != 0);
    });
}

int ftl_server_remove_factors(ftl_server* s, uint64_t fs) {
    return guarded([&] {
        FTL_REQUIRE(s, "null server");
        s->impl.remove_factors(fs);
    });
}

int ftl_server_submit_on(ftl_server* s, uint64_t fs, uint64_t J, uint32_t j, uint64_t seed, uint64_t* id_out) {
    return guarded([&] {
        FTL_REQUIRE(s && id_out, "null argument");
        *id_out = s->impl.submit_on(fs, J, j, seed);
    });
}

int ftl_server_wait(ftl_server* s, uint64_t id, double timeout_s, int* finished) {
    return guarded([&] {
        FTL_REQUIRE(s && finished, "null argument");
        FTL_REQUIRE(timeout_s >= 0.0, "timeout must be >= 0");
        *finished = s->impl.wait(id, timeout_s) ? 1 : 0;
    });
}

int ftl_server_progress(ftl_server* s, uint64_t* done_ids, uint64_t capacity, uint64_t* n_done) {
    return guarded([&] {
        FTL_REQUIRE(s && n_done, "null argument");
        std::vector<uint64_t> done;
        s->impl.progress(done, done_ids ? capacity : 0);
        for (size_t i = 0; i < done.size(); i++) done_ids[i] = done[i];
        *n_done = done.size();
    });
}

int ftl_server_result(ftl_server* s, uint64_t id, uint64_t* samples_host) {
    return guarded([&] {
        FTL_REQUIRE(s, "null server");
        s->impl.result(id, samples_host);
    });
}

int ftl_server_factor(ftl_server* s, uint64_t id, uint32_t k, double* U_host) {
    return guarded([&] {
        FTL_REQUIRE(s && U_host, "null argument");
        s->impl.factor(id, k, U_host);
    });
}

int ftl_server_cancel(ftl_server* s, uint64_t id) {
    return guarded([&] {
        FTL_REQUIRE(s, "null server");
        s->impl.cancel(id);
    });
}

int ftl_server_stats(ftl_server* s, double* out8) {
    return guarded([&] {
        FTL_REQUIRE(s && out8, "null argument");
        s->impl.stats(out8);
    });
}

}  // extern "C"

// serve.hip -- the draw-request server: requests for KRP leverage samples over
// different factor sets (uneven N, I, R, J) are admitted by HBM footprint into
// a fixed device arena, run concurrently on a pool of streams (factor
// generation, tree builds, computeM and the fused draw of one request are
// queued on its stream, so requests admitted while others run overlap on the
// device), and retired as they complete.  The reference has no server; its
// one-request surface is PySampler (als_module.cpp:30-82).
//
// A request's result depends only on its own seed (factors and uniforms are
// counter-based Philox streams of that seed), so it is the same served alone
// or in any mix.
#include <algorithm>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <string>
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
    Sampler* sampler = nullptr;
    std::vector<std::pair<void*, size_t>> blocks;
    double* U[kMaxModes] = {};
    uint64_t* samples = nullptr;
    hipEvent_t done = nullptr;
};

}  // namespace

struct Server {
    std::unique_ptr<ArenaAllocator> arena;
    uint32_t max_active = 0;
    std::vector<hipStream_t> streams;
    std::vector<int> free_slots;
    uint64_t next_id = 1;
    std::deque<Request*> waiting;
    std::vector<Request*> active;
    std::map<uint64_t, Request*> finished;
    std::vector<uint64_t> failed_ids;  // reported by the next progress()
    uint64_t n_admitted = 0, n_completed = 0, n_cancelled = 0, n_deferred = 0;

    ~Server() {
        for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
        for (Request* r : active) release(r);
        for (auto& kv : finished) release(kv.second);
        for (Request* r : waiting) delete r;
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
    }

    static size_t request_footprint(const Request& r) {
        size_t b = sampler_footprint(r.N, r.I, r.J, r.R, r.precision, ArenaAllocator::kAlign);
        for (uint32_t k = 0; k < r.N; k++) b += ArenaAllocator::round_up(r.I[k] * r.R * sizeof(double));
        b += ArenaAllocator::round_up(r.N * r.J * sizeof(uint64_t));
        return b;
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
        r->footprint = request_footprint(*r);
        if (r->footprint > arena->size())
            throw Error{FTL_ENOMEM, "request needs " + std::to_string(r->footprint) +
                                        " bytes, more than the whole HBM budget (" + std::to_string(arena->size()) +
                                        ")"};
        const uint64_t id = r->id;
        waiting.push_back(r.release());
        admit();
        return id;
    }

    void release(Request* r) {
        if (r->sampler) sampler_delete(r->sampler);
        for (auto& b : r->blocks) arena->release(b.first, b.second);
        if (r->done) (void)hipEventDestroy(r->done);
        if (r->slot >= 0) free_slots.push_back(r->slot);
        delete r;
    }

    // Starts r on a free stream; false (nothing held) if the arena cannot
    // place its blocks right now.
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
            for (uint32_t k = 0; k < r->N; k++) r->U[k] = (double*)take(r->I[k] * r->R * sizeof(double));
            r->samples = (uint64_t*)take(r->N * r->J * sizeof(uint64_t));
            for (uint32_t k = 0; k < r->N; k++) {
                const uint64_t n = r->I[k] * r->R;
                const unsigned g = (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, (n / 2 + 255) / 256), 16384);
                hipLaunchKernelGGL(k_normal_factor, dim3(g), dim3(256), 0, st, r->U[k], n, r->seed, k);
                FTL_HIP_CHECK(hipGetLastError());
            }
            FTL_HIP_CHECK(hipMemsetAsync(r->samples, 0, r->N * r->J * sizeof(uint64_t), st));
            r->sampler = sampler_new_async(r->N, r->I, r->J, r->R, r->U, r->precision, arena.get(), st);
            sampler_set_seed(r->sampler, r->seed, 0);
            sampler_draw(r->sampler, r->j, r->samples, nullptr, nullptr, nullptr, st);
            FTL_HIP_CHECK(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
            FTL_HIP_CHECK(hipEventRecord(r->done, st));
        } catch (const Error& e) {
            undo();
            if (r->done) {
                (void)hipEventDestroy(r->done);
                r->done = nullptr;
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
                    failed_ids.push_back(r->id);
                    continue;
                }
                break;
            }
            waiting.pop_front();
            active.push_back(r);
            n_admitted++;
        }
    }

    // Retires finished requests (ids appended to done), then admits.
    void progress(std::vector<uint64_t>& done) {
        done.insert(done.end(), failed_ids.begin(), failed_ids.end());
        failed_ids.clear();
        for (size_t i = 0; i < active.size();) {
            Request* r = active[i];
            const hipError_t q = hipEventQuery(r->done);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            FTL_HIP_CHECK(q);
            active.erase(active.begin() + (long)i);
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
            done.push_back(r->id);
        }
        admit();
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
        FTL_HIP_CHECK(hipMemcpy(out, r->U[k], r->I[k] * r->R * sizeof(double), hipMemcpyDeviceToHost));
    }

    void cancel(uint64_t id) {
        for (auto it = waiting.begin(); it != waiting.end(); ++it)
            if ((*it)->id == id) {
                delete *it;
                waiting.erase(it);
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

int ftl_server_progress(ftl_server* s, uint64_t* done_ids, uint64_t capacity, uint64_t* n_done) {
    return guarded([&] {
        FTL_REQUIRE(s && n_done, "null argument");
        std::vector<uint64_t> done;
        s->impl.progress(done);
        // a call reports at most max_active completions plus the requests
        // that could not be placed (capacity >= max_active + waiting)
        if (done_ids)
            for (size_t i = 0; i < done.size() && i < capacity; i++) done_ids[i] = done[i];
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

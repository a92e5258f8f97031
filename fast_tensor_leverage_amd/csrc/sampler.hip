// sampler.hip -- host side of the MI355X KRP leverage sampler: device memory,
// tree builds, computeM and draw launches, and the C ABI of include/ftl_hip.h.
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include <dlfcn.h>

#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>

#include "ftl_common.hpp"
#include "sampler_internal.hpp"
#include "sampler_kernels.hpp"
#include "tuning.hpp"

namespace ftl {

namespace {
thread_local std::string g_err;
}
void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

// Device blocks of destroyed samplers, kept for samplers built later: a
// block of the same rounded size on the same device is handed out again, so
// building samplers of one shape repeatedly maps and unmaps no memory
// (hipFree of a multi-GB block takes tens of ms).  Blocks are returned only
// by the sampler destructor, after its streams are synchronised.  At most
// FTL_POOL_CAP_GB GiB are kept (default: a quarter of the first device's
// memory; 0 keeps none), the oldest kept blocks released first to make room;
// an allocation that fails releases every kept block of the device and
// tries again.
namespace {
class BlockPool {
  public:
    static size_t rounded(size_t bytes) {
        const size_t g = bytes >= (2u << 20) ? (2u << 20) : 256u;
        return (bytes + g - 1) / g * g;
    }

    void* take(size_t bytes, int* dev_out) {
        int dev = 0;
        FTL_HIP_CHECK(hipGetDevice(&dev));
        *dev_out = dev;
        const size_t rb = rounded(bytes);
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = free_.find({dev, rb});
            if (it != free_.end()) {
                void* p = it->second;
                free_.erase(it);
                forget_age(p);
                held_ -= rb;
                return p;
            }
        }
        void* p = nullptr;
        if (hipMalloc(&p, rb) != hipSuccess) {
            (void)hipGetLastError();
            trim(dev);
            FTL_HIP_CHECK(hipMalloc(&p, rb));
        }
        return p;
    }

    void give(void* p, size_t bytes, int dev) {
        const size_t rb = rounded(bytes);
        std::vector<void*> out;
        {
            std::lock_guard<std::mutex> g(mu_);
            const size_t c = cap_locked();
            if (rb > c) {
                out.push_back(p);
            } else {
                evict_to(c - rb, out);  // the oldest kept blocks first
                free_.emplace(std::make_pair(dev, rb), p);
                age_.emplace(seq_++, std::make_pair(std::make_pair(dev, rb), p));
                held_ += rb;
            }
        }
        for (void* q : out) (void)hipFree(q);
    }

    // releases every kept block of device dev (-1: of every device)
    void trim(int dev) {
        std::vector<void*> out;
        {
            std::lock_guard<std::mutex> g(mu_);
            for (auto it = free_.begin(); it != free_.end();) {
                if (dev < 0 || it->first.first == dev) {
                    out.push_back(it->second);
                    forget_age(it->second);
                    held_ -= it->first.second;
                    it = free_.erase(it);
                } else {
                    ++it;
                }
            }
        }
        for (void* p : out) (void)hipFree(p);
    }

    size_t held() {
        std::lock_guard<std::mutex> g(mu_);
        return held_;
    }

    // a new cap; kept blocks beyond it are released, the oldest first
    void set_cap(size_t bytes) {
        std::vector<void*> out;
        {
            std::lock_guard<std::mutex> g(mu_);
            cap_ = bytes;
            cap_set_ = true;
            evict_to(cap_, out);
        }
        for (void* q : out) (void)hipFree(q);
    }

    size_t cap() {
        std::lock_guard<std::mutex> g(mu_);
        return cap_locked();
    }

  private:
    // caller holds mu_
    size_t cap_locked() {
        if (!cap_set_) {
            cap_set_ = true;
            if (const char* e = std::getenv("FTL_POOL_CAP_GB")) {
                const double gb = std::atof(e);
                cap_ = gb > 0 ? (size_t)(gb * 1073741824.0) : (size_t)0;
            } else {
                size_t free_b = 0, total = 0;
                if (hipMemGetInfo(&free_b, &total) != hipSuccess) {
                    (void)hipGetLastError();
                    total = 0;
                }
                cap_ = total / 4;
            }
        }
        return cap_;
    }
    // caller holds mu_: releases the oldest kept blocks (into `out`) until at
    // most `limit` bytes are kept
    void evict_to(size_t limit, std::vector<void*>& out) {
        while (held_ > limit && !age_.empty()) {
            const auto old = age_.begin();
            const auto range = free_.equal_range(old->second.first);
            for (auto it = range.first; it != range.second; ++it)
                if (it->second == old->second.second) {
                    free_.erase(it);
                    break;
                }
            held_ -= old->second.first.second;
            out.push_back(old->second.second);
            age_.erase(old);
        }
    }
    // caller holds mu_
    void forget_age(void* p) {
        for (auto it = age_.begin(); it != age_.end(); ++it)
            if (it->second.second == p) {
                age_.erase(it);
                return;
            }
    }
    std::mutex mu_;
    std::multimap<std::pair<int, size_t>, void*> free_;
    // kept blocks by the order they were kept in
    std::map<uint64_t, std::pair<std::pair<int, size_t>, void*>> age_;
    uint64_t seq_ = 0;
    size_t held_ = 0;
    size_t cap_ = 0;
    bool cap_set_ = false;
};

// never destroyed: samplers may be destroyed by static destructors at exit
BlockPool& block_pool() {
    static BlockPool* p = new BlockPool;
    return *p;
}
}  // namespace

void device_pool_trim() { block_pool().trim(-1); }
size_t device_pool_held() { return block_pool().held(); }
void device_pool_set_cap(size_t bytes) { block_pool().set_cap(bytes); }
size_t device_pool_cap() { return block_pool().cap(); }

TuningTable parse_tuning(const std::string& text) {
    TuningTable t;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string prec, opts;
        TuningRow r;
        if (!(ls >> prec)) continue;
        if (prec == "fp64") r.precision = FTL_FP64;
        else if (prec == "fp32") r.precision = FTL_FP32;
        else continue;
        if (!(ls >> r.r_lo >> r.r_hi >> r.d_lo >> r.d_hi >> r.j_lo >> r.j_hi >> opts)) continue;
        bool ok = true;
        std::istringstream os(opts);
        std::string kv;
        while (std::getline(os, kv, ',')) {
            const size_t eq = kv.find('=');
            if (eq == std::string::npos || eq == 0) {
                ok = false;
                break;
            }
            char* end = nullptr;
            const long long v = std::strtoll(kv.c_str() + eq + 1, &end, 10);
            if (!end || *end != '\0') {
                ok = false;
                break;
            }
            r.opts.push_back({kv.substr(0, eq), (int64_t)v});
        }
        if (ok) t.rows.push_back(std::move(r));
    }
    return t;
}

namespace {

std::string library_dir() {
    Dl_info info{};
    if (dladdr((const void*)&library_dir, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t slash = p.rfind('/');
        if (slash != std::string::npos) return p.substr(0, slash);
    }
    return ".";
}

}  // namespace

const TuningTable& tuning_table() {
    static TuningTable table;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* env = std::getenv("FTL_TUNING_TABLE");
        std::string path = env ? std::string(env) : library_dir() + "/ftl_tuning.tsv";
        if (path.empty()) return;
        std::ifstream f(path);
        if (!f) return;
        std::stringstream buf;
        buf << f.rdbuf();
        table = parse_tuning(buf.str());
        table.path = path;
    });
    return table;
}

namespace {

// Levels below this many from the bottom are stored FP32 in FP32 mode when
// no tuning row sets "fp32_levels".
constexpr uint32_t kDefaultFp32Levels = 5;

// Depth (levels below the root) of the Gram tree over I rows with leaf size
// min(R, I).
inline uint32_t tree_depth(uint64_t I, uint64_t R) {
    const Shape sh = make_shape(I, std::min<uint64_t>(R, I));
    return (uint32_t)sh.lfill_level + (sh.node_count > sh.nodes_upto_lfill ? 1 : 0);
}

inline uint32_t max_depth(uint32_t N, const uint64_t* I, uint64_t R) {
    uint32_t d = 0;
    for (uint32_t k = 0; k < N; k++) d = std::max(d, tree_depth(I[k], R));
    return d;
}

// fp32_levels of the tuning row for this shape (the default without one).
inline uint32_t tuned_fp32_levels(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, int precision) {
    const TuningTable& t = tuning_table();
    const int row = t.find(precision, R, max_depth(N, I, R), J);
    if (row >= 0)
        for (const auto& kv : t.rows[(size_t)row].opts)
            if (kv.first == "fp32_levels" && kv.second > 0 && kv.second <= 64) return (uint32_t)kv.second;
    return kDefaultFp32Levels;
}

// Zeroed entries after the last node of every tree beyond one packed node:
// covers the over-read of the widest draw kernel (pairs of entries over 512
// lanes per sample, 1024 entries) with a margin.
constexpr uint64_t kTreePadExtra = 1088;

struct ModeHost {
    uint64_t I = 0, F = 0;
    Shape sh{};
    uint32_t d64 = 0;
    uint64_t n64 = 0;
    double* U64 = nullptr;
    float* U32 = nullptr;
    double* tree64 = nullptr;
    float* tree32 = nullptr;
    uint64_t tree32_elems = 0;
    // FP32 sampler whose FP32 levels are held in FP32 only: tree64 holds the
    // n64 FP64 nodes (k_subtree_build writes both parts)
    bool compact = false;
    uint64_t tree64_nodes = 0;
};

// Levels below the root of a tree shape (the depth of its deepest nodes).
inline uint32_t shape_depth(const Shape& sh) {
    return (uint32_t)sh.lfill_level + (sh.node_count > sh.nodes_upto_lfill ? 1 : 0);
}

// LDS of one subtree build block: a slot per wave (its leaf) and one per
// depth below d64 - 1 (a left child waiting for its sibling).
inline size_t subtree_lds(uint64_t R, uint32_t depth, uint32_t d64, uint32_t waves) {
    return (size_t)(waves + depth - (d64 - 1)) * tri_stride64(R) * sizeof(double);
}

constexpr size_t kSubtreeLds = 150u * 1024u;

// LDS of a k_subtree_pairs block for subtrees L levels deep: a node per
// wave (2^(L-1) waves) in the accumulator layout and the (row, column) of
// each packed entry, or at R = 49-64 the packed triangle alone.
inline size_t pairs_lds(uint64_t R, uint32_t L) {
    const uint64_t nb = (R + 15) / 16;
    if (nb == 4) return (size_t)(1ull << (L - 1)) * pairs_slot_doubles<4>(R) * sizeof(double);
    return (size_t)(1ull << (L - 1)) * (nb * (nb + 1) / 2 * 256) * sizeof(double) + tri_size(R) * sizeof(uint16_t);
}

// Whether an FP32 tree is held compact: FP32 levels below d64 exist, the
// leaves go on the matrix cores (R <= 64), the lowest FP64 level has at
// least 64 nodes (one subtree block each: smaller trees build faster level
// by level, and their FP64 copies are small), a subtree's 2^(L-1) leaf pairs
// fit one block (L <= 5, L <= 4 above R = 32) and both subtree kernels fit
// the LDS budget.
inline bool compact_fits(uint64_t R, uint32_t depth, uint32_t d64) {
    if (R > 64 || d64 > depth || d64 < 7) return false;
    const uint32_t L = depth - (d64 - 1);
    if (L > (R <= 32 ? 5u : 4u)) return false;
    return pairs_lds(R, L) <= kSubtreeLds && subtree_lds(R, depth, d64, 4) <= kSubtreeLds;
}

inline unsigned grid_for(uint64_t work, unsigned bs) {
    uint64_t g = (work + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

struct Sampler {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t N = 0;
    uint64_t J = 0, R = 0;
    int precision = FTL_FP64;
    uint32_t ts64 = 0, ts32 = 0;
    uint32_t fp32_levels = kDefaultFp32Levels;
    // FP32 mode: leaf scans over the FP32 factor copy (option "fp32_leaf")
    uint32_t leaf32 = 1;
    // FP32 mode: leaf Grams on the FP64 matrix cores (option "leaf_mfma")
    uint32_t leaf_mfma = 1;
    // option "compact_fp32": FP32 samplers keep no FP64 copy of their FP32 levels
    uint32_t compact_opt = 1;
    uint64_t tree_pad = 0;  // zeroed entries after every tree (draw kernel over-read)
    std::vector<ModeHost> modes;
    // computeM outputs and scratch.  A sampler that owns its stream keeps two
    // sets and runs the computeM of call n on cm_stream into set n % 2 while
    // the draw kernels of call n-1 (set (n-1) % 2) still run on the draw
    // stream: cm_done orders a set's draw after its computeM, draw_done the
    // next computeM into the set after the draw that read it, and trees_ev
    // every computeM after the tree builds queued before it.
    struct MSet {
        double *pinv = nullptr, *W = nullptr, *E = nullptr, *etree = nullptr, *work = nullptr, *lam = nullptr;
        hipEvent_t cm_done = nullptr, draw_done = nullptr;
    };
    MSet msets[2];
    uint32_t n_msets = 1;
    uint64_t mset_next = 0;
    hipStream_t cm_stream = nullptr;
    hipEvent_t trees_ev = nullptr;
    bool trees_dirty = true;
    // Draws enqueued on a caller's stream (not the handle's): handle_ev,
    // recorded on the handle's stream before such a draw, orders it after the
    // factor copies and tree builds queued there; foreign_ev, recorded on the
    // caller's stream after it, makes the next factor copy or tree build wait
    // until that draw has read the old factors and trees.
    hipEvent_t handle_ev = nullptr, foreign_ev = nullptr;
    bool foreign_pending = false;
    // some draw ran on a caller's stream: the destructor then waits for the
    // whole device before its blocks go back to the pool (foreign_ev covers
    // only the last such stream)
    bool foreign_used = false;

    void wait_foreign_draws() {
        if (!foreign_pending) return;
        FTL_HIP_CHECK(hipStreamWaitEvent(stream, foreign_ev, 0));
        foreign_pending = false;
    }
    // option "overlap_compute_m" (only with a stream of the sampler's own)
    uint32_t overlap_cm = 1;
    // set 0 (compute_m_host and the single-set samplers read these)
    double *pinv = nullptr, *W = nullptr, *E = nullptr, *etree = nullptr, *work = nullptr, *lam = nullptr;
    uint32_t etree_nodes = 0;
    // staging for the host-pointer API
    uint64_t* samples_dev = nullptr;
    double *h_dev = nullptr, *w_dev = nullptr, *draws_dev = nullptr;
    uint64_t seed = 0x243F6A8885A308D3ull, call = 0;
    uint32_t sg_override = 0, grid_override = 0, jacobi_override = 0;
    // 1: one draw launch per sampled mode (h kept in h_state between them),
    // so a launch's working set is one tree; 0: all modes in one launch
    uint32_t mode_split = 1;
    // draw kernel: 0 = by rank (a group of lanes per sample for R <= 64, one
    // sample per block above), 1 = one sample per wave (R <= 32), 2 = one
    // sample per block (R > 32), 3 = a group of lanes per sample at any rank.
    // Measured at N=3, I=2^23, J=5e4 (FP64 draw): R=64 group 5.68 ms, block
    // 6.20 (128 threads) / 7.23 (256); R=128 group 172.5 ms (the 129
    // coefficients and node values per lane spill), block 27.8 (256 threads)
    uint32_t draw_kernel = 0;
    // threads per block of the one-sample-per-block kernel (0: by rank)
    uint32_t block_threads = 0;
    // 1: each mode's descent bucketed by level-3 subtree onto XCD-affine
    // blocks.  Measured at N=3, I=2^23, R=32 (FP64): prefix 167 + partition
    // 97 + suffix 367 us per mode against 539 us for the one-launch mode, so
    // off by default: the per-level cost does not depend on where the node
    // is cached.
    uint32_t xcd_buckets = 0;
    double *st_x = nullptr, *st_s = nullptr;
    uint32_t *bucket = nullptr, *blist = nullptr, *boff = nullptr;
    double* h_state = nullptr;
    bool timing = false;
    // events of every timed call since the last read-out: {start, after
    // gram_pinv, after eigen_modes, draw start, draw end}
    std::vector<std::array<hipEvent_t, 5>> pending;
    std::vector<hipEvent_t> event_pool;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double last_ms[3] = {0, 0, 0};
    uint64_t last_count = 0;

    hipEvent_t take_event() {
        hipEvent_t e;
        if (!event_pool.empty()) {
            e = event_pool.back();
            event_pool.pop_back();
        } else {
            FTL_HIP_CHECK(hipEventCreate(&e));
        }
        return e;
    }

    // Averages (ms) over the pending timed calls; blocks until they finished.
    void collect_timing() {
        double acc[3] = {0, 0, 0};
        for (auto& q : pending) {
            FTL_HIP_CHECK(hipEventSynchronize(q[4]));
            const int from[3] = {0, 1, 3}, to[3] = {1, 2, 4};
            for (int i = 0; i < 3; i++) {
                float t = 0;
                FTL_HIP_CHECK(hipEventElapsedTime(&t, q[from[i]], q[to[i]]));
                acc[i] += t;
            }
            for (auto e : q) event_pool.push_back(e);
        }
        last_count = pending.size();
        for (int i = 0; i < 3; i++) last_ms[i] = last_count ? acc[i] / (double)last_count : 0.0;
        pending.clear();
    }
    size_t bytes_held = 0;
    std::vector<std::pair<void*, size_t>> allocs;
    // device of the blocks taken from the block pool (hipMalloc otherwise)
    int pool_dev = -1;
    // optional sub-allocator (a serving arena); hipMalloc / hipFree otherwise
    DeviceAllocator* allocator = nullptr;
    bool own_stream = true;

    template <typename T>
    T* dalloc(uint64_t count) {
        void* p = nullptr;
        const size_t bytes = (size_t)count * sizeof(T);
        if (allocator) {
            p = allocator->alloc(bytes ? bytes : 16);
            if (!p) throw Error{FTL_ENOMEM, "serving arena exhausted"};
        } else {
            p = block_pool().take(bytes ? bytes : 16, &pool_dev);
        }
        allocs.push_back({p, bytes ? bytes : 16});
        bytes_held += bytes;
        return (T*)p;
    }

    void release_block(void* p, size_t bytes) {
        if (allocator) allocator->release(p, bytes);
        else (void)hipFree(p);
    }

    // at destruction, once the handle's streams are idle: back to the pool
    void retire_block(void* p, size_t bytes) {
        if (allocator) allocator->release(p, bytes);
        else block_pool().give(p, bytes, pool_dev);
    }

    void dfree(void* p, size_t bytes) {
        if (!p) return;
        for (size_t i = 0; i < allocs.size(); i++)
            if (allocs[i].first == p) {
                release_block(p, allocs[i].second);
                allocs.erase(allocs.begin() + (long)i);
                bytes_held -= bytes;
                return;
            }
    }

    ~Sampler() {
        if (foreign_pending) (void)hipEventSynchronize(foreign_ev);
        if (foreign_used) (void)hipDeviceSynchronize();
        if (handle_ev) (void)hipEventDestroy(handle_ev);
        if (foreign_ev) (void)hipEventDestroy(foreign_ev);
        if (cm_stream) (void)hipStreamSynchronize(cm_stream);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& ms : msets) {
            if (ms.cm_done) (void)hipEventDestroy(ms.cm_done);
            if (ms.draw_done) (void)hipEventDestroy(ms.draw_done);
        }
        if (trees_ev) (void)hipEventDestroy(trees_ev);
        if (cm_stream) (void)hipStreamDestroy(cm_stream);
        for (auto& a : allocs) retire_block(a.first, a.second);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& q : pending)
            for (auto e : q) (void)hipEventDestroy(e);
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }

    void set_mode_precision(ModeHost& m) {
        m.sh = make_shape(m.I, m.F);
        const uint32_t depth = shape_depth(m.sh);
        if (precision == FTL_FP32) {
            const uint32_t keep64 = depth + 1 > fp32_levels ? depth + 1 - fp32_levels : 0;
            m.d64 = std::max<uint32_t>(1, keep64);
        } else {
            m.d64 = depth + 1;
        }
        const uint64_t full = m.d64 >= 63 ? ~0ull : ((1ull << m.d64) - 1);
        m.n64 = std::min<uint64_t>(full, m.sh.node_count);
        m.compact = precision == FTL_FP32 && leaf_mfma && compact_opt && compact_fits(R, depth, m.d64);
    }

    // (Re)allocates the trees of mode k for its current split: tree64 with
    // every node, or the FP64 levels only when compact; tree32 with the FP32
    // levels.  The zeroed tail pad after each covers the draw kernels'
    // over-reads.
    void alloc_trees(ModeHost& m) {
        if (m.tree64) dfree(m.tree64, (m.tree64_nodes * ts64 + tree_pad) * sizeof(double));
        if (m.tree32) dfree(m.tree32, m.tree32_elems * sizeof(float));
        m.tree64 = nullptr;
        m.tree32 = nullptr;
        m.tree32_elems = 0;
        m.tree64_nodes = m.compact ? m.n64 : m.sh.node_count;
        m.tree64 = dalloc<double>(m.tree64_nodes * ts64 + tree_pad);
        FTL_HIP_CHECK(hipMemsetAsync(m.tree64 + m.tree64_nodes * ts64, 0, tree_pad * sizeof(double), stream));
        if (precision == FTL_FP32 && m.sh.node_count > m.n64) {
            const uint64_t n32 = (m.sh.node_count - m.n64) * ts32;
            m.tree32_elems = n32 + tree_pad;
            m.tree32 = dalloc<float>(m.tree32_elems);
            FTL_HIP_CHECK(hipMemsetAsync(m.tree32 + n32, 0, tree_pad * sizeof(float), stream));
        }
    }

    template <int NB>
    void launch_subtree_pairs(uint32_t L, dim3 grid, hipStream_t st, const BuildArgs& ba, uint64_t c0, uint64_t count,
                              uint64_t skip, uint32_t sdep, uint32_t depth) {
        hipLaunchKernelGGL((k_subtree_pairs<NB>), grid, dim3(64u << (L - 1)), pairs_lds(R, L), st, ba, c0, count, L,
                           skip);
        FTL_HIP_CHECK(hipGetLastError());
        if (skip != ~0ull) {
            // one block: as many waves (leaves per batch) as its slots allow
            // (the wider blocks only up to R = 32 / 48, whose registers fit them)
            const size_t l16 = subtree_lds(R, depth, sdep + 1, 16), l8 = subtree_lds(R, depth, sdep + 1, 8);
            bool done = false;
            if constexpr (NB <= 2) {
                if (l16 <= kSubtreeLds) {
                    hipLaunchKernelGGL((k_subtree_build<NB, 16>), dim3(1), dim3(1024), l16, st, ba, skip, (uint64_t)1,
                                       sdep);
                    done = true;
                }
            }
            if constexpr (NB <= 3) {
                if (!done && l8 <= kSubtreeLds) {
                    hipLaunchKernelGGL((k_subtree_build<NB, 8>), dim3(1), dim3(512), l8, st, ba, skip, (uint64_t)1, sdep);
                    done = true;
                }
            }
            if (!done)
                hipLaunchKernelGGL((k_subtree_build<NB, 4>), dim3(1), dim3(256), subtree_lds(R, depth, sdep + 1, 4), st,
                                   ba, skip, (uint64_t)1, sdep);
            FTL_HIP_CHECK(hipGetLastError());
        }
    }

    // Builds on a stream other than the sampler's own occur only for samplers
    // on a borrowed stream, which keep one computeM set on that stream.
    void build_mode(uint32_t k, hipStream_t st = nullptr) {
        if (!st) st = stream;
        if (st == stream) wait_foreign_draws();
        trees_dirty = true;
        ModeHost& m = modes[k];
        BuildArgs ba;
        ba.U = m.U64;
        ba.tree = m.tree64;
        ba.sh = m.sh;
        ba.R = (uint32_t)R;
        ba.ts64 = ts64;
        ba.first_leaf = m.sh.node_count - m.sh.leaf_count;
        // FP32 samplers: the FP32 factor copy and FP32 bottom levels are
        // written by the same kernels (the same (float) of each FP64 value)
        const bool f32 = precision == FTL_FP32;
        ba.U32 = f32 ? m.U32 : nullptr;
        ba.tree32 = f32 ? m.tree32 : nullptr;
        ba.n64 = m.n64;
        ba.ts32 = ts32;
        auto grid_nodes = [](uint64_t n) { return dim3((unsigned)std::min<uint64_t>(std::max<uint64_t>(n, 1), 1u << 20)); };
        if (m.compact) {
            // every node below depth s = d64 - 1 comes from the subtree
            // builds (FP32 only), the nodes of depth s in FP64; the complete
            // FP64 levels above are summed as in the full build
            const uint32_t sdep = m.d64 - 1;
            const uint32_t depth = shape_depth(m.sh);
            const uint32_t L = depth - sdep;
            const uint64_t c0 = (1ull << sdep) - 1, count = 1ull << sdep;
            // the one node of depth s whose subtree straddles a partial bottom
            // level (its leaves at two depths), if any: built by the fold
            const uint64_t span = 1ull << L, q = m.sh.node_count + 1;
            uint64_t skip = ~0ull;
            if (q % span != 0 && q / span >= 1 && q / span - 1 >= c0 && q / span - 1 < c0 + count) skip = q / span - 1;
            if (R <= 16) launch_subtree_pairs<1>(L, grid_nodes(count), st, ba, c0, count, skip, sdep, depth);
            else if (R <= 32) launch_subtree_pairs<2>(L, grid_nodes(count), st, ba, c0, count, skip, sdep, depth);
            else if (R <= 48) launch_subtree_pairs<3>(L, grid_nodes(count), st, ba, c0, count, skip, sdep, depth);
            else launch_subtree_pairs<4>(L, grid_nodes(count), st, ba, c0, count, skip, sdep, depth);
            for (int64_t lv = (int64_t)sdep - 1; lv >= 0;) {
                if (lv >= 2) {
                    const uint64_t cnt = 1ull << (lv - 2);
                    hipLaunchKernelGGL(k_level_sum3, grid_nodes(cnt), dim3(256), 0, st, m.tree64, cnt - 1, cnt, ts64,
                                       (float*)nullptr, m.n64, ts32);
                    lv -= 3;
                } else {
                    const uint64_t cnt = 1ull << lv;
                    hipLaunchKernelGGL(k_level_sum, grid_nodes(cnt), dim3(256), 0, st, m.tree64, cnt - 1, cnt, ts64,
                                       (float*)nullptr, m.n64, ts32);
                    lv -= 1;
                }
                FTL_HIP_CHECK(hipGetLastError());
            }
            return;
        }
        if (f32 && leaf_mfma && R <= 64) {
            // FP32 samplers: leaf Grams on the matrix cores, a wave per leaf
            const uint64_t lblocks = (m.sh.leaf_count + 3) / 4;
            FTL_REQUIRE(lblocks <= 0x7fffffffull, "too many leaves for one launch");
            if (R <= 16)
                hipLaunchKernelGGL(k_leaf_gram_mfma<1>, dim3((unsigned)lblocks), dim3(256), 0, st, ba);
            else if (R <= 32)
                hipLaunchKernelGGL(k_leaf_gram_mfma<2>, dim3((unsigned)lblocks), dim3(256), 0, st, ba);
            else if (R <= 48)
                hipLaunchKernelGGL(k_leaf_gram_mfma<3>, dim3((unsigned)lblocks), dim3(256), 0, st, ba);
            else
                hipLaunchKernelGGL(k_leaf_gram_mfma<4>, dim3((unsigned)lblocks), dim3(256), 0, st, ba);
        } else {
            // leaves per block: their row blocks within 32 KB of LDS (a power
            // of two <= 8, at least one leaf); entries bitwise the oracle's
            const uint64_t leaf_doubles = std::max<uint64_t>(1, m.F * R);
            uint32_t lpb = 1;
            while (lpb < 8 && 2 * lpb * leaf_doubles <= 4096) lpb *= 2;
            const size_t lds = (size_t)lpb * leaf_doubles * sizeof(double) + tri_size(R) * sizeof(uint16_t);
            const uint64_t lblocks = (m.sh.leaf_count + lpb - 1) / lpb;
            FTL_REQUIRE(lblocks <= 0x7fffffffull, "too many leaves for one launch");
            hipLaunchKernelGGL(k_leaf_gram<256>, dim3((unsigned)lblocks), dim3(256), lds, st, ba, lpb);
        }
        FTL_HIP_CHECK(hipGetLastError());
        // internal nodes of the lfill level (parents of a partial bottom level)
        const uint64_t start = m.sh.nodes_before_lfill, end = ba.first_leaf;
        if (end > start) {
            hipLaunchKernelGGL(k_level_sum, grid_nodes(end - start), dim3(256), 0, st, m.tree64, start, end - start,
                               ts64, ba.tree32, ba.n64, ts32);
            FTL_HIP_CHECK(hipGetLastError());
        }
        // levels lfill_level - 1 .. 0 are complete: three per launch
        for (int64_t lv = (int64_t)m.sh.lfill_level - 1; lv >= 0;) {
            if (lv >= 2) {
                const uint64_t count = 1ull << (lv - 2);
                hipLaunchKernelGGL(k_level_sum3, grid_nodes(count), dim3(256), 0, st, m.tree64, count - 1, count,
                                   ts64, ba.tree32, ba.n64, ts32);
                lv -= 3;
            } else {
                const uint64_t count = 1ull << lv;
                hipLaunchKernelGGL(k_level_sum, grid_nodes(count), dim3(256), 0, st, m.tree64, count - 1, count,
                                   ts64, ba.tree32, ba.n64, ts32);
                lv -= 1;
            }
            FTL_HIP_CHECK(hipGetLastError());
        }
    }

    // (Re)allocate and fill the FP32 copy of the levels below d64 of mode k.
    void make_fp32_copy(uint32_t k) {
        ModeHost& m = modes[k];
        const bool was_compact = m.compact;
        set_mode_precision(m);
        if (m.compact || was_compact) {
            // no FP64 copy of the old FP32 levels to convert from (or the new
            // split is compact): rebuild the mode from its FP64 factor
            wait_foreign_draws();
            alloc_trees(m);
            build_mode(k);
            return;
        }
        if (m.tree32) {
            dfree(m.tree32, m.tree32_elems * sizeof(float));
            m.tree32 = nullptr;
            m.tree32_elems = 0;
        }
        if (precision != FTL_FP32 || m.sh.node_count <= m.n64) return;
        wait_foreign_draws();
        const uint64_t n32 = (m.sh.node_count - m.n64) * ts32;
        m.tree32_elems = n32 + tree_pad;
        m.tree32 = dalloc<float>(m.tree32_elems);
        FTL_HIP_CHECK(hipMemsetAsync(m.tree32 + n32, 0, tree_pad * sizeof(float), stream));
        const uint64_t nodes = m.sh.node_count - m.n64;
        hipLaunchKernelGGL(k_tree_to_f32, dim3((unsigned)std::min<uint64_t>(nodes, 1u << 20)), dim3(256), 0, stream,
                           m.tree64, m.tree32, m.n64, nodes, ts64, ts32);
        FTL_HIP_CHECK(hipGetLastError());
    }

    void set_option(const std::string& key, int64_t value) {
        if (key == "group_lanes") {
            FTL_REQUIRE(value == 0 || value == 8 || value == 16 || value == 32 || value == 64,
                        "group_lanes must be 0, 8, 16, 32 or 64");
            sg_override = (uint32_t)value;
        } else if (key == "grid_blocks") {
            FTL_REQUIRE(value >= 0, "grid_blocks must be >= 0");
            grid_override = (uint32_t)value;
        } else if (key == "xcd_buckets") {
            FTL_REQUIRE(value == 0 || value == 1, "xcd_buckets must be 0 or 1");
            xcd_buckets = (uint32_t)value;
        } else if (key == "draw_kernel") {
            FTL_REQUIRE(value >= 0 && value <= 3, "draw_kernel must be 0, 1, 2 or 3");
            draw_kernel = (uint32_t)value;
        } else if (key == "block_threads") {
            FTL_REQUIRE(value == 0 || value == 128 || value == 256 || value == 512,
                        "block_threads must be 0, 128, 256 or 512");
            block_threads = (uint32_t)value;
        } else if (key == "mode_split") {
            FTL_REQUIRE(value == 0 || value == 1, "mode_split must be 0 or 1");
            mode_split = (uint32_t)value;
        } else if (key == "overlap_compute_m") {
            FTL_REQUIRE(value == 0 || value == 1, "overlap_compute_m must be 0 or 1");
            if (cm_stream) FTL_HIP_CHECK(hipStreamSynchronize(cm_stream));
            overlap_cm = (uint32_t)value;
        } else if (key == "jacobi_threads") {
            FTL_REQUIRE(value == 0 || value == 256 || value == 512 || value == 1024,
                        "jacobi_threads must be 0, 256, 512 or 1024");
            jacobi_override = (uint32_t)value;
        } else if (key == "fp32_leaf") {
            FTL_REQUIRE(value == 0 || value == 1, "fp32_leaf must be 0 or 1");
            leaf32 = (uint32_t)value;
        } else if (key == "leaf_mfma" || key == "compact_fp32") {
            FTL_REQUIRE(value == 0 || value == 1, key + " must be 0 or 1");
            (key == "leaf_mfma" ? leaf_mfma : compact_opt) = (uint32_t)value;
            // trees held compact, or due to be, are rebuilt in the new form
            // (none yet while the constructor applies the tuning table)
            for (uint32_t k = 0; k < (uint32_t)modes.size(); k++) {
                ModeHost& m = modes[k];
                const bool was = m.compact;
                set_mode_precision(m);
                if (m.compact != was) {
                    wait_foreign_draws();
                    alloc_trees(m);
                    build_mode(k);
                }
            }
            FTL_HIP_CHECK(hipStreamSynchronize(stream));
        } else if (key == "fp32_levels") {
            FTL_REQUIRE(value >= 0 && value <= 64, "fp32_levels must be in [0, 64]");
            fp32_levels = value == 0 ? kDefaultFp32Levels : (uint32_t)value;
            for (uint32_t k = 0; k < N; k++) make_fp32_copy(k);
            FTL_HIP_CHECK(hipStreamSynchronize(stream));
        } else {
            throw Error{FTL_EINVAL, "unknown option '" + key + "'"};
        }
    }

    // Options of the tuning-table row for this shape (set before any buffer
    // is allocated; later set_option calls override them).  Entries with an
    // unknown key or a value set_option rejects are skipped.
    int tuning_row = -1;
    void apply_tuning(const uint64_t* I) {
        const TuningTable& t = tuning_table();
        tuning_row = t.find(precision, R, max_depth(N, I, R), J);
        if (tuning_row < 0) return;
        for (const auto& kv : t.rows[(size_t)tuning_row].opts) {
            if (kv.first == "fp32_levels") {
                if (kv.second > 0 && kv.second <= 64) fp32_levels = (uint32_t)kv.second;
                continue;
            }
            try {
                set_option(kv.first, kv.second);
            } catch (const Error&) {
            }
        }
    }

    int64_t get_option(const std::string& key) const {
        if (key == "tuning_row") return tuning_row;
        if (key == "group_lanes") return sg_override;
        if (key == "grid_blocks") return grid_override;
        if (key == "mode_split") return mode_split;
        if (key == "draw_kernel") return draw_kernel;
        if (key == "block_threads") return block_threads;
        if (key == "overlap_compute_m") return overlap_cm && n_msets > 1;
        if (key == "xcd_buckets") return xcd_buckets;
        if (key == "jacobi_threads") return jacobi_override ? jacobi_override : jacobi_threads_for(R);
        if (key == "fp32_levels") return fp32_levels;
        if (key == "fp32_leaf") return leaf32;
        if (key == "leaf_mfma") return leaf_mfma;
        if (key == "compact_fp32") return compact_opt;
        if (key == "compact_modes") {
            int64_t c = 0;
            for (const auto& m : modes) c += m.compact ? 1 : 0;
            return c;
        }
        if (key == "fp64_levels") return modes.empty() ? 0 : modes[0].d64;
        throw Error{FTL_EINVAL, "unknown option '" + key + "'"};
    }

    // alias: U are device arrays that the sampler uses in place as its FP64
    // factors (neither copied nor freed; the owner keeps them alive)
    // sync_end: wait for the tree builds before returning (false: they stay
    // queued on the sampler's stream)
    // use_stream: run on this stream (borrowed, not destroyed) instead of a
    // stream of the sampler's own
    void init(uint32_t N_, const uint64_t* I, uint64_t J_, uint64_t R_, const double* const* U, bool on_device,
              int prec, bool alias = false, bool sync_end = true, hipStream_t use_stream = nullptr) {
        FTL_REQUIRE(N_ >= 1 && N_ <= (uint32_t)kMaxModes, "N must be in [1, 16]");
        FTL_REQUIRE(R_ >= 1 && R_ <= 128, "R must be in [1, 128]");
        FTL_REQUIRE(J_ >= 1 && J_ <= (1ull << 40), "J must be in [1, 2^40]");
        FTL_REQUIRE(prec == FTL_FP64 || prec == FTL_FP32, "precision must be FTL_FP64 or FTL_FP32");
        FTL_REQUIRE(I != nullptr && U != nullptr, "null factor list");
        N = N_;
        J = J_;
        R = R_;
        precision = prec;
        ts64 = (uint32_t)tri_stride64(R);
        ts32 = (uint32_t)tri_stride32(R);
        FTL_HIP_CHECK(hipGetDevice(&device));
        if (use_stream) {
            stream = use_stream;
            own_stream = false;
        } else {
            // a blocking stream: its work is ordered after the work callers
            // queued on the legacy default stream (where torch runs by
            // default), so device inputs written there are complete when read
            FTL_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamDefault));
            // computeM reads only the tree roots (ordered by trees_ev) and
            // writes its own buffer set, so its stream need not follow the
            // legacy default stream
            FTL_HIP_CHECK(hipStreamCreateWithFlags(&cm_stream, hipStreamNonBlocking));
            FTL_HIP_CHECK(hipEventCreateWithFlags(&trees_ev, hipEventDisableTiming));
            n_msets = 2;
        }
        // once per process (thread-safe: a throw leaves it to the next call)
        static std::once_flag attrs_once;
        std::call_once(attrs_once, [] {
            FTL_HIP_CHECK(hipFuncSetAttribute((const void*)k_leaf_gram<256>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            const void* sb[] = {(const void*)k_subtree_build<1, 4>,  (const void*)k_subtree_build<2, 4>,
                                (const void*)k_subtree_build<3, 4>,  (const void*)k_subtree_build<4, 4>,
                                (const void*)k_subtree_build<1, 8>,  (const void*)k_subtree_build<2, 8>,
                                (const void*)k_subtree_build<3, 8>,
                                (const void*)k_subtree_build<1, 16>, (const void*)k_subtree_build<2, 16>,
                                (const void*)k_subtree_pairs<1>,     (const void*)k_subtree_pairs<2>,
                                (const void*)k_subtree_pairs<3>,     (const void*)k_subtree_pairs<4>};
            for (const void* f : sb)
                FTL_HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSubtreeLds));
            // computeM blocks hold A ((R+1) R doubles) in LDS, plus V and the
            // eigenvectors when R <= 64: at most 132 KB of dynamic LDS
            const void* cm[] = {(const void*)k_gram_pinv<256>, (const void*)k_eigen_modes<256>,
                                (const void*)k_gram_pinv<512>, (const void*)k_eigen_modes<512>,
                                (const void*)k_gram_pinv<1024>, (const void*)k_eigen_modes<1024>,
                                (const void*)k_gram_pinv_batched<256>, (const void*)k_eigen_modes_batched<256>,
                                (const void*)k_gram_pinv_batched<512>, (const void*)k_eigen_modes_batched<512>,
                                (const void*)k_gram_pinv_batched<1024>, (const void*)k_eigen_modes_batched<1024>};
            for (const void* f : cm)
                FTL_HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 129 * 128 * 8));
        });
        // Tail pad of every tree: the draw kernels read 2 SG ceil(Tmax/(2 SG))
        // entries from a node start (SG up to 512 lanes per sample), Tmax
        // being the packed size for the kernel's compile-time rank bound (R
        // rounded up to a power of two).
        apply_tuning(I);
        uint64_t rmax = 8;
        while (rmax < R) rmax *= 2;
        const uint64_t pad = tri_size(rmax) + kTreePadExtra;
        tree_pad = pad;
        modes.resize(N);
        for (uint32_t k = 0; k < N; k++) {
            FTL_REQUIRE(I[k] >= 1 && I[k] < (1ull << 40), "factor row counts must be in [1, 2^40)");
            FTL_REQUIRE(U[k] != nullptr, "null factor matrix");
            ModeHost& m = modes[k];
            m.I = I[k];
            m.F = std::min<uint64_t>(R, I[k]);
            set_mode_precision(m);
            if (alias) {
                m.U64 = const_cast<double*>(U[k]);
            } else {
                m.U64 = dalloc<double>(m.I * R);
                FTL_HIP_CHECK(hipMemcpyAsync(m.U64, U[k], m.I * R * sizeof(double),
                                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream));
            }
            alloc_trees(m);
            if (precision == FTL_FP32) m.U32 = dalloc<float>(m.I * R);
            build_mode(k);
        }
        const uint64_t R2 = R * R;
        etree_nodes = (uint32_t)(2 * R - 1);
        for (uint32_t q = 0; q < n_msets; q++) {
            MSet& ms = msets[q];
            ms.pinv = dalloc<double>(R2);
            ms.W = dalloc<double>(R2);
            ms.E = dalloc<double>(N * R2);
            ms.etree = dalloc<double>((uint64_t)N * etree_nodes * ts64 + pad);
            ms.work = dalloc<double>((uint64_t)(N + 1) * 3 * (R + 1) * R);
            ms.lam = dalloc<double>((uint64_t)(N + 1) * R);
            FTL_HIP_CHECK(hipMemsetAsync(ms.E, 0, N * R2 * sizeof(double), stream));
            FTL_HIP_CHECK(
                hipMemsetAsync(ms.etree, 0, ((uint64_t)N * etree_nodes * ts64 + pad) * sizeof(double), stream));
            if (n_msets > 1) {
                FTL_HIP_CHECK(hipEventCreateWithFlags(&ms.cm_done, hipEventDisableTiming));
                FTL_HIP_CHECK(hipEventCreateWithFlags(&ms.draw_done, hipEventDisableTiming));
                // the first computeM into the set waits for the zero fills
                FTL_HIP_CHECK(hipEventRecord(ms.draw_done, stream));
            }
        }
        pinv = msets[0].pinv;
        W = msets[0].W;
        E = msets[0].E;
        etree = msets[0].etree;
        work = msets[0].work;
        lam = msets[0].lam;
        h_state = dalloc<double>(J * R);
        st_x = dalloc<double>(J * R);
        st_s = dalloc<double>(J * 4);
        bucket = dalloc<uint32_t>(J);
        blist = dalloc<uint32_t>(J);
        boff = dalloc<uint32_t>(kBuckets + 1);
        trees_dirty = true;
        if (sync_end) FTL_HIP_CHECK(hipStreamSynchronize(stream));
    }

    // Device bytes init(alias = true) allocates, each block rounded up to
    // `align` (the factors are the caller's).
    static size_t footprint(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, int prec, uint32_t fp32_lv,
                            size_t align) {
        auto up = [&](size_t b) { return (std::max<size_t>(b, 16) + align - 1) / align * align; };
        const uint64_t ts64 = tri_stride64(R), ts32 = tri_stride32(R);
        uint64_t rmax = 8;
        while (rmax < R) rmax *= 2;
        const uint64_t pad = tri_size(rmax) + kTreePadExtra;
        size_t total = 0;
        for (uint32_t k = 0; k < N; k++) {
            ModeHost m;
            m.I = I[k];
            m.F = std::min<uint64_t>(R, I[k]);
            m.sh = make_shape(m.I, m.F);
            const uint32_t depth = shape_depth(m.sh);
            uint32_t d64 = depth + 1;
            if (prec == FTL_FP32) d64 = std::max<uint32_t>(1, depth + 1 > fp32_lv ? depth + 1 - fp32_lv : 0);
            const uint64_t full = d64 >= 63 ? ~0ull : ((1ull << d64) - 1);
            const uint64_t n64 = std::min<uint64_t>(full, m.sh.node_count);
            // the default options (matrix-core leaves, compact FP32 trees)
            const bool compact = prec == FTL_FP32 && compact_fits(R, depth, d64);
            total += up(((compact ? n64 : m.sh.node_count) * ts64 + pad) * 8);
            if (prec == FTL_FP32) {
                total += up(m.I * R * 4);
                if (m.sh.node_count > n64) total += up(((m.sh.node_count - n64) * ts32 + pad) * 4);
            }
        }
        const uint64_t R2 = R * R;
        total += up(R2 * 8) * 2 + up(N * R2 * 8) + up(((uint64_t)N * (2 * R - 1) * ts64 + pad) * 8) +
                 up((uint64_t)(N + 1) * 3 * (R + 1) * R * 8) + up((uint64_t)(N + 1) * R * 8);
        total += up(J * R * 8) * 2 + up(J * 4 * 8) + up(J * 4) * 2 + up((kBuckets + 1) * 4);
        return total;
    }

    void update(uint32_t j, const double* Uj, bool on_device) {
        FTL_REQUIRE(j < N, "update_sampler: mode index out of range");
        FTL_REQUIRE(Uj != nullptr, "update_sampler: null factor");
        ModeHost& m = modes[j];
        wait_foreign_draws();
        FTL_HIP_CHECK(hipMemcpyAsync(m.U64, Uj, m.I * R * sizeof(double),
                                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream));
        build_mode(j);
    }

    std::vector<int> sampled_modes(uint32_t j) const {
        std::vector<int> out;
        for (uint32_t k = 0; k < N; k++)
            if (j >= N || k != j) out.push_back((int)k);
        return out;
    }

    ComputeMArgs computeM_args(const std::vector<int>& order, const MSet& ms) {
        ComputeMArgs a{};
        for (uint32_t k = 0; k < N; k++) a.roots[k] = modes[k].tree64;
        for (size_t t = 0; t < order.size(); t++) a.modes[t] = order[t];
        a.m = (int)order.size();
        a.R = (int)R;
        a.pinv = ms.pinv;
        a.W = ms.W;
        a.E = ms.E;
        a.etree = ms.etree;
        a.work = ms.work;
        a.lam_out = ms.lam;
        a.ts64 = ts64;
        a.etree_nodes = etree_nodes;
        return a;
    }

    void compute_m(uint32_t j, hipStream_t st, const MSet& ms, hipEvent_t* evs = nullptr) {
        std::vector<int> order = sampled_modes(j);
        FTL_REQUIRE(!order.empty(), "no mode left to sample (N == 1 and j == 0)");
        ComputeMArgs a = computeM_args(order, ms);
        const uint64_t jac = 3 * (R + 1) * R <= (uint64_t)kJacobiLdsDoubles ? 3 * (R + 1) * R : (R + 1) * R;
        const size_t lds = jac * sizeof(double);
        const uint32_t bs = jacobi_override ? jacobi_override : jacobi_threads_for(R);
        if (bs == 256) launch_cm<256>(a, lds, st, evs);
        else if (bs == 512) launch_cm<512>(a, lds, st, evs);
        else launch_cm<1024>(a, lds, st, evs);
    }

    template <int BS>
    void launch_cm(const ComputeMArgs& a, size_t lds, hipStream_t st, hipEvent_t* evs) {
        if (evs) FTL_HIP_CHECK(hipEventRecord(evs[0], st));
        hipLaunchKernelGGL(k_gram_pinv<BS>, dim3(1), dim3(BS), lds, st, a);
        FTL_HIP_CHECK(hipGetLastError());
        if (evs) FTL_HIP_CHECK(hipEventRecord(evs[1], st));
        hipLaunchKernelGGL(k_eigen_modes<BS>, dim3(a.m), dim3(BS), lds, st, a);
        FTL_HIP_CHECK(hipGetLastError());
        if (evs) FTL_HIP_CHECK(hipEventRecord(evs[2], st));
    }

    template <int RMAX, int SG, bool USE32>
    void launch_draw(const DrawArgs& a, hipStream_t st) {
        constexpr int BS = 256;
        const uint64_t per_block = BS / SG;
        uint64_t blocks = (J + per_block - 1) / per_block;
        if (grid_override) blocks = std::min<uint64_t>(blocks, grid_override);
        blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 20));
        hipLaunchKernelGGL((k_krp_draw<RMAX, SG, BS, USE32>), dim3((unsigned)blocks), dim3(BS), 0, st, a);
        FTL_HIP_CHECK(hipGetLastError());
    }

    template <int RMAX, int SG>
    void launch_draw_p(const DrawArgs& a, hipStream_t st) {
        if (precision == FTL_FP32)
            launch_draw<RMAX, SG, true>(a, st);
        else
            launch_draw<RMAX, SG, false>(a, st);
    }

    template <int RMAX, bool USE32>
    void launch_draw_wave(const DrawArgs& a, hipStream_t st) {
        constexpr int BS = 256;
        uint64_t blocks = (J + 3) / 4;
        if (grid_override) blocks = std::min<uint64_t>(blocks, grid_override);
        blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 20));
        hipLaunchKernelGGL((k_krp_draw_wave<RMAX, BS, USE32>), dim3((unsigned)blocks), dim3(BS), 0, st, a);
        FTL_HIP_CHECK(hipGetLastError());
    }

    // Calls f(RMAX, SG) (as integral constants) for the draw configuration of
    // this sampler.  Measured at N=3, I=2^23, R=32, J=5e4: SG=32 fastest for
    // both precisions (FP64 1.56 ms vs 1.64 at SG=16; FP32 1.15 vs 1.33).
    template <typename F>
    void with_cfg(F&& f) const {
        using std::integral_constant;
        const uint32_t sg = sg_override;
        if (R <= 8) {
            f(integral_constant<int, 8>{}, integral_constant<int, 8>{});
        } else if (R <= 16) {
            f(integral_constant<int, 16>{}, integral_constant<int, 16>{});
        } else if (R <= 32) {
            const uint32_t s = sg ? sg : 32u;
            if (s == 8) f(integral_constant<int, 32>{}, integral_constant<int, 8>{});
            else if (s == 16) f(integral_constant<int, 32>{}, integral_constant<int, 16>{});
            else if (s == 64) f(integral_constant<int, 32>{}, integral_constant<int, 64>{});
            else f(integral_constant<int, 32>{}, integral_constant<int, 32>{});
        } else if (R <= 64) {
            if (sg == 32) f(integral_constant<int, 64>{}, integral_constant<int, 32>{});
            else f(integral_constant<int, 64>{}, integral_constant<int, 64>{});
        } else {
            f(integral_constant<int, 128>{}, integral_constant<int, 64>{});
        }
    }

    // One-sample-per-block kernel: a grid of as many blocks as are resident
    // at once, each looping over samples i = blockIdx.x + k * gridDim.x.
    template <int RMAX, int BS, bool USE32>
    void launch_draw_block(const DrawArgs& a, hipStream_t st) {
        // blocks per CU and CU count, per instantiation (a function-local
        // static: initialised once, thread-safe; a throw leaves it unset)
        struct Occ {
            int resident, cus;
        };
        static const Occ occ = [] {
            int dev = 0, cus = 0, nb = 0;
            FTL_HIP_CHECK(hipGetDevice(&dev));
            FTL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            FTL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_krp_draw_block<RMAX, BS, USE32>, BS, 0));
            return Occ{std::max(1, nb), std::max(1, cus)};
        }();
        uint64_t blocks = std::min<uint64_t>(J, (uint64_t)occ.resident * (uint64_t)occ.cus);
        if (grid_override) blocks = std::min<uint64_t>(blocks, grid_override);
        blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 20));
        hipLaunchKernelGGL((k_krp_draw_block<RMAX, BS, USE32>), dim3((unsigned)blocks), dim3(BS), 0, st, a);
        FTL_HIP_CHECK(hipGetLastError());
    }

    template <int RMAX, bool USE32>
    void launch_draw_block_bs(const DrawArgs& a, hipStream_t st) {
        const uint32_t bs = block_threads ? block_threads : 256u;
        if constexpr (RMAX == 64) {
            if (bs == 128) launch_draw_block<64, 128, USE32>(a, st);
            else launch_draw_block<64, 256, USE32>(a, st);
        } else {
            if (bs == 128) launch_draw_block<128, 128, USE32>(a, st);
            else if (bs == 512) launch_draw_block<128, 512, USE32>(a, st);
            else launch_draw_block<128, 256, USE32>(a, st);
        }
    }

    void dispatch_draw(const DrawArgs& a, hipStream_t st) {
        const uint32_t sg = sg_override;
        if ((draw_kernel == 2 && R > 32) || (draw_kernel == 0 && R > 64)) {
            if (R <= 64) {
                if (precision == FTL_FP32) launch_draw_block_bs<64, true>(a, st);
                else launch_draw_block_bs<64, false>(a, st);
            } else {
                if (precision == FTL_FP32) launch_draw_block_bs<128, true>(a, st);
                else launch_draw_block_bs<128, false>(a, st);
            }
            return;
        }
        if (draw_kernel == 1 && R <= 32 && sg == 0) {
            if (R <= 8) {
                if (precision == FTL_FP32) launch_draw_wave<8, true>(a, st);
                else launch_draw_wave<8, false>(a, st);
            } else if (R <= 16) {
                if (precision == FTL_FP32) launch_draw_wave<16, true>(a, st);
                else launch_draw_wave<16, false>(a, st);
            } else {
                if (precision == FTL_FP32) launch_draw_wave<32, true>(a, st);
                else launch_draw_wave<32, false>(a, st);
            }
            return;
        }
        with_cfg([&](auto rmax, auto sgc) { launch_draw_p<decltype(rmax)::value, decltype(sgc)::value>(a, st); });
    }

    // The bucketed draw of sampled mode a.mi_begin: prefix, partition, suffix.
    void dispatch_bucketed(const DrawArgs& a, hipStream_t st) {
        with_cfg([&](auto rmax, auto sgc) {
            constexpr int RM = decltype(rmax)::value, SG = decltype(sgc)::value, BS = 256;
            const uint64_t per_block = BS / SG;
            const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((J + per_block - 1) / per_block, 1u << 20));
            hipLaunchKernelGGL((k_draw_prefix<RM, SG, BS>), dim3((unsigned)blocks), dim3(BS), 0, st, a);
            FTL_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_bucket_partition<1024>, dim3(1), dim3(1024), 0, st, a.bucket, J, a.blist, a.boff);
            FTL_HIP_CHECK(hipGetLastError());
            // kBuckets x the blocks one bucket of J / kBuckets samples needs
            const uint64_t per_bucket = (J + kBuckets - 1) / kBuckets;
            const uint64_t sblocks = kBuckets * std::max<uint64_t>(1, (per_bucket + per_block - 1) / per_block);
            if (precision == FTL_FP32)
                hipLaunchKernelGGL((k_draw_suffix<RM, SG, BS, true>), dim3((unsigned)std::min<uint64_t>(sblocks, 1u << 20)),
                                   dim3(BS), 0, st, a);
            else
                hipLaunchKernelGGL((k_draw_suffix<RM, SG, BS, false>), dim3((unsigned)std::min<uint64_t>(sblocks, 1u << 20)),
                                   dim3(BS), 0, st, a);
            FTL_HIP_CHECK(hipGetLastError());
        });
    }

    // Kernel arguments of one draw over the sampled modes `order` reading the
    // computeM outputs of set ms (all modes in one launch; the caller sets
    // mi_begin / mi_end / h_state for per-mode launches).
    DrawArgs make_draw_args(const std::vector<int>& order, uint64_t* samples, double* h, double* w,
                            const double* draws, const MSet& ms) const {
        DrawArgs a{};
        for (uint32_t k = 0; k < N; k++) {
            const ModeHost& m = modes[k];
            ModeDev& d = a.mode[k];
            d.tree64 = m.tree64;
            d.tree32 = m.tree32;
            d.U64 = m.U64;
            d.U32 = m.U32;
            d.etree = ms.etree + (uint64_t)k * etree_nodes * ts64;
            d.E = ms.E + (uint64_t)k * R * R;
            d.sh = m.sh;
            d.n64 = m.n64;
            d.d64 = m.d64;
            d.T = (uint32_t)tri_size(R);
        }
        for (size_t t = 0; t < order.size(); t++) a.order[t] = order[t];
        a.m = (int)order.size();
        a.N = (int)N;
        a.J = J;
        a.R = (int)R;
        a.ts64 = ts64;
        a.ts32 = ts32;
        a.esh = make_shape(R, 1);
        a.pinv = ms.pinv;
        a.leaf32 = (int)leaf32;
        a.draws = draws;
        a.seed = seed;
        a.call = call;
        a.samples = samples;
        a.h_out = h;
        a.w_out = w;
        a.split_level = kBucketLevel;
        a.st_x = st_x;
        a.st_s = st_s;
        a.bucket = bucket;
        a.blist = blist;
        a.boff = boff;
        a.mi_begin = 0;
        a.mi_end = a.m;
        a.h_state = nullptr;
        return a;
    }

    void draw_device(uint32_t j, uint64_t* samples, double* h, double* w, const double* draws, hipStream_t st) {
        FTL_REQUIRE(samples != nullptr, "null samples buffer");
        if (!st) st = stream;
        std::vector<int> order = sampled_modes(j);
        FTL_REQUIRE(!order.empty(), "no mode left to sample (N == 1 and j == 0)");
        const bool foreign = st != stream;
        if (foreign) {
            if (!handle_ev) FTL_HIP_CHECK(hipEventCreateWithFlags(&handle_ev, hipEventDisableTiming));
            if (!foreign_ev) FTL_HIP_CHECK(hipEventCreateWithFlags(&foreign_ev, hipEventDisableTiming));
            FTL_HIP_CHECK(hipEventRecord(handle_ev, stream));
            FTL_HIP_CHECK(hipStreamWaitEvent(st, handle_ev, 0));
        }
        std::array<hipEvent_t, 5> evs{};
        if (timing)
            for (auto& e : evs) e = take_event();
        const bool overlap = overlap_cm && n_msets > 1;
        const MSet& ms = msets[overlap ? (mset_next++ & 1) : 0];
        if (overlap) {
            if (trees_dirty) {
                FTL_HIP_CHECK(hipEventRecord(trees_ev, stream));
                trees_dirty = false;
            }
            FTL_HIP_CHECK(hipStreamWaitEvent(cm_stream, trees_ev, 0));
            FTL_HIP_CHECK(hipStreamWaitEvent(cm_stream, ms.draw_done, 0));
            compute_m(j, cm_stream, ms, timing ? evs.data() : nullptr);
            FTL_HIP_CHECK(hipEventRecord(ms.cm_done, cm_stream));
            FTL_HIP_CHECK(hipStreamWaitEvent(st, ms.cm_done, 0));
        } else {
            compute_m(j, st, ms, timing ? evs.data() : nullptr);
        }
        if (timing) FTL_HIP_CHECK(hipEventRecord(evs[3], st));
        DrawArgs a = make_draw_args(order, samples, h, w, draws, ms);
        if (mode_split && h_state) {
            a.h_state = h_state;
            for (int mi = 0; mi < a.m; mi++) {
                a.mi_begin = mi;
                a.mi_end = mi + 1;
                // bucketed when the tree is deep enough for the split to matter
                if (xcd_buckets && draw_kernel == 0 && R <= 64 && modes[a.order[mi]].sh.lfill_level >= 8)
                    dispatch_bucketed(a, st);
                else dispatch_draw(a, st);
            }
        } else {
            a.mi_begin = 0;
            a.mi_end = a.m;
            a.h_state = nullptr;
            dispatch_draw(a, st);
        }
        if (overlap) FTL_HIP_CHECK(hipEventRecord(ms.draw_done, st));
        if (foreign) {
            FTL_HIP_CHECK(hipEventRecord(foreign_ev, st));
            foreign_pending = true;
            foreign_used = true;
        }
        if (timing) {
            FTL_HIP_CHECK(hipEventRecord(evs[4], st));
            pending.push_back(evs);
        }
        if (!draws) call++;
    }

    void ensure_staging() {
        if (!samples_dev) {
            samples_dev = dalloc<uint64_t>(N * J);
            h_dev = dalloc<double>(J * R);
            w_dev = dalloc<double>(J);
            draws_dev = dalloc<double>(2 * N * J);
        }
    }

    void draw_host(uint32_t j, uint64_t* samples, double* h, double* w, const double* draws) {
        FTL_REQUIRE(samples != nullptr, "null samples array");
        ensure_staging();
        if (draws)
            FTL_HIP_CHECK(hipMemcpyAsync(draws_dev, draws, 2 * N * J * sizeof(double), hipMemcpyHostToDevice,
                                         stream));
        draw_device(j, samples_dev, h ? h_dev : nullptr, w ? w_dev : nullptr, draws ? draws_dev : nullptr,
                    stream);
        for (int k : sampled_modes(j))
            FTL_HIP_CHECK(hipMemcpyAsync(samples + (uint64_t)k * J, samples_dev + (uint64_t)k * J,
                                         J * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
        if (h) FTL_HIP_CHECK(hipMemcpyAsync(h, h_dev, J * R * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (w) FTL_HIP_CHECK(hipMemcpyAsync(w, w_dev, J * sizeof(double), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
    }

    void get_tree(uint32_t k, double* out, uint64_t capacity) {
        FTL_REQUIRE(k < N, "mode index out of range");
        const ModeHost& m = modes[k];
        FTL_REQUIRE(capacity >= m.sh.node_count, "output holds fewer nodes than the tree");
        // a compact FP32 tree: the FP64 levels, then the FP32 ones widened
        std::vector<double> packed(m.sh.node_count * ts64);
        FTL_HIP_CHECK(hipMemcpyAsync(packed.data(), m.tree64, m.tree64_nodes * ts64 * sizeof(double),
                                     hipMemcpyDeviceToHost, stream));
        std::vector<float> p32;
        if (m.tree64_nodes < m.sh.node_count) {
            p32.resize((m.sh.node_count - m.n64) * ts32);
            FTL_HIP_CHECK(hipMemcpyAsync(p32.data(), m.tree32, p32.size() * sizeof(float), hipMemcpyDeviceToHost,
                                         stream));
        }
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        for (uint64_t c = m.tree64_nodes; c < m.sh.node_count; c++)
            for (uint32_t e = 0; e < ts64; e++) packed[c * ts64 + e] = (double)p32[(c - m.n64) * ts32 + e];
        const uint64_t R2 = R * R;
        for (uint64_t c = 0; c < m.sh.node_count; c++) {
            double* dst = out + c * R2;
            std::fill(dst, dst + R2, 0.0);
            for (uint64_t a = 0; a < R; a++)
                for (uint64_t b = a; b < R; b++) dst[a * R + b] = packed[c * ts64 + tri_offset(R, a, b)];
        }
    }

    void compute_m_host(uint32_t j, double* pinv_out, double* E_out) {
        if (cm_stream) FTL_HIP_CHECK(hipStreamSynchronize(cm_stream));
        FTL_HIP_CHECK(hipMemsetAsync(E, 0, N * R * R * sizeof(double), stream));
        compute_m(j, stream, msets[0]);
        if (pinv_out)
            FTL_HIP_CHECK(hipMemcpyAsync(pinv_out, pinv, R * R * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (E_out)
            FTL_HIP_CHECK(hipMemcpyAsync(E_out, E, N * R * R * sizeof(double), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
    }
};

}  // namespace ftl

struct ftl_sampler {
    ftl::Sampler impl;
};

namespace ftl {

Sampler* sampler_new_async(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, double* const* U_dev, int precision,
                           DeviceAllocator* alloc, hipStream_t st) {
    std::unique_ptr<Sampler> s(new Sampler());
    s->allocator = alloc;
    s->init(N, I, J, R, U_dev, true, precision, true, false, st);
    return s.release();
}

hipStream_t sampler_stream(Sampler* s) { return s->stream; }

namespace {

constexpr size_t align_up(size_t b, size_t a) { return (b + a - 1) / a * a; }

// Byte layout of one packed launch's arguments (host staging and device copy).
struct PackLayout {
    size_t draw = 0, cm = 0, off = 0, breq = 0, bpos = 0, total = 0;
    PackLayout(uint32_t n, uint32_t blocks) {
        draw = 0;
        cm = align_up(draw + (size_t)n * sizeof(DrawArgs), 256);
        off = align_up(cm + (size_t)n * sizeof(ComputeMArgs), 256);
        breq = align_up(off + (size_t)(n + 1) * sizeof(uint64_t), 256);
        bpos = align_up(breq + (size_t)blocks * sizeof(uint32_t), 256);
        total = align_up(bpos + (size_t)blocks * sizeof(uint32_t), 256);
    }
};

}  // namespace

size_t packed_args_bytes(uint32_t max_req) { return PackLayout(max_req, max_req * kMaxModes).total; }

uint64_t sampler_capacity(const Sampler* s) { return s->J; }

void sampler_set_option(Sampler* s, const char* key, int64_t value) { s->set_option(key, value); }

bool sampler_packable(const Sampler* s) {
    return s->R <= 64 && (s->draw_kernel == 0 || s->draw_kernel == 3) && s->N >= 1;
}

uint64_t sampler_pack_key(const Sampler* s) {
    return (s->R << 8) | ((uint64_t)s->precision << 4) | (uint64_t)(s->sg_override >> 3) |
           ((uint64_t)s->leaf32 << 40) | ((uint64_t)(s->jacobi_override ? s->jacobi_override : jacobi_threads_for(s->R))
                                          << 44);
}

void sampler_draw_packed(const PackItem* items, uint32_t n, void* dev_args, void* host_stage, hipStream_t st) {
    FTL_REQUIRE(n >= 1, "empty pack");
    const Sampler* s0 = items[0].s;
    const uint64_t key = sampler_pack_key(s0);
    std::vector<DrawArgs> da(n);
    std::vector<ComputeMArgs> cm(n);
    std::vector<uint32_t> breq, bpos;
    std::vector<uint64_t> off(n + 1, 0);
    uint32_t sg = 0;
    s0->with_cfg([&](auto, auto sgc) { sg = (uint32_t)decltype(sgc)::value; });
    const uint64_t gpw = 64 / sg;
    for (uint32_t i = 0; i < n; i++) {
        Sampler* s = items[i].s;
        FTL_REQUIRE(sampler_packable(s) && sampler_pack_key(s) == key, "draws of one pack need the same R, precision "
                                                                         "and launch options");
        std::vector<int> order = s->sampled_modes(items[i].j);
        FTL_REQUIRE(!order.empty(), "no mode left to sample (N == 1 and j == 0)");
        cm[i] = s->computeM_args(order, s->msets[0]);
        da[i] = s->make_draw_args(order, items[i].samples, nullptr, nullptr, nullptr, s->msets[0]);
        FTL_REQUIRE(items[i].J >= 1 && items[i].J <= s->J, "a packed draw asks for more samples than its sampler holds");
        da[i].J = items[i].J;
        s->call++;
        for (size_t t = 0; t < order.size(); t++) {
            breq.push_back(i);
            bpos.push_back((uint32_t)t);
        }
        off[i + 1] = off[i] + (items[i].J + gpw - 1) / gpw;
    }
    const uint32_t blocks = (uint32_t)breq.size();
    const PackLayout L(n, blocks);
    char* hs = (char*)host_stage;
    std::memcpy(hs + L.draw, da.data(), n * sizeof(DrawArgs));
    std::memcpy(hs + L.cm, cm.data(), n * sizeof(ComputeMArgs));
    std::memcpy(hs + L.off, off.data(), (n + 1) * sizeof(uint64_t));
    std::memcpy(hs + L.breq, breq.data(), blocks * sizeof(uint32_t));
    std::memcpy(hs + L.bpos, bpos.data(), blocks * sizeof(uint32_t));
    char* dv = (char*)dev_args;
    FTL_HIP_CHECK(hipMemcpyAsync(dv, hs, L.total, hipMemcpyHostToDevice, st));
    const ComputeMArgs* cm_dev = (const ComputeMArgs*)(dv + L.cm);
    const uint64_t R = s0->R;
    const uint64_t jac = 3 * (R + 1) * R <= (uint64_t)kJacobiLdsDoubles ? 3 * (R + 1) * R : (R + 1) * R;
    const size_t lds = jac * sizeof(double);
    const uint32_t bs = s0->jacobi_override ? s0->jacobi_override : jacobi_threads_for(R);
    auto cm_launch = [&](auto bsc) {
        constexpr int BS = decltype(bsc)::value;
        hipLaunchKernelGGL(k_gram_pinv_batched<BS>, dim3(n), dim3(BS), lds, st, cm_dev);
        FTL_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_eigen_modes_batched<BS>, dim3(blocks), dim3(BS), lds, st, cm_dev,
                           (const uint32_t*)(dv + L.breq), (const uint32_t*)(dv + L.bpos));
        FTL_HIP_CHECK(hipGetLastError());
    };
    if (bs == 256) cm_launch(std::integral_constant<int, 256>{});
    else if (bs == 512) cm_launch(std::integral_constant<int, 512>{});
    else cm_launch(std::integral_constant<int, 1024>{});
    s0->with_cfg([&](auto rmax, auto sgc) {
        constexpr int RM = decltype(rmax)::value, SG = decltype(sgc)::value, BS = 256;
        const uint64_t blocks_d = std::max<uint64_t>(1, std::min<uint64_t>((off[n] + BS / 64 - 1) / (BS / 64), 1u << 20));
        if (s0->precision == FTL_FP32)
            hipLaunchKernelGGL((k_krp_draw_packed<RM, SG, BS, true>), dim3((unsigned)blocks_d), dim3(BS), 0, st,
                               (const DrawArgs*)(dv + L.draw), (const uint64_t*)(dv + L.off), n, (int)R);
        else
            hipLaunchKernelGGL((k_krp_draw_packed<RM, SG, BS, false>), dim3((unsigned)blocks_d), dim3(BS), 0, st,
                               (const DrawArgs*)(dv + L.draw), (const uint64_t*)(dv + L.off), n, (int)R);
        FTL_HIP_CHECK(hipGetLastError());
    });
}

size_t sampler_footprint(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, int precision, size_t align) {
    return Sampler::footprint(N, I, J, R, precision, tuned_fp32_levels(N, I, J, R, precision), align);
}

Sampler* sampler_new_aliased(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, double* const* U_dev,
                             int precision) {
    std::unique_ptr<Sampler> s(new Sampler());
    s->init(N, I, J, R, U_dev, true, precision, true);
    return s.release();
}

void sampler_delete(Sampler* s) { delete s; }

void sampler_draw(Sampler* s, uint32_t j, uint64_t* samples, double* h, double* w, const double* draws,
                  hipStream_t st) {
    s->draw_device(j, samples, h, w, draws, st);
}

void sampler_rebuild(Sampler* s, uint32_t j, hipStream_t st) {
    FTL_REQUIRE(j < s->N, "update_sampler: mode index out of range");
    s->build_mode(j, st);
}

SamplerView sampler_view(ftl_sampler* h) {
    Sampler& s = h->impl;
    // the factors are settled: no update queued on the handle's stream (or
    // a caller's draw stream) is still writing them
    s.wait_foreign_draws();
    FTL_HIP_CHECK(hipStreamSynchronize(s.stream));
    SamplerView v{};
    v.N = s.N;
    for (uint32_t k = 0; k < s.N; k++) {
        v.I[k] = s.modes[k].I;
        v.U[k] = s.modes[k].U64;
    }
    v.R = s.R;
    v.J = s.J;
    v.precision = s.precision;
    return v;
}

void sampler_set_seed(Sampler* s, uint64_t seed, uint64_t counter) {
    s->seed = seed;
    s->call = counter;
}

}  // namespace ftl

using ftl::guarded;

extern "C" {

const char* ftl_last_error(void) { return ftl::get_error(); }

const char* ftl_version(void) { return "ftl_hip 0.1 (gfx950)"; }

int ftl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ftl_device_pool_trim(void) {
    return guarded([&] { ftl::device_pool_trim(); });
}

int ftl_device_pool_held(uint64_t* bytes) {
    return guarded([&] {
        FTL_REQUIRE(bytes, "null argument");
        *bytes = (uint64_t)ftl::device_pool_held();
    });
}

int ftl_device_pool_set_cap(uint64_t bytes) {
    return guarded([&] { ftl::device_pool_set_cap((size_t)bytes); });
}

int ftl_device_pool_cap(uint64_t* bytes) {
    return guarded([&] {
        FTL_REQUIRE(bytes, "null argument");
        *bytes = (uint64_t)ftl::device_pool_cap();
    });
}

int ftl_tuning_lookup(const char* table_text, int precision, uint64_t R, uint64_t depth, uint64_t J,
                      char* out, uint64_t capacity) {
    return guarded([&] {
        FTL_REQUIRE(out && capacity > 0, "null output");
        ftl::TuningTable parsed;
        const ftl::TuningTable* t = &ftl::tuning_table();
        if (table_text) {
            parsed = ftl::parse_tuning(table_text);
            t = &parsed;
        }
        const int row = t->find(precision, R, depth, J);
        std::string s = t->path + "\t" + std::to_string(row) + "\t";
        if (row >= 0) {
            bool first = true;
            for (const auto& kv : t->rows[(size_t)row].opts) {
                if (!first) s += ",";
                s += kv.first + "=" + std::to_string(kv.second);
                first = false;
            }
        }
        const size_t n = std::min<size_t>(s.size(), (size_t)capacity - 1);
        std::memcpy(out, s.data(), n);
        out[n] = '\0';
    });
}

int ftl_eigensolver_capped(uint64_t* out) {
    return guarded([&] {
        FTL_REQUIRE(out, "null argument");
        *out = ftl::jacobi_capped_read() + ftl::jacobi_capped_als_tu();
    });
}

static int create_common(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, const double* const* U,
                         bool on_device, int precision, ftl_sampler** out) {
    return guarded([&] {
        FTL_REQUIRE(out != nullptr, "null output handle");
        *out = nullptr;
        int ndev = 0;
        FTL_HIP_CHECK(hipGetDeviceCount(&ndev));
        FTL_REQUIRE(ndev > 0, "no HIP device visible");
        std::unique_ptr<ftl_sampler> s(new ftl_sampler());
        s->impl.init(N, I, J, R, U, on_device, precision);
        *out = s.release();
    });
}

int ftl_sampler_create(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, const double* const* U_host,
                       int precision, ftl_sampler** out) {
    return create_common(N, I, J, R, U_host, false, precision, out);
}

int ftl_sampler_create_device(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, const double* const* U_dev,
                              int precision, ftl_sampler** out) {
    return create_common(N, I, J, R, U_dev, true, precision, out);
}

int ftl_sampler_destroy(ftl_sampler* s) {
    return guarded([&] { delete s; });
}

int ftl_sampler_update(ftl_sampler* s, uint32_t j, const double* Uj_host) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        s->impl.update(j, Uj_host, false);
        FTL_HIP_CHECK(hipStreamSynchronize(s->impl.stream));
    });
}

int ftl_sampler_update_device(ftl_sampler* s, uint32_t j, const double* Uj_dev) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        s->impl.update(j, Uj_dev, true);
    });
}

int ftl_sampler_draw(ftl_sampler* s, uint32_t j, uint64_t* samples_host, const double* draws_host) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        s->impl.draw_host(j, samples_host, nullptr, nullptr, draws_host);
    });
}

int ftl_sampler_draw_materialize(ftl_sampler* s, uint32_t j, uint64_t* samples_host, double* h_host,
                                 double* weights_host, const double* draws_host) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        FTL_REQUIRE(h_host && weights_host, "null h or weights array");
        s->impl.draw_host(j, samples_host, h_host, weights_host, draws_host);
    });
}

int ftl_sampler_draw_device(ftl_sampler* s, uint32_t j, uint64_t* samples_dev, double* h_dev, double* weights_dev,
                            const double* draws_dev, void* stream) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        s->impl.draw_device(j, samples_dev, h_dev, weights_dev, draws_dev, (hipStream_t)stream);
    });
}

int ftl_sampler_set_seed(ftl_sampler* s, uint64_t seed, uint64_t counter) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        s->impl.seed = seed;
        s->impl.call = counter;
    });
}

int ftl_sampler_get_tree(ftl_sampler* s, uint32_t k, double* G_host, uint64_t node_capacity) {
    return guarded([&] {
        FTL_REQUIRE(s && G_host, "null argument");
        s->impl.get_tree(k, G_host, node_capacity);
    });
}

int ftl_sampler_compute_m(ftl_sampler* s, uint32_t j, double* pinv_host, double* E_host) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        s->impl.compute_m_host(j, pinv_host, E_host);
    });
}

int ftl_sampler_info(ftl_sampler* s, uint64_t* out) {
    return guarded([&] {
        FTL_REQUIRE(s && out, "null argument");
        const auto& m = s->impl;
        out[0] = m.N;
        out[1] = m.J;
        out[2] = m.R;
        out[3] = (uint64_t)m.precision;
        out[4] = m.modes[0].d64;
        out[5] = m.bytes_held;
        out[6] = m.modes[0].sh.node_count;
        out[7] = m.modes[0].sh.lfill_level;
    });
}

void* ftl_sampler_stream(ftl_sampler* s) { return s ? (void*)s->impl.stream : nullptr; }

int ftl_sampler_set_option(ftl_sampler* s, const char* key, int64_t value) {
    return guarded([&] {
        FTL_REQUIRE(s && key, "null argument");
        s->impl.set_option(key, value);
    });
}

int ftl_sampler_get_option(ftl_sampler* s, const char* key, int64_t* value) {
    return guarded([&] {
        FTL_REQUIRE(s && key && value, "null argument");
        *value = s->impl.get_option(key);
    });
}

int ftl_sampler_enable_timing(ftl_sampler* s, int enable) {
    return guarded([&] {
        FTL_REQUIRE(s, "null sampler");
        s->impl.timing = enable != 0;
    });
}

int ftl_sampler_last_timing(ftl_sampler* s, double* out4) {
    return guarded([&] {
        FTL_REQUIRE(s && out4, "null argument");
        s->impl.collect_timing();
        for (int i = 0; i < 3; i++) out4[i] = s->impl.last_ms[i];
        out4[3] = (double)s->impl.last_count;
    });
}

}  // extern "C"

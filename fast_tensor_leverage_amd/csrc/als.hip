// als.hip -- host side of the device sparse tensor and the sampled CP-ALS
// step: sorted lookups, the ALS update pipeline, renormalization and the
// exact fit, and their C ABI (include/ftl_hip.h).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/functional.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <array>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "als_kernels.hpp"
#include "ftl_common.hpp"
#include "sampler_internal.hpp"

namespace ftl {

namespace {

inline unsigned grid_for(uint64_t work, unsigned per_block, unsigned cap = 65536) {
    uint64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

inline uint32_t bit_width(uint64_t x) { return x ? 64u - (uint32_t)__builtin_clzll(x) : 0u; }

// Key layout leaving out mode `skip`; fields hold values < 2^bits[k].
KeyLayout make_layout(uint32_t N, uint32_t skip, const uint32_t* bits) {
    KeyLayout L{};
    L.N = N;
    L.skip = skip;
    uint32_t sh = 0;
    for (int k = (int)N - 1; k >= 0; k--) {
        if ((uint32_t)k == skip) continue;
        L.shift[k] = sh;
        L.bits[k] = bits[k];
        sh += bits[k];
    }
    L.total_bits = sh;
    return L;
}

// Device allocations of one handle, freed together.
struct Arena {
    std::vector<std::pair<void*, size_t>> blocks;
    size_t bytes = 0;

    template <typename T>
    T* alloc(uint64_t count) {
        void* p = nullptr;
        const size_t b = (size_t)count * sizeof(T);
        FTL_HIP_CHECK(hipMalloc(&p, b ? b : 16));
        blocks.push_back({p, b});
        bytes += b;
        return (T*)p;
    }
    void release(void* p) {
        if (!p) return;
        for (size_t i = 0; i < blocks.size(); i++)
            if (blocks[i].first == p) {
                bytes -= blocks[i].second;
                blocks.erase(blocks.begin() + (long)i);
                break;
            }
        (void)hipFree(p);
    }
    ~Arena() {
        for (auto& b : blocks) (void)hipFree(b.first);
    }
};

// Waits for a stream when it goes out of scope: declared after an Arena whose
// buffers the stream's queued work uses, it keeps an early exit (exception)
// from freeing memory that kernels still read.
struct StreamDrain {
    hipStream_t st;
    ~StreamDrain() {
        if (st) (void)hipStreamSynchronize(st);
    }
};

// Growable temporary storage of the rocPRIM primitives.
struct PrimTemp {
    void* p = nullptr;
    size_t bytes = 0;
    void* ensure(size_t b) {
        if (b > bytes) {
            if (p) (void)hipFree(p);
            p = nullptr;
            FTL_HIP_CHECK(hipMalloc(&p, b));
            bytes = b;
        }
        return p;
    }
    ~PrimTemp() {
        if (p) (void)hipFree(p);
    }
};

// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits.
template <typename K, typename V>
void sort_pairs(PrimTemp& t, const K* kin, K* kout, const V* vin, V* vout, uint64_t n, uint32_t bits,
                hipStream_t st) {
    const int end_bit = (int)std::max<uint32_t>(1, std::min<uint32_t>(bits, 8 * sizeof(K)));
    size_t b = 0;
    FTL_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, b, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
    void* p = t.ensure(b);
    FTL_HIP_CHECK(rocprim::radix_sort_pairs(p, b, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
}

void inclusive_sum(PrimTemp& t, const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t st) {
    size_t b = 0;
    FTL_HIP_CHECK(rocprim::inclusive_scan(nullptr, b, in, out, (size_t)n, rocprim::plus<uint32_t>(), st));
    void* p = t.ensure(b);
    FTL_HIP_CHECK(rocprim::inclusive_scan(p, b, in, out, (size_t)n, rocprim::plus<uint32_t>(), st));
}

void inclusive_max(PrimTemp& t, const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t st) {
    size_t b = 0;
    FTL_HIP_CHECK(rocprim::inclusive_scan(nullptr, b, in, out, (size_t)n, rocprim::maximum<uint32_t>(), st));
    void* p = t.ensure(b);
    FTL_HIP_CHECK(rocprim::inclusive_scan(p, b, in, out, (size_t)n, rocprim::maximum<uint32_t>(), st));
}

void exclusive_sum(PrimTemp& t, const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t st) {
    size_t b = 0;
    FTL_HIP_CHECK(rocprim::exclusive_scan(nullptr, b, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), st));
    void* p = t.ensure(b);
    FTL_HIP_CHECK(rocprim::exclusive_scan(p, b, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), st));
}

#define FTL_LAUNCH_CHECK() FTL_HIP_CHECK(hipGetLastError())

// ids[m] = perm[lo[u] + m - off[u]] for the match numbering of k_expand.
__global__ void k_match_ids(const uint32_t* __restrict__ off, const uint32_t* __restrict__ lo, uint64_t n, uint64_t T,
                            const uint32_t* __restrict__ perm, uint64_t* __restrict__ ids) {
    for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < T; m += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t l = 0, h = n;
        while (h - l > 1) {
            const uint64_t mid = (l + h) >> 1;
            if (off[mid] <= m) l = mid;
            else h = mid;
        }
        ids[m] = perm[(uint64_t)lo[l] + (m - off[l])];
    }
}

// samples (n x N uint64, row-major) -> uint32 tuples (the reference's cast,
// sparse_tensor.hpp:95-115)
__global__ void k_tuples_u32(const uint64_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ out) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x)
        out[e] = (uint32_t)in[e];
}

}  // namespace

// ---------------------------------------------------------------------------
// matched-nonzero pipeline shared by the plain MTTKRP and the ALS update
// ---------------------------------------------------------------------------
struct Matches {
    uint64_t cap_n = 0, cap_t = 0, cap_r = 0;
    uint32_t *lo = nullptr, *cnt = nullptr, *off = nullptr;
    uint32_t *mrow = nullptr, *srow = nullptr, *flags = nullptr, *segid = nullptr;
    uint32_t *seg_start = nullptr, *seg_row = nullptr, *npieces = nullptr, *piece_off = nullptr;
    uint32_t *piece_seg = nullptr, *piece_mark = nullptr, *heavy = nullptr;
    uint64_t *mpay = nullptr, *spay = nullptr;
    double* partial = nullptr;
    uint32_t* counters = nullptr;       // device, kCntSlots
    uint32_t* host_counters = nullptr;  // pinned host mirror

    void ensure_n(Arena& ar, uint64_t n) {
        if (n <= cap_n) return;
        ar.release(lo);
        ar.release(cnt);
        ar.release(off);
        lo = ar.alloc<uint32_t>(n);
        cnt = ar.alloc<uint32_t>(n);
        off = ar.alloc<uint32_t>(n);
        cap_n = n;
    }

    void ensure_t(Arena& ar, uint64_t T, uint64_t R) {
        if (T <= cap_t && R <= cap_r) return;
        uint64_t c = std::max<uint64_t>(T, cap_t + cap_t / 2);
        for (void* p : {(void*)mrow, (void*)srow, (void*)flags, (void*)segid, (void*)seg_start, (void*)seg_row,
                        (void*)npieces, (void*)piece_off, (void*)piece_seg, (void*)piece_mark, (void*)heavy,
                        (void*)mpay, (void*)spay,
                        (void*)partial})
            ar.release(p);
        mrow = ar.alloc<uint32_t>(c);
        srow = ar.alloc<uint32_t>(c);
        flags = ar.alloc<uint32_t>(c);
        segid = ar.alloc<uint32_t>(c);
        seg_start = ar.alloc<uint32_t>(c + 1);
        seg_row = ar.alloc<uint32_t>(c);
        npieces = ar.alloc<uint32_t>(c);
        piece_off = ar.alloc<uint32_t>(c);
        mpay = ar.alloc<uint64_t>(c);
        spay = ar.alloc<uint64_t>(c);
        heavy = ar.alloc<uint32_t>(c);
        piece_seg = ar.alloc<uint32_t>(c + c / kSegPiece + 1);
        piece_mark = ar.alloc<uint32_t>(c + c / kSegPiece + 1);
        // pieces W <= D + T / kSegPiece <= T + T / kSegPiece + 1
        cap_r = std::max<uint64_t>(R, cap_r);
        partial = ar.alloc<double>((c + c / kSegPiece + 1) * cap_r);
        cap_t = c;
    }

    void init(Arena& ar) {
        counters = ar.alloc<uint32_t>(kCntSlots);
        FTL_HIP_CHECK(hipHostMalloc((void**)&host_counters, kCntSlots * sizeof(uint32_t), hipHostMallocDefault));
        for (int i = 0; i < kCntSlots; i++) host_counters[i] = 0;
    }

    ~Matches() {
        if (host_counters) (void)hipHostFree(host_counters);
    }
};

struct SparseDev {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t N = 0;
    uint64_t nnz = 0;
    uint64_t dims[kMaxModes] = {};
    uint32_t bits[kMaxModes] = {};
    KeyLayout layout[kMaxModes] = {};
    uint64_t* keys[kMaxModes] = {};
    uint32_t* perm[kMaxModes] = {};
    uint32_t* row[kMaxModes] = {};
    double* val[kMaxModes] = {};
    double normsq = 0.0;
    Arena mem;
    PrimTemp prim;
    Matches mt;
    uint32_t* tup = nullptr;
    uint64_t tup_cap = 0;
    // exact MTTKRP of mode j: lookup-j positions sorted stably by the mode-j
    // coordinate, the row segments (D rows) and their reduction pieces (W),
    // built on first use
    struct RowOrder {
        uint32_t *pos = nullptr, *seg_start = nullptr, *seg_row = nullptr, *piece_off = nullptr;
        uint32_t *piece_seg = nullptr, *heavy = nullptr;  // segment of each piece; rows of > kHeavyPieces pieces
        uint32_t* counters = nullptr;  // kCntSlots, kCntRows = D, kCntPieces = W
        uint64_t D = 0, W = 0;
    };
    RowOrder rows_of[kMaxModes];
    // lookup method of the probes: 0 = binary search of the sorted keys,
    // 1 = the per-mode coordinate hash (built by use_hash)
    int method = 0;
    HashSlot* table[kMaxModes] = {};
    uint64_t table_mask[kMaxModes] = {};

    // Builds the coordinate hash of every mode (once) and probes with it.
    void use_hash() {
        for (uint32_t j = 0; j < N; j++)
            if (layout[j].total_bits >= 64)
                throw Error{FTL_EUNSUPPORTED, "the coordinate hash needs keys of fewer than 64 bits (mode " +
                                                  std::to_string(j) + " uses 64)"};
        unsigned long long* cnt_dev = mem.alloc<unsigned long long>(1);
        for (uint32_t j = 0; j < N; j++) {
            if (table[j]) continue;
            FTL_HIP_CHECK(hipMemsetAsync(cnt_dev, 0, sizeof(unsigned long long), stream));
            hipLaunchKernelGGL(k_count_heads, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, stream, keys[j], nnz,
                               cnt_dev);
            FTL_LAUNCH_CHECK();
            unsigned long long distinct = 0;
            FTL_HIP_CHECK(hipMemcpyAsync(&distinct, cnt_dev, sizeof(distinct), hipMemcpyDeviceToHost, stream));
            FTL_HIP_CHECK(hipStreamSynchronize(stream));
            uint64_t slots = 2;
            while (slots < 2 * (uint64_t)distinct) slots *= 2;
            table[j] = mem.alloc<HashSlot>(slots);
            table_mask[j] = slots - 1;
            FTL_HIP_CHECK(hipMemsetAsync(table[j], 0xff, slots * sizeof(HashSlot), stream));
            hipLaunchKernelGGL(k_hash_build, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, stream, keys[j], nnz,
                               table[j], table_mask[j]);
            FTL_LAUNCH_CHECK();
        }
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        mem.release(cnt_dev);
        method = 1;
    }

    ~SparseDev() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    void build(uint32_t N_, uint64_t nnz_, const uint32_t* idx, const double* values, const uint64_t* dims_in,
               bool on_device) {
        FTL_REQUIRE(N_ >= 1 && N_ <= (uint32_t)kMaxModes, "N must be in [1, 16]");
        FTL_REQUIRE(nnz_ >= 1 && nnz_ < (1ull << 32), "nnz must be in [1, 2^32)");
        FTL_REQUIRE(idx && values, "null index or value array");
        N = N_;
        nnz = nnz_;
        FTL_HIP_CHECK(hipGetDevice(&device));
        // blocking stream: ordered after the caller's default-stream work
        FTL_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamDefault));
        mt.init(mem);
        // staging copies of the input (freed at the end of construction)
        uint32_t* d_idx = nullptr;
        double* d_val = nullptr;
        Arena stage;
        StreamDrain drain{stream};
        if (on_device) {
            d_idx = const_cast<uint32_t*>(idx);
            d_val = const_cast<double*>(values);
        } else {
            d_idx = stage.alloc<uint32_t>(nnz * N);
            d_val = stage.alloc<double>(nnz);
            FTL_HIP_CHECK(hipMemcpyAsync(d_idx, idx, nnz * N * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
            FTL_HIP_CHECK(hipMemcpyAsync(d_val, values, nnz * sizeof(double), hipMemcpyHostToDevice, stream));
        }
        // mode sizes: given, or max coordinate + 1
        uint32_t* d_max = stage.alloc<uint32_t>(kMaxModes);
        FTL_HIP_CHECK(hipMemsetAsync(d_max, 0, kMaxModes * sizeof(uint32_t), stream));
        hipLaunchKernelGGL(k_coord_max, dim3(grid_for(nnz, 256, 2048)), dim3(256), 0, stream, d_idx, nnz, N, d_max);
        FTL_LAUNCH_CHECK();
        uint32_t hmax[kMaxModes];
        FTL_HIP_CHECK(hipMemcpyAsync(hmax, d_max, sizeof(hmax), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        for (uint32_t k = 0; k < N; k++) {
            if (dims_in) {
                FTL_REQUIRE(dims_in[k] >= (uint64_t)hmax[k] + 1,
                            "a coordinate of mode " + std::to_string(k) + " is >= its given size");
                dims[k] = dims_in[k];
            } else {
                dims[k] = (uint64_t)hmax[k] + 1;
            }
            bits[k] = bit_width(dims[k] - 1);
        }
        for (uint32_t j = 0; j < N; j++) {
            layout[j] = make_layout(N, j, bits);
            if (layout[j].total_bits > 64)
                throw Error{FTL_EUNSUPPORTED, "coordinates of the modes other than " + std::to_string(j) +
                                                  " need more than 64 key bits"};
        }
        // per mode: stable sort of the nonzeros by their key, then the
        // mode-j coordinate and the value in that order
        uint64_t* kin = stage.alloc<uint64_t>(nnz);
        uint32_t* pin = stage.alloc<uint32_t>(nnz);
        const unsigned g = grid_for(nnz, 256, 8192);
        for (uint32_t j = 0; j < N; j++) {
            keys[j] = mem.alloc<uint64_t>(nnz);
            perm[j] = mem.alloc<uint32_t>(nnz);
            row[j] = mem.alloc<uint32_t>(nnz);
            val[j] = mem.alloc<double>(nnz);
            hipLaunchKernelGGL(k_nz_keys, dim3(g), dim3(256), 0, stream, d_idx, nnz, layout[j], kin, pin);
            FTL_LAUNCH_CHECK();
            sort_pairs(prim, kin, keys[j], pin, perm[j], nnz, layout[j].total_bits, stream);
            hipLaunchKernelGGL(k_nz_gather, dim3(g), dim3(256), 0, stream, d_idx, d_val, perm[j], nnz, N, j, row[j],
                               val[j]);
            FTL_LAUNCH_CHECK();
        }
        // ||X||^2
        double* part = stage.alloc<double>(1024 + 1);
        hipLaunchKernelGGL(k_sumsq_partial<256>, dim3(1024), dim3(256), 0, stream, d_val, nnz, part);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, stream, part, 1024u, part + 1024);
        FTL_LAUNCH_CHECK();
        FTL_HIP_CHECK(hipMemcpyAsync(&normsq, part + 1024, sizeof(double), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        // stage's buffers are freed here (after the stream drained)
    }

    uint32_t* tuples(uint64_t n) {
        if (n > tup_cap) {
            mem.release(tup);
            tup = mem.alloc<uint32_t>(n * N);
            tup_cap = n;
        }
        return tup;
    }

    // The row order of mode j (sort_lookup.hpp:199-254 visits the nonzeros in
    // lookup-j order and scatters to rows; here the same nonzeros are grouped
    // by row, keeping the lookup order within a row, so that each row is one
    // deterministic segmented sum).  One host wait, once per mode.
    const RowOrder& row_order(uint32_t j) {
        RowOrder& o = rows_of[j];
        if (o.pos) return o;
        Arena stage;
        StreamDrain drain{stream};
        uint32_t* iota = stage.alloc<uint32_t>(nnz);
        uint32_t* srow = stage.alloc<uint32_t>(nnz);
        uint32_t* flags = stage.alloc<uint32_t>(nnz);
        uint32_t* segid = stage.alloc<uint32_t>(nnz);
        uint32_t* start = stage.alloc<uint32_t>(nnz + 1);
        uint32_t* srows = stage.alloc<uint32_t>(nnz);
        uint32_t* npieces = stage.alloc<uint32_t>(nnz);
        uint32_t* poff = stage.alloc<uint32_t>(nnz);
        o.pos = mem.alloc<uint32_t>(nnz);
        o.counters = mem.alloc<uint32_t>(kCntSlots);
        FTL_HIP_CHECK(hipMemsetAsync(o.counters, 0, kCntSlots * sizeof(uint32_t), stream));
        const unsigned g = grid_for(nnz, 256, 8192);
        hipLaunchKernelGGL(k_iota, dim3(g), dim3(256), 0, stream, iota, nnz);
        FTL_LAUNCH_CHECK();
        sort_pairs(prim, row[j], srow, iota, o.pos, nnz, bits[j], stream);
        hipLaunchKernelGGL(k_head_flags<uint32_t>, dim3(g), dim3(256), 0, stream, srow, nnz, flags);
        FTL_LAUNCH_CHECK();
        inclusive_sum(prim, flags, segid, nnz, stream);
        hipLaunchKernelGGL(k_seg_heads, dim3(g), dim3(256), 0, stream, srow, segid, nnz, start, srows, o.counters);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_seg_pieces, dim3(g), dim3(256), 0, stream, start, o.counters, nnz, npieces);
        FTL_LAUNCH_CHECK();
        exclusive_sum(prim, npieces, poff, nnz, stream);
        hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, stream, poff, npieces, nnz, o.counters,
                           (int)kCntPieces);
        FTL_LAUNCH_CHECK();
        uint32_t hc[kCntSlots];
        FTL_HIP_CHECK(hipMemcpyAsync(hc, o.counters, sizeof(hc), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        o.D = hc[kCntRows];
        o.W = hc[kCntPieces];
        o.seg_start = mem.alloc<uint32_t>(o.D + 1);
        o.seg_row = mem.alloc<uint32_t>(o.D);
        o.piece_off = mem.alloc<uint32_t>(o.D);
        FTL_HIP_CHECK(hipMemcpyAsync(o.seg_start, start, (o.D + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                     stream));
        FTL_HIP_CHECK(hipMemcpyAsync(o.seg_row, srows, o.D * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
        FTL_HIP_CHECK(hipMemcpyAsync(o.piece_off, poff, o.D * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
        // segment of every piece (max-scan of the segment starts) and the heavy rows
        o.piece_seg = mem.alloc<uint32_t>(o.W);
        o.heavy = mem.alloc<uint32_t>(o.D);
        FTL_HIP_CHECK(hipMemsetAsync(flags, 0, o.W * sizeof(uint32_t), stream));
        hipLaunchKernelGGL(k_mark_segs, dim3(grid_for(o.D, 256, 8192)), dim3(256), 0, stream, o.piece_off, o.counters,
                           flags);
        FTL_LAUNCH_CHECK();
        inclusive_max(prim, flags, o.piece_seg, o.W, stream);
        hipLaunchKernelGGL(k_heavy_list, dim3(grid_for(o.D, 256, 2048)), dim3(256), 0, stream, o.piece_off, o.counters,
                           o.heavy);
        FTL_LAUNCH_CHECK();
        return o;
        // stage's buffers are freed after the stream drained
    }

    ProbeArgs probe_args(uint32_t j, const uint32_t* t, const uint32_t* n_dev, uint64_t n_host, uint64_t n_cap,
                         Matches& m) const {
        ProbeArgs a{};
        a.tup = t;
        a.n_dev = n_dev;
        a.n_host = n_host;
        a.n_cap = n_cap;
        a.keys = keys[j];
        a.nnz = nnz;
        a.L = layout[j];
        for (uint32_t k = 0; k < N; k++) a.dims[k] = dims[k];
        a.lo = m.lo;
        a.cnt = m.cnt;
        a.table = method == 1 ? table[j] : nullptr;
        a.mask = table_mask[j];
        return a;
    }
};

// Probe n tuples (n from n_dev when given) on mode j's lookup: per-sample
// ranges, their exclusive scan and the match total in counters[kCntMatches].
void probe_and_count(SparseDev& X, uint32_t j, const uint32_t* tup, const uint32_t* n_dev, uint64_t n_host,
                     uint64_t n_cap, Matches& m, Arena& ar, PrimTemp& prim, hipStream_t st) {
    m.ensure_n(ar, n_cap);
    ProbeArgs pa = X.probe_args(j, tup, n_dev, n_host, n_cap, m);
    hipLaunchKernelGGL(k_probe, dim3(grid_for(n_cap, 256)), dim3(256), 0, st, pa);
    FTL_LAUNCH_CHECK();
    exclusive_sum(prim, m.cnt, m.off, n_cap, st);
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, st, m.off, m.cnt, n_cap, m.counters, (int)kCntMatches);
    FTL_LAUNCH_CHECK();
}

// Returns the block count of the row-finalise kernel (the number of column
// partials written to colsq when it is given).
// Rows of more than kHeavyPieces pieces get their pieces summed by a block
// each (into the row's first piece) before the row finalise.
void reduce_heavy(const uint32_t* piece_off, uint32_t* counters, uint32_t* list, uint64_t D_cap, uint32_t R,
                  double* partial, hipStream_t st) {
    FTL_HIP_CHECK(hipMemsetAsync(counters + kCntHeavy, 0, sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_heavy_list, dim3(grid_for(D_cap, 256, 2048)), dim3(256), 0, st, piece_off, counters, list);
    FTL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_seg_heavy_sum, dim3(512), dim3(256), 0, st, piece_off, counters, list, R, partial);
    FTL_LAUNCH_CHECK();
}

// ldl: the stride of the lhs rows; an even stride from a 16-byte aligned
// base (16-byte aligned rows) lets the pieces read column pairs
// (k_item_partial_pairs, R <= 64).
template <int G, int CPL>
unsigned launch_reduce(const SparseDev& X, uint32_t j, Matches& m, uint64_t T, const double* lhs, uint32_t R,
                       uint32_t ldl, const double* pinv, double* out, double* colsq, hipEvent_t join,
                       hipStream_t st) {
    constexpr int GPB = 256 / G;
    const unsigned blocks = grid_for(T, GPB, 16384);
    if (ldl % 2 == 0 && R <= 64 && ((uintptr_t)lhs & 15) == 0) {
        if (R <= 32)
            hipLaunchKernelGGL((k_item_partial_pairs<16>), dim3(grid_for(T, 16, 16384)), dim3(256), 0, st, m.seg_start,
                               m.piece_off, m.piece_seg, m.counters, m.spay, X.val[j], lhs, R, ldl, m.partial);
        else
            hipLaunchKernelGGL((k_item_partial_pairs<32>), dim3(grid_for(T, 8, 16384)), dim3(256), 0, st, m.seg_start,
                               m.piece_off, m.piece_seg, m.counters, m.spay, X.val[j], lhs, R, ldl, m.partial);
    } else {
        hipLaunchKernelGGL((k_item_partial<G, CPL>), dim3(blocks), dim3(256), 0, st, m.seg_start, m.piece_off,
                           m.piece_seg, m.counters, m.spay, X.val[j], lhs, R, m.partial);
    }
    FTL_LAUNCH_CHECK();
    reduce_heavy(m.piece_off, m.counters, m.heavy, std::min<uint64_t>(T, X.dims[j]), R, m.partial, st);
    if (join) FTL_HIP_CHECK(hipStreamWaitEvent(st, join, 0));
    const unsigned fblocks = grid_for(T, GPB, kSegBlocksMax);
    if (pinv)
        hipLaunchKernelGGL((k_seg_final<G, CPL, true>), dim3(fblocks), dim3(256), 0, st, m.seg_row, m.piece_off,
                           m.counters, m.partial, R, pinv, out, colsq);
    else
        hipLaunchKernelGGL((k_seg_final<G, CPL, false>), dim3(fblocks), dim3(256), 0, st, m.seg_row, m.piece_off,
                           m.counters, m.partial, R, pinv, out, (double*)nullptr);
    FTL_LAUNCH_CHECK();
    return fblocks;
}

// Given probed ranges of n samples and T >= 1 matches: list the matches,
// sort them by output row (stable), form the row segments and their pieces,
// reduce, and write each touched row of out (times pinv when given).
// join (or null): an event the row finalise waits for before it writes out
// (out's zero fill on another stream).
unsigned reduce_matches(SparseDev& X, uint32_t j, uint64_t n, uint64_t T, const double* lhs, uint32_t R,
                        uint32_t ldl, const double* pinv, double* out, double* colsq, hipEvent_t join, Matches& m,
                        Arena& ar, PrimTemp& prim, hipStream_t st) {
    m.ensure_t(ar, T, R);
    // owner of every match (flags and segid are free until the row sort)
    FTL_HIP_CHECK(hipMemsetAsync(m.flags, 0, T * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_mark_owners, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, m.off, m.cnt, n, m.flags);
    FTL_LAUNCH_CHECK();
    inclusive_max(prim, m.flags, m.segid, T, st);
    hipLaunchKernelGGL(k_expand, dim3(grid_for(T, 256)), dim3(256), 0, st, m.segid, m.off, m.lo, T, X.row[j],
                       m.mrow, m.mpay);
    FTL_LAUNCH_CHECK();
    sort_pairs(prim, m.mrow, m.srow, m.mpay, m.spay, T, X.bits[j], st);
    hipLaunchKernelGGL(k_head_flags<uint32_t>, dim3(grid_for(T, 256)), dim3(256), 0, st, m.srow, T, m.flags);
    FTL_LAUNCH_CHECK();
    inclusive_sum(prim, m.flags, m.segid, T, st);
    hipLaunchKernelGGL(k_seg_heads, dim3(grid_for(T, 256)), dim3(256), 0, st, m.srow, m.segid, T, m.seg_start,
                       m.seg_row, m.counters);
    FTL_LAUNCH_CHECK();
    // bounds known on the host: D <= dcap (the segments are distinct rows of
    // mode j) and W <= wcap (a segment of n >= 1 matches has ceil(n /
    // kSegPiece) <= min(n, n / kSegPiece + 1) pieces); the scans below run
    // over these prefixes, which hold every entry that is read
    const uint64_t dcap = std::min<uint64_t>(T, X.dims[j]);
    const uint64_t wcap = std::min<uint64_t>(T, dcap + (T + kSegPiece - 1) / kSegPiece);
    hipLaunchKernelGGL(k_seg_pieces, dim3(grid_for(dcap, 256)), dim3(256), 0, st, m.seg_start, m.counters, dcap,
                       m.npieces);
    FTL_LAUNCH_CHECK();
    exclusive_sum(prim, m.npieces, m.piece_off, dcap, st);
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, st, m.piece_off, m.npieces, dcap, m.counters,
                       (int)kCntPieces);
    FTL_LAUNCH_CHECK();
    // segment of every piece
    FTL_HIP_CHECK(hipMemsetAsync(m.piece_mark, 0, wcap * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_mark_segs, dim3(grid_for(dcap, 256, 8192)), dim3(256), 0, st, m.piece_off, m.counters,
                       m.piece_mark);
    FTL_LAUNCH_CHECK();
    inclusive_max(prim, m.piece_mark, m.piece_seg, wcap, st);
    if (R <= 32) return launch_reduce<32, 1>(X, j, m, T, lhs, R, ldl, pinv, out, colsq, join, st);
    if (R <= 64) return launch_reduce<64, 1>(X, j, m, T, lhs, R, ldl, pinv, out, colsq, join, st);
    return launch_reduce<64, 2>(X, j, m, T, lhs, R, ldl, pinv, out, colsq, join, st);
}

// ---------------------------------------------------------------------------
// Larsen-Kolda leverage sampler (larsen_kolda_sampler.hpp:18-136): per factor
// k the row leverage scores of U_k (pinv of U_k^T U_k with the 1e-11 cut and
// target rank min(I_k, R), compute_pinv, common.h:271-323), their prefix sums
// and total; a draw samples every mode k != j independently from them.
// ---------------------------------------------------------------------------
// The one-block pinv kernels hold A ((R+1) R doubles) in LDS, plus V and
// the eigenvectors when R <= 64: allow up to 132 KB of dynamic LDS.
void pinv_lds_attrs() {
    static std::once_flag once;  // thread-safe; a throw leaves it to the next call
    std::call_once(once, [] {
        FTL_HIP_CHECK(hipFuncSetAttribute((const void*)k_pinv_partials<256>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 129 * 128 * 8));
        FTL_HIP_CHECK(hipFuncSetAttribute((const void*)k_pinv_partials<1024>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 129 * 128 * 8));
    });
}

template <int RMAX>
void launch_gram_rows(const double* X, uint64_t n, uint32_t R, double* parts, unsigned nb, hipStream_t st) {
    hipLaunchKernelGGL((k_gram_partial<256, RMAX>), dim3(nb), dim3(256), 0, st, X, (const uint32_t*)nullptr, n, R,
                       parts);
    FTL_LAUNCH_CHECK();
}

void gram_rows(const double* X, uint64_t n, uint32_t R, double* parts, unsigned nb, hipStream_t st) {
    if (R <= 8) launch_gram_rows<8>(X, n, R, parts, nb, st);
    else if (R <= 16) launch_gram_rows<16>(X, n, R, parts, nb, st);
    else if (R <= 32) launch_gram_rows<32>(X, n, R, parts, nb, st);
    else if (R <= 64) launch_gram_rows<64>(X, n, R, parts, nb, st);
    else launch_gram_rows<128>(X, n, R, parts, nb, st);
}

// rows of the hybrid sampler's candidate lists: lev[i] >= *cut
struct LevAtLeast {
    const double* lev;
    const double* cut;
    __device__ bool operator()(const uint64_t& i) const { return lev[i] >= *cut; }
};

struct LKSampler {
    uint32_t N = 0;
    uint64_t J = 0, R = 0;
    uint64_t I[kMaxModes] = {};
    double* U[kMaxModes] = {};  // device factors (owned, or the ALS handle's)
    double *lev[kMaxModes] = {}, *cum[kMaxModes] = {}, *off[kMaxModes] = {};
    double *total = nullptr, *tsum = nullptr, *gparts = nullptr, *gred = nullptr, *pinv = nullptr, *ppad = nullptr,
           *work = nullptr, *lam = nullptr;
    // segmented tile-offset scan: thread sums, segment sums, segment offsets
    // (+ one slot for the segment scan's total)
    double *hsum = nullptr, *segsum = nullptr, *segoff = nullptr;
    uint64_t nsegmax = 0;
    // hybrid sampler (larsen_kolda_hybrid.hpp): alpha_k, per-draw cuts,
    // candidate lists, enumeration block counts / offsets / p sums, the
    // deterministic count and mass, the rejection failure count
    bool hybrid = false;
    static constexpr uint64_t kHybBlocksMax = 1ull << 20;  // candidates <= 2^32
    double *alpha = nullptr, *cuts = nullptr, *bmax = nullptr, *hpp = nullptr, *hpoff = nullptr, *pdet = nullptr;
    uint64_t *cand[kMaxModes] = {}, *ncand = nullptr, *hoff = nullptr, *ndet = nullptr;
    uint32_t *hcnt = nullptr, *hfail = nullptr;
    PrimTemp sel;
    // last draw: deterministic samples, their mass, candidates, 1 if the
    // enumeration was skipped, samples whose rejection attempts ran out
    double info[5] = {};
    uint32_t* counters = nullptr;
    uint64_t seed = 0x243F6A8885A308D3ull, call = 0;
    // staging of the host-pointer API
    uint64_t* samples_dev = nullptr;
    double *h_dev = nullptr, *w_dev = nullptr, *draws_dev = nullptr;
    Arena mem;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // update_sampler runs on `stream`, draws on the caller's stream: upd_ev
    // (after an update) orders the next draw, draw_ev (after a draw on
    // another stream) the next update, so neither reads scores the other is
    // rewriting
    hipEvent_t upd_ev = nullptr, draw_ev = nullptr;
    bool draw_pending = false;

    ~LKSampler() {
        if (draw_pending && draw_ev) (void)hipEventSynchronize(draw_ev);
        if (stream) (void)hipStreamSynchronize(stream);
        if (upd_ev) (void)hipEventDestroy(upd_ev);
        if (draw_ev) (void)hipEventDestroy(draw_ev);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }

    // U: device factors aliased (alias) or host factors copied; st: the
    // caller's stream (borrowed) or null for one of the sampler's own
    void init(uint32_t N_, const uint64_t* I_, uint64_t J_, uint64_t R_, double* const* U_, bool alias,
              hipStream_t st, bool hybrid_ = false) {
        FTL_REQUIRE(N_ >= 1 && N_ <= (uint32_t)kMaxModes, "N must be in [1, 16]");
        hybrid = hybrid_;
        FTL_REQUIRE(R_ >= 1 && R_ <= 128, "R must be in [1, 128]");
        FTL_REQUIRE(J_ >= 1 && J_ < (1ull << 40), "J must be in [1, 2^40)");
        FTL_REQUIRE(I_ && U_, "null factor list");
        N = N_;
        J = J_;
        R = R_;
        if (st) {
            stream = st;
        } else {
            FTL_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamDefault));
            own_stream = true;
        }
        uint64_t imax = 1;
        for (uint32_t k = 0; k < N; k++) {
            FTL_REQUIRE(I_[k] >= 1 && I_[k] < (1ull << 40), "factor row counts must be in [1, 2^40)");
            FTL_REQUIRE(U_[k], "null factor matrix");
            I[k] = I_[k];
            imax = std::max(imax, I[k]);
            if (alias) {
                U[k] = U_[k];
            } else {
                U[k] = mem.alloc<double>(I[k] * R);
                FTL_HIP_CHECK(hipMemcpyAsync(U[k], U_[k], I[k] * R * sizeof(double), hipMemcpyHostToDevice, stream));
            }
            lev[k] = mem.alloc<double>(I[k]);
            cum[k] = mem.alloc<double>(I[k]);
            off[k] = mem.alloc<double>((I[k] + kLevTile - 1) / kLevTile);
        }
        pinv_lds_attrs();
        total = mem.alloc<double>(N);
        const uint64_t ntmax = (imax + kLevTile - 1) / kLevTile;
        tsum = mem.alloc<double>(ntmax);
        nsegmax = (ntmax + kOffSeg - 1) / kOffSeg;
        hsum = mem.alloc<double>(nsegmax * 256);
        segsum = mem.alloc<double>(nsegmax);
        segoff = mem.alloc<double>(nsegmax + 1);
        gred = mem.alloc<double>(8 * (R * (R + 1) / 2));
        ppad = mem.alloc<double>(64 * 64);
        if (hybrid) {
            FTL_REQUIRE(imax < (1ull << 31), "the hybrid sampler needs factor row counts below 2^31");
            alpha = mem.alloc<double>(N);
            cuts = mem.alloc<double>(N);
            bmax = mem.alloc<double>(kGramBlocksMax);
            ncand = mem.alloc<uint64_t>(N);
            for (uint32_t k = 0; k < N; k++) cand[k] = mem.alloc<uint64_t>(I[k]);
            hcnt = mem.alloc<uint32_t>(kHybBlocksMax);
            hoff = mem.alloc<uint64_t>(kHybBlocksMax);
            hpp = mem.alloc<double>(kHybBlocksMax);
            hpoff = mem.alloc<double>(kHybBlocksMax);
            pdet = mem.alloc<double>(1);
            ndet = mem.alloc<uint64_t>(1);
            hfail = mem.alloc<uint32_t>(1);
            size_t bytes = 0;
            FTL_HIP_CHECK(rocprim::select(nullptr, bytes, rocprim::counting_iterator<uint64_t>(0), cand[0], ncand,
                                          (size_t)imax, LevAtLeast{lev[0], cuts}, stream));
            sel.ensure(bytes);
        }
        gparts = mem.alloc<double>((uint64_t)kGramBlocksMax * (R * (R + 1) / 2));
        pinv = mem.alloc<double>(R * R);
        work = mem.alloc<double>(3 * (R + 1) * R);
        lam = mem.alloc<double>(R);
        counters = mem.alloc<uint32_t>(kCntSlots);
        for (uint32_t k = 0; k < N; k++) update(k);
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
    }

    // update_sampler(k) (larsen_kolda_sampler.hpp:45-100): Gram, pinv
    // (target rank min(I_k, R)), leverage scores with their tile sums, tile
    // offsets and total, prefix sums
    void update(uint32_t k) {
        FTL_REQUIRE(k < N, "update_sampler: mode index out of range");
        if (draw_pending) {
            FTL_HIP_CHECK(hipStreamWaitEvent(stream, draw_ev, 0));
            draw_pending = false;
        }
        const uint64_t n = I[k];
        const uint32_t r = (uint32_t)R;
        const unsigned nb = (unsigned)std::min<uint64_t>(kGramBlocksMax, std::max<uint64_t>(1, (n + 4095) / 4096));
        if (r <= 16)
            hipLaunchKernelGGL(k_gram_mfma<1>, dim3(nb), dim3(256), 0, stream, U[k], n, r, gparts);
        else if (r <= 32)
            hipLaunchKernelGGL(k_gram_mfma<2>, dim3(nb), dim3(256), 0, stream, U[k], n, r, gparts);
        else if (r <= 48)
            hipLaunchKernelGGL(k_gram_mfma<3>, dim3(nb), dim3(256), 0, stream, U[k], n, r, gparts);
        else if (r <= 64)
            hipLaunchKernelGGL(k_gram_mfma<4>, dim3(nb), dim3(256), 0, stream, U[k], n, r, gparts);
        else
            gram_rows(U[k], n, r, gparts, nb, stream);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, stream, counters + kCntUnique,
                           (uint32_t)std::min<uint64_t>(n, R));
        FTL_LAUNCH_CHECK();
        // the nb block partials in 8 ordered chunks, then the pinv
        const uint32_t T = r * (r + 1) / 2;
        hipLaunchKernelGGL(k_partials_to8, dim3((8 * T + 255) / 256), dim3(256), 0, stream, gparts, nb, T, gred);
        FTL_LAUNCH_CHECK();
        const uint64_t jac = 3 * (R + 1) * R <= (uint64_t)kJacobiLdsDoubles ? 3 * (R + 1) * R : (R + 1) * R;
        hipLaunchKernelGGL(k_pinv_partials<1024>, dim3(1), dim3(1024), jac * sizeof(double), stream, gred, 8u,
                           counters, r, pinv, work, lam);
        FTL_LAUNCH_CHECK();
        const uint64_t nt = (n + kLevTile - 1) / kLevTile;
        FTL_REQUIRE((nt + 3) / 4 <= 0x7fffffffull, "factor too tall for one launch");
        // tiles of 64 rows, four per block: k_lev_rows takes one per wave
        // (grid over all tiles); the scans stride over them
        const unsigned gl = (unsigned)((nt + 3) / 4);
        const unsigned g = grid_for(nt, 4, 2048);
        if (r <= 64) {
            const uint32_t rm = r <= 8 ? 8 : r <= 16 ? 16 : r <= 32 ? 32 : 64;
            hipLaunchKernelGGL(k_pad_square, dim3(1), dim3(256), 0, stream, pinv, r, rm, ppad);
            FTL_LAUNCH_CHECK();
            if (rm == 8)
                hipLaunchKernelGGL(k_lev_rows<8>, dim3(gl), dim3(256), 0, stream, U[k], n, r, ppad, lev[k], tsum);
            else if (rm == 16)
                hipLaunchKernelGGL(k_lev_rows<16>, dim3(gl), dim3(256), 0, stream, U[k], n, r, ppad, lev[k], tsum);
            else if (rm == 32)
                hipLaunchKernelGGL(k_lev_rows<32>, dim3(gl), dim3(256), 0, stream, U[k], n, r, ppad, lev[k], tsum);
            else
                hipLaunchKernelGGL(k_lev_rows<64>, dim3(gl), dim3(256), 0, stream, U[k], n, r, ppad, lev[k], tsum);
            FTL_LAUNCH_CHECK();
        } else {
            hipLaunchKernelGGL((k_row_leverage<64, 2>), dim3(grid_for(n, 4, 16384)), dim3(256), 0, stream, U[k], n,
                               r, pinv, lev[k]);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_tile_sums, dim3(g), dim3(256), 0, stream, lev[k], n, tsum);
            FTL_LAUNCH_CHECK();
        }
        const uint64_t nseg = (nt + kOffSeg - 1) / kOffSeg;
        hipLaunchKernelGGL(k_seg_tile_sums, dim3((unsigned)nseg), dim3(256), 0, stream, tsum, nt, hsum, segsum);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(1024), 0, stream, segsum, nseg, segoff, segoff + nseg);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_seg_tile_offsets, dim3((unsigned)nseg), dim3(256), 0, stream, tsum, nt, hsum, segoff,
                           off[k], total + k);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_cum_rows, dim3(g), dim3(256), 0, stream, lev[k], n, off[k], cum[k]);
        FTL_LAUNCH_CHECK();
        if (!upd_ev) FTL_HIP_CHECK(hipEventCreateWithFlags(&upd_ev, hipEventDisableTiming));
        if (hybrid) {
            // alpha_k = max lev / total (larsen_kolda_hybrid.hpp:101-116)
            const unsigned nbm = (unsigned)std::min<uint64_t>(kGramBlocksMax, std::max<uint64_t>(1, n / 16384));
            hipLaunchKernelGGL(k_block_max, dim3(nbm), dim3(256), 0, stream, lev[k], n, bmax);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_alpha, dim3(1), dim3(64), 0, stream, bmax, nbm, total, k, alpha);
            FTL_LAUNCH_CHECK();
        }
        FTL_HIP_CHECK(hipEventRecord(upd_ev, stream));
    }

    // KRPDrawSamples(j) (larsen_kolda_sampler.hpp:102-133): samples (N x J,
    // rows k != j written), h (J x R) and weights (J), each optional but
    // samples; draws: uniform of (sample i, mode k) at draws[k J + i], or null
    // for the built-in Philox stream
    void draw_device(uint32_t j, uint64_t* samples, double* h, double* w, const double* draws, hipStream_t st) {
        FTL_REQUIRE(samples != nullptr, "null samples buffer");
        FTL_REQUIRE(N > 1 || j >= N, "no mode left to sample (N == 1 and j == 0)");
        if (!st) st = stream;
        if (st != stream && upd_ev) FTL_HIP_CHECK(hipStreamWaitEvent(st, upd_ev, 0));
        LKArgs a{};
        for (uint32_t k = 0; k < N; k++) {
            a.cum[k] = cum[k];
            a.off[k] = off[k];
            a.lev[k] = lev[k];
            a.I[k] = I[k];
        }
        a.total = total;
        a.N = (int)N;
        a.j = j < N ? (int)j : -1;
        a.J = J;
        a.draws = draws;
        a.seed = seed;
        a.call = call;
        a.samples = samples;
        a.w = w;
        if (hybrid)
            draw_hybrid(a, st);
        else
            hipLaunchKernelGGL(k_lk_draw, dim3(grid_for(J, 256, 16384)), dim3(256), 0, st, a);
        FTL_LAUNCH_CHECK();
        if (h) {
            FactorPtrs F{};
            for (uint32_t k = 0; k < N; k++) F.U[k] = U[k];
            hipLaunchKernelGGL(k_fill_h, dim3(grid_for(J * R, 256, 16384)), dim3(256), 0, st, samples, J,
                               (uint32_t)R, (int)N, a.j, F, h);
            FTL_LAUNCH_CHECK();
        }
        if (!draws) call++;
        if (st != stream) {
            if (!draw_ev) FTL_HIP_CHECK(hipEventCreateWithFlags(&draw_ev, hipEventDisableTiming));
            FTL_HIP_CHECK(hipEventRecord(draw_ev, st));
            draw_pending = true;
        }
        if (hybrid) {
            // the deterministic count, its mass and the rejection failures
            // are checked once the draw is done
            uint64_t nd = 0;
            double pd = 0.0;
            uint32_t nf = 0;
            FTL_HIP_CHECK(hipMemcpyAsync(&nd, ndet, sizeof(nd), hipMemcpyDeviceToHost, st));
            FTL_HIP_CHECK(hipMemcpyAsync(&pd, pdet, sizeof(pd), hipMemcpyDeviceToHost, st));
            FTL_HIP_CHECK(hipMemcpyAsync(&nf, hfail, sizeof(nf), hipMemcpyDeviceToHost, st));
            FTL_HIP_CHECK(hipStreamSynchronize(st));
            info[0] = (double)nd;
            info[1] = pd;
            info[4] = (double)nf;
            FTL_REQUIRE(nd <= J, "hybrid sampler: more than J tuples reach probability 1/J");
        }
    }

    // KRPDrawSamples of the hybrid sampler (larsen_kolda_hybrid.hpp:139-246):
    // cuts, candidate lists, the candidate count on the host (it sizes the
    // enumeration), the kept tuples in candidate order, then rejection
    // sampling of the rest
    void draw_hybrid(const LKArgs& a, hipStream_t st) {
        const int j = a.j;
        hipLaunchKernelGGL(k_hyb_cuts, dim3(1), dim3(64), 0, st, alpha, total, (int)N, j, J, cuts);
        FTL_LAUNCH_CHECK();
        for (uint32_t k = 0; k < N; k++) {
            if ((int)k == j) continue;
            size_t bytes = sel.bytes;
            FTL_HIP_CHECK(rocprim::select(sel.p, bytes, rocprim::counting_iterator<uint64_t>(0), cand[k], ncand + k,
                                          (size_t)I[k], LevAtLeast{lev[k], cuts + k}, st));
        }
        uint64_t nc[kMaxModes] = {};
        FTL_HIP_CHECK(hipMemcpyAsync(nc, ncand, N * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        FTL_HIP_CHECK(hipStreamSynchronize(st));
        double logc = 0.0;
        uint64_t C = 1;
        bool zero = false;
        for (uint32_t k = 0; k < N; k++) {
            if ((int)k == j) continue;
            if (nc[k] == 0) zero = true;
            logc += std::log((double)nc[k]);
        }
        // the reference skips the deterministic part (with a warning) when
        // the candidate count reaches 2^62 (larsen_kolda_hybrid.hpp:171-173)
        const bool skip = !zero && logc >= 62.0 * std::log(2.0);
        if (zero) C = 0;
        if (!zero && !skip) {
            FTL_REQUIRE(logc < 32.0 * std::log(2.0) + 1e-9, "hybrid sampler: more than 2^32 candidate tuples");
            for (uint32_t k = 0; k < N; k++)
                if ((int)k != j) C *= nc[k];
            FTL_REQUIRE(C <= (1ull << 32), "hybrid sampler: more than 2^32 candidate tuples");
        }
        info[2] = zero ? 0.0 : std::exp(logc);
        info[3] = skip ? 1.0 : 0.0;
        HybArgs ha{};
        for (uint32_t k = 0; k < N; k++) {
            ha.cand[k] = cand[k];
            ha.lev[k] = lev[k];
        }
        ha.ncand = ncand;
        ha.total = total;
        ha.N = (int)N;
        ha.j = j;
        ha.J = J;
        ha.C = skip ? 0 : C;
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, hfail, 0u);
        FTL_LAUNCH_CHECK();
        if (ha.C > 0) {
            const unsigned nb = (unsigned)((ha.C + kHybChunk - 1) / kHybChunk);
            hipLaunchKernelGGL(k_hyb_count, dim3(nb), dim3(256), 0, st, ha, hcnt, hpp);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_hyb_offsets, dim3(1), dim3(64), 0, st, hcnt, nb, hoff, ndet);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(1024), 0, st, hpp, (uint64_t)nb, hpoff, pdet);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_hyb_write, dim3(nb), dim3(256), 0, st, ha, hoff, a.samples, a.w);
            FTL_LAUNCH_CHECK();
        } else {
            hipLaunchKernelGGL(k_hyb_zero, dim3(1), dim3(64), 0, st, ndet, pdet);
            FTL_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_hyb_reject, dim3(grid_for(J, 256, 16384)), dim3(256), 0, st, a, ndet, pdet, hfail);
    }

    void draw_host(uint32_t j, uint64_t* samples, double* h, double* w, const double* draws) {
        FTL_REQUIRE(samples != nullptr, "null samples array");
        if (!samples_dev) {
            samples_dev = mem.alloc<uint64_t>(N * J);
            h_dev = mem.alloc<double>(J * R);
            w_dev = mem.alloc<double>(J);
            draws_dev = mem.alloc<double>(2 * N * J);
        }
        if (draws)
            FTL_HIP_CHECK(hipMemcpyAsync(draws_dev, draws, 2 * N * J * sizeof(double), hipMemcpyHostToDevice,
                                         stream));
        draw_device(j, samples_dev, h ? h_dev : nullptr, w ? w_dev : nullptr, draws ? draws_dev : nullptr, stream);
        for (uint32_t k = 0; k < N; k++)
            if (k != j)
                FTL_HIP_CHECK(hipMemcpyAsync(samples + (uint64_t)k * J, samples_dev + (uint64_t)k * J,
                                             J * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
        if (h) FTL_HIP_CHECK(hipMemcpyAsync(h, h_dev, J * R * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (w) FTL_HIP_CHECK(hipMemcpyAsync(w, w_dev, J * sizeof(double), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
    }
};

// ---------------------------------------------------------------------------
// LowRankTensor + ALS (als.hpp:31-220, low_rank_tensor.hpp)
// ---------------------------------------------------------------------------
struct Als {
    SparseDev* X = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t N = 0;
    uint64_t R = 0;
    uint64_t I[kMaxModes] = {};
    uint32_t ibits[kMaxModes] = {};
    double* U[kMaxModes] = {};
    double *sigma = nullptr, *norms = nullptr, *colpart = nullptr, *scal = nullptr, *resid_part = nullptr,
           *gram_parts = nullptr, *gram_red = nullptr;
    Arena mem;
    PrimTemp prim;
    Matches mt;
    // sampled update state (initialize_ds_als): the exact sampler, or the
    // Larsen-Kolda sampler
    Sampler* sampler = nullptr;
    std::unique_ptr<LKSampler> lk;
    uint64_t J = 0;
    KeyLayout dlayout[kMaxModes] = {};
    uint64_t* samples = nullptr;
    double *h = nullptr, *w = nullptr, *draws_stage = nullptr;
    uint64_t *keys_a = nullptr, *keys_b = nullptr;
    uint32_t *vals_a = nullptr, *vals_b = nullptr, *flags = nullptr, *uid = nullptr;
    uint32_t *utup = nullptr, *ufirst = nullptr, *ustart = nullptr;
    double *hs = nullptr, *hdd = nullptr, *gpart = nullptr, *pinv = nullptr, *work = nullptr, *lam = nullptr;
    // exact update state
    double *gram_had = nullptr, *exact_partial = nullptr;
    uint64_t exact_partial_cap = 0;
    uint32_t* full_rank = nullptr;  // counters with kCntUnique = R (pinv at full target rank)
    // the zero fill of U_j in a sampled update runs on a second stream after
    // the draw (bandwidth-bound, it would be slowed), beside the dedup, probe
    // and match sort (none of which reads U_j)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_zero = nullptr;
    // mode excluded by the last sampled update (-1: none since the sampler
    // was set up)
    int last_sampled_j = -1;
    // diagnostics
    bool timing = false;
    std::array<hipEvent_t, 5> ev{};
    bool have_events = false;
    uint64_t last_unique = 0, last_found = 0;

    ~Als() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        if (sampler) sampler_delete(sampler);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_zero) (void)hipEventDestroy(ev_zero);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void init(SparseDev* X_, uint32_t N_, const uint64_t* I_, uint64_t R_, const double* const* U_host) {
        FTL_REQUIRE(N_ >= 1 && N_ <= (uint32_t)kMaxModes, "N must be in [1, 16]");
        FTL_REQUIRE(R_ >= 1 && R_ <= 128, "R must be in [1, 128]");
        FTL_REQUIRE(I_ && U_host, "null factor list");
        X = X_;
        N = N_;
        R = R_;
        if (X) FTL_REQUIRE(X->N == N, "the factor count differs from the sparse tensor's mode count");
        FTL_HIP_CHECK(hipGetDevice(&device));
        // blocking stream: ordered after the caller's default-stream work
        FTL_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamDefault));
        FTL_HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        FTL_HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        FTL_HIP_CHECK(hipEventCreateWithFlags(&ev_zero, hipEventDisableTiming));
        pinv_lds_attrs();
        mt.init(mem);
        for (uint32_t k = 0; k < N; k++) {
            FTL_REQUIRE(I_[k] >= 1 && I_[k] < (1ull << 32), "factor row counts must be in [1, 2^32)");
            FTL_REQUIRE(U_host[k], "null factor matrix");
            if (X)
                FTL_REQUIRE(I_[k] >= X->dims[k], "factor " + std::to_string(k) +
                                                     " has fewer rows than the sparse tensor's mode size");
            I[k] = I_[k];
            ibits[k] = bit_width(I[k] - 1);
            U[k] = mem.alloc<double>(I[k] * R);
            FTL_HIP_CHECK(hipMemcpyAsync(U[k], U_host[k], I[k] * R * sizeof(double), hipMemcpyHostToDevice, stream));
        }
        const uint64_t T = R * (R + 1) / 2;
        sigma = mem.alloc<double>(R);
        norms = mem.alloc<double>(R);
        colpart = mem.alloc<double>((uint64_t)std::max(kColBlocks, kSegBlocksMax) * R);
        scal = mem.alloc<double>(4);
        resid_part = mem.alloc<double>(kResidBlocks);
        gram_parts = mem.alloc<double>((uint64_t)N * kGramBlocksMax * T);
        gram_red = mem.alloc<double>((uint64_t)N * T);
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(128), 0, stream, sigma, R, 1.0);
        FTL_LAUNCH_CHECK();
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
    }

    // kind: FTL_SAMPLER_EFFICIENT (precision applies) or FTL_SAMPLER_LARSEN_KOLDA
    void initialize_ds(uint64_t J_, int precision, int kind = 0) {
        FTL_REQUIRE(X, "the sampled update needs a sparse tensor");
        FTL_REQUIRE(N >= 2, "the sampled update needs at least two modes");
        FTL_REQUIRE(J_ >= 1 && J_ < (1ull << 31), "J must be in [1, 2^31)");
        FTL_REQUIRE(!sampler && !lk, "initialize_ds_als was already called");
        FTL_REQUIRE(kind == 0 || kind == 1 || kind == 2, "unknown sampler kind");
        for (uint32_t j = 0; j < N; j++) {
            dlayout[j] = make_layout(N, j, ibits);
            if (dlayout[j].total_bits > 64)
                throw Error{FTL_EUNSUPPORTED, "sample coordinates of the modes other than " + std::to_string(j) +
                                                  " need more than 64 key bits"};
        }
        J = J_;
        if (kind == 1 || kind == 2) {
            lk.reset(new LKSampler());
            lk->init(N, I, J, R, U, true, stream, kind == 2);
        } else {
            sampler = sampler_new_aliased(N, I, J, R, U, precision);
        }
        samples = mem.alloc<uint64_t>(N * J);
        FTL_HIP_CHECK(hipMemsetAsync(samples, 0, N * J * sizeof(uint64_t), stream));
        h = mem.alloc<double>(J * R);
        w = mem.alloc<double>(J);
        keys_a = mem.alloc<uint64_t>(J);
        keys_b = mem.alloc<uint64_t>(J);
        vals_a = mem.alloc<uint32_t>(J);
        vals_b = mem.alloc<uint32_t>(J);
        flags = mem.alloc<uint32_t>(J);
        uid = mem.alloc<uint32_t>(J);
        utup = mem.alloc<uint32_t>(J * N);
        ufirst = mem.alloc<uint32_t>(J);
        ustart = mem.alloc<uint32_t>(J + 1);
        hs = mem.alloc<double>(J * R);
        hdd = mem.alloc<double>(J * (R + (R & 1)));  // rows of even stride (a zero pad column for odd R)
        gpart = mem.alloc<double>((uint64_t)kGramBlocks * (R * (R + 1) / 2));
        pinv = mem.alloc<double>(R * R);
        work = mem.alloc<double>(3 * (R + 1) * R);
        lam = mem.alloc<double>(R);
        mt.ensure_n(mem, J);
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
    }

    void record(int i) {
        if (!timing) return;
        if (!ev[i]) FTL_HIP_CHECK(hipEventCreate(&ev[i]));
        FTL_HIP_CHECK(hipEventRecord(ev[i], stream));
    }

    template <int RMAX>
    void launch_gram(const double* Xrows, const uint32_t* n_dev, uint64_t n_host, double* parts, unsigned nb) {
        hipLaunchKernelGGL((k_gram_partial<256, RMAX>), dim3(nb), dim3(256), 0, stream, Xrows, n_dev, n_host,
                           (uint32_t)R, parts);
        FTL_LAUNCH_CHECK();
    }

    // block partials (nb blocks) of the packed Gram of n rows
    void gram_partials(const double* Xrows, const uint32_t* n_dev, uint64_t n_host, double* parts, unsigned nb) {
        if (R <= 8) launch_gram<8>(Xrows, n_dev, n_host, parts, nb);
        else if (R <= 16) launch_gram<16>(Xrows, n_dev, n_host, parts, nb);
        else if (R <= 32) launch_gram<32>(Xrows, n_dev, n_host, parts, nb);
        else if (R <= 64) launch_gram<64>(Xrows, n_dev, n_host, parts, nb);
        else launch_gram<128>(Xrows, n_dev, n_host, parts, nb);
    }

    // The last sampled update's draw: samples (N x J, the excluded mode's row
    // zero), h (J x R) and weights (J), as the sampler wrote them.
    void last_draw(uint64_t* samples_host, double* h_host, double* w_host) {
        FTL_REQUIRE(sampler || lk, "initialize_ds_als has not been called");
        FTL_REQUIRE(last_sampled_j >= 0, "no sampled update has run");
        FTL_HIP_CHECK(hipMemcpyAsync(samples_host, samples, N * J * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
        if (h_host) FTL_HIP_CHECK(hipMemcpyAsync(h_host, h, J * R * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (w_host) FTL_HIP_CHECK(hipMemcpyAsync(w_host, w, J * sizeof(double), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        // the draw writes the sampled modes' rows only: the excluded mode's
        // row of the device buffer still holds an earlier update's samples
        std::fill(samples_host + (uint64_t)last_sampled_j * J, samples_host + (uint64_t)(last_sampled_j + 1) * J,
                  (uint64_t)0);
    }

    // execute_ds_als_update(j, renormalize, update_sampler) (als.hpp:57-220)
    void ds_update(uint32_t j, bool renormalize, bool update_sampler, const double* draws_dev) {
        FTL_REQUIRE(sampler || lk, "initialize_ds_als has not been called");
        FTL_REQUIRE(j < N, "mode index out of range");
        last_sampled_j = (int)j;
        have_events = timing;
        hipStream_t st = stream;
        record(0);
        // 1. draw J samples of the KRP of the modes != j
        if (lk) lk->draw_device(j, samples, h, w, draws_dev, st);
        else sampler_draw(sampler, j, samples, h, w, draws_dev, st);
        record(1);
        // U_j is zeroed on the side stream while steps 2-3 and the match
        // expansion and sort run (latency-bound; the draw before them is
        // bandwidth-bound and would be slowed)
        FTL_HIP_CHECK(hipEventRecord(ev_fork, st));
        FTL_HIP_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
        FTL_HIP_CHECK(hipMemsetAsync(U[j], 0, I[j] * R * sizeof(double), side));
        FTL_HIP_CHECK(hipEventRecord(ev_zero, side));
        // 2. sort / unique with counts, scaled rows, sampled Gram pinv
        const KeyLayout& DL = dlayout[j];
        const unsigned gJ = grid_for(J, 256);
        hipLaunchKernelGGL(k_sample_keys, dim3(gJ), dim3(256), 0, st, samples, J, DL, keys_a, vals_a);
        FTL_LAUNCH_CHECK();
        sort_pairs(prim, keys_a, keys_b, vals_a, vals_b, J, DL.total_bits, st);
        hipLaunchKernelGGL(k_head_flags<uint64_t>, dim3(gJ), dim3(256), 0, st, keys_b, J, flags);
        FTL_LAUNCH_CHECK();
        inclusive_sum(prim, flags, uid, J, st);
        hipLaunchKernelGGL(k_dedup_heads, dim3(gJ), dim3(256), 0, st, keys_b, vals_b, uid, J, DL, utup, ufirst,
                           ustart, mt.counters);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_dedup_rows, dim3(grid_for(J * R, 256)), dim3(256), 0, st, ufirst, ustart, mt.counters,
                           J, (uint32_t)R, h, w, hs, hdd);
        FTL_LAUNCH_CHECK();
        gram_partials(hs, mt.counters + kCntUnique, 0, gpart, kGramBlocks);
        const uint64_t jac = 3 * (R + 1) * R <= (uint64_t)kJacobiLdsDoubles ? 3 * (R + 1) * R : (R + 1) * R;
        if (jacobi_threads_for(R) == 256)
            hipLaunchKernelGGL(k_pinv_partials<256>, dim3(1), dim3(256), jac * sizeof(double), st, gpart,
                               kGramBlocks, mt.counters, (uint32_t)R, pinv, work, lam);
        else
            hipLaunchKernelGGL(k_pinv_partials<1024>, dim3(1), dim3(1024), jac * sizeof(double), st, gpart,
                               kGramBlocks, mt.counters, (uint32_t)R, pinv, work, lam);
        FTL_LAUNCH_CHECK();
        // 3. probe the unique tuples on mode j's lookup; the match count
        //    sizes the rest, so it is read back here (one wait per update)
        probe_and_count(*X, j, utup, mt.counters + kCntUnique, 0, J, mt, mem, prim, st);
        FTL_HIP_CHECK(hipMemcpyAsync(mt.host_counters, mt.counters, kCntSlots * sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, st));
        FTL_HIP_CHECK(hipStreamSynchronize(st));
        const uint64_t Jp = mt.host_counters[kCntUnique], T = mt.host_counters[kCntMatches];
        last_unique = Jp;
        last_found = T;
        record(2);
        // 4. U_j = MTTKRP(h'') pinv on the touched rows, zero elsewhere (the
        //    zero fill is joined in reduce_matches, before the row finalise)
        unsigned colparts = 0;
        if (T > 0)
            colparts = reduce_matches(*X, j, Jp, T, hdd, (uint32_t)R, (uint32_t)(R + (R & 1)), pinv, U[j],
                                      renormalize ? colpart : nullptr, ev_zero, mt,
                                      mem, prim, st);
        else {
            FTL_HIP_CHECK(hipMemsetAsync(mt.counters + kCntRows, 0, 2 * sizeof(uint32_t), st));
            FTL_HIP_CHECK(hipStreamWaitEvent(st, ev_zero, 0));
        }
        record(3);
        // 5. renormalize (only touched rows are nonzero) and rebuild tree j
        if (renormalize) {
            if (T > 0) {
                // column sums of squares of the touched rows came with the row finalise
                hipLaunchKernelGGL(k_colnorm_final, dim3(1), dim3(256), 0, st, colpart, colparts, (uint32_t)R, norms,
                                   sigma, 0);
                FTL_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(T * R, 256)), dim3(256), 0, st, U[j], mt.seg_row,
                                   mt.counters + kCntRows, (uint64_t)0, (uint32_t)R, norms);
                FTL_LAUNCH_CHECK();
            } else {
                hipLaunchKernelGGL(k_fill, dim3(1), dim3(128), 0, st, sigma, R, 1.0);
                FTL_LAUNCH_CHECK();
            }
        }
        if (update_sampler) {
            if (lk) lk->update(j);
            else sampler_rebuild(sampler, j, st);
        }
        record(4);
    }

    // execute_exact_als_update(j, renormalize, update_sampler) (als.hpp:222-270):
    // pinv of (*)_{k != j} U_k^T U_k at full target rank (compute_pinv_square,
    // common.h:228-269), U_j = exact MTTKRP * pinv on the rows that hold
    // nonzeros (zero elsewhere), then renormalize_columns(j).  update_sampler
    // is accepted and unused, as in the reference.
    template <int G, int CPL>
    unsigned launch_exact(const SparseDev::RowOrder& o, uint32_t j, double* colsq) {
        constexpr int GPB = 256 / G;
        FactorPtrs F{};
        for (uint32_t k = 0; k < N; k++) F.U[k] = U[k];
        hipLaunchKernelGGL((k_exact_item_partial<G, CPL>), dim3(grid_for(o.W, GPB, 16384)), dim3(256), 0, stream,
                           o.seg_start, o.piece_off, o.piece_seg, o.counters, o.pos, X->keys[j], X->val[j],
                           X->layout[j], F, (uint32_t)R, exact_partial);
        FTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_seg_heavy_sum, dim3(512), dim3(256), 0, stream, o.piece_off, o.counters, o.heavy,
                           (uint32_t)R, exact_partial);
        FTL_LAUNCH_CHECK();
        const unsigned fblocks = grid_for(o.D, GPB, kSegBlocksMax);
        hipLaunchKernelGGL((k_seg_final<G, CPL, true>), dim3(fblocks), dim3(256), 0, stream, o.seg_row, o.piece_off,
                           o.counters, exact_partial, (uint32_t)R, pinv, U[j], colsq);
        FTL_LAUNCH_CHECK();
        return fblocks;
    }

    void exact_update(uint32_t j, bool renormalize, bool /*update_sampler*/) {
        FTL_REQUIRE(X, "the exact update needs a sparse tensor");
        FTL_REQUIRE(j < N, "mode index out of range");
        const SparseDev::RowOrder& o = X->row_order(j);
        // X's stream built the row order; this handle's stream reads it
        FTL_HIP_CHECK(hipStreamSynchronize(X->stream));
        const uint64_t T = R * (R + 1) / 2;
        if (!pinv) {
            pinv = mem.alloc<double>(R * R);
            work = mem.alloc<double>(3 * (R + 1) * R);
            lam = mem.alloc<double>(R);
        }
        if (!gram_had) {
            gram_had = mem.alloc<double>(T);
            full_rank = mem.alloc<uint32_t>(kCntSlots);
            const uint32_t fr[kCntSlots] = {(uint32_t)R, 0, 0, 0};
            FTL_HIP_CHECK(hipMemcpyAsync(full_rank, fr, sizeof(fr), hipMemcpyHostToDevice, stream));
            FTL_HIP_CHECK(hipStreamSynchronize(stream));
        }
        if (o.W * R > exact_partial_cap) {
            mem.release(exact_partial);
            exact_partial = mem.alloc<double>(o.W * R);
            exact_partial_cap = o.W * R;
        }
        hipStream_t st = stream;
        factor_grams();
        hipLaunchKernelGGL(k_gram_hadamard, dim3(grid_for(T, 256)), dim3(256), 0, st, gram_red, N, (uint32_t)T, j,
                           gram_had);
        FTL_LAUNCH_CHECK();
        const uint64_t jac = 3 * (R + 1) * R <= (uint64_t)kJacobiLdsDoubles ? 3 * (R + 1) * R : (R + 1) * R;
        if (jacobi_threads_for(R) == 256)
            hipLaunchKernelGGL(k_pinv_partials<256>, dim3(1), dim3(256), jac * sizeof(double), st, gram_had, 1u,
                               full_rank, (uint32_t)R, pinv, work, lam);
        else
            hipLaunchKernelGGL(k_pinv_partials<1024>, dim3(1), dim3(1024), jac * sizeof(double), st, gram_had, 1u,
                               full_rank, (uint32_t)R, pinv, work, lam);
        FTL_LAUNCH_CHECK();
        FTL_HIP_CHECK(hipMemsetAsync(U[j], 0, I[j] * R * sizeof(double), st));
        double* colsq = renormalize ? colpart : nullptr;
        unsigned parts;
        if (R <= 32) parts = launch_exact<32, 1>(o, j, colsq);
        else if (R <= 64) parts = launch_exact<64, 1>(o, j, colsq);
        else parts = launch_exact<64, 2>(o, j, colsq);
        if (renormalize) {
            hipLaunchKernelGGL(k_colnorm_final, dim3(1), dim3(256), 0, st, colpart, parts, (uint32_t)R, norms, sigma, 0);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(o.D * R, 256)), dim3(256), 0, st, U[j], o.seg_row,
                               o.counters + kCntRows, (uint64_t)0, (uint32_t)R, norms);
            FTL_LAUNCH_CHECK();
        }
        last_unique = 0;
        last_found = X->nnz;
        have_events = false;
    }

    // renormalize_columns(j); j = -1: every mode, sigma = product of norms
    void renormalize(int j) {
        FTL_REQUIRE(j >= -1 && j < (int)N, "mode index out of range");
        hipStream_t st = stream;
        if (j == -1) {
            hipLaunchKernelGGL(k_fill, dim3(1), dim3(128), 0, st, sigma, R, 1.0);
            FTL_LAUNCH_CHECK();
        }
        for (uint32_t k = 0; k < N; k++) {
            if (j != -1 && (uint32_t)j != k) continue;
            hipLaunchKernelGGL(k_colsq_partial<256>, dim3(kColBlocks), dim3(256), 0, st, U[k], nullptr, nullptr, I[k],
                               (uint32_t)R, colpart);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_colnorm_final, dim3(1), dim3(256), 0, st, colpart, kColBlocks, (uint32_t)R, norms,
                               sigma, j == -1 ? 1 : 0);
            FTL_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(I[k] * R, 256)), dim3(256), 0, st, U[k], nullptr, nullptr,
                               I[k], (uint32_t)R, norms);
            FTL_LAUNCH_CHECK();
        }
        FTL_HIP_CHECK(hipStreamSynchronize(st));
    }

    // packed U_k^T U_k of every mode into gram_red (enqueued)
    void factor_grams() {
        const uint64_t T = R * (R + 1) / 2;
        uint64_t imax = 1;
        for (uint32_t k = 0; k < N; k++) imax = std::max(imax, I[k]);
        const unsigned nb = (unsigned)std::min<uint64_t>(kGramBlocksMax, std::max<uint64_t>(1, (imax + 4095) / 4096));
        for (uint32_t k = 0; k < N; k++) gram_partials(U[k], nullptr, I[k], gram_parts + (uint64_t)k * nb * T, nb);
        hipLaunchKernelGGL(k_gram_reduce, dim3(grid_for(N * T, 256)), dim3(256), 0, stream, gram_parts, nb,
                           (uint32_t)T, N, gram_red);
        FTL_LAUNCH_CHECK();
    }

    // ||[[sigma; U]]||^2 into scal[1] (enqueued)
    void lowrank_normsq_async() {
        factor_grams();
        hipLaunchKernelGGL(k_chain_final, dim3(1), dim3(256), 0, stream, gram_red, N, (uint32_t)R, sigma, scal + 1);
        FTL_LAUNCH_CHECK();
    }

    double lowrank_normsq() {
        lowrank_normsq_async();
        double out = 0;
        FTL_HIP_CHECK(hipMemcpyAsync(&out, scal + 1, sizeof(double), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        return out;
    }

    template <int G, int CPL>
    void launch_residual(const FactorPtrs& F, uint32_t s) {
        hipLaunchKernelGGL((k_residual_partial<G, CPL, 256>), dim3(kResidBlocks), dim3(256), 0, stream, X->keys[s],
                           X->row[s], X->val[s], X->nnz, X->layout[s], F, sigma, (uint32_t)R, resid_part);
        FTL_LAUNCH_CHECK();
    }

    // compute_residual_normsq(sigma, U) of the sparse tensor
    // (sort_lookup.hpp:134-197): sum over nonzeros + ||M||^2
    double residual_normsq() {
        FTL_REQUIRE(X, "no sparse tensor bound to this handle");
        FactorPtrs F{};
        for (uint32_t k = 0; k < N; k++) F.U[k] = U[k];
        // walk the lookup that leaves out the smallest mode: its factor rows
        // are the per-nonzero gathers
        uint32_t s = 0;
        for (uint32_t k = 1; k < N; k++)
            if (X->dims[k] < X->dims[s]) s = k;
        if (R <= 32) launch_residual<32, 1>(F, s);
        else if (R <= 64) launch_residual<64, 1>(F, s);
        else launch_residual<64, 2>(F, s);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, stream, resid_part, kResidBlocks, scal);
        FTL_LAUNCH_CHECK();
        lowrank_normsq_async();
        double out[2] = {0, 0};
        FTL_HIP_CHECK(hipMemcpyAsync(out, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        return out[0] + out[1];
    }

    void stats(double* out) {
        FTL_HIP_CHECK(hipStreamSynchronize(stream));
        FTL_HIP_CHECK(hipMemcpy(mt.host_counters, mt.counters, kCntSlots * sizeof(uint32_t), hipMemcpyDeviceToHost));
        out[0] = (double)last_unique;
        out[1] = (double)last_found;
        out[2] = last_found ? (double)mt.host_counters[kCntRows] : 0.0;
        out[3] = last_found ? (double)mt.host_counters[kCntPieces] : 0.0;
        for (int i = 0; i < 4; i++) {
            float t = 0.f;
            if (have_events && ev[i] && ev[i + 1]) FTL_HIP_CHECK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
            out[4 + i] = have_events ? (double)t : 0.0;
        }
    }
};

}  // namespace ftl

struct ftl_sparse {
    ftl::SparseDev impl;
};

struct ftl_als {
    ftl::Als impl;
};

using ftl::guarded;

namespace ftl {
uint64_t jacobi_capped_als_tu() { return jacobi_capped_read(); }

// ---------------------------------------------------------------------------
// Allocation-free residual: compute_residual_normsq on device factors with
// every intermediate in `scratch`, a file-scope device array of
// kNoAllocScratchDoubles doubles (9 MiB, held by each GPU the library runs
// on until the process ends).  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (scratch_base); calls
// that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kNoAllocScratchDoubles = 9ull << 17;  // 1,179,648 doubles = 9 MiB

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_base;  // per device, null until first looked up

double* scratch_base() {
    int dev = 0;
    FTL_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)dev >= g_scratch_base.size()) g_scratch_base.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_base[(size_t)dev]) {
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !p)
            throw Error{FTL_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc)};
        g_scratch_base[(size_t)dev] = (double*)p;
    }
    return g_scratch_base[(size_t)dev];
}

__global__ void k_add_pair(const double* __restrict__ two, double* __restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = two[0] + two[1];
}

template <int RMAX>
void noalloc_gram(const double* U, uint64_t n, uint32_t R, double* parts, unsigned nb, hipStream_t st) {
    hipLaunchKernelGGL((k_gram_partial<256, RMAX>), dim3(nb), dim3(256), 0, st, U, (const uint32_t*)nullptr, n, R,
                       parts);
    FTL_LAUNCH_CHECK();
}

}  // namespace

// The same kernels, in the same order and with the same grids, as
// Als::residual_normsq (so the result equals ftl_sparse_residual_normsq's
// bitwise), with the Gram partials of one mode at a time.
void residual_normsq_noalloc(SparseDev& X, uint64_t R, const uint64_t* I, const double* sigma,
                             const double* const* U, double* out, hipStream_t st) {
    FTL_REQUIRE(R >= 1 && R <= 128, "R must be in [1, 128]");
    FTL_REQUIRE(I && U && sigma && out, "null argument");
    const uint32_t N = X.N;
    const uint64_t T = R * (R + 1) / 2;
    uint64_t imax = 1;
    for (uint32_t k = 0; k < N; k++) {
        FTL_REQUIRE(U[k], "null factor");
        FTL_REQUIRE(I[k] >= X.dims[k], "factor has fewer rows than the tensor's mode size");
        imax = std::max(imax, I[k]);
    }
    const unsigned nb = (unsigned)std::min<uint64_t>(kGramBlocksMax, std::max<uint64_t>(1, (imax + 4095) / 4096));
    // layout: [scal 4][resid_part kResidBlocks][gram_red N T][gram_parts nb T]
    const uint64_t need = 4 + kResidBlocks + (uint64_t)N * T + (uint64_t)nb * T;
    if (need > kNoAllocScratchDoubles)
        throw Error{FTL_EUNSUPPORTED, "the residual needs " + std::to_string(need * 8) + " bytes of scratch for R = " +
                                          std::to_string(R) + " and " + std::to_string(imax) +
                                          " factor rows; the allocation-free entry point holds " +
                                          std::to_string(kNoAllocScratchDoubles * 8)};
    double* base = scratch_base();
    double* scal = base;
    double* resid_part = base + 4;
    double* gram_red = resid_part + kResidBlocks;
    double* gram_parts = gram_red + (uint64_t)N * T;
    FactorPtrs F{};
    for (uint32_t k = 0; k < N; k++) F.U[k] = U[k];
    uint32_t s = 0;
    for (uint32_t k = 1; k < N; k++)
        if (X.dims[k] < X.dims[s]) s = k;
    const uint32_t Ru = (uint32_t)R;
    if (R <= 32)
        hipLaunchKernelGGL((k_residual_partial<32, 1, 256>), dim3(kResidBlocks), dim3(256), 0, st, X.keys[s], X.row[s],
                           X.val[s], X.nnz, X.layout[s], F, sigma, Ru, resid_part);
    else if (R <= 64)
        hipLaunchKernelGGL((k_residual_partial<64, 1, 256>), dim3(kResidBlocks), dim3(256), 0, st, X.keys[s], X.row[s],
                           X.val[s], X.nnz, X.layout[s], F, sigma, Ru, resid_part);
    else
        hipLaunchKernelGGL((k_residual_partial<64, 2, 256>), dim3(kResidBlocks), dim3(256), 0, st, X.keys[s], X.row[s],
                           X.val[s], X.nnz, X.layout[s], F, sigma, Ru, resid_part);
    FTL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, resid_part, kResidBlocks, scal);
    FTL_LAUNCH_CHECK();
    for (uint32_t k = 0; k < N; k++) {
        if (R <= 8) noalloc_gram<8>(U[k], I[k], Ru, gram_parts, nb, st);
        else if (R <= 16) noalloc_gram<16>(U[k], I[k], Ru, gram_parts, nb, st);
        else if (R <= 32) noalloc_gram<32>(U[k], I[k], Ru, gram_parts, nb, st);
        else if (R <= 64) noalloc_gram<64>(U[k], I[k], Ru, gram_parts, nb, st);
        else noalloc_gram<128>(U[k], I[k], Ru, gram_parts, nb, st);
        hipLaunchKernelGGL(k_gram_reduce, dim3(grid_for(T, 256)), dim3(256), 0, st, gram_parts, nb, (uint32_t)T, 1u,
                           gram_red + (uint64_t)k * T);
        FTL_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_chain_final, dim3(1), dim3(256), 0, st, gram_red, N, Ru, sigma, scal + 1);
    FTL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_add_pair, dim3(1), dim3(64), 0, st, scal, out);
    FTL_LAUNCH_CHECK();
}

}  // namespace ftl

extern "C" {

static int sparse_create_common(uint32_t N, uint64_t nnz, const uint32_t* idx, const double* values,
                                const uint64_t* dims, bool on_device, ftl_sparse** out) {
    return guarded([&] {
        FTL_REQUIRE(out, "null output handle");
        *out = nullptr;
        std::unique_ptr<ftl_sparse> s(new ftl_sparse());
        s->impl.build(N, nnz, idx, values, dims, on_device);
        *out = s.release();
    });
}

int ftl_sparse_create(uint32_t N, uint64_t nnz, const uint32_t* idx_host, const double* values_host,
                      const uint64_t* dims, ftl_sparse** out) {
    return sparse_create_common(N, nnz, idx_host, values_host, dims, false, out);
}

int ftl_sparse_create_device(uint32_t N, uint64_t nnz, const uint32_t* idx_dev, const double* values_dev,
                             const uint64_t* dims, ftl_sparse** out) {
    return sparse_create_common(N, nnz, idx_dev, values_dev, dims, true, out);
}

int ftl_sparse_destroy(ftl_sparse* X) {
    return guarded([&] { delete X; });
}

int ftl_sparse_info(ftl_sparse* X, uint64_t* out) {
    return guarded([&] {
        FTL_REQUIRE(X && out, "null argument");
        out[0] = X->impl.N;
        out[1] = X->impl.nnz;
        for (uint32_t k = 0; k < X->impl.N; k++) out[2 + k] = X->impl.dims[k];
    });
}

int ftl_sparse_set_lookup(ftl_sparse* X, int method) {
    return guarded([&] {
        FTL_REQUIRE(X, "null sparse tensor");
        FTL_REQUIRE(method == 0 || method == 1, "lookup method must be 0 (sort) or 1 (hash)");
        if (method == 1) X->impl.use_hash();
        else X->impl.method = 0;
    });
}

int ftl_sparse_normsq(ftl_sparse* X, double* out) {
    return guarded([&] {
        FTL_REQUIRE(X && out, "null argument");
        *out = X->impl.normsq;
    });
}

// Probes host samples (Jp x N uint64) on mode j: leaves ranges and the match
// total in the tensor's match buffers; returns T.
static uint64_t probe_host_samples(ftl::SparseDev& S, uint32_t j, uint64_t Jp, const uint64_t* samples_host) {
    using namespace ftl;
    FTL_REQUIRE(j < S.N, "mode index out of range");
    FTL_REQUIRE(Jp >= 1 && Jp < (1ull << 31), "sample count must be in [1, 2^31)");
    FTL_REQUIRE(samples_host, "null samples");
    hipStream_t st = S.stream;
    uint32_t* tup = S.tuples(Jp);
    uint64_t* stage = S.mem.alloc<uint64_t>(Jp * S.N);
    FTL_HIP_CHECK(hipMemcpyAsync(stage, samples_host, Jp * S.N * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_tuples_u32, dim3(grid_for(Jp * S.N, 256)), dim3(256), 0, st, stage, Jp * S.N, tup);
    FTL_HIP_CHECK(hipGetLastError());
    probe_and_count(S, j, tup, nullptr, Jp, Jp, S.mt, S.mem, S.prim, st);
    hipLaunchKernelGGL(k_total_hi, dim3(1), dim3(1024), 0, st, S.mt.cnt, Jp, S.mt.counters, (int)kCntMatchesHi);
    FTL_HIP_CHECK(hipGetLastError());
    FTL_HIP_CHECK(hipMemcpyAsync(S.mt.host_counters, S.mt.counters, kCntSlots * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, st));
    FTL_HIP_CHECK(hipStreamSynchronize(st));
    S.mem.release(stage);
    // the per-sample offsets and the match buffers are indexed in 32 bits
    if (S.mt.host_counters[kCntMatchesHi] != 0)
        throw Error{FTL_EUNSUPPORTED, "the samples match 2^32 or more nonzeros in total (" +
                                          std::to_string(((uint64_t)S.mt.host_counters[kCntMatchesHi] << 32) |
                                                         S.mt.host_counters[kCntMatches]) +
                                          "); pass fewer or deduplicated samples"};
    return S.mt.host_counters[kCntMatches];
}

int ftl_sparse_mttkrp(ftl_sparse* X, uint32_t j, uint64_t Jp, const uint64_t* samples_host, const double* lhs_host,
                      uint64_t R, double* result_host, uint64_t result_rows, uint64_t* found) {
    return guarded([&] {
        using namespace ftl;
        FTL_REQUIRE(X && lhs_host && result_host, "null argument");
        SparseDev& S = X->impl;
        FTL_REQUIRE(R >= 1 && R <= 128, "R must be in [1, 128]");
        FTL_REQUIRE(j < S.N, "mode index out of range");
        FTL_REQUIRE(result_rows >= S.dims[j], "result has fewer rows than the tensor's mode size");
        const uint64_t T = probe_host_samples(S, j, Jp, samples_host);
        hipStream_t st = S.stream;
        Arena tmp;
        StreamDrain drain{st};
        double* lhs = tmp.alloc<double>(Jp * R);
        double* res = tmp.alloc<double>(result_rows * R);
        FTL_HIP_CHECK(hipMemcpyAsync(lhs, lhs_host, Jp * R * sizeof(double), hipMemcpyHostToDevice, st));
        FTL_HIP_CHECK(hipMemsetAsync(res, 0, result_rows * R * sizeof(double), st));
        if (T > 0)
            reduce_matches(S, j, Jp, T, lhs, (uint32_t)R, (uint32_t)R, nullptr, res, nullptr, nullptr, S.mt, S.mem,
                           S.prim, st);
        FTL_HIP_CHECK(hipMemcpyAsync(result_host, res, result_rows * R * sizeof(double), hipMemcpyDeviceToHost, st));
        FTL_HIP_CHECK(hipStreamSynchronize(st));
        if (found) *found = T;
    });
}

int ftl_sparse_lookup(ftl_sparse* X, uint32_t j, uint64_t Jp, const uint64_t* samples_host, uint64_t* counts_host,
                      uint64_t* ids_host, uint64_t ids_cap, uint64_t* found) {
    return guarded([&] {
        using namespace ftl;
        FTL_REQUIRE(X, "null sparse tensor");
        SparseDev& S = X->impl;
        const uint64_t T = probe_host_samples(S, j, Jp, samples_host);
        hipStream_t st = S.stream;
        if (counts_host) {
            std::vector<uint32_t> c(Jp);
            FTL_HIP_CHECK(hipMemcpyAsync(c.data(), S.mt.cnt, Jp * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            FTL_HIP_CHECK(hipStreamSynchronize(st));
            for (uint64_t s = 0; s < Jp; s++) counts_host[s] = c[s];
        }
        if (ids_host && T > 0) {
            Arena tmp;
            StreamDrain drain{st};
            uint64_t* ids = tmp.alloc<uint64_t>(T);
            hipLaunchKernelGGL(k_match_ids, dim3(grid_for(T, 256)), dim3(256), 0, st, S.mt.off, S.mt.lo, Jp, T,
                               S.perm[j], ids);
            FTL_HIP_CHECK(hipGetLastError());
            FTL_HIP_CHECK(hipMemcpyAsync(ids_host, ids, std::min(T, ids_cap) * sizeof(uint64_t),
                                         hipMemcpyDeviceToHost, st));
            FTL_HIP_CHECK(hipStreamSynchronize(st));
        }
        if (found) *found = T;
    });
}

static int als_create_common(ftl_sparse* X, uint32_t N, const uint64_t* I, uint64_t R, const double* const* U_host,
                             ftl_als** out) {
    return guarded([&] {
        FTL_REQUIRE(out, "null output handle");
        *out = nullptr;
        std::unique_ptr<ftl_als> a(new ftl_als());
        a->impl.init(X ? &X->impl : nullptr, N, I, R, U_host);
        *out = a.release();
    });
}

int ftl_sparse_residual_normsq_noalloc(ftl_sparse* X, uint64_t R, const uint64_t* I, const double* sigma_dev,
                                       const double* const* U_dev, double* out_dev, void* stream) {
    return guarded([&] {
        FTL_REQUIRE(X, "null sparse tensor");
        hipStream_t st = stream ? (hipStream_t)stream : X->impl.stream;
        ftl::residual_normsq_noalloc(X->impl, R, I, sigma_dev, U_dev, out_dev, st);
    });
}

int ftl_sparse_residual_normsq(ftl_sparse* X, uint64_t R, const uint64_t* I, const double* sigma_host,
                               const double* const* U_host, double* out) {
    return guarded([&] {
        FTL_REQUIRE(X && sigma_host && out, "null argument");
        ftl::Als a;
        a.init(&X->impl, X->impl.N, I, R, U_host);
        FTL_HIP_CHECK(hipMemcpy(a.sigma, sigma_host, R * sizeof(double), hipMemcpyHostToDevice));
        *out = a.residual_normsq();
    });
}

int ftl_als_create(ftl_sparse* X, uint32_t N, const uint64_t* I, uint64_t R, const double* const* U_host,
                   ftl_als** out) {
    return als_create_common(X, N, I, R, U_host, out);
}

int ftl_als_destroy(ftl_als* a) {
    return guarded([&] { delete a; });
}

int ftl_als_initialize_ds(ftl_als* a, uint64_t J, int precision) {
    return guarded([&] {
        FTL_REQUIRE(a, "null handle");
        FTL_REQUIRE(precision == FTL_FP64 || precision == FTL_FP32, "precision must be FTL_FP64 or FTL_FP32");
        a->impl.initialize_ds(J, precision);
    });
}

int ftl_als_initialize_ds_sampler(ftl_als* a, uint64_t J, int sampler_type, int precision) {
    return guarded([&] {
        FTL_REQUIRE(a, "null handle");
        FTL_REQUIRE(precision == FTL_FP64 || precision == FTL_FP32, "precision must be FTL_FP64 or FTL_FP32");
        FTL_REQUIRE(sampler_type == FTL_SAMPLER_EFFICIENT || sampler_type == FTL_SAMPLER_LARSEN_KOLDA ||
                        sampler_type == FTL_SAMPLER_LARSEN_KOLDA_HYBRID,
                    "sampler_type must be FTL_SAMPLER_EFFICIENT, FTL_SAMPLER_LARSEN_KOLDA or "
                    "FTL_SAMPLER_LARSEN_KOLDA_HYBRID");
        a->impl.initialize_ds(J, precision, sampler_type);
    });
}

int ftl_als_set_seed(ftl_als* a, uint64_t seed, uint64_t counter) {
    return guarded([&] {
        FTL_REQUIRE(a && (a->impl.sampler || a->impl.lk), "initialize_ds_als has not been called");
        if (a->impl.lk) {
            a->impl.lk->seed = seed;
            a->impl.lk->call = counter;
        } else {
            ftl::sampler_set_seed(a->impl.sampler, seed, counter);
        }
    });
}

int ftl_als_ds_update_device(ftl_als* a, uint32_t j, int renormalize, int update_sampler, const double* draws_dev) {
    return guarded([&] {
        FTL_REQUIRE(a, "null handle");
        a->impl.ds_update(j, renormalize != 0, update_sampler != 0, draws_dev);
    });
}

int ftl_als_ds_update(ftl_als* a, uint32_t j, int renormalize, int update_sampler, const double* draws_host) {
    return guarded([&] {
        FTL_REQUIRE(a, "null handle");
        ftl::Als& A = a->impl;
        FTL_REQUIRE(A.sampler || A.lk, "initialize_ds_als has not been called");
        const double* d = nullptr;
        if (draws_host) {
            if (!A.draws_stage) A.draws_stage = A.mem.alloc<double>(2 * A.N * A.J);
            FTL_HIP_CHECK(hipMemcpyAsync(A.draws_stage, draws_host, 2 * A.N * A.J * sizeof(double),
                                         hipMemcpyHostToDevice, A.stream));
            d = A.draws_stage;
        }
        A.ds_update(j, renormalize != 0, update_sampler != 0, d);
        FTL_HIP_CHECK(hipStreamSynchronize(A.stream));
    });
}

int ftl_als_exact_update(ftl_als* a, uint32_t j, int renormalize, int update_sampler) {
    return guarded([&] {
        FTL_REQUIRE(a, "null handle");
        a->impl.exact_update(j, renormalize != 0, update_sampler != 0);
        FTL_HIP_CHECK(hipStreamSynchronize(a->impl.stream));
    });
}

int ftl_als_get_factor(ftl_als* a, uint32_t k, double* U_host) {
    return guarded([&] {
        FTL_REQUIRE(a && U_host, "null argument");
        ftl::Als& A = a->impl;
        FTL_REQUIRE(k < A.N, "mode index out of range");
        FTL_HIP_CHECK(hipMemcpyAsync(U_host, A.U[k], A.I[k] * A.R * sizeof(double), hipMemcpyDeviceToHost, A.stream));
        FTL_HIP_CHECK(hipStreamSynchronize(A.stream));
    });
}

int ftl_als_set_factor(ftl_als* a, uint32_t k, const double* U_host) {
    return guarded([&] {
        FTL_REQUIRE(a && U_host, "null argument");
        ftl::Als& A = a->impl;
        FTL_REQUIRE(k < A.N, "mode index out of range");
        FTL_HIP_CHECK(hipMemcpyAsync(A.U[k], U_host, A.I[k] * A.R * sizeof(double), hipMemcpyHostToDevice, A.stream));
        FTL_HIP_CHECK(hipStreamSynchronize(A.stream));
    });
}

int ftl_als_update_sampler(ftl_als* a, uint32_t k) {
    return guarded([&] {
        FTL_REQUIRE(a && (a->impl.sampler || a->impl.lk), "initialize_ds_als has not been called");
        FTL_REQUIRE(k < a->impl.N, "mode index out of range");
        if (a->impl.lk) a->impl.lk->update(k);
        else ftl::sampler_rebuild(a->impl.sampler, k, a->impl.stream);
        FTL_HIP_CHECK(hipStreamSynchronize(a->impl.stream));
    });
}

// ---------------------------------------------------------------------------
// Larsen-Kolda sampler handle (Sampler(U, J, R, "larsen_kolda"))
// ---------------------------------------------------------------------------
struct ftl_lk_sampler {
    ftl::LKSampler impl;
};

int ftl_lk_sampler_create(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, const double* const* U_host,
                          ftl_lk_sampler** out) {
    return guarded([&] {
        FTL_REQUIRE(out, "null output handle");
        *out = nullptr;
        std::unique_ptr<ftl_lk_sampler> s(new ftl_lk_sampler());
        s->impl.init(N, I, J, R, const_cast<double* const*>(U_host), false, nullptr);
        *out = s.release();
    });
}

int ftl_lkh_sampler_create(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, const double* const* U_host,
                           ftl_lk_sampler** out) {
    return guarded([&] {
        FTL_REQUIRE(out, "null output handle");
        *out = nullptr;
        std::unique_ptr<ftl_lk_sampler> s(new ftl_lk_sampler());
        s->impl.init(N, I, J, R, const_cast<double* const*>(U_host), false, nullptr, true);
        *out = s.release();
    });
}

int ftl_lk_sampler_info(ftl_lk_sampler* s, double* out5) {
    return guarded([&] {
        FTL_REQUIRE(s && out5, "null argument");
        for (int i = 0; i < 5; i++) out5[i] = s->impl.info[i];
    });
}

int ftl_lk_sampler_destroy(ftl_lk_sampler* s) {
    return guarded([&] { delete s; });
}

int ftl_lk_sampler_draw(ftl_lk_sampler* s, uint32_t j, uint64_t* samples, double* h, double* weights,
                        const double* draws) {
    return guarded([&] {
        FTL_REQUIRE(s, "null handle");
        s->impl.draw_host(j, samples, h, weights, draws);
    });
}

int ftl_lk_sampler_draw_device(ftl_lk_sampler* s, uint32_t j, uint64_t* samples, double* h, double* weights,
                               const double* draws, void* stream) {
    return guarded([&] {
        FTL_REQUIRE(s, "null handle");
        s->impl.draw_device(j, samples, h, weights, draws, (hipStream_t)stream);
    });
}

int ftl_lk_sampler_update(ftl_lk_sampler* s, uint32_t j, const double* Uj_host) {
    return guarded([&] {
        FTL_REQUIRE(s, "null handle");
        ftl::LKSampler& L = s->impl;
        FTL_REQUIRE(j < L.N, "update_sampler: mode index out of range");
        if (Uj_host)
            FTL_HIP_CHECK(hipMemcpyAsync(L.U[j], Uj_host, L.I[j] * L.R * sizeof(double), hipMemcpyHostToDevice,
                                         L.stream));
        L.update(j);
        FTL_HIP_CHECK(hipStreamSynchronize(L.stream));
    });
}

void* ftl_lk_sampler_stream(ftl_lk_sampler* s) { return s ? (void*)s->impl.stream : nullptr; }

int ftl_lk_sampler_factor_device(ftl_lk_sampler* s, uint32_t k, double** out) {
    return guarded([&] {
        FTL_REQUIRE(s && out, "null argument");
        FTL_REQUIRE(k < s->impl.N, "mode index out of range");
        *out = s->impl.U[k];
    });
}

int ftl_lk_sampler_set_seed(ftl_lk_sampler* s, uint64_t seed, uint64_t counter) {
    return guarded([&] {
        FTL_REQUIRE(s, "null handle");
        s->impl.seed = seed;
        s->impl.call = counter;
    });
}

int ftl_lk_sampler_leverage(ftl_lk_sampler* s, uint32_t k, double* lev_host, double* total_host) {
    return guarded([&] {
        FTL_REQUIRE(s && lev_host, "null argument");
        ftl::LKSampler& L = s->impl;
        FTL_REQUIRE(k < L.N, "mode index out of range");
        FTL_HIP_CHECK(hipMemcpyAsync(lev_host, L.lev[k], L.I[k] * sizeof(double), hipMemcpyDeviceToHost, L.stream));
        if (total_host)
            FTL_HIP_CHECK(hipMemcpyAsync(total_host, L.total + k, sizeof(double), hipMemcpyDeviceToHost, L.stream));
        FTL_HIP_CHECK(hipStreamSynchronize(L.stream));
    });
}

int ftl_als_get_sigma(ftl_als* a, double* sigma_host) {
    return guarded([&] {
        FTL_REQUIRE(a && sigma_host, "null argument");
        FTL_HIP_CHECK(hipMemcpyAsync(sigma_host, a->impl.sigma, a->impl.R * sizeof(double), hipMemcpyDeviceToHost,
                                     a->impl.stream));
        FTL_HIP_CHECK(hipStreamSynchronize(a->impl.stream));
    });
}

int ftl_als_set_sigma(ftl_als* a, const double* sigma_host) {
    return guarded([&] {
        FTL_REQUIRE(a && sigma_host, "null argument");
        FTL_HIP_CHECK(hipMemcpyAsync(a->impl.sigma, sigma_host, a->impl.R * sizeof(double), hipMemcpyHostToDevice,
                                     a->impl.stream));
        FTL_HIP_CHECK(hipStreamSynchronize(a->impl.stream));
    });
}

int ftl_als_renormalize_columns(ftl_als* a, int j) {
    return guarded([&] {
        FTL_REQUIRE(a, "null handle");
        a->impl.renormalize(j);
    });
}

int ftl_als_residual_normsq(ftl_als* a, double* out) {
    return guarded([&] {
        FTL_REQUIRE(a && out, "null argument");
        *out = a->impl.residual_normsq();
    });
}

int ftl_als_lowrank_normsq(ftl_als* a, double* out) {
    return guarded([&] {
        FTL_REQUIRE(a && out, "null argument");
        *out = a->impl.lowrank_normsq();
    });
}

int ftl_als_factor_device(ftl_als* a, uint32_t k, double** out) {
    return guarded([&] {
        FTL_REQUIRE(a && out, "null argument");
        FTL_REQUIRE(k < a->impl.N, "mode index out of range");
        *out = a->impl.U[k];
    });
}

int ftl_als_last_stats(ftl_als* a, double* out8) {
    return guarded([&] {
        FTL_REQUIRE(a && out8, "null argument");
        a->impl.stats(out8);
    });
}

int ftl_als_last_draw(ftl_als* a, uint64_t* samples_host, double* h_host, double* w_host) {
    return guarded([&] {
        FTL_REQUIRE(a && samples_host, "null argument");
        a->impl.last_draw(samples_host, h_host, w_host);
    });
}

int ftl_als_enable_timing(ftl_als* a, int enable) {
    return guarded([&] {
        FTL_REQUIRE(a, "null handle");
        a->impl.timing = enable != 0;
    });
}

void* ftl_als_stream(ftl_als* a) { return a ? (void*)a->impl.stream : nullptr; }

}  // extern "C"

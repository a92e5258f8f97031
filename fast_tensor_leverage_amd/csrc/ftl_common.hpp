// ftl_common.hpp -- host/device helpers shared by the sampler and ALS sources
// of libftl_hip.so: status handling, tree shapes, packed-triangle sizes and
// the Philox4x32-10 counter-based generator.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ftl_hip.h"

namespace ftl {

// most modes a sampler or sparse tensor handles
constexpr int kMaxModes = 16;

// ---------------------------------------------------------------------------
// status handling: every entry point returns an ftl_status and records the
// cause in a thread-local string (the reference exits instead,
// als_module.cpp:57-61, sparse_tensor.hpp:64-66).
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* get_error();

struct Error {
    int code;
    std::string msg;
};

#define FTL_HIP_CHECK(expr)                                                               \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            throw ::ftl::Error{_e == hipErrorOutOfMemory ? FTL_ENOMEM : FTL_EHIP,         \
                               std::string(#expr) + ": " + hipGetErrorString(_e)};        \
        }                                                                                 \
    } while (0)

#define FTL_REQUIRE(cond, msg)                                                            \
    do {                                                                                  \
        if (!(cond)) throw ::ftl::Error{FTL_EINVAL, std::string(msg)};                    \
    } while (0)

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return FTL_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return FTL_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return FTL_EINVAL;
    }
}

// ---------------------------------------------------------------------------
// Tree shape (common.h:17-33, partition_tree.hpp:83-118): heap-ordered
// complete binary tree over ceil(n/F) leaf blocks of F rows.  Leaves on the
// partial bottom level hold the first row blocks, leaves on the last full
// level the rest, both in order.
// ---------------------------------------------------------------------------
struct Shape {
    uint64_t n, F, leaf_count, node_count;
    uint64_t lfill_level, lfill_count;
    uint64_t nodes_upto_lfill, nodes_before_lfill, complete_level_offset;
};

__host__ __device__ inline Shape make_shape(uint64_t n, uint64_t F) {
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
    s.complete_level_offset = s.nodes_before_lfill - (s.node_count - s.nodes_upto_lfill) / 2;
    return s;
}

__host__ __device__ inline uint64_t leaf_index(const Shape& s, uint64_t c) {
    return c >= s.nodes_upto_lfill ? c - s.nodes_upto_lfill : c - s.complete_level_offset;
}

// Depth of node c in the heap (root depth 0).
__host__ __device__ inline uint32_t node_depth(uint64_t c) {
    uint32_t d = 0;
    uint64_t v = c + 1;
    while (v > 1) {
        v >>= 1;
        d++;
    }
    return d;
}

// Packed upper triangle: element (a, b), a <= b, of an R x R symmetric matrix
// sits at a*R - a*(a-1)/2 + (b - a).  Node strides are rounded up so every
// node starts on a 128-byte line.
__host__ __device__ inline uint64_t tri_size(uint64_t R) { return R * (R + 1) / 2; }
__host__ __device__ inline uint64_t tri_stride64(uint64_t R) { return (tri_size(R) + 15) / 16 * 16; }
__host__ __device__ inline uint64_t tri_stride32(uint64_t R) { return (tri_size(R) + 31) / 32 * 32; }
__host__ __device__ inline uint64_t tri_offset(uint64_t R, uint64_t a, uint64_t b) {
    return a * R - a * (a - 1) / 2 + (b - a);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl increments
// 0x9E3779B9 / 0xBB67AE85, 10 rounds.  The uniform for (sample i, stage t) of
// call `ctr` is built from the first two output words as
// ((w0 >> 5) * 2^26 + (w1 >> 6)) * 2^-53, a double in [0, 1).
// ---------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ ctr[3] ^ k1;
        uint32_t n3 = lo0;
        ctr[0] = n0;
        ctr[1] = n1;
        ctr[2] = n2;
        ctr[3] = n3;
        k0 += W0;
        k1 += W1;
    }
}

__host__ __device__ inline double philox_uniform(uint64_t seed, uint64_t call, uint64_t sample,
                                                 uint32_t stage) {
    uint32_t c[4] = {(uint32_t)sample, (uint32_t)(sample >> 32) ^ (stage << 16), (uint32_t)call,
                     (uint32_t)(call >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = c[0] >> 5, b = c[1] >> 6;
    return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
}

}  // namespace ftl

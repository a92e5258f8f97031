// sampler_internal.hpp -- the part of the sampler that the ALS step drives
// from another translation unit: a sampler whose FP64 factors are device
// arrays owned by the caller (the ALS factor matrices, updated in place, as
// the reference's sampler aliases cp_decomp.U, als.hpp:39-55), its draw and
// its tree rebuild, all enqueued on a caller-chosen stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ftl {

struct Sampler;

// Device memory source for a sampler (a serving arena); alloc returns null
// when the request does not fit.
struct DeviceAllocator {
    virtual void* alloc(size_t bytes) = 0;
    virtual void release(void* p, size_t bytes) = 0;
    virtual ~DeviceAllocator() = default;
};

Sampler* sampler_new_aliased(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, double* const* U_dev,
                             int precision);
void sampler_delete(Sampler* s);
// Same, allocating from `alloc` and leaving the tree builds queued on stream
// `st` (borrowed; no host wait).
Sampler* sampler_new_async(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, double* const* U_dev, int precision,
                           DeviceAllocator* alloc, hipStream_t st);
hipStream_t sampler_stream(Sampler* s);
// Device bytes sampler_new_async allocates (blocks rounded up to `align`).
size_t sampler_footprint(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, int precision, size_t align);
// KRPDrawSamples(j) into device buffers (samples N x J, h J x R, w J); draws
// NULL -> the built-in Philox uniforms of the next call
void sampler_draw(Sampler* s, uint32_t j, uint64_t* samples, double* h, double* w, const double* draws,
                  hipStream_t st);
// update_sampler(j): rebuild tree j from the (aliased) factor j
void sampler_rebuild(Sampler* s, uint32_t j, hipStream_t st);
void sampler_set_seed(Sampler* s, uint64_t seed, uint64_t counter);
// Packed draws (one launch of each computeM kernel and one of the draw
// kernel for several samplers of the same R, precision and launch options,
// each drawing all its sampled modes with its built-in uniforms; the samples
// equal those of its own draw).  dev_args / host_stage hold
// packed_args_bytes(n) bytes (host_stage pinned, reusable once the copy
// queued on st has run); the samplers' trees must be complete in st's order.
struct PackItem {
    Sampler* s;
    uint32_t j;
    uint64_t* samples;  // N x J device
    uint64_t J;         // samples of this draw, at most the sampler's J
};
// Sampler's J (capacity of its draw buffers) and built-in generator state.
uint64_t sampler_capacity(const Sampler* s);
size_t packed_args_bytes(uint32_t max_req);
// ftl_sampler_set_option on an internal sampler.
void sampler_set_option(Sampler* s, const char* key, int64_t value);
bool sampler_packable(const Sampler* s);
uint64_t sampler_pack_key(const Sampler* s);
void sampler_draw_packed(const PackItem* items, uint32_t n, void* dev_args, void* host_stage, hipStream_t st);

// The factors of a public sampler handle (device arrays it owns) and its shape.
struct SamplerView {
    uint32_t N;
    uint64_t I[16];
    const double* U[16];
    uint64_t R, J;
    int precision;
};

// Eigensolves of als.hip's kernels that stopped at the Jacobi sweep cap on
// the current device (waits for the device).
uint64_t jacobi_capped_als_tu();

}  // namespace ftl

struct ftl_sampler;
namespace ftl {
SamplerView sampler_view(ftl_sampler* s);
}  // namespace ftl

/*
 * ftl_hip.h -- C ABI of libftl_hip.so, the MI355X (gfx950) implementation of
 * the exact Khatri-Rao-product leverage-score sampler and the sampled CP-ALS
 * step of fast_tensor_leverage.
 *
 * Plain C types only: host pointers are C-contiguous row-major arrays, device
 * pointers are HIP device addresses, streams are hipStream_t passed as void*.
 * Every entry point returns an ftl_status; on failure ftl_last_error() returns
 * a message (thread-local) describing the cause.  No entry point exits the
 * process (the reference calls exit(1) on bad arguments, als_module.cpp:57-61).
 *
 * Threading: a handle is not re-entrant (the reference shares one scratch
 * buffer per sampler, partition_tree.hpp:19-46); callers serialise calls on
 * one handle.  Different handles may be used from different threads.
 */
#ifndef FTL_HIP_H
#define FTL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ftl_status {
    FTL_OK = 0,
    FTL_EINVAL = 1,       /* bad argument (shape, index, null pointer) */
    FTL_EHIP = 2,         /* a HIP runtime call failed */
    FTL_ENOMEM = 3,       /* device allocation failed or over the HBM budget */
    FTL_EUNSUPPORTED = 4  /* size outside what the kernels are built for */
} ftl_status;

typedef enum ftl_precision {
    FTL_FP64 = 0, /* every tree level, leaf scan and update in FP64 */
    FTL_FP32 = 1  /* FP64 top tree levels, FP32 bottom levels and leaf scan */
} ftl_precision;

typedef struct ftl_sampler ftl_sampler;

/* Message of the last failed call on this thread ("" if none). */
const char* ftl_last_error(void);

/* Library version string. */
const char* ftl_version(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int ftl_device_count(void);

/*
 * Sampler(matrices, J, R, "efficient") -- replaces PySampler's ctor
 * (cpp_ext/als_module.cpp:34-44) and EfficientKRPSampler's ctor
 * (cpp_ext/efficient_krp_sampler.hpp:32-67): copies the N factor matrices
 * U_host[k] (I[k] x R, FP64, host) to HBM and builds the N Gram trees there.
 * max_rows_per_leaf F = min(R, I[k]) as in the reference.  precision is an
 * ftl_precision.  R <= 128, N <= 16, J >= 1.
 */
int ftl_sampler_create(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R,
                       const double* const* U_host, int precision, ftl_sampler** out);

/* Same, from device-resident factors (copied into the sampler's own buffers). */
int ftl_sampler_create_device(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R,
                              const double* const* U_dev, int precision, ftl_sampler** out);

/* Frees every device buffer of the handle.  Accepts NULL. */
int ftl_sampler_destroy(ftl_sampler* s);

/*
 * update_sampler(j) -- efficient_krp_sampler.hpp:73-75: replaces factor j
 * with Uj (I[j] x R, host or device per the _device suffix) and rebuilds
 * Gram tree j only.
 */
int ftl_sampler_update(ftl_sampler* s, uint32_t j, const double* Uj_host);
int ftl_sampler_update_device(ftl_sampler* s, uint32_t j, const double* Uj_dev);

/*
 * KRPDrawSamples(j, samples) -- als_module.cpp:64-68 and
 * efficient_krp_sampler.hpp:209-268.  Draws J rows of the KRP of all factors
 * except j (every j >= N samples all modes) with probability equal to their
 * exact leverage score.  samples_host is N x J uint64 row-major; row k holds
 * the row index drawn from factor k (row j is left untouched).
 * draws_host: NULL to use the sampler's counter-based generator (Philox4x32-10,
 * advanced on every call), or 2*N*J uniforms in [0,1) laid out as in
 * efficient_krp_sampler.hpp:221-229 (eigen draws of mode k at [k*J,(k+1)*J),
 * Gram draws at [N*J + k*J, ...)).
 */
int ftl_sampler_draw(ftl_sampler* s, uint32_t j, uint64_t* samples_host, const double* draws_host);

/*
 * KRPDrawSamples_materialize -- als_module.cpp:70-81: also returns the sampled
 * KRP rows h (J x R) and the weights R / (J * h^T pinv(G) h) (J).
 */
int ftl_sampler_draw_materialize(ftl_sampler* s, uint32_t j, uint64_t* samples_host,
                                 double* h_host, double* weights_host, const double* draws_host);

/*
 * Device-resident KRPDrawSamples: every pointer is a device address, work is
 * enqueued on `stream` (hipStream_t; NULL = the handle's own stream) and the
 * call returns without synchronising.  h_dev / weights_dev / draws_dev may be
 * NULL.  samples_dev is N x J.
 */
int ftl_sampler_draw_device(ftl_sampler* s, uint32_t j, uint64_t* samples_dev, double* h_dev,
                            double* weights_dev, const double* draws_dev, void* stream);

/* Seed and counter of the built-in uniform generator. */
int ftl_sampler_set_seed(ftl_sampler* s, uint64_t seed, uint64_t counter);

/* Read back the Gram tree of mode k: node_count x R x R FP64 (upper triangle,
 * lower zero, as PartitionTree::G, partition_tree.hpp:61) -- test surface. */
int ftl_sampler_get_tree(ftl_sampler* s, uint32_t k, double* G_host, uint64_t node_capacity);

/* computeM(j) outputs (efficient_krp_sampler.hpp:90-194): pinv (R x R,
 * symmetric) and E (N x R x R, the eigen matrix rows of each sampled mode,
 * zero for the excluded mode) -- test surface. */
int ftl_sampler_compute_m(ftl_sampler* s, uint32_t j, double* pinv_host, double* E_host);

/* Shape facts: out[0..7] = N, J, R, precision, fp64_levels, device bytes held,
 * node_count of mode 0, lfill_level of mode 0. */
int ftl_sampler_info(ftl_sampler* s, uint64_t* out);

/* Stream of the handle (hipStream_t) for callers that want to order work. */
void* ftl_sampler_stream(ftl_sampler* s);

/* Tuning knobs (the autotuner's surface; no counterpart in the reference).
 * Keys: "group_lanes" (lanes per sample in the draw kernel: 8/16/32/64,
 * 0 = default), "grid_blocks" (cap on draw-kernel blocks, 0 = one block per
 * 256/group_lanes samples), "fp32_levels" (FP32 mode:
 * number of bottom tree levels stored FP32; rebuilds the FP32 copies).
 * get_option also accepts "fp64_levels" (FP64 levels of mode 0's tree). */
int ftl_sampler_set_option(ftl_sampler* s, const char* key, int64_t value);
int ftl_sampler_get_option(ftl_sampler* s, const char* key, int64_t* value);

/* Kernel timing with HIP events recorded on the launch stream.  While enabled,
 * every draw records four events without synchronising; last_timing waits for
 * the recorded calls and returns average durations (ms) over them:
 * out[0] = k_gram_pinv, out[1] = k_eigen_modes, out[2] = k_krp_draw,
 * out[3] = number of calls averaged.  Reading clears the record. */
int ftl_sampler_enable_timing(ftl_sampler* s, int enable);
int ftl_sampler_last_timing(ftl_sampler* s, double* out4);

#ifdef __cplusplus
}
#endif

#endif /* FTL_HIP_H */

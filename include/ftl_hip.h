/*
 * ftl_hip.h -- C ABI of libftl_hip.so, the MI355X (gfx950) implementation of
 * the exact Khatri-Rao-product leverage-score sampler and the sampled CP-ALS
 * step of fast_tensor_leverage.
 *
 * Plain C types only: host pointers are C-contiguous row-major arrays, device
 * pointers are HIP device addresses, streams are hipStream_t passed as void*.
 * Handles run on blocking HIP streams of their own: their work is ordered
 * after work queued earlier on the legacy default stream (torch's default),
 * so device inputs written there are complete when read.  Inputs written on
 * other non-blocking streams must be synchronised by the caller.
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

/* Sampler kinds of ALS::initialize_ds_als (als.hpp:39-55, als_module.cpp:34-62). */
typedef enum ftl_sampler_type {
    FTL_SAMPLER_EFFICIENT = 0,    /* the exact KRP leverage sampler ("efficient") */
    FTL_SAMPLER_LARSEN_KOLDA = 1,  /* per-factor leverage product ("larsen_kolda") */
    FTL_SAMPLER_LARSEN_KOLDA_HYBRID = 2  /* deterministic + rejection ("larsen_kolda_hybrid") */
} ftl_sampler_type;

typedef struct ftl_sampler ftl_sampler;

/* Message of the last failed call on this thread ("" if none). */
const char* ftl_last_error(void);

/* Library version string. */
const char* ftl_version(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int ftl_device_count(void);

/* The sampler's tuning table (no counterpart in the reference): launch
 * shapes and the FP32/FP64 level split per (precision, R, tree depth, J)
 * bucket, written by the package's autotuner into ftl_tuning.tsv beside
 * libftl_hip.so (or the file named by FTL_TUNING_TABLE; empty disables it)
 * and applied when a sampler is created.  Looks up the row for a shape in
 * table_text (NULL: the loaded table) and writes "<file>\t<row>\t<k=v,...>"
 * to out (row -1 when none applies).  Host-only: needs no GPU. */
int ftl_tuning_lookup(const char* table_text, int precision, uint64_t R, uint64_t depth, uint64_t J,
                      char* out, uint64_t capacity);

/* Device blocks kept by the library (no counterpart in the reference): the
 * device memory of a destroyed sampler is kept, up to FTL_POOL_CAP_GB GiB
 * (environment; default a quarter of the device's memory; 0 keeps none; the
 * oldest kept blocks are released first to make room), and handed to
 * samplers built later with blocks of the same rounded size, so that
 * rebuilding samplers of one shape maps no new memory.
 * ftl_device_pool_trim releases every kept block (hipFree);
 * ftl_device_pool_held reports their bytes.  An allocation of the library
 * that fails first releases the kept blocks of its device. */
int ftl_device_pool_trim(void);
int ftl_device_pool_held(uint64_t* bytes);
/* The most bytes kept (replaces FTL_POOL_CAP_GB; 0 keeps none); kept blocks
 * beyond a lowered cap are released, the oldest first. */
int ftl_device_pool_set_cap(uint64_t bytes);
int ftl_device_pool_cap(uint64_t* bytes);

/* Eigensolves (computeM, the pseudo-inverses of the ALS and Larsen-Kolda
 * Grams) on the current device that stopped at the 40-sweep cap of the
 * cyclic Jacobi iteration before converging, since the library was loaded.
 * LAPACK's dsyev, which the reference calls (efficient_krp_sampler.hpp:118,
 * 168; common.h:257, 288), reports the same condition as info > 0.  Waits
 * for the device.  No counterpart entry point in the reference. */
int ftl_eigensolver_capped(uint64_t* out);

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
 * NULL.  samples_dev is N x J.  A draw on another stream is ordered after the
 * factor copies and tree builds queued before it on the handle's stream, and
 * the next update_sampler (or fp32_levels change) waits for it on the device.
 * Draws of one handle share its scratch buffers: draws on different streams
 * must not run concurrently (order them with events or one stream).
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
 * 256/group_lanes samples), "jacobi_threads" (block size of the computeM
 * eigensolver kernels: 256/512/1024, 0 = 1024), "mode_split" (1: one draw
 * launch per sampled mode with h carried in HBM between them, the default;
 * 0: every mode in one launch), "fp32_levels" (FP32 mode:
 * number of bottom tree levels stored FP32; rebuilds the FP32 copies),
 * "fp32_leaf" (FP32 mode: 1, the default, scans the leaf rows of the FP32
 * factor copy; 0 the FP64 factors), "leaf_mfma" (FP32 mode, R <= 64: 1, the
 * default, builds the leaf Grams on the FP64 matrix cores -- equal to the
 * sequential row sums to rounding; 0: the sequential sums, bitwise those of
 * an FP64 sampler; applies to later builds),
 * "draw_kernel" (0 = by rank: a group of lanes per sample for R <= 64, one
 * sample per block above; 1 = one sample per wave, R <= 32; 2 = one sample
 * per block, R > 32; 3 = a group of lanes per sample at any rank),
 * "block_threads" (threads per block of the one-sample-per-block kernel:
 * 128/256/512, 0 = 256), "xcd_buckets" (1: descents bucketed by level-3
 * subtree onto XCD-affine blocks, R <= 64), "overlap_compute_m" (1, the
 * default for a sampler with a stream of its own: computeM of a call runs on
 * a second stream into one of two buffer sets, overlapping the previous
 * call's draw kernels; 0: computeM and draw in order on one stream).
 * get_option also accepts "fp64_levels" (FP64 levels of mode 0's tree) and
 * "tuning_row" (the tuning-table row applied at creation, -1 for none). */
int ftl_sampler_set_option(ftl_sampler* s, const char* key, int64_t value);
int ftl_sampler_get_option(ftl_sampler* s, const char* key, int64_t* value);

/* Kernel timing with HIP events recorded on the launch streams.  While
 * enabled, every draw records five events without synchronising (computeM
 * start, after k_gram_pinv, after k_eigen_modes on the computeM stream; draw
 * start and end on the draw stream); last_timing waits for the recorded calls
 * and returns average durations (ms) over them: out[0] = k_gram_pinv,
 * out[1] = k_eigen_modes, out[2] = the draw kernels, out[3] = number of calls
 * averaged.  Reading clears the record. */
int ftl_sampler_enable_timing(ftl_sampler* s, int enable);
int ftl_sampler_last_timing(ftl_sampler* s, double* out4);


/* ------------------------------------------------------------------------
 * Sparse tensor and sampled CP-ALS
 * ------------------------------------------------------------------------ */
typedef struct ftl_sparse ftl_sparse;
typedef struct ftl_als ftl_als;

/*
 * SparseTensor(idx, values, "sort") -- replaces the SparseTensor ctor
 * (cpp_ext/sparse_tensor.hpp:46-77) and its per-mode SortIdxLookup ctor
 * (cpp_ext/sort_lookup.hpp:33-60), bound at cpp_ext/als_module.cpp:100-107.
 * idx: nnz x N uint32 row-major, values: nnz FP64 (host).  dims: NULL for
 * max coordinate + 1 per mode (the reference takes the sizes from the file,
 * tensors.py:85-93), else the N mode sizes.  For every mode j the nonzeros
 * are sorted in HBM by their coordinates on the other modes (ties keep input
 * order).  Limits: nnz < 2^32; per mode, the bit widths of the other modes'
 * sizes sum to <= 64 (FTL_EUNSUPPORTED otherwise).
 */
int ftl_sparse_create(uint32_t N, uint64_t nnz, const uint32_t* idx_host, const double* values_host,
                      const uint64_t* dims, ftl_sparse** out);

/* Same from device arrays (read during the call, not retained). */
int ftl_sparse_create_device(uint32_t N, uint64_t nnz, const uint32_t* idx_dev, const double* values_dev,
                             const uint64_t* dims, ftl_sparse** out);

int ftl_sparse_destroy(ftl_sparse* X);

/*
 * Lookup method of the tensor's probes: 0 = binary search of each mode's
 * sorted keys (SortIdxLookup, sort_lookup.hpp:33-132), 1 = a device
 * open-addressing coordinate hash per mode (linear probing, load factor
 * <= 1/2) mapping each distinct key to its run of the sorted nonzeros -- the
 * reference's HashIdxLookup (hash_lookup.hpp:13-162) in device form.  Both
 * find the same matches in the same order.  The hash is built on the first
 * call (needs per-mode keys of fewer than 64 bits: FTL_EUNSUPPORTED
 * otherwise).  SparseTensor(idx, values, "hash") selects 1.
 */
int ftl_sparse_set_lookup(ftl_sparse* X, int method);

/* out[0] = N, out[1] = nnz, out[2 + k] = size of mode k. */
int ftl_sparse_info(ftl_sparse* X, uint64_t* out);

/* ||X||^2 (sparse_tensor.hpp:70-74; SparseTensor::get_normsq). */
int ftl_sparse_normsq(ftl_sparse* X, double* out);

/*
 * execute_downsampled_mttkrp_py(samples, lhs, j, result) -- replaces
 * SparseTensor::execute_downsampled_mttkrp (sparse_tensor.hpp:95-115) and
 * SortIdxLookup::execute_spmm (sort_lookup.hpp:69-132): for each sample s
 * (Jp x N uint64, cast to uint32) every nonzero equal to it on all modes but
 * j adds value * lhs[s] (lhs Jp x R) to result row (its mode-j coordinate).
 * result (result_rows x R, result_rows >= size of mode j) is zero-filled
 * first.  *found = number of matched nonzeros (found_count).  Samples may
 * repeat (each repeat matches its nonzeros again); a total of 2^32 or more
 * matches fails with FTL_EUNSUPPORTED before anything is written.
 */
int ftl_sparse_mttkrp(ftl_sparse* X, uint32_t j, uint64_t Jp, const uint64_t* samples_host,
                      const double* lhs_host, uint64_t R, double* result_host, uint64_t result_rows,
                      uint64_t* found);

/*
 * The lookup half of the same call (sort_lookup.hpp:94-103): counts_host[s]
 * = matched nonzeros of sample s; ids_host (capacity ids_cap) = their input
 * positions, sample by sample, ascending within a sample.  Test surface.
 */
int ftl_sparse_lookup(ftl_sparse* X, uint32_t j, uint64_t Jp, const uint64_t* samples_host,
                      uint64_t* counts_host, uint64_t* ids_host, uint64_t ids_cap, uint64_t* found);

/*
 * compute_residual_normsq(sigma, U) -- replaces
 * SortIdxLookup::compute_residual_normsq (sort_lookup.hpp:134-197) with
 * ATB_chain_prod_sum (common.h:215-226), reached through
 * SparseTensor::compute_residual_normsq (sparse_tensor.hpp:117-119):
 * ||X||^2 - 2 <X, M> + ||M||^2 with M = [[sigma; U]].  U_host[k] is
 * I[k] x R FP64 (host), I[k] >= size of mode k.
 */
int ftl_sparse_residual_normsq(ftl_sparse* X, uint64_t R, const uint64_t* I, const double* sigma_host,
                               const double* const* U_host, double* out);

/*
 * Allocation-free ftl_sparse_residual_normsq on device arrays.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  sigma_dev (R), U_dev[k] (I[k] x R FP64, I[k] >= size of mode
 * k; the pointer list itself is a host array) and out_dev (one double) are
 * device addresses; the work is enqueued on `stream` (NULL: the tensor's
 * stream) without synchronising.  Intermediates live in a 9 MiB device array
 * of the library (one per GPU, held until the process ends): a shape whose
 * intermediates exceed it (4 + 4096 + (N + nb) R(R+1)/2 doubles, nb =
 * min(512, ceil(max I[k] / 4096))) fails with FTL_EUNSUPPORTED before any
 * work is queued.  Calls that overlap on one GPU share that array and race.
 * The result equals ftl_sparse_residual_normsq's bitwise.
 */
int ftl_sparse_residual_normsq_noalloc(ftl_sparse* X, uint64_t R, const uint64_t* I, const double* sigma_dev,
                                       const double* const* U_dev, double* out_dev, void* stream);

/*
 * ALS(LowRankTensor(R, U), X) -- replaces the LowRankTensor ctor
 * (low_rank_tensor.hpp:52-68, sigma = 1) and the ALS ctor (als.hpp:31-37),
 * bound at als_module.cpp:85-98, 111-115.  Copies the factors U_host[k]
 * (I[k] x R FP64) to HBM; they stay there and are updated in place.  X may be
 * NULL for a factor-only handle (renormalize / normsq only).
 */
int ftl_als_create(ftl_sparse* X, uint32_t N, const uint64_t* I, uint64_t R, const double* const* U_host,
                   ftl_als** out);
int ftl_als_destroy(ftl_als* a);

/* initialize_ds_als(J, "efficient") (als.hpp:39-55): the leverage sampler on
 * the handle's device factors (shared, not copied). */
int ftl_als_initialize_ds(ftl_als* a, uint64_t J, int precision);

/* initialize_ds_als(J, method) with the method as an ftl_sampler_type
 * (als.hpp:39-55): FTL_SAMPLER_EFFICIENT is ftl_als_initialize_ds;
 * FTL_SAMPLER_LARSEN_KOLDA builds the Larsen-Kolda sampler on the handle's
 * factors (precision does not apply to it). */
int ftl_als_initialize_ds_sampler(ftl_als* a, uint64_t J, int sampler_type, int precision);

/* Seed and counter of the sampler's built-in uniforms. */
int ftl_als_set_seed(ftl_als* a, uint64_t seed, uint64_t counter);

/*
 * execute_ds_als_update(j, renormalize, update_sampler) -- replaces
 * ALS::execute_ds_als_update (als.hpp:57-220): draw J samples excluding mode
 * j, sort/unique them with counts, w' = w c, h' = sqrt(w') h,
 * pinv(h'^T h') (cut 1e-11, target rank min(J', R)), h'' = sqrt(w') h',
 * sampled MTTKRP, U_j = MTTKRP pinv, optional renormalize_columns(j) and
 * update_sampler(j).  draws_host: NULL (built-in uniforms) or 2 N J uniforms
 * in the layout of efficient_krp_sampler.hpp:221-236.  Returns when done.
 */
int ftl_als_ds_update(ftl_als* a, uint32_t j, int renormalize, int update_sampler, const double* draws_host);

/* Same with the uniforms in HBM (or NULL), enqueued on ftl_als_stream(a);
 * waits once inside for the match count.  Returns before the tail finishes. */
int ftl_als_ds_update_device(ftl_als* a, uint32_t j, int renormalize, int update_sampler,
                             const double* draws_dev);

/*
 * execute_exact_als_update(j, renormalize, update_sampler) -- replaces
 * ALS::execute_exact_als_update (als.hpp:222-270): gram = Hadamard product of
 * U_k^T U_k over k != j, pinv = compute_pinv_square(gram, R) (cut 1e-11),
 * U_j = exact MTTKRP over every nonzero (sort_lookup.hpp:199-254) times pinv,
 * optional renormalize_columns(j).  update_sampler is accepted and unused, as
 * in the reference.  Needs no initialize_ds.  The first call for a mode sorts
 * the nonzeros by their mode-j coordinate (kept for later calls).  Returns
 * when done.
 */
int ftl_als_exact_update(ftl_als* a, uint32_t j, int renormalize, int update_sampler);

/* Factor k to / from host memory (I[k] x R FP64).  set_factor does not
 * rebuild the sampler's tree: call ftl_als_update_sampler (update_sampler,
 * efficient_krp_sampler.hpp:73-75). */
int ftl_als_get_factor(ftl_als* a, uint32_t k, double* U_host);
int ftl_als_set_factor(ftl_als* a, uint32_t k, const double* U_host);
int ftl_als_update_sampler(ftl_als* a, uint32_t k);

/* get_sigma / set_sigma (low_rank_tensor.hpp:181-193). */
int ftl_als_get_sigma(ftl_als* a, double* sigma_host);
int ftl_als_set_sigma(ftl_als* a, const double* sigma_host);

/* renormalize_columns(j), j = -1 for every mode (low_rank_tensor.hpp:203-260). */
int ftl_als_renormalize_columns(ftl_als* a, int j);

/* X.compute_residual_normsq(sigma, U) with the handle's device factors. */
int ftl_als_residual_normsq(ftl_als* a, double* out);

/* LowRankTensor::get_normsq (low_rank_tensor.hpp:71-78): ||[[sigma; U]]||^2. */
int ftl_als_lowrank_normsq(ftl_als* a, double* out);

/* Device address of factor k (I[k] x R FP64) for device-resident callers. */
int ftl_als_factor_device(ftl_als* a, uint32_t k, double** out);

/* Last update: out[0] = unique samples J', out[1] = matched nonzeros
 * (found_count), out[2] = touched rows of U_j, out[3] = reduction pieces;
 * with timing enabled, ms of: out[4] draw, out[5] dedup + pinv + probe,
 * out[6] MTTKRP + solve, out[7] renormalize + tree rebuild. */
int ftl_als_last_stats(ftl_als* a, double* out8);
int ftl_als_enable_timing(ftl_als* a, int enable);

/* The draw of the last sampled update (step 1 of execute_ds_als_update,
 * als.hpp:70-74): samples (N x J uint64, the excluded mode's row 0), h
 * (J x R, or NULL) and weights (J, or NULL), copied to the host.
 * FTL_EINVAL before the first sampled update. */
int ftl_als_last_draw(ftl_als* a, uint64_t* samples_host, double* h_host, double* w_host);

/* Stream of the handle (hipStream_t). */
void* ftl_als_stream(ftl_als* a);

/* ------------------------------------------------------------------------
 * Larsen-Kolda leverage sampler -- replaces LarsenKoldaSampler
 * (larsen_kolda_sampler.hpp:18-136; Sampler(U, J, R, "larsen_kolda"),
 * als_module.cpp:45-52): per factor the row leverage scores
 * u_i^T pinv(U^T U) u_i (compute_DAGAT, sampler.hpp:94-122; pinv with the
 * 1e-11 cut and target rank min(I_k, R)), and per draw every mode k != j
 * sampled independently in proportion to them, with weight
 * exp(-log J + sum_k (log total_k - log lev_k[row])).  The reference draws from
 * std::discrete_distribution on mt19937; here a row is the first whose
 * inclusive prefix sum exceeds u * total, u from draws[k J + i] when given
 * (the first N J of the 2 N J uniforms) or the built-in Philox stream.
 * ------------------------------------------------------------------------ */
typedef struct ftl_lk_sampler ftl_lk_sampler;

int ftl_lk_sampler_create(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, const double* const* U_host,
                          ftl_lk_sampler** out);
int ftl_lk_sampler_destroy(ftl_lk_sampler* s);
/* Hybrid variant -- replaces LarsenKoldaHybrid (larsen_kolda_hybrid.hpp:20-248;
 * Sampler(U, J, R, "larsen_kolda_hybrid"), als_module.cpp:45-50): tau = 1/J,
 * alpha_k = max lev_k / total_k; per draw the rows of mode k with
 * lev_k >= (tau alpha_k / prod_{k' != j} alpha_k') total_k are candidates
 * (ascending row order); the tuples of the candidate product (lowest mode's
 * digit fastest) whose probability p = prod lev_k / total_k reaches tau come
 * first, in that order, with weight 1; the other samples are drawn from the
 * product distribution until p < tau, weight (1 - p_det) / (J p), p_det the
 * kept tuples' mass.  Attempt 0 of sample i takes draws[k J + i] (or the
 * built-in stream at stage 2k), attempt a >= 1 the built-in stream at stage
 * 2k of call (call XOR a 2^40).  A sample whose 65536 attempts are all
 * rejected keeps its last tuple with weight 0 (the reference's loop never
 * ends, e.g. when the deterministic tuples hold all the mass).  Limits (the
 * reference has none): at most 2^32 candidate tuples (the reference skips the
 * deterministic part only from 2^62) and factor row counts below 2^31;
 * exceeding them fails the draw.  Draws with this handle synchronise its
 * stream. */
int ftl_lkh_sampler_create(uint32_t N, const uint64_t* I, uint64_t J, uint64_t R, const double* const* U_host,
                           ftl_lk_sampler** out);
/* Last draw of a hybrid handle: out[0] deterministic samples, out[1] their
 * probability mass, out[2] candidate tuples, out[3] 1 if the deterministic
 * part was skipped (2^62 candidates or more), out[4] samples given weight 0
 * after their rejection attempts ran out. */
int ftl_lk_sampler_info(ftl_lk_sampler* s, double* out5);
/* KRPDrawSamples / KRPDrawSamples_materialize: samples N x J (rows k != j),
 * h J x R and weights J optional (NULL), draws optional; returns when done. */
int ftl_lk_sampler_draw(ftl_lk_sampler* s, uint32_t j, uint64_t* samples, double* h, double* weights,
                        const double* draws);
/* Same on device buffers, enqueued on stream (NULL: the handle's).  A draw on
 * another stream waits for the last update_sampler (events, no host wait),
 * and the next update_sampler waits for the last such draw. */
int ftl_lk_sampler_draw_device(ftl_lk_sampler* s, uint32_t j, uint64_t* samples, double* h, double* weights,
                               const double* draws, void* stream);
/* update_sampler(j) (larsen_kolda_sampler.hpp:45-100) from a new factor j
 * (I_j x R, copied in), or with Uj_host NULL from the resident factor j (as
 * written through ftl_lk_sampler_factor_device).  Returns when done. */
int ftl_lk_sampler_update(ftl_lk_sampler* s, uint32_t j, const double* Uj_host);
/* Device address of the sampler's factor k (I_k x R FP64). */
int ftl_lk_sampler_factor_device(ftl_lk_sampler* s, uint32_t k, double** out);
/* Stream of the handle (hipStream_t). */
void* ftl_lk_sampler_stream(ftl_lk_sampler* s);
int ftl_lk_sampler_set_seed(ftl_lk_sampler* s, uint64_t seed, uint64_t counter);
/* Leverage scores of factor k (I_k) and their total (optional). */
int ftl_lk_sampler_leverage(ftl_lk_sampler* s, uint32_t k, double* lev_host, double* total_host);

/* ------------------------------------------------------------------------
 * Draw-request server (no counterpart in the reference, whose only surface is
 * one PySampler per factor set, als_module.cpp:30-82).  Requests for J
 * samples of the KRP of a factor set (N, I[k], R) are admitted FIFO by HBM
 * footprint into a device arena of hbm_budget_bytes, up to max_active at a
 * time, and retired as they finish.  A request either carries its own
 * factors, generated on the device (entries N(0,1) rounded to FP32 from a
 * Philox stream of its seed) and built on a stream of its own, or draws from
 * a registered factor set whose trees stay resident.  Every request whose
 * trees are complete is drawn in a packed launch with the others of the same
 * R and precision (R <= 64; one launch each of the two computeM kernels and
 * of the draw kernel).  The uniforms are the Philox stream of the request's
 * seed, so a result depends only on the request, alone or in any mix.
 * Not thread-safe: one host thread drives a server.
 * ------------------------------------------------------------------------ */
typedef struct ftl_server ftl_server;

int ftl_server_create(uint64_t hbm_budget_bytes, uint32_t max_active, ftl_server** out);
int ftl_server_destroy(ftl_server* s);

/* Queue a request; j as in KRPDrawSamples (j >= N: every mode).  FTL_ENOMEM
 * when the request alone exceeds the budget. */
int ftl_server_submit(ftl_server* s, uint32_t N, const uint64_t* I, uint64_t R, uint64_t J, uint32_t j,
                      uint64_t seed, int precision, uint64_t* id_out);

/* Register a caller's factor set: N factors U[k] (I[k] x R FP64, host or
 * device pointers per on_device) are copied into the arena and their Gram
 * trees built once (the builds are queued; the factors, host or device, are
 * read before the call returns and may be freed at once).
 * They stay resident until ftl_server_remove_factors; draws on the set ask
 * for at most J_max samples.  This is the server form of Sampler(matrices,
 * J, R, "efficient") (als_module.cpp:34-44).  FTL_ENOMEM when the set does
 * not fit next to what the arena holds. */
int ftl_server_add_factors(ftl_server* s, uint32_t N, const uint64_t* I, uint64_t R, const double* const* U,
                           int on_device, int precision, uint64_t J_max, uint64_t* fs_out);

/* Register the factors of an existing sampler handle (copied from its
 * device factors; J_max = its J, its precision). */
int ftl_server_add_sampler(ftl_server* s, ftl_sampler* sampler, uint64_t* fs_out);

/* update_sampler(k) of a registered set (efficient_krp_sampler.hpp:73-75):
 * after the set's queued draws, factor k is replaced by Uk and only tree k
 * is rebuilt; later draws see the new factor.  Uk is read before the call
 * returns (which therefore waits for the set's queued draws); the rebuild
 * stays queued. */
int ftl_server_update_factor(ftl_server* s, uint64_t fs, uint32_t k, const double* Uk, int on_device);

/* Release a registered set (deferred until its queued requests finish). */
int ftl_server_remove_factors(ftl_server* s, uint64_t fs);

/* Queue a draw of J <= J_max samples on a registered set, excluding mode j
 * (j >= N: every mode), uniforms from the Philox stream of seed.  Draws of
 * one set run one after another (they share its computeM buffers). */
int ftl_server_submit_on(ftl_server* s, uint64_t fs, uint64_t J, uint32_t j, uint64_t seed, uint64_t* id_out);

/* Drive the server until request id has finished (id 0: until any request
 * finishes that progress() has not reported yet) or timeout_s seconds
 * passed; *finished = 1 when it has.  The host sleeps between polls of the
 * device's completion events instead of spinning. */
int ftl_server_wait(ftl_server* s, uint64_t id, double timeout_s, int* finished);

/* Retire finished requests, admit waiting ones and launch the packed draws
 * of requests whose trees are complete.  done_ids receives up to capacity ids
 * of requests finished and not reported before (*n_done of them); the rest
 * are reported by later calls. */
int ftl_server_progress(ftl_server* s, uint64_t* done_ids, uint64_t capacity, uint64_t* n_done);

/* Samples (N x J uint64, row j zero when j < N) of a finished request, then
 * release it.  FTL_ENOMEM for a request that could not be placed. */
int ftl_server_result(ftl_server* s, uint64_t id, uint64_t* samples_host);

/* Factor k (I[k] x R FP64) of a finished, unreleased request (test surface). */
int ftl_server_factor(ftl_server* s, uint64_t id, uint32_t k, double* U_host);

/* Drop a request: waiting ones at once, running ones when their queued work
 * ends, finished ones with their result. */
int ftl_server_cancel(ftl_server* s, uint64_t id);

/* out[0..7] = active, waiting, finished-unreleased, arena bytes in use, peak
 * bytes in use, admitted, completed, cancelled. */
int ftl_server_stats(ftl_server* s, double* out8);

#ifdef __cplusplus
}
#endif

#endif /* FTL_HIP_H */

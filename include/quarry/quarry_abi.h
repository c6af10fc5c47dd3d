/* quarry_abi.h - C ABI of libquarry_hip.so, the MI355X (gfx950) BVH build and
 * query library.
 *
 * Plain C: opaque handles, POD structs, raw pointers and sizes; no HIP or torch
 * types.  Every entry point returns a qr_status; on failure qr_last_error()
 * (thread-local) holds the message.  The C++ API in <gpu/gpu_bvh.h> and the
 * Python module `quarry` are thin layers over these functions.
 *
 * Each declaration names the reference interface it replaces, as
 * src/gpu/gpu_bvh.h:<line> of ycl-berserk/quarry-gpu.
 *
 * Semantics shared by every call (reference: docs/api.md:146, :203):
 *   - calls are synchronous: results are ready on return;
 *   - returned milliseconds are device time measured with events;
 *   - work is issued on the stream set with qr_set_stream (default: null stream).
 *
 * Wire layouts (byte-identical to the reference):
 *   triangle  9 x f32 (a, b, c)                          36 B
 *   point     3 x f32                                    12 B
 *   Gaussian  mean 3 f32, quat xyzw 4 f32, scale 3 f32   40 B
 *   ray       origin 3 f32, dir 3 f32, tMin, tMax        32 B
 *   node      lo 3 f32, hi 3 f32, left i32, right i32    32 B (internal 0..n-2,
 *             leaf n-1+slot with left = -(slot+1), right = 1)
 *   closest/kNN result  {u32 dist2 bits, i32 index}       8 B (index -1: none)
 *   ray result u64 (hit << 32) | t bits, 0 for a miss
 *   radius CSR offsets i32[n+1], keys u64 (dist2 bits << 32 | index), ascending
 */
#ifndef QUARRY_ABI_H
#define QUARRY_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QR_API __attribute__((visibility("default")))

typedef struct qr_bvh qr_bvh;
typedef struct qr_rays qr_rays;
typedef struct qr_points qr_points;
typedef struct qr_csr qr_csr; /* radius-query CSR result */

typedef enum {
    QR_OK = 0,
    QR_ERR_INVALID = 1, /* bad argument (std::runtime_error in the reference API) */
    QR_ERR_DEVICE = 2,  /* HIP failure */
    QR_ERR_OOM = 3,     /* device allocation failed */
    QR_ERR_EDIT = 4,    /* invalid topology edit (std::invalid_argument) */
    QR_ERR_NO_DEVICE = 5
} qr_status;

typedef enum { QR_PRIM_TRIANGLE = 0, QR_PRIM_POINT = 1, QR_PRIM_GAUSSIAN = 2 } qr_prim;
typedef enum { QR_BUILDER_LBVH = 0, QR_BUILDER_SAH = 1 } qr_builder;
/* RayOrder / QueryOrder (src/gpu/gpu_bvh.h:224, :274) */
typedef enum { QR_ORDER_INPUT = 0, QR_ORDER_MORTON = 1, QR_ORDER_MORTON_DIRECTION = 2 } qr_order;

/* GpuBuildTimings (src/gpu/gpu_bvh.h:79-89) */
typedef struct {
    double morton_ms, sort_ms, tree_ms, scatter_ms, refit_ms;
    double sah_bin_ms, sah_split_ms, sah_part_ms, total_ms;
} qr_build_timings;

/* GpuRefitTimings (src/gpu/gpu_bvh.h:129-138) */
typedef struct {
    double upload_ms, gather_ms, tree_ms, mark_ms, merge_ms, wide_ms, total_ms;
    int updated;
} qr_refit_timings;

/* GpuTraversalStats / GpuQueryStats (src/gpu/gpu_bvh.h:240-245, :264-269) */
typedef struct {
    uint64_t node_tests, leaf_tests, stack_overflows;
    double sort_ms;
    double kernel_ms;      /* traversal kernel(s) alone (radius: count + fill passes) */
    uint64_t nodes_opened; /* interior nodes whose children were loaded (binary or 4-wide) */
} qr_traversal_stats;

/* RayStats (src/core/fingerprint.h:16-43) */
typedef struct {
    uint64_t rays, hits, hash;
    double sum_t;
} qr_ray_stats;

/* Public fields of GpuBVH (src/gpu/gpu_bvh.h:26-33) plus the device arrays. */
typedef struct {
    int prim_count, node_count, depth, wide_count, prim_kind, builder;
    void* d_nodes;      /* node[node_count] */
    void* d_items;      /* primitives in leaf order */
    void* d_prim_ids;   /* i32[prim_count], leaf slot -> scene index */
    void* d_wide_nodes; /* 128 B 4-wide nodes, or NULL */
} qr_bvh_info;

QR_API const char* qr_last_error(void);

/* gpuDescribe (src/gpu/gpu_bvh.h:100): device name into buf; QR_ERR_NO_DEVICE if none. */
QR_API int qr_describe(char* buf, int len);
/* warmupBuildGPU / warmupQueryGPU (src/gpu/gpu_bvh.h:122, :343) */
QR_API int qr_warmup(void);
/* Shape of the 4-wide traversal, process-wide: 0 automatic (closest / kNN below
 * 65,536 queries by query count, leaf kind and k; from 65,536 on, the first call
 * for a tree, k class and query-count octave times one lane, four lanes, 16 and 32
 * lanes and one wave per query on its own queries and keeps the fastest; radius 32
 * lanes per query (two queries per wave) below 131,072 queries, four lanes above;
 * rays one lane per ray), 1 one lane per query, 2 four lanes per query, 3 one wave
 * per query or ray, 4 as 0 but closest / kNN and rays tuned at every count, 5 / 6
 * closest and kNN with 16 / 32 lanes per query (four / two queries per wave; kNN
 * with k > 16 takes 32 lanes), radius as in mode 0 (5) or 32 lanes per query at
 * every count (6), other queries as in mode 0.  Results are identical in every
 * mode.  No reference counterpart (a tuning knob of this backend). */
QR_API int qr_set_traversal_mode(int mode);
/* LBVH build: how many frontier nodes a workgroup (or group of workgroups) may hand
 * up to the next level, process-wide, 0..32 (default 32).  Units with more send
 * their nodes through the device-wide arrival protocol instead; 0 sends every node
 * that way.  Trees are identical byte for byte at every setting.  No reference
 * counterpart (a tuning knob of this backend; tests use it to cover both paths). */
QR_API int qr_set_frontier_cap(int cap);
/* Which events a build records for its qr_build_timings, process-wide: 1 (the
 * default, as the reference fills them) the phase fields and total_ms; 0 total_ms
 * only, the phase fields left 0 (each event between two kernels costs ~12 us of idle
 * device time: a 2M-triangle LBVH build runs ~24 us faster); 2 also the SAH
 * per-level events behind sah_bin_ms, sah_split_ms and sah_part_ms (about 0.2 ms of
 * a 2M-triangle SAH build).  No reference counterpart (a profiling knob of this
 * backend). */
QR_API int qr_set_build_phase_timing(int on);
/* Stream for subsequent calls on this host thread (a hipStream_t; NULL = default). */
QR_API int qr_set_stream(void* hip_stream);

/* Device memory helpers for callers without their own HIP runtime use
 * (benchmarks, C consumers of the *_device entry points).  No reference
 * counterpart: the reference API only takes host arrays. */
QR_API int qr_dev_alloc(size_t bytes, void** out);
QR_API int qr_dev_free(void* ptr);
QR_API int qr_dev_upload(void* dst, const void* src_host, size_t bytes);
QR_API int qr_dev_download(void* dst_host, const void* src, size_t bytes);
QR_API int qr_dev_synchronize(void);
/* Device-to-device copy through a streaming kernel (16-byte aligned pointers);
   ms = kernel time.  No counterpart in the reference; used to calibrate HBM
   traffic counters against a copy of known size. */
QR_API int qr_dev_copy(void* dst, const void* src, size_t bytes, double* ms);

/* ---- build (src/gpu/gpu_bvh.h:103-118) ----------------------------------------
 * buildBVHGPU / buildPointBVHGPU / buildGaussianBVHGPU (builder LBVH) and
 * buildBVHSahGPU / buildPointBVHSahGPU / buildGaussianBVHSahGPU (builder SAH).
 * items_host: n primitives of kind `prim`, scene order.  n >= 2. */
QR_API int qr_bvh_build(int prim, int builder, const void* items_host, int n,
                        qr_build_timings* timings, qr_bvh** out);
/* Same with the primitives already in device memory (not retained). */
QR_API int qr_bvh_build_device(int prim, int builder, const void* items_dev, int n,
                               qr_build_timings* timings, qr_bvh** out);
/* destroyBVHGPU (src/gpu/gpu_bvh.h:215) */
QR_API int qr_bvh_destroy(qr_bvh* bvh);
QR_API int qr_bvh_get_info(const qr_bvh* bvh, qr_bvh_info* out);
/* buildWideNodesGPU (src/gpu/gpu_bvh.h:229) */
QR_API int qr_bvh_build_wide(qr_bvh* bvh, double* ms);

/* refit{,Point,Gaussian}BVHGPU (src/gpu/gpu_bvh.h:141-155).  items_host holds all n
 * primitives in scene order (topology unchanged).  dirty == NULL: full refit;
 * otherwise only those scene indices are re-uploaded and only their ancestors
 * are merged.  Wide nodes, if built, are re-emitted. */
QR_API int qr_bvh_refit(qr_bvh* bvh, const void* items_host, int n, const int* dirty, int n_dirty,
                        qr_refit_timings* timings, double* ms);
/* Full refit from device-resident scene-order primitives. */
QR_API int qr_bvh_refit_device(qr_bvh* bvh, const void* items_dev, int n,
                               qr_refit_timings* timings, double* ms);

/* GpuTopoUpdateTimings (src/gpu/gpu_bvh.h:190-206).  The fields keep the reference's
 * names, but an edit that changes the primitive count is not a subtree surgery here:
 * the new scene is composed on the device from the tree's items and the uploaded
 * ones, and the whole tree is rebuilt with its builder.  Such an update reports
 * full_rebuild = 1, surgery_roots = 1 (the root), rebuilt_leaves = n, plan_ms = the
 * composition, surgery_ms = the rebuild, relayout_ms = merge_ms = 0.  A pure-move
 * edit takes the dirty refit (refit_ms) and reports full_rebuild = 0. */
typedef struct {
    double upload_ms, plan_ms, surgery_ms, relayout_ms, merge_ms, wide_ms, total_ms;
    int deleted, moved, inserted, appended, surgery_roots, rebuilt_leaves, new_depth;
    int full_rebuild;
    double refit_ms;
} qr_topo_timings;

/* update{,Point,Gaussian}BVHGPU (src/gpu/gpu_bvh.h:208-213; edit script at :182-187).
 * items_host: the whole new scene (n primitives).  deleted: old indices, ascending,
 * unique; moved: new indices whose geometry changed; inserted: new indices,
 * ascending, unique, each < n - appended; appended: count at the end.
 * n must equal prim_count - n_deleted + n_inserted + appended.  An invalid edit
 * returns QR_ERR_EDIT.  Queries afterwards equal those on a fresh build of the
 * new scene bit for bit. */
QR_API int qr_bvh_update(qr_bvh* bvh, const void* items_host, int n, const int* deleted,
                         int n_deleted, const int* moved, int n_moved, const int* inserted,
                         int n_inserted, int appended, qr_topo_timings* timings, double* ms);

/* validate{,Point,Gaussian}BVHGPU (src/gpu/gpu_bvh.h:347-349): why receives "" when
 * every structural invariant holds. */
QR_API int qr_bvh_validate(qr_bvh* bvh, const void* items_host, int n, char* why, int len);
/* treeHashGPU (src/gpu/gpu_bvh.h:351-352): FNV-1a over the node array bytes. */
QR_API int qr_bvh_tree_hash(const qr_bvh* bvh, uint64_t* out);
/* Copies the node array (node_count x 32 B) and/or prim ids (prim_count) to the host. */
QR_API int qr_bvh_read(const qr_bvh* bvh, void* nodes_host, int* prim_ids_host);

/* ---- rays (src/gpu/gpu_bvh.h:217-246) ------------------------------------------ */
/* uploadRaysGPU */
QR_API int qr_rays_upload(const void* rays_host, int n, qr_rays** out);
/* destroyRayBufferGPU */
QR_API int qr_rays_destroy(qr_rays* rays);
/* raycastGPU: closest hit per ray into the buffer's device results. */
QR_API int qr_raycast(const qr_bvh* bvh, qr_rays* rays, int collect_stats, int order, int wide,
                      double* ms);
/* summarizeGPU: fingerprint of the per-ray results; per_ray (optional) gets the u64s. */
QR_API int qr_rays_summarize(const qr_rays* rays, uint64_t* per_ray, qr_ray_stats* out);
/* lastTraversalStatsGPU */
QR_API int qr_last_traversal_stats(qr_traversal_stats* out);
/* GpuRayBuffer::d_rays / d_results / count (src/gpu/gpu_bvh.h:91-96). */
QR_API int qr_rays_device_ptrs(const qr_rays* rays, void** rays_dev, uint64_t** results_dev,
                               int* count);
/* Device-pointer variant: rays_dev[n] -> results_dev[n] (u64). */
QR_API int qr_raycast_device(const qr_bvh* bvh, const void* rays_dev, int n, uint64_t* results_dev,
                             int order, int wide, double* ms);

/* ---- point queries (src/gpu/gpu_bvh.h:276-342) ---------------------------------- */
/* uploadPointsGPU: results_per_point = result slots per query (closest 1, kNN k). */
QR_API int qr_points_upload(const float* points_host, int n, int results_per_point,
                            qr_points** out);
/* destroyPointBufferGPU */
QR_API int qr_points_destroy(qr_points* pts);
/* closestPointGPU / closestVertexGPU / closestGaussianGPU (leaf kind from the tree) */
QR_API int qr_closest(const qr_bvh* bvh, qr_points* pts, int collect_stats, int order, int wide,
                      double* ms);
/* kNearestGPU / kNearestVertexGPU / kNearestGaussianGPU; 1 <= k <= 32. */
QR_API int qr_knn(const qr_bvh* bvh, qr_points* pts, int k, int collect_stats, int order, int wide,
                  double* ms);
/* radiusQueryGPU / radiusQueryVertexGPU / radiusQueryGaussianGPU; *out is a new CSR. */
QR_API int qr_radius(const qr_bvh* bvh, qr_points* pts, float radius, int collect_stats, int order,
                     int wide, qr_csr** out, double* ms);
/* readQueryResultsGPU: count * stride records. */
QR_API int qr_points_read(const qr_points* pts, void* results_host);
QR_API int qr_points_info(const qr_points* pts, int* count, int* stride);
/* readRadiusResultsGPU: offsets_host[count + 1], keys_host[total]. */
QR_API int qr_radius_info(const qr_csr* r, int* count, int* total);
QR_API int qr_radius_read(const qr_csr* r, int* offsets_host, uint64_t* keys_host);
/* GpuPointBuffer::d_points / d_results and GpuRadiusBuffer::d_offsets / d_keys
 * (src/gpu/gpu_bvh.h:255-262, :296-301). */
QR_API int qr_points_device_ptrs(const qr_points* pts, float** points_dev, void** results_dev);
QR_API int qr_radius_device_ptrs(const qr_csr* r, int** offsets_dev, uint64_t** keys_dev);
/* destroyRadiusBufferGPU */
QR_API int qr_radius_destroy(qr_csr* r);
/* lastQueryStatsGPU */
QR_API int qr_last_query_stats(qr_traversal_stats* out);
/* Work of the calling thread's last binned-SAH build (no reference counterpart: the
 * measurement behind the SAH build's bandwidth figure): split-loop levels, primitive
 * slots binned and partitioned summed over the levels, and slots finished by the
 * serial tails (tasks of at most 64 primitives). */
typedef struct qr_sah_stats {
    int levels;
    long long binned_slots, partitioned_slots, serial_slots;
} qr_sah_stats;
QR_API int qr_last_sah_stats(qr_sah_stats* out);
/* Device-pointer variants: points_dev[n] (3 f32) -> results_dev[n * k] (8 B each). */
QR_API int qr_closest_device(const qr_bvh* bvh, const float* points_dev, int n, void* results_dev,
                             int order, int wide, double* ms);
QR_API int qr_knn_device(const qr_bvh* bvh, const float* points_dev, int n, int k,
                         void* results_dev, int order, int wide, double* ms);
/* qr_knn_device without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (tight serving loops, code that
 * must not call hipMalloc).  The Morton-order scratch lives in a fixed device array
 * of about 23 MB per GPU, held until the process ends; n <= QR_NOALLOC_MAX_QUERIES
 * (larger n returns QR_ERR_INVALID).  Synchronous, no timing.  Calls that overlap on
 * one GPU share that array and race; results equal qr_knn_device's. */
#define QR_NOALLOC_MAX_QUERIES 1048576
QR_API int qr_knn_device_noalloc(const qr_bvh* bvh, const float* points_dev, int n, int k,
                                 void* results_dev, int order, int wide);

/* ---- batched mode ----------------------------------------------------------------
 * No reference counterpart (the reference builds and queries one scene per call,
 * src/gpu/gpu_bvh.h:103-118, :276-342).  Thousands of independent scenes of one
 * primitive kind are built with one segmented launch per stage, and queried with a
 * persistent work-queue kernel replayed from a captured hipGraph.  Every tree and
 * every result equals the plain per-scene path's (scene-local indices). */
typedef struct qr_batch qr_batch;
typedef struct qr_batch_plan qr_batch_plan;
/* items: scene-major primitives; offsets_host[scenes + 1] (0 first, >= 2 per scene). */
QR_API int qr_batch_build(int prim, const void* items_host, const int* offsets_host, int scenes,
                          qr_build_timings* timings, qr_batch** out);
QR_API int qr_batch_build_device(int prim, const void* items_dev, const int* offsets_host, int scenes,
                                 qr_build_timings* timings, qr_batch** out);
QR_API int qr_batch_destroy(qr_batch* b);
QR_API int qr_batch_info(const qr_batch* b, int* scenes, int* total_prims, int* max_depth);
/* One scene's node array (2 n_s - 1 nodes, local ids) and leaf slot -> local index. */
QR_API int qr_batch_read_scene(const qr_batch* b, int scene, void* nodes_host, int* prim_ids_host);
/* FNV-1a over one scene's node bytes: equals qr_bvh_tree_hash of that scene built alone. */
QR_API int qr_batch_tree_hash(const qr_batch* b, int scene, uint64_t* out);
/* Query plan over device points (scene-major, qoffsets_host[scenes + 1]); k = 0: closest,
 * 1..32: kNN.  Kernel variants v: 0..3 walk the binary nodes, 4..7 the 4-wide nodes
 * (collapsed for every scene on first use), each with the traversal stack in LDS with
 * 256 or 64 lanes per workgroup (v % 4 = 0, 1), in scratch (2), or its top 8 entries in
 * registers (3).  tune = 1 times every variant with and without per-scene Morton order
 * and keeps the fastest; tune = 0 takes variant 0 with Morton order; tune = 2 + 2 v + m
 * forces variant v with Morton order m.  The chosen pipeline is captured in a hipGraph.
 * The points must stay valid while the plan lives. */
QR_API int qr_batch_plan_create(qr_batch* b, const float* points_dev, const int* qoffsets_host, int k,
                                int tune, qr_batch_plan** out);
/* Replays the graph; results stay in the plan (n * max(k, 1) records, scene-local ids). */
QR_API int qr_batch_plan_run(qr_batch_plan* p, double* ms);
/* The same pipeline launched directly on the plan's stream (no graph). */
QR_API int qr_batch_plan_run_eager(qr_batch_plan* p, double* ms);
QR_API int qr_batch_plan_read(const qr_batch_plan* p, void* results_host);
/* variant: 0 LDS stack / 256 lanes, 1 LDS stack / 64 lanes, 2 scratch stack / 64 lanes. */
QR_API int qr_batch_plan_config(const qr_batch_plan* p, int* variant, int* morton, double* tuned_ms);
QR_API int qr_batch_plan_destroy(qr_batch_plan* p);

/* ---- differentiable mode ------------------------------------------------------------
 * No reference counterpart: the reference is forward-only and its ray kernel
 * discards the hit index (src/gpu/raycast.cu:108-111).  All pointers are device
 * pointers; work runs on the stream set with qr_set_stream.  Forward results are
 * the plain entry points' results bit for bit. */
/* Ray cast that also writes the hit triangle's scene index (-1: miss) per ray. */
QR_API int qr_raycast_save_device(const qr_bvh* bvh, const void* rays_dev, int n, uint64_t* results_dev,
                                  int* hit_prim_dev, int order, int wide, double* ms);
/* Saved word per closest result (from qr_closest_device): triangle -> closest-feature
 * region 0..9 (vertex A, B, edge AB, vertex C, edge AC, edge BC, face, degenerate
 * A/B/C), Gaussian -> bits of the Eberly root (-2: inside), point -> 0; -1: no result. */
QR_API int qr_closest_save_aux(qr_bvh* bvh, const float* points_dev, int n, const void* results_dev,
                               int* aux_dev);
/* Adjoint of dist2: grad_points (n x 3, overwritten, may be NULL) and grad_items
 * (prim_count x 9 / 3 / 10 floats in scene order, accumulated, may be NULL;
 * Gaussians: mean 3, quaternion 4, scale 3) from grad_dist2 (n). */
QR_API int qr_adjoint_closest(qr_bvh* bvh, const float* points_dev, int n, const void* results_dev,
                              const int* aux_dev, const float* grad_dist2_dev, float* grad_points_dev,
                              float* grad_items_dev);
/* Adjoint of hit t: grad_origins, grad_dirs (n x 3, overwritten, may be NULL),
 * grad_tris (prim_count x 9, accumulated, may be NULL) from grad_t (n). */
QR_API int qr_adjoint_ray(qr_bvh* bvh, const void* rays_dev, int n, const uint64_t* results_dev,
                          const int* hit_prim_dev, const float* grad_t_dev, float* grad_origins_dev,
                          float* grad_dirs_dev, float* grad_tris_dev);

#ifdef __cplusplus
}
#endif

#endif /* QUARRY_ABI_H */

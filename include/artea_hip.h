/*
 * artea_hip.h -- C-ABI of the MI355X (gfx950) ARTEA search and refine path.
 *
 * One shared library, artea_amd/libartea_hip.so.  Every entry point is
 * extern "C", takes plain pointers and sizes, returns an int status
 * (ARTEA_OK = 0) and leaves a thread-local message in artea_last_error() on
 * failure.  Handles are not thread-safe; each handle owns one HIP stream
 * created with the default priority.  Reference citations are relative to
 * include/artea/cpu/ of the reference checkout.
 */
#ifndef ARTEA_HIP_H
#define ARTEA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum artea_status {
    ARTEA_OK = 0,
    ARTEA_EINVAL = 1,   /* std::runtime_error / std::invalid_argument in the reference */
    ARTEA_EIO = 2,      /* HGRA snapshot/restore failures (ARTEA_ERROR -> runtime_error) */
    ARTEA_EHIP = 3,     /* a HIP runtime call failed */
    ARTEA_ENOMEM = 4,   /* device or host allocation failed */
    ARTEA_EOVERFLOW = 5 /* std::overflow_error in the reference */
};

/* DistanceMetricsT (framework/type_traits/computer_traits.hpp:38-43). */
enum artea_metric { ARTEA_L2SQR = 0, ARTEA_DOT = 1, ARTEA_COSINE = 2, ARTEA_L2 = 3 };

/* Query flags: template switches of HierarchicalGraphRouter::batch_query
 * (router/hierarchical_graph_router.hpp:293-319) and batch_query_l0_only
 * (:353-385). */
enum artea_query_flags { ARTEA_EARLY_STOP = 1u, ARTEA_L0_ONLY = 2u };

#define ARTEA_INVALID_VID 0x7FFFFFFFu   /* CandidateEntry::make_invalid_entry vid */
#define ARTEA_INVALID_LEVEL 0xFFFFFFFFu /* compact::HierarchicalGraph::invalid_level_id */

/* Thread-local message of the last failing call (ARTEA_ERROR's text). */
const char* artea_last_error(void);
/* Library build id string (compiler, offload arch). */
const char* artea_version(void);

/* ------------------------------------------------------------------ graph */
/* Host-side compact::hierarchical_graph_t (index/compact_structure/
 * hierarchical_graph.hpp:33-201) and its HGRA v1 persistence
 * (persistence/hierarchical_graph_file_manager.hpp:20-40).  vertex_info[2v]
 * = highest level (ARTEA_INVALID_LEVEL = unassigned), vertex_info[2v+1] =
 * local id in its group; group h holds count_h*(h+1)+1 uint64 offsets into
 * the neighbour array, row index local*(h+1)+(h-level) (:99-108). */
typedef struct artea_graph artea_graph;
/* Builder path (the builder-only mutators, :135-162): copies the arrays. */
int artea_graph_from_csr(uint32_t n, uint32_t top, uint32_t ul_max, uint32_t bl_max, uint32_t entry_vid,
                         const uint32_t* vertex_info, const uint64_t* const* group_offsets, const uint32_t* nbrs,
                         uint64_t n_nbrs, artea_graph** out);
/* HierarchicalGraphFileManager::restore (:152-324) / snapshot (:88-145):
 * same validation order and messages; ARTEA_EIO on failure. */
int artea_graph_restore(const char* path, artea_graph** out);
int artea_graph_snapshot(const artea_graph* g, const char* path);
/* get_num_vertices / top_occupied_level_id / ul,bl_max_nbr_size /
 * entry_point_vid (:83-132); top is ARTEA_INVALID_LEVEL for an empty graph. */
int artea_graph_info(const artea_graph* g, uint32_t* n, uint32_t* top, uint32_t* ul_max, uint32_t* bl_max,
                     uint32_t* entry_vid, uint64_t* n_nbrs);
/* Views valid until the graph is destroyed.  group_offsets: *count = the
 * group's vertices; the array holds count*(h+1)+1 offsets. */
const uint32_t* artea_graph_vertex_info(const artea_graph* g);
const uint64_t* artea_graph_group_offsets(const artea_graph* g, uint32_t h, uint64_t* count);
const uint32_t* artea_graph_nbrs(const artea_graph* g);
/* fetch_level_nbrs (:99-108); count 0 when level > highest. */
const uint32_t* artea_graph_level_nbrs(const artea_graph* g, uint32_t vid, uint32_t level, uint32_t* count);
int artea_graph_destroy(artea_graph* g);

/* ------------------------------------------------------------------ index */
typedef struct artea_index artea_index;

/* Device copy of the base vectors: vector_array_t rows of dim_pad floats
 * (containers/vector_array.hpp:26-506; dim_pad = the SIMD-padded dimension,
 * a multiple of 16, containers/vector_dataset.hpp:242-269).  `vecs` is a
 * host pointer; artea_index_create_device takes a device pointer instead
 * (copied device-to-device).  metric is the search metric of dist_func_t
 * (framework/type_context/default_context.hpp:147). */
int artea_index_create(const float* vecs, uint32_t n, uint32_t dim_pad, int metric, artea_index** out);
int artea_index_create_device(const float* d_vecs, uint32_t n, uint32_t dim_pad, int metric,
                              artea_index** out);

/* Attach a compact::hierarchical_graph_t given as its grouped CSR
 * (index/compact_structure/hierarchical_graph.hpp:33-201): vertex_info[2v] =
 * highest level (ARTEA_INVALID_LEVEL = unassigned), vertex_info[2v+1] =
 * local id in its group; group_offsets[h] (h = 0..top) holds
 * count_h*(h+1)+1 uint64 offsets into nbrs, row index local*(h+1)+(h-level)
 * (:99-108).  The graph is converted once into the device layout.  n must
 * equal the index's vector count. */
int artea_index_set_graph_csr(artea_index* idx, uint32_t top, uint32_t ul_max, uint32_t bl_max,
                              uint32_t entry_vid, const uint32_t* vertex_info,
                              const uint64_t* const* group_offsets, const uint32_t* nbrs,
                              uint64_t n_nbrs);
/* Attach a host graph (copied; the graph may be destroyed afterwards). */
int artea_index_set_graph(artea_index* idx, const artea_graph* g);
/* HierarchicalGraphFileManager::restore (persistence/
 * hierarchical_graph_file_manager.hpp:152-324) straight into the device
 * layout, with the same validation and error texts: the file's body is
 * copied to the device and parsed there (records found by a segmented parse
 * that the host verifies segment by segment, ids checked and rows written
 * into the ELL arrays by one wavefront per vertex). */
int artea_index_set_graph_hgra(artea_index* idx, const char* path);
/* The same with the segmented parse's settings (no counterpart in the
 * reference): segment_words = words per segment, trial_records = records a
 * guessed segment start must parse before it is taken (0 and 0 = the
 * defaults, 4096 and 4).  Every guess is verified against the previous
 * segment's end, so the result does not depend on them. */
int artea_index_set_graph_hgra_segmented(artea_index* idx, const char* path, uint32_t segment_words,
                                         uint32_t trial_records);
/* Copy of the attached graph as a host graph (snapshot, export). */
int artea_index_get_graph(const artea_index* idx, artea_graph** out);
/* HierarchicalGraphFileManager::snapshot (:88-145) of the attached graph,
 * laid out on the device (record lengths, their prefix sum, one wavefront
 * per record) and streamed to the file. */
int artea_index_snapshot_hgra(const artea_index* idx, const char* path);
/* The attached graph's device layout: ELL row strides (words) of level 0
 * and of the upper levels, and the number of upper-level rows.  No
 * counterpart in the reference (its layout is the grouped CSR). */
int artea_index_layout(const artea_index* idx, uint32_t* stride0, uint32_t* stride_upper, uint64_t* rows_upper);
/* Host-clock phases of the index's last device HGRA restore (load3: file
 * read and copy to the device, segmented parse and its walk, id checks and
 * ELL fill) and snapshot (store3: record layout on the device, copy back
 * and file write, 0), in seconds.  No counterpart in the reference. */
int artea_index_hgra_seconds(const artea_index* idx, double* load3, double* store3);
/* top_occupied_level_id / entry_point_vid / get_num_vertices (:83-132). */
int artea_index_info(const artea_index* idx, uint32_t* n, uint32_t* dim_pad, uint32_t* top,
                     uint32_t* entry_vid, uint32_t* ul_max, uint32_t* bl_max);
/* Device pointer of the base vectors (n x dim_pad floats). */
const float* artea_index_device_vectors(const artea_index* idx);
/* 8-bit rows (no counterpart in the reference; a storage choice of this
 * library).  When every base value is a whole number in [0, 255] and
 * dim_pad <= 256, the index also keeps a one-byte copy of each row, built
 * when the index is created.  The router then computes the distances of
 * every query whose values are of the same kind from the one-byte rows, in
 * exact integer arithmetic: for such values every FP32 operation of
 * dist_func_t is exact, so the distances, and hence the results, equal the
 * FP32 chains' bit for bit, from a quarter of the bytes.  Other queries of
 * the same batch use the FP32 rows.  enable = 0 drops the copy, enable = 1
 * builds it again when the values allow; *active (optional) tells whether
 * the index has it.  artea_index_u8_rows returns 1 when it has, else 0.
 * Launch graphs (artea_batched_router_capture) record the rows in use when
 * they are captured: capture again after dropping or rebuilding the copy. */
int artea_index_set_u8_rows(artea_index* idx, int enable, int* active);
int artea_index_u8_rows(const artea_index* idx);
int artea_index_destroy(artea_index* idx);

/* ----------------------------------------------------------------- router */
typedef struct artea_router artea_router;

/* HierarchicalGraphRouter ctor (router/hierarchical_graph_router.hpp:93-109):
 * fails with ARTEA_EINVAL "candidate_queue_size (Q) must be >= topk (K)". */
int artea_router_create(artea_index* idx, uint32_t topk, uint32_t queue, artea_router** out);
int artea_router_destroy(artea_router* r);

/* batch_query / batch_query_l0_only (:293-385), synchronous, host buffers:
 * queries nq x dim_pad floats; results row-major nq x topk, ascending by
 * (distance, vid), padded with (ARTEA_INVALID_VID, FLT_MAX).  thresholds
 * (read only with ARTEA_EARLY_STOP) holds one value per level, index =
 * level, already in query distance units (squared for L2SQR), as
 * _early_stop_thresholds produces them (:388-408). */
int artea_router_batch_query(artea_router* r, const float* queries, uint32_t nq, uint32_t flags,
                             const float* thresholds, uint32_t* out_vids, float* out_dists);

/* Same, with device buffers, enqueued on `stream` (a hipStream_t; NULL =
 * the router's own stream) and returning without a synchronisation.
 * Launches of one router (plain or batched, eager or a launch-graph replay)
 * share its counters, so they run one after another whichever streams they
 * are enqueued on: a launch on another stream than the router's previous
 * launch waits for that launch first.  Launches that should overlap need a
 * router each. */
int artea_router_batch_query_device(artea_router* r, const float* d_queries, uint32_t nq, uint32_t flags,
                                    const float* thresholds, uint32_t* d_out_vids, float* d_out_dists,
                                    void* stream);

/* Counters of the last launch on this router (device-counted with the
 * reference profiler's definitions, router/profiler/hg_router_profiler.hpp
 * :153-197): distance evaluations (NDC, entry included), beam expansions,
 * and tie-tail overflows (0 = the candidate list held every tied evictee).
 * Waits for that launch to complete. */
int artea_router_last_stats(artea_router* r, uint64_t* ndc, uint64_t* expansions, uint64_t* tail_overflows);

/* Queries of the last launch whose candidate list overflowed its tie tail
 * (no reference counterpart: the reference's heaps are unbounded).  Entries
 * evicted from the top-L set stay poppable in the reference only while their
 * distance equals the bound; the device list keeps L plus at least 16 such
 * entries, and a query that had to drop one more may return a result that
 * differs from the reference's.  *count = the number of such queries (0 in
 * every workload measured); qidx receives up to min(max, 255) of their batch
 * indices in ascending order.  artea_server_* re-runs such requests on a
 * larger list (artea_server_overflows).  Waits for the launch. */
int artea_router_last_overflows(artea_router* r, uint32_t max, uint32_t* qidx, uint32_t* count);

/* --------------------------------------------------------- batched router */
/* Persistent fused router: queries of mixed topk / queue size are pulled
 * from a device work queue.  topk and queue are per query (device arrays of
 * nq uint32); results use row stride max_topk.  No reference counterpart:
 * it is the batched form of batch_query for the same graph and dist_func_t,
 * and returns for each query exactly what batch_query returns for it. */
typedef struct artea_batched_router artea_batched_router;
int artea_batched_router_create(artea_index* idx, uint32_t max_topk, uint32_t max_queue,
                                artea_batched_router** out);
int artea_batched_router_destroy(artea_batched_router* r);
int artea_batched_router_query_device(artea_batched_router* r, const float* d_queries, uint32_t nq,
                                      const uint32_t* d_topk, const uint32_t* d_queue, uint32_t flags,
                                      const float* thresholds, uint32_t* d_out_vids, float* d_out_dists,
                                      void* stream);
int artea_batched_router_last_stats(artea_batched_router* r, uint64_t* ndc, uint64_t* expansions,
                                    uint64_t* tail_overflows);
int artea_batched_router_last_overflows(artea_batched_router* r, uint32_t max, uint32_t* qidx, uint32_t* count);

/* Allocation-free form of artea_batched_router_query_device.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It takes the index and the launch's
 * maxima instead of a router handle (max_topk <= max_queue <= 1008, as
 * artea_batched_router_create), runs the built-in default kernel variant,
 * and keeps the launch's work counter, counters and early-stop thresholds
 * (one per level, at most 64 levels) in a fixed 2 KiB device array of the
 * library, one per GPU, held until the process ends: it calls no hipMalloc,
 * hipFree or other allocating routine.  The caller passes the inputs and
 * outputs as for artea_batched_router_query_device; the results are the
 * same.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU race: enqueue them on one stream, or order them. */
int artea_batched_query_noalloc(artea_index* idx, uint32_t max_topk, uint32_t max_queue, const float* d_queries,
                                uint32_t nq, const uint32_t* d_topk, const uint32_t* d_queue, uint32_t flags,
                                const float* thresholds, uint32_t* d_out_vids, float* d_out_dists, void* stream);

/* Kernel variants of the batched router (no reference counterpart).  Slots
 * 0..3 read the FP32 rows: the distance gather width (16 or 32 candidate
 * rows per pass) and the register budget the kernel is compiled for (no
 * bound, 3 or 4 resident waves per SIMD).  While the index keeps one-byte
 * rows (artea_index_set_u8_rows) slots 4..7 exist too: the integer distance
 * path with 64, 32 or 16 rows per pass, and 64 under a 4-wave budget.
 * Every variant returns the same results; only the speed differs. */
#define ARTEA_ROUTER_VARIANTS 4
#define ARTEA_ROUTER_VARIANTS_U8 8
/* Autotuner: times each variant on this batch (one untimed and one timed
 * launch each, on the router's stream, synchronously) and keeps the fastest
 * for the router's later queries.  ms (optional) receives one launch time in
 * milliseconds per slot (ARTEA_ROUTER_VARIANTS_U8 of them when the index
 * keeps one-byte rows, else ARTEA_ROUTER_VARIANTS), *chosen (optional) the
 * slot kept.  The output buffers hold the batch's results after. */
int artea_batched_router_autotune(artea_batched_router* r, const float* d_queries, uint32_t nq,
                                  const uint32_t* d_topk, const uint32_t* d_queue, uint32_t flags,
                                  const float* thresholds, uint32_t* d_out_vids, float* d_out_dists, float* ms,
                                  int* chosen);
/* Pins a variant slot (0 .. 3, or 0 .. 7 with one-byte rows), or -1 for the
 * built-in default for the router's queue size; get returns the slot in use
 * (a one-byte slot reads as its FP32 counterpart once the rows are dropped). */
int artea_batched_router_set_variant(artea_batched_router* r, int variant);
int artea_batched_router_get_variant(artea_batched_router* r, int* variant);

/* HIP-graph form of artea_batched_router_query_device over fixed device
 * buffers (no reference counterpart: the reference has no launch graphs).
 * Capture runs the query once eagerly, then records the launch sequence
 * (counter-reset kernel + persistent router kernel, kernel nodes only) on
 * the router's stream; every replay re-runs the query on the buffers'
 * current contents and refreshes the router's last_stats.  Thresholds are
 * copied at capture time into a buffer the graph owns.  Replays are ordered
 * with the router's other launches (see artea_router_batch_query_device);
 * destroy the graph before its router. */
typedef struct artea_launch_graph artea_launch_graph;
int artea_batched_router_capture(artea_batched_router* r, const float* d_queries, uint32_t nq,
                                 const uint32_t* d_topk, const uint32_t* d_queue, uint32_t flags,
                                 const float* thresholds, uint32_t* d_out_vids, float* d_out_dists,
                                 artea_launch_graph** out);
int artea_launch_graph_launch(artea_launch_graph* g, void* stream);
int artea_launch_graph_destroy(artea_launch_graph* g);

/* ----------------------------------------------------------------- refine */
/* PropagateEngine over one dynamic::SingleLayerGraph
 * (refiner/propagate_engine.hpp:38-325, index/dynamic_structure/
 * single_layer_graph.hpp:29-227).  Rows are indexed by local id; entries
 * are Neighbor records (index/neighbor.hpp:43-67) packed as two uint32:
 * vid:31 | new:1, then the FP32 distance bits.  cap = the layer's
 * max_nbr_size (at most 128).  d_vecs is a device pointer (e.g.
 * artea_index_device_vectors); local_to_global is a host array of n_local
 * global ids, or NULL for the dense identity layer (n_local = n_global).
 * seed keys the prefill's counter-based random stream. */
typedef struct artea_refine artea_refine;
int artea_refine_create(const float* d_vecs, uint32_t n_global, uint32_t dim_pad, int metric, uint32_t cap,
                        const uint32_t* local_to_global, uint32_t n_local, uint64_t seed, artea_refine** out);
int artea_refine_destroy(artea_refine* r);
/* fill / export of rows: counts[n_local] and the rows' entries back to back
 * (2 uint32 per entry).  Upload expects rows sorted by (distance, vid) and
 * at most cap long, as fill_single_layer_graph_from_layer leaves them
 * (refiner/refiner_utils.hpp:73-110). */
int artea_refine_upload_rows(artea_refine* r, const uint32_t* counts, const uint32_t* entries);
int artea_refine_download_rows(artea_refine* r, uint32_t* counts, uint32_t* entries);
/* next_range(RandomUpdater(gen_size, start, end), row_begin, row_end)
 * (refiner/updaters/random_updater.hpp:97-142); `stream` selects the
 * random stream (the full build uses the level id). */
int artea_refine_prefill(artea_refine* r, uint32_t gen_size, uint32_t start, uint32_t end, uint32_t row_begin,
                         uint32_t row_end, uint64_t stream);
/* next(TriangleUpdater(scale_coeffs, shifted_coeffs))
 * (refiner/updaters/triangle_updater.hpp:114-237).  The merge of every
 * next() keeps at most cap entries per row, which equals the reference's
 * merge followed by the arc updater that refine_layer always runs next. */
int artea_refine_triangle(artea_refine* r, float scale_coeffs, float shifted_coeffs);
/* next_range(TriangleUpdater / ReverseUpdater, row_begin, row_end)
 * (refiner/propagate_engine.hpp:217-225): the updater runs on the rows
 * [row_begin, row_end) only (clamped to the row count); the logs it emits
 * are merged into every row. */
int artea_refine_triangle_range(artea_refine* r, float scale_coeffs, float shifted_coeffs, uint32_t row_begin,
                                uint32_t row_end);
int artea_refine_reverse_range(artea_refine* r, uint32_t row_begin, uint32_t row_end);
/* next(arc_updater_t) (refiner/updaters/truncate_updater.hpp:58-76). */
int artea_refine_truncate(artea_refine* r);
/* next(ReverseUpdater) (refiner/updaters/reverse_updater.hpp:72-100). */
int artea_refine_reverse(artea_refine* r);
/* refine_layer's loop (index/artea_graph/index_factory.hpp:385-393):
 * build_loops x (triu_iters x (triangle, arc), reverse except last, arc). */
int artea_refine_schedule(artea_refine* r, uint32_t build_loops, uint32_t triu_iters, float scale_coeffs,
                          float effective_shift);
/* refine_layer's routing loops (index/artea_graph/index_factory.hpp:378-403):
 * `loops` x (next(RoutingUpdater(topk, queue)), next(PruningUpdater(scale)))
 * with next(ReverseUpdater), next(arc) between loops.  RoutingUpdater
 * (refiner/updaters/routing_updater.hpp:84-114): per row, the warm-start
 * beam search of SingleLayerRouter (router/single_layer_router.hpp:259-282;
 * queue capacity max(topk, queue), seeded with the row) over the layer's
 * rows; its results except the row's own vertex, and except those not
 * closer than the row's entry cap - 1 when the row is full, are merged into
 * the row (no degree cap).  PruningUpdater (pruning_updater.hpp:63-177):
 * scaled rule, pairs of two old entries skipped, survivors old.  Rows keep
 * their first cap entries at the end (the writeback's truncation).  topk in
 * [1, 128], queue in [topk, 256] (the reference requires queue >= topk and
 * fails with "candidate_queue_size (q) must be >= topk (k)"). */
int artea_refine_routing_loops(artea_refine* r, uint32_t loops, uint32_t topk, uint32_t queue, float scale_coeffs);
/* next(PruningUpdater(scale)) alone (refiner/updaters/pruning_updater.hpp
 * :170-177): every row pruned in place, no logs. */
int artea_refine_prune(artea_refine* r, float scale_coeffs);
/* next_range(PruningUpdater(scale), row_begin, row_end): the rows of the
 * window only (clamped to the row count). */
int artea_refine_prune_range(artea_refine* r, float scale_coeffs, uint32_t row_begin, uint32_t row_end);
/* Since the last call: routing distances (router and pruning pairs),
 * router expansions, and rows whose candidate list dropped an entry tied with
 * its bound even with 512 entries (their results may differ from the
 * reference's; none expected on real data). */
int artea_refine_routing_stats(artea_refine* r, uint64_t* distances, uint64_t* expansions, uint64_t* unresolved);
/* Precision of the triangle join's pair distances (no reference
 * counterpart; the reference's join is FP32 only).  ARTEA_JOIN_FP32 (the
 * default): every needed pair by dist_func_t's FP32 chain.  ARTEA_JOIN_BF16:
 * the row's Gram matrix on BF16 MFMA tiles (two-term split, FP32
 * accumulation) places every pair's distance in a proven interval; pairs whose
 * interval straddles a pruning or recommendation threshold, and candidates
 * that may be the recommendation target, are recomputed by the FP32 chain, so
 * rows and logs are identical to ARTEA_JOIN_FP32's.  Layers of cap <= 64, and
 * rows whose vectors do not fit in LDS, use the FP32 chain either way. */
enum artea_join_precision { ARTEA_JOIN_FP32 = 0, ARTEA_JOIN_BF16 = 1 };
int artea_refine_set_join(artea_refine* r, int precision);
/* Entries of the BF16 join's per-row queue of pairs to recompute in FP32
 * (0..1024, default 1024).  Pairs beyond it are found again from a bitmap:
 * slower, same result.  A tuning knob; no reference counterpart. */
int artea_refine_set_recheck_queue(artea_refine* r, uint32_t entries);
/* Since the last call: log records merged, pair distances evaluated by the
 * FP32 chain, pair distances evaluated on the BF16 tiles. */
int artea_refine_last_stats(artea_refine* r, uint64_t* logs, uint64_t* exact_pairs, uint64_t* approx_pairs);

/* HierarchicalGraphCompactor::compact_graph (compactor/
 * hierarchical_graph_compactor.hpp:118-327) of a dynamic hierarchical graph,
 * attached to idx: highest[v] = source level (ARTEA_INVALID_LEVEL =
 * unassigned, at most max_level), row_off[l] (n + 1 offsets) and row_ids[l]
 * = the rows of level l (l = 0..max_level, global ids), order = the order
 * the vertices were assigned to their buckets (a permutation; NULL =
 * ascending vid).  The first bucket from the top with >= 32 vertices is the
 * top; higher vertices are demoted into it after its own, by ascending source
 * level, keeping their rows at levels <= top; entry = the top-bucket vertex
 * closest by `metric` to the FP64 centroid of the top bucket (computed on the
 * GPU; ties keep the earliest bucket position).  No assigned vertex: an empty
 * graph without entry. */
int artea_index_compact_graph(artea_index* idx, uint32_t max_level, uint32_t ul_max, uint32_t bl_max,
                              const uint32_t* highest, const uint64_t* const* row_off, const uint32_t* const* row_ids,
                              const uint32_t* order, int metric);

/* Hierarchy build on the GPU, attached to idx.  level_fraction = 1: every
 * vertex on levels 0..max_level (the r-net hierarchy when no vertex is
 * absorbed, index/stacked_rgraph/index_factory.hpp:364-381).  Below 1:
 * vertex v's source level is the highest h <= max_level with u(v) <
 * level_fraction^h (u a seeded uniform draw; a geometric stand-in for the
 * r-net insertion) and level l is refined as a sparse layer of the vertices
 * at level >= l.  Each level is refined by refine_layer with caps
 * floor(1.5 x max), prefill floor(cap x prefill_ratio) from the level's random
 * stream, shift = shifted_coeffs x l0_min_distance on L0 and 0 above; rows
 * are written back truncated to bl_max / ul_max and compacted
 * (compactor/hierarchical_graph_compactor.hpp:118-327: the first bucket from
 * the top with >= 32 vertices is the top, higher vertices are demoted, entry
 * = the top-bucket vertex nearest its FP64 centroid).  Build metric is L2 for
 * an L2SQR index; join_precision is an artea_join_precision.
 * seconds_per_level (optional) gets max_level+1 values. */
int artea_build_full(artea_index* idx, uint32_t ul_max, uint32_t bl_max, uint32_t max_level, uint32_t build_loops,
                     uint32_t triu_iters, float prefill_ratio, float scale_coeffs, float shifted_coeffs,
                     float l0_min_distance, uint64_t seed, int join_precision, float level_fraction,
                     double* seconds_per_level);
/* The same with `routing_loops` routing loops (routing topk and queue) after
 * each level's build loops (PropagateConfig's num_routing_loops,
 * routing_topk, routing_queue_size; index/conv_graph/configs.hpp:33-77,
 * defaults 0, 64, 96); artea_build_full = routing_loops 0. */
int artea_build_full_routing(artea_index* idx, uint32_t ul_max, uint32_t bl_max, uint32_t max_level,
                             uint32_t build_loops, uint32_t triu_iters, float prefill_ratio, float scale_coeffs,
                             float shifted_coeffs, float l0_min_distance, uint64_t seed, int join_precision,
                             float level_fraction, uint32_t routing_loops, uint32_t routing_topk,
                             uint32_t routing_queue, double* seconds_per_level);

/* --------------------------------------------------------- r-net insertion */
/* RGraphConfig (index/stacked_rgraph/configs.hpp:65-138) with the insertion
 * path's RNG scale (PruningConfig::scale_coeffs) and the wave schedule.
 * Vertices are inserted in vid order in waves: each vertex of a wave plans
 * its insertion (_insert_one's descent, covering rule and per-level select
 * searches with HierarchicalPruningUpdater, index_factory.hpp:281-478)
 * against the graph as it stood when the wave began; the wave is then
 * applied in vid order (levels, forward rows, reverse edges per target in
 * vid order, _write_reverse_edges :528-580).  A wave ends after its first
 * vertex that grows the hierarchy.  The first 10 vertices go one by one
 * (startup_points), then waves of min(wave_max, inserted / wave_div);
 * wave_max = 1 is the reference's serial insertion.  The apex of each
 * descent is the top bucket's entry at a seeded counter-based draw (the
 * reference draws it from std::random_device).  Queues: 1 <= search_qs <=
 * ul/bl_select_qs <= 224; max_nbr_size <= 128. */
typedef struct {
    float rnet_beta, tau_k, l0_min_distance, scale_coeffs;
    uint32_t search_nn_qs, ul_select_nbrs_qs, bl_select_nbrs_qs, ul_max_nbr_size, bl_max_nbr_size;
    uint32_t insert_on_l0;
    uint32_t wave_div, wave_max;
    uint64_t seed;
} artea_rnet_params;
typedef struct artea_rnet artea_rnet;
/* stacked_rgraph::IndexFactory::build_from_vectors (index_factory.hpp
 * :176-188) over idx's vectors with build metric `metric`. */
int artea_rnet_build(artea_index* idx, int metric, const artea_rnet_params* p, artea_rnet** out);
/* max_allowed_level_id, top occupied level (ARTEA_INVALID_LEVEL: none),
 * waves, distances evaluated, vertices whose candidate list dropped an
 * entry tied with its bound even with 512 entries (their plan may differ
 * from the reference's; none expected on real data). */
int artea_rnet_info(const artea_rnet* r, uint32_t* max_level, uint32_t* top, uint32_t* waves, uint64_t* distances,
                    uint64_t* unresolved);
/* highest[v] for every vertex. */
int artea_rnet_levels(const artea_rnet* r, uint32_t* highest);
/* Rows of `level` by vid as stored (forward entries, then appended reverse
 * entries): counts[v] (n values) and (vid_status, distance bits) pairs back
 * to back, at most max_entries pairs written; *total = all pairs. */
int artea_rnet_level_rows(const artea_rnet* r, uint32_t level, uint32_t* counts, uint32_t* entries,
                          uint64_t max_entries, uint64_t* total);
int artea_rnet_destroy(artea_rnet* r);
/* The artea_graph build (index/artea_graph/index_factory.hpp:135-410) on the
 * GPU, attached to idx: r-net insertion (insert_on_l0 as given), then
 * refine_layer for every occupied level with at least 32 vertices (rows
 * imported sorted by (distance, vid) into caps floor(1.5 x max); L0 topped
 * up by the random updater to floor(cap x prefill_ratio); the build loops
 * with scale_coeffs and, on L0, shift = shifted_coeffs x l0_min_distance;
 * routing loops), the writeback of the first max_nbr_size entries, and the
 * compactor (vid order = insertion order).  Build metric L2 for an L2SQR
 * index.  seconds (optional) = {r-net insertion, whole build}. */
int artea_build_full_rnet(artea_index* idx, const artea_rnet_params* p, uint32_t build_loops, uint32_t triu_iters,
                          float prefill_ratio, float shifted_coeffs, int join_precision, uint32_t routing_loops,
                          uint32_t routing_topk, uint32_t routing_queue, double* seconds);

/* -------------------------------------------------------------- utilities */
/* dist_func_t on the device (utils/simd_distance.hpp:64-79): dists[i*nb+j]
 * = dist(a_i, b_j) for device arrays, bit-identical to the AVX-512 kernel's
 * 16-lane FMA chain and reduce tree. */
int artea_distance_matrix_device(int metric, const float* d_a, uint32_t na, const float* d_b, uint32_t nb,
                                 uint32_t dim_pad, float* d_out, void* stream);

/* FP64 exact brute-force k nearest neighbours, the recall oracle of the
 * benchmarks (utils/recall_estimator.hpp:53-150 scores against such a
 * ground truth): device buffers, distances accumulated in FP64 from the FP32
 * inputs, ids ascending by (distance, vid).  k <= 256. */
int artea_bruteforce_knn_device(int metric, const float* d_base, uint32_t n, const float* d_queries, uint32_t nq,
                                uint32_t dim_pad, uint32_t k, uint32_t* d_out_vids, double* d_out_dists,
                                void* stream);

/* The same result (ids and FP64 distances identical to
 * artea_bruteforce_knn_device) through a BF16 MFMA filter: the FP64 kernel
 * on a prefix sample of the base bounds each query's k-th distance, a BF16
 * MFMA GEMM (two-term split operands, FP32 accumulation, L2 rows centred on
 * the base mean) keeps every pair whose proven error interval reaches that
 * bound, and the kept pairs are recomputed in FP64 and sorted by (distance,
 * vid); a query with more than 4096 kept pairs is recomputed by the FP64
 * kernel.  Synchronous on `stream`.  dim_pad > 128 or small bases go to the
 * FP64 kernel directly.  stats (optional, 3 values): pairs kept, queries
 * recomputed by the FP64 kernel, sample rows. */
int artea_bruteforce_knn_mfma_device(int metric, const float* d_base, uint32_t n, const float* d_queries,
                                     uint32_t nq, uint32_t dim_pad, uint32_t k, uint32_t* d_out_vids,
                                     double* d_out_dists, void* stream, uint64_t* stats);

/* ------------------------------------------------ scoring and data probes */
/* RecallEstimator::calculate_recall_at_k (utils/recall_estimator.hpp:54-230)
 * on device id rows: *matches = sum over the nq queries of the number of
 * distinct ids among the first k of the prediction row that occur among the
 * first k of the ground-truth row; Recall@K = matches / (nq * k).  k <=
 * 1024; a ground-truth row shorter than k fails as the reference does
 * ("Ground truth dimension ... is less than k=...").  Synchronous on
 * `stream`. */
int artea_recall_at_k_device(const uint32_t* d_pred, uint32_t pred_stride, const uint32_t* d_gt, uint32_t gt_stride,
                             uint32_t nq, uint32_t k, uint64_t* matches, void* stream);
/* ADREstimator::calculate_adr_at_1 (utils/adr_estimator.hpp): the mean over
 * queries of double(dist(q, pred[q][0])) / double(dist(q, gt[q][0])) with
 * dist_func_t's FP32 chain; queries whose ground-truth distance is <= 0
 * are skipped (a NaN one is not), as are queries whose top-1 slot holds no
 * vertex (the invalid entry); the terms are summed in query order.  *counted = the
 * queries summed; *adr = 0 when none.  Synchronous on `stream`. */
int artea_adr_at_1_device(int metric, const float* d_base, uint32_t n, const float* d_queries, uint32_t nq,
                          uint32_t dim_pad, const uint32_t* d_pred, uint32_t pred_stride, const uint32_t* d_gt,
                          uint32_t gt_stride, double* adr, uint64_t* counted, void* stream);
/* DatasetProber::_compute_knn_table (utils/dataset_prober.hpp:471-540): for
 * each sample vid (host array, every id < n), the 128 smallest
 * dist(base[vid], base[j]) over j != vid, ascending, into knn (host,
 * num_samples x 128); *min_nearest (optional) = the smallest first column
 * (probe's min_nearest, :253-306).  A build's auto-probed l0_min_distance
 * is the first column's entry at one quantile
 * (unit_tests/test_hierarchical_graph_persistence.cpp:125-137).  n must
 * exceed 128.  The reference draws
 * the samples from MKL streams; here the caller passes them. */
int artea_probe_knn_device(int metric, const float* d_base, uint32_t n, uint32_t dim_pad, const uint32_t* samples,
                           uint32_t num_samples, float* knn, float* min_nearest, void* stream);
/* DatasetProber::_compute_farthest (:542-585): per sample, the largest
 * dist(base[vid], base[j]) over j != vid, into farthest (host). */
int artea_probe_farthest_device(int metric, const float* d_base, uint32_t n, uint32_t dim_pad,
                                const uint32_t* samples, uint32_t num_samples, float* farthest, void* stream);
/* out[i * k + j] = dist(a_i, b[ids[i * ids_stride + j]]) for device arrays
 * (the query-to-ground-truth distances of DatasetProber::probe_query,
 * :327-432); an id at or beyond nb (the invalid vertex of a
 * padded result row) gives +inf. */
int artea_distance_rows_device(int metric, const float* d_a, uint32_t na, const float* d_b, uint32_t nb,
                               const uint32_t* d_ids, uint32_t ids_stride, uint32_t k, uint32_t dim_pad, float* d_out,
                               void* stream);

/* ---------------------------------------------------------------- serving */
/* Continuous-batching request server (no reference counterpart: the
 * reference's batch_query, router/hierarchical_graph_router.hpp:293-385, is
 * one synchronous call per batch; this queues requests across calls).  Each
 * request carries its own topk (1..max_topk) and queue size (topk..max_queue;
 * a queue below topk fails as the router ctor does, :104-108) and gets
 * exactly batch_query's result for it: ids ascending by (distance, vid),
 * padded to max_topk slots with (ARTEA_INVALID_VID, FLT_MAX).  `lanes`
 * streams (1..16) each run at most `budget` requests at once; requests wait
 * in one queue per queue-size class, and an idle lane takes the class of the
 * oldest waiting request, up to `budget` of its requests, oldest first.
 * submit copies the requests (queries n x dim_pad floats) and returns at
 * once; tickets are consecutive from *first_ticket.  poll never blocks: it
 * launches what idle lanes can take and returns up to max_results completed
 * requests (tickets[i], row i of out_vids / out_dists, max_topk entries per
 * row) in completion order.  status counts queued, in-flight and completed-
 * but-unreturned requests.  Not thread-safe per handle; after a failed call
 * other than an argument error, destroy the server. */
typedef struct artea_server artea_server;
int artea_server_create(artea_index* idx, uint32_t max_topk, uint32_t max_queue, uint32_t lanes, uint32_t budget,
                        artea_server** out);
/* The server with admission by device memory: the lanes' device buffers
 * take at most hbm_bytes, so a lane holds hbm_bytes / (lanes * (4 dim_pad +
 * 8 max_topk + 8)) requests (query row, result ids and distances, topk and
 * queue words); and a launch of one queue-size class takes at most
 * lds_rounds times the wavefronts that class's kernel keeps resident on the
 * GPU, which its LDS per wavefront (candidate-list mirror, visited cache,
 * query row) and registers decide.  Larger queues, more LDS per wavefront,
 * fewer requests per launch.  EINVAL when hbm_bytes hold less than one
 * request per lane. */
int artea_server_create_budget(artea_index* idx, uint32_t max_topk, uint32_t max_queue, uint32_t lanes,
                               uint64_t hbm_bytes, uint32_t lds_rounds, artea_server** out);
/* Admission limits: admit[c] requests per launch and resident[c] resident
 * wavefronts (0 for the count form) of each of the 6 queue-size classes
 * (candidate lists up to 48, 112, 240, 496, 816, 1008 entries), the requests
 * one lane holds, and the lanes' device bytes. */
int artea_server_admission(artea_server* s, uint32_t* admit, uint32_t* resident, uint32_t* lane_requests,
                           uint64_t* hbm_bytes);
int artea_server_destroy(artea_server* s);
int artea_server_submit(artea_server* s, const float* queries, uint32_t n, const uint32_t* topk,
                        const uint32_t* queue, uint64_t* first_ticket);
int artea_server_poll(artea_server* s, uint32_t max_results, uint64_t* tickets, uint32_t* out_vids,
                      float* out_dists, uint32_t* n_out);
/* Same, writing the result of ticket t into row t - base_ticket of the
 * caller's tables (table_rows rows of max_topk entries each); tickets[i]
 * lists the requests returned.  A ticket outside the table fails with
 * ARTEA_EINVAL (the requests returned before it are counted in *n_out). */
int artea_server_poll_into(artea_server* s, uint64_t base_ticket, uint64_t table_rows, uint32_t* table_vids,
                           float* table_dists, uint32_t max_results, uint64_t* tickets, uint32_t* n_out);
int artea_server_status(artea_server* s, uint64_t* queued, uint64_t* in_flight, uint64_t* ready);
/* Requests whose candidate list overflowed its tie tail
 * (artea_router_last_overflows) are run again, before they are returned, on
 * the next larger list class until none overflows: *rerun counts them;
 * *unresolved counts those that still overflowed in the largest class
 * (1024 entries) and were returned as computed. */
int artea_server_overflows(artea_server* s, uint64_t* rerun, uint64_t* unresolved);

/* Device plumbing for callers without a GPU framework of their own: plain
 * allocations, copies, streams (default priority) and timing events of the
 * library's HIP runtime. */
int artea_device_alloc(size_t bytes, void** d_ptr);
int artea_device_free(void* d_ptr);
int artea_memcpy_htod(void* d_dst, const void* src, size_t bytes, void* stream);
int artea_memcpy_dtoh(void* dst, const void* d_src, size_t bytes, void* stream);
int artea_memset_device(void* d_dst, int value, size_t bytes, void* stream);
/* Device-to-device copy by a plain kernel (bytes % 16 == 0); also the
 * calibration kernel for HBM byte counters. */
int artea_copy_device(void* d_dst, const void* d_src, size_t bytes, void* stream);
int artea_stream_create(void** stream);
int artea_stream_destroy(void* stream);
int artea_stream_synchronize(void* stream);
int artea_event_create(void** event);
int artea_event_destroy(void* event);
int artea_event_record(void* event, void* stream);
int artea_event_elapsed_ms(void* start, void* end, float* ms);
/* 1 when all work before the event has completed, 0 when not yet. */
int artea_event_query(void* event, int* done);
/* Page-locked host staging memory (hipHostMalloc). */
int artea_host_alloc(size_t bytes, void** h_ptr);
int artea_host_free(void* h_ptr);
int artea_device_info(char* name, size_t name_len, int* compute_units, size_t* total_bytes);

/* Seeded synthetic workloads, host side, identical on every machine:
 * kind 0 = integer-rounded 256-component Gaussian mixture (centres U[0,128],
 * sigma 12, clamped to [0,255]); 1 = power-law-spectrum Gaussian on a random
 * orthogonal basis, unit-normalised; 2 = 1024-component mixture (centres
 * N(0,1), sigma 0.3), unit-normalised.  which = 0 base, 1 queries.  Writes
 * n x dim_pad floats (zero tail beyond dim). */
int artea_generate(int kind, uint64_t seed, uint32_t which, uint64_t first, uint64_t n, uint32_t dim,
                   uint32_t dim_pad, float* out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

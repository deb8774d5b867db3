/*
 * gphi.h -- C ABI of libgphi, the MI355X-native tiled graph-traversal library.
 *
 * The reference (johnpzh/bfs_simd_benchmarks) has no library API: its boundary
 * is the process CLI, the on-disk 64-way text formats and the C-style
 * signatures inside each app's single translation unit, with sizes passed
 * through globals.  Each entry point below names the reference function whose
 * role it takes over (paths relative to the reference repository root).
 *
 * Conventions
 *   - every function returns int: GPHI_OK (0) or a negative GPHI_E_* code;
 *     gphi_last_error() gives the text.  The library never calls exit().
 *   - all vertex ids are 0-based uint32 (the files are 1-based, as the
 *     reference's readers subtract one: apps/bfs_simd/bfs.cpp:1049-1050).
 *   - a gphi_ctx owns one HIP device and one HIP stream; calls on different
 *     contexts may run concurrently; calls on one context are serialised.
 *   - device memory is owned by the library; host buffers are the caller's.
 */
#ifndef GPHI_H
#define GPHI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPHI_OK 0
#define GPHI_E_IO -1      /* missing or unreadable file */
#define GPHI_E_FORMAT -2  /* truncated / malformed file, header mismatch */
#define GPHI_E_RANGE -3   /* id or size out of range (e.g. S > SIDE, M >= 2^32) */
#define GPHI_E_OOM -4     /* host or device allocation failed */
#define GPHI_E_HIP -5     /* HIP runtime error */
#define GPHI_E_ARG -6     /* invalid argument */
#define GPHI_E_STATE -7   /* wrong object state (e.g. graph lacks weights) */

typedef struct gphi_ctx gphi_ctx;
typedef struct gphi_graph gphi_graph;

typedef struct gphi_opts {
	int device;          /* HIP device ordinal */
	uint32_t t_ratio;    /* direction rule |F|+sum(outdeg) > M/t_ratio; 0 -> 20
	                        (apps/bfs_simd/bfs.cpp:1171) */
} gphi_opts;

/* Tile order of the COO arrays. */
#define GPHI_TILES_COL_STRIPE 0 /* _col-S-coo-tiled-W (tools/column_major_tile/kcore.cpp:378-415) */
#define GPHI_TILES_ROW_MAJOR 1  /* _coo-tiled-W (tools/coo_tiling/page_rank.cpp:430) */

/* Host arrays of one loaded graph: exactly what graph_input hands to
 * graph_prepare (apps/bfs_simd/bfs.cpp:968-1118, :753-761), without globals.
 * Any of csr_edges / tile_heads+tile_tails may be NULL when the caller only
 * needs the other half (e.g. page_rank reads row-major tiles only). */
typedef struct gphi_graph_desc {
	uint32_t n;                   /* NNODES */
	uint64_t m;                   /* NEDGES (< 2^32, as the formats are %u) */
	uint32_t tile_width;          /* W */
	uint32_t stripe_length;       /* S */
	int tile_order;               /* GPHI_TILES_* */
	const uint32_t *degrees;      /* n out-degrees (the -nneibor file) */
	const uint32_t *csr_edges;    /* m tails in _untiled order (graph_edges) */
	const uint32_t *csr_weights;  /* m weights in _untiled order, or NULL */
	const uint32_t *tile_heads;   /* m heads in tile order (graph_heads) */
	const uint32_t *tile_tails;   /* m tails in tile order (graph_tails) */
	const uint32_t *tile_weights; /* m weights in tile order, or NULL */
	const uint32_t *tile_offsets; /* SIDE^2 tile offsets (tile_offsets), or NULL */
} gphi_graph_desc;

/* Per-step statistics of one traversal.  A step expands one frontier. */
#define GPHI_MAX_STEPS 512
typedef struct gphi_stats {
	uint32_t steps;                      /* frontiers expanded */
	uint32_t dir[GPHI_MAX_STEPS];        /* 0 = sparse (top-down CSR), 1 = dense (tiles) */
	uint32_t fsize[GPHI_MAX_STEPS];      /* |F| expanded by the step */
	uint32_t fdeg[GPHI_MAX_STEPS];       /* sum of out-degrees of F */
	float sweep_ms[GPHI_MAX_STEPS];      /* sweep kernel time (HIP events) */
	float select_ms[GPHI_MAX_STEPS];     /* chunk selection before a dense sweep (0 if sparse) */
	float update_ms[GPHI_MAX_STEPS];     /* frontier update kernel time */
	uint32_t chunks[GPHI_MAX_STEPS];     /* 2048-edge chunks a dense step swept (0 if sparse) */
	float init_ms;                       /* state initialisation (device) */
	float total_ms;                      /* init + all steps (device events) */
	uint64_t reached;                    /* vertices with a finite result */
	uint64_t edges_traversed;            /* sum of outdeg over reached vertices */
} gphi_stats;

/* ---- context: replaces the globals NNODES..T_RATIO (apps/bfs_simd/bfs.cpp:17-26) ---- */
int gphi_open(const gphi_opts *opts, gphi_ctx **out);
/* Per-level timing of the traversals on this context (gphi_stats sweep_ms,
 * select_ms, update_ms, init_ms from HIP events around each level's kernels):
 * on (1) by default.  Off (0), a traversal records only its total_ms, and the
 * kernels of consecutive levels run back to back (the per-level events cost
 * about 0.1 ms of a 2.6 ms scale-25 BFS).  The reference's counterpart is its
 * per-phase timers, apps/bfs_simd/bfs.cpp:733-751. */
int gphi_set_level_timing(gphi_ctx *ctx, int on);
void gphi_close(gphi_ctx *ctx);
const char *gphi_last_error(gphi_ctx *ctx);
const char *gphi_version(void);

/* ---- graphs ---- */
/* Upload host arrays (graph_input's outputs).  apps/bfs_simd/bfs.cpp:968-1118 */
int gphi_graph_from_desc(gphi_ctx *ctx, const gphi_graph_desc *desc, gphi_graph **out);
/* Read the reference's 64-way text files and upload:
 *   prefix_col-S-coo-tiled-W-{0..63},-offsets  and  prefix_untiled-{0..63},-nneibor
 * (apps/bfs_simd/bfs.cpp:983-1113; weighted: apps/sssp_simd/sssp.cpp:41-189).
 * `prefix` is the full name before the format suffix, e.g. "rmat16_reorder". */
int gphi_graph_load_text(gphi_ctx *ctx, const char *prefix, uint32_t W, uint32_t S,
                         int weighted, gphi_graph **out);
/* Read prefix_coo-tiled-W-{0..63},-offsets,-nneibor (row-major tiles only), as
 * apps/pageRank_simd/page_rank.cpp:468-568 and apps/bellman-ford_simd do. */
int gphi_graph_load_text_rowmajor(gphi_ctx *ctx, const char *prefix, uint32_t W,
                                  uint32_t S, gphi_graph **out);
int gphi_graph_info(const gphi_graph *g, uint32_t *n, uint64_t *m, uint32_t *W, uint32_t *S,
                    int *weighted);
uint64_t gphi_graph_device_bytes(const gphi_graph *g);
void gphi_graph_free(gphi_graph *g);

/* ---- traversals (device-resident graph, results copied to host buffers) ---- */
/* Hybrid BFS: graph_prepare, apps/bfs_simd/bfs.cpp:753-963.
 * levels[n]: hop level (-1 unreached, the h_cost array); parents[n]: BFS-tree
 * parent (-1 unreached, parent[src]=src).  Either output may be NULL. */
int gphi_bfs(gphi_graph *g, uint32_t src, int32_t *levels, int32_t *parents, gphi_stats *st);
/* Hybrid BFS that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * traversal and results as gphi_bfs, but the traversal state lives in one
 * fixed 128 MiB device array per GPU (held from gphi_open until the process
 * ends) instead of buffers the library allocates, and the results stay on the
 * device: levels_dev / parents_dev are device pointers of n int32 each (-1
 * unreached, parent[src]=src), either may be NULL.  A graph whose state does
 * not fit (about 25 bytes per vertex plus 8 per 1024 edges: N up to roughly
 * 5 million) gets GPHI_E_RANGE.  The calls on one GPU share its array, so
 * they are serialised by a process-wide lock (calls that overlapped would
 * race on it); each returns after its kernels are done. */
int gphi_bfs_noalloc(gphi_graph *g, uint32_t src, int32_t *levels_dev, int32_t *parents_dev,
                     gphi_stats *st);
/* Weighted hybrid SSSP: sssp_weighted, apps/sssp_simd/sssp.cpp:969-1112.
 * dist[n]: shortest distance, 2147483647 when unreached. */
int gphi_sssp(gphi_graph *g, uint32_t src, uint32_t *dist, gphi_stats *st);
/* PageRank sweeps: page_rank, apps/pageRank_simd/page_rank.cpp:362-456, one
 * sweep per iteration; rank starts at `init` everywhere, sum is cleared per
 * iteration; rank = (1-damping)/N + damping*sum in FP32. */
int gphi_pagerank(gphi_graph *g, uint32_t iters, float damping, float init, float *rank,
                  float *ms);
/* Hop relaxation: apps/bellman-ford_simd/bellman-ford.cpp:291-392 (init
 * INT_MAX-2, dist[src]=0, at most max_sweeps sweeps, stops at the fixpoint). */
int gphi_bellman_ford(gphi_graph *g, uint32_t src, uint32_t max_sweeps, uint32_t *dist,
                      uint32_t *sweeps_run, float *ms);

/* Multi-source BFS: k (1..256) sources in one traversal, bit-packed per
 * vertex; one edge stream serves all of them (no reference counterpart: the
 * batched form of graph_prepare, apps/bfs_simd/bfs.cpp:753-963, with its
 * CSR push and direction rule applied to the union frontier; dense steps pull
 * over the in-edges, and push the sources whose own frontiers are small).
 * levels: k*n uint8, source-major (levels[i*n + v]); 255 = unreached; NULL
 * keeps them in HBM.  Each row equals gphi_bfs's levels of that source; a
 * traversal deeper than 254 levels returns GPHI_E_RANGE.  st->reached and
 * st->edges_traversed are summed over the k sources. */
int gphi_bfs_batch(gphi_graph *g, const uint32_t *srcs, uint32_t k, uint8_t *levels,
                   gphi_stats *st);
/* Multi-source weighted SSSP: k (1..GPHI_SSSP_BATCH_MAX) sources in one
 * label-correcting traversal (no reference counterpart: sssp_weighted,
 * apps/sssp_simd/sssp.cpp:969-1112, with the 16 sources' distances of a
 * vertex side by side, relaxed together along every edge of the union
 * frontier).  dist: k*n uint32, source-major (dist[i*n + v]), 2147483647 when
 * unreached; NULL keeps them in HBM.  Each row equals gphi_sssp's distances
 * of that source.  st->steps and st->total_ms are set. */
#define GPHI_SSSP_BATCH_MAX 16
int gphi_sssp_batch(gphi_graph *g, const uint32_t *srcs, uint32_t k, uint32_t *dist, gphi_stats *st);

/* ---- serving: resident graphs, continuous batching (no reference counterpart) ----
 * Each added graph gets a BFS lane with the graph's own context (HIP stream)
 * and a worker thread; lanes run concurrently on the device.  A BFS lane
 * takes all BFS queries pending for its graph when it becomes free (up to
 * max_batch) and answers them with one multi-source traversal when that is
 * expected to beat running them one by one (running estimates of both).
 * That traversal runs level by level: BFS queries that arrive while it runs
 * join it at the next level boundary, each in a free source slot of its own
 * (up to the launch's mask width: 64, 128 or 256 sources), and every query
 * completes as soon as its own frontier is empty.  A weighted graph also
 * gets sssp_lanes SSSP lanes, each a gphi_graph_view on a context of its own,
 * running SSSP queries beside the BFS lane: the SSSP queries pending on a
 * lane (up to GPHI_SSSP_BATCH_MAX) share one multi-source SSSP traversal
 * when that is expected to beat running them alone (with max_batch 1, every
 * query runs alone).  A graph is
 * admitted only while its HBM bytes plus every lane's traversal state fit
 * hbm_budget (GPHI_E_OOM otherwise).  The server takes ownership of each
 * added graph and of its context (create one context per graph). */
typedef struct gphi_server gphi_server;
typedef struct gphi_server_opts {
	uint32_t max_batch;   /* BFS queries per launch, 1..256 (0 -> 64); 1 = no batching */
	uint64_t hbm_budget;  /* bytes; 0 -> 90% of device memory */
	uint32_t sssp_lanes;  /* SSSP lanes per weighted graph, 1..8 (0 -> 2) */
} gphi_server_opts;
#define GPHI_QUERY_BFS 0
#define GPHI_QUERY_SSSP 1
typedef struct gphi_query_info {
	double latency_us;      /* submit -> result */
	double queue_us;        /* submit -> start of its work (launch start, or the level it joined at) */
	double run_us;          /* start of its work -> result */
	uint32_t batch;         /* queries answered by that launch (so far, when it completed) */
	uint32_t joined_level;  /* level of the running traversal it joined at; 0 = started with it */
	uint64_t launch;        /* launch that answered it (per server, counting from 1) */
} gphi_query_info;
int gphi_server_create(const gphi_server_opts *opts, gphi_server **out);
int gphi_server_add_graph(gphi_server *s, gphi_graph *g, uint32_t *graph_id);
/* out: NULL, or n int32 levels (BFS, -1 unreached) / n uint32 distances (SSSP),
 * written before the query completes */
int gphi_server_submit(gphi_server *s, uint32_t graph_id, uint32_t kind, uint32_t src, void *out,
                       uint64_t *ticket);
int gphi_server_wait(gphi_server *s, uint64_t ticket, gphi_query_info *info);
int gphi_server_stats(gphi_server *s, uint32_t graph_id, uint64_t *batches, uint64_t *served,
                      uint64_t *hbm_bytes);
void gphi_server_destroy(gphi_server *s);

/* ---- synthetic inputs and the tool pipeline on the device ----
 * The reference's inputs came from external generators (PBBS rMat,
 * tools/adj2snap/adj2snap.cpp; Galois graph-convert).  gphi_graph_generate
 * draws an RMAT edge list (project spec: per-sample splitmix64 streams), drops
 * self-loops / ids >= nverts / duplicates, optionally permutes ids, relabels
 * by degree (tools/vertex_id_reorder_to_degrees/bfs.cpp:50-83, glibc rand()
 * tie-break) and builds the CSR (tools/csr_tiling) and tile order
 * (tools/coo_tiling, tools/column_major_tile) directly in HBM. */
typedef struct gphi_rmat_spec {
	uint32_t scale;          /* ids drawn from [0, 2^scale) */
	uint32_t nverts;         /* N; samples with an id >= N are dropped */
	uint64_t seed;
	uint64_t samples;        /* samples drawn (upper bound when target_unique > 0) */
	uint64_t target_unique;  /* 0: keep all unique edges; else the first this many */
	double a, b, c;          /* quadrant probabilities (d = 1-a-b-c) */
	uint64_t perm_seed;      /* 0: no id permutation before reordering */
} gphi_rmat_spec;

typedef struct gphi_weight_spec {
	uint64_t seed;           /* 0: unweighted */
	uint32_t lo, hi;         /* uniform integer weights in [lo, hi] */
} gphi_weight_spec;

#define GPHI_REORDER_NONE 0
#define GPHI_REORDER_DEGREE 1 /* tools/vertex_id_reorder_to_degrees */
#define GPHI_REORDER_BFS 2    /* tools/vertex_id_remap/bfs.cpp:168-247,465-561: BFS from vertex 0,
                                 ids level by level in ascending old id, unreached last */

int gphi_graph_generate(gphi_ctx *ctx, const gphi_rmat_spec *spec, uint32_t W, uint32_t S,
                        int reorder, const gphi_weight_spec *weights, int tile_order,
                        gphi_graph **out);
/* Copy a graph's arrays to host buffers (any may be NULL): degrees[n],
 * csr_edges[m], csr_weights[m], tile_heads[m], tile_tails[m], tile_weights[m]. */
int gphi_graph_download(gphi_graph *g, uint32_t *degrees, uint32_t *csr_edges,
                        uint32_t *csr_weights, uint32_t *tile_heads, uint32_t *tile_tails,
                        uint32_t *tile_weights);

/* ---- host-side graphs: the reference's text formats without a device ----
 * A gphi_host_graph holds graph_input's arrays in host memory.  Reading
 * validates every count and id (the reference readers never check fscanf,
 * apps/bfs_simd/bfs.cpp:1046-1051); writing produces the exact bytes of the
 * reference tools (tools/csr_tiling/page_rank.cpp:570-598,
 * tools/coo_tiling/page_rank.cpp:450-499, tools/column_major_tile/kcore.cpp:419-450).
 * Errors are reported through gphi_last_error(NULL). */
typedef struct gphi_host_graph gphi_host_graph;
/* row_major = 0: prefix_col-S-coo-tiled-W + prefix_untiled; 1: prefix_coo-tiled-W */
int gphi_host_read_text(const char *prefix, uint32_t W, uint32_t S, int weighted, int row_major,
                        gphi_host_graph **out);
int gphi_host_from_device(gphi_graph *g, gphi_host_graph **out);
/* Writes prefix_untiled-* (if it holds the CSR) and its tile files
 * (prefix_col-S-coo-tiled-W-* or prefix_coo-tiled-W-*, with -offsets, -nneibor). */
int gphi_host_write_text(const gphi_host_graph *h, const char *prefix);
int gphi_host_desc(const gphi_host_graph *h, gphi_graph_desc *out);
void gphi_host_free(gphi_host_graph *h);

/* A second handle on a resident graph: it shares the graph's device arrays
 * but has its own traversal state and runs on ctx's stream, so traversals of
 * the two handles can run at the same time (no reference counterpart; the
 * serving loop uses it to run SSSP beside batched BFS).  ctx must be on the
 * graph's device.  Free every view before the graph; re-tiling the graph
 * invalidates its views. */
int gphi_graph_view(gphi_ctx *ctx, gphi_graph *base, gphi_graph **out);

/* Rebuild the tile arrays from the graph's CSR with another W, S or order
 * (tools/coo_tiling + tools/column_major_tile again, on the device); used by
 * the (W, S) autotuner.  Traversal state is dropped and rebuilt lazily. */
int gphi_graph_retile(gphi_graph *g, uint32_t W, uint32_t S, int tile_order);

/* Binary tiled-graph cache: one file of raw uint32 arrays behind a 64-byte
 * header with a checksum.  Replaces re-parsing the 64 text parts that
 * graph_input reads on every run (apps/bfs_simd/bfs.cpp:968-1118) and the
 * per-(W, S) re-reads of tools/tile_size/process.sh:19-21.  gphi_graph_load
 * validates the header, the file size, the checksum and every id. */
int gphi_graph_save(gphi_graph *g, const char *path);
int gphi_graph_load(gphi_ctx *ctx, const char *path, gphi_graph **out);

/* Exclusive tile offsets in the stored tile order (the -offsets file);
 * out has SIDE^2 entries. */
int gphi_graph_tile_offsets(gphi_graph *g, uint32_t *out);

/* Testing hook (no reference counterpart): the device allocation `after`
 * allocations from now fails as if HBM were exhausted (GPHI_E_OOM); after < 0
 * clears it.  Lets the tests check that a failed allocation leaves a graph
 * that later calls can still use or that reports an error, never one that a
 * kernel would be launched on half-built. */
void gphi_debug_fail_alloc(int after);

/* Algorithmic HBM bytes of a finished BFS (per step, the byte model of the
 * design doc) from its levels and stats; bytes_per_step has st->steps entries. */
int gphi_bfs_alg_bytes(const gphi_graph *g, const int32_t *levels, const gphi_stats *st,
                       double *bytes_per_step);

#ifdef __cplusplus
}
#endif
#endif

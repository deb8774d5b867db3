// internal.h -- shared declarations of libgphi (host C++ and HIP translation units).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gphi.h"

namespace gphi {

// Edges of one sweep chunk.  The tiled COO arrays are padded to a multiple of
// this with sentinel edges (head = tail = n) whose frontier bit is never set.
constexpr uint32_t kChunk = 2048;
constexpr uint32_t kBlock = 256;
constexpr uint32_t kSparseTile = 1024;  // edges per block iteration of the CSR push
constexpr uint32_t kPadVerts = kBlock * 32;  // vertices per frontier-update block
constexpr uint32_t kPullLight = 64;     // pull steps: a lane per vertex up to this in-degree,
constexpr uint32_t kPullPiece = 4096;    // else a wavefront per piece of this many in-edges
constexpr uint8_t kVisited = 2;          // state byte of a vertex reached in an earlier step
constexpr int kNumXcd = 8;              // MI355X: 8 XCDs, blocks dispatched round-robin
// The per-XCD live-list counters, one 128-byte line apart so that their
// atomics do not queue on one line.
constexpr uint32_t kLiveCntStride = 32;
constexpr size_t kLiveCntBytes = (size_t)kNumXcd * kLiveCntStride * 4;

struct Error {
	int code;
	std::string msg;
};

// Per-chunk summary used to skip whole chunks of the tile sweep: the range of
// head tile-rows and tail tile-columns its (non-sentinel) edges touch.
struct ChunkMeta {
	uint32_t row_lo, row_hi, col_lo, col_hi;
};

// 16-bit forms of a chunk's heads and tails.  Heads: offsets from the first
// vertex of the aligned run of tile rows that holds the chunk's rows.  Tails:
// offsets from the smallest tail of each group of kPackGroup consecutive
// edges (a wavefront's quad row in the dense sweep).  flags bit 0 / bit 1:
// every head / tail of the chunk, padding included, fits in 16 bits, so the
// 16-bit form is valid.
constexpr uint32_t kPackGroup = 256;
struct ChunkPack {
	uint32_t hbase, flags, pad0, pad1;
	uint32_t tbase[8];  // per group of kPackGroup edges
};

// Control block of one frontier traversal (128 bytes).  The level loop runs
// on the device: every kernel of a level reads the direction, the level and
// the queue sizes here, and the host reads the block back only every few
// levels to see whether the traversal is done.
struct StepCtrl {
	uint32_t fsize;         // |F| of the frontier just produced
	uint32_t fdeg;          // sum of out-degrees of that frontier
	uint32_t qsize;         // queue entries (frontier vertices with outdeg > 0)
	uint32_t dir;           // direction chosen for expanding it
	uint32_t done;          // frontier empty
	uint32_t blocks_done;   // last-block detection
	uint32_t facc;          // accumulator: |F|
	uint32_t err;           // multi-source BFS: a level would not fit its uint8 levels
	unsigned long long qpack;  // accumulator: (queue count << 32) | degree sum
	uint32_t changed;       // hop relaxation: edges that lowered a distance
	uint32_t chunks;        // chunks on the live lists of the last dense sweep
	uint32_t level;         // the step being expanded (device-driven loop)
	uint32_t queued;        // BFS: the step expanded last built the next queue itself
	unsigned long long reached, traversed;  // sums of |F| and of sum deg(F) over the steps
	uint32_t col_stale;     // BFS: column counts miss the vertices of queue-building steps
	uint32_t pad[13];
};

// A BFS sparse step whose frontier's edges fill at most this many CSR-push
// tiles builds the next queue itself (its claims append to it); a larger one
// leaves the next frontier to the frontier update's pass over all vertices,
// which then costs less than the appends' reservations on one counter.
constexpr uint32_t kQueueStepTiles = 512;
__host__ __device__ inline bool queue_step(uint32_t dir, uint32_t fdeg)
{
	return dir == 0 && ((uint64_t)fdeg + kSparseTile - 1) / kSparseTile <= kQueueStepTiles;
}
static_assert(sizeof(StepCtrl) == 128, "StepCtrl is 128 bytes");

// The queue of a sparse step (vertices, degree prefix, first entry of each
// kSparseTile-edge tile) in two halves, one per level parity: the step that
// expands level L reads half L & 1, and the next frontier goes to the other.
struct QueueBufs {
	uint32_t *queue, *qoff, *tile_lo;  // half 0
	uint32_t qstride, tstride;         // half 1 starts this many entries later
};
inline uint32_t queue_stride(uint32_t n_pad) { return n_pad; }
inline uint32_t tile_lo_stride(uint64_t m) { return (uint32_t)(m / kSparseTile + 2); }

struct DevGraph {
	uint32_t n = 0;
	uint64_t m = 0;
	uint32_t W = 0, S = 0, side = 0;
	uint64_t ntiles = 0;
	int tile_order = GPHI_TILES_COL_STRIPE;
	uint32_t n_words = 0;      // 32-bit words of a vertex bitmap (covers id n)
	uint32_t n_pad = 0;        // padded vertex count (multiple of 32*kBlock, > n)
	uint64_t m_pad = 0;        // tile arrays padded to whole chunks
	uint64_t nchunks = 0;
	bool has_csr = false, has_tiles = false, weighted = false;
	// CSR (the _untiled files)
	uint32_t *row_off = nullptr;  // n+1
	uint32_t *cols = nullptr;     // m
	uint32_t *csr_w = nullptr;    // m or null
	uint32_t *deg = nullptr;      // n_pad
	// tiled COO (padded)
	uint32_t *heads = nullptr, *tails = nullptr, *coo_w = nullptr;
	ChunkMeta *meta = nullptr;
	uint16_t *heads16 = nullptr, *tails16 = nullptr;  // or null when no window fits (W too wide)
	ChunkPack *pack = nullptr;
	uint8_t *coo_w8 = nullptr;  // 8-bit tile-order weights when every weight is below 256
	uint64_t device_bytes = 0;
};

// PageRank head bands (build_pr_bands).  The heads are cut into bands
// [head_lo, head_hi); within a band, the in-edges of a tail whose heads lie
// in it form a pair (tail, first edge, count).  A small pair has 1..kPrSmall
// in-edges, a light pair up to kPullLight, a heavy pair more, split into
// pieces of kPullPiece; the heavy pairs of more than one piece are also
// listed apart (multi).  Each class is in tail order; indices are global
// over the band arrays.  The heads are copied band by band (pr_src): a
// band's small pairs' heads from edge_small on, its light pairs' from
// edge_light, its heavy pairs' from edge_heavy, each pair's ascending, so a
// band's sweep streams one contiguous run of the copy.
constexpr uint32_t kSsspBatchMax = GPHI_SSSP_BATCH_MAX;  // distances per vertex of the SSSP batch
constexpr uint32_t kBandHeads = 1u << 19;  // heads per band window: 2 MiB of FP32 contributions
constexpr uint32_t kPrSmall = 8;           // in-edges a lane sums alone
struct PrBand {
	uint32_t head_lo, head_hi;
	uint32_t small0, nsmall;
	uint32_t light0, nlight;
	uint32_t heavy0, nheavy;
	uint32_t piece0, npieces;
	uint32_t multi0, nmulti;
	uint32_t edge_small, edge_light, edge_heavy;
};

}  // namespace gphi

struct gphi_ctx {
	int device = 0;
	uint32_t t_ratio = 20;
	hipStream_t stream = nullptr;
	std::mutex mu;
	gphi::Error err{0, ""};
	std::string last;
	// pinned host control block for per-step readback
	gphi::StepCtrl *h_ctrl = nullptr;
	// per-step timing: [0] start, [1] init done, [2+2k] sweep k done, [3+2k] update k
	// done, [2*MAX+4+k] chunk selection of dense step k done
	hipEvent_t ev[3 * GPHI_MAX_STEPS + 4];
	bool ev_ok = false;
	bool level_timing = true;  // gphi_set_level_timing
};

struct gphi_graph {
	gphi_ctx *ctx = nullptr;
	gphi::DevGraph g;
	gphi_graph *base = nullptr;  // a view (gphi_graph_view) shares base's graph arrays
	// per-graph traversal state (allocated lazily, reused across calls)
	int32_t *level = nullptr;  // levels (BFS) ; dist (SSSP/BF) aliases via uint32
	int32_t *parent = nullptr;
	uint32_t *dist = nullptr;
	// per-vertex state byte: 0 not reached, 1 reached in the current step,
	// kVisited (BFS) reached in an earlier step; SSSP and multi-source use 0/1
	uint8_t *flag = nullptr;
	uint32_t *visited = nullptr;  // BFS: bitmap of the vertices reached in earlier steps
	uint32_t *cur = nullptr;      // frontier bitmap of the current step
	uint32_t *row_stamp = nullptr, *col_unvis = nullptr;
	// sparse-step queue, two halves each (QueueBufs): queue / qoff of
	// 2 * n_pad entries, tile_lo of 2 * (m / kSparseTile + 2)
	uint32_t *queue = nullptr, *qoff = nullptr, *tile_lo = nullptr;
	gphi::QueueBufs queue_bufs() const
	{
		return gphi::QueueBufs{queue, qoff, tile_lo, gphi::queue_stride(g.n_pad), gphi::tile_lo_stride(g.m)};
	}
	gphi::StepCtrl *ctrl = nullptr;
	uint4 *blk_cnt = nullptr;  // per update block: |F|, queue entries, degree sum
	uint2 *blk_base = nullptr; // per update block: queue and degree prefix
	uint2 *wave_cnt = nullptr; // per update wavefront: queue entries, degree sum
	uint32_t *live = nullptr;  // per-XCD lists of live chunks of a dense step
	uint32_t *live_cnt = nullptr;
	uint4 *trace = nullptr;    // per step: direction, |F|, sum deg(F), chunks swept
	// PageRank: one allocation holding rank, sum and contrib (n_pad each) and
	// the partial sums of the heavy pairs' pieces
	float *rank = nullptr, *sum = nullptr, *contrib = nullptr;
	// PageRank head bands (build_pr_bands; built once, independent of W and S):
	// the heads band by band (pr_src), small and light pairs (tail, first
	// edge in pr_src, count), heavy pairs {tail, first edge, count, first
	// piece}, each piece's heavy pair, and the multi-piece heavy pairs; all in
	// one allocation, pr_mem
	std::vector<gphi::PrBand> pr_bands;
	void *pr_mem = nullptr;
	uint32_t *pr_src = nullptr;
	uint32_t *pr_lt = nullptr, *pr_llo = nullptr;
	uint8_t *pr_lcnt = nullptr;
	uint4 *pr_hp = nullptr;
	uint32_t *pr_pc = nullptr, *pr_multi = nullptr;
	uint32_t pr_npieces = 0;
	uint64_t pr_bytes = 0;
	// multi-source BFS (batch.hip): per-vertex source masks and uint8 levels
	uint32_t batch_words = 0;  // 64-bit words per vertex the buffers were sized for
	// leading mask words that front / next (and the touched bitmap) are known
	// to hold as zero: a traversal that runs to its end leaves them so
	uint64_t batch_clean_words = 0;
	unsigned long long *b_front = nullptr, *b_seen = nullptr, *b_next = nullptr;
	uint8_t *b_level = nullptr;
	unsigned long long *b_sums = nullptr;
	unsigned long long *b_live = nullptr;  // sources with a non-empty frontier, then the per-block slots
	uint32_t *b_touch = nullptr;           // bitmap: vertices whose next mask got a bit this step
	uint32_t *b_scount = nullptr;          // in b_touch's allocation: new vertices per source, per slot
	uint32_t *b_aux = nullptr;             // sources to seed, their slots, per-slot start levels
	// multi-source SSSP (sssp_batch.hip): kSsspBatchMax distances per vertex
	// (vertex-major, n_pad rows), then the sources
	uint32_t *sb_dist = nullptr;
	// in-edge arrays of the pull steps (built once, independent of W and S)
	uint32_t *in_off = nullptr, *in_src = nullptr;
	// the weights of in_src's edges (build_in_weights): 1 byte when every
	// weight is below 256, else 4
	uint8_t *in_w8 = nullptr;
	uint32_t *in_w32 = nullptr;
	uint2 *pieces = nullptr;
	uint32_t npieces = 0;
	uint64_t in_bytes = 0;
};

namespace gphi {
int set_error(gphi_ctx *ctx, int code, const char *fmt, ...);
// true when an allocation failure injected by gphi_debug_fail_alloc is due now
bool take_injected_alloc_failure();
int hip_fail(gphi_ctx *ctx, hipError_t e, const char *what, const char *file, int line);

// device-side entry points (device.hip)
int upload_graph(gphi_ctx *ctx, const gphi_graph_desc *d, gphi_graph *gr);
void free_graph_device(gphi_graph *gr);
void free_traversal_state(gphi_graph *gr);
int retile_graph(gphi_graph *gr, uint32_t W, uint32_t S, int tile_order);
int build_in_edges(gphi_graph *gr);
int build_pr_bands(gphi_graph *gr);  // after build_in_edges
int build_in_weights(gphi_graph *gr);  // after build_in_edges; weighted graphs with CSR
int run_bfs(gphi_graph *gr, uint32_t src, int32_t *levels, int32_t *parents, gphi_stats *st);
int run_bfs_noalloc(gphi_graph *gr, uint32_t src, int32_t *levels_dev, int32_t *parents_dev, gphi_stats *st);
void noalloc_setup(int device);  // with `device` current, once per device (gphi_open)
int run_sssp(gphi_graph *gr, uint32_t src, uint32_t *dist, gphi_stats *st);
int run_pagerank(gphi_graph *gr, uint32_t iters, float damping, float init, float *rank,
                 float *ms);
int run_bellman_ford(gphi_graph *gr, uint32_t src, uint32_t max_sweeps, uint32_t *dist,
                     uint32_t *sweeps_run, float *ms);

int finish_tiles(gphi_ctx *ctx, gphi_graph *gr);
void free_chunk_summaries(DevGraph &G);  // meta and the 16-bit tile forms
int count_active_tiles(gphi_graph *gr, const int32_t *levels, const std::vector<uint32_t> &dense_steps,
                       std::vector<double> &e_act, std::vector<double> &t_act);
int tile_offsets(gphi_graph *gr, uint32_t *out);
int download_graph(gphi_graph *gr, uint32_t *deg, uint32_t *cols, uint32_t *cw, uint32_t *th,
                   uint32_t *tt, uint32_t *tw);

int ensure_traversal_state(gphi_graph *gr);
// the BFS level loop of run_bfs on gr's state buffers (set up by the caller)
int bfs_level_loop(gphi_graph *gr, uint32_t src, gphi_stats *st);
uint32_t live_list_capacity(uint64_t nchunks);  // entries of one XCD's live-chunk list
void launch_frontier_update_driven(gphi_graph *gr, uint32_t threshold);
int run_bfs_batch(gphi_graph *gr, const uint32_t *srcs, uint32_t k, uint8_t *levels, gphi_stats *st);
// the multi-source state of gr for nw mask words: traversal state, in-edge
// arrays, masks, levels (idempotent)
int batch_prepare(gphi_graph *gr, uint32_t nw);
// weighted SSSP from k <= kSsspBatchMax sources in one traversal
// (sssp_batch.hip); rows[i]: n distances of source i to copy out, or null
// (rows itself may be null)
int run_sssp_batch(gphi_graph *gr, const uint32_t *srcs, uint32_t k, uint32_t *const *rows, gphi_stats *st);
int sssp_batch_prepare(gphi_graph *gr);  // its state, in-edges and weights (idempotent)

// A multi-source traversal (batch.hip).  gphi_bfs_batch runs one to its end
// on the device; the serving loop advances one level by level, admitting new
// sources into free slots between levels and reading out the rows of the
// slots whose frontier has emptied.  Slot i is bit i of the masks.
constexpr uint32_t kBatchMaxSources = 256;
struct BatchRun {
	gphi_graph *gr = nullptr;
	uint32_t nw = 0;       // 64-bit mask words per vertex
	uint32_t cap = 0;      // slots (64 * nw)
	uint32_t used = 0;     // slots handed out (slots are not reused within a run)
	uint32_t levels = 0;   // levels expanded
	bool done = false;     // the union frontier is empty
	std::vector<uint32_t> slot_src, slot_start;  // per slot: source vertex, level it joined at
	uint64_t mask_words = 0;                     // n_pad * nw
	uint64_t prev_clean = 0;                     // gr->batch_clean_words when the run began
	unsigned long long live[4] = {0, 0, 0, 0};   // sources with a non-empty frontier (after step)
	unsigned long long pairs[2] = {0, 0};        // reached (source, vertex) pairs, their out-degrees
	uint32_t aux[3 * kBatchMaxSources] = {};     // staging of the sources to seed

	// clears the state and seeds k sources (slots 0..k-1); capacity picks the
	// mask width (1..256 sources)
	int begin(gphi_graph *g, uint32_t capacity, const uint32_t *srcs, uint32_t k);
	// one level, then k new sources join (slots used..used+k-1); reads back
	// the control block and the live mask
	int step(const uint32_t *srcs, uint32_t k);
	int run_to_end(gphi_stats *st);  // the remaining levels, on the device
	int read_totals();               // pairs[] from the device
	// levels of the given slots into host rows of n bytes (255 unreached)
	int copy_rows(const uint32_t *slots, uint32_t ns, uint8_t *const *rows);
	int seed(const uint32_t *srcs, uint32_t k, uint32_t start_level);
};

// generator + tool pipeline on the device (gen.hip)
int generate_graph(gphi_ctx *ctx, const gphi_rmat_spec *sp, uint32_t W, uint32_t S, int reorder,
                   const gphi_weight_spec *ws, int tile_order, gphi_graph *gr);
void glibc_rand_seq(uint32_t seed, uint64_t n, int32_t *out);
void random_perm_host(uint32_t n, uint64_t seed, uint32_t *perm);

// host text formats (formats.cpp)
// moves st[0..6] (degrees, csr, csr w, heads, tails, tile w, offsets) into a new host graph
int host_adopt(const gphi_graph_desc *d, std::vector<uint32_t> *st, gphi_host_graph **out);
int read_text_graph(gphi_ctx *ctx, const char *prefix, uint32_t W, uint32_t S, int weighted,
                    bool row_major, gphi_graph_desc *d, std::vector<uint32_t> *store);
}  // namespace gphi

#define GPHI_HIP(ctx, call)                                                    \
	do {                                                                       \
		hipError_t e_ = (call);                                                \
		if (e_ != hipSuccess) return gphi::hip_fail((ctx), e_, #call, __FILE__, __LINE__); \
	} while (0)

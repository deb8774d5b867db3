// noalloc.hip -- hybrid BFS without allocation (gphi_bfs_noalloc).
//
// No counterpart in the reference.  The traversal state of gphi_bfs (levels,
// parents, state bytes, bitmaps, queues, live-chunk lists, control block) is
// carved out of one fixed device array per GPU instead of buffers the library
// allocates, for callers that cannot call hipMalloc.  Every call on a GPU
// shares that GPU's array; a process-wide lock per device serialises the
// calls, and each call's kernels are done before it returns (the level loop
// reads its control block back), so the next call finds the array free.  The
// array keeps the last call's contents: everything that must start at zero is
// cleared by the traversal's own init kernels, none by hipMemsetAsync.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "internal.h"

namespace gphi {

constexpr size_t kNoAllocScratchBytes = 128ull << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mu;                        // guards g_scratch_ptr / g_scratch_err
void *g_scratch_ptr[kNoAllocMaxDevices];        // device address of `scratch` per device
std::string g_scratch_err[kNoAllocMaxDevices];  // why an address is missing
std::mutex g_scratch_use[kNoAllocMaxDevices];   // one no-alloc traversal per device at a time

// level and parent of every vertex < n into the caller's arrays; -1 for the
// vertices the traversal did not reach (state 0)
__global__ __launch_bounds__(kBlock) void bfs_copy_out_kernel(uint32_t n, const uint8_t *__restrict__ state,
                                                              const int32_t *__restrict__ level,
                                                              const int32_t *__restrict__ parent,
                                                              int32_t *__restrict__ level_out,
                                                              int32_t *__restrict__ parent_out)
{
	for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
		const bool reached = state[v] != 0;
		if (level_out) level_out[v] = reached ? level[v] : -1;
		if (parent_out) parent_out[v] = reached ? parent[v] : -1;
	}
}

// The state buffers of the BFS level loop laid out from `base` in 256-byte
// aligned pieces, the same sizes gphi_bfs allocates; returns the bytes they
// take.  base == nullptr only measures.
size_t carve_state(const DevGraph &G, uint8_t *base, gphi_graph *t)
{
	size_t off = 0;
	auto take = [&](size_t bytes) {
		void *p = base ? base + off : nullptr;
		off += (bytes + 255) / 256 * 256;
		return p;
	};
	const size_t nb = (size_t)G.n_pad * 4;
	const size_t nblk = G.n_pad / kPadVerts;
	t->level = (int32_t *)take(nb);
	t->parent = (int32_t *)take(nb);
	t->flag = (uint8_t *)take(G.n_pad);
	t->visited = (uint32_t *)take((size_t)G.n_words * 4);
	t->cur = (uint32_t *)take((size_t)G.n_words * 4);
	t->row_stamp = (uint32_t *)take(((size_t)G.side + 1) * 4);
	t->col_unvis = (uint32_t *)take(((size_t)G.side + 1) * 4);
	t->queue = (uint32_t *)take(2 * (size_t)queue_stride(G.n_pad) * 4);  // two halves (QueueBufs)
	t->qoff = (uint32_t *)take(2 * (size_t)queue_stride(G.n_pad) * 4);
	t->tile_lo = (uint32_t *)take(2 * (size_t)tile_lo_stride(G.m) * 4);
	t->ctrl = (StepCtrl *)take(sizeof(StepCtrl));
	t->live = (uint32_t *)take((size_t)live_list_capacity(G.nchunks) * kNumXcd * 4 + 4);
	t->live_cnt = (uint32_t *)take(kLiveCntBytes);
	t->trace = (uint4 *)take((GPHI_MAX_STEPS + 1) * sizeof(uint4));
	t->blk_cnt = (uint4 *)take(nblk * sizeof(uint4));
	t->blk_base = (uint2 *)take(nblk * sizeof(uint2));
	t->wave_cnt = (uint2 *)take(nblk * (kBlock / 64) * sizeof(uint2));
	return off;
}

}  // namespace

// Once per device, with that device current (called by gphi_open): the
// device address of `scratch`.  The array's host-side name is only a
// stand-in; kernels and copies use this address.
void noalloc_setup(int device)
{
	if (device < 0 || device >= kNoAllocMaxDevices) return;
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (g_scratch_ptr[device]) return;
	void *p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e != hipSuccess || !p) {
		g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
		return;
	}
	g_scratch_ptr[device] = p;
}

int run_bfs_noalloc(gphi_graph *gr, uint32_t src, int32_t *levels_dev, int32_t *parents_dev, gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	const DevGraph &G = gr->g;
	if (src >= G.n) return set_error(ctx, GPHI_E_RANGE, "source %u >= N %u", src, G.n);
	if (!G.has_csr || !G.has_tiles)
		return set_error(ctx, GPHI_E_STATE, "a hybrid traversal needs the CSR and the tile arrays");
	const int dev = ctx->device;
	if (dev < 0 || dev >= kNoAllocMaxDevices)
		return set_error(ctx, GPHI_E_STATE, "no scratch array for device %d", dev);
	uint8_t *base;
	{
		std::lock_guard<std::mutex> lk(g_scratch_mu);
		base = static_cast<uint8_t *>(g_scratch_ptr[dev]);
		if (!base)
			return set_error(ctx, GPHI_E_STATE, "scratch array of device %d unavailable: %s", dev,
			                 g_scratch_err[dev].empty() ? "not set up" : g_scratch_err[dev].c_str());
	}
	gphi_graph t = *gr;  // the graph's arrays, with the state buffers replaced below
	t.dist = nullptr;
	t.rank = t.sum = t.contrib = nullptr;
	t.b_front = t.b_seen = t.b_next = t.b_sums = t.b_live = nullptr;
	t.b_level = nullptr;
	const size_t need = carve_state(G, nullptr, &t);
	if (need > kNoAllocScratchBytes)
		return set_error(ctx, GPHI_E_RANGE,
		                 "graph (N=%u, M=%llu) needs %zu bytes of traversal state; the no-allocation BFS "
		                 "has a fixed %zu-byte scratch array per GPU (about 25 bytes per vertex)",
		                 G.n, (unsigned long long)G.m, need, kNoAllocScratchBytes);
	carve_state(G, base, &t);
	std::lock_guard<std::mutex> use(g_scratch_use[dev]);
	if (st) memset(st, 0, sizeof(*st));
	int rc = bfs_level_loop(&t, src, st);
	if (rc) return rc;
	hipStream_t s = ctx->stream;
	if (levels_dev || parents_dev) {
		const unsigned grid = (unsigned)std::min<uint64_t>(((uint64_t)G.n + kBlock - 1) / kBlock, 8192);
		bfs_copy_out_kernel<<<grid, kBlock, 0, s>>>(G.n, t.flag, t.level, t.parent, levels_dev, parents_dev);
	}
	GPHI_HIP(ctx, hipGetLastError());
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	return GPHI_OK;
}

}  // namespace gphi

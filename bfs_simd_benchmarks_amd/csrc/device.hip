// device.hip -- gfx950 kernels and their host drivers for libgphi.
//
// Hot path (reference apps/bfs_simd/bfs.cpp):
//   dense step  : tile sweep over the column-stripe COO (bfs_kernel_dense
//                 :181-245 driven by scheduler_dense/BFS_dense :246-388)
//   sparse step : CSR push from the frontier queue (BFS_kernel_sparse :510-711)
//   update      : frontier bookkeeping + direction rule (update_dense :91-179,
//                 update_sparse :467-509, to_dense :54-88, to_sparse :400-465,
//                 heuristic :829-834) fused into one vertex pass
// The same chunked tile sweep serves SSSP (apps/sssp_simd/sssp.cpp:344-445),
// PageRank (apps/pageRank_simd/page_rank.cpp:128-199) and hop relaxation
// (apps/bellman-ford_simd/bellman-ford.cpp:129-191).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>

#include "internal.h"

namespace gphi {

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
	for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
	return x;
}

// exclusive prefix sum across the wave; *total gets the wave sum
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t *total)
{
	uint32_t lane = lane_id();
	uint32_t inc = x;
	for (int o = 1; o < 64; o <<= 1) {
		uint32_t y = __shfl_up(inc, o, 64);
		if (lane >= (uint32_t)o) inc += y;
	}
	*total = __shfl(inc, 63, 64);
	return inc - x;
}

__device__ __forceinline__ bool test_bit(const uint32_t *__restrict__ bm, uint32_t v)
{
	return (bm[v >> 5] >> (v & 31)) & 1u;
}

// ---------------------------------------------------------------------------
// graph upload helpers
// ---------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t *p, uint64_t count, uint32_t value)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (; i < count; i += stride) p[i] = value;
}

// Chunk summary: range of head rows / tail columns of the real edges of each
// kChunk-edge chunk.  One block per chunk (grid-stride).
__global__ __launch_bounds__(kBlock) void chunk_meta_kernel(const uint32_t *__restrict__ heads,
                                                            const uint32_t *__restrict__ tails,
                                                            uint32_t n, uint32_t W, uint64_t nchunks,
                                                            ChunkMeta *__restrict__ meta)
{
	__shared__ uint32_t s[4][kBlock / 64];
	for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
		uint32_t rlo = UINT32_MAX, rhi = 0, clo = UINT32_MAX, chi = 0;
		const uint32_t *H = heads + c * kChunk, *T = tails + c * kChunk;
		for (uint32_t i = threadIdx.x; i < kChunk; i += kBlock) {
			uint32_t h = H[i];
			if (h >= n) continue;
			uint32_t r = h / W, col = T[i] / W;
			rlo = min(rlo, r);
			rhi = max(rhi, r);
			clo = min(clo, col);
			chi = max(chi, col);
		}
		for (int o = 32; o > 0; o >>= 1) {
			rlo = min(rlo, (uint32_t)__shfl_xor(rlo, o, 64));
			rhi = max(rhi, (uint32_t)__shfl_xor(rhi, o, 64));
			clo = min(clo, (uint32_t)__shfl_xor(clo, o, 64));
			chi = max(chi, (uint32_t)__shfl_xor(chi, o, 64));
		}
		uint32_t w = threadIdx.x >> 6;
		if (lane_id() == 0) {
			s[0][w] = rlo;
			s[1][w] = rhi;
			s[2][w] = clo;
			s[3][w] = chi;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			for (uint32_t k = 1; k < kBlock / 64; ++k) {
				rlo = min(rlo, s[0][k]);
				rhi = max(rhi, s[1][k]);
				clo = min(clo, s[2][k]);
				chi = max(chi, s[3][k]);
			}
			meta[c] = ChunkMeta{rlo, rhi, clo, chi};
		}
		__syncthreads();
	}
}

static int dmalloc(gphi_ctx *ctx, void **p, size_t bytes, uint64_t *acc)
{
	if (take_injected_alloc_failure()) {
		*p = nullptr;
		return set_error(ctx, GPHI_E_OOM, "hipMalloc(%zu bytes) failed: injected failure", bytes);
	}
	hipError_t e = hipMalloc(p, bytes ? bytes : 4);
	if (e != hipSuccess) {
		*p = nullptr;
		return set_error(ctx, GPHI_E_OOM, "hipMalloc(%zu bytes) failed: %s", bytes,
		                 hipGetErrorString(e));
	}
	if (acc) *acc += bytes;
	return GPHI_OK;
}

static unsigned grid_for(uint64_t items, unsigned per_block, unsigned cap)
{
	uint64_t g = (items + per_block - 1) / per_block;
	if (g < 1) g = 1;
	if (g > cap) g = cap;
	return (unsigned)g;
}

// A window: an aligned run of R tile rows (or columns) whose R*W vertices fit
// kWinWords bitmap words, i.e. 16-bit offsets (R*W <= 65536).
constexpr uint32_t kWinWords = 2048;

// tile rows per window for tile width W (0: no window fits)
static uint32_t dense_window_rows(uint32_t W)
{
	if (W == 0) return 0;
	if ((W & (W - 1)) == 0 && W <= kWinWords * 32) return kWinWords * 32 / W;
	return W <= (kWinWords - 4) * 32 ? (kWinWords - 4) * 32 / W : 0;
}

// first vertex of the window that holds tile row (column) lo
__host__ __device__ __forceinline__ uint32_t window_base(uint32_t lo, uint32_t R, uint32_t W)
{
	return (uint32_t)((uint64_t)(lo - lo % R) * W);
}

// 16-bit forms of the heads and tails of each chunk (head offsets from the
// window base of its first row, tail offsets from the smallest tail of each
// group of kPackGroup edges) and the flags saying which are exact.  One block
// per chunk; thread i handles edges i + kBlock * r, r = 0..7, so pass r is
// group r.
static_assert(kPackGroup == kBlock && kChunk == 8 * kPackGroup, "one group per pass of the block");
__global__ __launch_bounds__(kBlock) void pack_chunks_kernel(const uint32_t *__restrict__ heads,
                                                             const uint32_t *__restrict__ tails,
                                                             const ChunkMeta *__restrict__ meta, uint64_t nchunks,
                                                             uint32_t R, uint32_t W, uint16_t *__restrict__ h16,
                                                             uint16_t *__restrict__ t16, ChunkPack *__restrict__ pack)
{
	__shared__ uint32_t s_lo[8][kBlock / 64], s_hi[8][kBlock / 64], s_ok[kBlock / 64];
	__shared__ uint32_t s_base[8];
	const uint32_t wid = threadIdx.x >> 6;
	for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
		const ChunkMeta md = meta[c];
		const bool real = md.row_lo <= md.row_hi;
		const uint32_t hb = real ? window_base(md.row_lo, R, W) : 0u;
		uint32_t ok = real ? 1u : 0u;
		uint32_t t[8];
#pragma unroll
		for (int r = 0; r < 8; ++r) {
			const uint64_t e = c * kChunk + threadIdx.x + r * kBlock;
			const uint32_t ho = heads[e] - hb;
			if (ho > 0xFFFFu) ok = 0u;
			h16[e] = (uint16_t)ho;
			t[r] = tails[e];
			uint32_t lo = t[r], hi = t[r];
			for (int o = 32; o > 0; o >>= 1) {
				lo = min(lo, (uint32_t)__shfl_xor((int)lo, o, 64));
				hi = max(hi, (uint32_t)__shfl_xor((int)hi, o, 64));
			}
			if (lane_id() == 0) {
				s_lo[r][wid] = lo;
				s_hi[r][wid] = hi;
			}
		}
		for (int o = 32; o > 0; o >>= 1) ok &= (uint32_t)__shfl_xor((int)ok, o, 64);
		if (lane_id() == 0) s_ok[wid] = ok;
		__syncthreads();
		if (threadIdx.x < 8) {
			const uint32_t r = threadIdx.x;
			uint32_t lo = s_lo[r][0], hi = s_hi[r][0];
			for (int k = 1; k < (int)(kBlock / 64); ++k) {
				lo = min(lo, s_lo[r][k]);
				hi = max(hi, s_hi[r][k]);
			}
			s_base[r] = lo;
			s_hi[r][0] = hi - lo;  // span of the group
		}
		__syncthreads();
#pragma unroll
		for (int r = 0; r < 8; ++r) t16[c * kChunk + threadIdx.x + r * kBlock] = (uint16_t)(t[r] - s_base[r]);
		if (threadIdx.x == 0) {
			uint32_t okh = s_ok[0], okt = real ? 1u : 0u;
			for (int k = 1; k < (int)(kBlock / 64); ++k) okh &= s_ok[k];
			ChunkPack pk;
			pk.hbase = hb;
			pk.pad0 = pk.pad1 = 0u;
			for (int r = 0; r < 8; ++r) {
				pk.tbase[r] = s_base[r];
				if (s_hi[r][0] > 0xFFFFu) okt = 0u;
			}
			pk.flags = okh | (okt << 1);
			pack[c] = pk;
		}
		__syncthreads();
	}
}

// largest weight, then the 8-bit copy of the tile-order weights
__global__ __launch_bounds__(kBlock) void max_u32_kernel(const uint32_t *__restrict__ a, uint64_t n, uint32_t *out)
{
	uint32_t m = 0;
	for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
		m = max(m, a[i]);
	for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
	if (lane_id() == 0 && m) atomicMax(out, m);
}
__global__ __launch_bounds__(kBlock) void narrow_u8_kernel(const uint32_t *__restrict__ a, uint64_t n,
                                                           uint8_t *__restrict__ b)
{
	for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
		b[i] = (uint8_t)a[i];
}

void free_chunk_summaries(DevGraph &G)
{
	if (G.coo_w8) G.device_bytes -= G.m_pad, (void)hipFree(G.coo_w8);
	G.coo_w8 = nullptr;
	if (G.meta) G.device_bytes -= G.nchunks * sizeof(ChunkMeta) + 16, (void)hipFree(G.meta);
	if (G.pack) G.device_bytes -= G.nchunks * sizeof(ChunkPack) + 16, (void)hipFree(G.pack);
	if (G.heads16) G.device_bytes -= G.m_pad * 2, (void)hipFree(G.heads16);
	if (G.tails16) G.device_bytes -= G.m_pad * 2, (void)hipFree(G.tails16);
	G.meta = nullptr;
	G.pack = nullptr;
	G.heads16 = G.tails16 = nullptr;
}

// Pads the tile arrays to whole chunks with sentinel edges and builds the
// chunk summaries and the 16-bit tile forms.
int finish_tiles(gphi_ctx *ctx, gphi_graph *gr)
{
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	uint64_t padn = G.m_pad - G.m;
	if (padn) {
		fill_u32_kernel<<<grid_for(padn, kBlock, 1024), kBlock, 0, s>>>(G.heads + G.m, padn, G.n);
		fill_u32_kernel<<<grid_for(padn, kBlock, 1024), kBlock, 0, s>>>(G.tails + G.m, padn, G.n);
		if (G.coo_w) fill_u32_kernel<<<grid_for(padn, kBlock, 1024), kBlock, 0, s>>>(G.coo_w + G.m, padn, 0);
	}
	int rc;
	if ((rc = dmalloc(ctx, (void **)&G.meta, (size_t)G.nchunks * sizeof(ChunkMeta) + 16, &G.device_bytes)))
		return rc;
	if (G.nchunks)
		chunk_meta_kernel<<<grid_for(G.nchunks, 1, 65536), kBlock, 0, s>>>(G.heads, G.tails, G.n, G.W,
		                                                                  G.nchunks, G.meta);
	if ((rc = dmalloc(ctx, (void **)&G.pack, (size_t)G.nchunks * sizeof(ChunkPack) + 16, &G.device_bytes)))
		return rc;
	GPHI_HIP(ctx, hipMemsetAsync(G.pack, 0, (size_t)G.nchunks * sizeof(ChunkPack) + 16, s));
	const uint32_t R = dense_window_rows(G.W);
	if (R && G.nchunks) {
		if ((rc = dmalloc(ctx, (void **)&G.heads16, G.m_pad * 2, &G.device_bytes))) return rc;
		if ((rc = dmalloc(ctx, (void **)&G.tails16, G.m_pad * 2, &G.device_bytes))) return rc;
		pack_chunks_kernel<<<grid_for(G.nchunks, 1, 65536), kBlock, 0, s>>>(G.heads, G.tails, G.meta, G.nchunks, R,
		                                                                   G.W, G.heads16, G.tails16, G.pack);
	}
	if (G.coo_w && G.nchunks) {
		// SSSP sweeps read 1-byte weights when every weight fits
		uint32_t *d_max = nullptr, h_max = 0;
		if ((rc = dmalloc(ctx, (void **)&d_max, 4, nullptr))) return rc;
		GPHI_HIP(ctx, hipMemsetAsync(d_max, 0, 4, s));
		max_u32_kernel<<<grid_for(G.m_pad, kBlock, 4096), kBlock, 0, s>>>(G.coo_w, G.m_pad, d_max);
		GPHI_HIP(ctx, hipMemcpyAsync(&h_max, d_max, 4, hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		(void)hipFree(d_max);
		if (h_max < 256) {
			if ((rc = dmalloc(ctx, (void **)&G.coo_w8, G.m_pad, &G.device_bytes))) return rc;
			narrow_u8_kernel<<<grid_for(G.m_pad, kBlock, 8192), kBlock, 0, s>>>(G.coo_w, G.m_pad, G.coo_w8);
		}
	}
	GPHI_HIP(ctx, hipGetLastError());
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	return GPHI_OK;
}

int download_graph(gphi_graph *gr, uint32_t *deg, uint32_t *cols, uint32_t *cw, uint32_t *th,
                   uint32_t *tt, uint32_t *tw)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	size_t mb = (size_t)G.m * 4;
	if ((cols && !G.cols) || (cw && !G.csr_w) || (th && !G.heads) || (tw && !G.coo_w))
		return set_error(ctx, GPHI_E_STATE, "graph lacks a requested array");
	if (deg) GPHI_HIP(ctx, hipMemcpy(deg, G.deg, (size_t)G.n * 4, hipMemcpyDeviceToHost));
	if (cols) GPHI_HIP(ctx, hipMemcpy(cols, G.cols, mb, hipMemcpyDeviceToHost));
	if (cw) GPHI_HIP(ctx, hipMemcpy(cw, G.csr_w, mb, hipMemcpyDeviceToHost));
	if (th) GPHI_HIP(ctx, hipMemcpy(th, G.heads, mb, hipMemcpyDeviceToHost));
	if (tt) GPHI_HIP(ctx, hipMemcpy(tt, G.tails, mb, hipMemcpyDeviceToHost));
	if (tw) GPHI_HIP(ctx, hipMemcpy(tw, G.coo_w, mb, hipMemcpyDeviceToHost));
	return GPHI_OK;
}

// Byte-model counters: for each dense step k, edges (E) and distinct
// non-empty tiles (T) whose head row holds a level-k vertex and whose tail
// column still had an unvisited vertex when step k began.  Tiles are
// contiguous in the stored order, so a tile starts where (row, col) changes.
__global__ __launch_bounds__(kBlock) void active_tiles_kernel(
    const uint32_t *__restrict__ heads, const uint32_t *__restrict__ tails, uint64_t m, uint32_t W,
    const uint64_t *__restrict__ row_lv, const int64_t *__restrict__ col_max,
    const uint32_t *__restrict__ steps, uint32_t nsteps, unsigned long long *e_out,
    unsigned long long *t_out)
{
	__shared__ unsigned long long se[64], st_[64];
	for (uint32_t k = threadIdx.x; k < 64; k += blockDim.x) se[k] = st_[k] = 0;
	__syncthreads();
	for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
	     e += (uint64_t)gridDim.x * blockDim.x) {
		uint32_t r = heads[e] / W, c = tails[e] / W;
		bool start = e == 0 || heads[e - 1] / W != r || tails[e - 1] / W != c;
		uint64_t lv = row_lv[r];
		int64_t cm = col_max[c];
		for (uint32_t j = 0; j < nsteps; ++j) {
			uint32_t k = steps[j];
			if ((lv >> k & 1ull) && cm > (int64_t)k) {
				atomicAdd(&se[j], 1ull);
				if (start) atomicAdd(&st_[j], 1ull);
			}
		}
	}
	__syncthreads();
	for (uint32_t j = threadIdx.x; j < nsteps; j += blockDim.x) {
		if (se[j]) atomicAdd(e_out + j, se[j]);
		if (st_[j]) atomicAdd(t_out + j, st_[j]);
	}
}

int count_active_tiles(gphi_graph *gr, const int32_t *levels, const std::vector<uint32_t> &dense_steps,
                       std::vector<double> &e_act, std::vector<double> &t_act)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	const uint32_t ns = (uint32_t)dense_steps.size();
	e_act.assign(ns, 0.0);
	t_act.assign(ns, 0.0);
	if (!ns) return GPHI_OK;
	if (!G.has_tiles) return set_error(ctx, GPHI_E_STATE, "graph has no tile arrays");
	if (ns > 64) return set_error(ctx, GPHI_E_RANGE, "more than 64 dense steps");
	std::vector<uint64_t> row_lv(G.side, 0);
	std::vector<int64_t> col_max(G.side, -1);
	for (uint32_t v = 0; v < G.n; ++v) {
		int32_t l = levels[v];
		uint32_t b = v / G.W;
		if (l >= 0 && l < 64) row_lv[b] |= 1ull << l;
		int64_t key = l < 0 ? INT64_MAX : (int64_t)l;
		if (key > col_max[b]) col_max[b] = key;
	}
	uint64_t *d_lv = nullptr;
	int64_t *d_cm = nullptr;
	uint32_t *d_steps = nullptr;
	unsigned long long *d_cnt = nullptr;
	int rc;
	if ((rc = dmalloc(ctx, (void **)&d_lv, (size_t)G.side * 8, nullptr)) ||
	    (rc = dmalloc(ctx, (void **)&d_cm, (size_t)G.side * 8, nullptr)) ||
	    (rc = dmalloc(ctx, (void **)&d_steps, 64 * 4, nullptr)) ||
	    (rc = dmalloc(ctx, (void **)&d_cnt, 128 * 8, nullptr))) {
		(void)hipFree(d_lv);
		(void)hipFree(d_cm);
		(void)hipFree(d_steps);
		return rc;
	}
	hipError_t e = hipMemcpy(d_lv, row_lv.data(), (size_t)G.side * 8, hipMemcpyHostToDevice);
	if (e == hipSuccess) e = hipMemcpy(d_cm, col_max.data(), (size_t)G.side * 8, hipMemcpyHostToDevice);
	if (e == hipSuccess) e = hipMemcpy(d_steps, dense_steps.data(), ns * 4, hipMemcpyHostToDevice);
	if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, 128 * 8, s);
	std::vector<unsigned long long> h(128, 0);
	if (e == hipSuccess) {
		active_tiles_kernel<<<grid_for(G.m, kBlock, 8192), kBlock, 0, s>>>(G.heads, G.tails, G.m, G.W, d_lv,
		                                                               d_cm, d_steps, ns, d_cnt, d_cnt + 64);
		e = hipGetLastError();
	}
	if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_cnt, 128 * 8, hipMemcpyDeviceToHost, s);
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	(void)hipFree(d_lv);
	(void)hipFree(d_cm);
	(void)hipFree(d_steps);
	(void)hipFree(d_cnt);
	if (e != hipSuccess) return hip_fail(ctx, e, "active tile count", __FILE__, __LINE__);
	for (uint32_t j = 0; j < ns; ++j) {
		e_act[j] = (double)h[j];
		t_act[j] = (double)h[64 + j];
	}
	return GPHI_OK;
}

int upload_graph(gphi_ctx *ctx, const gphi_graph_desc *d, gphi_graph *gr)
{
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	G.n = d->n;
	G.m = d->m;
	G.W = d->tile_width;
	G.S = d->stripe_length;
	G.side = G.n % G.W ? G.n / G.W + 1 : G.n / G.W;
	G.ntiles = (uint64_t)G.side * G.side;
	G.tile_order = d->tile_order;
	G.n_pad = (uint32_t)(((uint64_t)G.n + 1 + kPadVerts - 1) / kPadVerts * kPadVerts);
	G.n_words = G.n_pad / 32;  // the update pass writes whole words up to n_pad
	G.nchunks = (G.m + kChunk - 1) / kChunk;
	G.m_pad = G.nchunks * kChunk;
	G.has_csr = d->csr_edges != nullptr;
	G.has_tiles = d->tile_heads != nullptr && d->tile_tails != nullptr;
	G.weighted = (d->csr_weights != nullptr || !G.has_csr) &&
	             (d->tile_weights != nullptr || !G.has_tiles) &&
	             (d->csr_weights || d->tile_weights);
	int rc;
	uint64_t *acc = &G.device_bytes;

	// degrees (padded) and CSR row offsets
	if ((rc = dmalloc(ctx, (void **)&G.deg, (size_t)G.n_pad * 4, acc))) return rc;
	GPHI_HIP(ctx, hipMemsetAsync(G.deg, 0, (size_t)G.n_pad * 4, s));
	GPHI_HIP(ctx, hipMemcpyAsync(G.deg, d->degrees, (size_t)G.n * 4, hipMemcpyHostToDevice, s));
	if (G.has_csr) {
		std::vector<uint32_t> off((size_t)G.n + 1);
		uint64_t a = 0;
		for (uint32_t v = 0; v < G.n; ++v) {
			off[v] = (uint32_t)a;
			a += d->degrees[v];
		}
		off[G.n] = (uint32_t)a;
		if (a != G.m)
			return set_error(ctx, GPHI_E_FORMAT, "sum of degrees %llu != M %llu",
			                 (unsigned long long)a, (unsigned long long)G.m);
		if ((rc = dmalloc(ctx, (void **)&G.row_off, off.size() * 4, acc))) return rc;
		GPHI_HIP(ctx, hipMemcpy(G.row_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
		if ((rc = dmalloc(ctx, (void **)&G.cols, (size_t)G.m * 4, acc))) return rc;
		GPHI_HIP(ctx, hipMemcpyAsync(G.cols, d->csr_edges, (size_t)G.m * 4,
		                             hipMemcpyHostToDevice, s));
		if (d->csr_weights) {
			if ((rc = dmalloc(ctx, (void **)&G.csr_w, (size_t)G.m * 4, acc))) return rc;
			GPHI_HIP(ctx, hipMemcpyAsync(G.csr_w, d->csr_weights, (size_t)G.m * 4,
			                             hipMemcpyHostToDevice, s));
		}
	}
	if (G.has_tiles) {
		size_t pb = (size_t)G.m_pad * 4;
		if ((rc = dmalloc(ctx, (void **)&G.heads, pb, acc))) return rc;
		if ((rc = dmalloc(ctx, (void **)&G.tails, pb, acc))) return rc;
		GPHI_HIP(ctx, hipMemcpyAsync(G.heads, d->tile_heads, (size_t)G.m * 4,
		                             hipMemcpyHostToDevice, s));
		GPHI_HIP(ctx, hipMemcpyAsync(G.tails, d->tile_tails, (size_t)G.m * 4,
		                             hipMemcpyHostToDevice, s));
		if (d->tile_weights) {
			if ((rc = dmalloc(ctx, (void **)&G.coo_w, pb, acc))) return rc;
			GPHI_HIP(ctx, hipMemcpyAsync(G.coo_w, d->tile_weights, (size_t)G.m * 4,
			                             hipMemcpyHostToDevice, s));
		}
		if ((rc = finish_tiles(ctx, gr))) return rc;
	}
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	return GPHI_OK;
}

// drops the lazily allocated traversal buffers (they depend on W and S)
void free_traversal_state(gphi_graph *gr)
{
	void **ps[] = {(void **)&gr->level, (void **)&gr->parent, (void **)&gr->dist, (void **)&gr->flag,
	               (void **)&gr->visited, (void **)&gr->cur, (void **)&gr->row_stamp, (void **)&gr->col_unvis,
	               (void **)&gr->queue, (void **)&gr->qoff, (void **)&gr->tile_lo, (void **)&gr->ctrl,
	               (void **)&gr->blk_cnt, (void **)&gr->blk_base, (void **)&gr->wave_cnt, (void **)&gr->live, (void **)&gr->live_cnt,
	               (void **)&gr->trace,
	               (void **)&gr->rank, (void **)&gr->b_front,
	               (void **)&gr->b_seen, (void **)&gr->b_next, (void **)&gr->b_level, (void **)&gr->b_sums,
	               (void **)&gr->b_live, (void **)&gr->b_touch, (void **)&gr->b_aux, (void **)&gr->sb_dist};
	for (void **p : ps) {
		if (*p) (void)hipFree(*p);
		*p = nullptr;
	}
	gr->sum = gr->contrib = nullptr;  // inside rank's allocation
	gr->b_scount = nullptr;  // inside b_touch's allocation
	gr->batch_words = 0;
	gr->batch_clean_words = 0;
}

void free_graph_device(gphi_graph *gr)
{
	DevGraph &G = gr->g;
	free_traversal_state(gr);
	void *in[] = {gr->in_off, gr->in_src, gr->pieces, gr->pr_mem, gr->in_w8, gr->in_w32};
	for (void *p : in)
		if (p) (void)hipFree(p);
	gr->in_off = gr->in_src = nullptr;
	gr->in_w8 = nullptr;
	gr->in_w32 = nullptr;
	gr->pieces = nullptr;
	gr->pr_mem = nullptr;
	gr->pr_src = nullptr;
	gr->pr_bands.clear();
	if (gr->base) return;  // a view: the graph arrays are its base's
	void *ps[] = {G.row_off, G.cols, G.csr_w, G.deg, G.heads, G.tails, G.coo_w, G.coo_w8, G.meta, G.pack, G.heads16, G.tails16};
	for (void *p : ps)
		if (p) (void)hipFree(p);
}


// Capacity of one XCD's live-chunk list: chunk selection hands out chunks in
// groups of 64 (a wavefront's), group g to XCD g % 8.
static uint32_t live_list_cap(uint64_t nchunks)
{
	const uint64_t groups = (nchunks + kBlock - 1) / kBlock;
	return (uint32_t)((groups + kNumXcd - 1) / kNumXcd * kBlock);
}

// ---------------------------------------------------------------------------
// traversal state
// ---------------------------------------------------------------------------
static int alloc_state(gphi_graph *gr);

// The traversal buffers, all or none: a failed allocation frees the ones
// already made, so no later call finds gr->level set beside a missing buffer
// and launches kernels on it.
static int ensure_state(gphi_graph *gr)
{
	if (gr->level) return GPHI_OK;
	const int rc = alloc_state(gr);
	if (rc) free_traversal_state(gr);
	return rc;
}

static int alloc_state(gphi_graph *gr)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	int rc;
	size_t nb = (size_t)G.n_pad * 4;
	if ((rc = dmalloc(ctx, (void **)&gr->level, nb, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->parent, nb, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->flag, G.n_pad, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->visited, (size_t)G.n_words * 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->cur, (size_t)G.n_words * 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->row_stamp, ((size_t)G.side + 1) * 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->col_unvis, ((size_t)G.side + 1) * 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->queue, 2 * (size_t)queue_stride(G.n_pad) * 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->qoff, 2 * (size_t)queue_stride(G.n_pad) * 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->tile_lo, 2 * (size_t)tile_lo_stride(G.m) * 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->ctrl, sizeof(StepCtrl), nullptr))) return rc;
	const size_t list_cap = live_list_cap(G.nchunks);
	if ((rc = dmalloc(ctx, (void **)&gr->live, list_cap * kNumXcd * 4 + 4, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->live_cnt, kLiveCntBytes, nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->trace, (GPHI_MAX_STEPS + 1) * sizeof(uint4), nullptr))) return rc;
	const size_t nblk = G.n_pad / kPadVerts;
	if ((rc = dmalloc(ctx, (void **)&gr->blk_cnt, nblk * sizeof(uint4), nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->blk_base, nblk * sizeof(uint2), nullptr))) return rc;
	if ((rc = dmalloc(ctx, (void **)&gr->wave_cnt, nblk * (kBlock / 64) * sizeof(uint2), nullptr))) return rc;
	GPHI_HIP(ctx, hipMemsetAsync(gr->flag, 0, G.n_pad, ctx->stream));
	GPHI_HIP(ctx, hipMemsetAsync(gr->ctrl, 0, sizeof(StepCtrl), ctx->stream));
	GPHI_HIP(ctx, hipMemsetAsync(gr->live_cnt, 0, kLiveCntBytes, ctx->stream));
	return GPHI_OK;
}

// One launch for the per-traversal clears: state bytes, visited and frontier
// bitmaps, live-list counters, and the tile-row stamps / column counts
// (col_unvis[c] = vertices of tail block c; row_stamp[r] = none).
__global__ __launch_bounds__(kBlock) void init_state_kernel(uint32_t n, uint32_t W, uint32_t side, uint32_t n_pad,
                                                            uint32_t n_words, uint4 *__restrict__ flag16,
                                                            uint32_t *__restrict__ visited, uint32_t *__restrict__ cur,
                                                            uint32_t *__restrict__ live_cnt,
                                                            uint32_t *__restrict__ row_stamp,
                                                            uint32_t *__restrict__ col_unvis)
{
	const uint32_t nf = n_pad / 16;  // n_pad is a multiple of 8192
	const uint32_t nl = (uint32_t)(kLiveCntBytes / 4);
	for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < max(max(nf, n_words), max(side, nl));
	     i += gridDim.x * kBlock) {
		if (i < nf) flag16[i] = make_uint4(0, 0, 0, 0);
		if (i < n_words) {
			cur[i] = 0;
			if (visited) visited[i] = 0;
		}
		if (i < side) {
			row_stamp[i] = UINT32_MAX;
			col_unvis[i] = (i + 1 == side) ? n - i * W : W;
		}
		if (i < nl) live_cnt[i] = 0;
	}
}

// Source set-up, one block: level/parent/dist of src, its bits, the first
// queue (the reference's first step is always sparse, bfs.cpp:795-817).
template <bool kBfs>
__global__ void init_source_kernel(uint32_t src, uint32_t W, const uint32_t *__restrict__ deg,
                                   int32_t *level, int32_t *parent, uint32_t *dist,
                                   uint8_t *flag, uint32_t *visited, uint32_t *cur, uint32_t *row_stamp,
                                   uint32_t *col_unvis, uint32_t *queue, uint32_t *qoff,
                                   uint32_t *tile_lo, StepCtrl *ctrl, uint4 *trace)
{
	uint32_t d = deg[src];
	uint32_t ntile = (d + kSparseTile - 1) / kSparseTile;
	for (uint32_t k = threadIdx.x; k < ntile + 1; k += blockDim.x) tile_lo[k] = 0;
	if (threadIdx.x == 0) {
		if (kBfs) {
			level[src] = 0;
			parent[src] = (int32_t)src;
			flag[src] = kVisited;
			visited[src >> 5] |= 1u << (src & 31);
			col_unvis[src / W] -= 1;
		} else {
			dist[src] = 0;
		}
		cur[src >> 5] |= 1u << (src & 31);
		row_stamp[src / W] = 0;
		queue[0] = src;
		qoff[0] = 0;
		ctrl->fsize = 1;
		ctrl->fdeg = d;
		ctrl->qsize = d ? 1 : 0;
		ctrl->dir = 0;
		ctrl->done = 0;
		ctrl->blocks_done = 0;
		ctrl->facc = 0;
		ctrl->qpack = 0;
		ctrl->changed = 0;
		ctrl->chunks = 0;
		ctrl->level = 0;
		ctrl->queued = 0;
		ctrl->col_stale = 0;
		ctrl->reached = 1;
		ctrl->traversed = d;
		trace[0] = make_uint4(0u, 1u, d, 0u);
	}
}

// ---------------------------------------------------------------------------
// dense step: chunked tile sweep
// ---------------------------------------------------------------------------
// Chunk selection (one lane per chunk): a chunk is live when a head row of
// it holds a frontier vertex (the reference's is_active_side test,
// bfs.cpp:275) and, for BFS, a tail column of it still has an unvisited
// vertex.  Live chunks are appended to one list per XCD (the XCD that runs
// blocks b with b % 8 == x), 64 consecutive chunks at a time.
template <bool kColSkip>
__device__ __forceinline__ void chunk_select_block(
    uint32_t bid, const ChunkMeta *__restrict__ meta, uint64_t nchunks, const uint32_t *__restrict__ row_stamp,
    const uint32_t *__restrict__ col_unvis, const StepCtrl *__restrict__ ctrl, uint32_t *__restrict__ lists,
    uint32_t list_cap, uint32_t *__restrict__ counts)
{
	// a sparse step (or a finished traversal) selects nothing; the counts
	// stay 0, so the dense sweep launched after it returns at once
	if (ctrl->done || ctrl->dir != 1) return;
	const uint32_t level = ctrl->level;
	const uint64_t c = (uint64_t)bid * kBlock + threadIdx.x;
	bool live = false;
	if (c < nchunks) {
		const ChunkMeta md = meta[c];
		if (md.row_lo <= md.row_hi) {
			// up to 16 rows / columns are checked with independent loads;
			// wider ranges count as live
			const uint32_t nr = md.row_hi - md.row_lo + 1;
			if (nr > 16) {
				live = true;
			} else {
				bool any = false;
#pragma unroll
				for (uint32_t k = 0; k < 16; ++k)
					any |= k < nr && row_stamp[md.row_lo + k] == level;
				live = any;
			}
			if (kColSkip && live) {
				const uint32_t nc = md.col_hi - md.col_lo + 1;
				if (nc <= 16) {
					bool any = false;
#pragma unroll
					for (uint32_t k = 0; k < 16; ++k)
						any |= k < nc && col_unvis[md.col_lo + k] != 0;
					live = any;
				}
			}
		}
	}
	// the 256 chunks of block b go to XCD b % 8: every region of the edge
	// order is shared by all XCDs, so the lists stay balanced however the live
	// chunks cluster (each XCD still walks its chunks in edge order), and a
	// block appends with one atomic
	__shared__ uint32_t s_cnt[kBlock / 64 + 1];
	const unsigned long long m = __ballot(live);
	const uint32_t wid = threadIdx.x >> 6;
	if (lane_id() == 0) s_cnt[wid] = (uint32_t)__popcll(m);
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t tot = 0;
		for (uint32_t k = 0; k < kBlock / 64; ++k) {
			const uint32_t n = s_cnt[k];
			s_cnt[k] = tot;
			tot += n;
		}
		s_cnt[kBlock / 64] = tot ? atomicAdd(counts + (bid % kNumXcd) * kLiveCntStride, tot) : 0u;
	}
	__syncthreads();
	if (live)
		lists[(uint64_t)(bid % kNumXcd) * list_cap + s_cnt[kBlock / 64] + s_cnt[wid] +
		      __popcll(m & ((1ull << lane_id()) - 1ull))] = (uint32_t)c;
}

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
// streaming (non-temporal) 16-byte load: the tile arrays are read once per
// sweep and should not evict the bitmaps and state arrays from L2 / MALL
__device__ __forceinline__ uint4 ld_stream(const uint4 *p)
{
	const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t *>(p));
	return make_uint4(v.x, v.y, v.z, v.w);
}

// Frontier-bit lookups of the dense sweep go through LDS.  A chunk's heads
// lie in its tile-row range [row_lo, row_hi] (a column-stripe chunk stays in
// one stripe), so the frontier bitmap of an aligned run of R tile rows
// (kWinWords words) is copied into LDS once and serves every chunk of the
// block inside it; a chunk outside any such run looks its bits up in global
// memory.  Gathers cost address-unit cycles per lane, LDS reads do not.
constexpr uint32_t kDenseRun = 8;  // consecutive live-list entries per block run

// window [wlo, wlo + wlen) of bitmap words holding tile rows row_lo..row_hi,
// or false; wlo is a multiple of 4 and wlen <= kWinWords
__device__ __forceinline__ bool window_of(uint32_t row_lo, uint32_t row_hi, uint32_t R, uint32_t W,
                                          uint32_t n_words, uint32_t &wlo, uint32_t &wlen)
{
	if (R == 0) return false;
	const uint32_t r0 = row_lo - row_lo % R;
	if (row_hi - r0 >= R) return false;
	const uint64_t b0 = (uint64_t)r0 * W, b1 = (uint64_t)(r0 + R) * W;
	wlo = (uint32_t)(b0 >> 5) & ~3u;
	const uint32_t whi = (uint32_t)min((b1 + 31) >> 5, (uint64_t)n_words);
	wlen = whi - wlo;
	return true;
}

// block-wide copy of bitmap words [wlo, wlo + wlen) into LDS (n_words is a
// multiple of 4, so the rounded-up last uint4 stays inside the bitmap)
__device__ __forceinline__ void load_window(uint4 *s_win, const uint32_t *cur, uint32_t wlo, uint32_t wlen,
                                            uint32_t tid)
{
	const uint4 *src = reinterpret_cast<const uint4 *>(cur + wlo);
	const uint32_t nq = (wlen + 3) / 4;
	// one uint4 in flight per lane: window changes are rare, registers are not
#pragma nounroll
	for (uint32_t q = tid; q < nq; q += kBlock) s_win[q] = src[q];
}

// frontier bits of 8 heads; in the window every real head of the chunk is
// inside it, a padding head (id n) outside it reads as not in the frontier
__device__ __forceinline__ uint32_t frontier_bits(const uint32_t (&hh)[8], bool inw, const uint32_t *s_win,
                                                  uint32_t wlo, uint32_t wlen, const uint32_t *cur)
{
	uint32_t cw[8];
	if (inw) {
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			const uint32_t wi = (hh[j] >> 5) - wlo;
			cw[j] = wi < wlen ? s_win[wi] : 0u;
		}
	} else {
#pragma unroll
		for (int j = 0; j < 8; ++j) cw[j] = cur[hh[j] >> 5];
	}
	uint32_t hit = 0;
#pragma unroll
	for (int j = 0; j < 8; ++j) hit |= ((cw[j] >> (hh[j] & 31)) & 1u) << j;
	return hit;
}

// Software-pipelined walk of this XCD's live-chunk list (runs of kDenseRun
// entries per block), 8 edges per lane per chunk.  While chunk k finishes
// (state bytes of its tails, writes), the frontier bits of chunk k+1 are
// looked up, and then its tails and the heads of chunk k+2 are requested, so
// each iteration waits on one HBM latency and one L2 round trip.  Every
// global load is unconditional (a lane with nothing to fetch reads a cached
// dummy word) so the compiler can count them.  Tails of a 4-edge quad are
// fetched only when one of its heads is in the frontier.
// Heads and tails come in their 16-bit forms wherever those are exact (half
// the bytes and memory requests); lane tid always owns edges 4tid..4tid+3
// and 1024+4tid..1024+4tid+3 of the chunk, in either form, so the tails of
// a wavefront's two quads decode from its two tail groups.
// One byte per vertex says whether a tail is new (state 0): "visited before
// this step" and "reached in this step" live in the same byte, so a
// candidate costs one lookup.  Every lane that sees state 0 writes parent and
// state 1 (benign races: every writer stores a valid frontier head and the
// same state; an earlier step's kVisited is never overwritten, since no lane
// of this step reads state 0 for such a vertex).
__device__ __forceinline__ void unpack8(const uint4 a, const uint4 b, uint32_t (&x)[8])
{
	x[0] = a.x, x[1] = a.y, x[2] = a.z, x[3] = a.w, x[4] = b.x, x[5] = b.y, x[6] = b.z, x[7] = b.w;
}

typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint2 ld_stream8(const uint2 *p)
{
	const u32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x2_t *>(p));
	return make_uint2(v.x, v.y);
}

// Raw edge words of one array of chunk c as loaded: 8 ids, or (16-bit form)
// 4 words of two offsets each.  q0 / q1: fetch the first / second quad (a
// lane that skips one reads the array's first bytes instead).
__device__ __forceinline__ void load_edges(bool p16, const uint32_t *__restrict__ a32,
                                           const uint16_t *__restrict__ a16, uint64_t c, uint32_t tid, bool q0,
                                           bool q1, uint32_t (&raw)[8])
{
	if (p16) {
		const uint2 *A = reinterpret_cast<const uint2 *>(a16 + c * kChunk);
		const uint2 *D = reinterpret_cast<const uint2 *>(a16);
		const uint2 x = ld_stream8(q0 ? A + tid : D), y = ld_stream8(q1 ? A + tid + kBlock : D);
		raw[0] = x.x, raw[1] = x.y, raw[2] = y.x, raw[3] = y.y;
	} else {
		const uint4 *A = reinterpret_cast<const uint4 *>(a32 + c * kChunk);
		const uint4 *D = reinterpret_cast<const uint4 *>(a32);
		unpack8(ld_stream(q0 ? A + tid : D), ld_stream(q1 ? A + tid + kBlock : D), raw);
	}
}

// raw edge words -> 8 vertex ids (in place)
// (base0 for the lane's first quad, base1 for its second)
__device__ __forceinline__ void decode_edges(bool p16, uint32_t base0, uint32_t base1, uint32_t (&x)[8])
{
	if (p16) {
		const uint32_t w0 = x[0], w1 = x[1], w2 = x[2], w3 = x[3];
		x[0] = base0 + (w0 & 0xFFFFu), x[1] = base0 + (w0 >> 16);
		x[2] = base0 + (w1 & 0xFFFFu), x[3] = base0 + (w1 >> 16);
		x[4] = base1 + (w2 & 0xFFFFu), x[5] = base1 + (w2 >> 16);
		x[6] = base1 + (w3 & 0xFFFFu), x[7] = base1 + (w3 >> 16);
	}
}

// what the sweep needs of a chunk's ChunkPack: head base, flags, and the
// tail bases of this wavefront's two edge groups (wv and 4 + wv)
struct PackSlot {
	uint32_t hbase, flags, tb0, tb1;
};
__device__ __forceinline__ PackSlot pack_slot(const ChunkPack *__restrict__ pack, uint32_t c, uint32_t wv)
{
	return PackSlot{pack[c].hbase, pack[c].flags, pack[c].tbase[wv], pack[c].tbase[4 + wv]};
}

__global__ __launch_bounds__(kBlock) void bfs_dense_kernel(
    const uint32_t *__restrict__ heads, const uint32_t *__restrict__ tails, const uint16_t *__restrict__ heads16,
    const uint16_t *__restrict__ tails16, const ChunkMeta *__restrict__ meta, const ChunkPack *__restrict__ pack,
    const uint32_t *__restrict__ lists, uint32_t list_cap, const uint32_t *__restrict__ counts,
    const uint32_t *__restrict__ cur, const uint32_t *__restrict__ visited, uint32_t n_words, uint32_t W,
    uint32_t R, uint8_t *__restrict__ state, int32_t *__restrict__ parent)
{
	__shared__ uint4 s_win4[kWinWords / 4];   // frontier bits of chunk k+1's rows
	__shared__ uint4 s_vwin4[kWinWords / 4];  // visited / claimed bits of chunk k's columns
	const uint32_t *s_win = reinterpret_cast<const uint32_t *>(s_win4);
	uint32_t *s_vwin = reinterpret_cast<uint32_t *>(s_vwin4);
	const uint32_t xcd = blockIdx.x % kNumXcd, nb = gridDim.x / kNumXcd;
	const uint32_t end = counts[xcd * kLiveCntStride];
	// runs of kDenseRun consecutive list entries, run r of the XCD to block
	// r % nb: chunks of a run share a window, while the blocks of an XCD still
	// sweep one narrow region of the edge arrays at a time (much longer runs
	// scatter the blocks over the arrays and are slower)
	const uint32_t skip = (nb - 1) * kDenseRun;
	auto nxt = [&](uint32_t p) { return ((p + 1) % kDenseRun) ? p + 1 : p + 1 + skip; };
	uint32_t pos = (blockIdx.x / kNumXcd) * kDenseRun;
	if (pos >= end) return;
	const uint32_t *lst = lists + (uint64_t)xcd * list_cap;
	const uint32_t tid = threadIdx.x;
	const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));  // wavefront in the block
	static_assert(kPackGroup == 4 * 64, "a wavefront's quad row is one tail group");
	// first words of the windows held in LDS (block-uniform)
	uint32_t win_lo = UINT32_MAX, vwin_lo = UINT32_MAX;
	// list entries: chunks k, k+1, k+2 (an absent one repeats chunk k's id, so
	// every address stays valid; its results are masked off)
	const uint32_t c0 = lst[pos];
	uint32_t p1 = nxt(pos), p2 = nxt(p1);
	bool v1 = p1 < end, v2 = p2 < end;
	uint32_t c1 = v1 ? lst[p1] : c0;
	uint32_t c2 = v2 ? lst[p2] : c0;
	// prologue: heads and frontier bits of chunk k, then heads of k+1 and tails of k
	PackSlot pk = pack_slot(pack, c0, wv), pk1 = pack_slot(pack, c1, wv);
	uint32_t h[8], hn[8], t[8];
	load_edges(pk.flags & 1u, heads, heads16, c0, tid, true, true, h);
	decode_edges(pk.flags & 1u, pk.hbase, pk.hbase, h);
	uint32_t hit;
	{
		const ChunkMeta m0 = meta[c0];
		uint32_t wlo = 0, wlen = 0;
		const bool inw = window_of(m0.row_lo, m0.row_hi, R, W, n_words, wlo, wlen);
		if (inw) {
			load_window(s_win4, cur, wlo, wlen, tid);
			__syncthreads();
			win_lo = wlo;
		}
		hit = frontier_bits(h, inw, s_win, wlo, wlen, cur);
	}
	load_edges(pk1.flags & 1u, heads, heads16, c1, tid, true, true, hn);
	load_edges(pk.flags & 2u, tails, tails16, c0, tid, hit & 0x0Fu, hit & 0xF0u, t);
	uint32_t ck = c0;  // chunk k
	for (;;) {
		// windows (block-uniform): frontier rows of chunk k+1, visited columns of chunk k
		bool inw_n = false;
		uint32_t wlo_n = 0, wlen_n = 0, vwlo = 0, vwlen = 0;
		if (v1) {
			const ChunkMeta mn = meta[c1];
			inw_n = window_of(mn.row_lo, mn.row_hi, R, W, n_words, wlo_n, wlen_n);
		}
		const ChunkMeta mk = meta[ck];
		const bool vin = window_of(mk.col_lo, mk.col_hi, R, W, n_words, vwlo, vwlen);
		const bool ld_c = inw_n && wlo_n != win_lo, ld_v = vin && vwlo != vwin_lo;
		if (ld_c || ld_v) {
			__syncthreads();  // every wave is done with the old windows
			if (ld_c) {
				load_window(s_win4, cur, wlo_n, wlen_n, tid);
				win_lo = wlo_n;
			}
			if (ld_v) {
				load_window(s_vwin4, visited, vwlo, vwlen, tid);
				vwin_lo = vwlo;
			}
			__syncthreads();
		}
		decode_edges(pk.flags & 2u, pk.tb0, pk.tb1, t);
		decode_edges(pk1.flags & 1u, pk1.hbase, pk1.hbase, hn);
		// stage A: states of chunk k's new frontier-edge tails, frontier bits of
		// chunk k+1.  In the column window each frontier edge claims its tail's
		// bit: a tail visited in an earlier step (bit set when the window was
		// loaded) or already claimed by another edge of this block in this step
		// needs no state lookup, so a block looks each new tail up once.
		uint32_t need = hit;
		if (vin) {
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				const uint32_t wi = (t[j] >> 5) - vwlo, b = 1u << (t[j] & 31);
				if (((hit >> j) & 1u) && wi < vwlen && (atomicOr(s_vwin + wi, b) & b)) need &= ~(1u << j);
			}
		}
		uint32_t sv[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) sv[j] = state[((need >> j) & 1u) ? t[j] : 0u];
		const uint32_t hit_n = v1 ? frontier_bits(hn, inw_n, s_win, wlo_n, wlen_n, cur) : 0u;
		uint32_t cand = 0;
#pragma unroll
		for (int j = 0; j < 8; ++j) cand |= (sv[j] == 0 ? 1u : 0u) << j;
		cand &= need;
		// stage B: chunk k's writes (few; issued first, so chunk k's heads and
		// tails are dead before the next requests take their registers)
#pragma unroll
		for (int j = 0; j < 8; ++j)
			if ((cand >> j) & 1u) {
				parent[t[j]] = (int32_t)h[j];
				state[t[j]] = 1;
			}
		// stage C: the HBM requests of the next iteration: tails of k+1, heads of k+2
		const PackSlot pk2 = pack_slot(pack, c2, wv);
		uint32_t tn[8], h2[8];
		load_edges(pk1.flags & 2u, tails, tails16, c1, tid, hit_n & 0x0Fu, hit_n & 0xF0u, tn);
		load_edges(pk2.flags & 1u, heads, heads16, c2, tid, true, true, h2);
		const uint32_t p3 = nxt(p2);
		const bool v3 = p3 < end;
		const uint32_t c3 = v3 ? lst[p3] : c0;
		if (!v1) break;
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			h[j] = hn[j];
			t[j] = tn[j];
			hn[j] = h2[j];
		}
		hit = hit_n;
		ck = c1;
		c1 = c2;
		c2 = c3;
		pk = pk1;
		pk1 = pk2;
		v1 = v2;
		v2 = v3;
		p2 = p3;
	}
}
static_assert(kChunk == 2 * 4 * kBlock, "dense sweep loads two uint4 per array per thread");

// SSSP relax (apps/sssp_simd/sssp.cpp:344-445): if mask[h] and dist[h]+w <
// dist[t] then lower dist[t] and mark t for the next frontier.  The walk of
// bfs_dense_kernel: runs of kDenseRun list entries per block, frontier bits
// from the LDS row window, heads and tails in their 16-bit forms where
// exact, and a three-chunk pipeline: while chunk k is relaxed, the frontier
// bits of chunk k+1 are looked up and its tails and weights (only for 4-edge
// quads with a frontier head) and the heads of chunk k+2 are requested.  A
// lane that sees dist[h]+w below the dist[t] it read issues an atomicMin (no
// return) and marks t: dist[t] then ends the step below the value read
// during it, i.e. t improved this step, which is exactly the condition for t
// to be in the next frontier.
// a lane's 8 weights of chunk c: 32-bit (8 registers) or, with kW8, the
// 1-byte copy (2 registers, 4 weights each); quads without a frontier head
// read the array's first bytes
template <bool kW8>
__device__ __forceinline__ void load_weights(const uint32_t *__restrict__ w32, const uint8_t *__restrict__ w8,
                                             uint64_t c, uint32_t tid, bool q0, bool q1,
                                             uint32_t (&w)[kW8 ? 2 : 8])
{
	if constexpr (kW8) {
		const uint32_t *A = reinterpret_cast<const uint32_t *>(w8 + c * kChunk);
		const uint32_t *D = reinterpret_cast<const uint32_t *>(w8);
		w[0] = q0 ? A[tid] : D[0];
		w[1] = q1 ? A[tid + kBlock] : D[0];
	} else {
		load_edges(false, w32, nullptr, c, tid, q0, q1, w);
	}
}
template <bool kW8>
__device__ __forceinline__ uint32_t weight_of(const uint32_t (&w)[kW8 ? 2 : 8], int j)
{
	if constexpr (kW8)
		return (w[j >> 2] >> (8 * (j & 3))) & 0xFFu;
	else
		return w[j];
}

template <bool kW8>
__global__ __launch_bounds__(kBlock) void sssp_dense_kernel(
    const uint32_t *__restrict__ heads, const uint32_t *__restrict__ tails, const uint16_t *__restrict__ heads16,
    const uint16_t *__restrict__ tails16, const uint32_t *__restrict__ wts, const uint8_t *__restrict__ wts8,
    const ChunkMeta *__restrict__ meta,
    const ChunkPack *__restrict__ pack, const uint32_t *__restrict__ lists, uint32_t list_cap,
    const uint32_t *__restrict__ counts, const uint32_t *__restrict__ cur, uint32_t n_words, uint32_t W, uint32_t R,
    uint32_t *dist, uint8_t *__restrict__ flag)
{
	__shared__ uint4 s_win4[kWinWords / 4];  // frontier bits of chunk k+1's rows
	const uint32_t *s_win = reinterpret_cast<const uint32_t *>(s_win4);
	const uint32_t xcd = blockIdx.x % kNumXcd, nb = gridDim.x / kNumXcd;
	const uint32_t end = counts[xcd * kLiveCntStride];
	const uint32_t skip = (nb - 1) * kDenseRun;
	auto nxt = [&](uint32_t p) { return ((p + 1) % kDenseRun) ? p + 1 : p + 1 + skip; };
	const uint32_t first = (blockIdx.x / kNumXcd) * kDenseRun;
	if (first >= end) return;
	const uint32_t *lst = lists + (uint64_t)xcd * list_cap;
	const uint32_t tid = threadIdx.x;
	const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));
	uint32_t win_lo = UINT32_MAX;  // first word of the window held in LDS (block-uniform)
	const uint32_t c0 = lst[first];
	uint32_t p2 = nxt(nxt(first));
	bool v1 = nxt(first) < end, v2 = p2 < end;
	uint32_t c1 = v1 ? lst[nxt(first)] : c0, c2 = v2 ? lst[p2] : c0;
	PackSlot pk = pack_slot(pack, c0, wv), pk1 = pack_slot(pack, c1, wv);
	uint32_t h[8], hn[8], t[8], w[kW8 ? 2 : 8];
	load_edges(pk.flags & 1u, heads, heads16, c0, tid, true, true, h);
	decode_edges(pk.flags & 1u, pk.hbase, pk.hbase, h);
	uint32_t hit;
	{
		const ChunkMeta m0 = meta[c0];
		uint32_t wlo = 0, wlen = 0;
		const bool inw = window_of(m0.row_lo, m0.row_hi, R, W, n_words, wlo, wlen);
		if (inw) {
			load_window(s_win4, cur, wlo, wlen, tid);
			__syncthreads();
			win_lo = wlo;
		}
		hit = frontier_bits(h, inw, s_win, wlo, wlen, cur);
	}
	{
		load_edges(pk1.flags & 1u, heads, heads16, c1, tid, true, true, hn);
		load_edges(pk.flags & 2u, tails, tails16, c0, tid, hit & 0x0Fu, hit & 0xF0u, t);
		load_weights<kW8>(wts, wts8, c0, tid, hit & 0x0Fu, hit & 0xF0u, w);
	}
	for (;;) {
		// window of chunk k+1 (a block-uniform branch)
		bool inw_n = false;
		uint32_t wlo_n = 0, wlen_n = 0;
		if (v1) {
			const ChunkMeta mn = meta[c1];
			inw_n = window_of(mn.row_lo, mn.row_hi, R, W, n_words, wlo_n, wlen_n);
			if (inw_n && wlo_n != win_lo) {
				__syncthreads();  // every wave is done with the old window
				load_window(s_win4, cur, wlo_n, wlen_n, tid);
				__syncthreads();
				win_lo = wlo_n;
			}
		}
		decode_edges(pk.flags & 2u, pk.tb0, pk.tb1, t);
		decode_edges(pk1.flags & 1u, pk1.hbase, pk1.hbase, hn);
		// stage A: distances at both ends of chunk k's frontier edges, frontier bits of chunk k+1
		uint32_t dh[8], dt[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			const bool on = (hit >> j) & 1u;
			dh[j] = dist[on ? h[j] : 0u];
			dt[j] = dist[on ? t[j] : 0u];
		}
		const uint32_t hit_n = v1 ? frontier_bits(hn, inw_n, s_win, wlo_n, wlen_n, cur) : 0u;
		uint32_t better = 0;
#pragma unroll
		for (int j = 0; j < 8; ++j) better |= (uint32_t)(dh[j] + weight_of<kW8>(w, j) < dt[j]) << j;
		better &= hit;
		// stage B: chunk k's relaxations
#pragma unroll
		for (int j = 0; j < 8; ++j)
			if ((better >> j) & 1u) {
				atomicMin(dist + t[j], dh[j] + weight_of<kW8>(w, j));
				flag[t[j]] = 1;
			}
		// stage C: the requests of the next iteration
		const PackSlot pk2 = pack_slot(pack, c2, wv);
		uint32_t tn[8], wn[kW8 ? 2 : 8], h2[8];
		load_edges(pk1.flags & 2u, tails, tails16, c1, tid, hit_n & 0x0Fu, hit_n & 0xF0u, tn);
		load_weights<kW8>(wts, wts8, c1, tid, hit_n & 0x0Fu, hit_n & 0xF0u, wn);
		load_edges(pk2.flags & 1u, heads, heads16, c2, tid, true, true, h2);
		const uint32_t p3 = nxt(p2);
		const bool v3 = p3 < end;
		const uint32_t c3 = v3 ? lst[p3] : c0;
		if (!v1) break;
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			h[j] = hn[j];
			t[j] = tn[j];
			hn[j] = h2[j];
		}
#pragma unroll
		for (int j = 0; j < (kW8 ? 2 : 8); ++j) w[j] = wn[j];
		hit = hit_n;
		c1 = c2;
		c2 = c3;
		pk = pk1;
		pk1 = pk2;
		v1 = v2;
		v2 = v3;
		p2 = p3;
	}
}

// ---------------------------------------------------------------------------
// sparse step: CSR push, edge-balanced over kSparseTile-edge tiles whose first
// queue entry (tile_lo) was written by the previous frontier update.
// ---------------------------------------------------------------------------
// A small BFS sparse step (queue_step) also builds the next frontier itself:
// a tail is claimed by setting its visited bit (the lane whose atomicOr finds
// the bit clear wins), and the winner writes its parent, level, state and
// tile-row stamp, then appends the tail (when it has out-edges) to the next
// queue -- position and degree prefix reserved with one atomic per block and
// tile -- and the first-entry index of every next-step tile its edges start.
// So the level needs no pass over all vertices.  Column counts are left to a
// recount before the next dense step (col_recount).  A large step
// claims through the state byte and leaves the next frontier to that pass.
template <bool kBfs>
__device__ __forceinline__ void sparse_push_block(
    uint32_t bid, uint32_t nbid, QueueBufs qb, StepCtrl *ctrl, const uint32_t *__restrict__ row_off,
    const uint32_t *__restrict__ cols, const uint32_t *__restrict__ csr_w, const uint32_t *__restrict__ deg,
    uint32_t *visited, uint8_t *__restrict__ flag, int32_t *__restrict__ parent, int32_t *__restrict__ level,
    uint32_t *dist, uint32_t W, uint32_t *__restrict__ row_stamp)
{
	__shared__ uint32_t s_off[kSparseTile + 2];
	__shared__ uint32_t s_u[kSparseTile + 2];
	__shared__ uint32_t s_base[kSparseTile + 2];
	__shared__ uint32_t s_res[kBlock / 64][3];  // BFS: per wavefront |F'|, queue entries, degree sum
	__shared__ uint32_t s_rbase[2];             // BFS: the block's queue position and degree offset
	if (ctrl->done || ctrl->dir != 0) return;  // a dense step, or a finished traversal
	const uint32_t qsize = ctrl->qsize, qdeg = ctrl->fdeg;
	const uint32_t L = ctrl->level, par = L & 1u;
	const bool qmode = kBfs && queue_step(0, qdeg);  // this step builds the next queue itself
	const uint32_t *__restrict__ queue = qb.queue + par * qb.qstride;
	const uint32_t *__restrict__ qoff = qb.qoff + par * qb.qstride;
	const uint32_t *__restrict__ tile_lo = qb.tile_lo + par * qb.tstride;
	uint32_t *nqueue = qb.queue + (par ^ 1u) * qb.qstride;
	uint32_t *nqoff = qb.qoff + (par ^ 1u) * qb.qstride;
	uint32_t *ntile_lo = qb.tile_lo + (par ^ 1u) * qb.tstride;
	uint32_t ntile = (qdeg + kSparseTile - 1) / kSparseTile;
	for (uint32_t k = bid; k < ntile; k += nbid) {
		uint32_t lo = tile_lo[k];
		uint32_t hi = (k + 1 < ntile) ? tile_lo[k + 1] : qsize - 1;
		if (hi > qsize - 1) hi = qsize - 1;
		uint32_t cnt = hi - lo + 1;
		for (uint32_t j = threadIdx.x; j < cnt; j += kBlock) {
			uint32_t u = queue[lo + j];
			s_off[j] = qoff[lo + j];
			s_u[j] = u;
			s_base[j] = row_off[u];
		}
		__syncthreads();
		uint32_t e0 = k * kSparseTile;
		uint32_t e1 = min(e0 + kSparseTile, qdeg);
		if (kBfs) {
			// the tile's 4 edges per lane in batches of dependent loads (tails,
			// then visited bits, then state bytes), so a lane waits on three
			// round trips per tile instead of twelve; the visited bitmap
			// (4 MiB, cache-resident) first: late sparse steps mostly reach
			// vertices visited in earlier steps
			constexpr int kPer = kSparseTile / kBlock;
			static_assert(kSparseTile % kBlock == 0, "whole edges per lane");
			uint32_t uu[kPer], vv[kPer], vw[kPer], sv[kPer];
			bool on[kPer];
#pragma unroll
			for (int r = 0; r < kPer; ++r) {
				const uint32_t e = e0 + threadIdx.x + r * kBlock;
				on[r] = e < e1;
				uint32_t a = 0, b = cnt;  // largest j with s_off[j] <= e
				while (b - a > 1) {
					const uint32_t mid = (a + b) >> 1;
					if (s_off[mid] <= e)
						a = mid;
					else
						b = mid;
				}
				uu[r] = s_u[a];
				vv[r] = on[r] ? cols[s_base[a] + (e - s_off[a])] : 0u;
			}
#pragma unroll
			for (int r = 0; r < kPer; ++r) vw[r] = on[r] ? visited[vv[r] >> 5] : 0xFFFFFFFFu;
			if (!qmode) {
				// a large step: a tail is claimed through its state byte, and the
				// frontier update's pass builds the next frontier from those
#pragma unroll
				for (int r = 0; r < kPer; ++r) {
					on[r] = on[r] && !((vw[r] >> (vv[r] & 31)) & 1u);
					sv[r] = on[r] ? flag[vv[r]] : 1u;
				}
#pragma unroll
				for (int r = 0; r < kPer; ++r)
					if (on[r] && sv[r] == 0) {
						parent[vv[r]] = (int32_t)uu[r];
						flag[vv[r]] = 1;
					}
				__syncthreads();
				continue;
			}
			// claims: the lane that sets a still-clear visited bit wins the tail
#pragma unroll
			for (int r = 0; r < kPer; ++r) {
				const uint32_t bit = 1u << (vv[r] & 31);
				on[r] = on[r] && !(vw[r] & bit);
				sv[r] = on[r] ? atomicOr(visited + (vv[r] >> 5), bit) & bit : bit;
			}
			const uint32_t next_level = L + 1;
			uint32_t nf = 0, nq = 0, nd = 0, dv[kPer];
#pragma unroll
			for (int r = 0; r < kPer; ++r) {
				dv[r] = 0;
				if (on[r] && sv[r] == 0) {
					const uint32_t v = vv[r];
					dv[r] = deg[v];
					parent[v] = (int32_t)uu[r];
					level[v] = (int32_t)next_level;
					flag[v] = kVisited;
					// many winners share a hub row: store the stamp only once
					if (dv[r] && row_stamp[v / W] != next_level) row_stamp[v / W] = next_level;
					nf += 1;
					nq += dv[r] > 0;
					nd += dv[r];
				}
			}
			// the winners' next-queue entries: one reservation per block and
			// tile (|F'| and the (entries, degree sum) pair, two atomics on
			// the control block), then each winner's position and degree
			// offset from the block's and the wavefront's prefixes
			uint32_t ftot, qtot, dtot;
			(void)wave_excl_scan(nf, &ftot);
			uint32_t qx = wave_excl_scan(nq, &qtot), dx = wave_excl_scan(nd, &dtot);
			const uint32_t wid = threadIdx.x >> 6;
			if (lane_id() == 0) {
				s_res[wid][0] = ftot;
				s_res[wid][1] = qtot;
				s_res[wid][2] = dtot;
			}
			__syncthreads();
			if (threadIdx.x == 0) {
				uint32_t f = 0, q = 0, d = 0;
				for (uint32_t w2 = 0; w2 < kBlock / 64; ++w2) {
					f += s_res[w2][0];
					q += s_res[w2][1];
					d += s_res[w2][2];
				}
				unsigned long long base = 0;
				if (f) atomicAdd(&ctrl->facc, f);
				if (q) base = atomicAdd(&ctrl->qpack, ((unsigned long long)q << 32) | d);
				s_rbase[0] = (uint32_t)(base >> 32);
				s_rbase[1] = (uint32_t)base;
			}
			__syncthreads();
			if (qtot) {
				uint32_t pos = s_rbase[0] + qx, off = s_rbase[1] + dx;
				for (uint32_t w2 = 0; w2 < wid; ++w2) {
					pos += s_res[w2][1];
					off += s_res[w2][2];
				}
#pragma unroll
				for (int r = 0; r < kPer; ++r)
					if (dv[r]) {
						nqueue[pos] = vv[r];
						nqoff[pos] = off;
						// next-step tiles whose first edge falls in [off, off + d)
						for (uint32_t k = (off + kSparseTile - 1) / kSparseTile;
						     (uint64_t)k * kSparseTile < (uint64_t)off + dv[r]; ++k)
							ntile_lo[k] = pos;
						pos += 1;
						off += dv[r];
					}
			}
		}
		for (uint32_t e = e0 + threadIdx.x; !kBfs && e < e1; e += kBlock) {
			// largest j with s_off[j] <= e
			uint32_t a = 0, b = cnt;
			while (b - a > 1) {
				uint32_t mid = (a + b) >> 1;
				if (s_off[mid] <= e)
					a = mid;
				else
					b = mid;
			}
			uint32_t u = s_u[a];
			uint32_t idx = s_base[a] + (e - s_off[a]);
			uint32_t v = cols[idx];
			{
				// as in the dense sweep: an edge that beats the dist[v] it read
				// lowers it (no return needed) and marks v, which then ends the
				// step below a value read during it, i.e. improved
				const uint32_t nd = dist[u] + csr_w[idx];
				if (nd < dist[v]) {
					atomicMin(dist + v, nd);
					flag[v] = 1;
				}
			}
		}
		__syncthreads();
	}
}


// The first kernel of a level: the CSR push of a sparse step on blocks
// [0, npush), or the chunk selection of a dense step on blocks [0, nsel)
// (block-uniform branches), so a level does not launch the one it does not
// need.
//
// A BFS dense step that follows a queue-building sparse step first sets the
// frontier bits of its queue (that step built the queue but not the bitmap,
// which the frontier update cleared); a frontier vertex without out-edges is
// in no queue and needs no bit, as it heads no edge of the sweep.
template <bool kBfs>
__global__ __launch_bounds__(kBlock) void level_front_kernel(
    uint32_t npush, uint32_t nsel, QueueBufs qb, StepCtrl *ctrl, const uint32_t *__restrict__ row_off,
    const uint32_t *__restrict__ cols, const uint32_t *__restrict__ csr_w, const uint32_t *__restrict__ deg,
    uint32_t *visited, uint8_t *__restrict__ flag, int32_t *__restrict__ parent, int32_t *__restrict__ level,
    uint32_t *dist, uint32_t W, const ChunkMeta *__restrict__ meta, uint64_t nchunks, uint32_t *row_stamp,
    uint32_t *col_unvis, uint32_t *__restrict__ lists, uint32_t list_cap, uint32_t *__restrict__ counts,
    uint32_t *cur)
{
	if (ctrl->done) return;
	if (ctrl->dir == 0) {
		if (blockIdx.x < npush)
			sparse_push_block<kBfs>(blockIdx.x, npush, qb, ctrl, row_off, cols, csr_w, deg, visited, flag, parent,
			                        level, dist, W, row_stamp);
		return;
	}
	if (kBfs && ctrl->queued) {
		const uint32_t *queue = qb.queue + (ctrl->level & 1u) * qb.qstride;
		const uint32_t qsize = ctrl->qsize;
		for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < qsize; i += gridDim.x * kBlock) {
			const uint32_t v = queue[i];
			atomicOr(cur + (v >> 5), 1u << (v & 31));
		}
	}
	if (blockIdx.x < nsel)
		chunk_select_block<kBfs>(blockIdx.x, meta, nchunks, row_stamp, col_unvis, ctrl, lists, list_cap, counts);
}

// ---------------------------------------------------------------------------
// frontier update (update_dense :91-179, update_sparse :467-509, to_dense
// :54-88, to_sparse :400-465 and the rule :829-834 of apps/bfs_simd/bfs.cpp,
// fused).  Three launches, no same-address atomics on the hot counters:
//   mark : one lane per 32-vertex word: state bytes -> next-frontier word,
//          state 1 -> kVisited (BFS) or 0, levels, row stamps, column counts;
//          per-block (|F|, queue, sum deg)
//   scan : one block: exclusive scan of the per-block counts, direction rule
//   emit : one lane per word: the sparse queue in ascending vertex order with
//          its degree prefix and the tile partition of the next CSR push
// ---------------------------------------------------------------------------

constexpr uint32_t kLevelFromCtrl = UINT32_MAX;  // mark kernel: take the level from the control block

template <bool kBfs>
__global__ __launch_bounds__(kBlock) void frontier_mark_kernel(
    uint32_t W, uint32_t next_level, uint8_t *__restrict__ flag, const uint32_t *__restrict__ deg,
    int32_t *__restrict__ level, uint32_t *__restrict__ visited, uint32_t *__restrict__ cur,
    uint32_t *__restrict__ row_stamp, uint32_t *__restrict__ col_unvis, uint4 *__restrict__ blk_cnt,
    uint2 *__restrict__ wave_cnt, const StepCtrl *__restrict__ ctrl)
{
	__shared__ uint32_t s_red[3][kBlock / 64];
	if (next_level == kLevelFromCtrl) {  // device-driven loop
		if (ctrl->done) return;
		// a small BFS sparse step built its next frontier itself (sparse_push_block)
		if (kBfs && queue_step(ctrl->dir, ctrl->fdeg)) return;
		next_level = ctrl->level + 1;
	}
	const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
	const uint32_t v0 = w * 32;
	uint4 *fp = reinterpret_cast<uint4 *>(flag + v0);
	const uint4 fa = fp[0], fb = fp[1];
	const uint32_t fw[8] = {fa.x, fa.y, fa.z, fa.w, fb.x, fb.y, fb.z, fb.w};
	uint32_t bits = 0;
#pragma unroll
	for (int j = 0; j < 8; ++j)
#pragma unroll
		for (int k = 0; k < 4; ++k)
			bits |= ((fw[j] >> (8 * k)) & 1u) << (4 * j + k);
	uint32_t F = 0, Q = 0, D = 0;
	cur[w] = bits;
	const bool row_uniform = (W & 31u) == 0;
	if (bits) {
		if (kBfs) {
			// state 1 (reached this step) -> kVisited; 0 and kVisited stay
			constexpr uint32_t kLow = 0x01010101u;
			fp[0] = make_uint4(fa.x + (fa.x & kLow), fa.y + (fa.y & kLow), fa.z + (fa.z & kLow), fa.w + (fa.w & kLow));
			fp[1] = make_uint4(fb.x + (fb.x & kLow), fb.y + (fb.y & kLow), fb.z + (fb.z & kLow), fb.w + (fb.w & kLow));
		} else {
			fp[0] = make_uint4(0, 0, 0, 0);
			fp[1] = make_uint4(0, 0, 0, 0);
		}
		if (kBfs) visited[w] |= bits;
		if (row_uniform) row_stamp[v0 / W] = next_level;
		F = (uint32_t)__popc(bits);
	}
	// Levels and degrees of the new vertices, vertex-major over the wave's
	// 2048 vertices: in iteration i lane l takes vertex 64i + l, whose bit is in
	// word 2i + l/32 of the wave, so every level store and degree load of a
	// iteration covers 64 consecutive vertices (a lane per word would scatter
	// them 128 bytes apart).  The 32 degree loads are all issued before the
	// first is summed (one memory latency per wave, not 32); the words come
	// from their lanes by readlane.
	{
		const uint32_t lane = lane_id();
		const uint32_t wv0 = (w - lane) * 32;
		if (__ballot(bits != 0)) {
			uint32_t dv[32];
#pragma unroll
			for (uint32_t i = 0; i < 32; ++i) {
				const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)bits, (int)(2 * i));
				const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)bits, (int)(2 * i + 1));
				const uint32_t wb = lane < 32 ? lo : hi;
				const uint32_t v = wv0 + 64 * i + lane;
				const bool mine = (wb >> (lane & 31)) & 1u;
				dv[i] = mine ? deg[v] : 0u;
				if (mine) {
					if (kBfs) level[v] = (int32_t)next_level;
					if (!row_uniform) row_stamp[v / W] = next_level;
				}
			}
#pragma unroll
			for (uint32_t i = 0; i < 32; ++i) {
				D += dv[i];
				Q += dv[i] > 0;
			}
		}
	}
	if (kBfs) {
		// column counts of unvisited vertices: one atomic per wave when the
		// wave's 2048 vertices share a tail block, else one per word
		const uint32_t wv0 = (w - lane_id()) * 32;
		const bool uniform = (W & 31u) == 0 && wv0 / W == (wv0 + 64 * 32 - 1) / W;
		if (uniform) {
			const uint32_t tot = wave_sum(F);
			if (lane_id() == 0 && tot) atomicSub(col_unvis + wv0 / W, tot);
		} else if (bits) {
			if (row_uniform) {
				atomicSub(col_unvis + v0 / W, F);
			} else {
				for (uint32_t b = bits; b; b &= b - 1) atomicSub(col_unvis + (v0 + __ffs(b) - 1) / W, 1u);
			}
		}
	}
	F = wave_sum(F);
	Q = wave_sum(Q);
	D = wave_sum(D);
	const uint32_t wid = threadIdx.x >> 6;
	if (lane_id() == 0) {
		s_red[0][wid] = F;
		s_red[1][wid] = Q;
		s_red[2][wid] = D;
		wave_cnt[blockIdx.x * (kBlock / 64) + wid] = make_uint2(Q, D);
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		uint4 r = make_uint4(0, 0, 0, 0);
		for (int k = 0; k < (int)(kBlock / 64); ++k) {
			r.x += s_red[0][k];
			r.y += s_red[1][k];
			r.z += s_red[2][k];
		}
		blk_cnt[blockIdx.x] = r;
	}
}

// One block: exclusive scan of per-block (queue, degree) counts, totals, and
// the direction rule |F| + sum(outdeg) > M / T_RATIO in uint32 arithmetic.
constexpr uint32_t kScanThreads = 1024;
// The end of a level, one thread: the next frontier's totals, the direction
// rule |F| + sum(outdeg) > M / T_RATIO in uint32 arithmetic, and (device-
// driven loop) the step's trace entry, the next level, and cleared live-list
// counts for its chunk selection.
__device__ void finish_level(StepCtrl *ctrl, uint4 *trace, uint32_t *live_cnt, uint32_t threshold, uint32_t fall,
                             uint32_t qall, uint32_t dall, bool queued)
{
	ctrl->fsize = fall;
	ctrl->fdeg = dall;
	ctrl->qsize = qall;
	ctrl->queued = queued;
	// the column counts were recounted before a dense step and kept by its
	// frontier pass; a queue-building step leaves them behind again
	if (ctrl->dir == 1) ctrl->col_stale = 0;
	if (queued) ctrl->col_stale = 1;
	ctrl->dir = (uint32_t)(fall + dall) > threshold ? 1u : 0u;
	ctrl->done = fall == 0;
	uint32_t chunks = 0;  // chunks the step's dense sweep took (its live lists)
	for (int x = 0; x < kNumXcd; ++x) chunks += live_cnt[x * kLiveCntStride];
	ctrl->chunks = chunks;
	if (trace) {
		const uint32_t L = ctrl->level;
		if (L <= GPHI_MAX_STEPS) trace[L].w = chunks;
		if (fall && L + 1 <= GPHI_MAX_STEPS) trace[L + 1] = make_uint4(ctrl->dir, fall, dall, 0u);
		ctrl->level = L + 1;
		ctrl->reached += fall;
		ctrl->traversed += dall;
		for (int x = 0; x < kNumXcd; ++x) live_cnt[x * kLiveCntStride] = 0;
	}
}

// queue_path: a small BFS sparse step (queue_step) counted its own next
// frontier (|F'|, queue entries and degree sum in the control block's
// accumulators), so there is nothing to scan.
__global__ __launch_bounds__(kScanThreads) void frontier_scan_kernel(uint32_t nblk, uint32_t threshold,
                                                                     const uint4 *__restrict__ blk_cnt,
                                                                     uint2 *__restrict__ blk_base,
                                                                     uint32_t *__restrict__ live_cnt,
                                                                     StepCtrl *ctrl, uint4 *__restrict__ trace,
                                                                     bool queue_path)
{
	if (trace && ctrl->done) return;  // device-driven loop: nothing left to do
	if (queue_path && queue_step(ctrl->dir, ctrl->fdeg)) {
		if (threadIdx.x == 0) {
			const uint32_t fall = ctrl->facc;
			const unsigned long long qp = ctrl->qpack;
			ctrl->facc = 0;
			ctrl->qpack = 0;
			finish_level(ctrl, trace, live_cnt, threshold, fall, (uint32_t)(qp >> 32), (uint32_t)qp, true);
		}
		return;
	}
	__shared__ uint32_t sq[kScanThreads / 64], sd[kScanThreads / 64], sf[kScanThreads / 64];
	const uint32_t per = (nblk + kScanThreads - 1) / kScanThreads;
	const uint32_t b0 = min(nblk, threadIdx.x * per), b1 = min(nblk, b0 + per);
	uint32_t q = 0, d = 0, f = 0;
	for (uint32_t b = b0; b < b1; ++b) {
		uint4 c = blk_cnt[b];
		f += c.x;
		q += c.y;
		d += c.z;
	}
	uint32_t qt, dt;
	uint32_t qx = wave_excl_scan(q, &qt), dx = wave_excl_scan(d, &dt);
	uint32_t ft = wave_sum(f);
	const uint32_t wid = threadIdx.x >> 6;
	if (lane_id() == 0) {
		sq[wid] = qt;
		sd[wid] = dt;
		sf[wid] = ft;
	}
	__syncthreads();
	uint32_t qb = 0, db = 0, qall = 0, dall = 0, fall = 0;
	for (uint32_t k = 0; k < kScanThreads / 64; ++k) {
		if (k < wid) {
			qb += sq[k];
			db += sd[k];
		}
		qall += sq[k];
		dall += sd[k];
		fall += sf[k];
	}
	qx += qb;
	dx += db;
	for (uint32_t b = b0; b < b1; ++b) {
		uint4 c = blk_cnt[b];
		blk_base[b] = make_uint2(qx, dx);
		qx += c.y;
		dx += c.z;
	}
	if (threadIdx.x == 0) finish_level(ctrl, trace, live_cnt, threshold, fall, qall, dall, false);
}

// Unvisited vertices per tail column, recounted from the visited bitmap
// before a BFS dense step when queue-building sparse steps came before it
// (their pushes leave the counts alone: many winners share a column, and one
// atomic each on it would queue); the dense step's chunk selection skips
// columns with none left.  A thread per column (a grid-stride walk), run by
// the emit kernel's threads: the two passes never both have work at a level.
__device__ __forceinline__ void col_recount(uint32_t n, uint32_t W, uint32_t side,
                                            const uint32_t *__restrict__ visited, uint32_t *__restrict__ col_unvis)
{
	for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < side; c += gridDim.x * kBlock) {
		const uint32_t lo = c * W, hi = min(lo + W, n);  // vertices [lo, hi)
		uint32_t seen = 0;
		for (uint32_t w = lo >> 5; w <= (hi - 1) >> 5; ++w) {
			uint32_t bits = visited[w];
			if (w == (lo >> 5) && (lo & 31)) bits &= ~0u << (lo & 31);
			if (w == ((hi - 1) >> 5) && (hi & 31)) bits &= ~0u >> (32 - (hi & 31));
			seen += (uint32_t)__popc(bits);
		}
		col_unvis[c] = (hi - lo) - seen;
	}
}

// The queue goes to the half of the next level's parity.  queue_path: after a
// queue-building BFS sparse step the push built the queue already; on a
// switch to a dense step the frontier bitmap, which such a step leaves as it
// was, is cleared here (one lane per word) and refilled from the queue by the
// next level's first kernel.
__global__ __launch_bounds__(kBlock) void frontier_emit_kernel(
    const StepCtrl *__restrict__ ctrl, const uint32_t *__restrict__ deg, uint32_t *__restrict__ cur,
    const uint4 *__restrict__ blk_cnt, const uint2 *__restrict__ blk_base, const uint2 *__restrict__ wave_cnt,
    QueueBufs qb, bool queue_path, uint32_t n, uint32_t W, uint32_t side, const uint32_t *__restrict__ visited,
    uint32_t *__restrict__ col_unvis)
{
	if (ctrl->done) return;
	// BFS, before a dense step that follows queue-building steps: the column counts
	if (col_unvis && ctrl->col_stale && ctrl->dir == 1) col_recount(n, W, side, visited, col_unvis);
	const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
	if (queue_path && ctrl->queued) {
		if (ctrl->dir == 1) cur[w] = 0u;
		return;
	}
	// the queue only feeds a sparse step; empty blocks have nothing to emit
	if (ctrl->dir != 0 || blk_cnt[blockIdx.x].y == 0) return;
	const uint32_t par = ctrl->level & 1u;
	uint32_t *__restrict__ queue = qb.queue + par * qb.qstride;
	uint32_t *__restrict__ qoff = qb.qoff + par * qb.qstride;
	uint32_t *__restrict__ tile_lo = qb.tile_lo + par * qb.tstride;
	const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
	const uint32_t bits = cur[w];
	// this wavefront's first queue entry and degree prefix: the block's, plus
	// the totals the mark kernel recorded for the block's earlier wavefronts
	uint2 base = blk_base[blockIdx.x];
	for (uint32_t k = 0; k < wid; ++k) {
		const uint2 c = wave_cnt[blockIdx.x * (kBlock / 64) + k];
		base.x += c.x;
		base.y += c.y;
	}
	// vertex-major over the wave's 2048 vertices, as in the mark kernel: in
	// iteration i lane l takes vertex 64i + l, so the entries come out in
	// ascending vertex order and each store covers consecutive entries.  The
	// 32 degree loads are issued before the first scan needs one.
	const uint32_t wv0 = (w - lane) * 32;
	const unsigned long long any = __ballot(bits != 0);
	const unsigned long long below = (1ull << lane) - 1ull;
	uint32_t dd[32];
#pragma unroll
	for (uint32_t i = 0; i < 32; ++i) {
		const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)bits, (int)(2 * i));
		const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)bits, (int)(2 * i + 1));
		const uint32_t wb = lane < 32 ? lo : hi;
		dd[i] = ((wb >> (lane & 31)) & 1u) ? deg[wv0 + 64 * i + lane] : 0u;
	}
	uint32_t pos = base.x, off = base.y;
#pragma unroll
	for (uint32_t i = 0; i < 32; ++i) {
		if (!((any >> (2 * i)) & 3ull)) continue;
		const uint32_t v = wv0 + 64 * i + lane;
		const uint32_t dv = dd[i];
		const unsigned long long qm = __ballot(dv > 0);
		uint32_t dtot;
		const uint32_t dx = wave_excl_scan(dv, &dtot);
		if (dv) {
			const uint32_t p = pos + (uint32_t)__popcll(qm & below), o = off + dx;
			queue[p] = v;
			qoff[p] = o;
			// tiles of the next push whose first edge falls in [o, o+dv)
			for (uint32_t k = (o + kSparseTile - 1) / kSparseTile; (uint64_t)k * kSparseTile < (uint64_t)o + dv; ++k)
				tile_lo[k] = p;
		}
		pos += (uint32_t)__popcll(qm);
		off += dtot;
	}
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
constexpr unsigned kBfsDenseGrid = 2048;  // BFS/SSSP sweeps: 8 blocks per CU, every block resident at once

// blocks of `block` threads of a kernel that fit the device at once, rounded
// down to a multiple of the XCD count (kBfsDenseGrid if the query fails)
static unsigned resident_grid(const void *kernel, unsigned block = kBlock)
{
	int dev = 0, ncu = 0, per = 0;
	if (hipGetDevice(&dev) != hipSuccess ||
	    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
	    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) != hipSuccess || ncu <= 0 || per <= 0)
		return kBfsDenseGrid;
	return std::max<unsigned>((unsigned)(ncu * per) / kNumXcd * kNumXcd, kNumXcd);
}
static constexpr unsigned kSparseGrid = 2048;
static constexpr uint32_t kLevelGroup = 8;  // levels enqueued per read-back of the control block

static float ev_ms(hipEvent_t a, hipEvent_t b)
{
	float ms = 0.f;
	if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
	return ms;
}

template <bool kBfs>
static int frontier_loop(gphi_graph *gr, uint32_t src, gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	if (!G.has_csr || !G.has_tiles)
		return set_error(ctx, GPHI_E_STATE, "a hybrid traversal needs the CSR and the tile arrays");
	const uint32_t threshold = (uint32_t)G.m / ctx->t_ratio;
	StepCtrl *hc = ctx->h_ctrl;
	hipEvent_t *ev = ctx->ev;
	// total time from two events; per-level events only with level timing on
	const bool total_timing = st != nullptr && ctx->ev_ok;
	const bool timing = total_timing && ctx->level_timing;
	constexpr uint32_t kEvEnd = 3 * GPHI_MAX_STEPS + 3;  // end of the last group (no level timing)
	if (total_timing) GPHI_HIP(ctx, hipEventRecord(ev[0], s));

	// state initialisation (the reference's untimed set-up, bfs.cpp:771-797).
	// BFS levels and parents are written for reached vertices only and never
	// read during the traversal: the -1 of unreached ones is filled in by
	// bfs_finalize_kernel when they are copied out.
	if (!kBfs) fill_u32_kernel<<<grid_for(G.n_pad, kBlock, 4096), kBlock, 0, s>>>(gr->dist, G.n_pad, INT_MAX);
	{
		const uint64_t items = std::max<uint64_t>({G.n_pad / 16, G.n_words, G.side, kLiveCntBytes / 4});
		init_state_kernel<<<grid_for(items, kBlock, 8192), kBlock, 0, s>>>(
		    G.n, G.W, G.side, G.n_pad, G.n_words, reinterpret_cast<uint4 *>(gr->flag), kBfs ? gr->visited : nullptr,
		    gr->cur, gr->live_cnt, gr->row_stamp, gr->col_unvis);
	}
	init_source_kernel<kBfs><<<1, kBlock, 0, s>>>(src, G.W, G.deg, gr->level, gr->parent, gr->dist,
	                                               gr->flag, gr->visited, gr->cur, gr->row_stamp, gr->col_unvis,
	                                               gr->queue, gr->qoff, gr->tile_lo, gr->ctrl, gr->trace);
	GPHI_HIP(ctx, hipGetLastError());
	if (timing) GPHI_HIP(ctx, hipEventRecord(ev[1], s));

	// The level loop runs on the device.  Every kernel of a level reads the
	// direction, the level and the queue sizes from the control block: the
	// CSR push of a dense step and the chunk selection and sweep of a sparse
	// step return at once (the first step is sparse, bfs.cpp:807), and every
	// kernel of a level past the last one returns at once.  So the host
	// enqueues kLevelGroup levels at a time and reads the control block back
	// once per group, not once per level.
	const unsigned nblk = G.n_pad / kPadVerts;
	const uint32_t list_cap = live_list_cap(G.nchunks);
	const unsigned sgrid = (unsigned)std::max<uint64_t>(1, (G.nchunks + kBlock - 1) / kBlock);
	// every block of the sweep resident at once (blocks of a run-interleaved
	// list walk all have the same work), a multiple of the XCD count
	static const unsigned resident = resident_grid(reinterpret_cast<const void *>(bfs_dense_kernel));
	static const unsigned resident_s8 = resident_grid(reinterpret_cast<const void *>(sssp_dense_kernel<true>));
	static const unsigned resident_s32 = resident_grid(reinterpret_cast<const void *>(sssp_dense_kernel<false>));
	const unsigned dense_cap = kBfs ? resident : G.coo_w8 ? resident_s8 : resident_s32;
	unsigned grid = (unsigned)std::min<uint64_t>(dense_cap, (G.nchunks + kNumXcd - 1) / kNumXcd * kNumXcd);
	grid = std::max<unsigned>(grid / kNumXcd * kNumXcd, kNumXcd);
	const uint32_t win_rows = dense_window_rows(G.W);
	const QueueBufs qb = gr->queue_bufs();
	uint32_t enq = 0;
	for (;;) {
		for (uint32_t q = 0; q < kLevelGroup; ++q, ++enq) {
			const bool ev_step = timing && enq < GPHI_MAX_STEPS - 1;
			level_front_kernel<kBfs><<<std::max(kSparseGrid, sgrid), kBlock, 0, s>>>(
			    kSparseGrid, sgrid, qb, gr->ctrl, G.row_off, G.cols, G.csr_w, G.deg, gr->visited, gr->flag,
			    gr->parent, gr->level, gr->dist, G.W, G.meta, G.nchunks, gr->row_stamp, gr->col_unvis, gr->live,
			    list_cap, gr->live_cnt, gr->cur);
			if (ev_step) GPHI_HIP(ctx, hipEventRecord(ev[2 * GPHI_MAX_STEPS + 4 + enq], s));
			if (kBfs)
				bfs_dense_kernel<<<grid, kBlock, 0, s>>>(G.heads, G.tails, G.heads16, G.tails16, G.meta, G.pack,
				                                       gr->live, list_cap, gr->live_cnt,
				                                       gr->cur, gr->visited, G.n_words, G.W, win_rows, gr->flag,
				                                       gr->parent);
			else
			{
				if (G.coo_w8)
					sssp_dense_kernel<true><<<grid, kBlock, 0, s>>>(
					    G.heads, G.tails, G.heads16, G.tails16, G.coo_w, G.coo_w8, G.meta, G.pack, gr->live, list_cap,
					    gr->live_cnt, gr->cur, G.n_words, G.W, win_rows, gr->dist, gr->flag);
				else
					sssp_dense_kernel<false><<<grid, kBlock, 0, s>>>(
					    G.heads, G.tails, G.heads16, G.tails16, G.coo_w, G.coo_w8, G.meta, G.pack, gr->live, list_cap,
					    gr->live_cnt, gr->cur, G.n_words, G.W, win_rows, gr->dist, gr->flag);
			}
			if (ev_step) GPHI_HIP(ctx, hipEventRecord(ev[2 + 2 * enq], s));
			frontier_mark_kernel<kBfs><<<nblk, kBlock, 0, s>>>(G.W, kLevelFromCtrl, gr->flag, G.deg, gr->level,
			                                                  gr->visited, gr->cur, gr->row_stamp, gr->col_unvis,
			                                                  gr->blk_cnt, gr->wave_cnt, gr->ctrl);
			frontier_scan_kernel<<<1, kScanThreads, 0, s>>>(nblk, threshold, gr->blk_cnt, gr->blk_base,
			                                              gr->live_cnt, gr->ctrl, gr->trace, kBfs);
			frontier_emit_kernel<<<nblk, kBlock, 0, s>>>(gr->ctrl, G.deg, gr->cur, gr->blk_cnt, gr->blk_base,
			                                            gr->wave_cnt, qb, kBfs, G.n, G.W, G.side, gr->visited,
			                                            kBfs ? gr->col_unvis : nullptr);
			if (ev_step) GPHI_HIP(ctx, hipEventRecord(ev[3 + 2 * enq], s));
		}
		GPHI_HIP(ctx, hipGetLastError());
		if (total_timing && !timing) GPHI_HIP(ctx, hipEventRecord(ev[kEvEnd], s));
		GPHI_HIP(ctx, hipMemcpyAsync(hc, gr->ctrl, sizeof(StepCtrl), hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		if (hc->done) break;
		if (enq > G.n + 2 + kLevelGroup) return set_error(ctx, GPHI_E_STATE, "frontier loop did not terminate");
	}
	const uint32_t steps = hc->level;  // frontiers expanded
	if (st) {
		const uint32_t ns = std::min<uint32_t>(steps, GPHI_MAX_STEPS);
		std::vector<uint4> tr(ns);
		if (ns) {
			GPHI_HIP(ctx, hipMemcpyAsync(tr.data(), gr->trace, ns * sizeof(uint4), hipMemcpyDeviceToHost, s));
			GPHI_HIP(ctx, hipStreamSynchronize(s));
		}
		st->steps = ns;
		for (uint32_t k = 0; k < ns; ++k) {
			st->dir[k] = tr[k].x;
			st->fsize[k] = tr[k].y;
			st->fdeg[k] = tr[k].z;
			st->chunks[k] = tr[k].x ? tr[k].w : 0u;
		}
		st->reached = hc->reached;
		st->edges_traversed = hc->traversed;
		if (timing) {
			st->init_ms = ev_ms(ev[0], ev[1]);
			hipEvent_t prev = ev[1];
			const uint32_t nt = std::min<uint32_t>(steps, GPHI_MAX_STEPS - 1);
			for (uint32_t k = 0; k < nt; ++k) {
				if (st->dir[k]) {
					hipEvent_t sel = ev[2 * GPHI_MAX_STEPS + 4 + k];
					st->select_ms[k] = ev_ms(prev, sel);
					st->sweep_ms[k] = ev_ms(sel, ev[2 + 2 * k]);
				} else {
					st->sweep_ms[k] = ev_ms(prev, ev[2 + 2 * k]);
				}
				st->update_ms[k] = ev_ms(ev[2 + 2 * k], ev[3 + 2 * k]);
				prev = ev[3 + 2 * k];
			}
			st->total_ms = nt ? ev_ms(ev[0], ev[1 + 2 * nt]) : st->init_ms;
		} else {
			memset(st->sweep_ms, 0, sizeof(st->sweep_ms));
			memset(st->select_ms, 0, sizeof(st->select_ms));
			memset(st->update_ms, 0, sizeof(st->update_ms));
			st->init_ms = 0.f;
			st->total_ms = total_timing ? ev_ms(ev[0], ev[kEvEnd]) : 0.f;
		}
	}
	return GPHI_OK;
}

// ---- helpers shared with other translation units (batch.hip) ----
int ensure_traversal_state(gphi_graph *gr) { return ensure_state(gr); }

// Frontier update from the state bytes (mark / scan / emit) in the
// device-driven form: nothing when the control block says the traversal is
// done; otherwise the next level's union frontier, its direction and queue,
// and ctrl->level advanced by one (from UINT32_MAX, the first update opens
// level 0).  No level writes.
void launch_frontier_update_driven(gphi_graph *gr, uint32_t threshold)
{
	DevGraph &G = gr->g;
	hipStream_t s = gr->ctx->stream;
	const unsigned nblk = G.n_pad / kPadVerts;
	frontier_mark_kernel<false><<<nblk, kBlock, 0, s>>>(G.W, kLevelFromCtrl, gr->flag, G.deg, gr->level,
	                                                   gr->visited, gr->cur, gr->row_stamp, gr->col_unvis,
	                                                   gr->blk_cnt, gr->wave_cnt, gr->ctrl);
	frontier_scan_kernel<<<1, kScanThreads, 0, s>>>(nblk, threshold, gr->blk_cnt, gr->blk_base, gr->live_cnt,
	                                              gr->ctrl, gr->trace, false);
	frontier_emit_kernel<<<nblk, kBlock, 0, s>>>(gr->ctrl, G.deg, gr->cur, gr->blk_cnt, gr->blk_base,
	                                            gr->wave_cnt, gr->queue_bufs(), false, G.n, G.W, G.side, gr->visited,
	                                            nullptr);
}

// level and parent -1 for the vertices a BFS did not reach (state 0; every
// reached vertex ends in kVisited)
__global__ __launch_bounds__(kBlock) void bfs_finalize_kernel(uint32_t n, const uint8_t *__restrict__ state,
                                                              int32_t *__restrict__ level,
                                                              int32_t *__restrict__ parent)
{
	for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
		if (state[v] == 0) {
			if (level) level[v] = -1;
			if (parent) parent[v] = -1;
		}
}

int run_bfs(gphi_graph *gr, uint32_t src, int32_t *levels, int32_t *parents, gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	if (src >= G.n) return set_error(ctx, GPHI_E_RANGE, "source %u >= N %u", src, G.n);
	int rc = ensure_state(gr);
	if (rc) return rc;
	if (st) memset(st, 0, sizeof(*st));
	if ((rc = frontier_loop<true>(gr, src, st))) return rc;
	hipStream_t s = ctx->stream;
	if (levels || parents)
		bfs_finalize_kernel<<<grid_for(G.n, kBlock, 8192), kBlock, 0, s>>>(
		    G.n, gr->flag, levels ? gr->level : nullptr, parents ? gr->parent : nullptr);
	if (levels) GPHI_HIP(ctx, hipMemcpyAsync(levels, gr->level, (size_t)G.n * 4, hipMemcpyDeviceToHost, s));
	if (parents) GPHI_HIP(ctx, hipMemcpyAsync(parents, gr->parent, (size_t)G.n * 4, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	return GPHI_OK;
}

int bfs_level_loop(gphi_graph *gr, uint32_t src, gphi_stats *st) { return frontier_loop<true>(gr, src, st); }

uint32_t live_list_capacity(uint64_t nchunks) { return live_list_cap(nchunks); }

int run_sssp(gphi_graph *gr, uint32_t src, uint32_t *dist, gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	if (src >= G.n) return set_error(ctx, GPHI_E_RANGE, "source %u >= N %u", src, G.n);
	if (!G.weighted) return set_error(ctx, GPHI_E_STATE, "graph carries no weights");
	int rc = ensure_state(gr);
	if (rc) return rc;
	if (!gr->dist && (rc = dmalloc(ctx, (void **)&gr->dist, (size_t)G.n_pad * 4, nullptr))) return rc;
	if (st) memset(st, 0, sizeof(*st));
	if ((rc = frontier_loop<false>(gr, src, st))) return rc;
	if (dist)
		GPHI_HIP(ctx, hipMemcpyAsync(dist, gr->dist, (size_t)G.n * 4, hipMemcpyDeviceToHost, ctx->stream));
	GPHI_HIP(ctx, hipStreamSynchronize(ctx->stream));
	if (st) {
		// reached = vertices with a finite distance (a vertex may be expanded
		// several times by a label-correcting sweep, so the trace overcounts)
		st->reached = 0;
		st->edges_traversed = 0;
	}
	return GPHI_OK;
}

// ---------------------------------------------------------------------------
// PageRank (FP32): contrib = rank/deg; sum[t] += contrib[h] over the
// in-edges; rank = (1-d)/N + d*sum   (apps/pageRank_simd/page_rank.cpp:128-199,428-446)
// ---------------------------------------------------------------------------
__global__ void pr_prepare_kernel(uint32_t n, const float *__restrict__ rank, const uint32_t *__restrict__ deg,
                                  float *__restrict__ contrib, float *__restrict__ sum)
{
	uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v < n) {
		uint32_t d = deg[v];
		contrib[v] = d ? rank[v] / (float)d : 0.f;
		sum[v] = 0.f;
	}
}

// PageRank sum (apps/pageRank_simd/page_rank.cpp:128-199 gathers
// rank[h]/deg[h] into sum[t] over stripes of tile rows, each tile column
// owned by one thread) over head bands (build_pr_bands), in band order, each
// band adding its in-edges of every tail to sum[t].  A band's gathers all
// fall in its window of contrib, which stays in L2, instead of spreading over
// the whole vector.  The FP32 additions run in a fixed order, so an
// iteration gives the same bits on every run: sum[t] = ((0 + band 0) +
// band 1) + ...; a small pair is summed by one lane in edge order; within a
// light pair kPullGroup lanes add in-edges g, g + kPullGroup, ... in turn and
// the group adds its lanes in a fixed butterfly; a heavy pair's pieces (a
// wavefront each, rounds of kPrUnroll * 64 in-edges) are summed the same way,
// a one-piece pair's straight into sum, the others' into psum, which
// pr_multi_kernel then adds in piece order (lane j takes pieces j, j + 64,
// ..., then the butterfly).  Every tail appears at most once per band, so
// the adds of one band never meet on a tail.  All loads of a lane are issued
// before any gathered value is used.
constexpr uint32_t kPullGroup = 8;
constexpr uint32_t kPrSlots = kPullLight / kPullGroup;  // in-edges per lane of a light pair
constexpr uint32_t kPrUnroll = 8;                       // loads in flight per lane of a heavy piece
constexpr uint32_t kPrLightPerBlock = kBlock / kPullGroup;
constexpr uint32_t kPrPiecesPerBlock = kBlock / 64;

__global__ __launch_bounds__(kBlock) void pr_small_kernel(PrBand b, const uint32_t *__restrict__ lt,
                                                          const uint32_t *__restrict__ llo,
                                                          const uint8_t *__restrict__ lcnt,
                                                          const uint32_t *__restrict__ src,
                                                          const float *__restrict__ contrib, float *__restrict__ sum)
{
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= b.nsmall) return;
	const uint32_t t = lt[b.small0 + i], lo = llo[b.small0 + i], cnt = lcnt[b.small0 + i];
	uint32_t h[kPrSmall];
#pragma unroll
	for (uint32_t j = 0; j < kPrSmall; ++j) h[j] = j < cnt ? src[lo + j] : 0xFFFFFFFFu;
	float c[kPrSmall];
#pragma unroll
	for (uint32_t j = 0; j < kPrSmall; ++j) c[j] = h[j] != 0xFFFFFFFFu ? contrib[h[j]] : 0.f;
	float acc = 0.f;
#pragma unroll
	for (uint32_t j = 0; j < kPrSmall; ++j) acc += c[j];
	sum[t] += acc;
}

__global__ __launch_bounds__(kBlock) void pr_light_kernel(PrBand b, const uint32_t *__restrict__ lt,
                                                          const uint32_t *__restrict__ llo,
                                                          const uint8_t *__restrict__ lcnt,
                                                          const uint32_t *__restrict__ src,
                                                          const float *__restrict__ contrib, float *__restrict__ sum)
{
	const uint32_t i = blockIdx.x * kPrLightPerBlock + threadIdx.x / kPullGroup, g = threadIdx.x % kPullGroup;
	uint32_t t = 0, lo = 0, cnt = 0;
	if (i < b.nlight) {
		t = lt[b.light0 + i];
		lo = llo[b.light0 + i];
		cnt = lcnt[b.light0 + i];
	}
	uint32_t h[kPrSlots];
#pragma unroll
	for (uint32_t j = 0; j < kPrSlots; ++j) {
		const uint32_t k = j * kPullGroup + g;
		h[j] = k < cnt ? src[lo + k] : 0xFFFFFFFFu;
	}
	float c[kPrSlots];
#pragma unroll
	for (uint32_t j = 0; j < kPrSlots; ++j) c[j] = h[j] != 0xFFFFFFFFu ? contrib[h[j]] : 0.f;
	float acc = 0.f;
#pragma unroll
	for (uint32_t j = 0; j < kPrSlots; ++j) acc += c[j];
#pragma unroll
	for (int o = kPullGroup / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kPullGroup);
	if (g == 0 && cnt) sum[t] += acc;
}

__device__ __forceinline__ float wave_sum_f32(float x)
{
	// xor butterfly: lanes l and l^o add the same two values, so every lane
	// ends with the same bits
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
	return x;
}

// One piece p of a heavy pair, by one wavefront.  With kHot, heads below
// `hot` read their contributions from s_hot (LDS) instead of contrib; the
// values and their order are the same either way.
template <bool kHot>
__device__ __forceinline__ void pr_piece(uint32_t p, uint32_t lane, const uint4 *__restrict__ hp,
                                         const uint32_t *__restrict__ pc, const uint32_t *__restrict__ src,
                                         const float *__restrict__ contrib, const float *s_hot, uint32_t hot,
                                         float *__restrict__ sum, float *__restrict__ psum)
{
	const uint32_t hi = pc[p];
	const uint4 r = hp[hi];  // tail, first edge, count, first piece
	const uint32_t k = p - r.w, e0 = r.y + k * kPullPiece, cnt = min(kPullPiece, r.z - k * kPullPiece);
	float acc = 0.f;
	for (uint32_t off = 0; off < cnt; off += 64 * kPrUnroll) {
		uint32_t hh[kPrUnroll];
#pragma unroll
		for (uint32_t u = 0; u < kPrUnroll; ++u) {
			const uint32_t j = off + u * 64 + lane;
			hh[u] = j < cnt ? src[e0 + j] : 0xFFFFFFFFu;
		}
		float c[kPrUnroll];
#pragma unroll
		for (uint32_t u = 0; u < kPrUnroll; ++u) {
			if (kHot && hh[u] < hot)
				c[u] = s_hot[hh[u]];
			else
				c[u] = hh[u] != 0xFFFFFFFFu ? contrib[hh[u]] : 0.f;
		}
#pragma unroll
		for (uint32_t u = 0; u < kPrUnroll; ++u) acc += c[u];
	}
	acc = wave_sum_f32(acc);
	if (lane == 0) {
		if (r.z <= kPullPiece)
			sum[r.x] += acc;  // the pair's only piece
		else
			psum[p] = acc;
	}
}

__global__ __launch_bounds__(kBlock) void pr_piece_kernel(PrBand b, const uint4 *__restrict__ hp,
                                                          const uint32_t *__restrict__ pc,
                                                          const uint32_t *__restrict__ src,
                                                          const float *__restrict__ contrib, float *__restrict__ sum,
                                                          float *__restrict__ psum)
{
	const uint32_t q = blockIdx.x * kPrPiecesPerBlock + (threadIdx.x >> 6);
	if (q >= b.npieces) return;  // uniform per wavefront
	pr_piece<false>(b.piece0 + q, threadIdx.x & 63, hp, pc, src, contrib, nullptr, 0, sum, psum);
}

// The pieces of the first band (heads from 0), whose lowest heads, the
// highest out-degrees after the degree reordering, carry a large share of
// its in-edges: every block first copies the contributions of heads below
// kPrHot into LDS (144 KiB, one block per CU), so those gathers skip L2,
// then walks the pieces, a wavefront each.
constexpr uint32_t kPrHot = 36864;
constexpr uint32_t kPrHotBlock = 1024;

__global__ __launch_bounds__(kPrHotBlock) void pr_piece_hot_kernel(PrBand b, const uint4 *__restrict__ hp,
                                                                   const uint32_t *__restrict__ pc,
                                                                   const uint32_t *__restrict__ src,
                                                                   const float *__restrict__ contrib,
                                                                   float *__restrict__ sum, float *__restrict__ psum)
{
	__shared__ float s_hot[kPrHot];
	const uint32_t hot = min(kPrHot, b.head_hi);  // head_lo is 0
	for (uint32_t i = threadIdx.x; i < hot; i += kPrHotBlock) s_hot[i] = contrib[i];
	__syncthreads();
	constexpr uint32_t kWaves = kPrHotBlock / 64;
	for (uint32_t q = blockIdx.x * kWaves + (threadIdx.x >> 6); q < b.npieces; q += gridDim.x * kWaves)
		pr_piece<true>(b.piece0 + q, threadIdx.x & 63, hp, pc, src, contrib, s_hot, hot, sum, psum);
}

// a wavefront per heavy pair of more than one piece: its pieces' sums in
// piece order
__global__ __launch_bounds__(kBlock) void pr_multi_kernel(PrBand b, const uint32_t *__restrict__ multi,
                                                          const uint4 *__restrict__ hp,
                                                          const float *__restrict__ psum, float *__restrict__ sum)
{
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t q = blockIdx.x * kPrPiecesPerBlock + (threadIdx.x >> 6);
	if (q >= b.nmulti) return;  // uniform per wavefront
	const uint4 r = hp[multi[b.multi0 + q]];
	const uint32_t np = (r.z + kPullPiece - 1) / kPullPiece;
	float s = 0.f;
	for (uint32_t j = lane; j < np; j += 64) s += psum[r.w + j];
	s = wave_sum_f32(s);
	if (lane == 0) sum[r.x] += s;
}

// rank = (1-d)/N + d*sum (page_rank.cpp:428-446), the next iteration's
// contributions rank/deg, and sum cleared for it, in one pass
__global__ void pr_apply_kernel(uint32_t n, float first, float damping, float *__restrict__ sum,
                                const uint32_t *__restrict__ deg, float *__restrict__ rank,
                                float *__restrict__ contrib)
{
	uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v < n) {
		const float r = first + damping * sum[v];
		sum[v] = 0.f;
		rank[v] = r;
		const uint32_t d = deg[v];
		contrib[v] = d ? r / (float)d : 0.f;
	}
}

int run_pagerank(gphi_graph *gr, uint32_t iters, float damping, float init, float *rank, float *ms)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	int rc;
	if ((rc = build_in_edges(gr))) return rc;
	if ((rc = build_pr_bands(gr))) return rc;
	if (!gr->rank) {
		// rank, sum, contrib (n_pad each), then the partial sums of the pieces
		const size_t bytes = 3 * (size_t)G.n_pad * 4 + ((size_t)gr->pr_npieces + 1) * 4;
		if ((rc = dmalloc(ctx, (void **)&gr->rank, bytes, nullptr))) return rc;
		gr->sum = gr->rank + G.n_pad;
		gr->contrib = gr->sum + G.n_pad;
	}
	float *psum = gr->contrib + G.n_pad;
	uint32_t bits;
	memcpy(&bits, &init, 4);
	fill_u32_kernel<<<grid_for(G.n, kBlock, 4096), kBlock, 0, s>>>((uint32_t *)gr->rank, G.n, bits);
	const float first = (1.0f - damping) / (float)G.n;
	unsigned vgrid = grid_for(G.n, kBlock, 1u << 30);
	static const unsigned hot_grid = resident_grid((const void *)pr_piece_hot_kernel, kPrHotBlock);
	if (ctx->ev_ok) GPHI_HIP(ctx, hipEventRecord(ctx->ev[0], s));
	pr_prepare_kernel<<<vgrid, kBlock, 0, s>>>(G.n, gr->rank, G.deg, gr->contrib, gr->sum);
	for (uint32_t it = 0; it < iters; ++it) {
		for (const PrBand &b : gr->pr_bands) {
			if (b.nsmall)
				pr_small_kernel<<<(b.nsmall + kBlock - 1) / kBlock, kBlock, 0, s>>>(
				    b, gr->pr_lt, gr->pr_llo, gr->pr_lcnt, gr->pr_src, gr->contrib, gr->sum);
			if (b.nlight)
				pr_light_kernel<<<(b.nlight + kPrLightPerBlock - 1) / kPrLightPerBlock, kBlock, 0, s>>>(
				    b, gr->pr_lt, gr->pr_llo, gr->pr_lcnt, gr->pr_src, gr->contrib, gr->sum);
			if (b.npieces && b.head_lo == 0)
				pr_piece_hot_kernel<<<std::min<unsigned>(hot_grid, (b.npieces + kPrHotBlock / 64 - 1) /
				                                                      (kPrHotBlock / 64)),
				                      kPrHotBlock, 0, s>>>(b, gr->pr_hp, gr->pr_pc, gr->pr_src, gr->contrib,
				                                           gr->sum, psum);
			else if (b.npieces)
				pr_piece_kernel<<<(b.npieces + kPrPiecesPerBlock - 1) / kPrPiecesPerBlock, kBlock, 0, s>>>(
				    b, gr->pr_hp, gr->pr_pc, gr->pr_src, gr->contrib, gr->sum, psum);
			if (b.nmulti)
				pr_multi_kernel<<<(b.nmulti + kPrPiecesPerBlock - 1) / kPrPiecesPerBlock, kBlock, 0, s>>>(
				    b, gr->pr_multi, gr->pr_hp, psum, gr->sum);
		}
		pr_apply_kernel<<<vgrid, kBlock, 0, s>>>(G.n, first, damping, gr->sum, G.deg, gr->rank, gr->contrib);
	}
	GPHI_HIP(ctx, hipGetLastError());
	if (ctx->ev_ok) GPHI_HIP(ctx, hipEventRecord(ctx->ev[1], s));
	if (rank) GPHI_HIP(ctx, hipMemcpyAsync(rank, gr->rank, (size_t)G.n * 4, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	if (ms) *ms = ctx->ev_ok ? ev_ms(ctx->ev[0], ctx->ev[1]) : 0.f;
	return GPHI_OK;
}

// ---------------------------------------------------------------------------
// hop relaxation (apps/bellman-ford_simd/bellman-ford.cpp:129-191): full
// sweeps of dist[t] = min(dist[t], dist[h]+1); a sweep that lowers nothing is
// the fixpoint, where further sweeps would change nothing.
// ---------------------------------------------------------------------------
// A sweep over the tile arrays in their stored (row-major) order, the
// reference's edge order: per edge, dist[t] = min(dist[t], dist[h]+1).  The
// tile order keeps both ends of a chunk's edges inside one row and one column
// block, so the distance reads hit cache; only an edge that lowers dist[t]
// issues an atomicMin.  Two chunks per block iteration with every load issued
// before any is used; blocks of XCD x (blockIdx % 8) walk the chunks c with
// c / 64 % 8 == x.  Heads and tails are read in their 16-bit forms where
// exact.  changed is set to 1 when some edge lowered a distance.
__global__ __launch_bounds__(kBlock) void hop_sweep_kernel(const uint32_t *__restrict__ heads,
                                                           const uint32_t *__restrict__ tails,
                                                           const uint16_t *__restrict__ heads16,
                                                           const uint16_t *__restrict__ tails16,
                                                           const ChunkPack *__restrict__ pack,
                                                           uint64_t nchunks, uint32_t n, uint32_t *dist,
                                                           uint32_t *changed)
{
	const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
	const uint32_t xcd = blockIdx.x % kNumXcd, slot = blockIdx.x / kNumXcd, per = gridDim.x / kNumXcd;
	const uint64_t groups = (nchunks + 63) / 64;
	const uint64_t mine = groups / kNumXcd + (xcd < groups % kNumXcd ? 1 : 0);  // groups g = xcd + 8i
	const uint64_t my_chunks = mine * 64;
	bool lowered = false;
	for (uint64_t i = slot; i < my_chunks; i += 2 * per) {
		uint64_t cc[2];
		bool ok[2];
#pragma unroll
		for (int q = 0; q < 2; ++q) {
			const uint64_t k = i + (uint64_t)q * per;  // k-th chunk of this XCD
			cc[q] = ((k / 64) * kNumXcd + xcd) * 64 + k % 64;
			ok[q] = k < my_chunks && cc[q] < nchunks;
		}
		uint32_t h[16], t[16];
#pragma unroll
		for (int q = 0; q < 2; ++q) {
			// heads and tails in their 16-bit forms where exact (pack is null
			// when the graph has none)
			const uint64_t c = ok[q] ? cc[q] : 0;
			const PackSlot pk = pack ? pack_slot(pack, (uint32_t)c, wv) : PackSlot{0u, 0u, 0u, 0u};
			uint32_t hv[8], tv[8];
			load_edges(pk.flags & 1u, heads, heads16, c, threadIdx.x, true, true, hv);
			load_edges(pk.flags & 2u, tails, tails16, c, threadIdx.x, true, true, tv);
			decode_edges(pk.flags & 1u, pk.hbase, pk.hbase, hv);
			decode_edges(pk.flags & 2u, pk.tb0, pk.tb1, tv);
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				// sentinel edges (head n) and absent chunks read vertex 0 and are masked
				const bool e = ok[q] && hv[j] < n;
				h[8 * q + j] = e ? hv[j] : 0u;
				t[8 * q + j] = e ? tv[j] : 0u;
			}
		}
		uint32_t dh[16], dt[16];
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			dh[j] = dist[h[j]];
			dt[j] = dist[t[j]];
		}
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			const bool e = ok[j / 8] && h[j] != t[j];
			const uint32_t nd = dh[j] + 1;
			if (e && (int32_t)nd < (int32_t)dt[j] && nd < atomicMin(dist + t[j], nd)) lowered = true;
		}
	}
	// a plain store of 1, not an atomic: every writer stores the same value,
	// and thousands of wavefronts would queue on one counter
	if (__ballot(lowered) && lane_id() == 0) *changed = 1u;
}

int run_bellman_ford(gphi_graph *gr, uint32_t src, uint32_t max_sweeps, uint32_t *dist,
                     uint32_t *sweeps_run, float *ms)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	if (src >= G.n) return set_error(ctx, GPHI_E_RANGE, "source %u >= N %u", src, G.n);
	if (!G.has_tiles) return set_error(ctx, GPHI_E_STATE, "graph has no tile arrays");
	hipStream_t s = ctx->stream;
	int rc;
	if (!gr->dist && (rc = dmalloc(ctx, (void **)&gr->dist, (size_t)G.n_pad * 4, nullptr))) return rc;
	if (!gr->ctrl && (rc = dmalloc(ctx, (void **)&gr->ctrl, sizeof(StepCtrl), nullptr))) return rc;
	// one "lowered something" counter per sweep of a batch; the batch size,
	// not max_sweeps, sizes the buffer
	constexpr uint32_t kBatch = 4;
	uint32_t *counts = nullptr;
	if ((rc = dmalloc(ctx, (void **)&counts, kBatch * 4, nullptr))) return rc;
	if (ctx->ev_ok) GPHI_HIP(ctx, hipEventRecord(ctx->ev[0], s));
	fill_u32_kernel<<<grid_for(G.n_pad, kBlock, 4096), kBlock, 0, s>>>(gr->dist, G.n_pad, INT_MAX - 2);
	uint32_t zero = 0;
	GPHI_HIP(ctx, hipMemcpyAsync(gr->dist + src, &zero, 4, hipMemcpyHostToDevice, s));
	const unsigned grid = kBfsDenseGrid;
	uint32_t done = 0, run = 0;
	uint32_t hcnt[kBatch];
	int err = GPHI_OK;
	while (run < max_sweeps && !done) {
		const uint32_t nb = std::min<uint32_t>(kBatch, max_sweeps - run);
		hipError_t e = hipMemsetAsync(counts, 0, kBatch * 4, s);
		for (uint32_t k = 0; k < nb && e == hipSuccess; ++k) {
			hop_sweep_kernel<<<grid, kBlock, 0, s>>>(G.heads, G.tails, G.heads16, G.tails16,
			                                       G.heads16 ? G.pack : nullptr, G.nchunks, G.n, gr->dist,
			                                       counts + k);
			e = hipGetLastError();
		}
		if (e == hipSuccess) e = hipMemcpyAsync(hcnt, counts, nb * 4, hipMemcpyDeviceToHost, s);
		if (e == hipSuccess) e = hipStreamSynchronize(s);
		if (e != hipSuccess) {
			err = hip_fail(ctx, e, "hop relaxation batch", __FILE__, __LINE__);
			break;
		}
		for (uint32_t k = 0; k < nb; ++k) {
			++run;
			if (hcnt[k] == 0) {
				done = 1;
				break;
			}
		}
	}
	(void)hipFree(counts);
	if (err) return err;
	if (ctx->ev_ok) GPHI_HIP(ctx, hipEventRecord(ctx->ev[1], s));
	if (dist) GPHI_HIP(ctx, hipMemcpyAsync(dist, gr->dist, (size_t)G.n * 4, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	if (ms) *ms = ctx->ev_ok ? ev_ms(ctx->ev[0], ctx->ev[1]) : 0.f;
	if (sweeps_run) *sweeps_run = run;
	return GPHI_OK;
}

}  // namespace gphi

// sssp_batch.hip -- weighted SSSP from up to 16 sources in one traversal.
//
// The label-correcting relaxation of apps/sssp_simd/sssp.cpp (dist[t] =
// min(dist[t], dist[h] + w) over the frontier's edges, the frontier being the
// vertices whose distance dropped in the last step; :968-1100) run for k
// sources at once.  Each vertex holds 16 distances side by side (64 bytes, one
// sector), so a frontier edge reads its head's and its tail's distances with
// one random access each, for all sources, where k single-source runs would
// make k of each.  The frontier is the union over the sources; a step relaxes
// every source's distance along every union-frontier edge (a source whose
// distance at the head did not drop relaxes to no effect), and the traversal
// ends when no distance drops.  Label correction reaches the unique shortest
// distances whatever the order of the relaxations, so each row equals that
// source's single-source result (and Dijkstra's).
//
// The steps reuse the single-source machinery: a direction rule on the
// union frontier (|F| + sum deg > M / 2, see kSsspBatchDenseDiv), the frontier update in its
// device-driven form (state byte 1 = some distance dropped) and the CSR push
// of a sparse step over the union queue.  A dense step is a pull over the
// in-edges and their weights (build_in_weights), skipping heads outside the
// frontier bitmap, so that a tail's 16 distances take one write, not an
// atomic per improved source.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "internal.h"

namespace gphi {

namespace {

constexpr uint32_t kInf = 2147483647u;  // unreached, as sssp.cpp:994-998
constexpr uint32_t kSsspLevelGroup = 8;  // levels enqueued per read-back of the control block
constexpr unsigned kSsspGrid = 8192;  // blocks of the push and pull launches (grid-stride loops)
// A step is dense when the union frontier's |F| + sum deg(F) exceeds M / 2,
// not M / T_RATIO as in the single-source rule: a push step relaxes 16
// distances per frontier edge with one row gather, while a pull step reads
// every in-edge.  Measured over divisors 20, 5, 4, 2 and 1: 2 is fastest on
// rmat25-ef36 (106 against 112 ms for 16 sources) and on the serving mix's
// rmat22 and rmat24 graphs (6.9 against 7.7 ms, 23.0 against 24.5 ms).
constexpr uint64_t kSsspBatchDenseDiv = 2;

__device__ __forceinline__ void load_row(const uint32_t *__restrict__ dist, uint32_t v, uint32_t (&d)[kSsspBatchMax])
{
	const uint4 *p = reinterpret_cast<const uint4 *>(dist + (uint64_t)v * kSsspBatchMax);
#pragma unroll
	for (uint32_t i = 0; i < kSsspBatchMax / 4; ++i) {
		const uint4 x = p[i];
		d[4 * i] = x.x;
		d[4 * i + 1] = x.y;
		d[4 * i + 2] = x.z;
		d[4 * i + 3] = x.w;
	}
}

// every source's dist[t] = min(dist[t], dist[h] + w); true when one dropped.
// A distance that beats the dist[t] read lowers it with a return-free
// atomicMin, as the single-source sweep does: dist[t] then ends the step below
// a value read during it, so t improved this step.  The sum is taken in 64
// bits: a head unreached by a source (kInf) or a wide weight cannot wrap.
__device__ __forceinline__ bool relax_row(const uint32_t (&dh)[kSsspBatchMax], uint32_t w, uint32_t t,
                                          uint32_t *dist)
{
	uint32_t dt[kSsspBatchMax];
	load_row(dist, t, dt);
	bool better = false;
#pragma unroll
	for (int s = 0; s < (int)kSsspBatchMax; ++s) {
		const uint64_t nd = (uint64_t)dh[s] + w;
		if (dh[s] != kInf && nd < dt[s]) {
			atomicMin(dist + (uint64_t)t * kSsspBatchMax + s, (uint32_t)nd);
			better = true;
		}
	}
	return better;
}

// all distances kInf, state bytes and frontier bitmap clear, control block
// set for the first frontier update (which opens level 0)
__global__ __launch_bounds__(kBlock) void ssspb_clear_kernel(uint32_t n_pad, uint32_t n_words, uint4 *dist4,
                                                             uint8_t *flag, uint32_t *cur, StepCtrl *ctrl)
{
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	const uint64_t i0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
	for (uint64_t i = i0; i < (uint64_t)n_pad * kSsspBatchMax / 4; i += stride)
		dist4[i] = make_uint4(kInf, kInf, kInf, kInf);
	for (uint64_t i = i0; i < n_pad / 16; i += stride) reinterpret_cast<uint4 *>(flag)[i] = make_uint4(0, 0, 0, 0);
	for (uint64_t i = i0; i < n_words; i += stride) cur[i] = 0;
	if (i0 == 0) {
		ctrl->level = UINT32_MAX;
		ctrl->done = 0;
		ctrl->err = 0;
		ctrl->dir = 0;
		ctrl->reached = 0;
		ctrl->traversed = 0;
	}
}

// source i at slot i: distance 0, in the first frontier
__global__ void ssspb_seed_kernel(const uint32_t *__restrict__ srcs, uint32_t k, uint32_t *dist, uint8_t *flag)
{
	for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
		dist[(uint64_t)srcs[i] * kSsspBatchMax + i] = 0;
		flag[srcs[i]] = 1;
	}
}

// Sparse step: the edge-balanced CSR push over the union queue (tiles of
// kSparseTile edges from tile_lo, as the single-source push).
__global__ __launch_bounds__(kBlock) void ssspb_sparse_kernel(const StepCtrl *__restrict__ ctrl, QueueBufs qb,
                                                              const uint32_t *__restrict__ row_off,
                                                              const uint32_t *__restrict__ cols,
                                                              const uint32_t *__restrict__ csr_w, uint32_t *dist,
                                                              uint8_t *flag)
{
	__shared__ uint32_t s_off[kSparseTile + 2];
	__shared__ uint32_t s_u[kSparseTile + 2];
	__shared__ uint32_t s_base[kSparseTile + 2];
	if (ctrl->done || ctrl->dir != 0) return;
	const uint32_t qsize = ctrl->qsize, qdeg = ctrl->fdeg;
	const uint32_t par = ctrl->level & 1u;  // the half the frontier update wrote
	const uint32_t *__restrict__ queue = qb.queue + par * qb.qstride;
	const uint32_t *__restrict__ qoff = qb.qoff + par * qb.qstride;
	const uint32_t *__restrict__ tile_lo = qb.tile_lo + par * qb.tstride;
	const uint32_t ntile = (qdeg + kSparseTile - 1) / kSparseTile;
	for (uint32_t k = blockIdx.x; k < ntile; k += gridDim.x) {
		const uint32_t lo = tile_lo[k];
		uint32_t hi = (k + 1 < ntile) ? tile_lo[k + 1] : qsize - 1;
		if (hi > qsize - 1) hi = qsize - 1;
		const uint32_t cnt = hi - lo + 1;
		for (uint32_t j = threadIdx.x; j < cnt; j += kBlock) {
			const uint32_t u = queue[lo + j];
			s_off[j] = qoff[lo + j];
			s_u[j] = u;
			s_base[j] = row_off[u];
		}
		__syncthreads();
		const uint32_t e0 = k * kSparseTile, e1 = min(e0 + kSparseTile, qdeg);
		for (uint32_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
			uint32_t a = 0, b = cnt;  // largest j with s_off[j] <= e
			while (b - a > 1) {
				const uint32_t mid = (a + b) >> 1;
				if (s_off[mid] <= e)
					a = mid;
				else
					b = mid;
			}
			const uint32_t idx = s_base[a] + (e - s_off[a]);
			const uint32_t v = cols[idx];
			uint32_t du[kSsspBatchMax];
			load_row(dist, s_u[a], du);
			if (relax_row(du, csr_w[idx], v, dist)) flag[v] = 1;
		}
		__syncthreads();
	}
}

// Dense step: a pull over the in-edges.  A tail takes, per source, the
// minimum of dist[h] + w over its in-edges from frontier heads, then lowers
// its own distances where that beats them; no atomics, where a push along
// the tiles makes one per improved (edge, source) pair (15-46 ms per dense
// step at scale 25 against 17 ms for a whole single-source query).  The sums
// are clamped to kInf, which never beats a distance.
constexpr uint32_t kPullRound = 4;  // in-edges per lane whose loads are issued together

template <bool kW8>
__device__ __forceinline__ void pull_edges(uint32_t e, uint32_t e_end, uint32_t stride,
                                           const uint32_t *__restrict__ in_src, const uint32_t *__restrict__ w32,
                                           const uint8_t *__restrict__ w8, const uint32_t *__restrict__ cur,
                                           const uint32_t *dist, uint32_t (&cand)[kSsspBatchMax])
{
	for (; e < e_end; e += kPullRound * stride) {
		uint32_t h[kPullRound], w[kPullRound];
		bool on[kPullRound];
#pragma unroll
		for (uint32_t u = 0; u < kPullRound; ++u) {
			const uint32_t x = e + u * stride;
			on[u] = x < e_end;
			h[u] = on[u] ? in_src[x] : 0u;
			w[u] = on[u] ? (kW8 ? (uint32_t)w8[x] : w32[x]) : 0u;
		}
#pragma unroll
		for (uint32_t u = 0; u < kPullRound; ++u) on[u] = on[u] && ((cur[h[u] >> 5] >> (h[u] & 31)) & 1u);
#pragma unroll
		for (uint32_t u = 0; u < kPullRound; ++u) {
			if (!on[u]) continue;
			uint32_t dh[kSsspBatchMax];
			load_row(dist, h[u], dh);
#pragma unroll
			for (int s = 0; s < (int)kSsspBatchMax; ++s) {
				const uint64_t nd = dh[s] == kInf ? (uint64_t)kInf : (uint64_t)dh[s] + w[u];
				const uint32_t c = nd < kInf ? (uint32_t)nd : kInf;
				cand[s] = c < cand[s] ? c : cand[s];
			}
		}
	}
}

// vertices with 1..kPullLight in-edges, a lane each (it owns the vertex:
// plain stores)
template <bool kW8>
__global__ __launch_bounds__(kBlock) void ssspb_pull_light_kernel(const StepCtrl *__restrict__ ctrl, uint32_t n,
                                                                  const uint32_t *__restrict__ in_off,
                                                                  const uint32_t *__restrict__ in_src,
                                                                  const uint32_t *__restrict__ w32,
                                                                  const uint8_t *__restrict__ w8,
                                                                  const uint32_t *__restrict__ cur, uint32_t *dist,
                                                                  uint8_t *flag)
{
	if (ctrl->done || ctrl->dir != 1) return;
	for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
		const uint32_t e0 = in_off[t], e1 = in_off[t + 1];
		if (e1 == e0 || e1 - e0 > kPullLight) continue;
		uint32_t cand[kSsspBatchMax];
#pragma unroll
		for (int s = 0; s < (int)kSsspBatchMax; ++s) cand[s] = kInf;
		pull_edges<kW8>(e0, e1, 1, in_src, w32, w8, cur, dist, cand);
		uint32_t dt[kSsspBatchMax];
		load_row(dist, t, dt);
		bool better = false;
#pragma unroll
		for (int s = 0; s < (int)kSsspBatchMax; ++s)
			if (cand[s] < dt[s]) {
				dist[(uint64_t)t * kSsspBatchMax + s] = cand[s];
				better = true;
			}
		if (better) flag[t] = 1;
	}
}

// pieces of up to kPullPiece in-edges of the other vertices, a wavefront
// each: lane minima, a wavefront minimum per source, then lane s lowers
// source s's distance (atomicMin: a vertex's pieces may run at once)
template <bool kW8>
__global__ __launch_bounds__(kBlock) void ssspb_pull_heavy_kernel(const StepCtrl *__restrict__ ctrl, uint32_t npieces,
                                                                  const uint2 *__restrict__ pieces,
                                                                  const uint32_t *__restrict__ in_off,
                                                                  const uint32_t *__restrict__ in_src,
                                                                  const uint32_t *__restrict__ w32,
                                                                  const uint8_t *__restrict__ w8,
                                                                  const uint32_t *__restrict__ cur, uint32_t *dist,
                                                                  uint8_t *flag)
{
	if (ctrl->done || ctrl->dir != 1) return;
	const uint32_t lane = threadIdx.x & 63, nwaves = (gridDim.x * kBlock) >> 6;
	for (uint32_t p = (blockIdx.x * kBlock + threadIdx.x) >> 6; p < npieces; p += nwaves) {
		const uint2 pc = pieces[p];
		const uint32_t t = pc.x, e1 = min(pc.y + kPullPiece, in_off[t + 1]);
		uint32_t cand[kSsspBatchMax];
#pragma unroll
		for (int s = 0; s < (int)kSsspBatchMax; ++s) cand[s] = kInf;
		pull_edges<kW8>(pc.y + lane, e1, 64, in_src, w32, w8, cur, dist, cand);
		uint32_t mine = kInf;  // lane s < 16: source s's minimum over the wavefront
#pragma unroll
		for (int s = 0; s < (int)kSsspBatchMax; ++s) {
			uint32_t c = cand[s];
#pragma unroll
			for (int o = 32; o > 0; o >>= 1) {
				const uint32_t x = (uint32_t)__shfl_xor((int)c, o, 64);
				c = x < c ? x : c;
			}
			if ((int)lane == s) mine = c;
		}
		bool better = false;
		if (lane < kSsspBatchMax) {
			const uint64_t i = (uint64_t)t * kSsspBatchMax + lane;
			if (mine < dist[i]) {
				atomicMin(dist + i, mine);
				better = true;
			}
		}
		if (__ballot(better) && lane == 0) flag[t] = 1;
	}
}

// row s (source s's distances) of the vertex-major array, for the copy-out
__global__ void ssspb_row_kernel(uint32_t n, uint32_t s, const uint32_t *__restrict__ dist, uint32_t *__restrict__ row)
{
	for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
		row[v] = dist[(uint64_t)v * kSsspBatchMax + s];
}

}  // namespace

// The traversal state, in-edge arrays with their weights, and the distance
// array of an SSSP batch on gr (idempotent).
int sssp_batch_prepare(gphi_graph *gr)
{
	gphi_ctx *ctx = gr->ctx;
	const DevGraph &G = gr->g;
	if (!G.weighted) return set_error(ctx, GPHI_E_STATE, "graph carries no weights");
	if (!G.has_csr || !G.has_tiles) return set_error(ctx, GPHI_E_STATE, "graph needs CSR and tile arrays");
	int rc = ensure_traversal_state(gr);
	if (rc) return rc;
	if ((rc = build_in_edges(gr)) || (rc = build_in_weights(gr))) return rc;
	if (!gr->sb_dist) {
		const size_t bytes = (size_t)G.n_pad * kSsspBatchMax * 4 + (size_t)kSsspBatchMax * 4;
		if (take_injected_alloc_failure())
			return set_error(ctx, GPHI_E_OOM, "hipMalloc(%zu bytes) failed: injected failure", bytes);
		if (hipMalloc((void **)&gr->sb_dist, bytes) != hipSuccess) {
			gr->sb_dist = nullptr;
			return set_error(ctx, GPHI_E_OOM, "hipMalloc(%zu bytes) for the SSSP batch failed", bytes);
		}
	}
	return GPHI_OK;
}

int run_sssp_batch(gphi_graph *gr, const uint32_t *srcs, uint32_t k, uint32_t *const *rows, gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	const DevGraph &G = gr->g;
	if (k == 0 || k > kSsspBatchMax)
		return set_error(ctx, GPHI_E_ARG, "SSSP batch size %u not in [1, %u]", k, kSsspBatchMax);
	if (!srcs) return set_error(ctx, GPHI_E_ARG, "sources are NULL");
	for (uint32_t i = 0; i < k; ++i)
		if (srcs[i] >= G.n) return set_error(ctx, GPHI_E_RANGE, "source %u >= N %u", srcs[i], G.n);
	int rc = sssp_batch_prepare(gr);
	if (rc) return rc;
	uint32_t *d_srcs = gr->sb_dist + (size_t)G.n_pad * kSsspBatchMax;
	if (st) memset(st, 0, sizeof(*st));
	hipStream_t s = ctx->stream;
	const bool timing = st != nullptr && ctx->ev_ok;
	if (timing) GPHI_HIP(ctx, hipEventRecord(ctx->ev[0], s));
	const uint64_t items = std::max<uint64_t>((uint64_t)G.n_pad * kSsspBatchMax / 4, G.n_words);
	ssspb_clear_kernel<<<(unsigned)std::min<uint64_t>((items + kBlock - 1) / kBlock, 16384), kBlock, 0, s>>>(
	    G.n_pad, G.n_words, reinterpret_cast<uint4 *>(gr->sb_dist), gr->flag, gr->cur, gr->ctrl);
	GPHI_HIP(ctx, hipMemcpyAsync(d_srcs, srcs, (size_t)k * 4, hipMemcpyHostToDevice, s));
	ssspb_seed_kernel<<<1, 64, 0, s>>>(d_srcs, k, gr->sb_dist, gr->flag);
	const uint32_t threshold = (uint32_t)(G.m / kSsspBatchDenseDiv);
	launch_frontier_update_driven(gr, threshold);  // opens level 0: the sources
	GPHI_HIP(ctx, hipGetLastError());
	const unsigned lgrid = std::min<unsigned>((G.n + kBlock - 1) / kBlock, kSsspGrid);
	const unsigned hgrid = std::max<unsigned>(1, std::min<unsigned>((gr->npieces + 3) / 4, kSsspGrid));
	StepCtrl *hc = ctx->h_ctrl;
	uint32_t enq = 0;
	for (;;) {
		for (uint32_t q = 0; q < kSsspLevelGroup; ++q, ++enq) {
			ssspb_sparse_kernel<<<kSsspGrid, kBlock, 0, s>>>(gr->ctrl, gr->queue_bufs(), G.row_off, G.cols,
			                                                 G.csr_w, gr->sb_dist, gr->flag);
			if (gr->in_w8) {
				ssspb_pull_light_kernel<true><<<lgrid, kBlock, 0, s>>>(gr->ctrl, G.n, gr->in_off, gr->in_src,
				                                                       nullptr, gr->in_w8, gr->cur, gr->sb_dist,
				                                                       gr->flag);
				if (gr->npieces)
					ssspb_pull_heavy_kernel<true><<<hgrid, kBlock, 0, s>>>(gr->ctrl, gr->npieces, gr->pieces,
					                                                       gr->in_off, gr->in_src, nullptr,
					                                                       gr->in_w8, gr->cur, gr->sb_dist, gr->flag);
			} else {
				ssspb_pull_light_kernel<false><<<lgrid, kBlock, 0, s>>>(gr->ctrl, G.n, gr->in_off, gr->in_src,
				                                                        gr->in_w32, nullptr, gr->cur, gr->sb_dist,
				                                                        gr->flag);
				if (gr->npieces)
					ssspb_pull_heavy_kernel<false><<<hgrid, kBlock, 0, s>>>(gr->ctrl, gr->npieces, gr->pieces,
					                                                        gr->in_off, gr->in_src, gr->in_w32,
					                                                        nullptr, gr->cur, gr->sb_dist, gr->flag);
			}
			launch_frontier_update_driven(gr, threshold);
		}
		GPHI_HIP(ctx, hipGetLastError());
		GPHI_HIP(ctx, hipMemcpyAsync(hc, gr->ctrl, sizeof(StepCtrl), hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		if (hc->done) break;
		// every step lowers some distance, and distances are bounded
		if (enq > (1u << 20)) return set_error(ctx, GPHI_E_STATE, "SSSP batch did not terminate");
	}
	if (timing) GPHI_HIP(ctx, hipEventRecord(ctx->ev[1], s));
	bool any_row = false;
	for (uint32_t i = 0; rows && i < k; ++i) any_row |= rows[i] != nullptr;
	if (any_row) {
		if (!gr->dist && hipMalloc((void **)&gr->dist, (size_t)G.n_pad * 4) != hipSuccess) {
			gr->dist = nullptr;
			return set_error(ctx, GPHI_E_OOM, "hipMalloc of the distance row failed");
		}
		// row i through the n-entry staging row (stream order keeps each copy
		// ahead of the next row's kernel)
		for (uint32_t i = 0; i < k; ++i) {
			if (!rows[i]) continue;
			ssspb_row_kernel<<<std::min<unsigned>((G.n + kBlock - 1) / kBlock, 8192), kBlock, 0, s>>>(
			    G.n, i, gr->sb_dist, gr->dist);
			GPHI_HIP(ctx, hipMemcpyAsync(rows[i], gr->dist, (size_t)G.n * 4, hipMemcpyDeviceToHost, s));
		}
	}
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	if (st) {
		st->steps = hc->level;
		if (timing) {
			float ms = 0.f;
			if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) st->total_ms = ms;
		}
	}
	return GPHI_OK;
}

}  // namespace gphi

// batch.hip -- multi-source BFS: k <= 256 sources in one traversal.
//
// Each vertex carries k-bit masks (NW 64-bit words): the sources that have
// reached it (seen), that reached it in the last step (front) and that reach
// it in the current step (next).  One edge stream serves every source:
//   dense step : a pull over the in-edges (next[t] = OR of front[h] over the
//                in-neighbours h, restricted to the sources t has not seen and
//                whose frontier is non-empty; a vertex stops scanning once all
//                of those are found) -- the bottom-up counterpart of the
//                reference's dense sweep (apps/bfs_simd/bfs.cpp:181-388), with
//                no atomics: a lane owns a vertex, or a wavefront owns a piece
//                of a high in-degree vertex
//   sparse step: the CSR push (:510-711) from the vertices whose front mask is
//                non-empty, next[t] |= front[h] & ~seen[t]
// and the direction rule |F| + sum(outdeg F) > M/T_RATIO (:829-834) is applied
// to the union frontier F.  A vertex v gets source s at level L+1 exactly when
// an in-neighbour had s in its front mask at level L and v had not seen s, so
// the level of every (source, vertex) pair equals that source's own BFS level.
// Levels are stored as uint8, source-major; 255 = unreached.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gphi {

namespace {

constexpr uint8_t kUnreached = 255;
typedef unsigned long long u64;

__device__ __forceinline__ bool test_bit32(const uint32_t *__restrict__ bm, uint32_t v)
{
	return (bm[v >> 5] >> (v & 31)) & 1u;
}

template <int NW>
struct Mask {
	u64 w[NW];
	__device__ __forceinline__ bool any() const
	{
		u64 a = 0;
#pragma unroll
		for (int i = 0; i < NW; ++i) a |= w[i];
		return a != 0;
	}
};

template <int NW>
__device__ __forceinline__ Mask<NW> load_mask(const u64 *__restrict__ p, uint32_t v)
{
	Mask<NW> m;
#pragma unroll
	for (int i = 0; i < NW; ++i) m.w[i] = p[(uint64_t)v * NW + i];
	return m;
}

// next[t] |= f & ~seen[t], skipping words already set in next
template <int NW>
__device__ __forceinline__ void push_mask(const Mask<NW> &f, uint32_t t, const u64 *__restrict__ seen,
                                          u64 *next)
{
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		if (!f.w[i]) continue;
		const u64 m = f.w[i] & ~seen[(uint64_t)t * NW + i];
		if (m && (m & ~next[(uint64_t)t * NW + i])) atomicOr(next + (uint64_t)t * NW + i, m);
	}
}

__device__ __forceinline__ u64 wave_or(u64 x)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
	return x;
}

// a value every lane holds, as a scalar (keeps loops over it uniform)
__device__ __forceinline__ u64 uniform64(u64 x)
{
	// readfirstlane returns int: go through uint32_t so the low half is not sign-extended
	const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
	const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
	return ((u64)hi << 32) | (u64)lo;
}

// acc |= front[h] (NW words) with 8- or 16-byte loads
template <int NW>
__device__ __forceinline__ void or_front(const u64 *__restrict__ front, uint32_t h, u64 (&acc)[NW])
{
	if constexpr (NW == 1) {
		acc[0] |= front[h];
	} else {
		const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(front) + (uint64_t)h * (NW / 2);
#pragma unroll
		for (int i = 0; i < NW / 2; ++i) {
			const ulonglong2 f = p[i];
			acc[2 * i] |= f.x;
			acc[2 * i + 1] |= f.y;
		}
	}
}

// acc |= front[h[j]] for the valid heads h[j] (0xFFFFFFFF = none).  A head
// outside the frontier has front = 0, so the frontier bitmap is only a
// filter: worth a 4-byte probe of a cache-resident word per edge when the
// frontier is small (most heads then skip the gather), not when it is large.
template <int NW, bool kFilter, uint32_t K>
__device__ __forceinline__ void gather_fronts(const uint32_t (&h)[K], const uint32_t *__restrict__ cur,
                                              const u64 *__restrict__ front, u64 (&acc)[NW])
{
	uint32_t use = 0;
#pragma unroll
	for (uint32_t j = 0; j < K; ++j)
		if (h[j] != 0xFFFFFFFFu && (!kFilter || test_bit32(cur, h[j]))) use |= 1u << j;
#pragma unroll
	for (uint32_t j = 0; j < K; ++j)
		if ((use >> j) & 1u) or_front<NW>(front, h[j], acc);
}

// Pull for vertices with 1..kPullLight in-edges.  Each lane first reads its
// own vertex (in-edge range, want = sources not yet seen there whose frontier
// is non-empty); the wavefront then serves its vertices that want something,
// kLightGroup at a time, with kLightGroup lanes per vertex taking every
// kLightGroup-th in-edge (at most kLightSlots each) and an OR across the
// group.  next[v] is zero on entry and written only by v's group.
constexpr uint32_t kLightGroup = 8;
constexpr uint32_t kLightSlots = kPullLight / kLightGroup;
// a wavefront serves one vertex per group per round and takes kLightGroup
// vertices per round, so there must be kLightGroup groups of kLightGroup lanes
static_assert(kLightGroup * kLightGroup == 64, "light pull: groups per wavefront == lanes per group");

__device__ __forceinline__ uint32_t shfl_u32(uint32_t x, uint32_t src) { return (uint32_t)__shfl((int)x, (int)src, 64); }

template <int NW, bool kFilter>
__global__ __launch_bounds__(kBlock) void batch_pull_light_kernel(
    uint32_t n, const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ in_src,
    const uint32_t *__restrict__ cur, const u64 *__restrict__ front, const u64 *__restrict__ seen,
    const u64 *__restrict__ live, u64 *__restrict__ next)
{
	const uint32_t lane = threadIdx.x & 63, g = lane % kLightGroup, grp = lane / kLightGroup;
	const uint32_t v0 = blockIdx.x * kBlock + (threadIdx.x & ~63u);  // first vertex of the wavefront
	const uint32_t v = v0 + lane;
	uint32_t e0 = 0, cnt = 0;
	u64 want[NW], any = 0;
	if (v < n) {
		e0 = in_off[v];
		cnt = in_off[v + 1] - e0;
		if (cnt > kPullLight) cnt = 0;
	}
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		want[i] = cnt ? ~seen[(uint64_t)v * NW + i] & live[i] : 0;
		any |= want[i];
	}
	u64 act = uniform64(__ballot(cnt != 0 && any != 0));
	while (act) {
		// the next kLightGroup wanting lanes of the wavefront, one per group
		uint32_t pick[kLightGroup];
#pragma unroll
		for (uint32_t q = 0; q < kLightGroup; ++q) {
			pick[q] = act ? (uint32_t)__ffsll((long long)act) - 1 : 64u;
			act &= act - 1;
		}
		uint32_t src = pick[0];
#pragma unroll
		for (uint32_t q = 1; q < kLightGroup; ++q) src = grp == q ? pick[q] : src;
		// every lane takes part in every shuffle (a lane left out of one would
		// hand 0 to the lanes reading it); groups without a vertex drop the result
		const uint32_t from = src & 63u;
		const uint32_t ge0 = shfl_u32(e0, from);
		const uint32_t ocnt = shfl_u32(cnt, from);
		const uint32_t gcnt = src < 64 ? ocnt : 0u;
		u64 gwant[NW], acc[NW];
#pragma unroll
		for (int i = 0; i < NW; ++i) {
			gwant[i] = __shfl(want[i], (int)from, 64);
			acc[i] = 0;
		}
		uint32_t h[kLightSlots];
#pragma unroll
		for (uint32_t j = 0; j < kLightSlots; ++j) {
			const uint32_t k = j * kLightGroup + g;
			h[j] = k < gcnt ? in_src[ge0 + k] : 0xFFFFFFFFu;
		}
		gather_fronts<NW, kFilter, kLightSlots>(h, cur, front, acc);
#pragma unroll
		for (int i = 0; i < NW; ++i)
#pragma unroll
			for (int o = kLightGroup / 2; o > 0; o >>= 1) acc[i] |= __shfl_xor(acc[i], o, kLightGroup);
		if (g == 0 && gcnt) {
			const uint64_t gv = v0 + src;
#pragma unroll
			for (int i = 0; i < NW; ++i) {
				const u64 x = acc[i] & gwant[i];
				if (x) next[gv * NW + i] = x;
			}
		}
	}
}

// Pull for vertices with more than kPullLight in-edges: a wavefront per piece
// of up to kPullPiece in-edges, 64 · kUnroll edges between checks of the stop rule.
// Pieces of one vertex combine with one atomicOr per word.
template <int NW, bool kFilter>
__global__ __launch_bounds__(kBlock) void batch_pull_heavy_kernel(
    uint32_t npieces, const uint2 *__restrict__ pieces, const uint32_t *__restrict__ in_off,
    const uint32_t *__restrict__ in_src, const uint32_t *__restrict__ cur, const u64 *__restrict__ front,
    const u64 *__restrict__ seen, const u64 *__restrict__ live, u64 *next)
{
	// loads in flight per lane between checks of the stop rule: a vertex stops
	// as soon as every wanted source is found, so short rounds (64 or 128
	// edges) beat deeper unrolling here
	constexpr uint32_t kUnroll = NW == 1 ? 1 : 2;
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t nwaves = (gridDim.x * kBlock) >> 6;
	for (uint32_t p = (blockIdx.x * kBlock + threadIdx.x) >> 6; p < npieces; p += nwaves) {
		const uint2 pc = pieces[p];
		const uint32_t v = pc.x;
		const uint32_t e1 = min(pc.y + kPullPiece, in_off[v + 1]);
		// want and red are the same in every lane (v is per wavefront); acc is
		// this lane's own, red the wavefront's OR of acc at the last check
		u64 want[NW], acc[NW], red[NW], any = 0;
#pragma unroll
		for (int i = 0; i < NW; ++i) {
			want[i] = uniform64(~seen[(uint64_t)v * NW + i] & live[i]);
			any |= want[i];
			acc[i] = 0;
			red[i] = 0;
		}
		if (!any) continue;
		const uint32_t cnt = e1 - pc.y;  // 1..kPullPiece
		for (uint32_t off = 0; off < cnt; off += 64 * kUnroll) {
			uint32_t h[kUnroll];
#pragma unroll
			for (uint32_t u = 0; u < kUnroll; ++u) {
				const uint32_t j = off + u * 64 + lane;
				h[u] = j < cnt ? in_src[pc.y + j] : 0xFFFFFFFFu;
			}
			gather_fronts<NW, kFilter, kUnroll>(h, cur, front, acc);
			u64 miss = 0;
#pragma unroll
			for (int i = 0; i < NW; ++i) {
				if (want[i] & ~red[i]) red[i] = uniform64(wave_or(acc[i]));  // a uniform branch
				miss |= want[i] & ~red[i];
			}
			if (!miss) break;
		}
		if (lane == 0) {
#pragma unroll
			for (int i = 0; i < NW; ++i) {
				const u64 x = red[i] & want[i];
				if (x) atomicOr(next + (uint64_t)v * NW + i, x);
			}
		}
	}
}

// CSR push from the union-frontier queue (same edge partition as the
// single-source push: tiles of kSparseTile edges starting at tile_lo)
template <int NW>
__global__ __launch_bounds__(kBlock) void batch_sparse_kernel(
    const uint32_t *__restrict__ queue, const uint32_t *__restrict__ qoff,
    const uint32_t *__restrict__ tile_lo, uint32_t qsize, uint32_t qdeg,
    const uint32_t *__restrict__ row_off, const uint32_t *__restrict__ cols,
    const u64 *__restrict__ front, const u64 *__restrict__ seen, u64 *next)
{
	__shared__ uint32_t s_off[kSparseTile + 2];
	__shared__ uint32_t s_u[kSparseTile + 2];
	__shared__ uint32_t s_base[kSparseTile + 2];
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
			uint32_t a = 0, b = cnt;
			while (b - a > 1) {
				const uint32_t mid = (a + b) >> 1;
				if (s_off[mid] <= e)
					a = mid;
				else
					b = mid;
			}
			const uint32_t u = s_u[a];
			const uint32_t v = cols[s_base[a] + (e - s_off[a])];
			push_mask<NW>(load_mask<NW>(front, u), v, seen, next);
		}
		__syncthreads();
	}
}

// Per vertex: new = next & ~seen; seen |= new; front = new; next = 0; the
// level of every new source is L+1; flag marks the union frontier for the
// shared mark/scan/emit update.  Level bytes are written one source at a time
// across the wavefront (64 consecutive vertices, one coalesced store), and each
// block publishes the sources it saw advance into one of kLiveSlots slots.
constexpr uint32_t kLiveSlots = 64;
constexpr uint32_t kLiveStride = 16;  // u64 words per slot (one 128-byte line)

template <int NW>
__global__ __launch_bounds__(kBlock) void batch_update_kernel(uint32_t n, uint32_t n_pitch, uint32_t next_level,
                                                              const uint32_t *__restrict__ cur,
                                                              u64 *__restrict__ front, u64 *__restrict__ seen,
                                                              u64 *__restrict__ next, uint8_t *__restrict__ level,
                                                              uint8_t *__restrict__ flag, u64 *slots)
{
	__shared__ u64 s_or[kBlock / 64][NW];
	const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
	const bool in = v < n;
	// next first: a vertex nothing reached this step only has to drop its
	// front mask, and only if it was in the (union) frontier, whose bitmap is
	// cur -- so most vertices cost one read of next
	u64 nx[NW], any_nx = 0;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		nx[i] = in ? next[(uint64_t)v * NW + i] : 0ull;
		any_nx |= nx[i];
	}
	const bool was_front = in && test_bit32(cur, v);
	u64 nwm[NW];
	bool any = false;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		nwm[i] = 0;
		const uint64_t idx = (uint64_t)v * NW + i;
		if (any_nx) {
			const u64 sn = seen[idx];
			const u64 nw = nx[i] & ~sn;
			if (nw || was_front) front[idx] = nw;
			if (nx[i]) next[idx] = 0;
			if (nw) {
				seen[idx] = sn | nw;
				any = true;
			}
			nwm[i] = nw;
		} else if (was_front) {
			front[idx] = 0;
		}
	}
	const uint32_t wid = threadIdx.x >> 6;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		u64 u = uniform64(wave_or(nwm[i]));
		if ((threadIdx.x & 63) == 0) s_or[wid][i] = u;
		while (u) {
			const uint32_t b = (uint32_t)__ffsll((long long)u) - 1;
			u &= u - 1;
			if ((nwm[i] >> b) & 1) level[(uint64_t)(i * 64 + b) * n_pitch + v] = (uint8_t)next_level;
		}
	}
	if (any) flag[v] = 1;
	__syncthreads();
	if (threadIdx.x < NW) {
		u64 u = 0;
#pragma unroll
		for (uint32_t w = 0; w < kBlock / 64; ++w) u |= s_or[w][threadIdx.x];
		u64 *slot = slots + (uint64_t)(blockIdx.x % kLiveSlots) * kLiveStride + threadIdx.x;
		// read the slot at L2 (not a stale L1 copy) and add only new sources
		const u64 have = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (u & ~have) atomicOr(slot, u);
	}
}

// live = OR of the slots (the sources whose new frontier is non-empty); the
// slots are cleared for the next step.  One wavefront.
template <int NW>
__global__ void batch_live_kernel(u64 *live, u64 *slots)
{
	const uint32_t lane = threadIdx.x;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		u64 *p = slots + (uint64_t)lane * kLiveStride + i;
		const u64 x = wave_or(*p);
		*p = 0;
		if (lane == 0) live[i] = x;
	}
}

__global__ void batch_seed_kernel(const uint32_t *__restrict__ srcs, uint32_t k, uint32_t nw,
                                  uint32_t n_pitch, u64 *front, u64 *seen, uint8_t *level,
                                  uint8_t *flag, u64 *live)
{
	for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
		const uint32_t v = srcs[i];
		atomicOr(front + (uint64_t)v * nw + i / 64, 1ull << (i % 64));
		atomicOr(seen + (uint64_t)v * nw + i / 64, 1ull << (i % 64));
		atomicOr(live + i / 64, 1ull << (i % 64));
		level[(uint64_t)i * n_pitch + v] = 0;
		flag[v] = 1;
	}
}

// level 255 for every (source, vertex) pair the traversal did not reach: the
// level bytes are written for reached pairs only, so this runs only when the
// levels are copied out.  One lane per vertex, one coalesced store per source.
template <int NW>
__global__ __launch_bounds__(kBlock) void batch_finalize_kernel(uint32_t n, uint32_t n_pitch, uint32_t k,
                                                                const u64 *__restrict__ seen, uint8_t *level)
{
	const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
	if (v >= n) return;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		const u64 sn = seen[(uint64_t)v * NW + i];
		for (uint32_t b = 0; b < 64 && i * 64 + b < k; ++b)
			if (!((sn >> b) & 1ull)) level[(uint64_t)(i * 64 + b) * n_pitch + v] = kUnreached;
	}
}

// totals: sum over v of popcount(seen[v]) and of deg[v] * popcount(seen[v])
template <int NW>
__global__ __launch_bounds__(kBlock) void batch_sums_kernel(uint32_t n, const u64 *__restrict__ seen,
                                                            const uint32_t *__restrict__ deg, u64 *sums)
{
	u64 r = 0, e = 0;
	for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
		uint32_t c = 0;
#pragma unroll
		for (int i = 0; i < NW; ++i) c += (uint32_t)__popcll(seen[(uint64_t)v * NW + i]);
		r += c;
		e += (u64)c * deg[v];
	}
	for (int o = 32; o > 0; o >>= 1) {
		r += __shfl_xor(r, o, 64);
		e += __shfl_xor(e, o, 64);
	}
	if ((threadIdx.x & 63) == 0) {
		atomicAdd(sums, r);
		atomicAdd(sums + 1, e);
	}
}

unsigned grid_n(uint64_t items) { return (unsigned)std::max<uint64_t>(1, (items + kBlock - 1) / kBlock); }

float ev_ms(hipEvent_t a, hipEvent_t b)
{
	float ms = 0.f;
	return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.f;
}

template <int NW, bool kFilter>
void launch_pull(gphi_graph *gr, const u64 *live)
{
	const DevGraph &G = gr->g;
	hipStream_t s = gr->ctx->stream;
	batch_pull_light_kernel<NW, kFilter><<<grid_n(G.n), kBlock, 0, s>>>(
	    G.n, gr->in_off, gr->in_src, gr->cur, gr->b_front, gr->b_seen, live, gr->b_next);
	if (gr->npieces)
		batch_pull_heavy_kernel<NW, kFilter><<<std::min<uint32_t>((gr->npieces + 3) / 4, 16384), kBlock, 0, s>>>(
		    gr->npieces, gr->pieces, gr->in_off, gr->in_src, gr->cur, gr->b_front, gr->b_seen, live, gr->b_next);
}

template <int NW>
int batch_loop(gphi_graph *gr, const uint32_t *d_srcs, uint32_t k, gphi_stats *st, u64 *sums_host)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	const uint32_t threshold = (uint32_t)G.m / ctx->t_ratio;
	StepCtrl *hc = ctx->h_ctrl;
	hipEvent_t *ev = ctx->ev;
	// total time from two events; per-step events only with level timing on
	const bool total_timing = st != nullptr && ctx->ev_ok;
	const bool timing = total_timing && ctx->level_timing;
	const size_t mask_bytes = (size_t)G.n_pad * NW * 8;
	if (total_timing) GPHI_HIP(ctx, hipEventRecord(ev[0], s));
	GPHI_HIP(ctx, hipMemsetAsync(gr->b_front, 0, mask_bytes, s));
	GPHI_HIP(ctx, hipMemsetAsync(gr->b_seen, 0, mask_bytes, s));
	GPHI_HIP(ctx, hipMemsetAsync(gr->b_next, 0, mask_bytes, s));
	GPHI_HIP(ctx, hipMemsetAsync(gr->flag, 0, G.n_pad, s));
	GPHI_HIP(ctx, hipMemsetAsync(gr->cur, 0, (size_t)G.n_words * 4, s));
	GPHI_HIP(ctx, hipMemsetAsync(gr->b_live, 0, (size_t)(1 + kLiveSlots) * kLiveStride * 8, s));
	u64 *live = gr->b_live, *slots = gr->b_live + kLiveStride;
	launch_init_blocks(gr);
	batch_seed_kernel<<<1, kBlock, 0, s>>>(d_srcs, k, NW, G.n_pad, gr->b_front, gr->b_seen, gr->b_level,
	                                       gr->flag, live);
	// the sources form the first (union) frontier
	launch_frontier_update_flags(gr, 0, threshold);
	GPHI_HIP(ctx, hipGetLastError());
	if (timing) GPHI_HIP(ctx, hipEventRecord(ev[1], s));
	GPHI_HIP(ctx, hipMemcpyAsync(hc, gr->ctrl, sizeof(StepCtrl), hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	uint32_t steps = 0;
	while (!hc->done) {
		const uint32_t L = steps;
		// the rule applies to the union frontier from the first step on
		const uint32_t dir = hc->dir;
		if (st && steps < GPHI_MAX_STEPS) {
			st->dir[steps] = dir;
			st->fsize[steps] = hc->fsize;
			st->fdeg[steps] = hc->fdeg;
		}
		const bool ev_step = timing && steps < GPHI_MAX_STEPS - 1;
		if (dir == 0) {
			if (hc->fdeg)
				batch_sparse_kernel<NW><<<(unsigned)std::min<uint64_t>((hc->fdeg + kSparseTile - 1) / kSparseTile, 4096),
				                          kBlock, 0, s>>>(gr->queue, gr->qoff, gr->tile_lo, hc->qsize, hc->fdeg,
				                                          G.row_off, G.cols, gr->b_front, gr->b_seen, gr->b_next);
		} else {
			// the frontier filter pays while the union frontier is under 1/8 of the vertices
			if (hc->fsize < G.n / 8)
				launch_pull<NW, true>(gr, live);
			else
				launch_pull<NW, false>(gr, live);
		}
		GPHI_HIP(ctx, hipGetLastError());
		if (ev_step) GPHI_HIP(ctx, hipEventRecord(ev[2 + 2 * steps], s));
		batch_update_kernel<NW><<<grid_n(G.n), kBlock, 0, s>>>(G.n, G.n_pad, L + 1, gr->cur, gr->b_front,
		                                                     gr->b_seen, gr->b_next, gr->b_level, gr->flag, slots);
		batch_live_kernel<NW><<<1, 64, 0, s>>>(live, slots);
		launch_frontier_update_flags(gr, L + 1, threshold);
		GPHI_HIP(ctx, hipGetLastError());
		if (ev_step) GPHI_HIP(ctx, hipEventRecord(ev[3 + 2 * steps], s));
		GPHI_HIP(ctx, hipMemcpyAsync(hc, gr->ctrl, sizeof(StepCtrl), hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		// a pair reached at level 255 would read as unreached: the expansion
		// of level 254 may run, but only if it reaches nothing new
		if (L + 1 >= kUnreached && !hc->done)
			return set_error(ctx, GPHI_E_RANGE, "a batched BFS deeper than 254 levels does not fit uint8 levels");
		++steps;
		if (steps > G.n + 2) return set_error(ctx, GPHI_E_STATE, "batched traversal did not terminate");
	}
	GPHI_HIP(ctx, hipMemsetAsync(gr->b_sums, 0, 16, s));
	batch_sums_kernel<NW><<<std::min<unsigned>(grid_n(G.n), 4096), kBlock, 0, s>>>(G.n, gr->b_seen, G.deg,
	                                                                               gr->b_sums);
	if (total_timing) GPHI_HIP(ctx, hipEventRecord(ev[2 * GPHI_MAX_STEPS + 2], s));
	GPHI_HIP(ctx, hipMemcpyAsync(sums_host, gr->b_sums, 16, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	if (st) {
		st->steps = std::min<uint32_t>(steps, GPHI_MAX_STEPS);
		if (timing) {
			st->init_ms = ev_ms(ev[0], ev[1]);
			hipEvent_t prev = ev[1];
			const uint32_t ns = std::min<uint32_t>(steps, GPHI_MAX_STEPS - 1);
			for (uint32_t q = 0; q < ns; ++q) {
				st->sweep_ms[q] = ev_ms(prev, ev[2 + 2 * q]);
				st->update_ms[q] = ev_ms(ev[2 + 2 * q], ev[3 + 2 * q]);
				prev = ev[3 + 2 * q];
			}
		} else {
			memset(st->sweep_ms, 0, sizeof(st->sweep_ms));
			memset(st->update_ms, 0, sizeof(st->update_ms));
			st->init_ms = 0.f;
		}
		st->total_ms = total_timing ? ev_ms(ev[0], ev[2 * GPHI_MAX_STEPS + 2]) : 0.f;
	}
	return GPHI_OK;
}

}  // namespace

int run_bfs_batch(gphi_graph *gr, const uint32_t *srcs, uint32_t k, uint8_t *levels, gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	if (k == 0 || k > 256) return set_error(ctx, GPHI_E_ARG, "batch size %u not in [1, 256]", k);
	if (!srcs) return set_error(ctx, GPHI_E_ARG, "sources are NULL");
	for (uint32_t i = 0; i < k; ++i)
		if (srcs[i] >= G.n) return set_error(ctx, GPHI_E_RANGE, "source %u >= N %u", srcs[i], G.n);
	if (!G.has_csr || !G.has_tiles) return set_error(ctx, GPHI_E_STATE, "graph needs CSR and tile arrays");
	int rc = ensure_traversal_state(gr);
	if (rc) return rc;
	if ((rc = build_in_edges(gr))) return rc;
	const uint32_t nw = k <= 64 ? 1 : k <= 128 ? 2 : 4;
	if (gr->batch_words < nw || !gr->b_level) {
		void *ps[] = {gr->b_front, gr->b_seen, gr->b_next, gr->b_level, gr->b_sums, gr->b_live};
		for (void *p : ps)
			if (p) (void)hipFree(p);
		gr->b_front = gr->b_seen = gr->b_next = gr->b_sums = gr->b_live = nullptr;
		gr->b_level = nullptr;
		gr->batch_words = 0;
		const size_t mb = (size_t)G.n_pad * nw * 8;
		hipError_t e = hipMalloc((void **)&gr->b_front, mb);
		if (e == hipSuccess) e = hipMalloc((void **)&gr->b_seen, mb);
		if (e == hipSuccess) e = hipMalloc((void **)&gr->b_next, mb);
		if (e == hipSuccess) e = hipMalloc((void **)&gr->b_level, (size_t)nw * 64 * G.n_pad);
		if (e == hipSuccess) e = hipMalloc((void **)&gr->b_sums, 16 + 256 * 4);  // totals, then the sources
		if (e == hipSuccess) e = hipMalloc((void **)&gr->b_live, (size_t)(1 + kLiveSlots) * kLiveStride * 8);
		if (e != hipSuccess)
			return set_error(ctx, GPHI_E_OOM, "batched BFS buffers (%zu MB): %s",
			                 (3 * mb + (size_t)nw * 64 * G.n_pad) >> 20, hipGetErrorString(e));
		gr->batch_words = nw;
	}
	hipStream_t s = ctx->stream;
	uint32_t *d_srcs = reinterpret_cast<uint32_t *>(gr->b_sums + 2);
	GPHI_HIP(ctx, hipMemcpyAsync(d_srcs, srcs, (size_t)k * 4, hipMemcpyHostToDevice, s));
	if (st) memset(st, 0, sizeof(*st));
	u64 sums[2] = {0, 0};
	// the buffers hold gr->batch_words >= nw words per vertex; a batch runs at
	// the width its own k needs (a 64-source batch after a 256-source one
	// would otherwise gather and update four words per vertex)
	switch (nw) {
	case 1: rc = batch_loop<1>(gr, d_srcs, k, st, sums); break;
	case 2: rc = batch_loop<2>(gr, d_srcs, k, st, sums); break;
	default: rc = batch_loop<4>(gr, d_srcs, k, st, sums); break;
	}
	if (rc) return rc;
	if (st) {
		st->reached = sums[0];
		st->edges_traversed = sums[1];
	}
	if (levels) {
		const unsigned fg = grid_n(G.n);
		if (nw == 1)
			batch_finalize_kernel<1><<<fg, kBlock, 0, s>>>(G.n, G.n_pad, k, gr->b_seen, gr->b_level);
		else if (nw == 2)
			batch_finalize_kernel<2><<<fg, kBlock, 0, s>>>(G.n, G.n_pad, k, gr->b_seen, gr->b_level);
		else
			batch_finalize_kernel<4><<<fg, kBlock, 0, s>>>(G.n, G.n_pad, k, gr->b_seen, gr->b_level);
		GPHI_HIP(ctx, hipGetLastError());
		GPHI_HIP(ctx, hipMemcpy2DAsync(levels, G.n, gr->b_level, G.n_pad, G.n, k, hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
	}
	return GPHI_OK;
}

}  // namespace gphi

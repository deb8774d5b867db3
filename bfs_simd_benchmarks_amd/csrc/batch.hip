// batch.hip -- multi-source BFS: up to 256 sources in one traversal.
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
//
// The level loop runs on the device, as the single-source one does: every
// kernel of a level reads the direction and sizes from the control block and
// returns at once when it has nothing to do, so the host enqueues groups of
// levels and reads the block back once per group.  A step marks the vertices
// whose next mask it set in a bitmap, and the per-level update visits only
// those (and the previous frontier), not every vertex.
//
// The union direction rule can make a dense step for sources whose own
// frontiers are still small.  A pull finds such a source only far into a
// vertex's in-edges, so wanting it keeps the heavy pieces scanning to their
// end.  The update counts each source's new vertices; a source whose
// frontier is estimated to have fewer than M / kSparseRatio out-edges is
// left out of the next dense step's pulls and pushed instead, from a queue of
// its frontier vertices.  Which sources go which way changes only the work,
// not the levels.
//
// Sources may also join a running traversal between two levels (the serving
// loop's continuous batching): a source that joins at level J gets a slot of
// its own, its vertex enters the union frontier, and the levels stored for
// it are the traversal's level minus J -- its own BFS levels, since its bit
// spreads exactly as if it had started alone at level J.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gphi {

namespace {

constexpr uint8_t kUnreached = 255;
typedef unsigned long long u64;

// b_aux layout (uint32 words): sources to seed, their slots, per-slot start level
constexpr uint32_t kAuxSrc = 0, kAuxSlot = kBatchMaxSources, kAuxStart = 2 * kBatchMaxSources;
constexpr uint32_t kAuxWords = 3 * kBatchMaxSources;
// the sources of a dense step whose frontiers have few out-edges (sparse
// sources, batch_live_kernel); their mask sits beside the live mask
constexpr uint32_t kSparseWord = 4;     // live[4..7]: the sparse sources of the next frontier
constexpr uint32_t kSparseCount = 8;    // live[8]: entries of the sparse-source queue
constexpr uint32_t kSparseOverflow = 9; // live[9]: the queue overflowed (the pulls take the sparse sources)
// sparse: fewer than M / 1536 frontier out-edges (estimated).  Measured at
// scale 25 over ratios 640..2560: 960..1920 is best, past it the pull keeps
// scanning for the sources left to it, below it the push takes too many
constexpr uint32_t kSparseRatio = 1536;
constexpr uint32_t kCountSample = 16;   // after a dense step, every 16th block counts new vertices per source

__device__ __forceinline__ bool test_bit32(const uint32_t *__restrict__ bm, uint32_t v)
{
	return (bm[v >> 5] >> (v & 31)) & 1u;
}

__device__ __forceinline__ void mark_touched(uint32_t *touched, uint32_t v)
{
	atomicOr(touched + (v >> 5), 1u << (v & 31));
}

template <int NW>
struct Mask {
	u64 w[NW];
};

template <int NW>
__device__ __forceinline__ Mask<NW> load_mask(const u64 *__restrict__ p, uint32_t v)
{
	Mask<NW> m;
#pragma unroll
	for (int i = 0; i < NW; ++i) m.w[i] = p[(uint64_t)v * NW + i];
	return m;
}

// next[t] |= f & ~seen[t], skipping words already set in next; whoever sets
// the first bit of a word marks t touched
template <int NW>
__device__ __forceinline__ void push_mask(const Mask<NW> &f, uint32_t t, const u64 *__restrict__ seen,
                                          u64 *next, uint32_t *touched)
{
	bool first = false;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		if (!f.w[i]) continue;
		const u64 m = f.w[i] & ~seen[(uint64_t)t * NW + i];
		if (m && (m & ~next[(uint64_t)t * NW + i])) first |= atomicOr(next + (uint64_t)t * NW + i, m) == 0ull;
	}
	if (first) mark_touched(touched, t);
}

__device__ __forceinline__ u64 wave_or(u64 x)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
	return x;
}

__device__ __forceinline__ u64 wave_add(u64 x)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
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

// The pull kernels of a dense step: a variant runs only on a dense step and
// only if its filter setting is the one the step's union frontier calls for
// (the frontier filter pays while the union frontier is under 1/8 of the
// vertices).
template <bool kFilter>
__device__ __forceinline__ bool pull_runs(const StepCtrl *ctrl, uint32_t n)
{
	return !ctrl->done && ctrl->dir == 1 && (ctrl->fsize < n / 8) == kFilter;
}

// The sources a dense step pulls: the live ones, less the sparse ones, which
// batch_sparse_src_push_kernel pushes (unless their queue overflowed).
template <int NW>
__device__ __forceinline__ void pulled_sources(const u64 *__restrict__ live, u64 (&pull)[NW])
{
	const bool pushed = live[kSparseOverflow] == 0;
#pragma unroll
	for (int i = 0; i < NW; ++i) pull[i] = live[i] & (pushed ? ~live[kSparseWord + i] : ~0ull);
}

// Pull for vertices with 1..kPullLight in-edges.  Each lane first reads its
// own vertex (in-edge range, want = sources not yet seen there whose frontier
// is non-empty).  A vertex with at most kLightSlots in-edges is served by its
// own lane; the wavefront then serves its other vertices that want something,
// kLightGroup at a time, with kLightGroup lanes per vertex taking every
// kLightGroup-th in-edge (at most kLightSlots each) and an OR across the
// group: first the vertex's first 2 * kLightGroup in-edges, then the rest only
// if a wanted source is still missing (after the first dense step most
// vertices find every source they want among their first in-edges).  next[v] is zero on entry and written only by v's lane or group; the
// wavefront's 64 vertices are two whole touched-bitmap words, which it
// writes with plain stores (the heavy pull, a later launch, ORs into them).
constexpr uint32_t kLightGroup = 8;
constexpr uint32_t kLightSlots = kPullLight / kLightGroup;
// a wavefront serves one vertex per group per round and takes kLightGroup
// vertices per round, so there must be kLightGroup groups of kLightGroup lanes
static_assert(kLightGroup * kLightGroup == 64, "light pull: groups per wavefront == lanes per group");

__device__ __forceinline__ uint32_t shfl_u32(uint32_t x, uint32_t src) { return (uint32_t)__shfl((int)x, (int)src, 64); }

template <int NW, bool kFilter>
__global__ __launch_bounds__(kBlock) void batch_pull_light_kernel(
    uint32_t n, const StepCtrl *__restrict__ ctrl, const uint32_t *__restrict__ in_off,
    const uint32_t *__restrict__ in_src, const uint32_t *__restrict__ cur, const u64 *__restrict__ front,
    const u64 *__restrict__ seen, const u64 *__restrict__ live, u64 *__restrict__ next,
    uint32_t *__restrict__ touched)
{
	if (!pull_runs<kFilter>(ctrl, n)) return;
	u64 pull[NW];
	pulled_sources<NW>(live, pull);
	const uint32_t lane = threadIdx.x & 63, g = lane % kLightGroup, grp = lane / kLightGroup;
	// blocks walk the vertices kBlock at a time: a resident-size grid, so a
	// launch on a step it has nothing to do in returns at once
	for (uint32_t vb = blockIdx.x * kBlock; vb < n; vb += gridDim.x * kBlock) {
	const uint32_t v0 = vb + (threadIdx.x & ~63u);  // first vertex of the wavefront
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
		want[i] = cnt ? ~seen[(uint64_t)v * NW + i] & pull[i] : 0;
		any |= want[i];
	}
	u64 got = 0;  // the wavefront's vertices (bit = lane) that got a source
	// a vertex with at most kLightSlots in-edges (most of them) is served by
	// its own lane in one pass: all its heads loaded, then all gathers
	const bool small = cnt != 0 && cnt <= kLightSlots && any != 0;
	if (__ballot(small)) {  // uniform
		uint32_t h[kLightSlots];
#pragma unroll
		for (uint32_t j = 0; j < kLightSlots; ++j) h[j] = small && j < cnt ? in_src[e0 + j] : 0xFFFFFFFFu;
		u64 acc[NW];
#pragma unroll
		for (int i = 0; i < NW; ++i) acc[i] = 0;
		gather_fronts<NW, kFilter, kLightSlots>(h, cur, front, acc);
		bool hit = false;
#pragma unroll
		for (int i = 0; i < NW; ++i) {
			const u64 x = acc[i] & want[i];
			if (small && x) {
				next[(uint64_t)v * NW + i] = x;
				hit = true;
			}
		}
		if (hit) got = 1ull << lane;
	}
	// the others, kLightGroup lanes per vertex
	u64 act = uniform64(__ballot(cnt > kLightSlots && any != 0));
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
		// the vertex's first 2 * kLightGroup in-edges, then the rest only for
		// a vertex that still misses a wanted source (its group decides; the
		// whole wavefront runs both rounds, so every lane joins the shuffles)
		constexpr uint32_t kFirst = 2;
		uint32_t h1[kFirst];
#pragma unroll
		for (uint32_t j = 0; j < kFirst; ++j) {
			const uint32_t k = j * kLightGroup + g;
			h1[j] = k < gcnt ? in_src[ge0 + k] : 0xFFFFFFFFu;
		}
		gather_fronts<NW, kFilter, kFirst>(h1, cur, front, acc);
#pragma unroll
		for (int i = 0; i < NW; ++i)
#pragma unroll
			for (int o = kLightGroup / 2; o > 0; o >>= 1) acc[i] |= __shfl_xor(acc[i], o, kLightGroup);
		bool miss = false;
#pragma unroll
		for (int i = 0; i < NW; ++i) miss |= (gwant[i] & ~acc[i]) != 0;
		if (__ballot(miss && gcnt > kFirst * kLightGroup)) {  // uniform
			uint32_t h2[kLightSlots - kFirst];
#pragma unroll
			for (uint32_t j = kFirst; j < kLightSlots; ++j) {
				const uint32_t k = j * kLightGroup + g;
				h2[j - kFirst] = miss && k < gcnt ? in_src[ge0 + k] : 0xFFFFFFFFu;
			}
			u64 acc2[NW];
#pragma unroll
			for (int i = 0; i < NW; ++i) acc2[i] = 0;
			gather_fronts<NW, kFilter, kLightSlots - kFirst>(h2, cur, front, acc2);
#pragma unroll
			for (int i = 0; i < NW; ++i) {
#pragma unroll
				for (int o = kLightGroup / 2; o > 0; o >>= 1) acc2[i] |= __shfl_xor(acc2[i], o, kLightGroup);
				acc[i] |= acc2[i];
			}
		}
		if (g == 0 && gcnt) {
			const uint64_t gv = v0 + src;
			bool hit = false;
#pragma unroll
			for (int i = 0; i < NW; ++i) {
				const u64 x = acc[i] & gwant[i];
				if (x) {
					next[gv * NW + i] = x;
					hit = true;
				}
			}
			if (hit) got |= 1ull << src;
		}
	}
	got = wave_or(got);
	if (lane == 0 && got) {
		const uint32_t w0 = v0 >> 5;
		if ((uint32_t)got) touched[w0] = (uint32_t)got;
		if ((uint32_t)(got >> 32)) touched[w0 + 1] = (uint32_t)(got >> 32);
	}
	}
}

// Pull for vertices with more than kPullLight in-edges: a wavefront per piece
// of up to kPullPiece in-edges, 64 · kUnroll edges between checks of the stop rule.
// Pieces of one vertex combine with one atomicOr per word.  The sparse
// sources (few frontier out-edges, batch_live_kernel) are left to
// batch_sparse_src_push_kernel: a pull finds them only far into a vertex's
// in-edges, so wanting them would keep most pieces scanning to their end.
template <int NW, bool kFilter>
__global__ __launch_bounds__(kBlock) void batch_pull_heavy_kernel(
    uint32_t n, const StepCtrl *__restrict__ ctrl, uint32_t npieces, const uint2 *__restrict__ pieces,
    const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ in_src, const uint32_t *__restrict__ cur,
    const u64 *__restrict__ front, const u64 *__restrict__ seen, const u64 *__restrict__ live, u64 *next,
    uint32_t *touched)
{
	if (!pull_runs<kFilter>(ctrl, n)) return;
	u64 pull[NW];
	pulled_sources<NW>(live, pull);
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
			want[i] = uniform64(~seen[(uint64_t)v * NW + i] & pull[i]);
			any |= want[i];
			acc[i] = 0;
			red[i] = 0;
		}
		if (!any) continue;
		const uint32_t cnt = e1 - pc.y;  // 1..kPullPiece
		// the heads of round r+1 are loaded before the fronts of round r are
		// gathered, so a round waits on one load, not two in sequence
		uint32_t hn[kUnroll];
#pragma unroll
		for (uint32_t u = 0; u < kUnroll; ++u) {
			const uint32_t j = u * 64 + lane;
			hn[u] = j < cnt ? in_src[pc.y + j] : 0xFFFFFFFFu;
		}
		for (uint32_t off = 0; off < cnt; off += 64 * kUnroll) {
			uint32_t h[kUnroll];
#pragma unroll
			for (uint32_t u = 0; u < kUnroll; ++u) {
				h[u] = hn[u];
				const uint32_t j = off + 64 * kUnroll + u * 64 + lane;
				hn[u] = j < cnt ? in_src[pc.y + j] : 0xFFFFFFFFu;
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
			bool hit = false;
#pragma unroll
			for (int i = 0; i < NW; ++i) {
				const u64 x = red[i] & want[i];
				if (x) {
					atomicOr(next + (uint64_t)v * NW + i, x);
					hit = true;
				}
			}
			if (hit) mark_touched(touched, v);
		}
	}
}

// CSR push from the union-frontier queue (same edge partition as the
// single-source push: tiles of kSparseTile edges starting at tile_lo)
template <int NW>
__global__ __launch_bounds__(kBlock) void batch_sparse_kernel(
    const StepCtrl *__restrict__ ctrl, QueueBufs qb, const uint32_t *__restrict__ row_off,
    const uint32_t *__restrict__ cols, const u64 *__restrict__ front, const u64 *__restrict__ seen, u64 *next,
    uint32_t *touched)
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
			push_mask<NW>(load_mask<NW>(front, u), v, seen, next, touched);
		}
		__syncthreads();
	}
}

// The level update, one lane per vertex.  A lane first reads the touched and
// frontier bitmap words of its vertex (one word serves 32 lanes); a vertex
// with neither bit has nothing to do, and a wavefront with none skips the
// rest, so a sparse step costs little beyond the two bitmaps.  Per vertex:
// new = next & ~seen; seen |= new; front = new; next = 0; a vertex of the
// old frontier with nothing new drops its front mask.  The level of every
// new (source, vertex) pair is the level minus the source's start level,
// written one source at a time across the wavefront (64 consecutive
// vertices, one coalesced byte store).  Blocks publish the sources they saw
// advance into one of kLiveSlots slots and add the new pairs (and their
// out-degrees) to the totals.
//
// Two launch shapes, each returning at once on the other's steps: after a
// dense step (kSpan = false) a block per kBlock vertices, since the pulls
// touch most vertices and independent blocks hide their latency best; after a
// sparse step (kSpan = true) a block per kUpdateSpan vertices, which reads
// their touched and frontier words in one round into LDS and then walks
// them, so that a step touching few vertices costs one load round per block,
// not one per kBlock vertices.
constexpr uint32_t kUpdateSpan = 4096;  // vertices per block after a sparse step (divides N_pad)
constexpr uint32_t kLiveSlots = 64;
constexpr uint32_t kLiveStride = 16;  // u64 words per slot (one 128-byte line)
constexpr uint32_t kSlotPairs = 8;    // slot words 8, 9: new pairs and their out-degrees

template <int NW, bool kSpan>
__global__ __launch_bounds__(kBlock) void batch_update_kernel(
    uint32_t n, uint32_t n_pitch, uint32_t next_level, StepCtrl *ctrl, const uint32_t *__restrict__ cur,
    uint32_t *__restrict__ touched, u64 *__restrict__ front, u64 *__restrict__ seen, u64 *__restrict__ next,
    uint8_t *__restrict__ level, uint8_t *__restrict__ flag, const uint32_t *__restrict__ slot_start,
    const uint32_t *__restrict__ deg, u64 *slots, uint32_t *scount)
{
	if (ctrl->done || (ctrl->dir == 0) != kSpan) return;
	__shared__ uint32_t s_words[kSpan ? 2 * kUpdateSpan / 32 : 1];  // touched words, then frontier words
	__shared__ u64 s_or[kBlock / 64][NW];
	__shared__ u64 s_cnt[kBlock / 64][2];
	__shared__ uint32_t s_src[kBlock / 64][NW * 64];  // new vertices per source, per wavefront
	const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	// every block after a sparse step, every kCountSample-th after a dense
	// one, counts the new vertices of each source (scaled up accordingly)
	const bool counting = kSpan || blockIdx.x % kCountSample == 0;
	if (counting) {
#pragma unroll
		for (int j = 0; j < NW; ++j) s_src[wid][j * 64 + lane] = 0;
	}
	u64 wor[NW];
	u64 pairs = 0, pdeg = 0;
	bool over = false;
#pragma unroll
	for (int j = 0; j < NW; ++j) wor[j] = 0;
	// lane l holds the level sources j*64 + l reach at this step, relative to
	// their start levels
	uint32_t rel_lane[NW];
#pragma unroll
	for (int j = 0; j < NW; ++j) rel_lane[j] = next_level - slot_start[j * 64 + lane];
	uint32_t v_lo, v_hi, v_step;
	if (kSpan) {
		// the span's touched words (cleared for the next step as they are
		// read) and frontier words, one load each per thread
		v_lo = blockIdx.x * kUpdateSpan;
		v_hi = min(n, v_lo + kUpdateSpan);
		v_step = kBlock;
		static_assert(2 * kUpdateSpan / 32 == kBlock, "one word per thread");
		const uint32_t i = threadIdx.x, w = v_lo / 32 + (i % (kUpdateSpan / 32));
		uint32_t x = 0;
		if (i < kUpdateSpan / 32) {
			x = touched[w];
			if (x) touched[w] = 0u;
		} else {
			x = cur[w];
		}
		s_words[i] = x;
		__syncthreads();
	} else {
		// blocks walk the vertices kBlock at a time (any grid size works)
		v_lo = blockIdx.x * kBlock;
		v_hi = n;
		v_step = gridDim.x * kBlock;
	}
	for (uint32_t vb = v_lo; vb < v_hi; vb += v_step) {
		const uint32_t v = vb + threadIdx.x;
		const bool in = v < v_hi;
		uint32_t tw, cw;
		if (kSpan) {
			const uint32_t k = (v - v_lo) / 32;
			tw = in ? s_words[k] : 0u;
			cw = in ? s_words[kUpdateSpan / 32 + k] : 0u;
		} else {
			tw = in ? touched[v >> 5] : 0u;
			cw = in ? cur[v >> 5] : 0u;
		}
		const bool t = (tw >> (v & 31)) & 1u, c = (cw >> (v & 31)) & 1u;
		if (!__ballot(t || c)) continue;  // uniform
		if (!kSpan && (v & 31) == 0 && tw) touched[v >> 5] = 0u;  // clear for the next step
		u64 nw[NW];
		u64 anyn = 0;
#pragma unroll
		for (int j = 0; j < NW; ++j) {
			nw[j] = 0;
			const uint64_t idx = (uint64_t)v * NW + j;
			if (t) {
				const u64 nx = next[idx];
				const u64 sn = seen[idx];
				nw[j] = nx & ~sn;
				front[idx] = nw[j];
				if (nx) next[idx] = 0;
				if (nw[j]) seen[idx] = sn | nw[j];
			} else if (c) {
				front[idx] = 0;
			}
			anyn |= nw[j];
		}
		if (anyn) {
			flag[v] = 1;
			uint32_t pc = 0;
#pragma unroll
			for (int j = 0; j < NW; ++j) pc += (uint32_t)__popcll(nw[j]);
			pairs += pc;
			pdeg += (u64)pc * deg[v];
		}
#pragma unroll
		for (int j = 0; j < NW; ++j) {
			u64 u = uniform64(wave_or(nw[j]));
			wor[j] |= u;
			while (u) {
				const uint32_t b = (uint32_t)__ffsll((long long)u) - 1;
				u &= u - 1;
				const uint32_t s = j * 64 + b;
				// source s's level, from lane b's register (no load in the loop)
				const uint32_t rel = (uint32_t)__builtin_amdgcn_readlane((int)rel_lane[j], (int)b);
				over |= rel >= kUnreached;
				const bool mine = (nw[j] >> b) & 1;
				if (mine) level[(uint64_t)s * n_pitch + v] = (uint8_t)rel;
				if (counting) {  // block-uniform; every lane takes part in the ballot
					const uint32_t c = (uint32_t)__popcll(__ballot(mine));
					if (lane == 0) s_src[wid][s] += c;
				}
			}
		}
	}
	pairs = wave_add(pairs);
	pdeg = wave_add(pdeg);
	if (lane == 0) {
#pragma unroll
		for (int j = 0; j < NW; ++j) s_or[wid][j] = wor[j];
		s_cnt[wid][0] = pairs;
		s_cnt[wid][1] = pdeg;
		// a level of 255 or more would read as unreached: the live kernel
		// stops the traversal (no block of this launch may see done change,
		// since every block of it reads done to decide whether to run)
		if (over) ctrl->err = 1;
	}
	__syncthreads();
	if (threadIdx.x < NW) {
		u64 u = 0;
#pragma unroll
		for (uint32_t k = 0; k < kBlock / 64; ++k) u |= s_or[k][threadIdx.x];
		u64 *slot = slots + (uint64_t)(blockIdx.x % kLiveSlots) * kLiveStride + threadIdx.x;
		// read the slot at L2 (not a stale L1 copy) and add only new sources
		if (u) {
			const u64 have = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (u & ~have) atomicOr(slot, u);
		}
	} else if (threadIdx.x == 64) {
		// the totals go to the same spread-out slot lines (words kSlotPairs,
		// kSlotPairs + 1); the live kernel adds them up
		u64 p = 0, d = 0;
#pragma unroll
		for (uint32_t k = 0; k < kBlock / 64; ++k) {
			p += s_cnt[k][0];
			d += s_cnt[k][1];
		}
		if (p) {
			u64 *slot = slots + (uint64_t)(blockIdx.x % kLiveSlots) * kLiveStride + kSlotPairs;
			atomicAdd(slot, p);
			atomicAdd(slot + 1, d);
		}
	}
	if (counting && threadIdx.x < NW * 64) {
		uint32_t c = 0;
#pragma unroll
		for (uint32_t k = 0; k < kBlock / 64; ++k) c += s_src[k][threadIdx.x];
		if (c)
			atomicAdd(scount + (blockIdx.x % kLiveSlots) * kBatchMaxSources + threadIdx.x,
			          kSpan ? c : c * kCountSample);
	}
}

// live = OR of the slots (the sources whose new frontier is non-empty); the
// slots are cleared for the next step.  Then the sparse sources of the next
// frontier (live[kSparseWord + i]): those whose new vertices, times the mean
// out-degree of this step's new (source, vertex) pairs, come to less than
// M / kSparseRatio edges.  A pull finds such a source only after scanning
// about kSparseRatio in-edges per vertex, so on a dense step the heavy pull
// leaves them to a push from their frontier.  One wavefront.
template <int NW>
__global__ void batch_live_kernel(StepCtrl *ctrl, u64 *live, u64 *slots, u64 *sums, uint32_t *scount, u64 m)
{
	if (ctrl->done) return;
	if (ctrl->err) {  // the update met a level past 254: stop here
		if (threadIdx.x == 0) ctrl->done = 1;
		return;
	}
	const uint32_t lane = threadIdx.x;
	u64 lv[NW];
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		u64 *p = slots + (uint64_t)lane * kLiveStride + i;
		lv[i] = wave_or(*p);
		*p = 0;
		if (lane == 0) live[i] = lv[i];
	}
	// the step's new pairs and their out-degrees, into the totals
	u64 *t = slots + (uint64_t)lane * kLiveStride + kSlotPairs;
	const u64 p = wave_add(t[0]), d = wave_add(t[1]);
	t[0] = 0;
	t[1] = 0;
	if (lane == 0) {
		sums[0] += p;
		sums[1] += d;
	}
	// lane l: source i * 64 + l; its count summed over the slots (and cleared)
	const double mean_deg = p ? (double)d / (double)p : 0.0;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		const uint32_t src = i * 64 + lane;
		uint32_t c = 0;
		for (uint32_t k = 0; k < kLiveSlots; ++k) {
			uint32_t *q = scount + k * kBatchMaxSources + src;
			c += *q;
			*q = 0;
		}
		const bool sparse = ((lv[i] >> lane) & 1ull) && (double)c * mean_deg * kSparseRatio < (double)m;
		const u64 sp = __ballot(sparse);
		if (lane == 0) live[kSparseWord + i] = sp;
	}
	if (lane == 0) {
		live[kSparseCount] = 0;
		live[kSparseOverflow] = 0;
	}
}

// inclusive prefix sum over the wavefront
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
	const uint32_t lane = threadIdx.x & 63;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
		if (lane >= (uint32_t)o) x += y;
	}
	return x;
}

// The queue of a dense step's sparse sources: an entry (vertex, chunk) per
// kSparseTile out-edges of each vertex whose front mask holds a sparse
// source, in the queue half of the step's parity (a dense step's frontier
// update leaves that half unused).  Runs after the frontier update has set
// the step's direction; nothing to do unless the step is dense and has
// sparse sources.  A lane per vertex, a block-stride walk; live[kSparseCount]
// counts the entries (cleared by the live kernel), and an entry past the
// half's capacity sets live[kSparseOverflow], which hands the sparse sources
// back to the pulls.
template <int NW>
__global__ __launch_bounds__(kBlock) void batch_sparse_queue_kernel(uint32_t n, uint32_t n_pad,
                                                                    const StepCtrl *__restrict__ ctrl, u64 *live,
                                                                    const u64 *__restrict__ front,
                                                                    const uint32_t *__restrict__ row_off,
                                                                    QueueBufs qb)
{
	if (ctrl->done || ctrl->dir != 1) return;
	u64 sp[NW], any = 0;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		sp[i] = live[kSparseWord + i];
		any |= sp[i];
	}
	if (!any) return;
	const uint32_t par = ctrl->level & 1u;
	uint32_t *__restrict__ queue = qb.queue + par * qb.qstride;
	uint32_t *__restrict__ qchunk = qb.qoff + par * qb.qstride;
	const uint32_t lane = threadIdx.x & 63;
	for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n_pad; v += gridDim.x * kBlock) {
		bool hit = false;
		if (v < n) {
#pragma unroll
			for (int i = 0; i < NW; ++i) hit |= (front[(uint64_t)v * NW + i] & sp[i]) != 0;
		}
		const uint32_t nc = hit ? (row_off[v + 1] - row_off[v] + kSparseTile - 1) / kSparseTile : 0u;
		const uint32_t incl = wave_incl_scan(nc);
		const uint32_t tot = (uint32_t)__shfl((int)incl, 63, 64);
		if (!tot) continue;  // uniform
		u64 base = 0;
		if (lane == 0) base = atomicAdd(live + kSparseCount, (u64)tot);
		base = (u64)__shfl((long long)base, 0, 64);
		if (base + tot > qb.qstride) {
			if (lane == 0) live[kSparseOverflow] = 1;
			continue;
		}
		for (uint32_t c = 0; c < nc; ++c) {
			const uint32_t i = (uint32_t)base + incl - nc + c;
			queue[i] = v;
			qchunk[i] = c;
		}
	}
}

// The sparse sources of a dense step, pushed along the out-edges of their
// frontier (the queue above): next[t] |= (front[h] & sparse) & ~seen[t].  A
// wavefront per entry, after the pulls (the light pull stores next and the
// touched words outright).
template <int NW>
__global__ __launch_bounds__(kBlock) void batch_sparse_src_push_kernel(
    const StepCtrl *__restrict__ ctrl, const u64 *__restrict__ live, QueueBufs qb,
    const uint32_t *__restrict__ row_off, const uint32_t *__restrict__ cols, const u64 *__restrict__ front,
    const u64 *__restrict__ seen, u64 *next, uint32_t *touched)
{
	if (ctrl->done || ctrl->dir != 1 || live[kSparseOverflow]) return;
	Mask<NW> sp;
	u64 any = 0;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		sp.w[i] = live[kSparseWord + i];
		any |= sp.w[i];
	}
	if (!any) return;
	const uint32_t cnt = (uint32_t)live[kSparseCount];
	const uint32_t par = ctrl->level & 1u;
	const uint32_t *__restrict__ queue = qb.queue + par * qb.qstride;
	const uint32_t *__restrict__ qchunk = qb.qoff + par * qb.qstride;
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t nwaves = (gridDim.x * kBlock) >> 6;
	for (uint32_t q = (blockIdx.x * kBlock + threadIdx.x) >> 6; q < cnt; q += nwaves) {
		const uint32_t h = queue[q];
		const uint32_t e0 = row_off[h] + qchunk[q] * kSparseTile;
		const uint32_t e1 = min(e0 + kSparseTile, row_off[h + 1]);
		Mask<NW> f = load_mask<NW>(front, h);
#pragma unroll
		for (int i = 0; i < NW; ++i) f.w[i] &= sp.w[i];
		for (uint32_t e = e0 + lane; e < e1; e += 64) push_mask<NW>(f, cols[e], seen, next, touched);
	}
}

// Sources joining the traversal (at its start, or between two levels): slot
// s gets vertex v at level 0 relative to start_level; v enters the union
// frontier of the next step through its state byte.  One block.
__global__ void batch_seed_kernel(const uint32_t *__restrict__ aux, uint32_t count, uint32_t start_level,
                                  uint32_t nw, uint32_t n_pitch, u64 *front, u64 *seen, uint8_t *level,
                                  uint8_t *flag, u64 *live, uint32_t *slot_start, const uint32_t *__restrict__ deg,
                                  u64 *sums)
{
	for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
		const uint32_t v = aux[kAuxSrc + i], s = aux[kAuxSlot + i];
		const u64 bit = 1ull << (s % 64);
		atomicOr(front + (uint64_t)v * nw + s / 64, bit);
		atomicOr(seen + (uint64_t)v * nw + s / 64, bit);
		atomicOr(live + s / 64, bit);
		level[(uint64_t)s * n_pitch + v] = 0;
		slot_start[s] = start_level;
		flag[v] = 1;
		atomicAdd(sums, 1ull);
		atomicAdd(sums + 1, (u64)deg[v]);
	}
}

// Per-traversal clears of the masks, state bytes, bitmaps, live words, slot
// start levels and totals; the control block is set up for the first
// frontier update (which opens level 0).
__global__ __launch_bounds__(kBlock) void batch_clear_kernel(uint64_t mask_words, uint32_t n_pad, uint32_t n_words,
                                                             u64 *front, u64 *seen, u64 *next, uint8_t *flag,
                                                             uint32_t *cur, uint32_t *touched, u64 *live_all,
                                                             uint32_t live_words, uint32_t *slot_start,
                                                             u64 *sums, uint32_t *scount, StepCtrl *ctrl, bool clean)
{
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	const uint64_t i0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
	// clean: the last traversal on these buffers ran to its end, which leaves
	// front, next and the touched bitmap all zero (only seen keeps its bits)
	for (uint64_t i = i0; i < mask_words; i += stride) {
		seen[i] = 0;
		if (!clean) {
			front[i] = 0;
			next[i] = 0;
		}
	}
	for (uint64_t i = i0; i < n_pad / 16; i += stride) reinterpret_cast<uint4 *>(flag)[i] = make_uint4(0, 0, 0, 0);
	for (uint64_t i = i0; i < n_words; i += stride) {
		cur[i] = 0;
		if (!clean) touched[i] = 0;
	}
	for (uint64_t i = i0; i < live_words; i += stride) live_all[i] = 0;
	for (uint64_t i = i0; i < kBatchMaxSources; i += stride) slot_start[i] = 0;
	for (uint64_t i = i0; i < kLiveSlots * kBatchMaxSources; i += stride) scount[i] = 0;
	if (i0 < 2) sums[i0] = 0;
	if (i0 == 0) {
		ctrl->level = UINT32_MAX;  // the first frontier update opens level 0
		ctrl->done = 0;
		ctrl->err = 0;
		ctrl->dir = 0;
		ctrl->reached = 0;
		ctrl->traversed = 0;
	}
}

// level 255 for every (slot, vertex) pair of the given slots the traversal
// did not reach: the level bytes are written for reached pairs only, so this
// runs only when levels are read out.  One lane per vertex, one coalesced
// store per slot.
template <int NW>
__global__ __launch_bounds__(kBlock) void batch_finalize_kernel(uint32_t n, uint32_t n_pitch, Mask<NW> which,
                                                                const u64 *__restrict__ seen, uint8_t *level)
{
	const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
	if (v >= n) return;
#pragma unroll
	for (int i = 0; i < NW; ++i) {
		const u64 miss = which.w[i] & ~seen[(uint64_t)v * NW + i];
		for (u64 u = which.w[i]; u; u &= u - 1) {
			const uint32_t b = (uint32_t)__ffsll((long long)u) - 1;
			if ((miss >> b) & 1ull) level[(uint64_t)(i * 64 + b) * n_pitch + v] = kUnreached;
		}
	}
}

unsigned grid_n(uint64_t items) { return (unsigned)std::max<uint64_t>(1, (items + kBlock - 1) / kBlock); }

float ev_ms(hipEvent_t a, hipEvent_t b)
{
	float ms = 0.f;
	return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.f;
}

constexpr unsigned kBatchSparseGrid = 2048;
constexpr unsigned kBatchWalkGrid = 8192;  // blocks of the kernels that walk every vertex
constexpr uint32_t kBatchLevelGroup = 8;  // levels enqueued per read-back of the control block

u64 *live_words(gphi_graph *gr) { return gr->b_live; }
u64 *live_slots(gphi_graph *gr) { return gr->b_live + kLiveStride; }
uint32_t *slot_starts(gphi_graph *gr) { return gr->b_aux + kAuxStart; }

// The expansion of one level and its update: the CSR push (sparse step) or
// the pulls (dense step) -- each kernel returns at once unless the control
// block says it is its step -- then the update of the vertices the step
// touched, then the live-source mask.
template <int NW>
void enqueue_expand(gphi_graph *gr)
{
	const DevGraph &G = gr->g;
	hipStream_t s = gr->ctx->stream;
	batch_sparse_kernel<NW><<<kBatchSparseGrid, kBlock, 0, s>>>(gr->ctrl, gr->queue_bufs(), G.row_off, G.cols,
	                                                           gr->b_front, gr->b_seen, gr->b_next, gr->b_touch);
	const unsigned lg = std::min<unsigned>(grid_n(G.n), kBatchWalkGrid);
	const u64 *live = live_words(gr);
	batch_pull_light_kernel<NW, true><<<lg, kBlock, 0, s>>>(G.n, gr->ctrl, gr->in_off, gr->in_src, gr->cur, gr->b_front,
	                                                       gr->b_seen, live, gr->b_next, gr->b_touch);
	batch_pull_light_kernel<NW, false><<<lg, kBlock, 0, s>>>(G.n, gr->ctrl, gr->in_off, gr->in_src, gr->cur, gr->b_front,
	                                                        gr->b_seen, live, gr->b_next, gr->b_touch);
	if (gr->npieces) {
		const unsigned hg = std::min<uint32_t>((gr->npieces + 3) / 4, 16384);
		batch_pull_heavy_kernel<NW, true><<<hg, kBlock, 0, s>>>(G.n, gr->ctrl, gr->npieces, gr->pieces, gr->in_off,
		                                                       gr->in_src, gr->cur, gr->b_front, gr->b_seen, live,
		                                                       gr->b_next, gr->b_touch);
		batch_pull_heavy_kernel<NW, false><<<hg, kBlock, 0, s>>>(G.n, gr->ctrl, gr->npieces, gr->pieces, gr->in_off,
		                                                        gr->in_src, gr->cur, gr->b_front, gr->b_seen, live,
		                                                        gr->b_next, gr->b_touch);
	}
	batch_sparse_src_push_kernel<NW><<<kBatchSparseGrid, kBlock, 0, s>>>(gr->ctrl, live, gr->queue_bufs(), G.row_off,
	                                                                   G.cols, gr->b_front, gr->b_seen, gr->b_next,
	                                                                   gr->b_touch);
}

template <int NW>
void enqueue_update(gphi_graph *gr, uint32_t next_level)
{
	const DevGraph &G = gr->g;
	hipStream_t s = gr->ctx->stream;
	// after a dense step, a block per kBlock vertices; after a sparse one, a
	// block per kUpdateSpan vertices (each launch returns at once on the
	// other's steps)
	batch_update_kernel<NW, false><<<grid_n(G.n), kBlock, 0, s>>>(
	    G.n, G.n_pad, next_level, gr->ctrl, gr->cur, gr->b_touch, gr->b_front, gr->b_seen, gr->b_next, gr->b_level,
	    gr->flag, slot_starts(gr), G.deg, live_slots(gr), gr->b_scount);
	batch_update_kernel<NW, true><<<G.n_pad / kUpdateSpan, kBlock, 0, s>>>(
	    G.n, G.n_pad, next_level, gr->ctrl, gr->cur, gr->b_touch, gr->b_front, gr->b_seen, gr->b_next, gr->b_level,
	    gr->flag, slot_starts(gr), G.deg, live_slots(gr), gr->b_scount);
	batch_live_kernel<NW><<<1, 64, 0, s>>>(gr->ctrl, live_words(gr), live_slots(gr), gr->b_sums, gr->b_scount,
	                                       (u64)G.m);
}

// the sparse-source queue of the next step (after the frontier update has
// set its direction)
template <int NW>
void enqueue_sparse_queue(gphi_graph *gr)
{
	const DevGraph &G = gr->g;
	batch_sparse_queue_kernel<NW><<<std::min<unsigned>(grid_n(G.n_pad), kBatchWalkGrid), kBlock, 0, gr->ctx->stream>>>(
	    G.n, G.n_pad, gr->ctrl, live_words(gr), gr->b_front, G.row_off, gr->queue_bufs());
}

template <int NW>
int launch_finalize(gphi_graph *gr, const u64 *which)
{
	const DevGraph &G = gr->g;
	Mask<NW> m;
	for (int i = 0; i < NW; ++i) m.w[i] = which[i];
	batch_finalize_kernel<NW><<<grid_n(G.n), kBlock, 0, gr->ctx->stream>>>(G.n, G.n_pad, m, gr->b_seen, gr->b_level);
	return hipGetLastError() == hipSuccess ? GPHI_OK : GPHI_E_HIP;
}

int finalize_rows(gphi_graph *gr, uint32_t nw, const u64 *which)
{
	int rc = nw == 1 ? launch_finalize<1>(gr, which) : nw == 2 ? launch_finalize<2>(gr, which) : launch_finalize<4>(gr, which);
	if (rc) return set_error(gr->ctx, rc, "batched BFS: finalize launch failed");
	return GPHI_OK;
}

void enqueue_level(gphi_graph *gr, uint32_t nw, uint32_t next_level, bool expand_half)
{
	if (expand_half) {
		if (nw == 1) enqueue_expand<1>(gr);
		else if (nw == 2) enqueue_expand<2>(gr);
		else enqueue_expand<4>(gr);
	} else {
		if (nw == 1) enqueue_update<1>(gr, next_level);
		else if (nw == 2) enqueue_update<2>(gr, next_level);
		else enqueue_update<4>(gr, next_level);
	}
}

// the frontier update of the union frontier (mark / scan / emit of the
// single-source loop), in its device-driven form: it opens the next level;
// then the next step's sparse-source queue
void enqueue_frontier(gphi_graph *gr, uint32_t nw)
{
	launch_frontier_update_driven(gr, (uint32_t)gr->g.m / gr->ctx->t_ratio);
	if (nw == 1) enqueue_sparse_queue<1>(gr);
	else if (nw == 2) enqueue_sparse_queue<2>(gr);
	else enqueue_sparse_queue<4>(gr);
}

}  // namespace

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

int batch_prepare(gphi_graph *gr, uint32_t nw)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	if (!G.has_csr || !G.has_tiles) return set_error(ctx, GPHI_E_STATE, "graph needs CSR and tile arrays");
	int rc = ensure_traversal_state(gr);
	if (rc) return rc;
	if ((rc = build_in_edges(gr))) return rc;
	if (gr->batch_words >= nw && gr->b_level) return GPHI_OK;
	void *ps[] = {gr->b_front, gr->b_seen, gr->b_next, gr->b_level, gr->b_sums, gr->b_live, gr->b_touch, gr->b_aux};
	for (void *p : ps)
		if (p) (void)hipFree(p);
	gr->b_front = gr->b_seen = gr->b_next = gr->b_sums = gr->b_live = nullptr;
	gr->b_level = nullptr;
	gr->b_touch = gr->b_aux = gr->b_scount = nullptr;
	gr->batch_words = 0;
	const size_t mb = (size_t)G.n_pad * nw * 8;
	hipError_t e = hipMalloc((void **)&gr->b_front, mb);
	if (e == hipSuccess) e = hipMalloc((void **)&gr->b_seen, mb);
	if (e == hipSuccess) e = hipMalloc((void **)&gr->b_next, mb);
	if (e == hipSuccess) e = hipMalloc((void **)&gr->b_level, (size_t)nw * 64 * G.n_pad);
	if (e == hipSuccess) e = hipMalloc((void **)&gr->b_sums, 16);
	if (e == hipSuccess) e = hipMalloc((void **)&gr->b_live, (size_t)(1 + kLiveSlots) * kLiveStride * 8);
	// the touched bitmap, then the per-source counts
	if (e == hipSuccess)
		e = hipMalloc((void **)&gr->b_touch, ((size_t)G.n_words + (size_t)kLiveSlots * kBatchMaxSources) * 4);
	if (e == hipSuccess) e = hipMalloc((void **)&gr->b_aux, (size_t)kAuxWords * 4);
	if (e != hipSuccess)
		return set_error(ctx, GPHI_E_OOM, "batched BFS buffers (%zu MB): %s",
		                 (3 * mb + (size_t)nw * 64 * G.n_pad) >> 20, hipGetErrorString(e));
	gr->b_scount = gr->b_touch + G.n_words;
	gr->batch_words = nw;
	gr->batch_clean_words = 0;
	return GPHI_OK;
}

int BatchRun::begin(gphi_graph *g, uint32_t capacity, const uint32_t *srcs, uint32_t k)
{
	gr = g;
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	if (capacity == 0 || capacity > kBatchMaxSources)
		return set_error(ctx, GPHI_E_ARG, "batch capacity %u not in [1, %u]", capacity, kBatchMaxSources);
	if (k > capacity) return set_error(ctx, GPHI_E_ARG, "%u sources exceed the batch capacity %u", k, capacity);
	if (k && !srcs) return set_error(ctx, GPHI_E_ARG, "sources are NULL");
	for (uint32_t i = 0; i < k; ++i)
		if (srcs[i] >= G.n) return set_error(ctx, GPHI_E_RANGE, "source %u >= N %u", srcs[i], G.n);
	nw = capacity <= 64 ? 1 : capacity <= 128 ? 2 : 4;
	cap = nw * 64;
	int rc = batch_prepare(gr, nw);
	if (rc) return rc;
	used = 0;
	levels = 0;
	done = false;
	slot_src.clear();
	slot_start.clear();
	memset(live, 0, sizeof live);
	memset(pairs, 0, sizeof pairs);
	hipStream_t s = ctx->stream;
	// the buffers hold gr->batch_words >= nw words per vertex; a traversal runs
	// at its own width nw (the masks are addressed with that pitch)
	mask_words = (uint64_t)G.n_pad * nw;
	const uint64_t items = std::max<uint64_t>({mask_words, G.n_pad / 16, G.n_words, kBatchMaxSources});
	// the front and next masks and the touched bitmap are left all zero by a
	// traversal that ran to its end; they need clearing only over words no
	// such traversal has covered since the buffers were made
	prev_clean = gr->batch_clean_words;
	const bool clean = prev_clean >= mask_words;
	gr->batch_clean_words = 0;  // until this traversal ends
	batch_clear_kernel<<<(unsigned)std::min<uint64_t>(grid_n(items), 16384), kBlock, 0, s>>>(
	    mask_words, G.n_pad, G.n_words, gr->b_front, gr->b_seen, gr->b_next, gr->flag, gr->cur, gr->b_touch,
	    gr->b_live, (1 + kLiveSlots) * kLiveStride, slot_starts(gr), gr->b_sums, gr->b_scount, gr->ctrl, clean);
	GPHI_HIP(ctx, hipGetLastError());
	if ((rc = seed(srcs, k, 0))) return rc;
	// the first frontier update: the sources form the union frontier of level 0
	enqueue_frontier(gr, nw);
	GPHI_HIP(ctx, hipGetLastError());
	return GPHI_OK;
}

int BatchRun::seed(const uint32_t *srcs, uint32_t k, uint32_t start_level)
{
	if (k == 0) return GPHI_OK;
	gphi_ctx *ctx = gr->ctx;
	const DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	for (uint32_t i = 0; i < k; ++i) {
		aux[kAuxSrc + i] = srcs[i];
		aux[kAuxSlot + i] = used + i;
	}
	// aux is not written again before the stream is synchronized: every
	// caller of seed synchronizes before its next seed
	GPHI_HIP(ctx, hipMemcpyAsync(gr->b_aux + kAuxSrc, aux + kAuxSrc, (size_t)k * 4, hipMemcpyHostToDevice, s));
	GPHI_HIP(ctx, hipMemcpyAsync(gr->b_aux + kAuxSlot, aux + kAuxSlot, (size_t)k * 4, hipMemcpyHostToDevice, s));
	batch_seed_kernel<<<1, kBlock, 0, s>>>(gr->b_aux, k, start_level, nw, G.n_pad, gr->b_front, gr->b_seen,
	                                        gr->b_level, gr->flag, live_words(gr), slot_starts(gr), G.deg,
	                                        gr->b_sums);
	GPHI_HIP(ctx, hipGetLastError());
	for (uint32_t i = 0; i < k; ++i) {
		slot_src.push_back(srcs[i]);
		slot_start.push_back(start_level);
	}
	used += k;
	return GPHI_OK;
}

// Runs levels on the device until the traversal ends, kBatchLevelGroup levels
// per read-back of the control block (no admissions).
int BatchRun::run_to_end(gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	hipStream_t s = ctx->stream;
	StepCtrl *hc = ctx->h_ctrl;
	hipEvent_t *ev = ctx->ev;
	const bool total_timing = st != nullptr && ctx->ev_ok;
	const bool timing = total_timing && ctx->level_timing;
	constexpr uint32_t kEvEnd = 3 * GPHI_MAX_STEPS + 3;
	if (timing) GPHI_HIP(ctx, hipEventRecord(ev[1], s));
	uint32_t enq = 0;
	for (;;) {
		for (uint32_t q = 0; q < kBatchLevelGroup; ++q, ++enq) {
			const bool ev_step = timing && enq < GPHI_MAX_STEPS - 1;
			enqueue_level(gr, nw, enq + 1, true);
			if (ev_step) GPHI_HIP(ctx, hipEventRecord(ev[2 + 2 * enq], s));
			enqueue_level(gr, nw, enq + 1, false);
			enqueue_frontier(gr, nw);
			if (ev_step) GPHI_HIP(ctx, hipEventRecord(ev[3 + 2 * enq], s));
		}
		GPHI_HIP(ctx, hipGetLastError());
		if (total_timing && !timing) GPHI_HIP(ctx, hipEventRecord(ev[kEvEnd], s));
		GPHI_HIP(ctx, hipMemcpyAsync(hc, gr->ctrl, sizeof(StepCtrl), hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		if (hc->err) return set_error(ctx, GPHI_E_RANGE, "a batched BFS deeper than 254 levels does not fit uint8 levels");
		if (hc->done) break;
		// every level ends the traversal, advances it, or (past level 254) stops it with err
		if (enq > 256 + kBatchLevelGroup)
			return set_error(ctx, GPHI_E_STATE, "batched traversal did not terminate");
	}
	levels = hc->level;
	done = true;
	gr->batch_clean_words = std::max(prev_clean, mask_words);
	if (st) {
		const uint32_t ns = std::min<uint32_t>(levels, GPHI_MAX_STEPS);
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
		}
		if (timing) {
			st->init_ms = ev_ms(ev[0], ev[1]);
			hipEvent_t prev = ev[1];
			const uint32_t nt = std::min<uint32_t>(levels, GPHI_MAX_STEPS - 1);
			for (uint32_t q = 0; q < nt; ++q) {
				st->sweep_ms[q] = ev_ms(prev, ev[2 + 2 * q]);
				st->update_ms[q] = ev_ms(ev[2 + 2 * q], ev[3 + 2 * q]);
				prev = ev[3 + 2 * q];
			}
			st->total_ms = nt ? ev_ms(ev[0], ev[1 + 2 * nt]) : st->init_ms;
		} else {
			st->total_ms = total_timing ? ev_ms(ev[0], ev[kEvEnd]) : 0.f;
		}
	}
	return read_totals();
}

int BatchRun::read_totals()
{
	gphi_ctx *ctx = gr->ctx;
	GPHI_HIP(ctx, hipMemcpyAsync(pairs, gr->b_sums, 16, hipMemcpyDeviceToHost, ctx->stream));
	GPHI_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return GPHI_OK;
}

// One level, then the admission of k new sources (they join the union
// frontier of the next level), then the frontier update; reads the control
// block and the live-source mask back.
int BatchRun::step(const uint32_t *srcs, uint32_t k)
{
	gphi_ctx *ctx = gr->ctx;
	hipStream_t s = ctx->stream;
	StepCtrl *hc = ctx->h_ctrl;
	if (done) return set_error(ctx, GPHI_E_STATE, "the batched traversal has ended");
	if (used + k > cap) return set_error(ctx, GPHI_E_ARG, "no free slot for %u more sources", k);
	const uint32_t next_level = levels + 1;
	enqueue_level(gr, nw, next_level, true);
	enqueue_level(gr, nw, next_level, false);
	int rc = seed(srcs, k, next_level);
	if (rc) return rc;
	enqueue_frontier(gr, nw);
	GPHI_HIP(ctx, hipGetLastError());
	GPHI_HIP(ctx, hipMemcpyAsync(hc, gr->ctrl, sizeof(StepCtrl), hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipMemcpyAsync(live, live_words(gr), (size_t)nw * 8, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	if (hc->err) return set_error(ctx, GPHI_E_RANGE, "a batched BFS deeper than 254 levels does not fit uint8 levels");
	levels = next_level;
	done = hc->done != 0;
	if (done) gr->batch_clean_words = std::max(prev_clean, mask_words);
	return GPHI_OK;
}

// Levels of the given slots (n bytes each, 255 unreached) into host rows.
int BatchRun::copy_rows(const uint32_t *slots, uint32_t ns, uint8_t *const *rows)
{
	gphi_ctx *ctx = gr->ctx;
	const DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	u64 which[4] = {0, 0, 0, 0};
	for (uint32_t i = 0; i < ns; ++i) {
		if (slots[i] >= used) return set_error(ctx, GPHI_E_ARG, "slot %u is not in use", slots[i]);
		which[slots[i] / 64] |= 1ull << (slots[i] % 64);
	}
	int rc = finalize_rows(gr, nw, which);
	if (rc) return rc;
	for (uint32_t i = 0; i < ns; ++i)
		GPHI_HIP(ctx, hipMemcpyAsync(rows[i], gr->b_level + (size_t)slots[i] * G.n_pad, G.n, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	return GPHI_OK;
}

int run_bfs_batch(gphi_graph *gr, const uint32_t *srcs, uint32_t k, uint8_t *levels, gphi_stats *st)
{
	gphi_ctx *ctx = gr->ctx;
	const DevGraph &G = gr->g;
	if (k == 0 || k > kBatchMaxSources)
		return set_error(ctx, GPHI_E_ARG, "batch size %u not in [1, %u]", k, kBatchMaxSources);
	if (!srcs) return set_error(ctx, GPHI_E_ARG, "sources are NULL");
	if (st) memset(st, 0, sizeof(*st));
	const bool total_timing = st != nullptr && ctx->ev_ok;
	if (total_timing) GPHI_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
	BatchRun run;
	int rc = run.begin(gr, k, srcs, k);
	if (rc) return rc;
	if ((rc = run.run_to_end(st))) return rc;
	if (st) {
		st->reached = run.pairs[0];
		st->edges_traversed = run.pairs[1];
	}
	if (levels) {
		hipStream_t s = ctx->stream;
		u64 which[4] = {0, 0, 0, 0};
		for (uint32_t i = 0; i < k; ++i) which[i / 64] |= 1ull << (i % 64);
		if ((rc = finalize_rows(gr, run.nw, which))) return rc;
		GPHI_HIP(ctx, hipMemcpy2DAsync(levels, G.n, gr->b_level, G.n_pad, G.n, k, hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
	}
	return GPHI_OK;
}

}  // namespace gphi

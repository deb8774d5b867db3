// gen.hip -- synthetic inputs and the reference tool pipeline, on the device.
//
// The reference takes its edge lists from external generators (PBBS rMat via
// tools/adj2snap/adj2snap.cpp, Galois graph-convert) and turns them into its
// formats with five CPU tools.  This file builds the same arrays in HBM:
//
//   edge list  --(csr_tiling, tools/csr_tiling/page_rank.cpp:513-570)-->  _untiled
//   _untiled   --(vertex_id_reorder_to_degrees, tools/.../bfs.cpp:50-83)--> _reorder list
//   _reorder   --(csr_tiling)--> CSR  --(coo_tiling, tools/coo_tiling/page_rank.cpp:396-496)
//              --> row-major tiles --(column_major_tile, kcore.cpp:360-417)--> column-stripe tiles
//
// Every tool step is a stable bucketing, done here as a stable LSD radix sort.
// RMAT spec (project-defined; the reference ships no generator): sample i of
// seed s draws, per bit level from the MSB, one 32-bit uniform from a
// splitmix64 stream started at (s << 40) ^ i; u < a -> (0,0), < a+b -> (0,1),
// < a+b+c -> (1,0), else (1,1).  Self-loops, ids >= nverts and duplicates drop.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <climits>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gphi {

__host__ __device__ __forceinline__ uint64_t sm64_next(uint64_t &s)
{
	uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ void rmat_sample(uint64_t seed, uint64_t i, uint32_t scale,
                                                     uint32_t t0, uint32_t t1, uint32_t t2,
                                                     uint32_t &h, uint32_t &t)
{
	uint64_t st = (seed << 40) ^ i;
	uint32_t hh = 0, tt = 0;
	uint64_t r = 0;
	for (uint32_t l = 0; l < scale; ++l) {
		uint32_t u;
		if ((l & 1) == 0) {
			r = sm64_next(st);
			u = (uint32_t)r;
		} else {
			u = (uint32_t)(r >> 32);
		}
		uint32_t bit = 1u << (scale - 1 - l);
		if (u < t0) {
		} else if (u < t1) {
			tt |= bit;
		} else if (u < t2) {
			hh |= bit;
		} else {
			hh |= bit;
			tt |= bit;
		}
	}
	h = hh;
	t = tt;
}

void rmat_thresholds(double a, double b, double c, uint32_t thr[3])
{
	const double two32 = 4294967296.0;
	double ab = a + b, abc = a + b + c;
	thr[0] = (uint32_t)(a * two32);
	thr[1] = ab >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(ab * two32);
	thr[2] = abc >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(abc * two32);
}

namespace {

constexpr unsigned kB = 256;

unsigned grid1(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kB - 1) / kB, 1u << 20)); }

#define GRID_STRIDE(i, n) \
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

__global__ void sample_kernel(uint64_t seed, uint64_t first, uint64_t count, uint32_t scale,
                              uint32_t nverts, uint32_t t0, uint32_t t1, uint32_t t2,
                              uint64_t *keys, uint32_t *idx)
{
	GRID_STRIDE(k, count)
	{
		uint32_t h, t;
		rmat_sample(seed, first + k, scale, t0, t1, t2, h, t);
		keys[k] = (h >= nverts || t >= nverts || h == t) ? UINT64_MAX : ((uint64_t)h << 32) | t;
		if (idx) idx[k] = (uint32_t)(first + k);
	}
}

// 1 at the first element of every run of equal keys (and never for invalid keys)
__global__ void run_heads_kernel(const uint64_t *keys, uint64_t n, uint8_t *flags)
{
	GRID_STRIDE(i, n)
	{
		uint64_t k = keys[i];
		flags[i] = k != UINT64_MAX && (i == 0 || keys[i - 1] != k);
	}
}

__global__ void split_keys_kernel(const uint64_t *keys, uint64_t m, uint32_t *h, uint32_t *t)
{
	GRID_STRIDE(i, m)
	{
		uint64_t k = keys[i];
		h[i] = (uint32_t)(k >> 32);
		t[i] = (uint32_t)k;
	}
}

__global__ void relabel_kernel(uint64_t m, const uint32_t *map, const uint32_t *h, const uint32_t *t,
                               uint32_t *oh, uint32_t *ot)
{
	GRID_STRIDE(i, m)
	{
		oh[i] = map[h[i]];
		ot[i] = map[t[i]];
	}
}

__global__ void iota_kernel(uint64_t m, uint32_t *p)
{
	GRID_STRIDE(i, m) p[i] = (uint32_t)i;
}

// out-degrees of an edge list sorted by head: deg[v] = (first position of
// head v+1) - (first position of head v), two binary searches per vertex (an
// atomic count per edge serializes on the hubs: 0.3 s at scale 25)
__device__ __forceinline__ uint64_t first_at_least(const uint32_t *__restrict__ h, uint64_t m, uint64_t v)
{
	uint64_t a = 0, b = m;
	while (a < b) {
		const uint64_t mid = (a + b) >> 1;
		if (h[mid] < v)
			a = mid + 1;
		else
			b = mid;
	}
	return a;
}

__global__ void degree_kernel(uint32_t n, uint64_t m, const uint32_t *__restrict__ h, uint32_t *__restrict__ deg)
{
	GRID_STRIDE(v, n) deg[v] = (uint32_t)(first_at_least(h, m, v + 1) - first_at_least(h, m, v));
}

__global__ void tile_first_kernel(uint64_t nt, uint64_t m, const uint32_t *__restrict__ key, uint32_t *__restrict__ off)
{
	GRID_STRIDE(p, nt) off[p] = (uint32_t)first_at_least(key, m, p);
}

// key(v) = (unsigned)((float)deg + float(rand())/RAND_MAX); RAND_MAX converts to
// 2^31 so the quotient is an exact scaling (tools/vertex_id_reorder_to_degrees/bfs.cpp:61)
__global__ void degree_key_kernel(uint32_t n, const uint32_t *deg, const int32_t *rnd, uint64_t *out)
{
	GRID_STRIDE(v, n)
	{
		float f = (float)rnd[v] * 4.656612873077392578125e-10f;  // * 2^-31, exact
		float s = (float)deg[v] + f;
		uint32_t key = (uint32_t)s;
		out[v] = ((uint64_t)key << 32) | (uint32_t)v;
	}
}

__global__ void rank_from_sorted_kernel(uint32_t n, const uint64_t *sorted, uint32_t *rank)
{
	GRID_STRIDE(r, n) rank[(uint32_t)sorted[r]] = (uint32_t)r;
}

// One level of the BFS that orders ids for the BFS-order relabelling, over
// the edge list (h, t): every edge out of a level-L vertex gives an unreached
// tail level L+1 (concurrent writers store the same value).
__global__ void bfs_order_level_kernel(uint64_t m, const uint32_t *__restrict__ h, const uint32_t *__restrict__ t,
                                       int32_t *level, int32_t L, uint32_t *changed)
{
	GRID_STRIDE(e, m)
	{
		if (level[h[e]] == L && level[t[e]] < 0) {
			level[t[e]] = L + 1;
			*changed = 1;
		}
	}
}

// Sort key of the BFS-order relabelling: (level, old id), unreached vertices
// after every level, in old-id order.
__global__ void bfs_order_key_kernel(uint32_t n, const int32_t *level, uint64_t *out)
{
	GRID_STRIDE(v, n)
	{
		const uint32_t l = level[v] < 0 ? 0xFFFFFFFFu : (uint32_t)level[v];
		out[v] = ((uint64_t)l << 32) | (uint32_t)v;
	}
}

__global__ void weights_kernel(uint64_t m, uint64_t seed, uint32_t lo, uint64_t span, uint32_t *w)
{
	GRID_STRIDE(e, m)
	{
		uint64_t st = (seed << 40) ^ e ^ 0x00A5A5A5A5ULL;
		w[e] = lo + (uint32_t)(sm64_next(st) % span);
	}
}

// position of tile (h/W, t/W) in the stored tile order
__global__ void tile_key_kernel(uint64_t m, const uint32_t *h, const uint32_t *t, uint32_t W,
                                uint32_t side, uint32_t S, int col_stripe, uint32_t *key)
{
	const uint32_t nfull = side / S, rem = side % S;
	GRID_STRIDE(i, m)
	{
		uint32_t r = h[i] / W, c = t[i] / W;
		uint32_t k;
		if (!col_stripe) {
			k = r * side + c;
		} else {
			uint32_t s = r / S;
			if (s < nfull)
				k = s * S * side + c * S + (r - s * S);
			else
				k = nfull * S * side + c * rem + (r - nfull * S);
		}
		key[i] = k;
	}
}

__global__ void gather_kernel(uint64_t m, const uint32_t *perm, const uint32_t *src, uint32_t *dst)
{
	GRID_STRIDE(i, m) dst[i] = src[perm[i]];
}

struct Scratch {
	gphi_ctx *ctx;
	std::vector<void *> ptrs;
	~Scratch()
	{
		for (void *p : ptrs) (void)hipFree(p);
	}
	template <class T>
	T *alloc(size_t count)
	{
		void *p = nullptr;
		if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
		ptrs.push_back(p);
		return (T *)p;
	}
	void release(void *p)
	{
		auto it = std::find(ptrs.begin(), ptrs.end(), p);
		if (it != ptrs.end()) {
			(void)hipFree(p);
			ptrs.erase(it);
		}
	}
	void keep(void *p)  // ownership moves to the graph
	{
		auto it = std::find(ptrs.begin(), ptrs.end(), p);
		if (it != ptrs.end()) ptrs.erase(it);
	}
};

#define OOM_IF(p) \
	if (!(p)) return set_error(ctx, GPHI_E_OOM, "device allocation failed in graph generation")

// stable sort of m uint32 keys with an index payload -> perm (sorted order of indices)
int stable_order(gphi_ctx *ctx, Scratch &sc, const uint32_t *keys, uint64_t m, int end_bit,
                 uint32_t *perm_out)
{
	hipStream_t s = ctx->stream;
	uint32_t *kout = sc.alloc<uint32_t>(m), *idx = sc.alloc<uint32_t>(m);
	OOM_IF(kout && idx);
	iota_kernel<<<grid1(m), kB, 0, s>>>(m, idx);
	size_t tb = 0;
	GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, kout, idx, perm_out, m, 0,
	                                                 end_bit, s));
	void *tmp = sc.alloc<uint8_t>(tb);
	OOM_IF(tmp);
	GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, kout, idx, perm_out, m, 0,
	                                                 end_bit, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	sc.release(tmp);
	sc.release(kout);
	sc.release(idx);
	return GPHI_OK;
}

int bits_for(uint64_t maxval)
{
	int b = 1;
	while (b < 64 && (maxval >> b)) ++b;
	return b;
}

}  // namespace

// glibc rand() (TYPE_3 additive feedback generator), the same published
// algorithm the oracle restates; the reference tool never seeds it (seed 1).
void glibc_rand_seq(uint32_t seed, uint64_t n, int32_t *out)
{
	int32_t r[34];
	uint32_t ring[34];
	if (seed == 0) seed = 1;
	r[0] = (int32_t)seed;
	for (int i = 1; i < 31; ++i) {
		int32_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
		int32_t word = 16807 * lo - 2836 * hi;
		if (word < 0) word += 2147483647;
		r[i] = word;
	}
	for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
	for (int i = 0; i < 34; ++i) ring[i] = (uint32_t)r[i];
	uint64_t k = 34;
	for (; k < 344; ++k) ring[k % 34] = ring[(k - 31) % 34] + ring[(k - 3) % 34];
	for (uint64_t j = 0; j < n; ++j, ++k) {
		uint32_t v = ring[(k - 31) % 34] + ring[(k - 3) % 34];
		ring[k % 34] = v;
		out[j] = (int32_t)(v >> 1);
	}
}

void random_perm_host(uint32_t n, uint64_t seed, uint32_t *perm)
{
	uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
	for (uint32_t i = 0; i < n; ++i) perm[i] = i;
	for (uint32_t i = n; i > 1; --i) {
		uint32_t j = (uint32_t)(sm64_next(st) % i);
		std::swap(perm[i - 1], perm[j]);
	}
}

// Builds the device graph of a generated RMAT edge list.
int generate_graph(gphi_ctx *ctx, const gphi_rmat_spec *sp, uint32_t W, uint32_t S, int reorder,
                   const gphi_weight_spec *ws, int tile_order, gphi_graph *gr)
{
	hipStream_t s = ctx->stream;
	Scratch sc{ctx, {}};
	const uint32_t n = sp->nverts;
	int rc = GPHI_OK;
	if (n == 0 || sp->scale == 0 || sp->scale > 32 || W == 0 || S == 0)
		return set_error(ctx, GPHI_E_ARG, "bad generator arguments");
	uint32_t thr[3];
	rmat_thresholds(sp->a, sp->b, sp->c, thr);

	// ---- 1. samples -> unique sorted (h,t) keys -------------------------------
	uint64_t m = 0;
	uint64_t *ukeys = nullptr;
	if (sp->target_unique == 0) {
		const uint64_t ns = sp->samples;
		uint64_t *k0 = sc.alloc<uint64_t>(ns), *k1 = sc.alloc<uint64_t>(ns);
		OOM_IF(k0 && k1);
		sample_kernel<<<grid1(ns), kB, 0, s>>>(sp->seed, 0, ns, sp->scale, n, thr[0], thr[1], thr[2], k0, nullptr);
		size_t tb = 0;
		int eb = 32 + bits_for(n);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, tb, k0, k1, ns, 0, 64, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(tmp, tb, k0, k1, ns, 0, 64, s));
		(void)eb;
		sc.release(tmp);
		uint8_t *fl = sc.alloc<uint8_t>(ns);
		uint64_t *cnt = sc.alloc<uint64_t>(1);
		OOM_IF(fl && cnt);
		run_heads_kernel<<<grid1(ns), kB, 0, s>>>(k1, ns, fl);
		tb = 0;
		GPHI_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, tb, k1, fl, k0, cnt, ns, s));
		tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceSelect::Flagged(tmp, tb, k1, fl, k0, cnt, ns, s));
		GPHI_HIP(ctx, hipMemcpyAsync(&m, cnt, 8, hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		sc.release(tmp);
		sc.release(fl);
		sc.release(k1);
		ukeys = k0;
	} else {
		// first `target_unique` distinct valid samples in sample order
		uint64_t ns = std::min<uint64_t>(sp->samples, std::max<uint64_t>(sp->target_unique * 2, 1024));
		for (;;) {
			uint64_t *k0 = sc.alloc<uint64_t>(ns), *k1 = sc.alloc<uint64_t>(ns);
			uint32_t *i0 = sc.alloc<uint32_t>(ns), *i1 = sc.alloc<uint32_t>(ns);
			uint8_t *fl = sc.alloc<uint8_t>(ns);
			uint64_t *cnt = sc.alloc<uint64_t>(1);
			OOM_IF(k0 && k1 && i0 && i1 && fl && cnt);
			sample_kernel<<<grid1(ns), kB, 0, s>>>(sp->seed, 0, ns, sp->scale, n, thr[0], thr[1], thr[2], k0, i0);
			size_t tb = 0;
			GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, i0, i1, ns, 0, 64, s));
			void *tmp = sc.alloc<uint8_t>(tb);
			OOM_IF(tmp);
			GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, i1, ns, 0, 64, s));
			sc.release(tmp);
			run_heads_kernel<<<grid1(ns), kB, 0, s>>>(k1, ns, fl);
			// keep (first index, key) of each run
			tb = 0;
			GPHI_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, tb, i1, fl, i0, cnt, ns, s));
			tmp = sc.alloc<uint8_t>(tb);
			OOM_IF(tmp);
			GPHI_HIP(ctx, hipcub::DeviceSelect::Flagged(tmp, tb, i1, fl, i0, cnt, ns, s));
			GPHI_HIP(ctx, hipStreamSynchronize(s));
			sc.release(tmp);
			tb = 0;
			GPHI_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, tb, k1, fl, k0, cnt, ns, s));
			tmp = sc.alloc<uint8_t>(tb);
			OOM_IF(tmp);
			GPHI_HIP(ctx, hipcub::DeviceSelect::Flagged(tmp, tb, k1, fl, k0, cnt, ns, s));
			uint64_t u = 0;
			GPHI_HIP(ctx, hipMemcpyAsync(&u, cnt, 8, hipMemcpyDeviceToHost, s));
			GPHI_HIP(ctx, hipStreamSynchronize(s));
			sc.release(tmp);
			if (u >= sp->target_unique || ns >= sp->samples) {
				// order the unique edges by first occurrence, keep the first target
				tb = 0;
				GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, i0, i1, k0, k1, u, 0, 32, s));
				tmp = sc.alloc<uint8_t>(tb);
				OOM_IF(tmp);
				GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, i0, i1, k0, k1, u, 0, 32, s));
				sc.release(tmp);
				m = std::min<uint64_t>(u, sp->target_unique);
				tb = 0;
				GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, tb, k1, k0, m, 0, 64, s));
				tmp = sc.alloc<uint8_t>(tb);
				OOM_IF(tmp);
				GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(tmp, tb, k1, k0, m, 0, 64, s));
				GPHI_HIP(ctx, hipStreamSynchronize(s));
				sc.release(tmp);
				sc.release(k1);
				sc.release(i0);
				sc.release(i1);
				sc.release(fl);
				ukeys = k0;
				break;
			}
			sc.release(k0);
			sc.release(k1);
			sc.release(i0);
			sc.release(i1);
			sc.release(fl);
			sc.release(cnt);
			ns = std::min<uint64_t>(ns * 2, sp->samples);
		}
	}
	if (m >= (1ull << 32)) return set_error(ctx, GPHI_E_RANGE, "too many edges (%llu)", (unsigned long long)m);

	// ---- 2. edge list (sorted by (h,t)), optional id permutation ------------
	uint32_t *h0 = sc.alloc<uint32_t>(m), *t0 = sc.alloc<uint32_t>(m);
	OOM_IF(h0 && t0);
	split_keys_kernel<<<grid1(m), kB, 0, s>>>(ukeys, m, h0, t0);
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	sc.release(ukeys);
	uint32_t *h1 = sc.alloc<uint32_t>(m), *t1 = sc.alloc<uint32_t>(m), *perm = sc.alloc<uint32_t>(m);
	OOM_IF(h1 && t1 && perm);
	const int vbits = bits_for(n);
	if (sp->perm_seed) {
		std::vector<uint32_t> hp(n);
		random_perm_host(n, sp->perm_seed, hp.data());
		uint32_t *dmap = sc.alloc<uint32_t>(n);
		OOM_IF(dmap);
		GPHI_HIP(ctx, hipMemcpy(dmap, hp.data(), (size_t)n * 4, hipMemcpyHostToDevice));
		relabel_kernel<<<grid1(m), kB, 0, s>>>(m, dmap, h0, t0, h1, t1);
		// _untiled: stable bucketing by (new) head
		if ((rc = stable_order(ctx, sc, h1, m, vbits, perm))) return rc;
		gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, h1, h0);
		gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, t1, t0);
		sc.release(dmap);
	}
	// (h0, t0) is now the _untiled order of the edge list

	// ---- 3. relabel (degree order or none) -> the `_reorder` edge list -------
	uint32_t *deg = sc.alloc<uint32_t>(n);
	OOM_IF(deg);
	if (reorder == GPHI_REORDER_DEGREE) {
		degree_kernel<<<grid1(n), kB, 0, s>>>(n, m, h0, deg);  // (h0, t0) is sorted by head here
		std::vector<int32_t> rnd(n);
		glibc_rand_seq(1, n, rnd.data());
		int32_t *drnd = sc.alloc<int32_t>(n);
		uint64_t *vk0 = sc.alloc<uint64_t>(n), *vk1 = sc.alloc<uint64_t>(n);
		uint32_t *rank = sc.alloc<uint32_t>(n);
		OOM_IF(drnd && vk0 && vk1 && rank);
		GPHI_HIP(ctx, hipMemcpyAsync(drnd, rnd.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
		degree_key_kernel<<<grid1(n), kB, 0, s>>>(n, deg, drnd, vk0);
		size_t tb = 0;
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeysDescending(nullptr, tb, vk0, vk1, n, 0, 64, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeysDescending(tmp, tb, vk0, vk1, n, 0, 64, s));
		rank_from_sorted_kernel<<<grid1(n), kB, 0, s>>>(n, vk1, rank);
		relabel_kernel<<<grid1(m), kB, 0, s>>>(m, rank, h0, t0, h1, t1);
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		sc.release(tmp);
		sc.release(drnd);
		sc.release(vk0);
		sc.release(vk1);
		sc.release(rank);
		std::swap(h0, h1);
		std::swap(t0, t1);
	} else if (reorder == GPHI_REORDER_BFS) {
		// tools/vertex_id_remap/bfs.cpp:168-247,465-491: BFS from vertex 0;
		// new ids level by level in ascending old id, unreached vertices last
		int32_t *lev = sc.alloc<int32_t>(n);
		uint32_t *changed = sc.alloc<uint32_t>(1);
		uint64_t *vk0 = sc.alloc<uint64_t>(n), *vk1 = sc.alloc<uint64_t>(n);
		uint32_t *rank = sc.alloc<uint32_t>(n);
		OOM_IF(lev && changed && vk0 && vk1 && rank);
		GPHI_HIP(ctx, hipMemsetAsync(lev, 0xFF, (size_t)n * 4, s));
		GPHI_HIP(ctx, hipMemsetAsync(lev, 0, 4, s));  // level[0] = 0
		for (int32_t L = 0; m > 0; ++L) {
			uint32_t hc = 0;
			GPHI_HIP(ctx, hipMemsetAsync(changed, 0, 4, s));
			bfs_order_level_kernel<<<grid1(m), kB, 0, s>>>(m, h0, t0, lev, L, changed);
			GPHI_HIP(ctx, hipMemcpyAsync(&hc, changed, 4, hipMemcpyDeviceToHost, s));
			GPHI_HIP(ctx, hipStreamSynchronize(s));
			if (!hc) break;
			if ((uint32_t)L >= n) return set_error(ctx, GPHI_E_STATE, "BFS-order relabelling did not terminate");
		}
		bfs_order_key_kernel<<<grid1(n), kB, 0, s>>>(n, lev, vk0);
		size_t tb = 0;
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, tb, vk0, vk1, n, 0, 64, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(tmp, tb, vk0, vk1, n, 0, 64, s));
		rank_from_sorted_kernel<<<grid1(n), kB, 0, s>>>(n, vk1, rank);
		relabel_kernel<<<grid1(m), kB, 0, s>>>(m, rank, h0, t0, h1, t1);
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		sc.release(tmp);
		sc.release(lev);
		sc.release(changed);
		sc.release(vk0);
		sc.release(vk1);
		sc.release(rank);
		std::swap(h0, h1);
		std::swap(t0, t1);
	} else if (reorder != GPHI_REORDER_NONE) {
		return set_error(ctx, GPHI_E_ARG, "unsupported reorder kind %d for generated graphs", reorder);
	}
	// (h0, t0) = the `_reorder` edge list; weights attach to this order
	uint32_t *w0 = nullptr, *w1 = nullptr;
	if (ws && ws->seed) {
		w0 = sc.alloc<uint32_t>(m);
		w1 = sc.alloc<uint32_t>(m);
		OOM_IF(w0 && w1);
		weights_kernel<<<grid1(m), kB, 0, s>>>(m, ws->seed, ws->lo, (uint64_t)ws->hi - ws->lo + 1, w0);
	}

	// ---- 4. CSR (_untiled of the _reorder list) ------------------------------
	if ((rc = stable_order(ctx, sc, h0, m, vbits, perm))) return rc;
	gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, h0, h1);
	gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, t0, t1);
	if (w0) gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, w0, w1);
	// (h1, t1, w1) in CSR order
	degree_kernel<<<grid1(n), kB, 0, s>>>(n, m, h1, deg);

	// ---- 5. tile order (stable by tile position, from CSR order) --------------
	const uint32_t side = n % W ? n / W + 1 : n / W;
	if ((uint64_t)side * side > (1ull << 32))
		return set_error(ctx, GPHI_E_RANGE, "SIDE^2 = %llu tiles exceeds 2^32", (unsigned long long)side * side);
	uint32_t *key = h0;  // reuse: h0/t0 become scratch
	tile_key_kernel<<<grid1(m), kB, 0, s>>>(m, h1, t1, W, side, S, tile_order == GPHI_TILES_COL_STRIPE, key);
	if ((rc = stable_order(ctx, sc, key, m, bits_for((uint64_t)side * side - 1), perm))) return rc;
	// tiled arrays go straight into padded graph buffers
	DevGraph &G = gr->g;
	G.n = n;
	G.m = m;
	G.W = W;
	G.S = S;
	G.side = side;
	G.ntiles = (uint64_t)side * side;
	G.tile_order = tile_order;
	G.n_pad = (uint32_t)(((uint64_t)n + 1 + kPadVerts - 1) / kPadVerts * kPadVerts);
	G.n_words = G.n_pad / 32;
	G.nchunks = (m + kChunk - 1) / kChunk;
	G.m_pad = G.nchunks * kChunk;
	G.has_csr = G.has_tiles = true;
	G.weighted = w0 != nullptr;
	uint64_t *acc = &G.device_bytes;
	auto galloc = [&](void **p, size_t bytes) -> int {
		if (hipMalloc(p, std::max<size_t>(bytes, 4)) != hipSuccess) {
			*p = nullptr;
			return set_error(ctx, GPHI_E_OOM, "hipMalloc(%zu) failed", bytes);
		}
		*acc += bytes;
		return GPHI_OK;
	};
	if ((rc = galloc((void **)&G.heads, G.m_pad * 4))) return rc;
	if ((rc = galloc((void **)&G.tails, G.m_pad * 4))) return rc;
	gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, h1, G.heads);
	gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, t1, G.tails);
	if (w0) {
		if ((rc = galloc((void **)&G.coo_w, G.m_pad * 4))) return rc;
		gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, w1, G.coo_w);
	}
	// CSR arrays: t1 / w1 move into the graph
	G.cols = t1;
	sc.keep(t1);
	if (w0) {
		G.csr_w = w1;
		sc.keep(w1);
	}
	*acc += m * 4 * (w0 ? 2 : 1);
	if ((rc = galloc((void **)&G.deg, (size_t)G.n_pad * 4))) return rc;
	GPHI_HIP(ctx, hipMemsetAsync(G.deg, 0, (size_t)G.n_pad * 4, s));
	GPHI_HIP(ctx, hipMemcpyAsync(G.deg, deg, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
	if ((rc = galloc((void **)&G.row_off, ((size_t)n + 1) * 4))) return rc;
	{
		size_t tb = 0;
		GPHI_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, G.deg, G.row_off, n + 1, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tb, G.deg, G.row_off, n + 1, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
	}
	return finish_tiles(ctx, gr);
}

namespace {
// heads of the CSR order: vertex v at row_off[v] (if deg > 0), then a max-scan
__global__ void row_start_kernel(uint32_t n, const uint32_t *row_off, const uint32_t *deg, uint32_t *h)
{
	GRID_STRIDE(v, n)
	if (deg[v]) h[row_off[v]] = (uint32_t)v;
}

struct MaxOp {
	__device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};
}  // namespace

// Re-tiles a graph from its CSR (the coo_tiling / column_major_tile steps
// again with another W, S or order); traversal state is dropped and rebuilt
// on the next call.
int retile_graph(gphi_graph *gr, uint32_t W, uint32_t S, int tile_order)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	if (!G.has_csr) return set_error(ctx, GPHI_E_STATE, "re-tiling needs the CSR arrays");
	if (W == 0 || S == 0) return set_error(ctx, GPHI_E_ARG, "tile width and stripe length must be > 0");
	if (tile_order != GPHI_TILES_COL_STRIPE && tile_order != GPHI_TILES_ROW_MAJOR)
		return set_error(ctx, GPHI_E_ARG, "unknown tile order %d", tile_order);
	const uint32_t n = G.n;
	const uint64_t m = G.m;
	const uint32_t side = n % W ? n / W + 1 : n / W;
	if ((uint64_t)side * side > (1ull << 32))
		return set_error(ctx, GPHI_E_RANGE, "SIDE^2 = %llu tiles exceeds 2^32", (unsigned long long)side * side);
	Scratch sc{ctx, {}};
	uint32_t *h = sc.alloc<uint32_t>(m), *key = sc.alloc<uint32_t>(m), *perm = sc.alloc<uint32_t>(m);
	OOM_IF(h && key && perm);
	int rc;
	if (m) {
		GPHI_HIP(ctx, hipMemsetAsync(h, 0, m * 4, s));
		row_start_kernel<<<grid1(n), kB, 0, s>>>(n, G.row_off, G.deg, h);
		size_t tb = 0;
		GPHI_HIP(ctx, hipcub::DeviceScan::InclusiveScan(nullptr, tb, h, h, MaxOp(), m, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceScan::InclusiveScan(tmp, tb, h, h, MaxOp(), m, s));
		tile_key_kernel<<<grid1(m), kB, 0, s>>>(m, h, G.cols, W, side, S, tile_order == GPHI_TILES_COL_STRIPE, key);
		if ((rc = stable_order(ctx, sc, key, m, bits_for((uint64_t)side * side - 1), perm))) return rc;
	}
	uint32_t *nh = sc.alloc<uint32_t>(G.m_pad), *nt = sc.alloc<uint32_t>(G.m_pad);
	uint32_t *nw = G.csr_w ? sc.alloc<uint32_t>(G.m_pad) : nullptr;
	OOM_IF(nh && nt && (nw || !G.csr_w));
	if (m) {
		gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, h, nh);
		gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, G.cols, nt);
		if (nw) gather_kernel<<<grid1(m), kB, 0, s>>>(m, perm, G.csr_w, nw);
	}
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	free_traversal_state(gr);
	const uint64_t arr = G.m_pad * 4;
	if (G.heads) G.device_bytes -= arr, (void)hipFree(G.heads);
	if (G.tails) G.device_bytes -= arr, (void)hipFree(G.tails);
	if (G.coo_w) G.device_bytes -= arr, (void)hipFree(G.coo_w);
	free_chunk_summaries(G);
	G.device_bytes += arr * (nw ? 3 : 2);
	G.heads = nh;
	G.tails = nt;
	G.coo_w = nw;
	sc.keep(nh);
	sc.keep(nt);
	if (nw) sc.keep(nw);
	G.W = W;
	G.S = S;
	G.side = side;
	G.ntiles = (uint64_t)side * side;
	G.tile_order = tile_order;
	G.has_tiles = true;
	const int frc = finish_tiles(ctx, gr);
	if (frc) {
		// the chunk summaries are incomplete: no sweep may run on these tiles
		free_chunk_summaries(G);
		G.has_tiles = false;
	}
	return frc;
}

namespace {
// off[v] = first position e of the tail-sorted keys with keys[e] >= v
__global__ void lower_bound_kernel(uint32_t n, const uint32_t *__restrict__ keys, uint64_t m,
                                   uint32_t *__restrict__ off)
{
	GRID_STRIDE(v, (uint64_t)n + 1)
	{
		uint64_t a = 0, b = m;
		while (a < b) {
			const uint64_t mid = (a + b) >> 1;
			if (keys[mid] < v)
				a = mid + 1;
			else
				b = mid;
		}
		off[v] = (uint32_t)a;
	}
}

__global__ void piece_count_kernel(uint32_t n, const uint32_t *__restrict__ off, uint32_t *__restrict__ pc)
{
	GRID_STRIDE(v, (uint64_t)n + 1)
	{
		const uint32_t d = v < n ? off[v + 1] - off[v] : 0;
		pc[v] = d > kPullLight ? (d + kPullPiece - 1) / kPullPiece : 0;
	}
}

__global__ void piece_emit_kernel(uint32_t n, const uint32_t *__restrict__ off, const uint32_t *__restrict__ pc,
                                  const uint32_t *__restrict__ base, uint2 *__restrict__ pieces)
{
	GRID_STRIDE(v, n)
	{
		for (uint32_t j = 0; j < pc[v]; ++j) pieces[base[v] + j] = make_uint2((uint32_t)v, off[v] + j * kPullPiece);
	}
}
}  // namespace

// In-edge arrays for the pull sweeps (multi-source BFS, PageRank): in_src
// holds the heads of every edge grouped by tail
// (ascending heads within a tail: a stable sort of the CSR order by tail, or
// of the tiles by head then tail for a graph without CSR), in_off[v] the first
// in-edge of v, and pieces the (vertex, first edge) of every kPullPiece-edge
// piece of a vertex with more than kPullLight in-edges.
int build_in_edges(gphi_graph *gr)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	if (gr->in_off) return GPHI_OK;
	if (!G.has_csr && !G.has_tiles) return set_error(ctx, GPHI_E_STATE, "graph has no edge arrays");
	const uint32_t n = G.n;
	const uint64_t m = G.m;
	Scratch sc{ctx, {}};
	uint32_t *off = sc.alloc<uint32_t>((uint64_t)n + 1), *src = sc.alloc<uint32_t>(m);
	uint32_t *h = sc.alloc<uint32_t>(m), *kout = sc.alloc<uint32_t>(m);
	uint32_t *pc = sc.alloc<uint32_t>((uint64_t)n + 1), *base = sc.alloc<uint32_t>((uint64_t)n + 1);
	OOM_IF(off && src && h && kout && pc && base);
	if (m && !G.has_csr) {
		// only the tile arrays (the row-major files PageRank and hop relaxation
		// read): sorted by head, then stably by tail, so heads ascend within a
		// tail whatever the order inside the tiles
		const int eb = bits_for(n ? n - 1 : 0);
		uint32_t *ts = sc.alloc<uint32_t>(m);
		OOM_IF(ts);
		size_t tb = 0;
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, G.heads, h, G.tails, ts, m, 0, eb, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, G.heads, h, G.tails, ts, m, 0, eb, s));
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, ts, kout, h, src, m, 0, eb, s));
		sc.release(tmp);
		sc.release(ts);
	} else if (m) {
		GPHI_HIP(ctx, hipMemsetAsync(h, 0, m * 4, s));
		row_start_kernel<<<grid1(n), kB, 0, s>>>(n, G.row_off, G.deg, h);
		size_t tb = 0;
		GPHI_HIP(ctx, hipcub::DeviceScan::InclusiveScan(nullptr, tb, h, h, MaxOp(), m, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceScan::InclusiveScan(tmp, tb, h, h, MaxOp(), m, s));
		sc.release(tmp);
		tb = 0;
		const int eb = bits_for(n ? n - 1 : 0);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, G.cols, kout, h, src, m, 0, eb, s));
		tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, G.cols, kout, h, src, m, 0, eb, s));
		sc.release(tmp);
	}
	lower_bound_kernel<<<grid1((uint64_t)n + 1), kB, 0, s>>>(n, kout, m, off);
	piece_count_kernel<<<grid1((uint64_t)n + 1), kB, 0, s>>>(n, off, pc);
	size_t tb = 0;
	GPHI_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, pc, base, (uint64_t)n + 1, s));
	void *tmp = sc.alloc<uint8_t>(tb);
	OOM_IF(tmp);
	GPHI_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tb, pc, base, (uint64_t)n + 1, s));
	uint32_t npieces = 0;
	GPHI_HIP(ctx, hipMemcpyAsync(&npieces, base + n, 4, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	uint2 *pieces = sc.alloc<uint2>((uint64_t)npieces + 1);
	OOM_IF(pieces);
	if (npieces) piece_emit_kernel<<<grid1(n), kB, 0, s>>>(n, off, pc, base, pieces);
	GPHI_HIP(ctx, hipGetLastError());
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	sc.keep(off);
	sc.keep(src);
	sc.keep(pieces);
	gr->in_off = off;
	gr->in_src = src;
	gr->pieces = pieces;
	gr->npieces = npieces;
	gr->in_bytes = ((uint64_t)n + 1) * 4 + m * 4 + ((uint64_t)npieces + 1) * 8;
	return GPHI_OK;
}

namespace {
__global__ void narrow_w8_kernel(uint64_t m, const uint32_t *__restrict__ w, uint8_t *__restrict__ w8)
{
	GRID_STRIDE(e, m) w8[e] = (uint8_t)w[e];
}
}  // namespace

// The weights of the in-edges, in in_src's order: the same stable sort of
// the CSR by tail that build_in_edges ran, with the CSR weights as payload
// (equal keys keep their CSR order in both sorts, so the two payloads line
// up).  One byte per edge when the tile weights have a 1-byte copy (every
// weight below 256).
int build_in_weights(gphi_graph *gr)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	if (gr->in_w8 || gr->in_w32) return GPHI_OK;
	if (!gr->in_off) return set_error(ctx, GPHI_E_STATE, "in-edge weights need the in-edge arrays");
	if (!G.has_csr || !G.csr_w) return set_error(ctx, GPHI_E_STATE, "in-edge weights need CSR weights");
	const uint64_t m = G.m;
	Scratch sc{ctx, {}};
	uint32_t *kout = sc.alloc<uint32_t>(m), *w = sc.alloc<uint32_t>(m);
	OOM_IF(kout && w);
	const int eb = bits_for(G.n ? G.n - 1 : 0);
	size_t tb = 0;
	if (m) {
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, G.cols, kout, G.csr_w, w, m, 0, eb, s));
		void *tmp = sc.alloc<uint8_t>(tb);
		OOM_IF(tmp);
		GPHI_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, G.cols, kout, G.csr_w, w, m, 0, eb, s));
		sc.release(tmp);
	}
	if (G.coo_w8) {
		uint8_t *w8 = sc.alloc<uint8_t>(m);
		OOM_IF(w8);
		if (m) narrow_w8_kernel<<<grid1(m), kB, 0, s>>>(m, w, w8);
		GPHI_HIP(ctx, hipGetLastError());
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		sc.keep(w8);
		gr->in_w8 = w8;
	} else {
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		sc.keep(w);
		gr->in_w32 = w;
	}
	return GPHI_OK;
}

namespace {
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *__restrict__ a, uint32_t lo, uint32_t hi,
                                                    uint32_t x)
{
	while (lo < hi) {
		const uint32_t mid = lo + ((hi - lo) >> 1);
		if (a[mid] < x)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

// t's in-edges with heads in [hlo, hhi): [*lo, *lo + cnt)
__device__ __forceinline__ uint32_t band_range(const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ in_src,
                                               uint32_t t, uint32_t hlo, uint32_t hhi, uint32_t *lo)
{
	const uint32_t a = in_off[t], b = in_off[t + 1];
	*lo = lower_bound_u32(in_src, a, b, hlo);
	return lower_bound_u32(in_src, *lo, b, hhi) - *lo;
}

// the classes of a band's pairs, each counted per tail: small, light and
// heavy pairs, pieces, multi-piece pairs, then the edges of the small, light
// and heavy pairs
constexpr int kBandClasses = 8;
__device__ __forceinline__ void band_classes(uint32_t c, uint32_t (&f)[kBandClasses])
{
	const uint32_t np = c > kPullLight ? (c + kPullPiece - 1) / kPullPiece : 0;
	f[0] = c >= 1 && c <= kPrSmall;
	f[1] = c > kPrSmall && c <= kPullLight;
	f[2] = c > kPullLight;
	f[3] = np;
	f[4] = np > 1;
	f[5] = f[0] ? c : 0;
	f[6] = f[1] ? c : 0;
	f[7] = f[2] ? c : 0;
}

// per tail (and a zero entry at n, so that an exclusive scan of n+1 entries
// ends with the totals): the pair counts of each class, class k at
// f + k * (n + 1)
__global__ void band_count_kernel(uint32_t n, const uint32_t *__restrict__ in_off,
                                  const uint32_t *__restrict__ in_src, uint32_t hlo, uint32_t hhi,
                                  uint32_t *__restrict__ f)
{
	GRID_STRIDE(t, (uint64_t)n + 1)
	{
		uint32_t lo, c = 0, cls[kBandClasses];
		if (t < n) c = band_range(in_off, in_src, (uint32_t)t, hlo, hhi, &lo);
		band_classes(c, cls);
#pragma unroll
		for (int k = 0; k < kBandClasses; ++k) f[(uint64_t)k * (n + 1) + t] = cls[k];
	}
}

__global__ void band_emit_kernel(uint32_t n, const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ in_src,
                                  const uint32_t *__restrict__ pos, PrBand b, uint32_t *__restrict__ src,
                                  uint32_t *__restrict__ lt, uint32_t *__restrict__ llo, uint8_t *__restrict__ lcnt,
                                  uint4 *__restrict__ hp, uint32_t *__restrict__ pc, uint32_t *__restrict__ multi)
{
	GRID_STRIDE(t, n)
	{
		uint32_t lo;
		const uint32_t c = band_range(in_off, in_src, (uint32_t)t, b.head_lo, b.head_hi, &lo);
		const uint64_t stride = (uint64_t)n + 1;
		uint32_t e = 0;  // first edge of the pair in src
		if (c >= 1 && c <= kPullLight) {
			const bool small = c <= kPrSmall;
			const uint32_t i = small ? b.small0 + pos[t] : b.light0 + pos[stride + t];
			e = small ? b.edge_small + pos[5 * stride + t] : b.edge_light + pos[6 * stride + t];
			lt[i] = (uint32_t)t;
			llo[i] = e;
			lcnt[i] = (uint8_t)c;
		} else if (c > kPullLight) {
			const uint32_t i = b.heavy0 + pos[2 * stride + t], p0 = b.piece0 + pos[3 * stride + t];
			e = b.edge_heavy + pos[7 * stride + t];
			hp[i] = make_uint4((uint32_t)t, e, c, p0);
			const uint32_t np = (c + kPullPiece - 1) / kPullPiece;
			for (uint32_t k = 0; k < np; ++k) pc[p0 + k] = i;
			if (np > 1) multi[b.multi0 + pos[4 * stride + t]] = i;
		}
		for (uint32_t k = 0; k < c; ++k) src[e + k] = in_src[lo + k];
	}
}
}  // namespace

// PageRank head bands.  The sum over in-edges gathers contrib[h] at random;
// swept band by band, every gather of a band falls in one window of
// contributions that stays in each XCD's L2.  Bands are cut from windows of
// kBandHeads heads: a window holding more than M/64 edges (its out-degree
// sum) is a band of its own, lighter windows merge with the next while the
// band stays under M/64 edges (their few gathers miss either way, and each
// band costs a launch).  Per band, a count pass and scans place the pairs in
// tail order, then an emit pass writes them and copies their heads into the
// band's run of pr_src (a sweep over the tail-major in_src would fetch each
// tail's cache lines once per band); the counts of all bands are taken first
// so that one allocation holds every band.
int build_pr_bands(gphi_graph *gr)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	if (gr->pr_mem || !gr->pr_bands.empty()) return GPHI_OK;
	if (!gr->in_off) return set_error(ctx, GPHI_E_STATE, "PageRank bands need the in-edge arrays");
	const uint32_t n = G.n;
	const uint64_t m = G.m;
	// band cuts from the out-degree sums of the windows
	std::vector<uint32_t> hdeg(n);
	if (n) GPHI_HIP(ctx, hipMemcpyAsync(hdeg.data(), G.deg, (size_t)n * 4, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	const uint32_t nwin = (uint32_t)(((uint64_t)n + kBandHeads - 1) / kBandHeads);
	std::vector<uint64_t> wsum(nwin, 0);
	for (uint32_t v = 0; v < n; ++v) wsum[v / kBandHeads] += hdeg[v];
	std::vector<uint32_t> cuts{0};
	const uint64_t cap = std::max<uint64_t>(m / 64, 1);
	for (uint32_t w = 0; w < nwin;) {
		uint64_t e = wsum[w++];
		while (w < nwin && e + wsum[w] <= cap) e += wsum[w++];
		cuts.push_back((uint32_t)std::min<uint64_t>((uint64_t)w * kBandHeads, n));
	}
	const uint64_t stride = (uint64_t)n + 1;
	Scratch sc{ctx, {}};
	uint32_t *f = sc.alloc<uint32_t>(kBandClasses * stride), *pos = sc.alloc<uint32_t>(kBandClasses * stride);
	OOM_IF(f && pos);
	size_t tb = 0;
	GPHI_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, f, pos, stride, s));
	void *tmp = sc.alloc<uint8_t>(tb);
	OOM_IF(tmp);
	// per-tail class counts of the band [hlo, hhi) and their exclusive scans
	auto count_band = [&](uint32_t hlo, uint32_t hhi) -> int {
		band_count_kernel<<<grid1(stride), kB, 0, s>>>(n, gr->in_off, gr->in_src, hlo, hhi, f);
		for (int k = 0; k < kBandClasses; ++k) {
			size_t t2 = tb;
			GPHI_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, t2, f + k * stride, pos + k * stride, stride, s));
		}
		return GPHI_OK;
	};
	const size_t nb = cuts.size() - 1;
	std::vector<std::array<uint32_t, kBandClasses>> tot(nb);
	uint64_t all[kBandClasses] = {};
	for (size_t b = 0; b < nb; ++b) {
		int rc = count_band(cuts[b], cuts[b + 1]);
		if (rc) return rc;
		for (int k = 0; k < kBandClasses; ++k)
			GPHI_HIP(ctx, hipMemcpyAsync(&tot[b][k], pos + k * stride + n, 4, hipMemcpyDeviceToHost, s));
		GPHI_HIP(ctx, hipStreamSynchronize(s));
		for (int k = 0; k < kBandClasses; ++k) all[k] += tot[b][k];
	}
	const uint64_t npairs = all[0] + all[1];  // small pairs of every band, then light pairs
	if (npairs + all[2] + all[3] >= (1ull << 32)) return set_error(ctx, GPHI_E_RANGE, "PageRank bands: too many pairs");
	std::vector<PrBand> bands(nb);
	uint64_t off[5] = {0, all[0], 0, 0, 0}, e_off = 0;
	for (size_t b = 0; b < nb; ++b) {
		const uint64_t el = e_off + tot[b][5], eh = el + tot[b][6];
		bands[b] = PrBand{cuts[b],         cuts[b + 1], (uint32_t)off[0], tot[b][0], (uint32_t)off[1],
		                  tot[b][1],       (uint32_t)off[2], tot[b][2], (uint32_t)off[3], tot[b][3],
		                  (uint32_t)off[4], tot[b][4], (uint32_t)e_off, (uint32_t)el, (uint32_t)eh};
		for (int k = 0; k < 5; ++k) off[k] += tot[b][k];
		e_off = eh + tot[b][7];
	}
	if (e_off != m)
		return set_error(ctx, GPHI_E_STATE, "PageRank bands cover %llu of %llu in-edges", (unsigned long long)e_off,
		                 (unsigned long long)m);
	// heads band by band, pair tails and firsts, heavy pairs, piece owners,
	// multi-piece pairs, pair counts
	auto up16 = [](uint64_t x) { return (x + 15) & ~15ull; };
	const uint64_t o_lt = up16(m * 4), o_llo = up16(o_lt + npairs * 4), o_hp = up16(o_llo + npairs * 4),
	               o_pc = o_hp + all[2] * 16, o_mu = up16(o_pc + all[3] * 4), o_lc = up16(o_mu + all[4] * 4),
	               bytes = up16(o_lc + npairs);
	uint8_t *mem = sc.alloc<uint8_t>(bytes);
	OOM_IF(mem);
	uint32_t *src = (uint32_t *)mem;
	uint32_t *lt = (uint32_t *)(mem + o_lt), *llo = (uint32_t *)(mem + o_llo), *pc = (uint32_t *)(mem + o_pc);
	uint32_t *multi = (uint32_t *)(mem + o_mu);
	uint4 *hp = (uint4 *)(mem + o_hp);
	uint8_t *lcnt = mem + o_lc;
	for (const PrBand &b : bands) {
		int rc = count_band(b.head_lo, b.head_hi);
		if (rc) return rc;
		band_emit_kernel<<<grid1(n), kB, 0, s>>>(n, gr->in_off, gr->in_src, pos, b, src, lt, llo, lcnt, hp, pc,
		                                         multi);
	}
	GPHI_HIP(ctx, hipGetLastError());
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	sc.keep(mem);
	gr->pr_mem = mem;
	gr->pr_src = src;
	gr->pr_lt = lt;
	gr->pr_llo = llo;
	gr->pr_lcnt = lcnt;
	gr->pr_hp = hp;
	gr->pr_pc = pc;
	gr->pr_multi = multi;
	gr->pr_npieces = (uint32_t)all[3];
	gr->pr_bytes = bytes;
	gr->pr_bands = std::move(bands);
	return GPHI_OK;
}

// -offsets of the stored tile order: histogram of tile positions + scan
int tile_offsets(gphi_graph *gr, uint32_t *out)
{
	gphi_ctx *ctx = gr->ctx;
	DevGraph &G = gr->g;
	hipStream_t s = ctx->stream;
	if (!G.has_tiles) return set_error(ctx, GPHI_E_STATE, "graph has no tile arrays");
	if (G.ntiles > (1ull << 32)) return set_error(ctx, GPHI_E_RANGE, "too many tiles");
	Scratch sc{ctx, {}};
	const uint64_t nt = G.ntiles, m = G.m;
	uint32_t *key = sc.alloc<uint32_t>(m), *off = sc.alloc<uint32_t>(nt);
	OOM_IF(key && off);
	tile_key_kernel<<<grid1(m), kB, 0, s>>>(m, G.heads, G.tails, G.W, G.side, G.S,
	                                       G.tile_order == GPHI_TILES_COL_STRIPE, key);
	// the keys are the tiles' positions in the stored order, so they do not
	// decrease along the edges: a tile's offset is the first edge at or past it
	tile_first_kernel<<<grid1(nt), kB, 0, s>>>(nt, m, key, off);
	GPHI_HIP(ctx, hipMemcpyAsync(out, off, nt * 4, hipMemcpyDeviceToHost, s));
	GPHI_HIP(ctx, hipStreamSynchronize(s));
	return GPHI_OK;
}

}  // namespace gphi

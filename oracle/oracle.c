/*
 * oracle.c -- CPU restatement of GraphPhi's tiled graph-traversal path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by smoke() and by
 * the cpu_baseline leg of bench.py; never by the product.  Reference citations
 * are file:line relative to the reference repository root.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <immintrin.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Synthetic input spec (the reference has no generator: its inputs came from
 * PBBS rMat / Galois, tools/adj2snap/adj2snap.cpp:26).  Parity unpinned.    */
/* ------------------------------------------------------------------------ */

static inline uint64_t sm64_next(uint64_t *s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

void orc_rmat_thresholds(double a, double b, double c, uint32_t thr[3])
{
	const double two32 = 4294967296.0;
	double ab = a + b, abc = a + b + c;
	thr[0] = (uint32_t)(a * two32);
	thr[1] = ab >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(ab * two32);
	thr[2] = abc >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(abc * two32);
}

/* One RMAT sample: per bit level (MSB first) one 32-bit uniform u picks the
 * quadrant: u<thr0 -> (0,0), u<thr1 -> (0,1), u<thr2 -> (1,0), else (1,1). */
void orc_rmat_sample(uint64_t seed, uint64_t i, uint32_t scale, const uint32_t thr[3],
                     uint32_t *h, uint32_t *t)
{
	uint64_t st = (seed << 40) ^ i;
	uint32_t hh = 0, tt = 0;
	uint64_t r = 0;
	for (uint32_t l = 0; l < scale; ++l) {
		uint32_t u;
		if ((l & 1) == 0) {
			r = sm64_next(&st);
			u = (uint32_t)r;
		} else {
			u = (uint32_t)(r >> 32);
		}
		uint32_t bit = 1u << (scale - 1 - l);
		if (u < thr[0]) {
		} else if (u < thr[1]) {
			tt |= bit;
		} else if (u < thr[2]) {
			hh |= bit;
		} else {
			hh |= bit;
			tt |= bit;
		}
	}
	*h = hh;
	*t = tt;
}

static int cmp_u64(const void *a, const void *b)
{
	uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return x < y ? -1 : (x > y);
}

uint64_t orc_rmat_edges(uint32_t scale, uint32_t nverts, uint64_t seed, uint64_t nsamples,
                        uint64_t target_unique, double a, double b, double c,
                        uint32_t *h_out, uint32_t *t_out, uint64_t cap)
{
	uint32_t thr[3];
	orc_rmat_thresholds(a, b, c, thr);
	uint64_t *keys;
	uint64_t nk = 0;
	if (target_unique == 0) {
		keys = (uint64_t *)malloc((nsamples ? nsamples : 1) * sizeof(uint64_t));
		if (!keys) return 0;
		for (uint64_t i = 0; i < nsamples; ++i) {
			uint32_t h, t;
			orc_rmat_sample(seed, i, scale, thr, &h, &t);
			if (h >= nverts || t >= nverts || h == t) continue;
			keys[nk++] = ((uint64_t)h << 32) | t;
		}
		qsort(keys, nk, sizeof(uint64_t), cmp_u64);
		uint64_t u = 0;
		for (uint64_t i = 0; i < nk; ++i)
			if (u == 0 || keys[i] != keys[u - 1]) keys[u++] = keys[i];
		nk = u;
	} else {
		/* first-occurrence dedup in sample order, open addressing */
		uint64_t hs = 1;
		while (hs < 2 * target_unique) hs <<= 1;
		uint64_t *tab = (uint64_t *)malloc(hs * sizeof(uint64_t));
		keys = (uint64_t *)malloc(target_unique * sizeof(uint64_t));
		if (!tab || !keys) {
			free(tab);
			free(keys);
			return 0;
		}
		memset(tab, 0xFF, hs * sizeof(uint64_t));
		for (uint64_t i = 0; i < nsamples && nk < target_unique; ++i) {
			uint32_t h, t;
			orc_rmat_sample(seed, i, scale, thr, &h, &t);
			if (h >= nverts || t >= nverts || h == t) continue;
			uint64_t k = ((uint64_t)h << 32) | t;
			uint64_t x = k * 0x9E3779B97F4A7C15ULL;
			uint64_t p = (x ^ (x >> 29)) & (hs - 1);
			int dup = 0;
			while (tab[p] != UINT64_MAX) {
				if (tab[p] == k) {
					dup = 1;
					break;
				}
				p = (p + 1) & (hs - 1);
			}
			if (dup) continue;
			tab[p] = k;
			keys[nk++] = k;
		}
		free(tab);
		qsort(keys, nk, sizeof(uint64_t), cmp_u64);
	}
	if (nk > cap) nk = cap;
	for (uint64_t i = 0; i < nk; ++i) {
		h_out[i] = (uint32_t)(keys[i] >> 32);
		t_out[i] = (uint32_t)keys[i];
	}
	free(keys);
	return nk;
}

void orc_random_perm(uint32_t n, uint64_t seed, uint32_t *perm)
{
	uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
	for (uint32_t i = 0; i < n; ++i) perm[i] = i;
	for (uint32_t i = n; i > 1; --i) {
		uint32_t j = (uint32_t)(sm64_next(&st) % i);
		uint32_t tmp = perm[i - 1];
		perm[i - 1] = perm[j];
		perm[j] = tmp;
	}
}

void orc_weights(uint64_t m, uint64_t seed, uint32_t lo, uint32_t hi, uint32_t *w)
{
	uint64_t span = (uint64_t)hi - lo + 1;
	for (uint64_t e = 0; e < m; ++e) {
		uint64_t st = (seed << 40) ^ e ^ 0x00A5A5A5A5ULL;
		w[e] = lo + (uint32_t)(sm64_next(&st) % span);
	}
}

void orc_relabel(uint64_t m, const uint32_t *map, uint32_t *h, uint32_t *t)
{
	for (uint64_t e = 0; e < m; ++e) {
		h[e] = map[h[e]];
		t[e] = map[t[e]];
	}
}

/* ------------------------------------------------------------------------ */
/* Reorderings                                                              */
/* ------------------------------------------------------------------------ */

/* glibc rand() (TYPE_3 additive feedback generator, r[i] = r[i-3] + r[i-31],
 * seeded through the 16807 LCG and discarding the first 310 outputs) -- the
 * published description of glibc's random_r.  srand is never called by the
 * reference tool, so the seed is 1 (tools/vertex_id_reorder_to_degrees/bfs.cpp:61). */
void orc_glibc_rand(uint32_t seed, uint64_t n, int32_t *out)
{
	int32_t r[34];
	uint32_t ring[34];
	if (seed == 0) seed = 1;
	r[0] = (int32_t)seed;
	for (int i = 1; i < 31; ++i) {
		int32_t hi = r[i - 1] / 127773;
		int32_t lo = r[i - 1] % 127773;
		int32_t word = 16807 * lo - 2836 * hi;
		if (word < 0) word += 2147483647;
		r[i] = word;
	}
	for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
	for (int i = 0; i < 34; ++i) ring[i] = (uint32_t)r[i];
	/* ring holds r[k-34 .. k-1] at index k % 34 */
	uint64_t k = 34;
	for (; k < 344; ++k) ring[k % 34] = ring[(k - 31) % 34] + ring[(k - 3) % 34];
	for (uint64_t j = 0; j < n; ++j, ++k) {
		uint32_t v = ring[(k - 31) % 34] + ring[(k - 3) % 34];
		ring[k % 34] = v;
		out[j] = (int32_t)(v >> 1);
	}
}

typedef struct {
	uint32_t key;
	uint32_t v;
} kv_t;

static int cmp_kv_desc(const void *a, const void *b)
{
	const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
	if (x->key != y->key) return x->key > y->key ? -1 : 1;
	if (x->v != y->v) return x->v > y->v ? -1 : 1;
	return 0;
}

/* tools/vertex_id_reorder_to_degrees/bfs.cpp:50-67: key = (unsigned)(deg +
 * float(rand())/RAND_MAX) built into pair<unsigned,unsigned>, sorted in
 * descending (key, id) order, new id = rank. */
void orc_degree_rank(uint32_t n, const uint32_t *deg, uint32_t *rank)
{
	int32_t *rv = (int32_t *)malloc((size_t)(n ? n : 1) * sizeof(int32_t));
	kv_t *kv = (kv_t *)malloc((size_t)(n ? n : 1) * sizeof(kv_t));
	orc_glibc_rand(1, n, rv);
	const int rand_max = 2147483647;
	for (uint32_t v = 0; v < n; ++v) {
		volatile float f = (float)rv[v] / (float)rand_max;
		volatile float s = (float)deg[v] + f;
		kv[v].key = (uint32_t)s;
		kv[v].v = v;
	}
	qsort(kv, n, sizeof(kv_t), cmp_kv_desc);
	for (uint32_t r = 0; r < n; ++r) rank[kv[r].v] = r;
	free(rv);
	free(kv);
}

/* tools/vertex_id_remap/bfs.cpp:230-247,465-491,493-561: level-synchronous BFS
 * from vertex 0; new ids handed out level by level in ascending old id
 * (add_mask2map), unreached vertices appended in old-id order. */
void orc_bfs_order_rank(uint32_t n, const uint64_t *row_off, const uint32_t *cols,
                        uint32_t *rank)
{
	if (n == 0) return;
	int32_t *lev = (int32_t *)malloc((size_t)n * sizeof(int32_t));
	orc_bfs_levels(n, row_off, cols, 0, lev);
	int32_t maxl = 0;
	for (uint32_t v = 0; v < n; ++v)
		if (lev[v] > maxl) maxl = lev[v];
	uint32_t top = 0;
	for (int32_t l = 0; l <= maxl; ++l)
		for (uint32_t v = 0; v < n; ++v)
			if (lev[v] == l) rank[v] = top++;
	for (uint32_t v = 0; v < n; ++v)
		if (lev[v] < 0) rank[v] = top++;
	free(lev);
}

/* ------------------------------------------------------------------------ */
/* Layouts                                                                  */
/* ------------------------------------------------------------------------ */

/* tools/csr_tiling/page_rank.cpp:513-570 (weighted :414-470): bucket edges by
 * head keeping input order within a head; nneibor = out-degree. */
void orc_untiled(uint32_t n, uint64_t m, const uint32_t *h, const uint32_t *t, const uint32_t *w,
                 uint32_t *oh, uint32_t *ot, uint32_t *ow, uint32_t *deg)
{
	uint64_t *pos = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
	memset(deg, 0, (size_t)n * sizeof(uint32_t));
	for (uint64_t e = 0; e < m; ++e) deg[h[e]]++;
	for (uint32_t v = 0; v < n; ++v) pos[v + 1] = pos[v] + deg[v];
	for (uint64_t e = 0; e < m; ++e) {
		uint64_t p = pos[h[e]]++;
		oh[p] = h[e];
		ot[p] = t[e];
		if (w && ow) ow[p] = w[e];
	}
	free(pos);
}

/* SIDE_LENGTH = ceil(N / W) (apps/bfs_simd/bfs.cpp:995-999) */
uint32_t orc_side(uint32_t n, uint32_t W) { return n % W ? n / W + 1 : n / W; }

/* tools/coo_tiling/page_rank.cpp:396-496: tile = (h/W)*SIDE + t/W, stable in
 * _untiled order; offsets are exclusive prefix sums over SIDE^2 tiles. */
void orc_coo_tile(uint32_t n, uint64_t m, uint32_t W, const uint32_t *h, const uint32_t *t,
                  const uint32_t *w, uint32_t *oh, uint32_t *ot, uint32_t *ow, uint32_t *offsets)
{
	uint64_t side = orc_side(n, W), nt = side * side;
	uint64_t *cnt = (uint64_t *)calloc((size_t)nt + 1, sizeof(uint64_t));
	for (uint64_t e = 0; e < m; ++e) cnt[(uint64_t)(h[e] / W) * side + t[e] / W + 1]++;
	for (uint64_t i = 0; i < nt; ++i) cnt[i + 1] += cnt[i];
	for (uint64_t i = 0; i < nt; ++i) offsets[i] = (uint32_t)cnt[i];
	for (uint64_t e = 0; e < m; ++e) {
		uint64_t p = cnt[(uint64_t)(h[e] / W) * side + t[e] / W]++;
		oh[p] = h[e];
		ot[p] = t[e];
		if (w && ow) ow[p] = w[e];
	}
	free(cnt);
}

/* tools/column_major_tile/kcore.cpp:360-417 (weighted :246-303): re-emit tiles
 * stripe by stripe (S rows), column by column, row by row; a last partial
 * stripe takes the remaining rows. */
void orc_col_major(uint32_t n, uint64_t m, uint32_t W, uint32_t S, const uint32_t *h,
                   const uint32_t *t, const uint32_t *w, const uint32_t *offsets,
                   uint32_t *oh, uint32_t *ot, uint32_t *ow, uint32_t *new_offsets)
{
	uint64_t side = orc_side(n, W), nt = side * side;
	uint64_t ei = 0, nti = 0, side_i = 0;
	while (side_i < side) {
		/* full stripes first, then the remainder rows as one last stripe */
		uint64_t rend = side_i + S <= side ? side_i + S : side;
		for (uint64_t col = 0; col < side; ++col) {
			for (uint64_t row = side_i; row < rend; ++row) {
				uint64_t tid = row * side + col;
				uint64_t b = offsets[tid];
				uint64_t e = tid != nt - 1 ? offsets[tid + 1] : m;
				new_offsets[nti++] = (uint32_t)ei;
				for (uint64_t k = b; k < e; ++k) {
					oh[ei] = h[k];
					ot[ei] = t[k];
					if (w && ow) ow[ei] = w[k];
					++ei;
				}
			}
		}
		side_i = rend;
	}
}

static int write_part(const char *fname, int header, uint32_t n, uint64_t m, const uint32_t *h,
                      const uint32_t *t, const uint32_t *w, uint64_t b, uint64_t e)
{
	FILE *f = fopen(fname, "w");
	if (!f) return -1;
	if (header) fprintf(f, "%u %u\n", n, (uint32_t)m);
	for (uint64_t k = b; k < e; ++k) {
		if (w)
			fprintf(f, "%u %u %u\n", h[k] + 1, t[k] + 1, w[k]);
		else
			fprintf(f, "%u %u\n", h[k] + 1, t[k] + 1);
	}
	return fclose(f);
}

/* 64-way split: part k < 63 holds [k*floor(M/64), (k+1)*floor(M/64)), part 63
 * the rest; header only in part 0 (tools/csr_tiling/page_rank.cpp:570-592). */
int orc_write_parts(const char *prefix, uint32_t n, uint64_t m, const uint32_t *h,
                    const uint32_t *t, const uint32_t *w, uint32_t parts)
{
	uint64_t eb = m / parts;
	char fname[4096];
	for (uint32_t k = 0; k < parts; ++k) {
		snprintf(fname, sizeof fname, "%s-%u", prefix, k);
		uint64_t b = k * eb, e = k != parts - 1 ? b + eb : m;
		if (write_part(fname, k == 0, n, m, h, t, w, b, e)) return -1;
	}
	return 0;
}

int orc_write_u32_lines(const char *path, const uint32_t *a, uint64_t count)
{
	FILE *f = fopen(path, "w");
	if (!f) return -1;
	for (uint64_t i = 0; i < count; ++i) fprintf(f, "%u\n", a[i]);
	return fclose(f);
}

int orc_write_edgelist(const char *path, uint32_t n, uint64_t m, const uint32_t *h,
                       const uint32_t *t, const uint32_t *w)
{
	return write_part(path, 1, n, m, h, t, w, 0, m);
}

/* ------------------------------------------------------------------------ */
/* Traversals                                                               */
/* ------------------------------------------------------------------------ */

/* Textbook BFS hop levels (the semantics of apps/bfs_naive/bfs.cpp:316-375). */
void orc_bfs_levels(uint32_t n, const uint64_t *row_off, const uint32_t *cols, uint32_t src,
                    int32_t *level)
{
	for (uint32_t v = 0; v < n; ++v) level[v] = -1;
	if (src >= n) return;
	uint32_t *q = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
	uint64_t qh = 0, qt = 0;
	level[src] = 0;
	q[qt++] = src;
	while (qh < qt) {
		uint32_t u = q[qh++];
		for (uint64_t k = row_off[u]; k < row_off[u + 1]; ++k) {
			uint32_t v = cols[k];
			if (level[v] < 0) {
				level[v] = level[u] + 1;
				q[qt++] = v;
			}
		}
	}
	free(q);
}

/* ---- hybrid BFS: restates apps/bfs_simd/bfs.cpp:54-963 with scalar lanes ---- */

typedef struct {
	uint32_t n, W, S, side;
	uint64_t m, ntiles;
	const uint32_t *deg, *cols, *heads, *tails, *toff;
	uint64_t *vtx;
	int32_t *mask, *updating, *cost, *active_side, *upd_side;
	uint32_t *parents;
} hyb_t;

/* to_dense, bfs.cpp:54-88 */
static void hyb_to_dense(hyb_t *g, const uint32_t *q, uint32_t qs)
{
	memset(g->mask, 0, (size_t)g->n * sizeof(int32_t));
	memset(g->active_side, 0, (size_t)g->side * sizeof(int32_t));
#pragma omp parallel for
	for (uint32_t i = 0; i < qs; ++i) {
		g->mask[q[i]] = 1;
		g->active_side[q[i] / g->W] = 1;
	}
}

/* update_dense, bfs.cpp:91-179 */
static void hyb_update_dense(hyb_t *g, uint32_t *fs_out, uint32_t *od_out)
{
	uint32_t fs = 0, od = 0;
#pragma omp parallel for reduction(+ : fs, od) schedule(dynamic, 16)
	for (uint32_t s = 0; s < g->side; ++s) {
		uint32_t b = s * g->W, e = s + 1 == g->side ? g->n : b + g->W;
		if (!g->upd_side[s]) {
			g->active_side[s] = 0;
			memset(g->mask + b, 0, (size_t)(e - b) * sizeof(int32_t));
			continue;
		}
		g->upd_side[s] = 0;
		g->active_side[s] = 1;
		for (uint32_t v = b; v < e; ++v) {
			if (g->updating[v]) {
				g->updating[v] = 0;
				g->mask[v] = 1;
				fs++;
				od += g->deg[v];
			} else {
				g->mask[v] = 0;
			}
		}
	}
	*fs_out = fs;
	*od_out = od;
}

/* bfs_kernel_dense over one tile's edges, bfs.cpp:181-245 (one lane at a time) */
static void hyb_dense_edges(hyb_t *g, uint64_t b, uint64_t e)
{
	for (uint64_t k = b; k < e; ++k) {
		uint32_t h = g->heads[k];
		if (!g->mask[h]) continue;
		uint32_t t = g->tails[k];
		if (g->parents[t] != UINT32_MAX) continue;
		g->cost[t] = g->cost[h] + 1;
		g->updating[t] = 1;
		g->upd_side[t / g->W] = 1;
		g->parents[t] = h;
	}
}

/* bfs_kernel_dense with AVX-512, bfs.cpp:181-245: 16 edges per vector; the
 * frontier flags of the heads, then the parents of the active edges' tails,
 * are gathered; the unvisited lanes scatter cost, the updating flag and the
 * parent.  Within a vector, lanes with the same tail write the same cost and
 * one of their heads as parent (the scatter order picks it, as in the
 * reference).  The updating flag of the tail's block, which the reference
 * computes with SVML's _mm512_div_epi32, is set lane by lane from the mask. */
__attribute__((target("avx512f"))) static void hyb_dense_edges_avx512(hyb_t *g, uint64_t b, uint64_t e)
{
	const __m512i minus1 = _mm512_set1_epi32(-1), one = _mm512_set1_epi32(1);
	for (uint64_t k = b; k < e; k += 16) {
		const uint64_t left = e - k;
		const __mmask16 in = left >= 16 ? (__mmask16)0xFFFF : (__mmask16)((1u << left) - 1u);
		const __m512i head = _mm512_maskz_loadu_epi32(in, g->heads + k);
		const __m512i act = _mm512_mask_i32gather_epi32(_mm512_setzero_si512(), in, head, g->mask, 4);
		const __mmask16 am = _mm512_mask_test_epi32_mask(in, act, minus1);
		if (!am) continue;
		const __m512i tail = _mm512_maskz_loadu_epi32(am, g->tails + k);
		const __m512i par = _mm512_mask_i32gather_epi32(_mm512_setzero_si512(), am, tail, g->parents, 4);
		const __mmask16 nv = _mm512_mask_cmpeq_epi32_mask(am, par, minus1);
		if (!nv) continue;
		const __m512i ch = _mm512_mask_i32gather_epi32(_mm512_setzero_si512(), nv, head, g->cost, 4);
		_mm512_mask_i32scatter_epi32(g->cost, nv, tail, _mm512_add_epi32(ch, one), 4);
		_mm512_mask_i32scatter_epi32(g->updating, nv, tail, one, 4);
		_mm512_mask_i32scatter_epi32(g->parents, nv, tail, head, 4);
		uint32_t t[16];
		_mm512_storeu_si512(t, tail);
		for (unsigned m = nv; m; m &= m - 1) g->upd_side[t[__builtin_ctz(m)] / g->W] = 1;
	}
}

/* 512: the dense sweep runs hyb_dense_edges_avx512 (the host has AVX-512F
 * and it is not switched off); 0: the one-lane-at-a-time hyb_dense_edges */
static int g_simd = -1;
int orc_simd_width(void)
{
	if (g_simd < 0) g_simd = __builtin_cpu_supports("avx512f") ? 512 : 0;
	return g_simd;
}
void orc_set_simd(int on) { g_simd = on && __builtin_cpu_supports("avx512f") ? 512 : 0; }

/* scheduler_dense / BFS_dense, bfs.cpp:246-388: stripes in sequence, columns of
 * a stripe in parallel (dynamic,1), each column owning its tail block. */
static void hyb_dense(hyb_t *g)
{
	const int simd = orc_simd_width() == 512;
	uint64_t side = g->side;
	uint64_t rem = side % g->S, bound = side - rem;
	for (uint64_t sr = 0; sr < side;) {
		uint64_t step = sr < bound ? g->S : rem;
		uint64_t st = sr * side, et = st + step * side;
#pragma omp parallel for schedule(dynamic, 1)
		for (uint64_t ti = st; ti < et; ti += step) {
			for (uint64_t tile = ti; tile < ti + step; ++tile) {
				uint64_t row = (tile - st) % step + sr;
				uint64_t b = g->toff[tile];
				uint64_t e = tile != g->ntiles - 1 ? g->toff[tile + 1] : g->m;
				if (b == e || !g->active_side[row]) continue;
				if (simd)
					hyb_dense_edges_avx512(g, b, e);
				else
					hyb_dense_edges(g, b, e);
			}
		}
		sr += step;
	}
}

/* to_sparse, bfs.cpp:400-465: ascending ids of the mask */
static uint32_t *hyb_to_sparse(hyb_t *g, uint32_t fs)
{
	uint32_t *q = (uint32_t *)malloc((size_t)(fs ? fs : 1) * sizeof(uint32_t));
	uint32_t k = 0;
	for (uint32_t v = 0; v < g->n; ++v)
		if (g->mask[v]) q[k++] = v;
	return q;
}

/* BFS_kernel_sparse, bfs.cpp:510-711: gather degrees, exclusive scan, expand
 * with CAS on parents, compact. Returns the new queue (NULL when empty). */
static uint32_t *hyb_sparse(hyb_t *g, const uint32_t *q, uint32_t *qs)
{
	uint32_t n_in = *qs;
	uint64_t *off = (uint64_t *)malloc(((size_t)n_in + 1) * sizeof(uint64_t));
	off[0] = 0;
	for (uint32_t i = 0; i < n_in; ++i) off[i + 1] = off[i] + g->deg[q[i]];
	uint64_t tot = off[n_in];
	if (tot == 0) {
		free(off);
		*qs = 0;
		return NULL;
	}
	uint32_t *tmp = (uint32_t *)malloc((size_t)tot * sizeof(uint32_t));
#pragma omp parallel for schedule(dynamic, 2048)
	for (uint32_t i = 0; i < n_in; ++i) {
		uint32_t u = q[i];
		uint64_t base = g->vtx[u];
		for (uint32_t k = 0; k < g->deg[u]; ++k) {
			uint32_t v = g->cols[base + k];
			uint32_t r = UINT32_MAX;
			if (g->parents[v] == UINT32_MAX &&
			    __sync_bool_compare_and_swap(g->parents + v, UINT32_MAX, u))
				r = v;
			tmp[off[i] + k] = r;
		}
	}
	uint64_t k = 0;
	for (uint64_t i = 0; i < tot; ++i)
		if (tmp[i] != UINT32_MAX) tmp[k++] = tmp[i];
	free(off);
	if (k == 0) {
		free(tmp);
		*qs = 0;
		return NULL;
	}
	*qs = (uint32_t)k;
	return tmp;
}

/* update_sparse, bfs.cpp:467-509 */
static uint32_t hyb_update_sparse(hyb_t *g, const uint32_t *q, uint32_t qs)
{
	uint32_t od = 0;
#pragma omp parallel for reduction(+ : od)
	for (uint32_t i = 0; i < qs; ++i) {
		uint32_t v = q[i];
		g->cost[v] = g->cost[g->parents[v]] + 1;
		od += g->deg[v];
	}
	return od;
}

/* graph_prepare, bfs.cpp:753-921: state init (untimed in the reference), first
 * step sparse, then the |F|+sum(outdeg) > M/T_RATIO direction rule in uint32
 * arithmetic (:829-834), until the frontier is empty.  Steps that expand an
 * empty frontier (only possible right after the first step) are not traced. */
int orc_bfs_hybrid(uint32_t n, uint64_t m, uint32_t W, uint32_t S, const uint32_t *csr_deg,
                   const uint32_t *cols, const uint32_t *heads, const uint32_t *tails,
                   const uint32_t *tile_offsets, uint32_t src, uint32_t t_ratio, int nthreads,
                   int32_t *level, int32_t *parent, uint32_t *trace, uint32_t trace_cap)
{
#ifdef _OPENMP
	if (nthreads > 0) omp_set_num_threads(nthreads);
#else
	(void)nthreads;
#endif
	hyb_t G, *g = &G;
	g->n = n;
	g->m = m;
	g->W = W;
	g->S = S;
	g->side = orc_side(n, W);
	g->ntiles = (uint64_t)g->side * g->side;
	g->deg = csr_deg;
	g->cols = cols;
	g->heads = heads;
	g->tails = tails;
	g->toff = tile_offsets;
	g->vtx = (uint64_t *)malloc((size_t)(n ? n : 1) * sizeof(uint64_t));
	g->mask = (int32_t *)calloc(n ? n : 1, sizeof(int32_t));
	g->updating = (int32_t *)calloc(n ? n : 1, sizeof(int32_t));
	g->active_side = (int32_t *)calloc(g->side ? g->side : 1, sizeof(int32_t));
	g->upd_side = (int32_t *)calloc(g->side ? g->side : 1, sizeof(int32_t));
	g->cost = level;
	g->parents = (uint32_t *)parent;
	if (!g->vtx || !g->mask || !g->updating || !g->active_side || !g->upd_side) return -1;
	uint64_t acc = 0;
	for (uint32_t v = 0; v < n; ++v) {
		g->vtx[v] = acc;
		acc += csr_deg[v];
	}
	for (uint32_t v = 0; v < n; ++v) {
		level[v] = -1;
		parent[v] = -1;
	}
	level[src] = 0;
	parent[src] = (int32_t)src;

	int steps = 0;
#define TRACE(d, f, o)                                   \
	do {                                                 \
		if ((uint32_t)steps < trace_cap) {               \
			trace[3 * steps] = (d);                      \
			trace[3 * steps + 1] = (f);                  \
			trace[3 * steps + 2] = (o);                  \
		}                                                \
		++steps;                                         \
	} while (0)

	uint32_t fs = 1;
	uint32_t *frontier = (uint32_t *)malloc(sizeof(uint32_t));
	frontier[0] = src;
	TRACE(0u, 1u, csr_deg[src]);
	uint32_t *nf = hyb_sparse(g, frontier, &fs);
	free(frontier);
	frontier = nf;
	uint32_t od = frontier ? hyb_update_sparse(g, frontier, fs) : 0;
	int last_dense = 0;
	uint32_t threshold = (uint32_t)m / t_ratio;
	while (1) {
		if ((uint32_t)(fs + od) > threshold) {
			if (!last_dense) hyb_to_dense(g, frontier, fs);
			if (fs) TRACE(1u, fs, od);
			hyb_dense(g);
			last_dense = 1;
		} else {
			if (last_dense) {
				free(frontier);
				frontier = hyb_to_sparse(g, fs);
			}
			if (fs) TRACE(0u, fs, od);
			nf = fs ? hyb_sparse(g, frontier, &fs) : NULL;
			if (!nf) fs = 0;
			free(frontier);
			frontier = nf;
			last_dense = 0;
		}
		if (last_dense) {
			hyb_update_dense(g, &fs, &od);
			if (fs == 0) break;
		} else {
			if (fs == 0) break;
			od = hyb_update_sparse(g, frontier, fs);
		}
	}
#undef TRACE
	free(frontier);
	free(g->vtx);
	free(g->mask);
	free(g->updating);
	free(g->active_side);
	free(g->upd_side);
	return steps;
}

/* Parent-tree invariant: parent[src]=src; reached v != src has an edge
 * (parent[v], v) with level[parent[v]] == level[v]-1; unreached has -1.
 * Returns 0 when valid, else 1 + the first offending vertex.  The edge test
 * is one pass over the CSR (every edge (u, w) with parent[w] == u marks w),
 * so hubs with many children cost their degree once, not once per child. */
int64_t orc_check_parents(uint32_t n, const uint64_t *row_off, const uint32_t *cols,
                          uint32_t src, const int32_t *level, const int32_t *parent)
{
	uint8_t *has_edge = (uint8_t *)calloc(n ? n : 1, 1);
	if (!has_edge) return -1;
#pragma omp parallel for schedule(dynamic, 1024)
	for (uint32_t u = 0; u < n; ++u)
		for (uint64_t k = row_off[u]; k < row_off[u + 1]; ++k) {
			const uint32_t w = cols[k];
			if (parent[w] == (int32_t)u) has_edge[w] = 1;
		}
	int64_t bad = 0;
	for (uint32_t v = 0; v < n && !bad; ++v) {
		if (v == src) {
			if (parent[v] != (int32_t)src || level[v] != 0) bad = (int64_t)v + 1;
			continue;
		}
		if (level[v] < 0) {
			if (parent[v] != -1) bad = (int64_t)v + 1;
			continue;
		}
		int32_t p = parent[v];
		if (p < 0 || (uint32_t)p >= n || level[p] != level[v] - 1 || !has_edge[v]) bad = (int64_t)v + 1;
	}
	free(has_edge);
	return bad;
}

/* Dijkstra with a binary heap; unreached = INT_MAX as the reference
 * initialises its distances (apps/sssp_simd/sssp.cpp:995-998). */
void orc_sssp_dijkstra(uint32_t n, const uint64_t *row_off, const uint32_t *cols,
                       const uint32_t *wts, uint32_t src, uint32_t *dist)
{
	for (uint32_t v = 0; v < n; ++v) dist[v] = INT_MAX;
	if (src >= n) return;
	uint64_t cap = 1024, sz = 0;
	uint64_t *heap = (uint64_t *)malloc(cap * sizeof(uint64_t)); /* (dist<<32)|v */
	dist[src] = 0;
	heap[sz++] = (uint64_t)src;
	while (sz) {
		uint64_t top = heap[0];
		heap[0] = heap[--sz];
		for (uint64_t i = 0;;) {
			uint64_t l = 2 * i + 1, r = l + 1, s = i;
			if (l < sz && heap[l] < heap[s]) s = l;
			if (r < sz && heap[r] < heap[s]) s = r;
			if (s == i) break;
			uint64_t tmp = heap[i];
			heap[i] = heap[s];
			heap[s] = tmp;
			i = s;
		}
		uint32_t d = (uint32_t)(top >> 32), u = (uint32_t)top;
		if (d != dist[u]) continue;
		for (uint64_t k = row_off[u]; k < row_off[u + 1]; ++k) {
			uint32_t v = cols[k];
			uint64_t nd = (uint64_t)d + wts[k];
			if (nd < dist[v]) {
				dist[v] = (uint32_t)nd;
				if (sz == cap) {
					cap *= 2;
					heap = (uint64_t *)realloc(heap, cap * sizeof(uint64_t));
				}
				uint64_t i = sz++;
				heap[i] = (nd << 32) | v;
				while (i && heap[(i - 1) / 2] > heap[i]) {
					uint64_t p = (i - 1) / 2, tmp = heap[p];
					heap[p] = heap[i];
					heap[i] = tmp;
					i = p;
				}
			}
		}
	}
	free(heap);
}

/* FP64 power iteration of the reference update rule
 * (apps/pageRank_simd/page_rank.cpp:128-199 accumulation, :428-446 update;
 * apps/pageRank_naive/page_rank.cpp:182-197): sum[t] += rank[h]/deg[h];
 * rank = (1-d)/N + d*sum; dangling mass is dropped, no normalisation. */
void orc_pagerank_f64(uint32_t n, uint64_t m, const uint32_t *heads, const uint32_t *tails,
                      const uint32_t *deg, uint32_t iters, double damping, double init,
                      double *rank)
{
	double *sum = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
	for (uint32_t v = 0; v < n; ++v) rank[v] = init;
	for (uint32_t it = 0; it < iters; ++it) {
		memset(sum, 0, (size_t)n * sizeof(double));
		for (uint64_t e = 0; e < m; ++e) sum[tails[e]] += rank[heads[e]] / (double)deg[heads[e]];
		for (uint32_t v = 0; v < n; ++v) rank[v] = (1.0 - damping) / (double)n + damping * sum[v];
	}
	free(sum);
}

/* apps/bellman-ford_simd/bellman-ford.cpp:129-191,291-392: in-place sweeps of
 * dist[t] = min(dist[t], dist[h]+1) in edge order from the source, init
 * INT_MAX-2 (:308-312), at most `sweeps` sweeps (99 in the reference, :322).
 * A sweep that changes nothing ends the loop early: later sweeps would be
 * identical, so the result is the same. */
void orc_bellman_ford_hops(uint32_t n, uint64_t m, const uint32_t *heads, const uint32_t *tails,
                           uint32_t src, uint32_t sweeps, uint32_t *dist)
{
	for (uint32_t v = 0; v < n; ++v) dist[v] = INT_MAX - 2;
	if (src < n) dist[src] = 0;
	for (uint32_t s = 0; s < sweeps; ++s) {
		int changed = 0;
		for (uint64_t e = 0; e < m; ++e) {
			int32_t nd = (int32_t)dist[heads[e]] + 1;
			if (nd < (int32_t)dist[tails[e]]) {
				dist[tails[e]] = (uint32_t)nd;
				changed = 1;
			}
		}
		if (!changed) break;
	}
}

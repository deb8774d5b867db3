/*
 * oracle.h -- CPU restatement of GraphPhi's tiled graph-traversal path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the
 * smoke() entry and the cpu_baseline leg of bench.py may load it, nothing
 * in the product (libgphi, the CLIs) links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   orc_glibc_rand        -> glibc rand() called through ctypes (same values)
 *   orc_degree_rank       -> reference tool vertex_id_reorder_to_degrees built in oracle/_ref
 *   orc_bfs_order_rank    -> reference tool vertex_id_remap built in oracle/_ref
 *   orc_untiled/coo_tile/col_major + writers -> byte-compare with reference
 *                            tools csr_tiling / coo_tiling / column_major_tile
 *   orc_bfs_levels        -> reference bfs_naive level dump (-DONEDEBUG) and
 *                            brute-force all-pairs on tiny graphs
 *   orc_bfs_hybrid        -> levels == orc_bfs_levels; direction trace == closed
 *                            form of the T_RATIO rule; parents validated
 *   orc_sssp_dijkstra     -> brute-force Bellman-Ford on tiny graphs, w==1 reduces
 *                            to BFS levels
 *   orc_pagerank_f64      -> closed form on regular graphs (cycle), reference
 *                            pageRank_naive rank dump (-DONEDEBUG, one iteration)
 *   orc_bellman_ford_hops -> reference bellman-ford_simd dist[0..19] print
 *   orc_rmat_*            -> parity unpinned (the reference ships no generator;
 *                            the spec is this project's own and is pinned only by
 *                            distribution invariants and product==oracle tests)
 */
#ifndef GPHI_ORACLE_H
#define GPHI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic inputs (project spec; parity unpinned) ---- */
void orc_rmat_thresholds(double a, double b, double c, uint32_t thr[3]);
void orc_rmat_sample(uint64_t seed, uint64_t i, uint32_t scale, const uint32_t thr[3],
                     uint32_t *h, uint32_t *t);
/* Returns the number of unique non-loop edges written (sorted by (h,t)).
 * target_unique == 0: all nsamples are drawn and deduplicated.
 * target_unique  > 0: samples are drawn in order until that many unique
 * edges with both ids < nverts exist (or nsamples are exhausted). */
uint64_t orc_rmat_edges(uint32_t scale, uint32_t nverts, uint64_t seed, uint64_t nsamples,
                        uint64_t target_unique, double a, double b, double c,
                        uint32_t *h_out, uint32_t *t_out, uint64_t cap);
void orc_random_perm(uint32_t n, uint64_t seed, uint32_t *perm);
void orc_weights(uint64_t m, uint64_t seed, uint32_t lo, uint32_t hi, uint32_t *w);
void orc_relabel(uint64_t m, const uint32_t *map, uint32_t *h, uint32_t *t);

/* ---- reorderings (tools/vertex_id_reorder_to_degrees, tools/vertex_id_remap) ---- */
void orc_glibc_rand(uint32_t seed, uint64_t n, int32_t *out);
void orc_degree_rank(uint32_t n, const uint32_t *deg, uint32_t *rank);
void orc_bfs_order_rank(uint32_t n, const uint64_t *row_off, const uint32_t *cols,
                        uint32_t *rank);

/* ---- layouts (tools/csr_tiling, tools/coo_tiling, tools/column_major_tile) ---- */
void orc_untiled(uint32_t n, uint64_t m, const uint32_t *h, const uint32_t *t, const uint32_t *w,
                 uint32_t *oh, uint32_t *ot, uint32_t *ow, uint32_t *deg);
uint32_t orc_side(uint32_t n, uint32_t W);
void orc_coo_tile(uint32_t n, uint64_t m, uint32_t W, const uint32_t *h, const uint32_t *t,
                  const uint32_t *w, uint32_t *oh, uint32_t *ot, uint32_t *ow,
                  uint32_t *offsets /* side*side */);
void orc_col_major(uint32_t n, uint64_t m, uint32_t W, uint32_t S, const uint32_t *h,
                   const uint32_t *t, const uint32_t *w, const uint32_t *offsets,
                   uint32_t *oh, uint32_t *ot, uint32_t *ow, uint32_t *new_offsets);
/* writes prefix-{0..parts-1}; header "n m" in part 0; ids written 1-based */
int orc_write_parts(const char *prefix, uint32_t n, uint64_t m, const uint32_t *h,
                    const uint32_t *t, const uint32_t *w, uint32_t parts);
int orc_write_u32_lines(const char *path, const uint32_t *a, uint64_t count);
int orc_write_edgelist(const char *path, uint32_t n, uint64_t m, const uint32_t *h,
                       const uint32_t *t, const uint32_t *w);

/* ---- traversal oracles ---- */
void orc_bfs_levels(uint32_t n, const uint64_t *row_off, const uint32_t *cols, uint32_t src,
                    int32_t *level);
/* trace: per expansion step k, trace[3k]=dir (0 sparse, 1 dense),
 * trace[3k+1]=|F| and trace[3k+2]=sum outdeg(F) of the frontier it expanded.
 * Returns the number of steps (or -1 on allocation failure). */
int orc_bfs_hybrid(uint32_t n, uint64_t m, uint32_t W, uint32_t S, const uint32_t *csr_deg,
                   const uint32_t *cols, const uint32_t *heads, const uint32_t *tails,
                   const uint32_t *tile_offsets, uint32_t src, uint32_t t_ratio, int nthreads,
                   int32_t *level, int32_t *parent, uint32_t *trace, uint32_t trace_cap);
/* vector width of orc_bfs_hybrid's dense sweep: 512 (AVX-512 restatement of
 * bfs.cpp:181-245) or 0 (one lane at a time); orc_set_simd(0) forces the latter */
int orc_simd_width(void);
void orc_set_simd(int on);
int64_t orc_check_parents(uint32_t n, const uint64_t *row_off, const uint32_t *cols,
                          uint32_t src, const int32_t *level, const int32_t *parent);
void orc_sssp_dijkstra(uint32_t n, const uint64_t *row_off, const uint32_t *cols,
                       const uint32_t *wts, uint32_t src, uint32_t *dist);
void orc_pagerank_f64(uint32_t n, uint64_t m, const uint32_t *heads, const uint32_t *tails,
                      const uint32_t *deg, uint32_t iters, double damping, double init,
                      double *rank);
void orc_bellman_ford_hops(uint32_t n, uint64_t m, const uint32_t *heads, const uint32_t *tails,
                           uint32_t src, uint32_t sweeps, uint32_t *dist);

#ifdef __cplusplus
}
#endif
#endif

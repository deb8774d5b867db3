// api.cpp -- extern "C" entry points of libgphi (include/gphi.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstring>
#include <new>

#include "internal.h"

namespace gphi {

static thread_local std::string g_orphan_error;

int set_error(gphi_ctx *ctx, int code, const char *fmt, ...)
{
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	if (ctx) {
		ctx->err.code = code;
		ctx->err.msg = buf;
	} else {
		g_orphan_error = buf;
	}
	return code;
}

// allocations left before the injected failure (gphi_debug_fail_alloc); < 0: none
static std::atomic<long> g_fail_alloc_in{-1};

bool take_injected_alloc_failure()
{
	long left = g_fail_alloc_in.load(std::memory_order_relaxed);
	while (left >= 0) {
		if (g_fail_alloc_in.compare_exchange_weak(left, left - 1, std::memory_order_relaxed))
			return left == 0;
	}
	return false;
}

int hip_fail(gphi_ctx *ctx, hipError_t e, const char *what, const char *file, int line)
{
	return set_error(ctx, GPHI_E_HIP, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file,
	                 line);
}

}  // namespace gphi

using gphi::set_error;

extern "C" {

const char *gphi_version(void) { return "gphi 0.1 (gfx950)"; }

void gphi_debug_fail_alloc(int after) { gphi::g_fail_alloc_in.store(after < 0 ? -1 : after); }

int gphi_open(const gphi_opts *opts, gphi_ctx **out)
{
	if (!out) return set_error(nullptr, GPHI_E_ARG, "gphi_open: out is NULL");
	*out = nullptr;
	gphi_ctx *ctx = new (std::nothrow) gphi_ctx();
	if (!ctx) return set_error(nullptr, GPHI_E_OOM, "gphi_open: out of host memory");
	ctx->device = opts ? opts->device : 0;
	ctx->t_ratio = (opts && opts->t_ratio) ? opts->t_ratio : 20;
	int ndev = 0;
	hipError_t e = hipGetDeviceCount(&ndev);
	if (e != hipSuccess || ndev <= ctx->device || ctx->device < 0) {
		int rc = set_error(nullptr, GPHI_E_HIP, "gphi_open: no HIP device %d (count %d, %s)",
		                   ctx->device, ndev, hipGetErrorString(e));
		delete ctx;
		return rc;
	}
	if ((e = hipSetDevice(ctx->device)) != hipSuccess ||
	    (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess ||
	    (e = hipHostMalloc((void **)&ctx->h_ctrl, sizeof(gphi::StepCtrl), hipHostMallocDefault)) !=
	        hipSuccess) {
		int rc = set_error(nullptr, GPHI_E_HIP, "gphi_open: %s", hipGetErrorString(e));
		if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
		delete ctx;
		return rc;
	}
	memset(ctx->h_ctrl, 0, sizeof(gphi::StepCtrl));
	gphi::noalloc_setup(ctx->device);  // device address of gphi_bfs_noalloc's scratch array
	ctx->ev_ok = true;
	for (auto &ev : ctx->ev)
		if (hipEventCreate(&ev) != hipSuccess) ctx->ev_ok = false;
	*out = ctx;
	return GPHI_OK;
}

int gphi_set_level_timing(gphi_ctx *ctx, int on)
{
	if (!ctx) return set_error(nullptr, GPHI_E_ARG, "gphi_set_level_timing: ctx is NULL");
	std::lock_guard<std::mutex> lk(ctx->mu);
	ctx->level_timing = on != 0;
	return GPHI_OK;
}

void gphi_close(gphi_ctx *ctx)
{
	if (!ctx) return;
	(void)hipSetDevice(ctx->device);
	if (ctx->ev_ok)
		for (auto &ev : ctx->ev) (void)hipEventDestroy(ev);
	if (ctx->h_ctrl) (void)hipHostFree(ctx->h_ctrl);
	if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

const char *gphi_last_error(gphi_ctx *ctx)
{
	if (!ctx) return gphi::g_orphan_error.c_str();
	ctx->last = ctx->err.msg;
	return ctx->last.c_str();
}

static int check_desc(gphi_ctx *ctx, const gphi_graph_desc *d)
{
	if (!d) return set_error(ctx, GPHI_E_ARG, "graph descriptor is NULL");
	if (d->n == 0) return set_error(ctx, GPHI_E_RANGE, "graph has no vertices");
	if (d->n > 0xFFFFFFFFu - 2 * gphi::kPadVerts) return set_error(ctx, GPHI_E_RANGE, "N=%u too large", d->n);
	if (d->m >= (1ull << 32)) return set_error(ctx, GPHI_E_RANGE, "M=%llu does not fit the %%u formats",
	                                           (unsigned long long)d->m);
	if (d->tile_width == 0) return set_error(ctx, GPHI_E_ARG, "tile width is 0");
	if (d->stripe_length == 0) return set_error(ctx, GPHI_E_ARG, "stripe length is 0");
	if (!d->degrees) return set_error(ctx, GPHI_E_ARG, "degrees are required");
	if (!d->csr_edges && !(d->tile_heads && d->tile_tails))
		return set_error(ctx, GPHI_E_ARG, "need CSR edges or tile arrays");
	if (d->tile_order != GPHI_TILES_COL_STRIPE && d->tile_order != GPHI_TILES_ROW_MAJOR)
		return set_error(ctx, GPHI_E_ARG, "unknown tile order %d", d->tile_order);
	const uint32_t n = d->n;
	// ids must be in range: a bad id would make a kernel access out of bounds
	uint64_t bad = UINT64_MAX;
	if (d->csr_edges) {
#pragma omp parallel for reduction(min : bad)
		for (uint64_t e = 0; e < d->m; ++e)
			if (d->csr_edges[e] >= n) bad = std::min<uint64_t>(bad, e);
	}
	if (bad != UINT64_MAX)
		return set_error(ctx, GPHI_E_RANGE, "CSR edge %llu: tail %u >= N %u", (unsigned long long)bad,
		                 d->csr_edges[bad], n);
	if (d->tile_heads && d->tile_tails) {
#pragma omp parallel for reduction(min : bad)
		for (uint64_t e = 0; e < d->m; ++e)
			if (d->tile_heads[e] >= n || d->tile_tails[e] >= n) bad = std::min<uint64_t>(bad, e);
	}
	if (bad != UINT64_MAX)
		return set_error(ctx, GPHI_E_RANGE, "tile edge %llu: (%u,%u) out of range N %u",
		                 (unsigned long long)bad, d->tile_heads[bad], d->tile_tails[bad], n);
	if (d->csr_edges) {
		uint64_t sum = 0;
		for (uint32_t v = 0; v < n; ++v) sum += d->degrees[v];
		if (sum != d->m)
			return set_error(ctx, GPHI_E_FORMAT, "sum of degrees %llu != M %llu",
			                 (unsigned long long)sum, (unsigned long long)d->m);
	}
	return GPHI_OK;
}

int gphi_graph_from_desc(gphi_ctx *ctx, const gphi_graph_desc *desc, gphi_graph **out)
{
	if (!ctx || !out) return set_error(ctx, GPHI_E_ARG, "gphi_graph_from_desc: NULL argument");
	std::lock_guard<std::mutex> lk(ctx->mu);
	*out = nullptr;
	int rc = check_desc(ctx, desc);
	if (rc) return rc;
	if (hipSetDevice(ctx->device) != hipSuccess)
		return set_error(ctx, GPHI_E_HIP, "hipSetDevice(%d) failed", ctx->device);
	gphi_graph *g = new (std::nothrow) gphi_graph();
	if (!g) return set_error(ctx, GPHI_E_OOM, "out of host memory");
	g->ctx = ctx;
	rc = gphi::upload_graph(ctx, desc, g);
	if (rc) {
		gphi::free_graph_device(g);
		delete g;
		return rc;
	}
	*out = g;
	return GPHI_OK;
}

static int load_text(gphi_ctx *ctx, const char *prefix, uint32_t W, uint32_t S, int weighted,
                     bool row_major, gphi_graph **out)
{
	if (!ctx || !prefix || !out) return set_error(ctx, GPHI_E_ARG, "NULL argument");
	*out = nullptr;
	if (W == 0 || S == 0) return set_error(ctx, GPHI_E_ARG, "tile width and stripe length must be > 0");
	gphi_graph_desc d;
	std::vector<uint32_t> store[8];
	int rc = gphi::read_text_graph(ctx, prefix, W, S, weighted, row_major, &d, store);
	if (rc) return rc;
	return gphi_graph_from_desc(ctx, &d, out);
}

int gphi_graph_load_text(gphi_ctx *ctx, const char *prefix, uint32_t W, uint32_t S, int weighted,
                         gphi_graph **out)
{
	return load_text(ctx, prefix, W, S, weighted, false, out);
}

int gphi_graph_load_text_rowmajor(gphi_ctx *ctx, const char *prefix, uint32_t W, uint32_t S,
                                  gphi_graph **out)
{
	return load_text(ctx, prefix, W, S, 0, true, out);
}

int gphi_graph_info(const gphi_graph *g, uint32_t *n, uint64_t *m, uint32_t *W, uint32_t *S,
                    int *weighted)
{
	if (!g) return set_error(nullptr, GPHI_E_ARG, "graph is NULL");
	if (n) *n = g->g.n;
	if (m) *m = g->g.m;
	if (W) *W = g->g.W;
	if (S) *S = g->g.S;
	if (weighted) *weighted = g->g.weighted ? 1 : 0;
	return GPHI_OK;
}

uint64_t gphi_graph_device_bytes(const gphi_graph *g) { return g ? g->g.device_bytes : 0; }

int gphi_graph_view(gphi_ctx *ctx, gphi_graph *base, gphi_graph **out)
{
	if (!ctx || !base || !out) return set_error(ctx, GPHI_E_ARG, "gphi_graph_view: NULL argument");
	*out = nullptr;
	if (base->base) base = base->base;
	if (ctx->device != base->ctx->device)
		return set_error(ctx, GPHI_E_ARG, "a view must live on its graph's device (%d, not %d)",
		                 base->ctx->device, ctx->device);
	gphi_graph *v = new (std::nothrow) gphi_graph();
	if (!v) return set_error(ctx, GPHI_E_OOM, "out of host memory");
	{
		std::lock_guard<std::mutex> lk(base->ctx->mu);
		v->g = base->g;
	}
	v->g.device_bytes = 0;
	v->ctx = ctx;
	v->base = base;
	*out = v;
	return GPHI_OK;
}

void gphi_graph_free(gphi_graph *g)
{
	if (!g) return;
	std::lock_guard<std::mutex> lk(g->ctx->mu);
	(void)hipSetDevice(g->ctx->device);
	gphi::free_graph_device(g);
	delete g;
}

#define GPHI_ENTER(g)                                                              \
	if (!(g)) return set_error(nullptr, GPHI_E_ARG, "graph is NULL");              \
	std::lock_guard<std::mutex> lk((g)->ctx->mu);                                  \
	if (hipSetDevice((g)->ctx->device) != hipSuccess)                              \
		return set_error((g)->ctx, GPHI_E_HIP, "hipSetDevice failed");

int gphi_bfs(gphi_graph *g, uint32_t src, int32_t *levels, int32_t *parents, gphi_stats *st)
{
	GPHI_ENTER(g);
	return gphi::run_bfs(g, src, levels, parents, st);
}

int gphi_bfs_noalloc(gphi_graph *g, uint32_t src, int32_t *levels_dev, int32_t *parents_dev, gphi_stats *st)
{
	GPHI_ENTER(g);
	return gphi::run_bfs_noalloc(g, src, levels_dev, parents_dev, st);
}

int gphi_bfs_batch(gphi_graph *g, const uint32_t *srcs, uint32_t k, uint8_t *levels, gphi_stats *st)
{
	GPHI_ENTER(g);
	return gphi::run_bfs_batch(g, srcs, k, levels, st);
}

int gphi_sssp(gphi_graph *g, uint32_t src, uint32_t *dist, gphi_stats *st)
{
	GPHI_ENTER(g);
	return gphi::run_sssp(g, src, dist, st);
}

int gphi_sssp_batch(gphi_graph *g, const uint32_t *srcs, uint32_t k, uint32_t *dist, gphi_stats *st)
{
	GPHI_ENTER(g);
	uint32_t *rows[GPHI_SSSP_BATCH_MAX] = {};
	for (uint32_t i = 0; dist && i < k && i < GPHI_SSSP_BATCH_MAX; ++i) rows[i] = dist + (size_t)i * g->g.n;
	return gphi::run_sssp_batch(g, srcs, k, dist ? rows : nullptr, st);
}

int gphi_pagerank(gphi_graph *g, uint32_t iters, float damping, float init, float *rank, float *ms)
{
	GPHI_ENTER(g);
	return gphi::run_pagerank(g, iters, damping, init, rank, ms);
}

int gphi_bellman_ford(gphi_graph *g, uint32_t src, uint32_t max_sweeps, uint32_t *dist,
                      uint32_t *sweeps_run, float *ms)
{
	GPHI_ENTER(g);
	return gphi::run_bellman_ford(g, src, max_sweeps, dist, sweeps_run, ms);
}

// Byte model of a finished BFS.  Sparse step expanding F:
//   4|F| + 8|F| + 4 D_F + 4 D_F + 12 |F'|
// dense step:  8 E_act + 4 T_act + 3 ceil(N/8) + 8 |F'|, E_act/T_act = edges /
// count of non-empty tiles whose head row holds a frontier vertex and whose
// tail column still had an unvisited vertex when the step began; a direction
// switch adds ceil(N/8) (to dense) or ceil(N/8) + 4|F| (to sparse).
int gphi_bfs_alg_bytes(const gphi_graph *gr, const int32_t *levels, const gphi_stats *st,
                       double *out)
{
	if (!gr || !levels || !st || !out) return set_error(nullptr, GPHI_E_ARG, "NULL argument");
	const gphi::DevGraph &G = gr->g;
	const uint32_t steps = st->steps;
	const double nbm = (double)((G.n + 7) / 8);
	std::vector<uint32_t> dense_steps;
	for (uint32_t k = 0; k < steps; ++k)
		if (st->dir[k]) dense_steps.push_back(k);
	std::vector<double> e_d, t_d;
	std::lock_guard<std::mutex> lk(gr->ctx->mu);
	if (hipSetDevice(gr->ctx->device) != hipSuccess) return set_error(gr->ctx, GPHI_E_HIP, "hipSetDevice failed");
	int rc = gphi::count_active_tiles(const_cast<gphi_graph *>(gr), levels, dense_steps, e_d, t_d);
	if (rc) return rc;
	std::vector<double> e_act(steps, 0.0), t_act(steps, 0.0);
	for (size_t j = 0; j < dense_steps.size(); ++j) {
		e_act[dense_steps[j]] = e_d[j];
		t_act[dense_steps[j]] = t_d[j];
	}
	for (uint32_t k = 0; k < steps; ++k) {
		double f = st->fsize[k], dF = st->fdeg[k];
		double fn = k + 1 < steps ? st->fsize[k + 1] : 0.0;
		double b;
		if (st->dir[k] == 0)
			b = 4 * f + 8 * f + 4 * dF + 4 * dF + 12 * fn;
		else
			b = 8 * e_act[k] + 4 * t_act[k] + 3 * nbm + 8 * fn;
		if (k > 0 && st->dir[k] != st->dir[k - 1]) b += st->dir[k] ? nbm : nbm + 4 * f;
		out[k] = b;
	}
	return GPHI_OK;
}

int gphi_graph_generate(gphi_ctx *ctx, const gphi_rmat_spec *spec, uint32_t W, uint32_t S,
                        int reorder, const gphi_weight_spec *weights, int tile_order,
                        gphi_graph **out)
{
	if (!ctx || !spec || !out) return set_error(ctx, GPHI_E_ARG, "gphi_graph_generate: NULL argument");
	std::lock_guard<std::mutex> lk(ctx->mu);
	*out = nullptr;
	if (tile_order != GPHI_TILES_COL_STRIPE && tile_order != GPHI_TILES_ROW_MAJOR)
		return set_error(ctx, GPHI_E_ARG, "unknown tile order %d", tile_order);
	if (weights && weights->seed && weights->lo > weights->hi)
		return set_error(ctx, GPHI_E_ARG, "weight range [%u,%u] is empty", weights->lo, weights->hi);
	if (hipSetDevice(ctx->device) != hipSuccess) return set_error(ctx, GPHI_E_HIP, "hipSetDevice failed");
	gphi_graph *g = new (std::nothrow) gphi_graph();
	if (!g) return set_error(ctx, GPHI_E_OOM, "out of host memory");
	g->ctx = ctx;
	int rc = gphi::generate_graph(ctx, spec, W, S, reorder, weights, tile_order, g);
	if (rc) {
		gphi::free_graph_device(g);
		delete g;
		return rc;
	}
	*out = g;
	return GPHI_OK;
}

int gphi_graph_download(gphi_graph *g, uint32_t *degrees, uint32_t *csr_edges, uint32_t *csr_weights,
                        uint32_t *tile_heads, uint32_t *tile_tails, uint32_t *tile_weights)
{
	GPHI_ENTER(g);
	return gphi::download_graph(g, degrees, csr_edges, csr_weights, tile_heads, tile_tails, tile_weights);
}

int gphi_host_from_device(gphi_graph *g, gphi_host_graph **out)
{
	if (!out) return set_error(nullptr, GPHI_E_ARG, "out is NULL");
	*out = nullptr;
	gphi_graph_desc d;
	std::vector<uint32_t> st[7];
	{
		GPHI_ENTER(g);
		const gphi::DevGraph &G = g->g;
		memset(&d, 0, sizeof d);
		d.n = G.n;
		d.m = G.m;
		d.tile_width = G.W;
		d.stripe_length = G.S;
		d.tile_order = G.tile_order;
		try {
			st[0].resize(G.n);
			if (G.has_csr) st[1].resize(G.m);
			if (G.csr_w) st[2].resize(G.m);
			if (G.has_tiles) {
				st[3].resize(G.m);
				st[4].resize(G.m);
				st[6].resize(G.ntiles);
			}
			if (G.coo_w) st[5].resize(G.m);
		} catch (const std::bad_alloc &) {
			return set_error(g->ctx, GPHI_E_OOM, "host allocation failed");
		}
		auto ptr = [](std::vector<uint32_t> &v) { return v.empty() ? nullptr : v.data(); };
		int rc = gphi::download_graph(g, ptr(st[0]), ptr(st[1]), ptr(st[2]), ptr(st[3]), ptr(st[4]), ptr(st[5]));
		if (!rc && G.has_tiles) rc = gphi::tile_offsets(g, st[6].data());
		if (rc) return rc;
	}
	return gphi::host_adopt(&d, st, out);
}

int gphi_graph_retile(gphi_graph *g, uint32_t W, uint32_t S, int tile_order)
{
	GPHI_ENTER(g);
	if (g->base) return set_error(g->ctx, GPHI_E_STATE, "a view cannot re-tile the graph it shares");
	return gphi::retile_graph(g, W, S, tile_order);
}

int gphi_graph_tile_offsets(gphi_graph *g, uint32_t *out)
{
	GPHI_ENTER(g);
	if (!out) return set_error(g->ctx, GPHI_E_ARG, "out is NULL");
	return gphi::tile_offsets(g, out);
}

}  // extern "C"

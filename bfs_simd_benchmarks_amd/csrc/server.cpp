// server.cpp -- query serving over resident graphs with continuous batching.
//
// No reference counterpart (the reference runs one traversal per process,
// apps/bfs_simd/bfs.cpp:1176-1199).  Each admitted graph gets a BFS lane: its
// own gphi_ctx (HIP stream) and worker thread, so the lanes' kernels overlap
// on the device.  A lane's worker takes every BFS query pending for its graph
// when it becomes free (up to max_batch) and answers them with one
// multi-source traversal when that is expected to beat running them one by
// one.  The traversal advances level by level; between two levels the worker
// admits the BFS queries that arrived meanwhile into free source slots (they
// join the running traversal, each from its own level 0), and every query
// completes as soon as its own frontier is empty: its levels are copied out
// while the traversal goes on for the others.  A weighted graph also gets SSSP
// lanes -- views of the graph (gphi_graph_view) on streams of their own --
// that run SSSP queries beside the BFS lane: the SSSP queries pending on a
// lane share one multi-source traversal (up to GPHI_SSSP_BATCH_MAX) when it
// is expected to take less time than running them alone.  Graphs are admitted
// while their HBM bytes plus traversal state fit the budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double now_us()
{
	return std::chrono::duration<double, std::micro>(Clock::now().time_since_epoch()).count();
}

struct Query {
	uint64_t ticket;
	uint32_t kind;  // GPHI_QUERY_*
	uint32_t src;
	void *out;      // caller buffer (int32 levels / uint32 distances), may be null
	double t_submit = 0, t_start = 0, t_done = 0;
	uint32_t batch = 0, joined_level = 0;
	uint64_t launch = 0;
	int status = GPHI_OK;
	bool done = false;
};

// A multi-source traversal costs a fixed amount per level (the update and
// frontier passes, a host round trip; for SSSP, 16 distances per vertex
// whatever the number of sources) plus a little per source, so a query's
// share of one falls as the group grows.  RunCosts keeps the time of a whole
// traversal per size class, k in [2^c, 2^(c+1)), the class of a BFS traversal
// counting the queries that joined it.
struct RunCosts {
	double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
	static int size_class(size_t k)
	{
		int c = 0;
		while (c < 8 && (size_t(2) << c) <= k) ++c;
		return c;
	}
	// the estimate for k's class, else that of the nearest larger class
	// measured (an upper bound for k); 0 when none is
	double upper(size_t k) const
	{
		for (int c = size_class(k); c < 9; ++c)
			if (t[c] > 0) return t[c];
		return 0;
	}
};

struct Lane {
	gphi_graph *g = nullptr;
	gphi_ctx *ctx = nullptr;
	uint32_t n = 0;
	std::deque<std::shared_ptr<Query>> pending;
	std::thread worker;
	uint64_t device_bytes = 0;
	uint64_t batches = 0, served = 0;
	// running estimates (microseconds) of one query run alone, and of a whole
	// multi-source traversal per size class (RunCosts); 0 = not measured yet
	double t_plain = 0, t_sssp = 0;
	RunCosts bfs_runs, sssp_runs;
	bool sssp_batch_ok = false;  // the lane's graph was prepared for SSSP batches
	bool busy = false;  // a launch is running (guarded by the server mutex)
};

void ema(double &x, double v) { x = x > 0 ? 0.8 * x + 0.2 * v : v; }

// k pending queries share one traversal when that lowers their mean latency:
// run one after another, the i-th of them completes after i runs alone, so
// their mean is (k + 1) / 2 runs; sharing, each completes with the traversal.
// (Comparing with all k runs instead maximises throughput but makes queries
// wait on wide traversals at loads the runs alone still keep up with.)
// Before both estimates exist: when at least `first` are pending.
bool share_pays(const RunCosts &runs, double t_alone, size_t k, size_t first)
{
	if (k <= 1) return false;
	const double run = runs.upper(k);
	if (t_alone > 0 && run > 0) return run < 0.5 * (double)(k + 1) * t_alone;
	return k >= first;
}

bool use_batch(const Lane &L, size_t k, uint32_t max_batch)
{
	return max_batch > 1 && share_pays(L.bfs_runs, L.t_plain, k, 4);
}

bool use_sssp_batch(const Lane &L, size_t k)
{
	return L.sssp_batch_ok && share_pays(L.sssp_runs, L.t_sssp, k, 8);
}

}  // namespace

struct gphi_server {
	gphi_server_opts opts{};
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::vector<std::unique_ptr<Lane>> lanes;
	std::vector<std::pair<Lane *, std::vector<Lane *>>> graphs;  // per graph id: BFS lane, SSSP lanes
	std::unordered_map<uint64_t, std::shared_ptr<Query>> queries;
	uint64_t next_ticket = 1;
	uint64_t next_launch = 1;
	uint64_t hbm_used = 0;
	bool stopping = false;
};

namespace gphi {
namespace {

// state a lane needs besides the graph: single-source arrays (~42 B per
// vertex), batched masks, levels and bitmaps for max_batch sources, and the
// in-edge arrays of the batched pull steps (4 B per edge and per vertex, pieces)
uint64_t lane_state_bytes(const DevGraph &G, uint32_t max_batch)
{
	const uint64_t nw = max_batch <= 64 ? 1 : max_batch <= 128 ? 2 : 4;
	const uint64_t in_edges = max_batch > 1 ? 4 * G.m + 12 * (uint64_t)G.n_pad + G.m / 256 : 0;
	return (uint64_t)G.n_pad * (42 + 3 * 8 * nw + 64 * nw + 1) + G.m / 256 + in_edges;
}

// what an SSSP lane adds for multi-source SSSP: 16 distances per vertex, the
// in-edge arrays and their weights (1 byte each below 256, else 4)
uint64_t sssp_batch_bytes(const DevGraph &G)
{
	return (uint64_t)G.n_pad * 4 * GPHI_SSSP_BATCH_MAX + 4 * G.m + 12 * (uint64_t)G.n_pad + G.m / 256 +
	       (G.coo_w8 ? 1 : 4) * G.m;
}

// BFS levels of a finished slot into the query's int32 buffer (-1 unreached)
void store_levels(const std::vector<uint8_t> &row, uint32_t n, Query &q)
{
	if (!q.out) return;
	int32_t *o = static_cast<int32_t *>(q.out);
	for (uint32_t v = 0; v < n; ++v) o[v] = row[v] == 255 ? -1 : (int32_t)row[v];
}

void complete(gphi_server *S, const std::vector<std::shared_ptr<Query>> &qs, int rc, double t)
{
	{
		std::lock_guard<std::mutex> lk(S->mu);
		for (auto &q : qs) {
			q->t_done = t;
			q->status = rc;
			q->done = true;
		}
	}
	S->cv_done.notify_all();
}

// A multi-source traversal with admission between levels.  `take` holds the
// queries it starts with; arrivals join while slots are free.
void run_continuous(gphi_server *S, Lane *L, std::vector<std::shared_ptr<Query>> take)
{
	const uint32_t cap = std::min<uint32_t>(S->opts.max_batch, std::max<uint32_t>((uint32_t)take.size(), 64u));
	BatchRun run;
	std::vector<std::shared_ptr<Query>> slot_q;  // per slot
	std::vector<uint8_t> finished;               // per slot: answered
	std::vector<uint32_t> srcs;
	std::vector<uint8_t> row(L->n);
	uint64_t launch;
	{
		std::lock_guard<std::mutex> lk(S->mu);
		launch = S->next_launch++;
	}
	const double t0 = now_us();
	for (auto &q : take) srcs.push_back(q->src);
	int rc = run.begin(L->g, cap, srcs.data(), (uint32_t)srcs.size());
	for (auto &q : take) {
		q->t_start = t0;
		q->launch = launch;
		slot_q.push_back(q);
		finished.push_back(0);
	}
	while (rc == GPHI_OK && !run.done) {
		// arrivals join at this level boundary, up to the free slots
		std::vector<std::shared_ptr<Query>> join;
		{
			std::lock_guard<std::mutex> lk(S->mu);
			const uint32_t free_slots = run.cap - run.used;
			for (auto it = L->pending.begin(); it != L->pending.end() && join.size() < free_slots;) {
				if ((*it)->kind == GPHI_QUERY_BFS) {
					join.push_back(*it);
					it = L->pending.erase(it);
				} else {
					++it;
				}
			}
		}
		srcs.clear();
		const double tj = now_us();
		for (auto &q : join) {
			srcs.push_back(q->src);
			q->t_start = tj;
			q->launch = launch;
			q->joined_level = run.levels + 1;
			slot_q.push_back(q);
			finished.push_back(0);
		}
		rc = run.step(srcs.data(), (uint32_t)srcs.size());
		if (rc) break;
		// slots whose frontier emptied: copy their levels out and answer them
		for (uint32_t s = 0; s < run.used && rc == GPHI_OK; ++s) {
			if (finished[s] || ((run.live[s / 64] >> (s % 64)) & 1ull)) continue;
			if (slot_q[s]->out) {
				uint8_t *r = row.data();
				rc = run.copy_rows(&s, 1, &r);
				if (rc == GPHI_OK) store_levels(row, L->n, *slot_q[s]);
			}
			finished[s] = 1;
			slot_q[s]->batch = run.used;
			if (rc == GPHI_OK) complete(S, {slot_q[s]}, GPHI_OK, now_us());
		}
	}
	// the traversal ended (or failed): answer whatever is left
	std::vector<std::shared_ptr<Query>> rest;
	for (uint32_t s = 0; s < slot_q.size(); ++s) {
		if (finished[s]) continue;
		if (rc == GPHI_OK && slot_q[s]->out) {
			uint8_t *r = row.data();
			rc = run.copy_rows(&s, 1, &r);
			if (rc == GPHI_OK) store_levels(row, L->n, *slot_q[s]);
		}
		slot_q[s]->batch = (uint32_t)slot_q.size();
		rest.push_back(slot_q[s]);
	}
	const double t1 = now_us();
	if (!rest.empty()) complete(S, rest, rc, t1);
	if (rc == GPHI_OK && !slot_q.empty()) ema(L->bfs_runs.t[RunCosts::size_class(slot_q.size())], t1 - t0);
	std::lock_guard<std::mutex> lk(S->mu);
	L->batches += 1;
	L->served += slot_q.size();
}

void lane_loop(gphi_server *S, Lane *L)
{
	std::vector<std::shared_ptr<Query>> take;
	for (;;) {
		take.clear();
		{
			std::unique_lock<std::mutex> lk(S->mu);
			S->cv_work.wait(lk, [&] { return S->stopping || !L->pending.empty(); });
			if (L->pending.empty()) return;  // stopping and drained
			// the pending SSSP queries (up to GPHI_SSSP_BATCH_MAX) when they
			// are to share a traversal, else one alone; or every pending BFS
			// query (up to max_batch)
			if (L->pending.front()->kind == GPHI_QUERY_SSSP) {
				size_t k = 0;
				for (auto &q : L->pending) k += q->kind == GPHI_QUERY_SSSP;
				const size_t lim = use_sssp_batch(*L, std::min<size_t>(k, GPHI_SSSP_BATCH_MAX))
				                       ? GPHI_SSSP_BATCH_MAX : 1;
				for (auto it = L->pending.begin(); it != L->pending.end() && take.size() < lim;) {
					if ((*it)->kind == GPHI_QUERY_SSSP) {
						take.push_back(*it);
						it = L->pending.erase(it);
					} else {
						++it;
					}
				}
			} else {
				for (auto it = L->pending.begin(); it != L->pending.end() && take.size() < S->opts.max_batch;) {
					if ((*it)->kind == GPHI_QUERY_BFS) {
						take.push_back(*it);
						it = L->pending.erase(it);
					} else {
						++it;
					}
				}
			}
			L->busy = true;
		}
		if (take[0]->kind == GPHI_QUERY_SSSP) {
			uint64_t launch;
			{
				std::lock_guard<std::mutex> lk(S->mu);
				launch = S->next_launch++;
			}
			const double t0 = now_us();
			for (auto &q : take) {
				q->t_start = t0;
				q->launch = launch;
				q->batch = (uint32_t)take.size();
			}
			int rc;
			if (take.size() == 1) {
				rc = gphi_sssp(L->g, take[0]->src, (uint32_t *)take[0]->out, nullptr);
			} else {
				uint32_t srcs[GPHI_SSSP_BATCH_MAX];
				uint32_t *rows[GPHI_SSSP_BATCH_MAX];
				for (size_t i = 0; i < take.size(); ++i) {
					srcs[i] = take[i]->src;
					rows[i] = (uint32_t *)take[i]->out;
				}
				std::lock_guard<std::mutex> ctx_lk(L->ctx->mu);
				rc = hipSetDevice(L->ctx->device) == hipSuccess
				         ? run_sssp_batch(L->g, srcs, (uint32_t)take.size(), rows, nullptr)
				         : GPHI_E_HIP;
			}
			const double t1 = now_us();
			complete(S, take, rc, t1);
			std::lock_guard<std::mutex> lk(S->mu);
			if (rc == GPHI_OK) {
				if (take.size() == 1)
					ema(L->t_sssp, t1 - t0);
				else
					ema(L->sssp_runs.t[RunCosts::size_class(take.size())], t1 - t0);
			}
			L->batches += 1;
			L->served += take.size();
		} else if (use_batch(*L, take.size(), S->opts.max_batch)) {
			std::lock_guard<std::mutex> ctx_lk(L->ctx->mu);  // the traversal's kernels on this context
			if (hipSetDevice(L->ctx->device) != hipSuccess) {
				complete(S, take, GPHI_E_HIP, now_us());
			} else {
				run_continuous(S, L, take);
			}
		} else {
			for (auto &q : take) {
				uint64_t launch;
				{
					std::lock_guard<std::mutex> lk(S->mu);
					launch = S->next_launch++;
				}
				q->t_start = now_us();
				q->launch = launch;
				q->batch = 1;
				const int r = gphi_bfs(L->g, q->src, (int32_t *)q->out, nullptr, nullptr);
				const double t = now_us();
				ema(L->t_plain, t - q->t_start);
				complete(S, {q}, r, t);
				std::lock_guard<std::mutex> lk(S->mu);
				L->batches += 1;
				L->served += 1;
			}
		}
		{
			std::lock_guard<std::mutex> lk(S->mu);
			L->busy = false;
		}
	}
}

}  // namespace
}  // namespace gphi

using gphi::set_error;

extern "C" {

int gphi_server_create(const gphi_server_opts *opts, gphi_server **out)
{
	if (!out) return set_error(nullptr, GPHI_E_ARG, "out is NULL");
	*out = nullptr;
	gphi_server *S = new (std::nothrow) gphi_server();
	if (!S) return set_error(nullptr, GPHI_E_OOM, "out of host memory");
	if (opts) S->opts = *opts;
	if (S->opts.max_batch == 0) S->opts.max_batch = 64;
	if (S->opts.max_batch > 256) {
		delete S;
		return set_error(nullptr, GPHI_E_ARG, "max_batch must be <= 256");
	}
	if (S->opts.sssp_lanes == 0) S->opts.sssp_lanes = 2;
	if (S->opts.sssp_lanes > 8) {
		delete S;
		return set_error(nullptr, GPHI_E_ARG, "sssp_lanes must be <= 8");
	}
	if (S->opts.hbm_budget == 0) {
		size_t fr = 0, tot = 0;
		if (hipMemGetInfo(&fr, &tot) == hipSuccess) S->opts.hbm_budget = (uint64_t)(tot * 0.9);
	}
	*out = S;
	return GPHI_OK;
}

int gphi_server_add_graph(gphi_server *S, gphi_graph *g, uint32_t *graph_id)
{
	if (!S || !g || !graph_id) return set_error(nullptr, GPHI_E_ARG, "NULL argument");
	if (g->base) return set_error(nullptr, GPHI_E_ARG, "add the graph itself, not a view of it");
	const bool weighted = g->g.weighted;
	const uint32_t nv = weighted ? S->opts.sssp_lanes : 0;
	// SSSP queries share traversals when the server batches and the graph has
	// the CSR and tile arrays the multi-source SSSP reads
	const bool sssp_batch = weighted && S->opts.max_batch > 1 && g->g.has_csr && g->g.has_tiles;
	const uint64_t need = g->g.device_bytes + gphi::lane_state_bytes(g->g, S->opts.max_batch) +
	                      nv * (gphi::lane_state_bytes(g->g, 1) + (sssp_batch ? gphi::sssp_batch_bytes(g->g) : 0));
	{
		// the check and the reservation in one critical section, so two
		// concurrent admissions cannot both fit into the same free bytes
		std::lock_guard<std::mutex> lk(S->mu);
		if (S->hbm_used + need > S->opts.hbm_budget)
			return set_error(nullptr, GPHI_E_OOM,
			                 "graph needs %llu MB of HBM; %llu MB of the %llu MB budget are in use",
			                 (unsigned long long)(need >> 20), (unsigned long long)(S->hbm_used >> 20),
			                 (unsigned long long)(S->opts.hbm_budget >> 20));
		S->hbm_used += need;
	}
	auto unreserve = [&] {
		std::lock_guard<std::mutex> lk(S->mu);
		S->hbm_used -= need;
	};
	// the lane's state is built at admission, not by its first query: the
	// in-edge sort alone takes tens of milliseconds on the larger graphs, and
	// every query queued behind the first traversal would wait for it
	{
		std::lock_guard<std::mutex> ctx_lk(g->ctx->mu);
		int rc = hipSetDevice(g->ctx->device) == hipSuccess ? GPHI_OK : GPHI_E_HIP;
		if (rc == GPHI_OK)
			rc = S->opts.max_batch > 1 && g->g.has_csr && g->g.has_tiles
			         ? gphi::batch_prepare(g, S->opts.max_batch <= 64 ? 1 : S->opts.max_batch <= 128 ? 2 : 4)
			         : gphi::ensure_traversal_state(g);
		if (rc != GPHI_OK) {
			unreserve();
			return rc;
		}
	}
	std::vector<gphi_graph *> views;
	std::vector<gphi_ctx *> vctxs;
	for (uint32_t i = 0; i < nv; ++i) {
		gphi_opts o{g->ctx->device, g->ctx->t_ratio};
		gphi_ctx *vc = nullptr;
		gphi_graph *v = nullptr;
		int rc = gphi_open(&o, &vc);
		if (rc == GPHI_OK) rc = gphi_graph_view(vc, g, &v);
		if (rc != GPHI_OK) {
			if (vc) gphi_close(vc);
			for (gphi_graph *x : views) gphi_graph_free(x);
			for (gphi_ctx *x : vctxs) gphi_close(x);
			unreserve();
			return rc;
		}
		if (sssp_batch) {
			{
				// built now, not by the first batch of SSSP queries
				std::lock_guard<std::mutex> ctx_lk(vc->mu);
				rc = hipSetDevice(vc->device) == hipSuccess ? gphi::sssp_batch_prepare(v) : GPHI_E_HIP;
			}
			if (rc != GPHI_OK) {
				gphi_graph_free(v);
				gphi_close(vc);
				for (gphi_graph *x : views) gphi_graph_free(x);
				for (gphi_ctx *x : vctxs) gphi_close(x);
				unreserve();
				return rc;
			}
		}
		views.push_back(v);
		vctxs.push_back(vc);
	}
	std::lock_guard<std::mutex> lk(S->mu);
	auto L = std::make_unique<Lane>();
	L->g = g;
	L->ctx = g->ctx;
	L->n = g->g.n;
	L->device_bytes = need;
	Lane *bl = L.get();
	std::vector<Lane *> sl;
	S->lanes.push_back(std::move(L));
	for (uint32_t i = 0; i < nv; ++i) {
		auto V = std::make_unique<Lane>();
		V->g = views[i];
		V->ctx = vctxs[i];
		V->n = g->g.n;
		V->sssp_batch_ok = sssp_batch;
		sl.push_back(V.get());
		S->lanes.push_back(std::move(V));
	}
	*graph_id = (uint32_t)S->graphs.size();
	S->graphs.emplace_back(bl, sl);
	bl->worker = std::thread(gphi::lane_loop, S, bl);
	for (Lane *v : sl) v->worker = std::thread(gphi::lane_loop, S, v);
	return GPHI_OK;
}

int gphi_server_submit(gphi_server *S, uint32_t graph_id, uint32_t kind, uint32_t src, void *out,
                       uint64_t *ticket)
{
	if (!S || !ticket) return set_error(nullptr, GPHI_E_ARG, "NULL argument");
	auto q = std::make_shared<Query>();
	{
		std::lock_guard<std::mutex> lk(S->mu);
		if (graph_id >= S->graphs.size()) return set_error(nullptr, GPHI_E_RANGE, "no graph %u", graph_id);
		if (kind != GPHI_QUERY_BFS && kind != GPHI_QUERY_SSSP)
			return set_error(nullptr, GPHI_E_ARG, "unknown query kind %u", kind);
		Lane *L = S->graphs[graph_id].first;
		if (kind == GPHI_QUERY_SSSP) {
			const auto &sl = S->graphs[graph_id].second;
			if (sl.empty()) return set_error(nullptr, GPHI_E_STATE, "graph %u carries no weights", graph_id);
			L = sl[0];  // the SSSP lane with the shortest queue
			auto load = [](const Lane *x) { return x->pending.size() + (x->busy ? 1 : 0); };
			for (Lane *v : sl)
				if (load(v) < load(L)) L = v;
		}
		if (src >= L->n) return set_error(nullptr, GPHI_E_RANGE, "source %u >= N %u", src, L->n);
		if (S->stopping) return set_error(nullptr, GPHI_E_STATE, "server is stopping");
		q->ticket = S->next_ticket++;
		q->kind = kind;
		q->src = src;
		q->out = out;
		q->t_submit = now_us();
		S->queries[q->ticket] = q;
		L->pending.push_back(q);
		*ticket = q->ticket;
	}
	S->cv_work.notify_all();
	return GPHI_OK;
}

int gphi_server_wait(gphi_server *S, uint64_t ticket, gphi_query_info *info)
{
	if (!S) return set_error(nullptr, GPHI_E_ARG, "server is NULL");
	std::unique_lock<std::mutex> lk(S->mu);
	auto it = S->queries.find(ticket);
	if (it == S->queries.end()) return set_error(nullptr, GPHI_E_ARG, "unknown ticket %llu", (unsigned long long)ticket);
	std::shared_ptr<Query> q = it->second;
	// one waiter per ticket: the ticket leaves the table now, so a second
	// wait on it (concurrent or later) is an unknown ticket, not a use of an
	// iterator the first waiter has invalidated
	S->queries.erase(it);
	S->cv_done.wait(lk, [&] { return q->done; });
	if (info) {
		info->latency_us = q->t_done - q->t_submit;
		info->queue_us = q->t_start - q->t_submit;
		info->run_us = q->t_done - q->t_start;
		info->batch = q->batch;
		info->joined_level = q->joined_level;
		info->launch = q->launch;
	}
	return q->status;
}

int gphi_server_stats(gphi_server *S, uint32_t graph_id, uint64_t *batches, uint64_t *served,
                      uint64_t *hbm_bytes)
{
	if (!S) return set_error(nullptr, GPHI_E_ARG, "server is NULL");
	std::lock_guard<std::mutex> lk(S->mu);
	if (graph_id >= S->graphs.size()) return set_error(nullptr, GPHI_E_RANGE, "no graph %u", graph_id);
	const Lane *L = S->graphs[graph_id].first;
	uint64_t b = L->batches, sv = L->served;
	for (const Lane *v : S->graphs[graph_id].second) {
		b += v->batches;
		sv += v->served;
	}
	if (batches) *batches = b;
	if (served) *served = sv;
	if (hbm_bytes) *hbm_bytes = L->device_bytes;
	return GPHI_OK;
}

void gphi_server_destroy(gphi_server *S)
{
	if (!S) return;
	{
		std::lock_guard<std::mutex> lk(S->mu);
		S->stopping = true;
	}
	S->cv_work.notify_all();
	for (auto &L : S->lanes)
		if (L->worker.joinable()) L->worker.join();
	// views first (they share their graph's arrays), then the graphs; every
	// context once
	std::vector<gphi_graph *> views, bases;
	std::vector<gphi_ctx *> ctxs;
	for (auto &L : S->lanes) {
		(L->g->base ? views : bases).push_back(L->g);
		if (std::find(ctxs.begin(), ctxs.end(), L->ctx) == ctxs.end()) ctxs.push_back(L->ctx);
	}
	for (gphi_graph *v : views) gphi_graph_free(v);
	for (gphi_graph *b : bases) gphi_graph_free(b);
	for (gphi_ctx *c : ctxs) gphi_close(c);
	delete S;
}

}  // extern "C"

// sssp -- drop-in for apps/sssp_simd's ./sssp <data> <tile_width> <stripe_length> [-w|--weighted]
//
// -w: reads <data>_weighted_reorder_col-S-coo-tiled-W-* and
// <data>_weighted_reorder_untiled-* (apps/sssp_simd/sssp.cpp:41-189) and runs
// the weighted hybrid SSSP from vertex 0 (:969-1112); prints tile_size /
// stripe_length (:1472-1473) and one "<gpus> <seconds>" line per run.
// Without -w the reference's unweighted path reads <data>_coo-tiled-W-*
// (:1125-1127) and computes hop distances from vertex 0 (unreached -1,
// :1294-1295); here that is the tile hop relaxation to its fixpoint.
// Extra options: -s <source>, -r <runs>, -d <device>, -o <file>.
#include <climits>
#include <vector>

#include "common.h"

int main(int argc, char **argv)
{
	cli::Args a;
	a.runs = 1;
	if (!cli::parse(argc, argv, a)) {
		printf("Usage: ./sssp <data> <tile_width> <stripe_length> [-w|--weighted]\n");
		exit(1);
	}
	gphi_ctx *ctx = nullptr;
	gphi_opts o{a.device, 20};
	int rc = gphi_open(&o, &ctx);
	if (rc) return cli::die(nullptr, "gphi_open", rc);
	printf("tile_size: %u\n", a.W);
	printf("stripe_length: %u\n", a.S);
	gphi_graph *g = nullptr;
	if (a.weighted) {
		const std::string prefix = std::string(a.data) + "_weighted_reorder";
		rc = gphi_graph_load_text(ctx, prefix.c_str(), a.W, a.S, 1, &g);
	} else {
		rc = gphi_graph_load_text_rowmajor(ctx, a.data, a.W, a.S, &g);
	}
	if (rc) return cli::die(ctx, "loading input", rc);
	uint32_t n = 0;
	gphi_graph_info(g, &n, nullptr, nullptr, nullptr, nullptr);
	std::vector<uint32_t> d(a.out ? n : 0);
	cli::Timing tm;
	for (int r = 0; r < a.runs; ++r) {
		uint32_t *outp = (r + 1 == a.runs && a.out) ? d.data() : nullptr;
		if (a.weighted) {
			gphi_stats st;
			if ((rc = gphi_sssp(g, a.source, outp, &st))) return cli::die(ctx, "gphi_sssp", rc);
			tm.record(st.total_ms * 1e-3);
		} else {
			float ms = 0;
			uint32_t sweeps = 0;
			// hop distances reach their fixpoint within n sweeps
			if ((rc = gphi_bellman_ford(g, a.source, n + 1, outp, &sweeps, &ms)))
				return cli::die(ctx, "gphi_bellman_ford", rc);
			tm.record(ms * 1e-3);
		}
	}
	tm.average();
	int bad = 0;
	if (a.out) {
		if (a.weighted) {
			bad = cli::dump(a.out, d.data(), n, "%u\n");
		} else {
			std::vector<int32_t> hops(n);
			for (uint32_t v = 0; v < n; ++v) hops[v] = d[v] == (uint32_t)(INT_MAX - 2) ? -1 : (int32_t)d[v];
			bad = cli::dump(a.out, hops.data(), n, "%d\n");
		}
	}
	gphi_graph_free(g);
	gphi_close(ctx);
	return bad;
}

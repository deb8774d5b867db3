// bfs -- drop-in for apps/bfs_simd's ./bfs <data> <tile_width> <stripe_length> [-w]
//
// Reads <data>_reorder_col-S-coo-tiled-W-* and <data>_reorder_untiled-*
// (apps/bfs_simd/bfs.cpp:983-1113), runs the hybrid BFS from vertex 0
// (:1159) on the GPU `runs` times and prints one "<gpus> <seconds>" line per
// run and "Average: <gpus> <seconds>" (:921, :1198).  The time is the device
// time of the whole traversal including its state initialisation.
// -w selects the weighted inputs <data>_weighted_reorder_* and runs SSSP, as
// sssp_simd -w does (apps/sssp_simd/sssp.cpp:1397-1418).
// Extra options: -s <source>, -r <runs>, -d <device>, -o <file> (levels or
// distances, one per line).
#include <vector>

#include "common.h"

int main(int argc, char **argv)
{
	cli::Args a;
	if (!cli::parse(argc, argv, a)) {
		printf("Usage: ./bfs <data> <tile_width> <stripe_length>\n");
		exit(1);
	}
	gphi_ctx *ctx = nullptr;
	gphi_opts o{a.device, 20};
	int rc = gphi_open(&o, &ctx);
	if (rc) return cli::die(nullptr, "gphi_open", rc);
	const std::string prefix = std::string(a.data) + (a.weighted ? "_weighted_reorder" : "_reorder");
	gphi_graph *g = nullptr;
	if ((rc = gphi_graph_load_text(ctx, prefix.c_str(), a.W, a.S, a.weighted, &g)))
		return cli::die(ctx, "loading input", rc);
	uint32_t n = 0;
	gphi_graph_info(g, &n, nullptr, nullptr, nullptr, nullptr);
	if (a.weighted) {
		printf("tile_size: %u\n", a.W);
		printf("stripe_length: %u\n", a.S);
	}
	std::vector<int32_t> lv(a.out ? n : 0);
	gphi_stats st;
	cli::Timing tm;
	for (int r = 0; r < a.runs; ++r) {
		const bool last = r + 1 == a.runs;
		if (a.weighted)
			rc = gphi_sssp(g, a.source, last && a.out ? (uint32_t *)lv.data() : nullptr, &st);
		else
			rc = gphi_bfs(g, a.source, last && a.out ? lv.data() : nullptr, nullptr, &st);
		if (rc) return cli::die(ctx, a.weighted ? "gphi_sssp" : "gphi_bfs", rc);
		tm.record(st.total_ms * 1e-3);
	}
	tm.average();
	int bad = 0;
	if (a.out)
		bad = a.weighted ? cli::dump(a.out, (const uint32_t *)lv.data(), n, "%u\n")
		                 : cli::dump(a.out, lv.data(), n, "%d\n");
	gphi_graph_free(g);
	gphi_close(ctx);
	return bad;
}

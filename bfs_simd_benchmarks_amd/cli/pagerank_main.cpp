// page_rank -- drop-in for apps/pageRank_simd's ./page_rank <data> <tile_width> <stripe_length>
//
// Reads <data>_coo-tiled-W-* (row-major tiles, -offsets, -nneibor;
// apps/pageRank_simd/page_rank.cpp:468-568).  Each timed run is one sweep from
// rank = 1 (:583-587): sum[t] += rank[h]/deg[h], then rank = 0.15/N + 0.85*sum
// (:428-446), FP32; one "<gpus> <seconds>" line per run and the average.
// The reference exits with status 3 when SIDE < S (:374-378); so does this.
// Extra options: -r <runs>, -d <device>, -o <file> (ranks after the last run),
// -i <iterations per run>.
#include <cstring>
#include <vector>

#include "common.h"

int main(int argc, char **argv)
{
	uint32_t iters = 1;
	std::vector<char *> rest;
	for (int i = 0; i < argc; ++i) {
		if ((!strcmp(argv[i], "-i") || !strcmp(argv[i], "--iterations")) && i + 1 < argc)
			iters = (uint32_t)strtoul(argv[++i], nullptr, 0);
		else
			rest.push_back(argv[i]);
	}
	cli::Args a;
	if (!cli::parse((int)rest.size(), rest.data(), a)) {
		printf("Usage: ./pageRank <data> <tile_width> <stripe_length>\n");
		exit(1);
	}
	gphi_ctx *ctx = nullptr;
	gphi_opts o{a.device, 20};
	int rc = gphi_open(&o, &ctx);
	if (rc) return cli::die(nullptr, "gphi_open", rc);
	gphi_graph *g = nullptr;
	if ((rc = gphi_graph_load_text_rowmajor(ctx, a.data, a.W, a.S, &g))) return cli::die(ctx, "loading input", rc);
	uint32_t n = 0;
	gphi_graph_info(g, &n, nullptr, nullptr, nullptr, nullptr);
	const uint32_t side = n % a.W ? n / a.W + 1 : n / a.W;
	if (side < a.S) {
		printf("Error: ROW_STEP (%u) is to large, larger than side_length (%u)\n", a.S, side);
		exit(3);
	}
	std::vector<float> rank(a.out ? n : 0);
	cli::Timing tm;
	for (int r = 0; r < a.runs; ++r) {
		float ms = 0;
		float *outp = (r + 1 == a.runs && a.out) ? rank.data() : nullptr;
		if ((rc = gphi_pagerank(g, iters, 0.85f, 1.0f, outp, &ms))) return cli::die(ctx, "gphi_pagerank", rc);
		tm.record(ms * 1e-3);
	}
	tm.average();
	int bad = a.out ? cli::dump(a.out, rank.data(), n, "%.9g\n") : 0;
	gphi_graph_free(g);
	gphi_close(ctx);
	return bad;
}

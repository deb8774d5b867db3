// bellman-ford -- drop-in for apps/bellman-ford_simd's ./bellman-ford <data> <tile_width> <stripe_length>
//
// Reads <data>_coo-tiled-W-* (apps/bellman-ford_simd/bellman-ford.cpp:400-495),
// relaxes dist[t] = min(dist[t], dist[h]+1) from vertex 0 with dist = INT_MAX-2
// elsewhere (:308-312) for at most 99 sweeps (:322), stopping at the fixpoint;
// prints "<gpus> <seconds>" (:380) and "i dist[i]" for i < 20 (:384-386).
// Exits with status 3 when SIDE < S (:296-300).  Extra: -d <device>, -o <file>.
#include <vector>

#include "common.h"

int main(int argc, char **argv)
{
	cli::Args a;
	if (!cli::parse(argc, argv, a)) {
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
	std::vector<uint32_t> d(n);
	float ms = 0;
	uint32_t sweeps = 0;
	if ((rc = gphi_bellman_ford(g, a.source, 99, d.data(), &sweeps, &ms)))
		return cli::die(ctx, "gphi_bellman_ford", rc);
	printf("%u %lf\n", 1u, ms * 1e-3);
	for (uint32_t i = 0; i < 20 && i < n; ++i) printf("%u %u\n", i, d[i]);
	int bad = a.out ? cli::dump(a.out, d.data(), n, "%u\n") : 0;
	gphi_graph_free(g);
	gphi_close(ctx);
	return bad;
}

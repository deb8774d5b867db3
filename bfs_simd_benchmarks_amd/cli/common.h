// common.h -- shared parts of the drop-in command lines (argv parsing in the
// reference's positional form, the BestPerform output lines, result dumps).
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gphi.h"

namespace cli {

struct Args {
	const char *data = nullptr;
	uint32_t W = 0, S = 0;
	bool weighted = false;
	uint32_t source = 0;
	int runs = 10;
	int device = 0;
	const char *out = nullptr;  // optional dump of the results, one value per line
};

// positional <data> <tile_width> <stripe_length> (strtoul base 0 as the
// reference, apps/bfs_simd/bfs.cpp:1126-1137), then options in any order
inline bool parse(int argc, char **argv, Args &a)
{
	std::vector<const char *> pos;
	for (int i = 1; i < argc; ++i) {
		const char *s = argv[i];
		if (!strcmp(s, "-w") || !strcmp(s, "--weighted")) {
			a.weighted = true;
		} else if ((!strcmp(s, "-s") || !strcmp(s, "--source")) && i + 1 < argc) {
			a.source = (uint32_t)strtoul(argv[++i], nullptr, 0);
		} else if ((!strcmp(s, "-r") || !strcmp(s, "--runs")) && i + 1 < argc) {
			a.runs = atoi(argv[++i]);
		} else if ((!strcmp(s, "-d") || !strcmp(s, "--device")) && i + 1 < argc) {
			a.device = atoi(argv[++i]);
		} else if ((!strcmp(s, "-o") || !strcmp(s, "--output")) && i + 1 < argc) {
			a.out = argv[++i];
		} else {
			pos.push_back(s);
		}
	}
	if (pos.size() < 3) return false;
	a.data = pos[0];
	a.W = (uint32_t)strtoul(pos[1], nullptr, 0);
	a.S = (uint32_t)strtoul(pos[2], nullptr, 0);
	if (a.runs < 1) a.runs = 1;
	return true;
}

inline int die(gphi_ctx *ctx, const char *what, int rc)
{
	fprintf(stderr, "%s failed (%d): %s\n", what, rc, gphi_last_error(ctx));
	return 1;
}

// "Average: <metric> <sec>" as BestPerform::print_average (include/peg_util.h:48-56);
// the metric column holds the GPU count where the reference prints threads
struct Timing {
	double total = 0;
	int count = 0;
	void record(double sec)
	{
		printf("%u %lf\n", 1u, sec);
		total += sec;
		++count;
	}
	void average() const { printf("Average: %u %f\n", 1u, count ? total / count : 0.0); }
};

template <class T>
inline int dump(const char *path, const T *v, uint32_t n, const char *fmt)
{
	FILE *f = fopen(path, "w");
	if (!f) {
		fprintf(stderr, "cannot create %s\n", path);
		return 1;
	}
	for (uint32_t i = 0; i < n; ++i) fprintf(f, fmt, v[i]);
	return fclose(f) ? 1 : 0;
}

}  // namespace cli

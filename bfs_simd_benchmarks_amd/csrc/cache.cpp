// cache.cpp -- binary tiled-graph cache.
//
// The reference rebuilds its inputs with the offline tools and parses 64
// text parts on every run (apps/bfs_simd/bfs.cpp:968-1118); tile-size sweeps
// re-read them once per (W, S) (tools/tile_size/process.sh:19-21).  A cache
// file holds one tiled graph as raw little-endian uint32 arrays behind a
// 64-byte header, so a load is a few large reads plus the upload.
//
//   offset 0   char[8]  magic "GPHIBIN1"
//          8   u32      version (1)
//         12   u32      N
//         16   u64      M
//         24   u32      W, S
//         32   i32      tile order (GPHI_TILES_*)
//         36   u32      flags: 1 = CSR present, 2 = tiles present, 4 = weighted
//         40   u64      sum of every uint32 word of the arrays, mod 2^64
//         48   u64[2]   zero
//         64   degrees[N], then CSR tails[M] (+ weights[M]) if flag 1,
//              then tile heads[M], tails[M] (+ weights[M]) if flag 2
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.h"

namespace {

constexpr char kMagic[8] = {'G', 'P', 'H', 'I', 'B', 'I', 'N', '1'};
constexpr uint32_t kVersion = 1;

struct Header {
	char magic[8];
	uint32_t version;
	uint32_t n;
	uint64_t m;
	uint32_t W, S;
	int32_t tile_order;
	uint32_t flags;
	uint64_t checksum;
	uint64_t zero[2];
};
static_assert(sizeof(Header) == 64, "cache header is 64 bytes");

uint64_t word_sum(const std::vector<uint32_t> &a)
{
	uint64_t s = 0;
	const size_t n = a.size();
#pragma omp parallel for reduction(+ : s)
	for (size_t i = 0; i < n; ++i) s += a[i];
	return s;
}

struct File {
	FILE *f = nullptr;
	~File()
	{
		if (f) fclose(f);
	}
};

bool write_all(FILE *f, const void *p, size_t bytes)
{
	const char *c = (const char *)p;
	while (bytes) {
		const size_t chunk = bytes < (1u << 30) ? bytes : (1u << 30);
		if (fwrite(c, 1, chunk, f) != chunk) return false;
		c += chunk;
		bytes -= chunk;
	}
	return true;
}

bool read_all(FILE *f, void *p, size_t bytes)
{
	char *c = (char *)p;
	while (bytes) {
		const size_t chunk = bytes < (1u << 30) ? bytes : (1u << 30);
		if (fread(c, 1, chunk, f) != chunk) return false;
		c += chunk;
		bytes -= chunk;
	}
	return true;
}

}  // namespace

using gphi::set_error;

extern "C" {

int gphi_graph_save(gphi_graph *g, const char *path)
{
	if (!g || !path) return set_error(nullptr, GPHI_E_ARG, "gphi_graph_save: NULL argument");
	const gphi::DevGraph &G = g->g;
	gphi_ctx *ctx = g->ctx;
	const uint64_t m = G.m;
	// weights are stored only when every present edge array carries them
	const bool w = G.weighted && (!G.has_csr || G.csr_w) && (!G.has_tiles || G.coo_w);
	std::vector<uint32_t> a[6];  // deg, cols, csr_w, heads, tails, tile_w
	try {
		a[0].resize(G.n);
		if (G.has_csr) {
			a[1].resize(m);
			if (w) a[2].resize(m);
		}
		if (G.has_tiles) {
			a[3].resize(m);
			a[4].resize(m);
			if (w) a[5].resize(m);
		}
	} catch (const std::bad_alloc &) {
		return set_error(ctx, GPHI_E_OOM, "host buffers for %llu edges", (unsigned long long)m);
	}
	auto ptr = [&](int i) { return a[i].empty() ? nullptr : a[i].data(); };
	int rc = gphi_graph_download(g, ptr(0), ptr(1), ptr(2), ptr(3), ptr(4), ptr(5));
	if (rc) return rc;
	Header h;
	memset(&h, 0, sizeof(h));
	memcpy(h.magic, kMagic, 8);
	h.version = kVersion;
	h.n = G.n;
	h.m = m;
	h.W = G.W;
	h.S = G.S;
	h.tile_order = G.tile_order;
	h.flags = (G.has_csr ? 1u : 0u) | (G.has_tiles ? 2u : 0u) | (w ? 4u : 0u);
	for (auto &v : a) h.checksum += word_sum(v);
	File f;
	f.f = fopen(path, "wb");
	if (!f.f) return set_error(ctx, GPHI_E_IO, "cannot create %s: %s", path, strerror(errno));
	bool ok = write_all(f.f, &h, sizeof(h));
	for (auto &v : a)
		if (ok && !v.empty()) ok = write_all(f.f, v.data(), v.size() * 4);
	if (ok) ok = fclose(f.f) == 0;
	f.f = nullptr;
	if (!ok) return set_error(ctx, GPHI_E_IO, "short write to %s", path);
	return GPHI_OK;
}

int gphi_graph_load(gphi_ctx *ctx, const char *path, gphi_graph **out)
{
	if (!ctx || !path || !out) return set_error(ctx, GPHI_E_ARG, "gphi_graph_load: NULL argument");
	*out = nullptr;
	File f;
	f.f = fopen(path, "rb");
	if (!f.f) return set_error(ctx, GPHI_E_IO, "cannot open %s: %s", path, strerror(errno));
	Header h;
	if (!read_all(f.f, &h, sizeof(h))) return set_error(ctx, GPHI_E_FORMAT, "%s: no header", path);
	if (memcmp(h.magic, kMagic, 8) != 0) return set_error(ctx, GPHI_E_FORMAT, "%s: not a graph cache", path);
	if (h.version != kVersion)
		return set_error(ctx, GPHI_E_FORMAT, "%s: cache version %u, expected %u", path, h.version, kVersion);
	if ((h.flags & ~7u) || !(h.flags & 3u) || h.zero[0] || h.zero[1])
		return set_error(ctx, GPHI_E_FORMAT, "%s: bad header flags", path);
	const bool csr = h.flags & 1u, tiles = h.flags & 2u, w = h.flags & 4u;
	const uint64_t m = h.m;
	if (m >= (1ull << 32)) return set_error(ctx, GPHI_E_FORMAT, "%s: M=%llu too large", path, (unsigned long long)m);
	// the file size must match the header before anything is allocated
	const uint64_t words = (uint64_t)h.n + (csr ? m * (w ? 2 : 1) : 0) + (tiles ? m * (w ? 3 : 2) : 0);
	if (fseeko(f.f, 0, SEEK_END) != 0) return set_error(ctx, GPHI_E_IO, "%s: seek failed", path);
	const uint64_t size = (uint64_t)ftello(f.f);
	if (size != sizeof(Header) + words * 4)
		return set_error(ctx, GPHI_E_FORMAT, "%s: %llu bytes, header implies %llu", path,
		                 (unsigned long long)size, (unsigned long long)(sizeof(Header) + words * 4));
	if (fseeko(f.f, sizeof(Header), SEEK_SET) != 0) return set_error(ctx, GPHI_E_IO, "%s: seek failed", path);
	std::vector<uint32_t> a[6];
	try {
		a[0].resize(h.n);
		if (csr) {
			a[1].resize(m);
			if (w) a[2].resize(m);
		}
		if (tiles) {
			a[3].resize(m);
			a[4].resize(m);
			if (w) a[5].resize(m);
		}
	} catch (const std::bad_alloc &) {
		return set_error(ctx, GPHI_E_OOM, "host buffers for %llu edges", (unsigned long long)m);
	}
	uint64_t sum = 0;
	for (auto &v : a) {
		if (!v.empty() && !read_all(f.f, v.data(), v.size() * 4))
			return set_error(ctx, GPHI_E_FORMAT, "%s: short read", path);
		sum += word_sum(v);
	}
	if (sum != h.checksum) return set_error(ctx, GPHI_E_FORMAT, "%s: checksum mismatch", path);
	gphi_graph_desc d;
	memset(&d, 0, sizeof(d));
	d.n = h.n;
	d.m = m;
	d.tile_width = h.W;
	d.stripe_length = h.S;
	d.tile_order = h.tile_order;
	auto ptr = [&](int i) { return a[i].empty() ? nullptr : a[i].data(); };
	d.degrees = ptr(0);
	d.csr_edges = ptr(1);
	d.csr_weights = ptr(2);
	d.tile_heads = ptr(3);
	d.tile_tails = ptr(4);
	d.tile_weights = ptr(5);
	return gphi_graph_from_desc(ctx, &d, out);
}

}  // extern "C"

// formats.cpp -- readers of the reference's 64-way text formats.
//
//   <prefix>_untiled-{0..63}, -nneibor        tools/csr_tiling/page_rank.cpp:570-598
//   <prefix>_coo-tiled-W-{0..63}, -offsets    tools/coo_tiling/page_rank.cpp:450-499
//   <prefix>_col-S-coo-tiled-W-{0..63}, ...   tools/column_major_tile/kcore.cpp:419-450
// Part 0 starts with "N M"; part k < 63 holds edges [k*floor(M/64),
// (k+1)*floor(M/64)), part 63 the rest; ids are 1-based ("h t" or "h t w").
// Unlike the reference readers (apps/bfs_simd/bfs.cpp:968-1118, which never
// check fscanf) every count and id is validated.
#include <algorithm>
#include <cerrno>
#include <new>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace gphi {

namespace {

constexpr uint32_t kParts = 64;

struct Text {
	std::vector<char> buf;
	size_t pos = 0;
	bool next(uint64_t &x)
	{
		const size_t n = buf.size();
		while (pos < n && (buf[pos] == ' ' || buf[pos] == '\n' || buf[pos] == '\r' || buf[pos] == '\t'))
			++pos;
		if (pos >= n || buf[pos] < '0' || buf[pos] > '9') return false;
		uint64_t v = 0;
		while (pos < n && buf[pos] >= '0' && buf[pos] <= '9') {
			v = v * 10 + (uint64_t)(buf[pos] - '0');
			if (v > 0xFFFFFFFFull) return false;
			++pos;
		}
		x = v;
		return true;
	}
};

int slurp(gphi_ctx *ctx, const std::string &path, Text &t)
{
	FILE *f = fopen(path.c_str(), "rb");
	if (!f) return set_error(ctx, GPHI_E_IO, "cannot open file: %s (%s)", path.c_str(), strerror(errno));
	if (fseek(f, 0, SEEK_END) != 0) {
		fclose(f);
		return set_error(ctx, GPHI_E_IO, "cannot seek: %s", path.c_str());
	}
	long sz = ftell(f);
	rewind(f);
	t.buf.resize(sz > 0 ? (size_t)sz : 0);
	size_t got = sz > 0 ? fread(t.buf.data(), 1, (size_t)sz, f) : 0;
	fclose(f);
	if ((long)got != sz) return set_error(ctx, GPHI_E_IO, "short read: %s", path.c_str());
	t.pos = 0;
	return GPHI_OK;
}

int read_header(gphi_ctx *ctx, const std::string &part0, uint32_t &n, uint64_t &m)
{
	Text t;
	int rc = slurp(ctx, part0, t);
	if (rc) return rc;
	uint64_t a, b;
	if (!t.next(a) || !t.next(b))
		return set_error(ctx, GPHI_E_FORMAT, "missing \"N M\" header in %s", part0.c_str());
	n = (uint32_t)a;
	m = b;
	return GPHI_OK;
}

// Reads the 64 edge parts into h/t(/w) (0-based).  heads may be null.
int read_parts(gphi_ctx *ctx, const std::string &prefix, uint32_t n, uint64_t m, bool weighted,
               uint32_t *heads, uint32_t *tails, uint32_t *wts)
{
	const uint64_t eb = m / kParts;
	int rc_all = GPHI_OK;
	std::vector<std::string> errs(kParts);
	std::vector<int> codes(kParts, GPHI_OK);
#pragma omp parallel for schedule(dynamic, 1)
	for (uint32_t k = 0; k < kParts; ++k) {
		std::string path = prefix + "-" + std::to_string(k);
		Text t;
		char msg[512];
		FILE *f = fopen(path.c_str(), "rb");
		if (!f) {
			codes[k] = GPHI_E_IO;
			snprintf(msg, sizeof msg, "cannot open file: %s", path.c_str());
			errs[k] = msg;
			continue;
		}
		fseek(f, 0, SEEK_END);
		long sz = ftell(f);
		rewind(f);
		t.buf.resize(sz > 0 ? (size_t)sz : 0);
		size_t got = sz > 0 ? fread(t.buf.data(), 1, (size_t)sz, f) : 0;
		fclose(f);
		if ((long)got != sz) {
			codes[k] = GPHI_E_IO;
			errs[k] = "short read: " + path;
			continue;
		}
		uint64_t a = 0, b = 0, c = 0;
		if (k == 0) {
			if (!t.next(a) || !t.next(b) || a != n || b != m) {
				codes[k] = GPHI_E_FORMAT;
				errs[k] = "header mismatch in " + path;
				continue;
			}
		}
		uint64_t lo = (uint64_t)k * eb, hi = (k + 1 < kParts) ? lo + eb : m;
		for (uint64_t e = lo; e < hi; ++e) {
			if (!t.next(a) || !t.next(b) || (weighted && !t.next(c))) {
				codes[k] = GPHI_E_FORMAT;
				snprintf(msg, sizeof msg, "%s: truncated at edge %llu (expected edges [%llu,%llu))",
				         path.c_str(), (unsigned long long)e, (unsigned long long)lo,
				         (unsigned long long)hi);
				errs[k] = msg;
				break;
			}
			if (a < 1 || a > n || b < 1 || b > n) {
				codes[k] = GPHI_E_RANGE;
				snprintf(msg, sizeof msg, "%s: edge %llu has id out of [1,%u]", path.c_str(),
				         (unsigned long long)e, n);
				errs[k] = msg;
				break;
			}
			if (heads) heads[e] = (uint32_t)(a - 1);
			tails[e] = (uint32_t)(b - 1);
			if (weighted) wts[e] = (uint32_t)c;
		}
	}
	for (uint32_t k = 0; k < kParts; ++k)
		if (codes[k] != GPHI_OK) {
			rc_all = set_error(ctx, codes[k], "%s", errs[k].c_str());
			break;
		}
	return rc_all;
}

int read_u32_lines(gphi_ctx *ctx, const std::string &path, uint64_t count, uint32_t *out)
{
	Text t;
	int rc = slurp(ctx, path, t);
	if (rc) return rc;
	for (uint64_t i = 0; i < count; ++i) {
		uint64_t x;
		if (!t.next(x))
			return set_error(ctx, GPHI_E_FORMAT, "%s: expected %llu values, got %llu", path.c_str(),
			                 (unsigned long long)count, (unsigned long long)i);
		out[i] = (uint32_t)x;
	}
	return GPHI_OK;
}

}  // namespace

// store[0] degrees, [1] csr tails, [2] csr weights, [3] tile heads,
// [4] tile tails, [5] tile weights, [6] tile offsets
int read_text_graph(gphi_ctx *ctx, const char *prefix_c, uint32_t W, uint32_t S, int weighted,
                    bool row_major, gphi_graph_desc *d, std::vector<uint32_t> *store)
{
	const std::string prefix(prefix_c);
	const std::string tiled = row_major ? prefix + "_coo-tiled-" + std::to_string(W)
	                                    : prefix + "_col-" + std::to_string(S) + "-coo-tiled-" +
	                                          std::to_string(W);
	const std::string untiled = prefix + "_untiled";
	uint32_t n = 0;
	uint64_t m = 0;
	int rc = read_header(ctx, tiled + "-0", n, m);
	if (rc) return rc;
	if (n == 0) return set_error(ctx, GPHI_E_FORMAT, "%s-0: N is 0", tiled.c_str());
	const uint64_t side = n % W ? n / W + 1 : n / W;
	if (!row_major && S > side)
		return set_error(ctx, GPHI_E_RANGE, "stripe length %u exceeds SIDE %llu", S,
		                 (unsigned long long)side);
	const uint64_t ntiles = side * side;
	memset(d, 0, sizeof(*d));
	d->n = n;
	d->m = m;
	d->tile_width = W;
	d->stripe_length = S;
	d->tile_order = row_major ? GPHI_TILES_ROW_MAJOR : GPHI_TILES_COL_STRIPE;
	try {
		store[0].resize(n);
		store[3].resize(m);
		store[4].resize(m);
		if (weighted) store[5].resize(m);
		store[6].resize(ntiles);
		if (!row_major) {
			store[1].resize(m);
			if (weighted) store[2].resize(m);
		}
	} catch (const std::bad_alloc &) {
		return set_error(ctx, GPHI_E_OOM, "host allocation for %llu edges failed", (unsigned long long)m);
	}
	if ((rc = read_u32_lines(ctx, tiled + "-offsets", ntiles, store[6].data()))) return rc;
	for (uint64_t i = 0; i < ntiles; ++i)
		if (store[6][i] > m || (i && store[6][i] < store[6][i - 1]))
			return set_error(ctx, GPHI_E_FORMAT, "%s-offsets: offset %llu not monotone within M",
			                 tiled.c_str(), (unsigned long long)i);
	if ((rc = read_parts(ctx, tiled, n, m, weighted, store[3].data(), store[4].data(),
	                     weighted ? store[5].data() : nullptr)))
		return rc;
	if (row_major) {
		if ((rc = read_u32_lines(ctx, tiled + "-nneibor", n, store[0].data()))) return rc;
	} else {
		uint32_t n2 = 0;
		uint64_t m2 = 0;
		if ((rc = read_header(ctx, untiled + "-0", n2, m2))) return rc;
		if (n2 != n || m2 != m)
			return set_error(ctx, GPHI_E_FORMAT, "%s-0 header (%u %llu) differs from tiles (%u %llu)",
			                 untiled.c_str(), n2, (unsigned long long)m2, n, (unsigned long long)m);
		if ((rc = read_u32_lines(ctx, untiled + "-nneibor", n, store[0].data()))) return rc;
		if ((rc = read_parts(ctx, untiled, n, m, weighted, nullptr, store[1].data(),
		                     weighted ? store[2].data() : nullptr)))
			return rc;
		d->csr_edges = store[1].data();
		d->csr_weights = weighted ? store[2].data() : nullptr;
	}
	d->degrees = store[0].data();
	d->tile_heads = store[3].data();
	d->tile_tails = store[4].data();
	d->tile_weights = weighted ? store[5].data() : nullptr;
	d->tile_offsets = store[6].data();
	return GPHI_OK;
}

// ---------------------------------------------------------------------------
// writers: the exact bytes of the reference tools' fprintf("%u %u\n") lines
// ---------------------------------------------------------------------------
namespace {

inline void put_u32(std::string &o, uint32_t x)
{
	char b[10];
	int k = 0;
	do {
		b[k++] = (char)('0' + x % 10);
		x /= 10;
	} while (x);
	while (k) o.push_back(b[--k]);
}

int write_file(const std::string &path, const std::string &data)
{
	FILE *f = fopen(path.c_str(), "wb");
	if (!f) return set_error(nullptr, GPHI_E_IO, "cannot create %s (%s)", path.c_str(), strerror(errno));
	size_t wr = fwrite(data.data(), 1, data.size(), f);
	int rc = fclose(f);
	if (wr != data.size() || rc != 0) return set_error(nullptr, GPHI_E_IO, "short write to %s", path.c_str());
	return GPHI_OK;
}

// heads==null: heads are implied by the CSR degrees (the _untiled order)
int write_parts(const std::string &prefix, uint32_t n, uint64_t m, const uint32_t *deg,
                const uint32_t *heads, const uint32_t *tails, const uint32_t *wts)
{
	const uint64_t eb = m / kParts;
	std::vector<uint64_t> row;  // CSR row starts, for implied heads
	if (!heads) {
		row.resize((size_t)n + 1);
		row[0] = 0;
		for (uint32_t v = 0; v < n; ++v) row[v + 1] = row[v] + deg[v];
	}
	std::vector<int> codes(kParts, GPHI_OK);
#pragma omp parallel for schedule(dynamic, 1)
	for (uint32_t k = 0; k < kParts; ++k) {
		const uint64_t lo = (uint64_t)k * eb, hi = (k + 1 < kParts) ? lo + eb : m;
		std::string o;
		o.reserve((size_t)(hi - lo) * (wts ? 24 : 16) + 32);
		if (k == 0) {
			put_u32(o, n);
			o.push_back(' ');
			put_u32(o, (uint32_t)m);
			o.push_back('\n');
		}
		uint32_t hv = 0;
		if (!heads && hi > lo) hv = (uint32_t)(std::upper_bound(row.begin(), row.end(), lo) - row.begin() - 1);
		for (uint64_t e = lo; e < hi; ++e) {
			uint32_t h;
			if (heads) {
				h = heads[e];
			} else {
				while (row[hv + 1] <= e) ++hv;
				h = hv;
			}
			put_u32(o, h + 1);
			o.push_back(' ');
			put_u32(o, tails[e] + 1);
			if (wts) {
				o.push_back(' ');
				put_u32(o, wts[e]);
			}
			o.push_back('\n');
		}
		codes[k] = write_file(prefix + "-" + std::to_string(k), o);
	}
	for (int c : codes)
		if (c) return c;
	return GPHI_OK;
}

int write_lines(const std::string &path, const uint32_t *a, uint64_t count)
{
	std::string o;
	o.reserve((size_t)count * 8);
	for (uint64_t i = 0; i < count; ++i) {
		put_u32(o, a[i]);
		o.push_back('\n');
	}
	return write_file(path, o);
}

}  // namespace

}  // namespace gphi

struct gphi_host_graph {
	gphi_graph_desc d;
	std::vector<uint32_t> store[8];
};

namespace gphi {
int host_adopt(const gphi_graph_desc *d, std::vector<uint32_t> *st, gphi_host_graph **out)
{
	gphi_host_graph *h = new (std::nothrow) gphi_host_graph();
	if (!h) return set_error(nullptr, GPHI_E_OOM, "out of host memory");
	h->d = *d;
	for (int i = 0; i < 7; ++i) h->store[i].swap(st[i]);
	auto ptr = [&](int i) -> const uint32_t * { return h->store[i].empty() ? nullptr : h->store[i].data(); };
	h->d.degrees = ptr(0);
	h->d.csr_edges = ptr(1);
	h->d.csr_weights = ptr(2);
	h->d.tile_heads = ptr(3);
	h->d.tile_tails = ptr(4);
	h->d.tile_weights = ptr(5);
	h->d.tile_offsets = ptr(6);
	if (d->m == 0) {  // empty edge arrays are still present
		static const uint32_t none = 0;
		if (d->csr_edges) h->d.csr_edges = &none;
		if (d->tile_heads) h->d.tile_heads = h->d.tile_tails = &none;
	}
	*out = h;
	return GPHI_OK;
}
}  // namespace gphi

using gphi::set_error;

extern "C" {

int gphi_host_read_text(const char *prefix, uint32_t W, uint32_t S, int weighted, int row_major,
                        gphi_host_graph **out)
{
	if (!prefix || !out) return set_error(nullptr, GPHI_E_ARG, "NULL argument");
	*out = nullptr;
	if (W == 0 || S == 0) return set_error(nullptr, GPHI_E_ARG, "tile width and stripe length must be > 0");
	gphi_host_graph *h = new (std::nothrow) gphi_host_graph();
	if (!h) return set_error(nullptr, GPHI_E_OOM, "out of host memory");
	int rc = gphi::read_text_graph(nullptr, prefix, W, S, weighted, row_major != 0, &h->d, h->store);
	if (rc) {
		delete h;
		return rc;
	}
	*out = h;
	return GPHI_OK;
}

int gphi_host_desc(const gphi_host_graph *h, gphi_graph_desc *out)
{
	if (!h || !out) return set_error(nullptr, GPHI_E_ARG, "NULL argument");
	*out = h->d;
	return GPHI_OK;
}

void gphi_host_free(gphi_host_graph *h) { delete h; }

int gphi_host_write_text(const gphi_host_graph *h, const char *prefix_c)
{
	if (!h || !prefix_c) return set_error(nullptr, GPHI_E_ARG, "NULL argument");
	const gphi_graph_desc &d = h->d;
	const std::string prefix(prefix_c);
	int rc;
	if (d.csr_edges) {
		const std::string p = prefix + "_untiled";
		if ((rc = gphi::write_parts(p, d.n, d.m, d.degrees, nullptr, d.csr_edges, d.csr_weights))) return rc;
		if ((rc = gphi::write_lines(p + "-nneibor", d.degrees, d.n))) return rc;
	}
	if (d.tile_heads && d.tile_tails) {
		if (!d.tile_offsets) return set_error(nullptr, GPHI_E_STATE, "tile offsets are missing");
		const std::string p =
		    d.tile_order == GPHI_TILES_ROW_MAJOR
		        ? prefix + "_coo-tiled-" + std::to_string(d.tile_width)
		        : prefix + "_col-" + std::to_string(d.stripe_length) + "-coo-tiled-" + std::to_string(d.tile_width);
		const uint64_t side = d.n % d.tile_width ? d.n / d.tile_width + 1 : d.n / d.tile_width;
		if ((rc = gphi::write_parts(p, d.n, d.m, d.degrees, d.tile_heads, d.tile_tails, d.tile_weights)))
			return rc;
		if ((rc = gphi::write_lines(p + "-offsets", d.tile_offsets, side * side))) return rc;
		if ((rc = gphi::write_lines(p + "-nneibor", d.degrees, d.n))) return rc;
	}
	return GPHI_OK;
}

}  // extern "C"

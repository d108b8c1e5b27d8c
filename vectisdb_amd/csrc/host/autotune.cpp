// autotune.cpp -- the sort variant choice: a measured static default, and an online autotuner.
//
// A radix sort here has four free parameters per shape: the digit width (8 or 11 bits, which fixes the number
// of passes), items per thread (16 or 32, which fixes the tile), and whether payload values are loaded before
// the ranking or just before their scatter (registers versus latency hiding). With the autotuner on, the
// first sort of a shape runs every variant twice on its own input, keeps the fastest, and caches it for the
// shape; every variant produces the same stable order, so the choice never changes a result.
#include "device_pool.h"
#include "operators.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

namespace vgpu {

namespace {

std::atomic<bool> g_autotune {std::getenv("VGPU_AUTOTUNE") != nullptr && std::atoi(std::getenv("VGPU_AUTOTUNE")) != 0};

struct ShapeKey {
	int passes8;
	int passes11;
	int log2rows;
	int npayload;
	int payload_bytes;
	bool hist8;
	bool operator<(const ShapeKey &o) const {
		return std::tie(passes8, passes11, log2rows, npayload, payload_bytes, hist8) <
		       std::tie(o.passes8, o.passes11, o.log2rows, o.npayload, o.payload_bytes, o.hist8);
	}
};

struct Choice {
	SortTuning tuning;
	double ms = 0;
	int candidates = 0;
};

std::mutex g_mutex;
std::map<ShapeKey, Choice> g_cache;

//! Below this many rows a sort is launch-bound and not worth tuning.
constexpr uint64_t kTuneMinRows = 1ull << 20;

int Log2Rows(uint64_t n) {
	int l = 0;
	while ((2ull << l) <= n) {
		l++;
	}
	return l;
}

//! Measured on MI355X (100M rows, 20-bit key + two 8-byte payloads; 256M rows, 25-bit key alone): sorts that
//! carry payload are fastest with the 8192-row tile as 512 threads x 16 items and the payload loaded at its
//! scatter (126 VGPRs, 4 waves per SIMD; 32 items on 256 threads needs 241, 2 waves: 0.98 -> 0.86 ms per
//! pass); key-only sorts with 16 items on 256 threads.
//! The variant used without the autotuner: for a payload sort, 16 items per thread without the prefetch in
//! tiles of 1,024 threads from 2^24 rows (16,384-row tiles: half the tiles, tickets and digit-run boundaries
//! of 512 threads; 100M rows with 16 B of payload: first pass 0.91 against 0.98 ms, step 2.60 against 2.67 ms)
//! and of 512 below (more tiles to fill the device); 256 threads for key-only sorts.
void StaticDefault(SortTuning &t, int npayload, uint64_t n) {
	t.items = 16;
	t.block = npayload > 0 ? (n >= (1ull << 24) ? 1024 : 512) : 256;
	t.prefetch = 0;
}

double TimeSort(const KeyPack &pack, const void *const *payload_in, const int *payload_width, int npayload,
                uint64_t n, const SortTuning &t, hipStream_t stream, hipEvent_t e0, hipEvent_t e1) {
	PoolArena tmp(stream);
	SortedRun run;
	VGPU_HIP(hipEventRecord(e0, stream));
	SortPackedWith(pack, payload_in, payload_width, npayload, n, t, run, tmp, stream);
	VGPU_HIP(hipEventRecord(e1, stream));
	VGPU_HIP(hipEventSynchronize(e1));
	float ms = 0;
	VGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
	return ms;
}

} // namespace

void SortAutotuneEnable(bool on) {
	g_autotune.store(on);
}

bool SortAutotuneEnabled() {
	return g_autotune.load();
}

std::string SortAutotuneReport() {
	std::lock_guard<std::mutex> lock(g_mutex);
	std::ostringstream o;
	o << "[";
	bool first = true;
	for (auto &kv : g_cache) {
		const auto &k = kv.first;
		const auto &c = kv.second;
		o << (first ? "" : ",") << "{\"passes8\":" << k.passes8 << ",\"log2rows\":" << k.log2rows
		  << ",\"npayload\":" << k.npayload << ",\"payload_bytes\":" << k.payload_bytes
		  << ",\"radix_bits\":" << c.tuning.radix_bits << ",\"items\":" << c.tuning.items
		  << ",\"block\":" << c.tuning.block
		  << ",\"prefetch\":" << c.tuning.prefetch << ",\"ms\":" << c.ms << ",\"candidates\":" << c.candidates
		  << "}";
		first = false;
	}
	o << "]";
	return o.str();
}

namespace {
//! The pass variant a tuning choice stands for (digit width, items per thread, threads per tile, payload
//! prefetch), applied to a call's own tuning: what belongs to the call (its histograms, the narrowing mask and
//! its overflow word, whether it leaves the last digit pass, the state it clears) is never taken from the call
//! that tuned the shape.
void ApplyVariant(SortTuning &dst, const SortTuning &variant) {
	dst.radix_bits = variant.radix_bits;
	dst.items = variant.items;
	dst.block = variant.block;
	dst.prefetch = variant.prefetch;
}
} // namespace

void SortPacked(const KeyPack &pack, const void *const *payload_in, const int *payload_width, int npayload,
                uint64_t n, const SortTuning &tuning, SortedRun &out, Scratch &scratch, hipStream_t stream,
                const uint64_t *n_dev) {
	SortTuning t = tuning;
	if (t.fixed) {
		SortPackedWith(pack, payload_in, payload_width, npayload, n, t, out, scratch, stream, n_dev);
		return;
	}
	StaticDefault(t, npayload, n);
	// a device-side row count is never timed (the host does not wait on it)
	if (!g_autotune.load() || n < kTuneMinRows || n_dev != nullptr) {
		SortPackedWith(pack, payload_in, payload_width, npayload, n, t, out, scratch, stream, n_dev);
		return;
	}
	const int bits = std::max(pack.total_bits, 1);
	int pbytes = 0;
	for (int p = 0; p < npayload; p++) {
		pbytes += payload_width[p];
	}
	const ShapeKey key {(bits + 7) / 8, (bits + 10) / 11, Log2Rows(n), npayload, pbytes, tuning.hist8 != nullptr};
	{
		std::lock_guard<std::mutex> lock(g_mutex);
		auto it = g_cache.find(key);
		if (it != g_cache.end()) {
			SortTuning c = t;
			ApplyVariant(c, it->second.tuning);
			SortPackedWith(pack, payload_in, payload_width, npayload, n, c, out, scratch, stream);
			return;
		}
	}
	std::vector<SortTuning> cands;
	for (int items : {16, 32}) {
		for (int pf : {1, 0}) {
			if (npayload == 0 && pf == 0) {
				continue; // the prefetch flag only moves payload loads
			}
			SortTuning c = tuning;
			c.radix_bits = 8;
			c.items = items;
			c.prefetch = pf;
			cands.push_back(c);
		}
	}
	for (int block : {512, 1024}) {
		// The 8192-row tile of 32 items per thread spread over 512 threads at 16 items each, and the
		// 16,384-row tile of 1,024 threads.
		SortTuning c = tuning;
		c.radix_bits = 8;
		c.items = 16;
		c.block = block;
		c.prefetch = 0;
		cands.push_back(c);
	}
	if (key.passes11 < key.passes8) {
		SortTuning c = tuning;
		c.radix_bits = 11;
		c.items = 16;
		c.prefetch = 1;
		cands.push_back(c);
	}
	hipEvent_t e0 = nullptr, e1 = nullptr;
	VGPU_HIP(hipEventCreate(&e0));
	VGPU_HIP(hipEventCreate(&e1));
	Choice best;
	best.ms = 1e300;
	try {
		for (const auto &c : cands) {
			double ms = 1e300;
			for (int rep = 0; rep < 2; rep++) {
				ms = std::min(ms, TimeSort(pack, payload_in, payload_width, npayload, n, c, stream, e0, e1));
			}
			if (ms < best.ms) {
				best.ms = ms;
				best.tuning = c;
			}
		}
	} catch (...) {
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		throw;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	best.candidates = (int)cands.size();
	best.tuning.hist8 = nullptr;
	if (std::getenv("VGPU_AUTOTUNE_VERBOSE") != nullptr) {
		std::fprintf(stderr,
		             "[autotune] passes8=%d rows=2^%d payload=%d/%dB -> radix %d items %d block %d prefetch %d "
		             "(%.3f ms)\n",
		             key.passes8, key.log2rows, npayload, pbytes, best.tuning.radix_bits, best.tuning.items,
		             best.tuning.block, best.tuning.prefetch, best.ms);
	}
	{
		std::lock_guard<std::mutex> lock(g_mutex);
		g_cache.emplace(key, best);
	}
	SortTuning c = t;
	ApplyVariant(c, best.tuning);
	SortPackedWith(pack, payload_in, payload_width, npayload, n, c, out, scratch, stream);
}

} // namespace vgpu

// test_host_components.cpp -- the host-side drop-in components through include/*.hpp, linked to libvgpu.so the
// way the DuckDB extension links the engine: KernelCache and its SHA-256 keys, GpuLogger, the routing
// functions and thresholds of gpu_engine.hpp, and the engine's parallel host copy. No GPU needed.
#include "harness.hpp"

#include "code_generator.hpp"
#include "gpu_engine.hpp"
#include "gpu_logger.hpp"
#include "kernel_cache.hpp"
// the engine's host-copy pool (an internal header: the staging of host-column plans uses it)
#include "../../vectisdb_amd/csrc/engine/host_copy.h"

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <thread>
#include <vector>

using namespace vector_gpu;
namespace fs = std::filesystem;

namespace {

// ---------------------------------------------------------------- KernelCache

std::vector<CompiledKernelHandle> &Deleted() {
	static std::vector<CompiledKernelHandle> deleted;
	return deleted;
}
void CountingDeleter(CompiledKernelHandle h) {
	Deleted().push_back(h);
}
void OtherDeleter(CompiledKernelHandle) {
}
CompiledKernelHandle H(uintptr_t v) {
	return reinterpret_cast<CompiledKernelHandle>(v);
}

} // namespace

VT_CASE(CacheKeepsTheMostRecentlyUsedEntries) {
	KernelCache cache(3);
	cache.Insert("k1", H(1));
	cache.Insert("k2", H(2));
	cache.Insert("k3", H(3));
	EXPECT(cache.Get("k1") == H(1)); // k1 becomes the most recent; k2 is now the oldest
	cache.Insert("k4", H(4));
	EXPECT(cache.Get("k2") == nullptr);
	EXPECT(cache.Get("k1") == H(1) && cache.Get("k3") == H(3) && cache.Get("k4") == H(4));
	cache.Insert("k5", H(5)); // oldest after those lookups: k1
	EXPECT(cache.Get("k1") == nullptr && cache.Get("k5") == H(5));
}

VT_CASE(CacheReplacementReleasesTheOldHandleOnly) {
	Deleted().clear();
	{
		KernelCache cache(2);
		cache.SetHandleDeleter(&CountingDeleter);
		cache.SetHandleDeleter(&CountingDeleter); // the same deleter again is accepted
		cache.Insert("a", H(10));
		cache.Insert("a", H(10)); // same handle: nothing released
		EXPECT(Deleted().empty());
		cache.Insert("a", H(11));
		EXPECT(Deleted() == std::vector<CompiledKernelHandle>{H(10)});
		cache.Insert("b", H(20));
		cache.Insert("c", H(30)); // evicts "a" (the oldest)
		EXPECT(Deleted().size() == 2 && Deleted()[1] == H(11));
		EXPECT_THROWS(cache.SetHandleDeleter(&OtherDeleter), "another deleter");
	}
	// destruction releases what is left
	std::set<CompiledKernelHandle> released(Deleted().begin(), Deleted().end());
	EXPECT(released.count(H(20)) == 1 && released.count(H(30)) == 1 && Deleted().size() == 4);
}

VT_CASE(CacheWithoutCapacityDeclinesEveryHandle) {
	Deleted().clear();
	KernelCache owning(0);
	owning.SetHandleDeleter(&CountingDeleter);
	owning.Insert("x", H(7));
	EXPECT(owning.Get("x") == nullptr);
	EXPECT(Deleted() == std::vector<CompiledKernelHandle>{H(7)});
	KernelCache borrowing(0); // no deleter: the caller keeps ownership, nothing is released
	borrowing.Insert("y", H(8));
	EXPECT(borrowing.Get("y") == nullptr && Deleted().size() == 1);
}

VT_CASE(HashSourceIsSha256Hex) {
	// FIPS 180-4 example messages
	EXPECT(KernelCache::HashSource("") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855");
	EXPECT(KernelCache::HashSource("abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad");
	EXPECT(KernelCache::HashSource("abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") ==
	       "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1");
	EXPECT(KernelCache::HashSource(std::string(1000000, 'a')) ==
	       "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0");
	// lengths around the 55/56/64-byte padding boundaries give distinct, stable 64-digit keys
	std::set<std::string> keys;
	for (size_t n : {54, 55, 56, 57, 63, 64, 65, 119, 120}) {
		const std::string k = KernelCache::HashSource(std::string(n, 'q'));
		EXPECT(k.size() == 64 && k.find_first_not_of("0123456789abcdef") == std::string::npos);
		EXPECT(k == KernelCache::HashSource(std::string(n, 'q')));
		keys.insert(k);
	}
	EXPECT(keys.size() == 9);
}

// ---------------------------------------------------------------- GpuLogger

namespace {

fs::path FreshDir(const char *tag) {
	fs::path dir = fs::temp_directory_path() / ("vgpu_log_" + std::string(tag) + "_" + std::to_string(::getpid()));
	fs::remove_all(dir);
	fs::create_directories(dir);
	return dir;
}

std::vector<std::string> Records(const fs::path &dir) {
	std::vector<std::string> out;
	for (const char *name : {"gpu_log.0.jsonl", "gpu_log.1.jsonl"}) {
		std::ifstream in(dir / name);
		for (std::string line; std::getline(in, line);) {
			out.push_back(line);
		}
	}
	return out;
}

void LoggerOff(const fs::path &dir) {
	GpuLogger::Instance().ConfigureForTesting(false, LogLevel::INFO, 1 << 20, dir.string());
	fs::remove_all(dir);
}

} // namespace

VT_CASE(LoggerWritesOneJsonObjectPerRecord) {
	const fs::path dir = FreshDir("records");
	auto &log = GpuLogger::Instance();
	log.ConfigureForTesting(true, LogLevel::INFO, 1 << 20, dir.string());
	EXPECT(log.Enabled(LogLevel::INFO) && !log.Enabled(LogLevel::DEBUG));
	VGPU_LOG(LogLevel::INFO, "plan_done", "\"rows\":42,\"op\":" + GpuLogger::Quote("GROUP_BY"));
	VGPU_LOG(LogLevel::ERROR, "plan_failed", std::string());
	log.FlushForTesting();
	const auto lines = Records(dir);
	EXPECT(lines.size() == 2);
	if (lines.size() == 2) {
		EXPECT(lines[0].front() == '{' && lines[0].back() == '}');
		EXPECT(lines[0].find("\"lvl\":\"info\"") != std::string::npos);
		EXPECT(lines[0].find("\"evt\":\"plan_done\"") != std::string::npos);
		EXPECT(lines[0].find("\"rows\":42,\"op\":\"GROUP_BY\"}") != std::string::npos);
		EXPECT(lines[0].find("\"ts_us\":") != std::string::npos && lines[0].find("\"thread\":") != std::string::npos);
		EXPECT(lines[1].find("\"lvl\":\"error\"") != std::string::npos && lines[1].find("plan_failed") != std::string::npos);
	}
	LoggerOff(dir);
	EXPECT(!log.Enabled(LogLevel::ERROR));
}

VT_CASE(LoggerDropsRecordsAboveItsLevelAndWhenOff) {
	const fs::path dir = FreshDir("levels");
	auto &log = GpuLogger::Instance();
	log.ConfigureForTesting(true, LogLevel::ERROR, 1 << 20, dir.string());
	int evaluated = 0;
	auto fields = [&] {
		evaluated++;
		return std::string("\"x\":1");
	};
	VGPU_LOG(LogLevel::DEBUG, "d", fields());
	VGPU_LOG(LogLevel::INFO, "i", fields());
	VGPU_LOG(LogLevel::ERROR, "e", fields());
	EXPECT(evaluated == 1); // fields of a filtered record are never built
	EXPECT(Records(dir).size() == 1);
	log.ConfigureForTesting(false, LogLevel::DEBUG, 1 << 20, dir.string());
	VGPU_LOG(LogLevel::ERROR, "off", fields());
	EXPECT(evaluated == 1);
	LoggerOff(dir);
}

VT_CASE(QuoteEscapesJsonSpecialCharacters) {
	EXPECT(GpuLogger::Quote("plain") == "\"plain\"");
	EXPECT(GpuLogger::Quote("a\"b\\c") == "\"a\\\"b\\\\c\"");
	EXPECT(GpuLogger::Quote("t\tn\nr\r") == "\"t\\tn\\nr\\r\"");
	EXPECT(GpuLogger::Quote(std::string("\x01\x1f", 2)) == "\"\\u0001\\u001f\"");
	EXPECT(GpuLogger::Quote(std::string("nul\0x", 5)) == "\"nul\\u0000x\"");
	EXPECT(GpuLogger::Quote("caf\xc3\xa9") == "\"caf\xc3\xa9\""); // UTF-8 passes through
}

VT_CASE(LoggerRecordsFromManyThreadsStayWhole) {
	const fs::path dir = FreshDir("threads");
	auto &log = GpuLogger::Instance();
	log.ConfigureForTesting(true, LogLevel::DEBUG, 64 << 20, dir.string());
	constexpr int kThreads = 6, kEach = 400;
	std::vector<std::thread> writers;
	for (int t = 0; t < kThreads; t++) {
		writers.emplace_back([t] {
			const std::string pad(100 + 37 * t, 'a' + t);
			for (int i = 0; i < kEach; i++) {
				VGPU_LOG(LogLevel::DEBUG, "w", "\"t\":" + std::to_string(t) + ",\"i\":" + std::to_string(i) +
				                                   ",\"pad\":\"" + pad + "\"");
			}
		});
	}
	for (auto &w : writers) {
		w.join();
	}
	const auto lines = Records(dir);
	EXPECT(lines.size() == static_cast<size_t>(kThreads * kEach));
	std::map<int, int> per_thread;
	bool whole = true;
	for (const auto &l : lines) {
		const auto tpos = l.find("\"t\":");
		if (tpos == std::string::npos || l.back() != '}') {
			whole = false;
			continue;
		}
		const int t = std::stoi(l.substr(tpos + 4));
		per_thread[t]++;
		whole = whole && l.find(std::string(100 + 37 * t, 'a' + t) + "\"}") != std::string::npos;
	}
	EXPECT(whole);
	EXPECT(per_thread.size() == static_cast<size_t>(kThreads));
	for (const auto &kv : per_thread) {
		EXPECT(kv.second == kEach);
	}
	LoggerOff(dir);
}

VT_CASE(LoggerStaysUnderItsCapAndKeepsTheNewestRecords) {
	const fs::path dir = FreshDir("cap");
	auto &log = GpuLogger::Instance();
	const size_t cap = 48 * 1024;
	log.ConfigureForTesting(true, LogLevel::INFO, cap, dir.string());
	const int records = 3000; // about ten times the cap
	for (int i = 0; i < records; i++) {
		VGPU_LOG(LogLevel::INFO, "rec", "\"seq\":" + std::to_string(i) + ",\"filler\":\"0123456789abcdef\"");
	}
	EXPECT(log.OnDiskBytesForTesting() <= cap);
	EXPECT(log.OnDiskBytesForTesting() > cap / 4);
	std::set<int> seen;
	for (const auto &l : Records(dir)) {
		seen.insert(std::stoi(l.substr(l.find("\"seq\":") + 6)));
	}
	EXPECT(seen.count(records - 1) == 1); // the newest survived
	EXPECT(seen.count(0) == 0);           // the oldest was rotated out
	EXPECT(!seen.empty() && *seen.rbegin() - *seen.begin() + 1 == static_cast<int>(seen.size())); // no gaps
	LoggerOff(dir);
}

// ---------------------------------------------------------------- routing (gpu_engine.hpp)

namespace {

std::shared_ptr<GpuPlanNode> Scan(uint64_t rows, uint64_t width, std::vector<std::string> cols) {
	auto s = std::make_shared<GpuPlanNode>();
	s->op_type = GpuOpType::SCAN;
	s->table = {"memory", "main", "t"};
	s->column_names = std::move(cols);
	s->estimated_rows = rows;
	s->estimated_row_width_bytes = width;
	return s;
}

GpuPlanNode Projection(std::shared_ptr<GpuPlanNode> child, double ops_per_row, uint64_t rows) {
	GpuPlanNode p;
	p.op_type = GpuOpType::PROJECTION;
	GpuExpr e;
	e.generated_cuda_source = "((long long*)out)[idx] = ((const long long*)inputs[0])[row] + 1LL;";
	e.estimated_ops_per_row = ops_per_row;
	p.expressions.push_back(e);
	p.children.push_back(std::move(child));
	p.estimated_rows = rows;
	p.estimated_row_width_bytes = 8;
	return p;
}

} // namespace

VT_CASE(DensityClassesFollowTheExportedThresholds) {
	EXPECT(kMinOpsPerScannedRow < kOffloadOpsPerRow);
	const uint64_t rows = 1 << 20;
	struct Probe {
		double ops;
		GpuWorkDensity want;
	} probes[] = {{kMinOpsPerScannedRow * 0.5, GpuWorkDensity::INSUFFICIENT},
	              {kMinOpsPerScannedRow, GpuWorkDensity::INCONCLUSIVE},
	              {(kMinOpsPerScannedRow + kOffloadOpsPerRow) / 2, GpuWorkDensity::INCONCLUSIVE},
	              {kOffloadOpsPerRow, GpuWorkDensity::SUFFICIENT},
	              {kOffloadOpsPerRow * 4, GpuWorkDensity::SUFFICIENT}};
	for (const auto &p : probes) {
		const GpuPlanShape shape = AnalyzePlanShape(Projection(Scan(rows, 8, {"a"}), p.ops, rows));
		EXPECT(shape.scanned_rows == rows && shape.scanned_bytes == rows * 8);
		EXPECT(shape.ops_per_scanned_row == p.ops);
		EXPECT(shape.density == p.want);
	}
	// no ops at all: nothing to judge
	GpuPlanNode bare = *Scan(rows, 8, {"a"});
	EXPECT(AnalyzePlanShape(bare).density == GpuWorkDensity::INCONCLUSIVE);
}

VT_CASE(RowIndependenceAndColumnSizes) {
	auto scan = Scan(3'000'000, 12, {"k", "v"});
	GpuPlanNode proj = Projection(scan, 2.0, 3'000'000);
	EXPECT(IsRowIndependent(proj));
	EXPECT(EstimateColumnsSize(proj) == 3'000'000ull * 12);
	for (GpuOpType op : {GpuOpType::GROUP_BY_AGGREGATE, GpuOpType::HASH_JOIN, GpuOpType::CROSS_PRODUCT}) {
		GpuPlanNode top;
		top.op_type = op;
		top.children.push_back(std::make_shared<GpuPlanNode>(proj));
		EXPECT(!IsRowIndependent(top));
	}
	// a small budget sends a large row-independent plan to the streamed path; a roomy one keeps it cached
	EXPECT(DetermineExecutionMode(proj, 1024, 0.75) == ExecutionMode::PIPELINE_STREAM_UNCACHED);
	EXPECT(DetermineExecutionMode(proj, 1ull << 40, 0.75) != ExecutionMode::PIPELINE_STREAM_UNCACHED);
	// row-independent row counts are capped at the chunk size
	const GpuWorkingSetEstimate capped = EstimateWorkingSet(proj, 1 << 16);
	const GpuWorkingSetEstimate whole = EstimateWorkingSet(proj, 0);
	EXPECT(capped.chunk_capped && !whole.chunk_capped);
	EXPECT(capped.total_bytes < whole.total_bytes && whole.total_bytes >= 3'000'000ull * 12);
}

VT_CASE(FastMathSwitchIsProcessWide) {
	const bool before = IsGpuFastMathEnabled();
	SetGpuFastMathEnabled(true);
	EXPECT(IsGpuFastMathEnabled());
	bool seen_on_other_thread = false;
	std::thread([&] { seen_on_other_thread = IsGpuFastMathEnabled(); }).join();
	EXPECT(seen_on_other_thread);
	SetGpuFastMathEnabled(false);
	EXPECT(!IsGpuFastMathEnabled());
	SetGpuFastMathEnabled(before);
}

// ---------------------------------------------------------------- host-copy pool

VT_CASE(ParallelHostCopyAtOddSizesWritesExactlyTheRange) {
	const size_t sizes[] = {0, 3, 4095, (1u << 20) + 1, (2u << 20) - 3, (24u << 20) + 11, (48u << 20) + 4097};
	std::vector<unsigned char> src((48u << 20) + 8192), dst(src.size());
	for (size_t i = 0; i < src.size(); i++) {
		src[i] = static_cast<unsigned char>((i * 0x9E3779B1u) >> 11);
	}
	for (size_t n : sizes) {
		std::fill(dst.begin(), dst.end(), 0x5C);
		vgpu::ParallelHostCopy(dst.data() + 1, src.data() + 3, n); // unaligned both sides
		EXPECT(dst[0] == 0x5C && dst[n + 1] == 0x5C);
		EXPECT(std::equal(dst.begin() + 1, dst.begin() + 1 + n, src.begin() + 3));
	}
}

VT_CASE(ParallelHostCopyServesConcurrentCallers) {
	std::vector<unsigned char> src(32u << 20);
	for (size_t i = 0; i < src.size(); i++) {
		src[i] = static_cast<unsigned char>(i * 131 + 7);
	}
	constexpr int kCallers = 5;
	std::vector<std::vector<unsigned char>> dsts(kCallers, std::vector<unsigned char>((6u << 20) + 333));
	std::atomic<int> done {0};
	std::vector<std::thread> callers;
	for (int c = 0; c < kCallers; c++) {
		callers.emplace_back([&, c] {
			for (int r = 0; r < 4; r++) {
				vgpu::ParallelHostCopy(dsts[c].data(), src.data() + 4097 * c, dsts[c].size());
			}
			done++;
		});
	}
	for (auto &t : callers) {
		t.join();
	}
	EXPECT(done == kCallers);
	for (int c = 0; c < kCallers; c++) {
		EXPECT(std::equal(dsts[c].begin(), dsts[c].end(), src.begin() + 4097 * c));
	}
}

VT_MAIN()

// test_host_components.cpp -- host-only drop-in components through the C++ headers, as the extension
// links them: KernelCache + SHA-256 (cases of cuda_engine/tests/test_kernel_cache.cpp) and GpuLogger
// (cases of cuda_engine/tests/test_gpu_logger.cpp), plus the routing functions of gpu_engine.hpp.
// Needs no GPU.
#include "check.hpp"
#include "gpu_engine.hpp"
#include "gpu_logger.hpp"
#include "kernel_cache.hpp"
// the engine's host-copy pool (an internal header: the staging ring of host-column plans uses it)
#include "../../vectisdb_amd/csrc/engine/host_copy.h"

#include <algorithm>
#include <filesystem>
#include <unistd.h>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

using namespace vector_gpu;

// ---------------------------------------------------------------- KernelCache

static void TestMissThenHit() {
	KernelCache cache(4);
	CHECK(cache.Get("a") == nullptr);
	int a = 1;
	cache.Insert("a", &a);
	CHECK(cache.Get("a") == &a);
}

static void TestOverwriteSameKey() {
	KernelCache cache(4);
	int a1 = 1, a2 = 2;
	cache.Insert("a", &a1);
	cache.Insert("a", &a2);
	CHECK(cache.Get("a") == &a2);
}

static void TestLruEvictionOrder() {
	KernelCache cache(2);
	int a = 1, b = 2, c = 3;
	cache.Insert("a", &a);
	cache.Insert("b", &b);
	CHECK(cache.Get("a") == &a); // a becomes most recently used
	cache.Insert("c", &c);       // evicts b
	CHECK(cache.Get("a") == &a);
	CHECK(cache.Get("b") == nullptr);
	CHECK(cache.Get("c") == &c);
}

static void TestEvictionAtExactCapacity() {
	KernelCache cache(1);
	int a = 1, b = 2;
	cache.Insert("a", &a);
	CHECK(cache.Get("a") == &a);
	cache.Insert("b", &b);
	CHECK(cache.Get("a") == nullptr);
	CHECK(cache.Get("b") == &b);
}

static void TestHashSourceDeterministicAndKnownVectors() {
	CHECK(KernelCache::HashSource("out[idx] = a + b;") == KernelCache::HashSource("out[idx] = a + b;"));
	CHECK(KernelCache::HashSource("out[idx] = a + b;") != KernelCache::HashSource("out[idx] = a - b;"));
	CHECK(KernelCache::HashSource("") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855");
	CHECK(KernelCache::HashSource("abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad");
	// FIPS 180-4 two-block message.
	CHECK(KernelCache::HashSource("abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") ==
	      "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1");
	CHECK(KernelCache::HashSource("x").size() == 64);
	CHECK(KernelCache::HashSource(std::string(10000, 'z')).size() == 64);
}

static void TestGetOnEmptyAndZeroCapacity() {
	KernelCache empty(4);
	CHECK(empty.Get("anything") == nullptr);
	KernelCache zero(0);
	int d = 1;
	zero.Insert("a", &d);
	zero.Insert("b", &d);
	CHECK(zero.Get("a") == nullptr);
	CHECK(zero.Get("b") == nullptr);
}

static std::vector<void *> g_deleted;
static void RecordingDeleter(CompiledKernelHandle h) {
	g_deleted.push_back(h);
}

static void TestDeleterOnEvictionReplacementDestruction() {
	g_deleted.clear();
	{
		KernelCache cache(1);
		cache.SetHandleDeleter(&RecordingDeleter);
		int a = 1, b = 2;
		cache.Insert("a", &a);
		cache.Insert("b", &b); // evicts a
		CHECK(g_deleted.size() == 1 && g_deleted[0] == &a);
		CHECK(cache.Get("b") == &b);
	}
	CHECK(g_deleted.size() == 2); // b freed by the destructor
	g_deleted.clear();
	KernelCache cache(4);
	cache.SetHandleDeleter(&RecordingDeleter);
	int a1 = 1, a2 = 2;
	cache.Insert("a", &a1);
	cache.Insert("a", &a2); // replacement frees a1
	CHECK(g_deleted.size() == 1 && g_deleted[0] == &a1);
	cache.Insert("a", &a2); // same handle again: nothing freed
	CHECK(g_deleted.size() == 1);
	CHECK(cache.Get("a") == &a2);
}

static void TestZeroCapacityDeletesDeclinedHandle() {
	g_deleted.clear();
	KernelCache cache(0);
	cache.SetHandleDeleter(&RecordingDeleter);
	int a = 1;
	cache.Insert("a", &a);
	CHECK(g_deleted.size() == 1 && g_deleted[0] == &a);
	CHECK(cache.Get("a") == nullptr);
}

static void TestNoDeleterMeansNoOwnership() {
	KernelCache cache(1);
	int a = 1, b = 2;
	cache.Insert("a", &a);
	cache.Insert("b", &b);
	CHECK(cache.Get("b") == &b);
}

static void TestSecondDifferentDeleterRefused() {
	KernelCache cache(2);
	cache.SetHandleDeleter(&RecordingDeleter);
	cache.SetHandleDeleter(&RecordingDeleter); // same one again is fine
	CHECK_THROWS(cache.SetHandleDeleter([](CompiledKernelHandle) {}));
}

// ---------------------------------------------------------------- GpuLogger

namespace {
std::filesystem::path LogDir() {
	auto dir = std::filesystem::temp_directory_path() / ("vgpu_logger_test_" + std::to_string(::getpid()));
	std::filesystem::create_directories(dir);
	return dir;
}

std::vector<std::string> ReadLines(const std::filesystem::path &dir) {
	std::vector<std::string> lines;
	for (int i = 0; i < 2; i++) {
		std::ifstream in(dir / ("gpu_log." + std::to_string(i) + ".jsonl"));
		std::string line;
		while (std::getline(in, line)) {
			if (!line.empty()) {
				lines.push_back(line);
			}
		}
	}
	return lines;
}

bool WholeJsonObject(const std::string &line) {
	if (line.size() < 2 || line.front() != '{' || line.back() != '}') {
		return false;
	}
	int depth = 0;
	bool in_str = false, esc = false;
	for (char c : line) {
		if (in_str) {
			if (esc) {
				esc = false;
			} else if (c == '\\') {
				esc = true;
			} else if (c == '"') {
				in_str = false;
			}
			continue;
		}
		if (c == '"') {
			in_str = true;
		} else if (c == '{') {
			depth++;
		} else if (c == '}' && --depth < 0) {
			return false;
		}
	}
	return depth == 0 && !in_str;
}

void ClearLogs(const std::filesystem::path &dir) {
	std::filesystem::remove(dir / "gpu_log.0.jsonl");
	std::filesystem::remove(dir / "gpu_log.1.jsonl");
}
} // namespace

static void TestLoggerDisabledWritesNothing() {
	auto dir = LogDir();
	ClearLogs(dir);
	GpuLogger::Instance().ConfigureForTesting(false, LogLevel::DEBUG, 1 << 20, dir.string());
	for (int i = 0; i < 100; i++) {
		VGPU_LOG(LogLevel::ERROR, "should_not_appear", "");
	}
	GpuLogger::Instance().FlushForTesting();
	CHECK(ReadLines(dir).empty());
}

static void TestLoggerRecordsRoundTripAsJsonLines() {
	auto dir = LogDir();
	ClearLogs(dir);
	GpuLogger::Instance().ConfigureForTesting(true, LogLevel::DEBUG, 1 << 20, dir.string());
	VGPU_LOG(LogLevel::INFO, "h2d", "\"column\":\"revenue\",\"bytes\":4096,\"us\":12");
	VGPU_LOG(LogLevel::INFO, "decline", std::string("\"reason\":") + GpuLogger::Quote("ORDER BY is not offloadable"));
	GpuLogger::Instance().FlushForTesting();
	auto lines = ReadLines(dir);
	CHECK(lines.size() == 2);
	for (auto &l : lines) {
		CHECK(WholeJsonObject(l));
	}
	CHECK(!lines.empty() && lines[0].find("\"evt\":\"h2d\"") != std::string::npos);
	CHECK(!lines.empty() && lines[0].find("\"ts_us\":") != std::string::npos);
	CHECK(!lines.empty() && lines[0].find("\"thread\":") != std::string::npos);
	CHECK(lines.size() > 1 && lines[1].find("ORDER BY is not offloadable") != std::string::npos);
}

static void TestLoggerLevelFiltering() {
	auto dir = LogDir();
	ClearLogs(dir);
	GpuLogger::Instance().ConfigureForTesting(true, LogLevel::ERROR, 1 << 20, dir.string());
	VGPU_LOG(LogLevel::ERROR, "kept", "");
	VGPU_LOG(LogLevel::INFO, "dropped", "");
	VGPU_LOG(LogLevel::DEBUG, "dropped", "");
	GpuLogger::Instance().FlushForTesting();
	auto lines = ReadLines(dir);
	CHECK(lines.size() == 1);
	CHECK(!lines.empty() && lines[0].find("\"evt\":\"kept\"") != std::string::npos);
}

static void TestLoggerQuoteEscapes() {
	CHECK(GpuLogger::Quote("a\"b\\c\nd") == "\"a\\\"b\\\\c\\nd\"");
}

static void TestLoggerConcurrentWritersProduceWholeLines() {
	auto dir = LogDir();
	ClearLogs(dir);
	GpuLogger::Instance().ConfigureForTesting(true, LogLevel::DEBUG, 32u << 20, dir.string());
	const int threads = 8, per_thread = 500;
	std::vector<std::thread> workers;
	for (int t = 0; t < threads; t++) {
		workers.emplace_back([t] {
			for (int i = 0; i < per_thread; i++) {
				VGPU_LOG(LogLevel::INFO, "launch",
				         "\"worker\":" + std::to_string(t) + ",\"i\":" + std::to_string(i) +
				             ",\"pad\":\"xxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxx\"");
			}
		});
	}
	for (auto &w : workers) {
		w.join();
	}
	GpuLogger::Instance().FlushForTesting();
	auto lines = ReadLines(dir);
	CHECK(lines.size() == static_cast<size_t>(threads * per_thread));
	bool whole = !lines.empty();
	for (auto &l : lines) {
		whole = whole && WholeJsonObject(l);
	}
	CHECK(whole);
}

static void TestLoggerCapEvictsOldestKeepsNewest() {
	auto dir = LogDir();
	ClearLogs(dir);
	const size_t cap = 256 * 1024;
	GpuLogger::Instance().ConfigureForTesting(true, LogLevel::DEBUG, cap, dir.string());
	const int records = 4000;
	for (int i = 0; i < records; i++) {
		VGPU_LOG(LogLevel::INFO, "seq", "\"n\":" + std::to_string(i) + ",\"pad\":\"yyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyyy\"");
	}
	GpuLogger::Instance().FlushForTesting();
	const size_t on_disk = GpuLogger::Instance().OnDiskBytesForTesting();
	CHECK(on_disk <= cap && on_disk > 0);
	auto lines = ReadLines(dir);
	bool newest = false, oldest = false;
	for (auto &l : lines) {
		newest = newest || l.find("\"n\":" + std::to_string(records - 1) + ",") != std::string::npos;
		oldest = oldest || l.find("\"n\":0,") != std::string::npos;
	}
	CHECK(newest);
	CHECK(!oldest);
	CHECK(lines.size() > 100);
	GpuLogger::Instance().ConfigureForTesting(false, LogLevel::INFO, 1 << 20, dir.string());
	std::filesystem::remove_all(dir);
}

// ---------------------------------------------------------------- routing through the C++ header

static void TestRoutingThroughHeader() {
	auto scan = std::make_shared<GpuPlanNode>();
	scan->op_type = GpuOpType::SCAN;
	scan->table = {"memory", "main", "bench"};
	scan->column_names = {"a", "b"};
	scan->estimated_rows = 20000000;
	scan->estimated_row_width_bytes = 16;
	GpuPlanNode filter;
	filter.op_type = GpuOpType::FILTER;
	GpuExpr e;
	e.generated_cuda_source = "((bool*)out)[idx] = true;";
	e.estimated_ops_per_row = 2.0;
	filter.expressions.push_back(e);
	filter.children.push_back(scan);
	filter.estimated_rows = 200000;
	filter.estimated_row_width_bytes = 16;
	auto shape = AnalyzePlanShape(filter);
	CHECK(shape.total_ops == 2.0 * 20000000.0);
	CHECK(shape.ops_per_scanned_row == 2.0);
	CHECK(shape.density == GpuWorkDensity::INSUFFICIENT);
	CHECK(IsRowIndependent(filter));
	CHECK(EstimateColumnsSize(filter) == 20000000ull * 16);
	CHECK(DetermineExecutionMode(filter, 100, 0.75) == ExecutionMode::PIPELINE_STREAM_UNCACHED);
	CHECK(EstimateWorkingSet(filter, 1 << 21).chunk_capped);
}

// ---------------------------------------------------------------- host-copy pool

static void TestParallelHostCopySizesAndConcurrency() {
	// sizes around the parallel threshold and the 4 KiB part rounding, odd tails
	const size_t sizes[] = {0, 1, (2u << 20) - 1, 2u << 20, (16u << 20) + 7, (64u << 20) + 4095};
	std::vector<uint8_t> src((64u << 20) + 4096), dst(src.size());
	for (size_t i = 0; i < src.size(); i++) {
		src[i] = (uint8_t)(i * 2654435761u >> 13);
	}
	for (size_t n : sizes) {
		std::fill(dst.begin(), dst.end(), 0xA5);
		vgpu::ParallelHostCopy(dst.data(), src.data(), n);
		CHECK(std::equal(dst.begin(), dst.begin() + n, src.begin()));
		CHECK(dst[n] == 0xA5); // nothing written past the end
	}
	// several callers at once (the pool serves them one after another)
	std::vector<std::vector<uint8_t>> outs(4, std::vector<uint8_t>((8u << 20) + 123));
	std::vector<std::thread> callers;
	for (int t = 0; t < 4; t++) {
		callers.emplace_back([&, t] {
			for (int r = 0; r < 5; r++) {
				vgpu::ParallelHostCopy(outs[t].data(), src.data() + t * 1000, outs[t].size());
			}
		});
	}
	for (auto &c : callers) {
		c.join();
	}
	for (int t = 0; t < 4; t++) {
		CHECK(std::equal(outs[t].begin(), outs[t].end(), src.begin() + t * 1000));
	}
}

int main() {
	RUN_TEST(TestMissThenHit);
	RUN_TEST(TestOverwriteSameKey);
	RUN_TEST(TestLruEvictionOrder);
	RUN_TEST(TestEvictionAtExactCapacity);
	RUN_TEST(TestHashSourceDeterministicAndKnownVectors);
	RUN_TEST(TestGetOnEmptyAndZeroCapacity);
	RUN_TEST(TestDeleterOnEvictionReplacementDestruction);
	RUN_TEST(TestZeroCapacityDeletesDeclinedHandle);
	RUN_TEST(TestNoDeleterMeansNoOwnership);
	RUN_TEST(TestSecondDifferentDeleterRefused);
	RUN_TEST(TestLoggerDisabledWritesNothing);
	RUN_TEST(TestLoggerRecordsRoundTripAsJsonLines);
	RUN_TEST(TestLoggerLevelFiltering);
	RUN_TEST(TestLoggerQuoteEscapes);
	RUN_TEST(TestLoggerConcurrentWritersProduceWholeLines);
	RUN_TEST(TestLoggerCapEvictsOldestKeepsNewest);
	RUN_TEST(TestRoutingThroughHeader);
	RUN_TEST(TestParallelHostCopySizesAndConcurrency);
	TEST_MAIN_EPILOGUE();
}

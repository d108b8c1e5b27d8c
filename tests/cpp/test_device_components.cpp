// test_device_components.cpp -- the device-side drop-in components through include/*.hpp, linked to libvgpu.so
// the way the DuckDB extension links the engine: FilterGreaterThanInt32, CodeGenerator, RmmPool,
// GpuMemoryPool, GpuColumnCache and GpuEngine::ExecutePlan. Needs a GPU. Test buffers come from the HIP runtime;
// the engine is reached only through the headers.
#include "harness.hpp"

#include "code_generator.hpp"
#include "gpu_column_cache.hpp"
#include "gpu_engine.hpp"
#include "gpu_filter_ops.hpp"
#include "gpu_memory_pool.hpp"
#include "kernel_cache.hpp"
#include "rmm_pool.hpp"

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <set>
#include <thread>
#include <vector>

using namespace vector_gpu;

namespace {

[[noreturn]] void Broken(const char *what, hipError_t st) {
	std::fprintf(stderr, "test harness HIP failure in %s: %s\n", what, hipGetErrorString(st));
	std::exit(2); // the harness is broken, not the engine
}

template <typename T>
struct DeviceBuffer {
	T *ptr = nullptr;
	size_t n = 0;
	explicit DeviceBuffer(size_t count) : n(count) {
		if (hipError_t st = hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T))) {
			Broken("hipMalloc", st);
		}
	}
	explicit DeviceBuffer(const std::vector<T> &host) : DeviceBuffer(host.size()) {
		if (!host.empty()) {
			if (hipError_t st = hipMemcpy(ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice)) {
				Broken("upload", st);
			}
		}
	}
	~DeviceBuffer() {
		if (ptr != nullptr) {
			(void)hipFree(ptr);
		}
	}
	DeviceBuffer(const DeviceBuffer &) = delete;
	std::vector<T> Read(size_t count) const {
		std::vector<T> out(count);
		if (count > 0) {
			if (hipError_t st = hipMemcpy(out.data(), ptr, count * sizeof(T), hipMemcpyDeviceToHost)) {
				Broken("download", st);
			}
		}
		return out;
	}
	T *release() {
		T *p = ptr;
		ptr = nullptr;
		return p;
	}
};

void Sync() {
	if (hipError_t st = hipDeviceSynchronize()) {
		Broken("hipDeviceSynchronize", st);
	}
}

std::vector<int32_t> KeptOnCpu(const std::vector<int32_t> &in, int32_t threshold) {
	std::vector<int32_t> out;
	for (int32_t v : in) {
		if (v > threshold) {
			out.push_back(v);
		}
	}
	return out;
}

std::vector<int32_t> KeptOnGpu(const std::vector<int32_t> &in, int32_t threshold) {
	DeviceBuffer<int32_t> din(in), dout(in.size());
	size_t count = SIZE_MAX;
	FilterGreaterThanInt32(din.ptr, in.size(), threshold, dout.ptr, &count);
	return dout.Read(count == SIZE_MAX ? 0 : count);
}

} // namespace

// ---------------------------------------------------------------- FilterGreaterThanInt32

VT_CASE(FilterKeepsOrderAcrossTileBoundaries) {
	std::mt19937 rng(2024);
	// sizes around the 8,192-row and 16,384-row tiles and the 2^22-row switch between them
	const size_t sizes[] = {1, 63, 64, 65, 8191, 8192, 8193, 100000, (1u << 22) - 1, (1u << 22) + 16385};
	for (size_t n : sizes) {
		std::vector<int32_t> in(n);
		for (auto &v : in) {
			v = static_cast<int32_t>(rng());
		}
		const int32_t t = static_cast<int32_t>(rng());
		EXPECT(KeptOnGpu(in, t) == KeptOnCpu(in, t));
	}
}

VT_CASE(FilterThresholdExtremesAndEmptyInput) {
	std::vector<int32_t> in = {INT32_MIN, -1, 0, 1, INT32_MAX, INT32_MIN + 1, INT32_MAX - 1};
	EXPECT(KeptOnGpu(in, INT32_MAX).empty());
	EXPECT(KeptOnGpu(in, INT32_MIN) == KeptOnCpu(in, INT32_MIN));
	EXPECT(KeptOnGpu(in, -1) == (std::vector<int32_t> {0, 1, INT32_MAX, INT32_MAX - 1}));
	size_t count = 77;
	FilterGreaterThanInt32(nullptr, 0, 5, nullptr, &count); // n == 0 needs no buffers
	EXPECT(count == 0);
}

VT_CASE(FilterRejectsBadArguments) {
	DeviceBuffer<int32_t> buf(std::vector<int32_t> {5, 6, 7});
	size_t count = 0;
	EXPECT_THROWS(FilterGreaterThanInt32(buf.ptr, 3, 0, buf.ptr, nullptr), "out_count");
	EXPECT_THROWS(FilterGreaterThanInt32(nullptr, 3, 0, buf.ptr, &count), "null");
	EXPECT_THROWS(FilterGreaterThanInt32(buf.ptr, 3, 0, buf.ptr, &count), "overlap");
}

VT_CASE(FilterIgnoresAnEarlierUnrelatedRuntimeError) {
	EXPECT(hipSetDevice(9999) != hipSuccess); // leaves an error behind in the runtime
	std::vector<int32_t> in = {9, -9, 10, 11};
	EXPECT(KeptOnGpu(in, 9) == (std::vector<int32_t> {10, 11}));
	(void)hipGetLastError();
}

// ---------------------------------------------------------------- CodeGenerator

namespace {
GpuExpr Expr(const std::string &body, GpuValueType type = GpuValueType::FLOAT64) {
	GpuExpr e;
	e.generated_cuda_source = body;
	e.output_type = type;
	return e;
}
const char *kPoly = "double x = ((const double*)inputs[0])[row];\n"
                    "    ((double*)out)[idx] = (x * x - 3.0) * x + 0.5;";
} // namespace

VT_CASE(CodegenCompilesRunsAndReusesTheKernel) {
	KernelCache cache(8);
	CodeGenerator gen(cache);
	const CompiledKernelHandle k = gen.CompileOrFetch(Expr(kPoly));
	EXPECT(k != nullptr);
	EXPECT(gen.CompileOrFetch(Expr(kPoly)) == k); // second request is a cache hit
	const size_t n = 300007;                        // several blocks, ragged tail
	std::vector<double> x(n);
	for (size_t i = 0; i < n; i++) {
		x[i] = static_cast<double>(static_cast<long>(i % 2001) - 1000) / 8.0;
	}
	DeviceBuffer<double> dx(x), dout(n);
	gen.Launch(k, {dx.ptr}, dout.ptr, n);
	Sync();
	const auto got = dout.Read(n);
	bool same = true;
	for (size_t i = 0; i < n; i++) {
		volatile double v = x[i];
		const double t = v * v;
		const double u = (t - 3.0) * v; // separate roundings, as the strict kernel computes them
		same = same && got[i] == u + 0.5;
	}
	EXPECT(same);
}

VT_CASE(CodegenKeysOnTheBodyNotOnACallerHash) {
	GpuExpr twice = Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] * 2.0;");
	GpuExpr halve = Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] * 0.5;");
	twice.source_hash = halve.source_hash = "caller-supplied-and-wrong";
	KernelCache cache(8);
	CodeGenerator gen(cache);
	const auto kt = gen.CompileOrFetch(twice), kh = gen.CompileOrFetch(halve);
	EXPECT(kt != kh);
	DeviceBuffer<double> in(std::vector<double> {8.0}), out(1);
	gen.Launch(kh, {in.ptr}, out.ptr, 1);
	Sync();
	EXPECT(out.Read(1)[0] == 4.0);
}

VT_CASE(CodegenReportsCompileAndLaunchErrors) {
	KernelCache cache(8);
	CodeGenerator gen(cache);
	EXPECT_THROWS(gen.CompileOrFetch(Expr("((double*)out)[idx] = undeclared_name + ;")), "");
	const auto k = gen.CompileOrFetch(Expr(kPoly));
	DeviceBuffer<double> buf(4);
	EXPECT_THROWS(gen.Launch(k, {buf.ptr}, buf.ptr, 0), "");      // zero rows
	EXPECT_THROWS(gen.Launch(nullptr, {buf.ptr}, buf.ptr, 4), ""); // no kernel
	EXPECT_THROWS(gen.Launch(k, {buf.ptr}, buf.ptr, 4, nullptr, nullptr, {nullptr, nullptr}), ""); // validity size
	KernelCache none(0);
	CodeGenerator refusing(none);
	EXPECT_THROWS(refusing.CompileOrFetch(Expr(kPoly)), "");
}

VT_CASE(CodegenIeeeEdgeCasesMatchTheHost) {
	KernelCache cache(8);
	CodeGenerator gen(cache);
	const auto kdiv = gen.CompileOrFetch(Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] / "
	                                          "((const double*)inputs[1])[row];"));
	const std::vector<double> a = {1.0, -1.0, 0.0, 1e308, 5e-324, 7.0}, b = {0.0, 0.0, 0.0, 1e-10, 2.0, 3.0};
	DeviceBuffer<double> da(a), db(b), dq(a.size());
	gen.Launch(kdiv, {da.ptr, db.ptr}, dq.ptr, a.size());
	Sync();
	const auto q = dq.Read(a.size());
	EXPECT(std::isinf(q[0]) && q[0] > 0 && std::isinf(q[1]) && q[1] < 0 && std::isnan(q[2]));
	EXPECT(std::isinf(q[3]) && q[4] == 0.0 && q[5] == 7.0 / 3.0);
	const auto kf = gen.CompileOrFetch(Expr("((float*)out)[idx] = ((const float*)inputs[0])[row] / "
	                                        "((const float*)inputs[1])[row];",
	                                        GpuValueType::FLOAT32));
	const std::vector<float> fa = {1.0f, 10.0f, 3.0e-38f, 6.0e37f, 0.1f}, fb = {3.0f, 7.0f, 3.0f, 0.25f, 3.0f};
	DeviceBuffer<float> dfa(fa), dfb(fb), dfq(fa.size());
	gen.Launch(kf, {dfa.ptr, dfb.ptr}, dfq.ptr, fa.size());
	Sync();
	const auto fq = dfq.Read(fa.size());
	for (size_t i = 0; i < fa.size(); i++) {
		volatile float n = fa[i], d = fb[i];
		const float host = n / d;
		EXPECT(std::memcmp(&host, &fq[i], 4) == 0); // correctly rounded, the subnormal quotient included
	}
}

VT_CASE(CodegenSelectionValidityAndPreloadedTables) {
	KernelCache cache(8);
	CodeGenerator gen(cache);
	const auto k = gen.CompileOrFetch(Expr("((long long*)out)[idx] = ((const long long*)inputs[0])[row] * "
	                                       "((const long long*)inputs[1])[row];",
	                                       GpuValueType::INT64));
	const std::vector<long long> a = {2, 3, 5, 7, 11, 13, 17}, b = {1, 10, 100, 1000, 10000, 100000, 1000000};
	const std::vector<uint8_t> valid_b = {1, 1, 1, 0, 1, 1, 1};
	const std::vector<uint32_t> sel = {6, 3, 1, 0};
	DeviceBuffer<long long> da(a), db(b), dout(sel.size());
	DeviceBuffer<uint8_t> dvb(valid_b), dvalid(sel.size());
	DeviceBuffer<uint32_t> dsel(sel);
	gen.Launch(k, {da.ptr, db.ptr}, dout.ptr, sel.size(), dsel.ptr, nullptr,
	           {nullptr, reinterpret_cast<const bool *>(dvb.ptr)}, reinterpret_cast<bool *>(dvalid.ptr));
	Sync();
	const auto r = dout.Read(sel.size());
	const auto v = dvalid.Read(sel.size());
	EXPECT(r[0] == 17000000 && r[2] == 30 && r[3] == 2);
	EXPECT(v == (std::vector<uint8_t> {1, 0, 1, 1}));
	const void *table[2] = {da.ptr, db.ptr};
	DeviceBuffer<const void *> dtable(std::vector<const void *>(table, table + 2));
	gen.LaunchPreloaded(k, dtable.ptr, 2, dout.ptr, sel.size(), dsel.ptr);
	Sync();
	EXPECT(dout.Read(sel.size()) == (std::vector<long long> {17000000, 7000, 30, 2}));
	EXPECT_THROWS(gen.LaunchPreloaded(k, nullptr, 2, dout.ptr, 4), "");
}

// ---------------------------------------------------------------- RmmPool

VT_CASE(RmmBudgetIsAFractionOfFreeHbm) {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	EXPECT(!pool.HasHeadroomFor(1));
	EXPECT(pool.Allocate(64) == nullptr); // nothing before a budget exists
	EXPECT_THROWS(pool.EnsureInitialized(0.0), "fraction");
	EXPECT_THROWS(pool.EnsureInitialized(1.5), "fraction");
	size_t free_bytes = 0, total = 0;
	EXPECT(hipMemGetInfo(&free_bytes, &total) == hipSuccess);
	pool.EnsureInitialized(0.25);
	const double ratio = static_cast<double>(pool.TotalPoolBytes()) / static_cast<double>(free_bytes);
	EXPECT(ratio > 0.2 && ratio < 0.3 && pool.TotalPoolBytes() % 256 == 0);
	pool.EnsureInitialized(0.9); // already initialized: unchanged
	EXPECT(static_cast<double>(pool.TotalPoolBytes()) / static_cast<double>(free_bytes) < 0.3);
	EXPECT(pool.HasHeadroomFor(pool.TotalPoolBytes()) && !pool.HasHeadroomFor(pool.TotalPoolBytes() + 1));
	EXPECT(pool.AllocatedBytes() == 0);
	pool.ResetForTesting();
}

VT_CASE(RmmBestFitSplitsAndMergesHoles) {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	pool.EnsureInitializedExactBytesForTesting(64 * 1024);
	auto *a = static_cast<char *>(pool.Allocate(8 * 1024));
	auto *b = static_cast<char *>(pool.Allocate(1000)); // rounds to 1024
	auto *c = static_cast<char *>(pool.Allocate(16 * 1024));
	auto *d = static_cast<char *>(pool.Allocate(4 * 1024));
	EXPECT(a && b && c && d && b == a + 8192 && c == b + 1024 && d == c + 16384);
	EXPECT(pool.AllocatedBytes() == 8192 + 1024 + 16384 + 4096);
	pool.Free(a, 8192);
	pool.Free(c, 16384);
	// holes: 8 KiB at a, 16 KiB at c, 34 KiB tail; the 6 KiB request takes the smallest that fits (a's)
	auto *e = static_cast<char *>(pool.Allocate(6 * 1024));
	EXPECT(e == a);
	auto *f = static_cast<char *>(pool.Allocate(12 * 1024)); // 2 KiB left at a+6K is too small; c's 16 KiB fits
	EXPECT(f == c);
	EXPECT(pool.HasHeadroomFor(34 * 1024 - 256) && !pool.HasHeadroomFor(40 * 1024));
	for (char *p : {b, d, e, f}) {
		pool.Free(p, 1);
	}
	EXPECT(pool.AllocatedBytes() == 0);
	EXPECT(pool.HasHeadroomFor(64 * 1024)); // every hole merged back into one
	EXPECT(pool.Allocate(64 * 1024) == a);
	pool.ResetForTesting();
}

VT_CASE(RmmIgnoresFreesItDidNotHandOut) {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	pool.EnsureInitializedExactBytesForTesting(8 * 1024);
	auto *p = static_cast<char *>(pool.Allocate(512));
	int on_host = 0;
	pool.Free(nullptr, 512);
	pool.Free(&on_host, 4);  // foreign
	pool.Free(p + 256, 256); // interior
	EXPECT(pool.AllocatedBytes() == 512);
	pool.Free(p, 12345); // the tracked size wins over the argument
	pool.Free(p, 512);   // second free of the same block
	EXPECT(pool.AllocatedBytes() == 0);
	EXPECT(pool.Allocate(0) == nullptr && pool.Allocate(8 * 1024 + 1) == nullptr);
	EXPECT(!pool.HasHeadroomFor(0) && !pool.HasHeadroomFor(SIZE_MAX));
	pool.ResetForTesting();
}

// ---------------------------------------------------------------- GpuMemoryPool

VT_CASE(RingHandsOutFourDistinctFullSlotsInOrder) {
	auto &ring = GpuMemoryPool::Instance();
	ring.ResetCursorForTesting();
	std::vector<GpuMemoryPoolSlot> held;
	std::set<void *> addresses;
	for (size_t i = 0; i < GpuMemoryPool::kSlotCount; i++) {
		held.push_back(ring.AcquireNext());
		EXPECT(held.back().valid() && held.back().capacity() == GpuMemoryPool::kSlotBytes);
		addresses.insert(held.back().data());
		// page-locked: the DMA engines copy straight from it
		std::memset(held.back().data(), static_cast<int>(i), 4096);
	}
	EXPECT(addresses.size() == GpuMemoryPool::kSlotCount);
	DeviceBuffer<unsigned char> dev(4096);
	EXPECT(hipMemcpy(dev.ptr, held[2].data(), 4096, hipMemcpyHostToDevice) == hipSuccess);
	EXPECT(dev.Read(4096) == std::vector<unsigned char>(4096, 2));
	void *first = held[0].data();
	held.clear();
	ring.ResetCursorForTesting();
	GpuMemoryPoolSlot again = ring.AcquireNext();
	EXPECT(again.data() == first); // cursor back at slot 0
}

VT_CASE(RingSlotsMoveWithoutReturningTwice) {
	auto &ring = GpuMemoryPool::Instance();
	ring.ResetCursorForTesting();
	GpuMemoryPoolSlot a = ring.AcquireNext();
	void *pa = a.data();
	GpuMemoryPoolSlot b(std::move(a));
	EXPECT(!a.valid() && b.data() == pa && a.capacity() == 0);
	GpuMemoryPoolSlot c = ring.AcquireNext();
	c = std::move(b); // c's old slot goes back; c now holds slot 0
	EXPECT(c.data() == pa && !b.valid());
	c = std::move(c); // self-move keeps the slot
	EXPECT(c.data() == pa);
}

VT_CASE(RingBlocksUntilTheNextSlotReturnsAndServesInArrivalOrder) {
	auto &ring = GpuMemoryPool::Instance();
	ring.ResetCursorForTesting();
	std::vector<GpuMemoryPoolSlot> held;
	for (size_t i = 0; i < GpuMemoryPool::kSlotCount; i++) {
		held.push_back(ring.AcquireNext());
	}
	std::atomic<int> got {0};
	void *taken = nullptr;
	std::thread waiter([&] {
		GpuMemoryPoolSlot s = ring.AcquireNext(); // slot 0 again, once it comes back
		taken = s.data();
		got = 1;
	});
	std::this_thread::sleep_for(std::chrono::milliseconds(100));
	EXPECT(got == 0);
	held[1] = GpuMemoryPoolSlot(); // returning slot 1 does not serve the waiter for slot 0
	std::this_thread::sleep_for(std::chrono::milliseconds(100));
	EXPECT(got == 0);
	void *slot0 = held[0].data();
	held[0] = GpuMemoryPoolSlot();
	waiter.join();
	EXPECT(got == 1 && taken == slot0);
	held.clear();
	ring.ResetCursorForTesting();
}

// ---------------------------------------------------------------- GpuColumnCache (the extension's calls)

namespace {
GpuColumnChunk Chunk(uint64_t rows, GpuValueType type = GpuValueType::INT64, bool with_validity = false) {
	GpuColumnChunk c;
	c.row_count = rows;
	c.type = type;
	c.data_bytes = rows * 8;
	DeviceBuffer<int64_t> data(rows);
	c.device_data = data.release();
	if (with_validity) {
		DeviceBuffer<uint8_t> valid(rows);
		c.validity_bytes = rows;
		c.device_validity = valid.release();
	}
	return c;
}
} // namespace

VT_CASE(ColumnCacheStagesCommitsAndReplays) {
	auto &cache = GpuColumnCache::Instance();
	const GpuTableRef t {"memory", "main", "facts"};
	cache.InvalidateAll();
	const uint64_t base = cache.CurrentBytes();
	const uint64_t gen0 = cache.TableGeneration(t);
	cache.BeginStaging(t);
	EXPECT(cache.StageChunk(t, "a", Chunk(1000)));
	EXPECT(cache.StageChunk(t, "b", Chunk(1000, GpuValueType::INT64, true)));
	EXPECT(cache.StageChunk(t, "a", Chunk(500)));
	EXPECT(cache.StageChunk(t, "b", Chunk(500)));
	EXPECT(cache.StageChunk(t, "short", Chunk(1000))); // will not match the scanned row count
	EXPECT(cache.CurrentBytes() == base + 1500 * 8 /* a */ + (1500 * 8 + 1000) /* b + validity */ + 1000 * 8 /* short */);
	EXPECT(cache.ReplayableChunkRows(t, {"a"}).empty()); // nothing is visible before the commit
	cache.CommitStaging(t, 1500);
	EXPECT(cache.ReplayableChunkRows(t, {"a", "b"}) == (std::vector<uint64_t> {1000, 500}));
	EXPECT(cache.ReplayableChunkRows(t, {"a", "short"}).empty());
	EXPECT(cache.ReplayableChunkRows(t, {}).empty());
	GpuColumnChunk got;
	EXPECT(cache.TryGetChunk(t, "b", 0, got) && got.row_count == 1000 && got.device_validity != nullptr);
	EXPECT(!cache.TryGetChunk(t, "b", 2, got) && !cache.TryGetChunk(t, "zzz", 0, got));
	EXPECT(cache.CurrentBytes() == base + 1500 * 8 * 2 + 1000);
	EXPECT(cache.TableGeneration(t) == gen0);
	cache.InvalidateTable(t);
	EXPECT(cache.TableGeneration(t) == gen0 + 1);
	EXPECT(cache.ReplayableChunkRows(t, {"a"}).empty() && cache.CurrentBytes() == base);
}

VT_CASE(ColumnCacheDropsFailedScans) {
	auto &cache = GpuColumnCache::Instance();
	const GpuTableRef t {"memory", "main", "dims"};
	cache.InvalidateAll();
	const uint64_t base = cache.CurrentBytes();
	cache.BeginStaging(t);
	EXPECT(cache.StageChunk(t, "k", Chunk(64)));
	GpuColumnChunk retyped = Chunk(64, GpuValueType::FLOAT64);
	EXPECT(!cache.StageChunk(t, "k", retyped)); // type changed mid-scan
	(void)hipFree(retyped.device_data);
	GpuColumnChunk unowned = Chunk(64);
	EXPECT(!cache.StageChunk(t, "k", unowned)); // the column stays refused for this scan
	(void)hipFree(unowned.device_data);         // a refused chunk stays the caller's
	EXPECT(cache.CurrentBytes() == base);
	cache.CommitStaging(t, 128);
	EXPECT(cache.ReplayableChunkRows(t, {"k"}).empty());
	cache.BeginStaging(t);
	EXPECT(cache.StageChunk(t, "k", Chunk(128)));
	cache.AbortStaging(t);
	EXPECT(cache.CurrentBytes() == base);
	GpuColumnChunk late = Chunk(8);
	EXPECT(!cache.StageChunk(t, "k", late)); // no open scan after the abort
	(void)hipFree(late.device_data);
	cache.BeginStaging(t);
	EXPECT(cache.StageChunk(t, "k", Chunk(10)));
	cache.CommitStaging(t, 0); // an empty scan commits nothing
	EXPECT(cache.ReplayableChunkRows(t, {"k"}).empty() && cache.CurrentBytes() == base);
	const uint64_t g = cache.TableGeneration(t);
	cache.InvalidateAll();
	EXPECT(cache.TableGeneration(t) == g + 1);
}

// ---------------------------------------------------------------- GpuEngine::ExecutePlan

VT_CASE(EngineRefusesWhatItCannotRun) {
	GpuPlanNode bogus;
	bogus.op_type = static_cast<GpuOpType>(200);
	GpuExecutionResult r = GpuEngine::Instance().ExecutePlan(bogus, {});
	EXPECT(!r.success && r.error_message.find("unknown GpuOpType") != std::string::npos);
	GpuPlanNode scan; // an empty scan of nothing succeeds with nothing
	r = GpuEngine::Instance().ExecutePlan(scan, {});
	EXPECT(r.success && r.columns.empty());
	for (GpuOpType op : {GpuOpType::HASH_JOIN, GpuOpType::GROUP_BY_AGGREGATE, GpuOpType::CROSS_PRODUCT}) {
		GpuPlanNode childless;
		childless.op_type = op;
		r = GpuEngine::Instance().ExecutePlan(childless, {});
		EXPECT(!r.success && !r.error_message.empty());
	}
}

VT_CASE(EngineGroupByWithHugeintSumsAndMinMax) {
	// SELECT g, SUM(v), MIN(v), COUNT(*) GROUP BY g over sums that overflow 64 bits
	const size_t n = 250000;
	std::vector<int32_t> g(n);
	std::vector<int64_t> v(n);
	for (size_t i = 0; i < n; i++) {
		g[i] = static_cast<int32_t>((i * 7919) % 41) - 20;
		v[i] = (i % 3 == 0 ? INT64_MAX - static_cast<int64_t>(i) : -static_cast<int64_t>(i * 977));
	}
	auto scan = std::make_shared<GpuPlanNode>();
	scan->column_names = {"g", "v"};
	GpuPlanNode plan;
	plan.op_type = GpuOpType::GROUP_BY_AGGREGATE;
	plan.group_keys = {"g"};
	GpuAggregate sum, mn, cnt;
	sum.kind = GpuAggregateKind::SUM;
	sum.input_column = "v";
	sum.output_type = GpuValueType::HUGEINT;
	mn.kind = GpuAggregateKind::MIN;
	mn.input_column = "v";
	mn.output_type = GpuValueType::INT64;
	cnt.kind = GpuAggregateKind::COUNT_STAR;
	cnt.output_type = GpuValueType::INT64;
	plan.aggregates = {sum, mn, cnt};
	plan.children.push_back(scan);
	GpuColumn cg {"g", GpuValueType::INT32, g.data(), nullptr, n, {}};
	GpuColumn cv {"v", GpuValueType::INT64, v.data(), nullptr, n, {}};
	const GpuExecutionResult r = GpuEngine::Instance().ExecutePlan(plan, {cg, cv});
	EXPECT(r.success);
	if (!r.success || r.columns.size() != 4) {
		std::fprintf(stderr, "%s\n", r.error_message.c_str());
		return;
	}
	EXPECT(r.columns[0].name == "g" && r.columns[1].name == "agg_0" && r.columns[3].name == "agg_2");
	EXPECT(r.columns[0].row_count == 41);
	std::vector<__int128> want_sum(41, 0);
	std::vector<int64_t> want_min(41, INT64_MAX), want_cnt(41, 0);
	for (size_t i = 0; i < n; i++) {
		const int k = g[i] + 20;
		want_sum[k] += v[i];
		want_min[k] = std::min(want_min[k], v[i]);
		want_cnt[k]++;
	}
	const auto *keys = static_cast<const int32_t *>(r.columns[0].data);
	const auto *sums = static_cast<const GpuHugeInt *>(r.columns[1].data);
	const auto *mins = static_cast<const int64_t *>(r.columns[2].data);
	const auto *cnts = static_cast<const int64_t *>(r.columns[3].data);
	bool same = true;
	for (size_t j = 0; j < 41; j++) {
		const int k = keys[j] + 20;
		const __int128 got = static_cast<__int128>((static_cast<unsigned __int128>(static_cast<uint64_t>(sums[j].upper))
		                                            << 64) |
		                                           sums[j].lower);
		same = same && k >= 0 && k < 41 && got == want_sum[k] && mins[j] == want_min[k] && cnts[j] == want_cnt[k];
		same = same && (j == 0 || keys[j] > keys[j - 1]); // ascending key order
	}
	EXPECT(same);
}

VT_MAIN()

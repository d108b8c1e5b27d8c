// test_device_components.cpp -- drop-in device components through the C++ headers, with the cases of the
// reference suites: FilterGreaterThanInt32 (cuda_engine/tests/test_gpu_filter.cu), CodeGenerator
// (test_code_generator.cpp), RmmPool (test_rmm_pool.cpp), GpuMemoryPool (test_gpu_memory_pool.cpp) and
// GpuEngine edge cases (test_gpu_engine.cpp). Needs a GPU. Test buffers are allocated with the HIP
// runtime; the engine itself is reached only through include/*.hpp.
#include "check.hpp"
#include "code_generator.hpp"
#include "gpu_engine.hpp"
#include "gpu_filter_ops.hpp"
#include "gpu_memory_pool.hpp"
#include "kernel_cache.hpp"
#include "rmm_pool.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <future>
#include <set>
#include <vector>

using namespace vector_gpu;

namespace {

void Hip(hipError_t status, const char *what) {
	if (status != hipSuccess) {
		std::cerr << "HIP error in " << what << ": " << hipGetErrorString(status) << "\n";
		std::exit(2); // the harness itself is broken, not a result
	}
}

template <typename T>
T *ToDevice(const std::vector<T> &host) {
	T *d = nullptr;
	Hip(hipMalloc(&d, std::max<size_t>(host.size(), 1) * sizeof(T)), "hipMalloc");
	if (!host.empty()) {
		Hip(hipMemcpy(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "H2D");
	}
	return d;
}

template <typename T>
std::vector<T> ToHost(const T *d, size_t n) {
	std::vector<T> out(n);
	if (n > 0) {
		Hip(hipMemcpy(out.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "D2H");
	}
	return out;
}

// ---------------------------------------------------------------- FilterGreaterThanInt32

std::vector<int32_t> RunFilter(const std::vector<int32_t> &in, int32_t threshold, size_t &count) {
	int32_t *din = nullptr, *dout = nullptr;
	if (!in.empty()) {
		din = ToDevice(in);
		Hip(hipMalloc(&dout, in.size() * sizeof(int32_t)), "hipMalloc out");
	}
	FilterGreaterThanInt32(din, in.size(), threshold, dout, &count);
	auto out = ToHost(dout, count);
	(void)hipFree(din);
	(void)hipFree(dout);
	return out;
}

std::vector<int32_t> CpuFilter(const std::vector<int32_t> &in, int32_t threshold) {
	std::vector<int32_t> out;
	std::copy_if(in.begin(), in.end(), std::back_inserter(out), [&](int32_t v) { return v > threshold; });
	return out;
}

} // namespace

static void TestFilterEmptyInput() {
	size_t count = 999;
	auto r = RunFilter({}, 5, count);
	CHECK(count == 0);
	CHECK(r.empty());
}

static void TestFilterAllNoneSomePass() {
	size_t count = 0;
	CHECK(RunFilter({10, 20, 30, 40}, 0, count) == CpuFilter({10, 20, 30, 40}, 0));
	CHECK(count == 4);
	CHECK(RunFilter({1, 2, 3, 4}, 100, count).empty());
	CHECK(count == 0);
	std::vector<int32_t> some = {5, -3, 10, 0, 7, -1, 100};
	CHECK(RunFilter(some, 4, count) == CpuFilter(some, 4));
	CHECK(count == CpuFilter(some, 4).size());
}

static void TestFilterSingleElementsAndBoundary() {
	size_t count = 0;
	auto r = RunFilter({42}, 10, count);
	CHECK(count == 1 && r.size() == 1 && r[0] == 42);
	CHECK(RunFilter({5}, 10, count).empty() && count == 0);
	RunFilter({10}, 10, count);
	CHECK(count == 0); // strictly greater
}

static void TestFilterLargeArrayMatchesCpu() {
	const size_t n = 1000003;
	std::vector<int32_t> in(n);
	for (size_t i = 0; i < n; i++) {
		in[i] = static_cast<int32_t>((i * 2654435761u) % 1000);
	}
	size_t count = 0;
	auto r = RunFilter(in, 500, count);
	auto expect = CpuFilter(in, 500);
	CHECK(count == expect.size());
	CHECK(r == expect);
}

static void TestFilterNegativeThresholdAndExtremes() {
	size_t count = 0;
	std::vector<int32_t> neg = {-10, -5, -1, 0, 1, 5};
	CHECK(RunFilter(neg, -5, count) == CpuFilter(neg, -5));
	auto r = RunFilter({INT32_MIN, INT32_MIN + 1, 0, INT32_MAX - 1, INT32_MAX}, INT32_MAX - 1, count);
	CHECK(count == 1 && r.size() == 1 && r[0] == INT32_MAX);
}

static void TestFilterArgumentErrors() {
	int32_t in = 1, out = 0;
	CHECK_THROWS(FilterGreaterThanInt32(&in, 1, 0, &out, nullptr));
	size_t count = 0;
	CHECK_THROWS(FilterGreaterThanInt32(nullptr, 1, 0, nullptr, &count));
	int32_t *buf = ToDevice(std::vector<int32_t> {1, 2, 3, 4, 5, 6, 7, 8});
	CHECK_THROWS(FilterGreaterThanInt32(buf, 8, 4, buf, &count));
	(void)hipFree(buf);
}

static void TestFilterLatchedUnrelatedErrorDoesNotPoison() {
	// A host-side argument error (no device work is issued) latches as the runtime's last error.
	CHECK(hipSetDevice(1 << 20) != hipSuccess);
	std::vector<int32_t> in = {1, 2, 3, 4, 5, 6, 7, 8};
	size_t count = 0;
	auto r = RunFilter(in, 4, count);
	CHECK(count == 4);
	CHECK(r == CpuFilter(in, 4));
	(void)hipGetLastError();
}

// ---------------------------------------------------------------- CodeGenerator

namespace {
GpuExpr MulSub() {
	GpuExpr e;
	e.generated_cuda_source = "double a = ((const double*)inputs[0])[idx];\n"
	                          "    double b = ((const double*)inputs[1])[idx];\n"
	                          "    double c = ((const double*)inputs[2])[idx];\n"
	                          "    ((double*)out)[idx] = a * b - c;";
	e.source_hash = KernelCache::HashSource(e.generated_cuda_source);
	return e;
}

std::vector<double> Run3(CodeGenerator &gen, CompiledKernelHandle k, const std::vector<double> &a,
                         const std::vector<double> &b, const std::vector<double> &c) {
	const size_t n = a.size();
	double *da = ToDevice(a), *db = ToDevice(b), *dc = ToDevice(c), *dout = nullptr;
	Hip(hipMalloc(&dout, n * sizeof(double)), "hipMalloc out");
	gen.Launch(k, {da, db, dc}, dout, n);
	Hip(hipDeviceSynchronize(), "sync");
	auto r = ToHost(dout, n);
	(void)hipFree(da);
	(void)hipFree(db);
	(void)hipFree(dc);
	(void)hipFree(dout);
	return r;
}
} // namespace

static void TestCodegenCompileAndRun() {
	KernelCache cache;
	CodeGenerator gen(cache);
	auto k = gen.CompileOrFetch(MulSub());
	CHECK(k != nullptr);
	std::vector<double> a = {2, 3, 4, 5}, b = {10, 10, 10, 10}, c = {1, 2, 3, 4};
	auto r = Run3(gen, k, a, b, c);
	for (size_t i = 0; i < a.size(); i++) {
		CHECK(r[i] == a[i] * b[i] - c[i]);
	}
	CHECK(Run3(gen, k, {7.0}, {6.0}, {2.0})[0] == 40.0);
}

static void TestCodegenCacheHitSameHandle() {
	KernelCache cache;
	CodeGenerator gen(cache);
	auto h1 = gen.CompileOrFetch(MulSub());
	auto h2 = gen.CompileOrFetch(MulSub());
	CHECK(h1 == h2);
	CHECK(Run3(gen, h2, {3.0}, {3.0}, {1.0})[0] == 8.0);
}

static void TestCodegenErrors() {
	KernelCache cache;
	CodeGenerator gen(cache);
	GpuExpr broken;
	broken.generated_cuda_source = "this is not valid C++ at all {{{ )(";
	CHECK_THROWS(gen.CompileOrFetch(broken));
	auto k = gen.CompileOrFetch(MulSub());
	double dummy = 0;
	CHECK_THROWS(gen.Launch(k, {}, &dummy, 0));
	CHECK_THROWS(gen.Launch(nullptr, {}, &dummy, 1));
	CHECK_THROWS(gen.Launch(k, {&dummy}, &dummy, 1, nullptr, nullptr, {nullptr, nullptr}));
}

static void TestCodegenDivisionByZeroIsInf() {
	GpuExpr e;
	e.generated_cuda_source = "((double*)out)[idx] = ((const double*)inputs[0])[idx] / ((const double*)inputs[1])[idx];";
	KernelCache cache;
	CodeGenerator gen(cache);
	auto k = gen.CompileOrFetch(e);
	double *da = ToDevice(std::vector<double> {10.0}), *db = ToDevice(std::vector<double> {0.0}), *dout = nullptr;
	Hip(hipMalloc(&dout, sizeof(double)), "hipMalloc");
	gen.Launch(k, {da, db}, dout, 1);
	Hip(hipDeviceSynchronize(), "sync");
	CHECK(std::isinf(ToHost(dout, 1)[0]));
	(void)hipFree(da);
	(void)hipFree(db);
	(void)hipFree(dout);
}

static void TestCodegenLargeMultiBlock() {
	KernelCache cache;
	CodeGenerator gen(cache);
	auto k = gen.CompileOrFetch(MulSub());
	const size_t n = 500003;
	std::vector<double> a(n), b(n), c(n);
	for (size_t i = 0; i < n; i++) {
		a[i] = double(i % 97);
		b[i] = double((i * 3) % 89);
		c[i] = double(i % 13);
	}
	auto r = Run3(gen, k, a, b, c);
	bool ok = r.size() == n;
	for (size_t i = 0; i < n && ok; i++) {
		ok = r[i] == a[i] * b[i] - c[i];
	}
	CHECK(ok);
}

static void TestCodegenFloat32DivisionBitExact() {
	GpuExpr e;
	e.generated_cuda_source = "((float*)out)[idx] = ((const float*)inputs[0])[idx] / ((const float*)inputs[1])[idx];";
	e.output_type = GpuValueType::FLOAT32;
	KernelCache cache;
	CodeGenerator gen(cache);
	auto k = gen.CompileOrFetch(e);
	std::vector<float> a = {1.0f, 2.0f, 7.0f, 1e-38f, 3.14159265f, 1e30f, 1.0f, 123456.789f};
	std::vector<float> b = {3.0f, 7.0f, 13.0f, 1e5f, 2.71828182f, 3e-8f, 49.0f, 0.0001f};
	float *da = ToDevice(a), *db = ToDevice(b), *dout = nullptr;
	Hip(hipMalloc(&dout, a.size() * sizeof(float)), "hipMalloc");
	gen.Launch(k, {da, db}, dout, a.size());
	Hip(hipDeviceSynchronize(), "sync");
	auto g = ToHost(dout, a.size());
	for (size_t i = 0; i < a.size(); i++) {
		volatile float num = a[i], den = b[i];
		float cpu = num / den;
		uint32_t cb, gb;
		std::memcpy(&cb, &cpu, 4);
		std::memcpy(&gb, &g[i], 4);
		CHECK(cb == gb); // correctly rounded, including the denormal quotient at i == 3
	}
	(void)hipFree(da);
	(void)hipFree(db);
	(void)hipFree(dout);
}

static void TestCodegenSharedBogusHashCannotServeWrongKernel() {
	GpuExpr add = MulSub(), mul = MulSub();
	add.generated_cuda_source = "double a = ((const double*)inputs[0])[idx];\n double b = ((const double*)inputs[1])[idx];\n"
	                            " double c = ((const double*)inputs[2])[idx];\n ((double*)out)[idx] = a + b + c;";
	mul.generated_cuda_source = "double a = ((const double*)inputs[0])[idx];\n double b = ((const double*)inputs[1])[idx];\n"
	                            " double c = ((const double*)inputs[2])[idx];\n ((double*)out)[idx] = a * b * c;";
	add.source_hash = mul.source_hash = "same-bogus-hash-for-both";
	KernelCache cache;
	CodeGenerator gen(cache);
	auto ka = gen.CompileOrFetch(add);
	auto km = gen.CompileOrFetch(mul);
	CHECK(ka != km);
	CHECK(Run3(gen, ka, {2.0}, {3.0}, {4.0})[0] == 9.0);
	CHECK(Run3(gen, km, {2.0}, {3.0}, {4.0})[0] == 24.0);
}

static void TestCodegenSelectionAndValidity() {
	// rows read through a selection vector, output validity = AND of the referenced inputs'
	GpuExpr e;
	e.generated_cuda_source = "((long long*)out)[idx] = ((const long long*)inputs[0])[row] + ((const long long*)inputs[1])[row];";
	e.output_type = GpuValueType::INT64;
	KernelCache cache;
	CodeGenerator gen(cache);
	auto k = gen.CompileOrFetch(e);
	std::vector<long long> a = {1, 2, 3, 4, 5, 6}, b = {10, 20, 30, 40, 50, 60};
	std::vector<uint8_t> va = {1, 1, 0, 1, 1, 1};
	std::vector<uint32_t> sel = {5, 2, 0};
	long long *da = ToDevice(a), *db = ToDevice(b), *dout = nullptr;
	uint8_t *dva = ToDevice(va);
	uint32_t *dsel = ToDevice(sel);
	bool *dvalid = nullptr;
	Hip(hipMalloc(&dout, 3 * sizeof(long long)), "hipMalloc");
	Hip(hipMalloc(&dvalid, 3), "hipMalloc");
	gen.Launch(k, {da, db}, dout, 3, dsel, nullptr, {reinterpret_cast<const bool *>(dva), nullptr}, dvalid);
	Hip(hipDeviceSynchronize(), "sync");
	auto r = ToHost(dout, 3);
	auto v = ToHost(reinterpret_cast<uint8_t *>(dvalid), 3);
	CHECK(r[0] == 66 && r[2] == 11);
	CHECK(v[0] == 1 && v[1] == 0 && v[2] == 1);
	// The same kernel through device-resident pointer tables.
	const void *tab[2] = {da, db};
	const void **dtab = nullptr;
	Hip(hipMalloc(&dtab, sizeof(tab)), "hipMalloc");
	Hip(hipMemcpy(dtab, tab, sizeof(tab), hipMemcpyHostToDevice), "H2D");
	gen.LaunchPreloaded(k, dtab, 2, dout, 3, dsel, nullptr, nullptr, nullptr);
	Hip(hipDeviceSynchronize(), "sync");
	r = ToHost(dout, 3);
	CHECK(r[0] == 66 && r[1] == 33 && r[2] == 11);
	CHECK_THROWS(gen.LaunchPreloaded(k, nullptr, 2, dout, 3));
	for (void *p : std::vector<void *> {da, db, dout, dva, dsel, dvalid, dtab}) {
		(void)hipFree(p);
	}
}

// ---------------------------------------------------------------- RmmPool

static void TestRmmInvalidFractionRejected() {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	CHECK_THROWS(pool.EnsureInitialized(0.0));
	CHECK_THROWS(pool.EnsureInitialized(-0.5));
	CHECK_THROWS(pool.EnsureInitialized(1.5));
	pool.ResetForTesting();
}

static void TestRmmBasicZeroUninitializedOver() {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	CHECK(pool.Allocate(1024) == nullptr); // uninitialized
	CHECK(!pool.HasHeadroomFor(1));
	pool.EnsureInitializedExactBytesForTesting(1 << 20);
	CHECK(pool.AllocatedBytes() == 0);
	void *p = pool.Allocate(4096);
	CHECK(p != nullptr && pool.AllocatedBytes() >= 4096);
	pool.Free(p, 4096);
	CHECK(pool.AllocatedBytes() == 0);
	CHECK(pool.Allocate(0) == nullptr);
	CHECK(pool.Allocate(2 << 20) == nullptr);
	pool.ResetForTesting();
}

static void TestRmmFragmentationAndCoalescing() {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	pool.EnsureInitializedExactBytesForTesting(3 * 4096);
	void *a = pool.Allocate(4096), *b = pool.Allocate(4096), *c = pool.Allocate(4096);
	CHECK(a && b && c);
	CHECK(pool.Allocate(1) == nullptr);
	pool.Free(b, 4096);
	CHECK(pool.HasHeadroomFor(4096));
	CHECK(!pool.HasHeadroomFor(8192));
	pool.Free(a, 4096);
	CHECK(pool.HasHeadroomFor(8192)); // coalesced
	pool.Free(c, 4096);
	CHECK(pool.AllocatedBytes() == 0);
	pool.ResetForTesting();
}

static void TestRmmForeignNullDoubleWrongSizeFrees() {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	pool.EnsureInitializedExactBytesForTesting(64 * 1024);
	int not_ours = 0;
	void *a = pool.Allocate(4096), *b = pool.Allocate(4096);
	pool.Free(&not_ours, 4);
	pool.Free(nullptr, 100);
	CHECK(pool.AllocatedBytes() == 8192);
	pool.Free(a, 4096);
	pool.Free(a, 4096); // double free: no-op
	CHECK(pool.AllocatedBytes() == 4096);
	void *c = pool.Allocate(4096), *d = pool.Allocate(4096);
	CHECK(c && d && c != d && c != b && d != b);
	pool.Free(c, 8192); // wrong size: the tracked size wins
	CHECK(pool.AllocatedBytes() == 8192);
	void *e = pool.Allocate(8192);
	if (e != nullptr) {
		auto *eb = static_cast<uint8_t *>(e);
		auto *bb = static_cast<uint8_t *>(b);
		auto *db2 = static_cast<uint8_t *>(d);
		CHECK(!(eb <= bb && bb < eb + 8192));
		CHECK(!(eb <= db2 && db2 < eb + 8192));
	}
	pool.ResetForTesting();
}

static void TestRmmAlignmentDoubleInitAndWrapBand() {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	pool.EnsureInitializedExactBytesForTesting(1 << 20);
	void *a = pool.Allocate(1), *b = pool.Allocate(1);
	CHECK(a && b && a != b && pool.AllocatedBytes() == 512);
	pool.Free(a, 1);
	pool.Free(b, 1);
	const size_t total = pool.TotalPoolBytes();
	pool.EnsureInitializedExactBytesForTesting(999999); // ignored
	CHECK(pool.TotalPoolBytes() == total);
	const size_t kMax = static_cast<size_t>(-1);
	CHECK(!pool.HasHeadroomFor(kMax));
	CHECK(!pool.HasHeadroomFor(kMax - 100));
	CHECK(!pool.HasHeadroomFor(kMax - 255));
	CHECK(!pool.HasHeadroomFor(2 << 20));
	CHECK(pool.HasHeadroomFor(4096));
	CHECK(pool.Allocate(kMax) == nullptr);
	CHECK(pool.Allocate(kMax - 254) == nullptr);
	CHECK(pool.AllocatedBytes() == 0);
	void *h = pool.Allocate(4096);
	CHECK(h != nullptr);
	pool.Free(h, 4096);
	pool.ResetForTesting();
}

static void TestRmmBudgetFromFreeHbm() {
	auto &pool = RmmPool::Instance();
	pool.ResetForTesting();
	pool.EnsureInitialized(0.5);
	size_t free_b = 0, total_b = 0;
	Hip(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
	CHECK(pool.TotalPoolBytes() > 0 && pool.TotalPoolBytes() <= total_b);
	CHECK(pool.AllocatedBytes() == 0); // budget recorded, nothing reserved
	pool.ResetForTesting();
}

// ---------------------------------------------------------------- GpuMemoryPool

static void TestRingSlotsDistinctFullCapacity() {
	auto &pool = GpuMemoryPool::Instance();
	pool.ResetCursorForTesting();
	std::vector<GpuMemoryPoolSlot> slots;
	std::set<void *> ptrs;
	for (size_t i = 0; i < GpuMemoryPool::kSlotCount; i++) {
		slots.push_back(pool.AcquireNext());
		CHECK(slots.back().valid() && slots.back().capacity() == GpuMemoryPool::kSlotBytes);
		ptrs.insert(slots.back().data());
	}
	CHECK(ptrs.size() == GpuMemoryPool::kSlotCount);
	for (auto &s : slots) {
		auto *p = static_cast<uint8_t *>(s.data());
		p[0] = 0xAB;
		p[s.capacity() - 1] = 0xCD;
		CHECK(p[0] == 0xAB && p[s.capacity() - 1] == 0xCD);
	}
}

static void TestRingCyclicOrderAndMoves() {
	auto &pool = GpuMemoryPool::Instance();
	pool.ResetCursorForTesting();
	std::vector<void *> first;
	{
		std::vector<GpuMemoryPoolSlot> slots;
		for (size_t i = 0; i < GpuMemoryPool::kSlotCount; i++) {
			slots.push_back(pool.AcquireNext());
			first.push_back(slots.back().data());
		}
	}
	pool.ResetCursorForTesting();
	bool same = true;
	{
		std::vector<GpuMemoryPoolSlot> slots;
		for (size_t i = 0; i < GpuMemoryPool::kSlotCount; i++) {
			slots.push_back(pool.AcquireNext());
			same = same && slots.back().data() == first[i];
		}
	}
	CHECK(same);
	pool.ResetCursorForTesting();
	auto original = pool.AcquireNext();
	void *data = original.data();
	GpuMemoryPoolSlot moved(std::move(original));
	CHECK(moved.valid() && moved.data() == data && !original.valid());
	GpuMemoryPoolSlot assigned;
	assigned = std::move(moved);
	CHECK(assigned.valid() && assigned.data() == data && !moved.valid());
}

static void TestRingReleaseAndBlocking() {
	auto &pool = GpuMemoryPool::Instance();
	pool.ResetCursorForTesting();
	{
		auto held = pool.AcquireNext();
		(void)held;
	}
	pool.ResetCursorForTesting();
	auto t0 = std::chrono::steady_clock::now();
	auto again = pool.AcquireNext();
	CHECK(again.valid());
	CHECK(std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count() < 100);
	again = GpuMemoryPoolSlot();
	pool.ResetCursorForTesting();
	auto held = pool.AcquireNext();
	for (size_t i = 1; i < GpuMemoryPool::kSlotCount; i++) {
		auto passing = pool.AcquireNext();
		(void)passing;
	}
	auto blocked = std::async(std::launch::async, [&pool] { return pool.AcquireNext(); });
	CHECK(blocked.wait_for(std::chrono::milliseconds(200)) != std::future_status::ready);
	held = GpuMemoryPoolSlot(); // releases slot 0 and unblocks the waiter
	CHECK(blocked.get().valid());
}

// ---------------------------------------------------------------- GpuEngine edge cases

static void TestEngineUnknownOpType() {
	GpuPlanNode plan;
	plan.op_type = static_cast<GpuOpType>(255);
	auto r = GpuEngine::Instance().ExecutePlan(plan, {});
	CHECK(!r.success);
	CHECK(r.error_message.find("unknown GpuOpType") != std::string::npos);
}

static void TestEngineDefaultPlanIsEmptyScan() {
	GpuPlanNode plan;
	auto r = GpuEngine::Instance().ExecutePlan(plan, {});
	CHECK(r.success && r.columns.empty() && r.error_message.empty());
}

static void TestEngineExecutableAndUnexecutableOps() {
	for (auto op : {GpuOpType::SCAN, GpuOpType::FILTER, GpuOpType::PROJECTION}) {
		GpuPlanNode plan;
		plan.op_type = op;
		if (op != GpuOpType::SCAN) {
			plan.children.push_back(std::make_shared<GpuPlanNode>());
		}
		auto r = GpuEngine::Instance().ExecutePlan(plan, {});
		CHECK(r.success && r.error_message.empty());
	}
	for (auto op : {GpuOpType::HASH_JOIN, GpuOpType::GROUP_BY_AGGREGATE, GpuOpType::CROSS_PRODUCT}) {
		GpuPlanNode plan;
		plan.op_type = op;
		auto r = GpuEngine::Instance().ExecutePlan(plan, {});
		CHECK(!r.success && !r.error_message.empty());
	}
}

static void TestEngineZeroCapacityCacheCompileFailsCleanly() {
	KernelCache cache(0);
	CodeGenerator gen(cache);
	GpuExpr e;
	e.generated_cuda_source = "((double*)out)[idx] = ((const double*)inputs[0])[idx] * 2.0;";
	CHECK_THROWS(gen.CompileOrFetch(e));
}

static void TestEngineGroupByThroughHeader() {
	// SELECT k, SUM(v) GROUP BY k with HUGEINT sums, as the extension would call it.
	const size_t n = 100000;
	std::vector<int64_t> k(n), v(n);
	for (size_t i = 0; i < n; i++) {
		k[i] = static_cast<int64_t>(i % 37) - 18;
		v[i] = static_cast<int64_t>(i) * 1000003 - 7;
	}
	GpuColumn ck {"k", GpuValueType::INT64, k.data(), nullptr, n, {}};
	GpuColumn cv {"v", GpuValueType::INT64, v.data(), nullptr, n, {}};
	auto scan = std::make_shared<GpuPlanNode>();
	scan->column_names = {"k", "v"};
	GpuPlanNode g;
	g.op_type = GpuOpType::GROUP_BY_AGGREGATE;
	g.group_keys = {"k"};
	GpuAggregate sum;
	sum.kind = GpuAggregateKind::SUM;
	sum.input_column = "v";
	sum.output_type = GpuValueType::HUGEINT;
	g.aggregates = {sum};
	g.children.push_back(scan);
	auto r = GpuEngine::Instance().ExecutePlan(g, {ck, cv});
	CHECK(r.success);
	if (!r.success || r.columns.size() != 2) {
		std::cerr << r.error_message << "\n";
		return;
	}
	CHECK(r.columns[0].row_count == 37 && r.columns[1].name == "agg_0");
	auto *keys = static_cast<const int64_t *>(r.columns[0].data);
	auto *sums = static_cast<const GpuHugeInt *>(r.columns[1].data);
	bool ok = true;
	for (size_t g_i = 0; g_i < 37; g_i++) {
		__int128 expect = 0;
		for (size_t i = 0; i < n; i++) {
			if (k[i] == keys[g_i]) {
				expect += v[i];
			}
		}
		__int128 got = (__int128)(((unsigned __int128)(uint64_t)sums[g_i].upper << 64) | sums[g_i].lower);
		ok = ok && got == expect;
	}
	CHECK(ok);
}

int main() {
	RUN_TEST(TestFilterEmptyInput);
	RUN_TEST(TestFilterAllNoneSomePass);
	RUN_TEST(TestFilterSingleElementsAndBoundary);
	RUN_TEST(TestFilterLargeArrayMatchesCpu);
	RUN_TEST(TestFilterNegativeThresholdAndExtremes);
	RUN_TEST(TestFilterArgumentErrors);
	RUN_TEST(TestFilterLatchedUnrelatedErrorDoesNotPoison);
	RUN_TEST(TestCodegenCompileAndRun);
	RUN_TEST(TestCodegenCacheHitSameHandle);
	RUN_TEST(TestCodegenErrors);
	RUN_TEST(TestCodegenDivisionByZeroIsInf);
	RUN_TEST(TestCodegenLargeMultiBlock);
	RUN_TEST(TestCodegenFloat32DivisionBitExact);
	RUN_TEST(TestCodegenSharedBogusHashCannotServeWrongKernel);
	RUN_TEST(TestCodegenSelectionAndValidity);
	RUN_TEST(TestRmmInvalidFractionRejected);
	RUN_TEST(TestRmmBasicZeroUninitializedOver);
	RUN_TEST(TestRmmFragmentationAndCoalescing);
	RUN_TEST(TestRmmForeignNullDoubleWrongSizeFrees);
	RUN_TEST(TestRmmAlignmentDoubleInitAndWrapBand);
	RUN_TEST(TestRmmBudgetFromFreeHbm);
	RUN_TEST(TestRingSlotsDistinctFullCapacity);
	RUN_TEST(TestRingCyclicOrderAndMoves);
	RUN_TEST(TestRingReleaseAndBlocking);
	RUN_TEST(TestEngineUnknownOpType);
	RUN_TEST(TestEngineDefaultPlanIsEmptyScan);
	RUN_TEST(TestEngineExecutableAndUnexecutableOps);
	RUN_TEST(TestEngineZeroCapacityCacheCompileFailsCleanly);
	RUN_TEST(TestEngineGroupByThroughHeader);
	TEST_MAIN_EPILOGUE();
}

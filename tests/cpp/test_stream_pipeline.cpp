// test_stream_pipeline.cpp -- GpuStreamPipeline through its C++ header: which plans it takes, schema and batch
// checks, a stream of ragged batches coming back bit-exact in submission order with several batches in flight,
// chained projection levels, NULL bitmaps, the input slot going back to the ring before results are read,
// RUNNING_SUM carried from batch to batch, and a level wider than one generated kernel. Needs a GPU.
#include "harness.hpp"

#include "gpu_memory_pool.hpp"
#include "gpu_stream_pipeline.hpp"

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <vector>

using namespace vector_gpu;

namespace {

constexpr uint64_t kMaxRows = 1 << 14;

std::shared_ptr<GpuPlanNode> Scan(std::vector<std::string> cols) {
	auto s = std::make_shared<GpuPlanNode>();
	s->op_type = GpuOpType::SCAN;
	s->table = {"memory", "main", "stream"};
	s->column_names = std::move(cols);
	return s;
}

GpuExpr E(std::string body, std::vector<std::string> inputs, GpuValueType type) {
	GpuExpr e;
	e.generated_cuda_source = std::move(body);
	e.input_columns = std::move(inputs);
	e.output_type = type;
	return e;
}

std::shared_ptr<GpuPlanNode> Project(std::shared_ptr<GpuPlanNode> child, std::vector<GpuExpr> exprs) {
	auto p = std::make_shared<GpuPlanNode>();
	p->op_type = GpuOpType::PROJECTION;
	p->children.push_back(std::move(child));
	p->expressions = std::move(exprs);
	return p;
}

// inputs: x INT64, y DOUBLE
const std::vector<std::string> kNames = {"x", "y"};
const std::vector<GpuValueType> kTypes = {GpuValueType::INT64, GpuValueType::FLOAT64};

// SELECT x * 3 - 7, y / 4.0 + x
std::shared_ptr<GpuPlanNode> TwoOutputPlan() {
	return Project(Scan(kNames),
	               {E("((long long*)out)[idx] = ((const long long*)inputs[0])[row] * 3LL - 7LL;", {"x"},
	                  GpuValueType::INT64),
	                E("((double*)out)[idx] = ((const double*)inputs[0])[row] / 4.0 + "
	                  "((double)(((const long long*)inputs[1])[row]));",
	                  {"y", "x"}, GpuValueType::FLOAT64)});
}

//! Fills one batch in a ring slot: x, then y, then (optionally) the bitmap of y.
struct Rows {
	std::vector<int64_t> x;
	std::vector<double> y;
	std::vector<uint8_t> y_valid; // Arrow bitmap, empty = no NULLs
};

GpuStreamBatch Pack(const Rows &r) {
	GpuStreamBatch b;
	b.slot = GpuMemoryPool::Instance().AcquireNext();
	b.row_count = r.x.size();
	auto *p = static_cast<uint8_t *>(b.slot.data());
	GpuColumn cx, cy;
	cx.name = "x";
	cx.type = GpuValueType::INT64;
	cx.row_count = r.x.size();
	cx.data = p;
	std::memcpy(p, r.x.data(), r.x.size() * 8);
	p += r.x.size() * 8;
	cy.name = "y";
	cy.type = GpuValueType::FLOAT64;
	cy.row_count = r.y.size();
	cy.data = p;
	std::memcpy(p, r.y.data(), r.y.size() * 8);
	p += r.y.size() * 8;
	if (!r.y_valid.empty()) {
		std::memcpy(p, r.y_valid.data(), r.y_valid.size());
		cy.validity = p;
	}
	b.columns = {cx, cy};
	return b;
}

Rows Make(uint64_t n, uint64_t seed) {
	Rows r;
	r.x.resize(n);
	r.y.resize(n);
	for (uint64_t i = 0; i < n; i++) {
		const uint64_t h = (i + 1) * 0x9E3779B97F4A7C15ull ^ seed;
		r.x[i] = static_cast<int64_t>(h >> 20) - (1ll << 43);
		r.y[i] = static_cast<double>(static_cast<int64_t>(h % 2000001) - 1000000) / 64.0;
	}
	return r;
}

template <typename T>
const T *As(const GpuColumn &c) {
	return static_cast<const T *>(c.data);
}

bool IsValid(const GpuColumn &c, uint64_t row) {
	return c.validity == nullptr || ((c.validity[row >> 3] >> (row & 7)) & 1) != 0;
}

} // namespace

VT_CASE(SupportsOnlyProjectionChainsOverOneScan) {
	EXPECT(GpuStreamPipeline::Supports(*TwoOutputPlan()));
	EXPECT(GpuStreamPipeline::Supports(*Project(TwoOutputPlan(), {E("((long long*)out)[idx] = "
	                                                                "((const long long*)inputs[0])[row];",
	                                                                {"col0"}, GpuValueType::INT64)})));
	EXPECT(!GpuStreamPipeline::Supports(*Scan(kNames))); // nothing to compute
	auto where = std::make_shared<GpuPlanNode>();
	where->op_type = GpuOpType::FILTER;
	where->children.push_back(Scan(kNames));
	where->expressions.push_back(E("((bool*)out)[idx] = ((const long long*)inputs[0])[row] > 0LL;", {"x"},
	                               GpuValueType::BOOLEAN));
	EXPECT(!GpuStreamPipeline::Supports(*where)); // a filter changes row counts
	auto grouped = std::make_shared<GpuPlanNode>();
	grouped->op_type = GpuOpType::GROUP_BY_AGGREGATE;
	grouped->children.push_back(Scan(kNames));
	EXPECT(!GpuStreamPipeline::Supports(*Project(grouped, {E("((long long*)out)[idx] = 1LL;", {}, GpuValueType::INT64)})));
	EXPECT(!GpuStreamPipeline::Supports(
	    *Project(Scan(kNames), {E("((double*)out)[idx] = ((const double*)inputs[0])[row];", {"nope"},
	                              GpuValueType::FLOAT64)})));
}

VT_CASE(ConstructorChecksTheInputSchema) {
	auto plan = TwoOutputPlan();
	EXPECT_THROWS(GpuStreamPipeline(*plan, {"x"}, {GpuValueType::INT64}, kMaxRows), "");
	EXPECT_THROWS(GpuStreamPipeline(*plan, kNames, {GpuValueType::INT64}, kMaxRows), "");
	EXPECT_THROWS(GpuStreamPipeline(*plan, kNames, kTypes, 0), "");
	GpuStreamPipeline ok(*plan, kNames, kTypes, kMaxRows);
	EXPECT(ok.output_types() == (std::vector<GpuValueType> {GpuValueType::INT64, GpuValueType::FLOAT64}));
	EXPECT(ok.in_flight() == 0);
}

VT_CASE(RaggedBatchesComeBackExactAndInOrder) {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipe(*TwoOutputPlan(), kNames, kTypes, kMaxRows);
	const uint64_t sizes[] = {kMaxRows, 1, 4097, 12345, 64, kMaxRows - 1, 999, 2048, 7, 16000, 3};
	const int total = sizeof(sizes) / sizeof(sizes[0]);
	int next = 0, taken = 0;
	size_t deepest = 0;
	bool exact = true, named = true;
	while (taken < total) {
		if (next < total && pipe.CanSubmit()) {
			pipe.Submit(Pack(Make(sizes[next], next)));
			next++;
			deepest = std::max(deepest, pipe.in_flight());
			continue;
		}
		GpuStreamResult r;
		if (!pipe.TryTake(r)) {
			EXPECT(pipe.TakeBlocking(r));
		}
		const Rows want = Make(sizes[taken], taken);
		exact = exact && r.row_count() == sizes[taken] && r.columns().size() == 2;
		named = named && r.columns()[0].name == "col0" && r.columns()[1].name == "col1";
		for (uint64_t i = 0; exact && i < r.row_count(); i++) {
			volatile double y = want.y[i];
			const double q = y / 4.0;
			exact = As<int64_t>(r.columns()[0])[i] == want.x[i] * 3 - 7 &&
			        As<double>(r.columns()[1])[i] == q + static_cast<double>(want.x[i]);
		}
		taken++;
	}
	EXPECT(exact && named);
	EXPECT(deepest == GpuStreamPipeline::kDepth); // the pipeline filled up
	EXPECT(pipe.in_flight() == 0);
	GpuStreamResult none;
	EXPECT(!pipe.TryTake(none));
}

VT_CASE(ChainedLevelsReadTheLevelBelow) {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	// level 1: col0 = x * 3 - 7, col1 = y / 4 + x; level 2: col0 = col0 + 1, col1 = col1 * col1; level 3: col0 - col1
	auto l2 = Project(TwoOutputPlan(),
	                  {E("((long long*)out)[idx] = ((const long long*)inputs[0])[row] + 1LL;", {"col0"},
	                     GpuValueType::INT64),
	                   E("((double*)out)[idx] = ((const double*)inputs[0])[row] * ((const double*)inputs[0])[row];",
	                     {"col1"}, GpuValueType::FLOAT64)});
	auto l3 = Project(l2, {E("((double*)out)[idx] = ((double)((const long long*)inputs[0])[row]) - "
	                         "((const double*)inputs[1])[row];",
	                         {"col0", "col1"}, GpuValueType::FLOAT64)});
	GpuStreamPipeline pipe(*l3, kNames, kTypes, kMaxRows);
	const Rows in = Make(5000, 99);
	pipe.Submit(Pack(in));
	GpuStreamResult r;
	EXPECT(pipe.TakeBlocking(r) && r.columns().size() == 1);
	bool same = true;
	for (uint64_t i = 0; i < 5000; i++) {
		volatile double y = in.y[i];
		const double q = y / 4.0;
		const double c1 = q + static_cast<double>(in.x[i]);
		const double sq = c1 * c1;
		same = same && As<double>(r.columns()[0])[i] == static_cast<double>(in.x[i] * 3 - 7 + 1) - sq;
	}
	EXPECT(same);
}

VT_CASE(NullsReachOnlyTheOutputsThatReadThem) {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipe(*TwoOutputPlan(), kNames, kTypes, kMaxRows);
	Rows in = Make(777, 5);
	in.y_valid.assign((777 + 7) / 8, 0xFF);
	for (uint64_t i = 3; i < 777; i += 11) {
		in.y_valid[i / 8] &= static_cast<uint8_t>(~(1u << (i % 8)));
	}
	pipe.Submit(Pack(in));
	GpuStreamResult r;
	EXPECT(pipe.TakeBlocking(r));
	const GpuColumn &c0 = r.columns()[0], &c1 = r.columns()[1];
	bool c0_all_valid = true, c1_matches = true, values = true;
	for (uint64_t i = 0; i < 777; i++) {
		const bool y_ok = i < 3 || (i - 3) % 11 != 0;
		c0_all_valid = c0_all_valid && IsValid(c0, i); // col0 reads only x
		c1_matches = c1_matches && IsValid(c1, i) == y_ok;
		values = values && As<int64_t>(c0)[i] == in.x[i] * 3 - 7;
	}
	EXPECT(c0_all_valid && c1_matches && values);
	EXPECT(c1.validity != nullptr);
	r.Reset();
	pipe.Submit(Pack(Make(777, 5))); // the next batch has no NULLs: no bitmap comes back
	EXPECT(pipe.TakeBlocking(r) && r.columns()[1].validity == nullptr);
}

VT_CASE(InputSlotGoesBackBeforeResultsAreTaken) {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipe(*TwoOutputPlan(), kNames, kTypes, kMaxRows);
	for (size_t i = 0; i < GpuStreamPipeline::kDepth; i++) {
		EXPECT(pipe.CanSubmit());
		pipe.Submit(Pack(Make(kMaxRows, 40 + i))); // ring tickets 0, 1, 2
	}
	EXPECT(pipe.in_flight() == GpuStreamPipeline::kDepth);
	EXPECT(hipDeviceSynchronize() == hipSuccess); // every upload has finished
	pipe.Poll();                                  // ... and the pipeline has seen it
	// Tickets 3 and 4 take slot 3 and slot 0; slot 0 is only free if the first batch's input went back,
	// with all three results still untaken. A waiter that never gets it must not hang the suite.
	auto got = std::make_shared<std::atomic<int>>(0);
	std::thread([got] {
		GpuMemoryPoolSlot a = GpuMemoryPool::Instance().AcquireNext();
		GpuMemoryPoolSlot b = GpuMemoryPool::Instance().AcquireNext();
		*got = a.valid() && b.valid() ? 1 : -1;
	}).detach();
	for (int i = 0; i < 1000 && *got == 0; i++) {
		std::this_thread::sleep_for(std::chrono::milliseconds(10));
	}
	EXPECT(*got == 1);
	EXPECT(pipe.in_flight() == GpuStreamPipeline::kDepth);
	while (pipe.in_flight() > 0) {
		GpuStreamResult r;
		EXPECT(pipe.TakeBlocking(r));
	}
	EXPECT(pipe.CanSubmit());
}

VT_CASE(SubmitRefusesMalformedBatches) {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipe(*TwoOutputPlan(), kNames, kTypes, kMaxRows);
	GpuStreamBatch too_big = Pack(Make(8, 1));
	too_big.row_count = kMaxRows + 1;
	EXPECT_THROWS(pipe.Submit(std::move(too_big)), "");
	GpuStreamBatch empty = Pack(Make(8, 1));
	empty.row_count = 0;
	EXPECT_THROWS(pipe.Submit(std::move(empty)), "");
	GpuStreamBatch wrong_name = Pack(Make(8, 1));
	wrong_name.columns[0].name = "w";
	EXPECT_THROWS(pipe.Submit(std::move(wrong_name)), "");
	GpuStreamBatch missing = Pack(Make(8, 1));
	missing.columns.pop_back();
	EXPECT_THROWS(pipe.Submit(std::move(missing)), "");
	EXPECT(pipe.in_flight() == 0);
}

VT_CASE(RunningSumCarriesFromBatchToBatch) {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	auto plan = Project(Scan(kNames),
	                    {E("((long long*)out)[idx] = ((const long long*)inputs[0])[row] % 1000LL;", {"x"},
	                       GpuValueType::INT64),
	                     E("((double*)out)[idx] = ((const double*)inputs[0])[row];", {"y"}, GpuValueType::FLOAT64)});
	plan->expressions[0].window_primitive = GpuWindowPrimitive::RUNNING_SUM;
	plan->expressions[1].window_primitive = GpuWindowPrimitive::RUNNING_SUM;
	EXPECT(GpuStreamPipeline::Supports(*plan));
	GpuStreamPipeline pipe(*plan, kNames, kTypes, kMaxRows);
	int64_t run_x = 0;
	double run_y = 0.0;
	bool ints = true, doubles = true;
	for (int b = 0; b < 6; b++) {
		Rows in = Make(1000 + 2500 * b, 300 + b);
		for (auto &y : in.y) {
			y = std::floor(y) / 2.0; // halves: every partial sum is exact in any order
		}
		pipe.Submit(Pack(in));
		GpuStreamResult r;
		EXPECT(pipe.TakeBlocking(r));
		for (uint64_t i = 0; i < in.x.size(); i++) {
			run_x += in.x[i] % 1000;
			run_y += in.y[i];
			ints = ints && As<int64_t>(r.columns()[0])[i] == run_x;
			doubles = doubles && As<double>(r.columns()[1])[i] == run_y;
		}
	}
	EXPECT(ints && doubles);
}

VT_CASE(WideLevelOfTwentyFourOutputs) {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	std::vector<GpuExpr> outs;
	for (int k = 0; k < 24; k++) {
		outs.push_back(E("((double*)out)[idx] = ((const double*)inputs[0])[row] - " + std::to_string(k) +
		                     ".0 * (double)((const long long*)inputs[1])[row];",
		                 {"y", "x"}, GpuValueType::FLOAT64));
	}
	GpuStreamPipeline pipe(*Project(Scan(kNames), outs), kNames, kTypes, kMaxRows);
	Rows in = Make(2222, 8);
	for (auto &x : in.x) {
		x %= 4096; // keeps every product exact
	}
	pipe.Submit(Pack(in));
	GpuStreamResult r;
	EXPECT(pipe.TakeBlocking(r) && r.columns().size() == 24);
	bool same = r.columns().size() == 24;
	for (int k = 0; same && k < 24; k++) {
		for (uint64_t i = 0; same && i < 2222; i++) {
			same = As<double>(r.columns()[k])[i] == in.y[i] - static_cast<double>(k) * static_cast<double>(in.x[i]);
		}
	}
	EXPECT(same);
}

VT_MAIN()

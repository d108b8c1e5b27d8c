// test_stream_pipeline.cpp -- GpuStreamPipeline through its C++ header, with the cases of
// cuda_engine/tests/test_gpu_stream_pipeline.cpp (bit-exact doubles, FIFO order, overlap, chained levels,
// NULL propagation, early input-slot release, bad batches) plus RUNNING_SUM carried across batches and a
// level wider than one generated kernel. Needs a GPU.
#include "check.hpp"
#include "gpu_memory_pool.hpp"
#include "gpu_stream_pipeline.hpp"

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace vector_gpu;

namespace {

constexpr uint64_t kMaxRows = 8192;

std::shared_ptr<GpuPlanNode> Scan(const std::vector<std::string> &cols) {
	auto s = std::make_shared<GpuPlanNode>();
	s->op_type = GpuOpType::SCAN;
	s->table = {"memory", "main", "t"};
	s->column_names = cols;
	return s;
}

GpuExpr Expr(const std::string &body, const std::vector<std::string> &inputs, GpuValueType t) {
	GpuExpr e;
	e.generated_cuda_source = body;
	e.input_columns = inputs;
	e.output_type = t;
	return e;
}

std::shared_ptr<GpuPlanNode> Projection(std::shared_ptr<GpuPlanNode> child, std::vector<GpuExpr> exprs) {
	auto p = std::make_shared<GpuPlanNode>();
	p->op_type = GpuOpType::PROJECTION;
	p->children.push_back(std::move(child));
	p->expressions = std::move(exprs);
	return p;
}

std::shared_ptr<GpuPlanNode> SumDiffPlan() {
	return Projection(
	    Scan({"a", "b"}),
	    {Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] * 2.0 + ((const double*)inputs[1])[row];",
	          {"a", "b"}, GpuValueType::FLOAT64),
	     Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] - ((const double*)inputs[1])[row];", {"a", "b"},
	          GpuValueType::FLOAT64)});
}

std::shared_ptr<GpuPlanNode> ChainedPlan() {
	auto inner = Projection(
	    Scan({"a", "b"}),
	    {Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] + ((const double*)inputs[1])[row];", {"a", "b"},
	          GpuValueType::FLOAT64),
	     Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] - ((const double*)inputs[1])[row];", {"a", "b"},
	          GpuValueType::FLOAT64)});
	return Projection(inner, {Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] * 3.0;", {"col0"},
	                               GpuValueType::FLOAT64),
	                          Expr("((double*)out)[idx] = ((const double*)inputs[0])[row];", {"col1"},
	                               GpuValueType::FLOAT64)});
}

GpuStreamBatch Batch(GpuMemoryPoolSlot slot, const std::vector<double> &a, const std::vector<double> &b,
                     const std::vector<uint8_t> *a_valid = nullptr) {
	GpuStreamBatch batch;
	batch.slot = std::move(slot);
	batch.row_count = a.size();
	auto *base = static_cast<uint8_t *>(batch.slot.data());
	size_t off = 0;
	const size_t bytes = a.size() * sizeof(double);
	std::memcpy(base + off, a.data(), bytes);
	GpuColumn ca;
	ca.name = "a";
	ca.type = GpuValueType::FLOAT64;
	ca.data = base + off;
	ca.row_count = a.size();
	off += bytes;
	std::memcpy(base + off, b.data(), bytes);
	GpuColumn cb;
	cb.name = "b";
	cb.type = GpuValueType::FLOAT64;
	cb.data = base + off;
	cb.row_count = b.size();
	off += bytes;
	if (a_valid != nullptr) {
		std::memcpy(base + off, a_valid->data(), a_valid->size());
		ca.validity = base + off;
	}
	batch.columns.push_back(ca);
	batch.columns.push_back(cb);
	return batch;
}

std::vector<double> Doubles(const GpuColumn &c) {
	std::vector<double> v(c.row_count);
	std::memcpy(v.data(), c.data, c.row_count * sizeof(double));
	return v;
}

bool ValidAt(const GpuColumn &c, uint64_t r) {
	return c.validity == nullptr || (c.validity[r / 8] >> (r % 8)) & 1;
}

} // namespace

static void TestSupportsOnlyProjectionChainsOverScan() {
	CHECK(GpuStreamPipeline::Supports(*SumDiffPlan()));
	CHECK(GpuStreamPipeline::Supports(*ChainedPlan()));
	CHECK(!GpuStreamPipeline::Supports(*Scan({"a"})));
	auto filter = std::make_shared<GpuPlanNode>();
	filter->op_type = GpuOpType::FILTER;
	filter->children.push_back(Scan({"a"}));
	filter->expressions.push_back(Expr("((bool*)out)[idx] = ((const double*)inputs[0])[row] > 0.0;", {"a"},
	                                   GpuValueType::BOOLEAN));
	CHECK(!GpuStreamPipeline::Supports(*filter));
	CHECK(!GpuStreamPipeline::Supports(*Projection(filter, {Expr("((double*)out)[idx] = ((const double*)inputs[0])[row];",
	                                                             {"a"}, GpuValueType::FLOAT64)})));
	CHECK(!GpuStreamPipeline::Supports(*Projection(Scan({"a"}), {Expr("((double*)out)[idx] = 0.0;", {"missing"},
	                                                                   GpuValueType::FLOAT64)})));
}

static void TestConstructorRejectsMismatchedSchema() {
	auto plan = SumDiffPlan();
	CHECK_THROWS(GpuStreamPipeline(*plan, {"a"}, {GpuValueType::FLOAT64}, kMaxRows));
	CHECK_THROWS(GpuStreamPipeline(*plan, {"a", "b"}, {GpuValueType::FLOAT64}, kMaxRows));
	CHECK_THROWS(GpuStreamPipeline(*plan, {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, 0));
}

static void TestStreamsManyBatchesInOrderAndOverlaps() {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipeline(*SumDiffPlan(), {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, kMaxRows);
	CHECK(pipeline.output_types().size() == 2);
	const int batches = 9;
	const uint64_t rows = 5000;
	std::vector<double> want_sum, want_diff, got_sum, got_diff;
	int submitted = 0;
	size_t max_in_flight = 0;
	GpuStreamResult result;
	auto producer = GpuMemoryPool::Instance().AcquireNext();
	while (submitted < batches || pipeline.in_flight() > 0) {
		if (submitted < batches && pipeline.CanSubmit()) {
			std::vector<double> a(rows), b(rows);
			for (uint64_t i = 0; i < rows; i++) {
				a[i] = submitted * 1000.0 + i * 0.5;
				b[i] = i * 0.25 - 3.0;
				want_sum.push_back(a[i] * 2.0 + b[i]);
				want_diff.push_back(a[i] - b[i]);
			}
			pipeline.Submit(Batch(std::move(producer), a, b));
			producer = GpuMemoryPool::Instance().AcquireNext();
			submitted++;
			max_in_flight = std::max(max_in_flight, pipeline.in_flight());
			continue;
		}
		if (!pipeline.TryTake(result) && !pipeline.TakeBlocking(result)) {
			break;
		}
		CHECK(result.columns().size() == 2 && result.columns()[0].name == "col0" && result.columns()[1].name == "col1");
		auto s = Doubles(result.columns()[0]);
		auto d = Doubles(result.columns()[1]);
		got_sum.insert(got_sum.end(), s.begin(), s.end());
		got_diff.insert(got_diff.end(), d.begin(), d.end());
		result.Reset();
	}
	CHECK(submitted == batches && pipeline.in_flight() == 0);
	CHECK(max_in_flight >= 2);
	CHECK(got_sum == want_sum); // bit-exact, in submission order
	CHECK(got_diff == want_diff);
}

static void TestChainedLevels() {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipeline(*ChainedPlan(), {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, kMaxRows);
	const uint64_t rows = 4096;
	std::vector<double> a(rows), b(rows);
	for (uint64_t i = 0; i < rows; i++) {
		a[i] = i * 0.125;
		b[i] = double(i % 17) - 8.0;
	}
	CHECK(pipeline.CanSubmit());
	pipeline.Submit(Batch(GpuMemoryPool::Instance().AcquireNext(), a, b));
	GpuStreamResult r;
	CHECK(pipeline.TakeBlocking(r));
	CHECK(r.row_count() == rows);
	auto x = Doubles(r.columns()[0]);
	auto y = Doubles(r.columns()[1]);
	bool ok = true;
	for (uint64_t i = 0; i < rows && ok; i++) {
		ok = x[i] == (a[i] + b[i]) * 3.0 && y[i] == a[i] - b[i];
	}
	CHECK(ok);
}

static void TestNullPropagation() {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipeline(*SumDiffPlan(), {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, kMaxRows);
	const uint64_t rows = 1000;
	std::vector<double> a(rows), b(rows);
	for (uint64_t i = 0; i < rows; i++) {
		a[i] = double(i);
		b[i] = i * 0.5;
	}
	std::vector<uint8_t> valid((rows + 7) / 8, 0xFF);
	for (uint64_t i = 0; i < rows; i += 7) {
		valid[i / 8] &= static_cast<uint8_t>(~(1u << (i % 8)));
	}
	CHECK(pipeline.CanSubmit());
	pipeline.Submit(Batch(GpuMemoryPool::Instance().AcquireNext(), a, b, &valid));
	GpuStreamResult r;
	CHECK(pipeline.TakeBlocking(r));
	auto s = Doubles(r.columns()[0]);
	bool nulls = true, values = true;
	for (uint64_t i = 0; i < rows; i++) {
		const bool want = i % 7 != 0;
		nulls = nulls && ValidAt(r.columns()[0], i) == want && ValidAt(r.columns()[1], i) == want;
		values = values && (!want || s[i] == a[i] * 2.0 + b[i]);
	}
	CHECK(nulls);
	CHECK(values);
	CHECK(r.columns()[1].validity != nullptr);
	r.Reset();
	// A batch without NULLs right after a nullable one runs the plain kernels and carries no validity.
	pipeline.Submit(Batch(GpuMemoryPool::Instance().AcquireNext(), a, b));
	CHECK(pipeline.TakeBlocking(r));
	CHECK(r.columns()[0].validity == nullptr && Doubles(r.columns()[0])[7] == a[7] * 2.0 + b[7]);
}

static void TestInputSlotReturnedBeforeResultsAreRead() {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipeline(*SumDiffPlan(), {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, kMaxRows);
	std::vector<double> a(512, 1.5), b(512, 2.5);
	auto producer = GpuMemoryPool::Instance().AcquireNext();
	CHECK(pipeline.CanSubmit());
	pipeline.Submit(Batch(std::move(producer), a, b));
	producer = GpuMemoryPool::Instance().AcquireNext();
	CHECK(pipeline.CanSubmit());
	pipeline.Submit(Batch(std::move(producer), a, b));
	producer = GpuMemoryPool::Instance().AcquireNext();
	CHECK(pipeline.in_flight() == 2);
	bool released = false;
	for (int i = 0; i < 1000000 && !released; i++) {
		released = pipeline.CanSubmit();
	}
	CHECK(released);
	CHECK(pipeline.in_flight() == 2);
	pipeline.Submit(Batch(std::move(producer), a, b));
	producer = GpuMemoryPool::Instance().AcquireNext();
	CHECK(producer.valid() && pipeline.in_flight() == 3);
	GpuStreamResult r;
	while (pipeline.in_flight() > 0) {
		CHECK(pipeline.TakeBlocking(r));
		r.Reset();
	}
}

static void TestSubmitRejectsBadBatches() {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	GpuStreamPipeline pipeline(*SumDiffPlan(), {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, kMaxRows);
	std::vector<double> a(16, 1.0), b(16, 2.0);
	auto oversized = Batch(GpuMemoryPool::Instance().AcquireNext(), a, b);
	oversized.row_count = kMaxRows + 1;
	CHECK_THROWS(pipeline.Submit(std::move(oversized)));
	auto renamed = Batch(GpuMemoryPool::Instance().AcquireNext(), a, b);
	renamed.columns[1].name = "z";
	CHECK_THROWS(pipeline.Submit(std::move(renamed)));
	auto empty = Batch(GpuMemoryPool::Instance().AcquireNext(), a, b);
	empty.row_count = 0;
	CHECK_THROWS(pipeline.Submit(std::move(empty)));
}

static void TestRunningSumCarriesAcrossBatches() {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	auto plan = Projection(Scan({"a", "b"}),
	                       {Expr("((long long*)out)[idx] = (long long)((const double*)inputs[0])[row];", {"a"},
	                             GpuValueType::INT64),
	                        Expr("((double*)out)[idx] = ((const double*)inputs[0])[row];", {"b"}, GpuValueType::FLOAT64)});
	plan->expressions[0].window_primitive = GpuWindowPrimitive::RUNNING_SUM;
	plan->expressions[1].window_primitive = GpuWindowPrimitive::RUNNING_SUM;
	CHECK(GpuStreamPipeline::Supports(*plan));
	GpuStreamPipeline pipeline(*plan, {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, kMaxRows);
	long long run_i = 0;
	double run_d = 0.0;
	bool ints = true, dbl = true;
	for (int batch = 0; batch < 5; batch++) {
		const uint64_t rows = 3000 + 1234 * batch;
		std::vector<double> a(rows), b(rows);
		for (uint64_t i = 0; i < rows; i++) {
			a[i] = double((i * 7 + batch) % 101) - 50.0;
			b[i] = 0.5 * double(i % 8); // exact in any order
		}
		pipeline.Submit(Batch(GpuMemoryPool::Instance().AcquireNext(), a, b));
		GpuStreamResult r;
		CHECK(pipeline.TakeBlocking(r));
		auto *gi = static_cast<const long long *>(r.columns()[0].data);
		auto gd = Doubles(r.columns()[1]);
		for (uint64_t i = 0; i < rows; i++) {
			run_i += (long long)a[i];
			run_d += b[i];
			ints = ints && gi[i] == run_i;
			dbl = dbl && gd[i] == run_d;
		}
	}
	CHECK(ints);
	CHECK(dbl);
}

static void TestWideLevelSplitsAcrossKernels() {
	GpuMemoryPool::Instance().ResetCursorForTesting();
	std::vector<GpuExpr> exprs;
	for (int i = 0; i < 20; i++) {
		exprs.push_back(Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] * " + std::to_string(i + 1) +
		                         ".0 + ((const double*)inputs[1])[row];",
		                     {"a", "b"}, GpuValueType::FLOAT64));
	}
	auto plan = Projection(Scan({"a", "b"}), exprs);
	GpuStreamPipeline pipeline(*plan, {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, kMaxRows);
	const uint64_t rows = 3333;
	std::vector<double> a(rows), b(rows);
	for (uint64_t i = 0; i < rows; i++) {
		a[i] = i * 0.25;
		b[i] = double(i % 3);
	}
	pipeline.Submit(Batch(GpuMemoryPool::Instance().AcquireNext(), a, b));
	GpuStreamResult r;
	CHECK(pipeline.TakeBlocking(r));
	CHECK(r.columns().size() == 20);
	bool ok = r.columns().size() == 20;
	for (int e = 0; e < 20 && ok; e++) {
		auto v = Doubles(r.columns()[static_cast<size_t>(e)]);
		for (uint64_t i = 0; i < rows && ok; i++) {
			ok = v[i] == a[i] * double(e + 1) + b[i];
		}
	}
	CHECK(ok);
}

int main() {
	RUN_TEST(TestSupportsOnlyProjectionChainsOverScan);
	RUN_TEST(TestConstructorRejectsMismatchedSchema);
	RUN_TEST(TestStreamsManyBatchesInOrderAndOverlaps);
	RUN_TEST(TestChainedLevels);
	RUN_TEST(TestNullPropagation);
	RUN_TEST(TestInputSlotReturnedBeforeResultsAreRead);
	RUN_TEST(TestSubmitRejectsBadBatches);
	RUN_TEST(TestRunningSumCarriesAcrossBatches);
	RUN_TEST(TestWideLevelSplitsAcrossKernels);
	TEST_MAIN_EPILOGUE();
}

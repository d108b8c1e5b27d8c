// stream_pipeline.cpp -- the benchmark program for throughput of GpuStreamPipeline (cuda_engine/src/gpu_stream_pipeline.cu's
// interface): host batches in page-locked pool slots go H2D, through the generated projection kernel, and
// back D2H, three batches in flight. Prints one JSON line; bench.py reports it as an operator line.
// The batch contents are written once per slot before the timed loop (the producer's own copy is not
// part of the pipeline), and the last result is checked bit-exact against the same expressions on the CPU.
#include "gpu_memory_pool.hpp"
#include "gpu_stream_pipeline.hpp"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace vector_gpu;

namespace {

GpuExpr Expr(const std::string &body, std::vector<std::string> inputs, GpuValueType type) {
	GpuExpr e;
	e.generated_cuda_source = body;
	e.input_columns = std::move(inputs);
	e.output_type = type;
	return e;
}

//! SELECT a * 2 + b, a - b FROM t (DOUBLE columns)
std::shared_ptr<GpuPlanNode> Plan() {
	auto scan = std::make_shared<GpuPlanNode>();
	scan->op_type = GpuOpType::SCAN;
	scan->table = {"memory", "main", "t"};
	scan->column_names = {"a", "b"};
	auto p = std::make_shared<GpuPlanNode>();
	p->op_type = GpuOpType::PROJECTION;
	p->children.push_back(scan);
	p->expressions = {
	    Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] * 2.0 + ((const double*)inputs[1])[row];",
	         {"a", "b"}, GpuValueType::FLOAT64),
	    Expr("((double*)out)[idx] = ((const double*)inputs[0])[row] - ((const double*)inputs[1])[row];", {"a", "b"},
	         GpuValueType::FLOAT64)};
	return p;
}

GpuStreamBatch Wrap(GpuMemoryPoolSlot slot, uint64_t rows) {
	GpuStreamBatch batch;
	batch.row_count = rows;
	auto *base = static_cast<uint8_t *>(slot.data());
	GpuColumn a, b;
	a.name = "a";
	b.name = "b";
	a.type = b.type = GpuValueType::FLOAT64;
	a.row_count = b.row_count = rows;
	a.data = base;
	b.data = base + rows * sizeof(double);
	batch.columns = {a, b};
	batch.slot = std::move(slot);
	return batch;
}

double ValueA(uint64_t slot, uint64_t i) {
	return (double)(slot * 7 + 1) + (double)i * 0.5;
}
double ValueB(uint64_t i) {
	return (double)i * 0.25 - 3.0;
}

} // namespace

int main(int argc, char **argv) {
	const uint64_t rows = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : (8ull << 20);
	const int batches = argc > 2 ? std::atoi(argv[2]) : 24;
	if (rows * 2 * sizeof(double) > GpuMemoryPool::kSlotBytes || batches < 1) {
		std::fprintf(stderr, "bench_stream_pipeline: batch does not fit a pool slot\n");
		return 2;
	}
	auto plan = Plan();
	GpuStreamPipeline pipeline(*plan, {"a", "b"}, {GpuValueType::FLOAT64, GpuValueType::FLOAT64}, rows);
	// fill every slot once, in the order AcquireNext hands them out
	GpuMemoryPool::Instance().ResetCursorForTesting();
	for (uint64_t s = 0; s < GpuMemoryPool::kSlotCount; s++) {
		GpuMemoryPoolSlot slot = GpuMemoryPool::Instance().AcquireNext();
		auto *a = static_cast<double *>(slot.data());
		double *b = a + rows;
		for (uint64_t i = 0; i < rows; i++) {
			a[i] = ValueA(s, i);
			b[i] = ValueB(i);
		}
	}
	GpuMemoryPool::Instance().ResetCursorForTesting();
	// one untimed round trip compiles the kernel and warms the streams
	{
		pipeline.Submit(Wrap(GpuMemoryPool::Instance().AcquireNext(), rows));
		GpuStreamResult r;
		if (!pipeline.TakeBlocking(r)) {
			std::fprintf(stderr, "bench_stream_pipeline: warm-up batch lost\n");
			return 1;
		}
	}
	GpuMemoryPool::Instance().ResetCursorForTesting();
	int submitted = 0, taken = 0;
	uint64_t last_slot = 0;
	bool ok = true;
	GpuStreamResult result;
	const auto t0 = std::chrono::steady_clock::now();
	while (taken < batches) {
		if (submitted < batches && pipeline.CanSubmit()) {
			pipeline.Submit(Wrap(GpuMemoryPool::Instance().AcquireNext(), rows));
			submitted++;
			continue;
		}
		if (!pipeline.TryTake(result) && !pipeline.TakeBlocking(result)) {
			ok = false;
			break;
		}
		last_slot = (uint64_t)taken % GpuMemoryPool::kSlotCount;
		taken++;
		if (taken == batches) {
			const auto &c = result.columns();
			ok = ok && c.size() == 2 && result.row_count() == rows;
			for (uint64_t i = 0; ok && i < rows; i += 4099) {
				const double a = ValueA(last_slot, i), b = ValueB(i);
				double s, d;
				std::memcpy(&s, static_cast<const double *>(c[0].data) + i, 8);
				std::memcpy(&d, static_cast<const double *>(c[1].data) + i, 8);
				ok = ok && s == a * 2.0 + b && d == a - b;
			}
		}
		result.Reset();
	}
	const auto t1 = std::chrono::steady_clock::now();
	const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
	const double total_rows = (double)rows * batches;
	const double host_bytes = total_rows * 2 * sizeof(double) * 2; // two columns up, two results down
	std::printf("{\"operator\": \"GpuStreamPipeline, SELECT a*2+b, a-b over %d batches of %llu DOUBLE rows "
	            "(PCIe-inclusive)\", \"rows\": %.0f, \"batches\": %d, \"ms\": %.3f, \"rows_per_s\": %.1f, "
	            "\"host_bytes_moved\": %.0f, \"host_gbs\": %.2f, \"results_checked\": %s}\n",
	            batches, (unsigned long long)rows, total_rows, batches, ms, total_rows / (ms * 1e-3), host_bytes,
	            host_bytes / (ms * 1e-3) / 1e9, ok ? "true" : "false");
	return ok ? 0 : 1;
}

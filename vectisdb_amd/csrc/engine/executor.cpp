// executor.cpp -- GpuEngine::ExecutePlan: the recursive plan executor behind gpu_engine.hpp.
//
// Behaviour follows cuda_engine/src/gpu_engine.cpp:222-259 (ExecutePlan) and
// cuda_engine/src/gpu_executor.cu:368-967 (ExecuteScan, FILTER, PROJECTION, ExecuteNode, ExecuteGpuPlan)
// with the operators of cuda_engine/src/operators/*.cu: SCAN leaves consume `inputs` in traversal order
// (children[0] before children[1]); projection outputs are named col<i>, aggregates agg_<i>, group keys
// keep their names, join and cross outputs follow output_columns; NULLs are modelled only by
// row-independent plans (a NULL predicate drops the row, a projection is NULL where any referenced input
// is); every failure becomes success=false with the reference's message.
//
// Execution is shaped for the device instead of mirrored:
//   * Every intermediate is dense. A FILTER is one generated kernel that evaluates all its predicates and
//     writes the surviving rows of the columns its consumer reads (wave64 ballot compaction, decoupled
//     look-back); a PROJECTION over a FILTER (or over a CROSS_PRODUCT, or a FILTER over one) is the same
//     single kernel, so no mask, index vector, selection or per-expression pass is ever materialized.
//   * Columns no consumer reads are pruned: they keep their name and type (so name resolution and its
//     errors are unchanged) but are never uploaded, gathered or written.
//   * A global aggregate over an INNER join runs as a sort-merge join-aggregate without materializing a
//     single pair.
//   * Device memory comes from a stream-ordered caching pool, not one hipMalloc/hipFree per buffer.
//   * NULL validity is packed into the Arrow bitmap on the device; only the bitmap crosses PCIe.
#include "host_copy.h"
#include "../../../include/code_generator.hpp"
#include "../../../include/gpu_column_cache.hpp"
#include "../../../include/gpu_engine.hpp"
#include "../../../include/gpu_logger.hpp"
#include "../../../include/gpu_semantics.hpp"
#include "../../../include/pinned_buffer_pool.hpp"
#include "../../../include/rmm_pool.hpp"
#include "../host/device_pool.h"
#include "../host/operators.h"
#include "extended.h"
#include "fused.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <exception>
#include <thread>
#include <vector>

namespace vector_gpu {

namespace {
std::atomic<int> g_extended_override {-1};
} // namespace

bool IsGpuExtendedSemanticsEnabled() {
	const int v = g_extended_override.load();
	if (v != -1) {
		return v != 0;
	}
	const char *env = std::getenv("VECTOR_GPU_EXTENDED_SEMANTICS");
	return env != nullptr && (std::strcmp(env, "1") == 0 || std::strcmp(env, "true") == 0);
}

void SetGpuExtendedSemanticsEnabled(bool enabled) {
	g_extended_override.store(enabled ? 1 : 0);
}

namespace {

using NameSet = std::set<std::string>;

size_t TypeSize(GpuValueType type) {
	switch (type) {
	case GpuValueType::INT16:
		return 2;
	case GpuValueType::INT32:
		return 4;
	case GpuValueType::INT64:
		return 8;
	case GpuValueType::HUGEINT:
		return 16;
	case GpuValueType::FLOAT32:
		return 4;
	case GpuValueType::FLOAT64:
		return 8;
	case GpuValueType::BOOLEAN:
		return 1;
	case GpuValueType::DICTIONARY_STRING:
		throw std::runtime_error("gpu_executor: dictionary-encoded string columns cannot be executed on GPU");
	}
	throw std::runtime_error("gpu_executor: unknown GpuValueType");
}

//! A column resident in device memory while a plan executes. Always dense. `valid` is one byte per row
//! (non-zero = valid) or nullptr when the column has no NULLs. A pruned column carries name, type and row
//! count only: no operator above it reads its values.
struct DeviceColumn {
	std::string name;
	GpuValueType type = GpuValueType::FLOAT64;
	void *data = nullptr;
	uint64_t rows = 0;
	uint8_t *valid = nullptr;
	bool pruned = false;
};
using Columns = std::vector<DeviceColumn>;

//! Host <-> device transfers of one plan through two page-locked slots used alternately: the host copy
//! into (or out of) one slot overlaps the DMA of the other, and a plan never holds more than two slots
//! whatever its column count (the pool blocks at its cap, so holding one slot per column until the plan
//! ends would stall a plan wider than the pool).
class Staging {
public:
	explicit Staging(hipStream_t stream) : stream_(stream) {
	}
	~Staging() {
		for (auto &l : lanes_) {
			if (l.busy) {
				(void)hipEventSynchronize(l.done);
			}
			if (l.done != nullptr) {
				(void)hipEventDestroy(l.done);
			}
		}
	}
	Staging(const Staging &) = delete;
	Staging &operator=(const Staging &) = delete;

	//! Copies `bytes` of host memory to the device; `src` may be reused as soon as this returns.
	void Upload(void *dst, const void *src, size_t bytes) {
		for (size_t off = 0; off < bytes; off += kPiece) {
			const size_t len = std::min(kPiece, bytes - off);
			Lane &l = Take();
			vgpu::ParallelHostCopy(l.slot.data(), static_cast<const uint8_t *>(src) + off, len);
			VGPU_HIP(hipMemcpyAsync(static_cast<uint8_t *>(dst) + off, l.slot.data(), len, hipMemcpyHostToDevice,
			                        stream_));
			Issued(l, nullptr, 0);
		}
	}
	//! Queues a device-to-host copy; `dst` holds the data once Drain() returns.
	void Download(void *dst, const void *src, size_t bytes) {
		for (size_t off = 0; off < bytes; off += kPiece) {
			const size_t len = std::min(kPiece, bytes - off);
			Lane &l = Take();
			VGPU_HIP(hipMemcpyAsync(l.slot.data(), static_cast<const uint8_t *>(src) + off, len,
			                        hipMemcpyDeviceToHost, stream_));
			Issued(l, static_cast<uint8_t *>(dst) + off, len);
		}
	}
	void Drain() {
		Finish(lanes_[next_]);
		Finish(lanes_[next_ ^ 1]);
	}

private:
	static constexpr size_t kPiece = 32u << 20;
	struct Lane {
		PinnedBufferSlot slot;
		bool have = false;
		hipEvent_t done = nullptr;
		bool busy = false;
		void *host_dst = nullptr;
		size_t host_bytes = 0;
	};
	Lane &Take() {
		Lane &l = lanes_[next_];
		next_ ^= 1;
		Finish(l);
		if (!l.have) {
			l.slot = PinnedBufferPool::Instance().Acquire(kPiece);
			l.have = true;
		}
		if (l.done == nullptr) {
			VGPU_HIP(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
		}
		return l;
	}
	void Issued(Lane &l, void *host_dst, size_t host_bytes) {
		VGPU_HIP(hipEventRecord(l.done, stream_));
		l.busy = true;
		l.host_dst = host_dst;
		l.host_bytes = host_bytes;
	}
	void Finish(Lane &l) {
		if (!l.busy) {
			return;
		}
		VGPU_HIP(hipEventSynchronize(l.done));
		l.busy = false;
		if (l.host_dst != nullptr) {
			vgpu::ParallelHostCopy(l.host_dst, l.slot.data(), l.host_bytes);
			l.host_dst = nullptr;
		}
	}
	hipStream_t stream_;
	Lane lanes_[2];
	int next_ = 0;
};

struct ExecContext {
	const std::vector<GpuColumn> &inputs;
	size_t cursor = 0;
	vgpu::PoolArena &arena;
	hipStream_t stream = nullptr;
	Staging &staging;
	int64_t cache_chunk_index = -1;
	bool serve_whole_scan_from_cache = false;
	bool stage_scan_to_cache = false;
	bool nulls_supported = false;
	bool fast_math = false;
	bool extended = false; // IsGpuExtendedSemanticsEnabled() for this plan

	void *Alloc(size_t bytes) {
		return arena.Alloc(bytes == 0 ? 1 : bytes);
	}
};

const DeviceColumn &FindColumn(const Columns &columns, const std::string &name) {
	for (auto &column : columns) {
		if (column.name == name) {
			if (column.pruned) {
				throw std::logic_error("gpu_executor: column '" + name + "' was pruned but is read");
			}
			return column;
		}
	}
	throw std::runtime_error("gpu_executor: expression references column '" + name +
	                         "' which is not produced by its input operator");
}

bool HasColumn(const Columns &columns, const std::string &name) {
	for (auto &column : columns) {
		if (column.name == name) {
			return true;
		}
	}
	return false;
}

//! Resolves an output name against the two inputs of a join or cross product (same errors as
//! cuda_engine/src/gpu_executor.cu:116-141).
const DeviceColumn &ResolveSideColumn(const std::string &name, const Columns &left, const Columns &right,
                                      bool &from_left) {
	const DeviceColumn *found = nullptr;
	for (auto &column : left) {
		if (column.name == name) {
			found = &column;
			from_left = true;
			break;
		}
	}
	for (auto &column : right) {
		if (column.name == name) {
			if (found != nullptr) {
				throw std::runtime_error("gpu_executor: output column '" + name +
				                         "' is produced by both inputs, so selecting it by name is ambiguous");
			}
			found = &column;
			from_left = false;
			break;
		}
	}
	if (found == nullptr) {
		throw std::runtime_error("gpu_executor: output column '" + name + "' is produced by neither input");
	}
	return *found;
}

bool Needed(const NameSet *need, const std::string &name) {
	return need == nullptr || need->count(name) != 0;
}

void AddInputs(NameSet &set, const std::vector<GpuExpr> &exprs) {
	for (auto &e : exprs) {
		set.insert(e.input_columns.begin(), e.input_columns.end());
	}
}

DeviceColumn PrunedLike(const DeviceColumn &column, uint64_t rows) {
	DeviceColumn p;
	p.name = column.name;
	p.type = column.type;
	p.rows = rows;
	p.pruned = true;
	return p;
}

DeviceColumn EmptyColumn(const std::string &name, GpuValueType type, ExecContext &ctx) {
	DeviceColumn c;
	c.name = name;
	c.type = type;
	c.rows = 0;
	c.data = ctx.Alloc(0);
	return c;
}

const char *CTypeName(GpuValueType type) {
	switch (type) {
	case GpuValueType::INT16:
		return "short";
	case GpuValueType::INT32:
		return "int";
	case GpuValueType::INT64:
		return "long long";
	case GpuValueType::FLOAT32:
		return "float";
	case GpuValueType::FLOAT64:
		return "double";
	case GpuValueType::BOOLEAN:
		return "bool";
	default:
		return nullptr;
	}
}

bool FusableType(GpuValueType type) {
	return CTypeName(type) != nullptr;
}

//! The body that copies `column` unchanged, in the same text contract as the extension's expressions.
GpuExpr IdentityExpr(const DeviceColumn &column) {
	if (!FusableType(column.type)) {
		throw std::runtime_error("gpu_executor: column '" + column.name +
		                         "' has a type no generated expression can read");
	}
	GpuExpr e;
	const std::string t = CTypeName(column.type);
	e.generated_cuda_source = "((" + t + "*)out)[idx] = ((const " + t + "*)inputs[0])[row];";
	e.input_columns = {column.name};
	e.output_type = column.type;
	return e;
}

// ------------------------------------------------------------------ SCAN

Columns ExecuteScan(const GpuPlanNode &node, ExecContext &ctx, const NameSet *need) {
	Columns columns;
	columns.reserve(node.column_names.size());

	if (ctx.cache_chunk_index >= 0) {
		const auto chunk_index = static_cast<uint64_t>(ctx.cache_chunk_index);
		for (auto &column_name : node.column_names) {
			GpuColumnChunk cached;
			if (!GpuColumnCache::Instance().TryGetChunk(node.table, column_name, chunk_index, cached)) {
				throw std::runtime_error("gpu_executor: chunk " + std::to_string(chunk_index) + " of column '" +
				                         column_name + "' of table '" + node.table.ToString() +
				                         "' was expected to be cache-resident but is no longer present "
				                         "(likely invalidated by a concurrent write) -- retry the query");
			}
			if (cached.device_validity != nullptr && !ctx.nulls_supported) {
				throw std::runtime_error("gpu_executor: cached column '" + column_name +
				                         "' contains NULLs, which this operator does not model");
			}
			DeviceColumn column;
			column.name = column_name;
			column.type = cached.type;
			column.rows = cached.row_count;
			column.data = cached.device_data;
			column.valid = static_cast<uint8_t *>(cached.device_validity);
			column.pruned = !Needed(need, column_name);
			VGPU_LOG(LogLevel::DEBUG, "cache_hit",
			         "\"table\":" + GpuLogger::Quote(node.table.ToString()) + ",\"column\":" +
			             GpuLogger::Quote(column_name) + ",\"chunk\":" + std::to_string(chunk_index) +
			             ",\"rows\":" + std::to_string(column.rows));
			columns.push_back(std::move(column));
		}
		return columns;
	}

	if (ctx.serve_whole_scan_from_cache) {
		auto chunk_rows = GpuColumnCache::Instance().ReplayableChunkRows(node.table, node.column_names);
		if (chunk_rows.empty()) {
			throw std::runtime_error("gpu_executor: table '" + node.table.ToString() +
			                         "' was expected to be fully cache-resident but is no longer present "
			                         "(likely invalidated by a concurrent write) -- retry the query");
		}
		uint64_t total_rows = 0;
		for (auto rows : chunk_rows) {
			total_rows += rows;
		}
		for (auto &column_name : node.column_names) {
			std::vector<GpuColumnChunk> chunks;
			bool any_validity = false;
			for (uint64_t index = 0; index < chunk_rows.size(); index++) {
				GpuColumnChunk cached;
				if (!GpuColumnCache::Instance().TryGetChunk(node.table, column_name, index, cached)) {
					throw std::runtime_error("gpu_executor: chunk " + std::to_string(index) + " of column '" +
					                         column_name + "' of table '" + node.table.ToString() +
					                         "' disappeared from the cache mid-gather -- retry the query");
				}
				if (cached.device_validity != nullptr) {
					if (!ctx.nulls_supported) {
						throw std::runtime_error("gpu_executor: cached column '" + column_name +
						                         "' contains NULLs, which this operator does not model");
					}
					any_validity = true;
				}
				chunks.push_back(cached);
			}
			const size_t width = TypeSize(chunks.front().type);
			DeviceColumn column;
			column.name = column_name;
			column.type = chunks.front().type;
			column.rows = total_rows;
			if (!Needed(need, column_name)) {
				column.pruned = true;
				columns.push_back(std::move(column));
				continue;
			}
			for (auto &cached : chunks) {
				if (cached.data_bytes != cached.row_count * width) {
					throw std::runtime_error("gpu_executor: cached chunk of column '" + column_name + "' holds " +
					                         std::to_string(cached.data_bytes) + " bytes for " +
					                         std::to_string(cached.row_count) +
					                         " rows, which does not match its declared type");
				}
			}
			if (chunks.size() == 1) {
				// The whole column is one resident chunk: borrow it. Nothing evicts while a plan is served
				// from the cache, and the borrow ends with this execution.
				column.data = chunks[0].device_data;
				column.valid = static_cast<uint8_t *>(chunks[0].device_validity);
			} else {
				column.data = ctx.Alloc(total_rows * width);
				if (any_validity) {
					column.valid = static_cast<uint8_t *>(ctx.Alloc(total_rows));
					vgpu::FillU8(column.valid, 1, total_rows, ctx.stream);
				}
				uint64_t offset = 0;
				for (auto &cached : chunks) {
					vgpu::CopyDevice(static_cast<uint8_t *>(column.data) + offset * width, cached.device_data,
					                        cached.row_count * width, ctx.stream);
					if (any_validity && cached.device_validity != nullptr) {
						vgpu::CopyDevice(column.valid + offset, cached.device_validity, cached.row_count, ctx.stream);
					}
					offset += cached.row_count;
				}
			}
			VGPU_LOG(LogLevel::INFO, "cache_gather",
			         "\"table\":" + GpuLogger::Quote(node.table.ToString()) + ",\"column\":" +
			             GpuLogger::Quote(column_name) + ",\"chunks\":" + std::to_string(chunks.size()) +
			             ",\"rows\":" + std::to_string(total_rows));
			columns.push_back(std::move(column));
		}
		return columns;
	}

	for (size_t i = 0; i < node.column_names.size(); i++) {
		if (ctx.cursor >= ctx.inputs.size()) {
			throw std::runtime_error("gpu_executor: plan expects more scanned columns than were supplied");
		}
		const GpuColumn &host = ctx.inputs[ctx.cursor++];
		if (host.validity != nullptr && !ctx.nulls_supported) {
			throw std::runtime_error("gpu_executor: column '" + host.name +
			                         "' contains NULLs, which this operator does not model");
		}
		DeviceColumn column;
		column.name = host.name;
		column.type = host.type;
		column.rows = host.row_count;
		const size_t bytes = static_cast<size_t>(host.row_count) * TypeSize(host.type);
		// Staging offers every column to the cache, read or not, so the cached table is complete.
		if (!ctx.stage_scan_to_cache && !Needed(need, host.name)) {
			column.pruned = true;
			columns.push_back(std::move(column));
			continue;
		}
		// Drawn from the pool directly rather than through the plan arena: the buffer may end up owned by
		// the column cache. Exactly one owner results either way (the cache, or the arena adopting it).
		column.data = vgpu::DevicePool::Instance().Alloc(bytes == 0 ? 1 : bytes, ctx.stream);
		if (bytes > 0 && host.data != nullptr) {
			auto start = std::chrono::steady_clock::now();
			ctx.staging.Upload(column.data, host.data, bytes);
			VGPU_LOG(LogLevel::INFO, "h2d",
			         "\"column\":" + GpuLogger::Quote(host.name) + ",\"bytes\":" + std::to_string(bytes) + ",\"us\":" +
			             std::to_string(std::chrono::duration_cast<std::chrono::microseconds>(
			                                std::chrono::steady_clock::now() - start)
			                                .count()));
		}
		uint64_t validity_bytes = 0;
		if (host.validity != nullptr) {
			const size_t packed_bytes = (static_cast<size_t>(host.row_count) + 7) / 8;
			auto *packed = static_cast<uint8_t *>(ctx.Alloc(packed_bytes));
			if (packed_bytes > 0) {
				ctx.staging.Upload(packed, host.validity, packed_bytes);
			}
			validity_bytes = host.row_count;
			column.valid = static_cast<uint8_t *>(
			    vgpu::DevicePool::Instance().Alloc(validity_bytes == 0 ? 1 : validity_bytes, ctx.stream));
			vgpu::UnpackValidity(packed, column.valid, host.row_count, ctx.stream);
		}
		bool cached = false;
		if (ctx.stage_scan_to_cache) {
			GpuColumnChunk chunk;
			chunk.device_data = column.data;
			chunk.device_validity = column.valid;
			chunk.row_count = column.rows;
			chunk.data_bytes = bytes;
			chunk.validity_bytes = validity_bytes;
			chunk.type = column.type;
			cached = GpuColumnCache::Instance().StageChunk(node.table, host.name, chunk);
		}
		if (!cached) {
			ctx.arena.Adopt(column.data);
			if (column.valid != nullptr) {
				ctx.arena.Adopt(column.valid);
			}
		}
		column.pruned = !Needed(need, host.name);
		columns.push_back(std::move(column));
	}
	return columns;
}

// ------------------------------------------------------------------ fused FILTER / PROJECTION kernels

//! One output of a fused kernel: an expression, or (`copy_of`) an unchanged copy of a column whose type no
//! generated kernel handles (HUGEINT), which is gathered instead.
struct NamedExpr {
	std::string name;
	const GpuExpr *expr = nullptr;
	const DeviceColumn *copy_of = nullptr;
};

//! Generated kernel(s) over dense `child` columns: AND of `preds` (a NULL predicate drops the row), then
//! every output of `outs` over the surviving rows, written densely. An output whose inputs are all
//! non-NULL carries no validity; otherwise its validity is the AND of the referenced inputs'. One kernel
//! when the argument tables allow (16 columns, 16 outputs, every output an expression); otherwise one
//! predicate kernel emits a row selection and the outputs are produced through it in groups.
Columns RunFused(const Columns &child, uint64_t rows, const std::vector<GpuExpr> &preds,
                 const std::vector<NamedExpr> &outs, ExecContext &ctx, uint64_t *kept_out = nullptr) {
	std::vector<const DeviceColumn *> slot_cols;
	std::map<std::string, int> slot_of;
	auto resolve = [&](const GpuExpr &e) {
		vgpu::FusedExpr f;
		f.body = e.generated_cuda_source;
		f.out_type = static_cast<int>(e.output_type);
		for (auto &name : e.input_columns) {
			const DeviceColumn &c = FindColumn(child, name);
			if (c.rows != rows) {
				throw std::runtime_error("gpu_executor: input column '" + name + "' row count mismatch");
			}
			auto it = slot_of.find(name);
			int slot;
			if (it == slot_of.end()) {
				slot = static_cast<int>(slot_cols.size());
				slot_of.emplace(name, slot);
				slot_cols.push_back(&c);
			} else {
				slot = it->second;
			}
			f.slots.push_back(slot);
		}
		return f;
	};
	std::vector<vgpu::FusedExpr> predicates;
	for (auto &p : preds) {
		predicates.push_back(resolve(p));
		if (p.output_type != GpuValueType::BOOLEAN) {
			throw std::runtime_error("gpu_executor: FILTER predicate did not evaluate to BOOLEAN");
		}
	}
	std::vector<vgpu::FusedExpr> projections(outs.size());
	bool all_expr = true;
	for (size_t e = 0; e < outs.size(); e++) {
		if (outs[e].expr != nullptr) {
			projections[e] = resolve(*outs[e].expr);
		} else {
			all_expr = false;
		}
	}

	Columns result;
	for (auto &o : outs) {
		DeviceColumn c;
		c.name = o.name;
		c.type = o.expr != nullptr ? o.expr->output_type : o.copy_of->type;
		result.push_back(std::move(c));
	}
	if (kept_out != nullptr) {
		*kept_out = 0;
	}
	if (rows == 0) {
		for (auto &c : result) {
			c.data = ctx.Alloc(0);
		}
		return result;
	}
	for (auto *c : slot_cols) {
		if (!FusableType(c->type)) {
			throw std::runtime_error("gpu_executor: column '" + c->name +
			                         "' has a type no generated expression can read");
		}
	}
	for (auto &o : outs) {
		if (o.expr != nullptr && !FusableType(o.expr->output_type)) {
			throw std::runtime_error("gpu_executor: expression output type is not supported by generated kernels");
		}
	}
	auto fused_col = [&](int slot) {
		vgpu::FusedColumn fc;
		fc.vtype = static_cast<int>(slot_cols[slot]->type);
		fc.has_valid = slot_cols[slot]->valid != nullptr;
		return fc;
	};
	auto nullable = [&](const vgpu::FusedExpr &f) {
		for (int s : f.slots) {
			if (slot_cols[s]->valid != nullptr) {
				return true;
			}
		}
		return false;
	};
	//! A kernel spec over a subset of the slots: re-slots `exprs` densely and returns the pointer tables.
	struct SubSpec {
		vgpu::FusedSpec spec;
		std::vector<const void *> ptrs;
		std::vector<const uint8_t *> valid;
		std::map<int, int> remap;
		void Add(std::vector<vgpu::FusedExpr> &dst, vgpu::FusedExpr f, const std::function<vgpu::FusedColumn(int)> &col,
		         const std::vector<const DeviceColumn *> &cols) {
			for (auto &s : f.slots) {
				auto it = remap.find(s);
				if (it == remap.end()) {
					const int ns = static_cast<int>(spec.cols.size());
					remap.emplace(s, ns);
					spec.cols.push_back(col(s));
					ptrs.push_back(cols[s]->data);
					valid.push_back(cols[s]->valid);
					s = ns;
				} else {
					s = it->second;
				}
			}
			dst.push_back(std::move(f));
		}
		size_t NewCols(const vgpu::FusedExpr &f) const {
			std::set<int> fresh;
			for (int s : f.slots) {
				if (!remap.count(s)) {
					fresh.insert(s);
				}
			}
			return fresh.size();
		}
	};

	const bool filtering = !predicates.empty();
	auto *count_dev = filtering ? static_cast<uint64_t *>(ctx.Alloc(8)) : nullptr;
	const size_t state_bytes = filtering ? vgpu::FusedKernel::StateBytes(rows) : 0;
	void *state = filtering ? ctx.Alloc(state_bytes) : nullptr;
	uint64_t kept = rows;
	auto read_kept = [&]() {
		VGPU_HIP(hipMemcpyAsync(&kept, count_dev, 8, hipMemcpyDeviceToHost, ctx.stream));
		VGPU_HIP(hipStreamSynchronize(ctx.stream));
	};

	if (all_expr && slot_cols.size() <= (size_t)vgpu::kFusedMaxCols && outs.size() <= (size_t)vgpu::kFusedMaxOuts) {
		vgpu::FusedSpec spec;
		spec.fast_math = ctx.fast_math;
		std::vector<const void *> ptrs;
		std::vector<const uint8_t *> valid;
		for (size_t s = 0; s < slot_cols.size(); s++) {
			spec.cols.push_back(fused_col(static_cast<int>(s)));
			ptrs.push_back(slot_cols[s]->data);
			valid.push_back(slot_cols[s]->valid);
		}
		spec.predicates = predicates;
		spec.projections = projections;
		auto kernel = vgpu::GetFusedKernel(spec);
		std::vector<void *> out_ptrs;
		std::vector<uint8_t *> out_valid;
		for (size_t e = 0; e < outs.size(); e++) {
			result[e].data = ctx.Alloc(rows * TypeSize(result[e].type));
			if (nullable(projections[e])) {
				result[e].valid = static_cast<uint8_t *>(ctx.Alloc(rows));
			}
			out_ptrs.push_back(result[e].data);
			out_valid.push_back(result[e].valid);
		}
		vgpu::FusedLaunch la;
		la.cols = ptrs.data();
		la.valid = valid.data();
		la.outs = out_ptrs.data();
		la.out_valid = out_valid.data();
		la.n = rows;
		la.count_device = count_dev;
		kernel->Launch(la, state, state_bytes, ctx.stream);
		if (filtering) {
			read_kept();
		}
		for (auto &c : result) {
			c.rows = kept;
		}
		if (kept_out != nullptr) {
			*kept_out = kept;
		}
		return result;
	}

	if (rows >= (1ull << 32)) {
		throw std::runtime_error("gpu_executor: a FILTER/PROJECTION of this width is limited to 2^32 input rows");
	}
	const uint32_t *sel = nullptr;
	if (filtering) {
		SubSpec ps;
		ps.spec.fast_math = ctx.fast_math;
		ps.spec.emit_selection = true;
		for (auto &f : predicates) {
			ps.Add(ps.spec.predicates, f, fused_col, slot_cols);
		}
		if (ps.spec.cols.size() > (size_t)vgpu::kFusedMaxCols) {
			throw std::runtime_error("gpu_executor: FILTER predicates read more than 16 columns");
		}
		auto *selection = static_cast<uint32_t *>(ctx.Alloc(rows * 4));
		vgpu::FusedLaunch la;
		la.cols = ps.ptrs.data();
		la.valid = ps.valid.data();
		la.n = rows;
		la.sel_out = selection;
		la.count_device = count_dev;
		vgpu::GetFusedKernel(ps.spec)->Launch(la, state, state_bytes, ctx.stream);
		read_kept();
		sel = selection;
	}
	for (auto &c : result) {
		c.rows = kept;
		c.data = ctx.Alloc(kept * TypeSize(c.type));
	}
	if (kept_out != nullptr) {
		*kept_out = kept;
	}
	if (kept == 0) {
		return result;
	}
	// Unchanged copies of columns no generated kernel reads.
	for (size_t e = 0; e < outs.size(); e++) {
		const DeviceColumn *src = outs[e].copy_of;
		if (src == nullptr) {
			continue;
		}
		const int width = static_cast<int>(TypeSize(src->type));
		if (sel != nullptr) {
			vgpu::Gather(src->data, width, sel, false, kept, result[e].data, ctx.stream);
		} else {
			vgpu::CopyDevice(result[e].data, src->data, kept * width, ctx.stream);
		}
		if (src->valid != nullptr) {
			result[e].valid = static_cast<uint8_t *>(ctx.Alloc(kept));
			if (sel != nullptr) {
				vgpu::Gather(src->valid, 1, sel, false, kept, result[e].valid, ctx.stream);
			} else {
				vgpu::CopyDevice(result[e].valid, src->valid, kept, ctx.stream);
			}
		}
	}
	// Expression outputs in groups that fit the argument tables.
	for (size_t e = 0; e < outs.size();) {
		if (outs[e].expr == nullptr) {
			e++;
			continue;
		}
		SubSpec os;
		os.spec.fast_math = ctx.fast_math;
		os.spec.selection_in = sel != nullptr;
		std::vector<void *> out_ptrs;
		std::vector<uint8_t *> out_valid;
		for (; e < outs.size() && os.spec.projections.size() < (size_t)vgpu::kFusedMaxOuts; e++) {
			if (outs[e].expr == nullptr) {
				continue;
			}
			if (os.spec.cols.size() + os.NewCols(projections[e]) > (size_t)vgpu::kFusedMaxCols) {
				if (os.spec.projections.empty()) {
					throw std::runtime_error("gpu_executor: an expression reads more than 16 columns");
				}
				break;
			}
			os.Add(os.spec.projections, projections[e], fused_col, slot_cols);
			if (nullable(projections[e])) {
				result[e].valid = static_cast<uint8_t *>(ctx.Alloc(kept));
			}
			out_ptrs.push_back(result[e].data);
			out_valid.push_back(result[e].valid);
		}
		vgpu::FusedLaunch la;
		la.cols = os.ptrs.data();
		la.valid = os.valid.data();
		la.outs = out_ptrs.data();
		la.out_valid = out_valid.data();
		la.n = kept;
		la.sel = sel;
		vgpu::GetFusedKernel(os.spec)->Launch(la, nullptr, 0, ctx.stream);
	}
	return result;
}

// ------------------------------------------------------------------ plan walk

Columns ExecuteNode(const GpuPlanNode &node, ExecContext &ctx, const NameSet *need);

//! CROSS_PRODUCT, optionally with the FILTER (`preds`) and/or PROJECTION (`projection`) directly above it
//! fused into the pair enumeration (cuda_engine/src/operators/gpu_cross_join.cu:111-192 for the
//! operator's own checks and left-major order).
Columns ExecuteCross(const GpuPlanNode &node, ExecContext &ctx, const NameSet *need, const std::vector<GpuExpr> *preds,
                     const GpuPlanNode *projection) {
	if (node.children.size() != 2) {
		throw std::runtime_error("gpu_executor: CROSS_PRODUCT expects exactly two children");
	}
	// Names the fused consumer reads from the cross product's outputs.
	NameSet reads;
	if (projection != nullptr) {
		AddInputs(reads, projection->expressions);
	} else {
		for (auto &name : node.output_columns) {
			if (Needed(need, name)) {
				reads.insert(name);
			}
		}
	}
	if (preds != nullptr) {
		AddInputs(reads, *preds);
	}
	NameSet side_need;
	for (auto &name : node.output_columns) {
		if (reads.count(name)) {
			side_need.insert(name);
		}
	}
	Columns left = ExecuteNode(*node.children[0], ctx, &side_need);
	Columns right = ExecuteNode(*node.children[1], ctx, &side_need);

	if (node.output_columns.empty()) {
		throw std::runtime_error("gpu_cross_join: node has no output columns");
	}
	if (left.empty() || right.empty()) {
		throw std::runtime_error("gpu_cross_join: both inputs must produce at least one column");
	}
	std::vector<const DeviceColumn *> sources;
	std::vector<bool> from_left;
	for (auto &name : node.output_columns) {
		bool l = false;
		sources.push_back(&ResolveSideColumn(name, left, right, l));
		from_left.push_back(l);
		if (sources.back()->valid != nullptr) {
			throw std::runtime_error("gpu_cross_join: column '" + name + "' contains NULLs, which this operator "
			                         "does not model");
		}
	}
	const uint64_t left_rows = left.front().rows;
	const uint64_t right_rows = right.front().rows;
	for (auto &c : left) {
		if (c.rows != left_rows) {
			throw std::runtime_error("gpu_cross_join: left input column '" + c.name + "' row count mismatch");
		}
	}
	for (auto &c : right) {
		if (c.rows != right_rows) {
			throw std::runtime_error("gpu_cross_join: right input column '" + c.name + "' row count mismatch");
		}
	}
	constexpr uint64_t kMaxCrossProductOutputRows = 1ull << 32;
	uint64_t total = 0;
	if (left_rows != 0 && right_rows != 0) {
		if (left_rows > kMaxCrossProductOutputRows / right_rows) {
			throw std::runtime_error("gpu_cross_join: cross product of " + std::to_string(left_rows) + " x " +
			                         std::to_string(right_rows) + " rows is too large for GPU execution");
		}
		total = left_rows * right_rows;
		if (total > kMaxCrossProductOutputRows) {
			throw std::runtime_error("gpu_cross_join: cross product result of " + std::to_string(total) +
			                         " rows is too large for GPU execution");
		}
	}

	// The cross product's own output columns (what a FILTER or PROJECTION above it sees).
	Columns cross_out;
	for (size_t i = 0; i < sources.size(); i++) {
		DeviceColumn c = *sources[i];
		c.name = node.output_columns[i];
		c.rows = total;
		cross_out.push_back(c);
	}

	std::vector<NamedExpr> outs;
	std::vector<GpuExpr> identity;
	std::vector<std::string> proj_names;
	if (projection != nullptr) {
		for (size_t i = 0; i < projection->expressions.size(); i++) {
			proj_names.push_back("col" + std::to_string(i));
		}
		for (size_t i = 0; i < projection->expressions.size(); i++) {
			outs.push_back({proj_names[i], &projection->expressions[i]});
		}
	} else if (preds != nullptr) {
		identity.reserve(cross_out.size());
		for (auto &c : cross_out) {
			if (!c.pruned && Needed(need, c.name)) {
				identity.push_back(IdentityExpr(c));
			}
		}
		size_t k = 0;
		for (auto &c : cross_out) {
			if (!c.pruned && Needed(need, c.name)) {
				outs.push_back({c.name, &identity[k++]});
			}
		}
	}

	if (total == 0) {
		if (projection != nullptr) {
			Columns r;
			for (size_t i = 0; i < projection->expressions.size(); i++) {
				r.push_back(EmptyColumn(proj_names[i], projection->expressions[i].output_type, ctx));
			}
			return r;
		}
		Columns r;
		for (auto &c : cross_out) {
			r.push_back(c.pruned || !Needed(need, c.name) ? PrunedLike(c, 0) : EmptyColumn(c.name, c.type, ctx));
		}
		return r;
	}

	if (preds == nullptr && projection == nullptr) {
		// Plain expansion, one write-bound pass per requested column.
		Columns r;
		for (size_t i = 0; i < cross_out.size(); i++) {
			const DeviceColumn &c = cross_out[i];
			if (c.pruned || !Needed(need, c.name)) {
				r.push_back(PrunedLike(c, total));
				continue;
			}
			DeviceColumn o;
			o.name = c.name;
			o.type = c.type;
			o.rows = total;
			o.data = ctx.Alloc(total * TypeSize(c.type));
			vgpu::CrossExpand(c.data, static_cast<int>(TypeSize(c.type)), right_rows, from_left[i], total, o.data,
			                  ctx.stream);
			r.push_back(std::move(o));
		}
		return r;
	}

	// Fused pair enumeration: slot per distinct source column, tagged with its side.
	vgpu::FusedSpec spec;
	spec.cross = true;
	spec.fast_math = ctx.fast_math;
	std::vector<const void *> ptrs;
	std::map<std::string, int> slot_of;
	auto resolve = [&](const GpuExpr &e) {
		vgpu::FusedExpr f;
		f.body = e.generated_cuda_source;
		f.out_type = static_cast<int>(e.output_type);
		for (auto &name : e.input_columns) {
			size_t idx = 0;
			for (; idx < cross_out.size(); idx++) {
				if (cross_out[idx].name == name) {
					break;
				}
			}
			if (idx == cross_out.size()) {
				throw std::runtime_error("gpu_executor: expression references column '" + name +
				                         "' which is not produced by its input operator");
			}
			const DeviceColumn &c = cross_out[idx];
			if (!FusableType(c.type)) {
				throw std::runtime_error("gpu_executor: column '" + name +
				                         "' has a type no generated expression can read");
			}
			auto it = slot_of.find(name);
			int slot;
			if (it == slot_of.end()) {
				slot = static_cast<int>(spec.cols.size());
				slot_of.emplace(name, slot);
				vgpu::FusedColumn fc;
				fc.vtype = static_cast<int>(c.type);
				fc.side = from_left[idx] ? 0 : 1;
				spec.cols.push_back(fc);
				ptrs.push_back(c.data);
			} else {
				slot = it->second;
			}
			f.slots.push_back(slot);
		}
		return f;
	};
	if (preds != nullptr) {
		for (auto &p : *preds) {
			auto f = resolve(p);
			if (p.output_type != GpuValueType::BOOLEAN) {
				throw std::runtime_error("gpu_executor: FILTER predicate did not evaluate to BOOLEAN");
			}
			spec.predicates.push_back(std::move(f));
		}
	}
	for (auto &o : outs) {
		if (!FusableType(o.expr->output_type)) {
			throw std::runtime_error("gpu_executor: expression output type is not supported by generated kernels");
		}
		spec.projections.push_back(resolve(*o.expr));
	}
	if (spec.cols.size() > (size_t)vgpu::kFusedMaxCols || spec.projections.size() > (size_t)vgpu::kFusedMaxOuts) {
		throw std::runtime_error("gpu_executor: a fused cross product reads or writes more than 16 columns");
	}
	const bool filtering = !spec.predicates.empty();
	uint64_t row_bytes = 0;
	for (auto &o : outs) {
		row_bytes += TypeSize(o.expr->output_type);
	}
	auto *count_dev = filtering ? static_cast<uint64_t *>(ctx.Alloc(8)) : nullptr;
	size_t state_bytes = filtering ? vgpu::FusedKernel::StateBytes(total) : 0;
	void *state = filtering ? ctx.Alloc(state_bytes) : nullptr;

	uint64_t capacity = total;
	if (filtering && row_bytes > 0 && total * row_bytes > (8ull << 30)) {
		// Outputs at the full pair count would be large: count the survivors first.
		vgpu::FusedSpec cspec = spec;
		cspec.projections.clear();
		auto ck = vgpu::GetFusedKernel(cspec);
		vgpu::FusedLaunch la;
		la.cols = ptrs.data();
		la.n = total;
		la.right_rows = right_rows;
		la.count_device = count_dev;
		ck->Launch(la, state, state_bytes, ctx.stream);
		VGPU_HIP(hipMemcpyAsync(&capacity, count_dev, 8, hipMemcpyDeviceToHost, ctx.stream));
		VGPU_HIP(hipStreamSynchronize(ctx.stream));
	}
	Columns r;
	std::vector<void *> out_ptrs;
	for (auto &o : outs) {
		DeviceColumn c;
		c.name = o.name;
		c.type = o.expr->output_type;
		c.data = ctx.Alloc(capacity * TypeSize(c.type));
		out_ptrs.push_back(c.data);
		r.push_back(std::move(c));
	}
	auto kernel = vgpu::GetFusedKernel(spec);
	vgpu::FusedLaunch la;
	la.cols = ptrs.data();
	la.outs = out_ptrs.data();
	la.n = total;
	la.right_rows = right_rows;
	la.count_device = count_dev;
	la.capacity = capacity;
	kernel->Launch(la, state, state_bytes, ctx.stream);
	uint64_t kept = total;
	if (filtering) {
		VGPU_HIP(hipMemcpyAsync(&kept, count_dev, 8, hipMemcpyDeviceToHost, ctx.stream));
		VGPU_HIP(hipStreamSynchronize(ctx.stream));
		if (kept > capacity) {
			throw std::logic_error("gpu_executor: filtered cross product exceeded its counted capacity");
		}
	}
	for (auto &c : r) {
		c.rows = kept;
	}
	if (projection == nullptr) {
		// FILTER output: every cross output column, the unread ones pruned.
		Columns full;
		size_t k = 0;
		for (auto &c : cross_out) {
			if (!c.pruned && Needed(need, c.name)) {
				full.push_back(r[k++]);
			} else {
				full.push_back(PrunedLike(c, kept));
			}
		}
		return full;
	}
	return r;
}

//! FILTER, optionally fused with the PROJECTION directly above it (`projection`).
Columns ExecuteFilter(const GpuPlanNode &node, ExecContext &ctx, const NameSet *need, const GpuPlanNode *projection) {
	if (node.children.size() != 1) {
		throw std::runtime_error("gpu_executor: FILTER expects exactly one child");
	}
	const GpuPlanNode &child = *node.children[0];
	if (child.op_type == GpuOpType::CROSS_PRODUCT && !node.expressions.empty()) {
		return ExecuteCross(child, ctx, need, &node.expressions, projection);
	}
	NameSet child_need;
	const NameSet *child_need_ptr = nullptr;
	if (projection != nullptr) {
		AddInputs(child_need, projection->expressions);
		AddInputs(child_need, node.expressions);
		child_need_ptr = &child_need;
	} else if (need != nullptr) {
		child_need = *need;
		AddInputs(child_need, node.expressions);
		child_need_ptr = &child_need;
	}
	Columns cols = ExecuteNode(child, ctx, child_need_ptr);
	const uint64_t rows = cols.empty() ? 0 : cols.front().rows;
	const bool filter_applies = !cols.empty() && !node.expressions.empty() && rows != 0;
	static const std::vector<GpuExpr> kNoPredicates;

	if (projection != nullptr) {
		std::vector<NamedExpr> outs;
		for (size_t i = 0; i < projection->expressions.size(); i++) {
			outs.push_back({"col" + std::to_string(i), &projection->expressions[i]});
		}
		return RunFused(cols, rows, filter_applies ? node.expressions : kNoPredicates, outs, ctx);
	}
	if (!filter_applies) {
		return cols;
	}
	std::vector<GpuExpr> identity;
	identity.reserve(cols.size());
	std::vector<NamedExpr> outs;
	for (auto &c : cols) {
		if (c.pruned || !Needed(need, c.name)) {
			continue;
		}
		if (FusableType(c.type)) {
			identity.push_back(IdentityExpr(c));
			outs.push_back({c.name, &identity.back(), nullptr});
		} else {
			outs.push_back({c.name, nullptr, &c});
		}
	}
	uint64_t kept = 0;
	Columns carried = RunFused(cols, rows, node.expressions, outs, ctx, &kept);
	Columns result;
	size_t j = 0;
	for (auto &c : cols) {
		if (!c.pruned && Needed(need, c.name)) {
			result.push_back(std::move(carried[j++]));
		} else {
			result.push_back(PrunedLike(c, kept));
		}
	}
	return result;
}

Columns ExecuteProjection(const GpuPlanNode &node, ExecContext &ctx) {
	if (node.children.size() != 1) {
		throw std::runtime_error("gpu_executor: PROJECTION expects exactly one child");
	}
	const GpuPlanNode &child = *node.children[0];
	if (child.op_type == GpuOpType::FILTER) {
		return ExecuteFilter(child, ctx, nullptr, &node);
	}
	if (child.op_type == GpuOpType::CROSS_PRODUCT) {
		return ExecuteCross(child, ctx, nullptr, nullptr, &node);
	}
	NameSet need;
	AddInputs(need, node.expressions);
	Columns cols = ExecuteNode(child, ctx, &need);
	const uint64_t rows = cols.empty() ? 0 : cols.front().rows;
	std::vector<NamedExpr> outs;
	for (size_t i = 0; i < node.expressions.size(); i++) {
		outs.push_back({"col" + std::to_string(i), &node.expressions[i]});
	}
	return RunFused(cols, rows, {}, outs, ctx);
}

// ------------------------------------------------------------------ GROUP BY

void CheckGroupKey(const DeviceColumn &column, const std::string &key_name, uint64_t rows) {
	if (column.rows != rows) {
		throw std::runtime_error("gpu_groupby: group key '" + key_name + "' row count mismatch");
	}
	switch (column.type) {
	case GpuValueType::INT16:
	case GpuValueType::INT32:
	case GpuValueType::INT64:
	case GpuValueType::BOOLEAN:
		return;
	default:
		throw std::runtime_error("gpu_groupby: group key '" + key_name +
		                         "' is not an integer/boolean column, which this GPU path requires");
	}
}

[[noreturn]] void ThrowGlobalEmpty() {
	throw std::runtime_error("gpu_groupby: global aggregate over an empty input needs NULL results, "
	                         "which this GPU path cannot represent");
}

void CheckGroupByShape(const GpuPlanNode &node) {
	if (node.group_keys.empty() && node.aggregates.empty()) {
		throw std::runtime_error("gpu_groupby: node has neither group keys nor aggregates");
	}
	if (node.aggregates.size() != node.aggregate_exprs.size() && !node.aggregate_exprs.empty()) {
		throw std::runtime_error("gpu_groupby: aggregates and aggregate_exprs disagree in length");
	}
}

//! The reference's aggregate rules (cuda_engine/src/operators/gpu_groupby.cu:224-274) with one
//! extension: SUM over an integer column produces HUGEINT, DuckDB's own result type (the reference refuses
//! it). MIN/MAX stay integer/boolean only, SUM over floating input produces DOUBLE.
void ValidateEngineAggregate(const GpuAggregate &aggregate, GpuValueType input) {
	const bool is_float = input == GpuValueType::FLOAT32 || input == GpuValueType::FLOAT64;
	const bool is_int = input == GpuValueType::INT16 || input == GpuValueType::INT32 || input == GpuValueType::INT64;
	switch (aggregate.kind) {
	case GpuAggregateKind::SUM:
		if (is_int) {
			if (aggregate.output_type != GpuValueType::HUGEINT) {
				throw std::runtime_error("gpu_groupby: SUM over an integer column must produce HUGEINT (DuckDB's "
				                         "result type)");
			}
			return;
		}
		if (!is_float || aggregate.output_type != GpuValueType::FLOAT64) {
			throw std::runtime_error("gpu_groupby: SUM on this GPU path supports only FLOAT/DOUBLE input producing "
			                         "DOUBLE, or integer input producing HUGEINT");
		}
		return;
	case GpuAggregateKind::MIN:
	case GpuAggregateKind::MAX:
		if (aggregate.output_type != input) {
			throw std::runtime_error("gpu_groupby: MIN/MAX output type must match its input column type");
		}
		if (!is_int && input != GpuValueType::BOOLEAN) {
			throw std::runtime_error("gpu_groupby: MIN/MAX on this GPU path supports only integer/boolean columns");
		}
		return;
	default:
		return;
	}
}

//! Aggregate requests over `value_cols` (the distinct aggregate inputs, resolved in `cols`), validated
//! with the GROUP BY rules.
std::vector<vgpu::AggReq> BuildAggRequests(const GpuPlanNode &node, const Columns &cols, uint64_t rows,
                                           std::vector<vgpu::DevCol> &values) {
	std::vector<vgpu::AggReq> reqs;
	std::map<std::string, int> value_of;
	for (auto &aggregate : node.aggregates) {
		vgpu::AggReq r;
		r.kind = static_cast<int>(aggregate.kind);
		r.out_type = static_cast<int>(aggregate.output_type);
		if (aggregate.kind != GpuAggregateKind::COUNT_STAR) {
			const DeviceColumn &c = FindColumn(cols, aggregate.input_column);
			if (c.rows != rows) {
				throw std::runtime_error("gpu_groupby: aggregate input column '" + aggregate.input_column +
				                         "' row count mismatch");
			}
			auto it = value_of.find(aggregate.input_column);
			if (it == value_of.end()) {
				vgpu::DevCol d;
				d.data = c.data;
				d.type = static_cast<int>(c.type);
				d.rows = c.rows;
				it = value_of.emplace(aggregate.input_column, static_cast<int>(values.size())).first;
				values.push_back(d);
			}
			r.input = it->second;
			ValidateEngineAggregate(aggregate, c.type);
		}
		vgpu::ValidateAggregateReq(r, values);
		reqs.push_back(r);
	}
	return reqs;
}

Columns ExecuteJoinAggregate(const GpuPlanNode &node, const GpuPlanNode &join, ExecContext &ctx);

bool JoinAggregateFusable(const GpuPlanNode &node) {
	if (!node.group_keys.empty() || node.aggregates.empty() || node.children.size() != 1 ||
	    node.children[0]->op_type != GpuOpType::HASH_JOIN) {
		return false;
	}
	for (auto &a : node.aggregates) {
		// MIN/MAX over a join run on 8-byte states only; the other widths take the materializing path.
		if (a.kind == GpuAggregateKind::MIN || a.kind == GpuAggregateKind::MAX) {
			if (a.output_type != GpuValueType::INT64 && a.output_type != GpuValueType::FLOAT64) {
				return false;
			}
		}
	}
	return node.aggregates.size() < 8;
}

bool IsFloatType(GpuValueType t) {
	return t == GpuValueType::FLOAT32 || t == GpuValueType::FLOAT64;
}

//! Bits of the value MIN (max = false) or MAX (max = true) must never prefer over a real one: the type's
//! largest / smallest value (NaN is DuckDB's largest floating value, -inf its smallest).
uint64_t ExtremeBits(GpuValueType t, bool largest) {
	switch (t) {
	case GpuValueType::INT16:
		return largest ? 0x7FFFull : 0x8000ull;
	case GpuValueType::INT32:
		return largest ? 0x7FFFFFFFull : 0x80000000ull;
	case GpuValueType::INT64:
		return largest ? 0x7FFFFFFFFFFFFFFFull : 0x8000000000000000ull;
	case GpuValueType::BOOLEAN:
		return largest ? 1 : 0;
	case GpuValueType::FLOAT32:
		return largest ? 0x7FC00000ull : 0xFF800000ull;
	case GpuValueType::FLOAT64:
		return largest ? 0x7FF8000000000000ull : 0xFFF0000000000000ull;
	default:
		throw std::runtime_error("gpu_groupby: no MIN/MAX over this column type");
	}
}

//! GROUP BY with the extended semantics (gpu_semantics.hpp): nullable keys become an "is NULL" key column
//! plus the key with NULLs zeroed; FLOAT/DOUBLE keys become their ordered integer image; an aggregate over
//! a nullable column reads a copy with NULLs replaced by the aggregate's neutral value, and a hidden
//! SUM of the column's validity (as INT32) counts its non-NULL values, from which its validity (and AVG)
//! is completed after the reduction. Columns without NULLs are read as they are.
Columns GroupByExtended(const GpuPlanNode &node, const Columns &cols, uint64_t rows, ExecContext &ctx) {
	std::vector<vgpu::DevCol> keys;
	struct KeyOut {
		GpuValueType type;
		int value_slot = -1;
		int null_slot = -1;          // "is NULL" key column of a nullable integer key
		uint64_t null_sentinel = 0;  // a nullable floating key: NULL is this ordered value (above NaN's)
		bool has_sentinel = false;
	};
	std::vector<KeyOut> key_info;
	for (auto &key_name : node.group_keys) {
		const DeviceColumn &c = FindColumn(cols, key_name);
		if (c.rows != rows) {
			throw std::runtime_error("gpu_groupby: group key '" + key_name + "' row count mismatch");
		}
		const bool is_float = IsFloatType(c.type);
		if (!is_float && c.type != GpuValueType::INT16 && c.type != GpuValueType::INT32 &&
		    c.type != GpuValueType::INT64 && c.type != GpuValueType::BOOLEAN) {
			throw std::runtime_error("gpu_groupby: group key '" + key_name +
			                         "' is not an integer, boolean or floating column");
		}
		KeyOut k;
		k.type = c.type;
		if (c.valid != nullptr && !is_float) {
			auto *flags = static_cast<uint8_t *>(ctx.Alloc(rows));
			vgpu::NullFlags(c.valid, rows, flags, ctx.stream);
			k.null_slot = static_cast<int>(keys.size());
			keys.push_back({flags, vgpu::VT_BOOLEAN, rows});
		}
		const int width = static_cast<int>(TypeSize(c.type));
		int key_type = static_cast<int>(c.type);
		const void *data = c.data;
		if (is_float) {
			void *ordered = ctx.Alloc(rows * width);
			vgpu::FloatKeyToOrdered(c.data, static_cast<int>(c.type), rows, ordered, ctx.stream);
			data = ordered;
			key_type = c.type == GpuValueType::FLOAT64 ? vgpu::VT_INT64 : vgpu::VT_INT32;
		}
		if (c.valid != nullptr) {
			// NULL -> 0 beside the "is NULL" column; a floating key's NULL takes the largest ordered value, which
			// no canonical value has (every NaN orders just above +inf), so it needs no extra column
			if (is_float) {
				k.has_sentinel = true;
				k.null_sentinel = width == 8 ? 0x7FFFFFFFFFFFFFFFull : 0x7FFFFFFFull;
			}
			void *masked = ctx.Alloc(rows * width);
			vgpu::MaskNulls(data, c.valid, width, k.null_sentinel, rows, masked, ctx.stream);
			data = masked;
		}
		k.value_slot = static_cast<int>(keys.size());
		keys.push_back({data, key_type, rows});
		key_info.push_back(k);
	}
	if (keys.size() > 4) {
		throw std::runtime_error("gpu_groupby: at most 4 packed key columns (a nullable key takes two)");
	}

	// Aggregate requests over value columns; nullable inputs get masked copies and a validity count.
	std::vector<vgpu::DevCol> values;
	std::map<std::pair<std::string, uint64_t>, int> value_of; // (column, fill or ~0 for as-is)
	std::map<std::string, int> count_req_of;                  // column -> hidden SUM(valid) request
	std::vector<vgpu::AggReq> reqs;
	auto value_slot = [&](const DeviceColumn &c, bool masked, uint64_t fill, GpuValueType type) {
		const auto key = std::make_pair(c.name, masked ? fill : ~0ull);
		auto it = value_of.find(key);
		if (it != value_of.end()) {
			return it->second;
		}
		const void *data = c.data;
		if (masked) {
			void *m = ctx.Alloc(rows * TypeSize(c.type));
			vgpu::MaskNulls(c.data, c.valid, static_cast<int>(TypeSize(c.type)), fill, rows, m, ctx.stream);
			data = m;
		}
		values.push_back({data, static_cast<int>(type), rows});
		const int slot = static_cast<int>(values.size()) - 1;
		value_of.emplace(key, slot);
		return slot;
	};
	struct Fix {
		size_t agg = 0;     // index into node.aggregates
		int count_req = -1; // hidden request holding the non-NULL count
		int sum_req = -1;   // AVG: hidden request holding the SUM
	};
	std::vector<Fix> fixes;
	std::vector<int> out_req(node.aggregates.size(), -1);
	for (size_t i = 0; i < node.aggregates.size(); i++) {
		const auto &a = node.aggregates[i];
		vgpu::AggReq r;
		r.kind = static_cast<int>(a.kind);
		r.out_type = static_cast<int>(a.output_type);
		if (a.kind == GpuAggregateKind::COUNT_STAR) {
			out_req[i] = static_cast<int>(reqs.size());
			reqs.push_back(r);
			continue;
		}
		const DeviceColumn &c = FindColumn(cols, a.input_column);
		if (c.rows != rows) {
			throw std::runtime_error("gpu_groupby: aggregate input column '" + a.input_column + "' row count mismatch");
		}
		if (c.valid == nullptr) {
			r.input = value_slot(c, false, 0, c.type);
			vgpu::ValidateAggregateReq(r, values);
			out_req[i] = static_cast<int>(reqs.size());
			reqs.push_back(r);
			continue;
		}
		auto count_req = [&]() {
			auto it = count_req_of.find(c.name);
			if (it != count_req_of.end()) {
				return it->second;
			}
			auto *flags = static_cast<int32_t *>(ctx.Alloc(rows * 4));
			vgpu::ValidAsInt32(c.valid, rows, flags, ctx.stream);
			values.push_back({flags, vgpu::VT_INT32, rows});
			vgpu::AggReq cr;
			cr.kind = vgpu::AGG_SUM;
			cr.input = static_cast<int>(values.size()) - 1;
			cr.out_type = vgpu::VT_INT64;
			const int idx = static_cast<int>(reqs.size());
			reqs.push_back(cr);
			count_req_of.emplace(c.name, idx);
			return idx;
		};
		switch (a.kind) {
		case GpuAggregateKind::COUNT:
			out_req[i] = count_req(); // the non-NULL count is the result (never NULL)
			if (a.output_type != GpuValueType::INT64) {
				throw std::runtime_error("gpu_groupby: COUNT must produce INT64");
			}
			continue;
		case GpuAggregateKind::SUM:
		case GpuAggregateKind::AVG:
			r.input = value_slot(c, true, 0, c.type);
			break;
		case GpuAggregateKind::MIN:
			r.input = value_slot(c, true, ExtremeBits(c.type, true), c.type);
			break;
		case GpuAggregateKind::MAX:
			r.input = value_slot(c, true, ExtremeBits(c.type, false), c.type);
			break;
		default:
			throw std::runtime_error("gpu_groupby: unknown aggregate kind");
		}
		Fix f;
		f.agg = i;
		if (a.kind == GpuAggregateKind::AVG) {
			// SUM then divide by the non-NULL count
			vgpu::AggReq sr;
			sr.kind = vgpu::AGG_SUM;
			sr.input = r.input;
			sr.out_type = IsFloatType(c.type) ? vgpu::VT_FLOAT64 : vgpu::VT_HUGEINT;
			vgpu::ValidateAggregateReq(r, values);
			out_req[i] = static_cast<int>(reqs.size());
			reqs.push_back(r); // the AVG itself only reserves the DOUBLE output column
			f.sum_req = static_cast<int>(reqs.size());
			reqs.push_back(sr);
		} else {
			vgpu::ValidateAggregateReq(r, values);
			out_req[i] = static_cast<int>(reqs.size());
			reqs.push_back(r);
		}
		f.count_req = count_req();
		fixes.push_back(f);
	}

	const bool global = node.group_keys.empty();
	if (rows == 0 && !global) {
		Columns r;
		for (size_t k = 0; k < key_info.size(); k++) {
			r.push_back(EmptyColumn(node.group_keys[k], key_info[k].type, ctx));
		}
		for (size_t i = 0; i < node.aggregates.size(); i++) {
			r.push_back(EmptyColumn("agg_" + std::to_string(i), node.aggregates[i].output_type, ctx));
		}
		return r;
	}
	if (rows == 0) {
		// keyless aggregate over no rows: one row, COUNT 0 and every other aggregate NULL
		Columns r;
		for (size_t i = 0; i < node.aggregates.size(); i++) {
			const auto &a = node.aggregates[i];
			DeviceColumn c;
			c.name = "agg_" + std::to_string(i);
			c.type = a.output_type;
			c.rows = 1;
			c.data = ctx.Alloc(16);
			vgpu::FillU64(static_cast<uint64_t *>(c.data), 0, 2, ctx.stream);
			if (a.kind != GpuAggregateKind::COUNT_STAR && a.kind != GpuAggregateKind::COUNT) {
				c.valid = static_cast<uint8_t *>(ctx.Alloc(8));
				vgpu::FillU64(reinterpret_cast<uint64_t *>(c.valid), 0, 1, ctx.stream);
			}
			r.push_back(std::move(c));
		}
		return r;
	}
	vgpu::KeyPack pack;
	const uint32_t *hist8 = nullptr;
	const uint64_t bound = vgpu::PlanGroupBy(keys, rows, pack, ctx.arena, ctx.stream, &hist8);
	vgpu::SortTuning tuning = vgpu::TuneSort(pack.total_bits, rows);
	tuning.hist8 = hist8;
	vgpu::OutAlloc alloc = [&](size_t bytes) { return ctx.Alloc(bytes); };
	// The reduction takes at most 4 distinct payload columns and 8 distinct component states per call; the
	// requests are split into runs that fit, each run groups the same keys (so its groups line up with the
	// other runs'), and the masked copies and validity counts make several runs common here.
	std::vector<std::vector<int>> runs(1);
	{
		std::set<int> payloads;
		std::set<std::pair<int, int>> comps;
		for (int q = 0; q < static_cast<int>(reqs.size()); q++) {
			const auto &rq = reqs[q];
			const bool counting = rq.kind == vgpu::AGG_COUNT_STAR || rq.kind == vgpu::AGG_COUNT;
			const int cls = rq.kind == vgpu::AGG_MIN ? 1 : (rq.kind == vgpu::AGG_MAX ? 2 : 0);
			const int newp = !counting && !payloads.count(rq.input) ? 1 : 0;
			const int newc = !counting && !comps.count({cls, rq.input}) ? 1 : 0;
			if (!runs.back().empty() && (payloads.size() + newp > 4 || 1 + comps.size() + newc > 8 ||
			                             runs.back().size() >= 16)) {
				runs.emplace_back();
				payloads.clear();
				comps.clear();
			}
			if (!counting) {
				payloads.insert(rq.input);
				comps.insert({cls, rq.input});
			}
			runs.back().push_back(q);
		}
	}
	std::vector<void *> key_out, agg_out(reqs.size(), nullptr);
	uint64_t groups = 0;
	for (size_t ri = 0; ri < runs.size(); ri++) {
		std::vector<vgpu::DevCol> run_values;
		std::map<int, int> local;
		std::vector<vgpu::AggReq> run_reqs;
		for (int q : runs[ri]) {
			vgpu::AggReq rq = reqs[q];
			if (rq.input >= 0) {
				auto it = local.find(rq.input);
				if (it == local.end()) {
					it = local.emplace(rq.input, static_cast<int>(run_values.size())).first;
					run_values.push_back(values[rq.input]);
				}
				rq.input = it->second;
			}
			run_reqs.push_back(rq);
		}
		std::vector<void *> run_keys, run_aggs;
		const uint64_t g = vgpu::GroupByDevice(keys, run_values, run_reqs, rows, pack, std::max<uint64_t>(bound, 1),
		                                       alloc, run_keys, run_aggs, tuning, ctx.arena, ctx.stream, true);
		if (ri == 0) {
			groups = g;
			key_out = run_keys;
		} else if (g != groups) {
			throw std::logic_error("gpu_groupby: aggregate runs over the same keys disagree on the group count");
		}
		for (size_t i = 0; i < runs[ri].size(); i++) {
			agg_out[runs[ri][i]] = run_aggs[i];
		}
	}
	if (groups == 0) {
		throw std::runtime_error("gpu_groupby: computed zero groups for a non-empty input");
	}
	Columns r;
	for (size_t k = 0; k < key_info.size(); k++) {
		const KeyOut &ki = key_info[k];
		DeviceColumn c;
		c.name = node.group_keys[k];
		c.type = ki.type;
		c.rows = groups;
		c.data = key_out[ki.value_slot];
		if (IsFloatType(ki.type)) {
			void *f = ctx.Alloc(groups * TypeSize(ki.type));
			vgpu::OrderedToFloatKey(c.data, static_cast<int>(ki.type), groups, f, ctx.stream);
			c.data = f;
		}
		if (ki.null_slot >= 0) {
			c.valid = static_cast<uint8_t *>(ctx.Alloc(groups));
			vgpu::ValidFromNullFlags(static_cast<const uint8_t *>(key_out[ki.null_slot]), groups, c.valid, ctx.stream);
		}
		if (ki.has_sentinel) {
			c.valid = static_cast<uint8_t *>(ctx.Alloc(groups));
			vgpu::ValidUnlessEqual(key_out[ki.value_slot], static_cast<int>(TypeSize(ki.type)), ki.null_sentinel, groups,
			                       c.valid, ctx.stream);
		}
		r.push_back(std::move(c));
	}
	std::vector<uint8_t *> agg_valid(node.aggregates.size(), nullptr);
	for (auto &f : fixes) {
		const auto &a = node.aggregates[f.agg];
		agg_valid[f.agg] = static_cast<uint8_t *>(ctx.Alloc(groups));
		vgpu::FinishNullableAggregate(static_cast<int>(a.kind), static_cast<int>(a.output_type), agg_out[out_req[f.agg]],
		                              agg_valid[f.agg], f.sum_req >= 0 ? agg_out[f.sum_req] : nullptr,
		                              f.sum_req >= 0 ? reqs[f.sum_req].out_type : 0,
		                              static_cast<const int64_t *>(agg_out[f.count_req]), groups, ctx.stream);
	}
	for (size_t i = 0; i < node.aggregates.size(); i++) {
		DeviceColumn c;
		c.name = "agg_" + std::to_string(i);
		c.type = node.aggregates[i].output_type;
		c.rows = groups;
		c.data = agg_out[out_req[i]];
		c.valid = agg_valid[i];
		r.push_back(std::move(c));
	}
	return r;
}

Columns ExecuteGroupBy(const GpuPlanNode &node, ExecContext &ctx) {
	if (node.children.size() != 1) {
		throw std::runtime_error("gpu_executor: GROUP_BY_AGGREGATE expects exactly one child");
	}
	if (JoinAggregateFusable(node)) {
		return ExecuteJoinAggregate(node, *node.children[0], ctx);
	}
	NameSet need(node.group_keys.begin(), node.group_keys.end());
	for (auto &a : node.aggregates) {
		if (a.kind != GpuAggregateKind::COUNT_STAR) {
			need.insert(a.input_column);
		}
	}
	Columns cols = ExecuteNode(*node.children[0], ctx, &need);
	if (cols.empty()) {
		throw std::runtime_error("gpu_groupby: input operator produced no columns to group");
	}
	CheckGroupByShape(node);
	const uint64_t rows = cols.front().rows;
	if (ctx.extended) {
		return GroupByExtended(node, cols, rows, ctx);
	}
	std::vector<vgpu::DevCol> keys;
	std::vector<const DeviceColumn *> key_cols;
	for (auto &key_name : node.group_keys) {
		const DeviceColumn &c = FindColumn(cols, key_name);
		CheckGroupKey(c, key_name, rows);
		vgpu::DevCol d;
		d.data = c.data;
		d.type = static_cast<int>(c.type);
		d.rows = c.rows;
		keys.push_back(d);
		key_cols.push_back(&c);
	}
	std::vector<vgpu::DevCol> values;
	auto reqs = BuildAggRequests(node, cols, rows, values);

	if (rows == 0) {
		if (node.group_keys.empty()) {
			ThrowGlobalEmpty();
		}
		Columns r;
		for (size_t k = 0; k < key_cols.size(); k++) {
			r.push_back(EmptyColumn(node.group_keys[k], key_cols[k]->type, ctx));
		}
		for (size_t i = 0; i < node.aggregates.size(); i++) {
			r.push_back(EmptyColumn("agg_" + std::to_string(i), node.aggregates[i].output_type, ctx));
		}
		return r;
	}

	vgpu::KeyPack pack;
	const uint32_t *hist8 = nullptr;
	const uint64_t bound = vgpu::PlanGroupBy(keys, rows, pack, ctx.arena, ctx.stream, &hist8);
	vgpu::SortTuning tuning = vgpu::TuneSort(pack.total_bits, rows);
	tuning.hist8 = hist8;
	std::vector<void *> key_out, agg_out;
	vgpu::OutAlloc alloc = [&](size_t bytes) { return ctx.Alloc(bytes); };
	const uint64_t groups = vgpu::GroupByDevice(keys, values, reqs, rows, pack, std::max<uint64_t>(bound, 1), alloc,
	                                            key_out, agg_out, tuning, ctx.arena, ctx.stream, true, nullptr,
	                                            nullptr, true);
	if (groups == 0) {
		throw std::runtime_error("gpu_groupby: computed zero groups for a non-empty input");
	}
	Columns r;
	for (size_t k = 0; k < key_cols.size(); k++) {
		DeviceColumn c;
		c.name = node.group_keys[k];
		c.type = key_cols[k]->type;
		c.rows = groups;
		c.data = key_out[k];
		r.push_back(std::move(c));
	}
	for (size_t i = 0; i < node.aggregates.size(); i++) {
		DeviceColumn c;
		c.name = "agg_" + std::to_string(i);
		c.type = node.aggregates[i].output_type;
		c.rows = groups;
		c.data = agg_out[i];
		r.push_back(std::move(c));
	}
	return r;
}

// ------------------------------------------------------------------ HASH_JOIN

void CheckJoinKeyType(const DeviceColumn &column, const char *side) {
	switch (column.type) {
	case GpuValueType::INT16:
	case GpuValueType::INT32:
	case GpuValueType::INT64:
	case GpuValueType::BOOLEAN:
		return;
	default:
		throw std::runtime_error(std::string("gpu_hash_join: ") + side +
		                         " join key is not an integer/boolean column, which this GPU path requires");
	}
}

struct JoinInputs {
	Columns left, right;
	const DeviceColumn *left_key = nullptr;
	const DeviceColumn *right_key = nullptr;
	std::vector<const DeviceColumn *> sources;
	std::vector<bool> from_left;
	//! Key columns built for the join (packed composite or ordered floating keys); heap-held so the pointers
	//! above stay valid when the struct moves.
	std::unique_ptr<DeviceColumn> packed_left, packed_right;
};

bool IsIntegerKeyType(GpuValueType t) {
	return t == GpuValueType::INT16 || t == GpuValueType::INT32 || t == GpuValueType::INT64 ||
	       t == GpuValueType::BOOLEAN;
}

//! Drops the rows of `side` whose value in any of `keys` is NULL (an INNER equi-join never matches them):
//! one generated predicate kernel over the nullable key columns emits the kept rows, every other column
//! is gathered through them with its validity. Pruned columns only take the new row count.
void DropNullKeyRows(Columns &side, const std::vector<std::string> &keys, ExecContext &ctx) {
	GpuExpr pred;
	pred.generated_cuda_source = "((bool*)out)[idx] = true;";
	pred.output_type = GpuValueType::BOOLEAN;
	for (auto &k : keys) {
		if (FindColumn(side, k).valid != nullptr) {
			pred.input_columns.push_back(k);
		}
	}
	if (pred.input_columns.empty() || side.empty()) {
		return;
	}
	const uint64_t rows = side.front().rows;
	std::vector<NamedExpr> outs;
	std::vector<size_t> where;
	for (size_t i = 0; i < side.size(); i++) {
		if (!side[i].pruned) {
			outs.push_back(NamedExpr {side[i].name, nullptr, &side[i]});
			where.push_back(i);
		}
	}
	uint64_t kept = 0;
	Columns compact = RunFused(side, rows, {pred}, outs, ctx, &kept);
	Columns out;
	size_t next = 0;
	for (size_t i = 0; i < side.size(); i++) {
		if (next < where.size() && where[next] == i) {
			DeviceColumn c = compact[next++];
			if (std::find(keys.begin(), keys.end(), c.name) != keys.end()) {
				c.valid = nullptr; // every kept key is non-NULL
			}
			out.push_back(std::move(c));
		} else {
			out.push_back(PrunedLike(side[i], kept));
		}
	}
	side = std::move(out);
}

//! Extended join keys (gpu_semantics.hpp): NULL-key rows dropped; FLOAT/DOUBLE keys replaced by their ordered
//! images; several key columns, or integer keys of different widths, packed into one 64-bit key per row with
//! the same packing on both sides (ranges over both sides; signed values compare after sign extension).
void PrepareJoinKeysExtended(const GpuPlanNode &node, ExecContext &ctx, JoinInputs &j) {
	const size_t k = node.build_keys.size();
	if (k == 0 || k != node.probe_keys.size()) {
		throw std::runtime_error("gpu_hash_join: build and probe key lists must have the same nonzero length");
	}
	DropNullKeyRows(j.left, node.build_keys, ctx);
	DropNullKeyRows(j.right, node.probe_keys, ctx);
	std::vector<DeviceColumn> lk, rk;
	bool same_single = k == 1;
	for (size_t i = 0; i < k; i++) {
		DeviceColumn l = FindColumn(j.left, node.build_keys[i]);
		DeviceColumn r = FindColumn(j.right, node.probe_keys[i]);
		const bool ints = IsIntegerKeyType(l.type) && IsIntegerKeyType(r.type);
		const bool floats = IsFloatType(l.type) && l.type == r.type;
		if (!ints && !floats) {
			throw std::runtime_error("gpu_hash_join: join keys '" + node.build_keys[i] + "' and '" + node.probe_keys[i] +
			                         "' are not both integer/boolean or both the same floating type");
		}
		same_single = same_single && l.type == r.type && ints;
		for (auto *c : {&l, &r}) {
			if (floats && c->rows > 0) {
				void *o = ctx.Alloc(c->rows * TypeSize(c->type));
				vgpu::FloatKeyToOrdered(c->data, static_cast<int>(c->type), c->rows, o, ctx.stream);
				c->data = o;
			}
			if (floats) {
				c->type = c->type == GpuValueType::FLOAT64 ? GpuValueType::INT64 : GpuValueType::INT32;
			}
		}
		lk.push_back(l);
		rk.push_back(r);
	}
	if (same_single) {
		j.left_key = &FindColumn(j.left, node.build_keys[0]);
		j.right_key = &FindColumn(j.right, node.probe_keys[0]);
		return;
	}
	if (k == 1 && lk[0].type == rk[0].type) {
		j.packed_left.reset(new DeviceColumn(lk[0]));
		j.packed_right.reset(new DeviceColumn(rk[0]));
	} else {
		// One packing for both sides: per key column, the range over both sides' values.
		uint64_t mins[vgpu::kMaxKeyCols], maxs[vgpu::kMaxKeyCols];
		if (k > (size_t)vgpu::kMaxKeyCols) {
			throw std::runtime_error("gpu_hash_join: at most 4 join key columns");
		}
		for (size_t i = 0; i < k; i++) {
			mins[i] = ~0ull;
			maxs[i] = 0;
			for (auto *c : {&lk[i], &rk[i]}) {
				if (c->rows == 0) {
					continue;
				}
				const void *cp[1] = {c->data};
				int ct[1] = {static_cast<int>(c->type)};
				uint64_t mn = 0, mx = 0;
				vgpu::KeyRanges(cp, ct, 1, c->rows, &mn, &mx, ctx.arena, ctx.stream);
				mins[i] = std::min(mins[i], mn);
				maxs[i] = std::max(maxs[i], mx);
			}
			if (mins[i] > maxs[i]) {
				mins[i] = maxs[i] = 0;
			}
		}
		auto pack_side = [&](std::vector<DeviceColumn> &cols, std::unique_ptr<DeviceColumn> &dst, const char *name) {
			std::vector<const void *> ptrs;
			std::vector<int> types;
			for (auto &c : cols) {
				ptrs.push_back(c.data);
				types.push_back(static_cast<int>(c.type));
			}
			vgpu::KeyPack pack;
			if (!vgpu::PlanKeyPack(ptrs.data(), types.data(), static_cast<int>(k), mins, maxs, pack)) {
				throw std::runtime_error("gpu_hash_join: the join key columns' value ranges need more than 64 bits "
				                         "once packed");
			}
			DeviceColumn d;
			d.name = name;
			d.type = GpuValueType::INT64;
			d.rows = cols[0].rows;
			d.data = ctx.Alloc(d.rows * 8);
			vgpu::PackKeys(pack, d.rows, static_cast<uint64_t *>(d.data), ctx.stream);
			dst.reset(new DeviceColumn(d));
		};
		pack_side(lk, j.packed_left, "$left_key");
		pack_side(rk, j.packed_right, "$right_key");
	}
	j.left_key = j.packed_left.get();
	j.right_key = j.packed_right.get();
}

//! Executes both join inputs and applies the operator's checks in the reference's order
//! (cuda_engine/src/operators/gpu_hash_join.cu:138-176).
JoinInputs PrepareJoin(const GpuPlanNode &node, ExecContext &ctx, const NameSet &reads) {
	if (node.children.size() != 2) {
		throw std::runtime_error("gpu_executor: HASH_JOIN expects exactly two children");
	}
	NameSet left_need = reads, right_need = reads;
	left_need.insert(node.build_keys.begin(), node.build_keys.end());
	right_need.insert(node.probe_keys.begin(), node.probe_keys.end());
	JoinInputs j;
	j.left = ExecuteNode(*node.children[0], ctx, &left_need);
	j.right = ExecuteNode(*node.children[1], ctx, &right_need);
	if (ctx.extended) {
		if (node.output_columns.empty()) {
			throw std::runtime_error("gpu_hash_join: node has no output columns");
		}
		PrepareJoinKeysExtended(node, ctx, j);
		for (auto &name : node.output_columns) {
			bool l = false;
			j.sources.push_back(&ResolveSideColumn(name, j.left, j.right, l));
			j.from_left.push_back(l);
		}
		return j;
	}
	if (node.build_keys.size() != 1 || node.probe_keys.size() != 1) {
		throw std::runtime_error("gpu_hash_join: this GPU path joins on exactly one equality key column");
	}
	if (node.output_columns.empty()) {
		throw std::runtime_error("gpu_hash_join: node has no output columns");
	}
	j.left_key = &FindColumn(j.left, node.build_keys[0]);
	j.right_key = &FindColumn(j.right, node.probe_keys[0]);
	CheckJoinKeyType(*j.left_key, "left");
	CheckJoinKeyType(*j.right_key, "right");
	if (j.left_key->type != j.right_key->type) {
		throw std::runtime_error("gpu_hash_join: join key columns must have the same type on both sides");
	}
	for (auto &name : node.output_columns) {
		bool l = false;
		j.sources.push_back(&ResolveSideColumn(name, j.left, j.right, l));
		j.from_left.push_back(l);
	}
	return j;
}

vgpu::DevCol ToDev(const DeviceColumn &c) {
	vgpu::DevCol d;
	d.data = c.data;
	d.type = static_cast<int>(c.type);
	d.rows = c.rows;
	return d;
}

//! The join's output columns from prepared inputs: row-id pairs, then one gather per needed column (and
//! per validity array).
Columns MaterializeJoin(const GpuPlanNode &node, const JoinInputs &j, ExecContext &ctx, const NameSet *need) {
	const uint64_t left_rows = j.left_key->rows;
	const uint64_t right_rows = j.right_key->rows;
	auto shaped = [&](uint64_t rows) {
		Columns r;
		for (size_t i = 0; i < j.sources.size(); i++) {
			if (rows == 0 && !j.sources[i]->pruned && Needed(need, node.output_columns[i])) {
				r.push_back(EmptyColumn(node.output_columns[i], j.sources[i]->type, ctx));
			} else {
				DeviceColumn p = PrunedLike(*j.sources[i], rows);
				p.name = node.output_columns[i];
				r.push_back(p);
			}
		}
		return r;
	};
	if (left_rows == 0 || right_rows == 0) {
		return shaped(0);
	}
	uint32_t *lidx = nullptr, *ridx = nullptr;
	const uint64_t total =
	    vgpu::JoinMaterializeDevice(ToDev(*j.left_key), ToDev(*j.right_key), &lidx, &ridx, ctx.arena, ctx.stream);
	if (total == 0) {
		return shaped(0);
	}
	Columns r = shaped(total);
	for (size_t i = 0; i < j.sources.size(); i++) {
		const DeviceColumn &src = *j.sources[i];
		if (src.pruned || !Needed(need, node.output_columns[i])) {
			continue;
		}
		DeviceColumn o;
		o.name = node.output_columns[i];
		o.type = src.type;
		o.rows = total;
		o.pruned = false;
		o.data = ctx.Alloc(total * TypeSize(src.type));
		vgpu::Gather(src.data, static_cast<int>(TypeSize(src.type)), j.from_left[i] ? lidx : ridx, false, total, o.data,
		             ctx.stream);
		if (src.valid != nullptr) {
			o.valid = static_cast<uint8_t *>(ctx.Alloc(total));
			vgpu::Gather(src.valid, 1, j.from_left[i] ? lidx : ridx, false, total, o.valid, ctx.stream);
		}
		r[i] = std::move(o);
	}
	return r;
}

Columns ExecuteJoin(const GpuPlanNode &node, ExecContext &ctx, const NameSet *need) {
	NameSet reads;
	for (auto &name : node.output_columns) {
		if (Needed(need, name)) {
			reads.insert(name);
		}
	}
	JoinInputs j = PrepareJoin(node, ctx, reads);
	return MaterializeJoin(node, j, ctx, need);
}

//! GROUP BY without keys directly over an INNER join: per-side sort + run reduction, merge of the
//! distinct keys, exact combination of the per-key states. No pair is materialized. Semantics are those of
//! materializing the join and aggregating it (including the refusal of an empty result).
Columns ExecuteJoinAggregate(const GpuPlanNode &node, const GpuPlanNode &join, ExecContext &ctx) {
	NameSet reads;
	for (auto &a : node.aggregates) {
		if (a.kind != GpuAggregateKind::COUNT_STAR) {
			reads.insert(a.input_column);
		}
	}
	JoinInputs j = PrepareJoin(join, ctx, reads);
	CheckGroupByShape(node);
	if (ctx.extended) {
		// NULLs in an aggregate input, or an empty join (one row of NULLs), take the materializing path.
		bool materialize = j.left_key->rows == 0 || j.right_key->rows == 0;
		for (auto *src : j.sources) {
			materialize = materialize || src->valid != nullptr;
		}
		if (materialize) {
			Columns joined = MaterializeJoin(join, j, ctx, &reads);
			const uint64_t rows = joined.empty() ? 0 : joined.front().rows;
			return GroupByExtended(node, joined, rows, ctx);
		}
	}
	// The join's output columns, as the aggregate sees them.
	Columns join_out;
	for (size_t i = 0; i < j.sources.size(); i++) {
		DeviceColumn c = *j.sources[i];
		c.name = join.output_columns[i];
		join_out.push_back(c);
	}
	std::vector<vgpu::DevCol> side_vals[2];
	std::vector<vgpu::JoinAggReq> reqs;
	std::vector<vgpu::DevCol> check_values;
	for (auto &a : node.aggregates) {
		vgpu::JoinAggReq r;
		r.kind = static_cast<int>(a.kind);
		r.out_type = static_cast<int>(a.output_type);
		vgpu::AggReq check;
		check.kind = r.kind;
		check.out_type = r.out_type;
		if (a.kind != GpuAggregateKind::COUNT_STAR) {
			size_t idx = 0;
			for (; idx < join_out.size() && join_out[idx].name != a.input_column; idx++) {
			}
			if (idx == join_out.size()) {
				throw std::runtime_error("gpu_executor: expression references column '" + a.input_column +
				                         "' which is not produced by its input operator");
			}
			const DeviceColumn &src = *j.sources[idx];
			const int side = j.from_left[idx] ? 0 : 1;
			r.side = side;
			r.input = static_cast<int>(side_vals[side].size());
			side_vals[side].push_back(ToDev(src));
			check.input = static_cast<int>(check_values.size());
			check_values.push_back(ToDev(src));
			if (!ctx.extended) {
				ValidateEngineAggregate(a, src.type);
			}
		}
		vgpu::ValidateAggregateReq(check, check_values);
		reqs.push_back(r);
	}
	if (j.left_key->rows == 0 || j.right_key->rows == 0) {
		ThrowGlobalEmpty();
	}
	const size_t count_slot = reqs.size();
	reqs.push_back(vgpu::JoinAggReq {vgpu::AGG_COUNT_STAR, -1, -1, vgpu::VT_INT64});
	auto vals = vgpu::JoinAggregateDevice(ToDev(*j.left_key), side_vals[0], ToDev(*j.right_key), side_vals[1], reqs,
	                                      ctx.arena, ctx.stream);
	const uint64_t pairs = vals[count_slot].lo;
	if (pairs == 0 && vals[count_slot].hi == 0) {
		if (ctx.extended) {
			return GroupByExtended(node, MaterializeJoin(join, j, ctx, &reads), 0, ctx);
		}
		ThrowGlobalEmpty();
	}
	std::vector<uint8_t> host(16 * node.aggregates.size(), 0);
	Columns r;
	for (size_t i = 0; i < node.aggregates.size(); i++) {
		vgpu::FinishJoinAggregate(vals[i], reqs[i].kind, reqs[i].out_type, pairs, host.data() + 16 * i);
		DeviceColumn c;
		c.name = "agg_" + std::to_string(i);
		c.type = node.aggregates[i].output_type;
		c.rows = 1;
		c.data = ctx.Alloc(TypeSize(c.type));
		ctx.staging.Upload(c.data, host.data() + 16 * i, TypeSize(c.type));
		r.push_back(std::move(c));
	}
	return r;
}

Columns ExecuteNode(const GpuPlanNode &node, ExecContext &ctx, const NameSet *need) {
	switch (node.op_type) {
	case GpuOpType::SCAN:
		return ExecuteScan(node, ctx, need);
	case GpuOpType::FILTER:
		return ExecuteFilter(node, ctx, need, nullptr);
	case GpuOpType::PROJECTION:
		return ExecuteProjection(node, ctx);
	case GpuOpType::GROUP_BY_AGGREGATE:
		return ExecuteGroupBy(node, ctx);
	case GpuOpType::HASH_JOIN:
		return ExecuteJoin(node, ctx, need);
	case GpuOpType::CROSS_PRODUCT:
		return ExecuteCross(node, ctx, need, nullptr, nullptr);
	default:
		throw std::runtime_error("gpu_executor: unknown GpuOpType value " +
		                         std::to_string(static_cast<int>(node.op_type)));
	}
}

hipStream_t SharedStream() {
	static hipStream_t stream = [] {
		hipStream_t s = nullptr;
		hipError_t status = hipStreamCreate(&s);
		if (status != hipSuccess) {
			(void)hipGetLastError();
			throw std::runtime_error(std::string("gpu_executor: hipStreamCreate failed: ") + hipGetErrorString(status));
		}
		return s;
	}();
	return stream;
}

//! Waits for the stream before anything the plan staged on it (page-locked slots) is released.
struct StreamDrain {
	hipStream_t stream;
	~StreamDrain() {
		(void)hipStreamSynchronize(stream);
	}
};

std::mutex &PlanMutex() {
	static std::mutex m;
	return m;
}

} // namespace

bool ExecuteGpuPlan(const GpuPlanNode &plan, const std::vector<GpuColumn> &inputs, const GpuExecuteOptions &options,
                    GpuExecutionResult &out) {
	try {
		std::lock_guard<std::mutex> plan_lock(PlanMutex());
		hipStream_t stream = SharedStream();
		vgpu::PoolArena arena(stream);
		StreamDrain drain {stream};
		Staging staging(stream);
		ExecContext ctx {inputs, 0, arena, stream, staging};
		const bool serving_from_cache = options.cache_chunk_index >= 0 || options.serve_whole_scan_from_cache;
		if (serving_from_cache && options.stage_scan_to_cache) {
			throw std::runtime_error("gpu_executor: cache replay and cache staging cannot be requested together");
		}
		if (options.cache_chunk_index >= 0 && options.serve_whole_scan_from_cache) {
			throw std::runtime_error("gpu_executor: cannot serve one cached chunk and the whole scan at once");
		}
		if (serving_from_cache && !inputs.empty()) {
			throw std::runtime_error("gpu_executor: cache replay expects no host inputs");
		}
		ctx.cache_chunk_index = options.cache_chunk_index;
		ctx.serve_whole_scan_from_cache = options.serve_whole_scan_from_cache;
		ctx.stage_scan_to_cache = options.stage_scan_to_cache;
		ctx.extended = IsGpuExtendedSemanticsEnabled();
		ctx.nulls_supported = IsRowIndependent(plan) || ctx.extended;
		ctx.fast_math = IsGpuFastMathEnabled();

		Columns columns = ExecuteNode(plan, ctx, nullptr);

		// Egress: every owned buffer is created first (pointers into them stay valid), then data and
		// device-packed validity stream through the staging ring straight into them.
		for (auto &c : columns) {
			if (c.pruned) {
				throw std::logic_error("gpu_executor: result column '" + c.name + "' was pruned");
			}
		}
		std::vector<size_t> data_index(columns.size()), valid_index(columns.size(), SIZE_MAX);
		std::vector<size_t> buffer_bytes;
		for (size_t i = 0; i < columns.size(); i++) {
			buffer_bytes.push_back(static_cast<size_t>(columns[i].rows) * TypeSize(columns[i].type));
			data_index[i] = buffer_bytes.size() - 1;
			if (columns[i].valid != nullptr && columns[i].rows > 0) {
				buffer_bytes.push_back((static_cast<size_t>(columns[i].rows) + 7) / 8);
				valid_index[i] = buffer_bytes.size() - 1;
			}
		}
		// The result's buffers are std::vector<uint8_t> (the reference's type), which zero-fills -- and so
		// page-faults -- every byte on construction; large ones are built on threads of their own so those
		// fills run side by side instead of one after another.
		out.owned_buffers.resize(buffer_bytes.size());
		{
			std::vector<std::thread> makers;
			std::vector<std::exception_ptr> failed(buffer_bytes.size());
			for (size_t b = 0; b < buffer_bytes.size(); b++) {
				if (buffer_bytes[b] >= (16u << 20)) {
					makers.emplace_back([&out, &buffer_bytes, &failed, b] {
						try {
							out.owned_buffers[b].resize(buffer_bytes[b]);
						} catch (...) {
							failed[b] = std::current_exception();
						}
					});
				} else {
					out.owned_buffers[b].resize(buffer_bytes[b]);
				}
			}
			for (auto &t : makers) {
				t.join();
			}
			for (auto &e : failed) {
				if (e) {
					std::rethrow_exception(e);
				}
			}
		}
		for (size_t i = 0; i < columns.size(); i++) {
			const DeviceColumn &c = columns[i];
			auto &data = out.owned_buffers[data_index[i]];
			auto start = std::chrono::steady_clock::now();
			if (!data.empty()) {
				staging.Download(data.data(), c.data, data.size());
			}
			if (valid_index[i] != SIZE_MAX) {
				auto &packed_host = out.owned_buffers[valid_index[i]];
				auto *packed = static_cast<uint8_t *>(ctx.Alloc(packed_host.size()));
				vgpu::PackValidity(c.valid, packed, c.rows, stream);
				staging.Download(packed_host.data(), packed, packed_host.size());
			}
			VGPU_LOG(LogLevel::INFO, "d2h",
			         "\"column\":" + GpuLogger::Quote(c.name) + ",\"bytes\":" + std::to_string(data.size()) +
			             ",\"us\":" +
			             std::to_string(std::chrono::duration_cast<std::chrono::microseconds>(
			                                std::chrono::steady_clock::now() - start)
			                                .count()));
		}
		staging.Drain();
		VGPU_HIP(hipStreamSynchronize(stream));
		for (size_t i = 0; i < columns.size(); i++) {
			auto &data = out.owned_buffers[data_index[i]];
			GpuColumn column;
			column.name = columns[i].name;
			column.type = columns[i].type;
			column.row_count = columns[i].rows;
			column.data = data.empty() ? nullptr : data.data();
			if (valid_index[i] != SIZE_MAX) {
				column.validity = out.owned_buffers[valid_index[i]].data();
			}
			out.columns.push_back(std::move(column));
		}
		out.success = true;
		return true;
	} catch (const std::exception &e) {
		out.success = false;
		out.columns.clear();
		out.owned_buffers.clear();
		out.error_message = e.what();
		return false;
	}
}

namespace {
int64_t FreeDeviceMiB() {
	size_t free_bytes = 0, total_bytes = 0;
	if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) {
		(void)hipGetLastError();
		return -1;
	}
	return static_cast<int64_t>(free_bytes >> 20);
}
constexpr int64_t kLowFreeMiB = 128;
} // namespace

GpuEngine &GpuEngine::Instance() {
	static GpuEngine instance;
	return instance;
}

GpuExecutionResult GpuEngine::ExecutePlan(const GpuPlanNode &plan, const std::vector<GpuColumn> &inputs,
                                          const GpuExecuteOptions &options) {
	GpuExecutionResult result;
	try {
		RmmPool::Instance().EnsureInitialized();
	} catch (const std::exception &e) {
		result.success = false;
		result.error_message = std::string("GpuEngine::ExecutePlan: GPU memory pool initialization failed "
		                                   "(is a GPU device available?): ") +
		                       e.what();
		return result;
	}
	const bool watermark = GpuLogger::Instance().Enabled(LogLevel::INFO);
	const int64_t before = watermark ? FreeDeviceMiB() : -1;
	ExecuteGpuPlan(plan, inputs, options, result);
	if (watermark) {
		const int64_t after = FreeDeviceMiB();
		if (before >= 0 && after >= 0) {
			const bool low = before < kLowFreeMiB || after < kLowFreeMiB;
			VGPU_LOG(LogLevel::INFO, "vram_watermark",
			         "\"free_before_mib\":" + std::to_string(before) + ",\"free_after_mib\":" + std::to_string(after) +
			             ",\"paging_risk\":" + (low ? "true" : "false") + ",\"threshold_mib\":" +
			             std::to_string(kLowFreeMiB));
		}
	}
	return result;
}

} // namespace vector_gpu

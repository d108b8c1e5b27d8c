// server.cpp -- the served form (request loop, memory-footprint admission, continuous batching).
#include "server.h"

#include "device_pool.h"
#include "../hip/join_kernels.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>

namespace vgpu {

namespace {

const char *CarryType(int vtype) {
	switch (vtype) {
	case VT_INT16:
		return "short";
	case VT_INT32:
		return "int";
	case VT_INT64:
		return "long long";
	case VT_FLOAT32:
		return "float";
	case VT_FLOAT64:
		return "double";
	case VT_BOOLEAN:
		return "bool";
	default:
		throw std::runtime_error("served group by: column type cannot be carried by a generated kernel");
	}
}

uint64_t SumWidths(const std::vector<DevCol> &cols, const std::vector<int> &idx) {
	uint64_t w = 0;
	for (int i : idx) {
		w += TypeWidth(cols[i].type);
	}
	return w;
}

//! Moves `bytes` of device memory into a block owned by `hold` (exact-size result buffers).
void *KeepCopy(const void *src, size_t bytes, BatchBuffers &hold, hipStream_t stream) {
	void *dst = DevicePool::Instance().Alloc(std::max<size_t>(bytes, 8), stream);
	hold.blocks.push_back(dst);
	if (bytes > 0) {
		CopyDevice(dst, src, bytes, stream);
	}
	return dst;
}

} // namespace

BatchBuffers::~BatchBuffers() {
	auto &pool = DevicePool::Instance();
	for (void *p : blocks) {
		pool.FreeAny(p);
	}
}

uint64_t Server::NowNs() {
	return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
	           std::chrono::steady_clock::now().time_since_epoch())
	    .count();
}

Server::Server(const ServerOptions &options) : options_(options) {
	if (options_.slots < 1 || options_.slots > 16) {
		throw std::runtime_error("server: 1 .. 16 executor slots");
	}
	if (options_.max_batch < 1) {
		throw std::runtime_error("server: max_batch must be at least 1");
	}
	if (!options_.batching) {
		options_.max_batch = 1;
	} else {
		options_.resident_join = true;
	}
	VGPU_HIP(hipGetDevice(&device_));
	if (options_.memory_budget == 0) {
		size_t free_bytes = 0, total = 0;
		VGPU_HIP(hipMemGetInfo(&free_bytes, &total));
		options_.memory_budget = free_bytes / 4 * 3;
	}
	stats_.budget = options_.memory_budget;
	for (int i = 0; i < options_.slots; i++) {
		hipStream_t s = nullptr;
		VGPU_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		streams_.push_back(s);
	}
	for (int i = 0; i < options_.slots; i++) {
		executors_.emplace_back(&Server::Executor, this, i);
	}
}

Server::~Server() {
	{
		std::lock_guard<std::mutex> lock(mutex_);
		stop_ = true;
	}
	work_cv_.notify_all();
	for (auto &t : executors_) {
		t.join();
	}
	done_.clear();
	templates_.clear();
	for (auto s : streams_) {
		DevicePool::Instance().ForgetStream(s);
		(void)hipStreamDestroy(s);
	}
}

int Server::AddTemplate(std::unique_ptr<ServedTemplate> t) {
	if (t->cols.empty()) {
		throw std::runtime_error("server template: no table columns");
	}
	t->rows = t->cols[0].rows;
	for (auto &c : t->cols) {
		if (c.rows != t->rows) {
			throw std::runtime_error("server template: table columns differ in length");
		}
	}
	t->fast_math = t->fast_math || options_.fast_math;
	switch (t->kind) {
	case TMPL_FILTER_PROJECT:
		if (t->projs.empty()) {
			throw std::runtime_error("server template: a filter/projection needs at least one projection");
		}
		break;
	case TMPL_GROUP_BY:
		if (t->key_col < 0 || t->key_col >= (int)t->cols.size() || !TypeIsInteger(t->cols[t->key_col].type)) {
			throw std::runtime_error("server template: GROUP BY key must be an integer column of the table");
		}
		for (int v : t->value_cols) {
			if (v < 0 || v >= (int)t->cols.size()) {
				throw std::runtime_error("server template: value column out of range");
			}
			CarryType(t->cols[v].type);
		}
		for (auto &a : t->aggs) {
			if (a.kind != AGG_COUNT_STAR && (a.input < 0 || a.input >= (int)t->value_cols.size())) {
				throw std::runtime_error("server template: aggregate input out of range");
			}
		}
		break;
	case TMPL_CROSS: {
		if (t->projs.empty() || t->left_cols.empty()) {
			throw std::runtime_error("server template: a cross product needs left columns and a projection");
		}
		for (auto &c : t->left_cols) {
			if (c.rows != t->left_cols[0].rows) {
				throw std::runtime_error("server template: left columns differ in length");
			}
		}
		const int slots = (int)(t->left_cols.size() + t->cols.size());
		if (slots > kFusedMaxCols || (int)t->projs.size() > kFusedMaxOuts) {
			throw std::runtime_error("server template: at most 16 cross-product columns and outputs");
		}
		for (auto *list : {&t->preds, &t->projs}) {
			for (auto &e : *list) {
				for (int sl : e.slots) {
					if (sl < 0 || sl >= slots) {
						throw std::runtime_error("server template: expression input out of range");
					}
				}
			}
		}
		break;
	}
	case TMPL_JOIN_AGG:
		if (t->cols[0].type != t->build_key.type) {
			throw std::runtime_error("server template: probe and build keys must have the same type");
		}
		if (options_.resident_join) {
			// The resident intermediate: the build side reduced once, probed by every batch. Built on a stream
			// of its own so it never interleaves with an executor's batch.
			hipStream_t prep = nullptr;
			VGPU_HIP(hipStreamCreateWithFlags(&prep, hipStreamNonBlocking));
			try {
				t->index = BuildJoinIndex(t->build_key, t->build_payload, prep);
			} catch (...) {
				DevicePool::Instance().ForgetStream(prep);
				(void)hipStreamDestroy(prep);
				throw;
			}
			DevicePool::Instance().ForgetStream(prep);
			(void)hipStreamDestroy(prep);
		}
		break;
	default:
		throw std::runtime_error("server template: unknown kind");
	}
	std::lock_guard<std::mutex> lock(mutex_);
	templates_.push_back(std::move(t));
	return (int)templates_.size() - 1;
}

uint64_t Server::Footprint(const ServedTemplate &t, uint64_t rows, uint32_t nqueries) const {
	const uint64_t slack = 16ull << 20; // offsets, look-back state, key ranges, small buffers
	switch (t.kind) {
	case TMPL_FILTER_PROJECT: {
		uint64_t w = 4; // segment ids
		for (auto &p : t.projs) {
			w += TypeWidth(p.out_type);
		}
		return rows * w + slack;
	}
	case TMPL_GROUP_BY: {
		const uint64_t kw = TypeWidth(t.cols[t.key_col].type);
		const uint64_t vw = SumWidths(t.cols, t.value_cols);
		uint64_t aw = 0;
		for (auto &a : t.aggs) {
			aw += TypeWidth(a.out_type);
		}
		// compacted rows, sort buffers (packed keys twice, permuted payload twice), group outputs
		return rows * (3 * (4 + kw + vw) + 32 + (4 + kw + aw)) + slack;
	}
	case TMPL_CROSS: {
		// outputs sized for every pair of the query, plus the look-back state
		const uint64_t pairs = rows * std::max<uint64_t>(t.left_cols[0].rows, 1);
		uint64_t w = 0;
		for (auto &p : t.projs) {
			w += TypeWidth(p.out_type);
		}
		return pairs * w + pairs / 512 + slack;
	}
	default: // JOIN_AGG
		if (t.index != nullptr) {
			return 24ull * nqueries + slack;
		}
		// plain path: both sides sorted and reduced per query
		return (rows + t.build_key.rows) * 64 + slack;
	}
}

uint64_t Server::Submit(int template_id, uint64_t row_start, uint64_t row_count, uint64_t submit_ns) {
	std::unique_lock<std::mutex> lock(mutex_);
	if (template_id < 0 || template_id >= (int)templates_.size()) {
		throw std::runtime_error("server: unknown template " + std::to_string(template_id));
	}
	const auto &t = *templates_[template_id];
	if (row_start > t.rows || row_count > t.rows - row_start) {
		throw std::runtime_error("server: row range past the table");
	}
	const uint64_t ticket = next_ticket_++;
	open_[ticket] = true;
	Pending q {ticket, template_id, row_start, row_count, submit_ns == 0 ? NowNs() : submit_ns};
	if (Footprint(t, row_count + 3, 1) > options_.memory_budget) {
		auto r = std::unique_ptr<ServedResult>(new ServedResult());
		r->status = 1;
		r->error = "server: the query's footprint exceeds the whole memory budget";
		r->kind = t.kind;
		r->submit_ns = r->start_ns = r->done_ns = q.submit_ns;
		done_[ticket] = std::move(r);
		stats_.rejected++;
		lock.unlock();
		done_cv_.notify_all();
		return ticket;
	}
	queue_.push_back(q);
	lock.unlock();
	work_cv_.notify_one();
	return ticket;
}

std::unique_ptr<ServedResult> Server::Wait(uint64_t ticket) {
	std::unique_lock<std::mutex> lock(mutex_);
	if (open_.find(ticket) == open_.end()) {
		throw std::runtime_error("server: unknown or already collected ticket");
	}
	done_cv_.wait(lock, [&] { return done_.find(ticket) != done_.end(); });
	auto r = std::move(done_[ticket]);
	done_.erase(ticket);
	open_.erase(ticket);
	return r;
}

ServerStats Server::Stats() const {
	std::lock_guard<std::mutex> lock(mutex_);
	return stats_;
}

void Server::Executor(int slot) {
	(void)hipSetDevice(device_);
	hipStream_t stream = streams_[slot];
	for (;;) {
		std::vector<Pending> batch;
		uint64_t reserved = 0;
		const ServedTemplate *tp = nullptr;
		{
			std::unique_lock<std::mutex> lock(mutex_);
			auto head_fits = [&] {
				if (queue_.empty()) {
					return false;
				}
				const auto &h = queue_.front();
				return Footprint(*templates_[h.tmpl], h.count + 3, 1) <= options_.memory_budget - reserved_;
			};
			work_cv_.wait(lock, [&] { return head_fits() || (stop_ && queue_.empty()); });
			if (queue_.empty()) {
				return; // stopping and drained
			}
			// The oldest query, then every later query of its template that still fits.
			const int tmpl = queue_.front().tmpl;
			const auto &t = *templates_[tmpl];
			tp = &t; // templates are never removed while the server lives; the vector may grow meanwhile
			const uint64_t avail = options_.memory_budget - reserved_;
			// a GROUP BY batch sorts fewer than 2^30 rows
			uint64_t max_rows = t.kind == TMPL_GROUP_BY ? (1ull << 30) - (1ull << 20) : ~0ull;
			if (options_.max_batch_rows != 0) {
				max_rows = std::min(max_rows, options_.max_batch_rows);
			}
			uint64_t rows = 0;
			const int max_batch = t.kind == TMPL_CROSS ? 1 : options_.max_batch;
			for (auto it = queue_.begin(); it != queue_.end() && (int)batch.size() < max_batch;) {
				if (it->tmpl != tmpl) {
					++it;
					continue;
				}
				const uint64_t r = rows + it->count + 3;
				const uint64_t fp = Footprint(t, r, (uint32_t)batch.size() + 1);
				if (!batch.empty() && (fp > avail || r > max_rows)) {
					break;
				}
				batch.push_back(*it);
				rows = r;
				reserved = fp;
				it = queue_.erase(it);
			}
			reserved_ += reserved;
			stats_.peak_reserved = std::max(stats_.peak_reserved, reserved_);
			stats_.batches++;
			stats_.queries += batch.size();
			stats_.max_batch = std::max<uint64_t>(stats_.max_batch, batch.size());
		}
		const uint64_t start = NowNs();
		std::vector<std::unique_ptr<ServedResult>> results(batch.size());
		for (auto &r : results) {
			r.reset(new ServedResult());
		}
		try {
			if (tp->kind == TMPL_CROSS) {
				RunCross(*tp, batch[0], stream, *results[0]);
			} else if (batch.size() > 1 || tp->index != nullptr) {
				// a resident join index is probed through the batch path, also for a batch of one; any other
				// query that found no company runs alone on the plain operators (no segment column, no
				// composite key: the shortest path for one query)
				RunBatch(*tp, batch, stream, results);
			} else {
				RunAlone(*tp, batch[0], stream, *results[0]);
			}
		} catch (const std::exception &e) {
			for (auto &r : results) {
				r->status = 1;
				r->error = e.what();
			}
			(void)hipStreamSynchronize(stream);
		}
		const uint64_t done = NowNs();
		{
			std::lock_guard<std::mutex> lock(mutex_);
			reserved_ -= reserved;
			for (size_t i = 0; i < batch.size(); i++) {
				auto &r = results[i];
				r->kind = tp->kind;
				r->batch_size = (uint32_t)batch.size();
				r->submit_ns = batch[i].submit_ns;
				r->start_ns = start;
				r->done_ns = done;
				done_[batch[i].ticket] = std::move(r);
			}
		}
		done_cv_.notify_all();
		work_cv_.notify_all();
	}
}

void Server::RunBatch(const ServedTemplate &t, const std::vector<Pending> &batch, hipStream_t stream,
                      std::vector<std::unique_ptr<ServedResult>> &results) {
	const uint32_t nq = (uint32_t)batch.size();
	std::vector<uint64_t> starts(nq), lens(nq);
	for (uint32_t g = 0; g < nq; g++) {
		starts[g] = batch[g].start;
		lens[g] = batch[g].count;
	}
	PoolArena arena(stream);
	auto hold = std::make_shared<BatchBuffers>();
	if (t.kind == TMPL_FILTER_PROJECT) {
		const DeviceRanges ranges = UploadRanges(starts.data(), lens.data(), nq, t.rows, arena, stream);
		PoolArena out_arena(stream);
		std::vector<void *> outs;
		std::vector<uint64_t> offs;
		BatchedFilterProjectDevice(t.cols, ranges, t.preds, t.projs, t.fast_math, outs, offs, out_arena, arena,
		                           stream);
		for (void *p : outs) {
			out_arena.Release(p);
			hold->blocks.push_back(p);
		}
		for (uint32_t g = 0; g < nq; g++) {
			auto &r = *results[g];
			r.rows = offs[g + 1] - offs[g];
			for (size_t i = 0; i < t.projs.size(); i++) {
				r.data.push_back(static_cast<const uint8_t *>(outs[i]) + offs[g] * TypeWidth(t.projs[i].out_type));
				r.types.push_back(t.projs[i].out_type);
			}
			r.hold = hold;
		}
		return;
	}
	if (t.kind == TMPL_GROUP_BY) {
		BatchedGroupByArgs a;
		a.cols = t.cols;
		a.rows = t.rows;
		a.nseg = nq;
		a.range_rows_host = starts.data();
		a.range_lens_host = lens.data();
		a.preds = t.preds;
		a.key_col = t.key_col;
		a.value_cols = t.value_cols;
		a.aggs = t.aggs;
		a.fast_math = t.fast_math;
		uint64_t cap = 0;
		for (uint64_t l : lens) {
			cap += l + 3;
		}
		cap = std::max<uint64_t>(cap, 1);
		const int kt = t.cols[t.key_col].type;
		a.seg_out = arena.Alloc(cap * 4);
		a.key_out = arena.Alloc(cap * TypeWidth(kt));
		for (auto &ag : t.aggs) {
			a.agg_out.push_back(arena.Alloc(cap * TypeWidth(ag.out_type)));
		}
		a.capacity = cap;
		std::vector<uint64_t> goffs(nq + 1);
		a.seg_group_offsets_host = goffs.data();
		const uint64_t groups = BatchedFilterGroupByDevice(a, arena, stream);
		// Exact-size result buffers; the batch's scratch goes back to the pool.
		const void *keys = KeepCopy(a.key_out, groups * TypeWidth(kt), *hold, stream);
		std::vector<const void *> aggs;
		for (size_t i = 0; i < t.aggs.size(); i++) {
			aggs.push_back(KeepCopy(a.agg_out[i], groups * TypeWidth(t.aggs[i].out_type), *hold, stream));
		}
		VGPU_HIP(hipStreamSynchronize(stream));
		for (uint32_t g = 0; g < nq; g++) {
			auto &r = *results[g];
			r.rows = goffs[g + 1] - goffs[g];
			r.data.push_back(static_cast<const uint8_t *>(keys) + goffs[g] * TypeWidth(kt));
			r.types.push_back(kt);
			for (size_t i = 0; i < t.aggs.size(); i++) {
				r.data.push_back(static_cast<const uint8_t *>(aggs[i]) + goffs[g] * TypeWidth(t.aggs[i].out_type));
				r.types.push_back(t.aggs[i].out_type);
			}
			r.hold = hold;
		}
		return;
	}
	// JOIN_AGG: one probe pass of every range against the resident index.
	const DeviceRanges ranges = UploadRanges(starts.data(), lens.data(), nq, t.rows, arena, stream);
	auto values = BatchedJoinAggDevice(t.cols[0], ranges, *t.index, arena, stream);
	for (uint32_t g = 0; g < nq; g++) {
		auto &r = *results[g];
		const uint64_t pairs = values[2 * g + 1].lo;
		FinishJoinAggregate(values[2 * g], AGG_SUM, VT_HUGEINT, pairs, r.scalars[0]);
		FinishJoinAggregate(values[2 * g + 1], AGG_COUNT_STAR, VT_INT64, pairs, r.scalars[1]);
		r.rows = 1;
		r.on_device = false;
		r.data = {r.scalars[0], r.scalars[1]};
		r.types = {VT_HUGEINT, VT_INT64};
		r.null = {(uint8_t)(pairs == 0), 0};
	}
}

void Server::RunAlone(const ServedTemplate &t, const Pending &q, hipStream_t stream, ServedResult &r) {
	PoolArena arena(stream);
	auto hold = std::make_shared<BatchBuffers>();
	auto &pool = DevicePool::Instance();
	auto keep_alloc = [&](size_t bytes) {
		void *p = pool.Alloc(std::max<size_t>(bytes, 8), stream);
		hold->blocks.push_back(p);
		return p;
	};
	// The table's columns viewed from the query's first row.
	std::vector<DevCol> view = t.cols;
	for (auto &c : view) {
		c.data = static_cast<const uint8_t *>(c.data) + q.start * TypeWidth(c.type);
		c.rows = q.count;
	}
	if (t.kind == TMPL_FILTER_PROJECT) {
		// The plain fused kernel over this query's rows.
		FusedSpec spec;
		spec.fast_math = t.fast_math;
		std::vector<const void *> ptrs;
		for (auto &c : view) {
			FusedColumn fc;
			fc.vtype = c.type;
			spec.cols.push_back(fc);
			ptrs.push_back(c.data);
		}
		spec.predicates = t.preds;
		spec.projections = t.projs;
		auto kernel = GetFusedKernel(spec);
		std::vector<void *> outs;
		for (auto &p : t.projs) {
			outs.push_back(keep_alloc(q.count * TypeWidth(p.out_type)));
		}
		uint64_t kept = q.count;
		auto *count = static_cast<uint64_t *>(arena.Alloc(8));
		const size_t state_bytes = FusedKernel::StateBytes(q.count);
		void *state = arena.Alloc(std::max<size_t>(state_bytes, 8));
		FusedLaunch la;
		la.cols = ptrs.data();
		la.outs = outs.data();
		la.n = q.count;
		la.count_device = t.preds.empty() ? nullptr : count;
		kernel->Launch(la, state, state_bytes, stream);
		if (!t.preds.empty()) {
			VGPU_HIP(hipMemcpyAsync(&kept, count, 8, hipMemcpyDeviceToHost, stream));
		}
		VGPU_HIP(hipStreamSynchronize(stream));
		r.rows = q.count == 0 ? 0 : kept;
		for (size_t i = 0; i < outs.size(); i++) {
			r.data.push_back(outs[i]);
			r.types.push_back(t.projs[i].out_type);
		}
		r.hold = hold;
		return;
	}
	if (t.kind == TMPL_GROUP_BY) {
		BatchedGroupByArgs a;
		a.cols = t.cols;
		a.rows = t.rows;
		a.nseg = 1;
		a.preds = t.preds;
		a.key_col = t.key_col;
		a.value_cols = t.value_cols;
		a.aggs = t.aggs;
		a.fast_math = t.fast_math;
		const uint64_t cap = std::max<uint64_t>(q.count, 1);
		const int kt = t.cols[t.key_col].type;
		a.seg_out = arena.Alloc(cap * 4);
		a.key_out = arena.Alloc(cap * TypeWidth(kt));
		for (auto &ag : t.aggs) {
			a.agg_out.push_back(arena.Alloc(cap * TypeWidth(ag.out_type)));
		}
		a.capacity = cap;
		const uint64_t offsets[2] = {q.start, q.start + q.count};
		const uint64_t groups = PerQueryFilterGroupByDevice(a, offsets, arena, stream);
		r.data.push_back(KeepCopy(a.key_out, groups * TypeWidth(kt), *hold, stream));
		r.types.push_back(kt);
		for (size_t i = 0; i < t.aggs.size(); i++) {
			r.data.push_back(KeepCopy(a.agg_out[i], groups * TypeWidth(t.aggs[i].out_type), *hold, stream));
			r.types.push_back(t.aggs[i].out_type);
		}
		VGPU_HIP(hipStreamSynchronize(stream));
		r.rows = groups;
		r.hold = hold;
		return;
	}
	// JOIN_AGG on the plain sort-merge join: both sides sorted and reduced for this query.
	std::vector<JoinAggReq> aggs(2);
	aggs[0].kind = AGG_SUM;
	aggs[0].side = 1;
	aggs[0].input = 0;
	aggs[0].out_type = VT_HUGEINT;
	aggs[1].kind = AGG_COUNT_STAR;
	aggs[1].out_type = VT_INT64;
	auto values = JoinAggregateDevice(view[0], {}, t.build_key, {t.build_payload}, aggs, arena, stream);
	const uint64_t pairs = values[1].lo;
	FinishJoinAggregate(values[0], AGG_SUM, VT_HUGEINT, pairs, r.scalars[0]);
	FinishJoinAggregate(values[1], AGG_COUNT_STAR, VT_INT64, pairs, r.scalars[1]);
	r.rows = 1;
	r.on_device = false;
	r.data = {r.scalars[0], r.scalars[1]};
	r.types = {VT_HUGEINT, VT_INT64};
	r.null = {(uint8_t)(pairs == 0), 0};
}

void Server::RunCross(const ServedTemplate &t, const Pending &q, hipStream_t stream, ServedResult &r) {
	// The fused cross kernel over the left side x rows [start, start + count) of the right side.
	PoolArena arena(stream);
	auto hold = std::make_shared<BatchBuffers>();
	auto &pool = DevicePool::Instance();
	const uint64_t nl = t.left_cols[0].rows;
	const uint64_t pairs = nl * q.count;
	FusedSpec spec;
	spec.cross = true;
	spec.fast_math = t.fast_math;
	std::vector<const void *> ptrs;
	for (auto &c : t.left_cols) {
		FusedColumn fc;
		fc.vtype = c.type;
		fc.side = 0;
		spec.cols.push_back(fc);
		ptrs.push_back(c.data);
	}
	for (auto &c : t.cols) {
		FusedColumn fc;
		fc.vtype = c.type;
		fc.side = 1;
		spec.cols.push_back(fc);
		ptrs.push_back(static_cast<const uint8_t *>(c.data) + q.start * TypeWidth(c.type));
	}
	spec.predicates = t.preds;
	spec.projections = t.projs;
	auto kernel = GetFusedKernel(spec);
	std::vector<void *> outs;
	for (auto &p : t.projs) {
		void *o = pool.Alloc(std::max<uint64_t>(pairs, 1) * TypeWidth(p.out_type), stream);
		hold->blocks.push_back(o);
		outs.push_back(o);
	}
	uint64_t kept = pairs;
	FusedLaunch la;
	la.cols = ptrs.data();
	la.outs = outs.data();
	la.n = pairs;
	la.right_rows = q.count;
	la.capacity = std::max<uint64_t>(pairs, 1);
	void *state = nullptr;
	size_t state_bytes = 0;
	if (!t.preds.empty()) {
		la.count_device = static_cast<uint64_t *>(arena.Alloc(8));
		state_bytes = FusedKernel::StateBytes(pairs);
		state = arena.Alloc(std::max<size_t>(state_bytes, 8));
	}
	kernel->Launch(la, state, state_bytes, stream);
	if (!t.preds.empty() && pairs > 0) {
		VGPU_HIP(hipMemcpyAsync(&kept, la.count_device, 8, hipMemcpyDeviceToHost, stream));
	} else if (!t.preds.empty()) {
		kept = 0;
	}
	VGPU_HIP(hipStreamSynchronize(stream));
	r.rows = kept;
	for (size_t i = 0; i < outs.size(); i++) {
		r.data.push_back(outs[i]);
		r.types.push_back(t.projs[i].out_type);
	}
	r.hold = hold;
}

double RunOpenLoop(Server &server, const int32_t *tmpl, const uint64_t *starts, const uint64_t *counts,
                   const double *arrival_s, uint64_t n, double *latency_s, uint32_t *batch_sizes) {
	if (n == 0) {
		return 0.0;
	}
	std::vector<uint64_t> tickets(n);
	std::vector<std::atomic<bool>> ready(n);
	for (auto &r : ready) {
		r.store(false);
	}
	const uint64_t t0 = Server::NowNs() + 2000000; // 2 ms to start the submitter
	std::string submit_error;
	std::thread submitter([&] {
		try {
			for (uint64_t i = 0; i < n; i++) {
				const uint64_t target = t0 + (uint64_t)(arrival_s[i] * 1e9);
				for (;;) {
					const uint64_t now = Server::NowNs();
					if (now >= target) {
						break;
					}
					if (target - now > 300000) {
						std::this_thread::sleep_for(std::chrono::nanoseconds(target - now - 200000));
					}
				}
				tickets[i] = server.Submit(tmpl[i], starts[i], counts[i], target);
				ready[i].store(true, std::memory_order_release);
			}
		} catch (const std::exception &e) {
			submit_error = e.what();
			for (uint64_t i = 0; i < n; i++) {
				if (!ready[i].load()) {
					tickets[i] = 0;
					ready[i].store(true, std::memory_order_release);
				}
			}
		}
	});
	uint64_t last = t0;
	std::string first_error;
	for (uint64_t i = 0; i < n; i++) {
		while (!ready[i].load(std::memory_order_acquire)) {
			std::this_thread::yield();
		}
		if (tickets[i] == 0) {
			continue;
		}
		auto r = server.Wait(tickets[i]);
		if (r->status != 0 && first_error.empty()) {
			first_error = r->error;
		}
		latency_s[i] = (double)(r->done_ns - r->submit_ns) * 1e-9;
		if (batch_sizes != nullptr) {
			batch_sizes[i] = r->batch_size;
		}
		last = std::max(last, r->done_ns);
	}
	submitter.join();
	if (!submit_error.empty()) {
		throw std::runtime_error(submit_error);
	}
	if (!first_error.empty()) {
		throw std::runtime_error("open loop: a query failed: " + first_error);
	}
	return (double)(last - t0) * 1e-9;
}

} // namespace vgpu

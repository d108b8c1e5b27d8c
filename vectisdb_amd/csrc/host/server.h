// server.h -- the served form: a request loop over device-resident tables with memory-footprint admission
// and continuous batching.
//
// Queries are instances of prepared templates (a filter/projection, a filter + GROUP BY, or a SUM/COUNT
// join aggregate over a resident table) applied to a row range of that table. `slots` executor threads
// each own a HIP stream. Whenever an executor is free it takes the oldest waiting query whose footprint
// fits the memory still unreserved and, with batching on, every later waiting query of the same template
// that still fits (up to max_batch): the batch runs as ONE pipeline (row-range fused kernel, composite-key
// GROUP BY, or one probe pass against the template's resident join index). Queries that arrive while
// others run are admitted as executors and memory free up, and batches of different executors overlap on
// the device. With batching off each query runs alone on the plain per-query operators. A cross-product
// query (the whole left side times a row range of the right side) always runs as its own launch.
#pragma once

#include "operators.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace vgpu {

enum TemplateKind : int { TMPL_FILTER_PROJECT = 0, TMPL_GROUP_BY = 1, TMPL_JOIN_AGG = 2, TMPL_CROSS = 3 };

struct ServedTemplate {
	int kind = TMPL_FILTER_PROJECT;
	std::vector<DevCol> cols;        // the table; JOIN_AGG: cols[0] is the probe key; CROSS: the right side
	std::vector<DevCol> left_cols;   // CROSS: the whole left side (expression slots 0..nleft-1)
	uint64_t rows = 0;
	std::vector<FusedExpr> preds;    // FILTER_PROJECT, GROUP_BY
	std::vector<FusedExpr> projs;    // FILTER_PROJECT, CROSS
	int key_col = -1;                // GROUP_BY
	std::vector<int> value_cols;     // GROUP_BY (aggregate inputs index this list)
	std::vector<AggReq> aggs;        // GROUP_BY
	DevCol build_key, build_payload; // JOIN_AGG
	std::unique_ptr<JoinIndex> index; // JOIN_AGG, built once when batching
	bool fast_math = false;
};

//! Device buffers of one executed batch, shared by its queries' results.
struct BatchBuffers {
	std::vector<void *> blocks;
	~BatchBuffers();
};

struct ServedResult {
	int status = 0; // 0 ok, else the error text is set
	std::string error;
	int kind = 0;
	uint64_t rows = 0;
	std::vector<const void *> data; // device pointers (host for JOIN_AGG scalars)
	std::vector<int> types;
	bool on_device = true;
	std::vector<uint8_t> null;      // JOIN_AGG: per scalar
	uint8_t scalars[2][16] = {};
	std::shared_ptr<BatchBuffers> hold;
	uint32_t batch_size = 0;
	uint64_t submit_ns = 0, start_ns = 0, done_ns = 0;
};

struct ServerOptions {
	int slots = 4;
	bool batching = true;
	int max_batch = 512;
	uint64_t memory_budget = 0; // 0: three quarters of the free device memory at creation
	bool fast_math = false;
	uint64_t max_batch_rows = 0; // 0: no limit (a GROUP BY batch still sorts fewer than 2^30 rows)
	//! A JOIN_AGG template reduces its build side once and every query probes that resident index, with
	//! batching off too (batching on always does).
	bool resident_join = false;
};

struct ServerStats {
	uint64_t queries = 0;
	uint64_t batches = 0;
	uint64_t max_batch = 0;
	uint64_t peak_reserved = 0;
	uint64_t budget = 0;
	uint64_t rejected = 0;
};

class Server {
public:
	explicit Server(const ServerOptions &options);
	~Server(); // drains the queue, then joins the executors
	Server(const Server &) = delete;
	Server &operator=(const Server &) = delete;

	int AddTemplate(std::unique_ptr<ServedTemplate> t);
	//! Queues a query; `submit_ns` (steady clock) defaults to now.
	uint64_t Submit(int template_id, uint64_t row_start, uint64_t row_count, uint64_t submit_ns = 0);
	//! Blocks until the query is done; the caller owns the result.
	std::unique_ptr<ServedResult> Wait(uint64_t ticket);
	ServerStats Stats() const;
	//! Bytes reserved for a query or batch of `rows` launch rows of template `t`.
	uint64_t Footprint(const ServedTemplate &t, uint64_t rows, uint32_t nqueries) const;

	static uint64_t NowNs();

private:
	struct Pending {
		uint64_t ticket;
		int tmpl;
		uint64_t start, count;
		uint64_t submit_ns;
	};
	void Executor(int slot);
	void RunBatch(const ServedTemplate &t, const std::vector<Pending> &batch, hipStream_t stream,
	              std::vector<std::unique_ptr<ServedResult>> &results);
	void RunAlone(const ServedTemplate &t, const Pending &q, hipStream_t stream, ServedResult &r);
	void RunCross(const ServedTemplate &t, const Pending &q, hipStream_t stream, ServedResult &r);

	ServerOptions options_;
	int device_ = 0;
	std::vector<std::unique_ptr<ServedTemplate>> templates_;
	mutable std::mutex mutex_;
	std::condition_variable work_cv_, done_cv_;
	std::deque<Pending> queue_;
	std::unordered_map<uint64_t, std::unique_ptr<ServedResult>> done_;
	std::unordered_map<uint64_t, bool> open_; // submitted, not yet collected
	uint64_t next_ticket_ = 1;
	uint64_t reserved_ = 0;
	bool stop_ = false;
	ServerStats stats_;
	std::vector<hipStream_t> streams_;
	std::vector<std::thread> executors_;
};

//! Open-loop driver: query i is submitted at t0 + arrival_s[i] whatever the server's state; its latency is
//! done - (t0 + arrival_s[i]). Fills latency_s[n] and returns the elapsed seconds from t0 to the last
//! completion. Results are checked for errors and released.
double RunOpenLoop(Server &server, const int32_t *tmpl, const uint64_t *starts, const uint64_t *counts,
                   const double *arrival_s, uint64_t n, double *latency_s, uint32_t *batch_sizes);

} // namespace vgpu

// server_capi.cpp -- extern "C" boundary of the served form.
#include "capi_util.h"
#include "server.h"

#include <cstring>
#include <memory>
#include <stdexcept>

struct vgpu_server {
	std::unique_ptr<vgpu::Server> impl;
};
struct vgpu_served_result {
	std::unique_ptr<vgpu::ServedResult> impl;
};

using namespace vgpu;

namespace {
std::vector<FusedExpr> ToExprs(const vgpu_expr *e, int32_t n, const char *what) {
	if (n < 0 || (n > 0 && e == nullptr)) {
		throw std::runtime_error(std::string("server template: bad ") + what);
	}
	std::vector<FusedExpr> out;
	for (int32_t i = 0; i < n; i++) {
		if (e[i].body == nullptr || (e[i].ninputs > 0 && e[i].inputs == nullptr)) {
			throw std::runtime_error(std::string("server template: ") + what + " without body or inputs");
		}
		FusedExpr f;
		f.body = e[i].body;
		f.out_type = e[i].output_type;
		f.slots.assign(e[i].inputs, e[i].inputs + e[i].ninputs);
		out.push_back(f);
	}
	return out;
}
} // namespace

extern "C" {

int vgpu_server_create(const vgpu_server_options *options, vgpu_server **server) {
	return Guard([&] {
		if (server == nullptr) {
			throw std::runtime_error("server: null output");
		}
		ServerOptions o;
		if (options != nullptr) {
			o.slots = options->slots == 0 ? 4 : options->slots;
			o.batching = options->batching != 0;
			o.max_batch = options->max_batch == 0 ? 512 : options->max_batch;
			o.memory_budget = options->memory_budget;
			o.fast_math = options->fast_math != 0;
			o.max_batch_rows = options->max_batch_rows;
			o.resident_join = options->resident_join != 0;
		}
		auto s = std::unique_ptr<vgpu_server>(new vgpu_server());
		s->impl.reset(new Server(o));
		*server = s.release();
	});
}

int vgpu_server_destroy(vgpu_server *server) {
	return Guard([&] { delete server; });
}

int vgpu_server_add_template(vgpu_server *server, const vgpu_template *tmpl, int32_t *template_id) {
	return Guard([&] {
		if (server == nullptr || tmpl == nullptr || template_id == nullptr || tmpl->ncols <= 0 ||
		    tmpl->cols == nullptr) {
			throw std::runtime_error("server template: bad arguments");
		}
		auto t = std::unique_ptr<ServedTemplate>(new ServedTemplate());
		t->kind = tmpl->kind;
		t->cols = ToDevCols(tmpl->cols, tmpl->ncols);
		t->preds = ToExprs(tmpl->preds, tmpl->npreds, "predicates");
		t->projs = ToExprs(tmpl->projs, tmpl->nprojs, "projections");
		t->key_col = tmpl->key_col;
		if (tmpl->nvalues < 0 || tmpl->naggs < 0 || (tmpl->nvalues > 0 && tmpl->value_cols == nullptr) ||
		    (tmpl->naggs > 0 && tmpl->aggs == nullptr)) {
			throw std::runtime_error("server template: bad value columns or aggregates");
		}
		t->value_cols.assign(tmpl->value_cols, tmpl->value_cols + tmpl->nvalues);
		for (int32_t i = 0; i < tmpl->naggs; i++) {
			t->aggs.push_back({tmpl->aggs[i].kind, tmpl->aggs[i].input, tmpl->aggs[i].output_type});
		}
		if (tmpl->kind == VGPU_TEMPLATE_CROSS) {
			if (tmpl->nleft <= 0 || tmpl->left_cols == nullptr) {
				throw std::runtime_error("server template: a cross product needs left columns");
			}
			t->left_cols = ToDevCols(tmpl->left_cols, tmpl->nleft);
		}
		if (tmpl->kind == VGPU_TEMPLATE_JOIN_AGG) {
			t->build_key = ToDevCols(&tmpl->build_key, 1)[0];
			t->build_payload = ToDevCols(&tmpl->build_payload, 1)[0];
		}
		*template_id = server->impl->AddTemplate(std::move(t));
	});
}

int vgpu_server_submit(vgpu_server *server, int32_t template_id, uint64_t row_start, uint64_t row_count,
                       uint64_t *ticket) {
	return Guard([&] { *ticket = server->impl->Submit(template_id, row_start, row_count); });
}

int vgpu_server_wait(vgpu_server *server, uint64_t ticket, vgpu_served_result **result) {
	return Guard([&] {
		auto r = std::unique_ptr<vgpu_served_result>(new vgpu_served_result());
		r->impl = server->impl->Wait(ticket);
		*result = r.release();
	});
}

int vgpu_server_stats_read(vgpu_server *server, vgpu_server_stats *stats) {
	return Guard([&] {
		const ServerStats s = server->impl->Stats();
		stats->queries = s.queries;
		stats->batches = s.batches;
		stats->max_batch = s.max_batch;
		stats->peak_reserved = s.peak_reserved;
		stats->budget = s.budget;
		stats->rejected = s.rejected;
	});
}

int vgpu_served_result_info(const vgpu_served_result *result, vgpu_served_info *info) {
	return Guard([&] {
		const auto &r = *result->impl;
		info->status = r.status;
		info->kind = r.kind;
		info->rows = r.rows;
		info->ncols = (int32_t)r.data.size();
		info->on_device = r.on_device ? 1 : 0;
		info->batch_size = r.batch_size;
		info->submit_ns = r.submit_ns;
		info->start_ns = r.start_ns;
		info->done_ns = r.done_ns;
	});
}

int vgpu_served_result_column(const vgpu_served_result *result, int32_t i, const void **data, int32_t *type,
                              int32_t *is_null) {
	return Guard([&] {
		const auto &r = *result->impl;
		if (i < 0 || i >= (int32_t)r.data.size()) {
			throw std::runtime_error("served result: column index out of range");
		}
		*data = r.data[i];
		*type = r.types[i];
		if (is_null != nullptr) {
			*is_null = (size_t)i < r.null.size() ? r.null[i] : 0;
		}
	});
}

int vgpu_served_result_copy(const vgpu_served_result *result, int32_t i, void *dst) {
	return Guard([&] {
		const auto &r = *result->impl;
		if (i < 0 || i >= (int32_t)r.data.size()) {
			throw std::runtime_error("served result: column index out of range");
		}
		const size_t bytes = (size_t)r.rows * TypeWidth(r.types[i]);
		if (bytes == 0) {
			return;
		}
		if (r.on_device) {
			VGPU_HIP(hipMemcpy(dst, r.data[i], bytes, hipMemcpyDeviceToHost));
		} else {
			std::memcpy(dst, r.data[i], bytes);
		}
	});
}

const char *vgpu_served_result_error(const vgpu_served_result *result) {
	return result == nullptr ? "" : result->impl->error.c_str();
}

void vgpu_served_result_free(vgpu_served_result *result) {
	delete result;
}

int vgpu_server_run_open_loop(vgpu_server *server, const int32_t *tmpl, const uint64_t *starts,
                              const uint64_t *counts, const double *arrival_s, uint64_t n, double *latency_s,
                              uint32_t *batch_sizes, double *elapsed_s) {
	return Guard([&] {
		*elapsed_s = RunOpenLoop(*server->impl, tmpl, starts, counts, arrival_s, n, latency_s, batch_sizes);
	});
}

} // extern "C"

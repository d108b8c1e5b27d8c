// kernels.h -- host-side launch interface of the vectis HIP kernels (internal to the library).
//
// The C++ engine (host/) and the extern "C" boundary (capi.cpp) call these; nothing here is visible to
// the DuckDB extension. All pointers are device pointers unless stated otherwise; every launcher is
// asynchronous on `stream` unless its comment says it synchronizes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace vgpu {

//! Value types, numbered exactly like vector_gpu::GpuValueType (gpu_engine.hpp).
enum VType : int {
	VT_INT16 = 0,
	VT_INT32 = 1,
	VT_INT64 = 2,
	VT_FLOAT32 = 3,
	VT_FLOAT64 = 4,
	VT_BOOLEAN = 5,
	VT_DICTIONARY_STRING = 6,
	VT_HUGEINT = 7,
};

int TypeWidth(int vtype); // bytes per element on the device; -1 for DICTIONARY_STRING
bool TypeIsSigned(int vtype);
bool TypeIsInteger(int vtype); // INT16/32/64 and BOOLEAN

//! Thrown by launchers; the engine turns it into success=false.
struct HipError : std::runtime_error {
	using std::runtime_error::runtime_error;
};

void Check(hipError_t status, const char *what);
#define VGPU_HIP(call) ::vgpu::Check((call), #call)

//! Scratch allocator the launchers draw temporaries from. The engine passes its per-plan arena; the C ABI
//! passes a pooled arena. Freed when the allocator is destroyed (after the caller's stream work is
//! ordered, see DevicePool).
struct Scratch {
	virtual ~Scratch() = default;
	virtual void *Alloc(size_t bytes) = 0;
};

// ------------------------------------------------------------------ profiling (HIP events per launch)

//! When enabled, every launcher brackets each kernel with hipEvents on its launch stream; totals are read
//! with ProfileSnapshot. Used by the bench to time the dominant kernel on its own stream.
void ProfileEnable(bool on);
bool ProfileEnabled();
//! Times only the scopes named `name` while enabled (nullptr or "": every scope). Each timed scope records
//! two events on its stream, so timing one kernel of an operator leaves the others' launches unchanged.
void ProfileOnly(const char *name);
struct ProfileEntry {
	std::string name;
	double total_ms = 0;
	uint64_t launches = 0;
	double bytes = 0; // algorithmic bytes the launcher attributes to these launches
};
std::vector<ProfileEntry> ProfileSnapshot(); // synchronizes the device
void ProfileReset();

//! RAII bracket used by launchers.
class ProfileScope {
public:
	ProfileScope(const char *name, hipStream_t stream, double bytes);
	~ProfileScope();

private:
	const char *name_;
	hipStream_t stream_;
	double bytes_;
	hipEvent_t start_ = nullptr;
	hipEvent_t stop_ = nullptr;
};

// ------------------------------------------------------------------ key packing

constexpr int kMaxKeyCols = 4;
constexpr int kMaxPayload = 8;
constexpr int kMaxComponents = 8;

//! Packs up to kMaxKeyCols integer key columns into one order-preserving unsigned key:
//! packed = sum_c (Ordered(col_c[row]) - min_c) << shift_c. The first column is most significant.
struct KeyPack {
	int ncols = 0;
	const void *col[kMaxKeyCols] = {};
	int width[kMaxKeyCols] = {};
	int is_signed[kMaxKeyCols] = {};
	uint64_t min[kMaxKeyCols] = {};
	int shift[kMaxKeyCols] = {};
	int bits[kMaxKeyCols] = {};
	int total_bits = 0; // sum of bits; <= 64
	uint64_t max_packed = 0; // largest packed key the ranges allow (0: not known, 2^total_bits - 1 is assumed)
};

//! Ordered min/max of each key column (synchronizes: the host needs them to plan the passes).
void KeyRanges(const void *const *cols, const int *vtypes, int ncols, uint64_t n, uint64_t *min_out,
               uint64_t *max_out, Scratch &scratch, hipStream_t stream);

//! Builds the packing from the ranges. Returns false when the packed key needs more than 64 bits.
bool PlanKeyPack(const void *const *cols, const int *vtypes, int ncols, const uint64_t *mins, const uint64_t *maxs,
                 KeyPack &pack);

//! Single integer key column, one read: ordered min/max plus the 256-bin histogram of each of the eight
//! byte-aligned digits of the ordered key (hist8_dev: 8 x 256 uint32 on the device, zeroed here).
//! Synchronizes (min/max are read back).
void KeyRangeHist8(const void *col, int vtype, uint64_t n, uint64_t *min_out, uint64_t *max_out,
                   uint32_t *hist8_dev, hipStream_t stream);

//! Packing for KeyRangeHist8: the varying span of the ordered key rounded up to whole bytes, minus the
//! common prefix, so digit p of the packed key is byte p of the ordered key and hist8 row p is exactly
//! that pass's histogram.
void PlanKeyPackAligned(const void *col, int vtype, uint64_t omin, uint64_t omax, KeyPack &pack);

// ------------------------------------------------------------------ radix sort

//! Up to four spans of 32-bit words that one launch clears (ZeroU32Spans): look-back words, tickets and flags
//! of consecutive steps of an operator, zeroed together instead of by one fill launch each.
struct ZeroSpans {
	static constexpr int kMax = 4;
	uint32_t *ptr[kMax] = {};
	uint64_t words[kMax] = {};
	int count = 0;
	//! Adds a span (empty spans are skipped); more than kMax spans throw.
	void Add(uint32_t *p, uint64_t n);
};

struct SortTuning {
	int radix_bits = 8;      // 8 or 11 (digit width upper bound)
	int passes_override = 0; // 0 = derive from total_bits
	int items = 16;          // items per thread of a pass with 8-bit digits: 16 or 32 (tile = block x items)
	int block = 256;         // threads per tile with 8-bit digits and 16 items: 256 or 512 (tile 4096 or 8192)
	int prefetch = 1;        // load payload values before ranking (1) or just before their scatter (0)
	bool fixed = false;      // variant forced (environment override): no defaults, no autotuning
	//! Optional precomputed per-byte histograms (KeyRangeHist8) for a byte-aligned pack: skips the
	//! histogram pass.
	const uint32_t *hist8 = nullptr;
	//! Payload narrowing (one key column only): bit p = payload p is an INT64 column whose values the
	//! caller expects in the INT32 range, given with payload_width 4. The first pass reads it whole and
	//! carries the low 32 bits; any value outside that range sets *overflow (device word, cleared by the
	//! caller), and the sort's payload is then wrong: the caller sorts again without narrowing.
	uint32_t narrow = 0;
	uint32_t *overflow = nullptr;
	//! Leave the last (most significant) digit pass to the caller when it is an 8-bit pass that is not the
	//! first: SortedRun then holds the keys ordered by every bit below that digit (see RunReduceLastPass).
	bool leave_last = false;
	//! The caller's spans that the sort clears in the same launch as its look-back state, before its first
	//! pass (the overflow word above, the fused last pass's state).
	ZeroSpans zero_also;
};

//! Output of SortPacked: the compact keys (uint32 when total_bits <= 32, else uint64) in ascending order and
//! each payload column permuted identically (stable). Buffers come from `scratch`.
struct SortedRun {
	void *keys = nullptr;
	bool keys64 = false;
	void *payload[kMaxPayload] = {};
	int payload_width[kMaxPayload] = {};
	int npayload = 0;
	int passes = 0;
	int radix_bits = 0;
	//! Set when the sort left its last digit pass (SortTuning::leave_last): that digit's bits, its row
	//! histogram on the device, and the largest packed key.
	bool last_pending = false;
	int last_shift = 0;
	int last_bits = 0;
	const uint32_t *last_hist = nullptr;
	uint64_t max_packed = 0;
};

//! LSD radix sort of the packed key with payload. payload_in[i] == nullptr with width 4 means "generate
//! row ids" (a uint32 permutation). Requires n < 2^30.
void SortPacked(const KeyPack &pack, const void *const *payload_in, const int *payload_width, int npayload,
                uint64_t n, const SortTuning &tuning, SortedRun &out, Scratch &scratch, hipStream_t stream,
                const uint64_t *n_dev = nullptr);
//! SortPacked with exactly the given variant (no defaults, no autotuning). With n_dev the row count is read
//! by the kernels from device memory (a count an earlier kernel on `stream` wrote, at most n): buffers and
//! grids are sized for n, blocks past the count leave at once, and the host never reads the count, so the
//! sort can be captured into a graph.
void SortPackedWith(const KeyPack &pack, const void *const *payload_in, const int *payload_width, int npayload,
                    uint64_t n, const SortTuning &tuning, SortedRun &out, Scratch &scratch, hipStream_t stream,
                    const uint64_t *n_dev = nullptr);

//! Sort autotuner: when on, the first sort of each shape (digit passes, rows rounded to a power of two,
//! payload columns and bytes) times every variant on that sort's own data and keeps the fastest for the
//! shape; later sorts of the shape use it. Off by default (the measured static choice is used).
void SortAutotuneEnable(bool on);
bool SortAutotuneEnabled();
//! JSON array of the shapes tuned so far and their choice.
std::string SortAutotuneReport();

// ------------------------------------------------------------------ segmented multi-aggregate reduction

//! Component kinds of a group's partial state. Aggregates map onto components (deduplicated), so
//! SUM(v)/AVG(v) share one SUM_DD(v) and every COUNT shares the one COUNT.
enum CompKind : int {
	COMP_COUNT = 0,   // uint64
	COMP_SUM_I128 = 1, // int128 over an integer column
	COMP_SUM_DD = 2,   // double-double over a FLOAT/DOUBLE column
	COMP_MIN_I64 = 3,  // signed min over an integer column (widened)
	COMP_MAX_I64 = 4,
	COMP_MIN_F64 = 5,  // floating min (DuckDB order: NaN is greatest)
	COMP_MAX_F64 = 6,
};

struct Component {
	int kind = COMP_COUNT;
	int payload = -1; // index into the sorted payload columns
	int in_type = VT_INT64;
};

//! Output aggregate kinds, numbered like vector_gpu::GpuAggregateKind.
enum AggKind : int { AGG_COUNT_STAR = 0, AGG_COUNT = 1, AGG_SUM = 2, AGG_MIN = 3, AGG_MAX = 4, AGG_AVG = 5 };

struct AggOut {
	int kind = AGG_COUNT_STAR;
	int comp = 0;        // main component
	int out_type = VT_INT64;
	void *out = nullptr; // device output column (capacity rows)
};

struct ReduceSpec {
	int ncomp = 0;
	Component comp[kMaxComponents];
	int naggs = 0;
	AggOut aggs[16];
	// group key reconstruction
	KeyPack pack;
	void *key_out[kMaxKeyCols] = {};
	uint64_t capacity = 0; // rows of every output (0: the input row count)
};

//! Reduces sorted keys/payload into one row per distinct key. `keys` may be nullptr: every row is then one
//! group (global aggregate). Writes the group count to *groups_device (device uint64). With n_dev the row
//! count is read on the device as in SortPackedWith (n is its bound).
void SegmentedReduce(const SortedRun &run, uint64_t n, const ReduceSpec &spec, uint64_t *groups_device,
                     Scratch &scratch, hipStream_t stream, const uint64_t *n_dev = nullptr);

//! The GROUP BY's last digit pass fused with the reduction, for a SortedRun whose sort left that pass
//! (last_pending): each 8192-row tile is ranked by the digit in LDS, which leaves its rows in full-key order,
//! and reduces its runs of equal keys to one partial state each; a small reduction then combines each key's
//! partials into its group. Same outputs as SegmentedReduce over the fully sorted run.
//! `state`, when given, is RunReduceStateWords(n) words already zero (cleared with the sort's own state);
//! otherwise the pass allocates and clears it.
void RunReduceLastPass(const SortedRun &run, uint64_t n, const ReduceSpec &spec, uint64_t *groups_device,
                       Scratch &scratch, hipStream_t stream, uint32_t *state = nullptr);
//! Look-back words and ticket of the fused last pass over n rows.
uint64_t RunReduceStateWords(uint64_t n);
//! The aggregate mixes the fused pass is compiled for (COUNT alone, or with one SUM, or the
//! COUNT/SUM/MIN/MAX + DOUBLE SUM mix, with the integer column carried at 4 or 8 bytes).
bool RunReduceSupported(const ReduceSpec &spec, int npayload, bool keys64);
//! Partial states the fused pass may write: per value of the last digit, at most one per distinct value of the
//! bits below it plus one per tile boundary.
uint64_t RunReduceRecordBound(uint64_t n, int shift, int bits, uint64_t max_packed);

// ------------------------------------------------------------------ join

//! Per-key state of one side after its own reduce-by-key: unique compact keys + component columns.
struct KeyedStates {
	void *keys = nullptr; // compact keys, ascending, unique
	bool keys64 = false;
	uint64_t groups = 0;
	int ncomp = 0;
	Component comp[kMaxComponents];
	void *state = nullptr; // groups x ncomp x 16 bytes (component slots)
};

//! Aggregate-only INNER join: merge-path over two unique-key state runs, accumulating
//! COUNT(*) = sum ca*cb and the per-side components; results land in host `result` (16 bytes per output).
struct JoinAggSpec {
	int naggs = 0;
	struct {
		int kind;     // AggKind
		int side;     // 0 = left (probe/children[0]), 1 = right, -1 = COUNT_STAR
		int comp;     // component index on that side
		int out_type; // VType
	} aggs[16];
};

// ------------------------------------------------------------------ misc primitives

//! out[i] = src[idx[i]] for fixed-width elements (1/2/4/8/16 bytes); idx is uint32 or uint64.
void Gather(const void *src, int width, const void *idx, bool idx64, uint64_t n, void *out, hipStream_t stream);

//! Dense byte validity from an Arrow LSB-first bitmap.
void UnpackValidity(const uint8_t *packed, uint8_t *out, uint64_t n, hipStream_t stream);

//! Arrow LSB-first bitmap ((n + 7) / 8 bytes) from dense byte validity.
void PackValidity(const uint8_t *dense, uint8_t *packed, uint64_t n, hipStream_t stream);

//! Zero-fill / value-fill without hipMemset (safe inside captured graphs).
void FillU8(uint8_t *dst, uint8_t value, uint64_t n, hipStream_t stream);
void FillU32(uint32_t *dst, uint32_t value, uint64_t n, hipStream_t stream);
void FillU64(uint64_t *dst, uint64_t value, uint64_t n, hipStream_t stream);
//! Clears every span of `spans` in one launch (nothing when it holds none).
void ZeroU32Spans(const ZeroSpans &spans, hipStream_t stream);
//! Small device-to-host read-back on `stream` that the host waits for: through a page-locked staging buffer
//! of the calling thread (up to 4 KiB; larger reads copy directly), with the host polling the stream instead
//! of sleeping in hipStreamSynchronize (a planning read-back sits between two kernels of one operator, so
//! its wake-up latency is on the critical path). Synchronizes `stream`.
void ReadBack(void *host, const void *dev, size_t bytes, hipStream_t stream);

//! Device-to-device copy on `stream` (a block-chunked copy kernel; falls back to hipMemcpyAsync when a
//! pointer is not 16-byte aligned).
void CopyDevice(void *dst, const void *src, uint64_t bytes, hipStream_t stream);

//! Left-major cross product expansion of one column: dst[i] = src[from_left ? i / m : i % m].
void CrossExpand(const void *src, int width, uint64_t m, bool from_left, uint64_t total, void *dst,
                 hipStream_t stream);

//! Seeded counter-based column generator (dist: 0 uniform int [lo, hi], 1 uniform double [0,1), 2 normal,
//! 3 Zipf via `cdf`/`perm` tables on the device, 4 sequence lo + i).
void GenerateColumn(void *out, int vtype, uint64_t n, int dist, long long lo, long long hi, uint64_t seed,
                    uint64_t stream_id, const double *cdf, uint64_t cdf_len, const long long *perm,
                    hipStream_t stream);

//! In-place running sum with a carried state (out[i] = state + x[0..i]; state = out[n-1]); integers wrap.
//! `scratch` holds RunningSumScratchBytes(n) bytes.
size_t RunningSumScratchBytes(uint64_t n);
void RunningSum(int vtype, void *data, uint64_t n, void *state, void *scratch, size_t scratch_bytes,
                hipStream_t stream);

//! SUM/COUNT of an INNER equi-join of many probe row ranges against a build-side index (the distinct build
//! keys ascending, each with its row count and int128 payload sum as lo/hi words). Range g covers launch
//! rows [seg[g], seg[g] + seg_len[g]) read from table rows seg_row[g] + (v - seg[g]); seg has nseg + 1
//! entries, seg[nseg] = nvirt. acc (3 x nseg uint64: matched rows, sum lo, sum hi) is zeroed here. With
//! `dense` (see BuildDenseIndex) a probe key is looked up directly instead of by binary search.
void RangeProbeAgg(int key_type, const void *probe_keys, const uint64_t *seg, const uint64_t *seg_row,
                   const uint64_t *seg_len, uint32_t nseg, uint64_t nvirt, const void *index_keys,
                   const uint64_t *index_count, const uint64_t *index_sum, uint64_t nd, const uint64_t *dense,
                   uint64_t dense_min, uint64_t dense_span, uint64_t *acc, hipStream_t stream);

//! SUM/COUNT of an equi-join of one probe column against a direct-address table (see BuildDenseIndex), for a
//! probe side much larger than the table, in two launches that may run on different streams: the probe
//! keys are grouped by key range (256 buckets, batch by batch; LaunchProbePartition), then each bucket
//! reads its slice of the table once, from LDS when the table has a 16-bit form (LaunchPartitionedProbe).
//! The partition needs only the table's key range, so it can run while the table is being built.
struct ProbePartition {
	int key_type = VT_INT64;
	uint64_t n = 0;
	bool has_payload = false;
	int payload_width = 0;
	uint64_t dense_min = 0; // ordered as uint64 of the sign-extended key
	uint64_t span = 0;
	int shift = 0;           // bucket = (key - dense_min) >> shift
	uint64_t batch = 0;      // probe rows per scatter batch
	uint64_t nbatches = 0;
	uint32_t *part = nullptr; // table offsets of the probe rows, grouped by bucket within each batch
	void *part_y = nullptr;   // their payload values (4 or 8 bytes), in the same order
	uint32_t *tab = nullptr;  // per batch: 257 bucket starts
	uint64_t *partials = nullptr;
};
//! Checks the shape and allocates the partition's buffers from `scratch`: about n * 4.03 bytes (n * 8 more
//! with a probe payload, an integer column of payload_width bytes).
ProbePartition PlanProbePartition(int key_type, uint64_t n, bool has_payload, int payload_width,
                                  uint64_t dense_min, uint64_t dense_span, Scratch &scratch);
//! Groups the probe keys (and payload) by key range into the plan's buffers.
void LaunchProbePartition(const ProbePartition &p, const void *probe_keys, const void *probe_payload,
                          hipStream_t stream);
//! Probes the partition against the table (dense16: its 16-bit form, or nullptr to read the 64-bit table
//! from L2). acc (5 words) = {matched build rows, build payload sum lo, hi, and -- with a probe payload --
//! the sum of each probe row's payload times its key's build row count, lo, hi}.
void LaunchPartitionedProbe(const ProbePartition &p, const uint64_t *dense, const uint16_t *dense16, uint64_t *acc,
                            hipStream_t stream);
//! The table's 16-bit form (4-bit count, 12-bit sum) when its span allows it (else nullptr); *nofit_dev
//! becomes nonzero when an entry does not fit, in which case the form must not be used. No host sync.
const uint16_t *LaunchNarrowDense(const uint64_t *dense, uint64_t dense_span, unsigned int *nofit_dev,
                                  Scratch &scratch, hipStream_t stream);

//! Direct-address table (BuildDenseIndex's format) straight from a sorted build side: keys are the packed
//! sort keys of one column (key = ((packed + pack_min) ^ flip) as sign-extended value; `flip` is the top
//! bit for signed types), payload the rows' payload in the same order (nullptr: count only; raw integer
//! of payload_width bytes, sign-extended except for width 1). Returns flags: bit 0 an entry does not
//! pack, bit 1 a key has more than 64 rows (the table is then incomplete). Synchronizes.
unsigned BuildDenseFromRun(const void *keys, bool keys64, const void *payload, int payload_width, uint64_t n,
                           uint64_t pack_min, uint64_t flip, uint64_t dense_min, uint64_t span, uint64_t *dense,
                           unsigned int *flags_dev, hipStream_t stream);
//! The same without the read-back: the flags stay in *flags_dev.
void LaunchDenseFromRun(const void *keys, bool keys64, const void *payload, int payload_width, uint64_t n,
                        uint64_t pack_min, uint64_t flip, uint64_t dense_min, uint64_t span, uint64_t *dense,
                        unsigned int *flags_dev, hipStream_t stream);

//! Direct-address form of a join index whose keys span [min, min + span): dense[key - min] packs the key's
//! row count (top 23 bits) and payload sum (41-bit two's complement); 0 = no build row. Returns false (the
//! table is then unusable) when a count or sum does not fit. Synchronizes.
bool BuildDenseIndex(int key_type, const void *keys, const uint64_t *count, const uint64_t *sum, uint64_t nd,
                     uint64_t min, uint64_t span, uint64_t *dense, unsigned int *overflow_dev, hipStream_t stream);

//! Order-preserving compaction of the signed keys with lo <= x <= hi into `out` (n slots); returns the kept
//! count. Synchronizes.
uint64_t KeepKeyRange(int vtype, const void *in, uint64_t n, long long lo, long long hi, void *out, Scratch &scratch,
                      hipStream_t stream);

//! Order-preserving compaction of x > threshold (the FilterGreaterThanInt32 primitive).
void FilterGtI32(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_device,
                 Scratch &scratch, hipStream_t stream);

//! The same compaction with its look-back words, tile ticket and count in a fixed per-device array instead of
//! allocated scratch: no allocation of any kind. *count_host receives the kept count (synchronizes). Throws
//! when n exceeds FilterGtI32NoAllocMaxRows(). Calls on one GPU share the array, so overlapping calls race.
void FilterGtI32NoAlloc(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_host,
                        hipStream_t stream);
uint64_t FilterGtI32NoAllocMaxRows();
//! Resolves the current device's copy of the fixed scratch array (once per device).
void NoAllocScratchInit();

} // namespace vgpu

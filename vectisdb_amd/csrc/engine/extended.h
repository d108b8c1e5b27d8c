// extended.h -- device helpers of the opt-in semantics beyond the reference operators: NULL-aware GROUP BY
// and join, FLOAT/DOUBLE group and join keys with DuckDB's equality, composite and mixed-width join keys.
// Each helper rewrites columns so the existing integer-key operators compute the extended result.
#pragma once

#include "../hip/kernels.h"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vgpu {

//! FLOAT (vtype VT_FLOAT32) or DOUBLE (VT_FLOAT64) key column -> INT32 / INT64 column whose signed order is
//! DuckDB's order of the values and whose equality is DuckDB's equality: -0.0 and +0.0 map to one value,
//! every NaN to one value above +inf. Written as the bits b of the canonical value with every bit but the
//! sign flipped when the sign is set, which makes the mapping its own inverse on canonical values.
void FloatKeyToOrdered(const void *src, int vtype, uint64_t n, void *dst, hipStream_t stream);
//! The inverse: ordered INT32 / INT64 -> the canonical FLOAT / DOUBLE value (+0.0 for zero, the default
//! quiet NaN for NaN).
void OrderedToFloatKey(const void *src, int vtype, uint64_t n, void *dst, hipStream_t stream);

//! dst[i] = valid[i] != 0 ? src[i] : fill, for `width`-byte values (1, 2, 4, 8); fill is the value's bits.
void MaskNulls(const void *src, const uint8_t *valid, int width, uint64_t fill, uint64_t n, void *dst,
               hipStream_t stream);
//! dst[i] = valid[i] != 0 ? 0 : 1 (a BOOLEAN "is NULL" column).
void NullFlags(const uint8_t *valid, uint64_t n, uint8_t *dst, hipStream_t stream);
//! dst[i] = valid[i] != 0 ? 1 : 0 as INT32 (summed per group, it counts the non-NULL values).
void ValidAsInt32(const uint8_t *valid, uint64_t n, int32_t *dst, hipStream_t stream);
//! valid[i] = flag[i] == 0 (the validity of a group key from its "is NULL" key column).
void ValidFromNullFlags(const uint8_t *flag, uint64_t n, uint8_t *valid, hipStream_t stream);
//! valid[i] = values[i] != sentinel, for `width`-byte values (4 or 8): the validity of a key whose NULL was
//! folded into one reserved value.
void ValidUnlessEqual(const void *values, int width, uint64_t sentinel, uint64_t n, uint8_t *valid,
                      hipStream_t stream);

//! Completes one aggregate over a nullable column for `groups` groups: out_valid[g] = count[g] > 0; where
//! the group has no non-NULL value the output is zeroed. AGG_AVG: out[g] = sum[g] / count[g], `sum` being the
//! group's SUM output (HUGEINT when sum_type is VT_HUGEINT, else DOUBLE); other kinds leave `out` as the
//! reduction wrote it.
void FinishNullableAggregate(int kind, int out_type, void *out, uint8_t *out_valid, const void *sum, int sum_type,
                             const int64_t *count, uint64_t groups, hipStream_t stream);

//! dst[i] = the packed key of row i (KeyPack: ordered, minimum subtracted, shifted, OR-ed over columns).
void PackKeys(const KeyPack &pack, uint64_t n, uint64_t *dst, hipStream_t stream);

} // namespace vgpu

// gpu_semantics.hpp -- opt-in query semantics beyond the reference engine's operator rules.
//
// No reference counterpart: the reference's GROUP BY and join refuse NULL inputs, FLOAT/DOUBLE keys,
// composite join keys and MIN/MAX over floating columns (cuda_engine/src/operators/gpu_groupby.cu:24-33,
// cuda_engine/src/operators/gpu_hash_join.cu:25-38), and its tests pin those refusals, so the drop-in
// ExecutePlan keeps them unless a caller asks for more. With the extended semantics on, ExecutePlan
// follows DuckDB instead:
//   * GROUP BY: NULL keys form one group per key column (output key NULL); COUNT(col) counts non-NULL
//     values; SUM/MIN/MAX/AVG skip NULLs and are NULL for a group without a non-NULL value; FLOAT/DOUBLE
//     keys group under DuckDB's equality (-0.0 = +0.0, NaN = NaN; the group's key is output as +0.0 / NaN);
//     MIN/MAX over FLOAT/DOUBLE order NaN above every value; a keyless aggregate over no rows returns one
//     row (COUNT 0, every other aggregate NULL).
//   * HASH_JOIN (INNER): rows whose key (any key column) is NULL match nothing; several key columns per
//     side are compared as a tuple; integer key columns of different widths compare by value; FLOAT/DOUBLE
//     keys compare under DuckDB's equality; NULLs in output columns are carried through.
//   * NULLs are accepted in every plan shape (a CROSS_PRODUCT over nullable columns is still refused).
#pragma once

namespace vector_gpu {

//! Off unless SetGpuExtendedSemanticsEnabled(true) was called or VECTOR_GPU_EXTENDED_SEMANTICS is "1" or
//! "true"; the setter overrides the environment for the whole process.
bool IsGpuExtendedSemanticsEnabled();
void SetGpuExtendedSemanticsEnabled(bool enabled);

} // namespace vector_gpu

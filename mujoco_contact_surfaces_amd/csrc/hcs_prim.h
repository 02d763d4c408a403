/*
 * hcs_prim.h -- the engine's hand-written integer primitives
 * (hcs_prim.hip): exclusive scan of unsigned counts and stable LSD radix
 * sort of unsigned keys with int32 values.  All are asynchronous launches
 * on the given stream (capture-safe: no memset, no allocation); the caller
 * provides the temporary storage.  They return 0, or -1 for a size beyond
 * their limits (nothing is launched then).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace hcs {

size_t scan_temp_bytes(long long n);
/* out[i] = in[0] + ... + in[i-1]; in == out is allowed. */
int exclusive_scan_u32(hipStream_t s, const unsigned* in, unsigned* out, long long n, void* tmp, size_t tmp_bytes);

size_t sort_temp_bytes(long long n, int key_bytes);
/* Sorts (kin, vin) by bits [begin_bit, end_bit) of the keys into (kout,
 * vout), stable; the inputs are left unchanged. */
int radix_sort_pairs_u32(hipStream_t s, const unsigned* kin, unsigned* kout, const int* vin, int* vout, long long n,
                         int begin_bit, int end_bit, void* tmp, size_t tmp_bytes);
int radix_sort_pairs_u64(hipStream_t s, const unsigned long long* kin, unsigned long long* kout, const int* vin,
                         int* vout, long long n, int begin_bit, int end_bit, void* tmp, size_t tmp_bytes);

}  // namespace hcs

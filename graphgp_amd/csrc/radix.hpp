// radix.hpp -- stable LSD radix sort of (key, uint32 value) pairs.
#pragma once
#include "common.hpp"

namespace ggp {

// Workspace (bytes) needed to sort n pairs.
size_t radix_workspace(int64_t n);

// Sorts n pairs by key bits [begin_bit, end_bit) in 8-bit passes.  The data ping-pong
// between (*keys, *vals) and (keys_alt, vals_alt); on return *keys / *vals point at
// the buffers that hold the sorted result.
template <typename K>
int radix_sort_pairs(K** keys, uint32_t** vals, K* keys_alt, uint32_t* vals_alt, int64_t n, int begin_bit,
                     int end_bit, void* ws, size_t ws_bytes, hipStream_t s);

}  // namespace ggp

// Internal entry points of the onesweep sort used by other translation units.
#pragma once

#include "vx_common.hpp"

namespace vx {
namespace sort {

size_t workspace_bytes(uint32_t n, uint32_t key_bytes, uint32_t key_bits, bool vals);
vx_status sort_u32(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* ws, size_t ws_size,
                   hipStream_t s);
vx_status sort_u64(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* ws, size_t ws_size,
                   hipStream_t s);

// Digit width, keys per thread and workgroup of a onesweep sort. The
// device-count drivers below take the tuning explicitly: a caller that sizes
// a workspace once and sorts into it later (a persistent executor) keeps the
// tuning it sized with, whatever later calls to set_tuning change.
struct Tuning {
  uint32_t rb, kpt, block;
};
uint32_t passes(uint32_t key_bits, const Tuning& tn);

// (u64 key, u32 value) pairs whose count lives in device memory (*n_ptr <=
// n_upper). The caller provides the second buffer pair; *in_alt reports
// whether the sorted result is in (alt_keys, alt_vals).
Tuning pairs_u64_dev_tuning();
size_t pairs_u64_dev_workspace_bytes(uint64_t n_upper, uint32_t key_bits, const Tuning& tn);
vx_status pairs_u64_dev(uint64_t* keys, uint32_t* vals, uint64_t* alt_keys, uint32_t* alt_vals, const uint32_t* n_ptr,
                        uint64_t n_upper, uint32_t key_bits, void* ws, const Tuning& tn, bool* in_alt, hipStream_t s);

// u64 keys only, sorted (stably) by bits [begin_bit, begin_bit + key_bits),
// count in device memory: for keys that pack a sort key above a record index.
Tuning keys_u64_dev_tuning(uint32_t key_bits);
size_t keys_u64_dev_workspace_bytes(uint64_t n_upper, uint32_t key_bits, const Tuning& tn);
vx_status keys_u64_dev(uint64_t* keys, uint64_t* alt_keys, const uint32_t* n_ptr, uint64_t n_upper, uint32_t begin_bit,
                       uint32_t key_bits, void* ws, const Tuning& tn, bool* in_alt, hipStream_t s);

vx_status set_tuning(uint32_t key_bytes, uint32_t rb, uint32_t kpt, uint32_t block);
vx_status get_tuning(uint32_t key_bytes, uint32_t* rb, uint32_t* kpt, uint32_t* block);

}  // namespace sort
}  // namespace vx

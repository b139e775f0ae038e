// tfmm_internal.h -- host-side declarations shared by the TapasFMM HIP
// translation units: the device buffer set of one context and the launch
// wrappers each .hip file provides.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace tfmm {

// Growable device buffer (capacity only grows; no frees inside an evaluation).
struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
};

hipError_t dev_reserve(DevBuf &b, size_t bytes);
void dev_release(DevBuf &b);
// accounting of the device bytes the library holds, against the limit set by
// TFMM_TUNE_DEVICE_LIMIT (0: none): whether releasing `release` bytes and
// allocating `want` stays within it, and the bookkeeping of a change
bool dev_fits(size_t release, size_t want);
void dev_account(size_t freed, size_t allocated);
void dev_set_limit(size_t bytes);
size_t dev_limit();

template <class T>
inline T *dp(DevBuf &b) { return reinterpret_cast<T *>(b.ptr); }

// ---------------------------------------------------------------------------
// scan / sort primitives (sort.hip)
// ---------------------------------------------------------------------------
// exclusive prefix sum of n uint32 values; tmp needs scan_tmp_bytes(n)
size_t scan_tmp_bytes(int64_t n);
hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, void *tmp,
                              hipStream_t s);

// stable LSD radix sort of (key, value) pairs on key bits [bit_lo, bit_hi);
// with dn, the number of pairs is *dn on the device and n its capacity
size_t radix_tmp_bytes(int64_t n);
hipError_t radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                            uint32_t *vals_alt, int64_t n, int bit_lo, int bit_hi, void *tmp,
                            hipStream_t s, bool *result_in_alt, const uint32_t *dn = nullptr);
hipError_t radix_sort_pairs32(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt,
                              uint32_t *vals_alt, int64_t n, int bit_lo, int bit_hi, void *tmp,
                              hipStream_t s, bool *result_in_alt, const uint32_t *dn = nullptr);

}  // namespace tfmm

struct tfmm_picker;
namespace tfmm {
// device weights of a picker (null for a null picker)
const float *picker_weights(tfmm_picker *p);
float picker_theta_max(const tfmm_picker *p);
}  // namespace tfmm

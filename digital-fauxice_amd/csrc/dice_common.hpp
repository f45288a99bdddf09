// dice_common.hpp -- host-side infrastructure shared by the C ABI
// (dice_host.cpp) and the batch / serving runtime (dice_runtime.cpp):
// error reporting, HIP buffers, the hash-pinned tables, SHA-256.
#pragma once

#include <hip/hip_runtime.h>
#include <openssl/evp.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/dice_hip.h"

namespace dice {

// Thread-local message of the last failing call (dice_last_error).
char* error_buffer();

inline int32_t set_error(int32_t code, const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(error_buffer(), 512, fmt, args);
  va_end(args);
  return code;
}

#define HIP_TRY(expr)                                                         \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return ::dice::set_error(DICE_INTERNAL, "%s failed: %s", #expr,         \
                               hipGetErrorString(_e));                        \
  } while (0)

#define DICE_TRY(expr)                  \
  do {                                  \
    int32_t _rc = (expr);               \
    if (_rc != DICE_OK) return _rc;     \
  } while (0)

// Waits for an event without holding a core.  hipEventSynchronize and
// hipStreamSynchronize keep the calling thread running for the whole wait,
// blocking-sync events included (measured on MI355X under ROCm 7: a 200 ms
// wait costs 200 ms of CPU), which takes cores from the writer chains and
// the hashes.  Here the first 200 us poll and yield (a sleep lasts at least
// one timer tick, too long for the streaming path's sub-millisecond block
// waits), then the thread sleeps 50 us between polls, about 1 % of a long
// wait in CPU.
inline hipError_t wait_event(hipEvent_t event) {
  const auto start = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipEventQuery(event);
    if (e != hipErrorNotReady) return e;
    if (std::chrono::steady_clock::now() - start < std::chrono::microseconds(200))
      std::this_thread::yield();
    else
      std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

inline double now_seconds() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
  }
  // Grows to at least `need` bytes (contents are not kept).  With a stream,
  // the old block is released and the new one taken in stream order from the
  // device's memory pool, so growing one job's workspace does not wait for
  // the work of other streams.
  hipError_t ensure(size_t need, hipStream_t stream = nullptr) {
    if (need <= bytes) return hipSuccess;
    const size_t grown = bytes ? need + need / 4 : need;
    const size_t rounded = (grown + 0xFFFFF) & ~(size_t)0xFFFFF;
    if (ptr) {
      if (stream)
        (void)hipFreeAsync(ptr, stream);
      else
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
    hipError_t e = stream ? hipMallocAsync(&ptr, rounded, stream) : hipMalloc(&ptr, rounded);
    if (e == hipSuccess) bytes = rounded;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

struct PinnedBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  ~PinnedBuffer() {
    if (ptr) (void)hipHostFree(ptr);
  }
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    const size_t grown = bytes ? need + need / 4 : need;
    const size_t rounded = (grown + 0xFFFF) & ~(size_t)0xFFFF;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&ptr, rounded, hipHostMallocDefault);
    if (e == hipSuccess) bytes = rounded;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// Incremental SHA-256 (OpenSSL EVP) with lowercase hex output.
class Sha256 {
 public:
  Sha256() : md_(EVP_MD_CTX_new()) {
    ok_ = md_ && EVP_DigestInit_ex(md_, EVP_sha256(), nullptr) == 1;
  }
  ~Sha256() {
    if (md_) EVP_MD_CTX_free(md_);
  }
  Sha256(const Sha256&) = delete;
  Sha256& operator=(const Sha256&) = delete;
  void update(const void* data, size_t bytes) {
    if (ok_ && bytes) ok_ = EVP_DigestUpdate(md_, data, bytes) == 1;
  }
  bool hex(char out[72]) {
    unsigned char digest[32];
    unsigned int length = 0;
    if (!ok_ || EVP_DigestFinal_ex(md_, digest, &length) != 1) return false;
    static const char* digits = "0123456789abcdef";
    for (int i = 0; i < 32; ++i) {
      out[2 * i] = digits[digest[i] >> 4];
      out[2 * i + 1] = digits[digest[i] & 15];
    }
    out[64] = 0;
    return true;
  }

 private:
  EVP_MD_CTX* md_;
  bool ok_;
};

inline bool sha256_hex(const void* data, size_t bytes, char out[72]) {
  Sha256 sha;
  sha.update(data, bytes);
  return sha.hex(out);
}

// The response LUT, the inverse factor tables and the no-op emit table
// (emit of LUT[i]), generated once and pinned by SHA-256.  The no-op emit
// stays within a few units of the raw value, so it also travels as 4-bit
// deltas, eight to a word (noop_delta, kNoopDeltaBias in dice_device.hpp).
struct Tables {
  std::vector<float> lut;
  std::vector<uint32_t> high, low;
  std::vector<uint16_t> noop_emit;
  std::vector<uint32_t> noop_delta;  // 8192 words; nibble r & 7 of word r >> 3
  int32_t status = -1;
};
const Tables& tables();
int32_t require_tables();

struct Callbacks {
  const dice_callbacks* cb = nullptr;
  void progress(int32_t phase, int64_t completed, int64_t total, int64_t attempted,
                int64_t written) const {
    if (cb && cb->progress) cb->progress(cb->user, phase, completed, total, attempted, written);
  }
  bool cancelled() const { return cb && cb->cancelled && cb->cancelled(cb->user) != 0; }
};

inline int32_t cancelled_error() {
  return set_error(DICE_CANCELLED, "portable correction was cancelled");
}

}  // namespace dice

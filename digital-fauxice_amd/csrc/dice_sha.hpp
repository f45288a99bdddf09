// SHA-256 of the replay outputs: one message per job, fed band by band as
// the output rows reach pinned host memory (the reference keeps one
// hashlib.sha256() per replay and updates it row by row,
// src/portable_digital_ice/streaming.py:425, :550).
//
// SHA-256 is serial within a message: each 64-byte block's compression
// needs the previous one's result.  On the x86 SHA extensions one
// compression is a chain of sha256rnds2 instructions whose latency, not
// their issue rate, bounds a single message, so one core compresses two
// independent messages interleaved for much less than twice the cost of
// one (update_interleaved).
#pragma once

#include <cstddef>
#include <cstdint>

namespace dice {

// True when the SHA extensions are present and not switched off with
// DICE_SHA_EXTENSIONS=0 (read once, at first use).  Without them every
// message goes through OpenSSL on its own.
bool sha_extensions();

class Sha256Stream {
 public:
  Sha256Stream();
  ~Sha256Stream();
  Sha256Stream(const Sha256Stream&) = delete;
  Sha256Stream& operator=(const Sha256Stream&) = delete;

  void update(const void* data, size_t bytes);
  // The next piece of each of `count` (1..4) distinct messages, their whole
  // blocks compressed interleaved.
  static void update_interleaved(Sha256Stream* const* streams, const uint8_t* const* data,
                                 const size_t* bytes, int count);
  // Pads, finishes and writes the lowercase hex digest; nothing may be fed
  // after it.
  bool hex(char out[72]);

 private:
  // The eight working words in the order the SHA extensions keep them in
  // two registers: {F, E, B, A} and {H, G, D, C}.
  alignas(16) uint32_t state_[8];
  uint8_t pending_[64];
  uint32_t pending_bytes_ = 0;
  uint64_t total_bytes_ = 0;
  void* openssl_ = nullptr;  // EVP_MD_CTX when the extensions are not used
  bool ok_ = true;
};

}  // namespace dice

// SHA-256 (FIPS 180-4) with the x86 SHA extensions, several messages
// interleaved per core.  Host code only: built by the host compiler,
// without HIP.
#include "dice_sha.hpp"

#include <cpuid.h>
#include <immintrin.h>
#include <openssl/evp.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace dice {

namespace {

// FIPS 180-4 section 4.2.2: the first 32 bits of the fractional parts of the
// cube roots of the first 64 primes.
alignas(16) const uint32_t kRound[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4,
    0xab1c5ed5, 0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe,
    0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f,
    0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7,
    0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc,
    0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b,
    0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116,
    0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
    0xc67178f2};

// Section 5.3.3's initial hash value A..H, laid out {F, E, B, A, H, G, D, C}.
const uint32_t kInitial[8] = {0x9b05688c, 0x510e527f, 0xbb67ae85, 0x6a09e667,
                              0x5be0cd19, 0x1f83d9ab, 0xa54ff53a, 0x3c6ef372};

// `blocks` consecutive 64-byte blocks of each of K messages.  The message
// schedule runs four words at a time: words 0..15 are the block's
// big-endian words, and every later group g is
//   sha256msg2(sha256msg1(W[g-4], W[g-3]) + W[g-1..g-2] shifted by one word, W[g-1]),
// i.e. W[t] = s1(W[t-2]) + W[t-7] + s0(W[t-15]) + W[t-16].  Each group's
// four rounds are two sha256rnds2, the second taking the upper two
// schedule words.  The K messages' instructions are independent, so the
// core overlaps their sha256rnds2 chains.
template <int K>
__attribute__((target("sha,sse4.1"))) void compress(uint32_t* const* states,
                                                    const uint8_t* const* data, size_t blocks) {
  const __m128i byteswap = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
  __m128i abef[K], cdgh[K];
  for (int m = 0; m < K; ++m) {
    abef[m] = _mm_load_si128(reinterpret_cast<const __m128i*>(states[m]));
    cdgh[m] = _mm_load_si128(reinterpret_cast<const __m128i*>(states[m] + 4));
  }
  for (size_t block = 0; block < blocks; ++block) {
    __m128i x[K], y[K], w[K][4];
    for (int m = 0; m < K; ++m) {
      x[m] = abef[m];
      y[m] = cdgh[m];
    }
#pragma GCC unroll 16
    for (int g = 0; g < 16; ++g) {
      const __m128i k = _mm_load_si128(reinterpret_cast<const __m128i*>(kRound + 4 * g));
#pragma GCC unroll 4
      for (int m = 0; m < K; ++m) {
        __m128i words;
        if (g < 4) {
          words = _mm_shuffle_epi8(
              _mm_loadu_si128(reinterpret_cast<const __m128i*>(data[m] + 64 * block + 16 * g)),
              byteswap);
        } else {
          const __m128i low = _mm_sha256msg1_epu32(w[m][g & 3], w[m][(g + 1) & 3]);
          const __m128i mid = _mm_alignr_epi8(w[m][(g + 3) & 3], w[m][(g + 2) & 3], 4);
          words = _mm_sha256msg2_epu32(_mm_add_epi32(low, mid), w[m][(g + 3) & 3]);
        }
        w[m][g & 3] = words;
        const __m128i wk = _mm_add_epi32(words, k);
        y[m] = _mm_sha256rnds2_epu32(y[m], x[m], wk);
        x[m] = _mm_sha256rnds2_epu32(x[m], y[m], _mm_shuffle_epi32(wk, 0x0E));
      }
    }
    for (int m = 0; m < K; ++m) {
      abef[m] = _mm_add_epi32(abef[m], x[m]);
      cdgh[m] = _mm_add_epi32(cdgh[m], y[m]);
    }
  }
  for (int m = 0; m < K; ++m) {
    _mm_store_si128(reinterpret_cast<__m128i*>(states[m]), abef[m]);
    _mm_store_si128(reinterpret_cast<__m128i*>(states[m] + 4), cdgh[m]);
  }
}

void compress_any(int count, uint32_t* const* states, const uint8_t* const* data, size_t blocks) {
  switch (count) {
    case 1: compress<1>(states, data, blocks); break;
    case 2: compress<2>(states, data, blocks); break;
    case 3: compress<3>(states, data, blocks); break;
    default: compress<4>(states, data, blocks); break;
  }
}

bool detect_sha_extensions() {
  const char* env = std::getenv("DICE_SHA_EXTENSIONS");
  if (env && env[0] == '0') return false;
  unsigned a = 0, b = 0, c = 0, d = 0;
  if (!__get_cpuid(1, &a, &b, &c, &d)) return false;
  const bool ssse3 = (c >> 9) & 1u, sse41 = (c >> 19) & 1u;
  if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return false;
  const bool sha = (b >> 29) & 1u;
  return ssse3 && sse41 && sha;
}

}  // namespace

bool sha_extensions() {
  static const bool on = detect_sha_extensions();
  return on;
}

Sha256Stream::Sha256Stream() {
  std::memcpy(state_, kInitial, sizeof state_);
  if (!sha_extensions()) {
    EVP_MD_CTX* md = EVP_MD_CTX_new();
    ok_ = md && EVP_DigestInit_ex(md, EVP_sha256(), nullptr) == 1;
    openssl_ = md;
  }
}

Sha256Stream::~Sha256Stream() {
  if (openssl_) EVP_MD_CTX_free(static_cast<EVP_MD_CTX*>(openssl_));
}

void Sha256Stream::update(const void* data, size_t bytes) {
  Sha256Stream* self = this;
  const uint8_t* p = static_cast<const uint8_t*>(data);
  update_interleaved(&self, &p, &bytes, 1);
}

void Sha256Stream::update_interleaved(Sha256Stream* const* streams, const uint8_t* const* data,
                                      const size_t* bytes, int count) {
  if (!sha_extensions()) {
    for (int i = 0; i < count; ++i) {
      Sha256Stream& s = *streams[i];
      if (s.ok_ && bytes[i])
        s.ok_ = EVP_DigestUpdate(static_cast<EVP_MD_CTX*>(s.openssl_), data[i], bytes[i]) == 1;
    }
    return;
  }
  const uint8_t* at[4];
  size_t left[4];
  for (int i = 0; i < count; ++i) {
    Sha256Stream& s = *streams[i];
    at[i] = data[i];
    left[i] = bytes[i];
    s.total_bytes_ += bytes[i];
    // complete a block begun by an earlier piece
    if (s.pending_bytes_ > 0 && left[i] > 0) {
      const size_t take = std::min<size_t>(64 - s.pending_bytes_, left[i]);
      std::memcpy(s.pending_ + s.pending_bytes_, at[i], take);
      s.pending_bytes_ += (uint32_t)take;
      at[i] += take;
      left[i] -= take;
      if (s.pending_bytes_ == 64) {
        uint32_t* state = s.state_;
        const uint8_t* block = s.pending_;
        compress<1>(&state, &block, 1);
        s.pending_bytes_ = 0;
      }
    }
  }
  // whole blocks: every message with blocks left, interleaved, for as many
  // blocks as the shortest of them has; then again over those still left
  for (;;) {
    uint32_t* states[4];
    const uint8_t* ptrs[4];
    int index[4], active = 0;
    size_t blocks = 0;
    for (int i = 0; i < count; ++i) {
      if (left[i] < 64) continue;
      const size_t have = left[i] / 64;
      blocks = active == 0 ? have : std::min(blocks, have);
      states[active] = streams[i]->state_;
      ptrs[active] = at[i];
      index[active++] = i;
    }
    if (active == 0) break;
    compress_any(active, states, ptrs, blocks);
    for (int a = 0; a < active; ++a) {
      at[index[a]] += 64 * blocks;
      left[index[a]] -= 64 * blocks;
    }
  }
  for (int i = 0; i < count; ++i) {
    if (left[i] == 0) continue;  // (a partial block implies pending_bytes_ == 0 here)
    Sha256Stream& s = *streams[i];
    std::memcpy(s.pending_, at[i], left[i]);
    s.pending_bytes_ = (uint32_t)left[i];
  }
}

bool Sha256Stream::hex(char out[72]) {
  static const char* digits = "0123456789abcdef";
  uint8_t digest[32];
  if (!sha_extensions()) {
    unsigned int length = 0;
    if (!ok_ || EVP_DigestFinal_ex(static_cast<EVP_MD_CTX*>(openssl_), digest, &length) != 1)
      return false;
  } else {
    // section 5.1.1: 0x80, zeros to 56 mod 64, the length in bits (big-endian)
    uint8_t tail[128] = {};
    std::memcpy(tail, pending_, pending_bytes_);
    tail[pending_bytes_] = 0x80;
    const size_t tail_bytes = pending_bytes_ + 9 <= 64 ? 64 : 128;
    const uint64_t bits = total_bytes_ * 8;
    for (int i = 0; i < 8; ++i) tail[tail_bytes - 1 - i] = (uint8_t)(bits >> (8 * i));
    uint32_t* state = state_;
    const uint8_t* block = tail;
    compress<1>(&state, &block, tail_bytes / 64);
    // A..H from {F, E, B, A, H, G, D, C}
    const uint32_t words[8] = {state_[3], state_[2], state_[7], state_[6],
                               state_[1], state_[0], state_[5], state_[4]};
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 4; ++j) digest[4 * i + j] = (uint8_t)(words[i] >> (24 - 8 * j));
  }
  for (int i = 0; i < 32; ++i) {
    out[2 * i] = digits[digest[i] >> 4];
    out[2 * i + 1] = digits[digest[i] & 15];
  }
  out[64] = 0;
  return true;
}

}  // namespace dice

// dice_routing.hip -- at-floor region routing on the device
// (hybrid/src/fauxce_hybrid/routing.py:447-687).
//
// Connected components use a lock-free union-find over the pixels:
//   * every set pixel starts as its own root;
//   * each pixel unions with its four backward 8-neighbours, always linking
//     the larger root under the smaller;
//   * so every component ends rooted at its smallest linear index, the first
//     pixel of the component in raster order.
// Ranking the roots in index order then reproduces the raster-order
// numbering of a sequential labeller.  Distances, areas, boxes and counts are
// integer atomics, so the whole routing is exact and order-free.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <tuple>
#include <vector>

#include "dice_routing.hpp"

#pragma clang fp contract(off)

namespace dice {
namespace {

constexpr int kRT = 256;
constexpr int32_t kFar = INT_MAX / 2;
constexpr int64_t kCropMargin = 128;
constexpr int64_t kMaxCropPixels = 4194304;
constexpr int64_t kTileEdge = 2048 - 2 * kCropMargin;  // isqrt(kMaxCropPixels) - 2 margins

unsigned grid_for(int64_t n) {
  const int64_t blocks = (n + kRT - 1) / kRT;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks, 1 << 20));
}

#define RT_LOOP(i, n)                                                            \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n);    \
       i += (int64_t)gridDim.x * blockDim.x)

__device__ __forceinline__ int32_t load_relaxed(const int32_t* p) {
  return __atomic_load_n(p, __ATOMIC_RELAXED);
}

__device__ __forceinline__ int32_t uf_find(const int32_t* parent, int32_t x) {
  for (;;) {
    const int32_t p = load_relaxed(parent + x);
    if (p == x) return x;
    x = p;
  }
}

// Links the larger of the two roots under the smaller; a link another
// thread made first is followed and the union retried.
__device__ __forceinline__ void uf_union(int32_t* parent, int32_t a, int32_t b) {
  for (;;) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a > b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    const int32_t old = atomicCAS(parent + b, b, a);
    if (old == b) return;
    b = old;
  }
}

// Restriction of the second (fragment) pass: pixels of a component marked
// `cut` only join within the same kTileEdge grid tile counted from the
// component's box origin.
struct Cut {
  const int32_t* raw = nullptr;    // component id per pixel (1-based)
  const uint8_t* cut = nullptr;    // per component id
  const int32_t* origin = nullptr; // per component id: y0, x0
};

__device__ __forceinline__ bool same_tile(const Cut& c, int32_t id, int64_t ya, int64_t xa,
                                          int64_t yb, int64_t xb) {
  const int64_t y0 = c.origin[2 * id], x0 = c.origin[2 * id + 1];
  return (ya - y0) / kTileEdge == (yb - y0) / kTileEdge &&
         (xa - x0) / kTileEdge == (xb - x0) / kTileEdge;
}

__global__ __launch_bounds__(kRT) void k_uf_init(const uint8_t* mask, int32_t* parent, int64_t n,
                                                  Cut c) {
  RT_LOOP(i, n) {
    if (c.cut) {  // fragment pass: only the pixels of cut components start over
      if (mask[i] && c.cut[c.raw[i]]) parent[i] = (int32_t)i;
    } else {
      parent[i] = mask[i] ? (int32_t)i : -1;
    }
  }
}

__global__ __launch_bounds__(kRT) void k_uf_merge(const uint8_t* mask, int32_t* parent,
                                                   int64_t height, int64_t width, Cut c) {
  const int64_t n = height * width;
  RT_LOOP(i, n) {
    if (!mask[i]) continue;
    int32_t id = 0;
    if (c.cut) {
      id = c.raw[i];
      if (!c.cut[id]) continue;
    }
    const int64_t y = i / width, x = i % width;
    const int64_t ny[4] = {y, y - 1, y - 1, y - 1};
    const int64_t nx[4] = {x - 1, x - 1, x, x + 1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (ny[k] < 0 || nx[k] < 0 || nx[k] >= width) continue;
      const int64_t j = ny[k] * width + nx[k];
      if (!mask[j]) continue;
      if (c.cut && !same_tile(c, id, y, x, ny[k], nx[k])) continue;
      uf_union(parent, (int32_t)i, (int32_t)j);
    }
  }
}

__global__ __launch_bounds__(kRT) void k_uf_flatten(const uint8_t* mask, int32_t* parent,
                                                     int64_t n) {
  RT_LOOP(i, n) {
    if (mask[i]) parent[i] = uf_find(parent, (int32_t)i);
  }
}

__global__ __launch_bounds__(kRT) void k_uf_roots(const uint8_t* mask, const int32_t* parent,
                                                   int64_t n, int32_t* list, uint32_t* count) {
  RT_LOOP(i, n) {
    if (mask[i] && parent[i] == (int32_t)i) list[atomicAdd(count, 1u)] = (int32_t)i;
  }
}

__global__ __launch_bounds__(kRT) void k_rank_roots(const int32_t* sorted, int64_t m,
                                                     int32_t* rank_at) {
  RT_LOOP(j, m) rank_at[sorted[j]] = (int32_t)(j + 1);
}

__global__ __launch_bounds__(kRT) void k_label(const uint8_t* mask, const int32_t* parent,
                                                const int32_t* rank_at, int64_t n,
                                                int32_t* label) {
  RT_LOOP(i, n) label[i] = mask[i] ? rank_at[parent[i]] : 0;
}

// Per label: area, half-open box, and the at-floor pixels (when `floor`).
__global__ __launch_bounds__(kRT) void k_stats(const int32_t* label, const uint8_t* floor,
                                                int64_t height, int64_t width,
                                                unsigned long long* area, int32_t* box,
                                                unsigned long long* floor_count) {
  const int64_t n = height * width;
  RT_LOOP(i, n) {
    const int32_t id = label[i];
    if (id <= 0) continue;
    const int32_t y = (int32_t)(i / width), x = (int32_t)(i % width);
    atomicAdd(area + id, 1ull);
    atomicMin(box + 4 * id + 0, y);
    atomicMin(box + 4 * id + 1, x);
    atomicMax(box + 4 * id + 2, y + 1);
    atomicMax(box + 4 * id + 3, x + 1);
    if (floor && floor[i]) atomicAdd(floor_count + id, 1ull);
  }
}

__global__ __launch_bounds__(kRT) void k_box_init(int32_t* box, int64_t ids) {
  RT_LOOP(i, ids) {
    box[4 * i + 0] = INT_MAX;
    box[4 * i + 1] = INT_MAX;
    box[4 * i + 2] = 0;
    box[4 * i + 3] = 0;
  }
}

// Horizontal chessboard reach: distance along the row to the nearest pixel
// off the floor (kFar when the row has none).
__global__ __launch_bounds__(kRT) void k_row_reach(const uint8_t* mask, int64_t height,
                                                    int64_t width, int32_t* reach) {
  RT_LOOP(y, height) {
    const uint8_t* row = mask + y * width;
    int32_t* out = reach + y * width;
    int64_t last = -1;
    for (int64_t x = 0; x < width; ++x) {
      if (!row[x]) last = x;
      out[x] = last < 0 ? kFar : (int32_t)(x - last);
    }
    last = -1;
    for (int64_t x = width - 1; x >= 0; --x) {
      if (!row[x]) last = x;
      if (last >= 0) out[x] = min(out[x], (int32_t)(last - x));
    }
  }
}

// Chessboard distance of every at-floor pixel to the nearest in-frame pixel
// off the floor: min over rows y' of max(|y - y'|, reach(y', x)), searched
// outwards until |y - y'| reaches the best so far; the largest per region.
__global__ __launch_bounds__(kRT) void k_radius(const int32_t* label, const int32_t* reach,
                                                 int64_t height, int64_t width,
                                                 unsigned long long* radius) {
  const int64_t n = height * width;
  RT_LOOP(i, n) {
    const int32_t id = label[i];
    if (id <= 0) continue;
    const int64_t y = i / width, x = i % width;
    int64_t best = reach[i];
    for (int64_t k = 1; k < best; ++k) {
      const bool up = y - k >= 0, down = y + k < height;
      if (!up && !down) break;
      if (up) best = min(best, max(k, (int64_t)reach[(y - k) * width + x]));
      if (down) best = min(best, max(k, (int64_t)reach[(y + k) * width + x]));
    }
    atomicMax(radius + id, (unsigned long long)best);
  }
}

__global__ __launch_bounds__(kRT) void k_perimeter(const int32_t* label, int64_t height,
                                                    int64_t width, int32_t corner,
                                                    unsigned long long* hits) {
  // perimeter pixels: the first and last rows, then the first and last
  // columns of the rows between (each pixel once)
  const int64_t inner = height > 2 ? height - 2 : 0;
  const int64_t rows = height > 1 ? 2 : 1, cols = width > 1 ? 2 : 1;
  const int64_t n = rows * width + cols * inner;
  RT_LOOP(t, n) {
    int64_t y, x;
    if (t < rows * width) {
      y = t < width ? 0 : height - 1;
      x = t % width;
    } else {
      const int64_t u = t - rows * width;
      y = 1 + u / cols;
      x = (u % cols) ? width - 1 : 0;
    }
    if (label[y * width + x] == corner) atomicAdd(hits, 1ull);
  }
}

__global__ __launch_bounds__(kRT) void k_routed_union(const int32_t* label, const uint8_t* routed,
                                                       int64_t n, uint8_t* out) {
  RT_LOOP(i, n) out[i] = label[i] > 0 && routed[label[i]] ? 1 : 0;
}

// Square dilation by separable windows: running counts along a column (or
// row), then "any set within r" from two counts.
__global__ __launch_bounds__(kRT) void k_column_counts(const uint8_t* in, int64_t height,
                                                        int64_t width, int32_t* count) {
  RT_LOOP(x, width) {
    int32_t c = 0;
    for (int64_t y = 0; y < height; ++y) {
      c += in[y * width + x] ? 1 : 0;
      count[y * width + x] = c;
    }
  }
}

__global__ __launch_bounds__(kRT) void k_row_counts(const uint8_t* in, int64_t height,
                                                     int64_t width, int32_t* count) {
  RT_LOOP(y, height) {
    int32_t c = 0;
    for (int64_t x = 0; x < width; ++x) {
      c += in[y * width + x] ? 1 : 0;
      count[y * width + x] = c;
    }
  }
}

__global__ __launch_bounds__(kRT) void k_window_any(const int32_t* count, int64_t height,
                                                    int64_t width, int64_t radius, bool vertical,
                                                    uint8_t* out) {
  const int64_t n = height * width;
  RT_LOOP(i, n) {
    const int64_t y = i / width, x = i % width;
    int32_t hi, lo;
    if (vertical) {
      hi = count[min(y + radius, height - 1) * width + x];
      lo = y - radius - 1 >= 0 ? count[(y - radius - 1) * width + x] : 0;
    } else {
      hi = count[y * width + min(x + radius, width - 1)];
      lo = x - radius - 1 >= 0 ? count[y * width + (x - radius - 1)] : 0;
    }
    out[i] = hi > lo ? 1 : 0;
  }
}

// Removes the perimeter component from the synthesis mask (counting the
// halo it had gained), and counts the at-floor, synthesis and non-floor
// synthesis pixels.
__global__ __launch_bounds__(kRT) void k_finish_mask(const uint8_t* floor, const int32_t* label,
                                                      int32_t excluded, int64_t n, uint8_t* synth,
                                                      unsigned long long* counts) {
  unsigned long long at_floor = 0, synthesized = 0, non_floor = 0, halo = 0;
  RT_LOOP(i, n) {
    uint8_t s = synth[i];
    if (excluded > 0 && label[i] == excluded && s) {
      ++halo;
      s = 0;
      synth[i] = 0;
    }
    at_floor += floor[i] ? 1 : 0;
    synthesized += s;
    non_floor += (s && !floor[i]) ? 1 : 0;
  }
  atomicAdd(counts + 0, at_floor);
  atomicAdd(counts + 1, synthesized);
  atomicAdd(counts + 2, non_floor);
  atomicAdd(counts + 3, halo);
}

// Per fragment (temporary id): its raw component and tile, for the order
// of creation.
__global__ __launch_bounds__(kRT) void k_fragment_keys(const int32_t* sorted_roots, int64_t m,
                                                        const int32_t* raw, const int32_t* origin,
                                                        const uint8_t* cut, int64_t width,
                                                        int32_t* keys) {
  RT_LOOP(j, m) {
    const int64_t i = sorted_roots[j];
    const int32_t id = raw[i];
    int32_t ty = 0, tx = 0;
    if (cut[id]) {
      ty = (int32_t)((i / width - origin[2 * id]) / kTileEdge);
      tx = (int32_t)((i % width - origin[2 * id + 1]) / kTileEdge);
    }
    keys[3 * j + 0] = id;
    keys[3 * j + 1] = ty;
    keys[3 * j + 2] = tx;
  }
}

__global__ __launch_bounds__(kRT) void k_count_clear(const uint8_t* mask, int64_t n,
                                                      unsigned long long* clear) {
  unsigned long long c = 0;
  RT_LOOP(i, n) c += mask[i] ? 0 : 1;
  if (c) atomicAdd(clear, c);
}

__global__ __launch_bounds__(kRT) void k_relabel(int32_t* label, const int32_t* map, int64_t n) {
  RT_LOOP(i, n) {
    if (label[i] > 0) label[i] = map[label[i]];
  }
}

// Distinct (provisional, final) pairs over the at-floor pixels of the
// synthesis mask, in an open-addressing set of 64-bit keys; the overlap
// counts per provisional region.
__global__ __launch_bounds__(kRT) void k_pairs(const int32_t* provisional, const int32_t* final_,
                                                int64_t n, unsigned long long* table,
                                                uint32_t capacity, unsigned long long* overlap,
                                                uint32_t* overflow) {
  RT_LOOP(i, n) {
    const int32_t p = provisional[i], f = final_[i];
    if (p <= 0 || f <= 0) continue;
    atomicAdd(overlap + p, 1ull);
    const unsigned long long key = ((unsigned long long)(uint32_t)p << 32) | (uint32_t)f;
    uint32_t slot = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & (capacity - 1);
    for (uint32_t probe = 0;; ++probe) {
      if (probe == capacity) {
        atomicOr(overflow, 1u);
        break;
      }
      const unsigned long long old = atomicCAS(table + slot, 0ull, key);
      if (old == 0ull || old == key) break;
      slot = (slot + 1) & (capacity - 1);
    }
  }
}

template <class T>
struct Dev {
  DeviceBuffer b;
  int32_t alloc(size_t n, hipStream_t s) {
    HIP_TRY(b.ensure(std::max<size_t>(n, 1) * sizeof(T), s));
    return DICE_OK;
  }
  T* p() const { return b.as<T>(); }
};

// Components of `mask`: labels 1..count in raster order of first pixels.
struct Components {
  int64_t count = 0;
  std::vector<int32_t> roots;  // sorted root (first pixel) indices
};

int32_t label_components(hipStream_t s, const uint8_t* mask, int64_t height, int64_t width,
                         int32_t* parent, int32_t* rank_at, int32_t* label, Dev<int32_t>& list,
                         Dev<uint32_t>& counter, Components& out, const Cut* fragments) {
  const int64_t n = height * width;
  const unsigned g = grid_for(n);
  const Cut none;
  const Cut& c = fragments ? *fragments : none;
  hipLaunchKernelGGL(k_uf_init, dim3(g), dim3(kRT), 0, s, mask, parent, n, c);
  hipLaunchKernelGGL(k_uf_merge, dim3(g), dim3(kRT), 0, s, mask, parent, height, width, c);
  hipLaunchKernelGGL(k_uf_flatten, dim3(g), dim3(kRT), 0, s, mask, parent, n);
  HIP_TRY(hipMemsetAsync(counter.p(), 0, 4, s));
  hipLaunchKernelGGL(k_uf_roots, dim3(g), dim3(kRT), 0, s, mask, parent, n, list.p(), counter.p());
  HIP_TRY(hipGetLastError());
  uint32_t m = 0;
  HIP_TRY(hipMemcpyAsync(&m, counter.p(), 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  out.count = m;
  out.roots.resize(m);
  if (m) {
    HIP_TRY(hipMemcpyAsync(out.roots.data(), list.p(), (size_t)m * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::sort(out.roots.begin(), out.roots.end());
    HIP_TRY(hipMemcpyAsync(list.p(), out.roots.data(), (size_t)m * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_rank_roots, dim3(grid_for(m)), dim3(kRT), 0, s, list.p(), (int64_t)m,
                       rank_at);
  }
  hipLaunchKernelGGL(k_label, dim3(g), dim3(kRT), 0, s, mask, parent, rank_at, n, label);
  HIP_TRY(hipGetLastError());
  return DICE_OK;
}

struct RegionStats {
  std::vector<unsigned long long> area, floor;
  std::vector<int32_t> box;
};

int32_t region_stats(hipStream_t s, const int32_t* label, const uint8_t* floor, int64_t height,
                     int64_t width, int64_t count, RegionStats& out) {
  Dev<unsigned long long> area, floor_count;
  Dev<int32_t> box;
  DICE_TRY(area.alloc(count + 1, s));
  DICE_TRY(floor_count.alloc(count + 1, s));
  DICE_TRY(box.alloc(4 * (count + 1), s));
  HIP_TRY(hipMemsetAsync(area.p(), 0, (count + 1) * 8, s));
  HIP_TRY(hipMemsetAsync(floor_count.p(), 0, (count + 1) * 8, s));
  hipLaunchKernelGGL(k_box_init, dim3(grid_for(count + 1)), dim3(kRT), 0, s, box.p(), count + 1);
  hipLaunchKernelGGL(k_stats, dim3(grid_for(height * width)), dim3(kRT), 0, s, label, floor,
                     height, width, area.p(), box.p(), floor_count.p());
  HIP_TRY(hipGetLastError());
  out.area.resize(count + 1);
  out.floor.resize(count + 1);
  out.box.resize(4 * (count + 1));
  HIP_TRY(hipMemcpyAsync(out.area.data(), area.p(), (count + 1) * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(out.floor.data(), floor_count.p(), (count + 1) * 8,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(out.box.data(), box.p(), 4 * (count + 1) * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return DICE_OK;
}

}  // namespace

int32_t route_at_floor(hipStream_t s, const uint8_t* d_mask, int64_t height, int64_t width,
                       int64_t min_area, int64_t min_radius, int64_t margin, RoutingOutput& out) {
  const int64_t n = height * width;
  const unsigned g = grid_for(n);
  out.height = height;
  out.width = width;
  {
    Dev<unsigned long long> clear;
    DICE_TRY(clear.alloc(1, s));
    HIP_TRY(hipMemsetAsync(clear.p(), 0, 8, s));
    hipLaunchKernelGGL(k_count_clear, dim3(std::min<unsigned>(g, 4096)), dim3(kRT), 0, s, d_mask,
                       n, clear.p());
    HIP_TRY(hipGetLastError());
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, clear.p(), 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (c == 0)
      return set_error(DICE_INVALID,
                       "at_floor_mask covers the entire frame; no healthy in-frame context "
                       "exists for bounded synthesis");
  }
  Dev<int32_t> parent, rank_at, reach, list, scratch;
  Dev<uint32_t> counter;
  DICE_TRY(parent.alloc(n, s));
  DICE_TRY(rank_at.alloc(n, s));
  DICE_TRY(reach.alloc(n, s));
  DICE_TRY(list.alloc(n, s));
  DICE_TRY(counter.alloc(4, s));
  HIP_TRY(out.provisional_labels.ensure(std::max<int64_t>(n, 1) * 4, s));
  HIP_TRY(out.synthesis_mask.ensure(std::max<int64_t>(n, 1), s));
  HIP_TRY(out.final_labels.ensure(std::max<int64_t>(n, 1) * 4, s));
  int32_t* provisional = out.provisional_labels.as<int32_t>();
  uint8_t* synth = out.synthesis_mask.as<uint8_t>();
  int32_t* final_ = out.final_labels.as<int32_t>();

  // ---- provisional regions: labels, areas, boxes, chessboard radii
  Components comps;
  DICE_TRY(label_components(s, d_mask, height, width, parent.p(), rank_at.p(), provisional, list,
                            counter, comps, nullptr));
  const int64_t C = comps.count;
  RegionStats prov;
  DICE_TRY(region_stats(s, provisional, nullptr, height, width, C, prov));
  Dev<unsigned long long> radius;
  DICE_TRY(radius.alloc(C + 1, s));
  HIP_TRY(hipMemsetAsync(radius.p(), 0, (C + 1) * 8, s));
  hipLaunchKernelGGL(k_row_reach, dim3(grid_for(height)), dim3(kRT), 0, s, d_mask, height, width,
                     reach.p());
  hipLaunchKernelGGL(k_radius, dim3(g), dim3(kRT), 0, s, provisional, reach.p(), height, width,
                     radius.p());
  std::vector<unsigned long long> radii(C + 1);
  HIP_TRY(hipMemcpyAsync(radii.data(), radius.p(), (C + 1) * 8, hipMemcpyDeviceToHost, s));
  int32_t corner = 0;
  HIP_TRY(hipMemcpyAsync(&corner, provisional, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));

  // ---- the component holding the whole perimeter, if any
  int64_t excluded = 0;
  if (corner > 0) {
    Dev<unsigned long long> hits;
    DICE_TRY(hits.alloc(1, s));
    HIP_TRY(hipMemsetAsync(hits.p(), 0, 8, s));
    hipLaunchKernelGGL(k_perimeter, dim3(grid_for(2 * (height + width))), dim3(kRT), 0, s,
                       provisional, height, width, corner, hits.p());
    HIP_TRY(hipGetLastError());
    unsigned long long h = 0;
    HIP_TRY(hipMemcpyAsync(&h, hits.p(), 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const int64_t perimeter = height == 1 ? width : width == 1 ? height : 2 * (height + width) - 4;
    if ((int64_t)h == perimeter) excluded = corner;
  }

  // ---- routed union, square dilation, perimeter removal
  std::vector<uint8_t> routed(C + 1, 0);
  for (int64_t id = 1; id <= C; ++id)
    routed[id] = id != excluded &&
                 ((int64_t)prov.area[id] >= min_area || (int64_t)radii[id] >= min_radius);
  Dev<uint8_t> d_routed, u8;
  DICE_TRY(d_routed.alloc(C + 1, s));
  DICE_TRY(u8.alloc(n, s));
  HIP_TRY(hipMemcpyAsync(d_routed.p(), routed.data(), C + 1, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_routed_union, dim3(g), dim3(kRT), 0, s, provisional, d_routed.p(), n,
                     margin > 0 ? u8.p() : synth);
  if (margin > 0) {
    const int64_t rv = std::min(margin, height - 1), rh = std::min(margin, width - 1);
    DICE_TRY(scratch.alloc(n, s));
    hipLaunchKernelGGL(k_column_counts, dim3(grid_for(width)), dim3(kRT), 0, s, u8.p(), height,
                       width, scratch.p());
    hipLaunchKernelGGL(k_window_any, dim3(g), dim3(kRT), 0, s, scratch.p(), height, width, rv,
                       true, synth);
    hipLaunchKernelGGL(k_row_counts, dim3(grid_for(height)), dim3(kRT), 0, s, synth, height,
                       width, scratch.p());
    hipLaunchKernelGGL(k_window_any, dim3(g), dim3(kRT), 0, s, scratch.p(), height, width, rh,
                       false, synth);
  }
  Dev<unsigned long long> counts;
  DICE_TRY(counts.alloc(4, s));
  HIP_TRY(hipMemsetAsync(counts.p(), 0, 32, s));
  hipLaunchKernelGGL(k_finish_mask, dim3(std::min<unsigned>(g, 4096)), dim3(kRT), 0, s, d_mask,
                     provisional, (int32_t)excluded, n, synth, counts.p());
  HIP_TRY(hipGetLastError());
  unsigned long long mask_counts[4];
  HIP_TRY(hipMemcpyAsync(mask_counts, counts.p(), 32, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));

  // ---- final regions: components of the synthesis mask, oversized ones
  // cut on the tile grid, then numbered by (box, order of creation)
  Components raw;
  DICE_TRY(label_components(s, synth, height, width, parent.p(), rank_at.p(), final_, list,
                            counter, raw, nullptr));
  RegionStats raw_stats;
  DICE_TRY(region_stats(s, final_, nullptr, height, width, raw.count, raw_stats));
  std::vector<uint8_t> cut(raw.count + 1, 0);
  std::vector<int32_t> origin(2 * (raw.count + 1), 0);
  bool any_cut = false;
  for (int64_t id = 1; id <= raw.count; ++id) {
    const int64_t y0 = raw_stats.box[4 * id], x0 = raw_stats.box[4 * id + 1];
    const int64_t y1 = raw_stats.box[4 * id + 2], x1 = raw_stats.box[4 * id + 3];
    const int64_t crop = (std::min(height, y1 + kCropMargin) - std::max<int64_t>(0, y0 - kCropMargin)) *
                         (std::min(width, x1 + kCropMargin) - std::max<int64_t>(0, x0 - kCropMargin));
    origin[2 * id] = (int32_t)y0;
    origin[2 * id + 1] = (int32_t)x0;
    if (crop > kMaxCropPixels) {
      cut[id] = 1;
      any_cut = true;
    }
  }
  Dev<int32_t> d_raw, d_origin, d_keys;
  Dev<uint8_t> d_cut;
  DICE_TRY(d_raw.alloc(n, s));
  DICE_TRY(d_origin.alloc(origin.size(), s));
  DICE_TRY(d_cut.alloc(cut.size(), s));
  HIP_TRY(hipMemcpyAsync(d_raw.p(), final_, n * 4, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(d_origin.p(), origin.data(), origin.size() * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d_cut.p(), cut.data(), cut.size(), hipMemcpyHostToDevice, s));
  Components fragments = raw;
  if (any_cut) {
    Cut c;
    c.raw = d_raw.p();
    c.cut = d_cut.p();
    c.origin = d_origin.p();
    // parent still holds every raw root; only the cut components start over
    const unsigned gn = grid_for(n);
    hipLaunchKernelGGL(k_uf_init, dim3(gn), dim3(kRT), 0, s, synth, parent.p(), n, c);
    hipLaunchKernelGGL(k_uf_merge, dim3(gn), dim3(kRT), 0, s, synth, parent.p(), height, width, c);
    hipLaunchKernelGGL(k_uf_flatten, dim3(gn), dim3(kRT), 0, s, synth, parent.p(), n);
    HIP_TRY(hipMemsetAsync(counter.p(), 0, 4, s));
    hipLaunchKernelGGL(k_uf_roots, dim3(gn), dim3(kRT), 0, s, synth, parent.p(), n, list.p(),
                       counter.p());
    HIP_TRY(hipGetLastError());
    uint32_t m = 0;
    HIP_TRY(hipMemcpyAsync(&m, counter.p(), 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    fragments.count = m;
    fragments.roots.resize(m);
    HIP_TRY(hipMemcpyAsync(fragments.roots.data(), list.p(), (size_t)m * 4, hipMemcpyDeviceToHost,
                           s));
    HIP_TRY(hipStreamSynchronize(s));
    std::sort(fragments.roots.begin(), fragments.roots.end());
    HIP_TRY(hipMemcpyAsync(list.p(), fragments.roots.data(), (size_t)m * 4, hipMemcpyHostToDevice,
                           s));
    hipLaunchKernelGGL(k_rank_roots, dim3(grid_for(m)), dim3(kRT), 0, s, list.p(), (int64_t)m,
                       rank_at.p());
    hipLaunchKernelGGL(k_label, dim3(gn), dim3(kRT), 0, s, synth, parent.p(), rank_at.p(), n,
                       final_);
    HIP_TRY(hipGetLastError());
  }
  const int64_t F = fragments.count;
  RegionStats frag;
  DICE_TRY(region_stats(s, final_, d_mask, height, width, F, frag));
  std::vector<int32_t> keys(3 * (F + 1), 0);
  if (F) {
    DICE_TRY(d_keys.alloc(3 * F, s));
    HIP_TRY(hipMemcpyAsync(list.p(), fragments.roots.data(), (size_t)F * 4, hipMemcpyHostToDevice,
                           s));
    hipLaunchKernelGGL(k_fragment_keys, dim3(grid_for(F)), dim3(kRT), 0, s, list.p(), F, d_raw.p(),
                       d_origin.p(), d_cut.p(), width, d_keys.p());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(keys.data() + 3, d_keys.p(), (size_t)F * 12, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  std::vector<int32_t> order(F);
  for (int64_t f = 0; f < F; ++f) order[f] = (int32_t)(f + 1);
  auto key = [&](int32_t f) {
    return std::make_tuple(frag.box[4 * f], frag.box[4 * f + 1], frag.box[4 * f + 2],
                           frag.box[4 * f + 3], keys[3 * f], keys[3 * f + 1], keys[3 * f + 2],
                           fragments.roots[f - 1]);
  };
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key(a) < key(b); });
  std::vector<int32_t> map(F + 1, 0);
  for (int64_t k = 0; k < F; ++k) map[order[k]] = (int32_t)(k + 1);
  if (F) {
    Dev<int32_t> d_map;
    DICE_TRY(d_map.alloc(F + 1, s));
    HIP_TRY(hipMemcpyAsync(d_map.p(), map.data(), (F + 1) * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_relabel, dim3(g), dim3(kRT), 0, s, final_, d_map.p(), n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
  }

  // ---- provisional/final pairs and overlap counts
  std::vector<unsigned long long> overlap(C + 1, 0);
  std::vector<int32_t> pairs;
  if (C && F) {
    uint32_t capacity = 1024;
    while (capacity < 4 * (uint64_t)(C + F)) capacity <<= 1;
    for (;;) {
      Dev<unsigned long long> table, d_overlap;
      Dev<uint32_t> overflow;
      DICE_TRY(table.alloc(capacity, s));
      DICE_TRY(d_overlap.alloc(C + 1, s));
      DICE_TRY(overflow.alloc(1, s));
      HIP_TRY(hipMemsetAsync(table.p(), 0, (size_t)capacity * 8, s));
      HIP_TRY(hipMemsetAsync(d_overlap.p(), 0, (C + 1) * 8, s));
      HIP_TRY(hipMemsetAsync(overflow.p(), 0, 4, s));
      hipLaunchKernelGGL(k_pairs, dim3(g), dim3(kRT), 0, s, provisional, final_, n, table.p(),
                         capacity, d_overlap.p(), overflow.p());
      HIP_TRY(hipGetLastError());
      uint32_t over = 0;
      HIP_TRY(hipMemcpyAsync(&over, overflow.p(), 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (over && capacity < (1u << 30)) {
        capacity <<= 1;
        continue;
      }
      if (over) return set_error(DICE_INTERNAL, "routing pair table overflow");
      std::vector<unsigned long long> h(capacity);
      HIP_TRY(hipMemcpyAsync(h.data(), table.p(), (size_t)capacity * 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(overlap.data(), d_overlap.p(), (C + 1) * 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      std::vector<unsigned long long> found;
      for (unsigned long long k : h)
        if (k) found.push_back(k);
      std::sort(found.begin(), found.end());
      for (unsigned long long k : found) {
        pairs.push_back((int32_t)(k >> 32));
        pairs.push_back((int32_t)(k & 0xFFFFFFFFu));
      }
      break;
    }
  }

  // ---- tables
  out.provisional.assign(C, dice_provisional_region{});
  for (int64_t id = 1; id <= C; ++id) {
    dice_provisional_region& r = out.provisional[id - 1];
    r.y0 = prov.box[4 * id];
    r.x0 = prov.box[4 * id + 1];
    r.y1 = prov.box[4 * id + 2];
    r.x1 = prov.box[4 * id + 3];
    r.area = (int64_t)prov.area[id];
    r.max_chessboard_radius = (int64_t)radii[id];
    r.synthesized_pixels = (int64_t)overlap[id];
    r.routed = routed[id];
  }
  out.finals.assign(F, dice_final_region{});
  for (int64_t k = 0; k < F; ++k) {
    const int32_t f = order[k];
    dice_final_region& r = out.finals[k];
    r.y0 = frag.box[4 * f];
    r.x0 = frag.box[4 * f + 1];
    r.y1 = frag.box[4 * f + 2];
    r.x1 = frag.box[4 * f + 3];
    r.area = (int64_t)frag.area[f];
    r.at_floor_pixels = (int64_t)frag.floor[f];
  }
  out.pairs = std::move(pairs);
  dice_routing_summary& S = out.summary;
  S.height = height;
  S.width = width;
  S.provisional_regions = C;
  S.final_regions = F;
  S.pairs = (int64_t)out.pairs.size() / 2;
  S.at_floor_pixels = (int64_t)mask_counts[0];
  S.synthesis_pixels = (int64_t)mask_counts[1];
  S.non_floor_synthesis_pixels = (int64_t)mask_counts[2];
  S.perimeter_suppressed_halo_pixels = (int64_t)mask_counts[3];
  S.perimeter_excluded_id = excluded;
  HIP_TRY(hipStreamSynchronize(s));
  return DICE_OK;
}

}  // namespace dice

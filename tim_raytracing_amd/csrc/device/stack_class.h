// Per-lane LDS stack sizes of the traversal kernels, chosen from the depth of the tree(s) a launch
// traces. A walk keeps at most one deferred far child per level of the path from the root to its
// current node (each entry is the sibling of a node on that path), and it pushes only when it steps
// from an inner node at depth k to a child at depth k + 1 <= the depth of the deepest leaf, so a tree
// whose leaves lie at depth <= d never holds more than d entries. Every launch takes its stack from
// here, and every kernel stops a lane instead of writing past its stack column, which this choice
// never lets happen. Plain C++ so that the mapping is checked on a CPU.
#pragma once
#include <cstdint>

namespace rtq {

constexpr uint32_t kMaxTreeDepth = 64;  // rtq_scene_upload rejects deeper trees

// Single-scene kernels: 22, 24, 32 or 64 entries (with 22 a 128-lane block takes 11 KB of LDS and
// the grid kernels reach their register limit of 7 waves per SIMD; with 24, 6.5). 0 if too deep.
inline uint32_t gridStackClass(uint32_t depth) {
  return depth <= 22 ? 22u : depth <= 24 ? 24u : depth <= 32 ? 32u : depth <= kMaxTreeDepth ? 64u : 0u;
}

// Batched and served kernels (one stack for many scenes): 24, 32 or 64 entries. 0 if too deep.
inline uint32_t batchStackClass(uint32_t depth) {
  return depth <= 24 ? 24u : depth <= 32 ? 32u : depth <= kMaxTreeDepth ? 64u : 0u;
}

// Whether a stack of `stack` entries holds every walk of a tree whose leaves lie at depth <= `depth`.
inline bool stackHolds(uint32_t stack, uint32_t depth) { return depth <= stack; }

}  // namespace rtq

/* CPU check of the identity the traversal's slab division rests on: for fp32 x and d with |d| in
 * [2^-30, 2^30] and |x| in [2^-60, 2^60], RN32(RN64(double(x) * RN64(1 / double(d)))) == RN32(x / d).
 * The host's IEEE fp64 division, multiply and fp64 -> fp32 conversion are the operations the device
 * performs. Random pairs (mantissas forced to the extremes a quarter of the time) and pairs whose
 * quotient lies next to a midpoint between fp32 numbers. Prints "ok <pairs>" or the first failure. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint64_t mix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static float fbits(uint32_t v) { float x; memcpy(&x, &v, 4); return x; }
static uint32_t ubits(float x) { uint32_t v; memcpy(&v, &x, 4); return v; }

static int check(float x, float d) {
  volatile float ieee = x / d;
  const float q = (float)((double)x * (1.0 / (double)d));
  if (ubits(ieee) != ubits(q)) {
    printf("mismatch x=%a d=%a ieee=%a fp64=%a\n", x, d, ieee, q);
    return 1;
  }
  return 0;
}

int main(void) {
  uint64_t s = 2024, pairs = 0;
  for (uint64_t it = 0; it < 30000000ull; ++it, ++pairs) {
    const uint64_t r = mix(&s), c = mix(&s);
    uint32_t md = r & 0x7FFFFFu, mx = (r >> 23) & 0x7FFFFFu;
    if ((c & 3u) == 0) md = (c & 4u) ? 0x7FFFFFu : 0u;
    if (((c >> 3) & 3u) == 0) mx = (c & 32u) ? 0x7FFFFFu : 0u;
    const uint32_t ed = 127u - 30u + (uint32_t)((c >> 9) % 60u), ex = 127u - 60u + (uint32_t)((c >> 15) % 120u);
    const float d = fbits((uint32_t)((r >> 62) & 1u) << 31 | ed << 23 | md);
    const float x = fbits((uint32_t)(r >> 63) << 31 | ex << 23 | mx);
    if (check(x, d)) return 1;
  }
  for (uint64_t it = 0; it < 30000000ull; ++it) {
    /* x / d next to the midpoint (2K+1) / 2^h of two fp32 significands K, K+1 */
    const uint64_t r = mix(&s);
    const uint64_t D = (1ull << 23) | (r & 0x7FFFFFu), K = (1ull << 23) | ((r >> 23) & 0x7FFFFFu);
    const unsigned __int128 P = (unsigned __int128)D * (2 * K + 1);
    int h = 0;
    while ((P >> h) >= (1ull << 24)) h++;
    uint64_t X = (uint64_t)((P + ((unsigned __int128)1 << (h - 1))) >> h) + ((r >> 46) & 3u) - 1u;
    if (X < (1ull << 23) || X >= (1ull << 24)) continue;
    float x = ldexpf((float)X, -23 + (int)((r >> 50) % 60u) - 30), d = ldexpf((float)D, -23 + (int)((r >> 56) % 40u) - 20);
    if ((r >> 48) & 1u) x = -x;
    if ((r >> 49) & 1u) d = -d;
    ++pairs;
    if (check(x, d)) return 1;
  }
  printf("ok %llu\n", (unsigned long long)pairs);
  return 0;
}

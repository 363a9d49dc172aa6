// Writes Philox4x32-10 (counter, key) -> output vectors computed by rocRAND's
// public source header (rocrand_philox4x32_10.h), as an independent pin for
// the oracle's Philox.  Host-only; build and run with gen_philox_rocrand.sh.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine {
  uint4 bijection(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

static uint32_t lcg(uint64_t &s) {
  s = s * 6364136223846793005ULL + 1442695040888963407ULL;
  return static_cast<uint32_t>(s >> 32);
}

int main() {
  Probe p;
  uint64_t s = 12345;
  std::printf("# ctr0 ctr1 ctr2 ctr3 key0 key1 -> out0 out1 out2 out3 (hex)\n");
  for (int i = 0; i < 64; ++i) {
    uint4 c;
    uint2 k;
    c.x = lcg(s); c.y = lcg(s); c.z = lcg(s); c.w = lcg(s);
    k.x = lcg(s); k.y = lcg(s);
    if (i == 0) { c = {0, 0, 0, 0}; k = {0, 0}; }
    if (i == 1) { c = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}; k = {0xffffffffu, 0xffffffffu}; }
    if (i == 2) { c = {5, 0, 7, 0}; k = {3, 0}; }
    uint4 o = p.bijection(c, k);
    std::printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c.x, c.y, c.z, c.w,
                k.x, k.y, o.x, o.y, o.z, o.w);
  }
  return 0;
}

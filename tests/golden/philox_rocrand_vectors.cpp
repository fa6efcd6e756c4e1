// Writes Philox4x32-10 vectors from rocRAND's implementation (an implementation
// independent of this repository's), compiled for the host with hipcc:
//   hipcc -O1 --offload-arch=gfx950 -o /tmp/philox_vec tests/golden/philox_rocrand_vectors.cpp && /tmp/philox_vec > tests/golden/philox_rocrand.txt
// Each line: c0 c1 c2 c3 k0 k1 -> r0 r1 r2 r3 (hex).
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

// ten_rounds is a protected member; a derived type exposes it.
struct philox_open : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main()
{
    philox_open eng;
    uint32_t s = 0x12345678u;
    auto next = [&]() { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; };
    const uint32_t fixed[][6] = {
        {0, 0, 0, 0, 0, 0},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1, 2, 3, 4, 5, 6},
    };
    for (auto& f : fixed) {
        uint4 c = {f[0], f[1], f[2], f[3]};
        uint2 k = {f[4], f[5]};
        uint4 r = eng.ten_rounds(c, k);
        std::printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", f[0], f[1], f[2], f[3],
                    f[4], f[5], r.x, r.y, r.z, r.w);
    }
    for (int i = 0; i < 60; ++i) {
        uint4 c = {next(), next(), next(), next()};
        uint2 k = {next(), next()};
        uint4 r = eng.ten_rounds(c, k);
        std::printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c.x, c.y, c.z, c.w, k.x,
                    k.y, r.x, r.y, r.z, r.w);
    }
    return 0;
}

// Writes Philox4x32-10 blocks computed by rocRAND's open-source, host-callable
// engine (rocrand/rocrand_philox4x32_10.h, MIT licence) for a fixed set of
// (counter, key) pairs, as an independent check of the engine's generator.
// Built and run by make_philox_rocrand.py; the output is committed as
// philox_rocrand.json.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine
{
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    Probe p;
    const unsigned int ctr[][4] = {{0u, 0u, 0u, 0u},          {1u, 0u, 0u, 0u},        {0u, 7u, 3u, 0u},
                                   {123456u, 99u, 5u, 1u},    {0xffffffffu, 1u, 2u, 3u}, {16383u, 99u, 999u, 0u},
                                   {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}};
    const unsigned int key[][2] = {{0u, 0u}, {0xdeadbeefu, 0x12345678u}, {1u, 2u}, {0xa4093822u, 0x299f31d0u}};
    std::printf("[\n");
    bool first = true;
    for (auto& c : ctr)
        for (auto& k : key)
        {
            const uint4 o = p.block(uint4{c[0], c[1], c[2], c[3]}, uint2{k[0], k[1]});
            std::printf("%s  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}", first ? "" : ",\n", c[0],
                        c[1], c[2], c[3], k[0], k[1], o.x, o.y, o.z, o.w);
            first = false;
        }
    std::printf("\n]\n");
    return 0;
}

#pragma once
// IEEE binary16 storage type standing in for sycl::half in NodeBox: 16 bits,
// convertible to float.
#include <cstdint>
#include <cstring>

namespace dyphur {

struct half {
    uint16_t bits;
    operator float() const noexcept {
        const uint32_t sign = static_cast<uint32_t>(bits & 0x8000u) << 16;
        const uint32_t exp = (bits >> 10) & 0x1fu, man = bits & 0x3ffu;
        uint32_t out;
        if (exp == 0x1fu) {
            out = sign | 0x7f800000u | (man << 13);  // inf / nan
        } else if (exp != 0) {
            out = sign | ((exp + 112u) << 23) | (man << 13);
        } else if (man == 0) {
            out = sign;
        } else {  // subnormal: renormalise
            uint32_t e = 113u, m = man;
            while (!(m & 0x400u)) { m <<= 1; --e; }
            out = sign | (e << 23) | ((m & 0x3ffu) << 13);
        }
        float f;
        std::memcpy(&f, &out, sizeof f);
        return f;
    }
};
static_assert(sizeof(half) == 2, "half is 16 bits");

}  // namespace dyphur

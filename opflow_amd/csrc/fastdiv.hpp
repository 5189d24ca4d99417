// Division of 32-bit cell indices by a grid extent fixed per launch, as a
// multiply-high and shifts (Granlund & Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, Fig. 4.1): for 1 <= d < 2^32 and
// any 32-bit x,
//   l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
//   t = mulhi(m, x), q = (t + ((x - t) >> min(l, 1))) >> max(l - 1, 0).
// Host- and device-callable; checked on the CPU (tests/cpp_host/fastdiv_check.cpp).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OPF_FD_HD __host__ __device__ __forceinline__
#else
#define OPF_FD_HD inline
#endif

namespace opf {

struct FastDiv {
    uint32_t d = 1, m = 1;
    int s1 = 0, s2 = 0;

    FastDiv() = default;
    explicit FastDiv(uint32_t dv) : d(dv) {
        int l = 0;
        while (l < 32 && (uint64_t(1) << l) < dv) ++l;
        m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - dv)) / dv + 1);
        s1 = l < 1 ? l : 1;
        s2 = l - 1 > 0 ? l - 1 : 0;
    }
    OPF_FD_HD uint32_t div(uint32_t x) const {
        uint32_t t = (uint32_t)(((uint64_t)m * x) >> 32);
        return (t + ((x - t) >> s1)) >> s2;
    }
};

}  // namespace opf

namespace opf {

// Workgroups of a launch are dispatched round-robin over the 8 XCDs (each
// with its own L2). xcd_block maps the hardware block id b of a G-block
// launch to a logical block so that every XCD receives one contiguous range
// of logical blocks: neighbouring cells (the rows and planes of a stencil)
// are then mostly worked on by one XCD and shared through its L2.
// A bijection of [0, G) for every G >= 1 (checked on the CPU).
OPF_FD_HD uint32_t xcd_block(uint32_t b, uint32_t G) {
    const uint32_t q = G >> 3, r = G & 7u, x = b & 7u, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

}  // namespace opf

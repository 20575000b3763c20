// vdb_count.h -- element counts of kernel launches (host and device).
//
// A count is either known on the host when the launch is queued, or it lives in a device word
// that an earlier kernel on the same stream wrote (a run count, a scan total, an emitted pair
// count). The second form lets a whole raster pipeline be captured into one HIP graph with no
// count read back between stages: launches are then sized by a capacity bound and every kernel
// reads the actual count when it runs, clamped to that bound. A device count may be scaled
// (mul: elements per counted item) or divided, rounding up (div: items per tile).
#ifndef VDB_COUNT_H
#define VDB_COUNT_H

#include <hip/hip_runtime.h>
#include <stdint.h>

struct DCount
{
    const void* dev; // device word holding the count (u32 or u64), or null
    uint64_t cap;    // the count itself when dev is null, else its upper bound
    uint32_t wide;   // dev points at a u64
    uint32_t mul;    // device count x mul (0: 1)
    uint32_t div;    // then divided by div, rounding up (0: 1)
    uint32_t pad;

    __device__ __forceinline__ uint64_t get() const
    {
        if (!dev)
            return cap;
        uint64_t v = wide ? *(const uint64_t*)dev : (uint64_t)*(const uint32_t*)dev;
        if (mul)
            v *= mul;
        if (div)
            v = (v + div - 1) / div;
        return v < cap ? v : cap;
    }
};

static inline DCount host_count(uint64_t n)
{
    return DCount{ nullptr, n, 0u, 0u, 0u, 0u };
}
static inline DCount dev_count(const uint32_t* p, uint64_t cap)
{
    return DCount{ p, cap, 0u, 0u, 0u, 0u };
}
static inline DCount dev_count(const uint64_t* p, uint64_t cap)
{
    return DCount{ p, cap, 1u, 0u, 0u, 0u };
}
// ceil(c / d) (c without a divisor of its own)
static inline DCount count_div(DCount c, uint32_t d)
{
    const uint64_t cap = (c.cap + d - 1) / d;
    if (!c.dev)
        return host_count(cap);
    c.cap = cap;
    c.div = d;
    return c;
}
// c x m (c without a scale or divisor of its own)
static inline DCount count_mul(DCount c, uint32_t m)
{
    c.cap *= m;
    if (c.dev)
        c.mul = m;
    return c;
}

#endif

// Wave-uniform base plus per-lane offset addressing for gfx950 kernels.
//
// A kernel whose lanes read many arrays at addresses (uniform base) + (lane's
// offset) can keep every base in scalar registers (readfirstlane) and one
// 32-bit byte offset per lane, instead of a 64-bit address per array and
// neighbour.  The base must be the same on every active lane of the wave;
// the address space stays global, so the loads and stores stay global ones.
#pragma once

#include <hip/hip_runtime.h>

#define EMS_GLOBAL __attribute__((address_space(1)))

namespace ems {

__device__ __forceinline__ unsigned long long uniform_bits(const void *p)
{
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Element at byte offset boff from the wave-uniform base.
template <typename T> __device__ __forceinline__ EMS_GLOBAL T &lane_at(T *base, unsigned boff)
{
    EMS_GLOBAL char *b = reinterpret_cast<EMS_GLOBAL char *>(uniform_bits(base));
    return *reinterpret_cast<EMS_GLOBAL T *>(b + boff);
}
template <typename T> __device__ __forceinline__ const EMS_GLOBAL T &lane_at(const T *base, unsigned boff)
{
    const EMS_GLOBAL char *b = reinterpret_cast<const EMS_GLOBAL char *>(uniform_bits(base));
    return *reinterpret_cast<const EMS_GLOBAL T *>(b + boff);
}

}  // namespace ems

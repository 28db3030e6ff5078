// Double-double (hi + lo) sums: TwoSum accumulation, so a reduced value is
// the correctly rounded sum of its terms whatever the summation order.
#pragma once

#include <hip/hip_runtime.h>

namespace mrfg {

struct DD {
    double hi, lo;
};

__device__ inline DD twoSum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}

__device__ inline DD ddAdd(DD x, DD y)
{
    const DD s = twoSum(x.hi, y.hi);
    const double e = s.lo + (x.lo + y.lo);
    const double hi = s.hi + e;
    return DD{hi, e - (hi - s.hi)};
}

__device__ inline DD waveSum(DD v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = ddAdd(v, DD{__shfl_xor(v.hi, o, 64), __shfl_xor(v.lo, o, 64)});
    return v;
}

}  // namespace mrfg

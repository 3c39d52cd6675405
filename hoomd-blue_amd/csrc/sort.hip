// sort.hip -- 64-bit key radix sort for the Hilbert particle order (hipCUB device radix sort).
// Runs once per sorter period (500 steps in the benchmark workload), so a library sort is used
// here rather than a hand-written kernel.
#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace hm
{

hipError_t sort_keys64(unsigned long long* keys_in, unsigned long long* keys_out, unsigned int N,
                       int end_bit, void* tmp, size_t* tmp_bytes, hipStream_t s)
{
    return hipcub::DeviceRadixSort::SortKeys(tmp, *tmp_bytes, keys_in, keys_out, (int)N, 0, end_bit,
                                             s);
}

} // namespace hm

// host_copy.cpp -- multi-threaded host memcpy for result read-back (the last
// leg of a device-to-host copy, from the pinned staging buffer into the
// caller's arrays).
#include <omp.h>

#include <algorithm>
#include <cstddef>
#include <cstring>

namespace cp {

void par_memcpy(void* dst, const void* src, size_t bytes) {
    constexpr size_t PIECE = size_t(256) << 10;
    const int nt = (int)std::min<size_t>((size_t)std::min(8, omp_get_max_threads()), bytes / PIECE);
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes + nt - 1) / nt;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int t = 0; t < nt; ++t) {
        const size_t a = (size_t)t * per, b = std::min(bytes, a + per);
        if (a < b) std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
    }
}

}  // namespace cp

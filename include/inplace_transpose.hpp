// inplace_transpose.hpp — C++ mirror of the reference interface over the C ABI.
//
// Same names and argument meaning as the reference's include/transpose.h and
// include/openmp.h, so a caller of the reference recompiles against this
// header and links libinplace_hip.so:
//
//   inplace::transpose<T>(T* data, int d1, int d2, int d3)      src/transpose.cu:99-124
//   inplace::c2r::transpose<T>(T* data, int d1, int d2, int d3) src/transpose.cu:30-47
//   inplace::r2c::transpose<T>(T* data, int d1, int d2, int d3) src/transpose.cu:70-89
//   inplace::openmp::transpose(bool row_major, T* data, int m, int n, T* tmp)
//                                                               include/openmp.h:7-8
//
// T is any 4- or 8-byte type (the reference instantiates int, long long, float,
// double).  For the GPU forms, `data` is device memory of the current device,
// managed memory (prefetched to the device first, as the reference does) or
// pinned host memory; pageable host memory is refused.  Calls are
// enqueued on the null stream, like the reference; the overloads taking a
// hipStream_t-compatible `void* stream` enqueue there instead.  The openmp
// form runs on the host CPU (ipt_cpu_transpose) on host memory, as the
// reference's declaration implies; `tmp` is optional caller scratch (NULL, or
// at least ipt_cpu_scratch_bytes(sizeof(T), d1, d2, 1, threads) bytes), and
// `threads` <= 0 uses one thread per hardware thread.
//
// Error behaviour: the reference prints and exits on CUDA errors
// (src/cudacheck.cu:4-10) and throws on bad divisors / old GPUs.  Here every
// failure throws inplace::error (a std::runtime_error carrying the ipt_status);
// nothing exits.  The reference declares the openmp form but never
// implements it; here it is the library's host path.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "inplace_transpose.h"

namespace inplace {

class error : public std::runtime_error {
  public:
    error(ipt_status s, const std::string& what)
        : std::runtime_error(what + ": " + ipt_status_string(s) + " " + ipt_last_error()),
          status(s) {}
    ipt_status status;
};

namespace detail {
inline void check(ipt_status s, const char* what) {
    if (s != IPT_OK) throw error(s, what);
}
template <typename T> constexpr int elem_bytes() {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte element types only");
    return (int)sizeof(T);
}
}  // namespace detail

template <typename T> void transpose(T* data, int d1, int d2, int d3, void* stream = nullptr) {
    detail::check(ipt_transpose_213(data, detail::elem_bytes<T>(), d1, d2, d3, stream),
                  "inplace::transpose");
}

namespace c2r {
template <typename T> void transpose(T* data, int d1, int d2, int d3, void* stream = nullptr) {
    detail::check(ipt_c2r_transpose(data, detail::elem_bytes<T>(), d1, d2, d3, stream),
                  "inplace::c2r::transpose");
}
}  // namespace c2r

namespace r2c {
template <typename T> void transpose(T* data, int d1, int d2, int d3, void* stream = nullptr) {
    detail::check(ipt_r2c_transpose(data, detail::elem_bytes<T>(), d1, d2, d3, stream),
                  "inplace::r2c::transpose");
}
}  // namespace r2c

// Any permutation of a rank-3 tensor in place: dims fastest first, output
// dim j is source dim perm[j] (the oracle's seq_tt convention).  No
// counterpart in the reference's in-place API.
template <typename T>
void permute3(T* data, int d1, int d2, int d3, const int (&perm)[3], void* stream = nullptr) {
    detail::check(ipt_permute3(data, detail::elem_bytes<T>(), d1, d2, d3, perm, stream),
                  "inplace::permute3");
}

namespace openmp {
template <typename T>
void transpose(bool row_major, T* data, int m, int n, T* tmp = nullptr, int threads = 0) {
    detail::check(ipt_cpu_transpose(row_major ? 1 : 0, data, detail::elem_bytes<T>(), m, n, tmp,
                                    threads),
                  "inplace::openmp::transpose");
}
}  // namespace openmp

}  // namespace inplace

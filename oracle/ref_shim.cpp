// ref_shim.cpp — C entry points over the reference's own host arithmetic.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
// reference's src/gcd.cpp and src/reduced_math.cpp, taken where they lie under
// the reference checkout (see oracle/Makefile), into oracle/_ref/libref_math.so.
// The tests use it to check the oracle's restatements of extended_gcd
// (src/gcd.cpp:3-30) and find_divisor (src/reduced_math.cpp:36-67) against
// the reference's compiled code.  Only declarations are written here.
#include <stdexcept>

namespace inplace {
void extended_gcd(int a, int b, int& gcd, int& mmi);
namespace detail {
void find_divisor(unsigned int denom, unsigned int& mul_coeff, unsigned int& shift_coeff);
void find_divisor(unsigned long long denom, unsigned long long& mul_coeff,
                  unsigned int& shift_coeff);
unsigned long long host_umulhi(unsigned long long x, unsigned long long y);
}  // namespace detail
}  // namespace inplace

extern "C" {

void ref_extended_gcd(int a, int b, int* gcd, int* mmi) {
    inplace::extended_gcd(a, b, *gcd, *mmi);
}

int ref_find_divisor32(unsigned int d, unsigned int* mul, unsigned int* shift) {
    try {
        inplace::detail::find_divisor(d, *mul, *shift);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    return 0;
}

int ref_find_divisor64(unsigned long long d, unsigned long long* mul, unsigned int* shift) {
    try {
        inplace::detail::find_divisor(d, *mul, *shift);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    return 0;
}

unsigned long long ref_umulhi64(unsigned long long x, unsigned long long y) {
    return inplace::detail::host_umulhi(x, y);
}

}  // extern "C"

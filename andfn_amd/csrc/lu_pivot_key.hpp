// Pivot key of the tournament LU's row search (aem_lu.hip), kept in a header of its own so the
// ordering can be checked on the host (tests/test_lu_pivot_key.py builds it with g++).
//
// A candidate row's key is the IEEE-754 bit pattern of |a| (sign bit cleared: unsigned order is
// then the order of the magnitudes, and a NaN ranks above every number, so a NaN row is still
// taken and the winner list stays full) with its lowest kIdxBits bits replaced by the inverted
// thread index.  One 64-bit max over the group then selects the largest |a|, ties (within
// 2^-(52 - kIdxBits) relative) going to the lowest thread.  The inverted index is at least
// 2^(kIdxBits - 1), so every active row has a nonzero key, also for a = 0 or a subnormal a, and
// 0 marks a row that does not take part.  Magnitudes below 2^kIdxBits subnormal steps (~5e-321)
// rank as zero.
#ifndef ANDFN_LU_PIVOT_KEY_HPP
#define ANDFN_LU_PIVOT_KEY_HPP

namespace aem {

constexpr int kPivotIdxBits = 10;  // rows per group <= 2^(kPivotIdxBits - 1) = 512
constexpr unsigned long long kPivotIdxMask = (1ull << kPivotIdxBits) - 1ull;

constexpr unsigned long long lu_pivot_key(double a, int thread) {
  const unsigned long long mag = __builtin_bit_cast(unsigned long long, a) & 0x7FFFFFFFFFFFFFFFull;
  return (mag & ~kPivotIdxMask) | (kPivotIdxMask - (unsigned long long)thread);
}

constexpr int lu_pivot_thread(unsigned long long key) { return (int)(kPivotIdxMask - (key & kPivotIdxMask)); }

}  // namespace aem

#endif

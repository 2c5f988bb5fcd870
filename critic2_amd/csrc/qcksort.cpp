// The tie order of critic2's quicksort (qcksort_r8, tools@proc.f90:83-168).
//
// yt_integrate ranks the grid points with qcksort (yt@proc.f90:85-93) and
// treats a neighbour as higher when its rank is higher, so between equal
// values the reference's answer is this sort's arrangement: a quicksort with
// a pseudo-random pivot (a linear congruential sequence that advances once
// per partition, in the order the partitions are taken), a partition that
// moves elements through a hole from the right and the left in turn, and an
// insertion sort below 7 elements.  Only the same steps give the same
// arrangement of equal keys, so the steps are the reference's; the library
// calls this only when a grid has equal values where the order matters.
#include <cmath>
#include <cstdint>

#include "common.hpp"

namespace c2 {

// iord[first..last] (0-based positions, element indices into arr) sorted by
// arr ascending, equal keys in qcksort_r8's arrangement.
void qcksort_order(const double *arr, uint32_t *iord, int64_t first, int64_t last) {
  constexpr int M = 7, NSTACK = 64;  // the reference stops at 50 (2 log2 of 2^25)
  constexpr double FM = 7875.0, FA = 211.0, FC = 1663.0, FMI = 1.2698413e-4;
  int64_t stack[NSTACK];
  int top = 0;
  int64_t l = first, ir = last;
  double fx = 0.0;
  for (;;) {
    if (ir - l < M) {
      for (int64_t j = l + 1; j <= ir; j++) {
        const uint32_t na = iord[j];
        const double a = arr[na];
        int64_t i = j - 1;
        while (i >= first && arr[iord[i]] > a) {
          iord[i + 1] = iord[i];
          i--;
        }
        iord[i + 1] = na;
      }
      if (top == 0) return;
      ir = stack[--top];
      l = stack[--top];
      continue;
    }
    int64_t i = l, j = ir;
    fx = std::fmod(fx * FA + FC, FM);
    const int64_t iq = l + (int64_t)((double)(ir - l + 1) * (fx * FMI));
    const uint32_t na = iord[iq];
    const double a = arr[na];
    iord[iq] = iord[l];
    for (;;) {
      while (j >= first && a < arr[iord[j]]) j--;
      if (j <= i) {
        iord[i] = na;
        break;
      }
      iord[i] = iord[j];
      i++;
      while (i <= last && a > arr[iord[i]]) i++;
      if (j <= i) {
        iord[j] = na;
        i = j;
        break;
      }
      iord[j] = iord[i];
      j--;
    }
    if (top + 2 > NSTACK) fail(C2GPU_ERR_ARG, "qcksort: stack exhausted");
    if (ir - i >= i - l) {  // the larger part waits on the stack
      stack[top++] = i + 1;
      stack[top++] = ir;
      ir = i - 1;
    } else {
      stack[top++] = l;
      stack[top++] = i - 1;
      l = i + 1;
    }
  }
}

}  // namespace c2

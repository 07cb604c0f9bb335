// Test harness (CPU): spx::div_rn (spectrum_amd/csrc/spx_vec.h), the FP64 kernels' division of several
// numerators by one denominator from a shared reciprocal, against the correctly rounded quotient a / b.
// Compiled for the host with hipcc (the header includes the HIP runtime header), without contraction.
#include <stdint.h>
#include <string.h>

#include "../../spectrum_amd/csrc/spx_vec.h"

static uint64_t bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }

// For each pair: 0 if div_rn(a, b, 1 / b) has the bits of a / b; the count of mismatches is returned and
// the first mismatching index written to *first.
extern "C" long div_rn_mismatches(const double* a, const double* b, long n, long* first)
{
   long bad = 0;
   *first = -1;
   for (long i = 0; i < n; i++) {
      const double y = 1.0 / b[i];
      const double q = div_rn(a[i], b[i], y);
      const double want = a[i] / b[i];
      if (bits(q) != bits(want) && !(q == 0.0 && want == 0.0)) {
         if (bad == 0) *first = i;
         bad++;
      }
   }
   return bad;
}

// The vector form the FP64 kernels use (v / a with the shared reciprocal)
extern "C" void v3_div(const double* v, double a, double* out)
{
   const v3<double> r = mk<double>(v[0], v[1], v[2]) / a;
   out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

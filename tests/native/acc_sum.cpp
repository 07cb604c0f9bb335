// Test harness (CPU): exact histogram sums of the engine (spectrum_amd/csrc/spx_acc.h) over a list of
// doubles, added limb by limb in a caller-chosen order, the way the device adds event weights.
#include <stddef.h>

#include "../../spectrum_amd/csrc/spx_acc.h"

extern "C" double acc_sum(const double* v, const long* order, size_t n)
{
   unsigned long long limbs[SPX_ACC_LIMBS] = {0};
   for (size_t i = 0; i < n; i++) {
      const spx::AccParts a = spx::acc_split(v[order ? order[i] : (long)i]);
      if (a.special) { limbs[SPX_ACC_FLAGS] |= (unsigned long long)a.special; continue; }
      limbs[a.idx] += (unsigned long long)a.c0;
      limbs[a.idx + 1] += (unsigned long long)a.c1;
      limbs[a.idx + 2] += (unsigned long long)a.c2;
   }
   return spx::acc_value(limbs);
}

extern "C" double acc_round_trip(double v)
{
   unsigned long long limbs[SPX_ACC_LIMBS];
   spx::acc_from_double(v, limbs);
   return spx::acc_value(limbs);
}

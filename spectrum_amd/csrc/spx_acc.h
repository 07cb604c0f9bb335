// spx_acc.h -- exact, order-independent FP64 histogram sums.
//
// Every component of every bin is a fixed-point integer wide enough for any finite double: limb i
// (i = 0 .. SPX_ACC_DIGITS-1) holds a signed sum of 32-bit digits of weight 2^(32 i - 1074), the
// next limb takes the carries of the final normalisation, and the last one collects NaN / +inf /
// -inf flags.  A double m * 2^(k - 1074) (m the 53-bit significand, k >= 0) adds at most three
// digits to limbs k/32 .. k/32 + 2.  Integer addition is associative, so the sum does not depend on
// the order in which events arrive (wave scheduling, launch shape, queries served beside it), and
// the value read back is the exact sum rounded once to nearest.  Each addition changes a limb by
// less than 2^32 per event, so a limb is exact for up to 2^31 events per bin component.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define SPX_ACC_HD __host__ __device__
#else
#define SPX_ACC_HD   // host-only compilation (the exactness test builds this header with g++)
#endif

#define SPX_ACC_DIGITS 66                  // limbs 0..65: bit 0 = 2^-1074 .. bit 2111
#define SPX_ACC_LIMBS 68                   // + carry limb + flag limb
#define SPX_ACC_FLAGS (SPX_ACC_LIMBS - 1)

namespace spx {

enum { ACC_NAN = 1, ACC_PINF = 2, ACC_NINF = 4 };

// A double split into the limb index of its lowest digit, three signed digits and special flags.
struct AccParts {
   int32_t idx, special;
   long long c0, c1, c2;
};

SPX_ACC_HD inline AccParts acc_split(double v)
{
   AccParts r{0, 0, 0, 0, 0};
   uint64_t bits;
   __builtin_memcpy(&bits, &v, 8);
   const int e = (int)((bits >> 52) & 0x7ffu);
   uint64_t m = bits & 0x000fffffffffffffull;
   const bool neg = (bits >> 63) != 0;
   if (e == 0x7ff) {
      r.special = m ? ACC_NAN : (neg ? ACC_NINF : ACC_PINF);
      return r;
   }
   int k = 0;                               // subnormal: m * 2^-1074
   if (e) { m |= 0x0010000000000000ull; k = e - 1; }
   r.idx = k >> 5;
   const int sh = k & 31;
   const uint64_t lo = m << sh;
   const uint64_t hi = sh ? (m >> (64 - sh)) : 0ull;
   r.c0 = (long long)(lo & 0xffffffffull);
   r.c1 = (long long)(lo >> 32);
   r.c2 = (long long)hi;
   if (neg) { r.c0 = -r.c0; r.c1 = -r.c1; r.c2 = -r.c2; }
   return r;
}

// Host side: the exact sum held by one component's limbs, rounded to nearest (ties to even).
inline double acc_value(const unsigned long long* limbs)
{
   const unsigned long long fl = limbs[SPX_ACC_FLAGS];
   if ((fl & ACC_NAN) || ((fl & ACC_PINF) && (fl & ACC_NINF))) return __builtin_nan("");
   if (fl & ACC_PINF) return __builtin_inf();
   if (fl & ACC_NINF) return -__builtin_inf();
   // normalise: digits 0..65 in [0, 2^32), the signed rest in limb 66
   long long d[SPX_ACC_DIGITS + 1];
   long long carry = 0;
   for (int i = 0; i < SPX_ACC_DIGITS; i++) {
      const __int128 t = (__int128)(long long)limbs[i] + carry;
      const long long low = (long long)(t & 0xffffffff);
      d[i] = low;
      carry = (long long)((t - low) >> 32);
   }
   d[SPX_ACC_DIGITS] = (long long)limbs[SPX_ACC_DIGITS] + carry;
   bool neg = d[SPX_ACC_DIGITS] < 0;
   if (neg) {   // two's complement of the digit string
      long long c = 1;
      for (int i = 0; i < SPX_ACC_DIGITS; i++) {
         const long long t = (0xffffffffll - d[i]) + c;
         d[i] = t & 0xffffffffll;
         c = t >> 32;
      }
      d[SPX_ACC_DIGITS] = -d[SPX_ACC_DIGITS] - 1 + c;
   }
   if (d[SPX_ACC_DIGITS] != 0) return neg ? -__builtin_inf() : __builtin_inf();   // beyond 2^1038
   int top = SPX_ACC_DIGITS - 1;
   while (top >= 0 && d[top] == 0) top--;
   if (top < 0) return 0.0;
   // a 96-bit window of the top digits and a sticky bit of everything below
   unsigned __int128 w = (unsigned __int128)(uint64_t)d[top] << 64;
   if (top >= 1) w |= (unsigned __int128)(uint64_t)d[top - 1] << 32;
   if (top >= 2) w |= (unsigned __int128)(uint64_t)d[top - 2];
   bool sticky = false;
   for (int i = top - 3; i >= 0; i--) sticky = sticky || d[i] != 0;
   int bl = 0;
   for (unsigned __int128 t = w; t; t >>= 1) bl++;
   int shift = bl - 53;
   uint64_t mant;
   if (shift > 0) {
      const unsigned __int128 rem = w & (((unsigned __int128)1 << shift) - 1);
      const unsigned __int128 half = (unsigned __int128)1 << (shift - 1);
      mant = (uint64_t)(w >> shift);
      if (rem > half || (rem == half && (sticky || (mant & 1ull)))) mant++;
      if (mant == (1ull << 53)) { mant >>= 1; shift++; }
   }
   else {
      mant = (uint64_t)w;
      shift = 0;
   }
   // the window's lowest bit has weight 2^(32 (top - 2) - 1074); sums below 2^-1022 are exact multiples
   // of 2^-1074 with fewer than 53 significant bits, so the scaling below rounds nothing
   const double r = __builtin_ldexp((double)mant, shift + 32 * (top - 2) - 1074);
   return neg ? -r : r;
}

// Host side: limbs holding exactly v (a restored checkpoint value).
inline void acc_from_double(double v, unsigned long long* limbs)
{
   for (int i = 0; i < SPX_ACC_LIMBS; i++) limbs[i] = 0ull;
   const AccParts p = acc_split(v);
   if (p.special) { limbs[SPX_ACC_FLAGS] = (unsigned long long)p.special; return; }
   limbs[p.idx] = (unsigned long long)p.c0;
   limbs[p.idx + 1] = (unsigned long long)p.c1;
   limbs[p.idx + 2] = (unsigned long long)p.c2;
}

}  // namespace spx

// spx_vec.h -- 3-vector and 3x3-matrix arithmetic for the trajectory kernels.
//
// The FP64 kernels must reproduce the reference's rounding, so every operation keeps the
// evaluation order of common/vectors.{hh,cc}, common/simple_array.hh and common/matrix.hh:
// dot products sum left to right, a*v scales each component, v/a divides each component,
// basis changes follow ChangeToBasis / ChangeFromBasis (common/vectors.cc:58-93).
// Built with -ffp-contract=off so that nothing is fused into multiply-adds.
#pragma once

#include <hip/hip_runtime.h>

#define SPX_HD __host__ __device__ __forceinline__

template <typename R>
struct v3 {
   R x, y, z;
   SPX_HD R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

template <typename R> SPX_HD v3<R> mk(R x, R y, R z) { v3<R> r; r.x = x; r.y = y; r.z = z; return r; }
template <typename R> SPX_HD v3<R> operator+(v3<R> a, v3<R> b) { return mk<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> SPX_HD v3<R> operator-(v3<R> a, v3<R> b) { return mk<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> SPX_HD v3<R> operator-(v3<R> a) { return mk<R>(-a.x, -a.y, -a.z); }
template <typename R> SPX_HD v3<R> operator*(R a, v3<R> b) { return mk<R>(a * b.x, a * b.y, a * b.z); }
template <typename R> SPX_HD v3<R> operator/(v3<R> b, R a) { return mk<R>(b.x / a, b.y / a, b.z / a); }
// a / b rounded as the division rounds it, from y = 1 / b (itself a correctly rounded division): q = a y,
// then twice q + (a - b q) y with the residual exact in a multiply-add.  The first correction makes q
// faithful; the second then rounds to nearest (Markstein's theorem: y within half an ulp of 1 / b and q
// within one ulp of a / b give RN(q + (a - b q) y) = RN(a / b); P. Markstein, IBM J. Res. Dev. 34 (1990)
// 111-119).  Five operations instead of a division's ~11 when several numerators share b.  For finite,
// nonzero b far from overflow and underflow (the fields of a step); a zero quotient comes out +0.
SPX_HD double div_rn(double a, double b, double y)
{
   double q = a * y;
   double r = __builtin_fma(-b, q, a);
   q = __builtin_fma(r, y, q);
   r = __builtin_fma(-b, q, a);
   return __builtin_fma(r, y, q);
}
template <> SPX_HD v3<double> operator/(v3<double> b, double a)
{
   const double y = 1.0 / a;
   return mk<double>(div_rn(b.x, a, y), div_rn(b.y, a, y), div_rn(b.z, a, y));
}
template <typename R> SPX_HD R dot(v3<R> a, v3<R> b) { R s = a.x * b.x; s += a.y * b.y; s += a.z * b.z; return s; }
template <typename R> SPX_HD R norm2(v3<R> a) { R s = a.x * a.x; s += a.y * a.y; s += a.z * a.z; return s; }
// sqrt(x) in the FP64 kernels: the correctly rounded sequence the compiler emits for sqrt (reciprocal square
// root seed, then g = x y, h = y / 2 refined by r = 1/2 - h g and two residual corrections d = x - g g),
// without its rescaling of arguments below 2^-767, which is the identity for every other argument, and
// with its pass-through of +-0 and +inf.  For x = 0, +inf or x >= 2^-767: squared lengths, fields and
// momenta of a step, never that small.  The host keeps sqrt.
SPX_HD double sqrt_rn(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
   const double y = __builtin_amdgcn_rsq(x);
   double g = x * y, h = y * 0.5;
   const double r = __builtin_fma(-h, g, 0.5);
   g = __builtin_fma(g, r, g);
   double d = __builtin_fma(-g, g, x);
   h = __builtin_fma(h, r, h);
   g = __builtin_fma(d, h, g);
   d = __builtin_fma(-g, g, x);
   g = __builtin_fma(d, h, g);
   return (x == 0.0 || x == __builtin_inf()) ? x : g;
#else
   return sqrt(x);
#endif
}
template <typename R> SPX_HD R norm(v3<R> a) { return sqrt(norm2(a)); }
template <> SPX_HD double norm(v3<double> a) { return sqrt_rn(norm2(a)); }
template <typename R> SPX_HD v3<R> unit(v3<R> a) { R n = norm(a); return a / n; }
template <typename R> SPX_HD v3<R> cross(v3<R> a, v3<R> b)
{
   return mk<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// ChangeToBasis: components along the three basis vectors
template <typename R> SPX_HD v3<R> to_basis(v3<R> a, const v3<R>* b)
{
   return mk<R>(a.x * b[0].x + a.y * b[0].y + a.z * b[0].z,
                a.x * b[1].x + a.y * b[1].y + a.z * b[1].z,
                a.x * b[2].x + a.y * b[2].y + a.z * b[2].z);
}
// ChangeFromBasis: back to the global frame
template <typename R> SPX_HD v3<R> from_basis(v3<R> a, const v3<R>* b)
{
   return mk<R>(a.x * b[0].x + a.y * b[1].x + a.z * b[2].x,
                a.x * b[0].y + a.y * b[1].y + a.z * b[2].y,
                a.x * b[0].z + a.y * b[1].z + a.z * b[2].z);
}
// GetSecondUnitVec (common/vectors.cc:453-468): cross with the axis of the smallest component
template <typename R> SPX_HD v3<R> second_unit(v3<R> f)
{
   R ax = fabs(f.x), ay = fabs(f.y), az = fabs(f.z);
   int k;
   if (ax < ay) k = (ax < az) ? 0 : 2;
   else k = (ay < az) ? 1 : 2;
   v3<R> e = mk<R>(k == 0 ? R(1) : R(0), k == 1 ? R(1) : R(0), k == 2 ? R(1) : R(0));
   return unit(cross(f, e));
}

template <typename R>
struct m3 {
   v3<R> r[3];
};
// M * v (common/matrix.hh:480-483)
template <typename R> SPX_HD v3<R> mul(const m3<R>& m, v3<R> a) { return mk<R>(dot(m.r[0], a), dot(m.r[1], a), dot(m.r[2], a)); }
// M1 * M2 (common/matrix.hh:509-519): row of M1 dotted with column of M2
template <typename R> SPX_HD m3<R> mul(const m3<R>& l, const m3<R>& rr)
{
   m3<R> o;
#pragma unroll
   for (int i = 0; i < 3; i++) {
      v3<R> c0 = mk<R>(rr.r[0].x, rr.r[1].x, rr.r[2].x);
      v3<R> c1 = mk<R>(rr.r[0].y, rr.r[1].y, rr.r[2].y);
      v3<R> c2 = mk<R>(rr.r[0].z, rr.r[1].z, rr.r[2].z);
      o.r[i] = mk<R>(dot(l.r[i], c0), dot(l.r[i], c1), dot(l.r[i], c2));
   }
   return o;
}
template <typename R> SPX_HD m3<R> transpose(const m3<R>& m)
{
   m3<R> t;
   t.r[0] = mk<R>(m.r[0].x, m.r[1].x, m.r[2].x);
   t.r[1] = mk<R>(m.r[0].y, m.r[1].y, m.r[2].y);
   t.r[2] = mk<R>(m.r[0].z, m.r[1].z, m.r[2].z);
   return t;
}
template <typename R> SPX_HD R sqr(R x) { return x * x; }

// Correctly rounded float32 division by a launch constant without the
// hardware division sequence.
//
// The stencils divide central differences by 2h (numpy.gradient's interior
// formula) and must give the IEEE quotient bit for bit.  With y = RN(1/b)
// computed once on the host, Markstein's theorem gives the correctly rounded
// quotient from fused multiply-adds (Markstein 1990, IBM J. Res. Dev. 34(1);
// Muller et al., Handbook of Floating-Point Arithmetic, 2nd ed., Th. 4.10):
//   if y is within half an ulp of 1/b and q is within one ulp of a/b, then
//   r = a - b q is exact and RN(q + r y) = RN(a / b).
// q0 = RN(a y) can be 1.5 ulp away, so one correction first makes a faithful
// q1 (error below half an ulp plus 2^-46 relative), then the theorem's step.
// The theorem needs r to stay normal and the quotient in range: callers take
// this path only for |b| in [2^-40, 2^40]; a nonzero |a| outside
// [2^-75, 2^75] (subnormal, huge, inf or NaN) takes the IEEE division.
//
// The functions are host and device (clang's force_cuda_host_device), so the
// host build that checks them against IEEE division compiles this same code.
#pragma once

#pragma clang force_cuda_host_device begin
namespace fsg {

// |b| range in which div_rn's fast path holds (checked on the host).
inline bool fdiv_ok(float b) {
  const float m = __builtin_fabsf(b);
  return m >= 0x1p-40f && m <= 0x1p40f;
}

// RN(a / b) given y = RN(1/b) and fdiv_ok(b).  A zero dividend returns
// a * y, the signed zero of the IEEE quotient.
inline float div_rn(float a, float b, float y) {
  const float q0 = a * y;
  const float r0 = __builtin_fmaf(-b, q0, a);
  const float q1 = __builtin_fmaf(r0, y, q0);
  const float r1 = __builtin_fmaf(-b, q1, a);
  const float q2 = __builtin_fmaf(r1, y, q1);
  const float m = __builtin_fabsf(a);
  if (m >= 0x1p-75f && m <= 0x1p75f) return q2;
  if (m == 0.0f) return q0;
  return a / b;
}

// The same quotient without the range guard, for a caller whose output is
// checked to a tolerance (curvature): equal to RN(a / b) for 2^-75 <= |a| <=
// 2^75, a zero (of either sign) for a = 0; outside that range it may be one
// ulp off (a subnormal remainder) or NaN where the quotient is infinite.
inline float div_rn_unguarded(float a, float b, float y) {
  const float q0 = a * y;
  const float r0 = __builtin_fmaf(-b, q0, a);
  const float q1 = __builtin_fmaf(r0, y, q0);
  const float r1 = __builtin_fmaf(-b, q1, a);
  return __builtin_fmaf(r1, y, q1);
}

}  // namespace fsg
#pragma clang force_cuda_host_device end

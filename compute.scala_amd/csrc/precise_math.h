// FP32 exp and tanh for the generated kernels' precise and accurate math modes.
//
// The reference's kernels call OpenCL's exp and tanh (OpenCLKernelBuilder.scala:295-325).
//
// Precise mode (each result within 2 ulp, checked over every FP32 input on the device):
//   csmi_tanhf_p(x), branch free: for |x| < 0.625 the odd polynomial x + x^3 P(x^2), P a
//            degree-4 weighted least-squares fit of (tanh(x) - x) / x^3 (about 0.86 ulp in
//            FP32 arithmetic); above, 1 - 2 / (E + 1) with E = e^{2|x|} from the hardware 2^t
//            on t = 2|x| log2(e) >= 1.8. The error of E reaches the result scaled by
//            2/(E + 1)^2 * E, small enough that the whole function stays within 2 ulp. Large
//            |x| needs no check: E overflows to +inf, its reciprocal is 0 and the result 1.
//   exp is the device library's expf (branch free, 1 ulp), or csmi_expf_np where the
//            argument cannot overflow (the same steps without the overflow check).
//
// Accurate mode (each result within about 0.7 ulp, checked likewise), FMA-rate only:
//   exp(x) = 2^n (1 + q),  n = rint(x log2 e),  r = x - n ln2 (Cody-Waite, ln2 in two
//            parts, r kept as rh + rl), q = r + r^2 P(r) with a degree-5 least-squares
//            polynomial P ~ (e^r - 1 - r) / r^2 on |r| <= ln2/2 carried as a double-float
//            (qh, ql), then 1 + q by Fast2Sum and one rounding.
//   tanh(x) = sign(x) E / (E + 2),  E = expm1(2|x|) = 2^n q + (2^n - 1) from the same
//            reduction as a double-float; the quotient is refined by one Newton step on its
//            residual. No branch on |x|: small arguments keep full relative accuracy
//            because E carries q directly.
#pragma once

__device__ __forceinline__ float csmi_tanhf_p(float x) {
#pragma clang fp contract(off) reassociate(off)  // the rounding steps are part of the algorithm
  const float a = fabsf(x);
  // |x| >= 0.625: 1 - 2 / (2^t + 1), t = 2a log2(e) >= 1.8; for t >= 128, 2^t is +inf and
  // the reciprocal 0, which gives the 1 that tanh rounds to beyond 9.01
  const float e = __builtin_amdgcn_exp2f(a * 2.88539004f);
  const float big = fmaf(__builtin_amdgcn_rcpf(e + 1.0f), -2.0f, 1.0f);
  // |x| < 0.625: a + a^3 P(a^2)
  const float s = a * a;
  float p = -0.005692214239388704f;
  p = fmaf(p, s, 0.02062651701271534f);
  p = fmaf(p, s, -0.0537353977560997f);
  p = fmaf(p, s, 0.13331381976604462f);
  p = fmaf(p, s, -0.3333328068256378f);
  const float small = fmaf(a * s, p, a);
  return copysignf(a < 0.625f ? small : big, x);  // NaN: a is NaN, so is big
}

// exp for an argument that cannot exceed 88 (the generated code uses it where the argument
// is the negation of a square or of an absolute value): 2^n 2^f with n = rint(x log2 e), the
// reduced argument f = x log2(e) - n from log2 e as a float pair L1 + L2 (FMA keeps x L1 - n
// exact before its one rounding), the hardware 2^f on |f| <= 1/2 and an exact scaling that
// produces subnormal results; below -104 the result rounds to +0 (as does -inf, for which f
// would be NaN). No overflow check is needed.
__device__ __forceinline__ float csmi_expf_np(float x) {
#pragma clang fp contract(off) reassociate(off)  // the rounding steps are part of the algorithm
  const float n = rintf(x * 1.44269502f);
  // f = x log2(e) - n, |f| <= 1/2 + 2^-17: one rounding for x L1 - n, one for the x L2 term
  const float f = fmaf(x, 1.92596299e-08f, fmaf(x, 1.44269502f, -n));
  const float r = ldexpf(__builtin_amdgcn_exp2f(f), static_cast<int>(n));
  return x < -104.0f ? 0.0f : r;  // NaN fails the compare and r is NaN
}

__device__ __forceinline__ void csmi_exp_reduce(float x, float &n, float &qh, float &ql) {
#pragma clang fp contract(off) reassociate(off)  // the rounding steps are part of the algorithm
  n = rintf(x * 1.44269502f);
  const float r1 = fmaf(n, -0.693145751953125f, x);  // n * L1 is exact (L1 has 15 bits)
  const float rh = fmaf(n, -1.42860677e-06f, r1);
  const float rl = fmaf(n, -1.42860677e-06f, r1 - rh);  // what rounding rh left behind
  float p = 0.000198495268705301f;
  p = fmaf(p, rh, 0.0013933626469224691f);
  p = fmaf(p, rh, 0.008333372883498669f);
  p = fmaf(p, rh, 0.04166646674275398f);
  p = fmaf(p, rh, 0.1666666716337204f);
  p = fmaf(p, rh, 0.5f);
  const float s = fmaf(rh * rh, p, rl);  // r^2 P(r) + rl, small next to rh
  qh = rh + s;
  ql = s - (qh - rh);  // Fast2Sum: |rh| >= |s|
}

// Branch free: the argument is clamped to [-104, 89] (exp(89) overflows to +inf and
// exp(-104) rounds to +0, as the unclamped values would), and NaN is selected at the end.
__device__ __forceinline__ float csmi_expf(float x) {
#pragma clang fp contract(off) reassociate(off)  // the rounding steps are part of the algorithm
  const float xc = fminf(fmaxf(x, -104.0f), 89.0f);
  float n, qh, ql;
  csmi_exp_reduce(xc, n, qh, ql);
  const float t = 1.0f + qh;
  const float e = (1.0f - t) + qh;  // Fast2Sum: |qh| < 1
  const float r = ldexpf(t + (e + ql), static_cast<int>(n));
  return x != x ? x : r;
}

__device__ __forceinline__ float csmi_tanhf(float x) {
#pragma clang fp contract(off) reassociate(off)  // the rounding steps are part of the algorithm
  const float a = fminf(fabsf(x), 9.5f);  // tanh rounds to 1 beyond 9.01; NaN selected below
  float n, qh, ql;
  csmi_exp_reduce(2.0f * a, n, qh, ql);
  const int k = static_cast<int>(n);
  // E = 2^n (1 + q) - 1 = 2^n q + (2^n - 1), as (eh, el)
  const float bh = ldexpf(qh, k), bl = ldexpf(ql, k);
  const float am = ldexpf(1.0f, k) - 1.0f;
  const float eh = am + bh;
  const float bv = eh - am;
  const float el = ((am - (eh - bv)) + (bh - bv)) + bl;  // TwoSum error plus the low part
  // D = E + 2, as (dh, dl)
  const float dh = eh + 2.0f;
  const float dv = dh - eh;
  const float dl = ((eh - (dh - dv)) + (2.0f - dv)) + el;
  // t = E / D with one correction of the residual
  const float inv = __builtin_amdgcn_rcpf(dh);
  const float t0 = eh * inv;
  const float res = fmaf(-t0, dh, eh) + fmaf(-t0, dl, el);
  const float t = fmaf(res, inv, t0);
  return x != x ? x : copysignf(t, x);
}

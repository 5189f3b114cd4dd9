// Scalar device code shared by the interpolation kernels: the quantizer, the
// interpolation stencils and the per-line stencil/rank rules.
//
// Bit-exactness with the reference x86-64 build: no FMA contraction (the
// pragma below and -ffp-contract=off), FP32 denormals preserved, every
// float/double operation in the reference's order.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <stdint.h>

#pragma clang fp contract(off)

namespace faz {

// (int) of a non-negative double the way x86-64 cvttsd2si converts it: values
// beyond INT_MAX (and NaN) become INT_MIN (IntegerQuantizer.hpp:81 relies on it).
__device__ __forceinline__ int32_t x86_trunc_i32(double v) {
    return (v > -2147483649.0 && v < 2147483648.0) ? (int32_t) v : INT32_MIN;
}

// The same for v = |diff| * recip with recip > 0 (v is >= 0 or NaN): one
// comparison sends NaN and everything from 2^31 up to INT_MIN.
__host__ __device__ __forceinline__ int32_t x86_trunc_i32_nonneg(double v) {
    return v < 2147483648.0 ? (int32_t) v : INT32_MIN;
}

// LinearQuantizer::quantize_and_overwrite with trimToZero == 0
// (IntegerQuantizer.hpp:69-108); recip = 1 / eb > 0.  Also built for the host
// (the RegressionPredictor's coefficient chain).  Returns the code; *rec gets the value the
// reference leaves in its data buffer (the original when unpredictable).
__host__ __device__ __forceinline__ int quantize_point(float orig, float pred, double eb, double recip, int radius,
                                              float *rec) {
    float diff = orig - pred;
    uint32_t qi = (uint32_t) x86_trunc_i32_nonneg((double) fabsf(diff) * recip) + 1u;
    if ((int32_t) qi < radius * 2) {
        int32_t q = ((int32_t) qi) >> 1;
        int32_t half = q;
        uint32_t qs = (uint32_t) q << 1;
        int32_t shifted;
        if (diff < 0) {
            qs = 0u - qs;
            shifted = radius - half;
        } else {
            shifted = radius + half;
        }
        float dec = (float) ((double) pred + (double) (int32_t) qs * eb);
        if ((double) fabsf(dec - orig) > eb) {
            *rec = orig;
            return 0;
        }
        *rec = dec;
        return shifted;
    }
    *rec = orig;
    return 0;
}

// quantize_point without branches: every step is computed and the outcome
// selected, so a wavefront never splits on the in-range test.  ebf is the
// largest float not above eb: for a float f, (double) f > eb exactly when
// f > ebf, so the final bound check needs no conversion to double.
__device__ __forceinline__ int quantize_select(float orig, float pred, double eb, double recip, float ebf, int radius,
                                               float *rec) {
    const float diff = orig - pred;
    const int32_t qi = (int32_t) ((uint32_t) x86_trunc_i32_nonneg((double) fabsf(diff) * recip) + 1u);
    const bool in_range = qi < radius * 2;
    const int32_t half = qi >> 1;
    const uint32_t qs = (uint32_t) qi & ~1u;  // (qi >> 1) << 1
    const bool neg = diff < 0;
    const int32_t qsigned = (int32_t) (neg ? 0u - qs : qs);
    const int32_t shifted = neg ? radius - half : radius + half;
    const float dec = (float) ((double) pred + (double) qsigned * eb);
    const bool ok = in_range && !(fabsf(dec - orig) > ebf);
    *rec = ok ? dec : orig;
    return ok ? shifted : 0;
}

// quantize_select when |diff| * recip is known to stay below 2^30 (the host
// bounds it from max |x| and the level's error bound): the truncation never
// overflows, so the x86 out-of-range path is dead, and the code follows from
// the signed quantization index as radius + qsigned / 2.
__device__ __forceinline__ int quantize_bounded(float orig, float pred, double eb, double recip, float ebf,
                                                int radius, float *rec) {
    const float diff = orig - pred;
    const int32_t qi = (int32_t) ((double) fabsf(diff) * recip) + 1;
    const bool in_range = qi < radius * 2;
    const int32_t qs = qi & ~1;  // (qi >> 1) << 1
    const int32_t qsigned = diff < 0 ? -qs : qs;
    const float dec = (float) ((double) pred + (double) qsigned * eb);
    const bool ok = in_range && !(fabsf(dec - orig) > ebf);
    *rec = ok ? dec : orig;
    return ok ? radius + (qsigned >> 1) : 0;
}

// LinearQuantizer::recover (IntegerQuantizer.hpp:160-175)
__host__ __device__ __forceinline__ float recover_point(float pred, int code, int radius, double eb) {
    return (float) ((double) pred + (double) (2 * (code - radius)) * eb);
}

// Interpolators.hpp:9-36, evaluated in float left to right; linear1 in double.
__device__ __forceinline__ float i_linear(float a, float b) { return (a + b) / 2; }
__device__ __forceinline__ float i_linear1(float a, float b) { return (float) (-0.5 * (double) a + 1.5 * (double) b); }
__device__ __forceinline__ float i_quad1(float a, float b, float c) { return (3 * a + 6 * b - c) / 8; }
__device__ __forceinline__ float i_quad2(float a, float b, float c) { return (-a + 6 * b + 3 * c) / 8; }
__device__ __forceinline__ float i_quad3(float a, float b, float c) { return (3 * a - 10 * b + 15 * c) / 8; }
__device__ __forceinline__ float i_cubic(float a, float b, float c, float d) { return (-a + 9 * b + 9 * c - d) / 16; }

// Stencil of block_interpolation_1d (SIC:1306-1417) for the odd local
// position i of a line of n points; nb(k) returns the value at i + k.
template<class F>
__device__ __forceinline__ float stencil_pred(uint32_t i, uint32_t n, int linear, F nb) {
    if (linear || n < 5) {
        if (i + 1 < n) return i_linear(nb(-1), nb(1));
        if (n < 4) return nb(-1);
        return i_linear1(nb(-3), nb(-1));
    }
    const uint32_t ex = 2 * ((n - 3) / 2) + 1;  // loop-exit position (quad_2)
    if (i == 1) return i_quad1(nb(-1), nb(1), nb(3));
    if (i == ex) return i_quad2(nb(-3), nb(-1), nb(1));
    if (i == n - 1) return i_quad3(nb(-5), nb(-3), nb(-1));
    return i_cubic(nb(-3), nb(-1), nb(1), nb(3));
}

// Emission rank of odd position i within its line (same rules): interior
// cubic points first, then i = 1, the loop-exit point and, for even n, n - 1.
__device__ __forceinline__ uint32_t line_rank(uint32_t i, uint32_t n, int linear) {
    if (linear || n < 5) return (i - 1) / 2;
    const uint32_t m = (n - 3) / 2 - 1;
    if (i == 1) return m;
    if (i == 3 + 2 * m) return m + 1;
    if (i == n - 1) return m + 2;
    return (i - 3) / 2;
}

}  // namespace faz

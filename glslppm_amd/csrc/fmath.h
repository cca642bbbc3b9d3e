// Strict fp32 arithmetic of the GLSL passes, for gfx950.
//
// Every expression keeps the operation order the GLSL source writes (left to right, no
// contraction: the library is built with -ffp-contract=off), division and sqrt are the
// correctly rounded IEEE operations, min/max reject NaN.  sin, cos, exp, log and pow are our own
// single-precision routines after the published Cephes library (S. L. Moshier, "Methods and
// Programs for Mathematical Functions", sinf.c/cosf.c/expf.c/logf.c): Cody-Waite range reduction
// and the Cephes minimax polynomials, so that a host build of the same steps yields the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PPM_DEV __device__ __forceinline__

struct f3 {
    float x, y, z;
};

PPM_DEV f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
PPM_DEV f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
PPM_DEV f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
PPM_DEV f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PPM_DEV f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
PPM_DEV f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
PPM_DEV f3 operator/(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
// scalar * vector and vector * scalar are the same componentwise products
PPM_DEV f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
PPM_DEV f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
PPM_DEV f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }

// dot(a, b) = ((a.x*b.x + a.y*b.y) + a.z*b.z)
PPM_DEV float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PPM_DEV f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PPM_DEV float length3(f3 a) { return sqrtf(dot3(a, a)); }
// normalize(v) = v * (1 / sqrt(dot(v, v)))
PPM_DEV f3 normalize3(f3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }
// reflect(I, N) = I - (2 dot(N, I)) N
PPM_DEV f3 reflect3(f3 i, f3 n) { return i - (2.0f * dot3(n, i)) * n; }
// refract(I, N, eta): GLSL 1.20 definition
PPM_DEV f3 refract3(f3 i, f3 n, float eta) {
    const float d = dot3(n, i);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return mk3(0.0f, 0.0f, 0.0f);
    return eta * i - (eta * d + sqrtf(k)) * n;
}
PPM_DEV float fract1(float x) { return x - floorf(x); }
PPM_DEV float sign1(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
PPM_DEV float min1(float a, float b) { return fminf(a, b); }
PPM_DEV float max1(float a, float b) { return fmaxf(a, b); }

// ---- Cephes-style single precision transcendentals ----

PPM_DEV float scale_pow2(float p, int n) {
    if (n > 127) {
        p = p * 1.7014118346046923e38f;  // 2^127
        n -= 127;
    }
    if (n < -126) {
        p = p * 1.1754943508222875e-38f;  // 2^-126
        n += 126;
    }
    if (n < -126) {
        p = p * 1.1754943508222875e-38f;
        n += 126;
    }
    return p * __int_as_float((n + 127) << 23);
}

// Shared reduction of sin and cos: x >= 0 reduced by j * pi/4.
PPM_DEV float trig_reduce(float x, int& j) {
    const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
    j = (int)(1.27323954473516f * x);
    float y = (float)j;
    if (j & 1) {
        j += 1;
        y = y + 1.0f;
    }
    j &= 7;
    return ((x - y * DP1) - y * DP2) - y * DP3;
}
PPM_DEV float poly_sin(float z) {
    const float zz = z * z;
    return ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
}
PPM_DEV float poly_cos(float z) {
    const float zz = z * z;
    return ((((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz) * zz -
            0.5f * zz) + 1.0f;
}
PPM_DEV float ppm_sinf(float x) {
    int neg = 0;
    if (x < 0.0f) {
        x = -x;
        neg = 1;
    }
    int j;
    const float z = trig_reduce(x, j);
    if (j > 3) {
        neg ^= 1;
        j -= 4;
    }
    const float r = (j == 1 || j == 2) ? poly_cos(z) : poly_sin(z);
    return neg ? -r : r;
}
PPM_DEV float ppm_cosf(float x) {
    if (x < 0.0f) x = -x;
    int j;
    const float z = trig_reduce(x, j);
    int neg = 0;
    if (j > 3) {
        neg ^= 1;
        j -= 4;
    }
    if (j > 1) neg ^= 1;
    const float r = (j == 1 || j == 2) ? poly_sin(z) : poly_cos(z);
    return neg ? -r : r;
}
// ppm_sinf(x) and ppm_cosf(x) from one reduction and one evaluation of each polynomial: the same
// operations, so the same bits, at about half the cost of the two calls.
PPM_DEV void ppm_sincosf(float x, float& s, float& c) {
    int sneg = 0;
    if (x < 0.0f) {
        x = -x;
        sneg = 1;
    }
    int j;
    const float z = trig_reduce(x, j);
    int cneg = 0;
    if (j > 3) {
        sneg ^= 1;
        cneg ^= 1;
        j -= 4;
    }
    if (j > 1) cneg ^= 1;
    const float ps = poly_sin(z), pc = poly_cos(z);
    const bool swap = j == 1 || j == 2;
    const float rs = swap ? pc : ps, rc = swap ? ps : pc;
    s = sneg ? -rs : rs;
    c = cneg ? -rc : rc;
}
PPM_DEV float ppm_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283905206835f) return __int_as_float(0x7f800000);
    if (x < -103.278929903431851103f) return 0.0f;
    const float z = floorf(1.44269504088896341f * x + 0.5f);
    x = (x - z * 0.693359375f) - z * -2.12194440e-4f;
    const int n = (int)z;
    const float zz = x * x;
    const float p = ((((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x +
                        4.1665795894e-2f) * x + 1.6666665459e-1f) * x + 5.0000001201e-1f) * zz + x) + 1.0f;
    return scale_pow2(p, n);
}
PPM_DEV float ppm_logf(float x) {
    if (x != x) return x;
    if (x <= 0.0f) return x == 0.0f ? __int_as_float(0xff800000) : __int_as_float(0x7fc00000);
    if (x == __int_as_float(0x7f800000)) return x;
    int e = 0;
    if (x < 1.1754943508222875e-38f) {  // subnormal: scale into the normal range
        x = x * 33554432.0f;            // 2^25
        e = -25;
    }
    const int b = __float_as_int(x);
    e += ((b >> 23) & 0xff) - 126;
    float m = __int_as_float((b & 0x807fffff) | 0x3f000000);  // [0.5, 1)
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    const float z = m * m;
    float y = (((((((((7.0376836292e-2f * m - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m -
                     1.2420140846e-1f) * m + 1.4249322787e-1f) * m - 1.6668057665e-1f) * m +
                  2.0000714765e-1f) * m - 2.4999993993e-1f) * m + 3.3333331174e-1f) * m) * z;
    const float fe = (float)e;
    y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    float r = m + y;
    r = r + 0.693359375f * fe;
    return r;
}
// pow(x, y) = exp(y log x) for x > 0; x == 0 gives 0 for y > 0.
PPM_DEV float ppm_powf(float x, float y) {
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : __int_as_float(0x7f800000);
    if (x < 0.0f) return __int_as_float(0x7fc00000);
    return ppm_expf(y * ppm_logf(x));
}

// ---- the MLCG of GPURnd and hash() (eyeraytrace.fs:326-347, hash.fs:21-48) ----

struct Mlcg {
    static constexpr float q0 = 1225.0f, q1 = 1585.0f, q2 = 2457.0f, q3 = 2098.0f;
    static constexpr float r0 = 1112.0f, r1 = 367.0f, r2 = 92.0f, r3 = 265.0f;
    static constexpr float a0 = 3423.0f, a1 = 2646.0f, a2 = 1707.0f, a3 = 1999.0f;
    static constexpr float m0 = 4194287.0f, m1 = 4194277.0f, m2 = 4194191.0f, m3 = 4194167.0f;
};

// One Schrage step of one lane: beta = floor(n/q); p = a (n - beta q) - beta r; p + (p<0 ? m : 0)
PPM_DEV float mlcg_lane(float n, float qi, float q, float r, float a, float m) {
    const float beta = floorf(n * qi);
    const float p = a * (n - beta * q) - beta * r;
    const float m2 = 0.5f * m;
    return p + (sign1(-p) * m2 + m2);
}

// GPURnd: advances the four lanes in place, returns fract(n0/m0 - n1/m1 + n2/m2 - n3/m3).
PPM_DEV float gpu_rnd(float4& n) {
    n.x = mlcg_lane(n.x, 1.0f / 1225.0f, 1225.0f, 1112.0f, 3423.0f, 4194287.0f);
    n.y = mlcg_lane(n.y, 1.0f / 1585.0f, 1585.0f, 367.0f, 2646.0f, 4194277.0f);
    n.z = mlcg_lane(n.z, 1.0f / 2457.0f, 2457.0f, 92.0f, 1707.0f, 4194191.0f);
    n.w = mlcg_lane(n.w, 1.0f / 2098.0f, 2098.0f, 265.0f, 1999.0f, 4194167.0f);
    const float s = ((n.x * (1.0f / 4194287.0f) - n.y * (1.0f / 4194277.0f)) + n.z * (1.0f / 4194191.0f)) -
                    n.w * (1.0f / 4194167.0f);
    return fract1(s);
}

// hash(): the cell through the same generator, scaled to H slots; returns the float slot number.
PPM_DEV float cell_hash(f3 idx, float grid_scale, float hash_scale, float hash_num) {
    float4 n = make_float4(idx.x * hash_scale, idx.y * hash_scale, idx.z * hash_scale,
                           (grid_scale * 0.5f) * hash_scale);
    n.x = mlcg_lane(n.x, 1.0f / 1225.0f, 1225.0f, 1112.0f, 3423.0f, 4194287.0f);
    n.y = mlcg_lane(n.y, 1.0f / 1585.0f, 1585.0f, 367.0f, 2646.0f, 4194277.0f);
    n.z = mlcg_lane(n.z, 1.0f / 2457.0f, 2457.0f, 92.0f, 1707.0f, 4194191.0f);
    n.w = mlcg_lane(n.w, 1.0f / 2098.0f, 2098.0f, 265.0f, 1999.0f, 4194167.0f);
    const float s = ((n.x * (1.0f / 4194287.0f) - n.y * (1.0f / 4194277.0f)) + n.z * (1.0f / 4194191.0f)) -
                    n.w * (1.0f / 4194167.0f);
    return floorf(fract1(s) * hash_num);
}

// Device math of the DFSPH liquid step for gfx950: smoothing kernels, the half-space wall volume, Newton's box /
// sphere / plane SDFs and transforms. Every formula keeps the operation order of the reference's Warp function it
// restates (cited per function) and the build turns FMA contraction off, so each result is the correctly rounded
// float32 sequence of those operations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfsph {

constexpr int ACTIVE = 1;
constexpr int GEO_PLANE = 1, GEO_SPHERE = 3, GEO_BOX = 7;
constexpr int CHUNK = 256;

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }  // correctly rounded (not v_sqrt_f32 alone)
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fcl(float v, float lo, float hi) { return fmn(fmx(v, lo), hi); }
__device__ __forceinline__ f3 normalize(f3 a) {
    float l = length(a);
    if (l > 0.0f) return a / l;
    return f3{0.0f, 0.0f, 0.0f};
}
__device__ __forceinline__ f3 ld3(const float* p, int64_t k) { return f3{p[3 * k], p[3 * k + 1], p[3 * k + 2]}; }
__device__ __forceinline__ void st3(float* p, int64_t k, f3 v) {
    p[3 * k] = v.x;
    p[3 * k + 1] = v.y;
    p[3 * k + 2] = v.z;
}

constexpr float PI_F = 3.14159265358979323846f;

// ------------------------------------------------------------------------------ correctly rounded pair arithmetic
// The pair loops spend most of their instructions on one square root and two divisions per pair, which the compiler
// expands for the whole float range (input scaling for tiny values, special-case fix-ups). For the operand ranges
// of a smoothing length h in [PAIR_H_MIN, PAIR_H_MAX] the shorter sequences below give the same correctly rounded
// results (the bounds keep every operand, quotient and residual normal and the quotient-to-divisor exponent gap
// under the 96 at which the hardware's division pre-scaling begins):
// - sqrt_rn: the hardware estimate (within 1 ulp) replaced by a neighbour when an exact residual of the midpoint
//   test says so; exact for x = 0 and x >= 2^-96 (a pair closer than 2^-48 m is inside the 1e-9 h cut-off anyway);
// - div_h: r / h from the correctly rounded 1 / h with two remainder corrections (Markstein's theorem for the
//   second, the first bringing the quotient within 1 ulp);
// - div_rn: a / d from the hardware reciprocal with one Newton step and two remainder corrections, the compiler's
//   own division sequence without its range scaling and fix-up, which change nothing in this range.
// tests/pair_math_check.c checks the three sequences against IEEE results on the CPU.
constexpr float PAIR_H_MIN = 1.0e-5f, PAIR_H_MAX = 100.0f;
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s0 = __builtin_amdgcn_sqrtf(x);
    const float dn = __int_as_float(__float_as_int(s0) - 1), up = __int_as_float(__float_as_int(s0) + 1);
    const float vp = fmaf(-dn, s0, x), vs = fmaf(-up, s0, x);
    const float s = vp <= 0.0f ? dn : s0;
    return vs > 0.0f ? up : s;
}
__device__ __forceinline__ float div_rn(float a, float d) {
    const float y0 = __builtin_amdgcn_rcpf(d);
    const float y = fmaf(fmaf(-d, y0, 1.0f), y0, y0);
    const float q0 = a * y;
    const float q1 = fmaf(fmaf(-d, q0, a), y, q0);
    return fmaf(fmaf(-d, q1, a), y, q1);
}
// The smoothing kernel's launch constants, built once per thread from the same float expressions the per-call
// forms below evaluate (h^3, the kind's normalisation, the 1e-9 h cut-off), so both give identical bits; `fast`
// selects the short pair arithmetic above (a launch-uniform branch).
struct KC {
    float h, rmin, knorm, inv_h;
    int kind, fast;
};
__device__ __forceinline__ KC make_kc(float h, int kind) {
    float h3 = h * h * h;
    KC K;
    K.h = h;
    K.kind = kind;
    K.rmin = 1.0e-9f * h;
    K.knorm = kind == 1 ? (21.0f / (2.0f * PI_F * h3)) : 8.0f / (PI_F * h3);
    K.inv_h = 1.0f / h;
    K.fast = h >= PAIR_H_MIN && h <= PAIR_H_MAX;
    return K;
}
__device__ __forceinline__ float div_h(float r, const KC& K) {
    const float q0 = r * K.inv_h;
    const float q1 = fmaf(fmaf(-K.h, q0, r), K.inv_h, q0);
    return fmaf(fmaf(-K.h, q1, r), K.inv_h, q1);
}
// |rv| and r / h, correctly rounded
__device__ __forceinline__ float dist(f3 rv, const KC& K) {
    if (K.fast) return sqrt_rn(dot(rv, rv));
    return length(rv);
}
__device__ __forceinline__ float over_h(float r, const KC& K) {
    if (K.fast) return div_h(r, K);
    return r / K.h;
}

// kernels.py:41-55
__device__ __forceinline__ float kernel_w(float r, const KC& K) {
    float q = over_h(r, K);
    if (q >= 1.0f) return 0.0f;
    if (K.kind == 1) {
        float s = 1.0f - q;
        return K.knorm * s * s * s * s * (1.0f + 4.0f * q);
    }
    if (q <= 0.5f) return K.knorm * (6.0f * q * q * q - 6.0f * q * q + 1.0f);
    float s = 1.0f - q;
    return K.knorm * 2.0f * s * s * s;
}

// kernels.py:58-77 as a scale: grad W(rv) = rv * dW/dq / (r h), with 0 standing for the zero vector of the cut-off
// cases (q >= 1 or r <= 1e-9 h). dW/dq vanishes inside (0, 1) nowhere, so a zero scale occurs only there.
// KIND: the kernel kind when the caller knows it (0 cubic, 1 Wendland), -1 to read K.kind
template <int KIND>
__device__ __forceinline__ float kernel_grad_scale_fast_kind(f3 rv, const KC& K) {
    // every branch evaluated and the result selected: no divergent control flow per pair
    const float r = sqrt_rn(dot(rv, rv));
    const float q = div_h(r, K);
    const float s = 1.0f - q;
    float dwdq;
    if (KIND == 1 || (KIND < 0 && K.kind == 1)) {
        dwdq = K.knorm * (-20.0f * q * s * s * s);
    } else {
        const float near = K.knorm * (18.0f * q * q - 12.0f * q);
        const float far = K.knorm * (-6.0f * s * s);
        dwdq = q <= 0.5f ? near : far;
    }
    const float g = div_rn(dwdq, r * K.h);
    return (q >= 1.0f || r <= K.rmin) ? 0.0f : g;
}
__device__ __forceinline__ float kernel_grad_scale_fast(f3 rv, const KC& K) {
    return kernel_grad_scale_fast_kind<-1>(rv, K);
}
__device__ __forceinline__ float kernel_grad_scale(f3 rv, const KC& K) {
    if (K.fast) return kernel_grad_scale_fast(rv, K);
    float r = length(rv);
    float q = r / K.h;
    if (q >= 1.0f || r <= K.rmin) return 0.0f;
    float dwdq;
    if (K.kind == 1) {
        float s = 1.0f - q;
        dwdq = K.knorm * (-20.0f * q * s * s * s);
    } else if (q <= 0.5f) {
        dwdq = K.knorm * (18.0f * q * q - 12.0f * q);
    } else {
        float s = 1.0f - q;
        dwdq = K.knorm * (-6.0f * s * s);
    }
    return dwdq / (r * K.h);
}
// kernel_w(dist(rv, K), K) and kernel_grad_scale(rv, K) with the variant known to the caller (KV 0: the cubic kernel
// at a regular h, K.fast and K.kind 0; KV 1: any), so a neighbour loop chooses once instead of per pair; the same
// operations either way
template <int KV>
__device__ __forceinline__ float kernel_w_rv(f3 rv, const KC& K) {
    if (KV == 0) {
        const float q = div_h(sqrt_rn(dot(rv, rv)), K);
        if (q >= 1.0f) return 0.0f;
        if (q <= 0.5f) return K.knorm * (6.0f * q * q * q - 6.0f * q * q + 1.0f);
        const float s = 1.0f - q;
        return K.knorm * 2.0f * s * s * s;
    }
    return kernel_w(dist(rv, K), K);
}
template <int KV>
__device__ __forceinline__ float kernel_grad_scale_v(f3 rv, const KC& K) {
    if (KV == 0) return kernel_grad_scale_fast_kind<0>(rv, K);
    return kernel_grad_scale(rv, K);
}
// the gradient from its scale: the same products kernel_grad forms, the exact zero vector for a zero scale
__device__ __forceinline__ f3 pair_grad(f3 rv, float s) { return s == 0.0f ? f3{0.0f, 0.0f, 0.0f} : rv * s; }
__device__ __forceinline__ f3 kernel_grad(f3 rv, const KC& K) { return pair_grad(rv, kernel_grad_scale(rv, K)); }

// kernels.py:41-55
__device__ __forceinline__ float kernel_w(float r, float h, int kind) {
    float q = r / h;
    if (q >= 1.0f) return 0.0f;
    float h3 = h * h * h;
    if (kind == 1) {
        float s = 1.0f - q;
        return (21.0f / (2.0f * PI_F * h3)) * s * s * s * s * (1.0f + 4.0f * q);
    }
    float k = 8.0f / (PI_F * h3);
    if (q <= 0.5f) return k * (6.0f * q * q * q - 6.0f * q * q + 1.0f);
    float s = 1.0f - q;
    return k * 2.0f * s * s * s;
}

// kernels.py:58-77
__device__ __forceinline__ f3 kernel_grad(f3 rv, float h, int kind) {
    float r = length(rv);
    float q = r / h;
    if (q >= 1.0f || r <= 1.0e-9f * h) return f3{0.0f, 0.0f, 0.0f};
    float h3 = h * h * h;
    float dwdq;
    if (kind == 1) {
        float s = 1.0f - q;
        dwdq = (21.0f / (2.0f * PI_F * h3)) * (-20.0f * q * s * s * s);
    } else {
        float k = 8.0f / (PI_F * h3);
        if (q <= 0.5f) {
            dwdq = k * (18.0f * q * q - 12.0f * q);
        } else {
            float s = 1.0f - q;
            dwdq = k * (-6.0f * s * s);
        }
    }
    return rv * (dwdq / (r * h));
}

// kernels.py:80-112
__device__ __forceinline__ float projected_p(float q, int kind) {
    if (q >= 1.0f) return 0.0f;
    float q2 = q * q;
    if (kind == 1) return 1.5f + q2 * (-10.5f + q2 * (52.5f + q * (-84.0f + q * (52.5f - 12.0f * q))));
    if (q <= 0.5f) return 1.4f + q2 * (-8.0f + q2 * (24.0f - 19.2f * q));
    return 1.6f + q2 * (-16.0f + q * (32.0f + q * (-24.0f + 6.4f * q)));
}

__device__ __forceinline__ float g0(float y, int kind) {
    float y2 = y * y;
    if (kind == 1) return y * (1.5f + y2 * (-3.5f + y2 * (10.5f + y * (-14.0f + y * (7.5f - 1.5f * y)))));
    if (y <= 0.5f) return y * (1.4f + y2 * (-8.0f / 3.0f + y2 * (4.8f - 3.2f * y)));
    return -1.0f / 30.0f + y * (1.6f + y2 * (-16.0f / 3.0f + y * (8.0f + y * (-4.8f + y * (16.0f / 15.0f)))));
}

__device__ __forceinline__ float g1(float y, int kind) {
    float y2 = y * y;
    if (kind == 1)
        return y2 * (0.75f + y2 * (-2.625f + y2 * (8.75f + y * (-12.0f + y * (6.5625f - (4.0f / 3.0f) * y)))));
    if (y <= 0.5f) return y2 * (0.7f + y2 * (-2.0f + y2 * (4.0f - (19.2f / 7.0f) * y)));
    return -1.0f / 280.0f + y2 * (0.8f + y2 * (-4.0f + y * (6.4f + y * (-4.0f + y * (6.4f / 7.0f)))));
}

// (Psi, dPsi/dd) of a solid half-space at signed distance d, mean curvature H (kernels.py:122-138)
__device__ __forceinline__ void halfspace_volume(float d, float h, float H, int kind, float& psi, float& dpsi) {
    if (d >= h) {
        psi = 0.0f;
        dpsi = 0.0f;
        return;
    }
    float z = fcl(-d / h, -1.0f, 1.0f);
    float y = fabsf(z);
    float sgn = 1.0f;
    if (z < 0.0f) sgn = -1.0f;
    float g1_full = g1(1.0f, kind);
    psi = 0.5f + sgn * g0(y, kind) - 0.5f * H * h * (g1(y, kind) - g1_full);
    dpsi = 0.0f;
    if (d > -h) dpsi = -projected_p(y, kind) * (1.0f - 0.5f * H * h * z) / h;
}

// ----------------------------------------------------------------------------------------- transforms (Warp)
struct Xf {
    f3 p;
    float qx, qy, qz, qw;
};

__device__ __forceinline__ f3 quat_rotate(float qx, float qy, float qz, float qw, f3 v) {
    f3 u{qx, qy, qz};
    f3 c = cross(u, v);
    return v * (2.0f * qw * qw - 1.0f) + c * qw * 2.0f + u * dot(u, v) * 2.0f;
}
__device__ __forceinline__ f3 xf_point(const Xf& X, f3 v) { return quat_rotate(X.qx, X.qy, X.qz, X.qw, v) + X.p; }
__device__ __forceinline__ f3 xf_vector(const Xf& X, f3 v) { return quat_rotate(X.qx, X.qy, X.qz, X.qw, v); }
__device__ __forceinline__ f3 xf_inv_point(const Xf& X, f3 v) {
    return quat_rotate(-X.qx, -X.qy, -X.qz, X.qw, v) + (-quat_rotate(-X.qx, -X.qy, -X.qz, X.qw, X.p));
}
__device__ __forceinline__ Xf load_xf(const float* x7) {
    return Xf{f3{x7[0], x7[1], x7[2]}, x7[3], x7[4], x7[5], x7[6]};
}

__device__ __forceinline__ float sgn1(float v) { return v < 0.0f ? -1.0f : 1.0f; }

// Newton's analytic SDFs: sphere/box (alloy/_vendor/newton/_src/geometry/kernels.py:277-370), infinite plane
// (soft_contacts_sdf.py:73-84); (phi, unit outward gradient)
__device__ __forceinline__ void shape_sdf(int geo, f3 s, f3 p, float& phi, f3& grad) {
    if (geo == GEO_SPHERE) {
        phi = length(p) - s.x;
        float l = length(p);
        grad = l > 1.0e-8f ? p / l : f3{0.0f, 0.0f, 1.0f};
        return;
    }
    if (geo == GEO_PLANE) {
        phi = p.z;
        grad = f3{0.0f, 0.0f, 1.0f};
        return;
    }
    float qx = fabsf(p.x) - s.x, qy = fabsf(p.y) - s.y, qz = fabsf(p.z) - s.z;
    f3 e{fmx(qx, 0.0f), fmx(qy, 0.0f), fmx(qz, 0.0f)};
    phi = length(e) + fmn(fmx(qx, fmx(qy, qz)), 0.0f);
    if (qx > 0.0f || qy > 0.0f || qz > 0.0f) {
        f3 c{fcl(p.x, -s.x, s.x), fcl(p.y, -s.y, s.y), fcl(p.z, -s.z, s.z)};
        grad = normalize(p - c);
        return;
    }
    if ((qx > qy && qx > qz) || (qy == 0.0f && qz == 0.0f)) {
        grad = f3{sgn1(p.x), 0.0f, 0.0f};
        return;
    }
    if ((qy > qx && qy > qz) || (qx == 0.0f && qz == 0.0f)) {
        grad = f3{0.0f, sgn1(p.y), 0.0f};
        return;
    }
    grad = f3{0.0f, 0.0f, sgn1(p.z)};
}

}  // namespace dfsph

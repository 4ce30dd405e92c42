// render.hip -- HIP kernels for gfx950 (MI355X): BIRP's per-pixel ray generation, fixed-step
// march through the emissivity cube (or the analytic CMEM model) and per-ray integration
// (reference Camera.cpp:140-229, DataCube.cpp:203-277, Cmem.cpp:117-288), plus the layout
// builders the renders read from.
//
// Two arithmetic families:
//  * exact  -- FP32 in the reference's operation order (no FMA contraction: the file is built
//              with -ffp-contract=off), correctly rounded divisions via birp_div_rn, the double
//              scale chain via birp_scale_sample, std::lerp's branch structure, and the image sum
//              in k order. Bit-identical to the reference CPU render for finite cubes.
//  * fast   -- FP64 camera basis, ray origin and cube entry; FP32 FMA march in index space;
//              BF16/F16/FP32 storage; FP32 Kahan sums. Batched in a persistent work-queue kernel.
//
// Wave mapping: a wave64 owns a tile of 64/S rays; S lanes ("sub-lanes") share one ray and take
// its samples k = kb + sub, kb advancing by S each trip. In the exact kernels every sub-lane then
// reads its quad-mates' samples through DPP quad_perm operands and all of them add the S samples
// in k order, so the sequential FP32 sum of the reference is kept while S lanes sample in
// parallel. A masked-off sample contributes -0.0f, which leaves every partial sum unchanged.
//
// Memory safety: cube records are read with buffer loads whose descriptor carries the buffer's
// byte size, so the hardware returns 0 for any offset past the end; indices are clamped before
// the tables are read.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <atomic>

#include "../../include/birp.h"
#include "birp_internal.h"
#include "exact_math.h"

namespace birp {

#define BIRP_BLOCK 256
#define BIRP_MAX_EXACT_SAMPLES 8192
#define BIRP_FAST_TAB 2048  // largest per-wave LDS table of the fast kernel (nx+ny+nz entries)
#define BIRP_RSRC_WORD3 0x00020000

// ------------------------------------------------------------------------------------------
// small device helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ unsigned wave_sum_u(unsigned v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// DPP row_shl:N: lane l receives lane l + N of its 16-lane row (0 past the row's end). With
// the exact kernels' lane layout, sub-lane 0 of a ray reads its sub-lane m at N = 4m.
template <int N>
__device__ __forceinline__ int row_next(int v) {
    if constexpr (N == 0) return v;
    else return __builtin_amdgcn_update_dpp(0, v, 0x100 + N, 0xF, 0xF, true);
}
template <int N>
__device__ __forceinline__ float row_next_f(float v) {
    return __int_as_float(row_next<N>(__float_as_int(v)));
}

// std::lerp (libstdc++ 11, cmath:1894-1908) for any finite or non-finite input.
__device__ __forceinline__ float lerp_general(float a, float b, float t) {
    bool straddle = (a <= 0 && b >= 0) || (a >= 0 && b <= 0);
    float s1 = t * b + (1 - t) * a;
    float x = a + t * (b - a);
    float r = ((t > 1) == (b > a)) ? (b < x ? x : b) : (b > x ? x : b);
    r = (t == 1) ? b : r;
    return straddle ? s1 : r;
}

// std::lerp when a, b > 0 and finite and 0 <= t < 1: a + t(b-a) clamped towards b, which is
// the median of (x, a, b) since x never passes a.
__device__ __forceinline__ float lerp_positive(float a, float b, float t) {
    float x = a + t * (b - a);
    return __builtin_amdgcn_fmed3f(x, a, b);
}

// Clamp in the float domain to [0, hi], then convert (NaN -> 0 in hardware).
__device__ __forceinline__ int to_index(float x, float hi) { return (int)__builtin_amdgcn_fmed3f(x, 0.0f, hi); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, BIRP_RSRC_WORD3);
}
__device__ __forceinline__ float4 load16(__amdgpu_buffer_rsrc_t r, unsigned off) {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    float4 f;
    __builtin_memcpy(&f, &v, 16);
    return f;
}
__device__ __forceinline__ uint4 load16u(__amdgpu_buffer_rsrc_t r, unsigned off) {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    uint4 f;
    __builtin_memcpy(&f, &v, 16);
    return f;
}

// ------------------------------------------------------------------------------------------
// exact per-pixel ray (Camera.cpp:148-168), FP32 IEEE operations in the reference's order
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void exact_ray_dir(const ExactView& ev, int i, int j, float& dx, float& dy, float& dz) {
    float device_x = ((float)i + 0.5f) / (float)ev.W;
    float device_y = ((float)j + 0.5f) / (float)ev.H;
    float screen_x = ((2 * device_x) - 1) * ev.tanx * ev.aspect;
    float screen_y = (1 - (2 * device_y)) * ev.tany;
    float nd = sqrtf(screen_x * screen_x + screen_y * screen_y + (-1.0f) * (-1.0f));
    float ux = screen_x / nd, uy = screen_y / nd, uz = -1.0f / nd;
    float w0 = (ev.rinv[0] * ux) + (ev.rinv[1] * uy) + (ev.rinv[2] * uz);
    float w1 = (ev.rinv[3] * ux) + (ev.rinv[4] * uy) + (ev.rinv[5] * uz);
    float w2 = (ev.rinv[6] * ux) + (ev.rinv[7] * uy) + (ev.rinv[8] * uz);
    float wd = sqrtf(w0 * w0 + w1 * w1 + w2 * w2);
    dx = w0 / wd;
    dy = w1 / wd;
    dz = w2 / wd;
}

// Index of the sample point on each axis: RN(RN(RN(d t) + p) - o) / s), exactly.
struct ExactIdx {
    float xi, yi, zi;  // ingress
    float xd, yd, zd;  // index
};

// GUARD: the double-product division is exact while the quotient is a normal float, which
// holds for |ingress| >= 2^-64 (spacing <= 2^30); below that (possible only when an origin
// component is within 2^-30 of 0) the IEEE division is used.
template <bool GUARD>
__device__ __forceinline__ ExactIdx exact_index(const DevVolume& v, float dx, float dy, float dz, float px, float py,
                                                float pz, float t) {
    ExactIdx r;
    float x = dx * t + px;
    float y = dy * t + py;
    float z = dz * t + pz;
    r.xi = x - v.ox;
    r.yi = y - v.oy;
    r.zi = z - v.oz;
    if constexpr (GUARD) {
        float m = fminf(fminf(fabsf(r.xi), fabsf(r.yi)), fabsf(r.zi));
        if (__builtin_expect(m < 0x1p-64f && m != 0.0f, 0)) {
            r.xd = r.xi / v.sx;
            r.yd = r.yi / v.sy;
            r.zd = r.zi / v.sz;
            return r;
        }
    }
    r.xd = birp_div_rn64(r.xi, v.rdx);
    r.yd = birp_div_rn64(r.yi, v.rdy);
    r.zd = birp_div_rn64(r.zi, v.rdz);
    return r;
}

// DataCube.cpp:214-218 negated: inside iff ingress >= 0 and index + 1 < n on every axis.
// (NaN compares false here and counts as outside; the reference would index out of range.)
__device__ __forceinline__ bool exact_inside(const DevVolume& v, const ExactIdx& r) {
    return (r.xi >= 0) && (r.xd + 1 < v.fnx) && (r.yi >= 0) && (r.yd + 1 < v.fny) && (r.zi >= 0) &&
           (r.zd + 1 < v.fnz);
}

// Trilinear sample from quad records (finite cubes) through the per-axis offset tables. Corner
// x1 = x0 + 1 always: for finite data std::lerp at t == 0 returns a, so this equals DataCube's
// ceil() corner choice.
struct QuadCtx {
    __amdgpu_buffer_rsrc_t rs;
    const unsigned* tx;
    const unsigned* ty;
    const unsigned* tz;
    float hx, hy, hz;  // n - 2 per axis
};

// For an in-cube sample the index is >= 0, so the lower corner floor(x) is the truncating
// conversion of the clamped index (the clamp only matters to lanes outside the cube, whose
// result is discarded) and the weight x - floor(x) is v_fract_f32 (exact: the subtraction of
// floor(x) from x >= 0 is exact, so fract's clamp below 1.0 never applies).
struct QuadFetch {
    float4 a, b;       // the z and z+1 quad records
    float fx, fy, fz;  // lerp weights
};

__device__ __forceinline__ QuadFetch quad_fetch(const QuadCtx& c, float xd, float yd, float zd) {
    QuadFetch q;
    q.fx = __builtin_amdgcn_fractf(xd);
    q.fy = __builtin_amdgcn_fractf(yd);
    q.fz = __builtin_amdgcn_fractf(zd);
    unsigned off = c.tx[to_index(xd, c.hx)] + c.ty[to_index(yd, c.hy)] + c.tz[to_index(zd, c.hz)];
    q.a = load16(c.rs, off);
    q.b = load16(c.rs, off + 1024u);
    return q;
}

// For a sample that passed the reference's bounds test (0 <= ingress, index + 1 < n): the index
// is in [0, n-1), so its truncation is the lower corner and needs no clamp.
__device__ __forceinline__ QuadFetch quad_fetch_inside(const QuadCtx& c, float xd, float yd, float zd) {
    QuadFetch q;
    q.fx = __builtin_amdgcn_fractf(xd);
    q.fy = __builtin_amdgcn_fractf(yd);
    q.fz = __builtin_amdgcn_fractf(zd);
    unsigned off = c.tx[(int)xd] + c.ty[(int)yd] + c.tz[(int)zd];
    q.a = load16(c.rs, off);
    q.b = load16(c.rs, off + 1024u);
    return q;
}

// x first, then y, then z (DataCube.cpp:260-270).
template <bool POS>
__device__ __forceinline__ float quad_lerp(const QuadFetch& q) {
    const float4 a = q.a, b = q.b;
    if constexpr (POS) {
        float ly = lerp_positive(lerp_positive(a.x, a.y, q.fx), lerp_positive(a.z, a.w, q.fx), q.fy);
        float hy = lerp_positive(lerp_positive(b.x, b.y, q.fx), lerp_positive(b.z, b.w, q.fx), q.fy);
        return lerp_positive(ly, hy, q.fz);
    } else {
        float ly = lerp_general(lerp_general(a.x, a.y, q.fx), lerp_general(a.z, a.w, q.fx), q.fy);
        float hy = lerp_general(lerp_general(b.x, b.y, q.fx), lerp_general(b.z, b.w, q.fx), q.fy);
        return lerp_general(ly, hy, q.fz);
    }
}

template <bool POS>
__device__ __forceinline__ float sample_quad(const QuadCtx& c, float xd, float yd, float zd) {
    return quad_lerp<POS>(quad_fetch(c, xd, yd, zd));
}

// The scale chain of Camera.cpp:201 for a positive cube whose scaled samples stay normal floats:
// the folded product, unless it lies within 16 double ulps of a float rounding midpoint.
// The chain is taken by about one sample in 10^7: it must stay a branch (if-converted, its two
// FP64 divisions would run for every sample).
// The test |m29 - 2^28| <= 16 on the low 29 mantissa bits m29 is done on m29 * 8 (the low word
// shifted left by 3): m29 * 8 + 2^31 + 128 wraps into [0, 256] exactly when it holds.
__device__ __forceinline__ float scale_pos(float vw) {
    double y = (double)vw * BIRP_SCALE_FOLDED;
    const uint32_t lo = (uint32_t)birp_f64_bits(y);
    float s = (float)y;
    if (__builtin_expect((lo << 3) + 0x80000080u < 257u, 0)) s = birp_scale_chain(vw);
    return s;
}

// acc += the S sub-lane samples of this ray, in sub-lane (= k) order. Exact on sub-lane 0 of
// the ray (the lane that stores the pixel); the other sub-lanes' sums are not used. With 8 lanes
// per ray the ray spans two 16-lane rows: sub-lanes 4..7 are read from the next row's copy.
template <int S>
__device__ __forceinline__ float add_in_order(float acc, float s) {
    acc = acc + s;
    if constexpr (S >= 2) acc = acc + row_next_f<4>(s);
    if constexpr (S >= 4) {
        acc = acc + row_next_f<8>(s);
        acc = acc + row_next_f<12>(s);
    }
#pragma unroll
    for (int r = 1; r < S / 4; ++r) {  // 8 and 16 lanes: sub-lanes 4r..4r+3 in row r
        const float sr = __shfl_down(s, 16 * r, 64);  // lane l + 16 r: row r's copy
        acc = acc + sr;
        acc = acc + row_next_f<4>(sr);
        acc = acc + row_next_f<8>(sr);
        acc = acc + row_next_f<12>(sr);
    }
    return acc;
}

// Trilinear sample from the [z][y][x] layout with floor/ceil corners (DataCube.cpp:224-270):
// exact for any data, including inf/nan.
__device__ __forceinline__ float sample_linear_tri(const DevVolume& v, const ExactIdx& r) {
    float x0f = floorf(r.xd), y0f = floorf(r.yd), z0f = floorf(r.zd);
    float fx = r.xd - x0f, fy = r.yd - y0f, fz = r.zd - z0f;
    int x0 = to_index(x0f, v.fnx - 2), y0 = to_index(y0f, v.fny - 2), z0 = to_index(z0f, v.fnz - 2);
    int x1 = x0 + (fx > 0.0f ? 1 : 0), y1 = y0 + (fy > 0.0f ? 1 : 0), z1 = z0 + (fz > 0.0f ? 1 : 0);
    const float* L = v.linear;
    size_t sy = (size_t)v.nx, sz = (size_t)v.nx * (size_t)v.ny;
    float s000 = L[z0 * sz + y0 * sy + x0], s001 = L[z0 * sz + y0 * sy + x1];
    float s010 = L[z0 * sz + y1 * sy + x0], s011 = L[z0 * sz + y1 * sy + x1];
    float s100 = L[z1 * sz + y0 * sy + x0], s101 = L[z1 * sz + y0 * sy + x1];
    float s110 = L[z1 * sz + y1 * sy + x0], s111 = L[z1 * sz + y1 * sy + x1];
    float ly = lerp_general(lerp_general(s000, s001, fx), lerp_general(s010, s011, fx), fy);
    float hy = lerp_general(lerp_general(s100, s101, fx), lerp_general(s110, s111, fx), fy);
    return lerp_general(ly, hy, fz);
}

// Nearest: slices[(int)z][(int)y][(int)x] (DataCube.cpp:273), truncation.
__device__ __forceinline__ float sample_linear_nearest(const DevVolume& v, const ExactIdx& r) {
    int x = to_index(truncf(r.xd), v.fnx - 2), y = to_index(truncf(r.yd), v.fny - 2), z = to_index(truncf(r.zd), v.fnz - 2);
    return v.linear[((size_t)z * v.ny + y) * v.nx + x];
}

// ------------------------------------------------------------------------------------------
// CMEM::GetSample (Cmem.cpp:117-288) on the device: the reference's float/double types and
// order; the transcendental functions are ROCm's device libm (within ~1 ulp of glibc's).
// ------------------------------------------------------------------------------------------
// (float)acos(c) for a double c, to float accuracy without an FP64 arc cosine: acos(c) =
// 2 asin(sqrt((1 - c) / 2)) for c >= 0 and pi - 2 asin(sqrt((1 + c) / 2)) below, with 1 -/+ c formed
// in double (exact near the ends, where acos is steep) and the rest in float; within a few float
// ulps of the double arc cosine rounded to float. |c| > 1 gives NaN (a negative square root), as
// acos does.
__device__ __forceinline__ float acos_to_float(double c) {
    if (c >= 0.0) return 2.0f * asinf(sqrtf((float)((1.0 - c) * 0.5)));
    return 3.14159265358979323846f - 2.0f * asinf(sqrtf((float)((1.0 + c) * 0.5)));
}

__device__ __forceinline__ float cmem_sample(const CmemParams& m, float x, float y, float z) {
    float r = sqrtf(x * x + y * y + z * z);
    const float ryz = sqrtf(y * y + z * z);
    float theta = (x != 0) ? acosf(x / r) : 0.f;
    // phi = acos(y / ryz) (0 when y == 0) enters only through its cosine and sine, and theta's
    // cosine and sine are needed too: all four from the ratios that are their arc cosines' arguments
    // (theta, phi in [0, pi], so both sines are >= 0; x == 0 and y == 0 give the reference's angle
    // 0), within an ulp or two of cos(acos(u)) and sin(acos(u)), with no trigonometric call
    const float ct = (x != 0) ? x / r : 1.f, st = (x != 0) ? ryz / r : 0.f;
    const float cos_phi = (y != 0) ? y / ryz : 1.f, sin_phi = (y != 0) ? fabsf(z) / ryz : 0.f;
    // LinScaled (Cmem.cpp:188-222)
    // cos(phi - pi/2) and cos(phi - 3 pi/2) in double (Cmem.cpp:205, 209) are sin(phi) and
    // -sin(phi), formed here in double from the coordinates
    const double sphi = (y != 0) ? (double)fabsf(z) / sqrt((double)y * y + (double)z * z) : 0.0;
    float phi_n = acos_to_float((double)(ct * m.cos_theta_n) + (double)(st * m.sin_theta_n) * sphi);
    float phi_s = acos_to_float((double)(ct * m.cos_theta_s) + (double)(st * m.sin_theta_s) * -sphi);
    float f = powf(cosf(theta / 2) + m.a5 * sinf(2 * theta) * (1 - expf(-theta)),
                   (m.p1 * (m.beta_c[0] + m.beta_c[1] * cos_phi + m.beta_c[3] * (sin_phi * sin_phi))));
    float q = m.p2 * m.charlie * expf(m.p3 * m.dn * (powf(phi_n, m.a21))) +
              m.p2 * m.charlie * expf(m.p3 * m.ds * (powf(phi_s, m.a21)));
    float radius_mp = m.p0 * m.r0_lin * f + q;
    float h = 0.05f / 2;
    if (r < radius_mp - h) return 0;  // inside the magnetopause: the bow shock is not needed
    // ShueModel (Cmem.cpp:280-288)
    float base = 2 / (1 + ct);
    float ry = m.bs * powf(base, m.ay_bs);
    float rz = (m.az_bs == m.ay_bs) ? ry : m.bs * powf(base, m.az_bs);  // equal unless set apart
    float rzc = rz * cos_phi, rys = ry * sin_phi;
    float radius_bs = (ry * rz) / sqrtf((rzc * rzc) + (rys * rys));
    // the reference's zone tests in its order (Cmem.cpp:136-176); the sheath emissivity is
    // evaluated only for the three zones that use it
    const bool z_mp = r >= radius_mp - h && r <= radius_mp + h;
    const bool z_sheath = !z_mp && r > radius_mp + h && r < radius_bs - h;
    const bool z_bs = !z_mp && !z_sheath && r >= radius_bs - h && r <= radius_bs + h;
    double sd = r / 10;
    if (!(z_mp || z_sheath || z_bs)) return (float)(m.A2 * (1 / (sd * sd * sd)));
    float half_t = theta / 2;
    float inner = m.A1 * (expf(-m.B * (half_t * half_t * half_t * half_t))) * powf(r / 10, (-m.alpha - (m.beta * (st * st))));
    if (z_mp) {
        float norm = (r - (radius_mp - h)) / (radius_mp + h);
        return lerp_general(0.f, inner, norm);
    }
    if (z_sheath) return inner;
    float norm = (r - (radius_bs - h)) / (radius_bs + h);
    float b2 = (float)(m.A2 * (1 / (sd * sd * sd)));
    return lerp_general(inner, b2, norm);
}

// ------------------------------------------------------------------------------------------
// cube entry: FP64 slab interval, widened by `margin`, mapped to a sample range, then walked
// with the exact FP32 predicate so that the first and last samples are the ones the
// reference's own bounds test accepts. The in-cube samples of a ray form one contiguous range
// because every term of the FP32 index is monotone in t.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void slab_axis(double p, double d, double lo, double hi, double& tin, double& tout) {
    if (d == 0.0) {
        if (p < lo || p > hi) { tin = 1e300; tout = -1e300; }
    } else {
        double inv = 1.0 / d;
        double t1 = (lo - p) * inv, t2 = (hi - p) * inv;
        tin = fmax(tin, fmin(t1, t2));
        tout = fmin(tout, fmax(t1, t2));
    }
}

__device__ void exact_clip(const DevVolume& v, const ExactView& ev, const float* s_t, float dx, float dy, float dz,
                           int& kf, int& kl) {
    double m = (double)ev.margin;
    double tin = -1e300, tout = 1e300;
    slab_axis(ev.px, dx, (double)v.ox - m, (double)v.ox + (double)(v.nx - 1) * (double)v.sx + m, tin, tout);
    slab_axis(ev.py, dy, (double)v.oy - m, (double)v.oy + (double)(v.ny - 1) * (double)v.sy + m, tin, tout);
    slab_axis(ev.pz, dz, (double)v.oz - m, (double)v.oz + (double)(v.nz - 1) * (double)v.sz + m, tin, tout);
    int n = ev.n;
    if (!(tin <= tout)) { kf = n; kl = -1; return; }
    double sc = (double)n / (double)ev.L;
    // The margin (2^-16 of the scene's scale) exceeds every FP32 rounding of a sample's index,
    // so a sample outside the widened box is outside the FP32 test too: the widened range starts
    // at or before the first accepted sample and ends at or after the last (FP64 rounding of the
    // two expressions can only move them outward), and the walks below only move inward.
    double klo = ceil(tin * sc - 0.5), khi = floor(tout * sc - 0.5);
    klo = fmin(fmax(klo, 0.0), (double)n);
    khi = fmin(fmax(khi, -1.0), (double)(n - 1));
    int a = (int)klo, b = (int)khi;
    while (a <= b && !exact_inside(v, exact_index<true>(v, dx, dy, dz, ev.px, ev.py, ev.pz, s_t[a]))) a++;
    while (b >= a && !exact_inside(v, exact_index<true>(v, dx, dy, dz, ev.px, ev.py, ev.pz, s_t[b]))) b--;
    kf = a;
    kl = b;
}

// Tile geometry of the exact kernels: a wave holds 64/S rays as TWxTH, a block 2x2 waves. Two
// lanes per ray take 4x8-ray waves (8x16-pixel blocks): on the 256^3 cube's 576x336 x 800 view
// 0.0900 ms against 0.0914 ms with 8x4 and 0.110 ms with 2x16 (4 samples in flight each), and
// the same 4.92 - 4.94 ms as 8x4 for the 256-view orbit batch.
template <int S> struct ExactTile {
    static constexpr int TW = (S == 1) ? 8 : (S == 16 ? 2 : 4);
    static constexpr int TH = (64 / S) / TW;
};

// The lanes of a wave and the rays they trace. The wave's TW x TH rays are cut into 2x2 pixel
// blocks; a block takes 4S consecutive lanes (one 16-lane row for S <= 4, two for S = 8, the
// whole wave for S = 16), its
// four pixels on the four lanes of a quad and its S sub-lanes on S consecutive quads. A 16-byte-per-lane record load is served a
// quad (64 B) per cycle, one cycle per distinct cache line in the quad: the four rays of a 2x2
// block sample the same k at points less than a voxel apart, mostly inside one line, where a
// quad of two rays' consecutive samples spans about two (on the 256^3 cube's 576x336 x 800 view,
// 22.3 line lookups per 64-lane load against 32.2). Sub-lane m of a ray sits 4m lanes above
// sub-lane 0 (row_next<4m>).
template <int S>
__device__ __forceinline__ void exact_lane(int lane, int& ti, int& tj, int& sub) {
    constexpr int TW = ExactTile<S>::TW;
    sub = (lane >> 2) & (S - 1);
    const int blk = lane / (4 * S);  // 2x2 block of the wave, row-major
    ti = (blk % (TW / 2)) * 2 + (lane & 1);
    tj = (blk / (TW / 2)) * 2 + ((lane >> 1) & 1);
}


// ------------------------------------------------------------------------------------------
// the exact render: one block renders one 2x2-wave tile `lb` of a view (wg_x tiles per row of
// tiles); every thread of the block calls it
// ------------------------------------------------------------------------------------------
template <int S, int MODE, int P>
__device__ __forceinline__ void exact_block(const DevVolume& v, const ExactView& ev, int lb, int wg_x) {
    constexpr bool QUAD = (MODE == SM_QUAD_POS || MODE == SM_QUAD_GEN);
    constexpr bool POS = (MODE == SM_QUAD_POS);
    extern __shared__ unsigned s_mem[];
    float* s_t = (float*)s_mem;
    const int n = ev.n;
    // ray distances t_k = L * ((k + 0.5f) / n) (Camera.cpp:73), IEEE division as on the host
    for (int k = threadIdx.x; k < n; k += BIRP_BLOCK) s_t[k] = ev.L * (((float)k + 0.5f) / (float)n);
    QuadCtx qc;
    if constexpr (QUAD) {
        unsigned* tab = s_mem + ((n + 3) & ~3);
        const int ntab = v.nx + v.ny + v.nz;
        for (int q = threadIdx.x; q < ntab; q += BIRP_BLOCK) tab[q] = v.qtab[q];
        qc.rs = make_rsrc(v.quad, v.quad_bytes);
        qc.tx = tab;
        qc.ty = tab + v.nx;
        qc.tz = tab + v.nx + v.ny;
        qc.hx = v.fnx - 2.0f;
        qc.hy = v.fny - 2.0f;
        qc.hz = v.fnz - 2.0f;
    }
    __syncthreads();

    constexpr int TW = ExactTile<S>::TW, TH = ExactTile<S>::TH;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    int ti, tj, sub;
    exact_lane<S>(lane, ti, tj, sub);
    const int i = ((lb % wg_x) * 2 + (wave & 1)) * TW + ti;
    const int j = ((lb / wg_x) * 2 + (wave >> 1)) * TH + tj;
    const bool valid = (i < ev.W) && (j < ev.H);

    float dx = 0.f, dy = 0.f, dz = 1.f;
    int kf = n, kl = -1;
    if (valid) {
        exact_ray_dir(ev, i, j, dx, dy, dz);
        if constexpr (MODE == SM_CMEM) {
            kf = 0;
            kl = n - 1;
        } else {
            exact_clip(v, ev, s_t, dx, dy, dz, kf, kl);
            if (ev.semantics == BIRP_SEM_REFERENCE && kf != 0) kl = -1;  // first sample outside: ray dead
        }
    }
    // Every ray walks its own range kf..kl: sub-lane `sub` takes k = kf + kb + sub, and the wave
    // makes as many trips as its longest ray needs. t is read at a clamped LDS byte offset.
    const int len = (kl >= kf) ? kl - kf + 1 : 0;
    const int trips = wave_max_i(len);
    const int lim = len - sub;  // sample kb + sub of this ray exists iff kb < lim
    const unsigned t_first = (unsigned)(kf + sub) * 4u, t_last = (unsigned)(n - 1) * 4u;

    float acc = 0.f;
    if (valid && ev.accumulate) acc = ev.out[(size_t)j * ev.W + i];
    bool alive = true;
    unsigned added = 0;
    const float w = ev.w;
    const float px = ev.px, py = ev.py, pz = ev.pz;

    if constexpr (POS) {
        // P samples per lane per iteration (k, k + S, ...), staged so that the later samples' LDS
        // and buffer reads are in flight while the first ones' lerps run. The sum order is
        // unchanged: the S samples at k, then the S samples at k + S, and so on.
        // Every lane samples inside the cube, so the corner indices need no clamp: a lane past
        // its ray's last in-cube sample kl re-reads kl's t (its result is masked), and a lane
        // whose ray has no in-cube sample is sent to the cube's origin (zero direction), which is
        // inside. The in-cube samples kf..kl are contiguous (the FP32 index is monotone in t).
        const bool has = len > 0;
        const float rx = has ? dx : 0.f, ry = has ? dy : 0.f, rz = has ? dz : 0.f;
        const float qx = has ? px : v.ox, qy = has ? py : v.oy, qz = has ? pz : v.oz;
        // t is read at LDS byte address min(t_lo + 4k, t_hi), both bounds including the LDS base
        typedef __attribute__((address_space(3))) const float* lds_float_ptr;
        const unsigned t_base = (unsigned)(uintptr_t)(lds_float_ptr)s_t;
        unsigned t_lo = t_base + (has ? t_first : 0u), t_hi = t_base + (has ? (unsigned)kl * 4u : 0u);
        // keep the selects above out of the loop (the compiler would otherwise redo them per sample)
        asm volatile("" : "+v"(t_lo), "+v"(t_hi));
        // Trips that every lane of the wave fills (kb + P S <= the wave's shortest range) need
        // neither the clamp of t's address nor the mask, and read their P t values at constant
        // offsets from one address. The arithmetic and the sum order are those of the loop below.
        const int full = wave_min_i(len);
        int kb = 0;
        for (; kb + P * S <= full; kb += P * S) {
            QuadFetch q[P];
            const unsigned tb = t_lo + (unsigned)kb * 4u;
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const float t = *(lds_float_ptr)(uintptr_t)(tb + (unsigned)(u * S) * 4u);
                const ExactIdx r = exact_index<false>(v, rx, ry, rz, qx, qy, qz, t);
                q[u] = quad_fetch_inside(qc, r.xd, r.yd, r.zd);
            }
            float sc[P];
#pragma unroll
            for (int u = 0; u < P; ++u) sc[u] = scale_pos(quad_lerp<true>(q[u]) * w);  // Camera.cpp:201
#pragma unroll
            for (int u = 0; u < P; ++u) acc = add_in_order<S>(acc, sc[u]);
        }
        for (; kb < trips; kb += P * S) {
            QuadFetch q[P];
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const float t = *(lds_float_ptr)(uintptr_t)min(t_lo + (unsigned)(kb + u * S) * 4u, t_hi);
                const ExactIdx r = exact_index<false>(v, rx, ry, rz, qx, qy, qz, t);
                q[u] = quad_fetch_inside(qc, r.xd, r.yd, r.zd);
            }
            float sc[P];
#pragma unroll
            for (int u = 0; u < P; ++u) {
                sc[u] = scale_pos(quad_lerp<true>(q[u]) * w);  // Camera.cpp:201
                sc[u] = (kb + u * S < lim) ? sc[u] : -0.0f;
            }
#pragma unroll
            for (int u = 0; u < P; ++u) acc = add_in_order<S>(acc, sc[u]);
        }
    }
    for (int kb = 0; !POS && kb < trips; kb += S) {
        const float t = *(const float*)((const char*)s_t + min(t_first + (unsigned)kb * 4u, t_last));
        const bool act = kb < lim;
        float val;
        if constexpr (MODE == SM_CMEM) {
            val = cmem_sample(v.cmem, dx * t + px, dy * t + py, dz * t + pz);
        } else {
            ExactIdx r = exact_index<true>(v, dx, dy, dz, px, py, pz, t);
            if constexpr (MODE == SM_QUAD_GEN) val = sample_quad<false>(qc, r.xd, r.yd, r.zd);
            else if constexpr (MODE == SM_LIN_TRI) val = sample_linear_tri(v, r);
            else val = sample_linear_nearest(v, r);
        }
        const float vw = val * w;  // Camera.cpp:201
        {
            const float s = birp_scale_sample(vw);
            if constexpr (S == 1) {
                bool add = act && alive;
                bool stop = add && (s < 0);  // Camera.cpp:207-210
                acc = (add && !stop) ? acc + s : acc;
                added += (add && !stop) ? 1u : 0u;
                alive = alive && !stop;
            } else {
                // sub-lane 0 gathers the ray's S samples in k order (the other sub-lanes'
                // sums and stops are not used)
                float sv[S];
                int av[S];
                const int ai = act ? 1 : 0;
                sv[0] = s;
                av[0] = ai;
                sv[1] = row_next_f<4>(s);
                av[1] = row_next<4>(ai);
                if constexpr (S >= 4) {
                    sv[2] = row_next_f<8>(s);
                    av[2] = row_next<8>(ai);
                    sv[3] = row_next_f<12>(s);
                    av[3] = row_next<12>(ai);
                }
#pragma unroll
                for (int rr = 1; rr < S / 4; ++rr) {  // sub-lanes 4rr..4rr+3: row rr
                    const float sr = __shfl_down(s, 16 * rr, 64);
                    const int ar = __shfl_down(ai, 16 * rr, 64);
                    sv[4 * rr] = sr;
                    av[4 * rr] = ar;
                    sv[4 * rr + 1] = row_next_f<4>(sr);
                    av[4 * rr + 1] = row_next<4>(ar);
                    sv[4 * rr + 2] = row_next_f<8>(sr);
                    av[4 * rr + 2] = row_next<8>(ar);
                    sv[4 * rr + 3] = row_next_f<12>(sr);
                    av[4 * rr + 3] = row_next<12>(ar);
                }
#pragma unroll
                for (int mm = 0; mm < S; ++mm) {
                    bool add = av[mm] && alive;
                    bool stop = add && (sv[mm] < 0);
                    acc = (add && !stop) ? acc + sv[mm] : acc;
                    added += (add && !stop) ? 1u : 0u;
                    alive = alive && !stop;
                }
            }
        }
    }
    if constexpr (POS) added = (unsigned)len;
    if (valid && sub == 0) ev.out[(size_t)j * ev.W + i] = acc;
    if (ev.counter || ev.wave_counts) {
        unsigned tot = wave_sum_u((valid && sub == 0) ? added : 0u);
        if (lane == 0) {
            if (ev.wave_counts) ev.wave_counts[lb * 4 + wave] = tot;  // a plain store, every wave
            else if (tot) atomicAdd(ev.counter, (unsigned long long)tot);
        }
    }
}

// One view: block b renders tile lb(b).
// Without a plan, lb = b: consecutive tiles go round-robin over the 8 XCDs, so every XCD renders
// a fair share of every part of the image (contiguous bands, one per XCD, left the XCDs with the
// middle rows of the cube's image working longest).
// With a plan (`order`: the tiles sorted heaviest first by k_exact_plan), slot b takes the
// sorted tile at its place in a snake over the CUs: round r of ncu slots runs forward when r is
// even and backward when odd, so that every CU draws one tile from each round and the heavy
// rounds are spread evenly (the longest-processing-time-first rule). The grid is then nblocks.
template <int S, int MODE, int P = 1>
__global__ __launch_bounds__(BIRP_BLOCK) void k_render_exact(DevVolume v, ExactView ev, int wg_x, int nblocks,
                                                             const int* __restrict__ order, int ncu) {
    if ((int)blockIdx.x >= nblocks) return;
    int lb = blockIdx.x;
    if (order) {
        const int b = blockIdx.x, r = b / ncu, base = r * ncu, c = b - base;
        const int nr = min(ncu, nblocks - base);
        lb = order[base + ((r & 1) ? nr - 1 - c : c)];
    }
    if (lb < 0 || lb >= nblocks) return;
    exact_block<S, MODE, P>(v, ev, lb, wg_x);
}

// ------------------------------------------------------------------------------------------
// Tile plan of a single exact view. A view of a few thousand tiles is resident all at once
// (a 576x336 view is 1512 blocks, six per CU): each CU keeps the blocks the dispatcher handed it,
// and a CU that drew several tiles across the middle of the cube's image finishes last while
// others idle. The plan estimates every tile's work -- the longest in-cube run of samples among
// five rays of each wave (its four corners and its centre), from a slab test against the cube's
// box in approximate FP32 -- counting-sorts the tiles on it, heaviest first, and writes the
// order. Orders change which CU renders a tile, never what it computes: the image is the same
// bit for bit whatever the order.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void plan_slab(float p, float rd, float lo, float hi, float& tin, float& tout) {
    const float t1 = (lo - p) * rd, t2 = (hi - p) * rd;  // rd = 1/d (inf for d = 0: NaN t drop out)
    tin = fmaxf(tin, fminf(t1, t2));
    tout = fminf(tout, fmaxf(t1, t2));
}

// samples of ray (i, j) inside the cube's box, approximately: an estimate of its trip count
__device__ __forceinline__ int plan_ray_len(const DevVolume& v, const ExactView& ev, int i, int j, float sc, float rW,
                                            float rH) {
    const float sx = (2.f * ((float)i + 0.5f) * rW - 1.f) * ev.tanx * ev.aspect;
    const float sy = (1.f - 2.f * ((float)j + 0.5f) * rH) * ev.tany;
    // world direction of the camera ray (sx, sy, -1), unnormalised: t along it is t / |(sx, sy, 1)|
    const float dx = ev.rinv[0] * sx + ev.rinv[1] * sy - ev.rinv[2];
    const float dy = ev.rinv[3] * sx + ev.rinv[4] * sy - ev.rinv[5];
    const float dz = ev.rinv[6] * sx + ev.rinv[7] * sy - ev.rinv[8];
    float tin = -3.0e38f, tout = 3.0e38f;
    plan_slab(ev.px, __builtin_amdgcn_rcpf(dx), v.ox, v.ox + (float)(v.nx - 1) * v.sx, tin, tout);
    plan_slab(ev.py, __builtin_amdgcn_rcpf(dy), v.oy, v.oy + (float)(v.ny - 1) * v.sy, tin, tout);
    plan_slab(ev.pz, __builtin_amdgcn_rcpf(dz), v.oz, v.oz + (float)(v.nz - 1) * v.sz, tin, tout);
    if (!(tin <= tout)) return 0;
    const float s = sc * __builtin_amdgcn_sqrtf(sx * sx + sy * sy + 1.f);  // samples per unit of t
    const float klo = fminf(fmaxf(ceilf(tin * s - 0.5f), 0.f), (float)ev.n);
    const float khi = fmaxf(fminf(floorf(tout * s - 0.5f), (float)(ev.n - 1)), -1.f);
    if (ev.semantics == BIRP_SEM_REFERENCE && klo > 0.f) return 0;  // first sample outside: dead ray
    return khi >= klo ? (int)(khi - klo) + 1 : 0;
}

// grid: ceil(4 nblocks / BIRP_BLOCK) blocks of BIRP_BLOCK threads, thread 4b + w estimating wave
// w of tile b (a tile's four waves sit in neighbouring lanes). The last block to take a ticket
// (which it resets for the next plan) sorts. The estimates go between blocks as agent-scope
// atomic stores and loads, and each block waits for its stores before it takes its ticket:
// no cache-wide fence. The sort deals out exactly the tiles 0..nblocks-1 whatever estimates it
// reads, so the order is a permutation in every case. plan = [order | estimates | ticket], each
// of the first two BIRP_PLAN_MAX_BLOCKS ints, BIRP_PLAN_INTS in all; the ticket starts at zero.
template <int S>
__global__ __launch_bounds__(BIRP_BLOCK) void k_exact_plan(DevVolume v, ExactView ev, int wg_x, int nblocks, int* __restrict__ plan) {
    constexpr int TW = ExactTile<S>::TW, TH = ExactTile<S>::TH, NB = BIRP_PLAN_BUCKETS, T = BIRP_BLOCK;
    constexpr int PER_LANE = NB / 64;  // buckets per lane of the scanning wave
    static_assert(NB % 64 == 0, "plan shape");
    static_assert(4 * BIRP_MAX_EXACT_SAMPLES <= 65535, "tile estimates are kept as u16");
    int* order = plan;
    int* cost = plan + BIRP_PLAN_MAX_BLOCKS;
    int* ticket = plan + 2 * BIRP_PLAN_MAX_BLOCKS;
    __shared__ unsigned short s_key[BIRP_PLAN_MAX_BLOCKS];
    __shared__ int s_start[NB];
    __shared__ int s_lo, s_hi, s_last;
    const int tid = threadIdx.x;
    const int q = blockIdx.x * T + tid, b = q >> 2, w = q & 3;
    int m = 0;
    if (b < nblocks) {
        const int i0 = ((b % wg_x) * 2 + (w & 1)) * TW, j0 = ((b / wg_x) * 2 + (w >> 1)) * TH;
        if (i0 < ev.W && j0 < ev.H) {
            const float sc = (float)ev.n / ev.L, rW = 1.f / (float)ev.W, rH = 1.f / (float)ev.H;
            const int i1 = min(i0 + TW - 1, ev.W - 1), j1 = min(j0 + TH - 1, ev.H - 1);
            m = plan_ray_len(v, ev, (i0 + i1) >> 1, (j0 + j1) >> 1, sc, rW, rH);
            m = max(m, plan_ray_len(v, ev, i0, j0, sc, rW, rH));
            m = max(m, plan_ray_len(v, ev, i1, j0, sc, rW, rH));
            m = max(m, plan_ray_len(v, ev, i0, j1, sc, rW, rH));
            m = max(m, plan_ray_len(v, ev, i1, j1, sc, rW, rH));
        }
    }
    m += __shfl_xor(m, 1);
    m += __shfl_xor(m, 2);
    if (b < nblocks && w == 0) __hip_atomic_store(cost + b, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // this thread's store has completed
    __syncthreads();
    if (tid == 0) s_last = atomicAdd(ticket, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;

    for (int k = tid; k < NB; k += T) s_start[k] = 0;
    if (tid == 0) { s_lo = 0x7fffffff; s_hi = 0; }
    int lo = 0x7fffffff, hi = 0;
    for (int t0 = 0; t0 < nblocks; t0 += 8 * T) {
        int c[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)  // eight loads in flight; the index is clamped into the area
            c[k] = __hip_atomic_load(cost + min(t0 + k * T + tid, BIRP_PLAN_MAX_BLOCKS - 1), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int t = t0 + k * T + tid;
            if (t < nblocks) {
                const int cv = min(max(c[k], 0), 4 * BIRP_MAX_EXACT_SAMPLES);
                s_key[t] = (unsigned short)cv;
                lo = min(lo, cv);
                hi = max(hi, cv);
            }
        }
    }
    __syncthreads();
    atomicMin(&s_lo, lo);
    atomicMax(&s_hi, hi);
    __syncthreads();
    // NB buckets over [lo, hi], the heaviest in bucket 0
    const int base = s_lo, span = s_hi - s_lo;
    for (int t = tid; t < nblocks; t += T) {
        const int qq = span > 0 ? (int)(((long long)(s_key[t] - base) * (NB - 1)) / span) : 0;
        s_key[t] = (unsigned short)(NB - 1 - qq);
        atomicAdd(&s_start[NB - 1 - qq], 1);
    }
    __syncthreads();
    // exclusive prefix over the buckets by the first wave: lane l owns PER_LANE buckets in a row
    if (tid < 64) {
        int cnt[PER_LANE], sum = 0;
#pragma unroll
        for (int k = 0; k < PER_LANE; ++k) {
            cnt[k] = s_start[tid * PER_LANE + k];
            sum += cnt[k];
        }
        int incl = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int up = __shfl_up(incl, off);
            if (tid >= off) incl += up;
        }
        int run = incl - sum;
#pragma unroll
        for (int k = 0; k < PER_LANE; ++k) {
            s_start[tid * PER_LANE + k] = run;
            run += cnt[k];
        }
    }
    __syncthreads();
    // every tile takes one place in its bucket's range: the places are 0..nblocks-1, each once
    for (int t = tid; t < nblocks; t += T) order[atomicAdd(&s_start[s_key[t]], 1)] = t;
    if (tid == 0) atomicExch(ticket, 0);  // ready for the next plan
}

// Many views (each with its own volume, image and sample-count slots) in one launch: items are
// the views' tiles, view-major (prefix[k] = first item of view k), and block g renders item g.
// Consecutive blocks go round-robin over the 8 XCDs, so every XCD renders a fair share of every
// view: contiguous bands, one per XCD, left whole XCDs with the heavy views (a 256-view orbit
// batch took 6.07 ms in bands and 4.89 ms interleaved). Every block renders one tile with exactly
// the arithmetic of k_render_exact, so each image is bit-identical to its single-view render; with
// more tiles in flight than one view has, the SIMDs stay fed while the longest rays of a view finish.
// `order` (may be null): block b renders item order[b] instead of item b (a batch's tile plan).
template <int S, int MODE, int P = 1>
__global__ __launch_bounds__(BIRP_BLOCK) void k_render_exact_multi(const DevVolume* __restrict__ vols,
                                                                   const ExactView* __restrict__ views,
                                                                   const int* __restrict__ prefix, int nview, int total,
                                                                   const int* __restrict__ order) {
    if ((int)blockIdx.x >= total) return;
    const int g = order ? order[blockIdx.x] : (int)blockIdx.x;
    if (g < 0 || g >= total) return;
    int lo = 0, hi = nview - 1;  // the view of item g: the last k with prefix[k] <= g
    while (lo < hi) {
        const int mid = __builtin_amdgcn_readfirstlane((lo + hi + 1) >> 1);
        if (prefix[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const ExactView& ev = views[lo];
    constexpr int TW = ExactTile<S>::TW;
    exact_block<S, MODE, P>(vols[ev.vol], ev, g - prefix[lo], (ev.W + 2 * TW - 1) / (2 * TW));
}

// Tile estimates of a multi-view item list (the batch form of k_exact_plan's first phase): thread
// 4g + w estimates wave w of item g; cost[g] = the sum over the item's four waves.
template <int S>
__global__ __launch_bounds__(BIRP_BLOCK) void k_plan_multi(const DevVolume* __restrict__ vols, const ExactView* __restrict__ views,
                                                           const int* __restrict__ prefix, int nview, int total, int* __restrict__ cost) {
    constexpr int TW = ExactTile<S>::TW, TH = ExactTile<S>::TH;
    const long long q = (long long)blockIdx.x * BIRP_BLOCK + threadIdx.x;
    const int g = (int)(q >> 2), w = (int)(q & 3);
    int m = 0;
    if (g < total) {
        int lo = 0, hi = nview - 1;  // the view of item g: the last k with prefix[k] <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= g) lo = mid;
            else hi = mid - 1;
        }
        const ExactView& ev = views[lo];
        const DevVolume& v = vols[ev.vol];
        const int lb = g - prefix[lo], wg_x = (ev.W + 2 * TW - 1) / (2 * TW);
        const int i0 = ((lb % wg_x) * 2 + (w & 1)) * TW, j0 = ((lb / wg_x) * 2 + (w >> 1)) * TH;
        if (i0 < ev.W && j0 < ev.H) {
            const float sc = (float)ev.n / ev.L, rW = 1.f / (float)ev.W, rH = 1.f / (float)ev.H;
            const int i1 = min(i0 + TW - 1, ev.W - 1), j1 = min(j0 + TH - 1, ev.H - 1);
            m = plan_ray_len(v, ev, (i0 + i1) >> 1, (j0 + j1) >> 1, sc, rW, rH);
            m = max(m, plan_ray_len(v, ev, i0, j0, sc, rW, rH));
            m = max(m, plan_ray_len(v, ev, i1, j0, sc, rW, rH));
            m = max(m, plan_ray_len(v, ev, i0, j1, sc, rW, rH));
            m = max(m, plan_ray_len(v, ev, i1, j1, sc, rW, rH));
        }
    }
    m += __shfl_xor(m, 1);
    m += __shfl_xor(m, 2);
    if (g < total && w == 0) cost[g] = m;
}

// ------------------------------------------------------------------------------------------
// FP64 render (every operation in double) for error reporting on large configurations
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double lerp64(double a, double b, double t) {
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
    if (t == 1) return b;
    double x = a + t * (b - a);
    return ((t > 1) == (b > a)) ? (b < x ? x : b) : (b > x ? x : b);
}

__device__ double cmem_sample64(const CmemParams& m, double x, double y, double z) {
    const double PI = 3.14159265358979323846;
    double r = sqrt(x * x + y * y + z * z);
    double theta = (x != 0) ? acos(x / r) : 0.0;
    double phi = (y != 0) ? acos(y / sqrt(y * y + z * z)) : 0.0;
    double phi_n = acos(cos(theta) * cos((double)m.theta_n) + sin(theta) * sin((double)m.theta_n) * cos(phi - PI / 2));
    double phi_s = acos(cos(theta) * cos((double)m.theta_s) + sin(theta) * sin((double)m.theta_s) * cos(phi - 3 * PI / 2));
    double sp = sin(phi);
    double f = pow(cos(theta / 2) + m.a5 * sin(2 * theta) * (1 - exp(-theta)),
                   m.p1 * (m.beta_c[0] + m.beta_c[1] * cos(phi) + (double)m.beta_c[3] * (sp * sp)));
    double q = (double)m.p2 * m.charlie * exp((double)m.p3 * m.dn * pow(phi_n, (double)m.a21)) +
               (double)m.p2 * m.charlie * exp((double)m.p3 * m.ds * pow(phi_s, (double)m.a21));
    double rmp = (double)m.p0 * m.r0_lin * f + q;
    double ry = m.bs * pow(2 / (1 + cos(theta)), (double)m.ay_bs);
    double rz = m.bs * pow(2 / (1 + cos(theta)), (double)m.az_bs);
    double rbs = (ry * rz) / sqrt((rz * cos(phi)) * (rz * cos(phi)) + (ry * sin(phi)) * (ry * sin(phi)));
    double h = 0.025, hs = theta / 2, ss = sin(theta);
    double inner = (double)m.A1 * exp(-(double)m.B * (hs * hs * hs * hs)) * pow(r / 10, -(double)m.alpha - (double)m.beta * (ss * ss));
    double outer = (double)m.A2 / ((r / 10) * (r / 10) * (r / 10));
    if (r < rmp - h) return 0.0;
    if (r <= rmp + h) return lerp64(0.0, inner, (r - (rmp - h)) / (rmp + h));
    if (r < rbs - h) return inner;
    if (r <= rbs + h) return lerp64(inner, outer, (r - (rbs - h)) / (rbs + h));
    return outer;
}

__global__ __launch_bounds__(BIRP_BLOCK) void k_render_f64(DevVolume v, F64View fv) {
    int pix = blockIdx.x * BIRP_BLOCK + threadIdx.x;
    if (pix >= fv.W * fv.H) return;
    int i = pix % fv.W, j = pix / fv.W;
    double sx = (2 * ((i + 0.5) / fv.W) - 1) * fv.tanx * fv.aspect;
    double sy = (1 - 2 * ((j + 0.5) / fv.H)) * fv.tany;
    double nd = sqrt(sx * sx + sy * sy + 1);
    double u0 = sx / nd, u1 = sy / nd, u2 = -1 / nd;
    double w0 = fv.rinv[0] * u0 + fv.rinv[1] * u1 + fv.rinv[2] * u2;
    double w1 = fv.rinv[3] * u0 + fv.rinv[4] * u1 + fv.rinv[5] * u2;
    double w2 = fv.rinv[6] * u0 + fv.rinv[7] * u1 + fv.rinv[8] * u2;
    double wd = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    double dx = w0 / wd, dy = w1 / wd, dz = w2 / wd;
    double acc = 0.0;
    const double K = BIRP_SCALE_K, C = BIRP_SCALE_C;
    for (int k = 0; k < fv.n; ++k) {
        double t = fv.L * ((k + 0.5) / fv.n);
        double x = dx * t + fv.px, y = dy * t + fv.py, z = dz * t + fv.pz;
        double val;
        if (v.kind == SK_CMEM) {
            val = cmem_sample64(v.cmem, x, y, z);
        } else {
            double xi = x - (double)v.ox, yi = y - (double)v.oy, zi = z - (double)v.oz;
            double xd = xi / (double)v.sx, yd = yi / (double)v.sy, zd = zi / (double)v.sz;
            bool in = (xi >= 0) && (xd + 1 < v.nx) && (yi >= 0) && (yd + 1 < v.ny) && (zi >= 0) && (zd + 1 < v.nz);
            if (!in) {
                if (fv.semantics == BIRP_SEM_REFERENCE) break;
                continue;
            }
            if (v.interp == BIRP_INTERP_TRILINEAR) {
                int x0 = (int)floor(xd), y0 = (int)floor(yd), z0 = (int)floor(zd);
                int x1 = (int)ceil(xd), y1 = (int)ceil(yd), z1 = (int)ceil(zd);
                double fx = xd - x0, fy = yd - y0, fz = zd - z0;
                const float* L = v.linear;
                size_t sy_ = v.nx, sz_ = (size_t)v.nx * v.ny;
                double lx1 = lerp64(L[z0 * sz_ + y0 * sy_ + x0], L[z0 * sz_ + y0 * sy_ + x1], fx);
                double lx2 = lerp64(L[z0 * sz_ + y1 * sy_ + x0], L[z0 * sz_ + y1 * sy_ + x1], fx);
                double hx1 = lerp64(L[z1 * sz_ + y0 * sy_ + x0], L[z1 * sz_ + y0 * sy_ + x1], fx);
                double hx2 = lerp64(L[z1 * sz_ + y1 * sy_ + x0], L[z1 * sz_ + y1 * sy_ + x1], fx);
                val = lerp64(lerp64(lx1, lx2, fy), lerp64(hx1, hx2, fy), fz);
            } else {
                val = v.linear[((size_t)(int)zd * v.ny + (int)yd) * v.nx + (int)xd];
            }
        }
        double s = val * fv.w * K / C / 10000;
        if (s < 0) break;
        acc += s;
    }
    fv.out[pix] = acc;
}

// ------------------------------------------------------------------------------------------
// Any other Space (its GetSample is the caller's host code): the rays' sample points, then the
// integration of the caller's sample values (Camera.cpp:140-229 around the call at :191)
// ------------------------------------------------------------------------------------------
// One thread per (pixel, k) of rows [row0, row0 + rows): X = d t_k + P in the reference's FP32
// order (Camera.cpp:183-189), stored at xyz[((pixel - row0 W) n + k) * 3].
__global__ __launch_bounds__(BIRP_BLOCK) void k_space_points(ExactView ev, int row0, int rows, float* xyz) {
    const long long q = (long long)blockIdx.x * BIRP_BLOCK + threadIdx.x;
    const long long total = (long long)rows * ev.W * ev.n;
    if (q >= total) return;
    const int k = (int)(q % ev.n);
    const long long pix = q / ev.n;
    const int i = (int)(pix % ev.W), j = row0 + (int)(pix / ev.W);
    float dx, dy, dz;
    exact_ray_dir(ev, i, j, dx, dy, dz);
    const float t = ev.L * (((float)k + 0.5f) / (float)ev.n);  // Camera.cpp:73
    xyz[3 * q] = dx * t + ev.px;
    xyz[3 * q + 1] = dy * t + ev.py;
    xyz[3 * q + 2] = dz * t + ev.pz;
}

// One thread per pixel: the reference's scale chain (Camera.cpp:201), the stop at the first
// negative sample (:207-210) and the FP32 sum in k order (:223) over vals[(pixel - row0 W) n + k].
__global__ __launch_bounds__(BIRP_BLOCK) void k_space_integrate(ExactView ev, int row0, int rows, const float* vals, float* img) {
    const long long pix = (long long)blockIdx.x * BIRP_BLOCK + threadIdx.x;
    if (pix >= (long long)rows * ev.W) return;
    float acc = ev.accumulate ? img[pix] : 0.f;
    const float* v = vals + pix * ev.n;
    for (int k = 0; k < ev.n; ++k) {
        const float s = birp_scale_sample(v[k] * ev.w);
        if (s < 0) break;
        acc = acc + s;
    }
    img[pix] = acc;
    (void)row0;
}

// ------------------------------------------------------------------------------------------
// GetSample at arbitrary points (DataCube.cpp:203-277 / Cmem.cpp:117-186), exact arithmetic
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BIRP_BLOCK) void k_sample_points(DevVolume v, const float* xyz, long long n, float* out) {
    long long p = (long long)blockIdx.x * BIRP_BLOCK + threadIdx.x;
    if (p >= n) return;
    float x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
    if (v.kind == SK_CMEM) {
        out[p] = cmem_sample(v.cmem, x, y, z);
        return;
    }
    ExactIdx r;
    r.xi = x - v.ox;
    r.xd = r.xi / v.sx;
    r.yi = y - v.oy;
    r.yd = r.yi / v.sy;
    r.zi = z - v.oz;
    r.zd = r.zi / v.sz;
    if (!exact_inside(v, r)) {
        out[p] = -1.f;
        return;
    }
    out[p] = (v.interp == BIRP_INTERP_TRILINEAR) ? sample_linear_tri(v, r) : sample_linear_nearest(v, r);
}

// ------------------------------------------------------------------------------------------
// layout builders and scans
// ------------------------------------------------------------------------------------------
// flags[0]: bit 0 = some value <= 0, bit 1 = some value non-finite; flags[1], flags[2]: min and
// max of |v| over the nonzero finite values as float bits (ordered as unsigned for positive floats).
__global__ __launch_bounds__(BIRP_BLOCK) void k_scan_class(const float* vals, size_t n, unsigned* flags) {
    unsigned f = 0, mn = 0x7F7FFFFFu, mx = 0u;
    for (size_t idx = (size_t)blockIdx.x * BIRP_BLOCK + threadIdx.x; idx < n; idx += (size_t)gridDim.x * BIRP_BLOCK) {
        float x = vals[idx];
        if (!(x > 0.f)) f |= 1u;
        if (!isfinite(x)) f |= 2u;
        float a = fabsf(x);
        if (a > 0.f && isfinite(a)) {
            unsigned u = __float_as_uint(a);
            mn = min(mn, u);
            mx = max(mx, u);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        f |= (unsigned)__shfl_xor((int)f, o, 64);
        mn = min(mn, (unsigned)__shfl_xor((int)mn, o, 64));
        mx = max(mx, (unsigned)__shfl_xor((int)mx, o, 64));
    }
    if (lane_id() == 0) {
        if (f) atomicOr(flags, f);
        atomicMin(flags + 1, mn);
        atomicMax(flags + 2, mx);
    }
}

// Quad records: one thread per (x, y, brick z, layer 0..8) slot; tz[bz*8] is the brick's base.
__global__ __launch_bounds__(BIRP_BLOCK) void k_build_quad(DevVolume v, unsigned char* quad) {
    const unsigned* tx = v.qtab;
    const unsigned* ty = v.qtab + v.nx;
    const unsigned* tz = v.qtab + v.nx + v.ny;
    size_t cx = v.nx - 1, cy = v.ny - 1;
    size_t total = cx * cy * (size_t)v.nbz * 9u;
    for (size_t r = (size_t)blockIdx.x * BIRP_BLOCK + threadIdx.x; r < total; r += (size_t)gridDim.x * BIRP_BLOCK) {
        int x = (int)(r % cx);
        int y = (int)((r / cx) % cy);
        int zr = (int)(r / (cx * cy));
        int bz = zr / 9, lz = zr % 9;
        int z = bz * 8 + lz;
        if (z >= v.nz) continue;
        const float* L = v.linear;
        size_t base = ((size_t)z * v.ny + y) * v.nx + x;
        float4 q = make_float4(L[base], L[base + 1], L[base + v.nx], L[base + v.nx + 1]);
        unsigned off = tx[x] + ty[y] + tz[bz * 8] + (unsigned)lz * 1024u;
        *(float4*)(quad + off) = q;
    }
}

__device__ __forceinline__ unsigned short f32_to_bf16_bits(float x) {
    __hip_bfloat16 h = __float2bfloat16(x);
    return *reinterpret_cast<unsigned short*>(&h);
}
__device__ __forceinline__ unsigned short f32_to_f16_bits(float x) {
    __half h = __float2half_rn(x);
    return *reinterpret_cast<unsigned short*>(&h);
}

// Per-brick power-of-two scale for F16: max|v| over the brick's corners maps to < 2^15. With
// v.f16_uniform every brick gets the volume's scale v.f16_scale instead.
__global__ __launch_bounds__(BIRP_BLOCK) void k_brick_scale(DevVolume v, float* bscale) {
    int brick = blockIdx.x;
    if (v.f16_uniform) {
        if (threadIdx.x == 0) {
            int bx = brick % v.nbx, by = (brick / v.nbx) % v.nby, bz = brick / (v.nbx * v.nby);
            const unsigned* t = v.otab;
            bscale[(t[bx * 8] + t[v.nx + by * 8] + t[v.nx + v.ny + bz * 8]) >> 13] = v.f16_scale;
        }
        return;
    }
    int bx = brick % v.nbx, by = (brick / v.nbx) % v.nby, bz = brick / (v.nbx * v.nby);
    float mx = 0.f;
    for (int t = threadIdx.x; t < 729; t += BIRP_BLOCK) {
        int lx = t % 9, ly = (t / 9) % 9, lz = t / 81;
        int x = bx * 8 + lx, y = by * 8 + ly, z = bz * 8 + lz;
        if (x < v.nx && y < v.ny && z < v.nz) mx = fmaxf(mx, fabsf(v.linear[((size_t)z * v.ny + y) * v.nx + x]));
    }
    __shared__ float red[BIRP_BLOCK / 64];
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (lane_id() == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        float s = 1.f;
        if (m > 0.f && isfinite(m)) {
            int e;
            frexpf(m, &e);           // m = f * 2^e, f in [0.5, 1)
            s = ldexpf(1.f, e - 15); // m / s < 2^15
        }
        const unsigned* t = v.otab;
        unsigned code = (t[bx * 8] + t[v.nx + by * 8] + t[v.nx + v.ny + bz * 8]) >> 13;
        bscale[code] = s;
    }
}

template <bool F16>
__global__ __launch_bounds__(BIRP_BLOCK) void k_build_oct(DevVolume v, unsigned char* oct, const float* bscale) {
    const unsigned* tx = v.otab;
    const unsigned* ty = v.otab + v.nx;
    const unsigned* tz = v.otab + v.nx + v.ny;
    size_t cx = v.nx - 1, cy = v.ny - 1, cz = v.nz - 1;
    size_t total = cx * cy * cz;
    for (size_t r = (size_t)blockIdx.x * BIRP_BLOCK + threadIdx.x; r < total; r += (size_t)gridDim.x * BIRP_BLOCK) {
        int x = (int)(r % cx), y = (int)((r / cx) % cy), z = (int)(r / (cx * cy));
        const float* L = v.linear;
        size_t sy = v.nx, sz = (size_t)v.nx * v.ny;
        size_t b0 = (size_t)z * sz + (size_t)y * sy + x;
        float c[8] = {L[b0], L[b0 + 1], L[b0 + sy], L[b0 + sy + 1], L[b0 + sz], L[b0 + sz + 1], L[b0 + sz + sy], L[b0 + sz + sy + 1]};
        unsigned off = tx[x] + ty[y] + tz[z];
        float inv = F16 ? 1.f / bscale[off >> 13] : 1.f;  // power of two: exact
        unsigned short h[8];
#pragma unroll
        for (int q = 0; q < 8; q++) h[q] = F16 ? f32_to_f16_bits(c[q] * inv) : f32_to_bf16_bits(c[q]);
        uint4 o;
        o.x = (unsigned)h[0] | ((unsigned)h[1] << 16);
        o.y = (unsigned)h[2] | ((unsigned)h[3] << 16);
        o.z = (unsigned)h[4] | ((unsigned)h[5] << 16);
        o.w = (unsigned)h[6] | ((unsigned)h[7] << 16);
        *(uint4*)(oct + off) = o;
    }
}

__global__ void k_reset_counter(unsigned* c) {
    if (blockIdx.x == 0 && threadIdx.x < 256) c[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(BIRP_BLOCK) void k_zero_f32(float* p, size_t n) {
    for (size_t idx = (size_t)blockIdx.x * BIRP_BLOCK + threadIdx.x; idx < n; idx += (size_t)gridDim.x * BIRP_BLOCK) p[idx] = 0.f;
}

// ------------------------------------------------------------------------------------------
// fast path: persistent work-queue kernel over (pose, tile) items
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float bf16lo(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16hi(unsigned u) { return __uint_as_float(u & 0xFFFF0000u); }

__device__ __forceinline__ float flerp(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Lerp of the FP16 pair (a, b) packed in one dword, in FP32: (a - t a) + t b as two fused
// multiply-adds, each reading its FP16 operand directly (v_fma_mix_f32, no conversions).
__device__ __forceinline__ float xlerp_f16(unsigned u, float t) {
    const half2_t h = __builtin_bit_cast(half2_t, u);
    const float a = (float)h.x, b = (float)h.y;
    return __builtin_fmaf(t, b, __builtin_fmaf(-t, a, a));
}

// Fast trilinear sample at index-space point (xd, yd, zd), already clamped into
// [0, nextbelow(n-1)] per axis, so the truncation is a valid cell index (0 .. n-2).
template <int DT>
__device__ __forceinline__ float fast_sample(const DevVolume& v, __amdgpu_buffer_rsrc_t rs, const unsigned* tab, float xd,
                                             float yd, float zd) {
    float fx = __builtin_amdgcn_fractf(xd), fy = __builtin_amdgcn_fractf(yd), fz = __builtin_amdgcn_fractf(zd);
    int ix = (int)xd, iy = (int)yd, iz = (int)zd;
    unsigned off = tab[ix] + tab[v.nx + iy] + tab[v.nx + v.ny + iz];
    if constexpr (DT == BIRP_DTYPE_F32) {
        float4 a = load16(rs, off), b = load16(rs, off + 1024u);
        float ly = flerp(flerp(a.x, a.y, fx), flerp(a.z, a.w, fx), fy);
        float hy = flerp(flerp(b.x, b.y, fx), flerp(b.z, b.w, fx), fy);
        return flerp(ly, hy, fz);
    } else {  // BF16, F16 with per-brick scale, F16 with one scale (BIRP_KDT_F16_UNIFORM)
        // each dword of the record holds the x-pair (x0, x1) of one (y, z) corner row
        uint4 o = load16u(rs, off);
        float c00, c10, c01, c11;  // the four x-lerps, at (y0,z0), (y1,z0), (y0,z1), (y1,z1)
        if constexpr (DT == BIRP_DTYPE_BF16) {
            c00 = flerp(bf16lo(o.x), bf16hi(o.x), fx);
            c10 = flerp(bf16lo(o.y), bf16hi(o.y), fx);
            c01 = flerp(bf16lo(o.z), bf16hi(o.z), fx);
            c11 = flerp(bf16lo(o.w), bf16hi(o.w), fx);
        } else {
            c00 = xlerp_f16(o.x, fx);
            c10 = xlerp_f16(o.y, fx);
            c01 = xlerp_f16(o.z, fx);
            c11 = xlerp_f16(o.w, fx);
        }
        float val = flerp(flerp(c00, c10, fy), flerp(c01, c11, fy), fz);
        if constexpr (DT == BIRP_DTYPE_F16) val *= v.bscale[off >> 13];
        return val;
    }
}

// ------------------------------------------------------------------------------------------
// CMEM::GetSample (Cmem.cpp:117-186) in the fast form: FP32 with the hardware exp2/log2, square
// root and reciprocal, cosines and sines of the angles as coordinate ratios and half-angle
// identities, and arc cosines from Abramowitz & Stegun 4.4.46 (|error| <= 2e-8). Zone logic as the
// reference's. A cusp angle whose cosine rounds past 1 gives 0 here (the reference: NaN).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float acos_fast(float x) {
    const float ax = fminf(fabsf(x), 1.0f);
    float p = -0.0012624911f;
    p = __builtin_fmaf(p, ax, 0.0066700901f);
    p = __builtin_fmaf(p, ax, -0.0170881256f);
    p = __builtin_fmaf(p, ax, 0.0308918810f);
    p = __builtin_fmaf(p, ax, -0.0501743046f);
    p = __builtin_fmaf(p, ax, 0.0889789874f);
    p = __builtin_fmaf(p, ax, -0.2145988016f);
    p = __builtin_fmaf(p, ax, 1.5707963050f);
    const float r = __builtin_amdgcn_sqrtf(1.0f - ax) * p;
    return x >= 0.f ? r : 3.14159265358979f - r;
}
__device__ __forceinline__ float pow_fast(float b, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(b)); }
__device__ __forceinline__ float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896f); }

__device__ __forceinline__ float cmem_fast(const CmemParams& m, float x, float y, float z) {
    const float ryz2 = y * y + z * z;
    const float r = __builtin_amdgcn_sqrtf(__builtin_fmaf(x, x, ryz2)), ryz = __builtin_amdgcn_sqrtf(ryz2);
    const float ir = __builtin_amdgcn_rcpf(r), iyz = __builtin_amdgcn_rcpf(ryz);
    // theta = acos(x / r), 0 when x == 0; phi = acos(y / |(y, z)|), 0 when y == 0 (Cmem.cpp:224-246)
    const bool xs = x != 0.f, ys = y != 0.f;
    const float ct = xs ? x * ir : 1.f, st = xs ? ryz * ir : 0.f;
    const float theta = xs ? acos_fast(ct) : 0.f;
    const float cphi = ys ? y * iyz : 1.f, sphi = ys ? fabsf(z) * iyz : 0.f;
    // LinScaled (Cmem.cpp:188-222): cos(phi - pi/2) = sin(phi), cos(phi - 3 pi/2) = -sin(phi)
    const float phi_n = acos_fast(ct * m.cos_theta_n + st * m.sin_theta_n * sphi);
    const float phi_s = acos_fast(ct * m.cos_theta_s - st * m.sin_theta_s * sphi);
    const float cos_half = __builtin_amdgcn_sqrtf(0.5f * (1.f + ct));  // cos(theta / 2), theta in [0, pi]
    const float sin_2t = 2.f * st * ct;
    const float f = pow_fast(cos_half + m.a5 * sin_2t * (1.f - exp_fast(-theta)),
                             m.p1 * (m.beta_c[0] + m.beta_c[1] * cphi + m.beta_c[3] * (sphi * sphi)));
    const float q = m.p2 * m.charlie * (exp_fast(m.p3 * m.dn * pow_fast(phi_n, m.a21)) + exp_fast(m.p3 * m.ds * pow_fast(phi_s, m.a21)));
    const float rmp = m.p0 * m.r0_lin * f + q;
    const float h = 0.025f;
    if (r < rmp - h) return 0.f;
    // ShueModel (Cmem.cpp:280-288)
    const float base = 2.f * __builtin_amdgcn_rcpf(1.f + ct);
    const float ry = m.bs * pow_fast(base, m.ay_bs);
    const float rz = (m.az_bs == m.ay_bs) ? ry : m.bs * pow_fast(base, m.az_bs);
    const float rzc = rz * cphi, rys = ry * sphi;
    const float rbs = ry * rz * __builtin_amdgcn_rsqf(rzc * rzc + rys * rys);
    // the zones of Cmem.cpp:132-176 (r >= rmp - h here)
    const bool z_mp = r <= rmp + h;
    const bool z_sheath = !z_mp && r < rbs - h;
    const bool z_bs = !z_mp && !z_sheath && r >= rbs - h && r <= rbs + h;
    const float sd = 0.1f * r;
    const float outer = m.A2 * __builtin_amdgcn_rcpf(sd * sd * sd);
    if (!(z_mp || z_sheath || z_bs)) return outer;
    const float ht = 0.5f * theta, ht2 = ht * ht;
    const float inner = m.A1 * exp_fast(-m.B * (ht2 * ht2)) * pow_fast(sd, -m.alpha - m.beta * (st * st));
    if (z_mp) return ((r - (rmp - h)) * __builtin_amdgcn_rcpf(rmp + h)) * inner;  // lerp(0, inner, t)
    if (z_sheath) return inner;
    return __builtin_fmaf((r - (rbs - h)) * __builtin_amdgcn_rcpf(rbs + h), outer - inner, inner);
}

// Keeps the FP32 march x(j) = RN(j d + i0), j = 0 .. last, inside [0, hi]. x(j) is one rounding of
// a linear function of j, so it is monotone in j and stays inside when x(0) and x(last) do. When
// either is outside (a ray end within FP32 rounding of a face), the start is clamped and the step
// recomputed, then shortened an ulp at a time until x(last) fits; the last resort is a zero step.
__device__ __forceinline__ void fit_march(float& i0, float& d, float hi, int last) {
    const float fl = (float)last;
    const float b = __builtin_fmaf(fl, d, i0);
    if (i0 >= 0.f && i0 <= hi && b >= 0.f && b <= hi) return;
    const float a = fminf(fmaxf(i0, 0.f), hi);
    const float bc = fminf(fmaxf(b, 0.f), hi);
    i0 = a;
    d = last > 0 ? (bc - a) / fl : 0.f;
    for (int it = 0; it < 8 && d != 0.f; ++it) {
        const float e = __builtin_fmaf(fl, d, i0);
        if (e >= 0.f && e <= hi) return;
        d = __uint_as_float(__float_as_uint(d) - 1u);  // one ulp toward zero
    }
    const float e = __builtin_fmaf(fl, d, i0);
    if (!(e >= 0.f && e <= hi)) d = 0.f;
}

// x folded into a Kahan sum (sum, nc), nc holding the negated compensation -c:
// y = x - c; t = sum + y; -c' = y - (t - sum); sum = t.
__device__ __forceinline__ void kahan_add(float x, float& sum, float& nc) {
    const float y = x + nc;
    const float t = sum + y;
    nc = y - (t - sum);
    sum = t;
}

// The lanes of a fast-kernel wave: a wave traces a tile of tw x th rays, S lanes per ray.
// Positive kernels put 2x2 pixel blocks on lane quads, the S lanes of a ray on S consecutive
// quads of a 16-lane row (as exact_lane: a quad's record loads then mostly share one cache
// line); signed kernels keep the S lanes of a ray adjacent, since their first-negative test
// reads the ray's S bits of one ballot as a group.
struct FastLane {
    int tw, th, ray, ti, tj, sub, sub_step;
};
template <int S, bool SIGNED>
__device__ __forceinline__ FastLane fast_lane(int lane, int tw_log2) {
    FastLane f;
    f.tw = 1 << tw_log2;
    f.th = (64 / S) >> tw_log2;
    const bool blocks = !SIGNED && f.tw >= 2 && f.th >= 2;
    f.ray = lane / S;
    if (blocks) {
        const int r16 = lane & 15, quad = r16 >> 2;
        f.sub = quad & (S - 1);
        const int blk = (lane >> 4) * (4 / S) + quad / S;
        f.ti = (blk & ((f.tw >> 1) - 1)) * 2 + (lane & 1);
        f.tj = (blk >> (tw_log2 - 1)) * 2 + ((lane >> 1) & 1);
    } else {
        f.sub = lane % S;
        f.ti = f.ray & (f.tw - 1);
        f.tj = f.ray >> tw_log2;
    }
    f.sub_step = blocks ? 4 : 1;  // lane distance between sub-lanes m and m + 1 of a ray
    return f;
}

// One wave renders tile `tile` (row-major over the image's tiles) of pose P over volume V, whose
// offset tables the wave holds in `tab`; every lane of the wave calls it.
template <int KV, int S, int U>
__device__ __forceinline__ void fast_tile(const DevVolume& V, const FastPose& P, int tile, const unsigned* tab, const FastLane& L) {
    constexpr bool CMEM = KV == BIRP_KDT_CMEM;
    constexpr int DT = CMEM ? BIRP_DTYPE_F32 : (KV & 3);
    constexpr bool SIGNED = CMEM || (KV & BIRP_KDT_SIGNED) != 0;
    const int tw = L.tw, th = L.th, ray = L.ray, ti = L.ti, tj = L.tj, sub = L.sub, sub_step = L.sub_step;
    const float K = (float)BIRP_SCALE_FOLDED;
    const __amdgpu_buffer_rsrc_t rs = (DT == BIRP_DTYPE_F32) ? make_rsrc(V.quad, V.quad_bytes) : make_rsrc(V.oct, V.oct_bytes);
    const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
    const int i = tx * tw + ti;
    const int j = ty * th + tj;
    const bool valid = (i < P.W) && (j < P.H);

    int kf = P.n, kl = -1;
    float ix0 = 0.f, iy0 = 0.f, iz0 = 0.f, dix = 0.f, diy = 0.f, diz = 0.f;
    if (valid) {
        // FP64 ray (the reference's formulas, in double)
        double sx = (2 * ((i + 0.5) * P.inv_w) - 1) * P.tanx_aspect;
        double sy = (1 - 2 * ((j + 0.5) * P.inv_h)) * P.tany;
        // (divisions by a norm or a spacing as products with its FP64 reciprocal: a relative
        // change of ~1e-16, far below the FP32 march)
        const double ind = 1.0 / sqrt(sx * sx + sy * sy + 1.0);
        double u0 = sx * ind, u1 = sy * ind, u2 = -ind;
        double w0 = P.rinv[0] * u0 + P.rinv[1] * u1 + P.rinv[2] * u2;
        double w1 = P.rinv[3] * u0 + P.rinv[4] * u1 + P.rinv[5] * u2;
        double w2 = P.rinv[6] * u0 + P.rinv[7] * u1 + P.rinv[8] * u2;
        const double iwd = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
        double dx = w0 * iwd, dy = w1 * iwd, dz = w2 * iwd;
        if constexpr (CMEM) {
            // the model has no bounds: every sample, marched in world coordinates
            kf = 0;
            kl = P.n - 1;
            const double dt = P.dt, t0 = 0.5 * dt;
            ix0 = (float)(P.px + dx * t0);
            iy0 = (float)(P.py + dy * t0);
            iz0 = (float)(P.pz + dz * t0);
            dix = (float)(dx * dt);
            diy = (float)(dy * dt);
            diz = (float)(dz * dt);
        } else {
            // FP64 cube entry/exit: inside iff 0 <= (x - o)/s < n - 1 per axis
            double tin = -1e300, tout = 1e300;
            slab_axis(P.px, dx, (double)V.ox, (double)V.ox + (double)(V.nx - 1) * V.sx, tin, tout);
            slab_axis(P.py, dy, (double)V.oy, (double)V.oy + (double)(V.ny - 1) * V.sy, tin, tout);
            slab_axis(P.pz, dz, (double)V.oz, (double)V.oz + (double)(V.nz - 1) * V.sz, tin, tout);
            if (tin <= tout) {
                const double sc = P.sc;
                double a = ceil(tin * sc - 0.5), b = ceil(tout * sc - 0.5) - 1.0;
                a = fmax(a, 0.0);
                b = fmin(b, (double)(P.n - 1));
                if (a <= b) {
                    kf = (int)a;
                    kl = (int)b;
                }
            }
            if (P.semantics == BIRP_SEM_REFERENCE && kf != 0) kl = -1;
            if (kf <= kl) {
                const double dt = P.dt, t0 = (kf + 0.5) * dt;
                ix0 = (float)((P.px + dx * t0 - V.ox) * V.rdx);
                iy0 = (float)((P.py + dy * t0 - V.oy) * V.rdy);
                iz0 = (float)((P.pz + dz * t0 - V.oz) * V.rdz);
                dix = (float)(dx * dt * V.rdx);
                diy = (float)(dy * dt * V.rdy);
                diz = (float)(dz * dt * V.rdz);
            }
        }
    }
    const int cnt = (kl >= kf) ? (kl - kf + 1) : 0;
    const int my = (cnt > sub) ? (cnt - sub + S - 1) / S : 0;  // samples of this sub-lane
    const int trips = wave_max_i(my);
    if (!CMEM && cnt > 0) {
        // every sample inside [0, nextbelow(n - 1)] per axis, so that the cell index is the
        // truncation (0 .. n-2) with no clamp per sample; a sample rounded onto the far face
        // keeps weight ~1 on the last grid point
        fit_march(ix0, dix, __uint_as_float(__float_as_uint(V.fnx - 1.0f) - 1u), cnt - 1);
        fit_march(iy0, diy, __uint_as_float(__float_as_uint(V.fny - 1.0f) - 1u), cnt - 1);
        fit_march(iz0, diz, __uint_as_float(__float_as_uint(V.fnz - 1.0f) - 1u), cnt - 1);
    }
    // the U samples of an iteration (neighbours on the ray, of similar size) are added plainly,
    // and their partial sum enters the running sum by one Kahan step. The trips that every lane
    // of the wave fills completely need no clamp and no mask; the rest of the trips run below
    // with both (the same samples in the same order and grouping either way).
    float sum = 0.f, nc = 0.f;
    if constexpr (SIGNED) {
        // Camera.cpp:207-210: a ray ends at its first negative sample. The S lanes of a ray take
        // samples q*S + sub in trip q, so the ray's first negative sample of a trip is the lowest
        // set bit of its S-bit group in the wave's ballot: a lane adds its sample unless its
        // group has a negative sample at or before it, and the ray dies with any in the group.
        const int gshift = ray * S;
        bool alive = true;  // the same in all S lanes of a ray
        for (int q = 0; q < trips; ++q) {
            const int js = q * S + sub;
            const bool act = alive && js < cnt;
            const float fj = (float)min(js, cnt - 1);
            const float xd = __builtin_fmaf(fj, dix, ix0);
            const float yd = __builtin_fmaf(fj, diy, iy0);
            const float zd = __builtin_fmaf(fj, diz, iz0);
            float x;
            if constexpr (CMEM) x = cmem_fast(V.cmem, xd, yd, zd);
            else x = fast_sample<DT>(V, rs, tab, xd, yd, zd);
            const unsigned long long neg = __ballot(act && x < 0.f);
            const unsigned grp = (unsigned)(neg >> gshift) & ((1u << S) - 1u);
            const bool cut = (grp & ((2u << sub) - 1u)) != 0u;
            kahan_add((act && !cut) ? x : 0.f, sum, nc);
            alive = alive && grp == 0u;
            if (__ballot(alive && (q + 1) * S < cnt) == 0ull) break;  // every ray of the wave is done
        }
    }
    const int tfull = SIGNED ? 0 : wave_min_i(my) / U * U;
    int q = SIGNED ? trips : 0;
    for (; q < tfull; q += U) {
        float part = 0.f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float fj = (float)((q + u) * S + sub);
            const float xd = __builtin_fmaf(fj, dix, ix0);
            const float yd = __builtin_fmaf(fj, diy, iy0);
            const float zd = __builtin_fmaf(fj, diz, iz0);
            const float x = fast_sample<DT>(V, rs, tab, xd, yd, zd);
            part = (u == 0) ? x : part + x;
        }
        kahan_add(part, sum, nc);
    }
    for (; q < trips; q += U) {
        float part = 0.f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int js = (q + u) * S + sub;  // sample offset from kf
            const bool act = js < cnt;
            // a lane past its ray's last sample re-reads that sample (its result is masked); an
            // empty ray has i0 = d = 0 and samples the grid origin
            const float fj = (float)min(js, cnt - 1);
            const float xd = __builtin_fmaf(fj, dix, ix0);
            const float yd = __builtin_fmaf(fj, diy, iy0);
            const float zd = __builtin_fmaf(fj, diz, iz0);
            const float x = fast_sample<DT>(V, rs, tab, xd, yd, zd);
            const float xa = act ? x : 0.f;
            part = (u == 0) ? xa : part + xa;
        }
        kahan_add(part, sum, nc);
    }
    // each sub-lane's Kahan total, then the S partials combined in sub order; the step weight
    // and the scale constant are applied once per ray (all in FP32)
    float tot = sum + nc;
    if constexpr (S > 1) {
        float t1 = __shfl_xor(tot, sub_step, 64);
        tot = (sub & 1) ? t1 + tot : tot + t1;
        if constexpr (S == 4) {
            float t2 = __shfl_xor(tot, 2 * sub_step, 64);
            tot = (sub & 2) ? t2 + tot : tot + t2;
        }
    }
    if (valid && sub == 0) P.out[(size_t)j * P.W + i] = (tot * P.w) * K;
}

// Work distribution: items (pose-major, tiles row-major) are split into 8 contiguous ranges, one
// per XCD queue. Block b starts on queue b % 8 (dispatch places consecutive blocks on successive
// XCDs, so a queue's tiles -- neighbouring rays of one image -- meet in one L2) and moves on to
// the other queues when its own is drained. Every wave leaves after finding all 8 queues empty.
// KV is a kernel variant (fast_kernel_dtype): the storage kernel in its low bits, plus
// BIRP_KDT_SIGNED for volumes with values <= 0, whose rays stop at their first negative sample.
// Each wave keeps its volume's offset tables in tab_stride words of dynamic LDS, sized per launch
// to the largest table it reads: 768 words for a 256^3 cube, 1536 for 512^3. With the tables sized
// so, six blocks fit a CU where fixed 2048-word tables let five (the F16 orbit batch: 2.69 -> 2.51
// ms), and the launch bound keeps the registers within six waves per SIMD.
template <int KV, int S, int U>
__global__ __launch_bounds__(BIRP_BLOCK, 6) void k_render_fast(const DevVolume* __restrict__ vols, const FastPose* __restrict__ poses,
                                                               const int* __restrict__ prefix, int npose, int total,
                                                               unsigned* counters, int tw_log2, int tab_stride) {
    constexpr bool CMEM = KV == BIRP_KDT_CMEM;
    constexpr int DT = CMEM ? BIRP_DTYPE_F32 : (KV & 3);
    constexpr bool SIGNED = CMEM || (KV & BIRP_KDT_SIGNED) != 0;
    static_assert(!SIGNED || U == 1, "signed kernels take one sample per lane per trip");
    extern __shared__ unsigned s_tab[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    unsigned* tab = s_tab + wave * tab_stride;
    const FastLane fl = fast_lane<S, SIGNED>(lane, tw_log2);
    int cached_vol = -1;
    int queue = blockIdx.x & 7;
    for (int tried = 0; tried < 8;) {
        const int q_lo = (int)(((long long)total * queue) >> 3), q_hi = (int)(((long long)total * (queue + 1)) >> 3);
        int item = 0;
        if (lane == 0) item = q_lo + (int)atomicAdd(&counters[queue * 32], 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= q_hi) {
            queue = (queue + 1) & 7;
            tried++;
            continue;
        }
        // pose = last p with prefix[p] <= item
        int lo = 0, hi = npose - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= item) lo = mid; else hi = mid - 1;
        }
        const int pidx = __builtin_amdgcn_readfirstlane(lo);
        const FastPose& P = poses[pidx];
        const DevVolume& V = vols[P.vol];
        if (!CMEM && P.vol != cached_vol) {
            const unsigned* src = (DT == BIRP_DTYPE_F32) ? V.qtab : V.otab;
            const int ntab = min(V.nx + V.ny + V.nz, tab_stride);  // the host sizes tab_stride to cover it
            for (int q = lane; q < ntab; q += 64) tab[q] = src[q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            cached_vol = P.vol;
        }
        const int tile = item - prefix[pidx];
        fast_tile<KV, S, U>(V, P, tile, tab, fl);
    }
}

// One pose as a plain grid, with the pose and its volume by value: the launch needs no descriptor
// upload and no work counters (single fast renders, a lone served request). Wave w of block b
// renders tile 4 b' + w, where blocks go round-robin over the 8 XCDs and b' numbers them XCD by
// XCD, so that each XCD renders a contiguous range of neighbouring tiles, as the work queues of
// k_render_fast do. The grid is a multiple of 8 blocks; waves past the last tile leave at once.
template <int KV, int S, int U>
__global__ __launch_bounds__(BIRP_BLOCK, 6) void k_render_fast_one(DevVolume V, FastPose P, int total, int tw_log2, int tab_stride) {
    constexpr bool CMEM = KV == BIRP_KDT_CMEM;
    constexpr int DT = CMEM ? BIRP_DTYPE_F32 : (KV & 3);
    constexpr bool SIGNED = CMEM || (KV & BIRP_KDT_SIGNED) != 0;
    static_assert(!SIGNED || U == 1, "signed kernels take one sample per lane per trip");
    extern __shared__ unsigned s_tab[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int b = (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
    const int item = b * (BIRP_BLOCK / 64) + wave;
    if (item >= total) return;
    unsigned* tab = s_tab + wave * tab_stride;
    if constexpr (!CMEM) {
        const unsigned* src = (DT == BIRP_DTYPE_F32) ? V.qtab : V.otab;
        const int ntab = min(V.nx + V.ny + V.nz, tab_stride);  // the host sizes tab_stride to cover it
        for (int q = lane; q < ntab; q += 64) tab[q] = src[q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    fast_tile<KV, S, U>(V, P, item, tab, fast_lane<S, SIGNED>(lane, tw_log2));
}

// ------------------------------------------------------------------------------------------
// the fixed scratch of birp_render_views_noalloc
// ------------------------------------------------------------------------------------------
// Each GPU holds its own copy for the life of the process. In host code `scratch` names a host-side
// stand-in, not a device address: host code uses only the address noalloc_scratch_address returns.
static __device__ __align__(256) unsigned char scratch[BIRP_NOALLOC_SCRATCH_BYTES];

static_assert(sizeof(DevVolume) <= BIRP_NOALLOC_POSE_OFF, "noalloc scratch: volume slot");
static_assert(BIRP_NOALLOC_SCRATCH_BYTES <= BIRP_NOALLOC_SCRATCH_KIB * 1024, "noalloc scratch: stated size");
static_assert(sizeof(FastPose) <= 256, "noalloc scratch: pose slot");
static_assert(3 * (BIRP_NOALLOC_MAX_VIEWS + 1) * 4 <= BIRP_NOALLOC_CTR_OFF - BIRP_NOALLOC_PREFIX_OFF, "noalloc scratch: prefix");

struct NoallocHead {
    DevVolume v;
    int prefix[3 * (BIRP_NOALLOC_MAX_VIEWS + 1)];
};
struct NoallocPoses {
    FastPose p[4];
    int first, count;
};

__global__ __launch_bounds__(BIRP_BLOCK) void k_noalloc_head(NoallocHead h, unsigned char* s) {
    if (threadIdx.x == 0) *reinterpret_cast<DevVolume*>(s) = h.v;
    int* pre = reinterpret_cast<int*>(s + BIRP_NOALLOC_PREFIX_OFF);
    for (int i = threadIdx.x; i < 3 * (BIRP_NOALLOC_MAX_VIEWS + 1); i += BIRP_BLOCK) pre[i] = h.prefix[i];
    unsigned* ctr = reinterpret_cast<unsigned*>(s + BIRP_NOALLOC_CTR_OFF);
    for (int i = threadIdx.x; i < 3 * 256; i += BIRP_BLOCK) ctr[i] = 0u;
}

__global__ void k_noalloc_poses(NoallocPoses c, unsigned char* s) {
    FastPose* dst = reinterpret_cast<FastPose*>(s + BIRP_NOALLOC_POSE_OFF) + c.first;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (threadIdx.x == q && q < c.count) dst[q] = c.p[q];
}

int noalloc_scratch_address(void** out) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return BIRP_ERR_DEVICE;
    }
    *out = p;
    return BIRP_OK;
}

int noalloc_stage_head(const DevVolume& v, const int* prefix, void* s, void* stream) {
    NoallocHead h;
    h.v = v;
    for (int i = 0; i < 3 * (BIRP_NOALLOC_MAX_VIEWS + 1); i++) h.prefix[i] = prefix[i];
    hipLaunchKernelGGL(k_noalloc_head, dim3(1), dim3(BIRP_BLOCK), 0, (hipStream_t)stream, h, (unsigned char*)s);
    return hipGetLastError() == hipSuccess ? BIRP_OK : BIRP_ERR_DEVICE;
}

int noalloc_stage_poses(const FastPose* poses, int first, int count, void* s, void* stream) {
    if (count < 1 || count > 4 || first < 0 || first + count > BIRP_NOALLOC_MAX_VIEWS) return BIRP_ERR_INVALID_ARGUMENT;
    NoallocPoses c;
    for (int q = 0; q < count; q++) c.p[q] = poses[q];
    c.first = first;
    c.count = count;
    hipLaunchKernelGGL(k_noalloc_poses, dim3(1), dim3(64), 0, (hipStream_t)stream, c, (unsigned char*)s);
    return hipGetLastError() == hipSuccess ? BIRP_OK : BIRP_ERR_DEVICE;
}

// ------------------------------------------------------------------------------------------
// host-side launch wrappers
// ------------------------------------------------------------------------------------------
static int hip_status(hipError_t e) { return e == hipSuccess ? BIRP_OK : BIRP_ERR_DEVICE; }

int exact_mode_of(const DevVolume& v, const ExactView& ev) {
    if (v.kind == SK_CMEM) return SM_CMEM;
    if (v.interp == BIRP_INTERP_NEAREST) return SM_LIN_NEAREST;
    if (v.vclass == VC_NONFINITE || !v.quad || !v.qtab) return SM_LIN_TRI;
    if (v.vclass == VC_POSITIVE && v.origin_far && ev.fast_guards) return SM_QUAD_POS;
    return SM_QUAD_GEN;
}

size_t exact_lds_bytes(const DevVolume& v, int n, int mode) {
    size_t lds = (size_t)((n + 3) & ~3) * sizeof(float);
    if (mode == SM_QUAD_POS || mode == SM_QUAD_GEN) lds += (size_t)(v.nx + v.ny + v.nz) * sizeof(unsigned);
    return lds;
}

template <int S>
static int launch_exact_s(const DevVolume& v, const ExactView& ev, int pipe, hipStream_t st, const ExactOrder* ord) {
    constexpr int TW = ExactTile<S>::TW, TH = ExactTile<S>::TH;
    int wg_x = (ev.W + 2 * TW - 1) / (2 * TW);
    int wg_y = (ev.H + 2 * TH - 1) / (2 * TH);
    int nblocks = wg_x * wg_y;
    const int grid = nblocks;
    const int mode = exact_mode_of(v, ev);
    const size_t lds = exact_lds_bytes(v, ev.n, mode);
    if (lds > 64 * 1024) return BIRP_ERR_UNSUPPORTED;
    const int* order = nullptr;
    int ncu = 0;
    if (ord && ord->order) {
        if (nblocks > BIRP_PLAN_MAX_BLOCKS || ord->ncu <= 0 || mode == SM_CMEM) return BIRP_ERR_INVALID_ARGUMENT;
        if (ord->replan) {
            hipLaunchKernelGGL((k_exact_plan<S>), dim3((4 * nblocks + BIRP_BLOCK - 1) / BIRP_BLOCK), dim3(BIRP_BLOCK), 0, st, v, ev,
                               wg_x, nblocks, ord->order);
            const int rc = hip_status(hipGetLastError());
            if (rc) return rc;
        }
        order = ord->order;
        ncu = ord->ncu;
    }
    switch (mode) {
        case SM_QUAD_POS:
            if (pipe >= 4) hipLaunchKernelGGL((k_render_exact<S, SM_QUAD_POS, 4>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu);
            else if (pipe == 3) hipLaunchKernelGGL((k_render_exact<S, SM_QUAD_POS, 3>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu);
            else if (pipe == 1) hipLaunchKernelGGL((k_render_exact<S, SM_QUAD_POS, 1>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu);
            else hipLaunchKernelGGL((k_render_exact<S, SM_QUAD_POS, 2>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu);
            break;
        case SM_QUAD_GEN: hipLaunchKernelGGL((k_render_exact<S, SM_QUAD_GEN>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu); break;
        case SM_LIN_TRI: hipLaunchKernelGGL((k_render_exact<S, SM_LIN_TRI>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu); break;
        case SM_LIN_NEAREST: hipLaunchKernelGGL((k_render_exact<S, SM_LIN_NEAREST>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu); break;
        default: hipLaunchKernelGGL((k_render_exact<S, SM_CMEM>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, ev, wg_x, nblocks, order, ncu); break;
    }
    return hip_status(hipGetLastError());
}

template <int S>
static int exact_blocks(int W, int H) {
    constexpr int TW = ExactTile<S>::TW, TH = ExactTile<S>::TH;
    return ((W + 2 * TW - 1) / (2 * TW)) * ((H + 2 * TH - 1) / (2 * TH));
}

int exact_tiles(int W, int H, int lanes_per_ray) {
    if (lanes_per_ray == 16) return exact_blocks<16>(W, H);
    if (lanes_per_ray == 8) return exact_blocks<8>(W, H);
    return lanes_per_ray == 4 ? exact_blocks<4>(W, H) : (lanes_per_ray == 1 ? exact_blocks<1>(W, H) : exact_blocks<2>(W, H));
}

int exact_wave_slots(int W, int H, int lanes_per_ray) { return 4 * exact_tiles(W, H, lanes_per_ray); }

template <int S>
static int launch_exact_multi_s(const DevVolume* vols, const ExactView* views, const int* prefix, int nview, int total, int mode,
                                size_t lds, hipStream_t st, const int* order, int pipe) {
    const int grid = total;
    switch (mode) {
        case SM_QUAD_POS:
            if (pipe >= 4) hipLaunchKernelGGL((k_render_exact_multi<S, SM_QUAD_POS, 4>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, views, prefix, nview, total, order);
            else hipLaunchKernelGGL((k_render_exact_multi<S, SM_QUAD_POS, 2>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, views, prefix, nview, total, order);
            break;
        case SM_QUAD_GEN: hipLaunchKernelGGL((k_render_exact_multi<S, SM_QUAD_GEN>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, views, prefix, nview, total, order); break;
        case SM_LIN_TRI: hipLaunchKernelGGL((k_render_exact_multi<S, SM_LIN_TRI>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, views, prefix, nview, total, order); break;
        case SM_LIN_NEAREST: hipLaunchKernelGGL((k_render_exact_multi<S, SM_LIN_NEAREST>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, views, prefix, nview, total, order); break;
        default: return BIRP_ERR_INVALID_ARGUMENT;
    }
    return hip_status(hipGetLastError());
}

int launch_exact_multi(const DevVolume* vols, const ExactView* views, const int* prefix, int nview, int total, int mode,
                       size_t lds, void* stream, const int* order, int lanes, int pipe) {
    if (nview <= 0 || total <= 0) return BIRP_OK;
    if (lds > 64 * 1024) return BIRP_ERR_UNSUPPORTED;
    hipStream_t st = (hipStream_t)stream;
    if (mode == SM_CMEM) {  // always 4 lanes per ray (exact_multi_lanes)
        hipLaunchKernelGGL((k_render_exact_multi<4, SM_CMEM>), dim3(total), dim3(BIRP_BLOCK), lds, st, vols, views, prefix, nview, total, order);
        return hip_status(hipGetLastError());
    }
    if (lanes == 8) return launch_exact_multi_s<8>(vols, views, prefix, nview, total, mode, lds, st, order, pipe);
    if (lanes == 4) return launch_exact_multi_s<4>(vols, views, prefix, nview, total, mode, lds, st, order, pipe);
    if (lanes == 2) return launch_exact_multi_s<2>(vols, views, prefix, nview, total, mode, lds, st, order, pipe);
    return BIRP_ERR_INVALID_ARGUMENT;
}

int launch_plan_multi(const DevVolume* vols, const ExactView* views, const int* prefix, int nview, int total, int mode, int* cost,
                      void* stream, int lanes) {
    if (nview <= 0 || total <= 0 || mode == SM_CMEM || (lanes != 2 && lanes != 4 && lanes != 8)) return BIRP_ERR_INVALID_ARGUMENT;
    const long long threads = 4LL * total;
    const dim3 grid((unsigned)((threads + BIRP_BLOCK - 1) / BIRP_BLOCK));
    if (lanes == 8) hipLaunchKernelGGL((k_plan_multi<8>), grid, dim3(BIRP_BLOCK), 0, (hipStream_t)stream, vols, views, prefix, nview, total, cost);
    else if (lanes == 4) hipLaunchKernelGGL((k_plan_multi<4>), grid, dim3(BIRP_BLOCK), 0, (hipStream_t)stream, vols, views, prefix, nview, total, cost);
    else hipLaunchKernelGGL((k_plan_multi<2>), grid, dim3(BIRP_BLOCK), 0, (hipStream_t)stream, vols, views, prefix, nview, total, cost);
    return hip_status(hipGetLastError());
}

int launch_exact(const DevVolume& v, const ExactView& ev, int lanes_per_ray, int pipe, void* stream, const ExactOrder* ord) {
    if (ev.n < 2 || ev.n > BIRP_MAX_EXACT_SAMPLES || ev.W <= 0 || ev.H <= 0) return BIRP_ERR_INVALID_ARGUMENT;
    hipStream_t st = (hipStream_t)stream;
    if (lanes_per_ray == 16) return launch_exact_s<16>(v, ev, pipe, st, ord);
    if (lanes_per_ray == 8) return launch_exact_s<8>(v, ev, pipe, st, ord);
    if (lanes_per_ray == 4) return launch_exact_s<4>(v, ev, pipe, st, ord);
    if (lanes_per_ray == 2) return launch_exact_s<2>(v, ev, pipe, st, ord);
    return launch_exact_s<1>(v, ev, pipe, st, ord);
}

template <int S>
static int launch_plan_s(const DevVolume& v, const ExactView& ev, int* order, hipStream_t st) {
    constexpr int TW = ExactTile<S>::TW, TH = ExactTile<S>::TH;
    const int wg_x = (ev.W + 2 * TW - 1) / (2 * TW);
    const int nblocks = wg_x * ((ev.H + 2 * TH - 1) / (2 * TH));
    if (nblocks > BIRP_PLAN_MAX_BLOCKS) return BIRP_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL((k_exact_plan<S>), dim3((4 * nblocks + BIRP_BLOCK - 1) / BIRP_BLOCK), dim3(BIRP_BLOCK), 0, st, v, ev, wg_x,
                       nblocks, order);
    return hip_status(hipGetLastError());
}

int launch_exact_plan(const DevVolume& v, const ExactView& ev, int lanes_per_ray, int* order, void* stream) {
    if (ev.n < 2 || ev.W <= 0 || ev.H <= 0 || !order) return BIRP_ERR_INVALID_ARGUMENT;
    hipStream_t st = (hipStream_t)stream;
    if (lanes_per_ray == 16) return launch_plan_s<16>(v, ev, order, st);
    if (lanes_per_ray == 8) return launch_plan_s<8>(v, ev, order, st);
    if (lanes_per_ray == 4) return launch_plan_s<4>(v, ev, order, st);
    if (lanes_per_ray == 2) return launch_plan_s<2>(v, ev, order, st);
    return launch_plan_s<1>(v, ev, order, st);
}

int launch_f64(const DevVolume& v, const F64View& fv, void* stream) {
    int npix = fv.W * fv.H;
    hipLaunchKernelGGL(k_render_f64, dim3((npix + BIRP_BLOCK - 1) / BIRP_BLOCK), dim3(BIRP_BLOCK), 0, (hipStream_t)stream, v, fv);
    return hip_status(hipGetLastError());
}

int launch_space_points(const ExactView& ev, int row0, int rows, float* xyz, void* stream) {
    const long long total = (long long)rows * ev.W * ev.n;
    if (total <= 0) return BIRP_OK;
    hipLaunchKernelGGL(k_space_points, dim3((unsigned)((total + BIRP_BLOCK - 1) / BIRP_BLOCK)), dim3(BIRP_BLOCK), 0, (hipStream_t)stream,
                       ev, row0, rows, xyz);
    return hip_status(hipGetLastError());
}

int launch_space_integrate(const ExactView& ev, int row0, int rows, const float* vals, float* img, void* stream) {
    const long long total = (long long)rows * ev.W;
    if (total <= 0) return BIRP_OK;
    hipLaunchKernelGGL(k_space_integrate, dim3((unsigned)((total + BIRP_BLOCK - 1) / BIRP_BLOCK)), dim3(BIRP_BLOCK), 0,
                       (hipStream_t)stream, ev, row0, rows, vals, img);
    return hip_status(hipGetLastError());
}

int launch_sample_points(const DevVolume& v, const float* xyz, long long n, float* out, void* stream) {
    if (n <= 0) return BIRP_OK;
    long long blocks = (n + BIRP_BLOCK - 1) / BIRP_BLOCK;
    hipLaunchKernelGGL(k_sample_points, dim3((unsigned)blocks), dim3(BIRP_BLOCK), 0, (hipStream_t)stream, v, xyz, n, out);
    return hip_status(hipGetLastError());
}

int launch_scan_class(const float* values, size_t n, unsigned* flags_dev, void* stream) {
    size_t blocks = (n + BIRP_BLOCK - 1) / BIRP_BLOCK;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_scan_class, dim3((unsigned)blocks), dim3(BIRP_BLOCK), 0, (hipStream_t)stream, values, n, flags_dev);
    return hip_status(hipGetLastError());
}

int launch_build_quad(const DevVolume& v, void* quad, void* stream) {
    size_t total = (size_t)(v.nx - 1) * (v.ny - 1) * v.nbz * 9u;
    size_t blocks = (total + BIRP_BLOCK - 1) / BIRP_BLOCK;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_build_quad, dim3((unsigned)blocks), dim3(BIRP_BLOCK), 0, (hipStream_t)stream, v, (unsigned char*)quad);
    return hip_status(hipGetLastError());
}

int launch_build_oct(const DevVolume& v, void* oct, float* bscale, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    int nb = v.nbx * v.nby * v.nbz;
    if (v.dtype == BIRP_DTYPE_F16) {
        hipLaunchKernelGGL(k_brick_scale, dim3(nb), dim3(BIRP_BLOCK), 0, st, v, bscale);
        if (hipGetLastError() != hipSuccess) return BIRP_ERR_DEVICE;
    }
    size_t total = (size_t)(v.nx - 1) * (v.ny - 1) * (v.nz - 1);
    size_t blocks = (total + BIRP_BLOCK - 1) / BIRP_BLOCK;
    if (blocks > 65536) blocks = 65536;
    if (v.dtype == BIRP_DTYPE_F16)
        hipLaunchKernelGGL((k_build_oct<true>), dim3((unsigned)blocks), dim3(BIRP_BLOCK), 0, st, v, (unsigned char*)oct, bscale);
    else
        hipLaunchKernelGGL((k_build_oct<false>), dim3((unsigned)blocks), dim3(BIRP_BLOCK), 0, st, v, (unsigned char*)oct, bscale);
    return hip_status(hipGetLastError());
}

int launch_reset_counter(unsigned* counter, void* stream) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(256), 0, (hipStream_t)stream, counter);
    return hip_status(hipGetLastError());
}

int launch_zero_f32(float* p, size_t n, void* stream) {
    size_t blocks = (n + BIRP_BLOCK - 1) / BIRP_BLOCK;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)blocks), dim3(BIRP_BLOCK), 0, (hipStream_t)stream, p, n);
    return hip_status(hipGetLastError());
}

template <int DT, int S, int U>
static void* fast_kernel_ptr() { return (void*)k_render_fast<DT, S, U>; }

template <int DT, int S>
static int launch_fast_su(int unroll, int grid, hipStream_t st, const DevVolume* vols, const FastPose* poses, const int* prefix,
                          int npose, int total, unsigned* counter, int twl, int ts) {
    const size_t lds = (size_t)(BIRP_BLOCK / 64) * ts * sizeof(unsigned);
    if constexpr ((DT & (BIRP_KDT_SIGNED | BIRP_KDT_CMEM)) != 0) {
        hipLaunchKernelGGL((k_render_fast<DT, S, 1>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, poses, prefix, npose, total, counter, twl, ts);
    } else {
        switch (unroll) {
            case 4: hipLaunchKernelGGL((k_render_fast<DT, S, 4>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, poses, prefix, npose, total, counter, twl, ts); break;
            case 2: hipLaunchKernelGGL((k_render_fast<DT, S, 2>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, poses, prefix, npose, total, counter, twl, ts); break;
            default: hipLaunchKernelGGL((k_render_fast<DT, S, 1>), dim3(grid), dim3(BIRP_BLOCK), lds, st, vols, poses, prefix, npose, total, counter, twl, ts); break;
        }
    }
    return hip_status(hipGetLastError());
}

template <int DT>
static int launch_fast_d(int S, int unroll, int grid, hipStream_t st, const DevVolume* vols, const FastPose* poses,
                         const int* prefix, int npose, int total, unsigned* counter, int twl, int ts) {
    if (S == 4) return launch_fast_su<DT, 4>(unroll, grid, st, vols, poses, prefix, npose, total, counter, twl, ts);
    if (S == 2) return launch_fast_su<DT, 2>(unroll, grid, st, vols, poses, prefix, npose, total, counter, twl, ts);
    return launch_fast_su<DT, 1>(unroll, grid, st, vols, poses, prefix, npose, total, counter, twl, ts);
}

int fast_tab_stride(int dtype, int tab_entries) {
    if (dtype == BIRP_KDT_CMEM) return 0;  // the model has no tables
    return ((std::max(tab_entries, 1) + 255) / 256) * 256;
}

// counters: 256 unsigned (8 queues, 128 B apart), reset by launch_reset_counter before each launch.
int launch_fast_batch(const DevVolume* vols_dev, int dtype, const FastPose* poses_dev, const int* tile_prefix_dev, int npose,
                      int total_tiles, unsigned* work_counter, int tile_w, int tile_h, int lanes_per_ray, int unroll,
                      int grid, void* stream, int tab_entries) {
    int twl = 0;
    while ((1 << twl) < tile_w) twl++;
    if ((1 << twl) != tile_w || tile_w * tile_h * lanes_per_ray != 64) return BIRP_ERR_INVALID_ARGUMENT;
    if (dtype < 0 || dtype >= BIRP_KDT_COUNT) return BIRP_ERR_INVALID_ARGUMENT;
    if (tab_entries < 0 || tab_entries > BIRP_FAST_TAB) return BIRP_ERR_INVALID_ARGUMENT;
    const int ts = fast_tab_stride(dtype, tab_entries);
    grid = ((grid + 7) / 8) * 8;
    hipStream_t st = (hipStream_t)stream;
    if (dtype & (BIRP_KDT_SIGNED | BIRP_KDT_CMEM)) unroll = 1;  // the signed and CMEM kernels take one sample per lane per trip
#define BIRP_LFD(KV) \
    case KV: return launch_fast_d<KV>(lanes_per_ray, unroll, grid, st, vols_dev, poses_dev, tile_prefix_dev, npose, total_tiles, work_counter, twl, ts);
    switch (dtype) {
        BIRP_LFD(0) BIRP_LFD(1) BIRP_LFD(2) BIRP_LFD(3) BIRP_LFD(4) BIRP_LFD(5) BIRP_LFD(6) BIRP_LFD(7) BIRP_LFD(8)
        default: return BIRP_ERR_INVALID_ARGUMENT;
    }
#undef BIRP_LFD
}

template <int DT, int S>
static int launch_one_su(int unroll, int grid, hipStream_t st, const DevVolume& v, const FastPose& p, int total, int twl, int ts) {
    const size_t lds = (size_t)(BIRP_BLOCK / 64) * ts * sizeof(unsigned);
    if constexpr ((DT & (BIRP_KDT_SIGNED | BIRP_KDT_CMEM)) != 0) {
        hipLaunchKernelGGL((k_render_fast_one<DT, S, 1>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, p, total, twl, ts);
    } else {
        switch (unroll) {
            case 4: hipLaunchKernelGGL((k_render_fast_one<DT, S, 4>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, p, total, twl, ts); break;
            case 2: hipLaunchKernelGGL((k_render_fast_one<DT, S, 2>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, p, total, twl, ts); break;
            default: hipLaunchKernelGGL((k_render_fast_one<DT, S, 1>), dim3(grid), dim3(BIRP_BLOCK), lds, st, v, p, total, twl, ts); break;
        }
    }
    return hip_status(hipGetLastError());
}

template <int DT>
static int launch_one_d(int S, int unroll, int grid, hipStream_t st, const DevVolume& v, const FastPose& p, int total, int twl, int ts) {
    if (S == 4) return launch_one_su<DT, 4>(unroll, grid, st, v, p, total, twl, ts);
    if (S == 2) return launch_one_su<DT, 2>(unroll, grid, st, v, p, total, twl, ts);
    return launch_one_su<DT, 1>(unroll, grid, st, v, p, total, twl, ts);
}

int launch_fast_one(const DevVolume& v, int dtype, const FastPose& pose, int total_tiles, int tile_w, int tile_h, int lanes_per_ray,
                    int unroll, void* stream, int tab_entries) {
    int twl = 0;
    while ((1 << twl) < tile_w) twl++;
    if ((1 << twl) != tile_w || tile_w * tile_h * lanes_per_ray != 64) return BIRP_ERR_INVALID_ARGUMENT;
    if (dtype < 0 || dtype >= BIRP_KDT_COUNT) return BIRP_ERR_INVALID_ARGUMENT;
    if (tab_entries < 0 || tab_entries > BIRP_FAST_TAB || total_tiles < 0) return BIRP_ERR_INVALID_ARGUMENT;
    if (total_tiles == 0) return BIRP_OK;
    const int ts = fast_tab_stride(dtype, tab_entries);
    const int blocks = (int)(((long long)total_tiles + BIRP_BLOCK / 64 - 1) / (BIRP_BLOCK / 64));
    const int grid = ((blocks + 7) / 8) * 8;
    hipStream_t st = (hipStream_t)stream;
    if (dtype & (BIRP_KDT_SIGNED | BIRP_KDT_CMEM)) unroll = 1;  // the signed and CMEM kernels take one sample per lane per trip
#define BIRP_LOD(KV) \
    case KV: return launch_one_d<KV>(lanes_per_ray, unroll, grid, st, v, pose, total_tiles, twl, ts);
    switch (dtype) {
        BIRP_LOD(0) BIRP_LOD(1) BIRP_LOD(2) BIRP_LOD(3) BIRP_LOD(4) BIRP_LOD(5) BIRP_LOD(6) BIRP_LOD(7) BIRP_LOD(8)
        default: return BIRP_ERR_INVALID_ARGUMENT;
    }
#undef BIRP_LOD
}

int fast_grid_for(int dtype, int lanes_per_ray, int unroll, int tiles, int tab_entries) {
    // back-to-back single renders with the tuned tiles: 576x336 at n = 800 took 0.10-0.11 ms per
    // render at one tile per wave, 0.12-0.13 ms at two waves per tile or on the full grid; 144x84
    // at n = 160 took 0.033-0.036 ms against 0.086 ms on the full grid
    return std::min(fast_grid_size(dtype, lanes_per_ray, unroll, tab_entries), (tiles + 3) / 4 + 1);
}

int fast_grid_size(int dtype, int lanes_per_ray, int unroll, int tab_entries) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    // cached per device, kernel variant and table size: the occupancy query costs microseconds
    if (dtype & (BIRP_KDT_SIGNED | BIRP_KDT_CMEM)) unroll = 1;
    const int ts = fast_tab_stride(dtype, std::min(std::max(tab_entries, 0), BIRP_FAST_TAB));
    const int tsb = ts / 256;  // 0..8
    static std::atomic<int> cache[16][BIRP_KDT_COUNT][5][5][BIRP_FAST_TAB / 256 + 1];
    const bool cacheable = dev >= 0 && dev < 16 && dtype >= 0 && dtype < BIRP_KDT_COUNT && lanes_per_ray <= 4 && unroll <= 4;
    if (cacheable) {
        const int c = cache[dev][dtype][lanes_per_ray][unroll][tsb].load(std::memory_order_relaxed);
        if (c > 0) return c;
    }
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
    int per_cu = 0;
    void* fn = nullptr;
#define BIRP_FK(DT)                                                                          \
    if (lanes_per_ray == 4) fn = unroll == 4 ? fast_kernel_ptr<DT, 4, 4>() : unroll == 2 ? fast_kernel_ptr<DT, 4, 2>() : fast_kernel_ptr<DT, 4, 1>(); \
    else if (lanes_per_ray == 2) fn = unroll == 4 ? fast_kernel_ptr<DT, 2, 4>() : unroll == 2 ? fast_kernel_ptr<DT, 2, 2>() : fast_kernel_ptr<DT, 2, 1>(); \
    else fn = unroll == 4 ? fast_kernel_ptr<DT, 1, 4>() : unroll == 2 ? fast_kernel_ptr<DT, 1, 2>() : fast_kernel_ptr<DT, 1, 1>();
#define BIRP_FKS(DT) \
    fn = lanes_per_ray == 4 ? fast_kernel_ptr<DT, 4, 1>() : lanes_per_ray == 2 ? fast_kernel_ptr<DT, 2, 1>() : fast_kernel_ptr<DT, 1, 1>();
    switch (dtype) {
        case 3: BIRP_FK(3) break;
        case 2: BIRP_FK(2) break;
        case 1: BIRP_FK(1) break;
        case 4: BIRP_FKS(4) break;
        case 5: BIRP_FKS(5) break;
        case 6: BIRP_FKS(6) break;
        case 7: BIRP_FKS(7) break;
        case 8: BIRP_FKS(8) break;
        default: BIRP_FK(0) break;
    }
#undef BIRP_FKS
#undef BIRP_FK
    const size_t lds = (size_t)(BIRP_BLOCK / 64) * ts * sizeof(unsigned);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BIRP_BLOCK, lds) != hipSuccess || per_cu <= 0) per_cu = 4;
    if (cacheable) cache[dev][dtype][lanes_per_ray][unroll][tsb].store(cus * per_cu, std::memory_order_relaxed);
    return cus * per_cu;
}

}  // namespace birp

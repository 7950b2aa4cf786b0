// kernels.hpp -- gfx950 kernels of the per-stage hot path.
//
// Layout in HBM: the conserved state is three FP64 plane-sets A, B, C, each
// [4][nx][ny] (x-major, j contiguous), used as RK ping-pong buffers; an
// optional uint8 cell-type plane [nx][ny]; immersed-boundary tables.
//
// LDS holds the staged state and the face fluxes as planes per variable
// (structure of arrays), so the lanes of a wave read consecutive doubles.
//
// Stage kernel (one launch per RK stage): a workgroup owns a TI x TJ cell
// tile (15 x 16 with 256 threads by default; the other tile shapes are
// below).  It stages the tile plus a 3-cell halo (the WENO5 stencil reach,
// minus the unused corners) in LDS, applying the physical boundary on the fly
// (one ghost cell per boundary face, as source/boundary/*Boundary.cpp;
// periodic wrap), then
//   phase 1: every thread computes one or two faces -- the 16x16 = 256
//            x-faces and the 15x17 = 255 y-faces of the tile, 511 in all,
//            two rounds of 256 with one idle slot -- WENO5 left/right
//            states, Euler fluxes and the HLL flux, in the face-normal frame
//            so x and y faces run the same code with no divergence; fluxes
//            go to LDS;
//   phase 2: every thread updates one cell: flux divergence and the TVD-RK
//            combination, written to the next plane.
// Every face is computed exactly once per tile (tile-edge faces once per
// adjacent tile: 511 faces per 240 cells).
//
// Arithmetic: FP64 kernels reproduce the reference's operations one for one
// in its source order (compiled with -ffp-contract=off; IEEE division and
// square root, or forms proven to give the same bits -- WOps), so their
// results equal a strict-IEEE CPU evaluation of the reference source bit for
// bit.  FP32-bulk (R = float) evaluates the reconstruction, fluxes and RHS in
// FP32 with fused multiply-adds and hardware reciprocals; the state, the RK
// combination, the immersed-boundary fill and the CFL bound stay FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace cfs {

constexpr int HALO = 3;            // WENO5 reach: stencil offsets -3..+2

// The state pointers a kernel reads from its StageArgs point to global
// memory.  The compiler infers that for pointers passed as kernel arguments,
// not for pointers read from device memory (the batched kernel's per-problem
// arguments), and then emits flat instructions; G() states it.
template <typename T> using gptr = __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> G(T* p) { return (gptr<T>)p; }

// Tile shape: a workgroup of NT threads owns TI x TJ cells (TI rows along x /
// i, TJ columns along y / j, j contiguous in memory) and stages SR x SC cells
// (the tile plus the halo) in LDS.  Its (TI + 1) TJ x-faces and TI (TJ + 1)
// y-faces take two rounds of NT threads; every cell update one thread.
template <int TI_, int TJ_, int NT_, bool DUAL_ = false> struct TileShape {
    static constexpr int TI = TI_, TJ = TJ_, NT = NT_;
    // DUAL: each thread computes its x-face and its y-face in one pass (two
    // independent dependency chains per thread: for grids too small to give
    // every SIMD enough waves, where the per-face latency, not the issue
    // rate, bounds the stage)
    static constexpr bool DUAL = DUAL_;
    static constexpr int SR = TI + 2 * HALO, SC = TJ + 2 * HALO;
    static constexpr int NXF = (TI + 1) * TJ, NYF = TI * (TJ + 1);
    static_assert(NXF + NYF <= 2 * NT && TI * TJ <= NT && NT % 64 == 0, "two face rounds, one cell per thread");
    static_assert(NXF >= NYF, "the y-faces fit in the x-faces' flux planes");
    // flux plane length in LDS: every face, or (HALF) the x-faces only, which
    // the y-faces overwrite once each cell holds its x part of the divergence
    template <bool HALF> static constexpr int flux_plane() { return HALF ? NXF : NXF + NYF; }
};
using Shape0 = TileShape<15, 16, 256>;    // 240 cells, 511 faces (256 + 255)
using Shape1 = TileShape<15, 32, 512>;    // 480 cells, 1007 faces (512 + 495)
using Shape2 = TileShape<31, 32, 1024>;   // 992 cells, 2047 faces (1024 + 1023)
using Shape3 = TileShape<15, 16, 256, true>;   // as Shape0, x- and y-face per thread in one pass
constexpr int NUM_SHAPES = 4;

enum { BC_NEUMANN = 0, BC_DIRICHLET = 1, BC_PERIODIC = 2 };

struct StageArgs {
    const double* in;     // state being differentiated [4][nx][ny]
    const double* u0;     // RK snapshot U0 (forms 1, 2)
    const double* u1;     // RK snapshot U1 (form 2)
    double* out;          // next state
    const uint8_t* type;  // cell types or nullptr (all fluid)
    long long plane;      // nx*ny
    int nx, ny;
    int tiles_j;          // tiles along j
    int ntiles;           // tiles of the grid
    int form;             // 0: c1*in + (c2*dt)L   1: (c0*u0 + c1*in) + (c2*dt)L   2: (c0*u0 + c1*u1) + (c2*dt)L
                          // 3: L itself on fluid cells, 0 elsewhere (no state is written)
    int positivity;       // inadmissible WENO face state -> adjacent cell state
    int hll_bounds;       // 0: reference (momentum, left state in SR), 1: velocity bounds
    int dx_pow2, dy_pow2; // dx (dy) is a power of two: divide by multiplying with 1/dx
    int stream_nt;        // the RK snapshots are read and the new state written with non-temporal
                          // accesses (state planes larger than the last-level cache: their lines
                          // are not reused before eviction, and would push out staged halo lines)
    int stage_early;      // the first staging pass is read before the time step is known (grids of
                          // a few tiles per CU, where a stage costs one tile's latency chain; a
                          // launch replayed past the end time then still reads that pass)
    double dx, dy, inv_dx, inv_dy, gamma;
    double c0, c1, c2;
    int bc[4];
    double bcs[4][4];
    int* exact_count;     // tiles of this stage that took the IEEE-division form of WENO5
    int* exact_other;     // the previous stage's counter: cleared by this stage
    // immersed boundary, filled while staging the tiles next to the body
    const uint8_t* tile_ib;   // per tile: its load region holds immersed-ghost cells
    const int* ig_index;      // per cell: row of the IB tables, -1 elsewhere
    const int* ig_donors;
    const double* ig_coef;
    const int* ig_transp;
    const double* ig_xy;
    int ib_interp;            // 0: reference LDL^T solve, 1: donor weights
    // time step
    int dt_mode;              // 0: *dt given; 1: first stage of a CFL step, dt from acc_in;
                              // 2: later stage of a CFL step, *dt and *halt from the first
    double* dt;
    const unsigned long long* acc_in;   // CFL bound of the state entering the step (bits)
    unsigned long long* acc_out;        // CFL bound of the state leaving the step
    int fuse_cfl;             // last stage: accumulate the CFL bound of the new state
    const double* t_in;       // time entering the step
    double* t_out;            // time leaving it
    int* halt;                // 1: this step does nothing (end time reached or bad bound)
    int* status;              // set to 1 on a non-finite or non-positive CFL bound
    int* steps;               // steps taken
    double cfl, t_end;        // t_end <= 0: no end time
    // viscous terms
    double mu, kappa, ma2;    // viscosity, mu/((gamma-1) Ma^2 Pr), Ma^2
};

// LDS of the viscous variant: primitive u, v, T of every staged cell and the
// cell gradients d/dy (x-face cross terms) and d/dx (y-face cross terms) of u,
// v where the faces read them.  An x-face at tile row fr reads d/dy at staged
// rows fr + 1 .. fr + 4 of its own column, so d/dy is kept for staged rows
// 1 .. TI + 4 of the tile's TJ columns (index (row - 1) * TJ + col - HALO);
// likewise d/dx for the tile's TI rows and staged columns 1 .. TJ + 4 (index
// (row - HALO) * (TJ + 4) + col - 1).
//
// SHARED (FP64, whose viscous kernel computes the x-faces, then the y-faces
// over the same flux planes): only x-faces read d/dy and only y-faces d/dx,
// so d/dx is computed once the x-faces are done, over d/dy: 38 KB of LDS per
// 15 x 16 tile instead of 52 KB, four workgroups per CU instead of three.
template <typename R, typename Sh> struct ViscLDS {
    static constexpr int N = Sh::SR * Sh::SC;
    static constexpr int GYW = Sh::TJ, NGY = (Sh::TI + 4) * Sh::TJ;        // d/dy: rows x columns
    static constexpr int GXW = Sh::TJ + 4, NGX = Sh::TI * (Sh::TJ + 4);    // d/dx
    static constexpr bool SHARED = sizeof(R) == 8;
    static constexpr int NG = SHARED ? (NGY > NGX ? NGY : NGX) : NGY + NGX;
    R pu[N], pv[N], pT[N];
    R gu[NG], gv[NG];   // d/dy of u, v at [0, NGY); d/dx at [0, NGX) if SHARED, else at [NGY, NGY + NGX)
    __device__ R* gyu() { return gu; }
    __device__ R* gyv() { return gv; }
    __device__ R* gxu() { return gu + (SHARED ? 0 : NGY); }
    __device__ R* gxv() { return gv + (SHARED ? 0 : NGY); }
    __device__ const R* gyu() const { return gu; }
    __device__ const R* gyv() const { return gv; }
    __device__ const R* gxu() const { return gu + (SHARED ? 0 : NGY); }
    __device__ const R* gxv() const { return gv + (SHARED ? 0 : NGY); }
};

// ---------------------------------------------------------------------------
// scalar building blocks (templated on the reconstruction/flux type)

// FP64 division and square root as the compiler's IEEE sequences compute them
// on gfx950, without their range handling, for operands where that handling
// changes nothing -- the same operations on the same values, so the same
// (correctly rounded) results.
//
// Division a/b: v_rcp_f64 and two Newton steps give r ~ 1/b, then q = a*r and
// RN(q + (a - b q) r); around it the compiler's sequence scales the operands
// (v_div_scale) only when b or 1/b is not normal, a/b is not normal,
// exponent(a) - exponent(b) >= 768 or |a| < 2^-969, and fixes up (v_div_fixup)
// only zeros, infinities and NaNs.  Taking the residual negated, b q - a, a
// zero numerator over b > 0 gives IEEE's signed zero as well (a = -0: q = -0,
// b q - a = +0, and -(+0) r + q = -0).
// Square root: v_rsq_f64 and the Newton steps of the compiler's sequence,
// which scales x by 2^256 only below 2^-767 and returns x itself for zero and
// infinity.
__device__ __forceinline__ double rcp64(double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
}
__device__ __forceinline__ double div64(double a, double b, double r) {
    const double q = a * r;
    const double e = __builtin_fma(b, q, -a);
    return __builtin_fma(-e, r, q);
}
__device__ __forceinline__ double sqrt64_core(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
// IEEE sqrt: the short form on [2^-767, 2^1000] (nearly every value), else
// the compiler's full sequence.  The empty asm keeps the second arm a branch,
// skipped when no lane takes it: the full sequence is one intrinsic call until
// instruction selection, and speculated as such it would run on every lane
// beside the short form (about 20 more VALU instructions per root).
__device__ __forceinline__ double sqrt64(double x) {
    double y;
    if (x >= 0x1p-767 && x <= 0x1p+1000) {
        y = sqrt64_core(x);
    } else {
        asm volatile("" : "+v"(x));
        y = __builtin_sqrt(x);
    }
    return y;
}

template <typename R> struct K;
template <> struct K<double> {
    static constexpr double eps = 1e-2, c1312 = 13.0 / 12.0, q = 0.25;
    static constexpr double d0 = 0.1, d1 = 0.6, d2 = 0.3, half = 0.5;
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double sqrt_(double a) { return sqrt64(a); }
    // a*b + c rounded twice, as written in the reference (no contraction)
    __device__ static double madd(double a, double b, double c) { return a * b + c; }
};
template <> struct K<float> {
    static constexpr float eps = 1e-2f, c1312 = 13.0f / 12.0f, q = 0.25f;
    static constexpr float d0 = 0.1f, d1 = 0.6f, d2 = 0.3f, half = 0.5f;
    // FP32-bulk: hardware reciprocal (1 ulp) instead of the IEEE division
    __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static float sqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
    // FP32-bulk fuses a*b + c
    __device__ static float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};

// Divisions of the WENO5 reconstruction.  WOps<double, true> is the fast form
// used on tiles whose staged values all lie in the range of fast_div_safe();
// WOps<double, false> is plain IEEE division, used on the other tiles.  Both
// give the IEEE quotient bit for bit on the tiles they run on.
//
// Division by the constants 3 and 6: 1/d as z = RN(1/d) plus z' = RN(1/d - z)
// (1/3 = 0x1.5555555555555p-2 + 0x1.5555555555555p-56 + 1.03e-33, 1/6 half of
// each), q = RN(x z + RN(x z')): one multiply and one fma.  For
// 2^-900 <= |x| <= 2^1000 (x z' normal), x z + RN(x z') differs from x/d by
// at most |x| (|z + z' - 1/d| + |z'| 2^-53) < |x| 2^-107, while x/d
// (d = 3*2^k) lies at least 2^(e-53)/3 > |x| 2^-59 from every rounding
// midpoint of its binade 2^e (x/d is a multiple of 1/3 half-ulp, and never a
// midpoint itself), so the fma's single rounding gives RN(x/d).  Zeros: x z'
// and x z carry the sign of x, and their sum keeps it.
//
// Division by a variable: v_rcp_f64 and two Newton steps give r ~ 1/b, then
// q = a*r, e = a - b*q (exact), RN(q + e*r) -- the sequence the compiler emits
// for a/b on gfx950 (v_div_scale / v_div_fmas / v_div_fixup around it), minus
// the scaling and fix-up instructions, which leave every operand unchanged
// when a, b > 0, a, b, 1/b and a/b are normal, |a| >= 2^-969 and
// exponent(a) - exponent(b) < 768.  In WENO5 the numerators are 0.1, 0.6,
// 0.3 and the alpha_k, the denominators (eps + beta)^2 >= 1e-4 and
// sum(alpha) <= 1e4; with every stencil value at most 2^200 in magnitude,
// beta <= 2^406, so alpha_k >= 2^-816 and every condition holds.  The three
// weights alpha_k / sum share one reciprocal (the same r the compiler's
// sequence computes for each of them).
template <typename R, bool FAST> struct WOps;
template <> struct WOps<double, true> {
    __device__ static double d3(double x) { return __builtin_fma(x, 0x1.5555555555555p-2, x * 0x1.5555555555555p-56); }
    __device__ static double d6(double x) { return __builtin_fma(x, 0x1.5555555555555p-3, x * 0x1.5555555555555p-57); }
    __device__ static double rcp(double b) {
        double r = __builtin_amdgcn_rcp(b);
        r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
        return __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    }
    __device__ static double div_r(double a, double b, double r) {
        const double q = a * r;
        return __builtin_fma(__builtin_fma(-b, q, a), r, q);
    }
    __device__ static double div(double a, double b) { return div_r(a, b, rcp(b)); }
    // x - k*y for k = 2, 4: k*y is exact below 2^1021, so one fma rounds once as well
    __device__ static double sub_pow2(double x, double k, double y) { return __builtin_fma(-k, y, x); }
};
template <> struct WOps<double, false> {
    __device__ static double d3(double x) { return x / 3.0; }
    __device__ static double d6(double x) { return x / 6.0; }
    __device__ static double rcp(double b) { return b; }
    __device__ static double div_r(double a, double b, double) { return a / b; }
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double sub_pow2(double x, double k, double y) { return x - k * y; }
};
template <bool FAST> struct WOps<float, FAST> {
    __device__ static float d3(float x) { return x * (1.0f / 3.0f); }
    __device__ static float d6(float x) { return x * (1.0f / 6.0f); }
    __device__ static float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
    __device__ static float div_r(float a, float, float r) { return a * r; }
    __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static float sub_pow2(float x, float k, float y) { return x - k * y; }
};

// range of the staged values for which WOps<double, true> is exact: every
// value, and its multiples by 5, 7 and 11, inside the constant-division range,
// and the smoothness indicators small enough for the unscaled division
__device__ __forceinline__ bool fast_div_safe(double v) {
    const double a = __builtin_fabs(v);
    return v == 0.0 || (a >= 0x1p-900 && a <= 0x1p+200);
}

// Jiang-Shu smoothness indicator 13/12 s^2 + 1/4 t^2.  FP64 folds the
// quarter into an fma: RN(A + t^2/4) equals the reference's RN(A + RN(t^2/4))
// unless t^2/4 is subnormal, and then the two can differ only when the sum is
// below 2^-967, where eps + beta -- the indicator's only use -- rounds to eps
// either way.
template <typename R>
__device__ __forceinline__ R weno5_beta(R s, R t) {
    if constexpr (sizeof(R) == 8) return __builtin_fma(t * t, 0.25, K<R>::c1312 * (s * s));
    else return K<R>::c1312 * (s * s) + K<R>::q * (t * t);
}

// WENO5 value at the face from five cell values, v2 the upwind cell
// (source/interpolation/Weno5th.cpp:4-28: Jiang-Shu smoothness indicators,
// eps = 1e-2, ideal weights 0.1/0.6/0.3).  Operation order as the reference.
//
// FP32-bulk evaluates the same function rearranged for fused multiply-adds
// (FP32 results are compared with FP64 within a tolerance, not bit for bit):
// value = (sum_k r_k D_k) / (sum_k d_k r_k), r_k = 1 / (eps' + beta'_k)^2,
// with beta'_k = (12/13) beta_k and eps' = (12/13) eps (the common factor
// (13/12)^2 of the r_k cancels) and D_k = d_k q_k, the ideal weights folded
// into the candidate stencils' coefficients.
constexpr float WENO_F32_K = 3.0f / 13.0f;           // (12/13)(1/4)
constexpr float WENO_F32_EPS = 1e-2f * 12.0f / 13.0f;  // (12/13) eps

// The smoothness indicator beta'_k = s^2 + k t^2 of one candidate stencil.
__device__ __forceinline__ float weno5_beta_f32(float s, float t) {
    return __builtin_fmaf(s, s, (WENO_F32_K * t) * t);
}

// value = (sum_k r_k D_k) / (sum_k d_k r_k), r_k = 1/E_k, E_k = (eps' + beta'_k)^2.
__device__ __forceinline__ float weno5_combine_f32(float b0, float b1, float b2, float D0, float D1, float D2) {
    constexpr float e = WENO_F32_EPS;
    const float e0 = e + b0, e1 = e + b1, e2 = e + b2;
    const float E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
    // r_k = 1/E_k multiplied through by E0 E1 E2: r_k -> the product of the
    // other two E's, one reciprocal instead of four.  E_k >= e^2 > 8e-5, so
    // with every E_k <= 2^40 the products lie in [7e-9, 2^80] and the
    // denominator is positive and finite; a numerator beyond the FP32 range
    // (values above ~1e13) shows as a non-finite result.  Either way out of
    // range, the reciprocal form.
    const float m = __builtin_fmaxf(E0, __builtin_fmaxf(E1, E2));
    const float P12 = E1 * E2, P02 = E0 * E2, P01 = E0 * E1;
    const float pnum = __builtin_fmaf(P01, D2, __builtin_fmaf(P02, D1, P12 * D0));
    const float pden = __builtin_fmaf(0.3f, P01, __builtin_fmaf(0.6f, P02, 0.1f * P12));
    const float val = pnum * __builtin_amdgcn_rcpf(pden);
    if (m <= 0x1p40f && __builtin_fabsf(val) <= 0x1.fffffep127f) return val;
    const float r0 = __builtin_amdgcn_rcpf(E0);
    const float r1 = __builtin_amdgcn_rcpf(E1);
    const float r2 = __builtin_amdgcn_rcpf(E2);
    const float num = __builtin_fmaf(r2, D2, __builtin_fmaf(r1, D1, r0 * D0));
    const float den = __builtin_fmaf(0.3f, r2, __builtin_fmaf(0.6f, r1, 0.1f * r0));
    return num * __builtin_amdgcn_rcpf(den);
}

// D_k = d_k q_k of the candidate stencils of (v0 .. v4), v2 the upwind cell.
__device__ __forceinline__ float weno5_d0_f32(float v0, float v1, float v2) {
    return __builtin_fmaf(0.1f * 11.0f / 6.0f, v2, __builtin_fmaf(-0.1f * 7.0f / 6.0f, v1, v0 * (0.1f / 3.0f)));
}
__device__ __forceinline__ float weno5_d1_f32(float v1, float v2, float v3) {
    return __builtin_fmaf(0.6f / 3.0f, v3, __builtin_fmaf(0.6f * 5.0f / 6.0f, v2, v1 * (-0.6f / 6.0f)));
}
__device__ __forceinline__ float weno5_d2_f32(float v2, float v3, float v4) {
    return __builtin_fmaf(-0.3f / 6.0f, v4, __builtin_fmaf(0.3f * 5.0f / 6.0f, v3, v2 * (0.3f / 3.0f)));
}

__device__ __forceinline__ float weno5_f32(float v0, float v1, float v2, float v3, float v4) {
    const float b0 = weno5_beta_f32(__builtin_fmaf(-2.0f, v1, v0) + v2,
                                    __builtin_fmaf(3.0f, v2, __builtin_fmaf(-4.0f, v1, v0)));
    const float b1 = weno5_beta_f32(__builtin_fmaf(-2.0f, v2, v1) + v3, v1 - v3);
    const float b2 = weno5_beta_f32(__builtin_fmaf(-2.0f, v3, v2) + v4,
                                    __builtin_fmaf(-4.0f, v3, __builtin_fmaf(3.0f, v2, v4)));
    return weno5_combine_f32(b0, b1, b2, weno5_d0_f32(v0, v1, v2), weno5_d1_f32(v1, v2, v3),
                             weno5_d2_f32(v2, v3, v4));
}

// Both states of a face from its six cells c0 .. c5 (cells f-3 .. f+2): the
// left state is weno5_f32(c0 .. c4), the right state weno5_f32(c5 .. c1).
// They share, up to rounding, the second differences of the triples
// (c1, c2, c3) and (c2, c3, c4) (a - 2b + c is symmetric under reversal) and
// two candidate values: 0.3 q2(c2, c3, c4) = 0.6 q1(c4, c3, c2) / 2 and
// 0.3 q2(c3, c2, c1) = 0.6 q1(c1, c2, c3) / 2.
__device__ __forceinline__ void weno5_pair_f32(const float c[6], float& L, float& Rr) {
    const float c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], c4 = c[4], c5 = c[5];
    const float s0 = __builtin_fmaf(-2.0f, c1, c0) + c2;
    const float s1 = __builtin_fmaf(-2.0f, c2, c1) + c3;
    const float s2 = __builtin_fmaf(-2.0f, c3, c2) + c4;
    const float s3 = __builtin_fmaf(-2.0f, c4, c3) + c5;
    const float bL0 = weno5_beta_f32(s0, __builtin_fmaf(3.0f, c2, __builtin_fmaf(-4.0f, c1, c0)));
    const float bL1 = weno5_beta_f32(s1, c1 - c3);
    const float bL2 = weno5_beta_f32(s2, __builtin_fmaf(-4.0f, c3, __builtin_fmaf(3.0f, c2, c4)));
    const float bR0 = weno5_beta_f32(s3, __builtin_fmaf(3.0f, c3, __builtin_fmaf(-4.0f, c4, c5)));
    const float bR1 = weno5_beta_f32(s2, c4 - c2);
    const float bR2 = weno5_beta_f32(s1, __builtin_fmaf(-4.0f, c2, __builtin_fmaf(3.0f, c3, c1)));
    const float DL1 = weno5_d1_f32(c1, c2, c3), DR1 = weno5_d1_f32(c4, c3, c2);
    L = weno5_combine_f32(bL0, bL1, bL2, weno5_d0_f32(c0, c1, c2), DL1, 0.5f * DR1);
    Rr = weno5_combine_f32(bR0, bR1, bR2, weno5_d0_f32(c5, c4, c3), DR1, 0.5f * DL1);
}

template <typename R, bool FAST>
__device__ __forceinline__ R weno5(R v0, R v1, R v2, R v3, R v4) {
    if constexpr (sizeof(R) == 4) return weno5_f32(v0, v1, v2, v3, v4);
    using C = K<R>;
    using D = WOps<R, FAST>;
    const R s0 = D::sub_pow2(v0, R(2), v1) + v2;
    const R t0 = D::sub_pow2(v0, R(4), v1) + R(3) * v2;
    const R b0 = weno5_beta<R>(s0, t0);
    const R s1 = D::sub_pow2(v1, R(2), v2) + v3;
    const R t1 = v1 - v3;
    const R b1 = weno5_beta<R>(s1, t1);
    const R s2 = D::sub_pow2(v2, R(2), v3) + v4;
    const R t2 = D::sub_pow2(R(3) * v2, R(4), v3) + v4;
    const R b2 = weno5_beta<R>(s2, t2);
    const R e0 = C::eps + b0, e1 = C::eps + b1, e2 = C::eps + b2;
    const R a0 = D::div(C::d0, e0 * e0);
    const R a1 = D::div(C::d1, e1 * e1);
    const R a2 = D::div(C::d2, e2 * e2);
    const R sum = (a0 + a1) + a2;
    const R rs = D::rcp(sum);
    const R w0 = D::div_r(a0, sum, rs), w1 = D::div_r(a1, sum, rs), w2 = D::div_r(a2, sum, rs);
    const R q0 = (D::d3(v0) - D::d6(R(7) * v1)) + D::d6(R(11) * v2);
    // d6(-x) == -d6(x) bit for bit (round-to-nearest is odd-symmetric): the
    // right state's d6(v1) is then the left state's d6(v4) and vice versa
    const R q1 = (-D::d6(v1) + D::d6(R(5) * v2)) + D::d3(v3);
    const R q2 = (D::d3(v2) + D::d6(R(5) * v3)) - D::d6(v4);
    return (w0 * q0 + w1 * q1) + w2 * q2;
}

// Left and right states of one variable from the face's six cells v[0..5]
// (cells f-3 .. f+2): weno5 of v[0..4] and of v[5..1].  FP32-bulk computes
// the two together (weno5_pair_f32).
template <typename R, bool FAST>
__device__ __forceinline__ void weno5_lr(const R v[6], R& L, R& Rr) {
    if constexpr (sizeof(R) == 4) {
        weno5_pair_f32(v, L, Rr);
    } else {
        L = weno5<R, FAST>(v[0], v[1], v[2], v[3], v[4]);
        Rr = weno5<R, FAST>(v[5], v[4], v[3], v[2], v[1]);
    }
}

// The divisions of a face state (r, n, t, E) by its density, as the
// reference computes them: pressure p = gm1 (E - (n^2 + t^2)/2 / r)
// (EulerPhase.cpp:132), u_n = n / r, n n / r, n t / r (EulerMod.cpp:35-50) and
// g p / r (RiemmanSolverHLL.cpp:18-23).
template <typename R>
__device__ __forceinline__ void state_terms(R r, R n, R t, R E, R g, R gm1, R& p, R& un, R& nn, R& nt, R& gp) {
    using C = K<R>;
    p = gm1 * (E - C::div(C::half * C::madd(n, n, t * t), r));
    un = C::div(n, r);
    nn = C::div(n * n, r);
    nt = C::div(n * t, r);
    gp = C::div(g * p, r);
}

// state_terms with one reciprocal of the density for its five divisions
// (div64).  In range when r in [2^-150, 2^150], n and t zero or of magnitude
// in [2^-200, 2^150], E zero or in [2^-500, 2^450] and (checked once per
// launch, exact_state_gamma) g in [2^-10, 2^10], g - 1 zero or of magnitude
// in [2^-60, 2^10]: then n, n^2, n t and (n^2 + t^2)/2 are zero or in
// [2^-401, 2^301], their quotients by r in [2^-551, 2^451]; E - (n^2+t^2)/(2r),
// a difference of multiples of 2^-603, is zero or in [2^-603, 2^452], so g p
// is zero or in [2^-673, 2^471] and g p / r normal: every division meets
// div64's conditions.
__device__ __forceinline__ bool exact_state_ok(double r, double n, double t, double E) {
    const double an = __builtin_fabs(n), at = __builtin_fabs(t), aE = __builtin_fabs(E);
    return r >= 0x1p-150 && r <= 0x1p+150 && an <= 0x1p+150 && (n == 0.0 || an >= 0x1p-200) && at <= 0x1p+150 &&
           (t == 0.0 || at >= 0x1p-200) && aE <= 0x1p+450 && (E == 0.0 || aE >= 0x1p-500);
}
__host__ __device__ inline bool exact_state_gamma(double g) {
    const double gm1 = g - 1.0, a = gm1 < 0.0 ? -gm1 : gm1;
    return g >= 0x1p-10 && g <= 0x1p+10 && (gm1 == 0.0 || (a >= 0x1p-60 && a <= 0x1p+10));
}
__device__ __forceinline__ void state_terms_exact(double r, double n, double t, double E, double g, double gm1,
                                                  double& p, double& un, double& nn, double& nt, double& gp) {
    const double ir = rcp64(r);
    const double n2 = n * n;
    p = gm1 * (E - div64(0.5 * (n2 + t * t), r, ir));
    un = div64(n, r, ir);
    nn = div64(n2, r, ir);
    nt = div64(n * t, r, ir);
    gp = div64(g * p, r, ir);
}

template <typename R>
__device__ __forceinline__ void riemann_flux(const R L[4], const R Rr[4], const R cl[4], const R cr[4], R g,
                                             int positivity, int hll_bounds, bool xg, R out[4]);

template <typename R, bool FAST>
__device__ __forceinline__ void face_flux(const R* __restrict__ S, int ns, const int cell[6], int nrm, int tng,
                                          R g, int positivity, int hll_bounds, bool xg, unsigned slow, R out[4]) {
    R L[4], Rr[4], cl[4], cr[4];
    const int comp[4] = {0, nrm, tng, 3};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        R v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = S[comp[q] * ns + cell[k]];
        if (!FAST && ((slow >> comp[q]) & 1u)) {   // this variable's IEEE form (uniform per tile)
            weno5_lr<R, false>(v, L[q], Rr[q]);
        } else {
            weno5_lr<R, true>(v, L[q], Rr[q]);
        }
        cl[q] = v[2];
        cr[q] = v[3];
    }
    riemann_flux<R>(L, Rr, cl, cr, g, positivity, hll_bounds, xg, out);
}

// Pressures, the positivity fallback, Euler fluxes and the HLL flux of N
// faces' two reconstructed states each (cl, cr: the adjacent cells' states).
// The faces are independent; each statement runs over all N before the next
// one, so N > 1 gives the scheduler N independent dependency chains.
template <typename R, int N>
__device__ __forceinline__ void riemann_flux_n(const R (*L)[4], const R (*Rr)[4], const R (*cl)[4], const R (*cr)[4],
                                               R g, int positivity, int hll_bounds, bool xg, R (*out)[4]) {
    using C = K<R>;
    const R gm1 = g - R(1);
    R rL[N], nL[N], tL[N], EL[N], rR[N], nR[N], tR[N], ER[N];
    R pL[N], unL[N], nnL[N], ntL[N], gpL[N], pR[N], unR[N], nnR[N], ntR[N], gpR[N];
#pragma unroll
    for (int f = 0; f < N; ++f) {
        rL[f] = L[f][0]; nL[f] = L[f][1]; tL[f] = L[f][2]; EL[f] = L[f][3];
        rR[f] = Rr[f][0]; nR[f] = Rr[f][1]; tR[f] = Rr[f][2]; ER[f] = Rr[f][3];
        // pressure and the state's divisions by its density (EulerPhase.cpp:132,
        // EulerMod.cpp:35-50, RiemmanSolverHLL.cpp:18-23)
        // (one face at a time only: two faces' both forms spill)
        bool exact = false;
        if constexpr (sizeof(R) == 8 && N == 1)
            exact = xg && exact_state_ok(rL[f], nL[f], tL[f], EL[f]) && exact_state_ok(rR[f], nR[f], tR[f], ER[f]);
        if (exact) {
            if constexpr (sizeof(R) == 8 && N == 1) {
                state_terms_exact(rL[f], nL[f], tL[f], EL[f], g, gm1, pL[f], unL[f], nnL[f], ntL[f], gpL[f]);
                state_terms_exact(rR[f], nR[f], tR[f], ER[f], g, gm1, pR[f], unR[f], nnR[f], ntR[f], gpR[f]);
            }
        } else {
            state_terms<R>(rL[f], nL[f], tL[f], EL[f], g, gm1, pL[f], unL[f], nnL[f], ntL[f], gpL[f]);
            state_terms<R>(rR[f], nR[f], tR[f], ER[f], g, gm1, pR[f], unR[f], nnR[f], ntR[f], gpR[f]);
        }
    }
    if (positivity) {
#pragma unroll
        for (int f = 0; f < N; ++f) {
            // an inadmissible reconstruction takes the state of the cell on its side
            if (!(rL[f] > R(0) && pL[f] > R(0))) {
                rL[f] = cl[f][0]; nL[f] = cl[f][1]; tL[f] = cl[f][2]; EL[f] = cl[f][3];
                state_terms<R>(rL[f], nL[f], tL[f], EL[f], g, gm1, pL[f], unL[f], nnL[f], ntL[f], gpL[f]);
            }
            if (!(rR[f] > R(0) && pR[f] > R(0))) {
                rR[f] = cr[f][0]; nR[f] = cr[f][1]; tR[f] = cr[f][2]; ER[f] = cr[f][3];
                state_terms<R>(rR[f], nR[f], tR[f], ER[f], g, gm1, pR[f], unR[f], nnR[f], ntR[f], gpR[f]);
            }
        }
    }
#pragma unroll
    for (int f = 0; f < N; ++f) {
        // Euler fluxes of both states
        const R FL0 = nL[f], FL1 = nnL[f] + pL[f], FL2 = ntL[f], FL3 = (EL[f] + pL[f]) * unL[f];
        const R FR0 = nR[f], FR1 = nnR[f] + pR[f], FR2 = ntR[f], FR3 = (ER[f] + pR[f]) * unR[f];
        // HLL with PVRS star-state speed estimates
        const R aL = C::sqrt_(gpL[f]);
        const R aR = C::sqrt_(gpR[f]);
        const R a_ave = C::half * (aL + aR);
        const R rho_ave = C::half * (rL[f] + rR[f]);
        const R p_star = C::madd(-((C::half * (unR[f] - unL[f])) * rho_ave), a_ave, C::half * (pL[f] + pR[f]));
        const R u_star = C::half * (unR[f] + unL[f]) - C::div(C::half * (pR[f] - pL[f]), rho_ave * a_ave);
        const R rsL = rL[f] + C::div((unL[f] - u_star) * rho_ave, a_ave);
        const R rsR = rR[f] - C::div((unR[f] - u_star) * rho_ave, a_ave);
        const R asL = C::sqrt_(C::div(g * p_star, rsL));
        const R asR = C::sqrt_(C::div(g * p_star, rsR));
        // reference bounds: normal momentum, left state in SR; else velocities
        const R x1 = (hll_bounds ? unL[f] : nL[f]) - aL, x2 = u_star - asL;
        const R SL = (x2 < x1) ? x2 : x1;           // std::min
        const R y1 = hll_bounds ? unR[f] + aR : nL[f] + aL, y2 = u_star + asR;
        const R SR = (y1 < y2) ? y2 : y1;           // std::max
        const bool left = SL >= R(0), right = SR <= R(0);
        if (left || right) {
            // supersonic: one side's Euler flux, the HLL average is not formed
            // (whole waves of a supersonic stream skip it)
            out[f][0] = left ? FL0 : FR0;
            out[f][1] = left ? FL1 : FR1;
            out[f][2] = left ? FL2 : FR2;
            out[f][3] = left ? FL3 : FR3;
        } else {
            const R inv = C::div(R(1), SR - SL);
            const R ss = SL * SR;
            out[f][0] = C::madd(ss, rR[f] - rL[f], C::madd(SR, FL0, -(SL * FR0))) * inv;
            out[f][1] = C::madd(ss, nR[f] - nL[f], C::madd(SR, FL1, -(SL * FR1))) * inv;
            out[f][2] = C::madd(ss, tR[f] - tL[f], C::madd(SR, FL2, -(SL * FR2))) * inv;
            out[f][3] = C::madd(ss, ER[f] - EL[f], C::madd(SR, FL3, -(SL * FR3))) * inv;
        }
    }
}

template <typename R>
__device__ __forceinline__ void riemann_flux(const R L[4], const R Rr[4], const R cl[4], const R cr[4], R g,
                                             int positivity, int hll_bounds, bool xg, R out[4]) {
    R o[1][4];
    riemann_flux_n<R, 1>((const R(*)[4])L, (const R(*)[4])Rr, (const R(*)[4])cl, (const R(*)[4])cr, g, positivity,
                         hll_bounds, xg, o);
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = o[0][q];
}

// Fetch one cell of the (halo-extended) tile, applying the physical boundary:
// an index outside [0, n) on a Neumann / Dirichlet side is the side's single
// ghost cell (the reference's face walk clamps to it, CellInterface.cpp:172-178):
// Neumann ghost = copy of the adjacent interior cell (NeummanBoundary.cpp:31-45),
// Dirichlet ghost = the fixed state (DirichletBoundary.cpp:32-46).
// Returns the linear index of the interior cell read, -1 for a ghost value.
__device__ __forceinline__ long long fetch_cell(const StageArgs& a, int gi, int gj, double v[4]) {
    int side = -1;
    if (gi < 0) {
        if (a.bc[0] == BC_PERIODIC) gi += a.nx; else side = 0;
    } else if (gi >= a.nx) {
        if (a.bc[1] == BC_PERIODIC) gi -= a.nx; else side = 1;
    }
    if (gj < 0) {
        if (a.bc[2] == BC_PERIODIC) gj += a.ny; else if (side < 0) side = 2;
    } else if (gj >= a.ny) {
        if (a.bc[3] == BC_PERIODIC) gj -= a.ny; else if (side < 0) side = 3;
    }
    gi = gi < 0 ? 0 : (gi >= a.nx ? a.nx - 1 : gi);
    gj = gj < 0 ? 0 : (gj >= a.ny ? a.ny - 1 : gj);
    if (side >= 0 && a.bc[side] == BC_DIRICHLET) {
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = a.bcs[side][q];
        return -1;
    }
    const long long c = (long long)gi * a.ny + gj;
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = G(a.in)[q * a.plane + c];
    return side >= 0 ? -1 : c;
}

// ---------------------------------------------------------------------------
// Immersed-boundary ghost value (Cell::comImmersedGhostPhase, Cell.cpp:172-181
// -> ImagePoint::InterpolatePhase, ImagePoint.cpp:20-82): 4 donor states; the
// reference interpolation solves the donor system through its precomputed
// pivoted LDL^T per variable and evaluates w0*x*y + w1*x + w2*y + w3 at the
// image point; the positivity-preserving one is sum_k g_k f_k.  The ghost is
// (rho, -rho*u, -rho*v, E).  FP64 in every precision mode.

__device__ __forceinline__ double pick4(const double x[4], int p) {
    return p == 0 ? x[0] : p == 1 ? x[1] : p == 2 ? x[2] : x[3];
}
__device__ __forceinline__ void put4(double x[4], int p, double v) {
    x[0] = (p == 0) ? v : x[0];
    x[1] = (p == 1) ? v : x[1];
    x[2] = (p == 2) ? v : x[2];
    x[3] = (p == 3) ? v : x[3];
}

// Solve with the pivoted LDL^T of the donor matrix, in the order of Eigen's
// LDLT::solve: transpositions, unit-lower solve (column sweep), D^-1 (zero
// below DBL_MIN), unit-upper (L^T) solve (row dot products), transpositions
// in reverse.
__device__ __forceinline__ void ldlt_solve(const double* l, const int* tr, double x[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = tr[k];
        const double xk = x[k], xp = pick4(x, p);
        put4(x, p, xk);
        x[k] = xp;
    }
    x[1] -= x[0] * l[0];
    x[2] -= x[0] * l[1];
    x[3] -= x[0] * l[3];
    x[2] -= x[1] * l[2];
    x[3] -= x[1] * l[4];
    x[3] -= x[2] * l[5];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double d = l[6 + i];
        x[i] = (__builtin_fabs(d) > 2.2250738585072014e-308) ? x[i] / d : 0.0;
    }
    x[2] -= l[5] * x[3];
    x[1] -= l[2] * x[2] + l[4] * x[3];
    x[0] -= (l[0] * x[1] + l[1] * x[2]) + l[3] * x[3];
#pragma unroll
    for (int k = 3; k >= 0; --k) {
        const int p = tr[k];
        const double xk = x[k], xp = pick4(x, p);
        put4(x, p, xk);
        x[k] = xp;
    }
}

__device__ __forceinline__ void ghost_value(const double* __restrict__ U, long long plane, int interp,
                                            const int* dn, const double* l, const int* tr, const double* xy,
                                            double v[4]) {
    // one variable at a time keeps the register footprint small
    const long long c0 = dn[0], c1 = dn[1], c2 = dn[2], c3 = dn[3];
    double val[4];
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const double* Uq = U + q * plane;
        double w[4] = {Uq[c0], Uq[c1], Uq[c2], Uq[c3]};
        if (interp == 0) {
            ldlt_solve(l, tr, w);
            const double x = xy[0], y = xy[1];
            val[q] = ((w[0] * x * y + w[1] * x) + w[2] * y) + w[3];
        } else {
            val[q] = ((l[0] * w[0] + l[1] * w[1]) + l[2] * w[2]) + l[3] * w[3];
        }
    }
    v[0] = val[0];
    v[1] = -val[1];
    v[2] = -val[2];
    v[3] = val[3];
}

// CFL bound of one cell: (a_x + c)/dx + (a_y + c)/dy, c = sqrt(gamma p/rho),
// a_x = max(|u|, |rho u|) with the reference HLL bounds (they use the
// momentum), |u| with velocity bounds.
__device__ __forceinline__ double cell_speed(double rho, double mu, double mv, double E, double gamma, double dx,
                                             double dy, int hll_bounds) {
    const double p = (gamma - 1) * (E - 0.5 * (mu * mu + mv * mv) / rho);
    const double cs = sqrt64(gamma * p / rho);
    const double u = mu / rho, v = mv / rho;
    double ax = __builtin_fabs(u), ay = __builtin_fabs(v);
    if (hll_bounds == 0) {
        if (ax < __builtin_fabs(mu)) ax = __builtin_fabs(mu);
        if (ay < __builtin_fabs(mv)) ay = __builtin_fabs(mv);
    }
    return (ax + cs) / dx + (ay + cs) / dy;
}

// The same bound for FP32-bulk stages, still in FP64 but with one refined
// reciprocal of rho (two Newton steps after the hardware reciprocal) and
// multiplications by 1/dx, 1/dy in place of the five IEEE divisions: within
// a few ulp of cell_speed.  A non-positive density or pressure gives a
// non-finite bound, as there.
__device__ __forceinline__ double cell_speed_bulk(double rho, double mu, double mv, double E, double gamma,
                                                  double inv_dx, double inv_dy, int hll_bounds) {
    double ir = __builtin_amdgcn_rcp(rho);
    ir = __builtin_fma(ir, __builtin_fma(-rho, ir, 1.0), ir);
    ir = __builtin_fma(ir, __builtin_fma(-rho, ir, 1.0), ir);
    const double u = mu * ir, v = mv * ir;
    const double p = (gamma - 1) * (E - 0.5 * (mu * u + mv * v));
    const double cs = sqrt64((gamma * p) * ir);
    double ax = __builtin_fabs(u), ay = __builtin_fabs(v);
    if (hll_bounds == 0) {
        if (ax < __builtin_fabs(mu)) ax = __builtin_fabs(mu);
        if (ay < __builtin_fabs(mv)) ay = __builtin_fabs(mv);
    }
    return (ax + cs) * inv_dx + (ay + cs) * inv_dy;
}

__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

// Block-wide max (NaN wins) -> atomicMax on the bit pattern.  The atomic is
// skipped when the accumulator already holds at least this block's value: the
// accumulator only grows, so a stale read can only cost an extra atomic, never
// lose a maximum -- and thousands of workgroups do not queue on one address.
__device__ __forceinline__ void block_max_to(double m, unsigned long long* acc, double* wmax) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = max_nan(m, __shfl_xor(m, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wmax[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) b = max_nan(b, wmax[k]);
        const unsigned long long mine = (unsigned long long)__double_as_longlong(b);
        const unsigned long long seen = __hip_atomic_load(acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (mine > seen) atomicMax(acc, mine);
    }
}

// ---------------------------------------------------------------------------
// One tile: stage it (with halo) in LDS, compute its 511 faces, update its
// cells.  Tiles holding a value outside the range of the fast WENO5 division
// forms (fast_div_safe) compute their faces with IEEE division and are
// counted in exact_count.
// ---------------------------------------------------------------------------
// Positivity-preserving flux limiter (positivity = 2; no reference
// counterpart).  Hu, Adams & Shu, J. Comput. Phys. 242 (2013) 169-180: each
// face flux is blended with the Lax-Friedrichs flux of the two adjacent cell
// states, F = F_LF + theta (F_hat - F_LF), theta in [0, 1] the largest value
// for which both cells' shares of the forward-Euler update stay admissible
// (rho, p >= min(1e-13, their Lax-Friedrichs value)).  A cell's update
// U - ln (F_r - F_l) - lt (G_t - G_b) (ln = dt/h_n) is split into one share per
// face, U -+ k (F_face - F(U)) with k = 2 (ln an + lt at) / an, an (at) the
// largest |u_n| + c (|u_t| + c) over the cell and its two neighbours along n
// (t); the Lax-Friedrichs shares are admissible when ln an + lt at <= 1/2,
// which the CFL bound gives for CFL <= 1/4 and, in practice, at 1/2.  With
// SSP Runge-Kutta stages (RK2, and RK3 except the reference's last stage) the
// new state keeps rho, p > 0.  All in the face-normal frame.  A face whose two
// shares taken with its own flux already keep rho and p above a fixed floor
// (FP64: rho >= 1e-13 and rho p >= 1e-13 rho; FP32: a thousandth of the cell's
// own density and pressure) keeps its flux (theta = 1) without the
// Lax-Friedrichs flux being formed: about 99.4% of the faces of a Mach-3
// cylinder flow, whose other faces run the full test per wave.

// The limiter's per-cell values in LDS.  FP64 (CellK, the LDS bounds the
// workgroups per CU): for the cells next to a tile face (rows -1..TI, columns
// -1..TJ of the tile, corners unused) 1/rho and the pressure, and the share
// factors k of the cells of x-faces (rows -1..TI, columns 0..TJ-1) and of
// y-faces (rows 0..TI-1, columns -1..TJ), each computed once for both faces
// of its cell; while the tile is staged, |u| + c and |v| + c of rows and
// columns -2..TI+1 / -2..TJ+1 live in the flux planes, not yet written.
// FP32 (FaceK): |u| + c, |v| + c, 1/rho and pressure of every staged cell in
// the state planes' indexing, k per face (fewer integer offsets and no
// second pass: the faster form where the LDS has room).
template <typename R, typename Sh> struct LimitCellK {
    static constexpr int PW = Sh::TJ + 2, NP = (Sh::TI + 2) * PW;     // 1/rho, p: row width, cells
    static constexpr int NKX = (Sh::TI + 2) * Sh::TJ, NKY = Sh::TI * PW;
    static constexpr int WW = Sh::TJ + 4, NW = (Sh::TI + 4) * WW;     // |u|+c, |v|+c: row width, cells
    static_assert(2 * NW <= 4 * Sh::NYF, "the wave speeds fit in any flux plane-set");
    R ir[NP], p[NP];
    R kx[NKX], ky[NKY];
};
template <typename R, typename Sh> struct LimitFaceK {
    R sx[Sh::SR * Sh::SC], sy[Sh::SR * Sh::SC];   // |u| + c and |v| + c of every staged cell
    R ir[Sh::SR * Sh::SC], p[Sh::SR * Sh::SC];    // 1/rho and the pressure of every staged cell
};
template <typename R, typename Sh>
using LimitLDS = typename std::conditional<sizeof(R) == 8, LimitCellK<R, Sh>, LimitFaceK<R, Sh>>::type;

template <typename R> __device__ __forceinline__ R maxr(R x, R y) { return (x < y) ? y : x; }

template <typename R>
__device__ __forceinline__ R pressure_n(R r, R n, R t, R E, R gm1) {
    return gm1 * (E - K<R>::div(K<R>::half * K<R>::madd(n, n, t * t), r));
}

// density times pressure without a division: rho p = gm1 (E rho - (n^2 + t^2)/2)
__device__ __forceinline__ double rho_p(double r, double n, double t, double E, double gm1) {
    return gm1 * (E * r - 0.5 * (n * n + t * t));
}

// largest theta in [0, 1] keeping B0 + theta*Dl admissible: density first,
// then pressure, which is concave along the density-limited segment.
// Floors: min(1e-13, Lax-Friedrichs value) in FP64; in FP32 a thousandth of
// the Lax-Friedrichs value, which stays clear of the rounding of FP32 fluxes.
// FP64 makes the pressure test p1 < floor on rho p, multiplied out (positive
// densities), so the common case -- nothing to limit -- has no division.
template <typename R>
__device__ __forceinline__ R share_theta(const R B0[4], const R Dl[4], R gm1) {
    using C = K<R>;
    const R r0 = B0[0];
    const R eps = R(1e-13);
    if constexpr (sizeof(R) == 8) {
        const double P0 = rho_p(B0[0], B0[1], B0[2], B0[3], gm1);
        const double er = (r0 < eps) ? r0 : eps;
        double th = 1.0;
        const double r1 = B0[0] + Dl[0];
        if (r1 < er) th = C::div(r0 - er, r0 - r1);
        double B1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) B1[q] = C::madd(th, Dl[q], B0[q]);
        const double P1 = rho_p(B1[0], B1[1], B1[2], B1[3], gm1);
        // p1 < min(p0, eps): p0 < eps ? p1 < p0 : p1 < eps
        const bool below = (P0 < eps * r0) ? (P1 * r0 < P0 * B1[0]) : (P1 < eps * B1[0]);
        if (below) {
            const double p0 = C::div(P0, r0), p1 = C::div(P1, B1[0]);
            const double ep = (p0 < eps) ? p0 : eps;
            th = th * C::div(p0 - ep, p0 - p1);
        }
        return th;
    } else {
        const R p0 = pressure_n<R>(B0[0], B0[1], B0[2], B0[3], gm1);
        const R er = r0 * R(1e-3), ep = p0 * R(1e-3);
        R th = R(1);
        const R r1 = B0[0] + Dl[0];
        if (r1 < er) th = C::div(r0 - er, r0 - r1);
        R B1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) B1[q] = C::madd(th, Dl[q], B0[q]);
        const R p1 = pressure_n<R>(B1[0], B1[1], B1[2], B1[3], gm1);
        if (p1 < ep) th = th * C::div(p0 - ep, p0 - p1);
        return th;
    }
}

// Euler flux of a cell state in the face-normal frame from the cell's 1/rho
// (ir) and pressure p, both computed once per staged cell: un = n*ir,
// ut = t*ir, p = gm1*(E - half*(rho u^2 + rho v^2)) with the two momentum
// terms in either order (IEEE addition commutes, so both directions see the
// same bits)
template <typename R>
__device__ __forceinline__ void euler_flux_n(R n, R t, R E, R ir, R p, R F[4]) {
    using C = K<R>;
    const R un = n * ir, ut = t * ir;
    F[0] = n;
    F[1] = C::madd(n, un, p);
    F[2] = n * ut;
    F[3] = (E + p) * un;
}

// The limiter's values of staged cell s (tile row r, column c; state rho, mx,
// my, E), computed as the cell is staged:
// |u| + c and |v| + c (the |u_n| + c of the limiter) and, for the cells next
// to a tile face, 1/rho and the pressure.  FP64 keeps the wave speeds of rows
// and columns -2..TI+1 / -2..TJ+1 in W (the flux planes, not yet written).
template <typename R, typename Sh>
__device__ __forceinline__ void limit_cell(int s, int r, int c, R rho, R mx, R my, R E, R g, R* __restrict__ W,
                                           LimitLDS<R, Sh>* Lm) {
    using C = K<R>;
    constexpr int TI = Sh::TI, TJ = Sh::TJ;
    const R ir = C::div(R(1), rho);
    const R ux = mx * ir, uy = my * ir;
    const R p = (g - R(1)) * (E - C::half * (mx * ux + my * uy));
    const R cs = C::sqrt_((g * p) * ir);
    const R sx = (ux < R(0) ? -ux : ux) + cs, sy = (uy < R(0) ? -uy : uy) + cs;
    if constexpr (sizeof(R) == 8) {
        using LL = LimitCellK<R, Sh>;
        if (r >= -2 && r <= TI + 1 && c >= -2 && c <= TJ + 1) {
            const int w = (r + 2) * LL::WW + c + 2;
            W[w] = sx;
            W[LL::NW + w] = sy;
            if (r >= -1 && r <= TI && c >= -1 && c <= TJ) {
                const int q = (r + 1) * LL::PW + c + 1;
                Lm->ir[q] = ir;
                Lm->p[q] = p;
            }
        }
    } else {
        Lm->sx[s] = sx;
        Lm->sy[s] = sy;
        Lm->ir[s] = ir;
        Lm->p[s] = p;
    }
}

// FP64: the share factors k = 2 (ln an + lt at) / an of the cells next to a
// tile face (an, at: the largest |u_n| + c, |u_t| + c over the cell and its
// two neighbours along n, t; ln, lt: dt/h along them), from the wave speeds
// in W, once the tile is staged
template <typename R, typename Sh>
__device__ __forceinline__ void limit_share(const R* __restrict__ W, LimitLDS<R, Sh>* Lm, R lam_x, R lam_y) {
    if constexpr (sizeof(R) == 8) {
        using LL = LimitCellK<R, Sh>;
        using C = K<R>;
        constexpr int TI = Sh::TI, TJ = Sh::TJ;
        const R* sx = W;
        const R* sy = W + LL::NW;
        for (int s = threadIdx.x; s < LL::NP; s += Sh::NT) {
            const int r = s / LL::PW, c = s - r * LL::PW;   // tile row r - 1, column c - 1
            const int w = (r + 1) * LL::WW + c + 1;
            const R mx = maxr<R>(maxr<R>(sx[w - LL::WW], sx[w]), sx[w + LL::WW]);
            const R my = maxr<R>(maxr<R>(sy[w - 1], sy[w]), sy[w + 1]);
            if (c >= 1 && c <= TJ) Lm->kx[r * TJ + c - 1] = C::div(R(2) * C::madd(lam_x, mx, lam_y * my), mx);
            if (r >= 1 && r <= TI) Lm->ky[(r - 1) * LL::PW + c] = C::div(R(2) * C::madd(lam_y, my, lam_x * mx), my);
        }
        __syncthreads();
    }
}

// The LDS indices of a face's two cells in the limiter's 1/rho and pressure
// planes (pL, pR) and its share factors (kL, kR).  cL: the left cell's index
// in the state planes; s: the face's index among the tile's faces, x-faces
// [0, NXF), then y-faces; lam_x, lam_y: dt/dx, dt/dy.
template <typename R, typename Sh>
__device__ __forceinline__ void limit_cells(const LimitLDS<R, Sh>& Lm, bool xf, int s, int cL, R lam_x, R lam_y,
                                            int& pL, int& pR, R& kL, R& kR) {
    constexpr int TJ = Sh::TJ, SC = Sh::SC;
    if constexpr (sizeof(R) == 8) {
        constexpr int PW = LimitCellK<R, Sh>::PW;
        int kl;
        if (xf) {
            const int fr = s / TJ, jl = s - fr * TJ;
            pL = fr * PW + jl + 1;
            pR = pL + PW;
            kl = s;
            kL = Lm.kx[kl];
            kR = Lm.kx[kl + TJ];
        } else {
            const int t = s - Sh::NXF, il = t / (TJ + 1), gl = t - il * (TJ + 1);
            pL = (il + 1) * PW + gl;
            pR = pL + 1;
            kl = t + il;
            kL = Lm.ky[kl];
            kR = Lm.ky[kl + 1];
        }
    } else {
        using C = K<R>;
        const int dn = xf ? SC : 1, dtg = xf ? 1 : SC;
        const R* sn = xf ? Lm.sx : Lm.sy;
        const R* st = xf ? Lm.sy : Lm.sx;
        const R ln = xf ? lam_x : lam_y, lt = xf ? lam_y : lam_x;
        pL = cL;
        pR = cL + dn;
        const R anL = maxr<R>(maxr<R>(sn[pL - dn], sn[pL]), sn[pL + dn]);
        const R atL = maxr<R>(maxr<R>(st[pL - dtg], st[pL]), st[pL + dtg]);
        const R anR = maxr<R>(maxr<R>(sn[pR - dn], sn[pR]), sn[pR + dn]);
        const R atR = maxr<R>(maxr<R>(st[pR - dtg], st[pR]), st[pR + dtg]);
        kL = C::div(R(2) * C::madd(ln, anL, lt * atL), anL);
        kR = C::div(R(2) * C::madd(ln, anR, lt * atR), anR);
    }
}

// cL, cR: LDS cells left and right of the face; pL, pR: the same cells in the
// 1/rho and pressure planes; kL, kR: their share factors
template <typename R, typename Sh>
__device__ __forceinline__ void limit_face(const R* __restrict__ S, int ns, const LimitLDS<R, Sh>& Lm, int cL, int cR,
                                           int pL, int pR, R kL, R kR, int nrm, int tng, R g, R fl[4]) {
    using C = K<R>;
    const R gm1 = g - R(1);
    const R UL[4] = {S[cL], S[nrm * ns + cL], S[tng * ns + cL], S[3 * ns + cL]};
    const R UR[4] = {S[cR], S[nrm * ns + cR], S[tng * ns + cR], S[3 * ns + cR]};
    const R irL = Lm.ir[pL], prL = Lm.p[pL], irR = Lm.ir[pR], prR = Lm.p[pR];
    R FL[4], FR[4];
    euler_flux_n<R>(UL[1], UL[2], UL[3], irL, prL, FL);
    euler_flux_n<R>(UR[1], UR[2], UR[3], irR, prR, FR);
    // Both cells' shares with the face's own flux, U -+ k (F - F(U)): when both
    // keep their density and pressure at or above a floor that every limited
    // share also respects, theta = 1 and the flux stays as it is (nearly every
    // face).  FP64: rho >= 1e-13 and rho p >= 1e-13 rho; FP32: a thousandth of
    // the cell's own density and pressure.
    R HL[4], HR[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        HL[q] = C::madd(-kL, fl[q] - FL[q], UL[q]);
        HR[q] = C::madd(kR, fl[q] - FR[q], UR[q]);
    }
    bool accL, accR;
    if constexpr (sizeof(R) == 8) {
        const double eps = 1e-13;
        accL = HL[0] >= eps && rho_p(HL[0], HL[1], HL[2], HL[3], gm1) >= eps * HL[0];
        accR = HR[0] >= eps && rho_p(HR[0], HR[1], HR[2], HR[3], gm1) >= eps * HR[0];
    } else {
        accL = HL[0] >= R(1e-3) * UL[0] && pressure_n<R>(HL[0], HL[1], HL[2], HL[3], gm1) >= R(1e-3) * prL;
        accR = HR[0] >= R(1e-3) * UR[0] && pressure_n<R>(HR[0], HR[1], HR[2], HR[3], gm1) >= R(1e-3) * prR;
    }
    if (accL && accR) return;
    // |u_n| + c of both cells: FP32 keeps them, FP64 computes them again as
    // limit_stage did
    R snL, snR;
    if constexpr (sizeof(R) == 8) {
        const R unL = UL[1] * irL, unR = UR[1] * irR;
        snL = (unL < R(0) ? -unL : unL) + C::sqrt_((g * prL) * irL);
        snR = (unR < R(0) ? -unR : unR) + C::sqrt_((g * prR) * irR);
    } else {
        snL = (nrm == 1 ? Lm.sx : Lm.sy)[pL];
        snR = (nrm == 1 ? Lm.sx : Lm.sy)[pR];
    }
    const R a = maxr<R>(snL, snR);
    R Flf[4], D[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        Flf[q] = C::madd(-(C::half * a), UR[q] - UL[q], C::half * (FL[q] + FR[q]));
        D[q] = fl[q] - Flf[q];
    }
    R B0[4], Dl[4];
    R thL = R(1), thR = R(1);
    if (!accL) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            B0[q] = C::madd(-kL, Flf[q] - FL[q], UL[q]);
            Dl[q] = -(kL * D[q]);
        }
        thL = share_theta<R>(B0, Dl, gm1);
    }
    if (!accR) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            B0[q] = C::madd(kR, Flf[q] - FR[q], UR[q]);
            Dl[q] = kR * D[q];
        }
        thR = share_theta<R>(B0, Dl, gm1);
    }
    R th = (thR < thL) ? thR : thL;
    th = (th < R(0)) ? R(0) : th;
    if (th < R(1)) {
#pragma unroll
        for (int q = 0; q < 4; ++q) fl[q] = C::madd(th, D[q], Flf[q]);
    }
}

// ---------------------------------------------------------------------------
// Viscous terms: textbook compressible Navier-Stokes (constant mu, Prandtl
// number, T = gamma Ma^2 p/rho as EulerPhase.cpp:137) on the stencil tables
// of source/gradient/FourthOrderCS.h:27-42 (whose own viscous code is never
// called, Run.cpp:177,188): 4th-order central cell gradients (theta_c),
// 6th-order staggered derivatives at i-3/2, i-1/2, i+1/2 (theta_r), 4-point
// interpolation of cell gradients and velocities to those points (zeta_r, the
// i-3/2 row corrected to its Lagrange weights [5,15,-5,1]/16), combined with
// (-1, 26, -1)/24 (FourthOrderCS.cpp:77-112).  Face-normal frame: n, t.

template <typename R> struct VT;
template <> struct VT<double> {
    static constexpr double thr[3][6] = {
        {71.0 / 1920, -141.0 / 128, 69.0 / 64, 1.0 / 192, -3.0 / 128, 3.0 / 640},
        {-3.0 / 640, 25.0 / 384, -75.0 / 64, 75.0 / 64, -25.0 / 384, 3.0 / 640},
        {-3.0 / 640, 3.0 / 128, -1.0 / 192, -69.0 / 64, 141.0 / 128, -71.0 / 1920}};
    static constexpr double thc[5] = {1.0 / 12, -8.0 / 12, 0.0, 8.0 / 12, -1.0 / 12};
    static constexpr double zr[3][4] = {
        {5.0 / 16, 15.0 / 16, -5.0 / 16, 1.0 / 16}, {-1.0 / 16, 9.0 / 16, 9.0 / 16, -1.0 / 16},
        {1.0 / 16, -5.0 / 16, 15.0 / 16, 5.0 / 16}};
};
template <> struct VT<float> {
    static constexpr float thr[3][6] = {
        {71.0f / 1920, -141.0f / 128, 69.0f / 64, 1.0f / 192, -3.0f / 128, 3.0f / 640},
        {-3.0f / 640, 25.0f / 384, -75.0f / 64, 75.0f / 64, -25.0f / 384, 3.0f / 640},
        {-3.0f / 640, 3.0f / 128, -1.0f / 192, -69.0f / 64, 141.0f / 128, -71.0f / 1920}};
    static constexpr float thc[5] = {1.0f / 12, -8.0f / 12, 0.0f, 8.0f / 12, -1.0f / 12};
    static constexpr float zr[3][4] = {
        {5.0f / 16, 15.0f / 16, -5.0f / 16, 1.0f / 16}, {-1.0f / 16, 9.0f / 16, 9.0f / 16, -1.0f / 16},
        {1.0f / 16, -5.0f / 16, 15.0f / 16, 5.0f / 16}};
};

// net viscous flux (0, normal-momentum, tangential-momentum, energy) of a face
// whose 6 stencil cells have LDS cell indices cs[k]; gs[m] indexes the
// gradient arrays at cell cs[1 + m]
template <typename R, typename Sh>
__device__ __forceinline__ void viscous_flux(const ViscLDS<R, Sh>& V, const int cs[6], const int gs[4], bool xface,
                                             R h, bool h_pow2, R inv_h, R mu, R kappa, R out[3]) {
    const R* un = xface ? V.pu : V.pv;
    const R* ut = xface ? V.pv : V.pu;
    const R* gn = xface ? V.gyu() : V.gxv();   // tangential derivative of the normal velocity
    const R* gt = xface ? V.gyv() : V.gxu();   // tangential derivative of the tangential velocity
    R acc[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        R dn_un = R(0), dn_ut = R(0), dn_T = R(0);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            dn_un += VT<R>::thr[r][k] * un[cs[k]];
            dn_ut += VT<R>::thr[r][k] * ut[cs[k]];
            dn_T += VT<R>::thr[r][k] * V.pT[cs[k]];
        }
        // x/h == x*(1/h) bit for bit when h is a power of two
        dn_un = h_pow2 ? dn_un * inv_h : K<R>::div(dn_un, h);
        dn_ut = h_pow2 ? dn_ut * inv_h : K<R>::div(dn_ut, h);
        dn_T = h_pow2 ? dn_T * inv_h : K<R>::div(dn_T, h);
        R dt_un = R(0), dt_ut = R(0), unf = R(0), utf = R(0);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            dt_un += VT<R>::zr[r][m] * gn[gs[m]];
            dt_ut += VT<R>::zr[r][m] * gt[gs[m]];
            unf += VT<R>::zr[r][m] * un[cs[1 + m]];
            utf += VT<R>::zr[r][m] * ut[cs[1 + m]];
        }
        const R tnn = mu * ((R(4.0) / R(3.0)) * dn_un - (R(2.0) / R(3.0)) * dt_ut);
        const R tnt = mu * (dt_un + dn_ut);
        acc[r][0] = tnn;
        acc[r][1] = tnt;
        acc[r][2] = (unf * tnn + utf * tnt) + kappa * dn_T;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
        out[q] = ((R(-1.0) / R(24.0)) * acc[0][q] + (R(26.0) / R(24.0)) * acc[1][q]) + (R(-1.0) / R(24.0)) * acc[2][q];
}

// One face of a tile, flux to LDS.  s indexes the tile's faces: x-faces
// [0, NXF) (face i0 + s / TJ at column s % TJ), then y-faces [NXF, NXF + NYF).
// DIR = 1 or 2 fixes the face's direction at compile time (its LDS strides and
// flux components become constants); DIR = 0 decides it from s at run time
// (one body for both directions: fewer registers, more index arithmetic).
// REG: no face of this direction in the tile is a degenerate boundary face
// (0 or n of a non-periodic axis), so every stencil is the plain 6 cells and
// its LDS offsets are constants.
template <typename R, bool VISC, bool LIMIT, bool FAST, typename Sh, int DIR, bool HALF, bool REG = false>
__device__ __forceinline__ void tile_face(const StageArgs& a, int s, int i0, int j0, R g, bool xg, unsigned slow,
                                          R lam_x, R lam_y, const R* S, R* F, const ViscLDS<R, Sh>* V,
                                          const LimitLDS<R, Sh>* Lm) {
    constexpr int TJ = Sh::TJ, SC = Sh::SC, NXF = Sh::NXF;
    constexpr int NS = Sh::SR * Sh::SC, NF = Sh::template flux_plane<HALF>();   // LDS plane lengths
    const bool xf = DIR == 0 ? s < NXF : DIR == 1;
    const int nrm = xf ? 1 : 2, tng = xf ? 2 : 1;
    int cell[6], gcell[4];
    if (xf) {
        const bool px = a.bc[0] == BC_PERIODIC;
        const int fr = s / TJ, jl = s - fr * TJ;
        const int f = i0 + fr;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int o = k - 3;
            // boundary faces: the walk stops at the boundary face, so the
            // stencil is [G,G,G,c0,c0,c0] at f = 0 and [c,c,c,G,G,G] at f = n
            if (!REG && !px && f == 0 && o > 0) o = 0;
            if (!REG && !px && f == a.nx && o < -1) o = -1;
            cell[k] = (fr + o + HALO) * SC + jl + HALO;
            if (VISC && k >= 1 && k <= 4) gcell[k - 1] = (fr + o + HALO - 1) * ViscLDS<R, Sh>::GYW + jl;
        }
    } else {
        const bool py = a.bc[2] == BC_PERIODIC;
        const int t = s - NXF;
        const int il = t / (TJ + 1), gl = t - il * (TJ + 1);
        const int gf = j0 + gl;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int o = k - 3;
            if (!REG && !py && gf == 0 && o > 0) o = 0;
            if (!REG && !py && gf == a.ny && o < -1) o = -1;
            cell[k] = (il + HALO) * SC + gl + o + HALO;
            if (VISC && k >= 1 && k <= 4) gcell[k - 1] = il * ViscLDS<R, Sh>::GXW + gl + o + HALO - 1;
        }
    }
    R fl[4];
    face_flux<R, FAST>(S, NS, cell, nrm, tng, g, a.positivity, a.hll_bounds, xg, slow, fl);
    if constexpr (LIMIT) {
        if (a.positivity == 2) {
            int pL, pR;
            R kL, kR;
            limit_cells<R, Sh>(*Lm, xf, s, cell[2], lam_x, lam_y, pL, pR, kL, kR);
            limit_face<R, Sh>(S, NS, *Lm, cell[2], cell[3], pL, pR, kL, kR, nrm, tng, g, fl);
        }
    }
    if constexpr (VISC) {   // net flux = inviscid - viscous (mass: minus zero)
        R fv[3];
        viscous_flux<R, Sh>(*V, cell, gcell, xf, xf ? (R)a.dx : (R)a.dy, xf ? a.dx_pow2 != 0 : a.dy_pow2 != 0,
                            xf ? (R)a.inv_dx : (R)a.inv_dy, (R)a.mu, (R)a.kappa, fv);
        fl[1] = fl[1] - fv[0];
        fl[2] = fl[2] - fv[1];
        fl[3] = fl[3] - fv[2];
    }
    // the face's slot: x-faces [0, NXF), y-faces after them or, with HALF
    // (the y-faces written over the x-faces once those are consumed), from 0
    const int e = (HALF && !xf) ? s - NXF : s;
    F[e] = fl[0];
    F[nrm * NF + e] = fl[1];
    F[tng * NF + e] = fl[2];
    F[3 * NF + e] = fl[3];
}

// x-face s and y-face s of a tile (the dual-face tile shape), computed
// together: the WENO5 reconstructions of both faces and then their HLL fluxes
// run statement by statement over the two faces, two independent dependency
// chains per thread.  Stencils, arithmetic and flux slots are tile_face's.
// The tile has one y-face fewer than x-faces: the last thread computes a copy
// of the y-face before it and does not store it.
template <typename R, bool LIMIT, bool FAST, typename Sh>
__device__ __forceinline__ void tile_face_pair(const StageArgs& a, int s, int i0, int j0, R g, bool xg, unsigned slow,
                                               R lam_x, R lam_y, const R* S, R* F, const LimitLDS<R, Sh>* Lm) {
    constexpr int TJ = Sh::TJ, SC = Sh::SC, NXF = Sh::NXF, NYF = Sh::NYF;
    constexpr int NS = Sh::SR * Sh::SC, NF = NXF + NYF;
    const int sy = s < NYF ? s : NYF - 1;
    int cell[2][6];
    {
        const bool px = a.bc[0] == BC_PERIODIC;
        const int fr = s / TJ, jl = s - fr * TJ;
        const int f = i0 + fr;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int o = k - 3;
            if (!px && f == 0 && o > 0) o = 0;
            if (!px && f == a.nx && o < -1) o = -1;
            cell[0][k] = (fr + o + HALO) * SC + jl + HALO;
        }
    }
    {
        const bool py = a.bc[2] == BC_PERIODIC;
        const int il = sy / (TJ + 1), gl = sy - il * (TJ + 1);
        const int gf = j0 + gl;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int o = k - 3;
            if (!py && gf == 0 && o > 0) o = 0;
            if (!py && gf == a.ny && o < -1) o = -1;
            cell[1][k] = (il + HALO) * SC + gl + o + HALO;
        }
    }
    // variables in each face's normal frame: (rho, m_n, m_t, E)
    constexpr int comp[2][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}};
    R L[2][4], Rr[2][4], cl[2][4], cr[2][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            R v[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = S[comp[f][q] * NS + cell[f][k]];
            if (!FAST && ((slow >> comp[f][q]) & 1u)) {
                weno5_lr<R, false>(v, L[f][q], Rr[f][q]);
            } else {
                weno5_lr<R, true>(v, L[f][q], Rr[f][q]);
            }
            cl[f][q] = v[2];
            cr[f][q] = v[3];
        }
    }
    R fl[2][4];
    riemann_flux_n<R, 2>(L, Rr, cl, cr, g, a.positivity, a.hll_bounds, xg, fl);
    if constexpr (LIMIT) {
        if (a.positivity == 2) {
            int pL[2], pR[2];
            R kL[2], kR[2];
            limit_cells<R, Sh>(*Lm, true, s, cell[0][2], lam_x, lam_y, pL[0], pR[0], kL[0], kR[0]);
            limit_cells<R, Sh>(*Lm, false, NXF + sy, cell[1][2], lam_x, lam_y, pL[1], pR[1], kL[1], kR[1]);
            limit_face<R, Sh>(S, NS, *Lm, cell[0][2], cell[0][3], pL[0], pR[0], kL[0], kR[0], 1, 2, g, fl[0]);
            limit_face<R, Sh>(S, NS, *Lm, cell[1][2], cell[1][3], pL[1], pR[1], kL[1], kR[1], 2, 1, g, fl[1]);
        }
    }
    F[s] = fl[0][0];
    F[NF + s] = fl[0][1];
    F[2 * NF + s] = fl[0][2];
    F[3 * NF + s] = fl[0][3];
    if (s < NYF) {
        const int e = NXF + s;
        F[e] = fl[1][0];
        F[2 * NF + e] = fl[1][1];
        F[NF + e] = fl[1][2];
        F[3 * NF + e] = fl[1][3];
    }
}

// x-face part of a cell's flux divergence, -(F(i+1/2) - F(i-1/2)) / dx, the
// first operand of Cell::comRHS's difference (Cell.cpp:110-125)
template <typename R, typename Sh, bool HALF>
__device__ __forceinline__ R rhs_x(const StageArgs& a, const R* F, int q, int il, int jl) {
    constexpr int TJ = Sh::TJ, NF = Sh::template flux_plane<HALF>();
    const R fl = F[q * NF + il * TJ + jl];
    const R fr = F[q * NF + (il + 1) * TJ + jl];
    // x/d == x*(1/d) bit for bit when d is a power of two
    return a.dx_pow2 ? (-(fr - fl)) * (R)a.inv_dx : K<R>::div(-(fr - fl), (R)a.dx);
}

// Phase 1 of a tile: its 511 faces, fluxes to LDS.  FAST selects the WENO5
// division forms (WOps); the caller picks it per tile.  SPLIT: x-faces, then
// y-faces, each direction its own code; else one loop over both.  HALF (with
// SPLIT): one flux plane-set of NXF faces, which the y-faces overwrite after
// each cell has taken the x part of its divergence into tx (the cells of
// threads tid < TI * TJ).
// 4th-order central cell gradients of u and v where the faces read them,
// entries [q0, q1) of d/dy ([0, NGY)) followed by d/dx ([NGY, NGY + NGX))
// (FourthOrderCS.cpp:120-151 stencil).  The 5-cell stencils stay inside the
// staged region: d/dy columns HALO +- 2 .. HALO + TJ - 1 + 2, d/dx rows
// likewise.
template <typename R, typename Sh>
__device__ __forceinline__ void visc_gradients(const StageArgs& a, ViscLDS<R, Sh>* V, int q0, int q1) {
    using VL = ViscLDS<R, Sh>;
    constexpr int SC = Sh::SC;
    const R rdx = (R)a.dx, rdy = (R)a.dy;
    for (int q = q0 + (int)threadIdx.x; q < q1; q += Sh::NT) {
        R du = R(0), dv = R(0);
        if (q < VL::NGY) {
            const int gr = q / VL::GYW, gc = q - gr * VL::GYW;
            const int s = (gr + 1) * SC + gc + HALO;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                du += VT<R>::thc[k] * V->pu[s - 2 + k];
                dv += VT<R>::thc[k] * V->pv[s - 2 + k];
            }
            V->gyu()[q] = a.dy_pow2 ? du * (R)a.inv_dy : K<R>::div(du, rdy);
            V->gyv()[q] = a.dy_pow2 ? dv * (R)a.inv_dy : K<R>::div(dv, rdy);
        } else {
            const int t = q - VL::NGY;
            const int gr = t / VL::GXW, gc = t - gr * VL::GXW;
            const int s = (gr + HALO) * SC + gc + 1;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                du += VT<R>::thc[k] * V->pu[s + (k - 2) * SC];
                dv += VT<R>::thc[k] * V->pv[s + (k - 2) * SC];
            }
            V->gxu()[t] = a.dx_pow2 ? du * (R)a.inv_dx : K<R>::div(du, rdx);
            V->gxv()[t] = a.dx_pow2 ? dv * (R)a.inv_dx : K<R>::div(dv, rdx);
        }
    }
}

template <typename R, bool VISC, bool LIMIT, bool FAST, typename Sh, bool SPLIT, bool HALF>
__device__ __forceinline__ void tile_faces(const StageArgs& a, int i0, int j0, double dt, unsigned slow, const R* S, R* F,
                                           ViscLDS<R, Sh>* V, const LimitLDS<R, Sh>* Lm, R tx[4]) {
    static_assert(!HALF || SPLIT, "the halved flux planes need the x-faces first");
    constexpr int NXF = Sh::NXF, NF = Sh::NXF + Sh::NYF;
    const int tid = threadIdx.x;
    const R g = (R)a.gamma;
    // the FP64 face states' divisions by their density with one reciprocal
    // where in range (state_terms_exact; the gamma condition checked here)
    const bool xg = sizeof(R) == 8 && exact_state_gamma(a.gamma);
    // the limiter's dt/dx and dt/dy (x/d == x*(1/d) bit for bit when d is a
    // power of two); the FP32 limiter's share factors are per face
    const R lam_x = (R)(a.dx_pow2 ? dt * a.inv_dx : dt / a.dx);
    const R lam_y = (R)(a.dy_pow2 ? dt * a.inv_dy : dt / a.dy);
    if constexpr (SPLIT && Sh::DUAL) {
        static_assert(!HALF && !VISC, "both directions' fluxes are live at once; inviscid only");
        for (int s = tid; s < NXF; s += Sh::NT) tile_face_pair<R, LIMIT, FAST, Sh>(a, s, i0, j0, g, xg, slow, lam_x, lam_y, S, F, Lm);
    } else if constexpr (SPLIT) {
        // a direction whose tile faces include no degenerate boundary face
        // (the fast division forms only: the IEEE form serves rare tiles)
        const bool xreg = FAST && (a.bc[0] == BC_PERIODIC || (i0 > 0 && i0 + Sh::TI < a.nx));
        const bool yreg = FAST && (a.bc[2] == BC_PERIODIC || (j0 > 0 && j0 + Sh::TJ < a.ny));
        if (xreg) {
            for (int s = tid; s < NXF; s += Sh::NT)
                tile_face<R, VISC, LIMIT, FAST, Sh, 1, HALF, FAST>(a, s, i0, j0, g, xg, slow, lam_x, lam_y, S, F, V, Lm);
        } else {
            for (int s = tid; s < NXF; s += Sh::NT)
                tile_face<R, VISC, LIMIT, FAST, Sh, 1, HALF>(a, s, i0, j0, g, xg, slow, lam_x, lam_y, S, F, V, Lm);
        }
        if constexpr (HALF) {
            __syncthreads();
            if (tid < Sh::TI * Sh::TJ) {
                const int il = tid / Sh::TJ, jl = tid - il * Sh::TJ;
#pragma unroll
                for (int q = 0; q < 4; ++q) tx[q] = rhs_x<R, Sh, HALF>(a, F, q, il, jl);
            }
            if constexpr (VISC && ViscLDS<R, Sh>::SHARED)   // d/dx over d/dy, which the x-faces have read
                visc_gradients<R, Sh>(a, V, ViscLDS<R, Sh>::NGY, ViscLDS<R, Sh>::NGY + ViscLDS<R, Sh>::NGX);
            __syncthreads();
        }
        if (yreg) {
            for (int s = NXF + tid; s < NF; s += Sh::NT)
                tile_face<R, VISC, LIMIT, FAST, Sh, 2, HALF, FAST>(a, s, i0, j0, g, xg, slow, lam_x, lam_y, S, F, V, Lm);
        } else {
            for (int s = NXF + tid; s < NF; s += Sh::NT)
                tile_face<R, VISC, LIMIT, FAST, Sh, 2, HALF>(a, s, i0, j0, g, xg, slow, lam_x, lam_y, S, F, V, Lm);
        }
    } else {
        for (int s = tid; s < NF; s += Sh::NT)
            tile_face<R, VISC, LIMIT, FAST, Sh, 0, HALF>(a, s, i0, j0, g, xg, slow, lam_x, lam_y, S, F, V, Lm);
    }
}

// The RK snapshots of cell c the stage's form combines (u0 for forms 1 and 2,
// u1 for form 2), read before the cell's flux divergence is formed so that
// the loads are in flight meanwhile.  NT: non-temporal loads.
template <bool NT>
__device__ __forceinline__ void load_snapshots(const StageArgs& a, long long c, double s0[4], double s1[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const gptr<const double> p0 = G(a.u0) + q * a.plane + c, p1 = G(a.u1) + q * a.plane + c;
        s0[q] = 0.0;
        s1[q] = 0.0;
        if (a.form == 1 || a.form == 2) s0[q] = NT ? __builtin_nontemporal_load(p0) : *p0;
        if (a.form == 2) s1[q] = NT ? __builtin_nontemporal_load(p1) : *p1;
    }
}

// The TVD-RK combination of one cell (TVD2rdRK.cpp:3-31, TVD3rdRK.cpp:3-48)
// from its state u, right-hand side L and the RK snapshots s0, s1, and the
// store of the new state (immersed ghosts were written while staging the
// tile).  NT: the state is written with non-temporal stores.
template <bool NT>
__device__ __forceinline__ void update_cell(const StageArgs& a, long long c, int ty, double dt, const double u[4],
                                            const double L[4], const double s0[4], const double s1[4],
                                            double res[4]) {
    const double c2dt = a.c2 * dt;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        res[q] = u[q];
        if (ty == 0) {
            if (a.form == 0) res[q] = a.c1 * u[q] + c2dt * L[q];
            else if (a.form == 1) res[q] = (a.c0 * s0[q] + a.c1 * u[q]) + c2dt * L[q];
            else if (a.form == 2) res[q] = (a.c0 * s0[q] + a.c1 * s1[q]) + c2dt * L[q];
            else res[q] = L[q];
        } else if (a.form == 3) {
            res[q] = 0.0;
        }
    }
    if (ty != 2 || a.form == 3) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const gptr<double> o = G(a.out) + q * a.plane + c;
            if (NT) __builtin_nontemporal_store(res[q], o);
            else *o = res[q];
        }
    }
}

struct NoHook {
    __device__ void operator()() const {}
};

// PER_VAR: a tile outside the fast WENO5 range takes the IEEE form for the
// variables that hold the out-of-range values only (else for all four: the
// batched kernel, whose registers do not take both forms of every variable).
// PEEL: the first staging pass is peeled off the staging loop, which keeps
// the FP32-bulk single-solve kernels free of register spills (0.0881 ->
// 0.0851 ms on the Mach-3 cylinder); the batched kernel, which also holds its
// problem's arguments, spills with it and keeps the plain loop.
// `staged` runs once the tile is staged (after the barrier that ends the
// staging, before the faces): the batched kernel takes its next tile there.
template <typename R, bool VISC, bool LIMIT, typename Sh, bool SPLIT = true, bool HALF = false, bool PER_VAR = true,
          bool PEEL = true, typename Hook = NoHook>
__device__ __forceinline__ void stage_tile(const StageArgs& a, int tile, R* S, R* F, double* wmax,
                                           ViscLDS<R, Sh>* V, LimitLDS<R, Sh>* Lm, const Hook& staged = Hook()) {
    constexpr int TI = Sh::TI, TJ = Sh::TJ, SR = Sh::SR, SC = Sh::SC, NXF = Sh::NXF, BLOCK = Sh::NT;
    constexpr int NS = SR * SC, NF = Sh::template flux_plane<HALF>();   // LDS plane lengths (state, fluxes)
    // y-faces' first slot in the flux planes
    constexpr int YF = HALF ? 0 : NXF;
    const int tid = threadIdx.x;
    const int ti = tile / a.tiles_j;
    const int tj = tile - ti * a.tiles_j;
    const int i0 = ti * TI, j0 = tj * TJ;

    if (tile == 0 && tid == 0) *a.exact_other = 0;   // the previous stage's counter, consumed

    bool unsafe_q[4] = {false, false, false, false};   // per variable (state plane)

    // ---- tile + halo -> LDS (corners are never read) ----
    const bool near_ib = a.tile_ib != nullptr && G(a.tile_ib)[tile];
    // a tile whose halo lies inside the grid reads its cells directly: no
    // boundary ghost, no periodic wrap (the loads fetch_cell makes there)
    const bool interior = i0 >= HALO && i0 + TI + HALO <= a.nx && j0 >= HALO && j0 + TJ + HALO <= a.ny;
    // the limiter's per-cell values (positivity = 2) are taken as each cell is staged
    const bool limit = LIMIT && a.positivity == 2;
    // FP64: the cells a thread stages come two per pass (s and s + BLOCK), both
    // read from memory before either is written to LDS, one memory latency per
    // pass (FP32-bulk kernels spill with both cells live: one per pass)
    constexpr int PER = sizeof(R) == 8 ? 2 : 1;
    auto needed = [&](int s) {   // viscous cross terms and the limiter read the corners
        const int r = s / SC, c = s - r * SC;
        return VISC || LIMIT || (c >= HALO && c < HALO + TJ) || (r >= HALO && r < HALO + TI);
    };
    auto load = [&](int s, double v[4]) -> long long {
        const int r = s / SC, c = s - r * SC;
        if (interior) {
            const long long cell = (long long)(i0 - HALO + r) * a.ny + (j0 - HALO + c);
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = G(a.in)[q * a.plane + cell];
            return cell;
        }
        return fetch_cell(a, i0 - HALO + r, j0 - HALO + c, v);
    };
    auto store = [&](int s, double v[4], long long cell) {
        const int r = s / SC, c = s - r * SC;
        const bool xs = (c >= HALO && c < HALO + TJ), ys = (r >= HALO && r < HALO + TI);
        if (near_ib && cell >= 0 && G(a.type)[cell] == 2) {
            // immersed ghost: its value at the start of this stage, from its donors
            const int k = a.ig_index[cell];
            ghost_value(a.in, a.plane, a.ib_interp, a.ig_donors + 4 * k, a.ig_coef + 10 * k, a.ig_transp + 4 * k,
                        a.ig_xy + 2 * k, v);
            if (a.form != 3 && xs && ys && i0 - HALO + r < a.nx && j0 - HALO + c < a.ny) {   // own cell
#pragma unroll
                for (int q = 0; q < 4; ++q) a.out[q * a.plane + cell] = v[q];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            S[q * NS + s] = (R)v[q];
            if (sizeof(R) == 8) unsafe_q[q] = unsafe_q[q] || !fast_div_safe(v[q]);
        }
        if constexpr (LIMIT) {
            if (limit)
                limit_cell<R, Sh>(s, r - HALO, c - HALO, (R)v[0], (R)v[1], (R)v[2], (R)v[3], (R)a.gamma, F, Lm);
        }
    };
    // the first staging pass (each thread's first cells), peeled off the loop
    double f0[4], f1[4];
    long long fc0 = -1, fc1 = -1;
    const bool fn0 = PEEL && needed(tid), fn1 = PEEL && PER == 2 && tid + BLOCK < SR * SC && needed(tid + BLOCK);
    const bool early = PEEL && a.stage_early;
    if (early) {
        if (fn0) fc0 = load(tid, f0);
        if (fn1) fc1 = load(tid + BLOCK, f1);
    }
    // ---- time step ----
    double dt;
    if (a.dt_mode == 1) {
        // first stage of a CFL step: every workgroup derives the same dt from
        // the bound accumulated by the previous step's last stage
        const double t = *a.t_in;
        // a halt is sticky until the host clears it (workgroup 0 rewrites the
        // flag with the value every workgroup derives, so a read racing with
        // that write sees the same decision)
        bool stop = *a.halt != 0 || (a.t_end > 0.0 && !(t < a.t_end));
        bool bad = false;
        double d = 0.0;
        if (!stop) {
            const double s = __longlong_as_double((long long)*a.acc_in);
            d = a.cfl / s;
            if (!(s > 0.0) || !(d < 1.0e300)) {
                bad = true;
                stop = true;
                d = 0.0;
            } else if (a.t_end > 0.0) {
                const double rem = a.t_end - t;
                if (!(d < rem)) d = rem > 0.0 ? rem : 0.0;
            }
        }
        if (tile == 0 && tid == 0) {
            *a.dt = d;
            *a.t_out = t + d;
            *a.halt = stop ? 1 : 0;
            if (bad) *a.status = 1;
            if (!stop) *a.steps += 1;
            *a.acc_out = 0ull;   // consumed two steps ago; this step's last stage fills it
        }
        if (stop) return;
        dt = d;
    } else if (a.dt_mode == 2) {
        if (*a.halt) return;
        dt = *a.dt;
    } else {
        dt = *a.dt;
    }

    // otherwise the first pass is read once the time step is known (a halted
    // stage reads nothing)
    if (!early) {
        if (fn0) fc0 = load(tid, f0);
        if (fn1) fc1 = load(tid + BLOCK, f1);
    }
    if (fn0) store(tid, f0, fc0);
    if (fn1) store(tid + BLOCK, f1, fc1);
    for (int s0 = tid + (PEEL ? PER * BLOCK : 0); s0 < SR * SC; s0 += PER * BLOCK) {
        const int s1 = s0 + BLOCK;
        const bool n0 = needed(s0), n1 = PER == 2 && s1 < SR * SC && needed(s1);
        double v0[4], v1[4];
        long long c0 = -1, c1 = -1;
        if (n0) c0 = load(s0, v0);
        if (n1) c1 = load(s1, v1);
        if (n0) store(s0, v0, c0);
        if (n1) store(s1, v1, c1);
    }
    // one barrier + OR over the tile: tiles outside the range of the fast
    // WENO5 division forms take the IEEE form (and are counted)
    const bool unsafe = unsafe_q[0] || unsafe_q[1] || unsafe_q[2] || unsafe_q[3];
    const bool tile_unsafe = __syncthreads_or(unsafe ? 1 : 0) != 0;
    // such a tile takes the IEEE form for the variables that hold the
    // out-of-range values only (bit p: state plane p), OR-ed over the block
    unsigned slow = PER_VAR ? 0u : 0xFu;
    if (sizeof(R) == 8 && PER_VAR && tile_unsafe) {
        unsigned wm = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) wm |= (__ballot(unsafe_q[q] ? 1 : 0) != 0ull) ? (1u << q) : 0u;
        unsigned* m = reinterpret_cast<unsigned*>(wmax);   // free until the block maximum at the end
        if ((tid & 63) == 0) m[tid >> 6] = wm;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) slow |= m[w];
        __syncthreads();
    }
    if (sizeof(R) == 8 && tile_unsafe && tid == 0) atomicAdd(a.exact_count, 1);
    staged();

    if constexpr (LIMIT) {
        if (a.positivity == 2) {
            // dt/dx and dt/dy (x/d == x*(1/d) bit for bit when d is a power of two)
            const R lam_x = (R)(a.dx_pow2 ? dt * a.inv_dx : dt / a.dx);
            const R lam_y = (R)(a.dy_pow2 ? dt * a.inv_dy : dt / a.dy);
            limit_share<R, Sh>(F, Lm, lam_x, lam_y);
        }
    }

    if constexpr (VISC) {
        // primitive variables of every staged cell (EulerPhase.cpp:130-137 forms)
        const R g = (R)a.gamma, gm1 = g - R(1), gma2 = g * (R)a.ma2;
        for (int s = tid; s < SR * SC; s += BLOCK) {
            const R rho = S[s], mx = S[NS + s], my = S[2 * NS + s], E = S[3 * NS + s];
            const R p = gm1 * (E - K<R>::div(K<R>::half * (mx * mx + my * my), rho));
            V->pu[s] = K<R>::div(mx, rho);
            V->pv[s] = K<R>::div(my, rho);
            V->pT[s] = K<R>::div(gma2 * p, rho);
        }
        __syncthreads();
        // cell gradients where the faces read them: d/dy and d/dx, or with
        // shared storage d/dy now and d/dx once the x-faces are done
        using VL = ViscLDS<R, Sh>;
        static_assert(!VL::SHARED || HALF, "shared gradient storage needs the x-faces first");
        visc_gradients<R, Sh>(a, V, 0, VL::SHARED ? VL::NGY : VL::NGY + VL::NGX);
        __syncthreads();
    }

    // ---- faces: the exact-division form on tiles outside the fast range ----
    R tx[4];   // with HALF: the x part of the cell's flux divergence
    if (sizeof(R) == 8 && tile_unsafe)
        tile_faces<R, VISC, LIMIT, false, Sh, SPLIT, HALF>(a, i0, j0, dt, slow, S, F, V, Lm, tx);
    else
        tile_faces<R, VISC, LIMIT, true, Sh, SPLIT, HALF>(a, i0, j0, dt, 0u, S, F, V, Lm, tx);
    __syncthreads();

    // ---- cell update: flux divergence (Cell::comRHS, Cell.cpp:110-125) and the
    // TVD-RK combination (TVD2rdRK.cpp:3-31, TVD3rdRK.cpp:3-48) ----
    double smax = 0.0;
    if (tid < TI * TJ) {
        const int il = tid / TJ, jl = tid - il * TJ;
        const int i = i0 + il, j = j0 + jl;
        if (i < a.nx && j < a.ny) {
            const long long c = (long long)i * a.ny + j;
            const int ty = (a.type == nullptr) ? 0 : G(a.type)[c];
            const R rdy = (R)a.dy, idy = (R)a.inv_dy;
            // the two forms of the snapshot reads and state writes stay apart
            // (the empty asm statements keep the compiler from merging them,
            // which would drop the non-temporal hint)
            double s0[4], s1[4];
            if (a.stream_nt) {
                asm volatile("" ::: "memory");
                load_snapshots<true>(a, c, s0, s1);
                asm volatile("" ::: "memory");
            } else {
                load_snapshots<false>(a, c, s0, s1);
            }
            double u[4], L[4], res[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // the FP64 tile holds the exact state; FP32-bulk re-reads it
                u[q] = (sizeof(R) == 8) ? (double)S[q * NS + (il + HALO) * SC + jl + HALO] : G(a.in)[q * a.plane + c];
                L[q] = 0.0;
                if (ty == 0) {
                    const R gb = F[q * NF + YF + il * (TJ + 1) + jl];
                    const R gt = F[q * NF + YF + il * (TJ + 1) + jl + 1];
                    const R txq = HALF ? tx[q] : rhs_x<R, Sh, HALF>(a, F, q, il, jl);
                    // x/d == x*(1/d) bit for bit when d is a power of two
                    const R tyy = a.dy_pow2 ? (gt - gb) * idy : K<R>::div(gt - gb, rdy);
                    L[q] = (double)(txq - tyy);
                }
            }
            if (a.stream_nt) {
                asm volatile("" ::: "memory");
                update_cell<true>(a, c, ty, dt, u, L, s0, s1, res);
                asm volatile("" ::: "memory");
            } else {
                update_cell<false>(a, c, ty, dt, u, L, s0, s1, res);
            }
            if (a.fuse_cfl && ty == 0)
                smax = (sizeof(R) == 8)
                           ? cell_speed(res[0], res[1], res[2], res[3], a.gamma, a.dx, a.dy, a.hll_bounds)
                           : cell_speed_bulk(res[0], res[1], res[2], res[3], a.gamma, a.inv_dx, a.inv_dy, a.hll_bounds);
        }
    }
    if (a.fuse_cfl) block_max_to(smax, a.acc_out, wmax);
}

// Waves per SIMD the stage kernels are compiled for: the second argument of
// __launch_bounds__ is the minimum number of waves per execution unit (SIMD)
// on this target, which sets the VGPR budget (512 / waves).  The wanted
// figure -- FP64 5, FP32 6, with the limiter FP64 5, viscous FP64
// 4 and FP32 6 -- is capped by the waves per SIMD that the
// workgroup's LDS (160 KB per CU) allows.
template <typename R, bool VISC, bool LIMIT, typename Sh> struct Occupancy {
    // the single-solve FP64 limiter and viscous kernels keep one flux
    // plane-set of x-faces (HALF): with the limiter 31.5 KB instead of 39.5 KB
    // of LDS per 15 x 16 tile, five workgroups per CU instead of four;
    // viscous (its gradient planes shared too) 38 KB instead of 52 KB, four
    // instead of three
    static constexpr bool half_flux = sizeof(R) == 8 && (LIMIT || VISC) && !Sh::DUAL;
    static constexpr int want = VISC ? (sizeof(R) == 8 ? 4 : 6)
                              : Sh::DUAL ? ((LIMIT && sizeof(R) == 8) ? 3 : 4)
                              : LIMIT ? (sizeof(R) == 8 ? 5 : 6) : (sizeof(R) == 8 ? 5 : 6);
    static constexpr int extra = VISC ? (int)sizeof(ViscLDS<R, Sh>) : LIMIT ? (int)sizeof(LimitLDS<R, Sh>) : 0;
    static constexpr int lds = (int)sizeof(R) * (Sh::SR * Sh::SC * 4 + Sh::template flux_plane<half_flux>() * 4) +
                               8 * (Sh::NT / 64) + extra;
    static constexpr int by_lds = (160 * 1024 / lds) * (Sh::NT / 64) / 4;
    static constexpr int waves_per_simd = want < by_lds ? want : (by_lds > 0 ? by_lds : 1);
    // the batched kernel also holds its problem's arguments, which at the
    // single-solve figure spill: at most five waves (FP32) or four (FP64)
    static constexpr int batch_cap = sizeof(R) == 8 ? 4 : 5;
    static constexpr int batch_waves_per_simd = waves_per_simd < batch_cap ? waves_per_simd : batch_cap;
    // faces of the batched kernel, one body per direction in FP64 (8,138 ->
    // 8,363 Mcell-updates/s on 256 Riemann problems); FP32 keeps one loop over
    // both (0.3% faster: split, it spills at five waves), and so does the
    // 1024-thread shape (split, it spills in FP64)
    static constexpr bool batch_split = sizeof(R) == 8 && Sh::NT <= 512;
};

// The stage kernel: one workgroup per tile.  With 15 x 16 tiles: FP64 at
// most 96 VGPRs so that five workgroups (20 waves) share a CU; FP32 six; the
// viscous variant's LDS (38 KB in FP64, HALF) allows four, the limiter's (31.5 KB, HALF) five.
//
// Tile order: workgroups are dispatched round-robin over the 8 XCDs
// (workgroup b to XCD b % 8), each with its own L2.  The grid is 8 x ceil(n/8)
// and workgroup b takes tile (b % 8) * ceil(n/8) + b / 8, so each XCD sweeps
// one contiguous band of tiles and neighbouring tiles, which share halo
// cells, are staged through the same L2.
template <typename R, bool VISC, bool LIMIT, typename Sh>
__global__ __launch_bounds__(Sh::NT, (Occupancy<R, VISC, LIMIT, Sh>::waves_per_simd)) void stage_kernel(StageArgs a) {
    constexpr bool HALF = Occupancy<R, VISC, LIMIT, Sh>::half_flux;
    __shared__ R S[Sh::SR * Sh::SC * 4];
    __shared__ R F[Sh::template flux_plane<HALF>() * 4];
    __shared__ double wmax[Sh::NT / 64];
    const int chunk = (a.ntiles + 7) >> 3;
    const int tile = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
    if (tile >= a.ntiles) return;
    if constexpr (VISC) {
        __shared__ ViscLDS<R, Sh> V;
        stage_tile<R, true, false, Sh, true, HALF>(a, tile, S, F, wmax, &V, nullptr);
    } else if constexpr (LIMIT) {
        __shared__ LimitLDS<R, Sh> Lm;
        stage_tile<R, false, true, Sh, true, HALF>(a, tile, S, F, wmax, nullptr, &Lm);
    } else {
        stage_tile<R, false, false, Sh>(a, tile, S, F, wmax, nullptr, nullptr);
    }
}

// Batched stage: a persistent grid walks a device-resident list of (problem
// slot, tile) pairs; the arguments of slot p for step parity q and sub-stage s
// are descs[(p * 2 + q) * 3 + s].  The list is cut into 8 contiguous ranges,
// one per XCD (workgroup b is dispatched to XCD b % 8), so tiles that share
// halo cells are staged through the same L2.  Every problem's arithmetic is
// the single-solve kernel's, so a problem's result does not depend on what
// else is in the batch, nor on which workgroup takes which tile.
//
// Tiles are taken in list order, so that the tiles in flight at any time are
// neighbours in the list and a tile's halo lines are still in the L2 when the
// next row of tiles reads them (with a fixed stride the workgroups drift
// rounds apart: 1.19x the algorithmic reads on 64 problems of 512^2, 1.27x on
// 16 of 1024^2, against 1.01x in list order).  Each range is cut into
// BATCH_SUBRANGES contiguous parts with a counter each, every counter on a
// cache line of its own (device-scope atomics on one line are served one at
// a time: two counters on one line were 7-9% slower than on two); more parts
// put more windows of tiles in flight in one L2 (eight: 1.28x the reads at
// 1024^2).  A part's workgroups take its first tiles by their index and the
// rest from its counter, a tile ahead (the atomic is issued once the current
// tile is staged, so its latency overlaps the faces).  Counters are kept per
// (step parity, sub-stage): a launch zeroes those of every other sub-stage,
// and consecutive launches are always of different sub-stages (a step runs
// sub-stages 0 .. rk - 1, rk >= 2), so each launch finds its own at zero.
constexpr int BATCH_SUBRANGES = 2;
constexpr int BATCH_SCHED_PAD = 32;   // ints per counter: one 128-byte line each
constexpr int BATCH_SCHED_INTS = 6 * 8 * BATCH_SUBRANGES * BATCH_SCHED_PAD;   // [parity * 3 + sub][range][part]

template <typename R, bool LIMIT, typename Sh>
__global__ __launch_bounds__(Sh::NT, (Occupancy<R, false, LIMIT, Sh>::batch_waves_per_simd)) void batch_stage_kernel(
    const StageArgs* __restrict__ descs, const int2* __restrict__ tiles, const int* __restrict__ ntiles, int parity,
    int sub, int* __restrict__ sched) {
    constexpr int NP = BATCH_SUBRANGES, PER_SET = 8 * NP;
    __shared__ R S[Sh::SR * Sh::SC * 4];
    __shared__ R F[(Sh::NXF + Sh::NYF) * 4];
    __shared__ double wmax[Sh::NT / 64];
    __shared__ int next_tile;
    using LimitT = typename std::conditional<LIMIT, LimitLDS<R, Sh>, char>::type;
    __shared__ LimitT Lm;
    const int set = parity * 3 + sub;
    if (blockIdx.x == 0) {   // the other sub-stages' counters, for their launches
        for (int i = (int)threadIdx.x; i < 5 * PER_SET; i += Sh::NT) {
            const int o = i / PER_SET, os = o < set ? o : o + 1;
            atomicExch(&sched[(os * PER_SET + i % PER_SET) * BATCH_SCHED_PAD], 0);
        }
    }
    const long long n = *ntiles;
    const int per = (int)(gridDim.x >> 3);                 // workgroups of a range
    const int np = per < NP ? per : NP;                    // parts: at least one workgroup each
    const int x = (int)(blockIdx.x & 7), l = (int)(blockIdx.x >> 3);
    const int r = l % np, q = l / np;                      // part, and rank among the part's workgroups
    const int wr = (per - r + np - 1) / np;                // workgroups of part r
    const long long lo = (n * x) >> 3, len = ((n * (x + 1)) >> 3) - lo;
    const int plo = (int)(lo + len * r / np), phi = (int)(lo + len * (r + 1) / np);
    int* ctr = &sched[(set * PER_SET + x * NP + r) * BATCH_SCHED_PAD];
    const bool counted = plo + wr < phi;                   // the part has tiles beyond the first round
    const bool lead = threadIdx.x == 0;
    int t = plo + q;
    while (t < phi) {
        int ahead = phi;
        bool taken = false;
        auto take = [&]() {
            if (lead && counted) ahead = plo + wr + atomicAdd(ctr, 1);
            taken = true;
        };
        const int2 pt = tiles[t];
        constexpr bool split = Occupancy<R, false, LIMIT, Sh>::batch_split;
        if constexpr (LIMIT)
            stage_tile<R, false, true, Sh, split, false, false, false>(descs[(pt.x * 2 + parity) * 3 + sub], pt.y,
                                                                       S, F, wmax, nullptr, &Lm, take);
        else
            stage_tile<R, false, false, Sh, split, false, false, false>(descs[(pt.x * 2 + parity) * 3 + sub], pt.y,
                                                                        S, F, wmax, nullptr, nullptr, take);
        if (!taken) take();   // a halted problem's stage returns before its tile is staged
        __syncthreads();      // the next tile reuses S, F and wmax; next_tile is read
        if (lead) next_tile = ahead;
        __syncthreads();
        t = next_tile;
    }
}

// ---------------------------------------------------------------------------
// Immersed-boundary fill of a whole state (the reference's prologue fill,
// Run.cpp:106-112); during stages the fill happens in the tile load.
struct IBArgs {
    double* U;
    long long plane;
    int n;
    int interp;           // 0: reference LDL^T solve, 1: donor weights
    const int* cells;
    const int* donors;
    const double* ipxy;
    const double* ldl;    // L10 L20 L21 L30 L31 L32 D0 D1 D2 D3, or 4 weights
    const int* transp;
};

static __global__ __launch_bounds__(256) void ib_fill_kernel(IBArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    double v[4];
    ghost_value(a.U, a.plane, a.interp, a.donors + 4 * k, a.ldl + 10 * k, a.transp + 4 * k, a.ipxy + 2 * k, v);
    const long long c = a.cells[k];
#pragma unroll
    for (int q = 0; q < 4; ++q) a.U[q * a.plane + c] = v[q];
}

// ---------------------------------------------------------------------------
// CFL reduction over a whole state (first step of an advance, cfs_compute_dt):
// wave64 shuffles, LDS, one atomicMax on the bit pattern (non-negative doubles
// order as unsigned integers; a NaN wins) -- exact, so order-free.
// bulk: the FP32-bulk form of the bound (cell_speed_bulk), the one its
// stages fold into their last sub-stage, so that an advance that starts from
// a stored state takes the steps an uninterrupted one takes.
static __global__ __launch_bounds__(256) void cfl_kernel(const double* __restrict__ U, long long plane,
                                                  const uint8_t* __restrict__ type, double gamma,
                                                  double dx, double dy, int hll_bounds, int bulk,
                                                  unsigned long long* acc) {
    __shared__ double wmax[4];
    double m = 0.0;
    const double inv_dx = 1.0 / dx, inv_dy = 1.0 / dy;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < plane;
         c += (long long)gridDim.x * blockDim.x) {
        if (type && type[c] != 0) continue;
        const double s = bulk ? cell_speed_bulk(U[c], U[plane + c], U[2 * plane + c], U[3 * plane + c], gamma,
                                                inv_dx, inv_dy, hll_bounds)
                              : cell_speed(U[c], U[plane + c], U[2 * plane + c], U[3 * plane + c], gamma, dx, dy,
                                           hll_bounds);
        m = max_nan(m, s);
    }
    block_max_to(m, acc, wmax);
}

// plain streaming copy (16-byte vectors), the byte-counter calibration kernel
static __global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

static __global__ void set_scalar_kernel(double* p, double v) { *p = v; }
static __global__ void set_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

// ---------------------------------------------------------------------------
// Fixed-order checksum tree: chunks of 256 summed by halving in LDS
// (a[t] += a[t+h], h = 128..1); the chunk sums form the next level.
static __global__ __launch_bounds__(256) void tree_sum_kernel(const double* __restrict__ src, long long n,
                                                       long long src_stride, const uint8_t* __restrict__ type,
                                                       double* __restrict__ dst, long long dst_stride) {
    __shared__ double buf[256];
    const int q = blockIdx.y;
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (k < n) {
        v = src[q * src_stride + k];
        if (type && type[k] != 0) v = 0.0;
    }
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) buf[threadIdx.x] = buf[threadIdx.x] + buf[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[q * dst_stride + blockIdx.x] = buf[0];
}

}  // namespace cfs

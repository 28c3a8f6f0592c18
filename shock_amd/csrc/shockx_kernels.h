// shockx_kernels.h -- gfx950 device code for the SHOCK RHS / RK-stage path (FP64 form).
//
// Layout: ghost-padded SoA, linear index i*ng1*ng2 + j*ng2 + k (k fastest), the reference's
// own layout (src/Functions.c:56-103), so host arrays cross the C ABI without a transpose.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shx {

// Physical and scheme constants shared by every kernel (kernel-argument struct).
struct Consts {
    double gamma, upsilon, psi, gam_heat, suth, eps;
    double gm1_ups;   // (gamma - 1) / Upsilon
    double mach;      // reference Mach number (post-processing)
    double zd_mid[10], zd_int[10], zd_der[11];
    int mms;
};

// Pointer to a device array (hipMalloc'd, so in the global address space).  Element accesses
// cast to that address space, so the compiler emits global loads and stores even where the
// pointer itself was read from memory (the per-case records of the batched kernels); a plain
// pointer read from memory is generic and gets flat instructions, which share the LDS wait
// counter and serialise against the LDS traffic of the sweeps.  Host code assigns plain pointers.
template <class T>
struct gptr {
    T *p;
    gptr() = default;
    __host__ __device__ gptr(T *q) : p(q) {}
    template <class U>
    __host__ __device__ gptr(const gptr<U> &o) : p(o.p) {}
    __device__ __forceinline__ T &operator[](const long i) const
    {
        return *(T *)&((__attribute__((address_space(1))) T *)p)[i];
    }
    __host__ __device__ explicit operator bool() const { return p != nullptr; }
};
using gdp = gptr<double>;
using gcdp = gptr<const double>;

struct Fields5 { gdp f[5]; };
struct CFields5 { gcdp f[5]; };
struct Fields10 { gdp f[10]; };
struct CFields10 { gcdp f[10]; };
struct CFields12 { gcdp f[12]; };
struct Fields12 { gdp f[12]; };

struct Geom {
    int ng[3];        // padded extents
    int n[3];         // interior counts (n[2] = 1 in 2D)
    int lo[3];        // padded index of the first interior node (g, or 0 for the 2D z axis)
    long st[3];       // strides; the contiguous axis is padded to a multiple of 16 doubles
    int per[3];       // 1 where the direction is periodic (ghosts wrap), 0 where face writers fill them
    int dims, g;
    long ntot;        // logical padded nodes ng0*ng1*ng2
    long nalloc;      // doubles allocated per field (with the row padding)
};

// linear index of the t-th logical padded node (t in [0, ntot)); ijk receives its indices
__device__ __forceinline__ long padded_node(const Geom &gm, const long t, int ijk[3])
{
    ijk[2] = (int)(t % gm.ng[2]);
    const long r = t / gm.ng[2];
    ijk[1] = (int)(r % gm.ng[1]);
    ijk[0] = (int)(r / gm.ng[1]);
    return (long)ijk[0] * gm.st[0] + (long)ijk[1] * gm.st[1] + (long)ijk[2] * gm.st[2];
}

// ---------------------------------------------------------------------------------------
// Characteristic eigenvectors of the Euler flux Jacobian in direction nh at the arithmetic
// interface average (CalcEigenVectorsIn*Direction, src/WENO.c:1085-1548).
__device__ __forceinline__ void eigvec(const double nx, const double ny, const double nz,
                                       const double u, const double v, const double w,
                                       const double rho, const double cs, const double gamma,
                                       double L[5][5], double R[5][5])
{
    const double s2 = 1.4142135623730951;  // sqrt(2.), correctly rounded
    const double q2 = (u * u + v * v + w * w);
    const double M = sqrt(q2) / cs;
    const double gc = ((gamma - 1.0) / (cs * cs));
    const double ics = (1. / (cs * s2));
    const double un = (u * nx + v * ny + w * nz);
    const double irho = 1. / rho;
    const double hm = 1.0 - 0.5 * (gamma - 1.0) * (M * M);
    const double h0 = q2 / 2. + cs * cs / (gamma - 1.);

    R[0][0] = nx; R[0][1] = ny; R[0][2] = nz; R[0][3] = ics; R[0][4] = ics;
    R[1][0] = u * nx; R[1][1] = u * ny - rho * nz; R[1][2] = u * nz + rho * ny;
    R[1][3] = ics * (u + nx * cs); R[1][4] = ics * (u - nx * cs);
    R[2][0] = v * nx + rho * nz; R[2][1] = v * ny; R[2][2] = v * nz - rho * nx;
    R[2][3] = ics * (v + ny * cs); R[2][4] = ics * (v - ny * cs);
    R[3][0] = w * nx - rho * ny; R[3][1] = w * ny + rho * nx; R[3][2] = w * nz;
    R[3][3] = ics * (w + nz * cs); R[3][4] = ics * (w - nz * cs);
    R[4][0] = q2 / 2. * nx + rho * (v * nz - w * ny);
    R[4][1] = q2 / 2. * ny + rho * (w * nx - u * nz);
    R[4][2] = q2 / 2. * nz + rho * (u * ny - v * nx);
    R[4][3] = ics * (h0 + cs * un);
    R[4][4] = ics * (h0 - cs * un);

    L[0][0] = nx * hm - irho * (v * nz - w * ny);
    L[0][1] = gc * u * nx;
    L[0][2] = gc * v * nx + nz * irho;
    L[0][3] = gc * w * nx - ny * irho;
    L[0][4] = -gc * nx;
    L[1][0] = ny * hm - irho * (w * nx - u * nz);
    L[1][1] = gc * u * ny - nz * irho;
    L[1][2] = gc * v * ny;
    L[1][3] = gc * w * ny + nx * irho;
    L[1][4] = -gc * ny;
    L[2][0] = nz * hm - irho * (u * ny - v * nx);
    L[2][1] = gc * u * nz + ny * irho;
    L[2][2] = gc * v * nz - nx * irho;
    L[2][3] = gc * w * nz;
    L[2][4] = -gc * nz;
    const double gs = (gamma - 1.0) / cs;
    L[3][0] = ics * ((gamma - 1.0) * q2 / 2. - cs * un);
    L[3][1] = (-gs * u + nx) / s2;
    L[3][2] = (-gs * v + ny) / s2;
    L[3][3] = (-gs * w + nz) / s2;
    L[3][4] = (gamma - 1.) * ics;
    L[4][0] = ics * ((gamma - 1.0) * q2 / 2. + cs * un);
    L[4][1] = (-gs * u - nx) / s2;
    L[4][2] = (-gs * v - ny) / s2;
    L[4][3] = (-gs * w - nz) / s2;
    L[4][4] = (gamma - 1.) * ics;
}

// Reciprocal from the hardware estimate (v_rcp_f64) refined by two Newton steps: within an ulp
// of 1/x for the normal, finite arguments the flux kernels pass, at a fraction of the cost of
// the IEEE-exact division sequence.
__device__ __forceinline__ double frcp(const double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Averaged interface state and the scalars the eigenvector entries share.
struct Eig {
    double nx, ny, nz, u, v, w, rho, cs;
    double gm1, q2, hm, gc, ics, irho, un, gs, h0;
};

__device__ __forceinline__ Eig make_eig(double nx, double ny, double nz, double u, double v, double w,
                                        double rho, double cs, double gamma)
{
    Eig e;
    e.nx = nx; e.ny = ny; e.nz = nz; e.u = u; e.v = v; e.w = w; e.rho = rho; e.cs = cs;
    e.gm1 = gamma - 1.0;
    e.q2 = u * u + v * v + w * w;
    const double ic = frcp(cs);
    const double M = sqrt(e.q2) * ic;
    e.hm = 1.0 - 0.5 * e.gm1 * (M * M);
    e.gc = e.gm1 * (ic * ic);
    e.ics = ic * 0.70710678118654752440;
    e.irho = frcp(rho);
    e.un = u * nx + v * ny + w * nz;
    e.gs = e.gm1 * ic;
    e.h0 = e.q2 / 2. + cs * cs * frcp(e.gm1);
    return e;
}

// Row C of the left eigenvector matrix (src/WENO.c:1208-1236).
template <int C>
__device__ __forceinline__ void left_row(const Eig &e, double L[5])
{
    constexpr double is2 = 0.70710678118654752440;
    if constexpr (C == 0) {
        L[0] = e.nx * e.hm - e.irho * (e.v * e.nz - e.w * e.ny);
        L[1] = e.gc * e.u * e.nx; L[2] = e.gc * e.v * e.nx + e.nz * e.irho;
        L[3] = e.gc * e.w * e.nx - e.ny * e.irho; L[4] = -e.gc * e.nx;
    } else if constexpr (C == 1) {
        L[0] = e.ny * e.hm - e.irho * (e.w * e.nx - e.u * e.nz);
        L[1] = e.gc * e.u * e.ny - e.nz * e.irho; L[2] = e.gc * e.v * e.ny;
        L[3] = e.gc * e.w * e.ny + e.nx * e.irho; L[4] = -e.gc * e.ny;
    } else if constexpr (C == 2) {
        L[0] = e.nz * e.hm - e.irho * (e.u * e.ny - e.v * e.nx);
        L[1] = e.gc * e.u * e.nz + e.ny * e.irho; L[2] = e.gc * e.v * e.nz - e.nx * e.irho;
        L[3] = e.gc * e.w * e.nz; L[4] = -e.gc * e.nz;
    } else if constexpr (C == 3) {
        L[0] = e.ics * (e.gm1 * e.q2 / 2. - e.cs * e.un);
        L[1] = (-e.gs * e.u + e.nx) * is2; L[2] = (-e.gs * e.v + e.ny) * is2;
        L[3] = (-e.gs * e.w + e.nz) * is2; L[4] = e.gm1 * e.ics;
    } else {
        L[0] = e.ics * (e.gm1 * e.q2 / 2. + e.cs * e.un);
        L[1] = (-e.gs * e.u - e.nx) * is2; L[2] = (-e.gs * e.v - e.ny) * is2;
        L[3] = (-e.gs * e.w - e.nz) * is2; L[4] = e.gm1 * e.ics;
    }
}

// Row M of the right eigenvector matrix (src/WENO.c:1176-1204): R[M][0..4].
template <int M>
__device__ __forceinline__ void right_row(const Eig &e, double R[5])
{
    if constexpr (M == 0) {
        R[0] = e.nx; R[1] = e.ny; R[2] = e.nz; R[3] = e.ics; R[4] = e.ics;
    } else if constexpr (M == 1) {
        R[0] = e.u * e.nx; R[1] = e.u * e.ny - e.rho * e.nz; R[2] = e.u * e.nz + e.rho * e.ny;
        R[3] = e.ics * (e.u + e.nx * e.cs); R[4] = e.ics * (e.u - e.nx * e.cs);
    } else if constexpr (M == 2) {
        R[0] = e.v * e.nx + e.rho * e.nz; R[1] = e.v * e.ny; R[2] = e.v * e.nz - e.rho * e.nx;
        R[3] = e.ics * (e.v + e.ny * e.cs); R[4] = e.ics * (e.v - e.ny * e.cs);
    } else if constexpr (M == 3) {
        R[0] = e.w * e.nx - e.rho * e.ny; R[1] = e.w * e.ny + e.rho * e.nx; R[2] = e.w * e.nz;
        R[3] = e.ics * (e.w + e.nz * e.cs); R[4] = e.ics * (e.w - e.nz * e.cs);
    } else {
        R[0] = e.q2 / 2. * e.nx + e.rho * (e.v * e.nz - e.w * e.ny);
        R[1] = e.q2 / 2. * e.ny + e.rho * (e.w * e.nx - e.u * e.nz);
        R[2] = e.q2 / 2. * e.nz + e.rho * (e.u * e.ny - e.v * e.nx);
        R[3] = e.ics * (e.h0 + e.cs * e.un); R[4] = e.ics * (e.h0 - e.cs * e.un);
    }
}

// ---- axis-aligned normals (diagonal metric) -----------------------------------------------------
// With n = s e_AX (s = +-1) the eigenvector rows above lose most entries.  eps3(a, b, c) is the
// Levi-Civita symbol; the rows below are the general ones with the zero components dropped, so
// they produce the same values.
__host__ __device__ constexpr int eps3(int a, int b, int c)
{
    return (a == b || b == c || a == c) ? 0 : (((a + 1) % 3 == b) ? 1 : -1);
}
// L[C][q] can be non-zero
__host__ __device__ constexpr bool lnz(int C, int AX, int q)
{
    return (C == AX || C >= 3) ? true : (q == 0 || q == 1 + (3 - C - AX));
}
// R[M][C] can be non-zero
__host__ __device__ constexpr bool rnz(int M, int AX, int C)
{
    return (C >= 3 || M == 4) ? true
         : (M == 0 ? C == AX : (C == AX || eps3(M - 1, C, AX) != 0));
}

template <int C, int AX>
__device__ __forceinline__ void left_row_ax(const Eig &e, const double s, double L[5])
{
    constexpr double is2 = 0.70710678118654752440;
    const double uu[3] = {e.u, e.v, e.w};
    if constexpr (C == AX) {
        L[0] = s * e.hm;
        L[1] = e.gc * e.u * s; L[2] = e.gc * e.v * s; L[3] = e.gc * e.w * s; L[4] = -e.gc * s;
    } else if constexpr (C < 3) {
        constexpr int T = 3 - C - AX;
        constexpr double sg = (double)eps3(C, AX, T);
        L[0] = sg * s * (e.irho * uu[T]);
        L[1 + T] = -sg * s * e.irho;
    } else {
        constexpr double pm = (C == 3) ? 1.0 : -1.0;
        L[0] = e.ics * (e.gm1 * e.q2 / 2. - pm * e.cs * e.un);
#pragma unroll
        for (int q = 0; q < 3; q++) L[1 + q] = (q == AX) ? (-e.gs * uu[q] + pm * s) * is2 : (-e.gs * uu[q]) * is2;
        L[4] = e.gm1 * e.ics;
    }
}

template <int M, int AX>
__device__ __forceinline__ void right_row_ax(const Eig &e, const double s, double R[5])
{
    const double uu[3] = {e.u, e.v, e.w};
    if constexpr (M == 0) {
        R[AX] = s; R[3] = e.ics; R[4] = e.ics;
    } else if constexpr (M < 4) {
        constexpr int q = M - 1;
#pragma unroll
        for (int C = 0; C < 3; C++) {
            if (C == AX) R[C] = uu[q] * s;
            else if (eps3(q, C, AX) != 0) R[C] = -(double)eps3(q, C, AX) * (e.rho * s);
        }
        R[3] = (q == AX) ? e.ics * (uu[q] + s * e.cs) : e.ics * uu[q];
        R[4] = (q == AX) ? e.ics * (uu[q] - s * e.cs) : e.ics * uu[q];
    } else {
#pragma unroll
        for (int C = 0; C < 3; C++) {
            if (C == AX) R[C] = e.q2 / 2. * s;
            else R[C] = e.rho * ((double)eps3(C, 3 - C - AX, AX) * uu[3 - C - AX] * s);
        }
        R[3] = e.ics * (e.h0 + e.cs * e.un);
        R[4] = e.ics * (e.h0 - e.cs * e.un);
    }
}

// Smoothness indicators s_k = IS_k + eps scaled by a common power of two so the largest lies in
// [1/2, 1), each clamped from below at 2^lo.  The nonlinear weights d_k / s_k^2 normalised to sum
// one are unchanged by the common scale; the clamp only touches stencils whose weight is below
// 2^(2 lo) of the largest one's.
template <int N>
__device__ __forceinline__ void scaled_sq(double s[N], const int lo)
{
    double mx = s[0];
#pragma unroll
    for (int i = 1; i < N; i++) mx = fmax(mx, s[i]);
    const int ex = __builtin_amdgcn_frexp_exp(mx);
    const double floor_ = __builtin_amdgcn_ldexp(1.0, lo);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double t = fmax(__builtin_amdgcn_ldexp(s[i], -ex), floor_);
        s[i] = t * t;
    }
}

// Jiang-Shu WENO5 deviation from the central flux (Phi_Function_W5, src/WENO.c:118-204) with one
// reciprocal for the weights: for q_k = s_k^2, w0 = q1 q2 / D and w2 = 3 q0 q1 / D with
// D = q1 q2 + 6 q0 q2 + 3 q0 q1.
__device__ __forceinline__ double phi_w5_fast(const double eps, const double d0, const double d1,
                                              const double d2, const double d3)
{
    const double t0 = d0 - d1, t1 = d0 - 3. * d1;
    const double t2 = d1 - d2, t3 = d1 + d2;
    const double t4 = d2 - d3, t5 = 3. * d2 - d3;
    double q[3] = {13. * t0 * t0 + 3. * t1 * t1 + eps, 13. * t2 * t2 + 3. * t3 * t3 + eps,
                   13. * t4 * t4 + 3. * t5 * t5 + eps};
    scaled_sq<3>(q, -250);
    const double P0 = q[1] * q[2], P2 = 3. * (q[0] * q[1]);
    const double r = frcp(P0 + 6. * (q[0] * q[2]) + P2);
    const double w0 = P0 * r, w2 = P2 * r;
    const double q1 = d0 - 2. * d1 + d2;
    const double qq2 = d1 - 2. * d2 + d3;
    return 1. / 3. * w0 * q1 + 1. / 6. * (w2 - 0.5) * qq2;
}

// ---- FP32 forms for the mixed-precision sweeps ------------------------------------------------
// Same formulas as phi_w5_fast / phi_w9 on locally offset FP32 inputs (node-to-node differences).  epsilon is dropped: FP32 cannot hold the
// reference's 10 sqrt(DBL_MIN), and the clamp in scaled_sq_f gives the ideal weights where all
// indicators vanish, which is what an epsilon far below every indicator does.
template <int N>
__device__ __forceinline__ void scaled_sq_f(float s[N], const int lo)
{
    float mx = s[0];
#pragma unroll
    for (int i = 1; i < N; i++) mx = fmaxf(mx, s[i]);
    const int ex = __builtin_amdgcn_frexp_expf(mx);
    const float floor_ = __builtin_amdgcn_ldexpf(1.0f, lo);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const float t = fmaxf(__builtin_amdgcn_ldexpf(s[i], -ex), floor_);
        s[i] = t * t;
    }
}

__device__ __forceinline__ float phi_w5_f(const float d0, const float d1, const float d2, const float d3)
{
    const float t0 = d0 - d1, t1 = d0 - 3.f * d1;
    const float t2 = d1 - d2, t3 = d1 + d2;
    const float t4 = d2 - d3, t5 = 3.f * d2 - d3;
    float q[3] = {13.f * t0 * t0 + 3.f * t1 * t1, 13.f * t2 * t2 + 3.f * t3 * t3, 13.f * t4 * t4 + 3.f * t5 * t5};
    scaled_sq_f<3>(q, -30);
    const float P0 = q[1] * q[2], P2 = 3.f * (q[0] * q[1]);
    const float r = __builtin_amdgcn_rcpf(P0 + 6.f * (q[0] * q[2]) + P2);
    const float w0 = P0 * r, w2 = P2 * r;
    const float q1 = d0 - 2.f * d1 + d2;
    const float qq2 = d1 - 2.f * d2 + d3;
    return (1.f / 3.f) * w0 * q1 + (1.f / 6.f) * (w2 - 0.5f) * qq2;
}

// The Balsara-Shu indicators rewritten as sums of squares of difference combinations (the
// LDL^T factorisation of each quadratic form in the four first differences of its sub-stencil,
// exact rationals): every term is non-negative, so FP32 keeps them to ~1e-6 relative where the
// point-value form loses 1e-4 to cancellation on smooth data.
__device__ __forceinline__ float phi_w9_f(const float *d)
{
    const float y00 = d[0] + (-163185.f / 45316.f) * d[1] + (100839.f / 22658.f) * d[2] + (-86329.f / 45316.f) * d[3];
    const float y01 = d[1] + (-41089198.f / 14249315.f) * d[2] + (1813059.f / 838195.f) * d[3];
    const float y02 = d[2] + (-350813921.f / 232768856.f) * d[3];
    const float y03 = d[3];
    const float is0 = 22658.f * y00 * y00 + (299235615.f / 90632.f) * y01 * y01 + (39105167808.f / 14249315.f) * y02 * y02 + (116566909305.f / 29096107.f) * y03 * y03;
    const float y10 = d[1] + (-47055.f / 13816.f) * d[2] + (26079.f / 6908.f) * d[3] + (-18079.f / 13816.f) * d[4];
    const float y11 = d[2] + (-12448558.f / 5880675.f) * d[3] + (5686883.f / 5880675.f) * d[4];
    const float y12 = d[3] + (-146761231.f / 232768856.f) * d[4];
    const float y13 = d[4];
    const float is1 = 6908.f * y10 * y10 + (123494175.f / 27632.f) * y11 * y11 + (13035055936.f / 1960225.f) * y12 * y12 + (116566909305.f / 29096107.f) * y13 * y13;
    const float y20 = d[2] + (-37185.f / 13816.f) * d[3] + (15369.f / 6908.f) * d[4] + (-8209.f / 13816.f) * d[5];
    const float y21 = d[3] + (-1032654.f / 838195.f) * d[4] + (5686883.f / 14249315.f) * d[5];
    const float y22 = d[4] + (-20636113.f / 51089708.f) * d[5];
    const float y23 = d[5];
    const float is2 = 6908.f * y20 * y20 + (299235615.f / 27632.f) * y21 * y21 + (8583070944.f / 838195.f) * y22 * y22 + (233133818610.f / 217131259.f) * y23 * y23;
    const float y30 = d[3] + (-94935.f / 45316.f) * d[4] + (35109.f / 22658.f) * d[5] + (-18079.f / 45316.f) * d[6];
    const float y31 = d[4] + (-106728718.f / 110863235.f) * d[5] + (30822003.f / 110863235.f) * d[6];
    const float y32 = d[5] + (-1745642471.f / 4217089556.f) * d[6];
    const float y33 = d[6];
    const float is3 = 22658.f * y30 * y30 + (2328127935.f / 90632.f) * y31 * y31 + (101210149344.f / 15837605.f) * y32 * y32 + (233133818610.f / 1054272389.f) * y33 * y33;
    const float y40 = d[4] + (-433665.f / 215836.f) * d[5] + (162579.f / 107918.f) * d[6] + (-86329.f / 215836.f) * d[7];
    const float y41 = d[5] + (-945305518.f / 890329635.f) * d[6] + (290071763.f / 890329635.f) * d[7];
    const float y42 = d[6] + (-8384962321.f / 17843974736.f) * d[7];
    const float y43 = d[7];
    const float is4 = 107918.f * y40 * y40 + (18696922335.f / 431672.f) * y41 * y41 + (999262585216.f / 296776545.f) * y42 * y42 + (16652415615.f / 318642406.f) * y43 * y43;
    // weights d_k / q_k with the indicators scaled by the largest and clamped at 2^-60 of it, so
    // every q_k >= 2^-120 and the sum stays finite (products of four q, as in the FP64 form,
    // would need a clamp near 2^-15 in FP32, which distorts weights that still matter)
    float q[5] = {is0, is1, is2, is3, is4};
    scaled_sq_f<5>(q, -60);
    const float A0 = (1.f / 126.f) * __builtin_amdgcn_rcpf(q[0]), A1 = (10.f / 63.f) * __builtin_amdgcn_rcpf(q[1]);
    const float A2 = (10.f / 21.f) * __builtin_amdgcn_rcpf(q[2]), A3 = (20.f / 63.f) * __builtin_amdgcn_rcpf(q[3]);
    const float A4 = (5.f / 126.f) * __builtin_amdgcn_rcpf(q[4]);
    const float ia = __builtin_amdgcn_rcpf(A0 + A1 + A2 + A3 + A4);
    const float w0 = A0 * ia, w2 = A2 * ia, w3 = A3 * ia, w4 = A4 * ia;
    const float q1 = d[0] - 4.f * d[1] + 6.f * d[2] - 4.f * d[3] + d[4];
    const float q2 = d[1] - 4.f * d[2] + 6.f * d[3] - 4.f * d[4] + d[5];
    const float q3 = d[2] - 4.f * d[3] + 6.f * d[4] - 4.f * d[5] + d[6];
    const float q4 = d[3] - 4.f * d[4] + 6.f * d[5] - 4.f * d[6] + d[7];
    return (1.f / 5.f) * w0 * q1 + (1.f / 20.f) * w2 * q2 + (1.f / 60.f) * w3 * (3.f * q2 - 2.f * q3)
         + (1.f / 60.f) * w4 * (3.f * q2 - 2.f * q3 + 3.f * q4) - (1.f / 840.f) * (39.f * q2 - 14.f * q3 + 3.f * q4);
}

// Balsara-Shu WENO9 (Phi_Function_W9, src/WENO.c:11-116); f: 9 point values, d: 8 differences.
// The smoothness indicators keep the reference's quadratic forms in the point values.
__device__ __forceinline__ double phi_w9(const double eps, const double *f, const double *d)
{
    const double is0 = f[0] * (22658. * f[0] - 208501. * f[1] + 364863. * f[2] - 288007. * f[3] + 86329. * f[4])
        + f[1] * (482963. * f[1] - 1704396. * f[2] + 1358458. * f[3] - 411487. * f[4])
        + f[2] * (1521393. * f[2] - 2462076. * f[3] + 758823. * f[4])
        + f[3] * (1020563. * f[3] - 649501. * f[4])
        + f[4] * (107918. * f[4]);
    const double is1 = f[1] * (6908. * f[1] - 60871. * f[2] + 99213. * f[3] - 70237. * f[4] + 18079. * f[5])
        + f[2] * (138563. * f[2] - 464976. * f[3] + 337018. * f[4] - 88297. * f[5])
        + f[3] * (406293. * f[3] - 611976. * f[4] + 165153. * f[5])
        + f[4] * (242723. * f[4] - 140251. * f[5])
        + f[5] * (22658. * f[5]);
    const double is2 = f[2] * (6908. * f[2] - 51001. * f[3] + 67923. * f[4] - 38947. * f[5] + 8209. * f[6])
        + f[3] * (104963. * f[3] - 299076. * f[4] + 179098. * f[5] - 38947. * f[6])
        + f[4] * (231153. * f[4] - 299076. * f[5] + 67923. * f[6])
        + f[5] * (104963. * f[5] - 51001. * f[6])
        + f[6] * (6908. * f[6]);
    const double is3 = f[3] * (22658. * f[3] - 140251. * f[4] + 165153. * f[5] - 88297. * f[6] + 18079. * f[7])
        + f[4] * (242723. * f[4] - 611976. * f[5] + 337018. * f[6] - 70237. * f[7])
        + f[5] * (406293. * f[5] - 464976. * f[6] + 99213. * f[7])
        + f[6] * (138563. * f[6] - 60871. * f[7])
        + f[7] * (6908. * f[7]);
    const double is4 = f[4] * (107918. * f[4] - 649501. * f[5] + 758823. * f[6] - 411487. * f[7] + 86329. * f[8])
        + f[5] * (1020563. * f[5] - 2462076. * f[6] + 1358458. * f[7] - 288007. * f[8])
        + f[6] * (1521393. * f[6] - 1704396. * f[7] + 364863. * f[8])
        + f[7] * (482963. * f[7] - 208501. * f[8])
        + f[8] * (22658. * f[8]);
    // weights d_k / q_k normalised with one reciprocal: w_k = d_k prod_{j != k} q_j / D
    double q[5] = {is0 + eps, is1 + eps, is2 + eps, is3 + eps, is4 + eps};
    scaled_sq<5>(q, -120);
    const double p01 = q[0] * q[1], p012 = p01 * q[2], s34 = q[3] * q[4], s234 = q[2] * s34;
    const double P0 = q[1] * s234, P1 = q[0] * s234, P2 = p01 * s34, P3 = p012 * q[4], P4 = p012 * q[3];
    const double A0 = (1. / 126.) * P0, A2 = (10. / 21.) * P2, A3 = (20. / 63.) * P3, A4 = (5. / 126.) * P4;
    const double ia = frcp(A0 + (10. / 63.) * P1 + A2 + A3 + A4);
    const double w0 = A0 * ia, w2 = A2 * ia, w3 = A3 * ia, w4 = A4 * ia;
    const double q1 = d[0] - 4. * d[1] + 6. * d[2] - 4. * d[3] + d[4];
    const double q2 = d[1] - 4. * d[2] + 6. * d[3] - 4. * d[4] + d[5];
    const double q3 = d[2] - 4. * d[3] + 6. * d[4] - 4. * d[5] + d[6];
    const double q4 = d[3] - 4. * d[4] + 6. * d[5] - 4. * d[6] + d[7];
    return 1. / 5. * w0 * q1 + 1. / 20. * w2 * q2 + 1. / 60. * w3 * (3. * q2 - 2. * q3)
         + 1. / 60. * w4 * (3. * q2 - 2. * q3 + 3. * q4) - 1. / 840. * (39. * q2 - 14. * q3 + 3. * q4);
}

// 30 viscous metric factors of one node for direction `dir` (CreateViscidMetric,
// src/Functions.c:6245-6386): the Newtonian stress contracted with metric rows a (= dir) and
// r_b (b = xi, eta, zeta), times J.
__device__ __forceinline__ void visc_fac(const int dir, const double m[10], double fac[30])
{
    const double J = m[9];
    const double a0 = m[3 * dir], a1 = m[3 * dir + 1], a2 = m[3 * dir + 2];
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const double r0 = m[3 * b], r1 = m[3 * b + 1], r2 = m[3 * b + 2];
        fac[0 + b] = J * (4.0 / 3.0 * r0 * a0 + r1 * a1 + r2 * a2);
        fac[3 + b] = J * (-2.0 / 3.0 * r1 * a0 + r0 * a1);
        fac[6 + b] = J * (-2.0 / 3.0 * r2 * a0 + r0 * a2);
        fac[9 + b] = J * (r1 * a0 - 2.0 / 3.0 * r0 * a1);
        fac[12 + b] = J * (r0 * a0 + 4.0 / 3.0 * r1 * a1 + r2 * a2);
        fac[15 + b] = J * (-2.0 / 3.0 * r2 * a1 + r1 * a2);
        fac[18 + b] = J * (r2 * a0 - 2.0 / 3.0 * r0 * a2);
        fac[21 + b] = J * (r2 * a1 - 2.0 / 3.0 * r1 * a2);
        fac[24 + b] = J * (r0 * a0 + r1 * a1 + 4.0 / 3.0 * r2 * a2);
        fac[27 + b] = J * (r0 * a0 + r1 * a1 + r2 * a2);
    }
}

__device__ __forceinline__ double pressure(const Consts &k, double rho, double u, double v, double w, double e)
{
    return fabs(rho * (k.gamma - 1) / k.upsilon * (e - 0.5 * ((u * u) + (v * v) + (w * w))));
}

// p = |rho (gamma-1)/Upsilon (e - |u|^2/2)| with the constant ratio folded (CalcValues,
// src/Functions.c:2445-2459)
__device__ __forceinline__ double pressure_fast(const Consts &k, double rho, double u, double v, double w, double e)
{
    return fabs(rho * k.gm1_ups * (e - 0.5 * ((u * u) + (v * v) + (w * w))));
}

}  // namespace shx
